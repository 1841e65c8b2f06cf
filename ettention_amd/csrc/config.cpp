// Parameter sources and parameter sets (see config.hpp).
#include "config.hpp"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <sstream>

#include "common.hpp"
#include "xml.hpp"

namespace etn {

std::string ParameterSource::path(const std::string& key) const { return normalize_path(string(key)); }

// ---------------------------------------------------------------------------
// XML
// ---------------------------------------------------------------------------

XmlParameterSource::XmlParameterSource(const std::string& filename)
    : filename_(filename), roots_(parse_xml_file(filename))
{
}

const XmlNode* XmlParameterSource::node(const std::string& key) const
{
    std::vector<std::string> steps;
    std::stringstream ss(key);
    std::string item;
    while (std::getline(ss, item, '.')) steps.push_back(item);
    if (steps.empty()) return nullptr;
    const XmlNode* n = nullptr;
    for (const auto& r : roots_)
        if (r->name == steps[0]) {
            n = r.get();
            break;
        }
    for (size_t i = 1; n && i < steps.size(); ++i) n = n->child(steps[i]);
    return n;
}

bool XmlParameterSource::exists(const std::string& key) const { return node(key) != nullptr; }

std::string XmlParameterSource::string(const std::string& key) const
{
    const XmlNode* n = node(key);
    if (!n) fail("xml config file entry does not exist: " + key);
    const std::string* v = n->attribute("value");
    if (!v) fail("xml config file entry " + key + " has no value attribute (" + filename_ + ")");
    return *v;
}

std::string XmlParameterSource::path(const std::string& key) const
{
    return join_path(parent_path(filename_), string(key));
}

// ---------------------------------------------------------------------------
// manual / command line / cascading
// ---------------------------------------------------------------------------

std::string ManualParameterSource::string(const std::string& key) const
{
    auto it = values_.find(key);
    if (it == values_.end()) throw ParameterNotFound(key);
    return it->second;
}

CommandLineParameterSource::CommandLineParameterSource(int argc, const char* const* argv)
{
    auto value_of = [&](int& i, const std::string& opt) -> std::string {
        if (i + 1 >= argc) fail("the required argument for option '" + opt + "' is missing");
        return argv[++i];
    };
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-c" || a == "--config") {
            set("config", value_of(i, a));
            mode_ = Mode::Reconstruct;
        } else if (a.rfind("-c", 0) == 0 && a.size() > 2 && a[1] != '-') {   // sticky -cFILE
            set("config", a.substr(2));
            mode_ = Mode::Reconstruct;
        } else if (a == "-e" || a == "--explicitDeviceId") {
            set("hardware.openclDeviceId", value_of(i, a));
        } else if (a == "-h" || a == "--help") {
            mode_ = Mode::Help;
        } else if (a == "-H" || a == "--parameters") {
            mode_ = Mode::HelpAndParameters;
        } else if (a == "-d" || a == "--devices" || a == "-D" || a == "--devicesAndExtensions") {
            mode_ = Mode::Devices;
        } else if (a.rfind("--", 0) == 0 && a.size() > 2) {
            const size_t eq = a.find('=');
            if (eq != std::string::npos) set(a.substr(2, eq - 2), a.substr(eq + 1));
            else set(a.substr(2), value_of(i, a));
        } else {
            fail("unrecognised command line argument '" + a + "'");
        }
    }
    // the Help / Devices states win over Reconstruct (parseConsoleParameters order)
}

bool CascadingParameterSource::exists(const std::string& key) const
{
    for (const auto& s : sources_)
        if (s->exists(key)) return true;
    return false;
}

std::string CascadingParameterSource::string(const std::string& key) const
{
    for (const auto& s : sources_)
        if (s->exists(key)) return s->string(key);
    throw ParameterNotFound(key);
}

std::string CascadingParameterSource::path(const std::string& key) const
{
    for (const auto& s : sources_)
        if (s->exists(key)) return s->path(key);
    throw ParameterNotFound(key);
}

// ---------------------------------------------------------------------------
// typed reads
// ---------------------------------------------------------------------------

namespace {

[[noreturn]] void bad_value(const std::string& key, const std::string& v, const char* type)
{
    fail("conversion of data to type \"" + std::string(type) + "\" failed for parameter " + key + " ('" + v + "')");
}

template <typename T>
bool read_exact(const std::string& v, T& out)
{
    std::istringstream is(v);
    is >> out;
    if (is.fail()) return false;
    is >> std::ws;
    return is.eof();
}

bool negative(const std::string& v)
{
    for (char c : v) {
        if (std::isspace((unsigned char)c)) continue;
        return c == '-';
    }
    return false;
}

}  // namespace

float get_float(const ParameterSource& s, const std::string& key)
{
    const std::string v = s.string(key);
    float out;
    if (!read_exact(v, out)) bad_value(key, v, "float");
    return out;
}

uint32_t get_uint(const ParameterSource& s, const std::string& key)
{
    const std::string v = s.string(key);
    unsigned long long out;
    if (negative(v) || !read_exact(v, out) || out > 0xffffffffull) bad_value(key, v, "unsigned int");
    return (uint32_t)out;
}

int32_t get_int(const ParameterSource& s, const std::string& key)
{
    const std::string v = s.string(key);
    int32_t out;
    if (!read_exact(v, out)) bad_value(key, v, "int");
    return out;
}

bool get_bool(const ParameterSource& s, const std::string& key)
{
    std::string v = s.string(key);
    const size_t b = v.find_first_not_of(" \t\r\n"), e = v.find_last_not_of(" \t\r\n");
    const std::string t = b == std::string::npos ? "" : v.substr(b, e - b + 1);
    if (t == "true" || t == "1") return true;
    if (t == "false" || t == "0") return false;
    bad_value(key, v, "bool");
}

void get_vec3f(const ParameterSource& s, const std::string& key, float out[3])
{
    const std::string v = s.string(key);
    std::istringstream is(v);
    is >> out[0] >> out[1] >> out[2];
    if (is.fail()) bad_value(key, v, "Vec3f");
    is >> std::ws;
    if (!is.eof()) bad_value(key, v, "Vec3f");
}

void get_vec3ui(const ParameterSource& s, const std::string& key, uint32_t out[3])
{
    const std::string v = s.string(key);
    std::istringstream is(v);
    long long t[3];
    is >> t[0] >> t[1] >> t[2];
    if (is.fail()) bad_value(key, v, "Vec3ui");
    is >> std::ws;
    if (!is.eof()) bad_value(key, v, "Vec3ui");
    for (int k = 0; k < 3; ++k) {
        if (t[k] < 0 || t[k] > 0xffffffffll) bad_value(key, v, "Vec3ui");
        out[k] = (uint32_t)t[k];
    }
}

std::vector<uint32_t> parse_skip_projections(const std::string& s)
{
    std::vector<uint32_t> out;
    if (s == "none") return out;
    auto one = [&](std::string t) -> uint32_t {
        const size_t b = t.find_first_not_of(" \t"), e = t.find_last_not_of(" \t");
        t = b == std::string::npos ? "" : t.substr(b, e - b + 1);
        unsigned long long v;
        if (t.empty() || negative(t) || !read_exact(t, v) || v == 0 || v > 0xffffffffull)
            fail("input.skipProjections: '" + t + "' is not a projection number (they count from 1)");
        return (uint32_t)(v - 1);
    };
    std::stringstream ss(s);
    std::string tok;
    while (std::getline(ss, tok, ',')) {
        if (tok.find_first_not_of(" \t") == std::string::npos) continue;
        const size_t dash = tok.find('-');
        if (dash == std::string::npos) {
            out.push_back(one(tok));
        } else {
            const uint32_t a = one(tok.substr(0, dash)), b = one(tok.substr(dash + 1));
            for (uint32_t j = a; j <= b; ++j) out.push_back(j);
        }
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    return out;
}

// ---------------------------------------------------------------------------
// parameter sets
// ---------------------------------------------------------------------------

ReconstructionConfig ReconstructionConfig::from(const ParameterSource& s)
{
    ReconstructionConfig c;

    // AlgorithmParameterSet: the block size follows the exact (case-sensitive)
    // names; the algorithm itself is chosen case-insensitively (Framework.cpp:142-152).
    c.algorithm = s.string("algorithm");
    c.block_size = 0xffffffffu;
    if (c.algorithm == "sart") c.block_size = 1;
    else if (c.algorithm == "sirt") c.block_size = 0xffffffffu;
    else if (s.exists("algorithm.blockSize")) c.block_size = get_uint(s, "algorithm.blockSize");
    std::string lower = c.algorithm;
    std::transform(lower.begin(), lower.end(), lower.begin(), [](unsigned char ch) { return (char)std::tolower(ch); });
    if (lower != "sart" && lower != "sirt" && lower != "blockiterative" && lower != "wbp")
        unsupported("reconstruction algorithm '" + c.algorithm +
                    "' is not available (this library provides sart, sirt, blockIterative and wbp)");
    if (c.block_size == 0) fail("Illegal ProjectionBlockSize!");
    c.iterations = get_uint(s, "algorithm.numberOfIterations");
    c.lambda = get_float(s, "algorithm.lambda");
    if (s.exists("algorithm.useLongObjectCompensation"))
        c.use_long_object_compensation = get_bool(s, "algorithm.useLongObjectCompensation");
    if (s.exists("algorithm.basisFunctions")) {
        const std::string b = s.string("algorithm.basisFunctions");
        if (b == "blobs") unsupported("the blob basis is not available (voxels only)");
        if (b != "voxels") fail("Illegal value for parameter \"algorithm.basisFunctions\": " + b + "!");
    }
    if (s.exists("algorithm.discreteReconstruction"))
        unsupported("discrete reconstruction (DART) is not available");

    if (s.exists("forwardProjection.samples")) c.fp_samples = get_uint(s, "forwardProjection.samples");
    if (s.exists("backProjection.samples")) c.bp_samples = get_uint(s, "backProjection.samples");

    if (s.exists("optimization.internalVoxelRepresentationType")) {
        c.voxel_representation = s.string("optimization.internalVoxelRepresentationType");
        if (c.voxel_representation == "uchar")
            fail("UChar is not supported as a valid internal voxel storage data type.");
        if (c.voxel_representation != "float" && c.voxel_representation != "half" &&
            c.voxel_representation != "bfloat16")
            fail("Illegal parameter value for voxel representation: " + c.voxel_representation);
    }
    if (s.exists("optimization.projectionIteration"))
        c.projection_iteration = s.string("optimization.projectionIteration");

    if (s.exists("hardware.subVolumeCount")) c.subvolume_count = get_uint(s, "hardware.subVolumeCount");
    if (s.exists("hardware.openclDeviceId")) {
        if (s.exists("hardware.openclDeviceType"))
            fail("parameters hardware.openclDeviceId and hardware.openclDeviceType are exclusive");
        if (s.exists("hardware.openclVendor"))
            fail("parameters hardware.openclDeviceId and hardware.openclVendor are exclusive");
        c.device_id = get_int(s, "hardware.openclDeviceId");
    }

    get_vec3f(s, "volume.dimension", c.dimension);
    if (s.exists("volume.origin")) get_vec3f(s, "volume.origin", c.origin);
    for (int k = 0; k < 3; ++k) {   // Vec3ui(Vec3f): truncation
        if (!(c.dimension[k] >= 0.0f) || c.dimension[k] > 4294967295.0f)
            fail("volume.dimension must be positive");
        c.resolution[k] = (uint32_t)c.dimension[k];
    }
    if (s.exists("volume.resolution")) get_vec3ui(s, "volume.resolution", c.resolution);
    if (s.exists("volume.initValue")) c.init_value = get_float(s, "volume.initValue");
    if (s.exists("volume.initFile")) c.init_file = s.path("volume.initFile");
    if (s.exists("volume.initFileOrientation"))
        c.init_file_orientation = parse_coordinate_order(s.string("volume.initFileOrientation"));

    c.projections = s.path("input.projections");
    if (s.exists("input.tiltAngles")) c.tilt_angles = s.path("input.tiltAngles");
    if (s.exists("input.xAxisTilt")) c.x_axis_tilt = get_float(s, "input.xAxisTilt");
    if (s.exists("input.logarithmize")) c.logarithmize = get_bool(s, "input.logarithmize");
    if (s.exists("input.skipProjections")) c.skip_projections = parse_skip_projections(s.string("input.skipProjections"));

    c.output_filename = s.path("output.filename");
    if (s.exists("output.format")) c.output_format = "." + s.string("output.format");
    if (s.exists("output.options.voxelType"))
        c.output_voxel_type = parse_voxel_value_type(s.string("output.options.voxelType"));
    if (s.exists("output.options.orientation"))
        c.output_orientation = parse_coordinate_order(s.string("output.options.orientation"));
    if (s.exists("output.options.invert")) c.output_invert = get_bool(s, "output.options.invert");
    if (s.exists("output.options.enableVoxelization") && get_bool(s, "output.options.enableVoxelization"))
        unsupported("output.options.enableVoxelization is not available");

    if (s.exists("priorknowledge.maskVolumeFilename")) c.mask_file = s.path("priorknowledge.maskVolumeFilename");

    if (s.exists("debug.profilingLevel")) {
        c.profiling_level = s.string("debug.profilingLevel");
        static const char* levels[] = {"none", "overall", "normal", "detail", "timing"};
        if (std::find_if(std::begin(levels), std::end(levels),
                         [&](const char* l) { return c.profiling_level == l; }) == std::end(levels))
            fail("Illegal value for parameter debug.profilingLevel: " + c.profiling_level);
    }
    if (s.exists("debug.writeIntermediateVolumes"))
        c.write_intermediate_volumes = get_bool(s, "debug.writeIntermediateVolumes");
    if (s.exists("debug.writeIntermediateVolumesAfterIterations"))
        c.write_intermediate_volumes_after_iterations = get_bool(s, "debug.writeIntermediateVolumesAfterIterations");
    if (s.exists("debug.infoPath")) c.debug_info_path = s.path("debug.infoPath");
    return c;
}

namespace {

std::string json_string(const std::string& s)
{
    std::string out = "\"";
    for (char ch : s) {
        if (ch == '"' || ch == '\\') {
            out += '\\';
            out += ch;
        } else if ((unsigned char)ch < 0x20) {
            char b[8];
            std::snprintf(b, sizeof b, "\\u%04x", (unsigned)(unsigned char)ch);
            out += b;
        } else out += ch;
    }
    return out + "\"";
}

std::string json_float(float f)
{
    char b[32];
    std::snprintf(b, sizeof b, "%.9g", f);
    return b;
}

}  // namespace

std::string ReconstructionConfig::to_json() const
{
    std::ostringstream o;
    auto v3f = [&](const float* v) { return "[" + json_float(v[0]) + ", " + json_float(v[1]) + ", " + json_float(v[2]) + "]"; };
    o << "{";
    o << "\"algorithm\": " << json_string(algorithm) << ", \"block_size\": " << block_size
      << ", \"iterations\": " << iterations << ", \"lambda\": " << json_float(lambda)
      << ", \"use_long_object_compensation\": " << (use_long_object_compensation ? "true" : "false")
      << ", \"fp_samples\": " << fp_samples << ", \"bp_samples\": " << bp_samples
      << ", \"voxel_representation\": " << json_string(voxel_representation)
      << ", \"projection_iteration\": " << json_string(projection_iteration)
      << ", \"subvolume_count\": " << subvolume_count << ", \"device_id\": " << device_id
      << ", \"dimension\": " << v3f(dimension) << ", \"origin\": " << v3f(origin) << ", \"resolution\": ["
      << resolution[0] << ", " << resolution[1] << ", " << resolution[2] << "]"
      << ", \"init_value\": " << json_float(init_value) << ", \"init_file\": " << json_string(init_file)
      << ", \"init_file_orientation\": " << json_string(coordinate_order_name(init_file_orientation))
      << ", \"projections\": " << json_string(projections) << ", \"tilt_angles\": " << json_string(tilt_angles)
      << ", \"x_axis_tilt\": " << json_float(x_axis_tilt)
      << ", \"logarithmize\": " << (logarithmize ? "true" : "false") << ", \"skip_projections\": [";
    for (size_t i = 0; i < skip_projections.size(); ++i) o << (i ? ", " : "") << skip_projections[i];
    o << "], \"output_filename\": " << json_string(output_filename)
      << ", \"output_format\": " << json_string(output_format)
      << ", \"output_voxel_type\": " << json_string(voxel_value_type_name(output_voxel_type))
      << ", \"output_orientation\": " << json_string(coordinate_order_name(output_orientation))
      << ", \"output_invert\": " << (output_invert ? "true" : "false")
      << ", \"mask_file\": " << json_string(mask_file) << ", \"profiling_level\": " << json_string(profiling_level)
      << ", \"write_intermediate_volumes\": " << (write_intermediate_volumes ? "true" : "false")
      << ", \"write_intermediate_volumes_after_iterations\": "
      << (write_intermediate_volumes_after_iterations ? "true" : "false")
      << ", \"debug_info_path\": " << json_string(debug_info_path) << "}";
    return o.str();
}

}  // namespace etn
