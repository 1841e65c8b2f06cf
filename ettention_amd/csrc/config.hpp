// Parameter sources and the typed parameter sets of the block-iterative
// reconstruction, with the reference's keys, defaults and precedence
// (ettention/libReconstruction/src/setup/parametersource/,
// src/setup/parameterset/).
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "io.hpp"
#include "xml.hpp"

namespace etn {

// Thrown when a required key is in no source (ParameterNotFoundException).
struct ParameterNotFound : std::runtime_error {
    std::string key;
    explicit ParameterNotFound(const std::string& k)
        : std::runtime_error("Missing required parameter " + k), key(k) {}
};

class ParameterSource {
public:
    virtual ~ParameterSource() = default;
    virtual bool exists(const std::string& key) const = 0;
    // The raw value text; throws ParameterNotFound when absent.
    virtual std::string string(const std::string& key) const = 0;
    // A file path: '\' becomes '/', normalised; sources may resolve relative paths.
    virtual std::string path(const std::string& key) const;
};

// XMLParameterSource (XMLParameterSource.cpp): key "a.b.c" names the element
// path <a><b><c value="..."/></b></a>; the first element of a name wins.
// Relative paths are taken against the XML file's directory.
class XmlParameterSource : public ParameterSource {
public:
    explicit XmlParameterSource(const std::string& filename);
    bool exists(const std::string& key) const override;
    std::string string(const std::string& key) const override;
    std::string path(const std::string& key) const override;
    const std::string& filename() const { return filename_; }

private:
    const XmlNode* node(const std::string& key) const;
    std::string filename_;
    std::vector<std::unique_ptr<XmlNode>> roots_;
};

// Key/value pairs set by code or parsed from a command line.
class ManualParameterSource : public ParameterSource {
public:
    void set(const std::string& key, const std::string& value) { values_[key] = value; }
    bool exists(const std::string& key) const override { return values_.count(key) != 0; }
    std::string string(const std::string& key) const override;
    const std::map<std::string, std::string>& values() const { return values_; }

private:
    std::map<std::string, std::string> values_;
};

// CommandLineToolParameterSource (frontend/src/setup/CommandLineToolParameterSource.cpp):
// --config / -c <file>, --help / -h, --parameters / -H, --devices / -d,
// --explicitDeviceId / -e <id> (= hardware.openclDeviceId), and any
// --<key> <value> or --<key>=<value>.
class CommandLineParameterSource : public ManualParameterSource {
public:
    enum class Mode { Default, Reconstruct, Help, HelpAndParameters, Devices };
    CommandLineParameterSource(int argc, const char* const* argv);
    Mode mode() const { return mode_; }

private:
    Mode mode_ = Mode::Default;
};

// CascadingParameterSource: the first source holding a key answers it;
// add_source puts a source in front (it then wins over the earlier ones).
class CascadingParameterSource : public ParameterSource {
public:
    void add_source(std::shared_ptr<ParameterSource> s) { sources_.insert(sources_.begin(), std::move(s)); }
    bool exists(const std::string& key) const override;
    std::string string(const std::string& key) const override;
    std::string path(const std::string& key) const override;

private:
    std::vector<std::shared_ptr<ParameterSource>> sources_;
};

// Typed reads with the reference's text formats (boost lexical reads):
// numbers, booleans true/false/1/0, vectors "x y z".
float get_float(const ParameterSource& s, const std::string& key);
uint32_t get_uint(const ParameterSource& s, const std::string& key);
int32_t get_int(const ParameterSource& s, const std::string& key);
bool get_bool(const ParameterSource& s, const std::string& key);
void get_vec3f(const ParameterSource& s, const std::string& key, float out[3]);
void get_vec3ui(const ParameterSource& s, const std::string& key, uint32_t out[3]);

// The parameter sets the block-iterative reconstruction reads, resolved.
struct ReconstructionConfig {
    // AlgorithmParameterSet (AlgorithmParameterSet.cpp:8-53)
    std::string algorithm;
    uint32_t block_size = 1;             // 0xffffffff: the whole stack (SIRT)
    uint32_t iterations = 0;
    float lambda = 0.0f;
    bool use_long_object_compensation = false;
    // ForwardProjectionParameterSet / BackprojectionParameterSet
    uint32_t fp_samples = 1, bp_samples = 1;
    // OptimizationParameterSet (OptimizationParameterSet.cpp:7-22)
    std::string voxel_representation = "float";   // float | half | bfloat16 (extension)
    std::string projection_iteration = "random";
    // HardwareParameterSet
    uint32_t subvolume_count = 0;        // 0: automatic
    int32_t device_id = -1;              // hardware.openclDeviceId as a HIP device ordinal
    // VolumeParameterSet (VolumeParameterSet.cpp:9-38)
    float dimension[3] = {0, 0, 0};
    float origin[3] = {0, 0, 0};
    uint32_t resolution[3] = {0, 0, 0};
    float init_value = 0.0f;
    std::string init_file;
    CoordinateOrder init_file_orientation = CoordinateOrder::XZY;
    // InputParameterSet (InputParameterSet.cpp:7-33)
    std::string projections;
    std::string tilt_angles;
    float x_axis_tilt = 0.0f;
    bool logarithmize = false;
    std::vector<uint32_t> skip_projections;   // 0-based
    // OutputParameterSet (OutputParameterSet.cpp:19-52)
    std::string output_filename;
    std::string output_format;           // ".mrc" when output.format is set
    VoxelValueType output_voxel_type = VoxelValueType::Float32;
    CoordinateOrder output_orientation = CoordinateOrder::XZY;
    bool output_invert = false;
    // PriorKnowledgeParameterSet
    std::string mask_file;
    // DebugParameterSet (DebugParameterSet.cpp:59-110)
    std::string profiling_level = "overall";
    bool write_intermediate_volumes = false;               // after every block
    bool write_intermediate_volumes_after_iterations = false;
    std::string debug_info_path = "debug";

    static ReconstructionConfig from(const ParameterSource& s);
    std::string to_json() const;
};

// InputParameterSet::parseSkipProjectionList: "none" or a comma list of
// 1-based indices and ranges "a-b"; returns 0-based indices.
std::vector<uint32_t> parse_skip_projections(const std::string& s);

}  // namespace etn
