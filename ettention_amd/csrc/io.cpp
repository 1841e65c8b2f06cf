// File formats of the reconstruction (see io.hpp).  Every conversion follows
// the reference's arithmetic in float, in the reference's order, so that a
// file read here holds the same values the reference would compute from it.
#include "io.hpp"

#include <sys/stat.h>
#include <zlib.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <memory>
#include <sstream>

#include "common.hpp"
#include "xml.hpp"

namespace etn {

// ---------------------------------------------------------------------------
// enumerations
// ---------------------------------------------------------------------------

VoxelValueType parse_voxel_value_type(const std::string& s)
{
    if (s == "unsigned8") return VoxelValueType::Unsigned8;
    if (s == "unsigned16") return VoxelValueType::Unsigned16;
    if (s == "float32") return VoxelValueType::Float32;
    if (s == "signed16") return VoxelValueType::Signed16;
    fail("ParseVoxelValueType: Unknown file mode " + s);
}

CoordinateOrder parse_coordinate_order(const std::string& s)
{
    if (s == "xzy_order") return CoordinateOrder::XZY;
    if (s == "xyz_order") return CoordinateOrder::XYZ;
    if (s == "yxz_order") return CoordinateOrder::YXZ;
    if (s == "yzx_order") return CoordinateOrder::YZX;
    if (s == "zxy_order") return CoordinateOrder::ZXY;
    if (s == "zyx_order") return CoordinateOrder::ZYX;
    fail("Unknown coordinate order: " + s + "!");
}

const char* voxel_value_type_name(VoxelValueType t)
{
    switch (t) {
        case VoxelValueType::Unsigned8: return "unsigned8";
        case VoxelValueType::Signed16: return "signed16";
        case VoxelValueType::Float32: return "float32";
        case VoxelValueType::Unsigned16: return "unsigned16";
    }
    return "?";
}

const char* coordinate_order_name(CoordinateOrder o)
{
    switch (o) {
        case CoordinateOrder::XYZ: return "xyz_order";
        case CoordinateOrder::XZY: return "xzy_order";
        case CoordinateOrder::YXZ: return "yxz_order";
        case CoordinateOrder::YZX: return "yzx_order";
        case CoordinateOrder::ZXY: return "zxy_order";
        case CoordinateOrder::ZYX: return "zyx_order";
    }
    return "?";
}

namespace {

// For a coordinate order: which volume axis (0 x, 1 y, 2 z) runs along the
// file's sections, rows and columns (IOEnumerations.cpp:79-126).
struct AxisMap {
    int section, row, column;
};

AxisMap axes_of(CoordinateOrder o)
{
    switch (o) {
        case CoordinateOrder::XYZ: return {2, 1, 0};
        case CoordinateOrder::XZY: return {1, 2, 0};
        case CoordinateOrder::YXZ: return {2, 0, 1};
        case CoordinateOrder::YZX: return {0, 2, 1};
        case CoordinateOrder::ZXY: return {1, 0, 2};
        case CoordinateOrder::ZYX: return {0, 1, 2};
    }
    fail("unknown coordinate order");
}

std::vector<char> read_file(const std::string& path, const std::string& what)
{
    std::ifstream f(path, std::ios::binary);
    if (!f.good()) fail("Could not open " + what + " " + path);
    f.seekg(0, std::ios::end);
    std::streamoff n = f.tellg();
    f.seekg(0, std::ios::beg);
    std::vector<char> buf((size_t)std::max<std::streamoff>(n, 0));
    if (n > 0) f.read(buf.data(), n);
    if (!f.good()) fail("Error reading " + what + " " + path);
    return buf;
}

}  // namespace

// ---------------------------------------------------------------------------
// half / bfloat16 host conversions
// ---------------------------------------------------------------------------

uint16_t float_to_half_truncate(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    const uint16_t sign = (uint16_t)((u >> 16) & 0x8000u);
    const uint32_t a = u & 0x7fffffffu;
    if (a > 0x7f800000u) return (uint16_t)(sign | 0x7c01u);          // NaN
    if (a > 0x477fe000u) return (uint16_t)(sign | 0x7c00u);          // above 65504: infinity
    if (a < 0x38800000u) {                                           // below the smallest normal half
        float m;
        std::memcpy(&m, &a, 4);
        return (uint16_t)(sign | (uint32_t)(m * 16777216.0f));       // subnormal, truncated
    }
    return (uint16_t)(sign | (((a >> 13) - (112u << 10)) & 0x7fffu));   // rebias exponent, drop 13 bits
}

float half_to_float(uint16_t h)
{
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    uint32_t u;
    if (e == 0) {
        float v = (float)m * 5.9604644775390625e-08f;   // m * 2^-24
        std::memcpy(&u, &v, 4);
        u |= sign;
    } else if (e == 31) {
        u = sign | 0x7f800000u | (m << 13);
    } else {
        u = sign | ((e + 112u) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

uint16_t float_to_bfloat16_rne(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)(((u >> 16) & 0x8000u) | 0x7fc0u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// ---------------------------------------------------------------------------
// paths
// ---------------------------------------------------------------------------

std::string normalize_path(const std::string& in)
{
    std::string p = in;
    std::replace(p.begin(), p.end(), '\\', '/');
    const bool absolute = !p.empty() && p[0] == '/';
    std::vector<std::string> parts;
    std::stringstream ss(p);
    std::string item;
    while (std::getline(ss, item, '/')) {
        if (item.empty() || item == ".") continue;
        if (item == ".." && !parts.empty() && parts.back() != "..") parts.pop_back();
        else if (item == ".." && absolute) continue;
        else parts.push_back(item);
    }
    std::string out = absolute ? "/" : "";
    for (size_t i = 0; i < parts.size(); ++i) out += (i ? "/" : "") + parts[i];
    if (out.empty()) out = ".";
    return out;
}

std::string join_path(const std::string& base_dir, const std::string& p)
{
    std::string q = p;
    std::replace(q.begin(), q.end(), '\\', '/');
    if (!q.empty() && q[0] == '/') return normalize_path(q);
    if (base_dir.empty()) return normalize_path(q);
    return normalize_path(base_dir + "/" + q);
}

std::string parent_path(const std::string& p)
{
    std::string q = p;
    std::replace(q.begin(), q.end(), '\\', '/');
    size_t k = q.find_last_of('/');
    if (k == std::string::npos) return "";
    if (k == 0) return "/";
    return q.substr(0, k);
}

bool is_directory(const std::string& p)
{
    struct stat st;
    return ::stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

bool is_regular_file(const std::string& p)
{
    struct stat st;
    return ::stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

std::string extension(const std::string& p)
{
    size_t slash = p.find_last_of("/\\");
    std::string name = slash == std::string::npos ? p : p.substr(slash + 1);
    size_t dot = name.find_last_of('.');
    if (dot == std::string::npos || dot == 0) return "";
    return name.substr(dot);
}

// ---------------------------------------------------------------------------
// tilt angles
// ---------------------------------------------------------------------------

std::vector<float> read_tilt_angles(const std::string& path, uint32_t expected)
{
    std::ifstream f(path);
    if (!f.good()) fail("Could not open file: " + path);
    std::vector<float> out;
    std::string tok;
    while (f >> tok) {
        char* end = nullptr;
        errno = 0;
        float a = std::strtof(tok.c_str(), &end);
        if (end == tok.c_str() || *end != '\0') fail("Malformed tilt angle '" + tok + "' in " + path);
        out.push_back(a);
    }
    if (out.size() != expected)
        fail("Number of projections DIFFERS from number of Tilt Angles given in " + path + " (" +
             std::to_string(out.size()) + " angles, " + std::to_string(expected) + " projections)");
    return out;
}

// ---------------------------------------------------------------------------
// MRC
// ---------------------------------------------------------------------------

namespace {

struct MrcFile {
    MrcHeader h{};
    std::vector<char> bytes;
    size_t data_offset = 0;
    size_t value_bytes = 0;
    size_t count = 0;   // nx * ny * nz
};

size_t mode_bytes(int32_t mode)
{
    switch (mode) {
        case 0: return 1;
        case 1: return 2;
        case 2: return 4;
        case 6: return 2;
    }
    return 0;
}

MrcFile open_mrc(const std::string& path)
{
    MrcFile m;
    m.bytes = read_file(path, "MRC stack");
    if (m.bytes.size() < sizeof(MrcHeader)) fail("Could not open MRC stack " + path + " (shorter than its header)");
    std::memcpy(&m.h, m.bytes.data(), sizeof(MrcHeader));
    m.value_bytes = mode_bytes(m.h.mode);
    if (!m.value_bytes)
        unsupported("Format of the MRC stack is either unknown or not supported (mode " + std::to_string(m.h.mode) +
                    ")");
    if (m.h.nx <= 0 || m.h.ny <= 0 || m.h.nz <= 0) fail("MRC stack " + path + " has a non-positive size");
    if (m.h.next < 0) fail("MRC stack " + path + " has a negative extended header size");
    m.data_offset = sizeof(MrcHeader) + (size_t)m.h.next;
    m.count = (size_t)m.h.nx * (size_t)m.h.ny * (size_t)m.h.nz;
    if (m.bytes.size() < m.data_offset + m.count * m.value_bytes)
        fail("Error reading " + std::to_string(m.count) + " voxel value(s) from " + path + "!");
    return m;
}

template <typename T>
float raw_value(const MrcFile& m, size_t i)
{
    T v;
    std::memcpy(&v, m.bytes.data() + m.data_offset + i * sizeof(T), sizeof(T));
    return (float)v;
}

float mrc_value(const MrcFile& m, size_t i)
{
    switch (m.h.mode) {
        case 0: return raw_value<uint8_t>(m, i);
        case 1: return raw_value<int16_t>(m, i);
        case 2: return raw_value<float>(m, i);
        case 6: return raw_value<uint16_t>(m, i);
    }
    return 0.0f;
}

// MRCStack::readImageData / convertValue over the whole stack, file order.
std::vector<float> mrc_values(const MrcFile& m, bool logarithmize)
{
    std::vector<float> out(m.count);
    if (!logarithmize) {
        for (size_t i = 0; i < m.count; ++i) out[i] = mrc_value(m, i);
        return out;
    }
    // MRCStack::getDataRange and the adjustments of MRCStack::init
    float dmax = -std::numeric_limits<float>::infinity();
    float dmin = std::numeric_limits<float>::infinity();
    for (size_t i = 0; i < m.count; ++i) {
        float v = mrc_value(m, i);
        dmax = std::max(v, dmax);
        dmin = std::min(v, dmin);
    }
    float min_value;
    if (m.h.mode == 1 || m.h.mode == 2) {
        dmax = dmax - dmin + 1.0f;
        min_value = dmin - 1.0f;
    } else {
        dmax = dmax + 1.0f;
        min_value = -1.0f;
    }
    for (size_t i = 0; i < m.count; ++i) out[i] = logf(dmax / (mrc_value(m, i) - min_value));
    return out;
}

}  // namespace

ProjectionStack read_mrc_stack(const std::string& path, const std::string& tilt_path, bool logarithmize)
{
    MrcFile m = open_mrc(path);
    ProjectionStack s;
    s.kind = "MRCStack";
    s.nu = (uint32_t)m.h.nx;
    s.nv = (uint32_t)m.h.ny;
    s.count = (uint32_t)m.h.nz;
    s.data = mrc_values(m, logarithmize);
    if (tilt_path.empty()) s.tilt_angles.assign(s.count, 0.0f);
    else s.tilt_angles = read_tilt_angles(tilt_path, s.count);
    const float nx = (float)m.h.nx, ny = (float)m.h.ny;
    const float src[3] = {nx * -0.5f, ny * -0.5f, nx * 1.0f};
    const float det[3] = {nx * -0.5f, ny * -0.5f, nx * -1.0f};
    for (int k = 0; k < 3; ++k) {
        s.base.source[k] = src[k];
        s.base.detector[k] = det[k];
        s.base.horizontal_pitch[k] = k == 0 ? 1.0f : 0.0f;
        s.base.vertical_pitch[k] = k == 1 ? 1.0f : 0.0f;
    }
    s.base.resolution[0] = s.nu;
    s.base.resolution[1] = s.nv;
    return s;
}

std::vector<float> read_mrc_volume(const std::string& path, CoordinateOrder order, uint32_t res[3])
{
    MrcFile m = open_mrc(path);
    std::vector<float> file = mrc_values(m, false);
    const uint32_t fres[3] = {(uint32_t)m.h.nx, (uint32_t)m.h.ny, (uint32_t)m.h.nz};
    // the file's column / row / section index is the volume coordinate named by the order
    const AxisMap a = axes_of(order);
    res[a.column] = fres[0];
    res[a.row] = fres[1];
    res[a.section] = fres[2];
    std::vector<float> out(m.count);
    uint32_t c[3];
    for (uint32_t fz = 0; fz < fres[2]; ++fz)
        for (uint32_t fy = 0; fy < fres[1]; ++fy)
            for (uint32_t fx = 0; fx < fres[0]; ++fx) {
                c[a.section] = fz;
                c[a.row] = fy;
                c[a.column] = fx;
                out[((size_t)c[2] * res[1] + c[1]) * res[0] + c[0]] =
                    file[((size_t)fz * fres[1] + fy) * fres[0] + fx];
            }
    return out;
}

namespace {

// VolumeStatistics::compute (VolumeStatistics.cpp:18-53): NaN voxels skipped,
// max starts at the smallest positive normal float, the mean is a float
// running sum scaled by 1/(float)(all voxels).
struct Stats {
    float min, max, mean;
};

Stats volume_statistics(const float* v, size_t n)
{
    float mn = std::numeric_limits<float>::max();
    float mx = std::numeric_limits<float>::min();
    float mean = 0.0f;
    for (size_t i = 0; i < n; ++i) {
        const float x = v[i];
        if (std::isnan(x)) continue;
        if (x < mn) mn = x;
        if (x > mx) mx = x;
        mean += x;
    }
    mean *= (1 / (float)n);
    return {mn, mx, mean};
}

// LinearRangeTransformation::transformRange (LinearRangeTransformation.cpp:14-21)
struct LinearRange {
    Stats s;
    float out_min, out_max;
    float operator()(float value) const
    {
        const float zero_to_max = value - s.min;
        const float zero_to_one = zero_to_max / (s.max - s.min);
        const float zero_to_new_max = zero_to_one * (out_max - out_min);
        return zero_to_new_max + out_min;
    }
};

// static_cast of an out-of-range or NaN float is undefined in C++; saturate instead
template <typename T>
T cast_saturating(float v)
{
    if (std::isnan(v)) return (T)0;
    if (v <= (float)std::numeric_limits<T>::lowest()) return std::numeric_limits<T>::lowest();
    if (v >= (float)std::numeric_limits<T>::max()) return std::numeric_limits<T>::max();
    return (T)v;
}

}  // namespace

void write_mrc_volume(const std::string& path, const float* volume, const uint32_t res[3], VoxelValueType type,
                      CoordinateOrder order, bool invert)
{
    const size_t n = (size_t)res[0] * res[1] * res[2];
    if (n == 0) fail("cannot write an empty volume");
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    if (!out.good()) fail("Failed to open output file " + path + " for writing!");

    const Stats st = volume_statistics(volume, n);
    // VolumeFileFormatWriter::createRangeTransformation (VolumeFileFormatWriter.cpp:112-144)
    bool identity = false;
    LinearRange tr{st, 0.0f, 0.0f};
    if (invert) {
        switch (type) {
            case VoxelValueType::Unsigned8: tr.out_min = 255.0f, tr.out_max = 0.0f; break;
            case VoxelValueType::Unsigned16: tr.out_min = 65535.0f, tr.out_max = 0.0f; break;
            case VoxelValueType::Signed16: tr.out_min = 32767.0f, tr.out_max = -32768.0f; break;
            case VoxelValueType::Float32: tr.out_min = st.max, tr.out_max = st.min; break;
        }
    } else {
        switch (type) {
            case VoxelValueType::Unsigned8: tr.out_min = 0.0f, tr.out_max = 255.0f; break;
            case VoxelValueType::Unsigned16: tr.out_min = 0.0f, tr.out_max = 65535.0f; break;
            case VoxelValueType::Signed16: tr.out_min = -32768.0f, tr.out_max = 32767.0f; break;
            case VoxelValueType::Float32: identity = true; break;
        }
    }
    auto transform = [&](float v) { return identity ? v : tr(v); };

    const AxisMap a = axes_of(order);
    const uint32_t ores[3] = {res[a.column], res[a.row], res[a.section]};

    MrcHeader h;
    std::memset(&h, 0, sizeof h);
    h.mode = (int32_t)type;
    h.nx = h.mx = (int32_t)ores[0];
    h.ny = h.my = (int32_t)ores[1];
    h.nz = h.mz = (int32_t)ores[2];
    h.cell_dim[0] = (float)ores[0];
    h.cell_dim[1] = (float)ores[1];
    h.cell_dim[2] = (float)ores[2];
    std::memcpy(h.map, "MAP ", 4);
    h.mapc = 1;
    h.mapr = 2;
    h.maps = 3;
    h.dmin = transform(st.min);
    h.dmax = transform(st.max);
    h.dmean = transform(st.mean);
    out.write(reinterpret_cast<const char*>(&h), sizeof h);

    // Float32 files are written slice by slice from the volume itself, so an
    // inverted float32 file carries inverted statistics over unchanged data
    // (VolumeFileFormatWriter.cpp:24-29, writeSlicewise); the other types write
    // the range-transformed values (writePixelwise).
    const size_t vb = type == VoxelValueType::Unsigned8 ? 1 : type == VoxelValueType::Float32 ? 4 : 2;
    std::vector<char> row((size_t)ores[0] * vb);
    uint32_t c[3];
    for (uint32_t i = 0; i < ores[2]; ++i) {
        for (uint32_t j = 0; j < ores[1]; ++j) {
            c[a.section] = i;
            c[a.row] = j;
            for (uint32_t k = 0; k < ores[0]; ++k) {
                c[a.column] = k;
                const float v = volume[((size_t)c[2] * res[1] + c[1]) * res[0] + c[0]];
                char* dst = row.data() + (size_t)k * vb;
                switch (type) {
                    case VoxelValueType::Float32: std::memcpy(dst, &v, 4); break;
                    case VoxelValueType::Signed16: {
                        int16_t s = cast_saturating<int16_t>(tr(v));
                        std::memcpy(dst, &s, 2);
                        break;
                    }
                    case VoxelValueType::Unsigned16: {
                        uint16_t s = cast_saturating<uint16_t>(tr(v));
                        std::memcpy(dst, &s, 2);
                        break;
                    }
                    case VoxelValueType::Unsigned8: {
                        uint8_t s = cast_saturating<uint8_t>(tr(v));
                        std::memcpy(dst, &s, 1);
                        break;
                    }
                }
            }
            out.write(row.data(), (std::streamsize)row.size());
        }
    }
    if (!out.good()) fail("Error encountered while writing to output file " + path + " !");
}

// ---------------------------------------------------------------------------
// TIFF
// ---------------------------------------------------------------------------

namespace {

class Tiff {
public:
    explicit Tiff(const std::string& path) : path_(path), b_(read_file(path, "image"))
    {
        if (b_.size() < 8) bad("shorter than a TIFF header");
        if (b_[0] == 'I' && b_[1] == 'I') le_ = true;
        else if (b_[0] == 'M' && b_[1] == 'M') le_ = false;
        else bad("not a TIFF file");
        const uint16_t magic = u16(2);
        if (magic == 43) unsupported("BigTIFF is not supported: " + path);
        if (magic != 42) bad("not a TIFF file");
        uint32_t off = u32(4);
        while (off) {
            if (ifds_.size() > 65536) bad("IFD loop");
            need(off, 2);
            ifds_.push_back(off);
            const uint16_t n = u16(off);
            need(off + 2, (size_t)n * 12 + 4);
            off = u32(off + 2 + (size_t)n * 12);
        }
        if (ifds_.empty()) bad("no image in file");
    }

    uint32_t pages() const { return (uint32_t)ifds_.size(); }

    std::vector<float> read(uint32_t page, uint32_t* width, uint32_t* height)
    {
        if (page >= ifds_.size()) fail("index out of bound: image stack does not have that many images: " + path_);
        load_tags(ifds_[page]);
        const uint32_t w = scalar(256), h = scalar(257);
        const uint32_t bits = scalar(258, 1), comp = scalar(259, 1), spp = scalar(277, 1);
        const uint32_t fmt = scalar(339, 1), pred = scalar(317, 1), planar = scalar(284, 1);
        if (w == 0 || h == 0) bad("empty image");
        if (spp != 1 || planar != 1) unsupported("only single-channel TIFF images are supported: " + path_);
        const bool is_float = fmt == 3 && bits == 32;
        const bool is_u8 = fmt == 1 && bits == 8;
        const bool is_u16 = fmt == 1 && bits == 16;
        if (!is_float && !is_u8 && !is_u16)
            unsupported("TIFF sample format (" + std::to_string(bits) + " bits, format " + std::to_string(fmt) +
                        ") cannot be converted to float (float32, uint8 and uint16 can): " + path_);
        const size_t bps = bits / 8;

        const bool tiled = tags_.count(322) != 0;
        const uint32_t bw = tiled ? scalar(322) : w;
        const uint32_t bh = tiled ? scalar(323) : std::min(scalar(278, h), h);
        if (bw == 0 || bh == 0) bad("zero tile or strip size");
        const std::vector<uint32_t> offs = array(tiled ? 324 : 273);
        const std::vector<uint32_t> lens = array(tiled ? 325 : 279);
        const uint32_t across = tiled ? (w + bw - 1) / bw : 1;
        const uint32_t down = (h + bh - 1) / bh;
        if (offs.size() < (size_t)across * down || lens.size() < offs.size()) bad("missing strips or tiles");

        std::vector<float> img((size_t)w * h);
        std::vector<uint8_t> block((size_t)bw * bh * bps);
        for (uint32_t ty = 0; ty < down; ++ty) {
            for (uint32_t tx = 0; tx < across; ++tx) {
                const size_t idx = (size_t)ty * across + tx;
                const uint32_t rows = tiled ? bh : std::min(bh, h - ty * bh);
                const size_t want = (size_t)bw * rows * bps;
                need(offs[idx], lens[idx]);
                decompress(comp, reinterpret_cast<const uint8_t*>(b_.data()) + offs[idx], lens[idx], block.data(),
                           want);
                // libtiff applies the predictor to LZW and Deflate data only
                const bool predicted = comp == 5 || comp == 8 || comp == 32946;
                if (predicted) undo_predictor(pred, block.data(), bw, rows, bps, is_float);
                for (uint32_t r = 0; r < rows; ++r) {
                    const uint32_t y = ty * bh + r;
                    if (y >= h) break;
                    for (uint32_t c = 0; c < bw; ++c) {
                        const uint32_t x = tx * bw + c;
                        if (x >= w) break;
                        const uint8_t* p = block.data() + ((size_t)r * bw + c) * bps;
                        img[(size_t)y * w + x] = sample(p, bps, is_float, predicted && pred == 3);
                    }
                }
            }
        }
        *width = w;
        *height = h;
        return img;
    }

private:
    [[noreturn]] void bad(const std::string& why) const { fail("unable to load image " + path_ + ": " + why); }
    void need(size_t off, size_t n) const
    {
        if (off > b_.size() || n > b_.size() - off) bad("truncated file");
    }
    uint16_t u16(size_t o) const
    {
        need(o, 2);
        const uint8_t* p = reinterpret_cast<const uint8_t*>(b_.data()) + o;
        return le_ ? (uint16_t)(p[0] | (p[1] << 8)) : (uint16_t)((p[0] << 8) | p[1]);
    }
    uint32_t u32(size_t o) const
    {
        need(o, 4);
        const uint8_t* p = reinterpret_cast<const uint8_t*>(b_.data()) + o;
        return le_ ? (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24)
                   : ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | (uint32_t)p[3];
    }

    struct Entry {
        uint16_t type;
        uint32_t count;
        size_t value_offset;   // where the values are
    };

    void load_tags(uint32_t ifd)
    {
        tags_.clear();
        const uint16_t n = u16(ifd);
        for (uint16_t i = 0; i < n; ++i) {
            const size_t e = ifd + 2 + (size_t)i * 12;
            Entry en{u16(e + 2), u32(e + 4), 0};
            const size_t size = en.type == 3 ? 2 : (en.type == 4 || en.type == 9 || en.type == 11) ? 4
                                                    : (en.type == 5 || en.type == 10 || en.type == 12) ? 8 : 1;
            en.value_offset = (size * en.count <= 4) ? e + 8 : u32(e + 8);
            tags_[u16(e)] = en;
        }
    }

    uint32_t value(const Entry& e, uint32_t i) const
    {
        switch (e.type) {
            case 1: need(e.value_offset + i, 1); return (uint8_t)b_[e.value_offset + i];
            case 3: return u16(e.value_offset + (size_t)i * 2);
            case 4: return u32(e.value_offset + (size_t)i * 4);
        }
        bad("unexpected tag type " + std::to_string(e.type));
    }

    uint32_t scalar(uint16_t tag, uint32_t def) const
    {
        auto it = tags_.find(tag);
        if (it == tags_.end() || it->second.count == 0) return def;
        return value(it->second, 0);
    }
    uint32_t scalar(uint16_t tag) const
    {
        auto it = tags_.find(tag);
        if (it == tags_.end() || it->second.count == 0) bad("missing tag " + std::to_string(tag));
        return value(it->second, 0);
    }
    std::vector<uint32_t> array(uint16_t tag) const
    {
        auto it = tags_.find(tag);
        if (it == tags_.end()) bad("missing tag " + std::to_string(tag));
        std::vector<uint32_t> v(it->second.count);
        for (uint32_t i = 0; i < it->second.count; ++i) v[i] = value(it->second, i);
        return v;
    }

    void decompress(uint32_t comp, const uint8_t* src, size_t n, uint8_t* dst, size_t want) const
    {
        std::fill(dst, dst + want, 0);
        switch (comp) {
            case 1:
                std::memcpy(dst, src, std::min(n, want));
                return;
            case 5:
                lzw(src, n, dst, want);
                return;
            case 32773:
                packbits(src, n, dst, want);
                return;
            case 8:
            case 32946: {
                uLongf out = (uLongf)want;
                std::vector<uint8_t> tmp(want);
                int rc = ::uncompress(tmp.data(), &out, src, (uLong)n);
                if (rc != Z_OK && rc != Z_BUF_ERROR) bad("corrupt Deflate data");
                std::memcpy(dst, tmp.data(), out);
                return;
            }
        }
        unsupported("TIFF compression " + std::to_string(comp) + " is not supported: " + path_);
    }

    // TIFF LZW: MSB-first codes of 9..12 bits, Clear 256, EndOfInformation
    // 257, the code width grows one code early (TIFF 6.0, section 13).
    void lzw(const uint8_t* src, size_t n, uint8_t* dst, size_t want) const
    {
        std::vector<uint16_t> prefix(4096);
        std::vector<uint8_t> suffix(4096), first(4096);
        std::vector<uint16_t> length(4096);
        for (int i = 0; i < 256; ++i) {
            prefix[i] = 0xffff;
            suffix[i] = first[i] = (uint8_t)i;
            length[i] = 1;
        }
        size_t bitpos = 0, o = 0;
        int width = 9, next = 258, prev = -1;
        const size_t total_bits = n * 8;
        auto emit = [&](int code) {
            const size_t len = length[code];
            size_t end = o + len;
            int c = code;
            for (size_t k = end; k > o; --k) {
                if (k - 1 < want) dst[k - 1] = suffix[c];
                c = prefix[c];
            }
            o = end;
        };
        while (bitpos + width <= total_bits && o < want) {
            uint32_t code = 0;
            for (int k = 0; k < width; ++k, ++bitpos)
                code = (code << 1) | ((src[bitpos >> 3] >> (7 - (bitpos & 7))) & 1u);
            if (code == 257) break;
            if (code == 256) {
                width = 9;
                next = 258;
                prev = -1;
                continue;
            }
            if (prev < 0) {
                if (code > 255) bad("corrupt LZW data");
                emit((int)code);
                prev = (int)code;
                continue;
            }
            if ((int)code > next || (int)code >= 4096) bad("corrupt LZW data");
            if (next < 4096) {   // a full table waits for the encoder's Clear code
                const uint8_t lead = (int)code < next ? first[code] : first[prev];
                prefix[next] = (uint16_t)prev;
                suffix[next] = lead;
                first[next] = first[prev];
                length[next] = (uint16_t)(length[prev] + 1);
                ++next;
            } else if ((int)code == next) {
                bad("corrupt LZW data");
            }
            emit((int)code);
            prev = (int)code;
            if (next >= (1 << width) - 1 && width < 12) ++width;
        }
    }

    void packbits(const uint8_t* src, size_t n, uint8_t* dst, size_t want) const
    {
        size_t i = 0, o = 0;
        while (i < n && o < want) {
            const int8_t c = (int8_t)src[i++];
            if (c >= 0) {
                const size_t k = std::min<size_t>((size_t)c + 1, std::min(n - i, want - o));
                std::memcpy(dst + o, src + i, k);
                i += (size_t)c + 1;
                o += k;
            } else if (c != -128) {
                if (i >= n) break;
                const size_t k = std::min<size_t>((size_t)(1 - c), want - o);
                std::memset(dst + o, src[i++], k);
                o += k;
            }
        }
    }

    void undo_predictor(uint32_t pred, uint8_t* p, uint32_t w, uint32_t rows, size_t bps, bool is_float) const
    {
        if (pred == 1) return;
        if (pred == 2) {
            if (is_float) unsupported("horizontal predictor on float samples: " + path_);
            for (uint32_t r = 0; r < rows; ++r) {
                uint8_t* row = p + (size_t)r * w * bps;
                if (bps == 1) {
                    for (uint32_t x = 1; x < w; ++x) row[x] = (uint8_t)(row[x] + row[x - 1]);
                } else {
                    uint16_t prev = read_int16(row);
                    for (uint32_t x = 1; x < w; ++x) {
                        uint16_t v = (uint16_t)(read_int16(row + 2 * x) + prev);
                        write_int16(row + 2 * x, v);
                        prev = v;
                    }
                }
            }
            return;
        }
        if (pred == 3) {
            // floating-point predictor: bytes differenced along the row, stored
            // most significant byte plane first
            std::vector<uint8_t> tmp((size_t)w * bps);
            for (uint32_t r = 0; r < rows; ++r) {
                uint8_t* row = p + (size_t)r * w * bps;
                for (size_t k = 1; k < (size_t)w * bps; ++k) row[k] = (uint8_t)(row[k] + row[k - 1]);
                for (uint32_t x = 0; x < w; ++x)
                    for (size_t b = 0; b < bps; ++b) tmp[(size_t)x * bps + b] = row[b * w + x];
                std::memcpy(row, tmp.data(), tmp.size());
            }
            return;
        }
        unsupported("TIFF predictor " + std::to_string(pred) + " is not supported: " + path_);
    }

    uint16_t read_int16(const uint8_t* q) const { return le_ ? (uint16_t)(q[0] | (q[1] << 8)) : (uint16_t)((q[0] << 8) | q[1]); }
    void write_int16(uint8_t* q, uint16_t v) const
    {
        if (le_) {
            q[0] = (uint8_t)v;
            q[1] = (uint8_t)(v >> 8);
        } else {
            q[0] = (uint8_t)(v >> 8);
            q[1] = (uint8_t)v;
        }
    }

    // one sample to float: FreeImage ConvertToFloat scales 8- and 16-bit grey to [0, 1]
    float sample(const uint8_t* q, size_t bps, bool is_float, bool big_endian_planes) const
    {
        if (is_float) {
            uint32_t u;
            if (big_endian_planes || !le_)
                u = ((uint32_t)q[0] << 24) | ((uint32_t)q[1] << 16) | ((uint32_t)q[2] << 8) | (uint32_t)q[3];
            else
                u = (uint32_t)q[0] | ((uint32_t)q[1] << 8) | ((uint32_t)q[2] << 16) | ((uint32_t)q[3] << 24);
            float f;
            std::memcpy(&f, &u, 4);
            return f;
        }
        if (bps == 1) return (float)q[0] / 255.0f;
        return (float)read_int16(q) / 65535.0f;
    }

    std::string path_;
    std::vector<char> b_;
    bool le_ = true;
    std::vector<uint32_t> ifds_;
    std::map<uint16_t, Entry> tags_;
};

}  // namespace

uint32_t tiff_page_count(const std::string& path) { return Tiff(path).pages(); }

std::vector<float> read_tiff(const std::string& path, uint32_t page, uint32_t* width, uint32_t* height)
{
    return Tiff(path).read(page, width, height);
}

void write_tiff(const std::string& path, const float* data, uint32_t width, uint32_t height)
{
    if (!width || !height) fail("cannot write an empty image");
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f.good()) fail("Failed to open image file " + path + " for writing!");
    const uint32_t nbytes = width * height * 4u;
    const uint16_t ntags = 11;
    const uint32_t ifd = 8;
    const uint32_t data_off = ifd + 2 + ntags * 12 + 4;
    auto put16 = [&](uint16_t v) { f.write(reinterpret_cast<const char*>(&v), 2); };
    auto put32 = [&](uint32_t v) { f.write(reinterpret_cast<const char*>(&v), 4); };
    auto tag = [&](uint16_t t, uint16_t type, uint32_t value) {
        put16(t);
        put16(type);
        put32(1);
        if (type == 3) {
            put16((uint16_t)value);
            put16(0);
        } else put32(value);
    };
    f.write("II", 2);
    put16(42);
    put32(ifd);
    put16(ntags);
    tag(256, 4, width);         // ImageWidth
    tag(257, 4, height);        // ImageLength
    tag(258, 3, 32);            // BitsPerSample
    tag(259, 3, 1);             // Compression: none
    tag(262, 3, 1);             // PhotometricInterpretation: BlackIsZero
    tag(273, 4, data_off);      // StripOffsets
    tag(277, 3, 1);             // SamplesPerPixel
    tag(278, 4, height);        // RowsPerStrip
    tag(279, 4, nbytes);        // StripByteCounts
    tag(284, 3, 1);             // PlanarConfiguration: chunky
    tag(339, 3, 3);             // SampleFormat: IEEE float
    put32(0);                   // no next IFD
    f.write(reinterpret_cast<const char*>(data), nbytes);
    if (!f.good()) fail("Error writing image file " + path);
}

// ---------------------------------------------------------------------------
// image-stack directories
// ---------------------------------------------------------------------------

namespace {

float parse_float(const std::string& s, const std::string& what)
{
    std::istringstream is(s);
    float v;
    is >> v;
    if (is.fail()) fail("cannot read " + what + " from '" + s + "'");
    is >> std::ws;
    if (!is.eof()) fail("cannot read " + what + " from '" + s + "'");
    return v;
}

void parse_vec3(const std::string& s, float out[3], const std::string& what)
{
    std::istringstream is(s);
    is >> out[0] >> out[1] >> out[2];
    if (is.fail()) fail("cannot read " + what + " from '" + s + "'");
    is >> std::ws;
    if (!is.eof()) fail("cannot read " + what + " from '" + s + "'");
}

const std::string& value_of(const XmlNode* n, const char* attr, const std::string& what)
{
    if (!n) fail("metadata entry missing: " + what);
    const std::string* v = n->attribute(attr);
    if (!v) fail("metadata entry " + what + " has no '" + attr + "' attribute");
    return *v;
}

}  // namespace

ProjectionStack read_image_stack_directory(const std::string& directory)
{
    const std::string meta = join_path(directory, "metadata.xml");
    auto roots = parse_xml_file(meta);
    struct Entry {
        std::string file;
        uint32_t page;
        bool stack;
        bool logscale;
        float tilt;
    };
    std::vector<Entry> entries;
    const XmlNode* geometry = nullptr;
    for (const auto& r : roots) {
        if (r->name == "geometry" && !geometry) geometry = r.get();
        if (r->name != "projection" && r->name != "projectionstack") continue;
        const std::string& file = value_of(r->child("image"), "filename", "projection.image");
        const float tilt = parse_float(value_of(r->child("tiltAngle"), "value", "projection.tiltAngle"), "tiltAngle");
        const bool logscale = r->child("logscale") != nullptr;
        if (r->name == "projection") {
            entries.push_back({file, 0, false, logscale, tilt});
        } else {
            const uint32_t pages = tiff_page_count(join_path(directory, file));
            for (uint32_t i = 0; i < pages; ++i) entries.push_back({file, i, true, logscale, tilt});
        }
    }
    if (entries.empty()) fail("no projection in " + meta);
    if (!geometry) fail("no <geometry> in " + meta);
    // Projections are indexed by tilt angle in order of first appearance; a
    // repeated angle makes the stack two-dimensional, which the block-iterative
    // loop cannot address (ImageStackDirectoryDatasource.cpp:206-231).
    for (size_t i = 0; i < entries.size(); ++i)
        for (size_t j = 0; j < i; ++j)
            if (entries[i].tilt == entries[j].tilt)
                unsupported("tilt angle " + std::to_string(entries[i].tilt) + " appears twice in " + meta +
                            ": multi-dimensional image stacks are not supported by the block-iterative reconstruction");

    ProjectionStack s;
    s.kind = "ImageStackDirectoryDatasource";
    s.count = (uint32_t)entries.size();
    for (size_t i = 0; i < entries.size(); ++i) {
        const Entry& e = entries[i];
        uint32_t w = 0, h = 0;
        std::vector<float> img = read_tiff(join_path(directory, e.file), e.page, &w, &h);
        if (i == 0) {
            s.nu = w;
            s.nv = h;
            s.data.resize((size_t)s.count * w * h);
        } else if (w != s.nu || h != s.nv) {
            fail("projection " + e.file + " is " + std::to_string(w) + "x" + std::to_string(h) + ", the first is " +
                 std::to_string(s.nu) + "x" + std::to_string(s.nv));
        }
        if (e.logscale)   // Image::convertToLinearScale
            for (float& v : img) v = std::exp(v);
        std::copy(img.begin(), img.end(), s.data.begin() + (ptrdiff_t)(i * (size_t)w * h));
        s.tilt_angles.push_back(e.tilt);
    }
    parse_vec3(value_of(geometry->child("sourcePosition"), "value", "geometry.sourcePosition"), s.base.source,
               "sourcePosition");
    parse_vec3(value_of(geometry->child("detectorPosition"), "value", "geometry.detectorPosition"), s.base.detector,
               "detectorPosition");
    parse_vec3(value_of(geometry->child("horizontalPitch"), "value", "geometry.horizontalPitch"),
               s.base.horizontal_pitch, "horizontalPitch");
    parse_vec3(value_of(geometry->child("verticalPitch"), "value", "geometry.verticalPitch"), s.base.vertical_pitch,
               "verticalPitch");
    s.base.resolution[0] = s.nu;
    s.base.resolution[1] = s.nv;
    return s;
}

ProjectionStack open_projection_stack(const std::string& path, const std::string& tilt_path, bool logarithmize)
{
    if (is_directory(path) && is_regular_file(join_path(path, "metadata.xml"))) return read_image_stack_directory(path);
    const std::string ext = extension(path);
    if (is_regular_file(path) && (ext == ".mrc" || ext == ".ali")) return read_mrc_stack(path, tilt_path, logarithmize);
    fail("Unable to open image stack " + path + "! Incorrect filename or plugin missing?");
}

void write_image_stack_directory(const std::string& directory, const float* data, uint32_t count, uint32_t nu,
                                 uint32_t nv, const float* tilt_angles, const etn_scanner_geometry& base,
                                 const uint32_t volume_resolution[3])
{
    if (!is_directory(directory)) fail("not a directory: " + directory);
    char buf[256];
    auto vec = [&](const float* v) {
        std::snprintf(buf, sizeof buf, "%.9g %.9g %.9g", v[0], v[1], v[2]);
        return std::string(buf);
    };
    std::ostringstream x;
    x << "<?xml version=\"1.0\" encoding=\"utf-8\"?>\n<geometry>\n";
    x << "    <detectorPosition value=\"" << vec(base.detector) << "\"/>\n";
    x << "    <sourcePosition value=\"" << vec(base.source) << "\"/>\n";
    x << "    <horizontalPitch value=\"" << vec(base.horizontal_pitch) << "\"/>\n";
    x << "    <verticalPitch value=\"" << vec(base.vertical_pitch) << "\"/>\n";
    x << "    <volumeDimensions value=\"" << volume_resolution[0] << " " << volume_resolution[1] << " "
      << volume_resolution[2] << "\"/>\n</geometry>\n";
    for (uint32_t i = 0; i < count; ++i) {
        const std::string name = "projection_" + std::to_string(i) + ".tif";
        write_tiff(join_path(directory, name), data + (size_t)i * nu * nv, nu, nv);
        std::snprintf(buf, sizeof buf, "%.9g", tilt_angles[i]);
        x << "<projection>\n    <image filename=\"" << xml_escape(name) << "\"/>\n    <tiltAngle value=\"" << buf
          << "\"/>\n</projection>\n";
    }
    std::ofstream f(join_path(directory, "metadata.xml"), std::ios::trunc);
    if (!f.good()) fail("cannot write metadata.xml in " + directory);
    f << x.str();
}

}  // namespace etn
