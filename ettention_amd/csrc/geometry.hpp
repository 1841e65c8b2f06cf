// Geometry of the parallel-beam single-axis tilt setup: vectors, the 4x4
// matrix, ParallelScannerGeometry, SingleAxisTiltRotator and
// VolumeProperties, with the float arithmetic of the reference host code.
// Compiled with -ffp-contract=off so every expression rounds as written.
#pragma once

#include <array>
#include <cstdint>
#include <vector>

#include "common.hpp"

namespace etn {

struct Vec3f {
    float x = 0.0f, y = 0.0f, z = 0.0f;
    Vec3f() = default;
    Vec3f(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}
    explicit Vec3f(const float* p) : x(p[0]), y(p[1]), z(p[2]) {}
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    void store(float* p) const { p[0] = x; p[1] = y; p[2] = z; }
};
inline Vec3f operator+(const Vec3f& a, const Vec3f& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3f operator-(const Vec3f& a, const Vec3f& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec3f operator*(float s, const Vec3f& a) { return {s * a.x, s * a.y, s * a.z}; }
inline Vec3f operator*(const Vec3f& a, float s) { return {a.x * s, a.y * s, a.z * s}; }

// Vec3::getLength (sqrt in double of the float dot product) and doNormalize
// (components divided in double, narrowed to float); throws on zero length.
// ettention/libReconstruction/src/framework/math/Vec3.h:80-90, Vec3.cpp:67-70
Vec3f normalized(const Vec3f& v);

// Matrix4x4: 16 floats, column-major (values[4*column + row]),
// ettention/libReconstruction/src/framework/math/Matrix4x4.cpp:121-129.
class Matrix4x4 {
public:
    Matrix4x4() { values_.fill(0.0f); }
    static Matrix4x4 from_columns(const std::array<float, 4>& c0, const std::array<float, 4>& c1,
                                  const std::array<float, 4>& c2, const std::array<float, 4>& c3);
    float& at(int column, int row) { return values_[4 * column + row]; }
    float at(int column, int row) const { return values_[4 * column + row]; }
    // cofactor inverse in float (Matrix4x4.cpp:90-119, 207-219)
    Matrix4x4 inverse() const;
    const float* data() const { return values_.data(); }

private:
    float minor_determinant(int dropped_column, int dropped_row) const;
    std::array<float, 16> values_;
};
// Matrix4x4 multiply (Matrix4x4.cpp:297-312)
Matrix4x4 operator*(const Matrix4x4& a, const Matrix4x4& b);

// ParallelScannerGeometry (ScannerGeometry.cpp, ParallelScannerGeometry.cpp).
class ParallelScannerGeometry {
public:
    ParallelScannerGeometry() = default;
    explicit ParallelScannerGeometry(const etn_scanner_geometry& g);
    etn_scanner_geometry to_c() const;

    const Vec3f& source_base() const { return source_; }
    const Vec3f& detector_base() const { return detector_; }
    const Vec3f& horizontal_pitch() const { return hpitch_; }
    const Vec3f& vertical_pitch() const { return vpitch_; }
    uint32_t nu() const { return res_[0]; }
    uint32_t nv() const { return res_[1]; }
    void set(const Vec3f& s, const Vec3f& d, const Vec3f& h, const Vec3f& v)
    {
        source_ = s; detector_ = d; hpitch_ = h; vpitch_ = v;
    }

    // getRayDirection = doNormalize(detector - source) (ParallelScannerGeometry.cpp:44-47)
    Vec3f ray_direction() const;
    // getProjectionMatrix: T(res/2 + 0.5) * [h v r e4]^-1 (ParallelScannerGeometry.cpp:49-62)
    Matrix4x4 projection_matrix() const;
    // getTraversalLengthThroughBoundingBox (ParallelScannerGeometry.cpp:64-69)
    float traversal_length_through(const Vec3f& bmin, const Vec3f& bmax) const;

private:
    Vec3f source_, detector_, hpitch_, vpitch_;
    uint32_t res_[2] = {0, 0};
};

// SingleAxisTiltRotator::createRotatedScannerGeometry / doRotate
// (ettention/libReconstruction/src/model/geometry/SingleAxisTiltRotator.cpp:42-78).
class SingleAxisTiltRotator {
public:
    explicit SingleAxisTiltRotator(const ParallelScannerGeometry& base) : base_(base) {}
    ParallelScannerGeometry rotated(float tilt_deg, float x_axis_tilt_deg = 0.0f) const;
    const ParallelScannerGeometry& base() const { return base_; }

private:
    ParallelScannerGeometry base_;
};

// The geometry of every projection of a reconstruction: desc.geometries when
// given (ImageStackDatasource::createScannerGeometry, ImageStackDatasource.h:26),
// otherwise the base geometry rotated by each tilt angle and the x-axis tilt
// (MRCStack::createScannerGeometry, MRCStack.cpp:281-293).  Every geometry
// takes the detector resolution of desc.base_geometry.
std::vector<ParallelScannerGeometry> projection_geometries(const etn_reconstruction_desc& d);

// One BP sample: offset from the voxel centre and the in-voxel chord length.
struct VoxelSample {
    float dx, dy, dz, length;
};
// ParallelBeamsBackProjectionKernel::generateSamplingPatternForPerVoxelSampling
// (ParallelBeamsBackProjectionKernel.cpp:69-103) with Ray::getInvertedNormalizedDirection
// (Ray.cpp:63-75) and GeometricAlgorithms::getClipRayAgainstBoundingBox (:40-51).
std::vector<VoxelSample> bp_sample_pattern(const ParallelScannerGeometry& g, const Vec3f& voxel_size,
                                           uint32_t samples);

// One z-slab of the volume.
struct SubVolume {
    uint32_t res[3];
    uint32_t first_slice;
    Vec3f bmin, bmax;
    size_t first_voxel;
};

// VolumeProperties (VolumeProperties.cpp:21-98).
class VolumeProperties {
public:
    explicit VolumeProperties(const etn_volume_geometry& g);
    const etn_volume_geometry& desc() const { return g_; }
    uint32_t nx() const { return g_.resolution[0]; }
    uint32_t ny() const { return g_.resolution[1]; }
    uint32_t nz() const { return g_.resolution[2]; }
    size_t voxel_count() const { return (size_t)nx() * ny() * nz(); }
    const Vec3f& voxel_size() const { return voxel_size_; }
    Vec3f bbox_min() const { return Vec3f(g_.bbox_min); }
    Vec3f bbox_max() const { return Vec3f(g_.bbox_max); }
    uint32_t subvolume_count() const { return (uint32_t)subs_.size(); }
    const SubVolume& subvolume(uint32_t i) const;

private:
    etn_volume_geometry g_;
    Vec3f voxel_size_;
    std::vector<SubVolume> subs_;
};

}  // namespace etn
