#include "geometry.hpp"

#include <cfloat>
#include <cmath>

namespace etn {

Vec3f normalized(const Vec3f& v)
{
    const float dot = v.x * v.x + v.y * v.y + v.z * v.z;
    const double length = std::sqrt(static_cast<double>(dot));
    if (length == 0.0) fail("Length of vector is 0!");
    return {static_cast<float>(static_cast<double>(v.x) / length),
            static_cast<float>(static_cast<double>(v.y) / length),
            static_cast<float>(static_cast<double>(v.z) / length)};
}

Matrix4x4 Matrix4x4::from_columns(const std::array<float, 4>& c0, const std::array<float, 4>& c1,
                                  const std::array<float, 4>& c2, const std::array<float, 4>& c3)
{
    Matrix4x4 m;
    const std::array<float, 4>* cols[4] = {&c0, &c1, &c2, &c3};
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) m.at(c, r) = (*cols[c])[r];
    return m;
}

float Matrix4x4::minor_determinant(int dc, int dr) const
{
    int c[3], r[3];
    for (int i = 0; i < 3; ++i) {
        c[i] = i < dc ? i : i + 1;
        r[i] = i < dr ? i : i + 1;
    }
    return at(c[0], r[0]) * (at(c[1], r[1]) * at(c[2], r[2]) - at(c[1], r[2]) * at(c[2], r[1])) -
           at(c[0], r[1]) * (at(c[1], r[0]) * at(c[2], r[2]) - at(c[1], r[2]) * at(c[2], r[0])) +
           at(c[0], r[2]) * (at(c[1], r[0]) * at(c[2], r[1]) - at(c[1], r[1]) * at(c[2], r[0]));
}

Matrix4x4 Matrix4x4::inverse() const
{
    const float det = (at(0, 0) * minor_determinant(0, 0)) - (at(0, 1) * minor_determinant(0, 1)) +
                      (at(0, 2) * minor_determinant(0, 2)) - (at(0, 3) * minor_determinant(0, 3));
    if (std::fabs(det) == 0.0f) fail("matrix not invertible");
    const float one_over = 1.0f / det;
    Matrix4x4 inv;
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            inv.at(r, c) = minor_determinant(c, r) * one_over;
            if ((r + c) % 2 == 1) inv.at(r, c) *= -1.0f;
        }
    return inv;
}

Matrix4x4 operator*(const Matrix4x4& a, const Matrix4x4& b)
{
    Matrix4x4 m;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            float v = 0.0f;
            for (int k = 0; k < 4; ++k) v += a.at(k, j) * b.at(i, k);
            m.at(i, j) = v;
        }
    return m;
}

ParallelScannerGeometry::ParallelScannerGeometry(const etn_scanner_geometry& g)
    : source_(g.source), detector_(g.detector), hpitch_(g.horizontal_pitch), vpitch_(g.vertical_pitch)
{
    res_[0] = g.resolution[0];
    res_[1] = g.resolution[1];
}

etn_scanner_geometry ParallelScannerGeometry::to_c() const
{
    etn_scanner_geometry g;
    source_.store(g.source);
    detector_.store(g.detector);
    hpitch_.store(g.horizontal_pitch);
    vpitch_.store(g.vertical_pitch);
    g.resolution[0] = res_[0];
    g.resolution[1] = res_[1];
    return g;
}

Vec3f ParallelScannerGeometry::ray_direction() const { return normalized(detector_ - source_); }

Matrix4x4 ParallelScannerGeometry::projection_matrix() const
{
    const Vec3f r = ray_direction();
    const Matrix4x4 basis = Matrix4x4::from_columns({hpitch_.x, hpitch_.y, hpitch_.z, 0.0f},
                                                    {vpitch_.x, vpitch_.y, vpitch_.z, 0.0f},
                                                    {r.x, r.y, r.z, 0.0f}, {0.0f, 0.0f, 0.0f, 1.0f});
    const Matrix4x4 m = basis.inverse();
    const float tx = 0.5f * static_cast<float>(res_[0]) + 0.5f;
    const float ty = 0.5f * static_cast<float>(res_[1]) + 0.5f;
    const Matrix4x4 coord = Matrix4x4::from_columns({1.0f, 0.0f, 0.0f, 0.0f}, {0.0f, 1.0f, 0.0f, 0.0f},
                                                    {0.0f, 0.0f, 1.0f, 0.0f}, {tx, ty, 0.0f, 1.0f});
    return coord * m;
}

float ParallelScannerGeometry::traversal_length_through(const Vec3f& bmin, const Vec3f& bmax) const
{
    const Vec3f r = ray_direction();
    const float z_in = (bmin.z - source_.z) / r.z;
    const float z_out = (bmax.z - source_.z) / r.z;
    return std::max(z_in, z_out) - std::min(z_in, z_out);
}

namespace {
// SingleAxisTiltRotator::doRotate (SingleAxisTiltRotator.cpp:62-78), float math.
Vec3f rotate_point(const Vec3f& c, float t, float xt)
{
    if (xt == 0.0f) {
        return {c.x * std::cos(t) - c.z * std::sin(t), c.y, c.x * std::sin(t) + c.z * std::cos(t)};
    }
    const Vec3f p(c.x, c.y * std::cos(xt) + c.z * std::sin(xt), -c.y * std::sin(xt) + c.z * std::cos(xt));
    const Vec3f a = normalized(Vec3f(0.0f, std::cos(xt), -std::sin(xt)));
    const float ct = std::cos(t), st = std::sin(t), omc = 1 - ct;
    return {(ct + omc * a.x * a.x) * p.x + (omc * a.x * a.y - st * a.z) * p.y + (omc * a.x * a.z + st * a.y) * p.z,
            (omc * a.x * a.y + st * a.z) * p.x + (ct + omc * a.y * a.y) * p.y + (omc * a.y * a.z - st * a.x) * p.z,
            (omc * a.x * a.z - st * a.y) * p.x + (omc * a.y * a.z + st * a.x) * p.y + (ct + omc * a.z * a.z) * p.z};
}
}  // namespace

ParallelScannerGeometry SingleAxisTiltRotator::rotated(float tilt_deg, float x_axis_tilt_deg) const
{
    const float t = -tilt_deg / 180.f * static_cast<float>(M_PI);
    const float xt = x_axis_tilt_deg / 180.f * static_cast<float>(M_PI);
    ParallelScannerGeometry g = base_;
    g.set(rotate_point(base_.source_base(), t, xt), rotate_point(base_.detector_base(), t, xt),
          rotate_point(base_.horizontal_pitch(), t, xt), rotate_point(base_.vertical_pitch(), t, xt));
    return g;
}

std::vector<ParallelScannerGeometry> projection_geometries(const etn_reconstruction_desc& d)
{
    std::vector<ParallelScannerGeometry> out;
    out.reserve(d.projection_count);
    if (d.geometries) {
        for (uint32_t i = 0; i < d.projection_count; ++i) {
            etn_scanner_geometry g = d.geometries[i];
            g.resolution[0] = d.base_geometry.resolution[0];
            g.resolution[1] = d.base_geometry.resolution[1];
            out.emplace_back(g);
        }
        return out;
    }
    const SingleAxisTiltRotator rotator{ParallelScannerGeometry(d.base_geometry)};
    for (uint32_t i = 0; i < d.projection_count; ++i) out.push_back(rotator.rotated(d.tilt_angles[i], d.x_axis_tilt));
    return out;
}

std::vector<VoxelSample> bp_sample_pattern(const ParallelScannerGeometry& g, const Vec3f& vs, uint32_t samples)
{
    if (samples == 0) fail("zero samples not supported for BackProjection");
    const uint32_t per_dir = static_cast<uint32_t>(std::sqrt(static_cast<float>(samples)));
    if (per_dir * per_dir != samples) fail("non-square root number of samples not supported for BackProjection");
    const float max_vs = std::fmax(std::fmax(vs.x, vs.y), vs.z);
    const Vec3f h = normalized(g.horizontal_pitch()) * max_vs;
    const Vec3f v = normalized(g.vertical_pitch()) * max_vs;
    // Ray::createFromPointAndDirection normalises the (already unit) direction again.
    const Vec3f dir = normalized(g.ray_direction());
    Vec3f inv;
    for (int k = 0; k < 3; ++k) inv[k] = dir[k] != 0.0f ? 1.0f / dir[k] : FLT_MAX;
    std::vector<VoxelSample> out;
    out.reserve(samples);
    for (float y = 0.0f; y < static_cast<float>(per_dir); y += 1.0f) {
        const float y_pos = ((y + 0.5f) / static_cast<float>(per_dir)) - 0.5f;
        for (float x = 0.0f; x < static_cast<float>(per_dir); x += 1.0f) {
            const float x_pos = ((x + 0.5f) / static_cast<float>(per_dir)) - 0.5f;
            const Vec3f p = (Vec3f(0.0f, 0.0f, 0.0f) + x_pos * h) + y_pos * v;
            float tmin[3], tmax[3];
            for (int k = 0; k < 3; ++k) {
                const float t0 = (vs[k] * -0.5f - p[k]) * inv[k];
                const float t1 = (vs[k] * 0.5f - p[k]) * inv[k];
                tmin[k] = std::min(t0, t1);
                tmax[k] = std::max(t0, t1);
            }
            const float t_in = std::fmax(std::fmax(tmin[0], tmin[1]), tmin[2]);
            const float t_out = std::fmin(std::fmin(tmax[0], tmax[1]), tmax[2]);
            out.push_back({p.x, p.y, p.z, t_out - t_in});
        }
    }
    return out;
}

VolumeProperties::VolumeProperties(const etn_volume_geometry& g) : g_(g)
{
    for (int k = 0; k < 3; ++k) {
        if (g.resolution[k] == 0) fail("volume resolution must be non-zero");
        voxel_size_[k] = (g.bbox_max[k] - g.bbox_min[k]) / static_cast<float>(g.resolution[k]);
        if (!(voxel_size_[k] > 0.0f)) fail("volume bounding box must have positive extent");
    }
    const uint32_t count = g.subvolume_count ? g.subvolume_count : 1;
    if (g.resolution[2] < count) fail("Volume cannot be split into more subvolumes than z slices");
    g_.subvolume_count = count;
    const uint32_t base = g.resolution[2] / count, remaining = g.resolution[2] - count * base;
    Vec3f next_min = bbox_min();
    uint32_t next_z = 0;
    for (uint32_t i = 0; i < count; ++i) {
        SubVolume s;
        s.res[0] = g.resolution[0];
        s.res[1] = g.resolution[1];
        s.res[2] = i < remaining ? base + 1 : base;
        s.first_slice = next_z;
        s.first_voxel = static_cast<size_t>(next_z) * g.resolution[0] * g.resolution[1];
        s.bmin = next_min;
        s.bmax = Vec3f(next_min.x + static_cast<float>(s.res[0]) * voxel_size_.x,
                       next_min.y + static_cast<float>(s.res[1]) * voxel_size_.y,
                       next_min.z + static_cast<float>(s.res[2]) * voxel_size_.z);
        next_min.z = s.bmax.z;
        next_z += s.res[2];
        subs_.push_back(s);
    }
}

const SubVolume& VolumeProperties::subvolume(uint32_t i) const
{
    if (i >= subs_.size()) fail("subvolume index out of range");
    return subs_[i];
}

}  // namespace etn
