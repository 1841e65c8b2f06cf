// Host side of the fused plane kernels (fused_kernel.hpp): the member plans,
// the passes of an iteration, the launch schedule, the autotuner, and the
// back-projection-only engine used by weighted back projection.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <tuple>
#include <vector>

#include "block_iterative.hpp"
#include "fused_kernel.hpp"

namespace etn {
namespace dev {

// vec: every member's nx is a multiple of 4; cap: some member's rays can reach
// the 2048-step limit (step_limit_can_bind);
// exact: the generic kernels' operations and rounding order; mask: the launch
// carries prior-knowledge masks; over: oversampled FP and BP
void launch_fused(etn_voxel_type vt, bool vec, bool cap, bool exact, bool mask, bool over, const FusedLaunch& l,
                  hipStream_t stream)
{
    if (exact) {
        if (vt == ETN_VOXEL_FLOAT32) launch_fused_exact_F32(mask, over, vec, cap, l, stream);
        else if (vt == ETN_VOXEL_FLOAT16) launch_fused_exact_F16(mask, over, vec, cap, l, stream);
        else launch_fused_exact_BF16(mask, over, vec, cap, l, stream);
    } else {
        if (vt == ETN_VOXEL_FLOAT32) launch_fused_fast_F32(mask, over, vec, cap, l, stream);
        else if (vt == ETN_VOXEL_FLOAT16) launch_fused_fast_F16(mask, over, vec, cap, l, stream);
        else launch_fused_fast_BF16(mask, over, vec, cap, l, stream);
    }
    if (!l.resident) hip_check(hipGetLastError(), "fused plane kernel launch");
}

}  // namespace dev

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

struct Tile {
    int planes = 1;
    int band = 1;
};

// One reconstruction driven by the engine: its plane/row tables, its
// residual images and the passes of one iteration under its current tile.
struct MemberPlan {
    BlockIterativeReconstructionOperator* op = nullptr;
    int nx = 0, ny = 0, nz = 0, nu = 0, nv = 0;
    size_t pixels = 0;
    float lambda = 0.0f;
    bool use_loc = false;
    int rows_per_plane = 0;
    int row_len = 0, pad_l = 0;        // staged detector columns, zero columns left of column 0
    int row_floats = 0;                // LDS floats of one plane's staged rows
    const uint8_t* mask = nullptr;     // prior-knowledge mask (device), or null
    int threads = dev::kMaxThreads;    // workgroup shape of the engine's launches
    int vbytes = 4;                    // bytes of a stored voxel (the 512-thread shape's LDS array depends on it)
    int rpt_cap = dev::kMaxRaysPerThread;  // rays per thread a pass may give its forward projections
    bool sr = false;                   // stochastic rounding of the back projections' updates to 16-bit storage
    bool lds_bp = false;               // back projections on the band in LDS, never in registers
    int fp_n = 1, fp_shift = 0;        // forward sub-rays per pixel axis; log2 of ray slots (columns si) per pixel
    int bp_samples = 1, rowsets = 1;   // back projection samples per voxel; pattern rows
    bool oversampled() const { return fp_n != 1 || bp_samples != 1; }
    int ndead = 0;
    uint32_t bmax = 0;                 // largest block
    DeviceBuffer rows;                 // row_lo[ny] | row_hi[ny] | dead rows
    DeviceBuffer images;               // residual + length per block slot, two parities
    Tile tile;
    std::vector<dev::FusedPass> passes;
    int offset = 0;                    // launch that carries the member's first pass
};

int select_rpt(int rays, int threads)
{
    int rpt = 1;
    while (rpt * threads < rays && rpt < dev::kMaxRaysPerThread) rpt *= 2;
    return rpt;
}

// Largest band for `planes` planes per workgroup of `threads`, 0 when the tile does not fit.
int max_band(const MemberPlan& m, int planes, int threads)
{
    if (planes < 1 || planes > dev::kMaxPlanes || (planes & (planes - 1)) || m.ny % planes) return 0;
    if ((planes * m.rows_per_plane * m.nu << m.fp_shift) > threads * dev::kMaxRaysPerThread) return 0;
    if (m.oversampled() && threads != dev::kMaxThreads) return 0;   // oversampled kernels: 512 threads only
    const bool mask = m.mask != nullptr;
    const int masked = mask ? dev::shape_band_floats(threads, true, m.vbytes) : 0;
    int band = std::min(m.nz, (dev::shape_lds_floats(threads, m.vbytes) - masked - 3 -
                               planes * m.rowsets * (m.row_floats + 2)) / (planes * m.nx));
    band = std::min(band, dev::shape_band_floats(threads, mask, m.vbytes) / (planes * m.nx));
    return std::max(band, 0);
}

int max_band(const MemberPlan& m, int planes) { return max_band(m, planes, m.threads); }

// The 80 KiB register back projection (dev::FusedLaunch::res == 4): LDS
// floats of member m's band of `band` rows with `sets` record sets staged, and
// the tallest band (at most the member's own) that fits; 0 when none does or
// when it holds less than half the 80 KiB path's band cap (with the records
// leaving only a row or two, the 160 KiB array's one workgroup per CU with its
// full bands is the faster choice).
size_t res2_lds_bytes(const MemberPlan& m, int band, int sets, bool mask)
{
    const int width = m.tile.planes * m.nx;
    return sizeof(float) * (size_t)((mask ? dev::kRes2Band / 2 : 0) + ((band * width + 3) & ~3) +
                                    m.tile.planes * m.rowsets * sets * (m.row_floats + 2));
}
int res2_band(const MemberPlan& m, int sets, bool mask)
{
    const int width = m.tile.planes * m.nx;
    const int cap = mask ? dev::kRes2Band / 2 : dev::kRes2Band;
    const int rec = m.tile.planes * m.rowsets * sets * (m.row_floats + 2);
    const int room = dev::kRes2Lds - (mask ? dev::kRes2Band / 2 : 0) - rec - 3;
    if (room < width) return 0;
    const int band = std::min(m.tile.band, std::min(cap / width, room / width));
    return 2 * band * width >= cap || band == m.tile.band ? band : 0;
}

// Dispatch order of a launch's workgroups.  The plane groups of a launch that
// the device cannot hold at once run in rounds; in plane order the first round
// then streams only the first part of every z row of the volume (a contiguous
// nx * ny * voxel-size slice per row), and with 4 MiB slices (1024 x 1024
// f32) that part alone left HBM half idle: 1024 x 1024 x 256 f32
// SART ran at 485-489 GUPS against 605-620 with this order.  Here the rounds' groups
// are interleaved: with r rounds of n groups each, dispatch i runs group
// (i mod r) * n + i / r, so every round spans the whole slice in r contiguous
// runs.  Every group still runs once and the groups are independent within a
// launch, so the volumes do not change.  A batch launch applies it to each
// member's groups alone (interleaving across members had cost the 64-member
// batch 11%).
void interleave_rounds(dev::FusedWork* work, int groups, int resident)
{
    if (resident < 1 || groups <= resident) return;
    const int rounds = (groups + resident - 1) / resident;
    const int per = (groups + rounds - 1) / rounds;
    std::vector<dev::FusedWork> in(work, work + groups), out;
    out.reserve(groups);
    for (int i = 0; i < per; ++i)
        for (int r = 0; r < rounds; ++r)
            if (r * per + i < groups) out.push_back(in[r * per + i]);
    std::copy(out.begin(), out.end(), work);
}

// The 160 KiB LDS back projection (dev::FusedLaunch::res == 3): the tallest
// band of member m (one record set per plane and row set staged) within the
// 512-thread shape's band cap, and its LDS bytes.
int res3_band(const MemberPlan& m, bool mask)
{
    const int width = m.tile.planes * m.nx;
    const int masked = mask ? dev::shape_band_floats(dev::kMaxThreads, true, m.vbytes) : 0;
    const int rec = m.tile.planes * m.rowsets * (m.row_floats + 2);
    const int room = dev::kResLds - masked - rec - 3;
    if (room < width) return 0;
    return std::min(m.nz, std::min(dev::shape_band_floats(dev::kMaxThreads, mask, m.vbytes) / width, room / width));
}
size_t res3_lds_bytes(const MemberPlan& m, int band, bool mask)
{
    return sizeof(float) * (size_t)dev::fused_lds_floats(m.tile.planes, band, m.nx, m.row_floats, mask,
                                                          dev::kMaxThreads, m.rowsets, m.vbytes);
}

bool tile_fits(const MemberPlan& m, const Tile& t)
{
    return t.band >= 1 && t.band <= max_band(m, t.planes) &&
           dev::fused_lds_floats(t.planes, t.band, m.nx, m.row_floats, m.mask != nullptr, m.threads, m.rowsets,
                                 m.vbytes) <= dev::shape_lds_floats(m.threads, m.vbytes);
}

// Without timing: the fewest planes whose rays fill a workgroup, the tallest band.
Tile default_tile(const MemberPlan& m)
{
    Tile best{1, max_band(m, 1)};
    for (int p = 1; p <= dev::kMaxPlanes; p *= 2) {
        const int b = max_band(m, p);
        if (b < 1) continue;
        best = {p, b};
        if ((p * m.rows_per_plane * m.nu << m.fp_shift) >= m.threads) break;
    }
    return best;
}

std::vector<Tile> candidate_tiles(const MemberPlan& m)
{
    std::vector<Tile> out;
    for (int p = 1; p <= dev::kMaxPlanes; p *= 2) {
        const int b = max_band(m, p);
        if (b < 1) continue;
        out.push_back({p, b});
        if (b >= 2) out.push_back({p, (b + 1) / 2});
    }
    return out;
}

// The fused kernels sweep the whole volume.  A z-slab split
// (hardware.subVolumeCount) restarts the reference's forward projection at
// every slab face and sums the slab parts; sweeping the volume in one pass
// takes the same steps and only regroups those sums and the voxel-centre
// arithmetic, so it is allowed with fast arithmetic and left to the generic
// kernels with exact arithmetic.
VolumeProperties whole_volume(const VolumeProperties& vp)
{
    etn_volume_geometry g = vp.desc();
    g.subvolume_count = 1;
    return VolumeProperties(g);
}

// The fast lookup skips the column clamp: stage every column u - 0.5 can
// floor to for a voxel centre (and every sample of the BP pattern) under any
// of the geometries, plus one column either side.  False when the span is
// unbounded.
bool fast_row_span(const VolumeProperties& vp, const std::vector<const ParallelScannerGeometry*>& geoms, int nu,
                   int bp_s, int& pad_l, int& row_len)
{
    const Vec3f vs = vp.voxel_size(), lo = vp.bbox_min(), hi = vp.bbox_max();
    double umin = 1e300, umax = -1e300;
    for (const ParallelScannerGeometry* g : geoms) {
        const Matrix4x4 pm = g->projection_matrix();
        const float* mv = pm.data();
        for (const VoxelSample& smp : bp_sample_pattern(*g, vs, (uint32_t)bp_s))
            for (int c = 0; c < 8; ++c) {
                const double px = ((c & 1) ? hi.x - 0.5 * vs.x : lo.x + 0.5 * vs.x) + smp.dx;
                const double py = ((c & 2) ? hi.y - 0.5 * vs.y : lo.y + 0.5 * vs.y) + smp.dy;
                const double pz = ((c & 4) ? hi.z - 0.5 * vs.z : lo.z + 0.5 * vs.z) + smp.dz;
                const double u = mv[0] * px + mv[4] * py + mv[8] * pz + mv[12];
                umin = std::min(umin, u);
                umax = std::max(umax, u);
            }
    }
    if (!(umin > -1e6 && umax < 1e6)) return false;
    const int bu_lo = (int)std::floor(umin - 0.5) - 1, bu_hi = (int)std::floor(umax - 0.5) + 1;
    pad_l = std::max(dev::kRowPad, -bu_lo);
    row_len = pad_l + std::max(nu + dev::kRowPad, bu_hi + 1);
    return true;
}

// Whether the reference's 2048-step limit (2048 voxels per ray and volume
// sweep) can bind for one reconstruction.  A ray confined to a y plane visits
// one voxel more than it crosses x and z voxel borders.  With a chord spanning
// dx voxels in x and dz in z (dx <= nx, dz <= nz, dx / dz fixed by the ray
// direction in voxel units) it crosses at most floor(dx) + 1 x borders and
// floor(dz) + 1 z borders, and never more than nx + nz - 2 in all.  Only
// when some projection's bound exceeds 2040 voxels (8 below the limit, for
// the traversal's float rounding) do the launches take the step-limited
// kernels; otherwise the limit is never reached and the kernels without it
// compute the same sums.  (2048-wide volumes at tilts up to 60 degrees,
// for example, stay far below it.)
bool step_limit_can_bind(BlockIterativeReconstructionOperator& op)
{
    const VolumeProperties vp = whole_volume(op.getReconstructedVolume()->properties());
    const double nx = (double)vp.nx(), nz = (double)vp.nz();
    if (nx + nz - 1.0 <= 2048.0) return false;
    const Vec3f vs = vp.voxel_size();
    for (uint32_t b = 0; b < op.GetBlockCount(); ++b)
        for (uint32_t p = 0; p < op.getBlockSize(b); ++p) {
            const Vec3f r = op.geometry(op.getPhysicalProjectionIndex(b, p)).ray_direction();
            const double ax = std::fabs((double)r.x) / (double)vs.x, az = std::fabs((double)r.z) / (double)vs.z;
            double dx = nx, dz = nz;
            if (az == 0.0) dz = 0.0;
            else if (ax == 0.0) dx = 0.0;
            else {
                dx = std::min(nx, nz * ax / az);
                dz = std::min(nz, nx * az / ax);
            }
            const double visits = std::min(nx + nz - 1.0, std::floor(dx) + std::floor(dz) + 3.0);
            if (visits > 2040.0) return true;
        }
    return false;
}

// Eligibility of one reconstruction and its plane/row tables; false when it
// has no fused form.
bool plan_member(BlockIterativeReconstructionOperator& op, const etn_reconstruction_desc& d, bool exact,
                 MemberPlan& out)
{
    if (exact && op.getReconstructedVolume()->properties().subvolume_count() != 1) return false;
    const VolumeProperties whole = whole_volume(op.getReconstructedVolume()->properties());
    const VolumeProperties& vp = whole;
    // oversampling: up to 4 x 4 forward sub-rays and a 4 x 4 back projection pattern
    const int fp_n = (int)(d.fp_samples ? d.fp_samples : 1);
    const int bp_s = (int)(d.bp_samples ? d.bp_samples : 1);
    const int bp_pd = (int)std::lround(std::sqrt((double)bp_s));
    if (fp_n > dev::kMaxFusedFpPerAxis || bp_s > dev::kMaxFusedBpSamples || bp_pd * bp_pd != bp_s) return false;
    // forward sub-rays restarted at every slab face (the reference's slab-wise
    // FP) differ from one sweep by up to 2.4e-4 relative L2 after two SART
    // iterations (4 x 4 sub-rays, 2 or 4 slabs): beyond the FP32 bound, so an
    // oversampled FP over slabs stays on the generic kernels
    if (fp_n != 1 && op.getReconstructedVolume()->properties().subvolume_count() != 1) return false;
    int fp_shift = 0;
    while ((1 << fp_shift) < fp_n) ++fp_shift;     // ray slots per pixel: sub-ray columns si
    if (reinterpret_cast<uintptr_t>(d.mask) % 4) return false;   // masks are read 4 bytes at a time
    const int nx = (int)vp.nx(), ny = (int)vp.ny(), nz = (int)vp.nz();
    const int nu = (int)d.base_geometry.resolution[0], nv = (int)d.base_geometry.resolution[1];

    // every projection: rays in y planes, constant detector row per plane
    std::vector<int> plane_of_row;
    for (uint32_t b = 0; b < op.GetBlockCount(); ++b)
        for (uint32_t p = 0; p < op.getBlockSize(b); ++p) {
            const ParallelScannerGeometry& g = op.geometry(op.getPhysicalProjectionIndex(b, p));
            const Vec3f r = g.ray_direction();
            if (r.y != 0.0f || g.horizontal_pitch().y != 0.0f) return false;
            // oversampled FP: sub-rays of one column differ only in y
            if (fp_n > 1 && (g.vertical_pitch().x != 0.0f || g.vertical_pitch().z != 0.0f)) return false;
            // nonzero x/z components must have finite reciprocals (the plane
            // DDA relies on every res value staying at or below FLT_MAX)
            for (float c : {r.x, r.z})
                if (c != 0.0f && std::fabs(c) < 1e-30f) return false;
            const Matrix4x4 m = g.projection_matrix();
            const float* mv = m.data();
            if (mv[1] != 0.0f || mv[9] != 0.0f || mv[3] != 0.0f || mv[7] != 0.0f || mv[11] != 0.0f || mv[15] != 1.0f)
                return false;
            // plane of each detector row, with the FP's own arithmetic for every
            // sub-ray row sj (all must agree); -1 for rows whose sources lie
            // outside the box in y by a margin that makes every ray's y slab
            // interval empty (t_out <= -2^-20 * FLT_MAX)
            std::vector<int> map(nv);
            const SubVolume& sub = vp.subvolume(0);
            const float ivs = 1.0f / vp.voxel_size().y;
            const float margin = 1.0f / 1048576.0f;
            const float sd = fp_n == 1 ? 1.0f : 1.0f / (float)fp_n;
            const float off = fp_n == 1 ? 0.5f : sd / 2.0f;
            for (int w = 0; w < nv; ++w) {
                for (int sj = 0; sj < fp_n; ++sj) {
                    const float fx = (float)0 + off + sd * 0.0f;
                    const float fy = (float)w + off + sd * (float)sj;
                    const float ru = g.horizontal_pitch().y * fx, rv = g.vertical_pitch().y * fy;
                    float y = g.source_base().y + ru + rv;
                    int jj;
                    if (sub.bmin.y - y >= margin || y - sub.bmax.y >= margin) {
                        jj = -1;
                    } else {
                        if (y < sub.bmin.y || y > sub.bmax.y) return false;   // too close to call: generic kernels
                        y = std::fmin(std::fmax(y, sub.bmin.y), sub.bmax.y);
                        jj = (int)((y - sub.bmin.y) * ivs);
                        jj = std::max(jj, 0);
                        jj = std::min(jj, ny - 1);
                    }
                    if (sj == 0) map[w] = jj;
                    else if (map[w] != jj) return false;   // the pixel's sub-rays cross planes
                }
            }
            if (plane_of_row.empty()) plane_of_row = map;
            else if (plane_of_row != map) return false;
        }
    std::vector<int> lo(ny, 0), hi(ny, 0), dead;
    std::vector<int> seen(ny, 0);
    int prev = -1;
    for (int w = 0; w < nv; ++w) {
        const int jj = plane_of_row[w];
        if (jj < 0) {
            dead.push_back(w);
            continue;
        }
        if (jj != prev) {
            if (seen[jj]) return false;        // rows of a plane must be contiguous
            seen[jj] = 1;
            lo[jj] = w;
            prev = jj;
        }
        hi[jj] = w + 1;
    }
    int rows_per_plane = 0;
    for (int jj = 0; jj < ny; ++jj) rows_per_plane = std::max(rows_per_plane, hi[jj] - lo[jj]);
    if (rows_per_plane < 1 || (rows_per_plane * nu << fp_shift) > dev::kMaxThreads * dev::kMaxRaysPerThread)
        return false;

    out.op = &op;
    out.nx = nx;
    out.ny = ny;
    out.nz = nz;
    out.nu = nu;
    out.nv = nv;
    out.pixels = (size_t)nu * nv;
    out.lambda = d.lambda;
    out.use_loc = d.use_long_object_compensation != 0;
    out.rows_per_plane = rows_per_plane;
    out.pad_l = dev::kRowPad;
    out.row_len = nu + 2 * dev::kRowPad;
    if (!exact) {
        std::vector<const ParallelScannerGeometry*> geoms;
        for (uint32_t b = 0; b < op.GetBlockCount(); ++b)
            for (uint32_t p = 0; p < op.getBlockSize(b); ++p)
                geoms.push_back(&op.geometry(op.getPhysicalProjectionIndex(b, p)));
        if (!fast_row_span(vp, geoms, nu, bp_s, out.pad_l, out.row_len)) return false;
    }
    out.row_floats = out.row_len * (exact ? 8 : 4);
    out.mask = d.mask;
    out.threads = dev::kMaxThreads;
    out.vbytes = (int)voxel_bytes(d.voxel_type);
    out.fp_n = fp_n;
    out.fp_shift = fp_shift;
    out.bp_samples = bp_s;
    out.rowsets = bp_pd;
    if (dev::fused_lds_floats(1, 1, nx, out.row_floats, out.mask != nullptr, out.threads, out.rowsets,
                              out.vbytes) > dev::shape_lds_floats(out.threads, out.vbytes))
        return false;   // a z row and the rows fit
    if (max_band(out, 1) < 1) return false;   // not even one z row of one plane fits
    out.ndead = (int)dead.size();
    std::vector<int> table(lo);
    table.insert(table.end(), hi.begin(), hi.end());
    table.insert(table.end(), dead.begin(), dead.end());
    out.rows.allocate(sizeof(int) * table.size());
    ETN_HIP(hipMemcpy(out.rows.get(), table.data(), sizeof(int) * table.size(), hipMemcpyHostToDevice));
    out.bmax = 0;
    for (uint32_t b = 0; b < op.GetBlockCount(); ++b) out.bmax = std::max(out.bmax, op.getBlockSize(b));
    out.images.allocate(sizeof(float) * out.pixels * 2 * 2 * out.bmax);
    out.tile = default_tile(out);
    return true;
}

// The passes of one iteration: pass i back-projects what is left of the
// previous block and forward-projects a group of the current block.
void build_passes(MemberPlan& mb)
{
    BlockIterativeReconstructionOperator& op = *mb.op;
    mb.passes.clear();
    const int max_fp = std::max(1, std::min(dev::kMaxFusedFp, mb.threads * mb.rpt_cap /
                                                                  ((mb.tile.planes * mb.rows_per_plane * mb.nu)
                                                                   << mb.fp_shift)));
    auto image = [&](uint32_t parity, uint32_t slot, int which) {
        return mb.images.as<float>() + mb.pixels * ((((size_t)parity * mb.bmax + slot) * 2) + which);
    };
    const VolumeProperties vp = whole_volume(op.getReconstructedVolume()->properties());
    std::vector<dev::FusedBp> pending;   // back projections of the previous block
    auto flush_bp_only = [&](size_t keep) {
        // emit BP-only passes until at most `keep` back projections are pending
        while (pending.size() > keep) {
            dev::FusedPass p{};
            p.zdir = 1;
            const size_t n = std::min(pending.size() - keep, (size_t)dev::kMaxFusedBp);
            for (size_t i = 0; i < n; ++i) p.bp[p.nbp++] = pending[i];
            pending.erase(pending.begin(), pending.begin() + n);
            mb.passes.push_back(p);
        }
    };
    const uint32_t nb = op.GetBlockCount();
    for (uint32_t b = 0; b < nb; ++b) {
        const uint32_t bs = op.getBlockSize(b);
        const uint32_t parity = b & 1u;
        // long-object compensation uses the block's last geometry (compensateLongObject)
        const ParallelScannerGeometry& last = op.geometry(op.getPhysicalProjectionIndex(b, bs - 1));
        const float loc_c = mb.use_loc ? last.traversal_length_through(vp.bbox_min(), vp.bbox_max()) : 0.0f;
        std::vector<dev::FusedFp> neg, pos;
        for (uint32_t p = 0; p < bs; ++p) {
            const uint32_t phys = op.getPhysicalProjectionIndex(b, p);
            const ParallelScannerGeometry& g = op.geometry(phys);
            dev::FusedFp f{};
            g.source_base().store(f.src);
            g.horizontal_pitch().store(f.h);
            g.vertical_pitch().store(f.v);
            g.ray_direction().store(f.ray);
            f.measured = op.realProjection(phys);
            f.measured_type = op.projectionType();
            f.residual = image(parity, p, 0);
            f.length = image(parity, p, 1);
            f.loc_c = loc_c;
            f.use_loc = mb.use_loc ? 1 : 0;
            (f.ray[2] > 0.0f ? pos : neg).push_back(f);
        }
        // FP groups: rays moving towards -z (or staying in z) and towards +z
        std::vector<std::pair<int, std::vector<dev::FusedFp>>> groups;
        for (auto* grp : {&neg, &pos}) {
            const int zdir = grp == &pos ? 1 : -1;
            for (size_t i = 0; i < grp->size(); i += (size_t)max_fp)
                groups.push_back({zdir, std::vector<dev::FusedFp>(grp->begin() + i,
                                                                  grp->begin() + std::min(grp->size(), i + (size_t)max_fp))});
        }
        flush_bp_only(dev::kMaxFusedBp);
        for (size_t gi = 0; gi < groups.size(); ++gi) {
            dev::FusedPass p{};
            p.zdir = groups[gi].first;
            for (const auto& f : groups[gi].second) p.fp[p.nfp++] = f;
            if (gi == 0)
                for (const auto& e : pending) p.bp[p.nbp++] = e;
            mb.passes.push_back(p);
        }
        pending.clear();
        for (uint32_t p = 0; p < bs; ++p) {
            const ParallelScannerGeometry& g = op.geometry(op.getPhysicalProjectionIndex(b, p));
            dev::FusedBp e{};
            std::memcpy(e.m, g.projection_matrix().data(), sizeof(e.m));
            e.lambda = mb.lambda / (float)bs;
            e.residual = image(parity, p, 0);
            e.length = image(parity, p, 1);
            const std::vector<VoxelSample> smp = bp_sample_pattern(g, vp.voxel_size(), (uint32_t)mb.bp_samples);
            e.sample = make_float4(smp[0].dx, smp[0].dy, smp[0].dz, smp[0].length);
            e.nsamples = (int)smp.size();
            e.per_dir = mb.rowsets;
            e.sr_salt = mb.sr ? op.getPhysicalProjectionIndex(b, p) + 1u : 0u;
            for (size_t k = 0; k < smp.size(); ++k)
                e.samples[k] = make_float4(smp[k].dx, smp[k].dy, smp[k].dz, smp[k].length);
            pending.push_back(e);
        }
    }
    flush_bp_only(0);
}

class FusedEngineImpl final : public FusedPlaneEngine {
public:
    FusedEngineImpl(std::vector<MemberPlan> members, etn_voxel_type vt, bool vec, bool cap, bool exact, bool mask,
                    bool over, hipStream_t stream);
    void enqueueIteration(hipStream_t stream) override;
    void check(hipStream_t stream) override;
    std::vector<LaunchTime> timeIteration(hipStream_t stream) override;
    void autotune(hipStream_t stream) override;
    void setTile(size_t member, int planes, int band, hipStream_t stream) override;
    void tile(size_t member, int* planes, int* band) const override;
    void setThreads(int threads, hipStream_t stream) override;
    int threads() const override { return threads_; }
    size_t memberCount() const override { return members_.size(); }
    double residualNorm(size_t member, double* scratch, hipStream_t stream) override;

private:
    struct Launch {
        size_t work;            // first work item
        int groups;
        int rpt;
        size_t lds;
        bool steady;
        int res;                // register back projection (dev::FusedLaunch::res)
        size_t lds_res;         // LDS bytes with every back projection's records staged
        int band4;              // z rows per band of the 80 KiB register path (res 4) or the
        size_t lds4;            // 160 KiB LDS path (res 3), 0: none; its LDS bytes
    };
    struct Schedule {
        std::vector<Launch> launches;
        DeviceBuffer work, tables;
    };
    void upload(hipStream_t stream);
    void level_offsets();
    void build_schedule(const std::vector<size_t>& ids, Schedule& out, hipStream_t stream) const;
    void enqueue(const Schedule& s, size_t first, size_t last, hipStream_t stream, std::vector<hipEvent_t>* ev) const;
    bool threadsFit(int threads) const;
    void useThreads(int threads);
    void useRptCap(int cap);
    size_t passCount() const;
    void tuneTiles(hipStream_t stream);
    float timeSchedule(const Schedule& s, hipStream_t stream) const;

    std::vector<MemberPlan> members_;
    etn_voxel_type vt_;
    bool vec_, cap_, exact_;
    bool mask_;                         // every member has a mask (the launches use the mask kernels)
    bool over_;                         // every member oversamples (the launches use the oversampled kernels)
    bool sr_;                           // stochastic rounding to 16-bit storage
    bool reg_bp_;                       // register back projection where it applies
    int threads_ = dev::kMaxThreads;    // workgroup shape of every launch
    int dev_cus_ = 0;                   // compute units of the device
    DeviceBuffer diag_, vols_, passes_;
    std::vector<size_t> pass_base_;     // each member's first pass in passes_
    Schedule sched_;
};

FusedEngineImpl::FusedEngineImpl(std::vector<MemberPlan> members, etn_voxel_type vt, bool vec, bool cap, bool exact,
                                 bool mask, bool over, hipStream_t stream)
    : members_(std::move(members)), vt_(vt), vec_(vec), cap_(cap), exact_(exact), mask_(mask), over_(over),
      sr_(!members_.empty() && members_[0].sr), reg_bp_(!members_.empty() && !members_[0].lds_bp)
{
    diag_.allocate(sizeof(int));
    ETN_HIP(hipMemsetAsync(diag_.get(), 0, sizeof(int), stream));
    int dev = 0;
    ETN_HIP(hipGetDevice(&dev));
    ETN_HIP(hipDeviceGetAttribute(&dev_cus_, hipDeviceAttributeMultiprocessorCount, dev));
    upload(stream);
}

// Passes and volume descriptors of every member under the current tiles, and
// the schedule of all members.
void FusedEngineImpl::upload(hipStream_t stream)
{
    std::vector<dev::FusedPass> all;
    std::vector<dev::FusedVolume> vols;
    pass_base_.clear();
    for (MemberPlan& m : members_) {
        build_passes(m);
        pass_base_.push_back(all.size());
        all.insert(all.end(), m.passes.begin(), m.passes.end());
        GpuVolume* vol = m.op->getReconstructedVolume();
        const VolumeProperties vp = whole_volume(vol->properties());
        dev::FusedVolume fv{};
        fv.data = vol->data();
        fv.slab = make_slab_params(vp, 0);
        fv.nx = m.nx;
        fv.ny = m.ny;
        fv.nz = m.nz;
        fv.nu = m.nu;
        fv.nv = m.nv;
        fv.planes = m.tile.planes;
        fv.band = m.tile.band;
        fv.rows_per_plane = m.rows_per_plane;
        fv.row_len = m.row_len;
        fv.pad_l = m.pad_l;
        fv.row_lo = m.rows.as<int>();
        fv.row_hi = m.rows.as<int>() + m.ny;
        fv.dead_rows = m.rows.as<int>() + 2 * m.ny;
        fv.ndead = m.ndead;
        fv.mask = m.mask;
        fv.fp_n = m.fp_n;
        fv.fp_shift = m.fp_shift;
        fv.rowsets = m.rowsets;
        fv.diag = diag_.as<int>();
        vols.push_back(fv);
    }
    ETN_HIP(hipStreamSynchronize(stream));   // no launch may still read the old tables
    passes_.allocate(sizeof(dev::FusedPass) * all.size());
    vols_.allocate(sizeof(dev::FusedVolume) * vols.size());
    ETN_HIP(hipMemcpyAsync(passes_.get(), all.data(), passes_.bytes(), hipMemcpyHostToDevice, stream));
    ETN_HIP(hipMemcpyAsync(vols_.get(), vols.data(), vols_.bytes(), hipMemcpyHostToDevice, stream));
    level_offsets();
    std::vector<size_t> ids(members_.size());
    std::iota(ids.begin(), ids.end(), 0);
    build_schedule(ids, sched_, stream);
    ETN_HIP(hipStreamSynchronize(stream));
}

// Members with fewer passes than the longest start later or earlier so that
// the voxels a launch covers stay as even as possible across launches: each
// member, largest total first, takes the offset whose busiest launch is
// least loaded.
void FusedEngineImpl::level_offsets()
{
    size_t length = 0;
    for (const MemberPlan& m : members_) length = std::max(length, m.passes.size());
    std::vector<size_t> order(members_.size());
    std::iota(order.begin(), order.end(), 0);
    auto voxels = [&](size_t i) { return (double)members_[i].nx * members_[i].ny * members_[i].nz; };
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        return voxels(a) * members_[a].passes.size() > voxels(b) * members_[b].passes.size();
    });
    std::vector<double> load(length, 0.0);
    for (size_t i : order) {
        const size_t n = members_[i].passes.size();
        size_t best = 0;
        double best_peak = 0.0, best_sum = 0.0;
        for (size_t o = 0; o + n <= length; ++o) {
            double peak = 0.0, sum = 0.0;
            for (size_t k = o; k < o + n; ++k) {
                peak = std::max(peak, load[k]);
                sum += load[k];
            }
            if (o == 0 || peak < best_peak || (peak == best_peak && sum < best_sum)) {
                best = o;
                best_peak = peak;
                best_sum = sum;
            }
        }
        members_[i].offset = (int)best;
        for (size_t k = best; k < best + n; ++k) load[k] += voxels(i);
    }
}

// Launch k carries pass k - offset of every member in `ids` that has one,
// members with the most work per workgroup first.
void FusedEngineImpl::build_schedule(const std::vector<size_t>& ids, Schedule& out, hipStream_t stream) const
{
    size_t length = 0;
    for (size_t i : ids) length = std::max(length, (size_t)members_[i].offset + members_[i].passes.size());
    std::vector<size_t> order(ids);
    auto group_work = [&](size_t i) {
        const MemberPlan& m = members_[i];
        return (double)m.tile.planes * m.nx * m.nz;
    };
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return group_work(a) > group_work(b); });
    const dev::FusedPass* base = passes_.as<dev::FusedPass>();
    std::vector<dev::FusedWork> work;
    std::vector<const dev::FusedPass*> tables(length * members_.size(), nullptr);
    std::map<long, int> resident;   // workgroups the device holds at once, per kernel
    out.launches.clear();
    for (size_t k = 0; k < length; ++k) {
        Launch l{work.size(), 0, 1, 0, true, 0, 0, 0, 0};
        std::vector<std::pair<size_t, size_t>> segments;   // (first work item, groups) per member
        bool fit_shape = true, fit_res = true;
        int band4 = 1 << 30;
        int max_bp = 0;
        for (size_t i : order) {
            const MemberPlan& m = members_[i];
            if (k < (size_t)m.offset || k >= (size_t)m.offset + m.passes.size()) continue;
            const size_t pi = k - m.offset;
            const dev::FusedPass& p = m.passes[pi];
            tables[k * members_.size() + i] = base + pass_base_[i] + pi;
            segments.push_back({work.size(), (size_t)(m.ny / m.tile.planes)});
            for (int j = 0; j < m.ny; j += m.tile.planes) work.push_back({(int)i, j});
            l.rpt = std::max(l.rpt, select_rpt((p.nfp * m.tile.planes * m.rows_per_plane * m.nu) << m.fp_shift,
                                               threads_));
            l.lds = std::max(l.lds, sizeof(float) * (size_t)dev::fused_lds_floats(m.tile.planes, m.tile.band, m.nx,
                                                                                   m.row_floats, mask_, threads_,
                                                                                   m.rowsets, m.vbytes));
            l.steady = l.steady && p.nfp > 0 && p.nbp > 0;
            max_bp = std::max(max_bp, p.nbp);
            const size_t res = sizeof(float) * (size_t)dev::fused_lds_floats(m.tile.planes, m.tile.band, m.nx,
                                                                             m.row_floats, mask_, threads_,
                                                                             m.rowsets * std::max(p.nbp, 1), m.vbytes);
            l.lds_res = std::max(l.lds_res, res);
            // the register path needs a thread's groups at one x in every band row
            const bool same_x = threads_ % (m.tile.planes * m.nx / 4) == 0;
            fit_shape = fit_shape && same_x && res <= sizeof(float) * (size_t)dev::shape_lds_floats(threads_, m.vbytes);
            fit_res = fit_res && same_x && res <= sizeof(float) * (size_t)dev::kResLds;
            band4 = same_x ? std::min(band4, res2_band(m, std::max(p.nbp, 1), mask_)) : 0;
        }
        if (band4 >= 1 && band4 < (1 << 30)) {
            l.band4 = band4;
            for (size_t i : order) {
                const MemberPlan& m = members_[i];
                if (k < (size_t)m.offset || k >= (size_t)m.offset + m.passes.size()) continue;
                const dev::FusedPass& p = m.passes[k - m.offset];
                l.lds4 = std::max(l.lds4, res2_lds_bytes(m, band4, std::max(p.nbp, 1), mask_));
            }
        }
        // passes with one back projection keep the LDS path: in registers it
        // gained nothing with f32 storage (638 against 636 GUPS at 512^3) and
        // lost 13% with bf16 (the band's values and the next band's loads
        // exceed the VGPR budget of four waves per SIMD)
        if (reg_bp_ && !exact_ && !over_ && vec_ && !cap_ && !sr_ && max_bp > 1)
            l.res = fit_shape ? 1 : threads_ == 512 && l.band4 > 0 ? 4 : threads_ == 512 && fit_res ? 2 : 0;
        l.groups = (int)(work.size() - l.work);
        auto resident_of = [&](const Launch& c) {
            const long key = ((long)c.rpt << 32) | ((long)c.res << 16) | (long)c.band4;
            auto it = resident.find(key);
            if (it == resident.end()) {
                int r = 0;
                const dev::FusedLaunch q{threads_, c.rpt, c.groups, nullptr, nullptr, nullptr,
                                         c.res == 4 || c.res == 3 ? c.lds4 : c.res ? c.lds_res : c.lds, sr_, c.res,
                                         c.res == 4 || c.res == 3 ? c.band4 : 0, &r};
                dev::launch_fused(vt_, vec_, cap_, exact_, mask_, over_, q, stream);
                it = resident.emplace(key, r).first;
            }
            return it->second;
        };
        // a launch on the LDS path whose registers hold one 512-thread
        // workgroup per CU takes the 160 KiB array where that gives taller
        // bands than the tiles (wide volumes, whose records fill most of the
        // 64 KiB shape)
        if (l.groups > 0 && l.res == 0 && threads_ == 512 && vec_ && !cap_ && !over_ && !sr_ &&
            resident_of(l) <= dev_cus_) {
            int band3 = 1 << 30, tile_band = 1 << 30;
            size_t lds3 = 0;
            for (const auto& sg : segments) {
                const MemberPlan& m = members_[work[sg.first].member];
                band3 = std::min(band3, res3_band(m, mask_));
                tile_band = std::min(tile_band, m.tile.band);
            }
            for (const auto& sg : segments)
                lds3 = std::max(lds3, res3_lds_bytes(members_[work[sg.first].member], band3, mask_));
            if (band3 > tile_band && band3 < (1 << 30)) {
                l.res = 3;
                l.band4 = band3;
                l.lds4 = lds3;
            }
        }
        if (l.groups > 0) {
            const int held = resident_of(l);
            // each member's groups on their own (a batch keeps its largest-first order)
            for (const auto& sg : segments) interleave_rounds(work.data() + sg.first, (int)sg.second, held);
        }
        out.launches.push_back(l);
    }
    ETN_HIP(hipStreamSynchronize(stream));
    out.work.allocate(sizeof(dev::FusedWork) * std::max<size_t>(work.size(), 1));
    out.tables.allocate(sizeof(const dev::FusedPass*) * std::max<size_t>(tables.size(), 1));
    ETN_HIP(hipMemcpyAsync(out.work.get(), work.data(), sizeof(dev::FusedWork) * work.size(), hipMemcpyHostToDevice,
                           stream));
    ETN_HIP(hipMemcpyAsync(out.tables.get(), tables.data(), sizeof(const dev::FusedPass*) * tables.size(),
                           hipMemcpyHostToDevice, stream));
    ETN_HIP(hipStreamSynchronize(stream));
}

void FusedEngineImpl::enqueue(const Schedule& s, size_t first, size_t last, hipStream_t stream,
                              std::vector<hipEvent_t>* ev) const
{
    const dev::FusedWork* work = s.work.as<dev::FusedWork>();
    const dev::FusedPass* const* tables = s.tables.as<const dev::FusedPass*>();
    for (size_t k = first; k < last; ++k) {
        const Launch& l = s.launches[k];
        if (l.groups == 0) continue;
        const dev::FusedLaunch fl{threads_, l.rpt, l.groups, vols_.as<dev::FusedVolume>(), work + l.work,
                                  tables + k * members_.size(),
                                  l.res == 4 || l.res == 3 ? l.lds4 : l.res ? l.lds_res : l.lds, sr_, l.res,
                                  l.res == 4 || l.res == 3 ? l.band4 : 0};
        if (ev) ETN_HIP(hipEventRecord((*ev)[2 * k], stream));
        dev::launch_fused(vt_, vec_, cap_, exact_, mask_, over_, fl, stream);
        if (ev) ETN_HIP(hipEventRecord((*ev)[2 * k + 1], stream));
    }
}

void FusedEngineImpl::enqueueIteration(hipStream_t stream) { enqueue(sched_, 0, sched_.launches.size(), stream, nullptr); }

double FusedEngineImpl::residualNorm(size_t member, double* scratch, hipStream_t stream)
{
    if (member >= members_.size()) fail("member index out of range");
    const MemberPlan& m = members_[member];
    BlockIterativeReconstructionOperator& op = *m.op;
    const std::vector<uint32_t>& measured = op.measuredProjections();
    const size_t M = members_.size();
    const int max_fp = std::max(1, std::min(dev::kMaxFusedFp, threads_ * dev::kMaxRaysPerThread /
                                                                  ((m.tile.planes * m.rows_per_plane * m.nu)
                                                                   << m.fp_shift)));
    // virtual projection and ray length images for the projections of one chunk
    DeviceBuffer images(sizeof(float) * m.pixels * 2 * (size_t)max_fp);
    auto slot = [&](int k, int which) { return images.as<float>() + m.pixels * (2 * (size_t)k + which); };
    // chunks of max_fp measured projections: forward-only passes by z direction
    std::vector<dev::FusedPass> passes;
    std::vector<size_t> chunk_end;     // passes of chunk c end at chunk_end[c]
    for (size_t c0 = 0; c0 < measured.size(); c0 += (size_t)max_fp) {
        dev::FusedPass neg{}, pos{};
        neg.zdir = -1;
        pos.zdir = 1;
        for (size_t k = c0; k < std::min(measured.size(), c0 + (size_t)max_fp); ++k) {
            const ParallelScannerGeometry& g = op.geometry(measured[k]);
            dev::FusedFp f{};
            g.source_base().store(f.src);
            g.horizontal_pitch().store(f.h);
            g.vertical_pitch().store(f.v);
            g.ray_direction().store(f.ray);
            f.measured = op.realProjection(measured[k]);
            f.measured_type = op.projectionType();
            f.residual = slot((int)(k - c0), 0);
            f.length = slot((int)(k - c0), 1);
            f.raw = 1;
            dev::FusedPass& p = f.ray[2] > 0.0f ? pos : neg;
            p.fp[p.nfp++] = f;
        }
        if (neg.nfp) passes.push_back(neg);
        if (pos.nfp) passes.push_back(pos);
        chunk_end.push_back(passes.size());
    }
    std::vector<dev::FusedWork> work;
    for (int j = 0; j < m.ny; j += m.tile.planes) work.push_back({(int)member, j});
    DeviceBuffer dpasses(sizeof(dev::FusedPass) * std::max<size_t>(passes.size(), 1));
    DeviceBuffer dwork(sizeof(dev::FusedWork) * work.size());
    DeviceBuffer dtables(sizeof(const dev::FusedPass*) * std::max<size_t>(passes.size() * M, 1));
    std::vector<const dev::FusedPass*> tables(passes.size() * M, nullptr);
    for (size_t k = 0; k < passes.size(); ++k) tables[k * M + member] = dpasses.as<dev::FusedPass>() + k;
    ETN_HIP(hipMemcpyAsync(dpasses.get(), passes.data(), sizeof(dev::FusedPass) * passes.size(),
                           hipMemcpyHostToDevice, stream));
    ETN_HIP(hipMemcpyAsync(dwork.get(), work.data(), dwork.bytes(), hipMemcpyHostToDevice, stream));
    ETN_HIP(hipMemcpyAsync(dtables.get(), tables.data(), sizeof(const dev::FusedPass*) * tables.size(),
                           hipMemcpyHostToDevice, stream));
    double* partials = scratch;
    double* total = scratch + dev::kSsdBlocks;
    ETN_HIP(hipMemsetAsync(total, 0, sizeof(double), stream));
    const size_t lds = sizeof(float) * (size_t)dev::fused_lds_floats(m.tile.planes, m.tile.band, m.nx, m.row_floats,
                                                                      mask_, threads_, m.rowsets, m.vbytes);
    size_t pass = 0;
    for (size_t c = 0; c < chunk_end.size(); ++c) {
        for (; pass < chunk_end[c]; ++pass) {
            const int rays = (passes[pass].nfp * m.tile.planes * m.rows_per_plane * m.nu) << m.fp_shift;
            const dev::FusedLaunch fl{threads_, select_rpt(rays, threads_), (int)work.size(),
                                      vols_.as<dev::FusedVolume>(), dwork.as<dev::FusedWork>(),
                                      dtables.as<const dev::FusedPass*>() + pass * M, lds};
            dev::launch_fused(vt_, vec_, cap_, exact_, mask_, over_, fl, stream);
        }
        // squared differences in projection order, as the unfused norm adds them
        const size_t c0 = c * (size_t)max_fp;
        for (size_t k = c0; k < std::min(measured.size(), c0 + (size_t)max_fp); ++k)
            dev::launch_sum_squared_difference(op.realProjection(measured[k]), op.projectionType(),
                                               slot((int)(k - c0), 0), m.pixels,
                                               partials, total, stream);
    }
    double host = 0.0;
    ETN_HIP(hipMemcpyAsync(&host, total, sizeof(double), hipMemcpyDeviceToHost, stream));
    ETN_HIP(hipStreamSynchronize(stream));
    return std::sqrt(host);
}

void FusedEngineImpl::check(hipStream_t stream)
{
    int bad = 0;
    ETN_HIP(hipMemcpyAsync(&bad, diag_.get(), sizeof(int), hipMemcpyDeviceToHost, stream));
    ETN_HIP(hipStreamSynchronize(stream));
    if (bad) throw Error(ETN_ERROR_INTERNAL, "fused plane kernel found rays outside their plane");
}

std::vector<FusedPlaneEngine::LaunchTime> FusedEngineImpl::timeIteration(hipStream_t stream)
{
    const size_t n = sched_.launches.size();
    std::vector<hipEvent_t> ev(2 * n);
    for (auto& e : ev) ETN_HIP(hipEventCreate(&e));
    enqueue(sched_, 0, n, stream, &ev);
    ETN_HIP(hipStreamSynchronize(stream));
    std::vector<LaunchTime> out;
    for (size_t k = 0; k < n; ++k) {
        if (sched_.launches[k].groups == 0) continue;
        float ms = 0.0f;
        ETN_HIP(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
        out.push_back({ms, sched_.launches[k].steady});
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    return out;
}

void FusedEngineImpl::setTile(size_t member, int planes, int band, hipStream_t stream)
{
    if (member >= members_.size()) fail("member index out of range");
    const Tile t{planes, band};
    if (!tile_fits(members_[member], t)) fail("tile does not fit this volume (planes must divide ny, LDS limit)");
    members_[member].tile = t;
    upload(stream);
}

void FusedEngineImpl::tile(size_t member, int* planes, int* band) const
{
    if (member >= members_.size()) fail("member index out of range");
    *planes = members_[member].tile.planes;
    *band = members_[member].tile.band;
}

// Every member keeps a fused form with `threads` per workgroup: its rays and
// a z row of one plane fit.
bool FusedEngineImpl::threadsFit(int threads) const
{
    if (threads != 256 && threads != 512) return false;
    for (const MemberPlan& m : members_)
        if (max_band(m, 1, threads) < 1) return false;
    return true;
}

// Switches every member to `threads`; tiles that no longer fit fall back to the default.
void FusedEngineImpl::useThreads(int threads)
{
    threads_ = threads;
    for (MemberPlan& m : members_) {
        m.threads = threads;
        if (!tile_fits(m, m.tile)) m.tile = default_tile(m);
    }
}

void FusedEngineImpl::useRptCap(int cap)
{
    for (MemberPlan& m : members_) m.rpt_cap = cap;
}

size_t FusedEngineImpl::passCount() const
{
    size_t n = 0;
    for (const MemberPlan& m : members_) n += m.passes.size();
    return n;
}

void FusedEngineImpl::setThreads(int threads, hipStream_t stream)
{
    if (!threadsFit(threads)) fail("workgroup size must be 256 or 512 threads and fit every member");
    useThreads(threads);
    upload(stream);
}

// Every launch of a schedule, after one untimed launch.
float FusedEngineImpl::timeSchedule(const Schedule& s, hipStream_t stream) const
{
    const size_t n = s.launches.size();
    if (n < 1) return -1.0f;
    hipEvent_t a, b;
    ETN_HIP(hipEventCreate(&a));
    ETN_HIP(hipEventCreate(&b));
    enqueue(s, 0, 1, stream, nullptr);
    ETN_HIP(hipEventRecord(a, stream));
    enqueue(s, 0, n, stream, nullptr);
    ETN_HIP(hipEventRecord(b, stream));
    ETN_HIP(hipEventSynchronize(b));
    float ms = 0.0f;
    ETN_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms;
}

// Members of one size share a tile; each size is timed on its own, under the
// current workgroup shape, over a whole iteration: the planes of a tile decide
// how many forward projections a pass takes (the rays per workgroup are
// capped), so the tiles cut an iteration into different passes, and a few
// launches of each would not compare equal work.
void FusedEngineImpl::tuneTiles(hipStream_t stream)
{
    // (the staged row length and row sets and the ray slots per pixel decide
    // which tiles fit as much as the sizes do)
    std::map<std::tuple<int, int, int, int, int, int, int, int, int>, std::vector<size_t>> classes;
    for (size_t i = 0; i < members_.size(); ++i) {
        const MemberPlan& m = members_[i];
        classes[{m.nx, m.ny, m.nz, m.nu, m.nv, m.rows_per_plane, m.row_floats, m.rowsets, m.fp_shift}].push_back(i);
    }
    for (auto& kv : classes) {
        const std::vector<size_t>& ids = kv.second;
        const std::vector<Tile> cands = candidate_tiles(members_[ids[0]]);
        if (cands.size() <= 1) continue;
        float best_ms = -1.0f;
        Tile best = members_[ids[0]].tile;
        for (const Tile& t : cands) {
            for (size_t i : ids) members_[i].tile = t;
            upload(stream);
            for (size_t i : ids) members_[i].offset = 0;
            Schedule s;
            build_schedule(ids, s, stream);
            const float ms = timeSchedule(s, stream);
            if (ms >= 0.0f && (best_ms < 0.0f || ms < best_ms)) {
                best_ms = ms;
                best = t;
            }
        }
        for (size_t i : ids) members_[i].tile = best;
    }
}

// For each workgroup shape that fits every member: tune the tiles, then time
// the engine's whole schedule; keep the fastest shape with its tiles.
void FusedEngineImpl::autotune(hipStream_t stream)
{
    // Candidates: workgroup shape x cap on the rays per thread.  A lower cap
    // needs fewer registers per thread (more waves per SIMD: two rays take
    // about 120 VGPRs, eight about 190) but splits a block's forward
    // projections over more passes, i.e. more sweeps of the volume; a cap
    // that leaves the passes unchanged is not a candidate of its own.
    struct Candidate {
        int threads, cap;
    };
    std::vector<Candidate> cands;
    for (int t : {512, 256}) {
        if (!threadsFit(t)) continue;
        useThreads(t);
        size_t passes_at_max = 0;
        for (int c : {dev::kMaxRaysPerThread, 4, 2}) {
            useRptCap(c);
            for (MemberPlan& m : members_) {
                m.tile = default_tile(m);
                build_passes(m);
            }
            const size_t n = passCount();
            if (c == dev::kMaxRaysPerThread) passes_at_max = n;
            else if (n == passes_at_max) continue;
            cands.push_back({t, c});
        }
    }
    bool any = cands.size() > 1;
    for (const MemberPlan& m : members_) any = any || candidate_tiles(m).size() > 1;
    if (!any) {
        useThreads(cands.empty() ? threads_ : cands[0].threads);
        useRptCap(cands.empty() ? dev::kMaxRaysPerThread : cands[0].cap);
        upload(stream);
        return;
    }
    // save the volumes: the trial passes update them
    std::vector<size_t> bytes(members_.size());
    size_t total = 0;
    for (size_t i = 0; i < members_.size(); ++i) {
        bytes[i] = members_[i].op->getReconstructedVolume()->bytes();
        total += bytes[i];
    }
    DeviceBuffer saved(total);
    size_t at = 0;
    for (size_t i = 0; i < members_.size(); ++i) {
        ETN_HIP(hipMemcpyAsync(saved.as<char>() + at, members_[i].op->getReconstructedVolume()->data(), bytes[i],
                               hipMemcpyDeviceToDevice, stream));
        at += bytes[i];
    }
    // the candidates split an iteration into different passes, so each is
    // timed over a whole iteration
    Candidate best = cands[0];
    std::vector<Tile> best_tiles;
    float best_ms = -1.0f;
    for (const Candidate& c : cands) {
        useThreads(c.threads);
        useRptCap(c.cap);
        for (MemberPlan& m : members_) m.tile = default_tile(m);
        tuneTiles(stream);
        upload(stream);
        const float ms = cands.size() > 1 ? timeSchedule(sched_, stream) : 0.0f;
        if (best_tiles.empty() || (ms >= 0.0f && ms < best_ms)) {
            best_ms = ms;
            best = c;
            best_tiles.clear();
            for (const MemberPlan& m : members_) best_tiles.push_back(m.tile);
        }
    }
    useThreads(best.threads);
    useRptCap(best.cap);
    for (size_t i = 0; i < members_.size(); ++i) members_[i].tile = best_tiles[i];
    at = 0;
    for (size_t i = 0; i < members_.size(); ++i) {
        ETN_HIP(hipMemcpyAsync(members_[i].op->getReconstructedVolume()->data(), saved.as<char>() + at, bytes[i],
                               hipMemcpyDeviceToDevice, stream));
        at += bytes[i];
    }
    upload(stream);
    check(stream);
}

}  // namespace

namespace {

class FusedBackProjectorImpl final : public FusedBackProjector {
public:
    FusedBackProjectorImpl(GpuVolume* volume, const std::vector<ParallelScannerGeometry>& geometries, MemberPlan plan)
        : volume_(volume), geometries_(geometries), m_(std::move(plan))
    {
        diag_.allocate(sizeof(int));
        ETN_HIP(hipMemset(diag_.get(), 0, sizeof(int)));
    }
    void backProject(const std::vector<uint32_t>& indices, const std::vector<const float*>& residual,
                     const std::vector<const float*>& length, float lambda, hipStream_t stream,
                     std::vector<float>* launch_ms) override;

    GpuVolume* volume_;
    std::vector<ParallelScannerGeometry> geometries_;
    MemberPlan m_;
    DeviceBuffer diag_;
};

void FusedBackProjectorImpl::backProject(const std::vector<uint32_t>& indices,
                                         const std::vector<const float*>& residual,
                                         const std::vector<const float*>& length, float lambda, hipStream_t stream,
                                         std::vector<float>* launch_ms)
{
    if (residual.size() != indices.size() || length.size() != indices.size())
        fail("fused back projector: one residual and one length image per projection");
    const VolumeProperties vp = whole_volume(volume_->properties());
    // passes of at most kMaxFusedBp projections, as even as the count allows
    // (121 = 15 x 8 + 1 would spend a whole volume sweep on the last one)
    const size_t npass = (indices.size() + dev::kMaxFusedBp - 1) / dev::kMaxFusedBp;
    std::vector<dev::FusedPass> passes;
    size_t pass_end = 0;
    for (size_t k = 0; k < indices.size(); ++k) {
        if (indices[k] >= geometries_.size()) fail("fused back projector: projection index out of range");
        if (k == pass_end) {
            // pass p holds floor(n / npass) projections, one more for the first n mod npass
            const size_t p = passes.size();
            pass_end += indices.size() / npass + (p < indices.size() % npass ? 1 : 0);
            passes.push_back(dev::FusedPass{});
            passes.back().zdir = 1;
        }
        const ParallelScannerGeometry& g = geometries_[indices[k]];
        dev::FusedBp e{};
        std::memcpy(e.m, g.projection_matrix().data(), sizeof(e.m));
        e.lambda = lambda;
        e.residual = residual[k];
        e.length = length[k];
        const std::vector<VoxelSample> smp = bp_sample_pattern(g, vp.voxel_size(), (uint32_t)m_.bp_samples);
        e.sample = make_float4(smp[0].dx, smp[0].dy, smp[0].dz, smp[0].length);
        e.nsamples = (int)smp.size();
        e.per_dir = m_.rowsets;
        for (size_t t = 0; t < smp.size(); ++t) e.samples[t] = make_float4(smp[t].dx, smp[t].dy, smp[t].dz, smp[t].length);
        dev::FusedPass& p = passes.back();
        p.bp[p.nbp++] = e;
    }
    if (passes.empty()) return;
    dev::FusedVolume fv{};
    fv.data = volume_->data();
    fv.slab = make_slab_params(vp, 0);
    fv.nx = m_.nx;
    fv.ny = m_.ny;
    fv.nz = m_.nz;
    fv.nu = m_.nu;
    fv.nv = m_.nv;
    fv.planes = m_.tile.planes;
    fv.band = m_.tile.band;
    fv.rows_per_plane = m_.rows_per_plane;
    fv.row_len = m_.row_len;
    fv.pad_l = m_.pad_l;
    fv.fp_n = 1;
    fv.rowsets = m_.rowsets;
    fv.diag = diag_.as<int>();
    std::vector<dev::FusedWork> work;
    for (int j = 0; j < m_.ny; j += m_.tile.planes) work.push_back({0, j});
    DeviceBuffer dvol(sizeof(dev::FusedVolume)), dpasses(sizeof(dev::FusedPass) * passes.size());
    DeviceBuffer dwork(sizeof(dev::FusedWork) * work.size()), dtables(sizeof(const dev::FusedPass*) * passes.size());
    const size_t lds = sizeof(float) * (size_t)dev::fused_lds_floats(m_.tile.planes, m_.tile.band, m_.nx,
                                                                      m_.row_floats, false, m_.threads, m_.rowsets,
                                                                      m_.vbytes);
    // pass k's launch (resident: only count the workgroups the device holds)
    auto pass_launch = [&](size_t k, int* resident) {
        // register back projection: every record set of the pass staged at once
        const size_t lds_res = sizeof(float) * (size_t)dev::fused_lds_floats(
                                   m_.tile.planes, m_.tile.band, m_.nx, m_.row_floats, false, m_.threads,
                                   m_.rowsets * std::max(passes[k].nbp, 1), m_.vbytes);
        int res = 0;
        const int band4 = m_.threads == 512 ? res2_band(m_, std::max(passes[k].nbp, 1), false) : 0;
        if (!m_.lds_bp && !m_.oversampled() && passes[k].nbp > 1 && m_.threads % (m_.tile.planes * m_.nx / 4) == 0)
            res = lds_res <= sizeof(float) * (size_t)dev::shape_lds_floats(m_.threads, m_.vbytes) ? 1
                  : band4 > 0                                                                      ? 4
                  : m_.threads == 512 && lds_res <= sizeof(float) * (size_t)dev::kResLds           ? 2
                                                                                                     : 0;
        const size_t used = res == 4 ? res2_lds_bytes(m_, band4, std::max(passes[k].nbp, 1), false)
                            : res    ? lds_res
                                     : lds;
        const dev::FusedLaunch fl{m_.threads, 1, (int)work.size(), dvol.as<dev::FusedVolume>(),
                                  dwork.as<dev::FusedWork>(), dtables.as<const dev::FusedPass*>() + k,
                                  used, false, res, res == 4 ? band4 : 0, resident};
        dev::launch_fused(volume_->voxel_type(), true, false, false, false, m_.oversampled(), fl, stream);
    };
    // the dispatch order the first pass's kernel (the many-projection one) favours
    int resident = 0;
    if (!passes.empty()) pass_launch(0, &resident);
    interleave_rounds(work.data(), (int)work.size(), resident);
    std::vector<const dev::FusedPass*> tables(passes.size());
    for (size_t k = 0; k < passes.size(); ++k) tables[k] = dpasses.as<dev::FusedPass>() + k;
    ETN_HIP(hipMemcpyAsync(dvol.get(), &fv, sizeof(fv), hipMemcpyHostToDevice, stream));
    ETN_HIP(hipMemcpyAsync(dpasses.get(), passes.data(), dpasses.bytes(), hipMemcpyHostToDevice, stream));
    ETN_HIP(hipMemcpyAsync(dwork.get(), work.data(), dwork.bytes(), hipMemcpyHostToDevice, stream));
    ETN_HIP(hipMemcpyAsync(dtables.get(), tables.data(), dtables.bytes(), hipMemcpyHostToDevice, stream));
    std::vector<hipEvent_t> ev(launch_ms ? 2 * passes.size() : 0);
    for (auto& e : ev) ETN_HIP(hipEventCreate(&e));
    for (size_t k = 0; k < passes.size(); ++k) {
        if (launch_ms) ETN_HIP(hipEventRecord(ev[2 * k], stream));
        pass_launch(k, nullptr);
        if (launch_ms) ETN_HIP(hipEventRecord(ev[2 * k + 1], stream));
    }
    ETN_HIP(hipStreamSynchronize(stream));   // the tables above are freed on return
    for (size_t k = 0; launch_ms && k < passes.size(); ++k) {
        float ms = 0.0f;
        ETN_HIP(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
        if (passes[k].nbp == dev::kMaxFusedBp) launch_ms->push_back(ms);
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
}

}  // namespace

std::unique_ptr<FusedBackProjector> FusedBackProjector::create(GpuVolume* volume,
                                                               const std::vector<ParallelScannerGeometry>& geometries,
                                                               uint32_t nu, uint32_t nv, uint32_t bp_samples,
                                                               bool lds_bp, hipStream_t)
{
    const VolumeProperties vp = whole_volume(volume->properties());
    MemberPlan m;
    m.nx = (int)vp.nx();
    m.ny = (int)vp.ny();
    m.nz = (int)vp.nz();
    m.nu = (int)nu;
    m.nv = (int)nv;
    if (m.nx % 4 || geometries.empty()) return nullptr;
    const int bp_s = (int)(bp_samples ? bp_samples : 1);
    const int pd = (int)std::lround(std::sqrt((double)bp_s));
    if (bp_s > dev::kMaxFusedBpSamples || pd * pd != bp_s) return nullptr;
    std::vector<const ParallelScannerGeometry*> geoms;
    for (const ParallelScannerGeometry& g : geometries) {
        // the v coordinate of a voxel depends on its y only, w == 1
        if (g.ray_direction().y != 0.0f || g.horizontal_pitch().y != 0.0f) return nullptr;
        const Matrix4x4 pm = g.projection_matrix();
        const float* mv = pm.data();
        if (mv[1] != 0.0f || mv[9] != 0.0f || mv[3] != 0.0f || mv[7] != 0.0f || mv[11] != 0.0f || mv[15] != 1.0f)
            return nullptr;
        geoms.push_back(&g);
    }
    if (!fast_row_span(vp, geoms, m.nu, bp_s, m.pad_l, m.row_len)) return nullptr;
    m.pixels = (size_t)nu * nv;
    m.rows_per_plane = 1;
    m.row_floats = m.row_len * 4;
    m.threads = dev::kMaxThreads;
    m.vbytes = (int)voxel_bytes(volume->voxel_type());
    m.bp_samples = bp_s;
    m.rowsets = pd;
    m.lds_bp = lds_bp;
    if (dev::fused_lds_floats(1, 1, m.nx, m.row_floats, false, m.threads, m.rowsets, m.vbytes) >
        dev::shape_lds_floats(m.threads, m.vbytes) || max_band(m, 1) < 1)
        return nullptr;
    m.tile = default_tile(m);
    return std::make_unique<FusedBackProjectorImpl>(volume, geometries, std::move(m));
}

std::unique_ptr<FusedPlaneEngine> FusedPlaneEngine::create(const std::vector<Member>& members, bool exact,
                                                           hipStream_t stream, std::vector<bool>* taken)
{
    taken->assign(members.size(), false);
    std::vector<MemberPlan> plans;
    etn_voxel_type vt = ETN_VOXEL_FLOAT32;
    bool vec = true, cap = false, mask = false, over = false, sr = false;
    for (size_t i = 0; i < members.size(); ++i) {
        const etn_reconstruction_desc& d = *members[i].desc;
        const bool d_over = (d.fp_samples ? d.fp_samples : 1) != 1 || (d.bp_samples ? d.bp_samples : 1) != 1;
        const bool d_sr = (d.flags & ETN_RECON_FLAG_STOCHASTIC_ROUNDING) != 0;
        const bool d_lds = (d.flags & ETN_RECON_FLAG_LDS_BACK_PROJECTION) != 0;
        if (!plans.empty() && (d.voxel_type != vt || (d.mask != nullptr) != mask || d_over != over || d_sr != sr ||
                               d_lds != plans[0].lds_bp ||
                               (members[i].op->getReconstructedVolume()->properties().nx() % 4 == 0) != vec))
            continue;
        // stochastic rounding exists on the fast 16-bit kernels without mask or oversampling
        if (d_sr && (exact || d.voxel_type == ETN_VOXEL_FLOAT32 || d.mask != nullptr || d_over ||
                     members[i].op->getReconstructedVolume()->properties().nx() % 4 != 0))
            continue;
        MemberPlan m;
        if (!plan_member(*members[i].op, d, exact, m)) continue;
        const bool m_cap = step_limit_can_bind(*members[i].op);
        if (d_sr && m_cap) continue;
        m.sr = d_sr;
        m.lds_bp = d_lds;
        if (plans.empty()) {
            vt = d.voxel_type;
            vec = m.nx % 4 == 0;
            mask = m.mask != nullptr;
            over = m.oversampled();
            sr = m.sr;
        }
        cap = cap || m_cap;
        plans.push_back(std::move(m));
        (*taken)[i] = true;
    }
    if (plans.empty()) return nullptr;
    return std::make_unique<FusedEngineImpl>(std::move(plans), vt, vec, cap, exact, mask, over, stream);
}

}  // namespace etn
