// Learned correction network kernels (see unet_kernels.hpp).
//
// The 3x3x3 convolution is an implicit GEMM on the matrix cores: output
// voxels are the GEMM rows, output channels the columns, and the reduction
// runs over k = tap * cin + ci.  Each wave owns MT tiles of 16 voxels and all
// cout channels (NT tiles of 16); per 32-wide k step a lane loads its 8
// channels of one tap for each voxel tile (one 16-byte load from the
// channels-last activation, or zeros outside the volume) and 8 weights per
// channel tile, then issues MT x NT v_mfma_f32_16x16x32_bf16.  Lane layouts
// (checked on the device): A[i][k] in lane i + 16 (k / 8), element k % 8;
// B[k][j] in lane j + 16 (k / 8); D[i][j] in lane j + 16 (i / 4), element i % 4.
//
// The weight gradient is the GEMM dW[co][(tap, ci)] = sum_m gy[m][co] x[m + d(tap)][ci]
// with the voxels as the reduction: a workgroup takes one tap and a chunk of
// voxels, stages gy and the tap-shifted input transposed in LDS (32 voxels at
// a time) and accumulates all (co, ci) tiles; the per-chunk partials are then
// summed in a fixed order, so every gradient is deterministic.
#include <algorithm>

#include "common.hpp"
#include "device_math.hpp"
#include "unet_kernels.hpp"

namespace etn {
namespace nn {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) { return dev::BF16::pack(f); }

__device__ __forceinline__ u32x4 load16(const uint16_t* p) { return *reinterpret_cast<const u32x4*>(p); }
__device__ __forceinline__ bf16x8 as_bf16x8(u32x4 v) { return __builtin_bit_cast(bf16x8, v); }

constexpr int kConvWaves = 4;
constexpr int kThreads = 256;

// ---- 3x3x3 convolution, implicit GEMM ---------------------------------------
template <int NT, int MT>
__global__ void __launch_bounds__(kThreads) conv3_fwd_kernel(ConvArgs a)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 15, kg = lane >> 4;
    const size_t M = (size_t)a.D * a.H * a.W;
    const size_t m0 = ((size_t)blockIdx.x * kConvWaves + wave) * (MT * 16);
    if (m0 >= M) return;   // the whole wave is past the volume (no barriers below)
    const int HW = a.H * a.W;
    int vz[MT], vy[MT], vx[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        const size_t m = m0 + mt * 16 + row;
        if (m < M) {
            vz[mt] = (int)(m / HW);
            const int r = (int)(m - (size_t)vz[mt] * HW);
            vy[mt] = r / a.W;
            vx[mt] = r - vy[mt] * a.W;
        } else {
            vz[mt] = -4;   // every tap falls outside
            vy[mt] = 0;
            vx[mt] = 0;
        }
    }
    floatx4 acc[MT][NT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};

    const int ca = a.src.ca, cin = a.src.ca + a.src.cb;
    const int Hh = a.H >> 1, Wh = a.W >> 1;
    const bf16x8 zero = as_bf16x8(u32x4{0u, 0u, 0u, 0u});
    for (int k0 = kg * 8; k0 < a.kpad; k0 += 32) {
        bf16x8 B[NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) B[nt] = as_bf16x8(load16(a.wt + (size_t)(nt * 16 + row) * a.kpad + k0));
        bf16x8 A[MT];
        {
            const int t = k0 / cin, ci = k0 - t * cin;
            const int dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                const int zz = vz[mt] + dz, yy = vy[mt] + dy, xx = vx[mt] + dx;
                A[mt] = zero;
                if (t < 27 && (unsigned)zz < (unsigned)a.D && (unsigned)yy < (unsigned)a.H &&
                    (unsigned)xx < (unsigned)a.W) {
                    const uint16_t* p;
                    if (ci < ca) {
                        const size_t v = a.src.a_half ? ((size_t)(zz >> 1) * Hh + (yy >> 1)) * Wh + (xx >> 1)
                                                      : ((size_t)zz * a.H + yy) * a.W + xx;
                        p = a.src.a + v * ca + ci;
                    } else {
                        p = a.src.b + (((size_t)zz * a.H + yy) * a.W + xx) * a.src.cb + (ci - ca);
                    }
                    A[mt] = as_bf16x8(load16(p));
                }
            }
        }
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[mt], B[nt], acc[mt][nt], 0, 0, 0);
    }
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int co = nt * 16 + row;
            const float b = a.bias ? a.bias[co] : 0.0f;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const size_t m = m0 + mt * 16 + kg * 4 + r;
                if (m < M) {
                    float v = acc[mt][nt][r] + b;
                    if (a.relu) v = fmaxf(v, 0.0f);
                    a.out[m * a.cout + co] = f2bf(v);
                }
            }
        }
}

// ---- 3x3x3 convolution to 16 or 32 channels, LDS-staged bricks ----------
// A workgroup computes a kBZ x kBY x kBX brick of output voxels, all
// NT x 16 output channels.  Its input halo (kBZ + 2) x (kBY + 2) x (kBX + 2)
// is staged in LDS eight input channels at a time (one 16-byte cell per halo
// voxel), so that the 27 taps read each staged voxel from LDS instead of from
// memory.  An MFMA k step takes four taps of the chunk: lane group kg's A
// operand is its voxel's 8 channels at tap 4 s + kg, and its B operand the 8
// weights wt[co][tap * cin + ci ..] of the same tap, contiguous in the forward
// weight layout (the k order inside a step only has to agree between A and
// B).  The next chunk's weights and halo are loaded into registers while the
// current chunk's MFMAs run.  Wave w owns the brick's output rows
// (z = w, y = 0..7), two 16-voxel tiles each: the LDS reads of a step share one
// base address and differ by constant offsets.  The result goes through LDS
// (one 16-channel slice at a time) so that each voxel's channels leave in
// 16-byte stores.
constexpr int kBZ = kConvWaves, kBY = 8, kBX = 32;
constexpr int kHZ = kBZ + 2, kHY = kBY + 2, kHX = kBX + 2;
constexpr int kHaloCells = kHZ * kHY * kHX;                    // 2040
constexpr int kBrickVoxels = kBZ * kBY * kBX;                  // 1024
constexpr int kBrickTiles = kBrickVoxels / 16 / kConvWaves;    // 16 tiles of 16 voxels per wave
constexpr int kStage = (kHaloCells + kThreads - 1) / kThreads; // halo cells per thread
constexpr int kChunkSteps = (27 + 3) / 4;                      // k steps per 8-channel chunk
constexpr int kBrickLds = 2 * kBrickVoxels;                    // cells: >= the halo, = one output slice

template <int NT>
__global__ void __launch_bounds__(kThreads) conv3_brick_kernel(ConvArgs a, int nbx, int nby)
{
    static_assert(kBrickLds >= kHaloCells, "the output slice reuses the halo's LDS");
    __shared__ u32x4 lds[kBrickLds];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int row = lane & 15, kg = lane >> 4;
    const int bxi = blockIdx.x % nbx, rest = blockIdx.x / nbx;
    const int byi = rest % nby, bzi = rest / nby;
    const int x0 = bxi * kBX, y0 = byi * kBY, z0 = bzi * kBZ;
    const int ca = a.src.ca, cin = a.src.ca + a.src.cb;
    const int Hh = a.H >> 1, Wh = a.W >> 1;
    const u32x4 zero4 = u32x4{0u, 0u, 0u, 0u};

    floatx4 acc[kBrickTiles][NT];
#pragma unroll
    for (int t = 0; t < kBrickTiles; ++t)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[t][nt] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};

    // one chunk's weights (step st, tap 4 st + kg, output channel nt * 16 + row)
    // and halo cells, loaded into registers; the next chunk's loads are in
    // flight while the current chunk's MFMAs run
    u32x4 B[kChunkSteps][NT], cell[kStage];
    auto load_chunk = [&](int c0) {
#pragma unroll
        for (int st = 0; st < kChunkSteps; ++st) {
            const int tap = 4 * st + kg;
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                B[st][nt] = tap < 27 ? load16(a.wt + (size_t)(nt * 16 + row) * a.kpad + tap * cin + c0) : zero4;
        }
#pragma unroll
        for (int k = 0; k < kStage; ++k) {
            const int e = k * kThreads + tid;
            cell[k] = zero4;
            if (e < kHaloCells) {
                const int hz = e / (kHY * kHX), r2 = e - hz * (kHY * kHX), hy = r2 / kHX, hx = r2 - hy * kHX;
                const int zz = z0 - 1 + hz, yy = y0 - 1 + hy, xx = x0 - 1 + hx;
                if ((unsigned)zz < (unsigned)a.D && (unsigned)yy < (unsigned)a.H && (unsigned)xx < (unsigned)a.W) {
                    const uint16_t* p;
                    if (c0 < ca) {
                        const size_t vi = a.src.a_half ? ((size_t)(zz >> 1) * Hh + (yy >> 1)) * Wh + (xx >> 1)
                                                       : ((size_t)zz * a.H + yy) * a.W + xx;
                        p = a.src.a + vi * ca + c0;
                    } else {
                        p = a.src.b + (((size_t)zz * a.H + yy) * a.W + xx) * a.src.cb + (c0 - ca);
                    }
                    cell[k] = load16(p);
                }
            }
        }
    };
    load_chunk(0);
    for (int c0 = 0; c0 < cin; c0 += 8) {
        __syncthreads();   // the previous chunk's reads of lds are done
#pragma unroll
        for (int k = 0; k < kStage; ++k) {
            const int e = k * kThreads + tid;
            if (e < kHaloCells) lds[e] = cell[k];
        }
        u32x4 Bc[kChunkSteps][NT];
#pragma unroll
        for (int st = 0; st < kChunkSteps; ++st)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) Bc[st][nt] = B[st][nt];
        __syncthreads();
        if (c0 + 8 < cin) load_chunk(c0 + 8);
#pragma unroll
        for (int st = 0; st < kChunkSteps; ++st) {
            const int tap = 4 * st + kg;
            const bool live = tap < 27;
            const int tc = live ? tap : 0;   // a valid address for the padding taps (B is zero there)
            const int dz = tc / 9, dy = (tc / 3) % 3, dx = tc % 3;
            const u32x4* base = lds + ((wave + dz) * kHY + dy) * kHX + dx + row;
#pragma unroll
            for (int t = 0; t < kBrickTiles; ++t) {
                u32x4 av = base[(t >> 1) * kHX + (t & 1) * 16];
                if (!live) av = zero4;
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    acc[t][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(av), as_bf16x8(Bc[st][nt]),
                                                                         acc[t][nt], 0, 0, 0);
            }
        }
    }
    // epilogue, one 16-channel slice at a time: bias, activation, bf16 into
    // LDS as [brick voxel][16 channels], then 16-byte stores of 8 channels
    uint16_t* slice = reinterpret_cast<uint16_t*>(lds);
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        const int co = nt * 16 + row;
        const float b = a.bias ? a.bias[co] : 0.0f;
        __syncthreads();   // the last chunk's reads (or the previous slice's copy) are done
#pragma unroll
        for (int t = 0; t < kBrickTiles; ++t) {
            const int v0 = (wave * kBY + (t >> 1)) * kBX + (t & 1) * 16 + kg * 4;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                float v = acc[t][nt][e] + b;
                if (a.relu) v = fmaxf(v, 0.0f);
                slice[(v0 + e) * 16 + row] = f2bf(v);
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kBrickLds / kThreads; ++k) {
            const int c = k * kThreads + tid;         // cell: voxel c / 2, channels 8 (c % 2) ..
            const int bv = c >> 1, bz = bv / (kBY * kBX), by = (bv / kBX) % kBY, bx = bv % kBX;
            const int z = z0 + bz, y = y0 + by, x = x0 + bx;
            if (z < a.D && y < a.H && x < a.W)
                *reinterpret_cast<u32x4*>(a.out + (((size_t)z * a.H + y) * a.W + x) * a.cout + nt * 16 + (c & 1) * 8) =
                    lds[c];
        }
    }
}

template <int NT, int MT>
void launch_conv(const ConvArgs& a, hipStream_t stream)
{
    const size_t M = (size_t)a.D * a.H * a.W;
    const size_t per_block = (size_t)kConvWaves * MT * 16;
    const dim3 grid((unsigned)((M + per_block - 1) / per_block));
    conv3_fwd_kernel<NT, MT><<<grid, kThreads, 0, stream>>>(a);
}

// ---- weight gradient ----------------------------------------------------------
constexpr int kLdsRow = 40;        // 32 voxels + 8 pad (bf16), 80-byte rows keep 16-byte alignment
constexpr int kMaxCout = 64;
constexpr int kMaxCin = 96;

struct WgradArgs {
    ConvSrc src;
    int D, H, W;
    const uint16_t* gy;
    int cout;
    int chunk;          // voxels per workgroup, multiple of 32
    float* partial;     // [chunks][27][cout][cin16]
};

__global__ void __launch_bounds__(kThreads) conv3_wgrad_kernel(WgradArgs w)
{
    __shared__ __attribute__((aligned(16))) uint16_t gyT[kMaxCout * kLdsRow];
    __shared__ __attribute__((aligned(16))) uint16_t xT[kMaxCin * kLdsRow];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 15, kg = lane >> 4;
    const int t = blockIdx.y;
    const int dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
    const size_t M = (size_t)w.D * w.H * w.W;
    const int HW = w.H * w.W;
    const int ca = w.src.ca, cin = w.src.ca + w.src.cb;
    const int cin16 = (cin + 15) / 16 * 16;
    const int nco = w.cout / 16, nci = cin16 / 16, ntiles = nco * nci;
    const int Hh = w.H >> 1, Wh = w.W >> 1;
    const size_t mbeg = (size_t)blockIdx.x * w.chunk;
    floatx4 acc[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};

    for (int s = 0; s < w.chunk; s += 32) {
        const size_t ms = mbeg + s;
        // gy rows of 32 voxels, transposed: gyT[co][voxel]
        const int gpv = w.cout / 8;
        for (int g = threadIdx.x; g < 32 * gpv; g += kThreads) {
            const int mi = g / gpv, c0 = (g - mi * gpv) * 8;
            const size_t m = ms + mi;
            u32x4 v = u32x4{0u, 0u, 0u, 0u};
            if (m < M) v = load16(w.gy + m * w.cout + c0);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                gyT[(c0 + 2 * e) * kLdsRow + mi] = (uint16_t)(v[e] & 0xffffu);
                gyT[(c0 + 2 * e + 1) * kLdsRow + mi] = (uint16_t)(v[e] >> 16);
            }
        }
        // tap-shifted input, transposed: xT[ci][voxel]
        {
            const int gpx = cin / 8;
            for (int g = threadIdx.x; g < 32 * gpx; g += kThreads) {
                const int mi = g / gpx, c0 = (g - mi * gpx) * 8;
                const size_t m = ms + mi;
                u32x4 v = u32x4{0u, 0u, 0u, 0u};
                if (m < M) {
                    const int z = (int)(m / HW), r = (int)(m - (size_t)z * HW), y = r / w.W, x = r - y * w.W;
                    const int zz = z + dz, yy = y + dy, xx = x + dx;
                    if ((unsigned)zz < (unsigned)w.D && (unsigned)yy < (unsigned)w.H && (unsigned)xx < (unsigned)w.W) {
                        if (c0 < ca) {
                            const size_t vi = w.src.a_half ? ((size_t)(zz >> 1) * Hh + (yy >> 1)) * Wh + (xx >> 1)
                                                           : ((size_t)zz * w.H + yy) * w.W + xx;
                            v = load16(w.src.a + vi * ca + c0);
                        } else {
                            v = load16(w.src.b + (((size_t)zz * w.H + yy) * w.W + xx) * w.src.cb + (c0 - ca));
                        }
                    }
                }
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    xT[(c0 + 2 * e) * kLdsRow + mi] = (uint16_t)(v[e] & 0xffffu);
                    xT[(c0 + 2 * e + 1) * kLdsRow + mi] = (uint16_t)(v[e] >> 16);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int tile = wave + 4 * i;
            if (tile < ntiles) {
                const int ct = tile / nci, it = tile - ct * nci;
                const bf16x8 A = as_bf16x8(*reinterpret_cast<const u32x4*>(&gyT[(ct * 16 + row) * kLdsRow + kg * 8]));
                const bf16x8 B = as_bf16x8(*reinterpret_cast<const u32x4*>(&xT[(it * 16 + row) * kLdsRow + kg * 8]));
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[i], 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int tile = wave + 4 * i;
        if (tile < ntiles) {
            const int ct = tile / nci, it = tile - ct * nci;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int co = ct * 16 + kg * 4 + r, ci = it * 16 + row;
                w.partial[(((size_t)blockIdx.x * 27 + t) * w.cout + co) * cin16 + ci] = acc[i][r];
            }
        }
    }
}

__global__ void wgrad_reduce_kernel(const float* partial, int chunks, int cout, int cin, int cin16, float* dw)
{
    const int n = cout * 27 * cin;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int co = i / (27 * cin), r = i - co * 27 * cin, t = r / cin, ci = r - t * cin;
    const size_t stride = (size_t)27 * cout * cin16;
    const float* p = partial + ((size_t)t * cout + co) * cin16 + ci;
    float s = 0.0f;
    for (int c = 0; c < chunks; ++c) s += p[c * stride];
    dw[i] = s;
}

__global__ void channel_sum_partial_kernel(const uint16_t* g, size_t voxels, int C, int chunk, float* partial)
{
    __shared__ float red[kThreads];
    const int lanes = kThreads / C;   // voxel lanes per channel
    const int ch = threadIdx.x % C, j = threadIdx.x / C;
    const size_t beg = (size_t)blockIdx.x * chunk, end = std::min(voxels, beg + chunk);
    float s = 0.0f;
    if (j < lanes)
        for (size_t m = beg + j; m < end; m += lanes) s += bf2f(g[m * C + ch]);
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < C) {
        float t = 0.0f;
        for (int q = 0; q < lanes; ++q) t += red[q * C + threadIdx.x];
        partial[(size_t)blockIdx.x * C + threadIdx.x] = t;
    }
}

__global__ void sum_partials_kernel(const float* partial, int chunks, int C, float* out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    float s = 0.0f;
    for (int k = 0; k < chunks; ++k) s += partial[(size_t)k * C + c];
    out[c] = s;
}

__global__ void pack_weights_kernel(const float* w, int cout, int cin, int cin_pad, uint16_t* wt, int kpad,
                                    uint16_t* wb, int kpad_b)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cout * kpad) {
        const int co = i / kpad, k = i - co * kpad, t = k / cin_pad, ci = k - t * cin_pad;
        wt[i] = (t < 27 && ci < cin) ? f2bf(w[((size_t)co * 27 + t) * cin + ci]) : (uint16_t)0;
    }
    if (wb && i < cin * kpad_b) {
        const int ci = i / kpad_b, k = i - ci * kpad_b, t = k / cout, co = k - t * cout;
        wb[i] = t < 27 ? f2bf(w[((size_t)co * 27 + (26 - t)) * cin + ci]) : (uint16_t)0;
    }
}

__global__ void relu_backward_kernel(uint16_t* g, const uint16_t* y, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint16_t v = y[i];
        if ((v & 0x8000u) || (v & 0x7fffu) == 0) g[i] = 0;
    }
}

__global__ void pool2_kernel(const uint16_t* x, int D, int H, int W, int C, uint16_t* y)
{
    const int Dl = D / 2, Hl = H / 2, Wl = W / 2;
    const size_t n = (size_t)Dl * Hl * Wl * C;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const size_t p = i / C;
        const int xl = (int)(p % Wl), yl = (int)((p / Wl) % Hl), zl = (int)(p / ((size_t)Wl * Hl));
        float s = 0.0f;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int e = 0; e < 2; ++e)
                    s += bf2f(x[((((size_t)(2 * zl + a) * H) + 2 * yl + b) * W + 2 * xl + e) * C + c]);
        y[i] = f2bf(s * 0.125f);
    }
}

__global__ void pool2_backward_kernel(const uint16_t* gl, int D, int H, int W, int C, uint16_t* g, int accumulate)
{
    const size_t n = (size_t)D * H * W * C;
    const int Hl = H / 2, Wl = W / 2;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const size_t m = i / C;
        const int x = (int)(m % W), y = (int)((m / W) % H), z = (int)(m / ((size_t)W * H));
        const float v = bf2f(gl[(((size_t)(z / 2) * Hl + y / 2) * Wl + x / 2) * C + c]) * 0.125f;
        g[i] = f2bf(accumulate ? bf2f(g[i]) + v : v);
    }
}

__global__ void up2_backward_kernel(const uint16_t* g, int stride, int c0, int D, int H, int W, int C, uint16_t* gl)
{
    const int Dl = D / 2, Hl = H / 2, Wl = W / 2;
    const size_t n = (size_t)Dl * Hl * Wl * C;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        const size_t p = i / C;
        const int xl = (int)(p % Wl), yl = (int)((p / Wl) % Hl), zl = (int)(p / ((size_t)Wl * Hl));
        float s = 0.0f;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int e = 0; e < 2; ++e)
                    s += bf2f(g[((((size_t)(2 * zl + a) * H) + 2 * yl + b) * W + 2 * xl + e) * stride + c0 + c]);
        gl[i] = f2bf(s);
    }
}

__global__ void take_channels_kernel(const uint16_t* g, int stride, int c0, size_t voxels, int C, uint16_t* dst,
                                     int accumulate)
{
    const size_t n = voxels * C;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t m = i / C;
        const int c = (int)(i - m * C);
        const float v = bf2f(g[m * stride + c0 + c]);
        dst[i] = f2bf(accumulate ? bf2f(dst[i]) + v : v);
    }
}

__global__ void pack_input_kernel(const float* x, float scale, const float* d, float dscale, size_t n, uint16_t* y)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t a = f2bf(x[i] * scale), b = d ? f2bf(d[i] * dscale) : 0u;
        *reinterpret_cast<u32x4*>(y + 8 * i) = u32x4{a | (b << 16), 0u, 0u, 0u};
    }
}

__global__ void head_forward_kernel(const uint16_t* d1, const float* x, float scale, const float* hw, const float* hb,
                                    size_t n, float* pred)
{
    for (size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x; m < n; m += (size_t)gridDim.x * blockDim.x) {
        const u32x4 a = load16(d1 + m * 16), b = load16(d1 + m * 16 + 8);
        float s = 0.0f;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            s += hw[2 * e] * bf2f((uint16_t)(a[e] & 0xffffu)) + hw[2 * e + 1] * bf2f((uint16_t)(a[e] >> 16));
            s += hw[8 + 2 * e] * bf2f((uint16_t)(b[e] & 0xffffu)) + hw[9 + 2 * e] * bf2f((uint16_t)(b[e] >> 16));
        }
        pred[m] = x[m] * scale + (s + hb[0]);
    }
}

__global__ void unscale_kernel(const float* pred, const float* x, float scale, size_t n, float* out)
{
    for (size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x; m < n; m += (size_t)gridDim.x * blockDim.x)
        out[m] = x[m] + (pred[m] - x[m] * scale) / scale;
}

constexpr int kRedBlocks = 1024;

__global__ void mse_partial_kernel(const float* pred, const float* target, float scale, size_t n, float* g,
                                   double* partial)
{
    __shared__ double red[kThreads];
    double s = 0.0;
    const float inv = 2.0f / (float)n;
    for (size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x; m < n; m += (size_t)gridDim.x * blockDim.x) {
        const float r = pred[m] - target[m] * scale;
        g[m] = r * inv;
        s += (double)r * (double)r;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w /= 2) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void mse_final_kernel(const double* partial, int blocks, size_t n, double* loss)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int b = 0; b < blocks; ++b) s += partial[b];
        *loss = s / (double)n;
    }
}

__global__ void head_backward_kernel(const uint16_t* d1, const float* gp, const float* hw, size_t n, uint16_t* gd1,
                                     float* partial)
{
    __shared__ float red[kThreads * 17];
    float s[17];
#pragma unroll
    for (int c = 0; c < 17; ++c) s[c] = 0.0f;
    for (size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x; m < n; m += (size_t)gridDim.x * blockDim.x) {
        const float g = gp[m];
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            s[c] += g * bf2f(d1[m * 16 + c]);
            gd1[m * 16 + c] = f2bf(g * hw[c]);
        }
        s[16] += g;
    }
#pragma unroll
    for (int c = 0; c < 17; ++c) red[c * kThreads + threadIdx.x] = s[c];
    __syncthreads();
    if (threadIdx.x < 17) {
        float t = 0.0f;
        for (int q = 0; q < kThreads; ++q) t += red[threadIdx.x * kThreads + q];
        partial[(size_t)blockIdx.x * 17 + threadIdx.x] = t;
    }
}

__global__ void adam_kernel(float* p, const float* g, float* m, float* v, size_t n, float lr, float b1, float b2,
                            float eps, float bc1, float bc2)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    }
}

unsigned grid_for(size_t n)
{
    return (unsigned)std::max<size_t>(1, std::min<size_t>((n + kThreads - 1) / kThreads, 65536));
}

}  // namespace

void conv3_forward(const ConvArgs& a, hipStream_t stream)
{
    if (a.src.ca < 8 || (a.src.ca % 8) || (a.src.cb % 8)) fail("conv3: input channels must be multiples of 8");
    if (a.kpad != kpad_for(a.src.ca + a.src.cb)) fail("conv3: weight row length mismatch");
    if (a.src.a_half && ((a.D | a.H | a.W) & 1)) fail("conv3: upsampled input needs even sizes");
    const int nbx = (a.W + kBX - 1) / kBX, nby = (a.H + kBY - 1) / kBY, nbz = (a.D + kBZ - 1) / kBZ;
    const unsigned bricks = (unsigned)(nbx * nby * nbz);
    switch (a.cout) {
        case 16: conv3_brick_kernel<1><<<bricks, kThreads, 0, stream>>>(a, nbx, nby); break;
        case 32: conv3_brick_kernel<2><<<bricks, kThreads, 0, stream>>>(a, nbx, nby); break;
        case 48: launch_conv<3, 2>(a, stream); break;
        case 64: launch_conv<4, 2>(a, stream); break;
        case 96: launch_conv<6, 2>(a, stream); break;
        default: fail("conv3: output channels must be 16, 32, 48, 64 or 96");
    }
    hip_check(hipGetLastError(), "conv3 forward launch");
}

void conv3_weight_grad(const ConvSrc& src, int D, int H, int W, const uint16_t* gy, int cout, float* partial,
                       hipStream_t stream)
{
    if (cout % 16 || cout > kMaxCout || src.ca + src.cb > kMaxCin || src.ca % 8 || src.cb % 8)
        fail("conv3 weight gradient: channel counts");
    WgradArgs w{src, D, H, W, gy, cout, kWgradChunk, partial};
    const size_t M = (size_t)D * H * W;
    const dim3 grid((unsigned)wgrad_chunks(M), 27);
    conv3_wgrad_kernel<<<grid, kThreads, 0, stream>>>(w);
    hip_check(hipGetLastError(), "conv3 weight gradient launch");
}

void conv3_weight_grad_reduce(const float* partial, int chunks, int cout, int cin, float* dw, hipStream_t stream)
{
    const int n = cout * 27 * cin;
    const int cin16 = (cin + 15) / 16 * 16;
    wgrad_reduce_kernel<<<(n + kThreads - 1) / kThreads, kThreads, 0, stream>>>(partial, chunks, cout, cin, cin16, dw);
    hip_check(hipGetLastError(), "weight gradient reduction launch");
}

void channel_sum(const uint16_t* g, size_t voxels, int channels, float* partial, float* out, hipStream_t stream)
{
    if (channels > kThreads || kThreads % channels) fail("channel_sum: channel count");
    const int chunks = wgrad_chunks(voxels);
    channel_sum_partial_kernel<<<chunks, kThreads, 0, stream>>>(g, voxels, channels, kWgradChunk, partial);
    sum_partials_kernel<<<1, 128, 0, stream>>>(partial, chunks, channels, out);
    hip_check(hipGetLastError(), "channel sum launch");
}

void pack_weights(const float* w, int cout, int cin, int cin_pad, uint16_t* wt, uint16_t* wb, hipStream_t stream)
{
    const int kpad = kpad_for(cin_pad), kpad_b = kpad_for(cout);
    const int n = std::max(cout * kpad, wb ? cin * kpad_b : 0);
    pack_weights_kernel<<<(n + kThreads - 1) / kThreads, kThreads, 0, stream>>>(w, cout, cin, cin_pad, wt, kpad, wb,
                                                                               kpad_b);
    hip_check(hipGetLastError(), "weight packing launch");
}

void relu_backward(uint16_t* g, const uint16_t* y, size_t count, hipStream_t stream)
{
    relu_backward_kernel<<<grid_for(count), kThreads, 0, stream>>>(g, y, count);
}

void pool2(const uint16_t* x, int D, int H, int W, int C, uint16_t* y, hipStream_t stream)
{
    pool2_kernel<<<grid_for((size_t)D * H * W * C / 8), kThreads, 0, stream>>>(x, D, H, W, C, y);
}

void pool2_backward(const uint16_t* g_low, int D, int H, int W, int C, uint16_t* g_full, int accumulate,
                    hipStream_t stream)
{
    pool2_backward_kernel<<<grid_for((size_t)D * H * W * C), kThreads, 0, stream>>>(g_low, D, H, W, C, g_full,
                                                                                    accumulate);
}

void up2_backward(const uint16_t* g, int stride, int c0, int D, int H, int W, int C, uint16_t* g_low,
                  hipStream_t stream)
{
    up2_backward_kernel<<<grid_for((size_t)D * H * W * C / 8), kThreads, 0, stream>>>(g, stride, c0, D, H, W, C,
                                                                                      g_low);
}

void take_channels(const uint16_t* g, int stride, int c0, size_t voxels, int C, uint16_t* dst, int accumulate,
                   hipStream_t stream)
{
    take_channels_kernel<<<grid_for(voxels * C), kThreads, 0, stream>>>(g, stride, c0, voxels, C, dst, accumulate);
}

void pack_input(const float* x, float scale, const float* d, float dscale, size_t count, uint16_t* y,
                hipStream_t stream)
{
    pack_input_kernel<<<grid_for(count), kThreads, 0, stream>>>(x, scale, d, dscale, count, y);
}

void head_forward(const uint16_t* d1, const float* x, float scale, const float* hw, const float* hb, size_t voxels,
                  float* pred, hipStream_t stream)
{
    head_forward_kernel<<<grid_for(voxels), kThreads, 0, stream>>>(d1, x, scale, hw, hb, voxels, pred);
}

void unscale_prediction(const float* pred, const float* x, float scale, size_t voxels, float* out, hipStream_t stream)
{
    unscale_kernel<<<grid_for(voxels), kThreads, 0, stream>>>(pred, x, scale, voxels, out);
}

void mse_loss(const float* pred, const float* target, float scale, size_t voxels, float* g_pred, double* partial,
              double* loss, hipStream_t stream)
{
    mse_partial_kernel<<<kRedBlocks, kThreads, 0, stream>>>(pred, target, scale, voxels, g_pred, partial);
    mse_final_kernel<<<1, 64, 0, stream>>>(partial, kRedBlocks, voxels, loss);
}

void head_backward(const uint16_t* d1, const float* g_pred, const float* hw, size_t voxels, uint16_t* g_d1,
                   float* partial, float* ghw_ghb, hipStream_t stream)
{
    head_backward_kernel<<<kRedBlocks, kThreads, 0, stream>>>(d1, g_pred, hw, voxels, g_d1, partial);
    sum_partials_kernel<<<1, 64, 0, stream>>>(partial, kRedBlocks, 17, ghw_ghb);
}

void adam_step(float* p, const float* g, float* m, float* v, size_t count, float lr, float beta1, float beta2,
               float eps, float bc1, float bc2, hipStream_t stream)
{
    adam_kernel<<<grid_for(count), kThreads, 0, stream>>>(p, g, m, v, count, lr, beta1, beta2, eps, bc1, bc2);
}

}  // namespace nn
}  // namespace etn
