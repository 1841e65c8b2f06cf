/*
 * etn_oracle.h -- CPU oracle for the block-iterative SART/SIRT hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (libetn.so) never links, loads or calls it.
 *
 * It restates, in plain C, the arithmetic of Ettention's OpenCL kernels and
 * the host code that feeds them (reference paths are relative to the
 * reference checkout root):
 *   - FP32-faithful mode: the same float operations in the same order as the
 *     OpenCL C source (no FMA contraction: built with -ffp-contract=off);
 *   - FP64 mode: the same algorithm evaluated in double (geometry too), the
 *     mathematical reference for the reduced-precision storage variants;
 *   - an independent exact line integral (Siddon) used only to pin the DDA.
 *
 * Every function cites the reference file:line it follows.
 */
#ifndef ETN_ORACLE_H
#define ETN_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Scanner geometry of one projection: ScannerGeometry's state
 * (ettention/libReconstruction/src/model/geometry/ScannerGeometry.h). */
typedef struct orc_scanner {
    float source[3];
    float detector[3];
    float hpitch[3];
    float vpitch[3];
    uint32_t res[2];            /* detector resolution (nu, nv) */
} orc_scanner;

/* Volume description: VolumeProperties (VolumeProperties.cpp:21-98). */
typedef struct orc_volume {
    uint32_t res[3];
    float bmin[3];              /* complete bounding box */
    float bmax[3];
    uint32_t slabs;             /* z-slab ("subvolume") count, >= 1 */
} orc_volume;

/* ---- geometry (host side of the reference) ---- */
void orc_rotate(const orc_scanner* base, float tilt_deg, float xtilt_deg, orc_scanner* out);
void orc_rotate_f64(const orc_scanner* base, double tilt_deg, double xtilt_deg, double out[12]);
void orc_ray_direction(const orc_scanner* g, float out[3]);
int  orc_projection_matrix(const orc_scanner* g, float out[16]);
float orc_traversal_length_bbox(const orc_scanner* g, const float bmin[3], const float bmax[3]);
int  orc_bp_sample_pattern(const orc_scanner* g, const float voxel_size[3], uint32_t samples, float* out4);
void orc_voxel_size(const orc_volume* v, float out[3]);
int  orc_slab(const orc_volume* v, uint32_t slab, uint32_t sub_res[3], uint32_t* z0,
              float smin[3], float smax[3]);

/* ---- kernels, FP32-faithful ---- */
/* voxel_type: 0 = float32, 1 = IEEE half (uint16), 2 = bfloat16 (uint16) */
int  orc_fp(const orc_volume* v, int voxel_type, const void* volume, const uint8_t* mask,
            uint32_t slab, const orc_scanner* g, uint32_t samples, float* proj, float* len);
void orc_compare(size_t n, const float* real, const float* virt, float* res);
void orc_long_object_compensate(size_t n, float* proj, float* len, float c);
void orc_ray_length(const orc_volume* v, const orc_scanner* g, float* out);
int  orc_bp(const orc_volume* v, int voxel_type, void* volume, const uint8_t* mask, uint32_t slab,
            const orc_scanner* g, uint32_t samples, const float* residual, const float* raylen,
            float lambda, uint32_t block_size);

/* ---- FP64 mathematical reference ---- */
int  orc_fp_f64(const orc_volume* v, const double* volume, uint32_t slab, const double geom[12],
                const uint32_t det_res[2], uint32_t samples, double* proj, double* len);
int  orc_bp_f64(const orc_volume* v, double* volume, uint32_t slab, const double geom[12],
                const uint32_t det_res[2], uint32_t samples, const double* residual,
                const double* raylen, double lambda, uint32_t block_size);
double orc_siddon(const orc_volume* v, const double* volume, const double origin[3],
                  const double dir[3], double* chord);

/* ---- storage conversions ---- */
uint16_t orc_f32_to_f16_rte(float f);
uint16_t orc_f32_to_f16_trunc(float f);
float    orc_f16_to_f32(uint16_t h);
uint16_t orc_f32_to_bf16_rne(float f);
float    orc_bf16_to_f32(uint16_t h);

/* ---- projection iteration schemes ---- */
uint32_t orc_mt19937_nth(uint32_t seed, uint32_t n);
/* engine state: orc_mt19937_state_words() uint32 words, seeded like mt19937(seed) */
uint32_t orc_mt19937_state_words(void);
void     orc_mt19937_seed(uint32_t* state, uint32_t seed);
uint32_t orc_mt19937_next(uint32_t* state);
uint32_t orc_boost_uniform_int(uint32_t* mt_state_and_index, uint32_t lo, uint32_t hi);
int  orc_projection_order(const char* type, uint32_t count, const uint32_t* skip, uint32_t nskip,
                          uint32_t* out, uint32_t* out_count);

/* ---- block-iterative loop ---- */
typedef struct orc_recon {
    orc_volume vol;
    int voxel_type;              /* 0 f32, 1 f16, 2 bf16 (FP32 arithmetic on all) */
    orc_scanner base;            /* base (untilted) scanner geometry */
    uint32_t nproj;
    const float* tilt_deg;       /* nproj angles */
    float xtilt_deg;
    const float* projections;    /* nproj * nu * nv measured projections */
    uint32_t block_size;         /* 1 = SART, >= nproj = SIRT */
    const char* order;           /* identity | random | maxangle | buckets */
    const uint32_t* skip; uint32_t nskip;
    float lambda;
    int use_loc;
    uint32_t fp_samples, bp_samples;
    const uint8_t* mask;         /* optional prior-knowledge mask (0..255) */
} orc_recon;

int  orc_block_iterative(const orc_recon* r, void* volume, uint32_t iterations);
int  orc_block_iterative_f64(const orc_recon* r, double* volume, uint32_t iterations);
double orc_residual_norm(const orc_recon* r, const void* volume);
int  orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
