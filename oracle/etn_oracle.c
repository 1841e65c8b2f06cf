/*
 * etn_oracle.c -- CPU oracle (test infrastructure only; see etn_oracle.h).
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).  The FP32
 * functions perform the float operations of the reference's OpenCL C kernels
 * in the order the source writes them; the FP64 functions run the same
 * algorithms in double.  Reference citations are relative to the reference
 * checkout root (ettention/...).
 */
#include "etn_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define EPS_F 1e-4f   /* ettention/libReconstruction/ocl_src/Math.hl:4 */
#define EPS_D 1e-4

/* ------------------------------------------------------------------------
 * small vector helpers (float and double)
 * ---------------------------------------------------------------------- */
static void v3sub(const float a[3], const float b[3], float o[3]) { o[0]=a[0]-b[0]; o[1]=a[1]-b[1]; o[2]=a[2]-b[2]; }

/* doNormalize: length in double of a float dot product, components divided
 * in double then narrowed.  ettention/libReconstruction/src/framework/math/Vec3.h:80-90,
 * Vec3.cpp:67-70 (getLength = sqrt((double)dotProduct)). */
static int normalize_f(const float in[3], float out[3])
{
    float dot = in[0]*in[0] + in[1]*in[1];
    dot = dot + in[2]*in[2];
    double len = sqrt((double)dot);
    if (len == 0.0) return -1;
    out[0] = (float)((double)in[0] / len);
    out[1] = (float)((double)in[1] / len);
    out[2] = (float)((double)in[2] / len);
    return 0;
}

/* ------------------------------------------------------------------------
 * geometry
 * ---------------------------------------------------------------------- */

/* SingleAxisTiltRotator::doRotate
 * ettention/libReconstruction/src/model/geometry/SingleAxisTiltRotator.cpp:62-78 */
static void rotate_point_f(float c[3], float t, float xt)
{
    if (xt == 0.0f) {
        float ct = cosf(t), st = sinf(t);
        float x = c[0]*ct - c[2]*st;
        float y = c[1];
        float z = c[0]*st + c[2]*ct;
        c[0] = x; c[1] = y; c[2] = z;
    } else {
        float cx = cosf(xt), sx = sinf(xt);
        float p0 = c[0];
        float p1 = c[1]*cx + c[2]*sx;
        float p2 = -c[1]*sx + c[2]*cx;
        float axr[3] = {0.0f, cx, -sx}, ax[3] = {0.0f, 0.0f, 0.0f};
        normalize_f(axr, ax);
        float ct = cosf(t), st = sinf(t), omc = 1 - ct;
        float nx = (ct + omc*ax[0]*ax[0])*p0 + (omc*ax[0]*ax[1] - st*ax[2])*p1 + (omc*ax[0]*ax[2] + st*ax[1])*p2;
        float ny = (omc*ax[0]*ax[1] + st*ax[2])*p0 + (ct + omc*ax[1]*ax[1])*p1 + (omc*ax[1]*ax[2] - st*ax[0])*p2;
        float nz = (omc*ax[0]*ax[2] - st*ax[1])*p0 + (omc*ax[1]*ax[2] + st*ax[0])*p1 + (ct + omc*ax[2]*ax[2])*p2;
        c[0] = nx; c[1] = ny; c[2] = nz;
    }
}

/* SingleAxisTiltRotator::createRotatedScannerGeometry: angle = -tilt/180*pi
 * in float, then source, detector, h and v pitch rotated.
 * ettention/libReconstruction/src/model/geometry/SingleAxisTiltRotator.cpp:42-59 */
void orc_rotate(const orc_scanner* base, float tilt_deg, float xtilt_deg, orc_scanner* out)
{
    float t = -tilt_deg / 180.f * (float)M_PI;
    float xt = xtilt_deg / 180.f * (float)M_PI;
    *out = *base;
    rotate_point_f(out->detector, t, xt);
    rotate_point_f(out->source, t, xt);
    rotate_point_f(out->hpitch, t, xt);
    rotate_point_f(out->vpitch, t, xt);
}

/* Same rotation in double (x-axis tilt 0 only is used by the FP64 reference;
 * the general branch follows the same formula). out = src, det, h, v. */
void orc_rotate_f64(const orc_scanner* base, double tilt_deg, double xtilt_deg, double out[12])
{
    double t = -tilt_deg / 180.0 * M_PI;
    double xt = xtilt_deg / 180.0 * M_PI;
    const float* in[4] = {base->source, base->detector, base->hpitch, base->vpitch};
    for (int k = 0; k < 4; ++k) {
        double c0 = in[k][0], c1 = in[k][1], c2 = in[k][2];
        if (xt == 0.0) {
            out[3*k+0] = c0*cos(t) - c2*sin(t);
            out[3*k+1] = c1;
            out[3*k+2] = c0*sin(t) + c2*cos(t);
        } else {
            double p0 = c0, p1 = c1*cos(xt) + c2*sin(xt), p2 = -c1*sin(xt) + c2*cos(xt);
            double a0 = 0.0, a1 = cos(xt), a2 = -sin(xt);
            double al = sqrt(a1*a1 + a2*a2); a1 /= al; a2 /= al;
            double ct = cos(t), st = sin(t), omc = 1 - ct;
            out[3*k+0] = (ct + omc*a0*a0)*p0 + (omc*a0*a1 - st*a2)*p1 + (omc*a0*a2 + st*a1)*p2;
            out[3*k+1] = (omc*a0*a1 + st*a2)*p0 + (ct + omc*a1*a1)*p1 + (omc*a1*a2 - st*a0)*p2;
            out[3*k+2] = (omc*a0*a2 - st*a1)*p0 + (omc*a1*a2 + st*a0)*p1 + (ct + omc*a2*a2)*p2;
        }
    }
}

/* ParallelScannerGeometry::getRayDirection = doNormalize(detector - source)
 * ettention/libReconstruction/src/model/geometry/ParallelScannerGeometry.cpp:44-47 */
void orc_ray_direction(const orc_scanner* g, float out[3])
{
    float d[3];
    v3sub(g->detector, g->source, d);
    if (normalize_f(d, out) != 0) { out[0] = out[1] = out[2] = 0.0f; }
}

/* Matrix4x4 (column-major, values[4*col+row]) helpers.
 * ettention/libReconstruction/src/framework/math/Matrix4x4.cpp:121-129 (at),
 * :207-219 (determinantOfMinor), :113-119 (determinant), :90-111 (inverse),
 * :297-312 (multiply). */
#define AT(m, c, r) ((m)[4*(c) + (r)])
static float minor_f(const float* m, int dc, int dr)
{
    int c[3], r[3];
    for (int i = 0; i < 3; i++) { c[i] = i < dc ? i : i + 1; r[i] = i < dr ? i : i + 1; }
    return AT(m,c[0],r[0]) * (AT(m,c[1],r[1]) * AT(m,c[2],r[2]) - AT(m,c[1],r[2]) * AT(m,c[2],r[1])) -
           AT(m,c[0],r[1]) * (AT(m,c[1],r[0]) * AT(m,c[2],r[2]) - AT(m,c[1],r[2]) * AT(m,c[2],r[0])) +
           AT(m,c[0],r[2]) * (AT(m,c[1],r[0]) * AT(m,c[2],r[1]) - AT(m,c[1],r[1]) * AT(m,c[2],r[0]));
}
static int inverse_f(const float* m, float* inv)
{
    float det = (AT(m,0,0) * minor_f(m,0,0)) - (AT(m,0,1) * minor_f(m,0,1))
              + (AT(m,0,2) * minor_f(m,0,2)) - (AT(m,0,3) * minor_f(m,0,3));
    if (fabsf(det) == 0.0f) return -1;
    float oneOver = 1.0f / det;
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) {
            AT(inv, r, c) = minor_f(m, c, r) * oneOver;
            if ((r + c) % 2 == 1) AT(inv, r, c) *= -1.0f;
        }
    return 0;
}
static void mul_f(const float* a, const float* b, float* o)
{
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            float v = 0.0f;
            for (int k = 0; k < 4; k++) v += AT(a,k,j) * AT(b,i,k);
            AT(o,i,j) = v;
        }
}

/* ParallelScannerGeometry::getProjectionMatrix: T(res/2 + 0.5) * [h v r e4]^-1
 * ettention/libReconstruction/src/model/geometry/ParallelScannerGeometry.cpp:49-62 */
int orc_projection_matrix(const orc_scanner* g, float out[16])
{
    float r[3];
    orc_ray_direction(g, r);
    float a[16] = { g->hpitch[0], g->hpitch[1], g->hpitch[2], 0.0f,
                    g->vpitch[0], g->vpitch[1], g->vpitch[2], 0.0f,
                    r[0], r[1], r[2], 0.0f,
                    0.0f, 0.0f, 0.0f, 1.0f };
    float inv[16];
    if (inverse_f(a, inv) != 0) return -1;
    float vx = 0.5f * (float)g->res[0] + 0.5f;
    float vy = 0.5f * (float)g->res[1] + 0.5f;
    float t[16] = { 1,0,0,0, 0,1,0,0, 0,0,1,0, vx,vy,0,1 };
    mul_f(t, inv, out);
    return 0;
}

/* ParallelScannerGeometry::getTraversalLengthThroughBoundingBox
 * ettention/libReconstruction/src/model/geometry/ParallelScannerGeometry.cpp:64-69 */
float orc_traversal_length_bbox(const orc_scanner* g, const float bmin[3], const float bmax[3])
{
    float r[3];
    orc_ray_direction(g, r);
    float zin = (bmin[2] - g->source[2]) / r[2];
    float zout = (bmax[2] - g->source[2]) / r[2];
    return (zin > zout ? zin : zout) - (zin < zout ? zin : zout);
}

/* ParallelBeamsBackProjectionKernel::generateSamplingPatternForPerVoxelSampling
 * ettention/libReconstruction/src/algorithm/projections/back/parallel/ParallelBeamsBackProjectionKernel.cpp:69-103
 * with Ray::createFromPointAndDirection (normalises again), Ray::getInvertedNormalizedDirection
 * (+FLT_MAX for zero components, ettention/libReconstruction/src/framework/geometry/Ray.cpp:63-75) and
 * GeometricAlgorithms::getClipRayAgainstBoundingBox (GeometricAlgorithms.cpp:40-51).
 * out4[4*i] = offset xyz, out4[4*i+3] = in-voxel chord length. */
int orc_bp_sample_pattern(const orc_scanner* g, const float vs[3], uint32_t samples, float* out4)
{
    if (samples == 0) return -1;
    uint32_t m = (uint32_t)sqrtf((float)samples);
    if (m * m != samples) return -1;
    float maxvs = fmaxf(fmaxf(vs[0], vs[1]), vs[2]);
    float hn[3], vn[3], h[3], v[3], rdir[3], dir[3];
    if (normalize_f(g->hpitch, hn) || normalize_f(g->vpitch, vn)) return -1;
    for (int k = 0; k < 3; k++) { h[k] = hn[k] * maxvs; v[k] = vn[k] * maxvs; }
    orc_ray_direction(g, rdir);
    if (normalize_f(rdir, dir)) return -1;
    float inv[3];
    for (int k = 0; k < 3; k++) inv[k] = dir[k] != 0.0f ? 1.0f / dir[k] : FLT_MAX;
    uint32_t idx = 0;
    for (float y = 0.0f; y < (float)m; y += 1.0f) {
        float yPos = ((y + 0.5f) / (float)m) - 0.5f;
        for (float x = 0.0f; x < (float)m; x += 1.0f) {
            float xPos = ((x + 0.5f) / (float)m) - 0.5f;
            float p[3];
            for (int k = 0; k < 3; k++) { float a = 0.0f + xPos * h[k]; p[k] = a + yPos * v[k]; }
            float tmn[3], tmx[3];
            for (int k = 0; k < 3; k++) {
                float bmn = vs[k] * -0.5f, bmx = vs[k] * 0.5f;
                float t0 = (bmn - p[k]) * inv[k];
                float t1 = (bmx - p[k]) * inv[k];
                tmn[k] = t0 < t1 ? t0 : t1;
                tmx[k] = t0 < t1 ? t1 : t0;
            }
            float tin = fmaxf(fmaxf(tmn[0], tmn[1]), tmn[2]);
            float tout = fminf(fminf(tmx[0], tmx[1]), tmx[2]);
            out4[4*idx+0] = p[0]; out4[4*idx+1] = p[1]; out4[4*idx+2] = p[2];
            out4[4*idx+3] = tout - tin;
            idx++;
        }
    }
    return 0;
}

/* VolumeProperties::init: voxelSize = bbox dimension / resolution
 * ettention/libReconstruction/src/model/volume/VolumeProperties.cpp:21-27 */
void orc_voxel_size(const orc_volume* v, float out[3])
{
    for (int k = 0; k < 3; k++) out[k] = (v->bmax[k] - v->bmin[k]) / (float)v->res[k];
}

/* VolumeProperties::setSubVolumeCount: z-slabs, the first res.z % S slabs one
 * slice larger; slab bbox = running min + subRes * voxelSize.
 * ettention/libReconstruction/src/model/volume/VolumeProperties.cpp:74-98 */
int orc_slab(const orc_volume* v, uint32_t slab, uint32_t sub_res[3], uint32_t* z0, float smin[3], float smax[3])
{
    uint32_t S = v->slabs ? v->slabs : 1;
    if (v->res[2] < S || slab >= S) return -1;
    uint32_t base = v->res[2] / S, rem = v->res[2] - S * base;
    float vs[3];
    orc_voxel_size(v, vs);
    float nmin[3] = {v->bmin[0], v->bmin[1], v->bmin[2]};
    uint32_t zc = 0;
    for (uint32_t i = 0; i <= slab; i++) {
        uint32_t rz = i < rem ? base + 1 : base;
        float mx[3];
        mx[0] = nmin[0] + (float)v->res[0] * vs[0];
        mx[1] = nmin[1] + (float)v->res[1] * vs[1];
        mx[2] = nmin[2] + (float)rz * vs[2];
        if (i == slab) {
            sub_res[0] = v->res[0]; sub_res[1] = v->res[1]; sub_res[2] = rz;
            *z0 = zc;
            memcpy(smin, nmin, sizeof(nmin));
            memcpy(smax, mx, sizeof(mx));
            return 0;
        }
        nmin[2] = mx[2];
        zc += rz;
    }
    return -1;
}

/* ------------------------------------------------------------------------
 * storage conversions
 * ---------------------------------------------------------------------- */
static uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float bitsf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

float orc_f16_to_f32(uint16_t h)
{
    uint32_t s = (uint32_t)(h & 0x8000u) << 16, e = (h >> 10) & 0x1f, m = h & 0x3ffu;
    if (e == 0) { float f = ldexpf((float)m, -24); return s ? -f : f; }
    if (e == 31) return bitsf(s | 0x7f800000u | (m << 13));
    return bitsf(s | ((e - 15 + 127) << 23) | (m << 13));
}

/* vstore_half: round to nearest even (ettention/libReconstruction/ocl_src/HalfFloatSupport.hl:8-9). */
uint16_t orc_f32_to_f16_rte(float f)
{
    uint32_t u = fbits(f), s = (u >> 16) & 0x8000u;
    float a = fabsf(f);
    if (isnan(f)) return (uint16_t)(s | 0x7e00u);
    if (a >= 65520.0f) return (uint16_t)(s | 0x7c00u);
    if (a < 6.103515625e-05f) {                 /* half subnormal range: multiples of 2^-24 */
        float q = rintf(a * 16777216.0f);       /* rintf rounds to nearest even */
        return (uint16_t)(s | (uint32_t)q);
    }
    uint32_t au = u & 0x7fffffffu;
    uint32_t e = ((au >> 23) - 127 + 15), m = au & 0x7fffffu;
    uint32_t hm = m >> 13, rest = m & 0x1fffu;
    uint32_t hv = (e << 10) | hm;
    if (rest > 0x1000u || (rest == 0x1000u && (hv & 1u))) hv += 1;   /* carries into exponent correctly */
    return (uint16_t)(s | hv);
}

/* Host-side Float16Compressor::compress truncates the mantissa toward zero and
 * maps values above the largest half to infinity.
 * ettention/libReconstruction/src/framework/math/Float16Compressor.cpp:7-23 */
uint16_t orc_f32_to_f16_trunc(float f)
{
    uint32_t u = fbits(f), s = (u >> 16) & 0x8000u;
    float a = fabsf(f);
    if (isnan(f)) return (uint16_t)(s | 0x7e00u);
    if (a > 65504.0f) return (uint16_t)(s | 0x7c00u);
    if (a < 6.103515625e-05f) return (uint16_t)(s | (uint32_t)(a * 16777216.0f));
    uint32_t au = u & 0x7fffffffu;
    uint32_t e = ((au >> 23) - 127 + 15), m = au & 0x7fffffu;
    return (uint16_t)(s | (e << 10) | (m >> 13));
}

uint16_t orc_f32_to_bf16_rne(float f)
{
    uint32_t u = fbits(f);
    if (isnan(f)) return (uint16_t)(((u >> 16) & 0x8000u) | 0x7fc0u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
float orc_bf16_to_f32(uint16_t h) { return bitsf((uint32_t)h << 16); }

static float load_voxel(int vt, const void* vol, size_t i)
{
    if (vt == 0) return ((const float*)vol)[i];
    if (vt == 1) return orc_f16_to_f32(((const uint16_t*)vol)[i]);
    return orc_bf16_to_f32(((const uint16_t*)vol)[i]);
}
static void store_voxel(int vt, void* vol, size_t i, float f)
{
    if (vt == 0) ((float*)vol)[i] = f;
    else if (vt == 1) ((uint16_t*)vol)[i] = orc_f32_to_f16_rte(f);
    else ((uint16_t*)vol)[i] = orc_f32_to_bf16_rne(f);
}

/* ------------------------------------------------------------------------
 * forward projection, FP32-faithful
 * ---------------------------------------------------------------------- */
typedef struct { uint32_t sres[3]; float bmin[3], bmax[3], vs[3]; size_t off; } slabinfo;

static int get_slab(const orc_volume* v, uint32_t slab, slabinfo* s)
{
    uint32_t z0;
    if (orc_slab(v, slab, s->sres, &z0, s->bmin, s->bmax)) return -1;
    orc_voxel_size(v, s->vs);
    s->off = (size_t)z0 * v->res[0] * v->res[1];
    return 0;
}

/* ddaTraverseRay and its helpers.
 * ettention/libReconstruction/ocl_src/DDARaytracer.hl:20-110 (getInitialVoxelIndex :20-31,
 * getRemainingDifferenceToVoxelBorder :33-50, calculateRayEntryPoint :52-57, loop :59-110),
 * RayUtils.hl:7-35 (entry/exit, intersection test, invertVectorSafely). */
static void dda_f32(const slabinfo* s, int vt, const void* vol, const uint8_t* mask,
                    const float src[3], const float ray[3], float* outv, float* outl)
{
    float inv[3], t0[3], t1[3], tmn[3], tmx[3];
    for (int k = 0; k < 3; k++) {
        inv[k] = 1.0f / ray[k];
        if (ray[k] == 0.0f) inv[k] = -FLT_MAX;
    }
    for (int k = 0; k < 3; k++) {
        t0[k] = (s->bmin[k] - src[k]) * inv[k];
        t1[k] = (s->bmax[k] - src[k]) * inv[k];
        tmn[k] = fminf(t0[k], t1[k]);
        tmx[k] = fmaxf(t0[k], t1[k]);
    }
    float tin = fmaxf(fmaxf(tmn[0], tmn[1]), tmn[2]);
    float tout = fminf(fminf(tmx[0], tmx[1]), tmx[2]);
    float ivs[3];
    for (int k = 0; k < 3; k++) ivs[k] = 1.0f / s->vs[k];
    float acc = 0.0f, accl = 0.0f;
    if (!(tout - tin > EPS_F)) { *outv = acc; *outl = accl; return; }
    float pos[3];
    int vox[3], sgn[3];
    float stepd[3], res[3];
    for (int k = 0; k < 3; k++) {
        float p = src[k] + tin * ray[k];
        p = fminf(fmaxf(p, s->bmin[k]), s->bmax[k]);
        pos[k] = p;
    }
    for (int k = 0; k < 3; k++) {
        float gpos = (pos[k] - s->bmin[k]) * ivs[k];
        int vi = (int)gpos;                      /* convert_int4_rtz */
        int ub = (int)s->sres[k] - 1;
        vi = vi < 0 ? 0 : vi;
        vi = vi > ub ? ub : vi;
        vox[k] = vi;
    }
    for (int k = 0; k < 3; k++) {
        float sg = ray[k] > 0.0f ? 1.0f : (ray[k] < 0.0f ? -1.0f : 0.0f);
        stepd[k] = s->vs[k] * inv[k] * sg;
        sgn[k] = (int)sg;
    }
    for (int k = 0; k < 3; k++) {
        float r = floorf(pos[k] * ivs[k]) * s->vs[k];
        r = (ray[k] > 0.0f) ? (r + s->vs[k] - pos[k]) * inv[k] : (r - pos[k]) * inv[k];
        r = (ray[k] == 0.0f) ? FLT_MAX : r;
        r = (r == 0.0f) ? stepd[k] : r;
        res[k] = r;
    }
    const size_t nx = s->sres[0], nxy = (size_t)s->sres[0] * s->sres[1];
    unsigned brk = 0;
    while (vox[0] >= 0 && vox[1] >= 0 && vox[2] >= 0 &&
           vox[0] < (int)s->sres[0] && vox[1] < (int)s->sres[1] && vox[2] < (int)s->sres[2] &&
           ++brk < 4096u) {
        float dt = fminf(fminf(res[0], res[1]), res[2]);
        size_t idx = s->off + (size_t)vox[2] * nxy + (size_t)vox[1] * nx + (size_t)vox[0];
        acc += dt * load_voxel(vt, vol, idx);
        accl += dt * (mask ? (float)mask[idx] / 255.0f : 1.0f);
        int ix = (res[0] <= res[1] && res[0] <= res[2]) ? 1 : 0;
        int iy = (res[1] <  res[0] && res[1] <= res[2]) ? 1 : 0;
        int iz = (res[2] <  res[0] && res[2] <  res[1]) ? 1 : 0;
        res[0] = ix ? stepd[0] : res[0] - dt;
        res[1] = iy ? stepd[1] : res[1] - dt;
        res[2] = iz ? stepd[2] : res[2] - dt;
        vox[0] += ix * sgn[0]; vox[1] += iy * sgn[1]; vox[2] += iz * sgn[2];
        ++brk;
    }
    *outv = acc; *outl = accl;
}

/* ParallelBeamsForwardProjection.cl:execute (+ getRaySourcePosition :12-24):
 * average of samples^2 sub-rays, accumulated (+=) into both outputs.
 * ettention/libReconstruction/ocl_src/ParallelBeamsForwardProjection.cl:26-58 */
int orc_fp(const orc_volume* v, int vt, const void* volume, const uint8_t* mask, uint32_t slab,
           const orc_scanner* g, uint32_t samples, float* proj, float* len)
{
    slabinfo s;
    if (get_slab(v, slab, &s)) return -1;
    if (samples == 0) samples = 1;     /* ForwardProjectionKernel.cpp:65-68 */
    float ray[3];
    orc_ray_direction(g, ray);
    const uint32_t nu = g->res[0], nv = g->res[1];
    const float total = (float)(samples * samples);
    const float sd = 1.0f / (float)samples, off = sd / 2.0f;
    #pragma omp parallel for schedule(dynamic, 4)
    for (long long py = 0; py < (long long)nv; py++) {
        for (uint32_t px = 0; px < nu; px++) {
            float ar = 0.0f, al = 0.0f;
            for (uint32_t sx = 0; sx < samples; sx++)
                for (uint32_t sy = 0; sy < samples; sy++) {
                    float fx = (float)px + off + sd * (float)sx;
                    float fy = (float)py + off + sd * (float)sy;
                    float src[3];
                    for (int k = 0; k < 3; k++) {
                        float ru = g->hpitch[k] * fx, rv = g->vpitch[k] * fy;
                        src[k] = g->source[k] + ru + rv;
                    }
                    float a, l;
                    dda_f32(&s, vt, volume, mask, src, ray, &a, &l);
                    ar += a; al += l;
                }
            size_t i = (size_t)py * nu + px;
            proj[i] += ar / total;
            len[i] += al / total;
        }
    }
    return 0;
}

/* Compare.cl:compare -- ettention/libReconstruction/ocl_src/Compare.cl:1-9 */
void orc_compare(size_t n, const float* real, const float* virt, float* res)
{
    for (size_t i = 0; i < n; i++) res[i] = real[i] - virt[i];
}

/* VirtualProjection.cl:compensate -- ettention/libReconstruction/ocl_src/VirtualProjection.cl:3-10 */
void orc_long_object_compensate(size_t n, float* proj, float* len, float c)
{
    for (size_t i = 0; i < n; i++) {
        proj[i] = len[i] > EPS_F ? proj[i] * c / len[i] : 0.0f;
        len[i] = c;
    }
}

/* ComputeRayLength.cl:execute: chord of the pixel-centre ray through the
 * complete bounding box.  ettention/libReconstruction/ocl_src/ComputeRayLength.cl:5-34 */
void orc_ray_length(const orc_volume* v, const orc_scanner* g, float* out)
{
    float ray[3], inv[3];
    orc_ray_direction(g, ray);
    for (int k = 0; k < 3; k++) { inv[k] = 1.0f / ray[k]; if (ray[k] == 0.0f) inv[k] = -FLT_MAX; }
    for (uint32_t y = 0; y < g->res[1]; y++)
        for (uint32_t x = 0; x < g->res[0]; x++) {
            float fx = ((float)x) + 0.5f, fy = ((float)y) + 0.5f;
            float src[3], tmn[3], tmx[3];
            for (int k = 0; k < 3; k++) {
                float ru = g->hpitch[k] * fx, rv = g->vpitch[k] * fy;
                src[k] = g->source[k] + ru + rv;
                float t0 = (v->bmin[k] - src[k]) * inv[k];
                float t1 = (v->bmax[k] - src[k]) * inv[k];
                tmn[k] = fminf(t0, t1); tmx[k] = fmaxf(t0, t1);
            }
            float tin = fmaxf(fmaxf(tmn[0], tmn[1]), tmn[2]);
            float tout = fminf(fminf(tmx[0], tmx[1]), tmx[2]);
            out[(size_t)y * g->res[0] + x] = fmaxf(tout - tin, 0.0f);
        }
}

/* ------------------------------------------------------------------------
 * back projection, FP32-faithful
 * ---------------------------------------------------------------------- */

/* OpenCL fract(x, &i): i = floor(x), returns fmin(x - floor(x), 0x1.fffffep-1f). */
static float cl_fract(float x, float* ip)
{
    float fl = floorf(x);
    *ip = fl;
    return fminf(x - fl, 0x1.fffffep-1f);
}
static float cl_mix(float a, float b, float t) { return a + (b - a) * t; }

/* ImageHandler.hl LinearInterpolation / BilinearInterpolation (zero outside).
 * ettention/libReconstruction/ocl_src/ImageHandler.hl:20-40 */
static float lin_f(const float* row, int n, float c)
{
    float b;
    float t = cl_fract(c - 0.5f, &b);
    int i0 = (int)b, i1 = i0 + 1;
    float v0 = (i0 >= 0 && i0 < n) ? row[i0] : 0.0f;
    float v1 = (i1 >= 0 && i1 < n) ? row[i1] : 0.0f;
    return cl_mix(v0, v1, t);
}
static float bilin_f(const float* img, int nu, int nv, float cx, float cy)
{
    float b;
    float t = cl_fract(cy - 0.5f, &b);
    int j0 = (int)b, j1 = j0 + 1;
    float v0 = (j0 >= 0 && j0 < nv) ? lin_f(img + (size_t)j0 * nu, nu, cx) : 0.0f;
    float v1 = (j1 >= 0 && j1 < nv) ? lin_f(img + (size_t)j1 * nu, nu, cx) : 0.0f;
    return cl_mix(v0, v1, t);
}

/* ParallelBeamsBackProjection.cl:correct, with GetVoxelCenter (VolumeGeometry.hl:17-20)
 * and TransformCoordG (Matrix4x4.hl:38-58); lambda is divided by the block size on the
 * host (BackProjectionKernel.cpp:94).
 * ettention/libReconstruction/ocl_src/ParallelBeamsBackProjection.cl:9-53 */
int orc_bp(const orc_volume* v, int vt, void* volume, const uint8_t* mask, uint32_t slab,
           const orc_scanner* g, uint32_t samples, const float* residual, const float* raylen,
           float lambda, uint32_t block_size)
{
    slabinfo s;
    if (get_slab(v, slab, &s)) return -1;
    float M[16];
    if (orc_projection_matrix(g, M)) return -1;
    float* pat = (float*)malloc(sizeof(float) * 4 * (samples ? samples : 1));
    if (orc_bp_sample_pattern(g, s.vs, samples, pat)) { free(pat); return -1; }
    const float lam = lambda / (float)block_size;
    const int nu = (int)g->res[0], nv = (int)g->res[1];
    const uint32_t sx = s.sres[0], sy = s.sres[1], sz = s.sres[2];
    #pragma omp parallel for schedule(static)
    for (long long z = 0; z < (long long)sz; z++)
        for (uint32_t y = 0; y < sy; y++)
            for (uint32_t x = 0; x < sx; x++) {
                size_t idx = s.off + ((size_t)z * sy + y) * sx + x;
                float acc = 0.0f;
                unsigned used = 0;
                float gc[3] = {(float)x, (float)y, (float)z};
                for (uint32_t si = 0; si < samples; si++) {
                    float p[3];
                    for (int k = 0; k < 3; k++) {
                        float c = s.bmin[k] + (gc[k] + 0.5f) * s.vs[k];
                        p[k] = c + pat[4*si + k];
                    }
                    float r4[4];
                    for (int rr = 0; rr < 4; rr++) {
                        float a = AT(M,0,rr) * p[0] + AT(M,1,rr) * p[1];
                        a = a + AT(M,2,rr) * p[2];
                        r4[rr] = a + AT(M,3,rr) * 1.0f;
                    }
                    float pu = r4[0] / r4[3], pv = r4[1] / r4[3];
                    float L = bilin_f(raylen, nu, nv, pu, pv);
                    float R = bilin_f(residual, nu, nv, pu, pv);
                    float w = pat[4*si + 3] / L;
                    float contrib = R * lam * w;
                    acc += (L > EPS_F) ? contrib : 0.0f;
                    used += (L > EPS_F) ? 1u : 0u;
                }
                float val = load_voxel(vt, volume, idx);
                if (used != 0) {
                    float mv = mask ? (float)mask[idx] / 255.0f : 1.0f;
                    val += mv * acc / (float)used;
                }
                store_voxel(vt, volume, idx, val);
            }
    free(pat);
    return 0;
}

/* ------------------------------------------------------------------------
 * FP64 mathematical reference (same algorithms in double)
 * ---------------------------------------------------------------------- */
typedef struct { uint32_t sres[3]; double bmin[3], bmax[3], vs[3]; size_t off; } slabinfo_d;

static int get_slab_d(const orc_volume* v, uint32_t slab, slabinfo_d* s)
{
    uint32_t S = v->slabs ? v->slabs : 1;
    if (v->res[2] < S || slab >= S) return -1;
    uint32_t base = v->res[2] / S, rem = v->res[2] - S * base, zc = 0;
    for (int k = 0; k < 3; k++) s->vs[k] = ((double)v->bmax[k] - v->bmin[k]) / v->res[k];
    for (uint32_t i = 0; i < slab; i++) zc += i < rem ? base + 1 : base;
    uint32_t rz = slab < rem ? base + 1 : base;
    s->sres[0] = v->res[0]; s->sres[1] = v->res[1]; s->sres[2] = rz;
    for (int k = 0; k < 2; k++) { s->bmin[k] = v->bmin[k]; s->bmax[k] = v->bmin[k] + v->res[k] * s->vs[k]; }
    s->bmin[2] = v->bmin[2] + zc * s->vs[2];
    s->bmax[2] = s->bmin[2] + rz * s->vs[2];
    s->off = (size_t)zc * v->res[0] * v->res[1];
    return 0;
}

static void dda_f64(const slabinfo_d* s, const double* vol, const double src[3], const double ray[3],
                    double* outv, double* outl)
{
    double inv[3], tmn[3], tmx[3];
    for (int k = 0; k < 3; k++) {
        inv[k] = ray[k] == 0.0 ? -DBL_MAX : 1.0 / ray[k];
        double t0 = (s->bmin[k] - src[k]) * inv[k], t1 = (s->bmax[k] - src[k]) * inv[k];
        tmn[k] = fmin(t0, t1); tmx[k] = fmax(t0, t1);
    }
    double tin = fmax(fmax(tmn[0], tmn[1]), tmn[2]);
    double tout = fmin(fmin(tmx[0], tmx[1]), tmx[2]);
    double acc = 0.0, accl = 0.0;
    *outv = 0.0; *outl = 0.0;
    if (!(tout - tin > EPS_D)) return;
    double pos[3], stepd[3], res[3];
    int vox[3], sgn[3];
    for (int k = 0; k < 3; k++) {
        double p = src[k] + tin * ray[k];
        pos[k] = fmin(fmax(p, s->bmin[k]), s->bmax[k]);
        int vi = (int)((pos[k] - s->bmin[k]) / s->vs[k]);
        int ub = (int)s->sres[k] - 1;
        vox[k] = vi < 0 ? 0 : (vi > ub ? ub : vi);
        double sg = ray[k] > 0 ? 1.0 : (ray[k] < 0 ? -1.0 : 0.0);
        stepd[k] = s->vs[k] * inv[k] * sg;
        sgn[k] = (int)sg;
        double r = floor(pos[k] / s->vs[k]) * s->vs[k];
        r = ray[k] > 0.0 ? (r + s->vs[k] - pos[k]) * inv[k] : (r - pos[k]) * inv[k];
        r = ray[k] == 0.0 ? DBL_MAX : r;
        r = r == 0.0 ? stepd[k] : r;
        res[k] = r;
    }
    const size_t nx = s->sres[0], nxy = (size_t)s->sres[0] * s->sres[1];
    unsigned it = 0;
    while (vox[0] >= 0 && vox[1] >= 0 && vox[2] >= 0 &&
           vox[0] < (int)s->sres[0] && vox[1] < (int)s->sres[1] && vox[2] < (int)s->sres[2] &&
           it < 2048u) {
        double dt = fmin(fmin(res[0], res[1]), res[2]);
        acc += dt * vol[s->off + (size_t)vox[2] * nxy + (size_t)vox[1] * nx + (size_t)vox[0]];
        accl += dt;
        int ix = (res[0] <= res[1] && res[0] <= res[2]);
        int iy = (res[1] <  res[0] && res[1] <= res[2]);
        int iz = (res[2] <  res[0] && res[2] <  res[1]);
        res[0] = ix ? stepd[0] : res[0] - dt;
        res[1] = iy ? stepd[1] : res[1] - dt;
        res[2] = iz ? stepd[2] : res[2] - dt;
        vox[0] += ix * sgn[0]; vox[1] += iy * sgn[1]; vox[2] += iz * sgn[2];
        it++;
    }
    *outv = acc; *outl = accl;
}

static void ray_dir_d(const double geom[12], double r[3])
{
    double d[3] = {geom[3]-geom[0], geom[4]-geom[1], geom[5]-geom[2]};
    double l = sqrt(d[0]*d[0] + d[1]*d[1] + d[2]*d[2]);
    r[0] = d[0]/l; r[1] = d[1]/l; r[2] = d[2]/l;
}

/* FP64 counterpart of orc_fp (geometry = src, det, h, v in double). */
int orc_fp_f64(const orc_volume* v, const double* volume, uint32_t slab, const double geom[12],
               const uint32_t det_res[2], uint32_t samples, double* proj, double* len)
{
    slabinfo_d s;
    if (get_slab_d(v, slab, &s)) return -1;
    if (samples == 0) samples = 1;
    double ray[3];
    ray_dir_d(geom, ray);
    const double total = (double)samples * samples, sd = 1.0 / samples, off = sd / 2;
    #pragma omp parallel for schedule(dynamic, 4)
    for (long long py = 0; py < (long long)det_res[1]; py++)
        for (uint32_t px = 0; px < det_res[0]; px++) {
            double ar = 0, al = 0;
            for (uint32_t sx = 0; sx < samples; sx++)
                for (uint32_t sy = 0; sy < samples; sy++) {
                    double fx = px + off + sd * sx, fy = py + off + sd * sy, src[3], a, l;
                    for (int k = 0; k < 3; k++) src[k] = geom[k] + geom[6+k] * fx + geom[9+k] * fy;
                    dda_f64(&s, volume, src, ray, &a, &l);
                    ar += a; al += l;
                }
            size_t i = (size_t)py * det_res[0] + px;
            proj[i] += ar / total;
            len[i] += al / total;
        }
    return 0;
}

static double lin_d(const double* row, int n, double c)
{
    double b = floor(c - 0.5), t = (c - 0.5) - b;
    int i0 = (int)b, i1 = i0 + 1;
    double v0 = (i0 >= 0 && i0 < n) ? row[i0] : 0.0, v1 = (i1 >= 0 && i1 < n) ? row[i1] : 0.0;
    return v0 + (v1 - v0) * t;
}
static double bilin_d(const double* img, int nu, int nv, double cx, double cy)
{
    double b = floor(cy - 0.5), t = (cy - 0.5) - b;
    int j0 = (int)b, j1 = j0 + 1;
    double v0 = (j0 >= 0 && j0 < nv) ? lin_d(img + (size_t)j0 * nu, nu, cx) : 0.0;
    double v1 = (j1 >= 0 && j1 < nv) ? lin_d(img + (size_t)j1 * nu, nu, cx) : 0.0;
    return v0 + (v1 - v0) * t;
}

/* 3x3 inverse of [h v r] in double: (u, v) = A^-1 p + (res/2 + 0.5). */
static int proj_matrix_d(const double geom[12], const uint32_t det_res[2], double M[2][4])
{
    double r[3];
    ray_dir_d(geom, r);
    const double* h = geom + 6; const double* vv = geom + 9;
    double a[3][3] = {{h[0], vv[0], r[0]}, {h[1], vv[1], r[1]}, {h[2], vv[2], r[2]}};
    double det = a[0][0]*(a[1][1]*a[2][2]-a[1][2]*a[2][1]) - a[0][1]*(a[1][0]*a[2][2]-a[1][2]*a[2][0])
               + a[0][2]*(a[1][0]*a[2][1]-a[1][1]*a[2][0]);
    if (det == 0.0) return -1;
    double i00 = (a[1][1]*a[2][2]-a[1][2]*a[2][1]) / det, i01 = (a[0][2]*a[2][1]-a[0][1]*a[2][2]) / det;
    double i02 = (a[0][1]*a[1][2]-a[0][2]*a[1][1]) / det, i10 = (a[1][2]*a[2][0]-a[1][0]*a[2][2]) / det;
    double i11 = (a[0][0]*a[2][2]-a[0][2]*a[2][0]) / det, i12 = (a[0][2]*a[1][0]-a[0][0]*a[1][2]) / det;
    M[0][0] = i00; M[0][1] = i01; M[0][2] = i02; M[0][3] = 0.5 * det_res[0] + 0.5;
    M[1][0] = i10; M[1][1] = i11; M[1][2] = i12; M[1][3] = 0.5 * det_res[1] + 0.5;
    return 0;
}

/* FP64 counterpart of orc_bp. */
int orc_bp_f64(const orc_volume* v, double* volume, uint32_t slab, const double geom[12],
               const uint32_t det_res[2], uint32_t samples, const double* residual,
               const double* raylen, double lambda, uint32_t block_size)
{
    slabinfo_d s;
    if (get_slab_d(v, slab, &s)) return -1;
    uint32_t m = (uint32_t)floor(sqrt((double)samples) + 0.5);
    if (samples == 0 || m * m != samples) return -1;
    double M[2][4];
    if (proj_matrix_d(geom, det_res, M)) return -1;
    double r[3], h[3], vv[3];
    ray_dir_d(geom, r);
    double maxvs = fmax(fmax(s.vs[0], s.vs[1]), s.vs[2]);
    double hl = sqrt(geom[6]*geom[6] + geom[7]*geom[7] + geom[8]*geom[8]);
    double vl = sqrt(geom[9]*geom[9] + geom[10]*geom[10] + geom[11]*geom[11]);
    for (int k = 0; k < 3; k++) { h[k] = geom[6+k] / hl * maxvs; vv[k] = geom[9+k] / vl * maxvs; }
    double* pat = (double*)malloc(sizeof(double) * 4 * samples);
    for (uint32_t yy = 0, idx = 0; yy < m; yy++)
        for (uint32_t xx = 0; xx < m; xx++, idx++) {
            double xp = (xx + 0.5) / m - 0.5, yp = (yy + 0.5) / m - 0.5, p[3], tmn = -DBL_MAX, tmx = DBL_MAX;
            for (int k = 0; k < 3; k++) {
                p[k] = xp * h[k] + yp * vv[k];
                if (r[k] != 0.0) {
                    double t0 = (-0.5 * s.vs[k] - p[k]) / r[k], t1 = (0.5 * s.vs[k] - p[k]) / r[k];
                    tmn = fmax(tmn, fmin(t0, t1)); tmx = fmin(tmx, fmax(t0, t1));
                }
                pat[4*idx+k] = p[k];
            }
            pat[4*idx+3] = tmx - tmn;
        }
    const double lam = lambda / block_size;
    const int nu = (int)det_res[0], nv = (int)det_res[1];
    #pragma omp parallel for schedule(static)
    for (long long z = 0; z < (long long)s.sres[2]; z++)
        for (uint32_t y = 0; y < s.sres[1]; y++)
            for (uint32_t x = 0; x < s.sres[0]; x++) {
                size_t idx = s.off + ((size_t)z * s.sres[1] + y) * s.sres[0] + x;
                double acc = 0; unsigned used = 0;
                double gc[3] = {(double)x, (double)y, (double)z};
                for (uint32_t si = 0; si < samples; si++) {
                    double p[3];
                    for (int k = 0; k < 3; k++) p[k] = s.bmin[k] + (gc[k] + 0.5) * s.vs[k] + pat[4*si+k];
                    double pu = M[0][0]*p[0] + M[0][1]*p[1] + M[0][2]*p[2] + M[0][3];
                    double pv = M[1][0]*p[0] + M[1][1]*p[1] + M[1][2]*p[2] + M[1][3];
                    double L = bilin_d(raylen, nu, nv, pu, pv);
                    if (L > EPS_D) {
                        acc += bilin_d(residual, nu, nv, pu, pv) * lam * (pat[4*si+3] / L);
                        used++;
                    }
                }
                if (used) volume[idx] += acc / used;
            }
    free(pat);
    return 0;
}

/* Exact line integral (Siddon): merge the parametric crossings of the ray
 * with every voxel plane inside [t_in, t_out] and sum segment length times the
 * value of the voxel holding the segment midpoint.  Independent of the DDA;
 * used only to pin it.  dir need not be normalised (result is per unit t). */
static int cmp_d(const void* a, const void* b)
{
    double x = *(const double*)a, y = *(const double*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}
double orc_siddon(const orc_volume* v, const double* volume, const double o[3], const double d[3], double* chord)
{
    double vs[3], tin = -DBL_MAX, tout = DBL_MAX;
    for (int k = 0; k < 3; k++) {
        vs[k] = ((double)v->bmax[k] - v->bmin[k]) / v->res[k];
        if (d[k] != 0.0) {
            double t0 = (v->bmin[k] - o[k]) / d[k], t1 = (v->bmax[k] - o[k]) / d[k];
            tin = fmax(tin, fmin(t0, t1)); tout = fmin(tout, fmax(t0, t1));
        } else if (o[k] < v->bmin[k] || o[k] > v->bmax[k]) { if (chord) *chord = 0; return 0.0; }
    }
    double dl = sqrt(d[0]*d[0] + d[1]*d[1] + d[2]*d[2]);
    if (!(tout > tin)) { if (chord) *chord = 0; return 0.0; }
    size_t cap = (size_t)v->res[0] + v->res[1] + v->res[2] + 8, n = 0;
    double* ts = (double*)malloc(sizeof(double) * cap);
    ts[n++] = tin; ts[n++] = tout;
    for (int k = 0; k < 3; k++) {
        if (d[k] == 0.0) continue;
        for (uint32_t i = 0; i <= v->res[k]; i++) {
            double t = (v->bmin[k] + i * vs[k] - o[k]) / d[k];
            if (t > tin && t < tout) ts[n++] = t;
        }
    }
    qsort(ts, n, sizeof(double), cmp_d);
    double acc = 0.0;
    for (size_t i = 0; i + 1 < n; i++) {
        double ta = ts[i], tb = ts[i+1];
        if (tb <= ta) continue;
        double tm = 0.5 * (ta + tb);
        int idx[3], ok = 1;
        for (int k = 0; k < 3; k++) {
            idx[k] = (int)floor((o[k] + tm * d[k] - v->bmin[k]) / vs[k]);
            if (idx[k] < 0 || idx[k] >= (int)v->res[k]) ok = 0;
        }
        if (ok) acc += (tb - ta) * dl *
            volume[((size_t)idx[2] * v->res[1] + idx[1]) * v->res[0] + idx[0]];
    }
    free(ts);
    if (chord) *chord = (tout - tin) * dl;
    return acc;
}

/* ------------------------------------------------------------------------
 * projection iteration schemes
 * ---------------------------------------------------------------------- */

/* std/boost mt19937 (32-bit Mersenne twister, MT19937 parameters). */
typedef struct { uint32_t mt[624]; uint32_t i; } mt_state;
static void mt_seed(mt_state* s, uint32_t seed)
{
    s->mt[0] = seed;
    for (uint32_t i = 1; i < 624; i++) s->mt[i] = 1812433253u * (s->mt[i-1] ^ (s->mt[i-1] >> 30)) + i;
    s->i = 624;
}
static uint32_t mt_next(mt_state* s)
{
    if (s->i >= 624) {
        for (uint32_t k = 0; k < 624; k++) {
            uint32_t y = (s->mt[k] & 0x80000000u) | (s->mt[(k+1) % 624] & 0x7fffffffu);
            s->mt[k] = s->mt[(k+397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        s->i = 0;
    }
    uint32_t y = s->mt[s->i++];
    y ^= y >> 11; y ^= (y << 7) & 0x9d2c5680u; y ^= (y << 15) & 0xefc60000u; y ^= y >> 18;
    return y;
}
uint32_t orc_mt19937_state_words(void) { return (uint32_t)(sizeof(mt_state) / sizeof(uint32_t)); }
void orc_mt19937_seed(uint32_t* st, uint32_t seed) { mt_seed((mt_state*)st, seed); }
uint32_t orc_mt19937_next(uint32_t* st) { return mt_next((mt_state*)st); }
uint32_t orc_mt19937_nth(uint32_t seed, uint32_t n)
{
    mt_state s; mt_seed(&s, seed);
    uint32_t v = 0;
    for (uint32_t k = 0; k < n; k++) v = mt_next(&s);
    return v;
}

/* boost::uniform_int<>(lo, hi) on a 32-bit engine (Boost >= 1.47
 * generate_uniform_int, engine range larger than the target range): bucket
 * size floor(2^32-1 / (range+1)), +1 when the remainder equals range; draw
 * until eng()/bucket <= range.  No engine call when range == 0.
 * Used by ProjectionSetRandom::reorder (ProjectionSetRandom.cpp:17-38). */
static uint32_t boost_uniform(mt_state* s, uint32_t lo, uint32_t hi)
{
    uint32_t range = hi - lo;
    if (range == 0) return lo;
    if (range == 0xffffffffu) return mt_next(s);
    uint32_t brange = 0xffffffffu;
    uint32_t bucket = brange / (range + 1u);
    if (brange % (range + 1u) == range) ++bucket;
    for (;;) {
        uint32_t r = mt_next(s) / bucket;
        if (r <= range) return lo + r;
    }
}
uint32_t orc_boost_uniform_int(uint32_t* st, uint32_t lo, uint32_t hi)
{
    return boost_uniform((mt_state*)st, lo, hi);
}

/* ProjectionSet::setProjectionCount (skip list) + reorder():
 *   identity  ProjectionSetIdentity.cpp:17-20
 *   maxangle  ProjectionSetMaxAngle.cpp:16-38
 *   buckets   ProjectionSetBuckets.cpp:16-42
 *   random    ProjectionSetRandom.cpp:17-38: default-seeded mt19937, the index
 *             list reset to 0..n-1 (skip list ignored), then std::random_shuffle
 *             (for i = 1..n-1: j = rand(i+1) in [0, i]; swap).
 * ettention/libReconstruction/src/model/projectionset/ProjectionSet.cpp:14-25 */
int orc_projection_order(const char* type, uint32_t count, const uint32_t* skip, uint32_t nskip,
                         uint32_t* out, uint32_t* out_count)
{
    uint32_t n = 0;
    for (uint32_t i = 0; i < count; i++) {
        int sk = 0;
        for (uint32_t j = 0; j < nskip; j++) if (skip[j] == i) sk = 1;
        if (!sk) out[n++] = i;
    }
    *out_count = n;
    if (!strcmp(type, "identity")) return 0;
    uint32_t* tmp = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    if (!strcmp(type, "maxangle")) {
        uint32_t j = n / 2, k = 0;
        for (uint32_t i = 0; i < n; i++) tmp[i] = (i % 2 == 0) ? out[j++] : out[k++];
        memcpy(out, tmp, sizeof(uint32_t) * n);
    } else if (!strcmp(type, "buckets")) {
        int cur[4];
        for (int i = 0; i < 4; i++) cur[i] = i * (int)n / 4;
        for (int i = 0; i < (int)n; i++) tmp[i] = out[cur[i % 4]++];
        memcpy(out, tmp, sizeof(uint32_t) * n);
    } else if (!strcmp(type, "random")) {
        mt_state s; mt_seed(&s, 5489u);
        for (uint32_t i = 0; i < n; i++) out[i] = i;
        for (uint32_t i = 1; i < n; i++) {
            uint32_t j = boost_uniform(&s, 0, i);
            if (i != j) { uint32_t t = out[i]; out[i] = out[j]; out[j] = t; }
        }
    } else { free(tmp); return -1; }
    free(tmp);
    return 0;
}

/* ------------------------------------------------------------------------
 * block-iterative loop
 * ---------------------------------------------------------------------- */

/* BlockIterativeProjectionAccess: blockCount = ceil((float)N / (float)B), the last
 * block holds the remainder (BlockIterativeProjectionAccess.cpp:46,55-70). */
static uint32_t block_count(uint32_t n, uint32_t B)
{
    return (uint32_t)ceilf((float)n / (float)B);
}
static uint32_t block_len(uint32_t n, uint32_t B, uint32_t nb, uint32_t b)
{
    if (b + 1 < nb) return B;
    return n - (nb - 1) * B;
}

/* BlockIterativeReconstructionOperator::handleOneIteration and its phases:
 * FP over slabs x block projections (accumulating), residual with optional
 * long-object compensation, BP over slabs in reverse order with lambda/B.
 * LOC uses the traversal length of the geometry last set on the geometric
 * setup, i.e. the block's last projection (compensateLongObject reads
 * GeometricSetup::getScannerGeometry()).
 * ettention/libReconstruction/src/algorithm/reconstruction/blockiterative/
 *   BlockIterativeReconstructionOperator.cpp:92-185; BlockIterativeState.cpp:29-48 */
int orc_block_iterative(const orc_recon* r, void* volume, uint32_t iterations)
{
    const uint32_t nu = r->base.res[0], nv = r->base.res[1];
    const size_t P = (size_t)nu * nv;
    uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * (r->nproj ? r->nproj : 1));
    uint32_t n = 0;
    if (orc_projection_order(r->order, r->nproj, r->skip, r->nskip, order, &n) || n == 0) { free(order); return -1; }
    uint32_t B = r->block_size ? r->block_size : 1;
    uint32_t nb = block_count(n, B);
    uint32_t S = r->vol.slabs ? r->vol.slabs : 1;
    uint32_t Bmax = B < n ? B : n;
    float* virt = (float*)malloc(sizeof(float) * P * Bmax);
    float* len = (float*)malloc(sizeof(float) * P * Bmax);
    float* res = (float*)malloc(sizeof(float) * P * Bmax);
    orc_scanner* geo = (orc_scanner*)malloc(sizeof(orc_scanner) * Bmax);
    int rc = 0;
    for (uint32_t it = 0; it < iterations && !rc; it++)
        for (uint32_t b = 0; b < nb && !rc; b++) {
            uint32_t bl = block_len(n, B, nb, b);
            memset(virt, 0, sizeof(float) * P * bl);
            memset(len, 0, sizeof(float) * P * bl);
            memset(res, 0, sizeof(float) * P * bl);
            for (uint32_t p = 0; p < bl; p++)
                orc_rotate(&r->base, r->tilt_deg[order[b * B + p]], r->xtilt_deg, &geo[p]);
            for (uint32_t sl = 0; sl < S; sl++)
                for (uint32_t p = 0; p < bl; p++)
                    rc |= orc_fp(&r->vol, r->voxel_type, volume, r->mask, sl, &geo[p], r->fp_samples,
                                 virt + P * p, len + P * p);
            for (uint32_t p = 0; p < bl; p++) {
                if (r->use_loc) {
                    float c = orc_traversal_length_bbox(&geo[bl - 1], r->vol.bmin, r->vol.bmax);
                    orc_long_object_compensate(P, virt + P * p, len + P * p, c);
                }
                orc_compare(P, r->projections + P * order[b * B + p], virt + P * p, res + P * p);
            }
            for (uint32_t si = 0; si < S; si++)
                for (uint32_t p = 0; p < bl; p++)
                    rc |= orc_bp(&r->vol, r->voxel_type, volume, r->mask, S - si - 1, &geo[p], r->bp_samples,
                                 res + P * p, len + P * p, r->lambda, bl);
        }
    free(virt); free(len); free(res); free(geo); free(order);
    return rc;
}

/* FP64 counterpart of orc_block_iterative (no mask, no storage rounding). */
int orc_block_iterative_f64(const orc_recon* r, double* volume, uint32_t iterations)
{
    const uint32_t nu = r->base.res[0], nv = r->base.res[1];
    const size_t P = (size_t)nu * nv;
    uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * (r->nproj ? r->nproj : 1));
    uint32_t n = 0;
    if (orc_projection_order(r->order, r->nproj, r->skip, r->nskip, order, &n) || n == 0) { free(order); return -1; }
    uint32_t B = r->block_size ? r->block_size : 1;
    uint32_t nb = block_count(n, B), S = r->vol.slabs ? r->vol.slabs : 1, Bmax = B < n ? B : n;
    double* virt = (double*)malloc(sizeof(double) * P * Bmax);
    double* len = (double*)malloc(sizeof(double) * P * Bmax);
    double* res = (double*)malloc(sizeof(double) * P * Bmax);
    double* geo = (double*)malloc(sizeof(double) * 12 * Bmax);
    int rc = 0;
    for (uint32_t it = 0; it < iterations && !rc; it++)
        for (uint32_t b = 0; b < nb && !rc; b++) {
            uint32_t bl = block_len(n, B, nb, b);
            memset(virt, 0, sizeof(double) * P * bl);
            memset(len, 0, sizeof(double) * P * bl);
            for (uint32_t p = 0; p < bl; p++)
                orc_rotate_f64(&r->base, r->tilt_deg[order[b * B + p]], r->xtilt_deg, geo + 12 * p);
            for (uint32_t sl = 0; sl < S; sl++)
                for (uint32_t p = 0; p < bl; p++)
                    rc |= orc_fp_f64(&r->vol, volume, sl, geo + 12 * p, r->base.res, r->fp_samples,
                                     virt + P * p, len + P * p);
            for (uint32_t p = 0; p < bl; p++) {
                double* gl = geo + 12 * (bl - 1);
                double rr[3]; ray_dir_d(gl, rr);
                double c = fabs((r->vol.bmax[2] - r->vol.bmin[2]) / rr[2]);
                const float* meas = r->projections + P * order[b * B + p];
                for (size_t i = 0; i < P; i++) {
                    double vp = virt[P * p + i];
                    if (r->use_loc) {
                        vp = len[P * p + i] > EPS_D ? vp * c / len[P * p + i] : 0.0;
                        len[P * p + i] = c;
                    }
                    res[P * p + i] = (double)meas[i] - vp;
                }
            }
            for (uint32_t si = 0; si < S; si++)
                for (uint32_t p = 0; p < bl; p++)
                    rc |= orc_bp_f64(&r->vol, volume, S - si - 1, geo + 12 * p, r->base.res, r->bp_samples,
                                     res + P * p, len + P * p, r->lambda, bl);
        }
    free(virt); free(len); free(res); free(geo); free(order);
    return rc;
}

/* ||A x - b||_2 over every (non-skipped) projection: A is the forward
 * projection above (all slabs, fp_samples), accumulated in double. */
double orc_residual_norm(const orc_recon* r, const void* volume)
{
    const size_t P = (size_t)r->base.res[0] * r->base.res[1];
    float* virt = (float*)malloc(sizeof(float) * P);
    float* len = (float*)malloc(sizeof(float) * P);
    uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * (r->nproj ? r->nproj : 1));
    uint32_t n = 0;
    double acc = 0.0;
    orc_projection_order("identity", r->nproj, r->skip, r->nskip, order, &n);
    for (uint32_t q = 0; q < n; q++) {
        orc_scanner g;
        orc_rotate(&r->base, r->tilt_deg[order[q]], r->xtilt_deg, &g);
        memset(virt, 0, sizeof(float) * P);
        memset(len, 0, sizeof(float) * P);
        for (uint32_t sl = 0; sl < (r->vol.slabs ? r->vol.slabs : 1); sl++)
            orc_fp(&r->vol, r->voxel_type, volume, r->mask, sl, &g, r->fp_samples, virt, len);
        const float* meas = r->projections + P * order[q];
        for (size_t i = 0; i < P; i++) { double d = (double)meas[i] - (double)virt[i]; acc += d * d; }
    }
    free(virt); free(len); free(order);
    return sqrt(acc);
}

/* Worker threads the OpenMP loops above use (reported with the CPU baseline). */
#ifdef _OPENMP
#include <omp.h>
#endif
int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
