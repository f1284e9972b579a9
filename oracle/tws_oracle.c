/*
 * tws_oracle.c -- CPU restatement of terrainwatersim's virtual-pipe water step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (libtws.so, the C++ host
 * runtime, bench.py's measured path) links, loads or calls this file.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and
 * only as the checker / the reported CPU baseline.
 *
 * What it restates (paths relative to the reference repository root):
 *   flux pass    terrainwatersim/shader/flowUpdate.comp:14-62
 *   apply pass   terrainwatersim/shader/flowApply.comp:16-52
 *   brush        terrainwatersim/shader/waterBrush.comp:20-31 (+ helper.glsl:68 saturate)
 *   constants    terrainwatersim/source/scene/Terrain.cpp:175-198,
 *                terrainwatersim/shader/simulationCommon.glsl:1-13
 *   step order   terrainwatersim/source/scene/Terrain.cpp:253-265 (update, then apply)
 *
 * Every expression keeps the shader's operation order (GLSL a+b+c+d is
 * ((a+b)+c)+d).  The file must be compiled with -ffp-contract=off so that no
 * multiply-add is fused behind our back; fused forms are written out
 * explicitly (fmaf/fma) where a variant asks for them.
 *
 * The same code is instantiated twice through ORACLE_REAL: float (the FP32
 * restatement, expected to match the HIP kernels bit for bit) and double (the
 * FP64 accuracy oracle).
 *
 * State layout (host): b, d are W*H arrays, row-major (x fastest, index
 * x + y*W, as Terrain.cpp:216-220); f is W*H*4 interleaved per cell in the
 * reference's Flow channel order (+x, -x, +y, -y); v is W*H*2 (x, y).
 *
 * Pins (tests/test_oracle.py): closed-form first steps on 1x2, 3x3 and 2x2
 * grids, lake-at-rest (exact zero flux), non-negativity, closed-wall volume
 * conservation, x-mirror bitwise equivariance, FP32 vs FP64 agreement, the
 * brush's closed form, and the llvmpipe golden fixtures of the unmodified
 * shaders where those exist.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define TWS_ORACLE_BOUNDARY_CLOSED 0
#define TWS_ORACLE_BOUNDARY_ZERO 1
#define TWS_ORACLE_BOUNDARY_DISPATCH 2

/* ---------------------------------------------------------------------------
 * Constants: Terrain::SetSimulationStepsPerSecond / SetFlowDamping /
 * SetFlowAcceleration (Terrain.cpp:175-198).  ezTime stores seconds as double
 * (Terrain.cpp:28,177); ezMath::Pow(float,float) is taken as powf (ezEngine is
 * an empty submodule in the reference, so this one call is parity unpinned and
 * the tests feed both sides identical F/K/C floats instead).
 * ------------------------------------------------------------------------- */
void oracle_params_reference(float steps_per_second, float damping, float acceleration,
                             float grid_world_size, unsigned grid_resolution, float out[3])
{
    double dt = (double)(1.0f / steps_per_second);                   /* :177 */
    float cell = grid_world_size / (float)grid_resolution;           /* :183,196 */
    out[0] = powf(damping, (float)dt);                               /* :190 F */
    out[1] = (float)(dt * (double)acceleration * (double)cell);      /* :197 K */
    out[2] = (float)(dt / (double)(cell * cell));                    /* :184 C */
}

/* Physical parameterisation named by the simulationCommon.glsl:5-12 comments:
 * F = damping^dt, K = dt*A*g/l, C = dt/l^2.  Products in double, one rounding. */
void oracle_params_physical(float dt, float gravity, float cell_size, float pipe_area,
                            float damping, float out[3])
{
    out[0] = (float)pow((double)damping, (double)dt);
    out[1] = (float)((double)dt * (double)pipe_area * (double)gravity / (double)cell_size);
    out[2] = (float)((double)dt / ((double)cell_size * (double)cell_size));
}

/* GLSL's max (GLSL 4.60 §8.3): "y if x < y, otherwise x".  The shaders call
 * it as max(0, t) (flowUpdate.comp:54, flowApply.comp:41), which gives +0 for
 * t = -0; glibc's fmaxf returns -0 there.  For t = NaN both give 0. */
static inline float glsl_maxf(float x, float y) { return x < y ? y : x; }
static inline double glsl_max(double x, double y) { return x < y ? y : x; }

/* ------------------------------------------------------------------------- */
#define ORACLE_REAL float
#define ORACLE_SFX f32
#define ORACLE_MAX glsl_maxf
#include "tws_oracle_body.h"
#undef ORACLE_REAL
#undef ORACLE_SFX
#undef ORACLE_MAX

#define ORACLE_REAL double
#define ORACLE_SFX f64
#define ORACLE_MAX glsl_max
#include "tws_oracle_body.h"
#undef ORACLE_REAL
#undef ORACLE_SFX
#undef ORACLE_MAX

/* Host threads used by the restatement (each cell's update is independent of
 * the order cells are visited in, so the result does not depend on it). */
void oracle_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int oracle_get_threads(void) { return omp_get_max_threads(); }

/* FP64 sum of a float plane, fixed order (row by row).  Volume diagnostic. */
double oracle_sum_f32(const float* a, int64_t n)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += (double)a[i];
    return s;
}

/* ---- terrain mip chain ------------------------------------------------------
 * Terrain::PerformSimulationStep regenerates the TerrainData mip chain after
 * every frame with steps (Terrain.cpp:270-276 -> Texture2D::GenMipMaps,
 * Texture2D.cpp:64-68: glGenerateMipmap).  The chain is full:
 * floor(log2(max(W,H))) + 1 levels (Texture.cpp:28-44), level l being
 * max(1, W>>l) x max(1, H>>l).
 *
 * Mesa's llvmpipe builds level l+1 by a bilinear blit of level l: each
 * texel samples level l at its own centre, u = (i+0.5)*ws/wd - 0.5, with
 * lerp(a, b, t) = a + t*(b - a), first along x, then along y, clamped to the
 * edge.  For halving (ws = 2*wd) this is t = 0.5 between texels 2i and 2i+1,
 * which the llvmpipe fixtures pin bit for bit.  For any other ratio llvmpipe
 * takes the sample position from its rasteriser's interpolation, with
 * rounding of up to ~1e-4 texel; here the position is computed exactly (in
 * double, t rounded once to float), so those levels agree with llvmpipe to
 * within that rounding only. */
static void mip_axis(int i, int ws, int wd, int* i0, int* i1, float* t)
{
    if (ws == 2 * wd) {
        *i0 = 2 * i;
        *i1 = 2 * i + 1;
        *t = 0.5f;
        return;
    }
    const double u = ((double)i + 0.5) * (double)ws / (double)wd - 0.5;
    double fl = floor(u);
    if (fl < 0.0) fl = 0.0;
    int a = (int)fl;
    if (a > ws - 1) a = ws - 1;
    *i0 = a;
    *i1 = a + 1 < ws ? a + 1 : ws - 1;
    *t = (float)(u - (double)a);
    if (*t < 0.0f) *t = 0.0f;
}

static float mip_lerp(float a, float b, float t) { return a + t * (b - a); }

/* One level: src ws x hs RGBA -> dst wd x hd RGBA. */
static void mip_level(const float* src, int ws, int hs, float* dst, int wd, int hd)
{
#pragma omp parallel for schedule(static)
    for (int j = 0; j < hd; ++j) {
        int y0, y1;
        float ty;
        mip_axis(j, hs, hd, &y0, &y1, &ty);
        for (int i = 0; i < wd; ++i) {
            int x0, x1;
            float tx;
            mip_axis(i, ws, wd, &x0, &x1, &tx);
            for (int c = 0; c < 4; ++c) {
                const float s00 = src[((size_t)y0 * ws + x0) * 4 + c];
                const float s01 = src[((size_t)y0 * ws + x1) * 4 + c];
                const float s10 = src[((size_t)y1 * ws + x0) * 4 + c];
                const float s11 = src[((size_t)y1 * ws + x1) * 4 + c];
                dst[((size_t)j * wd + i) * 4 + c] =
                    mip_lerp(mip_lerp(s00, s01, tx), mip_lerp(s10, s11, tx), ty);
            }
        }
    }
}

int oracle_mip_levels(int W, int H)
{
    int n = 0;
    while (W > 0 || H > 0) { W /= 2; H /= 2; ++n; }
    return n;
}

/* Levels 1..levels-1 of the chain whose level 0 is the RGBA texture
 * `rgba0` (W x H), written one after the other into `out`. */
void oracle_mips(int W, int H, const float* rgba0, float* out)
{
    const int levels = oracle_mip_levels(W, H);
    const float* src = rgba0;
    int ws = W, hs = H;
    float* dst = out;
    for (int l = 1; l < levels; ++l) {
        const int wd = W >> l > 0 ? W >> l : 1, hd = H >> l > 0 ? H >> l : 1;
        mip_level(src, ws, hs, dst, wd, hd);
        src = dst;
        dst += (size_t)wd * hd * 4;
        ws = wd;
        hs = hd;
    }
}
