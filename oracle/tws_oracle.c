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

/* ------------------------------------------------------------------------- */
#define ORACLE_REAL float
#define ORACLE_SFX f32
#define ORACLE_MAX fmaxf
#include "tws_oracle_body.h"
#undef ORACLE_REAL
#undef ORACLE_SFX
#undef ORACLE_MAX

#define ORACLE_REAL double
#define ORACLE_SFX f64
#define ORACLE_MAX fmax
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
