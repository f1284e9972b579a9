/*
 * tws.h -- C ABI of the MI355X-native virtual-pipe water simulation.
 *
 * This is the drop-in boundary for the simulation half of terrainwatersim's
 * Terrain class (terrainwatersim/source/scene/Terrain.h:21-39,75-83) and the
 * GPU contract it drives: image unit 0 TerrainData rgba32f (r = terrain b,
 * a = water depth d), image unit 1 Flow rgba32f (outgoing flux through the
 * +x, -x, +y, -y pipes), image unit 2 FlowMap rg16f (directed flow v), and
 * the SimulationParameters uniform block {F, K, C}
 * (terrainwatersim/shader/simulationCommon.glsl:1-13, flowUpdate.comp:5-6,
 * flowApply.comp:5-7).  Each entry point cites the reference interface it
 * replaces.
 *
 * Conventions
 *   - Plain C types only; no torch or HIP types.  Streams are passed as void*
 *     holding a hipStream_t (NULL = a stream the library creates).
 *   - Every function returns TWS_OK (0) or a negative TWS_ERR_* code; no
 *     exception crosses the ABI.  tws_last_error() returns a message for the
 *     calling thread's most recent failure.
 *   - Host arrays are dense and row-major, x fastest: cell (x, y) is index
 *     x + y*width, as the reference fills its textures (Terrain.cpp:216-220).
 *     Flux arrays are 4 floats per cell in the Flow channel order
 *     (+x, -x, +y, -y); velocity arrays are 2 floats per cell (x, y).
 *   - A handle is used by one host thread at a time; its work is ordered on
 *     its stream.  Different handles are independent.
 */
#ifndef TWS_H
#define TWS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TWS_ABI_VERSION 1

/* Status codes. */
#define TWS_OK 0
#define TWS_ERR_INVALID (-1)    /* bad argument or handle                   */
#define TWS_ERR_NOMEM (-2)      /* device or host allocation failed         */
#define TWS_ERR_DEVICE (-3)     /* a HIP call failed                        */
#define TWS_ERR_NONFINITE (-4)  /* NaN/Inf seen in state (tws_check_finite) */
#define TWS_ERR_STATE (-5)      /* call not valid in the handle's state     */
#define TWS_ERR_CAPACITY (-6)   /* request exceeds a configured limit       */

/* Boundary handling of cells outside the W x H domain. */
#define TWS_BOUNDARY_CLOSED 0   /* closed walls: no flow leaves or enters    */
#define TWS_BOUNDARY_ZERO 1     /* the reference shaders' own edge: texels
                                   outside read as r=g=b=0, a=ghost_alpha and
                                   water drains into them
                                   (flowUpdate.comp:14-18, flowApply.comp:16-20).
                                   Every cell is updated: identical to the
                                   shaders when W and H are multiples of 16
                                   (of 32 with rain or the brush), see below */
#define TWS_BOUNDARY_REFERENCE_DISPATCH 2
                                /* ZERO plus the reference's dispatch shapes,
                                   so every size steps exactly as the shaders
                                   do: the two passes run W/16 x H/16 work
                                   groups of 16 x 16 cells (Terrain.cpp:258,
                                   264) and the brush W/32 x H/32 of 32 x 32
                                   (Terrain.cpp:167), integer divisions.
                                   Cells in columns >= 16*(W/16) or rows >=
                                   16*(H/16) keep their depth and flux forever
                                   (neighbours still read them) and have no
                                   directed flow (0); the brush and the rain
                                   reach only columns < 32*(W/32) and rows
                                   < 32*(H/32).  Where W and H are multiples of
                                   32 this is ZERO.  Other sizes step on the
                                   two-pass kernels, single grids only (a batch
                                   refuses them). */

/* Per-step constants of the SimulationParameters block
 * (simulationCommon.glsl:1-13) plus the boundary mode.  F, K and C must be
 * finite and >= 0, as the reference derives them (Terrain.cpp:175-198), and
 * with closed walls K <= 1e8 (the wall's ghost cell stands 1e30 high and its
 * head difference times K must stay finite).  Closed walls are then exact
 * for any state: a wall-ward flux uploaded into an edge cell is dropped when
 * the step loads it. */
typedef struct tws_params {
    float flow_friction;            /* F = FlowFriction_perStep             */
    float water_acceleration;       /* K = WaterAcceleration_perStep        */
    float cell_area_inv_timescaled; /* C = CellAreaInv_timeScaled           */
    int32_t boundary;               /* TWS_BOUNDARY_*                        */
    float ghost_alpha;              /* alpha read outside the domain in
                                       TWS_BOUNDARY_ZERO mode (0 for Mesa) */
} tws_params;

/* Device-resident planes of one simulation (for interop with renderers and
 * other device code).  Row pitch is in floats.  Valid until the next call
 * that steps, uploads or destroys the handle. */
typedef struct tws_device_planes {
    const float* terrain;   /* b                                  */
    const float* depth;     /* d                                  */
    const float* flux[4];   /* +x, -x, +y, -y outgoing flux       */
    int64_t pitch;
    int32_t width, height;
} tws_device_planes;

typedef struct tws_sim tws_sim;

/* ---- parameters --------------------------------------------------------- */

/* F = damping^dt, K = dt*pipe_area*gravity/cell_size, C = dt/cell_size^2:
 * the physical meaning stated by simulationCommon.glsl:5-12. */
int tws_params_physical(float dt, float gravity, float cell_size, float pipe_area,
                        float damping, int32_t boundary, tws_params* out);

/* The reference's own derivation from its UI settings, as
 * Terrain::SetSimulationStepsPerSecond / SetFlowDamping / SetFlowAcceleration
 * compute it (Terrain.cpp:175-198; defaults 60 steps/s, damping 0.98,
 * acceleration 10, world size 1024, resolution 1024: Terrain.cpp:22-30,95). */
int tws_params_reference(float steps_per_second, float damping, float acceleration,
                         float grid_world_size, uint32_t grid_resolution, int32_t boundary,
                         tws_params* out);

/* ---- one simulation ------------------------------------------------------ */

/* Allocates the device state of a width x height grid (replaces the texture
 * allocation of Terrain::CreateHeightmapFromNoiseAndResetSim,
 * Terrain.cpp:205,230,237).  State starts at b = d = f = 0. */
int tws_create(int32_t width, int32_t height, const tws_params* params, void* stream,
               tws_sim** out);
int tws_destroy(tws_sim* sim);

/* Replaces Terrain::SetFlowDamping / SetFlowAcceleration /
 * SetSimulationStepsPerSecond uniform updates (Terrain.cpp:175-198). */
int tws_set_params(tws_sim* sim, const tws_params* params);
int tws_get_params(const tws_sim* sim, tws_params* out);

/* Host -> device copy of the state (replaces Texture2D::SetData of the
 * TerrainData and Flow textures, Terrain.cpp:223,233).  Any pointer may be
 * NULL: terrain/depth keep their device values, flux resets to zero. */
int tws_upload(tws_sim* sim, const float* terrain, const float* depth, const float* flux4);

/* Device -> host copy.  Any pointer may be NULL.  velocity2 receives the
 * directed flow derived from the current flux exactly as flowApply.comp:45-46
 * computes it (the FlowMap of the last step, without the rg16f rounding). */
int tws_download(tws_sim* sim, float* depth, float* flux4, float* velocity2);

/* Reference interleaved layout: terrain_rgba32f[W*H*4] is TerrainData
 * (r = b, g, b channels carried through, a = d) and flow_rgba32f[W*H*4] is
 * Flow (Terrain.cpp:205-237).  Upload reads r and a; download writes r, a and
 * g = b = the given fill (0.3 in the reference, Terrain.cpp:217-218).
 * flowmap_rg16f[W*H*2] receives FlowMap as IEEE half bits, rounded toward
 * zero: the conversion the reference's rg16f image store (flowApply.comp:52)
 * performs on Mesa, pinned by the llvmpipe fixtures.  NULL pointers are
 * skipped. */
int tws_upload_reference_layout(tws_sim* sim, const float* terrain_rgba32f,
                                const float* flow_rgba32f);
int tws_download_reference_layout(tws_sim* sim, float* terrain_rgba32f, float gb_fill,
                                  float* flow_rgba32f, uint16_t* flowmap_rg16f);

/* One radial brush application, d += saturate(1 - |c - p|^2 / R^2) * I over
 * the whole domain (replaces Terrain::ApplyRadialWaterBrush, Terrain.cpp:150-168,
 * and waterBrush.comp:18-32; the caller converts world to texel coordinates). */
int tws_apply_brush(tws_sim* sim, float center_x, float center_y, float intensity,
                    float radius_sq);

/* A brush applied before every step of tws_step (a steady source, "rain").
 * intensity == 0 switches it off. */
int tws_set_rain(tws_sim* sim, float center_x, float center_y, float intensity,
                 float radius_sq);

/* Advances n steps (replaces the hot loop of Terrain::PerformSimulationStep,
 * Terrain.cpp:253-265: per step the flowUpdate then the flowApply dispatch).
 * Asynchronous on the handle's stream. */
int tws_step(tws_sim* sim, int32_t n);

/* tws_step with a GPU timer around every kernel launch, the counterpart of the
 * reference's GL_TIME_ELAPSED query around the simulation block
 * (Scene.cpp:362-364, TimerQuery.cpp:29-72).  launch_ms[i] receives the
 * duration of launch i (HIP events on the handle's stream) for the first
 * `capacity` launches; *launches receives the number of launches.
 * Synchronises the stream. */
int tws_step_profiled(tws_sim* sim, int32_t n, float* launch_ms, int32_t capacity,
                      int32_t* launches);

/* Frame-time driver with the reference's accumulator and 10-step clamp
 * (Terrain::PerformSimulationStep, Terrain.cpp:240-249).  step_seconds is the
 * step length; *steps_done (may be NULL) receives the number of steps run.
 * With a mip target set (tws_set_frame_mips), a frame that ran at least one
 * step also regenerates the mip chain, as the reference does
 * (Terrain.cpp:270-276). */
int tws_perform_simulation_step(tws_sim* sim, double frame_seconds, double step_seconds,
                                int32_t* steps_done);

/* ---- terrain mip chain ---------------------------------------------------- */
/*
 * The full mip chain of the reference-layout TerrainData texture (r = terrain,
 * g = b = gb_fill, a = depth), regenerated by the reference after every frame
 * with steps (Terrain.cpp:270-276 -> Texture2D::GenMipMaps,
 * Texture2D.cpp:64-68, glGenerateMipmap).  Level 0 is the state itself
 * (tws_download_reference_layout); levels 1..levels-1 are stored one after
 * the other, level l being max(1, W>>l) x max(1, H>>l) RGBA32F texels, row
 * by row.  Each level is a bilinear reduction of the one above: bit-identical
 * to Mesa llvmpipe's glGenerateMipmap where a dimension halves exactly (every
 * level of a power-of-two grid); other ratios sample at the exact texel-centre
 * position, where llvmpipe's rasteriser rounds the position by up to ~1e-4
 * texel.
 */

/* Number of levels (floor(log2(max(W,H))) + 1, Texture.cpp:28-44) and the
 * floats levels 1.. take together.  Either output may be NULL. */
int tws_mip_chain_size(int32_t width, int32_t height, int32_t* levels, size_t* floats);

/* Writes levels 1.. into device memory (16-byte aligned, `capacity_floats`
 * floats).  Asynchronous on the handle's stream. */
int tws_generate_mips(tws_sim* sim, float gb_fill, float* device_rgba, size_t capacity_floats);

/* The same chain copied to host memory (the handle's idle state buffers
 * serve as staging).  Synchronises the stream. */
int tws_download_mips(tws_sim* sim, float gb_fill, float* host_rgba, size_t capacity_floats);

/* Device buffer that tws_perform_simulation_step keeps up to date (the
 * texture the reference's renderer samples); NULL switches it off.  The
 * buffer stays owned by the caller and must outlive its use. */
int tws_set_frame_mips(tws_sim* sim, float* device_rgba, size_t capacity_floats, float gb_fill);

/* Total water volume sum(d) over the domain, accumulated in FP64 in a fixed
 * order (deterministic).  Synchronises the stream. */
int tws_volume(tws_sim* sim, double* out);

/* TWS_OK if every d and f seen by the steps since the last call was finite,
 * TWS_ERR_NONFINITE otherwise.  Clears the flag.  Synchronises the stream. */
int tws_check_finite(tws_sim* sim);

int tws_sync(tws_sim* sim);
int tws_get_device_planes(const tws_sim* sim, tws_device_planes* out);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The FP64 volume sum(d) of a depth plane in device
 * memory (`height` rows of `pitch` floats, the first `width` of each used),
 * bit for bit what tws_volume returns for a handle's grid: the same kernel
 * and fixed-order reduction, with its 512 partial sums in a fixed 4 KiB
 * device array of the library (one per device, kept until the process ends)
 * instead of a handle's buffer.  Needs no handle and allocates nothing.
 * Runs on `stream` (NULL: the default stream) and synchronises it.
 * TWS_ERR_INVALID for NULL pointers or a plane beyond the limits (width and
 * height 1..2^20, width <= pitch <= 2^30).  Every call on a device writes all
 * 512 partials before reading them, so nothing carries over from the last
 * call; calls that overlap on one device (other threads or streams) share
 * the array and race: serialise them. */
int tws_volume_noalloc(const float* depth, int32_t width, int32_t height, int64_t pitch,
                       void* stream, double* out);

/* ---- launch configuration ------------------------------------------------ */

/* Kernel variants.  TWOPASS mirrors the reference's structure (one flux
 * kernel and one apply kernel per step, neighbours read from memory).  The
 * TILE variants are the fused, temporally blocked kernel: one launch runs up
 * to k steps on register/LDS-resident tiles. */
#define TWS_KERNEL_AUTO 0
#define TWS_KERNEL_TWOPASS 1
#define TWS_KERNEL_TILE_128x128 2   /* 1024 threads, 2 columns x 8 rows each */
#define TWS_KERNEL_TILE_128x64 3    /*  512 threads                          */
#define TWS_KERNEL_TILE_256x64 4    /* 1024 threads, 4 columns x 4 rows each */
#define TWS_KERNEL_TILE_64x64 5     /*  512 threads, 1 column x 8 rows each:
                                       more, smaller tiles for small grids,
                                       whose steps are bound by latency     */
#define TWS_KERNEL_TILE_64x64_16W 6 /* 1024 threads, 1 column x 4 rows each:
                                       the 64x64 tile with half the serial
                                       work per wave and level               */
#define TWS_KERNEL_TILE_64x48_16W 7 /* 1024 threads, 1 column x 3 rows each:
                                       64x48 tiles, so grids up to 384^2
                                       spread over more CUs                  */
#define TWS_KERNEL_STREAM_128 8     /* streaming strips: each wavefront sweeps
                                       a 128-column strip segment down once
                                       per launch with all k levels in a
                                       register window, so only the x-halo is
                                       recomputed; closed walls (grids on
                                       other edges, or with planes of 2 GiB or
                                       more, step on the tile kernel's
                                       launch); k = 2, 3, 4, 5, 6 or 8; also a
                                       batch size class */
#define TWS_KERNEL_COUNT 9

/* Selects a kernel variant and the number of steps per launch (1, 2, 3, 4,
 * 5, 6 or 8; TWOPASS requires 1, STREAM_128 at least 2).  TWS_KERNEL_AUTO
 * picks from the tuning table: tile kernels up to 2176^2, STREAM_128 at k = 6
 * for closed grids above (a launch entered with tws_tune_record first). */
int tws_set_launch(tws_sim* sim, int32_t kernel, int32_t steps_per_launch);
int tws_get_launch(const tws_sim* sim, int32_t* kernel, int32_t* steps_per_launch);
/* Order in which the step kernel walks a grid's tiles: groups of
 * `group_columns` tile columns, row by row inside a group, so the tiles in
 * flight together form a compact block that shares its halos in L2
 * (0 = row-major).  tws_create picks the measured best for the grid's default
 * launch (16, or 20 for 128x64 tiles at 5 steps per launch).  Results do not
 * depend on it. */
int tws_set_tile_order(tws_sim* sim, int32_t group_columns);

/* Host-side float -> IEEE half conversion rounding toward zero, the FlowMap
 * rounding of tws_download_reference_layout, for n values. */
int tws_float_to_half_rtz(const float* in, uint16_t* out, int64_t n);

/* ---- batches: many independent grids advanced together ------------------ */
/*
 * No counterpart in the reference (its Terrain owns exactly one grid); this is
 * the multi-grid form of the same step.  All grids of a batch advance in one
 * launch sequence per tws_batch_run: the tiles of every grid form one
 * flattened work list, each grid keeps its own size, constants, rain and
 * state, and the sequence is captured once as a hipGraph and replayed.  Grid
 * states are carved from one device allocation of `pool_bytes` made at
 * creation, so adding and removing grids never allocates device memory; a
 * grid that does not fit is refused with TWS_ERR_CAPACITY.  Results are
 * bit-identical to the same grid stepped alone with tws_step.
 */
typedef struct tws_batch tws_batch;

typedef struct tws_batch_stats {
    int32_t grids;          /* grids in the batch                            */
    int32_t max_grids;
    int64_t cells;          /* total cells of all grids                      */
    int64_t pool_bytes;     /* device bytes reserved at creation             */
    int64_t pool_used;      /* device bytes held by grid states              */
    int64_t pool_largest_free;
    int64_t graph_captures; /* launch sequences captured so far              */
    int64_t graph_replays;  /* captured sequences replayed so far            */
} tws_batch_stats;

/* Device bytes one grid of width x height takes in a batch (its admission
 * footprint): 11 FP32 planes (terrain, two depth and eight flux planes) with
 * rows padded to 256 bytes. */
int64_t tws_grid_footprint(int32_t width, int32_t height);

int tws_batch_create(int32_t max_grids, int64_t pool_bytes, void* stream, tws_batch** out);
int tws_batch_destroy(tws_batch* batch);
/* Adds a grid (state zero) and returns its id; TWS_ERR_CAPACITY if the pool
 * has no block of tws_grid_footprint() bytes or all max_grids ids are used. */
int tws_batch_add(tws_batch* batch, int32_t width, int32_t height, const tws_params* params,
                  int32_t* grid_id);
int tws_batch_remove(tws_batch* batch, int32_t grid_id);
int tws_batch_upload(tws_batch* batch, int32_t grid_id, const float* terrain, const float* depth,
                     const float* flux4);
int tws_batch_download(tws_batch* batch, int32_t grid_id, float* depth, float* flux4,
                       float* velocity2);
int tws_batch_set_params(tws_batch* batch, int32_t grid_id, const tws_params* params);
int tws_batch_set_rain(tws_batch* batch, int32_t grid_id, float center_x, float center_y,
                       float intensity, float radius_sq);
int tws_batch_volume(tws_batch* batch, int32_t grid_id, double* out);
int tws_batch_check_finite(tws_batch* batch, int32_t grid_id);
/* Steps grid_id will advance in the next tws_batch_run (0 = stays put). */
int tws_batch_set_steps(tws_batch* batch, int32_t grid_id, int32_t steps);
/* Advances every grid by its pending steps in one launch sequence, then
 * clears the pending counts.  Asynchronous on the batch stream. */
int tws_batch_run(tws_batch* batch);
/* tws_batch_set_steps(n) for every grid, then tws_batch_run. */
int tws_batch_step(tws_batch* batch, int32_t n);
/* Kernel variant and steps per launch for the whole batch (as tws_set_launch;
 * TWOPASS is not available for batches).  Same as tws_batch_set_launch_classes
 * with one class.  TWS_KERNEL_AUTO (the default; reported as kernel
 * TWS_KERNEL_AUTO, k 0) picks per run: TWS_KERNEL_STREAM_128 at k = 6 when the
 * grids with pending steps that the streaming kernel can step hold more than
 * 2^23 cells (2896^2) together; otherwise the tile launch TWS_KERNEL_AUTO
 * gives one grid of all their cells (on other edges).
 * tws_batch_get_launch reports the class of the largest grids. */
int tws_batch_set_launch(tws_batch* batch, int32_t kernel, int32_t steps_per_launch);
int tws_batch_get_launch(const tws_batch* batch, int32_t* kernel, int32_t* steps_per_launch);

#define TWS_BATCH_MAX_CLASSES 4
/* Size classes, each with its own kernel and steps per launch (or
 * TWS_KERNEL_AUTO, whose k is ignored and which picks per run as
 * tws_batch_set_launch describes, among that class's grids): class c
 * (0 <= c < n <= TWS_BATCH_MAX_CLASSES) steps the grids of at most
 * max_cells[c] cells that no earlier class takes; the last class takes every
 * larger grid (its max_cells entry is ignored).  Bounds must be positive and
 * ascending.  Each tile class with pending steps gets its own chain of
 * launches; the chains of a run are captured as parallel branches of one
 * hipGraph.  A class on TWS_KERNEL_STREAM_128 has a chain too: the strip
 * segments of all its grids are the tasks of each launch, with the segment
 * length planned for the class's grid shapes so the launch ends in whole
 * rounds.  Steps short of a full launch of k, and grids the streaming kernel
 * cannot step (such as those on the reference's edge, which take the tile
 * table's launch), run as direct launches after the graph.
 * Results do not depend on the classes: every launch configuration steps a
 * grid bit-identically.  Has no counterpart in the reference, which steps
 * one terrain at a time. */
int tws_batch_set_launch_classes(tws_batch* batch, int32_t n, const int64_t* max_cells,
                                 const int32_t* kernels, const int32_t* steps_per_launch);
/* The classes in effect: *n of them, the first min(*n, capacity) written to
 * the arrays (any of which may be NULL). */
int tws_batch_get_launch_classes(const tws_batch* batch, int32_t capacity, int64_t* max_cells,
                                 int32_t* kernels, int32_t* steps_per_launch, int32_t* n);
/* 1 (default): replay captured hipGraphs; 0: launch kernels directly. */
int tws_batch_use_graphs(tws_batch* batch, int32_t enable);
int tws_batch_stats_get(const tws_batch* batch, tws_batch_stats* out);
int tws_batch_sync(tws_batch* batch);

/* ---- autotuning ------------------------------------------------------------ */

typedef struct tws_tune_entry {
    int32_t kernel;             /* TWS_KERNEL_*                                */
    int32_t steps_per_launch;
    double cell_updates_per_s;
} tws_tune_entry;

/* Times every (kernel, steps per launch) candidate -- the six tile variants
 * and TWS_KERNEL_STREAM_128, each at k = 2, 3, 4, 5, 6 and 8 -- on scratch grids of
 * the n given shapes advanced together as one batch (synthetic state: seeded
 * fBm terrain under a 0.5 m water sheet, a rain disc of radius 16 at each
 * grid's centre), `steps` steps per trial after an untimed run of as many, and
 * returns the fastest in *best.  results[0..capacity) (may be NULL) receives
 * every trial; *count (may be NULL) their number.  Records nothing: the
 * launch TWS_KERNEL_AUTO picks changes only through tws_tune_record.  Pass
 * the step count of the runs being tuned for: a count that is not a multiple
 * of k ends in a partial launch, which costs a full launch's memory traffic.
 * Allocates and frees its own device memory; never touches existing
 * simulations. */
int tws_autotune(const int32_t* widths, const int32_t* heights, int32_t n, int32_t steps,
                 tws_tune_entry* best, tws_tune_entry* results, int32_t capacity,
                 int32_t* count);
/* Enters (kernel, steps_per_launch) for the size class (log2 of the cell
 * count) of a width x height grid: TWS_KERNEL_AUTO (tws_create,
 * tws_set_launch) picks it for grids of that class from then on in this
 * process (a streaming launch steps grids it cannot step on the tile
 * table's launch, as tws_set_launch does).  TWS_ERR_INVALID for an unknown
 * kernel or k. */
int tws_tune_record(int32_t width, int32_t height, int32_t kernel, int32_t steps_per_launch);
/* Drops every recorded launch: TWS_KERNEL_AUTO follows the built-in table. */
int tws_tune_forget(void);

/* ---- serving: a request loop over a batch --------------------------------- */
/*
 * No counterpart in the reference.  A server owns a batch and a thread.
 * Submitted solves (a grid, its initial state, a step count) wait in a queue;
 * the thread admits them in arrival order while their footprint
 * (tws_grid_footprint) fits the pool, keeps admitted state on the device,
 * and advances every resident request by min(remaining, chunk_steps) per
 * batched run, so requests that arrive while others run join at the next
 * chunk (continuous batching).  A finished request's outputs are copied to
 * the caller's buffers and its block returns to the pool.  Each result is
 * bit-identical to the same solve run alone with tws_step.
 * Input arrays are copied at submit; output buffers must stay valid until
 * the request is DONE, FAILED or CANCELLED.
 */
typedef struct tws_server tws_server;

typedef struct tws_server_config {
    int64_t pool_bytes;        /* device bytes for resident request state     */
    int32_t max_resident;      /* requests on the device at once              */
    int32_t chunk_steps;       /* steps per batched run (admission latency;
                                  0 = 48, whole launches of the batch default) */
    int32_t kernel;            /* TWS_KERNEL_* for the batch (0 = auto)       */
    int32_t steps_per_launch;  /* 0 = auto                                    */
    int32_t max_queue;         /* queued requests before submit refuses       */
} tws_server_config;

typedef struct tws_request {
    int32_t width, height;
    int32_t steps;
    tws_params params;
    const float* terrain;      /* width*height, required                      */
    const float* depth;        /* width*height, NULL = dry                    */
    const float* flux4;        /* width*height*4, NULL = zero                 */
    int32_t rain_on;           /* brush before every step when nonzero        */
    float rain[4];             /* centre x, centre y, intensity, radius^2     */
    float* out_depth;          /* outputs written when the request completes; */
    float* out_flux4;          /* any may be NULL                             */
    float* out_velocity2;
} tws_request;

#define TWS_REQ_QUEUED 0
#define TWS_REQ_RUNNING 1
#define TWS_REQ_DONE 2
#define TWS_REQ_FAILED 3
#define TWS_REQ_CANCELLED 4

typedef struct tws_request_status {
    int32_t state;             /* TWS_REQ_*                                   */
    int32_t error;             /* TWS_ERR_* when FAILED (NONFINITE: outputs
                                  were still written)                         */
    int32_t steps_done;
    int32_t pad_;
    double t_submit;           /* seconds on the server's monotonic clock     */
    double t_admit;
    double t_done;
} tws_request_status;

typedef struct tws_server_stats {
    int64_t submitted, completed, failed, cancelled;
    int64_t queued, resident;
    int64_t runs;              /* batched runs executed                       */
    int64_t cell_updates;      /* cell-updates completed                      */
    double busy_seconds;       /* wall time spent in batched runs             */
    int64_t pool_used, pool_bytes;
} tws_server_stats;

int tws_server_create(const tws_server_config* config, tws_server** out);
/* Queues a request; *ticket identifies it.  TWS_ERR_CAPACITY if the queue is
 * full or the grid can never fit the pool. */
int tws_server_submit(tws_server* server, const tws_request* request, uint64_t* ticket);
int tws_server_poll(tws_server* server, uint64_t ticket, tws_request_status* out);
/* Waits until the request leaves QUEUED/RUNNING or timeout_seconds pass
 * (TWS_ERR_STATE on timeout). */
int tws_server_wait(tws_server* server, uint64_t ticket, double timeout_seconds,
                    tws_request_status* out);
/* Cancels a queued or running request (it leaves the device at the next
 * chunk boundary; its outputs are not written). */
int tws_server_cancel(tws_server* server, uint64_t ticket);
/* Blocks until every submitted request has finished. */
int tws_server_drain(tws_server* server);
/* Forgets a finished request (DONE, FAILED or CANCELLED): its ticket is no
 * longer valid and the server frees its record.  TWS_ERR_STATE while it is
 * queued or running.  A long-lived server's callers release every ticket
 * they have consumed. */
int tws_server_release(tws_server* server, uint64_t ticket);
int tws_server_stats_get(tws_server* server, tws_server_stats* out);
/* Server monotonic clock, in the time base of tws_request_status. */
double tws_server_now(const tws_server* server);
/* Stops the thread (cancelling what is left) and frees everything. */
int tws_server_destroy(tws_server* server);

/* ---- initial state ---------------------------------------------------------- */

/* Seeded value-noise fBm heightfield into out[width*height] (host memory).
 * Replaces the terrain half of Terrain::CreateHeightmapFromNoiseAndResetSim
 * (Terrain.cpp:209-222: NoiseGenerator::GetValueNoise, quintic fade,
 * period-wrapped lattice, persistence 0.43) with a generator whose lattice is
 * splitmix64(seed, octave, i, j) instead of the reference's PRNG.  Octave o
 * has base_frequency*2^o lattice cells across the domain and amplitude
 * persistence^o; the normalised sum in [0,1] is scaled by height_scale. */
int tws_generate_fbm_terrain(int32_t width, int32_t height, uint64_t seed, int32_t octaves,
                             int32_t base_frequency, float persistence, float height_scale,
                             float* out);

/* The reference's own default scene, bit for bit: the terrain and lake of
 * Terrain::CreateHeightmapFromNoiseAndResetSim (Terrain.cpp:200-222) with its
 * NoiseGenerator (NoiseGenerator.cpp:6-97: 4096 white-noise values, 3-D value
 * noise over octaves 2..10, persistence 0.43, periodic) fed by its Mersenne-
 * Twister variant (Random.cpp:22-56) seeded with `seed` (Application.cpp:57
 * uses 231656522).  scene_index n gives the scene the reference builds on its
 * (n+1)-th call, after n NoiseGenerators have drawn 4096 values each.
 * terrain and depth (may be NULL) receive resolution^2 floats, row-major
 * (x + y*resolution, as Terrain.cpp:216-220).  The lake's pow(r^2, 2) is
 * taken as the correctly rounded square.  Host memory, threaded. */
int tws_generate_reference_scene(int32_t resolution, float height_scale, uint32_t seed,
                                 int32_t scene_index, float* terrain, float* depth);

/* The same fBm computed on the device straight into a simulation's terrain
 * plane (bit-identical to tws_generate_fbm_terrain), and a device-side water
 * initialiser: depth `depth` everywhere except rect_depth inside
 * [x0, x1) x [y0, y1) (an empty rectangle for a uniform sheet; a column for
 * a dam break), flux zero.  Replace the host loops of
 * Terrain::CreateHeightmapFromNoiseAndResetSim (Terrain.cpp:209-234) with no
 * host-to-device traffic.  Asynchronous on the handle's stream. */
int tws_init_terrain_fbm(tws_sim* sim, uint64_t seed, int32_t octaves, int32_t base_frequency,
                         float persistence, float height_scale);
int tws_init_water(tws_sim* sim, float depth, int32_t x0, int32_t y0, int32_t x1, int32_t y1,
                   float rect_depth);
int tws_batch_init_terrain_fbm(tws_batch* batch, int32_t grid_id, uint64_t seed, int32_t octaves,
                               int32_t base_frequency, float persistence, float height_scale);
int tws_batch_init_water(tws_batch* batch, int32_t grid_id, float depth, int32_t x0, int32_t y0,
                         int32_t x1, int32_t y1, float rect_depth);

/* The reference's default scene (as tws_generate_reference_scene, the same
 * bits) computed on the device into a square grid: terrain, lake depth and
 * zero flux, replacing the host loop and the uploads of
 * Terrain::CreateHeightmapFromNoiseAndResetSim (Terrain.cpp:209-234).  Only
 * the 4096-value noise table is drawn on the host; it is staged through the
 * grid's idle flux set, so grids must be at least 16 x 16.  Synchronises the
 * handle's stream.  TWS_ERR_INVALID for a non-square grid, a NaN
 * height_scale or scene_index < 0. */
int tws_init_reference_scene(tws_sim* sim, float height_scale, uint32_t seed,
                             int32_t scene_index);
int tws_batch_init_reference_scene(tws_batch* batch, int32_t grid_id, float height_scale,
                                   uint32_t seed, int32_t scene_index);

/* ---- diagnostics ---------------------------------------------------------- */

/* Lane-layout check of the step kernel's cross-lane moves: lane i of one
 * wavefront holds i + 1; out[i] receives what lane i reads from its left
 * neighbour lane and out[64 + i] what it reads from its right neighbour lane
 * (0 where there is no such lane). */
int tws_debug_lane_shift_probe(int32_t* out128);

/* Device-to-device STREAM copy of `bytes` (read + write), `iterations`
 * kernel launches on the default stream, in launch shape `shape` (0 <= shape
 * < tws_stream_copy_shapes(): loads in flight per lane, non-temporal hints,
 * work groups per CU; shape 0 keeps one load per lane in flight), or the
 * fastest of all shapes for shape = -1.  *gbps receives (2 * bytes *
 * iterations) / time: the measured HBM ceiling next to the 8 TB/s spec. */
int tws_stream_copy_bench(int64_t bytes, int32_t iterations, int32_t shape, double* gbps);
int tws_stream_copy_shapes(void);

const char* tws_last_error(void);
int tws_abi_version(void);
/* Number of HIP devices visible (0 on a machine without a GPU). */
int tws_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* TWS_H */
