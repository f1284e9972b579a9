// tws_batch.hip -- many independent grids advanced by one launch sequence.
//
// Each grid is a GridState carved from the batch's device pool.  A run
// flattens the tiles of every grid with pending steps into one work list
// (GridDesc table + {grids, tiles} header in device memory, ordered by
// tile_begin) and launches ceil(max_steps / k) tile kernels; launch j moves
// grid g by min(k, steps_g - j*k) steps, so grids with fewer steps simply
// drop out of the later launches.  The launch sequence reads its table at
// run time, so the captured hipGraph of a (kernel, k, launches, block bound)
// shape is replayed for any table that fits it.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <new>
#include <tuple>
#include <vector>

#include "tws.h"
#include "tws_grid.hpp"
#include "tws_pool.hpp"
#include "tws_runtime.hpp"

using namespace tws;

namespace {

struct Slot {
    bool used = false;
    GridState g;
    size_t off = 0, bytes = 0;
    int steps = 0;  // pending for the next run
};

struct GraphKey {
    int kernel, k, launches, blocks, mode;
    bool operator<(const GraphKey& o) const {
        return std::tie(kernel, k, launches, blocks, mode) <
               std::tie(o.kernel, o.k, o.launches, o.blocks, o.mode);
    }
};

}  // namespace

struct tws_batch {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    DevicePool pool;
    int max_grids = 0;
    std::vector<Slot> slots;
    GridDesc* table_dev = nullptr;   // max_grids entries
    int* hdr_dev = nullptr;          // {grids, tiles}
    unsigned* status_dev = nullptr;  // one word per slot
    double* partials_dev = nullptr;
    GridDesc* table_host = nullptr;  // pinned
    int* hdr_host = nullptr;         // pinned
    hipEvent_t table_copied = nullptr;
    int kernel = TWS_KERNEL_TILE_128x64;
    int k = 4;
    bool use_graphs = true;
    std::map<GraphKey, hipGraphExec_t> graphs;
    int64_t captures = 0, replays = 0;
};

#define TWS_HIP(call, what)                              \
    do {                                                 \
        hipError_t e_ = (call);                          \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

static Slot* slot_of(tws_batch* b, int32_t id) {
    if (!b || id < 0 || id >= b->max_grids || !b->slots[id].used) return nullptr;
    return &b->slots[id];
}

namespace tws {
// Device state of a batch grid, for library code that drives transfers on a
// stream of its own (the server overlaps uploads and downloads of grids that
// are not in the running launch sequence with that sequence).
GridState* batch_grid(tws_batch* b, int32_t id) {
    Slot* s = slot_of(b, id);
    return s ? &s->g : nullptr;
}
hipStream_t batch_stream(tws_batch* b) { return b ? b->stream : nullptr; }
}  // namespace tws

static void drop_graphs(tws_batch* b) {
    for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    b->graphs.clear();
}

// Launches (or replays) the tile kernels of one run.
static int run_launches(tws_batch* b, int launches, int blocks, int mode) {
    if (!b->use_graphs) {
        for (int j = 0; j < launches; ++j)
            TWS_HIP(launch_tile_batch(b->kernel, b->k, mode, b->table_dev, b->hdr_dev, blocks,
                                      j, b->stream), "batch: tile kernel");
        return TWS_OK;
    }
    const GraphKey key{b->kernel, b->k, launches, blocks, mode};
    auto it = b->graphs.find(key);
    if (it == b->graphs.end()) {
        if (b->graphs.size() >= 32) drop_graphs(b);
        hipGraph_t graph = nullptr;
        TWS_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal),
                "batch: begin capture");
        hipError_t e = hipSuccess;
        for (int j = 0; j < launches && e == hipSuccess; ++j)
            e = launch_tile_batch(b->kernel, b->k, mode, b->table_dev, b->hdr_dev, blocks, j,
                                  b->stream);
        hipError_t e2 = hipStreamEndCapture(b->stream, &graph);
        if (e != hipSuccess) { if (graph) (void)hipGraphDestroy(graph); return hip_fail(e, "batch: capture"); }
        if (e2 != hipSuccess) return hip_fail(e2, "batch: end capture");
        hipGraphExec_t exec = nullptr;
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return hip_fail(e, "batch: instantiate graph");
        try {
            it = b->graphs.emplace(key, exec).first;
        } catch (...) {
            (void)hipGraphExecDestroy(exec);
            return fail(TWS_ERR_NOMEM, "batch: graph cache");
        }
        ++b->captures;
    }
    TWS_HIP(hipGraphLaunch(it->second, b->stream), "batch: graph launch");
    ++b->replays;
    return TWS_OK;
}

extern "C" {

int64_t tws_grid_footprint(int32_t width, int32_t height) {
    if (width <= 0 || height <= 0) return 0;
    return (int64_t)DevicePool::round(GridState::bytes_for(width, height));
}

int tws_batch_create(int32_t max_grids, int64_t pool_bytes, void* stream, tws_batch** out) {
    if (!out) return fail(TWS_ERR_INVALID, "tws_batch_create: out is NULL");
    *out = nullptr;
    if (max_grids <= 0 || max_grids > (1 << 20) || pool_bytes < (int64_t)DevicePool::kAlign)
        return fail(TWS_ERR_INVALID, "tws_batch_create: invalid limits");
    tws_batch* b = new (std::nothrow) tws_batch();
    if (!b) return fail(TWS_ERR_NOMEM, "tws_batch_create: host allocation failed");
    b->max_grids = max_grids;
    try {
        b->slots.resize((size_t)max_grids);
    } catch (...) {
        delete b;
        return fail(TWS_ERR_NOMEM, "tws_batch_create: host allocation failed");
    }
    hipError_t e = hipSuccess;
    if (stream) {
        b->stream = (hipStream_t)stream;
    } else {
        e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
        if (e != hipSuccess) { delete b; return hip_fail(e, "tws_batch_create: stream"); }
        b->own_stream = true;
    }
    char* base = nullptr;
    const size_t cap = DevicePool::round((size_t)pool_bytes);
    e = hipMalloc(&base, cap);
    if (e != hipSuccess) { tws_batch_destroy(b); return hip_fail(e, "tws_batch_create: pool"); }
    b->pool.reset(base, cap);
    e = hipMalloc(&b->table_dev, sizeof(GridDesc) * (size_t)max_grids);
    if (e == hipSuccess) e = hipMalloc(&b->hdr_dev, 256);
    if (e == hipSuccess) e = hipMalloc(&b->status_dev, sizeof(unsigned) * (size_t)max_grids);
    if (e == hipSuccess) e = hipMalloc(&b->partials_dev, sizeof(double) * kVolumeParts);
    if (e == hipSuccess) e = hipHostMalloc(&b->table_host, sizeof(GridDesc) * (size_t)max_grids);
    if (e == hipSuccess) e = hipHostMalloc(&b->hdr_host, 256);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->table_copied, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMemsetAsync(b->status_dev, 0, sizeof(unsigned) * (size_t)max_grids, b->stream);
    if (e == hipSuccess) e = hipEventRecord(b->table_copied, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) { tws_batch_destroy(b); return hip_fail(e, "tws_batch_create: buffers"); }
    *out = b;
    return TWS_OK;
}

int tws_batch_destroy(tws_batch* b) {
    if (!b) return TWS_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graphs(b);
    if (b->pool.base()) (void)hipFree(b->pool.base());
    if (b->table_dev) (void)hipFree(b->table_dev);
    if (b->hdr_dev) (void)hipFree(b->hdr_dev);
    if (b->status_dev) (void)hipFree(b->status_dev);
    if (b->partials_dev) (void)hipFree(b->partials_dev);
    if (b->table_host) (void)hipHostFree(b->table_host);
    if (b->hdr_host) (void)hipHostFree(b->hdr_host);
    if (b->table_copied) (void)hipEventDestroy(b->table_copied);
    if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return TWS_OK;
}

int tws_batch_add(tws_batch* b, int32_t width, int32_t height, const tws_params* params,
                  int32_t* grid_id) {
    if (!b || !grid_id) return fail(TWS_ERR_INVALID, "tws_batch_add: NULL argument");
    if (width <= 0 || height <= 0 || width > (1 << 20) || height > (1 << 20))
        return fail(TWS_ERR_INVALID, "tws_batch_add: invalid size %d x %d", width, height);
    if (!valid_params(params)) return fail(TWS_ERR_INVALID, "tws_batch_add: invalid params");
    if (!batchable(width, height, *params))
        return fail(TWS_ERR_INVALID, "tws_batch_add: a reference-dispatch grid of %d x %d (not "
                    "multiples of 32) steps on the single-grid path only", width, height);
    int id = -1;
    for (int i = 0; i < b->max_grids; ++i)
        if (!b->slots[i].used) { id = i; break; }
    if (id < 0) return fail(TWS_ERR_CAPACITY, "tws_batch_add: all %d grid slots in use", b->max_grids);
    const size_t bytes = GridState::bytes_for(width, height);
    const size_t off = b->pool.alloc(bytes);
    if (off == SIZE_MAX)
        return fail(TWS_ERR_CAPACITY, "tws_batch_add: %zu bytes do not fit (pool %zu, used %zu, "
                    "largest free %zu)", bytes, b->pool.capacity(), b->pool.used(),
                    b->pool.largest_free());
    Slot& s = b->slots[id];
    s = Slot();
    s.used = true;
    s.off = off;
    s.bytes = bytes;
    s.g.bind(reinterpret_cast<float*>(b->pool.base() + off), width, height);
    s.g.params = *params;
    s.g.status = b->status_dev + id;
    int rc = grid_clear(s.g, b->stream);
    if (rc == TWS_OK) {
        hipError_t e = hipMemsetAsync(s.g.status, 0, sizeof(unsigned), b->stream);
        if (e != hipSuccess) rc = hip_fail(e, "tws_batch_add: clear status");
    }
    if (rc != TWS_OK) {
        b->pool.release(off, bytes);
        s.used = false;
        return rc;
    }
    *grid_id = id;
    return TWS_OK;
}

int tws_batch_remove(tws_batch* b, int32_t id) {
    Slot* s = slot_of(b, id);
    if (!s) return fail(TWS_ERR_INVALID, "tws_batch_remove: unknown grid %d", id);
    // queued work may still use the block: the next user writes it in stream order
    b->pool.release(s->off, s->bytes);
    s->used = false;
    return TWS_OK;
}

int tws_batch_upload(tws_batch* b, int32_t id, const float* terrain, const float* depth,
                     const float* flux4) {
    Slot* s = slot_of(b, id);
    if (!s) return fail(TWS_ERR_INVALID, "tws_batch_upload: unknown grid %d", id);
    return grid_upload(s->g, terrain, depth, flux4, b->stream);
}

int tws_batch_download(tws_batch* b, int32_t id, float* depth, float* flux4, float* velocity2) {
    Slot* s = slot_of(b, id);
    if (!s) return fail(TWS_ERR_INVALID, "tws_batch_download: unknown grid %d", id);
    return grid_download(s->g, depth, flux4, velocity2, b->stream);
}

int tws_batch_init_terrain_fbm(tws_batch* b, int32_t id, uint64_t seed, int32_t octaves,
                               int32_t base_frequency, float persistence, float height_scale) {
    Slot* s = slot_of(b, id);
    if (!s) return fail(TWS_ERR_INVALID, "tws_batch_init_terrain_fbm: unknown grid %d", id);
    return grid_init_terrain_fbm(s->g, seed, octaves, base_frequency, persistence, height_scale,
                                 b->stream);
}

int tws_batch_init_water(tws_batch* b, int32_t id, float depth, int32_t x0, int32_t y0,
                         int32_t x1, int32_t y1, float rect_depth) {
    Slot* s = slot_of(b, id);
    if (!s) return fail(TWS_ERR_INVALID, "tws_batch_init_water: unknown grid %d", id);
    return grid_init_water(s->g, depth, x0, y0, x1, y1, rect_depth, b->stream);
}

int tws_batch_init_reference_scene(tws_batch* b, int32_t id, float height_scale, uint32_t seed,
                                   int32_t scene_index) {
    Slot* s = slot_of(b, id);
    if (!s) return fail(TWS_ERR_INVALID, "tws_batch_init_reference_scene: unknown grid %d", id);
    return grid_init_reference_scene(s->g, height_scale, seed, scene_index, b->stream);
}

int tws_batch_set_params(tws_batch* b, int32_t id, const tws_params* p) {
    Slot* s = slot_of(b, id);
    if (!s) return fail(TWS_ERR_INVALID, "tws_batch_set_params: unknown grid %d", id);
    if (!valid_params(p)) return fail(TWS_ERR_INVALID, "tws_batch_set_params: invalid params");
    if (!batchable(s->g.W, s->g.H, *p))
        return fail(TWS_ERR_INVALID, "tws_batch_set_params: a reference-dispatch grid of %d x %d "
                    "(not multiples of 32) steps on the single-grid path only", s->g.W, s->g.H);
    s->g.params = *p;
    return TWS_OK;
}

int tws_batch_set_rain(tws_batch* b, int32_t id, float cx, float cy, float I, float R2) {
    Slot* s = slot_of(b, id);
    if (!s) return fail(TWS_ERR_INVALID, "tws_batch_set_rain: unknown grid %d", id);
    return grid_set_rain(s->g, cx, cy, I, R2);
}

int tws_batch_volume(tws_batch* b, int32_t id, double* out) {
    Slot* s = slot_of(b, id);
    if (!s || !out) return fail(TWS_ERR_INVALID, "tws_batch_volume: bad argument");
    return grid_volume(s->g, b->partials_dev, out, b->stream);
}

int tws_batch_check_finite(tws_batch* b, int32_t id) {
    Slot* s = slot_of(b, id);
    if (!s) return fail(TWS_ERR_INVALID, "tws_batch_check_finite: unknown grid %d", id);
    return grid_check_finite(s->g, b->stream);
}

int tws_batch_set_steps(tws_batch* b, int32_t id, int32_t steps) {
    Slot* s = slot_of(b, id);
    if (!s || steps < 0) return fail(TWS_ERR_INVALID, "tws_batch_set_steps: bad argument");
    s->steps = steps;
    return TWS_OK;
}

int tws_batch_run(tws_batch* b) {
    if (!b) return fail(TWS_ERR_INVALID, "tws_batch_run: NULL handle");
    // the pinned table may still be in flight from the previous run
    TWS_HIP(hipEventSynchronize(b->table_copied), "tws_batch_run: wait table");
    int n = 0, tiles = 0, max_steps = 0;
    int mode = kModePlain;
    for (int i = 0; i < b->max_grids; ++i) {
        Slot& s = b->slots[i];
        if (!s.used || s.steps <= 0) continue;
        GridDesc g = s.g.desc();
        if (!grid_tiles(s.g.W, s.g.H, b->kernel, b->k, &g.tiles_x, &g.tiles_y))
            return fail(TWS_ERR_INVALID, "tws_batch_run: bad launch config");
        g.tile_begin = tiles;
        g.steps = s.steps;
        mode = std::max(mode, kernel_mode(g));
        tiles += g.tiles_x * g.tiles_y;
        max_steps = std::max(max_steps, s.steps);
        b->table_host[n++] = g;
    }
    if (n == 0) return TWS_OK;
    b->hdr_host[0] = n;
    b->hdr_host[1] = tiles;
    TWS_HIP(hipMemcpyAsync(b->table_dev, b->table_host, sizeof(GridDesc) * (size_t)n,
                           hipMemcpyHostToDevice, b->stream), "tws_batch_run: table");
    TWS_HIP(hipMemcpyAsync(b->hdr_dev, b->hdr_host, 2 * sizeof(int), hipMemcpyHostToDevice,
                           b->stream), "tws_batch_run: header");
    TWS_HIP(hipEventRecord(b->table_copied, b->stream), "tws_batch_run: event");
    const int launches = (max_steps + b->k - 1) / b->k;
    // block bound rounded up so small changes of the tile count reuse a graph
    const int blocks = (tiles + 511) / 512 * 512;
    int rc = run_launches(b, launches, blocks, mode);
    if (rc != TWS_OK) return rc;
    for (int i = 0; i < b->max_grids; ++i) {
        Slot& s = b->slots[i];
        if (!s.used || s.steps <= 0) continue;
        const int gl = (s.steps + b->k - 1) / b->k;
        s.g.parity ^= (gl & 1);
        s.steps = 0;
    }
    return TWS_OK;
}

int tws_batch_step(tws_batch* b, int32_t n) {
    if (!b || n < 0) return fail(TWS_ERR_INVALID, "tws_batch_step: bad argument");
    for (auto& s : b->slots)
        if (s.used) s.steps = n;
    return tws_batch_run(b);
}

int tws_batch_set_launch(tws_batch* b, int32_t kernel, int32_t k) {
    if (!b) return fail(TWS_ERR_INVALID, "tws_batch_set_launch: NULL handle");
    if (kernel == TWS_KERNEL_AUTO) {
        kernel = TWS_KERNEL_TILE_128x64;
        k = 4;
    }
    TileShape sh{};
    if (!tile_shape(kernel, k, &sh))
        return fail(TWS_ERR_INVALID, "tws_batch_set_launch: kernel %d / k %d not available for "
                    "batches", kernel, k);
    b->kernel = kernel;
    b->k = k;
    return TWS_OK;
}

int tws_batch_get_launch(const tws_batch* b, int32_t* kernel, int32_t* k) {
    if (!b) return fail(TWS_ERR_INVALID, "tws_batch_get_launch: NULL handle");
    if (kernel) *kernel = b->kernel;
    if (k) *k = b->k;
    return TWS_OK;
}

int tws_batch_use_graphs(tws_batch* b, int32_t enable) {
    if (!b) return fail(TWS_ERR_INVALID, "tws_batch_use_graphs: NULL handle");
    b->use_graphs = enable != 0;
    return TWS_OK;
}

int tws_batch_stats_get(const tws_batch* b, tws_batch_stats* out) {
    if (!b || !out) return fail(TWS_ERR_INVALID, "tws_batch_stats_get: NULL argument");
    tws_batch_stats st{};
    for (const auto& s : b->slots)
        if (s.used) {
            ++st.grids;
            st.cells += (int64_t)s.g.W * s.g.H;
        }
    st.max_grids = b->max_grids;
    st.pool_bytes = (int64_t)b->pool.capacity();
    st.pool_used = (int64_t)b->pool.used();
    st.pool_largest_free = (int64_t)b->pool.largest_free();
    st.graph_captures = b->captures;
    st.graph_replays = b->replays;
    *out = st;
    return TWS_OK;
}

int tws_batch_sync(tws_batch* b) {
    if (!b) return fail(TWS_ERR_INVALID, "tws_batch_sync: NULL handle");
    TWS_HIP(hipStreamSynchronize(b->stream), "tws_batch_sync");
    return TWS_OK;
}

}  // extern "C"
