// tws_batch.hip -- many independent grids advanced by one launch sequence.
//
// Each grid is a GridState carved from the batch's device pool.  Grids fall
// into size classes (by cell count), each with its own tile kernel and steps
// per launch.  A run flattens the tiles of each class's grids with pending
// steps into one work list (a segment of the GridDesc table plus a {grids,
// tiles, offset} header in device memory, ordered by tile_begin) and launches
// ceil(max_steps / k) tile kernels for it; launch j moves grid g by
// min(k, steps_g - j*k) steps, so grids with fewer steps simply drop out of
// the later launches.  The classes' launch chains are independent (disjoint
// grids) and are captured as parallel branches of one hipGraph.  The chains
// read their tables at run time, so a captured graph of a shape (per class:
// kernel, k, launches, block bound, mode) is replayed for any table that fits it.
// A class on the streaming kernel has a chain of its own too: its grids'
// strip segments are the tasks of one table (k_stream's batch mode), and the
// chain advances them in full launches of k steps.  The segment cap is
// planned for the class's grid shapes (stream_batch_cap) and kept while they
// stay the same.  A grid's remaining steps (fewer than k), and grids the
// streaming kernel cannot step, run as direct launches after the graph.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
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

// A size class: grids of at most max_cells cells not taken by an earlier
// class.  kernel TWS_KERNEL_AUTO (k 0) picks its launch per run
// (resolve_class).
struct LaunchClass {
    int64_t max_cells;
    int kernel, k;
};

// One class's launch chain in a run.  A streaming chain also ends in tail
// launches (bit r of `tails`: one launch of r steps over the grids whose step
// count leaves r after the full launches; work groups: tail_blocks) and in
// tile launches for its grids the streaming kernel cannot step (alt_*: the
// fallback tile launch's count, work groups and mode).
struct Chain {
    int cls, kernel, k, launches, blocks, mode, tails, tail_blocks;
    int alt_launches, alt_blocks, alt_mode;
    bool operator<(const Chain& o) const {
        return std::tie(cls, kernel, k, launches, blocks, mode, tails, tail_blocks, alt_launches,
                        alt_blocks, alt_mode) <
               std::tie(o.cls, o.kernel, o.k, o.launches, o.blocks, o.mode, o.tails,
                        o.tail_blocks, o.alt_launches, o.alt_blocks, o.alt_mode);
    }
};
using GraphKey = std::vector<Chain>;

constexpr int kHdrInts = 4;  // per class: grids, tiles, table offset, unused

// Tail launches of a streaming chain run the kernel's built step counts below
// 7; a remainder of 7 (k = 8) steps on direct launches.  A streaming class's
// grids that the streaming kernel cannot step (edges other than closed walls,
// planes of 2 GiB or more) take the fallback tile launch in the same chain.
// Headers: the classes' own, kMaxTail per class for its tails, one per class
// for its fallback table.
constexpr int kMaxTail = 6;
constexpr int kAltKernel = TWS_KERNEL_TILE_128x64, kAltK = 4;
constexpr int kHdrCount = TWS_BATCH_MAX_CLASSES * (2 + kMaxTail);
inline int tail_hdr(int cls, int r) { return TWS_BATCH_MAX_CLASSES + cls * kMaxTail + (r - 1); }
inline int alt_hdr(int cls) { return TWS_BATCH_MAX_CLASSES * (1 + kMaxTail) + cls; }

// A streaming class's segment cap, kept while the shapes it was planned for
// (W, H of the chain's grids in table order) and k stay the same.
struct CapPlan {
    std::vector<int> shapes;
    int k = 0, cap = 0;
};

}  // namespace

struct tws_batch {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    DevicePool pool;
    int max_grids = 0;
    std::vector<Slot> slots;
    GridDesc* table_dev = nullptr;   // 2 max_grids entries (a grid and its tail)
    int* hdr_dev = nullptr;          // kHdrCount headers {grids, tiles, offset, 0}
    unsigned* status_dev = nullptr;  // one word per slot
    double* partials_dev = nullptr;
    GridDesc* table_host = nullptr;  // pinned
    int* hdr_host = nullptr;         // pinned
    hipEvent_t table_copied = nullptr;
    std::vector<LaunchClass> classes;
    CapPlan caps[TWS_BATCH_MAX_CLASSES];
    hipStream_t aux[TWS_BATCH_MAX_CLASSES - 1] = {};  // branch streams of the chains
    hipEvent_t fork = nullptr;
    hipEvent_t join[TWS_BATCH_MAX_CLASSES - 1] = {};
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

// Issues the chains of one run: the first on the batch stream, the others on
// branch streams forked from it and joined back, so they may run side by side.
static hipError_t issue_chains(tws_batch* b, const GraphKey& chains) {
    hipError_t e = hipSuccess;
    const int n = (int)chains.size();
    if (n > 1) {
        e = hipEventRecord(b->fork, b->stream);
        for (int c = 1; c < n && e == hipSuccess; ++c) e = hipStreamWaitEvent(b->aux[c - 1], b->fork, 0);
    }
    for (int c = 0; c < n && e == hipSuccess; ++c) {
        const Chain& ch = chains[c];
        hipStream_t st = c == 0 ? b->stream : b->aux[c - 1];
        const int* hdr = b->hdr_dev + kHdrInts * ch.cls;
        for (int j = 0; j < ch.launches && e == hipSuccess; ++j)
            e = ch.kernel == TWS_KERNEL_STREAM_128
                    ? launch_stream_batch(ch.k, ch.mode, b->table_dev, hdr, ch.blocks, j, st)
                    : launch_tile_batch(ch.kernel, ch.k, ch.mode, b->table_dev, hdr, ch.blocks, j, st);
        for (int r = 1; r <= kMaxTail && e == hipSuccess; ++r)
            if ((ch.tails >> r) & 1)
                e = launch_stream_batch(r, ch.mode, b->table_dev, b->hdr_dev + kHdrInts * tail_hdr(ch.cls, r),
                                        ch.tail_blocks, 0, st);
        for (int j = 0; j < ch.alt_launches && e == hipSuccess; ++j)
            e = launch_tile_batch(kAltKernel, kAltK, ch.alt_mode, b->table_dev,
                                  b->hdr_dev + kHdrInts * alt_hdr(ch.cls), ch.alt_blocks, j, st);
    }
    for (int c = 1; c < n && e == hipSuccess; ++c) {
        e = hipEventRecord(b->join[c - 1], b->aux[c - 1]);
        if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->join[c - 1], 0);
    }
    return e;
}

// Launches (or replays) the tile kernels of one run.
static int run_launches(tws_batch* b, const GraphKey& chains) {
    if (!b->use_graphs) {
        TWS_HIP(issue_chains(b, chains), "batch: tile kernels");
        return TWS_OK;
    }
    auto it = b->graphs.find(chains);
    if (it == b->graphs.end()) {
        if (b->graphs.size() >= 32) drop_graphs(b);
        hipGraph_t graph = nullptr;
        TWS_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal),
                "batch: begin capture");
        hipError_t e = issue_chains(b, chains);
        hipError_t e2 = hipStreamEndCapture(b->stream, &graph);
        if (e != hipSuccess) { if (graph) (void)hipGraphDestroy(graph); return hip_fail(e, "batch: capture"); }
        if (e2 != hipSuccess) return hip_fail(e2, "batch: end capture");
        hipGraphExec_t exec = nullptr;
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return hip_fail(e, "batch: instantiate graph");
        try {
            it = b->graphs.emplace(chains, exec).first;
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

// Index of the class that steps a grid of `cells` cells.
static int class_of(const tws_batch* b, int64_t cells) {
    const int n = (int)b->classes.size();
    for (int c = 0; c + 1 < n; ++c)
        if (cells <= b->classes[c].max_cells) return c;
    return n - 1;
}

// Whether slot s has steps pending in class c of the run.
static bool pending_in(const tws_batch* b, const Slot& s, int c) {
    return s.used && s.steps > 0 && class_of(b, (int64_t)s.g.W * s.g.H) == c;
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
    e = hipMalloc(&b->table_dev, sizeof(GridDesc) * 2 * (size_t)max_grids);
    if (e == hipSuccess) e = hipMalloc(&b->hdr_dev, sizeof(int) * kHdrInts * kHdrCount);
    if (e == hipSuccess) e = hipMalloc(&b->status_dev, sizeof(unsigned) * (size_t)max_grids);
    if (e == hipSuccess) e = hipMalloc(&b->partials_dev, sizeof(double) * kVolumeParts);
    if (e == hipSuccess) e = hipHostMalloc(&b->table_host, sizeof(GridDesc) * 2 * (size_t)max_grids);
    if (e == hipSuccess) e = hipHostMalloc(&b->hdr_host, sizeof(int) * kHdrInts * kHdrCount);
    if (e == hipSuccess) std::memset(b->hdr_host, 0, sizeof(int) * kHdrInts * kHdrCount);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->table_copied, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMemsetAsync(b->status_dev, 0, sizeof(unsigned) * (size_t)max_grids, b->stream);
    if (e == hipSuccess) e = hipEventRecord(b->table_copied, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) { tws_batch_destroy(b); return hip_fail(e, "tws_batch_create: buffers"); }
    int rc = tws_batch_set_launch(b, TWS_KERNEL_AUTO, 0);
    if (rc != TWS_OK) { tws_batch_destroy(b); return rc; }
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
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (int c = 0; c < TWS_BATCH_MAX_CLASSES - 1; ++c) {
        if (b->join[c]) (void)hipEventDestroy(b->join[c]);
        if (b->aux[c]) (void)hipStreamDestroy(b->aux[c]);
    }
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

// Whether the streaming class's full launches step this grid: the streaming
// kernel applies and the grid has at least one full launch of k steps.
static bool in_stream_chain(const Slot& s, int k) { return s.steps >= k && stream_usable(s.g); }

// Whether a grid of a streaming class of k steps per launch ends in a tail
// launch (rather than a direct launch) for the steps short of a full launch.
static bool in_stream_tail(const Slot& s, int k) {
    const int r = s.steps % k;
    return r > 0 && r <= kMaxTail && stream_usable(s.g);
}

// The launch class c runs this run.  An AUTO class takes the streaming kernel
// at k = 6 when its grids with pending steps that the streaming kernel can
// step hold more than 2^23 cells (2896^2) together.  MI355X, ms per 100 steps
// with a rain disc, streaming k = 6 vs 128x64 tiles at k = 4: 20 x 512^2
// 2.22 vs 2.01, 1 x 2304^2 2.16 vs 1.81; at 2^23 cells 32 x 512^2 2.92 vs
// 2.87, 8 x 1024^2 2.77 vs 2.86, 2 x 2048^2 2.89 vs 2.97; 1 x 3072^2 2.80 vs
// 3.06, 1 x 4096^2 4.51 vs 4.87; 256 grids of 128^2..2047^2 4.21e11 vs
// 3.84e11 cell-updates/s, and the saturated served mix 4.25e11 vs 3.79e11.
// Otherwise it takes the tile table's launch for one square grid of all its
// grids' cells (the rounds of tiles are what the table is tuned on): 1 x
// 256^2 0.20 ms vs 0.60, 64 x 128^2 0.74 vs 0.97, 16 x 256^2 0.70 vs 0.89.
static constexpr int64_t kAutoStreamCells = (int64_t)1 << 23;
static LaunchClass resolve_class(const tws_batch* b, int c) {
    LaunchClass lc = b->classes[c];
    if (lc.kernel != TWS_KERNEL_AUTO) return lc;
    int64_t cells = 0, all = 0;
    for (const Slot& s : b->slots)
        if (pending_in(b, s, c)) {
            all += (int64_t)s.g.W * s.g.H;
            if (stream_usable(s.g)) cells += (int64_t)s.g.W * s.g.H;
        }
    if (cells > kAutoStreamCells) {
        lc.kernel = TWS_KERNEL_STREAM_128;
        lc.k = 6;
    } else {
        const int side = (int)std::sqrt((double)std::max<int64_t>(all, 1));
        tile_launch(side, side, &lc.kernel, &lc.k);
    }
    return lc;
}

// Segment cap of streaming class c's chain in this run (0 if it cannot be
// planned for lack of host memory).
static int chain_cap(tws_batch* b, int c, int k) {
    try {
        std::vector<int> shapes;
        for (int i = 0; i < b->max_grids; ++i) {
            const Slot& s = b->slots[i];
            if (!pending_in(b, s, c) || !in_stream_chain(s, k)) continue;
            shapes.push_back(s.g.W);
            shapes.push_back(s.g.H);
        }
        CapPlan& cp = b->caps[c];
        if (cp.k != k || cp.cap <= 0 || cp.shapes != shapes) {
            cp.cap = 0;
            cp.cap = stream_batch_cap(shapes.data(), (int)shapes.size() / 2, k, device_cus());
            cp.k = k;
            cp.shapes.swap(shapes);
        }
        return cp.cap;
    } catch (const std::bad_alloc&) {
        return 0;
    }
}

// Writes header h = {grids, tasks or tiles, table offset, 0}.
static void set_header(tws_batch* b, int h, int grids, int tasks, int offset) {
    int* p = b->hdr_host + kHdrInts * h;
    p[0] = grids;
    p[1] = tasks;
    p[2] = offset;
    p[3] = 0;
}

// Appends the tail table of streaming class c for remainder r at
// table_host[*n]: the grids left with r steps after their full launches of
// lc.k, each entry advanced past those launches (parity) and holding exactly
// r steps, for one launch of r levels.  *tasks gets the table's task count;
// *mode is raised to its grids' mode.
static int append_stream_tail(tws_batch* b, int c, const LaunchClass& lc, int cap, int r, int* n,
                              int* tasks, int* mode) {
    const int first = *n;
    *tasks = 0;
    for (int i = 0; i < b->max_grids; ++i) {
        const Slot& s = b->slots[i];
        if (!pending_in(b, s, c) || s.steps % lc.k != r || !stream_usable(s.g)) continue;
        GridDesc g = s.g.desc();
        if (!stream_batch_plan(s.g.W, s.g.H, r, cap, &g.tiles_x, &g.tiles_y, &g.seg_rows))
            return fail(TWS_ERR_INVALID, "tws_batch_run: bad launch config");
        g.parity0 ^= (s.steps / lc.k) & 1;
        g.steps = r;
        g.tile_begin = *tasks;
        *tasks += g.tiles_x * g.tiles_y;
        *mode = std::max(*mode, g.rain_on ? kModeRain : kModePlain);
        b->table_host[(*n)++] = g;
    }
    set_header(b, tail_hdr(c, r), *n - first, *tasks, first);
    return TWS_OK;
}

// Appends the fallback tile table of streaming class c at table_host[*n]:
// its grids the streaming kernel cannot step, as a tile class of kAltKernel
// at kAltK would step them.  *max_steps, *tiles and *mode describe the table.
static int append_alt_table(tws_batch* b, int c, int* n, int* max_steps, int* tiles, int* mode) {
    const int first = *n;
    *max_steps = *tiles = 0;
    *mode = kModePlain;
    for (int i = 0; i < b->max_grids; ++i) {
        const Slot& s = b->slots[i];
        if (!pending_in(b, s, c) || stream_usable(s.g)) continue;
        GridDesc g = s.g.desc();
        if (!grid_tiles(s.g.W, s.g.H, kAltKernel, kAltK, &g.tiles_x, &g.tiles_y))
            return fail(TWS_ERR_INVALID, "tws_batch_run: bad launch config");
        *mode = std::max(*mode, kernel_mode(g));
        *max_steps = std::max(*max_steps, s.steps);
        g.tile_begin = *tiles;
        g.steps = s.steps;
        *tiles += g.tiles_x * g.tiles_y;
        b->table_host[(*n)++] = g;
    }
    set_header(b, alt_hdr(c), *n - first, *tiles, first);
    return TWS_OK;
}

int tws_batch_run(tws_batch* b) {
    if (!b) return fail(TWS_ERR_INVALID, "tws_batch_run: NULL handle");
    // the pinned table may still be in flight from the previous run
    TWS_HIP(hipEventSynchronize(b->table_copied), "tws_batch_run: wait table");
    const int nc = (int)b->classes.size();
    LaunchClass run_class[TWS_BATCH_MAX_CLASSES];
    for (int c = 0; c < nc; ++c) run_class[c] = resolve_class(b, c);
    int n = 0;
    bool direct = false;  // grids of streaming classes with steps outside the chains
    GraphKey chains;
    for (int c = 0; c < nc; ++c) {
        const LaunchClass& lc = run_class[c];
        const bool stream = lc.kernel == TWS_KERNEL_STREAM_128;
        const int first = n;
        int tiles = 0, max_steps = 0, mode = kModePlain;
        const int cap = stream ? chain_cap(b, c, lc.k) : 0;
        if (stream && cap <= 0) return fail(TWS_ERR_NOMEM, "tws_batch_run: host allocation failed");
        for (int i = 0; i < b->max_grids; ++i) {
            const Slot& s = b->slots[i];
            if (!pending_in(b, s, c)) continue;
            GridDesc g = s.g.desc();
            if (stream) {
                if (stream_usable(s.g) && s.steps % lc.k != 0 && !in_stream_tail(s, lc.k))
                    direct = true;
                if (!in_stream_chain(s, lc.k)) continue;
                if (!stream_batch_plan(s.g.W, s.g.H, lc.k, cap, &g.tiles_x, &g.tiles_y, &g.seg_rows))
                    return fail(TWS_ERR_INVALID, "tws_batch_run: bad launch config");
                mode = std::max(mode, g.rain_on ? kModeRain : kModePlain);
                max_steps = std::max(max_steps, s.steps / lc.k * lc.k);  // full launches
            } else {
                if (!grid_tiles(s.g.W, s.g.H, lc.kernel, lc.k, &g.tiles_x, &g.tiles_y))
                    return fail(TWS_ERR_INVALID, "tws_batch_run: bad launch config");
                mode = std::max(mode, kernel_mode(g));
                max_steps = std::max(max_steps, s.steps);
            }
            g.tile_begin = tiles;  // first tile, or first task of the streaming kernel
            g.steps = s.steps;
            tiles += g.tiles_x * g.tiles_y;
            b->table_host[n++] = g;
        }
        set_header(b, c, n - first, tiles, first);
        // block bounds rounded up so small changes of the tile (task) count
        // reuse a graph; spare blocks exit
        const int blocks = stream ? ((tiles + 3) / 4 + 63) / 64 * 64 : (tiles + 511) / 512 * 512;
        // a streaming class's tails (one launch per remainder r) and its
        // fallback tile table
        int tails = 0, tail_blocks = 0;
        int alt_steps = 0, alt_tiles = 0, alt_mode = kModePlain;
        if (stream) {
            for (int r = 1; r < lc.k && r <= kMaxTail; ++r) {
                int tasks = 0;
                const int rc = append_stream_tail(b, c, lc, cap, r, &n, &tasks, &mode);
                if (rc != TWS_OK) return rc;
                if (tasks == 0) continue;
                tails |= 1 << r;
                tail_blocks = std::max(tail_blocks, ((tasks + 3) / 4 + 63) / 64 * 64);
            }
            const int rc = append_alt_table(b, c, &n, &alt_steps, &alt_tiles, &alt_mode);
            if (rc != TWS_OK) return rc;
        }
        if (max_steps == 0 && tails == 0 && alt_steps == 0) continue;
        chains.push_back(Chain{c, lc.kernel, lc.k, (max_steps + lc.k - 1) / lc.k, blocks, mode,
                               tails, tail_blocks, (alt_steps + kAltK - 1) / kAltK,
                               (alt_tiles + 511) / 512 * 512, alt_mode});
    }
    if (n == 0 && !direct) return TWS_OK;
    if (n > 0) {
        TWS_HIP(hipMemcpyAsync(b->table_dev, b->table_host, sizeof(GridDesc) * (size_t)n,
                               hipMemcpyHostToDevice, b->stream), "tws_batch_run: table");
        TWS_HIP(hipMemcpyAsync(b->hdr_dev, b->hdr_host, sizeof(int) * kHdrInts * kHdrCount,
                               hipMemcpyHostToDevice, b->stream), "tws_batch_run: header");
    }
    TWS_HIP(hipEventRecord(b->table_copied, b->stream), "tws_batch_run: event");
    if (!chains.empty()) {
        const int rc = run_launches(b, chains);
        if (rc != TWS_OK) return rc;
    }
    for (int i = 0; i < b->max_grids; ++i) {
        Slot& s = b->slots[i];
        if (!s.used || s.steps <= 0) continue;
        const LaunchClass& lc = run_class[class_of(b, (int64_t)s.g.W * s.g.H)];
        if (lc.kernel != TWS_KERNEL_STREAM_128) {
            const int gl = (s.steps + lc.k - 1) / lc.k;
            s.g.parity ^= (gl & 1);
        } else {
            // the chain's full launches and its tail (a remainder of 7 then
            // on direct launches), or the fallback tile launches
            int rest = 0;
            if (stream_usable(s.g)) {
                s.g.parity ^= ((s.steps / lc.k) & 1);
                rest = s.steps % lc.k;
                if (in_stream_tail(s, lc.k)) {
                    s.g.parity ^= 1;
                    rest = 0;
                }
            } else {
                s.g.parity ^= ((s.steps + kAltK - 1) / kAltK) & 1;
            }
            if (rest > 0) {
                const int rc = step_grid(s.g, lc.kernel, lc.k, rest, b->stream, nullptr, 0, nullptr);
                if (rc != TWS_OK) return rc;
            }
        }
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
    // TWS_KERNEL_AUTO: one class that picks its launch per run (resolve_class)
    const int64_t max_cells = INT64_MAX;
    return tws_batch_set_launch_classes(b, 1, &max_cells, &kernel, &k);
}

int tws_batch_get_launch(const tws_batch* b, int32_t* kernel, int32_t* k) {
    if (!b) return fail(TWS_ERR_INVALID, "tws_batch_get_launch: NULL handle");
    if (kernel) *kernel = b->classes.back().kernel;
    if (k) *k = b->classes.back().k;
    return TWS_OK;
}

int tws_batch_set_launch_classes(tws_batch* b, int32_t n, const int64_t* max_cells,
                                 const int32_t* kernels, const int32_t* ks) {
    if (!b || !max_cells || !kernels || !ks)
        return fail(TWS_ERR_INVALID, "tws_batch_set_launch_classes: NULL argument");
    if (n < 1 || n > TWS_BATCH_MAX_CLASSES)
        return fail(TWS_ERR_INVALID, "tws_batch_set_launch_classes: %d classes (1..%d)", n,
                    TWS_BATCH_MAX_CLASSES);
    for (int c = 0; c < n; ++c) {
        TileShape sh{};
        int ox = 0;
        const bool ok = kernels[c] == TWS_KERNEL_AUTO ||
                        (kernels[c] == TWS_KERNEL_STREAM_128 ? ks[c] >= 2 && stream_shape(ks[c], &ox)
                                                             : tile_shape(kernels[c], ks[c], &sh));
        if (!ok)
            return fail(TWS_ERR_INVALID, "tws_batch_set_launch_classes: class %d: kernel %d / k %d "
                        "not available for batches", c, kernels[c], ks[c]);
        if (c + 1 < n && (max_cells[c] <= 0 || (c > 0 && max_cells[c] <= max_cells[c - 1])))
            return fail(TWS_ERR_INVALID, "tws_batch_set_launch_classes: class bounds must be "
                        "positive and ascending");
    }
    // branch streams and events for the chains beyond the first, made once
    for (int c = 0; c + 1 < n; ++c) {
        if (!b->aux[c]) TWS_HIP(hipStreamCreateWithFlags(&b->aux[c], hipStreamNonBlocking),
                                "tws_batch_set_launch_classes: stream");
        if (!b->join[c]) TWS_HIP(hipEventCreateWithFlags(&b->join[c], hipEventDisableTiming),
                                 "tws_batch_set_launch_classes: event");
    }
    if (n > 1 && !b->fork)
        TWS_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming),
                "tws_batch_set_launch_classes: event");

    std::vector<LaunchClass> cl;
    try {
        for (int c = 0; c < n; ++c)
            cl.push_back(LaunchClass{c + 1 < n ? max_cells[c] : INT64_MAX, kernels[c],
                                     kernels[c] == TWS_KERNEL_AUTO ? 0 : ks[c]});
    } catch (...) {
        return fail(TWS_ERR_NOMEM, "tws_batch_set_launch_classes: host allocation failed");
    }
    b->classes.swap(cl);
    return TWS_OK;
}

int tws_batch_get_launch_classes(const tws_batch* b, int32_t capacity, int64_t* max_cells,
                                 int32_t* kernels, int32_t* ks, int32_t* n) {
    if (!b || !n) return fail(TWS_ERR_INVALID, "tws_batch_get_launch_classes: NULL argument");
    const int nc = (int)b->classes.size();
    *n = nc;
    for (int c = 0; c < nc && c < capacity; ++c) {
        if (max_cells) max_cells[c] = b->classes[c].max_cells;
        if (kernels) kernels[c] = b->classes[c].kernel;
        if (ks) ks[c] = b->classes[c].k;
    }
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
