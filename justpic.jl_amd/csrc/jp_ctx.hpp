// jp_ctx.hpp -- host-side context shared by the entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>
#include <unordered_map>

#include "jp_common.hpp"

// FP32-split geometry verdicts this context has computed: the axes they are
// for (pointers and lengths) and the state of the verdict
struct GeomKey {
    int32_t ndim;
    int32_t with_vel;
    int64_t ni[3];
    jp_axis xv[3];
    jp_axis vel[3][3];
};
struct GeomEntry {
    GeomKey key;
    int state = 0;          // 0 empty, 1 accepted (read back), 2 launched under capture
    bool standard = false;  // r == 1 and o == (c == d ? 0 : 1/2) exactly
    bool pinned = false;    // read by a captured graph: never reused for other axes
    uint64_t gen = 0;       // jp_free generation when checked
    uint64_t used = 0;      // last use (least recently used slot is reused)
};
constexpr int kGeomSlots = 256;

struct jp_ctx {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    int sync = 1;
    bool capturing = false;
    std::string err;
    // small device area: [0] jp_stats, [64..] int flags
    void* dsmall = nullptr;
    // growable scratch (never reallocated during capture)
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    int num_cu = 256;
    // launch options (jp_ctx_set_option); 0 = chosen from the problem size
    int opt_span = 0;
    int opt_move_phases = 0;  // JP_OPT_MOVE_PHASES
    int opt_advect_slots = 0;  // JP_OPT_ADVECT_SLOTS
    int64_t opt_inject_waves = 0;
    int opt_inject_phases = 0;  // JP_OPT_INJECT_PHASES
    // split geometry: device verdicts (kGeomSlots SplitGeom), the host side
    // of each, the slot of the current call, JP_OPT_GEOMETRY_RECHECK
    void* geo_dev = nullptr;
    GeomEntry geo_cache[kGeomSlots];
    std::unordered_map<uint64_t, int> geo_index;  // hash of the key -> slot
    int geo_cur = 0;
    uint64_t geo_clock = 0;
    int opt_geom_recheck = 0;
    // device address of this GPU's fixed scratch array (JP_NOALLOC_MAX_CELLS
    // words), looked up once at context creation
    uint64_t* noalloc = nullptr;
};

namespace jp {

inline int fail(jp_ctx* ctx, int code, const char* fmt, const char* a = "", long long b = 0) {
    if (ctx) {
        char buf[512];
        snprintf(buf, sizeof buf, fmt, a, b);
        ctx->err = buf;
    }
    return code;
}

#define JP_HIP(ctx, call)                                                          \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess)                                                      \
            return ::jp::fail((ctx), JP_ERR_HIP, "HIP error %s (line %lld)",       \
                              hipGetErrorString(e_), (long long)__LINE__);         \
    } while (0)

#define JP_LAUNCHED(ctx) JP_HIP(ctx, hipGetLastError())

inline jp_stats* dstats(jp_ctx* ctx) { return reinterpret_cast<jp_stats*>(ctx->dsmall); }
inline int* dflags(jp_ctx* ctx) {
    return reinterpret_cast<int*>(static_cast<char*>(ctx->dsmall) + 256);
}
// the sticky rejection word of asynchronous split calls lives here
inline SplitGeom* dgeom(jp_ctx* ctx) {
    return reinterpret_cast<SplitGeom*>(static_cast<char*>(ctx->dsmall) + 1024);
}
// the verdict (and affine maps) of the current split call's axes
inline SplitGeom* geo_ptr(jp_ctx* ctx) {
    return static_cast<SplitGeom*>(ctx->geo_dev) + ctx->geo_cur;
}
// the current call's axes were accepted and have the standard staggering
inline bool geo_standard(const jp_ctx* ctx) {
    const GeomEntry& e = ctx->geo_cache[ctx->geo_cur];
    return e.state == 1 && e.standard;
}
// incremented by every jp_free (any context): verdicts of older axes are not reused
uint64_t free_generation();

// semilagrangian_advection: [0] set by this call's kernel when a backtracked
// point left the grid, [1] the same, sticky until reported
inline int* drange(jp_ctx* ctx) {
    return reinterpret_cast<int*>(static_cast<char*>(ctx->dsmall) + 3072);
}
// report (and clear) an out-of-grid backtrack as JP_ERR_ARG
int range_report(jp_ctx* ctx);
// set by the ordered 2D injection when a dependency wait gave up
inline int* dinject(jp_ctx* ctx) {
    return reinterpret_cast<int*>(static_cast<char*>(ctx->dsmall) + 3584);
}
// report (and clear) such a timeout as JP_ERR_STATE (synchronises)
int inject_wait_report(jp_ctx* ctx);
// Ordered injections on one device run one at a time: each one's wavefronts
// wait on each other and are sized to what the device holds at once, so two
// at once on different streams could each hold slots the other's waits need.
// gate_enter (before the launch) makes the context stream wait for the
// previous ordered launch on the device when that was on another stream;
// gate_leave (after it) records an event on this stream (now, while it
// exists) once more than one stream has used ordered launches on the device
// (until then a switch of streams waits for the device instead).  Not
// applied under graph capture.
int ordered_gate_enter(jp_ctx* ctx);
int ordered_gate_leave(jp_ctx* ctx);
// report one left by asynchronous semilagrangian calls
int range_sticky(jp_ctx* ctx);

// FP32-split entry points: the geometry check of the vertex axes (and, when vg
// is not null, the affine maps of the velocity axes) into geo_ptr(ctx).  Axes
// this context accepted before (same pointers and lengths, no jp_free since)
// are not checked again.  Otherwise the check runs and its verdict is read
// back: rejected axes return JP_ERR_ARG before any kernel of the call runs.
// Under graph capture the verdict cannot be read: the call's kernels skip all
// work when it fails, and the next jp_sync reports it.
int split_geometry(jp_ctx* ctx, const jp_cellgrid* g, const jp_velgrid* vg);
// finish() of a split entry point: a synchronous call reports a rejected
// grid as JP_ERR_ARG; an asynchronous one leaves it in the sticky word, which
// the next jp_sync or jp_memcpy_d2h reports.
int split_finish(jp_ctx* ctx);
// report (and clear) a rejection left by asynchronous split calls
int geom_report(jp_ctx* ctx);

// scratch of at least `bytes`; grows outside capture only
int ensure_scratch(jp_ctx* ctx, size_t bytes);
// Injection work lists: per-wavefront segments from the scan, joined per
// colour into list[q*cap + i] with total[q] entries (no atomics).
struct InjectLists {
    int32_t* total;
    int32_t* cnt;
    int32_t* seg;
    int32_t* list;
    int64_t nwave;
    int64_t cap;
    uint64_t* occ;  // occupancy word of every cell
    void* box;      // 2*D planes of C: per-cell min, max of each coordinate (2D)
    uint32_t* state;  // per cell, written by the scan: 1 listed, 0 not (ordered 2D injection)
};
int inject_lists(jp_ctx* ctx, int64_t C, int nq, int64_t cap, size_t box_bytes, InjectLists& L);
int join_segments(jp_ctx* ctx, int nq, const InjectLists& L);
// workgroups (4 wavefronts each) of persistent wavefronts for a work list of
// at most n entries: one wavefront per entry up to 4096
inline int64_t inject_blocks(const jp_ctx* ctx, int64_t n) {
    const int64_t w = ctx->opt_inject_waves > 0 ? ctx->opt_inject_waves : (n < 4096 ? n : 4096);
    return (w + 3) / 4 > 0 ? (w + 3) / 4 : 1;
}
// occupancy words (bit k = slot k active) of the C cells, one kernel
int occ_words(jp_ctx* ctx, const uint8_t* index, int64_t C, int K, uint64_t* occ);
// zero `n` ints/bytes with a kernel (no memset nodes inside graphs)
int zero_bytes(jp_ctx* ctx, void* p, size_t bytes);
// stats bracket: zero the device stats before, copy + accumulate after
int stats_begin(jp_ctx* ctx, jp_stats* host);
int stats_end(jp_ctx* ctx, jp_stats* host);
int finish(jp_ctx* ctx);  // synchronise when the context is synchronous

inline int64_t blocks_for(int64_t n, int b = kBlock) { return (n + b - 1) / b; }

// Slots per slot group (a multiple of 8) for a per-cell kernel of `blocks`
// workgroups over K slots: groups are added (blockIdx.y) until the launch has
// about `target` workgroups (default 8 per CU of the 256), at most one group
// per 8 slots.
// JP_OPT_SLOT_SPAN, when set, replaces the choice.
inline int slot_span(const jp_ctx* ctx, int64_t blocks, int K, int64_t target = 2048) {
    if (ctx->opt_span > 0) return ctx->opt_span;
    const int64_t chunks = (K + 7) / 8;
    int64_t g = (target + blocks - 1) / (blocks > 0 ? blocks : 1);
    g = g < 1 ? 1 : (g > chunks ? chunks : g);
    const int64_t per = (chunks + g - 1) / g;
    return (int)(per * 8);
}

int check_grid(jp_ctx* ctx, const jp_cellgrid* g);

template <int D>
Grid<D> to_grid(const jp_cellgrid* g) {
    Grid<D> G;
    G.C = 1;
    for (int d = 0; d < D; ++d) {
        G.ni[d] = g->ni[d];
        G.C *= g->ni[d];
        G.xv[d] = Axis{g->xv[d].x, g->xv[d].dx, g->xv[d].inv_dx, g->xv[d].n};
    }
    G.K = g->max_xcell;
    return G;
}

}  // namespace jp
