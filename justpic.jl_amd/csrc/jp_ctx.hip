// jp_ctx.hip -- context, memory, events and graph capture of the C-ABI.
#include "jp_ctx.hpp"

#include <atomic>
#include <mutex>

#include <cstring>

using namespace jp;

namespace {

__global__ void k_fill_u8(uint8_t* p, uint8_t v, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_f64(double* p, double v, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
// 64-bit occupancy word of every cell (bit k = mask of slot k)
__global__ void __launch_bounds__(kBlock) k_occ_words(const uint8_t* __restrict__ index, int64_t C,
                                                      int K, uint64_t* __restrict__ occ) {
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= C) return;
    occ[c] = active_slots(index, c, C, K);
}

// Joins the per-wavefront list segments of colour q = blockIdx.y into
// list[q*cap + i], in wavefront order, and writes the colour's total.  Each
// workgroup owns kJoinSeg consecutive segments: it sums the counts of all
// earlier segments (a few KB, read from L2) for its base offset, scans its own
// counts in one wavefront, and copies with one wavefront per segment (lane i
// moves entry i), all loads of a wavefront issued before its stores.
constexpr int kJoinSeg = 64;
constexpr int kJoinBlock = 256;
__global__ void __launch_bounds__(kJoinBlock) k_join_segments(const int32_t* __restrict__ seg,
                                                              const int32_t* __restrict__ cnt,
                                                              int64_t nwave,
                                                              int32_t* __restrict__ list,
                                                              int32_t* __restrict__ total,
                                                              int64_t cap) {
    __shared__ int32_t s_part[kJoinBlock / 64];
    __shared__ int32_t s_off[kJoinSeg], s_cnt[kJoinSeg];
    const int q = blockIdx.y, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int32_t* __restrict__ cq = cnt + q * nwave;
    const int64_t s0 = (int64_t)blockIdx.x * kJoinSeg;
    const int64_t s1 = s0 + kJoinSeg < nwave ? s0 + kJoinSeg : nwave;
    // base: counts of segments [0, s0)
    int32_t acc = 0;
#pragma unroll 8
    for (int64_t w = t; w < s0; w += kJoinBlock) acc += cq[w];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) acc += __shfl_xor(acc, m, 64);
    if (lane == 0) s_part[wv] = acc;
    // this workgroup's counts, exclusive scan in wavefront 0
    if (wv == 0) {
        const int32_t n = s0 + lane < s1 ? cq[s0 + lane] : 0;
        int32_t incl = n;
#pragma unroll
        for (int m = 1; m < 64; m <<= 1) {
            const int32_t v = __shfl_up(incl, m, 64);
            if (lane >= m) incl += v;
        }
        s_cnt[lane] = n;
        s_off[lane] = incl - n;
    }
    __syncthreads();
    int32_t base = 0;
#pragma unroll
    for (int i = 0; i < kJoinBlock / 64; ++i) base += s_part[i];
    if (t == 0 && s1 == nwave) total[q] = base + s_off[kJoinSeg - 1] + s_cnt[kJoinSeg - 1];
    // copy: wavefront wv moves segments s0 + wv, s0 + wv + 4, ...
    constexpr int kPer = kJoinSeg / (kJoinBlock / 64);
    int32_t* __restrict__ lq = list + q * cap;
    int32_t v[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int i = wv + j * (kJoinBlock / 64);
        const int64_t w = s0 + i;
        v[j] = (w < s1 && lane < s_cnt[i]) ? seg[(q * nwave + w) * 64 + lane] : 0;
    }
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int i = wv + j * (kJoinBlock / 64);
        if (s0 + i < s1 && lane < s_cnt[i]) lq[base + s_off[i] + lane] = v[j];
    }
}

__global__ void k_fill_f32(float* p, float v, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// CellArray layouts: the device (B = 0) data of C cells x K slots is element
// (c, k) at g[c + C k]; the CPU CellArray data (1, K, C) is element (k, c) at
// h[k + K c], so converting is a transpose (src/CellArrays/conversion.jl:32-43
// permutes the data by (3, 2, 1)).  One workgroup per 64 cells: the slot planes
// are read 64 cells at a time (coalesced), the block's K x 64 output is one
// contiguous run of the CPU layout.
template <typename T>
__global__ void __launch_bounds__(256) k_to_cpu_layout(const T* __restrict__ g, T* __restrict__ h,
                                                       int64_t C, int K) {
    __shared__ T tile[64][65];
    const int64_t c0 = (int64_t)blockIdx.x * 64;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t nc = C - c0 < 64 ? C - c0 : 64;
    for (int k = w; k < K; k += 4)
        if (lane < nc) tile[lane][k] = g[c0 + lane + C * (int64_t)k];
    __syncthreads();
    const int64_t total = nc * K;
    for (int64_t i = threadIdx.x; i < total; i += 256) {
        const int cl = (int)(i / K), k = (int)(i - (int64_t)cl * K);
        h[K * c0 + i] = tile[cl][k];
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_from_cpu_layout(const T* __restrict__ h,
                                                         T* __restrict__ g, int64_t C, int K) {
    __shared__ T tile[64][65];
    const int64_t c0 = (int64_t)blockIdx.x * 64;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t nc = C - c0 < 64 ? C - c0 : 64;
    const int64_t total = nc * K;
    for (int64_t i = threadIdx.x; i < total; i += 256) {
        const int cl = (int)(i / K), k = (int)(i - (int64_t)cl * K);
        tile[cl][k] = h[K * c0 + i];
    }
    __syncthreads();
    for (int k = w; k < K; k += 4)
        if (lane < nc) g[c0 + lane + C * (int64_t)k] = tile[lane][k];
}

template <typename T>
void launch_layout(bool to_cpu, const void* src, void* dst, int64_t C, int K, hipStream_t s) {
    const unsigned blocks = (unsigned)((C + 63) / 64);
    if (to_cpu)
        k_to_cpu_layout<T><<<blocks, 256, 0, s>>>(static_cast<const T*>(src), static_cast<T*>(dst),
                                                 C, K);
    else
        k_from_cpu_layout<T><<<blocks, 256, 0, s>>>(static_cast<const T*>(src),
                                                   static_cast<T*>(dst), C, K);
}

int64_t fill_grid(int64_t n) {
    int64_t b = blocks_for(n);
    return b < 1 ? 1 : (b > 65536 ? 65536 : b);
}

std::atomic<uint64_t> g_free_gen{1};

// Scratch of the allocation-free entry point (jp_particle2grid_noalloc): one
// occupancy word per cell, JP_NOALLOC_MAX_CELLS words (16 MiB) per GPU, held
// until the process ends.  Host code never passes this name to a kernel: it
// uses the device address hipGetSymbolAddress returns (jp_ctx::noalloc).
static __device__ __align__(256) uint64_t scratch[JP_NOALLOC_MAX_CELLS];

}  // namespace

namespace jp {

uint64_t free_generation() { return g_free_gen.load(std::memory_order_relaxed); }


int ensure_scratch(jp_ctx* ctx, size_t bytes) {
    if (bytes <= ctx->scratch_bytes) return JP_OK;
    if (ctx->capturing)
        return fail(ctx, JP_ERR_STATE,
                    "scratch of %s%lld bytes needed during graph capture; run the "
                    "captured sequence once before capturing",
                    "", (long long)bytes);
    if (ctx->scratch) JP_HIP(ctx, hipFree(ctx->scratch));
    ctx->scratch = nullptr;
    ctx->scratch_bytes = 0;
    size_t sz = (bytes + 4095) & ~(size_t)4095;
    JP_HIP(ctx, hipMalloc(&ctx->scratch, sz));
    ctx->scratch_bytes = sz;
    return JP_OK;
}

int occ_words(jp_ctx* ctx, const uint8_t* index, int64_t C, int K, uint64_t* occ) {
    k_occ_words<<<blocks_for(C), kBlock, 0, ctx->stream>>>(index, C, K, occ);
    JP_LAUNCHED(ctx);
    return JP_OK;
}

int inject_lists(jp_ctx* ctx, int64_t C, int nq, int64_t cap, size_t box_bytes, InjectLists& L) {
    // scratch: totals | counts | segments | lists | occupancy words | boxes | states
    L.nwave = (C + 63) / 64;
    const size_t a = 256;
    const size_t b = a + (((size_t)nq * L.nwave * 4 + 255) & ~(size_t)255);
    const size_t c = b + (size_t)nq * L.nwave * 64 * 4;
    const size_t d = c + (((size_t)nq * cap * 4 + 255) & ~(size_t)255);
    const size_t e = d + (((size_t)C * 8 + 255) & ~(size_t)255);
    const size_t f = e + ((box_bytes + 255) & ~(size_t)255);
    const size_t h = f + (size_t)C * 4;
    int rc = ensure_scratch(ctx, h);
    if (rc) return rc;
    char* base = static_cast<char*>(ctx->scratch);
    L.total = reinterpret_cast<int32_t*>(base);
    L.cnt = reinterpret_cast<int32_t*>(base + a);
    L.seg = reinterpret_cast<int32_t*>(base + b);
    L.list = reinterpret_cast<int32_t*>(base + c);
    L.occ = reinterpret_cast<uint64_t*>(base + d);
    L.box = box_bytes ? static_cast<void*>(base + e) : nullptr;
    L.state = reinterpret_cast<uint32_t*>(base + f);
    L.cap = cap;
    return JP_OK;
}

int join_segments(jp_ctx* ctx, int nq, const InjectLists& L) {
    const int64_t nb = (L.nwave + kJoinSeg - 1) / kJoinSeg;
    const dim3 grid((unsigned)(nb > 0 ? nb : 1), (unsigned)nq);  // one block writes total = 0
    k_join_segments<<<grid, kJoinBlock, 0, ctx->stream>>>(L.seg, L.cnt, L.nwave, L.list, L.total,
                                                          L.cap);
    JP_LAUNCHED(ctx);
    return JP_OK;
}

int zero_bytes(jp_ctx* ctx, void* p, size_t bytes) {
    if (!bytes) return JP_OK;
    int64_t n = (int64_t)bytes;
    k_fill_u8<<<fill_grid(n), kBlock, 0, ctx->stream>>>((uint8_t*)p, 0, n);
    JP_LAUNCHED(ctx);
    return JP_OK;
}

int stats_begin(jp_ctx* ctx, jp_stats* host) {
    if (!host) return JP_OK;
    if (ctx->capturing)
        return fail(ctx, JP_ERR_STATE, "stats%s cannot be returned during graph capture");
    return zero_bytes(ctx, ctx->dsmall, sizeof(jp_stats));
}

int stats_end(jp_ctx* ctx, jp_stats* host) {
    if (!host) return JP_OK;
    jp_stats tmp;
    JP_HIP(ctx, hipMemcpyAsync(&tmp, ctx->dsmall, sizeof tmp, hipMemcpyDeviceToHost, ctx->stream));
    JP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    host->moved += tmp.moved;
    host->deleted += tmp.deleted;
    host->dropped += tmp.dropped;
    host->far += tmp.far;
    host->injected += tmp.injected;
    host->orphan += tmp.orphan;
    return JP_OK;
}

namespace {
struct OrderedGate {
    std::mutex m;
    hipEvent_t ev[64] = {};
    hipStream_t last[64] = {};
    bool used[64] = {};
    bool multi[64] = {};  // a second stream has used ordered launches here
};
OrderedGate& ordered_gate() {
    static OrderedGate g;  // process lifetime (the events are never destroyed)
    return g;
}
}  // namespace

int ordered_gate_enter(jp_ctx* ctx) {
    if (ctx->capturing) return JP_OK;
    if (ctx->device < 0 || ctx->device >= 64) return fail(ctx, JP_ERR_STATE, "device index%s");
    OrderedGate& g = ordered_gate();
    g.m.lock();
    const int d = ctx->device;
    hipError_t e = hipSuccess;
    if (g.used[d] && g.last[d] != ctx->stream) {
        if (g.multi[d]) {
            // wait for the event recorded after the previous ordered launch
            // (on its stream, at its launch; valid after that stream is gone)
            e = hipStreamWaitEvent(ctx->stream, g.ev[d], 0);
        } else {
            // the first switch of streams on this device: no event was
            // recorded (a single stream pays nothing), so wait for the device
            // to finish what it has, then record from now on
            e = hipDeviceSynchronize();
            if (e == hipSuccess) e = hipEventCreateWithFlags(&g.ev[d], hipEventDisableTiming);
            if (e == hipSuccess) g.multi[d] = true;
        }
    }
    if (e != hipSuccess) {
        g.m.unlock();
        return fail(ctx, JP_ERR_HIP, "HIP error %s", hipGetErrorString(e));
    }
    return JP_OK;  // (the lock is held until ordered_gate_leave)
}

int ordered_gate_leave(jp_ctx* ctx) {
    if (ctx->capturing) return JP_OK;
    OrderedGate& g = ordered_gate();
    const int d = ctx->device;
    // recorded on this stream now, while it exists
    const hipError_t e = g.multi[d] ? hipEventRecord(g.ev[d], ctx->stream) : hipSuccess;
    if (e == hipSuccess) {
        g.last[d] = ctx->stream;
        g.used[d] = true;
    }
    g.m.unlock();
    return e == hipSuccess ? JP_OK : fail(ctx, JP_ERR_HIP, "HIP error %s", hipGetErrorString(e));
}

int inject_wait_report(jp_ctx* ctx) {
    int v = 0;
    JP_HIP(ctx, hipMemcpyAsync(&v, dinject(ctx), sizeof v, hipMemcpyDeviceToHost, ctx->stream));
    JP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (!v) return JP_OK;
    int rc = zero_bytes(ctx, dinject(ctx), sizeof v);
    if (rc) return rc;
    return fail(ctx, JP_ERR_STATE,
                "inject_particles!: a cell's wait for an earlier-colour neighbour timed out%s");
}

int finish(jp_ctx* ctx) {
    if (ctx->sync && !ctx->capturing) JP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return JP_OK;
}

int check_grid(jp_ctx* ctx, const jp_cellgrid* g) {
    if (!ctx) return JP_ERR_STATE;
    if (!g) return fail(ctx, JP_ERR_ARG, "null grid%s");
    if (g->ndim != 2 && g->ndim != 3)
        return fail(ctx, JP_ERR_ARG, "The dimension of the problem must be either 2 or 3%s (got %lld)",
                    "", g->ndim);
    if (g->max_xcell < 1 || g->max_xcell > 64)
        return fail(ctx, JP_ERR_ARG, "max_xcell must be in 1..64%s (got %lld)", "", g->max_xcell);
    for (int d = 0; d < g->ndim; ++d) {
        if (g->ni[d] < 1) return fail(ctx, JP_ERR_ARG, "ni must be positive%s");
        if (g->xv[d].n != g->ni[d] + 1)
            return fail(ctx, JP_ERR_ARG, "vertex axis length must be ni+1%s (dim %lld)", "", d);
        if (!g->xv[d].x || !g->xv[d].dx || !g->xv[d].inv_dx)
            return fail(ctx, JP_ERR_ARG, "null vertex axis pointer%s (dim %lld)", "", d);
    }
    int64_t S = g->max_xcell;
    for (int d = 0; d < g->ndim; ++d) S *= g->ni[d];
    if (S > ((int64_t)1 << 40)) return fail(ctx, JP_ERR_ARG, "too many slots%s");
    return JP_OK;
}

}  // namespace jp

extern "C" {

int jp_ctx_create(jp_ctx** out, int device) {
    if (!out) return JP_ERR_ARG;
    *out = nullptr;
    jp_ctx* ctx = new jp_ctx();
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->own, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&ctx->dsmall, 4096);
    if (e == hipSuccess) e = hipMemset(ctx->dsmall, 0, 4096);
    if (e == hipSuccess) e = hipMalloc(&ctx->geo_dev, kGeomSlots * sizeof(SplitGeom));
    if (e == hipSuccess) e = hipMemset(ctx->geo_dev, 0, kGeomSlots * sizeof(SplitGeom));
    // this device's copy of the fixed scratch array (device `device` is current)
    if (e == hipSuccess) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        ctx->noalloc = static_cast<uint64_t*>(p);
        if (e == hipSuccess && !p) e = hipErrorInvalidSymbol;
    }
    if (e == hipSuccess) {
        int cu = 0;
        if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
            cu > 0)
            ctx->num_cu = cu;
    }
    if (e != hipSuccess) {
        delete ctx;
        return JP_ERR_HIP;
    }
    ctx->stream = ctx->own;
    *out = ctx;
    return JP_OK;
}

int jp_ctx_destroy(jp_ctx* ctx) {
    if (!ctx) return JP_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->own) (void)hipStreamSynchronize(ctx->own);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    if (ctx->dsmall) (void)hipFree(ctx->dsmall);
    if (ctx->geo_dev) (void)hipFree(ctx->geo_dev);
    if (ctx->own) (void)hipStreamDestroy(ctx->own);
    delete ctx;
    return JP_OK;
}

const char* jp_last_error(const jp_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int jp_ctx_set_sync(jp_ctx* ctx, int synchronous) {
    if (!ctx) return JP_ERR_STATE;
    ctx->sync = synchronous ? 1 : 0;
    return JP_OK;
}

int jp_ctx_set_stream(jp_ctx* ctx, void* s) {
    if (!ctx) return JP_ERR_STATE;
    if (ctx->capturing) return fail(ctx, JP_ERR_STATE, "cannot change stream while capturing%s");
    ctx->stream = s ? static_cast<hipStream_t>(s) : ctx->own;
    return JP_OK;
}

void* jp_ctx_stream(jp_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int jp_ctx_set_option(jp_ctx* ctx, int option, int64_t value) {
    if (!ctx) return JP_ERR_STATE;
    if (ctx->capturing) return fail(ctx, JP_ERR_STATE, "cannot change options while capturing%s");
    switch (option) {
    case JP_OPT_SLOT_SPAN:
        if (value < 0 || value > 64 || value % 8)
            return fail(ctx, JP_ERR_ARG, "slot span must be 0 or a multiple of 8 up to 64%s");
        ctx->opt_span = (int)value;
        return JP_OK;
    case JP_OPT_INJECT_WAVES:
        if (value < 0 || value > (1 << 20))
            return fail(ctx, JP_ERR_ARG, "injection wavefronts must be in 0..2^20%s");
        ctx->opt_inject_waves = value;
        return JP_OK;
    case JP_OPT_GEOMETRY_RECHECK:
        if (value < 0 || value > 1)
            return fail(ctx, JP_ERR_ARG, "geometry recheck must be 0 or 1%s");
        ctx->opt_geom_recheck = (int)value;
        return JP_OK;
    case JP_OPT_INJECT_PHASES:
        if (value < 0 || value > 1) return fail(ctx, JP_ERR_ARG, "inject phases must be 0 or 1%s");
        ctx->opt_inject_phases = (int)value;
        return JP_OK;
    case JP_OPT_MOVE_PHASES:
        if (value < 0 || value > 1) return fail(ctx, JP_ERR_ARG, "move phases must be 0 or 1%s");
        ctx->opt_move_phases = (int)value;
        return JP_OK;
    case JP_OPT_ADVECT_SLOTS:
        if (value < 0 || value > 1)
            return fail(ctx, JP_ERR_ARG, "advection slots option must be 0 or 1%s");
        ctx->opt_advect_slots = (int)value;
        return JP_OK;
    default:
        return fail(ctx, JP_ERR_ARG, "unknown option%s %lld", "", option);
    }
}

// errors that asynchronous calls leave in device words: a split grid
// rejection, a semi-Lagrangian backtrack that left the grid
static int async_report(jp_ctx* ctx) {
    if (ctx->capturing) return JP_OK;
    int rc = geom_report(ctx);
    return rc ? rc : range_sticky(ctx);
}

int jp_sync(jp_ctx* ctx) {
    if (!ctx) return JP_ERR_STATE;
    JP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return async_report(ctx);
}

int jp_malloc(jp_ctx* ctx, void** p, size_t bytes) {
    if (!ctx || !p) return JP_ERR_ARG;
    *p = nullptr;
    if (!bytes) bytes = 1;
    JP_HIP(ctx, hipMalloc(p, bytes));
    return JP_OK;
}

int jp_free(jp_ctx* ctx, void* p) {
    if (!ctx) return JP_ERR_STATE;
    if (p) {
        // memory that held an axis may come back with other coordinates
        g_free_gen.fetch_add(1, std::memory_order_relaxed);
        JP_HIP(ctx, hipFree(p));
    }
    return JP_OK;
}

int jp_memcpy_h2d(jp_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return JP_ERR_STATE;
    if (!bytes) return JP_OK;
    JP_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    JP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return JP_OK;
}

int jp_memcpy_d2h(jp_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return JP_ERR_STATE;
    if (!bytes) return JP_OK;
    JP_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    JP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return async_report(ctx);
}

int jp_memcpy_d2d(jp_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return JP_ERR_STATE;
    if (!bytes) return JP_OK;
    JP_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return finish(ctx);
}

int jp_fill_u8(jp_ctx* ctx, uint8_t* p, uint8_t v, int64_t n) {
    if (!ctx) return JP_ERR_STATE;
    if (n <= 0) return JP_OK;
    k_fill_u8<<<fill_grid(n), kBlock, 0, ctx->stream>>>(p, v, n);
    JP_LAUNCHED(ctx);
    return finish(ctx);
}

int jp_fill_f64(jp_ctx* ctx, double* p, double v, int64_t n) {
    if (!ctx) return JP_ERR_STATE;
    if (n <= 0) return JP_OK;
    k_fill_f64<<<fill_grid(n), kBlock, 0, ctx->stream>>>(p, v, n);
    JP_LAUNCHED(ctx);
    return finish(ctx);
}

int jp_fill_f32(jp_ctx* ctx, float* p, float v, int64_t n) {
    if (!ctx) return JP_ERR_STATE;
    if (n <= 0) return JP_OK;
    k_fill_f32<<<fill_grid(n), kBlock, 0, ctx->stream>>>(p, v, n);
    JP_LAUNCHED(ctx);
    return finish(ctx);
}

static int cellarray_layout(jp_ctx* ctx, bool to_cpu, const void* src, void* dst, int64_t C,
                            int K, int elem_bytes) {
    if (!ctx) return JP_ERR_STATE;
    if (!src || !dst) return fail(ctx, JP_ERR_ARG, "null layout buffer%s");
    if (C < 0 || K < 1 || K > 64) return fail(ctx, JP_ERR_ARG, "layout needs C >= 0 and 1 <= K <= 64%s");
    if (src == dst) return fail(ctx, JP_ERR_ARG, "layout conversion cannot run in place%s");
    if (C == 0) return JP_OK;
    switch (elem_bytes) {
    case 1: launch_layout<uint8_t>(to_cpu, src, dst, C, K, ctx->stream); break;
    case 4: launch_layout<uint32_t>(to_cpu, src, dst, C, K, ctx->stream); break;
    case 8: launch_layout<uint64_t>(to_cpu, src, dst, C, K, ctx->stream); break;
    default: return fail(ctx, JP_ERR_ARG, "element size must be 1, 4 or 8 bytes%s");
    }
    JP_LAUNCHED(ctx);
    return finish(ctx);
}

int jp_cellarray_to_cpu_layout(jp_ctx* ctx, const void* src, void* dst, int64_t ncells, int K,
                               int elem_bytes) {
    return cellarray_layout(ctx, true, src, dst, ncells, K, elem_bytes);
}

int jp_cellarray_from_cpu_layout(jp_ctx* ctx, const void* src, void* dst, int64_t ncells, int K,
                                 int elem_bytes) {
    return cellarray_layout(ctx, false, src, dst, ncells, K, elem_bytes);
}

int jp_event_create(jp_ctx* ctx, void** ev) {
    if (!ctx || !ev) return JP_ERR_ARG;
    hipEvent_t e;
    JP_HIP(ctx, hipEventCreate(&e));
    *ev = e;
    return JP_OK;
}

int jp_event_destroy(jp_ctx* ctx, void* ev) {
    if (!ctx) return JP_ERR_STATE;
    if (ev) JP_HIP(ctx, hipEventDestroy(static_cast<hipEvent_t>(ev)));
    return JP_OK;
}

int jp_event_record(jp_ctx* ctx, void* ev) {
    if (!ctx || !ev) return JP_ERR_ARG;
    JP_HIP(ctx, hipEventRecord(static_cast<hipEvent_t>(ev), ctx->stream));
    return JP_OK;
}

int jp_event_wait(jp_ctx* ctx, void* ev) {
    if (!ctx || !ev) return JP_ERR_ARG;
    JP_HIP(ctx, hipStreamWaitEvent(ctx->stream, static_cast<hipEvent_t>(ev), 0));
    return JP_OK;
}

int jp_event_elapsed_ms(jp_ctx* ctx, void* a, void* b, float* ms) {
    if (!ctx || !a || !b || !ms) return JP_ERR_ARG;
    JP_HIP(ctx, hipEventSynchronize(static_cast<hipEvent_t>(b)));
    JP_HIP(ctx, hipEventElapsedTime(ms, static_cast<hipEvent_t>(a), static_cast<hipEvent_t>(b)));
    return JP_OK;
}

int jp_graph_begin(jp_ctx* ctx) {
    if (!ctx) return JP_ERR_STATE;
    if (ctx->capturing) return fail(ctx, JP_ERR_STATE, "already capturing%s");
    if (ctx->stream == nullptr)
        return fail(ctx, JP_ERR_STATE, "graph capture needs a non-null stream%s");
    JP_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->capturing = true;
    return JP_OK;
}

int jp_graph_end(jp_ctx* ctx, void** exec) {
    if (!ctx || !exec) return JP_ERR_ARG;
    if (!ctx->capturing) return fail(ctx, JP_ERR_STATE, "not capturing%s");
    ctx->capturing = false;
    hipGraph_t graph;
    JP_HIP(ctx, hipStreamEndCapture(ctx->stream, &graph));
    hipGraphExec_t ge;
    hipError_t e = hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    JP_HIP(ctx, e);
    *exec = ge;
    return JP_OK;
}

int jp_graph_launch(jp_ctx* ctx, void* exec) {
    if (!ctx || !exec) return JP_ERR_ARG;
    JP_HIP(ctx, hipGraphLaunch(static_cast<hipGraphExec_t>(exec), ctx->stream));
    return finish(ctx);
}

int jp_graph_destroy(jp_ctx* ctx, void* exec) {
    if (!ctx) return JP_ERR_STATE;
    if (exec) JP_HIP(ctx, hipGraphExecDestroy(static_cast<hipGraphExec_t>(exec)));
    return JP_OK;
}

void jp_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    philox4x32_10(c, key[0], key[1]);
    for (int i = 0; i < 4; ++i) out[i] = c[i];
}

}  // extern "C"
