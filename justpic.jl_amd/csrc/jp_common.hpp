// jp_common.hpp -- device-side building blocks shared by the gfx950 kernels.
//
// All floating-point code here is compiled with -ffp-contract=off: every fused
// multiply-add is an explicit fma() placed where JustPIC.jl fuses (Base.fma,
// Base.muladd, MuladdMacro.@muladd); every other operation is one correctly
// rounded binary64 (or binary32 in the split kernels) operation in the
// reference's evaluation order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "justpic_hip.h"

namespace jp {

constexpr int kBlock = 256;

struct Axis {
    const double* x;
    const double* dx;
    const double* idx;  // inverse spacing
    int64_t n;
};

template <int D>
struct Grid {
    int64_t ni[D];
    int64_t C;  // cells
    int K;      // slots per cell
    Axis xv[D];
};

template <int D>
struct Vel {
    Axis ax[D][D];
};

struct Ptr3 {
    double* p[3];
};
struct CPtr3 {
    const double* p[3];
};
struct ArgPtrs {
    double* p[JP_MAX_ARGS];
};
// 0-based colour offsets of one phase
struct Off3 {
    int o[3];
};

// Geometry of the FP32-split form, written by the geometry kernel at the start
// of every split entry point that needs it and read by that call's kernels.
// A position is the cell I of its slot plus an offset u in cell units of the
// uniform vertex axes.  Velocity component c is sampled on a uniform axis
// along each dimension d whose node coordinate at that position is
// (I + u) r[c][d] + o[c][d]: r = vertex spacing / node spacing, o = (first
// vertex - first node) / node spacing.  The standard staggering has r = 1 and
// o = 0 (vertex axis) or 1/2 (centre axis with one ghost node per side); the
// README's expand_range rounds the ghost ends to 5 significant digits
// (README.md:48-55), which moves r and o slightly away from those values.
struct SplitGeom {
    int32_t bad;     // this call: 0 accepted, bit 0 vertex axis, bit 1 velocity axis
    int32_t sticky;  // OR of `bad` since the last report (asynchronous calls)
    double r[3][3];
    double o[3][3];
};
constexpr int kGeomVertex = 1;
constexpr int kGeomVelocity = 2;
// accepted deviation of an axis from x[0] + k h, relative to h
constexpr double kGeomTol = 1.0e-9;

// 0-based cell tuple <-> column-major linear index
template <int D>
__host__ __device__ inline int64_t lin(const int64_t (&ni)[D], const int64_t* I) {
    int64_t c = I[D - 1];
#pragma unroll
    for (int d = D - 2; d >= 0; --d) c = c * ni[d] + I[d];
    return c;
}

template <int D>
__host__ __device__ inline void unlin(const int64_t (&ni)[D], int64_t c, int64_t* I) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
        I[d] = c % ni[d];
        c /= ni[d];
    }
}

// lerp(t, v0, v1) = muladd(t, v1, muladd(-t, v0, v0)) -- src/Interpolations/ndlerp.jl:15
__host__ __device__ inline double lerp1(double t, double v0, double v1) {
    return fma(t, v1, fma(-t, v0, v0));
}
__host__ __device__ inline float lerp1f(float t, float v0, float v1) {
    return fmaf(t, v1, fmaf(-t, v0, v0));
}

// find_parent_cell_bisection (src/Utils.jl:108-130), 0-based: lo=0, hi=n-1.
// The iteration cap only guards memory safety; for p inside [x[0], x[n-1]]
// the loop terminates in O(log n) steps.  Returns -1 if it did not.
__device__ inline int64_t bisect_exact(double p, const double* __restrict__ x, int64_t n,
                                       int64_t s) {
    int64_t lo = 0, hi = n - 1;
    for (int it = 0; it < 160; ++it) {
        double a = x[s], b = x[s + 1];
        if (a <= p && p <= b) return s;
        if (a < p) {
            lo = s;
            s = (hi + s) >> 1;
        } else {
            hi = s;
            s = (lo + s) >> 1;
        }
    }
    return -1;
}

// Same result as bisect_exact for any p in [x[0], x[n-1]]: if p lies strictly
// inside one interval, that interval is the only one whose closed bracket
// contains p, so the reference loop must stop there whatever its seed; only
// ties (p equal to a grid value) depend on the search path, and those run the
// reference loop itself.  Candidates: the seed, its neighbours, and the index a
// uniform spacing would give.
// Out-of-line part (rare: displacement beyond one cell, refined axes, ties).
__device__ __noinline__ int64_t find_cell_slow(double p, const double* __restrict__ x, int64_t n,
                                               int64_t s) {
    const int64_t last = n - 2;
    double x0 = x[0], xn = x[n - 1];
    double fg = (p - x0) / (xn - x0) * (double)(n - 1);
    int64_t g = (int64_t)fg;
    g = g < 0 ? 0 : (g > last ? last : g);
    for (int o = 0; o < 3; ++o) {
        int64_t h = g + (o == 0 ? 0 : (o == 1 ? -1 : 1));
        if (h >= 0 && h <= last && x[h] < p && p < x[h + 1]) return h;
    }
    int64_t r = bisect_exact(p, x, n, s);
    return r < 0 ? 0 : r;
}

__device__ inline int64_t find_cell(double p, const Axis& a, int64_t seed) {
    const double* __restrict__ x = a.x;
    const int64_t last = a.n - 2;
    int64_t s = seed < last ? seed : last;
    double xs = x[s], xs1 = x[s + 1];
    if (xs < p && p < xs1) return s;
    if (p < xs) {
        if (s > 0) {
            double xm = x[s - 1];
            if (xm < p && p < xs) return s - 1;
        }
    } else if (p > xs1 && s < last) {
        double xp = x[s + 2];
        if (xs1 < p && p < xp) return s + 1;
    }
    return find_cell_slow(p, x, a.n, s);
}

// Row mapping of cell-parallel kernels: each wavefront owns 64 consecutive
// cells along x of one grid row (fixed j [,k]); blocks hold kBlock/64 rows.
template <int D>
__host__ __device__ inline int64_t row_count(const Grid<D>& g) {
    int64_t r = (g.ni[0] + 63) / 64;
    for (int d = 1; d < D; ++d) r *= g.ni[d];
    return r;
}
template <int D>
inline int64_t row_blocks(const Grid<D>& g) {
    return (row_count<D>(g) + kBlock / 64 - 1) / (kBlock / 64);
}
template <int D>
__device__ inline bool row_cell(const Grid<D>& g, int64_t& c, int64_t (&I)[D]) {
    const int lane = threadIdx.x & 63;
    // wave-uniform row index (decoded with scalar arithmetic)
    int64_t row = (int64_t)blockIdx.x * (kBlock / 64) +
                  __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (row >= row_count<D>(g)) return false;
    const int64_t nb = (g.ni[0] + 63) / 64;
    I[0] = (row % nb) * 64 + lane;
    row /= nb;
#pragma unroll
    for (int d = 1; d < D; ++d) {
        I[d] = row % g.ni[d];
        row /= g.ni[d];
    }
    if (I[0] >= g.ni[0]) return false;
    c = lin<D>(g.ni, I);
    return true;
}

// Active slots of cell c as a bitmask (loads coalesced across neighbouring
// cells of a wavefront in the cell-plane layout).
__device__ inline uint64_t active_slots(const uint8_t* __restrict__ index, int64_t c, int64_t C,
                                        int K) {
    // 32 loads per round, issued together: the slot index is clamped to K-1
    // (always a valid address) instead of guarding each load, since a guarded
    // load is waited for on its own
    uint64_t bits = 0;
#pragma unroll
    for (int k0 = 0; k0 < 64; k0 += 32) {
        if (k0 >= K) break;
        uint8_t m[32];
#pragma unroll
        for (int v = 0; v < 32; ++v) {
            const int k = k0 + v < K ? k0 + v : K - 1;
            m[v] = index[c + C * k];
        }
#pragma unroll
        for (int v = 0; v < 32; ++v)
            bits |= (uint64_t)(m[v] != 0 && k0 + v < K) << (k0 + v);
    }
    return bits;
}

// Workgroups are dealt to the 8 XCDs round-robin (workgroup w to XCD w % 8),
// each XCD with its own L2.  Renumbered by this bijection, XCD x runs one
// contiguous range of logical workgroups, so neighbouring cells share an L2.
__device__ inline unsigned xcd_block(unsigned w, unsigned n) {
    const unsigned q = n / 8, r = n % 8, x = w % 8;
    return x * q + (x < r ? x : r) + w / 8;
}

// Active slots kb..ke-1 of cell c (bits at their slot positions), in rounds
// of 8 clamped loads (slot groups are multiples of 8).
__device__ inline uint64_t active_slots_in(const uint8_t* __restrict__ index, int64_t c,
                                           int64_t C, int kb, int ke) {
    uint64_t bits = 0;
    for (int k0 = kb; k0 < ke; k0 += 8) {
        uint8_t m[8];
#pragma unroll
        for (int v = 0; v < 8; ++v) {
            const int k = k0 + v < ke ? k0 + v : ke - 1;
            m[v] = index[c + C * k];
        }
#pragma unroll
        for (int v = 0; v < 8; ++v)
            if (k0 + v < ke) bits |= (uint64_t)(m[v] != 0) << (k0 + v);
    }
    return bits;
}

// Nearest-neighbour search of the injection on squared distances.
// wave_argmin: lexicographic minimum of (s, order) over the wavefront.
__device__ inline void wave_argmin(double& best, int& border) {
#pragma unroll
    for (int sh = 32; sh >= 1; sh >>= 1) {
        const double ob = __shfl_xor(best, sh, 64);
        const int oo = __shfl_xor(border, sh, 64);
        if (ob < best || (ob == best && oo < border)) {
            best = ob;
            border = oo;
        }
    }
}

// The reference takes the first element (in order) of minimal distance
// sqrt(s), with sqrt correctly rounded.  sqrt is monotonic, so the minimal
// distance is sqrt(best); an element with s > best has the same rounded sqrt
// only if s <= best (1 + 2^-51) (half an ulp either side), so elements above
// best (1 + 2^-50) never tie and the others are checked with sqrt itself.
// s9[nb] = squared distance of this lane's element of neighbour nb (inf when
// absent); order = nb * 64 + lane.
template <int NB>
__device__ inline void resolve_ties(const double (&s9)[NB], double best, int& border, int lane) {
    const double lim = best * (1.0 + 0x1p-50);
    bool near = false;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) near = near || (s9[nb] > best && s9[nb] <= lim);
    if (!__any(near)) return;
    const double m = sqrt(best);
    int o = 0x7fffffff;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
        if (s9[nb] > best && s9[nb] <= lim && sqrt(s9[nb]) == m) o = min(o, nb * 64 + lane);
#pragma unroll
    for (int sh = 32; sh >= 1; sh >>= 1) o = min(o, __shfl_xor(o, sh, 64));
    border = min(border, o);
}

// a neighbour cell whose squared bound exceeds best (1 + 2^-40) cannot hold
// an element within the tie range above (the margin covers the rounding of
// the bound)
constexpr double kPrune = 1.0 + 0x1p-40;

// Appends c to the list segment of its wavefront and colour without atomics:
// segment (q, wave) holds up to 64 cells at seg[(q*nwave + wave)*64 + i] and
// its length at cnt[q*nwave + wave]; a compaction kernel joins the segments
// in cell order, which the ordered injection relies on for speed (a cell's
// neighbours of earlier colours are dealt early; appending with atomics, in
// whatever order wavefronts arrive, measured 85 -> 125 us at 512^2).
// Every lane of the wavefront must call it (want = false: nothing to add).
template <int NQ>
__device__ inline void wave_segment(bool want, int col, int32_t c, int64_t wave, int64_t nwave,
                                    int32_t* __restrict__ seg, int32_t* __restrict__ cnt) {
    if (wave >= nwave) return;  // a whole wavefront past the last cell
    const int lane = threadIdx.x & 63;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const bool mine = want && col == q;
        const uint64_t m = __ballot(mine);
        if (mine) seg[(q * nwave + wave) * 64 + __popcll(m & below)] = c;
        if (lane == 0) cnt[q * nwave + wave] = __popcll(m);
    }
}

// Iterates the set bits of `bits` in ascending order, loading NV values per
// slot (from planes p[0..NV-1], element c + C*k) two slots ahead of use.
template <int NV>
struct SlotStream {
    const double* const* p;
    int64_t c, C;
    uint64_t bits;
    int k1, k2;
    double v1[NV], v2[NV];
    __device__ inline void load(int k, double (&v)[NV]) {
#pragma unroll
        for (int i = 0; i < NV; ++i) v[i] = p[i][c + C * k];
    }
    __device__ inline int pop() {
        if (!bits) return -1;
        int k = __builtin_ctzll(bits);
        bits &= bits - 1;
        return k;
    }
    __device__ inline SlotStream(const double* const* planes, int64_t c_, int64_t C_, uint64_t b)
        : p(planes), c(c_), C(C_), bits(b) {
        k1 = pop();
        if (k1 >= 0) load(k1, v1);
        k2 = pop();
        if (k2 >= 0) load(k2, v2);
    }
    // next slot and its values; -1 when done
    __device__ inline int next(double (&out)[NV]) {
        const int k = k1;
        if (k < 0) return -1;
#pragma unroll
        for (int i = 0; i < NV; ++i) out[i] = v1[i];
        k1 = k2;
#pragma unroll
        for (int i = 0; i < NV; ++i) v1[i] = v2[i];
        k2 = pop();
        if (k2 >= 0) load(k2, v2);
        return k;
    }
};

// ---- Philox4x32-10 (Salmon et al., SC11) --------------------------------
__host__ __device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[1] = (uint32_t)p1;
        c[3] = (uint32_t)p0;
        c[0] = n0;
        c[2] = n2;
    }
}

// uniform double in [0,1), 53 bits: counter (cell, slot|dimblock<<8|purpose<<16,
// counter lo, counter hi), key = seed.  dims 0/1: block 0 words (0,1)/(2,3);
// dim 2: block 1 words (0,1).
__host__ __device__ inline double uniform53(uint64_t seed, uint64_t cell, int slot, int dim,
                                            int purpose, uint64_t counter) {
    int block = dim >> 1, half = dim & 1;
    uint32_t c[4];
    c[0] = (uint32_t)cell;
    c[1] = (uint32_t)slot | ((uint32_t)block << 8) | ((uint32_t)purpose << 16);
    c[2] = (uint32_t)counter;
    c[3] = (uint32_t)(counter >> 32);
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    uint64_t w = half ? (((uint64_t)c[2] << 32) | c[3]) : (((uint64_t)c[0] << 32) | c[1]);
    return (double)(w >> 11) * 0x1.0p-53;
}

__device__ inline void stat_add(jp_stats* st, int field, int64_t v) {
    if (st && v) atomicAdd(reinterpret_cast<unsigned long long*>(&st->moved) + field,
                           (unsigned long long)v);
}
enum { ST_MOVED = 0, ST_DELETED, ST_DROPPED, ST_FAR, ST_INJECTED, ST_ORPHAN };

}  // namespace jp
