// jp_inject2.hpp -- 2D injection with at most one field, FP64 and FP32-split
// positions (inject_particles!, src/Particles/injection.jl:68-131,
// index_min_distance :330-393).
//
// One wavefront per listed cell, lanes = slots (K <= 64).  The reference
// injects one particle at a time and searches its nearest neighbour among
// the particles present at that moment, which include the cell's earlier
// injections, so a cell's searches form a chain.  Nothing in a search feeds
// the next one's inputs, though: which slots are filled, in which quadrant,
// follows from the quadrant counts alone; each new position is a function of
// (cell, slot); and the candidates of injection j are the particles present
// before the call plus the injections before j.  So the wavefront first plans
// the whole cell with ballots (quadrants in order, the lowest free slots, the
// count of each quadrant including what earlier quadrants added), draws every
// new position in its own lane, loads the 3x3 neighbourhood once (positions
// and field values), and then runs the searches back to back on registers:
// per injection a lane-parallel distance pass, a wavefront argmin on squared
// distances with exact tie resolution (first occurrence of the minimal
// rounded distance in (neighbour cell, slot) order), and the field copy,
// which reads the source lane's current value (an earlier injection's copy
// when the nearest particle is one of them).
//
// The reference runs the cells in four colour launches (offsets (1,1),
// (1,2), (2,1), (2,2), injection.jl:28-35) because a cell reads its 3x3
// neighbourhood and writes itself.  ORDERED runs all four colour lists in
// one launch (one at a time per device: ordered_gate_enter/leave): a listed
// cell first waits until every listed neighbour of an
// earlier colour has published its result (per-cell state word: 0 not
// listed, 1 listed, 2 done), which gives every read the value of the colour
// order.  A later-colour listed neighbour waits for this cell in turn, so it
// cannot write before this cell has read it.  Only the data that crosses
// cells this way is made coherent across the chip, access by access: the
// cell's stores of positions, field values and occupancy word are
// device-scope (written through the XCD's L2), the state is stored after a
// wait for them, and a cell reads a done neighbour's data with device-scope
// loads (an agent-scope release / acquire fence would write back or drop
// the whole L2 of the XCD for every cell).  The wavefronts
// are persistent and all resident at once (the launch is capped at the
// occupancy limit); wavefront w takes items w, w + W, ... of the colour-major
// item order, so the lowest unfinished item always belongs to a running
// wavefront whose dependencies are all lower items: every wait ends.  A wait
// is still bounded (kInjectWaitMax polls) and reports a timeout instead of
// spinning on.
#pragma once

#include <type_traits>

#include "jp_ctx.hpp"

namespace jp {

struct Inject2 {
    Grid<2> g;
    double* x[2];      // FP64 form: coordinates
    float* u[2];       // split form: offsets in cell units
    double h[2];       // split form: uniform vertex spacing
    uint8_t* index;
    double* arg;       // the field moved with the particles, or null
    int min_q;         // particles wanted per quadrant
    uint64_t seed, counter;
    const int32_t* list;   // listed cells (ORDERED: colour q's at list + q * cap)
    const int32_t* count;  // their number (ORDERED: one per colour)
    uint64_t* occw;        // occupancy word of every cell (updated here)
    jp_stats* st;
    int64_t cap;           // ORDERED: list stride between colours
    uint32_t* state;       // ORDERED: per cell 0 not listed, 1 listed, 2 done
    int* timeout;          // ORDERED: set when a dependency wait gave up
};

constexpr int kInjectWaitMax = 1 << 20;
// s_waitcnt immediate of gfx9 for vmcnt(0) with expcnt and lgkmcnt left free
constexpr int kInjVmcnt0 = 0x0F70;

// device-scope (coherent across the XCDs' L2 caches) loads and stores
__device__ inline uint32_t ld_dev(const uint32_t* p) {
    return __hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint64_t ld_dev(const uint64_t* p) {
    return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline float ld_dev(const float* p) {
    return __uint_as_float(ld_dev(reinterpret_cast<const uint32_t*>(p)));
}
__device__ inline double ld_dev(const double* p) {
    return __longlong_as_double((long long)ld_dev(reinterpret_cast<const uint64_t*>(p)));
}
__device__ inline void st_dev(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_dev(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_dev(float* p, float v) {
    st_dev(reinterpret_cast<uint32_t*>(p), __float_as_uint(v));
}
__device__ inline void st_dev(double* p, double v) {
    st_dev(reinterpret_cast<uint64_t*>(p), (uint64_t)__double_as_longlong(v));
}

// a uniform lane's double (v_readlane on both halves)
__device__ inline double readlane_d(double v, int l) {
    const uint64_t b = __double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)b, l);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(b >> 32), l);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ inline float readlane_f(float v, int l) {
    return __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(v), l));
}

// the lowest n set bits of m
__device__ inline uint64_t lowest_bits(uint64_t m, int n) {
    uint64_t t = 0;
    for (int k = 0; k < n && m; ++k) {
        const uint64_t b = m & (~m + 1);
        t |= b;
        m ^= b;
    }
    return t;
}

template <bool SPLIT, bool ORDERED>
__global__ void __launch_bounds__(kBlock) k_inject2(Inject2 a) {
    using P = typename std::conditional<SPLIT, float, double>::type;
    constexpr int NB = 9, CENTER = 4;
    const int lane = threadIdx.x & 63;
    const int64_t nwv = (int64_t)gridDim.x * (kBlock / 64);
    int64_t tq[4] = {0, 0, 0, 0};
    int64_t total = 0;
    if constexpr (ORDERED) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            tq[q] = a.count[q];
            total += tq[q];
        }
    } else {
        total = *a.count;
    }
    const Grid<2>& g = a.g;
    const int64_t C = g.C;
    const int K = g.K;
    const uint64_t full = K == 64 ? ~0ull : ((1ull << K) - 1);
    const bool slot = lane < K;
    const bool has_arg = a.arg != nullptr;
    int64_t n_inj = 0, n_orph = 0;
    for (int64_t w = blockIdx.x * (int64_t)(kBlock / 64) + (threadIdx.x >> 6); w < total;
         w += nwv) {
        int64_t at = w;
        if constexpr (ORDERED) {
            // item w in colour-major order: colour q, entry w - (counts before q)
            int64_t r = w;
            int q = 0;
            while (q < 3 && r >= tq[q]) {
                r -= tq[q];
                ++q;
            }
            at = q * a.cap + r;
        }
        const int64_t c = __builtin_amdgcn_readfirstlane(a.list[at]);
        int64_t I[2];
        unlin<2>(g.ni, c, I);
        int64_t cn[NB];
        bool okn[NB];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const int64_t J0 = I[0] + nb % 3 - 1, J1 = I[1] + nb / 3 - 1;
            okn[nb] = J0 >= 0 && J0 < g.ni[0] && J1 >= 0 && J1 < g.ni[1];
            cn[nb] = okn[nb] ? J0 + g.ni[0] * J1 : c;  // a valid cell either way
        }
        bool coh[NB];  // a done neighbour's data: device-scope loads
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) coh[nb] = false;
        if constexpr (ORDERED) {
            // lane nb watches neighbour nb when it has an earlier colour
            // (colour = 2 * (i mod 2) + (j mod 2), the launch order)
            const int mine = (int)(((I[0] & 1) << 1) | (I[1] & 1));
            int64_t wc = c;
            bool need = false;
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                const int64_t J0 = I[0] + nb % 3 - 1, J1 = I[1] + nb / 3 - 1;
                const int col = (int)(((J0 & 1) << 1) | (J1 & 1));
                if (lane == nb) {
                    wc = cn[nb];
                    need = okn[nb] && col < mine;
                }
            }
            uint32_t sv = need ? ld_dev(a.state + wc) : 0u;
            const uint64_t listed = __ballot(sv != 0u);
            if (listed) {  // an earlier-colour neighbour injects in this call
                int polls = 0;
                while (__ballot(sv == 1u)) {
                    if (++polls > kInjectWaitMax) {
                        if (lane == 0) atomicOr(a.timeout, 1);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                    if (sv == 1u) sv = ld_dev(a.state + wc);
                }
            }
            // (the loads below issue after the state values they depend on)
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) coh[nb] = (listed >> nb) & 1;
        }
        uint64_t o9[NB];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) o9[nb] = coh[nb] ? ld_dev(a.occw + cn[nb]) : a.occw[cn[nb]];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) o9[nb] = okn[nb] ? o9[nb] : 0ull;
        // positions of the neighbourhood's particles (active slots only: in
        // the slot-plane layout every slot is its own cache line) and this
        // cell's field values (every slot: an orphan injection keeps the value
        // its slot holds); a neighbour's value is read only when a search
        // picks it, below
        P q[NB][2];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const bool take = (o9[nb] >> lane) & 1;
            const int64_t e2 = cn[nb] + C * lane;
            if constexpr (SPLIT) {
                if (coh[nb]) {
                    q[nb][0] = take ? ld_dev(a.u[0] + e2) : 0.0f;
                    q[nb][1] = take ? ld_dev(a.u[1] + e2) : 0.0f;
                } else {
                    q[nb][0] = take ? a.u[0][e2] : 0.0f;
                    q[nb][1] = take ? a.u[1][e2] : 0.0f;
                }
            } else {
                if (coh[nb]) {
                    q[nb][0] = take ? ld_dev(a.x[0] + e2) : 0.0;
                    q[nb][1] = take ? ld_dev(a.x[1] + e2) : 0.0;
                } else {
                    q[nb][0] = take ? a.x[0][e2] : 0.0;
                    q[nb][1] = take ? a.x[1][e2] : 0.0;
                }
            }
        }
        double fc = (has_arg && slot) ? a.arg[c + C * lane] : 0.0;
        const bool act0 = (o9[CENTER] >> lane) & 1;
        double xv[2], dq[2];
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            xv[d] = g.xv[d].x[I[d]];
            dq[d] = g.xv[d].dx[I[d]] / 2.0;
        }
        // the random numbers of a new particle depend on (cell, slot, dim)
        // only, not on its quadrant: drawn once, in the free slots
        double r2[2] = {0.0, 0.0};
        if (slot && !act0) {
            r2[0] = uniform53(a.seed, (uint64_t)c, lane, 0, 2, a.counter);
            r2[1] = uniform53(a.seed, (uint64_t)c, lane, 1, 2, a.counter);
        }
        // plan: the quadrants in order, each topped up from the lowest free
        // slots until it holds min_q particles; the first full quadrant ends
        // the cell (injection.jl:87-126)
        uint64_t injm = 0;
        bool act = act0;
        for (int qd = 0; qd < 4; ++qd) {
            bool inside = act;
            if constexpr (SPLIT) {
                const float q0 = 0.5f * (float)(qd & 1), q1 = 0.5f * (float)((qd >> 1) & 1);
                inside = inside && (q0 < q[CENTER][0] && q[CENTER][0] < q0 + 0.5f) &&
                         (q1 < q[CENTER][1] && q[CENTER][1] < q1 + 0.5f);
            } else {
#pragma unroll
                for (int d = 0; d < 2; ++d) {
                    const double v = xv[d] + dq[d] * (double)((qd >> d) & 1);
                    inside = inside && (v < q[CENTER][d] && q[CENTER][d] < v + dq[d]);
                }
            }
            const int num = __popcll(__ballot(inside));
            if (num >= a.min_q) break;
            const uint64_t t = lowest_bits(~__ballot(act) & full, a.min_q - num);
            injm |= t;
            if ((t >> lane) & 1) {
                // new_particle: vertex + dq * muladd(0.95, rand, 0.05)
                act = true;
                if constexpr (SPLIT) {
                    const double q0 = 0.5 * (double)(qd & 1), q1 = 0.5 * (double)((qd >> 1) & 1);
                    q[CENTER][0] = (float)(q0 + 0.5 * fma(0.95, r2[0], 0.05));
                    q[CENTER][1] = (float)(q1 + 0.5 * fma(0.95, r2[1], 0.05));
                } else {
#pragma unroll
                    for (int d = 0; d < 2; ++d) {
                        const double v = xv[d] + dq[d] * (double)((qd >> d) & 1);
                        q[CENTER][d] = v + dq[d] * fma(0.95, r2[d], 0.05);
                    }
                }
            }
        }
        // the searches, in injection (= ascending slot) order; lane j of
        // `src` records injection j's nearest particle (order nb*64 + slot)
        int src = 0x7fffffff;
        int j = 0;
        for (uint64_t bits = injm; bits; bits &= bits - 1, ++j) {
            const int ip = __builtin_ctzll(bits);
            P pn[2];
            if constexpr (SPLIT) {
                pn[0] = readlane_f(q[CENTER][0], ip);
                pn[1] = readlane_f(q[CENTER][1], ip);
            } else {
                pn[0] = readlane_d(q[CENTER][0], ip);
                pn[1] = readlane_d(q[CENTER][1], ip);
            }
            // this cell's candidates: its particles before the call and the
            // injections before this one
            const bool cand_c = (act0 || (((injm >> lane) & 1) && lane < ip)) && lane != ip;
            double best = __builtin_inf();
            int border = 0x7fffffff;
            double s9[NB];
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                s9[nb] = __builtin_inf();
                const bool av = nb == CENTER ? cand_c : (bool)((o9[nb] >> lane) & 1);
                if (!av) continue;
                double sq;
                if constexpr (SPLIT) {
                    const double r0 =
                        ((double)(nb % 3 - 1) + (double)q[nb][0] - (double)pn[0]) * a.h[0];
                    const double r1 =
                        ((double)(nb / 3 - 1) + (double)q[nb][1] - (double)pn[1]) * a.h[1];
                    sq = r0 * r0 + r1 * r1;
                } else {
                    if (q[nb][0] != q[nb][0] || q[nb][1] != q[nb][1]) continue;  // 2D NaN skip
                    const double du = q[nb][0] - pn[0], dv = q[nb][1] - pn[1];
                    sq = du * du + dv * dv;
                }
                s9[nb] = sq;
                if (sq < best) {
                    best = sq;
                    border = nb * 64 + lane;
                }
            }
            wave_argmin(best, border);
            resolve_ties(s9, best, border, lane);
            if (lane == j) src = border;
        }
        const int ninj = j;
        if (has_arg && ninj) {
            // values of the sources in neighbour cells: one load per
            // injection, all in flight together
            double nv = 0.0;
            if (lane < ninj && src != 0x7fffffff && (src >> 6) != CENTER) {
                const int nbw = src >> 6;
                int64_t cs = cn[0];
                bool cc = coh[0];
#pragma unroll
                for (int nb = 1; nb < NB; ++nb) {
                    cs = nb == nbw ? cn[nb] : cs;
                    cc = nb == nbw ? coh[nb] : cc;
                }
                nv = cc ? ld_dev(a.arg + cs + C * (src & 63)) : a.arg[cs + C * (src & 63)];
            }
            // the copies in injection order: a source in this cell is read
            // from the current values (an earlier injection's copy included)
            j = 0;
            for (uint64_t bits = injm; bits; bits &= bits - 1, ++j) {
                const int ip = __builtin_ctzll(bits);
                const int sj = __builtin_amdgcn_readlane(src, j);
                if (sj == 0x7fffffff) continue;  // orphan: the slot keeps its value
                const double v = (sj >> 6) == CENTER ? readlane_d(fc, sj & 63) : readlane_d(nv, j);
                if (lane == ip) fc = v;
            }
        }
        n_orph += __popcll(__ballot(lane < ninj && src == 0x7fffffff));
        n_inj += __popcll(injm);
        if ((injm >> lane) & 1) {
            const int64_t e = c + C * lane;
            if constexpr (ORDERED) {
                if constexpr (SPLIT) {
                    st_dev(a.u[0] + e, q[CENTER][0]);
                    st_dev(a.u[1] + e, q[CENTER][1]);
                } else {
                    st_dev(a.x[0] + e, q[CENTER][0]);
                    st_dev(a.x[1] + e, q[CENTER][1]);
                }
                if (has_arg) st_dev(a.arg + e, fc);
            } else {
                if constexpr (SPLIT) {
                    a.u[0][e] = q[CENTER][0];
                    a.u[1][e] = q[CENTER][1];
                } else {
                    a.x[0][e] = q[CENTER][0];
                    a.x[1][e] = q[CENTER][1];
                }
                if (has_arg) a.arg[e] = fc;
            }
            a.index[e] = 1;  // (read by later launches only)
        }
        // later colours see this cell's final occupancy
        if constexpr (ORDERED) {
            if (lane == 0) st_dev(a.occw + c, o9[CENTER] | injm);
            // publish: this cell's stores, and its loads of the neighbourhood,
            // complete before its state says done
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            __builtin_amdgcn_s_waitcnt(kInjVmcnt0);
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            if (lane == 0) st_dev(a.state + c, 2u);
        } else {
            if (lane == 0) a.occw[c] = o9[CENTER] | injm;
        }
    }  // work list
    if (lane == 0) {
        stat_add(a.st, ST_INJECTED, n_inj);
        stat_add(a.st, ST_ORPHAN, n_orph);
    }
}

// The 2D injection's launches after the scan and the join: the four colours
// in one ordered launch (default), or one launch per colour
// (JP_OPT_INJECT_PHASES).  `a` holds everything but the lists.
inline bool inject2_ordered(const jp_ctx* ctx) { return ctx->opt_inject_phases == 0; }

template <bool SPLIT>
int launch_inject2(jp_ctx* ctx, Inject2 a, const InjectLists& L, int64_t cap) {
    if (inject2_ordered(ctx)) {
        // every wavefront of the launch resident at once (each may wait on
        // lower items of the others)
        static int per_cu = -1;
        if (per_cu < 0) {
            int n = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_inject2<SPLIT, true>, kBlock,
                                                             0) != hipSuccess ||
                n < 1)
                n = 1;
            per_cu = n;
        }
        const int64_t most = (int64_t)per_cu * ctx->num_cu;
        const int64_t want = inject_blocks(ctx, 4 * cap);
        a.list = L.list;
        a.count = L.total;
        a.cap = cap;
        a.state = L.state;
        a.timeout = dinject(ctx);
        int rc = ordered_gate_enter(ctx);
        if (rc) return rc;
        k_inject2<SPLIT, true>
            <<<(unsigned)(want < most ? want : most), kBlock, 0, ctx->stream>>>(a);
        const hipError_t e = hipGetLastError();
        rc = ordered_gate_leave(ctx);
        if (e != hipSuccess) return fail(ctx, JP_ERR_HIP, "HIP error %s", hipGetErrorString(e));
        return rc;
    }
    for (int ph = 0; ph < 4; ++ph) {
        a.list = L.list + (int64_t)ph * cap;
        a.count = L.total + ph;
        k_inject2<SPLIT, false><<<(unsigned)inject_blocks(ctx, cap), kBlock, 0, ctx->stream>>>(a);
        JP_LAUNCHED(ctx);
    }
    return JP_OK;
}

}  // namespace jp
