// jp_hotpath.hip -- FP64 drop-in kernels for JustPIC's particle hot path on
// gfx950: init, particle2grid, grid2particle and injection, 2D and 3D, on the
// B = 0 cell-plane layout data[c + C*k] (advection: jp_advect.hip, move:
// jp_move.hip).
#include "jp_ctx.hpp"
#include "jp_inject2.hpp"

#include <cmath>

using namespace jp;

namespace {

__device__ inline double dnan() { return __builtin_nan(""); }

// ============================================================ init
// fill_coords_index! (src/Particles/particles_utils.jl:168-194), one thread
// per cell writing its K slots (coalesced across the cell planes).
template <int D>
__global__ void __launch_bounds__(kBlock) k_init(Grid<D> g, Ptr3 coords, uint8_t* __restrict__ index,
                                                 int npq, uint64_t seed) {
    int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= g.C) return;
    int64_t I[D];
    unlin<D>(g.ni, c, I);
    double x0[D], dx[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        x0[d] = g.xv[d].x[I[d]];
        dx[d] = g.xv[d].dx[I[d]];
    }
    constexpr int nq = D == 2 ? 4 : 8;
    int l = 0;
    for (int iq = 0; iq < nq; ++iq) {
        double xc[D];
#pragma unroll
        for (int d = 0; d < D; ++d) xc[d] = x0[d] + (dx[d] * (double)((iq >> d) & 1)) / 2.0;
        for (int n = 0; n < npq; ++n, ++l) {
            int64_t e = c + g.C * l;
#pragma unroll
            for (int d = 0; d < D; ++d)
                coords.p[d][e] = xc[d] + (dx[d] / 2.0) * uniform53(seed, (uint64_t)c, l, d, 1, 0);
            index[e] = 1;
        }
    }
    for (; l < g.K; ++l) {
        int64_t e = c + g.C * l;
#pragma unroll
        for (int d = 0; d < D; ++d) coords.p[d][e] = dnan();
        index[e] = 0;
    }
}

// ============================================================ particle2grid
// _particle2grid! (src/Interpolations/particle_to_grid.jl:37-199): one thread
// per node, exact reference accumulation order (cells with the last dim
// outermost, slots ascending).  mode 0: wF/w (2D scalar), 1: wF*inv(w).
template <int D, int NF>
__global__ void __launch_bounds__(kBlock) k_p2g(Grid<D> g, int nf_rt, ArgPtrs F, ArgPtrs Fp,
                                                CPtr3 coords, const uint64_t* __restrict__ occw,
                                                int mode) {
    // NF == 1: one field, accumulator kept in a register; NF == 0: nf_rt fields
    const int nf = NF == 1 ? 1 : nf_rt;
    int64_t nn[D];
    int64_t NN = 1;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        nn[d] = g.ni[d] + 1;
        NN *= nn[d];
    }
    int64_t node = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (node >= NN) return;
    int64_t N[D];
    {
        int64_t r = node;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            N[d] = r % nn[d];
            r /= nn[d];
        }
    }
    double xn[D];
#pragma unroll
    for (int d = 0; d < D; ++d) xn[d] = g.xv[d].x[N[d]];
    double w = 0.0;
    double wF[JP_MAX_ARGS];
    for (int f = 0; f < nf; ++f) wF[f] = 0.0;
    const int64_t C = g.C;
    constexpr int NC = 1 << D;
    // the (up to) 2^D cells around the node, last dim outermost, and their
    // occupancy words (one round of loads)
    int64_t cell[NC];
    uint64_t bits[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        int64_t I[D];
        bool ok = true;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            I[d] = N[d] + (((q >> d) & 1) ? 0 : -1);
            ok = ok && (I[d] >= 0 && I[d] < g.ni[d]);
        }
        cell[q] = ok ? lin<D>(g.ni, I) : 0;
        bits[q] = ok ? occw[cell[q]] : 0ull;
    }
    // Reference order: cells, then slots ascending.  The (cell, chunk of B
    // slots) sequence is walked with the next chunk's loads in flight while
    // the current one is accumulated; all lanes of a wavefront are on the same
    // slot plane, so every load is coalesced across neighbouring cells.
    constexpr int B = D == 2 ? 8 : 4;  // two chunk buffers stay in registers
    const int nk = (g.K + B - 1) / B;
    const int total = NC * nk;
    auto chunk = [&](int t, int64_t& c, uint32_t& bq, int& k0) __attribute__((always_inline)) {
        const int q = t / nk;
        k0 = (t - q * nk) * B;
        c = cell[0];
        uint64_t bw = bits[0];
#pragma unroll
        for (int u = 1; u < NC; ++u) {
            c = q == u ? cell[u] : c;
            bw = q == u ? bits[u] : bw;
        }
        bq = (uint32_t)(bw >> k0) & ((1u << B) - 1);
    };
    auto load = [&](int t, double (&pv)[B][D + 1]) __attribute__((always_inline)) {
        int64_t c;
        uint32_t bq;
        int k0;
        chunk(t, c, bq, k0);
#pragma unroll
        for (int j = 0; j < B; ++j) {
            const bool on = (bq >> j) & 1;
            const int64_t e = c + C * (k0 + j);
#pragma unroll
            for (int d = 0; d < D; ++d) pv[j][d] = on ? coords.p[d][e] : 0.0;
            pv[j][D] = on ? Fp.p[0][e] : 0.0;
        }
    };
    auto accumulate = [&](int t, const double (&pv)[B][D + 1]) __attribute__((always_inline)) {
        int64_t c;
        uint32_t bq;
        int k0;
        chunk(t, c, bq, k0);
#pragma unroll
        for (int j = 0; j < B; ++j) {
            if (!((bq >> j) & 1)) continue;
            double a = xn[0] - pv[j][0];
            double s = a * a;
#pragma unroll
            for (int d = 1; d < D; ++d) {
                double b = xn[d] - pv[j][d];
                s = s + b * b;
            }
            const double dist = sqrt(s);
            const double wi = 1.0 / (dist * dist);
            w += wi;
            wF[0] = fma(wi, pv[j][D], wF[0]);
            for (int f = 1; f < nf; ++f) wF[f] = fma(wi, Fp.p[f][c + C * (k0 + j)], wF[f]);
        }
    };
    double b0[B][D + 1], b1[B][D + 1];
    load(0, b0);
#pragma unroll 1
    for (int t = 0; t < total; t += 2) {
        if (t + 1 < total) load(t + 1, b1);
        accumulate(t, b0);
        if (t + 1 >= total) break;
        if (t + 2 < total) load(t + 2, b0);
        accumulate(t + 1, b1);
    }
    if (mode == 0) {
        F.p[0][node] = wF[0] / w;
    } else {
        double iw = 1.0 / w;
        for (int f = 0; f < nf; ++f) F.p[f][node] = wF[f] * iw;
    }
}

// ============================================================ grid2particle
// grid2particle_classic! (src/Interpolations/grid_to_particle.jl:38-82).
template <int D>
__global__ void __launch_bounds__(kBlock) k_g2p(Grid<D> g, double* __restrict__ Fp,
                                                const double* __restrict__ F, CPtr3 coords,
                                                const uint8_t* __restrict__ index) {
    int64_t c, I[D];
    if (!row_cell<D>(g, c, I)) return;
    const int64_t e = c + g.C * (int64_t)blockIdx.y;
    if (!index[e]) return;
    double t[D];
#pragma unroll
    for (int d = 0; d < D; ++d)
        t[d] = (coords.p[d][e] - g.xv[d].x[I[d]]) * g.xv[d].idx[I[d]];
    const int64_t n0 = g.ni[0] + 1;
    if constexpr (D == 2) {
        int64_t b = I[0] + n0 * I[1];
        Fp[e] = lerp1(t[1], lerp1(t[0], F[b], F[b + 1]), lerp1(t[0], F[b + n0], F[b + n0 + 1]));
    } else {
        const int64_t n1 = g.ni[1] + 1;
        int64_t b = I[0] + n0 * (I[1] + n1 * I[2]);
        int64_t s = n0 * n1;
        double lo = lerp1(t[1], lerp1(t[0], F[b], F[b + 1]), lerp1(t[0], F[b + n0], F[b + n0 + 1]));
        double hi = lerp1(t[1], lerp1(t[0], F[b + s], F[b + s + 1]),
                          lerp1(t[0], F[b + s + n0], F[b + s + n0 + 1]));
        Fp[e] = lerp1(t[2], lo, hi);
    }
}

// ============================================================ injection
// The first quadrant decides whether a cell injects at all: the reference
// breaks out of the quadrant loop as soon as one quadrant holds min_q
// particles (src/Particles/injection.jl:100), and injection into a cell never
// changes the particles of other cells.  One coalesced pass over all cells
// therefore lists the only cells that can inject, per colour; the colour
// phases then run the reference loop on those cells only.  ANYQ (the phase
// variant, which moves on to the next quadrant instead of breaking,
// src/Particles/injection.jl:278): a cell is listed when any quadrant is short.
template <int D, bool ANYQ = false>
__global__ void __launch_bounds__(kBlock) k_inject_scan(Grid<D> g, CPtr3 coords,
                                                        const uint8_t* __restrict__ index,
                                                        int min_q, InjectLists L) {
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const bool valid = c < g.C;  // all lanes stay for the wavefront append
    int64_t I[D];
    unlin<D>(g.ni, valid ? c : 0, I);
    int num = 0;
    int cnt_q[1 << D] = {};  // ANYQ: particles strictly inside quadrants 1..2^D-1
    if (valid) {
        double lo[D], hq[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            lo[d] = g.xv[d].x[I[d]];
            hq[d] = g.xv[d].dx[I[d]] / 2.0;
        }
        // quadrant c11: vertex + dq*0 (src/Particles/injection.jl:442-448)
        double q0[D], q1[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            q0[d] = lo[d] + hq[d] * 0.0;
            q1[d] = q0[d] + hq[d];
        }
        const uint64_t act = active_slots(index, c, g.C, g.K);
        double bl[D], bh[D];  // box of the cell's particles (NaN coordinates ignored)
#pragma unroll
        for (int d = 0; d < D; ++d) {
            bl[d] = __builtin_inf();
            bh[d] = -__builtin_inf();
        }
        constexpr int B = 16;  // slots per chunk; lanes share the slot plane (coalesced)
        for (int k0 = 0; k0 < g.K; k0 += B) {
            double pb[B][D];
#pragma unroll
            for (int j = 0; j < B; ++j) {
                const bool on = (act >> (k0 + j)) & 1;
#pragma unroll
                for (int d = 0; d < D; ++d) pb[j][d] = on ? coords.p[d][c + g.C * (k0 + j)] : 0.0;
            }
#pragma unroll
            for (int j = 0; j < B; ++j) {
                const bool on = (act >> (k0 + j)) & 1;
                bool inside = on;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    inside = inside & (q0[d] < pb[j][d] && pb[j][d] < q1[d]);
                    if (D == 2 && on) {
                        bl[d] = fmin(bl[d], pb[j][d]);
                        bh[d] = fmax(bh[d], pb[j][d]);
                    }
                }
                num += inside ? 1 : 0;
                if constexpr (ANYQ) {
                    // the other quadrants (vertex + dq*bit, the wave kernel's bounds)
#pragma unroll
                    for (int q = 1; q < (1 << D); ++q) {
                        bool in_q = on;
#pragma unroll
                        for (int d = 0; d < D; ++d) {
                            const double v = lo[d] + hq[d] * (double)((q >> d) & 1);
                            in_q = in_q & (v < pb[j][d] && pb[j][d] < v + hq[d]);
                        }
                        cnt_q[q] += in_q ? 1 : 0;
                    }
                }
            }
        }
        L.occ[c] = act;
        if (D == 2 && L.box) {
            double* __restrict__ bx = static_cast<double*>(L.box);
#pragma unroll
            for (int d = 0; d < D; ++d) {
                bx[(2 * d) * g.C + c] = bl[d];
                bx[(2 * d + 1) * g.C + c] = bh[d];
            }
        }
    }
    bool want = valid && num < min_q;
    if constexpr (ANYQ) {
#pragma unroll
        for (int q = 1; q < (1 << D); ++q) want = want || (valid && cnt_q[q] < min_q);
    }
    if (valid && L.state) L.state[c] = want ? 1u : 0u;
    int col = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) col = col * 2 + (int)(I[d] & 1);
    wave_segment<D == 2 ? 4 : 8>(want, col, (int32_t)c, c >> 6, L.nwave, L.seg, L.cnt);
}

// _inject_particles! (src/Particles/injection.jl:68-131) with one wavefront per
// listed cell: lane k owns slot k of the cell (K <= 64).  Quadrant counts are
// ballots; injected particles fill the free slots in ascending order, and the
// nearest-neighbour search of index_min_distance (:330-393) evaluates the
// 3^D x K candidates lane-parallel, keeping the reference's first-occurrence
// tie rule (strict '<' in loop order: neighbour cell, then slot) through a
// (distance, order) argmin.
//
// PHASE: _inject_particles_phase! (src/Particles/injection.jl:233-325): every
// short quadrant is topped up (`continue` past a full one, :278), the new
// particle takes its nearest neighbour's phase, and each arg is interpolated
// from its field (cell-centred or nodal) and clamped to the corner extrema.
struct PhaseInj {
    double* phases;
    const double* fields[JP_MAX_ARGS];
    uint32_t centred;  // bit a: fields[a] is cell-centred (dims ni), else nodal (ni + 1)
    Axis xc[3];        // centre axes (particles.xci with di.center)
};

// Julia's min / max on Float64 (NaN-propagating, -0.0 below 0.0) and clamp
__device__ inline double jl_min(double x, double y) {
    if (x != x || y != y) return x != x ? x : y;
    return (x < y || (x == y && signbit(x) && !signbit(y))) ? x : y;
}
__device__ inline double jl_max(double x, double y) {
    if (x != x || y != y) return x != x ? x : y;
    return (x > y || (x == y && !signbit(x) && signbit(y))) ? x : y;
}

// _grid2particle at corner cell idx of F (dims nn), t = (p - x0) * inv(di),
// clamped to the extrema of the 2^D corner values (src/Particles/
// injection.jl:299-314, cell_field :395-409)
template <int D>
__device__ double g2p_clamped(const double* __restrict__ F, const int64_t (&nn)[D],
                              const int64_t (&idx)[D], const double (&p)[D],
                              const double (&x0)[D], const double (&inv)[D]) {
    double v[1 << D], t[D];
#pragma unroll
    for (int q = 0; q < (1 << D); ++q) {
        int64_t o = 0, stride = 1;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            o += (idx[d] + ((q >> d) & 1)) * stride;
            stride *= nn[d];
        }
        v[q] = F[o];
    }
#pragma unroll
    for (int d = 0; d < D; ++d) t[d] = (p[d] - x0[d]) * inv[d];
    double r = lerp1(t[1], lerp1(t[0], v[0], v[1]), lerp1(t[0], v[2], v[3]));
    if constexpr (D == 3)
        r = lerp1(t[2], r, lerp1(t[1], lerp1(t[0], v[4], v[5]), lerp1(t[0], v[6], v[7])));
    double lo = v[0], hi = v[0];
#pragma unroll
    for (int q = 1; q < (1 << D); ++q) {
        lo = jl_min(lo, v[q]);
        hi = jl_max(hi, v[q]);
    }
    return r > hi ? hi : (r < lo ? lo : r);
}

template <int D, bool PHASE = false>
__global__ void __launch_bounds__(kBlock) k_inject_wave(Grid<D> g, Ptr3 coords, uint8_t* index,
                                                        int nargs, ArgPtrs args, int min_q,
                                                        uint64_t seed, uint64_t counter,
                                                        const int32_t* __restrict__ list,
                                                        const int32_t* __restrict__ count,
                                                        jp_stats* st, PhaseInj pi) {
    const int lane = threadIdx.x & 63;
    // persistent wavefronts over the colour's list (every wavefront exits once
    // the list is exhausted)
    const int64_t nwv = (int64_t)gridDim.x * (kBlock / 64);
    const int64_t total = *count;
    int64_t n_inj = 0, n_orph = 0;
    for (int64_t w = blockIdx.x * (int64_t)(kBlock / 64) + (threadIdx.x >> 6); w < total;
         w += nwv) {
    const int64_t c = list[w];
    const int64_t C = g.C;
    const int K = g.K;
    const uint64_t full = K == 64 ? ~0ull : ((1ull << K) - 1);
    int64_t I[D];
    unlin<D>(g.ni, c, I);
    constexpr int NQ = D == 2 ? 4 : 8;
    constexpr int NB = D == 2 ? 9 : 27;
    constexpr int CENTER = (NB - 1) / 2;
    double xv[D], dq[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        xv[d] = g.xv[d].x[I[d]];
        dq[d] = g.xv[d].dx[I[d]] / 2.0;
    }
    // this cell's slots, one per lane, kept current in registers
    const bool own = lane < K;
    const int64_t e = c + C * lane;
    bool act = own && index[e];
    double p[D];
#pragma unroll
    for (int d = 0; d < D; ++d) p[d] = act ? coords.p[d][e] : 0.0;
    // 2D: the neighbour cells' masks and coordinates, loaded once for all of
    // this cell's searches (they do not change while it injects: the cells
    // of one colour are two apart); 3D loads them per search, in batches
    uint8_t m9[D == 2 ? NB : 1];
    double q9[D == 2 ? NB : 1][D];
    if constexpr (D == 2) {
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            int64_t J[D];
            bool ok = own && u != CENTER;
            int r = u;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                J[d] = I[d] + (r % 3) - 1;
                r /= 3;
                ok = ok && J[d] >= 0 && J[d] < g.ni[d];
            }
            const int64_t e2 = ok ? lin<D>(g.ni, J) + C * lane : 0;
            m9[u] = ok ? index[e2] : (uint8_t)0;
#pragma unroll
            for (int d = 0; d < D; ++d) q9[u][d] = ok ? coords.p[d][e2] : 0.0;
        }
    }
    for (int q = 0; q < NQ; ++q) {
        double vert[D];
#pragma unroll
        for (int d = 0; d < D; ++d) vert[d] = xv[d] + dq[d] * (double)((q >> d) & 1);
        bool inside = act;
#pragma unroll
        for (int d = 0; d < D; ++d) inside = inside & (vert[d] < p[d] && p[d] < vert[d] + dq[d]);
        int num = __popcll(__ballot(inside));
        if (num >= min_q) {
            if constexpr (PHASE) continue;
            break;
        }
        uint64_t freeS = ~__ballot(act) & full;
        bool done = false;
        while (freeS) {
            const int ip = __builtin_ctzll(freeS);
            freeS &= freeS - 1;
            ++num;
            double pn[D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const double r = uniform53(seed, (uint64_t)c, ip, d, 2, counter);
                pn[d] = vert[d] + dq[d] * fma(0.95, r, 0.05);
            }
            const int64_t enew = c + C * ip;
            if (lane == ip) {
                act = true;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    p[d] = pn[d];
                    coords.p[d][enew] = pn[d];
                }
                index[enew] = 1;
            }
            // index_min_distance over the 3^D neighbourhood
            double best = __builtin_inf();
            int border = 0x7fffffff;
            constexpr int NBB = 9;  // neighbour cells per batch: masks, then coordinates
#pragma unroll
            for (int nb0 = 0; nb0 < NB; nb0 += NBB) {
                bool a[NBB];
                int64_t e2[NBB];
#pragma unroll
                for (int u = 0; u < NBB; ++u) {
                    const int nb = nb0 + u;
                    int64_t J[D];
                    bool ok = own;
                    int r = nb;
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        J[d] = I[d] + (r % 3) - 1;
                        r /= 3;
                        ok = ok && J[d] >= 0 && J[d] < g.ni[d];
                    }
                    e2[u] = ok ? lin<D>(g.ni, J) + C * lane : 0;
                    a[u] = ok;
                }
                // masks and coordinates of the batch in one round of loads
                // (an empty slot holds NaN coordinates and is masked out below)
                uint8_t m[NBB];
                double pq[NBB][D];
#pragma unroll
                for (int u = 0; u < NBB; ++u) {
                    const bool other = a[u] && nb0 + u != CENTER;
                    if constexpr (D == 2) {
                        m[u] = other ? m9[nb0 + u] : (uint8_t)0;
#pragma unroll
                        for (int d = 0; d < D; ++d) pq[u][d] = other ? q9[nb0 + u][d] : p[d];
                    } else {
                        m[u] = other ? index[e2[u]] : (uint8_t)0;
#pragma unroll
                        for (int d = 0; d < D; ++d)
                            pq[u][d] = other ? coords.p[d][e2[u]] : p[d];
                    }
                }
#pragma unroll
                for (int u = 0; u < NBB; ++u)
                    a[u] = (nb0 + u == CENTER) ? (act && lane != ip) : (m[u] != 0);
#pragma unroll
                for (int u = 0; u < NBB; ++u) {
                    if (!a[u]) continue;
                    bool hasnan = false;
#pragma unroll
                    for (int d = 0; d < D; ++d) hasnan = hasnan || (pq[u][d] != pq[u][d]);
                    if (D == 2 && hasnan) continue;
                    double du = pq[u][0] - pn[0];
                    double s = du * du;
#pragma unroll
                    for (int d = 1; d < D; ++d) {
                        double dv = pq[u][d] - pn[d];
                        s = s + dv * dv;
                    }
                    const double dd = sqrt(s);
                    if (dd < best) {
                        best = dd;
                        border = (nb0 + u) * 64 + lane;
                    }
                }
            }
#pragma unroll
            for (int m = 32; m >= 1; m >>= 1) {
                const double ob = __shfl_xor(best, m, 64);
                const int oo = __shfl_xor(border, m, 64);
                if (ob < best || (ob == best && oo < border)) {
                    best = ob;
                    border = oo;
                }
            }
            if (border == 0x7fffffff) ++n_orph;
            if (lane == ip && border != 0x7fffffff) {
                int64_t J[D];
                int r = border >> 6;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    J[d] = I[d] + (r % 3) - 1;
                    r /= 3;
                }
                const int64_t esrc = lin<D>(g.ni, J) + C * (border & 63);
                if constexpr (PHASE)
                    pi.phases[enew] = pi.phases[esrc];
                else
                    for (int a = 0; a < nargs; ++a) args.p[a][enew] = args.p[a][esrc];
            }
            if constexpr (PHASE) {
                if (lane == ip) {
                    // args from fields: nodal at this cell's corners, cell-centred at
                    // the shifted index clamped to 1..ni-1 (1-based)
                    int64_t nnv[D], nnc[D], ic[D], Iv[D];
                    double xc0[D], ivc[D], xv0[D], ivv[D];
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        nnv[d] = g.ni[d] + 1;
                        nnc[d] = g.ni[d];
                        Iv[d] = I[d];
                        xv0[d] = xv[d];
                        ivv[d] = g.xv[d].idx[I[d]];
                        int64_t k = pn[d] < xv[d] + dq[d] ? I[d] - 1 : I[d];
                        k = k > g.ni[d] - 2 ? g.ni[d] - 2 : (k < 0 ? 0 : k);
                        ic[d] = k;
                        xc0[d] = pi.xc[d].x[k];
                        ivc[d] = pi.xc[d].idx[k];
                    }
                    for (int a = 0; a < nargs; ++a)
                        args.p[a][enew] = ((pi.centred >> a) & 1)
                                              ? g2p_clamped<D>(pi.fields[a], nnc, ic, pn, xc0, ivc)
                                              : g2p_clamped<D>(pi.fields[a], nnv, Iv, pn, xv0, ivv);
                }
            }
            // later searches of this wavefront read what lane ip just wrote
            __threadfence_block();
            ++n_inj;
            if (num >= min_q) {
                done = true;
                break;
            }
        }
        (void)done;
    }
    }  // work list
    if (lane == 0) {
        stat_add(st, ST_INJECTED, n_inj);
        stat_add(st, ST_ORPHAN, n_orph);
    }
}

// ============================================================ host side
int check_ptrs(jp_ctx* ctx, int n, const void* const* p, const char* what) {
    if (!p) return fail(ctx, JP_ERR_ARG, "null %s array", what);
    for (int i = 0; i < n; ++i)
        if (!p[i]) return fail(ctx, JP_ERR_ARG, "null %s pointer %lld", what, i);
    return JP_OK;
}

Ptr3 ptr3(double* const* c, int D) {
    Ptr3 p{{nullptr, nullptr, nullptr}};
    for (int d = 0; d < D; ++d) p.p[d] = c[d];
    return p;
}
CPtr3 cptr3(const double* const* c, int D) {
    CPtr3 p{{nullptr, nullptr, nullptr}};
    for (int d = 0; d < D; ++d) p.p[d] = c[d];
    return p;
}
ArgPtrs argptrs(double* const* a, int n) {
    ArgPtrs p;
    for (int i = 0; i < JP_MAX_ARGS; ++i) p.p[i] = i < n ? a[i] : nullptr;
    return p;
}

Off3 phase_offsets(int D, int ph, int base) {
    // for o1 in 1:base, o2 in 1:base [, o3 in 1:base] -- last offset fastest; 0-based result
    Off3 o{{0, 0, 0}};
    if (D == 2) {
        o.o[0] = ph / base;
        o.o[1] = ph % base;
    } else {
        o.o[0] = ph / (base * base);
        o.o[1] = (ph / base) % base;
        o.o[2] = ph % base;
    }
    return o;
}

template <int D>
int run_init(jp_ctx* ctx, const jp_cellgrid* g, double* const* coords, uint8_t* index, int npq,
             uint64_t seed) {
    Grid<D> G = to_grid<D>(g);
    k_init<D><<<blocks_for(G.C), kBlock, 0, ctx->stream>>>(G, ptr3(coords, D), index, npq, seed);
    JP_LAUNCHED(ctx);
    return finish(ctx);
}

// occ_fixed: caller-provided occupancy-word scratch of at least C words (the
// allocation-free entry point); null: the context's growable scratch
template <int D>
int run_p2g(jp_ctx* ctx, const jp_cellgrid* g, int nf, double* const* F, const double* const* Fp,
            const double* const* coords, const uint8_t* index, int mode,
            uint64_t* occ_fixed = nullptr) {
    Grid<D> G = to_grid<D>(g);
    int64_t NN = 1;
    for (int d = 0; d < D; ++d) NN *= G.ni[d] + 1;
    ArgPtrs Fa = argptrs(F, nf);
    ArgPtrs Fpa = argptrs(const_cast<double* const*>(Fp), nf);
    int rc;
    uint64_t* occ = occ_fixed;
    if (!occ) {
        if ((rc = ensure_scratch(ctx, (size_t)G.C * 8))) return rc;
        occ = static_cast<uint64_t*>(ctx->scratch);
    }
    if ((rc = occ_words(ctx, index, G.C, G.K, occ))) return rc;
    if (nf == 1)
        k_p2g<D, 1><<<blocks_for(NN), kBlock, 0, ctx->stream>>>(G, nf, Fa, Fpa, cptr3(coords, D),
                                                                occ, mode);
    else
        k_p2g<D, 0><<<blocks_for(NN), kBlock, 0, ctx->stream>>>(G, nf, Fa, Fpa, cptr3(coords, D),
                                                                occ, mode);
    JP_LAUNCHED(ctx);
    return finish(ctx);
}

template <int D>
int run_g2p(jp_ctx* ctx, const jp_cellgrid* g, double* Fp, const double* F,
            const double* const* coords, const uint8_t* index) {
    Grid<D> G = to_grid<D>(g);
    dim3 grid((unsigned)row_blocks<D>(G), (unsigned)G.K);
    k_g2p<D><<<grid, kBlock, 0, ctx->stream>>>(G, Fp, F, cptr3(coords, D), index);
    JP_LAUNCHED(ctx);
    return finish(ctx);
}

template <int D>
int run_inject(jp_ctx* ctx, const jp_cellgrid* gc, double* const* coords, uint8_t* index,
               int nargs, double* const* args, int min_xcell, uint64_t seed, uint64_t counter,
               jp_stats* stats) {
    Grid<D> G = to_grid<D>(gc);
    int rc;
    if ((rc = stats_begin(ctx, stats))) return rc;
    constexpr int NQ = D == 2 ? 4 : 8;
    const int min_q = (min_xcell + NQ - 1) / NQ;
    // cells per colour: at most prod(ceil(ni/2))
    int64_t cap = 1;
    for (int d = 0; d < D; ++d) cap *= (G.ni[d] + 1) / 2;
    InjectLists L;
    if ((rc = inject_lists(ctx, G.C, NQ, cap, 0, L))) return rc;
    const bool two = D == 2 && nargs <= 1;  // the 2D kernel (k_inject2)
    const bool ordered = two && inject2_ordered(ctx);
    if (!ordered) L.state = nullptr;
    k_inject_scan<D><<<blocks_for(G.C), kBlock, 0, ctx->stream>>>(G, cptr3(coords, D), index,
                                                                  min_q, L);
    JP_LAUNCHED(ctx);
    if ((rc = join_segments(ctx, NQ, L))) return rc;
    Ptr3 cp = ptr3(coords, D);
    ArgPtrs ap = argptrs(args, nargs);
    jp_stats* st = stats ? dstats(ctx) : nullptr;
    if constexpr (D == 2) {
        if (two) {
            Inject2 a{};
            a.g = G;
            a.x[0] = coords[0];
            a.x[1] = coords[1];
            a.index = index;
            a.arg = nargs ? args[0] : nullptr;
            a.min_q = min_q;
            a.seed = seed;
            a.counter = counter;
            a.occw = L.occ;
            a.st = st;
            if ((rc = launch_inject2<false>(ctx, a, L, cap))) return rc;
        }
    }
    // colour index of a phase: offsets (o_1 outer ... o_D inner); persistent
    // wavefronts, one listed cell at a time each
    for (int ph = 0; ph < NQ && !two; ++ph) {
        k_inject_wave<D><<<(unsigned)inject_blocks(ctx, cap), kBlock, 0, ctx->stream>>>(
            G, cp, index, nargs, ap, min_q, seed, counter, L.list + (int64_t)ph * cap,
            L.total + ph, st, PhaseInj{});
        JP_LAUNCHED(ctx);
    }
    if ((rc = stats_end(ctx, stats))) return rc;
    if (ordered && stats && (rc = inject_wait_report(ctx))) return rc;
    return finish(ctx);
}

// inject_particles_phase! (src/Particles/injection.jl:146-325): the same
// scan / colour structure, cells listed when any quadrant is short
template <int D>
int run_inject_phase(jp_ctx* ctx, const jp_cellgrid* gc, const jp_axis* xc,
                     double* const* coords, uint8_t* index, double* phases, int nargs,
                     double* const* args, const double* const* fields, uint32_t centred,
                     int min_xcell, uint64_t seed, uint64_t counter, jp_stats* stats) {
    Grid<D> G = to_grid<D>(gc);
    int rc;
    if ((rc = stats_begin(ctx, stats))) return rc;
    constexpr int NQ = D == 2 ? 4 : 8;
    const int min_q = (min_xcell + NQ - 1) / NQ;
    int64_t cap = 1;
    for (int d = 0; d < D; ++d) cap *= (G.ni[d] + 1) / 2;
    InjectLists L;
    if ((rc = inject_lists(ctx, G.C, NQ, cap, 0, L))) return rc;
    L.state = nullptr;  // colour launches only
    k_inject_scan<D, true><<<blocks_for(G.C), kBlock, 0, ctx->stream>>>(G, cptr3(coords, D),
                                                                        index, min_q, L);
    JP_LAUNCHED(ctx);
    if ((rc = join_segments(ctx, NQ, L))) return rc;
    PhaseInj pi{};
    pi.phases = phases;
    pi.centred = centred;
    for (int a = 0; a < nargs; ++a) pi.fields[a] = fields[a];
    for (int d = 0; d < D; ++d) pi.xc[d] = Axis{xc[d].x, xc[d].dx, xc[d].inv_dx, xc[d].n};
    Ptr3 cp = ptr3(coords, D);
    ArgPtrs ap = argptrs(args, nargs);
    jp_stats* st = stats ? dstats(ctx) : nullptr;
    for (int ph = 0; ph < NQ; ++ph) {
        k_inject_wave<D, true><<<(unsigned)inject_blocks(ctx, cap), kBlock, 0, ctx->stream>>>(
            G, cp, index, nargs, ap, min_q, seed, counter, L.list + (int64_t)ph * cap,
            L.total + ph, st, pi);
        JP_LAUNCHED(ctx);
    }
    if ((rc = stats_end(ctx, stats))) return rc;
    return finish(ctx);
}

// cell_index (src/particles.jl:205-217) for n points per dimension:
// abs(trunc((x - x0) / dx)) + 1 with dx = |xv[2] - xv[1]| and x0 = first(xv)
// (subtracted only when nonzero: x - 0.0 would turn -0.0 into 0.0, which the
// result does not see either way).  1-based, as the reference returns it.
struct CellIndexArgs {
    const double* x[3];
    int64_t* out[3];
    const double* xv[3];  // vertex axes (only their first two points are read)
};

__global__ void __launch_bounds__(kBlock) k_cell_index(CellIndexArgs a, int ndim, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int d = 0; d < ndim; ++d) {
        const double x0 = a.xv[d][0];
        const double dx = fabs(a.xv[d][1] - x0);
        const double x = a.x[d][i];
        const double q = x0 == 0.0 ? x / dx : (x - x0) / dx;
        // unsafe_trunc(Int, trunc(q)): finite in-grid quotients only
        const int64_t t = (int64_t)trunc(q);
        a.out[d][i] = (t < 0 ? -t : t) + 1;
    }
}

}  // namespace

extern "C" {

int jp_cell_index(jp_ctx* ctx, int ndim, const jp_axis* xv, int64_t n, const double* const* x,
                  int64_t* const* out) {
    if (!ctx) return JP_ERR_STATE;
    if (ndim < 1 || ndim > 3) return fail(ctx, JP_ERR_ARG, "ndim must be 1, 2 or 3%s");
    if (n < 0) return fail(ctx, JP_ERR_ARG, "negative point count%s");
    if (!xv || !x || !out) return fail(ctx, JP_ERR_ARG, "null argument%s");
    CellIndexArgs a{};
    for (int d = 0; d < ndim; ++d) {
        if (!x[d] || !out[d] || !xv[d].x || xv[d].n < 2)
            return fail(ctx, JP_ERR_ARG, "bad axis or array%s (dim %lld)", "", d);
        a.x[d] = x[d];
        a.out[d] = out[d];
        a.xv[d] = xv[d].x;
    }
    if (n == 0) return JP_OK;
    k_cell_index<<<(unsigned)blocks_for(n), kBlock, 0, ctx->stream>>>(a, ndim, n);
    JP_LAUNCHED(ctx);
    return finish(ctx);
}

int jp_init_particles(jp_ctx* ctx, const jp_cellgrid* g, double* const* coords, uint8_t* index,
                      int np_quadrant, uint64_t seed) {
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if ((rc = check_ptrs(ctx, g->ndim, (const void* const*)coords, "coords"))) return rc;
    if (!index) return fail(ctx, JP_ERR_ARG, "null index%s");
    int nq = g->ndim == 2 ? 4 : 8;
    if (np_quadrant < 0 || np_quadrant * nq > g->max_xcell)
        return fail(ctx, JP_ERR_ARG, "np_quadrant*2^D exceeds max_xcell%s");
    return g->ndim == 2 ? run_init<2>(ctx, g, coords, index, np_quadrant, seed)
                        : run_init<3>(ctx, g, coords, index, np_quadrant, seed);
}

int jp_particle2grid(jp_ctx* ctx, const jp_cellgrid* g, double* F, const double* Fp,
                     const double* const* coords, const uint8_t* index) {
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if ((rc = check_ptrs(ctx, g->ndim, (const void* const*)coords, "coords"))) return rc;
    if (!index || !F || !Fp) return fail(ctx, JP_ERR_ARG, "null field or index%s");
    double* Fa[1] = {F};
    const double* Fpa[1] = {Fp};
    int mode = g->ndim == 2 ? 0 : 1;
    return g->ndim == 2 ? run_p2g<2>(ctx, g, 1, Fa, Fpa, coords, index, mode)
                        : run_p2g<3>(ctx, g, 1, Fa, Fpa, coords, index, mode);
}

int jp_particle2grid_noalloc(jp_ctx* ctx, const jp_cellgrid* g, double* F, const double* Fp,
                             const double* const* coords, const uint8_t* index) {
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if ((rc = check_ptrs(ctx, g->ndim, (const void* const*)coords, "coords"))) return rc;
    if (!index || !F || !Fp) return fail(ctx, JP_ERR_ARG, "null field or index%s");
    int64_t C = 1;
    for (int d = 0; d < g->ndim; ++d) C *= g->ni[d];
    if (C > JP_NOALLOC_MAX_CELLS)
        return fail(ctx, JP_ERR_ARG,
                    "particle2grid without allocation supports at most JP_NOALLOC_MAX_CELLS%s "
                    "= 2097152 cells (got %lld)",
                    "", (long long)C);
    if (!ctx->noalloc) return fail(ctx, JP_ERR_STATE, "fixed scratch address unknown%s");
    double* Fa[1] = {F};
    const double* Fpa[1] = {Fp};
    const int mode = g->ndim == 2 ? 0 : 1;
    return g->ndim == 2 ? run_p2g<2>(ctx, g, 1, Fa, Fpa, coords, index, mode, ctx->noalloc)
                        : run_p2g<3>(ctx, g, 1, Fa, Fpa, coords, index, mode, ctx->noalloc);
}

int jp_particle2grid_multi(jp_ctx* ctx, const jp_cellgrid* g, int nfields, double* const* F,
                           const double* const* Fp, const double* const* coords,
                           const uint8_t* index) {
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if ((rc = check_ptrs(ctx, g->ndim, (const void* const*)coords, "coords"))) return rc;
    if (!index) return fail(ctx, JP_ERR_ARG, "null index%s");
    if (nfields < 1 || nfields > JP_MAX_ARGS)
        return fail(ctx, JP_ERR_ARG, "nfields out of range%s");
    if ((rc = check_ptrs(ctx, nfields, (const void* const*)F, "F"))) return rc;
    if ((rc = check_ptrs(ctx, nfields, (const void* const*)Fp, "Fp"))) return rc;
    return g->ndim == 2 ? run_p2g<2>(ctx, g, nfields, F, Fp, coords, index, 1)
                        : run_p2g<3>(ctx, g, nfields, F, Fp, coords, index, 1);
}

int jp_grid2particle(jp_ctx* ctx, const jp_cellgrid* g, double* Fp, const double* F,
                     const double* const* coords, const uint8_t* index) {
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if ((rc = check_ptrs(ctx, g->ndim, (const void* const*)coords, "coords"))) return rc;
    if (!index || !F || !Fp) return fail(ctx, JP_ERR_ARG, "null field or index%s");
    return g->ndim == 2 ? run_g2p<2>(ctx, g, Fp, F, coords, index)
                        : run_g2p<3>(ctx, g, Fp, F, coords, index);
}

int jp_inject_particles(jp_ctx* ctx, const jp_cellgrid* g, double* const* coords, uint8_t* index,
                        int nargs, double* const* args, int min_xcell, uint64_t seed,
                        uint64_t call_counter, jp_stats* stats) {
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if ((rc = check_ptrs(ctx, g->ndim, (const void* const*)coords, "coords"))) return rc;
    if (!index) return fail(ctx, JP_ERR_ARG, "null index%s");
    if (nargs < 0 || nargs > JP_MAX_ARGS) return fail(ctx, JP_ERR_ARG, "nargs out of range%s");
    if (nargs && (rc = check_ptrs(ctx, nargs, (const void* const*)args, "args"))) return rc;
    if (min_xcell < 0) return fail(ctx, JP_ERR_ARG, "min_xcell must be non-negative%s");
    return g->ndim == 2
               ? run_inject<2>(ctx, g, coords, index, nargs, args, min_xcell, seed, call_counter, stats)
               : run_inject<3>(ctx, g, coords, index, nargs, args, min_xcell, seed, call_counter, stats);
}

int jp_inject_particles_phase(jp_ctx* ctx, const jp_cellgrid* g, const jp_axis* xc,
                              double* const* coords, uint8_t* index, double* phases, int nargs,
                              double* const* args, const double* const* fields,
                              const int* centred, int min_xcell, uint64_t seed,
                              uint64_t call_counter, jp_stats* stats) {
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if ((rc = check_ptrs(ctx, g->ndim, (const void* const*)coords, "coords"))) return rc;
    if (!index || !phases || !xc) return fail(ctx, JP_ERR_ARG, "null index, phases or xc%s");
    if (nargs < 0 || nargs > JP_MAX_ARGS) return fail(ctx, JP_ERR_ARG, "nargs out of range%s");
    if (nargs && (rc = check_ptrs(ctx, nargs, (const void* const*)args, "args"))) return rc;
    if (nargs && (rc = check_ptrs(ctx, nargs, (const void* const*)fields, "fields"))) return rc;
    if (nargs && !centred) return fail(ctx, JP_ERR_ARG, "null centred flags%s");
    if (min_xcell < 0) return fail(ctx, JP_ERR_ARG, "min_xcell must be non-negative%s");
    uint32_t cmask = 0;
    bool any_centred = false;
    for (int a = 0; a < nargs; ++a)
        if (centred[a]) {
            cmask |= 1u << a;
            any_centred = true;
        }
    for (int d = 0; d < g->ndim; ++d) {
        if (any_centred && g->ni[d] < 2)
            return fail(ctx, JP_ERR_ARG, "a cell-centred field needs two cells per dimension%s");
        if (!xc[d].x || !xc[d].inv_dx || xc[d].n != g->ni[d])
            return fail(ctx, JP_ERR_ARG, "centre axis%s %lld must have ni points", "", d);
    }
    return g->ndim == 2
               ? run_inject_phase<2>(ctx, g, xc, coords, index, phases, nargs, args, fields, cmask,
                                     min_xcell, seed, call_counter, stats)
               : run_inject_phase<3>(ctx, g, xc, coords, index, phases, nargs, args, fields, cmask,
                                     min_xcell, seed, call_counter, stats);
}

}  // extern "C"
