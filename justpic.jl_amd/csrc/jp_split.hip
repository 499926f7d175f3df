// jp_split.hip -- FP32-split particle positions (build-only precision form).
//
// A particle's position is the cell of its slot (an integer, implicit in the
// storage) plus an offset u in cell units, stored as float32.  Velocity
// interpolation and the RK stages run in float32 in cell units; FP64 is kept
// for the conversions to and from absolute coordinates, for the offset
// renormalisation when a particle changes cell (move), and for the weight sums
// of particle2grid.  Requires uniform vertex axes and the standard staggering
// of the reference examples: component c on the vertex axis along c and on
// the centre axis extended by one ghost point on each side elsewhere
// (README.md:42-44).
#include "jp_ctx.hpp"
#include "jp_inject2.hpp"
#include "jp_splitmath.hpp"

#include <algorithm>
#include <cstring>
#include <type_traits>
#include <vector>

using namespace jp;

namespace {

struct FPtr3 {
    float* p[3];
};

template <int D>
struct SGrid {
    int ni[D];
    int64_t C;
    int K;
    const double* h[D];  // spacing arrays of the (uniform) vertex axes; h[d][0] is used
};

template <int D>
SGrid<D> sgrid(const Grid<D>& g) {
    SGrid<D> s;
    for (int d = 0; d < D; ++d) {
        s.ni[d] = (int)g.ni[d];
        s.h[d] = g.xv[d].dx;
    }
    s.C = g.C;
    s.K = g.K;
    return s;
}

__device__ inline float fnan() { return __builtin_nanf(""); }
__device__ inline double dnan() { return __builtin_nan(""); }

// ------------------------------------------------------------ geometry
// Deviation of an axis from uniform spacing, relative to its mean spacing h =
// (x[n-1] - x[0]) / (n-1): max over k of |x[k] - (x[0] + k h)| and, with
// with_dx, of |dx[k] - h| (the spacing the reference stores in di).  A NaN or
// a non-increasing axis gives +Inf.  All threads of the workgroup call it.
__device__ double axis_deviation(const Axis& a, bool with_dx, double& h, double* red) {
    const double x0 = a.x[0], xn = a.x[a.n - 1];
    h = (xn - x0) / (double)(a.n - 1);
    double m = 0.0;
    for (int64_t k = threadIdx.x; k < a.n; k += blockDim.x) {
        const double e = fabs(a.x[k] - fma((double)k, h, x0));
        m = e == e ? fmax(m, e) : __builtin_inf();
    }
    if (with_dx)
        for (int64_t k = threadIdx.x; k + 1 < a.n; k += blockDim.x) {
            const double e = fabs(a.dx[k] - h);
            m = e == e ? fmax(m, e) : __builtin_inf();
        }
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) m = fmax(m, __shfl_xor(m, s, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    double r = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r = fmax(r, red[w]);
    __syncthreads();
    return (h > 0.0 && h < __builtin_inf()) ? r / h : __builtin_inf();
}

// One workgroup: checks that the vertex axes are uniform (split offsets are in
// cell units of one spacing per dimension) and, with velocity axes, that each
// of them is uniform, and derives their affine maps (SplitGeom).
template <int D>
__global__ void __launch_bounds__(kBlock) k_split_geom(Grid<D> g, Vel<D> v, int with_vel,
                                                       SplitGeom* __restrict__ geo,
                                                       int32_t* __restrict__ sticky) {
    __shared__ double red[kBlock / 64];
    int bad = 0;
    double h[D];
#pragma unroll
    for (int d = 0; d < D; ++d)
        if (!(axis_deviation(g.xv[d], true, h[d], red) <= kGeomTol)) bad |= kGeomVertex;
    double r[D][D], o[D][D];
#pragma unroll
    for (int c = 0; c < D; ++c)
#pragma unroll
        for (int d = 0; d < D; ++d) {
            r[c][d] = 1.0;
            o[c][d] = 0.0;
            if (!with_vel) continue;
            double hv;
            if (!(axis_deviation(v.ax[c][d], false, hv, red) <= kGeomTol)) {
                bad |= kGeomVelocity;
                continue;
            }
            r[c][d] = h[d] / hv;
            o[c][d] = (g.xv[d].x[0] - v.ax[c][d].x[0]) / hv;
        }
    if (threadIdx.x == 0) {
        geo->bad = bad;
        *sticky |= bad;
#pragma unroll
        for (int c = 0; c < D; ++c)
#pragma unroll
            for (int d = 0; d < D; ++d) {
                geo->r[c][d] = r[c][d];
                geo->o[c][d] = o[c][d];
            }
    }
}

// ------------------------------------------------------------ conversions
template <int D>
__global__ void __launch_bounds__(kBlock) k_to_split(Grid<D> g, CPtr3 x, const uint8_t* index,
                                                     FPtr3 u) {
    int64_t c, I[D];
    if (!row_cell<D>(g, c, I)) return;
    const int64_t e = c + g.C * (int64_t)blockIdx.y;
    if (!index[e]) {
#pragma unroll
        for (int d = 0; d < D; ++d) u.p[d][e] = fnan();
        return;
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const double ud = (x.p[d][e] - g.xv[d].x[I[d]]) * g.xv[d].idx[I[d]];
        float uf = (float)ud;
        // a point inside its cell stays inside: 1 - 2^-25 < ud < 1 would round up to 1
        if (ud < 1.0 && uf >= 1.0f) uf = 0x1.fffffep-1f;
        u.p[d][e] = uf;
    }
}

template <int D>
__global__ void __launch_bounds__(kBlock) k_from_split(Grid<D> g, FPtr3 u, const uint8_t* index,
                                                       Ptr3 x) {
    int64_t c, I[D];
    if (!row_cell<D>(g, c, I)) return;
    const int64_t e = c + g.C * (int64_t)blockIdx.y;
    if (!index[e]) {
#pragma unroll
        for (int d = 0; d < D; ++d) x.p[d][e] = dnan();
        return;
    }
#pragma unroll
    for (int d = 0; d < D; ++d)
        x.p[d][e] = g.xv[d].x[I[d]] + (double)u.p[d][e] * g.xv[d].dx[I[d]];
}

// the per-particle arithmetic: jp_splitmath.hpp
using namespace jp::split;

// advection of the particles in slots kb..ke-1 of tile `tile` of one grid (all
// threads of the block call it; the LDS window sV is the block's)
template <int D, int M, class TS, bool STD = false>
__device__ inline void adv_split_tile(const SGrid<D>& g, const SVelIn<D>& V, const SMap<D>& m,
                                      double dt, const FPtr3& off,
                                      const uint8_t* __restrict__ index, const RKf& P, int tile,
                                      int kb, int ke, float (*sV)[TS::size()]) {
    int t0[D];
    {
        int b = tile;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int nt = (g.ni[d] + TS::n(d) - 1) / TS::n(d);
            t0[d] = (b % nt) * TS::n(d);
            b /= nt;
        }
    }
    double sc[D];
#pragma unroll
    for (int c = 0; c < D; ++c) sc[c] = dt / g.h[c][0];
    int w0[D][D];
#pragma unroll
    for (int c = 0; c < D; ++c)
#pragma unroll
        for (int d = 0; d < D; ++d) w0[c][d] = window_origin<D, STD>(m, c, d, t0[d]);
    const int tid = threadIdx.x;
    // window fill: every load issued before the LDS stores
    constexpr int NI = (TS::size() + kBlock - 1) / kBlock;
    float wv[D][NI];
#pragma unroll
    for (int c = 0; c < D; ++c) {
        const double* __restrict__ F = V.v[c];
#pragma unroll
        for (int it = 0; it < NI; ++it) {
            const int i = tid + it * kBlock;
            int r = i;
            int64_t node = 0, stride = 1;
            bool ok = i < TS::size();
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int k = w0[c][d] + r % TS::w(d);
                r /= TS::w(d);
                ok = ok && k >= 0 && k < V.n[c][d];
                node += k * stride;
                stride *= V.n[c][d];
            }
            wv[c][it] = ok ? (float)(F[ok ? node : 0] * sc[c]) : 0.0f;
        }
    }
#pragma unroll
    for (int c = 0; c < D; ++c)
#pragma unroll
        for (int it = 0; it < NI; ++it)
            if (tid + it * kBlock < TS::size()) sV[c][tid + it * kBlock] = wv[c][it];
    __syncthreads();
    // the cell of thread t of the tile, its linear index, and whether it is
    // in the grid
    auto cell_of = [&](int t, int (&J)[D], int64_t& cj) {
        bool in = true;
        int r = t;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            J[d] = t0[d] + r % TS::n(d);
            r /= TS::n(d);
            in = in && J[d] < g.ni[d];
        }
        cj = J[D - 1];
#pragma unroll
        for (int d = D - 2; d >= 0; --d) cj = cj * g.ni[d] + J[d];
        return in;
    };
    int I[D];
    int64_t c;
    const bool in = cell_of(tid, I, c);
    const int64_t C = g.C;
    // a thread outside the grid owns no particle but still takes work
    const uint64_t act = !in ? 0ull
                             : ((kb == 0 && ke == g.K) ? active_slots(index, c, C, g.K)
                                                       : active_slots_in(index, c, C, kb, ke));
    // Chunks of CH slot planes.  The wavefront's active particles of a chunk
    // are listed (plane-major, then lane) and dealt to its lanes in turn, so
    // that every lane advects about the same number of particles (a lane per
    // cell walking its own slots idles whenever its cell has fewer active
    // slots in the chunk than the fullest cell); consecutive list entries are
    // neighbouring cells of one plane, so the loads and stores stay
    // coalesced.  The list is written and read by the same wavefront (its LDS
    // operations complete in order).
    constexpr int CH = D == 2 ? 8 : 4;  // 3D: fewer positions in registers
    __shared__ uint16_t s_items[kBlock / 64][CH * 64];
    uint16_t* items = s_items[tid >> 6];
    const int lane = tid & 63, wbase = tid & ~63;
    for (int k0 = kb; k0 < ke; k0 += CH) {
        int n = 0;
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            const bool a = k0 + j < 64 && ((act >> (k0 + j)) & 1);
            const uint64_t bal = __ballot(a);
            const int below = (int)__builtin_amdgcn_mbcnt_hi(
                (uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
            if (a) items[n + below] = (uint16_t)((j << 6) | lane);
            n += __popcll(bal);
        }
        if (!n) continue;
        // this lane's entries lane, lane + 64, ...: positions loaded first
        float uq[CH][D];
        int dq[CH];
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const int item = lane + 64 * q;
            dq[q] = item < n ? items[item] : -1;
            int J[D];
            int64_t cj;
            cell_of(wbase + (dq[q] & 63), J, cj);
            const int64_t e = cj + C * (k0 + (dq[q] >> 6));
#pragma unroll
            for (int d = 0; d < D; ++d) uq[q][d] = dq[q] >= 0 ? off.p[d][e] : 0.0f;
        }
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            if (dq[q] < 0) break;
            int J[D];
            int64_t cj;
            cell_of(wbase + (dq[q] & 63), J, cj);
            CellMap<D> cm;
            cell_map<D, STD>(m, J, cm);
            float out[D];
            advect_split_one<D, M, TS>(sV, V, sc, w0, J, cm, P, uq[q], out);
            const int64_t e = cj + C * (k0 + (dq[q] >> 6));
#pragma unroll
            for (int d = 0; d < D; ++d) off.p[d][e] = out[d];
        }
    }
}

// blockIdx.y: slot group of `span` slots (a multiple of 8), so that grids with
// few tiles still fill the chip; each group stages its own velocity window
// STD: axes accepted with the standard staggering (the maps are constants)
template <int D, int M, bool STD>
__global__ void __launch_bounds__(kBlock) k_adv_split(SGrid<D> g, SVelIn<D> V,
                                                      const SplitGeom* __restrict__ geo, double dt,
                                                      FPtr3 off, const uint8_t* __restrict__ index,
                                                      RKf P, int span) {
    using TS = DefaultTile<D>;
    __shared__ float sV[D][TS::size()];
    if (geo->bad) return;  // grid rejected: nothing changes
    SMap<D> m;
#pragma unroll
    for (int c = 0; c < D; ++c)
#pragma unroll
        for (int d = 0; d < D; ++d) {
            m.r[c][d] = STD ? 1.0 : geo->r[c][d];
            m.o[c][d] = STD ? (c == d ? 0.0 : 0.5) : geo->o[c][d];
        }
    const int kb = blockIdx.y * span;
    const int ke = kb + span < g.K ? kb + span : g.K;
    adv_split_tile<D, M, TS, STD>(g, V, m, dt, off, index, P, blockIdx.x, kb, ke, sV);
}

// ------------------------------------------------------------ ensemble
// Many independent 2D grids advected in one persistent launch: block b walks
// the global tile list t = b, b + gridDim.x, ...; tile t belongs to grid
// tile_grid[t] and is that grid's tile t - tile0[grid].
struct EnsGrid {
    SGrid<2> g;
    SVelIn<2> V;
    SMap<2> m;
    double dt;
    FPtr3 off;
    const uint8_t* index;
};

template <int M, class TS>
__global__ void __launch_bounds__(kBlock) k_adv_split_ens(const EnsGrid* __restrict__ grids,
                                                          const uint16_t* __restrict__ tile_grid,
                                                          const int32_t* __restrict__ tile0,
                                                          int total, RKf P, int zero_dt) {
    __shared__ float sV[2][TS::size()];
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int i = tile_grid[t];
        const EnsGrid& E = grids[i];
        // zero_dt: the autotuner's timing runs (same work, positions unchanged)
        adv_split_tile<2, M, TS>(E.g, E.V, E.m, zero_dt ? 0.0 : E.dt, E.off, E.index, P, t - tile0[i],
                                 0, E.g.K,
                                 sV);
        __syncthreads();  // the window is refilled for the next tile
    }
}

// ------------------------------------------------------------ particle2grid
// Inverse-distance-squared weights in float32 from the offsets (hardware
// reciprocal, about 1 ulp), weight and weighted-field sums in FP64.  Two
// passes, so that every particle is read once (a per-node gather would read
// it from each of its cell's 2^D nodes):
//  - k_p2g_split_cells: one thread per cell (row mapping: a wavefront's 64
//    cells share each slot plane, so the loads coalesce) walks its active
//    slots in chunks, the next chunk's loads in flight, and accumulates
//    sum(w) and sum(w F) for each of the cell's 2^D corners; the partial sums
//    go to part[(2 q + {0, 1}) C + c];
//  - k_p2g_split_nodes: one thread per node adds the partials of its 2^D
//    cells (in the reference's cell order) and divides.  A node without
//    particles gets 0/0 = NaN.
template <int D>
__global__ void __launch_bounds__(kBlock) k_p2g_split_cells(Grid<D> gg, SGrid<D> g,
                                                            const double* __restrict__ Fp,
                                                            FPtr3 off,
                                                            const uint8_t* __restrict__ index,
                                                            double* __restrict__ part) {
    int64_t c, I[D];
    if (!row_cell<D>(gg, c, I)) return;
    const int64_t C = g.C;
    const int K = g.K;
    const uint64_t act = active_slots(index, c, C, K);
    float hf[D];
#pragma unroll
    for (int d = 0; d < D; ++d) hf[d] = (float)g.h[d][0];
    constexpr int NC = 1 << D;
    constexpr int B = D == 2 ? 8 : 4;
    double W[NC], S[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) W[q] = S[q] = 0.0;
    // inactive slots load offset 2 in every dimension (so s > 0) and add a
    // zero weight: the accumulation is branch-free
    auto load = [&](int k0, float (&ub)[B][D], double (&fb)[B]) __attribute__((always_inline)) {
        const uint32_t bq = (uint32_t)(act >> k0) & ((1u << B) - 1);
#pragma unroll
        for (int j = 0; j < B; ++j) {
            const bool on = (bq >> j) & 1;
            const int64_t e = c + C * (k0 + j);
#pragma unroll
            for (int d = 0; d < D; ++d) ub[j][d] = on ? off.p[d][e] : 2.0f;
            fb[j] = on ? Fp[e] : 0.0;
        }
    };
    auto accumulate = [&](int k0, const float (&ub)[B][D], const double (&fb)[B])
                          __attribute__((always_inline)) {
        const uint32_t bq = (uint32_t)(act >> k0) & ((1u << B) - 1);
#pragma unroll
        for (int j = 0; j < B; ++j) {
            const bool on = (bq >> j) & 1;
#pragma unroll
            for (int q = 0; q < NC; ++q) {
                float sq = 0.0f;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const float r = ((float)((q >> d) & 1) - ub[j][d]) * hf[d];
                    sq = fmaf(r, r, sq);
                }
                const double w = on ? (double)__builtin_amdgcn_rcpf(sq) : 0.0;
                W[q] += w;
                S[q] = fma(w, fb[j], S[q]);
            }
        }
    };
    // chunks that still hold an active slot
    auto more = [&](int k) __attribute__((always_inline)) {
        return k < K && (act >> k) != 0ull;
    };
    if (act != 0ull) {
        float u0[B][D], u1[B][D];
        double f0[B], f1[B];
        int k0 = 0;
        load(0, u0, f0);
#pragma unroll 1
        for (;;) {
            const int k1 = k0 + B;
            const bool m1 = more(k1);
            if (m1) load(k1, u1, f1);
            accumulate(k0, u0, f0);
            if (!m1) break;
            const int k2 = k1 + B;
            const bool m2 = more(k2);
            if (m2) load(k2, u0, f0);
            accumulate(k1, u1, f1);
            if (!m2) break;
            k0 = k2;
        }
    }
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        part[(2 * q) * C + c] = W[q];
        part[(2 * q + 1) * C + c] = S[q];
    }
}

template <int D>
__global__ void __launch_bounds__(kBlock) k_p2g_split_nodes(Grid<D> gg, SGrid<D> g,
                                                            const double* __restrict__ part,
                                                            const SplitGeom* __restrict__ geo,
                                                            double* __restrict__ F) {
    if (geo->bad) return;  // grid rejected: F unchanged
    int64_t nn[D];
    int64_t NN = 1;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        nn[d] = g.ni[d] + 1;
        NN *= nn[d];
    }
    const int64_t node = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
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
    const int64_t C = g.C;
    double W = 0.0, S = 0.0;
    // cells (N - 1 .. N) per dimension, dimension 0 fastest; the node is the
    // lower corner (bit 0) of cell N and the upper corner (bit 1) of cell N - 1
#pragma unroll
    for (int q = 0; q < (1 << D); ++q) {
        int64_t I[D];
        bool ok = true;
        int cq = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const bool lower = (q >> d) & 1;
            I[d] = lower ? N[d] : N[d] - 1;
            ok = ok && I[d] >= 0 && I[d] < g.ni[d];
            cq |= (lower ? 0 : 1) << d;
        }
        if (!ok) continue;
        const int64_t cell = lin<D>(gg.ni, I);
        W += part[(2 * cq) * C + cell];
        S += part[(2 * cq + 1) * C + cell];
    }
    F[node] = S / W;
}

// ------------------------------------------------------------ grid2particle
template <int D>
__global__ void __launch_bounds__(kBlock) k_g2p_split(Grid<D> g, double* __restrict__ Fp,
                                                      const double* __restrict__ F, FPtr3 off,
                                                      const uint8_t* __restrict__ index) {
    int64_t c, I[D];
    if (!row_cell<D>(g, c, I)) return;
    const int64_t e = c + g.C * (int64_t)blockIdx.y;
    if (!index[e]) return;
    double t[D];
#pragma unroll
    for (int d = 0; d < D; ++d) t[d] = (double)off.p[d][e];
    const int64_t n0 = g.ni[0] + 1;
    if constexpr (D == 2) {
        const int64_t b = I[0] + n0 * I[1];
        Fp[e] = lerp1(t[1], lerp1(t[0], F[b], F[b + 1]), lerp1(t[0], F[b + n0], F[b + n0 + 1]));
    } else {
        const int64_t n1 = g.ni[1] + 1, s = n0 * n1;
        const int64_t b = I[0] + n0 * (I[1] + n1 * I[2]);
        const double lo =
            lerp1(t[1], lerp1(t[0], F[b], F[b + 1]), lerp1(t[0], F[b + n0], F[b + n0 + 1]));
        const double hi = lerp1(t[1], lerp1(t[0], F[b + s], F[b + s + 1]),
                                lerp1(t[0], F[b + s + n0], F[b + s + n0 + 1]));
        Fp[e] = lerp1(t[2], lo, hi);
    }
}

// ------------------------------------------------------------ injection
// Same structure as the FP64 injection: a scan lists the cells whose first
// quadrant is short (0 < u < 1/2 in every dimension), then one wavefront per
// listed cell fills free slots with vertex + dq*muladd(0.95, rand, 0.05)
// (in cell units: q/2 + muladd(0.95, rand, 0.05)/2) and copies the fields of
// the nearest particle (distances in physical units, FP64).
template <int D>
__global__ void __launch_bounds__(kBlock) k_inj_scan_split(Grid<D> g, FPtr3 off,
                                                           const uint8_t* __restrict__ index,
                                                           int min_q, InjectLists L,
                                                           const SplitGeom* __restrict__ geo) {
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    // all lanes stay for the wavefront append; a rejected grid lists no cell
    const bool valid = c < g.C && !geo->bad;
    int64_t I[D];
    unlin<D>(g.ni, valid ? c : 0, I);
    int num = 0;
    if (valid) {
        const uint64_t act = active_slots(index, c, g.C, g.K);
        float bl[D], bh[D];  // box of the cell's offsets
#pragma unroll
        for (int d = 0; d < D; ++d) {
            bl[d] = __builtin_inff();
            bh[d] = -__builtin_inff();
        }
        constexpr int B = 16;
        for (int k0 = 0; k0 < g.K; k0 += B) {
            float ub[B][D];
#pragma unroll
            for (int j = 0; j < B; ++j) {
                const bool on = (act >> (k0 + j)) & 1;
#pragma unroll
                for (int d = 0; d < D; ++d) ub[j][d] = on ? off.p[d][c + g.C * (k0 + j)] : 0.0f;
            }
#pragma unroll
            for (int j = 0; j < B; ++j) {
                const bool on = (act >> (k0 + j)) & 1;
                bool inside = on;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    inside = inside & (0.0f < ub[j][d] && ub[j][d] < 0.5f);
                    if (D == 2 && on) {
                        bl[d] = fminf(bl[d], ub[j][d]);
                        bh[d] = fmaxf(bh[d], ub[j][d]);
                    }
                }
                num += inside ? 1 : 0;
            }
        }
        L.occ[c] = act;
        if (D == 2 && L.box) {
            float* __restrict__ bx = static_cast<float*>(L.box);
#pragma unroll
            for (int d = 0; d < D; ++d) {
                bx[(2 * d) * g.C + c] = bl[d];
                bx[(2 * d + 1) * g.C + c] = bh[d];
            }
        }
    }
    if (valid && L.state) L.state[c] = num < min_q ? 1u : 0u;
    int col = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) col = col * 2 + (int)(I[d] & 1);
    wave_segment<D == 2 ? 4 : 8>(valid && num < min_q, col, (int32_t)c, c >> 6, L.nwave, L.seg,
                                 L.cnt);
}

template <int D>
__global__ void __launch_bounds__(kBlock) k_inj_wave_split(Grid<D> g, SGrid<D> sg, FPtr3 off,
                                                           uint8_t* index, int nargs, ArgPtrs args,
                                                           int min_q, uint64_t seed,
                                                           uint64_t counter,
                                                           const int32_t* __restrict__ list,
                                                           const int32_t* __restrict__ count,
                                                           jp_stats* st) {
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
    const bool own = lane < K;
    const int64_t e = c + C * lane;
    bool act = own && index[e];
    float u[D];
#pragma unroll
    for (int d = 0; d < D; ++d) u[d] = act ? off.p[d][e] : 0.0f;
    for (int q = 0; q < NQ; ++q) {
        float q0[D];
#pragma unroll
        for (int d = 0; d < D; ++d) q0[d] = 0.5f * (float)((q >> d) & 1);
        bool inside = act;
#pragma unroll
        for (int d = 0; d < D; ++d) inside = inside & (q0[d] < u[d] && u[d] < q0[d] + 0.5f);
        int num = __popcll(__ballot(inside));
        if (num >= min_q) break;
        uint64_t freeS = ~__ballot(act) & full;
        while (freeS) {
            const int ip = __builtin_ctzll(freeS);
            freeS &= freeS - 1;
            ++num;
            float un[D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const double r = uniform53(seed, (uint64_t)c, ip, d, 2, counter);
                un[d] = (float)((double)q0[d] + 0.5 * fma(0.95, r, 0.05));
            }
            const int64_t enew = c + C * ip;
            if (lane == ip) {
                act = true;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    u[d] = un[d];
                    off.p[d][enew] = un[d];
                }
                index[enew] = 1;
            }
            double best = __builtin_inf();
            int border = 0x7fffffff;
            constexpr int NBB = 9;  // neighbour cells per batch, one round of loads each
#pragma unroll
            for (int nb0 = 0; nb0 < NB; nb0 += NBB) {
                bool a[NBB];
                int64_t e2[NBB];
                int o[NBB][D];
#pragma unroll
                for (int v = 0; v < NBB; ++v) {
                    int64_t J[D];
                    bool ok = own;
                    int r = nb0 + v;
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        o[v][d] = (r % 3) - 1;
                        J[d] = I[d] + o[v][d];
                        r /= 3;
                        ok = ok && J[d] >= 0 && J[d] < g.ni[d];
                    }
                    e2[v] = ok ? lin<D>(g.ni, J) + C * lane : 0;
                    a[v] = ok;
                }
                uint8_t m[NBB];
                float uq[NBB][D];
#pragma unroll
                for (int v = 0; v < NBB; ++v) {
                    const bool other = a[v] && nb0 + v != CENTER;
                    m[v] = other ? index[e2[v]] : (uint8_t)0;
#pragma unroll
                    for (int d = 0; d < D; ++d) uq[v][d] = other ? off.p[d][e2[v]] : u[d];
                }
#pragma unroll
                for (int v = 0; v < NBB; ++v) {
                    const bool av = (nb0 + v == CENTER) ? (act && lane != ip) : (m[v] != 0);
                    if (!av) continue;
                    double s2 = 0.0;
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        const double r2 =
                            ((double)o[v][d] + (double)uq[v][d] - (double)un[d]) * sg.h[d][0];
                        s2 = s2 + r2 * r2;
                    }
                    const double dd = sqrt(s2);
                    if (dd < best) {
                        best = dd;
                        border = (nb0 + v) * 64 + lane;
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
            if (border == 0x7fffffff) {
                ++n_orph;
            } else if (lane == ip) {
                int64_t J[D];
                int r = border >> 6;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    J[d] = I[d] + (r % 3) - 1;
                    r /= 3;
                }
                const int64_t esrc = lin<D>(g.ni, J) + C * (border & 63);
                for (int a = 0; a < nargs; ++a) args.p[a][enew] = args.p[a][esrc];
            }
            __threadfence_block();
            ++n_inj;
            if (num >= min_q) break;
        }
    }
    }  // work list
    if (lane == 0) {
        stat_add(st, ST_INJECTED, n_inj);
        stat_add(st, ST_ORPHAN, n_orph);
    }
}

// ------------------------------------------------------------ host side
}  // namespace

namespace jp {

namespace {
int geom_reject(jp_ctx* ctx, int bits);

template <int D>
void launch_geom(jp_ctx* ctx, const jp_cellgrid* gc, const jp_velgrid* vg, SplitGeom* slot) {
    Grid<D> G = to_grid<D>(gc);
    Vel<D> v{};
    if (vg)
        for (int c = 0; c < D; ++c)
            for (int d = 0; d < D; ++d)
                v.ax[c][d] = Axis{vg->ax[c][d].x, vg->ax[c][d].dx, vg->ax[c][d].inv_dx,
                                  vg->ax[c][d].n};
    k_split_geom<D><<<1, kBlock, 0, ctx->stream>>>(G, v, vg != nullptr, slot,
                                                   &dgeom(ctx)->sticky);
}
}  // namespace

int split_geometry(jp_ctx* ctx, const jp_cellgrid* gc, const jp_velgrid* vg) {
    const int D = gc->ndim;
    if (vg)
        for (int c = 0; c < D; ++c)
            for (int d = 0; d < D; ++d)
                if (!vg->ax[c][d].x || vg->ax[c][d].n < 2 || vg->ax[c][d].n > (1 << 30))
                    return fail(ctx, JP_ERR_ARG, "velocity axis%s %lld needs at least two points",
                                "", c * 3 + d);
    GeomKey key;
    std::memset(&key, 0, sizeof key);
    key.ndim = D;
    key.with_vel = vg != nullptr;
    for (int d = 0; d < D; ++d) {
        key.ni[d] = gc->ni[d];
        key.xv[d] = gc->xv[d];
    }
    if (vg)
        for (int c = 0; c < D; ++c)
            for (int d = 0; d < D; ++d) key.vel[c][d] = vg->ax[c][d];
    const uint64_t gen = free_generation();
    // FNV-1a over the key's bytes
    uint64_t hsh = 1469598103934665603ull;
    {
        const unsigned char* b = reinterpret_cast<const unsigned char*>(&key);
        for (size_t i = 0; i < sizeof key; ++i) hsh = (hsh ^ b[i]) * 1099511628211ull;
    }
    int slot = -1;
    {
        auto it = ctx->geo_index.find(hsh);
        if (it != ctx->geo_index.end()) {
            const GeomEntry& e = ctx->geo_cache[it->second];
            if (e.state != 0 && e.gen == gen && std::memcmp(&e.key, &key, sizeof key) == 0)
                slot = it->second;
        }
    }
    if (slot >= 0 && !ctx->opt_geom_recheck && ctx->geo_cache[slot].state == 1) {
        // accepted before: no check; a captured graph keeps reading this slot
        GeomEntry& e = ctx->geo_cache[slot];
        ctx->geo_cur = slot;
        e.used = ++ctx->geo_clock;
        if (ctx->capturing) e.pinned = true;
        return JP_OK;
    }
    if (slot < 0) {
        // an empty slot, else the least recently used one no graph reads
        slot = -1;
        for (int i = 0; i < kGeomSlots; ++i) {
            const GeomEntry& a = ctx->geo_cache[i];
            if (a.pinned) continue;
            if (a.state == 0) {
                slot = i;
                break;
            }
            if (slot < 0 || a.used < ctx->geo_cache[slot].used) slot = i;
        }
        if (slot < 0)
            return fail(ctx, JP_ERR_STATE,
                        "every split geometry slot is read by a captured graph%s (%lld)", "",
                        kGeomSlots);
        const GeomEntry& old = ctx->geo_cache[slot];
        if (old.state != 0) {
            uint64_t oh = 1469598103934665603ull;
            const unsigned char* b = reinterpret_cast<const unsigned char*>(&old.key);
            for (size_t i = 0; i < sizeof old.key; ++i) oh = (oh ^ b[i]) * 1099511628211ull;
            auto it = ctx->geo_index.find(oh);
            if (it != ctx->geo_index.end() && it->second == slot) ctx->geo_index.erase(it);
        }
    }
    ctx->geo_index[hsh] = slot;
    GeomEntry& e = ctx->geo_cache[slot];
    SplitGeom* dev = static_cast<SplitGeom*>(ctx->geo_dev) + slot;
    if (D == 2)
        launch_geom<2>(ctx, gc, vg, dev);
    else
        launch_geom<3>(ctx, gc, vg, dev);
    JP_LAUNCHED(ctx);
    e.key = key;
    e.gen = gen;
    e.used = ++ctx->geo_clock;
    e.standard = false;
    ctx->geo_cur = slot;
    if (ctx->capturing) {
        // the verdict stays on the device (the graph's kernels check it); the
        // graph writes and reads this slot, which no other axes take over (a
        // later check of the same axes writes the same values)
        e.state = 2;
        e.pinned = true;
        return JP_OK;
    }
    SplitGeom h;
    JP_HIP(ctx, hipMemcpyAsync(&h, dev, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
    JP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (h.bad) {
        e.state = 0;
        return geom_reject(ctx, h.bad);
    }
    e.state = 1;
    bool standard = vg != nullptr;
    for (int c = 0; c < D && standard; ++c)
        for (int d = 0; d < D; ++d)
            standard = standard && h.r[c][d] == 1.0 && h.o[c][d] == (c == d ? 0.0 : 0.5);
    e.standard = standard;
    return JP_OK;
}

namespace {
int geom_reject(jp_ctx* ctx, int bits) {
    // the rejection is reported now: clear the sticky word
    int rc = zero_bytes(ctx, &dgeom(ctx)->sticky, sizeof(int32_t));
    if (rc) return rc;
    JP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const char* what = (bits & kGeomVertex) && (bits & kGeomVelocity)
                           ? "vertex and velocity"
                           : ((bits & kGeomVertex) ? "vertex" : "velocity");
    return fail(ctx, JP_ERR_ARG,
                "the FP32-split form needs uniform %s axes (deviation from x[0] + k*h above "
                "1e-9 of the spacing h); the call changed nothing",
                what);
}
}  // namespace

int geom_reject_now(jp_ctx* ctx, int bits) { return geom_reject(ctx, bits); }

// checks this call's geometry verdict now, whatever the synchronisation mode
int split_finish_sync(jp_ctx* ctx) {
    if (ctx->capturing || ctx->geo_cache[ctx->geo_cur].state == 1) return JP_OK;
    int32_t bad = 0;
    JP_HIP(ctx, hipMemcpyAsync(&bad, &geo_ptr(ctx)->bad, sizeof bad, hipMemcpyDeviceToHost,
                               ctx->stream));
    JP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return bad ? geom_reject(ctx, bad) : JP_OK;
}

int split_finish(jp_ctx* ctx) {
    if (!ctx->sync || ctx->capturing) return JP_OK;
    if (ctx->geo_cache[ctx->geo_cur].state == 1) return finish(ctx);  // accepted axes
    int32_t bad = 0;
    JP_HIP(ctx, hipMemcpyAsync(&bad, &geo_ptr(ctx)->bad, sizeof bad, hipMemcpyDeviceToHost,
                               ctx->stream));
    JP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return bad ? geom_reject(ctx, bad) : JP_OK;
}

int geom_report(jp_ctx* ctx) {
    int32_t st = 0;
    JP_HIP(ctx, hipMemcpyAsync(&st, &dgeom(ctx)->sticky, sizeof st, hipMemcpyDeviceToHost,
                               ctx->stream));
    JP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return st ? geom_reject(ctx, st) : JP_OK;
}

}  // namespace jp

namespace {

int check_common(jp_ctx* ctx, const jp_cellgrid* g, const void* const* off, const void* index) {
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if (!off) return fail(ctx, JP_ERR_ARG, "null offsets%s");
    for (int d = 0; d < g->ndim; ++d)
        if (!off[d]) return fail(ctx, JP_ERR_ARG, "null offset plane%s %lld", "", d);
    if (!index) return fail(ctx, JP_ERR_ARG, "null index%s");
    return JP_OK;
}

template <int D>
FPtr3 fptr3(float* const* u) {
    FPtr3 p{{nullptr, nullptr, nullptr}};
    for (int d = 0; d < D; ++d) p.p[d] = u[d];
    return p;
}

template <int D>
int run_adv(jp_ctx* ctx, const jp_cellgrid* gc, float* const* off, const uint8_t* index,
            const double* const* V, const jp_velgrid* vg, int method, double alpha, double dt) {
    Grid<D> G = to_grid<D>(gc);
    SGrid<D> sg = sgrid<D>(G);
    SVelIn<D> sv;
    for (int c = 0; c < D; ++c) {
        sv.v[c] = V[c];
        for (int d = 0; d < D; ++d) sv.n[c][d] = (int)vg->ax[c][d].n;
    }
    int rc = split_geometry(ctx, gc, vg);
    if (rc) return rc;
    const SplitGeom* geo = geo_ptr(ctx);
    RKf P;
    P.alpha = (float)alpha;
    P.half = alpha == 0.5;
    const double ia = 1.0 / alpha;
    P.w0 = (float)std::fma(-0.5, ia, 1.0);
    P.w1 = (float)(0.5 * ia);
    int64_t tiles = 1;
    for (int d = 0; d < D; ++d)
        tiles *= (G.ni[d] + DefaultTile<D>::n(d) - 1) / DefaultTile<D>::n(d);
    FPtr3 fp = fptr3<D>(off);
    const int span = slot_span(ctx, tiles, G.K);
    const dim3 grid((unsigned)tiles, (unsigned)((G.K + span - 1) / span));
    if (geo_standard(ctx)) {
        if (method == JP_EULER)
            k_adv_split<D, JP_EULER, true><<<grid, kBlock, 0, ctx->stream>>>(sg, sv, geo, dt, fp,
                                                                             index, P, span);
        else if (method == JP_RK2)
            k_adv_split<D, JP_RK2, true><<<grid, kBlock, 0, ctx->stream>>>(sg, sv, geo, dt, fp,
                                                                           index, P, span);
        else
            k_adv_split<D, JP_RK4, true><<<grid, kBlock, 0, ctx->stream>>>(sg, sv, geo, dt, fp,
                                                                           index, P, span);
    } else {
        if (method == JP_EULER)
            k_adv_split<D, JP_EULER, false><<<grid, kBlock, 0, ctx->stream>>>(sg, sv, geo, dt, fp,
                                                                              index, P, span);
        else if (method == JP_RK2)
            k_adv_split<D, JP_RK2, false><<<grid, kBlock, 0, ctx->stream>>>(sg, sv, geo, dt, fp,
                                                                            index, P, span);
        else
            k_adv_split<D, JP_RK4, false><<<grid, kBlock, 0, ctx->stream>>>(sg, sv, geo, dt, fp,
                                                                            index, P, span);
    }
    JP_LAUNCHED(ctx);
    return split_finish(ctx);
}

template <int D>
int run_p2g(jp_ctx* ctx, const jp_cellgrid* gc, double* F, const double* Fp, float* const* off,
            const uint8_t* index) {
    Grid<D> G = to_grid<D>(gc);
    SGrid<D> sg = sgrid<D>(G);
    int64_t NN = 1;
    for (int d = 0; d < D; ++d) NN *= G.ni[d] + 1;
    int rc;
    if ((rc = ensure_scratch(ctx, (size_t)G.C * 8 * 2 * (1 << D)))) return rc;
    double* part = static_cast<double*>(ctx->scratch);
    if ((rc = split_geometry(ctx, gc, nullptr))) return rc;
    k_p2g_split_cells<D><<<(unsigned)row_blocks<D>(G), kBlock, 0, ctx->stream>>>(
        G, sg, Fp, fptr3<D>(off), index, part);
    JP_LAUNCHED(ctx);
    k_p2g_split_nodes<D><<<(unsigned)blocks_for(NN), kBlock, 0, ctx->stream>>>(G, sg, part,
                                                                              geo_ptr(ctx), F);
    JP_LAUNCHED(ctx);
    return split_finish(ctx);
}

template <int D>
int run_g2p(jp_ctx* ctx, const jp_cellgrid* gc, double* Fp, const double* F, float* const* off,
            const uint8_t* index) {
    Grid<D> G = to_grid<D>(gc);
    dim3 grid((unsigned)row_blocks<D>(G), (unsigned)G.K);
    k_g2p_split<D><<<grid, kBlock, 0, ctx->stream>>>(G, Fp, F, fptr3<D>(off), index);
    JP_LAUNCHED(ctx);
    return finish(ctx);
}

template <int D>
int run_inject(jp_ctx* ctx, const jp_cellgrid* gc, float* const* off, uint8_t* index, int nargs,
               double* const* args, int min_xcell, uint64_t seed, uint64_t counter,
               jp_stats* stats) {
    Grid<D> G = to_grid<D>(gc);
    SGrid<D> sg = sgrid<D>(G);
    int rc;
    if ((rc = stats_begin(ctx, stats))) return rc;
    constexpr int NQ = D == 2 ? 4 : 8;
    const int min_q = (min_xcell + NQ - 1) / NQ;
    int64_t cap = 1;
    for (int d = 0; d < D; ++d) cap *= (G.ni[d] + 1) / 2;
    InjectLists L;
    if ((rc = inject_lists(ctx, G.C, NQ, cap, 0, L))) return rc;
    const bool two = D == 2 && nargs <= 1;  // the 2D kernel (k_inject2)
    const bool ordered = two && inject2_ordered(ctx);
    if (!ordered) L.state = nullptr;
    FPtr3 fp = fptr3<D>(off);
    if ((rc = split_geometry(ctx, gc, nullptr))) return rc;
    k_inj_scan_split<D><<<blocks_for(G.C), kBlock, 0, ctx->stream>>>(G, fp, index, min_q, L,
                                                                     geo_ptr(ctx));
    JP_LAUNCHED(ctx);
    if ((rc = join_segments(ctx, NQ, L))) return rc;
    ArgPtrs ap;
    for (int i = 0; i < JP_MAX_ARGS; ++i) ap.p[i] = i < nargs ? args[i] : nullptr;
    jp_stats* st = stats ? dstats(ctx) : nullptr;
    if constexpr (D == 2) {
        if (two) {
            Inject2 a{};
            a.g = G;
            a.u[0] = off[0];
            a.u[1] = off[1];
            a.h[0] = sg.h[0][0];
            a.h[1] = sg.h[1][0];
            a.index = index;
            a.arg = nargs ? args[0] : nullptr;
            a.min_q = min_q;
            a.seed = seed;
            a.counter = counter;
            a.occw = L.occ;
            a.st = st;
            if ((rc = launch_inject2<true>(ctx, a, L, cap))) return rc;
        }
    }
    for (int ph = 0; ph < NQ && !two; ++ph) {
        k_inj_wave_split<D><<<(unsigned)inject_blocks(ctx, cap), kBlock, 0, ctx->stream>>>(
            G, sg, fp, index, nargs, ap, min_q, seed, counter, L.list + (int64_t)ph * cap,
            L.total + ph, st);
        JP_LAUNCHED(ctx);
    }
    if ((rc = stats_end(ctx, stats))) return rc;
    if (ordered && stats && (rc = inject_wait_report(ctx))) return rc;
    return split_finish(ctx);
}

// ------------------------------------------------------------ ensemble host
// Variants: tile shape (64x4, 32x8, 16x16, 128x2) x persistent blocks per CU (2, 8).
constexpr int kEnsShapes = 4;
constexpr int kEnsVariants = 8;

int ens_tx(int s) { return s == 0 ? 64 : (s == 1 ? 32 : (s == 2 ? 16 : 128)); }
int ens_ty(int s) { return kBlock / ens_tx(s); }

template <int M>
void ens_launch(hipStream_t st, int shape, unsigned blocks, const EnsGrid* grids,
                const uint16_t* tg, const int32_t* t0, int total, const RKf& P, int zero_dt) {
    switch (shape) {
    case 0:
        k_adv_split_ens<M, Tile<2, 64, 4>><<<blocks, kBlock, 0, st>>>(grids, tg, t0, total, P, zero_dt);
        break;
    case 1:
        k_adv_split_ens<M, Tile<2, 32, 8>><<<blocks, kBlock, 0, st>>>(grids, tg, t0, total, P, zero_dt);
        break;
    case 2:
        k_adv_split_ens<M, Tile<2, 16, 16>><<<blocks, kBlock, 0, st>>>(grids, tg, t0, total, P, zero_dt);
        break;
    default:
        k_adv_split_ens<M, Tile<2, 128, 2>><<<blocks, kBlock, 0, st>>>(grids, tg, t0, total, P, zero_dt);
        break;
    }
}

}  // namespace

struct jp_ensemble {
    int n = 0;
    int method = JP_RK2;
    RKf P{};
    EnsGrid* grids = nullptr;             // device
    uint16_t* tile_grid[kEnsShapes] = {};  // device, per tile shape
    int32_t* tile0[kEnsShapes] = {};       // device, per tile shape
    int total[kEnsShapes] = {};
    int variant = 1;  // 64x4 tiles, 8 blocks per CU (variant = 2*shape + (8 blocks per CU))
};

namespace {

int ens_run(jp_ctx* ctx, jp_ensemble* e, int variant, int zero_dt) {
    const int shape = variant >> 1;
    const int bpc = (variant & 1) ? 8 : 2;
    const int total = e->total[shape];
    if (total == 0) return JP_OK;
    const unsigned blocks = (unsigned)std::min<int64_t>(total, (int64_t)ctx->num_cu * bpc);
    if (e->method == JP_EULER)
        ens_launch<JP_EULER>(ctx->stream, shape, blocks, e->grids, e->tile_grid[shape],
                             e->tile0[shape], total, e->P, zero_dt);
    else if (e->method == JP_RK2)
        ens_launch<JP_RK2>(ctx->stream, shape, blocks, e->grids, e->tile_grid[shape],
                           e->tile0[shape], total, e->P, zero_dt);
    else
        ens_launch<JP_RK4>(ctx->stream, shape, blocks, e->grids, e->tile_grid[shape],
                           e->tile0[shape], total, e->P, zero_dt);
    JP_LAUNCHED(ctx);
    return JP_OK;
}

void ens_free(jp_ensemble* e) {
    if (!e) return;
    (void)hipFree(e->grids);
    for (int s = 0; s < kEnsShapes; ++s) {
        (void)hipFree(e->tile_grid[s]);
        (void)hipFree(e->tile0[s]);
    }
    delete e;
}

}  // namespace

extern "C" {

int jp_split_from_f64(jp_ctx* ctx, const jp_cellgrid* g, const double* const* coords,
                      const uint8_t* index, float* const* off) {
    int rc = check_common(ctx, g, (const void* const*)off, index);
    if (rc) return rc;
    if (!coords) return fail(ctx, JP_ERR_ARG, "null coords%s");
    // the split form is defined on uniform vertex axes only: reject others here
    if ((rc = split_geometry(ctx, g, nullptr))) return rc;
    if ((rc = split_finish_sync(ctx))) return rc;
    if (g->ndim == 2) {
        Grid<2> G = to_grid<2>(g);
        CPtr3 x{{coords[0], coords[1], nullptr}};
        k_to_split<2><<<dim3((unsigned)row_blocks<2>(G), (unsigned)G.K), kBlock, 0, ctx->stream>>>(
            G, x, index, fptr3<2>(off));
    } else {
        Grid<3> G = to_grid<3>(g);
        CPtr3 x{{coords[0], coords[1], coords[2]}};
        k_to_split<3><<<dim3((unsigned)row_blocks<3>(G), (unsigned)G.K), kBlock, 0, ctx->stream>>>(
            G, x, index, fptr3<3>(off));
    }
    JP_LAUNCHED(ctx);
    return finish(ctx);
}

int jp_split_to_f64(jp_ctx* ctx, const jp_cellgrid* g, float* const* off, const uint8_t* index,
                    double* const* coords) {
    int rc = check_common(ctx, g, (const void* const*)off, index);
    if (rc) return rc;
    if (!coords) return fail(ctx, JP_ERR_ARG, "null coords%s");
    if (g->ndim == 2) {
        Grid<2> G = to_grid<2>(g);
        Ptr3 x{{coords[0], coords[1], nullptr}};
        k_from_split<2><<<dim3((unsigned)row_blocks<2>(G), (unsigned)G.K), kBlock, 0,
                          ctx->stream>>>(G, fptr3<2>(off), index, x);
    } else {
        Grid<3> G = to_grid<3>(g);
        Ptr3 x{{coords[0], coords[1], coords[2]}};
        k_from_split<3><<<dim3((unsigned)row_blocks<3>(G), (unsigned)G.K), kBlock, 0,
                          ctx->stream>>>(G, fptr3<3>(off), index, x);
    }
    JP_LAUNCHED(ctx);
    return finish(ctx);
}

int jp_advection_split(jp_ctx* ctx, const jp_cellgrid* g, float* const* off, const uint8_t* index,
                       const double* const* V, const jp_velgrid* vg, int method, double alpha,
                       double dt) {
    int rc = check_common(ctx, g, (const void* const*)off, index);
    if (rc) return rc;
    if (!V || !vg) return fail(ctx, JP_ERR_ARG, "null velocity%s");
    for (int c = 0; c < g->ndim; ++c)
        if (!V[c]) return fail(ctx, JP_ERR_ARG, "null velocity component%s %lld", "", c);
    if (method < JP_EULER || method > JP_RK4)
        return fail(ctx, JP_ERR_ARG, "unknown integrator%s %lld", "", method);
    if (method == JP_RK2 && !(0.0 < alpha && alpha < 1.0))
        return fail(ctx, JP_ERR_ARG, "Only 0 < α < 1 is supported%s");
    return g->ndim == 2 ? run_adv<2>(ctx, g, off, index, V, vg, method, alpha, dt)
                        : run_adv<3>(ctx, g, off, index, V, vg, method, alpha, dt);
}

int jp_particle2grid_split(jp_ctx* ctx, const jp_cellgrid* g, double* F, const double* Fp,
                           float* const* off, const uint8_t* index) {
    int rc = check_common(ctx, g, (const void* const*)off, index);
    if (rc) return rc;
    if (!F || !Fp) return fail(ctx, JP_ERR_ARG, "null field%s");
    return g->ndim == 2 ? run_p2g<2>(ctx, g, F, Fp, off, index)
                        : run_p2g<3>(ctx, g, F, Fp, off, index);
}

int jp_grid2particle_split(jp_ctx* ctx, const jp_cellgrid* g, double* Fp, const double* F,
                           float* const* off, const uint8_t* index) {
    int rc = check_common(ctx, g, (const void* const*)off, index);
    if (rc) return rc;
    if (!F || !Fp) return fail(ctx, JP_ERR_ARG, "null field%s");
    return g->ndim == 2 ? run_g2p<2>(ctx, g, Fp, F, off, index)
                        : run_g2p<3>(ctx, g, Fp, F, off, index);
}

int jp_inject_particles_split(jp_ctx* ctx, const jp_cellgrid* g, float* const* off,
                              uint8_t* index, int nargs, double* const* args, int min_xcell,
                              uint64_t seed, uint64_t call_counter, jp_stats* stats) {
    int rc = check_common(ctx, g, (const void* const*)off, index);
    if (rc) return rc;
    if (nargs < 0 || nargs > JP_MAX_ARGS) return fail(ctx, JP_ERR_ARG, "nargs out of range%s");
    for (int a = 0; a < nargs; ++a)
        if (!args || !args[a]) return fail(ctx, JP_ERR_ARG, "null args pointer%s %lld", "", a);
    if (min_xcell < 0) return fail(ctx, JP_ERR_ARG, "min_xcell must be non-negative%s");
    return g->ndim == 2
               ? run_inject<2>(ctx, g, off, index, nargs, args, min_xcell, seed, call_counter, stats)
               : run_inject<3>(ctx, g, off, index, nargs, args, min_xcell, seed, call_counter, stats);
}


int jp_ensemble_create(jp_ctx* ctx, int n, const jp_cellgrid* grids, const jp_velgrid* vgs,
                       float* const* off, uint8_t* const* index, const double* const* V,
                       const double* dt, int method, double alpha, jp_ensemble** out) {
    if (!ctx) return JP_ERR_STATE;
    if (!out || !grids || !vgs || !off || !index || !V || !dt)
        return fail(ctx, JP_ERR_ARG, "null ensemble argument%s");
    *out = nullptr;
    if (n < 1 || n > 65535) return fail(ctx, JP_ERR_ARG, "ensemble size%s %lld out of range", "", n);
    if (method < JP_EULER || method > JP_RK4)
        return fail(ctx, JP_ERR_ARG, "unknown integrator%s %lld", "", method);
    if (method == JP_RK2 && !(0.0 < alpha && alpha < 1.0))
        return fail(ctx, JP_ERR_ARG, "Only 0 < α < 1 is supported%s");
    if (ctx->capturing) return fail(ctx, JP_ERR_STATE, "ensemble created during capture%s");
    std::vector<EnsGrid> host(n);
    for (int i = 0; i < n; ++i) {
        const jp_cellgrid* g = grids + i;
        if (g->ndim != 2) return fail(ctx, JP_ERR_ARG, "ensemble grid%s %lld is not 2D", "", i);
        float* const o2[2] = {off[2 * i], off[2 * i + 1]};
        int rc = check_common(ctx, g, (const void* const*)o2, index[i]);
        if (rc) return rc;
        if (!V[2 * i] || !V[2 * i + 1]) return fail(ctx, JP_ERR_ARG, "null velocity of grid%s %lld", "", i);
        Grid<2> G = to_grid<2>(g);
        EnsGrid& E = host[i];
        E.g = sgrid<2>(G);
        for (int c = 0; c < 2; ++c) {
            E.V.v[c] = V[2 * i + c];
            for (int d = 0; d < 2; ++d) E.V.n[c][d] = (int)vgs[i].ax[c][d].n;
        }
        E.dt = dt[i];
        E.off = FPtr3{{o2[0], o2[1], nullptr}};
        E.index = index[i];
        // axis maps of this grid, checked now (the ensemble keeps its arrays)
        if ((rc = split_geometry(ctx, g, vgs + i))) return rc;
        SplitGeom geo;
        JP_HIP(ctx, hipMemcpyAsync(&geo, geo_ptr(ctx), sizeof geo, hipMemcpyDeviceToHost,
                                   ctx->stream));
        JP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (geo.bad) return geom_reject_now(ctx, geo.bad);
        for (int c = 0; c < 2; ++c)
            for (int d = 0; d < 2; ++d) {
                E.m.r[c][d] = geo.r[c][d];
                E.m.o[c][d] = geo.o[c][d];
            }
    }
    jp_ensemble* e = new jp_ensemble();
    e->n = n;
    e->method = method;
    e->P.alpha = (float)alpha;
    e->P.half = alpha == 0.5;
    const double ia = 1.0 / alpha;
    e->P.w0 = (float)std::fma(-0.5, ia, 1.0);
    e->P.w1 = (float)(0.5 * ia);
    hipError_t err = hipMalloc(&e->grids, sizeof(EnsGrid) * n);
    if (err == hipSuccess)
        err = hipMemcpy(e->grids, host.data(), sizeof(EnsGrid) * n, hipMemcpyHostToDevice);
    for (int s = 0; s < kEnsShapes && err == hipSuccess; ++s) {
        std::vector<int32_t> t0(n);
        std::vector<uint16_t> tg;
        int64_t tot = 0;
        for (int i = 0; i < n; ++i) {
            t0[i] = (int32_t)tot;
            const int64_t tiles = ((host[i].g.ni[0] + ens_tx(s) - 1) / ens_tx(s)) *
                                  (int64_t)((host[i].g.ni[1] + ens_ty(s) - 1) / ens_ty(s));
            for (int64_t t = 0; t < tiles; ++t) tg.push_back((uint16_t)i);
            tot += tiles;
        }
        if (tot >= ((int64_t)1 << 31)) {
            ens_free(e);
            return fail(ctx, JP_ERR_ARG, "ensemble too large%s");
        }
        e->total[s] = (int)tot;
        err = hipMalloc(&e->tile_grid[s], sizeof(uint16_t) * std::max<int64_t>(tot, 1));
        if (err == hipSuccess) err = hipMalloc(&e->tile0[s], sizeof(int32_t) * n);
        if (err == hipSuccess && tot)
            err = hipMemcpy(e->tile_grid[s], tg.data(), sizeof(uint16_t) * tot, hipMemcpyHostToDevice);
        if (err == hipSuccess)
            err = hipMemcpy(e->tile0[s], t0.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice);
    }
    if (err != hipSuccess) {
        ens_free(e);
        return fail(ctx, JP_ERR_HIP, "ensemble allocation: %s", hipGetErrorString(err));
    }
    *out = e;
    return JP_OK;
}

int jp_ensemble_advect(jp_ctx* ctx, jp_ensemble* e) {
    if (!ctx) return JP_ERR_STATE;
    if (!e) return fail(ctx, JP_ERR_ARG, "null ensemble%s");
    int rc = ens_run(ctx, e, e->variant, 0);
    if (rc) return rc;
    return finish(ctx);
}

int jp_ensemble_set_variant(jp_ctx* ctx, jp_ensemble* e, int variant) {
    if (!ctx) return JP_ERR_STATE;
    if (!e) return fail(ctx, JP_ERR_ARG, "null ensemble%s");
    if (variant < 0 || variant >= kEnsVariants)
        return fail(ctx, JP_ERR_ARG, "variant%s %lld out of range", "", variant);
    e->variant = variant;
    return JP_OK;
}

int jp_ensemble_tune(jp_ctx* ctx, jp_ensemble* e, int reps, int* best, float* ms) {
    if (!ctx) return JP_ERR_STATE;
    if (!e || reps < 1) return fail(ctx, JP_ERR_ARG, "bad ensemble tuning arguments%s");
    if (ctx->capturing) return fail(ctx, JP_ERR_STATE, "tuning during capture%s");
    hipEvent_t a, b;
    JP_HIP(ctx, hipEventCreate(&a));
    JP_HIP(ctx, hipEventCreate(&b));
    int bv = 0;
    float bt = 0.0f;
    int rc = JP_OK;
    for (int v = 0; v < kEnsVariants && !rc; ++v) {
        // one untimed launch, then reps timed ones; dt = 0 leaves positions unchanged
        rc = ens_run(ctx, e, v, 1);
        if (!rc && hipEventRecord(a, ctx->stream) != hipSuccess) rc = fail(ctx, JP_ERR_HIP, "event%s");
        for (int r = 0; r < reps && !rc; ++r) rc = ens_run(ctx, e, v, 1);
        if (!rc && hipEventRecord(b, ctx->stream) != hipSuccess) rc = fail(ctx, JP_ERR_HIP, "event%s");
        if (!rc && hipEventSynchronize(b) != hipSuccess) rc = fail(ctx, JP_ERR_HIP, "event sync%s");
        float t = 0.0f;
        if (!rc && hipEventElapsedTime(&t, a, b) != hipSuccess) rc = fail(ctx, JP_ERR_HIP, "elapsed%s");
        t /= (float)reps;
        if (ms) ms[v] = t;
        if (!rc && (v == 0 || t < bt)) {
            bt = t;
            bv = v;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc) return rc;
    e->variant = bv;
    if (best) *best = bv;
    return JP_OK;
}

int jp_ensemble_destroy(jp_ctx* ctx, jp_ensemble* e) {
    if (!ctx) return JP_ERR_STATE;
    if (ctx->capturing) return fail(ctx, JP_ERR_STATE, "ensemble destroyed during capture%s");
    (void)hipStreamSynchronize(ctx->stream);
    ens_free(e);
    return JP_OK;
}
}  // extern "C"
