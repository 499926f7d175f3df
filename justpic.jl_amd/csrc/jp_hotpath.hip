// jp_hotpath.hip -- FP64 drop-in kernels for JustPIC's particle hot path on
// gfx950: init, advection (Euler/RK2/RK4), move, particle2grid, grid2particle
// and injection, 2D and 3D, on the B = 0 cell-plane layout data[c + C*k].
#include "jp_ctx.hpp"

#include <cmath>

using namespace jp;

namespace {

template <int D>
struct VPtrs {
    const double* v[D];
};

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

// ============================================================ advection
// interp_velocity2particle (src/Particles/Advection/advection.jl:93-148) for
// all components: +Inf outside the component's grid (check_local_limits,
// src/Advection/advection.jl:83-91), else bisection seeded at the particle's
// cell, t = (p - x_k)*inv(dx_k), n-linear lerp of the corner values.
template <int D>
__device__ inline void interp_vel(const Vel<D>& vg, const VPtrs<D>& V, const double (&p)[D],
                                  const int64_t (&I)[D], double (&v)[D]) {
#pragma unroll
    for (int c = 0; c < D; ++c) {
        const Axis* ax = vg.ax[c];
        bool in = true;
#pragma unroll
        for (int d = 0; d < D; ++d) in = in && (ax[d].x[0] <= p[d] && p[d] <= ax[d].x[ax[d].n - 1]);
        if (!in) {
            v[c] = __builtin_inf();
            continue;
        }
        int64_t k[D];
        double t[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            k[d] = find_cell(p[d], ax[d], I[d]);
            t[d] = (p[d] - ax[d].x[k[d]]) * ax[d].idx[k[d]];
        }
        const double* __restrict__ F = V.v[c];
        const int64_t n0 = ax[0].n;
        if constexpr (D == 2) {
            int64_t b = k[0] + n0 * k[1];
            double F00 = F[b], F10 = F[b + 1], F01 = F[b + n0], F11 = F[b + n0 + 1];
            v[c] = lerp1(t[1], lerp1(t[0], F00, F10), lerp1(t[0], F01, F11));
        } else {
            const int64_t n1 = ax[1].n;
            int64_t b = k[0] + n0 * (k[1] + n1 * k[2]);
            int64_t s = n0 * n1;
            double F000 = F[b], F100 = F[b + 1], F010 = F[b + n0], F110 = F[b + n0 + 1];
            double F001 = F[b + s], F101 = F[b + s + 1], F011 = F[b + s + n0],
                   F111 = F[b + s + n0 + 1];
            double lo = lerp1(t[1], lerp1(t[0], F000, F100), lerp1(t[0], F010, F110));
            double hi = lerp1(t[1], lerp1(t[0], F001, F101), lerp1(t[0], F011, F111));
            v[c] = lerp1(t[2], lo, hi);
        }
    }
}

struct RKParams {
    double dt;
    double adt;  // (s*alpha)*dt
    int half;    // alpha == 0.5
    double w0;   // muladd(-0.5, inv(alpha), 1)
    double w1;   // 0.5*inv(alpha)
};

// advect_particle: Euler (Euler.jl), RK2 (RK2.jl), RK4 (RK4.jl)
template <int D, int M>
__device__ inline void advect_one(const Vel<D>& vg, const VPtrs<D>& V, const RKParams& P,
                                  const double (&p0)[D], const int64_t (&I)[D], double (&out)[D]) {
    double v0[D];
    interp_vel<D>(vg, V, p0, I, v0);
    if constexpr (M == JP_EULER) {
#pragma unroll
        for (int d = 0; d < D; ++d) out[d] = fma(P.dt, v0[d], p0[d]);
    } else if constexpr (M == JP_RK2) {
        double p1[D], v1[D];
#pragma unroll
        for (int d = 0; d < D; ++d) p1[d] = fma(P.adt, v0[d], p0[d]);
        interp_vel<D>(vg, V, p1, I, v1);
        if (P.half) {
#pragma unroll
            for (int d = 0; d < D; ++d) out[d] = fma(P.dt, v1[d], p0[d]);
        } else {
#pragma unroll
            for (int d = 0; d < D; ++d) out[d] = fma(P.dt, fma(P.w1, v1[d], P.w0 * v0[d]), p0[d]);
        }
    } else {
        double q[D], k2[D], k3[D], k4[D];
#pragma unroll
        for (int d = 0; d < D; ++d) q[d] = p0[d] + (P.dt * v0[d]) / 2.0;
        interp_vel<D>(vg, V, q, I, k2);
#pragma unroll
        for (int d = 0; d < D; ++d) q[d] = p0[d] + (P.dt * k2[d]) / 2.0;
        interp_vel<D>(vg, V, q, I, k3);
#pragma unroll
        for (int d = 0; d < D; ++d) q[d] = p0[d] + P.dt * k3[d];
        interp_vel<D>(vg, V, q, I, k4);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            double sum = ((v0[d] + 2.0 * k2[d]) + 2.0 * k3[d]) + k4[d];
            out[d] = p0[d] + (P.dt * sum) / 6.0;
        }
    }
}

// advection_kernel! (advection.jl:66-91): one thread per slot; blockIdx.y is
// the slot plane, so a wavefront reads 64 consecutive cells of one plane.
template <int D, int M>
__global__ void __launch_bounds__(kBlock) k_advect(Grid<D> g, Vel<D> vg, VPtrs<D> V, Ptr3 coords,
                                                   const uint8_t* __restrict__ index, RKParams P) {
    int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= g.C) return;
    int64_t e = c + g.C * (int64_t)blockIdx.y;
    if (!index[e]) return;
    int64_t I[D];
    unlin<D>(g.ni, c, I);
    double p0[D], p[D];
#pragma unroll
    for (int d = 0; d < D; ++d) p0[d] = coords.p[d][e];
    advect_one<D, M>(vg, V, P, p0, I, p);
#pragma unroll
    for (int d = 0; d < D; ++d) coords.p[d][e] = p[d];
}

// ============================================================ move
template <int D>
__device__ inline bool in_cell(const Grid<D>& g, const int64_t (&I)[D], const double (&p)[D]) {
    bool b = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        double x = g.xv[d].x[I[d]], h = g.xv[d].dx[I[d]];
        b = b & (x < p[d] && p[d] < x + h);  // isincell, src/Particles/utils.jl:15
    }
    return b;
}

template <int D>
__device__ inline bool in_domain(const Grid<D>& g, const double (&p)[D]) {
#pragma unroll
    for (int d = 0; d < D; ++d)  // indomain, src/Particles/move_safe.jl:199-206
        if (!(g.xv[d].x[0] < p[d] && p[d] < g.xv[d].x[g.xv[d].n - 1])) return false;
    return true;
}

// Detects movers whose destination is not a neighbour (CFL > 1).  Their
// presence makes the colour phases' neighbourhoods overlap, so the phases
// then run their cells one after another in a single thread.
template <int D>
__global__ void __launch_bounds__(kBlock) k_move_far(Grid<D> g, CPtr3 coords,
                                                     const uint8_t* __restrict__ index, int* flag) {
    int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= g.C) return;
    int64_t e = c + g.C * (int64_t)blockIdx.y;
    if (!index[e]) return;
    int64_t I[D];
    unlin<D>(g.ni, c, I);
    double p[D];
#pragma unroll
    for (int d = 0; d < D; ++d) p[d] = coords.p[d][e];
    if (in_cell<D>(g, I, p) || !in_domain<D>(g, p)) return;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        int64_t nc = find_cell(p[d], g.xv[d], I[d]);
        if (nc - I[d] > 1 || I[d] - nc > 1) {
            atomicOr(flag, 1);
            return;
        }
    }
}

// move_kernel! (src/Particles/move_safe.jl:72-125) for one source cell.
template <int D>
__device__ void move_cell(const Grid<D>& g, const Ptr3& coords, uint8_t* __restrict__ index,
                          int nargs, const ArgPtrs& args, const int64_t (&I)[D], int64_t (&cnt)[4]) {
    const int64_t C = g.C;
    const int K = g.K;
    const int64_t c = lin<D>(g.ni, I);
    int sp = 0;  // starting_point, 0-based
    for (int ip = 0; ip < K; ++ip) {
        const int64_t e = c + C * ip;
        if (!index[e]) continue;
        double p[D];
#pragma unroll
        for (int d = 0; d < D; ++d) p[d] = coords.p[d][e];
        if (in_cell<D>(g, I, p)) continue;
        if (!in_domain<D>(g, p)) {
            index[e] = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) coords.p[d][e] = dnan();
            for (int a = 0; a < nargs; ++a) args.p[a][e] = dnan();
            cnt[1]++;
            continue;
        }
        int64_t nc[D];
        bool far = false;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            nc[d] = find_cell(p[d], g.xv[d], I[d]);
            far = far || (nc[d] - I[d] > 1 || I[d] - nc[d] > 1);
        }
        if (far) cnt[3]++;
        double cached[JP_MAX_ARGS];
        for (int a = 0; a < nargs; ++a) cached[a] = args.p[a][e];
        index[e] = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) coords.p[d][e] = dnan();
        for (int a = 0; a < nargs; ++a) args.p[a][e] = dnan();
        const int64_t cn = lin<D>(g.ni, nc);
        int fi = -1;
        for (int i = sp; i < K; ++i)
            if (!index[cn + C * i]) {
                fi = i;
                break;
            }
        if (fi < 0) {
            cnt[2]++;
            continue;
        }
        sp = fi;
        const int64_t f = cn + C * fi;
        index[f] = 1;
#pragma unroll
        for (int d = 0; d < D; ++d) coords.p[d][f] = p[d];
        for (int a = 0; a < nargs; ++a) args.p[a][f] = cached[a];
        cnt[0]++;
    }
}

struct Off3 {
    int o[3];
};

// move_particles_ps! (src/Particles/move_safe.jl:51-62): one thread per cell
// of the colour (cells 3*T + offset).
template <int D>
__global__ void __launch_bounds__(kBlock) k_move_phase(Grid<D> g, Ptr3 coords, uint8_t* index,
                                                       int nargs, ArgPtrs args, Off3 off,
                                                       const int* __restrict__ flag, jp_stats* st) {
    int64_t ncol[D];
    int64_t total = 1;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        ncol[d] = (g.ni[d] + 2) / 3;
        total *= ncol[d];
    }
    int64_t cnt[4] = {0, 0, 0, 0};
    int64_t t0, t1;
    if (*flag) {
        if (blockIdx.x != 0 || threadIdx.x != 0) return;
        t0 = 0;
        t1 = total;
    } else {
        t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
        if (t0 >= total) return;
        t1 = t0 + 1;
    }
    for (int64_t t = t0; t < t1; ++t) {
        int64_t r = t, I[D];
        bool ok = true;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            I[d] = 3 * (r % ncol[d]) + off.o[d];
            r /= ncol[d];
            ok = ok && I[d] < g.ni[d];
        }
        if (ok) move_cell<D>(g, coords, index, nargs, args, I, cnt);
    }
    stat_add(st, ST_MOVED, cnt[0]);
    stat_add(st, ST_DELETED, cnt[1]);
    stat_add(st, ST_DROPPED, cnt[2]);
    stat_add(st, ST_FAR, cnt[3]);
}

// ============================================================ particle2grid
// _particle2grid! (src/Interpolations/particle_to_grid.jl:37-199): one thread
// per node, exact reference accumulation order (cells with the last dim
// outermost, slots ascending).  mode 0: wF/w (2D scalar), 1: wF*inv(w).
template <int D, int NF>
__global__ void __launch_bounds__(kBlock) k_p2g(Grid<D> g, int nf_rt, ArgPtrs F, ArgPtrs Fp,
                                                CPtr3 coords, const uint8_t* __restrict__ index,
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
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        int64_t I[D];
        bool ok = true;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            I[d] = N[d] + (((q >> d) & 1) ? 0 : -1);
            ok = ok && (I[d] >= 0 && I[d] < g.ni[d]);
        }
        if (!ok) continue;
        const int64_t c = lin<D>(g.ni, I);
        for (int ip = 0; ip < g.K; ++ip) {
            const int64_t e = c + C * ip;
            if (!index[e]) continue;
            double s;
            {
                double a = xn[0] - coords.p[0][e];
                s = a * a;
#pragma unroll
                for (int d = 1; d < D; ++d) {
                    double b = xn[d] - coords.p[d][e];
                    s = s + b * b;
                }
            }
            double dist = sqrt(s);
            double wi = 1.0 / (dist * dist);
            w += wi;
            for (int f = 0; f < nf; ++f) wF[f] = fma(wi, Fp.p[f][e], wF[f]);
        }
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
    int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= g.C) return;
    int64_t e = c + g.C * (int64_t)blockIdx.y;
    if (!index[e]) return;
    int64_t I[D];
    unlin<D>(g.ni, c, I);
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
// index_min_distance (src/Particles/injection.jl:330-393)
template <int D>
__device__ int64_t nearest(const Grid<D>& g, const Ptr3& coords, const uint8_t* index,
                           const double (&pn)[D], int cur, const int64_t (&Ic)[D]) {
    int64_t best = -1;
    double dmin = __builtin_inf();
    constexpr int NB = D == 2 ? 9 : 27;
    for (int q = 0; q < NB; ++q) {
        int64_t I[D];
        int r = q;
        bool ok = true, same = true;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            I[d] = Ic[d] + (r % 3) - 1;
            r /= 3;
            ok = ok && (I[d] >= 0 && I[d] < g.ni[d]);
            same = same && (I[d] == Ic[d]);
        }
        if (!ok) continue;
        const int64_t c = lin<D>(g.ni, I);
        for (int ip = 0; ip < g.K; ++ip) {
            if (same && ip == cur) continue;
            const int64_t e = c + g.C * ip;
            if (!index[e]) continue;
            double p[D];
            bool hasnan = false;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                p[d] = coords.p[d][e];
                hasnan = hasnan || (p[d] != p[d]);
            }
            if (D == 2 && hasnan) continue;
            double a = p[0] - pn[0];
            double s = a * a;
#pragma unroll
            for (int d = 1; d < D; ++d) {
                double b = p[d] - pn[d];
                s = s + b * b;
            }
            double dd = sqrt(s);
            if (dd < dmin) {
                best = e;
                dmin = dd;
            }
        }
    }
    return best;
}

// _inject_particles! (src/Particles/injection.jl:68-131) for one cell
template <int D>
__device__ void inject_cell(const Grid<D>& g, const Ptr3& coords, uint8_t* index, int nargs,
                            const ArgPtrs& args, int min_q, uint64_t seed, uint64_t counter,
                            const int64_t (&I)[D], int64_t (&cnt)[2]) {
    const int64_t c = lin<D>(g.ni, I);
    constexpr int NQ = D == 2 ? 4 : 8;
    double xv[D], dq[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        xv[d] = g.xv[d].x[I[d]];
        dq[d] = g.xv[d].dx[I[d]] / 2.0;
    }
    for (int q = 0; q < NQ; ++q) {
        double vert[D];
#pragma unroll
        for (int d = 0; d < D; ++d) vert[d] = xv[d] + dq[d] * (double)((q >> d) & 1);
        int num = 0;
        for (int ip = 0; ip < g.K; ++ip) {
            const int64_t e = c + g.C * ip;
            if (!index[e]) continue;
            bool inside = true;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                double p = coords.p[d][e];
                inside = inside & (vert[d] < p && p < vert[d] + dq[d]);
            }
            num += inside ? 1 : 0;
        }
        if (num >= min_q) break;
        for (int ip = 0; ip < g.K; ++ip) {
            const int64_t e = c + g.C * ip;
            if (index[e]) continue;
            ++num;
            double pn[D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                double r = uniform53(seed, (uint64_t)c, ip, d, 2, counter);
                pn[d] = vert[d] + dq[d] * fma(0.95, r, 0.05);
            }
#pragma unroll
            for (int d = 0; d < D; ++d) coords.p[d][e] = pn[d];
            index[e] = 1;
            int64_t src = nearest<D>(g, coords, index, pn, ip, I);
            if (src >= 0) {
                for (int a = 0; a < nargs; ++a) args.p[a][e] = args.p[a][src];
            } else {
                cnt[1]++;
            }
            cnt[0]++;
            if (num >= min_q) break;
        }
    }
}

template <int D>
__global__ void __launch_bounds__(kBlock) k_inject_phase(Grid<D> g, Ptr3 coords, uint8_t* index,
                                                         int nargs, ArgPtrs args, int min_q,
                                                         uint64_t seed, uint64_t counter, Off3 off,
                                                         jp_stats* st) {
    int64_t ncol[D];
    int64_t total = 1;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        ncol[d] = (g.ni[d] + 1) / 2;
        total *= ncol[d];
    }
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= total) return;
    int64_t I[D];
    bool ok = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        I[d] = 2 * (t % ncol[d]) + off.o[d];
        t /= ncol[d];
        ok = ok && I[d] < g.ni[d];
    }
    if (!ok) return;
    int64_t cnt[2] = {0, 0};
    inject_cell<D>(g, coords, index, nargs, args, min_q, seed, counter, I, cnt);
    stat_add(st, ST_INJECTED, cnt[0]);
    stat_add(st, ST_ORPHAN, cnt[1]);
}

// ============================================================ host side
template <int D>
Vel<D> to_vel(const jp_velgrid* vg) {
    Vel<D> V;
    for (int c = 0; c < D; ++c)
        for (int d = 0; d < D; ++d) {
            const jp_axis& a = vg->ax[c][d];
            V.ax[c][d] = Axis{a.x, a.dx, a.inv_dx, a.n};
        }
    return V;
}

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

template <int D, int M>
int launch_advect(jp_ctx* ctx, const Grid<D>& G, const Vel<D>& vg, const VPtrs<D>& V, Ptr3 cp,
                  const uint8_t* index, const RKParams& P) {
    dim3 grid((unsigned)blocks_for(G.C), (unsigned)G.K);
    k_advect<D, M><<<grid, kBlock, 0, ctx->stream>>>(G, vg, V, cp, index, P);
    JP_LAUNCHED(ctx);
    return JP_OK;
}

template <int D>
int run_advect(jp_ctx* ctx, const jp_cellgrid* g, double* const* coords, const uint8_t* index,
               const double* const* Vp, const jp_velgrid* vgp, int method, const RKParams& P) {
    Grid<D> G = to_grid<D>(g);
    Vel<D> vg = to_vel<D>(vgp);
    VPtrs<D> V;
    for (int c = 0; c < D; ++c) V.v[c] = Vp[c];
    Ptr3 cp = ptr3(coords, D);
    int rc;
    if (method == JP_EULER)
        rc = launch_advect<D, JP_EULER>(ctx, G, vg, V, cp, index, P);
    else if (method == JP_RK2)
        rc = launch_advect<D, JP_RK2>(ctx, G, vg, V, cp, index, P);
    else
        rc = launch_advect<D, JP_RK4>(ctx, G, vg, V, cp, index, P);
    if (rc) return rc;
    return finish(ctx);
}

template <int D>
int run_move(jp_ctx* ctx, const jp_cellgrid* g, double* const* coords, uint8_t* index, int nargs,
             double* const* args, jp_stats* stats) {
    Grid<D> G = to_grid<D>(g);
    int* flag = dflags(ctx);
    int rc = zero_bytes(ctx, flag, sizeof(int));
    if (rc) return rc;
    if ((rc = stats_begin(ctx, stats))) return rc;
    dim3 grid((unsigned)blocks_for(G.C), (unsigned)G.K);
    CPtr3 cc = cptr3(coords, D);
    k_move_far<D><<<grid, kBlock, 0, ctx->stream>>>(G, cc, index, flag);
    JP_LAUNCHED(ctx);
    int64_t total = 1;
    for (int d = 0; d < D; ++d) total *= (G.ni[d] + 2) / 3;
    const int nphase = D == 2 ? 9 : 27;
    Ptr3 cp = ptr3(coords, D);
    ArgPtrs ap = argptrs(args, nargs);
    jp_stats* st = stats ? dstats(ctx) : nullptr;
    for (int ph = 0; ph < nphase; ++ph) {
        k_move_phase<D><<<blocks_for(total), kBlock, 0, ctx->stream>>>(
            G, cp, index, nargs, ap, phase_offsets(D, ph, 3), flag, st);
        JP_LAUNCHED(ctx);
    }
    if ((rc = stats_end(ctx, stats))) return rc;
    return finish(ctx);
}

template <int D>
int run_p2g(jp_ctx* ctx, const jp_cellgrid* g, int nf, double* const* F, const double* const* Fp,
            const double* const* coords, const uint8_t* index, int mode) {
    Grid<D> G = to_grid<D>(g);
    int64_t NN = 1;
    for (int d = 0; d < D; ++d) NN *= G.ni[d] + 1;
    ArgPtrs Fa = argptrs(F, nf);
    ArgPtrs Fpa = argptrs(const_cast<double* const*>(Fp), nf);
    if (nf == 1)
        k_p2g<D, 1><<<blocks_for(NN), kBlock, 0, ctx->stream>>>(G, nf, Fa, Fpa, cptr3(coords, D),
                                                                index, mode);
    else
        k_p2g<D, 0><<<blocks_for(NN), kBlock, 0, ctx->stream>>>(G, nf, Fa, Fpa, cptr3(coords, D),
                                                                index, mode);
    JP_LAUNCHED(ctx);
    return finish(ctx);
}

template <int D>
int run_g2p(jp_ctx* ctx, const jp_cellgrid* g, double* Fp, const double* F,
            const double* const* coords, const uint8_t* index) {
    Grid<D> G = to_grid<D>(g);
    dim3 grid((unsigned)blocks_for(G.C), (unsigned)G.K);
    k_g2p<D><<<grid, kBlock, 0, ctx->stream>>>(G, Fp, F, cptr3(coords, D), index);
    JP_LAUNCHED(ctx);
    return finish(ctx);
}

template <int D>
int run_inject(jp_ctx* ctx, const jp_cellgrid* g, double* const* coords, uint8_t* index, int nargs,
               double* const* args, int min_xcell, uint64_t seed, uint64_t counter,
               jp_stats* stats) {
    Grid<D> G = to_grid<D>(g);
    int rc;
    if ((rc = stats_begin(ctx, stats))) return rc;
    constexpr int NQ = D == 2 ? 4 : 8;
    int min_q = (min_xcell + NQ - 1) / NQ;
    int64_t total = 1;
    for (int d = 0; d < D; ++d) total *= (G.ni[d] + 1) / 2;
    Ptr3 cp = ptr3(coords, D);
    ArgPtrs ap = argptrs(args, nargs);
    jp_stats* st = stats ? dstats(ctx) : nullptr;
    for (int ph = 0; ph < NQ; ++ph) {
        k_inject_phase<D><<<blocks_for(total), kBlock, 0, ctx->stream>>>(
            G, cp, index, nargs, ap, min_q, seed, counter, phase_offsets(D, ph, 2), st);
        JP_LAUNCHED(ctx);
    }
    if ((rc = stats_end(ctx, stats))) return rc;
    return finish(ctx);
}

int check_velgrid(jp_ctx* ctx, const jp_cellgrid* g, const jp_velgrid* vg,
                  const double* const* V) {
    if (!vg) return fail(ctx, JP_ERR_ARG, "null velocity grid%s");
    if (!V) return fail(ctx, JP_ERR_ARG, "null velocity array%s");
    for (int c = 0; c < g->ndim; ++c) {
        if (!V[c]) return fail(ctx, JP_ERR_ARG, "null velocity component%s %lld", "", c);
        for (int d = 0; d < g->ndim; ++d) {
            const jp_axis& a = vg->ax[c][d];
            if (!a.x || !a.dx || !a.inv_dx)
                return fail(ctx, JP_ERR_ARG, "null velocity axis pointer%s (component %lld)", "", c);
            // the bisection is seeded with the particle's cell index, which
            // must address a bracket [x[s], x[s+1]] of the velocity axis
            if (a.n < g->ni[d] + 1)
                return fail(ctx, JP_ERR_ARG,
                            "velocity axis shorter than the cell grid%s (component %lld)", "", c);
        }
    }
    return JP_OK;
}

}  // namespace

extern "C" {

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

int jp_advection(jp_ctx* ctx, const jp_cellgrid* g, double* const* coords, const uint8_t* index,
                 const double* const* V, const jp_velgrid* vg, int method, double alpha,
                 double dt) {
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if ((rc = check_ptrs(ctx, g->ndim, (const void* const*)coords, "coords"))) return rc;
    if (!index) return fail(ctx, JP_ERR_ARG, "null index%s");
    if ((rc = check_velgrid(ctx, g, vg, V))) return rc;
    if (method < JP_EULER || method > JP_RK4)
        return fail(ctx, JP_ERR_ARG, "unknown integrator%s %lld", "", method);
    if (method == JP_RK2 && !(0.0 < alpha && alpha < 1.0))
        return fail(ctx, JP_ERR_ARG, "Only 0 < α < 1 is supported%s");
    RKParams P;
    P.dt = dt;
    P.adt = (1.0 * alpha) * dt;
    P.half = alpha == 0.5;
    double ia = 1.0 / alpha;
    P.w0 = std::fma(-0.5, ia, 1.0);
    P.w1 = 0.5 * ia;
    return g->ndim == 2 ? run_advect<2>(ctx, g, coords, index, V, vg, method, P)
                        : run_advect<3>(ctx, g, coords, index, V, vg, method, P);
}

int jp_move_particles(jp_ctx* ctx, const jp_cellgrid* g, double* const* coords, uint8_t* index,
                      int nargs, double* const* args, jp_stats* stats) {
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if ((rc = check_ptrs(ctx, g->ndim, (const void* const*)coords, "coords"))) return rc;
    if (!index) return fail(ctx, JP_ERR_ARG, "null index%s");
    if (nargs < 0 || nargs > JP_MAX_ARGS) return fail(ctx, JP_ERR_ARG, "nargs out of range%s");
    if (nargs && (rc = check_ptrs(ctx, nargs, (const void* const*)args, "args"))) return rc;
    return g->ndim == 2 ? run_move<2>(ctx, g, coords, index, nargs, args, stats)
                        : run_move<3>(ctx, g, coords, index, nargs, args, stats);
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

}  // extern "C"
