// jp_adjoint.hip -- reverse mode of advection! + particle2grid! on a flat
// particle list (no reference counterpart: JustPIC.jl has no adjoint).
//
// Forward model, per particle: nsteps steps of the reference's integrator
// (src/Particles/Advection/advection.jl:66-91 with Euler.jl, RK2.jl, RK4.jl;
// same operations and fma placement as jp_advect.hip), deletion when the new
// position is not strictly inside the domain (src/Particles/move_safe.jl:
// 99-103), then the particle2grid! gather (src/Interpolations/
// particle_to_grid.jl:37-68, 113-151) of the survivors: each node sums
// w = inv(distance^2) and w*T over the particles of its 2^D adjacent cells.
// Brackets are seeded at the uniform-spacing estimate; for a point strictly
// inside one interval every seed gives the same bracket.
//
// Kernels, one thread per particle:
//  - k_adj_fwd: the whole trajectory in registers; the position at the start
//    of every block of `every` steps goes to the checkpoint array; survivors
//    record their final cell and count themselves into it.
//  - particle2grid! of the survivors without float atomics: the per-cell
//    counts are scanned, particle ids are bucketed by cell and each bucket is
//    sorted by id; every node then merges the id-sorted buckets of its 2^D
//    cells and accumulates w and fma(w, T, .) in increasing particle id, the
//    order of the oracle's restatement, so F and W are reproducible and equal
//    to it bit for bit.
//  - k_adj_bwd: blocks of steps in reverse order; a block is recomputed from
//    its checkpoint into a tape (the position at the start of each step, in
//    global scratch), then reversed step by step.  x0bar and Tpbar are per-thread
//    values.  In 2D the velocity cotangent of a workgroup and block of steps
//    is accumulated in an LDS tile covering exactly the nodes the recompute
//    touched (LDS atomic adds), then added to Vbar once; when that footprint
//    exceeds the tile, and in 3D, it goes to Vbar with FP64 atomic adds.
// The adjoint is the exact derivative of the forward model on the branch the
// forward pass takes (brackets, cells and survivors fixed): the n-linear
// interpolation's weights and slopes, the integrator's stage weights, and for
// the gather dF/dw = (T - F)/W, dw/dx = -2 (x - x_node) w^2 (w taken as
// 1/|x - x_node|^2).  Operation order matches the oracle's restatement, so
// x0bar and Tpbar agree with it bit for bit given the same F, W and Fbar.
#include "jp_ctx.hpp"
#include "jp_splitmath.hpp"

#include <cmath>

using namespace jp;

namespace {

constexpr int kAdjBlock = 512;  // backward workgroup: one 64-cell patch, 8 slot planes

template <int D>
struct AdjArgs {
    Grid<D> G;
    Vel<D> vg;
    const double* V[D];
    double dt;
    double adt;  // (1*alpha)*dt, first RK2 stage
    double w0;   // muladd(-0.5, inv(alpha), 1)
    double w1;   // 0.5*inv(alpha)
    int half;    // alpha == 0.5
    int nsteps;
    int every;
    int64_t np;
    int64_t nn[D];  // nodes per dimension
};

// bracket of p, which lies in [x[0], x[n-1]], seeded at the uniform estimate
__device__ inline int64_t adj_bracket(double p, const Axis& A) {
    const int64_t last = A.n - 2;
    int64_t s = (int64_t)((p - A.x[0]) * A.idx[0]);
    s = s < 0 ? 0 : (s > last ? last : s);
    return find_cell(p, A, s);
}

template <int D>
__device__ inline bool in_limits(const Vel<D>& vg, int c, const double (&p)[D]) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const Axis& A = vg.ax[c][d];
        if (!(A.x[0] <= p[d] && p[d] <= A.x[A.n - 1])) return false;
    }
    return true;
}

template <int D>
struct Corners {
    int64_t at[1 << D];
    double v[1 << D];
    double t[D];
    double iv[D];
    int64_t k[D];  // bracket (lower corner node) per dimension
};

// bracket observers of gather: NoTrack, or BoxTrack (per component and
// dimension, the range of lower-corner nodes seen)
struct NoTrack {
    template <int D>
    __device__ void operator()(int, const int64_t (&)[D]) const {}
};
template <int D>
struct BoxTrack {
    int lo[D][D], hi[D][D];
    __device__ BoxTrack() {
#pragma unroll
        for (int c = 0; c < D; ++c)
#pragma unroll
            for (int d = 0; d < D; ++d) {
                lo[c][d] = 0x7fffffff;
                hi[c][d] = -1;
            }
    }
    __device__ void operator()(int c, const int64_t (&k)[D]) {
#pragma unroll
        for (int cc = 0; cc < D; ++cc)
            if (cc == c)
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    lo[cc][d] = min(lo[cc][d], (int)k[d]);
                    hi[cc][d] = max(hi[cc][d], (int)k[d]);
                }
    }
};

template <int D, class Tr>
__device__ inline void gather(const AdjArgs<D>& A, int c, const double (&p)[D], Corners<D>& K,
                              Tr& tr) {
    int64_t k[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const Axis& X = A.vg.ax[c][d];
        k[d] = adj_bracket(p[d], X);
        K.iv[d] = X.idx[k[d]];
        K.t[d] = (p[d] - X.x[k[d]]) * K.iv[d];
        K.k[d] = k[d];
    }
    tr(c, k);
    const int64_t n0 = A.vg.ax[c][0].n, n1 = A.vg.ax[c][1].n;
    const double* __restrict__ Vc = A.V[c];
#pragma unroll
    for (int q = 0; q < (1 << D); ++q) {
        const int64_t a = k[0] + (q & 1), b = k[1] + ((q >> 1) & 1);
        const int64_t e = D == 3 ? k[D - 1] + ((q >> 2) & 1) : 0;
        K.at[q] = a + n0 * (b + n1 * e);
        K.v[q] = Vc[K.at[q]];
    }
}

// lerp(v, t) of src/Interpolations/ndlerp.jl:11-15
template <int D>
__device__ inline double lerp_n(const Corners<D>& K) {
    const double* v = K.v;
    const double* t = K.t;
    if constexpr (D == 2) {
        return lerp1(t[1], lerp1(t[0], v[0], v[1]), lerp1(t[0], v[2], v[3]));
    } else {
        const double a = lerp1(t[1], lerp1(t[0], v[0], v[1]), lerp1(t[0], v[2], v[3]));
        const double b = lerp1(t[1], lerp1(t[0], v[4], v[5]), lerp1(t[0], v[6], v[7]));
        return lerp1(t[2], a, b);
    }
}

// velocity component c at p; +Inf outside its grid (check_local_limits,
// src/Advection/advection.jl:83-91)
template <int D, class Tr>
__device__ inline double vel_c(const AdjArgs<D>& A, int c, const double (&p)[D], Tr& tr) {
    if (!in_limits(A.vg, c, p)) return __builtin_inf();
    Corners<D> K;
    gather(A, c, p, K, tr);
    return lerp_n(K);
}

template <int D, class Tr>
__device__ inline void vel(const AdjArgs<D>& A, const double (&p)[D], double (&v)[D], Tr& tr) {
#pragma unroll
    for (int c = 0; c < D; ++c) v[c] = vel_c(A, c, p, tr);
}

template <int D>
__device__ inline void vel(const AdjArgs<D>& A, const double (&p)[D], double (&v)[D]) {
    NoTrack tr;
    vel(A, p, v, tr);
}

// destination of velocity cotangents: FP64 atomic adds to Vbar in global
// memory, or into the workgroup's LDS tile (component c's nodes lo[c] ..
// lo[c] + ext[c] - 1 per dimension, at tile offset off[c]) that is added to
// Vbar once per block of steps
template <int D>
struct VSink {
    double* g[D];
    double* tile;  // nullptr: global
    int lo[D][D], ext[D][D], off[D];
};

// adjoint of component c at p for the cotangent vb: Vbar[c] += weights*vb,
// pb += slope*vb.  Only called on points the forward pass found inside the
// component's grid.
template <int D>
__device__ inline void vel_c_bw(const AdjArgs<D>& A, int c, const double (&p)[D], double vb,
                                const VSink<D>& S, double (&pb)[D]) {
    if (vb == 0.0) return;
    Corners<D> K;
    NoTrack nt;
    gather(A, c, p, K, nt);
    const double* v = K.v;
    const double* t = K.t;
    double dd[D];
    if constexpr (D == 2) {
        dd[0] = (1.0 - t[1]) * (v[1] - v[0]) + t[1] * (v[3] - v[2]);
        dd[1] = (1.0 - t[0]) * (v[2] - v[0]) + t[0] * (v[3] - v[1]);
    } else {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                if ((q >> d) & 1) continue;
                const int q1 = q | (1 << d);
                double w = 1.0;
#pragma unroll
                for (int e = 0; e < 3; ++e) {
                    if (e == d) continue;
                    w *= ((q >> e) & 1) ? t[e] : 1.0 - t[e];
                }
                s += w * (v[q1] - v[q]);
            }
            dd[d] = s;
        }
    }
#pragma unroll
    for (int d = 0; d < D; ++d) pb[d] += dd[d] * K.iv[d] * vb;
    int base = 0;  // tile index of the lower corner
    if (S.tile) {
        int stride = 1;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            base += ((int)K.k[d] - S.lo[c][d]) * stride;
            stride *= S.ext[c][d];
        }
    }
#pragma unroll
    for (int q = 0; q < (1 << D); ++q) {
        double w = 1.0;
#pragma unroll
        for (int d = 0; d < D; ++d) w *= ((q >> d) & 1) ? t[d] : 1.0 - t[d];
        if (S.tile) {
            int o = base + (q & 1) + ((q >> 1) & 1) * S.ext[c][0];
            if constexpr (D == 3) o += ((q >> 2) & 1) * S.ext[c][0] * S.ext[c][1];
            unsafeAtomicAdd(S.tile + S.off[c] + o, w * vb);
        } else {
            unsafeAtomicAdd(S.g[c] + K.at[q], w * vb);
        }
    }
}

// one step (Euler.jl, RK2.jl with @muladd, RK4.jl unfused)
template <int D, int M, class Tr>
__device__ inline void step(const AdjArgs<D>& A, const double (&p0)[D], double (&out)[D], Tr& tr) {
    double v0[D];
    vel(A, p0, v0, tr);
    if constexpr (M == JP_EULER) {
#pragma unroll
        for (int d = 0; d < D; ++d) out[d] = fma(A.dt, v0[d], p0[d]);
    } else if constexpr (M == JP_RK2) {
        double p1[D], v1[D];
#pragma unroll
        for (int d = 0; d < D; ++d) p1[d] = fma(A.adt, v0[d], p0[d]);
        vel(A, p1, v1, tr);
        if (A.half) {
#pragma unroll
            for (int d = 0; d < D; ++d) out[d] = fma(A.dt, v1[d], p0[d]);
        } else {
#pragma unroll
            for (int d = 0; d < D; ++d) out[d] = fma(A.dt, fma(A.w1, v1[d], A.w0 * v0[d]), p0[d]);
        }
    } else {
        double q[D], k2[D], k3[D], k4[D];
#pragma unroll
        for (int d = 0; d < D; ++d) q[d] = p0[d] + (A.dt * v0[d]) / 2.0;
        vel(A, q, k2, tr);
#pragma unroll
        for (int d = 0; d < D; ++d) q[d] = p0[d] + (A.dt * k2[d]) / 2.0;
        vel(A, q, k3, tr);
#pragma unroll
        for (int d = 0; d < D; ++d) q[d] = p0[d] + A.dt * k3[d];
        vel(A, q, k4, tr);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const double sum = ((v0[d] + 2.0 * k2[d]) + 2.0 * k3[d]) + k4[d];
            out[d] = p0[d] + (A.dt * sum) / 6.0;
        }
    }
}

template <int D, int M>
__device__ inline void step(const AdjArgs<D>& A, const double (&p0)[D], double (&out)[D]) {
    NoTrack tr;
    step<D, M>(A, p0, out, tr);
}

// reverse of step at p0: pb is the output's cotangent on entry, p0's on exit
template <int D, int M>
__device__ inline void step_bw(const AdjArgs<D>& A, const double (&p0)[D], double (&pb)[D],
                               const VSink<D>& Vb) {
    double ob[D];
#pragma unroll
    for (int d = 0; d < D; ++d) ob[d] = pb[d];
    if constexpr (M == JP_EULER) {
#pragma unroll
        for (int c = 0; c < D; ++c) vel_c_bw(A, c, p0, A.dt * ob[c], Vb, pb);
    } else if constexpr (M == JP_RK2) {
        double v0[D], p1[D], v0b[D], v1b[D], p1b[D];
        vel(A, p0, v0);
#pragma unroll
        for (int d = 0; d < D; ++d) p1[d] = fma(A.adt, v0[d], p0[d]);
        if (A.half) {
#pragma unroll
            for (int d = 0; d < D; ++d) {
                v1b[d] = A.dt * ob[d];
                v0b[d] = 0.0;
            }
        } else {
#pragma unroll
            for (int d = 0; d < D; ++d) {
                v1b[d] = A.dt * A.w1 * ob[d];
                v0b[d] = A.dt * A.w0 * ob[d];
            }
        }
#pragma unroll
        for (int d = 0; d < D; ++d) p1b[d] = 0.0;
#pragma unroll
        for (int c = 0; c < D; ++c) vel_c_bw(A, c, p1, v1b[c], Vb, p1b);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            pb[d] += p1b[d];
            v0b[d] += A.adt * p1b[d];
        }
#pragma unroll
        for (int c = 0; c < D; ++c) vel_c_bw(A, c, p0, v0b[c], Vb, pb);
    } else {
        double k1[D], k2[D], k3[D], q2[D], q3[D], q4[D], kb[4][D], qb[D];
        vel(A, p0, k1);
#pragma unroll
        for (int d = 0; d < D; ++d) q2[d] = p0[d] + (A.dt * k1[d]) / 2.0;
        vel(A, q2, k2);
#pragma unroll
        for (int d = 0; d < D; ++d) q3[d] = p0[d] + (A.dt * k2[d]) / 2.0;
        vel(A, q3, k3);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            q4[d] = p0[d] + A.dt * k3[d];
            kb[0][d] = A.dt / 6.0 * ob[d];
            kb[1][d] = 2.0 * A.dt / 6.0 * ob[d];
            kb[2][d] = 2.0 * A.dt / 6.0 * ob[d];
            kb[3][d] = A.dt / 6.0 * ob[d];
        }
        // k4 = v(q4), q4 = p0 + dt*k3; k3 = v(q3), q3 = p0 + dt*k2/2; k2 = v(q2), q2 = p0 + dt*k1/2
#pragma unroll
        for (int st = 3; st >= 1; --st) {
            const double(&qs)[D] = st == 3 ? q4 : (st == 2 ? q3 : q2);
            const double fac = st == 3 ? A.dt : A.dt / 2.0;
#pragma unroll
            for (int d = 0; d < D; ++d) qb[d] = 0.0;
#pragma unroll
            for (int c = 0; c < D; ++c) vel_c_bw(A, c, qs, kb[st][c], Vb, qb);
#pragma unroll
            for (int d = 0; d < D; ++d) {
                pb[d] += qb[d];
                kb[st - 1][d] += fac * qb[d];
            }
        }
#pragma unroll
        for (int c = 0; c < D; ++c) vel_c_bw(A, c, p0, kb[0][c], Vb, pb);
    }
}

// strictly inside the domain (move_safe.jl:199-206); NaN and Inf fail
template <int D>
__device__ inline bool indomain(const Grid<D>& G, const double (&p)[D]) {
    bool in = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const Axis& X = G.xv[d];
        in = in && (X.x[0] < p[d] && p[d] < X.x[X.n - 1]);
    }
    return in;
}

// node q (bit d: +1 along d) of cell k: linear index, coordinates, inv(distance^2)
template <int D>
__device__ inline int64_t node_of(const AdjArgs<D>& A, const int64_t (&k)[D], int q,
                                  const double (&p)[D], double (&xn)[D], double& wi) {
    int64_t N[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        N[d] = k[d] + ((q >> d) & 1);
        xn[d] = A.G.xv[d].x[N[d]];
    }
    // distance (src/Interpolations/utils.jl:9-19), node minus particle
    double e = xn[0] - p[0];
    double s = e * e;
#pragma unroll
    for (int d = 1; d < D; ++d) {
        e = xn[d] - p[d];
        s = s + e * e;
    }
    const double dist = sqrt(s);
    wi = 1.0 / (dist * dist);
    return D == 2 ? N[0] + A.nn[0] * N[1] : N[0] + A.nn[0] * (N[1] + A.nn[1] * N[D - 1]);
}

template <int D, int M>
__global__ __launch_bounds__(kBlock) void k_adj_fwd(AdjArgs<D> A, CPtr3 x0,
                                                    double* __restrict__ ckpt,
                                                    uint8_t* __restrict__ alive, Ptr3 xN,
                                                    int32_t* __restrict__ cell,
                                                    int32_t* __restrict__ cnt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.np) return;
    double x[D];
#pragma unroll
    for (int d = 0; d < D; ++d) x[d] = x0.p[d][i];
    bool a = indomain(A.G, x);
#pragma unroll
    for (int d = 0; d < D; ++d) ckpt[(int64_t)d * A.np + i] = x[d];
    int64_t blk = 0;
    int left = A.every;
    for (int s = 0; s < A.nsteps; ++s) {
        if (left == 0) {
            ++blk;
            left = A.every;
#pragma unroll
            for (int d = 0; d < D; ++d) ckpt[(blk * D + d) * A.np + i] = x[d];
        }
        --left;
        if (!a) continue;
        double y[D];
        step<D, M>(A, x, y);
#pragma unroll
        for (int d = 0; d < D; ++d) x[d] = y[d];
        a = indomain(A.G, x);
    }
    alive[i] = a ? 1 : 0;
#pragma unroll
    for (int d = 0; d < D; ++d) xN.p[d][i] = x[d];
    int32_t c = -1;
    if (a) {
        int64_t k[D];
#pragma unroll
        for (int d = 0; d < D; ++d) k[d] = adj_bracket(x[d], A.G.xv[d]);
        c = (int32_t)lin<D>(A.G.ni, k);
        atomicAdd(cnt + c, 1);
    }
    cell[i] = c;
}

// The forward model in the FP32-split form (jp_adjoint_forward_split): each
// particle is a cell index plus a float32 offset in cell units, every step
// the split advection's per-particle arithmetic (jp_splitmath.hpp: velocity
// scaled to cells per step in float32, every corner read from global memory,
// the same values a split advection window holds) followed by the split
// move's renormalisation to the destination cell in FP64; a particle whose
// cell leaves the grid, or whose offset is not finite, is deleted.  The
// checkpoints, xN and the gather use the FP64 position x_I + u*h, so the
// FP64 backward recomputes each block from the split trajectory.
struct NoWindow {  // every corner from global memory
    __host__ __device__ static constexpr int n(int) { return 1; }
    __host__ __device__ static constexpr int w(int) { return 0; }
    __host__ __device__ static constexpr int size() { return 1; }
};

template <int D, int M>
__global__ __launch_bounds__(kBlock) void k_adj_fwd_split(
    AdjArgs<D> A, split::SVelIn<D> SV, const SplitGeom* __restrict__ geo, split::RKf P, CPtr3 x0,
    double* __restrict__ ckpt, uint8_t* __restrict__ alive, Ptr3 xN, int32_t* __restrict__ cell,
    int32_t* __restrict__ cnt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.np || geo->bad) return;  // a rejected grid is reported by the call
    split::SMap<D> m;
    double sc[D];
#pragma unroll
    for (int c = 0; c < D; ++c) {
        sc[c] = A.dt / A.G.xv[c].dx[0];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            m.r[c][d] = geo->r[c][d];
            m.o[c][d] = geo->o[c][d];
        }
    }
    double x[D];
#pragma unroll
    for (int d = 0; d < D; ++d) x[d] = x0.p[d][i];
    bool a = indomain(A.G, x);
    int I[D];
    float u[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        I[d] = 0;
        u[d] = 0.0f;
        if (a) {
            const Axis& X = A.G.xv[d];
            I[d] = (int)adj_bracket(x[d], X);
            const double ud = (x[d] - X.x[I[d]]) * X.idx[I[d]];
            float uf = (float)ud;
            if (ud < 1.0 && uf >= 1.0f) uf = 0x1.fffffep-1f;  // stays inside its cell
            u[d] = uf;
        }
    }
#pragma unroll
    for (int d = 0; d < D; ++d) ckpt[(int64_t)d * A.np + i] = x[d];
    const int w0[D][D] = {};
    int64_t blk = 0;
    int left = A.every;
    for (int s = 0; s < A.nsteps; ++s) {
        if (left == 0) {
            ++blk;
            left = A.every;
#pragma unroll
            for (int d = 0; d < D; ++d) ckpt[(blk * D + d) * A.np + i] = x[d];
        }
        --left;
        if (!a) continue;
        split::CellMap<D> cm;
        split::cell_map<D, false>(m, I, cm);
        float out[D];
        split::advect_split_one<D, M, NoWindow>(nullptr, SV, sc, w0, I, cm, P, u, out);
        bool gone = false;
        int sd[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const bool fin = fabsf(out[d]) < 1.0e6f;  // false for NaN/Inf
            sd[d] = fin ? (int)floorf(out[d]) : 0;
            const int J = I[d] + sd[d];
            gone = gone || !fin || J < 0 || J >= A.G.ni[d];
        }
        if (gone) {  // deleted: the last position, as the FP64 model keeps it
#pragma unroll
            for (int d = 0; d < D; ++d)
                x[d] = A.G.xv[d].x[I[d]] + (double)out[d] * A.G.xv[d].dx[I[d]];
            a = false;
            continue;
        }
#pragma unroll
        for (int d = 0; d < D; ++d) {
            I[d] += sd[d];
            u[d] = (float)((double)out[d] - (double)sd[d]);
            x[d] = A.G.xv[d].x[I[d]] + (double)u[d] * A.G.xv[d].dx[I[d]];
        }
    }
    alive[i] = a ? 1 : 0;
#pragma unroll
    for (int d = 0; d < D; ++d) xN.p[d][i] = x[d];
    int32_t c = -1;
    if (a) {
        int64_t k[D];
#pragma unroll
        for (int d = 0; d < D; ++d) k[d] = I[d];
        c = (int32_t)lin<D>(A.G.ni, k);
        atomicAdd(cnt + c, 1);
    }
    cell[i] = c;
}

// exclusive scan of the C cell counts into start[0..C] (one workgroup)
constexpr int kScan = 1024;
__global__ __launch_bounds__(kScan) void k_adj_scan(const int32_t* __restrict__ cnt,
                                                    int32_t* __restrict__ start, int64_t C) {
    __shared__ int32_t part[kScan];
    const int t = threadIdx.x;
    const int64_t per = (C + kScan - 1) / kScan;
    const int64_t lo = t * per < C ? t * per : C;
    const int64_t hi = lo + per < C ? lo + per : C;
    int32_t s = 0;
    for (int64_t c = lo; c < hi; ++c) s += cnt[c];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < kScan; off <<= 1) {
        const int32_t v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int32_t run = part[t] - s;  // exclusive prefix of this thread's chunk
    for (int64_t c = lo; c < hi; ++c) {
        start[c] = run;
        run += cnt[c];
    }
    if (t == kScan - 1) start[C] = part[t];
}

// bucket the survivors by cell (order inside a bucket fixed by k_adj_sort);
// consumes the counts
__global__ __launch_bounds__(kBlock) void k_adj_bucket(const int32_t* __restrict__ cell,
                                                      int32_t* __restrict__ cnt,
                                                      const int32_t* __restrict__ start,
                                                      int32_t* __restrict__ list, int64_t np) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const int32_t c = cell[i];
    if (c < 0) return;
    const int32_t slot = atomicSub(cnt + c, 1) - 1;
    list[start[c] + slot] = (int32_t)i;
}

// sort every bucket by particle id (insertion sort; buckets hold tens of ids)
__global__ __launch_bounds__(kBlock) void k_adj_sort(const int32_t* __restrict__ start,
                                                    int32_t* __restrict__ list, int64_t C) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const int32_t lo = start[c], hi = start[c + 1];
    for (int32_t j = lo + 1; j < hi; ++j) {
        const int32_t v = list[j];
        int32_t k = j - 1;
        while (k >= lo && list[k] > v) {
            list[k + 1] = list[k];
            --k;
        }
        list[k + 1] = v;
    }
}

// per node: merge the id-sorted buckets of its 2^D cells; w = inv(distance^2),
// W += w, S = fma(w, T, S) in increasing particle id; F = S / W (2D),
// S * inv(W) (3D); a node without particles gets 0/0 = NaN
template <int D>
__global__ __launch_bounds__(kBlock) void k_adj_gather(AdjArgs<D> A,
                                                      const int32_t* __restrict__ start,
                                                      const int32_t* __restrict__ list, CPtr3 xN,
                                                      const double* __restrict__ Tp,
                                                      double* __restrict__ F,
                                                      double* __restrict__ W, int64_t nodes) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nodes) return;
    int64_t N[D];
    int64_t r = n;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        N[d] = r % A.nn[d];
        r /= A.nn[d];
    }
    double xn[D];
#pragma unroll
    for (int d = 0; d < D; ++d) xn[d] = A.G.xv[d].x[N[d]];
    int32_t h[1 << D], e[1 << D];
#pragma unroll
    for (int q = 0; q < (1 << D); ++q) {
        int64_t I[D];
        bool ok = true;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            I[d] = N[d] - ((q >> d) & 1);
            ok = ok && I[d] >= 0 && I[d] < A.G.ni[d];
        }
        h[q] = e[q] = 0;
        if (ok) {
            const int64_t c = lin<D>(A.G.ni, I);
            h[q] = start[c];
            e[q] = start[c + 1];
        }
    }
    double w = 0.0, s = 0.0;
    for (;;) {
        int best = -1;
        int32_t id = 0;
#pragma unroll
        for (int q = 0; q < (1 << D); ++q) {
            if (h[q] < e[q]) {
                const int32_t v = list[h[q]];
                if (best < 0 || v < id) {
                    best = q;
                    id = v;
                }
            }
        }
        if (best < 0) break;
#pragma unroll
        for (int q = 0; q < (1 << D); ++q) h[q] += q == best ? 1 : 0;
        // distance (src/Interpolations/utils.jl:9-19), node minus particle
        double dl = xn[0] - xN.p[0][id];
        double ss = dl * dl;
#pragma unroll
        for (int d = 1; d < D; ++d) {
            dl = xn[d] - xN.p[d][id];
            ss = ss + dl * dl;
        }
        const double dist = sqrt(ss);
        const double wi = 1.0 / (dist * dist);
        w += wi;
        s = fma(wi, Tp[id], s);
    }
    W[n] = w;
    F[n] = D == 2 ? s / w : s * (1.0 / w);
}

__global__ __launch_bounds__(kBlock) void k_adj_zero(double* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// LDS tile for the velocity cotangents of one workgroup and block of steps
// (2D only), the 64 KB workgroup budget less the static box words.  A
// workgroup from a compact patch touches ~2-6 k nodes over ten steps; in 3D
// the footprint rarely fits and global atomics are used.
constexpr size_t kLdsBudget = 65536 - 256;
template <int D>
inline int tile_cap() {
    return D == 2 ? (int)(kLdsBudget / sizeof(double)) : 0;
}

template <int D, int M>
__global__ __launch_bounds__(kAdjBlock) void k_adj_bwd(
    AdjArgs<D> A, const double* __restrict__ ckpt, const uint8_t* __restrict__ alive,
    const double* __restrict__ Tp, const double* __restrict__ F, const double* __restrict__ W,
    const double* __restrict__ Fbar, Ptr3 Vb, Ptr3 x0b, double* __restrict__ Tpb,
    double* __restrict__ tape, int cap) {
    // tape (global scratch, [every][D][np]): step-start positions of the block
    // of steps being reversed; dynamic LDS: the cotangent tile of cap doubles
    extern __shared__ double tile[];
    __shared__ int sbox[D][D][2];
    const int lt = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * kAdjBlock + lt;
    // every thread reaches the block-wide barriers below
    const bool live = i < A.np && alive[i];
    VSink<D> S;
#pragma unroll
    for (int c = 0; c < D; ++c) S.g[c] = Vb.p[c];
    S.tile = nullptr;
    double pb[D];
#pragma unroll
    for (int d = 0; d < D; ++d) pb[d] = 0.0;
    double tb = 0.0;
    const int nblk = A.nsteps == 0 ? 1 : (A.nsteps + A.every - 1) / A.every;
    for (int b = nblk - 1; b >= 0; --b) {
        const int s0 = b * A.every;
        const int n = A.nsteps - s0 < A.every ? A.nsteps - s0 : A.every;
        BoxTrack<D> box;
        if (live) {
            double x[D];
#pragma unroll
            for (int d = 0; d < D; ++d) x[d] = ckpt[((int64_t)b * D + d) * A.np + i];
            for (int j = 0; j < n; ++j) {
#pragma unroll
                for (int d = 0; d < D; ++d) tape[((int64_t)j * D + d) * A.np + i] = x[d];
                double y[D];
                step<D, M>(A, x, y, box);
#pragma unroll
                for (int d = 0; d < D; ++d) x[d] = y[d];
            }
            if (b == nblk - 1) {
                // particle2grid! adjoint at the final position x
                int64_t k[D];
#pragma unroll
                for (int d = 0; d < D; ++d) k[d] = adj_bracket(x[d], A.G.xv[d]);
                const double t = Tp[i];
#pragma unroll
                for (int q = 0; q < (1 << D); ++q) {
                    double xn[D], wi;
                    const int64_t nd = node_of(A, k, q, x, xn, wi);
                    const double iw = 1.0 / W[nd];
                    const double fb = Fbar[nd];
                    tb += fb * wi * iw;
                    const double gw = fb * (t - F[nd]) * iw;
#pragma unroll
                    for (int d = 0; d < D; ++d) pb[d] += gw * (-2.0 * (x[d] - xn[d]) * wi * wi);
                }
            }
        }
        int total = 0;
        if constexpr (D == 2) {
            // the workgroup's node range per component and dimension
            if (lt < D * D) {
                (&sbox[0][0][0])[2 * lt] = 0x7fffffff;
                (&sbox[0][0][0])[2 * lt + 1] = -1;
            }
            __syncthreads();
            if (live && n > 0) {
#pragma unroll
                for (int c = 0; c < D; ++c)
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        atomicMin(&sbox[c][d][0], box.lo[c][d]);
                        atomicMax(&sbox[c][d][1], box.hi[c][d]);
                    }
            }
            __syncthreads();
            bool any = true;
#pragma unroll
            for (int c = 0; c < D; ++c) {
                int sz = 1;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    S.lo[c][d] = sbox[c][d][0];
                    const int e = sbox[c][d][1] - sbox[c][d][0] + 2;  // corners k .. k+1
                    any = any && sbox[c][d][1] >= sbox[c][d][0];
                    S.ext[c][d] = e;
                    sz *= e > 0 ? e : 0;
                }
                S.off[c] = total;
                total += sz;
            }
            if (any && total <= cap) {
                for (int t = lt; t < total; t += kAdjBlock) tile[t] = 0.0;
                S.tile = tile;
            } else {
                S.tile = nullptr;
                total = 0;
            }
            __syncthreads();
        }
        if (live) {
            for (int j = n - 1; j >= 0; --j) {
                double p0[D];
#pragma unroll
                for (int d = 0; d < D; ++d) p0[d] = tape[((int64_t)j * D + d) * A.np + i];
                step_bw<D, M>(A, p0, pb, S);
            }
        }
        if constexpr (D == 2) {
            __syncthreads();
            // the tile's nonzero entries go to Vbar
            for (int t = lt; t < total; t += kAdjBlock) {
                const double v = tile[t];
                if (v == 0.0) continue;
                int c = 0;
#pragma unroll
                for (int cc = 1; cc < D; ++cc) c = t >= S.off[cc] ? cc : c;
                int r = t - S.off[c];
                int64_t gi = 0, stride = 1;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const int64_t node = S.lo[c][d] + r % S.ext[c][d];
                    r /= S.ext[c][d];
                    gi += node * stride;
                    stride *= A.vg.ax[c][d].n;
                }
                unsafeAtomicAdd(S.g[c] + gi, v);
            }
            __syncthreads();  // the tile and the box are reused by the next block
            S.tile = nullptr;
        }
    }
    if (i < A.np) {
#pragma unroll
        for (int d = 0; d < D; ++d) x0b.p[d][i] = pb[d];
        Tpb[i] = tb;
    }
}

template <int D>
Vel<D> vel_of(const jp_velgrid* vg) {
    Vel<D> V;
    for (int c = 0; c < D; ++c)
        for (int d = 0; d < D; ++d) {
            const jp_axis& a = vg->ax[c][d];
            V.ax[c][d] = Axis{a.x, a.dx, a.inv_dx, a.n};
        }
    return V;
}

int check_args(jp_ctx* ctx, const jp_cellgrid* g, const jp_velgrid* vg, const double* const* V,
               int method, double alpha, int nsteps, int every, int64_t np) {
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if (!vg || !V) return fail(ctx, JP_ERR_ARG, "null velocity grid or array%s");
    for (int c = 0; c < g->ndim; ++c) {
        if (!V[c]) return fail(ctx, JP_ERR_ARG, "null velocity component%s %lld", "", c);
        for (int d = 0; d < g->ndim; ++d) {
            const jp_axis& a = vg->ax[c][d];
            if (!a.x || !a.dx || !a.inv_dx || a.n < 2)
                return fail(ctx, JP_ERR_ARG, "bad velocity axis%s (component %lld)", "", c);
        }
    }
    if (method < JP_EULER || method > JP_RK4)
        return fail(ctx, JP_ERR_ARG, "unknown integrator%s %lld", "", method);
    if (method == JP_RK2 && !(0.0 < alpha && alpha < 1.0))
        return fail(ctx, JP_ERR_ARG, "Only 0 < α < 1 is supported%s");
    if (nsteps < 0) return fail(ctx, JP_ERR_ARG, "negative step count%s %lld", "", nsteps);
    if (every < 1 || every > JP_ADJ_MAX_EVERY)
        return fail(ctx, JP_ERR_ARG, "checkpoint interval out of range%s %lld", "", every);
    if (np < 0) return fail(ctx, JP_ERR_ARG, "negative particle count%s %lld", "", np);
    if (np >= ((int64_t)1 << 31))  // 32-bit particle ids and launch size
        return fail(ctx, JP_ERR_ARG, "too many particles%s %lld", "", np);
    int64_t C = 1;
    for (int d = 0; d < g->ndim; ++d) C *= g->ni[d];
    if (C >= ((int64_t)1 << 31))  // 32-bit cell ids
        return fail(ctx, JP_ERR_ARG, "too many cells%s %lld", "", C);
    return JP_OK;
}

template <int D>
AdjArgs<D> make_args(const jp_cellgrid* g, const jp_velgrid* vg, const double* const* V,
                     double alpha, double dt, int nsteps, int every, int64_t np) {
    AdjArgs<D> A;
    A.G = to_grid<D>(g);
    A.vg = vel_of<D>(vg);
    for (int c = 0; c < D; ++c) A.V[c] = V[c];
    A.dt = dt;
    A.adt = (1.0 * alpha) * dt;
    A.half = alpha == 0.5;
    const double ia = 1.0 / alpha;
    A.w0 = std::fma(-0.5, ia, 1.0);
    A.w1 = 0.5 * ia;
    A.nsteps = nsteps;
    A.every = every;
    A.np = np;
    for (int d = 0; d < D; ++d) A.nn[d] = g->ni[d] + 1;
    return A;
}

int zero(jp_ctx* ctx, double* p, int64_t n) {
    if (n <= 0) return JP_OK;
    k_adj_zero<<<(unsigned)blocks_for(n), kBlock, 0, ctx->stream>>>(p, n);
    JP_LAUNCHED(ctx);
    return JP_OK;
}

template <int D>
int64_t node_count(const jp_cellgrid* g) {
    int64_t n = 1;
    for (int d = 0; d < D; ++d) n *= g->ni[d] + 1;
    return n;
}

// the trajectories (FP64, or the split form with its velocity description,
// RK parameters and the current geometry verdict), then the gather
template <int D>
struct SplitFwd {
    bool on = false;
    split::SVelIn<D> SV;
    split::RKf P;
    const SplitGeom* geo = nullptr;
};

template <int D, int M>
int forward(jp_ctx* ctx, const AdjArgs<D>& A, const double* const* x0, const double* Tp,
            double* ckpt, uint8_t* alive, double* const* xN, double* F, double* W,
            int64_t nodes, const SplitFwd<D>& sf = SplitFwd<D>()) {
    // scratch: counts (C) | bucket starts (C+1) | cell of each particle | buckets
    const int64_t C = A.G.C;
    const size_t a = ((size_t)C * 4 + 255) & ~(size_t)255;
    const size_t b = a + (((size_t)(C + 1) * 4 + 255) & ~(size_t)255);
    const size_t c = b + (((size_t)A.np * 4 + 255) & ~(size_t)255);
    int rc = ensure_scratch(ctx, c + (size_t)A.np * 4 + 4);
    if (rc) return rc;
    char* base = static_cast<char*>(ctx->scratch);
    int32_t* cnt = reinterpret_cast<int32_t*>(base);
    int32_t* start = reinterpret_cast<int32_t*>(base + a);
    int32_t* cell = reinterpret_cast<int32_t*>(base + b);
    int32_t* list = reinterpret_cast<int32_t*>(base + c);
    if ((rc = zero_bytes(ctx, cnt, (size_t)C * 4))) return rc;
    CPtr3 xp{{nullptr, nullptr, nullptr}}, xq{{nullptr, nullptr, nullptr}};
    Ptr3 xo{{nullptr, nullptr, nullptr}};
    for (int d = 0; d < D; ++d) {
        xp.p[d] = x0[d];
        xo.p[d] = xN[d];
        xq.p[d] = xN[d];
    }
    if (A.np > 0) {
        if (sf.on)
            k_adj_fwd_split<D, M><<<(unsigned)blocks_for(A.np), kBlock, 0, ctx->stream>>>(
                A, sf.SV, sf.geo, sf.P, xp, ckpt, alive, xo, cell, cnt);
        else
            k_adj_fwd<D, M><<<(unsigned)blocks_for(A.np), kBlock, 0, ctx->stream>>>(
                A, xp, ckpt, alive, xo, cell, cnt);
        JP_LAUNCHED(ctx);
    }
    k_adj_scan<<<1, kScan, 0, ctx->stream>>>(cnt, start, C);
    JP_LAUNCHED(ctx);
    if (A.np > 0) {
        k_adj_bucket<<<(unsigned)blocks_for(A.np), kBlock, 0, ctx->stream>>>(cell, cnt, start,
                                                                             list, A.np);
        JP_LAUNCHED(ctx);
    }
    k_adj_sort<<<(unsigned)blocks_for(C), kBlock, 0, ctx->stream>>>(start, list, C);
    JP_LAUNCHED(ctx);
    k_adj_gather<D><<<(unsigned)blocks_for(nodes), kBlock, 0, ctx->stream>>>(A, start, list, xq,
                                                                            Tp, F, W, nodes);
    JP_LAUNCHED(ctx);
    return sf.on ? split_finish(ctx) : finish(ctx);
}

template <int D, int M>
int backward(jp_ctx* ctx, const AdjArgs<D>& A, const jp_velgrid* vg, const double* ckpt,
             const uint8_t* alive, const double* Tp, const double* F, const double* W,
             const double* Fbar, double* const* Vbar, double* const* x0bar, double* Tpbar) {
    Ptr3 vb{{nullptr, nullptr, nullptr}}, xb{{nullptr, nullptr, nullptr}};
    for (int c = 0; c < D; ++c) {
        int64_t m = 1;
        for (int d = 0; d < D; ++d) m *= vg->ax[c][d].n;
        int rc = zero(ctx, Vbar[c], m);
        if (rc) return rc;
        vb.p[c] = Vbar[c];
        xb.p[c] = x0bar[c];
    }
    if (A.np > 0) {
        const int cap = tile_cap<D>();
        const size_t lds = (size_t)cap * sizeof(double);
        int rc = ensure_scratch(ctx, (size_t)A.every * D * A.np * sizeof(double));
        if (rc) return rc;
        double* tape = static_cast<double*>(ctx->scratch);
        k_adj_bwd<D, M><<<(unsigned)blocks_for(A.np, kAdjBlock), kAdjBlock, lds, ctx->stream>>>(
            A, ckpt, alive, Tp, F, W, Fbar, vb, xb, Tpbar, tape, cap);
        JP_LAUNCHED(ctx);
    }
    return finish(ctx);
}

int null_ptrs(jp_ctx* ctx, int D, int64_t np, const void* const* list, int n, const char* what) {
    if (np == 0) return JP_OK;
    if (!list) return fail(ctx, JP_ERR_ARG, "null %s", what);
    for (int d = 0; d < n; ++d)
        if (!list[d]) return fail(ctx, JP_ERR_ARG, "null %s pointer %lld", what, d);
    (void)D;
    return JP_OK;
}

}  // namespace

extern "C" int jp_adjoint_forward(jp_ctx* ctx, const jp_cellgrid* g, const jp_velgrid* vg,
                                  const double* const* V, int method, double alpha, double dt,
                                  int nsteps, int every, int64_t np, const double* const* x0,
                                  const double* Tp, double* ckpt, uint8_t* alive,
                                  double* const* xN, double* F, double* W) {
    int rc = check_args(ctx, g, vg, V, method, alpha, nsteps, every, np);
    if (rc) return rc;
    const int D = g->ndim;
    if (!F || !W) return fail(ctx, JP_ERR_ARG, "null nodal output%s");
    if ((rc = null_ptrs(ctx, D, np, (const void* const*)x0, D, "x0"))) return rc;
    if ((rc = null_ptrs(ctx, D, np, (const void* const*)xN, D, "xN"))) return rc;
    if (np > 0 && (!Tp || !ckpt || !alive))
        return fail(ctx, JP_ERR_ARG, "null particle array (Tp, ckpt or alive)%s");
    if (D == 2) {
        const auto A = make_args<2>(g, vg, V, alpha, dt, nsteps, every, np);
        const int64_t nodes = node_count<2>(g);
        if (method == JP_EULER)
            return forward<2, JP_EULER>(ctx, A, x0, Tp, ckpt, alive, xN, F, W, nodes);
        if (method == JP_RK2)
            return forward<2, JP_RK2>(ctx, A, x0, Tp, ckpt, alive, xN, F, W, nodes);
        return forward<2, JP_RK4>(ctx, A, x0, Tp, ckpt, alive, xN, F, W, nodes);
    }
    const auto A = make_args<3>(g, vg, V, alpha, dt, nsteps, every, np);
    const int64_t nodes = node_count<3>(g);
    if (method == JP_EULER)
        return forward<3, JP_EULER>(ctx, A, x0, Tp, ckpt, alive, xN, F, W, nodes);
    if (method == JP_RK2) return forward<3, JP_RK2>(ctx, A, x0, Tp, ckpt, alive, xN, F, W, nodes);
    return forward<3, JP_RK4>(ctx, A, x0, Tp, ckpt, alive, xN, F, W, nodes);
}

namespace {

template <int D>
int forward_split(jp_ctx* ctx, const jp_cellgrid* g, const jp_velgrid* vg,
                  const double* const* V, int method, double alpha, double dt, int nsteps,
                  int every, int64_t np, const double* const* x0, const double* Tp, double* ckpt,
                  uint8_t* alive, double* const* xN, double* F, double* W) {
    const auto A = make_args<D>(g, vg, V, alpha, dt, nsteps, every, np);
    int rc = split_geometry(ctx, g, vg);
    if (rc) return rc;
    SplitFwd<D> sf;
    sf.on = true;
    sf.geo = geo_ptr(ctx);
    for (int c = 0; c < D; ++c) {
        sf.SV.v[c] = V[c];
        for (int d = 0; d < D; ++d) sf.SV.n[c][d] = (int)vg->ax[c][d].n;
    }
    sf.P.alpha = (float)alpha;
    sf.P.half = alpha == 0.5;
    const double ia = 1.0 / alpha;
    sf.P.w0 = (float)std::fma(-0.5, ia, 1.0);
    sf.P.w1 = (float)(0.5 * ia);
    const int64_t nodes = node_count<D>(g);
    if (method == JP_EULER)
        return forward<D, JP_EULER>(ctx, A, x0, Tp, ckpt, alive, xN, F, W, nodes, sf);
    if (method == JP_RK2)
        return forward<D, JP_RK2>(ctx, A, x0, Tp, ckpt, alive, xN, F, W, nodes, sf);
    return forward<D, JP_RK4>(ctx, A, x0, Tp, ckpt, alive, xN, F, W, nodes, sf);
}

}  // namespace

extern "C" int jp_adjoint_forward_split(jp_ctx* ctx, const jp_cellgrid* g, const jp_velgrid* vg,
                                        const double* const* V, int method, double alpha,
                                        double dt, int nsteps, int every, int64_t np,
                                        const double* const* x0, const double* Tp, double* ckpt,
                                        uint8_t* alive, double* const* xN, double* F, double* W) {
    int rc = check_args(ctx, g, vg, V, method, alpha, nsteps, every, np);
    if (rc) return rc;
    if (!F || !W) return fail(ctx, JP_ERR_ARG, "null nodal output%s");
    if ((rc = null_ptrs(ctx, g->ndim, np, (const void* const*)x0, g->ndim, "x0"))) return rc;
    if ((rc = null_ptrs(ctx, g->ndim, np, (const void* const*)xN, g->ndim, "xN"))) return rc;
    if (np > 0 && (!Tp || !ckpt || !alive))
        return fail(ctx, JP_ERR_ARG, "null particle array (Tp, ckpt or alive)%s");
    return g->ndim == 2 ? forward_split<2>(ctx, g, vg, V, method, alpha, dt, nsteps, every, np,
                                           x0, Tp, ckpt, alive, xN, F, W)
                        : forward_split<3>(ctx, g, vg, V, method, alpha, dt, nsteps, every, np,
                                           x0, Tp, ckpt, alive, xN, F, W);
}

extern "C" int jp_adjoint_backward(jp_ctx* ctx, const jp_cellgrid* g, const jp_velgrid* vg,
                                   const double* const* V, int method, double alpha, double dt,
                                   int nsteps, int every, int64_t np, const double* ckpt,
                                   const uint8_t* alive, const double* Tp, const double* F,
                                   const double* W, const double* Fbar, double* const* Vbar,
                                   double* const* x0bar, double* Tpbar) {
    int rc = check_args(ctx, g, vg, V, method, alpha, nsteps, every, np);
    if (rc) return rc;
    const int D = g->ndim;
    if (!F || !W || !Fbar) return fail(ctx, JP_ERR_ARG, "null nodal input%s");
    if ((rc = null_ptrs(ctx, D, 1, (const void* const*)Vbar, D, "Vbar"))) return rc;
    if ((rc = null_ptrs(ctx, D, np, (const void* const*)x0bar, D, "x0bar"))) return rc;
    if (np > 0 && (!Tp || !ckpt || !alive || !Tpbar))
        return fail(ctx, JP_ERR_ARG, "null particle array (Tp, ckpt, alive or Tpbar)%s");
    if (D == 2) {
        const auto A = make_args<2>(g, vg, V, alpha, dt, nsteps, every, np);
        if (method == JP_EULER)
            return backward<2, JP_EULER>(ctx, A, vg, ckpt, alive, Tp, F, W, Fbar, Vbar, x0bar,
                                         Tpbar);
        if (method == JP_RK2)
            return backward<2, JP_RK2>(ctx, A, vg, ckpt, alive, Tp, F, W, Fbar, Vbar, x0bar,
                                       Tpbar);
        return backward<2, JP_RK4>(ctx, A, vg, ckpt, alive, Tp, F, W, Fbar, Vbar, x0bar, Tpbar);
    }
    const auto A = make_args<3>(g, vg, V, alpha, dt, nsteps, every, np);
    if (method == JP_EULER)
        return backward<3, JP_EULER>(ctx, A, vg, ckpt, alive, Tp, F, W, Fbar, Vbar, x0bar,
                                     Tpbar);
    if (method == JP_RK2)
        return backward<3, JP_RK2>(ctx, A, vg, ckpt, alive, Tp, F, W, Fbar, Vbar, x0bar, Tpbar);
    return backward<3, JP_RK4>(ctx, A, vg, ckpt, alive, Tp, F, W, Fbar, Vbar, x0bar, Tpbar);
}
