// jp_advect.hip -- advection! (src/Particles/Advection/advection.jl:35-91) on
// gfx950, FP64 drop-in, 2D and 3D, Euler / RK2 / RK4.
//
// One thread per slot; the velocity components are gathered from global
// memory, where the few MB of a velocity grid stay L2/MALL-resident.  A bracket
// lookup costs one guess (uniform spacing estimate) and one strict test against
// the stored axis, with the reference search for everything else (ties,
// refined axes).  An LDS-tiled variant (velocity and axis windows per tile of
// cells, a thread per cell walking its slots) was as fast in 2D and half as
// fast in 3D, where its register footprint left one wave per SIMD.
#include "jp_ctx.hpp"

using namespace jp;

namespace {

template <int D>
struct VPtrs {
    const double* v[D];
};

struct RKParams {
    double dt;
    double adt;  // (s*alpha)*dt
    int half;    // alpha == 0.5
    double w0;   // muladd(-0.5, inv(alpha), 1)
    double w1;   // 0.5*inv(alpha)
};

// interp_velocity2particle (advection.jl:93-148) for component c: +Inf outside
// the component's grid (check_local_limits, src/Advection/advection.jl:83-91),
// t = (p - x_k)*inv(dx_k) (normalize_coordinates, src/Interpolations/utils.jl:54-58),
// n-linear lerp of the corner values (src/Interpolations/ndlerp.jl:11-15) in
// extract_field_corners order (src/Advection/advection.jl:47-71).

// Row mapping: a wavefront holds 64 consecutive cells of one grid row at one
// slot plane (grid.y), so coordinate and mask accesses coalesce.  Brackets:
// the uniform-spacing guess confirmed with the strict test x[k] < p < x[k+1]
// (a strictly interior point has one bracket, the one the reference's
// bisection returns whatever its seed), else the reference search seeded at
// the particle's cell.
template <int D>
struct AxisGuess {
    double lo[D][D], hi[D][D], i0[D][D];  // x[0], x[n-1], inv_dx[0] of axis (c, d)
};

__device__ inline int64_t bracket_slot(double p, const Axis& A, double x0, double i0,
                                       int64_t seed) {
    const int64_t last = A.n - 2;
    int64_t gk = (int64_t)((p - x0) * i0);
    gk = gk < 0 ? 0 : (gk > last ? last : gk);
    if (A.x[gk] < p && p < A.x[gk + 1]) return gk;
    return find_cell(p, A, seed);
}

template <int D>
__device__ inline double interp_slot(const AxisGuess<D>& Gs, const Vel<D>& vg,
                                     const VPtrs<D>& V, int c, const double (&p)[D],
                                     const int64_t (&I)[D]) {
    bool in = true;
#pragma unroll
    for (int d = 0; d < D; ++d) in = in && (Gs.lo[c][d] <= p[d] && p[d] <= Gs.hi[c][d]);
    if (!in) return __builtin_inf();
    int64_t k[D];
    double t[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const Axis& A = vg.ax[c][d];
        k[d] = bracket_slot(p[d], A, Gs.lo[c][d], Gs.i0[c][d], I[d]);
        t[d] = (p[d] - A.x[k[d]]) * A.idx[k[d]];
    }
    const double* __restrict__ F = V.v[c];
    const int64_t n0 = vg.ax[c][0].n, n1 = vg.ax[c][1].n;
    int64_t b = k[0] + n0 * k[1];
    if constexpr (D == 3) b += n0 * n1 * k[2];
    double q[1 << D];
#pragma unroll
    for (int r = 0; r < (1 << D); ++r) {
        const int64_t o = (r & 1) + ((r >> 1) & 1) * n0 + (D == 3 ? ((r >> 2) & 1) * n0 * n1 : 0);
        q[r] = F[b + o];
    }
    const double lo = lerp1(t[1], lerp1(t[0], q[0], q[1]), lerp1(t[0], q[2], q[3]));
    if constexpr (D == 2) {
        return lo;
    } else {
        const double hi = lerp1(t[1], lerp1(t[0], q[4], q[5]), lerp1(t[0], q[6], q[7]));
        return lerp1(t[2], lo, hi);
    }
}

template <int D>
__device__ inline void interp_v_slot(const AxisGuess<D>& Gs, const Vel<D>& vg, const VPtrs<D>& V,
                                     const double (&p)[D], const int64_t (&I)[D],
                                     double (&v)[D]) {
#pragma unroll
    for (int c = 0; c < D; ++c) v[c] = interp_slot<D>(Gs, vg, V, c, p, I);
}

// advect_particle (Euler.jl, RK2.jl, RK4.jl) with slot-kernel interpolation
template <int D, int M>
__device__ inline void advect_slot(const AxisGuess<D>& Gs, const Vel<D>& vg, const VPtrs<D>& V,
                                   const RKParams& P, const double (&p0)[D],
                                   const int64_t (&I)[D], double (&out)[D]) {
    double v0[D];
    interp_v_slot<D>(Gs, vg, V, p0, I, v0);
    if constexpr (M == JP_EULER) {
#pragma unroll
        for (int d = 0; d < D; ++d) out[d] = fma(P.dt, v0[d], p0[d]);
    } else if constexpr (M == JP_RK2) {
        double p1[D], v1[D];
#pragma unroll
        for (int d = 0; d < D; ++d) p1[d] = fma(P.adt, v0[d], p0[d]);
        interp_v_slot<D>(Gs, vg, V, p1, I, v1);  // same seed I (RK2.jl:14,20)
        if (P.half) {
#pragma unroll
            for (int d = 0; d < D; ++d) out[d] = fma(P.dt, v1[d], p0[d]);
        } else {
#pragma unroll
            for (int d = 0; d < D; ++d) out[d] = fma(P.dt, fma(P.w1, v1[d], P.w0 * v0[d]), p0[d]);
        }
    } else {
        double q[D], k[D], sum[D];
#pragma unroll
        for (int d = 0; d < D; ++d) q[d] = p0[d] + (P.dt * v0[d]) / 2.0;
        interp_v_slot<D>(Gs, vg, V, q, I, k);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            sum[d] = v0[d] + 2.0 * k[d];
            q[d] = p0[d] + (P.dt * k[d]) / 2.0;
        }
        interp_v_slot<D>(Gs, vg, V, q, I, k);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            sum[d] = sum[d] + 2.0 * k[d];
            q[d] = p0[d] + P.dt * k[d];
        }
        interp_v_slot<D>(Gs, vg, V, q, I, k);
#pragma unroll
        for (int d = 0; d < D; ++d) out[d] = p0[d] + (P.dt * (sum[d] + k[d])) / 6.0;
    }
}

template <int D, int M>
__global__ void __launch_bounds__(kBlock) k_advect_slot(Grid<D> g, Vel<D> vg, VPtrs<D> V,
                                                        Ptr3 coords,
                                                        const uint8_t* __restrict__ index,
                                                        RKParams P) {
    int64_t c, I[D];
    if (!row_cell<D>(g, c, I)) return;
    const int64_t e = c + g.C * (int64_t)blockIdx.y;
    if (!index[e]) return;  // doskip (src/Utils.jl:32-33)
    AxisGuess<D> Gs;
#pragma unroll
    for (int cc = 0; cc < D; ++cc)
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const Axis& A = vg.ax[cc][d];
            Gs.lo[cc][d] = A.x[0];
            Gs.hi[cc][d] = A.x[A.n - 1];
            Gs.i0[cc][d] = A.idx[0];
        }
    double p0[D], out[D];
#pragma unroll
    for (int d = 0; d < D; ++d) p0[d] = coords.p[d][e];
    advect_slot<D, M>(Gs, vg, V, P, p0, I, out);
#pragma unroll
    for (int d = 0; d < D; ++d) coords.p[d][e] = out[d];
}

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

template <int D, int M>
int launch(jp_ctx* ctx, const Grid<D>& G, const Vel<D>& vg, const VPtrs<D>& V, Ptr3 cp,
           const uint8_t* index, const RKParams& P) {
    dim3 grid((unsigned)row_blocks<D>(G), (unsigned)G.K);
    k_advect_slot<D, M><<<grid, kBlock, 0, ctx->stream>>>(G, vg, V, cp, index, P);
    JP_LAUNCHED(ctx);
    return JP_OK;
}

template <int D>
int run(jp_ctx* ctx, const jp_cellgrid* g, double* const* coords, const uint8_t* index,
        const double* const* Vp, const jp_velgrid* vgp, int method, const RKParams& P) {
    Grid<D> G = to_grid<D>(g);
    Vel<D> vg = to_vel<D>(vgp);
    VPtrs<D> V;
    for (int c = 0; c < D; ++c) V.v[c] = Vp[c];
    Ptr3 cp{{nullptr, nullptr, nullptr}};
    for (int d = 0; d < D; ++d) cp.p[d] = coords[d];
    int rc;
    if (method == JP_EULER)
        rc = launch<D, JP_EULER>(ctx, G, vg, V, cp, index, P);
    else if (method == JP_RK2)
        rc = launch<D, JP_RK2>(ctx, G, vg, V, cp, index, P);
    else
        rc = launch<D, JP_RK4>(ctx, G, vg, V, cp, index, P);
    if (rc) return rc;
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

extern "C" int jp_advection(jp_ctx* ctx, const jp_cellgrid* g, double* const* coords,
                            const uint8_t* index, const double* const* V, const jp_velgrid* vg,
                            int method, double alpha, double dt) {
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if (!coords) return fail(ctx, JP_ERR_ARG, "null coords%s");
    for (int d = 0; d < g->ndim; ++d)
        if (!coords[d]) return fail(ctx, JP_ERR_ARG, "null coords pointer%s %lld", "", d);
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
    const double ia = 1.0 / alpha;
    P.w0 = std::fma(-0.5, ia, 1.0);
    P.w1 = 0.5 * ia;
    return g->ndim == 2 ? run<2>(ctx, g, coords, index, V, vg, method, P)
                        : run<3>(ctx, g, coords, index, V, vg, method, P);
}
