// jp_advect.hip -- advection! (src/Particles/Advection/advection.jl:35-91) on
// gfx950, FP64 drop-in, 2D and 3D, Euler / RK2 / RK4.
//
// One workgroup per tile of cells with the tile's velocity and axis windows
// in LDS (see the tiled kernel below); a bracket lookup costs one guess
// (uniform spacing estimate) and one strict test against the stored axis, with
// the reference search for everything else (ties, refined axes).
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

// ---------------------------------------------------------------- tiled kernel
// One workgroup per tile of cells (2D 64 x 4, 3D 16 x 4 x 4), one thread per
// cell, slots walked in chunks of 8 planes in lockstep across the wavefront.
// Per tile, LDS holds for every velocity component c the corner values of the
// node window [t0 - 1, t0 + T + 3) per dimension, and for every axis (c, d) the
// coordinates and inverse spacings of the same node window.  A bracket is the
// uniform-spacing guess confirmed by the strict test x[k] < p < x[k+1] on the
// stored coordinates (a strictly interior point has one bracket, the one the
// reference's bisection returns whatever its seed); ties and misses take the
// seeded bisection on the global axis.  Values read from LDS are the stored
// values, so results are identical to the global-memory path.
template <int D>
__host__ __device__ constexpr int tile_n(int d) {
    return D == 2 ? (d == 0 ? 64 : 4) : (d == 0 ? 16 : 4);
}
template <int D>
__host__ __device__ constexpr int win_n(int d) {
    return tile_n<D>(d) + 4;
}
template <int D>
__host__ __device__ constexpr int win_size() {
    return D == 2 ? win_n<2>(0) * win_n<2>(1) : win_n<3>(0) * win_n<3>(1) * win_n<3>(2);
}
template <int D>
__host__ __device__ constexpr int win_max() {
    return win_n<D>(0);
}

template <int D>
struct TileLDS {
    double V[D][win_size<D>()];
    double x[D][D][win_max<D>()];    // axis (c, d): x[k], k = t0[d] - 1 + local
    double idx[D][D][win_max<D>()];  // 1 / dx[k]
    double lim[D][D][2];             // first and last coordinate of axis (c, d)
    double g0[D][D][2];              // x[0] and inv_dx[0]: the uniform guess
};

// ties and guesses outside the window: the seeded search on the global axis
__device__ inline void bracket_global(double p, const Axis& A, int64_t seed, int64_t& k,
                                           double& xk, double& ik) {
    k = find_cell(p, A, seed);
    xk = A.x[k];
    ik = A.idx[k];
}

// bracket of p on axis (c, d): global node k, its coordinate and inverse spacing
template <int D>
__device__ inline void bracket_tile(const TileLDS<D>& L, const Axis& A, int c, int d, int base,
                                    double p, int64_t seed, int64_t& k, double& xk, double& ik,
                                    bool& inw) {
    const int64_t last = A.n - 2;
    int64_t gk = (int64_t)((p - L.g0[c][d][0]) * L.g0[c][d][1]);
    gk = gk < 0 ? 0 : (gk > last ? last : gk);
    const int64_t kl = gk - base;
    if (kl >= 0 && kl + 1 < win_n<D>(d)) {
        const double a = L.x[c][d][kl], b = L.x[c][d][kl + 1];
        if (a < p && p < b) {
            k = gk;
            xk = a;
            ik = L.idx[c][d][kl];
            inw = true;
            return;
        }
    }
    bracket_global(p, A, seed, k, xk, ik);
    const int64_t kk = k - base;
    inw = kk >= 0 && kk + 1 < win_n<D>(d);
}

template <int D>
__device__ inline double interp_tile(const TileLDS<D>& L, const Vel<D>& vg, const VPtrs<D>& V,
                                     int c, const int (&base)[D], const double (&p)[D],
                                     const int64_t (&I)[D]) {
    bool in = true;
#pragma unroll
    for (int d = 0; d < D; ++d) in = in && (L.lim[c][d][0] <= p[d] && p[d] <= L.lim[c][d][1]);
    if (!in) return __builtin_inf();
    int64_t k[D];
    double t[D];
    bool inw = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        double xk, ik;
        bool w;
        bracket_tile<D>(L, vg.ax[c][d], c, d, base[d], p[d], I[d], k[d], xk, ik, w);
        t[d] = (p[d] - xk) * ik;
        inw = inw && w;
    }
    double q[1 << D];
    if (inw) {
        constexpr int W0 = win_n<D>(0), W1 = win_n<D>(1);
        int b = (int)(k[0] - base[0]) + W0 * (int)(k[1] - base[1]);
        if constexpr (D == 3) b += W0 * W1 * (int)(k[2] - base[2]);
#pragma unroll
        for (int r = 0; r < (1 << D); ++r) {
            const int o = (r & 1) + ((r >> 1) & 1) * W0 + (D == 3 ? ((r >> 2) & 1) * W0 * W1 : 0);
            q[r] = L.V[c][b + o];
        }
    } else {
        const double* __restrict__ F = V.v[c];
        const int64_t n0 = vg.ax[c][0].n, n1 = vg.ax[c][1].n;
        int64_t b = k[0] + n0 * k[1];
        if constexpr (D == 3) b += n0 * n1 * k[2];
#pragma unroll
        for (int r = 0; r < (1 << D); ++r) {
            const int64_t o = (r & 1) + ((r >> 1) & 1) * n0 + (D == 3 ? ((r >> 2) & 1) * n0 * n1 : 0);
            q[r] = F[b + o];
        }
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
__device__ inline void interp_v_tile(const TileLDS<D>& L, const Vel<D>& vg, const VPtrs<D>& V,
                                     const int (&base)[D], const double (&p)[D],
                                     const int64_t (&I)[D], double (&v)[D]) {
#pragma unroll
    for (int c = 0; c < D; ++c) v[c] = interp_tile<D>(L, vg, V, c, base, p, I);
}

// advect_particle with the tile's LDS: Euler (Euler.jl:1-20), RK2 (RK2.jl:1-26,
// stages src/Advection/RK2.jl:1-44), RK4 (RK4.jl:1-19, no muladd)
template <int D, int M>
__device__ inline void advect_tile(const TileLDS<D>& L, const Vel<D>& vg, const VPtrs<D>& V,
                                   const RKParams& P, const int (&base)[D],
                                   const double (&p0)[D], const int64_t (&I)[D],
                                   double (&out)[D]) {
    double v0[D];
    interp_v_tile<D>(L, vg, V, base, p0, I, v0);
    if constexpr (M == JP_EULER) {
#pragma unroll
        for (int d = 0; d < D; ++d) out[d] = fma(P.dt, v0[d], p0[d]);
    } else if constexpr (M == JP_RK2) {
        double p1[D], v1[D];
#pragma unroll
        for (int d = 0; d < D; ++d) p1[d] = fma(P.adt, v0[d], p0[d]);
        interp_v_tile<D>(L, vg, V, base, p1, I, v1);  // same seed I (RK2.jl:14,20)
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
        interp_v_tile<D>(L, vg, V, base, q, I, k2);
#pragma unroll
        for (int d = 0; d < D; ++d) q[d] = p0[d] + (P.dt * k2[d]) / 2.0;
        interp_v_tile<D>(L, vg, V, base, q, I, k3);
#pragma unroll
        for (int d = 0; d < D; ++d) q[d] = p0[d] + P.dt * k3[d];
        interp_v_tile<D>(L, vg, V, base, q, I, k4);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const double sum = ((v0[d] + 2.0 * k2[d]) + 2.0 * k3[d]) + k4[d];
            out[d] = p0[d] + (P.dt * sum) / 6.0;
        }
    }
}

// advection_kernel! (advection.jl:66-91): doskip, advect, write back
template <int D, int M>
__global__ void __launch_bounds__(kBlock) k_advect(Grid<D> g, Vel<D> vg, VPtrs<D> V, Ptr3 coords,
                                                   const uint8_t* __restrict__ index, RKParams P) {
    __shared__ TileLDS<D> L;
    int t0[D], base[D];
    {
        int64_t b = blockIdx.x;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int64_t nt = (g.ni[d] + tile_n<D>(d) - 1) / tile_n<D>(d);
            t0[d] = (int)(b % nt) * tile_n<D>(d);
            b /= nt;
            base[d] = t0[d] - 1;
        }
    }
    const int tid = threadIdx.x;
    // window fills: every load issued before the LDS stores
    constexpr int NI = (win_size<D>() + kBlock - 1) / kBlock;
    double wv[D][NI];
#pragma unroll
    for (int c = 0; c < D; ++c) {
        const double* __restrict__ F = V.v[c];
#pragma unroll
        for (int it = 0; it < NI; ++it) {
            const int i = tid + it * kBlock;
            int r = i;
            int64_t node = 0, stride = 1;
            bool ok = i < win_size<D>();
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int64_t k = base[d] + r % win_n<D>(d);
                r /= win_n<D>(d);
                ok = ok && k >= 0 && k < vg.ax[c][d].n;
                node += k * stride;
                stride *= vg.ax[c][d].n;
            }
            wv[c][it] = ok ? F[ok ? node : 0] : 0.0;
        }
    }
    constexpr int NA = D * D * win_max<D>();
    constexpr int NJ = (NA + kBlock - 1) / kBlock;
    double ax[NJ], ai[NJ];
#pragma unroll
    for (int it = 0; it < NJ; ++it) {
        const int i = tid + it * kBlock;
        const int c = i / (D * win_max<D>());
        const int d = (i / win_max<D>()) % D;
        const int j = i % win_max<D>();
        const bool in = i < NA;
        const Axis& A = vg.ax[in ? c : 0][in ? d : 0];
        const int64_t k = base[in ? d : 0] + j;
        const bool okx = in && j < win_n<D>(d) && k >= 0 && k < A.n;
        const bool oki = in && j < win_n<D>(d) && k >= 0 && k < A.n - 1;
        ax[it] = okx ? A.x[okx ? k : 0] : __builtin_nan("");
        ai[it] = oki ? A.idx[oki ? k : 0] : __builtin_nan("");
    }
#pragma unroll
    for (int c = 0; c < D; ++c)
#pragma unroll
        for (int it = 0; it < NI; ++it)
            if (tid + it * kBlock < win_size<D>()) L.V[c][tid + it * kBlock] = wv[c][it];
#pragma unroll
    for (int it = 0; it < NJ; ++it) {
        const int i = tid + it * kBlock;
        if (i < NA) {
            const int c = i / (D * win_max<D>());
            const int d = (i / win_max<D>()) % D;
            const int j = i % win_max<D>();
            L.x[c][d][j] = ax[it];
            L.idx[c][d][j] = ai[it];
        }
    }
    if (tid < D * D) {
        const int c = tid / D, d = tid % D;
        const Axis& A = vg.ax[c][d];
        L.lim[c][d][0] = A.x[0];
        L.lim[c][d][1] = A.x[A.n - 1];
        L.g0[c][d][0] = A.x[0];
        L.g0[c][d][1] = A.idx[0];
    }
    __syncthreads();
    int64_t I[D];
    bool in = true;
    {
        int r = tid;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            I[d] = t0[d] + r % tile_n<D>(d);
            r /= tile_n<D>(d);
            in = in && I[d] < g.ni[d];
        }
    }
    if (!in) return;
    const int64_t c = lin<D>(g.ni, I);
    const int64_t C = g.C;
    const uint64_t act = active_slots(index, c, C, g.K);
    constexpr int B = 8;
    for (int k0 = 0; k0 < g.K; k0 += B) {
        const uint32_t ab = (uint32_t)(act >> k0) & 0xFFu;
        if (!ab) continue;
        double pb[B][D];
#pragma unroll
        for (int j = 0; j < B; ++j)
#pragma unroll
            for (int d = 0; d < D; ++d)
                pb[j][d] = ((ab >> j) & 1) ? coords.p[d][c + C * (k0 + j)] : 0.0;
        // one particle at a time (the chunk's loads are already in flight)
        uint32_t todo = ab;
        while (todo) {
            const int j = __builtin_ctz(todo);
            todo &= todo - 1;
            double p0[D], out[D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                double v = pb[0][d];
#pragma unroll
                for (int q = 1; q < B; ++q) v = j == q ? pb[q][d] : v;
                p0[d] = v;
            }
            advect_tile<D, M>(L, vg, V, P, base, p0, I, out);
#pragma unroll
            for (int d = 0; d < D; ++d) coords.p[d][c + C * (k0 + j)] = out[d];
        }
    }
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
    int64_t tiles = 1;
    for (int d = 0; d < D; ++d) tiles *= (G.ni[d] + tile_n<D>(d) - 1) / tile_n<D>(d);
    k_advect<D, M><<<(unsigned)tiles, kBlock, 0, ctx->stream>>>(G, vg, V, cp, index, P);
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
