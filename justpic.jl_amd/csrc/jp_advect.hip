// jp_advect.hip -- advection! (src/Particles/Advection/advection.jl:35-91),
// advection_MQS! / advection_LinP! (advection_MQS.jl, advection_LinP.jl) and
// semilagrangian_advection! with the same three velocity interpolations
// (backtracking.jl, backtracking_MQS.jl, backtracking_LinP.jl) on gfx950, FP64
// drop-in, 2D and 3D, Euler / RK2 / RK4.
//
// The linear scheme (advection!): one workgroup per tile of cells with each
// velocity component's window of nodes, and its axis windows, staged in LDS;
// one thread per cell walks the cell's slots in lockstep chunks of planes
// (k_advect_tile).  A stage point outside a window sends the particle's whole
// step to the slot path (inline in 2D, a flagged pass in 3D).  The
// slot path (k_advect_slot, and the MQS / LinP schemes): one thread per
// slot, the velocity gathered from global memory.  A bracket lookup costs one
// guess (uniform spacing estimate) and one strict test against the stored
// axis, with the reference search for everything else (ties, refined axes).
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

// velocity interpolation schemes
constexpr int kLinear = 0;  // interp_velocity2particle
constexpr int kMQS = 1;     // interp_velocity2particle_MQS (advection_MQS.jl:102-124)
constexpr int kLinP = 2;    // interp_velocity2particle_LinP (advection_LinP.jl:104-131)

// column-major view of one velocity component (k = 0 in 2D)
struct FView {
    const double* __restrict__ F;
    int64_t n0, n1;
    __device__ double at(int64_t i, int64_t j, int64_t k) const { return F[i + n0 * (j + n1 * k)]; }
};

// half * (t - half)^2 * (muladd(-2, v1, v0) + v2) (src/Interpolations/MQS.jl:13)
__device__ inline double mqs_corr(double t, double v0, double v1, double v2) {
    const double e = t - 0.5;
    return (0.5 * (e * e)) * (fma(-2.0, v1, v0) + v2);
}

// MQS along x of the corners (v00, v10, v01, v11) of one plane: each row's
// lerp along x corrected with the second difference through the node left of
// the cell (ta < 1/2) or right of it, then the lerp along the plane's second
// axis (MQS.jl:2-27, 73-98, 126-152).  Rows 0 and 1 take their extra node
// from (i-1 | i+2, jb0, kb0) and (i-1 | i+2, jb1, kb1).
__device__ inline double mqs_x4(const FView& F, double v0, double v1, double v2, double v3,
                                double ta, double tb, int64_t i, int64_t jb0, int64_t kb0,
                                int64_t jb1, int64_t kb1) {
    const double lb = lerp1(ta, v0, v1), lt = lerp1(ta, v2, v3);
    const bool left = ta < 0.5;
    const double cb = left ? mqs_corr(ta, F.at(i - 1, jb0, kb0), v0, v1)
                           : mqs_corr(ta, v0, v1, F.at(i + 2, jb0, kb0));
    const double ct = left ? mqs_corr(ta, F.at(i - 1, jb1, kb1), v2, v3)
                           : mqs_corr(ta, v2, v3, F.at(i + 2, jb1, kb1));
    return lerp1(tb, lb + cb, lt + ct);
}

// MQS along y (MQS.jl:29-55, 100-124): columns (v00, v01), (v10, v11) lerped
// along y with the correction through the node below or above, then along x.
__device__ inline double mqs_y4(const FView& F, double v0, double v1, double v2, double v3,
                                double t1, double t2, int64_t i, int64_t j, int64_t k) {
    const double ll = lerp1(t2, v0, v2), lr = lerp1(t2, v1, v3);
    const bool below = t2 < 0.5;
    const double cl = below ? mqs_corr(t2, F.at(i, j - 1, k), v0, v2)
                            : mqs_corr(t2, v0, v2, F.at(i, j + 2, k));
    const double cr = below ? mqs_corr(t2, F.at(i + 1, j - 1, k), v1, v3)
                            : mqs_corr(t2, v1, v3, F.at(i + 1, j + 2, k));
    return lerp1(t1, ll + cl, lr + cr);
}

// MQS of component c at corner (i, j[, k]) with corner values q.  3D: x and y
// combine the plane-k and plane-k+1 results along z, both with extra nodes of
// plane k (MQS.jl:58-64 pass the same i, j, k); z combines two MQS-x
// evaluations of the (x, z) planes of rows j and j+1 along y, both with extra
// nodes of row j (MQS.jl:67-71, 126-152).
template <int D>
__device__ inline double mqs(const FView& F, int c, const int64_t (&k0)[D], const double* q,
                             const double (&t)[D]) {
    const int64_t i = k0[0], j = k0[1];
    if constexpr (D == 2) {
        return c == 0 ? mqs_x4(F, q[0], q[1], q[2], q[3], t[0], t[1], i, j, 0, j + 1, 0)
                      : mqs_y4(F, q[0], q[1], q[2], q[3], t[0], t[1], i, j, 0);
    } else {
        const int64_t k = k0[2];
        if (c == 0) {
            const double bot = mqs_x4(F, q[0], q[1], q[2], q[3], t[0], t[1], i, j, k, j + 1, k);
            const double top = mqs_x4(F, q[4], q[5], q[6], q[7], t[0], t[1], i, j, k, j + 1, k);
            return lerp1(t[2], bot, top);
        }
        if (c == 1) {
            const double bot = mqs_y4(F, q[0], q[1], q[2], q[3], t[0], t[1], i, j, k);
            const double top = mqs_y4(F, q[4], q[5], q[6], q[7], t[0], t[1], i, j, k);
            return lerp1(t[2], bot, top);
        }
        const double front = mqs_x4(F, q[0], q[1], q[4], q[5], t[0], t[2], i, j, k, j, k + 1);
        const double back = mqs_x4(F, q[2], q[3], q[6], q[7], t[0], t[2], i, j, k, j, k + 1);
        return lerp1(t[1], front, back);
    }
}

// LinP's pressure-point value of component c (advection_LinP.jl:208-391):
// the corner index moves by one along c when p is past the cell's middle
// (offset_LinP); each row of three nodes along c around it (index clamped to
// the array) is averaged pairwise as (a + b)/2 onto the two pressure points;
// the rows run over the other axes (augment_offset: x and y components in
// axis order, the z component with y before x); y and z components swap the
// middle pairs (swap_F); t is measured from the corner moved by +-dx_c/2
// (correct_xci_to_pressure_point).
template <int D>
__device__ inline double linp(const FView& F, const int64_t (&nn)[D], int c,
                              const int64_t (&k0)[D], const double (&xc)[D],
                              const double (&dxi)[D], const double (&idx)[D],
                              const double (&p)[D]) {
    int64_t b[D];
#pragma unroll
    for (int d = 0; d < D; ++d) b[d] = k0[d] + ((d == c && p[d] > xc[d] + dxi[d] / 2.0) ? 1 : 0);
    constexpr int nrow = D == 2 ? 2 : 4;
    double av[2 * nrow];
#pragma unroll
    for (int r = 0; r < nrow; ++r) {
        double val[3];
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            int64_t id[3] = {0, 0, 0};
            int other = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                int64_t o;
                if (d == c) {
                    o = e - 1;
                } else {
                    const int bit = (D == 3 && c == 2) ? 1 - other : other;
                    o = (r >> bit) & 1;
                    ++other;
                }
                const int64_t v = b[d] + o;
                id[d] = v < 0 ? 0 : (v > nn[d] - 1 ? nn[d] - 1 : v);
            }
            val[e] = F.at(id[0], id[1], id[2]);
        }
        av[2 * r] = (val[0] + val[1]) / 2.0;
        av[2 * r + 1] = (val[2] + val[1]) / 2.0;
    }
    double w[2 * nrow];
#pragma unroll
    for (int m = 0; m < 2 * nrow; ++m) {
        // swap_F: (1, 3, 2, 4[, 5, 7, 6, 8]) for the y and z components
        const int src = (c == 0) ? m : ((m & ~3) | ((m & 1) << 1) | ((m >> 1) & 1));
        w[m] = av[src];
    }
    double t[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        double xp = xc[d];
        if (d == c) {
            const double off = 1.0 - 2.0 * (double)(p[d] < xc[d] + dxi[d] / 2.0);
            xp = xc[d] + (off * dxi[d]) / 2.0;
        }
        t[d] = (p[d] - xp) * idx[d];
    }
    const double lo = lerp1(t[1], lerp1(t[0], w[0], w[1]), lerp1(t[0], w[2], w[3]));
    if constexpr (D == 2) {
        return lo;
    } else {
        const double hi = lerp1(t[1], lerp1(t[0], w[4], w[5]), lerp1(t[0], w[6], w[7]));
        return lerp1(t[2], lo, hi);
    }
}

// Component c at p with scheme S.  LIM: advection's limit check (+Inf outside
// the component's grid); without it (semi-Lagrangian) a point outside the
// grid gives NaN, which the caller reports -- the reference's bisection does
// not terminate there.  MQS and LinP apply where the corner index is interior
// along every axis (1 < idx < size - 1, 1-based), else the linear value.
template <int D, int S, bool LIM>
__device__ inline double interp_slot(const AxisGuess<D>& Gs, const Vel<D>& vg,
                                     const VPtrs<D>& V, int c, const double (&p)[D],
                                     const int64_t (&I)[D]) {
    bool in = true;
#pragma unroll
    for (int d = 0; d < D; ++d) in = in && (Gs.lo[c][d] <= p[d] && p[d] <= Gs.hi[c][d]);
    if (!in) return LIM ? __builtin_inf() : __builtin_nan("");
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
    double vl = lerp1(t[1], lerp1(t[0], q[0], q[1]), lerp1(t[0], q[2], q[3]));
    if constexpr (D == 3) {
        const double hi = lerp1(t[1], lerp1(t[0], q[4], q[5]), lerp1(t[0], q[6], q[7]));
        vl = lerp1(t[2], vl, hi);
    }
    if constexpr (S == kLinear) {
        return vl;
    } else {
        int64_t nn[D];
        bool interior = true;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            nn[d] = vg.ax[c][d].n;
            interior = interior && k[d] >= 1 && k[d] <= nn[d] - 3;
        }
        if (!interior) return vl;
        const FView Fv{F, n0, n1};
        if constexpr (S == kMQS) {
            return mqs<D>(Fv, c, k, q, t);
        } else {
            double xc[D], dxi[D], idx[D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const Axis& A = vg.ax[c][d];
                xc[d] = A.x[k[d]];
                dxi[d] = A.dx[k[d]];
                idx[d] = A.idx[k[d]];
            }
            const double vp = linp<D>(Fv, nn, c, k, xc, dxi, idx, p);
            const double A = 2.0 / 3.0;
            return A * vl + (1.0 - A) * vp;
        }
    }
}

template <int D, int S, bool LIM>
__device__ inline void interp_v_slot(const AxisGuess<D>& Gs, const Vel<D>& vg, const VPtrs<D>& V,
                                     const double (&p)[D], const int64_t (&I)[D],
                                     double (&v)[D]) {
#pragma unroll
    for (int c = 0; c < D; ++c) v[c] = interp_slot<D, S, LIM>(Gs, vg, V, c, p, I);
}

// advect_particle (Euler.jl, RK2.jl, RK4.jl) for any velocity interpolation
// interp(p, v); the semi-Lagrangian backtracking (advect_particle_SML,
// RK2.jl:53-107, RK4.jl:44-87) is the same with P.dt = -dt and
// P.adt = (-alpha) dt
template <int D, int M, class Interp>
__device__ inline void advect_rk(const RKParams& P, const double (&p0)[D], double (&out)[D],
                                 Interp&& interp) {
    double v0[D];
    interp(p0, v0);
    if constexpr (M == JP_EULER) {
#pragma unroll
        for (int d = 0; d < D; ++d) out[d] = fma(P.dt, v0[d], p0[d]);
    } else if constexpr (M == JP_RK2) {
        double p1[D], v1[D];
#pragma unroll
        for (int d = 0; d < D; ++d) p1[d] = fma(P.adt, v0[d], p0[d]);
        interp(p1, v1);  // same seed I (RK2.jl:14,20)
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
        interp(q, k);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            sum[d] = v0[d] + 2.0 * k[d];
            q[d] = p0[d] + (P.dt * k[d]) / 2.0;
        }
        interp(q, k);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            sum[d] = sum[d] + 2.0 * k[d];
            q[d] = p0[d] + P.dt * k[d];
        }
        interp(q, k);
#pragma unroll
        for (int d = 0; d < D; ++d) out[d] = p0[d] + (P.dt * (sum[d] + k[d])) / 6.0;
    }
}

// with the slot-kernel interpolation (velocity gathered from global memory)
template <int D, int M, int S = kLinear, bool LIM = true>
__device__ inline void advect_slot(const AxisGuess<D>& Gs, const Vel<D>& vg, const VPtrs<D>& V,
                                   const RKParams& P, const double (&p0)[D],
                                   const int64_t (&I)[D], double (&out)[D]) {
    advect_rk<D, M>(P, p0, out, [&](const double (&p)[D], double (&v)[D]) {
        interp_v_slot<D, S, LIM>(Gs, vg, V, p, I, v);
    });
}

template <int D>
__device__ inline void axis_guess(const Vel<D>& vg, AxisGuess<D>& Gs) {
#pragma unroll
    for (int cc = 0; cc < D; ++cc)
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const Axis& A = vg.ax[cc][d];
            Gs.lo[cc][d] = A.x[0];
            Gs.hi[cc][d] = A.x[A.n - 1];
            Gs.i0[cc][d] = A.idx[0];
        }
}

template <int D, int M, int S>
__global__ void __launch_bounds__(kBlock) k_advect_slot(Grid<D> g, Vel<D> vg, VPtrs<D> V,
                                                        Ptr3 coords,
                                                        const uint8_t* __restrict__ index,
                                                        RKParams P) {
    int64_t c, I[D];
    if (!row_cell<D>(g, c, I)) return;
    const int64_t e = c + g.C * (int64_t)blockIdx.y;
    if (!index[e]) return;  // doskip (src/Utils.jl:32-33)
    AxisGuess<D> Gs;
    axis_guess<D>(vg, Gs);
    double p0[D], out[D];
#pragma unroll
    for (int d = 0; d < D; ++d) p0[d] = coords.p[d][e];
    advect_slot<D, M, S>(Gs, vg, V, P, p0, I, out);
#pragma unroll
    for (int d = 0; d < D; ++d) coords.p[d][e] = out[d];
}

// The linear scheme with the velocity staged in LDS: one workgroup per tile
// of cells (2D 64x4, 3D 16x4x4), one thread per cell walking the cell's slot
// planes in lockstep chunks (a wavefront covers whole 128-byte runs of a
// plane).  Each
// component's window holds the nodes from two below the tile's first bracket
// to kWinPad - 2 beyond its end, which covers the RK stage points of a step of
// up to one cell; its axis coordinates and inverse spacings are staged too.
// A bracket, and the corner values, come from the window when they lie in it
// and from global memory otherwise, with the same tests and the same values
// as the slot kernel, so the results are identical.
template <int D>
struct F64Tile;
template <>
struct F64Tile<2> {
    static constexpr int n[3] = {64, 4, 1};
};
template <>
struct F64Tile<3> {
    static constexpr int n[3] = {16, 4, 4};
};
constexpr int kWinPad = 6;

template <int D>
struct F64Win {
    static constexpr int w(int d) { return d < D ? F64Tile<D>::n[d] + kWinPad : 1; }
    static constexpr int size() { return w(0) * w(1) * w(2); }
    static constexpr int axis() { return w(0) > w(1) ? (w(0) > w(2) ? w(0) : w(2)) : (w(1) > w(2) ? w(1) : w(2)); }
};

// window entries past the end of an axis hold NaN (no bracket test passes)
template <int D>
struct TileLds {
    double v[D][F64Win<D>::size()];
    double x[D][D][F64Win<D>::axis()];
    double idx[D][D][F64Win<D>::axis()];
};

// Component c at p from the window, if every axis brackets p strictly inside
// it (that bracket is the axis's only one, the one the slot kernel finds, and
// it lies inside the component's grid); false otherwise.
template <int D>
__device__ inline bool interp_tile_f64(const TileLds<D>& L, int c, const double (&p)[D],
                                       double& v) {
    int w[D];
    double t[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const double g = (p[d] - L.x[c][d][0]) * L.idx[c][d][0];
        int k = g >= 0.0 ? (g < (double)(F64Win<D>::w(d) - 2) ? (int)g : F64Win<D>::w(d) - 2) : 0;
        const double xl = L.x[c][d][k], xr = L.x[c][d][k + 1];
        if (!(xl < p[d] && p[d] < xr)) return false;
        w[d] = k;
        t[d] = (p[d] - xl) * L.idx[c][d][k];
    }
    const int b = w[0] + F64Win<D>::w(0) * (w[1] + (D == 3 ? F64Win<D>::w(1) * w[D - 1] : 0));
    double q[1 << D];
#pragma unroll
    for (int r = 0; r < (1 << D); ++r) {
        const int o = (r & 1) + ((r >> 1) & 1) * F64Win<D>::w(0) +
                      (D == 3 ? ((r >> 2) & 1) * F64Win<D>::w(0) * F64Win<D>::w(1) : 0);
        q[r] = L.v[c][b + o];
    }
    double vl = lerp1(t[1], lerp1(t[0], q[0], q[1]), lerp1(t[0], q[2], q[3]));
    if constexpr (D == 3) {
        const double hi = lerp1(t[1], lerp1(t[0], q[4], q[5]), lerp1(t[0], q[6], q[7]));
        vl = lerp1(t[2], vl, hi);
    }
    v = vl;
    return true;
}

// B: slot planes per chunk; FLAGS: particles whose step leaves the windows
// are flagged (flags[slot], written for every slot of the tile's planes) for
// k_advect_flagged instead of being advected inline, and the workgroup
// records whether it flagged any (any[workgroup])
template <int D, int M, int B, bool FLAGS>
__global__ void __launch_bounds__(kBlock) k_advect_tile(Grid<D> g, Vel<D> vg, VPtrs<D> V,
                                                        Ptr3 coords,
                                                        const uint8_t* __restrict__ index,
                                                        uint8_t* __restrict__ flags,
                                                        uint8_t* __restrict__ any, RKParams P,
                                                        int span) {
    using TL = F64Tile<D>;
    using WN = F64Win<D>;
    __shared__ TileLds<D> L;
    int64_t t0[D];
    {
        int64_t b = blockIdx.x;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int64_t nt = (g.ni[d] + TL::n[d] - 1) / TL::n[d];
            t0[d] = (b % nt) * TL::n[d];
            b /= nt;
        }
    }
    const int kb = blockIdx.y * span;
    const int ke = kb + span < g.K ? kb + span : g.K;
    const int tid = threadIdx.x;
    // window origins: two nodes below the guessed bracket of the tile's first
    // vertex (only where the window sits depends on the guess, not a result)
    int64_t wlo[D][D];
#pragma unroll
    for (int c = 0; c < D; ++c)
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const Axis& A = vg.ax[c][d];
            const double x = g.xv[d].x[t0[d]];
            const int64_t last = A.n - 2;
            int64_t gk = (int64_t)((x - A.x[0]) * A.idx[0]);
            gk = gk < 0 ? 0 : (gk > last ? last : gk);
            wlo[c][d] = gk - 2 < 0 ? 0 : gk - 2;
        }
    const double qnan = __builtin_nan("");
#pragma unroll
    for (int c = 0; c < D; ++c) {
        const double* __restrict__ F = V.v[c];
        for (int i = tid; i < WN::size(); i += kBlock) {
            int r = i;
            int64_t node = 0, stride = 1;
            bool ok = true;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int64_t kk = wlo[c][d] + r % WN::w(d);
                r /= WN::w(d);
                ok = ok && kk < vg.ax[c][d].n;
                node += kk * stride;
                stride *= vg.ax[c][d].n;
            }
            L.v[c][i] = ok ? F[node] : qnan;
        }
#pragma unroll
        for (int d = 0; d < D; ++d)
            for (int w = tid; w < WN::w(d); w += kBlock) {
                const Axis& A = vg.ax[c][d];
                const int64_t kk = wlo[c][d] + w;
                L.x[c][d][w] = kk < A.n ? A.x[kk] : qnan;
                L.idx[c][d][w] = kk < A.n - 1 ? A.idx[kk] : qnan;
            }
    }
    __syncthreads();
    // the cell of thread t of the tile and whether it is in the grid
    auto cell_of = [&](int t, int64_t (&J)[D]) {
        bool in = true;
        int r = t;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            J[d] = t0[d] + r % TL::n[d];
            r /= TL::n[d];
            in = in && J[d] < g.ni[d];
        }
        return in;
    };
    int64_t I[D];
    const bool inside = cell_of(tid, I);
    const int64_t c = inside ? lin<D>(g.ni, I) : 0;
    const int64_t C = g.C;
    // the step from the windows; if any stage point leaves a window (or sits
    // on a node), the whole step again on the slot kernel's path: inline
    // (2D), or in the flagged pass that follows (3D, where inlining it
    // doubles the registers)
    bool win = true;
    auto interp = [&](const double (&p)[D], double (&v)[D]) {
#pragma unroll
        for (int cc = 0; cc < D; ++cc) win = interp_tile_f64<D>(L, cc, p, v[cc]) && win;
    };
    // a thread outside the grid owns no particle but still takes work
    const uint64_t act = !inside ? 0ull
                                 : ((kb == 0 && ke == g.K) ? active_slots(index, c, C, g.K)
                                                           : active_slots_in(index, c, C, kb, ke));
    // Chunks of B slot planes.  The wavefront's active particles of a chunk
    // are listed (plane-major, then lane) and dealt to its lanes in turn, so
    // that every lane advects about the same number of particles; consecutive
    // entries are neighbouring cells of one plane, so the loads and stores
    // stay coalesced.  The list is written and read by the same wavefront
    // (its LDS operations complete in order).
    __shared__ uint16_t s_items[kBlock / 64][B * 64];
    uint16_t* items = s_items[tid >> 6];
    const int lane = tid & 63, wbase = tid & ~63;
    bool missed = false;
    for (int k0 = kb; k0 < ke; k0 += B) {
        if constexpr (FLAGS) {  // every slot of the chunk: no miss yet
#pragma unroll
            for (int j = 0; j < B; ++j)
                if (inside && k0 + j < ke) flags[c + C * (k0 + j)] = 0;
        }
        int n = 0;
#pragma unroll
        for (int j = 0; j < B; ++j) {
            const bool a = k0 + j < 64 && ((act >> (k0 + j)) & 1);
            const uint64_t bal = __ballot(a);
            const int below = (int)__builtin_amdgcn_mbcnt_hi(
                (uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
            if (a) items[n + below] = (uint16_t)((j << 6) | lane);
            n += __popcll(bal);
        }
        if (!n) continue;
        // this lane's entries lane, lane + 64, ...: positions loaded first
        double pq[B][D];
        int dq[B];
#pragma unroll
        for (int q = 0; q < B; ++q) {
            const int item = lane + 64 * q;
            dq[q] = item < n ? items[item] : -1;
            int64_t J[D];
            cell_of(wbase + (dq[q] & 63), J);
            const int64_t e = lin<D>(g.ni, J) + C * (k0 + (dq[q] >> 6));
#pragma unroll
            for (int d = 0; d < D; ++d) pq[q][d] = dq[q] >= 0 ? coords.p[d][e] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < B; ++q) {
            if (dq[q] < 0) break;
            int64_t J[D];
            cell_of(wbase + (dq[q] & 63), J);
            const int64_t e = lin<D>(g.ni, J) + C * (k0 + (dq[q] >> 6));
            double out[D];
            win = true;
            advect_rk<D, M>(P, pq[q], out, interp);
            if (!win) {
                if constexpr (FLAGS) {
                    flags[e] = 1;  // the flagged pass advects it
                    missed = true;
                    continue;
                } else {
                    AxisGuess<D> Gs;
                    axis_guess<D>(vg, Gs);
                    advect_slot<D, M>(Gs, vg, V, P, pq[q], J, out);
                }
            }
#pragma unroll
            for (int d = 0; d < D; ++d) coords.p[d][e] = out[d];
        }
    }
    if constexpr (FLAGS) {
        const int m = __syncthreads_or(missed ? 1 : 0);
        if (tid == 0) any[blockIdx.x + (int64_t)gridDim.x * blockIdx.y] = m ? 1 : 0;
    }
}

// the flagged particles of the tiled kernel: the same grid of tiles and slot
// groups, a workgroup returning at once unless its tile flagged some; then
// one thread per cell of the tile walks its slots, advecting the flagged ones
// as k_advect_slot does
template <int D, int M>
__global__ void __launch_bounds__(kBlock) k_advect_flagged(Grid<D> g, Vel<D> vg, VPtrs<D> V,
                                                           Ptr3 coords,
                                                           const uint8_t* __restrict__ flags,
                                                           const uint8_t* __restrict__ any,
                                                           RKParams P, int span) {
    using TL = F64Tile<D>;
    if (!any[blockIdx.x + (int64_t)gridDim.x * blockIdx.y]) return;
    int64_t I[D];
    {
        int64_t b = blockIdx.x;
        int r = threadIdx.x;
        bool in = true;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int64_t nt = (g.ni[d] + TL::n[d] - 1) / TL::n[d];
            I[d] = (b % nt) * TL::n[d] + r % TL::n[d];
            b /= nt;
            r /= TL::n[d];
            in = in && I[d] < g.ni[d];
        }
        if (!in) return;
    }
    const int kb = blockIdx.y * span;
    const int ke = kb + span < g.K ? kb + span : g.K;
    const int64_t c = lin<D>(g.ni, I);
    AxisGuess<D> Gs;
    axis_guess<D>(vg, Gs);
    for (int k = kb; k < ke; ++k) {
        const int64_t e = c + g.C * (int64_t)k;
        if (!flags[e]) continue;
        double p0[D], out[D];
#pragma unroll
        for (int d = 0; d < D; ++d) p0[d] = coords.p[d][e];
        advect_slot<D, M>(Gs, vg, V, P, p0, I, out);
#pragma unroll
        for (int d = 0; d < D; ++d) coords.p[d][e] = out[d];
    }
}

// semilagrangian_advection! (backtracking.jl:15-101 and the MQS / LinP forms):
// one thread per interior node I of the nodal grid (x fastest).  The node is
// traced back (velocity brackets seeded at I), the point bracketed on the
// vertex axes seeded at I, and every field F[f][I] = lerp of F0[f]'s corners
// there, t = (p - x[Ib]) * inv(dx[Ib]) -- dx[I] with `at_node` (the linear
// NTuple method, backtracking.jl:79).  A point that left a velocity or vertex
// axis sets range[0] and range[1] and leaves the node unchanged.
template <int D>
struct NodeGrid {
    int64_t n[D];
    int64_t N;  // interior nodes
    Axis x[D];
};
struct SmlFields {
    double* F[JP_MAX_ARGS];
    const double* F0[JP_MAX_ARGS];
    int n;
};

template <int D, int M, int S>
__global__ void __launch_bounds__(kBlock) k_sml(NodeGrid<D> G, Vel<D> vg, VPtrs<D> V,
                                                SmlFields Fs, RKParams P, int at_node,
                                                int* __restrict__ range) {
    const int64_t tid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (tid >= G.N) return;
    int64_t I[D], r = tid;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const int64_t m = G.n[d] - 2;
        I[d] = r % m + 1;
        r /= m;
    }
    AxisGuess<D> Gs;
    axis_guess<D>(vg, Gs);
    double p[D], pb[D];
#pragma unroll
    for (int d = 0; d < D; ++d) p[d] = G.x[d].x[I[d]];
    advect_slot<D, M, S, false>(Gs, vg, V, P, p, I, pb);
    int64_t Ib[D];
    double t[D];
    bool in = true;
#pragma unroll
    for (int d = 0; d < D; ++d) in = in && (G.x[d].x[0] <= pb[d] && pb[d] <= G.x[d].x[G.n[d] - 1]);
    if (!in) {
        range[0] = 1;
        range[1] = 1;
        return;
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const Axis& A = G.x[d];
        Ib[d] = bracket_slot(pb[d], A, A.x[0], A.idx[0], I[d]);
        t[d] = (pb[d] - A.x[Ib[d]]) * A.idx[at_node ? I[d] : Ib[d]];
    }
    const int64_t n0 = G.n[0], n1 = G.n[1];
    int64_t b = Ib[0] + n0 * Ib[1], node = I[0] + n0 * I[1];
    if constexpr (D == 3) {
        b += n0 * n1 * Ib[2];
        node += n0 * n1 * I[2];
    }
    for (int f = 0; f < Fs.n; ++f) {
        const double* __restrict__ F0 = Fs.F0[f];
        double q[1 << D];
#pragma unroll
        for (int m = 0; m < (1 << D); ++m) {
            const int64_t o =
                (m & 1) + ((m >> 1) & 1) * n0 + (D == 3 ? ((m >> 2) & 1) * n0 * n1 : 0);
            q[m] = F0[b + o];
        }
        double v = lerp1(t[1], lerp1(t[0], q[0], q[1]), lerp1(t[0], q[2], q[3]));
        if constexpr (D == 3) {
            const double hi = lerp1(t[1], lerp1(t[0], q[4], q[5]), lerp1(t[0], q[6], q[7]));
            v = lerp1(t[2], v, hi);
        }
        Fs.F[f][node] = v;
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

template <int D, int M, int S>
int launch(jp_ctx* ctx, const Grid<D>& G, const Vel<D>& vg, const VPtrs<D>& V, Ptr3 cp,
           const uint8_t* index, const RKParams& P) {
    if constexpr (S == kLinear) {
        if (!ctx->opt_advect_slots) {
            int64_t tiles = 1;
            for (int d = 0; d < D; ++d)
                tiles *= (G.ni[d] + F64Tile<D>::n[d] - 1) / F64Tile<D>::n[d];
            // slot groups when the tiles alone leave the chip under-filled
            const int span = slot_span(ctx, tiles, G.K, 2048);
            const dim3 grid((unsigned)tiles, (unsigned)((G.K + span - 1) / span));
            if constexpr (D == 2) {
                k_advect_tile<D, M, 8, false><<<grid, kBlock, 0, ctx->stream>>>(
                    G, vg, V, cp, index, nullptr, nullptr, P, span);
                JP_LAUNCHED(ctx);
            } else {
                // scratch: a flag per slot | a flag per workgroup
                const size_t nf = ((size_t)G.C * G.K + 255) & ~(size_t)255;
                const size_t nwg = (size_t)grid.x * grid.y;
                int rc = ensure_scratch(ctx, nf + nwg);
                if (rc) return rc;
                uint8_t* flags = static_cast<uint8_t*>(ctx->scratch);
                uint8_t* any = flags + nf;
                k_advect_tile<D, M, 4, true><<<grid, kBlock, 0, ctx->stream>>>(
                    G, vg, V, cp, index, flags, any, P, span);
                JP_LAUNCHED(ctx);
                k_advect_flagged<D, M><<<grid, kBlock, 0, ctx->stream>>>(G, vg, V, cp, flags,
                                                                        any, P, span);
                JP_LAUNCHED(ctx);
            }
            return JP_OK;
        }
    }
    dim3 grid((unsigned)row_blocks<D>(G), (unsigned)G.K);
    k_advect_slot<D, M, S><<<grid, kBlock, 0, ctx->stream>>>(G, vg, V, cp, index, P);
    JP_LAUNCHED(ctx);
    return JP_OK;
}

template <int D, int S>
int launch_method(jp_ctx* ctx, const Grid<D>& G, const Vel<D>& vg, const VPtrs<D>& V, Ptr3 cp,
                  const uint8_t* index, int method, const RKParams& P) {
    if (method == JP_EULER) return launch<D, JP_EULER, S>(ctx, G, vg, V, cp, index, P);
    if (method == JP_RK2) return launch<D, JP_RK2, S>(ctx, G, vg, V, cp, index, P);
    return launch<D, JP_RK4, S>(ctx, G, vg, V, cp, index, P);
}

template <int D>
int run(jp_ctx* ctx, const jp_cellgrid* g, double* const* coords, const uint8_t* index,
        const double* const* Vp, const jp_velgrid* vgp, int method, int scheme,
        const RKParams& P) {
    Grid<D> G = to_grid<D>(g);
    Vel<D> vg = to_vel<D>(vgp);
    VPtrs<D> V;
    for (int c = 0; c < D; ++c) V.v[c] = Vp[c];
    Ptr3 cp{{nullptr, nullptr, nullptr}};
    for (int d = 0; d < D; ++d) cp.p[d] = coords[d];
    int rc;
    if (scheme == kMQS)
        rc = launch_method<D, kMQS>(ctx, G, vg, V, cp, index, method, P);
    else if (scheme == kLinP)
        rc = launch_method<D, kLinP>(ctx, G, vg, V, cp, index, method, P);
    else
        rc = launch_method<D, kLinear>(ctx, G, vg, V, cp, index, method, P);
    if (rc) return rc;
    return finish(ctx);
}

template <int D, int M>
int launch_sml(jp_ctx* ctx, const NodeGrid<D>& G, const Vel<D>& vg, const VPtrs<D>& V,
               const SmlFields& Fs, const RKParams& P, int scheme, int at_node) {
    const unsigned blocks = (unsigned)blocks_for(G.N);
    int* range = drange(ctx);
    if (scheme == kMQS)
        k_sml<D, M, kMQS><<<blocks, kBlock, 0, ctx->stream>>>(G, vg, V, Fs, P, at_node, range);
    else if (scheme == kLinP)
        k_sml<D, M, kLinP><<<blocks, kBlock, 0, ctx->stream>>>(G, vg, V, Fs, P, at_node, range);
    else
        k_sml<D, M, kLinear><<<blocks, kBlock, 0, ctx->stream>>>(G, vg, V, Fs, P, at_node, range);
    JP_LAUNCHED(ctx);
    return JP_OK;
}

template <int D>
int run_sml(jp_ctx* ctx, const jp_axis* grid, const jp_velgrid* vgp, const double* const* Vp,
            const SmlFields& Fs, int method, int scheme, int at_node, const RKParams& P) {
    NodeGrid<D> G;
    G.N = 1;
    for (int d = 0; d < D; ++d) {
        G.n[d] = grid[d].n;
        G.N *= grid[d].n - 2;
        G.x[d] = Axis{grid[d].x, grid[d].dx, grid[d].inv_dx, grid[d].n};
    }
    Vel<D> vg = to_vel<D>(vgp);
    VPtrs<D> V;
    for (int c = 0; c < D; ++c) V.v[c] = Vp[c];
    // this call's flag starts clear (a kernel, so the call can be captured)
    int rc = zero_bytes(ctx, drange(ctx), sizeof(int));
    if (rc) return rc;
    rc = method == JP_RK2 ? launch_sml<D, JP_RK2>(ctx, G, vg, V, Fs, P, scheme, at_node)
                          : launch_sml<D, JP_RK4>(ctx, G, vg, V, Fs, P, scheme, at_node);
    if (rc) return rc;
    if (!ctx->sync || ctx->capturing) return JP_OK;
    int bad = 0;
    JP_HIP(ctx, hipMemcpyAsync(&bad, drange(ctx), sizeof bad, hipMemcpyDeviceToHost, ctx->stream));
    JP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return bad ? range_report(ctx) : JP_OK;
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

RKParams rk_params(double alpha, double dt, double sign) {
    RKParams P;
    P.dt = sign * dt;
    P.adt = (sign * alpha) * dt;
    P.half = alpha == 0.5;
    const double ia = 1.0 / alpha;
    P.w0 = std::fma(-0.5, ia, 1.0);
    P.w1 = 0.5 * ia;
    return P;
}

int advect_entry(jp_ctx* ctx, const jp_cellgrid* g, double* const* coords, const uint8_t* index,
                 const double* const* V, const jp_velgrid* vg, int method, double alpha,
                 double dt, int scheme) {
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
    if (scheme < kLinear || scheme > kLinP)
        return fail(ctx, JP_ERR_ARG, "unknown velocity interpolation scheme%s %lld", "", scheme);
    const RKParams P = rk_params(alpha, dt, 1.0);
    return g->ndim == 2 ? run<2>(ctx, g, coords, index, V, vg, method, scheme, P)
                        : run<3>(ctx, g, coords, index, V, vg, method, scheme, P);
}

}  // namespace

namespace jp {
int range_report(jp_ctx* ctx) {
    // reported now: clear the call and sticky words
    int rc = zero_bytes(ctx, drange(ctx), 2 * sizeof(int));
    if (rc) return rc;
    JP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return fail(ctx, JP_ERR_ARG,
                "semilagrangian_advection: a backtracked point left the velocity or vertex "
                "grid%s (its node was left unchanged; reduce dt)");
}

int range_sticky(jp_ctx* ctx) {
    int st = 0;
    JP_HIP(ctx, hipMemcpyAsync(&st, drange(ctx) + 1, sizeof st, hipMemcpyDeviceToHost,
                               ctx->stream));
    JP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return st ? range_report(ctx) : JP_OK;
}
}  // namespace jp

extern "C" int jp_advection(jp_ctx* ctx, const jp_cellgrid* g, double* const* coords,
                            const uint8_t* index, const double* const* V, const jp_velgrid* vg,
                            int method, double alpha, double dt) {
    return advect_entry(ctx, g, coords, index, V, vg, method, alpha, dt, kLinear);
}

extern "C" int jp_advection_scheme(jp_ctx* ctx, const jp_cellgrid* g, double* const* coords,
                                   const uint8_t* index, const double* const* V,
                                   const jp_velgrid* vg, int method, double alpha, double dt,
                                   int scheme) {
    return advect_entry(ctx, g, coords, index, V, vg, method, alpha, dt, scheme);
}

extern "C" int jp_semilagrangian(jp_ctx* ctx, int ndim, const jp_axis* grid,
                                 const jp_velgrid* vg, const double* const* V, int nfields,
                                 double* const* F, const double* const* F0, int method,
                                 double alpha, double dt, int scheme, int di_at_node) {
    if (!ctx) return JP_ERR_STATE;
    if (ndim != 2 && ndim != 3) return fail(ctx, JP_ERR_ARG, "ndim must be 2 or 3%s");
    if (!grid) return fail(ctx, JP_ERR_ARG, "null grid%s");
    for (int d = 0; d < ndim; ++d) {
        if (!grid[d].x || !grid[d].dx || !grid[d].inv_dx)
            return fail(ctx, JP_ERR_ARG, "null grid axis pointer%s %lld", "", d);
        if (grid[d].n < 3) return fail(ctx, JP_ERR_ARG, "grid axis%s %lld has fewer than 3 nodes", "", d);
    }
    if (!vg) return fail(ctx, JP_ERR_ARG, "null velocity grid%s");
    if (!V) return fail(ctx, JP_ERR_ARG, "null velocity array%s");
    for (int c = 0; c < ndim; ++c) {
        if (!V[c]) return fail(ctx, JP_ERR_ARG, "null velocity component%s %lld", "", c);
        for (int d = 0; d < ndim; ++d) {
            const jp_axis& a = vg->ax[c][d];
            if (!a.x || !a.dx || !a.inv_dx)
                return fail(ctx, JP_ERR_ARG, "null velocity axis pointer%s (component %lld)", "", c);
            // the velocity brackets are seeded with the node index
            if (a.n < grid[d].n)
                return fail(ctx, JP_ERR_ARG,
                            "velocity axis shorter than the nodal grid%s (component %lld)", "", c);
        }
    }
    if (method != JP_RK2 && method != JP_RK4)
        return fail(ctx, JP_ERR_ARG, "semilagrangian_advection takes RungeKutta2 or RungeKutta4%s");
    if (method == JP_RK2 && !(0.0 < alpha && alpha < 1.0))
        return fail(ctx, JP_ERR_ARG, "Only 0 < α < 1 is supported%s");
    if (scheme < kLinear || scheme > kLinP)
        return fail(ctx, JP_ERR_ARG, "unknown velocity interpolation scheme%s %lld", "", scheme);
    if (nfields < 1 || nfields > JP_MAX_ARGS || !F || !F0)
        return fail(ctx, JP_ERR_ARG, "nfields must be 1..JP_MAX_ARGS%s");
    SmlFields Fs;
    Fs.n = nfields;
    for (int f = 0; f < nfields; ++f) {
        if (!F[f] || !F0[f]) return fail(ctx, JP_ERR_ARG, "null field%s %lld", "", f);
        for (int h = 0; h < nfields; ++h)
            if ((const double*)F[f] == F0[h])
                return fail(ctx, JP_ERR_ARG,
                            "F and F0 must be different arrays%s (F[%lld] is an F0)", "", f);
        Fs.F[f] = F[f];
        Fs.F0[f] = F0[f];
    }
    const RKParams P = rk_params(alpha, dt, -1.0);
    return ndim == 2 ? run_sml<2>(ctx, grid, vg, V, Fs, method, scheme, di_at_node != 0, P)
                     : run_sml<3>(ctx, grid, vg, V, Fs, method, scheme, di_at_node != 0, P);
}
