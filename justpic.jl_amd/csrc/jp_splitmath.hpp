// jp_splitmath.hpp -- the per-particle arithmetic of the FP32-split
// advection (cell index plus float offset, velocity scaled to cells per step
// in float32), shared by the split advection kernels and the split forward
// of the adjoint.
#pragma once

#include <type_traits>

#include "jp_ctx.hpp"

namespace jp {
namespace split {

// ------------------------------------------------------------ advection
// One workgroup per tile of cells (2D 64 x 4, 3D 32 x 4 x 2), one thread per
// cell.  The tile's window of every velocity component, nodes [t0 - 1,
// t0 + T + 3) per dimension, is staged in LDS already scaled to cells per
// step, float(V_c * dt / h_c).  That window covers every stage point within one
// cell of the tile (vertex axis: nodes J, J+1; centre axis: J+up, J+up+1);
// points further out read the same values from global memory.  Each thread
// walks its cell's slots in chunks of 8 planes, in lockstep with the
// neighbouring cells of its wavefront.
// Tile shapes: TX*TY*TZ = kBlock cells, one per thread; the LDS window holds
// nodes [t0 - 1, t0 + T + 3) per dimension.
template <int D, int TX, int TY, int TZ = 1>
struct Tile {
    static_assert(TX * TY * TZ == kBlock, "one thread per cell of the tile");
    __host__ __device__ static constexpr int n(int d) { return d == 0 ? TX : (d == 1 ? TY : TZ); }
    __host__ __device__ static constexpr int w(int d) { return n(d) + 4; }
    __host__ __device__ static constexpr int size() {
        return D == 2 ? (TX + 4) * (TY + 4) : (TX + 4) * (TY + 4) * (TZ + 4);
    }
};
template <int D>
using DefaultTile = typename std::conditional<D == 2, Tile<2, 64, 4>, Tile<3, 32, 4, 2>>::type;

// FP64 input velocity (component grids n[c][d]) and the vertex spacings
template <int D>
struct SVelIn {
    const double* v[D];
    int n[D][D];
};

// Affine maps of one grid's velocity axes (SplitGeom::r, o)
template <int D>
struct SMap {
    double r[D][D], o[D][D];
};

// Per cell I and velocity axis (c, d): the node coordinate of the point I + u
// is I + ei + (u r + ef), where e = I (r - 1) + o = ei + ef is split in FP64
// into an integer and a fraction once per cell, so that the float32
// arithmetic per stage only handles values of order one.
template <int D>
struct CellMap {
    int ei[D][D];
    float ef[D][D];
    float r[D][D];
};

// STD: the standard staggering, r = 1 and o = 0 (the component's own axis) or
// 1/2 (a centre axis with one ghost node), as compile-time constants
template <int D, bool STD = false>
__device__ inline void cell_map(const SMap<D>& m, const int (&I)[D], CellMap<D>& cm) {
#pragma unroll
    for (int c = 0; c < D; ++c)
#pragma unroll
        for (int d = 0; d < D; ++d) {
            if constexpr (STD) {
                cm.ei[c][d] = 0;
                cm.ef[c][d] = c == d ? 0.0f : 0.5f;
                cm.r[c][d] = 1.0f;
                continue;
            }
            const double e = fma((double)I[d], m.r[c][d] - 1.0, m.o[c][d]);
            const double fl = floor(e);
            cm.ei[c][d] = (int)fl;
            cm.ef[c][d] = (float)(e - fl);
            cm.r[c][d] = (float)m.r[c][d];
        }
}

// first node of the LDS window of component c along d for a tile at t0
template <int D, bool STD = false>
__device__ inline int window_origin(const SMap<D>& m, int c, int d, int t0) {
    if constexpr (STD) return t0 - 1;
    return t0 + (int)floor(fma((double)t0, m.r[c][d] - 1.0, m.o[c][d])) - 1;
}

// velocity of component c at cell I + offset u (u in cell units, any value),
// from the LDS window sw (first node w0 per dimension) when it holds the
// four/eight corners, else from global memory.  Outside the limits of the
// component's grid the result is +Inf, as in check_local_limits
// (src/Advection/advection.jl:39-47: first <= p <= last, inclusive).
template <int D, class TS>
__device__ inline float interp_tile(const float* __restrict__ sw, const SVelIn<D>& V, int c,
                                    double sc, const int (&w0)[D], const int (&I)[D],
                                    const CellMap<D>& cm, const float (&u)[D]) {
    int k[D];
    float t[D];
    bool inw = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        if (!(fabsf(u[d]) < 1.0e6f)) return __builtin_inff();
        const float s = fmaf(u[d], cm.r[c][d], cm.ef[c][d]);
        const float fl = floorf(s);
        int kk = I[d] + cm.ei[c][d] + (int)fl;
        float tt = s - fl;
        const int n = V.n[c][d];
        if (kk < 0 || kk > n - 2) {
            if (kk == n - 1 && tt == 0.0f) {  // on the last node
                kk = n - 2;
                tt = 1.0f;
            } else {
                return __builtin_inff();
            }
        }
        k[d] = kk;
        t[d] = tt;
        const int kl = kk - w0[d];
        inw = inw && kl >= 0 && kl + 1 < TS::w(d);
    }
    float q[1 << D];
    if (inw) {
        constexpr int W0 = TS::w(0), W1 = TS::w(1);
        int b = (k[0] - w0[0]) + W0 * (k[1] - w0[1]);
        if constexpr (D == 3) b += W0 * W1 * (k[2] - w0[2]);
#pragma unroll
        for (int r = 0; r < (1 << D); ++r) {
            const int o = (r & 1) + ((r >> 1) & 1) * W0 + (D == 3 ? ((r >> 2) & 1) * W0 * W1 : 0);
            q[r] = sw[b + o];
        }
    } else {
        const double* __restrict__ F = V.v[c];
        const int64_t n0 = V.n[c][0], n1 = V.n[c][1];
        int64_t b = k[0] + n0 * (int64_t)k[1];
        if constexpr (D == 3) b += n0 * n1 * (int64_t)k[2];
#pragma unroll
        for (int r = 0; r < (1 << D); ++r) {
            const int64_t o = (r & 1) + ((r >> 1) & 1) * n0 + (D == 3 ? ((r >> 2) & 1) * n0 * n1 : 0);
            q[r] = (float)(F[b + o] * sc);
        }
    }
    const float lo = lerp1f(t[1], lerp1f(t[0], q[0], q[1]), lerp1f(t[0], q[2], q[3]));
    if constexpr (D == 2) {
        return lo;
    } else {
        const float hi = lerp1f(t[1], lerp1f(t[0], q[4], q[5]), lerp1f(t[0], q[6], q[7]));
        return lerp1f(t[2], lo, hi);
    }
}

struct RKf {
    float alpha;
    int half;
    float w0, w1;
};

template <int D, int M, class TS>
__device__ inline void advect_split_one(const float (*sV)[TS::size()], const SVelIn<D>& V,
                                        const double (&sc)[D], const int (&w0)[D][D],
                                        const int (&I)[D], const CellMap<D>& cm, const RKf& P,
                                        const float (&u0)[D], float (&out)[D]) {
    float v0[D];
#pragma unroll
    for (int cc = 0; cc < D; ++cc) v0[cc] = interp_tile<D, TS>(sV[cc], V, cc, sc[cc], w0[cc], I, cm, u0);
    if constexpr (M == JP_EULER) {
#pragma unroll
        for (int d = 0; d < D; ++d) out[d] = u0[d] + v0[d];
    } else if constexpr (M == JP_RK2) {
        float u1[D], v1[D];
#pragma unroll
        for (int d = 0; d < D; ++d) u1[d] = fmaf(P.alpha, v0[d], u0[d]);
#pragma unroll
        for (int cc = 0; cc < D; ++cc) v1[cc] = interp_tile<D, TS>(sV[cc], V, cc, sc[cc], w0[cc], I, cm, u1);
        if (P.half) {
#pragma unroll
            for (int d = 0; d < D; ++d) out[d] = u0[d] + v1[d];
        } else {
#pragma unroll
            for (int d = 0; d < D; ++d) out[d] = u0[d] + fmaf(P.w1, v1[d], P.w0 * v0[d]);
        }
    } else {
        float q[D], k2[D], k3[D], k4[D];
#pragma unroll
        for (int d = 0; d < D; ++d) q[d] = u0[d] + 0.5f * v0[d];
#pragma unroll
        for (int cc = 0; cc < D; ++cc) k2[cc] = interp_tile<D, TS>(sV[cc], V, cc, sc[cc], w0[cc], I, cm, q);
#pragma unroll
        for (int d = 0; d < D; ++d) q[d] = u0[d] + 0.5f * k2[d];
#pragma unroll
        for (int cc = 0; cc < D; ++cc) k3[cc] = interp_tile<D, TS>(sV[cc], V, cc, sc[cc], w0[cc], I, cm, q);
#pragma unroll
        for (int d = 0; d < D; ++d) q[d] = u0[d] + k3[d];
#pragma unroll
        for (int cc = 0; cc < D; ++cc) k4[cc] = interp_tile<D, TS>(sV[cc], V, cc, sc[cc], w0[cc], I, cm, q);
#pragma unroll
        for (int d = 0; d < D; ++d)
            out[d] = u0[d] + (((v0[d] + 2.0f * k2[d]) + 2.0f * k3[d]) + k4[d]) / 6.0f;
    }
}


}  // namespace split
}  // namespace jp
