// jp_phase.hip -- update_phase_ratios! (src/PhaseRatios/*.jl) on gfx950.
//
// Phase ratios are the fractions of each material phase at a grid location,
// from the particles around it: every particle adds its bilinear weight
// (src/PhaseRatios/utils.jl:61-70: the product over d of
// muladd(-|a_d - b_d|, inv(di_d), 1)) to the running sum of its phase, and
// the sums are divided by their total (w .* inv(sum(w)), summed left to
// right).  Locations: cell centres, vertices, the staggered velocity faces
// and, in 3D, the edge midpoints.  Outputs use the device CellArray layout of
// a PhaseRatios field: location l, phase k at out[l + L*k].
//
// One thread per output location (centres, vertices) or per cell (faces and
// midpoints write the location(s) the reference assigns to that cell); a
// wavefront's lanes are consecutive cells, so every slot-plane load is
// coalesced.  Each location accumulates its particles in the reference's cell
// and slot order, so the result is bit-identical to the sequential oracle.
#include "jp_ctx.hpp"

using namespace jp;

namespace {

template <int D>
struct CAxes {
    Axis a[3];
};

template <int NP>
__device__ inline void add_phase(double (&w)[NP], int nph, double x, double ph) {
#pragma unroll
    for (int j = 0; j < NP; ++j) w[j] = (j < nph && ph == (double)(j + 1)) ? w[j] + x : w[j];
}

template <int D>
__device__ inline double bilinear_weight(const double (&a)[D], const double (&b)[D],
                                         const double (&inv)[D]) {
    double val = 1.0;
#pragma unroll
    for (int d = 0; d < D; ++d) val = val * fma(-fabs(a[d] - b[d]), inv[d], 1.0);
    return val;
}

// w .* inv(sum(w)); zero_nan: w * !isnan(w) (faces and midpoints)
template <int NP>
__device__ inline void normalise_write(const double (&w)[NP], int nph, double* __restrict__ out,
                                       int64_t L, int64_t loc, bool zero_nan) {
    double sum = w[0];
#pragma unroll
    for (int j = 1; j < NP; ++j) sum = j < nph ? sum + w[j] : sum;
    const double iv = 1.0 / sum;
#pragma unroll
    for (int j = 0; j < NP; ++j) {
        if (j >= nph) break;
        const double v = w[j] * iv;
        out[loc + L * j] = (zero_nan && v != v) ? 0.0 : v;
    }
}

template <int D>
__device__ inline bool any_nan(const double (&p)[D]) {
    bool r = false;
#pragma unroll
    for (int d = 0; d < D; ++d) r = r || (p[d] != p[d]);
    return r;
}

// The slots of cell c in order, body(p, phase) for each: the coordinates and
// phase of 8 slots loaded together before the bodies run (one round trip per
// 8 slots instead of one per slot; the bodies, and so every sum, keep the
// slot order).  Empty slots hold NaN coordinates; past K the loads are
// skipped and the coordinates are NaN.
template <int D, typename Body>
__device__ inline void for_slots(const Grid<D>& g, const CPtr3& coords,
                                 const double* __restrict__ phases, int64_t c, Body&& body) {
    for (int k0 = 0; k0 < g.K; k0 += 8) {
        double pb[8][D], phb[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const bool in = k0 + j < g.K;
            const int64_t e = c + g.C * (k0 + j);
#pragma unroll
            for (int d = 0; d < D; ++d) pb[j][d] = in ? coords.p[d][e] : __builtin_nan("");
            phb[j] = in ? phases[e] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) body(pb[j], phb[j]);
    }
}

// particles of cell c within half a cell (<=) of m, in slot order; the
// cell's vertex spacing is returned in di/inv (Julia keeps the loop's last
// assignment for the faces' boundary pass)
template <int D, int NP>
__device__ inline void gather_half(const Grid<D>& g, const CPtr3& coords,
                                   const double* __restrict__ phases, int nph, const int64_t (&J)[D],
                                   const double (&m)[D], double (&w)[NP], double (&di)[D],
                                   double (&inv)[D]) {
    const int64_t c = lin<D>(g.ni, J);
#pragma unroll
    for (int d = 0; d < D; ++d) {
        di[d] = g.xv[d].dx[J[d]];
        inv[d] = g.xv[d].idx[J[d]];
    }
    for_slots<D>(g, coords, phases, c, [&](const double (&p)[D], double ph) {
        if (any_nan<D>(p)) return;
        bool in = true;
#pragma unroll
        for (int d = 0; d < D; ++d) in = in && (fabs(p[d] - m[d]) <= di[d] / 2.0);
        if (!in) return;
        add_phase<NP>(w, nph, bilinear_weight<D>(m, p, inv), ph);
    });
}

// phase_ratios_center_kernel! (src/PhaseRatios/centers.jl:14-30): the
// particles of cell I whose first coordinate is not NaN, about xci[I]
template <int D, int NP>
__global__ void __launch_bounds__(kBlock) k_ratio_center(Grid<D> g, CAxes<D> xc, CPtr3 coords,
                                                         const double* __restrict__ phases,
                                                         int nph, double* __restrict__ out) {
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= g.C) return;
    int64_t I[D];
    unlin<D>(g.ni, c, I);
    double m[D], inv[D], w[NP];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        m[d] = xc.a[d].x[I[d]];
        inv[d] = g.xv[d].idx[I[d]];
    }
#pragma unroll
    for (int j = 0; j < NP; ++j) w[j] = 0.0;
    for_slots<D>(g, coords, phases, c, [&](const double (&p)[D], double ph) {
        if (p[0] != p[0]) return;
        add_phase<NP>(w, nph, bilinear_weight<D>(m, p, inv), ph);
    });
    normalise_write<NP>(w, nph, out, g.C, c, false);
}

// phase_ratios_vertex_kernel! (src/PhaseRatios/vertices.jl:14-107): the
// particles strictly within half a cell of vertex N, from the up to 2^D
// cells around it with the first dimension's offset outermost
template <int D, int NP>
__global__ void __launch_bounds__(kBlock) k_ratio_vertex(Grid<D> g, CPtr3 coords,
                                                         const double* __restrict__ phases,
                                                         int nph, double* __restrict__ out) {
    int64_t nn[D], NN = 1;
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
    double v[D], w[NP];
#pragma unroll
    for (int d = 0; d < D; ++d) v[d] = g.xv[d].x[N[d]];
#pragma unroll
    for (int j = 0; j < NP; ++j) w[j] = 0.0;
#pragma unroll
    for (int q = 0; q < (1 << D); ++q) {
        int64_t J[D];
        bool ok = true;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            J[d] = N[d] - 1 + ((q >> (D - 1 - d)) & 1);
            ok = ok && J[d] >= 0 && J[d] < g.ni[d];
        }
        if (!ok) continue;
        const int64_t c = lin<D>(g.ni, J);
        double di[D], inv[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            di[d] = g.xv[d].dx[J[d]];
            inv[d] = g.xv[d].idx[J[d]];
        }
        for_slots<D>(g, coords, phases, c, [&](const double (&p)[D], double ph) {
            if (any_nan<D>(p)) return;
            bool far = false;
#pragma unroll
            for (int d = 0; d < D; ++d) far = far || (fabs(p[d] - v[d]) >= di[d] / 2.0);
            if (far) return;
            add_phase<NP>(w, nph, bilinear_weight<D>(v, p, inv), ph);
        });
    }
    normalise_write<NP>(w, nph, out, NN, node, false);
}

template <int D>
__device__ inline int64_t lin_dims(const int64_t (&I)[D], const int64_t (&dims)[D]) {
    int64_t c = I[D - 1];
#pragma unroll
    for (int d = D - 2; d >= 0; --d) c = c * dims[d] + I[d];
    return c;
}

// phase_ratios_face_kernel! (src/PhaseRatios/midpoints.jl:24-80) for the
// faces normal to `dim`: cell I writes face I + e_dim from cells I and
// min(I + e_dim, ni) (the last cell counted twice); the first cell along dim
// also writes face I from its own particles about xci - di e_dim / 2, with
// the spacing of the last cell gathered
template <int D, int NP>
__global__ void __launch_bounds__(kBlock) k_ratio_face(Grid<D> g, CAxes<D> xc, CPtr3 coords,
                                                       const double* __restrict__ phases, int nph,
                                                       int dim, double* __restrict__ out) {
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= g.C) return;
    int64_t I[D], dims[D];
    unlin<D>(g.ni, c, I);
    double xcen[D], di[D], inv[D], f[D], w[NP];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        dims[d] = g.ni[d] + (d == dim ? 1 : 0);
        xcen[d] = xc.a[d].x[I[d]];
        di[d] = g.xv[d].dx[I[d]];
        f[d] = xcen[d] + (di[d] * (d == dim ? 1.0 : 0.0)) / 2.0;
    }
#pragma unroll
    for (int j = 0; j < NP; ++j) w[j] = 0.0;
#pragma unroll
    for (int o = 0; o < 2; ++o) {
        int64_t J[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int64_t k = I[d] + (o && d == dim ? 1 : 0);
            J[d] = k > g.ni[d] - 1 ? g.ni[d] - 1 : k;
        }
        gather_half<D, NP>(g, coords, phases, nph, J, f, w, di, inv);
    }
    const int64_t L = [&] {
        int64_t n = 1;
#pragma unroll
        for (int d = 0; d < D; ++d) n *= dims[d];
        return n;
    }();
    int64_t If[D];
#pragma unroll
    for (int d = 0; d < D; ++d) If[d] = I[d] + (d == dim ? 1 : 0);
    normalise_write<NP>(w, nph, out, L, lin_dims<D>(If, dims), true);
    if (I[dim] != 0) return;
    // boundary face of the first cell: di / inv are the last cell's
    double w2[NP];
#pragma unroll
    for (int j = 0; j < NP; ++j) w2[j] = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) f[d] = xcen[d] - (di[d] * (d == dim ? 1.0 : 0.0)) / 2.0;
    for_slots<D>(g, coords, phases, c, [&](const double (&p)[D], double ph) {
        if (any_nan<D>(p)) return;
        bool in = true;
#pragma unroll
        for (int d = 0; d < D; ++d) in = in && (fabs(p[d] - f[d]) <= di[d] / 2.0);
        if (!in) return;
        add_phase<NP>(w2, nph, bilinear_weight<D>(f, p, inv), ph);
    });
    normalise_write<NP>(w2, nph, out, L, lin_dims<D>(I, dims), true);
}

// _phase_ratios_midpoint_kernel! (src/PhaseRatios/midpoints.jl:123-206), 3D,
// offsets o: cell I writes edge I + o from the cells min(I + o .* mask, ni)
// for the masks (1,0,0), (0,1,0), (0,0,1), (1,1,1) about xci + di .* o / 2; a
// cell on a first boundary (I_d = 1 where o_d = 1, 1-based) also writes the
// edges I and I + b, b_d = (ni_d == o_d I_d), about xci - (di .* o .* (-b)) / 2
__device__ inline void midpoint_pass(const Grid<3>& g, const CPtr3& coords,
                                     const double* __restrict__ phases, int nph,
                                     const int64_t (&I)[3], const int (&o)[3],
                                     const double (&m)[3], double (&w)[16], int np) {
    constexpr int MASK[4][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 1}};
    (void)np;
#pragma unroll
    for (int mk = 0; mk < 4; ++mk) {
        int64_t J[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int64_t k = I[d] + o[d] * MASK[mk][d];
            J[d] = k > g.ni[d] - 1 ? g.ni[d] - 1 : k;
        }
        double di[3], inv[3];
        gather_half<3, 16>(g, coords, phases, nph, J, m, w, di, inv);
    }
}

__global__ void __launch_bounds__(kBlock) k_ratio_midpoint(Grid<3> g, CAxes<3> xc, CPtr3 coords,
                                                           const double* __restrict__ phases,
                                                           int nph, int o0, int o1, int o2,
                                                           double* __restrict__ out) {
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= g.C) return;
    const int o[3] = {o0, o1, o2};
    int64_t I[3], dims[3];
    unlin<3>(g.ni, c, I);
    double xcen[3], m[3], dm[3], w[16];
    int64_t L = 1;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        dims[d] = g.ni[d] + o[d];
        L *= dims[d];
        xcen[d] = xc.a[d].x[I[d]];
        dm[d] = g.xv[d].dx[I[d]];
        m[d] = xcen[d] + (dm[d] * (double)o[d]) / 2.0;
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) w[j] = 0.0;
    midpoint_pass(g, coords, phases, nph, I, o, m, w, 16);
    int64_t Io[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) Io[d] = I[d] + o[d];
    normalise_write<16>(w, nph, out, L, lin_dims<3>(Io, dims), true);
    bool first = false;
    int b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        first = first || (o[d] * (I[d] + 1) == 1);
        b[d] = g.ni[d] == (int64_t)o[d] * (I[d] + 1) ? 1 : 0;
    }
    if (!first) return;
    for (int it = 0; it < 2; ++it) {
        double w2[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) w2[j] = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) m[d] = xcen[d] - ((dm[d] * (double)o[d]) * (double)(0 - b[d])) / 2.0;
        midpoint_pass(g, coords, phases, nph, I, o, m, w2, 16);
        int64_t Ib[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) Ib[d] = I[d] + (it ? b[d] : 0);
        normalise_write<16>(w2, nph, out, L, lin_dims<3>(Ib, dims), true);
    }
}

template <int D, int NP>
int launch_ratios(jp_ctx* ctx, const Grid<D>& G, const CAxes<D>& xa, const CPtr3& cp,
                  const double* phases, int nph, double* center, double* vertex,
                  double* const* faces) {
    if (center) {
        k_ratio_center<D, NP><<<(unsigned)blocks_for(G.C), kBlock, 0, ctx->stream>>>(
            G, xa, cp, phases, nph, center);
        JP_LAUNCHED(ctx);
    }
    if (vertex) {
        int64_t NN = 1;
        for (int d = 0; d < D; ++d) NN *= G.ni[d] + 1;
        k_ratio_vertex<D, NP><<<(unsigned)blocks_for(NN), kBlock, 0, ctx->stream>>>(
            G, cp, phases, nph, vertex);
        JP_LAUNCHED(ctx);
    }
    for (int d = 0; d < D; ++d) {
        if (!faces || !faces[d]) continue;
        k_ratio_face<D, NP><<<(unsigned)blocks_for(G.C), kBlock, 0, ctx->stream>>>(
            G, xa, cp, phases, nph, d, faces[d]);
        JP_LAUNCHED(ctx);
    }
    return JP_OK;
}

template <int D>
int run_ratios(jp_ctx* ctx, const jp_cellgrid* gc, const jp_axis* xc, const double* const* coords,
               const double* phases, int nph, double* center, double* vertex,
               double* const* faces, double* const* midpoints) {
    Grid<D> G = to_grid<D>(gc);
    CAxes<D> xa{};
    for (int d = 0; d < D; ++d) xa.a[d] = Axis{xc[d].x, xc[d].dx, xc[d].inv_dx, xc[d].n};
    CPtr3 cp{{nullptr, nullptr, nullptr}};
    for (int d = 0; d < D; ++d) cp.p[d] = coords[d];
    int rc = nph <= 4 ? launch_ratios<D, 4>(ctx, G, xa, cp, phases, nph, center, vertex, faces)
                      : launch_ratios<D, 16>(ctx, G, xa, cp, phases, nph, center, vertex, faces);
    if (rc) return rc;
    if constexpr (D == 3) {
        // xy, yz, xz (src/PhaseRatios/midpoints.jl:214-225)
        constexpr int OFF[3][3] = {{1, 1, 0}, {0, 1, 1}, {1, 0, 1}};
        for (int k = 0; k < 3; ++k) {
            if (!midpoints || !midpoints[k]) continue;
            k_ratio_midpoint<<<(unsigned)blocks_for(G.C), kBlock, 0, ctx->stream>>>(
                G, xa, cp, phases, nph, OFF[k][0], OFF[k][1], OFF[k][2], midpoints[k]);
            JP_LAUNCHED(ctx);
        }
    }
    return finish(ctx);
}

}  // namespace

extern "C" int jp_update_phase_ratios(jp_ctx* ctx, const jp_cellgrid* g, const jp_axis* xc,
                                      const double* const* coords, const double* phases,
                                      int nphases, double* center, double* vertex,
                                      double* const* faces, double* const* midpoints) {
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if (!coords || !phases || !xc) return fail(ctx, JP_ERR_ARG, "null coords, phases or xc%s");
    for (int d = 0; d < g->ndim; ++d) {
        if (!coords[d]) return fail(ctx, JP_ERR_ARG, "null coordinate plane%s %lld", "", d);
        if (!xc[d].x || xc[d].n != g->ni[d])
            return fail(ctx, JP_ERR_ARG, "centre axis%s %lld must have ni points", "", d);
    }
    if (nphases < 1 || nphases > JP_MAX_PHASES)
        return fail(ctx, JP_ERR_ARG, "nphases must be in 1..JP_MAX_PHASES%s (got %lld)", "",
                    nphases);
    return g->ndim == 2 ? run_ratios<2>(ctx, g, xc, coords, phases, nphases, center, vertex, faces,
                                        nullptr)
                        : run_ratios<3>(ctx, g, xc, coords, phases, nphases, center, vertex, faces,
                                        midpoints);
}
