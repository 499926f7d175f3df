// jp_transfer.hip -- the other grid <-> particle transfers of the reference
// on gfx950: grid2particle_flip! (PIC/FLIP blend), particle2centroid! and
// centroid2particle! (cell-centred fields).  FP64, bit-identical to the
// reference's arithmetic order (explicit fma where it calls muladd).
#include "jp_ctx.hpp"

using namespace jp;

namespace {

template <int D>
struct CAxes {
    Axis a[3];
};

struct FieldPtrs {
    const double* p[JP_MAX_ARGS];
};

// lerp of F's 2^D corners (F column-major with dims nn) at corner cell k
template <int D>
__device__ inline double lerp_corners(const double* __restrict__ F, const int64_t (&nn)[D],
                                      const int64_t (&k)[D], const double (&t)[D]) {
    int64_t b = k[D - 1];
#pragma unroll
    for (int d = D - 2; d >= 0; --d) b = b * nn[d] + k[d];
    const int64_t n0 = nn[0];
    const double lo = lerp1(t[1], lerp1(t[0], F[b], F[b + 1]), lerp1(t[0], F[b + n0], F[b + n0 + 1]));
    if constexpr (D == 2) {
        return lo;
    } else {
        const int64_t s = nn[0] * nn[1];
        const double hi = lerp1(t[1], lerp1(t[0], F[b + s], F[b + s + 1]),
                                lerp1(t[0], F[b + s + n0], F[b + s + n0 + 1]));
        return lerp1(t[2], lo, hi);
    }
}

// h_d = abs(minimum(diff(xv_d))) (grid_size, src/Interpolations/utils.jl:61-63):
// one workgroup, written to h[0..D-1]
template <int D>
__global__ void __launch_bounds__(kBlock) k_min_spacing(Grid<D> g, double* __restrict__ h) {
    __shared__ double red[kBlock / 64];
    for (int d = 0; d < D; ++d) {
        double m = __builtin_inf();
        for (int64_t k = threadIdx.x; k + 1 < g.xv[d].n; k += blockDim.x)
            m = fmin(m, g.xv[d].x[k + 1] - g.xv[d].x[k]);
#pragma unroll
        for (int s = 32; s >= 1; s >>= 1) m = fmin(m, __shfl_xor(m, s, 64));
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            double r = red[0];
            for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = fmin(r, red[w]);
            h[d] = fabs(r);
        }
        __syncthreads();
    }
}

// grid2particle_flip! (src/Interpolations/grid_to_particle.jl:125-207): one
// thread per slot (row mapping); F_pic and F0_pic at the same t, F_flip =
// Fp + (F_pic - F0_pic), Fp = muladd(F_pic, α, F_flip * (1 - α))
template <int D>
__global__ void __launch_bounds__(kBlock) k_g2p_flip(Grid<D> g, int nf, ArgPtrs Fp, FieldPtrs F,
                                                     FieldPtrs F0, CPtr3 coords,
                                                     const uint8_t* __restrict__ index,
                                                     double alpha,
                                                     const double* __restrict__ h) {
    int64_t c, I[D];
    if (!row_cell<D>(g, c, I)) return;
    const int64_t e = c + g.C * (int64_t)blockIdx.y;
    if (!index[e]) return;
    int64_t nn[D], k[D];
    double t[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        nn[d] = g.ni[d] + 1;
        k[d] = I[d];
        t[d] = (coords.p[d][e] - g.xv[d].x[I[d]]) * (1.0 / h[d]);
    }
    const double om = 1.0 - alpha;
    for (int f = 0; f < nf; ++f) {
        const double Fpic = lerp_corners<D>(F.p[f], nn, k, t);
        const double F0pic = lerp_corners<D>(F0.p[f], nn, k, t);
        const double dF = Fpic - F0pic;
        const double Fflip = Fp.p[f][e] + dF;
        Fp.p[f][e] = fma(Fpic, alpha, Fflip * om);
    }
}

// particle2centroid! (src/Interpolations/particle_to_grid_centroid.jl:12-128):
// one thread per cell (lanes are consecutive cells: slot-plane loads
// coalesce); the reference's four methods: 2D scalar (bilinear weights, any
// NaN skips, divide), 2D tuple (inverse squared distance, any NaN, times the
// inverse), 3D scalar (bilinear, NaN first coordinate, inverse), 3D tuple
// (bilinear, any NaN, inverse)
template <int D>
__global__ void __launch_bounds__(kBlock) k_p2c(Grid<D> g, CAxes<D> xc, int nf, ArgPtrs F,
                                                FieldPtrs Fp, CPtr3 coords, int tuple) {
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= g.C) return;
    int64_t I[D];
    unlin<D>(g.ni, c, I);
    const bool dist = D == 2 && tuple;
    const bool first_only = D == 3 && !tuple;
    double m[D], inv[D], w = 0.0, wF[JP_MAX_ARGS];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        m[d] = xc.a[d].x[I[d]];
        inv[d] = g.xv[d].idx[I[d]];
    }
    for (int f = 0; f < nf; ++f) wF[f] = 0.0;
    for (int ip = 0; ip < g.K; ++ip) {
        const int64_t e = c + g.C * ip;
        double p[D];
#pragma unroll
        for (int d = 0; d < D; ++d) p[d] = coords.p[d][e];
        bool skip = p[0] != p[0];
        if (!first_only)
#pragma unroll
            for (int d = 1; d < D; ++d) skip = skip || (p[d] != p[d]);
        if (skip) continue;
        double wi;
        if (dist) {
            double a = m[0] - p[0];
            double s = a * a;
#pragma unroll
            for (int d = 1; d < D; ++d) {
                const double b = m[d] - p[d];
                s = s + b * b;
            }
            const double dd = sqrt(s);
            wi = 1.0 / (dd * dd);
        } else {
            wi = 1.0;
#pragma unroll
            for (int d = 0; d < D; ++d) wi = wi * fma(-fabs(m[d] - p[d]), inv[d], 1.0);
        }
        w += wi;
        for (int f = 0; f < nf; ++f) wF[f] = fma(wi, Fp.p[f][e], wF[f]);
    }
    if (D == 2 && !tuple) {
        F.p[0][c] = wF[0] / w;
    } else {
        const double iw = 1.0 / w;
        for (int f = 0; f < nf; ++f) F.p[f][c] = wF[f] * iw;
    }
}

// centroid2particle! (src/Interpolations/centroid_to_particle.jl:15-76): one
// thread per slot; shifted_index against xci[I], clamped to 1..ni-1
// (1-based), lerp of the centred field's corners
template <int D>
__global__ void __launch_bounds__(kBlock) k_c2p(Grid<D> g, CAxes<D> xc, int nf, ArgPtrs Fp,
                                                FieldPtrs F, CPtr3 coords) {
    int64_t c, I[D];
    if (!row_cell<D>(g, c, I)) return;
    const int64_t e = c + g.C * (int64_t)blockIdx.y;
    double p[D];
    bool nan = false;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        p[d] = coords.p[d][e];
        nan = nan || (p[d] != p[d]);
    }
    if (nan) return;
    int64_t nn[D], k[D];
    double t[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        nn[d] = g.ni[d];
        int64_t kk = p[d] < xc.a[d].x[I[d]] ? I[d] - 1 : I[d];
        kk = kk > g.ni[d] - 2 ? g.ni[d] - 2 : (kk < 0 ? 0 : kk);
        k[d] = kk;
        t[d] = (p[d] - xc.a[d].x[kk]) * xc.a[d].idx[kk];
    }
    for (int f = 0; f < nf; ++f) Fp.p[f][e] = lerp_corners<D>(F.p[f], nn, k, t);
}

template <int D>
CAxes<D> caxes(const jp_axis* xc) {
    CAxes<D> a{};
    for (int d = 0; d < D; ++d) a.a[d] = Axis{xc[d].x, xc[d].dx, xc[d].inv_dx, xc[d].n};
    return a;
}

ArgPtrs argptrs(double* const* a, int n) {
    ArgPtrs p{};
    for (int i = 0; i < n; ++i) p.p[i] = a[i];
    return p;
}

FieldPtrs fieldptrs(const double* const* a, int n) {
    FieldPtrs p{};
    for (int i = 0; i < n; ++i) p.p[i] = a[i];
    return p;
}

CPtr3 cptr3(const double* const* c, int D) {
    CPtr3 p{{nullptr, nullptr, nullptr}};
    for (int d = 0; d < D; ++d) p.p[d] = c[d];
    return p;
}

int check_arrays(jp_ctx* ctx, int n, const void* const* p, const char* what) {
    if (!p) return fail(ctx, JP_ERR_ARG, "null %s array", what);
    for (int i = 0; i < n; ++i)
        if (!p[i]) return fail(ctx, JP_ERR_ARG, "null %s pointer %lld", what, i);
    return JP_OK;
}

int check_centres(jp_ctx* ctx, const jp_cellgrid* g, const jp_axis* xc) {
    if (!xc) return fail(ctx, JP_ERR_ARG, "null centre axes%s");
    for (int d = 0; d < g->ndim; ++d)
        if (!xc[d].x || !xc[d].dx || !xc[d].inv_dx || xc[d].n != g->ni[d])
            return fail(ctx, JP_ERR_ARG, "centre axis%s %lld must have ni points", "", d);
    return JP_OK;
}

template <int D>
int run_flip(jp_ctx* ctx, const jp_cellgrid* gc, int nf, double* const* Fp, const double* const* F,
             const double* const* F0, const double* const* coords, const uint8_t* index,
             double alpha) {
    Grid<D> G = to_grid<D>(gc);
    double* h = reinterpret_cast<double*>(static_cast<char*>(ctx->dsmall) + 2048);
    k_min_spacing<D><<<1, kBlock, 0, ctx->stream>>>(G, h);
    JP_LAUNCHED(ctx);
    const dim3 grid((unsigned)row_blocks<D>(G), (unsigned)G.K);
    k_g2p_flip<D><<<grid, kBlock, 0, ctx->stream>>>(G, nf, argptrs(Fp, nf), fieldptrs(F, nf),
                                                    fieldptrs(F0, nf), cptr3(coords, D), index,
                                                    alpha, h);
    JP_LAUNCHED(ctx);
    return finish(ctx);
}

}  // namespace

extern "C" {

int jp_grid2particle_flip(jp_ctx* ctx, const jp_cellgrid* g, int nfields, double* const* Fp,
                          const double* const* F, const double* const* F0,
                          const double* const* coords, const uint8_t* index, double alpha) {
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if (nfields < 1 || nfields > JP_MAX_ARGS) return fail(ctx, JP_ERR_ARG, "nfields out of range%s");
    if ((rc = check_arrays(ctx, g->ndim, (const void* const*)coords, "coords"))) return rc;
    if ((rc = check_arrays(ctx, nfields, (const void* const*)Fp, "Fp"))) return rc;
    if ((rc = check_arrays(ctx, nfields, (const void* const*)F, "F"))) return rc;
    if ((rc = check_arrays(ctx, nfields, (const void* const*)F0, "F0"))) return rc;
    if (!index) return fail(ctx, JP_ERR_ARG, "null index%s");
    return g->ndim == 2 ? run_flip<2>(ctx, g, nfields, Fp, F, F0, coords, index, alpha)
                        : run_flip<3>(ctx, g, nfields, Fp, F, F0, coords, index, alpha);
}

int jp_particle2centroid(jp_ctx* ctx, const jp_cellgrid* g, const jp_axis* xc, int nfields,
                         double* const* F, const double* const* Fp, const double* const* coords,
                         int tuple) {
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if (nfields < 1 || nfields > JP_MAX_ARGS || (!tuple && nfields != 1))
        return fail(ctx, JP_ERR_ARG, "nfields out of range (1 unless tuple)%s");
    if ((rc = check_centres(ctx, g, xc))) return rc;
    if ((rc = check_arrays(ctx, g->ndim, (const void* const*)coords, "coords"))) return rc;
    if ((rc = check_arrays(ctx, nfields, (const void* const*)F, "F"))) return rc;
    if ((rc = check_arrays(ctx, nfields, (const void* const*)Fp, "Fp"))) return rc;
    if (g->ndim == 2) {
        Grid<2> G = to_grid<2>(g);
        k_p2c<2><<<(unsigned)blocks_for(G.C), kBlock, 0, ctx->stream>>>(
            G, caxes<2>(xc), nfields, argptrs(F, nfields), fieldptrs(Fp, nfields), cptr3(coords, 2),
            tuple);
    } else {
        Grid<3> G = to_grid<3>(g);
        k_p2c<3><<<(unsigned)blocks_for(G.C), kBlock, 0, ctx->stream>>>(
            G, caxes<3>(xc), nfields, argptrs(F, nfields), fieldptrs(Fp, nfields), cptr3(coords, 3),
            tuple);
    }
    JP_LAUNCHED(ctx);
    return finish(ctx);
}

int jp_centroid2particle(jp_ctx* ctx, const jp_cellgrid* g, const jp_axis* xc, int nfields,
                         double* const* Fp, const double* const* F, const double* const* coords) {
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if (nfields < 1 || nfields > JP_MAX_ARGS) return fail(ctx, JP_ERR_ARG, "nfields out of range%s");
    if ((rc = check_centres(ctx, g, xc))) return rc;
    for (int d = 0; d < g->ndim; ++d)
        if (g->ni[d] < 2) return fail(ctx, JP_ERR_ARG, "centroid2particle! needs two cells per dimension%s");
    if ((rc = check_arrays(ctx, g->ndim, (const void* const*)coords, "coords"))) return rc;
    if ((rc = check_arrays(ctx, nfields, (const void* const*)F, "F"))) return rc;
    if ((rc = check_arrays(ctx, nfields, (const void* const*)Fp, "Fp"))) return rc;
    if (g->ndim == 2) {
        Grid<2> G = to_grid<2>(g);
        k_c2p<2><<<dim3((unsigned)row_blocks<2>(G), (unsigned)G.K), kBlock, 0, ctx->stream>>>(
            G, caxes<2>(xc), nfields, argptrs(Fp, nfields), fieldptrs(F, nfields), cptr3(coords, 2));
    } else {
        Grid<3> G = to_grid<3>(g);
        k_c2p<3><<<dim3((unsigned)row_blocks<3>(G), (unsigned)G.K), kBlock, 0, ctx->stream>>>(
            G, caxes<3>(xc), nfields, argptrs(Fp, nfields), fieldptrs(F, nfields), cptr3(coords, 3));
    }
    JP_LAUNCHED(ctx);
    return finish(ctx);
}

}  // extern "C"
