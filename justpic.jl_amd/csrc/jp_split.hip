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

using namespace jp;

namespace {

struct FPtr3 {
    float* p[3];
};

// scaled velocity: V'_c = V_c * dt / h_c (cell displacement per step), float
template <int D>
struct SVel {
    const float* v[D];
    int n[D][D];  // dims of component c's grid
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

__global__ void k_scale(const double* __restrict__ V, float* __restrict__ Vs, int64_t n,
                        double dt, const double* __restrict__ h) {
    const double s = dt / h[0];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        Vs[i] = (float)(V[i] * s);
}

// ------------------------------------------------------------ advection
// velocity of component c at cell I + offset u (u in cell units, any value)
template <int D>
__device__ inline float interp_split(const SVel<D>& V, const SGrid<D>& g, int c, const int (&I)[D],
                                     const float (&u)[D]) {
    int k[D];
    float t[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        if (!(fabsf(u[d]) < 1.0e6f)) return __builtin_inff();
        const float s = floorf(u[d]);
        const int J = I[d] + (int)s;
        const float f = u[d] - s;
        if (d == c) {  // vertex axis of the component: limits [0, ni] cells
            if (J < 0 || J >= g.ni[d]) return __builtin_inff();
            k[d] = J;
            t[d] = f;
        } else {  // centre axis with ghosts: limits [-1/2, ni + 1/2] cells
            const bool up = f >= 0.5f;
            k[d] = J + (up ? 1 : 0);
            if (k[d] < 0 || k[d] > g.ni[d]) return __builtin_inff();
            t[d] = up ? f - 0.5f : f + 0.5f;
        }
    }
    const float* __restrict__ F = V.v[c];
    const int n0 = V.n[c][0];
    if constexpr (D == 2) {
        const int b = k[0] + n0 * k[1];
        return lerp1f(t[1], lerp1f(t[0], F[b], F[b + 1]), lerp1f(t[0], F[b + n0], F[b + n0 + 1]));
    } else {
        const int n1 = V.n[c][1], s2 = n0 * n1;
        const int b = k[0] + n0 * (k[1] + n1 * k[2]);
        const float lo =
            lerp1f(t[1], lerp1f(t[0], F[b], F[b + 1]), lerp1f(t[0], F[b + n0], F[b + n0 + 1]));
        const float hi = lerp1f(t[1], lerp1f(t[0], F[b + s2], F[b + s2 + 1]),
                                lerp1f(t[0], F[b + s2 + n0], F[b + s2 + n0 + 1]));
        return lerp1f(t[2], lo, hi);
    }
}

struct RKf {
    float alpha;
    int half;
    float w0, w1;
};

template <int D, int M>
__global__ void __launch_bounds__(kBlock) k_adv_split(Grid<D> gg, SGrid<D> g, SVel<D> V,
                                                      FPtr3 off, const uint8_t* __restrict__ index,
                                                      RKf P) {
    int64_t c, I64[D];
    if (!row_cell<D>(gg, c, I64)) return;
    const int64_t e = c + g.C * (int64_t)blockIdx.y;
    if (!index[e]) return;
    int I[D];
    float u0[D], v0[D], out[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        I[d] = (int)I64[d];
        u0[d] = off.p[d][e];
    }
#pragma unroll
    for (int cc = 0; cc < D; ++cc) v0[cc] = interp_split<D>(V, g, cc, I, u0);
    if constexpr (M == JP_EULER) {
#pragma unroll
        for (int d = 0; d < D; ++d) out[d] = u0[d] + v0[d];
    } else if constexpr (M == JP_RK2) {
        float u1[D], v1[D];
#pragma unroll
        for (int d = 0; d < D; ++d) u1[d] = fmaf(P.alpha, v0[d], u0[d]);
#pragma unroll
        for (int cc = 0; cc < D; ++cc) v1[cc] = interp_split<D>(V, g, cc, I, u1);
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
        for (int cc = 0; cc < D; ++cc) k2[cc] = interp_split<D>(V, g, cc, I, q);
#pragma unroll
        for (int d = 0; d < D; ++d) q[d] = u0[d] + 0.5f * k2[d];
#pragma unroll
        for (int cc = 0; cc < D; ++cc) k3[cc] = interp_split<D>(V, g, cc, I, q);
#pragma unroll
        for (int d = 0; d < D; ++d) q[d] = u0[d] + k3[d];
#pragma unroll
        for (int cc = 0; cc < D; ++cc) k4[cc] = interp_split<D>(V, g, cc, I, q);
#pragma unroll
        for (int d = 0; d < D; ++d)
            out[d] = u0[d] + (((v0[d] + 2.0f * k2[d]) + 2.0f * k3[d]) + k4[d]) / 6.0f;
    }
#pragma unroll
    for (int d = 0; d < D; ++d) off.p[d][e] = out[d];
}

// ------------------------------------------------------------ particle2grid
// Inverse-distance-squared weights in float32 from the offsets, weight and
// weighted-field sums in FP64; cells and slots in the reference order.
template <int D>
__global__ void __launch_bounds__(kBlock) k_p2g_split(Grid<D> gg, SGrid<D> g, double* __restrict__ F,
                                                      const double* __restrict__ Fp, FPtr3 off,
                                                      const uint8_t* __restrict__ index) {
    int64_t nn[D];
    int64_t NN = 1;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        nn[d] = g.ni[d] + 1;
        NN *= nn[d];
    }
    const int64_t node = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (node >= NN) return;
    int N[D];
    {
        int64_t r = node;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            N[d] = (int)(r % nn[d]);
            r /= nn[d];
        }
    }
    float hf[D];
#pragma unroll
    for (int d = 0; d < D; ++d) hf[d] = (float)g.h[d][0];
    const int64_t C = g.C;
    double W = 0.0, WF = 0.0;
    constexpr int NC = 1 << D;
#pragma unroll 1
    for (int q = 0; q < NC; ++q) {
        int64_t I[D];
        float a[D];  // node - cell corner, in cells: 0 or 1
        bool ok = true;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int o = ((q >> d) & 1) ? 0 : 1;
            I[d] = N[d] - o;
            a[d] = (float)o;
            ok = ok && (I[d] >= 0 && I[d] < g.ni[d]);
        }
        if (!ok) continue;
        const int64_t c = lin<D>(gg.ni, I);
        const uint64_t bq = active_slots(index, c, C, g.K);
        constexpr int B = 16;
        for (int k0 = 0; k0 < g.K; k0 += B) {
            float ub[B][D];
            double fb[B];
#pragma unroll
            for (int j = 0; j < B; ++j) {
                const bool on = (bq >> (k0 + j)) & 1;
                const int64_t e = c + C * (k0 + j);
#pragma unroll
                for (int d = 0; d < D; ++d) ub[j][d] = on ? off.p[d][e] : 0.0f;
                fb[j] = on ? Fp[e] : 0.0;
            }
#pragma unroll
            for (int j = 0; j < B; ++j) {
                if (!((bq >> (k0 + j)) & 1)) continue;
                float s = 0.0f;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const float r = (a[d] - ub[j][d]) * hf[d];
                    s = fmaf(r, r, s);
                }
                const double w = (double)(1.0f / s);
                W += w;
                WF = fma(w, fb[j], WF);
            }
        }
    }
    F[node] = WF / W;
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
                                                           int min_q, int32_t* __restrict__ lists,
                                                           int32_t* __restrict__ counts,
                                                           int64_t cap) {
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= g.C) return;
    int64_t I[D];
    unlin<D>(g.ni, c, I);
    const uint64_t act = active_slots(index, c, g.C, g.K);
    int num = 0;
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
            bool inside = (act >> (k0 + j)) & 1;
#pragma unroll
            for (int d = 0; d < D; ++d) inside = inside & (0.0f < ub[j][d] && ub[j][d] < 0.5f);
            num += inside ? 1 : 0;
        }
    }
    if (num >= min_q) return;
    int col = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) col = col * 2 + (int)(I[d] & 1);
    const int32_t slot = atomicAdd(&counts[col], 1);
    lists[col * cap + slot] = (int32_t)c;
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
    const int64_t w = blockIdx.x * (int64_t)(kBlock / 64) + (threadIdx.x >> 6);
    if (w >= *count) return;
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
    int64_t n_inj = 0, n_orph = 0;
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
            for (int nb = 0; nb < NB; ++nb) {
                int64_t J[D];
                int o[D];
                bool ok = own;
                int r = nb;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    o[d] = (r % 3) - 1;
                    J[d] = I[d] + o[d];
                    r /= 3;
                    ok = ok && J[d] >= 0 && J[d] < g.ni[d];
                }
                if (!ok) continue;
                bool a;
                float uq[D];
                if (nb == CENTER) {
                    a = act && lane != ip;
#pragma unroll
                    for (int d = 0; d < D; ++d) uq[d] = u[d];
                } else {
                    const int64_t e2 = lin<D>(g.ni, J) + C * lane;
                    a = index[e2] != 0;
#pragma unroll
                    for (int d = 0; d < D; ++d) uq[d] = a ? off.p[d][e2] : 0.0f;
                }
                if (!a) continue;
                double s = 0.0;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const double r2 = ((double)o[d] + (double)uq[d] - (double)un[d]) * sg.h[d][0];
                    s = s + r2 * r2;
                }
                const double dd = sqrt(s);
                if (dd < best) {
                    best = dd;
                    border = nb * 64 + lane;
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
    if (lane == 0) {
        stat_add(st, ST_INJECTED, n_inj);
        stat_add(st, ST_ORPHAN, n_orph);
    }
}

// ------------------------------------------------------------ host side
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
    int rc;
    SGrid<D> sg = sgrid<D>(G);
    SVel<D> sv;
    size_t tot = 0, offs[3];
    int64_t cnt[3];
    for (int c = 0; c < D; ++c) {
        int64_t n = 1;
        for (int d = 0; d < D; ++d) {
            sv.n[c][d] = (int)vg->ax[c][d].n;
            n *= vg->ax[c][d].n;
            // standard staggering: the vertex axis along c, centres + ghosts elsewhere
            const int64_t want = d == c ? G.ni[d] + 1 : G.ni[d] + 2;
            if (vg->ax[c][d].n != want)
                return fail(ctx, JP_ERR_ARG,
                            "split advection needs the standard staggered grids%s (component %lld)",
                            "", c);
        }
        offs[c] = tot;
        cnt[c] = n;
        tot += (size_t)n * 4;
        tot = (tot + 255) & ~(size_t)255;
    }
    if ((rc = ensure_scratch(ctx, tot))) return rc;
    for (int c = 0; c < D; ++c) {
        float* vs = reinterpret_cast<float*>(static_cast<char*>(ctx->scratch) + offs[c]);
        int64_t b = blocks_for(cnt[c]);
        k_scale<<<b > 65536 ? 65536 : b, kBlock, 0, ctx->stream>>>(V[c], vs, cnt[c], dt,
                                                                    gc->xv[c].dx);
        JP_LAUNCHED(ctx);
        sv.v[c] = vs;
    }
    RKf P;
    P.alpha = (float)alpha;
    P.half = alpha == 0.5;
    const double ia = 1.0 / alpha;
    P.w0 = (float)std::fma(-0.5, ia, 1.0);
    P.w1 = (float)(0.5 * ia);
    dim3 grid((unsigned)row_blocks<D>(G), (unsigned)G.K);
    FPtr3 fp = fptr3<D>(off);
    if (method == JP_EULER)
        k_adv_split<D, JP_EULER><<<grid, kBlock, 0, ctx->stream>>>(G, sg, sv, fp, index, P);
    else if (method == JP_RK2)
        k_adv_split<D, JP_RK2><<<grid, kBlock, 0, ctx->stream>>>(G, sg, sv, fp, index, P);
    else
        k_adv_split<D, JP_RK4><<<grid, kBlock, 0, ctx->stream>>>(G, sg, sv, fp, index, P);
    JP_LAUNCHED(ctx);
    return finish(ctx);
}

template <int D>
int run_p2g(jp_ctx* ctx, const jp_cellgrid* gc, double* F, const double* Fp, float* const* off,
            const uint8_t* index) {
    Grid<D> G = to_grid<D>(gc);
    SGrid<D> sg = sgrid<D>(G);
    int64_t NN = 1;
    for (int d = 0; d < D; ++d) NN *= G.ni[d] + 1;
    k_p2g_split<D><<<blocks_for(NN), kBlock, 0, ctx->stream>>>(G, sg, F, Fp, fptr3<D>(off), index);
    JP_LAUNCHED(ctx);
    return finish(ctx);
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
    const size_t counts_bytes = 256;
    if ((rc = ensure_scratch(ctx, counts_bytes + (size_t)NQ * cap * 4))) return rc;
    int32_t* counts = static_cast<int32_t*>(ctx->scratch);
    int32_t* lists = reinterpret_cast<int32_t*>(static_cast<char*>(ctx->scratch) + counts_bytes);
    if ((rc = zero_bytes(ctx, counts, counts_bytes))) return rc;
    FPtr3 fp = fptr3<D>(off);
    k_inj_scan_split<D><<<blocks_for(G.C), kBlock, 0, ctx->stream>>>(G, fp, index, min_q, lists,
                                                                     counts, cap);
    JP_LAUNCHED(ctx);
    ArgPtrs ap;
    for (int i = 0; i < JP_MAX_ARGS; ++i) ap.p[i] = i < nargs ? args[i] : nullptr;
    jp_stats* st = stats ? dstats(ctx) : nullptr;
    for (int ph = 0; ph < NQ; ++ph) {
        k_inj_wave_split<D><<<blocks_for(cap, kBlock / 64), kBlock, 0, ctx->stream>>>(
            G, sg, fp, index, nargs, ap, min_q, seed, counter, lists + (int64_t)ph * cap,
            counts + ph, st);
        JP_LAUNCHED(ctx);
    }
    if ((rc = stats_end(ctx, stats))) return rc;
    return finish(ctx);
}

}  // namespace

extern "C" {

int jp_split_from_f64(jp_ctx* ctx, const jp_cellgrid* g, const double* const* coords,
                      const uint8_t* index, float* const* off) {
    int rc = check_common(ctx, g, (const void* const*)off, index);
    if (rc) return rc;
    if (!coords) return fail(ctx, JP_ERR_ARG, "null coords%s");
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

}  // extern "C"
