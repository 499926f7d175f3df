/*
 * jp_oracle.c -- sequential FP64 restatement of JustPIC.jl v0.6.7's particle
 * hot path.  TEST INFRASTRUCTURE ONLY (parity checker and CPU baseline); the
 * product library never links or calls this file.
 *
 * Build with -ffp-contract=off: every fused multiply-add below is an explicit
 * fma() at the sites where the reference calls fma/muladd (or @muladd); every
 * other operation is a single correctly rounded IEEE-754 binary64 operation in
 * the reference's evaluation order.
 *
 * Indexing: the reference is 1-based (Julia); this file keeps 1-based cell and
 * slot numbers where the algorithm's control flow depends on them (bisection,
 * starting_point) and converts to 0-based only when addressing memory.
 */
#include "jp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NDIM_MAX 3

/* ------------------------------------------------------------------ */
/* primitives                                                          */
/* ------------------------------------------------------------------ */

/* lerp(t, v0, v1) = muladd(t, v1, muladd(-t, v0, v0))
 * src/Interpolations/ndlerp.jl:15 */
double jpo_lerp1(double t, double v0, double v1) {
    return fma(t, v1, fma(-t, v0, v0));
}

/* lerp(v, t::NTuple{2}) = L(t2, L(t1,v1,v2), L(t1,v3,v4))
 * src/Interpolations/ndlerp.jl:11-14 */
double jpo_lerp2(const double v[4], const double t[2]) {
    return jpo_lerp1(t[1], jpo_lerp1(t[0], v[0], v[1]), jpo_lerp1(t[0], v[2], v[3]));
}

/* lerp(v, t::NTuple{3}) = L(t3, lerp2(v[1:4]), lerp2(v[5:8]))
 * src/Interpolations/ndlerp.jl:11-14 */
double jpo_lerp3(const double v[8], const double t[3]) {
    return jpo_lerp1(t[2], jpo_lerp2(v, t), jpo_lerp2(v + 4, t));
}

/* find_parent_cell_bisection(px, x, lo=1, hi=length(x), seed), 1-based
 * src/Utils.jl:108-130.  Returns -1 instead of looping forever when px is
 * outside [x[1], x[end]] (the reference never calls it there: advection checks
 * the limits first, move checks the domain first). */
int64_t jpo_bisect(double px, const double* x, int64_t n, int64_t seed) {
    int64_t lo = 1, hi = n;
    if (!(x[0] <= px && px <= x[n - 1])) return -1;
    for (int64_t it = 0; it < 4 * 64 + n; ++it) {
        if (x[seed - 1] <= px && px <= x[seed]) return seed;
        if (x[seed - 1] < px) {
            lo = seed;
            seed = (hi + seed) / 2;
        } else {
            hi = seed;
            seed = (lo + seed) / 2;
        }
    }
    return -1;
}

/* first_stage: Euler  p + s*dt*v          (src/Advection/Euler.jl:1-6)
 *              RK2    p + s*alpha*dt*v    (src/Advection/RK2.jl:7-20)
 * under @muladd with s = 1 (no backtracking).  MuladdMacro splits the product
 * into (all factors but the last) x (last factor): fma(s*alpha*dt, v, p).
 * The factor split is parity unpinned (no Julia here); see the fused/unfused
 * note in the design notes. */
void jpo_first_stage(int method, double alpha, double dt, int ndim,
                     const double* v, const double* p, double* out) {
    const double s = 1.0;
    double f = (method == JP_EULER) ? s * dt : (s * alpha) * dt;
    for (int d = 0; d < ndim; ++d) out[d] = fma(f, v[d], p[d]);
}

/* second_stage (src/Advection/RK2.jl:22-44):
 *   alpha == 0.5 : p + s*dt*v1
 *   else         : p + s*dt*((1 - 0.5*inv(alpha))*v0 + 0.5*inv(alpha)*v1)
 * with @muladd: muladd(s*dt, muladd(0.5*inv(a), v1, (1-0.5*inv(a))*v0), p),
 * 1 - 0.5*inv(a) -> muladd(-0.5, inv(a), 1). */
void jpo_second_stage(double alpha, double dt, int ndim, const double* v0,
                      const double* v1, const double* p, double* out) {
    const double s = 1.0, half = 0.5;
    if (alpha == half) {
        for (int d = 0; d < ndim; ++d) out[d] = fma(s * dt, v1[d], p[d]);
    } else {
        double ia = 1.0 / alpha;
        double w0 = fma(-half, ia, 1.0);
        double w1 = half * ia;
        for (int d = 0; d < ndim; ++d) {
            double inner = fma(w1, v1[d], w0 * v0[d]);
            out[d] = fma(s * dt, inner, p[d]);
        }
    }
}

/* distance(a, b) = sqrt(((a1-b1)^2 + (a2-b2)^2) + (a3-b3)^2)
 * src/Interpolations/utils.jl:9-19 (x^2 lowers to x*x) */
double jpo_distance(int ndim, const double* a, const double* b) {
    double e = a[0] - b[0];
    double s = e * e;
    for (int d = 1; d < ndim; ++d) {
        e = a[d] - b[d];
        s = s + e * e;
    }
    return sqrt(s);
}

/* cell_index(x, xv) = abs(trunc((x - x0)/dx)) + 1, dx = |xv[2]-xv[1]|,
 * x0 subtracted only when nonzero.  src/particles.jl:205-217 */
int64_t jpo_cell_index(double x, const double* xv, int64_t n) {
    (void)n;
    double dx = fabs(xv[1] - xv[0]);
    double x0 = xv[0];
    double q = (x0 == 0.0) ? x / dx : (x - x0) / dx;
    int64_t t = (int64_t)trunc(q);
    return (t < 0 ? -t : t) + 1;
}

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
 * as 1, 2, 3", SC11).  Replaces Julia's task-local rand(), which is not
 * reproducible (src/Particles/particles_utils.jl:189, injection.jl:414). */
void jpo_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* One uniform double in [0,1) with 53 random bits.  Counter block:
 * (cell, slot | dimblock<<8 | purpose<<16, counter lo, counter hi),
 * key = seed; dims 0,1 come from block 0 (words 0-1 and 2-3), dim 2 from
 * block 1 (words 0-1). */
double jpo_uniform(uint64_t seed, uint64_t cell, int slot, int dim, int purpose,
                   uint64_t counter) {
    uint32_t ctr[4], key[2], out[4];
    int block = dim / 2, half = dim % 2;
    ctr[0] = (uint32_t)cell;
    ctr[1] = (uint32_t)slot | ((uint32_t)block << 8) | ((uint32_t)purpose << 16);
    ctr[2] = (uint32_t)counter;
    ctr[3] = (uint32_t)(counter >> 32);
    key[0] = (uint32_t)seed;
    key[1] = (uint32_t)(seed >> 32);
    jpo_philox(ctr, key, out);
    uint64_t w = ((uint64_t)out[2 * half] << 32) | out[2 * half + 1];
    return (double)(w >> 11) * 0x1.0p-53;
}

/* ------------------------------------------------------------------ */
/* helpers                                                             */
/* ------------------------------------------------------------------ */

typedef struct {
    int D;
    int K;
    int64_t ni[NDIM_MAX];
    int64_t C;
} shape_t;

static shape_t shape_of(const jp_cellgrid* g) {
    shape_t s;
    s.D = g->ndim;
    s.K = g->max_xcell;
    s.C = 1;
    for (int d = 0; d < NDIM_MAX; ++d) s.ni[d] = 1;
    for (int d = 0; d < s.D; ++d) {
        s.ni[d] = g->ni[d];
        s.C *= g->ni[d];
    }
    return s;
}

static int check_grid(const jp_cellgrid* g) {
    if (!g || (g->ndim != 2 && g->ndim != 3)) return JP_ERR_ARG;
    if (g->max_xcell < 1 || g->max_xcell > 64) return JP_ERR_ARG;
    for (int d = 0; d < g->ndim; ++d)
        if (g->ni[d] < 1 || g->xv[d].n != g->ni[d] + 1) return JP_ERR_ARG;
    return JP_OK;
}

/* 0-based cell tuple -> linear cell index (column-major) */
static int64_t lin(const shape_t* s, const int64_t* I) {
    int64_t c = I[s->D - 1];
    for (int d = s->D - 2; d >= 0; --d) c = c * s->ni[d] + I[d];
    return c;
}

static void unlin(const shape_t* s, int64_t c, int64_t* I) {
    for (int d = 0; d < s->D; ++d) {
        I[d] = c % s->ni[d];
        c /= s->ni[d];
    }
}

/* ------------------------------------------------------------------ */
/* init_particles                                                      */
/* ------------------------------------------------------------------ */

/* fill_coords_index! (src/Particles/particles_utils.jl:168-194): for each
 * quadrant iq in quadrant_masks order (:224-240, x offset fastest) and each of
 * np_quadrant particles: p = (x0 + dx*mask/2) + dx/2*rand().  Coordinates of
 * unused slots are NaN and their mask false (:157-158). */
int jpo_init_particles(const jp_cellgrid* g, double* const* coords,
                       uint8_t* index, int np_quadrant, uint64_t seed) {
    if (check_grid(g)) return JP_ERR_ARG;
    shape_t s = shape_of(g);
    int nq = (s.D == 2) ? 4 : 8;
    if (np_quadrant < 0 || np_quadrant * nq > s.K) return JP_ERR_ARG;
    int64_t S = s.C * s.K;
    for (int64_t e = 0; e < S; ++e) {
        index[e] = 0;
        for (int d = 0; d < s.D; ++d) coords[d][e] = NAN;
    }
    for (int64_t c = 0; c < s.C; ++c) {
        int64_t I[NDIM_MAX];
        unlin(&s, c, I);
        double x0[NDIM_MAX], dx[NDIM_MAX];
        for (int d = 0; d < s.D; ++d) {
            x0[d] = g->xv[d].x[I[d]];
            dx[d] = g->xv[d].dx[I[d]];
        }
        int l = 0;
        for (int iq = 0; iq < nq; ++iq) {
            double xc[NDIM_MAX];
            for (int d = 0; d < s.D; ++d) {
                int m = (iq >> d) & 1;
                xc[d] = x0[d] + (dx[d] * (double)m) / 2.0;
            }
            for (int n = 0; n < np_quadrant; ++n) {
                for (int d = 0; d < s.D; ++d) {
                    double r = jpo_uniform(seed, (uint64_t)c, l, d, 1, 0);
                    coords[d][c + s.C * l] = xc[d] + (dx[d] / 2.0) * r;
                }
                index[c + s.C * l] = 1;
                ++l;
            }
        }
    }
    return JP_OK;
}

/* ------------------------------------------------------------------ */
/* advection!                                                          */
/* ------------------------------------------------------------------ */

/* check_local_limits (src/Advection/advection.jl:83-91): inclusive, all dims,
 * limits = (first, last) of each axis (inner_limits :76-81). */
static int check_limits(const jp_axis* ax, int D, const double* p) {
    for (int d = 0; d < D; ++d) {
        double lo = ax[d].x[0], hi = ax[d].x[ax[d].n - 1];
        if (!(lo <= p[d] && p[d] <= hi)) return 0;
    }
    return 1;
}

/* interp_velocity2particle innermost (src/Particles/Advection/advection.jl:
 * 113-148): bisection seeded at the particle's cell I, corner gather
 * (extract_field_corners, src/Advection/advection.jl:47-71), t = (p - x)*inv(dx)
 * (normalize_coordinates, src/Interpolations/utils.jl:54-58), lerp. */
static double interp_component(const jp_axis* ax, int D, const double* F,
                               const double* p, const int64_t* I1) {
    int64_t k[NDIM_MAX];
    double t[NDIM_MAX];
    for (int d = 0; d < D; ++d) {
        k[d] = jpo_bisect(p[d], ax[d].x, ax[d].n, I1[d]); /* 1-based */
        double xc = ax[d].x[k[d] - 1];
        double dxd = ax[d].dx[k[d] - 1];
        t[d] = (p[d] - xc) * (1.0 / dxd);
    }
    int64_t n0 = ax[0].n, n1 = ax[1].n;
    int64_t i = k[0] - 1, j = k[1] - 1;
    if (D == 2) {
        double v[4];
        v[0] = F[i + n0 * j];
        v[1] = F[(i + 1) + n0 * j];
        v[2] = F[i + n0 * (j + 1)];
        v[3] = F[(i + 1) + n0 * (j + 1)];
        return jpo_lerp2(v, t);
    } else {
        int64_t l = k[2] - 1;
        double v[8];
        for (int c = 0; c < 8; ++c) {
            int64_t a = i + (c & 1), b = j + ((c >> 1) & 1), e = l + ((c >> 2) & 1);
            v[c] = F[a + n0 * (b + n1 * e)];
        }
        return jpo_lerp3(v, t);
    }
}

/* interp_velocity2particle outer (src/Particles/Advection/advection.jl:93-110):
 * +Inf for a component whose velocity grid does not contain p. */
static void interp_velocity(const jp_velgrid* vg, const double* const* V, int D,
                            const double* p, const int64_t* I1, double* v) {
    for (int c = 0; c < D; ++c) {
        if (check_limits(vg->ax[c], D, p))
            v[c] = interp_component(vg->ax[c], D, V[c], p, I1);
        else
            v[c] = INFINITY;
    }
}

/* advect_particle for Euler (Euler.jl:1-20), RK2 (RK2.jl:1-26), RK4
 * (RK4.jl:1-19).  RK4 has no muladd: k2 = interp(p0 + (dt*k1)/2), ...,
 * p = p0 + (dt*(((k1 + 2k2) + 2k3) + k4))/6. */
static void advect_particle(int method, double alpha, double dt, int D,
                            const jp_velgrid* vg, const double* const* V,
                            const double* p0, const int64_t* I1, double* out) {
    double v0[NDIM_MAX], v1[NDIM_MAX], p1[NDIM_MAX];
    if (method == JP_EULER) {
        interp_velocity(vg, V, D, p0, I1, v0);
        jpo_first_stage(JP_EULER, alpha, dt, D, v0, p0, out);
    } else if (method == JP_RK2) {
        interp_velocity(vg, V, D, p0, I1, v0);
        jpo_first_stage(JP_RK2, alpha, dt, D, v0, p0, p1);
        interp_velocity(vg, V, D, p1, I1, v1);
        jpo_second_stage(alpha, dt, D, v0, v1, p0, out);
    } else {
        double k1[NDIM_MAX], k2[NDIM_MAX], k3[NDIM_MAX], k4[NDIM_MAX], q[NDIM_MAX];
        interp_velocity(vg, V, D, p0, I1, k1);
        for (int d = 0; d < D; ++d) q[d] = p0[d] + (dt * k1[d]) / 2.0;
        interp_velocity(vg, V, D, q, I1, k2);
        for (int d = 0; d < D; ++d) q[d] = p0[d] + (dt * k2[d]) / 2.0;
        interp_velocity(vg, V, D, q, I1, k3);
        for (int d = 0; d < D; ++d) q[d] = p0[d] + dt * k3[d];
        interp_velocity(vg, V, D, q, I1, k4);
        for (int d = 0; d < D; ++d) {
            double sum = ((k1[d] + 2.0 * k2[d]) + 2.0 * k3[d]) + k4[d];
            out[d] = p0[d] + (dt * sum) / 6.0;
        }
    }
}

/* advection_kernel! (src/Particles/Advection/advection.jl:66-91): every cell,
 * every active slot, advect and write back. */
int jpo_advection(const jp_cellgrid* g, double* const* coords,
                  const uint8_t* index, const double* const* V,
                  const jp_velgrid* vg, int method, double alpha, double dt) {
    if (check_grid(g)) return JP_ERR_ARG;
    if (method < JP_EULER || method > JP_RK4) return JP_ERR_ARG;
    if (method == JP_RK2 && !(0.0 < alpha && alpha < 1.0)) return JP_ERR_ARG;
    shape_t s = shape_of(g);
    for (int64_t c = 0; c < s.C; ++c) {
        int64_t I[NDIM_MAX], I1[NDIM_MAX];
        unlin(&s, c, I);
        for (int d = 0; d < s.D; ++d) I1[d] = I[d] + 1;
        for (int ip = 0; ip < s.K; ++ip) {
            int64_t e = c + s.C * ip;
            if (!index[e]) continue; /* doskip, src/Utils.jl:32-33 */
            double p0[NDIM_MAX], p[NDIM_MAX];
            for (int d = 0; d < s.D; ++d) p0[d] = coords[d][e];
            advect_particle(method, alpha, dt, s.D, vg, V, p0, I1, p);
            for (int d = 0; d < s.D; ++d) coords[d][e] = p[d];
        }
    }
    return JP_OK;
}

/* ------------------------------------------------------------------ */
/* move_particles!                                                     */
/* ------------------------------------------------------------------ */

/* move_kernel! (src/Particles/move_safe.jl:72-125) for one source cell I
 * (0-based).  isincell uses the scalar/vector vertex spacing
 * (src/Particles/utils.jl:15), indomain is strict against extrema of the
 * vertex axes (move_safe.jl:199-206), the new cell comes from bisection
 * seeded at the source cell (:104), find_free_memory starts at starting_point
 * (:115, :192-197), and a full destination drops the particle (:116). */
static void move_cell(const jp_cellgrid* g, const shape_t* s,
                      double* const* coords, uint8_t* index, int nargs,
                      double* const* args, const int64_t* I, jp_stats* st) {
    int D = s->D, K = s->K;
    int64_t C = s->C;
    int64_t c = lin(s, I);
    double corner[NDIM_MAX], dxi[NDIM_MAX], lo[NDIM_MAX], hi[NDIM_MAX];
    for (int d = 0; d < D; ++d) {
        corner[d] = g->xv[d].x[I[d]];
        dxi[d] = g->xv[d].dx[I[d]];
        lo[d] = g->xv[d].x[0];
        hi[d] = g->xv[d].x[g->xv[d].n - 1];
    }
    int starting_point = 1;
    for (int ip = 1; ip <= K; ++ip) {
        int64_t e = c + C * (ip - 1);
        if (!index[e]) continue;
        double p[NDIM_MAX];
        for (int d = 0; d < D; ++d) p[d] = coords[d][e];
        int incell = 1;
        for (int d = 0; d < D; ++d)
            incell = incell & (corner[d] < p[d] && p[d] < corner[d] + dxi[d]);
        if (incell) continue;
        int indomain = 1;
        for (int d = 0; d < D; ++d)
            if (!(lo[d] < p[d] && p[d] < hi[d])) { indomain = 0; break; }
        if (!indomain) {
            index[e] = 0;
            for (int d = 0; d < D; ++d) coords[d][e] = NAN;
            for (int a = 0; a < nargs; ++a) args[a][e] = NAN;
            if (st) st->deleted++;
            continue;
        }
        int64_t nc[NDIM_MAX];
        int far = 0;
        for (int d = 0; d < D; ++d) {
            nc[d] = jpo_bisect(p[d], g->xv[d].x, g->xv[d].n, I[d] + 1) - 1;
            if (nc[d] - I[d] > 1 || I[d] - nc[d] > 1) far = 1;
        }
        if (far && st) st->far++;
        double cached[JP_MAX_ARGS];
        for (int a = 0; a < nargs; ++a) cached[a] = args[a][e];
        index[e] = 0;
        for (int d = 0; d < D; ++d) coords[d][e] = NAN;
        for (int a = 0; a < nargs; ++a) args[a][e] = NAN;
        int64_t cn = lin(s, nc);
        int free_idx = 0;
        for (int i = starting_point; i <= K; ++i)
            if (!index[cn + C * (i - 1)]) { free_idx = i; break; }
        if (free_idx == 0) {
            if (st) st->dropped++;
            continue;
        }
        starting_point = free_idx;
        int64_t f = cn + C * (free_idx - 1);
        index[f] = 1;
        for (int d = 0; d < D; ++d) coords[d][f] = p[d];
        for (int a = 0; a < nargs; ++a) args[a][f] = cached[a];
        if (st) st->moved++;
    }
}

/* move_particles! (src/Particles/move_safe.jl:23-49, move_particles_ps! :51-62):
 * 3^D colour phases in order (offset_i outer ... offset_last inner); within a
 * phase the cells 3(I-1)+offset are visited in column-major order (the
 * reference runs them in parallel; their neighbourhoods are disjoint whenever
 * no particle jumps more than one cell). */
int jpo_move_particles(const jp_cellgrid* g, double* const* coords,
                       uint8_t* index, int nargs, double* const* args,
                       jp_stats* stats) {
    if (check_grid(g)) return JP_ERR_ARG;
    if (nargs < 0 || nargs > JP_MAX_ARGS) return JP_ERR_ARG;
    shape_t s = shape_of(g);
    int D = s.D;
    int64_t ncol[NDIM_MAX] = {1, 1, 1};
    for (int d = 0; d < D; ++d) ncol[d] = (s.ni[d] + 2) / 3; /* ceil(ni/3) */
    int nphase = (D == 2) ? 9 : 27;
    for (int ph = 0; ph < nphase; ++ph) {
        int off[NDIM_MAX];
        /* for offset_1 in 1:3, offset_2 in 1:3 [, offset_3 in 1:3]: the last
         * offset varies fastest */
        if (D == 2) {
            off[0] = ph / 3 + 1;
            off[1] = ph % 3 + 1;
        } else {
            off[0] = ph / 9 + 1;
            off[1] = (ph / 3) % 3 + 1;
            off[2] = ph % 3 + 1;
        }
        int64_t T[NDIM_MAX] = {0, 0, 0};
        int64_t total = 1;
        for (int d = 0; d < D; ++d) total *= ncol[d];
        for (int64_t t = 0; t < total; ++t) {
            int64_t r = t;
            for (int d = 0; d < D; ++d) {
                T[d] = r % ncol[d];
                r /= ncol[d];
            }
            int64_t I[NDIM_MAX];
            int ok = 1;
            for (int d = 0; d < D; ++d) {
                I[d] = 3 * T[d] + off[d] - 1;
                if (I[d] >= s.ni[d]) ok = 0;
            }
            if (ok) move_cell(g, &s, coords, index, nargs, args, I, stats);
        }
    }
    return JP_OK;
}

/* ------------------------------------------------------------------ */
/* particle2grid!                                                      */
/* ------------------------------------------------------------------ */

/* _particle2grid! (src/Interpolations/particle_to_grid.jl:37-68 2D scalar,
 * :113-151 3D scalar, :70-109 / :153-199 tuples): per node, cells (i-1..i) x
 * (j-1..j) [x (k-1..k)] with the last dim outermost, slots ascending;
 * w = inv(distance(node, p)^2); w_sum += w; wF = fma/muladd(w, Fp, wF).
 * Scalar 2D divides (wF / w_sum), scalar 3D and all tuples multiply by
 * inv(w_sum).  A node with no particles gets 0/0 = NaN. */
static void p2g_core(const jp_cellgrid* g, int nfields, double* const* F,
                     const double* const* Fp, const double* const* coords,
                     const uint8_t* index, int use_div) {
    shape_t s = shape_of(g);
    int D = s.D, K = s.K;
    int64_t C = s.C;
    int64_t nn[NDIM_MAX] = {1, 1, 1};
    int64_t NN = 1;
    for (int d = 0; d < D; ++d) {
        nn[d] = s.ni[d] + 1;
        NN *= nn[d];
    }
    for (int64_t node = 0; node < NN; ++node) {
        int64_t N[NDIM_MAX] = {0, 0, 0};
        int64_t r = node;
        for (int d = 0; d < D; ++d) {
            N[d] = r % nn[d];
            r /= nn[d];
        }
        double xn[NDIM_MAX];
        for (int d = 0; d < D; ++d) xn[d] = g->xv[d].x[N[d]];
        double w = 0.0, wF[JP_MAX_ARGS];
        for (int f = 0; f < nfields; ++f) wF[f] = 0.0;
        int ncorner = (D == 2) ? 4 : 8;
        for (int q = 0; q < ncorner; ++q) {
            /* q bit d set -> offset 0, else -1; dim 0 fastest, last dim outermost */
            int64_t I[NDIM_MAX];
            int ok = 1;
            for (int d = 0; d < D; ++d) {
                I[d] = N[d] + (((q >> d) & 1) ? 0 : -1);
                if (!(0 <= I[d] && I[d] < nn[d] - 1)) ok = 0;
            }
            if (!ok) continue;
            int64_t c = lin(&s, I);
            for (int ip = 0; ip < K; ++ip) {
                int64_t e = c + C * ip;
                if (!index[e]) continue;
                double p[NDIM_MAX];
                for (int d = 0; d < D; ++d) p[d] = coords[d][e];
                double dist = jpo_distance(D, xn, p);
                double wi = 1.0 / (dist * dist);
                w += wi;
                for (int f = 0; f < nfields; ++f) wF[f] = fma(wi, Fp[f][e], wF[f]);
            }
        }
        if (use_div) {
            F[0][node] = wF[0] / w;
        } else {
            double iw = 1.0 / w;
            for (int f = 0; f < nfields; ++f) F[f][node] = wF[f] * iw;
        }
    }
}

int jpo_particle2grid(const jp_cellgrid* g, double* F, const double* Fp,
                      const double* const* coords, const uint8_t* index) {
    if (check_grid(g)) return JP_ERR_ARG;
    double* Fa[1] = {F};
    const double* Fpa[1] = {Fp};
    p2g_core(g, 1, Fa, Fpa, coords, index, g->ndim == 2);
    return JP_OK;
}

int jpo_particle2grid_multi(const jp_cellgrid* g, int nfields, double* const* F,
                            const double* const* Fp,
                            const double* const* coords, const uint8_t* index) {
    if (check_grid(g)) return JP_ERR_ARG;
    if (nfields < 1 || nfields > JP_MAX_ARGS) return JP_ERR_ARG;
    p2g_core(g, nfields, F, Fp, coords, index, 0);
    return JP_OK;
}

/* ------------------------------------------------------------------ */
/* grid2particle!                                                      */
/* ------------------------------------------------------------------ */

/* grid2particle_classic! (src/Interpolations/grid_to_particle.jl:38-82):
 * per cell, corners F[i,j],F[i+1,j],F[i,j+1],F[i+1,j+1] (field_corners,
 * src/Interpolations/utils.jl:98-104), t = (p - (xv[i],yv[j]))*inv(dx_cell),
 * Fp = lerp (_grid2particle :265-273) for every active slot. */
int jpo_grid2particle(const jp_cellgrid* g, double* Fp, const double* F,
                      const double* const* coords, const uint8_t* index) {
    if (check_grid(g)) return JP_ERR_ARG;
    shape_t s = shape_of(g);
    int D = s.D, K = s.K;
    int64_t C = s.C;
    int64_t n0 = s.ni[0] + 1, n1 = s.ni[1] + 1;
    for (int64_t c = 0; c < C; ++c) {
        int64_t I[NDIM_MAX];
        unlin(&s, c, I);
        double v[8], corner[NDIM_MAX], dxi[NDIM_MAX];
        int ncorner = (D == 2) ? 4 : 8;
        for (int q = 0; q < ncorner; ++q) {
            int64_t a = I[0] + (q & 1), b = I[1] + ((q >> 1) & 1);
            int64_t e = (D == 3) ? I[2] + ((q >> 2) & 1) : 0;
            v[q] = F[a + n0 * (b + n1 * e)];
        }
        for (int d = 0; d < D; ++d) {
            corner[d] = g->xv[d].x[I[d]];
            dxi[d] = g->xv[d].dx[I[d]];
        }
        for (int ip = 0; ip < K; ++ip) {
            int64_t e = c + C * ip;
            if (!index[e]) continue;
            double t[NDIM_MAX];
            for (int d = 0; d < D; ++d) t[d] = (coords[d][e] - corner[d]) * (1.0 / dxi[d]);
            Fp[e] = (D == 2) ? jpo_lerp2(v, t) : jpo_lerp3(v, t);
        }
    }
    return JP_OK;
}

/* ------------------------------------------------------------------ */
/* inject_particles!                                                   */
/* ------------------------------------------------------------------ */

/* index_min_distance (src/Particles/injection.jl:330-393): neighbourhood
 * cells with the last dim outermost and dim 0 innermost, then slots; skips the
 * slot being filled, inactive slots and (2D) NaN coordinates; strict '<'.
 * Returns the 0-based linear element index, or -1 when nothing qualifies
 * (the reference then reads args at slot 0: out of bounds). */
static int64_t index_min_distance(const shape_t* s, double* const* coords,
                                  const uint8_t* index, const double* pn,
                                  int current_slot, const int64_t* Icell) {
    int D = s->D, K = s->K;
    int64_t C = s->C;
    int64_t best = -1;
    double dist_min = INFINITY;
    int nb = (D == 2) ? 9 : 27;
    for (int q = 0; q < nb; ++q) {
        int64_t I[NDIM_MAX];
        int r = q, ok = 1, same = 1;
        for (int d = 0; d < D; ++d) {
            I[d] = Icell[d] + (r % 3) - 1;
            r /= 3;
            if (I[d] < 0 || I[d] >= s->ni[d]) ok = 0;
            if (I[d] != Icell[d]) same = 0;
        }
        if (!ok) continue;
        int64_t c = lin(s, I);
        for (int ip = 0; ip < K; ++ip) {
            if (same && ip == current_slot) continue;
            int64_t e = c + C * ip;
            if (!index[e]) continue;
            double p[NDIM_MAX];
            int hasnan = 0;
            for (int d = 0; d < D; ++d) {
                p[d] = coords[d][e];
                if (isnan(p[d])) hasnan = 1;
            }
            if (D == 2 && hasnan) continue;
            double dd = jpo_distance(D, p, pn);
            if (dd < dist_min) {
                best = e;
                dist_min = dd;
            }
        }
    }
    return best;
}

/* _inject_particles! (src/Particles/injection.jl:68-131) for one cell:
 * quadrants in quadrant_corners order (:442-461), count active particles
 * strictly inside the quadrant, BREAK out of the quadrant loop as soon as a
 * quadrant has enough (:100), otherwise fill free slots ascending with
 * vertex + dq*muladd(0.95, rand(), 0.05) (new_particle :411-417) and copy all
 * args from the nearest neighbour. */
static void inject_cell(const jp_cellgrid* g, const shape_t* s,
                        double* const* coords, uint8_t* index, int nargs,
                        double* const* args, int min_xcell, uint64_t seed,
                        uint64_t counter, const int64_t* I, jp_stats* st) {
    int D = s->D, K = s->K;
    int64_t C = s->C;
    int64_t c = lin(s, I);
    int nq = (D == 2) ? 4 : 8;
    int min_q = (min_xcell + nq - 1) / nq; /* cld */
    double xv[NDIM_MAX], dq[NDIM_MAX];
    for (int d = 0; d < D; ++d) {
        xv[d] = g->xv[d].x[I[d]];
        dq[d] = g->xv[d].dx[I[d]] / 2.0;
    }
    for (int q = 0; q < nq; ++q) {
        double vert[NDIM_MAX];
        for (int d = 0; d < D; ++d) vert[d] = xv[d] + dq[d] * (double)((q >> d) & 1);
        int num = 0;
        for (int ip = 0; ip < K; ++ip) {
            int64_t e = c + C * ip;
            if (!index[e]) continue;
            int inside = 1;
            for (int d = 0; d < D; ++d) {
                double p = coords[d][e];
                inside = inside & (vert[d] < p && p < vert[d] + dq[d]);
            }
            if (inside) ++num;
        }
        if (num >= min_q) break;
        for (int ip = 0; ip < K; ++ip) {
            int64_t e = c + C * ip;
            if (index[e]) continue;
            ++num;
            double pn[NDIM_MAX];
            for (int d = 0; d < D; ++d) {
                double r = jpo_uniform(seed, (uint64_t)c, ip, d, 2, counter);
                pn[d] = vert[d] + dq[d] * fma(0.95, r, 0.05);
            }
            for (int d = 0; d < D; ++d) coords[d][e] = pn[d];
            index[e] = 1;
            int64_t src = index_min_distance(s, coords, index, pn, ip, I);
            if (src >= 0) {
                for (int a = 0; a < nargs; ++a) args[a][e] = args[a][src];
            } else if (st) {
                st->orphan++;
            }
            if (st) st->injected++;
            if (num >= min_q) break;
        }
    }
}

/* inject_particles! (src/Particles/injection.jl:21-66): 2^D colour phases
 * (offset_1 outer, last offset inner), cells 2(I-1)+offset. */
int jpo_inject_particles(const jp_cellgrid* g, double* const* coords,
                         uint8_t* index, int nargs, double* const* args,
                         int min_xcell, uint64_t seed, uint64_t call_counter,
                         jp_stats* stats) {
    if (check_grid(g)) return JP_ERR_ARG;
    if (nargs < 0 || nargs > JP_MAX_ARGS) return JP_ERR_ARG;
    shape_t s = shape_of(g);
    int D = s.D;
    int64_t ncol[NDIM_MAX] = {1, 1, 1};
    for (int d = 0; d < D; ++d) ncol[d] = (s.ni[d] + 1) / 2; /* ceil(ni*0.5) */
    int nphase = (D == 2) ? 4 : 8;
    for (int ph = 0; ph < nphase; ++ph) {
        int off[NDIM_MAX];
        if (D == 2) {
            off[0] = ph / 2 + 1;
            off[1] = ph % 2 + 1;
        } else {
            off[0] = ph / 4 + 1;
            off[1] = (ph / 2) % 2 + 1;
            off[2] = ph % 2 + 1;
        }
        int64_t total = 1;
        for (int d = 0; d < D; ++d) total *= ncol[d];
        for (int64_t t = 0; t < total; ++t) {
            int64_t r = t, I[NDIM_MAX];
            int ok = 1;
            for (int d = 0; d < D; ++d) {
                I[d] = 2 * (r % ncol[d]) + off[d] - 1;
                r /= ncol[d];
                if (I[d] >= s.ni[d]) ok = 0;
            }
            if (ok)
                inject_cell(g, &s, coords, index, nargs, args, min_xcell, seed,
                            call_counter, I, stats);
        }
    }
    return JP_OK;
}

/* ------------------------------------------------------------------ */
/* Adjoint of advection + particle2grid on a flat particle list        */
/* ------------------------------------------------------------------ */

/* No reference counterpart: JustPIC.jl has no adjoint.  The forward model
 * being differentiated follows the reference for every particle that is not
 * lost to a full destination cell:
 *  - one step is advect_particle above (src/Particles/Advection/advection.jl:
 *    66-91 with Euler.jl, RK2.jl, RK4.jl), every bracket seeded at the
 *    uniform-spacing estimate (the seed changes nothing for a point strictly
 *    inside one interval; on a tie both this and the HIP library run the
 *    reference bisection from that seed);
 *  - a particle whose new position is not strictly inside the domain is
 *    deleted, as move_particles! deletes it (src/Particles/move_safe.jl:99-103);
 *  - after the last step, particle2grid! (src/Interpolations/
 *    particle_to_grid.jl:37-68, 113-151): every node accumulates w =
 *    inv(distance^2) and w*T over the particles of its 2^D adjacent cells,
 *    F = wT / w in 2D and wT * inv(w) in 3D, accumulated here in particle
 *    order.  A particle's cell is the bracket of its position on the vertex
 *    axes.
 * The adjoint is the exact derivative of this model on the branch the forward
 * pass takes (brackets and cells held fixed), with w taken as 1/s,
 * s = |x - x_node|^2.  It recomputes each trajectory in full (no
 * checkpoints), which makes it independent of the HIP library's checkpoint
 * schedule. */

/* bracket (0-based interval) of p in [x[0], x[n-1]], bisection seeded at the
 * uniform-spacing estimate */
static int64_t adj_bracket(const jp_axis* a, double p) {
    int64_t last = a->n - 2;
    int64_t s = (int64_t)((p - a->x[0]) * a->inv_dx[0]);
    if (s < 0) s = 0;
    if (s > last) s = last;
    return jpo_bisect(p, a->x, a->n, s + 1) - 1;
}

/* velocity component c at p (+Inf outside the component's grid, as
 * interp_velocity above); with Vb non-NULL also its adjoint for the cotangent
 * vb: Vb[c][corner] += weight*vb and pb[d] += dv/dp_d * vb. */
static double adj_interp(const jp_velgrid* vg, const double* const* V, int D, int c,
                         const double* p, double vb, double* const* Vb, double* pb) {
    const jp_axis* ax = vg->ax[c];
    if (!check_limits(ax, D, p)) return INFINITY;
    int64_t k[NDIM_MAX] = {0, 0, 0};
    double t[NDIM_MAX] = {0, 0, 0}, iv[NDIM_MAX] = {0, 0, 0};
    for (int d = 0; d < D; ++d) {
        k[d] = adj_bracket(&ax[d], p[d]);
        iv[d] = 1.0 / ax[d].dx[k[d]];
        t[d] = (p[d] - ax[d].x[k[d]]) * iv[d];
    }
    int64_t n0 = ax[0].n, n1 = ax[1].n;
    int nc = (D == 2) ? 4 : 8;
    int64_t at[8];
    double v[8];
    for (int q = 0; q < nc; ++q) {
        int64_t a = k[0] + (q & 1), b = k[1] + ((q >> 1) & 1);
        int64_t e = (D == 3) ? k[2] + ((q >> 2) & 1) : 0;
        at[q] = a + n0 * (b + n1 * e);
        v[q] = V[c][at[q]];
    }
    double val = (D == 2) ? jpo_lerp2(v, t) : jpo_lerp3(v, t);
    if (!Vb) return val;
    /* n-linear interpolation: weight of corner q is prod_d (bit ? t : 1-t) */
    double dd[NDIM_MAX] = {0, 0, 0};
    if (D == 2) {
        dd[0] = (1.0 - t[1]) * (v[1] - v[0]) + t[1] * (v[3] - v[2]);
        dd[1] = (1.0 - t[0]) * (v[2] - v[0]) + t[0] * (v[3] - v[1]);
    } else {
        for (int d = 0; d < 3; ++d) {
            double s = 0.0;
            for (int q = 0; q < 8; ++q) {
                if ((q >> d) & 1) continue;
                int q1 = q | (1 << d);
                double w = 1.0;
                for (int e = 0; e < 3; ++e) {
                    if (e == d) continue;
                    w *= ((q >> e) & 1) ? t[e] : 1.0 - t[e];
                }
                s += w * (v[q1] - v[q]);
            }
            dd[d] = s;
        }
    }
    for (int d = 0; d < D; ++d) pb[d] += dd[d] * iv[d] * vb;
    for (int q = 0; q < nc; ++q) {
        double w = 1.0;
        for (int d = 0; d < D; ++d) w *= ((q >> d) & 1) ? t[d] : 1.0 - t[d];
        Vb[c][at[q]] += w * vb;
    }
    return val;
}

static void adj_velocity(const jp_velgrid* vg, const double* const* V, int D,
                         const double* p, double* v) {
    for (int c = 0; c < D; ++c) v[c] = adj_interp(vg, V, D, c, p, 0.0, NULL, NULL);
}

/* one step: the arithmetic of advect_particle with adjoint-model brackets */
static void adj_step(int method, double alpha, double dt, int D, const jp_velgrid* vg,
                     const double* const* V, const double* p0, double* out) {
    double v0[NDIM_MAX] = {0, 0, 0}, v1[NDIM_MAX] = {0, 0, 0}, p1[NDIM_MAX] = {0, 0, 0};
    if (method == JP_EULER) {
        adj_velocity(vg, V, D, p0, v0);
        jpo_first_stage(JP_EULER, alpha, dt, D, v0, p0, out);
    } else if (method == JP_RK2) {
        adj_velocity(vg, V, D, p0, v0);
        jpo_first_stage(JP_RK2, alpha, dt, D, v0, p0, p1);
        adj_velocity(vg, V, D, p1, v1);
        jpo_second_stage(alpha, dt, D, v0, v1, p0, out);
    } else {
        double k1[NDIM_MAX], k2[NDIM_MAX], k3[NDIM_MAX], k4[NDIM_MAX], q[NDIM_MAX];
        adj_velocity(vg, V, D, p0, k1);
        for (int d = 0; d < D; ++d) q[d] = p0[d] + (dt * k1[d]) / 2.0;
        adj_velocity(vg, V, D, q, k2);
        for (int d = 0; d < D; ++d) q[d] = p0[d] + (dt * k2[d]) / 2.0;
        adj_velocity(vg, V, D, q, k3);
        for (int d = 0; d < D; ++d) q[d] = p0[d] + dt * k3[d];
        adj_velocity(vg, V, D, q, k4);
        for (int d = 0; d < D; ++d) {
            double sum = ((k1[d] + 2.0 * k2[d]) + 2.0 * k3[d]) + k4[d];
            out[d] = p0[d] + (dt * sum) / 6.0;
        }
    }
}

/* reverse of adj_step at p0: pb holds the cotangent of the step's output on
 * entry and that of p0 on return; velocity cotangents accumulate into Vb */
static void adj_step_bw(int method, double alpha, double dt, int D, const jp_velgrid* vg,
                        const double* const* V, const double* p0, double* pb,
                        double* const* Vb) {
    double ob[NDIM_MAX];
    for (int d = 0; d < D; ++d) ob[d] = pb[d];
    if (method == JP_EULER) {
        for (int c = 0; c < D; ++c) adj_interp(vg, V, D, c, p0, dt * ob[c], Vb, pb);
    } else if (method == JP_RK2) {
        double v0[NDIM_MAX], p1[NDIM_MAX], v0b[NDIM_MAX], v1b[NDIM_MAX];
        double p1b[NDIM_MAX] = {0, 0, 0};
        adj_velocity(vg, V, D, p0, v0);
        jpo_first_stage(JP_RK2, alpha, dt, D, v0, p0, p1);
        double adt = (1.0 * alpha) * dt;
        if (alpha == 0.5) {
            for (int d = 0; d < D; ++d) {
                v1b[d] = dt * ob[d];
                v0b[d] = 0.0;
            }
        } else {
            double ia = 1.0 / alpha, w0 = fma(-0.5, ia, 1.0), w1 = 0.5 * ia;
            for (int d = 0; d < D; ++d) {
                v1b[d] = dt * w1 * ob[d];
                v0b[d] = dt * w0 * ob[d];
            }
        }
        for (int c = 0; c < D; ++c) adj_interp(vg, V, D, c, p1, v1b[c], Vb, p1b);
        for (int d = 0; d < D; ++d) {
            pb[d] += p1b[d];
            v0b[d] += adt * p1b[d];
        }
        for (int c = 0; c < D; ++c) adj_interp(vg, V, D, c, p0, v0b[c], Vb, pb);
    } else {
        double k1[NDIM_MAX], k2[NDIM_MAX], k3[NDIM_MAX], q2[NDIM_MAX], q3[NDIM_MAX],
            q4[NDIM_MAX];
        double kb[4][NDIM_MAX], qb[NDIM_MAX];
        adj_velocity(vg, V, D, p0, k1);
        for (int d = 0; d < D; ++d) q2[d] = p0[d] + (dt * k1[d]) / 2.0;
        adj_velocity(vg, V, D, q2, k2);
        for (int d = 0; d < D; ++d) q3[d] = p0[d] + (dt * k2[d]) / 2.0;
        adj_velocity(vg, V, D, q3, k3);
        for (int d = 0; d < D; ++d) q4[d] = p0[d] + dt * k3[d];
        for (int d = 0; d < D; ++d) {
            kb[0][d] = dt / 6.0 * ob[d];
            kb[1][d] = 2.0 * dt / 6.0 * ob[d];
            kb[2][d] = 2.0 * dt / 6.0 * ob[d];
            kb[3][d] = dt / 6.0 * ob[d];
        }
        const double* qs[3] = {q2, q3, q4};
        const double fac[3] = {dt / 2.0, dt / 2.0, dt};
        for (int st = 3; st >= 1; --st) { /* k_{st+1} = v(q_{st+1}), q = p0 + fac*k_st */
            for (int d = 0; d < D; ++d) qb[d] = 0.0;
            for (int c = 0; c < D; ++c) adj_interp(vg, V, D, c, qs[st - 1], kb[st][c], Vb, qb);
            for (int d = 0; d < D; ++d) {
                pb[d] += qb[d];
                kb[st - 1][d] += fac[st - 1] * qb[d];
            }
        }
        for (int c = 0; c < D; ++c) adj_interp(vg, V, D, c, p0, kb[0][c], Vb, pb);
    }
}

/* strictly inside the domain (move_safe.jl:199-206); NaN and Inf fail */
static int adj_indomain(const jp_cellgrid* g, const double* p) {
    for (int d = 0; d < g->ndim; ++d) {
        const jp_axis* a = &g->xv[d];
        if (!(a->x[0] < p[d] && p[d] < a->x[a->n - 1])) return 0;
    }
    return 1;
}

static int adj_check(const jp_cellgrid* g, int method, double alpha, int nsteps, int every,
                     int64_t np) {
    if (check_grid(g)) return JP_ERR_ARG;
    if (method < JP_EULER || method > JP_RK4) return JP_ERR_ARG;
    if (method == JP_RK2 && !(0.0 < alpha && alpha < 1.0)) return JP_ERR_ARG;
    if (nsteps < 0 || np < 0) return JP_ERR_ARG;
    if (every < 1 || every > JP_ADJ_MAX_EVERY) return JP_ERR_ARG;
    return JP_OK;
}

static int64_t adj_nodes(const jp_cellgrid* g, int64_t* nn) {
    int64_t NN = 1;
    for (int d = 0; d < NDIM_MAX; ++d) nn[d] = 1;
    for (int d = 0; d < g->ndim; ++d) {
        nn[d] = g->ni[d] + 1;
        NN *= nn[d];
    }
    return NN;
}

/* node q (bit d: +1 along d) of the cell holding p: linear node index and
 * weight inv(distance^2) */
static int64_t adj_node(const jp_cellgrid* g, const int64_t* nn, const int64_t* k, int q,
                        const double* p, double* xn, double* wi) {
    int D = g->ndim;
    int64_t N[NDIM_MAX] = {0, 0, 0};
    for (int d = 0; d < D; ++d) {
        N[d] = k[d] + ((q >> d) & 1);
        xn[d] = g->xv[d].x[N[d]];
    }
    double dist = jpo_distance(D, xn, p);
    *wi = 1.0 / (dist * dist);
    return N[0] + nn[0] * (N[1] + nn[1] * N[2]);
}

int jpo_adjoint_forward(const jp_cellgrid* g, const jp_velgrid* vg, const double* const* V,
                        int method, double alpha, double dt, int nsteps, int every,
                        int64_t np, const double* const* x0, const double* Tp,
                        double* ckpt, uint8_t* alive, double* const* xN, double* F,
                        double* W) {
    if (adj_check(g, method, alpha, nsteps, every, np)) return JP_ERR_ARG;
    int D = g->ndim;
    int64_t nn[NDIM_MAX];
    int64_t NN = adj_nodes(g, nn);
    for (int64_t n = 0; n < NN; ++n) F[n] = W[n] = 0.0;
    for (int64_t i = 0; i < np; ++i) {
        double x[NDIM_MAX], y[NDIM_MAX];
        for (int d = 0; d < D; ++d) x[d] = x0[d][i];
        int a = adj_indomain(g, x);
        if (ckpt) /* block 0 always exists, also for nsteps == 0 */
            for (int d = 0; d < D; ++d) ckpt[(int64_t)d * np + i] = x[d];
        for (int s = 0; s < nsteps; ++s) {
            if (ckpt && s > 0 && s % every == 0)
                for (int d = 0; d < D; ++d) ckpt[((int64_t)(s / every) * D + d) * np + i] = x[d];
            if (!a) continue;
            adj_step(method, alpha, dt, D, vg, V, x, y);
            for (int d = 0; d < D; ++d) x[d] = y[d];
            a = adj_indomain(g, x);
        }
        alive[i] = (uint8_t)a;
        for (int d = 0; d < D; ++d) xN[d][i] = x[d];
        if (!a) continue;
        int64_t k[NDIM_MAX] = {0, 0, 0};
        for (int d = 0; d < D; ++d) k[d] = adj_bracket(&g->xv[d], x[d]);
        for (int q = 0; q < (1 << D); ++q) {
            double xn[NDIM_MAX], wi;
            int64_t n = adj_node(g, nn, k, q, x, xn, &wi);
            W[n] += wi;
            F[n] = fma(wi, Tp[i], F[n]);
        }
    }
    for (int64_t n = 0; n < NN; ++n) F[n] = (D == 2) ? F[n] / W[n] : F[n] * (1.0 / W[n]);
    return JP_OK;
}

int jpo_adjoint_backward(const jp_cellgrid* g, const jp_velgrid* vg, const double* const* V,
                         int method, double alpha, double dt, int nsteps, int64_t np,
                         const double* const* x0, const double* Tp, const double* F,
                         const double* W, const double* Fbar, double* const* Vbar,
                         double* const* x0bar, double* Tpbar) {
    if (adj_check(g, method, alpha, nsteps, 1, np)) return JP_ERR_ARG;
    int D = g->ndim;
    int64_t nn[NDIM_MAX];
    adj_nodes(g, nn);
    for (int c = 0; c < D; ++c) {
        int64_t m = 1;
        for (int d = 0; d < D; ++d) m *= vg->ax[c][d].n;
        for (int64_t e = 0; e < m; ++e) Vbar[c][e] = 0.0;
    }
    double* traj = (double*)malloc(sizeof(double) * (size_t)(nsteps + 1) * (size_t)D);
    if (!traj) return JP_ERR_ARG;
    for (int64_t i = 0; i < np; ++i) {
        for (int d = 0; d < D; ++d) traj[d] = x0[d][i];
        int a = adj_indomain(g, traj);
        for (int s = 0; a && s < nsteps; ++s) {
            adj_step(method, alpha, dt, D, vg, V, traj + s * D, traj + (s + 1) * D);
            a = adj_indomain(g, traj + (s + 1) * D);
        }
        double pb[NDIM_MAX] = {0, 0, 0}, tb = 0.0;
        if (a) {
            const double* x = traj + (int64_t)nsteps * D;
            int64_t k[NDIM_MAX] = {0, 0, 0};
            for (int d = 0; d < D; ++d) k[d] = adj_bracket(&g->xv[d], x[d]);
            for (int q = 0; q < (1 << D); ++q) {
                double xn[NDIM_MAX], wi;
                int64_t n = adj_node(g, nn, k, q, x, xn, &wi);
                double iw = 1.0 / W[n];
                tb += Fbar[n] * wi * iw;
                double gw = Fbar[n] * (Tp[i] - F[n]) * iw;
                for (int d = 0; d < D; ++d) pb[d] += gw * (-2.0 * (x[d] - xn[d]) * wi * wi);
            }
            for (int s = nsteps - 1; s >= 0; --s)
                adj_step_bw(method, alpha, dt, D, vg, V, traj + s * D, pb, Vbar);
        }
        for (int d = 0; d < D; ++d) x0bar[d][i] = pb[d];
        Tpbar[i] = tb;
    }
    free(traj);
    return JP_OK;
}
