/*
 * jp_oracle.c -- FP64 restatement of JustPIC.jl v0.6.7's particle hot path
 * (sequential; the OpenMP pragmas only act in the threaded baseline build,
 * where they run the reference's parallel loops on host threads).  TEST INFRASTRUCTURE ONLY (parity checker and CPU baseline); the
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
#ifdef _OPENMP
#include <omp.h>
#endif

int jpo_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

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

static void stats_add(jp_stats* to, const jp_stats* from) {
    to->moved += from->moved;
    to->deleted += from->deleted;
    to->dropped += from->dropped;
    to->far += from->far;
    to->injected += from->injected;
    to->orphan += from->orphan;
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

/* ---- MQS and LinP velocity interpolation ------------------------------ */

static double lerp_at(int D, const double* F, const int64_t* nn, const int64_t* idx,
                      const double* t);

/* half * (t - half)^2 * (muladd(-2, v1, v0) + v2) (src/Interpolations/MQS.jl) */
static double mqs_corr(double t, double v0, double v1, double v2) {
    double e = t - 0.5;
    return (0.5 * (e * e)) * (fma(-2.0, v1, v0) + v2);
}

/* F at 0-based (i, j[, k]) of an array with dims nn (F[a + n0 (b + n1 c)]) */
static double at(const double* F, const int64_t* nn, int64_t i, int64_t j, int64_t k) {
    return F[i + nn[0] * (j + nn[1] * k)];
}

/* MQS along x of 4 corners (v00, v10, v01, v11) in a plane: the lerps along
 * x corrected with the second difference of the three nodes around the
 * particle (one more node left of the cell when t1 < 1/2, else right), then
 * lerp along the second axis.  jb/kb select the rows of the extra nodes
 * (row 0: (jb0, kb0), row 1: (jb1, kb1)).  src/Interpolations/MQS.jl:2-27,
 * :73-98, :126-152. */
static double mqs_x4(const double* F, const int64_t* nn, const double* v, double ta, double tb,
                     int64_t i, int64_t jb0, int64_t kb0, int64_t jb1, int64_t kb1) {
    double lb = jpo_lerp1(ta, v[0], v[1]), lt = jpo_lerp1(ta, v[2], v[3]);
    double cb = ta < 0.5 ? mqs_corr(ta, at(F, nn, i - 1, jb0, kb0), v[0], v[1])
                         : mqs_corr(ta, v[0], v[1], at(F, nn, i + 2, jb0, kb0));
    double ct = ta < 0.5 ? mqs_corr(ta, at(F, nn, i - 1, jb1, kb1), v[2], v[3])
                         : mqs_corr(ta, v[2], v[3], at(F, nn, i + 2, jb1, kb1));
    return jpo_lerp1(tb, lb + cb, lt + ct);
}

/* MQS along y (src/Interpolations/MQS.jl:29-55, :100-124): columns
 * (v00, v01) and (v10, v11) lerped along y with the correction from the node
 * below or above, then lerp along x.  k: the plane of the extra nodes. */
static double mqs_y4(const double* F, const int64_t* nn, const double* v, double t1, double t2,
                     int64_t i, int64_t j, int64_t k) {
    double ll = jpo_lerp1(t2, v[0], v[2]), lr = jpo_lerp1(t2, v[1], v[3]);
    double cl = t2 < 0.5 ? mqs_corr(t2, at(F, nn, i, j - 1, k), v[0], v[2])
                         : mqs_corr(t2, v[0], v[2], at(F, nn, i, j + 2, k));
    double cr = t2 < 0.5 ? mqs_corr(t2, at(F, nn, i + 1, j - 1, k), v[1], v[3])
                         : mqs_corr(t2, v[1], v[3], at(F, nn, i + 1, j + 2, k));
    return jpo_lerp1(t1, ll + cl, lr + cr);
}

/* MQS of velocity component c (0-based) at corner cell (i, j[, k]) with the
 * corner values v (extract_field_corners order) and t.  3D: components x and
 * y combine the plane-k and plane-k+1 results along z but take the extra
 * nodes of both from plane k (MQS.jl:58-64 pass the same i, j, k); z
 * combines two MQS-x evaluations along y, the second (row j+1 corners) also
 * with extra nodes from row j (MQS.jl:67-71, 126-152). */
static double mqs(const double* F, const int64_t* nn, int D, int c, const int64_t* k0,
                  const double* v, const double* t) {
    int64_t i = k0[0], j = k0[1], k = D == 3 ? k0[2] : 0;
    if (D == 2) {
        return c == 0 ? mqs_x4(F, nn, v, t[0], t[1], i, j, 0, j + 1, 0)
                      : mqs_y4(F, nn, v, t[0], t[1], i, j, 0);
    }
    if (c == 0) {
        double bot = mqs_x4(F, nn, v, t[0], t[1], i, j, k, j + 1, k);
        double top = mqs_x4(F, nn, v + 4, t[0], t[1], i, j, k, j + 1, k);
        return jpo_lerp1(t[2], bot, top);
    }
    if (c == 1) {
        double bot = mqs_y4(F, nn, v, t[0], t[1], i, j, k);
        double top = mqs_y4(F, nn, v + 4, t[0], t[1], i, j, k);
        return jpo_lerp1(t[2], bot, top);
    }
    double vf[4] = {v[0], v[1], v[4], v[5]}, vb[4] = {v[2], v[3], v[6], v[7]};
    double front = mqs_x4(F, nn, vf, t[0], t[2], i, j, k, j, k + 1);
    double back = mqs_x4(F, nn, vb, t[0], t[2], i, j, k, j, k + 1);
    return jpo_lerp1(t[1], front, back);
}

/* LinP (src/Particles/Advection/advection_LinP.jl:104-391) of component c:
 * the velocities averaged to the pressure points around the particle
 * (interpolate_V_to_P: the corner index moves by one along c when p is past
 * the cell's middle, every index clamped to the array, pairs averaged as
 * (a + b)/2, y/z orders swapped by swap_F), lerped with t measured from the
 * shifted corner xc_c +- dx_c/2 (correct_xci_to_pressure_point). */
static double linp_vp(const double* F, const int64_t* nn, int D, int c, const int64_t* k0,
                      const double* xc, const double* dxi, const double* p) {
    int64_t b[NDIM_MAX];
    for (int d = 0; d < D; ++d) b[d] = k0[d] + ((d == c && p[d] > xc[d] + dxi[d] / 2.0) ? 1 : 0);
    /* augment_offset rows (advection_LinP.jl:364-391): along the component's
     * direction -1, 0, 1; the other directions take the row's bits: 2D rows
     * 0, 1 on the other axis; 3D x and y components rows (0,0), (1,0), (0,1),
     * (1,1) over the two other axes in order, the z component rows (0,0),
     * (0,1), (1,0), (1,1) over (x, y) */
    int nrow = D == 2 ? 2 : 4;
    double av[8];
    for (int r = 0; r < nrow; ++r) {
        double val[3];
        for (int e = 0; e < 3; ++e) {
            int64_t idx[NDIM_MAX];
            int other = 0;
            for (int d = 0; d < D; ++d) {
                int64_t o;
                if (d == c) {
                    o = e - 1;
                } else {
                    const int bit = (D == 3 && c == 2) ? 1 - other : other;
                    o = (r >> bit) & 1;
                    ++other;
                }
                int64_t q = b[d] + o;
                idx[d] = q < 0 ? 0 : (q > nn[d] - 1 ? nn[d] - 1 : q);
            }
            val[e] = at(F, nn, idx[0], idx[1], D == 3 ? idx[2] : 0);
        }
        av[2 * r] = (val[0] + val[1]) / 2.0;
        av[2 * r + 1] = (val[2] + val[1]) / 2.0;
    }
    /* lerp order: dimension 0 fastest.  The rows above run over the other
     * dimensions; swap_F puts the component's pair where its dimension is. */
    double w[8];
    if (c == 0) {
        for (int q = 0; q < 2 * nrow; ++q) w[q] = av[q];
    } else {
        /* F[1], F[3], F[2], F[4] (, F[5], F[7], F[6], F[8]) */
        static const int SW[8] = {0, 2, 1, 3, 4, 6, 5, 7};
        for (int q = 0; q < 2 * nrow; ++q) w[q] = av[SW[q]];
    }
    double xp[NDIM_MAX], t[NDIM_MAX];
    for (int d = 0; d < D; ++d) xp[d] = xc[d];
    {
        double off = 1.0 - 2.0 * (double)(p[c] < xc[c] + dxi[c] / 2.0);
        xp[c] = xc[c] + (off * dxi[c]) / 2.0;
    }
    for (int d = 0; d < D; ++d) t[d] = (p[d] - xp[d]) * (1.0 / dxi[d]);
    return D == 2 ? jpo_lerp2(w, t) : jpo_lerp3(w, t);
}

/* velocity component c at p with interpolation scheme 0 linear, 1 MQS,
 * 2 LinP: bracket by bisection seeded at I1 (1-based), t = (p - x)*inv(dx);
 * MQS and LinP only where the corner index is interior along every axis
 * (1 < idx < size - 1, 1-based), else the linear value. */
static double interp_scheme(const jp_axis* ax, int D, const double* F, const double* p,
                            const int64_t* I1, int c, int scheme) {
    int64_t k[NDIM_MAX], nn[NDIM_MAX] = {1, 1, 1};
    double t[NDIM_MAX], xc[NDIM_MAX], dxi[NDIM_MAX];
    int interior = 1;
    for (int d = 0; d < D; ++d) {
        k[d] = jpo_bisect(p[d], ax[d].x, ax[d].n, I1[d]) - 1;
        /* outside the axis (semi-Lagrangian, no limit check): the reference's
         * bisection does not terminate; NaN marks the point for the caller */
        if (k[d] < 0) return NAN;
        xc[d] = ax[d].x[k[d]];
        dxi[d] = ax[d].dx[k[d]];
        t[d] = (p[d] - xc[d]) * (1.0 / dxi[d]);
        nn[d] = ax[d].n;
        if (!(k[d] >= 1 && k[d] <= nn[d] - 3)) interior = 0;
    }
    double v[8];
    for (int q = 0; q < (1 << D); ++q)
        v[q] = at(F, nn, k[0] + (q & 1), k[1] + ((q >> 1) & 1), D == 3 ? k[2] + ((q >> 2) & 1) : 0);
    double vl = D == 2 ? jpo_lerp2(v, t) : jpo_lerp3(v, t);
    if (scheme == 0 || !interior) return vl;
    if (scheme == 1) return mqs(F, nn, D, c, k, v, t);
    double vp = linp_vp(F, nn, D, c, k, xc, dxi, p);
    const double A = 2.0 / 3.0;
    return A * vl + (1.0 - A) * vp;
}

/* all components; `limits`: +Inf outside a component's grid (advection);
 * the semi-Lagrangian interpolation has no limit check */
static void interp_velocity_s(const jp_velgrid* vg, const double* const* V, int D,
                              const double* p, const int64_t* I1, int scheme, int limits,
                              double* v) {
    for (int c = 0; c < D; ++c) {
        if (limits && !check_limits(vg->ax[c], D, p))
            v[c] = INFINITY;
        else
            v[c] = interp_scheme(vg->ax[c], D, V[c], p, I1, c, scheme);
    }
}

/* advect_particle with a scheme; back: backtracking (sign -1, semi-Lagrangian,
 * src/Particles/Advection/RK2.jl:53-107, RK4.jl:40-87) */
static void advect_particle_s(int method, double alpha, double dt, int D,
                              const jp_velgrid* vg, const double* const* V, const double* p0,
                              const int64_t* I1, int scheme, int limits, int back,
                              double* out) {
    double v0[NDIM_MAX], v1[NDIM_MAX], p1[NDIM_MAX];
    const double s = back ? -1.0 : 1.0;
    if (method == JP_EULER) {
        interp_velocity_s(vg, V, D, p0, I1, scheme, limits, v0);
        for (int d = 0; d < D; ++d) out[d] = fma(s * dt, v0[d], p0[d]);
    } else if (method == JP_RK2) {
        interp_velocity_s(vg, V, D, p0, I1, scheme, limits, v0);
        for (int d = 0; d < D; ++d) p1[d] = fma((s * alpha) * dt, v0[d], p0[d]);
        interp_velocity_s(vg, V, D, p1, I1, scheme, limits, v1);
        if (alpha == 0.5) {
            for (int d = 0; d < D; ++d) out[d] = fma(s * dt, v1[d], p0[d]);
        } else {
            double ia = 1.0 / alpha;
            double w0 = fma(-0.5, ia, 1.0), w1 = 0.5 * ia;
            for (int d = 0; d < D; ++d) out[d] = fma(s * dt, fma(w1, v1[d], w0 * v0[d]), p0[d]);
        }
    } else {
        double k1[NDIM_MAX], k2[NDIM_MAX], k3[NDIM_MAX], k4[NDIM_MAX], q[NDIM_MAX];
        const double sdt = s * dt;
        interp_velocity_s(vg, V, D, p0, I1, scheme, limits, k1);
        for (int d = 0; d < D; ++d) q[d] = p0[d] + (sdt * k1[d]) / 2.0;
        interp_velocity_s(vg, V, D, q, I1, scheme, limits, k2);
        for (int d = 0; d < D; ++d) q[d] = p0[d] + (sdt * k2[d]) / 2.0;
        interp_velocity_s(vg, V, D, q, I1, scheme, limits, k3);
        for (int d = 0; d < D; ++d) q[d] = p0[d] + sdt * k3[d];
        interp_velocity_s(vg, V, D, q, I1, scheme, limits, k4);
        for (int d = 0; d < D; ++d) {
            double sum = ((k1[d] + 2.0 * k2[d]) + 2.0 * k3[d]) + k4[d];
            out[d] = p0[d] + (sdt * sum) / 6.0;
        }
    }
}

/* advection_MQS! / advection_LinP! (src/Particles/Advection/advection_MQS.jl:
 * 30-124, advection_LinP.jl:27-131): advection! with the scheme's velocity
 * interpolation (scheme 0 is advection! itself). */
int jpo_advection_scheme(const jp_cellgrid* g, double* const* coords, const uint8_t* index,
                         const double* const* V, const jp_velgrid* vg, int method, double alpha,
                         double dt, int scheme) {
    if (check_grid(g)) return JP_ERR_ARG;
    if (method < JP_EULER || method > JP_RK4 || scheme < 0 || scheme > 2) return JP_ERR_ARG;
    if (method == JP_RK2 && !(0.0 < alpha && alpha < 1.0)) return JP_ERR_ARG;
    shape_t s = shape_of(g);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < s.C; ++c) {
        int64_t I[NDIM_MAX], I1[NDIM_MAX];
        unlin(&s, c, I);
        for (int d = 0; d < s.D; ++d) I1[d] = I[d] + 1;
        for (int ip = 0; ip < s.K; ++ip) {
            int64_t e = c + s.C * ip;
            if (!index[e]) continue;
            double p0[NDIM_MAX], p[NDIM_MAX];
            for (int d = 0; d < s.D; ++d) p0[d] = coords[d][e];
            advect_particle_s(method, alpha, dt, s.D, vg, V, p0, I1, scheme, 1, 0, p);
            for (int d = 0; d < s.D; ++d) coords[d][e] = p[d];
        }
    }
    return JP_OK;
}

/* semilagrangian_advection!(F, F0, method, V, grid_vi, grid, dt) and its MQS
 * and LinP forms (src/Particles/Advection/backtracking.jl:15-101,
 * backtracking_MQS.jl, backtracking_LinP.jl): every interior node I of the
 * nodal grid is traced back (RK2 or RK4 with sign -1, no limit check, the
 * velocity bisection seeded at the node index), the backtracked point is
 * bracketed on the vertex axes seeded at I, and F[I] = lerp of F0's corners
 * there with t = (p - x[Ib]) * inv(dx[Ib]) (dx[I], the node's own spacing,
 * for the linear NTuple method, backtracking.jl:79).  The spacings are the
 * axes' dx arrays, which the caller fills with compute_dx
 * (src/Particles/utils.jl:46-49: grid[3] - grid[2] for a range).  Boundary
 * nodes are left unchanged.  The reference's linear scalar method reads F
 * instead of F0 while writing F (backtracking.jl:68); this restatement reads
 * F0, as its NTuple and MQS/LinP methods do.  A backtracked point outside
 * the velocity or vertex grid (where the reference's bisection does not
 * terminate) leaves its node unchanged and the call returns JP_ERR_ARG. */
int jpo_semilagrangian(int ndim, const jp_axis* grid, const jp_velgrid* vg,
                       const double* const* V, int nfields, double* const* F,
                       const double* const* F0, int method, double alpha, double dt, int scheme,
                       int di_at_node) {
    if (ndim != 2 && ndim != 3) return JP_ERR_ARG;
    if (method != JP_RK2 && method != JP_RK4) return JP_ERR_ARG;
    if (method == JP_RK2 && !(0.0 < alpha && alpha < 1.0)) return JP_ERR_ARG;
    if (nfields < 1 || nfields > JP_MAX_ARGS || scheme < 0 || scheme > 2) return JP_ERR_ARG;
    int D = ndim;
    int64_t nn[NDIM_MAX] = {1, 1, 1}, N = 1;
    for (int d = 0; d < D; ++d) {
        if (grid[d].n < 3) return JP_ERR_ARG;
        nn[d] = grid[d].n;
        N *= nn[d];
    }
    /* the velocity brackets are seeded with the node index */
    for (int c = 0; c < D; ++c)
        for (int d = 0; d < D; ++d)
            if (vg->ax[c][d].n < grid[d].n) return JP_ERR_ARG;
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int64_t node = 0; node < N; ++node) {
        int64_t I[NDIM_MAX] = {0, 0, 0}, r = node;
        int interior = 1;
        for (int d = 0; d < D; ++d) {
            I[d] = r % nn[d];
            r /= nn[d];
            if (I[d] < 1 || I[d] > nn[d] - 2) interior = 0;
        }
        if (!interior) continue;
        int64_t I1[NDIM_MAX];
        double p[NDIM_MAX], pb[NDIM_MAX];
        for (int d = 0; d < D; ++d) {
            I1[d] = I[d] + 1;
            p[d] = grid[d].x[I[d]];
        }
        advect_particle_s(method, alpha, dt, D, vg, V, p, I1, scheme, 0, 1, pb);
        int64_t Ib[NDIM_MAX];
        double t[NDIM_MAX];
        int out = 0;
        for (int d = 0; d < D; ++d) {
            Ib[d] = jpo_bisect(pb[d], grid[d].x, grid[d].n, I1[d]) - 1;
            if (Ib[d] < 0) {
                out = 1;
                break;
            }
            double h = grid[d].dx[di_at_node ? I[d] : Ib[d]];
            t[d] = (pb[d] - grid[d].x[Ib[d]]) * (1.0 / h);
        }
        if (out) {
            bad = 1;
            continue;
        }
        for (int f = 0; f < nfields; ++f) F[f][node] = lerp_at(D, F0[f], nn, Ib, t);
    }
    return bad ? JP_ERR_ARG : JP_OK;
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
    /* cells are independent (the reference launches one thread per cell) */
#pragma omp parallel for schedule(static)
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
        int64_t total = 1;
        for (int d = 0; d < D; ++d) total *= ncol[d];
        /* the cells of one colour run in parallel in the threaded build, as
         * the reference's backend runs them; with no jump beyond one cell
         * their neighbourhoods are disjoint and the result is the sequential
         * one.  Outcome counts are kept per thread and added after the phase. */
#pragma omp parallel
        {
            jp_stats loc = {0, 0, 0, 0, 0, 0};
#pragma omp for schedule(static)
            for (int64_t t = 0; t < total; ++t) {
                int64_t r = t, T[NDIM_MAX] = {0, 0, 0};
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
                if (ok) move_cell(g, &s, coords, index, nargs, args, I, stats ? &loc : NULL);
            }
            if (stats) {
#pragma omp critical(jpo_stats)
                stats_add(stats, &loc);
            }
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
#pragma omp parallel for schedule(static)
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
#pragma omp parallel for schedule(static)
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
        /* cells of one colour in parallel in the threaded build (each reads
         * only its 3^D neighbourhood, written by no other cell of the colour) */
#pragma omp parallel
        {
            jp_stats loc = {0, 0, 0, 0, 0, 0};
#pragma omp for schedule(dynamic, 64)
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
                                call_counter, I, stats ? &loc : NULL);
            }
            if (stats) {
#pragma omp critical(jpo_stats)
                stats_add(stats, &loc);
            }
        }
    }
    return JP_OK;
}

/* ------------------------------------------------------------------ */
/* inject_particles_phase!                                             */
/* ------------------------------------------------------------------ */

/* Julia's min/max on Float64 (NaN-propagating; -0.0 below 0.0) and
 * clamp(x, lo, hi) = ifelse(x > hi, hi, ifelse(x < lo, lo, x)). */
static double jl_min(double x, double y) {
    if (isnan(x) || isnan(y)) return isnan(x) ? x : y;
    return (x < y || (x == y && signbit(x) && !signbit(y))) ? x : y;
}
static double jl_max(double x, double y) {
    if (isnan(x) || isnan(y)) return isnan(x) ? x : y;
    return (x > y || (x == y && !signbit(x) && signbit(y))) ? x : y;
}
static double jl_clamp(double x, double lo, double hi) {
    return x > hi ? hi : (x < lo ? lo : x);
}

/* _grid2particle(p, x, di, F, idx) (src/Interpolations/grid_to_particle.jl:
 * 265-273) clamped to the extrema of the same corner values (cell_field,
 * src/Particles/injection.jl:395-409): t_d = (p_d - x_d[idx_d]) * inv(di_d),
 * lerp of field_corners(F, idx) (src/Interpolations/utils.jl:98-117); F is
 * column-major with dims nn. */
static double g2p_clamped(int D, const double* F, const int64_t* nn, const int64_t* idx,
                          const double* p, const double* x0, const double* di) {
    double v[8], t[NDIM_MAX];
    int nc = 1 << D;
    for (int q = 0; q < nc; ++q) {
        int64_t a = idx[0] + (q & 1), b = idx[1] + ((q >> 1) & 1);
        int64_t c = (D == 3) ? idx[2] + ((q >> 2) & 1) : 0;
        v[q] = F[a + nn[0] * (b + nn[1] * c)];
    }
    for (int d = 0; d < D; ++d) t[d] = (p[d] - x0[d]) * (1.0 / di[d]);
    double tmp = (D == 2) ? jpo_lerp2(v, t) : jpo_lerp3(v, t);
    double lo = v[0], hi = v[0];
    for (int q = 1; q < nc; ++q) {
        lo = jl_min(lo, v[q]);
        hi = jl_max(hi, v[q]);
    }
    return jl_clamp(tmp, lo, hi);
}

/* _inject_particles_phase! (src/Particles/injection.jl:233-325) for one cell:
 * every quadrant whose strict count is short is topped up (`continue` to the
 * next quadrant when it has enough, :278, unlike inject_particles!'s break);
 * free slots ascending, new_particle (:411-417) from the shared Philox
 * stream; the new particle takes the phase of its nearest neighbour
 * (index_min_distance before the new particle is stored, :287-289) and each
 * arg j is interpolated from fields[j]: a cell-centred field (size ni) from
 * the centre grid at the shifted, clamped index (:299-308), a nodal field
 * from the cell's corners (:310-314), both clamped to the corner extrema.
 * A new particle without any neighbour keeps its phase value and counts as
 * an orphan (the reference reads phases at slot 0: out of bounds). */
static void inject_phase_cell(const jp_cellgrid* g, const shape_t* s, const jp_axis* xc,
                              double* const* coords, uint8_t* index, double* phases, int nargs,
                              double* const* args, const double* const* fields,
                              const int* centred, int min_xcell, uint64_t seed, uint64_t counter,
                              const int64_t* I, jp_stats* st) {
    int D = s->D, K = s->K;
    int64_t C = s->C;
    int64_t c = lin(s, I);
    int nq = (D == 2) ? 4 : 8;
    int min_q = (min_xcell + nq - 1) / nq; /* cld */
    double xv[NDIM_MAX], dq[NDIM_MAX], di[NDIM_MAX], xcell[NDIM_MAX];
    int64_t nnv[NDIM_MAX] = {1, 1, 1}, nnc[NDIM_MAX] = {1, 1, 1};
    for (int d = 0; d < D; ++d) {
        xv[d] = g->xv[d].x[I[d]];
        di[d] = g->xv[d].dx[I[d]];
        dq[d] = di[d] / 2.0;
        xcell[d] = xv[d] + dq[d]; /* xvi_quadrants[1] .+ di_quadrant */
        nnv[d] = s->ni[d] + 1;
        nnc[d] = s->ni[d];
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
        if (num >= min_q) continue;
        for (int ip = 0; ip < K; ++ip) {
            int64_t e = c + C * ip;
            if (index[e]) continue;
            ++num;
            double pn[NDIM_MAX];
            for (int d = 0; d < D; ++d) {
                double r = jpo_uniform(seed, (uint64_t)c, ip, d, 2, counter);
                pn[d] = vert[d] + dq[d] * fma(0.95, r, 0.05);
            }
            int64_t src = index_min_distance(s, coords, index, pn, ip, I);
            if (src >= 0)
                phases[e] = phases[src];
            else if (st)
                st->orphan++;
            for (int d = 0; d < D; ++d) coords[d][e] = pn[d];
            index[e] = 1;
            for (int a = 0; a < nargs; ++a) {
                if (centred[a]) {
                    int64_t ic[NDIM_MAX];
                    double x0[NDIM_MAX], dc[NDIM_MAX];
                    for (int d = 0; d < D; ++d) {
                        /* shifted_index, then clamp to 1..ni-1 (1-based) */
                        int64_t k = pn[d] < xcell[d] ? I[d] - 1 : I[d];
                        k = k > s->ni[d] - 2 ? s->ni[d] - 2 : (k < 0 ? 0 : k);
                        ic[d] = k;
                        x0[d] = xc[d].x[k];
                        dc[d] = xc[d].dx[k];
                    }
                    args[a][e] = g2p_clamped(D, fields[a], nnc, ic, pn, x0, dc);
                } else {
                    args[a][e] = g2p_clamped(D, fields[a], nnv, I, pn, xv, di);
                }
            }
            if (st) st->injected++;
            if (num >= min_q) break;
        }
    }
}

int jpo_inject_particles_phase(const jp_cellgrid* g, const jp_axis* xc, double* const* coords,
                               uint8_t* index, double* phases, int nargs, double* const* args,
                               const double* const* fields, const int* centred, int min_xcell,
                               uint64_t seed, uint64_t call_counter, jp_stats* stats) {
    if (check_grid(g)) return JP_ERR_ARG;
    if (nargs < 0 || nargs > JP_MAX_ARGS) return JP_ERR_ARG;
    shape_t s = shape_of(g);
    int D = s.D;
    for (int a = 0; a < nargs; ++a)
        if (centred[a])
            for (int d = 0; d < D; ++d)
                if (s.ni[d] < 2) return JP_ERR_ARG;
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
#pragma omp parallel
        {
            jp_stats loc = {0, 0, 0, 0, 0, 0};
#pragma omp for schedule(dynamic, 64)
            for (int64_t t = 0; t < total; ++t) {
                int64_t r = t, I[NDIM_MAX];
                int ok = 1;
                for (int d = 0; d < D; ++d) {
                    I[d] = 2 * (r % ncol[d]) + off[d] - 1;
                    r /= ncol[d];
                    if (I[d] >= s.ni[d]) ok = 0;
                }
                if (ok)
                    inject_phase_cell(g, &s, xc, coords, index, phases, nargs, args, fields,
                                      centred, min_xcell, seed, call_counter, I,
                                      stats ? &loc : NULL);
            }
            if (stats) {
#pragma omp critical(jpo_stats)
                stats_add(stats, &loc);
            }
        }
    }
    return JP_OK;
}

/* ------------------------------------------------------------------ */
/* update_phase_ratios!                                                */
/* ------------------------------------------------------------------ */

/* bilinear_weight (src/PhaseRatios/utils.jl:61-70): the product over d of
 * muladd(-|a_d - b_d|, inv(di_d), 1), accumulated from val = 1. */
static double bilinear_weight(int D, const double* a, const double* b, const double* di) {
    double val = 1.0;
    for (int d = 0; d < D; ++d) val = val * fma(-fabs(a[d] - b[d]), 1.0 / di[d], 1.0);
    return val;
}

/* w = w .+ x .* (phase == j) for j = 1..nph (src/PhaseRatios/utils.jl:44-57):
 * x*false is +0.0 for x >= 0, so only the matching phase changes */
static void add_phase(double* w, int nph, double x, double ph) {
    for (int j = 0; j < nph; ++j)
        if (ph == (double)(j + 1)) w[j] = w[j] + x;
}

/* w .* inv(sum(w)), sum folded left; `zero_nan`: w * !isnan(w) (faces and
 * midpoints write 0 where the sum was 0) */
static void normalise_write(const double* w, int nph, double* out, int64_t L, int64_t loc,
                            int zero_nan) {
    double sum = w[0];
    for (int j = 1; j < nph; ++j) sum = sum + w[j];
    double iv = 1.0 / sum;
    for (int j = 0; j < nph; ++j) {
        double v = w[j] * iv;
        out[loc + L * j] = (zero_nan && isnan(v)) ? 0.0 : v;
    }
}

static void cell_coords(const shape_t* s, double* const* coords, int64_t c, int ip, double* p) {
    for (int d = 0; d < s->D; ++d) p[d] = coords[d][c + s->C * ip];
}

static int any_nan(int D, const double* p) {
    for (int d = 0; d < D; ++d)
        if (isnan(p[d])) return 1;
    return 0;
}

/* |p - m| <= di/2 in every dimension (isinhalfcell, src/PhaseRatios/
 * PhaseRatios.jl:48-49) */
static int in_half_cell(int D, const double* p, const double* m, const double* di) {
    for (int d = 0; d < D; ++d)
        if (!(fabs(p[d] - m[d]) <= di[d] / 2.0)) return 0;
    return 1;
}

/* phase_ratios_center_kernel! (src/PhaseRatios/centers.jl:14-30): the
 * particles of cell I (first coordinate not NaN) weighted about the cell
 * centre xci[I] with the cell's vertex spacing. */
static void ratios_center(const jp_cellgrid* g, const shape_t* s, const jp_axis* xc,
                          double* const* coords, const double* phases, int nph, double* out) {
    int D = s->D;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < s->C; ++c) {
        int64_t I[NDIM_MAX];
        unlin(s, c, I);
        double m[NDIM_MAX], di[NDIM_MAX], w[JPO_MAX_PHASES] = {0};
        for (int d = 0; d < D; ++d) {
            m[d] = xc[d].x[I[d]];
            di[d] = g->xv[d].dx[I[d]];
        }
        for (int ip = 0; ip < s->K; ++ip) {
            double p[NDIM_MAX] = {0.0, 0.0, 0.0};
            cell_coords(s, coords, c, ip, p);
            if (isnan(p[0])) continue;
            add_phase(w, nph, bilinear_weight(D, m, p, di), phases[c + s->C * ip]);
        }
        normalise_write(w, nph, out, s->C, c, 0);
    }
}

/* phase_ratios_vertex_kernel! (src/PhaseRatios/vertices.jl:14-107): node N
 * gathers the particles within half a cell (strict) of the vertex from the
 * up to 2^D cells around it, the first dimension's offset outermost. */
static void ratios_vertex(const jp_cellgrid* g, const shape_t* s, double* const* coords,
                          const double* phases, int nph, double* out) {
    int D = s->D;
    int64_t nn[NDIM_MAX] = {1, 1, 1}, NN = 1;
    for (int d = 0; d < D; ++d) {
        nn[d] = s->ni[d] + 1;
        NN *= nn[d];
    }
#pragma omp parallel for schedule(static)
    for (int64_t node = 0; node < NN; ++node) {
        int64_t N[NDIM_MAX] = {0, 0, 0}, r = node;
        for (int d = 0; d < D; ++d) {
            N[d] = r % nn[d];
            r /= nn[d];
        }
        double vx[NDIM_MAX], w[JPO_MAX_PHASES] = {0};
        for (int d = 0; d < D; ++d) vx[d] = g->xv[d].x[N[d]];
        int nc = 1 << D;
        for (int q = 0; q < nc; ++q) {
            /* offsets (-1..0 per dim), dimension 0 outermost */
            int64_t I[NDIM_MAX];
            int ok = 1;
            for (int d = 0; d < D; ++d) {
                int bit = (q >> (D - 1 - d)) & 1;
                I[d] = N[d] - 1 + bit;
                if (I[d] < 0 || I[d] >= s->ni[d]) ok = 0;
            }
            if (!ok) continue;
            int64_t c = lin(s, I);
            double di[NDIM_MAX];
            for (int d = 0; d < D; ++d) di[d] = g->xv[d].dx[I[d]];
            for (int ip = 0; ip < s->K; ++ip) {
                double p[NDIM_MAX];
                cell_coords(s, coords, c, ip, p);
                if (any_nan(D, p)) continue;
                int far = 0;
                for (int d = 0; d < D; ++d)
                    if (fabs(p[d] - vx[d]) >= di[d] / 2.0) {
                        far = 1;
                        break;
                    }
                if (far) continue;
                add_phase(w, nph, bilinear_weight(D, vx, p, di), phases[c + s->C * ip]);
            }
        }
        normalise_write(w, nph, out, NN, node, 0);
    }
}

/* column-major linear index of I + off in an array of dims ni + off */
static int64_t lin_off(const shape_t* s, const int64_t* I, const int* off) {
    int64_t c = 0, stride = 1;
    for (int d = 0; d < s->D; ++d) {
        c += (I[d] + off[d]) * stride;
        stride *= s->ni[d] + off[d];
    }
    return c;
}

/* column-major linear index of I in an array of dims ni + dims_off */
static int64_t lin_in(const shape_t* s, const int64_t* I, const int* dims_off) {
    int64_t c = 0, stride = 1;
    for (int d = 0; d < s->D; ++d) {
        c += I[d] * stride;
        stride *= s->ni[d] + dims_off[d];
    }
    return c;
}

/* accumulate the particles of cell I (0-based) within half a cell (<=) of m */
static void gather_half(const jp_cellgrid* g, const shape_t* s, double* const* coords,
                        const double* phases, int nph, const int64_t* I, const double* m,
                        double* w, double* di_out) {
    int D = s->D;
    int64_t c = lin(s, I);
    double di[NDIM_MAX];
    for (int d = 0; d < D; ++d) di[d] = g->xv[d].dx[I[d]];
    for (int ip = 0; ip < s->K; ++ip) {
        double p[NDIM_MAX];
        cell_coords(s, coords, c, ip, p);
        if (any_nan(D, p)) continue;
        if (!in_half_cell(D, p, m, di)) continue;
        add_phase(w, nph, bilinear_weight(D, m, p, di), phases[c + s->C * ip]);
    }
    if (di_out)
        for (int d = 0; d < D; ++d) di_out[d] = di[d];
}

/* phase_ratios_face_kernel! (src/PhaseRatios/midpoints.jl:24-80) for the
 * faces normal to `dim`: cell I writes the face at I + e_dim from the
 * particles of cells I and min(I + e_dim, ni) (the last cell twice), and the
 * first cell along dim also writes face I from its own particles about
 * xci - di/2, with di left from the last cell gathered (Julia keeps the loop's
 * assignment).  0 where no particle contributes. */
static void ratios_face(const jp_cellgrid* g, const shape_t* s, const jp_axis* xc,
                        double* const* coords, const double* phases, int nph, int dim,
                        double* out) {
    int D = s->D;
    int off[NDIM_MAX] = {0, 0, 0};
    off[dim] = 1;
    int64_t L = 1;
    for (int d = 0; d < D; ++d) L *= s->ni[d] + off[d];
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < s->C; ++c) {
        int64_t I[NDIM_MAX];
        unlin(s, c, I);
        double di[NDIM_MAX], xcen[NDIM_MAX], face[NDIM_MAX], w[JPO_MAX_PHASES] = {0};
        for (int d = 0; d < D; ++d) {
            di[d] = g->xv[d].dx[I[d]];
            xcen[d] = xc[d].x[I[d]];
            face[d] = xcen[d] + (di[d] * (double)off[d]) / 2.0;
        }
        for (int o = 0; o < 2; ++o) {
            int64_t J[NDIM_MAX];
            for (int d = 0; d < D; ++d) {
                int64_t k = I[d] + (o ? off[d] : 0);
                J[d] = k > s->ni[d] - 1 ? s->ni[d] - 1 : k;
            }
            gather_half(g, s, coords, phases, nph, J, face, w, di);
        }
        normalise_write(w, nph, out, L, lin_off(s, I, off), 1);
        if (I[dim] == 0) {
            double w2[JPO_MAX_PHASES] = {0};
            for (int d = 0; d < D; ++d) face[d] = xcen[d] - (di[d] * (double)off[d]) / 2.0;
            double di_keep[NDIM_MAX];
            for (int d = 0; d < D; ++d) di_keep[d] = di[d];
            /* the boundary face uses the spacing left in di, for every particle */
            int64_t cc = lin(s, I);
            for (int ip = 0; ip < s->K; ++ip) {
                double p[NDIM_MAX];
                cell_coords(s, coords, cc, ip, p);
                if (any_nan(D, p)) continue;
                if (!in_half_cell(D, p, face, di_keep)) continue;
                add_phase(w2, nph, bilinear_weight(D, face, p, di_keep), phases[cc + s->C * ip]);
            }
            normalise_write(w2, nph, out, L, lin_in(s, I, off), 1);
        }
    }
}

/* _phase_ratios_midpoint_kernel! (src/PhaseRatios/midpoints.jl:123-206), 3D:
 * offsets o = (1,1,0) xy, (0,1,1) yz, (1,0,1) xz.  Cell I writes the edge at
 * I + o from the cells min(I + o .* mask, ni) for the masks (1,0,0),
 * (0,1,0), (0,0,1), (1,1,1), in that order, about xci + di .* o / 2.  A cell
 * on a first boundary (I_d == 1 where o_d == 1) also runs the boundary
 * branch: b_d = (ni_d == o_d I_d); for the offsets (0,0,0) and b (both run:
 * an Int tuple is never === false), it writes index I + offset about
 * xci - (di .* o .* (-b)) / 2 with di the spacing of cell min(I, ni),
 * gathering from the same four cells. */
static void midpoint_pass(const jp_cellgrid* g, const shape_t* s, double* const* coords,
                          const double* phases, int nph, const int64_t* I, const int* o,
                          const double* m, double* w, double* di) {
    static const int MASK[4][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 1}};
    for (int mk = 0; mk < 4; ++mk) {
        int64_t J[NDIM_MAX];
        for (int d = 0; d < 3; ++d) {
            int64_t k = I[d] + o[d] * MASK[mk][d];
            J[d] = k > s->ni[d] - 1 ? s->ni[d] - 1 : k;
        }
        gather_half(g, s, coords, phases, nph, J, m, w, di);
    }
}

static void ratios_midpoint(const jp_cellgrid* g, const shape_t* s, const jp_axis* xc,
                            double* const* coords, const double* phases, int nph,
                            const int* o, double* out) {
    int64_t L = 1;
    for (int d = 0; d < 3; ++d) L *= s->ni[d] + o[d];
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < s->C; ++c) {
        int64_t I[NDIM_MAX];
        unlin(s, c, I);
        double di[NDIM_MAX], xcen[NDIM_MAX], m[NDIM_MAX], w[JPO_MAX_PHASES] = {0};
        for (int d = 0; d < 3; ++d) {
            di[d] = g->xv[d].dx[I[d]];
            xcen[d] = xc[d].x[I[d]];
            m[d] = xcen[d] + (di[d] * (double)o[d]) / 2.0;
        }
        midpoint_pass(g, s, coords, phases, nph, I, o, m, w, di);
        normalise_write(w, nph, out, L, lin_off(s, I, o), 1);
        int first = 0;
        for (int d = 0; d < 3; ++d)
            if (o[d] * (I[d] + 1) == 1) first = 1;
        if (!first) continue;
        int b[NDIM_MAX];
        for (int d = 0; d < 3; ++d) b[d] = (s->ni[d] == (int64_t)o[d] * (I[d] + 1)) ? 1 : 0;
        for (int it = 0; it < 2; ++it) {
            int ob[NDIM_MAX];
            for (int d = 0; d < 3; ++d) ob[d] = it ? b[d] : 0;
            double dm[NDIM_MAX], w2[JPO_MAX_PHASES] = {0}, di2[NDIM_MAX];
            for (int d = 0; d < 3; ++d) {
                dm[d] = g->xv[d].dx[I[d]]; /* min(I, ni) == I */
                double flip = (double)(0 - b[d]);
                m[d] = xcen[d] - ((dm[d] * (double)o[d]) * flip) / 2.0;
            }
            midpoint_pass(g, s, coords, phases, nph, I, o, m, w2, di2);
            int64_t Ib[NDIM_MAX];
            for (int d = 0; d < 3; ++d) Ib[d] = I[d] + ob[d];
            normalise_write(w2, nph, out, L, lin_in(s, Ib, o), 1);
        }
    }
}

int jpo_update_phase_ratios(const jp_cellgrid* g, const jp_axis* xc, double* const* coords,
                            const double* phases, int nphases, double* center, double* vertex,
                            double* const* faces, double* const* midpoints) {
    if (check_grid(g)) return JP_ERR_ARG;
    if (nphases < 1 || nphases > JPO_MAX_PHASES) return JP_ERR_ARG;
    shape_t s = shape_of(g);
    if (center) ratios_center(g, &s, xc, coords, phases, nphases, center);
    if (vertex) ratios_vertex(g, &s, coords, phases, nphases, vertex);
    for (int d = 0; d < s.D; ++d)
        if (faces && faces[d]) ratios_face(g, &s, xc, coords, phases, nphases, d, faces[d]);
    if (s.D == 3 && midpoints) {
        /* xy, yz, xz (src/PhaseRatios/midpoints.jl:214-225) */
        static const int OFF[3][3] = {{1, 1, 0}, {0, 1, 1}, {1, 0, 1}};
        for (int k = 0; k < 3; ++k)
            if (midpoints[k]) ratios_midpoint(g, &s, xc, coords, phases, nphases, OFF[k],
                                              midpoints[k]);
    }
    return JP_OK;
}

/* ------------------------------------------------------------------ */
/* grid2particle_flip!, particle2centroid!, centroid2particle!         */
/* ------------------------------------------------------------------ */

/* lerp of the 2^D corners of F (dims nn) at corner cell idx */
static double lerp_at(int D, const double* F, const int64_t* nn, const int64_t* idx,
                      const double* t) {
    double v[8];
    for (int q = 0; q < (1 << D); ++q) {
        int64_t a = idx[0] + (q & 1), b = idx[1] + ((q >> 1) & 1);
        int64_t c = (D == 3) ? idx[2] + ((q >> 2) & 1) : 0;
        v[q] = F[a + nn[0] * (b + nn[1] * c)];
    }
    return (D == 2) ? jpo_lerp2(v, t) : jpo_lerp3(v, t);
}

/* grid2particle_flip!(Fp, xvi, F, F0, particles; α)
 * (src/Interpolations/grid_to_particle.jl:125-207): per active slot of cell I,
 * t = (p - xv[I]) * inv(h) with h = abs(minimum(diff(xv))) per dimension
 * (grid_size, src/Interpolations/utils.jl:61-63: one spacing for the whole
 * axis), F_pic and F0_pic interpolated at the same t, F_flip = Fp + (F_pic -
 * F0_pic), Fp = muladd(F_pic, α, F_flip * (1 - α)).  Every field of a tuple
 * the same way. */
int jpo_grid2particle_flip(const jp_cellgrid* g, int nfields, double* const* Fp,
                           const double* const* F, const double* const* F0,
                           const double* const* coords, const uint8_t* index, double alpha) {
    if (check_grid(g)) return JP_ERR_ARG;
    shape_t s = shape_of(g);
    int D = s.D;
    int64_t nn[NDIM_MAX] = {1, 1, 1};
    double h[NDIM_MAX];
    for (int d = 0; d < D; ++d) {
        nn[d] = s.ni[d] + 1;
        double m = INFINITY;
        for (int64_t k = 0; k + 1 < g->xv[d].n; ++k) {
            double e = g->xv[d].x[k + 1] - g->xv[d].x[k];
            if (e < m) m = e;
        }
        h[d] = fabs(m);
    }
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < s.C; ++c) {
        int64_t I[NDIM_MAX];
        unlin(&s, c, I);
        for (int ip = 0; ip < s.K; ++ip) {
            int64_t e = c + s.C * ip;
            if (!index[e]) continue; /* doskip */
            double t[NDIM_MAX];
            for (int d = 0; d < D; ++d) t[d] = (coords[d][e] - g->xv[d].x[I[d]]) * (1.0 / h[d]);
            for (int f = 0; f < nfields; ++f) {
                double Fpic = lerp_at(D, F[f], nn, I, t);
                double F0pic = lerp_at(D, F0[f], nn, I, t);
                double dF = Fpic - F0pic;
                double Fflip = Fp[f][e] + dF;
                Fp[f][e] = fma(Fpic, alpha, Fflip * (1.0 - alpha));
            }
        }
    }
    return JP_OK;
}

/* particle2centroid!(F, Fp, xci, particles, di)
 * (src/Interpolations/particle_to_grid_centroid.jl:12-128): per cell I, the
 * cell's slots (coordinates not NaN) weighted about the centre xci[I]:
 * ω += ω_i, ωF = muladd(ω_i, Fp, ωF).  The reference's four methods differ:
 *   2D scalar: bilinear weight, any NaN skips, F = ωF / ω            (:28-47)
 *   2D tuple:  ω_i = inv(distance^2) (order 2), ωF[f] * inv(ω)        (:49-75)
 *   3D scalar: bilinear weight, NaN first coordinate skips, ωF*inv(ω)(:79-99)
 *   3D tuple:  bilinear weight, any NaN skips, ωF[f] * inv(ω)        (:101-128)
 * The bilinear weight is src/PhaseRatios/utils.jl:61-70 with di = the
 * cell's vertex spacing.  tuple selects the tuple methods. */
int jpo_particle2centroid(const jp_cellgrid* g, const jp_axis* xc, int nfields,
                          double* const* F, const double* const* Fp,
                          const double* const* coords, int tuple) {
    if (check_grid(g)) return JP_ERR_ARG;
    if (nfields < 1 || nfields > JP_MAX_ARGS || (!tuple && nfields != 1)) return JP_ERR_ARG;
    shape_t s = shape_of(g);
    int D = s.D;
    int dist = (D == 2 && tuple);
    int first_only = (D == 3 && !tuple);
    int divide = (D == 2 && !tuple);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < s.C; ++c) {
        int64_t I[NDIM_MAX];
        unlin(&s, c, I);
        double m[NDIM_MAX], di[NDIM_MAX], w = 0.0, wF[JP_MAX_ARGS];
        for (int d = 0; d < D; ++d) {
            m[d] = xc[d].x[I[d]];
            di[d] = g->xv[d].dx[I[d]];
        }
        for (int f = 0; f < nfields; ++f) wF[f] = 0.0;
        for (int ip = 0; ip < s.K; ++ip) {
            int64_t e = c + s.C * ip;
            double p[NDIM_MAX] = {0.0, 0.0, 0.0};
            for (int d = 0; d < D; ++d) p[d] = coords[d][e];
            if (first_only ? isnan(p[0]) : any_nan(D, p)) continue;
            double wi;
            if (dist) {
                double dd = jpo_distance(D, m, p);
                wi = 1.0 / (dd * dd);
            } else {
                wi = bilinear_weight(D, m, p, di);
            }
            w += wi;
            for (int f = 0; f < nfields; ++f) wF[f] = fma(wi, Fp[f][e], wF[f]);
        }
        if (divide) {
            F[0][c] = wF[0] / w;
        } else {
            double iw = 1.0 / w;
            for (int f = 0; f < nfields; ++f) F[f][c] = wF[f] * iw;
        }
    }
    return JP_OK;
}

/* centroid2particle!(Fp, xci, F, particles, di)
 * (src/Interpolations/centroid_to_particle.jl:15-76): per slot of cell I with
 * no NaN coordinate, the centre cell is shifted_index(p, xci[I], I) (I - 1
 * where p < xci[I]) clamped to 1..ni-1 (1-based), t = (p - xci[k]) *
 * inv(di.center[k]), Fp = lerp of F's corners there (F: dims ni, one value
 * per cell). */
int jpo_centroid2particle(const jp_cellgrid* g, const jp_axis* xc, int nfields,
                          double* const* Fp, const double* const* F,
                          const double* const* coords) {
    if (check_grid(g)) return JP_ERR_ARG;
    if (nfields < 1 || nfields > JP_MAX_ARGS) return JP_ERR_ARG;
    shape_t s = shape_of(g);
    int D = s.D;
    int64_t nn[NDIM_MAX] = {1, 1, 1};
    for (int d = 0; d < D; ++d) {
        if (s.ni[d] < 2) return JP_ERR_ARG;
        nn[d] = s.ni[d];
    }
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < s.C; ++c) {
        int64_t I[NDIM_MAX];
        unlin(&s, c, I);
        for (int ip = 0; ip < s.K; ++ip) {
            int64_t e = c + s.C * ip;
            double p[NDIM_MAX] = {0.0, 0.0, 0.0};
            for (int d = 0; d < D; ++d) p[d] = coords[d][e];
            if (any_nan(D, p)) continue;
            int64_t k[NDIM_MAX];
            double t[NDIM_MAX];
            for (int d = 0; d < D; ++d) {
                int64_t kk = p[d] < xc[d].x[I[d]] ? I[d] - 1 : I[d];
                kk = kk > s.ni[d] - 2 ? s.ni[d] - 2 : (kk < 0 ? 0 : kk);
                k[d] = kk;
                t[d] = (p[d] - xc[d].x[kk]) * (1.0 / xc[d].dx[kk]);
            }
            for (int f = 0; f < nfields; ++f) Fp[f][e] = lerp_at(D, F[f], nn, k, t);
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
