/*
 * jp_oracle.h -- CPU FP64 oracle of JustPIC's particle hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a C restatement of the
 * reference algorithm (JuliaGeodynamics/JustPIC.jl v0.6.7) used as the parity
 * checker for the HIP library and as the CPU baseline; nothing in the product
 * path links, loads or calls it.  Arrays use the same layout as the library
 * (element (cell c, slot k) at data[c + C*k], 0-based column-major cells) but
 * live in HOST memory: every pointer inside jp_cellgrid / jp_velgrid passed to
 * these functions is a host pointer.
 *
 * Pinning (what fixes each function independently of this restatement) is
 * listed in the test files under tests/ that exercise it; functions without an
 * independent pin are marked "parity unpinned" in their comment.
 */
#ifndef JP_ORACLE_H
#define JP_ORACLE_H

#include "justpic_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* host threads the loops run on: 1 for the sequential build, the OpenMP
 * thread count for the threaded build (libjporacle_omp.so) */
int jpo_threads(void);

/* primitives, exposed for known-answer tests */
double jpo_lerp1(double t, double v0, double v1);
double jpo_lerp2(const double v[4], const double t[2]);
double jpo_lerp3(const double v[8], const double t[3]);
int64_t jpo_bisect(double px, const double* x, int64_t n, int64_t seed1);
void jpo_first_stage(int method, double alpha, double dt, int ndim,
                     const double* v, const double* p, double* out);
void jpo_second_stage(double alpha, double dt, int ndim, const double* v0,
                      const double* v1, const double* p, double* out);
double jpo_distance(int ndim, const double* a, const double* b);
void jpo_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double jpo_uniform(uint64_t seed, uint64_t cell, int slot, int dim,
                   int purpose, uint64_t counter);
int64_t jpo_cell_index(double x, const double* xv, int64_t n);

/* hot path */
int jpo_init_particles(const jp_cellgrid* g, double* const* coords,
                       uint8_t* index, int np_quadrant, uint64_t seed);
int jpo_advection(const jp_cellgrid* g, double* const* coords,
                  const uint8_t* index, const double* const* V,
                  const jp_velgrid* vg, int method, double alpha, double dt);
int jpo_move_particles(const jp_cellgrid* g, double* const* coords,
                       uint8_t* index, int nargs, double* const* args,
                       jp_stats* stats);
int jpo_particle2grid(const jp_cellgrid* g, double* F, const double* Fp,
                      const double* const* coords, const uint8_t* index);
int jpo_particle2grid_multi(const jp_cellgrid* g, int nfields, double* const* F,
                            const double* const* Fp,
                            const double* const* coords, const uint8_t* index);
int jpo_grid2particle(const jp_cellgrid* g, double* Fp, const double* F,
                      const double* const* coords, const uint8_t* index);
int jpo_inject_particles(const jp_cellgrid* g, double* const* coords,
                         uint8_t* index, int nargs, double* const* args,
                         int min_xcell, uint64_t seed, uint64_t call_counter,
                         jp_stats* stats);

/* inject_particles_phase!(particles, phases, args, fields)
 * (src/Particles/injection.jl:146-325): xc = centre axes (particles.xci with
 * di.center); centred[a] = 1 when fields[a] is a cell-centred field (dims ni),
 * 0 for a nodal field (dims ni+1). */
int jpo_inject_particles_phase(const jp_cellgrid* g, const jp_axis* xc, double* const* coords,
                               uint8_t* index, double* phases, int nargs, double* const* args,
                               const double* const* fields, const int* centred, int min_xcell,
                               uint64_t seed, uint64_t call_counter, jp_stats* stats);

/* update_phase_ratios!(phase_ratios, particles, phases)
 * (src/PhaseRatios/utils.jl:15-43): outputs in the device CellArray layout
 * (location l, phase k at out[l + L*k]); NULL outputs are skipped.
 * faces[d]: dims ni + e_d; midpoints (3D): xy (ni + (1,1,0)), yz (0,1,1),
 * xz (1,0,1). */
/* grid2particle_flip!, particle2centroid!, centroid2particle!
 * (src/Interpolations/grid_to_particle.jl:125-207,
 * particle_to_grid_centroid.jl:12-128, centroid_to_particle.jl:15-76);
 * xc: centre axes; centroid fields have dims ni.  tuple: the reference's
 * NTuple methods of particle2centroid! (2D: inverse-distance weights). */
int jpo_grid2particle_flip(const jp_cellgrid* g, int nfields, double* const* Fp,
                           const double* const* F, const double* const* F0,
                           const double* const* coords, const uint8_t* index, double alpha);
int jpo_particle2centroid(const jp_cellgrid* g, const jp_axis* xc, int nfields,
                          double* const* F, const double* const* Fp,
                          const double* const* coords, int tuple);
int jpo_centroid2particle(const jp_cellgrid* g, const jp_axis* xc, int nfields,
                          double* const* Fp, const double* const* F,
                          const double* const* coords);

/* advection_MQS! / advection_LinP! (scheme 1 / 2; 0 is advection!) and
 * semilagrangian_advection! with the same schemes
 * (src/Particles/Advection/advection_{MQS,LinP}.jl, backtracking*.jl).
 * semilagrangian: grid = ndim nodal axes whose dx arrays hold compute_dx
 * (src/Particles/utils.jl:46-49), vg likewise; F, F0: nodal fields;
 * di_at_node = 1 for the linear NTuple method (spacing at the node). */
int jpo_advection_scheme(const jp_cellgrid* g, double* const* coords, const uint8_t* index,
                         const double* const* V, const jp_velgrid* vg, int method, double alpha,
                         double dt, int scheme);
int jpo_semilagrangian(int ndim, const jp_axis* grid, const jp_velgrid* vg,
                       const double* const* V, int nfields, double* const* F,
                       const double* const* F0, int method, double alpha, double dt, int scheme,
                       int di_at_node);

#define JPO_MAX_PHASES 16
int jpo_update_phase_ratios(const jp_cellgrid* g, const jp_axis* xc, double* const* coords,
                            const double* phases, int nphases, double* center, double* vertex,
                            double* const* faces, double* const* midpoints);

/* adjoint of advection + particle2grid on a flat particle list (the model of
 * jp_adjoint_forward / jp_adjoint_backward in justpic_hip.h).  ckpt may be
 * NULL.  The backward recomputes whole trajectories from x0 (no checkpoints).
 * Pinned by central finite differences of jpo_adjoint_forward, by the
 * uniform-velocity identity sum(Vbar_c) = nsteps*dt*sum(x0bar_c), and the
 * forward by jpo_advection + jpo_move_particles + jpo_particle2grid. */
int jpo_adjoint_forward(const jp_cellgrid* g, const jp_velgrid* vg, const double* const* V,
                        int method, double alpha, double dt, int nsteps, int every,
                        int64_t np, const double* const* x0, const double* Tp,
                        double* ckpt, uint8_t* alive, double* const* xN, double* F,
                        double* W);
int jpo_adjoint_backward(const jp_cellgrid* g, const jp_velgrid* vg, const double* const* V,
                         int method, double alpha, double dt, int nsteps, int64_t np,
                         const double* const* x0, const double* Tp, const double* F,
                         const double* W, const double* Fbar, double* const* Vbar,
                         double* const* x0bar, double* Tpbar);

#ifdef __cplusplus
}
#endif

#endif
