/*
 * justpic_hip.h -- C-ABI of the MI355X (gfx950) particle hot-path library.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, returns an
 * int status (JP_OK == 0, negative on error; jp_last_error() gives the text) and
 * throws nothing across the ABI.  Array arguments are DEVICE pointers unless the
 * comment says otherwise; arrays of pointers (`double* const* coords`) are HOST
 * arrays whose elements are device pointers.
 *
 * Storage layout (identical to JustPIC's GPU CellArrays, block length B = 0):
 *   a per-particle quantity with K = max_xcell slots per cell over C = prod(ni)
 *   cells is one contiguous array of C*K elements, element (cell c, slot k) at
 *   data[c + C*k]; c = i + ni[0]*(j + ni[1]*l) is the column-major (Julia) cell
 *   index, all indices 0-based.  The active-slot mask ("index") is one byte per
 *   slot (Julia Bool).  Nodal fields are column-major with dims ni[d]+1.
 *
 * Reference: JuliaGeodynamics/JustPIC.jl v0.6.7.  Each function names the
 * reference interface it replaces.
 */
#ifndef JUSTPIC_HIP_H
#define JUSTPIC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JP_OK 0
#define JP_ERR_ARG (-1)      /* invalid argument (sizes, dims, method, alpha) */
#define JP_ERR_HIP (-2)      /* HIP runtime failure */
#define JP_ERR_STATE (-3)    /* context misuse */

#define JP_MAX_ARGS 16       /* particle fields moved together with the coordinates */

/* integrators (src/Advection/types.jl:113-145) */
#define JP_EULER 0
#define JP_RK2 1
#define JP_RK4 2

/* One grid axis: materialised coordinates plus spacing and inverse spacing.
 * For a range grid the reference stores a scalar spacing x[2]-x[1]
 * (src/Particles/particles_utils.jl:137-146); the caller replicates it into
 * dx[0..n-2].  For a vector grid the reference stores diff(x)
 * (src/Particles/particles_utils.jl:76-85).  inv_dx[k] must equal 1/dx[k]. */
typedef struct jp_axis {
    const double* x;      /* n points, strictly increasing */
    const double* dx;     /* n-1 spacings  (particles.di.*)  */
    const double* inv_dx; /* n-1 inverses  (particles._di.*) */
    int64_t n;
} jp_axis;

/* Cell grid of a Particles container: particles.xvi + di.vertex + _di.vertex,
 * cells ni = size(particles.index), K = particles.max_xcell. */
typedef struct jp_cellgrid {
    int32_t ndim;        /* 2 or 3 */
    int32_t max_xcell;   /* K, 1..64 */
    int64_t ni[3];       /* cells per dimension (ni[2] ignored in 2D) */
    jp_axis xv[3];       /* vertex axes, xv[d].n == ni[d] + 1 */
} jp_cellgrid;

/* Staggered velocity grids: particles.xi_vel[c][d] with di.velocity[c][d]. */
typedef struct jp_velgrid {
    jp_axis ax[3][3];    /* ax[c][d]: axis d of velocity component c */
} jp_velgrid;

/* Data-dependent outcomes the reference handles silently (move_safe.jl:95-117,
 * injection.jl:119-123).  Counts accumulate over the calls they are passed to. */
typedef struct jp_stats {
    int64_t moved;     /* particles re-inserted into a free slot              */
    int64_t deleted;   /* particles removed for leaving the domain            */
    int64_t dropped;   /* movers lost because the destination cell was full   */
    int64_t far;       /* movers whose destination is not a neighbouring cell */
    int64_t injected;  /* particles created by injection                      */
    int64_t orphan;    /* injected particles with no neighbour to copy from   */
} jp_stats;

typedef struct jp_ctx jp_ctx;

/* ---- context, memory, streams (no reference counterpart: they replace the
 *      AMDGPU.jl allocation the reference ext/JustPICAMDGPUExt.jl:26-47 uses) */
int jp_ctx_create(jp_ctx** out, int device);
int jp_ctx_destroy(jp_ctx* ctx);
const char* jp_last_error(const jp_ctx* ctx);
/* 1 (default): every entry point synchronises its stream before returning, as
 * the reference launch! does (src/launch.jl:65-69).  0: asynchronous. */
int jp_ctx_set_sync(jp_ctx* ctx, int synchronous);
/* Use a caller-owned hipStream_t (NULL restores the context's own stream). */
int jp_ctx_set_stream(jp_ctx* ctx, void* hip_stream);
void* jp_ctx_stream(jp_ctx* ctx);
/* Launch options.  Results never depend on them; they select launch shapes
 * that the library otherwise chooses from the problem size, so that small
 * problems can run the shapes large ones use (the tests compare every
 * setting with the oracle).
 *   JP_OPT_SLOT_SPAN    slots per slot group of the per-cell kernels (move
 *                       classify and scatter, FP32-split advection, FP64
 *                       advection with the linear interpolation): 0
 *                       (default) chooses from the grid size; otherwise a
 *                       multiple of 8 up to 64 (>= max_xcell: whole cells).
 *   JP_OPT_INJECT_WAVES persistent wavefronts per injection launch: 0
 *                       (default) is one per listed cell up to 4096 (the
 *                       ordered 2D launch is also capped at the number the
 *                       device holds at once).
 *   JP_OPT_INJECT_PHASES the 2D injection (at most one field): 0 (default)
 *                       the four colours in one launch, each listed cell
 *                       waiting for its listed neighbours of earlier
 *                       colours; 1 one launch per colour (the form the 3D
 *                       injection always takes).
 *   JP_OPT_GEOMETRY_RECHECK  the FP32-split axis check: 0 (default) once per
 *                       set of axes (pointers and lengths) until any jp_free,
 *                       1 on every call (for axes whose coordinates the
 *                       caller changes in place or frees outside jp_free).
 *   JP_OPT_MOVE_PHASES  the 2D move's colour phases: 0 (default) all nine in
 *                       one launch over tiles of cells, 1 one launch per
 *                       phase (the form the 3D move always takes).
 *   JP_OPT_ADVECT_SLOTS the FP64 advection with the linear interpolation: 0
 *                       (default) tiles of cells with the velocity staged
 *                       in LDS, 1 one thread per slot gathering from global
 *                       memory (the form the MQS and LinP schemes take).
 * No reference counterpart (the KernelAbstractions launch sizes of the
 * reference are chosen by its backend). */
#define JP_OPT_SLOT_SPAN 1
#define JP_OPT_INJECT_WAVES 2
#define JP_OPT_GEOMETRY_RECHECK 3
#define JP_OPT_MOVE_PHASES 4
#define JP_OPT_ADVECT_SLOTS 5
#define JP_OPT_INJECT_PHASES 6
int jp_ctx_set_option(jp_ctx* ctx, int option, int64_t value);
int jp_sync(jp_ctx* ctx);
int jp_malloc(jp_ctx* ctx, void** dptr, size_t bytes);
int jp_free(jp_ctx* ctx, void* dptr);
int jp_memcpy_h2d(jp_ctx* ctx, void* dst, const void* host_src, size_t bytes);
int jp_memcpy_d2h(jp_ctx* ctx, void* host_dst, const void* src, size_t bytes);
int jp_memcpy_d2d(jp_ctx* ctx, void* dst, const void* src, size_t bytes);
/* fill with a byte value, or with a double value (NaN planes, zero fields) */
int jp_fill_u8(jp_ctx* ctx, uint8_t* dptr, uint8_t value, int64_t count);
int jp_fill_f64(jp_ctx* ctx, double* dptr, double value, int64_t count);
int jp_fill_f32(jp_ctx* ctx, float* dptr, float value, int64_t count);

/* CellArray layout conversion for host copies and checkpoints.  The device
 * layout (B = 0, element (c, k) at data[c + C*k]) is the CPU CellArray data
 * (1, K, C) (element (k, c) at data[k + K*c]) permuted by (3, 2, 1): what
 * Array(particles) does before a JLD2 checkpoint (src/CellArrays/
 * conversion.jl:32-43, src/IO/JLD2.jl:98-127) and what ROCArray(particles)
 * undoes (ext/JustPICAMDGPUExt.jl:53-75).  Device to device, elements of 1,
 * 4 or 8 bytes, src != dst. */
int jp_cellarray_to_cpu_layout(jp_ctx* ctx, const void* src, void* dst, int64_t ncells,
                               int K, int elem_bytes);
int jp_cellarray_from_cpu_layout(jp_ctx* ctx, const void* src, void* dst, int64_t ncells,
                                 int K, int elem_bytes);

/* timing on the context stream (hipEvents) */
int jp_event_create(jp_ctx* ctx, void** ev);
int jp_event_destroy(jp_ctx* ctx, void* ev);
int jp_event_record(jp_ctx* ctx, void* ev);
/* the context stream waits for an event recorded on another context's stream
 * (fork/join of independent work; inside a capture it joins the other stream
 * into the graph) */
int jp_event_wait(jp_ctx* ctx, void* ev);
int jp_event_elapsed_ms(jp_ctx* ctx, void* ev_start, void* ev_stop, float* ms);

/* hipGraph capture of whatever entry points are called between begin and end
 * on the context stream (asynchronous mode is forced while capturing). */
int jp_graph_begin(jp_ctx* ctx);
int jp_graph_end(jp_ctx* ctx, void** graph_exec);
int jp_graph_launch(jp_ctx* ctx, void* graph_exec);
int jp_graph_destroy(jp_ctx* ctx, void* graph_exec);

/* ---- hot path, FP64 drop-in ------------------------------------------- */

/* init_particles, random quadrant-stratified fill
 * (src/Particles/particles_utils.jl:108-194, fill_coords_index! :168-194).
 * rand() is replaced by the Philox4x32-10 stream (seed, cell, slot, dim).
 * coords[d] and index must be pre-sized C*K; every slot is written. */
int jp_init_particles(jp_ctx* ctx, const jp_cellgrid* g, double* const* coords,
                      uint8_t* index, int np_quadrant, uint64_t seed);

/* advection!(particles, method, V, grid_vi, dt, dxi)
 * (src/Particles/Advection/advection.jl:35-91, RK2.jl:1-26, RK4.jl:1-19,
 * Euler.jl:1-20).  V[c] is column-major with dims vg->ax[c][d].n. */
int jp_advection(jp_ctx* ctx, const jp_cellgrid* g, double* const* coords,
                 const uint8_t* index, const double* const* V,
                 const jp_velgrid* vg, int method, double alpha, double dt);

/* velocity interpolation schemes */
#define JP_INTERP_LINEAR 0 /* interp_velocity2particle (advection.jl:93-148) */
#define JP_INTERP_MQS 1    /* MQS (src/Interpolations/MQS.jl:2-158) */
#define JP_INTERP_LINP 2   /* LinP (src/Particles/Advection/advection_LinP.jl:104-391) */

/* advection_MQS!(particles, method, V, grid_vi, dt, dxi) and
 * advection_LinP!(...) (src/Particles/Advection/advection_MQS.jl:30-100,
 * advection_LinP.jl:27-88): jp_advection with the scheme's velocity
 * interpolation, which applies where the bracket is interior along every
 * axis of the component's grid (1 < idx < n - 1, 1-based) and falls back to
 * the linear value elsewhere.  JP_INTERP_LINEAR is jp_advection. */
int jp_advection_scheme(jp_ctx* ctx, const jp_cellgrid* g, double* const* coords,
                        const uint8_t* index, const double* const* V,
                        const jp_velgrid* vg, int method, double alpha, double dt,
                        int scheme);

/* semilagrangian_advection!(F, F0, method, V, grid_vi, grid, dt) and its MQS
 * and LinP forms (src/Particles/Advection/backtracking.jl:15-101,
 * backtracking_MQS.jl:1-103, backtracking_LinP.jl:1-102).  grid: ndim nodal
 * axes; F[f], F0[f]: nodal fields of dims grid[d].n, F[f] distinct from every
 * F0.  Every interior node is traced back with RK2 or RK4 (velocity brackets
 * seeded at the node index, no limit check) and F[f] there becomes the
 * n-linear interpolation of F0[f] at the backtracked point; boundary nodes
 * are not written.  The axes' dx arrays hold compute_dx
 * (src/Particles/utils.jl:46-49: x[3] - x[2] for a range, diff(x) for a
 * vector) and inv_dx = 1/dx.  di_at_node = 1 takes the normalisation spacing
 * at the node instead of the backtracked cell (the linear NTuple method,
 * backtracking.jl:79).  The reference's linear and LinP scalar methods read
 * F while writing it (backtracking.jl:68, backtracking_LinP.jl:68); every
 * method here reads F0.  A backtracked point outside the velocity or vertex
 * grid, where the reference's bisection does not terminate, leaves its node
 * unchanged and makes the call (or, asynchronously, the next jp_sync /
 * jp_memcpy_d2h) return JP_ERR_ARG. */
int jp_semilagrangian(jp_ctx* ctx, int ndim, const jp_axis* grid, const jp_velgrid* vg,
                      const double* const* V, int nfields, double* const* F,
                      const double* const* F0, int method, double alpha, double dt,
                      int scheme, int di_at_node);

/* move_particles!(particles, grid, args, dxi)
 * (src/Particles/move_safe.jl:23-125).  stats may be NULL (no host sync). */
int jp_move_particles(jp_ctx* ctx, const jp_cellgrid* g, double* const* coords,
                      uint8_t* index, int nargs, double* const* args,
                      jp_stats* stats);

/* particle2grid!(F, Fp, particles) for one scalar field
 * (src/Interpolations/particle_to_grid.jl:23-68 2D, :113-151 3D). */
int jp_particle2grid(jp_ctx* ctx, const jp_cellgrid* g, double* F,
                     const double* Fp, const double* const* coords,
                     const uint8_t* index);

/* particle2grid!(F, Fp, particles) for one scalar field, allocating nothing.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under hipGraph capture, serving loops, harnesses that
 * forbid hipMalloc).  Same arguments and same result, bit for bit, as
 * jp_particle2grid.  Its only scratch, one 8-byte occupancy word per cell,
 * lives in a fixed device array of JP_NOALLOC_MAX_CELLS words per GPU (16 MiB),
 * kept until the process ends and shared by every context on that GPU: calls
 * that overlap on one GPU (two streams or two contexts at once) race.  Grids
 * with more than JP_NOALLOC_MAX_CELLS cells return JP_ERR_ARG. */
#define JP_NOALLOC_MAX_CELLS (1 << 21)
int jp_particle2grid_noalloc(jp_ctx* ctx, const jp_cellgrid* g, double* F,
                             const double* Fp, const double* const* coords,
                             const uint8_t* index);

/* particle2grid!(F::NTuple, Fp::NTuple, particles)
 * (src/Interpolations/particle_to_grid.jl:70-109 2D, :153-199 3D). */
int jp_particle2grid_multi(jp_ctx* ctx, const jp_cellgrid* g, int nfields,
                           double* const* F, const double* const* Fp,
                           const double* const* coords, const uint8_t* index);

/* grid2particle!(Fp, xvi, F, particles, di)
 * (src/Interpolations/grid_to_particle.jl:26-82, _grid2particle :265-273). */
int jp_grid2particle(jp_ctx* ctx, const jp_cellgrid* g, double* Fp,
                     const double* F, const double* const* coords,
                     const uint8_t* index);

/* cell_index(x::NTuple, xv::NTuple) for n points (src/particles.jl:205-224):
 * per dimension d, out[d][i] = abs(trunc((x[d][i] - x0) / dx)) + 1 with
 * x0 = xv[d].x[0] and dx = |xv[d].x[1] - xv[d].x[0]|, the 1-based index the
 * reference returns.  x[d], out[d]: device arrays of n values; ndim 1..3;
 * only the first two points of each axis are read. */
int jp_cell_index(jp_ctx* ctx, int ndim, const jp_axis* xv, int64_t n,
                  const double* const* x, int64_t* const* out);

/* grid2particle_flip!(Fp, xvi, F, F0, particles; α)
 * (src/Interpolations/grid_to_particle.jl:125-207): for every active slot,
 * F_pic and F0_pic interpolated from the nodal F and F0 at
 * t = (p - xv[I]) * inv(h), h = abs(minimum(diff(xv))) per dimension
 * (grid_size, src/Interpolations/utils.jl:61-63), then
 * Fp = muladd(F_pic, α, (Fp + F_pic - F0_pic) * (1 - α)); α = 1 is PIC,
 * α = 0 FLIP.  nfields > 1 is the NTuple method (each field alike). */
int jp_grid2particle_flip(jp_ctx* ctx, const jp_cellgrid* g, int nfields, double* const* Fp,
                          const double* const* F, const double* const* F0,
                          const double* const* coords, const uint8_t* index, double alpha);

/* particle2centroid!(F, Fp, xci, particles, di)
 * (src/Interpolations/particle_to_grid_centroid.jl:12-128): cell-centred F
 * (dims ni) from the particles of each cell weighted about its centre
 * (xc: ndim centre axes, particles.xci, xc[d].n == ni[d]).  tuple = 0 is the
 * scalar method (nfields == 1), tuple = 1 the NTuple method, which the
 * reference weights differently in 2D (inverse squared distance instead of
 * the bilinear weight). */
int jp_particle2centroid(jp_ctx* ctx, const jp_cellgrid* g, const jp_axis* xc, int nfields,
                         double* const* F, const double* const* Fp,
                         const double* const* coords, int tuple);

/* centroid2particle!(Fp, xci, F, particles, di)
 * (src/Interpolations/centroid_to_particle.jl:15-76): every slot whose
 * coordinates are not NaN interpolates the cell-centred F (dims ni) on the
 * centre grid at the cell shifted towards the particle (clamped inside). */
int jp_centroid2particle(jp_ctx* ctx, const jp_cellgrid* g, const jp_axis* xc, int nfields,
                         double* const* Fp, const double* const* F,
                         const double* const* coords);

/* inject_particles!(particles, args, grid, di)
 * (src/Particles/injection.jl:21-131, index_min_distance :330-393,
 * new_particle :411-417).  rand() is replaced by Philox4x32-10 keyed by
 * (seed) with counter (cell, slot, dim-block, call_counter). */
int jp_inject_particles(jp_ctx* ctx, const jp_cellgrid* g, double* const* coords,
                        uint8_t* index, int nargs, double* const* args,
                        int min_xcell, uint64_t seed, uint64_t call_counter,
                        jp_stats* stats);

/* inject_particles_phase!(particles, phases, args, fields, grid, grid_center,
 * di, di_center) (src/Particles/injection.jl:146-325, kernel :200-231, body
 * :233-325).  Same colours, quadrants, free-slot order and Philox stream as
 * jp_inject_particles, but every short quadrant is topped up (the reference
 * moves on to the next quadrant instead of breaking), a new particle takes
 * the phase (phases: one double per slot) of its nearest neighbour, and each
 * args[a] is interpolated from fields[a] at the new position, clamped to the
 * extrema of the corner values: centred[a] = 1 for a cell-centred field
 * (dims ni, read on the centre grid xc at the shifted index clamped to
 * 1..ni-1), 0 for a nodal field (dims ni+1, the cell's corners).  xc: ndim
 * centre axes (particles.xci with di.center), xc[d].n == ni[d].  A new
 * particle with no neighbour keeps its phase value and counts as an orphan
 * (the reference reads out of bounds there). */
int jp_inject_particles_phase(jp_ctx* ctx, const jp_cellgrid* g, const jp_axis* xc,
                              double* const* coords, uint8_t* index, double* phases,
                              int nargs, double* const* args, const double* const* fields,
                              const int* centred, int min_xcell, uint64_t seed,
                              uint64_t call_counter, jp_stats* stats);

/* update_phase_ratios!(phase_ratios, particles, phases)
 * (src/PhaseRatios/utils.jl:15-43; centers.jl:14-30, vertices.jl:14-107,
 * midpoints.jl:1-206).  Each output is one PhaseRatios field in the device
 * CellArray layout, location l and phase k (1-based phase value k+1) at
 * out[l + L*k] with L its location count:
 *   center: dims ni;  vertex: dims ni+1;
 *   faces[d] (Vx, Vy[, Vz]): dims ni + e_d;
 *   midpoints (3D only; xy, yz, xz): dims ni + (1,1,0), (0,1,1), (1,0,1).
 * A NULL output (or faces/midpoints array, or element) is skipped.  Face and
 * midpoint locations without particles get 0, centres and vertices NaN, as
 * in the reference.  xc: centre axes as for jp_inject_particles_phase. */
#define JP_MAX_PHASES 16
int jp_update_phase_ratios(jp_ctx* ctx, const jp_cellgrid* g, const jp_axis* xc,
                           const double* const* coords, const double* phases, int nphases,
                           double* center, double* vertex, double* const* faces,
                           double* const* midpoints);

/* ---- FP32-split form ---------------------------------------------------
 * No reference counterpart: a build-only precision form of the same path.
 * A position is the slot's cell (implicit) plus a float32 offset u in cell
 * units per dimension; `off` are D float planes in the CellArray layout.
 * Velocity interpolation and RK stages run in float32 in cell units; FP64 is
 * used for the conversions below, for the offset renormalisation when a
 * particle changes cell, and for the particle2grid weight sums.  Requires
 * uniform vertex axes and uniform velocity axes (any origin and spacing: each
 * velocity axis is mapped affinely from its end points, so
 * the README's expand_range ghost axes, README.md:48-55, are read as they
 * are); every split entry point checks this on the device (deviation from
 * x[0] + k*h at most 1e-9 of h) and a rejected grid changes nothing and
 * returns JP_ERR_ARG.  Cell assignment is integer: a particle stays in its
 * cell while 0 <= u < 1 in every dimension. */
int jp_split_from_f64(jp_ctx* ctx, const jp_cellgrid* g, const double* const* coords,
                      const uint8_t* index, float* const* off);
int jp_split_to_f64(jp_ctx* ctx, const jp_cellgrid* g, float* const* off,
                    const uint8_t* index, double* const* coords);
/* advection! on split positions (same methods as jp_advection) */
int jp_advection_split(jp_ctx* ctx, const jp_cellgrid* g, float* const* off,
                       const uint8_t* index, const double* const* V,
                       const jp_velgrid* vg, int method, double alpha, double dt);
/* move_particles! on split positions: same slot rules (colour phases,
 * starting_point, drops) with integer cell assignment */
int jp_move_particles_split(jp_ctx* ctx, const jp_cellgrid* g, float* const* off,
                            uint8_t* index, int nargs, double* const* args,
                            jp_stats* stats);
int jp_particle2grid_split(jp_ctx* ctx, const jp_cellgrid* g, double* F,
                           const double* Fp, float* const* off, const uint8_t* index);
int jp_grid2particle_split(jp_ctx* ctx, const jp_cellgrid* g, double* Fp,
                           const double* F, float* const* off, const uint8_t* index);
int jp_inject_particles_split(jp_ctx* ctx, const jp_cellgrid* g, float* const* off,
                              uint8_t* index, int nargs, double* const* args,
                              int min_xcell, uint64_t seed, uint64_t call_counter,
                              jp_stats* stats);

/* ---- batched ensemble (no reference counterpart) -----------------------
 * n independent 2D grids in the FP32-split form, advected in one persistent
 * launch: the tiles of all grids form one work list walked by a fixed grid
 * of workgroups.  Grid i uses grids[i], vgs[i] (standard staggering),
 * offsets off[2i], off[2i+1], mask index[i], velocities V[2i], V[2i+1] and
 * time step dt[i].  The pointers are stored: the arrays must stay allocated
 * until jp_ensemble_destroy.  Eight launch variants (tile shape 64x4, 32x8,
 * 16x16, 128x2 cells x 2 or 8 workgroups per CU); jp_ensemble_tune times
 * them on the ensemble itself (with dt = 0, so no particle moves) and keeps
 * the fastest.  jp_ensemble_advect can be captured in a hipGraph together
 * with the per-grid move and particle2grid calls of a time step. */
#define JP_ENS_VARIANTS 8
typedef struct jp_ensemble jp_ensemble;
int jp_ensemble_create(jp_ctx* ctx, int n, const jp_cellgrid* grids, const jp_velgrid* vgs,
                       float* const* off, uint8_t* const* index, const double* const* V,
                       const double* dt, int method, double alpha, jp_ensemble** out);
int jp_ensemble_advect(jp_ctx* ctx, jp_ensemble* ens);
int jp_ensemble_set_variant(jp_ctx* ctx, jp_ensemble* ens, int variant);
/* ms (optional, JP_ENS_VARIANTS floats): average launch time per variant */
int jp_ensemble_tune(jp_ctx* ctx, jp_ensemble* ens, int reps, int* best, float* ms);
int jp_ensemble_destroy(jp_ctx* ctx, jp_ensemble* ens);

/* ---- adjoint of advection + particle2grid (no reference counterpart) ----
 * A flat list of np particles (positions x0[d][i], tracer Tp[i], FP64,
 * device memory) is advected nsteps times with method/alpha/dt exactly as
 * advection! advects one particle (src/Particles/Advection/advection.jl:
 * 66-91); a particle whose position leaves the open domain is deleted, as
 * move_particles! deletes it (src/Particles/move_safe.jl:99-103); the
 * survivors are then gathered onto the nodal grid as particle2grid! does
 * (src/Interpolations/particle_to_grid.jl:37-68, 113-151): F = sum(w T) /
 * sum(w), w = inv(distance^2), over the particles of the 2^D cells around a
 * node.  The list has no slot capacity, so it does not model particles that
 * move_particles! drops when a destination cell is full.
 *
 * jp_adjoint_forward writes the positions at steps 0, every, 2*every, ...
 * (ckpt, max(1, ceil(nsteps/every)) * ndim * np doubles, at
 * [(s/every)*ndim + d]*np + i; block 0 holds x0 also when nsteps == 0),
 * alive[i] (0 once deleted), the final positions xN, F and the weight sums W
 * (nodal, (ni+1) column-major; a node without particles gets NaN).
 * jp_adjoint_backward takes the cotangent Fbar of F and returns the
 * cotangents of V (Vbar[c], overwritten), x0 and Tp (0 for deleted
 * particles), recomputing every block of `every` steps from its checkpoint.
 * The derivative is taken on the branch the forward pass takes (brackets and
 * cells fixed), with w = 1/|x - x_node|^2.  Vbar is accumulated with atomic
 * adds, so its rounding depends on their order; x0bar and Tpbar do not.
 * Methods: JP_EULER, JP_RK2, JP_RK4; 1 <= every <= JP_ADJ_MAX_EVERY. */
#define JP_ADJ_MAX_EVERY 16
int jp_adjoint_forward(jp_ctx* ctx, const jp_cellgrid* g, const jp_velgrid* vg,
                       const double* const* V, int method, double alpha, double dt,
                       int nsteps, int every, int64_t np, const double* const* x0,
                       const double* Tp, double* ckpt, uint8_t* alive, double* const* xN,
                       double* F, double* W);
/* jp_adjoint_forward with the trajectories in the FP32-split form: each
 * particle a cell index plus a float32 offset, every step the split
 * advection's arithmetic (jp_advection_split) and the split move's
 * renormalisation, deletion when the cell leaves the grid; checkpoints, xN
 * and the gather use the FP64 position x_I + u*h, so jp_adjoint_backward
 * recomputes each block in FP64 from the split trajectory.  The vertex axes
 * must be uniform and the velocity axes affine (as for the other split entry
 * points); JP_ERR_ARG otherwise.  No reference counterpart. */
int jp_adjoint_forward_split(jp_ctx* ctx, const jp_cellgrid* g, const jp_velgrid* vg,
                             const double* const* V, int method, double alpha, double dt,
                             int nsteps, int every, int64_t np, const double* const* x0,
                             const double* Tp, double* ckpt, uint8_t* alive, double* const* xN,
                             double* F, double* W);
int jp_adjoint_backward(jp_ctx* ctx, const jp_cellgrid* g, const jp_velgrid* vg,
                        const double* const* V, int method, double alpha, double dt,
                        int nsteps, int every, int64_t np, const double* ckpt,
                        const uint8_t* alive, const double* Tp, const double* F,
                        const double* W, const double* Fbar, double* const* Vbar,
                        double* const* x0bar, double* Tpbar);

/* Philox4x32-10 block (host helper, for tests): out[4] <- philox(ctr[4], key[2]) */
void jp_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

#ifdef __cplusplus
}
#endif

#endif /* JUSTPIC_HIP_H */
