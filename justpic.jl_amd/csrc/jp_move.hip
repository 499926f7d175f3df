// jp_move.hip -- move_particles! (src/Particles/move_safe.jl:23-125) on gfx950.
//
// The reference visits cells in 3^D colour phases; inside a cell it walks the
// slots in order, and a particle that left its cell goes to the first free
// slot >= starting_point of its destination (starting_point shared by all
// destinations of one source cell).  The outcome depends only on which slots
// are occupied, so it is computed on 64-bit occupancy words and the particle
// data are moved afterwards in one parallel pass:
//
//   1. classify (one thread per cell, or per cell and group of slots when the
//      cells alone do not fill the chip; coalesced over the cell planes):
//      in-cell test, domain test, destination (bisection seeded at the cell)
//      as a neighbour code, occupancy/mover words, and a copy of every
//      mover's coordinates and fields into a staging area indexed by its
//      original slot;
//   2. 3^D colour phases (one thread per source cell of the colour): the
//      reference's slot loop replayed on the occupancy words of the 3^D
//      neighbourhood (kept in LDS), recording for every slot that received a
//      particle where that particle came from;
//   3. scatter (threads as in classify): arrivals copy their staged data
//      into their final slot, vacated slots get NaN and mask false.
//
// A particle that jumps more than one cell breaks the colour independence;
// classify raises a flag (and records the largest jump), the regular phases
// return immediately, and a one-workgroup kernel runs the same slot loops on
// the same words, cell by cell in levels that keep the reference's order
// wherever two cells interact, finding the far destinations from the staged
// positions, and clears the flag; the scatter then moves the data as usual.
// The split classify raises another bit of the same flag when the grid's
// geometry was rejected: then nothing moves.
#include "jp_ctx.hpp"

#include <type_traits>

using namespace jp;

namespace {

constexpr uint8_t kDel = 0xFE;
// a mover whose destination is found from its staged position when its cell
// is processed (far-jump path only: movers that jump more than one cell, and
// arrivals there that fail the strict test of their new cell)
constexpr uint8_t kFar = 0xFC;
constexpr int kFlagFar = 1;       // a mover jumps more than one cell
constexpr int kFlagRejected = 2;  // split geometry rejected
constexpr uint8_t kGap = 0x80;  // arrival fails the strict in-cell test of its destination
// s_waitcnt immediate of gfx9 for vmcnt(0) with expcnt and lgkmcnt left free
constexpr int kVmcnt0 = 0x0F70;

__device__ inline double dnan() { return __builtin_nan(""); }

struct MoveScratch {
    uint64_t* occ0;  // occupancy before the call
    uint64_t* occ;   // running occupancy
    uint64_t* mov;   // slots holding a particle that moves when its cell's phase runs
    uint64_t* arr;   // slots holding a particle that arrived during this call
    uint64_t* tch;   // slots that received a particle at some point
    uint8_t* code;   // per slot, 8 per word (word j*C + c): neighbour code / kDel, movers only
    uint32_t* org;   // per slot: staged index of an arrival's record (layout: org_at)
    char* stg;       // staged movers (layout: stg_coord; index: stage_index)
    uint64_t* stm;   // 3D: each cell's staged movers (bit k), written by classify
    int64_t S;       // slots
    int64_t stride;  // 3D: record bytes (multiple of 8)
    int64_t foff;    // 3D: byte offset of the fields in a record; 2D: of the field planes
                     // (after the interleaved coordinates)
};

// Position of (cell, slot) in the origin array.  2D: plane-major like the
// particle data (c + C k).  3D: cell-major (c K + k): each mover of a colour
// phase stores one origin into a neighbour cell, and cell-major keeps a
// wavefront's stores inside the few lines of its cells' neighbourhoods
// (plane-major scatters them over an array larger than the last-level cache:
// 452 MB at 128^3 with K = 54).
template <int D>
__device__ inline uint32_t org_at(uint32_t c, uint32_t k, uint32_t C, uint32_t K) {
    return D == 2 ? c + C * k : c * K + k;
}

template <int D>
constexpr int kNB = D == 2 ? 9 : 27;
template <int D>
constexpr int kCenter = (kNB<D> - 1) / 2;

// colour phase of a cell: offsets (o_1 outer ... o_D inner), 0-based
template <int D>
__device__ inline int phase_of(const int64_t* I) {
    int ph = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) ph = ph * 3 + (int)(I[d] % 3);
    return ph;
}

// Staged coordinate d and field a of the mover from original slot o.  The
// scatter gathers movers by origin, one cache line per array touched, so the
// arrays are few.  2D: the D coordinates interleaved in one array, fields in
// planes (classify's stores stay contiguous across a wavefront's cells).  3D:
// one record per slot, the D coordinates then the fields (3D move -7% split,
// -20% FP64 against planes; full records were 3% slower in 2D).
// Staged movers: one record per mover (coordinates, then the fields at
// byte offset foff), the r-th staged mover of cell c (slot order) at record
// c K + r.  Packed per cell, a destination's arrivals from one origin cell
// share cache lines in the scatter's gathers, and classify writes only the
// movers' bytes (indexed by slot, every record was a cache line of its own
// and the staging writes touched every line of a slot-sized array).
// `below` is the cell's staged movers under the mover's slot.
template <int D>
__device__ inline int64_t stage_index(int64_t c, int K, uint64_t below) {
    return c * K + __popcll(below);
}
template <int D, typename T>
__device__ inline T& stg_coord(const MoveScratch& ms, int64_t o, int d) {
    return reinterpret_cast<T*>(ms.stg + o * ms.stride)[d];
}
template <int D>
__device__ inline double& stg_field(const MoveScratch& ms, int64_t o, int a) {
    return reinterpret_cast<double*>(ms.stg + o * ms.stride + ms.foff)[a];
}
// The 2D split record with one field is 16 bytes (two float offsets, the
// field): written and read with one 16-byte access each.
__device__ inline void stg_put16(const MoveScratch& ms, int64_t o, float u0, float u1, double f) {
    const uint64_t b = (uint64_t)__double_as_longlong(f);
    uint4 v;
    v.x = __float_as_uint(u0);
    v.y = __float_as_uint(u1);
    v.z = (uint32_t)b;
    v.w = (uint32_t)(b >> 32);
    *reinterpret_cast<uint4*>(ms.stg + o * 16) = v;
}
__device__ inline void stg_get16(const MoveScratch& ms, int64_t o, float& u0, float& u1,
                                 double& f) {
    const uint4 v = *reinterpret_cast<const uint4*>(ms.stg + o * 16);
    u0 = __uint_as_float(v.x);
    u1 = __uint_as_float(v.y);
    f = __longlong_as_double((long long)(((uint64_t)v.w << 32) | v.z));
}

// ------------------------------------------------------------ 1. classify
// Fields a0.. of the movers of one chunk of 8 slots, copied to the staging
// area in one batch of independent loads (stg: chunk-local mover bits).
// (o0: the record of the chunk's first staged mover)
template <int D>
__device__ inline void stage_args(int64_t c, int64_t C, int k0, uint32_t stg, int64_t o0,
                                  int nargs, const ArgPtrs& args, const MoveScratch& ms,
                                  int a0 = 0) {
    if (!stg) return;
    for (int a = a0; a < nargs; ++a) {
        const double* __restrict__ src = args.p[a];
        double v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = ((stg >> j) & 1) ? src[c + C * (k0 + j)] : 0.0;
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if ((stg >> j) & 1) stg_field<D>(ms, o0 + __popc(stg & ((1u << j) - 1)), a) = v[j];
    }
}

// Classify's mapping.  CHUNKS (2D, and the 3D split form): a workgroup per
// 64 cells and one wavefront per 8-slot chunk (enough wavefronts to fill the
// chip at 512^2, loads coalesced along the slot planes, no chunk's loads
// behind an earlier chunk's stores); the chunk-threads of a cell add up
// their staged counts in LDS for the packed record index.  Otherwise (the
// 3D FP64 form, whose 200 registers per lane do not fit such workgroups
// well): one thread per cell, whole cells (chunks of 8 slot planes in
// lockstep across the wavefront).
constexpr int kClassifyMax = 512;  // 2D: 64 x up to 8 chunks
template <int D>
struct ClassifySlots {
    int64_t c;
    int kb, ke;
    bool live;
};
template <int D, bool CHUNKS>
__device__ inline ClassifySlots<D> classify_slots(const Grid<D>& g) {
    ClassifySlots<D> r;
    if constexpr (CHUNKS) {
        r.c = blockIdx.x * (int64_t)64 + (threadIdx.x & 63);
        r.kb = (int)(threadIdx.x >> 6) * 8;
        r.ke = r.kb + 8 < g.K ? r.kb + 8 : g.K;
        r.live = r.c < g.C;
    } else {
        r.c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
        r.kb = 0;
        r.ke = g.K;
        r.live = r.c < g.C;
    }
    return r;
}
// The record index of the chunk's first staged mover.  CHUNKS: the counts
// of the cell's lower chunks, through LDS (every thread of the block calls
// this once); else the running count of the thread's own cell.
template <bool CHUNKS>
__device__ inline int64_t chunk_record(int64_t c, int K, int k0, uint32_t stg, uint64_t staged,
                                       int (&s_cnt)[8][64]) {
    if constexpr (CHUNKS) {
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        s_cnt[w][lane] = __popc(stg);
        __syncthreads();
        int base = 0;
        for (int v = 0; v < w; ++v) base += s_cnt[v][lane];
        (void)k0;
        (void)staged;
        return c * K + base;
    } else {
        (void)s_cnt;
        (void)k0;
        (void)stg;
        return c * K + __popcll(staged);
    }
}

// The per-cell words of slot group kb..ke-1 (kb, and ke unless it is K, are
// multiples of 8): whole words for a single group, else the group's bytes
// (the last group also writes the zero bytes above K).
__device__ inline void store_cell_words(const MoveScratch& ms, int64_t c, int kb, int ke, int K,
                                        uint64_t occ, uint64_t mov) {
    if (kb == 0 && ke == K) {
        ms.occ0[c] = occ;
        ms.occ[c] = occ;
        ms.mov[c] = mov;
        ms.arr[c] = 0;
        ms.tch[c] = 0;
        return;
    }
    uint8_t* __restrict__ o0 = reinterpret_cast<uint8_t*>(ms.occ0 + c);
    uint8_t* __restrict__ o1 = reinterpret_cast<uint8_t*>(ms.occ + c);
    uint8_t* __restrict__ mv = reinterpret_cast<uint8_t*>(ms.mov + c);
    const int b1 = ke == K ? 8 : ke >> 3;
    for (int b = kb >> 3; b < b1; ++b) {
        o0[b] = (uint8_t)(occ >> (8 * b));
        o1[b] = (uint8_t)(occ >> (8 * b));
        mv[b] = (uint8_t)(mov >> (8 * b));
    }
    if (kb == 0) {
        ms.arr[c] = 0;
        ms.tch[c] = 0;
    }
}

// Destination codes are packed 8 per 64-bit word, word j of cell c at word
// index j*C + c (coalesced stores here), so the colour phases load a cell's
// codes in one round of word loads.
template <int D>
__global__ void __launch_bounds__(kClassifyMax) k_mv_classify(Grid<D> g, CPtr3 coords,
                                                              const uint8_t* __restrict__ index,
                                                              int nargs, ArgPtrs args,
                                                              MoveScratch ms, int* flag) {
    constexpr bool CH = D == 2;
    __shared__ int s_cnt[8][64];
    const ClassifySlots<D> cs = classify_slots<D, CH>(g);
    if (!CH && !cs.live) return;  // (CH: every thread reaches the block's barrier)
    const int64_t c = cs.live ? cs.c : 0;
    const int kb = cs.kb, ke = cs.ke;  // slot chunk(s) kb..ke-1
    const int64_t C = g.C;
    int64_t I[D];
    unlin<D>(g.ni, c, I);
    // vertices x[I-1..I+2] and spacings dx[I-1..I+1] per dimension, loaded once:
    // the strict brackets of the three candidate cells are register compares
    // (a strictly interior point has a unique bracket, so this equals the
    // bisection seeded at the cell; ties and longer jumps take find_cell)
    double xa[D][4], hx[D][3], dlo[D], dhi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const Axis& A = g.xv[d];
        const int64_t i = I[d], n = A.n;
        xa[d][0] = i > 0 ? A.x[i - 1] : -__builtin_inf();
        xa[d][1] = A.x[i];
        xa[d][2] = A.x[i + 1];
        xa[d][3] = i + 2 < n ? A.x[i + 2] : __builtin_inf();
        hx[d][0] = i > 0 ? A.dx[i - 1] : 0.0;
        hx[d][1] = A.dx[i];
        hx[d][2] = i + 1 < n - 1 ? A.dx[i + 1] : 0.0;
        dlo[d] = A.x[0];
        dhi[d] = A.x[n - 1];
    }
    const uint64_t occ = !cs.live ? 0ull
                         : (kb == 0 && ke == g.K) ? active_slots(index, c, C, g.K)
                                                  : active_slots_in(index, c, C, kb, ke);
    uint64_t mov = 0, staged = 0;
    uint64_t* __restrict__ cw = reinterpret_cast<uint64_t*>(ms.code) + c;
    // chunks of slot planes walked in lockstep by the wavefront (coalesced
    // loads and stores across neighbouring cells)
    constexpr int B = 8;
    for (int k0 = kb; k0 < ke; k0 += B) {
        double pb[B][D];
#pragma unroll
        for (int j = 0; j < B; ++j) {
            const bool on = (occ >> (k0 + j)) & 1;
#pragma unroll
            for (int d = 0; d < D; ++d) pb[j][d] = on ? coords.p[d][c + C * (k0 + j)] : 0.0;
        }
        uint64_t word = 0;
        uint32_t mv8 = 0, stg = 0;
#pragma unroll
        for (int j = 0; j < B; ++j) {
            const int k = k0 + j;
            if (!((occ >> k) & 1)) continue;
            double p[D];
#pragma unroll
            for (int d = 0; d < D; ++d) p[d] = pb[j][d];
            bool inc = true, indom = true;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                inc = inc & (xa[d][1] < p[d] && p[d] < xa[d][1] + hx[d][1]);  // isincell
                indom = indom & (dlo[d] < p[d] && p[d] < dhi[d]);  // indomain
            }
            if (inc) continue;
            mv8 |= 1u << j;
            uint8_t cd;
            if (!indom) {
                cd = kDel;
            } else {
                int n = 0, mult = 1;
                bool far = false, inJ = true;
                int64_t jump = 0;  // cells jumped (largest over the dimensions)
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    int r;
                    if (xa[d][0] < p[d] && p[d] < xa[d][1]) r = -1;
                    else if (xa[d][1] < p[d] && p[d] < xa[d][2]) r = 0;
                    else if (xa[d][2] < p[d] && p[d] < xa[d][3]) r = 1;
                    else {
                        const int64_t t = find_cell(p[d], g.xv[d], I[d]) - I[d];
                        r = t < -1 ? -2 : (t > 1 ? 2 : (int)t);
                        jump = t < -jump ? -t : (t > jump ? t : jump);
                    }
                    far = far || r < -1 || r > 1;
                    n += (r + 1) * mult;
                    mult *= 3;
                    const double lo = r < 0 ? xa[d][0] : (r == 0 ? xa[d][1] : xa[d][2]);
                    const double h = r < 0 ? hx[d][0] : (r == 0 ? hx[d][1] : hx[d][2]);
                    inJ = inJ & (lo < p[d] && p[d] < lo + h);
                }
                if (far) {  // the far-jump path finds the destination again
                    atomicOr(flag, kFlagFar);
                    atomicMax(flag + 1, (int)(jump < (1 << 30) ? jump : (1 << 30)));
                    cd = kFar;
                } else {
                    cd = (uint8_t)n;
                    if (n != kCenter<D> && !inJ) cd |= kGap;
                }
                stg |= 1u << j;
            }
            word |= (uint64_t)cd << (8 * j);
        }
        if (mv8) cw[(k0 >> 3) * C] = word;
        mov |= (uint64_t)mv8 << k0;
        // the chunk's staged movers, packed after the cell's earlier ones
        const int64_t o0 = chunk_record<CH>(c, g.K, k0, stg, staged, s_cnt);
        if (cs.live) {
#pragma unroll
            for (int j = 0; j < B; ++j)
                if ((stg >> j) & 1) {
                    const int64_t o = o0 + __popc(stg & ((1u << j) - 1));
#pragma unroll
                    for (int d = 0; d < D; ++d) stg_coord<D, double>(ms, o, d) = pb[j][d];
                }
            stage_args<D>(c, C, k0, stg, o0, nargs, args, ms);
        }
        staged |= (uint64_t)stg << k0;
    }
    if (!cs.live) return;
    store_cell_words(ms, c, kb, ke, g.K, occ, mov);
    if constexpr (CH)
        reinterpret_cast<uint8_t*>(ms.stm + c)[kb >> 3] = (uint8_t)(staged >> kb);
    else
        ms.stm[c] = staged;
}

// ------------------------------------------------------------ 2. colour phases
template <int D>
__device__ inline int64_t nb_cell(const Grid<D>& g, const int64_t* I, int n, bool& ok) {
    int64_t J[D];
    ok = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        J[d] = I[d] + (n % 3) - 1;
        n /= 3;
        ok = ok && J[d] >= 0 && J[d] < g.ni[d];
    }
    return ok ? lin<D>(g.ni, J) : -1;
}

// Per thread in LDS: for each cell of the 3^D neighbourhood its occupancy
// word and cell number (one 16-byte entry), and the cell's code words.  The
// mover loop is one serial chain per thread with under two wavefronts per
// SIMD, so it is kept short: per mover one LDS byte read (the code), then one
// read of the destination's entry, 32-bit slot arithmetic (a move has fewer
// than 2^32 slots), and this cell's own word in a register.
struct alignas(16) NbEntry {
    uint64_t occ;
    uint32_t cell;  // 0xFFFFFFFF outside the grid (never a destination)
    uint32_t pad;
};

// The reference slot loop of cell I in phase `phase`.  `stop`: the far-jump
// path runs instead (checked after the cell's loads are issued).  FAR (the
// far-jump path, cells in an order that keeps apart any two that interact):
// kFar codes are resolved here from the staged positions, and a destination
// beyond the 3^D neighbourhood is updated in the global words directly.
template <int D, int BS, typename T, bool FAR>
__device__ inline void phase_cell(const Grid<D>& g, const MoveScratch& ms, const int64_t* I,
                                  int phase, bool stop, NbEntry (&s_nb)[kNB<D>][BS],
                                  uint64_t (&s_cw)[8][BS], int& n_moved, int& n_del,
                                  int& n_drop, int& n_far) {
    const int64_t S = lin<D>(g.ni, I);
    const int tid = threadIdx.x;
    const int K = g.K;
    const int64_t C = g.C;
    const uint32_t S32 = (uint32_t)S, C32 = (uint32_t)C;
    const uint64_t full = K == 64 ? ~0ull : ((1ull << K) - 1);
    // everything this cell needs, in one round of independent loads: its
    // mover/arrival words and code words, and occupancy words of its 3^D
    // neighbourhood; in 2D also the neighbours' arrival and touched words.
    // Cells of one colour are 3 apart, so no other thread of this phase reads
    // or writes these neighbourhood words and the values loaded now are still
    // current at the end (loading them late waits for this thread's earlier
    // stores).  In 3D the 26 extra pairs would cost the occupancy they save.
    constexpr bool kPre = D == 2;
    uint64_t pend = ms.mov[S];
    uint64_t arrS = ms.arr[S];
    const uint64_t tchS = ms.tch[S];
    const uint64_t stmS = ms.stm[S];  // staged movers (their record index)
    uint64_t cw[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // movers' code bytes, 8 per word
    {
        const uint64_t* __restrict__ cp = reinterpret_cast<const uint64_t*>(ms.code) + S;
        const int nw = (K + 7) >> 3;
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (j < nw) cw[j] = cp[j * C];
    }
    int64_t nbc[kNB<D>];
    uint64_t was[kNB<D>], arrN[kNB<D>], tchN[kNB<D>];
#pragma unroll
    for (int n = 0; n < kNB<D>; ++n) {
        bool v;
        nbc[n] = nb_cell<D>(g, I, n, v);
        const bool side = kPre && v && n != kCenter<D>;
        was[n] = v ? ms.occ[nbc[n]] : ~0ull;  // outside the grid: never free
        arrN[n] = side ? ms.arr[nbc[n]] : 0ull;
        tchN[n] = side ? ms.tch[nbc[n]] : 0ull;
    }
    if (stop || !pend) return;  // far-jump path, or nothing leaves or re-slots here
#pragma unroll
    for (int n = 0; n < kNB<D>; ++n) s_nb[n][tid] = NbEntry{was[n], (uint32_t)nbc[n], 0u};
#pragma unroll
    for (int j = 0; j < 8; ++j) s_cw[j][tid] = cw[j];
    uint64_t occC = was[kCenter<D>];  // this cell's occupancy
    uint64_t selfm = 0;  // slots that received a re-slotted particle of this cell
    uint64_t selfins = 0;
    int sp = 0;
    while (pend) {
        const int ip = __builtin_ctzll(pend);
        pend &= pend - 1;
        const uint64_t bit = 1ull << ip;
        const uint8_t code = reinterpret_cast<const uint8_t*>(&s_cw[ip >> 3][tid])[ip & 7];
        // a mover of this cell: its staged record; an arrival: the record
        // its origin staged
        uint32_t origin = (uint32_t)stage_index<D>(S, K, stmS & (bit - 1));
        if (arrS & bit) {  // rare: an arrival of this call moves on
            origin = ms.org[org_at<D>(S32, (uint32_t)ip, C32, (uint32_t)K)];
            // wait here, on the rare path: otherwise the loop head waits for
            // the load on every iteration, and with it (one counter for loads
            // and stores) for the previous iteration's store
            __builtin_amdgcn_s_waitcnt(kVmcnt0);
        }
        uint8_t cd = (selfm & bit) ? (uint8_t)kCenter<D> : code;
        occC &= ~bit;  // empty_particle! on the source slot
        arrS &= ~bit;
        selfm &= ~bit;
        if (cd == kDel) {
            ++n_del;
            continue;
        }
        if constexpr (FAR) {
            if (cd == kFar) {
                // the destination from the staged position: bisection seeded
                // at this cell (FP64), the cell + floor(u) (split, whose staged
                // offset is renormalised to it here, for the scatter)
                int64_t J[D];
                bool near = true, inJ = true;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const T q = stg_coord<D, T>(ms, origin, d);
                    if constexpr (std::is_same<T, float>::value) {
                        const int sd = (int)floorf(q);
                        J[d] = I[d] + sd;
                        stg_coord<D, T>(ms, origin, d) = (float)((double)q - (double)sd);
                    } else {
                        J[d] = find_cell(q, g.xv[d], I[d]);
                        const double xl = g.xv[d].x[J[d]];
                        inJ = inJ && xl < q && q < xl + g.xv[d].dx[J[d]];
                    }
                    near = near && J[d] - I[d] <= 1 && I[d] - J[d] <= 1;
                }
                if (near) {  // inside the neighbourhood: the usual path
                    int nn = 0, mult = 1;
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        nn += (int)(J[d] - I[d] + 1) * mult;
                        mult *= 3;
                    }
                    cd = (uint8_t)nn;
                    if (nn != kCenter<D> && !inJ) cd |= kGap;
                } else {
                    ++n_far;
                    const int64_t dcn = lin<D>(g.ni, J);
                    const uint64_t od = ms.occ[dcn];
                    const uint64_t fr = ~od & full & (~0ull << sp);
                    if (!fr) {
                        ++n_drop;
                        continue;
                    }
                    const int f = __builtin_ctzll(fr);
                    sp = f;
                    const uint64_t fb = 1ull << f;
                    ++n_moved;
                    ms.org[org_at<D>((uint32_t)dcn, (uint32_t)f, C32, (uint32_t)K)] = origin;
                    ms.occ[dcn] = od | fb;
                    atomicOr(reinterpret_cast<unsigned long long*>(ms.arr + dcn),
                             (unsigned long long)fb);
                    atomicOr(reinterpret_cast<unsigned long long*>(ms.tch + dcn),
                             (unsigned long long)fb);
                    // an arrival that fails the strict test of its new cell is
                    // looked up again when that cell comes later: in a later
                    // phase, or later in this one (a jump by a multiple of 3
                    // cells keeps the colour; cells of a phase go in linear
                    // order)
                    const int pj = phase_of<D>(J);
                    if (!inJ && (pj > phase || (pj == phase && dcn > S))) {
                        atomicOr(reinterpret_cast<unsigned long long*>(ms.mov + dcn),
                                 (unsigned long long)fb);
                        ms.code[((f >> 3) * C + dcn) * 8 + (f & 7)] = kFar;
                    }
                    continue;
                }
            }
        }
        const int n = cd & 0x7F;
        const NbEntry nb = s_nb[n][tid];
        const uint32_t dcn = nb.cell;
        const uint64_t od = n == kCenter<D> ? occC : nb.occ;
        const uint64_t fr = ~od & full & (~0ull << sp);
        // (classify only gives destinations inside the grid; testing the cell
        // number here makes its LDS read go out with the occupancy read)
        if (!fr || dcn == 0xFFFFFFFFu) {  // find_free_memory found nothing: dropped
            ++n_drop;
            continue;
        }
        const int f = __builtin_ctzll(fr);
        sp = f;  // starting_point = free_idx
        const uint64_t fb = 1ull << f;
        s_nb[n][tid].occ = od | fb;  // (the centre's LDS copy is not read)
        ++n_moved;
        ms.org[org_at<D>(dcn, (uint32_t)f, C32, (uint32_t)K)] = origin;
        if (n == kCenter<D>) {
            occC |= fb;
            arrS |= fb;
            selfins |= fb;
            if (f > ip) {  // the slot loop will reach this particle again
                pend |= fb;
                selfm |= fb;
            }
        } else if (cd & kGap) {
            // the arrival fails the strict test of its new cell: when that cell's
            // phase comes later, it is re-slotted there (bisection returns that cell)
            const int64_t Dc = dcn;
            int64_t J[D];
            unlin<D>(g.ni, Dc, J);
            if (phase_of<D>(J) > phase) {
                ms.mov[Dc] |= fb;
                // (the far-jump path looks the destination up again, as the
                // reference does)
                ms.code[((f >> 3) * C + Dc) * 8 + (f & 7)] = FAR ? kFar : (uint8_t)kCenter<D>;
            }
        }
    }
#pragma unroll
    for (int n = 0; n < kNB<D>; ++n) {
        if (n == kCenter<D>) continue;
        const NbEntry nb = s_nb[n][tid];
        const uint32_t dn = nb.cell;
        if (dn == 0xFFFFFFFFu) continue;
        const uint64_t now = nb.occ;
        if (now != was[n]) {
            const uint64_t ins = now & ~was[n];
            ms.occ[dn] = now;
            if constexpr (kPre) {
                ms.arr[dn] = arrN[n] | ins;
                ms.tch[dn] = tchN[n] | ins;
            } else {
                // no other thread of the phase touches these words: an atomic
                // OR without return updates them with no load to wait for (a
                // load here would also wait for every earlier store)
                atomicOr(reinterpret_cast<unsigned long long*>(ms.arr + dn),
                         (unsigned long long)ins);
                atomicOr(reinterpret_cast<unsigned long long*>(ms.tch + dn),
                         (unsigned long long)ins);
            }
        }
    }
    ms.occ[S] = occC;
    ms.arr[S] = arrS;
    ms.tch[S] = tchS | selfins;
    ms.mov[S] = 0;
}

// one colour phase, one thread per colour cell
template <int D, int BS>
__global__ void __launch_bounds__(BS) k_mv_phase(Grid<D> g, MoveScratch ms, Off3 off, int phase,
                                                 const int* __restrict__ flag, jp_stats* st) {
    __shared__ NbEntry s_nb[kNB<D>][BS];
    __shared__ uint64_t s_cw[8][BS];
    const int stop = *flag;
    int n_moved = 0, n_del = 0, n_drop = 0, n_far = 0;
    int64_t t = blockIdx.x * (int64_t)BS + threadIdx.x, total = 1, ncol[D], I[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        ncol[d] = (g.ni[d] + 2) / 3;
        total *= ncol[d];
    }
    bool ok = t < total;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        I[d] = 3 * (t % ncol[d]) + off.o[d];
        t /= ncol[d];
        ok = ok && I[d] < g.ni[d];
    }
    if (ok)
        phase_cell<D, BS, double, false>(g, ms, I, phase, stop != 0, s_nb, s_cw, n_moved, n_del,
                                         n_drop, n_far);
    stat_add(st, ST_MOVED, n_moved);
    stat_add(st, ST_DELETED, n_del);
    stat_add(st, ST_DROPPED, n_drop);
}

// ------------------------------------------------------------ 2'. 2D phases in tiles
// All nine colour phases of the 2D move in one launch.  A workgroup owns a
// tile of cells and replays the phases on the occupancy, pending, arrival and
// touched words of the tile plus a halo, all in LDS, with a barrier between
// phases; it keeps the results of its own cells.
//
// The halo is the dependency cone of the phases.  The processing of cell A in
// phase p reads and writes the words of A's 3x3 neighbourhood, so it depends
// on the processings, in earlier phases, of the cells B within 2 of A (their
// neighbourhoods overlap A's), and on theirs in turn.  A step from A to B
// moves, per dimension, by a multiple of 3 plus the difference of their
// colours, within [-2, 2]; taking the phases down from p, the offsets that can
// be reached lie in lo_p..hi_p with, for phase p = 3 c0 + c1 (c0 = I0 mod 3,
// the outer digit of the phase, c1 = I1 mod 3):
//   dimension 0: -c0 .. 2 c0,  dimension 1: -p .. 2 p.
// A processing is exact here when A + [lo_p - 1, hi_p + 1] lies inside the
// region (or leaves the grid there), and a processing that changes an exact
// one is itself exact, so only those run.  The final words of an owned cell
// come from the processings of its 3x3 neighbourhood, so the region is the
// tile widened by the largest cone plus 2: 4 cells below and 6 above in
// dimension 0, 10 below and 18 above in dimension 1.
//
// Tiles overlap, so a halo processing also runs in the tile that owns it;
// both compute the same slots from the same exact inputs.  Every processing
// here stores its origins (equal values where tiles overlap), since a later
// phase of this workgroup reads the origin of an arrival that moves on;
// counts and final words come from owned cells only.
//
// A pending slot that holds an arrival of this call is always re-slotted in
// its own cell (an arrival that fails the strict test of its new cell, or a
// particle re-slotted higher in its own cell), so the destination code of
// such a slot is the centre and the code words are read-only here.
constexpr int kTileHalo0Lo = 4, kTileHalo0Hi = 6, kTileHalo1Lo = 10, kTileHalo1Hi = 18;

__device__ inline int cone_lo(int p, int d) { return d == 0 ? -(p / 3) : -p; }
__device__ inline int cone_hi(int p, int d) { return d == 0 ? 2 * (p / 3) : 2 * p; }

template <int TX, int TY, int BS, bool STATS>
__global__ void __launch_bounds__(BS) k_mv_tiles2(Grid<2> g, MoveScratch ms,
                                                  const int* __restrict__ flag, jp_stats* st) {
    constexpr int RX = TX + kTileHalo0Lo + kTileHalo0Hi;
    constexpr int RY = TY + kTileHalo1Lo + kTileHalo1Hi;
    constexpr int R = RX * RY;
    static_assert(((RX + 2) / 3) * ((RY + 2) / 3) <= BS, "one cell per thread and phase");
    __shared__ uint64_t s_occ[R], s_mov[R], s_arr[R], s_tch[R];
    // the code bytes of the cell each thread is processing: a row of 68
    // bytes per thread (slot k at byte k; 17 dwords, so the lanes' rows start
    // on different banks)
    constexpr int kRow = 68;
    __shared__ uint32_t s_code[BS * kRow / 4];
    if (*flag) return;  // far jump or rejected grid: nothing moves here
    const int n0 = (int)g.ni[0], n1 = (int)g.ni[1];
    const int ntx = (n0 + TX - 1) / TX;
    // neighbouring tiles read each other's halo words: consecutive tiles of
    // a row on one XCD
    const int tile = (int)xcd_block(blockIdx.x, gridDim.x);
    const int tx0 = (tile % ntx) * TX, ty0 = (tile / ntx) * TY;
    const int tx1 = (tx0 + TX < n0 ? tx0 + TX : n0) - 1;
    const int ty1 = (ty0 + TY < n1 ? ty0 + TY : n1) - 1;
    const int gx0 = tx0 - kTileHalo0Lo > 0 ? tx0 - kTileHalo0Lo : 0;
    const int gx1 = tx1 + kTileHalo0Hi < n0 - 1 ? tx1 + kTileHalo0Hi : n0 - 1;
    const int gy0 = ty0 - kTileHalo1Lo > 0 ? ty0 - kTileHalo1Lo : 0;
    const int gy1 = ty1 + kTileHalo1Hi < n1 - 1 ? ty1 + kTileHalo1Hi : n1 - 1;
    const int rw = gx1 - gx0 + 1, rh = gy1 - gy0 + 1;
    for (int i = threadIdx.x; i < rw * rh; i += BS) {
        const int x = i % rw, y = i / rw;
        const uint32_t c = (uint32_t)(gx0 + x) + (uint32_t)n0 * (uint32_t)(gy0 + y);
        const int r = x + RX * y;
        s_occ[r] = ms.occ0[c];
        s_mov[r] = ms.mov[c];
        s_arr[r] = 0;
        s_tch[r] = 0;
    }
    __syncthreads();
    const int K = g.K;
    const uint32_t C32 = (uint32_t)g.C;
    const uint64_t full = K == 64 ? ~0ull : ((1ull << K) - 1);
    const int nw = (K + 7) >> 3;
    const uint8_t* __restrict__ sc = reinterpret_cast<const uint8_t*>(s_code) + threadIdx.x * kRow;
    int n_moved = 0, n_del = 0, n_drop = 0;
    // the colour-p cells whose processing is exact in the region: x = vx +
    // 3i (i < cx), y = vy + 3j (j < cy)
    auto box = [&](int p, int& vx, int& vy, int& cx, int& cy) {
        const int ax = gx0 == 0 ? 0 : gx0 - cone_lo(p, 0) + 1;
        const int bx = gx1 == n0 - 1 ? n0 - 1 : gx1 - cone_hi(p, 0) - 1;
        const int ay = gy0 == 0 ? 0 : gy0 - cone_lo(p, 1) + 1;
        const int by = gy1 == n1 - 1 ? n1 - 1 : gy1 - cone_hi(p, 1) - 1;
        vx = ax + (p / 3 - ax % 3 + 3) % 3;
        vy = ay + (p % 3 - ay % 3 + 3) % 3;
        cx = vx <= bx ? (bx - vx) / 3 + 1 : 0;
        cy = vy <= by ? (by - vy) / 3 + 1 : 0;
    };
    auto load_codes = [&](uint32_t S32, uint64_t (&w)[8]) {
        const uint64_t* __restrict__ cp = reinterpret_cast<const uint64_t*>(ms.code) + S32;
#pragma unroll
        for (int j = 0; j < 8; ++j) w[j] = j < nw ? cp[(int64_t)j * C32] : 0ull;
    };
    auto cell_at = [&](int vx, int vy, int cx, int t) {
        return (uint32_t)(vx + 3 * (t % cx)) + (uint32_t)n0 * (uint32_t)(vy + 3 * (t / cx));
    };
    // the code words of each thread's first cell of a phase are loaded one
    // phase ahead, so that they are in flight during the phase before
    uint64_t nxt[8];
    {
        int vx, vy, cx, cy;
        box(0, vx, vy, cx, cy);
        if ((int)threadIdx.x < cx * cy) load_codes(cell_at(vx, vy, cx, threadIdx.x), nxt);
    }
    for (int p = 0; p < kNB<2>; ++p) {
        const int col0 = p / 3, col1 = p % 3;
        int vx, vy, cx, cy;
        box(p, vx, vy, cx, cy);
        if ((int)threadIdx.x < cx * cy) {
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (j < nw) {
                    s_code[threadIdx.x * (kRow / 4) + 2 * j] = (uint32_t)nxt[j];
                    s_code[threadIdx.x * (kRow / 4) + 2 * j + 1] = (uint32_t)(nxt[j] >> 32);
                }
        }
        if (p + 1 < kNB<2>) {
            int wx, wy, dx_, dy_;
            box(p + 1, wx, wy, dx_, dy_);
            if ((int)threadIdx.x < dx_ * dy_) load_codes(cell_at(wx, wy, dx_, threadIdx.x), nxt);
        }
        const int t = threadIdx.x;  // at most one cell of the colour per thread
        uint64_t pend = 0;
        int x = 0, y = 0, r = 0;
        if (t < cx * cy) {
            x = vx + 3 * (t % cx);
            y = vy + 3 * (t / cx);
            r = (x - gx0) + RX * (y - gy0);
            pend = s_mov[r];
        }
        if (pend) {
            const bool own = tx0 <= x && x <= tx1 && ty0 <= y && y <= ty1;
            const uint32_t S32 = (uint32_t)x + (uint32_t)n0 * (uint32_t)y;
            // neighbours inside the grid (bit n)
            uint32_t ingm = 0;
#pragma unroll
            for (int n = 0; n < kNB<2>; ++n) {
                const int jx = x + n % 3 - 1, jy = y + n / 3 - 1;
                ingm |= (uint32_t)(jx >= 0 && jx < n0 && jy >= 0 && jy < n1) << n;
            }
            uint64_t occC = s_occ[r], arrS = s_arr[r], tchS = s_tch[r];
            uint64_t free_sp = full;  // slots >= starting_point
            // a mover of this cell: its packed record, the cell's staged
            // movers below it (the loop visits the cell's own movers in slot
            // order, each once; the deleted ones are not staged)
            uint32_t rec = S32 * (uint32_t)K;
            while (pend) {
                const int ip = __builtin_ctzll(pend);
                pend &= pend - 1;
                uint32_t cd = sc[ip];
                const uint64_t bit = 1ull << ip;
                occC &= ~bit;  // empty_particle! on the source slot
                uint32_t origin = rec;
                rec += (!(arrS & bit) && cd != kDel) ? 1u : 0u;
                if (arrS & bit) {  // rare: an arrival of this call is re-slotted
                    origin = ms.org[org_at<2>(S32, (uint32_t)ip, C32, (uint32_t)K)];
                    // waited for here, on the rare path: a wait at the loop
                    // head would also wait for the loop's origin stores (one
                    // counter for vector loads and stores)
                    __builtin_amdgcn_s_waitcnt(kVmcnt0);
                    arrS &= ~bit;
                    cd = kCenter<2>;
                }
                if (cd == kDel) {
                    if (STATS) n_del += own;
                    continue;
                }
                const int n = cd & 0x7F;
                const int q3 = (n * 11) >> 5;  // n / 3 for n < 9
                const int dx = n - 3 * q3 - 1, dy = q3 - 1;
                int rn = r + dx + RX * dy;  // inside the region when in the grid
                rn = rn < 0 ? 0 : (rn > R - 1 ? R - 1 : rn);
                const uint64_t ol = s_occ[rn];
                const bool centre = n == kCenter<2>;
                // outside the grid: never free
                const uint64_t od = centre ? occC : (((ingm >> n) & 1) ? ol : ~0ull);
                const uint64_t fr = ~od & free_sp;
                if (!fr) {  // find_free_memory found nothing: dropped
                    if (STATS) n_drop += own;
                    continue;
                }
                const int f = __builtin_ctzll(fr);
                const uint64_t fb = 1ull << f;
                free_sp = full & (~0ull << f);  // starting_point = free_idx
                if (STATS) n_moved += own;
                const uint32_t dcn = S32 + (uint32_t)(dx + n0 * dy);
                ms.org[org_at<2>(dcn, (uint32_t)f, C32, (uint32_t)K)] = origin;
                if (centre) {
                    occC |= fb;
                    arrS |= fb;
                    tchS |= fb;
                    if (f > ip) pend |= fb;  // the slot loop reaches it again
                } else {
                    s_occ[rn] = od | fb;
                    atomicOr(reinterpret_cast<unsigned long long*>(&s_arr[rn]),
                             (unsigned long long)fb);
                    atomicOr(reinterpret_cast<unsigned long long*>(&s_tch[rn]),
                             (unsigned long long)fb);
                    // an arrival that fails the strict test of its new cell is
                    // re-slotted there when that cell's phase comes later
                    if ((cd & kGap) && ((col0 + dx + 3) % 3) * 3 + (col1 + dy + 3) % 3 > p)
                        atomicOr(reinterpret_cast<unsigned long long*>(&s_mov[rn]),
                                 (unsigned long long)fb);
                }
            }
            s_occ[r] = occC;
            s_arr[r] = arrS;
            s_tch[r] = tchS;
            s_mov[r] = 0;
        }
        __syncthreads();
    }
    const int ow = tx1 - tx0 + 1, oh = ty1 - ty0 + 1;
    for (int i = threadIdx.x; i < ow * oh; i += BS) {
        const int x = i % ow, y = i / ow;
        const uint32_t c = (uint32_t)(tx0 + x) + (uint32_t)n0 * (uint32_t)(ty0 + y);
        const int r = (tx0 - gx0) + x + RX * ((ty0 - gy0) + y);
        ms.occ[c] = s_occ[r];
        ms.arr[c] = s_arr[r];
        ms.tch[c] = s_tch[r];
    }
    if (STATS) {
        stat_add(st, ST_MOVED, n_moved);
        stat_add(st, ST_DELETED, n_del);
        stat_add(st, ST_DROPPED, n_drop);
    }
}

// ------------------------------------------------------------ 3. scatter
// T: coordinate type (double: FP64 drop-in, float: split offsets)
template <int D, typename T>
struct TPtr3 {
    T* p[3];
};

// REC16: 2D split with one field (16-byte records, one load per arrival)
template <int D, typename T, bool REC16 = false>
__global__ void __launch_bounds__(kBlock) k_mv_scatter(Grid<D> g, TPtr3<D, T> coords,
                                                       uint8_t* __restrict__ index, int nargs,
                                                       ArgPtrs args, MoveScratch ms,
                                                       const SplitGeom* __restrict__ geo,
                                                       int span) {
    // one thread per cell and slot group (blockIdx.y, `span` slots), chunks of
    // 8 slot planes in lockstep across the wavefront (coalesced origin loads
    // and final stores); only the reads of the staged data are scattered.
    // Arrivals copy their staged data, vacated slots are emptied
    // (empty_particle!, src/Particles/move_safe.jl:243-251).  Nothing moves
    // when the split grid was rejected (geo: null in the FP64 form).
    if (geo && geo->bad) return;
    // 2D: XCD-contiguous cell ranges (a destination's sources are its
    // neighbour cells, whose staged lines other threads of the same L2 also
    // read; a row of cells above or below is only a few workgroups away, which
    // round-robin dealing puts on other XCDs); slot groups of one cell block
    // stay together.  3D keeps the dealt order: there the plane above or below
    // is a multiple of 8 workgroups away, already on the same XCD, and the
    // renumbering measured slower.
    const unsigned lb = D == 2 ? xcd_block(blockIdx.x + gridDim.x * blockIdx.y,
                                           gridDim.x * gridDim.y)
                               : blockIdx.x * gridDim.y + blockIdx.y;
    const int64_t c = (lb / gridDim.y) * (int64_t)blockDim.x + threadIdx.x;
    if (c >= g.C) return;
    const int kb = (int)(lb % gridDim.y) * span;
    const int ke = kb + span < g.K ? kb + span : g.K;
    const int64_t C = g.C;
    // the first chunk's origins go out with the cell's words: unguarded loads
    // (slot clamped to K - 1, values used for arrivals only), one round trip
    // fewer than loading them after the arrival bits are known
    uint32_t o0[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int k = kb + j < g.K ? kb + j : g.K - 1;
        o0[j] = ms.org[org_at<D>((uint32_t)c, (uint32_t)k, (uint32_t)C, (uint32_t)g.K)];
    }
    const uint64_t arr = ms.arr[c], occ = ms.occ[c], occ0 = ms.occ0[c], tch = ms.tch[c];
    const uint64_t work = arr | (~arr & ~occ & (occ0 | tch));
    if (!(work >> kb)) return;
    for (int k0 = kb; k0 < ke; k0 += 8) {
        const uint32_t wb = (uint32_t)(work >> k0) & 0xFFu;
        if (!wb) continue;
        const uint32_t ab = (uint32_t)(arr >> k0) & 0xFFu;
        uint32_t o[8];
        if (k0 == kb) {
#pragma unroll
            for (int j = 0; j < 8; ++j) o[j] = o0[j];
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j)
                o[j] = ((ab >> j) & 1) ? ms.org[org_at<D>((uint32_t)c, (uint32_t)(k0 + j),
                                                          (uint32_t)C, (uint32_t)g.K)]
                                       : 0u;
        }
        if constexpr (REC16) {
            float u0[8], u1[8];
            double fv[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                u0[j] = u1[j] = __builtin_nanf("");
                fv[j] = dnan();
                if ((ab >> j) & 1) stg_get16(ms, o[j], u0[j], u1[j], fv[j]);
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                if (!((wb >> j) & 1)) continue;
                const int64_t e = c + C * (k0 + j);
                coords.p[0][e] = (T)u0[j];
                coords.p[D - 1][e] = (T)u1[j];
                index[e] = (uint8_t)((ab >> j) & 1);
                args.p[0][e] = fv[j];
            }
            continue;
        }
        T v[8][D];
#pragma unroll
        for (int j = 0; j < 8; ++j)
#pragma unroll
            for (int d = 0; d < D; ++d)
                v[j][d] = ((ab >> j) & 1) ? stg_coord<D, T>(ms, o[j], d) : (T)dnan();
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if (!((wb >> j) & 1)) continue;
            const int64_t e = c + C * (k0 + j);
#pragma unroll
            for (int d = 0; d < D; ++d) coords.p[d][e] = v[j][d];
            index[e] = (uint8_t)((ab >> j) & 1);
        }
        for (int a = 0; a < nargs; ++a) {
            double av[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) av[j] = ((ab >> j) & 1) ? stg_field<D>(ms, o[j], a) : dnan();
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if ((wb >> j) & 1) args.p[a][c + C * (k0 + j)] = av[j];
        }
    }
}

// ------------------------------------------------------------ far jumps
// The colour phases when some particle jumps more than one cell (flag[0] has
// kFlagFar, flag[1] holds the largest jump R in cells), in one workgroup on
// the same words, codes and staged data as the regular phases; the scatter
// then moves the data as usual.  A cell's processing touches only cells
// within R of it, so two cells of one colour phase (3t + o) interact only
// when their colour coordinates t differ by at most q = floor(2R/3) in every
// dimension.  Within a phase the reference visits the cells in linear order
// of t; here they go in levels L(t) = sum_d m_d t_d with m_0 = 1 and
// m_d = 1 + sum_{e<d} m_e q_e (q_e = q capped at the colour count - 1): any
// two interacting cells have their linear order as level order, and cells of
// one level do not interact, so each level runs in parallel between
// barriers and the result is the reference's.  With a small R that is
// 3^D * (about (1 + m_1 + m_2) * the colour count per dimension) levels; a
// jump of the grid's size makes every level one cell (the serial order).
constexpr int kFarBlock = 256;

template <int D, typename T>
__global__ void __launch_bounds__(kFarBlock) k_mv_far(Grid<D> g, MoveScratch ms,
                                                      int* __restrict__ flag, jp_stats* st) {
    __shared__ NbEntry s_nb[kNB<D>][kFarBlock];
    __shared__ uint64_t s_cw[8][kFarBlock];
    const int fl = flag[0];
    if (!fl) return;
    const int R = flag[1];
    __syncthreads();  // every thread has read the words before they are cleared
    if (!(fl & kFlagFar) || (fl & kFlagRejected)) {
        if (threadIdx.x == 0) flag[0] = flag[1] = 0;
        return;
    }
    int64_t ncol[D], m[D];
#pragma unroll
    for (int d = 0; d < D; ++d) ncol[d] = (g.ni[d] + 2) / 3;
    const int64_t q = (2 * (int64_t)R) / 3;
    int64_t lmax = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        m[d] = 1;
        for (int e = 0; e < d; ++e) m[d] += m[e] * (q < ncol[e] - 1 ? q : ncol[e] - 1);
        lmax += m[d] * (ncol[d] - 1);
    }
    // the (t_1, ..., t_{D-1}) combinations, one per thread and round
    int64_t outer = 1;
#pragma unroll
    for (int d = 1; d < D; ++d) outer *= ncol[d];
    int n_moved = 0, n_del = 0, n_drop = 0, n_far = 0;
    for (int ph = 0; ph < kNB<D>; ++ph) {
        int o[D];
        int rr = ph;
#pragma unroll
        for (int d = D - 1; d >= 0; --d) {
            o[d] = rr % 3;
            rr /= 3;
        }
        for (int64_t lv = 0; lv <= lmax; ++lv) {
            for (int64_t i = threadIdx.x; i < outer; i += kFarBlock) {
                int64_t t[D], rest = i, l0 = lv;
#pragma unroll
                for (int d = 1; d < D; ++d) {
                    t[d] = rest % ncol[d];
                    rest /= ncol[d];
                    l0 -= m[d] * t[d];
                }
                t[0] = l0;
                if (t[0] < 0 || t[0] >= ncol[0]) continue;
                int64_t I[D];
                bool ok = true;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    I[d] = 3 * t[d] + o[d];
                    ok = ok && I[d] < g.ni[d];
                }
                if (ok)
                    phase_cell<D, kFarBlock, T, true>(g, ms, I, ph, false, s_nb, s_cw, n_moved,
                                                      n_del, n_drop, n_far);
            }
            __syncthreads();
        }
    }
    stat_add(st, ST_MOVED, n_moved);
    stat_add(st, ST_DELETED, n_del);
    stat_add(st, ST_DROPPED, n_drop);
    stat_add(st, ST_FAR, n_far);
    if (threadIdx.x == 0) flag[0] = flag[1] = 0;  // ready for the next move
}

// ------------------------------------------------------------ split classify
// Split positions: the slot's cell plus an offset u in cell units.  A particle
// whose floor(u) is zero in every dimension stays; otherwise its destination
// is the cell + floor(u) (deleted when that lies outside the grid), and the
// staged offset is renormalised to the destination, u - floor(u), in FP64.
template <int D>
__global__ void __launch_bounds__(kClassifyMax)
    k_mv_classify_split(Grid<D> g, TPtr3<D, float> off, const uint8_t* __restrict__ index,
                        int nargs, ArgPtrs args, MoveScratch ms, int* flag,
                        const SplitGeom* __restrict__ geo) {
    __shared__ int s_cnt[8][64];
    const ClassifySlots<D> cs = classify_slots<D, true>(g);
    if (geo->bad) {  // grid rejected: the later kernels of this move return at once
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(flag, kFlagRejected);
        return;  // (uniform over the block)
    }
    const int64_t c = cs.live ? cs.c : 0;
    const int kb = cs.kb, ke = cs.ke;  // slot chunk(s) kb..ke-1
    const int64_t C = g.C;
    int64_t I[D];
    unlin<D>(g.ni, c, I);
    const uint64_t occ = !cs.live ? 0ull
                         : (kb == 0 && ke == g.K) ? active_slots(index, c, C, g.K)
                                                  : active_slots_in(index, c, C, kb, ke);
    uint64_t mov = 0, staged = 0;
    uint64_t* __restrict__ cw = reinterpret_cast<uint64_t*>(ms.code) + c;
    constexpr int B = 8;
    const double* __restrict__ f0 = nargs ? args.p[0] : nullptr;
    for (int k0 = kb; k0 < ke; k0 += B) {
        // the offsets and the first field of the chunk's particles in one round
        // (the field of a particle that stays is loaded for nothing; loading it
        // after the movers are known would cost a round trip)
        float ub[B][D];
        double fb[B];
#pragma unroll
        for (int j = 0; j < B; ++j) {
            const bool on = (occ >> (k0 + j)) & 1;
#pragma unroll
            for (int d = 0; d < D; ++d) ub[j][d] = on ? off.p[d][c + C * (k0 + j)] : 0.0f;
            fb[j] = (on && f0) ? f0[c + C * (k0 + j)] : 0.0;
        }
        uint64_t word = 0;
        uint32_t mv8 = 0, stg = 0;
#pragma unroll
        for (int j = 0; j < B; ++j) {
            const int k = k0 + j;
            if (!((occ >> k) & 1)) continue;
            int s[D];
            bool stay = true, out = false, far = false;
            int n = 0, mult = 1, jump = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const bool fin = fabsf(ub[j][d]) < 1.0e6f;  // false for NaN/Inf
                s[d] = fin ? (int)floorf(ub[j][d]) : 0;
                out = out || !fin;
                const int64_t Jd = I[d] + s[d];
                out = out || Jd < 0 || Jd >= g.ni[d];
                stay = stay && s[d] == 0;
                far = far || s[d] > 1 || s[d] < -1;
                jump = s[d] < -jump ? -s[d] : (s[d] > jump ? s[d] : jump);
                n += (s[d] + 1) * mult;
                mult *= 3;
            }
            if (stay && !out) continue;
            mv8 |= 1u << j;
            uint8_t cd;
            if (out) {
                cd = kDel;
            } else if (far) {  // staged as it is: the far-jump path renormalises it
                atomicOr(flag, kFlagFar);
                atomicMax(flag + 1, jump);
                cd = kFar;
                stg |= 1u << j;
            } else {
                cd = (uint8_t)n;
#pragma unroll
                for (int d = 0; d < D; ++d)  // the staged value: the offset in the new cell
                    ub[j][d] = (float)((double)ub[j][d] - (double)s[d]);
                stg |= 1u << j;
            }
            word |= (uint64_t)cd << (8 * j);
        }
        if (mv8) cw[(k0 >> 3) * C] = word;
        mov |= (uint64_t)mv8 << k0;
        // the chunk's staged movers, packed after the cell's earlier ones
        const int64_t o0 = chunk_record<true>(c, g.K, k0, stg, staged, s_cnt);
        if (cs.live) {
            if (D == 2 && nargs == 1) {  // 16-byte records (the scatter reads them whole)
#pragma unroll
                for (int j = 0; j < B; ++j)
                    if ((stg >> j) & 1)
                        stg_put16(ms, o0 + __popc(stg & ((1u << j) - 1)), ub[j][0],
                                  ub[j][D - 1], fb[j]);
            } else {
#pragma unroll
                for (int j = 0; j < B; ++j)
                    if ((stg >> j) & 1) {
                        const int64_t o = o0 + __popc(stg & ((1u << j) - 1));
#pragma unroll
                        for (int d = 0; d < D; ++d) stg_coord<D, float>(ms, o, d) = ub[j][d];
                        if (f0) stg_field<D>(ms, o, 0) = fb[j];
                    }
                stage_args<D>(c, C, k0, stg, o0, nargs, args, ms, 1);
            }
        }
        staged |= (uint64_t)stg << k0;
    }
    if (!cs.live) return;
    store_cell_words(ms, c, kb, ke, g.K, occ, mov);
    reinterpret_cast<uint8_t*>(ms.stm + c)[kb >> 3] = (uint8_t)(staged >> kb);
}

ArgPtrs argptrs(double* const* a, int n) {
    ArgPtrs p;
    for (int i = 0; i < JP_MAX_ARGS; ++i) p.p[i] = i < n ? a[i] : nullptr;
    return p;
}

template <int D, typename T>
TPtr3<D, T> tptr3(T* const* c) {
    TPtr3<D, T> p{{nullptr, nullptr, nullptr}};
    for (int d = 0; d < D; ++d) p.p[d] = c[d];
    return p;
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// classify -> colour phases (2D: one tiled launch; 3D: 3^D launches) -> far-
// jump phases (one workgroup; returns at once unless a particle jumped more
// than one cell, in which case the regular phases returned at once) ->
// scatter
template <int D, typename T>
int run_move(jp_ctx* ctx, const jp_cellgrid* gc, T* const* coords, uint8_t* index, int nargs,
             double* const* args, jp_stats* stats) {
    Grid<D> g = to_grid<D>(gc);
    const int64_t C = g.C, S = C * g.K;
    if (S >= ((int64_t)1 << 32))
        return fail(ctx, JP_ERR_ARG, "more than 2^32 slots%s");
    MoveScratch ms;
    size_t off = 0;
    const size_t o_occ0 = off; off = align_up(off + C * 8, 256);
    const size_t o_occ = off;  off = align_up(off + C * 8, 256);
    const size_t o_mov = off;  off = align_up(off + C * 8, 256);
    const size_t o_arr = off;  off = align_up(off + C * 8, 256);
    const size_t o_tch = off;  off = align_up(off + C * 8, 256);
    const size_t o_code = off; off = align_up(off + (size_t)C * 64, 256);
    const size_t o_org = off;  off = align_up(off + (size_t)S * 4, 256);
    const size_t o_stm = off;  off = align_up(off + C * 8, 256);
    ms.S = S;
    ms.foff = (int64_t)align_up(sizeof(T) * D, 8);
    ms.stride = ms.foff + 8 * nargs;
    const size_t stg_bytes = (size_t)S * (size_t)ms.stride;
    const size_t o_stg = off;  off = align_up(off + stg_bytes, 256);
    int rc = ensure_scratch(ctx, off);
    if (rc) return rc;
    char* base = static_cast<char*>(ctx->scratch);
    ms.occ0 = (uint64_t*)(base + o_occ0);
    ms.occ = (uint64_t*)(base + o_occ);
    ms.mov = (uint64_t*)(base + o_mov);
    ms.arr = (uint64_t*)(base + o_arr);
    ms.tch = (uint64_t*)(base + o_tch);
    ms.code = (uint8_t*)(base + o_code);
    ms.org = (uint32_t*)(base + o_org);
    ms.stm = (uint64_t*)(base + o_stm);
    ms.stg = base + o_stg;

    int* flag = dflags(ctx);
    // the flag words are zero here: a new context starts with them cleared,
    // the far-jump kernel, which runs on every move after the regular phases,
    // clears them after use, and a launch that fails after classify clears
    // them here
    auto launched = [&]() -> int {
        const hipError_t e = hipGetLastError();
        if (e == hipSuccess) return JP_OK;
        if (!ctx->capturing) {
            (void)hipStreamSynchronize(ctx->stream);
            (void)hipMemset(flag, 0, 2 * sizeof(int));
        }
        return fail(ctx, JP_ERR_HIP, "HIP error %s", hipGetErrorString(e));
    };
    if ((rc = stats_begin(ctx, stats))) return rc;
    jp_stats* st = stats ? dstats(ctx) : nullptr;
    ArgPtrs ap = argptrs(args, nargs);
    TPtr3<D, T> cp = tptr3<D, T>(coords);
    // per-cell kernels: slot groups on blockIdx.y when the cells alone do not
    // fill the chip (up to 24 workgroups per CU: with one 8-slot chunk per
    // thread, no chunk's loads wait for an earlier chunk's stores)
    const int span = slot_span(ctx, blocks_for(C), g.K, 6144);
    const dim3 cgrid((unsigned)blocks_for(C), (unsigned)((g.K + span - 1) / span));
    // classify (classify_slots): a workgroup of one wavefront per 8-slot chunk
    // for each 64 cells (2D, 3D split), else a thread per cell
    const bool chunks = D == 2 || !std::is_same<T, double>::value;
    const unsigned kblocks = (unsigned)(chunks ? (C + 63) / 64 : blocks_for(C));
    const unsigned kthreads = (unsigned)(chunks ? 64 * ((g.K + 7) / 8) : kBlock);
    if constexpr (std::is_same<T, double>::value) {
        CPtr3 cc{{nullptr, nullptr, nullptr}};
        for (int d = 0; d < D; ++d) cc.p[d] = coords[d];
        k_mv_classify<D><<<kblocks, kthreads, 0, ctx->stream>>>(g, cc, index, nargs, ap, ms,
                                                               flag);
    } else {
        if ((rc = split_geometry(ctx, gc, nullptr))) return rc;
        k_mv_classify_split<D><<<kblocks, kthreads, 0, ctx->stream>>>(g, cp, index, nargs, ap,
                                                                     ms, flag, geo_ptr(ctx));
    }
    if ((rc = launched())) return rc;
    int64_t total = 1;
    for (int d = 0; d < D; ++d) total *= (g.ni[d] + 2) / 3;
    constexpr int BS = D == 2 ? 128 : 64;  // measured: 256 and 64 (2D), 128 (3D) slower
    const int nphase = (D == 2 && !ctx->opt_move_phases) ? 0 : kNB<D>;
    if (D == 2 && !ctx->opt_move_phases) {
        constexpr int TX = 32, TY = 16;
        const int64_t tiles = ((g.ni[0] + TX - 1) / TX) * ((g.ni[1] + TY - 1) / TY);
        if (tiles >= ((int64_t)1 << 31) || g.ni[0] >= ((int64_t)1 << 30) ||
            g.ni[1] >= ((int64_t)1 << 30))
            return fail(ctx, JP_ERR_ARG, "grid too large%s");
        Grid<2> g2 = to_grid<2>(gc);
        if (st)
            k_mv_tiles2<TX, TY, 256, true><<<(unsigned)tiles, 256, 0, ctx->stream>>>(g2, ms, flag,
                                                                                   st);
        else
            k_mv_tiles2<TX, TY, 256, false><<<(unsigned)tiles, 256, 0, ctx->stream>>>(g2, ms,
                                                                                    flag, st);
        if ((rc = launched())) return rc;
    }
    for (int ph = 0; ph < nphase; ++ph) {
        Off3 o{{0, 0, 0}};
        int r = ph;
        for (int d = D - 1; d >= 0; --d) {
            o.o[d] = r % 3;
            r /= 3;
        }
        k_mv_phase<D, BS><<<(unsigned)((total + BS - 1) / BS), BS, 0, ctx->stream>>>(g, ms, o, ph,
                                                                                    flag, st);
        if ((rc = launched())) return rc;
    }
    k_mv_far<D, T><<<1, kFarBlock, 0, ctx->stream>>>(g, ms, flag, st);
    if ((rc = launched())) return rc;
    const SplitGeom* geo = std::is_same<T, double>::value ? nullptr : geo_ptr(ctx);
    bool rec16 = false;  // 2D split, one field: 16-byte records
    if constexpr (D == 2 && std::is_same<T, float>::value) {
        rec16 = nargs == 1;
        if (rec16)
            k_mv_scatter<D, T, true><<<cgrid, kBlock, 0, ctx->stream>>>(g, cp, index, nargs, ap,
                                                                        ms, geo, span);
    }
    if (!rec16)
        k_mv_scatter<D, T><<<cgrid, kBlock, 0, ctx->stream>>>(g, cp, index, nargs, ap, ms, geo,
                                                             span);
    if ((rc = launched())) return rc;

    if ((rc = stats_end(ctx, stats))) return rc;
    return std::is_same<T, double>::value ? finish(ctx) : split_finish(ctx);
}

int check_ptrs(jp_ctx* ctx, int n, const void* const* p, const char* what) {
    if (!p) return fail(ctx, JP_ERR_ARG, "null %s array", what);
    for (int i = 0; i < n; ++i)
        if (!p[i]) return fail(ctx, JP_ERR_ARG, "null %s pointer %lld", what, i);
    return JP_OK;
}

template <typename T>
int move_entry(jp_ctx* ctx, const jp_cellgrid* g, T* const* coords, uint8_t* index, int nargs,
               double* const* args, jp_stats* stats) {
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if ((rc = check_ptrs(ctx, g->ndim, (const void* const*)coords, "coords"))) return rc;
    if (!index) return fail(ctx, JP_ERR_ARG, "null index%s");
    if (nargs < 0 || nargs > JP_MAX_ARGS) return fail(ctx, JP_ERR_ARG, "nargs out of range%s");
    if (nargs && (rc = check_ptrs(ctx, nargs, (const void* const*)args, "args"))) return rc;
    return g->ndim == 2 ? run_move<2, T>(ctx, g, coords, index, nargs, args, stats)
                        : run_move<3, T>(ctx, g, coords, index, nargs, args, stats);
}

}  // namespace

extern "C" int jp_move_particles(jp_ctx* ctx, const jp_cellgrid* g, double* const* coords,
                                 uint8_t* index, int nargs, double* const* args, jp_stats* stats) {
    return move_entry<double>(ctx, g, coords, index, nargs, args, stats);
}

extern "C" int jp_move_particles_split(jp_ctx* ctx, const jp_cellgrid* g, float* const* off,
                                       uint8_t* index, int nargs, double* const* args,
                                       jp_stats* stats) {
    return move_entry<float>(ctx, g, off, index, nargs, args, stats);
}
