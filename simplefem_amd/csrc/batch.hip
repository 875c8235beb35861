// batch.hip -- persistent multi-system Rainbow CG-SSOR engine.
//
// Many independent systems (each a matrix + one right-hand side, a "slot")
// advance together.  Every slot runs a small program of phases:
//
//   restart:   RESID  FWD_R[0..Gf)  BWD_R[0..Gb)  DINIT
//   iteration: SPMV   FWD[0..Gf)    BWD[0..Gb)    DIR          (repeat)
//   mixed:     RES64 -> CONVERT -> (restart, iterations in FP32) -> CORRECT -> RES64 ...
//
// A mixed slot's inner solves after the first continue the previous one's
// search direction (residual replacement): their DINIT forms z.r only, and a
// DIR phase with beta = z.r / z.r_prev makes d = z + beta d_prev -- or, when
// the new z.r exceeds 4x the recursive one the last inner solve ended with
// (its FP32 residual drifted), beta = 0, a restart.  A continued inner solve
// stops after max(100, 3 K_prev) iterations.
//
// the same phases as the single-system kernels, with the same per-row
// arithmetic.  A "step" executes the current phase of every active slot; a
// phase is split into work units of 64 rows (one SELL-64 slice, or 64
// vector rows), each run by one wavefront.  A step is four kernels:
//
//   k_batch_step    (two launches: the units of complete interleaved groups
//                   of four right-hand sides, then all others) every unit of
//                   the step's schedule (each wavefront grid-strides over
//                   units on its own); each unit stores its partial sum in
//                   its slot's partials[j];
//   k_batch_fold    one block per slot: sums the slot's unit partials in unit
//                   order, runs the slot's scalar step (alpha / beta / exit
//                   tests) and advances its program, then counts the units
//                   of its next phase;
//   k_batch_prefix  one block: prefix sums of the unit counts -> the next
//                   step's schedule, and the step counter.
//
// Kernel boundaries order the four, so no kernel waits on another block:
// there are no atomics, fences or spin loops.  The host replays a captured
// graph of a fixed number of steps; a step with no scheduled units is a
// no-op, so replays past the end change nothing.  Between replays the host
// admits new slots and collects finished ones: a new slot is counted by the
// next k_batch_fold and joins the schedule that k_batch_prefix builds.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/simplefem_amd.h"
#include "device_common.hpp"
#include "host_common.hpp"
#include "kernels.hpp"
#include "launch.hpp"
#include "plan.hpp"

namespace sfa {
namespace {

using namespace dev;

constexpr int kUnitBlock = 256;             // threads per block of the engine's kernels
constexpr int kUnitRows = 64;               // rows per work unit: one wavefront, one slice
constexpr int kStepWaves = kUnitBlock / 64; // wavefronts per block
constexpr int kRiders = 4;                  // most adjacent slots one unit serves
constexpr int kXcdChunk = 32;               // consecutive units one XCD takes at a time
constexpr int kMaxSlots = 1024;

enum Phase : int {
    PH_RESID = 0, PH_FWD_R, PH_BWD_R, PH_DINIT,
    PH_SPMV, PH_FWD, PH_BWD, PH_DIR,
    PH_RES64, PH_CONVERT, PH_CORRECT,
    PH_DONE
};

template <typename T>
struct BMat {
    SellViewT<T> full, lower, upper;
    SellViewT<double> full64;       // FP64 copy of `full` (mixed precision)
    const int* fslice;              // gf+1 slice offsets of the forward groups
    const int* frow0;               // first row of each forward group (-1: row list)
    const int* frows;               // rows per forward group
    const int* bslice;
    const int* brow0;
    const int* brows;
    const T* diag;
    int gf, gb;
    int n, n_active;
    int full_slices, full_row0;
    int spmv_lanes;                 // lanes per row of the iteration's SpMV (1, 2, 4, 8)
    T omega, c2;
};

template <typename T>
struct BSlot {
    int mat;
    int kind, group;
    int units;          // units of the phase in the step being scheduled
    int mixed;
    int cont;           // mixed: inner solves continue the previous direction
    int K, maxK, exit_reason, outer, max_outer, inner_total;
    double zr, dAd, alpha, beta, tol2, rel2, target64, res64;
    double zr_cont;     // > 0: z.r of the iterate that formed d, which the next inner solve continues
    int k_cont;         // iterations of the inner solve that left that direction
    // The slot's vectors are lane `lane` of a group of `ls` slots submitted
    // together, stored interleaved: element i of each vector at v[i * ls]
    // (v points at the lane's first element).  A slot submitted alone has
    // ls = 1.  vgroup: the group's first slot.
    int ls, lane, vgroup;
    T *x, *r, *b, *d, *z, *q, *m;
    double *x64, *b64, *r64;
    double* partials;
    double* grp;
};

struct BSched {
    unsigned step;      // steps completed
    int nslots;
    int W[2];           // units of the step of parity p: slots outside complete groups
    int WG[2];          // and leaders of complete interleaved groups (see k_batch_prefix)
    int active;         // slots with units in the next step
    int pad;
    int start[2][kMaxSlots + 1];  // unit offsets of the slots in the step of parity p
    int startG[2][kMaxSlots + 1]; // the same for the complete groups' units
    int ran[2][kMaxSlots];        // units of the phase each slot runs in that step
    int riders[2][kMaxSlots];     // slots a leading slot's units serve (itself included), else 0
};

template <typename T>
struct BArgs {
    BMat<T>* mats;
    BSlot<T>* slots;
    BSched* sched;
};

__device__ inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// Units of a phase; reduction phases always get at least one unit.
template <typename T>
__device__ int phase_units(const BMat<T>& M, int kind, int g)
{
    switch (kind) {
    case PH_RESID: return M.full_slices;
    case PH_SPMV: return max(1, M.full_slices * M.spmv_lanes);
    case PH_RES64: return max(1, M.full_slices);
    case PH_FWD_R:
    case PH_FWD: return M.fslice[g + 1] - M.fslice[g];
    case PH_BWD_R: return M.bslice[g + 1] - M.bslice[g];
    case PH_BWD: return max(1, M.bslice[g + 1] - M.bslice[g]);
    case PH_DINIT: return max(1, cdiv(M.n_active, kUnitRows));
    case PH_DIR:
    case PH_CONVERT:
    case PH_CORRECT: return cdiv(M.n_active, kUnitRows);
    default: return 0;
    }
}

// End of the T solve: mixed slots apply their correction, others stop.
template <typename T>
__device__ void finish_solve(BSlot<T>& S)
{
    const bool failed = S.exit_reason == SFA_EXIT_BREAKDOWN || S.exit_reason == SFA_EXIT_SMALL;
    if (S.mixed && !failed) {
        S.kind = PH_CORRECT;
        S.group = 0;
    } else {
        S.kind = PH_DONE;
    }
}

// Program order after a phase without a decision to make.
template <typename T>
__device__ void advance(BSlot<T>& S, const BMat<T>& M)
{
    const int g = S.group;
    switch (S.kind) {
    case PH_RESID:
        if (M.gf > 0) S.kind = PH_FWD_R, S.group = 0;
        else if (M.gb > 0) S.kind = PH_BWD_R, S.group = 0;
        else S.kind = PH_DINIT;
        break;
    case PH_FWD_R:
        if (g + 1 < M.gf) S.group = g + 1;
        else if (M.gb > 0) S.kind = PH_BWD_R, S.group = 0;
        else S.kind = PH_DINIT;
        break;
    case PH_BWD_R:
        if (g + 1 < M.gb) S.group = g + 1;
        else S.kind = PH_DINIT, S.group = 0;
        break;
    case PH_FWD:
        if (g + 1 < M.gf) S.group = g + 1;
        else S.kind = PH_BWD, S.group = 0;
        break;
    case PH_DIR: S.kind = PH_SPMV, S.group = 0; break;
    case PH_CONVERT: S.kind = PH_RESID, S.group = 0; break;
    case PH_CORRECT: S.kind = PH_RES64, S.group = 0; break;
    default: break;
    }
}

// Scalar step of a slot after its phase's last unit; `total` is the phase's
// reduction (unit partials summed in unit order).
template <typename T>
__device__ void phase_complete(BSlot<T>& S, const BMat<T>& M, double total)
{
    switch (S.kind) {
    case PH_DINIT: {  // solver/sparsesolver.cpp:1156-1190
        const double zr_end = S.zr;  // recursive z.r the last inner solve ended with
        S.zr = total;
        if (S.rel2 > 0.) S.tol2 = S.rel2 * ref_absolute(total);
        if (ref_absolute(total) < S.tol2) {
            S.zr_cont = 0.;
            S.exit_reason = SFA_EXIT_CONVERGED;
            finish_solve(S);
        } else if (S.zr_cont > 0.) {  // d still holds the previous direction
            if (total <= 1.25 * zr_end) {  // continue it
                S.beta = total / S.zr_cont;
                S.maxK = min(M.n, max(100, 3 * S.k_cont));
            } else {  // the FP32 residual drifted: d = z
                S.beta = 0.;
            }
            S.zr_cont = 0.;
            S.kind = PH_DIR, S.group = 0;
        } else {
            S.kind = PH_SPMV, S.group = 0;
        }
        return;
    }
    case PH_SPMV: {  // :1215-1231
        S.dAd = total;
        const bool num_bad = !isfinite(S.zr) || S.zr < 0.;
        const bool den_bad = !isfinite(total) || total < 0.;
        const bool small = sqrt(ref_absolute(total)) < 2.220446049250313e-16;
        if (num_bad || den_bad) {
            S.exit_reason = SFA_EXIT_BREAKDOWN;
            finish_solve(S);
        } else if (small) {
            S.exit_reason = SFA_EXIT_SMALL;
            finish_solve(S);
        } else {
            S.alpha = S.zr / total;
            if (M.gf > 0) S.kind = PH_FWD, S.group = 0;
            else S.kind = PH_BWD, S.group = 0;
        }
        return;
    }
    case PH_BWD: {  // :1263-1307, summed per colour then over colours
        S.grp[S.group] = total;
        if (S.group + 1 < M.gb) {
            S.group++;
            return;
        }
        double zr_new = 0.;
        for (int g = 0; g < M.gb; g++) zr_new += S.grp[g];
        const double zr_old = S.zr;
        S.beta = zr_new / S.zr;
        S.zr = zr_new;
        S.K++;
        if (S.mixed) S.inner_total++;
        if (S.K >= S.maxK) {
            S.exit_reason = SFA_EXIT_CAP;
            finish_solve(S);
        } else if (ref_absolute(zr_new) < S.tol2) {
            S.exit_reason = SFA_EXIT_CONVERGED;
            if (S.mixed && S.cont && zr_old > 0. && isfinite(zr_old)) {
                S.zr_cont = zr_old;
                S.k_cont = S.K;
            }
            finish_solve(S);
        } else {
            S.kind = PH_DIR, S.group = 0;
        }
        return;
    }
    case PH_RES64:
        S.res64 = sqrt(total);
        if (!(S.res64 > S.target64) || S.outer >= S.max_outer) {
            S.kind = PH_DONE;
        } else {
            S.kind = PH_CONVERT, S.group = 0;
        }
        return;
    case PH_CONVERT:  // a new inner solve starts from x = 0
        S.K = 0;
        S.maxK = M.n;
        S.exit_reason = SFA_EXIT_CAP;
        advance(S, M);
        return;
    case PH_CORRECT:
        S.outer++;
        advance(S, M);
        return;
    default:
        advance(S, M);
        return;
    }
}

// acc[k] (+|-)= sum_e v_e * g[k][c_e] over one SELL row for the kRiders
// lanes of one interleaved group (lane k's element i at base[i * kRiders +
// k]): the four values of a column come in one vector load.  Each rider's sum runs over the entries in order, as sell_row forms
// it (the unroll groups loads only).
template <bool SUB, typename TV>
__device__ inline void sell_row_group(const SellViewT<TV>& A, int slice, int lane,
                                      const TV* __restrict__ base, TV* acc)
{
    typedef TV V4 __attribute__((ext_vector_type(kRiders)));
    const V4* g4 = reinterpret_cast<const V4*>(base);
    const int beg = A.off[slice] + lane;
    const int end = A.off[slice + 1];
    // entries per unrolled step: the gathered values held are U * kRiders
    // (FP64 U = 2 and FP32 U = 4 keep the group kernel at 68 / 57 VGPRs)
    constexpr int U = sizeof(TV) == 8 ? 2 : 4;
    int e = beg;
    for (; e + (U - 1) * kWave < end; e += U * kWave) {
        unsigned c[U];
        TV v[U];
        V4 g[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            c[u] = (unsigned)__builtin_nontemporal_load(&A.cols[e + u * kWave]);
            v[u] = __builtin_nontemporal_load(&A.vals[e + u * kWave]);
        }
#pragma unroll
        for (int u = 0; u < U; u++) g[u] = g4[c[u]];
#pragma unroll
        for (int u = 0; u < U; u++) {
#pragma unroll
            for (int k = 0; k < kRiders; k++) {
                if (SUB) acc[k] -= v[u] * g[u][k];
                else     acc[k] += v[u] * g[u][k];
            }
        }
    }
    for (; e < end; e += kWave) {
        const unsigned c = (unsigned)__builtin_nontemporal_load(&A.cols[e]);
        const TV v = __builtin_nontemporal_load(&A.vals[e]);
        const V4 g = g4[c];
#pragma unroll
        for (int k = 0; k < kRiders; k++) {
            if (SUB) acc[k] -= v * g[k];
            else     acc[k] += v * g[k];
        }
    }
}

// acc[k] (+|-)= sum_e v_e * vec[k][c_e * ls] over one SELL row for riders
// k < nr whose vectors have stride ls, each entry's column and value loaded
// once for all riders; every rider's sum runs over the entries in order, as
// sell_row forms it.  G: acc[] is indexed by the lanes of an interleaved
// group whose base address (lane 0 of row 0) is vec[0] (see run_unit),
// which takes sell_row_group.
template <bool SUB, bool G, typename TV>
__device__ inline void sell_row_multi(const SellViewT<TV>& A, int slice, int lane,
                                      const TV* const* vec, int nr, int ls, TV* acc)
{
    if constexpr (G) {
        sell_row_group<SUB>(A, slice, lane, vec[0], acc);
        return;
    }
    const unsigned st = (unsigned)ls;
    const int beg = A.off[slice] + lane;
    const int end = A.off[slice + 1];
    constexpr int U = 8;
    int e = beg;
    for (; e + (U - 1) * kWave < end; e += U * kWave) {
        unsigned c[U];
        TV v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            c[u] = (unsigned)__builtin_nontemporal_load(&A.cols[e + u * kWave]) * st;
            v[u] = __builtin_nontemporal_load(&A.vals[e + u * kWave]);
        }
#pragma unroll
        for (int k = 0; k < kRiders; k++) {
            if (k < nr) {
                TV g[U];
#pragma unroll
                for (int u = 0; u < U; u++) g[u] = vec[k][c[u]];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    if (SUB) acc[k] -= v[u] * g[u];
                    else     acc[k] += v[u] * g[u];
                }
            }
        }
    }
    for (; e < end; e += kWave) {
        const unsigned c = (unsigned)__builtin_nontemporal_load(&A.cols[e]) * st;
        const TV v = __builtin_nontemporal_load(&A.vals[e]);
#pragma unroll
        for (int k = 0; k < kRiders; k++) {
            if (k < nr) {
                if (SUB) acc[k] -= v * vec[k][c];
                else     acc[k] += v * vec[k][c];
            }
        }
    }
}

// q = A d over the rows of SpMV task `task` of a matrix whose SpMV runs K
// lanes per row (task i: slice i/K, part i%K of 64/K rows), for vectors
// stored with stride ls; returns the lane's share of d.Ad (the row's first
// lane holds the row's product).
template <int K, typename T>
__device__ double spmv_task(const BMat<T>& M, const T* __restrict__ d, T* __restrict__ q,
                            int task, int lane, int ls)
{
    const int slice = task / K;
    if (slice >= M.full_slices) return 0.;
    const int row = group_row_at(M.full, slice, slice, sellk_slice_row<K>(task % K, lane),
                                 M.full_row0, M.n_active);
    if (row < 0) return 0.;  // every lane of a row agrees
    T acc;
    if constexpr (K == 1) {  // the step kernel's one-lane SpMV code, for one rider
        const T* ds[kRiders] = {d, d, d, d};
        T a[kRiders] = {T(0), T(0), T(0), T(0)};
        sell_row_multi<false, false>(M.full, slice, lane, ds, 1, ls, a);
        acc = a[0];
    } else {
        SellChunk<T, 8> R;
        sellk_prefetch<K>(M.full, slice, slice, 0, -1, task % K, lane, R);
        acc = sellk_finish<K, false>(M.full, R, [&](unsigned c) { return d[c * (unsigned)ls]; },
                                     T(0), lane);
    }
    if (lane % K != 0) return 0.;
    const unsigned i = (unsigned)row * (unsigned)ls;
    q[i] = acc;
    return (double)(d[i] * acc);
}

template <typename T>
__device__ double spmv_task_any(const BMat<T>& M, const T* d, T* q, int task, int lane, int ls)
{
    switch (M.spmv_lanes) {
    case 2: return spmv_task<2>(M, d, q, task, lane, ls);
    case 4: return spmv_task<4>(M, d, q, task, lane, ls);
    case 8: return spmv_task<8>(M, d, q, task, lane, ls);
    default: return spmv_task<1>(M, d, q, task, lane, ls);
    }
}

// Row `row` of the vector that member vp of each rider points to (T or
// double), for riders k < nr with vectors of stride ls.  G: the riders are
// lanes lo .. lo+nr-1 of an interleaved group of kRiders lanes, SS[0] the
// one in lane lo; the group's four values of the row come in one vector
// access, indexed by lane (lanes that do not ride are read and ignored).
// The group's base address is taken from SS[0]'s own pointer, never from
// the slot of lane 0, which may have been collected and reused.
template <bool G, typename V, typename T>
__device__ inline void row_load(const BSlot<T>* SS, int lo, int nr, int ls, V* BSlot<T>::*vp,
                                int row, V* out)
{
    if constexpr (G) {
        typedef V V4 __attribute__((ext_vector_type(kRiders)));
        const V* base = SS[0].*vp - lo;
        const V4 v = *reinterpret_cast<const V4*>(base + (unsigned)row * kRiders);
#pragma unroll
        for (int k = 0; k < kRiders; k++) out[k] = v[k];
    } else {
#pragma unroll
        for (int k = 0; k < kRiders; k++)
            if (k < nr) out[k] = (SS[k].*vp)[(unsigned)row * (unsigned)ls];
    }
}

// Stores in[k] for the riders (G: lanes) lo <= k < hi.  G with all four
// lanes riding: one vector store.
template <bool G, typename V, typename T>
__device__ inline void row_store(const BSlot<T>* SS, int lo, int hi, int ls, V* BSlot<T>::*vp,
                                 int row, const V* in)
{
    if constexpr (G) {
        V* base = SS[0].*vp - lo + (unsigned)row * kRiders;
        if (lo == 0 && hi == kRiders) {
            typedef V V4 __attribute__((ext_vector_type(kRiders)));
            V4 v;
#pragma unroll
            for (int k = 0; k < kRiders; k++) v[k] = in[k];
            *reinterpret_cast<V4*>(base) = v;
        } else {
#pragma unroll
            for (int k = 0; k < kRiders; k++)
                if (k >= lo && k < hi) base[k] = in[k];
        }
    } else {
#pragma unroll
        for (int k = 0; k < kRiders; k++)
            if (k < hi) (SS[k].*vp)[(unsigned)row * (unsigned)ls] = in[k];
    }
}

// One work unit of the current phase of slots SS[0..nr) -- adjacent slots of
// one matrix in the same phase ("riders"; nr = 1 for a slot on its own),
// whose vectors all have the stride ls -- run by one wavefront: slice j of
// the phase's part (SpMV task j), or vector rows 64j .. 64j+63.  Matrix
// entries are loaded once for all riders; each rider's rows get exactly the
// arithmetic they get alone.  part[k]: the lane's share of rider k's
// reduction.
//
// G: the riders are lanes lo .. lo+nr-1 of an interleaved group of kRiders
// (lo = SS[0].lane).  The unit then works in lane space: every loaded row and
// gathered column brings all four lanes in one vector access, the
// arithmetic runs for all four (a lane that does not ride computes values
// nobody stores, from the scalars of SS[0]), only the riding lanes' values
// are stored, and part[] is indexed by lane.  Never with a multi-lane SpMV
// (k_batch_prefix leaves those to G = false).
template <typename T, bool WIDE, bool G>
__device__ void run_unit(const BSlot<T>* SS, int nr, const BMat<T>& M, int kind, int g, int j,
                         double* part)
{
    using S_ = BSlot<T>;
    const int lane = threadIdx.x & 63;
    const int lo = G ? SS[0].lane : 0;  // riders (G: lanes) lo .. hi-1 are stored
    const int hi = lo + nr;
    const int ls = G ? kRiders : SS[0].ls;
    const int nl = G ? kRiders : nr;    // riders (G: lanes) computed
    // the slot of rider (G: lane) k, for its scalars; SS[0] for a lane that
    // does not ride
    auto slot = [&](int k) -> const S_& { return SS[(k >= lo && k < hi) ? k - lo : 0]; };
    // the vectors gathered by sell_row_multi: per rider, or (G) the group's base
    auto gathered = [&](auto vp, auto* out) {
#pragma unroll
        for (int k = 0; k < kRiders; k++)
            out[k] = G ? SS[0].*vp - lo : (k < nr ? SS[k].*vp : SS[0].*vp);
    };
#pragma unroll
    for (int k = 0; k < kRiders; k++) part[k] = 0.;
    switch (kind) {
    case PH_RESID: {
        const int local = j;
        if (local < M.full_slices) {
            const int row = group_row(M.full, local, local, lane, M.full_row0, M.n_active);
            if (row >= 0) {
                const T* xs[kRiders];
                T acc[kRiders];
                gathered(&S_::x, xs);
#pragma unroll
                for (int k = 0; k < kRiders; k++) acc[k] = T(0);
                row_load<G>(SS, lo, nr, ls, &S_::b, row, acc);
                sell_row_multi<true, G>(M.full, local, lane, xs, nr, ls, acc);
                row_store<G>(SS, lo, hi, ls, &S_::r, row, acc);
            }
        }
        break;
    }
    case PH_SPMV:
        if (G || !WIDE || M.spmv_lanes == 1) {
            const int local = j;
            if (local < M.full_slices) {
                const int row = group_row(M.full, local, local, lane, M.full_row0, M.n_active);
                if (row >= 0) {
                    const T* ds[kRiders];
                    T acc[kRiders], dr[kRiders];
                    gathered(&S_::d, ds);
#pragma unroll
                    for (int k = 0; k < kRiders; k++) acc[k] = dr[k] = T(0);
                    sell_row_multi<false, G>(M.full, local, lane, ds, nr, ls, acc);
                    row_store<G>(SS, lo, hi, ls, &S_::q, row, acc);
                    row_load<G>(SS, lo, nr, ls, &S_::d, row, dr);
#pragma unroll
                    for (int k = 0; k < kRiders; k++)
                        if (k < nl) part[k] = (double)(dr[k] * acc[k]);
                }
            }
        } else if constexpr (WIDE && !G) {
            for (int k = 0; k < nr; k++) part[k] = spmv_task_any(M, SS[k].d, SS[k].q, j, lane, ls);
        }
        break;
    case PH_RES64: {
        if constexpr (G) break;  // never scheduled here (k_batch_prefix)
        const int local = j;
        if (local < M.full_slices) {
            const int row = group_row(M.full64, local, local, lane, M.full_row0, M.n_active);
            if (row >= 0) {
                const double* xs[kRiders];
                double acc[kRiders];
                gathered(&S_::x64, xs);
#pragma unroll
                for (int k = 0; k < kRiders; k++) acc[k] = 0.;
                row_load<G>(SS, lo, nr, ls, &S_::b64, row, acc);
                sell_row_multi<true, G>(M.full64, local, lane, xs, nr, ls, acc);
                row_store<G>(SS, lo, hi, ls, &S_::r64, row, acc);
#pragma unroll
                for (int k = 0; k < kRiders; k++)
                    if (k < nl) part[k] = acc[k] * acc[k];
            }
        }
        break;
    }
    case PH_FWD_R:
    case PH_FWD: {
        const int local = j;
        const int slice = M.fslice[g] + local;
        if (slice < M.fslice[g + 1]) {
            const int row = group_row(M.lower, slice, local, lane, M.frow0[g], M.frows[g]);
            if (row >= 0) {
                const T* ms[kRiders];
                T acc[kRiders];
                gathered(&S_::m, ms);
#pragma unroll
                for (int k = 0; k < kRiders; k++) acc[k] = T(0);
                row_load<G>(SS, lo, nr, ls, &S_::r, row, acc);
                if (kind == PH_FWD) {
                    T xv[kRiders], dv[kRiders], qv[kRiders];
#pragma unroll
                    for (int k = 0; k < kRiders; k++) xv[k] = dv[k] = qv[k] = T(0);
                    row_load<G>(SS, lo, nr, ls, &S_::x, row, xv);
                    row_load<G>(SS, lo, nr, ls, &S_::d, row, dv);
                    row_load<G>(SS, lo, nr, ls, &S_::q, row, qv);
#pragma unroll
                    for (int k = 0; k < kRiders; k++) {
                        if (k < nl) {
                            const T alpha = (T)slot(k).alpha;
                            xv[k] += alpha * dv[k];
                            acc[k] = acc[k] - alpha * qv[k];
                        }
                    }
                    row_store<G>(SS, lo, hi, ls, &S_::x, row, xv);
                    row_store<G>(SS, lo, hi, ls, &S_::r, row, acc);
                }
                sell_row_multi<true, G>(M.lower, slice, lane, ms, nr, ls, acc);
                const T dg = M.diag[row];
#pragma unroll
                for (int k = 0; k < kRiders; k++) acc[k] = acc[k] * M.omega / dg;
                row_store<G>(SS, lo, hi, ls, &S_::m, row, acc);
            }
        }
        break;
    }
    case PH_BWD_R:
    case PH_BWD: {
        const int local = j;
        const int slice = M.bslice[g] + local;
        if (slice < M.bslice[g + 1]) {
            const int row = group_row(M.upper, slice, local, lane, M.brow0[g], M.brows[g]);
            if (row >= 0) {
                const T dg = M.diag[row];
                const T* zs[kRiders];
                T acc[kRiders];
                gathered(&S_::z, zs);
#pragma unroll
                for (int k = 0; k < kRiders; k++) acc[k] = T(0);
                row_load<G>(SS, lo, nr, ls, &S_::m, row, acc);
#pragma unroll
                for (int k = 0; k < kRiders; k++) {
                    T t = acc[k];
                    t *= dg;
                    t *= M.c2;
                    acc[k] = t;
                }
                sell_row_multi<true, G>(M.upper, slice, lane, zs, nr, ls, acc);
#pragma unroll
                for (int k = 0; k < kRiders; k++) acc[k] = acc[k] * M.omega / dg;
                row_store<G>(SS, lo, hi, ls, &S_::z, row, acc);
                if (kind == PH_BWD) {
                    T rv[kRiders];
#pragma unroll
                    for (int k = 0; k < kRiders; k++) rv[k] = T(0);
                    row_load<G>(SS, lo, nr, ls, &S_::r, row, rv);
#pragma unroll
                    for (int k = 0; k < kRiders; k++)
                        if (k < nl) part[k] = (double)(acc[k] * rv[k]);
                }
            }
        }
        break;
    }
    case PH_DINIT: {  // a slot that continues its direction keeps d (its DIR phase follows)
        const int row = j * kUnitRows + lane;
        if (row < M.n_active) {
            const unsigned i = (unsigned)row * (unsigned)ls;
            for (int k = lo; k < hi; k++) {
                const S_& S = SS[k - lo];
                const T zz = S.z[i];
                if (!(S.zr_cont > 0.)) S.d[i] = zz;
                part[k] = (double)(zz * S.r[i]);
            }
        }
        break;
    }
    case PH_DIR: {
        const int row = j * kUnitRows + lane;
        if (row < M.n_active) {
            T zv[kRiders], dv[kRiders];
#pragma unroll
            for (int k = 0; k < kRiders; k++) zv[k] = dv[k] = T(0);
            row_load<G>(SS, lo, nr, ls, &S_::z, row, zv);
            row_load<G>(SS, lo, nr, ls, &S_::d, row, dv);
#pragma unroll
            for (int k = 0; k < kRiders; k++)
                if (k < nl) dv[k] = zv[k] + (T)slot(k).beta * dv[k];
            row_store<G>(SS, lo, hi, ls, &S_::d, row, dv);
        }
        break;
    }
    case PH_CONVERT: {
        const int row = j * kUnitRows + lane;
        if (row < M.n_active) {
            const unsigned i = (unsigned)row * (unsigned)ls;
            for (int k = lo; k < hi; k++) {
                const S_& S = SS[k - lo];
                S.b[i] = (T)S.r64[i];
                S.x[i] = T(0);
                S.r[i] = T(0);
            }
        }
        break;
    }
    case PH_CORRECT: {
        const int row = j * kUnitRows + lane;
        if (row < M.n_active) {
            const unsigned i = (unsigned)row * (unsigned)ls;
            for (int k = lo; k < hi; k++) {
                const S_& S = SS[k - lo];
                S.x64[i] += (double)S.x[i];
            }
        }
        break;
    }
    default: break;
    }
}

__device__ inline double block_sum(double v, double* ws)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) ws[w] = v;
    __syncthreads();
    double t = 0.;
    for (int i = 0; i < kUnitBlock / 64; i++) t += ws[i];
    return t;
}

// Units of the slot's current phase, skipping phases that have no units and
// no decision to make.  One thread.
template <typename T>
__device__ int next_units(BSlot<T>& S, const BMat<T>& M)
{
    int units = 0;
    for (int guard = 0; guard < 64; guard++) {
        units = phase_units(M, S.kind, S.group);
        if (units > 0 || S.kind == PH_DONE) break;
        advance(S, M);  // a phase with no rows and nothing to decide
    }
    return S.kind == PH_DONE ? 0 : units;
}

// The slot s with start[s] <= u < start[s+1] (binary search, start in LDS).
__device__ inline int unit_owner(const int* start, int ns, int u)
{
    int lo = 0, hi = ns;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (start[mid] <= u) lo = mid;
        else hi = mid;
    }
    return lo;
}

// WIDE: the SpMV phase may run a matrix with 2, 4 or 8 lanes per row.  Those
// code paths need about twice the registers of everything else (117 vs 66
// VGPRs in FP64: 4 vs 7 wavefronts per SIMD), so the engine runs the variant
// without them whenever every bucket uses one lane per row.
// G: this launch runs the units of riders in interleaved groups of four
// (their own unit table, startG / WG), with the group code paths only;
// G = false runs every other unit.  Each variant's register count then follows its own
// paths (FP64: the group paths in the same kernel took 70 -> 96 VGPRs, 7 -> 5
// wavefronts per SIMD, for all units).
template <typename T, bool WIDE, bool G>
__global__ void __launch_bounds__(kUnitBlock) k_batch_step(BArgs<T> a)
{
    __shared__ int start[kMaxSlots + 1];  // the step's unit offsets, read once per block
    const BSched* sc = a.sched;
    const int p = (int)(sc->step & 1u);
    const int W = G ? sc->WG[p] : sc->W[p];
    const int* ustart = G ? sc->startG[p] : sc->start[p];
    const int ns = sc->nslots;
    // Workgroups reach the XCDs round robin (block i on XCD i % 8), each XCD
    // with its own L2: the step's units are dealt to the XCDs in chunks of
    // kXcdChunk consecutive units (chunk c to XCD c % 8), so neighbouring
    // slices of one matrix, which gather the same vector lines, mostly share
    // an L2, while heavy and light slots still spread over all XCDs.  Which
    // wavefront runs a unit changes nothing it computes.
    const int xcd = blockIdx.x % kXcds, kb = blockIdx.x / kXcds;
    const int per_xcd = gridDim.x / kXcds;
    // the v-th unit of this XCD (increasing in v)
    auto unit_of = [&](int v) { return ((v / kXcdChunk) * kXcds + xcd) * kXcdChunk + v % kXcdChunk; };
    if (unit_of(kb * kStepWaves) >= W) return;
    for (int i = threadIdx.x; i <= ns; i += kUnitBlock) start[i] = ustart[i];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // Each wavefront runs its own units, strided over its XCD's range; the
    // four wavefronts of a block take consecutive units, which mostly belong
    // to one slot, so the slot record and matrix descriptor (wave-uniform,
    // read by scalar loads) come from the scalar cache after the first.  No
    // wavefront waits for another: each unit's partial is a wavefront sum.
    // A unit of a leading slot also serves its riders (see k_batch_prefix).
    const int* riders = sc->riders[p];
    for (int v = kb * kStepWaves + wave;; v += per_xcd * kStepWaves) {
        const int u = unit_of(v);
        if (u >= W) break;
        const int lo = __builtin_amdgcn_readfirstlane(unit_owner(start, ns, u));
        const BSlot<T>* SS = a.slots + lo;
        const int nr = riders[lo];
        const int j = u - start[lo];
        double part[kRiders];
        run_unit<T, WIDE && !G, G>(SS, nr, a.mats[SS[0].mat], SS[0].kind, SS[0].group, j, part);
        // G: part[] is indexed by lane, the riders are lanes k0 .. k0+nr-1
        const int k0 = G ? SS[0].lane : 0;
#pragma unroll
        for (int k = 0; k < kRiders; k++) {
            if (k >= k0 && k < k0 + nr) {
                const double t = wave_sum(part[k]);
                if (lane == 0) SS[k - k0].partials[j] = t;
            }
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(kUnitBlock) k_batch_fold(BArgs<T> a)
{
    __shared__ double ws[kUnitBlock / 64];
    const BSched* sc = a.sched;
    const int s = blockIdx.x;
    if (s >= sc->nslots) return;
    BSlot<T>& S = a.slots[s];
    if (S.mat < 0 || S.kind == PH_DONE) {
        if (threadIdx.x == 0) S.units = 0;
        return;
    }
    const BMat<T>& M = a.mats[S.mat];
    const int p = (int)(sc->step & 1u);
    const int ran = sc->ran[p][s];
    if (ran > 0) {  // the slot's phase ran in this step
        double acc = 0.;
        for (int i = threadIdx.x; i < ran; i += kUnitBlock) acc += S.partials[i];
        const double total = block_sum(acc, ws);
        if (threadIdx.x == 0) phase_complete(S, M, total);
    }
    if (threadIdx.x == 0) S.units = next_units(S, M);
}

constexpr int kPrefixBlock = kMaxSlots;

// Inclusive scans over the block (one value per thread): sum and maximum.
__device__ inline int block_scan_add(int v, int* wtot)
{
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int w = __shfl_up(v, o, 64);
        if (lane >= o) v += w;
    }
    if (lane == 63) wtot[wave] = v;
    __syncthreads();
    if (t == 0) {
        int acc = 0;
        for (int w = 0; w < kPrefixBlock / 64; w++) {
            const int x = wtot[w];
            wtot[w] = acc;
            acc += x;
        }
    }
    __syncthreads();
    v += wtot[wave];
    __syncthreads();
    return v;
}

__device__ inline int block_scan_max(int v, int* wtot)
{
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int w = __shfl_up(v, o, 64);
        if (lane >= o) v = max(v, w);
    }
    if (lane == 63) wtot[wave] = v;
    __syncthreads();
    if (t == 0) {
        int acc = -1;
        for (int w = 0; w < kPrefixBlock / 64; w++) {
            const int x = wtot[w];
            wtot[w] = acc;
            acc = max(acc, x);
        }
    }
    __syncthreads();
    v = max(v, wtot[wave]);
    __syncthreads();
    return v;
}

// The next step's schedule.  Adjacent running slots of one matrix in the
// same phase form runs -- slots submitted alone (e.g. the right-hand sides
// of one system, submitted one after another) or lanes of one interleaved
// group, never a mix of the two; each run is cut into groups of up to
// kRiders slots, and only a group's first slot (its leader) gets units,
// which serve the whole group: the matrix is read once for all its
// right-hand sides.  Which slots ride together changes nothing in any slot's
// arithmetic.  Leaders of riders in interleaved groups of four go to the
// second unit table (startG, WG), run by the group variant of the step
// kernel -- except in an SpMV phase with more than
// one lane per row, which only the other variant has, and the FP64 residual
// phase of mixed solves (once per correction; its FP64 group gathers would
// set the group variant's register count).
template <typename T>
__global__ void __launch_bounds__(kPrefixBlock) k_batch_prefix(BArgs<T> a)
{
    __shared__ int wtot[kPrefixBlock / 64];
    __shared__ int kmat[kPrefixBlock], kkind[kPrefixBlock], kgrp[kPrefixBlock];
    __shared__ int kunits[kPrefixBlock], krun[kPrefixBlock], kvg[kPrefixBlock];
    __shared__ int active;
    BSched* sc = a.sched;
    const int t = threadIdx.x;
    const int ns = sc->nslots;
    const unsigned step = sc->step;
    const int q = (int)((step + 1u) & 1u);
    int u = 0, mat = -1, kind = -1, grp = -1, vg = -1, ls = 1;
    if (t < ns) {
        const BSlot<T>& S = a.slots[t];
        ls = S.ls;
        u = S.units;
        mat = S.mat;
        kind = S.kind;
        grp = S.group;
        vg = S.ls > 1 ? S.vgroup : -1;  // -1: a slot submitted alone
    }
    kmat[t] = mat;
    kvg[t] = vg;
    kkind[t] = kind;
    kgrp[t] = grp;
    kunits[t] = u;
    if (t == 0) active = 0;
    __syncthreads();
    // A lane of an interleaved group about to start an inner solve (mixed
    // engine, PH_CONVERT) waits while another lane of its group is still in
    // its previous one, so the group's inner solves start, and mostly run,
    // together.  Waiting changes when a slot's phases run, never what they
    // compute; the lanes waited for are running and reach PH_CONVERT or
    // PH_DONE, so every wait ends.
    bool wait = false;
    if (u > 0 && ls > 1 && kind == PH_CONVERT)
        for (int l = vg; l < vg + ls; l++)
            if (l != t && kvg[l] == vg && kkind[l] != PH_CONVERT && kkind[l] != PH_DONE)
                wait = true;
    __syncthreads();
    if (wait) {
        kunits[t] = 0;
        u = 0;
    }
    if (u > 0 || wait) atomicAdd(&active, 1);
    __syncthreads();
    // a run starts at a running slot whose predecessor is idle or differs
    const bool joins = u > 0 && t > 0 && kunits[t - 1] > 0 && kmat[t - 1] == mat &&
                       kkind[t - 1] == kind && kgrp[t - 1] == grp && kvg[t - 1] == vg;
    const int run = block_scan_max(joins ? -1 : t, wtot);  // first slot of t's run
    krun[t] = run;
    __syncthreads();
    const bool leader = u > 0 && (t - run) % kRiders == 0;
    int nr = 0;
    if (leader) {
        nr = 1;
        for (int k = 1; k < kRiders && t + k < ns; k++)
            if (kunits[t + k] > 0 && krun[t + k] == run) nr++;
    }
    const bool group = leader && ls == kRiders && kind != PH_RES64 &&
                       !(kind == PH_SPMV && a.mats[mat].spmv_lanes > 1);
    const int v = block_scan_add(leader && !group ? u : 0, wtot);
    const int vG = block_scan_add(group ? u : 0, wtot);
    if (t < ns) {
        sc->start[q][t + 1] = v;
        sc->startG[q][t + 1] = vG;
        sc->ran[q][t] = u;
        sc->riders[q][t] = nr;
    }
    if (t == ns - 1) {
        sc->W[q] = v;
        sc->WG[q] = vG;
    }
    if (t == 0) {
        sc->start[q][0] = 0;
        sc->startG[q][0] = 0;
        sc->active = active;
        sc->step = step + 1u;
    }
}

// SpMV autotuning probe (nm <= kMaxSlots): the PH_SPMV units of the matrices mats[0..nm) with
// K lanes per row, grid-stride as k_batch_step runs them (unit u belongs to
// the matrix m with ustart[m] <= u < ustart[m+1]); d and q are scratch
// vectors of at least n+1 entries for every listed matrix.
template <typename T, int K>
__global__ void __launch_bounds__(kUnitBlock) k_batch_spmv_probe(const BMat<T>* mats,
                                                                 const int* ustart, int nm,
                                                                 const T* d, T* q,
                                                                 double* partials)
{
    __shared__ int start[kMaxSlots + 1];
    const int W = ustart[nm];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // the step kernel's unit-to-XCD chunks
    const int xcd = blockIdx.x % kXcds, kb = blockIdx.x / kXcds;
    const int per_xcd = gridDim.x / kXcds;
    auto unit_of = [&](int v) { return ((v / kXcdChunk) * kXcds + xcd) * kXcdChunk + v % kXcdChunk; };
    if (unit_of(kb * kStepWaves) >= W) return;
    for (int i = threadIdx.x; i <= nm; i += kUnitBlock) start[i] = ustart[i];
    __syncthreads();
    for (int v = kb * kStepWaves + wave;; v += per_xcd * kStepWaves) {
        const int u = unit_of(v);
        if (u >= W) break;
        const int lo = __builtin_amdgcn_readfirstlane(unit_owner(start, nm, u));
        const int j = u - start[lo];
        const double part = wave_sum(spmv_task<K>(mats[lo], d, q, j, lane, 1));
        if (lane == 0) partials[u] = part;
    }
}

// ====================================================================== host

constexpr int kStepBlocks = 2048;  // grid of k_batch_step: 8 blocks of 256 threads per CU
constexpr int kGraphSteps = 32;    // steps per replay of the captured graph

template <typename T>
struct HostMat {
    bool live = false;
    int n = 0, n_active = 0;
    long long bytes = 0;
    std::vector<int> perm;
    DevPartT<T> full, lower, upper;
    DevPartT<double> full64;
    int *fslice = nullptr, *frow0 = nullptr, *frows = nullptr;
    int *bslice = nullptr, *brow0 = nullptr, *brows = nullptr;
    T* diag = nullptr;
    int max_units = 1, gb = 0;
    int slots_in_use = 0;
    int bucket = 0;   // SpMV row-length bucket of the matrix (its full part's main one)
    BMat<T> bm{};     // host copy of the device table entry

    void release()
    {
        full.release();
        lower.release();
        upper.release();
        full64.release();
        for (void* p : {(void*)fslice, (void*)frow0, (void*)frows, (void*)bslice, (void*)brow0,
                        (void*)brows, (void*)diag})
            dev_free(p);
        fslice = frow0 = frows = bslice = brow0 = brows = nullptr;
        diag = nullptr;
        live = false;
    }
};

// The Krylov vectors of a group of `lanes` slots submitted together (device
// order, n+1 rows), interleaved: row i of lane k at v[i * lanes + k].  Freed
// with the group's last slot.
template <typename T>
struct VecBlock {
    T *x = nullptr, *r = nullptr, *b = nullptr, *d = nullptr, *z = nullptr, *q = nullptr,
      *m = nullptr;
    double *x64 = nullptr, *b64 = nullptr, *r64 = nullptr;
    int lanes = 1;
    long long bytes = 0;
    VecBlock() = default;
    VecBlock(const VecBlock&) = delete;
    VecBlock& operator=(const VecBlock&) = delete;
    ~VecBlock()
    {
        for (void* p : {(void*)x, (void*)r, (void*)b, (void*)d, (void*)z, (void*)q, (void*)m,
                        (void*)x64, (void*)b64, (void*)r64})
            dev_free(p);
    }
};

// One slot's storage: its lane of a vector block, and its own unit partials
// and per-colour sums.
template <typename T>
struct SlotVectors {
    std::shared_ptr<VecBlock<T>> block;
    int lane = 0;
    double *partials = nullptr, *grp = nullptr;
    long long bytes = 0;  // partials and grp
    void release()
    {
        block.reset();
        dev_free(partials);
        dev_free(grp);
        *this = SlotVectors();
    }
};

int* upload_ints(const std::vector<int>& v)
{
    int* p = dev_alloc<int>(v.size());
    if (!v.empty()) SFA_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice));
    return p;
}

template <typename T>
struct Engine {
    bool mixed = false;
    bool cont = false;  // mixed: inner solves continue the previous direction
    int max_slots = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<HostMat<T>> mats;
    std::vector<BSlot<T>> slots;              // host mirror
    std::vector<SlotVectors<T>> vecs;
    std::vector<int> slot_used;
    BMat<T>* d_mats = nullptr;
    int d_mats_cap = 0;
    BSlot<T>* d_slots = nullptr;
    BSched* d_sched = nullptr;
    BSched* h_sched = nullptr;                // pinned
    unsigned step = 0;
    hipGraphExec_t graph = nullptr;
    std::unique_ptr<GraphRecorder> graph_rec;  // the recorded graph and its arguments
    BMat<T>* graph_mats = nullptr;            // d_mats the graph was captured with
    bool graph_wide = false;                  // the graph's step kernel variant
    int spmv_table[kSpmvBuckets] = {1, 1, 1, 1};  // lanes per row of the SpMV per bucket

    ~Engine() { release(); }

    void release()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
        graph_rec.reset();
        for (auto& v : vecs) v.release();
        for (auto& m : mats) m.release();
        dev_free(d_mats);
        dev_free(d_slots);
        dev_free(d_sched);
        pinned_free(h_sched);
        if (stream) (void)hipStreamDestroy(stream);
        d_mats = nullptr;
        d_slots = nullptr;
        d_sched = nullptr;
        h_sched = nullptr;
        stream = nullptr;
    }

    void create(int mixed_, int max_slots_)
    {
        if (max_slots_ <= 0 || max_slots_ > kMaxSlots)
            throw Error{SFA_ERR_INVALID, "max_slots must be in 1..1024"};
        ensure_device();
        SFA_HIP(hipGetDevice(&device));
        mixed = mixed_ != 0;
        cont = mixed_ == 1;
        max_slots = max_slots_;
        SFA_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        slots.assign(max_slots, BSlot<T>{});
        for (auto& sl : slots) {
            sl.mat = -1;
            sl.kind = PH_DONE;
        }
        vecs.assign(max_slots, SlotVectors<T>{});
        slot_used.assign(max_slots, 0);
        d_slots = dev_alloc<BSlot<T>>(max_slots);
        SFA_HIP(hipMemcpy(d_slots, slots.data(), sizeof(BSlot<T>) * max_slots, hipMemcpyHostToDevice));
        d_sched = dev_alloc<BSched>(1);
        h_sched = pinned_alloc<BSched>(1);
        std::memset(h_sched, 0, sizeof(BSched));
        h_sched->nslots = max_slots;
        SFA_HIP(hipMemcpy(d_sched, h_sched, sizeof(BSched), hipMemcpyHostToDevice));
    }

    BArgs<T> args() const { return BArgs<T>{d_mats, d_slots, d_sched}; }

    int add_matrix(int n, const int* A, const int* C, const double* V, const double* diagonal,
                   double omega, int nc, const int* F, const int* B, const int* R)
    {
        Plan plan;
        bool colouring_error = false;
        std::string err = build_plan(plan, SFA_METHOD_RAINBOW, n, A, C, V, diagonal, omega, nc, F,
                                     B, R, &colouring_error);
        if (!err.empty()) throw Error{colouring_error ? SFA_ERR_COLOURING : SFA_ERR_INVALID, err};
        return add_plan(plan);
    }

    // Uploads a plan built on the host and registers it; returns its id.
    int add_plan(const Plan& plan)
    {
        const int n = plan.n;
        const double omega = plan.omega;
        SFA_HIP(hipSetDevice(device));
        int id = -1;
        for (size_t i = 0; i < mats.size(); i++)
            if (!mats[i].live) id = (int)i;
        if (id < 0) {
            id = (int)mats.size();
            mats.emplace_back();
        }
        HostMat<T>& M = mats[id];
        M = HostMat<T>();
        M.live = true;
        M.n = n;
        M.n_active = plan.n_active;
        M.perm = plan.perm;
        M.full.upload(plan.full);
        M.lower.upload(plan.lower);
        M.upper.upload(plan.upper);
        if (mixed) M.full64.upload(plan.full);
        M.fslice = upload_ints(plan.lower.group_slice);
        M.frow0 = upload_ints(plan.lower.group_row0);
        std::vector<int> fr(plan.lower.group_rows.begin(), plan.lower.group_rows.end());
        M.frows = upload_ints(fr);
        M.bslice = upload_ints(plan.upper.group_slice);
        M.brow0 = upload_ints(plan.upper.group_row0);
        std::vector<int> br(plan.upper.group_rows.begin(), plan.upper.group_rows.end());
        M.brows = upload_ints(br);
        M.diag = dev_alloc<T>(n + 1);
        std::vector<T> dg(plan.diag.begin(), plan.diag.end());
        dg.push_back(T(0));
        SFA_HIP(hipMemcpy(M.diag, dg.data(), dg.size() * sizeof(T), hipMemcpyHostToDevice));
        M.gb = plan.upper.ngroups();
        int mu = std::max(1, plan.full.nslices);
        mu = std::max(mu, (plan.n_active + kUnitRows - 1) / kUnitRows);
        for (const SellPart* part : {&plan.lower, &plan.upper})
            for (int g = 0; g < part->ngroups(); g++)
                mu = std::max(mu, part->group_slice[g + 1] - part->group_slice[g]);
        M.bucket = M.full.main_bucket();
        mu = std::max(mu, plan.full.nslices * 8);  // SpMV tasks of any mapping
        M.max_units = mu;
        M.bytes = (long long)(plan.full.vals.size() + plan.lower.vals.size() + plan.upper.vals.size()) *
                      (sizeof(T) + 4) + (long long)(n + 1) * sizeof(T) +
                  (mixed ? (long long)plan.full.vals.size() * 12 : 0);

        BMat<T> bm{};
        bm.full = M.full.view();
        bm.lower = M.lower.view();
        bm.upper = M.upper.view();
        bm.full64 = mixed ? M.full64.view() : SellViewT<double>{nullptr, nullptr, nullptr, nullptr};
        bm.fslice = M.fslice;
        bm.frow0 = M.frow0;
        bm.frows = M.frows;
        bm.bslice = M.bslice;
        bm.brow0 = M.brow0;
        bm.brows = M.brows;
        bm.diag = M.diag;
        bm.gf = plan.lower.ngroups();
        bm.gb = plan.upper.ngroups();
        bm.n = n;
        bm.n_active = plan.n_active;
        bm.full_slices = plan.full.nslices;
        bm.full_row0 = plan.full.group_row0.empty() ? 0 : plan.full.group_row0[0];
        bm.spmv_lanes = spmv_table[M.bucket];
        bm.omega = T(omega);
        bm.c2 = T((2. - omega) / omega);
        M.bm = bm;
        if (id >= d_mats_cap) {  // grow the device table (no kernel is running)
            const int cap = std::max(16, 2 * d_mats_cap);
            BMat<T>* nm = dev_alloc<BMat<T>>(cap);
            if (d_mats) {
                SFA_HIP(hipMemcpy(nm, d_mats, sizeof(BMat<T>) * d_mats_cap, hipMemcpyDeviceToDevice));
                dev_free(d_mats);
            }
            d_mats = nm;
            d_mats_cap = cap;
        }
        SFA_HIP(hipMemcpy(d_mats + id, &bm, sizeof(BMat<T>), hipMemcpyHostToDevice));
        return id;
    }

    bool any_running()
    {
        running();
        for (int s = 0; s < max_slots; s++)
            if (slot_used[s] && slots[s].kind != PH_DONE) return true;
        return false;
    }

    // Installs lanes per row for each bucket in every matrix (the step kernel
    // reads them from the device table at each SpMV phase).  Only between
    // solves: a phase's unit count is fixed when the phase is scheduled.
    void set_spmv(const int* lanes)
    {
        for (int b = 0; b < kSpmvBuckets; b++)
            if (lanes[b] != 1 && lanes[b] != 2 && lanes[b] != 4 && lanes[b] != 8)
                throw Error{SFA_ERR_INVALID, "lanes per row must be 1, 2, 4 or 8"};
        SFA_HIP(hipSetDevice(device));
        if (any_running()) throw Error{SFA_ERR_INVALID, "solves are running; set the SpMV mapping between solves"};
        for (int b = 0; b < kSpmvBuckets; b++) spmv_table[b] = lanes[b];
        for (size_t id = 0; id < mats.size(); id++) {
            HostMat<T>& M = mats[id];
            if (!M.live) continue;
            M.bm.spmv_lanes = spmv_table[M.bucket];
            SFA_HIP(hipMemcpy(d_mats + id, &M.bm, sizeof(BMat<T>), hipMemcpyHostToDevice));
        }
    }

    // For each bucket that holds matrices: times the SpMV phase of all its
    // matrices under each mapping (k_batch_spmv_probe, the step kernel's SpMV
    // code and grid) and installs the fastest.
    void tune_spmv(sfa_spmv_tuning* out)
    {
        SFA_HIP(hipSetDevice(device));
        if (any_running()) throw Error{SFA_ERR_INVALID, "solves are running; tune between solves"};
        sfa_spmv_tuning t{};
        std::vector<int> chosen(spmv_table, spmv_table + kSpmvBuckets);
        std::vector<std::vector<int>> ids(kSpmvBuckets);
        int nmax = 1;
        long long umax = 1;
        for (size_t id = 0; id < mats.size(); id++) {
            const HostMat<T>& M = mats[id];
            if (!M.live) continue;
            ids[M.bucket].push_back((int)id);
            nmax = std::max(nmax, M.n + 1);
            for (int b = 0; b < kSpmvBuckets; b++) t.rows[b] += M.full.bucket_rows[b];
        }
        for (int b = 0; b < kSpmvBuckets; b++) {
            long long u = 0;
            for (int id : ids[b]) u += mats[id].full.nslices * 8LL + 1;
            umax = std::max(umax, u);
        }
        if (umax > INT_MAX / 2) throw Error{SFA_ERR_INVALID, "too many SpMV units to tune"};
        T* vd = dev_alloc<T>(nmax);
        T* vq = dev_alloc<T>(nmax);
        double* part = dev_alloc<double>(umax);
        BMat<T>* dm = nullptr;
        int* dstart = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        auto cleanup = [&] {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            for (void* v : {(void*)vd, (void*)vq, (void*)part, (void*)dm, (void*)dstart}) dev_free(v);
        };
        try {
            launch_fill<T>(stream, vd, T(1), nmax);
            SFA_HIP(hipEventCreate(&e0));
            SFA_HIP(hipEventCreate(&e1));
            const int cand[SFA_SPMV_MAPPINGS] = {1, 2, 4, 8};
            for (int b = 0; b < kSpmvBuckets; b++) {
                // at most kMaxSlots matrices per probe (its unit table lives in LDS)
                const int nm = std::min((int)ids[b].size(), kMaxSlots);
                if (nm == 0) continue;
                std::vector<BMat<T>> hm;
                for (int k = 0; k < nm; k++) hm.push_back(mats[ids[b][k]].bm);
                dev_free(dm);
                dev_free(dstart);
                dm = nullptr;
                dstart = nullptr;
                dm = dev_alloc<BMat<T>>(nm);
                dstart = dev_alloc<int>(nm + 1);
                SFA_HIP(hipMemcpy(dm, hm.data(), sizeof(BMat<T>) * nm, hipMemcpyHostToDevice));
                double best_us = 0.;
                for (int i = 0; i < SFA_SPMV_MAPPINGS; i++) {
                    const int K = cand[i];
                    std::vector<int> us(nm + 1, 0);
                    for (int m = 0; m < nm; m++)
                        us[m + 1] = us[m] + std::max(1, hm[m].full_slices * K);
                    SFA_HIP(hipMemcpy(dstart, us.data(), sizeof(int) * (nm + 1), hipMemcpyHostToDevice));
                    auto once = [&] {
                        switch (K) {
                        case 1: launch(k_batch_spmv_probe<T, 1>, dim3(kStepBlocks), dim3(kUnitBlock), stream, (const BMat<T>*)dm, (const int*)dstart, nm, (const T*)vd, vq, part); break;
                        case 2: launch(k_batch_spmv_probe<T, 2>, dim3(kStepBlocks), dim3(kUnitBlock), stream, (const BMat<T>*)dm, (const int*)dstart, nm, (const T*)vd, vq, part); break;
                        case 4: launch(k_batch_spmv_probe<T, 4>, dim3(kStepBlocks), dim3(kUnitBlock), stream, (const BMat<T>*)dm, (const int*)dstart, nm, (const T*)vd, vq, part); break;
                        default: launch(k_batch_spmv_probe<T, 8>, dim3(kStepBlocks), dim3(kUnitBlock), stream, (const BMat<T>*)dm, (const int*)dstart, nm, (const T*)vd, vq, part); break;
                        }
                    };
                    for (int w = 0; w < 3; w++) once();
                    constexpr int kReps = 10;
                    float ms_min = 0.f;
                    for (int rep = 0; rep < 3; rep++) {
                        SFA_HIP(hipEventRecord(e0, stream));
                        for (int k = 0; k < kReps; k++) once();
                        SFA_HIP(hipEventRecord(e1, stream));
                        SFA_HIP(hipEventSynchronize(e1));
                        float ms = 0.f;
                        SFA_HIP(hipEventElapsedTime(&ms, e0, e1));
                        if (rep == 0 || ms < ms_min) ms_min = ms;
                    }
                    SFA_HIP(hipGetLastError());
                    const double us_per = 1000.0 * ms_min / kReps;
                    t.us[b][i] = us_per;
                    if (i == 0 || us_per < best_us) {
                        best_us = us_per;
                        chosen[b] = K;
                    }
                }
            }
            SFA_HIP(hipStreamSynchronize(stream));
            cleanup();
        } catch (...) {
            (void)hipStreamSynchronize(stream);
            cleanup();
            throw;
        }
        set_spmv(chosen.data());
        for (int b = 0; b < kSpmvBuckets; b++) t.lanes_per_row[b] = t.rows[b] ? chosen[b] : 0;
        *out = t;
    }

    void remove_matrix(int id)
    {
        if (id < 0 || id >= (int)mats.size() || !mats[id].live)
            throw Error{SFA_ERR_INVALID, "no such matrix"};
        if (mats[id].slots_in_use > 0)
            throw Error{SFA_ERR_INVALID, "matrix still has slots; collect them first"};
        SFA_HIP(hipStreamSynchronize(stream));
        mats[id].release();
    }

    long long bytes_in_use() const
    {
        long long b = 0;
        for (const auto& m : mats)
            if (m.live) b += m.bytes;
        std::vector<const VecBlock<T>*> seen;
        for (const auto& v : vecs) {
            b += v.bytes;
            const VecBlock<T>* vb = v.block.get();
            if (vb && std::find(seen.begin(), seen.end(), vb) == seen.end()) {
                seen.push_back(vb);
                b += vb->bytes;
            }
        }
        return b;
    }

    int submit(int id, const double* bvec, const double* x0, double precision, double inner_rtol,
               int max_outer)
    {
        int s = -1;
        submit_group(id, 1, bvec, x0, precision, inner_rtol, max_outer, &s);
        return s;
    }

    // nrhs right-hand sides of matrix id (rhs k at bvec + k n, initial guess
    // at x0 + k n, or zero) in nrhs consecutive slots whose vectors are one
    // interleaved block; slot ids to out[0..nrhs).
    void submit_group(int id, int nrhs, const double* bvec, const double* x0, double precision,
                      double inner_rtol, int max_outer, int* out)
    {
        if (id < 0 || id >= (int)mats.size() || !mats[id].live)
            throw Error{SFA_ERR_INVALID, "no such matrix"};
        if (nrhs < 1 || nrhs > kRiders)
            throw Error{SFA_ERR_INVALID, "a group holds 1 to 4 right-hand sides"};
        if (!bvec) throw Error{SFA_ERR_INVALID, "null right-hand side"};
        if (!(precision > 0)) throw Error{SFA_ERR_INVALID, "precision must be positive"};
        if (mixed && !(inner_rtol > 0 && inner_rtol < 1))
            throw Error{SFA_ERR_INVALID, "inner_rtol must be in (0, 1)"};
        HostMat<T>& M = mats[id];
        const int n = M.n;
        if ((long long)(n + 1) * nrhs > INT_MAX)
            throw Error{SFA_ERR_INVALID, "group vectors exceed 2^31 entries"};
        int s0 = -1;  // first of nrhs consecutive free slots
        for (int i = 0, run = 0; i < max_slots; i++) {
            run = slot_used[i] ? 0 : run + 1;
            if (run == nrhs) {
                s0 = i - nrhs + 1;
                break;
            }
        }
        if (s0 < 0) throw Error{SFA_ERR_INVALID, "no free slot"};
        SFA_HIP(hipSetDevice(device));
        const size_t len = (size_t)(n + 1) * nrhs;
        auto blk = std::make_shared<VecBlock<T>>();
        blk->lanes = nrhs;
        T** tv[] = {&blk->x, &blk->r, &blk->b, &blk->d, &blk->z, &blk->q, &blk->m};
        for (T** p : tv) {
            *p = dev_alloc<T>(len);
            SFA_HIP(hipMemsetAsync(*p, 0, len * sizeof(T), stream));
        }
        blk->bytes = 7LL * (long long)len * sizeof(T);
        if (mixed) {
            for (double** p : {&blk->x64, &blk->b64, &blk->r64}) {
                *p = dev_alloc<double>(len);
                SFA_HIP(hipMemsetAsync(*p, 0, len * sizeof(double), stream));
            }
            blk->bytes += 3LL * (long long)len * 8;
        }
        // inputs in device order, lanes interleaved (row n, the padding row,
        // stays zero); each tolerance as the reference computes it
        std::vector<double> bI((size_t)n * nrhs), xI((size_t)n * nrhs, 0.0);
        std::vector<double> tol(nrhs);
        for (int k = 0; k < nrhs; k++) {
            const double* bk = bvec + (size_t)k * n;
            double bb = 0.;
            for (int i = 0; i < n; i++) bb += bk[i] * bk[i];
            for (int p = 0; p < n; p++) bI[(size_t)p * nrhs + k] = bk[M.perm[p]];
            if (x0) {
                const double* xk = x0 + (size_t)k * n;
                for (int p = 0; p < n; p++) xI[(size_t)p * nrhs + k] = xk[M.perm[p]];
            }
            double t = precision * std::sqrt(bb);
            const double floor_ = 1.4901161193847656e-08;
            if (!(t >= floor_) && t <= floor_) t = floor_;
            tol[k] = t;
        }
        const size_t m = (size_t)n * nrhs;
        std::vector<T> bt, xt;
        if (mixed) {
            SFA_HIP(hipMemcpyAsync(blk->b64, bI.data(), m * sizeof(double), hipMemcpyHostToDevice, stream));
            SFA_HIP(hipMemcpyAsync(blk->x64, xI.data(), m * sizeof(double), hipMemcpyHostToDevice, stream));
        } else {
            bt.assign(bI.begin(), bI.end());
            xt.assign(xI.begin(), xI.end());
            SFA_HIP(hipMemcpyAsync(blk->b, bt.data(), m * sizeof(T), hipMemcpyHostToDevice, stream));
            SFA_HIP(hipMemcpyAsync(blk->x, xt.data(), m * sizeof(T), hipMemcpyHostToDevice, stream));
        }
        std::vector<BSlot<T>> sls(nrhs);
        for (int k = 0; k < nrhs; k++) {
            const int s = s0 + k;
            SlotVectors<T>& v = vecs[s];
            v.release();
            v.block = blk;
            v.lane = k;
            v.partials = dev_alloc<double>(M.max_units);
            v.grp = dev_alloc<double>(std::max(1, M.gb));
            v.bytes = (long long)M.max_units * 8 + (long long)std::max(1, M.gb) * 8;
            BSlot<T> sl{};
            sl.mat = id;
            sl.mixed = mixed ? 1 : 0;
            sl.cont = cont ? 1 : 0;
            sl.maxK = n;
            sl.exit_reason = SFA_EXIT_CAP;
            sl.max_outer = max_outer;
            sl.ls = nrhs;
            sl.lane = k;
            sl.vgroup = s0;
            sl.x = blk->x + k; sl.r = blk->r + k; sl.b = blk->b + k; sl.d = blk->d + k;
            sl.z = blk->z + k; sl.q = blk->q + k; sl.m = blk->m + k;
            sl.x64 = mixed ? blk->x64 + k : nullptr;
            sl.b64 = mixed ? blk->b64 + k : nullptr;
            sl.r64 = mixed ? blk->r64 + k : nullptr;
            sl.partials = v.partials;
            sl.grp = v.grp;
            if (mixed) {
                sl.kind = PH_RES64;
                sl.target64 = tol[k];
                sl.rel2 = inner_rtol * inner_rtol;
            } else {
                sl.kind = PH_RESID;
                sl.tol2 = tol[k] * tol[k];
            }
            sls[k] = sl;
        }
        SFA_HIP(hipMemcpyAsync(d_slots + s0, sls.data(), sizeof(BSlot<T>) * nrhs,
                               hipMemcpyHostToDevice, stream));
        SFA_HIP(hipStreamSynchronize(stream));  // host vectors go out of scope
        for (int k = 0; k < nrhs; k++) {
            slots[s0 + k] = sls[k];
            slot_used[s0 + k] = 1;
            out[k] = s0 + k;
        }
        M.slots_in_use += nrhs;
    }

    void read_sched()
    {
        SFA_HIP(hipMemcpyAsync(h_sched, d_sched, sizeof(BSched), hipMemcpyDeviceToHost, stream));
        SFA_HIP(hipStreamSynchronize(stream));
        step = h_sched->step;
    }

    // Whether some bucket runs its SpMV with more than one lane per row (the
    // step kernel then needs those code paths).
    bool wide() const
    {
        for (int b = 0; b < kSpmvBuckets; b++)
            if (spmv_table[b] != 1) return true;
        return false;
    }

    // The two step kernels: complete interleaved groups, then every other unit.
    void launch_step(const BArgs<T>& a, bool w)
    {
        launch(k_batch_step<T, false, true>, dim3(kStepBlocks), dim3(kUnitBlock), stream, a);
        if (w) launch(k_batch_step<T, true, false>, dim3(kStepBlocks), dim3(kUnitBlock), stream, a);
        else launch(k_batch_step<T, false, false>, dim3(kStepBlocks), dim3(kUnitBlock), stream, a);
    }

    // Records kGraphSteps steps (k_batch_step twice, k_batch_fold, k_batch_prefix)
    // into a graph, node by node; rebuilt when the device matrix table has
    // moved or the step kernel variant changed.
    void build_graph()
    {
        const bool w = wide();
        if (graph && graph_mats == d_mats && graph_wide == w) return;
        if (graph) SFA_HIP(hipGraphExecDestroy(graph));
        graph = nullptr;
        graph_rec.reset();
        const BArgs<T> a = args();
        auto rec = std::make_unique<GraphRecorder>();
        for (int i = 0; i < kGraphSteps; i++) {
            launch_step(a, w);
            launch(k_batch_fold<T>, dim3(max_slots), dim3(kUnitBlock), stream, a);
            launch(k_batch_prefix<T>, dim3(1), dim3(kPrefixBlock), stream, a);
        }
        hipGraphExec_t exec = nullptr;
        const hipError_t e = rec->instantiate(&exec);
        graph_rec = std::move(rec);
        if (e != hipSuccess)
            throw Error{SFA_ERR_HIP, std::string("batch step graph: ") + hipGetErrorString(e)};
        graph = exec;
        graph_mats = d_mats;
        graph_wide = w;
    }

    // Runs about max_steps steps (whole graph replays, at least one; 0: until
    // no solve is running).  Returns the number of solves still running.
    int run(int max_steps)
    {
        SFA_HIP(hipSetDevice(device));
        if (!d_mats) return running();
        build_graph();
        if (max_steps > 0) {
            const int reps = std::max(1, max_steps / kGraphSteps);
            for (int i = 0; i < reps; i++) SFA_HIP(hipGraphLaunch(graph, stream));
        } else {
            // the first replay also schedules slots admitted since the last one
            do {
                SFA_HIP(hipGraphLaunch(graph, stream));
                read_sched();
            } while (h_sched->active > 0);
        }
        read_sched();
        return running();
    }

    // Runs `steps` steps with the kernels launched one by one, each
    // between two HIP events; adds each kernel's summed time (milliseconds)
    // to ms[0..3) (both step launches, fold, prefix).
    void profile(int steps, double* ms)
    {
        SFA_HIP(hipSetDevice(device));
        if (!d_mats || steps <= 0) return;
        const BArgs<T> a = args();
        std::vector<hipEvent_t> ev(4 * steps, nullptr);
        auto cleanup = [&] {
            for (hipEvent_t e : ev)
                if (e) (void)hipEventDestroy(e);
        };
        try {
            for (auto& e : ev) SFA_HIP(hipEventCreate(&e));
            for (int i = 0; i < steps; i++) {
                SFA_HIP(hipEventRecord(ev[4 * i], stream));
                launch_step(a, wide());
                SFA_HIP(hipEventRecord(ev[4 * i + 1], stream));
                launch(k_batch_fold<T>, dim3(max_slots), dim3(kUnitBlock), stream, a);
                SFA_HIP(hipEventRecord(ev[4 * i + 2], stream));
                launch(k_batch_prefix<T>, dim3(1), dim3(kPrefixBlock), stream, a);
                SFA_HIP(hipEventRecord(ev[4 * i + 3], stream));
            }
            SFA_HIP(hipGetLastError());
            SFA_HIP(hipEventSynchronize(ev.back()));
            for (int i = 0; i < steps; i++)
                for (int k = 0; k < 3; k++) {
                    float t = 0.f;
                    SFA_HIP(hipEventElapsedTime(&t, ev[4 * i + k], ev[4 * i + k + 1]));
                    ms[k] += t;
                }
            cleanup();
        } catch (...) {
            (void)hipStreamSynchronize(stream);
            cleanup();
            throw;
        }
        read_sched();
    }

    int running()
    {
        SFA_HIP(hipMemcpyAsync(slots.data(), d_slots, sizeof(BSlot<T>) * max_slots,
                               hipMemcpyDeviceToHost, stream));
        SFA_HIP(hipStreamSynchronize(stream));
        int n = 0;
        for (int s = 0; s < max_slots; s++)
            if (slot_used[s] && slots[s].kind != PH_DONE) n++;
        return n;
    }

    int finished(int* out, int cap)
    {
        running();
        int n = 0;
        for (int s = 0; s < max_slots && n < cap; s++)
            if (slot_used[s] && slots[s].kind == PH_DONE) out[n++] = s;
        return n;
    }

    void collect(int s, double* x, sfa_batch_info* info)
    {
        if (s < 0 || s >= max_slots || !slot_used[s]) throw Error{SFA_ERR_INVALID, "no such slot"};
        SFA_HIP(hipMemcpyAsync(&slots[s], d_slots + s, sizeof(BSlot<T>), hipMemcpyDeviceToHost, stream));
        SFA_HIP(hipStreamSynchronize(stream));
        const BSlot<T>& sl = slots[s];
        if (sl.kind != PH_DONE) throw Error{SFA_ERR_INVALID, "slot has not finished"};
        HostMat<T>& M = mats[sl.mat];
        const int n = M.n;
        if (x) {  // the slot's lane of its group's block
            const VecBlock<T>& blk = *vecs[s].block;
            const int L = blk.lanes, k = vecs[s].lane;
            const size_t m = (size_t)n * L;
            std::vector<double> xd(n);
            if (mixed) {
                std::vector<double> xa(m);
                SFA_HIP(hipMemcpy(xa.data(), blk.x64, m * sizeof(double), hipMemcpyDeviceToHost));
                for (int i = 0; i < n; i++) xd[i] = xa[(size_t)i * L + k];
            } else {
                std::vector<T> xa(m);
                SFA_HIP(hipMemcpy(xa.data(), blk.x, m * sizeof(T), hipMemcpyDeviceToHost));
                for (int i = 0; i < n; i++) xd[i] = (double)xa[(size_t)i * L + k];
            }
            for (int p = 0; p < n; p++) x[M.perm[p]] = xd[p];
        }
        if (info) {
            sfa_batch_info o{};
            o.iterations = mixed ? sl.inner_total : sl.K;
            o.outer = sl.outer;
            o.exit_reason = sl.exit_reason;
            o.zr = sl.zr;
            o.residual_norm = mixed ? sl.res64 : std::sqrt(sl.zr);
            o.success = mixed ? (sl.res64 <= sl.target64 ? 1 : 0)
                              : (std::sqrt(sl.zr) < std::sqrt(sl.tol2) ? 1 : 0);
            *info = o;
        }
        BSlot<T> cleared{};
        cleared.mat = -1;
        cleared.kind = PH_DONE;
        SFA_HIP(hipMemcpy(d_slots + s, &cleared, sizeof(BSlot<T>), hipMemcpyHostToDevice));
        slots[s] = cleared;
        slot_used[s] = 0;
        vecs[s].release();
        M.slots_in_use--;
    }
};

}  // namespace

struct BatchHandle {
    std::unique_ptr<Engine<double>> e64;
    std::unique_ptr<Engine<float>> e32;
    template <class Fn>
    void visit(Fn&& fn)
    {
        if (e64) fn(*e64);
        else fn(*e32);
    }
};

}  // namespace sfa

struct sfa_batch_s {
    sfa::BatchHandle h;
};

struct sfa_prepared_s {
    sfa::Plan plan;
};

extern "C" {

int sfa_batch_create(int precision_mode, int max_slots, sfa_batch* out)
{
    return sfa::guarded([&] {
        if (!out) throw sfa::Error{SFA_ERR_INVALID, "null output handle"};
        *out = nullptr;
        std::unique_ptr<sfa_batch_s> b(new sfa_batch_s());
        if (precision_mode == SFA_BATCH_FP64) {
            b->h.e64.reset(new sfa::Engine<double>());
            b->h.e64->create(0, max_slots);
        } else if (precision_mode == SFA_BATCH_MIXED || precision_mode == SFA_BATCH_MIXED_RESTARTS) {
            b->h.e32.reset(new sfa::Engine<float>());
            b->h.e32->create(precision_mode == SFA_BATCH_MIXED ? 1 : 2, max_slots);
        } else {
            throw sfa::Error{SFA_ERR_INVALID,
                             "precision_mode must be SFA_BATCH_FP64, SFA_BATCH_MIXED or SFA_BATCH_MIXED_RESTARTS"};
        }
        *out = b.release();
    });
}

int sfa_batch_add_matrix(sfa_batch b, int N, const int* csrrows, const int* csrcolumns,
                         const double* csrvalues, const double* diagonal, double omega,
                         int num_colors, const int* F, const int* B, const int* R, int* matrix_id)
{
    return sfa::guarded([&] {
        if (!b || !matrix_id) throw sfa::Error{SFA_ERR_INVALID, "null argument"};
        b->h.visit([&](auto& e) {
            *matrix_id = e.add_matrix(N, csrrows, csrcolumns, csrvalues, diagonal, omega,
                                      num_colors, F, B, R);
        });
    });
}

int sfa_batch_prepare_matrix(int N, const int* csrrows, const int* csrcolumns,
                             const double* csrvalues, const double* diagonal, double omega,
                             int num_colors, const int* F, const int* B, const int* R,
                             sfa_prepared* out)
{
    return sfa::guarded([&] {
        if (!out) throw sfa::Error{SFA_ERR_INVALID, "null output handle"};
        *out = nullptr;
        std::unique_ptr<sfa_prepared_s> p(new sfa_prepared_s());
        bool colouring_error = false;
        std::string err = sfa::build_plan(p->plan, SFA_METHOD_RAINBOW, N, csrrows, csrcolumns,
                                          csrvalues, diagonal, omega, num_colors, F, B, R,
                                          &colouring_error);
        if (!err.empty())
            throw sfa::Error{colouring_error ? SFA_ERR_COLOURING : SFA_ERR_INVALID, err};
        *out = p.release();
    });
}

int sfa_batch_add_prepared(sfa_batch b, sfa_prepared prepared, int* matrix_id)
{
    return sfa::guarded([&] {
        if (!b || !prepared || !matrix_id) throw sfa::Error{SFA_ERR_INVALID, "null argument"};
        b->h.visit([&](auto& e) { *matrix_id = e.add_plan(prepared->plan); });
    });
}

int sfa_prepared_destroy(sfa_prepared prepared)
{
    return sfa::guarded([&] { delete prepared; });
}

int sfa_batch_remove_matrix(sfa_batch b, int matrix_id)
{
    return sfa::guarded([&] {
        if (!b) throw sfa::Error{SFA_ERR_INVALID, "null batch"};
        b->h.visit([&](auto& e) { e.remove_matrix(matrix_id); });
    });
}

int sfa_batch_submit(sfa_batch b, int matrix_id, const double* rhs, const double* x0,
                     double precision, double inner_rtol, int max_outer, int* slot)
{
    return sfa::guarded([&] {
        if (!b || !slot) throw sfa::Error{SFA_ERR_INVALID, "null argument"};
        b->h.visit([&](auto& e) {
            *slot = e.submit(matrix_id, rhs, x0, precision, inner_rtol, max_outer);
        });
    });
}

int sfa_batch_submit_group(sfa_batch b, int matrix_id, int nrhs, const double* rhs,
                           const double* x0, double precision, double inner_rtol,
                           int max_outer, int* slots)
{
    return sfa::guarded([&] {
        if (!b || !slots) throw sfa::Error{SFA_ERR_INVALID, "null argument"};
        b->h.visit([&](auto& e) {
            e.submit_group(matrix_id, nrhs, rhs, x0, precision, inner_rtol, max_outer, slots);
        });
    });
}

int sfa_batch_run(sfa_batch b, int max_steps, int* running)
{
    return sfa::guarded([&] {
        if (!b) throw sfa::Error{SFA_ERR_INVALID, "null batch"};
        b->h.visit([&](auto& e) {
            const int r = e.run(max_steps);
            if (running) *running = r;
        });
    });
}

int sfa_batch_finished(sfa_batch b, int* slots, int capacity, int* count)
{
    return sfa::guarded([&] {
        if (!b || !count || (capacity > 0 && !slots)) throw sfa::Error{SFA_ERR_INVALID, "null argument"};
        b->h.visit([&](auto& e) { *count = e.finished(slots, capacity); });
    });
}

int sfa_batch_collect(sfa_batch b, int slot, double* x, sfa_batch_info* info)
{
    return sfa::guarded([&] {
        if (!b) throw sfa::Error{SFA_ERR_INVALID, "null batch"};
        b->h.visit([&](auto& e) { e.collect(slot, x, info); });
    });
}

int sfa_batch_memory(sfa_batch b, long long* bytes)
{
    return sfa::guarded([&] {
        if (!b || !bytes) throw sfa::Error{SFA_ERR_INVALID, "null argument"};
        b->h.visit([&](auto& e) { *bytes = e.bytes_in_use(); });
    });
}

int sfa_batch_profile(sfa_batch b, int steps, double* ms)
{
    return sfa::guarded([&] {
        if (!b || !ms) throw sfa::Error{SFA_ERR_INVALID, "null argument"};
        if (steps < 0 || steps > 100000) throw sfa::Error{SFA_ERR_INVALID, "steps must be in 0..100000"};
        b->h.visit([&](auto& e) { e.profile(steps, ms); });
    });
}

int sfa_batch_steps(sfa_batch b, long long* steps)
{
    return sfa::guarded([&] {
        if (!b || !steps) throw sfa::Error{SFA_ERR_INVALID, "null argument"};
        b->h.visit([&](auto& e) {
            e.read_sched();
            *steps = (long long)e.h_sched->step;
        });
    });
}

int sfa_batch_tune_spmv(sfa_batch b, sfa_spmv_tuning* out)
{
    return sfa::guarded([&] {
        if (!b || !out) throw sfa::Error{SFA_ERR_INVALID, "null argument"};
        b->h.visit([&](auto& e) { e.tune_spmv(out); });
    });
}

int sfa_batch_set_spmv(sfa_batch b, const int* lanes_per_row)
{
    return sfa::guarded([&] {
        if (!b || !lanes_per_row) throw sfa::Error{SFA_ERR_INVALID, "null argument"};
        b->h.visit([&](auto& e) { e.set_spmv(lanes_per_row); });
    });
}

int sfa_batch_destroy(sfa_batch b)
{
    return sfa::guarded([&] { delete b; });
}

}  // extern "C"

