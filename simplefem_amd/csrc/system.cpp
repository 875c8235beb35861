// system.cpp -- device-resident CG systems and the C-ABI of libsimplefem_amd.
//
// A system owns, in HBM: the SELL-64 matrix parts of its plan, the device
// order permutation, the diagonal, the Krylov vectors (n+1 entries each, the
// last a constant 0), the SolveState and the reduction scratch.  A solve runs
//   gather inputs -> state/tolerance -> restart kernels -> chunks of
//   iterations replayed from a captured hipGraph until state.done -> scatter
// and reads the state back once per chunk.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/simplefem_amd.h"
#include "host_common.hpp"
#include "kernels.hpp"
#include "launch.hpp"
#include "plan.hpp"

namespace sfa {

std::string& last_error()
{
    thread_local std::string msg;
    return msg;
}

namespace {

const char* method_title(int method)
{
    switch (method) {
    case SFA_METHOD_CG: return "Conjugate Gradient CSR";
    case SFA_METHOD_DIAGONAL: return "Conjugate Gradient CSR (Diag)";
    case SFA_METHOD_SSOR: return "Conjugate Gradient CSR (SSOR)";
    case SFA_METHOD_SSOR_EISENSTAT: return "Conjugate Gradient CSR (SSOR-Eisenstat)";
    case SFA_METHOD_RAINBOW: return "Conjugate Gradient CSR (SSOR-Rainbow)";
    default: return "Conjugate Gradient CSR (SSOR-Rainbow-Eisenstat)";
    }
}

}  // namespace

template <typename T>
struct SystemT {
    int device = 0;
    int method = 0;
    int n = 0;
    int n_active = 0;
    long long nnz = 0;
    double omega = 1.0;
    bool eisen = false;
    bool sweeps = false;
    bool rainbow = false;
    bool fuse_first_backward = false;  // see forward_sweep
    bool merge_first_forward = false;  // see forward_sweep
    // The iteration's sweeps (SSOR / Rainbow): one launch per entry, a group
    // or, with pair, the group and the next one (k_forward_pair /
    // k_backward_pair).  defer_bwd[g]: backward group g makes its rows' x and
    // r updates (their forward ran as the first group of a pair, or merged).
    // spmv_tail: the groups of the last backward launch, whose z.r partials
    // the next SpMV sums itself.
    struct SweepLaunch {
        int g;
        bool pair;
    };
    bool pair_sweeps = false;  // true: pairs of groups where possible (SFA_FLAG_PAIRED_SWEEPS)
    // device-pointer solves: sum ||b|| for the tolerance in row order, one
    // lane (the reference's bits; for short vectors) instead of by blocks
    bool row_order_tol = false;
    // mixed solves: restart every inner CG from d = z (SFA_FLAG_INNER_RESTARTS)
    // instead of continuing the previous inner solve's direction
    bool inner_restarts = false;
    std::vector<SweepLaunch> fwd_sched, bwd_sched;
    std::vector<int> defer_bwd;
    std::vector<int> spmv_tail;
    int spmv_lanes = 1;                // lanes per row of the iteration's SpMV (1, 2, 4, 8)

    DevPartT<T> full, lower, upper, lower_restart;
    bool has_lower_restart = false;
    DevPartT<T> apply_part;  // all rows, identity order (sfa_system_apply), built lazily
    bool has_apply_part = false;
    std::vector<int> host_A, host_C;  // kept only to build apply_part later
    std::vector<double> host_V;
    // the CSR arrays apply_part is built from: the copies above, or the
    // caller's arrays (SFA_FLAG_BORROW_CSR); null for an FP32 copy
    const int* csr_A = nullptr;
    const int* csr_C = nullptr;
    const double* csr_V = nullptr;
    // create(): host plan, host CSR copy, FP64 device set-up, FP32 copy
    double setup_ms[4] = {0., 0., 0., 0.};

    int* perm = nullptr;
    bool shares_perm = false;  // perm belongs to the FP64 system (FP32 copy)
    T* diag = nullptr;
    // d and d1: the search direction, double-buffered by iteration parity
    T *x = nullptr, *r = nullptr, *b = nullptr, *d = nullptr, *d1 = nullptr, *z = nullptr,
      *q = nullptr, *m = nullptr;
    T* io = nullptr;  // n staging entries for host <-> device copies
    T* io2 = nullptr;
    T* io3 = nullptr;
    T* dkeep = nullptr;  // the direction an inner solve continues (mixed precision), lazily
    SolveState* state = nullptr;
    double* hist = nullptr;
    // block partials of the deferred reductions: SpMV / combine / tolerance
    // (ps), update / restart (pu), backward groups by slice (pb); per-group z.r
    double* ps = nullptr;
    double* pu = nullptr;
    double* pb = nullptr;
    double* grp = nullptr;
    SolveState* host_state = nullptr;  // pinned

    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t null_done = nullptr;  // see order_after_null_stream
    hipGraphExec_t chunk_exec = nullptr;
    std::unique_ptr<GraphRecorder> chunk_graph;  // the recorded graph and its arguments, kept with the exec
    int chunk_iters = 0;
    Profiler* prof = nullptr;  // set only while sfa_system_profile runs
    // FP32 copy of the plan for the mixed-precision solve (T = double only)
    std::unique_ptr<SystemT<float>> inner;

    // Launches fn(); in a profiled solve, between two events on stream s.
    template <class Fn>
    void run(hipStream_t s, int klass, double bytes, Fn&& fn)
    {
        if (!prof) {
            fn();
            return;
        }
        const size_t i = prof->klass.size();
        SFA_HIP(hipEventRecord(prof->event(2 * i), s));
        fn();
        SFA_HIP(hipEventRecord(prof->event(2 * i + 1), s));
        prof->klass.push_back(klass);
        prof->bytes.push_back(bytes);
    }

    ~SystemT() { release(); }

    void release()
    {
        // The buffers go back to the block cache, which does not wait for the
        // device: nothing queued on the system's stream may still use them.
        if (stream) (void)hipStreamSynchronize(stream);
        inner.reset();
        if (chunk_exec) (void)hipGraphExecDestroy(chunk_exec);
        chunk_exec = nullptr;
        chunk_graph.reset();
        full.release();
        lower.release();
        upper.release();
        lower_restart.release();
        apply_part.release();
        if (!shares_perm) dev_free(perm);
        shares_perm = false;
        for (void* p : {(void*)diag, (void*)x, (void*)r, (void*)b, (void*)d, (void*)d1,
                        (void*)z, (void*)q, (void*)m, (void*)io, (void*)io2, (void*)io3,
                        (void*)dkeep, (void*)state, (void*)hist, (void*)ps, (void*)pu, (void*)pb, (void*)grp})
            dev_free(p);
        perm = nullptr;
        diag = x = r = b = d = d1 = z = q = m = io = io2 = io3 = dkeep = nullptr;
        hist = ps = pu = pb = grp = nullptr;
        state = nullptr;
        pinned_free(host_state);
        host_state = nullptr;
        if (own_stream) (void)hipStreamDestroy(own_stream);
        own_stream = stream = nullptr;
        if (null_done) (void)hipEventDestroy(null_done);
        null_done = nullptr;
    }

    // Device-pointer inputs may have been written by work queued on the null
    // stream (hipMemset or hipMemcpy there, a framework's default stream),
    // which the system's own stream, created non-blocking, does not wait for:
    // it waits for everything queued there so far.  A stream the caller set
    // (sfa_system_set_stream) is ordered by the caller.
    void order_after_null_stream(hipStream_t s)
    {
        if (s != own_stream) return;
        if (!null_done) SFA_HIP(hipEventCreateWithFlags(&null_done, hipEventDisableTiming));
        SFA_HIP(hipEventRecord(null_done, nullptr));
        SFA_HIP(hipStreamWaitEvent(s, null_done, 0));
    }

    // borrow: keep A, C, V as the caller's pointers (which stay valid and
    // unchanged until the system is destroyed) instead of a copy.
    void create(int method_, int n_, const int* A, const int* C, const double* V, const double* aux,
                double omega_, int nc, const int* F, const int* B, const int* R,
                bool with_fp32 = false, bool borrow = false)
    {
        // Arguments are checked on the host before the device is touched.
        using clock = std::chrono::steady_clock;
        auto ms_since = [](clock::time_point t0) {
            return std::chrono::duration<double, std::milli>(clock::now() - t0).count();
        };
        auto t0 = clock::now();
        Plan plan;
        bool colouring_error = false;
        std::string err = build_plan(plan, method_, n_, A, C, V, aux, omega_, nc, F, B, R,
                                     &colouring_error);
        if (!err.empty()) throw Error{colouring_error ? SFA_ERR_COLOURING : SFA_ERR_INVALID, err};
        setup_ms[0] = ms_since(t0);
        t0 = clock::now();
        if (borrow) {
            csr_A = A;
            csr_C = C;
            csr_V = V;
        } else {
            host_A.assign(A, A + n_ + 1);
            host_C.assign(C, C + A[n_]);
            host_V.assign(V, V + A[n_]);
            csr_A = host_A.data();
            csr_C = host_C.data();
            csr_V = host_V.data();
        }
        setup_ms[1] = ms_since(t0);
        t0 = clock::now();
        create_from_plan(plan, A, C, V);
        setup_ms[2] = ms_since(t0);
        if (with_fp32) {
            t0 = clock::now();
            inner.reset(new SystemT<float>());
            inner->pair_sweeps = pair_sweeps;
            inner->create_from_plan(plan, A, C, V, this);
            inner->stream = stream;
            setup_ms[3] = ms_since(t0);
        }
    }

    // share: for the FP32 copy of an FP64 system, that system.  The copy then
    // takes its SELL offsets, row ids, columns and permutation (which the
    // FP64 system owns and outlives the copy with), converts its values on
    // the device instead of uploading them, and keeps no host CSR (it is
    // never applied on its own).
    void create_from_plan(const Plan& plan, const int* A, const int* C, const double* V,
                          const SystemT<double>* share = nullptr)
    {
        const int method_ = plan.method, n_ = plan.n;
        const double omega_ = plan.omega;
        ensure_device();
        SFA_HIP(hipGetDevice(&device));
        method = method_;
        n = n_;
        n_active = plan.n_active;
        nnz = plan.nnz;
        omega = omega_;
        rainbow = method == SFA_METHOD_RAINBOW || method == SFA_METHOD_EISENSTAT_RAINBOW;
        eisen = method == SFA_METHOD_SSOR_EISENSTAT || method == SFA_METHOD_EISENSTAT_RAINBOW;
        sweeps = method >= SFA_METHOD_SSOR;
        SFA_HIP(hipStreamCreateWithFlags(&own_stream, hipStreamNonBlocking));
        stream = own_stream;

        auto part = [&](DevPartT<T>& dst, const SellPart& src, const DevPartT<double>* from) {
            if constexpr (std::is_same<T, float>::value) {
                if (from) {
                    dst.share_from(*from, own_stream);
                    return;
                }
            }
            dst.upload(src);
        };
        part(full, plan.full, share ? &share->full : nullptr);
        if (sweeps) {
            part(lower, plan.lower, share ? &share->lower : nullptr);
            part(upper, plan.upper, share ? &share->upper : nullptr);
        }
        has_lower_restart = plan.has_lower_restart;
        if (sweeps && !eisen && lower.ngroups() > 0 && upper.ngroups() > 0) {
            const int gl = lower.ngroups() - 1;
            fuse_first_backward = lower.row0(gl) >= 0 && upper.row0(0) == lower.row0(gl) &&
                                  upper.nrows_of(0) == lower.nrows_of(gl) &&
                                  upper.group_count(0) == lower.group_count(gl) &&
                                  upper.group_entries[0] == 0;
            // The first forward group (no lower entries) runs inside the second
            // one's kernel; its x, r updates move to its own backward kernel,
            // the last one, which must cover the same rows.
            const int gu = upper.ngroups() - 1;
            merge_first_forward = rainbow && lower.ngroups() >= 3 && lower.group_entries[0] == 0 &&
                                  lower.row0(0) >= 0 && upper.row0(gu) == lower.row0(0) &&
                                  upper.nrows_of(gu) == lower.nrows_of(0) &&
                                  upper.group_count(gu) == lower.group_count(0);
        }
        if (sweeps && !eisen) build_sweep_schedule(plan);
        if (has_lower_restart)
            part(lower_restart, plan.lower_restart, share ? &share->lower_restart : nullptr);

        if (share) {
            perm = share->perm;
            shares_perm = true;
            SFA_HIP(hipStreamSynchronize(own_stream));  // the converted values
        } else {
            perm = dev_alloc<int>(n);
            SFA_HIP(hipMemcpy(perm, plan.perm.data(), n * sizeof(int), hipMemcpyHostToDevice));
        }
        diag = dev_alloc<T>(n + 1);
        std::vector<T> dg(plan.diag.begin(), plan.diag.end());
        dg.push_back(T(0));
        SFA_HIP(hipMemcpy(diag, dg.data(), (n + 1) * sizeof(T), hipMemcpyHostToDevice));

        T** vecs[] = {&x, &r, &b, &d, &d1, &z, &q, &m};
        for (T** v : vecs) {
            *v = dev_alloc<T>(n + 1);
            SFA_HIP(hipMemset(*v, 0, (n + 1) * sizeof(T)));
        }
        io = dev_alloc<T>(n);
        io2 = dev_alloc<T>(n);
        io3 = dev_alloc<T>(n);
        state = dev_alloc<SolveState>(1);
        hist = dev_alloc<double>((size_t)n + 1);
        ps = dev_alloc<double>(std::max(kSpmvMaxBlocks, kVecReduceBlocks));
        pu = dev_alloc<double>(std::max({full.nslices, kVecReduceBlocks, 1}));
        pb = dev_alloc<double>(std::max(sweeps ? upper.nslices : 0, 1));
        int ngr = 1;
        if (sweeps) ngr = std::max({ngr, lower.ngroups(), upper.ngroups()});
        grp = dev_alloc<double>(ngr);
        SolveState init{};
        init.hist = hist;
        init.maxK = n;
        SFA_HIP(hipMemcpy(state, &init, sizeof(SolveState), hipMemcpyHostToDevice));
        host_state = pinned_alloc<SolveState>(1);
    }

    // ---------------------------------------------------------------- phases

    T c2() const { return T((2. - omega) / omega); }
    T c3() const { return T(2. / omega - 1.); }
    T w() const { return T(omega); }

    // Whether groups a and a+1 of part P run in one pair launch.  Possible
    // when a's rows are consecutive, a+1's slices hold at most kPairWidth
    // entries per lane and every row of a+1 reads at most kPairRefs rows of
    // a; worth it when the rows of a that a wavefront of a+1 forms again are
    // cheap: (most rows of a read by one row) x (entries per row of a) at
    // most kPairCost.  Measured on the 64^3 Kuhn Poisson matrix, pairs up to
    // 12 ran faster than their two launches, pairs from 24 up slower (the
    // rows of a are gathered one per lane, uncoalesced).
    static constexpr int kPairCost = 16;
    static bool pairable(const SellPart& P, int a)
    {
        if (a + 1 >= P.ngroups() || P.group_row0[a] < 0) return false;
        const unsigned r0 = (unsigned)P.group_row0[a], nr = (unsigned)P.group_rows[a];
        int refs = 0;
        for (int sl = P.group_slice[a + 1]; sl < P.group_slice[a + 2]; sl++) {
            const size_t off = (size_t)P.slice_off[sl];
            const int w = (P.slice_off[sl + 1] - P.slice_off[sl]) / kSliceRows;
            if (w > kPairWidth) return false;
            for (int l = 0; l < kSliceRows; l++) {
                int cnt = 0;
                for (int k = 0; k < w; k++)
                    cnt += (unsigned)P.cols[off + (size_t)k * kSliceRows + l] - r0 < nr ? 1 : 0;
                if (cnt > kPairRefs) return false;
                refs = std::max(refs, cnt);
            }
        }
        int wa = 0;  // entries per lane of a's slices
        for (int sl = P.group_slice[a]; sl < P.group_slice[a + 1]; sl++)
            wa = std::max(wa, (P.slice_off[sl + 1] - P.slice_off[sl]) / kSliceRows);
        return refs * std::max(wa, 1) <= kPairCost;
    }

    // Backward group with the rows of forward group lg, or -1.
    int upper_of(int lg) const
    {
        for (int ug = 0; ug < upper.ngroups(); ug++)
            if (lower.row0(lg) >= 0 && upper.row0(ug) == lower.row0(lg) &&
                upper.nrows_of(ug) == lower.nrows_of(lg))
                return ug;
        return -1;
    }

    void build_sweep_schedule(const Plan& plan)
    {
        const int GL = lower.ngroups(), GU = upper.ngroups();
        fwd_sched.clear();
        bwd_sched.clear();
        defer_bwd.assign(std::max(GU, 1), 0);
        if (merge_first_forward) defer_bwd[GU - 1] = 1;
        const bool pairs = pair_sweeps && rainbow;
        const int g0 = merge_first_forward ? 1 : 0;
        for (int g = g0; g < GL;) {
            // the first launch forms alpha and runs alone
            const int ug = g > g0 && pairs && pairable(plan.lower, g) ? upper_of(g) : -1;
            if (ug >= 0) {
                fwd_sched.push_back({g, true});
                defer_bwd[ug] = 1;
                g += 2;
            } else {
                fwd_sched.push_back({g, false});
                g += 1;
            }
        }
        std::vector<int> pend;
        if (fuse_first_backward) pend.push_back(0);
        for (int g = fuse_first_backward ? 1 : 0; g < GU;) {
            if (pairs && pairable(plan.upper, g)) {
                bwd_sched.push_back({g, true});
                pend = {g, g + 1};
                g += 2;
            } else {
                bwd_sched.push_back({g, false});
                pend = {g};
                g += 1;
            }
        }
        spmv_tail = pend;
    }

    SweepGroup sweep_group(const DevPartT<T>& P, int g) const
    {
        return SweepGroup{P.group_begin(g), P.group_count(g), P.row0(g),
                          P.nrows_of(g),    P.base(g),        P.width(g)};
    }

    // The iteration's forward sweep (FWD_FIRST, then FWD_UPDATE), by the
    // schedule; dd: the iteration's direction.
    void forward_iteration(hipStream_t s, int p, const T* dd)
    {
        const int nspmv = spmv_blocks(full.nslices * spmv_lanes);
        const int G = lower.ngroups();
        const int g0 = merge_first_forward ? 1 : 0;
        for (const SweepLaunch& e : fwd_sched) {
            const int g = e.g;
            const int last = e.pair ? g + 1 : g;
            const bool fuse = fuse_first_backward && last == G - 1;
            double* zpart = fuse ? pb + upper.group_begin(0) : nullptr;
            if (e.pair) {
                // A: r, q, diag, m; B: x r/w, d, r r/w, q, diag, m (+ z)
                const double bytes = lower.bytes(g, 32.0) + lower.bytes(g + 1, fuse ? 72.0 : 64.0);
                run(s, SFA_KERNEL_FORWARD, bytes, [&] {
                    launch_forward_pair(s, lower.view(), sweep_group(lower, g),
                                        sweep_group(lower, g + 1), state, x, r, dd, q, m, diag,
                                        w(), fuse ? z : (T*)nullptr, c2(), zpart);
                });
                continue;
            }
            const int gm = g == g0 ? FWD_FIRST : FWD_UPDATE;
            const bool merged = merge_first_forward && g == 1;
            const double bytes = lower.bytes(g, 64.0 + (fuse ? 8.0 : 0.0)) +
                                 (merged ? lower.bytes(0, 32.0) : 0.0);
            run(s, SFA_KERNEL_FORWARD, bytes, [&] {
                launch_forward(s, lower.view(), lower.group_begin(g), lower.group_count(g),
                               lower.row0(g), lower.nrows_of(g), state, gm, p, x, r, dd, q, m,
                               diag, w(), ps, nspmv, fuse ? z : (T*)nullptr, c2(), zpart,
                               lower.base(g), lower.width(g), merged ? lower.row0(0) : 0,
                               merged ? lower.nrows_of(0) : 0);
            });
        }
    }

    // The iteration's backward sweep with the z.r partials, by the schedule:
    // every launch sums the partials of the groups of the launch before it.
    void backward_iteration(hipStream_t s, const T* dd)
    {
        std::vector<int> pend;
        if (fuse_first_backward) pend.push_back(0);
        for (const SweepLaunch& e : bwd_sched) {
            const int g = e.g;
            const int fa = pend.size() > 0 ? pend[0] : -1;
            const int fb = pend.size() > 1 ? pend[1] : -1;
            const int f0a = fa >= 0 ? upper.group_begin(fa) : 0, fna = fa >= 0 ? upper.group_count(fa) : 0;
            const int f0b = fb >= 0 ? upper.group_begin(fb) : 0, fnb = fb >= 0 ? upper.group_count(fb) : 0;
            double* sa = fa >= 0 ? grp + fa : (double*)nullptr;
            double* sb = fb >= 0 ? grp + fb : (double*)nullptr;
            if (e.pair) {
                const double bytes = upper.bytes(g, defer_bwd[g] ? 64.0 : 32.0) +
                                     upper.bytes(g + 1, defer_bwd[g + 1] ? 64.0 : 32.0);
                run(s, SFA_KERNEL_BACKWARD, bytes, [&] {
                    launch_backward_pair(s, upper.view(), sweep_group(upper, g),
                                         sweep_group(upper, g + 1), state, r, m, z, diag, w(), c2(),
                                         pb, (int)pend.size(), f0a, fna, sa, f0b, fnb, sb,
                                         defer_bwd[g], defer_bwd[g + 1], x, dd, q);
                });
                pend = {g, g + 1};
                continue;
            }
            run(s, SFA_KERNEL_BACKWARD, upper.bytes(g, defer_bwd[g] ? 64.0 : 32.0), [&] {
                launch_backward(s, upper.view(), upper.group_begin(g), upper.group_count(g),
                                upper.row0(g), upper.nrows_of(g), state, 1, r, m, z, diag, w(),
                                c2(), pb, f0a, fna, sa, f0b, fnb, sb, upper.base(g),
                                upper.width(g), defer_bwd[g], x, dd, q);
            });
            pend = {g};
        }
    }

    // mode FWD_RESTART / FWD_UPDATE; the first group of an iteration runs as
    // FWD_FIRST (alpha from the SpMV's partials).
    // The last forward group also runs the first backward group when that
    // group has the same rows and no upper entries (fuse_first_backward); the
    // first forward group runs inside the second one's kernel
    // (merge_first_forward).
    void forward_sweep(hipStream_t s, const DevPartT<T>& L, int mode, int p, const T* dd)
    {
        // vector streams per row: x r/w, d, r r/w, q, diag, m (iteration);
        // r, diag, m (restart); + z when fused
        const double per_row = mode == FWD_RESTART ? 24.0 : 64.0;
        const int nspmv = spmv_blocks(full.nslices * spmv_lanes);
        const int G = L.ngroups();
        const int g0 = merge_first_forward ? 1 : 0;
        for (int g = g0; g < G; g++) {
            const int gm = (mode != FWD_RESTART && g == g0) ? FWD_FIRST : mode;
            const bool fuse = fuse_first_backward && g == G - 1;
            const bool merged = merge_first_forward && g == 1;
            double* zpart = fuse && mode != FWD_RESTART ? pb + upper.group_begin(0) : nullptr;
            const double bytes = L.bytes(g, per_row + (fuse ? 8.0 : 0.0)) +
                                 (merged ? L.bytes(0, mode == FWD_RESTART ? 24.0 : 32.0) : 0.0);
            run(s, SFA_KERNEL_FORWARD, bytes, [&] {
                launch_forward(s, L.view(), L.group_begin(g), L.group_count(g), L.row0(g),
                               L.nrows_of(g), state, gm, p, x, r, dd, q, m, diag, w(), ps, nspmv,
                               fuse ? z : (T*)nullptr, c2(), zpart, L.base(g), L.width(g),
                               merged ? L.row0(0) : 0, merged ? L.nrows_of(0) : 0);
            });
        }
    }

    // dd: the iteration's direction (for the deferred x update of a merged
    // first forward group, made by the last backward kernel).
    void backward_sweep(hipStream_t s, int reduce, const T* dd)
    {
        const int G = upper.ngroups();
        for (int g = fuse_first_backward ? 1 : 0; g < G; g++) {  // m, diag, z, r
            const bool fold_prev = reduce && g > 0;
            const bool defer = merge_first_forward && reduce && g == G - 1;
            run(s, SFA_KERNEL_BACKWARD, upper.bytes(g, defer ? 64.0 : 32.0), [&] {
                launch_backward(s, upper.view(), upper.group_begin(g), upper.group_count(g),
                                upper.row0(g), upper.nrows_of(g), state, reduce, r, m, z, diag,
                                w(), c2(), pb, fold_prev ? upper.group_begin(g - 1) : 0,
                                fold_prev ? upper.group_count(g - 1) : 0,
                                fold_prev ? grp + (g - 1) : (double*)nullptr, 0, 0,
                                (double*)nullptr, upper.base(g), upper.width(g), defer ? 1 : 0,
                                x, dd, q);
            });
        }
    }

    void eisenstat_sweep(hipStream_t s, const DevPartT<T>& P, int with_t, const T* in,
                         const T* t, T* out)
    {
        const double per_row = with_t ? 32.0 : 24.0;  // in, diag, out [, t]
        for (int g = 0; g < P.ngroups(); g++)
            run(s, SFA_KERNEL_TRIANGULAR, P.bytes(g, per_row), [&] {
                launch_eisenstat_solve(s, P.view(), P.group_begin(g), P.group_count(g), P.row0(g),
                                       P.nrows_of(g), state, with_t, in, t, out, diag, w(),
                                       c3(), P.base(g), P.width(g));
            });
    }

    // Restart (K = 0) of each variant; the direction goes to d (parity 0).
    void restart(hipStream_t s)
    {
        const SellViewT<T> F = full.view();
        const double na = n_active;
        switch (method) {
        case SFA_METHOD_CG:
            run(s, SFA_KERNEL_RESTART, full.bytes_all(32.0), [&] {
                launch_residual(s, RES_PLAIN, F, full.nslices, full.row0(0), full.nrows_of(0),
                                state, b, x, r, d, (T*)nullptr, (const T*)nullptr, pu);
            });
            break;
        case SFA_METHOD_DIAGONAL:
            run(s, SFA_KERNEL_RESTART, full.bytes_all(48.0), [&] {
                launch_residual(s, RES_JACOBI, F, full.nslices, full.row0(0), full.nrows_of(0),
                                state, b, x, r, d, z, diag, pu);
            });
            break;
        case SFA_METHOD_SSOR:
        case SFA_METHOD_RAINBOW:
            run(s, SFA_KERNEL_RESTART, full.bytes_all(16.0), [&] {
                launch_residual(s, RES_ONLY, F, full.nslices, full.row0(0), full.nrows_of(0),
                                state, b, x, r, (T*)nullptr, (T*)nullptr, (const T*)nullptr, pu);
            });
            forward_sweep(s, lower, FWD_RESTART, 0, d);
            backward_sweep(s, 0, d);
            run(s, SFA_KERNEL_RESTART, 24.0 * na,
                [&] { launch_direction_init(s, state, n_active, z, r, d, pu); });
            break;
        default:  // Eisenstat variants: q = b - A x, then r = lower solve of q
            run(s, SFA_KERNEL_RESTART, full.bytes_all(16.0), [&] {
                launch_residual(s, RES_ONLY, F, full.nslices, full.row0(0), full.nrows_of(0),
                                state, b, x, q, (T*)nullptr, (T*)nullptr, (const T*)nullptr, pu);
            });
            eisenstat_sweep(s, has_lower_restart ? lower_restart : lower, 0, q, (const T*)nullptr, r);
            run(s, SFA_KERNEL_RESTART, 32.0 * na, [&] {
                launch_eisenstat_init(s, state, n_active, r, diag, w(), z, d, pu);
            });
            break;
        }
    }

    // One CG iteration of parity p: reads the direction of slot p, writes
    // slot 1-p.  The z.r of the previous iteration is folded in by this
    // iteration's first kernel.
    void iteration(hipStream_t s, int p)
    {
        const SellViewT<T> F = full.view();
        const double na = n_active;
        T* dold = p ? d1 : d;
        T* dnew = p ? d : d1;
        const double spmv_bytes = full.bytes_all(32.0);  // src, dold, dnew, q
        const int nspmv = spmv_blocks(full.nslices * spmv_lanes);
        const int nvec = vec_reduce_blocks(n_active);
        switch (method) {
        case SFA_METHOD_CG:
        case SFA_METHOD_DIAGONAL: {
            const bool jac = method == SFA_METHOD_DIAGONAL;
            run(s, SFA_KERNEL_SPMV, spmv_bytes, [&] {
                launch_spmv_dir(s, F, full.nslices, full.row0(0), full.nrows_of(0), n_active,
                                state, p, jac ? z : r, dold, dnew, q, ps, pu, nvec,
                                (const double*)nullptr, 0, (const double*)nullptr, 0,
                                full.width(0), spmv_lanes);
            });
            run(s, SFA_KERNEL_UPDATE, (jac ? 64.0 : 48.0) * na, [&] {
                launch_update(s, jac ? UPD_JACOBI : UPD_PLAIN, state, p, n_active, x, r, dnew, q,
                              jac ? z : (T*)nullptr, jac ? diag : (const T*)nullptr, w(), ps,
                              nspmv, pu);
            });
            break;
        }
        case SFA_METHOD_SSOR:
        case SFA_METHOD_RAINBOW: {
            // z.r of the last backward launch's groups comes from their block
            // partials, the other groups' sums from grp
            const int G = upper.ngroups();
            const int nt = (int)spmv_tail.size();
            const int ta = nt > 0 ? spmv_tail[0] : -1, tb = nt > 1 ? spmv_tail[1] : -1;
            run(s, SFA_KERNEL_SPMV, spmv_bytes, [&] {
                launch_spmv_dir(s, F, full.nslices, full.row0(0), full.nrows_of(0), n_active,
                                state, p, z, dold, dnew, q, ps,
                                ta >= 0 ? pb + upper.group_begin(ta) : pb,
                                ta >= 0 ? upper.group_count(ta) : 0,
                                tb >= 0 ? pb + upper.group_begin(tb) : (const double*)nullptr,
                                tb >= 0 ? upper.group_count(tb) : 0, grp, G - nt,
                                full.width(0), spmv_lanes);
            });
            forward_iteration(s, p, dnew);
            backward_iteration(s, dnew);
            break;
        }
        default:  // Eisenstat: t (in m) = upper solve of d, q = lower solve
            for (int g = 0; g < upper.ngroups(); g++)  // z, dold, dnew, diag, m
                run(s, SFA_KERNEL_TRIANGULAR, upper.bytes(g, 40.0), [&] {
                    launch_eisenstat_upper(s, upper.view(), upper.group_begin(g),
                                           upper.group_count(g), upper.row0(g), upper.nrows_of(g),
                                           state, g == 0, p, z, dold, dnew, m, diag, w(), pu,
                                           nvec, upper.base(g), upper.width(g));
                });
            eisenstat_sweep(s, lower, 1, dnew, m, q);
            run(s, SFA_KERNEL_COMBINE, 32.0 * na,
                [&] { launch_eisenstat_combine(s, state, n_active, q, m, dnew, ps); });
            run(s, SFA_KERNEL_UPDATE, 64.0 * na, [&] {
                launch_update(s, UPD_EISENSTAT, state, p, n_active, x, r, m, q, z, diag, w(), ps,
                              nvec, pu);
            });
            break;
        }
    }

    int kernels_per_iteration() const
    {
        switch (method) {
        case SFA_METHOD_CG:
        case SFA_METHOD_DIAGONAL: return 2;
        case SFA_METHOD_SSOR:
        case SFA_METHOD_RAINBOW: return 1 + (int)fwd_sched.size() + (int)bwd_sched.size();
        default: return 2 + lower.ngroups() + upper.ngroups();
        }
    }

    void build_chunk_graph(int iters)
    {
        if (chunk_exec && chunk_iters == iters) return;
        if (chunk_exec) {
            SFA_HIP(hipGraphExecDestroy(chunk_exec));
            chunk_exec = nullptr;
        }
        // Recorded node by node (no stream capture), so other threads may use
        // the runtime meanwhile.
        chunk_graph.reset();
        auto rec = std::make_unique<GraphRecorder>();
        for (int i = 0; i < iters; i++) iteration(own_stream, i & 1);
        const hipError_t e = rec->instantiate(&chunk_exec);
        chunk_graph = std::move(rec);
        if (e != hipSuccess) {
            chunk_exec = nullptr;
            throw Error{SFA_ERR_HIP, std::string("iteration graph: ") + hipGetErrorString(e)};
        }
        chunk_iters = iters;
    }

    void read_state()
    {
        SFA_HIP(hipMemcpyAsync(host_state, state, sizeof(SolveState), hipMemcpyDeviceToHost, stream));
        SFA_HIP(hipStreamSynchronize(stream));
    }

    // Fills the scratch vectors the way the reference allocates them:
    // zeros for SSOR / Rainbow (:493-494, :1056-1057), NaN otherwise.
    void init_scratch(hipStream_t s)
    {
        const T init = (method == SFA_METHOD_SSOR || method == SFA_METHOD_RAINBOW) ? T(0) : T(NAN);
        for (T* v : {d, d1, z, q, m}) launch_fill(s, v, init, n);
    }

    // ---------------------------------------------------------------- solve

    // Copies the inputs into device order, resets the state and runs the
    // restart.  Returns the stream the solve runs on.
    hipStream_t begin_solve(T* ux, const T* ub, T* ures, double precision, int print_modulo,
                            int device_pointers, double rel2 = 0.)
    {
        if (!(precision > 0)) throw Error{SFA_ERR_INVALID, "precision must be positive"};
        SFA_HIP(hipSetDevice(device));
        hipStream_t s = stream;
        const T* src_x = ux;
        const T* src_b = ub;
        const T* src_r = ures;
        double host_tol = 0.;
        int use_host_tol = 0;
        if (!device_pointers) {
            // solver/sparsesolver.cpp:469-472 on the host, in row order
            double t = 0.;
            for (int i = 0; i < n; i++) t += (double)ub[i] * (double)ub[i];
            t = precision * std::sqrt(t);
            const double floor_ = 1.4901161193847656e-08;
            if (!(t >= floor_) && t <= floor_) t = floor_;
            host_tol = t;
            use_host_tol = 1;
            SFA_HIP(hipMemcpyAsync(io, ux, n * sizeof(T), hipMemcpyHostToDevice, s));
            SFA_HIP(hipMemcpyAsync(io2, ub, n * sizeof(T), hipMemcpyHostToDevice, s));
            SFA_HIP(hipMemcpyAsync(io3, ures, n * sizeof(T), hipMemcpyHostToDevice, s));
            src_x = io;
            src_b = io2;
            src_r = io3;
        } else {
            order_after_null_stream(s);
        }
        launch_gather(s, x, src_x, perm, n);
        launch_gather(s, b, src_b, perm, n);
        launch_gather(s, r, src_r, perm, n);
        // device b: ||b|| in the reference's row order when asked for
        const T* rows_b = device_pointers && row_order_tol ? src_b : nullptr;
        start(s, precision, print_modulo, host_tol, use_host_tol, rel2, rows_b);
        return s;
    }

    // State, scratch and restart for x, b, r already in device order.
    void start(hipStream_t s, double precision, int print_modulo, double host_tol,
               int use_host_tol, double rel2, const T* rows_b = nullptr)
    {
        init_scratch(s);
        const int small_breaks = (method == SFA_METHOD_SSOR_EISENSTAT) ? (print_modulo >= -2) : 1;
        launch_init_state(s, state, b, n, precision, host_tol, use_host_tol, n, small_breaks,
                          rel2, ps, rows_b);
        restart(s);
        read_state();
    }

    void finish_solve(hipStream_t s, T* ux, T* ures, int print_modulo, int device_pointers,
                      sfa_solve_info* info)
    {
        T* dst_x = device_pointers ? ux : io;
        T* dst_r = device_pointers ? ures : io3;
        launch_scatter(s, dst_x, x, perm, n);
        launch_scatter(s, dst_r, r, perm, n);
        if (!device_pointers) {
            SFA_HIP(hipMemcpyAsync(ux, io, n * sizeof(T), hipMemcpyDeviceToHost, s));
            SFA_HIP(hipMemcpyAsync(ures, io3, n * sizeof(T), hipMemcpyDeviceToHost, s));
        }
        SFA_HIP(hipStreamSynchronize(s));
        SFA_HIP(hipGetLastError());

        const SolveState& st = *host_state;
        sfa_solve_info out{};
        out.iterations = st.K;
        out.zr = st.zr;
        out.tolerance = st.tol;
        out.exit_reason = st.exit_reason;
        out.success = std::sqrt(st.zr) < st.tol ? 1 : 0;
        if (info) *info = out;
        if (print_modulo >= -2) print_log(print_modulo, st);
    }

    void solve(T* ux, const T* ub, T* ures, double precision, int print_modulo,
               int device_pointers, sfa_solve_info* info)
    {
        hipStream_t s = begin_solve(ux, ub, ures, precision, print_modulo, device_pointers);
        iterate(s);
        finish_solve(s, ux, ures, print_modulo, device_pointers, info);
    }

    // Runs iterations until the state says the loop has ended.
    void iterate(hipStream_t s)
    {
        // Iterations: replay a captured chunk, predicting from the z.r history
        // how many chunks to queue before the next look at the state.
        const int G = 4;
        build_chunk_graph(G);
        int queued = 4;
        while (!host_state->done) {
            for (int i = 0; i < queued; i++) SFA_HIP(hipGraphLaunch(chunk_exec, s));
            read_state();
            if (host_state->done) break;
            queued = predict_chunks(G);
        }
    }

    // The same solve, every kernel launched on its own between two events.
    void profile(T* ux, const T* ub, T* ures, double precision, int device_pointers,
                 sfa_kernel_profile* out, sfa_solve_info* info)
    {
        Profiler p;
        prof = &p;
        try {
            hipStream_t s = begin_solve(ux, ub, ures, precision, -3, device_pointers);
            for (int p = 0; !host_state->done; p ^= 1) {
                iteration(s, p);
                read_state();
            }
            finish_solve(s, ux, ures, -3, device_pointers, info);
        } catch (...) {
            prof = nullptr;
            throw;
        }
        prof = nullptr;
        for (int c = 0; c < SFA_KERNEL_CLASSES; c++) out[c] = sfa_kernel_profile{};
        for (size_t i = 0; i < p.klass.size(); i++) {
            float ms = 0.f;
            SFA_HIP(hipEventElapsedTime(&ms, p.event(2 * i), p.event(2 * i + 1)));
            sfa_kernel_profile& o = out[p.klass[i]];
            o.launches += 1;
            o.milliseconds += ms;
            o.bytes += p.bytes[i];
        }
    }

    // Microseconds per launch of the iteration's forward (or backward) sweep
    // kernels as they run back to back inside a recorded graph: the solve of
    // the last right-hand side is restarted from x = 0 and run one iteration,
    // then `reps` copies of one iteration's launches of that class are
    // recorded and the graph is replayed between two HIP events (after one
    // replay to warm up).  The kernels repeat their own arithmetic on
    // evolving values; the next solve restarts, so nothing of this remains.
    double time_sweeps(int klass, int reps)
    {
        if (method != SFA_METHOD_SSOR && method != SFA_METHOD_RAINBOW)
            throw Error{SFA_ERR_INVALID, "sweep timing needs an SSOR or Rainbow system"};
        if (klass != SFA_KERNEL_FORWARD && klass != SFA_KERNEL_BACKWARD)
            throw Error{SFA_ERR_INVALID, "kernel class must be SFA_KERNEL_FORWARD or _BACKWARD"};
        if (reps < 1 || reps > 100000) throw Error{SFA_ERR_INVALID, "reps must be 1..100000"};
        SFA_HIP(hipSetDevice(device));
        hipStream_t s = stream;
        launch_fill(s, x, T(0), n);
        start(s, 1e-10, -3, 0., 0, 0.);
        if (host_state->done) throw Error{SFA_ERR_INVALID, "the restart already converged"};
        iteration(s, 0);
        const size_t per = klass == SFA_KERNEL_FORWARD ? fwd_sched.size() : bwd_sched.size();
        hipGraphExec_t exec = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        std::unique_ptr<GraphRecorder> rec;
        auto cleanup = [&] {
            if (exec) (void)hipGraphExecDestroy(exec);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
        };
        try {
            rec = std::make_unique<GraphRecorder>();
            for (int i = 0; i < reps; i++) {  // the launches of the iteration of parity 1
                if (klass == SFA_KERNEL_FORWARD) forward_iteration(s, 1, d);
                else backward_iteration(s, d);
            }
            const hipError_t e = rec->instantiate(&exec);
            if (e != hipSuccess) throw Error{SFA_ERR_HIP, std::string("sweep graph: ") + hipGetErrorString(e)};
            SFA_HIP(hipEventCreate(&e0));
            SFA_HIP(hipEventCreate(&e1));
            SFA_HIP(hipGraphLaunch(exec, s));
            SFA_HIP(hipEventRecord(e0, s));
            SFA_HIP(hipGraphLaunch(exec, s));
            SFA_HIP(hipEventRecord(e1, s));
            SFA_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            SFA_HIP(hipEventElapsedTime(&ms, e0, e1));
            cleanup();
            return ms * 1e3 / ((double)reps * (double)per);
        } catch (...) {
            cleanup();
            throw;
        }
    }

    int predict_chunks(int G)
    {
        // Geometric model of the last few z.r values: remaining iterations
        // until |z.r| < tol^2, capped to keep the look-ahead short.
        const int K = host_state->K;
        const double zr = host_state->zr, tol2 = host_state->tol2;
        const int look = std::min(K, 8);
        int want = 2 * G;
        if (look >= 2 && zr > 0) {
            double hk[2];
            SFA_HIP(hipMemcpyAsync(&hk[0], hist + (K - look), sizeof(double),
                                   hipMemcpyDeviceToHost, stream));
            SFA_HIP(hipStreamSynchronize(stream));
            hk[1] = zr;
            if (hk[0] > zr && hk[0] > 0) {
                const double rate = std::log(zr / hk[0]) / look;  // < 0
                const double need = std::log(tol2 / zr) / rate;
                if (need > 0 && std::isfinite(need)) want = (int)std::ceil(need) + 1;
                else want = 1;
            }
        }
        int chunks = (want + G - 1) / G;
        return std::max(1, std::min(chunks, 16));
    }

    void print_log(int pm, const SolveState& st)
    {
        // Lines and formats of solver/sparsesolver.cpp:498-499,586-588,593-596,630-641,713-714.
        const int K = st.K;
        if (pm >= -1) std::printf("(%d/%d)     BEGIN: %s\n", 0, n, method_title(method));
        std::vector<double> h(K + 1);
        if (pm >= 0 && K + 1 > 0) {
            SFA_HIP(hipMemcpyAsync(h.data(), hist, (K + 1) * sizeof(double),
                                   hipMemcpyDeviceToHost, stream));
            SFA_HIP(hipStreamSynchronize(stream));
        }
        if (pm >= 0)
            std::printf("(%d/%d)   RESTART: Residual norm is %.9le < %.9le\n", 0, n, std::sqrt(h[0]), st.tol);
        if (pm >= 1) {
            const int last = st.exit_reason == SFA_EXIT_CAP ? K - 1 : K;
            for (int k = 0; k <= last; k += pm)
                std::printf("(%d/%d)   INTERIM: Residual norm is %.9le < %.9le\n", k, n, std::sqrt(h[k]), st.tol);
        }
        if (st.exit_reason == SFA_EXIT_BREAKDOWN)
            std::printf("(%d/%d) BREAKDOWN: Gradient energy is unreasonable with %.9le\n", K, n, st.dAd);
        if (st.exit_reason == SFA_EXIT_SMALL) {
            std::printf("(%d/%d)   WARNING: Residual norm is %.9le < %.9le\n", K, n, std::sqrt(st.zr), st.tol);
            std::printf("(%d/%d)   WARNING: Gradient energy is small with %.9le\n", K, n, st.dAd);
        }
        if (pm >= -1)
            std::printf("(%d/%d) %9s: Residual norm is %.9le < %.9le\n", K, n,
                        std::sqrt(st.zr) < st.tol ? "SUCCESS" : "FAILED", std::sqrt(st.zr), st.tol);
        std::fflush(stdout);
    }

    void apply(T* dest, const T* src, double scaling, int device_pointers)
    {
        SFA_HIP(hipSetDevice(device));
        if (!has_apply_part) {
            if (!csr_A) throw Error{SFA_ERR_INVALID, "no matrix kept for apply"};
            SellPart p;
            build_full_identity(p, n, n, csr_A, csr_C, csr_V);
            apply_part.upload(p);
            has_apply_part = true;
        }
        hipStream_t s = stream;
        // source with a zero slot at n: use the x-sized buffer m
        T* srcbuf = q;
        if (device_pointers) {
            order_after_null_stream(s);
            SFA_HIP(hipMemcpyAsync(srcbuf, src, n * sizeof(T), hipMemcpyDeviceToDevice, s));
        } else
            SFA_HIP(hipMemcpyAsync(srcbuf, src, n * sizeof(T), hipMemcpyHostToDevice, s));
        launch_fill(s, srcbuf + n, T(0), 1);
        T* out = device_pointers ? dest : io;
        launch_apply(s, apply_part.view(), apply_part.nslices, apply_part.row0(0), apply_part.nrows_of(0),
                     srcbuf, out, T(scaling), spmv_lanes);
        if (!device_pointers)
            SFA_HIP(hipMemcpyAsync(dest, io, n * sizeof(T), hipMemcpyDeviceToHost, s));
        SFA_HIP(hipStreamSynchronize(s));
    }

    // FP64 iterative refinement around FP32 inner solves of the same variant:
    // r = b - A x (FP64) -> e = A^-1 r in FP32 to a reduction inner_rtol of
    // the preconditioned residual -> x += e (FP64), until ||b - A x||_2 <=
    // max(rtol * ||b||_2, 2^-26) or max_outer corrections.
    //
    // Unless inner_restarts is set, an inner solve after the first continues
    // the previous one's search direction instead of starting from d = z:
    // the FP32 CG runs on as one Krylov iteration whose residual is replaced
    // by the FP64 one at every correction (residual replacement, van der
    // Vorst and Ye, SIAM J. Sci. Comput. 22 (2000) 1716-1728), so the
    // corrections do not throw away the Krylov space built so far.
    // A continued inner solve needs the restart's z.r within this factor of
    // the z.r the previous inner solve ended with.
    static constexpr double kContinueGap = 1.25;

    void solve_mixed(T* ux, const T* ub, double rtol, double inner_rtol, int max_outer,
                     int device_pointers, sfa_refine_info* info)
    {
        if (!inner) throw Error{SFA_ERR_INVALID, "system was created without SFA_FLAG_FP32_INNER"};
        if (!(rtol > 0) || !(inner_rtol > 0) || !(inner_rtol < 1) || max_outer < 0)
            throw Error{SFA_ERR_INVALID, "need rtol > 0, 0 < inner_rtol < 1, max_outer >= 0"};
        SFA_HIP(hipSetDevice(device));
        hipStream_t s = stream;
        inner->stream = s;
        const T* src_x = ux;
        const T* src_b = ub;
        double bnorm2 = 0.;
        if (!device_pointers) {
            for (int i = 0; i < n; i++) bnorm2 += (double)ub[i] * (double)ub[i];
            SFA_HIP(hipMemcpyAsync(io, ux, n * sizeof(T), hipMemcpyHostToDevice, s));
            SFA_HIP(hipMemcpyAsync(io2, ub, n * sizeof(T), hipMemcpyHostToDevice, s));
            src_x = io;
            src_b = io2;
        } else {
            order_after_null_stream(s);
        }
        launch_gather(s, x, src_x, perm, n);
        launch_gather(s, b, src_b, perm, n);
        if (device_pointers) {  // ||b|| on the device (k_init_state with precision 1)
            launch_init_state(s, state, b, n, 1.0, 0., 0, n, 1, 0., ps);
            read_state();
            bnorm2 = host_state->tol * host_state->tol;
        }
        double target = rtol * std::sqrt(bnorm2);
        const double floor_ = 1.4901161193847656e-08;
        if (!(target >= floor_)) target = floor_;
        const bool can_continue = !inner_restarts && inner->has_spmv();
        if (can_continue && !inner->dkeep) inner->dkeep = dev_alloc<float>(n);
        int outer = 0, continued = 0;
        long long inner_total = 0;
        double res = 0.;
        double zr_prev = 0.;  // > 0: the last inner solve left a direction to continue
        double zr_end = 0.;   // the recursive z.r that solve ended with
        int k_end = 0;        // and its iterations
        for (;;) {
            launch_residual_norm(s, full.view(), full.nslices, full.row0(0), full.nrows_of(0),
                                 state, b, x, r, pu);
            read_state();
            res = std::sqrt(host_state->result);
            if (!(res > target) || outer >= max_outer) break;
            launch_to_float(s, inner->b, r, n_active);
            launch_fill(s, inner->x, 0.0f, n);
            launch_fill(s, inner->r, 0.0f, n);
            if (zr_prev > 0.) {
                // the direction of the last inner iteration: slot K % 2 (the
                // iteration of parity p reads slot p)
                const float* dlast = (inner->host_state->K & 1) ? inner->d1 : inner->d;
                SFA_HIP(hipMemcpyAsync(inner->dkeep, dlast, n * sizeof(float),
                                       hipMemcpyDeviceToDevice, s));
            }
            inner->start(s, 1.0, -3, 0., 0, inner_rtol * inner_rtol);
            // Continue only where the FP64 residual agrees with the recursive
            // one the last inner solve ended with: a drifted FP32 residual
            // makes the old direction unfit (beta far too large).
            if (zr_prev > 0. && !inner->host_state->done &&
                inner->host_state->zr <= kContinueGap * zr_end) {
                launch_direction_continue(s, inner->state, inner->n_active, inner->d,
                                          (const float*)inner->dkeep, zr_prev,
                                          (int)std::min<long long>(inner->n, std::max(100LL, 3LL * k_end)));
                continued++;
            }
            inner->iterate(s);
            const SolveState& ist = *inner->host_state;
            if (ist.exit_reason == SFA_EXIT_BREAKDOWN) break;
            // z.r of the iterate that formed the last direction (the stopping
            // iteration read it from zrp[K % 2])
            zr_prev = (can_continue && ist.exit_reason == SFA_EXIT_CONVERGED && ist.K >= 1)
                          ? ist.zrp[ist.K & 1]
                          : 0.;
            if (!(zr_prev > 0.) || !std::isfinite(zr_prev)) zr_prev = 0.;
            zr_end = ist.zr;
            k_end = ist.K;
            inner_total += ist.K;
            launch_add_correction(s, x, inner->x, n_active);
            outer++;
        }
        T* dst_x = device_pointers ? ux : io;
        launch_scatter(s, dst_x, x, perm, n);
        if (!device_pointers)
            SFA_HIP(hipMemcpyAsync(ux, io, n * sizeof(T), hipMemcpyDeviceToHost, s));
        SFA_HIP(hipStreamSynchronize(s));
        SFA_HIP(hipGetLastError());
        if (info) {
            sfa_refine_info o{};
            o.outer = outer;
            o.inner_iterations = (int)inner_total;
            o.success = res <= target ? 1 : 0;
            o.continued = continued;
            o.residual_norm = res;
            o.relative_residual = bnorm2 > 0 ? res / std::sqrt(bnorm2) : res;
            *info = o;
        }
    }

    bool has_spmv() const
    {
        return method == SFA_METHOD_CG || method == SFA_METHOD_DIAGONAL ||
               method == SFA_METHOD_SSOR || method == SFA_METHOD_RAINBOW;
    }

    void set_spmv(int lanes)
    {
        if (lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8)
            throw Error{SFA_ERR_INVALID, "lanes per row must be 1, 2, 4 or 8"};
        if (lanes == spmv_lanes) return;
        spmv_lanes = lanes;
        if (chunk_exec) SFA_HIP(hipGraphExecDestroy(chunk_exec));  // recorded with the old mapping
        chunk_exec = nullptr;
        chunk_graph.reset();
        chunk_iters = 0;
    }

    // Times each mapping on the iteration's SpMV kernel, as it runs in a
    // solve (a fold, two gathers per entry), against a scratch state whose
    // loop never ends, and installs the fastest.
    void tune_spmv(sfa_spmv_tuning* out)
    {
        sfa_spmv_tuning t{};
        for (int bk = 0; bk < kSpmvBuckets; bk++) t.rows[bk] = full.bucket_rows[bk];
        auto report = [&](int lanes) {
            for (int bk = 0; bk < kSpmvBuckets; bk++) t.lanes_per_row[bk] = t.rows[bk] ? lanes : 0;
        };
        if (!has_spmv() || full.nslices == 0) {
            report(1);
            *out = t;
            return;
        }
        const int G = sweeps ? upper.ngroups() : 0;
        const int nfold = G > 0 ? upper.group_count(G - 1) : vec_reduce_blocks(n_active);
        const int ngroups = G > 0 ? G - 1 : 0;
        const int nscratch = std::max({kSpmvMaxBlocks, nfold, ngroups, 1});
        SolveState* ts = dev_alloc<SolveState>(1);
        double* th = dev_alloc<double>(2);
        double* tz = dev_alloc<double>(nscratch);  // zero partials to fold
        double* tp = dev_alloc<double>(kSpmvMaxBlocks);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        auto cleanup = [&] {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            for (void* v : {(void*)ts, (void*)th, (void*)tz, (void*)tp}) dev_free(v);
        };
        try {
            SolveState h{};
            h.tol2 = -1.;  // never converged
            h.maxK = INT_MAX;
            h.fold[0] = 1;
            h.zrp[0] = 1.;
            h.hist = th;
            SFA_HIP(hipMemcpyAsync(ts, &h, sizeof(h), hipMemcpyHostToDevice, stream));
            launch_fill<double>(stream, tz, 0., nscratch);
            SFA_HIP(hipEventCreate(&e0));
            SFA_HIP(hipEventCreate(&e1));
            const SellViewT<T> F = full.view();
            auto once = [&](int lanes) {
                launch_spmv_dir(stream, F, full.nslices, full.row0(0), full.nrows_of(0), n_active,
                                ts, 0, z, d, d1, q, tp, tz, nfold, (const double*)nullptr, 0,
                                tz, ngroups, full.width(0), lanes);
            };
            const int cand[SFA_SPMV_MAPPINGS] = {1, 2, 4, 8};
            const int main = full.main_bucket();
            constexpr int kReps = 20;
            int best = 1;
            double best_us = 0.;
            for (int i = 0; i < SFA_SPMV_MAPPINGS; i++) {
                const int lanes = cand[i];
                for (int w = 0; w < 3; w++) once(lanes);
                float ms_min = 0.f;
                for (int rep = 0; rep < 3; rep++) {
                    SFA_HIP(hipEventRecord(e0, stream));
                    for (int k = 0; k < kReps; k++) once(lanes);
                    SFA_HIP(hipEventRecord(e1, stream));
                    SFA_HIP(hipEventSynchronize(e1));
                    float ms = 0.f;
                    SFA_HIP(hipEventElapsedTime(&ms, e0, e1));
                    if (rep == 0 || ms < ms_min) ms_min = ms;
                }
                SFA_HIP(hipGetLastError());
                const double us = 1000.0 * ms_min / kReps;
                t.us[main][i] = us;
                if (i == 0 || us < best_us) {
                    best = lanes;
                    best_us = us;
                }
            }
            SFA_HIP(hipStreamSynchronize(stream));
            cleanup();
            set_spmv(best);
            report(best);
            *out = t;
        } catch (...) {
            (void)hipStreamSynchronize(stream);
            cleanup();
            throw;
        }
    }

    void query(sfa_system_stats* st) const
    {
        sfa_system_stats o{};
        o.n = n;
        o.n_active = n_active;
        o.nnz = nnz;
        o.groups_forward = sweeps ? lower.ngroups() : 0;
        o.groups_backward = sweeps ? upper.ngroups() : 0;
        o.stored_entries = full.stored + lower.stored + upper.stored;
        // Algorithmic bytes, the same model the profiled launches use.
        const double na = n_active;
        const double spmv = full.bytes_all(32.0);
        double it = 0.;
        if (method == SFA_METHOD_SSOR || method == SFA_METHOD_RAINBOW)
            it = spmv + lower.bytes_all(64.0) + upper.bytes_all(32.0);
        else if (method == SFA_METHOD_CG)
            it = spmv + 48.0 * na;
        else if (method == SFA_METHOD_DIAGONAL)
            it = spmv + 64.0 * na;
        else
            it = upper.bytes_all(40.0) + lower.bytes_all(32.0) + (32.0 + 64.0) * na;
        o.bytes_spmv = spmv;
        o.bytes_per_iteration = it;
        o.kernels_per_iteration = kernels_per_iteration();
        o.spmv_lanes_per_row = spmv_lanes;
        for (int i = 0; i < 4; i++) o.setup_ms[i] = setup_ms[i];
        *st = o;
    }
};

using System = SystemT<double>;

}  // namespace sfa

using sfa::Error;
using sfa::System;

struct sfa_system_s {
    System sys;
};

namespace {

// The reference's callers make the same raw call many times with one matrix
// (a call per right-hand side, per time step, per inner solve).  The last
// few Systems built by drop-in calls are kept: a call whose arguments equal
// a kept System's -- device, method, N, omega, and the CSR arrays, the
// diagonal (or precon) and the colouring byte for byte, compared with
// memcmp against host copies -- reuses its plan, device arrays and
// recorded graphs; anything else builds a new System and the least recently
// used one is dropped.  Reuse changes no result (every solve re-reads x, b
// and residual and resets its state).  SFA_DROPIN_CACHE sets how many are
// kept (default 4, 0: none).  A System is used by one call at a time; a
// call that finds its match busy builds another.
// The device parts of the last few matrices applied through
// sfa_MatrixCSR_apply (MatrixCSR::apply of the drop-in), each with host copies
// of its arrays: a call whose arrays hold the same bytes as a kept entry's
// reuses its part, so repeated applies of one matrix skip the SELL build and
// the upload.  A call holds the lock for its whole apply; the apply waits for
// its result, so nothing is queued on a part when it is evicted.
struct ApplyCache {
    struct Entry {
        int device = 0, rows = 0, cols = 0;
        std::vector<int> A, C;
        std::vector<double> V;
        sfa::DevPart part;
        unsigned long long last_use = 0;
        ~Entry() { part.release(); }
    };
    std::mutex mu;
    std::vector<std::unique_ptr<Entry>> entries;
    unsigned long long clock = 0;
    int capacity = 4;

    ApplyCache()
    {
        if (const char* e = std::getenv("SFA_DROPIN_CACHE")) capacity = std::max(0, std::atoi(e));
    }

    static bool matches(const Entry& e, int device, int rows, int cols, const int* A,
                        const int* C, const double* V)
    {
        if (e.device != device || e.rows != rows || e.cols != cols) return false;
        const size_t nnz = (size_t)A[rows];
        return e.A.size() == (size_t)rows + 1 && e.C.size() == nnz &&
               sfa::bytes_equal(e.A.data(), A, sizeof(int) * ((size_t)rows + 1)) &&
               (nnz == 0 || (sfa::bytes_equal(e.C.data(), C, sizeof(int) * nnz) &&
                             sfa::bytes_equal(e.V.data(), V, sizeof(double) * nnz)));
    }

    // The part for these arrays: a kept one, or a new one kept in place of
    // the least recently used (held by `fresh` alone when nothing is kept).
    const sfa::DevPart& part(int device, int rows, int cols, const int* A, const int* C,
                             const double* V, std::unique_ptr<Entry>* fresh)
    {
        for (auto& e : entries)
            if (matches(*e, device, rows, cols, A, C, V)) {
                e->last_use = ++clock;
                return e->part;
            }
        std::unique_ptr<Entry> e(new Entry());
        e->device = device;
        e->rows = rows;
        e->cols = cols;
        sfa::SellPart p;
        sfa::build_full_identity(p, rows, cols, A, C, V);
        e->part.upload(p);
        e->last_use = ++clock;
        if (capacity == 0) {
            *fresh = std::move(e);
            return (*fresh)->part;
        }
        const size_t nnz = (size_t)A[rows];
        e->A.assign(A, A + rows + 1);
        e->C.assign(C, C + nnz);
        e->V.assign(V, V + nnz);
        if ((int)entries.size() >= capacity) {
            auto lru = std::min_element(entries.begin(), entries.end(), [](const auto& a, const auto& b) {
                return a->last_use < b->last_use;
            });
            entries.erase(lru);
        }
        entries.push_back(std::move(e));
        return entries.back()->part;
    }
};

ApplyCache& apply_cache()
{
    static ApplyCache* c = new ApplyCache();  // lives to the end of the process
    return *c;
}

struct DropinCache {
    struct Entry {
        int device = 0, method = 0, n = 0, nc = 0;
        double omega = 0.;
        std::vector<double> aux;
        std::vector<int> F, B, R;
        std::unique_ptr<System> sys;
        unsigned long long last_use = 0;
        bool busy = false;
    };
    std::mutex mu;
    std::vector<std::unique_ptr<Entry>> entries;
    unsigned long long clock = 0;
    int capacity = 4;

    DropinCache()
    {
        if (const char* e = std::getenv("SFA_DROPIN_CACHE")) capacity = std::max(0, std::atoi(e));
    }

    template <class T>
    static bool same(const std::vector<T>& kept, const T* p, size_t count)
    {
        if (kept.size() != count) return false;
        return count == 0 || (p && sfa::bytes_equal(kept.data(), p, count * sizeof(T)));
    }

    static bool matches(const Entry& e, int device, int method, int n, const int* A,
                        const int* C, const double* V, const double* aux, double omega, int nc,
                        const int* F, const int* B, const int* R)
    {
        const System& s = *e.sys;
        if (e.device != device || e.method != method || e.n != n || e.nc != nc) return false;
        if (std::memcmp(&e.omega, &omega, sizeof omega) != 0) return false;
        if (!same(s.host_A, A, (size_t)n + 1)) return false;
        const size_t nnz = (size_t)A[n];
        return same(s.host_C, C, nnz) && same(s.host_V, V, nnz) &&
               same(e.aux, aux, aux ? (size_t)n : 0) && same(e.F, F, F ? (size_t)n : 0) &&
               same(e.B, B, B ? (size_t)nc + 1 : 0) && same(e.R, R, R ? (size_t)n : 0);
    }

    // A System for the call: a kept one (marked busy) or a new one.
    Entry* acquire(int method, int n, const int* A, const int* C, const double* V,
                   const double* aux, double omega, int nc, const int* F, const int* B,
                   const int* R, std::unique_ptr<Entry>* fresh)
    {
        int device = 0;
        SFA_HIP(hipGetDevice(&device));
        {
            std::lock_guard<std::mutex> lock(mu);
            for (auto& e : entries)
                if (!e->busy && matches(*e, device, method, n, A, C, V, aux, omega, nc, F, B, R)) {
                    e->busy = true;
                    return e.get();
                }
        }
        std::unique_ptr<Entry> e(new Entry());
        e->sys.reset(new System());
        e->sys->create(method, n, A, C, V, aux, omega, nc, F, B, R);
        e->device = device;
        e->method = method;
        e->n = n;
        e->nc = nc;
        e->omega = omega;
        if (aux) e->aux.assign(aux, aux + n);
        if (F) e->F.assign(F, F + n);
        if (B) e->B.assign(B, B + nc + 1);
        if (R) e->R.assign(R, R + n);
        e->busy = true;
        *fresh = std::move(e);
        return fresh->get();
    }

    // Ends a call: a new System is kept (dropping the least recently used
    // idle one beyond the capacity), a kept one is marked idle.  A System
    // whose solve failed is not kept.
    void release(Entry* e, std::unique_ptr<Entry> fresh, bool ok)
    {
        std::vector<std::unique_ptr<Entry>> drop;
        {
            std::lock_guard<std::mutex> lock(mu);
            if (fresh) {
                if (ok && capacity > 0) {
                    fresh->busy = false;
                    fresh->last_use = ++clock;
                    entries.push_back(std::move(fresh));
                }
            } else {
                for (size_t i = 0; i < entries.size(); i++)
                    if (entries[i].get() == e) {
                        if (ok) {
                            e->busy = false;
                            e->last_use = ++clock;
                        } else {
                            drop.push_back(std::move(entries[i]));
                            entries.erase(entries.begin() + i);
                        }
                        break;
                    }
            }
            while ((int)entries.size() > capacity) {
                int victim = -1;
                for (size_t i = 0; i < entries.size(); i++)
                    if (!entries[i]->busy &&
                        (victim < 0 || entries[i]->last_use < entries[victim]->last_use))
                        victim = (int)i;
                if (victim < 0) break;
                drop.push_back(std::move(entries[victim]));
                entries.erase(entries.begin() + victim);
            }
        }
        // fresh (if not kept) and dropped Systems are destroyed here, unlocked
    }
};

DropinCache& dropin_cache()
{
    static DropinCache* cache = new DropinCache();  // kept until the process ends
    return *cache;
}

int solve_host(int method, int N, double* x, const double* b, const int* A, const int* C,
               const double* V, double* residual, double precision, int print_modulo,
               const double* aux, double omega, int nc, const int* F, const int* B, const int* R,
               sfa_solve_info* info)
{
    return sfa::guarded([&] {
        if (!x || !b || !residual) throw Error{SFA_ERR_INVALID, "null x, b or residual"};
        // every argument is checked on the host before the device or a kept
        // System is touched (a kept one is then compared with valid arrays)
        bool colouring_error = false;
        const std::string err = sfa::check_arguments(method, N, A, C, V, aux, nc, F, B, R,
                                                     &colouring_error);
        if (!err.empty()) throw Error{colouring_error ? SFA_ERR_COLOURING : SFA_ERR_INVALID, err};
        sfa::ensure_device();
        DropinCache& cache = dropin_cache();
        std::unique_ptr<DropinCache::Entry> fresh;
        DropinCache::Entry* e =
            cache.acquire(method, N, A, C, V, aux, omega, nc, F, B, R, &fresh);
        try {
            e->sys->solve(x, b, residual, precision, print_modulo, 0, info);
        } catch (...) {
            cache.release(e, std::move(fresh), false);
            throw;
        }
        cache.release(e, std::move(fresh), true);
    });
}

}  // namespace

extern "C" {

const char* sfa_last_error(void) { return sfa::last_error().c_str(); }

int sfa_api_version(void) { return SFA_API_VERSION; }

int sfa_ConjugateGradientSolverCSR(int N, double* x, const double* b, const int* csrrows,
                                   const int* csrcolumns, const double* csrvalues,
                                   double* residual, double precision, int print_modulo,
                                   sfa_solve_info* info)
{
    return solve_host(SFA_METHOD_CG, N, x, b, csrrows, csrcolumns, csrvalues, residual, precision,
                      print_modulo, nullptr, 1.0, 0, nullptr, nullptr, nullptr, info);
}

int sfa_ConjugateGradientSolverCSR_DiagonalPreconditioner(
    int N, double* x, const double* b, const int* csrrows, const int* csrcolumns,
    const double* csrvalues, double* residual, double precision, int print_modulo,
    const double* precon, sfa_solve_info* info)
{
    return solve_host(SFA_METHOD_DIAGONAL, N, x, b, csrrows, csrcolumns, csrvalues, residual,
                      precision, print_modulo, precon, 1.0, 0, nullptr, nullptr, nullptr, info);
}

int sfa_ConjugateGradientSolverCSR_SSOR(int N, double* x, const double* b, const int* csrrows,
                                        const int* csrcolumns, const double* csrvalues,
                                        double* residual, double precision, int print_modulo,
                                        const double* diagonal, double omega,
                                        sfa_solve_info* info)
{
    return solve_host(SFA_METHOD_SSOR, N, x, b, csrrows, csrcolumns, csrvalues, residual,
                      precision, print_modulo, diagonal, omega, 0, nullptr, nullptr, nullptr,
                      info);
}

int sfa_ConjugateGradientSolverCSR_SSOR_Eisenstat(int N, double* x, const double* b,
                                                  const int* csrrows, const int* csrcolumns,
                                                  const double* csrvalues, double* residual,
                                                  double precision, int print_modulo,
                                                  const double* diagonal, double omega,
                                                  sfa_solve_info* info)
{
    return solve_host(SFA_METHOD_SSOR_EISENSTAT, N, x, b, csrrows, csrcolumns, csrvalues,
                      residual, precision, print_modulo, diagonal, omega, 0, nullptr, nullptr,
                      nullptr, info);
}

int sfa_ConjugateGradientSolverCSR_Rainbow(int N, double* x, const double* b, const int* csrrows,
                                           const int* csrcolumns, const double* csrvalues,
                                           double* residual, double precision, int print_modulo,
                                           const double* diagonal, double omega, int num_colors,
                                           const int* F, const int* B, const int* R,
                                           sfa_solve_info* info)
{
    return solve_host(SFA_METHOD_RAINBOW, N, x, b, csrrows, csrcolumns, csrvalues, residual,
                      precision, print_modulo, diagonal, omega, num_colors, F, B, R, info);
}

int sfa_ConjugateGradientSolverCSR_Eisenstat_Rainbow(
    int N, double* x, const double* b, const int* csrrows, const int* csrcolumns,
    const double* csrvalues, double* residual, double precision, int print_modulo,
    const double* diagonal, double omega, int num_colors, const int* F, const int* B,
    const int* R, sfa_solve_info* info)
{
    return solve_host(SFA_METHOD_EISENSTAT_RAINBOW, N, x, b, csrrows, csrcolumns, csrvalues,
                      residual, precision, print_modulo, diagonal, omega, num_colors, F, B, R,
                      info);
}

int sfa_MatrixCSR_apply(int rows, int columns, const int* A, const int* C, const double* V,
                        double* dest, const double* src, double scaling)
{
    return sfa::guarded([&] {
        std::string err = sfa::check_csr_rect(rows, columns, A, C);
        if (!err.empty()) throw Error{SFA_ERR_INVALID, err};
        if (!V || !dest || !src) throw Error{SFA_ERR_INVALID, "null pointer"};
        sfa::ensure_device();
        int device = 0;
        SFA_HIP(hipGetDevice(&device));
        ApplyCache& cache = apply_cache();
        std::lock_guard<std::mutex> lock(cache.mu);
        std::unique_ptr<ApplyCache::Entry> fresh;
        const sfa::DevPart& dp = cache.part(device, rows, columns, A, C, V, &fresh);
        double *dsrc = nullptr, *ddst = nullptr;
        try {
            dsrc = sfa::dev_alloc<double>(columns + 1);
            ddst = sfa::dev_alloc<double>(rows);
            SFA_HIP(hipMemcpy(dsrc, src, columns * sizeof(double), hipMemcpyHostToDevice));
            SFA_HIP(hipMemset(dsrc + columns, 0, sizeof(double)));
            sfa::launch_apply(nullptr, dp.view(), dp.nslices, dp.row0(0), dp.nrows_of(0), dsrc, ddst,
                              scaling);
            SFA_HIP(hipGetLastError());
            SFA_HIP(hipMemcpy(dest, ddst, rows * sizeof(double), hipMemcpyDeviceToHost));
        } catch (...) {
            (void)hipStreamSynchronize(nullptr);  // the kernel may still be queued
            sfa::dev_free(dsrc);
            sfa::dev_free(ddst);
            throw;
        }
        sfa::dev_free(dsrc);
        sfa::dev_free(ddst);
    });
}

int sfa_MatrixCSR_getDiagonal(int n, const int* A, const int* C, const double* V, double* diag)
{
    return sfa::guarded([&] {
        std::string err = sfa::check_csr(n, A, C);
        if (!err.empty()) throw Error{SFA_ERR_INVALID, err};
        if (!V || !diag) throw Error{SFA_ERR_INVALID, "null pointer"};
        for (int r = 0; r < n; r++) {
            double s = 0.;
            for (int k = A[r]; k < A[r + 1]; k++)
                if (C[k] == r) s += V[k];
            diag[r] = s;
        }
    });
}

int sfa_Rainbow(int n, const int* A, const int* C, int* F, int* B, int* R, int* num_colors)
{
    return sfa::guarded([&] {
        std::string err = sfa::check_csr(n, A, C);
        if (!err.empty()) throw Error{SFA_ERR_INVALID, err};
        if (!F || !B || !R || !num_colors) throw Error{SFA_ERR_INVALID, "null pointer"};
        *num_colors = sfa::rainbow_colour(n, A, C, F, B, R);
    });
}

int sfa_system_create(int method, int N, const int* csrrows, const int* csrcolumns,
                      const double* csrvalues, const double* aux, double omega, int num_colors,
                      const int* F, const int* B, const int* R, sfa_system* out)
{
    return sfa::guarded([&] {
        if (!out) throw Error{SFA_ERR_INVALID, "null output handle"};
        *out = nullptr;
        std::unique_ptr<sfa_system_s> h(new sfa_system_s());
        const bool fp32 = (method & SFA_FLAG_FP32_INNER) != 0;
        h->sys.pair_sweeps = (method & SFA_FLAG_PAIRED_SWEEPS) != 0;
        h->sys.inner_restarts = (method & SFA_FLAG_INNER_RESTARTS) != 0;
        const bool borrow = (method & SFA_FLAG_BORROW_CSR) != 0;
        h->sys.create(method & ~(SFA_FLAG_FP32_INNER | SFA_FLAG_PAIRED_SWEEPS |
                                 SFA_FLAG_INNER_RESTARTS | SFA_FLAG_BORROW_CSR),
                      N, csrrows, csrcolumns, csrvalues, aux, omega, num_colors, F, B, R, fp32,
                      borrow);
        *out = h.release();
    });
}

int sfa_system_solve(sfa_system sys, double* x, const double* b, double* residual,
                     double precision, int print_modulo, int device_pointers,
                     sfa_solve_info* info)
{
    return sfa::guarded([&] {
        if (!sys || !x || !b || !residual) throw Error{SFA_ERR_INVALID, "null argument"};
        sys->sys.solve(x, b, residual, precision, print_modulo, device_pointers, info);
    });
}

int sfa_system_solve_mixed(sfa_system sys, double* x, const double* b, double rtol,
                           double inner_rtol, int max_outer, int device_pointers,
                           sfa_refine_info* info)
{
    return sfa::guarded([&] {
        if (!sys || !x || !b) throw Error{SFA_ERR_INVALID, "null argument"};
        sys->sys.solve_mixed(x, b, rtol, inner_rtol, max_outer, device_pointers, info);
    });
}

int sfa_system_apply(sfa_system sys, double* dest, const double* src, double scaling,
                     int device_pointers)
{
    return sfa::guarded([&] {
        if (!sys || !dest || !src) throw Error{SFA_ERR_INVALID, "null argument"};
        sys->sys.apply(dest, src, scaling, device_pointers);
    });
}

int sfa_system_set_stream(sfa_system sys, void* hip_stream)
{
    return sfa::guarded([&] {
        if (!sys) throw Error{SFA_ERR_INVALID, "null system"};
        hipStream_t next = hip_stream ? (hipStream_t)hip_stream : sys->sys.own_stream;
        // what is queued on the old stream finishes first: the system's
        // buffers are released after a wait on the stream it runs on
        if (sys->sys.stream && sys->sys.stream != next)
            SFA_HIP(hipStreamSynchronize(sys->sys.stream));
        sys->sys.stream = next;
    });
}

}  // extern "C"

namespace sfa {
void system_row_order_tolerance(sfa_system sys, bool on)
{
    if (sys) sys->sys.row_order_tol = on;
}
}  // namespace sfa

extern "C" {

int sfa_system_query(sfa_system sys, sfa_system_stats* stats)
{
    return sfa::guarded([&] {
        if (!sys || !stats) throw Error{SFA_ERR_INVALID, "null argument"};
        sys->sys.query(stats);
    });
}

int sfa_system_profile(sfa_system sys, double* x, const double* b, double* residual,
                       double precision, int device_pointers, sfa_kernel_profile* out,
                       sfa_solve_info* info)
{
    return sfa::guarded([&] {
        if (!sys || !x || !b || !residual || !out) throw Error{SFA_ERR_INVALID, "null argument"};
        sys->sys.profile(x, b, residual, precision, device_pointers, out, info);
    });
}

int sfa_system_time_sweeps(sfa_system sys, int kernel_class, int reps, double* us_per_launch)
{
    return sfa::guarded([&] {
        if (!sys || !us_per_launch) throw Error{SFA_ERR_INVALID, "null argument"};
        *us_per_launch = sys->sys.time_sweeps(kernel_class, reps);
    });
}

int sfa_system_tune_spmv(sfa_system sys, int fp32_inner, sfa_spmv_tuning* out)
{
    return sfa::guarded([&] {
        if (!sys || !out) throw Error{SFA_ERR_INVALID, "null argument"};
        if (fp32_inner) {
            if (!sys->sys.inner) throw Error{SFA_ERR_INVALID, "the system has no FP32 inner system"};
            sys->sys.inner->tune_spmv(out);
        } else {
            sys->sys.tune_spmv(out);
        }
    });
}

int sfa_system_set_spmv(sfa_system sys, int lanes_per_row)
{
    return sfa::guarded([&] {
        if (!sys) throw Error{SFA_ERR_INVALID, "null system"};
        sys->sys.set_spmv(lanes_per_row);
        if (sys->sys.inner) sys->sys.inner->set_spmv(lanes_per_row);
    });
}

int sfa_system_destroy(sfa_system sys)
{
    return sfa::guarded([&] { delete sys; });
}

}  // extern "C"
