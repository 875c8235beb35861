// hodge.hip -- HodgeConjugateResidualSolverCSR_SSOR on the GPU
// (solver/systemsparsesolver.cpp:426-797), and the same loop as
// HodgeConjugateResidualSolverCSR_diagonal (:52-424) and
// HodgeConjugateResidualSolverCSR_textbook (:798-1121).
//
// Conjugate residuals on the Schur complement (B A^-1 Bt + C) x = b of a
// block system [A Bt; B -C], every A^-1 an inner ConjugateGradientSolverCSR_SSOR
// (natural-order SSOR, omega 1, the diagonal of A) or, for the diagonal
// variant, an inner ConjugateGradientSolverCSR_DiagonalPreconditioner with
// precon = 1 / the diagonal of A (:76-89), or, for the textbook variant, a
// plain ConjugateGradientSolverCSR (its Mr.r equals the Md.r here: Mres =
// Mdir at the restart, :975-977, and both update alike).  A stays resident as an
// sfa_system (its level schedules and SELL parts built once); B, Bt and C are
// SELL-64 parts in their own row order.  Each outer iteration is one product
// Bt Mdir, one inner solve, and two fused row kernels:
//
//   k_hodge_rows     vil = B aux2 + C Mdir (B's entries, then C's, in CSR
//                    order: the reference's per-row sum), then the x, res,
//                    Mres updates and the block partials of Mres.res;
//   k_hodge_dir      dir = res + beta dir, Mdir = Mres + beta Mdir and the
//                    block partials of Mdir.Mdir.
//
// The scalars (alpha, beta, the exit tests) run on the host from the block
// partials, summed in block order; the host already waits for every inner
// solve, so this costs no extra round trip per iteration beyond two small
// copies.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
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

constexpr int kHodgeWaves = 4;                    // slices (wavefronts) per block
constexpr int kHodgeBlock = kHodgeWaves * kWave;

enum HodgeRows : int {
    HR_RESIDUAL = 0,  // res = b - B aux2 - C x (term by term), dir = res     (:556-570)
    HR_PRODUCT = 1,   // Mdir = B aux2 + C dir, Mres = Mdir; Mdir.res, Mdir.Mdir (:600-619)
    HR_UPDATE = 2     // vil = B aux2 + C Mdir; x, res, Mres updates; Mres.res   (:691-721)
};

__device__ inline double block_sum(double v, double* ws)
{
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    v = wave_sum(v);
    if (lane == 0) ws[w] = v;
    __syncthreads();
    double t = 0.;
    for (int i = 0; i < kHodgeWaves; i++) t += ws[i];
    return t;
}

// One row per lane: row = slice*64 + lane of B and C (both in row order).
__global__ void __launch_bounds__(kHodgeBlock) k_hodge_rows(
    int mode, SellViewT<double> B, SellViewT<double> C, int N, const double* aux2,
    const double* cvec, const double* b, double* x, double* res, double* dir, double* Mdir,
    double* Mres, double alpha, double* part0, double* part1)
{
    __shared__ double ws[kHodgeWaves];
    const int lane = threadIdx.x % kWave;
    const int slice = blockIdx.x * kHodgeWaves + threadIdx.x / kWave;
    const int row = slice * kWave + lane;
    double p0 = 0., p1 = 0.;
    if (row < N) {
        const int b0 = B.off[slice] + lane, b1 = B.off[slice + 1];
        const int c0 = C.off[slice] + lane, c1 = C.off[slice + 1];
        if (mode == HR_RESIDUAL) {
            double r = b[row];
            for (int e = b0; e < b1; e += kWave) r -= 1. * B.vals[e] * aux2[B.cols[e]];
            for (int e = c0; e < c1; e += kWave) r -= C.vals[e] * cvec[C.cols[e]];
            res[row] = r;
            dir[row] = r;
        } else {
            double acc = 0.;
            for (int e = b0; e < b1; e += kWave) acc += 1. * B.vals[e] * aux2[B.cols[e]];
            for (int e = c0; e < c1; e += kWave) acc += C.vals[e] * cvec[C.cols[e]];
            if (mode == HR_PRODUCT) {
                Mdir[row] = acc;
                Mres[row] = acc;
                p0 = acc * res[row];
                p1 = acc * acc;
            } else {
                x[row] = x[row] + alpha * dir[row];
                const double rr = res[row] - alpha * Mdir[row];
                res[row] = rr;
                const double mr = Mres[row] - alpha * acc;
                Mres[row] = mr;
                p0 = mr * rr;
            }
        }
    }
    if (mode == HR_RESIDUAL) return;
    const double s0 = block_sum(p0, ws);
    if (threadIdx.x == 0) part0[blockIdx.x] = s0;
    if (mode == HR_PRODUCT) {
        __syncthreads();
        const double s1 = block_sum(p1, ws);
        if (threadIdx.x == 0) part1[blockIdx.x] = s1;
    }
}

// dir = res + beta dir; Mdir = Mres + beta Mdir; block partials of Mdir.Mdir (:725-739).
__global__ void __launch_bounds__(kHodgeBlock) k_hodge_dir(int N, double beta, const double* res,
                                                           const double* Mres, double* dir,
                                                           double* Mdir, double* part)
{
    __shared__ double ws[kHodgeWaves];
    const int i = blockIdx.x * kHodgeBlock + threadIdx.x;
    double p = 0.;
    if (i < N) {
        dir[i] = res[i] + beta * dir[i];
        const double m = Mres[i] + beta * Mdir[i];
        Mdir[i] = m;
        p = m * m;
    }
    const double s = block_sum(p, ws);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

int cdiv(int a, int b) { return (a + b - 1) / b; }

struct HodgeSolver {
    int N = 0, L = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    sfa_system inner = nullptr;       // A, natural-order SSOR or Jacobi CG, resident
    int inner_kind = SFA_HODGE_INNER_SSOR;
    DevPart Bp, Btp, Cp;              // B (N x L), Bt (L x N), C (N x N)
    double *x = nullptr, *b = nullptr, *res = nullptr, *dir = nullptr, *Mdir = nullptr,
           *Mres = nullptr, *aux1 = nullptr, *aux2 = nullptr, *auxR = nullptr;
    double *part0 = nullptr, *part1 = nullptr;  // block partials
    double* hpart = nullptr;                    // pinned copies of them
    int nblocks = 0;

    ~HodgeSolver() { release(); }

    void release()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        if (inner) sfa_system_destroy(inner);
        inner = nullptr;
        Bp.release();
        Btp.release();
        Cp.release();
        for (void* p : {(void*)x, (void*)b, (void*)res, (void*)dir, (void*)Mdir, (void*)Mres,
                        (void*)aux1, (void*)aux2, (void*)auxR, (void*)part0, (void*)part1})
            dev_free(p);
        x = b = res = dir = Mdir = Mres = aux1 = aux2 = auxR = part0 = part1 = nullptr;
        pinned_free(hpart);
        hpart = nullptr;
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }

    void create(int N_, int L_, const int* Ar, const int* Ac, const double* Av, const int* Br,
                const int* Bc, const double* Bv, const int* Btr, const int* Btc,
                const double* Btv, const int* Cr, const int* Cc, const double* Cv,
                int inner_choice)
    {
        if (inner_choice != SFA_HODGE_INNER_SSOR && inner_choice != SFA_HODGE_INNER_DIAGONAL &&
            inner_choice != SFA_HODGE_INNER_TEXTBOOK)
            throw Error{SFA_ERR_INVALID,
                        "inner must be SFA_HODGE_INNER_SSOR, _DIAGONAL or _TEXTBOOK"};
        if (N_ <= 0 || L_ <= 0) throw Error{SFA_ERR_INVALID, "N and L must be positive"};
        // A block without entries (the zero matrix C of the reference's Darcy
        // tests) may come with null column and value arrays.
        auto block = [](int rows, int cols, const int* r, const int* c, const double* v,
                        const char* name) {
            if (!r) throw Error{SFA_ERR_INVALID, std::string("null row offsets of ") + name};
            if (r[rows] == 0 && r[0] == 0) {
                for (int i = 0; i < rows; i++)
                    if (r[i + 1] != 0) throw Error{SFA_ERR_INVALID, "csrrows must be non-decreasing"};
                return;
            }
            if (!v) throw Error{SFA_ERR_INVALID, std::string("null values of ") + name};
            const std::string e = check_csr_rect(rows, cols, r, c);
            if (!e.empty()) throw Error{SFA_ERR_INVALID, std::string(name) + ": " + e};
        };
        block(L_, L_, Ar, Ac, Av, "A");
        block(N_, L_, Br, Bc, Bv, "B");
        block(L_, N_, Btr, Btc, Btv, "Bt");
        block(N_, N_, Cr, Cc, Cv, "C");
        N = N_;
        L = L_;
        inner_kind = inner_choice;
        const bool diagonal_variant = inner_choice == SFA_HODGE_INNER_DIAGONAL;
        // the diagonal of A as the reference sums it (:498-511, :76-83); the
        // diagonal variant inverts it (:85, no guard: the reference asserts > 0)
        std::vector<double> diag(L, 0.);
        for (int c = 0; c < L; c++) {
            for (int d = Ar[c]; d < Ar[c + 1]; d++)
                if (Ac[d] == c) diag[c] += Av[d];
            if (diagonal_variant) diag[c] = 1. / diag[c];
        }
        ensure_device();
        SFA_HIP(hipGetDevice(&device));
        SFA_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        const int method = inner_choice == SFA_HODGE_INNER_SSOR       ? SFA_METHOD_SSOR
                           : inner_choice == SFA_HODGE_INNER_DIAGONAL ? SFA_METHOD_DIAGONAL
                                                               : SFA_METHOD_CG;
        check(sfa_system_create(method, L, Ar, Ac, Av,
                                method == SFA_METHOD_CG ? nullptr : diag.data(), 1.0, 0, nullptr,
                                nullptr, nullptr, &inner));
        check(sfa_system_set_stream(inner, stream));
        // each inner tolerance from ||Bt v|| summed in row order, as the
        // reference's inner call sums it (solver/sparsesolver.cpp:469-472)
        system_row_order_tolerance(inner, true);
        SellPart part;
        build_full_identity(part, N, L, Br, Bc, Bv);
        Bp.upload(part);
        build_full_identity(part, L, N, Btr, Btc, Btv);
        Btp.upload(part);
        build_full_identity(part, N, N, Cr, Cc, Cv);
        Cp.upload(part);
        // x, dir and Mdir are gathered through Bt and C, aux2 through B:
        // each has its zero slot (index N or L) for the padding entries
        for (double** v : {&x, &dir, &Mdir}) *v = dev_alloc<double>(N + 1);
        for (double** v : {&b, &res, &Mres}) *v = dev_alloc<double>(N);
        aux2 = dev_alloc<double>(L + 1);
        aux1 = dev_alloc<double>(L);
        auxR = dev_alloc<double>(L);
        nblocks = cdiv(cdiv(N, kWave), kHodgeWaves);
        part0 = dev_alloc<double>(nblocks);
        part1 = dev_alloc<double>(nblocks);
        hpart = pinned_alloc<double>(2 * (size_t)nblocks);
        for (double* v : {x, dir, Mdir}) launch_fill(stream, v, 0.0, N + 1);
        launch_fill(stream, aux2, 0.0, L + 1);
        SFA_HIP(hipStreamSynchronize(stream));
    }

    static void check(int status)
    {
        if (status != SFA_OK) throw Error{status, sfa_last_error()};
    }

    // aux2 = A^-1 (Bt src) by the inner CG, from aux2 (zero it first
    // for a cold start).
    void inner_solve(const double* src, bool cold, double precision, int print_modulo)
    {
        launch_apply(stream, Btp.view(), Btp.nslices, 0, L, src, aux1, 1.0, 1);
        if (cold) launch_fill(stream, aux2, 0.0, L);
        sfa_solve_info info{};
        check(sfa_system_solve(inner, aux2, aux1, auxR, precision, print_modulo, 1, &info));
    }

    void rows(int mode, const double* cvec, double alpha)
    {
        const int nsl = Bp.nslices;
        launch(k_hodge_rows, dim3(cdiv(nsl, kHodgeWaves)), dim3(kHodgeBlock), stream, mode,
               Bp.view(), Cp.view(), N, (const double*)aux2, cvec, (const double*)b, x, res, dir,
               Mdir, Mres, alpha, part0, part1);
    }

    // Sums of the block partials in block order (one or two arrays).
    void sums(int count, double* s0, double* s1)
    {
        SFA_HIP(hipMemcpyAsync(hpart, part0, nblocks * sizeof(double), hipMemcpyDeviceToHost, stream));
        if (count > 1)
            SFA_HIP(hipMemcpyAsync(hpart + nblocks, part1, nblocks * sizeof(double),
                                   hipMemcpyDeviceToHost, stream));
        SFA_HIP(hipStreamSynchronize(stream));
        *s0 = 0.;
        for (int i = 0; i < nblocks; i++) *s0 += hpart[i];
        if (count > 1) {
            *s1 = 0.;
            for (int i = 0; i < nblocks; i++) *s1 += hpart[nblocks + i];
        }
    }

    void solve(double* ux, const double* ub, double* ures, double precision, int print_modulo,
               double inner_precision, int inner_print_modulo, int device_pointers,
               sfa_solve_info* info)
    {
        if (!(precision > 0)) throw Error{SFA_ERR_INVALID, "precision must be positive"};
        SFA_HIP(hipSetDevice(device));
        const hipMemcpyKind in = device_pointers ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        if (device_pointers) {
            // inputs written on the null stream, which this non-blocking
            // stream does not wait for by itself
            hipEvent_t ev = nullptr;
            SFA_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
            const hipError_t e1 = hipEventRecord(ev, nullptr);
            const hipError_t e2 = e1 == hipSuccess ? hipStreamWaitEvent(stream, ev, 0) : e1;
            (void)hipEventDestroy(ev);
            SFA_HIP(e2);
        }
        SFA_HIP(hipMemcpyAsync(x, ux, N * sizeof(double), in, stream));
        SFA_HIP(hipMemcpyAsync(b, ub, N * sizeof(double), in, stream));
        // tolerance from ||b|| (:457-461), summed in row order on the host
        std::vector<double> hb(N);
        SFA_HIP(hipMemcpyAsync(hb.data(), b, N * sizeof(double), hipMemcpyDeviceToHost, stream));
        SFA_HIP(hipStreamSynchronize(stream));
        double tol = 0.;
        for (int i = 0; i < N; i++) tol += hb[i] * hb[i];
        tol = precision * std::sqrt(tol);
        if (!(tol >= kDesiredPrecision) && tol <= kDesiredPrecision) tol = kDesiredPrecision;

        // the diagonal and textbook variants gate their breakdown, warning and
        // finish lines by print_modulo (:91-96, :280-286, :375); the SSOR one
        // prints the first two always and the last from print_modulo >= 0
        // (:465-470, :746)
        const bool dv = inner_kind != SFA_HODGE_INNER_SSOR;
        const bool print_hard = !dv || print_modulo >= -2;
        const bool print_finish = dv ? print_modulo >= -1 : print_modulo >= 0;
        if (print_modulo >= -1)
            std::printf("(%d/%d)     BEGIN: Hodge Conjugate Residual (%s) CSR\n", 0, N,
                        inner_kind == SFA_HODGE_INNER_SSOR       ? "SSOR"
                        : inner_kind == SFA_HODGE_INNER_DIAGONAL ? "diag"
                                                                 : "textbook");
        double Md_r = NAN, Md_Md = NAN;
        int K = 0, reason = SFA_EXIT_CAP;
        const double tol2 = tol * tol;
        while (K < N) {
            if (K == 0) {  // :529-626
                inner_solve(x, true, inner_precision, inner_print_modulo);
                rows(HR_RESIDUAL, x, 0.);
                inner_solve(dir, true, inner_precision, inner_print_modulo);
                rows(HR_PRODUCT, dir, 0.);
                sums(2, &Md_r, &Md_Md);
                if (print_modulo >= 0)
                    std::printf("(%d/%d)   RESTART: Residual norm is %.9le < %.9le\n", K, N,
                                std::sqrt(Md_r), tol);
            }
            if (print_modulo > 0 && K % print_modulo == 0)
                std::printf("(%d/%d)   INTERIM: Residual norm is %.9le < %.9le\n", K, N,
                            std::sqrt(Md_r), tol);
            if (!std::isfinite(Md_r) || Md_r < 0.) {  // :632-637
                if (print_modulo >= -2)
                    std::printf("(%d/%d) BREAKDOWN: Residual energy is unreasonable with %.9le\n",
                                K, N, Md_r);
                reason = SFA_EXIT_BREAKDOWN;
                break;
            }
            if (std::fabs(Md_r) < tol2) {  // :639-642
                reason = SFA_EXIT_CONVERGED;
                break;
            }
            if (!std::isfinite(Md_Md) || Md_Md < 0.) {  // :646-652
                if (print_hard)
                    std::printf("(%d/%d) BREAKDOWN: Gradient double energy is unreasonable with %.9le\n",
                                K, N, Md_Md);
                reason = SFA_EXIT_BREAKDOWN;
                break;
            }
            if (std::sqrt(std::fabs(Md_Md)) < kMachineEpsilon) {  // :654-658
                if (print_hard) {
                    std::printf("(%d/%d)   WARNING: Residual norm is %.9le < %.9le\n", K, N,
                                std::sqrt(Md_r), tol);
                    std::printf("(%d/%d)   WARNING: Gradient double energy is small with %.9le\n",
                                K, N, Md_Md);
                }
                reason = SFA_EXIT_SMALL;
                break;
            }
            const double alpha = Md_r / Md_Md;  // :663
            inner_solve(Mdir, false, inner_precision, inner_print_modulo);  // warm start, :679-689
            rows(HR_UPDATE, Mdir, alpha);
            double new_Mr_r = 0., unused = 0.;
            sums(1, &new_Mr_r, &unused);
            const double beta = new_Mr_r / Md_r;  // :723
            launch(k_hodge_dir, dim3(nblocks), dim3(kHodgeBlock), stream, N, beta,
                   (const double*)res, (const double*)Mres, dir, Mdir, part0);
            sums(1, &Md_Md, &unused);
            Md_r = new_Mr_r;
            K++;
        }
        if (print_finish)
            std::printf("(%d/%d) %9s: Residual norm is %.9le < %.9le\n", K, N,
                        std::sqrt(Md_r) < tol ? "SUCCESS" : "FAILED", std::sqrt(Md_r), tol);
        std::fflush(stdout);
        const hipMemcpyKind out = device_pointers ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
        SFA_HIP(hipMemcpyAsync(ux, x, N * sizeof(double), out, stream));
        SFA_HIP(hipMemcpyAsync(ures, res, N * sizeof(double), out, stream));
        SFA_HIP(hipStreamSynchronize(stream));
        SFA_HIP(hipGetLastError());
        if (info) {
            sfa_solve_info o{};
            o.iterations = K;
            o.zr = Md_r;
            o.tolerance = tol;
            o.exit_reason = reason;
            o.success = std::sqrt(Md_r) < tol ? 1 : 0;
            *info = o;
        }
    }
};

}  // namespace
}  // namespace sfa

struct sfa_hodge_s {
    sfa::HodgeSolver h;
};

static int hodge_once(int inner, int N, int L, double* x, const double* b, const int* Arows,
               const int* Acolumns, const double* Avalues, const int* Brows, const int* Bcolumns,
               const double* Bvalues, const int* Btrows, const int* Btcolumns,
               const double* Btvalues, const int* Crows, const int* Ccolumns,
               const double* Cvalues, double* residual, double precision, int print_modulo,
               double inneriteration_precision, int inneriteration_print_modulo,
               sfa_solve_info* info)
{
    return sfa::guarded([&] {
        if (!x || !b || !residual) throw sfa::Error{SFA_ERR_INVALID, "null x, b or residual"};
        sfa::HodgeSolver h;
        h.create(N, L, Arows, Acolumns, Avalues, Brows, Bcolumns, Bvalues, Btrows, Btcolumns,
                 Btvalues, Crows, Ccolumns, Cvalues, inner);
        h.solve(x, b, residual, precision, print_modulo, inneriteration_precision,
                inneriteration_print_modulo, 0, info);
    });
}

extern "C" {

int sfa_hodge_create_inner(int N, int L, const int* Arows, const int* Acolumns,
                           const double* Avalues, const int* Brows, const int* Bcolumns,
                           const double* Bvalues, const int* Btrows, const int* Btcolumns,
                           const double* Btvalues, const int* Crows, const int* Ccolumns,
                           const double* Cvalues, int inner, sfa_hodge* out)
{
    return sfa::guarded([&] {
        if (!out) throw sfa::Error{SFA_ERR_INVALID, "null output handle"};
        *out = nullptr;
        std::unique_ptr<sfa_hodge_s> h(new sfa_hodge_s());
        h->h.create(N, L, Arows, Acolumns, Avalues, Brows, Bcolumns, Bvalues, Btrows, Btcolumns,
                    Btvalues, Crows, Ccolumns, Cvalues, inner);
        *out = h.release();
    });
}

int sfa_hodge_create(int N, int L, const int* Arows, const int* Acolumns, const double* Avalues,
                     const int* Brows, const int* Bcolumns, const double* Bvalues,
                     const int* Btrows, const int* Btcolumns, const double* Btvalues,
                     const int* Crows, const int* Ccolumns, const double* Cvalues, sfa_hodge* out)
{
    return sfa_hodge_create_inner(N, L, Arows, Acolumns, Avalues, Brows, Bcolumns, Bvalues,
                                  Btrows, Btcolumns, Btvalues, Crows, Ccolumns, Cvalues,
                                  SFA_HODGE_INNER_SSOR, out);
}

int sfa_hodge_solve(sfa_hodge h, double* x, const double* b, double* residual, double precision,
                    int print_modulo, double inneriteration_precision,
                    int inneriteration_print_modulo, int device_pointers, sfa_solve_info* info)
{
    return sfa::guarded([&] {
        if (!h || !x || !b || !residual) throw sfa::Error{SFA_ERR_INVALID, "null argument"};
        h->h.solve(x, b, residual, precision, print_modulo, inneriteration_precision,
                   inneriteration_print_modulo, device_pointers, info);
    });
}

int sfa_hodge_destroy(sfa_hodge h)
{
    delete h;
    return SFA_OK;
}


int sfa_HodgeConjugateResidualSolverCSR_SSOR(
    int N, int L, double* x, const double* b, const int* Arows, const int* Acolumns,
    const double* Avalues, const int* Brows, const int* Bcolumns, const double* Bvalues,
    const int* Btrows, const int* Btcolumns, const double* Btvalues, const int* Crows,
    const int* Ccolumns, const double* Cvalues, double* residual, double precision,
    int print_modulo, double inneriteration_precision, int inneriteration_print_modulo,
    sfa_solve_info* info)
{
    return hodge_once(SFA_HODGE_INNER_SSOR, N, L, x, b, Arows, Acolumns, Avalues, Brows,
                      Bcolumns, Bvalues, Btrows, Btcolumns, Btvalues, Crows, Ccolumns, Cvalues,
                      residual, precision, print_modulo, inneriteration_precision,
                      inneriteration_print_modulo, info);
}

int sfa_HodgeConjugateResidualSolverCSR(
    int N, int L, double* x, const double* b, const int* Arows, const int* Acolumns,
    const double* Avalues, const int* Brows, const int* Bcolumns, const double* Bvalues,
    const int* Btrows, const int* Btcolumns, const double* Btvalues, const int* Crows,
    const int* Ccolumns, const double* Cvalues, double* residual, double precision,
    int print_modulo, double inneriteration_precision, int inneriteration_print_modulo,
    sfa_solve_info* info)
{
    // the reference's plain variant is the SSOR one (solver/systemsparsesolver.cpp:18-49)
    return sfa_HodgeConjugateResidualSolverCSR_SSOR(
        N, L, x, b, Arows, Acolumns, Avalues, Brows, Bcolumns, Bvalues, Btrows, Btcolumns,
        Btvalues, Crows, Ccolumns, Cvalues, residual, precision, print_modulo,
        inneriteration_precision, inneriteration_print_modulo, info);
}

int sfa_HodgeConjugateResidualSolverCSR_textbook(
    int N, int L, double* x, const double* b, const int* Arows, const int* Acolumns,
    const double* Avalues, const int* Brows, const int* Bcolumns, const double* Bvalues,
    const int* Btrows, const int* Btcolumns, const double* Btvalues, const int* Crows,
    const int* Ccolumns, const double* Cvalues, double* residual, double precision,
    int print_modulo, double inneriteration_precision, int inneriteration_print_modulo,
    sfa_solve_info* info)
{
    return hodge_once(SFA_HODGE_INNER_TEXTBOOK, N, L, x, b, Arows, Acolumns, Avalues, Brows,
                      Bcolumns, Bvalues, Btrows, Btcolumns, Btvalues, Crows, Ccolumns, Cvalues,
                      residual, precision, print_modulo, inneriteration_precision,
                      inneriteration_print_modulo, info);
}

int sfa_HodgeConjugateResidualSolverCSR_diagonal(
    int N, int L, double* x, const double* b, const int* Arows, const int* Acolumns,
    const double* Avalues, const int* Brows, const int* Bcolumns, const double* Bvalues,
    const int* Btrows, const int* Btcolumns, const double* Btvalues, const int* Crows,
    const int* Ccolumns, const double* Cvalues, double* residual, double precision,
    int print_modulo, double inneriteration_precision, int inneriteration_print_modulo,
    sfa_solve_info* info)
{
    return hodge_once(SFA_HODGE_INNER_DIAGONAL, N, L, x, b, Arows, Acolumns, Avalues, Brows,
                      Bcolumns, Bvalues, Btrows, Btcolumns, Btvalues, Crows, Ccolumns, Cvalues,
                      residual, precision, print_modulo, inneriteration_precision,
                      inneriteration_print_modulo, info);
}

}  // extern "C"
