// noalloc.hip -- Rainbow CG-SSOR on caller-owned device CSR arrays that
// allocates no device memory.
//
// sfa_ConjugateGradientSolverCSR_Rainbow_noalloc runs the iteration of
// ConjugateGradientSolverCSR_Rainbow (solver/sparsesolver.cpp:1002-1335) on
// the caller's CSR arrays and colouring as they lie in HBM (original row
// order, one thread per row): no device-order plan, no SELL parts, no graph.
// Its Krylov vectors (direction, z, m, q), the z.r history, the reduction
// partials and the solve's scalars live in `scratch`, a static device array
// of fixed size that each GPU holds from the first call until the process
// ends.  Host code never passes `scratch` itself to a kernel (in host code
// the name is a stand-in, not a device address): the first call on a device
// looks up the array's address on that device with hipGetSymbolAddress and
// every launch uses that address.
//
// All calls on one GPU share its array, and each call leaves its contents
// behind: every call fills what must start at zero.  Calls that overlap on
// one GPU race and must not be made.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <mutex>

#include "../../include/simplefem_amd.h"
#include "device_common.hpp"
#include "host_common.hpp"

namespace sfa {
namespace noalloc {

constexpr int kBlock = 256;
constexpr int kMaxRows = SFA_NOALLOC_MAX_ROWS;
constexpr int kMaxColours = SFA_NOALLOC_MAX_COLORS;
constexpr int kMaxDevices = 64;

// Layout of the scratch array, in doubles.  Partials: one per block of 256
// rows, for the whole-vector kernels (region A) and for the backward sweep's
// per-colour blocks (region B, at most one extra partial per colour).
constexpr size_t kStateDoubles = 64;
constexpr size_t kPartsA = (size_t)kMaxRows / kBlock;
constexpr size_t kPartsB = (size_t)kMaxRows / kBlock + kMaxColours;
constexpr size_t kColourTable = (kMaxColours + 2) / 2;  // kMaxColours + 1 ints
constexpr size_t kOffState = 0;
constexpr size_t kOffPartsA = kOffState + kStateDoubles;
constexpr size_t kOffPartsB = kOffPartsA + kPartsA;
constexpr size_t kOffTable = kOffPartsB + kPartsB;
constexpr size_t kOffVectors = kOffTable + kColourTable;
constexpr size_t kVectors = 5;  // direction, z, m, q (also the validation marks), history
constexpr size_t kScratchDoubles = kOffVectors + kVectors * ((size_t)kMaxRows + 1);
static_assert(kScratchDoubles * sizeof(double) <= (size_t)SFA_NOALLOC_SCRATCH_BYTES,
              "the scratch layout must fit SFA_NOALLOC_SCRATCH_BYTES");

}  // namespace noalloc
}  // namespace sfa

// The solver's whole device workspace (see the layout above).
static __device__ __attribute__((aligned(256))) double scratch[SFA_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace sfa {
namespace noalloc {

struct NaState {
    double tol, tol2, zr, dAd, alpha, beta;
    int K, done, exit_reason, bad_csr, bad_colouring, reserved;
};
static_assert(sizeof(NaState) <= kStateDoubles * sizeof(double), "state slot");

struct Work {
    NaState* st;
    double* pa;   // partials of whole-vector kernels
    double* pb;   // partials of the backward sweep, per colour
    int* coff;    // first partial of each colour in pb (num_colors + 1)
    double* d;    // direction
    double* z;    // zirconium
    double* m;    // mittlerer
    double* q;    // auxiliary; during validation: the position of each row in R
    double* hist; // z.r after the restart and after every iteration
};

Work carve(double* base)
{
    Work w;
    w.st = reinterpret_cast<NaState*>(base + kOffState);
    w.pa = base + kOffPartsA;
    w.pb = base + kOffPartsB;
    w.coff = reinterpret_cast<int*>(base + kOffTable);
    double* v = base + kOffVectors;
    const size_t stride = (size_t)kMaxRows + 1;
    w.d = v;
    w.z = v + stride;
    w.m = v + 2 * stride;
    w.q = v + 3 * stride;
    w.hist = v + 4 * stride;
    return w;
}

using dev::block_partial;
using dev::kDesiredPrecision;
using dev::kMachineEpsilon;
using dev::ref_absolute;
using dev::wave_fold;

__device__ inline int row_id() { return blockIdx.x * kBlock + threadIdx.x; }

// ---- validation: the CSR structure and the colouring, read where they lie

__global__ void __launch_bounds__(kBlock) k_check_csr(int N, const int* __restrict__ A,
                                                     const int* __restrict__ C, NaState* st)
{
    const int i = row_id();
    if (i >= N) return;
    const int nnz = A[N];
    const int a = A[i], e = A[i + 1];
    if (a < 0 || e < a || e > nnz) {
        st->bad_csr = 1;
        return;
    }
    for (int k = a; k < e; k++) {
        const int c = C[k];
        if (c < 0 || c >= N) st->bad_csr = 1;
    }
}

__global__ void __launch_bounds__(kBlock) k_fill_int(int* v, int value, int n)
{
    const int i = row_id();
    if (i < n) v[i] = value;
}

// Position i of R lies in colour c with B[c] <= i < B[c+1]; its row must be
// in range and carry colour c.  pos[row] = i.
__global__ void __launch_bounds__(kBlock) k_check_rows(int N, int nc, const int* __restrict__ F,
                                                      const int* __restrict__ B,
                                                      const int* __restrict__ R, int* pos,
                                                      NaState* st)
{
    const int i = row_id();
    if (i >= N) return;
    int lo = 0, hi = nc;  // last c with B[c] <= i
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (B[mid] <= i) lo = mid;
        else hi = mid;
    }
    const int r = R[i];
    if (r < 0 || r >= N) {
        st->bad_colouring = 1;
        return;
    }
    if (F[r] != lo) st->bad_colouring = 1;
    pos[r] = i;
}

// With pos filled from -1: R is a permutation iff every row has a position
// whose entry names it (a repeated row leaves another row without one).
__global__ void __launch_bounds__(kBlock) k_check_permutation(int N, const int* __restrict__ R,
                                                             const int* __restrict__ pos,
                                                             NaState* st)
{
    const int r = row_id();
    if (r >= N) return;
    const int p = pos[r];
    if (p < 0 || p >= N || R[p] != r) st->bad_colouring = 1;
}

// ---- restart

__global__ void __launch_bounds__(kBlock) k_zero(Work w, int n)
{
    const int i = row_id();
    if (i >= n) return;
    w.d[i] = 0.;  // solver/sparsesolver.cpp:1046-1057: zero-initialised scratch
    w.z[i] = 0.;
    w.m[i] = 0.;
    w.q[i] = 0.;
}

// partials of b.b over all rows (the tolerance, :1030-1033)
__global__ void __launch_bounds__(kBlock) k_sumsq(int N, const double* __restrict__ b, Work w)
{
    const int i = row_id();
    const double v = i < N ? b[i] * b[i] : 0.;
    block_partial<kBlock>(v, w.pa + blockIdx.x);
}

__global__ void __launch_bounds__(dev::kWave) k_start(int nparts, double precision, Work w)
{
    const double s = wave_fold(w.pa, nparts);
    if (threadIdx.x != 0) return;
    NaState* st = w.st;
    double t = precision * sqrt(s);
    if (!(t >= kDesiredPrecision) && t <= kDesiredPrecision) t = kDesiredPrecision;  // Maximum
    st->tol = t;
    st->tol2 = t * t;
    st->zr = NAN;
    st->dAd = 0.;
    st->alpha = 0.;
    st->beta = 0.;
    st->K = 0;
    st->done = 0;
    st->exit_reason = SFA_EXIT_CAP;
}

// r = b - A x on rows with diagonal != 0 (:1075-1087)
__global__ void __launch_bounds__(kBlock) k_residual(int N, const int* __restrict__ A,
                                                    const int* __restrict__ C,
                                                    const double* __restrict__ V,
                                                    const double* __restrict__ diag,
                                                    const double* __restrict__ b,
                                                    const double* __restrict__ x, double* r)
{
    const int c = row_id();
    if (c >= N || diag[c] == 0.) return;
    double acc = b[c];
    for (int k = A[c]; k < A[c + 1]; k++) acc -= V[k] * x[C[k]];
    r[c] = acc;
}

// One colour of the forward sweep (:1091-1116 at the restart, :1234-1261 in
// the iteration, where the row's x and r updates come first).
template <bool UPDATE>
__global__ void __launch_bounds__(kBlock) k_forward(int row_begin, int nrows,
                                                   const int* __restrict__ A,
                                                   const int* __restrict__ C,
                                                   const double* __restrict__ V,
                                                   const double* __restrict__ diag,
                                                   const int* __restrict__ F,
                                                   const int* __restrict__ R, double omega,
                                                   Work w, double* x, double* r)
{
    if (w.st->done) return;
    const int i = row_id();
    if (i >= nrows) return;
    const int c = R[row_begin + i];
    const double dc = diag[c];
    if (dc == 0.) return;
    double aux;
    if constexpr (UPDATE) {
        const double alpha = w.st->alpha;
        x[c] += alpha * w.d[c];
        aux = r[c] - alpha * w.q[c];
        r[c] = aux;
    } else {
        aux = r[c];
    }
    const int fc = F[c];
    for (int k = A[c]; k < A[c + 1]; k++) {
        const int j = C[k];
        if (F[j] < fc) aux -= V[k] * w.m[j];
    }
    w.m[c] = aux * omega / dc;
}

// One colour of the backward sweep (:1130-1154 at the restart, after the
// scaling of :1118-1128; :1263-1303 in the iteration): m *= diag,
// m *= (2 - omega) / omega, z = (m - U z) omega / diag, and the block's part
// of this colour's z.r.
__global__ void __launch_bounds__(kBlock) k_backward(int row_begin, int nrows, int part0,
                                                    const int* __restrict__ A,
                                                    const int* __restrict__ C,
                                                    const double* __restrict__ V,
                                                    const double* __restrict__ diag,
                                                    const int* __restrict__ F,
                                                    const int* __restrict__ R, double omega,
                                                    Work w, const double* __restrict__ r)
{
    if (w.st->done) return;
    const int i = row_id();
    double zr = 0.;
    if (i < nrows) {
        const int c = R[row_begin + i];
        const double dc = diag[c];
        if (dc != 0.) {
            double mc = w.m[c];
            mc *= dc;
            mc *= (2. - omega) / omega;
            w.m[c] = mc;
            double aux = mc;
            const int fc = F[c];
            for (int k = A[c]; k < A[c + 1]; k++) {
                const int j = C[k];
                if (F[j] > fc) aux -= V[k] * w.z[j];
            }
            const double zc = aux * omega / dc;
            w.z[c] = zc;
            zr = zc * r[c];
        }
    }
    block_partial<kBlock>(zr, w.pb + part0 + blockIdx.x);
}

// d = z and the block partials of z.r over active rows in row order (:1156-1169)
__global__ void __launch_bounds__(kBlock) k_direction_init(int N, const double* __restrict__ diag,
                                                          const double* __restrict__ r, Work w)
{
    if (w.st->done) return;
    const int c = row_id();
    double v = 0.;
    if (c < N && diag[c] != 0.) {
        const double zc = w.z[c];
        w.d[c] = zc;
        v = zc * r[c];
    }
    block_partial<kBlock>(v, w.pa + blockIdx.x);
}

// z.r of the restart and the convergence test at the top of the loop (:1187-1190)
__global__ void __launch_bounds__(dev::kWave) k_restart_zr(int nparts, Work w)
{
    const double zr = wave_fold(w.pa, nparts);
    if (threadIdx.x != 0) return;
    NaState* st = w.st;
    st->zr = zr;
    w.hist[0] = zr;
    if (ref_absolute(zr) < st->tol2) {
        st->exit_reason = SFA_EXIT_CONVERGED;
        st->done = 1;
    }
}

// ---- iteration

// q = A d on active rows (0 elsewhere) and the block partials of d.q (:1197-1212)
__global__ void __launch_bounds__(kBlock) k_spmv(int N, const int* __restrict__ A,
                                                const int* __restrict__ C,
                                                const double* __restrict__ V,
                                                const double* __restrict__ diag, Work w)
{
    if (w.st->done) return;
    const int c = row_id();
    double v = 0.;
    if (c < N) {
        double acc = 0.;
        if (diag[c] != 0.) {
            for (int k = A[c]; k < A[c + 1]; k++) acc += V[k] * w.d[C[k]];
            v = w.d[c] * acc;
        }
        w.q[c] = acc;
    }
    block_partial<kBlock>(v, w.pa + blockIdx.x);
}

// d.Ad, the breakdown tests (:1215-1228) and alpha (:1231)
__global__ void __launch_bounds__(dev::kWave) k_alpha(int nparts, Work w)
{
    NaState* st = w.st;
    if (st->done) return;
    const double dAd = wave_fold(w.pa, nparts);
    if (threadIdx.x != 0) return;
    const double zr = st->zr;
    st->dAd = dAd;
    const bool num_bad = !isfinite(zr) || zr < 0.;
    const bool den_bad = !isfinite(dAd) || dAd < 0.;
    if (num_bad || den_bad) {
        st->exit_reason = SFA_EXIT_BREAKDOWN;
        st->done = 1;
        return;
    }
    if (sqrt(ref_absolute(dAd)) < kMachineEpsilon) {
        st->exit_reason = SFA_EXIT_SMALL;
        st->done = 1;
        return;
    }
    st->alpha = zr / dAd;
}

// z.r: each colour's partials summed, then the colours from the last to the
// first (:1263-1303); beta (:1305-1307); K++ and the loop's exit tests.
constexpr int kBetaBlock = 1024;
__global__ void __launch_bounds__(kBetaBlock) k_beta(int nc, int N, Work w)
{
    NaState* st = w.st;
    if (st->done) return;
    __shared__ double csum[kMaxColours];
    const int wave = threadIdx.x / dev::kWave, lane = threadIdx.x % dev::kWave;
    constexpr int kWaves = kBetaBlock / dev::kWave;
    for (int c = wave; c < nc; c += kWaves) {
        const int p0 = w.coff[c], p1 = w.coff[c + 1];
        const double s = wave_fold(w.pb + p0, p1 - p0);
        if (lane == 0) csum[c] = s;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double zr_new = 0.;
    for (int c = nc - 1; c >= 0; c--) zr_new += csum[c];
    st->beta = zr_new / st->zr;
    st->zr = zr_new;
    const int K = st->K + 1;
    st->K = K;
    w.hist[K] = zr_new;
    if (K >= N) {  // while (K < N)
        st->exit_reason = SFA_EXIT_CAP;
        st->done = 1;
    } else if (ref_absolute(zr_new) < st->tol2) {
        st->exit_reason = SFA_EXIT_CONVERGED;
        st->done = 1;
    }
}

// d = z + beta d on active rows (:1309-1318)
__global__ void __launch_bounds__(kBlock) k_direction(int N, const double* __restrict__ diag, Work w)
{
    if (w.st->done) return;
    const int c = row_id();
    if (c >= N || diag[c] == 0.) return;
    w.d[c] = w.z[c] + w.st->beta * w.d[c];
}

// ---- host side

inline unsigned blocks_for(int n) { return (unsigned)((n + kBlock - 1) / kBlock); }

struct DeviceScratch {
    std::mutex mu;
    double* ptr[kMaxDevices] = {};
};

DeviceScratch& device_scratch()
{
    static DeviceScratch s;
    return s;
}

// The address of this device's copy of `scratch`, looked up once per device
// with that device current.
double* scratch_on(int device)
{
    if (device < 0 || device >= kMaxDevices) throw Error{SFA_ERR_INVALID, "device id out of range"};
    DeviceScratch& ds = device_scratch();
    std::lock_guard<std::mutex> lock(ds.mu);
    if (!ds.ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error{SFA_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
        ds.ptr[device] = static_cast<double*>(p);
    }
    return ds.ptr[device];
}

const char* const kTitle = "Conjugate Gradient CSR (SSOR-Rainbow)";

// Log lines of solver/sparsesolver.cpp:1064,1172,1182,1220,1225-1226,1325;
// the z.r history is read in pieces into a stack buffer.
void print_log(int pm, int N, const NaState& st, const double* hist)
{
    const int K = st.K;
    if (pm >= -1) std::printf("(%d/%d)     BEGIN: %s\n", 0, N, kTitle);
    if (pm >= 0) {
        constexpr int kPiece = 256;
        double h[kPiece];
        const int last = pm >= 1 ? (st.exit_reason == SFA_EXIT_CAP ? K - 1 : K) : 0;
        for (int k0 = 0; k0 <= last; k0 += kPiece) {
            const int cnt = std::min(kPiece, last + 1 - k0);
            SFA_HIP(hipMemcpy(h, hist + k0, cnt * sizeof(double), hipMemcpyDeviceToHost));
            for (int j = 0; j < cnt; j++) {
                const int k = k0 + j;
                if (k == 0)
                    std::printf("(%d/%d)   RESTART: Residual norm is %.9le < %.9le\n", 0, N,
                                std::sqrt(h[j]), st.tol);
                if (pm >= 1 && k % pm == 0)
                    std::printf("(%d/%d)   INTERIM: Residual norm is %.9le < %.9le\n", k, N,
                                std::sqrt(h[j]), st.tol);
            }
        }
    }
    if (st.exit_reason == SFA_EXIT_BREAKDOWN)
        std::printf("(%d/%d) BREAKDOWN: Gradient energy is unreasonable with %.9le\n", K, N, st.dAd);
    if (st.exit_reason == SFA_EXIT_SMALL) {
        std::printf("(%d/%d)   WARNING: Residual norm is %.9le < %.9le\n", K, N, std::sqrt(st.zr), st.tol);
        std::printf("(%d/%d)   WARNING: Gradient energy is small with %.9le\n", K, N, st.dAd);
    }
    if (pm >= -1)
        std::printf("(%d/%d) %9s: Residual norm is %.9le < %.9le\n", K, N,
                    std::sqrt(st.zr) < st.tol ? "SUCCESS" : "FAILED", std::sqrt(st.zr), st.tol);
    std::fflush(stdout);
}

void read_state(const Work& w, NaState* out)
{
    SFA_HIP(hipMemcpy(out, w.st, sizeof(NaState), hipMemcpyDeviceToHost));
}

void solve(int N, double* x, const double* b, const int* A, const int* C, const double* V,
           double* residual, double precision, int print_modulo, const double* diag,
           double omega, int nc, const int* F, const int* B, const int* R, sfa_solve_info* info)
{
    // Sizes and pointers first: nothing is launched for a call that is
    // rejected here.
    if (N < 1) throw Error{SFA_ERR_INVALID, "N must be positive"};
    if (N > kMaxRows)
        throw Error{SFA_ERR_INVALID, "N = " + std::to_string(N) + " exceeds SFA_NOALLOC_MAX_ROWS = " +
                                         std::to_string(kMaxRows) + " (the scratch array's size)"};
    if (nc < 1 || nc > kMaxColours)
        throw Error{SFA_ERR_INVALID, "num_colors must be 1.." + std::to_string(kMaxColours) +
                                         " (SFA_NOALLOC_MAX_COLORS)"};
    if (!x || !b || !A || !C || !V || !residual || !diag || !F || !B || !R)
        throw Error{SFA_ERR_INVALID, "null pointer argument"};
    if (!(precision > 0)) throw Error{SFA_ERR_INVALID, "precision must be positive"};
    ensure_device();
    int device = 0;
    SFA_HIP(hipGetDevice(&device));
    double* base = scratch_on(device);
    const Work w = carve(base);
    const hipStream_t s = nullptr;  // the null stream: ordered after the caller's prior work

    // Colour offsets to the host (stack arrays, no allocation) and the
    // colour table of partial offsets to the device.
    int hB[kMaxColours + 1];
    int hoff[kMaxColours + 1];
    SFA_HIP(hipMemcpy(hB, B, (nc + 1) * sizeof(int), hipMemcpyDeviceToHost));
    if (hB[0] != 0 || hB[nc] != N) throw Error{SFA_ERR_COLOURING, "B must run from 0 to N"};
    hoff[0] = 0;
    for (int c = 0; c < nc; c++) {
        if (hB[c + 1] < hB[c]) throw Error{SFA_ERR_COLOURING, "B must be non-decreasing"};
        hoff[c + 1] = hoff[c] + (int)blocks_for(hB[c + 1] - hB[c]);
    }
    int hA0[1];
    SFA_HIP(hipMemcpy(hA0, A, sizeof(int), hipMemcpyDeviceToHost));
    if (hA0[0] != 0) throw Error{SFA_ERR_INVALID, "csrrows[0] must be 0"};
    SFA_HIP(hipMemcpy(w.coff, hoff, (nc + 1) * sizeof(int), hipMemcpyHostToDevice));

    // Validation on the device, before any value is read: the CSR
    // structure, then the colouring (R a permutation consistent with F, B).
    const unsigned nb = blocks_for(N);
    {
        NaState clean{};
        SFA_HIP(hipMemcpy(w.st, &clean, sizeof(NaState), hipMemcpyHostToDevice));
        int* pos = reinterpret_cast<int*>(w.q);
        hipLaunchKernelGGL(k_check_csr, dim3(nb), dim3(kBlock), 0, s, N, A, C, w.st);
        hipLaunchKernelGGL(k_fill_int, dim3(nb), dim3(kBlock), 0, s, pos, -1, N);
        hipLaunchKernelGGL(k_check_rows, dim3(nb), dim3(kBlock), 0, s, N, nc, F, B, R, pos, w.st);
        hipLaunchKernelGGL(k_check_permutation, dim3(nb), dim3(kBlock), 0, s, N, R,
                           (const int*)pos, w.st);
        SFA_HIP(hipGetLastError());
        NaState chk;
        read_state(w, &chk);
        if (chk.bad_csr)
            throw Error{SFA_ERR_INVALID, "CSR arrays invalid (row offsets or a column out of range)"};
        if (chk.bad_colouring)
            throw Error{SFA_ERR_COLOURING, "R must be a permutation of the rows and F agree with B/R"};
    }

    // Restart (K = 0).
    hipLaunchKernelGGL(k_zero, dim3(nb), dim3(kBlock), 0, s, w, N);
    hipLaunchKernelGGL(k_sumsq, dim3(nb), dim3(kBlock), 0, s, N, b, w);
    hipLaunchKernelGGL(k_start, dim3(1), dim3(dev::kWave), 0, s, (int)nb, precision, w);
    hipLaunchKernelGGL(k_residual, dim3(nb), dim3(kBlock), 0, s, N, A, C, V, diag, b,
                       (const double*)x, residual);
    for (int c = 0; c < nc; c++) {
        const int rows = hB[c + 1] - hB[c];
        if (rows > 0)
            hipLaunchKernelGGL(k_forward<false>, dim3(blocks_for(rows)), dim3(kBlock), 0, s, hB[c],
                               rows, A, C, V, diag, F, R, omega, w, x, residual);
    }
    for (int c = nc - 1; c >= 0; c--) {
        const int rows = hB[c + 1] - hB[c];
        if (rows > 0)
            hipLaunchKernelGGL(k_backward, dim3(blocks_for(rows)), dim3(kBlock), 0, s, hB[c], rows,
                               hoff[c], A, C, V, diag, F, R, omega, w, (const double*)residual);
    }
    hipLaunchKernelGGL(k_direction_init, dim3(nb), dim3(kBlock), 0, s, N, diag,
                       (const double*)residual, w);
    hipLaunchKernelGGL(k_restart_zr, dim3(1), dim3(dev::kWave), 0, s, (int)nb, w);
    SFA_HIP(hipGetLastError());

    // Iterations, a few between looks at the state; every kernel returns at
    // once after the state says the loop has ended.
    NaState st;
    read_state(w, &st);
    constexpr int kPerLook = 8;
    while (!st.done) {
        for (int it = 0; it < kPerLook; it++) {
            hipLaunchKernelGGL(k_spmv, dim3(nb), dim3(kBlock), 0, s, N, A, C, V, diag, w);
            hipLaunchKernelGGL(k_alpha, dim3(1), dim3(dev::kWave), 0, s, (int)nb, w);
            for (int c = 0; c < nc; c++) {
                const int rows = hB[c + 1] - hB[c];
                if (rows > 0)
                    hipLaunchKernelGGL(k_forward<true>, dim3(blocks_for(rows)), dim3(kBlock), 0, s,
                                       hB[c], rows, A, C, V, diag, F, R, omega, w, x, residual);
            }
            for (int c = nc - 1; c >= 0; c--) {
                const int rows = hB[c + 1] - hB[c];
                if (rows > 0)
                    hipLaunchKernelGGL(k_backward, dim3(blocks_for(rows)), dim3(kBlock), 0, s,
                                       hB[c], rows, hoff[c], A, C, V, diag, F, R, omega, w,
                                       (const double*)residual);
            }
            hipLaunchKernelGGL(k_beta, dim3(1), dim3(kBetaBlock), 0, s, nc, N, w);
            hipLaunchKernelGGL(k_direction, dim3(nb), dim3(kBlock), 0, s, N, diag, w);
        }
        SFA_HIP(hipGetLastError());
        read_state(w, &st);
    }

    sfa_solve_info out{};
    out.iterations = st.K;
    out.zr = st.zr;
    out.tolerance = st.tol;
    out.exit_reason = st.exit_reason;
    out.success = std::sqrt(st.zr) < st.tol ? 1 : 0;
    if (info) *info = out;
    if (print_modulo >= -2) print_log(print_modulo, N, st, w.hist);
}

}  // namespace noalloc
}  // namespace sfa

extern "C" int sfa_ConjugateGradientSolverCSR_Rainbow_noalloc(
    int N, double* x, const double* b, const int* csrrows, const int* csrcolumns,
    const double* csrvalues, double* residual, double precision, int print_modulo,
    const double* diagonal, double omega, int num_colors, const int* F, const int* B,
    const int* R, sfa_solve_info* info)
{
    return sfa::guarded([&] {
        sfa::noalloc::solve(N, x, b, csrrows, csrcolumns, csrvalues, residual, precision,
                            print_modulo, diagonal, omega, num_colors, F, B, R, info);
    });
}
