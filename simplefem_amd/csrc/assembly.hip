// assembly.hip -- the reference's CSR products and transpose with their
// row-wise sort-and-compress on the GPU, bit for bit:
// MatrixCSRMultiplication (sparse/matcsr.cpp:767-849),
// MatrixCSRMultiplication_reduced (:857-962) and MatrixCSR::getTranspose
// (:291-350), all ending in the compression of :585-687.
//
// The compression is a per-row selection sort that merges an entry into the
// first one of its column as it meets it and then moves zeros to the end; the
// order of those merges decides how duplicate products round.  It is replayed
// here in the reference's own loop order, one thread per row, so every sum is
// formed in the same order.  Its cost grows with the square of a row's
// length, and that per-row work is what runs in parallel on the GPU.  The
// products are formed on the device, one wavefront per row, in the
// reference's (entry of m1, entry of m2) order.  The transpose's counting
// placement is a single O(nnz) pass and stays on the host, as in the
// reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <memory>
#include <string>
#include <vector>

#include "../../include/simplefem_amd.h"
#include "host_common.hpp"
#include "launch.hpp"

// A CSR matrix computed by this file (host arrays), handed out as sfa_csr.
struct sfa_csr_s {
    int rows = 0, cols = 0;
    std::vector<int> A, C;
    std::vector<double> V;
};

namespace sfa {
namespace {

constexpr int kRowBlock = 256;               // threads per block, one row each
constexpr int kWaveRows = kRowBlock / 64;    // rows per block, one wavefront each

// Products per row of m1 * m2: the lengths of the rows of m2 its entries name.
__global__ void __launch_bounds__(kRowBlock) k_product_counts(int rows, const int* __restrict__ A1,
                                                              const int* __restrict__ C1,
                                                              const int* __restrict__ A2,
                                                              long long* __restrict__ counts)
{
    const int r = blockIdx.x * kRowBlock + threadIdx.x;
    if (r >= rows) return;
    long long n = 0;
    for (int i = A1[r]; i < A1[r + 1]; i++) n += A2[C1[i] + 1] - A2[C1[i]];
    counts[r] = n;
}

// Row r's products from position P[r]: for each entry (r, c1) of m1 in order
// and each entry (c1, c2) of m2 in order, column c2 and value v1 * v2
// (sparse/matcsr.cpp:806-832).  One wavefront per row; its lanes copy one
// row of m2 at a time.
__global__ void __launch_bounds__(kRowBlock) k_product_fill(int rows, const int* __restrict__ A1,
                                                            const int* __restrict__ C1,
                                                            const double* __restrict__ V1,
                                                            const int* __restrict__ A2,
                                                            const int* __restrict__ C2,
                                                            const double* __restrict__ V2,
                                                            const long long* __restrict__ P,
                                                            int* __restrict__ C,
                                                            double* __restrict__ V)
{
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * kWaveRows + (threadIdx.x >> 6);
    if (r >= rows) return;
    long long at = P[r];
    for (int i = A1[r]; i < A1[r + 1]; i++) {
        const int c1 = C1[i];
        const double v1 = V1[i];
        const int k0 = A2[c1], len = A2[c1 + 1] - k0;
        for (int t = lane; t < len; t += 64) {
            C[at + t] = C2[k0 + t];
            V[at + t] = v1 * V2[k0 + t];
        }
        at += len;
    }
}

// Products per row of the reduced product: pairs whose two values are both
// nonzero (sparse/matcsr.cpp:889-892).
__global__ void __launch_bounds__(kRowBlock) k_product_counts_nz(int rows, const int* __restrict__ A1,
                                                                 const int* __restrict__ C1,
                                                                 const double* __restrict__ V1,
                                                                 const int* __restrict__ A2,
                                                                 const double* __restrict__ V2,
                                                                 long long* __restrict__ counts)
{
    const int r = blockIdx.x * kRowBlock + threadIdx.x;
    if (r >= rows) return;
    long long n = 0;
    for (int i = A1[r]; i < A1[r + 1]; i++) {
        if (V1[i] == 0.) continue;
        for (int k = A2[C1[i]]; k < A2[C1[i] + 1]; k++) n += V2[k] != 0. ? 1 : 0;
    }
    counts[r] = n;
}

// Row r's products with both values nonzero, from position P[r], in the
// reference's (entry of m1, entry of m2) order.  One wavefront per row: its
// lanes take one row of m2 at a time and pack the kept products with a
// ballot, lanes in order.
__global__ void __launch_bounds__(kRowBlock) k_product_fill_nz(int rows, const int* __restrict__ A1,
                                                               const int* __restrict__ C1,
                                                               const double* __restrict__ V1,
                                                               const int* __restrict__ A2,
                                                               const int* __restrict__ C2,
                                                               const double* __restrict__ V2,
                                                               const long long* __restrict__ P,
                                                               int* __restrict__ C,
                                                               double* __restrict__ V)
{
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * kWaveRows + (threadIdx.x >> 6);
    if (r >= rows) return;
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    long long at = P[r];
    for (int i = A1[r]; i < A1[r + 1]; i++) {
        const double v1 = V1[i];
        if (v1 == 0.) continue;  // wave-uniform
        const int c1 = C1[i];
        const int k0 = A2[c1], len = A2[c1 + 1] - k0;
        for (int t0 = 0; t0 < len; t0 += 64) {
            const int t = t0 + lane;
            const double v2 = t < len ? V2[k0 + t] : 0.;
            const bool keep = t < len && v2 != 0.;
            const unsigned long long mask = __ballot(keep);
            if (keep) {
                const long long to = at + __popcll(mask & below);
                C[to] = C2[k0 + t];
                V[to] = v1 * v2;
            }
            at += __popcll(mask);
        }
    }
}

// The reduced product's row before its compression (sparse/matcsr.cpp:
// 886-950): the distinct columns in increasing order, each value 0 plus its
// products in the order they were formed.  The row's products are sorted by
// column with an insertion sort, which keeps equal columns in their order,
// and each run is summed from 0; the row's distinct entries end up at its
// start and their count in distinct[r].  One thread per row.
__global__ void __launch_bounds__(kRowBlock) k_merge_columns(int rows, const long long* __restrict__ P,
                                                             int* __restrict__ C,
                                                             double* __restrict__ V,
                                                             int* __restrict__ distinct)
{
    const int r = blockIdx.x * kRowBlock + threadIdx.x;
    if (r >= rows) return;
    const long long b = P[r], e = P[r + 1];
    for (long long i = b + 1; i < e; i++) {
        const int c = C[i];
        const double v = V[i];
        long long j = i;
        for (; j > b && C[j - 1] > c; j--) {
            C[j] = C[j - 1];
            V[j] = V[j - 1];
        }
        C[j] = c;
        V[j] = v;
    }
    long long out = b;
    for (long long i = b; i < e;) {
        const int c = C[i];
        double s = 0.;
        for (; i < e && C[i] == c; i++) s += V[i];
        C[out] = c;
        V[out] = s;
        out++;
    }
    distinct[r] = (int)(out - b);
}

// The reference's compression of one row, positions b .. e-1
// (sparse/matcsr.cpp:585-687): a selection sort in which an entry meeting
// one of its own column adds that entry's value to itself and zeroes it;
// then zeros are moved to the end, other entries keeping their order; the
// row keeps its leading nonzero entries.  One thread per row.
//
// Each wavefront's 64 rows are first copied to scratch interleaved by lane
// (entry k of lane l's row at wofs + 64 k + l), so that the sort's accesses,
// which all lanes make at the same step, are one cache line per 8 lanes
// instead of one per lane.  Step j of the sort's inner loop writes position j
// only, so the next kAhead positions are loaded together and then taken in
// order: the same additions in the same order.  The reference moves the
// zeros with bubble passes that never let two nonzero entries cross, and
// runs enough of them to finish; the entries it keeps are therefore the
// row's nonzero entries in order, which one pass gathers back into the row.
__global__ void __launch_bounds__(kRowBlock) k_sort_compress(int rows, const long long* __restrict__ P,
                                                             const long long* __restrict__ wofs,
                                                             int* __restrict__ C,
                                                             double* __restrict__ V,
                                                             int* __restrict__ sC,
                                                             double* __restrict__ sV,
                                                             int* __restrict__ kept)
{
    constexpr int kAhead = 8;
    const int r = blockIdx.x * kRowBlock + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const long long w = (long long)blockIdx.x * (kRowBlock / 64) + (threadIdx.x >> 6);
    if (r >= rows) return;
    const long long b = P[r];
    const int n = (int)(P[r + 1] - b);
    int* c = sC + wofs[w] + lane;
    double* v = sV + wofs[w] + lane;
    for (int k = 0; k < n; k++) {
        c[64 * (long long)k] = C[b + k];
        v[64 * (long long)k] = V[b + k];
    }
    for (int i = 0; i < n; i++) {
        int ci = c[64 * (long long)i];
        double vi = v[64 * (long long)i];
        for (int j0 = i + 1; j0 < n; j0 += kAhead) {
            int cj[kAhead];
            double vj[kAhead];
#pragma unroll
            for (int u = 0; u < kAhead; u++) {
                if (j0 + u < n) {
                    cj[u] = c[64 * (long long)(j0 + u)];
                    vj[u] = v[64 * (long long)(j0 + u)];
                }
            }
#pragma unroll
            for (int u = 0; u < kAhead; u++) {
                if (j0 + u >= n) break;
                const long long at = 64 * (long long)(j0 + u);
                if (ci == cj[u]) {
                    vi += vj[u];
                    v[at] = 0.;
                } else if (ci > cj[u]) {
                    c[at] = ci;
                    v[at] = vi;
                    ci = cj[u];
                    vi = vj[u];
                }
            }
        }
        c[64 * (long long)i] = ci;
        v[64 * (long long)i] = vi;
    }
    int k = 0;
    for (int i = 0; i < n; i++) {
        const double x = v[64 * (long long)i];
        if (x != 0.) {
            C[b + k] = c[64 * (long long)i];
            V[b + k] = x;
            k++;
        }
    }
    kept[r] = k;
}

// The same compression in place, for rows too uneven in length to share a
// wavefront's scratch (see compress_rows).
__global__ void __launch_bounds__(kRowBlock) k_sort_compress_in_place(int rows,
                                                                      const long long* __restrict__ P,
                                                                      int* __restrict__ C,
                                                                      double* __restrict__ V,
                                                                      int* __restrict__ kept)
{
    constexpr int kAhead = 8;
    const int r = blockIdx.x * kRowBlock + threadIdx.x;
    if (r >= rows) return;
    const long long b = P[r], e = P[r + 1];
    for (long long i = b; i < e; i++) {
        int ci = C[i];
        double vi = V[i];
        for (long long j0 = i + 1; j0 < e; j0 += kAhead) {
            int cj[kAhead];
            double vj[kAhead];
#pragma unroll
            for (int u = 0; u < kAhead; u++) {
                if (j0 + u < e) {
                    cj[u] = C[j0 + u];
                    vj[u] = V[j0 + u];
                }
            }
#pragma unroll
            for (int u = 0; u < kAhead; u++) {
                if (j0 + u >= e) break;
                if (ci == cj[u]) {
                    vi += vj[u];
                    V[j0 + u] = 0.;
                } else if (ci > cj[u]) {
                    C[j0 + u] = ci;
                    V[j0 + u] = vi;
                    ci = cj[u];
                    vi = vj[u];
                }
            }
        }
        C[i] = ci;
        V[i] = vi;
    }
    long long k = b;
    for (long long i = b; i < e; i++) {
        const double x = V[i];
        if (x != 0.) {
            C[k] = C[i];
            V[k] = x;
            k++;
        }
    }
    kept[r] = (int)(k - b);
}

// Row r's kept entries, positions P[r] .. P[r] + (Q[r+1] - Q[r]), to Q[r].
// One wavefront per row.
__global__ void __launch_bounds__(kRowBlock) k_compact(int rows, const long long* __restrict__ P,
                                                       const int* __restrict__ Q,
                                                       const int* __restrict__ C,
                                                       const double* __restrict__ V,
                                                       int* __restrict__ Cout,
                                                       double* __restrict__ Vout)
{
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * kWaveRows + (threadIdx.x >> 6);
    if (r >= rows) return;
    const long long from = P[r];
    const int to = Q[r], n = Q[r + 1] - Q[r];
    for (int t = lane; t < n; t += 64) {
        Cout[to + t] = C[from + t];
        Vout[to + t] = V[from + t];
    }
}

int blocks_for(long long items, int per_block)
{
    const long long b = (items + per_block - 1) / per_block;
    if (b > INT_MAX) throw Error{SFA_ERR_INVALID, "too many rows"};
    return (int)std::max<long long>(1, b);
}

// Device buffers freed on every path.
struct Buffers {
    std::vector<void*> held;
    template <class T>
    T* get(size_t count)
    {
        T* p = dev_alloc<T>(count);
        held.push_back(p);
        return p;
    }
    ~Buffers()
    {
        // the kernels run on the null stream; on an error path one may still
        // be queued when the buffers go back to the cache
        (void)hipStreamSynchronize(nullptr);
        for (void* p : held) dev_free(p);
    }
};

template <class T>
T* upload(Buffers& buf, const T* host, size_t count)
{
    T* d = buf.get<T>(count);
    if (count) SFA_HIP(hipMemcpy(d, host, count * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

// The compressed rows of a row-ordered entry list on the device: the rows of
// entries (dC, dV) at offsets hP (rows + 1) are compressed in place and
// compacted into out (host arrays).
void compress_rows(int rows, int cols, const std::vector<long long>& hP, int* dC, double* dV,
                   Buffers& buf, sfa_csr_s* out)
{
    long long* dP = upload(buf, hP.data(), hP.size());
    int* dkept = buf.get<int>(rows);
    if (rows > 0) {
        // scratch for each wavefront's 64 rows, interleaved: 64 x its longest row
        const long long nwaves = ((long long)rows + 63) / 64;
        std::vector<long long> wofs(nwaves + 1, 0);
        for (long long w = 0; w < nwaves; w++) {
            long long widest = 0;
            for (long long r = w * 64; r < std::min<long long>(rows, (w + 1) * 64); r++)
                widest = std::max(widest, hP[r + 1] - hP[r]);
            wofs[w + 1] = wofs[w] + 64 * widest;
        }
        // rows of very different lengths in one wavefront would make the
        // scratch many times the entries: those are compressed in place
        if (wofs[nwaves] <= 4 * hP[rows] + 64 * 1024) {
            long long* dwofs = upload(buf, wofs.data(), wofs.size());
            int* sC = buf.get<int>((size_t)wofs[nwaves]);
            double* sV = buf.get<double>((size_t)wofs[nwaves]);
            launch(k_sort_compress, dim3(blocks_for(rows, kRowBlock)), dim3(kRowBlock), nullptr,
                   rows, (const long long*)dP, (const long long*)dwofs, dC, dV, sC, sV, dkept);
        } else {
            launch(k_sort_compress_in_place, dim3(blocks_for(rows, kRowBlock)), dim3(kRowBlock),
                   nullptr, rows, (const long long*)dP, dC, dV, dkept);
        }
    }
    std::vector<int> kept(rows);
    if (rows > 0) SFA_HIP(hipMemcpy(kept.data(), dkept, rows * sizeof(int), hipMemcpyDeviceToHost));
    out->rows = rows;
    out->cols = cols;
    out->A.assign(rows + 1, 0);
    for (int r = 0; r < rows; r++) out->A[r + 1] = out->A[r] + kept[r];
    const int nnz = out->A[rows];
    out->C.resize(nnz);
    out->V.resize(nnz);
    if (nnz > 0) {
        int* dQ = upload(buf, out->A.data(), out->A.size());
        int* dCo = buf.get<int>(nnz);
        double* dVo = buf.get<double>(nnz);
        launch(k_compact, dim3(blocks_for(rows, kWaveRows)), dim3(kRowBlock), nullptr, rows,
               (const long long*)dP, (const int*)dQ, (const int*)dC, (const double*)dV, dCo, dVo);
        SFA_HIP(hipMemcpy(out->C.data(), dCo, nnz * sizeof(int), hipMemcpyDeviceToHost));
        SFA_HIP(hipMemcpy(out->V.data(), dVo, nnz * sizeof(double), hipMemcpyDeviceToHost));
    }
    SFA_HIP(hipGetLastError());
}

}  // namespace
}  // namespace sfa

using sfa::Error;  // for SFA_HIP in the entry points

extern "C" {

int sfa_MatrixCSRMultiplication(int rows1, int cols1, const int* A1, const int* C1,
                                const double* V1, int rows2, int cols2, const int* A2,
                                const int* C2, const double* V2, sfa_csr* out)
{
    return sfa::guarded([&] {
        if (!out) throw sfa::Error{SFA_ERR_INVALID, "null output handle"};
        *out = nullptr;
        if (cols1 != rows2) throw sfa::Error{SFA_ERR_INVALID, "columns of m1 != rows of m2"};
        std::string err = sfa::check_csr_rect(rows1, cols1, A1, C1);
        if (err.empty()) err = sfa::check_csr_rect(rows2, cols2, A2, C2);
        if (!err.empty()) throw sfa::Error{SFA_ERR_INVALID, err};
        if ((A1[rows1] > 0 && !V1) || (A2[rows2] > 0 && !V2))
            throw sfa::Error{SFA_ERR_INVALID, "null values"};
        sfa::ensure_device();
        std::unique_ptr<sfa_csr_s> res(new sfa_csr_s());
        sfa::Buffers buf;
        const int nnz1 = A1[rows1], nnz2 = A2[rows2];
        int* dA1 = sfa::upload(buf, A1, rows1 + 1);
        int* dC1 = sfa::upload(buf, C1, nnz1);
        double* dV1 = sfa::upload(buf, V1, nnz1);
        int* dA2 = sfa::upload(buf, A2, rows2 + 1);
        int* dC2 = sfa::upload(buf, C2, nnz2);
        double* dV2 = sfa::upload(buf, V2, nnz2);
        long long* dcount = buf.get<long long>(rows1);
        if (rows1 > 0)
            sfa::launch(sfa::k_product_counts, dim3(sfa::blocks_for(rows1, sfa::kRowBlock)),
                        dim3(sfa::kRowBlock), nullptr, rows1, (const int*)dA1, (const int*)dC1,
                        (const int*)dA2, dcount);
        std::vector<long long> P(rows1 + 1, 0);
        if (rows1 > 0)
            SFA_HIP(hipMemcpy(P.data() + 1, dcount, rows1 * sizeof(long long), hipMemcpyDeviceToHost));
        for (int r = 0; r < rows1; r++) P[r + 1] += P[r];
        // the reference stores the product list with int offsets
        if (P[rows1] > INT_MAX) throw sfa::Error{SFA_ERR_INVALID, "more than 2^31 - 1 products"};
        const long long np = P[rows1];
        int* dC = buf.get<int>(np);
        double* dV = buf.get<double>(np);
        if (rows1 > 0 && np > 0) {
            long long* dP = sfa::upload(buf, P.data(), P.size());
            sfa::launch(sfa::k_product_fill, dim3(sfa::blocks_for(rows1, sfa::kWaveRows)),
                        dim3(sfa::kRowBlock), nullptr, rows1, (const int*)dA1, (const int*)dC1,
                        (const double*)dV1, (const int*)dA2, (const int*)dC2, (const double*)dV2,
                        (const long long*)dP, dC, dV);
        }
        sfa::compress_rows(rows1, cols2, P, dC, dV, buf, res.get());
        SFA_HIP(hipDeviceSynchronize());
        *out = res.release();
    });
}

int sfa_MatrixCSRMultiplication_reduced(int rows1, int cols1, const int* A1, const int* C1,
                                        const double* V1, int rows2, int cols2, const int* A2,
                                        const int* C2, const double* V2, sfa_csr* out)
{
    return sfa::guarded([&] {
        if (!out) throw sfa::Error{SFA_ERR_INVALID, "null output handle"};
        *out = nullptr;
        if (cols1 != rows2) throw sfa::Error{SFA_ERR_INVALID, "columns of m1 != rows of m2"};
        std::string err = sfa::check_csr_rect(rows1, cols1, A1, C1);
        if (err.empty()) err = sfa::check_csr_rect(rows2, cols2, A2, C2);
        if (!err.empty()) throw sfa::Error{SFA_ERR_INVALID, err};
        if ((A1[rows1] > 0 && !V1) || (A2[rows2] > 0 && !V2))
            throw sfa::Error{SFA_ERR_INVALID, "null values"};
        sfa::ensure_device();
        std::unique_ptr<sfa_csr_s> res(new sfa_csr_s());
        sfa::Buffers buf;
        const int nnz1 = A1[rows1], nnz2 = A2[rows2];
        int* dA1 = sfa::upload(buf, A1, rows1 + 1);
        int* dC1 = sfa::upload(buf, C1, nnz1);
        double* dV1 = sfa::upload(buf, V1, nnz1);
        int* dA2 = sfa::upload(buf, A2, rows2 + 1);
        int* dC2 = sfa::upload(buf, C2, nnz2);
        double* dV2 = sfa::upload(buf, V2, nnz2);
        long long* dcount = buf.get<long long>(rows1);
        if (rows1 > 0)
            sfa::launch(sfa::k_product_counts_nz, dim3(sfa::blocks_for(rows1, sfa::kRowBlock)),
                        dim3(sfa::kRowBlock), nullptr, rows1, (const int*)dA1, (const int*)dC1,
                        (const double*)dV1, (const int*)dA2, (const double*)dV2, dcount);
        std::vector<long long> P(rows1 + 1, 0);
        if (rows1 > 0)
            SFA_HIP(hipMemcpy(P.data() + 1, dcount, rows1 * sizeof(long long), hipMemcpyDeviceToHost));
        for (int r = 0; r < rows1; r++) P[r + 1] += P[r];
        const long long np = P[rows1];
        int* dC = buf.get<int>(np);
        double* dV = buf.get<double>(np);
        int* ddistinct = buf.get<int>(rows1);
        std::vector<long long> Q(rows1 + 1, 0);  // offsets of the distinct entries
        if (rows1 > 0 && np > 0) {
            long long* dP = sfa::upload(buf, P.data(), P.size());
            sfa::launch(sfa::k_product_fill_nz, dim3(sfa::blocks_for(rows1, sfa::kWaveRows)),
                        dim3(sfa::kRowBlock), nullptr, rows1, (const int*)dA1, (const int*)dC1,
                        (const double*)dV1, (const int*)dA2, (const int*)dC2, (const double*)dV2,
                        (const long long*)dP, dC, dV);
            sfa::launch(sfa::k_merge_columns, dim3(sfa::blocks_for(rows1, sfa::kRowBlock)),
                        dim3(sfa::kRowBlock), nullptr, rows1, (const long long*)dP, dC, dV,
                        ddistinct);
            std::vector<int> distinct(rows1);
            SFA_HIP(hipMemcpy(distinct.data(), ddistinct, rows1 * sizeof(int),
                              hipMemcpyDeviceToHost));
            for (int r = 0; r < rows1; r++) Q[r + 1] = Q[r] + distinct[r];
        }
        // the reference stores the distinct entries with int offsets
        if (Q[rows1] > INT_MAX) throw sfa::Error{SFA_ERR_INVALID, "more than 2^31 - 1 entries"};
        const long long nd = Q[rows1];
        int* dCd = buf.get<int>(nd);
        double* dVd = buf.get<double>(nd);
        if (nd > 0) {
            std::vector<int> Qi(Q.begin(), Q.end());
            long long* dP = sfa::upload(buf, P.data(), P.size());
            int* dQ = sfa::upload(buf, Qi.data(), Qi.size());
            sfa::launch(sfa::k_compact, dim3(sfa::blocks_for(rows1, sfa::kWaveRows)),
                        dim3(sfa::kRowBlock), nullptr, rows1, (const long long*)dP,
                        (const int*)dQ, (const int*)dC, (const double*)dV, dCd, dVd);
        }
        // then the compression of sparse/matcsr.cpp:959 (:690-693 calls :585-687)
        sfa::compress_rows(rows1, cols2, Q, dCd, dVd, buf, res.get());
        SFA_HIP(hipDeviceSynchronize());
        *out = res.release();
    });
}

int sfa_MatrixCSR_getTranspose(int rows, int cols, const int* A, const int* C, const double* V,
                               sfa_csr* out)
{
    return sfa::guarded([&] {
        if (!out) throw sfa::Error{SFA_ERR_INVALID, "null output handle"};
        *out = nullptr;
        std::string err = sfa::check_csr_rect(rows, cols, A, C);
        if (!err.empty()) throw sfa::Error{SFA_ERR_INVALID, err};
        const int nnz = A[rows];
        if (nnz > 0 && !V) throw sfa::Error{SFA_ERR_INVALID, "null values"};
        sfa::ensure_device();
        // the reference's counting placement (sparse/matcsr.cpp:303-335):
        // each column's entries fill its new row from the back, in row order
        std::vector<long long> B(cols + 1, 0);
        std::vector<int> Z(cols, 0), D(nnz);
        std::vector<double> W(nnz);
        for (int i = 0; i < nnz; i++) Z[C[i]]++;
        for (int c = 0; c < cols; c++) B[c + 1] = B[c] + Z[c];
        for (int r = 0; r < rows; r++)
            for (int i = A[r]; i < A[r + 1]; i++) {
                const long long at = B[C[i]] + (--Z[C[i]]);
                D[at] = r;
                W[at] = V[i];
            }
        std::unique_ptr<sfa_csr_s> res(new sfa_csr_s());
        sfa::Buffers buf;
        int* dD = sfa::upload(buf, D.data(), D.size());
        double* dW = sfa::upload(buf, W.data(), W.size());
        sfa::compress_rows(cols, rows, B, dD, dW, buf, res.get());
        SFA_HIP(hipDeviceSynchronize());
        *out = res.release();
    });
}

int sfa_csr_sizes(sfa_csr m, int* rows, int* cols, int* nnz)
{
    return sfa::guarded([&] {
        if (!m || !rows || !cols || !nnz) throw sfa::Error{SFA_ERR_INVALID, "null argument"};
        *rows = m->rows;
        *cols = m->cols;
        *nnz = m->A.empty() ? 0 : m->A.back();
    });
}

int sfa_csr_copy(sfa_csr m, int* A, int* C, double* V)
{
    return sfa::guarded([&] {
        if (!m || !A || (!m->C.empty() && (!C || !V)))
            throw sfa::Error{SFA_ERR_INVALID, "null argument"};
        std::copy(m->A.begin(), m->A.end(), A);
        std::copy(m->C.begin(), m->C.end(), C);
        std::copy(m->V.begin(), m->V.end(), V);
    });
}

int sfa_csr_destroy(sfa_csr m)
{
    return sfa::guarded([&] { delete m; });
}

}  // extern "C"
