// plan.hpp -- host-side preparation of a CSR system for the gfx950 kernels.
//
// A plan fixes, once per matrix and solver variant:
//   * a row permutation into "device order": the rows the variant works on
//     ("active" rows, diagonal != 0 for the SSOR family, precon != 0 for the
//     Jacobi variant, every row for plain CG) come first, so that every
//     row-wise vector kernel runs over the contiguous range [0, n_active);
//     for the Rainbow variants the active rows are grouped by colour;
//   * the sweep schedules: a list of row groups whose rows are independent
//     (colours for the Rainbow variants, dependency levels of the triangular
//     factor for the natural-order variants);
//   * the matrix parts each kernel reads, stored as SELL-64 slices: 64 rows
//     (one wavefront, one lane per row) per slice, entries column-major
//     inside a slice so that lane l reads entry k of its row at
//     off + 64*k + l (coalesced), each row's entries in the reference's CSR
//     order (so the per-row sums are formed in the same order), padded with
//     value 0 on column n, a vector slot that always holds 0.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace sfa {

constexpr int kSliceRows = 64;  // one gfx950 wavefront

// Leaves new elements uninitialised instead of zero-filling them: the SELL
// arrays are written in full by parallel loops, which then also take the
// first touch of their pages (a serial zero fill of a 10^7-entry part costs
// about as much as building it).
template <class T>
struct NoInitAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = NoInitAlloc<U>;
    };
    NoInitAlloc() = default;
    template <class U>
    NoInitAlloc(const NoInitAlloc<U>&) noexcept {}
    template <class U>
    void construct(U* p) noexcept { ::new ((void*)p) U; }
    template <class U, class... Args>
    void construct(U* p, Args&&... args) { ::new ((void*)p) U(std::forward<Args>(args)...); }
};
template <class T>
using HostArray = std::vector<T, NoInitAlloc<T>>;

struct SellPart {
    int nslices = 0;
    std::vector<int> slice_off;    // nslices+1 offsets into cols/vals (entries)
    HostArray<int> slice_rows;     // nslices*64 device-order row ids, -1 = pad
    HostArray<int> cols;           // device-order column ids, n = zero slot
    HostArray<double> vals;
    std::vector<int> group_slice;  // ngroups+1 slice ranges, in launch order
    std::vector<long long> group_entries;  // real entries per group
    std::vector<long long> group_rows;     // real rows per group
    // first device row of each group whose rows are consecutive in slice
    // order (the kernels then compute row ids instead of loading them), else -1
    std::vector<int> group_row0;
    // entries per lane of every slice of a group when they are all padded to
    // the same width (the kernels then compute slice offsets), else -1
    std::vector<int> group_width;
    int zero_slot = 0;             // padding column = the vectors' zero slot
    long long entries = 0;         // real (unpadded) entries
    long long rows = 0;            // real rows
    int ngroups() const { return (int)group_slice.size() - 1; }
};

struct Plan {
    int method = 0;
    int n = 0;
    int n_active = 0;
    long long nnz = 0;
    double omega = 1.0;
    std::vector<int> perm;        // perm[p] = original row of device row p
    std::vector<double> diag;     // device order: diagonal (or precon)
    // matrix parts, all in device order
    SellPart full;        // SpMV / restart residual over active rows
    SellPart lower;       // forward sweep (iteration)
    SellPart upper;       // backward sweep
    SellPart lower_restart;  // forward sweep of the restart when it differs
    bool has_lower_restart = false;
};

// SpMV row-length buckets: the widest row of a slice has <= 8, <= 16, <= 32
// or more entries.
constexpr int kSpmvBuckets = 4;
inline int spmv_bucket(int width)
{
    return width <= 8 ? 0 : width <= 16 ? 1 : width <= 32 ? 2 : 3;
}
// Rows and stored entries (padding included) of a part's slices per bucket.
void spmv_bucket_counts(const SellPart& part, long long rows[kSpmvBuckets],
                        long long stored[kSpmvBuckets]);

// Builds the plan.  aux is the diagonal (SSOR family) or precon (Jacobi), may
// be null for plain CG; F/B/R/num_colors only for the Rainbow variants.
// Returns "" on success, else an error message; *colouring_error is set when
// the message is about F/B/R.
// The host checks of a solver call's arguments (CSR structure, null arrays,
// the colouring's consistency), without building anything; "" when valid.
std::string check_arguments(int method, int n, const int* A, const int* C, const double* V,
                            const double* aux, int num_colors, const int* F, const int* B,
                            const int* R, bool* colouring_error);

std::string build_plan(Plan& plan, int method, int n, const int* A, const int* C,
                       const double* V, const double* aux, double omega, int num_colors,
                       const int* F, const int* B, const int* R, bool* colouring_error);

// SELL-64 part over all rows in identity order with every entry, for
// MatrixCSR::apply on a rows x columns matrix (padding points at column
// `columns`, the zero slot of the source vector).
void build_full_identity(SellPart& part, int rows, int columns, const int* A, const int* C,
                         const double* V);

// Whether two host arrays hold the same bytes (compared by several threads
// when large).
bool bytes_equal(const void* a, const void* b, size_t bytes);

// Checks a rows x columns CSR matrix; returns "" when it is well formed.
std::string check_csr_rect(int rows, int columns, const int* A, const int* C);

// Checks that A/C describe an n x n CSR matrix; returns "" when they do.
std::string check_csr(int n, const int* A, const int* C);

// Reference greedy colouring (sparse/rainbow.cpp:14-130).  B gets nc+1 entries.
int rainbow_colour(int n, const int* A, const int* C, int* F, int* B, int* R);

}  // namespace sfa
