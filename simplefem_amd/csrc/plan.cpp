// plan.cpp -- device order, sweep schedules and SELL-64 parts (host code).
#include "plan.hpp"

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>

#include "../../include/simplefem_amd.h"

namespace sfa {

namespace {

// Parallel passes over host arrays (a plan's parts, argument checks, the
// comparison with a kept plan) at this moment, over all host threads.  Each
// OpenMP parallel loop called from another host thread gets a team of its
// own; when several threads do this side by side (a server's workers),
// full-size teams would put many times more spinning threads than cores on
// the host.  So each takes its share of the threads: all of them when alone.
std::atomic<int> g_builds{0};

struct BuildShare {
    int threads;
    BuildShare()
    {
        const int k = ++g_builds;
        threads = std::max(1, omp_get_max_threads() / k);
    }
    ~BuildShare() { --g_builds; }
};

// Below this many elements a check runs on the calling thread alone.
constexpr long long kParallelMin = 1LL << 18;

// Whether pred(i) holds for any i in [0, count).
template <class Pred>
bool any_index(long long count, Pred pred)
{
    if (count < kParallelMin) {
        for (long long i = 0; i < count; i++)
            if (pred(i)) return true;
        return false;
    }
    const BuildShare share;
    int hit = 0;
#pragma omp parallel for schedule(static) reduction(| : hit) num_threads(share.threads)
    for (long long i = 0; i < count; i++) hit |= pred(i) ? 1 : 0;
    return hit != 0;
}

}  // namespace

bool bytes_equal(const void* a, const void* b, size_t bytes)
{
    constexpr size_t kChunk = size_t(1) << 20;
    const char* x = static_cast<const char*>(a);
    const char* y = static_cast<const char*>(b);
    if (bytes <= 4 * kChunk) return std::memcmp(x, y, bytes) == 0;
    const long long chunks = (long long)((bytes + kChunk - 1) / kChunk);
    const BuildShare share;
    int differ = 0;
#pragma omp parallel for schedule(static) reduction(| : differ) num_threads(share.threads)
    for (long long c = 0; c < chunks; c++) {
        const size_t off = (size_t)c * kChunk;
        differ |= std::memcmp(x + off, y + off, std::min(kChunk, bytes - off)) != 0;
    }
    return differ == 0;
}

std::string check_csr_rect(int rows, int columns, const int* A, const int* C)
{
    if (rows <= 0 || columns <= 0) return "matrix dimensions must be positive";
    if (!A || !C) return "null CSR pointer";
    if (A[0] != 0) return "csrrows[0] must be 0";
    if (any_index(rows, [&](long long i) { return A[i + 1] < A[i]; }))
        return "csrrows must be non-decreasing";
    const int nnz = A[rows];
    if (any_index(nnz, [&](long long k) { return C[k] < 0 || C[k] >= columns; }))
        return "column index out of range";
    return "";
}

std::string check_csr(int n, const int* A, const int* C)
{
    if (n <= 0) return "N must be positive";
    return check_csr_rect(n, n, A, C);
}

// Greedy colouring with the reference's rule (sparse/rainbow.cpp:48-76): row r
// takes the least colour no other column of its row carries, rows not yet
// visited counting as colour 0.  Written as a stamp-array mex instead of the
// reference's restart scan; both give that least colour.  B and R follow
// sparse/rainbow.cpp:80-117 (R lists each colour's rows back to front).
int rainbow_colour(int n, const int* A, const int* C, int* F, int* B, int* R)
{
    int num_colors = 1;
    std::vector<int> stamp(16, -1);
    for (int r = 0; r < n; r++) F[r] = 0;
    for (int r = 0; r < n; r++) {
        for (int k = A[r]; k < A[r + 1]; k++) {
            const int j = C[k];
            if (j == r) continue;
            const int c = F[j];
            if (c >= (int)stamp.size()) stamp.resize(2 * c + 2, -1);
            stamp[c] = r;
        }
        int c = 0;
        while (c < (int)stamp.size() && stamp[c] == r) c++;
        F[r] = c;
        if (c == num_colors) num_colors++;
    }
    std::vector<int> count(num_colors, 0);
    for (int r = 0; r < n; r++) count[F[r]]++;
    B[0] = 0;
    for (int c = 0; c < num_colors; c++) B[c + 1] = B[c] + count[c];
    for (int r = 0; r < n; r++) {
        const int c = F[r];
        R[B[c] + --count[c]] = r;
    }
    return num_colors;
}

namespace {

void start_part(SellPart& part)
{
    part = SellPart();
    part.slice_off.push_back(0);
    part.group_slice.push_back(0);
}

// Builds a part from groups of rows (device-order ids, in launch order): each
// group's rows fill consecutive SELL-64 slices.  keep(orow, k) decides whether
// CSR entry k of original row orow goes into the part.  Two passes, each
// parallel over the slices of all groups at once: the widths, then (after
// their prefix sum fixes every slice's offset) the entries and padding.
template <class Keep>
void build_part(SellPart& part, const std::vector<std::vector<int>>& groups,
                const std::vector<int>& perm, const std::vector<int>& iperm, int n, const int* A,
                const int* C, const double* V, Keep keep)
{
    start_part(part);
    part.zero_slot = n;
    const int G = (int)groups.size();
    std::vector<int> gs(G + 1, 0);  // first slice of each group
    for (int g = 0; g < G; g++)
        gs[g + 1] = gs[g] + ((int)groups[g].size() + kSliceRows - 1) / kSliceRows;
    const int ns = gs[G];
    std::vector<int> sg(ns);  // group of each slice
    for (int g = 0; g < G; g++)
        for (int sl = gs[g]; sl < gs[g + 1]; sl++) sg[sl] = g;

    const BuildShare share;
    std::vector<int> width(ns, 0);
#pragma omp parallel for schedule(dynamic, 16) num_threads(share.threads)
    for (int sl = 0; sl < ns; sl++) {
        const std::vector<int>& rows = groups[sg[sl]];
        const int s0 = (sl - gs[sg[sl]]) * kSliceRows;
        const int s1 = std::min((int)rows.size(), s0 + kSliceRows);
        int wmax = 0;
        for (int i = s0; i < s1; i++) {
            const int orow = perm[rows[i]];
            int w = 0;
            for (int k = A[orow]; k < A[orow + 1]; k++) w += keep(orow, k) ? 1 : 0;
            wmax = std::max(wmax, w);
        }
        width[sl] = wmax;
    }

    // Per group: one width for all its slices when that pads at most 10% more
    // entries; the first row when its rows are consecutive.
    for (int g = 0; g < G; g++) {
        const std::vector<int>& rows = groups[g];
        const int count = (int)rows.size();
        bool consecutive = true;
        for (int i = 1; i < count && consecutive; i++) consecutive = rows[i] == rows[0] + i;
        part.group_row0.push_back((count > 0 && consecutive) ? rows[0] : (count == 0 ? 0 : -1));
        long long wsum = 0;
        int wmax = 0;
        for (int sl = gs[g]; sl < gs[g + 1]; sl++) wsum += width[sl], wmax = std::max(wmax, width[sl]);
        const int nsl = gs[g + 1] - gs[g];
        if (nsl == 0) {
            part.group_width.push_back(0);
        } else if (((long long)wmax * nsl - wsum) * 10 <= wsum) {
            for (int sl = gs[g]; sl < gs[g + 1]; sl++) width[sl] = wmax;
            part.group_width.push_back(wmax);
        } else {
            part.group_width.push_back(-1);
        }
    }
    part.slice_off.resize((size_t)ns + 1);
    for (int sl = 0; sl < ns; sl++) part.slice_off[sl + 1] = part.slice_off[sl] + width[sl] * kSliceRows;
    const size_t total = (size_t)part.slice_off[ns];
    part.cols.resize(total);
    part.vals.resize(total);
    part.slice_rows.resize((size_t)ns * kSliceRows);

    std::vector<long long> stored(ns, 0);  // real entries per slice
#pragma omp parallel for schedule(dynamic, 16) num_threads(share.threads)
    for (int sl = 0; sl < ns; sl++) {
        const std::vector<int>& rows = groups[sg[sl]];
        const int s0 = (sl - gs[sg[sl]]) * kSliceRows;
        const int count = std::min((int)rows.size() - s0, kSliceRows);
        const size_t base = (size_t)part.slice_off[sl];
        const int wsl = width[sl];
        long long e = 0;
        for (int lane = 0; lane < kSliceRows; lane++) {
            int w = 0;
            if (lane < count) {
                const int p = rows[s0 + lane];
                const int orow = perm[p];
                part.slice_rows[(size_t)sl * kSliceRows + lane] = p;
                for (int k = A[orow]; k < A[orow + 1]; k++) {
                    if (!keep(orow, k)) continue;
                    const size_t at = base + (size_t)w * kSliceRows + lane;
                    part.cols[at] = iperm[C[k]];
                    part.vals[at] = V[k];
                    w++;
                }
            } else {
                part.slice_rows[(size_t)sl * kSliceRows + lane] = -1;
            }
            e += w;
            for (; w < wsl; w++) {  // padding: the zero slot, value 0
                const size_t at = base + (size_t)w * kSliceRows + lane;
                part.cols[at] = n;
                part.vals[at] = 0.0;
            }
        }
        stored[sl] = e;
    }

    for (int g = 0; g < G; g++) {
        long long e = 0;
        for (int sl = gs[g]; sl < gs[g + 1]; sl++) e += stored[sl];
        part.group_entries.push_back(e);
        part.group_rows.push_back((long long)groups[g].size());
        part.group_slice.push_back(gs[g + 1]);
        part.entries += e;
        part.rows += (long long)groups[g].size();
    }
    part.nslices = ns;
}


// Dependency levels of a triangular sweep over the active rows: rows of one
// level read no row of the same or a later level.  dep(orow, k) tells whether
// CSR entry k of row orow is a dependency; rows are visited in `order`.
template <class Dep>
std::vector<int> sweep_levels(int n, const int* A, const int* C, const std::vector<char>& active,
                              const std::vector<int>& order, Dep dep, int* nlevels)
{
    std::vector<int> level(n, -1);
    int top = 0;
    for (int orow : order) {
        if (!active[orow]) continue;
        int lv = 0;
        for (int k = A[orow]; k < A[orow + 1]; k++) {
            const int j = C[k];
            if (j == orow || !active[j] || !dep(orow, k)) continue;
            lv = std::max(lv, level[j] + 1);
        }
        level[orow] = lv;
        top = std::max(top, lv + 1);
    }
    *nlevels = top;
    return level;
}

}  // namespace

void spmv_bucket_counts(const SellPart& part, long long rows[kSpmvBuckets],
                        long long stored[kSpmvBuckets])
{
    for (int b = 0; b < kSpmvBuckets; b++) rows[b] = stored[b] = 0;
    for (int g = 0; g < part.ngroups(); g++) {
        for (int sl = part.group_slice[g]; sl < part.group_slice[g + 1]; sl++) {
            const int n = part.slice_off[sl + 1] - part.slice_off[sl];
            const int b = spmv_bucket(n / kSliceRows);
            int r = 0;
            if (part.group_row0[g] >= 0) {
                const long long left = part.group_rows[g] -
                                       (long long)(sl - part.group_slice[g]) * kSliceRows;
                r = (int)std::min<long long>(kSliceRows, std::max<long long>(0, left));
            } else {
                for (int l = 0; l < kSliceRows; l++) r += part.slice_rows[(size_t)sl * kSliceRows + l] >= 0;
            }
            rows[b] += r;
            stored[b] += n;
        }
    }
}

void build_full_identity(SellPart& part, int rows, int columns, const int* A, const int* C,
                         const double* V)
{
    std::vector<int> perm(std::max(rows, columns) + 1);
    std::vector<std::vector<int>> list(1, std::vector<int>(rows));
    for (int i = 0; i < (int)perm.size(); i++) perm[i] = i;
    for (int i = 0; i < rows; i++) list[0][i] = i;
    build_part(part, list, perm, perm, columns, A, C, V, [](int, int) { return true; });
}

namespace {

constexpr size_t kSortWindow = 512;  // rows per length-sorting window (8 slices)

// Breadth-first (Cuthill-McKee) order of the graph of the matrix: from row 0,
// then from the lowest unvisited row; neighbours in column order.
std::vector<int> bfs_order(int n, const int* A, const int* C)
{
    std::vector<int> order;
    order.reserve(n);
    std::vector<char> seen(n, 0);
    for (int s = 0; s < n; s++) {
        if (seen[s]) continue;
        seen[s] = 1;
        size_t head = order.size();
        order.push_back(s);
        while (head < order.size()) {
            const int r = order[head++];
            for (int k = A[r]; k < A[r + 1]; k++) {
                const int c = C[k];
                if (!seen[c]) {
                    seen[c] = 1;
                    order.push_back(c);
                }
            }
        }
    }
    return order;
}

// A locality order of the rows: rank[row] is its place, order[place] the row.
struct Locality {
    std::vector<int> rank, order;
};

// The breadth-first order if its summed |rank[row] - rank[col]| over the
// entries is smaller than the original order's, else the original order.
// The sums are exact integers, so the threads may add them in any order.
Locality locality_rank(int n, const int* A, const int* C)
{
    Locality bfs;
    bfs.order = bfs_order(n, A, C);
    bfs.rank.resize(n);
    for (int i = 0; i < n; i++) bfs.rank[bfs.order[i]] = i;
    const BuildShare share;
    long long sum_bfs = 0, sum_ident = 0;
    const int* rk = bfs.rank.data();
#pragma omp parallel for schedule(static) reduction(+ : sum_bfs, sum_ident) num_threads(share.threads)
    for (int r = 0; r < n; r++)
        for (int k = A[r]; k < A[r + 1]; k++) {
            sum_bfs += std::abs(rk[r] - rk[C[k]]);
            sum_ident += std::abs(r - C[k]);
        }
    if (sum_bfs < sum_ident) return bfs;
    Locality ident;
    ident.rank.resize(n);
    for (int i = 0; i < n; i++) ident.rank[i] = i;
    ident.order = ident.rank;
    return ident;
}

}  // namespace

std::string check_arguments(int method, int n, const int* A, const int* C, const double* V,
                            const double* aux, int num_colors, const int* F, const int* B,
                            const int* R, bool* colouring_error)
{
    *colouring_error = false;
    if (method < SFA_METHOD_CG || method > SFA_METHOD_EISENSTAT_RAINBOW) return "unknown method";
    std::string err = check_csr(n, A, C);
    if (!err.empty()) return err;
    if (!V) return "null csrvalues";
    if (method != SFA_METHOD_CG && !aux) return "null diagonal/precon";
    if (method == SFA_METHOD_RAINBOW || method == SFA_METHOD_EISENSTAT_RAINBOW) {
        *colouring_error = true;
        if (num_colors < 0 || !F || !B || !R) return "num_colors < 0 or null F/B/R";
        if (B[0] != 0 || B[num_colors] != n) return "B must run from 0 to N";
        for (int c = 0; c < num_colors; c++)
            if (B[c + 1] < B[c]) return "B must be non-decreasing";
        std::vector<unsigned char> seen(n, 0);
        if (n >= kParallelMin) {
            // one parallel pass says whether all is well; if not, the serial
            // pass below names the first problem in row order
            const BuildShare share;
            int bad = 0;
            unsigned char* sp = seen.data();
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad) num_threads(share.threads)
            for (int c = 0; c < num_colors; c++)
                for (int i = B[c]; i < B[c + 1]; i++) {
                    const int r = R[i];
                    if (r < 0 || r >= n || F[r] != c ||
                        __atomic_exchange_n(&sp[r], (unsigned char)1, __ATOMIC_RELAXED))
                        bad = 1;
                }
            if (!bad) {
                *colouring_error = false;
                return "";
            }
            std::fill(seen.begin(), seen.end(), (unsigned char)0);
        }
        for (int c = 0; c < num_colors; c++)
            for (int i = B[c]; i < B[c + 1]; i++) {
                const int r = R[i];
                if (r < 0 || r >= n || seen[r]) return "R must be a permutation of the rows";
                if (F[r] != c) return "F disagrees with B/R";
                seen[r] = 1;
            }
        *colouring_error = false;
    }
    return "";
}

std::string build_plan(Plan& plan, int method, int n, const int* A, const int* C, const double* V,
                       const double* aux, double omega, int num_colors, const int* F,
                       const int* B, const int* R, bool* colouring_error)
{
    const std::string err = check_arguments(method, n, A, C, V, aux, num_colors, F, B, R,
                                            colouring_error);
    if (!err.empty()) return err;
    const bool rainbow = method == SFA_METHOD_RAINBOW || method == SFA_METHOD_EISENSTAT_RAINBOW;
    const bool eisen = method == SFA_METHOD_SSOR_EISENSTAT || method == SFA_METHOD_EISENSTAT_RAINBOW;
    const bool sweeps = method >= SFA_METHOD_SSOR;

    plan = Plan();
    plan.method = method;
    plan.n = n;
    plan.nnz = A[n];
    plan.omega = omega;

    std::vector<char> active(n, 1);
    if (method != SFA_METHOD_CG)
        for (int i = 0; i < n; i++) active[i] = aux[i] != 0. ? 1 : 0;  // NaN counts as active

    // ---- locality order ----
    // Rows whose entries a wavefront gathers should sit close together in
    // every vector.  The original order does that for a mesh numbered along
    // its axes, not for one numbered in refinement order (a refined
    // vertex's neighbours were numbered levels apart); a breadth-first
    // (Cuthill-McKee) order over the matrix graph does it for both.  Of the
    // two, the one with the smaller mean index distance over the entries
    // ranks the rows: rows independent of one another (a colour or level,
    // the SpMV) are laid out by rank.  Only where a row sits changes, not
    // what it computes.
    // Lists filled by walking the rows in rank order are in rank order.
    const Locality loc = locality_rank(n, A, C);
    const std::vector<int>& by_rank = loc.order;

    // ---- device order and the sweep groups (as lists of original rows) ----
    std::vector<std::vector<int>> fwd_groups, bwd_groups;  // original row ids, launch order
    std::vector<int> order;                                // active rows in device order
    if (rainbow) {
        fwd_groups.assign(num_colors, {});
        for (int r : by_rank)
            if (active[r]) fwd_groups[F[r]].push_back(r);
        for (auto& g : fwd_groups) order.insert(order.end(), g.begin(), g.end());
        bwd_groups.assign(fwd_groups.rbegin(), fwd_groups.rend());
    } else if (sweeps) {
        std::vector<int> asc(n), desc(n);
        for (int i = 0; i < n; i++) asc[i] = i, desc[i] = n - 1 - i;
        int nf = 0, nb = 0;
        auto is_lower = [&](int orow, int k) { return C[k] < orow; };
        auto is_upper = [&](int orow, int k) { return C[k] > orow; };
        std::vector<int> lf = sweep_levels(n, A, C, active, asc, is_lower, &nf);
        std::vector<int> lb = sweep_levels(n, A, C, active, desc, is_upper, &nb);
        fwd_groups.assign(nf, {});
        bwd_groups.assign(nb, {});
        for (int r : by_rank)
            if (active[r]) {
                fwd_groups[lf[r]].push_back(r);
                bwd_groups[lb[r]].push_back(r);
            }
        for (auto& g : fwd_groups) order.insert(order.end(), g.begin(), g.end());
    } else {
        for (int r : by_rank)
            if (active[r]) order.push_back(r);
    }
    plan.n_active = (int)order.size();
    plan.perm = order;
    for (int r = 0; r < n; r++)
        if (!active[r]) plan.perm.push_back(r);
    std::vector<int> iperm(n + 1);
    for (int p = 0; p < n; p++) iperm[plan.perm[p]] = p;
    iperm[n] = n;

    plan.diag.assign(n, 0.0);
    if (aux)
        for (int p = 0; p < n; p++) plan.diag[p] = aux[plan.perm[p]];

    auto to_dev = [&](const std::vector<int>& orig) {
        std::vector<int> dev(orig.size());
        for (size_t i = 0; i < orig.size(); i++) dev[i] = iperm[orig[i]];
        return dev;
    };
    auto guarded = [&](int k) { return aux[C[k]] != 0.; };

    // ---- full part: SpMV of the iteration / residual of the restart ----
    // Its rows are independent, so they are listed in rank order, not in
    // device order: a range of slices then covers a piece of the mesh, whose
    // neighbours lie in a few short runs of each colour's rows (device order
    // groups rows by colour), and a GPU die that runs a contiguous range of
    // slices gathers them from its own cache instead of from all of each vector.
    //
    // Inside windows of kSortWindow rank-ordered rows, its rows are then
    // ordered by decreasing length (stable), so that a 64-row slice holds rows
    // of similar length and is padded little (SELL-C-sigma): for degree-2/3
    // Lagrange and Whitney matrices, whose vertex, edge and interior rows
    // differ in length, this cuts the stored entries of this part from
    // 1.16-1.73x to 1.00-1.07x the real ones; rows of equal length keep their
    // rank order.  (The colour groups are left in rank order: their parts are
    // padded little anyway, about 1.01x in 2D.)
    std::vector<int> active_rows;
    active_rows.reserve(plan.n_active);
    for (int r : by_rank)
        if (active[r]) active_rows.push_back(r);
    for (size_t w = 0; w < active_rows.size(); w += kSortWindow)
        std::stable_sort(active_rows.begin() + w,
                         active_rows.begin() + std::min(active_rows.size(), w + kSortWindow),
                         [&](int a, int b) { return A[a + 1] - A[a] > A[b + 1] - A[b]; });
    const std::vector<std::vector<int>> active_dev(1, to_dev(active_rows));
    if (eisen)  // solver/sparsesolver.cpp:806-807,1420-1421 skip columns with diagonal 0
        build_part(plan.full, active_dev, plan.perm, iperm, n, A, C, V,
                   [&](int, int k) { return guarded(k); });
    else
        build_part(plan.full, active_dev, plan.perm, iperm, n, A, C, V,
                   [](int, int) { return true; });

    if (!sweeps) return "";

    // ---- triangular parts ----
    auto lower_nat = [&](int orow, int k) { return C[k] < orow; };
    auto upper_nat = [&](int orow, int k) { return C[k] > orow; };
    auto lower_rb = [&](int orow, int k) { return F[C[k]] < F[orow]; };
    auto upper_rb = [&](int orow, int k) { return F[C[k]] > F[orow]; };

    std::vector<std::vector<int>> fwd_dev, bwd_dev;
    for (const auto& g : fwd_groups) fwd_dev.push_back(to_dev(g));
    for (const auto& g : bwd_groups) bwd_dev.push_back(to_dev(g));
    if (method == SFA_METHOD_SSOR) {
        build_part(plan.lower, fwd_dev, plan.perm, iperm, n, A, C, V, lower_nat);
        build_part(plan.upper, bwd_dev, plan.perm, iperm, n, A, C, V, upper_nat);
    } else if (method == SFA_METHOD_SSOR_EISENSTAT) {  // :824,903 and :886
        build_part(plan.lower, fwd_dev, plan.perm, iperm, n, A, C, V,
                   [&](int o, int k) { return lower_nat(o, k) && guarded(k); });
        build_part(plan.upper, bwd_dev, plan.perm, iperm, n, A, C, V,
                   [&](int o, int k) { return upper_nat(o, k) && guarded(k); });
    } else {
        build_part(plan.lower, fwd_dev, plan.perm, iperm, n, A, C, V, lower_rb);
        build_part(plan.upper, bwd_dev, plan.perm, iperm, n, A, C, V, upper_rb);
    }
    if (method == SFA_METHOD_EISENSTAT_RAINBOW) {
        // The restart's lower solve skips columns with diagonal 0 (:1448), the
        // iteration's does not (:1550-1552).
        plan.has_lower_restart = true;
        build_part(plan.lower_restart, fwd_dev, plan.perm, iperm, n, A, C, V,
                   [&](int o, int k) { return lower_rb(o, k) && guarded(k); });
    }
    return "";
}

}  // namespace sfa
