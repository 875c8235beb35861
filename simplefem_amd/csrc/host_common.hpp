// host_common.hpp -- host-side helpers shared by the library's sources:
// error type and HIP checks, device allocation, SELL parts in device memory.
#pragma once

#include <algorithm>
#include <type_traits>

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/simplefem_amd.h"
#include "kernels.hpp"
#include "plan.hpp"

namespace sfa {

// message of the calling thread's last failed C-ABI call
std::string& last_error();

// A system's device-pointer solves sum ||b|| for their tolerance in row
// order (the reference's bits) instead of by blocks; for the inner solves of
// the block solvers, whose right-hand sides are short.
void system_row_order_tolerance(sfa_system sys, bool on);


struct Error {
    int code;
    std::string msg;
};

#define SFA_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            int code_ = e_ == hipErrorOutOfMemory ? SFA_ERR_OUT_OF_MEMORY : SFA_ERR_HIP; \
            throw Error{code_, std::string(#call) + ": " + hipGetErrorString(e_)};     \
        }                                                                              \
    } while (0)

// Cached blocks (pool.cpp): device memory on the current device, pinned host
// memory.  release_block returns a block to the cache without waiting for
// the device, so its caller must first have synchronised the stream(s) whose
// queued work uses it.
void* device_block(size_t bytes);
void* pinned_block(size_t bytes);
void release_block(void* p);

template <class T>
T* dev_alloc(size_t count)
{
    if (count == 0) count = 1;
    return static_cast<T*>(device_block(count * sizeof(T)));
}

// No queued GPU work may still use p (see release_block).
inline void dev_free(void* p) { release_block(p); }

template <class T>
T* pinned_alloc(size_t count)
{
    if (count == 0) count = 1;
    return static_cast<T*>(pinned_block(count * sizeof(T)));
}

inline void pinned_free(void* p) { release_block(p); }

inline void ensure_device()
{
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
        throw Error{SFA_ERR_NO_DEVICE, "no HIP device available"};
}

template <typename T>
struct DevPartT {
    int* off = nullptr;
    int* rows = nullptr;
    int* cols = nullptr;
    T* vals = nullptr;
    int nslices = 0;
    std::vector<int> group_slice;
    std::vector<long long> group_entries, group_rows;
    std::vector<int> group_row0;
    std::vector<int> group_width;  // uniform entries per lane, or -1
    std::vector<int> group_base;   // first entry of each group
    long long stored = 0;
    long long entries = 0;
    long long nrows = 0;
    int zero_slot = 0;
    // rows and stored entries of the slices in each SpMV row-length bucket
    long long bucket_rows[kSpmvBuckets] = {};
    long long bucket_stored[kSpmvBuckets] = {};
    bool shares_structure = false;  // off, rows, cols belong to another part (share_from)

    void upload(const SellPart& p)
    {
        spmv_bucket_counts(p, bucket_rows, bucket_stored);
        nslices = p.nslices;
        zero_slot = p.zero_slot;
        group_slice = p.group_slice;
        group_entries = p.group_entries;
        group_rows = p.group_rows;
        group_row0 = p.group_row0;
        group_width = p.group_width;
        group_base.clear();
        for (size_t g = 0; g + 1 < p.group_slice.size(); g++)
            group_base.push_back(p.slice_off[p.group_slice[g]]);
        stored = (long long)p.vals.size();
        entries = p.entries;
        nrows = p.rows;
        off = dev_alloc<int>(p.slice_off.size());
        rows = dev_alloc<int>(p.slice_rows.size());
        cols = dev_alloc<int>(p.cols.size());
        vals = dev_alloc<T>(p.vals.size());
        SFA_HIP(hipMemcpy(off, p.slice_off.data(), p.slice_off.size() * sizeof(int), hipMemcpyHostToDevice));
        if (!p.slice_rows.empty())
            SFA_HIP(hipMemcpy(rows, p.slice_rows.data(), p.slice_rows.size() * sizeof(int), hipMemcpyHostToDevice));
        if (!p.cols.empty()) {
            SFA_HIP(hipMemcpy(cols, p.cols.data(), p.cols.size() * sizeof(int), hipMemcpyHostToDevice));
            if constexpr (std::is_same<T, double>::value) {
                SFA_HIP(hipMemcpy(vals, p.vals.data(), p.vals.size() * sizeof(double), hipMemcpyHostToDevice));
            } else {
                HostArray<T> tv(p.vals.size());
                const long long nv = (long long)p.vals.size();
#pragma omp parallel for schedule(static)
                for (long long i = 0; i < nv; i++) tv[i] = (T)p.vals[i];
                SFA_HIP(hipMemcpy(vals, tv.data(), tv.size() * sizeof(T), hipMemcpyHostToDevice));
            }
        }
    }
    // The part of src (same slices, groups and padding) in precision T: the
    // offsets, row ids and columns are src's arrays, which must outlive this
    // part; the values are src's, converted on the device on stream s.
    void share_from(const DevPartT<double>& src, hipStream_t s)
    {
        static_assert(std::is_same<T, float>::value, "an FP32 copy of an FP64 part");
        nslices = src.nslices;
        zero_slot = src.zero_slot;
        group_slice = src.group_slice;
        group_entries = src.group_entries;
        group_rows = src.group_rows;
        group_row0 = src.group_row0;
        group_width = src.group_width;
        group_base = src.group_base;
        stored = src.stored;
        entries = src.entries;
        nrows = src.nrows;
        for (int b = 0; b < kSpmvBuckets; b++) {
            bucket_rows[b] = src.bucket_rows[b];
            bucket_stored[b] = src.bucket_stored[b];
        }
        off = src.off;
        rows = src.rows;
        cols = src.cols;
        shares_structure = true;
        vals = dev_alloc<T>((size_t)stored);
        constexpr long long kChunk = 1LL << 30;  // the conversion kernel counts in int
        for (long long i = 0; i < stored; i += kChunk)
            launch_to_float(s, vals + i, src.vals + i, (int)std::min(kChunk, stored - i));
        SFA_HIP(hipGetLastError());
    }
    void release()
    {
        if (!shares_structure) {
            dev_free(off);
            dev_free(rows);
            dev_free(cols);
        }
        dev_free(vals);
        off = rows = cols = nullptr;
        vals = nullptr;
        shares_structure = false;
    }
    SellViewT<T> view() const { return SellViewT<T>{off, rows, cols, vals, zero_slot}; }
    // the bucket holding most of the part's stored entries
    int main_bucket() const
    {
        int b = 0;
        for (int i = 1; i < kSpmvBuckets; i++)
            if (bucket_stored[i] > bucket_stored[b]) b = i;
        return b;
    }
    int ngroups() const { return (int)group_slice.size() - 1; }
    int group_begin(int g) const { return group_slice[g]; }
    int group_count(int g) const { return group_slice[g + 1] - group_slice[g]; }
    int row0(int g) const { return group_row0[g]; }
    int width(int g) const { return group_width[g]; }
    int base(int g) const { return group_base[g]; }
    int nrows_of(int g) const { return (int)group_rows[g]; }
    // Algorithmic bytes of one launch over group g: (sizeof(T) + 4) B per
    // entry (value + column), 4 B per row id, `per_row` B of vector traffic
    // per row (callers give it for T = double and it scales with sizeof(T)).
    double bytes(int g, double per_row) const
    {
        return (sizeof(T) + 4.0) * group_entries[g] + (4.0 + per_row * sizeof(T) / 8.0) * group_rows[g];
    }
    double bytes_all(double per_row) const
    {
        return (sizeof(T) + 4.0) * entries + (4.0 + per_row * sizeof(T) / 8.0) * nrows;
    }
};
using DevPart = DevPartT<double>;

// Per-launch HIP events of a profiled solve.
struct Profiler {
    std::vector<hipEvent_t> pool;
    std::vector<int> klass;
    std::vector<double> bytes;
    ~Profiler()
    {
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
    hipEvent_t event(size_t i)
    {
        while (pool.size() <= i) {
            hipEvent_t e;
            SFA_HIP(hipEventCreate(&e));
            pool.push_back(e);
        }
        return pool[i];
    }
};


// Runs fn, turning exceptions into the C-ABI status and message.
template <class Fn>
int guarded(Fn&& fn)
{
    try {
        fn();
        last_error().clear();
        return SFA_OK;
    } catch (const Error& e) {
        last_error() = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        last_error() = "host allocation failed";
        return SFA_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        last_error() = e.what();
        return SFA_ERR_INVALID;
    }
}

}  // namespace sfa
