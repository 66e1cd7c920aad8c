// array.h -- Array4-compatible device views, boxes and launch helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "camr_hydro.h"

namespace camr {

struct Box {
    int lo[3], hi[3];
    __host__ __device__ int len(int d) const { return hi[d] - lo[d] + 1; }
    __host__ __device__ int64_t npts() const {
        return (int64_t)len(0) * len(1) * len(2);
    }
    __host__ __device__ bool ok() const {
        return hi[0] >= lo[0] && hi[1] >= lo[1] && hi[2] >= lo[2];
    }
    __host__ __device__ bool contains(int i, int j, int k) const {
        return i >= lo[0] && i <= hi[0] && j >= lo[1] && j <= hi[1] &&
               k >= lo[2] && k <= hi[2];
    }
    __host__ __device__ bool contains(const Box &b) const {
        for (int d = 0; d < 3; ++d)
            if (b.lo[d] < lo[d] || b.hi[d] > hi[d]) return false;
        return true;
    }
};

inline Box make_box(const int32_t lo[3], const int32_t hi[3]) {
    Box b;
    for (int d = 0; d < 3; ++d) { b.lo[d] = lo[d]; b.hi[d] = hi[d]; }
    return b;
}
__host__ __device__ inline Box grow(Box b, int n) {
    for (int d = 0; d < 3; ++d) { b.lo[d] -= n; b.hi[d] += n; }
    return b;
}
__host__ __device__ inline Box grow_dir(Box b, int d, int n) { b.lo[d] -= n; b.hi[d] += n; return b; }
__host__ __device__ inline Box grow_hi(Box b, int d, int n) { b.hi[d] += n; return b; }
__host__ __device__ inline Box nodes(Box b, int d) { b.hi[d] += 1; return b; }

// Device view with Array4 indexing.
struct Arr {
    double *p;
    int64_t js, ks, ns;
    int lo0, lo1, lo2;
    __host__ __device__ double &operator()(int i, int j, int k, int n) const {
        return p[(int64_t)(i - lo0) + (int64_t)(j - lo1) * js +
                 (int64_t)(k - lo2) * ks + (int64_t)n * ns];
    }
    __host__ __device__ double *ptr(int i, int j, int k) const {
        return p + ((int64_t)(i - lo0) + (int64_t)(j - lo1) * js +
                    (int64_t)(k - lo2) * ks);
    }
};

inline Arr view(const camr_array4 &a) {
    Arr v;
    v.p = a.p; v.js = a.jstride; v.ks = a.kstride; v.ns = a.nstride;
    v.lo0 = a.lo[0]; v.lo1 = a.lo[1]; v.lo2 = a.lo[2];
    return v;
}
inline Box box_of(const camr_array4 &a) { return make_box(a.lo, a.hi); }

// A dense scratch array over box b with nc components (Array4 layout).
inline Arr dense(double *p, const Box &b) {
    Arr v;
    v.p = p;
    v.js = b.len(0);
    v.ks = (int64_t)b.len(0) * b.len(1);
    v.ns = b.npts();
    v.lo0 = b.lo[0]; v.lo1 = b.lo[1]; v.lo2 = b.lo[2];
    return v;
}

// Index decomposition for kernels that cover a box with one thread per
// point, x fastest (coalesced along the unit-stride Array4 axis).
struct Range {
    int lo0, lo1, lo2, n0, n1;
    int64_t n;
};
__host__ __device__ inline Range range_of(const Box &b) {
    Range r;
    r.lo0 = b.lo[0]; r.lo1 = b.lo[1]; r.lo2 = b.lo[2];
    r.n0 = b.len(0); r.n1 = b.len(1);
    r.n = b.npts();
    return r;
}
// The blocks of one launch over a table of boxes of uneven sizes, in one
// dimension: entry u owns blocks [off[u], off[u + 1]).  n = 0: a 2-D grid,
// blockIdx.y the entry and blockIdx.x its block (for longer tables).
constexpr int kFlatEntries = 64;
struct FlatBlocks {
    int n;
    int off[kFlatEntries + 1];
};
// entry u and block lb within it of this block
__device__ __forceinline__ void flat_entry(const FlatBlocks &f, int &u, int &lb) {
    if (f.n == 0) {
        u = blockIdx.y;
        lb = blockIdx.x;
        return;
    }
    const int b = blockIdx.x;
    int lo = 0, hi = f.n;  // the last u with off[u] <= b (entries of no block skipped)
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (f.off[mid] <= b) lo = mid;
        else hi = mid;
    }
    u = lo;
    lb = b - f.off[lo];
}
// Host: the flattened blocks of n entries with cnt(u) blocks each, or a 2-D
// grid of the largest count when n exceeds kFlatEntries; returns the grid.
template <typename C>
inline dim3 flat_blocks(int n, C cnt, FlatBlocks &f) {
    if (n <= kFlatEntries) {
        f.n = n;
        f.off[0] = 0;
        for (int u = 0; u < n; ++u) f.off[u + 1] = f.off[u] + (int)cnt(u);
        return dim3((unsigned)f.off[n], 1);
    }
    f.n = 0;
    int64_t m = 0;
    for (int u = 0; u < n; ++u) m = m > cnt(u) ? m : cnt(u);
    return dim3((unsigned)m, (unsigned)n);
}

__device__ __forceinline__ bool range_point(const Range &r, int64_t t, int &i,
                                            int &j, int &k) {
    if (t >= r.n) return false;
    const int64_t pl = (int64_t)r.n0 * r.n1;
    const int kk = (int)(t / pl);
    const int64_t rem = t - (int64_t)kk * pl;
    const int jj = (int)(rem / r.n0);
    i = r.lo0 + (int)(rem - (int64_t)jj * r.n0);
    j = r.lo1 + jj;
    k = r.lo2 + kk;
    return true;
}

constexpr int kBlock = 256;

// Workgroups are dealt round-robin to the 8 XCDs of an MI355X, each with its
// own L2.  Renumber them so that XCD x owns a contiguous run of logical
// blocks: neighbouring rows / planes of a stencil then share one L2.
constexpr unsigned kXCD = 8;
__device__ __forceinline__ int64_t logical_thread() {
#ifndef CAMR_XCD_REMAP
    return (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
#endif
    const unsigned nb = gridDim.x, b = blockIdx.x;
    const unsigned q = nb / kXCD, r = nb % kXCD, x = b % kXCD;
    const unsigned lb = x * q + (x < r ? x : r) + b / kXCD;
    return (int64_t)lb * blockDim.x + threadIdx.x;
}
inline unsigned grid_for(int64_t n, int block = kBlock) {
    int64_t g = (n + block - 1) / block;
    return (unsigned)(g > 0 ? g : 1);
}

// Simple per-device scratch pool: one growing allocation, bump-allocated
// per call, released at the end of the call.  Calls on one pool must be
// serialised (the level object owns its own pool).
class Scratch {
  public:
    Scratch() = default;
    ~Scratch();
    // Ensure capacity of `bytes`; returns false on allocation failure.
    bool reserve(size_t bytes);
    void reset() { used_ = 0; }
    double *take(size_t ndouble);
    size_t capacity() const { return cap_; }
    double *base() const { return (double *)base_; }

  private:
    char *base_ = nullptr;
    size_t cap_ = 0;
    size_t used_ = 0;
};

// thread-local error message used by camr_last_error()
void set_error(const char *fmt, ...);

} // namespace camr
