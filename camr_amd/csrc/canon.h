// canon.h -- the canonical per-box layout of the fused engine's arrays and
// their buffer-addressed device views.
//
// Every intermediate of one box (Sborder copy, primitives, face states,
// fluxes, divu) is an Array4-shaped array over the same index box
// C = grow(bx, 4) extended by one node on the high side of each direction,
// so one 32-bit byte offset per thread addresses all of them: components
// and neighbours differ by uniform (scalar-register) offsets.  Loads and
// stores go through buffer resources (base in SGPRs, 32-bit VGPR offset),
// which removes the 64-bit address arithmetic of plain pointers and
// bounds each access to its array (out-of-range loads read 0, stores drop).
#pragma once

#include "array.h"

namespace camr {

// Layout of C in elements.  Host side: built by canon_of(); device side:
// passed by value.
struct Canon {
    int lo0, lo1, lo2;  // C.lo
    int js, ks;         // x-row and xy-plane strides
    unsigned cs;        // elements per component (C.npts())
    __device__ __forceinline__ unsigned off(int i, int j, int k) const {
        return (unsigned)((i - lo0) + (j - lo1) * js + (k - lo2) * ks) * 8u;
    }
    __device__ __forceinline__ unsigned sy() const { return (unsigned)js * 8u; }
    __device__ __forceinline__ unsigned sz() const { return (unsigned)ks * 8u; }
};

inline Box canon_box(const Box &bx) {
    Box c = grow(bx, CAMR_NGROW);
    for (int d = 0; d < 3; ++d) c.hi[d] += 1;
    return c;
}
inline Canon canon_of(const Box &bx) {
    const Box c = canon_box(bx);
    Canon L;
    L.lo0 = c.lo[0]; L.lo1 = c.lo[1]; L.lo2 = c.lo[2];
    L.js = c.len(0);
    L.ks = c.len(0) * c.len(1);
    L.cs = (unsigned)c.npts();
    return L;
}
// Largest component count of one canonical array (the compact primitives).
constexpr int kCanonMaxComp = 12;
// A box fits the canonical scheme when every array's byte size fits the
// 32-bit range of a buffer resource.
inline bool canon_fits(const Box &bx) {
    return (double)canon_box(bx).npts() * kCanonMaxComp * 8.0 < 4294967295.0;
}
// The same layout as an Array4 view (for kernels that take Arr).
inline Arr canon_arr(double *p, const Box &bx) { return dense(p, canon_box(bx)); }

// Device view of one canonical array with nc components.
struct CView {
    __amdgpu_buffer_rsrc_t r;
    unsigned cs8;  // component stride in bytes
    Canon L;
    __device__ __forceinline__ double ld(unsigned v, int n) const {
        return __builtin_bit_cast(
            double, __builtin_amdgcn_raw_buffer_load_b64(r, v, (unsigned)n * cs8, 0));
    }
    __device__ __forceinline__ void st(unsigned v, int n, double x) const {
        typedef unsigned u2 __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, x), r, v,
                                              (unsigned)n * cs8, 0);
    }
    // Array4-style read, for shared helpers templated on the accessor
    __device__ __forceinline__ double operator()(int i, int j, int k, int n) const {
        return ld(L.off(i, j, k), n);
    }
};

// The base and sizes are wave-uniform by construction; readfirstlane makes
// that explicit so that the descriptor always lives in scalar registers
// (a descriptor the compiler believes divergent costs a loop per access).
__device__ __forceinline__ CView cview(const Canon &L, const double *p, int nc) {
    const uint64_t a = (uint64_t)p;
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)a);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(a >> 32));
    const unsigned cs = (unsigned)__builtin_amdgcn_readfirstlane((int)L.cs);
    CView v;
    v.r = __builtin_amdgcn_make_buffer_rsrc((void *)(((uint64_t)hi << 32) | lo), 0,
                                            (int)((unsigned)nc * cs * 8u), 0x00020000);
    v.cs8 = cs * 8u;
    v.L = L;
    return v;
}

} // namespace camr
