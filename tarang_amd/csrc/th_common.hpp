// th_common.hpp -- shared device/host definitions for the gfx950 pseudo-spectral kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace th {

template <typename R> struct alignas(2 * sizeof(R)) cplx { R x, y; };
using c32 = cplx<float>;
using c64 = cplx<double>;

template <typename R> __host__ __device__ __forceinline__ cplx<R> mk(R a, R b) { cplx<R> c; c.x = a; c.y = b; return c; }
template <typename R> __host__ __device__ __forceinline__ cplx<R> operator+(cplx<R> a, cplx<R> b) { return mk<R>(a.x + b.x, a.y + b.y); }
template <typename R> __host__ __device__ __forceinline__ cplx<R> operator-(cplx<R> a, cplx<R> b) { return mk<R>(a.x - b.x, a.y - b.y); }
template <typename R> __host__ __device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> b) { return mk<R>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
template <typename R> __host__ __device__ __forceinline__ cplx<R> cmulc(cplx<R> a, cplx<R> b) { return mk<R>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y); } // a*conj(b)
template <typename R> __host__ __device__ __forceinline__ cplx<R> cscale(cplx<R> a, R s) { return mk<R>(a.x * s, a.y * s); }
template <typename R> __host__ __device__ __forceinline__ cplx<R> conjg(cplx<R> a) { return mk<R>(a.x, -a.y); }
template <typename R> __host__ __device__ __forceinline__ cplx<R> mul_i(cplx<R> a) { return mk<R>(-a.y, a.x); }      // i*a
// s + i w d in four fused multiply-adds (s + mul_i(cmul(w, d)) rounds the product first: two more instructions)
template <typename R> __device__ __forceinline__ cplx<R> add_iwd(cplx<R> s, cplx<R> w, cplx<R> d)
{
    return mk<R>(fma(-w.x, d.y, fma(-w.y, d.x, s.x)), fma(w.x, d.x, fma(-w.y, d.y, s.y)));
}
template <typename R> __host__ __device__ __forceinline__ cplx<R> mul_mi(cplx<R> a) { return mk<R>(a.y, -a.x); }     // -i*a
template <typename R> __host__ __device__ __forceinline__ cplx<R> czero() { return mk<R>(R(0), R(0)); }
// base + a 32-bit byte offset: with a wave-uniform base this is the SGPR-base + 32-bit VGPR-offset form of a
// global access (no 64-bit address per lane).  Offsets within one component of a fused-state array stay below
// 4 GB for every supported grid (N <= 512 per axis).
template <typename T> __device__ __forceinline__ T* at_byte(T* base, unsigned off)
{
    return (T*)((char*)base + off);
}
// Non-temporal store of one complex value: the pipeline buffers are written once and read by the next pass
// only after every workgroup of this one has finished, and at the grid sizes that matter they exceed the
// last-level cache, so their lines need not be kept on the way out.
template <typename R> __device__ __forceinline__ void st_nt(cplx<R>* p, cplx<R> v)
{
    typedef R v2 __attribute__((ext_vector_type(2)));
    v2 x;
    x.x = v.x;
    x.y = v.y;
    __builtin_nontemporal_store(x, (v2*)p);
}
// Non-temporal load of one complex value (an input line read once by this pass)
template <typename R> __device__ __forceinline__ cplx<R> ld_nt(const cplx<R>* p)
{
    typedef R v2 __attribute__((ext_vector_type(2)));
    const v2 x = __builtin_nontemporal_load((const v2*)p);
    return mk<R>(x.x, x.y);
}

// Wavenumber helpers (trunk/lib/four/four_inline.h:88-138, trunk/lib/scft/scft_inline.h:86)
__host__ __device__ __forceinline__ int kx_four(int l1, int N1) { return (l1 <= N1 / 2) ? l1 : l1 - N1; }
__host__ __device__ __forceinline__ int ky_of(int l2, int N2) { return (l2 <= N2 / 2) ? l2 : l2 - N2; }

// Dealias_array's 2/3 rule (universal_misc.cc:118-159): true when the mode survives.
__host__ __device__ __forceinline__ bool kept_mode(int basis, int dealias, int N1, int N2, int N3, int l1, int l2, int l3)
{
    if (!dealias) return true;
    int kx = basis == 1 ? l1 : kx_four(l1, N1);
    if (N1 >= 3) {
        int ak = kx < 0 ? -kx : kx;
        if (basis == 0 ? (ak >= N1 / 3) : (ak >= 2 * N1 / 3)) return false;
    }
    if (N2 >= 3 && l2 >= N2 / 3 && l2 <= 2 * N2 / 3 + 1) return false;
    if (N3 >= 3 && l3 >= N3 / 3 && l3 <= N3 / 2) return false;
    return true;
}

}  // namespace th
