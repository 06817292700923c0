// Complex division in the planes of its operands, shared by the shift LU
// (lhl_shift.hip) and the complex-J kernels (lhl_cplx.hip): Baudin & Smith,
// "A Robust Complex Division in Scilab" (2012), the unscaled fast path, the
// division Julia's ComplexF64 `/` performs for operands of ordinary range.
// No contraction: every product and sum is rounded on its own.
#pragma once

#include <hip/hip_runtime.h>

namespace lhl {

template <typename T> __device__ __forceinline__ T cabs1(T x) { return x < 0 ? -x : x; }

template <typename T>
__device__ __forceinline__ void cdiv1(T a, T b, T c, T d, T &p, T &q) {
#pragma clang fp contract(off)
    const T r = d / c;
    const T t = (T)1 / (c + d * r);
    if (r != (T)0) {
        const T br = b * r;
        p = (br != (T)0) ? (a + br) * t : a * t + (b * t) * r;
        const T ar = -a * r;
        q = (ar != (T)0) ? (b + ar) * t : b * t + (-a * t) * r;
    } else {
        p = (a + d * (b / c)) * t;
        q = (b + d * (-a / c)) * t;
    }
}
template <typename T>
__device__ __forceinline__ void cdiv(T a, T b, T c, T d, T &p, T &q) {
    if (cabs1(d) <= cabs1(c)) cdiv1(a, b, c, d, p, q);
    else { cdiv1(b, a, d, c, p, q); q = -q; }
}

// cdiv with selects in place of its branches: each chosen result is computed
// by the same operations as cdiv's branch, so the two agree bit for bit.
template <typename T>
__device__ __forceinline__ void cdiv_sel(T a, T b, T c, T d, T &p, T &q) {
#pragma clang fp contract(off)
    const bool sw = cabs1(d) > cabs1(c);
    const T A = sw ? b : a, B = sw ? a : b, C = sw ? d : c, D = sw ? c : d;
    const T r = D / C;
    const T t = (T)1 / (C + D * r);
    const T br = B * r;
    const T p1 = (A + br) * t, p2 = A * t + (B * t) * r, p3 = (A + D * (B / C)) * t;
    const T ar = -A * r;
    const T q1 = (B + ar) * t, q2 = B * t + (-A * t) * r, q3 = (B + D * (-A / C)) * t;
    p = (r != (T)0) ? ((br != (T)0) ? p1 : p2) : p3;
    const T qq = (r != (T)0) ? ((ar != (T)0) ? q1 : q2) : q3;
    q = sw ? -qq : qq;
}

}  // namespace lhl
