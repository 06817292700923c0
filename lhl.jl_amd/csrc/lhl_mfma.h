// 64x64 output tiles on the CDNA matrix cores (gfx950), FP32 and FP64.
//
// v_mfma_{f32,f64}_16x16x4: lane l supplies A[l%16][l/16] of a 16x4 A block
// and B[l/16][l%16] of a 4x16 B block.  The 16x16 result lands in four
// registers per lane: C[4*(l/16) + r][l%16] for f32 and C[4*r + l/16][l%16]
// for f64 (checked on the hardware with a one-wave probe).
//
// A workgroup of 256 threads (4 waves) owns a 64x64 tile; wave w computes the
// 32x32 quadrant (rows 32*(w&1), columns 32*(w>>1)) as 2x2 blocks.  Operands
// are staged in LDS k-major: As[k][i] = A(i, k), Bs[k][j] = B(k, j), so each
// group of 16 lanes reads 16 consecutive words.
#pragma once

#include <hip/hip_runtime.h>

namespace lhl {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
template <> struct Mfma<float> {
    typedef f32x4 acc_t;
    static __device__ __forceinline__ acc_t zero() { return acc_t{0.f, 0.f, 0.f, 0.f}; }
    static __device__ __forceinline__ acc_t mac(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};
template <> struct Mfma<double> {
    typedef f64x4 acc_t;
    static __device__ __forceinline__ acc_t zero() { return acc_t{0.0, 0.0, 0.0, 0.0}; }
    static __device__ __forceinline__ acc_t mac(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * r + (lane >> 4); }
};

constexpr int kTilePad = 4;   // LDS row padding of the staged operands

// Four consecutive elements through one (float) or two (double) vector loads
// that need only element alignment.
template <typename T> struct V4T;
template <> struct V4T<float> { typedef float type __attribute__((ext_vector_type(4), aligned(4))); };
template <> struct V4T<double> { typedef double type __attribute__((ext_vector_type(4), aligned(8))); };
template <typename T> using V4 = typename V4T<T>::type;
template <typename T> __device__ __forceinline__ V4<T> ld4u(const T *p) { return *(const V4<T> *)p; }

// acc[bi][bj] += As[0:kk, quadrant rows] * Bs[0:kk, quadrant cols]; kk % 4 == 0.
template <typename T>
__device__ __forceinline__ void mfma_tile64(const T (*As)[64 + kTilePad], const T (*Bs)[64 + kTilePad], int kk,
                                            typename Mfma<T>::acc_t acc[2][2]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wr = (w & 1) * 32, wc = (w >> 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
#pragma unroll 4
    for (int k0 = 0; k0 < kk; k0 += 4) {
        const int k = k0 + lk;
        const T a0 = As[k][wr + li], a1 = As[k][wr + 16 + li];
        const T b0 = Bs[k][wc + li], b1 = Bs[k][wc + 16 + li];
        acc[0][0] = Mfma<T>::mac(a0, b0, acc[0][0]);
        acc[0][1] = Mfma<T>::mac(a0, b1, acc[0][1]);
        acc[1][0] = Mfma<T>::mac(a1, b0, acc[1][0]);
        acc[1][1] = Mfma<T>::mac(a1, b1, acc[1][1]);
    }
}

// Visit every element of a tile held in acc: f(row, col, value) with row/col in [0, 64).
template <typename T, typename F>
__device__ __forceinline__ void mfma_tile64_each(const typename Mfma<T>::acc_t acc[2][2], F f) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wr = (w & 1) * 32, wc = (w >> 1) * 32;
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int bj = 0; bj < 2; bj++)
#pragma unroll
            for (int r = 0; r < 4; r++) f(wr + 16 * bi + Mfma<T>::row(lane, r), wc + 16 * bj + (lane & 15), acc[bi][bj][r]);
}

template <typename T>
__device__ __forceinline__ void mfma_tile64_zero(typename Mfma<T>::acc_t acc[2][2]) {
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int bj = 0; bj < 2; bj++) acc[bi][bj] = Mfma<T>::zero();
}

}  // namespace lhl
