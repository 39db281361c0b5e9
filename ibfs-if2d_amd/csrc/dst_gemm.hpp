// dst_gemm.hpp -- 1-D DST-I passes as dense products on the matrix cores.
//
// The DST-I of length N-1 (src/myfft.f90:41, FFTW RODFT00, unnormalised) is
// the product with the symmetric matrix S(a, b) = 2 sin(pi (a+1)(b+1) / N),
// a, b = 0..N-2.  Over a vertex field X ((m-1) x (n-1), i fastest, i.e. a
// row-major (n-1) x (m-1) matrix) the pass along i is Y = X S_m and the pass
// along j is Y = S_n X.  Both are C = A B with row-major operands.
//
// This is the dense alternative to the FFT passes of dst_core.hpp: 2 (N-1)^3
// flops per pass against O(N^2 log N), so it can only win where the FFT
// pass is latency-bound, i.e. at small N.  It is measured against them
// (ibfs_time_dst_gemm) and is not used by the time step.
//
// Kernel: one 16 x 16 tile of C per workgroup, KS wave64s that split the
// K dimension, each accumulating its slice with v_mfma_{f64,f32}_16x16x4
// (operand lane layout: A(row l%16, k l/16), B(k l/16, col l%16); result
// lane l, register r holds C(4 (l/16) + r, l%16) for f32 and
// C(l/16 + 4 r, l%16) for f64, as measured on gfx950).  Operand loads run one
// chunk of KC MFMAs ahead of the arithmetic (two register buffers); the KS
// partial tiles are summed through LDS and stored row by row.
#pragma once
#include <hip/hip_runtime.h>

namespace ibfs {

typedef double v4d_t __attribute__((ext_vector_type(4)));
typedef float v4f_t __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma16;
template <>
struct Mfma16<double> {
  using acc_t = v4d_t;
  static __device__ __forceinline__ acc_t step(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mfma16<float> {
  using acc_t = v4f_t;
  static __device__ __forceinline__ acc_t step(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

constexpr int kGemmKC = 8;  // MFMAs per operand chunk

// C (M x N) = A (M x K) B (K x N), row-major with leading dimensions.
// grid (ceil(N/16), ceil(M/16), batch), block 64*KS; batch strides sa, sb,
// sc (0 for an operand shared by the batch).
template <typename T, int KS>
__global__ __launch_bounds__(64 * KS) void k_gemm16(const T* __restrict__ A, int lda, long sa,
                                                    const T* __restrict__ B, int ldb, long sb, T* __restrict__ C,
                                                    int ldc, long sc, int M, int N, int K) {
  using MF = Mfma16<T>;
  using acc_t = typename MF::acc_t;
  __shared__ T part[KS][256];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int r0 = blockIdx.y * 16, c0 = blockIdx.x * 16;
  const T* Ab = A + blockIdx.z * sa;
  const T* Bb = B + blockIdx.z * sb;
  // K slice of this wave, in steps of 4 (one MFMA)
  const int nstep = (K + 3) / 4;
  const int per = (nstep + KS - 1) / KS;
  const int s0 = w * per, s1 = min(nstep, s0 + per);
  const int ar = r0 + (lane & 15), bc = c0 + (lane & 15), kl = lane >> 4;
  const bool arow = ar < M, bcol = bc < N;
  auto load = [&](int s, T& a, T& b) {
    const int k = 4 * s + kl;
    const bool kin = s < s1 && k < K;
    a = (kin && arow) ? Ab[(long)ar * lda + k] : T(0);
    b = (kin && bcol) ? Bb[(long)k * ldb + bc] : T(0);
  };
  acc_t acc = {T(0), T(0), T(0), T(0)};
  T a0[kGemmKC], b0[kGemmKC], a1[kGemmKC], b1[kGemmKC];
#pragma unroll
  for (int q = 0; q < kGemmKC; q++) load(s0 + q, a0[q], b0[q]);
  for (int s = s0; s < s1; s += 2 * kGemmKC) {
#pragma unroll
    for (int q = 0; q < kGemmKC; q++) load(s + kGemmKC + q, a1[q], b1[q]);
#pragma unroll
    for (int q = 0; q < kGemmKC; q++) acc = MF::step(a0[q], b0[q], acc);
    if (s + kGemmKC >= s1) break;
#pragma unroll
    for (int q = 0; q < kGemmKC; q++) load(s + 2 * kGemmKC + q, a0[q], b0[q]);
#pragma unroll
    for (int q = 0; q < kGemmKC; q++) acc = MF::step(a1[q], b1[q], acc);
  }
#pragma unroll
  for (int r = 0; r < 4; r++) part[w][MF::row(lane, r) * 16 + (lane & 15)] = acc[r];
  __syncthreads();
  for (int e = threadIdx.x; e < 256; e += 64 * KS) {
    T v = part[0][e];
#pragma unroll
    for (int q = 1; q < KS; q++) v += part[q][e];
    const int r = r0 + (e >> 4), c = c0 + (e & 15);
    if (r < M && c < N) C[blockIdx.z * sc + (long)r * ldc + c] = v;
  }
}

}  // namespace ibfs
