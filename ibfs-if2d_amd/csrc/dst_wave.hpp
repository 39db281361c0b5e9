// dst_wave.hpp -- register-blocked DST-I passes for 2N = 256, 512, 1024.
//
// The size-specialised passes of dst_core.hpp run one radix-4 butterfly per
// thread and stage, with an LDS round trip and a workgroup barrier between
// any two stages (five for 2N = 1024).  Here one FFT of length N2 = 16 TPF
// is held by TPF lanes of one wave, 16 values per lane, and two consecutive
// radix-4 Stockham stages run in registers: lane t owns the four butterflies
// t + TPF r' of stage 0, whose outputs are exactly the inputs of the four
// butterflies 4t + k' of stage 1; after the pair the lane holds the 16
// consecutive places 16t..16t+15.  One LDS exchange (wave-local, no
// workgroup barrier) brings every lane back to the places t + TPF q, the
// next pair (stages 2 and 3) runs the same way, and for 2N = 512 / 1024 a
// second exchange feeds the last stage (radix 2 / radix 4, ns = 256), whose
// outputs land on the lane's own places t + TPF q again.  So a lane reads
// its inputs, and writes its outputs, at the places t + TPF q of the
// (oddly extended) sequence: the first 8 straight from the loader, the
// mirrored 8 (z_{N2-p} = -z_p) from lane TPF - t by a cross-lane read.
//
// Every butterfly is the one the Stockham stage of dst_core.hpp computes:
// the same inputs, the same twiddle (the per-stage table entry w, with w^2
// and w^3 formed by the same complex products) and the same dft_small, so
// the results are bit-identical to the LDS passes for every tile shape.
//
// Operators: the load / store / store_pre / mid / mid_coef protocol of the
// LDS passes (no side work, gathered loaders or side stores).  A workgroup
// is W waves; FFT p of the workgroup (P = W 64 / TPF of them) transforms
// local sequences 2p and 2p+1, as there.
#pragma once
#include "dst_core.hpp"

namespace ibfs {
namespace wave {

// Op::kNoWave: a loader with several loads per element (16 elements per
// lane here would keep too many of them in flight), run on the LDS passes
template <class Op, class = void>
struct op_no_wave : std::false_type {};
template <class Op>
struct op_no_wave<Op, std::void_t<decltype(Op::kNoWave)>> : std::bool_constant<Op::kNoWave> {};
// operators the register-blocked passes can run
template <class Op>
constexpr bool kOpOk =
    !op_side<Op>::value && !op_gather<Op>::value && !op_load_post<Op>::value && !op_no_wave<Op>::value;

// LDS place of value p of one FFT: one pad entry per 16 values, so that the
// 16 consecutive values a lane writes after a stage pair start on
// different banks for neighbouring lanes
__device__ __forceinline__ int lx(int p) { return p + (p >> 4); }
template <int N2>
constexpr int kLdsPerFft = N2 + N2 / 16;

// the exchanges are wave-local: every value is written and read back by
// lanes of one wave (the workgroup's other waves own other FFTs)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// v[1..3] times w, w^2, w^3 (w^2 = w w, w^3 = w^2 w, as tw_load forms them)
template <typename T>
__device__ __forceinline__ void twiddle4(cplx<T>* v, cplx<T> w) {
  const cplx<T> w2 = cmul(w, w);
  const cplx<T> w3 = cmul(w2, w);
  v[1] = cmul(v[1], w);
  v[2] = cmul(v[2], w2);
  v[3] = cmul(v[3], w3);
}

// Twiddles of lane t (place in its FFT), loaded once at kernel start:
// stage 1 (ns 4): k' = 0..3; stage 2 (ns 16): t mod 16; stage 3 (ns 64):
// t mod 16 + 16 r2; last stage (ns 256): t + TPF u.
template <typename T, int N2>
struct Tw {
  static constexpr int TPF = N2 / 16;
  static constexpr spec::Sched sc = spec::make_sched(N2);
  static constexpr int NLAST = N2 == 1024 ? 4 : N2 == 512 ? 8 : 0;
  cplx<T> w1[4], w2, w3[4], wl[NLAST > 0 ? NLAST : 1];
  __device__ void load(const cplx<T>* __restrict__ tw, int t) {
#pragma unroll
    for (int k = 0; k < 4; k++) w1[k] = tw[sc.toff[1] + k];
    w2 = tw[sc.toff[2] + (t & 15)];
#pragma unroll
    for (int r = 0; r < 4; r++) w3[r] = tw[sc.toff[3] + (t & 15) + 16 * r];
#pragma unroll
    for (int u = 0; u < NLAST; u++) wl[u] = tw[sc.toff[4] + t + TPF * u];
  }
};

// Stages 0 and 1 from the places t + TPF q (v[q]) to the places 16 t + q
// (o[q]).
template <typename T>
__device__ __forceinline__ void pair_a(const cplx<T>* v, cplx<T>* o, const cplx<T>* w1) {
  cplx<T> a[4][4];
#pragma unroll
  for (int r = 0; r < 4; r++) {  // stage 0 (ns 1, no twiddles): butterfly t + TPF r
    a[r][0] = v[r];
    a[r][1] = v[r + 4];
    a[r][2] = v[r + 8];
    a[r][3] = v[r + 12];
    dft_small<T, 4>(a[r]);
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {  // stage 1 (ns 4): butterfly 4 t + k
    cplx<T> u[4] = {a[0][k], a[1][k], a[2][k], a[3][k]};
    twiddle4(u, w1[k]);
    dft_small<T, 4>(u);
#pragma unroll
    for (int r = 0; r < 4; r++) o[k + 4 * r] = u[r];
  }
}

// Stages 2 and 3 from the places t + TPF q (v[q]) to the places
// 256 (t / 16) + t mod 16 + 16 q (o[q]).
template <typename T>
__device__ __forceinline__ void pair_b(const cplx<T>* v, cplx<T>* o, cplx<T> w2, const cplx<T>* w3) {
  cplx<T> a[4][4];
#pragma unroll
  for (int r = 0; r < 4; r++) {  // stage 2 (ns 16): butterfly t + TPF r, k = t mod 16
    a[r][0] = v[r];
    a[r][1] = v[r + 4];
    a[r][2] = v[r + 8];
    a[r][3] = v[r + 12];
    twiddle4(a[r], w2);
    dft_small<T, 4>(a[r]);
  }
#pragma unroll
  for (int r2 = 0; r2 < 4; r2++) {  // stage 3 (ns 64): butterfly 64 (t/16) + t mod 16 + 16 r2
    cplx<T> u[4] = {a[0][r2], a[1][r2], a[2][r2], a[3][r2]};
    twiddle4(u, w3[r2]);
    dft_small<T, 4>(u);
#pragma unroll
    for (int r4 = 0; r4 < 4; r4++) o[r2 + 4 * r4] = u[r4];
  }
}

// One complex FFT of length N2 (forward), lane t of its TPF lanes, values
// v[q] at the places t + TPF q in and out.  zf: the FFT's LDS region.
template <typename T, int N2>
__device__ __forceinline__ void fft(cplx<T>* v, cplx<T>* __restrict__ zf, const Tw<T, N2>& tw, int t) {
  constexpr int TPF = N2 / 16;
  static_assert(N2 == 256 || N2 == 512 || N2 == 1024, "register-blocked FFT sizes");
  cplx<T> o[16];
  pair_a<T>(v, o, tw.w1);
  wave_sync();  // the previous use of zf (an earlier transform) is done
#pragma unroll
  for (int q = 0; q < 16; q++) zf[lx(16 * t + q)] = o[q];
  wave_sync();
#pragma unroll
  for (int q = 0; q < 16; q++) v[q] = zf[lx(t + TPF * q)];
  pair_b<T>(v, o, tw.w2, tw.w3);
  if constexpr (N2 == 256) {
#pragma unroll
    for (int q = 0; q < 16; q++) v[q] = o[q];  // 256 (t/16) + t mod 16 + 16 q = t + 16 q
  } else {
    const int base = 256 * (t >> 4) + (t & 15);
    wave_sync();
#pragma unroll
    for (int q = 0; q < 16; q++) zf[lx(base + 16 * q)] = o[q];
    wave_sync();
#pragma unroll
    for (int q = 0; q < 16; q++) v[q] = zf[lx(t + TPF * q)];
    if constexpr (N2 == 1024) {  // stage 4 (ns 256, radix 4): butterfly t + 64 u
#pragma unroll
      for (int u = 0; u < 4; u++) {
        cplx<T> x[4] = {v[u], v[u + 4], v[u + 8], v[u + 12]};
        twiddle4(x, tw.wl[u]);
        dft_small<T, 4>(x);
#pragma unroll
        for (int r = 0; r < 4; r++) o[u + 4 * r] = x[r];
      }
    } else {  // stage 4 (ns 256, radix 2): butterfly t + 32 u
#pragma unroll
      for (int u = 0; u < 8; u++) {
        cplx<T> x[2] = {v[u], cmul(v[u + 8], tw.wl[u])};
        dft_small<T, 2>(x);
        o[u] = x[0];
        o[u + 8] = x[1];
      }
    }
#pragma unroll
    for (int q = 0; q < 16; q++) v[q] = o[q];
  }
}

// v[8..15] (the places N2/2 .. N2-1 of lane t) from the first half of the
// odd extension: z_{N2-p} = -z_p.  Place t + TPF q (q >= 8) mirrors place
// (TPF - t) + TPF (15 - q) of lane TPF - t; lane 0 mirrors its own places
// TPF (16 - q), and its place N2/2 (q = 8) is 0.  fl: the lane index of
// this FFT's lane 0 in the wave.
template <typename T, int TPF>
__device__ __forceinline__ void mirror_upper(cplx<T>* v, int t, int fl) {
  const int src = fl + ((TPF - t) & (TPF - 1));
#pragma unroll
  for (int q = 8; q < 16; q++) {
    const T re = __shfl(v[15 - q].re, src, 64);
    const T im = __shfl(v[15 - q].im, src, 64);
    const cplx<T> own = q == 8 ? cplx<T>{T(0), T(0)} : v[16 - q];
    v[q] = t == 0 ? cplx<T>{-own.re, -own.im} : cplx<T>{-re, -im};
  }
}

// Loads of FFT p at lane t: the places t + TPF q, q < 8, of the odd
// extension of local sequences 2p (real) and 2p+1 (imaginary): z_0 = 0,
// z_j = x_{j-1}.  Then the mirrored places.
template <typename T, int N2, class Op>
__device__ __forceinline__ void load(const Op& op, cplx<T>* v, int b, int g, int p, int P, int t, int fl) {
  constexpr int TPF = N2 / 16;
  T xr[8], xi[8];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int idx = t + TPF * q - 1;
    const int ic = idx < 0 ? 0 : idx;  // every load at an existing place
    xr[q] = op.load(b, g, 2 * p, P, ic);
    xi[q] = op.load(b, g, 2 * p + 1, P, ic);
  }
#pragma unroll
  for (int q = 0; q < 8; q++) v[q] = (t + TPF * q == 0) ? cplx<T>{T(0), T(0)} : cplx<T>{xr[q], xi[q]};
  mirror_upper<T, TPF>(v, t, fl);
}

// grid.x = groups of sequences (XCD-contiguous order), grid.y = batch;
// W waves per workgroup.
template <typename T, int N2, int W, int AXIS, class Op>
__global__ __launch_bounds__(64 * W) void dst_pass_w(Op op, const cplx<T>* __restrict__ tw) {
  constexpr int TPF = N2 / 16, P = W * 64 / TPF;
  constexpr bool kPre = op_store_pre<Op>::value;
  __shared__ cplx<T> zs[P * kLdsPerFft<N2>];
  const int lane = threadIdx.x & 63, p = threadIdx.x / TPF, t = threadIdx.x % TPF;
  const int fl = lane - t;
  const int b = blockIdx.y;
  int g = blockIdx.x;
  if ((gridDim.x & 7) == 0) g = (g & 7) * (gridDim.x >> 3) + (g >> 3);
  T oa[kPre ? 8 : 1], ob[kPre ? 8 : 1];
  if constexpr (kPre) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int k = t + TPF * q;
      oa[q] = ob[q] = T(0);
      if (k >= 1) op.store_pre(b, g, p, P, k, oa[q], ob[q]);
    }
  }
  Tw<T, N2> w;
  w.load(tw, t);
  cplx<T> v[16];
  load<T, N2, Op>(op, v, b, g, p, P, t, fl);
  fft<T, N2>(v, zs + p * kLdsPerFft<N2>, w, t);
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int k = t + TPF * q;
    if (k >= 1) {
      if constexpr (kPre) op.store_put(b, g, p, P, k, -v[q].im, v[q].re, oa[q], ob[q]);
      else op.store(b, g, p, P, k, -v[q].im, v[q].re);
    }
  }
}

// fused column pair (dst_col2 of dst_core.hpp) with register-blocked FFTs:
// the first transform's outputs at the lane's places t + TPF q are the
// second transform's inputs at the same places, so the middle step needs no
// exchange beyond the mirrored half
template <typename T, int N2, int W, class Op>
__global__ __launch_bounds__(64 * W) void dst_col2_w(Op op, const cplx<T>* __restrict__ tw) {
  constexpr int TPF = N2 / 16, P = W * 64 / TPF;
  constexpr bool kCoef = op_has_mid_coef<Op>::value;
  constexpr bool kPre = op_store_pre<Op>::value;
  __shared__ cplx<T> zs[P * kLdsPerFft<N2>];
  const int lane = threadIdx.x & 63, p = threadIdx.x / TPF, t = threadIdx.x % TPF;
  const int fl = lane - t;
  const int b = blockIdx.y;
  int g = blockIdx.x;
  if ((gridDim.x & 7) == 0) g = (g & 7) * (gridDim.x >> 3) + (g >> 3);
  T ca[kCoef ? 8 : 1], cb[kCoef ? 8 : 1];
  if constexpr (kCoef) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int k = t + TPF * q;
      ca[q] = cb[q] = T(0);
      if (k >= 1) op.mid_coef(b, g, p, P, k, ca[q], cb[q]);
    }
  }
  T oa[kPre ? 8 : 1], ob[kPre ? 8 : 1];
  if constexpr (kPre) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int k = t + TPF * q;
      oa[q] = ob[q] = T(0);
      if (k >= 1) op.store_pre(b, g, p, P, k, oa[q], ob[q]);
    }
  }
  Tw<T, N2> w;
  w.load(tw, t);
  cplx<T> v[16];
  cplx<T>* zf = zs + p * kLdsPerFft<N2>;
  load<T, N2, Op>(op, v, b, g, p, P, t, fl);
  fft<T, N2>(v, zf, w, t);
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int k = t + TPF * q;
    T va = T(0), vb = T(0);
    if (k >= 1) {
      if constexpr (kCoef) op.mid_apply(b, g, p, P, k, -v[q].im, v[q].re, ca[q], cb[q], va, vb);
      else op.mid(b, g, p, P, k, -v[q].im, v[q].re, va, vb);
    }
    v[q] = {va, vb};
  }
  mirror_upper<T, TPF>(v, t, fl);
  fft<T, N2>(v, zf, w, t);
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int k = t + TPF * q;
    if (k >= 1) {
      if constexpr (kPre) op.store_put(b, g, p, P, k, -v[q].im, v[q].re, oa[q], ob[q]);
      else op.store(b, g, p, P, k, -v[q].im, v[q].re);
    }
  }
}

// ---------------------------------------------------------------------------
// Column passes with staged loads and stores.  The register-blocked FFT reads
// its 16 places per lane at 16 different rows when the sequences are
// columns, so one wave's loads touch one cache line per lane for 16 bytes of
// it.  Here the four waves of a workgroup (P = 4 64 / TPF FFTs, 2P adjacent
// columns) first load the 2P columns together, P pairs of adjacent values
// per row (64 bytes for 2m = 1024, a whole 128-byte line for 512), into LDS,
// each wave then takes its places from there; the outputs go back through
// LDS the same way.  The staging area is the FFTs' exchange area (it fits:
// 2P (N-1) values against P (N2 + N2/16) complex ones), so it costs three
// workgroup barriers and no extra LDS.  Any loader (several loads per
// element included) and any store; not the prefetched read-modify-write
// stores (store_pre), which run on the other passes.
template <class Op>
constexpr bool kStagedOk = !op_side<Op>::value && !op_gather<Op>::value && !op_load_post<Op>::value &&
                           !op_store_pre<Op>::value;
constexpr int kStagedWaves = 4;

// the 2P sequences of the workgroup, rows 0..LEN-1, into stage[s LEN + r]
template <typename T, int N2, class Op>
__device__ __forceinline__ void staged_load(const Op& op, T* stage, int b, int g) {
  constexpr int TPF = N2 / 16, P = kStagedWaves * 64 / TPF, LEN = N2 / 2 - 1, NT = kStagedWaves * 64;
  constexpr int NE = P * LEN, NL = (NE + NT - 1) / NT;
  T xa[NL], xb[NL];
#pragma unroll
  for (int i = 0; i < NL; i++) {  // element e: pair e mod P of row e / P (consecutive threads: one row)
    const int e = threadIdx.x + NT * i;
    const int ec = e < NE ? e : NE - 1;
    const int pr = ec % P, r = ec / P;
    xa[i] = op.load(b, g, 2 * pr, P, r);
    xb[i] = op.load(b, g, 2 * pr + 1, P, r);
  }
#pragma unroll
  for (int i = 0; i < NL; i++) {
    const int e = threadIdx.x + NT * i;
    if (e < NE) {
      const int pr = e % P, r = e / P;
      stage[(2 * pr) * LEN + r] = xa[i];
      stage[(2 * pr + 1) * LEN + r] = xb[i];
    }
  }
}
// places t + TPF q (q < 8) of FFT p from the staging area, then the mirror
template <typename T, int N2>
__device__ __forceinline__ void staged_take(const T* stage, cplx<T>* v, int p, int t, int fl) {
  constexpr int TPF = N2 / 16, LEN = N2 / 2 - 1;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int idx = t + TPF * q - 1;
    const int ic = idx < 0 ? 0 : idx;
    const T a = stage[(2 * p) * LEN + ic], c = stage[(2 * p + 1) * LEN + ic];
    v[q] = idx < 0 ? cplx<T>{T(0), T(0)} : cplx<T>{a, c};
  }
  mirror_upper<T, TPF>(v, t, fl);
}
// outputs Y(k) = (-Im Z_k, Re Z_k) of FFT p at the lane's places, into the
// staging area (row k-1 of sequences 2p, 2p+1)
template <typename T, int N2>
__device__ __forceinline__ void staged_put(T* stage, const cplx<T>* v, int p, int t) {
  constexpr int TPF = N2 / 16, LEN = N2 / 2 - 1;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int k = t + TPF * q;
    if (k >= 1) {
      stage[(2 * p) * LEN + k - 1] = -v[q].im;
      stage[(2 * p + 1) * LEN + k - 1] = v[q].re;
    }
  }
}
template <typename T, int N2, class Op>
__device__ __forceinline__ void staged_store(const Op& op, const T* stage, int b, int g) {
  constexpr int TPF = N2 / 16, P = kStagedWaves * 64 / TPF, LEN = N2 / 2 - 1, NT = kStagedWaves * 64;
  constexpr int NE = P * LEN, NL = (NE + NT - 1) / NT;
#pragma unroll
  for (int i = 0; i < NL; i++) {
    const int e = threadIdx.x + NT * i;
    if (e < NE) {
      const int pr = e % P, r = e / P;
      op.store(b, g, pr, P, r + 1, stage[(2 * pr) * LEN + r], stage[(2 * pr + 1) * LEN + r]);
    }
  }
}

// one DST-I along the columns (AXIS 1 pass)
template <typename T, int N2, class Op>
__global__ __launch_bounds__(kStagedWaves * 64) void dst_col_ws(Op op, const cplx<T>* __restrict__ tw) {
  constexpr int TPF = N2 / 16, P = kStagedWaves * 64 / TPF;
  static_assert(2 * P * (N2 / 2 - 1) * sizeof(T) <= P * kLdsPerFft<N2> * sizeof(cplx<T>), "staging area");
  __shared__ cplx<T> zs[P * kLdsPerFft<N2>];
  T* const stage = reinterpret_cast<T*>(zs);
  const int lane = threadIdx.x & 63, p = threadIdx.x / TPF, t = threadIdx.x % TPF;
  const int fl = lane - t;
  const int b = blockIdx.y;
  int g = blockIdx.x;
  if ((gridDim.x & 7) == 0) g = (g & 7) * (gridDim.x >> 3) + (g >> 3);
  Tw<T, N2> w;
  w.load(tw, t);
  staged_load<T, N2, Op>(op, stage, b, g);
  __syncthreads();
  cplx<T> v[16];
  staged_take<T, N2>(stage, v, p, t, fl);
  __syncthreads();  // the exchange areas overlap other FFTs' staged columns
  fft<T, N2>(v, zs + p * kLdsPerFft<N2>, w, t);
  __syncthreads();
  staged_put<T, N2>(stage, v, p, t);
  __syncthreads();
  staged_store<T, N2, Op>(op, stage, b, g);
}

// the fused column pair (dst_col2 of dst_core.hpp) with staged columns
template <typename T, int N2, class Op>
__global__ __launch_bounds__(kStagedWaves * 64) void dst_col2_ws(Op op, const cplx<T>* __restrict__ tw) {
  constexpr int TPF = N2 / 16, P = kStagedWaves * 64 / TPF;
  static_assert(2 * P * (N2 / 2 - 1) * sizeof(T) <= P * kLdsPerFft<N2> * sizeof(cplx<T>), "staging area");
  constexpr bool kCoef = op_has_mid_coef<Op>::value;
  __shared__ cplx<T> zs[P * kLdsPerFft<N2>];
  T* const stage = reinterpret_cast<T*>(zs);
  const int lane = threadIdx.x & 63, p = threadIdx.x / TPF, t = threadIdx.x % TPF;
  const int fl = lane - t;
  const int b = blockIdx.y;
  int g = blockIdx.x;
  if ((gridDim.x & 7) == 0) g = (g & 7) * (gridDim.x >> 3) + (g >> 3);
  T ca[kCoef ? 8 : 1], cb[kCoef ? 8 : 1];
  if constexpr (kCoef) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int k = t + TPF * q;
      ca[q] = cb[q] = T(0);
      if (k >= 1) op.mid_coef(b, g, p, P, k, ca[q], cb[q]);
    }
  }
  Tw<T, N2> w;
  w.load(tw, t);
  staged_load<T, N2, Op>(op, stage, b, g);
  __syncthreads();
  cplx<T> v[16];
  staged_take<T, N2>(stage, v, p, t, fl);
  __syncthreads();
  cplx<T>* zf = zs + p * kLdsPerFft<N2>;
  fft<T, N2>(v, zf, w, t);
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int k = t + TPF * q;
    T va = T(0), vb = T(0);
    if (k >= 1) {
      if constexpr (kCoef) op.mid_apply(b, g, p, P, k, -v[q].im, v[q].re, ca[q], cb[q], va, vb);
      else op.mid(b, g, p, P, k, -v[q].im, v[q].re, va, vb);
    }
    v[q] = {va, vb};
  }
  mirror_upper<T, TPF>(v, t, fl);
  fft<T, N2>(v, zf, w, t);
  __syncthreads();
  staged_put<T, N2>(stage, v, p, t);
  __syncthreads();
  staged_store<T, N2, Op>(op, stage, b, g);
}

}  // namespace wave
}  // namespace ibfs
