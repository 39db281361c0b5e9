// dst_core.hpp -- batched 1-D DST-I passes for gfx950.
//
// A 2-D DST-I of an (m-1) x (n-1) vertex field (reference src/myfft.f90:83-94,
// FFTW RODFT00 on both axes, unnormalised) is two 1-D passes: one along the
// contiguous axis i ("row pass") and one along j ("column pass").  Both passes
// read and write the natural i-fastest layout, so the field never needs a
// transposed copy.
//
// One 1-D DST-I of length N-1 is
//     Y_k = 2 sum_{j=1}^{N-1} x_j sin(pi j k / N),   k = 1..N-1.
// Two real sequences a, b are packed as z = a + i b and extended oddly to
// length 2N (z_0 = z_N = 0, z_{2N-j} = -z_j).  Then the forward DFT gives
// Z_k = -i Y(a)_k + Y(b)_k, so Y(b)_k = Re Z_k and Y(a)_k = -Im Z_k with no
// post-processing.  The complex DFT is the Stockham autosort formulation of
// Govindaraju et al., SC08, done in LDS with radix 4/2/3/5 stages.
//
// A workgroup (256 threads = 4 wave64) holds P = EMAX / (2N) such FFTs in LDS
// (EMAX complex values).  Every stage first pulls all butterfly inputs of a
// thread into registers, then (after a barrier) writes the outputs back in
// place, so one LDS buffer suffices.  The radix schedule and the twiddle
// table (exp(-2 pi i k / 2N), k < 2N) are runtime data, so one kernel
// instantiation serves every supported size; the per-radix stage code is a
// template so that register arrays are statically indexed.
//
// Fusion: the pass takes an operator object with
//     T    load(int b, int g, int s, int P, int idx)
//              value of local sequence s (0..2P-1) of group g, batch b, at
//              position idx (0..N-2); 0 for sequences that do not exist
//     void store(int b, int g, int p, int P, int k, T ya, T yb)
//              outputs of FFT p (0..P-1) at k (1..N-1): ya is the transform
//              of local sequence 2p, yb that of local sequence 2p+1
// Sequences 2p and 2p+1 share FFT p whatever the number P of FFTs per
// workgroup, so the results do not depend on P (the tile shape).
// so that pointwise prologue/epilogue work (the AB2 combine, boundary
// terms, eigenvalue scaling, the CN combine) runs inside the pass.  The row
// pass (AXIS 0) walks loads/stores along idx/k, which is contiguous in
// memory; the column pass (AXIS 1) walks them along s/p, i.e. across
// adjacent columns of one row.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

namespace ibfs {

// Optional operator hooks used by the size-specialised passes to issue
// every global load of a workgroup before its first FFT stage:
//   mid_coef(b, g, p, P, k, ca, cb) / mid_apply(..., ya, yb, ca, cb, va, vb):
//       mid split into the table reads and the arithmetic;
//   kStorePre, store_pre(b, g, p, P, k, oa, ob) / store_put(..., ya, yb, oa, ob):
//       a read-modify-write store split into the read and the write.
template <class Op, class = void>
struct op_has_mid_coef : std::false_type {};
template <class Op>
struct op_has_mid_coef<Op, std::void_t<decltype(&Op::mid_coef)>> : std::true_type {};
template <class Op, class = void>
struct op_store_pre : std::false_type {};
template <class Op>
struct op_store_pre<Op, std::void_t<decltype(Op::kStorePre)>> : std::bool_constant<Op::kStorePre> {};
//   kLoadPost, load_aux(b, g, s, P, idx, aux) / load_post(..., aux):
//       a loader with a side store split into the loads and the store.
template <class Op, class = void>
struct op_load_post : std::false_type {};
template <class Op>
struct op_load_post<Op, std::void_t<decltype(Op::kLoadPost)>> : std::bool_constant<Op::kLoadPost> {};
//   kGather, Gathered, gather(b, g, s, P, idx, G) / combine(..., G, aux):
//       a side-store loader in two phases, the loads of every element of a
//       thread issued before any value is formed.
//   kSide, side_groups_for(nt), side(b, x, t): independent work that the
//       pass's first side_groups_for(blockDim) workgroups do instead of a
//       transform (a multiple of 8 of them, so that the transform groups
//       keep their XCD order); 0: none.
template <class Op, class = void>
struct op_side : std::false_type {};
template <class Op>
struct op_side<Op, std::void_t<decltype(Op::kSide)>> : std::bool_constant<Op::kSide> {};
template <class Op, class = void>
struct op_gather : std::false_type {};
template <class Op>
struct op_gather<Op, std::void_t<decltype(Op::kGather)>> : std::bool_constant<Op::kGather> {};

constexpr int kThreads = 256;
constexpr int kMaxStages = 16;

// Workgroup x of a row of n workgroups runs on XCD x % 8 (round-robin
// dispatch over the 8 XCDs).  Returns x's place in an order that gives XCD k
// the contiguous range [k q + min(k, r), ...) of q + (k < r) workgroups,
// q = n / 8, r = n % 8: a bijection of [0, n) for every n, so that
// neighbouring tiles, which share cache lines, share one XCD's L2.
__device__ __forceinline__ int xcd_contiguous(int x, int n) {
  const int q = n >> 3, r = n & 7, k = x & 7;
  return k * q + min(k, r) + (x >> 3);
}

template <typename T>
struct cplx {
  T re, im;
};

template <typename T>
struct FftPlan {
  int n2;                  // complex FFT length 2N
  int nst;                 // number of stages
  int rad[kMaxStages];     // radix of each stage
  const cplx<T>* tw;       // n2 twiddles exp(-2 pi i k / n2)
  int nfft;                // FFTs per workgroup, set per launch (nfft * n2 <= EMAX)
};

// complex product with explicit fused multiply-adds, so that every kernel
// that forms it (whatever the compiler would contract around it) rounds it
// the same way: the register-blocked and LDS passes stay bit-identical
template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {fma(a.re, b.re, -(a.im * b.im)), fma(a.re, b.im, a.im * b.re)};
}

// In-register DFT of radix R (forward, exp(-2 pi i q r / R)).  The radix-2
// and radix-4 butterflies only add and subtract, and are kept so: no product
// computed before them (a loader's value, a spectral factor, a twiddle) may
// be fused into their additions, which would round differently in kernels
// whose inputs come from registers than in those whose inputs pass through
// LDS.
template <typename T, int R>
__device__ __forceinline__ void dft_small(cplx<T>* v) {
  if constexpr (R == 2) {
#pragma clang fp contract(off)
    cplx<T> a = v[0], b = v[1];
    v[0] = {a.re + b.re, a.im + b.im};
    v[1] = {a.re - b.re, a.im - b.im};
  } else if constexpr (R == 4) {
#pragma clang fp contract(off)
    cplx<T> t0 = {v[0].re + v[2].re, v[0].im + v[2].im};
    cplx<T> t1 = {v[0].re - v[2].re, v[0].im - v[2].im};
    cplx<T> t2 = {v[1].re + v[3].re, v[1].im + v[3].im};
    cplx<T> t3 = {v[1].re - v[3].re, v[1].im - v[3].im};
    v[0] = {t0.re + t2.re, t0.im + t2.im};
    v[2] = {t0.re - t2.re, t0.im - t2.im};
    v[1] = {t1.re + t3.im, t1.im - t3.re};  // t1 - i t3
    v[3] = {t1.re - t3.im, t1.im + t3.re};  // t1 + i t3
  } else if constexpr (R == 3) {
    // c = cos(2 pi/3) = -1/2, s = sin(2 pi/3) = sqrt(3)/2
    const T s3 = T(0.86602540378443864676372317075293618);
    cplx<T> a = v[0], b = v[1], c = v[2];
    cplx<T> sum = {b.re + c.re, b.im + c.im};
    cplx<T> dif = {b.re - c.re, b.im - c.im};
    cplx<T> mid = {a.re - T(0.5) * sum.re, a.im - T(0.5) * sum.im};
    v[0] = {a.re + sum.re, a.im + sum.im};
    // v1 = mid - i s dif, v2 = mid + i s dif
    v[1] = {mid.re + s3 * dif.im, mid.im - s3 * dif.re};
    v[2] = {mid.re - s3 * dif.im, mid.im + s3 * dif.re};
  } else if constexpr (R == 5) {
    const T c1 = T(0.30901699437494742410229341718281906);   // cos(2 pi/5)
    const T c2 = T(-0.80901699437494742410229341718281906);  // cos(4 pi/5)
    const T s1 = T(0.95105651629515357211643933337938214);   // sin(2 pi/5)
    const T s2 = T(0.58778525229247312916870595463907277);   // sin(4 pi/5)
    cplx<T> a = v[0];
    cplx<T> p1 = {v[1].re + v[4].re, v[1].im + v[4].im};
    cplx<T> m1 = {v[1].re - v[4].re, v[1].im - v[4].im};
    cplx<T> p2 = {v[2].re + v[3].re, v[2].im + v[3].im};
    cplx<T> m2 = {v[2].re - v[3].re, v[2].im - v[3].im};
    v[0] = {a.re + p1.re + p2.re, a.im + p1.im + p2.im};
    cplx<T> r1 = {a.re + c1 * p1.re + c2 * p2.re, a.im + c1 * p1.im + c2 * p2.im};
    cplx<T> r2 = {a.re + c2 * p1.re + c1 * p2.re, a.im + c2 * p1.im + c1 * p2.im};
    // i-parts: q1 = s1 m1 + s2 m2, q2 = s2 m1 - s1 m2 ; v1 = r1 - i q1, v4 = r1 + i q1
    cplx<T> q1 = {s1 * m1.re + s2 * m2.re, s1 * m1.im + s2 * m2.im};
    cplx<T> q2 = {s2 * m1.re - s1 * m2.re, s2 * m1.im - s1 * m2.im};
    v[1] = {r1.re + q1.im, r1.im - q1.re};
    v[4] = {r1.re - q1.im, r1.im + q1.re};
    v[2] = {r2.re + q2.im, r2.im - q2.re};
    v[3] = {r2.re - q2.im, r2.im + q2.re};
  }
}

// One Stockham stage of radix R over nfft FFTs of length n2 held in z.
template <typename T, int R, int EMAX>
__device__ __forceinline__ void fft_stage(cplx<T>* z, int nfft, int n2, int ns, const cplx<T>* __restrict__ tw) {
  constexpr int NB = (EMAX / R + kThreads - 1) / kThreads;
  const int nr = n2 / R;
  const int nbf = nfft * nr;
  const int twstep = n2 / (ns * R);
  cplx<T> v[NB][R];
  int dst[NB];
#pragma unroll
  for (int t = 0; t < NB; t++) {
    const int b = threadIdx.x + t * kThreads;
    dst[t] = -1;
    if (b < nbf) {
      const int p = b / nr;
      const int j = b - p * nr;
      const int k = j % ns;
      const cplx<T>* src = z + p * n2 + j;
#pragma unroll
      for (int r = 0; r < R; r++) v[t][r] = src[r * nr];
#pragma unroll
      for (int r = 1; r < R; r++) v[t][r] = cmul(v[t][r], tw[k * r * twstep]);
      dst[t] = p * n2 + (j - k) * R + k;
    }
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < NB; t++) {
    if (dst[t] >= 0) {
      dft_small<T, R>(v[t]);
#pragma unroll
      for (int r = 0; r < R; r++) z[dst[t] + r * ns] = v[t][r];
    }
  }
  __syncthreads();
}

template <typename T, int EMAX>
__device__ __forceinline__ void fft_run(cplx<T>* z, int nfft, const FftPlan<T>& plan) {
  int ns = 1;
  for (int s = 0; s < plan.nst; s++) {
    const int R = plan.rad[s];
    switch (R) {
      case 4: fft_stage<T, 4, EMAX>(z, nfft, plan.n2, ns, plan.tw); break;
      case 2: fft_stage<T, 2, EMAX>(z, nfft, plan.n2, ns, plan.tw); break;
      case 3: fft_stage<T, 3, EMAX>(z, nfft, plan.n2, ns, plan.tw); break;
      default: fft_stage<T, 5, EMAX>(z, nfft, plan.n2, ns, plan.tw); break;
    }
    ns *= R;
  }
}

// AXIS 0: sequences are contiguous rows; AXIS 1: sequences are columns.
// grid.x = groups of sequences, grid.y = batch.
template <typename T, int EMAX, int AXIS, class Op>
__global__ __launch_bounds__(kThreads) void dst_pass(Op op, FftPlan<T> plan) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* z = reinterpret_cast<cplx<T>*>(smem_raw);
  const int n2 = plan.n2;
  const int N = n2 / 2;
  const int len = N - 1;
  const int P = plan.nfft;          // FFTs per workgroup
  const int b = blockIdx.y;
  int g = blockIdx.x;
  if constexpr (op_side<Op>::value) {
    const int sg = op.side_groups_for(blockDim.x);
    if (g < sg) {
      op.side(b, g, threadIdx.x);
      return;
    }
    g -= sg;
  }

  // zero the two fixed points of every odd extension
  for (int p = threadIdx.x; p < P; p += kThreads) {
    z[p * n2] = {T(0), T(0)};
    z[p * n2 + N] = {T(0), T(0)};
  }
  // load: local sequence s -> FFT s/2, real part if s is even else imaginary
  const int tot = 2 * P * len;
  for (int e = threadIdx.x; e < tot; e += kThreads) {
    int s, idx;
    if (AXIS == 0) { s = e / len; idx = e - s * len; }
    else { idx = e / (2 * P); s = e - idx * (2 * P); }
    const T x = op.load(b, g, s, P, idx);
    const int p = s >> 1;
    const int h = s & 1;
    reinterpret_cast<T*>(z + p * n2 + idx + 1)[h] = x;
    reinterpret_cast<T*>(z + p * n2 + n2 - idx - 1)[h] = -x;
  }
  __syncthreads();
  fft_run<T, EMAX>(z, P, plan);
  // store: Y(a)_k = -Im Z_k, Y(b)_k = Re Z_k
  const int tot2 = P * len;
  for (int e = threadIdx.x; e < tot2; e += kThreads) {
    int p, k;
    if (AXIS == 0) { p = e / len; k = e - p * len + 1; }
    else { k = e / P + 1; p = e - (k - 1) * P; }
    const cplx<T> Z = z[p * n2 + k];
    op.store(b, g, p, P, k, -Z.im, Z.re);
  }
}

// Fused column pair: DST-I along j, a pointwise spectral operation, and a
// second DST-I along j, with one HBM round trip.  For a separable 2-D
// transform S = Sx Sy, the sequence Sx . Sy . diag(tab) . Sy . Sx keeps the
// middle three factors inside one column, so S diag(tab) S costs a row pass,
// this pass and a row pass.  The operator adds
//     void mid(int b, int g, int p, int P, int k, T ya, T yb, T& va, T& vb)
// which maps the first transform's outputs of FFT p at k (as store would
// receive them) to the inputs of the second transform at the same place.
template <typename T, class Op>
__device__ __forceinline__ void col2_mid(const Op& op, cplx<T>* z, int b, int g, int P, int n2, int t, int nt) {
  const int N = n2 / 2, len = N - 1;
  for (int e = t; e < P * len; e += nt) {
    const int k = e / P + 1, p = e - (k - 1) * P;
    const cplx<T> Z = z[p * n2 + k];
    T va, vb;
    op.mid(b, g, p, P, k, -Z.im, Z.re, va, vb);
    z[p * n2 + k] = {va, vb};
    z[p * n2 + n2 - k] = {-va, -vb};
  }
  for (int p = t; p < P; p += nt) {
    z[p * n2] = {T(0), T(0)};
    z[p * n2 + N] = {T(0), T(0)};
  }
  __syncthreads();
}

template <typename T, int EMAX, class Op>
__global__ __launch_bounds__(kThreads) void dst_col2(Op op, FftPlan<T> plan) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* z = reinterpret_cast<cplx<T>*>(smem_raw);
  const int n2 = plan.n2;
  const int N = n2 / 2;
  const int len = N - 1;
  const int P = plan.nfft;
  const int b = blockIdx.y;
  const int g = blockIdx.x;
  for (int p = threadIdx.x; p < P; p += kThreads) {
    z[p * n2] = {T(0), T(0)};
    z[p * n2 + N] = {T(0), T(0)};
  }
  const int tot = 2 * P * len;
  for (int e = threadIdx.x; e < tot; e += kThreads) {
    const int idx = e / (2 * P), s = e - idx * (2 * P);
    const T x = op.load(b, g, s, P, idx);
    const int p = s >> 1;
    const int h = s & 1;
    reinterpret_cast<T*>(z + p * n2 + idx + 1)[h] = x;
    reinterpret_cast<T*>(z + p * n2 + n2 - idx - 1)[h] = -x;
  }
  __syncthreads();
  fft_run<T, EMAX>(z, P, plan);
  col2_mid<T, Op>(op, z, b, g, P, n2, threadIdx.x, kThreads);
  fft_run<T, EMAX>(z, P, plan);
  for (int e = threadIdx.x; e < P * len; e += kThreads) {
    const int k = e / P + 1, p = e - (k - 1) * P;
    const cplx<T> Z = z[p * n2 + k];
    op.store(b, g, p, P, k, -Z.im, Z.re);
  }
}

// ---------------------------------------------------------------------------
// Size-specialised pass (the sizes of the production cases).  N2 = 2N is a
// compile-time constant, so the radix schedule, every index division and the
// twiddle-table offsets fold into shifts and constants.  One FFT is served by
// TPF = N2/4 threads (one radix-4 butterfly per thread per stage); a
// workgroup holds P FFTs.  Twiddles come from a per-stage table of
// w = exp(-2 pi i k / (ns R)), one load per butterfly and stage, the other
// powers formed in registers (TwLayout, tw_load).  Tiles of at most 32 KB
// of LDS ping-pong between two buffers (one barrier per stage).  The
// workgroup order is made XCD-aware: consecutive workgroups are
// dispatched round-robin over the 8 XCDs, so group g is remapped to give
// every XCD one contiguous range of sequences.  Column segments shared by
// neighbouring column groups, and the neighbouring rows that the stencil
// loaders of row passes read, then stay in that XCD's L2.
// ---------------------------------------------------------------------------
namespace spec {

struct Sched {
  int nst;
  int rad[kMaxStages];
  int ns[kMaxStages];   // product of the radices before stage s
  int toff[kMaxStages]; // offset of stage s in the twiddle table
  int ntw;              // table length
};

constexpr Sched make_sched(int n2) {
  Sched s{};
  int r = n2, k = 0;
  while (r % 4 == 0) { s.rad[k++] = 4; r /= 4; }
  while (r % 2 == 0) { s.rad[k++] = 2; r /= 2; }
  while (r % 3 == 0) { s.rad[k++] = 3; r /= 3; }
  while (r % 5 == 0) { s.rad[k++] = 5; r /= 5; }
  s.nst = k;
  int ns = 1, off = 0;
  for (int i = 0; i < k; i++) {
    s.ns[i] = ns;
    s.toff[i] = off;
    off += ns;  // one factor exp(-2 pi i k / (ns R)) per butterfly index k
    ns *= s.rad[i];
  }
  s.ntw = off;
  return s;
}

// Twiddles of one thread, all stages, held in registers: stage s with
// ns > 1 owns NBT(s) * (R(s)-1) factors starting at off(s).  The table holds
// only w = exp(-2 pi i k / (ns R)) for each k; a thread loads one factor per
// butterfly and stage at kernel start and forms w^2 .. w^(R-1) by complex
// products, so that no stage waits on a memory load and the loads are a
// third of those of a table of every power.
template <int N2>
struct TwLayout {
  static constexpr Sched sc = make_sched(N2);
  static constexpr int TPF = N2 / 4;
  static constexpr int nbt(int s) { return (N2 / sc.rad[s] + TPF - 1) / TPF; }
  static constexpr int off(int s) {
    int o = 0;
    for (int q = 0; q < s; q++)
      if (sc.ns[q] > 1) o += nbt(q) * (sc.rad[q] - 1);
    return o;
  }
  static constexpr int total() { return off(sc.nst) > 0 ? off(sc.nst) : 1; }
};

template <typename T, int N2, int S>
__device__ __forceinline__ void tw_load(cplx<T>* wr, const cplx<T>* __restrict__ tw, int lt) {
  constexpr Sched sc = make_sched(N2);
  if constexpr (S < sc.nst) {
    constexpr int R = sc.rad[S], NS = sc.ns[S], NR = N2 / R, TPF = N2 / 4;
    constexpr int NBT = (NR + TPF - 1) / TPF;
    if constexpr (NS > 1) {
      constexpr int O = TwLayout<N2>::off(S);
#pragma unroll
      for (int i = 0; i < NBT; i++) {
        const int j = lt + i * TPF;
        const int k = (j < NR ? j : 0) % NS;
        const cplx<T> w = tw[sc.toff[S] + k];
        cplx<T> pw = w;
        wr[O + i * (R - 1)] = w;
#pragma unroll
        for (int r = 2; r < R; r++) {
          pw = cmul(pw, w);
          wr[O + i * (R - 1) + r - 1] = pw;
        }
      }
    }
    tw_load<T, N2, S + 1>(wr, tw, lt);
  }
}

// One Stockham stage from zi to zo.  In place (zi == zo, DB false) the
// stage reads, waits at a barrier, then writes; with two buffers (DB) the
// writes go to the other buffer at once, and the stage ends at its only
// barrier.
template <typename T, int N2, int P, int S, bool DB>
__device__ __forceinline__ void stage(const cplx<T>* zi, cplx<T>* zo, const cplx<T>* wr, int t) {
  constexpr Sched sc = make_sched(N2);
  constexpr int R = sc.rad[S];
  constexpr int NS = sc.ns[S];
  constexpr int NR = N2 / R;       // butterflies per FFT
  constexpr int TPF = N2 / 4;
  constexpr int NBT = (NR + TPF - 1) / TPF;
  constexpr int O = TwLayout<N2>::off(S);
  const int p = t / TPF, lt = t % TPF;
  const cplx<T>* zp = zi + p * N2;
  cplx<T>* zq = zo + p * N2;
  cplx<T> v[NBT][R];
#pragma unroll
  for (int i = 0; i < NBT; i++) {
    const int j = lt + i * TPF;
    if (NBT * TPF == NR || j < NR) {
#pragma unroll
      for (int r = 0; r < R; r++) v[i][r] = zp[j + r * NR];
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; r++) v[i][r] = cmul(v[i][r], wr[O + i * (R - 1) + r - 1]);
      }
    }
  }
  if constexpr (!DB) __syncthreads();
#pragma unroll
  for (int i = 0; i < NBT; i++) {
    const int j = lt + i * TPF;
    if (NBT * TPF == NR || j < NR) {
      dft_small<T, R>(v[i]);
      const int k = j % NS;
      const int d = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; r++) zq[d + r * NS] = v[i][r];
    }
  }
  __syncthreads();
}

// All stages from za; returns the buffer holding the result (za, or zb
// after an odd number of stages with two buffers).
template <typename T, int N2, int P, int S, bool DB>
__device__ __forceinline__ cplx<T>* run(cplx<T>* za, cplx<T>* zb, const cplx<T>* wr, int t) {
  constexpr Sched sc = make_sched(N2);
  if constexpr (S < sc.nst) {
    if constexpr (DB) {
      stage<T, N2, P, S, true>(za, zb, wr, t);
      return run<T, N2, P, S + 1, true>(zb, za, wr, t);
    } else {
      stage<T, N2, P, S, false>(za, za, wr, t);
      return run<T, N2, P, S + 1, false>(za, za, wr, t);
    }
  } else {
    return za;
  }
}

// two LDS buffers when they fit in 32 KB (the single-case tiles), one
// otherwise, so that the larger tiles keep their occupancy
template <typename T, int N2, int P>
constexpr bool kTwoBuf = 2L * P * N2 * (long)sizeof(cplx<T>) <= 32768;

// Loads of one pass: every thread first issues all of its NL loads (values
// kept in registers), then writes them into the odd extensions, so the loads
// share one memory round trip.
template <typename T, int N2, int P, int AXIS, class Op>
__device__ __forceinline__ void load_seqs(const Op& op, cplx<T>* z, int b, int g, int t) {
  constexpr int N = N2 / 2, LEN = N - 1, NT = P * N2 / 4;
  constexpr int TOT = 2 * P * LEN, NL = (TOT + NT - 1) / NT;
  constexpr bool kPost = op_load_post<Op>::value;
  T x[NL], aux[kPost ? NL : 1];
  if constexpr (op_gather<Op>::value) {
    // every element's loads first, then every element's value
    typename Op::Gathered gv[NL];
#pragma unroll
    for (int q = 0; q < NL; q++) {
      const int e = t + q * NT;
      const int ec = (NL * NT == TOT || e < TOT) ? e : TOT - 1;
      int s, idx;
      if (AXIS == 0) { s = ec / LEN; idx = ec - s * LEN; }
      else { idx = ec / (2 * P); s = ec - idx * (2 * P); }
      op.gather(b, g, s, P, idx, gv[q]);
    }
#pragma unroll
    for (int q = 0; q < NL; q++) {
      const int e = t + q * NT;
      const bool in = NL * NT == TOT || e < TOT;
      const int ec = in ? e : TOT - 1;
      int s, idx;
      if (AXIS == 0) { s = ec / LEN; idx = ec - s * LEN; }
      else { idx = ec / (2 * P); s = ec - idx * (2 * P); }
      T a = T(0);
      const T v = op.combine(b, g, s, P, idx, gv[q], a);
      if constexpr (kPost) aux[q] = a;
      x[q] = in ? v : T(0);
    }
  }
#pragma unroll
  for (int q = 0; q < NL && !op_gather<Op>::value; q++) {
    const int e = t + q * NT;
    int s = 0, idx = 0;
    if (AXIS == 0) { s = e / LEN; idx = e - s * LEN; }
    else { idx = e / (2 * P); s = e - idx * (2 * P); }
    if constexpr (kPost) {
      // evaluated at an existing place whatever the test, then selected: the
      // loader's loads of every element are issued without a branch
      const bool in = NL * NT == TOT || e < TOT;
      const int ec = in ? e : TOT - 1;
      if (AXIS == 0) { s = ec / LEN; idx = ec - s * LEN; }
      else { idx = ec / (2 * P); s = ec - idx * (2 * P); }
      const T v = op.load_aux(b, g, s, P, idx, aux[q]);
      x[q] = in ? v : T(0);
    } else {
      x[q] = (NL * NT == TOT || e < TOT) ? op.load(b, g, s, P, idx) : T(0);
    }
  }
#pragma unroll
  for (int q = 0; q < NL; q++) {
    const int e = t + q * NT;
    if (NL * NT == TOT || e < TOT) {
      int s, idx;
      if (AXIS == 0) { s = e / LEN; idx = e - s * LEN; }
      else { idx = e / (2 * P); s = e - idx * (2 * P); }
      const int p = s >> 1;
      const int h = s & 1;
      reinterpret_cast<T*>(z + p * N2 + idx + 1)[h] = x[q];
      reinterpret_cast<T*>(z + p * N2 + N2 - idx - 1)[h] = -x[q];
      if constexpr (kPost) op.load_post(b, g, s, P, idx, aux[q]);
    }
  }
}

// One workgroup's part of a size-specialised pass: group g of ng (before
// the XCD-contiguous remap), LDS buffer z of (DB ? 2 : 1) P N2 entries.
template <typename T, int N2, int P, int AXIS, class Op>
__device__ __forceinline__ void dst_pass_body(const Op& op, const cplx<T>* __restrict__ tw, cplx<T>* z, int g,
                                              int ng) {
  constexpr int N = N2 / 2, LEN = N - 1, NT = P * N2 / 4;
  constexpr int NQ = (P * LEN + NT - 1) / NT;  // outputs (FFT, k) per thread
  constexpr bool kPre = op_store_pre<Op>::value;
  constexpr bool DB = kTwoBuf<T, N2, P>;
  const int t = threadIdx.x;
  const int b = blockIdx.y;
  if ((ng & 7) == 0) g = (g & 7) * (ng >> 3) + (g >> 3);
  auto out_at = [&](int e, int& p, int& k) {
    if (AXIS == 0) { p = e / LEN; k = e - p * LEN + 1; }
    else { k = e / P + 1; p = e - (k - 1) * P; }
  };
  // old values of a read-modify-write destination, read with the inputs
  T oa[kPre ? NQ : 1], ob[kPre ? NQ : 1];
  if constexpr (kPre) {
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const int e = t + q * NT;
      oa[q] = ob[q] = T(0);
      if (e < P * LEN) {
        int p, k;
        out_at(e, p, k);
        op.store_pre(b, g, p, P, k, oa[q], ob[q]);
      }
    }
  }
  cplx<T> wr[TwLayout<N2>::total()];
  tw_load<T, N2, 0>(wr, tw, t % (N2 / 4));
  if (t < P) {
    z[t * N2] = {T(0), T(0)};
    z[t * N2 + N] = {T(0), T(0)};
  }
  load_seqs<T, N2, P, AXIS, Op>(op, z, b, g, t);
  __syncthreads();
  const cplx<T>* zr = run<T, N2, P, 0, DB>(z, DB ? z + P * N2 : z, wr, t);
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const int e = t + q * NT;
    if (e < P * LEN) {
      int p, k;
      out_at(e, p, k);
      const cplx<T> Z = zr[p * N2 + k];
      if constexpr (kPre) op.store_put(b, g, p, P, k, -Z.im, Z.re, oa[q], ob[q]);
      else op.store(b, g, p, P, k, -Z.im, Z.re);
    }
  }
}

template <typename T, int N2, int P, int AXIS, class Op>
__global__ __launch_bounds__(P * N2 / 4) void dst_pass_n(Op op, const cplx<T>* __restrict__ tw) {
  constexpr int NT = P * N2 / 4;
  constexpr bool DB = kTwoBuf<T, N2, P>;
  __shared__ cplx<T> z[(DB ? 2 : 1) * P * N2];
  int g = blockIdx.x, ng = gridDim.x;
  if constexpr (op_side<Op>::value) {
    const int sg = op.side_groups_for(NT);
    if (g < sg) {
      op.side(blockIdx.y, g, threadIdx.x);
      return;
    }
    g -= sg;
    ng -= sg;
  }
  dst_pass_body<T, N2, P, AXIS, Op>(op, tw, z, g, ng);
}

// Two independent passes of one FFT length in one launch: workgroups
// x < ga run the row pass A, the others the column pass B (ops without
// side work; they must write disjoint data).  Latency-bound pairs of small
// passes then cost one launch and one load-transform-store chain.
template <typename T, int N2, int P, class OpA, class OpB>
__global__ __launch_bounds__(P * N2 / 4) void dst_pass_dual(OpA a, OpB bq, const cplx<T>* __restrict__ twa,
                                                            const cplx<T>* __restrict__ twb, int ga, int nya,
                                                            int nyb) {
  constexpr bool DB = kTwoBuf<T, N2, P>;
  __shared__ cplx<T> z[(DB ? 2 : 1) * P * N2];
  const int x = blockIdx.x;
  // batch entries blockIdx.y < nya of pass A, < nyb of pass B
  if (x < ga) {
    if ((int)blockIdx.y < nya) dst_pass_body<T, N2, P, 0, OpA>(a, twa, z, x, ga);
  } else if ((int)blockIdx.y < nyb) {
    dst_pass_body<T, N2, P, 1, OpB>(bq, twb, z, x - ga, (int)gridDim.x - ga);
  }
}

// fused column pair (see dst_col2) at a compile-time size
template <typename T, int N2, int P, class Op>
__global__ __launch_bounds__(P * N2 / 4) void dst_col2_n(Op op, const cplx<T>* __restrict__ tw) {
  constexpr int N = N2 / 2, LEN = N - 1, NT = P * N2 / 4;
  constexpr int NQ = (P * LEN + NT - 1) / NT;  // (FFT, k) places per thread
  constexpr bool kCoef = op_has_mid_coef<Op>::value;
  constexpr bool kPre = op_store_pre<Op>::value;
  constexpr bool DB = kTwoBuf<T, N2, P>;
  __shared__ cplx<T> z[(DB ? 2 : 1) * P * N2];
  cplx<T>* const z2 = DB ? z + P * N2 : z;
  const int t = threadIdx.x;
  const int b = blockIdx.y;
  int g = blockIdx.x;
  if ((gridDim.x & 7) == 0) g = (g & 7) * (gridDim.x >> 3) + (g >> 3);
  // the spectral table entries of this thread's places, read with the inputs
  T ca[kCoef ? NQ : 1], cb[kCoef ? NQ : 1];
  if constexpr (kCoef) {
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const int e = t + q * NT;
      ca[q] = cb[q] = T(0);
      if (e < P * LEN) {
        const int k = e / P + 1, p = e - (k - 1) * P;
        op.mid_coef(b, g, p, P, k, ca[q], cb[q]);
      }
    }
  }
  // old values of a read-modify-write destination, read with the inputs
  T oa[kPre ? NQ : 1], ob[kPre ? NQ : 1];
  if constexpr (kPre) {
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const int e = t + q * NT;
      oa[q] = ob[q] = T(0);
      if (e < P * LEN) {
        const int k = e / P + 1, p = e - (k - 1) * P;
        op.store_pre(b, g, p, P, k, oa[q], ob[q]);
      }
    }
  }
  cplx<T> wr[TwLayout<N2>::total()];
  tw_load<T, N2, 0>(wr, tw, t % (N2 / 4));
  if (t < P) {
    z[t * N2] = {T(0), T(0)};
    z[t * N2 + N] = {T(0), T(0)};
  }
  load_seqs<T, N2, P, 1, Op>(op, z, b, g, t);
  __syncthreads();
  // the middle step works in place on the first transform's result (each
  // place (p, k) is read and rewritten by one thread; the mirrored places
  // N2-k and the zeros at 0 and N are read by no one)
  cplx<T>* const zm = run<T, N2, P, 0, DB>(z, z2, wr, t);
  if constexpr (kCoef) {
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const int e = t + q * NT;
      if (e < P * LEN) {
        const int k = e / P + 1, p = e - (k - 1) * P;
        const cplx<T> Z = zm[p * N2 + k];
        T va, vb;
        op.mid_apply(b, g, p, P, k, -Z.im, Z.re, ca[q], cb[q], va, vb);
        zm[p * N2 + k] = {va, vb};
        zm[p * N2 + N2 - k] = {-va, -vb};
      }
    }
    if (t < P) {
      zm[t * N2] = {T(0), T(0)};
      zm[t * N2 + N] = {T(0), T(0)};
    }
    __syncthreads();
  } else {
    col2_mid<T, Op>(op, zm, b, g, P, N2, t, NT);
  }
  const cplx<T>* zr = run<T, N2, P, 0, DB>(zm, zm == z ? z2 : z, wr, t);
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const int e = t + q * NT;
    if (e < P * LEN) {
      const int k = e / P + 1, p = e - (k - 1) * P;
      const cplx<T> Z = zr[p * N2 + k];
      if constexpr (kPre) op.store_put(b, g, p, P, k, -Z.im, Z.re, oa[q], ob[q]);
      else op.store(b, g, p, P, k, -Z.im, Z.re);
    }
  }
}

}  // namespace spec

}  // namespace ibfs
