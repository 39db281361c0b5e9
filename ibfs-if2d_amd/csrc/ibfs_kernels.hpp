// ibfs_kernels.hpp -- device side of one multi-domain IBFS time step for a
// group of B independent cases that share the grid size m x n.
//
// Index conventions (reference src/grid.f90:119-134, src/variables.f90:46):
//   vertex field  w[ii + jj*(m-1)], ii = i-2, jj = j-2 for the reference's
//                 interior vertex (i,j), i = 2..m, j = 2..n;
//   flux faces    u(i,j) -> (i-1) + (j-1)(m+1), i = 1..m+1, j = 1..n;
//                 v(i,j) -> (m+1)n + (i-1) + (j-1)m, i = 1..m, j = 1..n+1;
//   edge vector   the reference's 1-based position pos lives at pos-1, with
//                 segment bases left = 0, right = n+1, bottom = 2(n+1),
//                 top = 2(n+1)+m+1 (src/grid.f90:73-76).
// Multi-level fields of case b are stored level after level in Lmax slots:
// field(b, lev) = base + (b*Lmax + lev)*size.  Case b uses levels
// 0..Lc[b]-1; a kernel launched for level slot lev skips cases with
// lev >= Lc[b], and level Lc[b]-1 is that case's coarsest (zero far field).
// The element type T is double (plain form) or float (mixed form); the force
// system, smearing weights and all reductions stay in double.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "dst_core.hpp"

namespace ibfs {

struct Geo {
  int m, n, Lmax;
  int mm1, nm1;  // m-1, n-1
  int nw, nq, nbc;
  int left, right, bottom, top;
};

__device__ __forceinline__ int uix(const Geo& g, int i, int j) { return (i - 1) + (j - 1) * (g.m + 1); }
__device__ __forceinline__ int vix(const Geo& g, int i, int j) { return (g.m + 1) * g.n + (i - 1) + (j - 1) * g.m; }
__device__ __forceinline__ int wix(const Geo& g, int i, int j) { return (i - 2) + (j - 2) * g.mm1; }

// Lc == nullptr: every case has all g.Lmax levels (no per-case load, so no
// kernel waits on one before its first field load)
__device__ __forceinline__ bool active(const int* Lc, int b, int lev) { return Lc == nullptr || lev < Lc[b]; }
__device__ __forceinline__ int levels(const Geo& g, const int* Lc, int b) { return Lc == nullptr ? g.Lmax : Lc[b]; }
__device__ __forceinline__ bool has_coarser(const Geo& g, const int* Lc, int b, int lev) {
  return lev + 1 < levels(g, Lc, b);
}

// get_bc (src/operators.f90:336-367) evaluated at one 1-based edge position.
// r is the coarser level's vertex field; the reference's assumed-shape dummy
// gives r(a,b) = omega(a+1,b+1) = w[(a-1) + (b-1)(m-1)].
template <typename T>
__device__ __forceinline__ T getbc_at(const Geo& g, const T* __restrict__ r, int pos, T fac) {
  auto R1 = [&](int a, int b) -> T { return r[(a - 1) + (b - 1) * g.mm1]; };
  const int m = g.m, n = g.n;
  T v;
  if (pos > g.bottom) {  // bottom or top: index along i
    const bool top = pos > g.top;
    const int i0 = pos - (top ? g.top : g.bottom) - 1;  // 0..m
    const int jb = top ? 3 * n / 4 : n / 4;
    if ((i0 & 1) == 0) v = R1(m / 4 + i0 / 2, jb);
    else v = T(0.5) * (R1(m / 4 + (i0 + 1) / 2, jb) + R1(m / 4 - 1 + (i0 + 1) / 2, jb));
  } else {  // left or right: index along j
    const bool right = pos > g.right;
    const int j0 = pos - (right ? g.right : g.left) - 1;  // 0..n
    const int ib = right ? 3 * m / 4 : m / 4;
    if ((j0 & 1) == 0) v = R1(ib, n / 4 + j0 / 2);
    else v = T(0.5) * (R1(ib, n / 4 + (j0 + 1) / 2) + R1(ib, n / 4 - 1 + (j0 + 1) / 2));
  }
  return v * fac;
}

// apply_bc (src/operators.f90:370-390) contribution at vertex (i,j) for an
// edge vector given as get_bc(rc, gfac) of the coarser field rc.
// Reference order: bottom/top first (i loop), then left/right (j loop).
template <typename T>
__device__ __forceinline__ T applybc_add(const Geo& g, T acc, int i, int j, const T* __restrict__ rc, T gfac, T fac) {
  if (j == 2) acc = acc + fac * getbc_at<T>(g, rc, g.bottom + i, gfac);
  if (j == g.n) acc = acc + fac * getbc_at<T>(g, rc, g.top + i, gfac);
  if (i == 2) acc = acc + fac * getbc_at<T>(g, rc, g.left + j, gfac);
  if (i == g.m) acc = acc + fac * getbc_at<T>(g, rc, g.right + j, gfac);
  return acc;
}

// same for an explicit edge vector bc (positions are 1-based: bc[pos-1])
template <typename T>
__device__ __forceinline__ T applybc_vec(const Geo& g, T acc, int i, int j, const T* __restrict__ bc, T fac) {
  if (j == 2) acc = acc + fac * bc[g.bottom + i - 1];
  if (j == g.n) acc = acc + fac * bc[g.top + i - 1];
  if (i == 2) acc = acc + fac * bc[g.left + j - 1];
  if (i == g.m) acc = acc + fac * bc[g.right + j - 1];
  return acc;
}

// Branch-free forms of getbc_at, applybc_add and applybc_vec for loaders
// that issue all their loads before using any: every candidate value is
// loaded from an in-range place and the terms are then selected, with the
// same operands, order and arithmetic, so the results are identical.  (The
// bottom/top and left/right conditions exclude each other for m, n > 2.)
template <typename T>
__device__ __forceinline__ T getbc_sel(const Geo& g, const T* __restrict__ r, int pos, T fac) {
  const bool horiz = pos > g.bottom;
  const bool hi = horiz ? pos > g.top : pos > g.right;
  const int k0 = pos - (horiz ? (hi ? g.top : g.bottom) : (hi ? g.right : g.left)) - 1;
  const bool odd = (k0 & 1) != 0;
  const int c1 = (horiz ? g.m : g.n) / 4 + (odd ? (k0 + 1) / 2 : k0 / 2);
  const int fx = horiz ? (hi ? 3 * g.n / 4 : g.n / 4) : (hi ? 3 * g.m / 4 : g.m / 4);
  auto ix = [&](int a, int c) { return horiz ? (a - 1) + (c - 1) * g.mm1 : (c - 1) + (a - 1) * g.mm1; };
  const int a1 = ix(c1, fx), a2 = odd ? ix(c1 - 1, fx) : a1;
  // both values are used whatever odd is (0.5 (v1 + v1) == v1 exactly), so
  // neither load is moved into a branch
  const T v1 = r[a1], v2 = r[a2];
  const T v = T(0.5) * (v1 + (odd ? v2 : v1));
  return v * fac;
}
// getbc_sel in two parts: the two places it reads (the same place twice
// when it needs one), and the value from what was read there
__device__ __forceinline__ void getbc_places(const Geo& g, int pos, int& a1, int& a2) {
  const bool horiz = pos > g.bottom;
  const bool hi = horiz ? pos > g.top : pos > g.right;
  const int k0 = pos - (horiz ? (hi ? g.top : g.bottom) : (hi ? g.right : g.left)) - 1;
  const bool odd = (k0 & 1) != 0;
  const int c1 = (horiz ? g.m : g.n) / 4 + (odd ? (k0 + 1) / 2 : k0 / 2);
  const int fx = horiz ? (hi ? 3 * g.n / 4 : g.n / 4) : (hi ? 3 * g.m / 4 : g.m / 4);
  auto ix = [&](int a, int c) { return horiz ? (a - 1) + (c - 1) * g.mm1 : (c - 1) + (a - 1) * g.mm1; };
  a1 = ix(c1, fx);
  a2 = odd ? ix(c1 - 1, fx) : a1;
}
template <typename T>
__device__ __forceinline__ T getbc_value(const Geo& g, int pos, T v1, T v2, T fac) {
  const bool horiz = pos > g.bottom;
  const bool hi = horiz ? pos > g.top : pos > g.right;
  const int k0 = pos - (horiz ? (hi ? g.top : g.bottom) : (hi ? g.right : g.left)) - 1;
  const bool odd = (k0 & 1) != 0;
  const T v = T(0.5) * (v1 + (odd ? v2 : v1));
  return v * fac;
}
template <typename T>
__device__ __forceinline__ T applybc_add_sel(const Geo& g, T acc, int i, int j, const T* __restrict__ rc, T gfac,
                                             T fac) {
  const bool ej = j == 2 || j == g.n, ei = i == 2 || i == g.m;
  const T gj = getbc_sel<T>(g, rc, (j == g.n ? g.top : g.bottom) + i, gfac);
  const T gi = getbc_sel<T>(g, rc, (i == g.m ? g.right : g.left) + j, gfac);
  acc = ej ? acc + fac * gj : acc;
  acc = ei ? acc + fac * gi : acc;
  return acc;
}
template <typename T>
__device__ __forceinline__ T applybc_vec_sel(const Geo& g, T acc, int i, int j, const T* __restrict__ bc, T fac) {
  const bool ej = j == 2 || j == g.n, ei = i == 2 || i == g.m;
  const T vj = bc[(j == g.n ? g.top : g.bottom) + i - 1];
  const T vi = bc[(i == g.m ? g.right : g.left) + j - 1];
  acc = ej ? acc + fac * vj : acc;
  acc = ei ? acc + fac * vi : acc;
  return acc;
}

// rot(nonlinear(omega, q, q0, bc)) at vertex (i,j):
// nonlinear src/operators.f90:465-510, rot src/operators.f90:291-304.
// kFree: the freestream form of q0 (src/variables.f90:59-63), u0 on every u
// face and 0 on every v face of the level (q0 is not read), with the same
// additions as the array form, so the same result.
template <typename T, bool kFree = false>
__device__ __forceinline__ T rot_nonlinear(const Geo& g, int i, int j, const T* __restrict__ om,
                                           const T* __restrict__ q, const T* __restrict__ q0,
                                           const T* __restrict__ bc, T u0 = T(0)) {
  const int m = g.m, n = g.n;
  // the edge value or the field value, one load from the selected place
  auto OmU = [&](int ii, int jj) -> T {  // along j at column ii (2..m)
    const T* p = jj == 1 ? bc + (g.bottom + ii - 1) : jj == n + 1 ? bc + (g.top + ii - 1) : om + wix(g, ii, jj);
    return *p;
  };
  auto OmV = [&](int ii, int jj) -> T {  // along i at row jj (2..n)
    const T* p = ii == 1 ? bc + (g.left + jj - 1) : ii == m + 1 ? bc + (g.right + jj - 1) : om + wix(g, ii, jj);
    return *p;
  };
  auto vavg = [&](int ii, int jj) -> T {
    const int a = vix(g, ii, jj), b = vix(g, ii - 1, jj);
    if constexpr (kFree) return T(0.5) * (q[a] + q[b] + T(0) + T(0));
    else return T(0.5) * (q[a] + q[b] + q0[a] + q0[b]);
  };
  auto uavg = [&](int ii, int jj) -> T {
    const int a = uix(g, ii, jj), b = uix(g, ii, jj - 1);
    if constexpr (kFree) return T(0.5) * (q[a] + q[b] + u0 + u0);
    else return T(0.5) * (q[a] + q[b] + q0[a] + q0[b]);
  };
  auto fqu = [&](int ii, int jj) -> T {  // fq(u(ii,jj)), ii = 2..m, jj = 1..n
    return T(0.5) * (vavg(ii, jj + 1) * OmU(ii, jj + 1) + vavg(ii, jj) * OmU(ii, jj));
  };
  auto fqv = [&](int ii, int jj) -> T {  // fq(v(ii,jj)), ii = 1..m, jj = 2..n
    return T(-0.5) * (uavg(ii + 1, jj) * OmV(ii + 1, jj) + uavg(ii, jj) * OmV(ii, jj));
  };
  return fqv(i, j) - fqv(i - 1, j) - fqu(i, j) + fqu(i, j - 1);
}

// ---------------------------------------------------------------------------
// smearing tables (E / E^T) on compact per-case tables, see the smearing section
// ---------------------------------------------------------------------------
// Active faces of case b: face_id[fbase+a] (0-based face of that case's q),
// entries face_ptr[pbase+a]..face_ptr[pbase+a+1] of (face_pt, face_w): the
// case-local force index (x or y component) and the weight, in increasing
// IB-point order (the ELL order of src/grid.f90:197-228).
// Force gather: force f of case b has entries frc_ptr[rbase+f]..+1 of
// (frc_face, frc_w) in increasing face order (the accumulation order of
// regT, src/operators.f90:541-545).
// Active vertices: vtx_id[vbase+v] (0-based W index) and
// vtx_face[4(vbase+v)..+3] = active-face slots (case-local) of v(i,j),
// v(i-1,j), u(i,j), u(i,j-1) (-1: face has no weight).
struct CaseSm {
  int nact, fbase, pbase;  // faces
  int nf, rbase;           // forces
  int nvtx, vbase;         // vertices
  int foff;                // offset of the case's force vectors
  int nb;                  // IB points
  long loff;               // offset of the case's Nf x Nf matrices
  long goff;               // offset of the case's nvtx x Nf matrix rot(reg(A^-1))
  int nbody, bmin, ncol, hoff;  // diagnostics row layout
  double dt, delta;
};

struct Smear {
  const int* face_id;
  const int* face_ptr;
  const int* face_pt;
  const double* face_w;
  const int* frc_ptr;
  const int* frc_face;
  const double* frc_w;
  const int* frc_ia;  // per force entry: curl(s) = s[ia] - s[ib] at an interior face, ia = -1 on the edge
  const int* frc_ib;
  const int* vtx_id;
  const int* vtx_face;
  const CaseSm* cs;
  // a group of one case also carries its record by value, so the kernels
  // of the step start their table loads without first loading it
  CaseSm c0;
  int single;
  // the force gather padded to kFrcK entries per force, slot-major: entry k
  // of force foff + f at k * f9_n + foff + f (zero weight and index 0 in the
  // unused slots), so a force's entries are loaded without first loading
  // its range and a wave's loads of one slot are contiguous; nullptr when
  // some force has more entries
  const int* f9_face;
  const double* f9_w;
  const int* f9_ia;
  const int* f9_ib;
  int f9_n;
};
__device__ __forceinline__ CaseSm case_sm(const Smear& sm, int b) {
  if (sm.single) return sm.c0;
  return sm.cs[b];
}

// Arguments of the diagnostics row (k_diag), also run as side work of the
// next step's first pass.  pend == nullptr: the row goes to slot counter[b],
// which is then incremented.  Otherwise the row is deferred: k_tail sets
// pend[b] and increments counter[b], and the row (slot counter[b] - 1) is
// written only while pend[b] is set; clear: pend[b] is then cleared (the
// flush after the last step; the side work leaves it to the flush).
template <typename T>
struct DiagArgs {
  Smear sm;
  const T* q; long qbs;
  const double* motion; const double* fb; const double* wsum; const int* bdyno;
  const double* partial; int nchunk;
  const int* Lc;
  double* out_base; int rowlen;
  int* counter; int capacity;
  int* pend; int clear;
};
template <typename T>
__device__ void diag_row(const Geo& g, const DiagArgs<T>& a, int b, double* red, int part);

// The head of a level-batched step: lastbc of every level (as k_lastbc) in
// workgroups x < ceil(nbc Lmax / 256), and with has_diag the previous step's
// pending diagnostics row in three more (diag_row parts 0-2), whose inputs
// no launch of the step changes before k_tail.  grid (.., B), 256 threads.
template <typename T>
__global__ __launch_bounds__(256) void k_step_head(Geo g, const T* __restrict__ om, T* __restrict__ lastbc,
                                                   const int* __restrict__ Lc, DiagArgs<T> dg, int has_diag) {
  const int nlb = (g.nbc * g.Lmax + 255) / 256, b = blockIdx.y;
  if ((int)blockIdx.x >= nlb) {
    __shared__ double red[16];
    if (has_diag) diag_row<T>(g, dg, b, red, blockIdx.x - nlb);
    return;
  }
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= g.nbc * g.Lmax) return;
  const int lev = t / g.nbc, p = t - lev * g.nbc;
  const long fb = (long)b * g.Lmax;
  T v = T(0);
  if (has_coarser(g, Lc, b, lev)) v = getbc_at<T>(g, om + (fb + lev + 1) * g.nw, p + 1, T(0.25));
  lastbc[(fb + lev) * g.nbc + p] = v;
}

// ---------------------------------------------------------------------------
// pass operators (see dst_core.hpp for the load/store protocol)
// ---------------------------------------------------------------------------
template <typename T>
struct SrcField {
  const T* a; long bs;
  __device__ T operator()(const Geo& g, int b, int ii, int jj) const { return a[b * bs + ii + jj * g.mm1]; }
};

// omega_k + apply_bc(get_bc(s_{k+1}, 1), 1): the vort of vort2flux
// (src/operators.f90:230-234); no edge term on a case's coarsest level.
// TS: element type of the result (FP64 for the FP64 solves of an FP32
// engine, whose omega is FP32); TC: element type of the coarser level's
// streamfunction sc.
template <typename T, typename TS = T, typename TC = TS>
struct SrcVortBC {
  const T* om; const TC* sc; long bs; const int* Lc; int lev;
  // two phases for the size-specialised passes (see SrcStep1::gather): the
  // vertex value and the four get_bc values an edge vertex may add
  static constexpr bool kGather = true;
  struct Gathered {
    T v;
    TC gj1, gj2, gi1, gi2;
  };
  __device__ void gather(const Geo& g, int b, int ii, int jj, Gathered& G) const {
    const int i = ii + 2, j = jj + 2;
    G.v = om[b * bs + ii + jj * g.mm1];
    G.gj1 = G.gj2 = G.gi1 = G.gi2 = TC(0);
    // only edge vertices add get_bc values (a branch around loads alone
    // waits for nothing)
    if (sc != nullptr && (i == 2 || i == g.m || j == 2 || j == g.n)) {
      const TC* rc = sc + b * bs;
      int a1, a2;
      getbc_places(g, (j == g.n ? g.top : g.bottom) + i, a1, a2);
      G.gj1 = rc[a1];
      G.gj2 = rc[a2];
      getbc_places(g, (i == g.m ? g.right : g.left) + j, a1, a2);
      G.gi1 = rc[a1];
      G.gi2 = rc[a2];
    }
  }
  // the value of operator() from the gathered loads (applybc_add_sel order)
  __device__ TS combine(const Geo& g, int b, int ii, int jj, const Gathered& G, TS& aux) const {
    TS v = TS(G.v);
    if (sc == nullptr || !has_coarser(g, Lc, b, lev)) return v;
    const int i = ii + 2, j = jj + 2;
    const bool ej = j == 2 || j == g.n, ei = i == 2 || i == g.m;
    const TC gj = getbc_value<TC>(g, (j == g.n ? g.top : g.bottom) + i, G.gj1, G.gj2, TC(1));
    const TC gi = getbc_value<TC>(g, (i == g.m ? g.right : g.left) + j, G.gi1, G.gi2, TC(1));
    if constexpr (std::is_same_v<TS, TC>) {
      v = ej ? v + TS(1) * gj : v;
      v = ei ? v + TS(1) * gi : v;
    } else {
      TC acc = TC(0);
      acc = ej ? acc + TC(1) * gj : acc;
      acc = ei ? acc + TC(1) * gi : acc;
      v = (ej || ei) ? v + TS(acc) : v;
    }
    return v;
  }
  __device__ TS operator()(const Geo& g, int b, int ii, int jj) const {
    TS v = TS(om[b * bs + ii + jj * g.mm1]);
    if (sc == nullptr || !has_coarser(g, Lc, b, lev)) return v;
    const int i = ii + 2, j = jj + 2;
    if (i == 2 || i == g.m || j == 2 || j == g.n) {
      if constexpr (std::is_same_v<TS, TC>) v = applybc_add<TS>(g, v, i, j, sc + b * bs, TS(1), TS(1));
      else v = v + TS(applybc_add<TC>(g, TC(0), i, j, sc + b * bs, TC(1), TC(1)));
    }
    return v;
  }
};

template <typename T>
struct DstField {
  T* a; long bs;
  __device__ void operator()(const Geo& g, int b, int ii, int jj, T v) const { a[b * bs + ii + jj * g.mm1] = v; }
};
// a = a - v (projection, src/operators.f90:174).  Read-modify-write
// destinations also offer pre (the old value, read by the size-specialised
// passes together with their inputs) and put (the store given that value),
// which give the same result as operator().
// (skip: per batch entry, nonzero leaves that entry unwritten)
template <typename T>
struct DstSub {
  T* a; long bs; const int* skip = nullptr;
  __device__ void operator()(const Geo& g, int b, int ii, int jj, T v) const {
    if (skip != nullptr && skip[b]) return;
    const long e = b * bs + ii + jj * g.mm1;
    a[e] = a[e] - v;
  }
  __device__ T pre(const Geo& g, int b, int ii, int jj) const { return a[b * bs + ii + jj * g.mm1]; }
  __device__ void put(const Geo& g, int b, int ii, int jj, T v, T old) const {
    if (skip != nullptr && skip[b]) return;
    a[b * bs + ii + jj * g.mm1] = old - v;
  }
};

// Dst::pre present?
template <class D, class = void>
struct dst_has_pre : std::false_type {};
template <class D>
struct dst_has_pre<D, std::void_t<decltype(&D::pre)>> : std::true_type {};
// Src::post present (a loader with a side store)?
// Src::kNoWave: keep this loader off the register-blocked passes
template <class S, class = void>
struct src_no_wave : std::false_type {};
template <class S>
struct src_no_wave<S, std::void_t<decltype(S::kNoWave)>> : std::bool_constant<S::kNoWave> {};
template <class S, class = void>
struct src_has_post : std::false_type {};
template <class S>
struct src_has_post<S, std::void_t<decltype(&S::post)>> : std::true_type {};
// Src::side present (independent work done by extra workgroups of the pass)?
template <class S, class = void>
struct src_has_side : std::false_type {};
template <class S>
struct src_has_side<S, std::void_t<decltype(&S::side)>> : std::true_type {};
// Src::gather/combine present and enabled (a loader in two phases)?
template <class S, class = void>
struct src_gathered {
  using type = char;
  static constexpr bool value = false;
};
template <class S>
struct src_gathered<S, std::void_t<typename S::Gathered>> {
  using type = typename S::Gathered;
  static constexpr bool value = S::kGather;
};

// single-field row pass: rows jj; FFT p pairs rows 2Pg+2p and 2Pg+2p+1.
// Cases whose level count is <= lev are skipped (Lc == nullptr: none).
template <typename T, class Src, class Dst>
struct RowOp {
  static constexpr bool kNoWave = src_no_wave<Src>::value;
  Geo g; Src src; Dst dst; const int* Lc; int lev;
  __device__ T load(int b, int gr, int s, int P, int idx) const {
    const int row = gr * 2 * P + s;
    return (row < g.nm1 && active(Lc, b, lev)) ? src(g, b, idx, row) : T(0);
  }
  // loaders with a side store (Src::value/post): load without the store,
  // then load_post once every load of the thread has been issued
  static constexpr bool kLoadPost = src_has_post<Src>::value;
  __device__ T load_aux(int b, int gr, int s, int P, int idx, T& aux) const {
    aux = T(0);
    if constexpr (kLoadPost) {
      // evaluated on a row that exists whatever the test, so no branch holds
      // the loads back
      const int row = gr * 2 * P + s;
      const T v = src.value(g, b, idx, row < g.nm1 ? row : g.nm1 - 1, aux);
      return (row < g.nm1 && active(Lc, b, lev)) ? v : T(0);
    } else {
      return load(b, gr, s, P, idx);
    }
  }
  __device__ void load_post(int b, int gr, int s, int P, int idx, T aux) const {
    if constexpr (kLoadPost) {
      const int row = gr * 2 * P + s;
      if (row < g.nm1 && active(Lc, b, lev)) src.post(g, b, idx, row, aux);
    }
  }
  // side work of the loader (Src::side), done by extra workgroups
  static constexpr bool kSide = src_has_side<Src>::value;
  __host__ __device__ int side_groups_for(int nt) const {
    if constexpr (kSide) return src.side_groups_for(nt);
    else return 0;
  }
  __device__ void side(int b, int x, int t) const {
    if constexpr (kSide) src.side(g, b, x, t);
  }
  // loaders in two phases (Src::gather/combine): the loads of an element on
  // a row that exists whatever the test, then the value
  static constexpr bool kGather = src_gathered<Src>::value;
  using Gathered = typename src_gathered<Src>::type;
  __device__ void gather(int b, int gr, int s, int P, int idx, Gathered& G) const {
    const int row = gr * 2 * P + s;
    src.gather(g, b, idx, row < g.nm1 ? row : g.nm1 - 1, G);
  }
  __device__ T combine(int b, int gr, int s, int P, int idx, const Gathered& G, T& aux) const {
    const int row = gr * 2 * P + s;
    const T v = src.combine(g, b, idx, row < g.nm1 ? row : g.nm1 - 1, G, aux);
    return (row < g.nm1 && active(Lc, b, lev)) ? v : T(0);
  }

  __device__ void store(int b, int gr, int p, int P, int k, T ya, T yb) const {
    if (!active(Lc, b, lev)) return;
    const int r0 = gr * 2 * P + 2 * p, r1 = r0 + 1;
    if (r0 < g.nm1) dst(g, b, k - 1, r0, ya);
    if (r1 < g.nm1) dst(g, b, k - 1, r1, yb);
  }
  // prefetched form of store (only when Dst has pre/put)
  static constexpr bool kStorePre = dst_has_pre<Dst>::value;
  __device__ void store_pre(int b, int gr, int p, int P, int k, T& oa, T& ob) const {
    oa = ob = T(0);
    if constexpr (kStorePre) {
      if (!active(Lc, b, lev)) return;
      const int r0 = gr * 2 * P + 2 * p, r1 = r0 + 1;
      if (r0 < g.nm1) oa = dst.pre(g, b, k - 1, r0);
      if (r1 < g.nm1) ob = dst.pre(g, b, k - 1, r1);
    }
  }
  __device__ void store_put(int b, int gr, int p, int P, int k, T ya, T yb, T oa, T ob) const {
    if constexpr (kStorePre) {
      if (!active(Lc, b, lev)) return;
      const int r0 = gr * 2 * P + 2 * p, r1 = r0 + 1;
      if (r0 < g.nm1) dst.put(g, b, k - 1, r0, ya, oa);
      if (r1 < g.nm1) dst.put(g, b, k - 1, r1, yb, ob);
    }
  }
};
// single-field column pass: columns ii
template <typename T, class Src, class Dst>
struct ColOp {
  static constexpr bool kNoWave = src_no_wave<Src>::value;
  Geo g; Src src; Dst dst; const int* Lc; int lev;
  __device__ T load(int b, int gr, int s, int P, int idx) const {
    const int col = gr * 2 * P + s;
    return (col < g.mm1 && active(Lc, b, lev)) ? src(g, b, col, idx) : T(0);
  }
  __device__ void store(int b, int gr, int p, int P, int k, T ya, T yb) const {
    if (!active(Lc, b, lev)) return;
    const int c0 = gr * 2 * P + 2 * p, c1 = c0 + 1;
    if (c0 < g.mm1) dst(g, b, c0, k - 1, ya);
    if (c1 < g.mm1) dst(g, b, c1, k - 1, yb);
  }
};

// spectral factors of the fused column pair (value at column ii, row jj)
// Their tables are stored pair-major: the entries of columns 2p and 2p+1 at
// row jj are adjacent, at (p (n-1) + jj) * 2, so the workgroup of a column
// pair reads its entries contiguously (pm_size(g) entries per table, the
// last column padded when m-1 is odd).
__host__ __device__ __forceinline__ long pm_ix(const Geo& g, int ii, int jj) {
  return ((long)(ii >> 1) * g.nm1 + jj) * 2 + (ii & 1);
}
__host__ __device__ __forceinline__ long pm_size(const Geo& g) { return (long)((g.mm1 + 1) / 2) * g.nm1 * 2; }
// tab * v: ctci's laminv (src/myfft.f90:106), shared by the group
// coef gives the table entry and apply the factor given that entry (the
// size-specialised column pair reads the entries together with its inputs);
// apply(coef(...), v) == operator()(..., v).
template <typename T>
struct MidScale {
  const T* tab;
  __device__ T operator()(const Geo& g, int b, int ii, int jj, T v) const { return tab[pm_ix(g, ii, jj)] * v; }
  __device__ T coef(const Geo& g, int b, int ii, int jj) const { return tab[pm_ix(g, ii, jj)]; }
  __device__ T apply(T c, T v) const { return c * v; }
};
// tab * v / norm: ainv's lam1i(:,:,1) / normalize (src/myfft.f90:121), per case
template <typename T>
struct MidScaleDiv {
  const T* tab; T norm; long tbs;
  __device__ T operator()(const Geo& g, int b, int ii, int jj, T v) const {
    return tab[b * tbs + pm_ix(g, ii, jj)] * v / norm;
  }
  __device__ T coef(const Geo& g, int b, int ii, int jj) const { return tab[b * tbs + pm_ix(g, ii, jj)]; }
  __device__ T apply(T c, T v) const { return c * v / norm; }
};

// fused column pair: columns ii, Sy . diag(mid) . Sy
template <typename T, class Src, class Mid, class Dst>
struct Col2Op {
  Geo g; Src src; Mid md; Dst dst; const int* Lc; int lev;
  __device__ T load(int b, int gr, int s, int P, int idx) const {
    const int col = gr * 2 * P + s;
    return (col < g.mm1 && active(Lc, b, lev)) ? src(g, b, col, idx) : T(0);
  }
  __device__ void mid(int b, int gr, int p, int P, int k, T ya, T yb, T& va, T& vb) const {
    const int c0 = gr * 2 * P + 2 * p, c1 = c0 + 1;
    const bool on = active(Lc, b, lev);
    va = (on && c0 < g.mm1) ? md(g, b, c0, k - 1, ya) : T(0);
    vb = (on && c1 < g.mm1) ? md(g, b, c1, k - 1, yb) : T(0);
  }
  // mid in two parts: the table entries of FFT p at k (read early), then
  // the factors applied to the transform outputs
  __device__ void mid_coef(int b, int gr, int p, int P, int k, T& ca, T& cb) const {
    const int c0 = gr * 2 * P + 2 * p, c1 = c0 + 1;
    const bool on = active(Lc, b, lev);
    ca = (on && c0 < g.mm1) ? md.coef(g, b, c0, k - 1) : T(0);
    cb = (on && c1 < g.mm1) ? md.coef(g, b, c1, k - 1) : T(0);
  }
  __device__ void mid_apply(int b, int gr, int p, int P, int k, T ya, T yb, T ca, T cb, T& va, T& vb) const {
    const int c0 = gr * 2 * P + 2 * p, c1 = c0 + 1;
    const bool on = active(Lc, b, lev);
    va = (on && c0 < g.mm1) ? md.apply(ca, ya) : T(0);
    vb = (on && c1 < g.mm1) ? md.apply(cb, yb) : T(0);
  }
  __device__ void store(int b, int gr, int p, int P, int k, T ya, T yb) const {
    if (!active(Lc, b, lev)) return;
    const int c0 = gr * 2 * P + 2 * p, c1 = c0 + 1;
    if (c0 < g.mm1) dst(g, b, c0, k - 1, ya);
    if (c1 < g.mm1) dst(g, b, c1, k - 1, yb);
  }
};

// Step 1 at level slot lev (src/operators.f90:103-134) in increment form.
// The reference forms omega* = S[lam1i * (S X + lam1 * S omega)] with
// X = con1*rhs + con2*rhs_old + rhsbc and rhs = rot(nonlinear(omega_k, q_k,
// q0_k, lastbc_k)).  S S = 4mn I and (with a = dt/(2 Re del^2))
// lam1 - 1/(4mn lam1i) = -2 a lam/(4mn), and S(-D0 w) = lam * S w for the
// 5-point Laplacian D0 with zero edge values, so the same omega* is
//   omega* = omega + S[lam1i * S(X + 2 vfac D0 omega)],   vfac = a/(4mn).
// The transforms carry only the O(dt) increment: their rounding scales with
// the increment, not with omega (with FP32 transforms the direct form puts
// an error of order eps*|omega| into the slip that the force system then
// removes, a 1e-4-relative bias of the mean forces at 1024^2), and one field
// is transformed instead of two.
// SrcStep1 gives X + 2 vfac D0 omega at (ii, jj) and stores rhs_old <- rhs.
template <typename T, bool kFree = false>
struct SrcStep1 {
  const T* om;      // omega at level lev (old)        stride fbs
  const T* omc;     // omega at level lev+1 (new)      stride fbs; nullptr if lev+1 == Lmax
  const T* q; const T* q0;  //                         stride qbs; not read when kFree
  const T* lastbc;  // get_bc of old omega_{lev+1}, x0.25  stride bbs
  T* rhs_old;       //                                 stride fbs
  const T* con1; const T* con2; const T* vfac;  // [b*Lmax + lev]
  const int* Lc;
  long fbs, qbs, bbs;
  int lev, first;   // first: itime == 1, explicit Euler start
  // Tangent-linear step: base state at level lev (same strides).  When bom
  // is set, (om, q, lastbc) are perturbations, q0 is zero, and rhs is the
  // linearised term rot(nl(d omega, q+q0, d bc)) + rot(nl(omega, d q, bc)).
  const T* bom = nullptr;
  const T* bq = nullptr;
  const T* bq0 = nullptr;
  const T* blb = nullptr;
  const T* q0c = nullptr;  // [b*Lmax + lev]: the u-face value of the freestream q0 (kFree)
  static constexpr bool kGather = kFree;  // the freestream form loads in two phases
  // lastbc of every level as side work (the job of k_lastbc), for the launch
  // of the coarsest level slot, which neither reads those levels' lastbc nor
  // writes any omega: lastbc_all[(b Lmax + k) nbc + p] = get_bc(omega_all of
  // level k+1, 0.25) for the levels k with a coarser level; the others are
  // not written (they hold zeros).  side_count = Lmax * nbc, 0 for none;
  // side_groups_for(nt) workgroups (a multiple of 8) do the side work.
  T* lastbc_all = nullptr;
  const T* om_all = nullptr;
  int side_count = 0;
  // With has_diag, three more side workgroups write the pending diagnostics
  // row of the previous step (diag_row parts 0-2; nt must be a multiple of
  // 64), which reads q, fb, the circulation partials and the motion: no
  // launch of the step before k_tail writes them.
  DiagArgs<T> dg{};
  int has_diag = 0;
  // Level-batched form (lev_items = levels per case > 0): the items are the
  // flattened (case, level) pairs of one field each, and the edge values of
  // the freestream form come from the coarser level's old vorticity (the
  // next item) through get_bc(., 0.25), as lastbc would hold them; zero on a
  // case's coarsest level.  The diagnostics row is then side work of the
  // items of level 0, written for case b / lev_items with the geometry gdiag
  // of all levels.
  int lev_items = 0;
  Geo gdiag{};
  __host__ __device__ int side_groups_for(int nt) const {
    const int n = (side_count + nt - 1) / nt + (has_diag ? 3 : 0);
    return (n + 7) & ~7;
  }
  __device__ void side(const Geo& g, int b, int x, int t) const {
    if (lev_items > 0) {
      if (has_diag && x < 3 && b % lev_items == 0) {
        __shared__ double red[16];
        diag_row<T>(gdiag, dg, b / lev_items, red, x);
      }
      return;
    }
    const int nl = (side_count + (int)blockDim.x - 1) / (int)blockDim.x;
    if (x >= nl) {  // the row's circulation, slip and forces, one workgroup each
      if (has_diag && x - nl < 3) {
        __shared__ double red[16];
        diag_row<T>(g, dg, b, red, x - nl);
      }
      return;
    }
    const int u = x * (int)blockDim.x + t;
    if (u >= side_count) return;
    const int k = u / g.nbc, p = u - k * g.nbc;
    if (!has_coarser(g, Lc, b, k)) return;
    const long fb = (long)b * g.Lmax;
    lastbc_all[(fb + k) * g.nbc + p] = getbc_at<T>(g, om_all + (fb + k + 1) * g.nw, p + 1, T(0.25));
  }
  __device__ T operator()(const Geo& g, int b, int ii, int jj) const {
    T rhs;
    const T v = value(g, b, ii, jj, rhs);
    post(g, b, ii, jj, rhs);
    return v;
  }
  // value without the rhs_old store (rhs is returned for post): the
  // size-specialised passes evaluate every element of a thread before any
  // store, so that all the loads are in flight together
  __device__ T value(const Geo& g, int b, int ii, int jj, T& rhs) const {
    const T* w = om + b * fbs;
    const int e = ii + jj * g.mm1;
    const int i = ii + 2, j = jj + 2;
    if constexpr (kFree) return value_free(g, b, ii, jj, rhs);
    if constexpr (!kFree) {
      rhs = rot_nonlinear<T>(g, i, j, w, bom ? bq + b * qbs : q + b * qbs, bom ? bq0 + b * qbs : q0 + b * qbs,
                             lastbc + b * bbs);
      if (bom) rhs = rhs + rot_nonlinear<T>(g, i, j, bom + b * fbs, q + b * qbs, q0 + b * qbs, blb + b * bbs);
    }
    // every load below is unconditional (in-range places, terms selected),
    // so a thread's elements issue their loads together
    const T ro = rhs_old[b * fbs + e];
    const T rold = first ? rhs : ro;
    const int c = b * g.Lmax + lev;
    // edge terms (zero away from the edge rows and columns)
    T bcs = applybc_vec_sel<T>(g, T(0), i, j, lastbc + b * bbs, vfac[c]);
    const bool hc = omc != nullptr && has_coarser(g, Lc, b, lev);
    const T bcc = applybc_add_sel<T>(g, bcs, i, j, omc != nullptr ? omc + b * fbs : w, T(0.25), vfac[c]);
    bcs = hc ? bcc : bcs;
    const T wl = w[ii > 0 ? e - 1 : e], wr = w[ii + 1 < g.mm1 ? e + 1 : e];
    const T wd = w[jj > 0 ? e - g.mm1 : e], wu = w[jj + 1 < g.nm1 ? e + g.mm1 : e];
    const T lap = (ii > 0 ? wl : T(0)) + (ii + 1 < g.mm1 ? wr : T(0)) + (jj > 0 ? wd : T(0)) +
                  (jj + 1 < g.nm1 ? wu : T(0)) - T(4) * w[e];
    return con1[c] * rhs + con2[c] * rold + bcs + T(2) * vfac[c] * lap;
  }
  // The freestream form in two phases.  gather loads every value of the
  // element from an in-range place (the vorticity and its four neighbours,
  // their edge replacements, 12 fluxes, rhs_old, the edge terms) and
  // combine selects the edge values and does the arithmetic of
  // rot_nonlinear<T, true> and of the general path above, term for term.
  // The size-specialised passes gather all elements of a thread before
  // combining any, so their loads are in flight together.
  struct Gathered {
    T C, wl, wr, wd, wu, bU, bD, bR, bL;
    T v_a, v_b, v_c, v_d, v_e, v_f, u_a, u_b, u_c, u_d, u_e, u_f;
    T ro, lj, li, gj1, gj2, gi1, gi2;
  };
  __device__ void gather(const Geo& g, int b, int ii, int jj, Gathered& G) const {
    const T* w = om + b * fbs;
    const T* qq = q + b * qbs;
    const T* bc = lastbc + b * bbs;
    const bool below = lev_items > 0 && b % lev_items != lev_items - 1;  // batched, with a coarser level
    const T* rc = omc != nullptr ? omc + b * fbs : (below ? w + g.nw : w);
    const int e = ii + jj * g.mm1;
    const int i = ii + 2, j = jj + 2;
    G.C = w[e];
    G.wl = w[ii > 0 ? e - 1 : e];
    G.wr = w[ii + 1 < g.mm1 ? e + 1 : e];
    G.wd = w[jj > 0 ? e - g.mm1 : e];
    G.wu = w[jj + 1 < g.nm1 ? e + g.mm1 : e];
    G.v_a = qq[vix(g, i, j + 1)];
    G.v_b = qq[vix(g, i - 1, j + 1)];
    G.v_c = qq[vix(g, i, j)];
    G.v_d = qq[vix(g, i - 1, j)];
    G.v_e = qq[vix(g, i, j - 1)];
    G.v_f = qq[vix(g, i - 1, j - 1)];
    G.u_a = qq[uix(g, i + 1, j)];
    G.u_b = qq[uix(g, i + 1, j - 1)];
    G.u_c = qq[uix(g, i, j)];
    G.u_d = qq[uix(g, i, j - 1)];
    G.u_e = qq[uix(g, i - 1, j)];
    G.u_f = qq[uix(g, i - 1, j - 1)];
    G.ro = rhs_old[b * fbs + e];
    // the edge values and the applybc_vec_sel / applybc_add_sel operands
    // (get_bc places of getbc_sel), loaded for edge vertices only: a branch
    // around loads alone waits for nothing; the others are never selected
    G.bU = G.bD = G.bR = G.bL = G.lj = G.li = G.gj1 = G.gj2 = G.gi1 = G.gi2 = T(0);
    if (i == 2 || i == g.m || j == 2 || j == g.n) {
      if (lev_items == 0) {
        G.bU = bc[g.top + i - 1];
        G.bD = bc[g.bottom + i - 1];
        G.bR = bc[g.right + j - 1];
        G.bL = bc[g.left + j - 1];
        G.lj = bc[(j == g.n ? g.top : g.bottom) + i - 1];
        G.li = bc[(i == g.m ? g.right : g.left) + j - 1];
      }
      int a1, a2;
      getbc_places(g, (j == g.n ? g.top : g.bottom) + i, a1, a2);
      G.gj1 = rc[a1];
      G.gj2 = rc[a2];
      getbc_places(g, (i == g.m ? g.right : g.left) + j, a1, a2);
      G.gi1 = rc[a1];
      G.gi2 = rc[a2];
    }
  }
  __device__ T combine(const Geo& g, int b, int ii, int jj, const Gathered& G, T& rhs) const {
    const int i = ii + 2, j = jj + 2, m = g.m, n = g.n;
    const int c = b * g.Lmax + lev;
    const T u0 = q0c[c];
    const T fac = vfac[c];
    const T gj = getbc_value(g, (j == n ? g.top : g.bottom) + i, G.gj1, G.gj2, T(0.25));
    const T gi = getbc_value(g, (i == m ? g.right : g.left) + j, G.gi1, G.gi2, T(0.25));
    // the edge values: lastbc, or in the batched form get_bc of the coarser
    // level read here (gj, gi), on the line (gj) and column (gi) of the
    // vertex; only the one of the vertex's own edge is used
    T bU = G.bU, bD = G.bD, bR = G.bR, bL = G.bL, lj = G.lj, li = G.li;
    if (lev_items > 0) {
      const bool below = b % lev_items != lev_items - 1;
      bU = bD = lj = below ? gj : T(0);
      bR = bL = li = below ? gi : T(0);
    }
    // OmU(i, j+1), OmU(i, j-1), OmV(i+1, j), OmV(i-1, j) of rot_nonlinear
    const T nU = j == n ? bU : G.wu, nD = j == 2 ? bD : G.wd;
    const T nR = i == m ? bR : G.wr, nL = i == 2 ? bL : G.wl;
    const T va_up = T(0.5) * (G.v_a + G.v_b + T(0) + T(0));  // vavg(i, j+1)
    const T va_c = T(0.5) * (G.v_c + G.v_d + T(0) + T(0));   // vavg(i, j)
    const T va_dn = T(0.5) * (G.v_e + G.v_f + T(0) + T(0));  // vavg(i, j-1)
    const T ua_r = T(0.5) * (G.u_a + G.u_b + u0 + u0);       // uavg(i+1, j)
    const T ua_c = T(0.5) * (G.u_c + G.u_d + u0 + u0);       // uavg(i, j)
    const T ua_l = T(0.5) * (G.u_e + G.u_f + u0 + u0);       // uavg(i-1, j)
    const T fqu_c = T(0.5) * (va_up * nU + va_c * G.C);      // fqu(i, j)
    const T fqu_d = T(0.5) * (va_c * G.C + va_dn * nD);      // fqu(i, j-1)
    const T fqv_c = T(-0.5) * (ua_r * nR + ua_c * G.C);      // fqv(i, j)
    const T fqv_l = T(-0.5) * (ua_c * G.C + ua_l * nL);      // fqv(i-1, j)
    rhs = fqv_c - fqv_l - fqu_c + fqu_d;
    const T rold = first ? rhs : G.ro;
    // applybc_vec_sel, then applybc_add_sel when a coarser level exists
    const bool ej = j == 2 || j == n, ei = i == 2 || i == m;
    T bcs = T(0);
    bcs = ej ? bcs + fac * lj : bcs;
    bcs = ei ? bcs + fac * li : bcs;
    T bcc = bcs;
    bcc = ej ? bcc + fac * gj : bcc;
    bcc = ei ? bcc + fac * gi : bcc;
    const bool hc = omc != nullptr && has_coarser(g, Lc, b, lev);
    bcs = hc ? bcc : bcs;
    const T lap = (ii > 0 ? G.wl : T(0)) + (ii + 1 < g.mm1 ? G.wr : T(0)) + (jj > 0 ? G.wd : T(0)) +
                  (jj + 1 < g.nm1 ? G.wu : T(0)) - T(4) * G.C;
    return con1[c] * rhs + con2[c] * rold + bcs + T(2) * fac * lap;
  }
  __device__ T value_free(const Geo& g, int b, int ii, int jj, T& rhs) const {
    Gathered G;
    gather(g, b, ii, jj, G);
    return combine(g, b, ii, jj, G, rhs);
  }
  __device__ void post(const Geo& g, int b, int ii, int jj, T rhs) const { rhs_old[b * fbs + ii + jj * g.mm1] = rhs; }
};
// tab * v with a per-case table (step 1's lam1i of level lev), pair-major
template <typename T>
struct MidTab {
  const T* tab; long tbs;
  __device__ T operator()(const Geo& g, int b, int ii, int jj, T v) const { return tab[b * tbs + pm_ix(g, ii, jj)] * v; }
  __device__ T coef(const Geo& g, int b, int ii, int jj) const { return tab[b * tbs + pm_ix(g, ii, jj)]; }
  __device__ T apply(T c, T v) const { return c * v; }
};
// Level-batched step 1, the correction for the new coarser edge values.
// Level k's right-hand side gets vfac_k apply_bc(get_bc(omega_{k+1}, 0.25))
// (src/operators.f90:103-107,122-123) on its first and last interior rows
// and columns; through S lam1i S that is the field
//   d = beta (I + beta C^T C)^-1 e,  beta = dt / (2 Re Delta_k^2)
// (4mn vfac_k = beta) with e the edge source.  C^T C separates in sines:
// for the bottom edge with values e_t (t = 0..m-2 along x),
//   d(ii, jj) = sum_k c_k rho_k^(jj+1) sin(pi k (ii+1) / m),
//   c_k = (2/m) sum_t e_t sin(pi k (t+1) / m),
//   rho_k + 1/rho_k = (1 + beta (lx_k + 2)) / beta, lx_k = 2 - 2 cos(pi k / m)
// (exact but for the reflection at the far edge, rho^(2n-2)), and likewise
// for the other three edges (top: jj -> n-2-jj; left/right: the roles of x
// and y swapped).  rho_k < rho_1 < 1e-17^(1/w): only the w rows and columns
// next to each edge are written (the host picks w per level).
//
// BandOp runs that as a fused column pair (dst_col2_n): for band depth d
// the first DST gives the spectra 2 sum_t e_t sin(pi k (t+1) / N) = N c_k of
// two opposite edge vectors (packed as one complex sequence), the middle
// step scales them by rho_k^(d+1) / (2N), and the second DST gives the
// correction on the two band lines at depth d, added into omega.  Sequence
// q of a launch: d = q / eper, edge e0 + q % eper (0 bottom, 1 top, 2 left,
// 3 right; one launch per axis, eper = 2: the row and column bands cross in
// the corners, where two launches add in a fixed order).  The edge values come from the coarser
// level's bulk solve; its own correction lies in its bands, w < min(m,n)/4
// cells from its edges, and get_bc reads it on lines min(m,n)/4 cells in,
// so the two never meet within the launch.  Items of a case's coarsest
// level and depths d >= w load 0 and store nothing.
// rho_k of every (case, level) item b for an axis of N cells:
// rho[b (N-1) + k-1], k = 1..N-1, from beta[b] (set-up; BandOp reads the
// table, so every pass kernel that runs it gets the same factors)
__global__ void k_band_rho(const double* __restrict__ beta, int N, int items, double* __restrict__ rho) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)items * (N - 1)) return;
  const int b = (int)(t / (N - 1)), k = (int)(t - (long)b * (N - 1)) + 1;
  const double bt = beta[b];
  const double lx = fma(-2.0, cospi((double)k / (double)N), 2.0);
  const double th = fma(bt, lx + 2.0, 1.0) / (2.0 * bt);
  rho[t] = 1.0 / (th + sqrt(fma(th, th, -1.0)));
}
template <typename T>
struct BandOp {
  Geo gr;  // the level geometry
  T* om; int lm; const double* rho; const int* wlev;  // rho: k_band_rho's table for this axis
  int e0, eper, N;
  __device__ void place(int q, int& d, int& e) const {
    d = q / eper;
    e = e0 + q % eper;
  }
  __device__ bool on(int b, int d) const {
    const int lev = b % lm;
    return lev != lm - 1 && d < wlev[lev];
  }
  __device__ T load(int b, int grp, int s, int P, int idx) const {
    int d, e;
    place(grp * 2 * P + s, d, e);
    if (!on(b, d)) return T(0);
    const int base = e == 0 ? gr.bottom : e == 1 ? gr.top : e == 2 ? gr.left : gr.right;
    return getbc_sel<T>(gr, om + (long)(b + 1) * gr.nw, base + idx + 2, T(0.25));
  }
  __device__ void mid_coef(int b, int grp, int p, int P, int k, T& ca, T& cb) const {
    int d, e;
    place(grp * 2 * P + 2 * p, d, e);
    ca = cb = T(0);
    if (!on(b, d)) return;
    const double r = rho[(long)b * (N - 1) + k - 1];
    double pw = r;
    for (int t = 0; t < d; t++) pw *= r;
    ca = cb = T(pw / (2.0 * N));
  }
  __device__ void mid_apply(int, int, int, int, int, T ya, T yb, T ca, T cb, T& va, T& vb) const {
    va = ya * ca;
    vb = yb * cb;
  }
  __device__ void mid(int b, int grp, int p, int P, int k, T ya, T yb, T& va, T& vb) const {
    T ca, cb;
    mid_coef(b, grp, p, P, k, ca, cb);
    mid_apply(b, grp, p, P, k, ya, yb, ca, cb, va, vb);
  }
  __device__ void add(int b, int e, int d, int pos, T v) const {
    const int ii = e < 2 ? pos : (e == 2 ? d : gr.mm1 - 1 - d);
    const int jj = e < 2 ? (e == 0 ? d : gr.nm1 - 1 - d) : pos;
    T* f = om + (long)b * gr.nw + ii + (long)jj * gr.mm1;
    *f = *f + v;
  }
  __device__ void store(int b, int grp, int p, int P, int k, T ya, T yb) const {
    int d, e;
    place(grp * 2 * P + 2 * p, d, e);
    if (!on(b, d)) return;
    add(b, e, d, k - 1, ya);
    add(b, e + 1, d, k - 1, yb);
  }
  // the add in two parts: the old band values read with the inputs, then
  // old + correction (no other workgroup of the launch writes them)
  static constexpr bool kStorePre = true;
  __device__ T* at(int b, int e, int d, int pos) const {
    const int ii = e < 2 ? pos : (e == 2 ? d : gr.mm1 - 1 - d);
    const int jj = e < 2 ? (e == 0 ? d : gr.nm1 - 1 - d) : pos;
    return om + (long)b * gr.nw + ii + (long)jj * gr.mm1;
  }
  __device__ void store_pre(int b, int grp, int p, int P, int k, T& oa, T& ob) const {
    int d, e;
    place(grp * 2 * P + 2 * p, d, e);
    oa = ob = T(0);
    if (!on(b, d)) return;
    oa = *at(b, e, d, k - 1);
    ob = *at(b, e + 1, d, k - 1);
  }
  __device__ void store_put(int b, int grp, int p, int P, int k, T ya, T yb, T oa, T ob) const {
    int d, e;
    place(grp * 2 * P + 2 * p, d, e);
    if (!on(b, d)) return;
    *at(b, e, d, k - 1) = oa + ya;
    *at(b, e + 1, d, k - 1) = ob + yb;
  }
};

// The same band correction in real space.  With the edge values e_t
// (t = 0..N-2) and their odd extension of period 2N (^e(-1) = ^e(N-1) = 0,
// ^e(-2-t) = ^e(2N-2-t) = -e_t), the correction at depth d, position pos is
//   v = (2/N) sum_{k=1}^{N-1} rho_k^(d+1) sum_t e_t sin(pi k (t+1)/N) sin(pi k (pos+1)/N)
//     = sum over one period of ^e(pos - x) H_d(x),
//   H_d(x) = (1/2N) sum_{k=0}^{2N-1} rho_k^(d+1) cos(pi k x / N)
// (the k = 0 and k = N terms, absent from the sine series, add a constant and
// an alternating part to H_d whose sums against one period of ^e vanish).
// rho_k is analytic in cos(pi k / N), so H_d decays geometrically in |x|;
// the host tabulates H_d(x) for |x| <= R, beyond which the dropped terms sum
// below 1e-17 (the tolerance of the band depth), and v is a (2R+1)-term sum.
// One launch for all four bands: a workgroup takes 64 positions of a band
// and all its depths (16 depth lanes), stages ^e over its positions and
// their reach, and adds v on its band vertices (each sum in two chains, even
// and odd offsets: the sums are latency-bound).
// A corner vertex lies in two bands; the bottom or top workgroup adds both
// terms there, its own first (the order of the spectral form's two
// launches), and the left and right workgroups skip it.  The edge values
// come from the coarser level, whose own bands are w < min(m,n)/4 cells from
// its edges and so never meet the get_bc lines (min(m,n)/4 cells in).
// hk[((b 2 + axis) wmax + d) (R+1) + x], axis 0 for the bottom and top edges
// (N = m), 1 for the left and right ones (N = n); rows d >= w are zero.
// grid (2 ceil((m-1)/64) + 2 ceil((n-1)/64), items), 64 kBandDL threads.
constexpr int kBandRmax = 32;    // the widest kernel the direct form takes
constexpr int kBandWmax = 31;    // the deepest band it takes
constexpr int kBandHmax = 1024;  // w (R+1) of one axis, staged in LDS
constexpr int kBandDL = 16;                                  // depth lanes of a workgroup
constexpr int kBandPer = (kBandWmax + kBandDL - 1) / kBandDL;  // depths per thread
template <typename T>
__device__ __forceinline__ double band_ehat(const Geo& g, const T* __restrict__ omc, int base, int N, int t) {
  const int P = 2 * N;
  int tt = t % P;
  tt += tt < 0 ? P : 0;
  if (tt == N - 1 || tt == P - 1) return 0.0;
  const bool neg = tt >= N;
  const int s = neg ? P - 2 - tt : tt;
  const double v = (double)getbc_sel<T>(g, omc, base + s + 2, T(0.25));
  return neg ? -v : v;
}
template <typename T>
__global__ __launch_bounds__(64 * kBandDL) void k_band_direct(Geo g, T* __restrict__ om, int lm,
                                                              const int* __restrict__ wlev,
                                                     const double* __restrict__ hk, int wmax, int R, int nbt,
                                                     int nlr) {
  __shared__ double es[64 + 2 * kBandRmax];
  __shared__ double hs[kBandHmax];
  // a bottom or top workgroup at a corner: the left and right edges' values
  // next to its band rows, and the left/right kernels
  __shared__ double cs[2][kBandWmax + 2 * kBandRmax];
  __shared__ double hc[kBandHmax];
  const int b = blockIdx.y, lev = b % lm;
  if (lev == lm - 1) return;  // a case's coarsest level has no coarser edge values
  const int w = wlev[lev];
  int x = blockIdx.x, e, chunk;
  if (x < 2 * nbt) {
    e = x / nbt;
    chunk = x - e * nbt;
  } else {
    x -= 2 * nbt;
    e = 2 + x / nlr;
    chunk = x - (e - 2) * nlr;
  }
  const bool horiz = e < 2;
  const int N = horiz ? g.m : g.n, PC = 64, p0 = chunk * PC;
  const int base = e == 0 ? g.bottom : e == 1 ? g.top : e == 2 ? g.left : g.right;
  const T* omc = om + (long)(b + 1) * g.nw;
  const double* H = hk + (long)(b * 2 + (horiz ? 0 : 1)) * wmax * (R + 1);
  // corners this workgroup owns: positions < w (left) or > m-2-w (right)
  const bool cl = horiz && p0 < w, cr = horiz && p0 + PC - 1 > g.mm1 - 1 - w;
  const int jlo = e == 0 ? 0 : g.nm1 - w;  // the band rows of a bottom or top workgroup
  const int nc = w + 2 * R;
  // this thread's band vertices (depths dl, dl + DL, ...) and their old
  // values, loaded with the staging loads below rather than after the sums
  // bottom/top: consecutive lanes take consecutive positions (one band row);
  // left/right: consecutive lanes take consecutive depths (one grid row)
  const int lp = horiz ? threadIdx.x & 63 : threadIdx.x / kBandDL;
  const int dl = horiz ? threadIdx.x >> 6 : threadIdx.x % kBandDL, DL = kBandDL;
  const int pos = p0 + lp;
  // a left or right workgroup leaves the corners to the bottom or top one
  const bool act = pos <= N - 2 && (horiz || (pos >= w && pos <= N - 2 - w));
  T old[kBandPer];
#pragma unroll
  for (int i = 0; i < kBandPer; i++) {
    const int d = dl + DL * i;
    old[i] = T(0);
    if (act && d < w) {
      const int ii = horiz ? pos : (e == 2 ? d : g.mm1 - 1 - d);
      const int jj = horiz ? (e == 0 ? d : g.nm1 - 1 - d) : pos;
      old[i] = om[(long)b * g.nw + ii + (long)jj * g.mm1];
    }
  }
  for (int i = threadIdx.x; i < PC + 2 * R; i += 64 * kBandDL) es[i] = band_ehat<T>(g, omc, base, N, p0 - R + i);
  for (int i = threadIdx.x; i < w * (R + 1); i += 64 * kBandDL) hs[i] = H[i];
  if (cl || cr) {
    const double* Hc = hk + (long)(b * 2 + 1) * wmax * (R + 1);
    for (int i = threadIdx.x; i < w * (R + 1); i += 64 * kBandDL) hc[i] = Hc[i];
    for (int i = threadIdx.x; i < 2 * nc; i += 64 * kBandDL) {
      const int side = i / nc, t = jlo - R + (i - side * nc);
      if (side == 0 ? cl : cr) cs[side][i - side * nc] = band_ehat<T>(g, omc, side == 0 ? g.left : g.right, g.n, t);
    }
  }
  __syncthreads();
  if (!act) return;
#pragma unroll
  for (int i = 0; i < kBandPer; i++) {
    const int d = dl + DL * i;
    if (d >= w) break;
    const double* hd = hs + d * (R + 1);
    const double* ec = es + lp + R;
    double v0 = hd[0] * ec[0], v1 = 0.0;  // even and odd k in two chains
    int k = 1;
#pragma unroll 4
    for (; k + 1 <= R; k += 2) {
      v1 = fma(hd[k], ec[-k] + ec[k], v1);
      v0 = fma(hd[k + 1], ec[-k - 1] + ec[k + 1], v0);
    }
    if (k <= R) v1 = fma(hd[k], ec[-k] + ec[k], v1);
    const double v = v0 + v1;
    const int ii = horiz ? pos : (e == 2 ? d : g.mm1 - 1 - d);
    const int jj = horiz ? (e == 0 ? d : g.nm1 - 1 - d) : pos;
    T out = old[i] + T(v);
    if (horiz && (ii < w || ii > g.mm1 - 1 - w)) {
      // the left or right band's term at this corner vertex: depth ii (or
      // mm1-1-ii) of that band, position jj along its edge
      const bool left = ii < w;
      const double* hl = hc + (left ? ii : g.mm1 - 1 - ii) * (R + 1);
      const double* el = cs[left ? 0 : 1] + (jj - jlo) + R;
      double u0 = hl[0] * el[0], u1 = 0.0;
      int kc = 1;
#pragma unroll 4
      for (; kc + 1 <= R; kc += 2) {
        u1 = fma(hl[kc], el[-kc] + el[kc], u1);
        u0 = fma(hl[kc + 1], el[-kc - 1] + el[kc + 1], u0);
      }
      if (kc <= R) u1 = fma(hl[kc], el[-kc] + el[kc], u1);
      out = out + T(u0 + u1);
    }
    om[(long)b * g.nw + ii + (long)jj * g.mm1] = out;
  }
}

// ---------------------------------------------------------------------------
// vort2flux with the levels batched (src/operators.f90:191-240).  The
// reference solves s_k = ctci(omega_k + apply_bc(get_bc(s_{k+1}, 1), 1))
// coarse to fine.  ctci is the inverse of the 5-point operator C^T C, so
//   s_k = s0_k + H(sbc_k),  s0_k = ctci(omega_k),  H(x) = ctci(apply_bc(0, x, 1)),
// with H(x) the discrete harmonic extension of the edge values x.  get_bc
// reads s_{k+1} only on its ring: the lines j = n/4, 3n/4 and i = m/4, 3m/4
// between those corners, R_{k+1} = ring(s_{k+1}), sbc_k = Interp(R_{k+1})
// (the averages of get_bc).  So
//   R_k = ring(s0_k) + T R_{k+1},   T = ring . H . Interp,
//   R_k = sum_{j >= 0} T^j ring(s0_{k+j})  (T^0 = I, up to the coarsest level),
// one product with each power of T, which depends on m and n only.  Every
// level's s0 is one batched ctci, the ring vectors of all levels are one
// kernel, and H(Interp(R_{k+1})) is added in closed form.  H separates in
// sines: for bottom edge values x_b (and top values x_t at row n-2-jj),
//   H(ii, jj) = sum_k (DST(x_b)_k g_k(jj) + DST(x_t)_k g_k(n-2-jj)) / (2m)
//               * 2 sin(pi k (ii+1) / m),
//   g_k(jj) = (rho^(jj+1) - rho^(2n-1-jj)) / (1 - rho^(2n)),
//   rho + 1/rho = 2 + lx_k,  lx_k = 4 sin^2(pi k / (2m)),
// the 1-D solve across the grid with Dirichlet ends; left and right edges
// swap the axes.  DST(x) = 2 sum_t x_t sin(pi k (t+1) / N).
//
// Ring layout: [bottom line (m/2+1), top line (m/2+1), left (n/2+1),
// right (n/2+1)], each from its first corner; entry c of the bottom line is
// s(m/4-1+c, n/4-1) (0-based interior indices), of the left line
// s(m/4-1, n/4-1+c).  Corners appear on two lines.
__host__ __device__ __forceinline__ int ring_count(int m, int n) { return 2 * (m / 2 + 1) + 2 * (n / 2 + 1); }
__host__ __device__ __forceinline__ void ring_pos(int m, int n, int p, int& ii, int& jj) {
  const int nb = m / 2 + 1, nl = n / 2 + 1;
  if (p < nb) { ii = m / 4 - 1 + p; jj = n / 4 - 1; }
  else if (p < 2 * nb) { ii = m / 4 - 1 + (p - nb); jj = 3 * n / 4 - 1; }
  else if (p < 2 * nb + nl) { ii = m / 4 - 1; jj = n / 4 - 1 + (p - 2 * nb); }
  else { ii = 3 * m / 4 - 1; jj = n / 4 - 1 + (p - 2 * nb - nl); }
}
__host__ __device__ __forceinline__ int ring_edge0(int m, int n, int e) {
  const int nb = m / 2 + 1, nl = n / 2 + 1;
  return e == 0 ? 0 : e == 1 ? nb : e == 2 ? 2 * nb : 2 * nb + nl;
}
// The symmetries of the ring places: g = 4a + 2b + c applies the flip
// x -> m-2-x if c (positions along x reverse, left <-> right), then the flip
// y -> n-2-y if b (bottom <-> top, positions along y reverse), then, on
// square grids, the transpose if a (bottom <-> left, top <-> right, the
// same position).  The ring map T is built from the sine transforms, the
// harmonic extension and get_bc's averages, all symmetric under these, so
// T(g r, g c) = T(r, c) up to rounding.
__host__ __device__ __forceinline__ int ring_sym(int m, int n, int g, int p) {
  const int nb = m / 2 + 1, nl = n / 2 + 1;
  int side, pos;
  if (p < nb) side = 0, pos = p;
  else if (p < 2 * nb) side = 1, pos = p - nb;
  else if (p < 2 * nb + nl) side = 2, pos = p - 2 * nb;
  else side = 3, pos = p - 2 * nb - nl;
  if (g & 1) {
    if (side < 2) pos = nb - 1 - pos;
    else side ^= 1;
  }
  if (g & 2) {
    if (side < 2) side ^= 1;
    else pos = nl - 1 - pos;
  }
  if (g & 4) side ^= 2;
  return side < 2 ? side * nb + pos : 2 * nb + (side - 2) * nl + pos;
}
// Interp(R) at interior place t (0-based, t = 0..N-2) of edge e (0 bottom,
// 1 top, 2 left, 3 right): get_bc's value at 1-based place i = t+1 of that
// edge, taken from the ring vector R (the same two terms in the same order)
__device__ __forceinline__ double ring_interp(const double* __restrict__ R, int m, int n, int e, int t) {
  const int r0 = ring_edge0(m, n, e), i = t + 1;
  return (i & 1) ? 0.5 * (R[r0 + (i + 1) / 2] + R[r0 + (i - 1) / 2]) : R[r0 + i / 2];
}

// tab[t] = sin(pi t / N), t < 2N
__global__ void k_sinpi_table(int N, double* __restrict__ tab) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 2 * N) tab[t] = sinpi((double)t / (double)N);
}
// g_k(jj) / (2 Nal) in G[jj * (Nal-1) + k-1]: modes k = 1..Nal-1 along the
// edge, jj = 0..Nac-2 across.  Grid over (Nal-1)(Nac-1) entries.
__global__ void k_ring_gtab(int Nal, int Nac, double* __restrict__ G) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long K = Nal - 1;
  if (e >= K * (Nac - 1)) return;
  const int jj = (int)(e / K), k = (int)(e - jj * K) + 1;
  const double sn = sinpi((double)k / (2.0 * Nal));
  const double lx = 4.0 * sn * sn;
  const double rho = 0.5 * (2.0 + lx - sqrt(lx * (lx + 4.0)));
  const double g = (pow(rho, (double)(jj + 1)) - pow(rho, (double)(2 * Nac - 1 - jj))) / (1.0 - pow(rho, 2.0 * Nac));
  G[e] = g / (2.0 * Nal);
}
// The two factors of T's block of edge e: B[p][k-1] = G(k, d_p) 2 sin(pi k
// (a_p+1) / Nal) for ring point p (d_p its distance from edge e across the
// grid, a_p its place along it), and A[c][k-1] = DST of Interp(unit c) on
// edge e = 2 sum_i w_i sin(pi k i / Nal) over get_bc's places i = 2c-1, 2c,
// 2c+1 inside 1..Nal-1 (weights 1/2, 1, 1/2).  T[p][r0_e + c] =
// sum_k B[p][k] A[c][k].  Grid over NR (Nal-1) entries; A in the first
// (Nal/2+1)(Nal-1).
__global__ void k_ring_factors(int m, int n, int e, const double* __restrict__ G, double* __restrict__ Bm,
                               double* __restrict__ Am) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const bool xe = e < 2;
  const int Nal = xe ? m : n, Nac = xe ? n : m, K = Nal - 1, NR = ring_count(m, n);
  if (t < (long)NR * K) {
    const int p = (int)(t / K), k = (int)(t - (long)p * K) + 1;
    int ii, jj;
    ring_pos(m, n, p, ii, jj);
    const int along = xe ? ii : jj, across = xe ? jj : ii;
    const int d = (e == 0 || e == 2) ? across : Nac - 2 - across;
    Bm[t] = G[(long)d * K + k - 1] * 2.0 * sinpi((double)k * (along + 1) / (double)Nal);
  }
  if (t < (long)(Nal / 2 + 1) * K) {
    const int c = (int)(t / K), k = (int)(t - (long)c * K) + 1;
    double a = 0.0;
    for (int i = 2 * c - 1; i <= 2 * c + 1; i++)
      if (i >= 1 && i <= Nal - 1) a += (i == 2 * c ? 1.0 : 0.5) * 2.0 * sinpi((double)k * i / (double)Nal);
    Am[t] = a;
  }
}
// C[i][j] = sum_k X[i xi + k xk] Y[k yk + j yj] (+ C when acc), 16 x 16
// tiles through LDS; set-up only (the powers of T).  grid (ceil(nj/16),
// ceil(ni/16)), 256 threads.
__global__ __launch_bounds__(256) void k_gemm64(int ni, int nj, int nk, const double* __restrict__ X, long xi,
                                                long xk, const double* __restrict__ Y, long yk, long yj,
                                                double* __restrict__ C, long ci, int acc) {
  __shared__ double xs[16][17], ys[16][17];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int i = blockIdx.y * 16 + ty, j = blockIdx.x * 16 + tx;
  double sum = 0.0;
  for (int k0 = 0; k0 < nk; k0 += 16) {
    const int kx = k0 + tx, ky = k0 + ty;
    xs[ty][tx] = (i < ni && kx < nk) ? X[i * xi + kx * xk] : 0.0;
    ys[ty][tx] = (ky < nk && j < nj) ? Y[ky * yk + j * yj] : 0.0;
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; kk++) sum += xs[ty][kk] * ys[kk][tx];
    __syncthreads();
  }
  if (i < ni && j < nj) C[i * ci + j] = (acc ? C[i * ci + j] : 0.0) + sum;
}

// Batch index -> (case, level) item: items of levels 0..per-1 of each case
// (per = lm: all of them).
struct ItemMap {
  int per, lm;
  __host__ __device__ int operator()(int b) const { return (b / per) * lm + b % per; }
};

// The ring vectors of s0_k = DST_i(Y_k) for k = 1..L-1 of case b, from Y_k
// (the column pair's output, s0 before its last transform along i):
// V[(b lm + k) NR + p].  s0(ii, jj) = 2 sum_kk Y(kk-1, jj) sin(pi kk (ii+1)/m);
// the left and right lines (ii+1 = m/4, 3m/4) take one wave per row jj of
// the ring, the bottom and top lines (rows n/4-1, 3n/4-1) one wave per 8
// places.  sintab[t] = sin(pi t / m), t < 2m, staged in LDS (2m doubles of
// dynamic LDS).  grid (ceil(tasks/4), B (lm-1)), 256 threads.
constexpr int kRingRowOut = 8;
__host__ __device__ __forceinline__ int ring_fromY_tasks(int m, int n) {
  return (n / 2 + 1) + 2 * ((m / 2 + 1 + kRingRowOut - 1) / kRingRowOut);
}
__device__ __forceinline__ double wave_sum(double v);
// Stages n doubles from src into LDS dst with every thread's loads issued
// before its stores (kStageU per thread per round), so that the staging
// costs one memory round trip per round rather than one per element.
constexpr int kStageU = 16;
template <class F>
__device__ __forceinline__ void stage_lds(double* dst, int n, F&& load) {
  for (int base = 0; base < n; base += 256 * kStageU) {
    double r[kStageU];
#pragma unroll
    for (int u = 0; u < kStageU; u++) {
      const int e = base + threadIdx.x + 256 * u;
      r[u] = e < n ? load(e) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kStageU; u++) {
      const int e = base + threadIdx.x + 256 * u;
      if (e < n) dst[e] = r[u];
    }
  }
}

// With fld set, every ring value is also written at its place of the
// field fld + case * fbs (the corners once, by the bottom and top lines).
__device__ __forceinline__ void ring_fromY_waves(const Geo& g, const double* __restrict__ Y, int lm, int NR,
                                                 const double* __restrict__ sintab, double* __restrict__ V,
                                                 bool lr_only, double* __restrict__ fld = nullptr, long fbs = 0) {
  extern __shared__ double sn[];  // sin(pi t / m), t < 2m
  const int b = blockIdx.y / (lm - 1), k = 1 + blockIdx.y % (lm - 1);
  const double* y = Y + ((long)b * lm + k) * g.nw;
  const int m = g.m, n = g.n, K = g.mm1, m2 = 2 * m;
  const int nb = m / 2 + 1, nl = n / 2 + 1, nrt = (nb + kRingRowOut - 1) / kRingRowOut;
  const int task = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  double* out = V + ((long)b * lm + k) * NR;
  const bool live = task < (lr_only ? nl : nl + 2 * nrt);  // whole waves
  const bool lr = task < nl;
  const int t2 = task - nl, line = lr ? 0 : t2 / nrt, c0 = lr ? 0 : (t2 - line * nrt) * kRingRowOut;
  const int jj = !live ? 0 : lr ? n / 4 - 1 + task : (line == 0 ? n / 4 - 1 : 3 * n / 4 - 1);
  // this lane's first modes kk = 1 + lane + 64 u, loaded before the table
  // is staged
  constexpr int UY = 8;
  double yv[UY];
#pragma unroll
  for (int u = 0; u < UY; u++) {
    const int kk = 1 + lane + 64 * u;
    yv[u] = (live && kk <= K) ? y[(kk - 1) + (long)jj * K] : 0.0;
  }
  stage_lds(sn, m2, [&](int e) { return sintab[e]; });
  __syncthreads();
  if (!live) return;
  // the sine places kk c mod 2m advance by 64 c mod 2m from one u to the
  // next (no division in the loop)
  double a[kRingRowOut];
#pragma unroll
  for (int f = 0; f < kRingRowOut; f++) a[f] = 0.0;
  if (lr) {
    int fL = ((1 + lane) * (m / 4)) % m2, fR = ((1 + lane) * (3 * m / 4)) % m2;
    const int sL = (64 * (m / 4)) % m2, sR = (64 * (3 * m / 4)) % m2;
    for (int k0 = 0; k0 < K; k0 += 64 * UY) {
      if (k0 > 0) {
#pragma unroll
        for (int u = 0; u < UY; u++) {
          const int kk = k0 + 1 + lane + 64 * u;
          yv[u] = kk <= K ? y[(kk - 1) + (long)jj * K] : 0.0;
        }
      }
#pragma unroll
      for (int u = 0; u < UY; u++) {
        a[0] += yv[u] * sn[fL];
        a[1] += yv[u] * sn[fR];
        fL += sL;
        fL -= fL >= m2 ? m2 : 0;
        fR += sR;
        fR -= fR >= m2 ? m2 : 0;
      }
    }
    const double aL = wave_sum(a[0]), aR = wave_sum(a[1]);
    if (lane == 0) {
      out[2 * nb + task] = 2.0 * aL;
      out[2 * nb + nl + task] = 2.0 * aR;
      if (fld != nullptr && task > 0 && task < nl - 1) {
        fld[b * fbs + (m / 4 - 1) + (long)jj * K] = 2.0 * aL;
        fld[b * fbs + (3 * m / 4 - 1) + (long)jj * K] = 2.0 * aR;
      }
    }
  } else {
    int fr[kRingRowOut], st[kRingRowOut];
#pragma unroll
    for (int f = 0; f < kRingRowOut; f++) {
      const int c = m / 4 + min(c0 + f, nb - 1);
      fr[f] = ((1 + lane) * c) % m2;
      st[f] = (64 * c) % m2;
    }
    for (int k0 = 0; k0 < K; k0 += 64 * UY) {
      if (k0 > 0) {
#pragma unroll
        for (int u = 0; u < UY; u++) {
          const int kk = k0 + 1 + lane + 64 * u;
          yv[u] = kk <= K ? y[(kk - 1) + (long)jj * K] : 0.0;
        }
      }
#pragma unroll
      for (int u = 0; u < UY; u++)
#pragma unroll
        for (int f = 0; f < kRingRowOut; f++) {
          a[f] += yv[u] * sn[fr[f]];
          fr[f] += st[f];
          fr[f] -= fr[f] >= m2 ? m2 : 0;
        }
    }
#pragma unroll
    for (int u = 0; u < kRingRowOut; u++) {
      const double t = wave_sum(a[u]);
      if (lane == 0 && c0 + u < nb) {
        out[line * nb + c0 + u] = 2.0 * t;
        if (fld != nullptr) fld[b * fbs + (m / 4 - 1 + c0 + u) + (long)jj * K] = 2.0 * t;
      }
    }
  }
}
__global__ __launch_bounds__(256) void k_ring_fromY(Geo g, const double* __restrict__ Y, int lm, int NR,
                                                    const double* __restrict__ sintab, double* __restrict__ V,
                                                    double* __restrict__ fld, long fbs) {
  ring_fromY_waves(g, Y, lm, NR, sintab, V, false, fld, fbs);
}

// The bottom and top lines of the ring vectors as one DST-I of the two rows
// n/4-1 and 3n/4-1 of Y_k (row 0 and row 1 of a two-row pass): every place
// of the two rows in one FFT, in place of 2 ceil((m/2+1)/8) waves of direct
// sums whose sine lookups conflict in LDS banks.  The item is the batch
// index as in k_ring_fromY; the outputs at ii = m/4-1 .. 3m/4-1 go to V at
// row nb + ii - (m/4-1), nb = m/2+1.
struct SrcRingLines {
  static constexpr bool kNoWave = true;
  const double* Y; int lm; Geo gf;
  __device__ double operator()(const Geo&, int b, int ii, int row) const {
    const int c = b / (lm - 1), k = 1 + b % (lm - 1);
    const int jj = row == 0 ? gf.n / 4 - 1 : 3 * gf.n / 4 - 1;
    return Y[((long)c * lm + k) * gf.nw + ii + (long)jj * gf.mm1];
  }
};
struct DstRingLines {
  double* V; int lm, NR; Geo gf;
  double* fld = nullptr;  // also at the places of fld + case * fbs when set
  long fbs = 0;
  __device__ void operator()(const Geo&, int b, int ii, int row, double v) const {
    const int nb = gf.m / 2 + 1, pos = ii - (gf.m / 4 - 1);
    if (pos < 0 || pos >= nb) return;
    const int c = b / (lm - 1), k = 1 + b % (lm - 1);
    V[((long)c * lm + k) * NR + row * nb + pos] = v;
    if (fld != nullptr) {
      const int jj = row == 0 ? gf.n / 4 - 1 : 3 * gf.n / 4 - 1;
      fld[c * fbs + ii + (long)jj * gf.mm1] = v;
    }
  }
};
// k_ring_fromY with the bottom and top lines from one FFT per item: the
// workgroups x < nlrg take the left and right lines (one wave per ring row,
// as k_ring_fromY), workgroup nlrg the two-row DST-I (P = 1024 / N2 FFTs of
// the size-specialised pass, the second and later ones of zeros).
// grid (nlrg + 1, B (lm-1)), 256 threads, 2m doubles of dynamic LDS.
template <int N2>
__global__ __launch_bounds__(256) void k_ring_fromY_fft(Geo g, const double* __restrict__ Y, int lm, int NR,
                                                        const double* __restrict__ sintab, double* __restrict__ V,
                                                        const cplx<double>* __restrict__ tw, int nlrg,
                                                        double* __restrict__ fld, long fbs) {
  constexpr int P = 1024 / N2;
  static_assert(P * N2 / 4 == 256, "k_ring_fromY_fft: one pass per workgroup of 256 threads");
  if ((int)blockIdx.x >= nlrg) {
    __shared__ cplx<double> z[(spec::kTwoBuf<double, N2, P> ? 2 : 1) * P * N2];
    Geo g2 = g;
    g2.nm1 = 2;  // the pass's two rows
    const RowOp<double, SrcRingLines, DstRingLines> op{g2, SrcRingLines{Y, lm, g},
                                                       DstRingLines{V, lm, NR, g, fld, fbs}, nullptr, 0};
    spec::dst_pass_body<double, N2, P, 0, RowOp<double, SrcRingLines, DstRingLines>>(op, tw, z, 0, 1);
    return;
  }
  ring_fromY_waves(g, Y, lm, NR, sintab, V, true, fld, fbs);
}

// R_k = V_k + sum_{j=1}^{L-1-k} T^j V_{k+j} for k = 1..L-1 of case b
// (L = NP + 2 levels, Tp: T^1..T^NP, NR x NR each) into Rb (layout of V).
// Workgroup: one ring row p, its columns split over the four waves.  Each
// lane loads its U columns of every power and of V_1..V_{L-1} before any
// product (a 1028-entry ring, 512^2, is one chunk: one memory round trip),
// and the four waves' sums are added in a fixed order.  The powers are the
// kernel's whole HBM traffic (NP NR^2 doubles); one row per workgroup keeps
// sixteen waves of loads in flight per CU.  With sring, R_1 is also written
// at the ring places of level 1 of sring (a field in the streamfunction's
// layout) for get_bc(s_1).  With sym set (sym[l] the row of workgroup l in
// XCD-contiguous order, sym[NR + p] the representative r of row p's orbit
// under ring_sym, sym[2 NR + p] a symmetry g with g r = p), row p is
// sum_q T^j(r, q) V(g q): the powers are read on the representatives' rows
// only (an eighth of them on square grids), and the rows of one orbit are
// adjacent in the order, so they run on one XCD and share its L2.
// grid (NR, B), 256 threads.
constexpr int kRingMaxLev = 8;
constexpr int chain_u(int np) { return np <= 3 ? 5 : 3; }
template <int NP>
__global__ __launch_bounds__(256) void k_ring_chain(Geo g, long fbs, int NR, const double* __restrict__ V,
                                                    const double* __restrict__ Tp, double* __restrict__ Rb,
                                                    double* __restrict__ sring, const int* __restrict__ sym) {
  constexpr int L = NP + 2, U = chain_u(NP), NJ = NP > 0 ? NP : 1;
  __shared__ double red[L][4];
  const int b = blockIdx.y;
  int p = blockIdx.x, r = p, gs = 0;
  if (sym != nullptr) {
    p = sym[xcd_contiguous(blockIdx.x, NR)];
    r = sym[NR + p];
    gs = sym[2 * NR + p];
  }
  const double* vb = V + ((long)b * L + 1) * NR;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double acc[L];
#pragma unroll
  for (int k = 0; k < L; k++) acc[k] = 0.0;
  if constexpr (NP > 0) {
    for (int q0 = 0; q0 < NR; q0 += 256 * U) {
      double t[NJ][U], vv[L - 1][U];
#pragma unroll
      for (int j = 0; j < NP; j++)
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int q = q0 + threadIdx.x + 256 * u;
          t[j][u] = q < NR ? Tp[(long)j * NR * NR + (long)r * NR + q] : 0.0;
        }
      int qv[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int q = q0 + threadIdx.x + 256 * u;
        qv[u] = q < NR ? (gs != 0 ? ring_sym(g.m, g.n, gs, q) : q) : -1;
      }
#pragma unroll
      for (int k = 1; k < L; k++)
#pragma unroll
        for (int u = 0; u < U; u++) vv[k - 1][u] = qv[u] >= 0 ? vb[(k - 1) * NR + qv[u]] : 0.0;
#pragma unroll
      for (int j = 0; j < NP; j++)  // T^(j+1) V_{k+j+1}
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
          for (int k = 1; k + j + 1 <= L - 1; k++) acc[k] += t[j][u] * vv[k + j][u];
    }
  }
#pragma unroll
  for (int k = 1; k < L; k++) {
    const double a = wave_sum(acc[k]);
    if (lane == 0) red[k][w] = a;
  }
  __syncthreads();
  if (threadIdx.x < L - 1) {
    const int k = threadIdx.x + 1;
    const double r = vb[(k - 1) * NR + p] + ((red[k][0] + red[k][1]) + (red[k][2] + red[k][3]));
    Rb[((long)b * L + k) * NR + p] = r;
    if (sring != nullptr && k == 1) {
      int ii, jj;
      ring_pos(g.m, g.n, p, ii, jj);
      sring[b * fbs + g.nw + ii + (long)jj * g.mm1] = r;
    }
  }
}

// The edge spectra of sbc_k = Interp(R_{k+1}) for the items of im:
// X[((item 4 + e) Kx) + kk-1] = DST(x_e)_kk = 2 sum_t x_e(t) sin(pi kk (t+1)
// / N_e), e = 0 bottom, 1 top (N = m), 2 left, 3 right (N = n).  Workgroup
// (edge, 32 modes) of one item: 8 partial sums per mode over t = part mod 8,
// the edge vector and the sine table in LDS (N-1 + 2N doubles), summed in
// a fixed order.  sin_m[t] = sin(pi t / m), t < 2m, and sin_n likewise.
// grid (4 ceil(Kx / 32), items), 256 threads.
__global__ __launch_bounds__(256) void k_ring_spectra(Geo g, const double* __restrict__ Rb, int NR, ItemMap im,
                                                      int Kx, const double* __restrict__ sin_m,
                                                      const double* __restrict__ sin_n, double* __restrict__ X) {
  extern __shared__ double sh[];
  __shared__ double red[256];
  const int item = im(blockIdx.y);
  const int nch = (Kx + 31) / 32, e = blockIdx.x / nch, ch = blockIdx.x - e * nch;
  const int N = e < 2 ? g.m : g.n, K = N - 1;
  if (ch * 32 >= K) return;  // the whole workgroup is past this edge's modes
  double* xs = sh;
  double* st = sh + K;
  const double* R = Rb + (long)(item + 1) * NR;
  const double* sn = e < 2 ? sin_m : sin_n;
  stage_lds(st, 2 * N, [&](int t) { return sn[t]; });
  stage_lds(xs, K, [&](int t) { return ring_interp(R, g.m, g.n, e, t); });
  __syncthreads();
  const int kk = ch * 32 + (threadIdx.x & 31) + 1, part = threadIdx.x >> 5;
  double acc = 0.0;
  if (kk <= K) {
    // terms t = part, part + 8, ... in four interleaved chains (t mod 32), so
    // that four LDS reads and FMAs are in flight; the sine place kk (t+1)
    // mod 2N advances by 8 kk mod 2N per term
    const int N2 = 2 * N;
    const int step = (kk * 8) % N2;
    int idx = (kk * (part + 1)) % N2;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int t = part;
    for (; t + 24 < K; t += 32) {
      int i1 = idx + step;
      i1 -= i1 >= N2 ? N2 : 0;
      int i2 = i1 + step;
      i2 -= i2 >= N2 ? N2 : 0;
      int i3 = i2 + step;
      i3 -= i3 >= N2 ? N2 : 0;
      a0 += xs[t] * st[idx];
      a1 += xs[t + 8] * st[i1];
      a2 += xs[t + 16] * st[i2];
      a3 += xs[t + 24] * st[i3];
      idx = i3 + step;
      idx -= idx >= N2 ? N2 : 0;
    }
    for (; t < K; t += 8) {
      a0 += xs[t] * st[idx];
      idx += step;
      idx -= idx >= N2 ? N2 : 0;
    }
    acc = (a0 + a1) + (a2 + a3);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  if (part == 0 && kk <= K) {
    const int l = threadIdx.x;
    const double sum = ((red[l] + red[l + 32]) + (red[l + 64] + red[l + 96])) +
                       ((red[l + 128] + red[l + 160]) + (red[l + 192] + red[l + 224]));
    X[((long)item * 4 + e) * Kx + kk - 1] = 2.0 * sum;
  }
}

// The edge spectra of k_ring_spectra by FFT: the four edge vectors
// x_e = Interp(R_{k+1}) of an item are the four "rows" of a pass (bottom and
// top in one complex FFT, left and right in the next), each transformed in
// one DST-I, for square grids (one FFT length for all four edges).  The
// same values as the direct sums but for rounding.  Item im(b) of batch
// entry b; grid (ceil(4 / (2P)), items), 256 threads, P = 1024 / N2.
struct SrcRingEdges {
  static constexpr bool kNoWave = true;
  const double* Rb; int NR; ItemMap im; Geo gf;
  __device__ double operator()(const Geo&, int b, int t, int e) const {
    return ring_interp(Rb + (long)(im(b) + 1) * NR, gf.m, gf.n, e, t);
  }
};
struct DstRingEdges {
  double* X; ItemMap im; int Kx;
  __device__ void operator()(const Geo&, int b, int kk1, int e, double v) const {
    X[((long)im(b) * 4 + e) * Kx + kk1] = v;
  }
};
template <int N2>
__global__ __launch_bounds__(256) void k_ring_spectra_fft(Geo g, const double* __restrict__ Rb, int NR, ItemMap im,
                                                          int Kx, double* __restrict__ X,
                                                          const cplx<double>* __restrict__ tw) {
  constexpr int P = 1024 / N2;
  static_assert(P * N2 / 4 == 256, "k_ring_spectra_fft: one pass per workgroup of 256 threads");
  __shared__ cplx<double> z[(spec::kTwoBuf<double, N2, P> ? 2 : 1) * P * N2];
  Geo g4 = g;
  g4.nm1 = 4;  // the four edges
  const RowOp<double, SrcRingEdges, DstRingEdges> op{g4, SrcRingEdges{Rb, NR, im, g}, DstRingEdges{X, im, Kx},
                                                     nullptr, 0};
  spec::dst_pass_body<double, N2, P, 0, RowOp<double, SrcRingEdges, DstRingEdges>>(op, tw, z, blockIdx.x,
                                                                                     gridDim.x);
}

// The last transform of the batched ctci with the bottom and top edge terms
// (a RowOp loader, batch index b -> item im(b)): the spectral row jj of s_k
// is Y(kk-1, jj) + (X_b(kk) G(kk, jj) + X_t(kk) G(kk, n-2-jj)), G = g / (2m)
// along i; nothing is added on a case's coarsest level.
struct SrcRingRow {
  static constexpr bool kNoWave = true;  // five loads per element
  const double* Y; const double* X; const double* G; ItemMap im; int Kx;
  __device__ double operator()(const Geo& g, int b, int ii, int jj) const {
    const int item = im(b);
    const double y = Y[(long)item * g.nw + ii + (long)jj * g.mm1];
    if (item % im.lm == im.lm - 1) return y;
    const double* xb = X + (long)item * 4 * Kx;
    const long K = g.mm1;
    // explicit fma: the same rounding in every pass kernel that inlines it
    return y + fma(xb[ii], G[jj * K + ii], xb[Kx + ii] * G[(g.nm1 - 1 - jj) * K + ii]);
  }
};
// The left and right edge terms as a column pass (ColOp loader): column ii
// of H along j has the spectrum X_l(l) G'(l, ii) + X_r(l) G'(l, m-2-ii),
// G' = g / (2n) along j.
struct SrcRingCol {
  static constexpr bool kNoWave = true;  // four loads per element
  const double* X; const double* G; ItemMap im; int Kx;
  __device__ double operator()(const Geo& g, int b, int ii, int l1) const {
    const double* xl = X + ((long)im(b) * 4 + 2) * Kx;
    const long K = g.nm1;
    return fma(xl[l1], G[ii * K + l1], xl[Kx + l1] * G[(g.mm1 - 1 - ii) * K + l1]);
  }
};
// a = v or a += v on item im(b)
struct DstRingItem {
  double* a; ItemMap im; int add;
  __device__ void operator()(const Geo& g, int b, int ii, int jj, double v) const {
    double* f = a + (long)im(b) * g.nw + ii + (long)jj * g.mm1;
    *f = add ? *f + v : v;
  }
};

// A window of a pass: the rows j0 .. j0 + rows - 1 of a row pass (its Geo
// has nm1 = rows) or the columns i0 .. i0 + cols - 1 of a column pass (mm1 =
// cols), with the loader and the store of the full field (geometry gf) at
// the full field's places.  A column window keeps the rows j0 .. j1 of its
// output.  Each row or column is transformed as in the full pass, so the
// values in the window are the full pass's.
template <class S>
struct SrcRowWin {
  static constexpr bool kNoWave = true;
  S s; Geo gf; int j0;
  __device__ double operator()(const Geo&, int b, int ii, int jj) const { return s(gf, b, ii, jj + j0); }
};
template <class D>
struct DstRowWin {
  D d; Geo gf; int j0;
  __device__ void operator()(const Geo&, int b, int ii, int jj, double v) const { d(gf, b, ii, jj + j0, v); }
};
template <class S>
struct SrcColWin {
  static constexpr bool kNoWave = true;
  S s; Geo gf; int i0;
  __device__ double operator()(const Geo&, int b, int ii, int l1) const { return s(gf, b, ii + i0, l1); }
};
template <class D>
struct DstColWin {
  D d; Geo gf; int i0, j0, j1;
  __device__ void operator()(const Geo&, int b, int ii, int jj, double v) const {
    if (jj >= j0 && jj <= j1) d(gf, b, ii + i0, jj, v);
  }
};

// a = a + scale * v (the step-1 increment; adjoint accumulations)
template <typename T>
struct DstAdd {
  T* a; long bs; T scale;
  __device__ void operator()(const Geo& g, int b, int ii, int jj, T v) const {
    const long e = b * bs + ii + jj * g.mm1;
    a[e] = a[e] + scale * v;
  }
  __device__ T pre(const Geo& g, int b, int ii, int jj) const { return a[b * bs + ii + jj * g.mm1]; }
  __device__ void put(const Geo& g, int b, int ii, int jj, T v, T old) const {
    a[b * bs + ii + jj * g.mm1] = old + scale * v;
  }
};

// The value a row-pass loader Src would hand the transform, evaluated at
// every vertex of item blockIdx.y and written to out (same layout), with the
// loader's side store (post) and its side work (side, in the first
// side_groups workgroups) done as the pass would do them.  The level-batched
// step 1 runs its source this way: one vertex per thread keeps every load of
// the stencil in flight with few registers, and the row pass that follows
// reads one value per element.  The element blocks run in XCD-contiguous
// order (gridDim.x - side_groups is a multiple of 8, surplus blocks exit), so
// the rows a block's stencil shares with its neighbours stay in one L2.
// Batch entries without level lev (per-case level counts Lc, null: all)
// get zeros.  grid (side_groups + ceil(nw / 256) rounded up to a multiple of
// 8, items), 256 threads; side_groups a multiple of 8.
template <typename T, class Src>
__global__ __launch_bounds__(256) void k_src_eval(Geo g, Src src, T* __restrict__ out, int side_groups,
                                                  const int* __restrict__ Lc, int lev) {
  const int b = blockIdx.y;
  int x = blockIdx.x;
  if constexpr (src_has_side<Src>::value) {
    if (x < side_groups) {
      src.side(g, b, x, threadIdx.x);
      return;
    }
    x -= side_groups;
  }
  x = xcd_contiguous(x, (int)gridDim.x - side_groups);
  const int e = x * 256 + threadIdx.x;
  if (e >= g.nw) return;
  const int jj = e / g.mm1, ii = e - jj * g.mm1;
  // a case without level lev (Lc) gets zeros and no side store, as the row
  // pass's loader gives it; the value is evaluated either way (every place
  // it reads exists)
  const bool act = active(Lc, b, lev);
  T aux = T(0), v;
  if constexpr (src_gathered<Src>::value) {
    typename src_gathered<Src>::type G;
    src.gather(g, b, ii, jj, G);
    v = src.combine(g, b, ii, jj, G, aux);
  } else if constexpr (src_has_post<Src>::value) {
    v = src.value(g, b, ii, jj, aux);
  } else {
    v = src(g, b, ii, jj);
  }
  if constexpr (src_has_post<Src>::value)
    if (act) src.post(g, b, ii, jj, aux);
  out[(long)b * g.nw + e] = act ? v : T(0);
}

// ---------------------------------------------------------------------------
// small kernels (grid.y = case unless stated)
// ---------------------------------------------------------------------------
// lastbc(:,k) = get_bc(omega(:,:,k+1), 0.25) below a case's coarsest level,
// 0 on it (src/operators.f90:97-100).  grid.x over Lmax*nbc
template <typename T>
__global__ void k_lastbc(Geo g, const T* __restrict__ om, T* __restrict__ lastbc, const int* __restrict__ Lc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.nbc * g.Lmax) return;
  const int b = blockIdx.y;
  const int lev = t / g.nbc, p = t - lev * g.nbc;
  const long fb = (long)b * g.Lmax;
  T v = T(0);
  if (has_coarser(g, Lc, b, lev)) v = getbc_at<T>(g, om + (fb + lev + 1) * g.nw, p + 1, T(0.25));
  lastbc[(fb + lev) * g.nbc + p] = v;
}

// full-weighting restriction of the fine values F around fine vertex
// (ii, jj), in the expression order of src/operators.f90:325-329 (the 0.5
// and 0.25 products are exact, so FMA contraction cannot change the result)
template <typename T, class F>
__device__ __forceinline__ T restrict_at(const F& f, int ii, int jj) {
  return f(ii, jj) + T(0.5) * (f(ii - 1, jj) + f(ii + 1, jj) + f(ii, jj - 1) + f(ii, jj + 1)) +
         T(0.25) * (f(ii - 1, jj - 1) + f(ii - 1, jj + 1) + f(ii + 1, jj - 1) + f(ii + 1, jj + 1));
}

// coarsify (src/operators.f90:307-333) of level lev from level lev-1, in
// place.  grid.x over (m/2-1)(n/2-1)
template <typename T>
__global__ void k_coarsify(Geo g, T* __restrict__ coarse, const T* __restrict__ fine, long bs,
                           const int* __restrict__ Lc, int lev) {
  const int wi = g.m / 2 - 1, wj = g.n / 2 - 1;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (t >= wi * wj || !active(Lc, b, lev)) return;
  const int i = g.m / 4 + 2 + t % wi, j = g.n / 4 + 2 + t / wi;
  const int ii = 2 * i - g.m / 2 - 1, jj = 2 * j - g.n / 2 - 1;
  const T* r = fine + b * bs;
  auto F = [&](int a, int c) -> T { return r[wix(g, a, c)]; };
  coarse[b * bs + wix(g, i, j)] = restrict_at<T>(F, ii, jj);
}

// The same chain two levels per launch: workgroups x < nblk coarsify level
// lev from lev-1 as k_coarsify; the next nblk coarsify level lev+1 from
// level lev, where every value of level lev inside its own centre region is
// recomputed from level lev-1 (restrict_at, the same expression, so the
// result is bit-identical to two k_coarsify launches) rather than read, as
// those entries are being overwritten by the first half of the grid.
// grid.x = 2*nblk (nblk = ceil((m/2-1)(n/2-1)/256)) or nblk when
// lev+1 == Lmax.
// STEP 3 near the bodies (src/operators.f90:174): omega_1 -= ainv(rr) in
// real space.  ainv = S diag(lam1i) S / (4mn) inverts the implicit viscous
// operator 1 + beta L (L the 5-point stencil, beta = dt / (2 Re Delta^2));
// written over one period of the sine form's odd extension its kernel is
// translation invariant plus images in the domain's edges,
//   K(a, b) = (1/2n) sum_{ky=0}^{2n-1} cos(pi ky b / n) r_ky^|a| / sqrt(A_ky^2 - 4 beta^2),
//   A_ky = 1 + beta (4 - 2 cos(pi ky / n)),  r_ky = (A_ky - sqrt(A_ky^2 - 4 beta^2)) / (2 beta)
// (the sum over kx in closed form: the lattice Green's function of a
// three-point operator, whose periodic images at distance 2m are below
// 1e-300), and it decays geometrically in |a| and |b|.  rr is zero but on
// the active vertices, so omega changes only within R cells of them: the
// host tabulates K for 0 <= a, b <= R (the dropped tail below 1e-17 of a unit
// source) and takes a case this way only when its active vertices are at
// least R cells from every edge (then every image term is past R).
// Workgroup: a kAinvTX x kAinvTY block of the case's target box; it stages
// rr over the block and R around it (zero outside the field), and K.  Each
// thread owns four neighbouring places of one block row; the sixteen groups
// of 32 threads take the window rows dy = grp mod 16 (two rows each at
// R = 14: short serial chains for the latency-bound single case), and for
// each row a thread
// streams the 2R + 4 staged values its four places read once from LDS,
// feeding four chains of FMAs (each place's taps in ascending dx), so a
// staged value serves up to four places.  The groups' partial sums are
// added in a fixed tree.  R is the largest reach of the cases taken; a
// case's table is zero past its own reach, and the row groups and tap order
// do not depend on R, so a case in a batch gives its single-case result bit
// for bit.
// om: level 1 of case b at b bs; rr (one level per case) at b rbs.
// box[b] = {i0, j0, nx, ny} (target box), K per case ((R+1)^2, row b = |dy|),
// loc[b] != 0 for the cases taken.  grid (max tiles, B), 512 threads.
constexpr int kAinvRmax = 20;
constexpr int kAinvTX = 16, kAinvTY = 8, kAinvPX = 4, kAinvGroups = 16, kAinvThreads = 32 * kAinvGroups;
constexpr int kAinvWX = kAinvTX + 2 * kAinvRmax, kAinvWY = kAinvTY + 2 * kAinvRmax;
template <typename T>
__global__ __launch_bounds__(kAinvThreads) void k_ainv_local(Geo g, T* __restrict__ om, long bs,
                                                             const T* __restrict__ rr, long rbs,
                                                             const int* __restrict__ box, const double* __restrict__ K,
                                                             const int* __restrict__ loc, int R) {
  constexpr int NP = kAinvTX * kAinvTY;  // places per block
  __shared__ double rs[kAinvWX * kAinvWY];
  __shared__ double ks[(kAinvRmax + 1) * (kAinvRmax + 1)];
  __shared__ double red[kAinvGroups][NP];
  const int b = blockIdx.y;
  if (!loc[b]) return;
  const int bi0 = box[4 * b], bj0 = box[4 * b + 1], nx = box[4 * b + 2], ny = box[4 * b + 3];
  const int ntx = (nx + kAinvTX - 1) / kAinvTX, tx = blockIdx.x % ntx, ty = blockIdx.x / ntx;
  if (ty * kAinvTY >= ny) return;
  const int x0 = bi0 + tx * kAinvTX, y0 = bj0 + ty * kAinvTY, W = kAinvTX + 2 * R, H = kAinvTY + 2 * R, R1 = R + 1;
  const int tid = threadIdx.x, grp = tid >> 5, lt = tid & 31, ly = lt >> 2, lq = lt & 3;
  // the place this thread writes at the end (threads < NP): block place tid
  const int wx = tid % kAinvTX, wy = tid / kAinvTX;
  const bool mine = tid < NP && tx * kAinvTX + wx < nx && ty * kAinvTY + wy < ny;
  T* o = om + b * bs + (x0 + wx) + (long)(y0 + wy) * g.mm1;
  const T old = mine ? *o : T(0);
  const T* rb = rr + b * rbs;
  // every staging load of the thread is issued before the first store
  constexpr int NT = kAinvThreads;
  constexpr int U = (kAinvWX * kAinvWY + NT - 1) / NT, UK = ((kAinvRmax + 1) * (kAinvRmax + 1) + NT - 1) / NT;
  double v[U], kv[UK];
#pragma unroll
  for (int u = 0; u < U; u++) {
    const int e = tid + NT * u, y = e / W, x = e - y * W, ii = x0 - R + x, jj = y0 - R + y;
    v[u] = (e < W * H && ii >= 0 && ii < g.mm1 && jj >= 0 && jj < g.nm1) ? (double)rb[ii + (long)jj * g.mm1] : 0.0;
  }
#pragma unroll
  for (int u = 0; u < UK; u++) {
    const int e = tid + NT * u;
    kv[u] = e < R1 * R1 ? K[(long)b * R1 * R1 + e] : 0.0;
  }
#pragma unroll
  for (int u = 0; u < U; u++) {
    const int e = tid + NT * u;
    if (e < W * H) rs[e] = v[u];
  }
#pragma unroll
  for (int u = 0; u < UK; u++) {
    const int e = tid + NT * u;
    if (e < R1 * R1) ks[e] = kv[u];
  }
  __syncthreads();
  // group grp takes the rows dy = grp mod 16 in ascending order.  For place
  // p of the thread (staged centre column c = 4 lq + p + R) and row dy the
  // sum is two chains: a0 = K(0) rr(c) + the even dx, a1 = the odd dx, each
  // term K(dx) (rr(c - dx) + rr(c + dx)), dx ascending; then acc += a0 + a1.
  // The left and right values of the four places are two windows that slide
  // apart by one column per dx (one new LDS value each).
  double acc[kAinvPX];
#pragma unroll
  for (int p = 0; p < kAinvPX; p++) acc[p] = 0.0;
  for (int dy = -R + ((grp + R) & (kAinvGroups - 1)); dy <= R; dy += kAinvGroups) {
    const double* kr = ks + (dy < 0 ? -dy : dy) * R1;
    const double* rrow = rs + (ly + R - dy) * W + kAinvPX * lq + R;  // centre of place p = 0
    double a0[kAinvPX], a1[kAinvPX], lw[kAinvPX], rw[kAinvPX];
#pragma unroll
    for (int p = 0; p < kAinvPX; p++) {
      a0[p] = kr[0] * rrow[p];
      a1[p] = 0.0;
    }
    if (R >= 1) {
#pragma unroll
      for (int p = 0; p < kAinvPX; p++) {  // dx = 1
        lw[p] = rrow[p - 1];
        rw[p] = rrow[p + 1];
      }
    }
    // lw[p] = rr(c_p - dx), rw[p] = rr(c_p + dx); the step to dx + 1 reads
    // rr(c_0 - dx - 1) and rr(c_3 + dx + 1) only
    auto slide = [&](int dx) {
#pragma unroll
      for (int p = kAinvPX - 1; p > 0; p--) lw[p] = lw[p - 1];
#pragma unroll
      for (int p = 0; p + 1 < kAinvPX; p++) rw[p] = rw[p + 1];
      lw[0] = rrow[-dx - 1];
      rw[kAinvPX - 1] = rrow[kAinvPX - 1 + dx + 1];
    };
    int dx = 1;
    for (; dx + 1 <= R; dx += 2) {
      const double k1 = kr[dx], k2 = kr[dx + 1];
#pragma unroll
      for (int p = 0; p < kAinvPX; p++) a1[p] = fma(k1, lw[p] + rw[p], a1[p]);
      slide(dx);
#pragma unroll
      for (int p = 0; p < kAinvPX; p++) a0[p] = fma(k2, lw[p] + rw[p], a0[p]);
      if (dx + 2 <= R) slide(dx + 1);
    }
    if (dx <= R) {
      const double k1 = kr[dx];
#pragma unroll
      for (int p = 0; p < kAinvPX; p++) a1[p] = fma(k1, lw[p] + rw[p], a1[p]);
    }
#pragma unroll
    for (int p = 0; p < kAinvPX; p++) acc[p] += a0[p] + a1[p];
  }
#pragma unroll
  for (int p = 0; p < kAinvPX; p++) red[grp][ly * kAinvTX + kAinvPX * lq + p] = acc[p];
  __syncthreads();
  if (mine) {
    double q[kAinvGroups];
#pragma unroll
    for (int x = 0; x < kAinvGroups; x++) q[x] = red[x][tid];
#pragma unroll
    for (int h = kAinvGroups / 2; h > 0; h /= 2)
#pragma unroll
      for (int x = 0; x < h; x++) q[x] = q[x] + q[x + h];
    *o = old - T(q[0]);
  }
}

// The same two levels per launch in LDS tiles.  Workgroup (tx, ty) takes a
// kCoTX x kCoTY block of level lev+1's centre.  It stages the level lev
// places its stencils read ((2 kCoTX + 1) x (2 kCoTY + 1)) and, when those
// places meet level lev's own centre, the level lev-1 block they are
// coarsified from ((4 kCoTX + 3) x (4 kCoTY + 3)), with every global load
// issued before the first store into LDS.  It then recomputes the level lev
// values inside that centre (restrict_at, k_coarsify's expression), writes
// those of its own (2 kCoTX) x (2 kCoTY) places to level lev (each centre
// place is owned by exactly one workgroup), and coarsifies its level lev+1
// block from the staged values.  Every value is the one k_coarsify computes,
// so the result is bit-identical to two k_coarsify launches.  Stored level
// lev values inside the centre are never read (they are being overwritten
// by other workgroups): those staging loads read place (2, 2) instead, and
// the recomputed value replaces them.  grid (ntx nty, B), 256 threads;
// lev + 1 < Lmax.  With tiles set, case b runs only the tiles
// tiles[4b..4b+3] = {tx0, ty0, ntx_b, nty_b} (grid.x = the largest
// ntx_b nty_b): the places whose values can differ from those already
// stored (the rest would be recomputed from unchanged inputs).
constexpr int kCoTX = 32, kCoTY = 8;
constexpr int kCoMX = 2 * kCoTX + 1, kCoMY = 2 * kCoTY + 1, kCoFX = 4 * kCoTX + 3, kCoFY = 4 * kCoTY + 3;
constexpr int kCoFU = (kCoFX * kCoFY + 255) / 256, kCoMU = (kCoMX * kCoMY + 255) / 256;
template <typename T>
__global__ __launch_bounds__(256) void k_coarsify2t(Geo g, T* __restrict__ om, long bs, const int* __restrict__ Lc,
                                                    int lev, int ntx, const int* __restrict__ tiles) {
  __shared__ T fs[kCoFY][kCoFX];
  __shared__ T ms[kCoMY][kCoMX];
  const int b = blockIdx.y;
  if (!active(Lc, b, lev)) return;
  const bool up = active(Lc, b, lev + 1);
  int tx = blockIdx.x % ntx, ty = blockIdx.x / ntx;
  if (tiles != nullptr) {  // the case's own rectangle of tiles
    const int* tb = tiles + 4 * b;
    const int nx = tb[2];
    if ((int)blockIdx.x >= nx * tb[3]) return;
    tx = tb[0] + (int)blockIdx.x % nx;
    ty = tb[1] + (int)blockIdx.x / nx;
  }
  const int i0 = g.m / 4 + 2 + tx * kCoTX, j0 = g.n / 4 + 2 + ty * kCoTY;
  const int ma = 2 * i0 - g.m / 2 - 2, mc = 2 * j0 - g.n / 2 - 2;  // level lev place of ms[0][0]
  const int fa = 2 * ma - g.m / 2 - 2, fc = 2 * mc - g.n / 2 - 2;  // level lev-1 place of fs[0][0]
  const int clo_x = g.m / 4 + 2, chi_x = 3 * g.m / 4, clo_y = g.n / 4 + 2, chi_y = 3 * g.n / 4;
  auto centre = [&](int a, int c) { return a >= clo_x && a <= chi_x && c >= clo_y && c <= chi_y; };
  // does this block's level lev region meet level lev's centre?
  const bool deep = ma + kCoMX - 1 >= clo_x && ma <= chi_x && mc + kCoMY - 1 >= clo_y && mc <= chi_y;
  const T* rF = om + b * bs + (long)(lev - 1) * g.nw;
  const T* rM = om + b * bs + (long)lev * g.nw;
  T vf[kCoFU], vm[kCoMU];
#pragma unroll
  for (int u = 0; u < kCoMU; u++) {
    const int e = threadIdx.x + 256 * u, y = e / kCoMX, x = e - y * kCoMX;
    int a = min(ma + x, g.m), c = min(mc + y, g.n);
    if (centre(a, c)) a = c = 2;
    vm[u] = e < kCoMX * kCoMY ? rM[wix(g, a, c)] : T(0);
  }
  if (deep) {
#pragma unroll
    for (int u = 0; u < kCoFU; u++) {
      const int e = threadIdx.x + 256 * u, y = e / kCoFX, x = e - y * kCoFX;
      const int a = min(max(fa + x, 2), g.m), c = min(max(fc + y, 2), g.n);
      vf[u] = e < kCoFX * kCoFY ? rF[wix(g, a, c)] : T(0);
    }
#pragma unroll
    for (int u = 0; u < kCoFU; u++) {
      const int e = threadIdx.x + 256 * u;
      if (e < kCoFX * kCoFY) (&fs[0][0])[e] = vf[u];
    }
  }
#pragma unroll
  for (int u = 0; u < kCoMU; u++) {
    const int e = threadIdx.x + 256 * u;
    if (e < kCoMX * kCoMY) (&ms[0][0])[e] = vm[u];
  }
  __syncthreads();
  if (deep) {
    T* wM = om + b * bs + (long)lev * g.nw;
    auto Fl = [&](int x, int y) -> T { return fs[y][x]; };
#pragma unroll
    for (int u = 0; u < kCoMU; u++) {
      const int e = threadIdx.x + 256 * u, y = e / kCoMX, x = e - y * kCoMX;
      if (e < kCoMX * kCoMY && centre(ma + x, mc + y)) {
        const T v = restrict_at<T>(Fl, 2 * x + 1, 2 * y + 1);
        ms[y][x] = v;
        if (x >= 1 && x <= 2 * kCoTX && y >= 1 && y <= 2 * kCoTY) wM[wix(g, ma + x, mc + y)] = v;
      }
    }
    __syncthreads();
  }
  if (!up) return;
  const int li = threadIdx.x % kCoTX, lj = threadIdx.x / kCoTX;
  const int i = i0 + li, j = j0 + lj;
  if (i > chi_x || j > chi_y) return;
  auto Ml = [&](int x, int y) -> T { return ms[y][x]; };
  om[b * bs + (long)(lev + 1) * g.nw + wix(g, i, j)] = restrict_at<T>(Ml, 1 + 2 * li, 1 + 2 * lj);
}

template <typename T>
__global__ void k_coarsify2(Geo g, T* __restrict__ om, long bs, const int* __restrict__ Lc, int lev, int nblk) {
  const int wi = g.m / 2 - 1, wj = g.n / 2 - 1;
  // XCD-contiguous, so the fine rows shared by the stencils of neighbouring
  // groups stay in one L2
  const int x = xcd_contiguous(blockIdx.x, gridDim.x);
  const int upper = x >= nblk;
  const int t = (x - upper * nblk) * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  const int lc = lev + upper;  // level written
  if (t >= wi * wj || !active(Lc, b, lc)) return;
  const int i = g.m / 4 + 2 + t % wi, j = g.n / 4 + 2 + t / wi;
  const int ii = 2 * i - g.m / 2 - 1, jj = 2 * j - g.n / 2 - 1;
  const T* r1 = om + b * bs + (long)(lc - 1) * g.nw;  // level lc-1 as stored
  auto F1 = [&](int a, int c) -> T { return r1[wix(g, a, c)]; };
  if (!upper) {
    om[b * bs + (long)lc * g.nw + wix(g, i, j)] = restrict_at<T>(F1, ii, jj);
    return;
  }
  // level lev values: inside its centre, coarsified from level lev-1 here.
  // The nine restrictions read one 7 x 7 block of level lev-1 around
  // (X, Y), the level lev-1 place of (ii, jj); the block (places clamped
  // into the field, used only where they are in it) and the nine stored
  // level lev values (read at (2, 2), outside the centre being written, for
  // places inside it, whose value is recomputed) are all loaded before any
  // is used.  The arithmetic is restrict_at's, so the result is unchanged.
  const T* r0 = om + b * bs + (long)(lev - 1) * g.nw;
  const int X = 2 * ii - g.m / 2 - 1, Y = 2 * jj - g.n / 2 - 1;
  T blk[7][7];
#pragma unroll
  for (int dy = 0; dy < 7; dy++)
#pragma unroll
    for (int dx = 0; dx < 7; dx++) {
      const int xa = min(max(X - 3 + dx, 2), g.m), ya = min(max(Y - 3 + dy, 2), g.n);
      blk[dy][dx] = r0[wix(g, xa, ya)];
    }
  bool in[3][3];
  T f1v[3][3];
#pragma unroll
  for (int dc = 0; dc < 3; dc++)
#pragma unroll
    for (int da = 0; da < 3; da++) {
      const int a = ii - 1 + da, c = jj - 1 + dc;
      in[dc][da] = a >= g.m / 4 + 2 && a <= 3 * g.m / 4 && c >= g.n / 4 + 2 && c <= 3 * g.n / 4;
      f1v[dc][da] = F1(in[dc][da] ? 2 : a, in[dc][da] ? 2 : c);
    }
  // (ii-1+da, jj-1+dc) is level lev-1 place (X+2da-2, Y+2dc-2), block entry
  // (1+2da, 1+2dc); the indices below are all constants
  auto B0 = [&](int a, int c) -> T { return blk[c][a]; };
  T fm[3][3];
#pragma unroll
  for (int dc = 0; dc < 3; dc++)
#pragma unroll
    for (int da = 0; da < 3; da++)
      fm[dc][da] = in[dc][da] ? restrict_at<T>(B0, 1 + 2 * da, 1 + 2 * dc) : f1v[dc][da];
  auto Fm = [&](int a, int c) -> T { return fm[c][a]; };
  om[b * bs + (long)lc * g.nw + wix(g, i, j)] = restrict_at<T>(Fm, 1, 1);
}

// curl face value (src/operators.f90:243-288) for 0-based face f; the edge
// vector is get_bc(sc, 1) (sc == nullptr: zero edge values), sc of element
// type TC.
template <typename T, typename TC = T>
__device__ __forceinline__ T curl_face(const Geo& g, int f, const T* __restrict__ s, const TC* __restrict__ sc) {
  const int m = g.m, n = g.n;
  auto X = [&](int i, int j) -> T { return s[wix(g, i, j)]; };
  auto B = [&](int pos) -> T { return sc ? T(getbc_at<TC>(g, sc, pos, TC(1))) : T(0); };
  const int nu = (m + 1) * n;
  if (f < nu) {
    const int i = f % (m + 1) + 1, j = f / (m + 1) + 1;
    if (i == 1) return B(g.left + j + 1) - B(g.left + j);
    if (i == m + 1) return B(g.right + j + 1) - B(g.right + j);
    if (j == 1) return X(i, 2) - B(g.bottom + i);
    if (j == n) return B(g.top + i) - X(i, n);
    return X(i, j + 1) - X(i, j);
  } else {
    const int fv = f - nu;
    const int i = fv % m + 1, j = fv / m + 1;
    if (j == 1) return B(g.bottom + i) - B(g.bottom + i + 1);
    if (j == n + 1) return B(g.top + i) - B(g.top + i + 1);
    if (i == 1) return -X(2, j) + B(g.left + j);
    if (i == m) return -B(g.right + j) + X(m, j);
    return X(i, j) - X(i + 1, j);
  }
}

// curl_face of the streamfunction s + c (c may be null) with the coarser
// level's sc + cc (cc may be null): every place's two parts are added before
// get_bc's averages, so the values are curl_face's of the summed fields.
template <typename T, typename TC = T>
__device__ __forceinline__ T curl_face_sum(const Geo& g, int f, const T* __restrict__ s, const T* __restrict__ c,
                                           const TC* __restrict__ sc, const TC* __restrict__ cc) {
  const int m = g.m, n = g.n;
  auto X = [&](int i, int j) -> T {
    const int e = wix(g, i, j);
    return c ? T(s[e] + c[e]) : s[e];
  };
  auto R1 = [&](int a, int b) -> TC {
    const int e = (a - 1) + (b - 1) * g.mm1;
    return cc ? TC(sc[e] + cc[e]) : sc[e];
  };
  auto B = [&](int pos) -> T {
    if (!sc) return T(0);
    TC v;
    if (pos > g.bottom) {  // getbc_at's places and averages
      const bool top = pos > g.top;
      const int i0 = pos - (top ? g.top : g.bottom) - 1;
      const int jb = top ? 3 * n / 4 : n / 4;
      if ((i0 & 1) == 0) v = R1(m / 4 + i0 / 2, jb);
      else v = TC(0.5) * (R1(m / 4 + (i0 + 1) / 2, jb) + R1(m / 4 - 1 + (i0 + 1) / 2, jb));
    } else {
      const bool right = pos > g.right;
      const int j0 = pos - (right ? g.right : g.left) - 1;
      const int ib = right ? 3 * m / 4 : m / 4;
      if ((j0 & 1) == 0) v = R1(ib, n / 4 + j0 / 2);
      else v = TC(0.5) * (R1(ib, n / 4 + (j0 + 1) / 2) + R1(ib, n / 4 - 1 + (j0 + 1) / 2));
    }
    return T(v * TC(1));
  };
  const int nu = (m + 1) * n;
  if (f < nu) {
    const int i = f % (m + 1) + 1, j = f / (m + 1) + 1;
    if (i == 1) return B(g.left + j + 1) - B(g.left + j);
    if (i == m + 1) return B(g.right + j + 1) - B(g.right + j);
    if (j == 1) return X(i, 2) - B(g.bottom + i);
    if (j == n) return B(g.top + i) - X(i, n);
    return X(i, j + 1) - X(i, j);
  } else {
    const int fv = f - nu;
    const int i = fv % m + 1, j = fv / m + 1;
    if (j == 1) return B(g.bottom + i) - B(g.bottom + i + 1);
    if (j == n + 1) return B(g.top + i) - B(g.top + i + 1);
    if (i == 1) return -X(2, j) + B(g.left + j);
    if (i == m) return -B(g.right + j) + X(m, j);
    return X(i, j) - X(i + 1, j);
  }
}

// vel = curl(s_lev, get_bc(s_{lev+1})) over all faces (zero edges on a case's
// coarsest level).  grid.x over nq; s/sc stride bs, vel stride vbs
template <typename T>
__global__ void k_curl(Geo g, const T* __restrict__ s, const T* __restrict__ sc, T* __restrict__ vel, long bs,
                       long vbs, const int* __restrict__ Lc, int lev) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (f >= g.nq || !active(Lc, b, lev)) return;
  const T* c = (sc != nullptr && has_coarser(g, Lc, b, lev)) ? sc + b * bs : nullptr;
  vel[b * vbs + f] = curl_face<T>(g, f, s + b * bs, c);
}

// k_curl on every level at once, once all streamfunctions are final:
// grid (nq / 256, Lmax, batch); s stride bs per case, nw per level; vel
// stride vbs per case, nq per level.
template <typename T>
__global__ void k_curl_levels(Geo g, const T* __restrict__ s, T* __restrict__ vel, long bs, long vbs,
                              const int* __restrict__ Lc) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  const int lev = blockIdx.y, b = blockIdx.z;
  if (f >= g.nq || !active(Lc, b, lev)) return;
  const T* sl = s + b * bs + (long)lev * g.nw;
  const bool coarser = lev + 1 < g.Lmax && has_coarser(g, Lc, b, lev);
  vel[b * vbs + (long)lev * g.nq + f] = curl_face<T>(g, f, sl, coarser ? sl + g.nw : nullptr);
}

// rot(x): Nq -> W (src/operators.f90:291-304).  grid.x over nw
template <typename T>
__global__ void k_rot(Geo g, const T* __restrict__ x, T* __restrict__ r) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.nw) return;
  const int i = e % g.mm1 + 2, j = e / g.mm1 + 2;
  r[e] = x[vix(g, i, j)] - x[vix(g, i - 1, j)] - x[uix(g, i, j)] + x[uix(g, i, j - 1)];
}

// rot(nonlinear(...)) as a standalone field (operator test entry)
template <typename T>
__global__ void k_rot_nonlinear(Geo g, const T* __restrict__ om, const T* __restrict__ q, const T* __restrict__ q0,
                                const T* __restrict__ bc, T* __restrict__ r) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.nw) return;
  const int i = e % g.mm1 + 2, j = e / g.mm1 + 2;
  r[e] = rot_nonlinear<T>(g, i, j, om, q, q0, bc);
}

// ---------------------------------------------------------------------------
// smearing (E / E^T) on compact per-case tables
// ---------------------------------------------------------------------------
// s = sum over entries e = lo..hi-1, in order, of w[e] * value(idx[e]).
// The entries are taken K at a time: all K index/weight loads are issued,
// then all K value loads, then the K products are added in entry order, so
// a list of up to K entries costs three dependent memory latencies instead
// of one per entry.  Loads past hi are not issued.
template <int K, class F>
__device__ __forceinline__ double gather_sum(int lo, int hi, const int* __restrict__ idx,
                                             const double* __restrict__ w, F value) {
  double s = 0.0;
  for (int base = lo; base < hi; base += K) {
    int id[K];
    double ww[K], v[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
      const bool ok = base + k < hi;
      id[k] = ok ? idx[base + k] : 0;
      ww[k] = ok ? w[base + k] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = (base + k < hi) ? value(id[k]) : 0.0;
#pragma unroll
    for (int k = 0; k < K; k++)
      if (base + k < hi) s = s + ww[k] * v[k];
  }
  return s;
}
// wave64 and workgroup (256 threads = 4 waves) reductions, fixed order
__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;  // complete in lane 0
}
__device__ __forceinline__ double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  return v;
}
// workgroup-wide sum / max (blockDim.x a multiple of 64, at most 1024),
// returned to every thread; waves are combined in a fixed order
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = red[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); w++) s += red[w];
  return s;
}
__device__ __forceinline__ double block_max(double v, double* red) {
  v = wave_max(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = red[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); w++) s = fmax(s, red[w]);
  return s;
}
constexpr int kFrcK = 9;   // entries of one force: at most 3 x 3 faces of the delta support
constexpr int kFaceK = 6;  // entries of one face: the IB points whose support covers it

template <typename H>
__device__ __forceinline__ double reg_face(const Smear& sm, const CaseSm& c, int a, const H* __restrict__ h0) {
  return gather_sum<kFaceK>(sm.face_ptr[c.pbase + a], sm.face_ptr[c.pbase + a + 1], sm.face_pt, sm.face_w,
                            [&](int p) { return double(h0[p]); });
}

// regT(curl(s1, get_bc(sc))) - sub for force f of case c (sub may be null);
// sc may have another element type (the FP32 coarse level of a mixed case)
template <typename T, typename TC = T>
__device__ __forceinline__ double regT_curl_at(const Geo& g, const Smear& sm, const CaseSm& c, int f,
                                               const T* __restrict__ s1, const TC* __restrict__ sc,
                                               const double* __restrict__ sub, const T* __restrict__ s1c = nullptr) {
  // Interior faces take curl = s1[ia] - s1[ib] from the table, so the K
  // entries' index, weight and field loads are each issued together; faces
  // on the domain edge (ia < 0, only for bodies touching the edge) fall back
  // to curl_face with its get_bc terms.  Same values as curl_face.  With s1c
  // the streamfunction is s1 + s1c, added here (interior faces only: the
  // caller passes s1c only when no stencil face lies on the edge).
  double s = 0.0;
  if (sm.f9_w != nullptr) {  // padded: the slots past a force's entries add +0
    const int base = c.foff + f;
    int ia[kFrcK], ib[kFrcK];
    double ww[kFrcK];
    T va[kFrcK], vb[kFrcK];
#pragma unroll
    for (int k = 0; k < kFrcK; k++) {
      ia[k] = sm.f9_ia[k * sm.f9_n + base];
      ib[k] = sm.f9_ib[k * sm.f9_n + base];
      ww[k] = sm.f9_w[k * sm.f9_n + base];
    }
#pragma unroll
    for (int k = 0; k < kFrcK; k++) {
      va[k] = s1[ia[k] < 0 ? 0 : ia[k]];
      vb[k] = s1[ib[k] < 0 ? 0 : ib[k]];
    }
    if (s1c != nullptr) {
      T ca[kFrcK], cb[kFrcK];
#pragma unroll
      for (int k = 0; k < kFrcK; k++) {
        ca[k] = s1c[ia[k] < 0 ? 0 : ia[k]];
        cb[k] = s1c[ib[k] < 0 ? 0 : ib[k]];
      }
#pragma unroll
      for (int k = 0; k < kFrcK; k++) {
        va[k] = va[k] + ca[k];
        vb[k] = vb[k] + cb[k];
      }
    }
#pragma unroll
    for (int k = 0; k < kFrcK; k++) {
      const double v =
          ia[k] >= 0 ? double(va[k] - vb[k]) : double(curl_face<T, TC>(g, sm.f9_face[k * sm.f9_n + base], s1, sc));
      s = s + ww[k] * v;
    }
    if (sub) s = s - sub[c.foff + f];
    return s;
  }
  const int lo = sm.frc_ptr[c.rbase + f], hi = sm.frc_ptr[c.rbase + f + 1];
  for (int base = lo; base < hi; base += kFrcK) {
    int ia[kFrcK], ib[kFrcK];
    double ww[kFrcK];
    T va[kFrcK], vb[kFrcK];
#pragma unroll
    for (int k = 0; k < kFrcK; k++) {
      const bool ok = base + k < hi;
      ia[k] = ok ? sm.frc_ia[base + k] : 0;
      ib[k] = ok ? sm.frc_ib[base + k] : 0;
      ww[k] = ok ? sm.frc_w[base + k] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < kFrcK; k++) {
      va[k] = s1[ia[k] < 0 ? 0 : ia[k]];
      vb[k] = s1[ib[k] < 0 ? 0 : ib[k]];
    }
    if (s1c != nullptr) {
#pragma unroll
      for (int k = 0; k < kFrcK; k++) {
        va[k] = va[k] + s1c[ia[k] < 0 ? 0 : ia[k]];
        vb[k] = vb[k] + s1c[ib[k] < 0 ? 0 : ib[k]];
      }
    }
#pragma unroll
    for (int k = 0; k < kFrcK; k++) {
      if (base + k >= hi) break;
      const double v =
          ia[k] >= 0 ? double(va[k] - vb[k]) : double(curl_face<T, TC>(g, sm.frc_face[base + k], s1, sc));
      s = s + ww[k] * v;
    }
  }
  if (sub) s = s - sub[c.foff + f];
  return s;
}

// rr = rot(reg(h0)) on the active vertices only; all other entries of rr are
// never written and stay zero.  h0 of case b at h0 + foff (or b*hbs when
// hbs > 0, for batches of force vectors of one case).  grid.x over max nvtx
template <typename T, typename H>
__global__ void k_rotreg(Smear sm, const H* __restrict__ h0, long hbs, T* __restrict__ rr, long bs, int one_case) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  const CaseSm c = sm.cs[one_case >= 0 ? one_case : b];
  if (v >= c.nvtx) return;
  const H* h = hbs > 0 ? h0 + b * hbs : h0 + c.foff;
  double fv[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int a = sm.vtx_face[4 * (c.vbase + v) + k];
    fv[k] = a >= 0 ? reg_face<H>(sm, c, a, h) : 0.0;
  }
  rr[b * bs + sm.vtx_id[c.vbase + v]] = T(fv[0] - fv[1] - fv[2] + fv[3]);
}

// h = reg(h0) on the active faces of case 0 (operator test entry)
template <typename T>
__global__ void k_reg_active(Smear sm, const double* __restrict__ h0, T* __restrict__ h) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const CaseSm c = sm.cs[0];
  if (a >= c.nact) return;
  h[sm.face_id[c.fbase + a]] = T(reg_face<double>(sm, c, a, h0));
}

// h0 = regT(h) for case 0 (operator test entry)
template <typename T>
__global__ void k_regT(Smear sm, const T* __restrict__ h, double* __restrict__ h0) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  const CaseSm c = sm.cs[0];
  if (f >= c.nf) return;
  h0[f] = gather_sum<kFrcK>(sm.frc_ptr[c.rbase + f], sm.frc_ptr[c.rbase + f + 1], sm.frc_face, sm.frc_w,
                            [&](int face) { return double(h[face]); });
}

// out = regT(curl(s1, get_bc(s2))) - sub (sub may be null): the right-hand
// side of the force system (src/operators.f90:163-164), and the columns of
// a_times (src/operators.f90:459-461) for batches of one case.  s2 is used
// only for cases with more than one level; s1c (may be null) is added to
// s1 (the same layout).  grid.x over max nf.
template <typename T>
__global__ void k_regT_curl(Geo g, Smear sm, const T* __restrict__ s1, const T* __restrict__ s2, long bs,
                            const double* __restrict__ sub, double* __restrict__ out, long obs,
                            const int* __restrict__ Lc, int one_case, const T* __restrict__ s1c) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  const int ci = one_case >= 0 ? one_case : b;
  const CaseSm c = sm.cs[ci];
  if (f >= c.nf) return;
  const T* sc = (s2 != nullptr && levels(g, Lc, b) > 1) ? s2 + b * bs : nullptr;  // Lc indexed by batch entry
  const long o = obs > 0 ? b * obs + f : c.foff + f;
  out[o] = regT_curl_at<T>(g, sm, c, f, s1 + b * bs, sc, sub, s1c != nullptr ? s1c + b * bs : nullptr);
}

// motion_eff = motion_in - regT(q0_1) (src/operators.f90:155), per case;
// regT(q0_1) is kept in rq0 for the actuator points, whose motion is
// replaced every step (k_actuation)
template <typename T>
__global__ void k_motion(Smear sm, const T* __restrict__ q0, long qbs, const double* __restrict__ motion_in,
                         double* __restrict__ motion, double* __restrict__ rq0) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  const CaseSm c = sm.cs[b];
  if (f >= c.nf) return;
  const T* q0b = q0 + b * qbs;
  const double s = gather_sum<kFrcK>(sm.frc_ptr[c.rbase + f], sm.frc_ptr[c.rbase + f + 1], sm.frc_face, sm.frc_w,
                                     [&](int face) { return double(q0b[face]); });
  motion[c.foff + f] = motion_in[c.foff + f] - s;
  rq0[c.foff + f] = s;
}

// One point of an actuator slot (src/grid.f90:579-633, 636-718): at step it
// the jet velocity is (u_j + u_j_osc sin(omg_j dt it)) (cos th, sin th) with
// th = pi/180 theta_j.
struct ActPt {
  int f;             // the point's x force in the case (its y force is f + nb)
  double uj, ujo;    // steady and oscillating amplitude
  double omgdt;      // omg_j * dt
  double cth, sth;   // cos th, sin th
};

// actuation (src/operators.f90:150-152 with src/grid.f90:579-633): the
// motion of every actuator point for this step, u_jet * delta - regT(q0_1),
// replacing what the caller's motion vector held there (as actuation
// overwrites move_body's result).  One workgroup per case; itime[b] holds the
// case's last step and is advanced here, once per step, so a captured step
// replays with the right time.  The products are rounded one by one, in the
// reference's order.
__global__ void k_actuation(Smear sm, const int* __restrict__ aptr, const ActPt* __restrict__ pts,
                            const double* __restrict__ rq0, int* __restrict__ itime,
                            double* __restrict__ motion) {
  const int b = blockIdx.x;
  const CaseSm c = sm.cs[b];
  const int it = itime[b] + 1;
  __syncthreads();  // every thread has read itime[b] before it is advanced
  for (int p = aptr[b] + threadIdx.x; p < aptr[b + 1]; p += blockDim.x) {
    const ActPt a = pts[p];
    const double u = __dadd_rn(a.uj, __dmul_rn(a.ujo, sin(__dmul_rn(a.omgdt, (double)it))));
    const int fx = c.foff + a.f, fy = fx + c.nb;
    motion[fx] = __dmul_rn(__dmul_rn(u, a.cth), c.delta) - rq0[fx];
    motion[fy] = __dmul_rn(__dmul_rn(u, a.sth), c.delta) - rq0[fy];
  }
  if (threadIdx.x == 0) itime[b] = it;
}

// ---------------------------------------------------------------------------
// force solve: fb = cholsl(rhsf) as two triangular products with the
// explicit inverse Li = L^{-1} of the Cholesky factor (column-major, lower):
// y = Li r, then x = Li^T y.  One wave per output row, 4 rows per block.
// ---------------------------------------------------------------------------
static __global__ void k_trmv_lower(Smear sm, const double* __restrict__ Li, const double* __restrict__ r,
                             double* __restrict__ y) {
  const CaseSm c = sm.cs[blockIdx.y];
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= c.nf) return;
  const double* L = Li + c.loff;
  const double* rr = r + c.foff;
  double s = 0.0;
  for (int k = lane; k <= i; k += 64) s += L[i + (long)k * c.nf] * rr[k];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) y[c.foff + i] = s;
}
// x[i] (+)= sum_{k>=i} Li(k,i) y[k]   (column i of Li, contiguous)
static __global__ void k_trmv_upper_t(Smear sm, const double* __restrict__ Li, const double* __restrict__ y,
                               double* __restrict__ x, int accumulate) {
  const CaseSm c = sm.cs[blockIdx.y];
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= c.nf) return;
  const double* col = Li + c.loff + (long)i * c.nf;
  const double* yy = y + c.foff;
  double s = 0.0;
  for (int k = i + lane; k < c.nf; k += 64) s += col[k] * yy[k];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) x[c.foff + i] = accumulate ? x[c.foff + i] + s : s;
}
// refinement residual r = b - A x with the symmetric A the factor was built
// from (upper triangle of the ibfs.chd matrix, stored here as a full
// matrix, so row i is read as the contiguous column i)
static __global__ void k_symv_residual(Smear sm, const double* __restrict__ A, const double* __restrict__ bvec,
                                const double* __restrict__ x, double* __restrict__ r) {
  const CaseSm c = sm.cs[blockIdx.y];
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= c.nf) return;
  const double* Ai = A + c.loff + (long)i * c.nf;
  double s = 0.0;
  for (int k = lane; k < c.nf; k += 64) s += Ai[k] * x[c.foff + k];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) r[c.foff + i] = bvec[c.foff + i] - s;
}

// A^-1 = Li^T Li of case ci (column-major, both triangles written from one
// product, so the stored matrix is exactly symmetric).  grid over nf^2
static __global__ void k_make_ainv(Smear sm, int ci, const double* __restrict__ Li, double* __restrict__ Ai) {
  const CaseSm c = sm.cs[ci];
  const int nf = c.nf;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)nf * nf) return;
  const int i = (int)(t % nf), j = (int)(t / nf);
  if (i > j) return;
  const double* ci_col = Li + c.loff + (long)i * nf;  // Li(k, i), zero for k < i
  const double* cj_col = Li + c.loff + (long)j * nf;
  double s = 0.0;
  for (int k = j; k < nf; k++) s += ci_col[k] * cj_col[k];
  Ai[c.loff + i + (long)j * nf] = s;
  Ai[c.loff + j + (long)i * nf] = s;
}

// G = rot(reg(A^-1)) of case ci on its active vertices: G[v*nf + j] is the
// step-3 source at vertex v for the unit force j (row-major, nvtx x nf).
// grid over nvtx*nf
static __global__ void k_make_g(Smear sm, int ci, const double* __restrict__ Ai, double* __restrict__ G) {
  const CaseSm c = sm.cs[ci];
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)c.nvtx * c.nf) return;
  const int j = (int)(t % c.nf), v = (int)(t / c.nf);
  const double* col = Ai + c.loff + (long)j * c.nf;
  double fv[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int a = sm.vtx_face[4 * (c.vbase + v) + k];
    fv[k] = a >= 0 ? reg_face<double>(sm, c, a, col) : 0.0;
  }
  G[c.goff + (long)v * c.nf + j] = fv[0] - fv[1] - fv[2] + fv[3];
}

// Force system of the step, one launch.  Every workgroup of case b
// evaluates the whole right-hand side rhsf = regT(curl(s_1, get_bc(s_2))) -
// motion (src/operators.f90:163-164) into LDS and multiplies kForceRows
// rows of [A^-1; G] with it: rows < nf give fb = A^-1 rhsf (the cholsl of
// src/operators.f90:166, with the explicit inverse of the factored matrix),
// the others give the step-3 source rot(reg(fb)) (src/operators.f90:174) at
// the active vertices; every other entry of rr stays zero.
//   mode 0: as above.
//   mode 1: rows < nf only; rhsf is also written out (first solve of the
//           refined form).
//   mode 2: the refinement sweep with the residual res = rhsf - A fb:
//           fb += A^-1 res and rr = G (rhsf + res).
//   mode 3: as mode 0 with rhsf read from memory (written by k_regT_curl
//           beforehand), so that the right-hand side is evaluated once per
//           case instead of once per workgroup.
// Latency: the right-hand side takes two forces per thread and pass, whose
// gathers are independent; each wave holds kRowsPerWave rows and walks the
// columns once, so the row loads of one column block are all in flight
// together.
// grid (ceil(max rows / kForceRows), B), kForceThreads (512) threads, LDS 2*nf_max doubles.
constexpr int kForceThreads = 512;
constexpr int kRowsPerWave = 2;
constexpr int kColBlock = 10;
constexpr int kForceRows = kForceThreads / 64 * kRowsPerWave;

template <typename T, typename TS = T>
__global__ __launch_bounds__(kForceThreads) void k_force_solve(Geo g, Smear sm, const TS* __restrict__ s1,
                                                     const TS* __restrict__ s2, long bs,
                                                     const double* __restrict__ motion, const int* __restrict__ Lc,
                                                     const double* __restrict__ Ainv, const double* __restrict__ G,
                                                     double* __restrict__ fb, T* __restrict__ rr, long rbs,
                                                     double* __restrict__ rhsf, const double* __restrict__ res,
                                                     int mode) {
  extern __shared__ double vec[];
  const int b = blockIdx.y, tid = threadIdx.x;
  const CaseSm c = case_sm(sm, b);
  const int nf = c.nf;
  const int nrows = mode == 1 ? nf : nf + c.nvtx;
  const int row0 = blockIdx.x * kForceRows;
  if (row0 >= nrows) return;
  const int NT = blockDim.x;
  const int lane = tid & 63;
  const int rbase = row0 + (tid >> 6) * kRowsPerWave;
  const double* Mr[kRowsPerWave];
  int xo[kRowsPerWave];
  double acc[kRowsPerWave];
#pragma unroll
  for (int q = 0; q < kRowsPerWave; q++) {
    const int r = min(rbase + q, nrows - 1);  // rows past the end repeat the last row and are not stored
    const bool frow = r < nf;
    Mr[q] = frow ? Ainv + c.loff + (long)r * nf : G + c.goff + (long)(r - nf) * nf;
    xo[q] = (!frow && mode == 2) ? nf : 0;
    acc[q] = 0.0;
  }
  // the first kColBlock columns per lane of this wave's rows, loaded before
  // the right-hand side is evaluated (they do not depend on it)
  double mv[kColBlock][kRowsPerWave];
  auto load_block = [&](int k0) {
#pragma unroll
    for (int u = 0; u < kColBlock; u++) {
      const int k = min(k0 + 64 * u, nf - 1);
#pragma unroll
      for (int q = 0; q < kRowsPerWave; q++) mv[u][q] = k >= 0 ? Mr[q][k] : 0.0;
    }
  };
  load_block(lane);
  if (mode < 2) {
    const TS* s1b = s1 + b * bs;
    const TS* sc = (s2 != nullptr && levels(g, Lc, b) > 1) ? s2 + b * bs : nullptr;
    for (int f0 = tid; f0 < nf; f0 += 2 * NT) {
      const int f1 = f0 + NT;
      const double v0 = regT_curl_at<TS>(g, sm, c, f0, s1b, sc, motion);
      const double v1 = f1 < nf ? regT_curl_at<TS>(g, sm, c, f1, s1b, sc, motion) : 0.0;
      vec[f0] = v0;
      if (f1 < nf) vec[f1] = v1;
      if (mode == 1 && blockIdx.x == 0) {
        rhsf[c.foff + f0] = v0;
        if (f1 < nf) rhsf[c.foff + f1] = v1;
      }
    }
  } else if (mode == 2) {
    for (int f = tid; f < nf; f += NT) {
      const double r = res[c.foff + f];
      vec[f] = r;
      vec[nf + f] = rhsf[c.foff + f] + r;
    }
  } else {
    for (int f = tid; f < nf; f += NT) vec[f] = rhsf[c.foff + f];
  }
  __syncthreads();
  // the sums keep the column order k = lane, lane+64, ... of a plain loop
  for (int k0 = lane;;) {
#pragma unroll
    for (int u = 0; u < kColBlock; u++) {
      const int k = k0 + 64 * u;
      if (k < nf) {
#pragma unroll
        for (int q = 0; q < kRowsPerWave; q++) acc[q] += mv[u][q] * vec[xo[q] + k];
      }
    }
    k0 += 64 * kColBlock;
    if (k0 >= nf) break;
    load_block(k0);
  }
#pragma unroll
  for (int q = 0; q < kRowsPerWave; q++) {
    const double s = wave_sum(acc[q]);
    const int r = rbase + q;
    if (lane == 0 && r < nrows) {
      if (r < nf) fb[c.foff + r] = mode == 2 ? fb[c.foff + r] + s : s;
      else rr[b * rbs + sm.vtx_id[c.vbase + r - nf]] = T(s);
    }
  }
}

// ---------------------------------------------------------------------------
// conjugate gradients on the force system: cjgr (src/operators.f90:393-441)
// for one case, its vectors (nf doubles) in device memory and its scalars in
// CgState.  Each kernel is one workgroup; the dot products are summed in
// index order by one thread, as DOT_PRODUCT, and the vector updates are
// not contracted into FMAs, so they round as the reference's statements.
// ---------------------------------------------------------------------------
// done: the reference's loop test (iter < cg_max_iter and delta_new > eps)
// has failed; every later CG kernel is then a no-op, so iterations can be
// replayed in batches without changing the iterate.
struct CgState {
  double delta_new, delta_old, alpha, beta, eps;
  int iter, max_iter, done, pad;
};

// a . b summed in index order; called by every thread of the workgroup
// (the vectors are staged through LDS in chunks by all threads, one memory
// round trip per chunk, and thread 0 sums the chunk); the sum is returned
// to thread 0
constexpr int kDotChunk = 1024;
// a . b in a fixed order that does not depend on timing: each thread sums
// its strided terms in index order, then the 256 partial sums are added in a
// fixed tree (wave shuffles, then the four waves' sums in order); returned
// to every thread.  256 threads.
__device__ double dot_tree(int nf, const double* __restrict__ a, const double* __restrict__ b) {
#pragma clang fp contract(off)
  __shared__ double wsum[4];
  double s = 0.0;
  for (int f = threadIdx.x; f < nf; f += 256) s = s + a[f] * b[f];
  for (int off = 32; off > 0; off >>= 1) s = s + __shfl_xor(s, off, 64);
  __syncthreads();  // wsum is free (an earlier call may still read it)
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
  __syncthreads();
  return ((wsum[0] + wsum[1]) + (wsum[2] + wsum[3]));
}
__device__ double dot_in_order(int nf, const double* __restrict__ a, const double* __restrict__ b) {
#pragma clang fp contract(off)
  __shared__ double sa[kDotChunk], sb[kDotChunk];
  double s = 0.0;
  for (int c0 = 0; c0 < nf; c0 += kDotChunk) {
    const int cn = min(kDotChunk, nf - c0);
    for (int f = threadIdx.x; f < cn; f += blockDim.x) {
      sa[f] = a[c0 + f];
      sb[f] = b[c0 + f];
    }
    __syncthreads();
    if (threadIdx.x == 0)
      for (int f = 0; f < cn; f++) s = s + sa[f] * sb[f];
    __syncthreads();
  }
  return s;
}

// r = b - Ax; d = r; delta_new = r . d; iter = 0 (src/operators.f90:407-413)
// columns per slice of k_gemv_dense (nslice = ceil(nf / 32)): the single
// workgroup of k_cg_iter sums nslice partials per force, so few slices keep
// its chain of loads short (522 forces, CG at 512^2 x 5: 8 columns 320
// steps/s, 16 406, 32 409, 64 344, too few workgroups for the product)
constexpr int kGemvCols = 32;
// q = A x with a dense A (column-major, nf x nf) for the CG form when no
// body moves: workgroup (r, s) takes 256 rows and the s-th of nslice column
// slices, and writes its partial sums to part[s * nf + row]; the CG kernels
// add the slices in slice order (form_q), so the product is deterministic.
// grid (ceil(nf/256), nslice), 256 threads.
// With cs set, the product is skipped once the CG loop is done and, with
// only_resid, in every iteration but the ones that recompute the residual
// from b - A x (every 50th), so one captured graph serves every iteration.
static __global__ __launch_bounds__(256) void k_gemv_dense(int nf, const double* __restrict__ A,
                                                    const double* __restrict__ x, double* __restrict__ part,
                                                    int nslice, const CgState* __restrict__ cs, int only_resid) {
  if (cs && (cs->done || (only_resid && cs->iter % 50 != 0))) return;
  const int i = blockIdx.x * 256 + threadIdx.x, sl = blockIdx.y;
  const int cw = (nf + nslice - 1) / nslice, j0 = sl * cw, j1 = min(nf, j0 + cw);
  if (i >= nf) return;
  // a slice is at most kGemvCols columns: every load of a thread is issued
  // before the first product (out-of-range columns load column j0)
  double a[kGemvCols], xv[kGemvCols];
#pragma unroll
  for (int u = 0; u < kGemvCols; u++) {
    const int j = j0 + u < j1 ? j0 + u : j0;
    a[u] = A[i + (long)j * nf];
    xv[u] = x[j];
  }
  double acc = 0.0;
#pragma unroll
  for (int u = 0; u < kGemvCols; u++)
    if (j0 + u < j1) acc += a[u] * xv[u];
  part[(long)sl * nf + i] = acc;
}
// q = the sum of the slices of k_gemv_dense, in slice order (by the
// threads of the calling single-workgroup CG kernel)
__device__ __forceinline__ void form_q(int nf, const double* __restrict__ part, int nslice, double* __restrict__ q) {
  for (int f = threadIdx.x; f < nf; f += blockDim.x) {
    double acc = 0.0;
#pragma unroll 16
    for (int sl = 0; sl < nslice; sl++) acc += part[(long)sl * nf + f];
    q[f] = acc;
  }
  __syncthreads();
}

// part != nullptr: Ax is first formed from the slices of k_gemv_dense
// With part set (the dense form) the products come from k_gemv_dense's
// slices and the dot products use the fixed tree (dot_tree); otherwise (the
// a_times form) they are summed in index order like DOT_PRODUCT.
static __global__ __launch_bounds__(256) void k_cg_start(int nf, const double* __restrict__ b,
                                                  double* __restrict__ Ax, double* __restrict__ r,
                                                  double* __restrict__ d, CgState* __restrict__ cs,
                                                  const double* __restrict__ part, int nslice) {
#pragma clang fp contract(off)
  if (part) form_q(nf, part, nslice, Ax);
  for (int f = threadIdx.x; f < nf; f += blockDim.x) {
    const double v = b[f] - Ax[f];
    r[f] = v;
    d[f] = v;
  }
  __syncthreads();
  const double dn = part ? dot_tree(nf, r, d) : dot_in_order(nf, r, d);
  if (threadIdx.x == 0) {
    cs->delta_new = dn;
    cs->iter = 0;
    cs->done = !(0 < cs->max_iter && dn > cs->eps);
  }
}

// iter += 1; alpha = delta_new / (d . q); x = x + alpha d  (q = A d;
// src/operators.f90:415-418)
static __global__ __launch_bounds__(256) void k_cg_alpha(int nf, const double* __restrict__ d,
                                                  double* __restrict__ q, double* __restrict__ x,
                                                  CgState* __restrict__ cs, const double* __restrict__ part,
                                                  int nslice) {
#pragma clang fp contract(off)
  __shared__ double alpha;
  if (cs->done) return;
  if (part) form_q(nf, part, nslice, q);
  const double dq = part ? dot_tree(nf, d, q) : dot_in_order(nf, d, q);
  if (threadIdx.x == 0) {
    alpha = cs->delta_new / dq;
    cs->alpha = alpha;
    cs->iter = cs->iter + 1;
  }
  __syncthreads();
  for (int f = threadIdx.x; f < nf; f += blockDim.x) x[f] = x[f] + alpha * d[f];
}

// the residual, every 50th iteration from scratch (q = A x: r = b - q),
// otherwise r = r - alpha q (q = A d); then delta_old = delta_new,
// delta_new = r . r, beta = delta_new / delta_old, d = r + beta d
// (src/operators.f90:420-431)
// recompute < 0: the device decides (iteration iter % 50 == 0, as cjgr)
static __global__ __launch_bounds__(256) void k_cg_resid(int nf, const double* __restrict__ b,
                                                  double* __restrict__ q, double* __restrict__ r,
                                                  double* __restrict__ d, CgState* __restrict__ cs, int recompute,
                                                  const double* __restrict__ part, int nslice) {
#pragma clang fp contract(off)
  __shared__ double beta;
  if (cs->done) return;
  const bool dense = recompute < 0;
  if (dense) recompute = cs->iter % 50 == 0;
  if (part && recompute) form_q(nf, part, nslice, q);  // the recomputed A x of this iteration
  const double alpha = cs->alpha;
  for (int f = threadIdx.x; f < nf; f += blockDim.x) r[f] = recompute ? b[f] - q[f] : r[f] - alpha * q[f];
  __syncthreads();
  const double rr = dense ? dot_tree(nf, r, r) : dot_in_order(nf, r, r);
  if (threadIdx.x == 0) {
    const double dold = cs->delta_new;
    const double dnew = rr;
    cs->delta_old = dold;
    cs->delta_new = dnew;
    beta = dnew / dold;
    cs->beta = beta;
    cs->done = !(cs->iter < cs->max_iter && dnew > cs->eps);
  }
  __syncthreads();
  for (int f = threadIdx.x; f < nf; f += blockDim.x) d[f] = r[f] + beta * d[f];
}

// One CG iteration that does not recompute the residual, after q = A d
// (k_gemv_dense slices): k_cg_alpha and then k_cg_resid with r = r - alpha q,
// in one workgroup, with the same arithmetic in the same order, so the
// iterate is bit-identical to the two kernels'.  The dense form's graph runs
// it for 49 of every 50 iterations (iter % 50 != 0 after the increment);
// every 50th runs the two kernels with the recomputed residual.
static __global__ __launch_bounds__(256) void k_cg_iter(int nf, double* __restrict__ d, double* __restrict__ q,
                                                 double* __restrict__ x, double* __restrict__ r,
                                                 CgState* __restrict__ cs, const double* __restrict__ part,
                                                 int nslice) {
#pragma clang fp contract(off)
  __shared__ double alpha, beta;
  if (cs->done) return;
  form_q(nf, part, nslice, q);
  const double dq = dot_tree(nf, d, q);
  if (threadIdx.x == 0) {
    alpha = cs->delta_new / dq;
    cs->alpha = alpha;
    cs->iter = cs->iter + 1;
  }
  __syncthreads();
  for (int f = threadIdx.x; f < nf; f += blockDim.x) {
    x[f] = x[f] + alpha * d[f];
    r[f] = r[f] - alpha * q[f];
  }
  __syncthreads();
  const double rr = dot_tree(nf, r, r);
  if (threadIdx.x == 0) {
    const double dold = cs->delta_new;
    const double dnew = rr;
    cs->delta_old = dold;
    cs->delta_new = dnew;
    beta = dnew / dold;
    cs->beta = beta;
    cs->done = !(cs->iter < cs->max_iter && dnew > cs->eps);
  }
  __syncthreads();
  for (int f = threadIdx.x; f < nf; f += blockDim.x) d[f] = r[f] + beta * d[f];
}

// ---------------------------------------------------------------------------
// tail of the step: fluxes, circulation and diagnostics
// ---------------------------------------------------------------------------
// Vertex indices of the two streamfunction values whose difference is the
// curl at interior face f (u faces X(i,j+1) - X(i,j), v faces X(i,j) -
// X(i+1,j), as curl_face); false for a face on the domain edge, whose value
// takes the edge vector (curl_face).
__device__ __forceinline__ bool curl_interior(const Geo& g, int f, int& ia, int& ib) {
  const int m = g.m, n = g.n, nu = (m + 1) * n;
  if (f < nu) {
    const int i = f % (m + 1) + 1, j = f / (m + 1) + 1;
    ia = wix(g, i, j + 1);
    ib = wix(g, i, j);
    return i > 1 && i < m + 1 && j > 1 && j < n;
  }
  const int fv = f - nu;
  const int i = fv % m + 1, j = fv / m + 1;
  ia = wix(g, i, j);
  ib = wix(g, i + 1, j);
  return j > 1 && j < n + 1 && i > 1 && i < m;
}

// One launch after the last streamfunction of the step is final:
//  * workgroups x < ncurl = ceil((m+1)(n+1)/(256*kTailPos))*Lmax: q =
//    curl(s) on level x / ceil(...) (src/operators.f90:236-238, as
//    k_curl_levels).  A thread takes kTailPos places p = (i-1) + (j-1)(m+1),
//    i <= m+1, j <= n+1, and at each the u face (i, j) (j <= n) and the v
//    face (i, j) (i <= m): s(i, j), s(i, j+1) and s(i+1, j) serve both, so
//    a workgroup reads two rows of s once (u and v faces of one row used to
//    be far apart in the grid and read those rows twice).  Every interior
//    load is issued before the first is used; faces on the domain edge take
//    curl_face.  With s1d set, the streamfunctions of every level are s1d
//    (FP64).  These groups run in XCD-contiguous order, so the rows shared
//    by neighbouring groups stay in one L2;
//  * the next nchunk*Lmax: partial sums of omega for write_circ
//    (src/operators.f90:675-684), chunk (x-ncurl) % nchunk of level
//    (x-ncurl) / nchunk, at partial[((b*Lmax)+lev)*nchunk + chunk].
// With pend set, pend[b] = 1 and counter[b] += 1: the step's diagnostics
// row (slot counter[b] - 1) is left to the next launch that writes pending
// rows (diag_row).  With C set (same layout as s), the streamfunction of
// every level but a case's coarsest is s + C.
// grid (ncurl + nchunk*Lmax, B), 256 threads.
constexpr int kTailPos = 2;
__host__ __device__ __forceinline__ int tail_groups(const Geo& g) {
  return ((g.m + 1) * (g.n + 1) + 256 * kTailPos - 1) / (256 * kTailPos);
}
template <typename T, typename TS, typename TC = T>
__device__ __forceinline__ void tail_curl(const Geo& g, int p0, const TS* __restrict__ sl, const TC* __restrict__ sc,
                                          T* __restrict__ ql, const TS* __restrict__ cl = nullptr,
                                          const TC* __restrict__ ccs = nullptr) {
  const int m = g.m, n = g.n, nu = (m + 1) * n, np = (m + 1) * (n + 1);
  int ic[kTailPos], iu[kTailPos], ir[kTailPos];
  bool hu[kTailPos], hv[kTailPos], inu[kTailPos], inv[kTailPos];
#pragma unroll
  for (int u = 0; u < kTailPos; u++) {
    const int p = p0 + 256 * u, i = p % (m + 1) + 1, j = p / (m + 1) + 1;
    hu[u] = p < np && j <= n;
    hv[u] = p < np && i <= m;
    inu[u] = hu[u] && i > 1 && i < m + 1 && j > 1 && j < n;
    inv[u] = hv[u] && j > 1 && j < n + 1 && i > 1 && i < m;
    const bool c = inu[u] || inv[u];
    ic[u] = c ? wix(g, i, j) : 0;
    iu[u] = inu[u] ? wix(g, i, j + 1) : 0;
    ir[u] = inv[u] ? wix(g, i + 1, j) : 0;
  }
  TS vc[kTailPos], vu[kTailPos], vr[kTailPos];
#pragma unroll
  for (int u = 0; u < kTailPos; u++) {
    vc[u] = sl[ic[u]];
    vu[u] = sl[iu[u]];
    vr[u] = sl[ir[u]];
  }
  if (cl != nullptr) {  // the streamfunction is sl + cl: the parts added first
    TS c0[kTailPos], c1[kTailPos], c2[kTailPos];
#pragma unroll
    for (int u = 0; u < kTailPos; u++) {
      c0[u] = cl[ic[u]];
      c1[u] = cl[iu[u]];
      c2[u] = cl[ir[u]];
    }
#pragma unroll
    for (int u = 0; u < kTailPos; u++) {
      vc[u] = vc[u] + c0[u];
      vu[u] = vu[u] + c1[u];
      vr[u] = vr[u] + c2[u];
    }
  }
#pragma unroll
  for (int u = 0; u < kTailPos; u++) {
    const int p = p0 + 256 * u, i = p % (m + 1) + 1, j = p / (m + 1) + 1;
    if (hu[u])
      ql[p] = inu[u] ? T(vu[u] - vc[u])
                     : T(cl != nullptr || ccs != nullptr ? curl_face_sum<TS, TC>(g, p, sl, cl, sc, ccs)
                                                         : curl_face<TS, TC>(g, p, sl, sc));
    if (hv[u]) {
      const int fv = nu + (i - 1) + (j - 1) * m;
      ql[fv] = inv[u] ? T(vc[u] - vr[u])
                      : T(cl != nullptr || ccs != nullptr ? curl_face_sum<TS, TC>(g, fv, sl, cl, sc, ccs)
                                                          : curl_face<TS, TC>(g, fv, sl, sc));
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_tail(Geo g, const T* __restrict__ s, T* __restrict__ q,
                                              const T* __restrict__ om, long bs, long qbs,
                                              double* __restrict__ partial, int nchunk, const int* __restrict__ Lc,
                                              const double* __restrict__ s1d, int* __restrict__ pend,
                                              int* __restrict__ counter, const T* __restrict__ C) {
  __shared__ double red[4];
  const int tid = threadIdx.x, b = blockIdx.y;
  if (pend != nullptr && blockIdx.x == 0 && tid == 0) {  // the row is deferred to the next pass
    pend[b] = 1;
    counter[b] += 1;
  }
  const int ncb = tail_groups(g), ncurl = ncb * g.Lmax;
  if ((int)blockIdx.x < ncurl) {
    const int x = xcd_contiguous(blockIdx.x, ncurl);
    const int lev = x / ncb, f0 = (x - lev * ncb) * 256 * kTailPos + tid;
    if (!active(Lc, b, lev)) return;
    const bool hc = has_coarser(g, Lc, b, lev);
    if (s1d != nullptr) {  // every level solved in FP64: the differences in FP64, rounded once
      const double* sl = s1d + b * bs + (long)lev * g.nw;
      tail_curl<T, double, double>(g, f0, sl, hc ? sl + g.nw : nullptr, q + b * qbs + (long)lev * g.nq);
    } else {
      const T* sl = s + b * bs + (long)lev * g.nw;
      // C: the left/right terms of every level but a case's coarsest, kept
      // apart from s (the same sum s + C)
      const int L = levels(g, Lc, b);
      const T* cl = (C != nullptr && lev + 1 < L) ? C + b * bs + (long)lev * g.nw : nullptr;
      const T* ccs = (C != nullptr && lev + 2 < L) ? C + b * bs + (long)(lev + 1) * g.nw : nullptr;
      tail_curl<T, T>(g, f0, sl, hc ? sl + g.nw : nullptr, q + b * qbs + (long)lev * g.nq, cl, ccs);
    }
    return;
  }
  const int xc = blockIdx.x - ncurl, lev = xc / nchunk, ch = xc - lev * nchunk;
  if (!active(Lc, b, lev)) return;
  const long per = (g.nw + nchunk - 1) / nchunk;
  const long lo = ch * per, hi = min((long)g.nw, lo + per);
  const T* w = om + b * bs + (long)lev * g.nw;
  double acc = 0.0;
  for (long e = lo + tid; e < hi; e += 256) acc += double(w[e]);
  acc = block_sum(acc, red);
  if (tid == 0) partial[((long)b * g.Lmax + lev) * nchunk + ch] = acc;
}

// The diagnostics row of case b (see k_diag and DiagArgs), by one workgroup
// whose size is a multiple of 64 (red: LDS for one double per wave).  part
// -1 writes the whole row; 0, 1, 2 its circulation, slip and forces (three
// workgroups that write disjoint entries).
template <typename T>
__device__ void diag_row(const Geo& g, const DiagArgs<T>& a, int b, double* red, int part) {
  const Smear& sm = a.sm;
  const int tid = threadIdx.x, lane = tid & 63, NT = blockDim.x;
  const bool deferred = a.pend != nullptr;
  const int pending = deferred ? a.pend[b] : 1;  // tested after the first loads are issued
  const CaseSm c = case_sm(sm, b);
  const int L = levels(g, a.Lc, b);
  const int slot0 = deferred ? a.counter[b] - 1 : a.counter[b];
  const int slot = slot0 < a.capacity ? slot0 : a.capacity - 1;
  double* out = a.out_base + (long)slot * a.rowlen + c.hoff;
  // the force terms of this thread's first points, loaded here with the
  // partials and the slip tables rather than after the slip's reduction
  // (the first kDiagPts points of the thread, p = tid + i NT)
  constexpr int kDiagPts = 3;
  const int pbase = c.foff / 2;
  int bdp[kDiagPts];
  double fvp[kDiagPts][2];
#pragma unroll
  for (int i = 0; i < kDiagPts; i++) {
    const int p = tid + i * NT;
    bdp[i] = -1;
    fvp[i][0] = fvp[i][1] = 0.0;
    if ((part < 0 || part == 2) && p < c.nb) {
      bdp[i] = a.bdyno[pbase + p];
#pragma unroll
      for (int comp = 0; comp < 2; comp++) {
        const int f = comp * c.nb + p;
        fvp[i][comp] = a.wsum[c.foff + f] * (a.fb[c.foff + f] / c.dt);
      }
    }
  }
  if (!pending) return;
  // circulation: wave w sums the partials of level w
  for (int lv = tid >> 6; (part < 0 || part == 0) && lv < L; lv += NT >> 6) {
    const double* pp = a.partial + ((long)b * g.Lmax + lv) * a.nchunk;
    double acc = 0.0;
    for (int k = lane; k < a.nchunk; k += 64) acc += pp[k];
    acc = wave_sum(acc);
    if (lane == 0) out[2 * c.nbody + lv] = acc;
  }
  // slip, three forces per thread and pass
  if (part < 0 || part == 1) {
    const T* q1 = a.q + b * a.qbs;
    auto qv = [&](int face) { return double(q1[face]); };
    auto slip_at = [&](int f) -> double {
      if (f >= c.nf) return 0.0;
      double r;
      if (sm.f9_w != nullptr) {  // padded: the slots past a force's entries add +0
        const int base = c.foff + f;
        int id[kFrcK];
        double ww[kFrcK], v[kFrcK];
#pragma unroll
        for (int k = 0; k < kFrcK; k++) {
          id[k] = sm.f9_face[k * sm.f9_n + base];
          ww[k] = sm.f9_w[k * sm.f9_n + base];
        }
#pragma unroll
        for (int k = 0; k < kFrcK; k++) v[k] = qv(id[k]);
        r = 0.0;
#pragma unroll
        for (int k = 0; k < kFrcK; k++) r = r + ww[k] * v[k];
      } else {
        r = gather_sum<kFrcK>(sm.frc_ptr[c.rbase + f], sm.frc_ptr[c.rbase + f + 1], sm.frc_face, sm.frc_w, qv);
      }
      return fabs(r - a.motion[c.foff + f]);
    };
    double mx = 0.0;
    for (int f0 = tid; f0 < c.nf; f0 += 3 * NT) {
      const double e0 = slip_at(f0), e1 = slip_at(f0 + NT), e2 = slip_at(f0 + 2 * NT);
      mx = fmax(mx, fmax(e0, fmax(e1, e2)));
    }
    mx = block_max(mx, red);
    if (tid == 0) out[2 * c.nbody + L] = mx;
  }
  // forces, one body and component at a time
  for (int bd = 0; (part < 0 || part == 2) && bd < c.nbody; bd++)
    for (int comp = 0; comp < 2; comp++) {
      double acc = 0.0;
#pragma unroll
      for (int i = 0; i < kDiagPts; i++)
        if (bdp[i] == c.bmin + bd) acc += fvp[i][comp];
      for (int p = tid + kDiagPts * NT; p < c.nb; p += NT)
        if (a.bdyno[pbase + p] == c.bmin + bd) {
          const int f = comp * c.nb + p;
          acc += a.wsum[c.foff + f] * (a.fb[c.foff + f] / c.dt);
        }
      acc = block_sum(acc, red);
      if (tid == 0) out[comp * c.nbody + bd] = acc * 2.0 * c.delta;
    }
  if (tid == 0) {
    if (!deferred) a.counter[b] = slot0 + 1;
    else if (a.clear) a.pend[b] = 0;  // every thread read pend[b] before the reductions' barriers
  }
}

// One workgroup per case: the diagnostics row of the step,
// [fx(bmin..bmax), fy(bmin..bmax), circ(1..L), slip] at
// out_base + slot*rowlen + hoff, slot = counter[b] (then incremented):
//   force: write_force (src/operators.f90:583-636) = 2*delta * sum over
//          faces of reg(fb/dt restricted to the body), evaluated as
//          2*delta * sum_p wsum(p) fb(p)/dt with wsum(p) the sum of the
//          point's weights on u (x) or v (y) faces;
//   circ:  the sum of the level's partial sums from k_tail;
//   slip:  write_slip (src/operators.f90:564) = max |regT(q_1) - motion|.
template <typename T>
__global__ __launch_bounds__(1024) void k_diag(Geo g, DiagArgs<T> a) {
  __shared__ double red[16];
  diag_row<T>(g, a, blockIdx.x, red, -1);
}

}  // namespace ibfs
