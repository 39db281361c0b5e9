// ibfs_kernels.hpp -- device side of one multi-domain IBFS time step.
//
// Index conventions (reference src/grid.f90:119-134, src/variables.f90:46):
//   vertex field  w[ii + jj*(m-1)], ii = i-2, jj = j-2 for the reference's
//                 interior vertex (i,j), i = 2..m, j = 2..n;
//   flux faces    u(i,j) -> (i-1) + (j-1)(m+1), i = 1..m+1, j = 1..n;
//                 v(i,j) -> (m+1)n + (i-1) + (j-1)m, i = 1..m, j = 1..n+1;
//   edge vector   the reference's 1-based position pos lives at pos-1, with
//                 segment bases left = 0, right = n+1, bottom = 2(n+1),
//                 top = 2(n+1)+m+1 (src/grid.f90:73-76).
// Levels are stored one after another (level k at offset (k-1)*size).
#pragma once
#include <hip/hip_runtime.h>

#include "dst_core.hpp"

namespace ibfs {

struct Geo {
  int m, n, L;
  int mm1, nm1;   // m-1, n-1
  int nw, nq, nbc;
  int left, right, bottom, top;
};

__device__ __forceinline__ int uix(const Geo& g, int i, int j) { return (i - 1) + (j - 1) * (g.m + 1); }
__device__ __forceinline__ int vix(const Geo& g, int i, int j) { return (g.m + 1) * g.n + (i - 1) + (j - 1) * g.m; }
__device__ __forceinline__ int wix(const Geo& g, int i, int j) { return (i - 2) + (j - 2) * g.mm1; }

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
// edge vector given as a get_bc of the coarser field rc (rc == nullptr: 0).
// Reference order: bottom/top first (i loop), then left/right (j loop).
template <typename T>
__device__ __forceinline__ T applybc_add(const Geo& g, T acc, int i, int j, const T* __restrict__ rc, T gfac, T fac) {
  if (rc == nullptr) return acc;
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

// rot(nonlinear(omega, q, q0, bc)) at vertex (i,j):
// nonlinear src/operators.f90:465-510, rot src/operators.f90:291-304.
template <typename T>
__device__ __forceinline__ T rot_nonlinear(const Geo& g, int i, int j, const T* __restrict__ om,
                                           const T* __restrict__ q, const T* __restrict__ q0,
                                           const T* __restrict__ bc) {
  const int m = g.m, n = g.n;
  // vorticity with the edge values of bc outside the interior
  auto OmU = [&](int ii, int jj) -> T {  // along j at column ii (2..m)
    if (jj == 1) return bc[g.bottom + ii - 1];
    if (jj == n + 1) return bc[g.top + ii - 1];
    return om[wix(g, ii, jj)];
  };
  auto OmV = [&](int ii, int jj) -> T {  // along i at row jj (2..n)
    if (ii == 1) return bc[g.left + jj - 1];
    if (ii == m + 1) return bc[g.right + jj - 1];
    return om[wix(g, ii, jj)];
  };
  auto vavg = [&](int ii, int jj) -> T {
    const int a = vix(g, ii, jj), b = vix(g, ii - 1, jj);
    return T(0.5) * (q[a] + q[b] + q0[a] + q0[b]);
  };
  auto uavg = [&](int ii, int jj) -> T {
    const int a = uix(g, ii, jj), b = uix(g, ii, jj - 1);
    return T(0.5) * (q[a] + q[b] + q0[a] + q0[b]);
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
// pass operators
// ---------------------------------------------------------------------------
// plain field source/destination (batch stride bs)
template <typename T>
struct SrcField {
  const T* a; long bs;
  __device__ T operator()(const Geo& g, int b, int ii, int jj) const { return a[b * bs + ii + jj * g.mm1]; }
};

// omega_k + apply_bc(get_bc(s_{k+1}, 1), 1): the vort of vort2flux
// (src/operators.f90:230-234).  sc == nullptr on the coarsest level.
template <typename T>
struct SrcVortBC {
  const T* om; const T* sc; long bs;
  __device__ T operator()(const Geo& g, int b, int ii, int jj) const {
    T v = om[b * bs + ii + jj * g.mm1];
    if (sc == nullptr) return v;
    const int i = ii + 2, j = jj + 2;
    if (i == 2 || i == g.m || j == 2 || j == g.n) v = applybc_add<T>(g, v, i, j, sc + b * bs, T(1), T(1));
    return v;
  }
};

template <typename T>
struct DstField {
  T* a; long bs;
  __device__ void operator()(const Geo& g, int b, int ii, int jj, T v) const { a[b * bs + ii + jj * g.mm1] = v; }
};
// a = tab * v (ctci: laminv * in, src/myfft.f90:106)
template <typename T>
struct DstScale {
  T* a; const T* tab; long bs;
  __device__ void operator()(const Geo& g, int b, int ii, int jj, T v) const {
    const int e = ii + jj * g.mm1;
    a[b * bs + e] = tab[e] * v;
  }
};
// a = tab * v / norm (ainv: lam1i(:,:,1) * in / normalize, src/myfft.f90:121)
template <typename T>
struct DstScaleDiv {
  T* a; const T* tab; T norm; long bs;
  __device__ void operator()(const Geo& g, int b, int ii, int jj, T v) const {
    const int e = ii + jj * g.mm1;
    a[b * bs + e] = tab[e] * v / norm;
  }
};
// a = a - v (projection, src/operators.f90:174)
template <typename T>
struct DstSub {
  T* a; long bs;
  __device__ void operator()(const Geo& g, int b, int ii, int jj, T v) const {
    const int e = b * bs + ii + jj * g.mm1;
    a[e] = a[e] - v;
  }
};

// single-field row pass: rows jj; FFT p pairs rows 2Pg+p and 2Pg+P+p
template <typename T, class Src, class Dst>
struct RowOp {
  Geo g; Src src; Dst dst;
  __device__ T load(int b, int gr, int s, int P, int idx) const {
    const int row = gr * 2 * P + s;
    return row < g.nm1 ? src(g, b, idx, row) : T(0);
  }
  __device__ void store(int b, int gr, int p, int P, int k, T ya, T yb) const {
    const int r0 = gr * 2 * P + p, r1 = r0 + P;
    if (r0 < g.nm1) dst(g, b, k - 1, r0, ya);
    if (r1 < g.nm1) dst(g, b, k - 1, r1, yb);
  }
};
// single-field column pass: columns ii
template <typename T, class Src, class Dst>
struct ColOp {
  Geo g; Src src; Dst dst;
  __device__ T load(int b, int gr, int s, int P, int idx) const {
    const int col = gr * 2 * P + s;
    return col < g.mm1 ? src(g, b, col, idx) : T(0);
  }
  __device__ void store(int b, int gr, int p, int P, int k, T ya, T yb) const {
    const int c0 = gr * 2 * P + p, c1 = c0 + P;
    if (c0 < g.mm1) dst(g, b, c0, k - 1, ya);
    if (c1 < g.mm1) dst(g, b, c1, k - 1, yb);
  }
};

// Step 1, first pass at level k: X = con1*rhs + con2*rhs_old + rhsbc with
// rhs = rot(nonlinear(omega_k, q_k, q0_k, lastbc_k)) and Y = omega_k / (4mn),
// rows of both transformed together (FFT p = X row + i Y row).
// rhs_old_k <- rhs.  src/operators.f90:110-133.
// Y is pre-divided by the normalisation 4mn so that both halves of a packed
// FFT have the scale of the final result: an FFT's rounding is relative to
// the larger half, and packing omega at full scale against X (which carries
// the 1/(4mn) of con1/con2) would cost a factor 4mn in the accuracy of S(X).
template <typename T>
struct Step1RowOp {
  Geo g;
  const T* om;      // omega_k (old)
  const T* q; const T* q0;
  const T* lastbc;  // edge vector of level k (get_bc of old omega_{k+1}, x0.25)
  const T* omc;     // new omega_{k+1} (nullptr on the coarsest level)
  T* rhs_old;
  T* tx; T* ty;     // row-transformed X and Y
  T con1, con2, vfac, norm;
  int first;        // itime == 1: explicit Euler start
  __device__ T load(int b, int gr, int s, int P, int idx) const {
    const int row = gr * P + (s < P ? s : s - P);
    if (row >= g.nm1) return T(0);
    const int e = idx + row * g.mm1;
    if (s >= P) return om[e] / norm;
    const int i = idx + 2, j = row + 2;
    const T rhs = rot_nonlinear<T>(g, i, j, om, q, q0, lastbc);
    const T ro = first ? rhs : rhs_old[e];
    rhs_old[e] = rhs;
    T bcs = T(0);
    if (i == 2 || i == g.m || j == 2 || j == g.n) {
      bcs = applybc_vec<T>(g, bcs, i, j, lastbc, vfac);
      bcs = applybc_add<T>(g, bcs, i, j, omc, T(0.25), vfac);
    }
    return con1 * rhs + con2 * ro + bcs;
  }
  __device__ void store(int b, int gr, int p, int P, int k, T ya, T yb) const {
    const int row = gr * P + p;
    if (row >= g.nm1) return;
    const int e = (k - 1) + row * g.mm1;
    tx[e] = ya;
    ty[e] = yb;
  }
};

// Step 1, second pass: columns of X~ and Y~ together, then
// Z = lam1i_k * (X^ + lam1_k * S(omega)) (src/operators.f90:126-130), where
// lam1_k * S(omega) = lam1n_k * S(omega / 4mn) with lam1n = lam1 * 4mn.
template <typename T>
struct Step1ColOp {
  Geo g;
  const T* tx; const T* ty;
  const T* lam1n; const T* lam1i;  // level-k tables (lam1n = lam1 * 4mn)
  T* z;
  __device__ T load(int b, int gr, int s, int P, int idx) const {
    const int col = gr * P + (s < P ? s : s - P);
    if (col >= g.mm1) return T(0);
    const int e = col + idx * g.mm1;
    return s < P ? tx[e] : ty[e];
  }
  __device__ void store(int b, int gr, int p, int P, int k, T ya, T yb) const {
    const int col = gr * P + p;
    if (col >= g.mm1) return;
    const int e = col + (k - 1) * g.mm1;
    z[e] = lam1i[e] * (ya + lam1n[e] * yb);
  }
};

// ---------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------
// lastbc(:,k) = get_bc(omega(:,:,k+1), 0.25) for k < L; 0 for k = L
// (src/operators.f90:97-100).  grid: ceil(nbc*L/256)
template <typename T>
__global__ void k_lastbc(Geo g, const T* __restrict__ om, T* __restrict__ lastbc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.nbc * g.L) return;
  const int lev = t / g.nbc, p = t - lev * g.nbc;
  lastbc[t] = (lev < g.L - 1) ? getbc_at<T>(g, om + (long)(lev + 1) * g.nw, p + 1, T(0.25)) : T(0);
}

// coarsify (src/operators.f90:307-333), in place on level k from level k-1.
// grid: ceil(((m/2-1)(n/2-1))/256) x batch
template <typename T>
__global__ void k_coarsify(Geo g, T* __restrict__ coarse, const T* __restrict__ fine, long bs) {
  const int wi = g.m / 2 - 1, wj = g.n / 2 - 1;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= wi * wj) return;
  const int b = blockIdx.y;
  const int i = g.m / 4 + 2 + t % wi, j = g.n / 4 + 2 + t / wi;
  const int ii = 2 * i - g.m / 2 - 1, jj = 2 * j - g.n / 2 - 1;
  const T* r = fine + b * bs;
  auto F = [&](int a, int c) -> T { return r[wix(g, a, c)]; };
  coarse[b * bs + wix(g, i, j)] =
      F(ii, jj) + T(0.5) * (F(ii - 1, jj) + F(ii + 1, jj) + F(ii, jj - 1) + F(ii, jj + 1)) +
      T(0.25) * (F(ii - 1, jj - 1) + F(ii - 1, jj + 1) + F(ii + 1, jj - 1) + F(ii + 1, jj + 1));
}

// curl face value (src/operators.f90:243-288) for 0-based face f; the edge
// vector is get_bc(sc, 1) (sc == nullptr: zero edge values).
template <typename T>
__device__ __forceinline__ T curl_face(const Geo& g, int f, const T* __restrict__ s, const T* __restrict__ sc) {
  const int m = g.m, n = g.n;
  auto X = [&](int i, int j) -> T { return s[wix(g, i, j)]; };
  auto B = [&](int pos) -> T { return sc ? getbc_at<T>(g, sc, pos, T(1)) : T(0); };
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

// vel = curl(s, get_bc(sc)) over all faces.  grid: ceil(nq/256) x batch
template <typename T>
__global__ void k_curl(Geo g, const T* __restrict__ s, const T* __restrict__ sc, T* __restrict__ vel, long bs,
                       long vbs) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= g.nq) return;
  const int b = blockIdx.y;
  vel[b * vbs + f] = curl_face<T>(g, f, s + b * bs, sc ? sc + b * bs : nullptr);
}

// rot(x): Nq -> W (src/operators.f90:291-304).  grid: ceil(nw/256)
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
// smearing (E / E^T) on compact tables
// ---------------------------------------------------------------------------
// Active faces: face_id[a] (0-based face), entries face_ptr[a]..face_ptr[a+1]
// of (face_pt, face_w): the force index (0-based, x or y component) and the
// weight, in increasing IB-point order (the ELL order of src/grid.f90:197-228).
// Force gather: for force f, frc_ptr[f]..frc_ptr[f+1] of (frc_face, frc_w) in
// increasing face order (the accumulation order of regT, src/operators.f90:541-545).
// Active vertices: vtx_id[v] (0-based W index) and vtx_face[4v..4v+3] = active
// face slots of v(i,j), v(i-1,j), u(i,j), u(i,j-1) (-1: face has no weight).
struct Smear {
  int nact;            // active faces
  const int* face_id;
  const int* face_ptr;
  const int* face_pt;
  const double* face_w;
  int nf;
  const int* frc_ptr;
  const int* frc_face;
  const double* frc_w;
  int nvtx;
  const int* vtx_id;
  const int* vtx_face;
};

template <typename T>
__device__ __forceinline__ T reg_face(const Smear& sm, int a, const T* __restrict__ h0) {
  T s = T(0);
  for (int e = sm.face_ptr[a]; e < sm.face_ptr[a + 1]; e++) s = s + T(sm.face_w[e]) * h0[sm.face_pt[e]];
  return s;
}

// rr = rot(reg(h0)) on the active vertices only; all other entries of rr are
// never written and stay zero.  grid: ceil(nvtx/256) x batch
template <typename T>
__global__ void k_rotreg(Smear sm, const T* __restrict__ h0, long hbs, T* __restrict__ rr, long bs) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= sm.nvtx) return;
  const int b = blockIdx.y;
  const T* h = h0 + b * hbs;
  T fv[4];
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int a = sm.vtx_face[4 * v + c];
    fv[c] = a >= 0 ? reg_face<T>(sm, a, h) : T(0);
  }
  rr[b * bs + sm.vtx_id[v]] = fv[0] - fv[1] - fv[2] + fv[3];
}

// h = reg(h0) on all faces (operator test entry): zero, then active faces
template <typename T>
__global__ void k_reg_active(Smear sm, const T* __restrict__ h0, T* __restrict__ h) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= sm.nact) return;
  h[sm.face_id[a]] = reg_face<T>(sm, a, h0);
}

// h0 = regT(h)  (grid: ceil(nf/256))
template <typename T>
__global__ void k_regT(Smear sm, const T* __restrict__ h, T* __restrict__ h0) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= sm.nf) return;
  T s = T(0);
  for (int e = sm.frc_ptr[f]; e < sm.frc_ptr[f + 1]; e++) s = s + T(sm.frc_w[e]) * h[sm.frc_face[e]];
  h0[f] = s;
}

// out[f] = regT(curl(s1, get_bc(s2)))[f] - sub[f]  (sub may be null), batched:
// the right-hand side of the force system (src/operators.f90:163-164) and
// the columns of a_times (src/operators.f90:459-461).
template <typename T>
__global__ void k_regT_curl(Geo g, Smear sm, const T* __restrict__ s1, const T* __restrict__ s2, long bs,
                            const double* __restrict__ sub, double* __restrict__ out, long obs) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= sm.nf) return;
  const int b = blockIdx.y;
  T s = T(0);
  for (int e = sm.frc_ptr[f]; e < sm.frc_ptr[f + 1]; e++)
    s = s + T(sm.frc_w[e]) * curl_face<T>(g, sm.frc_face[e], s1 + b * bs, s2 ? s2 + b * bs : nullptr);
  double r = double(s);
  if (sub) r = r - sub[f];
  out[b * obs + f] = r;
}

// motion_eff = motion_in - regT(q0_1) (src/operators.f90:155)
template <typename T>
__global__ void k_motion(Smear sm, const T* __restrict__ q0, const double* __restrict__ motion_in,
                         double* __restrict__ motion) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= sm.nf) return;
  T s = T(0);
  for (int e = sm.frc_ptr[f]; e < sm.frc_ptr[f + 1]; e++) s = s + T(sm.frc_w[e]) * q0[sm.frc_face[e]];
  motion[f] = (motion_in ? motion_in[f] : 0.0) - double(s);
}

// ---------------------------------------------------------------------------
// force solve: fb = cholsl(rhsf) as two triangular products with the
// explicit inverse Li = L^{-1} of the Cholesky factor (column-major, lower).
// y = Li r  then  x = Li^T y.  One wave per 64 output rows.
// ---------------------------------------------------------------------------
// y[i] = sum_{k<=i} Li(i,k) r[k]      grid: ceil(nf/4) blocks of 256 (4 rows per block, one wave each)
__global__ void k_trmv_lower(int nf, const double* __restrict__ Li, const double* __restrict__ r,
                             double* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= nf) return;
  double s = 0.0;
  for (int k = lane; k <= i; k += 64) s += Li[i + (long)k * nf] * r[k];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) y[i] = s;
}
// x[i] = sum_{k>=i} Li(k,i) y[k]   (column i of Li, contiguous)
__global__ void k_trmv_upper_t(int nf, const double* __restrict__ Li, const double* __restrict__ y,
                               double* __restrict__ x) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= nf) return;
  double s = 0.0;
  const double* col = Li + (long)i * nf;
  for (int k = i + lane; k < nf; k += 64) s += col[k] * y[k];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) x[i] = s;
}

// ---------------------------------------------------------------------------
// diagnostics
// ---------------------------------------------------------------------------
// per-level partial sums of omega (write_circ, src/operators.f90:675-684):
// grid (nchunk, L); partial[lev*nchunk + c]
template <typename T>
__global__ void k_circ_partial(Geo g, const T* __restrict__ om, double* __restrict__ partial, int nchunk) {
  __shared__ double red[kThreads];
  const int lev = blockIdx.y, c = blockIdx.x;
  const long per = (g.nw + nchunk - 1) / nchunk;
  const long lo = c * per, hi = min((long)g.nw, lo + per);
  double s = 0.0;
  const T* w = om + (long)lev * g.nw;
  for (long e = lo + threadIdx.x; e < hi; e += kThreads) s += double(w[e]);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = kThreads / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[lev * nchunk + c] = red[0];
}

// One workgroup: slip, body forces and circulation of this step.
// out = [fx(bmin..bmax), fy(bmin..bmax), circ(1..L), slip]
// force: write_force (src/operators.f90:583-636) as 2*delta * sum over the
// faces of reg(fb/dt restricted to the body); slip: write_slip
// (src/operators.f90:564) = max |regT(q_1) - motion|.
template <typename T>
__global__ __launch_bounds__(1024) void k_diag(Geo g, Smear sm, const T* __restrict__ q1,
                                               const double* __restrict__ motion, const double* __restrict__ fb,
                                               const int* __restrict__ bdyno, int nb, int bmin, int nbody,
                                               const double* __restrict__ partial, int nchunk, double dt,
                                               double delta, double* __restrict__ out_base,
                                               int* __restrict__ counter, int capacity) {
  __shared__ double red[1024];
  const int tid = threadIdx.x;
  const int slot = counter ? *counter : 0;
  double* out = out_base + (long)(slot < capacity ? slot : capacity - 1) * (2 * nbody + g.L + 1);
  // slip
  double mx = 0.0;
  for (int f = tid; f < sm.nf; f += 1024) {
    T s = T(0);
    for (int e = sm.frc_ptr[f]; e < sm.frc_ptr[f + 1]; e++) s = s + T(sm.frc_w[e]) * q1[sm.frc_face[e]];
    mx = fmax(mx, fabs(double(s) - motion[f]));
  }
  red[tid] = mx;
  __syncthreads();
  for (int off = 512; off > 0; off >>= 1) {
    if (tid < off) red[tid] = fmax(red[tid], red[tid + off]);
    __syncthreads();
  }
  if (tid == 0) out[2 * nbody + g.L] = red[0];
  __syncthreads();
  // forces per body and component
  const int nu = (g.m + 1) * g.n;
  for (int bd = 0; bd < nbody; bd++) {
    for (int comp = 0; comp < 2; comp++) {
      double s = 0.0;
      for (int a = tid; a < sm.nact; a += 1024) {
        const bool isu = sm.face_id[a] < nu;
        if (isu != (comp == 0)) continue;
        double fv = 0.0;
        for (int e = sm.face_ptr[a]; e < sm.face_ptr[a + 1]; e++) {
          const int f = sm.face_pt[e];
          const int pt = f < nb ? f : f - nb;
          if (bdyno[pt] == bmin + bd) fv = fv + sm.face_w[e] * (fb[f] / dt);
        }
        s += fv;
      }
      red[tid] = s;
      __syncthreads();
      for (int off = 512; off > 0; off >>= 1) {
        if (tid < off) red[tid] += red[tid + off];
        __syncthreads();
      }
      if (tid == 0) out[comp * nbody + bd] = red[0] * 2.0 * delta;
      __syncthreads();
    }
  }
  // circulation
  if (tid < g.L) {
    double s = 0.0;
    for (int c = 0; c < nchunk; c++) s += partial[tid * nchunk + c];
    out[2 * nbody + tid] = s;
  }
  __syncthreads();
  if (tid == 0 && counter) *counter = slot + 1;
}

}  // namespace ibfs
