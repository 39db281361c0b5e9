// ibfs_adjoint.hpp -- device kernels of the discrete adjoint of one time step.
//
// The step (reference src/operators.f90:62-188) is linear in the state except
// for the bilinear nonlinear term; its adjoint applies the transpose of every
// operator in reverse order.  Transposes used here (all exact, gather form,
// no atomics):
//   S^T = S (2-D DST-I), ctci and ainv are symmetric, cholsl is symmetric,
//   rot^T = curl(., 0), curl_s^T = rot (including the edge rows),
//   (regT)^T = reg,
//   get_bc^T    : edge vector -> the four box lines of the coarser level,
//   apply_bc^T  : first/last interior rows/columns -> edge vector,
//   curl_bc^T   : face adjoint -> edge vector,
//   coarsify^T  : centre of the coarse level -> full-weighting scatter onto
//                 the finer level, and the overwritten centre gets no adjoint,
//   nonlinear^T : for N = rot(nl(Omega, qt)) with qt = q + q0, the
//                 transposes of the two bilinear slots (Omega and its edge
//                 values, and qt).
#pragma once
#include "ibfs_kernels.hpp"

namespace ibfs {

// get_bc^T: contribution of the edge vector abc (of the finer level) to the
// coarse vertex (i,j), times fac.  getbc_at reads R1(a,b) = w(a+1,b+1).
template <typename T>
__device__ __forceinline__ T gt_at(const Geo& g, const T* __restrict__ abc, int i, int j, T fac) {
  const int a = i - 1, b = j - 1, m = g.m, n = g.n;
  T acc = T(0);
  for (int seg = 0; seg < 2; seg++) {  // bottom (b == n/4), top (b == 3n/4): along a
    const int jb = seg == 0 ? n / 4 : 3 * n / 4;
    const int base = seg == 0 ? g.bottom : g.top;
    if (b == jb && a >= m / 4 && a <= 3 * m / 4) {
      const int ie = 2 * (a - m / 4);
      acc += abc[base + ie];
      if (ie - 1 >= 1) acc += T(0.5) * abc[base + ie - 1];
      if (ie + 1 <= m - 1) acc += T(0.5) * abc[base + ie + 1];
    }
  }
  for (int seg = 0; seg < 2; seg++) {  // left (a == m/4), right (a == 3m/4): along b
    const int ib = seg == 0 ? m / 4 : 3 * m / 4;
    const int base = seg == 0 ? g.left : g.right;
    if (a == ib && b >= n / 4 && b <= 3 * n / 4) {
      const int je = 2 * (b - n / 4);
      acc += abc[base + je];
      if (je - 1 >= 1) acc += T(0.5) * abc[base + je - 1];
      if (je + 1 <= n - 1) acc += T(0.5) * abc[base + je + 1];
    }
  }
  return acc * fac;
}

__device__ __forceinline__ bool on_box_lines(const Geo& g, int i, int j) {
  const int a = i - 1, b = j - 1;
  const bool ab = a >= g.m / 4 && a <= 3 * g.m / 4, bb = b >= g.n / 4 && b <= 3 * g.n / 4;
  return (ab && (b == g.n / 4 || b == 3 * g.n / 4)) || (bb && (a == g.m / 4 || a == 3 * g.m / 4));
}

// curl_bc^T (aq) + fac * apply_bc^T (t) at 0-based edge index p (1-based
// position p+1).  aq or t may be null.
template <typename T>
__device__ __forceinline__ T bc_adj_at(const Geo& g, int p, const T* __restrict__ aq, const T* __restrict__ t,
                                       T fac) {
  const int m = g.m, n = g.n, pos = p + 1;
  auto U = [&](int i, int j) -> T { return aq[uix(g, i, j)]; };
  auto V = [&](int i, int j) -> T { return aq[vix(g, i, j)]; };
  auto W = [&](int i, int j) -> T { return t[wix(g, i, j)]; };
  T r = T(0);
  if (pos > g.bottom) {
    const bool top = pos > g.top;
    const int ii = pos - (top ? g.top : g.bottom);  // 1..m+1
    if (aq) {
      const int jv = top ? n + 1 : 1;
      if (ii <= m) r += V(ii, jv);
      if (ii >= 2) r -= V(ii - 1, jv);
      if (ii >= 2 && ii <= m) r += top ? U(ii, n) : -U(ii, 1);
    }
    if (t && ii >= 2 && ii <= m) r += fac * W(ii, top ? n : 2);
  } else {
    const bool right = pos > g.right;
    const int jj = pos - (right ? g.right : g.left);  // 1..n+1
    if (aq) {
      const int iu = right ? m + 1 : 1;
      if (jj >= 2) r += U(iu, jj - 1);
      if (jj <= n) r -= U(iu, jj);
      if (jj >= 2 && jj <= n) r += right ? -V(m, jj) : V(1, jj);
    }
    if (t && jj >= 2 && jj <= n) r += fac * W(right ? m : 2, jj);
  }
  return r;
}

// abc = curl_bc^T(aq) + fac * apply_bc^T(t); per case (grid.y), strides bs/qs
template <typename T>
__global__ void k_bc_adj(Geo g, const T* __restrict__ aq, long qs, const T* __restrict__ t, long ts, T fac,
                         T* __restrict__ abc, long cs, const int* __restrict__ Lc, int lev) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (p >= g.nbc || !active(Lc, b, lev)) return;
  abc[b * cs + p] = bc_adj_at<T>(g, p, aq ? aq + b * qs : nullptr, t ? t + b * ts : nullptr, fac);
}

// w[box lines] += fac * get_bc^T(abc) on the coarse level (one thread per
// box-line vertex: 2 lines of m/2+1 and 2 of n/2+1 vertices)
template <typename T>
__global__ void k_gt_add(Geo g, const T* __restrict__ abc, long cs, T fac, T* __restrict__ w, long ws,
                         const int* __restrict__ Lc, int coarse_lev) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  const int nx = g.m / 2 + 1, ny = g.n / 2 + 1;
  if (t >= 2 * nx + 2 * ny || !active(Lc, b, coarse_lev)) return;
  int a, bb;  // R1 indices
  if (t < nx) { a = g.m / 4 + t; bb = g.n / 4; }
  else if (t < 2 * nx) { a = g.m / 4 + t - nx; bb = 3 * g.n / 4; }
  else if (t < 2 * nx + ny) { a = g.m / 4; bb = g.n / 4 + t - 2 * nx; if (bb == g.n / 4 || bb == 3 * g.n / 4) return; }
  else { a = 3 * g.m / 4; bb = g.n / 4 + t - 2 * nx - ny; if (bb == g.n / 4 || bb == 3 * g.n / 4) return; }
  const int i = a + 1, j = bb + 1;
  T* wp = w + b * ws + wix(g, i, j);
  *wp = *wp + gt_at<T>(g, abc + b * cs, i, j, fac);
}

// coarsify^T: fine[ii,jj] += sum over coarse centre vertices whose 3x3
// stencil covers (ii,jj) of weight * coarse; one thread per fine vertex of
// the covered region ii in [3-1, m-1+1].
template <typename T>
__global__ void k_coarsify_adj(Geo g, const T* __restrict__ coarse, T* __restrict__ fine, long bs,
                               const int* __restrict__ Lc, int coarse_lev) {
  const int wi = g.m - 1, t = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (t >= wi * (g.n - 1) || !active(Lc, b, coarse_lev)) return;
  const int ii = t % wi + 2, jj = t / wi + 2;
  const int i_lo = g.m / 4 + 2, i_hi = 3 * g.m / 4, j_lo = g.n / 4 + 2, j_hi = 3 * g.n / 4;
  const T* c = coarse + b * bs;
  T acc = T(0);
  bool any = false;
  for (int di = -1; di <= 1; di++) {
    const int s = ii + g.m / 2 + 1 - di;
    if (s & 1) continue;
    const int i = s / 2;
    if (i < i_lo || i > i_hi) continue;
    const T wi_ = di == 0 ? T(1) : T(0.5);
    for (int dj = -1; dj <= 1; dj++) {
      const int r = jj + g.n / 2 + 1 - dj;
      if (r & 1) continue;
      const int j = r / 2;
      if (j < j_lo || j > j_hi) continue;
      const T wj = dj == 0 ? T(1) : T(0.5);
      acc += wi_ * wj * c[wix(g, i, j)];
      any = true;
    }
  }
  if (any) {
    T* f = fine + b * bs + wix(g, ii, jj);
    *f = *f + acc;
  }
}

__device__ __forceinline__ bool in_centre(const Geo& g, int i, int j) {
  return i >= g.m / 4 + 2 && i <= 3 * g.m / 4 && j >= g.n / 4 + 2 && j <= 3 * g.n / 4;
}

// --------------------------------------------------------------- functors
// rot(aq) + get_bc^T(abc_fine): the adjoint streamfunction source of one
// level in the reverse multi-domain Poisson chain (aq, abc may be null)
template <typename T>
struct SrcAdjS {
  const T* aq; long qs; const T* abc; long cs;
  __device__ T operator()(const Geo& g, int b, int ii, int jj) const {
    const int i = ii + 2, j = jj + 2;
    T v = T(0);
    if (aq) {
      const T* x = aq + b * qs;
      v = x[vix(g, i, j)] - x[vix(g, i - 1, j)] - x[uix(g, i, j)] + x[uix(g, i, j - 1)];
    }
    if (abc && on_box_lines(g, i, j)) v += gt_at<T>(g, abc + b * cs, i, j, T(1));
    return v;
  }
};
// a field read with its coarsify centre treated as zero (adjoint of the
// overwritten values); zero_centre false on the finest level
template <typename T>
struct SrcNoCentre {
  const T* a; long bs; int zero_centre;
  __device__ T operator()(const Geo& g, int b, int ii, int jj) const {
    if (zero_centre && in_centre(g, ii + 2, jj + 2)) return T(0);
    return a[b * bs + ii + jj * g.mm1];
  }
};
// Transpose of step 1 in increment form.  The forward step is
// omega* = omega + S lam1i S (X + 2 vfac D0 omega) (D0: 5-point Laplacian
// with zero edge values, symmetric; S and diag(lam1i) symmetric), so with
// aX = S lam1i S aw* (the adjoint of X, one transform pair):
// aw_old += aw* + 2 vfac D0 aX.  grid.x over nw, grid.y = case
template <typename T>
__global__ void k_step1_adj_incr(Geo g, const T* __restrict__ aws, long fbs, const T* __restrict__ aX,
                                 T* __restrict__ awo, const T* __restrict__ vfac, const int* __restrict__ Lc,
                                 int lev) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (e >= g.nw || !active(Lc, b, lev)) return;
  const int ii = e % g.mm1, jj = e / g.mm1;
  const T* x = aX + (long)b * g.nw;
  const T lap = (ii > 0 ? x[e - 1] : T(0)) + (ii + 1 < g.mm1 ? x[e + 1] : T(0)) + (jj > 0 ? x[e - g.mm1] : T(0)) +
                (jj + 1 < g.nm1 ? x[e + g.mm1] : T(0)) - T(4) * x[e];
  T* d = awo + b * fbs + e;
  *d = *d + aws[b * fbs + e] + T(2) * vfac[b * g.Lmax + lev] * lap;
}

// ------------------------------------------------------ nonlinear adjoint
// Face adjoint of N = rot(fq): afq = curl(aN, 0) on the faces the forward
// computes (u: i = 2..m; v: j = 2..n).
template <typename T>
__device__ __forceinline__ T afq_u(const Geo& g, const T* __restrict__ aN, int i, int j) {
  // afq(u(i,j)) = aN(i,j+1) - aN(i,j), aN = 0 outside 2..m x 2..n
  const T up = (j + 1 <= g.n) ? aN[wix(g, i, j + 1)] : T(0);
  const T dn = (j >= 2) ? aN[wix(g, i, j)] : T(0);
  return up - dn;
}
template <typename T>
__device__ __forceinline__ T afq_v(const Geo& g, const T* __restrict__ aN, int i, int j) {
  // afq(v(i,j)) = aN(i,j) - aN(i+1,j)
  const T here = (i >= 2) ? aN[wix(g, i, j)] : T(0);
  const T nxt = (i + 1 <= g.m) ? aN[wix(g, i + 1, j)] : T(0);
  return here - nxt;
}

// aN = c1*aX + aro_in (+ c2*aX on the Euler start); aro_out = c2*aX
// (0 on the Euler start).  aro_in and aro_out may alias.
template <typename T>
__global__ void k_step1_adj_combine(Geo g, const T* __restrict__ aX, T* aro, long fbs, T* __restrict__ aN,
                                    const T* __restrict__ c1, const T* __restrict__ c2, const int* __restrict__ Lc,
                                    int lev, int first) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (e >= g.nw || !active(Lc, b, lev)) return;
  const int c = b * g.Lmax + lev;
  const T x = aX[(long)b * g.nw + e];
  T* r = aro + b * fbs + e;
  const T rin = *r;
  aN[(long)b * g.nw + e] = first ? (c1[c] + c2[c]) * x + rin : c1[c] * x + rin;
  *r = first ? T(0) : c2[c] * x;
}

// Vertex part of the nonlinear adjoint: aw_old += nl_Omega^T(afq)
template <typename T>
__global__ void k_nl_adj_vertex(Geo g, const T* __restrict__ aN, const T* __restrict__ q, const T* __restrict__ q0,
                                long qbs, T* __restrict__ aw, long fbs, const int* __restrict__ Lc, int lev) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (e >= g.nw || !active(Lc, b, lev)) return;
  const int i = e % g.mm1 + 2, j = e / g.mm1 + 2;
  const T* an = aN + (long)b * g.nw;
  const T* qq = q + b * qbs;
  const T* qq0 = q0 + b * qbs;
  auto qt = [&](int f) -> T { return qq[f] + qq0[f]; };
  const T vavg = T(0.5) * (qt(vix(g, i, j)) + qt(vix(g, i - 1, j)));
  const T uavg = T(0.5) * (qt(uix(g, i, j)) + qt(uix(g, i, j - 1)));
  const T v = T(0.5) * vavg * (afq_u<T>(g, an, i, j) + afq_u<T>(g, an, i, j - 1)) -
              T(0.5) * uavg * (afq_v<T>(g, an, i, j) + afq_v<T>(g, an, i - 1, j));
  T* o = aw + b * fbs + e;
  *o = *o + v;
}

// Face part: aq_old = nl_qt^T(afq) (assigned: the old fluxes enter the step
// only through the nonlinear term).  Omega's edge values are lb.
template <typename T>
__global__ void k_nl_adj_face(Geo g, const T* __restrict__ aN, const T* __restrict__ om, long fbs,
                              const T* __restrict__ lb, long bbs, T* __restrict__ aq, long qbs,
                              const int* __restrict__ Lc, int lev) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (f >= g.nq || !active(Lc, b, lev)) return;
  const int m = g.m, n = g.n;
  const T* an = aN + (long)b * g.nw;
  const T* w = om + b * fbs;
  const T* bc = lb + b * bbs;
  auto OmU = [&](int ii, int jj) -> T {  // Omega along j at column ii (2..m)
    if (jj == 1) return bc[g.bottom + ii - 1];
    if (jj == n + 1) return bc[g.top + ii - 1];
    return w[wix(g, ii, jj)];
  };
  auto OmV = [&](int ii, int jj) -> T {  // Omega along i at row jj (2..n)
    if (ii == 1) return bc[g.left + jj - 1];
    if (ii == m + 1) return bc[g.right + jj - 1];
    return w[wix(g, ii, jj)];
  };
  // avavg(i,j), i = 2..m, j = 1..n+1 ; auavg(i,j), i = 1..m+1, j = 2..n
  auto avavg = [&](int i, int j) -> T {
    T s = T(0);
    if (j <= n) s += afq_u<T>(g, an, i, j);
    if (j >= 2) s += afq_u<T>(g, an, i, j - 1);
    return T(0.5) * OmU(i, j) * s;
  };
  auto auavg = [&](int i, int j) -> T {
    T s = T(0);
    if (i <= m) s += afq_v<T>(g, an, i, j);
    if (i >= 2) s += afq_v<T>(g, an, i - 1, j);
    return T(-0.5) * OmV(i, j) * s;
  };
  const int nu = (m + 1) * n;
  T r = T(0);
  if (f < nu) {  // u(a,bb): in uavg(a,bb) (bb >= 2) and uavg(a,bb+1) (bb+1 <= n)
    const int a = f % (m + 1) + 1, bb = f / (m + 1) + 1;
    if (bb >= 2) r += auavg(a, bb);
    if (bb + 1 <= n) r += auavg(a, bb + 1);
  } else {  // v(a,bb): in vavg(a,bb) (a >= 2) and vavg(a+1,bb) (a+1 <= m)
    const int fv = f - nu;
    const int a = fv % m + 1, bb = fv / m + 1;
    if (a >= 2) r += avavg(a, bb);
    if (a + 1 <= m) r += avavg(a + 1, bb);
  }
  aq[b * qbs + f] = T(0.5) * r;
}

// Edge part: alb = vfac * apply_bc^T(aX) + nl_bc^T(afq)
template <typename T>
__global__ void k_nl_adj_bc(Geo g, const T* __restrict__ aN, const T* __restrict__ aX, const T* __restrict__ q,
                            const T* __restrict__ q0, long qbs, const T* __restrict__ vfac, T* __restrict__ alb,
                            long bbs, T* __restrict__ anew, const int* __restrict__ Lc, int lev) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (p >= g.nbc || !active(Lc, b, lev)) return;
  const int m = g.m, n = g.n, pos = p + 1;
  const T* an = aN + (long)b * g.nw;
  const T* qq = q + b * qbs;
  const T* qq0 = q0 + b * qbs;
  auto qt = [&](int f) -> T { return qq[f] + qq0[f]; };
  T nl = T(0);
  if (pos > g.bottom) {
    const bool top = pos > g.top;
    const int i = pos - (top ? g.top : g.bottom);
    if (i >= 2 && i <= m) {
      const int jv = top ? n + 1 : 1;  // Omega(i, jv) = edge value
      const T vavg = T(0.5) * (qt(vix(g, i, jv)) + qt(vix(g, i - 1, jv)));
      nl = T(0.5) * vavg * afq_u<T>(g, an, i, top ? n : 1);
    }
  } else {
    const bool right = pos > g.right;
    const int j = pos - (right ? g.right : g.left);
    if (j >= 2 && j <= n) {
      const int iu = right ? m + 1 : 1;
      const T uavg = T(0.5) * (qt(uix(g, iu, j)) + qt(uix(g, iu, j - 1)));
      nl = T(-0.5) * uavg * afq_v<T>(g, an, right ? m : 1, j);
    }
  }
  const T ab = bc_adj_at<T>(g, p, nullptr, aX + (long)b * g.nw, vfac[b * g.Lmax + lev]);
  alb[b * bbs + p] = ab + nl;
  anew[b * bbs + p] = ab;  // the part that multiplies the new omega_{k+1} edge (x0.25 applied later)
}

// Linearised rot(nonlinear) for the tangent step:
// rot(nl(dOmega, q+q0, dbc)) + rot(nl(Omega, dq, bc)).
template <typename T>
__device__ __forceinline__ T rot_nonlinear_tlm(const Geo& g, int i, int j, const T* dom, const T* bq,
                                               const T* bq0, const T* dlb, const T* bom, const T* dq,
                                               const T* zero_q0, const T* blb) {
  return rot_nonlinear<T>(g, i, j, dom, bq, bq0, dlb) + rot_nonlinear<T>(g, i, j, bom, dq, zero_q0, blb);
}

// afb = wfx * cd + wfy * cl per point (J = mean over the window of
// cd * sum of body x forces + cl * sum of y forces): write afb into out.
static __global__ void k_force_adj(Smear sm, const double* __restrict__ wsum, double cd, double cl, double scale,
                            double* __restrict__ afb) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  const CaseSm c = sm.cs[blockIdx.y];
  if (f >= c.nf) return;
  const double coef = f < c.nb ? cd : cl;
  afb[c.foff + f] = scale * coef * 2.0 * c.delta / c.dt * wsum[c.foff + f];
}

// The adjoint of one step's diagnostics row (write_force and write_circ,
// src/operators.f90:583-684) with the row's cotangent grow, laid out like a
// history row: afb of point force f = (2 delta / dt) wsum_f times the
// cotangent of its body's x or y force column (body bdyno_p - bmin).
static __global__ void k_force_adj_row(Smear sm, const double* __restrict__ wsum, const int* __restrict__ bdyno,
                                       const double* __restrict__ grow, double* __restrict__ afb) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  const CaseSm c = sm.cs[blockIdx.y];
  if (f >= c.nf) return;
  const int p = f < c.nb ? f : f - c.nb, comp = f < c.nb ? 0 : 1;
  const double g = grow[c.hoff + comp * c.nbody + (bdyno[c.foff / 2 + p] - c.bmin)];
  afb[c.foff + f] = g * 2.0 * c.delta / c.dt * wsum[c.foff + f];
}

// circ(k) = sum of omega_k (write_circ): aw_k += the cotangent of circ(k),
// read from the row grow at the case's circulation columns
template <typename T>
__global__ void k_circ_adj(Geo g, Smear sm, const double* __restrict__ grow, T* __restrict__ aw, long bs,
                           const int* __restrict__ Lc) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int b = blockIdx.y / g.Lmax, lev = blockIdx.y % g.Lmax;
  if (e >= g.nw || !active(Lc, b, lev)) return;
  const CaseSm c = sm.cs[b];
  const double gc = grow[c.hoff + 2 * c.nbody + lev];
  if (gc != 0.0) aw[b * bs + (long)lev * g.nw + e] += T(gc);
}

// au[f] += delta * z[f]
static __global__ void k_axpy_case(Smear sm, const double* __restrict__ z, double a, double* __restrict__ y, int use_delta) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  const CaseSm c = sm.cs[blockIdx.y];
  if (f >= c.nf) return;
  y[c.foff + f] += (use_delta ? c.delta : 1.0) * a * z[c.foff + f];
}

}  // namespace ibfs
