// ibfs_capi.hip -- host side of the gfx950 IBFS library: set-up tables, device
// memory, the step schedule (captured once into a HIP graph) and the C-ABI
// declared in include/ibfs.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/ibfs.h"
#include "ibfs_kernels.hpp"

using namespace ibfs;

namespace {

constexpr int kEmax = 4096;  // complex values per workgroup in a DST pass
constexpr double kPiRef = 3.1415926535897932384;  // src/parameters.f90:35

struct HipError {
  std::string msg;
};

#define HIPCHK(expr)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) throw HipError{std::string(#expr) + ": " + hipGetErrorString(e_)}; \
  } while (0)

struct ArgError {
  int code;
  std::string msg;
};

// deltafnc (src/grid.f90:233-249), Roma, Peskin & Berger (1999)
double deltafnc(double r, double r0, double dr) {
  const double d = r - r0;
  if (std::fabs(d) <= 0.5 * dr) return (1.0 + std::sqrt(-3.0 * (d / dr) * (d / dr) + 1.0)) / (3.0 * dr);
  if (std::fabs(d) <= 1.5 * dr) {
    const double a = 1.0 - std::fabs(d / dr);
    return (5.0 - 3.0 * std::fabs(d / dr) - std::sqrt(-3.0 * a * a + 1.0)) / (6.0 * dr);
  }
  return 0.0;
}

bool factor_235(int n, std::vector<int>& rad) {
  rad.clear();
  int r = n;
  while (r % 4 == 0) { rad.push_back(4); r /= 4; }
  while (r % 2 == 0) { rad.push_back(2); r /= 2; }
  while (r % 3 == 0) { rad.push_back(3); r /= 3; }
  while (r % 5 == 0) { rad.push_back(5); r /= 5; }
  return r == 1 && (int)rad.size() <= kMaxStages;
}

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    release();
    n = count;
    if (count) HIPCHK(hipMalloc(&p, sizeof(T) * count));
  }
  void zero(hipStream_t s) {
    if (n) HIPCHK(hipMemsetAsync(p, 0, sizeof(T) * n, s));
  }
  void upload(const T* h, size_t count, hipStream_t s, size_t off = 0) {
    if (count) HIPCHK(hipMemcpyAsync(p + off, h, sizeof(T) * count, hipMemcpyHostToDevice, s));
  }
  void download(T* h, size_t count, hipStream_t s, size_t off = 0) const {
    if (count) HIPCHK(hipMemcpyAsync(h, p + off, sizeof(T) * count, hipMemcpyDeviceToHost, s));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
};

}  // namespace

struct ibfs_handle {
  // sizes and parameters
  int m = 0, n = 0, L = 0, nb = 0, nf = 0, nq = 0, nw = 0, nbc = 0, nbody = 0, bmin = 0, bmax = 0, nsmear = 0;
  double dt = 0, re = 0, len = 0, offx = 0, offy = 0, delta = 0, normalize = 0;
  int precision = 0;
  Geo geo{};
  std::string err;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;

  // FFT plans
  FftPlan<double> pm{}, pn{};
  DevBuf<cplx<double>> tw_m, tw_n;

  // tables
  std::vector<double> con1, con2, vfac;
  DevBuf<double> laminv, lam1, lam1i, lam1n;

  // state (reference layout)
  DevBuf<double> omega, rhs_old, q, q0, fb;
  // scratch
  DevBuf<double> S, T1, T2, T3, RR, lastbc, partial, diag;
  DevBuf<double> motion_in, motion, rhsf, ytmp;
  DevBuf<int> counter;
  int diag_cap = 0;
  int nchunk = 64;

  // smearing tables
  std::vector<int> h_face_id, h_face_ptr, h_face_pt, h_frc_ptr, h_frc_face, h_vtx_id, h_vtx_face;
  std::vector<double> h_face_w, h_frc_w;
  DevBuf<int> face_id, face_ptr, face_pt, frc_ptr, frc_face, vtx_id, vtx_face, bdyno_d;
  DevBuf<double> face_w, frc_w;
  Smear sm{};
  std::vector<int> bdyno;

  // force system
  bool have_chol = false;
  std::vector<double> cholmat, cholvec;
  DevBuf<double> Linv;

  // graph of one (non-first) step
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;

  ~ibfs_handle() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }

  int P_m() const { return kEmax / pm.n2; }
  int P_n() const { return kEmax / pn.n2; }
  static size_t lds() { return kEmax * sizeof(cplx<double>); }

  // ------------------------------------------------------------------ passes
  // FFTs per workgroup of the size-specialised passes (row pass / column pass)
  static constexpr int spec_p(int n2, int axis) {
    return axis == 0 ? (n2 >= 1024 ? 1 : (n2 >= 512 ? 2 : 4)) : (n2 >= 2048 ? 1 : (n2 >= 1024 ? 2 : 4));
  }
  DevBuf<cplx<double>> stw_m, stw_n;  // per-stage twiddles of the specialised sizes
  bool spec_m = false, spec_n = false;

  // Launch one pass.  nseq sequences along the other axis; mult = 2 when a
  // workgroup's P FFTs take 2P sequences of one field, 1 when they pair two
  // fields (X + iY).
  template <int AXIS, class Op>
  void launch(const Op& op, int nseq, int mult, int batch) {
    const int n2 = AXIS == 0 ? pm.n2 : pn.n2;
    const bool sp = AXIS == 0 ? spec_m : spec_n;
    const cplx<double>* stw = AXIS == 0 ? stw_m.p : stw_n.p;
    if (sp) {
      switch (n2) {
#define IBFS_SPEC_CASE(NN)                                                                                  \
  case NN: {                                                                                                \
    constexpr int P = spec_p(NN, AXIS);                                                                     \
    const int groups = (nseq + mult * P - 1) / (mult * P);                                                  \
    hipLaunchKernelGGL((spec::dst_pass_n<double, NN, P, AXIS, Op>), dim3(groups, batch), dim3(P * NN / 4), \
                       0, stream, op, stw);                                                                 \
    return;                                                                                                 \
  }
        IBFS_SPEC_CASE(400)
        IBFS_SPEC_CASE(512)
        IBFS_SPEC_CASE(640)
        IBFS_SPEC_CASE(768)
        IBFS_SPEC_CASE(1024)
        IBFS_SPEC_CASE(2048)
#undef IBFS_SPEC_CASE
        default: break;
      }
    }
    const FftPlan<double>& plan = AXIS == 0 ? pm : pn;
    const int P = kEmax / n2;
    const int groups = (nseq + mult * P - 1) / (mult * P);
    hipLaunchKernelGGL((dst_pass<double, kEmax, AXIS, Op>), dim3(groups, batch), dim3(kThreads), lds(), stream, op,
                       plan);
  }

  template <class Src, class Dst>
  void row_pass(const Src& src, const Dst& dst, int batch) {
    launch<0>(RowOp<double, Src, Dst>{geo, src, dst}, geo.nm1, 2, batch);
  }
  template <class Src, class Dst>
  void col_pass(const Src& src, const Dst& dst, int batch) {
    launch<1>(ColOp<double, Src, Dst>{geo, src, dst}, geo.mm1, 2, batch);
  }

  // out = S * diag(tab) * S (src) with the given first-pass source and final
  // destination functors; t1/t2 are scratch of batch*nw.
  template <class Src, class Dst>
  void ctci_like(const Src& src, const double* tab, bool ainv_form, const Dst& dst, double* t1, double* t2, int batch,
                 long bs_scratch) {
    row_pass(src, DstField<double>{t1, bs_scratch}, batch);
    if (ainv_form)
      col_pass(SrcField<double>{t1, bs_scratch}, DstScaleDiv<double>{t2, tab, normalize, bs_scratch}, batch);
    else
      col_pass(SrcField<double>{t1, bs_scratch}, DstScale<double>{t2, tab, bs_scratch}, batch);
    row_pass(SrcField<double>{t2, bs_scratch}, DstField<double>{t1, bs_scratch}, batch);
    col_pass(SrcField<double>{t1, bs_scratch}, dst, batch);
  }

  // vort2flux (src/operators.f90:191-240) on a batch of multi-level fields
  // om (batch stride bs = L*nw); streamfunctions go to s (same layout); curl
  // written to vel (level stride nq, batch stride vbs) for levels < nlev.
  void vort2flux(double* om, double* s, double* vel, long vbs, int nlev, int batch, double* t1, double* t2) {
    const long bs = (long)L * nw;
    const int wcnt = (m / 2 - 1) * (n / 2 - 1);
    for (int lev = 1; lev < L; lev++)
      hipLaunchKernelGGL(k_coarsify<double>, dim3((wcnt + 255) / 256, batch), dim3(256), 0, stream, geo,
                         om + (long)lev * nw, om + (long)(lev - 1) * nw, bs);
    for (int lev = L - 1; lev >= 0; lev--) {
      const double* sc = (lev == L - 1) ? nullptr : s + (long)(lev + 1) * nw;
      ctci_like(SrcVortBC<double>{om + (long)lev * nw, sc, bs}, laminv.p, false,
                DstField<double>{s + (long)lev * nw, bs}, t1, t2, batch, nw);
      if (vel && lev < nlev)
        hipLaunchKernelGGL(k_curl<double>, dim3((nq + 255) / 256, batch), dim3(256), 0, stream, geo,
                           s + (long)lev * nw, sc, vel + (long)lev * nq, bs, vbs);
    }
  }

  // --------------------------------------------------------------- one step
  void enqueue_step(int first) {
    hipLaunchKernelGGL(k_lastbc<double>, dim3((nbc * L + 255) / 256), dim3(256), 0, stream, geo, omega.p, lastbc.p);
    // STEP 1 (src/operators.f90:103-134), levels L..1
    for (int lev = L - 1; lev >= 0; lev--) {
      Step1RowOp<double> r1{geo,
                            omega.p + (long)lev * nw,
                            q.p + (long)lev * nq,
                            q0.p + (long)lev * nq,
                            lastbc.p + (long)lev * nbc,
                            lev < L - 1 ? omega.p + (long)(lev + 1) * nw : nullptr,
                            rhs_old.p + (long)lev * nw,
                            T1.p,
                            T1.p + nw,
                            con1[lev],
                            con2[lev],
                            vfac[lev],
                            normalize,
                            first};
      launch<0>(r1, geo.nm1, 1, 1);
      Step1ColOp<double> c1{geo, T1.p, T1.p + nw, lam1n.p + (long)lev * nw, lam1i.p + (long)lev * nw, T2.p};
      launch<1>(c1, geo.mm1, 1, 1);
      row_pass(SrcField<double>{T2.p, 0}, DstField<double>{T3.p, 0}, 1);
      col_pass(SrcField<double>{T3.p, 0}, DstField<double>{omega.p + (long)lev * nw, 0}, 1);
    }
    // STEP 2 (src/operators.f90:142-171)
    hipLaunchKernelGGL(k_motion<double>, dim3((nf + 255) / 256), dim3(256), 0, stream, sm, q0.p, motion_in.p,
                       motion.p);
    vort2flux(omega.p, S.p, nullptr, 0, 0, 1, T1.p, T2.p);
    hipLaunchKernelGGL(k_regT_curl<double>, dim3((nf + 255) / 256, 1), dim3(256), 0, stream, geo, sm, S.p,
                       L > 1 ? S.p + nw : nullptr, 0L, motion.p, rhsf.p, 0L);
    hipLaunchKernelGGL(k_trmv_lower, dim3((nf + 3) / 4), dim3(256), 0, stream, nf, Linv.p, rhsf.p, ytmp.p);
    hipLaunchKernelGGL(k_trmv_upper_t, dim3((nf + 3) / 4), dim3(256), 0, stream, nf, Linv.p, ytmp.p, fb.p);
    // STEP 3 (src/operators.f90:174)
    hipLaunchKernelGGL(k_rotreg<double>, dim3((sm.nvtx + 255) / 256, 1), dim3(256), 0, stream, sm, fb.p, 0L, RR.p,
                       0L);
    ctci_like(SrcField<double>{RR.p, 0}, lam1i.p, true, DstSub<double>{omega.p, 0}, T1.p, T2.p, 1, 0);
    // STEP 4 (src/operators.f90:177)
    vort2flux(omega.p, S.p, q.p, 0, L, 1, T1.p, T2.p);
    // diagnostics (src/operators.f90:184-186)
    hipLaunchKernelGGL(k_circ_partial<double>, dim3(nchunk, L), dim3(kThreads), 0, stream, geo, omega.p, partial.p,
                       nchunk);
    hipLaunchKernelGGL(k_diag<double>, dim3(1), dim3(1024), 0, stream, geo, sm, q.p, motion.p, fb.p, bdyno_d.p, nb,
                       bmin, nbody, partial.p, nchunk, dt, delta, diag.p, counter.p, diag_cap);
  }

  void ensure_diag(int cap) {
    if (cap <= diag_cap) return;
    HIPCHK(hipStreamSynchronize(stream));
    if (graph_exec) { (void)hipGraphExecDestroy(graph_exec); graph_exec = nullptr; }
    if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
    diag.alloc((size_t)cap * (2 * nbody + L + 1));
    diag_cap = cap;
  }

  void capture() {
    if (graph_exec) return;
    HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    enqueue_step(0);
    HIPCHK(hipStreamEndCapture(stream, &graph));
    HIPCHK(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
  }

  // ---------------------------------------------------------- Cholesky (A)
  // a_times (src/operators.f90:444-462) for the unit vectors c0..c0+B-1,
  // columns written to acol (nf x nf, column-major, device).
  void build_columns(double* acol) {
    const int B = std::max(1, std::min(nf, 64));
    const long bsL = (long)L * nw;
    DevBuf<double> F, RRb, T1b, T2b, VB, SB;
    F.alloc((size_t)B * nf);
    RRb.alloc((size_t)B * nw);
    T1b.alloc((size_t)B * nw);
    T2b.alloc((size_t)B * nw);
    VB.alloc((size_t)B * bsL);
    SB.alloc((size_t)B * bsL);
    RRb.zero(stream);
    VB.zero(stream);
    std::vector<double> hF((size_t)B * nf);
    for (int c0 = 0; c0 < nf; c0 += B) {
      const int bb = std::min(B, nf - c0);
      std::fill(hF.begin(), hF.end(), 0.0);
      for (int b = 0; b < bb; b++) hF[(size_t)b * nf + c0 + b] = 1.0;
      F.upload(hF.data(), (size_t)bb * nf, stream);
      hipLaunchKernelGGL(k_rotreg<double>, dim3((sm.nvtx + 255) / 256, bb), dim3(256), 0, stream, sm, F.p, (long)nf,
                         RRb.p, (long)nw);
      ctci_like(SrcField<double>{RRb.p, nw}, lam1i.p, true, DstField<double>{VB.p, bsL}, T1b.p, T2b.p, bb, nw);
      vort2flux(VB.p, SB.p, nullptr, 0, 0, bb, T1b.p, T2b.p);
      hipLaunchKernelGGL(k_regT_curl<double>, dim3((nf + 255) / 256, bb), dim3(256), 0, stream, geo, sm, SB.p,
                         L > 1 ? SB.p + nw : nullptr, bsL, (const double*)nullptr, acol + (long)c0 * nf, (long)nf);
      HIPCHK(hipStreamSynchronize(stream));  // hF is reused
    }
  }

  // choldc (src/operators.f90:766-791): upper triangle keeps A, strict lower
  // gets L, diagonal in cholvec.  Returns 0 or the failing row (1-based).
  int choldc() {
    double* a = cholmat.data();
    auto CM = [&](int i, int j) -> double& { return a[(size_t)(i - 1) + (size_t)(j - 1) * nf]; };
    for (int i = 1; i <= nf; i++)
      for (int j = i; j <= nf; j++) {
        double sum = CM(i, j);
        for (int k = i - 1; k >= 1; k--) sum = sum - CM(i, k) * CM(j, k);
        if (i == j) {
          if (sum <= 0.0) return i;
          cholvec[i - 1] = std::sqrt(sum);
        } else {
          CM(j, i) = sum / cholvec[i - 1];
        }
      }
    return 0;
  }

  // explicit inverse of the lower factor: Li = L^-1 (column-major, lower)
  void upload_inverse() {
    std::vector<double> Li((size_t)nf * nf, 0.0);
    const double* a = cholmat.data();
    auto Lf = [&](int i, int k) -> double { return i == k ? cholvec[i] : a[(size_t)i + (size_t)k * nf]; };
    std::vector<double> col(nf);
    for (int j = 0; j < nf; j++) {
      std::fill(col.begin(), col.end(), 0.0);
      col[j] = 1.0 / Lf(j, j);
      for (int i = j + 1; i < nf; i++) {
        double s = 0.0;
        for (int k = j; k < i; k++) s += Lf(i, k) * col[k];
        col[i] = -s / Lf(i, i);
      }
      std::copy(col.begin(), col.end(), Li.begin() + (size_t)j * nf);
    }
    Linv.upload(Li.data(), Li.size(), stream);
    HIPCHK(hipStreamSynchronize(stream));
    have_chol = true;
  }
};

namespace {

// Compact smearing tables: the entries of setup_reg (src/grid.f90:142-230),
// found per IB point inside its support box instead of by the O(Nq Nb) scan,
// with the same coordinates, the same strict support test and the same
// weight expression, kept in the reference's per-face point order.
void build_smear(ibfs_handle* h, const double* xb, const double* yb) {
  const int m = h->m, n = h->n, nb = h->nb;
  const double delta = h->delta, d2 = 0.5 * delta, sup = 3.0 * d2;
  std::map<int, std::vector<std::pair<int, double>>> faces;  // face -> (force, w)
  for (int k = 0; k < nb; k++) {
    // u faces: x = delta*(i-1) - offx, y = delta*(j-1) - offy + d2
    const int ic = (int)std::floor((xb[k] + h->offx) / delta) + 1;
    const int jc = (int)std::floor((yb[k] + h->offy) / delta) + 1;
    for (int i = std::max(1, ic - 3); i <= std::min(m + 1, ic + 3); i++)
      for (int j = std::max(1, jc - 3); j <= std::min(n, jc + 3); j++) {
        const double x = delta * (double)(i - 1) - h->offx;
        const double y = delta * (double)(j - 1) - h->offy + d2;
        if (std::fabs(x - xb[k]) < sup && std::fabs(y - yb[k]) < sup) {
          const int face = (i - 1) + (j - 1) * (m + 1);
          faces[face].push_back({k, (delta * delta) * deltafnc(x, xb[k], delta) * deltafnc(y, yb[k], delta)});
        }
      }
    for (int i = std::max(1, ic - 3); i <= std::min(m, ic + 3); i++)
      for (int j = std::max(1, jc - 3); j <= std::min(n + 1, jc + 3); j++) {
        const double x = delta * (double)(i - 1) - h->offx + d2;
        const double y = delta * (double)(j - 1) - h->offy;
        if (std::fabs(x - xb[k]) < sup && std::fabs(y - yb[k]) < sup) {
          const int face = (m + 1) * n + (i - 1) + (j - 1) * m;
          faces[face].push_back({nb + k, (delta * delta) * deltafnc(x, xb[k], delta) * deltafnc(y, yb[k], delta)});
        }
      }
  }
  h->nsmear = 0;
  h->h_face_id.clear(); h->h_face_ptr.assign(1, 0); h->h_face_pt.clear(); h->h_face_w.clear();
  std::vector<std::vector<std::pair<int, double>>> frc(h->nf);
  std::map<int, int> slot_of_face;
  for (auto& kv : faces) {
    const int a = (int)h->h_face_id.size();
    slot_of_face[kv.first] = a;
    h->h_face_id.push_back(kv.first);
    for (auto& e : kv.second) {
      h->h_face_pt.push_back(e.first);
      h->h_face_w.push_back(e.second);
      frc[e.first].push_back({kv.first, e.second});
    }
    h->h_face_ptr.push_back((int)h->h_face_pt.size());
    h->nsmear = std::max(h->nsmear, (int)kv.second.size());
  }
  h->h_frc_ptr.assign(1, 0); h->h_frc_face.clear(); h->h_frc_w.clear();
  for (int f = 0; f < h->nf; f++) {
    for (auto& e : frc[f]) { h->h_frc_face.push_back(e.first); h->h_frc_w.push_back(e.second); }
    h->h_frc_ptr.push_back((int)h->h_frc_face.size());
  }
  // active vertices: (i,j) reads v(i,j), v(i-1,j), u(i,j), u(i,j-1)
  std::map<int, std::array<int, 4>> vtx;
  auto touch = [&](int i, int j, int c, int a) {
    if (i < 2 || i > m || j < 2 || j > n) return;
    const int w = (i - 2) + (j - 2) * (m - 1);
    auto it = vtx.find(w);
    if (it == vtx.end()) it = vtx.emplace(w, std::array<int, 4>{-1, -1, -1, -1}).first;
    it->second[c] = a;
  };
  const int nu = (m + 1) * n;
  for (auto& kv : slot_of_face) {
    const int face = kv.first, a = kv.second;
    if (face < nu) {
      const int i = face % (m + 1) + 1, j = face / (m + 1) + 1;
      touch(i, j, 2, a);
      touch(i, j + 1, 3, a);
    } else {
      const int fv = face - nu;
      const int i = fv % m + 1, j = fv / m + 1;
      touch(i, j, 0, a);
      touch(i + 1, j, 1, a);
    }
  }
  h->h_vtx_id.clear(); h->h_vtx_face.clear();
  for (auto& kv : vtx) {
    h->h_vtx_id.push_back(kv.first);
    for (int c = 0; c < 4; c++) h->h_vtx_face.push_back(kv.second[c]);
  }
}

template <typename V>
void up(DevBuf<V>& d, const std::vector<V>& hv, hipStream_t s) {
  d.alloc(std::max<size_t>(hv.size(), 1));
  d.upload(hv.data(), hv.size(), s);
}

void make_plan(int n2, FftPlan<double>& plan, DevBuf<cplx<double>>& tw, hipStream_t s) {
  std::vector<int> rad;
  if (!factor_235(n2, rad) || n2 > kEmax)
    throw ArgError{IBFS_EINVAL, "grid size not supported: 2*m and 2*n must factor into 2, 3, 5 and be <= " +
                                    std::to_string(kEmax)};
  plan.n2 = n2;
  plan.nst = (int)rad.size();
  for (int i = 0; i < kMaxStages; i++) plan.rad[i] = i < plan.nst ? rad[i] : 1;
  std::vector<cplx<double>> t(n2);
  for (int k = 0; k < n2; k++) {
    long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n2;
    t[k] = {(double)cosl(a), (double)sinl(a)};
  }
  tw.alloc(n2);
  tw.upload(t.data(), n2, s);
  plan.tw = tw.p;
}

// Per-stage twiddles of a size-specialised pass: stage s, butterfly index
// k < ns, factor r = 1..R-1 holds exp(-2 pi i k r / (ns R)).
bool make_spec_twiddles(int n2, DevBuf<cplx<double>>& tw, hipStream_t st) {
  static const int kSizes[] = {400, 512, 640, 768, 1024, 2048};
  if (std::find(std::begin(kSizes), std::end(kSizes), n2) == std::end(kSizes)) return false;
  const spec::Sched sc = spec::make_sched(n2);
  std::vector<cplx<double>> t(std::max(sc.ntw, 1));
  for (int s = 0; s < sc.nst; s++) {
    const int R = sc.rad[s], ns = sc.ns[s];
    for (int k = 0; k < ns; k++)
      for (int r = 1; r < R; r++) {
        long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)(k * r) / (long double)(ns * R);
        t[sc.toff[s] + k * (R - 1) + (r - 1)] = {(double)cosl(a), (double)sinl(a)};
      }
  }
  tw.alloc(t.size());
  tw.upload(t.data(), t.size(), st);
  return true;
}

int fail(ibfs_handle* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  return code;
}

template <class F>
int guarded(ibfs_handle* h, F&& f) {
  if (!h) return IBFS_EINVAL;
  try {
    f();
    return IBFS_OK;
  } catch (const HipError& e) {
    return fail(h, IBFS_EHIP, e.msg);
  } catch (const ArgError& e) {
    return fail(h, e.code, e.msg);
  } catch (const std::exception& e) {
    return fail(h, IBFS_EHIP, e.what());
  }
}

}  // namespace

// ============================================================================
// C-ABI
// ============================================================================
extern "C" {

int ibfs_create(int m, int n, int mgridlev, int nb, double dt, double re, double len, double offsetx,
                double offsety, const double* xb, const double* yb, const int* bdyno, int unif_flow, int precision,
                ibfs_handle** out) {
  if (!out) return IBFS_EINVAL;
  *out = nullptr;
  if (m % 4 != 0 || n % 4 != 0 || m < 8 || n < 8 || mgridlev < 1 || nb < 0 || (nb > 0 && (!xb || !yb)) ||
      !(dt > 0) || !(re > 0) || !(len > 0))
    return IBFS_EINVAL;
  if (precision != IBFS_PLAIN_FP64) return IBFS_EINVAL;
  ibfs_handle* h = new ibfs_handle();
  int rc = guarded(h, [&] {
    h->m = m; h->n = n; h->nb = nb; h->nf = 2 * nb;
    // grid.f90:79-82 (single-precision LOG, as in the reference)
    const int nlevmax = (int)std::floor((double)logf((float)(std::min(m, n) / 2)) / std::log(2.0)) - 1;
    h->L = std::min(mgridlev, nlevmax);
    if (h->L < 1) throw ArgError{IBFS_EINVAL, "grid too small for one level"};
    h->nq = (m + 1) * n + (n + 1) * m;
    h->nw = (m - 1) * (n - 1);
    h->nbc = 2 * (m + 1) + 2 * (n + 1);
    h->dt = dt; h->re = re; h->len = len; h->offx = offsetx; h->offy = offsety;
    h->delta = len / (double)(float)m;  // grid.f90:64
    h->precision = precision;
    h->geo = Geo{m, n, h->L, m - 1, n - 1, h->nw, h->nq, h->nbc, 0, n + 1, 2 * (n + 1), 2 * (n + 1) + m + 1};
    h->bdyno.assign(std::max(nb, 1), 0);
    h->bmin = h->bmax = 0;
    for (int k = 0; k < nb; k++) {
      h->bdyno[k] = bdyno ? bdyno[k] : 0;
      if (k == 0 || h->bdyno[k] < h->bmin) h->bmin = h->bdyno[k];
      if (k == 0 || h->bdyno[k] > h->bmax) h->bmax = h->bdyno[k];
    }
    h->nbody = h->bmax - h->bmin + 1;
    HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&h->ev0));
    HIPCHK(hipEventCreate(&h->ev1));
    hipStream_t s = h->stream;
    make_plan(2 * m, h->pm, h->tw_m, s);
    make_plan(2 * n, h->pn, h->tw_n, s);
    h->spec_m = make_spec_twiddles(2 * m, h->stw_m, s);
    h->spec_n = make_spec_twiddles(2 * n, h->stw_n, s);

    // setup_fft tables (src/myfft.f90:42-69), same expressions and order
    const int L = h->L;
    const size_t nw = h->nw;
    h->normalize = 4.0 * (double)(float)((long)m * (long)n);
    h->con1.resize(L); h->con2.resize(L); h->vfac.resize(L);
    const double del2 = h->delta * h->delta;
    for (int k = 1; k <= L; k++) {
      const double del22 = del2 * std::pow(4.0, k - 1);
      h->vfac[k - 1] = 0.5 * dt / re / del22 / h->normalize;
      h->con1[k - 1] = 1.5 * dt / del22 / h->normalize;
      h->con2[k - 1] = -0.5 * dt / del22 / h->normalize;
    }
    std::vector<double> laminv(nw), lam1(nw * L), lam1i(nw * L), lam1n(nw * L);
    for (int j = 1; j <= n - 1; j++)
      for (int i = 1; i <= m - 1; i++) {
        const size_t e = (size_t)(i - 1) + (size_t)(j - 1) * (m - 1);
        const double lam = -2.0 * (std::cos(kPiRef * (double)(float)i / (double)(float)m) +
                                   std::cos(kPiRef * (double)(float)j / (double)(float)n) - 2.0);
        laminv[e] = 1.0 / lam / h->normalize;
        for (int k = 1; k <= L; k++) {
          const double del22 = del2 * std::pow(4.0, k - 1);
          lam1n[e + nw * (k - 1)] = 1.0 - 0.5 * dt * lam / del22 / re;
          lam1[e + nw * (k - 1)] = lam1n[e + nw * (k - 1)] / h->normalize;
          lam1i[e + nw * (k - 1)] = 1.0 / (1.0 + 0.5 * dt * lam / del22 / re);
        }
      }
    up(h->laminv, laminv, s);
    up(h->lam1, lam1, s);
    up(h->lam1i, lam1i, s);
    up(h->lam1n, lam1n, s);

    build_smear(h, xb, yb);
    up(h->face_id, h->h_face_id, s);
    up(h->face_ptr, h->h_face_ptr, s);
    up(h->face_pt, h->h_face_pt, s);
    up(h->face_w, h->h_face_w, s);
    up(h->frc_ptr, h->h_frc_ptr, s);
    up(h->frc_face, h->h_frc_face, s);
    up(h->frc_w, h->h_frc_w, s);
    up(h->vtx_id, h->h_vtx_id, s);
    up(h->vtx_face, h->h_vtx_face, s);
    up(h->bdyno_d, h->bdyno, s);
    h->sm = Smear{(int)h->h_face_id.size(), h->face_id.p, h->face_ptr.p, h->face_pt.p, h->face_w.p, h->nf,
                  h->frc_ptr.p, h->frc_face.p, h->frc_w.p, (int)h->h_vtx_id.size(), h->vtx_id.p, h->vtx_face.p};

    // state and initial condition (src/variables.f90:54-67)
    const size_t nq = h->nq;
    h->omega.alloc(nw * L); h->omega.zero(s);
    h->rhs_old.alloc(nw * L); h->rhs_old.zero(s);
    h->q.alloc(nq * L); h->q.zero(s);
    h->q0.alloc(nq * L);
    h->fb.alloc(std::max(h->nf, 1)); h->fb.zero(s);
    std::vector<double> q0(nq * L, 0.0);
    if (unif_flow)
      for (int k = 1; k <= L; k++) {
        const double v = h->delta * std::pow(2.0, k - 1);
        for (size_t e = 0; e < (size_t)(m + 1) * n; e++) q0[e + nq * (k - 1)] = v;
      }
    h->q0.upload(q0.data(), q0.size(), s);
    // scratch
    h->S.alloc(nw * L);
    h->T1.alloc(2 * nw);
    h->T2.alloc(nw);
    h->T3.alloc(nw);
    h->RR.alloc(nw); h->RR.zero(s);
    h->lastbc.alloc((size_t)h->nbc * L);
    h->partial.alloc((size_t)h->nchunk * L);
    const size_t nff = std::max(h->nf, 1);
    h->motion_in.alloc(nff); h->motion_in.zero(s);
    h->motion.alloc(nff);
    h->rhsf.alloc(nff);
    h->ytmp.alloc(nff);
    h->Linv.alloc(nff * nff);
    h->counter.alloc(1);
    h->ensure_diag(1);
    HIPCHK(hipStreamSynchronize(s));
  });
  if (rc) {
    // keep the handle so the caller can read the message, but report failure
    *out = h;
    return rc;
  }
  *out = h;
  return IBFS_OK;
}

int ibfs_dims(const ibfs_handle* h, int* d) {
  if (!h || !d) return IBFS_EINVAL;
  const int v[12] = {h->m, h->n, h->L, h->nb, h->nf, h->nq, h->nw, h->nbc, h->nbody, h->bmin, h->bmax, h->nsmear};
  std::memcpy(d, v, sizeof(v));
  return IBFS_OK;
}

int ibfs_set_state(ibfs_handle* h, const double* omega, const double* rhs_old, const double* q, const double* q0,
                   const double* fb) {
  return guarded(h, [&] {
    const size_t W = (size_t)h->nw * h->L, Q = (size_t)h->nq * h->L;
    if (omega) h->omega.upload(omega, W, h->stream);
    if (rhs_old) h->rhs_old.upload(rhs_old, W, h->stream);
    if (q) h->q.upload(q, Q, h->stream);
    if (q0) h->q0.upload(q0, Q, h->stream);
    if (fb) h->fb.upload(fb, h->nf, h->stream);
    HIPCHK(hipStreamSynchronize(h->stream));
  });
}

int ibfs_get_state(ibfs_handle* h, double* omega, double* rhs_old, double* q, double* q0, double* fb) {
  return guarded(h, [&] {
    const size_t W = (size_t)h->nw * h->L, Q = (size_t)h->nq * h->L;
    if (omega) h->omega.download(omega, W, h->stream);
    if (rhs_old) h->rhs_old.download(rhs_old, W, h->stream);
    if (q) h->q.download(q, Q, h->stream);
    if (q0) h->q0.download(q0, Q, h->stream);
    if (fb) h->fb.download(fb, h->nf, h->stream);
    HIPCHK(hipStreamSynchronize(h->stream));
  });
}

int ibfs_set_cholesky(ibfs_handle* h, const double* cholmat, const double* cholvec) {
  return guarded(h, [&] {
    if (!cholmat || !cholvec) throw ArgError{IBFS_EINVAL, "null factor"};
    const size_t nf = h->nf;
    h->cholmat.assign(cholmat, cholmat + nf * nf);
    h->cholvec.assign(cholvec, cholvec + nf);
    for (size_t i = 0; i < nf; i++)
      if (!(h->cholvec[i] > 0.0)) throw ArgError{IBFS_ECHOL, "cholvec has a non-positive entry"};
    h->upload_inverse();
  });
}

int ibfs_build_cholesky(ibfs_handle* h, double* cholmat, double* cholvec) {
  return guarded(h, [&] {
    const size_t nf = h->nf;
    if (nf == 0) { h->have_chol = true; return; }
    DevBuf<double> A;
    A.alloc(nf * nf);
    h->build_columns(A.p);
    h->cholmat.assign(nf * nf, 0.0);
    h->cholvec.assign(nf, 0.0);
    A.download(h->cholmat.data(), nf * nf, h->stream);
    HIPCHK(hipStreamSynchronize(h->stream));
    const int bad = h->choldc();
    if (bad) throw ArgError{IBFS_ECHOL, "choldc failed at row " + std::to_string(bad)};
    h->upload_inverse();
    if (cholmat) std::memcpy(cholmat, h->cholmat.data(), sizeof(double) * nf * nf);
    if (cholvec) std::memcpy(cholvec, h->cholvec.data(), sizeof(double) * nf);
  });
}

int ibfs_advance(ibfs_handle* h, int* itime, int nsteps, const double* motion, double* hist) {
  if (!h || !itime || nsteps < 0) return IBFS_EINVAL;
  return guarded(h, [&] {
    if (!h->have_chol && h->nf > 0) throw ArgError{IBFS_ESTATE, "no Cholesky factor: call ibfs_build_cholesky or "
                                                                "ibfs_set_cholesky before the first step"};
    if (nsteps == 0) return;
    hipStream_t s = h->stream;
    h->ensure_diag(nsteps);
    if (motion) h->motion_in.upload(motion, h->nf, s);
    else h->motion_in.zero(s);
    const int zero = 0;
    HIPCHK(hipMemcpyAsync(h->counter.p, &zero, sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipEventRecord(h->ev0, s));
    for (int k = 0; k < nsteps; k++) {
      const int it = *itime + 1;
      if (it == 1) {
        h->enqueue_step(1);
      } else {
        h->capture();
        HIPCHK(hipGraphLaunch(h->graph_exec, s));
      }
      *itime = it;
    }
    HIPCHK(hipEventRecord(h->ev1, s));
    if (hist) h->diag.download(hist, (size_t)nsteps * (2 * h->nbody + h->L + 1), s);
    HIPCHK(hipStreamSynchronize(s));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->last_ms = ms;
  });
}

int ibfs_step(ibfs_handle* h, int* itime, const double* motion, double* fbo, double* force, double* circ,
              double* slip) {
  if (!h) return IBFS_EINVAL;
  std::vector<double> row(2 * h->nbody + h->L + 1);
  int rc = ibfs_advance(h, itime, 1, motion, row.data());
  if (rc) return rc;
  if (force) std::memcpy(force, row.data(), sizeof(double) * 2 * h->nbody);
  if (circ) std::memcpy(circ, row.data() + 2 * h->nbody, sizeof(double) * h->L);
  if (slip) *slip = row[2 * h->nbody + h->L];
  if (fbo) return ibfs_get_state(h, nullptr, nullptr, nullptr, nullptr, fbo);
  return IBFS_OK;
}

int ibfs_last_advance_ms(const ibfs_handle* h, double* ms) {
  if (!h || !ms) return IBFS_EINVAL;
  *ms = h->last_ms;
  return IBFS_OK;
}

int ibfs_time_dst_pass(ibfs_handle* h, int axis, int reps, double* ms_per_launch) {
  if (!h || !ms_per_launch || reps < 1 || (axis != 0 && axis != 1)) return IBFS_EINVAL;
  return guarded(h, [&] {
    DevBuf<double> a, b;
    a.alloc(h->nw); b.alloc(h->nw);
    a.zero(h->stream);
    auto launch = [&] {
      if (axis == 0) h->row_pass(SrcField<double>{a.p, 0}, DstField<double>{b.p, 0}, 1);
      else h->col_pass(SrcField<double>{a.p, 0}, DstField<double>{b.p, 0}, 1);
    };
    launch();  // warm
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    for (int r = 0; r < reps; r++) launch();
    HIPCHK(hipEventRecord(h->ev1, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    *ms_per_launch = ms / reps;
  });
}

// ----------------------------------------------------------------- operators
int ibfs_op_dst(ibfs_handle* h, const double* in, double* out) {
  return guarded(h, [&] {
    DevBuf<double> a, b;
    a.alloc(h->nw); b.alloc(h->nw);
    a.upload(in, h->nw, h->stream);
    h->row_pass(SrcField<double>{a.p, 0}, DstField<double>{b.p, 0}, 1);
    h->col_pass(SrcField<double>{b.p, 0}, DstField<double>{a.p, 0}, 1);
    a.download(out, h->nw, h->stream);
    HIPCHK(hipStreamSynchronize(h->stream));
  });
}

static int op_ctci(ibfs_handle* h, const double* in, double* out, bool ainv) {
  return guarded(h, [&] {
    DevBuf<double> a, b, c, t1, t2;
    a.alloc(h->nw); b.alloc(h->nw); t1.alloc(h->nw); t2.alloc(h->nw);
    a.upload(in, h->nw, h->stream);
    h->ctci_like(SrcField<double>{a.p, 0}, ainv ? h->lam1i.p : h->laminv.p, ainv, DstField<double>{b.p, 0}, t1.p,
                 t2.p, 1, 0);
    b.download(out, h->nw, h->stream);
    HIPCHK(hipStreamSynchronize(h->stream));
  });
}
int ibfs_op_ctci(ibfs_handle* h, const double* in, double* out) { return op_ctci(h, in, out, false); }
int ibfs_op_ainv(ibfs_handle* h, const double* in, double* out) { return op_ctci(h, in, out, true); }

int ibfs_op_curl(ibfs_handle* h, const double* x, const double* coarse, double* vel) {
  return guarded(h, [&] {
    DevBuf<double> a, c, v;
    a.alloc(h->nw); c.alloc(h->nw); v.alloc(h->nq);
    a.upload(x, h->nw, h->stream);
    if (coarse) c.upload(coarse, h->nw, h->stream);
    hipLaunchKernelGGL(k_curl<double>, dim3((h->nq + 255) / 256, 1), dim3(256), 0, h->stream, h->geo, a.p,
                       coarse ? c.p : nullptr, v.p, 0L, 0L);
    v.download(vel, h->nq, h->stream);
    HIPCHK(hipStreamSynchronize(h->stream));
  });
}

int ibfs_op_rot(ibfs_handle* h, const double* x, double* out) {
  return guarded(h, [&] {
    DevBuf<double> a, r;
    a.alloc(h->nq); r.alloc(h->nw);
    a.upload(x, h->nq, h->stream);
    hipLaunchKernelGGL(k_rot<double>, dim3((h->nw + 255) / 256), dim3(256), 0, h->stream, h->geo, a.p, r.p);
    r.download(out, h->nw, h->stream);
    HIPCHK(hipStreamSynchronize(h->stream));
  });
}

int ibfs_op_coarsify(ibfs_handle* h, const double* crhs, const double* rhs, double* out) {
  return guarded(h, [&] {
    DevBuf<double> c, f;
    c.alloc(h->nw); f.alloc(h->nw);
    c.upload(crhs, h->nw, h->stream);
    f.upload(rhs, h->nw, h->stream);
    const int wcnt = (h->m / 2 - 1) * (h->n / 2 - 1);
    hipLaunchKernelGGL(k_coarsify<double>, dim3((wcnt + 255) / 256, 1), dim3(256), 0, h->stream, h->geo, c.p, f.p,
                       0L);
    c.download(out, h->nw, h->stream);
    HIPCHK(hipStreamSynchronize(h->stream));
  });
}

namespace {
__global__ void k_getbc_all(Geo g, const double* r, double fac, double* rbc) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < g.nbc) rbc[p] = getbc_at<double>(g, r, p + 1, fac);
}
}  // namespace

int ibfs_op_get_bc(ibfs_handle* h, const double* r, double fac, double* rbc) {
  return guarded(h, [&] {
    DevBuf<double> a, b;
    a.alloc(h->nw); b.alloc(h->nbc);
    a.upload(r, h->nw, h->stream);
    hipLaunchKernelGGL(k_getbc_all, dim3((h->nbc + 255) / 256), dim3(256), 0, h->stream, h->geo, a.p, fac, b.p);
    b.download(rbc, h->nbc, h->stream);
    HIPCHK(hipStreamSynchronize(h->stream));
  });
}

int ibfs_op_rot_nonlinear(ibfs_handle* h, const double* omega, const double* q, const double* q0, const double* bc,
                          double* out) {
  return guarded(h, [&] {
    DevBuf<double> a, b, c, d, r;
    a.alloc(h->nw); b.alloc(h->nq); c.alloc(h->nq); d.alloc(h->nbc); r.alloc(h->nw);
    a.upload(omega, h->nw, h->stream);
    b.upload(q, h->nq, h->stream);
    c.upload(q0, h->nq, h->stream);
    d.upload(bc, h->nbc, h->stream);
    hipLaunchKernelGGL(k_rot_nonlinear<double>, dim3((h->nw + 255) / 256), dim3(256), 0, h->stream, h->geo, a.p,
                       b.p, c.p, d.p, r.p);
    r.download(out, h->nw, h->stream);
    HIPCHK(hipStreamSynchronize(h->stream));
  });
}

int ibfs_op_vort2flux(ibfs_handle* h, double* omega, double* vel, int nlev) {
  return guarded(h, [&] {
    if (nlev < 0 || nlev > h->L) throw ArgError{IBFS_EINVAL, "nlev out of range"};
    const size_t W = (size_t)h->nw * h->L;
    DevBuf<double> om, s, v, t1, t2;
    om.alloc(W); s.alloc(W); v.alloc((size_t)h->nq * std::max(nlev, 1)); t1.alloc(h->nw); t2.alloc(h->nw);
    om.upload(omega, W, h->stream);
    h->vort2flux(om.p, s.p, nlev > 0 ? v.p : nullptr, 0, nlev, 1, t1.p, t2.p);
    om.download(omega, W, h->stream);
    if (nlev > 0 && vel) v.download(vel, (size_t)h->nq * nlev, h->stream);
    HIPCHK(hipStreamSynchronize(h->stream));
  });
}

int ibfs_op_reg(ibfs_handle* h, const double* h0, double* hq) {
  return guarded(h, [&] {
    DevBuf<double> a, b;
    a.alloc(std::max(h->nf, 1)); b.alloc(h->nq);
    a.upload(h0, h->nf, h->stream);
    b.zero(h->stream);
    if (h->sm.nact)
      hipLaunchKernelGGL(k_reg_active<double>, dim3((h->sm.nact + 255) / 256), dim3(256), 0, h->stream, h->sm, a.p,
                         b.p);
    b.download(hq, h->nq, h->stream);
    HIPCHK(hipStreamSynchronize(h->stream));
  });
}

int ibfs_op_regT(ibfs_handle* h, const double* hq, double* h0) {
  return guarded(h, [&] {
    DevBuf<double> a, b;
    a.alloc(h->nq); b.alloc(std::max(h->nf, 1));
    a.upload(hq, h->nq, h->stream);
    if (h->nf)
      hipLaunchKernelGGL(k_regT<double>, dim3((h->nf + 255) / 256), dim3(256), 0, h->stream, h->sm, a.p, b.p);
    b.download(h0, h->nf, h->stream);
    HIPCHK(hipStreamSynchronize(h->stream));
  });
}

int ibfs_op_a_times(ibfs_handle* h, const double* x, double* out) {
  return guarded(h, [&] {
    const int nf = h->nf;
    const long bsL = (long)h->L * h->nw;
    DevBuf<double> F, RRb, t1, t2, VB, SB, o;
    F.alloc(std::max(nf, 1)); RRb.alloc(h->nw); t1.alloc(h->nw); t2.alloc(h->nw);
    VB.alloc(bsL); SB.alloc(bsL); o.alloc(std::max(nf, 1));
    RRb.zero(h->stream); VB.zero(h->stream);
    F.upload(x, nf, h->stream);
    hipLaunchKernelGGL(k_rotreg<double>, dim3((h->sm.nvtx + 255) / 256, 1), dim3(256), 0, h->stream, h->sm, F.p,
                       0L, RRb.p, 0L);
    h->ctci_like(SrcField<double>{RRb.p, 0}, h->lam1i.p, true, DstField<double>{VB.p, 0}, t1.p, t2.p, 1, 0);
    h->vort2flux(VB.p, SB.p, nullptr, 0, 0, 1, t1.p, t2.p);
    hipLaunchKernelGGL(k_regT_curl<double>, dim3((nf + 255) / 256, 1), dim3(256), 0, h->stream, h->geo, h->sm, SB.p,
                       h->L > 1 ? SB.p + h->nw : nullptr, 0L, (const double*)nullptr, o.p, 0L);
    o.download(out, nf, h->stream);
    HIPCHK(hipStreamSynchronize(h->stream));
  });
}

int ibfs_op_cholsl(ibfs_handle* h, const double* b, double* x) {
  return guarded(h, [&] {
    if (!h->have_chol) throw ArgError{IBFS_ESTATE, "no Cholesky factor"};
    const int nf = h->nf;
    DevBuf<double> r, y, o;
    r.alloc(std::max(nf, 1)); y.alloc(std::max(nf, 1)); o.alloc(std::max(nf, 1));
    r.upload(b, nf, h->stream);
    hipLaunchKernelGGL(k_trmv_lower, dim3((nf + 3) / 4), dim3(256), 0, h->stream, nf, h->Linv.p, r.p, y.p);
    hipLaunchKernelGGL(k_trmv_upper_t, dim3((nf + 3) / 4), dim3(256), 0, h->stream, nf, h->Linv.p, y.p, o.p);
    o.download(x, nf, h->stream);
    HIPCHK(hipStreamSynchronize(h->stream));
  });
}

const char* ibfs_last_error(const ibfs_handle* h) { return h ? h->err.c_str() : "null handle"; }

int ibfs_destroy(ibfs_handle* h) {
  if (!h) return IBFS_EINVAL;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  delete h;
  return IBFS_OK;
}

}  // extern "C"
