// ibfs_engine.hpp -- host side of the step for one group of cases that share
// the grid size m x n: device memory, set-up tables and the launch schedule
// of one time step (advance, reference src/operators.f90:62-188).
//
// Engine<double> is the plain FP64 form.  Engine<float> is the mixed form:
// fields, transforms, the nonlinear term and the smeared fields in FP32; the
// force system (right-hand side accumulation, Cholesky solve, one refinement
// sweep) and every reduction in FP64.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "ibfs_kernels.hpp"

namespace ibfs {

struct HipError {
  std::string msg;
};
struct ArgError {
  int code;
  std::string msg;
};

#define HIPCHK(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) throw ::ibfs::HipError{std::string(#expr) + ": " + hipGetErrorString(e_)}; \
  } while (0)

constexpr int kEmax = 4096;                       // complex values per workgroup, generic pass
constexpr double kPiRef = 3.1415926535897932384;  // src/parameters.f90:35

template <typename V>
struct DevBuf {
  V* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  void alloc(size_t count) {
    release();
    n = count;
    if (count) HIPCHK(hipMalloc(&p, sizeof(V) * count));
  }
  void zero(hipStream_t s) {
    if (n) HIPCHK(hipMemsetAsync(p, 0, sizeof(V) * n, s));
  }
  void upload(const V* h, size_t count, hipStream_t s, size_t off = 0) {
    if (count) HIPCHK(hipMemcpyAsync(p + off, h, sizeof(V) * count, hipMemcpyHostToDevice, s));
  }
  void download(V* h, size_t count, hipStream_t s, size_t off = 0) const {
    if (count) HIPCHK(hipMemcpyAsync(h, p + off, sizeof(V) * count, hipMemcpyDeviceToHost, s));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
};

template <typename V>
void upvec(DevBuf<V>& d, const std::vector<V>& hv, hipStream_t s) {
  d.alloc(std::max<size_t>(hv.size(), 1));
  d.upload(hv.data(), hv.size(), s);
}

// deltafnc (src/grid.f90:233-249), Roma, Peskin & Berger (1999)
inline double deltafnc(double r, double r0, double dr) {
  const double d = r - r0;
  if (std::fabs(d) <= 0.5 * dr) return (1.0 + std::sqrt(-3.0 * (d / dr) * (d / dr) + 1.0)) / (3.0 * dr);
  if (std::fabs(d) <= 1.5 * dr) {
    const double a = 1.0 - std::fabs(d / dr);
    return (5.0 - 3.0 * std::fabs(d / dr) - std::sqrt(-3.0 * a * a + 1.0)) / (6.0 * dr);
  }
  return 0.0;
}

// One case as given at the C-ABI.
struct CaseDesc {
  int m, n, mgridlev, nb;
  double dt, re, len, offx, offy;
  std::vector<double> xb, yb;
  std::vector<int> bdyno;
  int unif_flow;
};

// Per-case data derived on the host.
struct CaseHost {
  CaseDesc d;
  int L = 0, nf = 0, nbody = 0, bmin = 0, bmax = 0, nsmear = 0, ncol = 0;
  double delta = 0, normalize = 0;
  std::vector<double> con1, con2, vfac;
  // compact smearing tables (case-local indices)
  std::vector<int> face_id, face_ptr, face_pt, frc_ptr, frc_face, vtx_id, vtx_face;
  // per force entry: curl(s) at an interior face is s[frc_ia] - s[frc_ib]
  // (vertex indices); -1 marks a face on the domain edge (get_bc terms)
  std::vector<int> frc_ia, frc_ib;
  std::vector<double> face_w, frc_w, wsum;
  // force system
  bool have_chol = false;
  std::vector<double> cholmat, cholvec;
};

inline int nlevmax(int m, int n) {
  // src/grid.f90:79 (single-precision LOG, as in the reference)
  return (int)std::floor((double)logf((float)(std::min(m, n) / 2)) / std::log(2.0)) - 1;
}

// Compact smearing tables: the entries of setup_reg (src/grid.f90:142-230),
// found per IB point inside its support box instead of by the O(Nq Nb) scan,
// with the same coordinates, the same strict support test and the same
// weight expression, kept in the reference's per-face point order.
inline void build_smear(CaseHost& c) {
  const int m = c.d.m, n = c.d.n, nb = c.d.nb;
  const double delta = c.delta, d2 = 0.5 * delta, sup = 3.0 * d2;
  std::map<int, std::vector<std::pair<int, double>>> faces;  // face -> (force, w)
  for (int k = 0; k < nb; k++) {
    const double xk = c.d.xb[k], yk = c.d.yb[k];
    const int ic = (int)std::floor((xk + c.d.offx) / delta) + 1;
    const int jc = (int)std::floor((yk + c.d.offy) / delta) + 1;
    for (int i = std::max(1, ic - 3); i <= std::min(m + 1, ic + 3); i++)
      for (int j = std::max(1, jc - 3); j <= std::min(n, jc + 3); j++) {
        const double x = delta * (double)(i - 1) - c.d.offx;
        const double y = delta * (double)(j - 1) - c.d.offy + d2;
        if (std::fabs(x - xk) < sup && std::fabs(y - yk) < sup)
          faces[(i - 1) + (j - 1) * (m + 1)].push_back(
              {k, (delta * delta) * deltafnc(x, xk, delta) * deltafnc(y, yk, delta)});
      }
    for (int i = std::max(1, ic - 3); i <= std::min(m, ic + 3); i++)
      for (int j = std::max(1, jc - 3); j <= std::min(n + 1, jc + 3); j++) {
        const double x = delta * (double)(i - 1) - c.d.offx + d2;
        const double y = delta * (double)(j - 1) - c.d.offy;
        if (std::fabs(x - xk) < sup && std::fabs(y - yk) < sup)
          faces[(m + 1) * n + (i - 1) + (j - 1) * m].push_back(
              {nb + k, (delta * delta) * deltafnc(x, xk, delta) * deltafnc(y, yk, delta)});
      }
  }
  c.nsmear = 0;
  c.face_id.clear(); c.face_ptr.assign(1, 0); c.face_pt.clear(); c.face_w.clear();
  std::vector<std::vector<std::pair<int, double>>> frc(c.nf);
  std::map<int, int> slot_of_face;
  for (auto& kv : faces) {
    slot_of_face[kv.first] = (int)c.face_id.size();
    c.face_id.push_back(kv.first);
    for (auto& e : kv.second) {
      c.face_pt.push_back(e.first);
      c.face_w.push_back(e.second);
      frc[e.first].push_back({kv.first, e.second});
    }
    c.face_ptr.push_back((int)c.face_pt.size());
    c.nsmear = std::max(c.nsmear, (int)kv.second.size());
  }
  c.frc_ptr.assign(1, 0); c.frc_face.clear(); c.frc_w.clear(); c.frc_ia.clear(); c.frc_ib.clear();
  const int nu = (m + 1) * n;
  auto w_at = [&](int i, int j) { return (i - 2) + (j - 2) * (m - 1); };
  // the two vertices of curl at an interior face (src/operators.f90:254-286)
  auto curl_pair = [&](int face, int& a, int& b) {
    a = b = -1;
    if (face < nu) {  // u(i,j) = s(i,j+1) - s(i,j)
      const int i = face % (m + 1) + 1, j = face / (m + 1) + 1;
      if (i >= 2 && i <= m && j >= 2 && j <= n - 1) { a = w_at(i, j + 1); b = w_at(i, j); }
    } else {          // v(i,j) = s(i,j) - s(i+1,j)
      const int fv = face - nu, i = fv % m + 1, j = fv / m + 1;
      if (i >= 2 && i <= m - 1 && j >= 2 && j <= n) { a = w_at(i, j); b = w_at(i + 1, j); }
    }
  };
  c.wsum.assign(std::max(c.nf, 1), 0.0);
  for (int f = 0; f < c.nf; f++) {
    for (auto& e : frc[f]) {
      c.frc_face.push_back(e.first);
      c.frc_w.push_back(e.second);
      int a, b;
      curl_pair(e.first, a, b);
      c.frc_ia.push_back(a);
      c.frc_ib.push_back(b);
      c.wsum[f] += e.second;
    }
    c.frc_ptr.push_back((int)c.frc_face.size());
  }
  std::map<int, std::array<int, 4>> vtx;
  auto touch = [&](int i, int j, int k, int a) {
    if (i < 2 || i > m || j < 2 || j > n) return;
    const int w = (i - 2) + (j - 2) * (m - 1);
    auto it = vtx.find(w);
    if (it == vtx.end()) it = vtx.emplace(w, std::array<int, 4>{-1, -1, -1, -1}).first;
    it->second[k] = a;
  };
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
  c.vtx_id.clear(); c.vtx_face.clear();
  for (auto& kv : vtx) {
    c.vtx_id.push_back(kv.first);
    for (int k = 0; k < 4; k++) c.vtx_face.push_back(kv.second[k]);
  }
}

inline CaseHost make_case(const CaseDesc& d) {
  CaseHost c;
  c.d = d;
  c.L = std::min(d.mgridlev, nlevmax(d.m, d.n));
  if (c.L < 1) throw ArgError{1, "grid too small for one level"};
  c.nf = 2 * d.nb;
  c.delta = d.len / (double)(float)d.m;  // src/grid.f90:64
  c.normalize = 4.0 * (double)(float)((long)d.m * (long)d.n);
  c.bmin = c.bmax = 0;
  for (int k = 0; k < d.nb; k++) {
    const int b = d.bdyno[k];
    if (k == 0 || b < c.bmin) c.bmin = b;
    if (k == 0 || b > c.bmax) c.bmax = b;
  }
  c.nbody = c.bmax - c.bmin + 1;
  c.ncol = 2 * c.nbody + c.L + 1;
  // factors of setup_fft (src/myfft.f90:45-51), same expressions and order
  const double del2 = c.delta * c.delta;
  c.con1.resize(c.L); c.con2.resize(c.L); c.vfac.resize(c.L);
  for (int k = 1; k <= c.L; k++) {
    const double del22 = del2 * std::pow(4.0, k - 1);
    c.vfac[k - 1] = 0.5 * d.dt / d.re / del22 / c.normalize;
    c.con1[k - 1] = 1.5 * d.dt / del22 / c.normalize;
    c.con2[k - 1] = -0.5 * d.dt / del22 / c.normalize;
  }
  build_smear(c);
  return c;
}

bool factor_235(int n, std::vector<int>& rad);

// Interface of a group, independent of the field type.
struct GroupBase {
  virtual ~GroupBase() = default;
  int B = 0, m = 0, n = 0, Lmax = 0, nw = 0, nq = 0, nbc = 0, rowlen = 0;
  long nwp = 0;  // entries of one pair-major spectral table (pm_size)
  std::vector<CaseHost> cs;
  std::vector<int> hoff;  // row offset of each case in the group's diagnostics row
  virtual void enqueue_step(int first) = 0;
  // one step with the conjugate-gradient force solve; returns its iterations
  virtual int step_cg(int first, double cgtol, int cg_max_iter) = 0;
  // new point positions of case c and the rebuilt smearing tables
  virtual void move_body(int c, const double* xb, const double* yb) = 0;
  virtual void enqueue_motion() = 0;
  // actuator slots of case c and the step counter their jets are timed by
  virtual void set_actuators(int c, int n_act, const int* npt, const double* par) = 0;
  virtual bool has_actuators() const = 0;
  virtual void set_itime(int itime) = 0;
  virtual void set_motion(int c, const double* motion) = 0;
  virtual void set_state(int c, const double* om, const double* ro, const double* q, const double* q0,
                         const double* fb) = 0;
  virtual void get_state(int c, double* om, double* ro, double* q, double* q0, double* fb) = 0;
  virtual void upload_factor(int c) = 0;
  virtual void reset_counters() = 0;
  // defer each step's diagnostics row to the next step's first pass, and
  // write the rows still pending
  virtual void set_side_diag(bool on) = 0;
  virtual void flush_diag() = 0;
  virtual void ensure_diag(int cap) = 0;
  virtual void download_diag(double* dst, int nsteps, int ldrow, int col0) = 0;
  virtual bool is_double() const = 0;
  // FFTs per workgroup chosen by the tuner: row, column, column pair,
  // step-1 row, step-1 column pair (0: generic pass)
  virtual void tile_shapes(int* out) const = 0;
  virtual void level_tile_shapes(int* out) const = 0;
};

}  // namespace ibfs
