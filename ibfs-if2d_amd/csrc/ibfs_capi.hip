// ibfs_capi.hip -- the C-ABI declared in include/ibfs.h.  A handle owns one
// HIP stream and one or more groups of cases (one group per grid size); a
// single case is an ensemble of one.  Every step after the first is replayed
// from one captured HIP graph covering all groups.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#include "../../include/ibfs.h"
#include "dst_gemm.hpp"
#include "ibfs_adjoint_host.hpp"
#include "ibfs_engine_t.hpp"

using namespace ibfs;

namespace ibfs {
bool factor_235(int n, std::vector<int>& rad) {
  rad.clear();
  int r = n;
  while (r % 4 == 0) { rad.push_back(4); r /= 4; }
  while (r % 2 == 0) { rad.push_back(2); r /= 2; }
  while (r % 3 == 0) { rad.push_back(3); r /= 3; }
  while (r % 5 == 0) { rad.push_back(5); r /= 5; }
  return r == 1 && (int)rad.size() <= kMaxStages;
}
}  // namespace ibfs

// The forward sweep of an adjoint run (ibfs_adjoint_forward), kept in the
// handle for the reverse sweep: the state (omega, rhs_old, q) every ckpt
// steps, the final state and the motion every step used.
struct TapeBase {
  virtual ~TapeBase() = default;
  int N = 0, ckpt = 1, nck = 0, it0 = 0;
  std::vector<int> ck_it;
  std::vector<double> mo;  // delta * u_B, concatenated per case
};
template <typename T>
struct Tape : TapeBase {
  DevBuf<T> CW, CRO, CQ, FW, FRO, FQ;
};

struct ibfs_handle {
  int precision = 0;
  std::string err;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  std::vector<std::unique_ptr<GroupBase>> groups;
  std::vector<std::pair<int, int>> where;  // case -> (group, index in group)
  std::vector<int> group_col0;             // offset of each group's block in a history row
  int rowlen = 0;
  int ncases = 0;
  int diag_cap = 0;
  bool use_graph = true;  // IBFS_NO_GRAPH=1: eager launches (for counter-collection profiling)
  // force solve (src/parameters.f90:19-20,31): the factor (stationary) or
  // cjgr warm-started from the previous fb, with the CG iterations of each
  // step of the last advance call
  int stationary = 1;
  double cgtol = 1.0e-7;
  int cg_max_iter = 3000;
  std::vector<int> cg_iters;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  // the forward sweep of ibfs_adjoint_forward, kept until ibfs_adjoint_reverse
  std::unique_ptr<TapeBase> tape;

  ~ibfs_handle() {
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graph();
    groups.clear();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
  }
  CaseHost& ch(int c) { return groups[where[c].first]->cs[where[c].second]; }
  GroupBase& gr(int c) { return *groups[where[c].first]; }
  Engine<double>* e64() { return dynamic_cast<Engine<double>*>(groups[0].get()); }
  bool all_factors() {
    for (auto& g : groups)
      for (auto& c : g->cs)
        if (!c.have_chol && c.nf > 0) return false;
    return true;
  }
  void ensure_diag(int cap) {
    if (cap <= diag_cap) return;
    HIPCHK(hipStreamSynchronize(stream));
    drop_graph();  // the graph holds the old buffer
    for (auto& g : groups) g->ensure_diag(cap);
    diag_cap = cap;
  }
  void enqueue_all(int first) {
    for (auto& g : groups) g->enqueue_step(first);
  }
  void capture() {
    if (gexec) return;
    HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    enqueue_all(0);
    HIPCHK(hipStreamEndCapture(stream, &graph));
    HIPCHK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
  }
  // builds the factor of case c on the device (FP64 even for the mixed form)
  void build_factor(int c) {
    CaseHost& h = ch(c);
    const int nf = h.nf;
    if (nf == 0) { h.have_chol = true; return; }
    GroupBase& g = gr(c);
    Engine<double>* ed = dynamic_cast<Engine<double>*>(&g);
    std::unique_ptr<Engine<double>> tmp;
    int idx = where[c].second;
    if (!ed) {  // mixed form: a one-case FP64 engine builds A
      tmp.reset(new Engine<double>(std::vector<CaseHost>{h}, stream));
      ed = tmp.get();
      idx = 0;
    }
    DevBuf<double> A;
    A.alloc((size_t)nf * nf);
    ed->build_columns(idx, A.p);
    h.cholmat.assign((size_t)nf * nf, 0.0);
    h.cholvec.assign(nf, 0.0);
    A.download(h.cholmat.data(), (size_t)nf * nf, stream);
    HIPCHK(hipStreamSynchronize(stream));
    // choldc (src/operators.f90:766-791): upper triangle keeps A, strict
    // lower gets L, diagonal in cholvec
    double* a = h.cholmat.data();
    auto CM = [&](int i, int j) -> double& { return a[(size_t)(i - 1) + (size_t)(j - 1) * nf]; };
    for (int i = 1; i <= nf; i++)
      for (int j = i; j <= nf; j++) {
        double sum = CM(i, j);
        for (int k = i - 1; k >= 1; k--) sum = sum - CM(i, k) * CM(j, k);
        if (i == j) {
          if (sum <= 0.0) throw ArgError{IBFS_ECHOL, "choldc failed at row " + std::to_string(i)};
          h.cholvec[i - 1] = std::sqrt(sum);
        } else {
          CM(j, i) = sum / h.cholvec[i - 1];
        }
      }
    g.upload_factor(where[c].second);
  }
};

namespace {

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

bool valid_case(const CaseDesc& d) {
  return d.m % 4 == 0 && d.n % 4 == 0 && d.m >= 8 && d.n >= 8 && d.mgridlev >= 1 && d.nb >= 0 && d.dt > 0 &&
         d.re > 0 && d.len > 0;
}

int create(std::vector<CaseDesc> cases, int precision, ibfs_handle** out) {
  if (!out) return IBFS_EINVAL;
  *out = nullptr;
  if (cases.empty() || (precision != IBFS_PLAIN_FP64 && precision != IBFS_MIXED)) return IBFS_EINVAL;
  for (auto& d : cases)
    if (!valid_case(d)) return IBFS_EINVAL;
  ibfs_handle* h = new ibfs_handle();
  *out = h;
  return guarded(h, [&] {
    h->precision = precision;
    h->ncases = (int)cases.size();
    HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&h->ev0));
    HIPCHK(hipEventCreate(&h->ev1));
    // group by grid size, keeping the input order inside each group
    std::map<std::pair<int, int>, std::vector<int>> by_size;
    std::vector<std::pair<int, int>> order;
    for (int c = 0; c < (int)cases.size(); c++) {
      auto key = std::make_pair(cases[c].m, cases[c].n);
      if (!by_size.count(key)) order.push_back(key);
      by_size[key].push_back(c);
    }
    h->where.resize(cases.size());
    int col0 = 0;
    for (auto& key : order) {
      std::vector<CaseHost> hs;
      for (int c : by_size[key]) {
        h->where[c] = {(int)h->groups.size(), (int)hs.size()};
        hs.push_back(make_case(cases[c]));
      }
      if (precision == IBFS_PLAIN_FP64) h->groups.emplace_back(new Engine<double>(std::move(hs), h->stream));
      else h->groups.emplace_back(new Engine<float>(std::move(hs), h->stream));
      h->group_col0.push_back(col0);
      col0 += h->groups.back()->rowlen;
    }
    h->rowlen = col0;
    h->diag_cap = 1;
    const char* ng = std::getenv("IBFS_NO_GRAPH");
    h->use_graph = !(ng && ng[0] == '1');
  });
}

CaseDesc make_desc(int m, int n, int mgridlev, int nb, double dt, double re, double len, double offx, double offy,
                   const double* xb, const double* yb, const int* bdyno, int unif_flow) {
  CaseDesc d{m, n, mgridlev, nb, dt, re, len, offx, offy, {}, {}, {}, unif_flow};
  if (nb > 0) {
    d.xb.assign(xb, xb + nb);
    d.yb.assign(yb, yb + nb);
  }
  d.bdyno.assign(std::max(nb, 0), 0);
  if (bdyno)
    for (int k = 0; k < nb; k++) d.bdyno[k] = bdyno[k];
  return d;
}

}  // namespace

// the message of the calling thread's last failed handle-free call
const char* noalloc_last_error();

// ============================================================================
// C-ABI
// ============================================================================
extern "C" {

int ibfs_create(int m, int n, int mgridlev, int nb, double dt, double re, double len, double offsetx,
                double offsety, const double* xb, const double* yb, const int* bdyno, int unif_flow, int precision,
                ibfs_handle** out) {
  if (nb > 0 && (!xb || !yb)) return IBFS_EINVAL;
  return create({make_desc(m, n, mgridlev, nb, dt, re, len, offsetx, offsety, xb, yb, bdyno, unif_flow)}, precision,
                out);
}

int ibfs_ensemble_create(int ncases, const int* m, const int* n, const int* mgridlev, const int* nb,
                         const double* dt, const double* re, const double* len, const double* offsetx,
                         const double* offsety, const double* xb, const double* yb, const int* bdyno, int unif_flow,
                         int precision, ibfs_handle** out) {
  if (ncases < 1 || !m || !n || !mgridlev || !nb || !dt || !re || !len || !offsetx || !offsety) return IBFS_EINVAL;
  std::vector<CaseDesc> cs;
  long p = 0;
  for (int c = 0; c < ncases; c++) {
    if (nb[c] > 0 && (!xb || !yb)) return IBFS_EINVAL;
    cs.push_back(make_desc(m[c], n[c], mgridlev[c], nb[c], dt[c], re[c], len[c], offsetx[c], offsety[c],
                           xb ? xb + p : nullptr, yb ? yb + p : nullptr, bdyno ? bdyno + p : nullptr, unif_flow));
    p += nb[c];
  }
  return create(std::move(cs), precision, out);
}

int ibfs_case_dims(const ibfs_handle* hc, int c, int* d) {
  if (!hc || !d) return IBFS_EINVAL;
  ibfs_handle* h = const_cast<ibfs_handle*>(hc);
  if (c < 0 || c >= h->ncases) return IBFS_EINVAL;
  const CaseHost& k = h->ch(c);
  const GroupBase& g = h->gr(c);
  const int v[12] = {k.d.m, k.d.n, k.L, k.d.nb, k.nf, g.nq, g.nw, g.nbc, k.nbody, k.bmin, k.bmax, k.nsmear};
  std::memcpy(d, v, sizeof(v));
  return IBFS_OK;
}

int ibfs_dims(const ibfs_handle* h, int* d) { return ibfs_case_dims(h, 0, d); }

int ibfs_ensemble_info(const ibfs_handle* h, int* ncases, int* rowlen, int* col0) {
  if (!h) return IBFS_EINVAL;
  if (ncases) *ncases = h->ncases;
  if (rowlen) *rowlen = h->rowlen;
  if (col0) {
    ibfs_handle* hh = const_cast<ibfs_handle*>(h);
    for (int c = 0; c < h->ncases; c++)
      col0[c] = h->group_col0[h->where[c].first] + hh->gr(c).hoff[h->where[c].second];
  }
  return IBFS_OK;
}

int ibfs_case_set_state(ibfs_handle* h, int c, const double* omega, const double* rhs_old, const double* q,
                        const double* q0, const double* fb) {
  return guarded(h, [&] {
    if (c < 0 || c >= h->ncases) throw ArgError{IBFS_EINVAL, "case index out of range"};
    h->gr(c).set_state(h->where[c].second, omega, rhs_old, q, q0, fb);
    if (q0) h->drop_graph();  // the step's kernels depend on the form of q0 (freestream or general)
  });
}
int ibfs_case_get_state(ibfs_handle* h, int c, double* omega, double* rhs_old, double* q, double* q0, double* fb) {
  return guarded(h, [&] {
    if (c < 0 || c >= h->ncases) throw ArgError{IBFS_EINVAL, "case index out of range"};
    h->gr(c).get_state(h->where[c].second, omega, rhs_old, q, q0, fb);
  });
}
int ibfs_set_state(ibfs_handle* h, const double* omega, const double* rhs_old, const double* q, const double* q0,
                   const double* fb) {
  return ibfs_case_set_state(h, 0, omega, rhs_old, q, q0, fb);
}
int ibfs_get_state(ibfs_handle* h, double* omega, double* rhs_old, double* q, double* q0, double* fb) {
  return ibfs_case_get_state(h, 0, omega, rhs_old, q, q0, fb);
}

int ibfs_case_set_cholesky(ibfs_handle* h, int c, const double* cholmat, const double* cholvec) {
  return guarded(h, [&] {
    if (c < 0 || c >= h->ncases) throw ArgError{IBFS_EINVAL, "case index out of range"};
    if (!cholmat || !cholvec) throw ArgError{IBFS_EINVAL, "null factor"};
    CaseHost& k = h->ch(c);
    const size_t nf = k.nf;
    k.cholmat.assign(cholmat, cholmat + nf * nf);
    k.cholvec.assign(cholvec, cholvec + nf);
    for (size_t i = 0; i < nf; i++)
      if (!(k.cholvec[i] > 0.0)) throw ArgError{IBFS_ECHOL, "cholvec has a non-positive entry"};
    h->gr(c).upload_factor(h->where[c].second);
  });
}
int ibfs_set_cholesky(ibfs_handle* h, const double* cholmat, const double* cholvec) {
  return ibfs_case_set_cholesky(h, 0, cholmat, cholvec);
}

int ibfs_case_build_cholesky(ibfs_handle* h, int c, double* cholmat, double* cholvec) {
  return guarded(h, [&] {
    if (c < 0 || c >= h->ncases) throw ArgError{IBFS_EINVAL, "case index out of range"};
    h->build_factor(c);
    const CaseHost& k = h->ch(c);
    if (cholmat) std::memcpy(cholmat, k.cholmat.data(), sizeof(double) * k.cholmat.size());
    if (cholvec) std::memcpy(cholvec, k.cholvec.data(), sizeof(double) * k.cholvec.size());
  });
}
int ibfs_build_cholesky(ibfs_handle* h, double* cholmat, double* cholvec) {
  return ibfs_case_build_cholesky(h, 0, cholmat, cholvec);
}
int ibfs_case_set_actuators(ibfs_handle* h, int c, int n_act, const int* npt, const double* par) {
  return guarded(h, [&] {
    if (c < 0 || c >= h->ncases) throw ArgError{IBFS_EINVAL, "case index out of range"};
    if (n_act < 0 || (n_act > 0 && (!npt || !par))) throw ArgError{IBFS_EINVAL, "bad actuator arguments"};
    HIPCHK(hipStreamSynchronize(h->stream));
    h->drop_graph();  // the captured step has no actuation launch, or old tables
    h->gr(c).set_actuators(h->where[c].second, n_act, npt, par);
  });
}
int ibfs_set_actuators(ibfs_handle* h, int n_act, const int* npt, const double* par) {
  return ibfs_case_set_actuators(h, 0, n_act, npt, par);
}
int ibfs_ensemble_build_cholesky(ibfs_handle* h) {
  return guarded(h, [&] {
    for (int c = 0; c < h->ncases; c++) h->build_factor(c);
  });
}

}  // extern "C"

namespace {
// nsteps steps of every case; motion (concatenated per case, scaled by
// delta) is set first.  The body of ibfs_advance.
void advance_impl(ibfs_handle* h, int* itime, int nsteps, const double* motion, double* hist) {
  if (!h->stationary) {
    if (h->ncases != 1 || !h->e64())
      throw ArgError{IBFS_EINVAL, "the conjugate-gradient force solve needs a one-case plain FP64 handle"};
  } else if (!h->all_factors())
    throw ArgError{IBFS_ESTATE, "no Cholesky factor: call ibfs_build_cholesky or ibfs_set_cholesky before the "
                                "first step"};
  if (nsteps == 0) return;
  hipStream_t s = h->stream;
  long p = 0;
  if (motion) {
    long nft = 0;
    for (int c = 0; c < h->ncases; c++) nft += h->ch(c).nf;
    for (long f = 0; f < nft; f++)
      if (!std::isfinite(motion[f])) throw ArgError{IBFS_EINVAL, "motion has a non-finite entry"};
  }
  h->ensure_diag(nsteps);
  for (int c = 0; c < h->ncases; c++) {
    h->gr(c).set_motion(h->where[c].second, motion ? motion + p : nullptr);
    p += h->ch(c).nf;
  }
  for (auto& g : h->groups) {
    g->reset_counters();
    g->enqueue_motion();
    g->set_itime(*itime);
  }
  HIPCHK(hipEventRecord(h->ev0, s));
  if (!h->stationary) h->cg_iters.assign(nsteps, 0);
  // each step's diagnostics row is written by the next step's first pass;
  // the last one is flushed after the loop
  for (auto& g : h->groups) g->set_side_diag(true);
  struct SideDiagOff {  // other paths (the adjoint's forward steps) write their rows in the step
    ibfs_handle* h;
    ~SideDiagOff() {
      for (auto& g : h->groups) g->set_side_diag(false);
    }
  } side_diag_off{h};
  for (int k = 0; k < nsteps; k++) {
    const int it = *itime + 1;
    if (!h->stationary) {
      h->cg_iters[k] = h->groups[0]->step_cg(it == 1, h->cgtol, h->cg_max_iter);
    } else if (it == 1) {
      h->enqueue_all(1);
    } else if (h->use_graph) {
      h->capture();
      HIPCHK(hipGraphLaunch(h->gexec, s));
    } else {
      h->enqueue_all(0);
    }
    *itime = it;
  }
  for (auto& g : h->groups) g->flush_diag();
  HIPCHK(hipEventRecord(h->ev1, s));
  HIPCHK(hipStreamSynchronize(s));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->last_ms = ms;
  if (hist)
    for (size_t g = 0; g < h->groups.size(); g++)
      h->groups[g]->download_diag(hist, nsteps, h->rowlen, h->group_col0[g]);
}

// motion = delta * u_B per case (u_B concatenated per case)
std::vector<double> scaled_motion(ibfs_handle* h, const double* ub) {
  std::vector<double> mo;
  long p = 0;
  for (int c = 0; c < h->ncases; c++) {
    const CaseHost& k = h->ch(c);
    for (int f = 0; f < k.nf; f++) mo.push_back(ub ? k.delta * ub[p + f] : 0.0);
    p += k.nf;
  }
  return mo;
}

// adds (1/nwin) * (cd * sum fx + cl * sum fy) of the rows inside the window
void accumulate_objective(ibfs_handle* h, const double* hist, int nrows, int it_first, int win_start, int nwin,
                          double cd, double cl, double* J) {
  std::vector<int> col0(h->ncases);
  ibfs_ensemble_info(h, nullptr, nullptr, col0.data());
  for (int r = 0; r < nrows; r++) {
    if (it_first + r <= win_start) continue;  // window = steps win_start+1 .. win_start+nwin
    const double* row = hist + (size_t)r * h->rowlen;
    for (int c = 0; c < h->ncases; c++) {
      const CaseHost& k = h->ch(c);
      double s = 0.0;
      for (int b = 0; b < k.nbody; b++) s += cd * row[col0[c] + b] + cl * row[col0[c] + k.nbody + b];
      J[c] += s / nwin;
    }
  }
}

// the one group of a handle whose cases share a grid size
GroupBase* single_group(ibfs_handle* h) {
  if (h->groups.size() != 1) throw ArgError{IBFS_EINVAL, "tangent/adjoint runs need all cases of one grid size"};
  // the control u_B drives every point; actuator jets would override it
  if (h->groups[0]->has_actuators()) throw ArgError{IBFS_EINVAL, "tangent/adjoint runs take no actuators"};
  return h->groups[0].get();
}

// tangent-linear run of ibfs_tangent_linear on the single group E
template <typename T>
void tangent_impl(ibfs_handle* h, Engine<T>* E, int* itime, int nspin, int nwin, const double* ub, const double* dub,
                  double cd, double cl, double* J, double* dJ) {
  std::vector<CaseHost> tc = E->cs;
  for (auto& c : tc) c.d.unif_flow = 0;  // the perturbation has no underlying potential flow
  Engine<T> Tg(tc, h->stream);
  Tg.base = E;
  for (int c = 0; c < Tg.B; c++) Tg.upload_factor(c);
  std::vector<double> mo = scaled_motion(h, ub), dmo = scaled_motion(h, dub);
  const int N = nspin + nwin, it0 = *itime;
  h->ensure_diag(N);
  Tg.ensure_diag(N);
  for (int c = 0; c < E->B; c++) {
    E->set_motion(c, mo.data() + E->hcs[c].foff);
    Tg.set_motion(c, dmo.data() + Tg.hcs[c].foff);
  }
  E->reset_counters();
  Tg.reset_counters();
  E->enqueue_motion();
  Tg.enqueue_motion();
  for (int k = 0; k < N; k++) {
    const int first = (*itime + 1 == 1);
    Tg.enqueue_step(first);  // linearised about the state before the base step
    E->enqueue_step(first);
    *itime += 1;
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  std::vector<double> hb((size_t)N * h->rowlen), ht((size_t)N * h->rowlen);
  E->download_diag(hb.data(), N, h->rowlen, 0);
  Tg.download_diag(ht.data(), N, h->rowlen, 0);
  for (int c = 0; c < h->ncases; c++) J[c] = dJ[c] = 0.0;
  accumulate_objective(h, hb.data(), N, it0 + 1, it0 + nspin, nwin, cd, cl, J);
  accumulate_objective(h, ht.data(), N, it0 + 1, it0 + nspin, nwin, cd, cl, dJ);
}

// the transposes assume no smearing weight on the two outermost rings of
// faces of the finest level (bodies well inside the domain)
template <typename T>
void check_adjoint_geometry(Engine<T>* E) {
  for (int c = 0; c < E->B; c++) {
    const CaseHost& k = E->cs[c];
    const int m = k.d.m, n = k.d.n, nu = (m + 1) * n;
    for (int f : k.face_id) {
      int i, j;
      if (f < nu) { i = f % (m + 1) + 1; j = f / (m + 1) + 1; }
      else { i = (f - nu) % m + 1; j = (f - nu) / m + 1; }
      if (i < 3 || j < 3 || i > m - 2 || j > n - 2)
        throw ArgError{IBFS_EINVAL, "adjoint: a body smears onto the outer rings of the finest grid"};
    }
  }
}

// forward sweep of an adjoint run: nsteps steps with motion = delta * u_B,
// the state checkpointed every ckpt steps into the handle's tape, the
// history rows into hist (nsteps x rowlen, may be NULL)
template <typename T>
void adjoint_forward_impl(ibfs_handle* h, Engine<T>* E, int* itime, int nsteps, int ckpt, const double* ub,
                          double* hist) {
  check_adjoint_geometry(E);
  hipStream_t st = h->stream;
  auto tp = std::make_unique<Tape<T>>();
  Tape<T>& t = *tp;
  t.N = nsteps;
  t.ckpt = ckpt;
  t.nck = (nsteps + ckpt - 1) / ckpt;
  t.it0 = *itime;
  t.mo = scaled_motion(h, ub);
  const size_t W = (size_t)E->B * E->fbs(), Q = (size_t)E->B * E->qbs();
  t.CW.alloc(W * t.nck); t.CRO.alloc(W * t.nck); t.CQ.alloc(Q * t.nck);
  t.FW.alloc(W); t.FRO.alloc(W); t.FQ.alloc(Q);
  auto d2d = [&](T* dst, const T* src, size_t n) {
    HIPCHK(hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyDeviceToDevice, st));
  };
  std::vector<double> rows((size_t)ckpt * h->rowlen);
  t.ck_it.resize(t.nck);
  int it = *itime;
  for (int s = 0; s < t.nck; s++) {
    d2d(t.CW.p + W * s, E->omega.p, W);
    d2d(t.CRO.p + W * s, E->rhs_old.p, W);
    d2d(t.CQ.p + Q * s, E->q.p, Q);
    t.ck_it[s] = it;
    const int ns = std::min(ckpt, nsteps - s * ckpt);
    advance_impl(h, &it, ns, t.mo.data(), rows.data());
    if (hist) std::memcpy(hist + (size_t)s * ckpt * h->rowlen, rows.data(), sizeof(double) * ns * h->rowlen);
  }
  d2d(t.FW.p, E->omega.p, W);
  d2d(t.FRO.p, E->rhs_old.p, W);
  d2d(t.FQ.p, E->q.p, Q);
  HIPCHK(hipStreamSynchronize(st));
  *itime = it;
  h->tape = std::move(tp);
}

// reverse sweep over the handle's tape: ghist holds the cotangents of the
// tape's nsteps history rows; each segment's states are recomputed from its
// checkpoint, then the segment is stepped back.  grad = dJ/du_B.  The handle
// ends at the forward sweep's final state.
template <typename T>
void adjoint_reverse_impl(ibfs_handle* h, Engine<T>* E, const double* ghist, double* grad) {
  auto* tp = dynamic_cast<Tape<T>*>(h->tape.get());
  if (!tp) throw ArgError{IBFS_ESTATE, "no forward sweep to reverse: call ibfs_adjoint_forward first"};
  Tape<T>& t = *tp;
  hipStream_t st = h->stream;
  const int rl = h->rowlen;
  const size_t W = (size_t)E->B * E->fbs(), Q = (size_t)E->B * E->qbs();
  DevBuf<T> SW, SQ;
  SW.alloc(W * t.ckpt);
  SQ.alloc(Q * t.ckpt);
  DevBuf<double> GH;
  GH.alloc((size_t)t.N * rl);
  GH.upload(ghist, (size_t)t.N * rl, st);
  std::vector<char> nonzero(t.N, 0);
  for (int r = 0; r < t.N; r++)
    for (int c = 0; c < rl && !nonzero[r]; c++) nonzero[r] = ghist[(size_t)r * rl + c] != 0.0;
  auto d2d = [&](T* dst, const T* src, size_t n) {
    HIPCHK(hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyDeviceToDevice, st));
  };
  // the motion of the forward sweep, for the recomputed steps
  for (int c = 0; c < E->B; c++) E->set_motion(c, t.mo.data() + E->hcs[c].foff);
  E->enqueue_motion();
  AdjointRunner<T> R(*E);
  for (int s = t.nck - 1; s >= 0; s--) {
    d2d(E->omega.p, t.CW.p + W * s, W);
    d2d(E->rhs_old.p, t.CRO.p + W * s, W);
    d2d(E->q.p, t.CQ.p + Q * s, Q);
    const int ns = std::min(t.ckpt, t.N - s * t.ckpt);
    for (int j = 0; j < ns; j++) {
      d2d(SW.p + W * j, E->omega.p, W);
      d2d(SQ.p + Q * j, E->q.p, Q);
      E->enqueue_step(t.ck_it[s] + j + 1 == 1);
    }
    for (int j = ns - 1; j >= 0; j--) {
      const int n = t.ck_it[s] + j + 1;  // itime of this step
      const int r = n - t.it0 - 1;       // its row in the tape
      R.step(SW.p + W * j, SQ.p + Q * j, n == 1, nonzero[r] ? GH.p + (size_t)r * rl : nullptr);
    }
    HIPCHK(hipStreamSynchronize(st));
  }
  // back to the final forward state
  d2d(E->omega.p, t.FW.p, W);
  d2d(E->rhs_old.p, t.FRO.p, W);
  d2d(E->q.p, t.FQ.p, Q);
  R.AU.download(grad, E->nf_tot, st);
  HIPCHK(hipStreamSynchronize(st));
}

// the history-row cotangent of the mean-force objective J (cd * x forces +
// cl * y forces of every body, averaged over the window rows)
std::vector<double> mean_force_rows(ibfs_handle* h, int N, int it0, int nspin, int nwin, double cd, double cl) {
  std::vector<double> g((size_t)N * h->rowlen, 0.0);
  std::vector<int> col0(h->ncases);
  ibfs_ensemble_info(h, nullptr, nullptr, col0.data());
  for (int r = 0; r < N; r++) {
    if (it0 + r + 1 <= it0 + nspin) continue;
    for (int c = 0; c < h->ncases; c++) {
      const CaseHost& k = h->ch(c);
      for (int b = 0; b < k.nbody; b++) {
        g[(size_t)r * h->rowlen + col0[c] + b] = cd / nwin;
        g[(size_t)r * h->rowlen + col0[c] + k.nbody + b] = cl / nwin;
      }
    }
  }
  return g;
}

// checkpointed adjoint run of ibfs_adjoint_gradient on the single group E:
// the forward sweep, the mean-force cotangent, the reverse sweep
template <typename T>
void adjoint_impl(ibfs_handle* h, Engine<T>* E, int* itime, int nspin, int nwin, int ckpt, const double* ub,
                  double cd, double cl, double* J, double* grad) {
  const int N = nspin + nwin, it0 = *itime;
  std::vector<double> hist((size_t)N * h->rowlen);
  adjoint_forward_impl(h, E, itime, N, ckpt, ub, hist.data());
  for (int c = 0; c < h->ncases; c++) J[c] = 0.0;
  accumulate_objective(h, hist.data(), N, it0 + 1, it0 + nspin, nwin, cd, cl, J);
  const std::vector<double> g = mean_force_rows(h, N, it0, nspin, nwin, cd, cl);
  adjoint_reverse_impl(h, E, g.data(), grad);
  h->tape.reset();
}

// vector-Jacobian product of one step of a one-case handle (see ibfs_step_vjp)
template <typename T>
void step_vjp_impl(ibfs_handle* h, Engine<T>* E, int itime, const double* omega, const double* q,
                   const double* a_omega, const double* a_rhs_old, const double* a_q, const double* a_row,
                   double* g_omega, double* g_rhs_old, double* g_q, double* g_motion) {
  if (h->ncases != 1) throw ArgError{IBFS_EINVAL, "ibfs_step_vjp takes a one-case handle"};
  check_adjoint_geometry(E);
  hipStream_t st = h->stream;
  const size_t W = (size_t)E->fbs(), Q = (size_t)E->qbs();
  DevBuf<T> SW, SQ;
  SW.alloc(W);
  SQ.alloc(Q);
  E->put(SW, 0, omega, W);
  E->put(SQ, 0, q, Q);
  AdjointRunner<T> R(*E);
  if (a_omega) E->put(R.AW, 0, a_omega, W);
  if (a_rhs_old) E->put(R.ARO, 0, a_rhs_old, W);
  if (a_q) E->put(R.AQ, 0, a_q, Q);
  DevBuf<double> GR;
  if (a_row) {
    GR.alloc(std::max(h->rowlen, 1));
    GR.upload(a_row, h->rowlen, st);
  }
  R.step(SW.p, SQ.p, itime == 1, a_row ? GR.p : nullptr);
  if (g_omega) E->get(R.AW, 0, g_omega, W);
  if (g_rhs_old) E->get(R.ARO, 0, g_rhs_old, W);
  if (g_q) E->get(R.AQ, 0, g_q, Q);
  if (g_motion) R.Z.download(g_motion, E->nf_tot, st);
  HIPCHK(hipStreamSynchronize(st));
}

}  // namespace

extern "C" {

int ibfs_objective(ibfs_handle* h, int* itime, int nspin, int nwin, const double* ub, double cd, double cl,
                   double* J) {
  if (!h || !itime || !J || nspin < 0 || nwin < 1) return IBFS_EINVAL;
  return guarded(h, [&] {
    std::vector<double> mo = scaled_motion(h, ub);
    const int N = nspin + nwin, it0 = *itime;
    std::vector<double> hist((size_t)N * h->rowlen);
    for (int c = 0; c < h->ncases; c++) J[c] = 0.0;
    advance_impl(h, itime, N, mo.data(), hist.data());
    accumulate_objective(h, hist.data(), N, it0 + 1, it0 + nspin, nwin, cd, cl, J);
  });
}

int ibfs_tangent_linear(ibfs_handle* h, int* itime, int nspin, int nwin, const double* ub, const double* dub,
                        double cd, double cl, double* J, double* dJ) {
  if (!h || !itime || !J || !dJ || !dub || nspin < 0 || nwin < 1) return IBFS_EINVAL;
  return guarded(h, [&] {
    if (!h->all_factors()) throw ArgError{IBFS_ESTATE, "no Cholesky factor"};
    GroupBase* G = single_group(h);
    if (auto* E = dynamic_cast<Engine<double>*>(G)) tangent_impl(h, E, itime, nspin, nwin, ub, dub, cd, cl, J, dJ);
    else tangent_impl(h, dynamic_cast<Engine<float>*>(G), itime, nspin, nwin, ub, dub, cd, cl, J, dJ);
  });
}

int ibfs_adjoint_forward(ibfs_handle* h, int* itime, int nsteps, int ckpt, const double* ub, double* hist) {
  if (!h || !itime || nsteps < 1 || ckpt < 1) return IBFS_EINVAL;
  return guarded(h, [&] {
    if (!h->all_factors()) throw ArgError{IBFS_ESTATE, "no Cholesky factor"};
    h->tape.reset();
    GroupBase* G = single_group(h);
    if (auto* E = dynamic_cast<Engine<double>*>(G)) adjoint_forward_impl(h, E, itime, nsteps, ckpt, ub, hist);
    else adjoint_forward_impl(h, dynamic_cast<Engine<float>*>(G), itime, nsteps, ckpt, ub, hist);
  });
}

int ibfs_adjoint_reverse(ibfs_handle* h, const double* ghist, double* grad) {
  if (!h || !ghist || !grad) return IBFS_EINVAL;
  return guarded(h, [&] {
    GroupBase* G = single_group(h);
    if (auto* E = dynamic_cast<Engine<double>*>(G)) adjoint_reverse_impl(h, E, ghist, grad);
    else adjoint_reverse_impl(h, dynamic_cast<Engine<float>*>(G), ghist, grad);
  });
}

int ibfs_step_vjp(ibfs_handle* h, int itime, const double* omega, const double* q, const double* a_omega,
                  const double* a_rhs_old, const double* a_q, const double* a_row, double* g_omega,
                  double* g_rhs_old, double* g_q, double* g_motion) {
  if (!h || itime < 1 || !omega || !q) return IBFS_EINVAL;
  return guarded(h, [&] {
    if (!h->all_factors()) throw ArgError{IBFS_ESTATE, "no Cholesky factor"};
    GroupBase* G = single_group(h);
    if (auto* E = dynamic_cast<Engine<double>*>(G))
      step_vjp_impl(h, E, itime, omega, q, a_omega, a_rhs_old, a_q, a_row, g_omega, g_rhs_old, g_q, g_motion);
    else
      step_vjp_impl(h, dynamic_cast<Engine<float>*>(G), itime, omega, q, a_omega, a_rhs_old, a_q, a_row, g_omega,
                    g_rhs_old, g_q, g_motion);
  });
}

int ibfs_adjoint_gradient(ibfs_handle* h, int* itime, int nspin, int nwin, int ckpt, const double* ub, double cd,
                          double cl, double* J, double* grad) {
  if (!h || !itime || !J || !grad || nspin < 0 || nwin < 1 || ckpt < 1) return IBFS_EINVAL;
  return guarded(h, [&] {
    if (!h->all_factors()) throw ArgError{IBFS_ESTATE, "no Cholesky factor"};
    GroupBase* G = single_group(h);
    if (auto* E = dynamic_cast<Engine<double>*>(G)) adjoint_impl(h, E, itime, nspin, nwin, ckpt, ub, cd, cl, J, grad);
    else adjoint_impl(h, dynamic_cast<Engine<float>*>(G), itime, nspin, nwin, ckpt, ub, cd, cl, J, grad);
  });
}

int ibfs_advance(ibfs_handle* h, int* itime, int nsteps, const double* motion, double* hist) {
  if (!h || !itime || nsteps < 0) return IBFS_EINVAL;
  return guarded(h, [&] { advance_impl(h, itime, nsteps, motion, hist); });
}

int ibfs_step(ibfs_handle* h, int* itime, const double* motion, double* fbo, double* force, double* circ,
              double* slip) {
  if (!h) return IBFS_EINVAL;
  std::vector<double> row(std::max(h->rowlen, 1));
  int rc = ibfs_advance(h, itime, 1, motion, row.data());
  if (rc) return rc;
  const CaseHost& k = h->ch(0);
  if (force) std::memcpy(force, row.data(), sizeof(double) * 2 * k.nbody);
  if (circ) std::memcpy(circ, row.data() + 2 * k.nbody, sizeof(double) * k.L);
  if (slip) *slip = row[2 * k.nbody + k.L];
  if (fbo) return ibfs_get_state(h, nullptr, nullptr, nullptr, nullptr, fbo);
  return IBFS_OK;
}

int ibfs_set_solver(ibfs_handle* h, int stationary, double cgtol, int cg_max_iter) {
  if (!h || cg_max_iter < 0 || !(cgtol >= 0.0)) return IBFS_EINVAL;
  h->stationary = stationary ? 1 : 0;
  h->cgtol = cgtol;
  h->cg_max_iter = cg_max_iter;
  return IBFS_OK;
}

int ibfs_move_body(ibfs_handle* h, const double* xb, const double* yb) {
  if (!xb || !yb) return IBFS_EINVAL;
  return guarded(h, [&] {
    if (h->ncases != 1) throw ArgError{IBFS_EINVAL, "moving bodies need a one-case handle"};
    HIPCHK(hipStreamSynchronize(h->stream));
    h->drop_graph();  // the captured step holds the old tables
    h->groups[0]->move_body(0, xb, yb);
  });
}

int ibfs_cg_iterations(const ibfs_handle* h, int nsteps, int* iters) {
  if (!h || !iters || nsteps < 0) return IBFS_EINVAL;
  for (int k = 0; k < nsteps; k++) iters[k] = k < (int)h->cg_iters.size() ? h->cg_iters[k] : 0;
  return IBFS_OK;
}

int ibfs_last_advance_ms(const ibfs_handle* h, double* ms) {
  if (!h || !ms) return IBFS_EINVAL;
  *ms = h->last_ms;
  return IBFS_OK;
}

// ----------------------------------------------------------------- operators
// The operator entry points run on case 0 of an FP64 handle.
#define NEED64()                                                                     \
  Engine<double>* E = h->e64();                                                      \
  if (!E) throw ArgError{IBFS_EINVAL, "operator entry points need an FP64 handle"}; \
  hipStream_t st = h->stream;                                                        \
  const int nw = E->nw, nq = E->nq;                                                  \
  (void)nq;                                                                          \
  (void)st

namespace {
// plain streaming copy of n doubles (the calibration kernel for HBM counters)
__global__ void k_copy(const double* __restrict__ a, double* __restrict__ b, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = a[i];
}
}  // namespace

int ibfs_time_step_kernel(ibfs_handle* h, int nsteps, double* us_mean, int* launches, int* fields) {
  if (!h || nsteps < 1 || !us_mean) return IBFS_EINVAL;
  return guarded(h, [&] {
    Engine<double>* E = h->e64();
    if (!E || h->groups.size() != 1) throw ArgError{IBFS_EINVAL, "the step-kernel probe needs an FP64 one-group handle"};
    if (!h->stationary || !h->all_factors()) throw ArgError{IBFS_ESTATE, "the probe steps need the factor form"};
    typename Engine<double>::Probe pr;
    pr.ev.resize((size_t)nsteps * 4 * E->Lmax + 8);
    for (auto& e : pr.ev) {
      HIPCHK(hipEventCreate(&e.first));
      HIPCHK(hipEventCreate(&e.second));
    }
    struct Cleanup {
      ibfs_handle* h; Engine<double>* E; typename Engine<double>::Probe* pr; bool g;
      ~Cleanup() {
        E->probe = nullptr;
        h->use_graph = g;
        for (auto& e : pr->ev) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
      }
    } cleanup{h, E, &pr, h->use_graph};
    std::vector<double> hist((size_t)nsteps * h->rowlen);
    h->use_graph = false;  // eager launches: events cannot ride in a captured graph
    E->probe = &pr;
    int itime = 1;  // the steps are never the Euler start
    advance_impl(h, &itime, nsteps, nullptr, hist.data());
    HIPCHK(hipStreamSynchronize(h->stream));
    double sum = 0.0;
    for (size_t k = 0; k < pr.next; k++) {
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, pr.ev[k].first, pr.ev[k].second));
      sum += ms;
    }
    *us_mean = pr.next ? 1e3 * sum / (double)pr.next : 0.0;
    if (launches) *launches = (int)pr.next;
    if (fields) *fields = (int)pr.fields;
  });
}

int ibfs_time_copy(ibfs_handle* h, long n, int reps, double* ms_per_launch) {
  if (!h || !ms_per_launch || reps < 1 || n < 1) return IBFS_EINVAL;
  return guarded(h, [&] {
    hipStream_t st = h->stream;
    DevBuf<double> a, b;
    a.alloc(n); b.alloc(n);
    a.zero(st);
    const int blocks = (int)std::min<long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, st, a.p, b.p, n);
    HIPCHK(hipEventRecord(h->ev0, st));
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, st, a.p, b.p, n);
    HIPCHK(hipEventRecord(h->ev1, st));
    HIPCHK(hipStreamSynchronize(st));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    *ms_per_launch = ms / reps;
  });
}

int ibfs_tile_shapes(const ibfs_handle* h, int group, int* ngroups, int* shapes) {
  if (!h || group < 0) return IBFS_EINVAL;
  if (ngroups) *ngroups = (int)h->groups.size();
  if (group >= (int)h->groups.size()) return IBFS_EINVAL;
  if (shapes) h->groups[group]->tile_shapes(shapes);
  return IBFS_OK;
}

int ibfs_level_tile_shapes(const ibfs_handle* h, int group, int* shapes) {
  if (!h || group < 0 || !shapes || group >= (int)h->groups.size()) return IBFS_EINVAL;
  h->groups[group]->level_tile_shapes(shapes);
  return IBFS_OK;
}

extern "C++" {
namespace {
// mean duration (ms) of reps launches of f on the handle's stream
template <class F>
double time_launches(ibfs_handle* h, int reps, F&& f) {
  f();
  HIPCHK(hipEventRecord(h->ev0, h->stream));
  for (int r = 0; r < reps; r++) f();
  HIPCHK(hipEventRecord(h->ev1, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  return ms / reps;
}

// one FFT pass over zero W fields, one per case of group 0 (the group's
// batch and tuned tile shape), in the engine's element type
template <typename T>
double time_fft_pass(ibfs_handle* h, Engine<T>* E, int axis, int reps) {
  const int B = E->B;
  const long nw = E->nw;
  DevBuf<T> a, b;
  a.alloc((size_t)B * nw); b.alloc((size_t)B * nw);
  a.zero(h->stream);
  return time_launches(h, reps, [&] {
    if (axis == 0) E->row_pass(SrcField<T>{a.p, nw}, DstField<T>{b.p, nw}, B);
    else if (axis == 1) E->col_pass(SrcField<T>{a.p, nw}, DstField<T>{b.p, nw}, B);
    else
      E->launch_col2(Col2Op<T, SrcField<T>, MidScale<T>, DstField<T>>{E->geo, SrcField<T>{a.p, nw},
                                                                        MidScale<T>{E->laminv.p}, DstField<T>{b.p, nw},
                                                                        nullptr, 0},
                     E->geo.mm1, 2, B);
  });
}

// DST-I matrix S(a, b) = 2 sin(pi (a+1)(b+1) / N), a, b < N-1, with the
// argument reduced exactly modulo 2N
template <typename T>
std::vector<T> dst_matrix(int N) {
  const int L = N - 1;
  std::vector<T> S((size_t)L * L);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int a = 0; a < L; a++)
    for (int b = 0; b < L; b++) {
      const long r = ((long)(a + 1) * (b + 1)) % (2L * N);
      S[(size_t)a * L + b] = T(2.0L * sinl(pi * (long double)r / (long double)N));
    }
  return S;
}

// a dense DST-I pass (Y = X S_m along i, Y = S_n X along j) of src into dst
template <typename T>
void gemm_pass(hipStream_t st, const Geo& g, int axis, int ks, const T* S, const T* src, T* dst) {
  const int M = g.nm1, N = g.mm1, K = axis == 0 ? g.mm1 : g.nm1;
  const dim3 grid((N + 15) / 16, (M + 15) / 16, 1);
  const T* A = axis == 0 ? src : S;
  const T* B = axis == 0 ? S : src;
  const int lda = axis == 0 ? g.mm1 : g.nm1, ldb = axis == 0 ? g.mm1 : g.mm1;
  switch (ks) {
    case 1: hipLaunchKernelGGL((k_gemm16<T, 1>), grid, dim3(64), 0, st, A, lda, 0L, B, ldb, 0L, dst, g.mm1, 0L, M, N, K); break;
    case 2: hipLaunchKernelGGL((k_gemm16<T, 2>), grid, dim3(128), 0, st, A, lda, 0L, B, ldb, 0L, dst, g.mm1, 0L, M, N, K); break;
    case 4: hipLaunchKernelGGL((k_gemm16<T, 4>), grid, dim3(256), 0, st, A, lda, 0L, B, ldb, 0L, dst, g.mm1, 0L, M, N, K); break;
    default: hipLaunchKernelGGL((k_gemm16<T, 8>), grid, dim3(512), 0, st, A, lda, 0L, B, ldb, 0L, dst, g.mm1, 0L, M, N, K); break;
  }
  HIPCHK(hipGetLastError());
}

template <typename T>
double time_gemm_pass(ibfs_handle* h, const Geo& g, int axis, int ks, int reps) {
  const int L = axis == 0 ? g.mm1 : g.nm1;
  const std::vector<T> hs = dst_matrix<T>(L + 1);
  DevBuf<T> S, a, b;
  S.alloc(hs.size());
  S.upload(hs.data(), hs.size(), h->stream);
  a.alloc(g.nw); b.alloc(g.nw);
  a.zero(h->stream);
  return time_launches(h, reps, [&] { gemm_pass<T>(h->stream, g, axis, ks, S.p, a.p, b.p); });
}
}  // namespace
}  // extern "C++"

int ibfs_time_dst_pass(ibfs_handle* h, int axis, int reps, double* ms_per_launch) {
  if (!h || !ms_per_launch || reps < 1 || axis < 0 || axis > 2) return IBFS_EINVAL;
  return guarded(h, [&] {
    if (auto* E = dynamic_cast<Engine<double>*>(h->groups[0].get())) *ms_per_launch = time_fft_pass(h, E, axis, reps);
    else if (auto* F = dynamic_cast<Engine<float>*>(h->groups[0].get()))
      *ms_per_launch = time_fft_pass(h, F, axis, reps);
    else throw ArgError{IBFS_EINVAL, "no engine"};
  });
}

int ibfs_time_dst_gemm(ibfs_handle* h, int axis, int ks, int reps, double* ms_per_launch) {
  if (!h || !ms_per_launch || reps < 1 || axis < 0 || axis > 1 || (ks != 1 && ks != 2 && ks != 4 && ks != 8))
    return IBFS_EINVAL;
  return guarded(h, [&] {
    if (auto* E = dynamic_cast<Engine<double>*>(h->groups[0].get()))
      *ms_per_launch = time_gemm_pass<double>(h, E->geo, axis, ks, reps);
    else if (auto* F = dynamic_cast<Engine<float>*>(h->groups[0].get()))
      *ms_per_launch = time_gemm_pass<float>(h, F->geo, axis, ks, reps);
    else throw ArgError{IBFS_EINVAL, "no engine"};
  });
}

int ibfs_op_dst_gemm(ibfs_handle* h, int ks, const double* in, double* out) {
  if (ks != 1 && ks != 2 && ks != 4 && ks != 8) return IBFS_EINVAL;
  return guarded(h, [&] {
    NEED64();
    const std::vector<double> hm = dst_matrix<double>(E->geo.m), hn = dst_matrix<double>(E->geo.n);
    DevBuf<double> Sm, Sn, a, b;
    Sm.alloc(hm.size()); Sn.alloc(hn.size());
    Sm.upload(hm.data(), hm.size(), st);
    Sn.upload(hn.data(), hn.size(), st);
    a.alloc(nw); b.alloc(nw);
    a.upload(in, nw, st);
    gemm_pass<double>(st, E->geo, 0, ks, Sm.p, a.p, b.p);
    gemm_pass<double>(st, E->geo, 1, ks, Sn.p, b.p, a.p);
    a.download(out, nw, st);
    HIPCHK(hipStreamSynchronize(st));
  });
}

int ibfs_op_dst(ibfs_handle* h, const double* in, double* out) {
  return guarded(h, [&] {
    NEED64();
    DevBuf<double> a, b;
    a.alloc(nw); b.alloc(nw);
    a.upload(in, nw, st);
    E->row_pass(SrcField<double>{a.p, 0}, DstField<double>{b.p, 0}, 1);
    E->col_pass(SrcField<double>{b.p, 0}, DstField<double>{a.p, 0}, 1);
    a.download(out, nw, st);
    HIPCHK(hipStreamSynchronize(st));
  });
}

static int op_ctci(ibfs_handle* h, const double* in, double* out, bool ainv) {
  return guarded(h, [&] {
    NEED64();
    DevBuf<double> a, b, t1, t2;
    a.alloc(nw); b.alloc(nw); t1.alloc(nw); t2.alloc(nw);
    a.upload(in, nw, st);
    E->ctci_like(SrcField<double>{a.p, 0}, ainv, DstField<double>{b.p, 0}, t1.p, t2.p, 1, nullptr, 0, E->lam1i.p, 0);
    b.download(out, nw, st);
    HIPCHK(hipStreamSynchronize(st));
  });
}
int ibfs_op_ctci(ibfs_handle* h, const double* in, double* out) { return op_ctci(h, in, out, false); }
int ibfs_op_ainv(ibfs_handle* h, const double* in, double* out) { return op_ctci(h, in, out, true); }

int ibfs_op_curl(ibfs_handle* h, const double* x, const double* coarse, double* vel) {
  return guarded(h, [&] {
    NEED64();
    DevBuf<double> a, c, v;
    a.alloc(nw); c.alloc(nw); v.alloc(nq);
    a.upload(x, nw, st);
    if (coarse) c.upload(coarse, nw, st);
    hipLaunchKernelGGL(k_curl<double>, dim3((nq + 255) / 256, 1), dim3(256), 0, st, E->geo, a.p,
                       coarse ? c.p : nullptr, v.p, 0L, 0L, (const int*)nullptr, 0);
    v.download(vel, nq, st);
    HIPCHK(hipStreamSynchronize(st));
  });
}

int ibfs_op_rot(ibfs_handle* h, const double* x, double* out) {
  return guarded(h, [&] {
    NEED64();
    DevBuf<double> a, r;
    a.alloc(nq); r.alloc(nw);
    a.upload(x, nq, st);
    hipLaunchKernelGGL(k_rot<double>, dim3((nw + 255) / 256), dim3(256), 0, st, E->geo, a.p, r.p);
    r.download(out, nw, st);
    HIPCHK(hipStreamSynchronize(st));
  });
}

int ibfs_op_coarsify(ibfs_handle* h, const double* crhs, const double* rhs, double* out) {
  return guarded(h, [&] {
    NEED64();
    DevBuf<double> c, f;
    c.alloc(nw); f.alloc(nw);
    c.upload(crhs, nw, st);
    f.upload(rhs, nw, st);
    const int wcnt = (E->m / 2 - 1) * (E->n / 2 - 1);
    hipLaunchKernelGGL(k_coarsify<double>, dim3((wcnt + 255) / 256, 1), dim3(256), 0, st, E->geo, c.p, f.p, 0L,
                       (const int*)nullptr, 1);
    c.download(out, nw, st);
    HIPCHK(hipStreamSynchronize(st));
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
    NEED64();
    DevBuf<double> a, b;
    a.alloc(nw); b.alloc(E->nbc);
    a.upload(r, nw, st);
    hipLaunchKernelGGL(k_getbc_all, dim3((E->nbc + 255) / 256), dim3(256), 0, st, E->geo, a.p, fac, b.p);
    b.download(rbc, E->nbc, st);
    HIPCHK(hipStreamSynchronize(st));
  });
}

int ibfs_op_rot_nonlinear(ibfs_handle* h, const double* omega, const double* q, const double* q0, const double* bc,
                          double* out) {
  return guarded(h, [&] {
    NEED64();
    DevBuf<double> a, b, c, d, r;
    a.alloc(nw); b.alloc(nq); c.alloc(nq); d.alloc(E->nbc); r.alloc(nw);
    a.upload(omega, nw, st);
    b.upload(q, nq, st);
    c.upload(q0, nq, st);
    d.upload(bc, E->nbc, st);
    hipLaunchKernelGGL(k_rot_nonlinear<double>, dim3((nw + 255) / 256), dim3(256), 0, st, E->geo, a.p, b.p, c.p,
                       d.p, r.p);
    r.download(out, nw, st);
    HIPCHK(hipStreamSynchronize(st));
  });
}

int ibfs_op_vort2flux(ibfs_handle* h, double* omega, double* vel, int nlev) {
  return guarded(h, [&] {
    NEED64();
    const int L = E->cs[0].L;
    if (nlev < 0 || nlev > L) throw ArgError{IBFS_EINVAL, "nlev out of range"};
    const size_t W = (size_t)E->Lmax * nw, Q = (size_t)E->Lmax * nq;
    DevBuf<double> om, s, v, t1, t2;
    om.alloc(W); s.alloc(W); v.alloc(Q); t1.alloc(nw); t2.alloc(nw);
    om.zero(st);
    om.upload(omega, (size_t)L * nw, st);
    E->vort2flux(om.p, s.p, nlev > 0 ? v.p : nullptr, 1, E->Lc.p, t1.p, t2.p);
    om.download(omega, (size_t)L * nw, st);
    if (nlev > 0 && vel) v.download(vel, (size_t)nq * nlev, st);
    HIPCHK(hipStreamSynchronize(st));
  });
}

int ibfs_op_reg(ibfs_handle* h, const double* h0, double* hq) {
  return guarded(h, [&] {
    NEED64();
    const int nf = E->cs[0].nf;
    DevBuf<double> a, b;
    a.alloc(std::max(nf, 1)); b.alloc(nq);
    a.upload(h0, nf, st);
    b.zero(st);
    const int nact = E->hcs[0].nact;
    if (nact)
      hipLaunchKernelGGL(k_reg_active<double>, dim3((nact + 255) / 256), dim3(256), 0, st, E->sm, a.p, b.p);
    b.download(hq, nq, st);
    HIPCHK(hipStreamSynchronize(st));
  });
}

int ibfs_op_regT(ibfs_handle* h, const double* hq, double* h0) {
  return guarded(h, [&] {
    NEED64();
    const int nf = E->cs[0].nf;
    DevBuf<double> a, b;
    a.alloc(nq); b.alloc(std::max(nf, 1));
    a.upload(hq, nq, st);
    if (nf) hipLaunchKernelGGL(k_regT<double>, dim3((nf + 255) / 256), dim3(256), 0, st, E->sm, a.p, b.p);
    b.download(h0, nf, st);
    HIPCHK(hipStreamSynchronize(st));
  });
}

int ibfs_op_a_times(ibfs_handle* h, const double* x, double* out) {
  return guarded(h, [&] {
    NEED64();
    const int nf = E->cs[0].nf;
    const long bsL = E->fbs();
    DevBuf<double> F, RRb, t1, t2, VB, SB, o;
    F.alloc(std::max(nf, 1)); RRb.alloc(nw); t1.alloc(nw); t2.alloc(nw);
    VB.alloc(bsL); SB.alloc(bsL); o.alloc(std::max(nf, 1));
    RRb.zero(st); VB.zero(st);
    F.upload(x, nf, st);
    hipLaunchKernelGGL((k_rotreg<double, double>), dim3((E->hcs[0].nvtx + 255) / 256, 1), dim3(256), 0, st, E->sm,
                       F.p, (long)nf, RRb.p, 0L, 0);
    E->ctci_like(SrcField<double>{RRb.p, 0}, true, DstField<double>{VB.p, 0}, t1.p, t2.p, 1, nullptr, 0, E->lam1i.p,
                 0);
    E->vort2flux(VB.p, SB.p, nullptr, 1, E->Lc.p, t1.p, t2.p);
    hipLaunchKernelGGL(k_regT_curl<double>, dim3((nf + 255) / 256, 1), dim3(256), 0, st, E->geo, E->sm, SB.p,
                       E->Lmax > 1 ? SB.p + nw : nullptr, 0L, (const double*)nullptr, o.p, (long)nf, E->Lc.p, 0,
                       (const double*)nullptr);
    o.download(out, nf, st);
    HIPCHK(hipStreamSynchronize(st));
  });
}

int ibfs_op_cholsl(ibfs_handle* h, const double* b, double* x) {
  return guarded(h, [&] {
    NEED64();
    if (!E->cs[0].have_chol) throw ArgError{IBFS_ESTATE, "no Cholesky factor"};
    const int nf = E->cs[0].nf;
    DevBuf<double> r, y, o;
    r.alloc(std::max(E->nf_tot, 1)); y.alloc(std::max(E->nf_tot, 1)); o.alloc(std::max(E->nf_tot, 1));
    r.upload(b, nf, st);
    hipLaunchKernelGGL(k_trmv_lower, dim3((nf + 3) / 4, 1), dim3(256), 0, st, E->sm, E->Linv.p, r.p, y.p);
    hipLaunchKernelGGL(k_trmv_upper_t, dim3((nf + 3) / 4, 1), dim3(256), 0, st, E->sm, E->Linv.p, y.p, o.p, 0);
    o.download(x, nf, st);
    HIPCHK(hipStreamSynchronize(st));
  });
}

const char* ibfs_last_error(const ibfs_handle* h) { return h ? h->err.c_str() : noalloc_last_error(); }

int ibfs_destroy(ibfs_handle* h) {
  if (!h) return IBFS_EINVAL;
  delete h;
  return IBFS_OK;
}

}  // extern "C"

// ============================================================================
// ibfs_ctci_noalloc: ctci (src/myfft.f90:97-109) on caller-owned device
// buffers, with no allocation at all.  Its intermediates, the pair-major
// eigenvalue table and the twiddles of both axes live in one file-scope
// device array, `scratch`, whose address on each device is looked up once
// (hipGetSymbolAddress) and kept; the array's host-side name is never passed
// to a kernel.  Every call on a device shares that device's array (calls that
// overlap on one device race); every entry a call reads is written by the
// same call first, so nothing is left over from an earlier call.
// ============================================================================
namespace {

constexpr int kNaMaxMN = 1024;                                   // m, n <= 1024
constexpr long kNaMaxW = (long)(kNaMaxMN - 1) * (kNaMaxMN - 1);  // (m-1)(n-1)
constexpr long kNaMaxPM = (long)kNaMaxMN * (kNaMaxMN - 1);        // pm_size <= m (n-1)
constexpr int kNaMaxTw = 2048;                                    // spec twiddles per axis (1365 at 2m = 2048)
// doubles: T1 (W) | T2 (W) | laminv, pair-major | twiddles along i | along j (complex)
constexpr long kNaScratch = 2 * kNaMaxW + kNaMaxPM + 2L * 2 * kNaMaxTw;
constexpr int kNaMaxDev = 64;

__device__ __align__(256) double scratch[kNaScratch];

std::mutex na_mu;
double* na_ptr[kNaMaxDev] = {};       // the array's address on each device
thread_local char na_err[256] = "";   // message of the calling thread's last failure

// the size-specialised FFT lengths; the no-alloc entry takes only these
constexpr int kNaSizes[] = {256, 400, 512, 600, 640, 768, 1024, 1200, 2048};
bool na_spec(int n2) {
  for (int s : kNaSizes)
    if (s == n2) return true;
  return false;
}

// ctci's table laminv = 1/(lam 4mn) (src/myfft.f90:57-59,106) in the
// pair-major layout of the column pair (padding entries 0), and the per-stage
// twiddles exp(-2 pi i k/(ns R)) of the two axes' schedules
__global__ void k_na_tables(Geo g, double norm, double* __restrict__ lpm, spec::Sched sm,
                            cplx<double>* __restrict__ twm, spec::Sched sn, cplx<double>* __restrict__ twn) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long npm = pm_size(g);
  if (t < npm) {
    const long pr = t >> 1, ii = 2 * (pr / g.nm1) + (t & 1), jj = pr % g.nm1;
    double v = 0.0;
    if (ii < g.mm1) {
      const double lam = -2.0 * (cos(kPiRef * (double)(float)(ii + 1) / (double)(float)g.m) +
                                 cos(kPiRef * (double)(float)(jj + 1) / (double)(float)g.n) - 2.0);
      v = 1.0 / lam / norm;
    }
    lpm[t] = v;
    return;
  }
  long u = t - npm;
  for (int ax = 0; ax < 2; ax++) {
    const spec::Sched& sc = ax == 0 ? sm : sn;
    if (u < sc.ntw) {
      int s = 0;
      while (s + 1 < sc.nst && u >= sc.toff[s + 1]) s++;
      const long k = u - sc.toff[s];
      double sn_, cs_;
      sincospi(-2.0 * (double)k / (double)(sc.ns[s] * sc.rad[s]), &sn_, &cs_);
      (ax == 0 ? twm : twn)[u] = cplx<double>{cs_, sn_};
      return;
    }
    u -= sc.ntw;
  }
}

template <int NN>
void na_row(const RowOp<double, SrcField<double>, DstField<double>>& op, const cplx<double>* tw, hipStream_t st) {
  hipLaunchKernelGGL((spec::dst_pass_n<double, NN, 1, 0, RowOp<double, SrcField<double>, DstField<double>>>),
                     dim3((op.g.nm1 + 1) / 2, 1), dim3(NN / 4), 0, st, op, tw);
}
template <int NN>
void na_col2(const Col2Op<double, SrcField<double>, MidScale<double>, DstField<double>>& op, const cplx<double>* tw,
             hipStream_t st) {
  hipLaunchKernelGGL((spec::dst_col2_n<double, NN, 1, Col2Op<double, SrcField<double>, MidScale<double>,
                                                            DstField<double>>>),
                     dim3((op.g.mm1 + 1) / 2, 1), dim3(NN / 4), 0, st, op, tw);
}
#define IBFS_NA_DISPATCH(FN, N2, ...) \
  switch (N2) {                       \
    case 256: FN<256>(__VA_ARGS__); break;   \
    case 400: FN<400>(__VA_ARGS__); break;   \
    case 512: FN<512>(__VA_ARGS__); break;   \
    case 600: FN<600>(__VA_ARGS__); break;   \
    case 640: FN<640>(__VA_ARGS__); break;   \
    case 768: FN<768>(__VA_ARGS__); break;   \
    case 1024: FN<1024>(__VA_ARGS__); break; \
    case 1200: FN<1200>(__VA_ARGS__); break; \
    default: FN<2048>(__VA_ARGS__); break;   \
  }

int na_fail(int code, const char* msg) {
  std::snprintf(na_err, sizeof(na_err), "%s", msg);
  return code;
}

}  // namespace

const char* noalloc_last_error() { return na_err[0] ? na_err : "null handle"; }

extern "C" {

int ibfs_ctci_noalloc(int m, int n, const double* omega, double* s, void* stream) {
  if (m < 4 || n < 4 || m > kNaMaxMN || n > kNaMaxMN || !na_spec(2 * m) || !na_spec(2 * n))
    return na_fail(IBFS_EINVAL, "ibfs_ctci_noalloc: m and n must be one of 128, 200, 256, 300, 320, 384, 512, "
                                "600, 1024 (the size-specialised transforms; scratch holds up to 1024 x 1024)");
  if (!omega || !s) return na_fail(IBFS_EINVAL, "ibfs_ctci_noalloc: null field");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNaMaxDev)
    return na_fail(IBFS_EHIP, "ibfs_ctci_noalloc: no current device");
  double* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(na_mu);
    if (!na_ptr[dev]) {  // once per device, with that device current
      void* p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p) return na_fail(IBFS_EHIP, hipGetErrorString(e));
      na_ptr[dev] = static_cast<double*>(p);
    }
    base = na_ptr[dev];
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  const Geo g{m, n, 1, m - 1, n - 1, (m - 1) * (n - 1), (m + 1) * n + (n + 1) * m, 2 * (m + 1) + 2 * (n + 1),
              0, n + 1, 2 * (n + 1), 2 * (n + 1) + m + 1};
  const long W = g.nw, npm = pm_size(g);
  double* t1 = base;
  double* t2 = base + W;
  double* lpm = base + 2 * W;
  cplx<double>* twm = reinterpret_cast<cplx<double>*>(base + 2 * kNaMaxW + kNaMaxPM);
  cplx<double>* twn = twm + kNaMaxTw;
  const spec::Sched sm = spec::make_sched(2 * m), sn = spec::make_sched(2 * n);
  const double norm = 4.0 * (double)(float)((long)m * (long)n);  // src/myfft.f90:42
  const long ntab = npm + sm.ntw + sn.ntw;
  hipLaunchKernelGGL(k_na_tables, dim3((unsigned)((ntab + 255) / 256)), dim3(256), 0, st, g, norm, lpm, sm, twm, sn,
                     twn);
  // S . diag(laminv) . S as a row pass, the fused column pair and a row pass
  const RowOp<double, SrcField<double>, DstField<double>> r1{g, SrcField<double>{omega, 0}, DstField<double>{t1, 0},
                                                             nullptr, 0};
  const Col2Op<double, SrcField<double>, MidScale<double>, DstField<double>> c2{
      g, SrcField<double>{t1, 0}, MidScale<double>{lpm}, DstField<double>{t2, 0}, nullptr, 0};
  const RowOp<double, SrcField<double>, DstField<double>> r2{g, SrcField<double>{t2, 0}, DstField<double>{s, 0},
                                                             nullptr, 0};
  IBFS_NA_DISPATCH(na_row, 2 * m, r1, twm, st)
  IBFS_NA_DISPATCH(na_col2, 2 * n, c2, twn, st)
  IBFS_NA_DISPATCH(na_row, 2 * m, r2, twm, st)
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return na_fail(IBFS_EHIP, hipGetErrorString(e));
  return IBFS_OK;
}

}  // extern "C"
