// ibfs_engine_t.hpp -- Engine<T>: one group of equal-size cases on the device.
#pragma once
#include "dst_wave.hpp"
#include "ibfs_engine.hpp"

namespace ibfs {

// Per-stage twiddles of a size-specialised pass: stage s, butterfly index
// k < ns holds exp(-2 pi i k / (ns R)) (the kernels form the other powers).
template <typename T>
bool make_spec_twiddles(int n2, DevBuf<cplx<T>>& tw, hipStream_t st) {
  static const int kSizes[] = {256, 400, 512, 600, 640, 768, 1024, 1200, 2048};
  if (std::find(std::begin(kSizes), std::end(kSizes), n2) == std::end(kSizes)) return false;
  const spec::Sched sc = spec::make_sched(n2);
  std::vector<cplx<T>> t(std::max(sc.ntw, 1));
  for (int s = 0; s < sc.nst; s++) {
    const int R = sc.rad[s], ns = sc.ns[s];
    for (int k = 0; k < ns; k++) {
      long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)(ns * R);
      t[sc.toff[s] + k] = {(T)cosl(a), (T)sinl(a)};
    }
  }
  tw.alloc(t.size());
  tw.upload(t.data(), t.size(), st);
  return true;
}

template <typename T>
void make_plan(int n2, FftPlan<T>& plan, DevBuf<cplx<T>>& tw, hipStream_t s) {
  std::vector<int> rad;
  if (!factor_235(n2, rad) || n2 > kEmax)
    throw ArgError{1, "grid size not supported: 2*m and 2*n must factor into 2, 3, 5 and be <= " +
                          std::to_string(kEmax)};
  plan.n2 = n2;
  plan.nfft = kEmax / n2;
  plan.nst = (int)rad.size();
  for (int i = 0; i < kMaxStages; i++) plan.rad[i] = i < plan.nst ? rad[i] : 1;
  std::vector<cplx<T>> t(n2);
  for (int k = 0; k < n2; k++) {
    long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n2;
    t[k] = {(T)cosl(a), (T)sinl(a)};
  }
  tw.alloc(n2);
  tw.upload(t.data(), n2, s);
  plan.tw = tw.p;
}

// Plans and twiddles of both axes in element type V.
template <typename V>
struct FftSet {
  FftPlan<V> pm{}, pn{};
  DevBuf<cplx<V>> tw_m, tw_n, stw_m, stw_n;
  bool spec_m = false, spec_n = false;
  void init(int m, int n, hipStream_t st) {
    make_plan<V>(2 * m, pm, tw_m, st);
    make_plan<V>(2 * n, pn, tw_n, st);
    spec_m = make_spec_twiddles<V>(2 * m, stw_m, st);
    spec_n = make_spec_twiddles<V>(2 * n, stw_n, st);
  }
};

// element type of a pass operator (what its load returns)
template <class Op>
using op_value_t = std::decay_t<decltype(std::declval<const Op&>().load(0, 0, 0, 1, 0))>;

template <typename T>
struct Engine : GroupBase {
  hipStream_t st = nullptr;
  Geo geo{};
  T norm = T(1);
  FftSet<T> fs;
  // An FP32 engine solves the step-2 Poisson problems (the streamfunction
  // whose curl is the right-hand side of the force system) in FP64: with
  // FP32 transforms there, their rounding, of order eps*|psi|, reaches the
  // slip at the body, and the force that removes it carries per-step noise
  // of ~1e-4 of the force scale (mean forces are unaffected).  Step 4 solves
  // every level in FP64 too: the stored fluxes are differences of nearby
  // streamfunction values, so an FP32 streamfunction would give them a
  // relative error of about eps*|psi|/|dpsi| ~ eps*m/pi (1.6e-5 measured on
  // level 2 at 512^2), against FP32 rounding of the flux itself in FP64.
  static constexpr bool kSolve64 = sizeof(T) == 4;
  FftSet<double> fs64;
  DevBuf<double> laminv64, S64, U1, U2;
  // step 2 of the FP64 streamfunctions needs the levels below the finest
  // only on their rings (get_bc of the next finer level, the force system's
  // get_bc(s_2)): there the last row pass gives way to the ring lines of the
  // column pair's output (ctci64 with ring; r2V: their ring vectors, r2Sin:
  // sin(pi t / m)).  Step 4 computes every level whole.  IBFS_STEP2_RING=0
  // keeps the row pass.
  bool ring2 = true;
  DevBuf<double> r2V, r2Sin, r2Y;  // r2Y: the FP64 engine's column pairs, two fields per case

  DevBuf<int> Lc;
  const int* gate = nullptr;  // Lc.p, or nullptr when every case has Lmax levels
  DevBuf<T> con1, con2, vfac;
  DevBuf<T> laminv, lam1i;
  DevBuf<T> omega, rhs_old, q, q0, S, T1, T2, T3, RR, lastbc;
  // q0 of every case in the freestream form (src/variables.f90:59-63: one
  // value on the u faces of a level, 0 on its v faces)?  Then the step-1
  // loader takes the per-(case, level) value from q0c instead of loading q0.
  DevBuf<T> q0c;
  std::vector<T> hq0c;
  std::vector<char> q0u;
  bool q0_uniform() const {
    for (char u : q0u)
      if (!u) return false;
    return true;
  }

  DevBuf<int> face_id, face_ptr, face_pt, frc_ptr, frc_face, frc_ia, frc_ib, vtx_id, vtx_face, bdyno;
  DevBuf<double> face_w, frc_w, wsum;
  DevBuf<int> f9_face, f9_ia, f9_ib;  // the force gather padded to kFrcK entries per force
  DevBuf<double> f9_w;
  DevBuf<CaseSm> csm;
  std::vector<CaseSm> hcs;
  Smear sm{};

  DevBuf<double> fb, motion_in, motion, rq0, rhsf, ytmp, rtmp, Linv, Asym;
  // actuator slots (src/grid.f90:579-718): the points of every case's
  // actuators (act_ptr: the per-case ranges), and each case's step counter,
  // which the step's k_actuation advances
  std::vector<std::vector<ActPt>> hact;
  DevBuf<ActPt> act_pts;
  DevBuf<int> act_ptr, dit;
  bool has_act = false;
  DevBuf<double> partial, diag;
  DevBuf<double> Ainv, Gm;  // A^-1 (Nf x Nf per case) and rot(reg(A^-1)) (nvtx x Nf per case)
  DevBuf<int> counter;
  DevBuf<int> dpend;        // per case: a diagnostics row is pending (deferred to the next step's first pass)
  bool side_diag = false;   // defer the rows (set by the advance loop, which flushes the last one)
  int diag_cap = 0, nchunk = 64;
  int nf_tot = 0, nf_max = 0, nvtx_max = 0;
  long mat_tot = 0, gmat_tot = 0;
  bool refine = false;
  Engine* base = nullptr;  // set on a tangent-linear engine: the state it is linearised about
  // Step 1 with the levels batched (enqueue_step1_levels): every level's
  // bulk solve in one set of launches over the flattened (case, level)
  // items (geoL: one level per item), then every level's correction for the
  // new coarser edge values in another.  T1L, T2L: one field per item.
  bool lev_batch = false;
  // the freestream step-1 source of the level-batched form (and of the
  // level-by-level FP64 form) runs as its own elementwise kernel
  // (k_src_eval) before a plain row pass, instead of as the loader of that
  // pass; IBFS_STEP1_SPLIT=0 keeps it in the loader
  bool src_split = true;
  // the two-level coarsify launches in LDS tiles; IBFS_COARSIFY_TILES=0
  // keeps one place per thread (the same values)
  bool coarsify_tiles = true;
  // the ring vectors' bottom and top lines and the edge spectra by FFT
  // (k_ring_fromY_fft, k_ring_spectra_fft); IBFS_RING_FFT=0 keeps the
  // direct sums
  bool ring_fft_lines = true;
  // step 4's batched vort2flux keeps the left/right terms apart from S (in
  // ringS) for the tail to add; IBFS_SPLIT_S=0 adds them into S
  bool s_split = true;
  Geo geoL{};
  DevBuf<T> T1L, T2L;
  // vort2flux with the levels batched (see v2f_can_batch): ringN ring
  // entries per level, ringT the powers T^1..T^(Lmax-2) of the ring map,
  // ringG / ringGp the closed-form tables along i / j, ringR the ring
  // vectors of every (case, level), ringS a streamfunction-shaped field that
  // holds the level-1 ring values of step 2 (its get_bc for the force
  // right-hand side)
  bool v2f_batch = false;
  // the rows swin[0..1] and columns swin[2..3] of level 0 that hold every
  // vertex the force stencils read (upload_smear); swin_ok when there are
  // such vertices and no stencil face lies on the domain edge.
  // IBFS_S_WINDOW=0: step 2 transforms the whole of level 0.
  bool swin_ok = false;
  int swin[4] = {0, 0, 0, 0};
  // step 3 in real space (k_ainv_local), per case: taken (ainv_loc), the
  // target box, the kernel K(a, b) for 0 <= a, b <= ainv_R; set in
  // upload_smear.  IBFS_STEP3_LOCAL=0 keeps the spectral form everywhere.
  DevBuf<int> ainv_loc, ainv_box;
  DevBuf<double> ainv_K;
  int ainv_nloc = 0, ainv_R = 0, ainv_tiles = 0;
  // step 4's coarsify after step 3 in real space: between step 2's coarsify
  // and step 4's only level 1's box around the body changed, so each
  // two-level tile launch runs only the tiles whose places have a stencil
  // reaching a changed place (co_tiles: per tile launch, per case {tx0, ty0,
  // ntx, nty}; the whole grid for a case on the spectral step 3; co_grid:
  // each launch's largest tile count).  The other places would be
  // recomputed from unchanged inputs: the result is bit-identical.
  // IBFS_COARSIFY_INC=0 runs the whole chain.
  bool coarsify_inc = true;
  DevBuf<int> co_tiles;
  std::vector<int> co_grid;
  int ringN = 0;
  DevBuf<double> ringT, ringG, ringGp, ringR, ringS, ringV, ringX, ringSin;
  // the ring chain on the representatives' rows of T's powers (k_ring_chain
  // with sym; set in build_ring when T commutes with the ring's symmetries);
  // IBFS_RING_SYM=0 reads every row
  DevBuf<int> ringSym;
  bool ring_sym_on = false;
  // the band correction (BandOp): beta per item, the band width per level
  DevBuf<double> band_beta, band_rho_m, band_rho_n;
  DevBuf<int> band_w;
  std::vector<int> hband_w;
  int band_wmax = 0;
  // the real-space form of the band correction (k_band_direct): H_d(x) per
  // item, axis and depth, |x| <= band_R; IBFS_BAND_FFT=1 keeps BandOp
  bool band_direct = false;
  int band_R = 0;
  DevBuf<double> band_h;

  long fbs() const { return (long)Lmax * nw; }
  long tbs() const { return (long)Lmax * nwp; }  // case stride of lam1i (pair-major tables)
  long qbs() const { return (long)Lmax * nq; }
  bool is_double() const override { return sizeof(T) == 8; }

  ~Engine() override { cg_release(); }
  Engine(std::vector<CaseHost> cases, hipStream_t stream) {
    st = stream;
    cs = std::move(cases);
    B = (int)cs.size();
    m = cs[0].d.m;
    n = cs[0].d.n;
    Lmax = 0;
    for (auto& c : cs) Lmax = std::max(Lmax, c.L);
    nw = (m - 1) * (n - 1);
    nq = (m + 1) * n + (n + 1) * m;
    nbc = 2 * (m + 1) + 2 * (n + 1);
    geo = Geo{m, n, Lmax, m - 1, n - 1, nw, nq, nbc, 0, n + 1, 2 * (n + 1), 2 * (n + 1) + m + 1};
    nwp = pm_size(geo);
    norm = T(cs[0].normalize);
    refine = sizeof(T) == 4;
    fs.init(m, n, st);
    if constexpr (kSolve64) fs64.init(m, n, st);

    // per-case scalars and eigenvalue tables (src/myfft.f90:42-69)
    std::vector<int> hL(B);
    std::vector<T> hc1((size_t)B * Lmax, T(0)), hc2(hc1), hvf(hc1);
    // the column pair's tables are pair-major (pm_ix), zero in the padding
    std::vector<T> hlaminv(nwp, T(0)), hl1i((size_t)B * Lmax * nwp, T(0));
    std::vector<double> hlaminv64(nwp, 0.0);
    std::vector<double> lam(nw);
    for (int j = 1; j <= n - 1; j++)
      for (int i = 1; i <= m - 1; i++) {
        const size_t e = (size_t)(i - 1) + (size_t)(j - 1) * (m - 1);
        const long ep = pm_ix(geo, i - 1, j - 1);
        lam[e] = -2.0 * (std::cos(kPiRef * (double)(float)i / (double)(float)m) +
                         std::cos(kPiRef * (double)(float)j / (double)(float)n) - 2.0);
        hlaminv[ep] = T(1.0 / lam[e] / cs[0].normalize);
        hlaminv64[ep] = 1.0 / lam[e] / cs[0].normalize;
      }
    for (int b = 0; b < B; b++) {
      const CaseHost& c = cs[b];
      hL[b] = c.L;
      const double del2 = c.delta * c.delta;
      for (int k = 0; k < c.L; k++) {
        hc1[b * Lmax + k] = T(c.con1[k]);
        hc2[b * Lmax + k] = T(c.con2[k]);
        hvf[b * Lmax + k] = T(c.vfac[k]);
        const double del22 = del2 * std::pow(4.0, k);
        const size_t base = ((size_t)b * Lmax + k) * nwp;
        for (int jj = 0; jj < n - 1; jj++)
          for (int ii = 0; ii < m - 1; ii++) {
            const size_t e = (size_t)ii + (size_t)jj * (m - 1);
            hl1i[base + pm_ix(geo, ii, jj)] = T(1.0 / (1.0 + 0.5 * c.d.dt * lam[e] / del22 / c.d.re));
          }
      }
    }
    upvec(Lc, hL, st);
    // the step's kernels take no level gate when every case has all levels
    gate = Lc.p;
    if (std::all_of(hL.begin(), hL.end(), [&](int l) { return l == Lmax; })) gate = nullptr;
    upvec(con1, hc1, st);
    upvec(con2, hc2, st);
    upvec(vfac, hvf, st);
    upvec(laminv, hlaminv, st);
    if constexpr (kSolve64) upvec(laminv64, hlaminv64, st);
    upvec(lam1i, hl1i, st);

    upload_smear();

    // state and initial condition (src/variables.f90:54-67)
    const size_t W = (size_t)B * Lmax * nw, Q = (size_t)B * Lmax * nq;
    omega.alloc(W); omega.zero(st);
    rhs_old.alloc(W); rhs_old.zero(st);
    S.alloc(W); S.zero(st);
    q.alloc(Q); q.zero(st);
    q0.alloc(Q);
    std::vector<T> hq0(Q, T(0));
    for (int b = 0; b < B; b++) {
      const CaseHost& c = cs[b];
      if (!c.d.unif_flow) continue;
      for (int k = 0; k < c.L; k++) {
        const T v = T(c.delta * std::pow(2.0, k));
        const size_t base = ((size_t)b * Lmax + k) * nq;
        for (size_t e = 0; e < (size_t)(m + 1) * n; e++) hq0[base + e] = v;
      }
    }
    q0.upload(hq0.data(), Q, st);
    // the freestream form of q0, per case and level (u faces, v faces 0)
    hq0c.assign((size_t)B * Lmax, T(0));
    q0u.assign(B, 1);
    for (int b = 0; b < B; b++)
      for (int k = 0; k < Lmax; k++) hq0c[(size_t)b * Lmax + k] = hq0[((size_t)b * Lmax + k) * nq];
    q0c.alloc(hq0c.size());
    q0c.upload(hq0c.data(), hq0c.size(), st);
    T1.alloc((size_t)B * nw);
    if constexpr (kSolve64) {
      S64.alloc(W);
      S64.zero(st);
      U1.alloc((size_t)B * nw);
      U2.alloc((size_t)2 * B * nw);  // two fields per case for ctci64's ring form
    }
    if (const char* r2 = std::getenv("IBFS_STEP2_RING")) ring2 = r2[0] != '0';
    ring2 = ring2 && m % 4 == 0 && n % 4 == 0 && Lmax >= 2;
    if (ring2) {
      r2V.alloc((size_t)2 * B * ring_count(m, n));
      r2Sin.alloc((size_t)2 * m);
      hipLaunchKernelGGL(k_sinpi_table, dim3((2 * m + 255) / 256), dim3(256), 0, st, m, r2Sin.p);
      if constexpr (!kSolve64) r2Y.alloc((size_t)2 * B * nw);
    }
    T2.alloc((size_t)B * nw);
    T3.alloc((size_t)B * nw);
    RR.alloc((size_t)B * nw); RR.zero(st);
    lastbc.alloc((size_t)B * Lmax * nbc);
    lastbc.zero(st);  // the levels without a coarser level keep zero edge values
    const size_t nff = std::max(nf_tot, 1);
    fb.alloc(nff); fb.zero(st);
    motion_in.alloc(nff); motion_in.zero(st);
    motion.alloc(nff); motion.zero(st);
    rq0.alloc(nff); rq0.zero(st);
    hact.assign(B, {});
    rhsf.alloc(nff);
    ytmp.alloc(nff);
    rtmp.alloc(nff);
    Linv.alloc(std::max<long>(mat_tot, 1));
    Ainv.alloc(std::max<long>(mat_tot, 1));
    Gm.alloc(std::max<long>(gmat_tot, 1));
    if (refine) Asym.alloc(std::max<long>(mat_tot, 1));
    partial.alloc((size_t)B * Lmax * nchunk);
    counter.alloc(B);
    counter.zero(st);
    dpend.alloc(B);
    dpend.zero(st);
    ensure_diag(1);
    HIPCHK(hipStreamSynchronize(st));
    tune();
    lev_batch = levels_can_batch();
    if (const char* sp = std::getenv("IBFS_STEP1_SPLIT")) src_split = sp[0] != '0';
    if (const char* ct = std::getenv("IBFS_COARSIFY_TILES")) coarsify_tiles = ct[0] != '0';
    if (const char* ci = std::getenv("IBFS_COARSIFY_INC")) coarsify_inc = ci[0] != '0';
    if (const char* rf = std::getenv("IBFS_RING_FFT")) ring_fft_lines = rf[0] != '0';
    if (const char* sp = std::getenv("IBFS_SPLIT_S")) s_split = sp[0] != '0';
    if (lev_batch) {
      geoL = geo;
      geoL.Lmax = 1;
      T1L.alloc((size_t)B * Lmax * nw);
      T2L.alloc((size_t)B * Lmax * nw);
      std::vector<double> hb((size_t)B * Lmax, 1.0);
      for (int b = 0; b < B; b++)
        for (int k = 0; k < Lmax; k++) {
          const double dk = cs[b].delta * std::pow(2.0, k);
          hb[(size_t)b * Lmax + k] = 0.5 * cs[b].d.dt / (cs[b].d.re * dk * dk);
        }
      upvec(band_beta, hb, st);
      upvec(band_w, hband_w, st);
      const int BL = B * Lmax;
      band_rho_m.alloc((size_t)BL * (m - 1));
      band_rho_n.alloc((size_t)BL * (n - 1));
      hipLaunchKernelGGL(k_band_rho, dim3((unsigned)(((long)BL * (m - 1) + 255) / 256)), dim3(256), 0, st,
                         band_beta.p, m, BL, band_rho_m.p);
      hipLaunchKernelGGL(k_band_rho, dim3((unsigned)(((long)BL * (n - 1) + 255) / 256)), dim3(256), 0, st,
                         band_beta.p, n, BL, band_rho_n.p);
      HIPCHK(hipGetLastError());
      band_wmax = *std::max_element(hband_w.begin(), hband_w.end());
      const char* bf = std::getenv("IBFS_BAND_FFT");
      band_direct = !(bf && bf[0] == '1') && build_band_kernels(hb);
      HIPCHK(hipStreamSynchronize(st));
    }
    v2f_batch = v2f_can_batch();
    if (v2f_batch) build_ring();
    if (lev_batch || v2f_batch) tune_levels();
  }

  // Whether vort2flux runs with the levels batched (the kernels after
  // k_ring_gtab): the decomposition s_k = ctci(omega_k) + H(sbc_k) is exact
  // for any size, so this asks only that the FP64 engine solve one or two
  // cases with all levels (batches are throughput-bound), at most
  // kRingMaxLev levels and at most 1 GiB of powers of T;
  // IBFS_V2F_LEVELS=0 switches it off.
  bool v2f_can_batch() const {
    const char* env = std::getenv("IBFS_V2F_LEVELS");
    if (env && env[0] == '0') return false;
    if constexpr (!std::is_same_v<T, double>) return false;
    if (gate != nullptr || Lmax < 2 || B > 2 || Lmax > kRingMaxLev) return false;
    const double nr = ring_count(m, n);
    return (double)std::max(Lmax - 2, 1) * nr * nr * 8.0 <= double(1L << 30);
  }
  // The orbits of the ring places under ring_sym (8 symmetries on square
  // grids, 4 otherwise): each place's representative (the smallest place of
  // its orbit) and a symmetry that maps the representative to it, and the
  // places ordered orbit by orbit.  Taken only when the built T commutes
  // with every symmetry to 1e-12 of its largest entry (it does up to the
  // rounding of its GEMMs; the chain then uses the representative's row,
  // the same matrix to rounding).
  void ring_sym_tables() {
    ring_sym_on = false;
    const char* env = std::getenv("IBFS_RING_SYM");
    if ((env && env[0] == '0') || Lmax < 3) return;
    const int NR = ringN, ng = m == n ? 8 : 4;
    std::vector<int> perm((size_t)ng * NR);
    for (int gi = 0; gi < ng; gi++)
      for (int p = 0; p < NR; p++) perm[(size_t)gi * NR + p] = ring_sym(m, n, gi, p);
    std::vector<double> Th((size_t)NR * NR);
    HIPCHK(hipMemcpyAsync(Th.data(), ringT.p, sizeof(double) * Th.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    double mx = 0.0, worst = 0.0;
    for (double v : Th) mx = std::max(mx, std::fabs(v));
    for (int gi = 1; gi < ng; gi++) {
      const int* pg = perm.data() + (size_t)gi * NR;
      for (int r = 0; r < NR; r++) {
        const double* tr = Th.data() + (size_t)r * NR;
        const double* tg = Th.data() + (size_t)pg[r] * NR;
        for (int c = 0; c < NR; c++) worst = std::max(worst, std::fabs(tg[pg[c]] - tr[c]));
      }
    }
    if (!(worst <= 1e-12 * mx)) return;
    std::vector<int> tab((size_t)3 * NR), ord(NR);
    for (int p = 0; p < NR; p++) {
      int r = p;
      for (int gi = 1; gi < ng; gi++) r = std::min(r, perm[(size_t)gi * NR + p]);
      int gp = -1;
      for (int gi = 0; gi < ng && gp < 0; gi++)
        if (perm[(size_t)gi * NR + r] == p) gp = gi;
      if (gp < 0) return;  // not a group action: keep every row
      tab[NR + p] = r;
      tab[2 * NR + p] = gp;
      ord[p] = p;
    }
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return tab[NR + a] < tab[NR + b]; });
    std::copy(ord.begin(), ord.end(), tab.begin());
    upvec(ringSym, tab, st);
    ring_sym_on = true;
  }
  // the tables of the batched vort2flux: g / (2N) along each axis, T from its
  // four edge blocks (T[:, r0_e + c] = B_e A_e^T), and T's powers
  void build_ring() {
    ringN = ring_count(m, n);
    const int NR = ringN;
    if (!lev_batch) {
      T1L.alloc((size_t)B * Lmax * nw);
      T2L.alloc((size_t)B * Lmax * nw);
      geoL = geo;
      geoL.Lmax = 1;
    }
    ringG.alloc((size_t)(m - 1) * (n - 1));
    hipLaunchKernelGGL(k_ring_gtab, dim3((unsigned)(((long)(m - 1) * (n - 1) + 255) / 256)), dim3(256), 0, st, m,
                       n, ringG.p);
    if (m != n) {
      ringGp.alloc((size_t)(m - 1) * (n - 1));
      hipLaunchKernelGGL(k_ring_gtab, dim3((unsigned)(((long)(m - 1) * (n - 1) + 255) / 256)), dim3(256), 0, st, n,
                         m, ringGp.p);
    }
    ringT.alloc((size_t)std::max(Lmax - 2, 1) * NR * NR);
    const int Kmax = std::max(m, n) - 1;
    DevBuf<double> Bm, Am;
    Bm.alloc((size_t)NR * Kmax);
    Am.alloc((size_t)(Kmax / 2 + 2) * Kmax);
    for (int e = 0; e < 4; e++) {
      const int Nal = e < 2 ? m : n, K = Nal - 1, cnt = Nal / 2 + 1;
      const double* G = (e < 2 || m == n) ? ringG.p : ringGp.p;
      hipLaunchKernelGGL(k_ring_factors, dim3((unsigned)(((long)NR * K + 255) / 256)), dim3(256), 0, st, m, n, e, G,
                         Bm.p, Am.p);
      hipLaunchKernelGGL(k_gemm64, dim3((cnt + 15) / 16, (NR + 15) / 16), dim3(256), 0, st, NR, cnt, K, Bm.p, (long)K,
                         1L, Am.p, 1L, (long)K, ringT.p + ring_edge0(m, n, e), (long)NR, 0);
    }
    for (int j = 2; j <= Lmax - 2; j++)  // T^j = T^(j-1) T
      hipLaunchKernelGGL(k_gemm64, dim3((NR + 15) / 16, (NR + 15) / 16), dim3(256), 0, st, NR, NR, NR,
                         ringT.p + (size_t)(j - 2) * NR * NR, (long)NR, 1L, ringT.p, (long)NR, 1L,
                         ringT.p + (size_t)(j - 1) * NR * NR, (long)NR, 0);
    ring_sym_tables();
    ringR.alloc((size_t)B * Lmax * NR);
    ringR.zero(st);
    ringV.alloc((size_t)B * Lmax * NR);
    ringV.zero(st);
    ringX.alloc((size_t)B * Lmax * 4 * (std::max(m, n) - 1));
    ringX.zero(st);
    ringS.alloc((size_t)B * fbs());
    ringS.zero(st);
    ringSin.alloc((size_t)2 * (m + n));  // sin(pi t / m), t < 2m, then sin(pi t / n), t < 2n
    hipLaunchKernelGGL(k_sinpi_table, dim3((2 * m + 255) / 256), dim3(256), 0, st, m, ringSin.p);
    hipLaunchKernelGGL(k_sinpi_table, dim3((2 * n + 255) / 256), dim3(256), 0, st, n, ringSin.p + 2 * m);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));  // Bm, Am are freed on return
  }

  // Whether step 1 may run with the levels batched.  The reference solves
  // the levels coarse to fine because level k's edge values come from the
  // new vorticity of level k+1 (src/operators.f90:103-107).  Those edge
  // values enter linearly, so level k's new vorticity is its bulk solve
  // (old edge values, nonlinear term) plus the response
  // (I + beta_k C^T C)^-1 applied to the edge source, and level k samples
  // level k+1 (get_bc) only on the lines i = m/4, 3m/4, j = n/4, 3n/4 of
  // level k+1, at least min(m,n)/4 cells from its edges.  Level k+1's own
  // edge response decays from its edges by at most r = exp(-acosh(1 +
  // 1/(2 beta))) per cell, beta = dt/(2 Re Delta_{k+1}^2) (the slowest mode
  // of the 1-D recurrence normal to an edge), so it is below r^(min(m,n)/4)
  // of the edge source there: the batched form takes the bulk solve of level
  // k+1 for level k's edge values when that factor is below 1e-17 on every
  // level with a finer level, for every case.  It is used for one or two
  // cases with all levels (batches are throughput-bound; one level's extra
  // transform pair would cost more than the launches it saves) and can be
  // switched off with IBFS_STEP1_LEVELS=0.
  // The correction itself is written on the w_k rows and columns next to
  // each edge of level k, w_k = ceil(ln(1e17) / kappa_k) for the level's own
  // beta (the dropped terms are below 1e-17 of the edge source), which must
  // stay under a quarter of the grid (hband_w; filled here).
  // The kernels of k_band_direct, in long double from the same rho_k
  // formula as k_band_rho: H_d(x) = (1/2N) sum_{k=0}^{2N-1} rho_k^(d+1)
  // cos(pi k x / N) for x <= X = min(N, 64), and the reach R = the largest
  // over items, axes and depths d < w of the smallest R_d with
  // 2 sum_{R_d < x <= X} |H_d(x)| < 1e-17 (H_d(X) itself below 1e-20, so
  // that the rest of the period adds nothing).  Declines (the spectral
  // BandOp form runs) when R would pass kBandRmax or a depth's table
  // would not fit the kernel's LDS.
  bool build_band_kernels(const std::vector<double>& beta) {
    const int BL = B * Lmax;
    if (band_wmax > kBandWmax || band_wmax * (kBandRmax + 1) > kBandHmax) return false;
    const long double pi = 3.141592653589793238462643383279502884L;
    std::vector<long double> tab((size_t)BL * 2 * band_wmax * 65, 0.0L);
    int R = 0;
    for (int axis = 0; axis < 2; axis++) {
      const int N = axis == 0 ? m : n, P = 2 * N, X = std::min(N, 64);
      std::vector<long double> cs2(P), rho(P);
      for (int j = 0; j < P; j++) cs2[j] = cosl(pi * j / N);
      for (int b = 0; b < BL; b++) {
        const int lev = b % Lmax;
        if (lev == Lmax - 1) continue;
        const long double bt = beta[b];
        for (int k = 0; k < P; k++) {
          const long double lx = 2.0L - 2.0L * cs2[k];
          const long double th = (1.0L + bt * (lx + 2.0L)) / (2.0L * bt);
          rho[k] = 1.0L / (th + sqrtl(th * th - 1.0L));
        }
        const int w = hband_w[lev];
        std::vector<long double> pw(rho);
        for (int d = 0; d < w; d++) {
          long double* h = &tab[(((size_t)b * 2 + axis) * band_wmax + d) * 65];
          for (int xx = 0; xx <= X; xx++) {
            long double acc = 0.0L;
            for (int k = 0; k < P; k++) acc += pw[k] * cs2[(size_t)((long)k * xx % P)];
            h[xx] = acc / P;
          }
          if (fabsl(h[X]) > 1e-20L) return false;
          long double tail = 0.0L;
          int rd = X;
          while (rd > 0 && tail + 2.0L * fabsl(h[rd]) < 1e-17L) tail += 2.0L * fabsl(h[rd--]);
          R = std::max(R, rd);
          for (int k = 0; k < P; k++) pw[k] *= rho[k];
        }
      }
    }
    if (R > kBandRmax) return false;
    band_R = R;
    std::vector<double> hh((size_t)BL * 2 * band_wmax * (R + 1), 0.0);
    for (size_t r = 0; r < (size_t)BL * 2 * band_wmax; r++)
      for (int xx = 0; xx <= R; xx++) hh[r * (R + 1) + xx] = (double)tab[r * 65 + xx];
    upvec(band_h, hh, st);
    return true;
  }
  bool levels_can_batch() {
    const char* env = std::getenv("IBFS_STEP1_LEVELS");
    if (env && env[0] == '0') return false;
    if (gate != nullptr || Lmax < 2 || B > 2) return false;
    const double need = std::log(1e17);
    hband_w.assign(Lmax, 1);
    for (const CaseHost& c : cs)
      for (int k = 0; k < c.L; k++) {
        const double dk = c.delta * std::pow(2.0, k);
        const double beta = 0.5 * c.d.dt / (c.d.re * dk * dk);
        const double kappa = std::acosh(1.0 + 1.0 / (2.0 * beta));
        if (k >= 1 && kappa * (std::min(m, n) / 4) < need) return false;
        if (k + 1 < c.L) hband_w[k] = std::max(hband_w[k], (int)std::ceil(need / kappa));
      }
    for (int w : hband_w)
      if (4 * w >= std::min(m, n)) return false;
    return true;
  }
  void tile_shapes(int* out) const override {
    for (int k = 0; k < kKinds; k++) {  // 0: the kind runs the generic pass
      const bool rowtype = k == kRow || k == kStep1Row;
      out[k] = (rowtype ? fs.spec_m : fs.spec_n) ? ptile[k] : 0;
    }
  }
  void level_tile_shapes(int* out) const override {
    for (int k = 0; k < kKinds; k++) {  // 0: no level-batched launch of this kind, or the generic pass
      const bool rowtype = k == kRow || k == kStep1Row;
      const bool used = lev_tiles && k != kStep1Row;
      out[k] = used && (rowtype ? fs.spec_m : fs.spec_n) ? ptileL[k] : 0;
    }
  }

  // ------------------------------------------------------------------ passes
  // Tile shapes: FFTs per workgroup (P) of the size-specialised kernels, per
  // kind of pass, chosen at construction by timing every candidate on this
  // group's batch size (tune()).  Results do not depend on P.
  enum Kind { kRow = 0, kCol = 1, kCol2 = 2, kStep1Row = 3, kStep1Col2 = 4, kKinds = 5 };
  int ptile[kKinds] = {1, 1, 1, 1, 1};
  // tile shapes of the level-batched launches (batch B Lmax), tuned on that
  // batch when the step batches its levels (lev_tiles)
  int ptileL[kKinds] = {1, 1, 1, 1, 1};
  bool lev_tiles = false;
  // tile code of the register-blocked kernels (dst_wave.hpp): one FFT per
  // 64 / TPF lanes, one wave per workgroup; bit-identical to P = 1
  static constexpr int kWaveTile = 8;
  // tile code of the column passes with staged loads and stores
  // (dst_col_ws, dst_col2_ws): four waves, four FFTs at 2m = 1024
  static constexpr int kStagedTile = 16;
  static constexpr bool wave_size(int n2) { return n2 == 256 || n2 == 512 || n2 == 1024; }
  template <typename V = T>
  static constexpr bool p_fits(int n2, int P) {
    return P * n2 / 4 <= 1024 && (long)P * n2 * (long)sizeof(cplx<V>) <= 65536;
  }
  // the transforms of element type V (the FP64 set of an FP32 engine serves
  // its FP64 step-2 solves); passes of the other type use P = 1
  template <typename V>
  const FftSet<V>& fset() const {
    if constexpr (std::is_same_v<V, T>) return fs;
    else return fs64;
  }
  template <typename V>
  int tile(int kind, int batch) const {
    if (!std::is_same_v<V, T>) return 1;
    return (lev_tiles && batch == B * Lmax && batch != B) ? ptileL[kind] : ptile[kind];
  }

  template <int NN, int P, int AXIS, class Op>
  void launch_spec(const Op& op, const cplx<op_value_t<Op>>* stw, int nseq, int mult, int batch) {
    using V = op_value_t<Op>;
    if constexpr (p_fits<V>(NN, P)) {
      int groups = (nseq + mult * P - 1) / (mult * P);
      if constexpr (op_side<Op>::value)
        groups += op.side_groups_for(P * NN / 4);
      hipLaunchKernelGGL((spec::dst_pass_n<V, NN, P, AXIS, Op>), dim3(groups, batch), dim3(P * NN / 4), 0, st, op,
                         stw);
    }
  }
  template <int NN, int P, class Op>
  void launch_spec_col2(const Op& op, int nseq, int mult, int batch, int ax = 1) {
    using V = op_value_t<Op>;
    if constexpr (p_fits<V>(NN, P)) {
      const int groups = (nseq + mult * P - 1) / (mult * P);
      if constexpr (std::is_same_v<Op, Col2Op<T, SrcField<T>, MidScale<T>, DstField<T>>>) {
        if (probe && probe->next < probe->ev.size()) {  // measurement: events on the dispatch itself
          auto& e = probe->ev[probe->next++];
          probe->fields += batch;
          hipExtLaunchKernelGGL((spec::dst_col2_n<V, NN, P, Op>), dim3(groups, batch), dim3(P * NN / 4), 0, st,
                                e.first, e.second, 0, op, fset<V>().stw_n.p);
          return;
        }
      }
      hipLaunchKernelGGL((spec::dst_col2_n<V, NN, P, Op>), dim3(groups, batch), dim3(P * NN / 4), 0, st, op,
                         ax ? fset<V>().stw_n.p : fset<V>().stw_m.p);
    }
  }
  // Measurement only: with probe set (eager steps, never a captured graph),
  // every launch of the ctci column pair (the step's dominant kernel) carries
  // a start and a stop event on its own dispatch, so their difference is the
  // kernel's duration inside the step; fields counts the W fields those
  // launches transform (their grid.y: one per level when the levels are
  // batched).
  struct Probe {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    size_t next = 0;
    long fields = 0;
  };
  Probe* probe = nullptr;

  // the register-blocked passes: one wave per workgroup, 64 / (N2 / 16)
  // FFTs per wave
  template <int NN, int AXIS, class Op>
  void launch_wave(const Op& op, const cplx<op_value_t<Op>>* stw, int nseq, int mult, int batch) {
    using V = op_value_t<Op>;
    constexpr int P = 64 / (NN / 16);
    const int groups = (nseq + mult * P - 1) / (mult * P);
    hipLaunchKernelGGL((wave::dst_pass_w<V, NN, 1, AXIS, Op>), dim3(groups, batch), dim3(64), 0, st, op, stw);
  }
  template <int NN, class Op>
  void launch_wave_col2(const Op& op, const cplx<op_value_t<Op>>* stw, int nseq, int mult, int batch) {
    using V = op_value_t<Op>;
    constexpr int P = 64 / (NN / 16);
    const int groups = (nseq + mult * P - 1) / (mult * P);
    if constexpr (std::is_same_v<Op, Col2Op<T, SrcField<T>, MidScale<T>, DstField<T>>>) {
      if (probe && probe->next < probe->ev.size()) {  // measurement: events on the dispatch itself
        auto& e = probe->ev[probe->next++];
        probe->fields += batch;
        hipExtLaunchKernelGGL((wave::dst_col2_w<V, NN, 1, Op>), dim3(groups, batch), dim3(64), 0, st, e.first,
                              e.second, 0, op, stw);
        return;
      }
    }
    hipLaunchKernelGGL((wave::dst_col2_w<V, NN, 1, Op>), dim3(groups, batch), dim3(64), 0, st, op, stw);
  }

  // the staged column passes: four waves per workgroup, 4 64 / (N2 / 16)
  // FFTs per workgroup
  template <int NN, class Op>
  void launch_staged(const Op& op, const cplx<op_value_t<Op>>* stw, int nseq, int mult, int batch) {
    using V = op_value_t<Op>;
    constexpr int P = wave::kStagedWaves * 64 / (NN / 16);
    const int groups = (nseq + mult * P - 1) / (mult * P);
    hipLaunchKernelGGL((wave::dst_col_ws<V, NN, Op>), dim3(groups, batch), dim3(wave::kStagedWaves * 64), 0, st, op,
                       stw);
  }
  template <int NN, class Op>
  void launch_staged_col2(const Op& op, const cplx<op_value_t<Op>>* stw, int nseq, int mult, int batch) {
    using V = op_value_t<Op>;
    constexpr int P = wave::kStagedWaves * 64 / (NN / 16);
    const int groups = (nseq + mult * P - 1) / (mult * P);
    if constexpr (std::is_same_v<Op, Col2Op<T, SrcField<T>, MidScale<T>, DstField<T>>>) {
      if (probe && probe->next < probe->ev.size()) {  // measurement: events on the dispatch itself
        auto& e = probe->ev[probe->next++];
        probe->fields += batch;
        hipExtLaunchKernelGGL((wave::dst_col2_ws<V, NN, Op>), dim3(groups, batch), dim3(wave::kStagedWaves * 64), 0,
                              st, e.first, e.second, 0, op, stw);
        return;
      }
    }
    hipLaunchKernelGGL((wave::dst_col2_ws<V, NN, Op>), dim3(groups, batch), dim3(wave::kStagedWaves * 64), 0, st,
                       op, stw);
  }

  // Launch one pass over nseq sequences along the other axis; mult = 2 when
  // a workgroup's P FFTs take 2P sequences of one field, 1 when they pair
  // two fields (X + iY).  kind selects the tile shape.
  // A row pass (batch nya) and a column pass (batch nyb) of the same FFT
  // length in one launch (dst_pass_dual), two sequences per FFT, when the
  // row pass runs a size-specialised LDS pass with P = 1, 2 or 4 FFTs per
  // workgroup at its batch (the column pass takes the same P: the results do
  // not depend on it); false (nothing launched) otherwise.
  template <class OpA, class OpB>
  bool launch_dual(const OpA& a, int nseqa, const OpB& bq, int nseqb, int nya, int nyb) {
    using V = op_value_t<OpA>;
    static_assert(std::is_same_v<V, op_value_t<OpB>>, "launch_dual: one value type");
    const FftSet<V>& F = fset<V>();
    if (!F.spec_m || !F.spec_n || F.pm.n2 != F.pn.n2) return false;
    int P = tile<V>(0, nya);
    if (P == kWaveTile || P == kStagedTile) {
      if (wave::kOpOk<OpA> && wave_size(F.pm.n2)) return false;  // the row pass is register-blocked
      P = 1;
    }
    if (P != 1 && P != 2 && P != 4) return false;
    const int ga = (nseqa + 2 * P - 1) / (2 * P), gb = (nseqb + 2 * P - 1) / (2 * P);
    const int batch = std::max(nya, nyb);
    switch (F.pm.n2 * 8 + P) {
#define IBFS_DUAL_CASE(NN, PP)                                                                                 \
  case NN * 8 + PP:                                                                                            \
    if constexpr (p_fits<V>(NN, PP)) {                                                                         \
      hipLaunchKernelGGL((spec::dst_pass_dual<V, NN, PP, OpA, OpB>), dim3(ga + gb, batch), dim3(PP * NN / 4), 0, \
                         st, a, bq, F.stw_m.p, F.stw_n.p, ga, nya, nyb);                                       \
      return true;                                                                                             \
    }                                                                                                          \
    return false;
#define IBFS_DUAL_SIZE(NN) IBFS_DUAL_CASE(NN, 1) IBFS_DUAL_CASE(NN, 2) IBFS_DUAL_CASE(NN, 4)
      IBFS_DUAL_SIZE(256)
      IBFS_DUAL_SIZE(400)
      IBFS_DUAL_SIZE(512)
      IBFS_DUAL_SIZE(600)
      IBFS_DUAL_SIZE(640)
      IBFS_DUAL_SIZE(768)
      IBFS_DUAL_SIZE(1024)
      IBFS_DUAL_SIZE(1200)
      IBFS_DUAL_SIZE(2048)
#undef IBFS_DUAL_SIZE
#undef IBFS_DUAL_CASE
      default: return false;
    }
  }
  // step 2 on the force-stencil rectangle (see vort2flux_levels)
  bool step2_window() const { return swin_ok && !refine; }
  template <int AXIS, class Op>
  void launch(const Op& op, int nseq, int mult, int batch, int kind = AXIS) {
    using V = op_value_t<Op>;
    const FftSet<V>& F = fset<V>();
    const int n2 = AXIS == 0 ? F.pm.n2 : F.pn.n2;
    const bool sp = AXIS == 0 ? F.spec_m : F.spec_n;
    const cplx<V>* stw = AXIS == 0 ? F.stw_m.p : F.stw_n.p;
    int P = tile<V>(kind, batch);
    if (P == kStagedTile) {
      if constexpr (AXIS == 1 && wave::kStagedOk<Op>) {
        switch (n2) {
          case 256: launch_staged<256>(op, stw, nseq, mult, batch); return;
          case 512: launch_staged<512>(op, stw, nseq, mult, batch); return;
          case 1024: launch_staged<1024>(op, stw, nseq, mult, batch); return;
          default: break;
        }
      }
      P = kWaveTile;
    }
    if (P == kWaveTile) {
      if constexpr (wave::kOpOk<Op>) {
        switch (n2) {
          case 256: launch_wave<256, AXIS>(op, stw, nseq, mult, batch); return;
          case 512: launch_wave<512, AXIS>(op, stw, nseq, mult, batch); return;
          case 1024: launch_wave<1024, AXIS>(op, stw, nseq, mult, batch); return;
          default: break;
        }
      }
      P = 1;  // the same results on the LDS pass
    }
    if (sp) {
      switch (n2 * 8 + P) {
#define IBFS_SPEC_CASE(NN, PP)                                 \
  case NN * 8 + PP:                                            \
    launch_spec<NN, PP, AXIS, Op>(op, stw, nseq, mult, batch); \
    return;
#define IBFS_SPEC_SIZE(NN) IBFS_SPEC_CASE(NN, 1) IBFS_SPEC_CASE(NN, 2) IBFS_SPEC_CASE(NN, 4)
        IBFS_SPEC_SIZE(256)
        IBFS_SPEC_SIZE(400)
        IBFS_SPEC_SIZE(512)
        IBFS_SPEC_SIZE(600)
        IBFS_SPEC_SIZE(640)
        IBFS_SPEC_SIZE(768)
        IBFS_SPEC_SIZE(1024)
        IBFS_SPEC_SIZE(1200)
        IBFS_SPEC_SIZE(2048)
#undef IBFS_SPEC_SIZE
#undef IBFS_SPEC_CASE
        default: break;
      }
    }
    FftPlan<V> plan = AXIS == 0 ? F.pm : F.pn;
    plan.nfft = generic_fanout(n2, nseq, mult, batch);
    int groups = (nseq + mult * plan.nfft - 1) / (mult * plan.nfft);
    if constexpr (op_side<Op>::value)
      groups += op.side_groups_for(kThreads);
    hipLaunchKernelGGL((dst_pass<V, kEmax, AXIS, Op>), dim3(groups, batch), dim3(kThreads),
                       plan.nfft * n2 * sizeof(cplx<V>), st, op, plan);
  }

  // FFTs per workgroup of a generic pass over nseq sequences: as many as
  // fit in kEmax values while the launch still has a workgroup for every
  // CU, else one.  (Filling kEmax regardless left a 127-row field to four
  // workgroups: 16 us per pass instead of 6.)
  static int generic_fanout(int n2, int nseq, int mult, int batch) {
    const long per = (long)nseq * batch / (mult * 256L);
    return (int)std::max(1L, std::min((long)(kEmax / n2), per));
  }

  // fused column pair (dst_col2) over nseq columns
  template <class Op>
  void launch_col2(const Op& op, int nseq, int mult, int batch, int kind = kCol2, int ax = 1) {
    using V = op_value_t<Op>;
    const FftSet<V>& F = fset<V>();
    const int n2 = ax ? F.pn.n2 : F.pm.n2;
    int P = tile<V>(kind, batch);
    const cplx<V>* stw = ax ? F.stw_n.p : F.stw_m.p;
    if (P == kStagedTile) {
      if constexpr (wave::kStagedOk<Op>) {
        if (ax == 1) {
          switch (n2) {
            case 256: launch_staged_col2<256>(op, stw, nseq, mult, batch); return;
            case 512: launch_staged_col2<512>(op, stw, nseq, mult, batch); return;
            case 1024: launch_staged_col2<1024>(op, stw, nseq, mult, batch); return;
            default: break;
          }
        }
      }
      P = kWaveTile;
    }
    if (P == kWaveTile) {
      if constexpr (wave::kOpOk<Op>) {
        switch (n2) {
          case 256: launch_wave_col2<256>(op, stw, nseq, mult, batch); return;
          case 512: launch_wave_col2<512>(op, stw, nseq, mult, batch); return;
          case 1024: launch_wave_col2<1024>(op, stw, nseq, mult, batch); return;
          default: break;
        }
      }
      P = 1;
    }
    if (ax ? F.spec_n : F.spec_m) {
      switch (n2 * 8 + P) {
#define IBFS_SPEC_CASE(NN, PP)                               \
  case NN * 8 + PP:                                          \
    launch_spec_col2<NN, PP, Op>(op, nseq, mult, batch, ax); \
    return;
#define IBFS_SPEC_SIZE(NN) IBFS_SPEC_CASE(NN, 1) IBFS_SPEC_CASE(NN, 2) IBFS_SPEC_CASE(NN, 4)
        IBFS_SPEC_SIZE(256)
        IBFS_SPEC_SIZE(400)
        IBFS_SPEC_SIZE(512)
        IBFS_SPEC_SIZE(600)
        IBFS_SPEC_SIZE(640)
        IBFS_SPEC_SIZE(768)
        IBFS_SPEC_SIZE(1024)
        IBFS_SPEC_SIZE(1200)
        IBFS_SPEC_SIZE(2048)
#undef IBFS_SPEC_SIZE
#undef IBFS_SPEC_CASE
        default: break;
      }
    }
    FftPlan<V> plan = ax ? F.pn : F.pm;
    plan.nfft = generic_fanout(n2, nseq, mult, batch);
    const int groups = (nseq + mult * plan.nfft - 1) / (mult * plan.nfft);
    hipLaunchKernelGGL((dst_col2<V, kEmax, Op>), dim3(groups, batch), dim3(kThreads),
                       plan.nfft * n2 * sizeof(cplx<V>), st, op, plan);
  }

  // ------------------------------------------------------------ autotuner
  // For every kind of pass, launches a representative pass over this
  // group's batch (zero inputs in scratch buffers, the state untouched) with
  // each tile shape that fits (P = 1, 2, 4 FFTs per workgroup), times 20
  // launches with HIP events in three rounds that interleave the candidates
  // (so clock ramp-up and noise hit all of them), and keeps the fastest.  The choice is cached
  // per (kind, m, n, element size, batch).  IBFS_TILE=P forces the largest
  // fitting value <= P instead.
  // IBFS_TILE: one forced shape for every kind, or five comma-separated
  // ones (row, column, column pair, step-1 row, step-1 column pair; 0 = tune);
  // IBFS_TILE_LEVELS likewise for the level-batched launches (default: as
  // IBFS_TILE)
  static std::array<int, kKinds> forced_tiles(const char* var = "IBFS_TILE") {
    std::array<int, kKinds> f{};
    const char* env = std::getenv(var);
    if (!env && std::strcmp(var, "IBFS_TILE") != 0) env = std::getenv("IBFS_TILE");
    if (!env) return f;
    const char* c = env;
    for (int k = 0; k < kKinds; k++) {
      f[k] = std::atoi(c);
      const char* nx = std::strchr(c, ',');
      if (!nx) {
        if (k == 0)
          for (int j = 1; j < kKinds; j++) f[j] = f[0];
        break;
      }
      c = nx + 1;
    }
    return f;
  }
  void tune() {
    static std::mutex mu;
    static std::map<std::array<int, 5>, int> cache;
    const std::array<int, kKinds> ftile = forced_tiles();
    DevBuf<T> qz, bz;
    bool scratch = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    for (int k = 0; k < kKinds; k++) {
      const bool rowtype = k == kRow || k == kStep1Row;
      const int n2 = rowtype ? fs.pm.n2 : fs.pn.n2;
      if (!(rowtype ? fs.spec_m : fs.spec_n)) continue;  // generic pass: P chosen per launch (generic_fanout)
      std::vector<int> cand;
      for (int P : {1, 2, 4})
        if (p_fits<T>(n2, P)) cand.push_back(P);
      if (wave_size(n2) && k != kStep1Row) cand.push_back(kWaveTile);  // its loader is not register-blocked
      if (wave_size(n2) && !rowtype) cand.push_back(kStagedTile);
      const int fp = ftile[k];
      if (fp > 0) {
        int best = cand[0];
        for (int P : cand)
          if (P <= fp) best = P;
        ptile[k] = best;
        continue;
      }
      const std::array<int, 5> key{k, m, n, (int)sizeof(T), B};
      {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) {
          ptile[k] = it->second;
          continue;
        }
      }
      if (!scratch) {
        qz.alloc((size_t)B * nq);
        qz.zero(st);
        bz.alloc((size_t)B * nbc);
        bz.zero(st);
        T2.zero(st);
        T3.zero(st);
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        scratch = true;
      }
      std::vector<float> tmin(cand.size(), 1e30f);
      for (int round = 0; round < 3; round++)  // rounds interleave the candidates
        for (size_t ci = 0; ci < cand.size(); ci++) {
        ptile[k] = cand[ci];
        auto run = [&] {
          switch (k) {
            case kRow: row_pass(SrcField<T>{T2.p, nw}, DstField<T>{T3.p, nw}, B); break;
            case kCol: col_pass(SrcField<T>{T2.p, nw}, DstField<T>{T3.p, nw}, B); break;
            case kCol2:
              launch_col2(Col2Op<T, SrcField<T>, MidScale<T>, DstField<T>>{geo, SrcField<T>{T2.p, nw},
                                                                            MidScale<T>{laminv.p},
                                                                            DstField<T>{T3.p, nw}, nullptr, 0},
                          geo.mm1, 2, B);
              break;
            case kStep1Row:
              {  // the freestream form, as the step launches it
                SrcStep1<T, true> s1f{T2.p, nullptr, qz.p, nullptr, bz.p, T3.p, con1.p, con2.p, vfac.p, gate,
                                      (long)nw, (long)nq, (long)nbc, 0, 0};
                s1f.q0c = q0c.p;
                launch<0>(RowOp<T, SrcStep1<T, true>, DstField<T>>{geo, s1f, DstField<T>{T1.p, nw}, gate, 0},
                          geo.nm1, 2, B, kStep1Row);
              }
              break;
            default:
              launch_col2(Col2Op<T, SrcField<T>, MidTab<T>, DstField<T>>{geo, SrcField<T>{T1.p, nw},
                                                                         MidTab<T>{lam1i.p, tbs()},
                                                                         DstField<T>{T2.p, nw}, gate, 0},
                          geo.mm1, 2, B, kStep1Col2);
              break;
          }
        };
        run();
        run();
        HIPCHK(hipEventRecord(e0, st));
        constexpr int reps = 20;
        for (int r = 0; r < reps; r++) run();
        HIPCHK(hipEventRecord(e1, st));
        HIPCHK(hipStreamSynchronize(st));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        tmin[ci] = std::min(tmin[ci], ms);
      }
      int best = cand[0];
      float best_ms = tmin[0];
      for (size_t ci = 1; ci < cand.size(); ci++)
        if (tmin[ci] < best_ms) {
          best_ms = tmin[ci];
          best = cand[ci];
        }
      ptile[k] = best;
      std::lock_guard<std::mutex> lk(mu);
      cache[key] = best;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    HIPCHK(hipStreamSynchronize(st));
  }

  // Tile shapes of the launches that transform every level of the cases at
  // once (batch B Lmax, one field per item): the kinds those launches use,
  // timed as tune() does on that batch (T1L in, T2L out).
  void tune_levels() {
    static std::mutex mu;
    static std::map<std::array<int, 5>, int> cache;
    const int BL = B * Lmax;
    lev_tiles = true;
    for (int k = 0; k < kKinds; k++) ptileL[k] = ptile[k];
    const std::array<int, kKinds> ftile = forced_tiles("IBFS_TILE_LEVELS");
    T1L.zero(st);
    T2L.zero(st);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    for (int k : {(int)kRow, (int)kCol, (int)kCol2, (int)kStep1Col2}) {
      const bool rowtype = k == kRow;
      const int n2 = rowtype ? fs.pm.n2 : fs.pn.n2;
      if (!(rowtype ? fs.spec_m : fs.spec_n)) continue;
      std::vector<int> cand;
      for (int P : {1, 2, 4})
        if (p_fits<T>(n2, P)) cand.push_back(P);
      if (wave_size(n2)) cand.push_back(kWaveTile);
      if (wave_size(n2) && !rowtype) cand.push_back(kStagedTile);
      const int fp = ftile[k];
      if (fp > 0) {
        int best = cand[0];
        for (int P : cand)
          if (P <= fp) best = P;
        ptileL[k] = best;
        continue;
      }
      const std::array<int, 5> key{k, m, n, (int)sizeof(T), BL};
      {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) {
          ptileL[k] = it->second;
          continue;
        }
      }
      if (!e0) {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
      }
      auto run = [&] {
        switch (k) {
          case kRow:
            launch<0>(RowOp<T, SrcField<T>, DstField<T>>{geoL, SrcField<T>{T1L.p, nw}, DstField<T>{T2L.p, nw},
                                                         nullptr, 0},
                      geo.nm1, 2, BL, kRow);
            break;
          case kCol:
            launch<1>(ColOp<T, SrcField<T>, DstField<T>>{geoL, SrcField<T>{T1L.p, nw}, DstField<T>{T2L.p, nw},
                                                         nullptr, 0},
                      geo.mm1, 2, BL, kCol);
            break;
          case kCol2:
            launch_col2(Col2Op<T, SrcField<T>, MidScale<T>, DstField<T>>{geoL, SrcField<T>{T1L.p, nw},
                                                                          MidScale<T>{laminv.p},
                                                                          DstField<T>{T2L.p, nw}, nullptr, 0},
                        geo.mm1, 2, BL, kCol2);
            break;
          default:
            launch_col2(Col2Op<T, SrcField<T>, MidTab<T>, DstField<T>>{geoL, SrcField<T>{T1L.p, nw},
                                                                       MidTab<T>{lam1i.p, nwp},
                                                                       DstField<T>{T2L.p, nw}, nullptr, 0},
                        geo.mm1, 2, BL, kStep1Col2);
            break;
        }
      };
      std::vector<float> tmin(cand.size(), 1e30f);
      for (int round = 0; round < 3; round++)
        for (size_t ci = 0; ci < cand.size(); ci++) {
          ptileL[k] = cand[ci];
          run();
          run();
          HIPCHK(hipEventRecord(e0, st));
          constexpr int reps = 20;
          for (int r = 0; r < reps; r++) run();
          HIPCHK(hipEventRecord(e1, st));
          HIPCHK(hipStreamSynchronize(st));
          float ms = 0.f;
          HIPCHK(hipEventElapsedTime(&ms, e0, e1));
          tmin[ci] = std::min(tmin[ci], ms);
        }
      size_t bi = 0;
      for (size_t ci = 1; ci < cand.size(); ci++)
        if (tmin[ci] < tmin[bi]) bi = ci;
      ptileL[k] = cand[bi];
      std::lock_guard<std::mutex> lk(mu);
      cache[key] = cand[bi];
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    HIPCHK(hipStreamSynchronize(st));
  }

  template <class Src, class Dst>
  void row_pass(const Src& src, const Dst& dst, int batch, const int* gate = nullptr, int lev = 0) {
    launch<0>(RowOp<T, Src, Dst>{geo, src, dst, gate, lev}, geo.nm1, 2, batch);
  }
  template <class Src, class Dst>
  void col_pass(const Src& src, const Dst& dst, int batch, const int* gate = nullptr, int lev = 0) {
    launch<1>(ColOp<T, Src, Dst>{geo, src, dst, gate, lev}, geo.mm1, 2, batch);
  }

  // out = S * diag(tab) * S (src): ctci (tab = laminv) or ainv (tab = the
  // level-1 lam1i of each case, divided by 4mn), as Sx . (Sy tab Sy) . Sx:
  // a row pass, the fused column pair and a row pass.  t1/t2: scratch,
  // stride nw; dst may alias either.
  template <class Src, class Dst>
  void ctci_like(const Src& src, bool ainv_form, const Dst& dst, T* t1, T* t2, int batch, const int* gate, int lev,
                 const T* lam1i_l1 = nullptr, long tbs = 0) {
    row_pass(src, DstField<T>{t1, nw}, batch, gate, lev);
    if (ainv_form)
      launch_col2(Col2Op<T, SrcField<T>, MidScaleDiv<T>, DstField<T>>{geo, SrcField<T>{t1, nw},
                                                                       MidScaleDiv<T>{lam1i_l1, norm, tbs},
                                                                       DstField<T>{t2, nw}, gate, lev},
                  geo.mm1, 2, batch);
    else
      launch_col2(Col2Op<T, SrcField<T>, MidScale<T>, DstField<T>>{geo, SrcField<T>{t1, nw}, MidScale<T>{laminv.p},
                                                                    DstField<T>{t2, nw}, gate, lev},
                  geo.mm1, 2, batch);
    row_pass(SrcField<T>{t2, nw}, dst, batch, gate, lev);
  }

  // the in-place coarsify chain of vort2flux (src/operators.f90:224-226):
  // two levels per launch, in LDS tiles (k_coarsify2t) or one place per
  // thread (k_coarsify2, also the single last level of an even level count).
  // inc: step 4 after step 3 in real space (the tiles of co_tiles only)
  void coarsify(T* om, int batch, const int* gate, bool inc = false) {
    const int wcnt = (m / 2 - 1) * (n / 2 - 1), nblk = (wcnt + 255) / 256;
    inc = inc && coarsify_inc && coarsify_tiles && batch == B && !co_grid.empty();
    for (int lev = 1, t = 0; lev < Lmax; lev += 2) {
      if (lev + 1 < Lmax && coarsify_tiles) {
        const int ntx = (m / 2 - 1 + kCoTX - 1) / kCoTX, nty = (n / 2 - 1 + kCoTY - 1) / kCoTY;
        if (!inc)
          hipLaunchKernelGGL(k_coarsify2t<T>, dim3(ntx * nty, batch), dim3(256), 0, st, geo, om, fbs(), gate, lev,
                             ntx, (const int*)nullptr);
        else if (co_grid[t] > 0)
          hipLaunchKernelGGL(k_coarsify2t<T>, dim3(co_grid[t], batch), dim3(256), 0, st, geo, om, fbs(), gate, lev,
                             ntx, (const int*)co_tiles.p + 4L * B * t);
        t++;
        continue;
      }
      const int nlev = lev + 1 < Lmax ? 2 : 1;
      hipLaunchKernelGGL(k_coarsify2<T>, dim3(nblk * nlev, batch), dim3(256), 0, st, geo, om, fbs(), gate, lev,
                         nblk);
    }
  }

  // vort2flux (src/operators.f90:191-240) on a batch of multi-level fields om
  // (stride Lmax*nw); streamfunctions go to s (same layout); when vel is set,
  // curl(s) is written on every level (stride Lmax*nq).
  // ring (FP64 engines): the levels below the finest only on their rings,
  // as ctci64's ring form (step 2)
  void vort2flux(T* om, T* s, T* vel, int batch, const int* gate, T* t1, T* t2, bool inc = false,
                 bool ring = false) {
    const long bs = fbs();
    coarsify(om, batch, gate, inc);
    for (int lev = Lmax - 1; lev >= 0; lev--) {
      T* sc = (lev + 1 < Lmax) ? s + (long)(lev + 1) * nw : nullptr;
      if constexpr (!kSolve64) {
        if (ring && (lev > 0 || swin_ok)) {
          row_pass(SrcVortBC<T>{om + (long)lev * nw, sc, bs, gate, lev}, DstField<T>{t1, nw}, batch, gate, lev);
          launch_col2(Col2Op<T, SrcField<T>, MidScale<T>, DstField<T>>{geo, SrcField<T>{t1, nw}, MidScale<T>{laminv.p},
                                                                        DstField<T>{r2Y.p + nw, 2L * nw}, gate,
                                                                        lev},
                      geo.mm1, 2, batch);
          if (lev > 0) ring_lines(r2Y.p, 2, batch, r2V.p, r2Sin.p, s + (long)lev * nw, bs);
          else last_row_window(r2Y.p + nw, 2L * nw, s, bs, batch, gate);
          continue;
        }
      }
      ctci_like(SrcVortBC<T>{om + (long)lev * nw, sc, bs, gate, lev}, false, DstField<T>{s + (long)lev * nw, bs}, t1,
                t2, batch, gate, lev);
    }
    if (vel)
      hipLaunchKernelGGL(k_curl_levels<T>, dim3((nq + 255) / 256, Lmax, batch), dim3(256), 0, st, geo, s, vel, bs,
                         qbs(), gate);
  }

  // ctci of level lev in FP64 from the FP32 vorticity om and the coarser
  // level's streamfunction sc (element type TC), into S64 level lev: the
  // three passes of ctci_like in FP64 with the FP64 laminv.
  // the last row pass of level 0 in step 2 on the force stencils' rows
  // swin[0]..swin[1] only (the force right-hand side reads s_1 nowhere
  // else; the window starts on an even row and ends on an odd one, so it
  // pairs rows into FFTs as the whole pass does: the same values)
  // (Y: the column pair's output, ys apart per case)
  void last_row_window(const double* Y, long ys, double* s0, long bs, int batch, const int* gate) {
    Geo gr = geo;
    gr.nm1 = swin[1] - swin[0] + 1;
    launch<0>(RowOp<double, SrcRowWin<SrcField<double>>, DstRowWin<DstField<double>>>{
                  gr, {SrcField<double>{Y, ys}, geo, swin[0]}, {DstField<double>{s0, bs}, geo, swin[0]}, gate, 0},
              gr.nm1, 2, batch);
  }
  // ring: only the ring places of S64 level lev, from the column pair's
  // output by the ring lines (which read it as level 1 of two-level items:
  // the pair writes at stride 2 nw); on level 0 the last row pass on the
  // force stencils' rows
  template <typename TC>
  void ctci64(const T* om, const TC* sc, int batch, const int* gate, int lev, bool ring = false) {
    const long bs = fbs();
    launch<0>(RowOp<double, SrcVortBC<T, double, TC>, DstField<double>>{
                  geo, SrcVortBC<T, double, TC>{om + (long)lev * nw, sc, bs, gate, lev}, DstField<double>{U1.p, nw},
                  gate, lev},
              geo.nm1, 2, batch);
    const bool two = ring && (lev > 0 || swin_ok);  // the pair at stride 2 nw
    launch_col2(Col2Op<double, SrcField<double>, MidScale<double>, DstField<double>>{
                    geo, SrcField<double>{U1.p, nw}, MidScale<double>{laminv64.p},
                    DstField<double>{two ? U2.p + nw : U2.p, two ? 2L * nw : (long)nw}, gate, lev},
                geo.mm1, 2, batch);
    if (two) {
      if (lev > 0) ring_lines(U2.p, 2, batch, r2V.p, r2Sin.p, S64.p + (long)lev * nw, bs);
      else last_row_window(U2.p + nw, 2L * nw, S64.p, bs, batch, gate);
      return;
    }
    launch<0>(RowOp<double, SrcField<double>, DstField<double>>{geo, SrcField<double>{U2.p, nw},
                                                                DstField<double>{S64.p + (long)lev * nw, bs}, gate,
                                                                lev},
              geo.nm1, 2, batch);
  }
  // vort2flux with the levels batched (see k_ring_gtab): the coarsified
  // vorticity's row transform and column pair on every (case, level) item
  // at once, the ring vectors of s0 on every level from the column pair's
  // output, the chain R_k = sum_j T^j ring(s0_{k+j}), the edge spectra of
  // Interp(R_{k+1}), then the last row transform with the bottom and top
  // edge terms and a column pass that adds the left and right ones: on every
  // level (all), or on level 0 only with the level-1 ring values in ringS for
  // get_bc (step 2, which needs s_1 nowhere else).
  // the ring vectors of s0 on the levels below the finest, from the column
  // pair's output T2L: the bottom and top lines by one FFT per item where
  // the row length has a size-specialised pass of 256 threads
  // (k_ring_fromY_fft), by direct sums otherwise (k_ring_fromY)
  void ring_fromY() { ring_lines(T2L.p, Lmax, B * (Lmax - 1), ringV.p, ringSin.p, nullptr, 0); }
  // the ring lines of the items of Y (lm levels per case, levels 1..lm-1;
  // nitems = cases (lm - 1)) into V (ring_count entries per level), and at
  // their places of fld + case * fb when fld is set
  void ring_lines(const double* Y, int lm, int nitems, double* V, const double* sintab, double* fld, long fb) {
    const int NR = ring_count(m, n), n2 = fset<double>().pm.n2;
    const dim3 blk(256);
    const size_t lds = sizeof(double) * 2 * m;
    const int nlrg = (n / 2 + 1 + 3) / 4;
    const cplx<double>* tw = fset<double>().stw_m.p;
    if (fset<double>().spec_m && ring_fft_lines) {
      switch (n2) {
        case 256:
          hipLaunchKernelGGL(k_ring_fromY_fft<256>, dim3(nlrg + 1, nitems), blk, lds, st, geo, Y, lm, NR, sintab, V,
                             tw, nlrg, fld, fb);
          return;
        case 512:
          hipLaunchKernelGGL(k_ring_fromY_fft<512>, dim3(nlrg + 1, nitems), blk, lds, st, geo, Y, lm, NR, sintab, V,
                             tw, nlrg, fld, fb);
          return;
        case 1024:
          hipLaunchKernelGGL(k_ring_fromY_fft<1024>, dim3(nlrg + 1, nitems), blk, lds, st, geo, Y, lm, NR, sintab, V,
                             tw, nlrg, fld, fb);
          return;
        default: break;
      }
    }
    hipLaunchKernelGGL(k_ring_fromY, dim3((ring_fromY_tasks(m, n) + 3) / 4, nitems), blk, lds, st, geo, Y, lm, NR,
                       sintab, V, fld, fb);
  }
  // the edge spectra of Interp(R_{k+1}) for the items of im: by FFT on
  // square grids whose row length has a size-specialised pass of 256
  // threads (k_ring_spectra_fft), by direct sums otherwise (k_ring_spectra)
  void ring_spectra(ItemMap im, int nitems, int Kx) {
    const int NR = ringN, n2 = fset<double>().pm.n2;
    const cplx<double>* tw = fset<double>().stw_m.p;
    if (m == n && fset<double>().spec_m && ring_fft_lines) {
      const int P = n2 > 0 ? 1024 / n2 : 0, groups = P > 0 ? (4 + 2 * P - 1) / (2 * P) : 0;
      switch (n2) {
        case 256:
          hipLaunchKernelGGL(k_ring_spectra_fft<256>, dim3(groups, nitems), dim3(256), 0, st, geo, ringR.p, NR, im,
                             Kx, ringX.p, tw);
          return;
        case 512:
          hipLaunchKernelGGL(k_ring_spectra_fft<512>, dim3(groups, nitems), dim3(256), 0, st, geo, ringR.p, NR, im,
                             Kx, ringX.p, tw);
          return;
        case 1024:
          hipLaunchKernelGGL(k_ring_spectra_fft<1024>, dim3(groups, nitems), dim3(256), 0, st, geo, ringR.p, NR, im,
                             Kx, ringX.p, tw);
          return;
        default: break;
      }
    }
    hipLaunchKernelGGL(k_ring_spectra, dim3(4 * ((Kx + 31) / 32), nitems), dim3(256),
                       sizeof(double) * (Kx + 2 * (Kx + 1)), st, geo, ringR.p, NR, im, Kx, ringSin.p,
                       ringSin.p + 2 * m, ringX.p);
  }
  void vort2flux_levels(bool all, bool inc = false) {
    if constexpr (std::is_same_v<T, double>) {
      const int BL = B * Lmax, NR = ringN, Kx = std::max(m, n) - 1;
      coarsify(omega.p, B, gate, inc);
      launch<0>(RowOp<T, SrcField<T>, DstField<T>>{geoL, SrcField<T>{omega.p, nw}, DstField<T>{T1L.p, nw}, nullptr,
                                                   0},
                geo.nm1, 2, BL);
      launch_col2(Col2Op<T, SrcField<T>, MidScale<T>, DstField<T>>{geoL, SrcField<T>{T1L.p, nw},
                                                                   MidScale<T>{laminv.p}, DstField<T>{T2L.p, nw},
                                                                   nullptr, 0},
                  geo.mm1, 2, BL);
      ring_fromY();
      double* sring = all ? nullptr : ringS.p;
      switch (Lmax - 2) {
#define IBFS_RING_CHAIN(NP)                                                                                   \
  case NP:                                                                                                    \
    hipLaunchKernelGGL(k_ring_chain<NP>, dim3(NR, B), dim3(256), 0, st, geo, fbs(), NR, ringV.p, ringT.p,     \
                       ringR.p, sring, ring_sym_on ? (const int*)ringSym.p : nullptr);                        \
    break;
        IBFS_RING_CHAIN(0)
        IBFS_RING_CHAIN(1)
        IBFS_RING_CHAIN(2)
        IBFS_RING_CHAIN(3)
        IBFS_RING_CHAIN(4)
        IBFS_RING_CHAIN(5)
        IBFS_RING_CHAIN(6)
#undef IBFS_RING_CHAIN
        default: throw std::logic_error("vort2flux_levels: level count");
      }
      const ItemMap below{all ? Lmax - 1 : 1, Lmax}, out{all ? Lmax : 1, Lmax};
      const int nbelow = all ? B * (Lmax - 1) : B, nout = all ? BL : B;
      ring_spectra(below, nbelow, Kx);
      const double* Gy = m == n ? ringG.p : ringGp.p;
      if (!all && step2_window()) {
        // step 2 reads s_1 only on the vertices of the force stencils: the
        // last two transforms of level 0 on the rows and the columns of
        // their bounding rectangle (the rest of s_1 is rewritten in step 4
        // before anything reads it).  The row pass writes S, the column pass
        // the left/right terms into level 0 of ringS, which the force
        // right-hand side adds (the same sum as the column pass adding into
        // S), so the two passes are independent: one launch when both are
        // the same size-specialised pass.
        Geo gr = geoL, gc = geoL;
        gr.nm1 = swin[1] - swin[0] + 1;
        gc.mm1 = swin[3] - swin[2] + 1;
        const RowOp<T, SrcRowWin<SrcRingRow>, DstRowWin<DstRingItem>> rop{
            gr, {SrcRingRow{T2L.p, ringX.p, ringG.p, out, Kx}, geoL, swin[0]},
            {DstRingItem{S.p, out, 0}, geoL, swin[0]}, nullptr, 0};
        const ColOp<T, SrcColWin<SrcRingCol>, DstColWin<DstRingItem>> cop{
            gc, {SrcRingCol{ringX.p, Gy, below, Kx}, geoL, swin[2]},
            {DstRingItem{ringS.p, below, 0}, geoL, swin[2], swin[0], swin[1]}, nullptr, 0};
        if (!launch_dual(rop, gr.nm1, cop, gc.mm1, nout, nbelow)) {
          launch<0>(rop, gr.nm1, 2, nout);
          launch<1>(cop, gc.mm1, 2, nbelow, kCol);
        }
        return;
      }
      if (all && s_split) {
        // step 4: the last row pass writes S and the left/right column pass
        // writes its terms into ringS (free after step 2), which the tail
        // adds (the same sum s = S + C as the column pass adding into S);
        // the two passes are then independent: one launch where it can be
        const RowOp<T, SrcRingRow, DstRingItem> rop{geoL, SrcRingRow{T2L.p, ringX.p, ringG.p, out, Kx},
                                                    DstRingItem{S.p, out, 0}, nullptr, 0};
        const ColOp<T, SrcRingCol, DstRingItem> cop{geoL, SrcRingCol{ringX.p, Gy, below, Kx},
                                                    DstRingItem{ringS.p, below, 0}, nullptr, 0};
        if (!launch_dual(rop, geo.nm1, cop, geo.mm1, nout, nbelow)) {
          launch<0>(rop, geo.nm1, 2, nout);
          launch<1>(cop, geo.mm1, 2, nbelow, kCol);
        }
        return;
      }
      launch<0>(RowOp<T, SrcRingRow, DstRingItem>{geoL, SrcRingRow{T2L.p, ringX.p, ringG.p, out, Kx},
                                                  DstRingItem{S.p, out, 0}, nullptr, 0},
                geo.nm1, 2, nout);
      launch<1>(ColOp<T, SrcRingCol, DstRingItem>{geoL, SrcRingCol{ringX.p, Gy, below, Kx},
                                                  DstRingItem{S.p, below, 1}, nullptr, 0},
                geo.mm1, 2, nbelow, kCol);
    }
  }
  // vort2flux streamfunctions of an FP32 engine with every level solved in
  // FP64 into S64 (steps 2 and 4: the force system and the stored fluxes)
  void vort2flux64(T* om, int batch, const int* gate, bool inc = false, bool ring = false) {
    coarsify(om, batch, gate, inc);
    for (int lev = Lmax - 1; lev >= 0; lev--)
      ctci64<double>(om, (lev + 1 < Lmax) ? S64.p + (long)(lev + 1) * nw : nullptr, batch, gate, lev, ring);
  }
  // --------------------------------------------------------------- one step
  void enqueue_step(int first) override {
    enqueue_step1(first);
    enqueue_step2_cholesky();
    enqueue_step34();
  }

  // STEP 1 with the levels batched (see levels_can_batch): the previous
  // step's pending diagnostics row (with lastbc of every level in the head
  // kernel, or as side work of the first pass), the bulk solve of every
  // level with its old edge values (a row
  // pass, the lam1i column pair and an adding row pass over all (case,
  // level) items), then every level's correction for the new edge values of
  // its coarser level, from that level's bulk solve (BandOp)
  void enqueue_step1_levels(int first) {
    const int BL = B * Lmax;
    const bool defer = diag_deferred();
    // The freestream loader reads the edge values itself and takes the
    // diagnostics row as side work when its workgroups are whole waves;
    // otherwise the head kernel writes lastbc and the row first.
    const int nt = fs.spec_m ? ptile[kStep1Row] * fs.pm.n2 / 4 : kThreads;
    const bool headless = q0_uniform() && (src_split || nt % 64 == 0);
    if (!headless) {
      const int nlb = (nbc * Lmax + 255) / 256;
      hipLaunchKernelGGL(k_step_head<T>, dim3(nlb + (defer ? 3 : 0), B), dim3(256), 0, st, geo, omega.p, lastbc.p,
                         gate, diag_args(dpend.p), defer ? 1 : 0);
    }
    if (q0_uniform()) {
      SrcStep1<T, true> s1f{omega.p, nullptr, q.p, nullptr, lastbc.p, rhs_old.p, con1.p, con2.p, vfac.p,
                            nullptr, (long)nw, (long)nq, (long)nbc, 0, first};
      s1f.q0c = q0c.p;
      if (headless) {
        s1f.lev_items = Lmax;
        s1f.gdiag = geo;
        if (defer) {
          s1f.has_diag = 1;
          s1f.dg = diag_args(dpend.p);
        }
      }
      if (src_split) {
        // the source of every level as its own kernel (into T2L), then a
        // plain row pass
        const int sg = s1f.side_groups_for(256);  // a multiple of 8
        const int eg = ((nw + 255) / 256 + 7) & ~7;
        hipLaunchKernelGGL((k_src_eval<T, SrcStep1<T, true>>), dim3(sg + eg, BL), dim3(256), 0, st, geoL, s1f, T2L.p,
                           sg, (const int*)nullptr, 0);
        launch<0>(RowOp<T, SrcField<T>, DstField<T>>{geoL, SrcField<T>{T2L.p, nw}, DstField<T>{T1L.p, nw}, nullptr,
                                                     0},
                  geo.nm1, 2, BL);
      } else {
        launch<0>(RowOp<T, SrcStep1<T, true>, DstField<T>>{geoL, s1f, DstField<T>{T1L.p, nw}, nullptr, 0},
                  geo.nm1, 2, BL, kStep1Row);
      }
    } else {
      SrcStep1<T> s1{omega.p, nullptr, q.p, q0.p, lastbc.p, rhs_old.p, con1.p, con2.p, vfac.p,
                     nullptr, (long)nw, (long)nq, (long)nbc, 0, first};
      launch<0>(RowOp<T, SrcStep1<T>, DstField<T>>{geoL, s1, DstField<T>{T1L.p, nw}, nullptr, 0}, geo.nm1, 2, BL,
                kStep1Row);
    }
    step1_levels_add();
    // the correction for the new coarser edge values, on the bands only: one
    // fused column pair per band depth and pair of opposite edges, the
    // bottom and top bands first, then the left and right ones (the bands
    // cross in the corners, so the two adds there are kept in order)
    if (band_direct) {
      const int nbt = (m - 1 + 63) / 64, nlr = (n - 1 + 63) / 64;
      hipLaunchKernelGGL(k_band_direct<T>, dim3(2 * nbt + 2 * nlr, BL), dim3(64 * kBandDL), 0, st, geoL, omega.p, Lmax,
                         band_w.p, band_h.p, band_wmax, band_R, nbt, nlr);
      return;
    }
    launch_col2(BandOp<T>{geoL, omega.p, Lmax, band_rho_m.p, band_w.p, 0, 2, m}, 2 * band_wmax, 2, BL, kCol2, 0);
    launch_col2(BandOp<T>{geoL, omega.p, Lmax, band_rho_n.p, band_w.p, 2, 2, n}, 2 * band_wmax, 2, BL, kCol2, 1);
  }
  // omega += S (lam1i * S T1L) over every (case, level) item
  void step1_levels_add() {
    const int BL = B * Lmax;
    launch_col2(Col2Op<T, SrcField<T>, MidTab<T>, DstField<T>>{geoL, SrcField<T>{T1L.p, nw}, MidTab<T>{lam1i.p, nwp},
                                                               DstField<T>{T2L.p, nw}, nullptr, 0},
                geo.mm1, 2, BL, kStep1Col2);
    launch<0>(RowOp<T, SrcField<T>, DstAdd<T>>{geoL, SrcField<T>{T2L.p, nw}, DstAdd<T>{omega.p, nw, T(1)}, nullptr,
                                               0},
              geo.nm1, 2, BL);
  }

  // lastbc and STEP 1 of advance
  void enqueue_step1(int first) {
    if (lev_batch && !base) {
      enqueue_step1_levels(first);
      return;
    }
    const long FB = fbs(), QB = qbs(), BB = (long)Lmax * nbc;
    // freestream form: lastbc is side work of the first row pass (the
    // coarsest slot, which reads no level's lastbc that it writes)
    const bool fold_lastbc = !base && q0_uniform() && Lmax > 1;
    if (base)  // tangent step: the edge values of the base state before its own step
      hipLaunchKernelGGL(k_lastbc<T>, dim3((nbc * Lmax + 255) / 256, B), dim3(256), 0, st, geo, base->omega.p,
                         base->lastbc.p, gate);
    if (!fold_lastbc)
      hipLaunchKernelGGL(k_lastbc<T>, dim3((nbc * Lmax + 255) / 256, B), dim3(256), 0, st, geo, omega.p, lastbc.p,
                         gate);
    // STEP 1 (src/operators.f90:103-134), level slots Lmax..1, in increment
    // form: omega += S[lam1i * S(X + 2 vfac D0 omega)] (see SrcStep1)
    for (int lev = Lmax - 1; lev >= 0; lev--) {
      SrcStep1<T> s1{omega.p + (long)lev * nw,
                     lev + 1 < Lmax ? omega.p + (long)(lev + 1) * nw : nullptr,
                     q.p + (long)lev * nq,
                     q0.p + (long)lev * nq,
                     lastbc.p + (long)lev * nbc,
                     rhs_old.p + (long)lev * nw,
                     con1.p,
                     con2.p,
                     vfac.p,
                     gate,
                     FB,
                     QB,
                     BB,
                     lev,
                     first};
      if (base) {
        s1.bom = base->omega.p + (long)lev * nw;
        s1.bq = base->q.p + (long)lev * nq;
        s1.bq0 = base->q0.p + (long)lev * nq;
        s1.blb = base->lastbc.p + (long)lev * nbc;
      }
      if (!base && q0_uniform()) {  // freestream q0: the loader reads no q0
        SrcStep1<T, true> s1f{s1.om, s1.omc, s1.q, nullptr, s1.lastbc, s1.rhs_old, s1.con1, s1.con2, s1.vfac,
                              s1.Lc, s1.fbs, s1.qbs, s1.bbs, s1.lev, s1.first};
        s1f.q0c = q0c.p;
        if (fold_lastbc && lev == Lmax - 1) {
          s1f.lastbc_all = lastbc.p;
          s1f.om_all = omega.p;
          s1f.side_count = Lmax * nbc;
          if (diag_deferred()) {  // the previous step's pending row
            s1f.has_diag = 1;
            s1f.dg = diag_args(dpend.p);
          }
        }
        if (src_split && sizeof(T) == 8) {
          // the source as its own kernel (into T2, free until the column
          // pair below), then a plain row pass.  FP64 only: the FP32 loader
          // reads half the bytes and keeps its stencil in flight itself
          // (FP32 batch of 16: 9874 case-steps/s fused, 9680 split)
          const int sg = s1f.side_groups_for(256);  // a multiple of 8
          const int eg = ((nw + 255) / 256 + 7) & ~7;
          hipLaunchKernelGGL((k_src_eval<T, SrcStep1<T, true>>), dim3(sg + eg, B), dim3(256), 0, st, geo, s1f, T2.p,
                             sg, gate, lev);
          launch<0>(RowOp<T, SrcField<T>, DstField<T>>{geo, SrcField<T>{T2.p, nw}, DstField<T>{T1.p, nw}, gate, lev},
                    geo.nm1, 2, B);
        } else {
          launch<0>(RowOp<T, SrcStep1<T, true>, DstField<T>>{geo, s1f, DstField<T>{T1.p, nw}, gate, lev}, geo.nm1,
                    2, B, kStep1Row);
        }
      } else {
        launch<0>(RowOp<T, SrcStep1<T>, DstField<T>>{geo, s1, DstField<T>{T1.p, nw}, gate, lev}, geo.nm1, 2, B,
                  kStep1Row);
      }
      launch_col2(Col2Op<T, SrcField<T>, MidTab<T>, DstField<T>>{geo, SrcField<T>{T1.p, nw},
                                                                 MidTab<T>{lam1i.p + (long)lev * nwp, tbs()},
                                                                 DstField<T>{T2.p, nw}, gate, lev},
                  geo.mm1, 2, B, kStep1Col2);
      row_pass(SrcField<T>{T2.p, nw}, DstAdd<T>{omega.p + (long)lev * nw, FB, T(1)}, B, gate, lev);
    }
  }

  // STEP 2 (src/operators.f90:142-171) with the factor; motion is prepared
  // per call (and per step at the actuator points).  The force kernel also
  // writes the step-3 source RR = rot(reg(fb)).
  void enqueue_step2_cholesky() {
    enqueue_actuation();
    const long FB = fbs();
    const size_t lds = sizeof(double) * 2 * std::max(nf_max, 1);
    const dim3 gall(std::max(1, (nf_max + nvtx_max + kForceRows - 1) / kForceRows), B);
    const auto force = [&](auto s1) {
      using TS = std::remove_const_t<std::remove_pointer_t<decltype(s1)>>;
      // get_bc(s_2) reads level 1's ring only; the batched vort2flux keeps
      // those values in ringS (level 1 of S is not completed in step 2)
      const TS* s2 = Lmax > 1 ? s1 + nw : nullptr;
      const TS* s1c = nullptr;  // the left/right terms of s_1 on the force-stencil rectangle
      if constexpr (std::is_same_v<TS, double> && std::is_same_v<T, double>) {
        if (v2f_batch) s2 = ringS.p + nw;
        if (v2f_batch && step2_window()) s1c = ringS.p;
      }
      if (!refine) {
        // the right-hand side once per case, then the rows of [A^-1; G]
        if constexpr (std::is_same_v<TS, T>)
          hipLaunchKernelGGL((k_regT_curl<T>), dim3((std::max(nf_max, 1) + 255) / 256, B), dim3(256), 0, st, geo, sm,
                             s1, s2, FB, motion.p, rhsf.p, 0L, gate, -1, s1c);
        hipLaunchKernelGGL((k_force_solve<T, TS>), gall, dim3(kForceThreads), lds, st, geo, sm, s1, s2, FB, motion.p,
                           gate, Ainv.p, Gm.p, fb.p, RR.p, (long)nw, rhsf.p, (const double*)nullptr,
                           std::is_same_v<TS, T> ? 3 : 0);
      } else {  // one FP64 refinement sweep: fb += A^-1 (rhsf - A fb)
        const dim3 gf(std::max(1, (nf_max + kForceRows - 1) / kForceRows), B);
        hipLaunchKernelGGL((k_force_solve<T, TS>), gf, dim3(kForceThreads), lds, st, geo, sm, s1, s2, FB, motion.p,
                           gate, Ainv.p, Gm.p, fb.p, RR.p, (long)nw, rhsf.p, (const double*)nullptr, 1);
        hipLaunchKernelGGL(k_symv_residual, dim3((nf_max + 3) / 4, B), dim3(256), 0, st, sm, Asym.p, rhsf.p, fb.p,
                           rtmp.p);
        hipLaunchKernelGGL((k_force_solve<T, TS>), gall, dim3(kForceThreads), lds, st, geo, sm, s1, s2, FB, motion.p,
                           gate, Ainv.p, Gm.p, fb.p, RR.p, (long)nw, rhsf.p, (const double*)rtmp.p, 2);
      }
    };
    if constexpr (kSolve64) {
      vort2flux64(omega.p, B, gate, false, ring2);
      force((const double*)S64.p);
    } else if (v2f_batch) {
      vort2flux_levels(false);
      force((const T*)S.p);
    } else {
      vort2flux(omega.p, S.p, nullptr, B, gate, T2.p, T3.p, false, ring2);
      force((const T*)S.p);
    }
  }

  // STEPS 3 and 4, fluxes and diagnostics
  // (local3: step 3 may take the real-space form; the CG form keeps the
  // spectral one, whose iterate near the stopping test follows rounding)
  void enqueue_step34(bool local3 = true) {
    const long FB = fbs(), QB = qbs();
    // STEP 3 (src/operators.f90:174): in real space for the cases whose
    // active vertices are far enough from the edges (k_ainv_local), as the
    // spectral ctci for the others
    const int nloc = base == nullptr && local3 ? ainv_nloc : 0;
    if (nloc < B)
      ctci_like(SrcField<T>{RR.p, nw}, true, DstSub<T>{omega.p, FB, nloc > 0 ? ainv_loc.p : nullptr}, T2.p, T3.p, B,
                nullptr, 0, lam1i.p, tbs());
    if (nloc > 0)
      hipLaunchKernelGGL(k_ainv_local<T>, dim3(ainv_tiles, B), dim3(kAinvThreads), 0, st, geo, omega.p, FB,
                         (const T*)RR.p,
                         (long)nw, ainv_box.p, ainv_K.p, ainv_loc.p, ainv_R);
    // STEP 4 (src/operators.f90:177): streamfunctions here, fluxes in k_tail
    // (the coarsify chain only where step 3 changed its inputs when some
    // case took the real-space form)
    const bool inc = nloc > 0;
    if constexpr (kSolve64) vort2flux64(omega.p, B, gate, inc);
    else if (v2f_batch) vort2flux_levels(true, inc);
    else vort2flux(omega.p, S.p, nullptr, B, gate, T2.p, T3.p, inc);
    // fluxes and diagnostics (src/operators.f90:184-186)
    const int ncurl = tail_groups(geo) * Lmax;
    const bool defer = diag_deferred();
    const T* tailC = nullptr;  // the left/right terms kept apart by the batched vort2flux
    if constexpr (std::is_same_v<T, double>)
      if (v2f_batch && s_split) tailC = ringS.p;
    hipLaunchKernelGGL(k_tail<T>, dim3(ncurl + nchunk * Lmax, B), dim3(256), 0, st, geo, S.p, q.p, omega.p, FB, QB,
                       partial.p, nchunk, gate, kSolve64 ? (const double*)S64.p : nullptr,
                       defer ? dpend.p : (int*)nullptr, counter.p, tailC);
    if (!defer)
      hipLaunchKernelGGL(k_diag<T>, dim3(B), dim3(1024), 0, st, geo, diag_args(nullptr));
  }

  // the arguments of the diagnostics row (pend: the pending flags, or
  // nullptr to write the row unconditionally)
  DiagArgs<T> diag_args(int* pend, int clear = 0) const {
    return DiagArgs<T>{sm, q.p, qbs(), motion.p, fb.p, wsum.p, bdyno.p, partial.p, nchunk, gate,
                       diag.p, rowlen, counter.p, diag_cap, pend, clear};
  }
  // The row of a step is written by the next step's first row pass (side
  // work, three more workgroups) when the advance loop asks for it and that
  // pass exists: the freestream form with lastbc as side work, and a
  // workgroup size of a multiple of 64 for the row's reductions.
  bool diag_deferred() const {
    if (!side_diag || base) return false;
    if (lev_batch) return true;  // the head kernel of the level-batched step 1 writes it
    if (!q0_uniform() || Lmax < 2) return false;
    if (!fs.spec_m) return true;  // the generic pass: 256 threads
    const int nt = ptile[kStep1Row] * fs.pm.n2 / 4;
    return nt % 64 == 0;
  }
  void set_side_diag(bool on) override {
    if (on) dpend.zero(st);  // no row left pending by an interrupted call
    side_diag = on;
  }
  // writes the pending rows (after the last step of an advance call)
  void flush_diag() override {
    hipLaunchKernelGGL(k_diag<T>, dim3(B), dim3(1024), 0, st, geo, diag_args(dpend.p, 1));
  }

  // The tables of k_ainv_local: for each case the kernel of ainv on level 1
  // in long double (the closed form over kx, the sum over ky) for
  // 0 <= a, b <= R, and the box of its active vertices widened by R.  The
  // reach R comes from a bound rather than from the tabulated values, whose
  // last digits are rounding: every term of the ky sum has r_ky <= r_0 and
  // sqrt(A_ky^2 - 4 beta^2) >= that of ky = 0, so |K(a, b)| <= r_0^|a| / s_0,
  // and likewise in b (the lattice kernel is symmetric), so the places with
  // max(|a|, |b|) = t (8t of them) add at most 8t r_0^t / s_0; R is the
  // smallest with the sum of that over t > R below 1e-17.  A case is taken
  // when its active box keeps R cells from every edge and R <= kAinvRmax.
  // One R (the largest over the cases taken) for the launch.
  // Tangent-linear engines (base set) keep the spectral form
  // (enqueue_step34), so that their step stays the exact transpose of the
  // adjoint's.
  void ainv_tables(const std::vector<int>& vid) {
    ainv_nloc = 0;
    co_grid.clear();
    const char* env = std::getenv("IBFS_STEP3_LOCAL");
    if (env && env[0] == '0') return;
    const long double pi = 3.141592653589793238462643383279502884L;
    std::vector<std::vector<long double>> Ks(B);
    std::vector<int> Rb(B, -1), loc(B, 0), boxv(4 * B, 0);
    std::vector<long double> cy(2 * n);
    for (int k = 0; k < 2 * n; k++) cy[k] = cosl(pi * k / n);
    for (int b = 0; b < B; b++) {
      const CaseHost& c = cs[b];
      const CaseSm& h = hcs[b];
      if (h.nvtx == 0) continue;
      const long double bt = 0.5L * c.d.dt / (c.d.re * c.delta * c.delta);
      const long double A0 = 1.0L + 2.0L * bt, s0 = sqrtl(A0 * A0 - 4.0L * bt * bt), r0 = (A0 - s0) / (2.0L * bt);
      int R = 0;
      for (;; R++) {
        long double tail = 0.0L, pw = powl(r0, (long double)(R + 1));
        for (int t = R + 1; t < R + 400 && pw > 1e-40L; t++, pw *= r0) tail += 8.0L * t * pw / s0;
        if (tail < 1e-17L || R > kAinvRmax) break;
      }
      if (R > kAinvRmax) continue;
      std::vector<long double> Kt((size_t)(R + 1) * (R + 1), 0.0L);
      for (int k = 0; k < 2 * n; k++) {
        const long double A = 1.0L + bt * (4.0L - 2.0L * cy[k]), sq = sqrtl(A * A - 4.0L * bt * bt);
        const long double r = (A - sq) / (2.0L * bt);
        long double ra = 1.0L / sq;
        for (int aa = 0; aa <= R; aa++) {
          for (int bb = 0; bb <= R; bb++) Kt[(size_t)bb * (R + 1) + aa] += ra * cy[(size_t)((long)k * bb % (2 * n))];
          ra *= r;
        }
      }
      for (long double& v : Kt) v /= 2.0L * n;
      int i0 = m, i1 = -1, j0 = n, j1 = -1;
      for (int e = h.vbase; e < h.vbase + h.nvtx; e++) {
        const int a = vid[e], ii = a % (m - 1), jj = a / (m - 1);
        i0 = std::min(i0, ii), i1 = std::max(i1, ii), j0 = std::min(j0, jj), j1 = std::max(j1, jj);
      }
      if (i0 < R || j0 < R || i1 > m - 2 - R || j1 > n - 2 - R) continue;
      Ks[b] = std::move(Kt);
      Rb[b] = R;
      loc[b] = 1;
      boxv[4 * b] = i0, boxv[4 * b + 1] = j0, boxv[4 * b + 2] = i1, boxv[4 * b + 3] = j1;
    }
    int R = 0, tiles = 0, nloc = 0;
    for (int b = 0; b < B; b++)
      if (loc[b]) R = std::max(R, Rb[b]), nloc++;
    if (nloc == 0) return;
    // the target box at the launch's R, inside the field (places past the
    // case's own reach get exact zeros)
    for (int b = 0; b < B; b++)
      if (loc[b]) {
        const int i0 = std::max(boxv[4 * b] - R, 0), j0 = std::max(boxv[4 * b + 1] - R, 0);
        const int i1 = std::min(boxv[4 * b + 2] + R, m - 2), j1 = std::min(boxv[4 * b + 3] + R, n - 2);
        boxv[4 * b] = i0, boxv[4 * b + 1] = j0, boxv[4 * b + 2] = i1 - i0 + 1, boxv[4 * b + 3] = j1 - j0 + 1;
        const int ntx = (boxv[4 * b + 2] + kAinvTX - 1) / kAinvTX, nty = (boxv[4 * b + 3] + kAinvTY - 1) / kAinvTY;
        tiles = std::max(tiles, ntx * nty);
      }
    std::vector<double> kh((size_t)B * (R + 1) * (R + 1), 0.0);
    for (int b = 0; b < B; b++)
      if (loc[b])
        for (int bb = 0; bb <= Rb[b]; bb++)
          for (int a = 0; a <= Rb[b]; a++)
            kh[((size_t)b * (R + 1) + bb) * (R + 1) + a] = (double)Ks[b][(size_t)bb * (Rb[b] + 1) + a];
    upvec(ainv_K, kh, st);
    upvec(ainv_loc, loc, st);
    upvec(ainv_box, boxv, st);
    ainv_R = R;
    ainv_tiles = tiles;
    ainv_nloc = nloc;
    coarsify_tables(loc, boxv);
  }
  // co_tiles and co_grid from the target boxes of step 3 (boxv, interior
  // indices {i0, j0, nx, ny}; vertex i = index + 2).  Level k's centre place
  // (i, j) reads level k-1's places 2i - m/2 - 2 .. 2i - m/2 (and likewise
  // in j), so the changed range [lo, hi] of level k-1 makes level k's
  // [ceil((lo + m/2) / 2), floor((hi + m/2 + 2) / 2)] within the centre
  // [m/4 + 2, 3m/4].  A tile launch over levels lev and lev+1 runs the tiles
  // of level lev+1's range: a level lev place p is written by the tile of
  // level lev+1's place floor((p + m/2 + 1) / 2), whose stencil holds p, so
  // those tiles hold every changed place of level lev too.
  void coarsify_tables(const std::vector<int>& loc, const std::vector<int>& boxv) {
    const int ntx = (m / 2 - 1 + kCoTX - 1) / kCoTX, nty = (n / 2 - 1 + kCoTY - 1) / kCoTY;
    int nlaunch = 0;
    for (int lev = 1; lev + 1 < Lmax; lev += 2) nlaunch++;
    if (nlaunch == 0) return;
    std::vector<int> tv((size_t)4 * B * nlaunch, 0);
    co_grid.assign(nlaunch, 0);
    for (int b = 0; b < B; b++) {
      int r[4] = {0, 0, 0, 0};  // the changed range of the current level {ilo, ihi, jlo, jhi}
      bool any = loc[b] != 0;
      if (any) {
        r[0] = boxv[4 * b] + 2, r[1] = boxv[4 * b] + boxv[4 * b + 2] + 1;
        r[2] = boxv[4 * b + 1] + 2, r[3] = boxv[4 * b + 1] + boxv[4 * b + 3] + 1;
      }
      for (int lev = 1, t = 0; lev < Lmax; lev++) {
        if (any) {
          r[0] = std::max((r[0] + m / 2 + 1) / 2, m / 4 + 2), r[1] = std::min((r[1] + m / 2 + 2) / 2, 3 * m / 4);
          r[2] = std::max((r[2] + n / 2 + 1) / 2, n / 4 + 2), r[3] = std::min((r[3] + n / 2 + 2) / 2, 3 * n / 4);
          any = r[0] <= r[1] && r[2] <= r[3];
        }
        if (lev % 2 == 0 && t < nlaunch) {  // the launch over levels lev-1 and lev: level lev's tiles
          int* o = tv.data() + 4 * ((size_t)t * B + b);
          if (!loc[b]) {
            o[0] = o[1] = 0, o[2] = ntx, o[3] = nty;
          } else if (any) {
            o[0] = (r[0] - (m / 4 + 2)) / kCoTX, o[1] = (r[2] - (n / 4 + 2)) / kCoTY;
            o[2] = (r[1] - (m / 4 + 2)) / kCoTX - o[0] + 1, o[3] = (r[3] - (n / 4 + 2)) / kCoTY - o[1] + 1;
          }
          co_grid[t] = std::max(co_grid[t], o[2] * o[3]);
          t++;
        }
      }
    }
    upvec(co_tiles, tv, st);
  }
  // the compact smearing tables of every case, concatenated, on the device
  // (and the per-case offsets of the force system that depend on them)
  void upload_smear() {
    nf_max = nvtx_max = 0;
    std::vector<int> fid, fptr, fpt, rptr, rface, ria, rib, vid, vface, bd;
    std::vector<double> fw, rw, ws;
    hcs.resize(B);
    long loff = 0, goff = 0;
    int foff = 0, hof = 0;
    hoff.resize(B);
    for (int b = 0; b < B; b++) {
      const CaseHost& c = cs[b];
      CaseSm& h = hcs[b];
      h.nact = (int)c.face_id.size();
      h.fbase = (int)fid.size();
      h.pbase = (int)fptr.size();
      const int ebase = (int)fpt.size();
      fid.insert(fid.end(), c.face_id.begin(), c.face_id.end());
      for (int v : c.face_ptr) fptr.push_back(v + ebase);
      fpt.insert(fpt.end(), c.face_pt.begin(), c.face_pt.end());
      fw.insert(fw.end(), c.face_w.begin(), c.face_w.end());
      h.nf = c.nf;
      h.rbase = (int)rptr.size();
      const int rebase = (int)rface.size();
      for (int v : c.frc_ptr) rptr.push_back(v + rebase);
      rface.insert(rface.end(), c.frc_face.begin(), c.frc_face.end());
      ria.insert(ria.end(), c.frc_ia.begin(), c.frc_ia.end());
      rib.insert(rib.end(), c.frc_ib.begin(), c.frc_ib.end());
      rw.insert(rw.end(), c.frc_w.begin(), c.frc_w.end());
      h.nvtx = (int)c.vtx_id.size();
      h.vbase = (int)vid.size();
      vid.insert(vid.end(), c.vtx_id.begin(), c.vtx_id.end());
      vface.insert(vface.end(), c.vtx_face.begin(), c.vtx_face.end());
      h.foff = foff;
      h.nb = c.d.nb;
      h.loff = loff;
      h.goff = goff;
      h.nbody = c.nbody;
      h.bmin = c.bmin;
      h.ncol = c.ncol;
      h.hoff = hof;
      hoff[b] = hof;
      h.dt = c.d.dt;
      h.delta = c.delta;
      ws.insert(ws.end(), c.wsum.begin(), c.wsum.begin() + c.nf);
      bd.insert(bd.end(), c.d.bdyno.begin(), c.d.bdyno.end());
      foff += c.nf;
      loff += (long)c.nf * c.nf;
      goff += (long)h.nvtx * c.nf;
      hof += c.ncol;
      nf_max = std::max(nf_max, c.nf);
      nvtx_max = std::max(nvtx_max, h.nvtx);
    }
    {
      const char* env = std::getenv("IBFS_S_WINDOW");
      swin_ok = !(env && env[0] == '0') && !ria.empty();
      swin[0] = swin[2] = std::max(m, n);
      swin[1] = swin[3] = -1;
      for (const std::vector<int>* v : {&ria, &rib})
        for (int a : *v) {
          if (a < 0) {
            swin_ok = false;
            break;
          }
          const int ii = a % (m - 1), jj = a / (m - 1);
          swin[0] = std::min(swin[0], jj);
          swin[1] = std::max(swin[1], jj);
          swin[2] = std::min(swin[2], ii);
          swin[3] = std::max(swin[3], ii);
        }
      if (swin[1] < swin[0] || swin[1] >= n - 1 || swin[3] >= m - 1) swin_ok = false;
      // sequences 2p and 2p+1 of a pass share one complex FFT, whose
      // rounding mixes the two: the window starts on an even row and column
      // and ends on an odd one (or the last), so that it pairs them as the
      // whole-field pass does
      swin[0] &= ~1;
      swin[2] &= ~1;
      swin[1] = std::min(swin[1] | 1, n - 2);
      swin[3] = std::min(swin[3] | 1, m - 2);
    }
    ainv_tables(vid);
    nf_tot = foff;
    mat_tot = loff;
    gmat_tot = goff;
    rowlen = hof;
    upvec(face_id, fid, st);
    upvec(face_ptr, fptr, st);
    upvec(face_pt, fpt, st);
    upvec(face_w, fw, st);
    upvec(frc_ptr, rptr, st);
    upvec(frc_face, rface, st);
    upvec(frc_w, rw, st);
    upvec(frc_ia, ria, st);
    upvec(frc_ib, rib, st);
    upvec(vtx_id, vid, st);
    upvec(vtx_face, vface, st);
    upvec(wsum, ws, st);
    upvec(bdyno, bd, st);
    upvec(csm, hcs, st);
    sm = Smear{face_id.p, face_ptr.p, face_pt.p, face_w.p, frc_ptr.p, frc_face.p, frc_w.p, frc_ia.p, frc_ib.p,
               vtx_id.p,   vtx_face.p, csm.p, hcs[0], B == 1};
    // the padded force gather (slot-major), when no force has more than kFrcK entries
    const size_t nslots = (size_t)std::max(foff, 1) * kFrcK;
    std::vector<int> p9f(nslots, 0), p9a(nslots, 0), p9b(nslots, 0);
    std::vector<double> p9w(nslots, 0.0);
    bool pad_ok = true;
    for (int b = 0; b < B && pad_ok; b++) {
      const CaseHost& c = cs[b];
      for (int f = 0; f < c.nf && pad_ok; f++) {
        const int lo = c.frc_ptr[f], hi = c.frc_ptr[f + 1];
        if (hi - lo > kFrcK) {
          pad_ok = false;
          break;
        }
        for (int e = lo; e < hi; e++) {
          const size_t slot = (size_t)(e - lo) * std::max(foff, 1) + hcs[b].foff + f;
          p9f[slot] = c.frc_face[e];
          p9w[slot] = c.frc_w[e];
          p9a[slot] = c.frc_ia[e];
          p9b[slot] = c.frc_ib[e];
        }
      }
    }
    if (pad_ok) {
      upvec(f9_face, p9f, st);
      upvec(f9_w, p9w, st);
      upvec(f9_ia, p9a, st);
      upvec(f9_ib, p9b, st);
      sm.f9_face = f9_face.p;
      sm.f9_w = f9_w.p;
      sm.f9_ia = f9_ia.p;
      sm.f9_ib = f9_ib.p;
      sm.f9_n = std::max(foff, 1);
    }
  }

  // new positions of the points of case c (one-case handles): what advance
  // does after move_body when a body moves (src/operators.f90:142-144):
  // setup_reg's tables rebuilt (src/grid.f90:142-230) and uploaded.  The
  // factor of the old geometry no longer applies; the step-3 source is
  // cleared because the set of vertices next to the body changes.
  void move_body(int c, const double* xb, const double* yb) override {
    if (B != 1) throw ArgError{IBFS_EINVAL, "moving bodies need a one-case handle"};
    CaseHost& h = cs[c];
    for (int k = 0; k < h.d.nb; k++) {
      h.d.xb[k] = xb[k];
      h.d.yb[k] = yb[k];
    }
    HIPCHK(hipStreamSynchronize(st));
    build_smear(h);
    upload_smear();
    cg_drop_graphs();  // they hold the old tables
    cg_moved = true;   // A follows the body from now on: a_times in every CG product
    cg_have_dense = false;
    Gm.alloc(std::max<long>(gmat_tot, 1));  // rot(reg(A^-1)) rows follow the new vertex set
    h.have_chol = false;
    RR.zero(st);
    HIPCHK(hipStreamSynchronize(st));
  }

  // ------------------------------------------------ conjugate-gradient step
  // advance with the force system solved by cjgr (stationary = F,
  // src/operators.f90:169-171) instead of the factor: one case (B == 1),
  // warm-started from the current fb.  The loop runs on the host, which
  // reads the CG scalars back after every iteration and applies the
  // reference's stopping test (iter < cg_max_iter and delta_new > cgtol^2).
  DevBuf<T> cgRR, cgVB, cgSB, cgT1, cgT2;
  DevBuf<double> cg_r, cg_d, cg_q;
  DevBuf<CgState> cg_dev;
  CgState* cg_host = nullptr;  // pinned
  // one CG iteration as a graph: [0] r -= alpha q, [1] the residual
  // recomputed from b - A x (every 50th iteration)
  hipGraphExec_t cg_exec[2] = {nullptr, nullptr};
  static constexpr int kCgBatch = 16;  // a_times form: iterations replayed between two read-backs
  // When no point moves, A is the same every step: it is built once on the
  // device (the columns of preprocess, build_columns) and each CG product is
  // a dense GEMV (k_gemv_dense) instead of the a_times composition.
  DevBuf<double> cgA, cgPart;
  int cg_nslice = 1;
  bool cg_have_dense = false;
  bool cg_moved = false;  // move_body was called: A changes, a_times every product
  bool cg_dense() const {
    if (cg_moved) return false;
    for (int b : cs[0].d.bdyno)
      if (b > 0) return false;
    return true;
  }

  void cg_alloc() {
    if (cg_host) return;
    if (B != 1) throw ArgError{IBFS_EINVAL, "the conjugate-gradient force solve runs one case per handle"};
    cgRR.alloc(nw); cgRR.zero(st);
    cgVB.alloc(fbs()); cgVB.zero(st);
    cgSB.alloc(fbs());
    cgT1.alloc(nw);
    cgT2.alloc(nw);
    const size_t nfv = std::max(nf_max, 1);
    cg_r.alloc(nfv); cg_d.alloc(nfv); cg_q.alloc(nfv);
    cg_dev.alloc(1);
    HIPCHK(hipHostMalloc((void**)&cg_host, sizeof(CgState), hipHostMallocDefault));
  }
  void cg_drop_graphs() {
    for (auto& e : cg_exec) {
      if (e) (void)hipGraphExecDestroy(e);
      e = nullptr;
    }
  }
  void cg_release() {
    cg_drop_graphs();
    if (cg_host) (void)hipHostFree(cg_host);
    cg_host = nullptr;
  }
  // The a_times form: one graph per iteration kind ([0] r -= alpha q, [1]
  // r = b - A x).  The dense form: one graph of kCgDense = 50 iterations
  // aligned with cjgr's residual recompute: 49 iterations of the product
  // A d and the fused iteration kernel (k_cg_iter), then the 50th with the
  // product A d, alpha, the product A x and the recomputed residual.  The
  // kernels read the done flag on the device, and every replay starts at an
  // iteration count that is a multiple of 50.
  static constexpr int kCgDense = 50;
  void cg_capture() {
    const int nf = cs[0].nf;
    if (cg_have_dense) {
      if (cg_exec[0]) return;
      hipGraph_t gr = nullptr;
      const dim3 gg((nf + 255) / 256, cg_nslice);
      HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      for (int k = 0; k < kCgDense; k++) {
        hipLaunchKernelGGL(k_gemv_dense, gg, dim3(256), 0, st, nf, cgA.p, cg_d.p, cgPart.p, cg_nslice,
                           (const CgState*)cg_dev.p, 0);
        if (k + 1 < kCgDense) {
          hipLaunchKernelGGL(k_cg_iter, dim3(1), dim3(256), 0, st, nf, cg_d.p, cg_q.p, fb.p, cg_r.p, cg_dev.p,
                             (const double*)cgPart.p, cg_nslice);
          continue;
        }
        hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(256), 0, st, nf, cg_d.p, cg_q.p, fb.p, cg_dev.p,
                           (const double*)cgPart.p, cg_nslice);
        hipLaunchKernelGGL(k_gemv_dense, gg, dim3(256), 0, st, nf, cgA.p, (const double*)fb.p, cgPart.p, cg_nslice,
                           (const CgState*)cg_dev.p, 1);
        hipLaunchKernelGGL(k_cg_resid, dim3(1), dim3(256), 0, st, nf, rhsf.p, cg_q.p, cg_r.p, cg_d.p, cg_dev.p, -1,
                           (const double*)cgPart.p, cg_nslice);
      }
      HIPCHK(hipStreamEndCapture(st, &gr));
      HIPCHK(hipGraphInstantiate(&cg_exec[0], gr, nullptr, nullptr, 0));
      HIPCHK(hipGraphDestroy(gr));
      return;
    }
    for (int v = 0; v < 2; v++) {
      if (cg_exec[v]) continue;
      hipGraph_t gr = nullptr;
      HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      a_times_one(cg_d.p, cg_q.p);
      hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(256), 0, st, nf, cg_d.p, cg_q.p, fb.p, cg_dev.p,
                         (const double*)nullptr, 1);
      if (v) a_times_one(fb.p, cg_q.p);
      hipLaunchKernelGGL(k_cg_resid, dim3(1), dim3(256), 0, st, nf, rhsf.p, cg_q.p, cg_r.p, cg_d.p, cg_dev.p, v,
                         (const double*)nullptr, 1);
      HIPCHK(hipStreamEndCapture(st, &gr));
      HIPCHK(hipGraphInstantiate(&cg_exec[v], gr, nullptr, nullptr, 0));
      HIPCHK(hipGraphDestroy(gr));
    }
  }
  // A of the current geometry on the device, once (the graphs of the
  // a_times form are dropped)
  void cg_build_dense() {
    if (cg_have_dense || !cg_dense()) return;
    const int nf = cs[0].nf;
    cgA.alloc((size_t)nf * nf);
    build_columns(0, cgA.p);
    cg_nslice = (nf + kGemvCols - 1) / kGemvCols;  // at most kGemvCols columns per slice
    cgPart.alloc((size_t)cg_nslice * nf);
    cg_drop_graphs();
    cg_have_dense = true;
  }

  // out = a_times(x) (src/operators.f90:444-462) for case 0: ainv(rot(reg(x)))
  // on level 1, vort2flux, regT of the level-1 fluxes
  void a_times_one(const double* x, double* out) {
    const int nf = cs[0].nf;
    hipLaunchKernelGGL((k_rotreg<T, double>), dim3((hcs[0].nvtx + 255) / 256, 1), dim3(256), 0, st, sm, x, 0L,
                       cgRR.p, (long)nw, 0);
    ctci_like(SrcField<T>{cgRR.p, nw}, true, DstField<T>{cgVB.p, fbs()}, cgT1.p, cgT2.p, 1, nullptr, 0, lam1i.p, 0);
    vort2flux(cgVB.p, cgSB.p, nullptr, 1, gate, cgT1.p, cgT2.p);
    hipLaunchKernelGGL(k_regT_curl<T>, dim3((nf + 255) / 256, 1), dim3(256), 0, st, geo, sm, cgSB.p,
                       Lmax > 1 ? cgSB.p + nw : nullptr, fbs(), (const double*)nullptr, out, (long)nf, gate, 0,
                       (const T*)nullptr);
  }

  CgState cg_read() {
    HIPCHK(hipMemcpyAsync(cg_host, cg_dev.p, sizeof(CgState), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return *cg_host;
  }

  // fb <- cjgr(fb, rhsf); returns the iteration count.  Iterations are
  // replayed from the captured graphs kCgBatch at a time, then the state is
  // read back; replays after the loop test fails are no-ops (CgState::done),
  // so the iterate is the one the loop would stop at.  A replay that is not
  // a no-op performs iteration c.iter + k + 1, which fixes its variant.
  int cg_solve(double cgtol, int cg_max_iter) {
    const int nf = cs[0].nf;
    if (nf == 0) return 0;
    cg_build_dense();
    cg_capture();
    *cg_host = CgState{};
    cg_host->eps = cgtol * cgtol;
    cg_host->max_iter = cg_max_iter;
    HIPCHK(hipMemcpyAsync(cg_dev.p, cg_host, sizeof(CgState), hipMemcpyHostToDevice, st));
    if (cg_have_dense) {
      hipLaunchKernelGGL(k_gemv_dense, dim3((nf + 255) / 256, cg_nslice), dim3(256), 0, st, nf, cgA.p,
                         (const double*)fb.p, cgPart.p, cg_nslice, (const CgState*)nullptr, 0);
      hipLaunchKernelGGL(k_cg_start, dim3(1), dim3(256), 0, st, nf, rhsf.p, cg_q.p, cg_r.p, cg_d.p, cg_dev.p,
                         (const double*)cgPart.p, cg_nslice);
      CgState c = cg_read();
      for (int guard = 0; !c.done && guard <= cg_max_iter; guard += kCgDense) {
        HIPCHK(hipGraphLaunch(cg_exec[0], st));
        c = cg_read();
      }
      return c.iter;
    }
    a_times_one(fb.p, cg_q.p);
    hipLaunchKernelGGL(k_cg_start, dim3(1), dim3(256), 0, st, nf, rhsf.p, cg_q.p, cg_r.p, cg_d.p, cg_dev.p,
                       (const double*)nullptr, 1);
    CgState c = cg_read();
    for (int guard = 0; !c.done && guard <= cg_max_iter; guard += kCgBatch) {
      for (int k = 0; k < kCgBatch; k++) HIPCHK(hipGraphLaunch(cg_exec[(c.iter + k + 1) % 50 == 0], st));
      c = cg_read();
    }
    return c.iter;
  }

  // one step with the CG force solve (eager launches); returns the
  // iteration count
  int step_cg(int first, double cgtol, int cg_max_iter) override {
    if constexpr (!std::is_same_v<T, double>)
      throw ArgError{IBFS_EINVAL, "the conjugate-gradient force solve needs the plain FP64 form"};
    cg_alloc();
    enqueue_step1(first);
    enqueue_actuation();
    // STEP 2 (src/operators.f90:155-164): rhsf = regT(q_1) - motion
    vort2flux(omega.p, S.p, nullptr, B, gate, T2.p, T3.p);
    hipLaunchKernelGGL(k_regT_curl<T>, dim3((nf_max + 255) / 256, B), dim3(256), 0, st, geo, sm, S.p,
                       Lmax > 1 ? S.p + nw : nullptr, fbs(), (const double*)motion.p, rhsf.p, 0L, gate, -1,
                       (const T*)nullptr);
    const int it = cg_solve(cgtol, cg_max_iter);
    // the step-3 source rot(reg(fb)) on the active vertices
    hipLaunchKernelGGL((k_rotreg<T, double>), dim3((nvtx_max + 255) / 256, B), dim3(256), 0, st, sm, fb.p, 0L, RR.p,
                       (long)nw, -1);
    enqueue_step34(false);
    return it;
  }

  void enqueue_motion() override {
    hipLaunchKernelGGL(k_motion<T>, dim3((nf_max + 255) / 256, B), dim3(256), 0, st, sm, q0.p, qbs(), motion_in.p,
                       motion.p, rq0.p);
  }

  // The actuators of case c (src/grid.f90:579-633): n_act slots whose points
  // are the last sum(npt) points of the case, in slot order (as actuation
  // addresses them, nb - n_act_pt + count); par holds each slot's namelist
  // read_act_parameters (x_j, y_j, h_j, u_j, u_j_osc, theta_j, omg_j).
  void set_actuators(int c, int n_act, const int* npt, const double* par) override {
    const CaseHost& h = cs[c];
    int tot = 0;
    for (int a = 0; a < n_act; a++) {
      if (npt[a] < 1) throw ArgError{IBFS_EINVAL, "an actuator slot needs at least one point"};
      tot += npt[a];
    }
    if (tot > h.d.nb) throw ArgError{IBFS_EINVAL, "actuator slots have more points than the case"};
    std::vector<ActPt> pts;
    int count = 0;
    for (int a = 0; a < n_act; a++) {
      const double* p = par + 7 * a;
      for (int k = 0; k < 7; k++)
        if (!std::isfinite(p[k])) throw ArgError{IBFS_EINVAL, "an actuator parameter is not finite"};
      const double th = kPiRef / 180.0 * p[5];  // u_jet: pi/180.d0*theta_j
      for (int k = 0; k < npt[a]; k++, count++)
        pts.push_back(ActPt{h.d.nb - tot + count, p[3], p[4], p[6] * h.d.dt, std::cos(th), std::sin(th)});
    }
    hact[c] = std::move(pts);
    std::vector<ActPt> all;
    std::vector<int> ptr{0};
    for (int b = 0; b < B; b++) {
      all.insert(all.end(), hact[b].begin(), hact[b].end());
      ptr.push_back((int)all.size());
    }
    has_act = !all.empty();
    upvec(act_pts, all, st);
    upvec(act_ptr, ptr, st);
    if (!dit.p) dit.alloc(B);
    HIPCHK(hipStreamSynchronize(st));
  }
  bool has_actuators() const override { return has_act; }
  // the step counter k_actuation reads (the itime of the step before the next)
  void set_itime(int itime) override {
    if (!has_act) return;
    std::vector<int> v(B, itime);
    dit.upload(v.data(), B, st);
  }
  void enqueue_actuation() {
    if (has_act)
      hipLaunchKernelGGL(k_actuation, dim3(B), dim3(256), 0, st, sm, act_ptr.p, act_pts.p, rq0.p, dit.p, motion.p);
  }
  void set_motion(int c, const double* mo) override {
    if (mo) motion_in.upload(mo, cs[c].nf, st, hcs[c].foff);
    else if (cs[c].nf) HIPCHK(hipMemsetAsync(motion_in.p + hcs[c].foff, 0, sizeof(double) * cs[c].nf, st));
  }

  void ensure_diag(int cap) override {
    if (cap <= diag_cap) return;
    diag.alloc((size_t)cap * std::max(rowlen, 1));
    diag_cap = cap;
  }
  void reset_counters() override { counter.zero(st); }
  void download_diag(double* dst, int nsteps, int ldrow, int col0) override {
    std::vector<double> h((size_t)nsteps * rowlen);
    diag.download(h.data(), h.size(), st);
    HIPCHK(hipStreamSynchronize(st));
    for (int s = 0; s < nsteps; s++)
      std::memcpy(dst + (size_t)s * ldrow + col0, h.data() + (size_t)s * rowlen, sizeof(double) * rowlen);
  }

  // ----------------------------------------------------------------- state
  template <typename Src, typename Dst>
  static void convert(const Src* a, Dst* b, size_t n) {
    for (size_t i = 0; i < n; i++) b[i] = Dst(a[i]);
  }
  void put(DevBuf<T>& d, size_t off, const double* h, size_t cnt) {
    if (!h || !cnt) return;
    if (sizeof(T) == 8) {
      d.upload(reinterpret_cast<const T*>(h), cnt, st, off);
      HIPCHK(hipStreamSynchronize(st));
    } else {
      std::vector<T> t(cnt);
      convert(h, t.data(), cnt);
      d.upload(t.data(), cnt, st, off);
      HIPCHK(hipStreamSynchronize(st));
    }
  }
  void get(const DevBuf<T>& d, size_t off, double* h, size_t cnt) {
    if (!h || !cnt) return;
    if (sizeof(T) == 8) {
      d.download(reinterpret_cast<T*>(h), cnt, st, off);
      HIPCHK(hipStreamSynchronize(st));
    } else {
      std::vector<T> t(cnt);
      d.download(t.data(), cnt, st, off);
      HIPCHK(hipStreamSynchronize(st));
      convert(t.data(), h, cnt);
    }
  }
  void set_state(int c, const double* om, const double* ro, const double* qq, const double* qq0,
                 const double* f) override {
    const size_t W = (size_t)cs[c].L * nw, Q = (size_t)cs[c].L * nq;
    put(omega, (size_t)c * fbs(), om, W);
    put(rhs_old, (size_t)c * fbs(), ro, W);
    put(q, (size_t)c * qbs(), qq, Q);
    put(q0, (size_t)c * qbs(), qq0, Q);
    if (qq0) {  // is the new q0 in the freestream form (compared in the engine's element type)?
      const size_t nu = (size_t)(m + 1) * n;
      bool uni = true;
      for (int k = 0; k < cs[c].L && uni; k++) {
        const double* lv = qq0 + (size_t)k * nq;
        const T u0 = T(lv[0]);
        for (size_t e = 0; e < nu && uni; e++) uni = T(lv[e]) == u0;
        for (size_t e = nu; e < (size_t)nq && uni; e++) uni = T(lv[e]) == T(0);
        hq0c[(size_t)c * Lmax + k] = u0;
      }
      q0u[c] = uni;
      q0c.upload(hq0c.data(), hq0c.size(), st);
      HIPCHK(hipStreamSynchronize(st));
    }
    if (f && cs[c].nf) {
      fb.upload(f, cs[c].nf, st, hcs[c].foff);
      HIPCHK(hipStreamSynchronize(st));
    }
  }
  void get_state(int c, double* om, double* ro, double* qq, double* qq0, double* f) override {
    const size_t W = (size_t)cs[c].L * nw, Q = (size_t)cs[c].L * nq;
    get(omega, (size_t)c * fbs(), om, W);
    get(rhs_old, (size_t)c * fbs(), ro, W);
    get(q, (size_t)c * qbs(), qq, Q);
    get(q0, (size_t)c * qbs(), qq0, Q);
    if (f && cs[c].nf) {
      fb.download(f, cs[c].nf, st, hcs[c].foff);
      HIPCHK(hipStreamSynchronize(st));
    }
  }

  // ---------------------------------------------------------- force system
  // Li = L^{-1} of the factor of case c (and, for the refinement sweep, the
  // symmetric matrix the factor was built from: the upper triangle of the
  // ibfs.chd matrix, src/operators.f90:779).
  void upload_factor(int c) override {
    CaseHost& h = cs[c];
    const int nf = h.nf;
    if (nf == 0) { h.have_chol = true; return; }
    std::vector<double> Li((size_t)nf * nf, 0.0), col(nf);
    const double* a = h.cholmat.data();
    auto Lf = [&](int i, int k) -> double { return i == k ? h.cholvec[i] : a[(size_t)i + (size_t)k * nf]; };
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
    Linv.upload(Li.data(), Li.size(), st, hcs[c].loff);
    const long nn = (long)nf * nf, ng = (long)hcs[c].nvtx * nf;
    hipLaunchKernelGGL(k_make_ainv, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, sm, c, Linv.p, Ainv.p);
    if (ng > 0)
      hipLaunchKernelGGL(k_make_g, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, st, sm, c, Ainv.p, Gm.p);
    HIPCHK(hipGetLastError());
    if (refine) {
      std::vector<double> As((size_t)nf * nf);
      for (int j = 0; j < nf; j++)
        for (int i = 0; i < nf; i++) As[i + (size_t)j * nf] = i <= j ? a[i + (size_t)j * nf] : a[j + (size_t)i * nf];
      Asym.upload(As.data(), As.size(), st, hcs[c].loff);
    }
    HIPCHK(hipStreamSynchronize(st));
    h.have_chol = true;
  }

  // a_times (src/operators.f90:444-462) for the unit vectors of case c,
  // batched over columns; A (nf x nf, column-major) written to acol (device).
  void build_columns(int c, double* acol) {
    const int nf = cs[c].nf, L = cs[c].L;
    const int Bc = std::max(1, std::min(nf, 64));
    const long bsL = fbs();
    DevBuf<double> F;
    DevBuf<T> RRb, t1, t2, VB, SB;
    DevBuf<int> gate;
    F.alloc((size_t)Bc * nf);
    RRb.alloc((size_t)Bc * nw);
    t1.alloc((size_t)Bc * nw);
    t2.alloc((size_t)Bc * nw);
    VB.alloc((size_t)Bc * bsL);
    SB.alloc((size_t)Bc * bsL);
    RRb.zero(st);
    VB.zero(st);
    std::vector<int> hg(Bc, L);
    upvec(gate, hg, st);
    std::vector<double> hF((size_t)Bc * nf);
    const T* l1i = lam1i.p + (long)c * tbs();  // level 1 of case c
    for (int c0 = 0; c0 < nf; c0 += Bc) {
      const int bb = std::min(Bc, nf - c0);
      std::fill(hF.begin(), hF.end(), 0.0);
      for (int b = 0; b < bb; b++) hF[(size_t)b * nf + c0 + b] = 1.0;
      F.upload(hF.data(), (size_t)bb * nf, st);
      hipLaunchKernelGGL((k_rotreg<T, double>), dim3((hcs[c].nvtx + 255) / 256, bb), dim3(256), 0, st, sm, F.p,
                         (long)nf, RRb.p, (long)nw, c);
      ctci_like(SrcField<T>{RRb.p, nw}, true, DstField<T>{VB.p, bsL}, t1.p, t2.p, bb, nullptr, 0, l1i, 0);
      vort2flux(VB.p, SB.p, nullptr, bb, gate.p, t1.p, t2.p);
      hipLaunchKernelGGL(k_regT_curl<T>, dim3((nf + 255) / 256, bb), dim3(256), 0, st, geo, sm, SB.p,
                         Lmax > 1 ? SB.p + nw : nullptr, bsL, (const double*)nullptr, acol + (long)c0 * nf,
                         (long)nf, gate.p, c, (const T*)nullptr);
      HIPCHK(hipStreamSynchronize(st));  // hF is reused
    }
  }
};

}  // namespace ibfs
