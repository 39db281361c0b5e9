// ibfs_adjoint_host.hpp -- launch schedule of the discrete adjoint step and
// of its checkpointed reverse sweep, on a group of cases: field adjoints in
// the group's element type T, force-system adjoints in FP64.
//
// Objective: J = (1/nwin) sum over the last nwin of nspin+nwin steps of
// (cd * sum of body x forces + cl * sum of body y forces), i.e. the mean
// drag/lift of write_force (src/operators.f90:583-636).  Control: the
// prescribed surface velocity u_B at each IB point, entering every step as
// motion = delta * u_B (src/grid.f90:573-574, src/operators.f90:143-155).
#pragma once
#include "ibfs_adjoint.hpp"
#include "ibfs_engine_t.hpp"

namespace ibfs {

// Writes v to t and adds scale * v to a (the multi-level reverse chain needs
// each level's solution twice: for the level itself and for its edge).
template <typename T>
struct DstAddCopy {
  T* a; long bs; T scale; T* t;
  __device__ void operator()(const Geo& g, int b, int ii, int jj, T v) const {
    const int e = ii + jj * g.mm1;
    a[b * bs + e] += scale * v;
    t[(long)b * g.nw + e] = v;
  }
};

template <typename T>
__global__ void k_zero_centre(Geo g, T* __restrict__ w, long bs, const int* __restrict__ Lc, int lev) {
  const int wi = g.m / 2 - 1, wj = g.n / 2 - 1;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (t >= wi * wj || !active(Lc, b, lev)) return;
  const int i = g.m / 4 + 2 + t % wi, j = g.n / 4 + 2 + t / wi;
  w[b * bs + wix(g, i, j)] = T(0);
}

template <typename T>
struct AdjointRunner {
  Engine<T>& E;
  hipStream_t st;
  int B, Lmax, nw, nq, nbc;
  long FB, QB, BB;
  DevBuf<T> AW, AWO, ARO, AQ, LB, AN, AX, TK, ABC, ALB, ANEW, URR, U, T1, T2;
  DevBuf<double> AFB, Z, Wv, Y, AU;

  explicit AdjointRunner(Engine<T>& e) : E(e), st(e.st) {
    B = E.B; Lmax = E.Lmax; nw = E.nw; nq = E.nq; nbc = E.nbc;
    FB = (long)Lmax * nw; QB = (long)Lmax * nq; BB = (long)Lmax * nbc;
    const size_t W = (size_t)B * FB, Q = (size_t)B * QB, C = (size_t)B * BB, F = (size_t)B * nw;
    AW.alloc(W); AWO.alloc(W); ARO.alloc(W); AQ.alloc(Q); LB.alloc(C); ALB.alloc(C); ANEW.alloc(C);
    AN.alloc(F); AX.alloc(F); TK.alloc(F); URR.alloc(F); U.alloc(F); T1.alloc(F); T2.alloc(F);
    ABC.alloc((size_t)B * nbc);
    const size_t nf = std::max(E.nf_tot, 1);
    AFB.alloc(nf); Z.alloc(nf); Wv.alloc(nf); Y.alloc(nf); AU.alloc(nf);
    URR.zero(st);
    reset();
  }

  void reset() {
    AW.zero(st); ARO.zero(st); AQ.zero(st); AU.zero(st);
  }

  dim3 grid(long n) const { return dim3((unsigned)((n + 255) / 256), B); }

  // Reverse of one multi-domain Poisson solve (vort2flux) whose outputs were
  // the fluxes q_k = curl(s_k, get_bc(s_{k+1})) on every level (aq != null,
  // level stride QB) or only the level-1 vertex source urr (the step-2 solve,
  // whose fluxes feed regT); then the reverse of its in-place coarsify
  // chain.  The adjoint of the vorticity accumulates into AW with sign.
  void mp_reverse(const T* aq, const T* urr, T sign) {
    const int* Lc = E.Lc.p;
    for (int k = 0; k < Lmax; k++) {
      DstAddCopy<T> dst{AW.p + (long)k * nw, FB, sign, TK.p};
      if (urr && k == 0) {
        E.ctci_like(SrcField<T>{urr, nw}, false, dst, T1.p, T2.p, B, Lc, k);
      } else {
        SrcAdjS<T> src{aq ? aq + (long)k * nq : nullptr, QB, k > 0 ? ABC.p : nullptr, (long)nbc};
        E.ctci_like(src, false, dst, T1.p, T2.p, B, Lc, k);
      }
      if (k + 1 < Lmax)
        hipLaunchKernelGGL(k_bc_adj<T>, grid(nbc), dim3(256), 0, st, E.geo, aq ? aq + (long)k * nq : nullptr, QB,
                           TK.p, (long)nw, T(1), ABC.p, (long)nbc, Lc, k);
    }
    const int wcnt = (E.m / 2 - 1) * (E.n / 2 - 1);
    for (int k = Lmax - 1; k >= 1; k--) {
      hipLaunchKernelGGL(k_coarsify_adj<T>, grid((long)nw), dim3(256), 0, st, E.geo, AW.p + (long)k * nw,
                         AW.p + (long)(k - 1) * nw, FB, Lc, k);
      hipLaunchKernelGGL(k_zero_centre<T>, grid(wcnt), dim3(256), 0, st, E.geo, AW.p + (long)k * nw, FB, Lc, k);
    }
  }

  // One adjoint step.  In: AW, ARO, AQ = adjoints of the step's outputs
  // (omega, rhs_old, q); grow = the cotangent of the step's diagnostics row
  // on the device (history-row layout, the cases at their row offsets: body
  // x and y forces and circulations; the slip entry is not differentiated),
  // or nullptr; wold/qold = the forward state at the start of the step.
  // Out: AW, ARO, AQ = adjoints of the step's inputs; Z = the adjoint of the
  // step's motion argument; AU += delta * Z (dJ/du_B).
  void step(const T* wold, const T* qold, int first, const double* grow) {
    const int* Lc = E.Lc.p;
    const Geo& g = E.geo;
    hipLaunchKernelGGL(k_lastbc<T>, dim3((nbc * Lmax + 255) / 256, B), dim3(256), 0, st, g, wold, LB.p, Lc);
    if (grow)  // circulations are sums of the step's final vorticity
      hipLaunchKernelGGL(k_circ_adj<T>, dim3((nw + 255) / 256, B * Lmax), dim3(256), 0, st, g, E.sm, grow, AW.p,
                         FB, Lc);
    // STEP 4 reverse: fluxes and vorticity of the final vort2flux
    mp_reverse(AQ.p, nullptr, T(1));
    // STEP 3 reverse: afb_total = afb_J - regT(curl(ainv(aw_1), 0)); z = A^-1 (regT(..) - afb_J) = -adrhsf
    E.ctci_like(SrcField<T>{AW.p, FB}, true, DstField<T>{U.p, nw}, T1.p, T2.p, B, nullptr, 0, E.lam1i.p, E.tbs());
    if (grow)
      hipLaunchKernelGGL(k_force_adj_row, grid(E.nf_max), dim3(256), 0, st, E.sm, E.wsum.p, E.bdyno.p, grow,
                         AFB.p);
    else
      AFB.zero(st);
    hipLaunchKernelGGL(k_regT_curl<T>, grid(E.nf_max), dim3(256), 0, st, g, E.sm, U.p, (const T*)nullptr,
                       (long)nw, AFB.p, Wv.p, 0L, Lc, -1, (const T*)nullptr);
    const dim3 g4((E.nf_max + 3) / 4, B);
    hipLaunchKernelGGL(k_trmv_lower, g4, dim3(256), 0, st, E.sm, E.Linv.p, Wv.p, Y.p);
    hipLaunchKernelGGL(k_trmv_upper_t, g4, dim3(256), 0, st, E.sm, E.Linv.p, Y.p, Z.p, 0);
    // dJ/du_B: rhsf = E q - (delta u_B - E q0), so au -= delta * adrhsf = delta * z
    hipLaunchKernelGGL(k_axpy_case, grid(E.nf_max), dim3(256), 0, st, E.sm, Z.p, 1.0, AU.p, 1);
    // STEP 2 reverse: aq*_1 = reg(adrhsf) = -reg(z); its level-1 vertex source is -rot(reg(z))
    hipLaunchKernelGGL((k_rotreg<T, double>), dim3((E.nvtx_max + 255) / 256, B), dim3(256), 0, st, E.sm, Z.p,
                       0L, URR.p, (long)nw, -1);
    mp_reverse(nullptr, URR.p, T(-1));
    // STEP 1 reverse, level by level (the forward ran coarse to fine)
    AWO.zero(st);
    for (int k = 0; k < Lmax; k++) {
      // aX = S lam1i S aw*, then aw_old += aw* + 2 vfac D0 aX (k_step1_adj_incr)
      E.row_pass(SrcField<T>{AW.p + (long)k * nw, FB}, DstField<T>{T1.p, nw}, B, Lc, k);
      E.launch_col2(Col2Op<T, SrcField<T>, MidTab<T>, DstField<T>>{g, SrcField<T>{T1.p, nw},
                                                                   MidTab<T>{E.lam1i.p + (long)k * E.nwp, E.tbs()},
                                                                   DstField<T>{T2.p, nw}, Lc, k},
                    g.mm1, 2, B, Engine<T>::kStep1Col2);
      E.row_pass(SrcField<T>{T2.p, nw}, DstField<T>{AX.p, nw}, B, Lc, k);
      hipLaunchKernelGGL(k_step1_adj_incr<T>, grid(nw), dim3(256), 0, st, g, AW.p + (long)k * nw, FB, AX.p,
                         AWO.p + (long)k * nw, E.vfac.p, Lc, k);
      hipLaunchKernelGGL(k_step1_adj_combine<T>, grid(nw), dim3(256), 0, st, g, AX.p, ARO.p + (long)k * nw, FB,
                         AN.p, E.con1.p, E.con2.p, Lc, k, first);
      hipLaunchKernelGGL(k_nl_adj_vertex<T>, grid(nw), dim3(256), 0, st, g, AN.p, qold + (long)k * nq,
                         E.q0.p + (long)k * nq, QB, AWO.p + (long)k * nw, FB, Lc, k);
      hipLaunchKernelGGL(k_nl_adj_face<T>, grid(nq), dim3(256), 0, st, g, AN.p, wold + (long)k * nw, FB,
                         LB.p + (long)k * nbc, BB, AQ.p + (long)k * nq, QB, Lc, k);
      if (k + 1 < Lmax) {
        hipLaunchKernelGGL(k_nl_adj_bc<T>, grid(nbc), dim3(256), 0, st, g, AN.p, AX.p, qold + (long)k * nq,
                           E.q0.p + (long)k * nq, QB, E.vfac.p, ALB.p + (long)k * nbc, BB, ANEW.p + (long)k * nbc,
                           Lc, k);
        hipLaunchKernelGGL(k_gt_add<T>, grid(g.m + g.n + 4), dim3(256), 0, st, g, ANEW.p + (long)k * nbc, BB, T(0.25),
                           AW.p + (long)(k + 1) * nw, FB, Lc, k + 1);
      }
    }
    for (int k = 0; k + 1 < Lmax; k++)
      hipLaunchKernelGGL(k_gt_add<T>, grid(g.m + g.n + 4), dim3(256), 0, st, g, ALB.p + (long)k * nbc, BB, T(0.25),
                         AWO.p + (long)(k + 1) * nw, FB, Lc, k + 1);
    std::swap(AW.p, AWO.p);
  }
};

}  // namespace ibfs
