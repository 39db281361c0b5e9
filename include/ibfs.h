/*
 * ibfs.h -- C-ABI of the MI355X (gfx950) IBFS time-step library.
 *
 * The reference has no FFI: its seam is the Fortran call
 *     CALL advance(it)                       (src/main.f90:47-50)
 * with every input and output held in module globals (src/variables.f90:13-23,
 * src/grid.f90:12-28, src/myfft.f90:12-17).  This library owns those arrays on
 * the device; a Fortran host binds these entry points with ISO_C_BINDING and
 * keeps the namelist input, the advance loop and the output files.
 *
 * Conventions
 *   - Every function returns 0 on success and a non-zero IBFS_E* code on
 *     failure; ibfs_last_error() gives the message (the Fortran shim turns it
 *     into STOP, as the reference's own failures are STOPs).
 *   - All arrays are host memory in the reference's layouts (Fortran
 *     column-major, 1-based in Fortran):
 *       omega, rhs_old : (m-1)*(n-1)*L, i fastest      (variables.f90:46)
 *       q, q0          : Nq*L, u faces then v faces     (grid.f90:119-134)
 *       fb, motion     : Nf = 2*Nb, x comps then y comps (grid.f90:105-116)
 *       cholmat        : Nf*Nf column-major, cholvec: Nf (variables.f90:22-23)
 *     They are copied, never retained past the call.
 *   - One handle owns one HIP stream and its device memory.  Calls on one
 *     handle are not thread-safe (the reference is single-threaded too).
 */
#ifndef IBFS_H
#define IBFS_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ibfs_handle ibfs_handle;

enum {
  IBFS_OK = 0,
  IBFS_EINVAL = 1,  /* bad argument or unsupported size */
  IBFS_EHIP = 2,    /* a HIP runtime call failed */
  IBFS_ECHOL = 3,   /* Cholesky factorisation failed (matrix not SPD) */
  IBFS_ESTATE = 4   /* call out of order (e.g. a step before the factor exists) */
};

/* Precision modes. */
enum {
  IBFS_PLAIN_FP64 = 0, /* FP64 throughout: the drop-in */
  IBFS_MIXED = 1       /* FP32 transforms/nonlinear/smearing; FP64 force solve,
                          reductions and one refinement sweep */
};

/* Creates the solver for one case.  Replaces the set-up chain
 * input -> setup -> setup_variables (src/main.f90:42-44):
 *   grid spacing delta = len/m and the %4 check     (src/grid.f90:64-69)
 *   level clamp mgridlev <= nlevmax                 (src/grid.f90:79-82)
 *   index maps and smearing tables (compact form)   (src/grid.f90:99-137,142-230)
 *   eigenvalue / AB2 / CN tables                    (src/myfft.f90:21-70)
 *   zero-flow or uniform-flow initial condition     (src/variables.f90:54-67)
 * xb, yb, bdyno: the Nb immersed-boundary points as read from geom.inp
 * (src/grid.f90:270-276).  bdyno may be NULL (all 0). */
int ibfs_create(int m, int n, int mgridlev, int nb, double dt, double re, double len, double offsetx,
                double offsety, const double* xb, const double* yb, const int* bdyno, int unif_flow,
                int precision, ibfs_handle** out);

/* dims[0..11] = m, n, L (after the clamp), Nb, Nf, Nq, W=(m-1)(n-1),
 * nbc=2(m+1)+2(n+1), number of bodies, min bdyno, max bdyno, Nsmear. */
int ibfs_dims(const ibfs_handle* h, int* dims);

/* Replaces read_variables / write_variables' state exchange
 * (src/variables.f90:141-175).  NULL pointers are skipped. */
int ibfs_set_state(ibfs_handle* h, const double* omega, const double* rhs_old, const double* q, const double* q0,
                   const double* fb);
int ibfs_get_state(ibfs_handle* h, double* omega, double* rhs_old, double* q, double* q0, double* fb);

/* Replaces read_cholesky (src/variables.f90:124-137): loads a factor in the
 * ibfs.chd layout (strict lower triangle = L, diagonal in cholvec). */
int ibfs_set_cholesky(ibfs_handle* h, const double* cholmat, const double* cholvec);

/* Replaces preprocess (src/operators.f90:29-59): builds
 * A = E C (C^T C)^-1_multidomain ainv C^T E^T column by column on the device
 * (a_times, src/operators.f90:444-462, batched over columns) and factors it
 * as choldc does (src/operators.f90:766-791).  If cholmat/cholvec are not
 * NULL the factor is returned in the ibfs.chd layout for the host to cache. */
int ibfs_build_cholesky(ibfs_handle* h, double* cholmat, double* cholvec);

/* One time step: replaces advance (src/operators.f90:62-188) for stationary
 * bodies.  *itime is incremented on entry as in the reference (line 77); at
 * *itime == 1 the explicit Euler start is taken (lines 115-117).  motion
 * (Nf, may be NULL = 0) is what move_body/actuation would return, already
 * scaled by delta (src/grid.f90:573-574); the library subtracts regT(q0)
 * (line 155).  A non-finite motion entry is rejected (IBFS_EINVAL) before
 * any step runs.  Outputs (each may be NULL):
 *   fb    : Nf forces, the reference's fb
 *   force : [fx(bmin..bmax), fy(bmin..bmax)] as written by write_force (line 184)
 *   circ  : L circulations as written by write_circ (line 185)
 *   slip  : max |E q - motion| as written by write_slip (line 186) */
int ibfs_step(ibfs_handle* h, int* itime, const double* motion, double* fb, double* force, double* circ,
              double* slip);

/* The force solve of advance (src/operators.f90:166-171; namelist
 * stationary, cgtol, cg_max_iter of src/parameters.f90:19-20,31).
 * stationary != 0 (the default): fb = cholsl(rhsf) with the factor of
 * ibfs_build_cholesky / ibfs_set_cholesky.  stationary == 0: cjgr
 * (src/operators.f90:393-441), conjugate gradients on the device a_times,
 * warm-started from the previous fb, stopping at iter == cg_max_iter or
 * r.r <= cgtol^2; no factor is needed.  The CG form runs one case per
 * handle in the plain FP64 form, with eager launches (the iteration count is
 * data-dependent). */
int ibfs_set_solver(ibfs_handle* h, int stationary, double cgtol, int cg_max_iter);

/* New positions of the body points (one-case handles): replaces the
 * setup_reg call that follows move_body when a body moves
 * (src/operators.f90:142-144; tables of src/grid.f90:142-230).  The point
 * velocities enter through the motion argument of the next step
 * (move_body's result, scaled by delta).  The factor of the old geometry is
 * dropped, so a moving body is stepped with ibfs_set_solver(h, 0, ...) as in
 * the reference (stationary = F), or with a factor built anew. */
int ibfs_move_body(ibfs_handle* h, const double* xb, const double* yb);

/* Actuator slots: replaces actuation and u_jet (src/grid.f90:579-633),
 * which advance calls every step when actuators exist
 * (src/operators.f90:150-152).  The n_act slots own the last
 * npt[0] + ... + npt[n_act-1] points of the case's point list, in slot order
 * (actuator_slot appends them with bdyno = -1, -2, ..., src/grid.f90:636-718;
 * actuation addresses them as nb - n_act_pt + count).  par holds 7 values per
 * slot, its namelist read_act_parameters in order: x_j, y_j, h_j, u_j,
 * u_j_osc, theta_j (degrees), omg_j.  Every step, on the device, each point
 * of slot a gets the motion
 *   (u_j + u_j_osc sin(omg_j dt itime)) (cos, sin)(pi/180 theta_j) * delta
 * in place of what the motion argument of ibfs_advance holds there.  The
 * reference re-reads the namelist files every step; the host reads them once
 * and passes them here.  n_act = 0 removes the actuators.  Handles with
 * actuators take no tangent or adjoint runs (their control drives every
 * point). */
int ibfs_set_actuators(ibfs_handle* h, int n_act, const int* npt, const double* par);
int ibfs_case_set_actuators(ibfs_handle* h, int c, int n_act, const int* npt, const double* par);

/* The CG iteration counts of the steps of the last ibfs_advance / ibfs_step
 * call (what cjgr writes to ibfs.inf, src/operators.f90:439); 0 for steps
 * solved with the factor.  iters receives nsteps values. */
int ibfs_cg_iterations(const ibfs_handle* h, int nsteps, int* iters);

/* nsteps consecutive calls of advance with a constant motion vector.  hist
 * (may be NULL) receives nsteps rows of [fx.., fy.., circ(1..L), slip]
 * (2*nbody + L + 1 values per row), i.e. the lines of force.dat, circ.dat
 * and slip.dat.  Steps after the first are replayed from one captured HIP
 * graph. */
int ibfs_advance(ibfs_handle* h, int* itime, int nsteps, const double* motion, double* hist);

/* ---------------------------------------------------------------------------
 * Ensembles (no counterpart in the reference, which runs one case per
 * process): ncases independent stationary-body cases advanced in lockstep by
 * one handle.  Cases may differ in grid size, levels, Re, dt, domain and
 * body; cases of equal m x n are batched into the same kernel launches, and
 * one HIP graph replays the whole ensemble step.  Per-case arrays are
 * indexed by case; xb, yb, bdyno are the cases' point lists concatenated
 * (nb[c] entries each).  The motion vector of ibfs_advance is the cases'
 * Nf vectors concatenated; a history row holds every case's row
 * [fx.., fy.., circ(1..L), slip] at the column given by col0[c] of
 * ibfs_ensemble_info.  The single-case calls above address case 0.
 * ------------------------------------------------------------------------- */
int ibfs_ensemble_create(int ncases, const int* m, const int* n, const int* mgridlev, const int* nb,
                         const double* dt, const double* re, const double* len, const double* offsetx,
                         const double* offsety, const double* xb, const double* yb, const int* bdyno, int unif_flow,
                         int precision, ibfs_handle** out);
int ibfs_ensemble_info(const ibfs_handle* h, int* ncases, int* rowlen, int* col0);
int ibfs_case_dims(const ibfs_handle* h, int c, int* dims);
int ibfs_case_set_state(ibfs_handle* h, int c, const double* omega, const double* rhs_old, const double* q,
                        const double* q0, const double* fb);
int ibfs_case_get_state(ibfs_handle* h, int c, double* omega, double* rhs_old, double* q, double* q0, double* fb);
int ibfs_case_set_cholesky(ibfs_handle* h, int c, const double* cholmat, const double* cholvec);
int ibfs_case_build_cholesky(ibfs_handle* h, int c, double* cholmat, double* cholvec);
int ibfs_ensemble_build_cholesky(ibfs_handle* h);

/* ---------------------------------------------------------------------------
 * Sensitivities (no counterpart in the reference).  Objective of every case:
 *   J = (1/nwin) * sum over steps itime+nspin+1 .. itime+nspin+nwin of
 *       (cd * sum over bodies of the x force + cl * sum of the y force)
 * with the forces of write_force (src/operators.f90:583-636), after
 * nspin + nwin steps from the handle's current state.  The control is the
 * prescribed surface velocity u_B at every IB point (Nf values per case,
 * x components then y components, concatenated over cases), active in every
 * step as motion = delta * u_B (src/grid.f90:573-574).  All three advance
 * the handle's state and *itime by nspin + nwin steps.  J, dJ: one value per
 * case.  Tangent and adjoint runs need a handle whose cases share one grid
 * size; on a mixed handle they run with FP32 fields and the FP64 force system
 * (the FP64 handle is the reference for their gradients).
 * ------------------------------------------------------------------------- */
/* forward only: J */
int ibfs_objective(ibfs_handle* h, int* itime, int nspin, int nwin, const double* ub, double cd, double cl,
                   double* J);
/* FP64 tangent-linear model along du_B: J and dJ = dJ/du_B . du_B */
int ibfs_tangent_linear(ibfs_handle* h, int* itime, int nspin, int nwin, const double* ub, const double* dub,
                        double cd, double cl, double* J, double* dJ);
/* discrete adjoint: J and grad = dJ/du_B (concatenated per case), with the
 * forward states checkpointed every ckpt steps and recomputed per segment */
int ibfs_adjoint_gradient(ibfs_handle* h, int* itime, int nspin, int nwin, int ckpt, const double* ub, double cd,
                          double cl, double* J, double* grad);
/* The same adjoint for any functional of the force and circulation history
 * (no counterpart in the reference, which has no adjoint; for callers that
 * differentiate their own functional of the rows of write_force and
 * write_circ, src/operators.f90:583-659, 662-699).
 * ibfs_adjoint_forward: nsteps steps from the current state with
 * motion = delta * u_B, the state checkpointed every ckpt steps in the
 * handle's tape; hist (nsteps x rowlen, may be NULL) gets the history rows
 * [fx of each body, fy of each body, circulation of each level, slip]
 * (concatenated over cases).  ibfs_adjoint_reverse: the reverse sweep of
 * the last forward sweep with ghist = dJ/d(hist) (the slip entries are not
 * differentiated and must be 0); grad = dJ/du_B (Nf values per case); the
 * state stays at the forward sweep's end.  IBFS_ESTATE without a forward
 * sweep. */
int ibfs_adjoint_forward(ibfs_handle* h, int* itime, int nsteps, int ckpt, const double* ub, double* hist);
int ibfs_adjoint_reverse(ibfs_handle* h, const double* ghist, double* grad);
/* Vector-Jacobian product of one step (advance, src/operators.f90:62-188)
 * of a one-case handle, from the state omega
 * (Lmax nw), q (Lmax nq) at the start of step itime (itime == 1 takes the
 * Euler start), with the handle's Cholesky factor: the cotangents of the
 * step's outputs (a_omega, a_rhs_old: Lmax nw; a_q: Lmax nq; a_row: one
 * history row; NULL means 0) give those of its inputs g_omega, g_rhs_old,
 * g_q and of its motion argument g_motion (Nf).  Outputs may be NULL.  The
 * handle's state is not touched.  No counterpart in the reference: for
 * callers that run their own reverse sweep. */
int ibfs_step_vjp(ibfs_handle* h, int itime, const double* omega, const double* q, const double* a_omega,
                  const double* a_rhs_old, const double* a_q, const double* a_row, double* g_omega,
                  double* g_rhs_old, double* g_q, double* g_motion);

/* Operator entry points, one per reference module procedure; used by the
 * parity tests.  Each copies its inputs in, runs the device kernels on the
 * handle's stream, and copies the result out; none touches the state. */
int ibfs_op_dst(ibfs_handle* h, const double* in, double* out);                 /* src/myfft.f90:83-94   */
int ibfs_op_ctci(ibfs_handle* h, const double* in, double* out);                /* src/myfft.f90:97-109  */
int ibfs_op_ainv(ibfs_handle* h, const double* in, double* out);                /* src/myfft.f90:112-124 */
int ibfs_op_curl(ibfs_handle* h, const double* x, const double* coarse, double* vel);
/* curl (src/operators.f90:243-288) with xbc = get_bc(coarse, 1.0); coarse == NULL means xbc = 0 */
int ibfs_op_rot(ibfs_handle* h, const double* x, double* out);                  /* src/operators.f90:291-304 */
int ibfs_op_coarsify(ibfs_handle* h, const double* crhs, const double* rhs, double* out); /* :307-333 */
int ibfs_op_get_bc(ibfs_handle* h, const double* r, double fac, double* rbc);   /* src/operators.f90:336-367 */
int ibfs_op_rot_nonlinear(ibfs_handle* h, const double* omega, const double* q, const double* q0, const double* bc,
                          double* out); /* rot(nonlinear(...)), src/operators.f90:465-510 + 291-304 */
int ibfs_op_vort2flux(ibfs_handle* h, double* omega, double* vel, int nlev);     /* src/operators.f90:191-240 */
int ibfs_op_reg(ibfs_handle* h, const double* h0, double* hq);                  /* src/operators.f90:513-528 */
int ibfs_op_regT(ibfs_handle* h, const double* hq, double* h0);                 /* src/operators.f90:531-546 */
int ibfs_op_a_times(ibfs_handle* h, const double* x, double* out);              /* src/operators.f90:444-462 */
int ibfs_op_cholsl(ibfs_handle* h, const double* b, double* x);                 /* src/operators.f90:794-819 */

/* Device time of the last ibfs_advance call in milliseconds (HIP events on
 * the handle's stream), for measurement. */
int ibfs_last_advance_ms(const ibfs_handle* h, double* ms);

/* Tile shapes (no counterpart in the reference): the FFTs per workgroup that
 * the autotuner chose for group `group` of the handle (groups = cases of one
 * grid size), for the row pass, column pass, fused column pair, step-1 row
 * pass and step-1 column pair (0: the generic pass of a size without a
 * specialised kernel; 8: the register-blocked kernel of FFT lengths 256,
 * 512 and 1024, one wave per workgroup).  *ngroups (may be NULL) receives
 * the group count.  The tuner times every fitting shape on the group's
 * batch when the handle is created; IBFS_TILE=P in the environment forces
 * the largest fitting shape <= P instead.  Results do not depend on the
 * tile shape. */
int ibfs_tile_shapes(const ibfs_handle* h, int group, int* ngroups, int* shapes);

/* The same five shapes for the launches that transform every level of the
 * group's cases at once (one or two cases with all levels), tuned on that
 * batch; 0 for a kind without such a launch. */
int ibfs_level_tile_shapes(const ibfs_handle* h, int group, int* shapes);

/* Measurement only (no counterpart in the reference): launches one 1-D
 * DST-I pass over a W field per case of the handle's first group (one field
 * for a single-case handle, B for a batch of B equal-size cases) reps times
 * on the handle's stream and returns the mean launch duration from HIP
 * events.  axis 0 = along i (rows), 1 = along j (columns), 2 = the fused
 * column pair of ctci (DST-I along j, times laminv, DST-I along j: one read
 * and one write of the field). */
int ibfs_time_dst_pass(ibfs_handle* h, int axis, int reps, double* ms_per_launch);

/* Measurement only (no counterpart in the reference): the dense form of a
 * DST-I pass on the matrix cores, Y = X S_m (axis 0, along i) or Y = S_n X
 * (axis 1, along j) with S the (N-1)x(N-1) DST-I matrix, one 16x16 output
 * tile per workgroup and ks (1, 2, 4, 8) wave64s splitting the inner
 * dimension (v_mfma_f64_16x16x4 on an FP64 handle, v_mfma_f32_16x16x4 on a
 * mixed one).  Timed like ibfs_time_dst_pass (which follows the handle's
 * precision too), so the two forms of the pass compare directly. */
int ibfs_time_dst_gemm(ibfs_handle* h, int axis, int ks, int reps, double* ms_per_launch);

/* Test entry (no counterpart in the reference): the 2-D DST-I of
 * src/myfft.f90:83-94 computed with the two dense passes above (FP64). */
int ibfs_op_dst_gemm(ibfs_handle* h, int ks, const double* in, double* out);

/* Measurement only (no counterpart in the reference): nsteps further steps
 * of an FP64 one-group handle with the factor (the state advances; zero
 * motion; the step counter is not updated), launched eagerly with a start
 * and a stop event attached to every dispatch of the fused ctci column pair,
 * the step's dominant kernel.  us_mean receives the mean duration of those
 * dispatches inside the step (microseconds), launches their number and
 * fields (may be NULL) the number of W fields they transform in total (one
 * per level and case of each launch). */
int ibfs_time_step_kernel(ibfs_handle* h, int nsteps, double* us_mean, int* launches, int* fields);

/* Measurement only (no counterpart in the reference): a plain streaming copy
 * of n doubles, launched reps times; the reference point that calibrates
 * HBM byte counters (FETCH_SIZE / WRITE_SIZE) against known traffic. */
int ibfs_time_copy(ibfs_handle* h, long n, int reps, double* ms_per_launch);

/* No counterpart in the reference: an entry point that allocates nothing,
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It is
 * ibfs_op_ctci (the Poisson solve C^T C s = omega with zero edge values,
 * src/myfft.f90:97-109) on device pointers the caller owns: omega and s
 * hold (m-1)*(n-1) doubles each, i fastest (they may be the same buffer);
 * stream is a hipStream_t (NULL: the default stream), and the call only
 * enqueues work on it.  No handle is needed.  Intermediates, the eigenvalue
 * table and the twiddles live in one library-owned device array of about
 * 25 MB per GPU (held until the process ends), so m and n are limited to the
 * size-specialised transforms with m, n <= 1024: 128, 200, 256, 300, 320,
 * 384, 512, 600, 1024; other sizes return IBFS_EINVAL.  Calls that overlap
 * on one GPU share that array and race: serialise them (one stream, or
 * events).  Errors: ibfs_last_error(NULL) gives the calling thread's last
 * message. */
int ibfs_ctci_noalloc(int m, int n, const double* omega, double* s, void* stream);

/* The message of the last failed call on h; with h == NULL, of the calling
 * thread's last failed handle-free call. */
const char* ibfs_last_error(const ibfs_handle* h);
int ibfs_destroy(ibfs_handle* h);

#ifdef __cplusplus
}
#endif

#endif /* IBFS_H */
