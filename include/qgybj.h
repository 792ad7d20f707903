/*
 * qgybj.h -- C ABI of the MI355X QG-YBJ+ time step.
 *
 * Drop-in for the body of the reference's integration loop
 * (oliasselin/QG-YBJp-model main_waqg.f90:326-467, with its Euler prologue
 * :197-317).  The reference has no library interface: its loop body is a
 * sequence of Fortran calls over static arrays.  Each entry point below names
 * the reference routine(s) it replaces.  A Fortran driver binds these
 * through ISO_C_BINDING (qg-ybjp-model_amd/fortran/qgybj_c.f90).
 *
 * Conventions
 *  - Spectral arrays are host arrays of interleaved (re, im) FP64 pairs in the
 *    reference's column-major layout X(iktx, ikty, n3), iktx = n1/2+1,
 *    ikty = n2, ky order [0..n2/2-1, Nyquist, -n2/2+1..-1] (init.f90:35-41);
 *    no vertical halo levels.
 *  - Real (physical) arrays are X(ldx, n2, nz) with ldx >= n1 (the reference
 *    passes ldx = n1+2, main_waqg.f90:15-110).
 *  - Every call returns QG_OK (0) or an error code; the library never exits.
 *    qg_last_error() gives the message of the last failure on a context.
 *  - One context = one member, or a batch of `members` members with the same
 *    grid and parameters (qg_create_batch), on one HIP stream of the current
 *    device.  For a batch every array argument carries a trailing member
 *    dimension: X(iktx, ikty, n3, members), X(ldx, n2, nz) with nz counting
 *    the levels of all members (member-major).  Calls on one context are
 *    serialised by the caller; distinct contexts are independent.
 */
#ifndef QGYBJ_H
#define QGYBJ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  QG_OK = 0,
  QG_ERR_ARG = 1,       /* invalid argument or unsupported size */
  QG_ERR_HIP = 2,       /* HIP runtime error */
  QG_ERR_NOMEM = 3,     /* device allocation failed */
  QG_ERR_NONFINITE = 4, /* a NaN/Inf appeared in the updated fields */
  QG_ERR_STATE = 5      /* call not valid in the current state */
};

/* special-case switches; same meaning as the reference's integer parameters
 * fixed_flow, no_feedback, passive_scalar, linear, no_dispersion, inviscid
 * (parameters_test1.f90:22-29).  QG_YBJ_REGULAR is ybj_plus = 0: the regular
 * YBJ operator (sumB, compute_sigma, compute_A, derivatives.f90:1037-1231;
 * main_waqg.f90:298-310, 449-461) instead of the YBJ+ inversion
 * A_solver_ybj_plus; without it ybj_plus = 1.  Other bits are rejected. */
enum {
  QG_FIXED_FLOW = 1,
  QG_NO_FEEDBACK = 2,
  QG_PASSIVE_SCALAR = 4,
  QG_LINEAR = 8,
  QG_NO_DISPERSION = 16,
  QG_INVISCID = 32,
  QG_YBJ_REGULAR = 64
};

/* Runtime form of the reference's compile-time parameters module
 * (parameters_test1.f90:1-230).  All values are the exact FP64 values the
 * step uses; the driver forms them (including the REAL(4) literals and powers
 * of the reference) and passes their bits. */
typedef struct qg_params {
  int32_t n1, n2, n3;   /* n1 == n2, a power of two in [16, 512]; n3 >= 2   */
  int32_t flags;        /* QG_* switches                                     */
  double dt;            /* delt                                              */
  double gamma;         /* Robert-Asselin coefficient                        */
  double Ro, Bu, W2F;   /* Rossby, (inverse) Burger, (Uw/U)^2                */
  double nuz;           /* vertical viscosity of q (derivatives.f90:1877)    */
  const double* ifac_q; /* (iktx, ikty): int_factor   (main_waqg.f90:371)    */
  const double* ifac_w; /* (iktx, ikty): int_factor_w (main_waqg.f90:372)    */
  const double* N2_u;   /* n3: N^2 at z_u(k) = k dz       (init.f90:232)     */
  const double* N2_s;   /* n3: N^2 at z_s(k) = (k-1/2) dz (init.f90:233)     */
} qg_params;

typedef struct qg_ctx qg_ctx;

/* field ids for qg_device_field */
enum { QG_PSI = 0, QG_Q, QG_QO, QG_BR, QG_BI, QG_BRO, QG_BIO, QG_AR, QG_AI, QG_NFIELDS };

/* Set-up and tear-down.  Replaces initialize_fftw (fft.f90:32-52),
 * init_arrays (init.f90:17-171) and the coefficient part of init_base_state
 * (init.f90:177-264): allocates the member's device state, twiddles, masks,
 * integrating-factor exponentials and vertical operator. */
int qg_create(const qg_params* p, qg_ctx** out);
/* A batch of `members` independent members with the same qg_params (their
 * states differ): every launch of the step covers all members, which fills
 * the GPU when one member alone cannot.  qg_create(p, out) is
 * qg_create_batch(p, 1, out).  No counterpart in the reference (one member
 * per run). */
int qg_create_batch(const qg_params* p, int32_t members, qg_ctx** out);
int qg_members(const qg_ctx* ctx, int32_t* members);
int qg_destroy(qg_ctx* ctx);
const char* qg_last_error(const qg_ctx* ctx);

/* Load a spectral state (host arrays, see conventions).  NULL arrays load
 * zeros.  iter = 0 makes the next step the Euler first step
 * (main_waqg.f90:197-317); iter > 0 continues with leapfrog steps.
 * Replaces the reference's start-up assignments main_waqg.f90:125-151. */
int qg_set_state(qg_ctx* ctx, const double* psik, const double* qk, const double* qok, const double* BRk,
                 const double* BIk, const double* BRok, const double* BIok, const double* ARk, const double* AIk,
                 int32_t iter);

/* Start from physical initial conditions: un-dealiased r2c of psi, BR, BI
 * (ncread_psi IO_ncf.f90:276-287, ncread_la :353-366), q = init_q(psi) or 0
 * under QG_FIXED_FLOW (main_waqg.f90:142-146), A = 0, old fields 0, iter = 0. */
int qg_set_state_physical(qg_ctx* ctx, const double* psir, const double* BRr, const double* BIr, int32_t ldx);

/* Copy the state out (NULL arrays skipped).  C = A_z as A_solver_ybj_plus
 * forms it (elliptic.f90:437-450).  Synchronises the context's stream. */
int qg_get_state(qg_ctx* ctx, double* psik, double* qk, double* qok, double* BRk, double* BIk, double* BRok,
                 double* BIok, double* ARk, double* AIk, double* CRk, double* CIk);

/* Advance nsteps time steps: the loop body main_waqg.f90:326-467 (Euler
 * prologue :197-317 when iter = 0): nonlinear terms (convol_waqg,
 * refraction_waqg), the fused update + Robert filter, compute_qw, psi_solver,
 * A_solver_ybj_plus, compute_velo.  Returns QG_ERR_NONFINITE if a NaN/Inf
 * appeared. */
int qg_step(qg_ctx* ctx, int32_t nsteps);
int qg_iter(const qg_ctx* ctx, int32_t* iter);

/* FP64 diagnostics of the current state (6 values per member, member-major),
 * nondimensional:
 * out[0] KE and out[1] PE (diag_zentrum ks / ps_quad, diagnostics.f90:127-158),
 * out[2] WKE, out[3] WPE, out[4] WCE (wave_energy k_p, p_p, c_p,
 * diagnostics.f90:683-712, without the Uw_scale^2 factor),
 * out[5] enstrophy 0.5<zeta^2> = sum kh2^2 |psi|^2; all averaged over z. */
int qg_diag(qg_ctx* ctx, double* out);

/* hspec_waves (diagnostics.f90:547-638): horizontal spectra of the waves at
 * nlev levels (1-based, 1 <= levels[i] <= n3, as the reference's height
 * array), from the current state in FP64 (the reference sums in REAL(4)).
 * For shells mode = floor(|k| + 1/2) = 0 .. n1/2 - 1 of the L = 1 modes:
 * out[((member * nlev + i) * (n1/2) + mode) * 3 + j] holds j = 0 the kinetic
 * part 1/2 sum |BR + i BI|^2, j = 1 the potential part
 * 1/2 sum 1/4 kh^2/Bu (|C|^2/N^2 at the interfaces below and above the level;
 * the one above only at level 1), j = 2 the number of modes (2 per mode);
 * nondimensional, as the reference writes them to h<i>w.dat. */
int qg_hspec_waves(qg_ctx* ctx, const int32_t* levels, int32_t nlev, double* out);

/* Omega equation (omega_eqn_rhs, derivatives.f90:1469-1586; omega_equation,
 * elliptic.f90:463-567): the QG vertical velocity w from the current psi, as
 * a host array shaped like a state field ((iktx, ikty, n3) per member).
 * Level k holds w at the unstaggered level z_u(k) = k dz: the interior
 * interfaces k = 1 .. n3-1 solve
 *   w_{k-1} - (2 + dz^2 kh2 N^2_u(k)/Bu) w_k + w_{k+1} = dz^2 rhs_k
 * (w = 0 below the bottom and at the top level k = n3) with
 * rhs = 2 J(psi_z, lap psi) formed by transforms; w = 0 at kh2 = 0 and
 * outside the mask.  Nondimensional, as the reference's wak. */
int qg_omega(qg_ctx* ctx, double* wk);

/* we_vave (diagnostics.f90:750-813): the vertically averaged wave energy
 * density (1/n3) sum_z 0.5 (BR^2 + BI^2) of the current state at every
 * physical point, from the c2r of BR and BI; out holds n1 * n2 values per
 * member, member-major, x fastest (the reference's WE_ave(ix, iy)), in FP64
 * (the reference sums in REAL(4)) and without the Uw_scale^2 factor. */
int qg_wave_vave(qg_ctx* ctx, double* out);

/* Keep the wave tendencies of every step (on != 0) or stop keeping them.
 * The reference keeps its tendency arrays nBRk, nBIk (convol_waqg,
 * derivatives.f90:607-803) and rBRk, rBIk (refraction_waqg :810-900) from
 * the last step, and its conversion diagnostics read them after the step
 * (main_waqg.f90:506-507).  The fused step never forms them separately, so
 * while this is on each step also writes F(zeta BR), F(zeta BI) planes and
 * the combined tendencies and splits them afterwards (four more fields of
 * device memory, allocated on the first call; the step costs more).  A new
 * state (qg_set_state*, qg_tune) discards kept tendencies. */
int qg_keep_tendencies(qg_ctx* ctx, int32_t on);

/* The last step's nBRk, nBIk, rBRk, rBIk as host arrays shaped like state
 * fields, masked by L; under QG_LINEAR nB = 0, under QG_PASSIVE rB = 0
 * (main_waqg.f90:338-358).  QG_ERR_STATE unless a step has run with
 * tendencies kept since the last state change.  NULL arrays are skipped. */
int qg_get_tendencies(qg_ctx* ctx, double* nBRk, double* nBIk, double* rBRk, double* rBIk);

/* Wave energy conversion terms from the kept tendencies of the last step and
 * the current A, B (as the reference evaluates them after the step,
 * main_waqg.f90:506-507), 5 values per member, member-major, in FP64 (the
 * reference sums in REAL(4)) and without the dimensional factors:
 * out[0], out[1]: we_conversion ca, cr (diagnostics.f90:967-1076), to be
 *   multiplied by Uw_scale^2;
 * out[2], out[3], out[4]: gamma_conversion ca, cr, cd (:1080-1256), to be
 *   multiplied by Uw_scale^2 U_scale / L_scale.
 * With D = lap A = -kh2 A and volume means <.> over the n1 n2 n3 points:
 * ca = <DR nBR + DI nBI> / (2 Bu), cr = <DI rBR - DR rBI> / (4 Bu),
 * gamma ca = -<DR nBR + DI nBI> / (2 Bu), gamma cr = -<DI rBR - DR rBI> / (4 Bu),
 * cd = <DR dR + DI dI> / Bu with d = -(1/2)(nuh1w (kx^2p1 + ky^2p1) +
 * nuh2w (kx^2p2 + ky^2p2)) B, the wave integrating-factor operator without
 * delt.  QG_ERR_STATE as for qg_get_tendencies. */
int qg_wave_conversion(qg_ctx* ctx, double* out);

/* init_q (init.f90:591-652) on host arrays: q = T(kh2) psi / dz^2 at L = 1. */
int qg_init_q(qg_ctx* ctx, const double* psik, double* qk);

/* fft_r2c / fft_c2r (fft.f90:55-83) on host arrays of nz <= n3 * members levels.
 * r2c is unmasked and scaled by 1/(n1 n2); c2r applies realit and is
 * unnormalised.  Unlike the reference, the c2r input is not modified. */
int qg_fft_r2c(qg_ctx* ctx, const double* r, int32_t ldx, double* k, int32_t nz);
int qg_fft_c2r(qg_ctx* ctx, const double* k, double* r, int32_t ldx, int32_t nz);

/* Device-resident access: device address of a state field (complex FP64,
 * (iktx, ikty, n3) with rows padded to qg_device_pitch complex elements so
 * that rows start on 128-byte lines) and the context's HIP stream
 * (hipStream_t). */
int qg_device_field(qg_ctx* ctx, int32_t which, void** dptr);
int qg_device_pitch(const qg_ctx* ctx, int32_t* pitch);
int qg_stream(qg_ctx* ctx, void** stream);
int qg_sync(qg_ctx* ctx);

/* Per-kernel timing with HIP events on the stream each kernel runs on.
 * mask: bit k times kernel id k (0 col_inv, 1 row_nl, 2 col_upd, 3 row_qw,
 * 4 col_qw, 5 solve (psi, and A when the A inversions run on the member's
 * stream), 6 solve_a (A alone), 7 aux); QG_TIME_ALL times every kernel, 0
 * stops timing.  qg_kernel_times reads accumulated milliseconds and launch
 * counts per kernel id and resets the accumulators.  Each timed launch adds
 * two event records to its stream, so time only what is needed where the
 * throughput itself is being measured. */
enum { QG_TIME_ALL = 0xFF };
int qg_timing(qg_ctx* ctx, int32_t mask);
/* Time only every `every`-th launch of each selected kernel (default 1: all),
 * counting from the call; fewer event records on the stream while the
 * average launch duration is still measured inside the timed run.  No
 * counterpart in the reference. */
int qg_timing_sample(qg_ctx* ctx, int32_t every);
int qg_kernel_times(qg_ctx* ctx, double* ms, int32_t* counts, int32_t n);
const char* qg_kernel_name(int32_t id);

/* Kernel variants, v[0..n-1] with n <= QG_NVARIANTS:
 *  v[0..4] the five transform kernels (ps_inv, row_nl, col_upd, row_qw,
 *          col_qw).  ps_inv: 0 = 8 elements per thread, 1 = 16.  col_upd,
 *          col_qw: 0 = 8, 1 = 16 with two LDS exchange buffers per column,
 *          2 = 8, 3 = 16 with one (half the LDS, more barriers).  Row
 *          kernels (row_nl, row_qw): 0 = 8 elements per thread in 64-thread
 *          blocks, 1 = 16/64, 2 = 8/128, 3 = 16/32.
 *  v[5]    shape of the partitioned vertical solve: 0..3 = (levels per
 *          chunk, columns per block) of (8,16), (8,32), (16,16), (16,32).
 *  v[6]    1: the A inversions run in the psi inversion's launch on the
 *          member's stream; 0: on a side stream, overlapped with q^w.
 * Set by the tuner; results agree to FP64 rounding.  Out-of-range values are
 * rejected (QG_ERR_ARG) and leave the settings unchanged. */
enum { QG_NVARIANTS = 7 };
int qg_set_variants(qg_ctx* ctx, const int32_t* v, int32_t n);
int qg_get_variants(const qg_ctx* ctx, int32_t* v, int32_t n);

/* Ensembles: contexts (members or batches, any mix of sizes and parameters,
 * one device) advanced together.  qg_ensemble_step(e, n) advances every
 * context n steps: after the first steps (Euler, full-width transforms,
 * clearing) it replays one captured HIP graph per k steps (k = 8, or
 * qg_ensemble_set_graph_steps), in which every context's kernels run on its
 * own streams, forked from and joined to one origin stream, so the members'
 * kernels overlap on the device.  Results are those of qg_step on each
 * context.  A context may belong to one ensemble at a time and must not be
 * stepped by other calls while an ensemble step runs.  Changing a context's
 * variants makes the ensemble capture its graphs again.  No counterpart in
 * the reference. */
typedef struct qg_ens qg_ens;
int qg_ensemble_create(qg_ctx* const* members, int32_t n, qg_ens** out);
int qg_ensemble_destroy(qg_ens* e);
const char* qg_ensemble_error(const qg_ens* e);
int qg_ensemble_set_graph_steps(qg_ens* e, int32_t k);
int qg_ensemble_step(qg_ens* e, int32_t nsteps);

/* Tuner: coordinate search over the QG_NVARIANTS kernel variants of one
 * context, timing `steps` whole steps per trial with HIP events from the
 * context's current state (saved before and restored after, so the state is
 * unchanged).  Leaves the fastest variants set; best (QG_NVARIANTS values) and
 * best_ms (per step) may be NULL.  Needs device memory for one copy of the
 * state.  No counterpart in the reference. */
int qg_tune(qg_ctx* ctx, int32_t steps, int32_t* best, double* best_ms);

/* Tangent-linear and adjoint of the step, for gradients with respect to the
 * initial streamfunction.  An adjoint context is bound to a context (one
 * member, or a batch whose members get independent gradients in the same
 * launches) with nuz = 0 and flags 0 or QG_NO_FEEDBACK, whose state it uses
 * as the forward trajectory (left at the final state of the last run).  J
 * and dJ return one value per member.  Objective: J = 1/2 * (2 pi)^2 dz * sum over the levels with
 * z_s < z_deep of sum over L = 1 modes of W kh2 (|AR|^2 + |AI|^2) after
 * nsteps, i.e. the domain integral of |grad_h A|^2 / 2 below the depth
 * L3 - z_deep (W: 2 per half-spectrum mode, 1 at kx = ky = 0).  Controls and
 * gradients are physical fields (ldx, n2, n3) like qg_set_state_physical's;
 * the forward run starts from psir, BRr, BIr with A = 0.  The adjoint sweep
 * recomputes the forward trajectory from checkpoints stored every
 * checkpoint_every steps.  No counterpart in the reference.
 *  qg_adj_objective: J only.
 *  qg_adj_tlm:       directional derivative dJ along dpsir, and (if dstate is
 *                    not NULL) the perturbation of the final state, 9 spectral
 *                    fields (psi, q, qo, BR, BI, BRo, BIo, AR, AI) in the
 *                    qg_get_state layout, one after the other.
 *  qg_adj_apply:     transpose of that map: the physical psi adjoint of a
 *                    given final-state adjoint (same 9-field layout).
 *  qg_adj_gradient:  J and dJ/dpsir. */
typedef struct qg_adj qg_adj;
int qg_adj_create(qg_ctx* ctx, int32_t checkpoint_every, double z_deep, qg_adj** out);
int qg_adj_destroy(qg_adj* a);
const char* qg_adj_error(const qg_adj* a);
/* 1 (default): the adjoint step runs on the fused transform kernels; 0: on
 * generic pointwise kernels and separate transforms (the reference form the
 * fused kernels are tested against).  Both give the same results to FP64
 * rounding. */
int qg_adj_set_fused(qg_adj* a, int32_t fused);
int qg_adj_objective(qg_adj* a, const double* psir, const double* BRr, const double* BIr, int32_t ldx,
                     int32_t nsteps, double* J);
int qg_adj_tlm(qg_adj* a, const double* psir, const double* BRr, const double* BIr, int32_t ldx, int32_t nsteps,
               const double* dpsir, double* dJ, double* dstate);
int qg_adj_apply(qg_adj* a, const double* psir, const double* BRr, const double* BIr, int32_t ldx, int32_t nsteps,
                 const double* state_bar, double* psi_bar);
int qg_adj_gradient(qg_adj* a, const double* psir, const double* BRr, const double* BIr, int32_t ldx,
                    int32_t nsteps, double* J, double* grad);

/* Host-only helpers (no GPU needed): the dealiasing mask L(iktx, ikty) as the
 * library evaluates it, and the box geometry {kmax, KX, nlo, nhi}. */
int qg_mask(int32_t n1, int32_t n2, int32_t* L);
int qg_box(int32_t n1, int32_t n2, int32_t out[4]);
const char* qg_version(void);

#ifdef __cplusplus
}
#endif

#endif /* QGYBJ_H */
