/*
 * periodicflow C ABI: the Fourier-Galerkin right-hand side and time stepping
 * of periodicflow on one MI355X (gfx950), as hand-written HIP kernels.
 *
 * The reference is a pure-Python package with no native interface; these
 * entry points replace the array work behind its Python API:
 *
 *   pf_plan_create           FourierSpace.__init__  src/periodicflow/basis.py:34-92
 *                            SpaceSolver.__init__   src/periodicflow/space_solver.py:107-150
 *                            TimeIntegrator.__init__ src/periodicflow/time_integrator.py:17-29
 *   pf_set_linear_operator   Solver.set_linear_operator src/periodicflow/simulation.py:182-184
 *                            TimeIntegrator.update_linear_operator time_integrator.py:31-33
 *   pf_set_correlation       Solver.set_correlation_function simulation.py:186-187
 *   pf_forward / pf_backward SpaceSolver.forward / backward space_solver.py:215-225
 *   pf_project               math.leray_projection (as called by Solver.initialize,
 *                            simulation.py:170-177)
 *   pf_rhs_nonlinear         SpaceSolver.compute_rhs_nonlinear space_solver.py:286-300
 *   pf_step                  the loop body of Solver.solve simulation.py:205-213:
 *                            u0 = u_hat; random_fields(); per stage compute_rhs(a)
 *                            (space_solver.py:302-310) then the integrator's stepping
 *   pf_compute_rhs           SpaceSolver.compute_rhs space_solver.py:302-310
 *   pf_stage_update          <integrator>.stepping time_integrator.py:51-318
 *   pf_draw_noise            SpaceSolver.random_fields space_solver.py:227-238
 *   pf_diagnostics, pf_spectrum, pf_accumulate_sk
 *                            FP64 reductions; no counterpart in the reference
 *
 * Conventions: every array argument is a device pointer in the plan's
 * precision (float / double; complex = interleaved pairs), C order, one
 * member after another.  u_hat is (batch, dim, N0, [N1,] N_last/2+1), the
 * physical velocity (batch, dim, N0, [N1,] N_last).  All work is enqueued
 * on the given stream (0 = null stream).  Return value 0 on success or a
 * negative PF_E* code; pf_last_error() describes the last failure of the
 * calling thread.  A plan is not thread safe.
 */
#ifndef PERIODICFLOW_H
#define PERIODICFLOW_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PF_OK 0
#define PF_EINVAL (-1)
#define PF_ENOMEM (-2)
#define PF_EHIP (-3)
#define PF_EUNSUPPORTED (-4)

enum { PF_FP32 = 0, PF_FP64 = 1 };
enum { PF_DEALIAS_PAD32 = 0, PF_DEALIAS_TRUNC23 = 1 };
enum { PF_EXPLICIT_PC = 0, PF_IMPLICIT_PC = 1, PF_RK2 = 2, PF_RK3 = 3, PF_RK4 = 4 };
enum { PF_NOISE_THERMAL = 0, PF_NOISE_CORRELATED = 1 };
enum { PF_FORCING_NONE = 0, PF_FORCING_CONSTANT_POWER = 1 };

typedef struct pf_desc {
  int32_t dim;          /* 2 or 3 */
  int32_t n[3];         /* grid points per axis (even; 2^a 3^b) */
  double length[3];     /* box lengths; wavenumbers are scaled by 2 pi / length */
  int32_t precision;    /* PF_FP32 / PF_FP64 */
  int32_t dealias;      /* PF_DEALIAS_PAD32 ('3/2') / PF_DEALIAS_TRUNC23 ('2/3') */
  int32_t mask_nyquist; /* zero the Nyquist planes of the nonlinear term */
  int32_t integrator;   /* PF_EXPLICIT_PC .. PF_RK4 */
  int32_t nonlinear;    /* 0: Stokes, 1: Navier-Stokes */
  int32_t noise_type;   /* PF_NOISE_THERMAL / PF_NOISE_CORRELATED */
  int32_t forcing;      /* PF_FORCING_NONE / PF_FORCING_CONSTANT_POWER */
  int32_t batch;        /* members advanced together (same shape and parameters) */
  double viscosity;
  double dt;
  double noise_mag;
  double forcing_power;  /* P of f = P / (2 E_band) u_hat */
  double forcing_kmin;   /* band kmin <= |k| <= kmax */
  double forcing_kmax;
} pf_desc;

typedef struct pf_plan pf_plan;

typedef struct pf_plan_info {
  int64_t n_fourier;       /* complex modes per component per member */
  int64_t n_physical;      /* real points per component per member */
  int32_t padded[3];       /* transform lengths of the padded grid */
  int32_t noise_components;/* d*d (thermal) or d (correlated) */
  int64_t workspace_bytes; /* device memory owned by the plan */
  int64_t state_bytes;     /* bytes of one u_hat (all members) */
} pf_plan_info;

int pf_plan_create(const pf_desc* desc, pf_plan** plan);
int pf_plan_destroy(pf_plan* plan);
int pf_plan_get_info(const pf_plan* plan, pf_plan_info* info);

/* Per-member RNG seeds (host array of `batch` values) for on-device noise. */
int pf_set_seeds(pf_plan* plan, const uint64_t* seeds);
/* Step counter: pf_get_step_counter returns the number of steps pf_step
 * has enqueued; pf_set_step_counter sets it and every member's device-side
 * RNG step counter (starts at 0). */
int pf_set_step_counter(pf_plan* plan, int64_t step);
int64_t pf_get_step_counter(const pf_plan* plan);
/* Adds delta to the host step counter only.  For callers that capture
 * pf_step into a HIP graph: capture enqueues (and counts) a step that does
 * not run, and each replay advances the device counters but not this one. */
int pf_add_step_counter(pf_plan* plan, int64_t delta);
/* Slot management for continuous batching (no reference counterpart):
 * the on-device noise of member m is keyed by (seed[m], its own step
 * counter), which advances on the device once per step while the member is
 * active, so a member's result does not depend on its neighbours or on
 * when it joined.  pf_set_member sets one slot's seed and step counter;
 * pf_set_active limits every call to members [0, n) (default: batch);
 * pf_get_member_steps copies the batch counters to the host. */
int pf_set_member(pf_plan* plan, int32_t member, uint64_t seed, int64_t step);
int pf_set_active(pf_plan* plan, int32_t n);
int pf_get_member_steps(pf_plan* plan, int64_t* out);

/* L: device array (N0, [N1,] nh) of reals, shared by all members; NULL
 * restores -nu |k|^2.  Copied into the plan. */
int pf_set_linear_operator(pf_plan* plan, const void* L, void* stream);
/* C: device array (N0, [N1,] nh) of complex values.  Copied into the plan.
 * A correlated-noise plan starts with C = 0 (the reference's zero
 * correlation_func, space_solver.py:148): its noise is zero until set. */
int pf_set_correlation(pf_plan* plan, const void* C, void* stream);

/* Unpadded transforms of `ncomp` components per member.
 * forward: normalised by 1/prod(N); backward: unnormalised. */
int pf_forward(pf_plan* plan, const void* u, void* u_hat, int32_t ncomp, void* stream);
int pf_backward(pf_plan* plan, const void* u_hat, void* u, int32_t ncomp, void* stream);

/* In-place Leray projection of u_hat (all members). */
int pf_project(pf_plan* plan, void* u_hat, void* stream);
/* Mask the k = 0 mode of u_hat in place (Solver.initialize mask_zero_mode). */
int pf_mask_zero_mode(pf_plan* plan, void* u_hat, void* stream);

/* Nonlinear term P(F(u x omega)) with the plan's dealiasing (Nyquist mask
 * applied when mask_nyquist is set).  rhs may not alias u_hat. */
int pf_rhs_nonlinear(pf_plan* plan, const void* u_hat, void* rhs, void* stream);

/* pf_rhs_nonlinear without a plan and without allocating: no counterpart in
 * the reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * desc describes the grid, batch, precision and dealiasing as for
 * pf_plan_create (desc->nonlinear must be 1); u_hat and rhs are laid out as
 * for a plan of that descriptor.  Twiddle tables and work buffers live in a
 * static device array of 128 MiB per GPU (held from library load until the
 * process ends); descriptors that need more are rejected with PF_EINVAL
 * (FP32: up to 96^3 under the 3/2 rule, 128^3 under the 2/3 rule, 1024^2;
 * FP64: 64^3 under the 3/2 rule).  Calls on one GPU share that array:
 * calls that overlap in time on one GPU (other streams, other threads) race
 * and must be ordered by the caller. */
int pf_rhs_nonlinear_noalloc(const pf_desc* desc, const void* u_hat, void* rhs, void* stream);

/* Advance u_hat by n_steps time steps in place.
 * noise == NULL: noise drawn on the device (Philox4x32-10 keyed by the
 *   member seed, the step counter, the mode and the component).
 * noise != NULL: raw standard-normal draws, n_steps blocks of
 *   (active, 2, noise_components, nk) complex values (w1 then w2), the
 *   quantities np.random.randn fills in random_fields before scaling.
 * The step counter advances by n_steps. */
int pf_step(pf_plan* plan, void* u_hat, int64_t n_steps, const void* noise, void* stream);

/* pf_step of n_steps for several plans at once (an ensemble whose members
 * have different grids, one plan per grid): every stage issues each pass of
 * all plans together, and two plans whose transform lengths form a compiled
 * pair (FP32, default launch variants: 192 + 144, 128 + 96, 144 + 96,
 * 256 + 192) share one launch of that pass.  Results are those of pf_step on
 * each plan alone.  Plans that cannot go pass by pass together (Stokes flow,
 * other integrators or dimension, no fused stage pass, profiling on) step
 * one after another.  Noise is the on-device stream.  paired (nullable)
 * receives the number of paired launches issued.  No counterpart in the
 * reference, which runs one solve per MPI job (simulation.py:205-213). */
int pf_step_multi(pf_plan* const* plans, void* const* u_hats, int32_t n_plans, int64_t n_steps, int32_t* paired,
                  void* stream);

/* pf_step of n_steps with the static structure factor accumulated after
 * every step: acc (device doubles, nk) += sum over the active members and
 * components of |u_hat|^2 (pf_accumulate_sk after each step).  Stokes flow
 * with implicit PC, no forcing and on-device noise runs the steps with the
 * state held in registers and acc summed in FP64 registers (one launch per
 * 200 steps).  The loop of Solver.solve (simulation.py:205-213) with a
 * reduction the reference lacks (its thermal-noise example writes snapshots,
 * examples/fluctuating_stokes_thermal_noise.py:24-36, for statistics
 * afterwards). */
int pf_step_accumulate_sk(pf_plan* plan, void* u_hat, int64_t n_steps, double* acc, void* stream);

/* compute_rhs(a) of one stage (space_solver.py:302-310): writes the
 * projected (and Nyquist-masked) nonlinear term, L u_hat, the forcing and
 * the projected noise built from w1 + a w2 into four arrays shaped like
 * u_hat.  noise: raw draws (batch, 2, noise_components, nk) of the current
 * step, or NULL for the on-device stream at the current step counter. */
int pf_compute_rhs(pf_plan* plan, void* u_hat, double a, const void* noise, void* rhs_nonlinear,
                   void* rhs_linear, void* forcing, void* noise_out, void* stream);
/* The plan integrator's stepping(unew, u0, stage, ...) from given terms
 * (time_integrator.py:51-318).  implicit_pc overwrites rhs_linear with
 * L u0 at its last stage, as the reference does (:116-117). */
int pf_stage_update(pf_plan* plan, int32_t stage, void* unew, const void* u0, const void* rhs_nonlinear,
                    void* rhs_linear, const void* forcing, const void* noise, void* stream);
/* random_fields (space_solver.py:227-238) on the device: raw unscaled
 * draws (batch, 2, noise_components, nk) for the current step counter. */
int pf_draw_noise(pf_plan* plan, void* out, void* stream);

/* FP64 diagnostics.  out: device doubles (batch, 2) = energy 1/2<|u|^2>,
 * enstrophy 1/2<|omega|^2>.  Overwritten. */
int pf_diagnostics(pf_plan* plan, const void* u_hat, double* out, void* stream);
/* Shell spectrum E(k), shells round(|k|): out device doubles (batch, nshell), overwritten. */
int pf_spectrum(pf_plan* plan, const void* u_hat, double* out, int32_t nshell, void* stream);
/* acc (device doubles, nk) += sum over members and components |u_hat|^2. */
int pf_accumulate_sk(pf_plan* plan, const void* u_hat, double* acc, void* stream);

/* Per-pass timing with HIP events recorded on the launch stream around
 * every launch of pf_step / pf_rhs_nonlinear while enabled.  Pass ids:
 * PF_PASS_INV0 axis-0 inverse (+curl), PF_PASS_INV1 axis-1 inverse,
 * PF_PASS_ROW last-axis c2r / cross product / r2c, PF_PASS_FWD1 axis-1
 * forward, PF_PASS_FWD0 axis-0 forward, PF_PASS_STAGE stage update,
 * PF_PASS_BAND forcing-band energy, PF_PASS_FWD0_STAGE axis-0 forward
 * fused with the stage update (pf_step's default).  Enabling (or disabling) clears what
 * was recorded; pf_profile_read synchronises and returns, per id, the
 * summed milliseconds and the launch count.  Tooling; no counterpart in the
 * reference. */
#define PF_PASS_INV0 0
#define PF_PASS_INV1 1
#define PF_PASS_ROW 2
#define PF_PASS_FWD1 3
#define PF_PASS_FWD0 4
#define PF_PASS_STAGE 5
#define PF_PASS_BAND 6
#define PF_PASS_FWD0_STAGE 7
#define PF_NPASS 8
int pf_profile_enable(pf_plan* plan, int32_t enable);
int pf_profile_read(pf_plan* plan, double* total_ms, int64_t* count, int32_t n);
/* Passes (bit 1 << PF_PASS_*) that record events while profiling is on;
 * default all.  Each recorded pass adds two event records to its stream. */
int pf_profile_select(pf_plan* plan, uint32_t pass_mask);
/* 1 (default): pf_step fuses the axis-0 forward pass with the stage update
 * where the tile fits in LDS; 0: separate passes (the unfused sequence the
 * per-stage API uses).  Results agree to rounding. */
int pf_set_fusion(pf_plan* plan, int32_t fuse_stage);

/* Launch plans per pass (no counterpart in the reference, whose transforms
 * are planned by shenfun/FFTW: fft_plan="FFTW_MEASURE", pf/simulation.py:76,
 * pf/basis.py:63-65).  tiles[PF_NPASS], indexed by PF_PASS_*: 0 is the
 * default tile width (column passes) or row-pair count (row passes); FP32
 * plans also have 1 (half) and 2 (twice) where they fit in LDS, column
 * variants 3 (two tiles per block), 4 (stage kinds: 1024-thread blocks of
 * twice the width), 5 (the radix-2/3 plan), 6 (nontemporal stores, whole
 * 128-byte lines stored by the axis-0 inverse and axis-1 forward passes) and
 * 7 (nontemporal stores and loads), and row variants 3 (the radix-2/3 plan)
 * and 4 (nontemporal stores).  Plans of Stokes flow with implicit PC and no
 * forcing have PF_PASS_STAGE variant 1: one launch per step instead of many
 * steps with the state in registers; other plans have PF_PASS_STAGE variant
 * 7: the separate stage kernel with nontemporal loads and stores.  A choice
 * other than a radix plan changes only how work is mapped to threads and
 * cached: results are bit-identical.
 * pf_set_tiles fails with PF_EUNSUPPORTED for a variant that does not exist
 * at this plan's lengths; pf_get_tiles returns the plan's choices and
 * whether the stage pass is fused. */
int pf_set_tiles(pf_plan* plan, const int32_t* tiles);
int pf_get_tiles(const pf_plan* plan, int32_t* tiles, int32_t* fuse_stage);
/* Times every variant of every pass of the nonlinear right-hand side and
 * the fused against the unfused axis-0 forward pass on a copy of u_hat
 * (left unchanged), `reps` evaluations each, and keeps the fastest; then
 * refines that plan on the whole launch sequence (other variants of each
 * pass, the stage kernel's cache policy and the fusion choice, each kept
 * where consecutive stages run at least 0.5% faster), since a pass's stores
 * decide what the next pass finds in the last-level cache.  The
 * FFTW_MEASURE counterpart.  Enqueued on `stream` and synchronised;
 * ms_out[PF_NPASS] (nullable) receives the chosen passes' average times. */
int pf_autotune(pf_plan* plan, const void* u_hat, int32_t reps, double* ms_out, void* stream);

const char* pf_last_error(void);
const char* pf_version(void);

#ifdef __cplusplus
}
#endif
#endif
