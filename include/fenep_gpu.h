/*
 * fenep_gpu.h -- C-ABI of the MI355X FENE-P / IBM SMAC time step.
 *
 * The reference (machapin/fortran8) has no plugin or FFI: the step is a set of
 * Fortran module subroutines with explicit-shape real(8) dummies, called in a
 * fixed order from `program main` (fenep.f90:2477-2527).  This header is the
 * boundary that replaces that loop.  Every entry point takes plain pointers and
 * sizes, returns 0 on success and a nonzero FP_E* code on error (the message is
 * then available from fp_last_error).  A context owns all of its device memory
 * and one HIP stream; contexts are independent, a single context is not
 * re-entrant.
 *
 * Array layouts at the boundary are the reference's (Fortran column-major, x
 * fastest):
 *   scalars   U,V,W,P  : (0:NX+1, 0:NY+1, 0:NZ+1)  ghosted, real(8)
 *   tensor    C        : (6, 0:NX+1, 0:NY+1, 0:NZ+1) components xx,xy,xz,yy,yz,zz
 *   histories A0,T0,B0 : (1:NX, 1:NY, 1:NZ) per component
 *   markers            : (NC, NL, NZ) flattened, l (cylinder) fastest
 * Ghost values passed in are ignored; the library applies the boundary rules
 * itself (periodic copy `PBM`/`C_PBM`, fenep.f90:147-166, or the y-wall rules
 * `U_kabe`..`C_kabe`, fenep_2d.f90:1098-1143).
 */
#ifndef FENEP_GPU_H
#define FENEP_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define FP_OK            0
#define FP_E_ARG         1   /* invalid argument / unsupported parameter      */
#define FP_E_HIP         2   /* HIP runtime error                             */
#define FP_E_STATE       3   /* call out of order (e.g. step before set_state)*/
#define FP_E_NAN         4   /* NaN detected in the velocity field            */
#define FP_E_NOMEM       5

/* method: fenep.f90:17 */
#define FP_METHOD_EXPLICIT 0  /* navier (AB2) + poisson + march               */
#define FP_METHOD_FFT      1  /* fft_navier (Crank-Nicolson) + fft_poisson    */
#define FP_METHOD_IBM      2  /* ibm_preliminary/Helmholtz/predict + poisson  */

#define FP_POISSON_FFT 0      /* spectral (+ tridiagonal line solve with walls) */
#define FP_POISSON_SOR 1      /* red-black SOR, fenep.f90:848-905               */

#define FP_PREC_F64   0       /* plain FP64, parity path                       */
#define FP_PREC_MIXED 1       /* FP32 bulk, FP64 defect correction             */

/* Run parameters.  Names follow the reference's module constants
 * (fenep.f90:4-49); Lp is L itself (f = (Lp^2-3)/(Lp^2-trC), fenep.f90:513). */
typedef struct fp_params {
  int32_t nx, ny, nz;          /* NX, NY, NZ (powers of two, >= 4)            */
  double  lx, ly, lz;          /* domain lengths; dX = lx/NX etc.             */
  double  dt, re, beta, wi, lp;
  int32_t method;              /* FP_METHOD_*                                 */
  int32_t walls_y;             /* 0: periodic in y; 1: no-slip walls at y=0,Ly */
  int32_t poisson_variant;     /* FP_POISSON_*                                */
  int32_t precision;           /* FP_PREC_*                                   */
  int32_t newton_itrmax;       /* Lyapunov Newton, fenep.f90:527 (10000)      */
  double  newton_eps;          /* fenep.f90:528 (1e-6)                        */
  double  newton_dc;           /* FD Jacobian step, fenep.f90:526 (1e-5)      */
  int32_t sor_itrmax;          /* fenep.f90:857 (10000)                       */
  double  sor_eps;             /* fenep.f90:856 (1e-5)                        */
  double  sor_omega;           /* <=0: 2/(1+sin(pi*dX)), fenep.f90:855         */
  double  body_force[3];       /* constant body force (mean pressure gradient)*/
  int32_t n_dc_max;            /* mixed: max FP64 defect-correction sweeps     */
  double  dc_tol;              /* mixed: relative defect tolerance             */
  int32_t helm_sweeps;         /* mixed: fixed-point sweeps of the CN Helmholtz
                                * solve; 0 = as many as bring the error below
                                * FP32 rounding, -1 = the spectral solve       */
  int32_t reserved0;           /* 0                                           */
  double  ibm_weight;          /* IBM spreading weight per marker and the
                                * weight of the drag sum: <= 0 = dX^3
                                * (fenep.f90:2125-2162, the default); > 0 =
                                * that volume, e.g. smac.f90's dV = pi DC/NL
                                * dX^2 (smac.f90:1129, 1428)                   */
  uint64_t seed;               /* seed of the run's synthetic inputs
                                * (fp_seeded_uniform); carried with the
                                * parameters, the step itself draws nothing   */
  int32_t reserved[2];
} fp_params;

/* Diagnostics (log_progress, fenep.f90:2266-2342, plus drag and divergence). */
typedef struct fp_diag {
  double ke;          /* kinetic energy on the regular grid / (NX NY NZ)     */
  double cfl;         /* max(u dt/dX, v dt/dY, w dt/dZ) * 6                  */
  double trc_min, trc_max, trc_mean, trc_std;
  double div_max;     /* max |div_h u| over cells                            */
  double drag;        /* -sum_markers Fxc * dX^3                             */
  double lift;        /* -sum_markers Fyc * dX^3                             */
  double newton_itr_mean;
  double sor_itr;     /* iterations of the last SOR solve (or -1)            */
  double n_dc;        /* defect-correction sweeps of the last mixed solve    */
  double dc_residual; /* mixed: |b - L phi| / |b| before the last correction */
  double sel_guarded; /* mixed: C+- selections re-decided in FP64, last step */
  double newton_unconverged; /* cells whose Lyapunov Newton stopped at its
                              * iteration cap without meeting the tolerance */
  int32_t nan_flag;   /* nonzero: NaN in U (driver prints 'NaN value')        */
  int32_t step;       /* steps taken by this context                         */
} fp_diag;

/* Fill params with the reference defaults (fenep.f90:4-49, method=1, beta=0.9). */
int fp_default_params(fp_params* p);

/* Replaces the module-level state and plan set-up: fft_init (fenep.f90:1374),
 * the allocations of program main (fenep.f90:2424-2446).  FP_E_ARG for axes
 * that are not powers of two up to 1024, and for the mixed precision with
 * the explicit method or 2^30 cells. */
int fp_create(const fp_params* p, void** ctx);
int fp_destroy(void* ctx);
const char* fp_last_error(void* ctx);

/* Replaces init/input (fenep.f90:52-145, 1186-1213): copy a full state in.
 * Every setter (fp_set_state, fp_set_state_device, fp_set_body_force,
 * fp_set_markers, fp_set_physics) first waits for steps still in flight
 * after fp_step(..., sync = 0).
 * hist may be NULL; otherwise it holds 9 interior arrays A0x,A0y,A0z,T0x,T0y,
 * T0z,B0x,B0y,B0z.  has_history=0 reproduces the step-1 rule of fft_navier /
 * ibm_preliminary / navier (fenep.f90:1445, 1985, 787): the first step copies
 * the current A,T(,B) into the histories. */
int fp_set_state(void* ctx, const double* U, const double* V, const double* W,
                 const double* P, const double* C6, const double* hist,
                 int has_history);
/* Copy out (output(), fenep.f90:1215-1233): ghosts filled by the BC rules. */
int fp_get_state(void* ctx, double* U, double* V, double* W, double* P,
                 double* C6, double* hist);
/* Pressure-correction potential Phi, (0:NX+1,0:NY+1,0:NZ+1). */
int fp_get_phi(void* ctx, double* Phi);

/* Body-force field Fx,Fy,Fz (1:NX,1:NY,1:NZ) (init, fenep.f90:96-120); NULL
 * pointers clear it.  The constant params.body_force is added on top. */
int fp_set_body_force(void* ctx, const double* fx, const double* fy,
                      const double* fz);

/* IBM markers (ibm_init, fenep.f90:1769-1900): n = NC*NL*NZ points with
 * positions xc,yc,zc and target velocities uc,vc,wc. nc, nl as in the
 * reference; the marker of plane n sits at zc(n) and interpolates from planes
 * n-1..n+1. */
int fp_set_markers(void* ctx, int nc, int nl, const double* xc,
                   const double* yc, const double* zc, const double* uc,
                   const double* vc, const double* wc);
/* Marker forces Fxc,Fyc,Fzc of the last step, (NC,NL,NZ). */
int fp_get_marker_forces(void* ctx, double* fxc, double* fyc, double* fzc);

/* Advance nsteps time steps (the body of fenep.f90:2477-2527).  sync != 0
 * waits for completion; otherwise returns after enqueueing on the context's
 * stream. */
int fp_step(void* ctx, int nsteps, int sync);

int fp_diagnostics(void* ctx, fp_diag* d);

/* fp_diagnostics without any allocation.  No counterpart in the reference
 * (its diagnostics are plain loops over host arrays); it is for callers that
 * cannot allocate device memory: code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  The same reductions as
 * fp_diagnostics (bit-identical results) run in a static device array of the
 * library, one per GPU, held until the process ends.  Grids above
 * FP_DIAG_STATIC_MAX_CELLS cells return FP_E_ARG.  Calls that overlap on one
 * GPU (from several host threads) share that array and race. */
#define FP_DIAG_STATIC_MAX_CELLS (1L << 24)
int fp_diagnostics_static(void* ctx, fp_diag* d);

/* Device-pointer variants for callers that keep the state resident (the
 * PyTorch wrapper passes tensor.data_ptr()).  Layout: interior only, x
 * fastest, FP64: U,V,W,P are nx*ny*nz each, C6 is 6 consecutive planes
 * (structure of arrays).  `stream` is a hipStream_t or NULL for the
 * context's own stream.  has_history of fp_set_state_device: 0 and 1 as in
 * fp_set_state (the potential Phi starts from 0, as after init,
 * fenep.f90:95); 2 as 1 and Phi kept, for a state that continues the
 * context's own run (the SOR variant's sweeps start from Phi). */
int fp_set_state_device(void* ctx, const double* U, const double* V,
                        const double* W, const double* P, const double* C6,
                        int has_history, void* stream);
int fp_get_state_device(void* ctx, double* U, double* V, double* W, double* P,
                        double* C6, void* stream);

/* Per-kernel timing of the dominant kernels over the last fp_step call with
 * timing enabled (milliseconds, summed over steps). names is filled with a
 * comma-separated list; returns the number of entries written to ms.
 * on: 0 off, 1 every phase, 2 only the predictor phase (two events per
 * step, for timing runs), 3 the predictor phase of the steps whose index
 * (fp_diag.step before the step) is a multiple of 8. */
int fp_enable_timing(void* ctx, int on);
int fp_kernel_times(void* ctx, double* ms, int max_n, char* names,
                    int names_len);

/* Time-series statistics of all_time_dif (fenep_mpi.f90:2200-2348) of the
 * current state, of the velocity less the mean set by fp_mean_apply (none by
 * default: the Log/time_same.d record; otherwise Log/time_diff.d), and
 * with unit reference scales (U_C = L_C = 1, nu = 1/Re).  out[13] = step,
 * step*dt, K, K_staggered, lambda, Re_lambda, epsilon, eta, tau_eta, v_eta,
 * Ep, trC mean, trC std (the column order of fenep_mpi.f90:2343-2344).
 * Reductions in FP64 on the device. */
int fp_time_stats(void* ctx, double* out);

/* Shell-summed kinetic-energy spectrum of the regular-grid velocity less the
 * mean set by fp_mean_apply (fft_energy_dif, fenep_mpi.f90:2566-2665): forward FFT
 * / (NX NY NZ), |u^|^2 of the three components summed into shells
 * nint(|k|) with half weight on the kx = 0 and Nyquist planes.  E[0..nbins)
 * receives shells 0..nbins-1 (the reference writes 0..NX-1); periodic grids
 * only.  sum(E) over all shells equals K of fp_time_stats (Parseval). */
int fp_energy_spectrum(void* ctx, double* E, int nbins);

/* all_time_dif_map (fenep_mpi.f90:2350-2564) of the current state: the
 * statistics of fp_time_stats per (i, j) column, averaged over z, and their
 * plane averages over the cells outside 1.5 cylinder radii of the given
 * centres (the reference masks four fixed centres, fenep_mpi.f90:2530-2537).
 * xy: ncyl (x, y) pairs; diameter: the cylinders' diameter DC.
 * map18 (NX*NY*18, may be NULL): the Map/{same,diff}_<step>.bin records, j
 * outer and i inner: x, y, K, lambda, Re_lambda, epsilon, eta, tau_eta,
 * v_eta, Ep, trC mean, trC std, mask, 0 x 5.  out13 (may be NULL): the
 * Log/time_{same,diff}_mask.d record, the column order of fp_time_stats
 * with K_staggered = 0.  FP_E_ARG if the mask covers every column. */
int fp_time_stats_map(void* ctx, int ncyl, const double* xy, double diameter,
                      double* map18, double* out13);

/* fft_energy_dif_map (fenep_mpi.f90:2670-2797) of the current state: each
 * (i, j) column's regular-grid velocity transformed along z (/ NZ), energy
 * (|u^|^2 + |v^|^2 + |w^|^2)/2 per |kz|, doubled for 0 < kz < NZ/2.
 * Emask[NZ]: the plane average over the unmasked columns (mask as in
 * fp_time_stats_map); Efirst[NZ]: the column (1, 1).  Entries above NZ/2 are
 * zero.  The reference writes their averages over the records of a Gstep
 * period (Log/energy_*_mask_<step>.d, Log/energy_*_z_<step>.d). */
int fp_energy_spectrum_z(void* ctx, int ncyl, const double* xy, double diameter,
                         double* Emask, double* Efirst);

/* Time mean of the velocity for the mean-subtracted statistics (the second
 * pass of input_type = 2, fenep_mpi.f90:2969-2971, 2999-3067).
 * fp_mean_accumulate adds the current U, V, W to a device accumulator.
 * fp_mean_apply(mode): 0 clears the subtracted mean (the 'same' records);
 * 1 makes the accumulated mean, averaged over z (ave_mean,
 * fenep_mpi.f90:2010-2033), the mean that fp_time_stats, fp_time_stats_map,
 * fp_energy_spectrum and fp_energy_spectrum_z subtract (the 'diff' records);
 * 2 the same without the z average.  Modes 1 and 2 need at least one
 * accumulated state; every mode empties the accumulator. */
int fp_mean_accumulate(void* ctx);
int fp_mean_apply(void* ctx, int mode);

/* Mixed precision: the C+- selection of the current stored (FP32) field as
 * the next step makes it (FP32 decisions that carry an error certificate, the
 * rest re-decided with the reference's formula in FP64), compared with the
 * reference's selection (fenep.f90:186-221 with the Vieta test of
 * fenep.f90:425-444, in FP64) of every cell and axis.  out[3] = decisions
 * re-decided in FP64, decisions that differ from the reference's, decisions
 * in total (3 NX NY NZ).  The state is not changed. */
int fp_check_selection(void* ctx, double* out);

/* Mixed precision: the C+- selection that fp_check_selection checks, as
 * bytes: out[i + NX (j + NY k)] over the interior cells, bits 2d and 2d+1 the
 * choice along axis d (x, y, z): 0, 1, 2 the candidate levels of
 * fenep.f90:179-184 (maxloc - 1 of fenep.f90:186-221), 3 the zero slope.
 * With beta = 1 (no polymer stress, no selection) out is set to zero.  The
 * state is not changed.  No counterpart in the reference, whose selection
 * stays inside Cstar: it lets a caller compare the selection with the
 * reference formula on the host. */
int fp_get_selection(void* ctx, unsigned char* out);

/* ---- Sensitivities (periodic grids, FFT Poisson, methods 1 and 2).  The
 * reference has no derivative code: these entry points serve the discrete
 * adjoint of the step it advances (fenep.f90:2477-2527) with respect to Wi,
 * beta, L^2 = Lp^2 (fenep.f90:23-26 module constants) and the body-force field
 * F of init (fenep.f90:96-120).  On an FP64 context the adjoint is that of the
 * recorded FP64 steps.  On a mixed-precision context the recording runs the
 * mixed step (its drag and its FP32 state, widened, as checkpoints) and the
 * reverse sweep is the FP64 adjoint along mixed-step recomputations of each
 * checkpoint segment (their states widened to FP64), with the C+- selections
 * those steps make. */

/* Change Wi, beta and Lp of a context between runs. */
int fp_set_physics(void* ctx, double wi, double beta, double lp);

/* Advance nsteps steps from the current state while recording: the state is
 * checkpointed every `every` steps, the C+- selections of every step and the
 * drag D_s = -sum_markers Fxc dX^3 of every step (IBM method; 0 otherwise)
 * are kept; drag_out (host, nsteps, may be NULL) receives D_s.  replay != 0
 * reuses the selections of the previous recording instead of computing them
 * (the same discrete branch; for finite-difference checks; FP64 contexts). */
int fp_adj_record(void* ctx, int nsteps, int every, int replay, double* drag_out);

/* Reverse sweep over the last recording for the objective
 *   J = sum_s weights[s] D_s + <bar_final, state after the last step>.
 * weights: host, nsteps (NULL = 0).  bar_final / bar_initial: device arrays
 * of 16 planes of NX*NY*NZ doubles (interior, x fastest): U V W P C(6) A0(3)
 * T0(3); NULL = zero / not wanted.  bar_params (host, 3): dJ/dWi, dJ/dbeta,
 * dJ/d(L^2).  bar_force (device, 3 planes): dJ/dF.  The selections are those
 * of the recording (frozen branch).  The context's state is left at the
 * state after the recording. */
int fp_adj_backward(void* ctx, const double* weights, const double* bar_final,
                    double* bar_initial, double* bar_params, double* bar_force);

/* Identity of the sources the library was built from: 16 hex digits of a
 * SHA-256 over the HIP sources, this header and the compile flags.  No
 * counterpart in the reference (it is compiled in one piece with the program);
 * host bindings compare it with the tree they run from and refuse a library
 * built from other sources. */
const char* fp_build_id(void);

/* Deterministic seeded uniform numbers in [0,1): SplitMix64 of
 * (seed, offset+i), top 53 bits.  Host-side utility so that the Fortran
 * driver, the Python wrapper and the tests produce identical inputs. */
int fp_seeded_uniform(uint64_t seed, uint64_t offset, size_t n, double* out);

#ifdef __cplusplus
}
#endif
#endif /* FENEP_GPU_H */
