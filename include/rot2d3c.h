/*
 * rot2d3c.h -- C ABI of the MI355X (gfx950) implementation of the Rot2D3C
 * pseudospectral right-hand side and its 4-stage Runge-Kutta integrator.
 *
 * Every entry point takes plain pointers and sizes.  Device arrays use the
 * reference's layout for one MPI rank, so a reference array and a device
 * buffer hold identical bytes:
 *   spectral field : complex128, (n/2+1) x n, index [kx][ky], ky fastest, ky
 *                    in FFT order, values unnormalised (n^2 x coefficient)
 *                    == Fortran DOUBLE COMPLEX C(n, 1:n/2+1)
 *                    (src/Rot2D3C.para.f90:107-117, src/fftp2D.f90:153)
 *   real field     : float64, n x n, index [y][x], x fastest
 *                    == Fortran DOUBLE PRECISION R(n, 1:n)
 * Complex arrays are passed as `double *` to interleaved (re, im) pairs.
 *
 * All device work is asynchronous on the caller's HIP stream (passed as
 * `void *`, i.e. a hipStream_t; NULL = the default stream).  Device memory
 * is allocated only by r2d3c_plan_create, r2d3c_autotune (scratch, freed
 * before it returns), r2d3c_ensemble_create, r2d3c_ensemble_set_closure,
 * r2d3c_ensemble_run (when the forcing table grows), r2d3c_cnn_create and
 * r2d3c_cnn_set_forcing_gain (its first call); the plan operators,
 * r2d3c_step and r2d3c_run allocate nothing, and r2d3c_run_noalloc allocates
 * nothing at all.  A plan is not thread-safe: use one host thread per plan.  Every function returns 0 on
 * success, a negative R2D3C_E* code on an argument error, or
 * -(1000 + hipError_t) when the HIP runtime fails; r2d3c_last_error()
 * returns a description of the last failure on the calling thread.
 */
#ifndef ROT2D3C_H
#define ROT2D3C_H

#ifdef __cplusplus
extern "C" {
#endif

#define R2D3C_OK 0
#define R2D3C_EINVAL (-1)     /* bad argument (size, pointer, exponent) */
#define R2D3C_ENOTSUP (-2)    /* grid size without a compiled kernel    */
#define R2D3C_EFAIL (-3)      /* a device-side check failed             */
#define R2D3C_EHIP_BASE 1000  /* -(1000 + hipError_t)                   */

typedef struct r2d3c_plan r2d3c_plan;

/* Dissipation and rotation parameters, parameter.inp lines 12-20
 * (src/Rot2D3C.para.f90:147-155). */
typedef struct r2d3c_params {
  double nu, hnu;   /* viscosity and hypo-viscosity of psi          */
  double nuv, hnuv; /* viscosity and hypo-viscosity of vz           */
  double omega;     /* rotation rate                                */
  int nn, mm;       /* orders: nu*k^(2nn), hnu*k^(-2mm) for psi     */
  int nnv, mmv;     /* same for vz                                  */
} r2d3c_params;

/* One draw of the iflow=3 random forcing for one field, as produced by the
 * host part of rand_force (src/pseudospec2D_Rot2D3C.f90:1086-1097):
 * wavevector (kx, ky) rounded to integers, phase in radians, injection rate
 * f0 (fp0 or fz0) and kupk = kup**kin.  The amplitude sqrt(2*f0)/(sqrt(dt)*kupk)
 * is formed on the device from the step's dt.  enabled = 0 means no forcing. */
typedef struct r2d3c_force_draw {
  double kx, ky, phase, f0, kupk;
  int enabled;
} r2d3c_force_draw;

/* Forcing of one step: a rand_force draw per field (iflow = 3) and/or a
 * static dense spectral array per field (iflow = 1, src/Rot2D3C.para.f90:343-371;
 * NULL = none), added as fp / fz in the stage update. */
typedef struct r2d3c_forcing {
  r2d3c_force_draw psi; /* fp: kin = 1 */
  r2d3c_force_draw vz;  /* fz: kin = 0 */
  const double *dense_psi;
  const double *dense_vz;
  /* iflow = 2: fp = const_inj(C1, ci_kdn, ci_kup, ci_fp0, 1) and
   * fz = const_inj(C2, ci_kdn, ci_kup, ci_fz0, 0) at every stage
   * (src/Rot2D3C.para.f90:565-569); the step then runs stage by stage.
   * ci_nprocs: the MPI rank count whose random numbers to reproduce
   * (<= 1: one rank).  In one r2d3c_run call every step must have the same
   * ci_enabled and (with iflow = 2) the same ci_kdn, ci_kup, ci_nprocs.  The
   * first iflow = 2 step of a plan, and the first after a change of band or
   * rank count, uploads the random-number table and synchronizes the stream
   * (not capturable); later ones only launch kernels. */
  int ci_enabled;
  double ci_kdn, ci_kup, ci_fp0, ci_fz0;
  int ci_nprocs;
} r2d3c_forcing;

#define R2D3C_DT_FIXED 0 /* dt given by the caller (no reference counterpart) */
#define R2D3C_DT_CFL 1   /* dt from CFL_condition on the pre-step state     */

/* ---- plan ------------------------------------------------------------- */

/* Creates the plan of an n x n grid on the current HIP device: twiddle tables
 * and the workspace of the fused step.  Replaces fftp2d_create_plan
 * (src/fftp2D.f90:30-63) and the wavenumber set-up (src/Rot2D3C.para.f90:213-228).
 * n: power of two, 16 <= n <= 4096, or 16384.  At 16384 the step runs stage
 * by stage from full 2D transforms (no fused kernels, iflow = 2, ensembles or
 * closure); the plan then holds about 41 GB. */
int r2d3c_plan_create(r2d3c_plan **plan, int n);
/* Replaces fftp2d_destroy_plan (src/fftp2D.f90:66-86). */
int r2d3c_plan_destroy(r2d3c_plan *plan);
/* Device bytes held by the plan. */
long long r2d3c_plan_workspace_bytes(const r2d3c_plan *plan);
int r2d3c_plan_size(const r2d3c_plan *plan);
const char *r2d3c_last_error(void);

/* ---- transforms ------------------------------------------------------- */

/* Unnormalised inverse 2D FFT, e^{+i}; the imaginary parts at kx = 0 and
 * kx = n/2 are ignored.  Replaces fftp2d_complex_to_real
 * (src/fftp2D.f90:214-290); unlike the reference the input is preserved. */
int r2d3c_c2r(r2d3c_plan *plan, const double *in_spec, double *out_real, void *stream);
/* Unnormalised forward 2D FFT, e^{-i}.  Replaces fftp2d_real_to_complex
 * (src/fftp2D.f90:133-211). */
int r2d3c_r2c(r2d3c_plan *plan, const double *in_real, double *out_spec, void *stream);

/* ---- spectral operators ----------------------------------------------- */

/* dir 1: i*kx*a, dir 2: i*ky*a, dir 0: copy; zero where kx^2+ky^2 > (n/3)^2.
 * Replaces derivk2 (src/pseudospec2D_Rot2D3C.f90:6-68). */
int r2d3c_derivk2(r2d3c_plan *plan, const double *a, double *b, int dir, void *stream);
/* b = -k^2 a.  Replaces laplak2 (src/pseudospec2D_Rot2D3C.f90:71-95). */
int r2d3c_laplak2(r2d3c_plan *plan, const double *a, double *b, void *stream);
/* b = -a/k^2 (k != 0), 0 at k = 0.  Replaces inv_laplak2
 * (src/pseudospec2D_Rot2D3C.f90:98-127). */
int r2d3c_inv_laplak2(r2d3c_plan *plan, const double *a, double *b, void *stream);
/* c = {a, b} = r2c[(da/dx db/dy - da/dy db/dx)/n^4] with dealiased inputs.
 * Replaces poisson (src/pseudospec2D_Rot2D3C.f90:169-226). */
int r2d3c_poisson(r2d3c_plan *plan, const double *a, const double *b, double *c,
                  void *stream);

/* ---- reductions (result written to device memory, deterministic) ------ */

/* sum two*|a|^2*k^(2 kin)/n^4.  Replaces energy (src/pseudospec2D_Rot2D3C.f90:336-396). */
int r2d3c_energy(r2d3c_plan *plan, const double *a, int kin, double *d_out, void *stream);
/* sum two*k^(2 kin)*Re(b conj a)/n^4 over the dealiased disc.  Replaces
 * inerprod (src/pseudospec2D_Rot2D3C.f90:400-451). */
int r2d3c_inerprod(r2d3c_plan *plan, const double *a, const double *b, int kin,
                   double *d_out, void *stream);

/* Constant-injection forcing of iflow = 2: fp = ps/(|ps|^2+1) on the band
 * kdn^2 <= k^2 <= kup^2 rescaled by fp0/inerprod(ps, ., kin), plus
 * i*randu*|k|*ps.  The random numbers are those of an nprocs-rank run
 * (<= 1: one rank): each rank owns the kx columns of its slab and restarts
 * randu at seed = rank on every call.  Replaces const_inj
 * (src/pseudospec2D_Rot2D3C.f90:977-1060). */
int r2d3c_const_inj(r2d3c_plan *plan, const double *ps, double kdn, double kup, double fp0,
                    int kin, int nprocs, double *fp, void *stream);

/* Copies the fp, fz of the last stage of the most recent iflow = 2 step (the
 * forcing cond_check sees at the next iteration) into fp, fz (either may be
 * NULL).  No reference counterpart: the reference keeps them in its arrays. */
int r2d3c_const_inj_last(r2d3c_plan *plan, double *fp, double *fz, void *stream);

/* ---- diagnostics (device results, deterministic) ----------------------- */

/* The balance sums of cond_check (src/pseudospec2D_Rot2D3C.f90:455-612) in
 * d_out[26]: Ep1 injp1 Dsp Hds Epv injz Dspv Hdsv coup Ep2 injp2 Dspw Hdsw Eph
 * injh1 injh2 Dsph1 Hdsh1 Dsph2 Hdsh2 Efp Efpz, then max and min of
 * c2r(laplak2(psi)) and of c2r(derivk2(vz,0)).  fp, fz: dense forcing arrays
 * of the previous step (zeros without forcing).  The extrema are those of an
 * nprocs-rank run (<= 1: one rank): each rank reduces its slab of real-space
 * rows and both are combined with MPI_MAX (:578,586), so "min" is the largest
 * slab minimum.  Writing the four balance files is left to the host. */
int r2d3c_cond_check(r2d3c_plan *plan, const r2d3c_params *p, double kup, const double *psi,
                     const double *vz, const double *fp, const double *fz, int nprocs,
                     double *d_out, void *stream);
/* Shell spectra (src/pseudospec2D_Rot2D3C.f90:615-712): d_out[2*(n/2+1)] =
 * Ek(k) then Ezk(k), k = 1..n/2+1. */
int r2d3c_spectrum(r2d3c_plan *plan, const double *psi, const double *vz, double *d_out,
                   void *stream);
/* Shell transfers (src/pseudospec2D_Rot2D3C.f90:715-837): d_out[3*(n/2+1)] =
 * Fl0 (enstrophy), Fl1 (2D energy), Fl2 (vz energy) per shell; the fluxes
 * are their running sums from k = 1. */
int r2d3c_transfers(r2d3c_plan *plan, const double *psi, const double *vz, double *d_out,
                    void *stream);

/* ---- time step -------------------------------------------------------- */

/* dt = cfl/(kcut*max|grad psi| + nu*kcut^(2nn) + nuv*kcut^(2nnv) + 2*omega),
 * kcut = n/3.  Replaces CFL_condition (src/pseudospec2D_Rot2D3C.f90:840-899). */
int r2d3c_cfl(r2d3c_plan *plan, const r2d3c_params *p, double cfl, const double *psi,
              double *d_dt, void *stream);

/* One time step of the 4-stage scheme, in place on psi and vz: C1=psi, C2=vz;
 * for o = 4..1 the semi-implicit update with the Poisson brackets
 * J(psi, lap psi), J(psi, vz), the 2*Omega*d/dy coupling and the forcing;
 * then psi = mask(C1), vz = mask(C2).  Replaces the loop body
 * src/Rot2D3C.para.f90:399 (CFL) and :545-611 (stages).
 * dt_mode R2D3C_DT_CFL: dt_or_cfl is the CFL number; R2D3C_DT_FIXED: dt.
 * The dt used is written to d_dt (device, may be NULL).  forcing may be NULL. */
int r2d3c_step(r2d3c_plan *plan, const r2d3c_params *p, double *psi, double *vz,
               const r2d3c_forcing *forcing, int dt_mode, double dt_or_cfl,
               double *d_dt, void *stream);

/* n_steps consecutive time steps in place, equal bit for bit to n_steps calls
 * of r2d3c_step.  forcing: array of n_steps draws (one per step, NULL = no
 * forcing); d_dt: device array of n_steps doubles receiving each step's dt
 * (may be NULL).  Consecutive steps share work: the last stage of a step also
 * produces the next step's column transforms, and after the first step only
 * the dealiased disc of the state is written back (the rest is already
 * zero).  No reference counterpart: the reference loops over single steps
 * (src/Rot2D3C.para.f90:393-615); use this when no diagnostics or output
 * fall between the steps. */
int r2d3c_run(r2d3c_plan *plan, const r2d3c_params *p, double *psi, double *vz,
              const r2d3c_forcing *forcing, int n_steps, int dt_mode, double dt_or_cfl,
              double *d_dt, void *stream);

/* r2d3c_run without a plan and without any device allocation, for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  Same
 * arguments and results as r2d3c_run on a plan of size n with the default
 * schedules (the same kernels, bit for bit), except: 16 <= n <=
 * R2D3C_NOALLOC_MAX_N, and iflow = 2 (ci_enabled) is rejected.  The
 * workspace is a static device array of about 45 MB that every device the
 * library runs on holds until the process ends.  The first call on a device
 * looks the array up and copies the twiddle tables into it synchronously, so
 * make it outside stream capture; later calls only launch kernels (a call
 * whose n differs from the previous call's on that device also clears the
 * fields it uses with a fill kernel).  All calls on one device share the
 * array: calls that overlap in time on one device (different streams or
 * host threads without ordering) race and give wrong results. */
#define R2D3C_NOALLOC_MAX_N 1024
int r2d3c_run_noalloc(int n, const r2d3c_params *p, double *psi, double *vz,
                      const r2d3c_forcing *forcing, int n_steps, int dt_mode, double dt_or_cfl,
                      double *d_dt, void *stream);

/* Dense forcing arrays fp (kin = 1) and fz (kin = 0) of one draw for a step
 * of length dt (host value), as rand_force fills them
 * (src/pseudospec2D_Rot2D3C.f90:1100-1122).  Either output may be NULL. */
int r2d3c_forcing_fill(r2d3c_plan *plan, const r2d3c_forcing *forcing, double dt,
                       double *fp, double *fz, void *stream);

/* ---- kernel timing of r2d3c_step (no reference counterpart) ------------ */

/* Kernel classes of r2d3c_step: 0 column inverse of the state, 1 row
 * (brackets) kernel, 2 fused column kernel of stages 1-3, 3 fused column
 * kernel of the last stage (state write). */
#define R2D3C_NKERNELS 4
/* While enabled, every kernel launched by r2d3c_step is bracketed by HIP
 * events stamped at the kernel's start and end.  Enabling or disabling
 * discards what was recorded. */
int r2d3c_timing_enable(r2d3c_plan *plan, int enable);
/* Waits for the recorded events, returns the summed milliseconds and the
 * launch count per kernel class (arrays of R2D3C_NKERNELS), and clears them. */
int r2d3c_timing_read(r2d3c_plan *plan, double *ms, long long *launches);

/* Streaming device copy of `bytes` (a multiple of 16; 16-byte aligned
 * pointers) from src to dst on
 * `stream`, the bandwidth yardstick of the benchmark (no reference
 * counterpart): 2 x bytes of HBM traffic per call. */
int r2d3c_copy_probe(void *dst, const void *src, long long bytes, void *stream);

/* ---- host random numbers (no device work) ----------------------------- */

/* Park-Miller minimal standard with the reference's XOR mask; float32 value in
 * [-1, 1).  Replaces randu (src/pseudospec2D_mod.f90:52-71). */
float r2d3c_randu(int *idum);
/* The draw of rand_force: theta = randu(seed)*pi (advances *seed),
 * phase = randu(seed+1)*pi (does not), kx/ky = floor(kup*cos/sin(theta)+0.5)
 * (src/pseudospec2D_Rot2D3C.f90:1086-1091). */
void r2d3c_rand_force_draw(double kup, int *seed, double *kx, double *ky, double *phase);

/* ---- schedules and autotuning (no reference counterpart) --------------- */

/* Schedule of the row kernel of r2d3c_step / r2d3c_run and of ensembles for
 * n >= 256: blocks per CU 0 selects the non-persistent kernel (one block per
 * 2 rows; rows must be 2); 1..4 select the persistent kernel with that many
 * blocks per CU and 2 rows per block, or 1 when n >= 512.  Schedules that
 * differ only in blocks per CU run the same kernel and give bit-identical
 * results; the three kernels (non-persistent, persistent with 2 or 1 rows)
 * agree within rounding (the compiler may fuse their multiply-adds
 * differently).  Default 2, 0. */
int r2d3c_plan_set_schedule(r2d3c_plan *plan, int row_rows, int row_blocks_per_cu);
int r2d3c_plan_get_schedule(const r2d3c_plan *plan, int *row_rows, int *row_blocks_per_cu);
/* Radix schedule of the column transforms of r2d3c_step / r2d3c_run and of
 * ensembles for n >= 512: 0 (default) the four-step form, 512-point
 * wave-local sub-transforms and one radix n/512 exchange; 1 radix-8 Stockham
 * passes with a last radix 2, 4 or 8 (below 512 the two coincide).  The two
 * agree within rounding. */
int r2d3c_plan_set_column_radix(r2d3c_plan *plan, int form);
int r2d3c_plan_get_column_radix(const r2d3c_plan *plan, int *form);
/* Split emission of the fused column kernel in stages 1-3 of r2d3c_step /
 * r2d3c_run and of ensembles: each kx column's block is repeated for each
 * emitted field (up to 4), every repeat doing the forward transform and the
 * update and emitting one field, so a latency-bound launch (a small grid)
 * finishes sooner.  -1 (default): split when the launch holds fewer threads
 * than the GPU has SIMDs x 64; 0 never; 1 always.  Results are bit-identical
 * either way. */
int r2d3c_plan_set_column_split(r2d3c_plan *plan, int mode);
int r2d3c_plan_get_column_split(const r2d3c_plan *plan, int *mode);
/* Form of the non-persistent row kernel (schedule (2, 0)) for n <= 512, in
 * r2d3c_step / r2d3c_run and ensembles: 1 the three field pairs' inverse
 * transforms run in parallel by three thread groups of one block (a
 * latency-bound launch finishes sooner), 0 one after another; -1 (default)
 * automatic: parallel for n <= 512.  The two forms agree within rounding;
 * larger grids always use the sequential form. */
int r2d3c_plan_set_row_split(r2d3c_plan *plan, int mode);
int r2d3c_plan_get_row_split(const r2d3c_plan *plan, int *mode);
/* Form of the n = 16384 step (other sizes ignore it): 1 (default) the fused
 * bracket pass, where both Jacobians of a stage and their row transforms are
 * formed in one pass over the rows and the second one's column transform
 * carries the update; 0 stage by stage through six real derivative fields.
 * The two agree within rounding. */
int r2d3c_plan_set_bracket_pass(r2d3c_plan *plan, int fused);
int r2d3c_plan_get_bracket_pass(const r2d3c_plan *plan, int *fused);
/* Times every valid schedule (row kernel and launch shape x column radix
 * schedule) on scratch state (best of three trials of `reps` fixed-dt steps
 * each), keeps the fastest in the plan and returns its milliseconds per step (may be NULL).
 * Allocates and frees its scratch; synchronizes the stream. */
int r2d3c_autotune(r2d3c_plan *plan, int reps, double *ms_per_step, void *stream);

/* ---- ensembles of independent runs (no reference counterpart) ---------- */

/* One run of an ensemble: grid size, parameters, fixed time step and its
 * device state (caller-owned, reference layout, advanced in place). */
typedef struct r2d3c_member {
  int n;
  r2d3c_params params;
  double dt;
  double *psi;
  double *vz;
} r2d3c_member;
typedef struct r2d3c_ensemble r2d3c_ensemble;

/* Groups the runs by grid size and allocates each run's workspace.  Runs of
 * one size advance together: one launch per kernel and stage covers all of
 * them; each size runs on a stream of its own, replaying captured HIP
 * graphs of its steps (or the sizes follow each other in one graph per time
 * step, r2d3c_ensemble_set_concurrent).
 * Each run's result equals r2d3c_run of that run alone (fixed dt, same
 * forcing) bit for bit. */
int r2d3c_ensemble_create(r2d3c_ensemble **ens, const r2d3c_member *members, int count);
int r2d3c_ensemble_destroy(r2d3c_ensemble *ens);
/* r2d3c_plan_set_schedule for the runs of size n. */
int r2d3c_ensemble_set_schedule(r2d3c_ensemble *ens, int n, int row_rows, int row_blocks_per_cu);
/* r2d3c_plan_set_column_radix for the runs of size n. */
int r2d3c_ensemble_set_column_radix(r2d3c_ensemble *ens, int n, int form);
/* How the size classes of a time step are ordered: 0 one after another,
 * each class's launches having the whole GPU; 1 as concurrent branches of
 * the step graph; 2 (default) each class on a stream of its own for the
 * whole run call, replaying graphs of its own steps (classes then overlap on
 * the device; ensembles with a closure run, and the all-size stage launches,
 * take 0 instead).  Results are identical in every mode. */
int r2d3c_ensemble_set_concurrent(r2d3c_ensemble *ens, int concurrent);
/* How a time step's stages are launched: 0 (default) the per-size kernels
 * (a row and a fused column launch per size and stage); 1 one persistent
 * launch per RK4 stage for every run of every size, whose blocks, all
 * resident, walk a work-list of (size, run, tile) units: the row transforms
 * of every run, a grid barrier, then the column transforms and updates;
 * 2 the same work-list as one launch per phase (rows of every size, then
 * columns of every size).  Forms 1 and 2: sizes 64 to 1024, at most 5 sizes,
 * runs without closures; they set every size's schedule to their arithmetic
 * (sequential row kernel, radix-8 Stockham columns), so each run still
 * equals its solo r2d3c_run with that schedule bit for bit.  Synchronizes
 * the ensemble.  No reference counterpart (the reference runs one
 * simulation per job). */
int r2d3c_ensemble_set_stage_launch(r2d3c_ensemble *ens, int form);
/* The stage launch form and, for form 1, its grid size (blocks). */
int r2d3c_ensemble_get_stage_launch(const r2d3c_ensemble *ens, int *form, int *grid);
/* Waits for the ensemble's work; R2D3C_EFAIL if a grid barrier of the
 * persistent launch (form 1) timed out since the last call (not every block
 * was resident: the state advanced by that run call is invalid). */
int r2d3c_ensemble_sync(r2d3c_ensemble *ens);
/* r2d3c_plan_set_row_split for the runs of size n. */
int r2d3c_ensemble_set_row_split(r2d3c_ensemble *ens, int n, int mode);
/* r2d3c_plan_set_column_split for the runs of size n. */
int r2d3c_ensemble_set_column_split(r2d3c_ensemble *ens, int n, int mode);
/* Advances every run n_steps steps.  forcing: NULL, or an array of
 * n_steps*count forcing structs, step-major with the runs in creation order,
 * holding rand_force draws only; the table is uploaded before the steps, so the steps run without host
 * round trips.  Asynchronous with respect to `stream`. */
int r2d3c_ensemble_run(r2d3c_ensemble *ens, const r2d3c_forcing *forcing, int n_steps,
                       void *stream);

/* ---- learned closure (no reference counterpart) ------------------------ */

/* Coarse-graining: box filter of width `width` (transfer function
 * sinc(kx w/2) sinc(ky w/2); 0 = none) and truncation of a spectral field of
 * plan `fine` to the dealiased modes of plan `coarse` (k^2 <= (nc/3)^2),
 * coefficients rescaled to the coarse grid's unnormalised convention. */
int r2d3c_filter(r2d3c_plan *fine, r2d3c_plan *coarse, const double *in, double *out,
                 double width, void *stream);
/* Network inputs on plan p's grid: fp32 [n*n][3] = (laplacian of psi, vz, psi)
 * in real space (physical units) times scale[3] (host array). */
int r2d3c_closure_inputs(r2d3c_plan *plan, const double *psi, const double *vz,
                         const double *scale, float *out, void *stream);
/* Training targets: fp32 [n*n][2] = real-space (f5 - c5, f6 - c6) restricted to
 * the update mask and |k| >= kmin, times scale[2]; f5, f6 the coarse-grained
 * fine brackets poisson(psi, laplak2(psi)) and poisson(psi, vz), c5, c6 the
 * same brackets of the coarse-grained fields. */
int r2d3c_closure_targets(r2d3c_plan *plan, const double *f5, const double *c5, const double *f6,
                          const double *c6, const double *scale, double kmin, float *out,
                          void *stream);
/* Dense forcing of the closure from a prediction fp32 [n*n][2] (times
 * scale[2]): fp = -r2c(pred0)/k^2 and fz = r2c(pred1) on the update mask and
 * |k| >= kmin (the subgrid parts of -C5/k^2 and C6 in the stage update). */
int r2d3c_closure_forcing(r2d3c_plan *plan, const float *pred, const double *scale, double kmin,
                          double *fp, double *fz, void *stream);

/* The closure network: 6 periodic 3x3 conv layers, 3 -> 32 -> ... -> 32 -> 2
 * channels, GELU between layers, BF16 weights and activations, FP32
 * accumulation (MFMA), on batches of up to `batch` images of height x width
 * (width a multiple of 64).  Parameters: fp32, layer by layer W[9][co][ci] then
 * b[co] with ci, co padded to 16/32 (r2d3c_cnn_layer_offsets; the padding stays
 * zero).  Images: fp32 [batch][height][width][channels]. */
typedef struct r2d3c_cnn r2d3c_cnn;
int r2d3c_cnn_param_count(void);
int r2d3c_cnn_layer_offsets(int *w_offsets, int *b_offsets, int *cin, int *cout);
int r2d3c_cnn_create(r2d3c_cnn **net, int batch, int height, int width);
int r2d3c_cnn_destroy(r2d3c_cnn *net);
/* Copies device fp32 parameters in (and resets the optimiser) or out. */
int r2d3c_cnn_set_params(r2d3c_cnn *net, const float *d_params, void *stream);
int r2d3c_cnn_get_params(r2d3c_cnn *net, float *d_params, void *stream);
/* out [batch][h][w][2] = network(in [batch][h][w][3]). */
int r2d3c_cnn_forward(r2d3c_cnn *net, const float *d_in, int batch, float *d_out, void *stream);
/* The network's own input and output buffers, for callers that fill and read
 * them in place: input bf16 [batch][h][w][16] (channels 0..2 used, 3..15
 * zero) and the last layer's fp32 [batch][h][w][16] (channels 0, 1 are the
 * prediction); r2d3c_cnn_forward_buffers runs the layers on them. */
int r2d3c_cnn_buffers(r2d3c_cnn *net, void **input_bf16, const float **output_f32);
int r2d3c_cnn_shape(const r2d3c_cnn *net, int *batch, int *height, int *width);
int r2d3c_cnn_forward_buffers(r2d3c_cnn *net, int batch, void *stream);
/* Inference calls (r2d3c_cnn_forward, r2d3c_cnn_forward_buffers) saturate
 * the outputs to [-clamp, clamp] when clamp > 0 (0, the default: none), which
 * bounds the network's response to inputs far outside its training data;
 * gradients and training steps use the unsaturated network. */
int r2d3c_cnn_set_output_clamp(r2d3c_cnn *net, double clamp);
/* Inference sign clipping of strength s in [0, 1] (0, the default: off):
 * output channels 0 and 1 are multiplied by 1 - s wherever they have the sign
 * of input channels 0 and 1 (after the clamp); at s = 1 the prediction can
 * only damp those input fields pointwise.  The closure uses it to limit local
 * production of vorticity^2 and vz^2.  Gradients and training steps are
 * unaffected. */
int r2d3c_cnn_set_sign_clip(r2d3c_cnn *net, double strength);
/* Gain per shell of the closure forcing made from this network's prediction
 * (r2d3c_closure_net_forcing, and the ensemble's closure runs): the forcing at
 * a mode of shell int(|k| + 0.5) < nshells is multiplied by gain[shell]
 * (0 <= gain <= 16); other shells keep gain 1.  gain = NULL or nshells = 0
 * removes the table (the default).  nshells <= max(height, width)/2 + 2.
 * The first call allocates the table's device memory and the copy is
 * synchronous; set it before r2d3c_ensemble_set_closure, whose step graph
 * keeps the table size it was captured with. */
int r2d3c_cnn_set_forcing_gain(r2d3c_cnn *net, const double *gain, int nshells);
/* Forward, MSE loss (mean over pixels and both output channels, in d_loss),
 * backward; the parameter gradient goes to d_grad when it is not NULL. */
int r2d3c_cnn_gradients(r2d3c_cnn *net, const float *d_in, const float *d_target, int batch,
                        float *d_grad, double *d_loss, void *stream);
/* r2d3c_cnn_gradients followed by one Adam step (beta 0.9, 0.999, eps 1e-8). */
int r2d3c_cnn_train_step(r2d3c_cnn *net, const float *d_in, const float *d_target, int batch,
                         float lr, double *d_loss, void *stream);

/* The closure's input and forcing steps fused around a network's own
 * buffers (see r2d3c_cnn_buffers): r2d3c_closure_net_inputs writes the
 * network input of (psi, vz) (as r2d3c_closure_inputs, rounded to bf16);
 * after r2d3c_cnn_forward_buffers, r2d3c_closure_net_forcing turns the
 * prediction into fp, fz (as r2d3c_closure_forcing).  n >= 64. */
int r2d3c_closure_net_inputs(r2d3c_plan *plan, const double *psi, const double *vz,
                             const double *scale, r2d3c_cnn *net, void *stream);
int r2d3c_closure_net_forcing(r2d3c_plan *plan, r2d3c_cnn *net, const double *scale,
                              double kmin, double *fp, double *fz, void *stream);
/* Backscatter clipping of a closure forcing, in place, on the modes
 * |k| >= kclip (kclip > 0): at each such mode where fp injects energy into psi
 * (Re(conj(psi) fp) > 0) its component along psi is removed,
 * fp -= Re(conj(psi) fp)/|psi|^2 psi, and likewise fz against vz, so that no
 * such mode gains energy from the closure. */
int r2d3c_closure_clip(r2d3c_plan *plan, const double *psi, const double *vz, double kclip,
                       double *fp, double *fz, void *stream);

/* Makes run `run` of an ensemble a closure run: each of its steps starts with
 * r2d3c_closure_net_inputs(scale_in) -> r2d3c_cnn_forward_buffers ->
 * r2d3c_closure_net_forcing(scale_out, kmin) (then r2d3c_closure_clip(kclip)
 * if kclip > 0) on the step's state, and the stages add that forcing (held for
 * the four stages), all inside the step's graph.  net = NULL removes the
 * closure (the plain solver).  The net must take 1 image of n x n, serve
 * no other run (its buffers hold this run's activations) and stay alive while
 * the ensemble uses it. */
/* The closure runs' forcing is evaluated on every `period`-th step of each
 * r2d3c_ensemble_run call (its first step, then steps period, 2 period, ...)
 * and held in between.  Default 1 (every step). */
int r2d3c_ensemble_set_closure_period(r2d3c_ensemble *ens, int period);
int r2d3c_ensemble_set_closure(r2d3c_ensemble *ens, int run, r2d3c_cnn *net,
                               const double *scale_in, const double *scale_out, double kmin,
                               double kclip);

#ifdef __cplusplus
}
#endif
#endif /* ROT2D3C_H */
