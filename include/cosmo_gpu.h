/*
 * cosmo_gpu.h -- C ABI of the MI355X BSSN hot path (libcosmo_gpu.so).
 *
 * The reference (CosmoGRaPH unigrid) has no FFI: its sims call C++ methods
 * of BSSN, RK4Register, Static, Dust and Lambda in-process.  Each entry point
 * below replaces one of those methods; the reference method it stands in
 * for is cited beside it.  Plain C: opaque handle, int status (0 = ok, < 0 =
 * error, message from cg_last_error), host buffers borrowed for the call
 * only, device memory owned by the context.  One host thread per context;
 * work is asynchronous on the context's HIP stream, and the calls that
 * return data to the host synchronise.
 *
 * Field names are the reference's field-map keys (cosmo_macros.h:154-187,
 * bssn_macros.h:77-112): "<field>_<p|a|c|f>" for the RK4 registers of
 * DIFFgamma11..DIFFgamma33, DIFFphi, A11..A33, DIFFK, Gamma1..3, DIFFalpha;
 * "DIFFD_a" for the static-dust density (components/static/static.h:67).
 * Grid layout is the reference's: index (i*ny + j)*nz + k, periodic.
 */
#ifndef COSMO_GPU_H
#define COSMO_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cg_ctx cg_ctx;

enum { CG_OK = 0, CG_ERR_ARG = -1, CG_ERR_HIP = -2, CG_ERR_STATE = -3, CG_ERR_NAN = -10 };
/* CG_FP32_SPLIT: A_ij, K, Gamma^i, alpha, D in FP32, the conformal metric
 * (DIFFgamma_ij, DIFFphi) in FP64.  CG_FP32_ALL: every BSSN field in FP32,
 * the metric included (BSSN + static dust or vacuum, single grids).
 * CG_FP32_FIX: as CG_FP32_ALL, but the conformal metric in 32-bit fixed
 * point (value = word x 2^e, one power-of-two scale per field, raised by
 * cg_rk4_steps as the field grows) and differentiated in FP64; steps run
 * through cg_rk4_steps only (cg_rk_evolve / cg_k_finalize return
 * CG_ERR_STATE: a right-hand side does not fit the field's scale).  All
 * keep the RK4 accumulation, the FRW background and the reductions in FP64. */
enum { CG_FP64 = 0, CG_FP32_SPLIT = 1, CG_FP32_ALL = 2, CG_FP32_FIX = 3 };
enum { CG_MATTER_NONE = 0, CG_MATTER_STATIC = 1, CG_MATTER_DUST = 2 };
/* the lapse gauges of BSSNGaugeHandler (BSSNGaugeHandler.h:96-111) */
enum {
  CG_LAPSE_STATIC = 0, CG_LAPSE_HARMONIC = 1, CG_LAPSE_CONFORMAL_FLRW = 2, CG_LAPSE_ANHARMONIC = 3,
  CG_LAPSE_ONE_PLUS_LOG = 4, CG_LAPSE_CONFORMAL_SYNC = 5, CG_LAPSE_AWA_GAUGE_WAVE = 6,
  CG_LAPSE_AWA_SHIFTED_WAVE = 7, CG_LAPSE_DAMPED_WAVE = 8, CG_LAPSE_TEST_AIJ_DRIVER = 9,
  CG_LAPSE_TEST_K_DRIVER = 10
};
/* The shift (BSSNGaugeHandler.h:113-135).  CG_SHIFT_NONE is the reference's
 * default build (USE_BSSN_SHIFT false, cosmo_macros.h:43-44: beta = 0, no
 * shift registers).  Any other value is the USE_BSSN_SHIFT build with that
 * shift gauge: registers "beta1..3_<r>" and "expN_<r>" are added, and every
 * field is advected with the upwind stencils of utils/math.h:1220-1252
 * (identically zero at orders 6 and 8, where the reference's lopsided
 * stencils are stubs).  CG_SHIFT_GAMMA_DRIVER is also the USE_GAMMA_DRIVER
 * build (registers "auxB1..3_<r>", bssn_macros.h:38-47).  Not with the dust
 * fluid, nor in a batch. */
enum {
  CG_SHIFT_NONE = 0, CG_SHIFT_STATIC = 1, CG_SHIFT_GAMMA_DRIVER = 2, CG_SHIFT_DAMPED_WAVE = 3,
  CG_SHIFT_AWA_SHIFTED_WAVE = 4, CG_SHIFT_AIJ_DRIVER = 5
};

/* Construction parameters.  Mirrors the compile-time macros and config keys
 * the reference reads: COSMO_N/NX/NY/NZ and STENCIL_ORDER
 * (cosmo_macros.h:12-40), dx and dt_frac (cosmo.cc:41-42),
 * KO_damping_coefficient, a_adj_amp, k_damping_amp, normalize_metric
 * (bssn.cc:15-22), lapse (BSSNGaugeHandler.h:175), the matter type
 * (simulation_type, cosmo.cc:52-81) and Lambda (lambda.cc:16-24). */
typedef struct {
  int nx, ny, nz;
  int stencil_order;      /* 2, 4, 6, 8 */
  int precision;          /* CG_FP64 or CG_FP32_SPLIT */
  double dx, dt;
  double ko_coeff;
  double a_adj_amp;
  double k_damping_amp;
  int normalize_metric;
  int lapse;
  int matter;
  double lambda;
  int device;             /* HIP device ordinal */
  int tile_i, tile_j, tile_k; /* 0 = autotuned/default */
  int shift;              /* CG_SHIFT_* (config key `shift`, BSSNGaugeHandler.h:176) */
} cg_params;

/* BSSN::BSSN(ConfigParser*, Fourier*) (components/bssn/bssn.cc:13-44):
 * allocate all registers, zero-filled.  CG_ERR_ARG for a grid of more than
 * 2^29 points (about 812^3): the kernels address each array with a 32-bit
 * byte offset, and the FP64 registers have 8-byte elements. */
int cg_create(const cg_params* params, cg_ctx** out);
/* BSSN::~BSSN (bssn.cc:46-54) */
int cg_destroy(cg_ctx* ctx);
const char* cg_last_error(const cg_ctx* ctx);

/* Writes / reads through the reference's `fields` map (bssn.h:43):
 * n must equal nx*ny*nz.  FP32-split contexts convert at the boundary. */
int cg_field_upload(cg_ctx* ctx, const char* name, const double* host, size_t n);
int cg_field_download(cg_ctx* ctx, const char* name, double* host, size_t n);

/* FRW<real_t> reference integrator owned by BSSN (bssn.h:46, utils/FRW.h):
 * set_phi/set_K (FRW.h:48-55), addFluid (FRW.h:60-76), getters (FRW.h:79-83). */
int cg_frw_set(cg_ctx* ctx, double phi, double K);
int cg_frw_add_fluid(cg_ctx* ctx, double rho, double w);
int cg_frw_get(const cg_ctx* ctx, double out_phi_K_rho_S[4]);
/* Lambda::setLambda (components/Lambda/lambda.cc:16-19) */
int cg_set_lambda(cg_ctx* ctx, double lambda);
/* Gauge parameters of BSSNGaugeHandler::_initDefaultParameters
 * (BSSNGaugeHandler.h:139-149): gd_c (OnePlusLog, default 1), dw_mu_l and
 * dw_p (DampedWave, default 0). */
int cg_set_gauge_params(cg_ctx* ctx, double gd_c, double dw_mu_l, double dw_p);
/* The remaining gauge parameters: gauge_wave_dir (AwAShiftedWave shift,
 * default 1), gd_eta (Gamma driver, bssn.cc:18, default 0), dw_mu_s
 * (DampedWave shift, default 0) and k_driver_coeff (TestKDriverLapse and
 * AijDriverShift, default 0.04), BSSNGaugeHandler.h:139-149. */
int cg_set_shift_params(cg_ctx* ctx, int gauge_wave_dir, double gd_eta, double dw_mu_s, double k_driver_coeff);

/* Method-level drop-ins (unfused; each streams the grid once):
 * BSSN::stepInit (bssn.cc:219-226), BSSN::RKEvolve with the matter source
 * of the context folded in (bssn.cc:231-243 after clearSrc/addBSSNSrc,
 * static.cc:19-36, lambda.cc:26-43), BSSN::K{n}Finalize (bssn.cc:267-305). */
int cg_step_init(cg_ctx* ctx);
int cg_rk_evolve(cg_ctx* ctx);
int cg_k_finalize(cg_ctx* ctx, int n);

/* One or more whole RK4 steps in the StaticSim sequence
 * (sims/static.cc:74-170: initStaticStep + runStaticStep, without output),
 * fused into four stage kernels per step.  Returns CG_ERR_NAN if DIFFphi
 * turned NaN (the guard of sims/sim.cc:179-183). */
int cg_rk4_steps(cg_ctx* ctx, int nsteps);

/* BSSN::setConstraintCalcs (bssn.cc:965-1071) on the _a register as the
 * reference's output step sees it: mean, stdev, max, sum of scales, mean
 * scaled, stdev scaled, max scaled of the Hamiltonian (H), momentum (M),
 * Christoffel (G), trace-free A (A) and unit-determinant (S) constraints;
 * H[7], M[7] = sqrt(sum C^2 dx^3).  FP64 arithmetic and deterministic FP64
 * reductions in every precision.  Also fills "ricci_a" and "AijAij_a" for
 * cg_field_download (prepBSSNOutput, sims/sim.cc:214-231), and the sources
 * of the evaluated state as "DIFFr_a" / "DIFFS_a" (what initStaticStep and
 * initDustStep leave there before the output step, sims/static.cc:74-82,
 * sims/dust.cc:53-63; with the dust fluid "DIFFr_a" otherwise holds the last
 * stage's source). */
int cg_constraints(cg_ctx* ctx, double H[8], double M[8], double G[7], double A[7], double S[7]);

/* Bardeen::setPotentials + getSVTViolations (components/bssn/bardeen.cc:22-467)
 * on the state cg_constraints reads (prepBSSNOutput, sims/sim.cc:214-231):
 * the BSSN right-hand side there (left in the _c registers, as RKEvolve
 * leaves it), the metric perturbation and its time derivatives, the SVT
 * potentials through inverse Laplacians, the Bardeen potentials, and the 17
 * linearised-constraint violations io_svt_violation writes (IO/io.cc:1116-1131)
 * into viols.  Afterwards "Bardeen_Phi", "Bardeen_Psi", "Bardeen_A",
 * "Bardeen_dt_A", "Bardeen_d2t_A", "Bardeen_B", "Bardeen_dt_B",
 * "Bardeen_d2t_B", "Bardeen_E", "Bardeen_F", "Bardeen_G1..3",
 * "Bardeen_C1..3", "Bardeen_dt_C1..3", "Bardeen_Vmag" and "Bardeen_D11..33"
 * are readable with cg_field_download (bardeen.h:127-156).  elapsed_t is the
 * simulation time; use_ml_scale_factor and omega_L are the config keys
 * use_ML_scale_factor and Omega_L (sims/sim.cc:82-95).  CG_ERR_STATE with
 * an FRW reference in use (K or phi non-zero), where the reference declines
 * to compute them (bardeen.cc:26-31); nz must be even. */
int cg_bardeen(cg_ctx* ctx, double elapsed_t, int use_ml_scale_factor, double omega_L, double viols[17]);

/* cg_constraints without any allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * constraint kernel reduces each workgroup's points into one row of a
 * fixed device array (0.98 MB per device, 4096 rows: the grid's tiles are
 * evaluated in chunks of 4096, each folded into running totals in a fixed
 * order), in two passes (sums and maxima, then the squared deviations from
 * the means for the N-1 standard deviations), instead of the 22 per-point
 * FP64 arrays cg_constraints allocates on first use (2.9 GB at 256^3).  Any
 * grid cg_create accepts fits.  Same outputs as cg_constraints to FP64
 * rounding (different summation order; maxima equal); "ricci_a",
 * "AijAij_a", "DIFFS_a" and the constraint strips are not filled.  The array
 * is shared by every context on a device: calls that overlap on one device
 * race. */
int cg_constraints_noalloc(cg_ctx* ctx, double H[8], double M[8], double G[7], double A[7], double S[7]);

/* BSSN::set1DConstraintOutput (bssn.cc:903-938), read by the
 * 1D_hamiltonian_constraints / 1D_momentum_constraints strips of
 * io_bssn_constraint_violation (IO/io.cc:186-235): H and the momentum
 * constraint component M_axis at n = N_axis points.  As in the reference the
 * running index is the i index for every axis (points (x mod NX, n1, n2));
 * offsets wrap periodically.  Evaluates the state as cg_constraints does. */
int cg_constraint_strip(cg_ctx* ctx, int axis, long n1, long n2, double* H, double* M, size_t n);

/* The values io_dump_strip writes (IO/io.cc:719-771): field(i, n1, n2),
 * field(n1, j, n2) or field(n1, n2, k) for axis 1, 2, 3, n = N_axis values,
 * offsets wrapped as CosmoArray::idx wraps them (utils/Array.h:149-159).
 * Any name cg_field_download accepts. */
int cg_field_strip(cg_ctx* ctx, const char* name, int axis, long n1, long n2, double* out, size_t n);

/* The DETAILS statistics of io_bssn_dump_statistics (IO/io.cc:8-14,
 * 341-403): out = {conformal_average, conformal_standard_deviation,
 * volume_average} (utils/math.h:1382-1415, 1458-1470) of a field-map array,
 * weighted by exp(6 (DIFFphi_a + phi_FRW)) with the current FRW phi.  FP64
 * sums, reduced in a fixed order (deterministic). */
int cg_field_stats(cg_ctx* ctx, const char* name, double out[3]);

/* Fourier::powerDump (utils/Fourier.h:147-240), written by
 * io_bssn_fields_powerdump (IO/io.cc:165-178) for DIFFphi_a and DIFFr_a:
 * the r2c transform of a field-map array (hipFFT, FP64), |F_k|^2 averaged in
 * bins of floor(|k|) (k in lattice units, each interior kz mode counted for
 * its conjugate too).  out[nbins], nbins = (int)(sqrt(NX^2+NY^2+NZ^2)/2) + 1;
 * empty bins are 0.  The per-bin sums use LDS atomics, so their last bits
 * may differ between runs.  Needs an even NZ. */
int cg_power_spectrum(cg_ctx* ctx, const char* name, double* out, int nbins);

/* BSSN::setExtraFieldData results (bssn.cc:165-170; the public members
 * K_min, K_avg, avg_vol, rho_avg of bssn.h:48): out = {K_min, K_avg,
 * avg_vol, rho_avg} as of the last stepInit / K{n}Finalize, i.e. after a
 * fused step those of its K4Finalize.  K_avg is the conformal average of
 * DIFFK (the perturbation, math.h:1403-1415), which the ConformalFLRW lapse
 * reads on the device every stage.  rho_avg averages the source DIFFr of the
 * stage that produced the state; a stepInit with no stage since creation
 * keeps the previous value (0 initially, like a zeroed DIFFr_a). */
int cg_extra_field_data(cg_ctx* ctx, double out[4]);

/* numNaNs (utils/math.h:1524-1541) of one field-map array: number of points
 * whose float-rounded value is NaN (CosmoSim::simNumNaNs, sims/sim.cc). */
int cg_num_nans(cg_ctx* ctx, const char* name, long long* count);

/* Static-dust initial data on the device (hipFFT), this build's recipe (the
 * reference's static ICs go from a chosen phi to rho, static_ic.cc:185-199,
 * and its constraint solver is an absent submodule); the host version is
 * cosmograph_amd/ics.py.  `modes` is the half-spectrum [nx][ny][nz/2+1] of
 * complex amplitudes (re, im interleaved; n_modes = 2*nx*ny*(nz/2+1)), as the
 * reference's set_gaussian_random_Phi_N fills its Fourier field
 * (ICs/ICs.cc:38-134).  delta = its inverse real FFT (the kz = 0 and nz/2
 * planes taken Hermitian), mean removed, scaled to `rms`; drho = rho_bar delta
 * (rho_bar = 3/(8 pi)).  lap(u) = -2 pi u^5 (drho - c) with <u^5 (drho - c)> = 0
 * is solved by spectral fixed-point iteration until max|du| < tol (at most
 * max_iter).  Writes DIFFphi_p = log u, DIFFD_a = e^{6 phi}(rho_bar + drho - c)
 * - rho_bar, zero in every other _p register, and the FRW background phi = 0,
 * K = -sqrt(24 pi rho_bar), one w = 0 fluid of rho_bar (static_ic.cc:372-378).
 * A dust-fluid context gets the same geometry as full fields with an empty
 * FRW reference, as the reference's dust ICs leave it (dust_ic.cc): dust_D_p
 * = e^{6 phi}(rho_bar + drho - c), S_i = 0, DIFFK_p = -sqrt(24 pi (rho_bar +
 * lambda)) with the Lambda energy density set by cg_set_lambda beforehand;
 * with cg_set_dust_background on, DIFFK_p = 0 and the FRW reference holds
 * phi = 0, that K and the fluids rho_bar (w = 0) and lambda (w = -1).
 * shift and iters (optional) receive c and the iteration count. */
int cg_ics_static_dust(cg_ctx* ctx, const double* modes, size_t n_modes, double rms, double tol, int max_iter,
                       double* shift, int* iters);

/* Dust fluid as perturbations (on = 1; default 0 = the reference's form).
 * The reference's dust sources (dust.cc:226-264) are the full density and
 * stress with an empty FRW reference (dust_ic.cc), so the evolved DIFFK is
 * the full K.  With the background on, the dust's DIFFr and DIFFS are taken
 * minus the FRW totals rho_F and S_F: the FRW fluids (e.g. the mean dust,
 * w = 0, and Lambda as a w = -1 fluid of the context's Lambda density) carry
 * the mean, and DIFFK, DIFFr stay perturbations, as FP32 storage needs.
 * With an empty FRW both forms are the same.  No counterpart in the
 * reference. */
int cg_set_dust_background(cg_ctx* ctx, int on);

/* Synchronise the context's stream. */
int cg_sync(cg_ctx* ctx);

/* Kernel resource usage of the stage kernel this context launches. */
int cg_kernel_info(const cg_ctx* ctx, int* vgprs, int* scratch_bytes, int* lds_bytes,
                   int* tile_i, int* tile_j, int* tile_k);

/* Project autotuner: times every compiled stage-kernel variant (tile shape x
 * register budget) that fits this grid on the context's own data and keeps
 * the fastest.  The evolved state is untouched; must be called between whole
 * steps.  `report` (optional) receives a JSON list of the timings. */
int cg_autotune(cg_ctx* ctx, int reps, char* report, size_t report_len);

/* Halo staging of the stage kernel.  on = 1 (default): a launch without
 * stage-1 metric normalisation, on a grid with nz % 4 == 0, uses the
 * kernel's LDS-DMA twin where one is compiled: halo tiles go from memory
 * straight into LDS (global_load_lds), two fields ahead of the one being
 * differentiated; the autotuner times both paths and keeps the faster.
 * on = 0: register staging everywhere.  The two paths agree to rounding
 * (the kernels may contract a few multiply-adds differently).
 * No counterpart in the reference (its stencils read memory directly). */
int cg_set_dma(cg_ctx* ctx, int on);

/* Tuning knob, no counterpart in the reference.  keep = 0 (default): fused
 * steps (cg_rk4_steps, and cg_batch_step of the members) may run RK4 without
 * the _f accumulator: stages 1-3 write only the next stage input and stage 4
 * rebuilds K3Finalize's accumulator from the three stage outputs
 * ((a1-p) + 2(a2-p) + (a3-p))/3, 14 register sweeps per step instead of the
 * 17 of RK4Register.h:118-176 (the dust fluid's registers included).  The
 * result differs from the reference's accumulation by rounding (about one
 * unit in the last place of _p per step).  The autotuner times both forms
 * and keeps the faster; without it the accumulator-free form runs.  KO at
 * orders 2 and 8 with metric normalisation always uses the reference's form.
 * keep = 1: the reference's accumulator. */
int cg_set_fused_accumulator(cg_ctx* ctx, int keep);

/* Traversal order of the stage kernel's tiles (the autotuner sets it):
 * 0, 1, 2 = the j, k, i axis slowest; 3 = k fastest, then the (i, j) tile
 * columns in Morton order (where the tile counts along i and j are powers
 * of two; order 0 otherwise); 4, 5 = k fastest, then strips of 2 or 4 tile
 * rows along j, traversed along i.  Results do not depend on it.  No
 * counterpart in the reference. */
int cg_set_tile_order(cg_ctx* ctx, int order);

/* Average duration of the last `n` stage-kernel launches recorded with HIP
 * events on the context's stream (ms); enable with cg_set_timing(ctx, 1). */
int cg_set_timing(cg_ctx* ctx, int on);
int cg_stage_time_ms(cg_ctx* ctx, double* avg_ms, int* count);

/* ------------------------------------------------------------------------
 * Batched ensemble.  The reference runs one realisation per process
 * (sims/sim.cc:101-131); a batch advances many independent grids together:
 * one stage-kernel launch per RK4 stage covers the tiles of every grid, and
 * the launches of a call are captured in a HIP graph and replayed.  Every
 * realisation is an ordinary context (cg_batch_member) on which the
 * single-grid entry points above work: field uploads/downloads, FRW set-up,
 * constraints, extra field data, numNaNs.  Member handles are owned by the
 * batch (never cg_destroy them).  All realisations share precision, stencil
 * order, device and whether KO damping is on; grid sizes, dx, dt, lambda,
 * FRW backgrounds and damping amplitudes may differ.  The ConformalFLRW
 * lapse (a per-grid reduction every stage) and KO at orders 2/8 with metric
 * normalisation are single-grid only (CG_ERR_ARG here).
 * ---------------------------------------------------------------------- */
typedef struct cg_batch cg_batch;

/* n realisations with params[0..n-1]; on failure *out may still hold a batch
 * for cg_batch_last_error, to be released with cg_batch_destroy. */
int cg_batch_create(const cg_params* params, int n, cg_batch** out);
int cg_batch_destroy(cg_batch* batch);
const char* cg_batch_last_error(const cg_batch* batch);
int cg_batch_size(const cg_batch* batch);
int cg_batch_member(cg_batch* batch, int r, cg_ctx** out);

/* nsteps whole RK4 steps of every realisation (each as cg_rk4_steps).
 * Returns CG_ERR_NAN if any realisation's DIFFphi turned NaN; the message
 * lists them.  Work queued on member streams is waited for first. */
int cg_batch_step(cg_batch* batch, int nsteps);
/* Capture the launches of a call in a HIP graph and replay it (default 1),
 * or launch them one by one (0). */
int cg_batch_set_graph(cg_batch* batch, int on);
/* GPU time of the last cg_batch_step call (HIP events around the launches). */
int cg_batch_step_time_ms(const cg_batch* batch, double* ms);
/* cg_constraints of every realisation: H, M [n][8]; G, A, S [n][7]. */
int cg_batch_constraints(cg_batch* batch, double* H, double* M, double* G, double* A, double* S);
/* Times every compiled batched variant that fits all grids and keeps the
 * fastest (state untouched; between whole steps). */
int cg_batch_autotune(cg_batch* batch, int reps, char* report, size_t report_len);
/* cg_set_dma for the batched kernel (drops a captured graph). */
int cg_batch_set_dma(cg_batch* batch, int on);
int cg_batch_kernel_info(const cg_batch* batch, int* vgprs, int* scratch_bytes, int* tile_i, int* tile_j,
                         int* tile_k);

#ifdef __cplusplus
}
#endif
#endif
