// bssn_stage.h -- shared declarations between the host runtime and the
// fused BSSN RK4 stage kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

namespace cg {

constexpr int NF = 18;  // evolved fields, reference order (bssn_macros.h:77-99)
enum Field {
  G11 = 0, G12, G13, G22, G23, G33, PHI, A11, A12, A13, A22, A23, A33, KK, GAM1, GAM2, GAM3, ALPHA
};
enum Matter { MATTER_NONE = 0, MATTER_STATIC = 1, MATTER_DUST = 2 };
constexpr int NDUST = 4;     // D, S1, S2, S3
constexpr int NFLUX = 12;    // alpha D v^i (3), alpha S_j v^i (9)
// stage-kernel modes: bit 0 constraint diagnostics, bit 1 dust fluid
enum { MODE_STAGE = 0, MODE_CONS = 1, MODE_DUST = 2, MODE_DUST_CONS = 3 };
// bit 2: all-FP32 storage (the conformal metric in FP32 as well); bit 3
// (with bit 2): the conformal metric in 32-bit fixed point instead, value =
// integer word x 2^e with one power-of-two scale per field
enum { MODE_M32 = 4, MODE_M32_CONS = 5, MODE_MQ = 8, MODE_MQ_STAGE = 12, MODE_MQ_CONS = 13 };
// fixed-point metric word (MODE_MQ): value / 2^e rounded to the nearest
// integer, saturated at MQ_WMAX = 2^30 - 1 in magnitude.  After each
// exponent check the host keeps |word| below 2^MQ_BITS = 2^29, so a field
// can double before the next check; saturation means it grew faster, and
// the stage kernels count it.  With |word| < 2^30 every first difference
// of two words fits in int32, and the second and mixed differences do for
// any field resolved finely enough to be differenced (the stencils form
// them in wrap-around arithmetic, exact whenever the result fits).
constexpr double MQ_WMAX = 1073741823.0;
__device__ __forceinline__ int mq_word(double v, double inv) {
  const double t = fmin(fmax(v * inv, -MQ_WMAX), MQ_WMAX);
  return (int)rint(t);
}
constexpr int MQ_BITS = 29;        // words stay below 2^MQ_BITS after each exponent check
constexpr int MQ_FLOOR = 6;        // exponents stay within 6 of the largest metric field's

enum Lapse {
  LAPSE_STATIC = 0, LAPSE_HARMONIC = 1, LAPSE_CONFORMAL_FLRW = 2, LAPSE_ANHARMONIC = 3, LAPSE_ONE_PLUS_LOG = 4,
  LAPSE_CONFORMAL_SYNC = 5, LAPSE_AWA_GAUGE_WAVE = 6, LAPSE_AWA_SHIFTED_WAVE = 7, LAPSE_DAMPED_WAVE = 8,
  LAPSE_TEST_AIJ_DRIVER = 9, LAPSE_TEST_K_DRIVER = 10, LAPSE_N = 11
};
// shift gauges (BSSNGaugeHandler.h:113-135); SHIFT_NONE: no shift registers
enum Shift {
  SHIFT_NONE = 0, SHIFT_STATIC = 1, SHIFT_GAMMA_DRIVER = 2, SHIFT_DAMPED_WAVE = 3, SHIFT_AWA_SHIFTED_WAVE = 4,
  SHIFT_AIJ_DRIVER = 5, SHIFT_N = 6
};
// registers of the shift build: beta1..3, expN, auxB1..3 (bssn_macros.h:18-47)
constexpr int NXF = 7;
enum XField { XBETA1 = 0, XEXPN = 3, XAUXB1 = 4 };
constexpr int n_xfields(int shift) { return shift == SHIFT_NONE ? 0 : shift == SHIFT_GAMMA_DRIVER ? NXF : XAUXB1; }
// gauges that read K_avg of setExtraFieldData every stage
__host__ __device__ constexpr bool lapse_needs_kavg(int l) {
  return l == LAPSE_CONFORMAL_FLRW || l == LAPSE_ANHARMONIC || l == LAPSE_ONE_PLUS_LOG;
}

// One fused RK4 stage over the whole grid: stencils of the `in` register,
// BSSN right-hand side, matter sources, and the RK4Register update
// (utils/RK4Register.h:118-176) in the epilogue.  Register pointers are
// resolved on the host; the DIFFphi registers are always FP64 (see
// bssn_stage.hip), every other field uses the context's storage type.
struct StageArgs {
  // Registers: field f, slot s lives at met + (f*4 + s)*rstride (f < 7,
  // DIFFgamma_ij and DIFFphi, FP64) or regs + ((f-7)*4 + s)*rstride (storage
  // type).  The slots of this stage: stencil input (_a), _p (step start), _f
  // accumulator, output (next _a, or _p at stage 4, or _c at stage 0).
  double* met;
  void* regs;
  long rstride;
  int s_in, s_p, s_f, s_out;
  // Stage 1 with the metric normalisation precomputed: DIFFgamma_ij and A_ij
  // are read from slot s_in_norm, the other fields from s_in (-1: off).
  // Stage 4: also write set_DIFFgamma_Aij_norm of the new state
  // (bssn.cc:61-119) into slot s_norm_out for the next step's stage 1
  // (-1: off).
  int s_in_norm, s_norm_out;
  // Accumulator-free RK4 (fused steps): stages 1-3 write only the next
  // input a_n = p + ca*c (no _f); stage 4 rebuilds the accumulator of
  // K3Finalize from the three stage outputs, f = ((a1-p) + 2(a2-p) + (a3-p))/3
  // with a1 in slot s_f, a2 in slot s_a2 and a3 the stencil input, then
  // p' = f + (cf*c + p).  13 register sweeps per step instead of 17.
  int incr, s_a2;
  const void* DIFFD;      // static dust: conserved density perturbation
  // dust fluid (components/dust_fluid/dust.h:32-46): registers of D, S1..S3
  // at dreg + (q*4 + slot)*rstride (storage type); the 12 fluxes
  // alpha D v^i, alpha S_j v^i of the stencil input at flux + q*rstride;
  // the matter sources use W, detg, g_ij of the previous stage's input
  // (the reference's lag, dust.cc:226-264) read at slot s_lag, its
  // DIFFgamma_ij at slot s_lag_met (a normalised copy, when the previous
  // stage read one)
  void* dreg;
  const void* flux;
  int s_lag, s_lag_met;
  int lag_mode;           // 0: lag = stencil input; 1: slots s_lag / s_lag_met; 2: the same, metric normalised
  double phi_F_lag;       // FRW phi the lagged state was seen with
  double* rho_out;        // dust stage: DIFFr per point (null: not kept)
  int nx, ny, nz;
  int ntx, nty, ntz;      // tile counts per axis
  int per_xcd;            // tiles handled per XCD (XCD-aware block remap)
  int tile_order;         // traversal order of the tiles (0, 1, 2: j, k, i slowest; 3: Morton columns; 4, 5: j strips)
  int stage;              // 0: RHS only into out; 1..4: fused K{n}Finalize
  int normalize;          // stage 1: det(gamma)=1 rescale, trace-free A on the fly
  int matter;
  int lapse;
  int use_ko;
  int use_H;              // a_adj_amp or k_damping_amp non-zero
  double inv_dx, inv_dx2, dt;
  double cf, ca;          // epilogue: f += cf*c ; a' = p + ca*c  (stage 4: p' = f + (cf*c + p))
  double phi_F, K_F, rho_F, S_F;  // FRW values seen by this stage
  const double* ext_in;   // [K_min, K_avg, avg_vol, rho_avg] seen by this stage (device)
  double* ext_part;       // per-wave partial sums of setExtraFieldData over the output (null: off)
  double phi_F_next;      // FRW phi after this stage's FRW step (weights of the averages)
  double lambda;
  int dust_bg;            // dust fluid: its sources minus the FRW totals rho_F, S_F (perturbations)
  double ko_pref;         // KO prefactor (sign included), per unit stencil sum
  double gd_c, dw_mu_l, dw_p;  // gauge parameters
  double a_adj, k_damp;
  double H_frw;           // 2/3 K_F^2 - 16 pi rho_F (for the split-precision H)
  unsigned int* nan_count;
  double* cons;           // constraint mode: 17 per-point FP64 arrays (see CONS_*)
  long cons_stride;
  // Constraint mode without per-point arrays: with cons_rows set, every
  // block reduces its points into one row of cons_rows (row = tile -
  // tile_base; pass 0: CONS_ROW_W sums and maxima, pass 1: the 10 squared
  // deviations from cons_means).  tile_base/tile_count select a chunk of the
  // grid's tiles (tile_count 0: all of them).
  double* cons_rows;
  const double* cons_means;
  int cons_pass;
  int tile_base, tile_count;
  // fixed-point metric (MODE_MQ): scale 2^e and its inverse of field f < 7,
  // the scale times 1/dx and 1/dx^2 rounded to FP32 (the stencils' factors),
  // and the count of saturated words
  double msc[7], minv[7];
  float msd1[7], msd2[7];
  unsigned int* mq_sat;
};

// row of the reduced constraint mode, per constraint q (H, M, G, A, S):
// 6q + {sum val, sum scale, sum scaled, sum val^2, max|val|, max|scaled|};
// pass 1: 2q + {sum (val - mean)^2, sum (scaled - mean scaled)^2}
constexpr int CONS_ROW_W = 30;
constexpr int CONS_ROW_W1 = 10;
// the 5 constraints' value, scale, scaled of one point
constexpr int CONS_PT = 15;

// FRW values one stage sees (utils/FRW.h getters), the split-precision
// background part of H, and the FRW phi after the stage's FRW step
struct FrwV {
  double phi, K, rho, S, H, phi_next;
};
constexpr int FRW_N = 6;

// One launch over the tiles of many grids (batched ensemble).
struct BatchArgs {
  const StageArgs* __restrict__ tab;  // [nreal] arguments of each grid for this stage
  const int* __restrict__ tstart;     // [nreal] first global tile of each grid
  const double* __restrict__ frw;     // [nreal][FRW_N] FRW values of this stage
  int nreal;
  int total;                          // tiles of all grids
  int per_xcd;                        // tiles per XCD (XCD-aware block remap)
};

// per-point outputs of the constraint mode (BSSN::setConstraintCalcs,
// bssn.cc:965-1071): value, scale and value/scale of H, M, G, A, S, then the
// ricci_a and AijAij_a arrays of prepBSSNOutput (sims/sim.cc:214-231), the
// three momentum-constraint components M_I (set1DConstraintOutput,
// bssn.cc:903-938) and the sources DIFFr, DIFFS of the evaluated state
enum { CONS_H = 0, CONS_M = 3, CONS_G = 6, CONS_A = 9, CONS_S = 12, CONS_RICCI = 15, CONS_AIJAIJ = 16,
       CONS_MI = 17, CONS_RHO = 20, CONS_SRCS = 21, CONS_N = 22 };
int launch_constraints(int prec, int order, bool live_alpha, bool dust, const StageArgs& a, hipStream_t s);

// BSSN::setExtraFieldData (bssn.cc:165-170) partial sums, one row per
// workgroup: sum w*DIFFK, sum w, sum w*DIFFphi, sum w*DIFFr, min DIFFK with
// w = exp(6 (DIFFphi + phi_FRW)) (utils/math.h:1382-1415, 1506-1518)
enum { EXT_SWK = 0, EXT_SW, EXT_SWPHI, EXT_SWR, EXT_KMIN, EXT_W };
enum { EXTV_KMIN = 0, EXTV_KAVG, EXTV_AVGVOL, EXTV_RHOAVG };

// deterministic wavefront sum / min (fixed butterfly order)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const double o = __shfl_xor(v, m, 64);
    v = o < v ? o : v;
  }
  return v;
}
// One partial row per workgroup (fixed order: butterfly within each wave,
// waves in index order).  Every thread of the block must call this: it holds
// a barrier.
__device__ __forceinline__ void block_ext_put(double* part, long row, bool on, double phi, double K, double rho,
                                              double phiF) {
  __shared__ double sh[16][EXT_W];
  const double w = on ? exp(6.0 * (phi + phiF)) : 0.0;
  const double s0 = wave_sum(on ? w * K : 0.0), s1 = wave_sum(w), s2 = wave_sum(on ? w * phi : 0.0),
               s3 = wave_sum(on ? w * rho : 0.0), mn = wave_min(on ? K : __builtin_huge_val());
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sh[wv][EXT_SWK] = s0; sh[wv][EXT_SW] = s1; sh[wv][EXT_SWPHI] = s2; sh[wv][EXT_SWR] = s3; sh[wv][EXT_KMIN] = mn;
  }
  __syncthreads();
  if (threadIdx.x < EXT_W) {
    const int q = threadIdx.x, nw = (int)(blockDim.x >> 6);
    double acc = sh[0][q];
    for (int k = 1; k < nw; ++k) acc = q == EXT_KMIN ? (sh[k][q] < acc ? sh[k][q] : acc) : acc + sh[k][q];
    part[row * EXT_W + q] = acc;
  }
}

// deterministic wavefront max (fixed butterfly order) of non-NaN values
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const double o = __shfl_xor(v, m, 64);
    v = o > v ? o : v;
  }
  return v;
}
// One row of the reduced constraint mode per workgroup (fixed order:
// butterfly within each wave, waves in index order).  cv holds the point's
// value, scale, scaled of H, M, G, A, S; `on` is false for threads outside
// the grid.  Every thread of the block must call this: it holds barriers.
__device__ __forceinline__ void block_cons_put(double* row, int pass, const double* means, bool on,
                                               const double cv[CONS_PT]) {
  __shared__ double sh[16][CONS_ROW_W];
  const int wv = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
  const int w = pass == 0 ? CONS_ROW_W : CONS_ROW_W1;
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    const double v = on ? cv[3 * q] : 0.0, sc = on ? cv[3 * q + 1] : 0.0, sd = on ? cv[3 * q + 2] : 0.0;
    if (pass == 0) {
      const double r0 = wave_sum(v), r1 = wave_sum(sc), r2 = wave_sum(sd), r3 = wave_sum(v * v);
      // the reference's `if (fabs(C) > max)` from 0: NaN never wins
      const double r4 = wave_max(on && fabs(v) > 0.0 ? fabs(v) : 0.0);
      const double r5 = wave_max(on && fabs(sd) > 0.0 ? fabs(sd) : 0.0);
      if ((threadIdx.x & 63) == 0) {
        sh[wv][6 * q] = r0; sh[wv][6 * q + 1] = r1; sh[wv][6 * q + 2] = r2;
        sh[wv][6 * q + 3] = r3; sh[wv][6 * q + 4] = r4; sh[wv][6 * q + 5] = r5;
      }
    } else {
      const double d0 = on ? v - means[2 * q] : 0.0, d1 = on ? sd - means[2 * q + 1] : 0.0;
      const double r0 = wave_sum(d0 * d0), r1 = wave_sum(d1 * d1);
      if ((threadIdx.x & 63) == 0) { sh[wv][2 * q] = r0; sh[wv][2 * q + 1] = r1; }
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < w) {
    const int col = threadIdx.x;
    const bool ismax = pass == 0 && (col % 6) >= 4;
    double acc = sh[0][col];
    for (int k = 1; k < nw; ++k) acc = ismax ? (sh[k][col] > acc ? sh[k][col] : acc) : acc + sh[k][col];
    row[col] = acc;
  }
}

// One compiled instantiation of the stage kernel.
struct StageVariant {
  int prec;               // 0 = FP64, 1 = FP32 split
  int order, halo;
  int ti, tj, tk;         // tile extents (i slowest, k fastest)
  int wpe;                // waves per SIMD the register budget allows
  bool live_alpha;        // false: lapse frozen at 1 (synchronous gauge)
  bool dust;              // evolves the dust fluid with the metric
  int (*launch)(const StageArgs&, hipStream_t);
  int (*info)(int* vgprs, int* scratch, int* lds);
  int (*launch_batch)(const BatchArgs&, hipStream_t);  // null: no batched kernel compiled
  // the same tile with LDS-DMA halo staging (null: not compiled); used for
  // launches without stage-1 normalisation on grids with nz % 4 == 0
  int (*launch_dma)(const StageArgs&, hipStream_t);
  int (*info_dma)(int* vgprs, int* scratch, int* lds);
  int (*launch_batch_dma)(const BatchArgs&, hipStream_t);
};

std::vector<StageVariant>& stage_variants();
struct StageVariantReg {
  explicit StageVariantReg(const StageVariant& v) { stage_variants().push_back(v); }
};

}  // namespace cg
