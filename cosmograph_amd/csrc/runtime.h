// runtime.h -- host-side state shared by the translation units of the
// runtime (single-grid contexts, batched ensembles).  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/cosmo_gpu.h"
#include "bssn_stage.h"

namespace cg {

// ---------------------------------------------------------------- FRW --
// FP64 background integrator with the stage structure of utils/FRW.h:93-254
// (RK4; the getters after stage n are what the RHS of stage n+1 sees).
// The 4*pi factor is formed in long double like the reference's PI_L
// literal (FRW.h:4), so the two agree bitwise.
struct Frw {
  static constexpr long double kPi = 3.141592653589793238462643383279502884L;
  double phi = 0, K = 0;
  std::vector<double> rho, w;
  double kphi[4] = {0, 0, 0, 0}, kK[4] = {0, 0, 0, 0};
  std::vector<double> krho[4];
  double g_phi = 0, g_K = 0, g_rho = 0, g_S = 0;

  void set(double p, double k) { phi = g_phi = p; K = g_K = k; }
  void add_fluid(double r, double ww) {
    rho.push_back(r);
    w.push_back(ww);
    for (auto& v : krho) v.resize(rho.size());
    sum_get(rho);
  }
  void sum_get(const std::vector<double>& rr) {
    g_rho = 0.0;
    g_S = 0.0;
    for (size_t n = 0; n < rr.size(); ++n) {
      g_rho += rr[n];
      g_S += 3.0 * rr[n] * w[n];
    }
  }
  // stage n = 1..4 (P1_step, P2_step, P3_step, RK_total_step)
  void stage(int n, double h) {
    const size_t nf = rho.size();
    double Kin = K, frac = 0.0;
    const std::vector<double>* prev = nullptr;
    if (n == 2) { Kin = K + (h / 2.0) * kK[0]; prev = &krho[0]; frac = h / 2.0; }
    if (n == 3) { Kin = K + (h / 2.0) * kK[1]; prev = &krho[1]; frac = h / 2.0; }
    if (n == 4) { Kin = K + h * kK[2]; prev = &krho[2]; frac = h; }
    std::vector<double> rin(nf);
    kphi[n - 1] = -1.0 * Kin / 6.0;
    double src = 0.0;
    for (size_t q = 0; q < nf; ++q) {
      rin[q] = prev ? rho[q] + frac * (*prev)[q] : rho[q];
      src += 4.0 * kPi * rin[q] * (1.0 + 3.0 * w[q]);
    }
    kK[n - 1] = 1.0 / 3.0 * Kin * Kin + src;
    for (size_t q = 0; q < nf; ++q) krho[n - 1][q] = Kin * rin[q] * (1.0 + w[q]);
    std::vector<double> rg(nf);
    if (n <= 2) {
      g_phi = phi + (h / 2.0) * kphi[n - 1];
      g_K = K + (h / 2.0) * kK[n - 1];
      for (size_t q = 0; q < nf; ++q) rg[q] = rho[q] + (h / 2.0) * krho[n - 1][q];
    } else if (n == 3) {
      g_phi = phi + h * kphi[2];
      g_K = K + h * kK[2];
      for (size_t q = 0; q < nf; ++q) rg[q] = rho[q] + h * krho[2][q];
    } else {
      phi += h / 6.0 * (kphi[0] + 2.0 * kphi[1] + 2.0 * kphi[2] + kphi[3]);
      K += h / 6.0 * (kK[0] + 2.0 * kK[1] + 2.0 * kK[2] + kK[3]);
      for (size_t q = 0; q < nf; ++q)
        rho[q] += h / 6.0 * (krho[0][q] + 2.0 * krho[1][q] + 2.0 * krho[2][q] + krho[3][q]);
      g_phi = phi;
      g_K = K;
      rg = rho;
    }
    sum_get(rg);
  }
};

// sources of the stage whose _a is given (static.cc:28-35, lambda.cc:32-41)
struct SrcArgs {
  const void* DIFFD;
  int matter;
  double phi_F, rho_F, lambda;
  const double* rho;  // DIFFr per point as the stage kernel formed it (dust fluid), else null
};

}  // namespace cg

namespace cg {
struct BrdWork;
}

// ------------------------------------------------------------ context --
struct cg_ctx {
  using Frw = cg::Frw;
  using SrcArgs = cg::SrcArgs;
  using StageVariant = cg::StageVariant;
  using StageArgs = cg::StageArgs;
  cg_params p{};
  int prec = 0;
  size_t ts = 8;          // storage bytes
  int R = 2, H = 2;
  int req_ti = 0, req_tj = 0, req_tk = 0;  // requested tile (0 = automatic)
  const StageVariant* var = nullptr;     // stage kernel in use
  int tile_order = 0;                     // tile traversal order (autotuned)
  // LDS-DMA halo staging: allowed by the caller (cg_set_dma) and picked by
  // the autotuner; used where the variant has it and the grid fits
  bool dma_allowed = true, dma_pick = true;
  double gd_c = 1.0, dw_mu_l = 0.0, dw_p = 0.0;  // gauge parameters (BSSNGaugeHandler.h:139-149)
  // shift build (cg_params::shift): FP64 registers [nxf][4][pad] of beta1..3,
  // expN (and auxB1..3 with the gamma driver), in the slots of the BSSN
  // fields; and the remaining gauge parameters
  double* xreg = nullptr;
  int nxf = 0;
  cg::BrdWork* brd = nullptr;             // Bardeen potentials workspace (cg_bardeen, lazy)
  int gauge_wave_dir = 1;
  double gd_eta = 0.0, dw_mu_s = 0.0, k_driver = 0.04;
  bool dust_bg = false;                   // dust sources as perturbations of the FRW fluids (cg_set_dust_background)
  bool keep_accumulator = false;          // fused steps with the reference's _f accumulator (cg_set_fused_accumulator)
  bool incr_pick = true;                  // the autotuner's choice between the two RK4 forms
  bool alpha_nonzero = false;             // a non-zero DIFFalpha was uploaded
  bool met32 = false;                     // all-FP32 storage: the metric in regs too (CG_FP32_ALL, CG_FP32_FIX)
  // fixed-point metric (CG_FP32_FIX): field f < 7 holds words w, value w * 2^mq_exp[f]
  bool mq = false;
  int mq_exp[7] = {-40, -40, -40, -40, -40, -40, -40};
  double* mq_tmp = nullptr;               // FP64 staging of one field (lazy)
  unsigned* mq_part = nullptr;            // per-block maxima (lazy)
  unsigned* mq_sat = nullptr;             // count of saturated words written by the stage kernels
  double* cons = nullptr;                 // [CONS_N][pad] per-point constraint values (lazy)
  double* cons_part = nullptr;            // reduction partials
  bool cons_valid = false;                // ricci_a / AijAij_a filled
  double* obuf = nullptr;                 // device staging of strips and partial sums (lazy)
  size_t obuf_n = 0;
  long npts = 0, fstride = 0, rstride = 0;
  void* regs = nullptr;   // [11][4][npts] A_ij, DIFFK, Gamma^i, DIFFalpha in the storage type
  double* met = nullptr;  // [7][4][npts] DIFFgamma11..33, DIFFphi registers, always FP64
  void* DIFFD = nullptr;
  // dust fluid (components/dust_fluid/dust.h:32-46): registers of D, S1..S3
  // [4][4][pad] in the storage type, the fluxes of the current stage input
  // [12][pad], and DIFFr of the last stage [pad] (what the reference keeps
  // in DIFFr_a; the averages of setExtraFieldData read it)
  void* dreg = nullptr;
  void* flux = nullptr;
  double* rho_src = nullptr;
  int dust_rk = 0;            // method-level RKEvolve calls since stepInit (source lag)
  double dust_lag_phiF = 0;   // FRW phi the previous RKEvolve's sources were formed with
  unsigned* nan_dev = nullptr;
  unsigned nan_seen = 0;
  // setExtraFieldData (bssn.cc:165-170): [K_min, K_avg, avg_vol, rho_avg] on
  // the device (read by the ConformalFLRW lapse of the next stage), the
  // per-block partial rows, and the first-level combine rows
  double* ext = nullptr;
  double* ext_part = nullptr;
  double* ext_mid = nullptr;
  long ext_rows = 0;
  long last_rows = 0;                     // partial rows written by the last stage launch
  bool ext_fresh = false;                 // ext holds the averages of the current _p
  // after a fused step without the ConformalFLRW lapse the averages of its
  // K4Finalize are formed on demand from _p and the stage-4 input register
  bool ext_lazy = false;
  int ext_lazy_in = 0;
  SrcArgs ext_lazy_src{};
  double ext_lazy_phiF = 0;
  unsigned long long* nan_cnt = nullptr;  // numNaNs counter
  hipStream_t stream = nullptr;
  Frw frw;
  double lambda = 0;
  // register roles (physical slot of each reference register)
  int sP = 0, sF = 1, sA = 2, sC = 3;
  bool a_is_p = true;     // after a fused step _a and _f hold the same data as _p
  std::string err;
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
};

namespace cg {
inline int fail(cg_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}
#define HIPCHK(ctx, expr)                                                              \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) return fail(ctx, CG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

void* reg_ptr(cg_ctx* c, int f, int slot);
int halo_for(int order, double ko);
bool variant_fits(const cg_ctx* c, const StageVariant& v);
bool dma_fits(const cg_ctx* c, const StageVariant& v);
bool morton_fits(int ntx, int nty);
bool alpha_live(const cg_ctx* c);
StageArgs make_args(cg_ctx* c, int stage, int rin, int rp, int rf, int rout, int normalize);
int ext_resolve(cg_ctx* c);
// device array of a field-map name for reading (err 1: unknown, 3: filled by cg_constraints)
const void* field_ptr(cg_ctx* c, const char* name, size_t* esz, int* err);
// the gauge pass (gauge.hip): shift terms, shift registers and the lapses
// that read the complete right-hand side of other fields
inline bool gauge_pass(const cg_ctx* c) { return c->nxf > 0 || c->p.lapse == LAPSE_TEST_K_DRIVER; }
int gauge_rhs(cg_ctx* c, int s_in, int s_out);  // after the stage kernel's RHS into s_out: adds to it, fills the shift s_out
int gauge_step_init(cg_ctx* c);           // RK4Register::stepInit of the shift registers
int gauge_finalize(cg_ctx* c, int n);     // RK4Register::K{n}Finalize of the shift registers (no swap)
// the register the output step's passes read (stepInit's _a, materialised
// after a fused step); resolves the pending averages first
int cons_input(cg_ctx* c, int* rin_out);
// BSSN::RKEvolve at the state in slot rin into slot rout (matter sources of
// that state, the gauge pass of the shift build)
int rhs_at(cg_ctx* c, int rin, int rout);
void bardeen_free(cg_ctx* c);
// fixed-point metric (CG_FP32_FIX): scales 2^e, 2^-e of the seven fields;
// FP64 device values x into words of field f, slot (raising e to fit);
// words back to FP64; the exponent check of cg_rk4_steps
void mq_scales(const cg_ctx* c, double* sc, double* inv);
int mq_store(cg_ctx* c, int f, int slot, const double* x);
int mq_load(cg_ctx* c, int f, int slot, double* x);
int mq_check(cg_ctx* c);
void mq_free(cg_ctx* c);
// a "Bardeen_<name>" array of the last cg_bardeen (err 1: unknown name, 4: not computed)
const double* bardeen_field(const cg_ctx* c, const std::string& name, int* err);
// KO at orders 2/8 with metric normalisation materialises stepInit's _a
inline bool ko_norm_path(const cg_ctx* c) {
  return c->p.ko_coeff != 0.0 && c->p.normalize_metric && (c->p.stencil_order == 2 || c->p.stencil_order == 8);
}
// fused steps may run in the accumulator-free form (StageArgs::incr)
inline bool fused_incr(const cg_ctx* c) { return !ko_norm_path(c) && !c->keep_accumulator; }
// stage slots of an accumulator-free fused step whose stage 1 reads the
// normalised copy in slot n (1..3), or n = 0 with none: a1, a2, a3 follow n
// in the rotation 1 -> 2 -> 3 -> 1, and stage 4 writes the next copy over a1
inline void incr_slots(int n, int& s1, int& s2, int& s3) {
  s1 = n % 3 + 1;
  s2 = s1 % 3 + 1;
  s3 = s2 % 3 + 1;
}
}  // namespace cg
