// gauge.hip -- the gauge pass of the shift build (USE_BSSN_SHIFT,
// cosmo_macros.h:43-44) and of the lapses that read other fields' complete
// right-hand sides.
//
// A stage of a shift context runs the stage kernel first (every term with
// beta = 0: BSSN::RKEvolve of the default build into _c), then this pass,
// which adds what the shift contributes and forms the shift registers'
// right-hand sides:
//   * the advection of every field, sum_d upwind_derivative(d, f, beta^d)
//     (bssn_macros.h:296-370, bssn.cc:786-828, utils/math.h:1220-1252);
//   * the d_j beta^i couplings of DIFFgamma_ij, A_ij (bssn_macros.h:292-333),
//     the Gt^i shift terms with the second derivatives of beta
//     (bssn_macros.h:363-382) and the divergence in d_t DIFFphi
//     (bssn.cc:806-812);
//   * beta.d alpha in the OnePlusLog and DampedWave lapses
//     (BSSNGaugeHandler.cc:64-68, 151-155) and the TestKDriverLapse (:93-102),
//     which reads the complete d_t DIFFK;
//   * d_t beta^i (shift gauge minus KO, bssn.cc:852-869 with
//     BSSNGaugeHandler.cc:113-273), d_t expN (bssn.cc:871-875) and
//     d_t auxB^i (bssn.cc:878-891).
// Every term is additive to what the stage kernel wrote, so the sum equals
// the reference's expression up to the association of the additions.
//
// One workgroup per 4x8x8 tile; each field goes through an LDS tile with a
// halo of 3 (the lopsided stencils reach 3 points at orders 2 and 4; order
// 6 needs 3 for its centred stencils) or 5 at order 8 (KO of beta).  All
// arithmetic is FP64; FP32-stored fields are widened on the way into LDS
// and their _c rounded back on the way out.  This pass runs only for the
// shift build, which the reference compiles out by default, so it is built
// for clarity over speed: one field in LDS at a time, register staging.
#include <hip/hip_runtime.h>

#include <cmath>

#include "bssn_stage.h"
#include "runtime.h"
#include "stencil.h"

namespace cg {
namespace {

constexpr int GTI = 4, GTJ = 8, GTK = 8, GNT = GTI * GTJ * GTK;

struct GaugeArgs {
  double* met;      // metric registers (FP64): field f, slot s at met + (f*4 + s)*rstride
  void* regs;       // other BSSN registers (storage type): field f >= 7 at regs + ((f-7)*4 + s)*rstride
  double* xreg;     // shift registers (FP64): q at xreg + (q*4 + s)*rstride
  long rstride;
  int s_in, s_out;  // _a (stencil input) and _c (right-hand side)
  int nx, ny, nz, ntx, nty, ntz;
  double inv_dx, inv_dx2;
  double ko_pref;   // KO prefactor of the stencil order (0: none)
  int shift, lapse, nxf, wave_dir;
  double gd_eta, dw_mu_s, dw_mu_l, dw_p, k_driver;
  double phi_F, K_F;
  const double* ext;  // [K_min, K_avg, avg_vol, rho_avg] (device)
};

__device__ __forceinline__ int sym(int a, int b) {
  return a == b ? (a == 0 ? 0 : a == 1 ? 3 : 5) : (a + b == 1 ? 1 : a + b == 2 ? 2 : 4);
}
__device__ __forceinline__ int wrapi(int v, int n) {
  v %= n;
  return v < 0 ? v + n : v;
}

// upwind_derivative (math.h:1247-1252) on an LDS tile at c with stride st:
// the lopsided first derivative plus the dissipation stencil of the order
// (order 2: math.h:166-230 with 1087-1130; order 4: 265-337; orders 6 and
// 8: stubs returning 0), forward for a positive speed, else backward
template <int R>
__device__ __forceinline__ double upwind(const double* T, int c, int st, double v, double i1) {
  if constexpr (R == 1) {
    const double dx = 1.0 / i1;
    if (v > 0)
      return v * ((-0.5 * T[c + 2 * st] + 2.0 * T[c + st] - 1.5 * T[c]) * i1 +
                  -0.5 * dx * ((2.0 * T[c] - 5.0 * T[c + st] + 4.0 * T[c + 2 * st] - T[c + 3 * st]) * i1 * i1));
    return v * ((0.5 * T[c - 2 * st] - 2.0 * T[c - st] + 1.5 * T[c]) * i1 +
                0.5 * dx * ((2.0 * T[c] - 5.0 * T[c - st] + 4.0 * T[c - 2 * st] - T[c - 3 * st]) * i1 * i1));
  } else if constexpr (R == 2) {
    if (v > 0)
      return v * ((1.0 / 12.0) * T[c + 3 * st] - 0.5 * T[c + 2 * st] + 1.5 * T[c + st] - (5.0 / 6.0) * T[c] -
                  0.25 * T[c - st]) * i1;
    return v * ((-1.0 / 12.0) * T[c - 3 * st] + 0.5 * T[c - 2 * st] - 1.5 * T[c - st] + (5.0 / 6.0) * T[c] +
                0.25 * T[c + st]) * i1;
  } else {
    return v * 0.0;
  }
}

template <typename TS, int R>
__global__ void __launch_bounds__(GNT) k_gauge_rhs(GaugeArgs a) {
  constexpr int H = R == 4 ? 5 : 3;                  // lopsided reach 3; KO of order 8 needs 5
  constexpr int KO = R == 1 ? 2 : R == 4 ? 8 : 0;    // KO_dissipation_Q exists at orders 2 and 8
  constexpr int HI = GTI + 2 * H, HJ = GTJ + 2 * H, HK = GTK + 2 * H, HV = HI * HJ * HK;
  constexpr int SI = HJ * HK, SJ = HK, SK = 1;
  __shared__ double T[HV];

  const int b = blockIdx.x;
  const int tz = b % a.ntz, tx = (b / a.ntz) % a.ntx, ty = b / (a.ntz * a.ntx);
  const int i0 = tx * GTI, j0 = ty * GTJ, k0 = tz * GTK;
  const int tid = threadIdx.x;
  const int lk = tid % GTK, lj = (tid / GTK) % GTJ, li = tid / (GTK * GTJ);
  const int px = i0 + li, py = j0 + lj, pz = k0 + lk;
  const bool valid = px < a.nx && py < a.ny && pz < a.nz;
  const long idx = valid ? ((long)px * a.ny + py) * a.nz + pz : 0;
  const int c = ((li + H) * HJ + (lj + H)) * HK + (lk + H);
  const double i1 = a.inv_dx, i2 = a.inv_dx2;

  auto reg = [&](int f, int s) -> const void* {
    return f < 7 ? (const void*)(a.met + ((long)f * 4 + s) * a.rstride)
                 : (const void*)((const TS*)a.regs + ((long)(f - 7) * 4 + s) * a.rstride);
  };
  auto xr = [&](int q, int s) -> double* { return a.xreg + ((long)q * 4 + s) * a.rstride; };
  // the halo tile of one array into T (every thread; barriers on both sides)
  auto stage = [&](const auto* src) {
    __syncthreads();
    for (int e = tid; e < HV; e += GNT) {
      const int hk = e % HK, hj = (e / HK) % HJ, hi = e / (HK * HJ);
      const int x = wrapi(i0 + hi - H, a.nx), y = wrapi(j0 + hj - H, a.ny), z = wrapi(k0 + hk - H, a.nz);
      T[e] = (double)src[((long)x * a.ny + y) * a.nz + z];
    }
    __syncthreads();
  };
  auto stage_f = [&](int f) {
    if (f < 7) stage((const double*)reg(f, a.s_in));
    else stage((const TS*)reg(f, a.s_in));
  };

  // ---- beta^m: values, first and second derivatives, KO (zero in the
  // lapse-only pass, which has no shift registers)
  const bool sh_build = a.nxf > 0;  // uniform over the grid
  double bv[3] = {0, 0, 0}, db[3][3] = {}, ddb[3][6] = {}, kob[3] = {0, 0, 0}, dN[3] = {0, 0, 0};
#pragma unroll
  for (int m = 0; m < 3 && sh_build; ++m) {
    stage((const double*)xr(XBETA1 + m, a.s_in));
    bv[m] = T[c];
    db[0][m] = fd1<R>(T, c, SI) * i1;
    db[1][m] = fd1<R>(T, c, SJ) * i1;
    db[2][m] = fd1<R>(T, c, SK) * i1;
    ddb[m][0] = fd2<R>(T, c, SI) * i2;
    ddb[m][3] = fd2<R>(T, c, SJ) * i2;
    ddb[m][5] = fd2<R>(T, c, SK) * i2;
    ddb[m][1] = fdm<R>(T, c, SI, SJ) * i2;
    ddb[m][2] = fdm<R>(T, c, SI, SK) * i2;
    ddb[m][4] = fdm<R>(T, c, SJ, SK) * i2;
    kob[m] = a.ko_pref * ko_sum<KO, double>(T, c, SI, SJ, SK);
  }
  // upwind advection of the tile in T along beta
  auto adv = [&]() -> double {
    return upwind<R>(T, c, SI, bv[0], i1) + upwind<R>(T, c, SJ, bv[1], i1) + upwind<R>(T, c, SK, bv[2], i1);
  };
  if (sh_build) {
    stage((const double*)xr(XEXPN, a.s_in));
    dN[0] = fd1<R>(T, c, SI) * i1;
    dN[1] = fd1<R>(T, c, SJ) * i1;
    dN[2] = fd1<R>(T, c, SK) * i1;
  }

  // ---- BSSN fields: point values, the derivatives the shift terms need,
  // and their advection
  double v[NF], ad[NF], dg[3][6], dphi[3], dal[3];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    stage_f(f);
    v[f] = T[c];
    ad[f] = adv();
    if (f <= G33) {
      dg[0][f] = fd1<R>(T, c, SI) * i1;
      dg[1][f] = fd1<R>(T, c, SJ) * i1;
      dg[2][f] = fd1<R>(T, c, SK) * i1;
    } else if (f == PHI || f == ALPHA) {
      double* d = f == PHI ? dphi : dal;
      d[0] = fd1<R>(T, c, SI) * i1;
      d[1] = fd1<R>(T, c, SJ) * i1;
      d[2] = fd1<R>(T, c, SK) * i1;
    }
  }
  if (!valid) return;  // no barriers below

  // inverse conformal metric from the DIFF variables (bssn.cc:447-455)
  double gi[6];
  {
    const double g11 = v[G11], g12 = v[G12], g13 = v[G13], g22 = v[G22], g23 = v[G23], g33 = v[G33];
    gi[0] = 1.0 + g22 + g33 - g23 * g23 + g22 * g33;
    gi[3] = 1.0 + g11 + g33 - g13 * g13 + g11 * g33;
    gi[5] = 1.0 + g11 + g22 - g12 * g12 + g11 * g22;
    gi[1] = g13 * g23 - g12 * (1.0 + g33);
    gi[2] = g12 * g23 - g13 * (1.0 + g22);
    gi[4] = g12 * g13 - g23 * (1.0 + g11);
  }
  auto GI = [&](int p, int q) { return gi[sym(p, q)]; };
  auto GM = [&](int p, int q) { return v[G11 + sym(p, q)] + (p == q ? 1.0 : 0.0); };
  auto AM = [&](int p, int q) { return v[A11 + sym(p, q)]; };
  // Gt^i_d = gamma^{jk} Gt^i_jk of the metric (bssn_macros.h:221-229, bssn.cc:584-589)
  double Gd[3];
#pragma unroll
  for (int I = 0; I < 3; ++I) {
    double G[6];
#pragma unroll
    for (int J = 0; J < 3; ++J)
#pragma unroll
      for (int K = J; K < 3; ++K) {
        double acc = 0.0;
#pragma unroll
        for (int l = 0; l < 3; ++l) acc += GI(I, l) * (dg[J][sym(K, l)] + dg[K][sym(J, l)] - dg[l][sym(J, K)]);
        G[sym(J, K)] = 0.5 * acc;
      }
    Gd[I] = G[0] * gi[0] + G[3] * gi[3] + G[5] * gi[5] + 2.0 * (G[1] * gi[1] + G[2] * gi[2] + G[4] * gi[4]);
  }
  const double divb = db[0][0] + db[1][1] + db[2][2];

  // ---- add the shift terms to _c
  double cv[NF];
  auto ldc = [&](int f) -> double {
    return f < 7 ? ((const double*)reg(f, a.s_out))[idx] : (double)((const TS*)reg(f, a.s_out))[idx];
  };
  auto stc = [&](int f, double x) {
    if (f < 7) ((double*)reg(f, a.s_out))[idx] = x;
    else ((TS*)reg(f, a.s_out))[idx] = (TS)x;
  };
#pragma unroll
  for (int f = 0; f < NF; ++f) cv[f] = ldc(f);
#pragma unroll
  for (int I = 0; I < 3; ++I)
#pragma unroll
    for (int J = I; J < 3; ++J) {
      const int s = sym(I, J);
      cv[G11 + s] += ad[G11 + s] + GM(I, 0) * db[J][0] + GM(I, 1) * db[J][1] + GM(I, 2) * db[J][2] +
                     GM(J, 0) * db[I][0] + GM(J, 1) * db[I][1] + GM(J, 2) * db[I][2] - (2.0 / 3.0) * GM(I, J) * divb;
      cv[A11 + s] += ad[A11 + s] + AM(I, 0) * db[J][0] + AM(I, 1) * db[J][1] + AM(I, 2) * db[J][2] +
                     AM(J, 0) * db[I][0] + AM(J, 1) * db[I][1] + AM(J, 2) * db[I][2] - (2.0 / 3.0) * AM(I, J) * divb;
    }
#pragma unroll
  for (int I = 0; I < 3; ++I) {
    double t13 = 0.0;  // gamma^{Ij} d_j d_k beta^k
#pragma unroll
    for (int j = 0; j < 3; ++j) t13 += GI(I, j) * (ddb[0][sym(0, j)] + ddb[1][sym(1, j)] + ddb[2][sym(2, j)]);
    cv[GAM1 + I] += ad[GAM1 + I] - Gd[0] * db[0][I] - Gd[1] * db[1][I] - Gd[2] * db[2][I] +
                    (2.0 / 3.0) * Gd[I] * divb + (1.0 / 3.0) * t13 +
                    (gi[0] * ddb[I][0] + gi[3] * ddb[I][3] + gi[5] * ddb[I][5] +
                     2.0 * (gi[1] * ddb[I][1] + gi[2] * ddb[I][2] + gi[4] * ddb[I][4]));
  }
  cv[KK] += ad[KK];
  cv[PHI] += (1.0 / 6.0) * divb + ad[PHI];
  const double alpha = 1.0 + v[ALPHA], K = v[KK] + a.K_F, phi = v[PHI] + a.phi_F;
  const double bda = bv[0] * dal[0] + bv[1] * dal[1] + bv[2] * dal[2];
  double lapse = 0.0;
  if (a.lapse == LAPSE_ONE_PLUS_LOG || a.lapse == LAPSE_DAMPED_WAVE) {
    lapse = bda;
  } else if (a.lapse == LAPSE_TEST_K_DRIVER) {
    const double Kavg = a.ext[EXTV_KAVG], ravg = a.ext[EXTV_RHOAVG];
    lapse = -a.k_driver * ((cv[KK] - (Kavg * Kavg / 3.0 + 4.0 * M_PI * ravg)) + (K - Kavg));
  }
  cv[ALPHA] += lapse + ad[ALPHA];
  if (a.nxf == 0) {
    // lapse-only gauge pass (TestKDriverLapse without the shift build)
    stc(ALPHA, cv[ALPHA]);
    return;
  }
#pragma unroll
  for (int f = 0; f < NF; ++f) stc(f, cv[f]);

  // ---- right-hand sides of the shift registers
#pragma unroll
  for (int I = 0; I < 3; ++I) {
    double sh = 0.0;
    if (a.shift == SHIFT_GAMMA_DRIVER) {
      sh = xr(XAUXB1 + I, a.s_in)[idx];
    } else if (a.shift == SHIFT_DAMPED_WAVE) {
      // the y component reads Gamma3 where Gamma2 is meant (BSSNGaugeHandler.cc:185), kept
      const double Gq = v[GAM1 + (I == 0 ? 0 : 2)];
      sh = bv[0] * db[0][I] + bv[1] * db[1][I] + bv[2] * db[2][I] - a.dw_mu_s * alpha * bv[I] +
           alpha * (-a.dw_mu_l * (12.0 * phi * a.dw_p - log(alpha)) * bv[I] +
                    exp(-4.0 * phi) * (-GI(I, 0) * dal[0] - GI(I, 1) * dal[1] - GI(I, 2) * dal[2] +
                                       alpha * (Gq - 2.0 * (GI(I, 0) * dphi[0] + GI(I, 1) * dphi[1] +
                                                            GI(I, 2) * dphi[2]))));
    } else if (a.shift == SHIFT_AWA_SHIFTED_WAVE) {
      sh = a.wave_dir == I + 1 ? -2.0 * K * alpha : 0.0;
    } else if (a.shift == SHIFT_AIJ_DRIVER) {
      sh = -a.k_driver * (fabs(cv[A11 + sym(I, 0)] + AM(I, 0)) + fabs(cv[A11 + sym(I, 1)] + AM(I, 1)) +
                          fabs(cv[A11 + sym(I, 2)] + AM(I, 2)));
    }
    xr(XBETA1 + I, a.s_out)[idx] = sh - kob[I];
  }
  xr(XEXPN, a.s_out)[idx] = bv[0] * dN[0] + bv[1] * dN[1] + bv[2] * dN[2] - alpha * K / 3.0;
  if (a.nxf > XAUXB1)
#pragma unroll
    for (int I = 0; I < 3; ++I)
      xr(XAUXB1 + I, a.s_out)[idx] = 0.75 * cv[GAM1 + I] - a.gd_eta * xr(XAUXB1 + I, a.s_in)[idx];
}

// RK4Register::stepInit (RK4Register.h:118-127) of the shift registers
__global__ void k_xstep_init(double* x, long rstride, int nxf, int sp, int sa, int sf, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * nxf) return;
  const int q = (int)(t / n);
  const long e = t - (long)q * n;
  double* r = x + (long)q * 4 * rstride;
  r[(long)sa * rstride + e] = r[(long)sp * rstride + e];
  r[(long)sf * rstride + e] = 0.0;
}

// RK4Register::K{n}Finalize (RK4Register.h:129-176) of the shift registers,
// without the pointer swap
__global__ void k_xfinalize(double* x, long rstride, int nxf, int sp, int sc, int sf, long n, int stage,
                            double dt) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * nxf) return;
  const int q = (int)(t / n);
  const long e = t - (long)q * n;
  double* r = x + (long)q * 4 * rstride;
  double *P = r + (long)sp * rstride, *C = r + (long)sc * rstride, *F = r + (long)sf * rstride;
  const double cc = C[e], p = P[e], fo = F[e];
  if (stage == 1) { F[e] = fo + dt * cc / 6.0; C[e] = p + dt * cc / 2.0; }
  else if (stage == 2) { F[e] = fo + dt * cc / 3.0; C[e] = p + dt * cc / 2.0; }
  else if (stage == 3) { F[e] = fo + dt * cc / 3.0; C[e] = p + dt * cc; }
  else { const double fn = fo + (dt * cc / 6.0 + p); F[e] = fn; C[e] = fn; P[e] = fn; }
}

template <typename TS, int R>
int launch_gauge(const GaugeArgs& g, hipStream_t s) {
  const long nb = (long)g.ntx * g.nty * g.ntz;
  hipLaunchKernelGGL((k_gauge_rhs<TS, R>), dim3((unsigned)nb), dim3(GNT), 0, s, g);
  return (int)hipGetLastError();
}

template <typename TS>
int launch_gauge_order(int order, const GaugeArgs& g, hipStream_t s) {
  switch (order) {
    case 2: return launch_gauge<TS, 1>(g, s);
    case 4: return launch_gauge<TS, 2>(g, s);
    case 6: return launch_gauge<TS, 3>(g, s);
    default: return launch_gauge<TS, 4>(g, s);
  }
}

}  // namespace

int gauge_rhs(cg_ctx* c, int s_in, int s_out) {
  GaugeArgs g{};
  g.met = c->met;
  g.regs = c->regs;
  g.xreg = c->xreg;
  g.rstride = c->rstride;
  g.s_in = s_in;
  g.s_out = s_out;
  g.nx = c->p.nx; g.ny = c->p.ny; g.nz = c->p.nz;
  g.ntx = (c->p.nx + GTI - 1) / GTI; g.nty = (c->p.ny + GTJ - 1) / GTJ; g.ntz = (c->p.nz + GTK - 1) / GTK;
  g.inv_dx = 1.0 / c->p.dx;
  g.inv_dx2 = 1.0 / (c->p.dx * c->p.dx);
  const double dx = c->p.dx, ko = c->p.ko_coeff;
  // KO_dissipation_Q (math.h:36-72): ko*dx^3/64 * sum/dx^4 ; -ko*dx^9/1024 * sum/dx^10
  g.ko_pref = c->p.stencil_order == 2 ? ko / (64.0 * dx) : c->p.stencil_order == 8 ? -ko / (1024.0 * dx) : 0.0;
  g.shift = c->p.shift; g.lapse = c->p.lapse; g.nxf = c->nxf; g.wave_dir = c->gauge_wave_dir;
  g.gd_eta = c->gd_eta; g.dw_mu_s = c->dw_mu_s; g.dw_mu_l = c->dw_mu_l; g.dw_p = c->dw_p; g.k_driver = c->k_driver;
  g.phi_F = c->frw.g_phi; g.K_F = c->frw.g_K;
  g.ext = c->ext;
  // the shift registers exist only with the shift build (nxf > 0); the
  // lapse-only pass (nxf == 0) reads and writes none of them
  int rc = c->prec == CG_FP64 ? launch_gauge_order<double>(c->p.stencil_order, g, c->stream)
                              : launch_gauge_order<float>(c->p.stencil_order, g, c->stream);
  if (rc) return fail(c, CG_ERR_HIP, std::string("gauge pass: ") + hipGetErrorString((hipError_t)rc));
  return CG_OK;
}

int gauge_step_init(cg_ctx* c) {
  if (!c->nxf) return CG_OK;
  const long n = c->npts, tot = n * c->nxf;
  hipLaunchKernelGGL(k_xstep_init, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c->stream, c->xreg,
                     c->rstride, c->nxf, c->sP, c->sA, c->sF, n);
  HIPCHK(c, hipGetLastError());
  return CG_OK;
}

int gauge_finalize(cg_ctx* c, int n) {
  if (!c->nxf) return CG_OK;
  const long np = c->npts, tot = np * c->nxf;
  hipLaunchKernelGGL(k_xfinalize, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c->stream, c->xreg,
                     c->rstride, c->nxf, c->sP, c->sC, c->sF, np, n, c->p.dt);
  HIPCHK(c, hipGetLastError());
  return CG_OK;
}

}  // namespace cg
