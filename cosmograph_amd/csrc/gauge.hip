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
// One workgroup per 8x8x8 tile.  The fields whose derivatives are needed go
// through a double-buffered LDS tile with a halo of 3 (the lopsided
// stencils reach 3 points at orders 2 and 4; order 6 needs 3 for its
// centred stencils) or 5 at order 8 (KO of beta); each thread loads its
// share of the next field's halo into registers while the current one is
// differentiated (halo offsets computed once per thread).  upwind_derivative
// is identically zero at orders 6 and 8 (its lopsided stencils are stubs
// there), so at those orders A_ij, K and Gt^i are not staged at all and
// phi and alpha only where a gauge reads their gradients.  The terms are
// accumulated field by field in the staging order (beta, expN, metric, phi,
// A_ij, K, Gt^i, alpha), so that only per-point sums stay live: the
// contracted Christoffel symbols, the Laplacian and gradient of div beta,
// and the shift right-hand sides.  All arithmetic is FP64; FP32-stored
// fields are widened on the way into LDS and their _c rounded back.
#include <hip/hip_runtime.h>

#include <cmath>

#include "bssn_stage.h"
#include "runtime.h"
#include "stencil.h"

namespace cg {
namespace {

constexpr int GT = 8, GNT = GT * GT * GT;

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
  int need_phi, need_alpha;  // phi / alpha staged (their gradients are read)
  double gd_eta, dw_mu_s, dw_mu_l, dw_p, k_driver;
  double phi_F, K_F;
  const double* ext;  // [K_min, K_avg, avg_vol, rho_avg] (device)
  int nseq;           // number of staged arrays (see staged())
};

// Materialises a value in registers at this point: the per-point sums
// below are accumulated field by field, and without this the compiler sinks
// their arithmetic to the last use, keeping every input of it live.
__device__ __forceinline__ void pin(double& x) { asm volatile("" : "+v"(x)); }

__device__ __forceinline__ constexpr int sym(int a, int b) {
  return a == b ? (a == 0 ? 0 : a == 1 ? 3 : 5) : (a + b == 1 ? 1 : a + b == 2 ? 2 : 4);
}

// upwind_derivative (math.h:1247-1252) on an LDS tile at c with stride st:
// the lopsided first derivative plus the dissipation stencil of the order
// (order 2: math.h:166-230 with 1087-1130; order 4: 265-337), forward for a
// positive speed, else backward.  Orders 6 and 8 are stubs returning 0 in
// the reference; the kernel does not call this there.
template <int R>
__device__ __forceinline__ double upwind(const double* T, int c, int st, double v, double i1) {
  if constexpr (R == 1) {
    const double dx = 1.0 / i1;
    if (v > 0)
      return v * ((-0.5 * T[c + 2 * st] + 2.0 * T[c + st] - 1.5 * T[c]) * i1 +
                  -0.5 * dx * ((2.0 * T[c] - 5.0 * T[c + st] + 4.0 * T[c + 2 * st] - T[c + 3 * st]) * i1 * i1));
    return v * ((0.5 * T[c - 2 * st] - 2.0 * T[c - st] + 1.5 * T[c]) * i1 +
                0.5 * dx * ((2.0 * T[c] - 5.0 * T[c - st] + 4.0 * T[c - 2 * st] - T[c - 3 * st]) * i1 * i1));
  } else {
    if (v > 0)
      return v * ((1.0 / 12.0) * T[c + 3 * st] - 0.5 * T[c + 2 * st] + 1.5 * T[c + st] - (5.0 / 6.0) * T[c] -
                  0.25 * T[c - st]) * i1;
    return v * ((-1.0 / 12.0) * T[c - 3 * st] + 0.5 * T[c - 2 * st] - 1.5 * T[c - st] + (5.0 / 6.0) * T[c] +
                0.25 * T[c + st]) * i1;
  }
}

template <typename TS, int R>
__global__ void __launch_bounds__(GNT) __attribute__((amdgpu_waves_per_eu(2))) k_gauge_rhs(GaugeArgs a) {
  constexpr int H = R == 4 ? 5 : 3;                  // lopsided reach 3; KO of order 8 needs 5
  constexpr int KO = R == 1 ? 2 : R == 4 ? 8 : 0;    // KO_dissipation_Q exists at orders 2 and 8
  constexpr bool ADV = R <= 2;                       // upwind_derivative vanishes at orders 6 and 8
  constexpr int HW = GT + 2 * H, HV = HW * HW * HW, NS = (HV + GNT - 1) / GNT;
  constexpr int SI = HW * HW, SJ = HW, SK = 1;
  __shared__ double T2[2 * HV];

  const int b = blockIdx.x;
  const int tz = b % a.ntz, tx = (b / a.ntz) % a.ntx, ty = b / (a.ntz * a.ntx);
  const int i0 = tx * GT, j0 = ty * GT, k0 = tz * GT;
  const int tid = threadIdx.x;
  const int lk = tid % GT, lj = (tid / GT) % GT, li = tid / (GT * GT);
  const int px = i0 + li, py = j0 + lj, pz = k0 + lk;
  const bool valid = px < a.nx && py < a.ny && pz < a.nz;
  const long idx = valid ? ((long)px * a.ny + py) * a.nz + pz : 0;
  const int c = ((li + H) * HW + (lj + H)) * HW + (lk + H);
  const double i1 = a.inv_dx, i2 = a.inv_dx2;

  // halo offsets of this thread's share of the tile (grid points < 2^31)
  int off[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int e = tid + s * GNT;
    off[s] = 0;  // past the tile: a harmless in-bounds load, never stored to LDS
    if (e < HV) {
      const int hk = e % HW, hj = (e / HW) % HW, hi = e / (HW * HW);
      int x = i0 + hi - H, y = j0 + hj - H, z = k0 + hk - H;
      if ((unsigned)x >= (unsigned)a.nx) x = (x % a.nx + a.nx) % a.nx;
      if ((unsigned)y >= (unsigned)a.ny) y = (y % a.ny + a.ny) % a.ny;
      if ((unsigned)z >= (unsigned)a.nz) z = (z % a.nz + a.nz) % a.nz;
      off[s] = (x * a.ny + y) * a.nz + z;
    }
  }
  // the q-th staged array, in the order of the code below: beta^1..3,
  // expN, the six metric fields, then phi (need_phi), A_ij, K and Gt^i
  // (ADV) and alpha (need_alpha).  Uniform arithmetic, no table.
  double pre[NS];
  auto load = [&](int q) {
    const double* pd = nullptr;
    const TS* ps = nullptr;
    if (q < 4) {
      pd = a.xreg + ((long)(q < 3 ? XBETA1 + q : XEXPN) * 4 + a.s_in) * a.rstride;
    } else if (q < 10 + (a.need_phi ? 1 : 0)) {
      pd = a.met + ((long)(q - 4) * 4 + a.s_in) * a.rstride;
    } else {
      const int r = q - 10 - (a.need_phi ? 1 : 0);   // 0..9: A11..Gt^3 (ADV); then alpha
      const int f = ADV && r < 10 ? r : ALPHA - 7;
      ps = (const TS*)a.regs + ((long)f * 4 + a.s_in) * a.rstride;
    }
#pragma unroll
    for (int s = 0; s < NS; ++s)
      pre[s] = pd ? pd[off[s]] : (double)ps[off[s]];
  };
  // stage k: the prefetched halo into buffer k&1, one barrier, then the
  // next array's loads are issued before the caller differentiates this one
  int k = 0;
  const double* T = T2;
  auto next = [&]() {
    double* B = T2 + (k & 1) * HV;
#pragma unroll
    for (int s = 0; s < NS; ++s)
      if (tid + s * GNT < HV) B[tid + s * GNT] = pre[s];
    __syncthreads();
    if (k + 1 < a.nseq) load(k + 1);
    T = B;
    ++k;
  };

  auto reg = [&](int f, int s) -> const void* {
    return f < 7 ? (const void*)(a.met + ((long)f * 4 + s) * a.rstride)
                 : (const void*)((const TS*)a.regs + ((long)(f - 7) * 4 + s) * a.rstride);
  };
  auto xr = [&](int q, int s) -> double* { return a.xreg + ((long)q * 4 + s) * a.rstride; };
  auto ldv = [&](int f, int s) -> double {
    return f < 7 ? ((const double*)reg(f, s))[idx] : (double)((const TS*)reg(f, s))[idx];
  };
  auto stc = [&](int f, double x) {
    if (!valid) return;
    if (f < 7) ((double*)reg(f, a.s_out))[idx] = x;
    else ((TS*)reg(f, a.s_out))[idx] = (TS)x;
  };

  load(0);

  // point values read before their fields are staged
  double gm[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) gm[q] = ldv(G11 + q, a.s_in);
  const double alpha = 1.0 + ldv(ALPHA, a.s_in), K = ldv(KK, a.s_in) + a.K_F, phi = ldv(PHI, a.s_in) + a.phi_F;
  // inverse conformal metric from the DIFF variables (bssn.cc:447-455)
  double gi[6];
  gi[0] = 1.0 + gm[3] + gm[5] - gm[4] * gm[4] + gm[3] * gm[5];
  gi[3] = 1.0 + gm[0] + gm[5] - gm[2] * gm[2] + gm[0] * gm[5];
  gi[5] = 1.0 + gm[0] + gm[3] - gm[1] * gm[1] + gm[0] * gm[3];
  gi[1] = gm[2] * gm[4] - gm[1] * (1.0 + gm[5]);
  gi[2] = gm[1] * gm[4] - gm[2] * (1.0 + gm[3]);
  gi[4] = gm[1] * gm[2] - gm[4] * (1.0 + gm[0]);
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    pin(gi[q]);
    pin(gm[q]);
  }
  auto GI = [&](int p, int q) { return gi[sym(p, q)]; };
  auto GM = [&](int p, int q) { return gm[sym(p, q)] + (p == q ? 1.0 : 0.0); };
  const bool dw = a.shift == SHIFT_DAMPED_WAVE;
  const double e4 = dw ? exp(-4.0 * phi) : 0.0;

  // ---- beta^m: values, first derivatives, and what the Gt^i terms need of
  // the second: gamma^{jk} d_j d_k beta^m and d_j (div beta).  sh[m]
  // accumulates d_t beta^m, starting from minus its KO term.
  double bv[3], db[3][3], lapb[3], ddivb[3] = {0, 0, 0}, sh[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    next();
    bv[m] = T[c];
    db[0][m] = fd1<R>(T, c, SI) * i1;
    db[1][m] = fd1<R>(T, c, SJ) * i1;
    db[2][m] = fd1<R>(T, c, SK) * i1;
    double h[6];
    h[0] = fd2<R>(T, c, SI) * i2;
    h[3] = fd2<R>(T, c, SJ) * i2;
    h[5] = fd2<R>(T, c, SK) * i2;
    h[1] = fdm<R>(T, c, SI, SJ) * i2;
    h[2] = fdm<R>(T, c, SI, SK) * i2;
    h[4] = fdm<R>(T, c, SJ, SK) * i2;
    lapb[m] = gi[0] * h[0] + gi[3] * h[3] + gi[5] * h[5] + 2.0 * (gi[1] * h[1] + gi[2] * h[2] + gi[4] * h[4]);
#pragma unroll
    for (int j = 0; j < 3; ++j) ddivb[j] += h[sym(m, j)];
    sh[m] = -(a.ko_pref * ko_sum<KO, double>(T, c, SI, SJ, SK));
  }
  double divb = db[0][0] + db[1][1] + db[2][2];
  pin(divb);
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    pin(bv[m]);
#pragma unroll
    for (int j = 0; j < 3; ++j) pin(db[j][m]);
  }
  // Gt^i shift terms without the Christoffel part:
  // gamma^{jk} d_j d_k beta^i + 1/3 gamma^{ij} d_j d_k beta^k
  double gx[3];
#pragma unroll
  for (int I = 0; I < 3; ++I)
    gx[I] = lapb[I] + (1.0 / 3.0) * (GI(I, 0) * ddivb[0] + GI(I, 1) * ddivb[1] + GI(I, 2) * ddivb[2]);
#pragma unroll
  for (int I = 0; I < 3; ++I) {
    pin(gx[I]);
    pin(sh[I]);
  }

  // shift gauge terms that need no other field's derivatives
  // (BSSNGaugeHandler.cc:113-273); the rest are added where they appear
#pragma unroll
  for (int I = 0; I < 3; ++I) {
    if (a.shift == SHIFT_GAMMA_DRIVER) {
      sh[I] += xr(XAUXB1 + I, a.s_in)[idx];
    } else if (dw) {
      sh[I] += bv[0] * db[0][I] + bv[1] * db[1][I] + bv[2] * db[2][I] - a.dw_mu_s * alpha * bv[I] +
               alpha * (-a.dw_mu_l * (12.0 * phi * a.dw_p - log(alpha)) * bv[I]);
    } else if (a.shift == SHIFT_AWA_SHIFTED_WAVE) {
      sh[I] += a.wave_dir == I + 1 ? -2.0 * K * alpha : 0.0;
    }
  }
#pragma unroll
  for (int I = 0; I < 3; ++I) pin(sh[I]);
  auto adv = [&]() -> double {
    return upwind<R>(T, c, SI, bv[0], i1) + upwind<R>(T, c, SJ, bv[1], i1) + upwind<R>(T, c, SK, bv[2], i1);
  };

  // ---- expN (bssn.cc:871-875)
  next();
  if (valid)
    xr(XEXPN, a.s_out)[idx] = bv[0] * fd1<R>(T, c, SI) * i1 + bv[1] * fd1<R>(T, c, SJ) * i1 +
                              bv[2] * fd1<R>(T, c, SK) * i1 - alpha * K / 3.0;

  // ---- metric: advection, d_j beta^i couplings (bssn_macros.h:292-312),
  // and Gd^I = gamma^{JK} Gt^I_JK accumulated per component:
  // sum_m d_m g_pq (g^{mp} g^{Iq} + [p!=q] g^{mq} g^{Ip}) - 1/2 w g^{pq} g^{Il} d_l g_pq
  double Gd[3] = {0, 0, 0};
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int q = p; q < 3; ++q) {
      const int f = G11 + sym(p, q);
      next();
      double d[3];
      d[0] = fd1<R>(T, c, SI) * i1;
      d[1] = fd1<R>(T, c, SJ) * i1;
      d[2] = fd1<R>(T, c, SK) * i1;
      const double w = p == q ? 0.5 : 1.0;
#pragma unroll
      for (int I = 0; I < 3; ++I) {
        double t = -w * GI(p, q) * (GI(I, 0) * d[0] + GI(I, 1) * d[1] + GI(I, 2) * d[2]);
#pragma unroll
        for (int m = 0; m < 3; ++m) t += d[m] * (GI(m, p) * GI(I, q) + (p != q ? GI(m, q) * GI(I, p) : 0.0));
        Gd[I] += t;
        pin(Gd[I]);
      }
      double cv = ldv(f, a.s_out);
      if constexpr (ADV) cv += adv();
      cv += GM(p, 0) * db[q][0] + GM(p, 1) * db[q][1] + GM(p, 2) * db[q][2] + GM(q, 0) * db[p][0] +
            GM(q, 1) * db[p][1] + GM(q, 2) * db[p][2] - (2.0 / 3.0) * GM(p, q) * divb;
      stc(f, cv);
    }

  // ---- phi (bssn.cc:806-812)
  {
    double cv = ldv(PHI, a.s_out) + (1.0 / 6.0) * divb;
    if (a.need_phi) {
      next();
      if constexpr (ADV) cv += adv();
      if (dw) {
        const double d0 = fd1<R>(T, c, SI) * i1, d1 = fd1<R>(T, c, SJ) * i1, d2 = fd1<R>(T, c, SK) * i1;
#pragma unroll
        for (int I = 0; I < 3; ++I)
          sh[I] += -2.0 * alpha * alpha * e4 * (GI(I, 0) * d0 + GI(I, 1) * d1 + GI(I, 2) * d2);
      }
    }
    stc(PHI, cv);
  }

  // ---- A_ij (bssn_macros.h:313-333); AijDriverShift sums |d_t A_ij + A_ij|
  {
    double am[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) am[q] = ldv(A11 + q, a.s_in);
    auto AM = [&](int p, int q) { return am[sym(p, q)]; };
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int q = p; q < 3; ++q) {
        const int f = A11 + sym(p, q);
        double cv = ldv(f, a.s_out);
        if constexpr (ADV) {
          next();
          cv += adv();
        }
        cv += AM(p, 0) * db[q][0] + AM(p, 1) * db[q][1] + AM(p, 2) * db[q][2] + AM(q, 0) * db[p][0] +
              AM(q, 1) * db[p][1] + AM(q, 2) * db[p][2] - (2.0 / 3.0) * AM(p, q) * divb;
        stc(f, cv);
        if (a.shift == SHIFT_AIJ_DRIVER) {
          const double t = a.k_driver * fabs(cv + AM(p, q));
          sh[p] -= t;
          if (p != q) sh[q] -= t;
        }
      }
  }

  // ---- DIFFK
  double cvK = ldv(KK, a.s_out);
  if constexpr (ADV) {
    next();
    cvK += adv();
    stc(KK, cvK);
  }

  // ---- Gt^i (bssn_macros.h:363-382); the Gamma driver's d_t B^i
  // (bssn.cc:878-891); DampedWaveShift2 reads Gt^3 where Gt^2 is meant
  // (BSSNGaugeHandler.cc:185), kept
#pragma unroll
  for (int I = 0; I < 3; ++I) {
    double cv = ldv(GAM1 + I, a.s_out), gq;
    if constexpr (ADV) {
      next();
      cv += adv();
      gq = T[c];
    } else {
      gq = dw ? ldv(GAM1 + I, a.s_in) : 0.0;
    }
    cv += -Gd[0] * db[0][I] - Gd[1] * db[1][I] - Gd[2] * db[2][I] + (2.0 / 3.0) * Gd[I] * divb + gx[I];
    stc(GAM1 + I, cv);
    if (a.shift == SHIFT_GAMMA_DRIVER && valid)
      xr(XAUXB1 + I, a.s_out)[idx] = 0.75 * cv - a.gd_eta * xr(XAUXB1 + I, a.s_in)[idx];
    if (dw) {
      const double t = alpha * alpha * e4 * gq;
      if (I == 0) sh[0] += t;
      if (I == 2) { sh[1] += t; sh[2] += t; }
    }
  }

  // ---- the lapse: beta.d alpha in OnePlusLog and DampedWave
  // (BSSNGaugeHandler.cc:64-68, 151-155), TestKDriverLapse (:93-102)
  {
    double lapse = 0.0;
    if (a.lapse == LAPSE_TEST_K_DRIVER) {
      const double Kavg = a.ext[EXTV_KAVG], ravg = a.ext[EXTV_RHOAVG];
      lapse = -a.k_driver * ((cvK - (Kavg * Kavg / 3.0 + 4.0 * M_PI * ravg)) + (K - Kavg));
    }
    if (a.need_alpha) {
      next();
      const double d0 = fd1<R>(T, c, SI) * i1, d1 = fd1<R>(T, c, SJ) * i1, d2 = fd1<R>(T, c, SK) * i1;
      if (a.lapse == LAPSE_ONE_PLUS_LOG || a.lapse == LAPSE_DAMPED_WAVE) lapse = bv[0] * d0 + bv[1] * d1 + bv[2] * d2;
      double cv = ldv(ALPHA, a.s_out) + lapse;
      if constexpr (ADV) cv += adv();
      stc(ALPHA, cv);
      if (dw)
#pragma unroll
        for (int I = 0; I < 3; ++I) sh[I] += -alpha * e4 * (GI(I, 0) * d0 + GI(I, 1) * d1 + GI(I, 2) * d2);
    } else if (a.lapse == LAPSE_TEST_K_DRIVER) {
      stc(ALPHA, ldv(ALPHA, a.s_out) + lapse);
    }
  }

  // ---- d_t beta^i (bssn.cc:852-869)
  if (valid)
#pragma unroll
    for (int I = 0; I < 3; ++I) xr(XBETA1 + I, a.s_out)[idx] = sh[I];
}

// TestKDriverLapse without the shift build: d_t alpha gains
// -k ((d_t K - (K_avg^2/3 + 4 pi rho_avg)) + (K - K_avg)) (BSSNGaugeHandler.cc:93-102),
// which reads the complete d_t DIFFK; nothing else changes
template <typename TS>
__global__ void k_kdriver_lapse(const TS* Kin, const TS* Kc, TS* Ac, const double* ext, double K_F, double kd,
                                long n) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double Kavg = ext[EXTV_KAVG], ravg = ext[EXTV_RHOAVG];
  const double K = (double)Kin[e] + K_F;
  Ac[e] = (TS)((double)Ac[e] - kd * (((double)Kc[e] - (Kavg * Kavg / 3.0 + 4.0 * M_PI * ravg)) + (K - Kavg)));
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
  const bool f32 = c->prec != CG_FP64;
  const size_t w = f32 ? sizeof(float) : sizeof(double);
  auto reg = [&](int f, int slot) -> void* {
    return f < 7 ? (void*)(c->met + ((long)f * 4 + slot) * c->rstride)
                 : (void*)((char*)c->regs + ((long)(f - 7) * 4 + slot) * c->rstride * w);
  };
  if (c->nxf == 0) {
    // TestKDriverLapse without the shift build: a pointwise update of d_t alpha
    const long n = c->npts;
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (f32)
      hipLaunchKernelGGL(k_kdriver_lapse<float>, dim3(nb), dim3(256), 0, c->stream, (const float*)reg(KK, s_in),
                         (const float*)reg(KK, s_out), (float*)reg(ALPHA, s_out), c->ext, c->frw.g_K, c->k_driver, n);
    else
      hipLaunchKernelGGL(k_kdriver_lapse<double>, dim3(nb), dim3(256), 0, c->stream, (const double*)reg(KK, s_in),
                         (const double*)reg(KK, s_out), (double*)reg(ALPHA, s_out), c->ext, c->frw.g_K, c->k_driver,
                         n);
    HIPCHK(c, hipGetLastError());
    return CG_OK;
  }
  GaugeArgs g{};
  g.met = c->met;
  g.regs = c->regs;
  g.xreg = c->xreg;
  g.rstride = c->rstride;
  g.s_in = s_in;
  g.s_out = s_out;
  g.nx = c->p.nx; g.ny = c->p.ny; g.nz = c->p.nz;
  g.ntx = (c->p.nx + GT - 1) / GT; g.nty = (c->p.ny + GT - 1) / GT; g.ntz = (c->p.nz + GT - 1) / GT;
  g.inv_dx = 1.0 / c->p.dx;
  g.inv_dx2 = 1.0 / (c->p.dx * c->p.dx);
  const double dx = c->p.dx, ko = c->p.ko_coeff;
  // KO_dissipation_Q (math.h:36-72): ko*dx^3/64 * sum/dx^4 ; -ko*dx^9/1024 * sum/dx^10
  g.ko_pref = c->p.stencil_order == 2 ? ko / (64.0 * dx) : c->p.stencil_order == 8 ? -ko / (1024.0 * dx) : 0.0;
  g.shift = c->p.shift; g.lapse = c->p.lapse; g.nxf = c->nxf; g.wave_dir = c->gauge_wave_dir;
  g.gd_eta = c->gd_eta; g.dw_mu_s = c->dw_mu_s; g.dw_mu_l = c->dw_mu_l; g.dw_p = c->dw_p; g.k_driver = c->k_driver;
  g.phi_F = c->frw.g_phi; g.K_F = c->frw.g_K;
  g.ext = c->ext;
  // the staging sequence, in the kernel's order: beta^1..3, expN, the
  // metric, then phi, A_ij, K, Gt^i and alpha where their derivatives or
  // their advection are needed (advection only at orders 2 and 4)
  const bool adv = c->p.stencil_order <= 4;
  const bool dw = c->p.shift == SHIFT_DAMPED_WAVE;
  g.need_phi = adv || dw;
  g.need_alpha = adv || dw || c->p.lapse == LAPSE_ONE_PLUS_LOG || c->p.lapse == LAPSE_DAMPED_WAVE;
  g.nseq = 4 + 6 + (g.need_phi ? 1 : 0) + (adv ? 10 : 0) + (g.need_alpha ? 1 : 0);
  int rc = f32 ? launch_gauge_order<float>(c->p.stencil_order, g, c->stream)
               : launch_gauge_order<double>(c->p.stencil_order, g, c->stream);
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
