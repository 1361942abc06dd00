// cosmo_gpu.hip -- host runtime and C ABI for the MI355X BSSN hot path.
//
// Owns the device registers of one periodic grid, the FP64 FRW reference
// integrator (host side, scalar), and the launch sequence of the fused RK4
// stage kernels.  See include/cosmo_gpu.h for the reference method each
// entry point replaces.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/cosmo_gpu.h"
#include "bssn_stage.h"
#include "runtime.h"

using namespace cg;

namespace {


const char* kFieldNames[NF] = {"DIFFgamma11", "DIFFgamma12", "DIFFgamma13", "DIFFgamma22", "DIFFgamma23",
                               "DIFFgamma33", "DIFFphi",     "A11",         "A12",         "A13",
                               "A22",         "A23",         "A33",         "DIFFK",       "Gamma1",
                               "Gamma2",      "Gamma3",      "DIFFalpha"};

// ------------------------------------------------- elementwise kernels --
// Register table: slot pointers of every field (DIFFphi always FP64).
struct MetScale {
  double s[7];
};

struct RegTable {
  void* r[NF][4];
  double ms[7], mi[7];  // fixed-point metric: scale 2^e and 2^-e per field
  unsigned* sat;        // ... and its count of saturated words
};

// the conformal metric (DIFFgamma_ij, DIFFphi) has its own storage type MT:
// FP64 in the drop-in and the split mode, FP32 in the all-FP32 mode, int
// (32-bit fixed point, scaled per field) in the fixed-point mode
__device__ __host__ inline bool is_metric(int f) { return f <= G33 || f == PHI; }

template <typename TS, typename MT>
__device__ __forceinline__ double ld(const RegTable& t, int f, int s, long x) {
  if (is_metric(f)) {
    if constexpr (std::is_same<MT, int>::value) return (double)((const int*)t.r[f][s])[x] * t.ms[f < 7 ? f : 0];
    return (double)((const MT*)t.r[f][s])[x];
  }
  return (double)((const TS*)t.r[f][s])[x];
}
template <typename TS, typename MT>
__device__ __forceinline__ void st(const RegTable& t, int f, int s, long x, double v) {
  if (is_metric(f)) {
    if constexpr (std::is_same<MT, int>::value) {
      if (fabs(v * t.mi[f < 7 ? f : 0]) > MQ_WMAX) atomicAdd(t.sat, 1u);
      ((int*)t.r[f][s])[x] = mq_word(v, t.mi[f < 7 ? f : 0]);
    }
    else ((MT*)t.r[f][s])[x] = (MT)v;
  } else {
    ((TS*)t.r[f][s])[x] = (TS)v;
  }
}

// Elementwise kernels run 256 threads per block; threads past the end keep
// going with point 0 and store nothing, so whole blocks reach the
// setExtraFieldData reduction (one partial row per block, see bssn_stage.h).
template <typename TS, typename MT>
__global__ void k_step_init(RegTable t, int rp, int ra, int rf, long n, int normalize, double* ext_part,
                            double phiF) {
  const long xg = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = xg < n;
  const long x = valid ? xg : 0;
  // RK4Register::stepInit (RK4Register.h:118-127)
  double v[NF];
  for (int f = 0; f < NF; ++f) {
    v[f] = ld<TS, MT>(t, f, rp, x);
    if (valid) st<TS, MT>(t, f, rf, x, 0.0);
  }
  // BSSN::stepInit calls setExtraFieldData before the normalisation, which
  // leaves DIFFphi and DIFFK alone (bssn.cc:219-226); DIFFr is not touched
  if (ext_part) block_ext_put(ext_part, blockIdx.x, valid, v[PHI], v[KK], 0.0, phiF);
  if (!valid) return;
  if (normalize) {
    // BSSN::set_DIFFgamma_Aij_norm (bssn.cc:61-119)
    double g11 = v[G11], g12 = v[G12], g13 = v[G13], g22 = v[G22], g23 = v[G23], g33 = v[G33];
    double omd = -1.0 * (g11 + g22 + g33 - g12 * g12 - g13 * g13 - g23 * g23 + g11 * g22 + g11 * g33 + g22 * g33 -
                         g23 * g23 * g11 - g13 * g13 * g22 - g12 * g12 * g33 + 2.0 * g12 * g13 * g23 + g11 * g22 * g33);
    double tn = -1.0 * expm1(log1p(-1.0 * omd) / 3.0);
    g11 = (g11 + tn) / (1.0 - tn); g22 = (g22 + tn) / (1.0 - tn); g33 = (g33 + tn) / (1.0 - tn);
    g12 = g12 / (1.0 - tn); g13 = g13 / (1.0 - tn); g23 = g23 / (1.0 - tn);
    v[G11] = g11; v[G12] = g12; v[G13] = g13; v[G22] = g22; v[G23] = g23; v[G33] = g33;
    double i11 = 1.0 + g22 + g33 - g23 * g23 + g22 * g33, i22 = 1.0 + g11 + g33 - g13 * g13 + g11 * g33;
    double i33 = 1.0 + g11 + g22 - g12 * g12 + g11 * g22, i12 = g13 * g23 - g12 * (1.0 + g33);
    double i13 = g12 * g23 - g13 * (1.0 + g22), i23 = g12 * g13 - g23 * (1.0 + g11);
    double trA = i11 * v[A11] + i22 * v[A22] + i33 * v[A33] + 2.0 * (i12 * v[A12] + i13 * v[A13] + i23 * v[A23]);
    v[A11] = (v[A11] - 1.0 / 3.0 * (1.0 + g11) * trA) / (1.0 - tn);
    v[A22] = (v[A22] - 1.0 / 3.0 * (1.0 + g22) * trA) / (1.0 - tn);
    v[A33] = (v[A33] - 1.0 / 3.0 * (1.0 + g33) * trA) / (1.0 - tn);
    v[A12] = (v[A12] - 1.0 / 3.0 * g12 * trA) / (1.0 - tn);
    v[A13] = (v[A13] - 1.0 / 3.0 * g13 * trA) / (1.0 - tn);
    v[A23] = (v[A23] - 1.0 / 3.0 * g23 * trA) / (1.0 - tn);
  }
  for (int f = 0; f < NF; ++f) st<TS, MT>(t, f, ra, x, v[f]);
}


template <typename TS>
__device__ __forceinline__ double diff_r(const SrcArgs& s, double dphi, long x) {
  if (s.rho) return s.rho[x];
  double r = 0.0;
  if (s.matter == MATTER_STATIC)
    r = exp(-6.0 * (dphi + s.phi_F)) * (double)((const TS*)s.DIFFD)[x] + s.rho_F * expm1(-6.0 * dphi);
  return r + s.lambda;
}

// RK4Register::K{n}Finalize (RK4Register.h:129-176) without the pointer swap,
// then setExtraFieldData (bssn.cc:267-305) of the new _a: ra is the _a of the
// finished stage, whose DIFFphi gives that stage's DIFFr
template <typename TS, typename MT>
__global__ void k_finalize(RegTable t, int rp, int rc, int rf, int ra, long n, int stage, double dt,
                           double* ext_part, SrcArgs src, double phiF_next) {
  const long xg = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = xg < n;
  const long x = valid ? xg : 0;
  double phi_new = 0.0, K_new = 0.0;
  const double phi_old = ld<TS, MT>(t, PHI, ra, x);
  for (int f = 0; f < NF; ++f) {
    const double c = ld<TS, MT>(t, f, rc, x), p = ld<TS, MT>(t, f, rp, x), fo = ld<TS, MT>(t, f, rf, x);
    double fn, cn;
    if (stage == 1) { fn = fo + dt * c / 6.0; cn = p + dt * c / 2.0; }
    else if (stage == 2) { fn = fo + dt * c / 3.0; cn = p + dt * c / 2.0; }
    else if (stage == 3) { fn = fo + dt * c / 3.0; cn = p + dt * c; }
    else { fn = fo + (dt * c / 6.0 + p); cn = fn; }
    if (f == PHI) phi_new = (double)(MT)cn;
    if (f == KK) K_new = is_metric(f) ? cn : (double)(TS)cn;
    if (!valid) continue;
    st<TS, MT>(t, f, rf, x, fn);
    st<TS, MT>(t, f, rc, x, cn);
    if (stage == 4) st<TS, MT>(t, f, rp, x, fn);
  }
  if (ext_part) block_ext_put(ext_part, blockIdx.x, valid, phi_new, K_new, diff_r<TS>(src, phi_old, x), phiF_next);
}

// setExtraFieldData after a fused step's K4Finalize: the new _p, and the
// stage-4 input register for that stage's DIFFr
// (phi words times ms: 1 except for the fixed-point metric)
template <typename TS, typename MT>
__global__ void k_ext_post(const MT* __restrict__ phi, const TS* __restrict__ K,
                           const MT* __restrict__ phi_in, long n, double* ext_part, SrcArgs src, double phiF,
                           double ms) {
  const long xg = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = xg < n;
  const long x = valid ? xg : 0;
  block_ext_put(ext_part, blockIdx.x, valid, (double)phi[x] * ms, (double)K[x],
                diff_r<TS>(src, (double)phi_in[x] * ms, x), phiF);
}

// setExtraFieldData over a register alone (DIFFr part left out): used when
// the averages are needed before any stage has produced them
template <typename TS, typename MT>
__global__ void k_ext_state(const MT* __restrict__ phi, const TS* __restrict__ K, long n, double* ext_part,
                            double phiF, double ms) {
  const long xg = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = xg < n;
  const long x = valid ? xg : 0;
  block_ext_put(ext_part, blockIdx.x, valid, (double)phi[x] * ms, (double)K[x], 0.0, phiF);
}

// Deterministic combine of partial rows: block b reduces the contiguous
// range [b*per, (b+1)*per) with a fixed per-thread stride and a fixed tree;
// with gridDim.x == 1 and ext != null it also forms the four averages of
// setExtraFieldData (math.h:1382-1415: sum/POINTS/vol, vol = sum w/POINTS).
constexpr int kExtThreads = 256;
constexpr int kExtMid = 64;  // first-level combine blocks
__global__ void k_ext_reduce(const double* __restrict__ in, long rows, double* __restrict__ out, double* ext,
                             double npts, int keep_rho) {
  __shared__ double sh[EXT_W][kExtThreads];
  const long per = (rows + gridDim.x - 1) / gridDim.x;
  const long lo = (long)blockIdx.x * per, hi = lo + per < rows ? lo + per : rows;
  double acc[EXT_W] = {0, 0, 0, 0, __builtin_huge_val()};
  for (long r = lo + threadIdx.x; r < hi; r += kExtThreads)
    for (int q = 0; q < EXT_W; ++q) {
      const double v = in[r * EXT_W + q];
      acc[q] = q == EXT_KMIN ? (v < acc[q] ? v : acc[q]) : acc[q] + v;
    }
  for (int q = 0; q < EXT_W; ++q) sh[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int w = kExtThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int q = 0; q < EXT_W; ++q) {
        const double o = sh[q][threadIdx.x + w];
        sh[q][threadIdx.x] = q == EXT_KMIN ? (o < sh[q][threadIdx.x] ? o : sh[q][threadIdx.x]) : sh[q][threadIdx.x] + o;
      }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    for (int q = 0; q < EXT_W; ++q) out[(long)blockIdx.x * EXT_W + q] = sh[q][0];
    if (ext) {
      const double vol = sh[EXT_SW][0] / npts;
      ext[EXTV_KMIN] = sh[EXT_KMIN][0];
      ext[EXTV_KAVG] = sh[EXT_SWK][0] / npts / vol;
      ext[EXTV_AVGVOL] = exp(6.0 * (sh[EXT_SWPHI][0] / npts / vol));
      if (!keep_rho) ext[EXTV_RHOAVG] = sh[EXT_SWR][0] / npts / vol;
    }
  }
}

// ------------------------------------------------------------ dust fluid --
// Dust::populateDerivedFields, the flux part (dust.cc:109-131, 177-200): at
// every point of the stage input, v^i = e^{-10 phi} gamma~^{ij} S_j and the
// fluxes alpha D v^i, alpha S_j v^i (flux q*3 + i; q = 0 D, 1..3 S_j).  At
// stage 1 of a fused step the metric is normalised on the fly as the stage
// kernel does (bssn.cc:61-119).
// (MT: the metric's storage, FP64 or 32-bit fixed-point words times ms[f];
// either way field f, slot s at met + (f*4 + s)*rstride)
template <typename TS, typename MT>
__global__ void k_dust_flux(const MT* __restrict__ met, MetScale ms, const TS* __restrict__ alpha,
                            const TS* __restrict__ dreg, long rstride, int slot, int mslot, int normalize,
                            double phiF, TS* __restrict__ flux, long n) {
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  // DIFFgamma_ij at slot mslot (a normalised copy of slot's state, or slot)
  auto M = [&](int f) { return (double)met[((long)f * 4 + (f < PHI ? mslot : slot)) * rstride + x] * ms.s[f]; };
  double g11 = M(G11), g12 = M(G12), g13 = M(G13), g22 = M(G22), g23 = M(G23), g33 = M(G33);
  if (normalize) {
    const double omd = -1.0 * (g11 + g22 + g33 - g12 * g12 - g13 * g13 - g23 * g23 + g11 * g22 + g11 * g33 +
                               g22 * g33 - g23 * g23 * g11 - g13 * g13 * g22 - g12 * g12 * g33 +
                               2.0 * g12 * g13 * g23 + g11 * g22 * g33);
    const double tn = -1.0 * expm1(log1p(-1.0 * omd) / 3.0);
    g11 = (g11 + tn) / (1.0 - tn); g22 = (g22 + tn) / (1.0 - tn); g33 = (g33 + tn) / (1.0 - tn);
    g12 = g12 / (1.0 - tn); g13 = g13 / (1.0 - tn); g23 = g23 / (1.0 - tn);
  }
  // set_gammai_values (bssn.cc:447-455)
  const double i11 = 1.0 + g22 + g33 - g23 * g23 + g22 * g33, i22 = 1.0 + g11 + g33 - g13 * g13 + g11 * g33;
  const double i33 = 1.0 + g11 + g22 - g12 * g12 + g11 * g22, i12 = g13 * g23 - g12 * (1.0 + g33);
  const double i13 = g12 * g23 - g13 * (1.0 + g22), i23 = g12 * g13 - g23 * (1.0 + g11);
  const double phi = M(PHI) + phiF;
  const double al = (double)alpha[(long)slot * rstride + x] + 1.0;
  double q[4];
  for (int f = 0; f < NDUST; ++f) q[f] = (double)dreg[((long)f * 4 + slot) * rstride + x];
  const double S1 = q[1], S2 = q[2], S3 = q[3];
  const double em4phi = exp(-4.0 * phi), e6phi = exp(6.0 * phi);
  const double v[3] = {em4phi / e6phi * (i11 * S1 + i12 * S2 + i13 * S3),
                       em4phi / e6phi * (i12 * S1 + i22 * S2 + i23 * S3),
                       em4phi / e6phi * (i13 * S1 + i23 * S2 + i33 * S3)};
  for (int f = 0; f < NDUST; ++f)
    for (int d = 0; d < 3; ++d) flux[(long)(f * 3 + d) * rstride + x] = (TS)(al * q[f] * v[d]);
}

// RK4Register::stepInit of the dust registers (dust.cc:65-73)
template <typename TS>
__global__ void k_dust_reg_init(TS* __restrict__ dreg, long rstride, int rp, int ra, int rf, long n) {
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  for (int f = 0; f < NDUST; ++f) {
    TS* r = dreg + (long)f * 4 * rstride;
    r[(long)ra * rstride + x] = r[(long)rp * rstride + x];
    r[(long)rf * rstride + x] = TS(0);
  }
}

// Dust::K{n}Finalize (dust.cc:78-106): RK4Register::K{n}Finalize of D, S_j
template <typename TS>
__global__ void k_dust_finalize(TS* __restrict__ dreg, long rstride, int rp, int rc, int rf, long n, int stage,
                                double dt) {
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  for (int f = 0; f < NDUST; ++f) {
    TS* r = dreg + (long)f * 4 * rstride;
    const double c = (double)r[(long)rc * rstride + x], p = (double)r[(long)rp * rstride + x],
                 fo = (double)r[(long)rf * rstride + x];
    double fn, cn;
    if (stage == 1) { fn = fo + dt * c / 6.0; cn = p + dt * c / 2.0; }
    else if (stage == 2) { fn = fo + dt * c / 3.0; cn = p + dt * c / 2.0; }
    else if (stage == 3) { fn = fo + dt * c / 3.0; cn = p + dt * c; }
    else { fn = fo + (dt * c / 6.0 + p); cn = fn; }
    r[(long)rf * rstride + x] = (TS)fn;
    r[(long)rc * rstride + x] = (TS)cn;
    if (stage == 4) r[(long)rp * rstride + x] = (TS)fn;
  }
}

// numNaNs (math.h:1524-1541): NaN test of the float-rounded value
template <typename TS>
__global__ void k_num_nans(const TS* __restrict__ f, long n, unsigned long long* count) {
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const float v = (float)f[x];
  if ((__float_as_uint(v) << 1) > 0xff000000u) atomicAdd(count, 1ull);
}

// ---------------------------------------------- constraint statistics --
// Deterministic FP64 reductions of the per-point constraint values: each
// block reduces a fixed chunk, a single block then combines the partials in
// block order.  Max uses `>` so that NaN values never win, like the
// reference's `if(fabs(C_val) > max_C)` (bssn_macros.h:581-585).
constexpr int kRedThreads = 256;
constexpr int kStats = 6;  // sum val, sum scale, sum scaled, sum val^2, max|val|, max|scaled|

__device__ __forceinline__ double red_max(double a, double b) { return b > a ? b : a; }

__global__ void k_cons_partial(const double* __restrict__ cons, long stride, long n, double* __restrict__ part) {
  __shared__ double sh[kRedThreads];
  const long per = (n + gridDim.x - 1) / gridDim.x;
  const long lo = (long)blockIdx.x * per, hi = lo + per < n ? lo + per : n;
  for (int c = 0; c < 5; ++c) {
    const double* val = cons + (long)(3 * c) * stride;
    const double* sca = val + stride;
    const double* scd = val + 2 * stride;
    double acc[kStats] = {0, 0, 0, 0, 0, 0};
    for (long x = lo + threadIdx.x; x < hi; x += kRedThreads) {
      const double v = val[x], sc = sca[x], sd = scd[x];
      acc[0] += v; acc[1] += sc; acc[2] += sd; acc[3] += v * v;
      acc[4] = red_max(acc[4], fabs(v)); acc[5] = red_max(acc[5], fabs(sd));
    }
    for (int q = 0; q < kStats; ++q) {
      sh[threadIdx.x] = acc[q];
      __syncthreads();
      for (int w = kRedThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
          sh[threadIdx.x] = q < 4 ? sh[threadIdx.x] + sh[threadIdx.x + w] : red_max(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
      }
      if (threadIdx.x == 0) part[((long)blockIdx.x * 5 + c) * kStats + q] = sh[0];
      __syncthreads();
    }
  }
}

// second pass: sum (val - mean)^2 and (scaled - mean_scaled)^2
__global__ void k_cons_var(const double* __restrict__ cons, long stride, long n, const double* __restrict__ means,
                           double* __restrict__ part) {
  __shared__ double sh[kRedThreads];
  const long per = (n + gridDim.x - 1) / gridDim.x;
  const long lo = (long)blockIdx.x * per, hi = lo + per < n ? lo + per : n;
  for (int c = 0; c < 5; ++c) {
    const double* val = cons + (long)(3 * c) * stride;
    const double* scd = val + 2 * stride;
    const double m0 = means[2 * c], m1 = means[2 * c + 1];
    double a0 = 0, a1 = 0;
    for (long x = lo + threadIdx.x; x < hi; x += kRedThreads) {
      const double d0 = val[x] - m0, d1 = scd[x] - m1;
      a0 += d0 * d0;
      a1 += d1 * d1;
    }
    for (int q = 0; q < 2; ++q) {
      sh[threadIdx.x] = q ? a1 : a0;
      __syncthreads();
      for (int w = kRedThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
      }
      if (threadIdx.x == 0) part[((long)blockIdx.x * 5 + c) * 2 + q] = sh[0];
      __syncthreads();
    }
  }
}

// one block: combine nb partial rows of width w in block order; columns whose
// index mod kStats is set in maxmask take the max, the others the sum
__global__ void k_combine(const double* __restrict__ part, int nb, int w, unsigned maxmask, double* __restrict__ out) {
  for (int col = threadIdx.x; col < w; col += blockDim.x) {
    const bool ismax = (maxmask >> (col % kStats)) & 1u;
    double acc = 0.0;
    for (int b = 0; b < nb; ++b) {
      const double v = part[(long)b * w + col];
      acc = ismax ? red_max(acc, v) : acc + v;
    }
    out[col] = acc;
  }
}

// ------------------------------------- constraints without allocation --
// Scratch of cg_constraints_noalloc: the rows of one chunk of kNaRows tiles
// (one row per workgroup of the reduced constraint mode), then the running
// totals of pass 0 (CONS_ROW_W) and pass 1 (CONS_ROW_W1) and the 10 means.
// One copy per device, held until the process ends; host code reaches it
// only through the address hipGetSymbolAddress returned for that device.
constexpr int kNaRows = 4096;
constexpr long kNaTot = (long)kNaRows * CONS_ROW_W;
constexpr long kNaScratch = kNaTot + 64;  // doubles: 0.98 MB
static_assert(CONS_ROW_W == 5 * kStats && CONS_ROW_W + 2 * CONS_ROW_W1 <= 64, "scratch layout");
constexpr int kNaDevices = 64;
double* g_scratch_dev[kNaDevices];  // per device, set once at context creation
std::mutex g_scratch_mu;

constexpr int kFoldThreads = 256;
// acc[col] = (first ? 0 : acc[col]) (+ or max) the rows[r * stride + col],
// r < n, in a fixed order: thread t takes rows t, t + 256, ..., then a fixed
// tree.  Columns 6q + 4, 6q + 5 of pass 0 are maxima of non-negative values.
__global__ void __launch_bounds__(kFoldThreads) k_rows_fold(const double* __restrict__ rows, int n, int stride,
                                                           int width, int pass, int first, double* __restrict__ acc) {
  __shared__ double sh[kFoldThreads];
  for (int col = 0; col < width; ++col) {
    const bool ismax = pass == 0 && (col % 6) >= 4;
    double a = 0.0;
    for (int r = threadIdx.x; r < n; r += kFoldThreads) {
      const double v = rows[(long)r * stride + col];
      a = ismax ? red_max(a, v) : a + v;
    }
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int w = kFoldThreads / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w)
        sh[threadIdx.x] = ismax ? red_max(sh[threadIdx.x], sh[threadIdx.x + w]) : sh[threadIdx.x] + sh[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) acc[col] = first ? sh[0] : (ismax ? red_max(acc[col], sh[0]) : acc[col] + sh[0]);
    __syncthreads();
  }
}

// means of value and scaled per constraint for the second pass
__global__ void k_cons_means(const double* __restrict__ acc0, double n, double* __restrict__ means) {
  const int q = threadIdx.x;
  if (q < 5) {
    means[2 * q] = acc0[6 * q] / n;
    means[2 * q + 1] = acc0[6 * q + 2] / n;
  }
}

}  // namespace

// the scratch array itself (file scope; see kNaScratch above)
static __device__ __align__(256) double scratch[kNaScratch];

namespace cg {

void* reg_ptr(cg_ctx* c, int f, int slot) {
  // all-FP32 mode: one storage-type allocation of all 18 fields
  if (c->met32) return (char*)c->regs + ((long)f * c->fstride + (long)slot * c->rstride) * (long)c->ts;
  if (is_metric(f)) return c->met + ((long)f * 4 + slot) * c->rstride;  // f = 0..6
  return (char*)c->regs + ((long)(f - 7) * c->fstride + (long)slot * c->rstride) * (long)c->ts;  // f = 7..17
}

int halo_for(int order, double ko) {
  int R = order / 2;
  if (ko != 0.0 && order == 2) return 2;
  if (ko != 0.0 && order == 8) return 5;
  return R;
}

// the single conditional wrap in the halo loader needs n > tile + H - 2
bool variant_fits(const cg_ctx* c, const StageVariant& v) {
  const int t[3] = {v.ti, v.tj, v.tk}, n[3] = {c->p.nx, c->p.ny, c->p.nz};
  for (int d = 0; d < 3; ++d)
    if (n[d] < v.halo || n[d] <= t[d] + v.halo - 2) return false;
  return true;
}

// the LDS-DMA twin of v applies to this grid: compiled, nz % 4 == 0 (16-byte
// groups never straddle the periodic wrap) and one wrap of the k halo
// (rounded up to 4) suffices; the 4-byte dust kernels address the three
// fluxes of a quantity from the first one, and the kernels with every field
// in 4 bytes Gamma^1 and Gamma^3 from K and Gamma^2, with a 32-bit byte
// offset; their halo offsets are formed with 24-bit multiplies (nx * ny < 2^24)
bool dma_fits(const cg_ctx* c, const StageVariant& v) {
  if (!v.launch_dma) return false;
  if ((long long)c->p.nx * c->p.ny >= (1ll << 24)) return false;
  const unsigned long long rs4 = 4ull * (unsigned long long)c->rstride;
  if (v.dust && v.prec != CG_FP64 && 3ull * rs4 > 0xFFFFFFFFull) return false;
  if ((v.prec == CG_FP32_ALL || v.prec == CG_FP32_FIX) && 5ull * rs4 > 0xFFFFFFFFull) return false;
  const int kh = (v.halo + 3) / 4 * 4;
  return c->p.nz % 4 == 0 && c->p.nz >= kh && c->p.nz > v.tk + kh - 2;
}

// tile order 3 (Morton columns) needs power-of-two tile counts along i and j
bool morton_fits(int ntx, int nty) { return ntx > 0 && nty > 0 && !(ntx & (ntx - 1)) && !(nty & (nty - 1)); }

bool alpha_live(const cg_ctx* c) { return c->p.lapse != CG_LAPSE_STATIC || c->alpha_nonzero; }

}  // namespace cg

static RegTable reg_table(cg_ctx* c) {
  RegTable t;
  for (int f = 0; f < NF; ++f)
    for (int s = 0; s < 4; ++s) t.r[f][s] = reg_ptr(c, f, s);
  mq_scales(c, t.ms, t.mi);
  t.sat = c->mq_sat;
  return t;
}

static bool variant_matches(const cg_ctx* c, const StageVariant& v) {
  return v.prec == c->prec && v.order == c->p.stencil_order && v.halo == c->H && v.live_alpha == alpha_live(c) &&
         v.dust == (c->p.matter == CG_MATTER_DUST) && variant_fits(c, v);
}

// default: the requested tile, else the 4x8x8 tile with the register budget
// of 3 (FP32) / 2 (FP64) waves per SIMD when compiled, else any fitting tile
static const StageVariant* pick_variant(const cg_ctx* c) {
  const StageVariant* best = nullptr;
  int best_score = -1;
  for (const auto& v : stage_variants()) {
    if (!variant_matches(c, v)) continue;
    int score = 0;
    if (c->req_ti) {
      if (v.ti != c->req_ti || v.tj != c->req_tj || v.tk != c->req_tk) continue;
      score += 100;
    } else if (v.ti == 4 && v.tj == 8 && v.tk == 8) {
      score += 100;
    }
    score += (v.wpe == (c->prec == CG_FP64 ? 2 : 3)) ? 10 : (v.wpe == 1 ? 5 : 0);
    if (score > best_score) { best = &v; best_score = score; }
  }
  return best;
}

// The definitions below take C linkage from their declarations in cosmo_gpu.h.

const char* cg_last_error(const cg_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int cg_create(const cg_params* params, cg_ctx** out) {
  if (!params || !out) return CG_ERR_ARG;
  *out = nullptr;
  const cg_params& p = *params;
  if (p.nx <= 0 || p.ny <= 0 || p.nz <= 0) return CG_ERR_ARG;
  // The stage kernels address every array as a uniform base plus a 32-bit
  // byte offset (element offset * element size, unsigned), and the FP64
  // registers (the conformal metric in every precision) have 8-byte elements:
  // the largest element offset times 8 must stay below 2^32.
  if ((long long)p.nx * p.ny * p.nz > (1ll << 29)) return CG_ERR_ARG;
  if (p.stencil_order != 2 && p.stencil_order != 4 && p.stencil_order != 6 && p.stencil_order != 8) return CG_ERR_ARG;
  if (p.precision != CG_FP64 && p.precision != CG_FP32_SPLIT && p.precision != CG_FP32_ALL &&
      p.precision != CG_FP32_FIX)
    return CG_ERR_ARG;
  // all-FP32 and fixed-point storage: the shift-free gauges the stage kernel
  // forms by itself, on single grids; BSSN with static dust or vacuum (and,
  // with the fixed-point metric, the dust fluid)
  if ((p.precision == CG_FP32_ALL || p.precision == CG_FP32_FIX) &&
      (p.shift != CG_SHIFT_NONE || p.lapse == LAPSE_TEST_K_DRIVER))
    return CG_ERR_ARG;
  if (p.precision == CG_FP32_ALL && p.matter == CG_MATTER_DUST) return CG_ERR_ARG;
  if (p.matter != CG_MATTER_NONE && p.matter != CG_MATTER_STATIC && p.matter != CG_MATTER_DUST) return CG_ERR_ARG;
  if (p.lapse < 0 || p.lapse >= LAPSE_N) return CG_ERR_ARG;
  if (p.shift < 0 || p.shift >= SHIFT_N) return CG_ERR_ARG;
  if (p.shift != CG_SHIFT_NONE && p.matter == CG_MATTER_DUST) return CG_ERR_ARG;  // shift build: BSSN + static dust
  if (!(p.dx > 0) || !std::isfinite(p.dt)) return CG_ERR_ARG;
  cg_ctx* c = new cg_ctx();
  c->p = p;
  c->prec = p.precision;
  c->ts = p.precision == CG_FP64 ? 8 : 4;
  c->met32 = p.precision == CG_FP32_ALL || p.precision == CG_FP32_FIX;
  c->mq = p.precision == CG_FP32_FIX;
  c->R = p.stencil_order / 2;
  c->H = halo_for(p.stencil_order, p.ko_coeff);
  c->lambda = p.lambda;
  if (p.tile_i > 0) { c->req_ti = p.tile_i; c->req_tj = p.tile_j; c->req_tk = p.tile_k; }
  c->npts = (long)p.nx * p.ny * p.nz;
  const long pad = ((c->npts + 63) / 64) * 64;  // 256/512-byte aligned registers
  c->rstride = pad;
  c->fstride = 4 * pad;
  *out = c;
  HIPCHK(c, hipSetDevice(p.device));
  if (p.device >= 0 && p.device < kNaDevices) {
    // the device address of this device's copy of `scratch`, once per device
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[p.device]) {
      void* ptr = nullptr;
      HIPCHK(c, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
      g_scratch_dev[p.device] = (double*)ptr;
    }
  }
  HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  const int nreg = c->met32 ? NF : NF - 7;  // storage-type fields
  HIPCHK(c, hipMalloc(&c->regs, (size_t)nreg * 4 * pad * c->ts));
  HIPCHK(c, hipMemsetAsync(c->regs, 0, (size_t)nreg * 4 * pad * c->ts, c->stream));
  if (!c->met32) {
    HIPCHK(c, hipMalloc((void**)&c->met, (size_t)7 * 4 * pad * sizeof(double)));
    HIPCHK(c, hipMemsetAsync(c->met, 0, (size_t)7 * 4 * pad * sizeof(double), c->stream));
  }
  HIPCHK(c, hipMalloc(&c->DIFFD, (size_t)pad * c->ts));
  HIPCHK(c, hipMemsetAsync(c->DIFFD, 0, (size_t)pad * c->ts, c->stream));
  HIPCHK(c, hipMalloc(&c->nan_dev, sizeof(unsigned)));
  HIPCHK(c, hipMemsetAsync(c->nan_dev, 0, sizeof(unsigned), c->stream));
  if (c->mq) {
    HIPCHK(c, hipMalloc((void**)&c->mq_sat, sizeof(unsigned)));
    HIPCHK(c, hipMemsetAsync(c->mq_sat, 0, sizeof(unsigned), c->stream));
  }
  // partial rows: one per tile of any compiled stage variant, or one per
  // 256-point block of the elementwise kernels
  long rows = (c->npts + 255) / 256;
  for (const auto& v : stage_variants()) {
    const long nt = (long)((p.nx + v.ti - 1) / v.ti) * ((p.ny + v.tj - 1) / v.tj) * ((p.nz + v.tk - 1) / v.tk);
    rows = nt > rows ? nt : rows;
  }
  c->ext_rows = rows;
  HIPCHK(c, hipMalloc((void**)&c->ext_part, (size_t)rows * EXT_W * sizeof(double)));
  HIPCHK(c, hipMalloc((void**)&c->ext_mid, (size_t)(kExtMid + 1) * EXT_W * sizeof(double)));
  HIPCHK(c, hipMalloc((void**)&c->ext, 4 * sizeof(double)));
  HIPCHK(c, hipMemsetAsync(c->ext, 0, 4 * sizeof(double), c->stream));
  HIPCHK(c, hipMalloc((void**)&c->nan_cnt, sizeof(unsigned long long)));
  c->nxf = n_xfields(p.shift);
  if (c->nxf) {
    HIPCHK(c, hipMalloc((void**)&c->xreg, (size_t)c->nxf * 4 * pad * sizeof(double)));
    HIPCHK(c, hipMemsetAsync(c->xreg, 0, (size_t)c->nxf * 4 * pad * sizeof(double), c->stream));
  }
  if (p.matter == CG_MATTER_DUST) {
    HIPCHK(c, hipMalloc(&c->dreg, (size_t)NDUST * 4 * pad * c->ts));
    HIPCHK(c, hipMemsetAsync(c->dreg, 0, (size_t)NDUST * 4 * pad * c->ts, c->stream));
    HIPCHK(c, hipMalloc(&c->flux, (size_t)NFLUX * pad * c->ts));
    HIPCHK(c, hipMalloc((void**)&c->rho_src, (size_t)pad * sizeof(double)));
    HIPCHK(c, hipMemsetAsync(c->rho_src, 0, (size_t)pad * sizeof(double), c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CG_OK;
}

int cg_destroy(cg_ctx* c) {
  if (!c) return CG_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& e : c->ev) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
  if (c->regs) (void)hipFree(c->regs);
  if (c->DIFFD) (void)hipFree(c->DIFFD);
  if (c->met) (void)hipFree(c->met);
  if (c->cons) (void)hipFree(c->cons);
  if (c->cons_part) (void)hipFree(c->cons_part);
  if (c->obuf) (void)hipFree(c->obuf);
  if (c->nan_dev) (void)hipFree(c->nan_dev);
  if (c->ext) (void)hipFree(c->ext);
  if (c->ext_part) (void)hipFree(c->ext_part);
  if (c->ext_mid) (void)hipFree(c->ext_mid);
  if (c->nan_cnt) (void)hipFree(c->nan_cnt);
  if (c->dreg) (void)hipFree(c->dreg);
  if (c->flux) (void)hipFree(c->flux);
  if (c->rho_src) (void)hipFree(c->rho_src);
  if (c->xreg) (void)hipFree(c->xreg);
  if (c->mq_sat) (void)hipFree(c->mq_sat);
  mq_free(c);
  bardeen_free(c);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return CG_OK;
}

// a device pointer that is a register slot of a fixed-point metric field:
// its field and slot
static bool mq_register(const cg_ctx* c, const void* d, int* f, int* slot) {
  if (!c->mq || !d || d < c->regs) return false;
  const long off = ((const char*)d - (const char*)c->regs) / (long)c->ts;
  if (off >= 7 * c->fstride) return false;
  *f = (int)(off / c->fstride);
  *slot = (int)((off % c->fstride) / c->rstride);
  return true;
}

// resolve a reference field-map key to a device pointer and element size
static void* resolve(cg_ctx* c, const char* name, bool for_write, int* err, size_t* esz) {
  *err = 0;
  *esz = c->ts;
  std::string s(name ? name : "");
  if (s == "DIFFD_a") return c->DIFFD;
  if (s == "DIFFr_a" || s == "DIFFS_a") {  // sources (read only)
    if (for_write) { *err = 1; return nullptr; }
    *esz = 8;
    // dust: the source of the last stage, or of the state cg_constraints
    // evaluated (what initDustStep leaves in DIFFr_a, sims/dust.cc:53-63)
    if (s == "DIFFr_a" && c->rho_src) return c->rho_src;
    // otherwise the sources of the state cg_constraints evaluated
    // (initStaticStep's clearSrc + addBSSNSrc, sims/static.cc:74-82)
    if (!c->cons_valid) { *err = 3; return nullptr; }
    return c->cons + (long)(s == "DIFFr_a" ? CONS_RHO : CONS_SRCS) * c->rstride;
  }
  if (s.compare(0, 8, "Bardeen_") == 0) {  // Bardeen::Bardeen's field-map entries (bardeen.h:127-156)
    *esz = 8;
    int e4 = 0;
    const double* p = bardeen_field(c, s.substr(8), &e4);
    if (for_write) { *err = 1; return nullptr; }
    *err = e4;
    return (void*)p;
  }
  if (s == "ricci_a" || s == "AijAij_a") {
    if (for_write || !c->cons_valid) { *err = 3; return nullptr; }
    *esz = 8;
    return c->cons + (long)(s == "ricci_a" ? CONS_RICCI : CONS_AIJAIJ) * c->rstride;
  }
  if (s.size() < 3 || s[s.size() - 2] != '_') { *err = 1; return nullptr; }
  const std::string base = s.substr(0, s.size() - 2);
  const char r = s.back();
  auto slot_of = [&](char rr) {
    return rr == 'p' ? c->sP : rr == 'f' ? (c->a_is_p ? c->sP : c->sF) : rr == 'a' ? (c->a_is_p ? c->sP : c->sA)
           : rr == 'c' ? c->sC : -1;
  };
  static const char* kX[NXF] = {"beta1", "beta2", "beta3", "expN", "auxB1", "auxB2", "auxB3"};
  for (int q = 0; q < c->nxf; ++q) {
    if (base != kX[q]) continue;
    const int slot = slot_of(r);
    if (slot < 0) { *err = 1; return nullptr; }
    if (for_write && c->a_is_p && (r == 'a' || r == 'f')) { *err = 2; return nullptr; }
    *esz = 8;
    return c->xreg + ((long)q * 4 + slot) * c->rstride;
  }
  static const char* kDust[NDUST] = {"dust_D", "dust_S1", "dust_S2", "dust_S3"};
  for (int q = 0; q < NDUST && c->dreg; ++q) {
    if (base != kDust[q]) continue;
    const int slot = slot_of(r);
    if (slot < 0) { *err = 1; return nullptr; }
    if (for_write && c->a_is_p && (r == 'a' || r == 'f')) { *err = 2; return nullptr; }
    return (char*)c->dreg + ((long)q * 4 + slot) * c->rstride * (long)c->ts;
  }
  for (int f = 0; f < NF; ++f) {
    if (base != kFieldNames[f]) continue;
    int slot;
    if (r == 'p') slot = c->sP;
    else if (r == 'f') slot = c->a_is_p ? c->sP : c->sF;
    else if (r == 'a') slot = c->a_is_p ? c->sP : c->sA;
    else if (r == 'c') slot = c->sC;
    else { *err = 1; return nullptr; }
    if (for_write && c->a_is_p && (r == 'a' || r == 'f')) { *err = 2; return nullptr; }
    if (is_metric(f) && !c->met32) *esz = 8;
    return reg_ptr(c, f, slot);
  }
  *err = 1;
  return nullptr;
}

namespace cg {
const void* field_ptr(cg_ctx* c, const char* name, size_t* esz, int* err) {
  const void* d = resolve(c, name, false, err, esz);
  int f, s;
  if (mq_register(c, d, &f, &s)) {  // words, not values: read through cg_field_download
    *err = 5;
    return nullptr;
  }
  return d;
}
}  // namespace cg

int cg_field_upload(cg_ctx* c, const char* name, const double* host, size_t n) {
  if (!c || !host) return CG_ERR_ARG;
  if ((long)n != c->npts) return fail(c, CG_ERR_ARG, "size mismatch");
  int e;
  size_t esz;
  void* d = resolve(c, name, true, &e, &esz);
  if (!d) return fail(c, e == 2 ? CG_ERR_STATE : CG_ERR_ARG,
                      e == 2 ? "_a/_f alias _p after a fused step; write _p" : std::string("unknown field ") + name);
  if (int rc = ext_resolve(c)) return rc;  // the averages of the state before this write
  c->ext_fresh = false;
  if (!strncmp(name, "DIFFalpha_", 10))
    for (size_t x = 0; x < n; ++x)
      if (host[x] != 0.0) { c->alpha_nonzero = true; break; }
  int mf, ms;
  if (mq_register(c, d, &mf, &ms)) {
    // fixed-point metric: the FP64 values through the device, scaled there
    if (!c->mq_tmp) HIPCHK(c, hipMalloc((void**)&c->mq_tmp, (size_t)c->rstride * sizeof(double)));
    HIPCHK(c, hipMemcpyAsync(c->mq_tmp, host, n * 8, hipMemcpyHostToDevice, c->stream));
    return mq_store(c, mf, ms, c->mq_tmp);
  }
  if (esz == 8) {
    HIPCHK(c, hipMemcpyAsync(d, host, n * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  } else {
    std::vector<float> tmp(n);
    for (size_t x = 0; x < n; ++x) tmp[x] = (float)host[x];
    HIPCHK(c, hipMemcpyAsync(d, tmp.data(), n * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return CG_OK;
}

int cg_field_download(cg_ctx* c, const char* name, double* host, size_t n) {
  if (!c || !host) return CG_ERR_ARG;
  if ((long)n != c->npts) return fail(c, CG_ERR_ARG, "size mismatch");
  int e;
  size_t esz;
  void* d = resolve(c, name, false, &e, &esz);
  if (!d) return fail(c, e >= 3 ? CG_ERR_STATE : CG_ERR_ARG,
                      e == 3 ? std::string(name) + " is filled by cg_constraints"
                      : e == 4 ? std::string(name) + " is filled by cg_bardeen" : std::string("unknown field ") + name);
  int mf, ms;
  if (mq_register(c, d, &mf, &ms)) {
    if (!c->mq_tmp) HIPCHK(c, hipMalloc((void**)&c->mq_tmp, (size_t)c->rstride * sizeof(double)));
    if (int rc = mq_load(c, mf, ms, c->mq_tmp)) return rc;
    HIPCHK(c, hipMemcpyAsync(host, c->mq_tmp, n * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CG_OK;
  }
  if (esz == 8) {
    HIPCHK(c, hipMemcpyAsync(host, d, n * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  } else {
    std::vector<float> tmp(n);
    HIPCHK(c, hipMemcpyAsync(tmp.data(), d, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (size_t x = 0; x < n; ++x) host[x] = (double)tmp[x];
  }
  return CG_OK;
}

int cg_frw_set(cg_ctx* c, double phi, double K) {
  if (!c) return CG_ERR_ARG;
  c->frw.set(phi, K);
  c->ext_fresh = false;
  return CG_OK;
}
int cg_frw_add_fluid(cg_ctx* c, double rho, double w) {
  if (!c) return CG_ERR_ARG;
  c->frw.add_fluid(rho, w);
  return CG_OK;
}
int cg_frw_get(const cg_ctx* c, double out[4]) {
  if (!c || !out) return CG_ERR_ARG;
  out[0] = c->frw.g_phi; out[1] = c->frw.g_K; out[2] = c->frw.g_rho; out[3] = c->frw.g_S;
  return CG_OK;
}
int cg_set_lambda(cg_ctx* c, double lambda) {
  if (!c) return CG_ERR_ARG;
  c->lambda = lambda;
  return CG_OK;
}

int cg_set_dust_background(cg_ctx* c, int on) {
  if (!c) return CG_ERR_ARG;
  if (int rc = ext_resolve(c)) return rc;
  c->dust_bg = on != 0;
  return CG_OK;
}

int cg_set_tile_order(cg_ctx* c, int order) {
  if (!c || order < 0 || order > 5) return CG_ERR_ARG;
  c->tile_order = order;
  return CG_OK;
}

int cg_set_fused_accumulator(cg_ctx* c, int keep) {
  if (!c) return CG_ERR_ARG;
  c->keep_accumulator = keep != 0;
  return CG_OK;
}

int cg_set_dma(cg_ctx* c, int on) {
  if (!c) return CG_ERR_ARG;
  c->dma_allowed = on != 0;
  return CG_OK;
}

int cg_set_gauge_params(cg_ctx* c, double gd_c, double dw_mu_l, double dw_p) {
  if (!c) return CG_ERR_ARG;
  c->gd_c = gd_c; c->dw_mu_l = dw_mu_l; c->dw_p = dw_p;
  return CG_OK;
}

int cg_set_shift_params(cg_ctx* c, int gauge_wave_dir, double gd_eta, double dw_mu_s, double k_driver_coeff) {
  if (!c || gauge_wave_dir < 1 || gauge_wave_dir > 3) return CG_ERR_ARG;
  c->gauge_wave_dir = gauge_wave_dir;
  c->gd_eta = gd_eta; c->dw_mu_s = dw_mu_s; c->k_driver = k_driver_coeff;
  return CG_OK;
}

namespace cg {
StageArgs make_args(cg_ctx* c, int stage, int rin, int rp, int rf, int rout, int normalize) {
  StageArgs a{};
  a.met = c->met;
  // all-FP32 mode: regs points at field 7 so that the kernel's
  // regs + (f - 7)*4*rstride addresses the metric fields too
  a.regs = c->met32 ? (void*)((char*)c->regs + 7 * c->fstride * (long)c->ts) : c->regs;
  a.rstride = c->rstride;
  mq_scales(c, a.msc, a.minv);
  a.mq_sat = c->mq_sat;
  a.s_in = rin; a.s_p = rp; a.s_f = rf; a.s_out = rout;
  a.s_in_norm = -1; a.s_norm_out = -1;
  a.incr = 0; a.s_a2 = -1;
  a.DIFFD = c->DIFFD;
  a.nx = c->p.nx; a.ny = c->p.ny; a.nz = c->p.nz;
  a.tile_order = c->tile_order;

  a.stage = stage;
  a.normalize = normalize;
  a.matter = c->p.matter;
  a.lapse = c->p.lapse;
  a.use_ko = c->p.ko_coeff != 0.0 && (c->p.stencil_order == 2 || c->p.stencil_order == 8);
  a.use_H = (c->p.a_adj_amp != 0.0 || c->p.k_damping_amp != 0.0);
  a.inv_dx = 1.0 / c->p.dx;
  a.inv_dx2 = 1.0 / (c->p.dx * c->p.dx);
  for (int f = 0; f < 7; ++f) {
    a.msd1[f] = (float)(a.msc[f] * a.inv_dx);
    a.msd2[f] = (float)(a.msc[f] * a.inv_dx2);
  }
  a.dt = c->p.dt;
  // RK4Register weights (RK4Register.h:129-176)
  const double dt = c->p.dt;
  a.cf = (stage == 1 || stage == 4) ? dt / 6.0 : dt / 3.0;
  a.ca = (stage == 3) ? dt : dt / 2.0;
  a.phi_F = c->frw.g_phi; a.K_F = c->frw.g_K; a.rho_F = c->frw.g_rho; a.S_F = c->frw.g_S;
  a.ext_in = c->ext;
  a.ext_part = lapse_needs_kavg(c->p.lapse) ? c->ext_part : nullptr;
  a.gd_c = c->gd_c; a.dw_mu_l = c->dw_mu_l; a.dw_p = c->dw_p;
  a.phi_F_next = c->frw.g_phi;
  a.lambda = c->lambda;
  a.dust_bg = c->dust_bg ? 1 : 0;
  a.dreg = c->dreg;
  a.flux = c->flux;
  a.s_lag = a.s_lag_met = rin;
  a.lag_mode = 0;
  a.phi_F_lag = c->frw.g_phi;
  a.rho_out = c->rho_src;
  const double dx = c->p.dx, ko = c->p.ko_coeff;
  // KO_dissipation_Q (math.h:36-72): ko*dx^3/64 * sum/dx^4 ; -ko*dx^9/1024 * sum/dx^10
  a.ko_pref = c->p.stencil_order == 2 ? ko / (64.0 * dx) : c->p.stencil_order == 8 ? -ko / (1024.0 * dx) : 0.0;
  a.a_adj = c->p.a_adj_amp;
  a.k_damp = c->p.k_damping_amp;
  a.H_frw = 2.0 / 3.0 * c->frw.g_K * c->frw.g_K - 16.0 * M_PI * c->frw.g_rho;
  a.nan_count = c->nan_dev;
  return a;
}
}  // namespace cg

static int launch(cg_ctx* c, StageArgs a) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (c->timing) {
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    HIPCHK(c, hipEventRecord(e0, c->stream));
  }
  const bool live = alpha_live(c);
  if (!c->var || c->var->live_alpha != live) {
    c->var = pick_variant(c);
    if (!c->var) return fail(c, CG_ERR_ARG, "no stage kernel variant fits this grid/order");
  }
  a.ntx = (c->p.nx + c->var->ti - 1) / c->var->ti;
  a.nty = (c->p.ny + c->var->tj - 1) / c->var->tj;
  a.ntz = (c->p.nz + c->var->tk - 1) / c->var->tk;
  if (a.tile_order == 3 && !morton_fits(a.ntx, a.nty)) a.tile_order = 0;
  a.per_xcd = (a.ntx * a.nty * a.ntz + 7) / 8;
  c->last_rows = (long)a.ntx * a.nty * a.ntz;
  if (c->last_rows > c->ext_rows) return fail(c, CG_ERR_STATE, "extra-field partial buffer too small");
  // stage-1 normalisation needs the register path (it transforms the metric on its way into LDS)
  const bool dma = c->dma_allowed && c->dma_pick && !a.normalize && dma_fits(c, *c->var);
  if (a.s_norm_out >= 0 && !dma) return fail(c, CG_ERR_STATE, "normalised copy without DMA staging");
  int rc = (dma ? c->var->launch_dma : c->var->launch)(a, c->stream);
  if (rc != 0) return fail(c, CG_ERR_HIP, std::string("stage launch: ") + hipGetErrorString((hipError_t)rc));
  if (c->timing) {
    HIPCHK(c, hipEventRecord(e1, c->stream));
    c->ev.emplace_back(e0, e1);
  }
  return CG_OK;
}

// combine `rows` partial rows into c->ext (two deterministic passes)
static int ext_combine(cg_ctx* c, long rows, bool keep_rho) {
  const double n = (double)c->npts;
  hipLaunchKernelGGL(k_ext_reduce, dim3(kExtMid), dim3(kExtThreads), 0, c->stream, c->ext_part, rows, c->ext_mid,
                     (double*)nullptr, n, 0);
  hipLaunchKernelGGL(k_ext_reduce, dim3(1), dim3(kExtThreads), 0, c->stream, c->ext_mid, (long)kExtMid,
                     c->ext_mid + (long)kExtMid * EXT_W, c->ext, n, keep_rho ? 1 : 0);
  HIPCHK(c, hipGetLastError());
  return CG_OK;
}

// setExtraFieldData of _p alone (rho_avg kept), when no stage has produced it
static int ext_from_state(cg_ctx* c) {
  const int bs = 256;
  const long nb = (c->npts + bs - 1) / bs;
  const void* phi = reg_ptr(c, PHI, c->sP);
  double ms[7], mi[7];
  mq_scales(c, ms, mi);
  if (c->prec == CG_FP64)
    hipLaunchKernelGGL((k_ext_state<double, double>), dim3(nb), dim3(bs), 0, c->stream, (const double*)phi,
                       (const double*)reg_ptr(c, KK, c->sP), c->npts, c->ext_part, c->frw.g_phi, 1.0);
  else if (c->mq)
    hipLaunchKernelGGL((k_ext_state<float, int>), dim3(nb), dim3(bs), 0, c->stream, (const int*)phi,
                       (const float*)reg_ptr(c, KK, c->sP), c->npts, c->ext_part, c->frw.g_phi, ms[PHI]);
  else if (c->met32)
    hipLaunchKernelGGL((k_ext_state<float, float>), dim3(nb), dim3(bs), 0, c->stream, (const float*)phi,
                       (const float*)reg_ptr(c, KK, c->sP), c->npts, c->ext_part, c->frw.g_phi, 1.0);
  else
    hipLaunchKernelGGL((k_ext_state<float, double>), dim3(nb), dim3(bs), 0, c->stream, (const double*)phi,
                       (const float*)reg_ptr(c, KK, c->sP), c->npts, c->ext_part, c->frw.g_phi, 1.0);
  HIPCHK(c, hipGetLastError());
  int rc = ext_combine(c, nb, true);
  if (!rc) c->ext_fresh = true;
  return rc;
}

// form the pending K4Finalize averages of the last fused step
namespace cg {
int ext_resolve(cg_ctx* c) {
  if (!c->ext_lazy) return CG_OK;
  c->ext_lazy = false;
  const int bs = 256;
  const long nb = (c->npts + bs - 1) / bs;
  const void* phi = reg_ptr(c, PHI, c->sP);
  const void* phi_in = reg_ptr(c, PHI, c->ext_lazy_in);
  double ms[7], mi[7];
  mq_scales(c, ms, mi);
  if (c->prec == CG_FP64)
    hipLaunchKernelGGL((k_ext_post<double, double>), dim3(nb), dim3(bs), 0, c->stream, (const double*)phi,
                       (const double*)reg_ptr(c, KK, c->sP), (const double*)phi_in, c->npts, c->ext_part,
                       c->ext_lazy_src, c->ext_lazy_phiF, 1.0);
  else if (c->mq)
    hipLaunchKernelGGL((k_ext_post<float, int>), dim3(nb), dim3(bs), 0, c->stream, (const int*)phi,
                       (const float*)reg_ptr(c, KK, c->sP), (const int*)phi_in, c->npts, c->ext_part,
                       c->ext_lazy_src, c->ext_lazy_phiF, ms[PHI]);
  else if (c->met32)
    hipLaunchKernelGGL((k_ext_post<float, float>), dim3(nb), dim3(bs), 0, c->stream, (const float*)phi,
                       (const float*)reg_ptr(c, KK, c->sP), (const float*)phi_in, c->npts, c->ext_part,
                       c->ext_lazy_src, c->ext_lazy_phiF, 1.0);
  else
    hipLaunchKernelGGL((k_ext_post<float, double>), dim3(nb), dim3(bs), 0, c->stream, (const double*)phi,
                       (const float*)reg_ptr(c, KK, c->sP), (const double*)phi_in, c->npts, c->ext_part,
                       c->ext_lazy_src, c->ext_lazy_phiF, 1.0);
  HIPCHK(c, hipGetLastError());
  return ext_combine(c, nb, false);
}
}  // namespace cg

template <typename TS, typename MT>
static int step_init_t(cg_ctx* c) {
  const int bs = 256;
  const long nb = (c->npts + bs - 1) / bs;
  hipLaunchKernelGGL((k_step_init<TS, MT>), dim3(nb), dim3(bs), 0, c->stream, reg_table(c), c->sP, c->sA, c->sF, c->npts,
                     c->p.normalize_metric, c->ext_part, c->frw.g_phi);
  HIPCHK(c, hipGetLastError());
  int rc = ext_combine(c, nb, true);
  if (!rc) c->ext_fresh = true;
  return rc;
}

// Dust::stepInit's register part (dust.cc:65-73) into slot ra
static int dust_reg_init(cg_ctx* c, int rp, int ra, int rf) {
  const int bs = 256;
  const long nb = (c->npts + bs - 1) / bs;
  if (c->prec == CG_FP64)
    hipLaunchKernelGGL(k_dust_reg_init<double>, dim3(nb), dim3(bs), 0, c->stream, (double*)c->dreg, c->rstride, rp,
                       ra, rf, c->npts);
  else
    hipLaunchKernelGGL(k_dust_reg_init<float>, dim3(nb), dim3(bs), 0, c->stream, (float*)c->dreg, c->rstride, rp,
                       ra, rf, c->npts);
  HIPCHK(c, hipGetLastError());
  return CG_OK;
}

// fluxes of the dust fluid at every point of register slot `slot`
static int dust_flux(cg_ctx* c, int slot, int normalize, int mslot = -1) {
  if (mslot < 0) mslot = slot;
  const int bs = 256;
  const long nb = (c->npts + bs - 1) / bs;
  MetScale ms, mi;
  for (int f = 0; f < 7; ++f) ms.s[f] = 1.0;
  if (c->mq) mq_scales(c, ms.s, mi.s);
  if (c->prec == CG_FP64)
    hipLaunchKernelGGL((k_dust_flux<double, double>), dim3(nb), dim3(bs), 0, c->stream, (const double*)c->met, ms,
                       (const double*)reg_ptr(c, ALPHA, 0), (const double*)c->dreg, c->rstride, slot, mslot, normalize,
                       c->frw.g_phi, (double*)c->flux, c->npts);
  else if (c->mq)
    hipLaunchKernelGGL((k_dust_flux<float, int>), dim3(nb), dim3(bs), 0, c->stream, (const int*)reg_ptr(c, 0, 0), ms,
                       (const float*)reg_ptr(c, ALPHA, 0), (const float*)c->dreg, c->rstride, slot, mslot, normalize,
                       c->frw.g_phi, (float*)c->flux, c->npts);
  else
    hipLaunchKernelGGL((k_dust_flux<float, double>), dim3(nb), dim3(bs), 0, c->stream, (const double*)c->met, ms,
                       (const float*)reg_ptr(c, ALPHA, 0), (const float*)c->dreg, c->rstride, slot, mslot, normalize,
                       c->frw.g_phi, (float*)c->flux, c->npts);
  HIPCHK(c, hipGetLastError());
  return CG_OK;
}

int cg_step_init(cg_ctx* c) {
  if (!c) return CG_ERR_ARG;
  c->ext_lazy = false;  // superseded by this stepInit's averages
  if (c->a_is_p) { c->a_is_p = false; c->sA = 2; c->sC = 3; }
  if (c->dreg) {
    c->dust_rk = 0;
    if (int rc = dust_reg_init(c, c->sP, c->sA, c->sF)) return rc;
  }
  if (int rc = gauge_step_init(c)) return rc;
  return c->prec == CG_FP64 ? step_init_t<double, double>(c)
         : c->mq             ? step_init_t<float, int>(c)
         : c->met32          ? step_init_t<float, float>(c)
                             : step_init_t<float, double>(c);
}

int cg_rk_evolve(cg_ctx* c) {
  if (!c) return CG_ERR_ARG;
  if (c->a_is_p) return fail(c, CG_ERR_STATE, "RKEvolve before stepInit");
  if (c->mq)
    return fail(c, CG_ERR_STATE, "the fixed-point metric holds states, not right-hand sides: step with cg_rk4_steps");
  StageArgs a = make_args(c, 0, c->sA, c->sP, c->sF, c->sC, 0);
  if (c->dreg) {
    // Dust::RKEvolve (dust.cc:160-174) with the sources addBSSNSrc formed
    // from the previous RKEvolve's derived fields: after swap_a_c that
    // stage's input sits in _c, read at each point before it is overwritten
    if (int rc = dust_flux(c, c->sA, 0)) return rc;
    if (c->dust_rk > 0) {
      a.lag_mode = 1;
      a.s_lag = a.s_lag_met = c->sC;
      a.phi_F_lag = c->dust_lag_phiF;
    }
    c->dust_lag_phiF = c->frw.g_phi;
    ++c->dust_rk;
  }
  if (int rc = launch(c, a)) return rc;
  // the shift build's terms (and the lapses that read other fields'
  // complete right-hand sides) on top of the stage kernel's _c
  return gauge_pass(c) ? gauge_rhs(c, c->sA, c->sC) : CG_OK;
}

namespace cg {
int rhs_at(cg_ctx* c, int rin, int rout) {
  StageArgs a = make_args(c, 0, rin, c->sP, c->sF, rout, 0);
  if (c->dreg && dust_flux(c, rin, 0)) return CG_ERR_HIP;  // sources from this state itself (lag_mode 0)
  if (int rc = launch(c, a)) return rc;
  return gauge_pass(c) ? gauge_rhs(c, rin, rout) : CG_OK;
}
}  // namespace cg

template <typename TS, typename MT>
static int finalize_t(cg_ctx* c, int n, const SrcArgs& src) {
  const int bs = 256;
  const long nb = (c->npts + bs - 1) / bs;
  hipLaunchKernelGGL((k_finalize<TS, MT>), dim3(nb), dim3(bs), 0, c->stream, reg_table(c), c->sP, c->sC, c->sF, c->sA,
                     c->npts, n, c->p.dt, c->ext_part, src, c->frw.g_phi);
  HIPCHK(c, hipGetLastError());
  return ext_combine(c, nb, false);
}

int cg_k_finalize(cg_ctx* c, int n) {
  if (!c || n < 1 || n > 4) return CG_ERR_ARG;
  if (c->a_is_p) return fail(c, CG_ERR_STATE, "K-finalize before stepInit");
  if (c->mq)
    return fail(c, CG_ERR_STATE, "the fixed-point metric holds states, not right-hand sides: step with cg_rk4_steps");
  c->ext_lazy = false;
  // the finished stage's sources were formed with the FRW values before this step
  const SrcArgs src{c->DIFFD, c->p.matter, c->frw.g_phi, c->frw.g_rho, c->lambda, c->rho_src};
  c->frw.stage(n, c->p.dt);  // bssn.cc:269: FRW stage first
  if (c->dreg) {
    const int bs = 256;
    const long nb = (c->npts + bs - 1) / bs;
    if (c->prec == CG_FP64)
      hipLaunchKernelGGL(k_dust_finalize<double>, dim3(nb), dim3(bs), 0, c->stream, (double*)c->dreg, c->rstride,
                         c->sP, c->sC, c->sF, c->npts, n, c->p.dt);
    else
      hipLaunchKernelGGL(k_dust_finalize<float>, dim3(nb), dim3(bs), 0, c->stream, (float*)c->dreg, c->rstride,
                         c->sP, c->sC, c->sF, c->npts, n, c->p.dt);
    HIPCHK(c, hipGetLastError());
  }
  if (int rc = gauge_finalize(c, n)) return rc;
  int rc = c->prec == CG_FP64 ? finalize_t<double, double>(c, n, src)
           : c->met32          ? finalize_t<float, float>(c, n, src)
                               : finalize_t<float, double>(c, n, src);  // (not the fixed-point mode: rejected above)
  if (rc) return rc;
  std::swap(c->sA, c->sC);  // swap_a_c
  c->ext_fresh = (n == 4);
  return CG_OK;
}

constexpr int kMqCheck = 2;

int cg_rk4_steps(cg_ctx* c, int nsteps) {
  if (!c || nsteps < 0) return CG_ERR_ARG;
  if (gauge_pass(c)) {
    // The shift build runs the reference's method sequence
    // (sims/static.cc:74-170): stepInit, the NaN guard, then RKEvolve (stage
    // kernel + gauge pass) and K{n}Finalize per stage.
    for (int s = 0; s < nsteps; ++s) {
      int rc;
      if ((rc = cg_step_init(c))) return rc;
      long long nn = 0;
      if ((rc = cg_num_nans(c, "DIFFphi_a", &nn))) return rc;
      if (nn) return fail(c, CG_ERR_NAN, "NAN detected in DIFFphi");
      for (int n = 1; n <= 4; ++n) {
        if ((rc = cg_rk_evolve(c))) return rc;
        if ((rc = cg_k_finalize(c, n))) return rc;
      }
    }
    return CG_OK;
  }
  // fixed physical slots for the fused path: P=0, F=1, X=2, Y=3
  c->sP = 0; c->sF = 1;
  const int X = 2, Y = 3;
  const bool ko_norm = ko_norm_path(c);
  // one fused stage: sources and RHS with the current FRW values, then the
  // FRW stage step (bssn.cc:269), whose phi weights the averages of the output
  const bool flrw = lapse_needs_kavg(c->p.lapse);  // averages needed inside the step
  SrcArgs src4{};
  // dust: the sources of stage n use W, detg, g_ij of stage n-1's input
  // (its DIFFgamma_ij from the normalised copy when that stage read one)
  int prev_in = 0, prev_met = 0, prev_norm = 0;
  double prev_phiF = 0.0;
  // Accumulator-free form (no _f sweeps; the KO path with normalisation
  // keeps the reference's accumulator)
  const bool incr = fused_incr(c) && c->incr_pick;
  auto stage = [&](int n, int rin, int rout, int normalize, int in_norm = -1, int norm_out = -1, int sf = 1,
                   int sa2 = -1) -> int {
    StageArgs a = make_args(c, n, rin, 0, sf, rout, normalize);
    a.s_in_norm = in_norm;
    a.s_norm_out = norm_out;
    a.incr = incr ? 1 : 0;
    a.s_a2 = sa2;
    if (c->dreg) {
      const int met_in = in_norm >= 0 ? in_norm : rin;
      if (int rc = dust_flux(c, rin, normalize, met_in)) return rc;
      if (n > 1) {
        a.lag_mode = prev_norm ? 2 : 1;
        a.s_lag = prev_in;
        a.s_lag_met = prev_met;
        a.phi_F_lag = prev_phiF;
      }
      prev_in = rin;
      prev_met = met_in;
      prev_norm = normalize;
      prev_phiF = c->frw.g_phi;
    }
    src4 = SrcArgs{c->DIFFD, c->p.matter, c->frw.g_phi, c->frw.g_rho, c->lambda, c->rho_src};
    c->frw.stage(n, c->p.dt);
    a.phi_F_next = c->frw.g_phi;
    int rc = launch(c, a);
    return (rc || !flrw) ? rc : ext_combine(c, c->last_rows, false);
  };
  c->ext_lazy = false;
  // With DMA staging, stage 4 also writes the normalised copy of the new
  // state (stepInit's _a) into the dead slot Y, so that the next step's
  // stage 1 runs the DMA kernel instead of normalising halos on the fly.
  bool have_norm = false;
  int norm_slot = Y;
  for (int s = 0; s < nsteps; ++s) {
    int rc;
    c->sP = 0; c->sF = 1;
    // fixed-point metric: exponents checked every kMqCheck steps (a field's
    // words stay below 2^MQ_BITS; twice that saturates)
    if (c->mq && s % kMqCheck == 0 && (rc = mq_check(c))) return rc;
    const bool pre = !ko_norm && c->p.normalize_metric && s + 1 < nsteps && c->var &&
                     c->dma_allowed && c->dma_pick && dma_fits(c, *c->var);
    if (incr) {
      // a1, a2, a3 rotate over slots 1..3 so that the normalised copy of the
      // last stage 4 (written over a1) is the next stage 1's input
      if (flrw && !c->ext_fresh && (rc = ext_from_state(c))) return rc;
      int s1, s2, s3;
      incr_slots(have_norm ? norm_slot : 0, s1, s2, s3);
      if ((rc = have_norm ? stage(1, 0, s1, 0, norm_slot) : stage(1, 0, s1, c->p.normalize_metric))) return rc;
      if ((rc = stage(2, s1, s2, 0))) return rc;
      if ((rc = stage(3, s2, s3, 0))) return rc;
      if ((rc = stage(4, s3, 0, 0, -1, pre ? s1 : -1, s1, s2))) return rc;
      have_norm = pre;
      norm_slot = s1;
      c->sA = 0; c->sC = s3; c->a_is_p = true;
      c->ext_fresh = flrw;
      if (!flrw) {
        c->ext_lazy = true;
        c->ext_lazy_in = s3;
        c->ext_lazy_src = src4;
        c->ext_lazy_phiF = c->frw.g_phi;
      }
      continue;
    }
    if (ko_norm) {
      // KO needs normalised A_ij at halo points: materialise stepInit first
      c->a_is_p = false; c->sA = X; c->sC = Y;
      if ((rc = cg_step_init(c))) return rc;
      if ((rc = stage(1, X, Y, 0))) return rc;
      if ((rc = stage(2, Y, X, 0))) return rc;
    } else {
      // stepInit's setExtraFieldData equals the last K4Finalize's unless the
      // state changed in between; only the ConformalFLRW lapse reads it
      if (flrw && !c->ext_fresh && (rc = ext_from_state(c))) return rc;
      if ((rc = have_norm ? stage(1, 0, X, 0, Y) : stage(1, 0, X, c->p.normalize_metric))) return rc;
      if ((rc = stage(2, X, Y, 0))) return rc;
    }
    const int in3 = ko_norm ? X : Y, out3 = ko_norm ? Y : X;
    if ((rc = stage(3, in3, out3, 0))) return rc;
    if ((rc = stage(4, out3, 0, 0, -1, pre ? Y : -1))) return rc;
    have_norm = pre;
    c->sA = 0; c->sC = out3; c->a_is_p = true;
    c->ext_fresh = flrw;
    if (!flrw) {
      c->ext_lazy = true;
      c->ext_lazy_in = out3;
      c->ext_lazy_src = src4;
      c->ext_lazy_phiF = c->frw.g_phi;
    }
  }
  unsigned nn = 0;
  HIPCHK(c, hipMemcpyAsync(&nn, c->nan_dev, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (nn != c->nan_seen) {
    c->nan_seen = nn;
    return fail(c, CG_ERR_NAN, "NAN detected in DIFFphi");
  }
  return CG_OK;
}

int cg_extra_field_data(cg_ctx* c, double out[4]) {
  if (!c || !out) return CG_ERR_ARG;
  if (int rc = ext_resolve(c)) return rc;
  HIPCHK(c, hipMemcpyAsync(out, c->ext, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CG_OK;
}

int cg_num_nans(cg_ctx* c, const char* name, long long* count) {
  if (!c || !count) return CG_ERR_ARG;
  int e;
  size_t esz;
  void* d = resolve(c, name, false, &e, &esz);
  if (!d) return fail(c, CG_ERR_ARG, std::string("unknown field ") + (name ? name : ""));
  int mf, ms;
  if (mq_register(c, d, &mf, &ms)) {
    // fixed-point words hold no NaN: the stage kernels count a NaN DIFFphi
    // before rounding it (cg_rk4_steps reports it)
    *count = 0;
    return CG_OK;
  }
  HIPCHK(c, hipMemsetAsync(c->nan_cnt, 0, sizeof(unsigned long long), c->stream));
  const int bs = 256;
  const long nb = (c->npts + bs - 1) / bs;
  if (esz == 8)
    hipLaunchKernelGGL(k_num_nans<double>, dim3(nb), dim3(bs), 0, c->stream, (const double*)d, c->npts, c->nan_cnt);
  else
    hipLaunchKernelGGL(k_num_nans<float>, dim3(nb), dim3(bs), 0, c->stream, (const float*)d, c->npts, c->nan_cnt);
  HIPCHK(c, hipGetLastError());
  unsigned long long n = 0;
  HIPCHK(c, hipMemcpyAsync(&n, c->nan_cnt, sizeof n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *count = (long long)n;
  return CG_OK;
}

int cg_sync(cg_ctx* c) {
  if (!c) return CG_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CG_OK;
}

int cg_kernel_info(const cg_ctx* cc, int* vgprs, int* scratch, int* lds, int* ti, int* tj, int* tk) {
  cg_ctx* c = const_cast<cg_ctx*>(cc);
  if (!c) return CG_ERR_ARG;
  if (!c->var || c->var->live_alpha != alpha_live(c)) c->var = pick_variant(c);
  if (!c->var) return fail(c, CG_ERR_ARG, "no stage kernel variant fits this grid/order");
  const bool dma = c->dma_allowed && c->dma_pick && dma_fits(c, *c->var);
  int rc = (dma ? c->var->info_dma : c->var->info)(vgprs, scratch, lds);
  *ti = c->var->ti; *tj = c->var->tj; *tk = c->var->tk;
  return rc ? CG_ERR_HIP : CG_OK;
}

// the per-point constraint pass (prepBSSNOutput's set_bd_values on every
// point, sims/sim.cc:214-231, with the constraint values of
// setConstraintCalcs) into c->cons
// The register the constraint pass reads: _a as the reference's output step
// sees it, after stepInit, i.e. the normalised _p when the last call was a
// fused step.  The momentum constraint differentiates A_ij, so the
// normalised register is materialised (into the spare stage register and the
// dead _f) rather than normalised on the fly.  Allocates nothing.
int cg::cons_input(cg_ctx* c, int* rin_out) {
  // the pending K4Finalize averages read the dust source this pass rewrites
  if (int rc = ext_resolve(c)) return rc;
  if (c->p.nx <= 4 + c->R - 2 || c->p.ny <= 8 + c->R - 2 || c->p.nz <= 8 + c->R - 2)
    return fail(c, CG_ERR_ARG, "grid too small for the constraint kernel tile");
  int rin = c->sA;
  if (c->a_is_p) {
    rin = (c->sC == 3) ? 2 : 3;  // the spare of the two stage registers
    const int bs = 256;
    const long nbk = (c->npts + bs - 1) / bs;
    if (c->prec == CG_FP64)
      hipLaunchKernelGGL((k_step_init<double, double>), dim3(nbk), dim3(bs), 0, c->stream, reg_table(c), c->sP, rin,
                         c->sF, c->npts, c->p.normalize_metric, (double*)nullptr, 0.0);
    else if (c->mq)
      hipLaunchKernelGGL((k_step_init<float, int>), dim3(nbk), dim3(bs), 0, c->stream, reg_table(c), c->sP, rin,
                         c->sF, c->npts, c->p.normalize_metric, (double*)nullptr, 0.0);
    else if (c->met32)
      hipLaunchKernelGGL((k_step_init<float, float>), dim3(nbk), dim3(bs), 0, c->stream, reg_table(c), c->sP, rin,
                         c->sF, c->npts, c->p.normalize_metric, (double*)nullptr, 0.0);
    else
      hipLaunchKernelGGL((k_step_init<float, double>), dim3(nbk), dim3(bs), 0, c->stream, reg_table(c), c->sP, rin,
                         c->sF, c->npts, c->p.normalize_metric, (double*)nullptr, 0.0);
    HIPCHK(c, hipGetLastError());
    if (c->dreg && dust_reg_init(c, c->sP, rin, c->sF)) return CG_ERR_HIP;
  }
  *rin_out = rin;
  return CG_OK;
}

// arguments of a constraint pass over register rin with the 4x8x8 tile of
// the constraint kernel; dust: sources from this state itself
// (initDustStep's populate, sims/dust.cc:53-63)
static StageArgs cons_args(cg_ctx* c, int rin) {
  StageArgs a = make_args(c, 0, rin, c->sP, c->sF, c->sC, 0);
  a.ntx = (c->p.nx + 3) / 4;
  a.nty = (c->p.ny + 7) / 8;
  a.ntz = (c->p.nz + 7) / 8;
  a.per_xcd = (a.ntx * a.nty * a.ntz + 7) / 8;
  return a;
}

static int cons_eval(cg_ctx* c) {
  int rin;
  if (int rc = cons_input(c, &rin)) return rc;
  const long pad = c->rstride;
  const int nb = 1024;
  if (!c->cons) {
    HIPCHK(c, hipMalloc((void**)&c->cons, (size_t)CONS_N * pad * sizeof(double)));
    HIPCHK(c, hipMalloc((void**)&c->cons_part, (size_t)nb * 5 * kStats * sizeof(double) + 64 * sizeof(double)));
  }
  StageArgs a = cons_args(c, rin);
  a.cons = c->cons;
  a.cons_stride = pad;
  int rc = launch_constraints(c->prec, c->p.stencil_order, alpha_live(c), c->dreg != nullptr, a, c->stream);
  if (rc) return fail(c, CG_ERR_HIP, std::string("constraint launch: ") + hipGetErrorString((hipError_t)rc));
  c->cons_valid = true;
  return CG_OK;
}

// setConstraintCalcs' outputs from the totals: tot[6q + {0..5}] = sum val,
// sum scale, sum scaled, sum val^2, max|val|, max|scaled|; means[2q + {0, 1}]
// of val and scaled; var[2q + {0, 1}] their summed squared deviations
static void cons_stats(const cg_ctx* c, const double* tot, const double* means, const double* var, double* H,
                       double* M, double* G, double* A, double* S) {
  const double N = (double)c->npts;
  double* outs[5] = {H, M, G, A, S};
  const double dx3 = c->p.dx * c->p.dx * c->p.dx;
  for (int q = 0; q < 5; ++q) {
    double* o = outs[q];
    if (!o) continue;
    o[0] = means[2 * q];
    o[1] = sqrt(var[2 * q] / (N - 1.0));           // BSSN_NORMALIZE_STDEV
    o[2] = tot[q * kStats + 4];
    o[3] = tot[q * kStats + 1];                    // mean_scale is left as a sum
    o[4] = means[2 * q + 1];
    o[5] = sqrt(var[2 * q + 1] / (N - 1.0));
    o[6] = tot[q * kStats + 5];
    if (q < 2) o[7] = sqrt(tot[q * kStats + 3] * dx3);  // H_L2, M_L2 (bssn.cc:989, 994, 1067)
  }
}

int cg_constraints(cg_ctx* c, double H[8], double M[8], double G[7], double A[7], double S[7]) {
  if (!c) return CG_ERR_ARG;
  if (int rc = cons_eval(c)) return rc;
  const long pad = c->rstride;
  const int nb = 1024;
  double* part = c->cons_part;
  double* fin = part + (size_t)nb * 5 * kStats;  // 30 + 10 + 10 doubles
  hipLaunchKernelGGL(k_cons_partial, dim3(nb), dim3(kRedThreads), 0, c->stream, c->cons, pad, c->npts, part);
  hipLaunchKernelGGL(k_combine, dim3(1), dim3(64), 0, c->stream, part, nb, 5 * kStats, 0x30u, fin);
  HIPCHK(c, hipGetLastError());
  double tot[5 * kStats];
  HIPCHK(c, hipMemcpyAsync(tot, fin, sizeof tot, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const double N = (double)c->npts;
  double means[10];
  for (int q = 0; q < 5; ++q) {
    means[2 * q] = tot[q * kStats + 0] / N;      // BSSN_NORMALIZE_MEAN
    means[2 * q + 1] = tot[q * kStats + 2] / N;
  }
  HIPCHK(c, hipMemcpyAsync(fin + 30, means, sizeof means, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_cons_var, dim3(nb), dim3(kRedThreads), 0, c->stream, c->cons, pad, c->npts, fin + 30, part);
  hipLaunchKernelGGL(k_combine, dim3(1), dim3(64), 0, c->stream, part, nb, 10, 0u, fin + 40);
  HIPCHK(c, hipGetLastError());
  double var[10];
  HIPCHK(c, hipMemcpyAsync(var, fin + 40, sizeof var, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  cons_stats(c, tot, means, var, H, M, G, A, S);
  return CG_OK;
}

int cg_constraints_noalloc(cg_ctx* c, double H[8], double M[8], double G[7], double A[7], double S[7]) {
  if (!c) return CG_ERR_ARG;
  double* sc = (c->p.device >= 0 && c->p.device < kNaDevices) ? g_scratch_dev[c->p.device] : nullptr;
  if (!sc) return fail(c, CG_ERR_STATE, "no constraint scratch on this device (device ordinal >= 64)");
  int rin;
  if (int rc = cons_input(c, &rin)) return rc;
  StageArgs a = cons_args(c, rin);
  double* acc0 = sc + kNaTot;             // pass-0 totals [CONS_ROW_W]
  double* acc1 = acc0 + CONS_ROW_W;       // pass-1 totals [CONS_ROW_W1]
  double* dmeans = acc1 + CONS_ROW_W1;    // [10]
  a.cons = nullptr;
  a.cons_rows = sc;
  a.cons_means = dmeans;
  const int total = a.ntx * a.nty * a.ntz;
  for (int pass = 0; pass < 2; ++pass) {
    a.cons_pass = pass;
    for (int base = 0; base < total; base += kNaRows) {
      a.tile_base = base;
      a.tile_count = std::min(kNaRows, total - base);
      a.per_xcd = (a.tile_count + 7) / 8;
      int rc = launch_constraints(c->prec, c->p.stencil_order, alpha_live(c), c->dreg != nullptr, a, c->stream);
      if (rc) return fail(c, CG_ERR_HIP, std::string("constraint launch: ") + hipGetErrorString((hipError_t)rc));
      hipLaunchKernelGGL(k_rows_fold, dim3(1), dim3(kFoldThreads), 0, c->stream, (const double*)sc, a.tile_count,
                         CONS_ROW_W, pass == 0 ? CONS_ROW_W : CONS_ROW_W1, pass, base == 0 ? 1 : 0,
                         pass == 0 ? acc0 : acc1);
      HIPCHK(c, hipGetLastError());
    }
    if (pass == 0) {
      hipLaunchKernelGGL(k_cons_means, dim3(1), dim3(64), 0, c->stream, (const double*)acc0, (double)c->npts, dmeans);
      HIPCHK(c, hipGetLastError());
    }
  }
  double res[CONS_ROW_W + 2 * CONS_ROW_W1];
  HIPCHK(c, hipMemcpyAsync(res, acc0, sizeof res, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  cons_stats(c, res, res + CONS_ROW_W + CONS_ROW_W1, res + CONS_ROW_W, H, M, G, A, S);
  return CG_OK;
}

// ------------------------------------------------ output-step readers --
// device staging of small results (strips, partial sums), grown on demand
static double* out_buf(cg_ctx* c, size_t n) {
  if (c->obuf_n < n) {
    if (c->obuf) {
      (void)hipStreamSynchronize(c->stream);
      (void)hipFree(c->obuf);
      c->obuf = nullptr;
      c->obuf_n = 0;
    }
    if (hipMalloc((void**)&c->obuf, n * sizeof(double)) != hipSuccess) {
      fail(c, CG_ERR_HIP, "out of device memory for output staging");
      return nullptr;
    }
    c->obuf_n = n;
  }
  return c->obuf;
}

static long wrap_index(long v, long n) { return ((v % n) + n) % n; }

// out[x] = f[base + (x mod wrap) * run], x < n
template <typename T>
__global__ void k_strip(const T* __restrict__ f, long base, long run, long wrap, long n, double* __restrict__ out) {
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x < n) out[x] = (double)f[base + (x % wrap) * run];
}

static int strip_gather(cg_ctx* c, const void* d, size_t esz, long base, long run, long wrap, long n, double* dev) {
  const int bs = 256;
  const long nb = (n + bs - 1) / bs;
  if (esz == 8)
    hipLaunchKernelGGL(k_strip<double>, dim3(nb), dim3(bs), 0, c->stream, (const double*)d, base, run, wrap, n, dev);
  else
    hipLaunchKernelGGL(k_strip<float>, dim3(nb), dim3(bs), 0, c->stream, (const float*)d, base, run, wrap, n, dev);
  HIPCHK(c, hipGetLastError());
  return CG_OK;
}

int cg_field_strip(cg_ctx* c, const char* name, int axis, long n1, long n2, double* out, size_t n) {
  if (!c || !out || axis < 1 || axis > 3) return CG_ERR_ARG;
  const long nx = c->p.nx, ny = c->p.ny, nz = c->p.nz;
  const long ext[3] = {nx, ny, nz};
  if ((long)n != ext[axis - 1]) return fail(c, CG_ERR_ARG, "a strip holds the grid's extent along its axis");
  int e;
  size_t esz;
  const void* d = field_ptr(c, name, &esz, &e);
  if (!d) return fail(c, e == 3 || e == 5 ? CG_ERR_STATE : CG_ERR_ARG,
                      e == 3   ? std::string(name) + " is filled by cg_constraints"
                      : e == 5 ? std::string(name) + " is stored in fixed point: read it with cg_field_download"
                               : std::string("unknown field ") + name);
  // field(i, n1, n2), field(n1, j, n2), field(n1, n2, k) with the offsets
  // wrapped as CosmoArray::idx wraps them (utils/Array.h:149-159)
  long base, run;
  if (axis == 1) { base = wrap_index(n1, ny) * nz + wrap_index(n2, nz); run = ny * nz; }
  else if (axis == 2) { base = wrap_index(n1, nx) * ny * nz + wrap_index(n2, nz); run = nz; }
  else { base = (wrap_index(n1, nx) * ny + wrap_index(n2, ny)) * nz; run = 1; }
  double* dev = out_buf(c, n);
  if (!dev) return CG_ERR_HIP;
  if (int rc = strip_gather(c, d, esz, base, run, (long)n, (long)n, dev)) return rc;
  HIPCHK(c, hipMemcpyAsync(out, dev, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CG_OK;
}

int cg_constraint_strip(cg_ctx* c, int axis, long n1, long n2, double* H, double* M, size_t n) {
  if (!c || !H || !M || axis < 1 || axis > 3) return CG_ERR_ARG;
  const long nx = c->p.nx, ny = c->p.ny, nz = c->p.nz;
  const long ext[3] = {nx, ny, nz};
  if ((long)n != ext[axis - 1]) return fail(c, CG_ERR_ARG, "a strip holds the grid's extent along its axis");
  if (int rc = cons_eval(c)) return rc;
  // set1DConstraintOutput passes the running index as i for every axis
  // (bssn.cc:918-934), so the points are (x, n1, n2) with x wrapped in i
  const long base = wrap_index(n1, ny) * nz + wrap_index(n2, nz), run = ny * nz;
  double* dev = out_buf(c, 2 * n);
  if (!dev) return CG_ERR_HIP;
  const double* cons = c->cons;
  if (int rc = strip_gather(c, cons + (long)CONS_H * c->rstride, 8, base, run, nx, (long)n, dev)) return rc;
  if (int rc = strip_gather(c, cons + (long)(CONS_MI + axis - 1) * c->rstride, 8, base, run, nx, (long)n, dev + n))
    return rc;
  HIPCHK(c, hipMemcpyAsync(H, dev, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(M, dev + n, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CG_OK;
}

// conformal_average / conformal_standard_deviation / volume_average partial
// sums (utils/math.h:1382-1415, 1458-1470): one row per block of
// sum w*f, sum w (pass 0) or sum w*(avg - f)^2 (pass 1), w = exp(6(DIFFphi + phi_FRW))
constexpr int kCwThreads = 256;
template <typename T>
__global__ void __launch_bounds__(kCwThreads) k_conformal_sums(const T* __restrict__ f, const double* __restrict__ dphi,
                                                             double phiF, long n, int pass, double avg,
                                                             double* __restrict__ part) {
  __shared__ double s0[kCwThreads], s1[kCwThreads];
  double a0 = 0.0, a1 = 0.0;
  for (long x = (long)blockIdx.x * kCwThreads + threadIdx.x; x < n; x += (long)gridDim.x * kCwThreads) {
    const double w = exp(6.0 * (dphi[x] + phiF));
    const double v = (double)f[x];
    if (pass == 0) {
      a0 += w * v;
      a1 += w;
    } else {
      const double d = avg - v;
      a0 += w * (d * d);
    }
  }
  s0[threadIdx.x] = a0;
  s1[threadIdx.x] = a1;
  __syncthreads();
  for (int h = kCwThreads / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      s0[threadIdx.x] += s0[threadIdx.x + h];
      s1[threadIdx.x] += s1[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s0[0];
    part[2 * blockIdx.x + 1] = s1[0];
  }
}

int cg_field_stats(cg_ctx* c, const char* name, double out[3]) {
  if (!c || !out) return CG_ERR_ARG;
  if (c->met32) return fail(c, CG_ERR_ARG, "conformal statistics read an FP64 DIFFphi (not in the all-FP32 mode)");
  int e;
  size_t esz, pesz;
  const void* d = resolve(c, name, false, &e, &esz);
  if (!d) return fail(c, e == 3 ? CG_ERR_STATE : CG_ERR_ARG,
                      e == 3 ? std::string(name) + " is filled by cg_constraints" : std::string("unknown field ") + name);
  const double* dphi = (const double*)resolve(c, "DIFFphi_a", false, &e, &pesz);
  const double phiF = c->frw.g_phi;
  const long n = c->npts;
  const int nb = (int)std::min<long>(1024, (n + kCwThreads - 1) / kCwThreads);
  double* dev = out_buf(c, 2 * (size_t)nb);
  if (!dev) return CG_ERR_HIP;
  std::vector<double> part(2 * (size_t)nb);
  auto pass = [&](int which, double avg) -> int {
    if (esz == 8)
      hipLaunchKernelGGL(k_conformal_sums<double>, dim3(nb), dim3(kCwThreads), 0, c->stream, (const double*)d, dphi,
                         phiF, n, which, avg, dev);
    else
      hipLaunchKernelGGL(k_conformal_sums<float>, dim3(nb), dim3(kCwThreads), 0, c->stream, (const float*)d, dphi,
                         phiF, n, which, avg, dev);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(part.data(), dev, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CG_OK;
  };
  if (int rc = pass(0, 0.0)) return rc;
  double swf = 0.0, sw = 0.0;
  for (int b = 0; b < nb; ++b) { swf += part[2 * b]; sw += part[2 * b + 1]; }  // block order: deterministic
  const double vol = sw / (double)n;
  const double avg = swf / (double)n / vol;
  if (int rc = pass(1, avg)) return rc;
  double sdev = 0.0;
  for (int b = 0; b < nb; ++b) sdev += part[2 * b];
  out[0] = avg;
  out[1] = sqrt(sdev / (double)(n - 1) / vol);
  out[2] = vol;
  return CG_OK;
}

int cg_autotune(cg_ctx* c, int reps, char* report, size_t report_len) {
  if (!c || reps < 1) return CG_ERR_ARG;
  if (!c->a_is_p) return fail(c, CG_ERR_STATE, "autotune between method-level stages would clobber _a/_c");
  if (int rc = ext_resolve(c)) return rc;
  // Times stage-1, stage-2 and stage-4 shaped launches per candidate, in
  // each RK4 form the fused step may use, so the evolved state is untouched:
  // outputs into the spare slot 3, the NaN count into scratch, and dead
  // slots as the other inputs (stage 1: stencils of _p; stages 2 and 4:
  // stencils of slot 2, as their real input is not _p, whose point values
  // would otherwise come from cache; stage 4 reads its a2 from slot 3 just
  // before writing it).  The candidate's step
  // estimate is t1 + 2 t2 + t4 (stage 3 has stage 2's shape); the forms load
  // their stages differently, so the best form depends on the grid.
  std::string rep = "[";
  const StageVariant* best = nullptr;
  double best_ms = 1e30;
  // events and the per-launch timing switch are restored on every return
  struct Guard {
    cg_ctx* c;
    bool timing;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Guard() {
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      c->timing = timing;
    }
  } guard{c, c->timing};
  HIPCHK(c, hipEventCreate(&guard.e0));
  HIPCHK(c, hipEventCreate(&guard.e1));
  hipEvent_t e0 = guard.e0, e1 = guard.e1;
  c->timing = false;
  int best_order = 0;
  bool best_dma = false, best_incr = false;
  if (c->dreg) {  // dust: the fluxes of the state the timed launches read
    if (int rc = dust_flux(c, 0, 0)) return rc;
  }
  unsigned* nan_scratch = reinterpret_cast<unsigned*>(out_buf(c, 1));
  if (!nan_scratch) return CG_ERR_HIP;
  const bool can_incr = fused_incr(c);
  for (const auto& v : stage_variants()) {
    if (!variant_matches(c, v)) continue;
    c->var = &v;
    for (int d = 0; d < 2; ++d) {
      if (d == 1 && !(c->dma_allowed && dma_fits(c, v))) continue;
      c->dma_pick = d == 1;
      for (int order = 0; order < 6; ++order) {
        if (order == 3 && !morton_fits((c->p.nx + v.ti - 1) / v.ti, (c->p.ny + v.tj - 1) / v.tj)) continue;
        c->tile_order = order;
        for (int form = can_incr ? 1 : 0; form >= 0; --form) {
          double tq[3] = {0.0, 0.0, 0.0};
          for (int q = 0; q < 3; ++q) {
            const int st = q == 0 ? 1 : q == 1 ? 2 : 4;
            StageArgs a = make_args(c, st, st == 1 ? 0 : 2, 0, 1, 3, 0);
            if (st == 1 && c->p.normalize_metric) {
              // steady state of the fused step: the normalised copy with
              // DMA staging (slot 2 stands in), on the fly without
              if (d == 1 && !ko_norm_path(c)) a.s_in_norm = 2;
              else a.normalize = 1;
            }
            a.rho_out = nullptr;  // dust: DIFFr_a keeps the last real stage's source
            a.incr = form;
            a.s_a2 = 3;
            a.nan_count = nan_scratch;
            if (c->mq) a.mq_sat = nan_scratch;  // timing launches count nothing real
            int rc = launch(c, a);
            if (rc) return rc;
            HIPCHK(c, hipEventRecord(e0, c->stream));
            for (int r = 0; r < reps; ++r)
              if ((rc = launch(c, a))) return rc;
            HIPCHK(c, hipEventRecord(e1, c->stream));
            HIPCHK(c, hipEventSynchronize(e1));
            float ms = 0;
            HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
            tq[q] = ms / reps;
          }
          const double avg = 0.25 * (tq[0] + 2.0 * tq[1] + tq[2]);  // mean stage of a step
          int vg = 0, sc = 0, ld = 0;
          (d ? v.info_dma : v.info)(&vg, &sc, &ld);
          char buf[320];
          snprintf(buf, sizeof buf,
                   "%s{\"tile\":[%d,%d,%d],\"wpe\":%d,\"dma\":%d,\"order\":%d,\"incr\":%d,\"ms\":%.4f,"
                   "\"ms1\":%.4f,\"ms2\":%.4f,\"ms4\":%.4f,\"vgprs\":%d,\"scratch\":%d}",
                   rep.size() > 1 ? "," : "", v.ti, v.tj, v.tk, v.wpe, d, order, form, avg, tq[0], tq[1], tq[2], vg,
                   sc);
          rep += buf;
          if (avg < best_ms) {
            best_ms = avg; best = &v; best_order = order; best_dma = d == 1; best_incr = form == 1;
          }
        }
      }
    }
  }
  c->tile_order = best_order;
  c->dma_pick = best_dma || !best;
  c->incr_pick = best_incr || !best;
  rep += "]";
  c->var = best;
  if (report && report_len) {
    strncpy(report, rep.c_str(), report_len - 1);
    report[report_len - 1] = 0;
  }
  if (!best) return fail(c, CG_ERR_ARG, "no stage kernel variant fits this grid/order");
  c->req_ti = best->ti; c->req_tj = best->tj; c->req_tk = best->tk;
  return CG_OK;
}

int cg_set_timing(cg_ctx* c, int on) {
  if (!c) return CG_ERR_ARG;
  c->timing = on != 0;
  return CG_OK;
}

int cg_stage_time_ms(cg_ctx* c, double* avg, int* count) {
  if (!c || !avg || !count) return CG_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double tot = 0;
  int n = 0;
  for (auto& e : c->ev) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, e.first, e.second) == hipSuccess) { tot += ms; ++n; }
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  c->ev.clear();
  *avg = n ? tot / n : 0.0;
  *count = n;
  return CG_OK;
}

