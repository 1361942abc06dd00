// bardeen.hip -- Bardeen potentials and the SVT decomposition of the metric
// on the device (components/bssn/bardeen.cc:22-460, bardeen.h:16-236),
// the consumer of the evolved fields behind the reference's `use_bardeen`.
//
// cg_bardeen evaluates, on the state the reference's output step sees (the
// register cg_constraints reads), what Bardeen::setPotentials does:
//   1. the BSSN right-hand side of every field at that state (the stage
//      kernel in RKEvolve form, plus the gauge pass of the shift build),
//      i.e. the ev_* values the reference calls per point, into _c;
//   2. second time derivatives of gamma~_ij and phi, the metric perturbation
//      h_ij = e^{4 phi} gamma~_ij - a^2 delta_ij with its first two time
//      derivatives, h_0i and E (bardeen.cc:75-216);
//   3. the SVT potentials A, B, F, C_i and their time derivatives through
//      inverse Laplacians (hipFFT r2c / c2r, each mode divided by
//      -(2 pi |p| / L)^2 with integer lattice momenta p and L = nx dx, the
//      mean zeroed: Fourier::inverseLaplacian, utils/Fourier.h:252-300), the
//      vector potentials G_i, the tensor part D_ij and the Bardeen
//      potentials Phi, Psi (bardeen.cc:219-409);
//   4. the linearised-constraint violations, 17 numbers
//      (bardeen.cc:412-455), from deterministic FP64 reductions.
// The reference's own quirks are kept: h_0i holds its time derivative and
// dt h_0i stays zero (bardeen.cc:198-208); E uses the conformal metric in
// the beta^2 term; violation ratios 8-13 use the matter+Lambda scale factor.
//
// Everything runs in FP64 on one thread per point, with stencils read from
// global memory through L1/L2 (this runs at output steps, not in the
// evolution loop); FP32-split registers are widened into the workspace
// first.  The workspace (about 60 FP64 arrays) is allocated on first use.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "bssn_stage.h"
#include "runtime.h"
#include "stencil.h"

namespace cg {

// workspace array indices
enum BrdArr {
  W_D2TG = 0, W_D2TPHI = 6, W_H = 7, W_DTH = 13, W_D2TH = 19, W_H0 = 25, W_E = 28,
  W_A = 29, W_DTA, W_D2TA, W_B, W_DTB, W_D2TB, W_F, W_DTF, W_G1, W_G2, W_G3, W_C1, W_C2, W_C3,
  W_DTC1, W_DTC2, W_DTC3, W_VMAG, W_D11, W_D12, W_D13, W_D22, W_D23, W_D33, W_PHI, W_PSI,
  W_LV, W_LVMAG, W_LVDMAG, W_LVD,
  // FP64 copies of storage-type registers (FP32 split): state and RHS
  W_SA = W_LVD + 1,              // A11..A33 (6), DIFFK, DIFFalpha of the state
  W_CA = W_SA + 8,               // the same of the right-hand side
  W_N = W_CA + 8
};

struct BrdWork {
  double* arr = nullptr;        // [W_N][pad]
  double2* spec = nullptr;      // [nx][ny][nz/2+1]
  double* part = nullptr;       // reduction partials
  hipfftHandle r2c = 0, c2r = 0;
  bool plans = false;
  bool valid = false;           // potentials of the last call present
};

namespace {

constexpr int kBT = 256;        // threads per block
constexpr int kRedBlocks = 1024;

__device__ __forceinline__ int wrap1(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// periodic stencils on a global array (coefficients of utils/math.h as in
// stencil.h), 1/dx and 1/dx^2 applied
template <int R>
struct Sten {
  int nx, ny, nz;
  double i1, i2;
  __device__ long at(int x, int y, int z) const {
    return ((long)wrap1(x, nx) * ny + wrap1(y, ny)) * nz + wrap1(z, nz);
  }
  __device__ double v(const double* f, int i, int j, int k, int d, int t) const {
    return d == 0 ? f[at(i + t, j, k)] : d == 1 ? f[at(i, j + t, k)] : f[at(i, j, k + t)];
  }
  __device__ double d1(const double* f, int i, int j, int k, int d) const {
    double acc = 0.0;
#pragma unroll
    for (int m = 1; m <= R; ++m) acc += c1<R>(m) * (v(f, i, j, k, d, m) - v(f, i, j, k, d, -m));
    return acc * i1;
  }
  __device__ double v2(const double* f, int i, int j, int k, int a, int ta, int b, int tb) const {
    int x = i, y = j, z = k;
    (a == 0 ? x : a == 1 ? y : z) += ta;
    (b == 0 ? x : b == 1 ? y : z) += tb;
    return f[at(x, y, z)];
  }
  __device__ double dd(const double* f, int i, int j, int k, int a, int b) const {
    if (a == b) {
      double acc = c2<R>(0) * f[at(i, j, k)];
#pragma unroll
      for (int m = 1; m <= R; ++m) acc += c2<R>(m) * (v(f, i, j, k, a, m) + v(f, i, j, k, a, -m));
      return acc * i2;
    }
    double acc = 0.0;
#pragma unroll
    for (int m = 1; m <= R; ++m)
      acc += cm<R>(m) * ((v2(f, i, j, k, a, m, b, m) - v2(f, i, j, k, a, m, b, -m)) -
                         (v2(f, i, j, k, a, -m, b, m) - v2(f, i, j, k, a, -m, b, -m)));
    return acc * i2;
  }
  __device__ double abs_dder(const double* f, int i, int j, int k) const {
    return fabs(dd(f, i, j, k, 0, 0)) + fabs(dd(f, i, j, k, 0, 1)) + fabs(dd(f, i, j, k, 0, 2)) +
           fabs(dd(f, i, j, k, 1, 1)) + fabs(dd(f, i, j, k, 1, 2)) + fabs(dd(f, i, j, k, 2, 2));
  }
};

__device__ __forceinline__ int sy2(int a, int b) {
  return a == b ? (a == 0 ? 0 : a == 1 ? 3 : 5) : (a + b == 1 ? 1 : a + b == 2 ? 2 : 4);
}

// device pointers of the inputs (all FP64): state and right-hand side
struct BrdIn {
  const double* g[6];    // DIFFgamma_ij (state)
  const double* cg[6];   // ev_DIFFgamma_ij
  const double* phi;     // DIFFphi
  const double* cphi;    // ev_DIFFphi
  const double* A[6];    // A_ij
  const double* cA[6];   // ev_A_ij
  const double* K;       // DIFFK
  const double* cK;      // ev_DIFFK
  const double* al;      // DIFFalpha
  const double* cal;     // ev_DIFFalpha
  const double* beta[3]; // beta^i (null without the shift build)
  const double* cbeta[3];
};

struct BrdBg {
  double a, dadt, H, d2adt2;
};

#define PT_INDEX()                                                  \
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;       \
  if (x >= n) return;                                               \
  const int k = (int)(x % S.nz), j = (int)((x / S.nz) % S.ny), i = (int)(x / ((long)S.nz * S.ny));

template <typename T>
__global__ void k_widen(const T* __restrict__ src, double* __restrict__ dst, long n) {
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x < n) dst[x] = (double)src[x];
}

// d^2/dt^2 gamma~_ij and phi (bardeen.cc:102-143)
template <int R>
__global__ void k_brd_d2t(Sten<R> S, BrdIn in, double* W, long pad, long n) {
  PT_INDEX();
  double b[3] = {0, 0, 0}, db[3] = {0, 0, 0}, dbj[3][3] = {}, ddb[3][3] = {};  // dbj[j][m] = d_j beta^m
  if (in.beta[0]) {
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      b[m] = in.beta[m][x];
      db[m] = in.cbeta[m][x];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        dbj[d][m] = S.d1(in.beta[m], i, j, k, d);
        ddb[d][m] = S.d1(in.cbeta[m], i, j, k, d);
      }
    }
  }
  const double dkbetak = dbj[0][0] + dbj[1][1] + dbj[2][2];
  const double dkdtbetak = ddb[0][0] + ddb[1][1] + ddb[2][2];
  const double dtalpha = in.cal[x], alpha = 1.0 + in.al[x];
  auto G = [&](int p, int q) { return in.g[sy2(p, q)][x] + (p == q ? 1.0 : 0.0); };
  auto DG = [&](int p, int q) { return in.cg[sy2(p, q)][x]; };
#pragma unroll
  for (int I = 0; I < 3; ++I)
#pragma unroll
    for (int J = I; J < 3; ++J) {
      const int s = sy2(I, J);
      double v = -2.0 * dtalpha * in.A[s][x] - 2.0 * alpha * in.cA[s][x];
      if (in.beta[0]) {
        v += db[0] * S.d1(in.g[s], i, j, k, 0) + db[1] * S.d1(in.g[s], i, j, k, 1) + db[2] * S.d1(in.g[s], i, j, k, 2) +
             b[0] * S.d1(in.cg[s], i, j, k, 0) + b[1] * S.d1(in.cg[s], i, j, k, 1) + b[2] * S.d1(in.cg[s], i, j, k, 2) +
             DG(I, 0) * dbj[J][0] + DG(I, 1) * dbj[J][1] + DG(I, 2) * dbj[J][2] + DG(J, 0) * dbj[I][0] +
             DG(J, 1) * dbj[I][1] + DG(J, 2) * dbj[I][2] + G(I, 0) * ddb[J][0] + G(I, 1) * ddb[J][1] +
             G(I, 2) * ddb[J][2] + G(J, 0) * ddb[I][0] + G(J, 1) * ddb[I][1] + G(J, 2) * ddb[I][2] -
             2.0 / 3.0 * (DG(I, J) * dkbetak + G(I, J) * dkdtbetak);
      }
      W[(W_D2TG + s) * pad + x] = v;
    }
  double v = -1.0 / 6.0 * (dtalpha * in.K[x] + alpha * in.cK[x]);
  if (in.beta[0]) {
#pragma unroll
    for (int d = 0; d < 3; ++d) v += db[d] * S.d1(in.phi, i, j, k, d) + b[d] * S.d1(in.cphi, i, j, k, d);
    v += 1.0 / 6.0 * dkdtbetak;
  }
  W[W_D2TPHI * pad + x] = v;
}

// h_ij, d_t h_ij, d_t^2 h_ij, h_0i (= its time derivative, as the
// reference leaves it), E (bardeen.cc:151-216)
__global__ void k_brd_h(BrdIn in, BrdBg bg, double* W, long pad, long n) {
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const double phi = in.phi[x], e4phi = exp(4.0 * phi), dtphi = in.cphi[x], d2tphi = W[W_D2TPHI * pad + x];
  const double a = bg.a, dadt = bg.dadt, d2adt2 = bg.d2adt2;
  for (int s = 0; s < 6; ++s) {
    const bool diag = s == 0 || s == 3 || s == 5;
    const double g = in.g[s][x] + (diag ? 1.0 : 0.0), dtg = in.cg[s][x], d2tg = W[(W_D2TG + s) * pad + x];
    W[(W_H + s) * pad + x] = e4phi * g - (diag ? a * a : 0.0);
    W[(W_DTH + s) * pad + x] = e4phi * (4.0 * dtphi * g + dtg) - (diag ? 2.0 * a * dadt : 0.0);
    W[(W_D2TH + s) * pad + x] =
        e4phi * ((4.0 * dtphi) * (4.0 * dtphi) * g + 8.0 * dtphi * dtg + 4.0 * d2tphi * g + d2tg) -
        (diag ? 2.0 * (dadt * dadt + a * d2adt2) : 0.0);
  }
  double b[3] = {0, 0, 0}, db[3] = {0, 0, 0};
  if (in.beta[0])
    for (int m = 0; m < 3; ++m) { b[m] = in.beta[m][x]; db[m] = in.cbeta[m][x]; }
  auto G = [&](int p, int q) { return in.g[sy2(p, q)][x] + (p == q ? 1.0 : 0.0); };
  auto DG = [&](int p, int q) { return in.cg[sy2(p, q)][x]; };
  for (int I = 0; I < 3; ++I) {
    const double h0 = e4phi * (G(I, 0) * b[0] + G(I, 1) * b[1] + G(I, 2) * b[2]);
    W[(W_H0 + I) * pad + x] = 4.0 * dtphi * h0 + e4phi * (DG(I, 0) * b[0] + DG(I, 1) * b[1] + DG(I, 2) * b[2] +
                                                          G(I, 0) * db[0] + G(I, 1) * db[1] + G(I, 2) * db[2]);
  }
  const double alpha = 1.0 + in.al[x];
  W[W_E * pad + x] = alpha * alpha - 1.0 - G(0, 0) * b[0] * b[0] - G(1, 1) * b[1] * b[1] - G(2, 2) * b[2] * b[2] -
                     2.0 * (G(0, 1) * b[0] * b[1] + G(0, 2) * b[0] * b[2] + G(1, 2) * b[1] * b[2]);
}

// d_i d_j of h, dt h, d2t h summed (bardeen.cc:221-239) into A, dt_A, d2t_A
template <int R>
__global__ void k_brd_ddh(Sten<R> S, double* W, long pad, long n) {
  PT_INDEX();
  for (int q = 0; q < 3; ++q) {
    const int base = q == 0 ? W_H : q == 1 ? W_DTH : W_D2TH;
    auto T = [&](int s) { return W + (base + s) * pad; };
    W[(W_A + q) * pad + x] = S.dd(T(0), i, j, k, 0, 0) + S.dd(T(3), i, j, k, 1, 1) + S.dd(T(5), i, j, k, 2, 2) +
                             2.0 * (S.dd(T(1), i, j, k, 0, 1) + S.dd(T(2), i, j, k, 0, 2) + S.dd(T(4), i, j, k, 1, 2));
  }
}

// (A.3) of bardeen.cc:243-261
__global__ void k_brd_A(BrdBg bg, double* W, long pad, long n) {
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const double a = bg.a, H = bg.H, d2adt2 = bg.d2adt2;
  auto tr = [&](int base) { return W[(base + 0) * pad + x] + W[(base + 3) * pad + x] + W[(base + 5) * pad + x]; };
  const double A = (tr(W_H) - W[W_A * pad + x]) / a / a / 2.0;
  const double dtA = -2.0 * A * H + (tr(W_DTH) - W[W_DTA * pad + x]) / a / a / 2.0;
  const double d2tA = -2.0 * A * (H * H + d2adt2 / a) - 4.0 * H * dtA + (tr(W_D2TH) - W[W_D2TA * pad + x]) / a / a / 2.0;
  W[W_A * pad + x] = A;
  W[W_DTA * pad + x] = dtA;
  W[W_D2TA * pad + x] = d2tA;
}

// monopole fix of A (bardeen.cc:262-285), then the right sides of B, dt B,
// d2t B (:290-304) and of F, dt F (:312-322)
template <int R>
__global__ void k_brd_B(Sten<R> S, BrdBg bg, double dA0, double dA1, double dA2, double* W, long pad, long n) {
  PT_INDEX();
  const double a = bg.a, dadt = bg.dadt, d2adt2 = bg.d2adt2, H = bg.H;
  const double A = W[W_A * pad + x] + dA0, dtA = W[W_DTA * pad + x] + dA1, d2tA = W[W_D2TA * pad + x] + dA2;
  W[W_A * pad + x] = A;
  W[W_DTA * pad + x] = dtA;
  W[W_D2TA * pad + x] = d2tA;
  auto tr = [&](int base) { return W[(base + 0) * pad + x] + W[(base + 3) * pad + x] + W[(base + 5) * pad + x]; };
  const double h_tr = tr(W_H), dt_h_tr = tr(W_DTH), d2t_h_tr = tr(W_D2TH);
  W[W_B * pad + x] = h_tr / a / a - 3.0 * A;
  W[W_DTB * pad + x] = dt_h_tr / a / a - 2.0 / a / a / a * dadt * h_tr - 3.0 * dtA;
  W[W_D2TB * pad + x] = d2t_h_tr / a / a - 4.0 / a / a / a * dadt * dt_h_tr - 2.0 / a / a / a * d2adt2 * h_tr +
                        6.0 / a / a / a / a * dadt * dadt * h_tr - 3.0 * d2tA;
  const double F = (S.d1(W + (W_H0 + 0) * pad, i, j, k, 0) + S.d1(W + (W_H0 + 1) * pad, i, j, k, 1) +
                    S.d1(W + (W_H0 + 2) * pad, i, j, k, 2)) / a;
  W[W_F * pad + x] = F;
  // d_t h_0i is never set in the reference (zero): its divergence is 0
  W[W_DTF * pad + x] = (0.0 + 0.0 + 0.0) / a - H * F;
}

// Phi, Psi (bardeen.cc:329-340), G_i and the right sides of C_i, dt C_i
// (:344-382)
template <int R>
__global__ void k_brd_C(Sten<R> S, BrdBg bg, double* W, long pad, long n) {
  PT_INDEX();
  const double a = bg.a, dadt = bg.dadt, H = bg.H;
  auto w = [&](int q) { return W + (long)q * pad; };
  const double A = w(W_A)[x], F = w(W_F)[x], dtF = w(W_DTF)[x], dtB = w(W_DTB)[x], d2tB = w(W_D2TB)[x];
  w(W_PSI)[x] = -1.0 / 2.0 * (A + 2.0 * H * (a * F - 0.5 * a * a * dtB));
  w(W_PHI)[x] = 1.0 / 2.0 * (w(W_E)[x] + 2.0 * (dadt * F + a * dtF) - 2.0 * a * dadt * dtB - a * a * d2tB);
  for (int I = 0; I < 3; ++I) {
    const int p = I == 0 ? 1 : 0, q = I == 2 ? 1 : 2;
    w(W_G1 + I)[x] = S.d1(w(W_F), i, j, k, I) - w(W_H0 + I)[x] / a;
    const double dA = S.d1(w(W_A), i, j, k, I);
    const double c = (S.d1(w(W_H + sy2(I, p)), i, j, k, p) + S.d1(w(W_H + sy2(I, q)), i, j, k, q) -
                      S.d1(w(W_H + sy2(p, p)), i, j, k, I) - S.d1(w(W_H + sy2(q, q)), i, j, k, I)) / a / a +
                     2.0 * dA;
    w(W_C1 + I)[x] = c;
    w(W_DTC1 + I)[x] = (S.d1(w(W_DTH + sy2(I, p)), i, j, k, p) + S.d1(w(W_DTH + sy2(I, q)), i, j, k, q) -
                        S.d1(w(W_DTH + sy2(p, p)), i, j, k, I) - S.d1(w(W_DTH + sy2(q, q)), i, j, k, I)) / a / a -
                       2.0 * H * (c - 2.0 * dA) + 2.0 * S.d1(w(W_DTA), i, j, k, I);
  }
}

// Vmag, D_ij (bardeen.cc:384-409), lin_viol and its magnitude (:412-420)
template <int R>
__global__ void k_brd_D(Sten<R> S, BrdBg bg, double* W, long pad, long n) {
  PT_INDEX();
  const double a = bg.a, dadt = bg.dadt;
  auto w = [&](int q) { return W + (long)q * pad; };
  double vm = 0.0;
  for (int I = 0; I < 3; ++I) {
    const double t = w(W_G1 + I)[x] - a * w(W_DTC1 + I)[x];
    vm += t * t;
  }
  w(W_VMAG)[x] = sqrt(vm);
  for (int I = 0; I < 3; ++I)
    for (int J = I; J < 3; ++J)
      w(W_D11 + sy2(I, J))[x] = w(W_H + sy2(I, J))[x] / a / a - (I == J ? 1.0 : 0.0) * w(W_A)[x] -
                                S.dd(w(W_B), i, j, k, I, J) - S.d1(w(W_C1 + J), i, j, k, I) -
                                S.d1(w(W_C1 + I), i, j, k, J);
  const double E = w(W_E)[x], A = w(W_A)[x], d2tB = w(W_D2TB)[x], dtB = w(W_DTB)[x], dtF = w(W_DTF)[x],
               F = w(W_F)[x];
  w(W_LV)[x] = E + A - a * a * d2tB - 3.0 * a * dadt * dtB + 2.0 * a * dtF + 4.0 * dadt * F;
  const double t2 = a * a * d2tB, t3 = 3.0 * a * dadt * dtB, t4 = 2.0 * a * dtF, t5 = 4.0 * dadt * F;
  w(W_LVMAG)[x] = sqrt(E * E + A * A + t2 * t2 + t3 * t3 + t4 * t4 + t5 * t5);
  w(W_LVDMAG)[x] = S.abs_dder(w(W_E), i, j, k) + S.abs_dder(w(W_A), i, j, k) + a * a * S.abs_dder(w(W_D2TB), i, j, k) +
                   fabs(3.0 * a * dadt) * S.abs_dder(w(W_DTB), i, j, k) + 2.0 * a * S.abs_dder(w(W_DTF), i, j, k) +
                   fabs(4.0 * dadt) * S.abs_dder(w(W_F), i, j, k);
}

template <int R>
__global__ void k_brd_lvd(Sten<R> S, double* W, long pad, long n) {
  PT_INDEX();
  W[W_LVD * pad + x] = S.abs_dder(W + W_LV * pad, i, j, k);
}

// Fourier::inverseLaplacian's mode division: -(2 pi |p| / L)^2, mean zeroed
__global__ void k_inv_lap_modes(double2* s, int nx, int ny, int nz, double box, double npts) {
  const int nzc = nz / 2 + 1;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)nx * ny * nzc) return;
  const int kz = (int)(t % nzc), jy = (int)((t / nzc) % ny), ix = (int)(t / ((long)nzc * ny));
  const double px = ix <= nx / 2 ? ix : ix - nx, py = jy <= ny / 2 ? jy : jy - ny, pz = kz;
  const double pmag = sqrt(px * px + py * py + pz * pz) * 2.0 * M_PI / box;
  const double d = -pmag * pmag * npts;
  if (t == 0) {
    s[t] = double2{0.0, 0.0};
  } else {
    s[t].x /= d;
    s[t].y /= d;
  }
}

// deterministic partial reductions: op 0 sum x, 1 sum w x, 2 sum w, 3 max x,
// 4 sum (m - x)^2; one value per block, combined on the host in block order
__global__ void __launch_bounds__(kBT) k_brd_red(const double* __restrict__ x, const double* __restrict__ w, long n,
                                                 int op, double m, double* __restrict__ part) {
  __shared__ double sh[kBT];
  const long per = (n + gridDim.x - 1) / gridDim.x;
  const long lo = (long)blockIdx.x * per, hi = lo + per < n ? lo + per : n;
  double acc = op == 3 ? -__builtin_huge_val() : 0.0;
  for (long e = lo + threadIdx.x; e < hi; e += kBT) {
    const double v = x ? x[e] : 0.0;
    if (op == 0) acc += v;
    else if (op == 1) acc += exp(6.0 * w[e]) * v;
    else if (op == 2) acc += exp(6.0 * w[e]);
    else if (op == 3) acc = v > acc ? v : acc;
    else acc += (m - v) * (m - v);
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kBT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const double o = sh[threadIdx.x + s];
      sh[threadIdx.x] = op == 3 ? (o > sh[threadIdx.x] ? o : sh[threadIdx.x]) : sh[threadIdx.x] + o;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// ------------------------------------------------------ matter + Lambda --
// bardeen.h:176-229 (units H_I = 1, a_I = 1)
double ml_initial_time(double OL) {
  if (OL == 0) return 2.0 / 3.0;
  const double s = std::sqrt(OL);
  return std::log1p(2.0 / (1.0 - s) * s) / (3.0 * s);
}
double ml_scale_factor(double t, double OL) {
  if (OL == 0) {
    const double tI = ml_initial_time(OL);
    return std::pow((tI + t) / tI, 2.0 / 3.0);
  }
  const double s = std::sqrt(OL);
  return std::pow(std::cosh(3.0 * s * t / 2.0) + std::sinh(3.0 * s * t / 2.0) / s, 2.0 / 3.0);
}
double ml_hubble(double t, double OL) {
  const double a = ml_scale_factor(t, OL);
  return std::sqrt((1 - OL) * std::pow(a, -3.0) + OL);
}
double ml_d2adt2(double t, double OL) {
  const double a = ml_scale_factor(t, OL), H = ml_hubble(t, OL);
  return a * H * H - (1.0 - OL) * 3.0 / 2.0 / a / a;
}

#define FFTCHK(ctx, expr)                                                                          \
  do {                                                                                             \
    hipfftResult r_ = (expr);                                                                      \
    if (r_ != HIPFFT_SUCCESS) return fail(ctx, CG_ERR_HIP, std::string(#expr) + " failed: " + std::to_string((int)r_)); \
  } while (0)

struct Red {
  cg_ctx* c;
  BrdWork* w;
  int nb;
  long n;
  // one reduction (see k_brd_red), synchronous
  int run(const double* x, const double* wgt, int op, double m, double* out) {
    hipLaunchKernelGGL(k_brd_red, dim3(nb), dim3(kBT), 0, c->stream, x, wgt, n, op, m, w->part);
    HIPCHK(c, hipGetLastError());
    std::vector<double> p(nb);
    HIPCHK(c, hipMemcpyAsync(p.data(), w->part, nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double acc = op == 3 ? -HUGE_VAL : 0.0;
    for (int b = 0; b < nb; ++b) acc = op == 3 ? (p[b] > acc ? p[b] : acc) : acc + p[b];
    *out = acc;
    return CG_OK;
  }
  double avg(const double* x, int* rc) {
    double s = 0;
    *rc = run(x, nullptr, 0, 0, &s);
    return s / (double)n;
  }
  // conformal_average (math.h:1382-1415) with phi_FRW = 0
  double cavg(const double* x, const double* dphi, int* rc) {
    double s = 0, v = 0;
    if ((*rc = run(x, dphi, 1, 0, &s))) return 0;
    if ((*rc = run(nullptr, dphi, 2, 0, &v))) return 0;
    return s / (double)n / (v / (double)n);
  }
  double vmax(const double* x, int* rc) {
    double s = 0;
    *rc = run(x, nullptr, 3, 0, &s);
    return s;
  }
  double sdev(const double* x, double m, int* rc) {
    double s = 0;
    *rc = run(x, nullptr, 4, m, &s);
    return std::sqrt(s / (double)(n - 1));
  }
};

template <int R>
int bardeen_t(cg_ctx* c, BrdWork* w, const BrdIn& in, double t, int use_ml, double OL, double viols[17]) {
  const long n = c->npts, pad = c->rstride;
  double* W = w->arr;
  Sten<R> S{c->p.nx, c->p.ny, c->p.nz, 1.0 / c->p.dx, 1.0 / (c->p.dx * c->p.dx)};
  const dim3 grid((unsigned)((n + kBT - 1) / kBT)), blk(kBT);
  Red red{c, w, (int)std::min<long>(kRedBlocks, (n + kBT - 1) / kBT), n};
  int rc = 0;
  const double box = c->p.nx * c->p.dx;  // H_LEN_FRAC
  auto invlap = [&](int q) -> int {
    FFTCHK(c, hipfftExecD2Z(w->r2c, W + (long)q * pad, (hipfftDoubleComplex*)w->spec));
    const long nm = (long)c->p.nx * c->p.ny * (c->p.nz / 2 + 1);
    hipLaunchKernelGGL(k_inv_lap_modes, dim3((unsigned)((nm + kBT - 1) / kBT)), blk, 0, c->stream, w->spec, c->p.nx,
                       c->p.ny, c->p.nz, box, (double)n);
    HIPCHK(c, hipGetLastError());
    FFTCHK(c, hipfftExecZ2D(w->c2r, (hipfftDoubleComplex*)w->spec, W + (long)q * pad));
    return CG_OK;
  };

  // background (bardeen.cc:41-73)
  const double alpha_avg = 1.0 + red.cavg(in.al, in.phi, &rc);
  if (rc) return rc;
  const double K_avg = red.cavg(in.K, in.phi, &rc);
  if (rc) return rc;
  const double phi_avg = red.cavg(in.phi, in.phi, &rc);
  if (rc) return rc;
  BrdBg bg{};
  if (use_ml) {
    bg.a = ml_scale_factor(t, OL);
    bg.H = ml_hubble(t, OL);
    bg.dadt = bg.H * bg.a;
    bg.d2adt2 = ml_d2adt2(t, OL);
  } else {
    bg.a = std::exp(2.0 * phi_avg);
    bg.dadt = -1.0 / 3.0 * bg.a * alpha_avg * K_avg;
    bg.H = bg.dadt / bg.a;
  }
  hipLaunchKernelGGL((k_brd_d2t<R>), grid, blk, 0, c->stream, S, in, W, pad, n);
  HIPCHK(c, hipGetLastError());
  const double cavg_d2tphi = red.cavg(W + W_D2TPHI * pad, in.phi, &rc);
  if (rc) return rc;
  if (!use_ml) bg.d2adt2 = bg.H * bg.dadt + 2.0 * bg.a * cavg_d2tphi;
  hipLaunchKernelGGL(k_brd_h, grid, blk, 0, c->stream, in, bg, W, pad, n);
  hipLaunchKernelGGL((k_brd_ddh<R>), grid, blk, 0, c->stream, S, W, pad, n);
  HIPCHK(c, hipGetLastError());
  for (int q = 0; q < 3; ++q)
    if ((rc = invlap(W_A + q))) return rc;
  hipLaunchKernelGGL(k_brd_A, grid, blk, 0, c->stream, bg, W, pad, n);
  HIPCHK(c, hipGetLastError());
  // monopoles (bardeen.cc:262-285)
  const double a = bg.a, dadt = bg.dadt, d2adt2 = bg.d2adt2;
  double avgA[3], avgtr[3];
  for (int q = 0; q < 3; ++q) {
    avgA[q] = red.avg(W + (W_A + q) * pad, &rc);
    if (rc) return rc;
    const int base = q == 0 ? W_H : q == 1 ? W_DTH : W_D2TH;
    double s = 0;
    for (int d : {0, 3, 5}) {
      double sd = 0;
      if ((rc = red.run(W + (base + d) * pad, nullptr, 0, 0, &sd))) return rc;
      s += sd;
    }
    avgtr[q] = s / (double)n;
  }
  const double dA0 = avgtr[0] / 3.0 / a / a - avgA[0];
  const double dA1 = avgtr[1] / 3.0 / a / a - 2.0 * avgtr[0] * dadt / 3.0 / a / a / a - avgA[1];
  const double dA2 = avgtr[2] / 3.0 / a / a - 2.0 * dadt * avgtr[1] / 3.0 / a / a / a +
                     6.0 * avgtr[0] * dadt * dadt / 3.0 / a / a / a / a - 2.0 * avgtr[0] * d2adt2 / 3.0 / a / a / a -
                     2.0 * avgtr[1] * dadt / 3.0 / a / a / a - avgA[2];
  hipLaunchKernelGGL((k_brd_B<R>), grid, blk, 0, c->stream, S, bg, dA0, dA1, dA2, W, pad, n);
  HIPCHK(c, hipGetLastError());
  for (int q : {W_B, W_DTB, W_D2TB, W_F, W_DTF})
    if ((rc = invlap(q))) return rc;
  hipLaunchKernelGGL((k_brd_C<R>), grid, blk, 0, c->stream, S, bg, W, pad, n);
  HIPCHK(c, hipGetLastError());
  for (int q = 0; q < 3; ++q) {
    if ((rc = invlap(W_C1 + q))) return rc;
    if ((rc = invlap(W_DTC1 + q))) return rc;
  }
  hipLaunchKernelGGL((k_brd_D<R>), grid, blk, 0, c->stream, S, bg, W, pad, n);
  hipLaunchKernelGGL((k_brd_lvd<R>), grid, blk, 0, c->stream, S, W, pad, n);
  HIPCHK(c, hipGetLastError());

  // violations (bardeen.cc:422-455)
  const double mean_viol_der = red.avg(W + W_LVD * pad, &rc);
  if (rc) return rc;
  const double viol_scale_der = red.avg(W + W_LVDMAG * pad, &rc);
  if (rc) return rc;
  const double max_viol_der = red.vmax(W + W_LVD * pad, &rc);
  if (rc) return rc;
  const double std_viol_der = red.sdev(W + W_LVD * pad, mean_viol_der, &rc);
  if (rc) return rc;
  const double mean_viol = red.avg(W + W_LV * pad, &rc);
  if (rc) return rc;
  const double viol_scale = red.avg(W + W_LVMAG * pad, &rc);
  if (rc) return rc;
  const double max_viol = red.vmax(W + W_LV * pad, &rc);
  if (rc) return rc;
  const double std_viol = red.sdev(W + W_LV * pad, mean_viol, &rc);
  if (rc) return rc;
  const double avg_phi = red.avg(in.phi, &rc);
  if (rc) return rc;
  const double avg_K = red.avg(in.K, &rc);
  if (rc) return rc;
  const double avg_d2tphi = red.avg(W + W_D2TPHI * pad, &rc);
  if (rc) return rc;
  const double mlA = ml_scale_factor(t, OL), mlH = ml_hubble(t, OL), mld2 = ml_d2adt2(t, OL);
  viols[0] = mean_viol_der / viol_scale_der;
  viols[1] = std_viol_der / viol_scale_der;
  viols[2] = max_viol_der / viol_scale_der;
  viols[3] = max_viol_der;
  viols[4] = mean_viol / viol_scale;
  viols[5] = std_viol / viol_scale;
  viols[6] = max_viol / viol_scale;
  viols[7] = max_viol;
  viols[8] = 1.0 - mlA / std::exp(2.0 * phi_avg);
  viols[9] = 1.0 - mlH / (-1.0 / 3.0 * a * alpha_avg * K_avg);
  viols[10] = 1.0 - mld2 / (bg.H * dadt + 2.0 * a * cavg_d2tphi);
  viols[11] = 1.0 - mlA / std::exp(2.0 * avg_phi);
  viols[12] = 1.0 - mlH / (-1.0 / 3.0 * a * alpha_avg * avg_K);
  viols[13] = 1.0 - mld2 / (bg.H * dadt + 2.0 * a * avg_d2tphi);
  viols[14] = red.avg(W + W_VMAG * pad, &rc);
  if (rc) return rc;
  viols[15] = red.sdev(W + W_VMAG * pad, viols[14], &rc);
  if (rc) return rc;
  viols[16] = red.vmax(W + W_VMAG * pad, &rc);
  return rc;
}

const char* kBrdNames[] = {"Phi", "Psi", "A", "dt_A", "d2t_A", "B", "dt_B", "d2t_B", "E", "F", "G1", "G2", "G3",
                           "C1", "C2", "C3", "dt_C1", "dt_C2", "dt_C3", "Vmag", "D11", "D12", "D13", "D22", "D23",
                           "D33"};
const int kBrdIdx[] = {W_PHI, W_PSI, W_A, W_DTA, W_D2TA, W_B, W_DTB, W_D2TB, W_E, W_F, W_G1, W_G2, W_G3,
                       W_C1, W_C2, W_C3, W_DTC1, W_DTC2, W_DTC3, W_VMAG, W_D11, W_D12, W_D13, W_D22, W_D23,
                       W_D33};

}  // namespace

void bardeen_free(cg_ctx* c) {
  BrdWork* w = c->brd;
  if (!w) return;
  if (w->plans) {
    (void)hipfftDestroy(w->r2c);
    (void)hipfftDestroy(w->c2r);
  }
  if (w->arr) (void)hipFree(w->arr);
  if (w->spec) (void)hipFree(w->spec);
  if (w->part) (void)hipFree(w->part);
  delete w;
  c->brd = nullptr;
}

const double* bardeen_field(const cg_ctx* c, const std::string& name, int* err) {
  *err = 0;
  for (size_t q = 0; q < sizeof(kBrdIdx) / sizeof(kBrdIdx[0]); ++q) {
    if (name != kBrdNames[q]) continue;
    if (!c->brd || !c->brd->valid) { *err = 4; return nullptr; }
    return c->brd->arr + (long)kBrdIdx[q] * c->rstride;
  }
  *err = 1;
  return nullptr;
}

// the _c register of the evaluated state holds the BSSN right-hand side at
// that state (RKEvolve's output); used by cg_bardeen
int rhs_at(cg_ctx* c, int rin, int rout);

}  // namespace cg

using namespace cg;

int cg_bardeen(cg_ctx* c, double elapsed_t, int use_ml_scale_factor, double omega_L, double viols[17]) {
  if (!c || !viols) return CG_ERR_ARG;
  if (c->met32) return fail(c, CG_ERR_ARG, "the Bardeen potentials read the FP64 metric (not in the all-FP32 mode)");
  if (use_ml_scale_factor && (omega_L < 0 || omega_L > 1)) return fail(c, CG_ERR_ARG, "Invalid Omega_Lambda specified.");
  if (c->frw.g_K != 0.0 || c->frw.g_phi != 0.0)
    return fail(c, CG_ERR_STATE,
                "Bardeen variables incompatible with use of reference metric. SVT and Bardeen fields will not be "
                "computed.");
  if (c->p.nz % 2) return fail(c, CG_ERR_ARG, "the Bardeen potentials need an even nz");
  // the state the output step sees, and the right-hand side there into _c
  int rin;
  if (int rc = cons_input(c, &rin)) return rc;
  const int rout = c->sC;
  if (int rc = rhs_at(c, rin, rout)) return rc;
  const long pad = c->rstride, n = c->npts;
  if (!c->brd) {
    BrdWork* w = new BrdWork();
    c->brd = w;
    HIPCHK(c, hipMalloc((void**)&w->arr, (size_t)W_N * pad * sizeof(double)));
    HIPCHK(c, hipMalloc((void**)&w->spec, (size_t)c->p.nx * c->p.ny * (c->p.nz / 2 + 1) * sizeof(double2)));
    HIPCHK(c, hipMalloc((void**)&w->part, kRedBlocks * sizeof(double)));
    FFTCHK(c, hipfftPlan3d(&w->r2c, c->p.nx, c->p.ny, c->p.nz, HIPFFT_D2Z));
    FFTCHK(c, hipfftPlan3d(&w->c2r, c->p.nx, c->p.ny, c->p.nz, HIPFFT_Z2D));
    w->plans = true;
  }
  BrdWork* w = c->brd;
  w->valid = false;
  FFTCHK(c, hipfftSetStream(w->r2c, c->stream));
  FFTCHK(c, hipfftSetStream(w->c2r, c->stream));
  // inputs in FP64: the metric registers directly, the storage-type ones
  // widened into the workspace in split precision
  BrdIn in{};
  auto met = [&](int f, int s) { return (const double*)reg_ptr(c, f, s); };
  for (int q = 0; q < 6; ++q) { in.g[q] = met(G11 + q, rin); in.cg[q] = met(G11 + q, rout); }
  in.phi = met(PHI, rin);
  in.cphi = met(PHI, rout);
  const int sto[8] = {A11, A12, A13, A22, A23, A33, KK, ALPHA};
  const double* st_in[8];
  const double* st_out[8];
  for (int q = 0; q < 8; ++q) {
    if (c->prec == CG_FP64) {
      st_in[q] = (const double*)reg_ptr(c, sto[q], rin);
      st_out[q] = (const double*)reg_ptr(c, sto[q], rout);
    } else {
      double* di = w->arr + (long)(W_SA + q) * pad;
      double* dO = w->arr + (long)(W_CA + q) * pad;
      hipLaunchKernelGGL(k_widen<float>, dim3((unsigned)((n + kBT - 1) / kBT)), dim3(kBT), 0, c->stream,
                         (const float*)reg_ptr(c, sto[q], rin), di, n);
      hipLaunchKernelGGL(k_widen<float>, dim3((unsigned)((n + kBT - 1) / kBT)), dim3(kBT), 0, c->stream,
                         (const float*)reg_ptr(c, sto[q], rout), dO, n);
      st_in[q] = di;
      st_out[q] = dO;
    }
  }
  HIPCHK(c, hipGetLastError());
  for (int q = 0; q < 6; ++q) { in.A[q] = st_in[q]; in.cA[q] = st_out[q]; }
  in.K = st_in[6]; in.cK = st_out[6];
  in.al = st_in[7]; in.cal = st_out[7];
  if (c->nxf) {
    for (int m = 0; m < 3; ++m) {
      in.beta[m] = c->xreg + ((long)(XBETA1 + m) * 4 + rin) * pad;
      in.cbeta[m] = c->xreg + ((long)(XBETA1 + m) * 4 + rout) * pad;
    }
  }
  int rc;
  switch (c->p.stencil_order) {
    case 2: rc = bardeen_t<1>(c, w, in, elapsed_t, use_ml_scale_factor, omega_L, viols); break;
    case 4: rc = bardeen_t<2>(c, w, in, elapsed_t, use_ml_scale_factor, omega_L, viols); break;
    case 6: rc = bardeen_t<3>(c, w, in, elapsed_t, use_ml_scale_factor, omega_L, viols); break;
    default: rc = bardeen_t<4>(c, w, in, elapsed_t, use_ml_scale_factor, omega_L, viols); break;
  }
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  w->valid = true;
  return CG_OK;
}
