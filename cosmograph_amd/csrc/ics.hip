// ics.hip -- static-dust initial data on the device (hipFFT).
//
// The reference builds its Gaussian random field on the host and inverts it
// with FFTW (ICs/ICs.cc:38-134, set_gaussian_random_Phi_N); its static-dust
// ICs then go from a chosen phi to rho (components/static/static_ic.cc) and
// its constraint solver is an absent submodule.  This build's recipe (also
// in cosmograph_amd/ics.py, the host version the tests compare with):
//   1. delta = inverse real FFT of the caller's half-spectrum, with the
//      kz = 0 and kz = nz/2 planes replaced by their Hermitian part (what a
//      real inverse transform of those planes means), mean removed, scaled
//      to the requested rms; drho = rho_bar * delta;
//   2. conformally flat metric e^{4 phi} delta_ij, K = K_FRW, A_ij = 0:
//      the Hamiltonian constraint (bssn.cc:1081-1083) is
//      lap(u) = -2 pi u^5 (drho - c), u = e^phi, with the shift c making
//      <u^5 (drho - c)> = 0; fixed-point iteration
//      u <- 1 + lap^{-1}[-2 pi u^5 (drho - c)] (spectral inverse Laplacian)
//      until max|du| < tol;
//   3. DIFFphi_p = log u, DIFFD_a = e^{6 phi}(rho_bar + drho - c) - rho_bar
//      (the conserved density of components/static/static.cc:32-33), every
//      other _p register zero, FRW phi = 0, K = -sqrt(24 pi rho_bar), one
//      w = 0 fluid of density rho_bar (components/static/static_ic.cc:372-378);
//      for a dust-fluid context the same geometry as full fields with an
//      empty FRW reference (dust_ic.cc): dust_D_p = e^{6 phi} rho at rest and
//      DIFFK_p = -sqrt(24 pi (rho_bar + lambda)); with the dust background on,
//      DIFFK_p = 0 and that K with the fluids rho_bar (w = 0) and lambda
//      (w = -1) in the FRW reference.
// Reductions are per-block partials in a fixed order, summed on the host in
// block order: the result is deterministic.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/cosmo_gpu.h"
#include "runtime.h"

using namespace cg;

namespace {

constexpr int kT = 256;     // threads per block
constexpr int kB = 1024;    // blocks of the reduction kernels
constexpr double kPi = 3.14159265358979323846;

__device__ __forceinline__ void block_reduce2(double s0, double s1, double* part) {
  __shared__ double sh[2][kT];
  sh[0][threadIdx.x] = s0;
  sh[1][threadIdx.x] = s1;
  __syncthreads();
  for (int w = kT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sh[0][0];
    part[2 * blockIdx.x + 1] = sh[1][0];
  }
}

// planes kz of the half-spectrum [nx][ny][nzc] replaced by their Hermitian
// part: A(k) <- (A(k) + conj A(-k)) / 2 over (kx, ky); each pair is handled
// by the thread of its lower index
__global__ void k_hermitize(double2* A, int nx, int ny, int nzc, int kz) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= (long)nx * ny) return;
  const int kx = (int)(q / ny), ky = (int)(q % ny);
  const long p = (long)((nx - kx) % nx) * ny + (ny - ky) % ny;
  if (p < q) return;
  double2* a = A + (q * nzc + kz);
  double2* b = A + (p * nzc + kz);
  const double2 va = *a, vb = *b;
  const double2 h = make_double2(0.5 * (va.x + vb.x), 0.5 * (va.y - vb.y));
  *a = h;
  *b = make_double2(h.x, -h.y);
}

// sum of x
__global__ void k_sum(const double* __restrict__ x, long n, double* part) {
  double s = 0.0;
  for (long i = blockIdx.x * (long)kT + threadIdx.x; i < n; i += (long)kT * gridDim.x) s += x[i];
  block_reduce2(s, 0.0, part);
}

// x -= m, sum of the centred squares
__global__ void k_center(double* __restrict__ x, long n, double m, double* part) {
  double s = 0.0;
  for (long i = blockIdx.x * (long)kT + threadIdx.x; i < n; i += (long)kT * gridDim.x) {
    const double v = x[i] - m;
    x[i] = v;
    s += v * v;
  }
  block_reduce2(s, 0.0, part);
}

__global__ void k_scale(double* __restrict__ x, long n, double f) {
  for (long i = blockIdx.x * (long)kT + threadIdx.x; i < n; i += (long)kT * gridDim.x) x[i] *= f;
}

__global__ void k_fill(double* __restrict__ x, long n, double v) {
  for (long i = blockIdx.x * (long)kT + threadIdx.x; i < n; i += (long)kT * gridDim.x) x[i] = v;
}

// sums of u^5 drho and u^5
__global__ void k_u5(const double* __restrict__ u, const double* __restrict__ drho, long n, double* part) {
  double s0 = 0.0, s1 = 0.0;
  for (long i = blockIdx.x * (long)kT + threadIdx.x; i < n; i += (long)kT * gridDim.x) {
    const double v = u[i], u5 = v * v * v * v * v;
    s0 += u5 * drho[i];
    s1 += u5;
  }
  block_reduce2(s0, s1, part);
}

// rhs = -2 pi u^5 (drho - c)
__global__ void k_rhs(const double* __restrict__ u, const double* __restrict__ drho, double c, long n,
                      double* __restrict__ rhs) {
  for (long i = blockIdx.x * (long)kT + threadIdx.x; i < n; i += (long)kT * gridDim.x) {
    const double v = u[i], u5 = v * v * v * v * v;
    rhs[i] = -2.0 * kPi * u5 * (drho[i] - c);
  }
}

// spectrum *= -1/k^2 / N (k = 0 -> 0); integer wavenumbers in FFT order
__global__ void k_green(double2* __restrict__ S, int nx, int ny, int nz, double wx, double wy, double wz,
                        double inv_n) {
  const int nzc = nz / 2 + 1;
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= (long)nx * ny * nzc) return;
  const int kz = (int)(q % nzc), ky = (int)((q / nzc) % ny), kx = (int)(q / ((long)nzc * ny));
  const double fx = kx <= nx / 2 ? kx : kx - nx, fy = ky <= ny / 2 ? ky : ky - ny, fz = kz;
  const double k2 = wx * fx * fx + wy * fy * fy + wz * fz * fz;
  const double g = k2 > 0.0 ? -1.0 / k2 * inv_n : 0.0;
  S[q] = make_double2(S[q].x * g, S[q].y * g);
}

// u <- 1 + x, partial max |du|
__global__ void k_update(double* __restrict__ u, const double* __restrict__ x, long n, double* part) {
  double m = 0.0;
  for (long i = blockIdx.x * (long)kT + threadIdx.x; i < n; i += (long)kT * gridDim.x) {
    const double v = 1.0 + x[i];
    const double d = fabs(v - u[i]);
    m = d > m ? d : m;
    u[i] = v;
  }
  __shared__ double sh[kT];
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int w = kT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[2 * blockIdx.x] = sh[0];
}

// DIFFphi = log u and the conserved density e^{6 phi} rho - sub (static dust:
// sub = rho_bar, DIFFD; dust fluid at rest: sub = 0, D); kk (dust fluid
// only, else null) = the uniform extrinsic curvature
template <typename TS, typename MT>
__global__ void k_out(const double* __restrict__ u, const double* __restrict__ drho, double c, double rho_bar,
                      double sub, long n, MT* __restrict__ phi, TS* __restrict__ D, TS* __restrict__ kk, double K) {
  for (long i = blockIdx.x * (long)kT + threadIdx.x; i < n; i += (long)kT * gridDim.x) {
    const double p = log(u[i]);
    phi[i] = (MT)p;
    D[i] = (TS)(exp(6.0 * p) * ((rho_bar + drho[i]) - c) - sub);
    if (kk) kk[i] = (TS)K;
  }
}

struct Scratch {
  double* u = nullptr;
  double* drho = nullptr;
  double* work = nullptr;
  double2* spec = nullptr;
  double* part = nullptr;
  hipfftHandle r2c = 0, c2r = 0;
  bool have_r2c = false, have_c2r = false;
  ~Scratch() {
    if (have_r2c) (void)hipfftDestroy(r2c);
    if (have_c2r) (void)hipfftDestroy(c2r);
    for (void* p : {(void*)u, (void*)drho, (void*)work, (void*)spec, (void*)part})
      if (p) (void)hipFree(p);
  }
};

}  // namespace

#define FFTCHK(ctx, expr)                                                                          \
  do {                                                                                             \
    hipfftResult r_ = (expr);                                                                      \
    if (r_ != HIPFFT_SUCCESS) return fail(ctx, CG_ERR_HIP, std::string(#expr) + ": hipFFT error " + std::to_string((int)r_)); \
  } while (0)

int cg_ics_static_dust(cg_ctx* c, const double* modes, size_t n_modes, double rms, double tol, int max_iter,
                       double* shift, int* iters) {
  if (!c || !modes || rms < 0.0 || !(tol > 0.0) || max_iter < 1) return CG_ERR_ARG;
  const int nx = c->p.nx, ny = c->p.ny, nz = c->p.nz, nzc = nz / 2 + 1;
  if (nz % 2) return fail(c, CG_ERR_ARG, "the static-dust ICs need an even nz");
  const long n = c->npts, nspec = (long)nx * ny * nzc;
  if ((long)n_modes != 2 * nspec) return fail(c, CG_ERR_ARG, "modes must hold nx*ny*(nz/2+1) complex values");
  if (int rc = ext_resolve(c)) return rc;
  const double rho_bar = 3.0 / (8.0 * kPi);
  hipStream_t s = c->stream;
  Scratch w;
  HIPCHK(c, hipMalloc((void**)&w.u, n * sizeof(double)));
  HIPCHK(c, hipMalloc((void**)&w.drho, n * sizeof(double)));
  HIPCHK(c, hipMalloc((void**)&w.work, n * sizeof(double)));
  HIPCHK(c, hipMalloc((void**)&w.spec, nspec * sizeof(double2)));
  HIPCHK(c, hipMalloc((void**)&w.part, 2 * kB * sizeof(double)));
  FFTCHK(c, hipfftPlan3d(&w.r2c, nx, ny, nz, HIPFFT_D2Z));
  w.have_r2c = true;
  FFTCHK(c, hipfftPlan3d(&w.c2r, nx, ny, nz, HIPFFT_Z2D));
  w.have_c2r = true;
  FFTCHK(c, hipfftSetStream(w.r2c, s));
  FFTCHK(c, hipfftSetStream(w.c2r, s));
  std::vector<double> part(2 * kB);
  auto sums = [&](double* s0, double* s1) -> int {
    HIPCHK(c, hipMemcpyAsync(part.data(), w.part, part.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    double a = 0.0, b = 0.0;
    for (int q = 0; q < kB; ++q) { a += part[2 * q]; b += part[2 * q + 1]; }
    *s0 = a;
    if (s1) *s1 = b;
    return CG_OK;
  };
  const int gb = (int)((nspec + kT - 1) / kT);

  // 1. delta: inverse FFT of the Hermitian-projected spectrum, centred, scaled to rms
  HIPCHK(c, hipMemcpyAsync(w.spec, modes, nspec * sizeof(double2), hipMemcpyHostToDevice, s));
  const int gh = (int)(((long)nx * ny + kT - 1) / kT);
  hipLaunchKernelGGL(k_hermitize, dim3(gh), dim3(kT), 0, s, w.spec, nx, ny, nzc, 0);
  if (nz / 2 != 0) hipLaunchKernelGGL(k_hermitize, dim3(gh), dim3(kT), 0, s, w.spec, nx, ny, nzc, nz / 2);
  FFTCHK(c, hipfftExecZ2D(w.c2r, (hipfftDoubleComplex*)w.spec, w.drho));
  hipLaunchKernelGGL(k_sum, dim3(kB), dim3(kT), 0, s, w.drho, n, w.part);
  double tot = 0.0, sq = 0.0;
  if (int rc = sums(&tot, nullptr)) return rc;
  hipLaunchKernelGGL(k_center, dim3(kB), dim3(kT), 0, s, w.drho, n, tot / (double)n, w.part);
  if (int rc = sums(&sq, nullptr)) return rc;
  const double sd = std::sqrt(sq / (double)n);
  hipLaunchKernelGGL(k_scale, dim3(kB), dim3(kT), 0, s, w.drho, n, sd > 0.0 ? rho_bar * (rms / sd) : rho_bar);
  HIPCHK(c, hipGetLastError());

  // 2. the conformal factor
  const double Lx = nx * c->p.dx, Ly = ny * c->p.dx, Lz = nz * c->p.dx;
  const double wx = std::pow(2.0 * kPi / Lx, 2), wy = std::pow(2.0 * kPi / Ly, 2), wz = std::pow(2.0 * kPi / Lz, 2);
  hipLaunchKernelGGL(k_fill, dim3(kB), dim3(kT), 0, s, w.u, n, 1.0);
  double cs = 0.0;
  int it = 0;
  for (it = 1; it <= max_iter; ++it) {
    double s0 = 0.0, s1 = 0.0;
    hipLaunchKernelGGL(k_u5, dim3(kB), dim3(kT), 0, s, w.u, w.drho, n, w.part);
    if (int rc = sums(&s0, &s1)) return rc;
    cs = s0 / s1;
    hipLaunchKernelGGL(k_rhs, dim3(kB), dim3(kT), 0, s, w.u, w.drho, cs, n, w.work);
    FFTCHK(c, hipfftExecD2Z(w.r2c, w.work, (hipfftDoubleComplex*)w.spec));
    hipLaunchKernelGGL(k_green, dim3(gb), dim3(kT), 0, s, w.spec, nx, ny, nz, wx, wy, wz, 1.0 / (double)n);
    FFTCHK(c, hipfftExecZ2D(w.c2r, (hipfftDoubleComplex*)w.spec, w.work));
    hipLaunchKernelGGL(k_update, dim3(kB), dim3(kT), 0, s, w.u, w.work, n, w.part);
    HIPCHK(c, hipMemcpyAsync(part.data(), w.part, part.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    double d = 0.0;
    for (int q = 0; q < kB; ++q) d = part[2 * q] > d ? part[2 * q] : d;
    if (d < tol) break;
  }
  if (it > max_iter) it = max_iter;

  // 3. registers and background.  Static dust: the perturbation around the
  // FRW background.  Dust fluid: the full fields with an empty FRW reference,
  // as the reference's dust ICs leave it (components/dust_fluid/dust_ic.cc):
  // dust_D = e^{6 phi} rho (the fluid at rest, S_i = 0) and
  // DIFFK = -sqrt(24 pi (rho_bar + lambda)), lambda the Lambda energy density.
  const bool fluid = c->dreg != nullptr;
  const double K0 = -std::sqrt(24.0 * kPi * (rho_bar + (fluid ? c->lambda : 0.0)));
  // dust fluid as perturbations (cg_set_dust_background): the mean goes to
  // the FRW fluids (the dust and, for Lambda, a w = -1 fluid) and DIFFK = 0
  const bool bg = fluid && c->dust_bg;
  for (int f = 0; f < NF; ++f)
    if (f != PHI)
      HIPCHK(c, hipMemsetAsync(reg_ptr(c, f, c->sP), 0, n * (f <= PHI && !c->met32 ? sizeof(double) : c->ts), s));
  if (fluid)
    for (int q = 0; q < NDUST; ++q)
      HIPCHK(c, hipMemsetAsync((char*)c->dreg + ((long)q * 4 + c->sP) * c->rstride * (long)c->ts, 0, n * c->ts, s));
  else
    HIPCHK(c, hipMemsetAsync(c->DIFFD, 0, n * c->ts, s));
  void* phi = reg_ptr(c, PHI, c->sP);
  void* D = fluid ? (void*)((char*)c->dreg + (long)c->sP * c->rstride * (long)c->ts) : c->DIFFD;
  void* kk = fluid && !bg ? reg_ptr(c, KK, c->sP) : nullptr;
  const double sub = fluid ? 0.0 : rho_bar;
  if (c->mq) {
    // fixed-point metric: phi in FP64 first, then words at a scale that fits it
    if (!c->mq_tmp) HIPCHK(c, hipMalloc((void**)&c->mq_tmp, (size_t)c->rstride * sizeof(double)));
    hipLaunchKernelGGL((k_out<float, double>), dim3(kB), dim3(kT), 0, s, w.u, w.drho, cs, rho_bar, sub, n,
                       c->mq_tmp, (float*)D, (float*)kk, K0);
    HIPCHK(c, hipGetLastError());
    if (int rc = mq_store(c, PHI, c->sP, c->mq_tmp)) return rc;
  } else if (c->ts == 8)
    hipLaunchKernelGGL((k_out<double, double>), dim3(kB), dim3(kT), 0, s, w.u, w.drho, cs, rho_bar, sub, n,
                       (double*)phi, (double*)D, (double*)kk, K0);
  else if (c->met32)
    hipLaunchKernelGGL((k_out<float, float>), dim3(kB), dim3(kT), 0, s, w.u, w.drho, cs, rho_bar, sub, n,
                       (float*)phi, (float*)D, (float*)kk, K0);
  else
    hipLaunchKernelGGL((k_out<float, double>), dim3(kB), dim3(kT), 0, s, w.u, w.drho, cs, rho_bar, sub, n,
                       (double*)phi, (float*)D, (float*)kk, K0);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(s));
  c->frw = Frw{};
  if (!fluid || bg) {
    c->frw.set(0.0, K0);
    c->frw.add_fluid(rho_bar, 0.0);
    if (bg && c->lambda != 0.0) c->frw.add_fluid(c->lambda, -1.0);
  }
  c->ext_fresh = false;
  c->alpha_nonzero = false;
  c->cons_valid = false;
  if (shift) *shift = cs;
  if (iters) *iters = it;
  return CG_OK;
}

// ------------------------------------------------------ power spectrum --
namespace {

template <typename T>
__global__ void k_to_double(const T* __restrict__ f, long n, double* __restrict__ out) {
  for (long x = (long)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (long)gridDim.x * blockDim.x)
    out[x] = (double)f[x];
}

// Fourier::powerDump's angular binning (utils/Fourier.h:159-207) of the r2c
// half spectrum [nx][ny][nz/2+1]: |F|^2 of each mode into bin
// floor(|p|), counted twice for 0 < pz < nz/2 (the mode and its conjugate)
// and once at pz = 0 and pz = nz/2.  One row of nbins sums and counts per
// block (LDS atomics), summed on the host in block order.
constexpr int kPowThreads = 256;
__global__ void __launch_bounds__(kPowThreads) k_power_bins(const double2* __restrict__ spec, int nx, int ny, int nz,
                                                           int nbins, double* __restrict__ part_f,
                                                           unsigned* __restrict__ part_n) {
  extern __shared__ double sh[];
  double* f2 = sh;
  unsigned* cnt = reinterpret_cast<unsigned*>(sh + nbins);
  for (int q = threadIdx.x; q < nbins; q += kPowThreads) { f2[q] = 0.0; cnt[q] = 0u; }
  __syncthreads();
  const int nzc = nz / 2 + 1;
  const long total = (long)nx * ny * nzc;
  for (long x = (long)blockIdx.x * kPowThreads + threadIdx.x; x < total; x += (long)gridDim.x * kPowThreads) {
    const int k = (int)(x % nzc), j = (int)((x / nzc) % ny), i = (int)(x / ((long)nzc * ny));
    const int px = i <= nx / 2 ? i : i - nx, py = j <= ny / 2 ? j : j - ny, pz = k;
    const int bin = (int)sqrt((double)(px * px + py * py + pz * pz));
    const double2 v = spec[x];
    const double fp2 = v.x * v.x + v.y * v.y;
    const bool twice = k >= 1 && k < nz / 2;
    if (bin < nbins) {
      atomicAdd(&f2[bin], twice ? 2.0 * fp2 : fp2);
      atomicAdd(&cnt[bin], twice ? 2u : 1u);
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < nbins; q += kPowThreads) {
    part_f[(long)blockIdx.x * nbins + q] = f2[q];
    part_n[(long)blockIdx.x * nbins + q] = cnt[q];
  }
}

}  // namespace

int cg_power_spectrum(cg_ctx* c, const char* name, double* out, int nbins) {
  if (!c || !out) return CG_ERR_ARG;
  const int nx = c->p.nx, ny = c->p.ny, nz = c->p.nz;
  const int want = (int)(sqrt((double)nx * nx + (double)ny * ny + (double)nz * nz) / 2.0) + 1;
  if (nbins != want) return fail(c, CG_ERR_ARG, "nbins must be (int)(sqrt(NX^2+NY^2+NZ^2)/2) + 1 = " + std::to_string(want));
  if (nz % 2) return fail(c, CG_ERR_ARG, "the power spectrum needs an even nz");
  size_t esz;
  int e;
  const void* d = field_ptr(c, name, &esz, &e);
  if (!d) return fail(c, e == 3 || e == 5 ? CG_ERR_STATE : CG_ERR_ARG,
                      e == 3   ? std::string(name) + " is filled by cg_constraints"
                      : e == 5 ? std::string(name) + " is stored in fixed point: read it with cg_field_download"
                               : std::string("unknown field ") + name);
  const long n = c->npts, nspec = (long)nx * ny * (nz / 2 + 1);
  constexpr int kBlocks = 512;
  hipStream_t s = c->stream;
  Scratch w;
  HIPCHK(c, hipMalloc((void**)&w.work, n * sizeof(double)));
  HIPCHK(c, hipMalloc((void**)&w.spec, nspec * sizeof(double2)));
  HIPCHK(c, hipMalloc((void**)&w.part, (size_t)kBlocks * nbins * (sizeof(double) + sizeof(unsigned))));
  FFTCHK(c, hipfftPlan3d(&w.r2c, nx, ny, nz, HIPFFT_D2Z));
  w.have_r2c = true;
  FFTCHK(c, hipfftSetStream(w.r2c, s));
  const int nb = (int)std::min<long>(4096, (n + 255) / 256);
  if (esz == 8)
    hipLaunchKernelGGL(k_to_double<double>, dim3(nb), dim3(256), 0, s, (const double*)d, n, w.work);
  else
    hipLaunchKernelGGL(k_to_double<float>, dim3(nb), dim3(256), 0, s, (const float*)d, n, w.work);
  HIPCHK(c, hipGetLastError());
  FFTCHK(c, hipfftExecD2Z(w.r2c, w.work, (hipfftDoubleComplex*)w.spec));
  double* pf = w.part;
  unsigned* pn = reinterpret_cast<unsigned*>(w.part + (size_t)kBlocks * nbins);
  const size_t shmem = (size_t)nbins * (sizeof(double) + sizeof(unsigned));
  hipLaunchKernelGGL(k_power_bins, dim3(kBlocks), dim3(kPowThreads), shmem, s, w.spec, nx, ny, nz, nbins, pf, pn);
  HIPCHK(c, hipGetLastError());
  std::vector<double> hf((size_t)kBlocks * nbins);
  std::vector<unsigned> hn((size_t)kBlocks * nbins);
  HIPCHK(c, hipMemcpyAsync(hf.data(), pf, hf.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipMemcpyAsync(hn.data(), pn, hn.size() * sizeof(unsigned), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  for (int q = 0; q < nbins; ++q) {
    double f = 0.0;
    unsigned long long cnt = 0;
    for (int b = 0; b < kBlocks; ++b) {
      f += hf[(size_t)b * nbins + q];
      cnt += hn[(size_t)b * nbins + q];
    }
    out[q] = cnt ? f / (double)cnt : 0.0;  // utils/Fourier.h:209-220
  }
  return CG_OK;
}

