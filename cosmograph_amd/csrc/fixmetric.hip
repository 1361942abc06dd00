// fixmetric.hip -- the fixed-point conformal metric of CG_FP32_FIX.
//
// DIFFgamma_ij and DIFFphi are stored as 32-bit integer words w with one
// power-of-two scale per field: value = w * 2^e.  The stage kernels
// differentiate the words exactly in FP64 and write the RK4 results back
// rounded to the nearest word.  The host keeps every field's words below
// 2^MQ_BITS: cg_rk4_steps checks the largest |w| of each field every few
// steps and raises e (halving the words, exactly up to the rounding of the
// last bit) where a field has grown past it, or lowers e (exactly) where a
// field's words have fallen below 2^(MQ_BITS-3).  The values checked are
// both the state's and those of its unit-determinant normalisation (the
// stencil input of the next stage 1): normalisation shifts each diagonal
// component by t = 1 - det^(1/3), which can exceed the component's own
// size when another component is large.  Uploads and the device initial
// data set e to fit the values they write.  No exponent stays
// more than MQ_FLOOR below the largest one of the fields that hold data,
// so that a field starting from zero can grow without saturating before
// the next check.  Exponents start at kExpDefault.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "../../include/cosmo_gpu.h"
#include "bssn_stage.h"
#include "bssn_stage_kernel.h"  // norm_factor, diag6
#include "runtime.h"

namespace cg {
namespace {

constexpr int kMqBlocks = 1024, kMqThreads = 256;

// largest |x| of an FP64 array: one value per block
__global__ void k_absmax_f64(const double* __restrict__ x, long n, double* __restrict__ part) {
  __shared__ double sh[kMqThreads];
  double m = 0.0;
  for (long i = (long)blockIdx.x * kMqThreads + threadIdx.x; i < n; i += (long)kMqThreads * gridDim.x)
    m = fmax(m, fabs(x[i]));
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int w = kMqThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// largest |w| of na <= 7 word arrays at base + a*stride: part[block * 7 + a]
__global__ void k_absmax_words(const int* __restrict__ base, long stride, int na, long n, unsigned* __restrict__ part) {
  __shared__ unsigned sh[kMqThreads];
  for (int f = 0; f < na; ++f) {
    const int* w = base + (long)f * stride;
    unsigned m = 0u;
    for (long i = (long)blockIdx.x * kMqThreads + threadIdx.x; i < n; i += (long)kMqThreads * gridDim.x) {
      const int v = w[i];
      const unsigned a = v < 0 ? 0u - (unsigned)v : (unsigned)v;
      m = a > m ? a : m;
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = kMqThreads / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) sh[threadIdx.x] = sh[threadIdx.x] > sh[threadIdx.x + s] ? sh[threadIdx.x] : sh[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x * 7 + f] = sh[0];
    __syncthreads();
  }
}

struct Scales {
  double s[7];
};

// largest |value| of each metric field in one slot, raw and after the
// unit-determinant normalisation (bssn.cc:61-119): part[block * 7 + f]
__global__ void k_absmax_metric(const int* __restrict__ base, long fstride, long n, Scales sc,
                                double* __restrict__ part) {
  __shared__ double sh[7][kMqThreads];
  double m[7] = {0, 0, 0, 0, 0, 0, 0};
  for (long i = (long)blockIdx.x * kMqThreads + threadIdx.x; i < n; i += (long)kMqThreads * gridDim.x) {
    double g[7];
#pragma unroll
    for (int f = 0; f < 7; ++f) g[f] = (double)base[(long)f * fstride + i] * sc.s[f];
    const double t = norm_factor<double>(g[0], g[1], g[2], g[3], g[4], g[5]);
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      const double gn = (g[f] + (diag6(f) ? t : 0.0)) / (1.0 - t);
      m[f] = fmax(m[f], fmax(fabs(g[f]), fabs(gn)));
    }
    m[6] = fmax(m[6], fabs(g[6]));
  }
  for (int f = 0; f < 7; ++f) sh[f][threadIdx.x] = m[f];
  __syncthreads();
  for (int w = kMqThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int f = 0; f < 7; ++f) sh[f][threadIdx.x] = fmax(sh[f][threadIdx.x], sh[f][threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x < 7) part[blockIdx.x * 7 + threadIdx.x] = sh[threadIdx.x][0];
}

// words of FP64 values at scale 2^e (inv = 2^-e)
__global__ void k_quantize(const double* __restrict__ x, long n, double inv, int* __restrict__ w) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)blockDim.x * gridDim.x)
    w[i] = mq_word(x[i], inv);
}

// FP64 values of words at scale sc
__global__ void k_dequantize(const int* __restrict__ w, long n, double sc, double* __restrict__ x) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)blockDim.x * gridDim.x)
    x[i] = (double)w[i] * sc;
}

// words re-expressed at a scale 2^k times larger (k >= 1), rounded to nearest
__global__ void k_rescale(int* __restrict__ w, long n, double f) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)blockDim.x * gridDim.x)
    w[i] = (int)rint((double)w[i] * f);
}

int* words(cg_ctx* c, int f, int slot) { return (int*)reg_ptr(c, f, slot); }

int ensure_buffers(cg_ctx* c) {
  if (!c->mq_tmp) HIPCHK(c, hipMalloc((void**)&c->mq_tmp, (size_t)c->rstride * sizeof(double)));
  if (!c->mq_part) HIPCHK(c, hipMalloc((void**)&c->mq_part, (size_t)kMqBlocks * 7 * sizeof(double)));
  return CG_OK;
}

// field f re-expressed at exponent e: all four slots, words times
// 2^(old - e) (exact when e is lower, rounded to nearest when higher)
int set_exp(cg_ctx* c, int f, int e) {
  if (e == c->mq_exp[f]) return CG_OK;
  const int k = c->mq_exp[f] - e;
  const double fac = k < -40 ? 0.0 : std::ldexp(1.0, k);
  hipLaunchKernelGGL(k_rescale, dim3(kMqBlocks), dim3(kMqThreads), 0, c->stream, words(c, f, 0), 4 * c->rstride, fac);
  HIPCHK(c, hipGetLastError());
  c->mq_exp[f] = e;
  return CG_OK;
}

// exponent that keeps a largest magnitude m below 2^MQ_BITS words
int exp_for(double m) {
  int k;
  std::frexp(m, &k);  // m < 2^k
  return k - MQ_BITS;
}

int bit_length(unsigned m) { return m ? 32 - __builtin_clz(m) : 0; }

}  // namespace

void mq_scales(const cg_ctx* c, double* sc, double* inv) {
  for (int f = 0; f < 7; ++f) {
    const int e = c->mq ? c->mq_exp[f] : 0;
    sc[f] = std::ldexp(1.0, e);
    inv[f] = std::ldexp(1.0, -e);
  }
}

int mq_store(cg_ctx* c, int f, int slot, const double* x) {
  if (int rc = ensure_buffers(c)) return rc;
  double* part = (double*)c->mq_part;
  hipLaunchKernelGGL(k_absmax_f64, dim3(kMqBlocks), dim3(kMqThreads), 0, c->stream, x, c->npts, part);
  HIPCHK(c, hipGetLastError());
  std::vector<double> h(kMqBlocks);
  HIPCHK(c, hipMemcpyAsync(h.data(), part, kMqBlocks * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double m = 0.0;
  for (double v : h) m = v > m ? v : m;
  if (!std::isfinite(m)) return fail(c, CG_ERR_ARG, "non-finite value for the fixed-point metric");
  // the exponent that fits both the new values and the words the field's
  // other slots keep (lowering it re-expresses those exactly)
  hipLaunchKernelGGL(k_absmax_words, dim3(kMqBlocks), dim3(kMqThreads), 0, c->stream, (const int*)words(c, f, 0),
                     c->rstride, 4, c->npts, c->mq_part);
  HIPCHK(c, hipGetLastError());
  std::vector<unsigned> hw((size_t)kMqBlocks * 7);
  HIPCHK(c, hipMemcpyAsync(hw.data(), c->mq_part, hw.size() * sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  unsigned wo = 0;
  for (int b = 0; b < kMqBlocks; ++b)
    for (int q = 0; q < 4; ++q)
      if (q != slot) wo = hw[(size_t)b * 7 + q] > wo ? hw[(size_t)b * 7 + q] : wo;
  int e = m > 0.0 ? exp_for(m) : c->mq_exp[f];
  if (wo) {
    const int eo = c->mq_exp[f] + bit_length(wo) - MQ_BITS;  // their words back to MQ_BITS bits
    e = eo > e ? eo : e;
  }
  if (int rc = set_exp(c, f, e)) return rc;
  double sc[7], inv[7];
  mq_scales(c, sc, inv);
  hipLaunchKernelGGL(k_quantize, dim3(kMqBlocks), dim3(kMqThreads), 0, c->stream, x, c->npts, inv[f], words(c, f, slot));
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CG_OK;
}

int mq_load(cg_ctx* c, int f, int slot, double* x) {
  double sc[7], inv[7];
  mq_scales(c, sc, inv);
  hipLaunchKernelGGL(k_dequantize, dim3(kMqBlocks), dim3(kMqThreads), 0, c->stream, (const int*)words(c, f, slot),
                     c->npts, sc[f], x);
  HIPCHK(c, hipGetLastError());
  return CG_OK;
}

int mq_check(cg_ctx* c) {
  if (int rc = ensure_buffers(c)) return rc;
  unsigned sat = 0;
  HIPCHK(c, hipMemcpyAsync(&sat, c->mq_sat, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
  Scales sc;
  double inv[7];
  mq_scales(c, sc.s, inv);
  double* part = (double*)c->mq_part;
  hipLaunchKernelGGL(k_absmax_metric, dim3(kMqBlocks), dim3(kMqThreads), 0, c->stream, (const int*)words(c, 0, c->sP),
                     c->fstride, c->npts, sc, part);
  HIPCHK(c, hipGetLastError());
  std::vector<double> h((size_t)kMqBlocks * 7);
  HIPCHK(c, hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (sat) return fail(c, CG_ERR_STATE, "the fixed-point metric saturated: a field grew past its scale between checks");
  // target exponents: the largest value (raw or normalised) below
  // 2^MQ_BITS words, raised as needed; lowered to below 2^(MQ_BITS-1)
  // words when it has fallen below 2^(MQ_BITS-3); then the floor below the
  // largest exponent of the fields that hold data
  int target[7], top = INT32_MIN;
  for (int f = 0; f < 7; ++f) {
    double m = 0.0;
    for (int b = 0; b < kMqBlocks; ++b) m = h[(size_t)b * 7 + f] > m ? h[(size_t)b * 7 + f] : m;
    if (!std::isfinite(m)) return fail(c, CG_ERR_NAN, "non-finite value in the fixed-point metric");
    target[f] = c->mq_exp[f];
    if (m > 0.0) {
      const int need = exp_for(m);
      if (need > target[f]) target[f] = need;
      else if (need + 3 < target[f]) target[f] = need + 1;
      if (target[f] > top) top = target[f];
    }
  }
  for (int f = 0; f < 7; ++f) {
    if (top != INT32_MIN && target[f] < top - MQ_FLOOR) target[f] = top - MQ_FLOOR;
    if (int rc = set_exp(c, f, target[f])) return rc;
  }
  return CG_OK;
}

void mq_free(cg_ctx* c) {
  if (c->mq_tmp) (void)hipFree(c->mq_tmp);
  if (c->mq_part) (void)hipFree(c->mq_part);
  c->mq_tmp = nullptr;
  c->mq_part = nullptr;
}

}  // namespace cg
