// bssn_stage_kernel.h -- fused BSSN RK4 stage kernel for gfx950 (template;
// instantiated by the stage_inst_*.hip translation units).
//
// One launch = one RK4 stage over the whole periodic grid:
//   * a 3D tile of points per workgroup; the tile plus a halo of H points is
//     staged into LDS one stencil field at a time (double-buffered: the
//     global loads of the next field are in flight while the current one is
//     differentiated),
//   * the stencil results are folded into per-point tensors immediately
//     (e.g. gamma^{kl} d_k d_l gamma_ij per metric field), which keeps the
//     live register set near 60 values instead of ~110 raw derivatives,
//   * the BSSN right-hand side (components/bssn/bssn.cc:361-831 and the
//     macros in components/bssn/bssn_macros.h) is evaluated in registers,
//   * the RK4Register update (utils/RK4Register.h:118-176) is the epilogue,
//     so each stage streams every field register once.
// Stage 1 also applies BSSN::set_DIFFgamma_Aij_norm (bssn.cc:61-119) to the
// _p register on the fly: the unit-determinant factor is computed for every
// halo point and applied while staging the metric into LDS.
//
// TS is the storage type, TC the stencil/algebra type.  FP64 drop-in: both
// double.  FP32 split: both float, with the RK4 accumulation in double and
// FRW scalars passed in double.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "bssn_stage.h"
#include "stencil.h"

namespace cg {

__device__ __host__ constexpr int sy(int a, int b) {
  return a == b ? (a == 0 ? 0 : a == 1 ? 3 : 5) : (a + b == 1 ? 1 : a + b == 2 ? 2 : 4);
}
__device__ __host__ constexpr bool diag6(int s) { return s == 0 || s == 3 || s == 5; }

template <typename T>
__device__ __forceinline__ T norm_factor(T g11, T g12, T g13, T g22, T g23, T g33) {
  // bssn.cc:75-88:  1 - det(1 + DIFFgamma), then -(det^(1/3) - 1)
  T omd = -(g11 + g22 + g33 - g12 * g12 - g13 * g13 - g23 * g23 + g11 * g22 + g11 * g33 + g22 * g33 -
            g23 * g23 * g11 - g13 * g13 * g22 - g12 * g12 * g33 + T(2) * g12 * g13 * g23 + g11 * g22 * g33);
  // 1 - (1 - x)^(1/3) for |x| < 2^-10 from its binomial series,
  // x/3 + x^2/9 + 5x^3/81 + 10x^4/243 + 22x^5/729 + 154x^6/6561 + ..., whose
  // first omitted term is below 2^-52 of the sum there (the unit-determinant
  // drift of one step is far smaller); log1p/expm1 otherwise
  if (fabs(omd) < T(0x1p-10))
    return omd * (T(1.0 / 3.0) +
                  omd * (T(1.0 / 9.0) +
                         omd * (T(5.0 / 81.0) + omd * (T(10.0 / 243.0) + omd * (T(22.0 / 729.0) + omd * T(154.0 / 6561.0))))));
  return -expm1(log1p(-omd) / T(3));
}

// element load at a 32-bit element offset from a uniform base pointer, so the
// address is formed as SGPR base + 32-bit VGPR byte offset
template <typename T>
__device__ __forceinline__ T ldo(const T* __restrict__ base, unsigned eoff) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (size_t)(eoff * (unsigned)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ void sto(T* __restrict__ base, unsigned eoff, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + (size_t)(eoff * (unsigned)sizeof(T))) = v;
}

// two consecutive elements from an element-aligned address (one 16-byte /
// 8-byte load; global memory allows element alignment on gfx950)
typedef double d2u __attribute__((ext_vector_type(2))) __attribute__((aligned(8)));
typedef float f2u __attribute__((ext_vector_type(2))) __attribute__((aligned(4)));
__device__ __forceinline__ void ld2(const double* __restrict__ base, unsigned eoff, double& v0, double& v1) {
  const d2u v = *reinterpret_cast<const d2u*>(reinterpret_cast<const char*>(base) + (size_t)(eoff * 8u));
  v0 = v.x;
  v1 = v.y;
}
__device__ __forceinline__ void ld2(const float* __restrict__ base, unsigned eoff, float& v0, float& v1) {
  const f2u v = *reinterpret_cast<const f2u*>(reinterpret_cast<const char*>(base) + (size_t)(eoff * 4u));
  v0 = v.x;
  v1 = v.y;
}


template <typename T>
__device__ __forceinline__ bool nan_as_float(T v) {
  // numNaNs tests the float-rounded value (utils/math.h:1524-1541)
  float f = (float)v;
  unsigned u = __float_as_uint(f);
  return (u << 1) > 0xff000000u;
}

// The conformal metric (DIFFgamma_ij, DIFFphi) is always stored and
// differentiated in FP64: its second derivatives enter the Ricci tensor with
// a cancellation of order 1/(k dx)^2 (~1600 for the longest mode at 256^3),
// so FP32 storage would put its rounding noise straight into A_ij.
__device__ __host__ constexpr bool metric_field(int f) { return f <= G33 || f == PHI; }
// ... except in the all-FP32 mode (MODE_M32), where the metric is stored
// and differentiated in the storage type like every other field
template <bool M32>
__device__ __host__ constexpr bool mfield(int f) { return !M32 && metric_field(f); }

// codes of the dust fluxes in the staging sequence (beyond the field codes):
// flux q*3 + i alone, or the three fluxes of quantity q in one tile
constexpr int FLUX0 = 32;
constexpr int FLUXG0 = 48;
// codes of the two-field tiles of first-derivative fields: (K, Gamma^1),
// (Gamma^2, Gamma^3)
constexpr int KG0 = 60, KG1 = 61;

// order in which fields are staged through LDS:
// PHI, G11..G33, [ALPHA], KK, GAM1..GAM3, [12 dust fluxes], A11..A33 (the
// A_ij only for KO and the constraints)
template <bool LA, bool DU>
__device__ __host__ constexpr int stencil_seq(int s) {
  if (s == 0) return PHI;
  if (s <= 6) return G11 + (s - 1);
  if (LA) {
    if (s == 7) return ALPHA;
    s -= 1;
  }
  if (s == 7) return KK;
  if (s <= 10) return GAM1 + (s - 8);
  if (DU) {
    if (s <= 10 + NFLUX) return FLUX0 + (s - 11);
    s -= NFLUX;
  }
  return A11 + (s - 11 < 5 ? s - 11 : 5);
}

// Staging order of the DMA path in split precision.  Position s is staged in
// LDS buffer s % 3, and buffer 1 is FP32-sized, so every FP64 (metric) field
// sits at a position s % 3 != 1:
// PHI, KK, G11, G12, GAM1, G13, G22, GAM2, G23, G33, GAM3, [ALPHA],
// [12 dust fluxes, or 4 tiles of three], A11..A33 (everything after the head
// is FP32).  KG (every field 4 bytes): K and Gamma^i in two two-field tiles
// at positions 2 and 5 (buffer 2, the larger one), the head
// PHI, G11, (K, GAM1), G12, G13, (GAM2, GAM3), G22, G23, G33.
template <bool LA, bool DU = false, bool GRP = false, bool KG = false>
__device__ __host__ constexpr int stencil_seq_split(int s) {
  constexpr int head[11] = {PHI, KK, G11, G12, GAM1, G13, G22, GAM2, G23, G33, GAM3};
  constexpr int headk[9] = {PHI, G11, KG0, G12, G13, KG1, G22, G23, G33};
  constexpr int NH = KG ? 9 : 11;
  if (s < NH) return KG ? headk[s] : head[s];
  if (LA) {
    if (s == NH) return ALPHA;
    s -= 1;
  }
  if (DU && GRP) {
    if (s < NH + NDUST) return FLUXG0 + (s - NH);
    s -= NDUST;
  } else if (DU) {
    if (s < NH + NFLUX) return FLUX0 + (s - NH);
    s -= NFLUX;
  }
  return A11 + (s - NH < 5 ? s - NH : 5);
}

// LDS DMA of 16 bytes per lane: lane l copies 16 bytes at base + voff into
// LDS byte lds + 16*l (global_load_lds_dwordx4; lds and base wave-uniform,
// 16-byte aligned).  Written as asm so that the compiler does not track these
// loads: its own LDS wait analysis would otherwise wait for every
// outstanding DMA before the first read of any tile.  The caller waits for
// them explicitly (dma_wait_barrier).
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"  // m0 is reserved: set and used inside the statement
__device__ __forceinline__ void dma16(const void* base, unsigned voff, unsigned lds) {
  // base and lds are wave-uniform; readfirstlane puts them in SGPRs where
  // the compiler holds them in VGPRs (it folds away for SGPR values)
  // (readfirstlane returns int: each half goes through unsigned before it
  // is widened, or a low half with bit 31 set would sign-extend into the
  // high half)
  const unsigned long long bb = (unsigned long long)base;
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(bb >> 32));
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)bb);
  const unsigned long long bs = ((unsigned long long)hi << 32) | (unsigned long long)lo;
  const unsigned ls = (unsigned)__builtin_amdgcn_readfirstlane((int)lds);
  asm volatile("s_mov_b32 m0, %2\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, %1" ::"v"(voff), "s"(bs), "s"(ls)
               : "memory", "m0");
}
// The same with a per-lane 64-bit address (lanes may read different arrays).
__device__ __forceinline__ void dma16v(const void* addr, unsigned lds) {
  const unsigned ls = (unsigned)__builtin_amdgcn_readfirstlane((int)lds);
  asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(addr), "s"(ls)
               : "memory", "m0");
}
#pragma clang diagnostic pop

// e / D for 0 <= e < 2^13 as one 24-bit multiply and a shift, full-rate
// instructions where a division by a constant becomes the quarter-rate
// 32-bit high multiply; the reciprocal is checked for every e at compile time
constexpr int ceil_log2(unsigned d) { return d <= 1 ? 0 : 1 + ceil_log2((d + 1) / 2); }
template <unsigned D>
struct SmallDiv {
  static constexpr unsigned S = 13 + ceil_log2(D);
  static constexpr unsigned M = ((1u << S) + D - 1) / D;
  static constexpr bool exact() {
    for (unsigned e = 0; e < (1u << 13); ++e)
      if ((e * M) >> S != e / D) return false;
    return M < (1u << 24);
  }
};
template <unsigned D>
__device__ __forceinline__ unsigned div13(unsigned e) {
  static_assert(SmallDiv<D>::exact(), "24-bit reciprocal exact below 2^13");
  return (unsigned)__umul24(e, SmallDiv<D>::M) >> SmallDiv<D>::S;
}

// Wrapped global element offset of element e (< 2^13) of a box of rows of
// RK elements along k, RJ rows along j, starting at (x0, y0, z0) (at most
// one period outside the grid).  24-bit multiplies: the launcher of the DMA
// kernels keeps nx * ny below 2^24.
template <int RK, int RJ>
__device__ __forceinline__ unsigned box_off(int e, int x0, int y0, int z0, int nx, int ny, int nz) {
  const unsigned r = div13<RK>((unsigned)e), hi = div13<RJ>(r);
  int x = x0 + (int)hi, y = y0 + (int)(r - (unsigned)__umul24(hi, RJ)), z = z0 + (int)((unsigned)e - (unsigned)__umul24(r, RK));
  x += x < 0 ? nx : 0; x -= x >= nx ? nx : 0;
  y += y < 0 ? ny : 0; y -= y >= ny ? ny : 0;
  z += z < 0 ? nz : 0; z -= z >= nz ? nz : 0;
  return (unsigned)__umul24((unsigned)__umul24(x, ny) + y, nz) + (unsigned)z;
}

// End of one step of the DMA staging: this wave's vector memory operations
// older than the newest K have completed (they retire in order), then the
// block barrier.  No fence: the DMA loads of the next fields stay in flight
// across it.
template <int K>
__device__ __forceinline__ void dma_wait_barrier() {
  static_assert(K >= 0 && K < 64, "vmcnt range");
  asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(0)\n\ts_barrier" ::"n"(K) : "memory");
}

// Morton (Z-order) index m -> tile column (x, y) on a 2^lx x 2^ly grid: the
// low 2*min(lx, ly) bits of m interleave x (even bits) and y (odd bits), the
// remaining high bits extend the longer side.
__device__ __forceinline__ void morton_xy(int m, int lx, int ly, int& x, int& y) {
  const int l = lx < ly ? lx : ly;
  x = 0;
  y = 0;
  for (int b = 0; b < l; ++b) {
    x |= ((m >> (2 * b)) & 1) << b;
    y |= ((m >> (2 * b + 1)) & 1) << b;
  }
  const int hi = m >> (2 * l);
  if (lx > ly) x |= hi << l; else y |= hi << l;
}

// WPE: minimum waves per SIMD the register allocation must allow.
// LIVE_ALPHA = false: synchronous gauge with DIFFalpha identically zero (Static
// lapse, zero initial lapse perturbation).  Every lapse derivative is then an
// exact zero, so the lapse stencil, D_iD_j alpha and the lapse registers are
// skipped; the result is bitwise the same as differentiating zeros.
// MODE bit 0: constraint diagnostics (TC = double), which also
// differentiates A_ij for the momentum constraint and writes per-point
// values instead of the RK4 epilogue.  MODE bit 1: the dust fluid of
// components/dust_fluid/dust.cc is evolved with the metric and sources it.
// MODE bit 2 (all-FP32 storage): the conformal metric is one more storage-
// type field; the host points a.regs at field 7 of one 18-field allocation,
// so RS(f) addresses the metric fields f < 7 too.
//
// stage_body is one tile t (< a.ntx*a.nty*a.ntz) of one grid; fv holds the
// FRW values of the stage.  Every thread of the block must call it (it holds
// barriers).
//
// DMA = true: the halo tiles go from memory straight into LDS
// (global_load_lds, 16 bytes per lane), two fields ahead of the one being
// differentiated, in three LDS buffers.  No registers hold the loads in
// flight.  Requires nz % 4 == 0 and no stage-1 normalisation (a.normalize is
// ignored); the launcher falls back to DMA = false otherwise.
template <typename TS, typename TC, int R, int H, int TI, int TJ, int TK, bool LIVE_ALPHA, int MODE, bool DMA = false>
__device__ __forceinline__ void stage_body(const StageArgs& a, const int t, const FrwV fv) {
  constexpr int NT = TI * TJ * TK;
  // Halo along k: H; with DMA staging rounded up so that every 16-byte group
  // of a halo row starts at a multiple of its element count in z and stays
  // contiguous across the periodic wrap: to even for FP64 pairs (KH), to a
  // multiple of 4 for FP32 quads (KH4, split precision).  The two tile
  // geometries differ only along k.
  constexpr bool SPLIT = sizeof(TS) == 4;
  constexpr int KH = DMA ? (H + 1) / 2 * 2 : H;
  constexpr int KH4 = DMA && SPLIT ? (H + 3) / 4 * 4 : KH;
  constexpr int HJ = TJ + 2 * H, HK = TK + 2 * KH, HI = TI + 2 * H, HK4 = TK + 2 * KH4;
  constexpr int HV = HI * HJ * HK, HV4 = HI * HJ * HK4;
  constexpr int SI = HJ * HK, SJ = HK, SK = 1;
  constexpr int SI4 = HJ * HK4, SJ4 = HK4;  // FP32 tiles of the DMA path

  // DMA path geometry.  Halo tiles move in 16-byte groups (FP64 pairs, FP32
  // quads), 64 lanes = one 1 KB chunk per instruction; the three staging
  // buffers are padded to whole chunks.  After the stencil phase the same
  // LDS holds each wave's prefetch of the epilogue's _p and _f values: one
  // 1 KB DMA per FP64-stored field (lanes 0-31 _p, 32-63 _f of the wave's 64
  // points) and one per two FP32 fields.
  constexpr int NW = NT / 64;                                   // waves per block
  constexpr int NPR = HV / 2, NQD = HV4 / 4;                    // FP64 pairs, FP32 quads per tile
  constexpr int CPR = (NPR + 63) / 64, CQD = (NQD + 63) / 64;   // chunks per field
  constexpr int P64 = CPR * 1024, P32 = SPLIT ? CQD * 1024 : P64;  // staging buffer bytes
  constexpr bool M32 = (MODE & MODE_M32) != 0;                  // metric in the storage type
  constexpr bool MQ = M32 && (MODE & MODE_MQ) != 0;             // ... as 32-bit fixed point
  // metric tiles hold the fixed-point words (FP32 compute); with FP64
  // compute (constraint mode) they hold the scaled FP64 values instead
  constexpr bool MQW = MQ && sizeof(TC) == 4;
  static_assert(!MQ || sizeof(TS) == 4, "the fixed-point metric shares the 4-byte storage");
  // epilogue prefetch: the stage kernels, and with every field in 4 bytes
  // the dust ones too (which then load the fluid's point values first)
  constexpr bool PF = DMA && ((MODE & ~(MODE_M32 | MODE_MQ)) == 0 ||
                              (M32 && (MODE & ~(MODE_M32 | MODE_MQ | MODE_DUST)) == 0));
  constexpr bool PFD = PF && (MODE & MODE_DUST) != 0;
  constexpr int F32 = M32 ? 0 : 7;                              // first field of the FP32 prefetch chunks
  constexpr int N64 = M32 ? 0 : SPLIT ? 7 : (LIVE_ALPHA ? NF : NF - 1);  // FP64-stored fields updated (ALPHA last)
  constexpr int N32 = M32 ? (LIVE_ALPHA ? NF : NF - 1) : SPLIT ? (LIVE_ALPHA ? NF - 7 : NF - 8) : 0;  // FP32 ones
  // every field 4 bytes, frozen lapse: stage 4 of the accumulator-free form
  // also prefetches a2 of the first NA2 fields, in the groups (16 lanes, 64
  // points, one field) of the staging LDS left after the _p/_f groups (a
  // 52 KB layout with every a2, issued from one loop with _p/_f, measured
  // stage 4 0.21 ms faster but stages 1-3 0.12 ms slower at 256^3)
  constexpr int LSTAGE = 2 * P64 + P32;
  constexpr bool PF2 = PF && M32 && !LIVE_ALPHA;
  constexpr int CPF = (2 * N32 + 3) / 4;                        // _p/_f chunks of the FP32 fields
  constexpr int NA2_ = PF2 ? (LSTAGE / (NW * 1024) - N64 - CPF) * 4 : 0;
  constexpr int NA2 = NA2_ < 0 ? 0 : NA2_ > N32 ? N32 : NA2_;   // a2 fields, from chunk N64 + CPF on
  constexpr int A2F = N32 - NA2;  // the last NA2 fields (the metric's first ones read a3 from memory anyway)
  constexpr int PFR = (N64 + CPF + (NA2 + 3) / 4) * 1024;       // prefetch bytes per wave
  constexpr int LPF = PF ? NW * PFR : 0;
  __shared__ __align__(16) char ldma[DMA ? (LSTAGE > LPF ? LSTAGE : LPF) : 16];

  // Register path: the halo tile is loaded in pairs of k-neighbours (HK is even, so a pair
  // never straddles a row): one 16-byte (FP64) or 8-byte (FP32) load per
  // pair, or two element loads where the row wraps around the periodic k
  // boundary between the two.
  static_assert(HK % 2 == 0, "halo rows hold whole pairs");
  constexpr int HP = HV / 2;
  constexpr int NPP = (HP + NT - 1) / NT;
  __shared__ __align__(16) double ldsraw[2][HV];  // double-buffered field tile (FP32 fields use the first half)
  __shared__ double ldn[HV];                      // stage-1 unit-determinant factor t at every halo point

  // Tile order: the ~100 tiles one XCD has in flight hold each tile's
  // neighbours along the two faster axes, so their halos are re-read from L2
  // and only the halo along the slowest axis comes back from memory.  The
  // slowest axis should be the one with the smallest halo share (the
  // autotuner picks): 0 = j slowest (k, i, j), 1 = k slowest (i, j, k),
  // 2 = i slowest (k, j, i).  3 = k fastest, then the (i, j) tile columns in
  // Morton (Z) order (tile counts along i and j powers of two): a column's
  // i and j neighbours were visited shortly before or after it at every
  // scale, so both its i and j halos tend to come from L2.
  int tx, ty, tz;
  const bool pow2 = !(a.ntx & (a.ntx - 1)) && !(a.nty & (a.nty - 1));
  if (a.tile_order == 0 || (a.tile_order == 3 && !pow2)) {
    tz = t % a.ntz; tx = (t / a.ntz) % a.ntx; ty = t / (a.ntz * a.ntx);
  } else if (a.tile_order == 1) {
    tx = t % a.ntx; ty = (t / a.ntx) % a.nty; tz = t / (a.ntx * a.nty);
  } else if (a.tile_order == 2) {
    tz = t % a.ntz; ty = (t / a.ntz) % a.nty; tx = t / (a.ntz * a.nty);
  } else if (a.tile_order >= 4) {
    // strips of W = 2 (order 4) or 4 (order 5) tile rows along j: k fastest,
    // then j inside the strip, then i, then the strips.  The j halos inside
    // a strip come from tiles in flight; only the strip edges re-read theirs
    const int W = a.tile_order == 4 ? 2 : 4;
    tz = t % a.ntz;
    const int r = t / a.ntz, s = r / (W * a.ntx), w = r - s * W * a.ntx;
    const int wc = a.nty - s * W < W ? a.nty - s * W : W;
    ty = s * W + w % wc; tx = w / wc;
  } else {
    tz = t % a.ntz;
    morton_xy(t / a.ntz, __builtin_ctz((unsigned)a.ntx), __builtin_ctz((unsigned)a.nty), tx, ty);
  }
  const int i0 = tx * TI, j0 = ty * TJ, k0 = tz * TK;

  const int tid = threadIdx.x;
  __builtin_assume(tid >= 0 && tid < NT);
  const int lk = tid % TK, lj = (tid / TK) % TJ, li = tid / (TK * TJ);
  const int px = i0 + li, py = j0 + lj, pz = k0 + lk;
  const bool valid = (px < a.nx) && (py < a.ny) && (pz < a.nz);
  const unsigned idx = valid ? (unsigned)((px * a.ny + py) * a.nz + pz) : 0u;
  const int c = ((li + H) * HJ + (lj + H)) * HK + (lk + KH);
  const int c4 = ((li + H) * HJ + (lj + H)) * HK4 + (lk + KH4);  // in the FP32 tiles

  // wrapped global offset of the first element of each of this thread's
  // halo pairs (pair p = tid + q*NT, elements 2p and 2p+1); the same share is
  // used for every field, so per-element data can stay in registers.  Bit q
  // of `split` marks a pair whose first element is the last of its row
  // (z = nz-1) and whose second is therefore at z = 0.
  unsigned goff[NPP];
  unsigned split = 0u;
  if constexpr (!DMA) {
#pragma unroll
  for (int q = 0; q < NPP; ++q) {
    const int pp = tid + q * NT;
    const int e = 2 * (pp < HP ? pp : 0);
    const int hk = e % HK, hj = (e / HK) % HJ, hi = e / (HK * HJ);
    int x = i0 + hi - H, y = j0 + hj - H, z = k0 + hk - H;
    if (z == -1 || z == a.nz - 1) split |= 1u << q;
    x += x < 0 ? a.nx : 0; x -= x >= a.nx ? a.nx : 0;
    y += y < 0 ? a.ny : 0; y -= y >= a.ny ? a.ny : 0;
    z += z < 0 ? a.nz : 0; z -= z >= a.nz ? a.nz : 0;
    goff[q] = (unsigned)((x * a.ny + y) * a.nz + z);
  }
  }
  const unsigned wrapk = (unsigned)(a.nz - 1);
  // both elements of pair q of a field
  auto pair = [&](const auto* __restrict__ src, int q, auto& v0, auto& v1) {
    if (!((split >> q) & 1u)) {
      ld2(src, goff[q], v0, v1);
    } else {
      v0 = ldo(src, goff[q]);
      v1 = ldo(src, goff[q] - wrapk);
    }
  };

  auto RM = [&](int f, int slot) -> double* { return a.met + ((long)f * 4 + slot) * a.rstride; };
  auto RS = [&](int f, int slot) -> TS* { return (TS*)a.regs + ((long)(f - 7) * 4 + slot) * a.rstride; };
  // stencil / point input of field f: DIFFgamma_ij and A_ij come from the
  // precomputed normalised copy when there is one
  auto INM = [&](int f) -> const double* { return RM(f, (f <= G33 && a.s_in_norm >= 0) ? a.s_in_norm : a.s_in); };
  auto INS = [&](int f) -> const TS* {
    return f >= FLUX0 ? (const TS*)a.flux + (long)(f - FLUX0) * a.rstride
                      : RS(f, (((f >= A11 && f <= A33) || (M32 && f <= G33)) && a.s_in_norm >= 0) ? a.s_in_norm
                                                                                                 : a.s_in);
  };
  auto RD = [&](int q, int slot) -> TS* { return (TS*)a.dreg + ((long)q * 4 + slot) * a.rstride; };
  constexpr bool CONS = (MODE & MODE_CONS) != 0;
  constexpr bool DUST = (MODE & MODE_DUST) != 0;
  const bool norm = !DMA && a.normalize != 0;

  // stage-1 unit-determinant factor at the halo points (bssn.cc:75-99):
  // gamma_ii -> (gamma_ii + t)/(1 - t), gamma_ij -> gamma_ij/(1 - t).  Each
  // thread writes and later reads only its own elements, so no barrier.
  if (norm) {
#pragma unroll
    for (int q = 0; q < NPP; ++q) {
      const int pp = tid + q * NT;
      if (pp < HP) {
        double g[6][2];
        if constexpr (M32) {
#pragma unroll
          for (int m = 0; m < 6; ++m) {
            TS g0, g1;
            pair(INS(G11 + m), q, g0, g1);
            g[m][0] = MQ ? (double)__float_as_int(g0) * a.msc[G11 + m] : (double)g0;
            g[m][1] = MQ ? (double)__float_as_int(g1) * a.msc[G11 + m] : (double)g1;
          }
        } else {
#pragma unroll
          for (int m = 0; m < 6; ++m) pair(INM(G11 + m), q, g[m][0], g[m][1]);
        }
#pragma unroll
        for (int h = 0; h < 2; ++h)
          ldn[2 * pp + h] = norm_factor<double>(g[0][h], g[1][h], g[2][h], g[3][h], g[4][h], g[5][h]);
      }
    }
  }

  double pd[NPP][2];
  TS pf[NPP][2];
  auto fetch = [&](int f) {
    if (mfield<M32>(f)) {
      const double* __restrict__ src = INM(f);
#pragma unroll
      for (int q = 0; q < NPP; ++q)
        if (tid + q * NT < HP) pair(src, q, pd[q][0], pd[q][1]);
    } else {
      const TS* __restrict__ src = INS(f);
#pragma unroll
      for (int q = 0; q < NPP; ++q)
        if (tid + q * NT < HP) pair(src, q, pf[q][0], pf[q][1]);
    }
  };
  auto store = [&](int buf, int f) {
    if (mfield<M32>(f)) {
      double* __restrict__ dst = ldsraw[buf];
#pragma unroll
      for (int q = 0; q < NPP; ++q) {
        const int pp = tid + q * NT;
        if (pp < HP) {
          double v0 = pd[q][0], v1 = pd[q][1];
          if (norm && f <= G33) {
            const double t0 = ldn[2 * pp], t1 = ldn[2 * pp + 1];
            v0 = (v0 + (diag6(f) ? t0 : 0.0)) / (1.0 - t0);
            v1 = (v1 + (diag6(f) ? t1 : 0.0)) / (1.0 - t1);
          }
          *reinterpret_cast<d2u*>(dst + 2 * pp) = d2u{v0, v1};
        }
      }
    } else {
      TC* __restrict__ dst = reinterpret_cast<TC*>(ldsraw[buf]);
#pragma unroll
      for (int q = 0; q < NPP; ++q) {
        const int pp = tid + q * NT;
        if (pp < HP) {
          if (MQ && !MQW && f < 7) {
            // FP64 compute: the scaled values (normalised at stage 1)
            const double sc = a.msc[f < 7 ? f : 0];
            double w0 = (double)__float_as_int(pf[q][0]) * sc, w1 = (double)__float_as_int(pf[q][1]) * sc;
            if (norm && f <= G33) {
              const double t0 = ldn[2 * pp], t1 = ldn[2 * pp + 1];
              w0 = (w0 + (diag6(f) ? t0 : 0.0)) / (1.0 - t0);
              w1 = (w1 + (diag6(f) ? t1 : 0.0)) / (1.0 - t1);
            }
            dst[2 * pp] = (TC)w0;
            dst[2 * pp + 1] = (TC)w1;
          } else if (MQW && norm && f <= G33) {
            // the normalised halo value, back in the field's fixed point
            const double t0 = ldn[2 * pp], t1 = ldn[2 * pp + 1], sc = a.msc[f < 7 ? f : 0], iv = a.minv[f < 7 ? f : 0];
            const double w0 = ((double)__float_as_int(pf[q][0]) * sc + (diag6(f) ? t0 : 0.0)) / (1.0 - t0);
            const double w1 = ((double)__float_as_int(pf[q][1]) * sc + (diag6(f) ? t1 : 0.0)) / (1.0 - t1);
            if (fmax(fabs(w0), fabs(w1)) * iv > MQ_WMAX) atomicAdd(a.mq_sat, 1u);
            dst[2 * pp] = __int_as_float(mq_word(w0, iv));
            dst[2 * pp + 1] = __int_as_float(mq_word(w1, iv));
          } else if (M32 && norm && f <= G33) {
            const double t0 = ldn[2 * pp], t1 = ldn[2 * pp + 1];
            dst[2 * pp] = (TC)(((double)pf[q][0] + (diag6(f) ? t0 : 0.0)) / (1.0 - t0));
            dst[2 * pp + 1] = (TC)(((double)pf[q][1] + (diag6(f) ? t1 : 0.0)) / (1.0 - t1));
          } else {
            dst[2 * pp] = (TC)pf[q][0];
            dst[2 * pp + 1] = (TC)pf[q][1];
          }
        }
      }
    }
  };

  // ---------------------------------------------------------- point values
  TC v[NF];
  double gc[7];  // metric at the point in FP64 (DIFFgamma_ij, DIFFphi)
#pragma unroll
  for (int f = 0; f < 6; ++f)  // idx = 0 outside the grid
    gc[f] = MQ    ? (double)ldo((const int*)INS(f), idx) * a.msc[f]
            : M32 ? (double)ldo(INS(f), idx)
                  : ldo(INM(f), idx);
  const double phi_c = MQ    ? (double)ldo((const int*)INS(PHI), idx) * a.msc[PHI]
                       : M32 ? (double)ldo(INS(PHI), idx)
                             : ldo(INM(PHI), idx);
  gc[6] = phi_c;
#pragma unroll
  for (int f = 0; f < NF; ++f)
    if (!metric_field(f)) v[f] = (f == ALPHA && !LIVE_ALPHA) ? TC(0) : (TC)ldo(INS(f), idx);
  if (norm) {
    // bssn.cc:94-117 at this point
    const double tn = norm_factor<double>(gc[0], gc[1], gc[2], gc[3], gc[4], gc[5]);
    const double rden = 1.0 / (1.0 - tn);
#pragma unroll
    for (int s = 0; s < 6; ++s) gc[s] = (gc[s] + (diag6(s) ? tn : 0.0)) * rden;
    const double g11 = gc[0], g12 = gc[1], g13 = gc[2], g22 = gc[3], g23 = gc[4], g33 = gc[5];
    const double i11 = 1.0 + g22 + g33 - g23 * g23 + g22 * g33;
    const double i22 = 1.0 + g11 + g33 - g13 * g13 + g11 * g33;
    const double i33 = 1.0 + g11 + g22 - g12 * g12 + g11 * g22;
    const double i12 = g13 * g23 - g12 * (1.0 + g33);
    const double i13 = g12 * g23 - g13 * (1.0 + g22);
    const double i23 = g12 * g13 - g23 * (1.0 + g11);
    const double trA = i11 * v[A11] + i22 * v[A22] + i33 * v[A33] +
                       2.0 * (i12 * v[A12] + i13 * v[A13] + i23 * v[A23]);
#pragma unroll
    for (int s = 0; s < 6; ++s)
      v[A11 + s] = (TC)(((double)v[A11 + s] - (1.0 / 3.0) * ((diag6(s) ? 1.0 : 0.0) + gc[s]) * trA) * rden);
  }
#pragma unroll
  for (int f = 0; f < 6; ++f) v[f] = (TC)gc[f];
  v[PHI] = (TC)phi_c;

  // static dust density: the DMA kernels load it with the point values, so
  // that waiting for it later does not also wait for the epilogue prefetch
  // (vector memory operations retire in order)
  TC Dpt = TC(0);
  if constexpr (DMA) {
    if (a.matter == MATTER_STATIC) Dpt = (TC)ldo((const TS*)a.DIFFD, idx);
  }
  // dust fluid with the epilogue prefetch: its stage input, RK4 inputs and
  // lagged state at the point, loaded after the staging and waited for
  // before the prefetch is issued (vector memory operations retire in order:
  // waiting for them in the algebra would also wait for the prefetch)
  static_assert(!PFD || MQ, "the prefetching dust kernels hold the metric as fixed-point words");
  constexpr int ND = PFD ? NDUST : 1;
  TS pdi[ND], pdp[ND], pd1[ND], pd2[ND], pdl[ND];
  int plw[PFD ? 7 : 1];
  auto dust_preload = [&]() {
    if constexpr (PFD) {
      const bool rk = a.stage >= 1, f1 = a.incr ? a.stage == 4 : a.stage >= 2, f2 = a.incr && a.stage == 4;
      const bool lag = a.lag_mode != 0;
#pragma unroll
      for (int q = 0; q < NDUST; ++q) {
        pdi[q] = ldo((const TS*)RD(q, a.s_in), idx);
        pdp[q] = rk ? ldo((const TS*)RD(q, a.s_p), idx) : TS(0);
        pd1[q] = f1 ? ldo((const TS*)RD(q, a.s_f), idx) : TS(0);
        pd2[q] = f2 ? ldo((const TS*)RD(q, a.s_a2), idx) : TS(0);
        pdl[q] = lag ? ldo((const TS*)RD(q, a.s_lag), idx) : TS(0);
      }
#pragma unroll
      for (int q = 0; q < 6; ++q) plw[q] = lag ? ldo((const int*)RS(q, a.s_lag_met), idx) : 0;
      plw[6] = lag ? ldo((const int*)RS(PHI, a.s_lag), idx) : 0;
#pragma unroll
      for (int q = 0; q < NDUST; ++q) asm volatile("" ::"v"(pdi[q]), "v"(pdp[q]), "v"(pd1[q]), "v"(pd2[q]), "v"(pdl[q]));
#pragma unroll
      for (int q = 0; q < 7; ++q) asm volatile("" ::"v"(plw[q]));
    }
  };

  // inverse conformal metric from DIFF variables (bssn.cc:447-455)
  TC gi[6];
  {
    const TC g11 = v[G11], g12 = v[G12], g13 = v[G13], g22 = v[G22], g23 = v[G23], g33 = v[G33];
    gi[0] = TC(1) + g22 + g33 - g23 * g23 + g22 * g33;
    gi[3] = TC(1) + g11 + g33 - g13 * g13 + g11 * g33;
    gi[5] = TC(1) + g11 + g22 - g12 * g12 + g11 * g22;
    gi[1] = g13 * g23 - g12 * (TC(1) + g33);
    gi[2] = g12 * g23 - g13 * (TC(1) + g22);
    gi[4] = g12 * g13 - g23 * (TC(1) + g11);
  }
  auto GI = [&](int p, int q) -> TC { return gi[sy(p, q)]; };
  auto GM = [&](int p, int q) -> TC { return v[G11 + sy(p, q)] + (p == q ? TC(1) : TC(0)); };
  auto AM = [&](int p, int q) -> TC { return v[A11 + sy(p, q)]; };

  // ------------------------------------------------ stencil phase (LDS)
  const double i1 = a.inv_dx, i2 = a.inv_dx2;
  const TC idx1 = (TC)a.inv_dx;
  TC dg[3][6];      // d_k gamma_ij
  TC T1[6];         // gamma^{kl} d_k d_l gamma_ij
  TC dphi[3], ddphi[6], dal[3], ddal[6], dK[3], dG[3][3];
  TC dA[3][6];      // d_j A_kl, constraint mode only
  TC ddust[NDUST];  // sum_i d_i of the dust fluxes
#pragma unroll
  for (int q = 0; q < NDUST; ++q) ddust[q] = TC(0);
  TC ko[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) ko[f] = TC(0);
  const bool use_ko = a.use_ko != 0;
  constexpr int KO_ORDER = (H > R) ? 2 * R : 0;

  // stencil fields in order; A_ij only when KO is active or for the
  // constraints; the dust fluxes only in the dust stage
  constexpr bool FLUXES = DUST && !CONS;
  // 4-byte DMA staging: the three fluxes of a quantity share one LDS tile,
  // each with the halo of its own derivative axis only (x part: HI x TJ x
  // TK, y part: TI x HJ x TK, z part: TI x TJ x HK4)
  constexpr bool FGRP = DMA && FLUXES && SPLIT;
  constexpr int FGX = HI * TJ * TK, FGY = TI * HJ * TK, FGZ = TI * TJ * HK4;
  // every field 4 bytes, no KO: K and the Gamma^i, which only have first
  // derivatives, go two to a tile, each field in the same three parts, where
  // such a tile fits the larger staging buffers (the halo of 3 of order 6)
  constexpr bool KGRP = DMA && M32 && KO_ORDER == 0 && !CONS &&
                        (2 * (FGX + FGY + FGZ) / 4 + 63) / 64 <= (HV / 2 + 63) / 64;
  constexpr int NSEQ = (LIVE_ALPHA ? 18 : 17) - (KGRP ? 2 : 0) + (FLUXES ? (FGRP ? NDUST : NFLUX) : 0);
  const int nseq = (use_ko || CONS) ? NSEQ : NSEQ - 6;
  if constexpr (!LIVE_ALPHA) {
#pragma unroll
    for (int d = 0; d < 3; ++d) dal[d] = TC(0);
#pragma unroll
    for (int d = 0; d < 6; ++d) ddal[d] = TC(0);
  }

  // derivatives of field f, staged in the LDS tile BUF, at this thread's
  // point, folded into the per-point tensors (a macro: the same code as a
  // lambda allocates registers differently and spills)
#define CG_DERIVE(f, BUF) \
    if (mfield<M32>(f)) {                                                                                               \
      const double* __restrict__ L = reinterpret_cast<const double*>(BUF);                                               \
      if (KO_ORDER && use_ko) ko[f] = (TC)(a.ko_pref * ko_sum<KO_ORDER, double>(L, c, SI, SJ, SK));                      \
      const double d0 = fd1<R>(L, c, SI) * i1, d1 = fd1<R>(L, c, SJ) * i1, d2 = fd1<R>(L, c, SK) * i1;                   \
      const double e00 = fd2<R>(L, c, SI) * i2, e11 = fd2<R>(L, c, SJ) * i2, e22 = fd2<R>(L, c, SK) * i2;                \
      const double e01 = fdm<R>(L, c, SI, SJ) * i2, e02 = fdm<R>(L, c, SI, SK) * i2,                                     \
                   e12 = fdm<R>(L, c, SJ, SK) * i2;                                                                      \
      if (f == PHI) {                                                                                                    \
        dphi[0] = (TC)d0; dphi[1] = (TC)d1; dphi[2] = (TC)d2;                                                            \
        ddphi[0] = (TC)e00; ddphi[1] = (TC)e01; ddphi[2] = (TC)e02;                                                      \
        ddphi[3] = (TC)e11; ddphi[4] = (TC)e12; ddphi[5] = (TC)e22;                                                      \
      } else {                                                                                                           \
        dg[0][f] = (TC)d0; dg[1][f] = (TC)d1; dg[2][f] = (TC)d2;                                                         \
        T1[f] = gi[0] * (TC)e00 + gi[3] * (TC)e11 + gi[5] * (TC)e22 +                                                    \
                TC(2) * (gi[1] * (TC)e01 + gi[2] * (TC)e02 + gi[4] * (TC)e12);                                           \
      }                                                                                                                  \
    } else {                                                                                                             \
      const TC* __restrict__ L = reinterpret_cast<const TC*>(BUF);                                                       \
      const int* __restrict__ LQ = reinterpret_cast<const int*>(BUF);                                                    \
      if (KO_ORDER && use_ko && f < NF) {                                                                                \
        if (MQW && (f <= G33 || f == PHI))                                                                               \
          ko[f] = (TC)(a.ko_pref * a.msc[f < 7 ? f : 0] * ko_sumq<KO_ORDER>(LQ, c4, SI4, SJ4, SK));                      \
        else                                                                                                             \
          ko[f] = (TC)a.ko_pref * ko_sum<KO_ORDER, TC>(L, c4, SI4, SJ4, SK);                                             \
      }                                                                                                                  \
      if (MQW && (f <= G33 || f == PHI)) {                                                                               \
        /* exact integer differences of the words, FP32 sums, the field's scale */                                      \
        const TC s1 = a.msd1[f < 7 ? f : 0], s2 = a.msd2[f < 7 ? f : 0];                                                \
        const TC d0 = fd1qf<R>(LQ, c4, SI4) * s1, d1 = fd1qf<R>(LQ, c4, SJ4) * s1, d2 = fd1qf<R>(LQ, c4, SK) * s1;       \
        const TC e00 = fd2qf<R>(LQ, c4, SI4) * s2, e11 = fd2qf<R>(LQ, c4, SJ4) * s2, e22 = fd2qf<R>(LQ, c4, SK) * s2;    \
        const TC e01 = fdmqf<R>(LQ, c4, SI4, SJ4) * s2, e02 = fdmqf<R>(LQ, c4, SI4, SK) * s2,                           \
                 e12 = fdmqf<R>(LQ, c4, SJ4, SK) * s2;                                                                   \
        if (f == PHI) {                                                                                                  \
          dphi[0] = d0; dphi[1] = d1; dphi[2] = d2;                                                                      \
          ddphi[0] = e00; ddphi[1] = e01; ddphi[2] = e02; ddphi[3] = e11; ddphi[4] = e12; ddphi[5] = e22;                \
        } else {                                                                                                         \
          dg[0][f] = d0; dg[1][f] = d1; dg[2][f] = d2;                                                                   \
          T1[f] = gi[0] * e00 + gi[3] * e11 + gi[5] * e22 + TC(2) * (gi[1] * e01 + gi[2] * e02 + gi[4] * e12);          \
        }                                                                                                                \
      } else if (M32 && (f <= G33 || f == PHI)) {                                                                        \
        const TC idx2 = (TC)a.inv_dx2;                                                                                   \
        const TC d0 = fd1<R>(L, c4, SI4) * idx1, d1 = fd1<R>(L, c4, SJ4) * idx1, d2 = fd1<R>(L, c4, SK) * idx1;          \
        const TC e00 = fd2<R>(L, c4, SI4) * idx2, e11 = fd2<R>(L, c4, SJ4) * idx2, e22 = fd2<R>(L, c4, SK) * idx2;       \
        const TC e01 = fdm<R>(L, c4, SI4, SJ4) * idx2, e02 = fdm<R>(L, c4, SI4, SK) * idx2,                              \
                 e12 = fdm<R>(L, c4, SJ4, SK) * idx2;                                                                    \
        if (f == PHI) {                                                                                                  \
          dphi[0] = d0; dphi[1] = d1; dphi[2] = d2;                                                                      \
          ddphi[0] = e00; ddphi[1] = e01; ddphi[2] = e02; ddphi[3] = e11; ddphi[4] = e12; ddphi[5] = e22;                \
        } else {                                                                                                         \
          dg[0][f] = d0; dg[1][f] = d1; dg[2][f] = d2;                                                                   \
          T1[f] = gi[0] * e00 + gi[3] * e11 + gi[5] * e22 + TC(2) * (gi[1] * e01 + gi[2] * e02 + gi[4] * e12);          \
        }                                                                                                                \
      } else if (f == ALPHA) {                                                                                           \
        const TC idx2 = (TC)a.inv_dx2;                                                                                   \
        dal[0] = fd1<R>(L, c4, SI4) * idx1; dal[1] = fd1<R>(L, c4, SJ4) * idx1; dal[2] = fd1<R>(L, c4, SK) * idx1;       \
        ddal[0] = fd2<R>(L, c4, SI4) * idx2; ddal[3] = fd2<R>(L, c4, SJ4) * idx2; ddal[5] = fd2<R>(L, c4, SK) * idx2;    \
        ddal[1] = fdm<R>(L, c4, SI4, SJ4) * idx2; ddal[2] = fdm<R>(L, c4, SI4, SK) * idx2;                               \
        ddal[4] = fdm<R>(L, c4, SJ4, SK) * idx2;                                                                         \
      } else if (f == KK) {                                                                                              \
        dK[0] = fd1<R>(L, c4, SI4) * idx1; dK[1] = fd1<R>(L, c4, SJ4) * idx1; dK[2] = fd1<R>(L, c4, SK) * idx1;          \
      } else if (f >= GAM1 && f <= GAM3) {                                                                               \
        const int kk = f - GAM1;                                                                                         \
        dG[0][kk] = fd1<R>(L, c4, SI4) * idx1; dG[1][kk] = fd1<R>(L, c4, SJ4) * idx1;                                    \
        dG[2][kk] = fd1<R>(L, c4, SK) * idx1;                                                                            \
      } else if (CONS && f >= A11 && f <= A33) {                                                                         \
        const int q = f - A11;                                                                                           \
        dA[0][q] = fd1<R>(L, c4, SI4) * idx1; dA[1][q] = fd1<R>(L, c4, SJ4) * idx1; dA[2][q] = fd1<R>(L, c4, SK) * idx1; \
      } else if (KGRP && f >= KG0) {                                                                                     \
        /* two first-derivative fields in three parts each; the same points as the full tile */                         \
        const int cx = ((li + H) * TJ + lj) * TK + lk, cy = FGX + (li * HJ + lj + H) * TK + lk;                          \
        const int cz = FGX + FGY + (li * TJ + lj) * HK4 + lk + KH4;                                                      \
        TC dd[2][3];                                                                                                     \
        _Pragma("unroll") for (int h = 0; h < 2; ++h) {                                                                  \
          dd[h][0] = fd1<R>(L, h * (FGX + FGY + FGZ) + cx, TJ * TK) * idx1;                                              \
          dd[h][1] = fd1<R>(L, h * (FGX + FGY + FGZ) + cy, TK) * idx1;                                                   \
          dd[h][2] = fd1<R>(L, h * (FGX + FGY + FGZ) + cz, 1) * idx1;                                                    \
        }                                                                                                                \
        if (f == KG0) {                                                                                                  \
          _Pragma("unroll") for (int d = 0; d < 3; ++d) { dK[d] = dd[0][d]; dG[d][0] = dd[1][d]; }                      \
        } else {                                                                                                         \
          _Pragma("unroll") for (int d = 0; d < 3; ++d) { dG[d][1] = dd[0][d]; dG[d][2] = dd[1][d]; }                   \
        }                                                                                                                \
      } else if (FGRP && f >= FLUXG0) {                                                                                  \
        /* the three fluxes of quantity q, each along its own axis, summed in the order x, y, z */                       \
        const int q = f - FLUXG0;                                                                                        \
        ddust[q] += fd1<R>(L, ((li + H) * TJ + lj) * TK + lk, TJ * TK) * idx1;                                           \
        ddust[q] += fd1<R>(L, FGX + (li * HJ + lj + H) * TK + lk, TK) * idx1;                                            \
        ddust[q] += fd1<R>(L, FGX + FGY + (li * TJ + lj) * HK4 + lk + KH4, 1) * idx1;                                    \
      } else if (FLUXES && f >= FLUX0) {                                                                                 \
        /* dust.cc:203-224: flux q*3 + i is differentiated along i only */                                               \
        const int q = (f - FLUX0) / 3, dir = (f - FLUX0) % 3;                                                            \
        ddust[q] += fd1<R>(L, c4, dir == 0 ? SI4 : dir == 1 ? SJ4 : SK) * idx1;                                          \
      }                                                                                                                  \
    }

  if constexpr (DMA) {
    static_assert(std::is_same<TS, TC>::value, "DMA staging differentiates the stored type");
    static_assert(NT % 64 == 0 && HK % 2 == 0 && HK4 % 4 == 0 && TK % 4 == 0,
                  "whole waves, whole 16-byte groups per row");
    constexpr int MPR = (CPR + NW - 1) / NW, MQD = (CQD + NW - 1) / NW;  // chunks per wave
    // Three field tiles; in split precision buffer 1 only ever holds FP32
    // fields.  Every wave issues the same number of chunks with all lanes
    // active: surplus chunks repeat the last chunk (the same bytes to the
    // same place), surplus lanes of the last chunk load element 0 into the
    // padding.
    char* const lb0 = ldma;
    char* const lb1 = ldma + P64;
    char* const lb2 = ldma + P64 + P32;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6), ln = tid & 63;
    // wrapped global element offset of halo element e of a tile with rows of
    // hk_n elements starting kh before the tile (e % 4 == 0 for quads, e % 2
    // == 0 for pairs: the group never leaves its row or wraps inside)
    static_assert(HV < (1 << 13) && 2 * (FGX + FGY + FGZ) < (1 << 13), "tile elements below 2^13");
    // chunk m of this wave is chunk min(wv + m*NW, last) of the field: lane l
    // takes group 64*chunk + l, which lands at LDS byte 16*(64*chunk + l)
    auto pchunk = [&](int m) { const int ch = wv + m * NW; return ch < CPR ? ch : CPR - 1; };
    auto qchunk = [&](int m) { const int ch = wv + m * NW; return ch < CQD ? ch : CQD - 1; };
    unsigned poff[MPR];
#pragma unroll
    for (int m = 0; m < MPR; ++m) {
      const int g = pchunk(m) * 64 + ln;
      poff[m] = box_off<HK, HJ>(2 * (g < NPR ? g : 0), i0 - H, j0 - H, k0 - KH, a.nx, a.ny, a.nz);
    }
    unsigned qoff[SPLIT ? MQD : 1];
    if constexpr (SPLIT) {
#pragma unroll
      for (int m = 0; m < MQD; ++m) {
        const int g = qchunk(m) * 64 + ln;
        qoff[m] = box_off<HK4, HJ>(4 * (g < NQD ? g : 0), i0 - H, j0 - H, k0 - KH4, a.nx, a.ny, a.nz);
      }
    }
    // flux tiles: quad g of the x, y and z parts in turn; its element offset
    // counts from flux 3q (x) of the quantity, the y and z fluxes being the
    // next two arrays (the launcher keeps 3 * rstride * 4 bytes below 2^32)
    constexpr int NQG = (FGX + FGY + FGZ) / 4, CFG = (NQG + 63) / 64, MFG = (CFG + NW - 1) / NW;
    static_assert(!FGRP || (FGX % 4 == 0 && FGY % 4 == 0 && CFG <= CQD), "flux tiles: whole quads, one FP32 buffer");
    auto fchunk = [&](int m) { const int ch = wv + m * NW; return ch < CFG ? ch : CFG - 1; };
    unsigned foff[FGRP ? MFG : 1];
    if constexpr (FGRP) {
#pragma unroll
      for (int m = 0; m < MFG; ++m) {
        const int g = fchunk(m) * 64 + ln;
        const int e = 4 * (g < NQG ? g : 0);
        unsigned o;
        int d;
        if (e < FGX) {
          d = 0; o = box_off<TK, TJ>(e, i0 - H, j0, k0, a.nx, a.ny, a.nz);
        } else if (e < FGX + FGY) {
          d = 1; o = box_off<TK, HJ>(e - FGX, i0, j0 - H, k0, a.nx, a.ny, a.nz);
        } else {
          d = 2; o = box_off<HK4, TJ>(e - FGX - FGY, i0, j0, k0 - KH4, a.nx, a.ny, a.nz);
        }
        foff[m] = (unsigned)d * (unsigned)a.rstride + o;
      }
    }
    // two-field tiles of K and Gamma^i: quad g of field h = g / (NCR / 4),
    // then its x, y and z parts as a flux tile; the second field's register
    // is 4 * rstride elements after the first's (the launcher keeps
    // 5 * rstride * 4 bytes below 2^32)
    constexpr int NCR = FGX + FGY + FGZ, NQK = 2 * NCR / 4, CKG = (NQK + 63) / 64, MKG = (CKG + NW - 1) / NW;
    static_assert(!KGRP || (FGX % 4 == 0 && FGY % 4 == 0 && CKG <= CPR), "two-field tiles: whole quads, buffer 2");
    auto kchunk = [&](int m) { const int ch = wv + m * NW; return ch < CKG ? ch : CKG - 1; };
    unsigned koff[KGRP ? MKG : 1];
    if constexpr (KGRP) {
#pragma unroll
      for (int m = 0; m < MKG; ++m) {
        const int g = kchunk(m) * 64 + ln;
        int e = 4 * (g < NQK ? g : 0);
        const int h = e >= NCR ? 1 : 0;
        e -= h * NCR;
        unsigned o;
        if (e < FGX) o = box_off<TK, TJ>(e, i0 - H, j0, k0, a.nx, a.ny, a.nz);
        else if (e < FGX + FGY) o = box_off<TK, HJ>(e - FGX, i0, j0 - H, k0, a.nx, a.ny, a.nz);
        else o = box_off<HK4, TJ>(e - FGX - FGY, i0, j0, k0 - KH4, a.nx, a.ny, a.nz);
        koff[m] = (unsigned)h * 4u * (unsigned)a.rstride + o;
      }
    }
    typedef __attribute__((address_space(3))) char lds_char;
    const unsigned lds0 = (unsigned)(size_t)(lds_char*)lb0, lds1 = lds0 + P64, lds2 = lds0 + P64 + P32;
    auto dma = [&](const int f, const int b) {
      const unsigned dst = b == 0 ? lds0 : b == 1 ? lds1 : lds2;
      if (KGRP && f >= KG0) {
        const void* src = (const void*)INS(f == KG0 ? KK : GAM2);
#pragma unroll
        for (int m = 0; m < MKG; ++m) dma16(src, koff[m] * 4u, dst + kchunk(m) * 1024);
      } else if (FGRP && f >= FLUXG0) {
        const void* src = (const void*)((const TS*)a.flux + (long)(f - FLUXG0) * 3 * a.rstride);
#pragma unroll
        for (int m = 0; m < MFG; ++m) dma16(src, foff[m] * 4u, dst + fchunk(m) * 1024);
      } else if (!SPLIT || mfield<M32>(f)) {
        const void* src = mfield<M32>(f) ? (const void*)INM(f) : (const void*)INS(f);
#pragma unroll
        for (int m = 0; m < MPR; ++m) dma16(src, poff[m] * 8u, dst + pchunk(m) * 1024);
      } else {
        const void* src = (const void*)INS(f);
#pragma unroll
        for (int m = 0; m < MQD; ++m) dma16(src, qoff[m] * 4u, dst + qchunk(m) * 1024);
      }
    };
    auto seq = [&](int s) {
      return SPLIT ? stencil_seq_split<LIVE_ALPHA, FLUXES, FGRP, KGRP>(s) : stencil_seq<LIVE_ALPHA, FLUXES>(s);
    };
    dma(seq(0), 0);
    if (nseq > 1) dma(seq(1), 1);
#pragma unroll
    for (int s = 0; s < NSEQ; ++s) {
      if (s < nseq) {
        // position s has landed for this wave once at most the chunks of
        // position s + 1 are outstanding; for every wave after the barrier
        const int f1 = seq(s + 1 < NSEQ ? s + 1 : 0);
        if (s + 1 < nseq) {
          if (!SPLIT || mfield<M32>(f1)) dma_wait_barrier<MPR>();
          else if (KGRP && f1 >= KG0) dma_wait_barrier<MKG>();
          else if (FGRP && f1 >= FLUXG0) dma_wait_barrier<MFG>();
          else dma_wait_barrier<MQD>();
        } else {
          dma_wait_barrier<0>();
        }
        // buffer (s + 2) % 3 was last read at position s - 1, before the barrier
        if (s + 2 < nseq) dma(seq(s + 2 < NSEQ ? s + 2 : 0), (s + 2) % 3);
        const int f = seq(s);
        const void* const buf = s % 3 == 0 ? (const void*)lb0 : s % 3 == 1 ? (const void*)lb1 : (const void*)lb2;
        if (valid) { CG_DERIVE(f, buf); }
      }
    }
    // Every point value has arrived by now (the last staging wait drained
    // the vector memory counter).  The compiler does not see the asm waits:
    // touching the values here puts its own waits for them here, where they
    // cost nothing, instead of at their first use in the algebra, where the
    // epilogue prefetch is in flight (vector memory operations retire in
    // order, so those waits would also wait for the prefetch).
    dust_preload();
  #pragma unroll
    for (int f = 0; f < NF; ++f) asm volatile("" ::"v"(v[f]));
    asm volatile("" ::"v"(Dpt));
    if constexpr (PF) {
      if (a.stage >= 1) {
        // The staging buffers are free once every wave has differentiated
        // the last field.  Each wave then loads the _p (and from stage 2 the
        // _f) values of its own 64 points into its own LDS region; it waits
        // for them (vmcnt) before its epilogue, while the algebra runs.
        dma_wait_barrier<0>();
        auto poff = [&](int q) -> unsigned {  // element offset of block point q, 0 outside the grid
          const int qk = q % TK, qj = (q / TK) % TJ, qi = q / (TK * TJ);
          const int x = i0 + qi, y = j0 + qj, z = k0 + qk;
          return (x < a.nx && y < a.ny && z < a.nz) ? (unsigned)__umul24((unsigned)__umul24(x, a.ny) + y, a.nz) + z
                                                    : 0u;
        };
        const unsigned o2 = poff(64 * wv + 2 * (ln & 31)), o4 = poff(64 * wv + 4 * (ln & 15));
        const unsigned reg = lds0 + wv * PFR;
        const bool need_f = a.incr ? a.stage == 4 : a.stage >= 2;
        // each lane's address: a wave-uniform field base plus the lane's
        // element offset (its slot, and for FP32 chunks its field of the
        // pair), formed once
        const bool f64 = (ln >> 5) != 0;  // FP64 chunk: lanes 32-63 load _f
        const long l64 = (long)(f64 ? a.s_f : a.s_p) * a.rstride + o2;
  #pragma unroll
        for (int m = 0; m < N64; ++m) {
          const char* b = m < 7 ? (const char*)RM(m, 0) : (const char*)RS(m, 0);
          if (!f64 || need_f) dma16v(b + l64 * 8, reg + m * 1024);
        }
        const bool f32 = ((ln >> 4) & 1) != 0;  // FP32 chunk: lanes 16-31 and 48-63 load _f
        const long l32 = (long)((ln >> 5) * 4 + (f32 ? a.s_f : a.s_p)) * a.rstride + o4;
  #pragma unroll
        for (int i = 0; i < (N32 + 1) / 2; ++i) {
          const int n = 2 * i + (ln >> 5);  // field F32 + n, 4 * rstride elements after field F32 + 2i for odd lanes
          const char* b = (const char*)RS(F32 + 2 * i, 0);
          if (n < N32 && (!f32 || need_f)) dma16v(b + l32 * 4, reg + (N64 + i) * 1024);
        }
        if (PF2 && a.incr && a.stage == 4) {
          // a2 of field F32 + A2F + 4i + lane/16 in chunk N64 + CPF + i
          const long la = (long)((ln >> 4) * 4 + a.s_a2) * a.rstride + o4;
  #pragma unroll
          for (int i = 0; i < (NA2 + 3) / 4; ++i) {
            const char* b = (const char*)RS(F32 + A2F + 4 * i, 0);
            if (4 * i + (ln >> 4) < NA2) dma16v(b + la * 4, reg + (N64 + CPF + i) * 1024);
          }
        }
      }
    }
  } else {
    fetch(PHI);
    store(0, PHI);
    __syncthreads();

  #pragma unroll
    for (int s = 0; s < NSEQ; ++s) {
      if (s < nseq) {
        const bool more = (s + 1) < nseq;
        const int fn = stencil_seq<LIVE_ALPHA, FLUXES>((s + 1) < NSEQ ? s + 1 : 0);
        if (more) fetch(fn);
        const int f = stencil_seq<LIVE_ALPHA, FLUXES>(s);
        if (valid) { CG_DERIVE(f, ldsraw[s & 1]); }
        if (more) store((s + 1) & 1, fn);
        __syncthreads();
      }
    }
  }
#undef CG_DERIVE
  // Threads outside the grid (ragged edge tiles) skip the algebra but stay
  // alive for the block-wide setExtraFieldData reduction at the end.
  double phi_new = 0.0, K_new = 0.0, rho_new = 0.0;  // stage output and source seen by the reduction
  double cv[CONS_PT];  // constraint mode: value, scale, scaled of H, M, G, A, S at this point
#pragma unroll
  for (int q = 0; q < CONS_PT; ++q) cv[q] = 0.0;
  if (valid) {
    // ---------------------------------------------------------- algebra
    const TC phiF = (TC)fv.phi, KF = (TC)fv.K, rhoF = (TC)fv.rho, SF = (TC)fv.S;
    const TC PI = TC(3.14159265358979323846);
    const TC alpha = v[ALPHA] + TC(1);
    const TC phi = (TC)(phi_c + fv.phi);
    const TC K = v[KK] + KF;

    // sources (components/static/static.cc:28-35, components/Lambda/lambda.cc:32-41)
    TC DIFFr = TC(0), DIFFS = TC(0);
    TC Si[3] = {TC(0), TC(0), TC(0)}, STF[6] = {TC(0), TC(0), TC(0), TC(0), TC(0), TC(0)};
    TC Dc = TC(0), Sc[3] = {TC(0), TC(0), TC(0)}, Wc = TC(1), vc[3] = {TC(0), TC(0), TC(0)};
    if (a.matter == MATTER_STATIC) {
      const TC D = DMA ? Dpt : (TC)ldo((const TS*)a.DIFFD, idx);
      DIFFr = exp((TC)(-6.0 * (phi_c + fv.phi))) * D + rhoF * expm1(TC(-6) * v[PHI]);
    }
    if constexpr (DUST) {
      // Dust::getDustData at the point (dust.cc:177-200)
      Dc = PFD ? (TC)pdi[0] : (TC)ldo((const TS*)RD(0, a.s_in), idx);
  #pragma unroll
      for (int j = 0; j < 3; ++j) Sc[j] = PFD ? (TC)pdi[1 + j] : (TC)ldo((const TS*)RD(1 + j, a.s_in), idx);
      {
        const TC em4 = exp(TC(-4) * phi), e6 = exp(TC(6) * phi);
        const TC giSS = em4 * (gi[0] * Sc[0] * Sc[0] + gi[3] * Sc[1] * Sc[1] + gi[5] * Sc[2] * Sc[2] +
                               TC(2) * (gi[1] * Sc[0] * Sc[1] + gi[2] * Sc[0] * Sc[2] + gi[4] * Sc[1] * Sc[2]));
        Wc = sqrt(TC(1) + giSS / Dc / Dc);
  #pragma unroll
        for (int i = 0; i < 3; ++i) vc[i] = em4 / e6 * (GI(i, 0) * Sc[0] + GI(i, 1) * Sc[1] + GI(i, 2) * Sc[2]);
      }
      // Dust::addBSSNSrc (dust.cc:226-264): D, S_i of this stage with W,
      // detg, g_ij of the state the previous populateDerivedFields saw
      TC Wl = Wc, detg = exp(TC(6) * phi), e4l = exp(TC(4) * phi);
      TC gl[6];
  #pragma unroll
      for (int q = 0; q < 6; ++q) gl[q] = v[G11 + q] + (diag6(q) ? TC(1) : TC(0));
      if (a.lag_mode != 0) {
        double m[6];
  #pragma unroll
        for (int q = 0; q < 6; ++q)
          m[q] = PFD  ? (double)plw[PFD ? q : 0] * a.msc[q]
                 : MQ ? (double)ldo((const int*)RS(q, a.s_lag_met), idx) * a.msc[q]
                      : ldo((const double*)RM(q, a.s_lag_met), idx);
        if (a.lag_mode == 2) {
          const double tn = norm_factor<double>(m[0], m[1], m[2], m[3], m[4], m[5]);
  #pragma unroll
          for (int q = 0; q < 6; ++q) m[q] = (m[q] + (diag6(q) ? tn : 0.0)) / (1.0 - tn);
        }
        const TC phl = (TC)((PFD  ? (double)plw[PFD ? 6 : 0] * a.msc[PHI]
                             : MQ ? (double)ldo((const int*)RS(PHI, a.s_lag), idx) * a.msc[PHI]
                                  : ldo((const double*)RM(PHI, a.s_lag), idx)) +
                            a.phi_F_lag);
        const TC Dl = PFD ? (TC)pdl[0] : (TC)ldo((const TS*)RD(0, a.s_lag), idx);
        TC Sl[3];
  #pragma unroll
        for (int j = 0; j < 3; ++j) Sl[j] = PFD ? (TC)pdl[PFD ? 1 + j : 0] : (TC)ldo((const TS*)RD(1 + j, a.s_lag), idx);
        const TC g11 = (TC)m[0], g12 = (TC)m[1], g13 = (TC)m[2], g22 = (TC)m[3], g23 = (TC)m[4], g33 = (TC)m[5];
        const TC i11 = TC(1) + g22 + g33 - g23 * g23 + g22 * g33, i22 = TC(1) + g11 + g33 - g13 * g13 + g11 * g33;
        const TC i33 = TC(1) + g11 + g22 - g12 * g12 + g11 * g22, i12 = g13 * g23 - g12 * (TC(1) + g33);
        const TC i13 = g12 * g23 - g13 * (TC(1) + g22), i23 = g12 * g13 - g23 * (TC(1) + g11);
        const TC giSS = exp(TC(-4) * phl) * (i11 * Sl[0] * Sl[0] + i22 * Sl[1] * Sl[1] + i33 * Sl[2] * Sl[2] +
                                             TC(2) * (i12 * Sl[0] * Sl[1] + i13 * Sl[0] * Sl[2] + i23 * Sl[1] * Sl[2]));
        Wl = sqrt(TC(1) + giSS / Dl / Dl);
        detg = exp(TC(6) * phl);
        e4l = exp(TC(4) * phl);
  #pragma unroll
        for (int q = 0; q < 6; ++q) gl[q] = (TC)m[q] + (diag6(q) ? TC(1) : TC(0));
      }
      DIFFr += Wl * Dc / detg;
      DIFFS += Dc / detg * (Wl * Wl - TC(1)) / Wl;
  #pragma unroll
      for (int j = 0; j < 3; ++j) Si[j] = Sc[j] / detg;
  #pragma unroll
      for (int I = 0; I < 3; ++I)
  #pragma unroll
        for (int J = I; J < 3; ++J)
          STF[sy(I, J)] = Sc[I] * Sc[J] / Wl / Dc / detg - TC(1.0 / 3.0) * (e4l * gl[sy(I, J)]) * DIFFS;
      if (a.dust_bg) {
        // perturbation form: the FRW fluids carry the mean (their Lambda
        // fluid cancels the Lambda source added below); the trace-free
        // stress above is the dust's own
        DIFFr -= rhoF;
        DIFFS -= SF;
      }
    }
    DIFFr += (TC)a.lambda;
    DIFFS += TC(-3) * (TC)a.lambda;
    if constexpr (DUST) {
      if (a.rho_out) a.rho_out[idx] = (double)DIFFr;
    }
    if constexpr (DUST && !CONS) {
      // dust right-hand sides (dust.cc:203-224): flux divergence plus the
      // momentum source of populateDerivedFields (dust.cc:139-150), and the
      // fluid's RK4 update, done here, before the BSSN algebra, so that none
      // of their inputs (fluxes' divergence, v^i, W, D) stays live through it
      TC drhs[NDUST];
      drhs[0] = ddust[0];
      const TC pre = TC(0.5) * alpha * Wc * Dc * exp(TC(4) * phi);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const TC dp4 = TC(4) * dphi[k];
        const TC acc = vc[0] * vc[0] * (dp4 * GM(0, 0) + dg[k][0]) + vc[1] * vc[1] * (dp4 * GM(1, 1) + dg[k][3]) +
                       vc[2] * vc[2] * (dp4 * GM(2, 2) + dg[k][5]) +
                       TC(2) * (vc[0] * vc[1] * (dp4 * GM(0, 1) + dg[k][1]) + vc[0] * vc[2] * (dp4 * GM(0, 2) + dg[k][2]) +
                                vc[1] * vc[2] * (dp4 * GM(1, 2) + dg[k][4]));
        drhs[1 + k] = ddust[1 + k] + (pre * acc - Wc * Dc * dal[k]);
      }
#pragma unroll
      for (int q = 0; q < NDUST; ++q) {
        const double cc = (double)drhs[q];
        TS* __restrict__ o = RD(q, a.s_out);
        if (a.stage == 0) {
          sto(o, idx, (TS)cc);
          continue;
        }
        const int qq = PFD ? q : 0;
        const double p = PFD ? (double)pdp[qq] : (double)ldo((const TS*)RD(q, a.s_p), idx);
        TS* __restrict__ fr = RD(q, a.s_f);
        double fo = 0.0;
        if (!a.incr) {
          if (a.stage >= 2) fo = PFD ? (double)pd1[qq] : (double)ldo((const TS*)fr, idx);
        } else if (a.stage == 4) {
          // accumulator-free form: K3Finalize's accumulator from the three stage outputs
          const double a1 = PFD ? (double)pd1[qq] : (double)ldo((const TS*)fr, idx);
          const double a2 = PFD ? (double)pd2[qq] : (double)ldo((const TS*)RD(q, a.s_a2), idx);
          const double a3 = PFD ? (double)pdi[qq] : (double)ldo((const TS*)RD(q, a.s_in), idx);
          fo = ((a1 - p) + 2.0 * (a2 - p) + (a3 - p)) * (1.0 / 3.0);
        }
        const double inc = a.cf * cc;
        sto(o, idx, a.stage == 4 ? (TS)(fo + (inc + p)) : (TS)(p + a.ca * cc));
        if (a.stage != 4 && !a.incr) sto(fr, idx, (TS)(fo + inc));
      }
    }

    // A^ij and A_ij A^ij (bssn.cc:463-472)
    TC Ac[6];
  #pragma unroll
    for (int I = 0; I < 3; ++I)
  #pragma unroll
      for (int J = I; J < 3; ++J) {
        TC acc = TC(0);
  #pragma unroll
        for (int p = 0; p < 3; ++p) {
          const TC w = GI(I, p);
          acc += w * (GI(J, 0) * AM(p, 0) + GI(J, 1) * AM(p, 1) + GI(J, 2) * AM(p, 2));
        }
        Ac[sy(I, J)] = acc;
      }
    const TC AijAij = Ac[0] * v[A11] + Ac[3] * v[A22] + Ac[5] * v[A33] +
                      TC(2) * (Ac[1] * v[A12] + Ac[2] * v[A13] + Ac[4] * v[A23]);

    // Christoffels (bssn_macros.h:221-229):  GL_ijk = 1/2(d_j g_ki + d_k g_ji - d_i g_jk),
    // G^i_jk = gamma^{im} GL_mjk
    TC GL[3][6], G[3][6];
  #pragma unroll
    for (int I = 0; I < 3; ++I)
  #pragma unroll
      for (int J = 0; J < 3; ++J)
  #pragma unroll
        for (int Kx = J; Kx < 3; ++Kx)
          GL[I][sy(J, Kx)] = TC(0.5) * (dg[J][sy(Kx, I)] + dg[Kx][sy(J, I)] - dg[I][sy(J, Kx)]);
  #pragma unroll
    for (int I = 0; I < 3; ++I)
  #pragma unroll
      for (int s = 0; s < 6; ++s) G[I][s] = GI(I, 0) * GL[0][s] + GI(I, 1) * GL[1][s] + GI(I, 2) * GL[2][s];
    TC Gd[3];
  #pragma unroll
    for (int I = 0; I < 3; ++I)
      Gd[I] = G[I][0] * gi[0] + G[I][3] * gi[3] + G[I][5] * gi[5] + TC(2) * (G[I][1] * gi[1] + G[I][2] * gi[2] + G[I][4] * gi[4]);

    // D_i D_j phi (bssn.cc:592-602)
    TC DDphi[6];
  #pragma unroll
    for (int s = 0; s < 6; ++s) DDphi[s] = ddphi[s] - (G[0][s] * dphi[0] + G[1][s] * dphi[1] + G[2][s] * dphi[2]);

    // (D_i D_j alpha)^TF and its trace (bssn.cc:604-626)
    TC DDa[6], DDaTR = TC(0);
  #pragma unroll
    for (int q = 0; q < 6; ++q) DDa[q] = TC(0);
    if constexpr (LIVE_ALPHA) {
      TC gild[3];
  #pragma unroll
      for (int I = 0; I < 3; ++I) gild[I] = GI(I, 0) * dphi[0] + GI(I, 1) * dphi[1] + GI(I, 2) * dphi[2];
  #pragma unroll
      for (int I = 0; I < 3; ++I)
  #pragma unroll
        for (int J = I; J < 3; ++J) {
          const int s = sy(I, J);
          TC acc = TC(0);
  #pragma unroll
          for (int m = 0; m < 3; ++m)
            acc += (G[m][s] + TC(2) * ((m == I ? dphi[J] : TC(0)) + (m == J ? dphi[I] : TC(0)) - GM(I, J) * gild[m])) * dal[m];
          DDa[s] = ddal[s] - acc;
        }
      DDaTR = gi[0] * DDa[0] + gi[3] * DDa[3] + gi[5] * DDa[5] + TC(2) * (gi[1] * DDa[1] + gi[2] * DDa[2] + gi[4] * DDa[4]);
  #pragma unroll
      for (int s = 0; s < 6; ++s) DDa[s] -= TC(1.0 / 3.0) * (v[G11 + s] + (diag6(s) ? TC(1) : TC(0))) * DDaTR;
    }
    const TC em4phi = exp(TC(-4) * phi);
    DDaTR *= em4phi;

    // conformal Ricci (bssn_macros.h:248-277).  TERM4 is regrouped as
    //   sum_{K,M} X[K][I][M] Y[J][K][M] + (I<->J),
    //   X[K][I][M] = gamma^{ML} G^K_{LI},  Y[J][K][M] = GL_{J,KM} + 1/2 GL_{K,MJ}
    TC Z[3][3];
  #pragma unroll
    for (int I = 0; I < 3; ++I)
  #pragma unroll
      for (int J = 0; J < 3; ++J) Z[I][J] = TC(0);
  #pragma unroll
    for (int Kx = 0; Kx < 3; ++Kx) {
      TC X[3][3], Y[3][3];
  #pragma unroll
      for (int I = 0; I < 3; ++I)
  #pragma unroll
        for (int M = 0; M < 3; ++M)
          X[I][M] = GI(M, 0) * G[Kx][sy(0, I)] + GI(M, 1) * G[Kx][sy(1, I)] + GI(M, 2) * G[Kx][sy(2, I)];
  #pragma unroll
      for (int J = 0; J < 3; ++J)
  #pragma unroll
        for (int M = 0; M < 3; ++M) Y[J][M] = GL[J][sy(Kx, M)] + TC(0.5) * GL[Kx][sy(M, J)];
  #pragma unroll
      for (int I = 0; I < 3; ++I)
  #pragma unroll
        for (int J = 0; J < 3; ++J) Z[I][J] += X[I][0] * Y[J][0] + X[I][1] * Y[J][1] + X[I][2] * Y[J][2];
    }
    TC ric[6];
    {
      // phi part (bssn.cc:639-655)
      const TC expression =
          gi[0] * (DDphi[0] + TC(2) * dphi[0] * dphi[0]) + gi[3] * (DDphi[3] + TC(2) * dphi[1] * dphi[1]) +
          gi[5] * (DDphi[5] + TC(2) * dphi[2] * dphi[2]) +
          TC(2) * (gi[1] * (DDphi[1] + TC(2) * dphi[0] * dphi[1]) + gi[2] * (DDphi[2] + TC(2) * dphi[0] * dphi[2]) +
                   gi[4] * (DDphi[4] + TC(2) * dphi[1] * dphi[2]));
  #pragma unroll
      for (int I = 0; I < 3; ++I)
  #pragma unroll
        for (int J = I; J < 3; ++J) {
          const int s = sy(I, J);
          TC t2 = TC(0), t3 = TC(0);
  #pragma unroll
          for (int k = 0; k < 3; ++k) {
            t2 += GM(k, I) * dG[J][k] + GM(k, J) * dG[I][k];
            t3 += Gd[k] * (GL[I][sy(J, k)] + GL[J][sy(I, k)]);
          }
          const TC U = TC(-0.5) * T1[s] + TC(0.5) * t2 + TC(0.5) * t3 + (Z[I][J] + Z[J][I]);
          ric[s] = U - TC(2) * (DDphi[s] - TC(2) * dphi[I] * dphi[J] + GM(I, J) * expression);
        }
    }
    const TC ricci = em4phi * (ric[0] * gi[0] + ric[3] * gi[3] + ric[5] * gi[5] +
                               TC(2) * (ric[1] * gi[1] + ric[2] * gi[2] + ric[4] * gi[4]));
    const TC e4phi = exp(TC(4) * phi);

    // Hamiltonian constraint (bssn.cc:1073-1084), only when a damping term uses it
    TC Hc = TC(0);
    if (a.use_H || CONS) {
      const TC rr = DIFFr + rhoF;
      if constexpr (sizeof(TC) == 8) {
        Hc = -exp(TC(5) * phi) / TC(8) * (ricci + TC(2.0 / 3.0) * K * K - AijAij - TC(16) * PI * rr);
      } else {
        // split precision: (2/3)K^2 - 16 pi rho in difference form around FRW
        const TC dK2 = TC(2.0 / 3.0) * (TC(2) * KF + v[KK]) * v[KK];
        Hc = -exp(TC(5) * phi) / TC(8) * (ricci + ((TC)fv.H + dK2) - AijAij - TC(16) * PI * DIFFr);
      }
    }

    if constexpr (CONS) {
      // setConstraintCalcs (bssn.cc:965-1071) per-point quantities, FP64
      double* __restrict__ co = a.cons;
      const long cs = a.cons_stride;
      auto put = [&](int q, double val) {
        if (q < CONS_PT) cv[q] = val;
        if (!a.cons_rows) co[(long)q * cs + idx] = val;
      };
      const double rr = (double)DIFFr + fv.rho;
      const double e5 = exp(5.0 * (double)phi), e6 = exp(6.0 * (double)phi);
      // hamiltonianConstraintScale (bssn.cc:1086-1100)
      const double Hs = (e5 / 8.0) * sqrt((double)ricci * ricci + (double)AijAij * AijAij +
                                          pow(2.0 / 3.0 * (double)K * K, 2) + pow(16.0 * 3.14159265358979323846 * rr, 2));
      put(CONS_H, Hc); put(CONS_H + 1, Hs); put(CONS_H + 2, Hc / Hs);
      // BSSN_MI / BSSN_MI_SCALE (bssn_macros.h:487-527)
      double M2 = 0, Ms2 = 0;
  #pragma unroll
      for (int I = 0; I < 3; ++I) {
        double t6 = 0, td = 0;
  #pragma unroll
        for (int n = 0; n < 3; ++n)
  #pragma unroll
          for (int m = 0; m < 3; ++m) {
            t6 += GI(m, n) * AM(m, I) * dphi[n];
            td += GI(m, n) * dA[n][sy(m, I)];
          }
  #pragma unroll
        for (int m = 0; m < 3; ++m) td -= Gd[m] * AM(m, I);
  #pragma unroll
        for (int n = 0; n < 3; ++n)
  #pragma unroll
          for (int m = 0; m < 3; ++m) td -= GL[m][sy(n, I)] * Ac[sy(m, n)];
        const double PI8S = 8.0 * 3.14159265358979323846 * (double)Si[I];
        const double MI = e6 * (-2.0 / 3.0 * dK[I] - PI8S + 6.0 * t6 + td);
        const double MS = e6 * (fabs(2.0 / 3.0 * dK[I]) + fabs(PI8S) + 6.0 * fabs(t6) + fabs(td));
        M2 += MI * MI;
        Ms2 += MS * MS;
        put(CONS_MI + I, MI);  // momentumConstraintCalc(bd, I + 1) (bssn.cc:1102-1118)
      }
      const double Mv = sqrt(M2), Msc = sqrt(Ms2);
      put(CONS_M, Mv); put(CONS_M + 1, Msc); put(CONS_M + 2, Mv / Msc);
      // BSSN_GI_CALC / BSSN_GI_SCALE (bssn_macros.h:530-536)
      double G2 = 0, Gs2 = 0;
  #pragma unroll
      for (int I = 0; I < 3; ++I) {
        const double tr = gi[0] * G[I][0] + gi[3] * G[I][3] + gi[5] * G[I][5] +
                          2.0 * (gi[1] * G[I][1] + gi[2] * G[I][2] + gi[4] * G[I][4]);
        const double gc_ = v[GAM1 + I] - tr;
        const double gs_ = fabs((double)v[GAM1 + I]) +
                           fabs(gi[0] * G[I][0] - gi[3] * G[I][3] - gi[5] * G[I][5] -
                                2.0 * (gi[1] * G[I][1] + gi[2] * G[I][2] + gi[4] * G[I][4]));
        G2 += gc_ * gc_;
        Gs2 += gs_ * gs_;
      }
      put(CONS_G, sqrt(G2)); put(CONS_G + 1, sqrt(Gs2)); put(CONS_G + 2, sqrt(G2) / sqrt(Gs2));
      // AijTFConstraintCalc / Scale (bssn.cc:1182-1192)
      const double At = gi[0] * v[A11] + gi[3] * v[A22] + gi[5] * v[A33] +
                        2.0 * (gi[1] * v[A12] + gi[2] * v[A13] + gi[4] * v[A23]);
      const double As = fabs(gi[0] * v[A11]) + fabs(gi[3] * v[A22]) + fabs(gi[5] * v[A33]) +
                        2.0 * (fabs(gi[1] * v[A12]) + fabs(gi[2] * v[A13]) + fabs(gi[4] * v[A23]));
      put(CONS_A, At); put(CONS_A + 1, As); put(CONS_A + 2, At / As);
      // unitDetConstraintCalc / Scale (bssn.cc:1199-1219)
      {
        const double g11 = v[G11], g12 = v[G12], g13 = v[G13], g22 = v[G22], g23 = v[G23], g33 = v[G33];
        const double Sv = fabs(g11 + g22 + g33 - g12 * g12 - g13 * g13 - g23 * g23 + g11 * g22 + g11 * g33 + g22 * g33 -
                               g23 * g23 * g11 - g13 * g13 * g22 - g12 * g12 * g33 + 2.0 * g12 * g13 * g23 +
                               g11 * g22 * g33);
        const double Ss = fabs(g11) + fabs(g22) + fabs(g33) + g12 * g12 + g13 * g13 + g23 * g23 + fabs(g11 * g22) +
                          fabs(g11 * g33) + fabs(g22 * g33) + g23 * g23 * g11 + g13 * g13 * g22 + g12 * g12 * g33 +
                          2.0 * fabs(g12 * g13 * g23) + fabs(g11 * g22 * g33);
        put(CONS_S, Sv); put(CONS_S + 1, Ss); put(CONS_S + 2, Sv / Ss);
      }
      put(CONS_RICCI, ricci);
      put(CONS_AIJAIJ, AijAij);
      put(CONS_RHO, DIFFr);   // the sources the output step sees (DIFFr_a, DIFFS_a)
      put(CONS_SRCS, DIFFS);
    } else {

      // ------------------------------------------------ right-hand sides
      TC rhs[NF];
      const TC dt = (TC)a.dt;
      const TC adj = (TC)a.a_adj;
    #pragma unroll
      for (int s = 0; s < 6; ++s)
        rhs[G11 + s] = TC(-2) * alpha * v[A11 + s] + TC(0.5) * adj * dt * Hc * v[G11 + s] - ko[G11 + s];
      // bssn_macros.h:309-338
    #pragma unroll
      for (int I = 0; I < 3; ++I)
    #pragma unroll
        for (int J = I; J < 3; ++J) {
          const int s = sy(I, J);
          TC AA = TC(0);
    #pragma unroll
          for (int m = 0; m < 3; ++m)
            AA += AM(m, I) * (GI(m, 0) * AM(0, J) + GI(m, 1) * AM(1, J) + GI(m, 2) * AM(2, J));
          const TC ricTF = ric[s] - TC(1.0 / 3.0) * e4phi * GM(I, J) * ricci;
          rhs[A11 + s] = em4phi * (alpha * (ricTF - TC(8) * PI * STF[s]) - DDa[s]) +
                         alpha * (K * v[A11 + s] - TC(2) * AA) - adj * dt * v[A11 + s] * Hc - ko[A11 + s];
        }
      // bssn_macros.h:340-361
    #pragma unroll
      for (int I = 0; I < 3; ++I) {
        const TC second = G[I][0] * Ac[0] + G[I][3] * Ac[3] + G[I][5] * Ac[5] +
                          TC(2) * (G[I][1] * Ac[1] + G[I][2] * Ac[2] + G[I][4] * Ac[4]) +
                          TC(6) * (Ac[sy(I, 0)] * dphi[0] + Ac[sy(I, 1)] * dphi[1] + Ac[sy(I, 2)] * dphi[2]);
        rhs[GAM1 + I] = TC(-2) * (Ac[sy(I, 0)] * dal[0] + Ac[sy(I, 1)] * dal[1] + Ac[sy(I, 2)] * dal[2]) +
                        TC(2) * alpha *
                            (second - TC(2.0 / 3.0) * (GI(I, 0) * dK[0] + GI(I, 1) * dK[1] + GI(I, 2) * dK[2]) -
                             TC(8) * PI * (GI(I, 0) * Si[0] + GI(I, 1) * Si[1] + GI(I, 2) * Si[2])) -
                        ko[GAM1 + I];
      }
      // bssn.cc:766-795
      rhs[KK] = -DDaTR + alpha * (TC(1.0 / 3.0) * v[KK] * (v[KK] + TC(2) * KF) + AijAij) +
                TC(4) * PI * alpha * (DIFFr + DIFFS) + TC(4) * PI * v[ALPHA] * (rhoF + SF) -
                (TC)a.k_damp * Hc * exp(TC(-5) * phi) - ko[KK];
      // bssn.cc:797-820
      rhs[PHI] = TC(0.1) * adj * dt * Hc - TC(1.0 / 6.0) * (alpha * v[KK] + v[ALPHA] * KF) - ko[PHI];
      // BSSNGaugeHandler.cc (the shift-free gauges; beta = 0)
      TC lapse = TC(0);
      switch (a.lapse) {
        case LAPSE_HARMONIC: lapse = -alpha * K; break;                                          // :32-35
        case LAPSE_CONFORMAL_FLRW: lapse = TC(-1.0 / 3.0) * alpha * (TC)a.ext_in[EXTV_KAVG]; break;  // :24-27
        case LAPSE_ANHARMONIC: lapse = alpha * alpha * (K - (TC)a.ext_in[EXTV_KAVG]); break;     // :50-58
        case LAPSE_ONE_PLUS_LOG:                                                                 // :64-68
          lapse = TC(-2) * alpha * (K - (TC)a.ext_in[EXTV_KAVG]) * (TC)a.gd_c;
          break;
        case LAPSE_CONFORMAL_SYNC: lapse = TC(-1.0 / 3.0) * alpha * KF; break;                   // :84-87
        case LAPSE_AWA_GAUGE_WAVE: lapse = -alpha * alpha * v[KK]; break;                        // :209-212
        case LAPSE_AWA_SHIFTED_WAVE: lapse = -v[KK]; break;                                      // :217-220
        case LAPSE_DAMPED_WAVE:                                                                  // :151-155
          lapse = alpha * alpha * ((TC)a.dw_mu_l * (TC(12) * phi * (TC)a.dw_p - log(alpha)) - K);
          break;
        default: break;  // Static, TestAijDriverLapse
      }
      rhs[ALPHA] = lapse - ko[ALPHA];

      // ------------------------------------------------ RK4Register epilogue
      // K{n}Finalize (RK4Register.h:129-176): f += cf*c ; a' = p + ca*c, and at
      // stage 4 p' = f + (cf*c + p).  Accumulation in FP64 for every precision.
      const int st = a.stage;
      rho_new = (double)DIFFr;
      double gn[6], An[6];  // stage 4: the new DIFFgamma_ij and A_ij as stored
      // DMA kernels: _p and _f of this point from the wave's LDS prefetch
      const char* const pfr = ldma + (tid >> 6) * PFR;
      const int pq = tid & 63;
      auto lpf = [&](int f, int kind) -> double {
        if (!SPLIT || (!M32 && f < 7)) return *reinterpret_cast<const double*>(pfr + f * 1024 + kind * 512 + pq * 8);
        const int n = f - F32;
        const char* w = pfr + (N64 + (n >> 1)) * 1024 + (n & 1) * 512 + kind * 256 + pq * 4;
        if (MQ && f < 7) return (double)*reinterpret_cast<const int*>(w) * a.msc[f < 7 ? f : 0];
        return (double)*reinterpret_cast<const TS*>(w);
      };
      // stage 4: a2 of field f (f - F32 >= A2F) from the prefetch
      auto la2 = [&](int f) -> double {
        const int n = f - F32 - A2F;
        const char* w = pfr + (N64 + CPF + (n >> 2)) * 1024 + (n & 3) * 256 + pq * 4;
        if (MQ && f < 7) return (double)*reinterpret_cast<const int*>(w) * a.msc[f < 7 ? f : 0];
        return (double)*reinterpret_cast<const TS*>(w);
      };
      if constexpr (PF) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      // one loop per RK4 form, each with only its own loads and stores
      auto epilogue = [&](auto incr_tag) {
        constexpr bool INC = decltype(incr_tag)::value;
        // The epilogue's loads from memory (stage 4 of the accumulator-free
        // form: a2 and a3 of the FP64 metric; without the LDS prefetch also p
        // and f / a1).  The compiler cannot move a load above a store into the
        // same register allocation, so loads issued field by field would each
        // wait out a memory latency; they are issued AHEAD fields before their
        // use instead (a few registers per field).
        constexpr int AHEAD = PF ? 4 : 2;
        double a2v[NF], a3m[7], pv[NF], qv[NF];
        auto pre = [&](int g) {
          if (g >= NF || st == 0 || (g == ALPHA && !LIVE_ALPHA)) return;
          // a storage-type word as FP64 (fixed-point metric words scaled)
          auto lw = [&](int slot) -> double {
            if (MQ && g < 7) return (double)ldo((const int*)RS(g, slot), idx) * a.msc[g < 7 ? g : 0];
            return (double)ldo((const TS*)RS(g, slot), idx);
          };
          if constexpr (!PF) {
            const bool m = mfield<M32>(g);
            pv[g] = m ? ldo((const double*)RM(g, a.s_p), idx) : lw(a.s_p);
            if (INC ? st == 4 : st >= 2) qv[g] = m ? ldo((const double*)RM(g, a.s_f), idx) : lw(a.s_f);
          }
          if constexpr (INC) {
            if (st == 4) {
              if (mfield<M32>(g)) {
                a2v[g] = ldo((const double*)RM(g, a.s_a2), idx);
                a3m[g] = ldo(INM(g), idx);
              } else {
                a2v[g] = PF2 && g - F32 >= A2F ? la2(g) : lw(a.s_a2);
                // the fixed-point metric's stage-4 input, re-read here rather
                // than kept in registers through the stencil phase
                if (MQ && g < 7) a3m[g < 7 ? g : 0] = lw(a.s_in);
              }
            }
          }
        };
    #pragma unroll
        for (int g = 0; g < AHEAD; ++g) pre(g);
    #pragma unroll
        for (int f = 0; f < NF; ++f) {
          pre(f + AHEAD);
          if (f == ALPHA && !LIVE_ALPHA) continue;  // lapse registers stay identically zero
          const double cc = (double)rhs[f];
          if (mfield<M32>(f)) {
            double* __restrict__ o = RM(f, a.s_out);
            if (st == 0) {
              sto(o, idx, cc);
              continue;
            }
            const double p = PF ? lpf(f, 0) : pv[f];
            double* __restrict__ fr = RM(f, a.s_f);
            double fo = 0.0;
            if constexpr (!INC) {
              if (st >= 2) fo = PF ? lpf(f, 1) : qv[f];
            } else if (st == 4) {
              const double a1 = PF ? lpf(f, 1) : qv[f];
              fo = ((a1 - p) + 2.0 * (a2v[f] - p) + (a3m[f] - p)) * (1.0 / 3.0);
            }
            const double inc = a.cf * cc;
            const double nv = st == 4 ? fo + (inc + p) : p + a.ca * cc;
            if (f == PHI) phi_new = nv;
            if (f <= G33) gn[f] = nv;
            sto(o, idx, nv);
            if (st == 4) {
              if (f == PHI && nan_as_float(nv)) atomicAdd(a.nan_count, 1u);
            } else if (!INC) {
              sto(fr, idx, fo + inc);
            }
          } else if (MQ && f < 7) {
            // fixed-point metric: the RK4 value in FP64, stored as the
            // field's word; phi_new and gn are the stored values
            int* __restrict__ o = (int*)RS(f, a.s_out);
            const double sc = a.msc[f < 7 ? f : 0], iv = a.minv[f < 7 ? f : 0];
            if (st == 0) {
              sto(o, idx, mq_word(cc, iv));
              continue;
            }
            const double p = PF ? lpf(f, 0) : pv[f];
            double fo = 0.0;
            if constexpr (!INC) {
              if (st >= 2) fo = PF ? lpf(f, 1) : qv[f];
            } else if (st == 4) {
              const double a1 = PF ? lpf(f, 1) : qv[f];
              fo = ((a1 - p) + 2.0 * (a2v[f] - p) + (a3m[f < 7 ? f : 0] - p)) * (1.0 / 3.0);
            }
            const double inc = a.cf * cc;
            const double nd = st == 4 ? fo + (inc + p) : p + a.ca * cc;
            if (fabs(nd * iv) > MQ_WMAX) atomicAdd(a.mq_sat, 1u);
            const int w = mq_word(nd, iv);
            const double nv = (double)w * sc;
            if (f == PHI) {
              phi_new = nv;
              if (st == 4 && nan_as_float(nd)) atomicAdd(a.nan_count, 1u);
            } else {
              gn[f < 6 ? f : 0] = nv;
            }
            sto(o, idx, w);
            if (st != 4 && !INC) sto((int*)RS(f, a.s_f), idx, mq_word(fo + inc, iv));
          } else {
            TS* __restrict__ o = RS(f, a.s_out);
            if (st == 0) {
              sto(o, idx, (TS)cc);
              continue;
            }
            const double p = PF ? lpf(f, 0) : pv[f];
            TS* __restrict__ fr = RS(f, a.s_f);
            double fo = 0.0;
            if constexpr (!INC) {
              if (st >= 2) fo = PF ? lpf(f, 1) : qv[f];
            } else if (st == 4) {
              const double a1 = PF ? lpf(f, 1) : qv[f];
              const double a2 = a2v[f];
              // the stored value: v holds it exactly when storage and compute types agree
              const double a3 = std::is_same<TS, TC>::value ? (double)v[f] : (double)ldo(INS(f), idx);
              fo = ((a1 - p) + 2.0 * (a2 - p) + (a3 - p)) * (1.0 / 3.0);
            }
            const double inc = a.cf * cc;
            const TS nv = st == 4 ? (TS)(fo + (inc + p)) : (TS)(p + a.ca * cc);
            if (f == KK) K_new = (double)nv;
            if (f >= A11 && f <= A33) An[f - A11] = (double)nv;
            if (M32 && f == PHI) {
              phi_new = (double)nv;
              if (st == 4 && nan_as_float(nv)) atomicAdd(a.nan_count, 1u);
            }
            if (M32 && f <= G33) gn[f] = (double)nv;
            sto(o, idx, nv);
            if (st != 4 && !INC) sto(fr, idx, (TS)(fo + inc));
          }
        }
      };
      if (a.incr) epilogue(std::true_type{});
      else epilogue(std::false_type{});
      if (DMA && st == 4 && a.s_norm_out >= 0) {
        // set_DIFFgamma_Aij_norm (bssn.cc:61-119) of the new state, as the
        // next step's stepInit applies it to _a (DMA kernels only: the
        // launcher pairs the copy with DMA staging)
        const double tn = norm_factor<double>(gn[0], gn[1], gn[2], gn[3], gn[4], gn[5]);
        double g[6];
    #pragma unroll
        for (int s = 0; s < 6; ++s) g[s] = (gn[s] + (diag6(s) ? tn : 0.0)) / (1.0 - tn);
        const double i11 = 1.0 + g[3] + g[5] - g[4] * g[4] + g[3] * g[5];
        const double i22 = 1.0 + g[0] + g[5] - g[2] * g[2] + g[0] * g[5];
        const double i33 = 1.0 + g[0] + g[3] - g[1] * g[1] + g[0] * g[3];
        const double i12 = g[2] * g[4] - g[1] * (1.0 + g[5]);
        const double i13 = g[1] * g[4] - g[2] * (1.0 + g[3]);
        const double i23 = g[1] * g[2] - g[4] * (1.0 + g[0]);
        const double trA = i11 * An[0] + i22 * An[3] + i33 * An[5] + 2.0 * (i12 * An[1] + i13 * An[2] + i23 * An[4]);
    #pragma unroll
        for (int s = 0; s < 6; ++s) {
          if constexpr (MQ) {
            if (fabs(g[s] * a.minv[s]) > MQ_WMAX) atomicAdd(a.mq_sat, 1u);
            sto((int*)RS(G11 + s, a.s_norm_out), idx, mq_word(g[s], a.minv[s]));
          }
          else if constexpr (M32) sto(RS(G11 + s, a.s_norm_out), idx, (TS)g[s]);
          else sto(RM(G11 + s, a.s_norm_out), idx, g[s]);
          sto(RS(A11 + s, a.s_norm_out), idx,
              (TS)((An[s] - (1.0 / 3.0) * ((diag6(s) ? 1.0 : 0.0) + g[s]) * trA) / (1.0 - tn)));
        }
      }
    }  // constraint mode / stage
  }  // valid
  if constexpr (CONS) {
    if (a.cons_rows)
      block_cons_put(a.cons_rows + (long)(t - a.tile_base) * CONS_ROW_W, a.cons_pass, a.cons_means, valid, cv);
    return;
  }

  // setExtraFieldData after K{n}Finalize (bssn.cc:267-305): the averages of
  // the new _a weighted with the new FRW phi; DIFFr is this stage's source.
  // Only the ConformalFLRW lapse reads them inside a step, so only the
  // live-lapse variants carry the reduction (the host computes the values
  // of the last stage on demand otherwise).
  if constexpr (LIVE_ALPHA) {
    if (a.ext_part && a.stage >= 1) block_ext_put(a.ext_part, t, valid, phi_new, K_new, rho_new, fv.phi_next);
  }
}

template <typename TS, typename TC, int R, int H, int TI, int TJ, int TK, int WPE, bool LIVE_ALPHA, int MODE = 0,
          bool DMA = false>
__global__ void __launch_bounds__(TI* TJ* TK) __attribute__((amdgpu_waves_per_eu(WPE)))
stage_kernel(const StageArgs a) {
  const int b = blockIdx.x;
  const int t = (b & 7) * a.per_xcd + (b >> 3);  // consecutive tiles stay on one XCD
  if (t >= (a.tile_count ? a.tile_count : a.ntx * a.nty * a.ntz)) return;
  stage_body<TS, TC, R, H, TI, TJ, TK, LIVE_ALPHA, MODE, DMA>(
      a, a.tile_base + t, FrwV{a.phi_F, a.K_F, a.rho_F, a.S_F, a.H_frw, a.phi_F_next});
}

// Batched ensemble: one launch over the tiles of many grids.  Tiles are
// numbered grid after grid (ba.tstart[r] = first tile of grid r) and mapped
// to blocks with the same XCD-aware remap as a single grid, so each XCD's L2
// sees neighbouring tiles of one grid.  (A persistent loop over tiles would
// let the compiler hoist per-thread tile geometry out of the loop and spill.)
template <typename TS, typename TC, int R, int H, int TI, int TJ, int TK, int WPE, bool LIVE_ALPHA, bool DMA = false>
__global__ void __launch_bounds__(TI* TJ* TK) __attribute__((amdgpu_waves_per_eu(WPE)))
stage_kernel_batch(const BatchArgs ba) {
  const int b = blockIdx.x;
  const int g = (b & 7) * ba.per_xcd + (b >> 3);
  if (g >= ba.total) return;
  // The tables are read through the constant address space: the kernel never
  // writes them, so the reads are scalar loads, as for kernel arguments.  (As
  // generic loads they would be vector loads made uniform, reissued after
  // every asm wait of the DMA staging, which clobbers memory.)  The host
  // compilation pass only checks this body and has no such address space.
#if defined(__HIP_DEVICE_COMPILE__)
  typedef __attribute__((address_space(4))) const int CInt;
  typedef __attribute__((address_space(4))) const double CDouble;
  typedef __attribute__((address_space(4))) const StageArgs CArgs;
#else
  typedef const int CInt;
  typedef const double CDouble;
  typedef const StageArgs CArgs;
#endif
  CInt* const ts = (CInt*)ba.tstart;
  int r = 0, n = ba.nreal;  // last grid whose first tile is <= g
  while (n > 1) {
    const int h = n >> 1;
    if (ts[r + h] <= g) { r += h; n -= h; } else { n = h; }
  }
  CDouble* const f = (CDouble*)ba.frw + (long)r * FRW_N;
  const StageArgs a = ((CArgs*)ba.tab)[r];
  stage_body<TS, TC, R, H, TI, TJ, TK, LIVE_ALPHA, 0, DMA>(a, g - ts[r], FrwV{f[0], f[1], f[2], f[3], f[4], f[5]});
}

// ------------------------------------------------------------ variants
template <typename TS, typename TC, int R, int H, int TI, int TJ, int TK, int WPE, bool LA, int MODE = MODE_STAGE,
          bool DMA = false>
int stage_launch_t(const StageArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((stage_kernel<TS, TC, R, H, TI, TJ, TK, WPE, LA, MODE, DMA>), dim3(8 * a.per_xcd),
                     dim3(TI * TJ * TK), 0, s, a);
  return (int)hipGetLastError();
}

template <typename TS, typename TC, int R, int H, int TI, int TJ, int TK, int WPE, bool LA, bool DMA = false>
int stage_launch_batch_t(const BatchArgs& ba, hipStream_t s) {
  hipLaunchKernelGGL((stage_kernel_batch<TS, TC, R, H, TI, TJ, TK, WPE, LA, DMA>), dim3(8 * ba.per_xcd),
                     dim3(TI * TJ * TK), 0, s, ba);
  return (int)hipGetLastError();
}

template <typename TS, typename TC, int R, int H, int TI, int TJ, int TK, int WPE, bool LA, int MODE = MODE_STAGE,
          bool DMA = false>
int stage_info_t(int* vgprs, int* spills, int* lds) {
  hipFuncAttributes attr;
  hipError_t e =
      hipFuncGetAttributes(&attr, (const void*)stage_kernel<TS, TC, R, H, TI, TJ, TK, WPE, LA, MODE, DMA>);
  if (e != hipSuccess) return (int)e;
  *vgprs = attr.numRegs;
  *spills = (int)attr.localSizeBytes;
  *lds = (int)attr.sharedSizeBytes;
  return 0;
}

template <typename TS, int R, bool LA, int MODE = MODE_CONS>
int cons_launch_t(const StageArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((stage_kernel<TS, double, R, R, 4, 8, 8, 1, LA, MODE>), dim3(8 * a.per_xcd), dim3(256), 0, s, a);
  return (int)hipGetLastError();
}

#define CG_REGISTER_STAGE_X(TS, TC, PREC, R, H, TI, TJ, TK, WPE, LA, MODE, BATCH, DMA_L, DMA_I, DMA_B) \
  static StageVariantReg CG_CAT(reg_, __COUNTER__)(                                                    \
      StageVariant{PREC, 2 * R, H, TI, TJ, TK, WPE, LA, (MODE & MODE_DUST) != 0,                     \
                   &stage_launch_t<TS, TC, R, H, TI, TJ, TK, WPE, LA, MODE>,                          \
                   &stage_info_t<TS, TC, R, H, TI, TJ, TK, WPE, LA, MODE>, BATCH, DMA_L, DMA_I, DMA_B});
#define CG_DMA_L(TS, TC, R, H, TI, TJ, TK, WPE, LA, MODE) (&stage_launch_t<TS, TC, R, H, TI, TJ, TK, WPE, LA, MODE, true>)
#define CG_DMA_I(TS, TC, R, H, TI, TJ, TK, WPE, LA, MODE) (&stage_info_t<TS, TC, R, H, TI, TJ, TK, WPE, LA, MODE, true>)
// single-grid kernel only
#define CG_REGISTER_STAGE(TS, TC, PREC, R, H, TI, TJ, TK, WPE, LA) \
  CG_REGISTER_STAGE_X(TS, TC, PREC, R, H, TI, TJ, TK, WPE, LA, MODE_STAGE, nullptr, nullptr, nullptr, nullptr)
// single-grid and batched-ensemble kernels
#define CG_REGISTER_STAGE_B(TS, TC, PREC, R, H, TI, TJ, TK, WPE, LA)                                   \
  CG_REGISTER_STAGE_X(TS, TC, PREC, R, H, TI, TJ, TK, WPE, LA, MODE_STAGE,                            \
                      (&stage_launch_batch_t<TS, TC, R, H, TI, TJ, TK, WPE, LA>), nullptr, nullptr, nullptr)
// single-grid kernel, with its LDS-DMA twin
#define CG_REGISTER_STAGE_D(TS, TC, PREC, R, H, TI, TJ, TK, WPE, LA)                                   \
  CG_REGISTER_STAGE_X(TS, TC, PREC, R, H, TI, TJ, TK, WPE, LA, MODE_STAGE, nullptr,                   \
                      CG_DMA_L(TS, TC, R, H, TI, TJ, TK, WPE, LA, MODE_STAGE),                        \
                      CG_DMA_I(TS, TC, R, H, TI, TJ, TK, WPE, LA, MODE_STAGE), nullptr)
// single-grid and batched-ensemble kernels, each with its LDS-DMA twin
#define CG_REGISTER_STAGE_BD(TS, TC, PREC, R, H, TI, TJ, TK, WPE, LA)                                  \
  CG_REGISTER_STAGE_X(TS, TC, PREC, R, H, TI, TJ, TK, WPE, LA, MODE_STAGE,                            \
                      (&stage_launch_batch_t<TS, TC, R, H, TI, TJ, TK, WPE, LA>),                     \
                      CG_DMA_L(TS, TC, R, H, TI, TJ, TK, WPE, LA, MODE_STAGE),                        \
                      CG_DMA_I(TS, TC, R, H, TI, TJ, TK, WPE, LA, MODE_STAGE),                        \
                      (&stage_launch_batch_t<TS, TC, R, H, TI, TJ, TK, WPE, LA, true>))
// metric + dust fluid (single grid)
#define CG_REGISTER_DUST(TS, TC, PREC, R, H, TI, TJ, TK, WPE, LA) \
  CG_REGISTER_STAGE_X(TS, TC, PREC, R, H, TI, TJ, TK, WPE, LA, MODE_DUST, nullptr, nullptr, nullptr, nullptr)
// metric + dust fluid (single grid), with its LDS-DMA twin
#define CG_REGISTER_DUST_D(TS, TC, PREC, R, H, TI, TJ, TK, WPE, LA)                                    \
  CG_REGISTER_STAGE_X(TS, TC, PREC, R, H, TI, TJ, TK, WPE, LA, MODE_DUST, nullptr,                    \
                      CG_DMA_L(TS, TC, R, H, TI, TJ, TK, WPE, LA, MODE_DUST),                         \
                      CG_DMA_I(TS, TC, R, H, TI, TJ, TK, WPE, LA, MODE_DUST), nullptr)
#define CG_CAT2(a, b) a##b
#define CG_CAT(a, b) CG_CAT2(a, b)

}  // namespace cg
