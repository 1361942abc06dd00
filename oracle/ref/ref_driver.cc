// ref_driver.cc -- exposes pieces of the reference CosmoGRaPH sources that
// build from their own files (no FFTW/HDF5/boost needed) behind a C ABI, so
// tests can compare the oracle against them value by value.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile (target `ref`) from
// the reference tree in place; outputs go to oracle/_ref/ only.  Grid size and
// stencil order are compile-time in the reference (cosmo_macros.h:12-40), so
// one library is built per (order, NX, NY, NZ).
//
// Pieces used:  utils/math.h (stencils, KO, reductions), utils/RK4Register.h,
// utils/FRW.h, components/static/static.cc (Static::addBSSNSrc).
#include "cosmo_includes.h"
#include "cosmo_types.h"
#include "cosmo_globals.h"
#include "utils/math.h"
#include "utils/FRW.h"
#include "utils/RK4Register.h"
#include "components/static/static.h"

using namespace cosmo;

// globals the reference headers declare extern (cosmo.cc:17-24 defines them)
real_t dt;
real_t dx;

static void load(arr_t &a, const double *src) {
  for (idx_t x = 0; x < POINTS; ++x) a[x] = src[x];
}

extern "C" {

int ref_info(int out[4]) {
  out[0] = NX; out[1] = NY; out[2] = NZ; out[3] = STENCIL_ORDER;
  return 0;
}

void ref_set_steps(double dx_in, double dt_in) { dx = dx_in; dt = dt_in; }

// derivative(i,j,k,d) for d=1..3 at every point: out[(d-1)*P + idx]
void ref_derivatives(const double *f, double *out) {
  arr_t a(NX, NY, NZ);
  load(a, f);
  idx_t i, j, k;
  for (int d = 1; d <= 3; ++d)
    LOOP3(i, j, k) out[(d - 1) * POINTS + NP_INDEX(i, j, k)] = derivative(i, j, k, d, a);
}

// double_derivative for (1,1),(1,2),(1,3),(2,2),(2,3),(3,3), and the swapped
// mixed arguments (2,1),(3,1),(3,2) in slots 6..8.
void ref_double_derivatives(const double *f, double *out) {
  arr_t a(NX, NY, NZ);
  load(a, f);
  static const int P[9][2] = {{1, 1}, {1, 2}, {1, 3}, {2, 2}, {2, 3}, {3, 3}, {2, 1}, {3, 1}, {3, 2}};
  idx_t i, j, k;
  for (int s = 0; s < 9; ++s)
    LOOP3(i, j, k) out[s * POINTS + NP_INDEX(i, j, k)] = double_derivative(i, j, k, P[s][0], P[s][1], a);
}

// upwind_derivative(i,j,k,d,f,c) with the speed c taken per point, d=1..3:
// out[(d-1)*P + idx]
void ref_upwind(const double *f, const double *c, double *out) {
  arr_t a(NX, NY, NZ);
  load(a, f);
  idx_t i, j, k;
  for (int d = 1; d <= 3; ++d)
    LOOP3(i, j, k) out[(d - 1) * POINTS + NP_INDEX(i, j, k)] = upwind_derivative(i, j, k, d, a, c[NP_INDEX(i, j, k)]);
}

void ref_ko(const double *f, double coeff, double *out) {
  arr_t a(NX, NY, NZ);
  load(a, f);
  idx_t i, j, k;
  LOOP3(i, j, k) out[NP_INDEX(i, j, k)] = KO_dissipation_Q(i, j, k, a, coeff);
}

// out: average, volume_average, conformal_average, min, max, numNaNs,
// conformal_standard_deviation
void ref_reductions(const double *f, const double *dphi, double phiF, double out[7]) {
  arr_t a(NX, NY, NZ), p(NX, NY, NZ);
  load(a, f);
  load(p, dphi);
  out[0] = average(a);
  out[1] = volume_average(p, phiF);
  out[2] = conformal_average(a, p, phiF);
  out[3] = cosmo::min(a);
  out[4] = cosmo::max(a);
  out[5] = (double)numNaNs(a);
  out[6] = conformal_standard_deviation(a, p, phiF);
}

// Static::addBSSNSrc on a cleared DIFFr with the given FRW (phi, rho via one
// w=0 fluid): writes DIFFr to out.
void ref_static_src(const double *DIFFphi, const double *DIFFD, double phi_frw, double rho_frw, double *out) {
  Static st;
  load(*st.fields["DIFFD_a"], DIFFD);
  arr_t DIFFr_a(NX, NY, NZ), DIFFphi_a(NX, NY, NZ);
  load(DIFFphi_a, DIFFphi);
  map_t fields;
  fields["DIFFr_a"] = &DIFFr_a;
  fields["DIFFphi_a"] = &DIFFphi_a;
  FRW<real_t> frw(phi_frw, 0.0);
  frw.addFluid(rho_frw, 0.0);
  st.addBSSNSrc(fields, &frw);
  for (idx_t x = 0; x < POINTS; ++x) out[x] = DIFFr_a[x];
}

// RK4Register driven with caller-supplied stage derivatives: x' = f(x) with
// f(x) = cos(x) * c0 + c1 * x, pointwise, for nsteps; writes the _p register
// after every step (nsteps*n) plus the stage _a inputs of the last step (4*n).
void ref_rk4(const double *x0, long n, double h, double c0, double c1, int nsteps, double *traj, double *last_stage_a) {
  RK4Register<long, double> r;
  r.init(n, 1, 1, h);
  for (long x = 0; x < n; ++x) r._array_p[x] = x0[x];
  for (int s = 0; s < nsteps; ++s) {
    r.stepInit();
    for (int st = 1; st <= 4; ++st) {
      for (long x = 0; x < n; ++x) {
        double a = r._array_a[x];
        if (s == nsteps - 1) last_stage_a[(st - 1) * n + x] = a;
        r._array_c[x] = std::cos(a) * c0 + c1 * a;
      }
      if (st == 1) r.K1Finalize();
      else if (st == 2) r.K2Finalize();
      else if (st == 3) r.K3Finalize();
      else r.K4Finalize();
    }
    for (long x = 0; x < n; ++x) traj[(long)s * n + x] = r._array_p[x];
  }
}

// FRW<double>: fluids (rho_n, w_n), nsteps of P1,P2,P3,RK_total with step h.
// Writes (phi, K, rho, S) after every stage: out[(4*s + stage)*4 + q].
void ref_frw(double phi0, double K0, int nfl, const double *rho, const double *w, double h, int nsteps, double *out) {
  FRW<double> frw(phi0, K0);
  for (int n = 0; n < nfl; ++n) frw.addFluid(rho[n], w[n]);
  for (int s = 0; s < nsteps; ++s)
    for (int st = 0; st < 4; ++st) {
      if (st == 0) frw.P1_step(h);
      else if (st == 1) frw.P2_step(h);
      else if (st == 2) frw.P3_step(h);
      else frw.RK_total_step(h);
      double *o = out + ((long)4 * s + st) * 4;
      o[0] = frw.get_phi(); o[1] = frw.get_K(); o[2] = frw.get_rho(); o[3] = frw.get_S();
    }
}

}  // extern "C"
