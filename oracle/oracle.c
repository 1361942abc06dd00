/*
 * oracle.c -- CPU restatement of CosmoGRaPH's BSSN hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Built with
 *   gcc -O2 -fopenmp -ffp-contract=off -fno-fast-math
 * so that every expression rounds exactly as written.  Expressions follow
 * the association order of the reference line they cite, so that pieces
 * which the reference lets us build from source (stencils, RK4 registers,
 * FRW, the static-dust source) agree bitwise.
 *
 * Parity unpinned (no check beyond the reference's own formula, since
 * BSSNGaugeHandler.cc and bssn.cc do not build here): the Anharmonic,
 * ConformalSync, TestKDriverLapse and TestAijDriverLapse lapses, the mu_L
 * and p terms of the DampedWave lapse, the mu_S term of the DampedWave
 * shift, the AijDriverShift, and the scale (normalisation) arrays of the
 * H, M, G, A, S constraint statistics.  The tests fix their wiring (which
 * K, phi and K_avg each stage sees; homogeneous closed forms) but restate
 * the formula itself.
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define PI (4.0 * atan(1.0)) /* cosmo_macros.h:107 */
#define PW2(x) ((x) * (x))
#define HALO 8

/* symmetric index map (a,b) -> 0..5 : 11 12 13 22 23 33 */
static const int SY[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
/* field slots, reference order bssn_macros.h:77-99 */
enum {
  F_G11 = 0, F_PHI = 6, F_A11 = 7, F_K = 13, F_GAM1 = 14, F_ALPHA = 17
};
static const char *FIELD_NAMES[OR_NF] = {
    "DIFFgamma11", "DIFFgamma12", "DIFFgamma13", "DIFFgamma22", "DIFFgamma23",
    "DIFFgamma33", "DIFFphi",     "A11",         "A12",         "A13",
    "A22",         "A23",         "A33",         "DIFFK",       "Gamma1",
    "Gamma2",      "Gamma3",      "DIFFalpha"};
static const char *SRC_NAMES[OR_NSRC] = {"DIFFr", "DIFFS", "S1",    "S2",
                                         "S3",    "STF11", "STF12", "STF13",
                                         "STF22", "STF23", "STF33"};
static const char *DUST_NAMES[OR_NDUST] = {"D", "S1", "S2", "S3"};
/* shift-build fields, bssn_macros.h:18-47 (order of BSSN_APPLY_TO_FIELDS) */
static const char *X_NAMES[OR_NX] = {"beta1", "beta2", "beta3", "expN", "auxB1", "auxB2", "auxB3"};
/* extra fields a build carries: none by default, beta and expN with
 * USE_BSSN_SHIFT, plus auxB with USE_GAMMA_DRIVER */
static int n_extra(const or_params *p) {
  return p->shift == OR_SHIFT_NONE ? 0 : p->shift == OR_SHIFT_GAMMA_DRIVER ? OR_NX : OR_X_AUXB1;
}
/* Dust derived arrays, components/dust_fluid/dust.h:37-46 */
enum {
  DD_ADV1 = 0, DD_AS1V1 = 3, DD_S1SRC = 12, DD_DETG = 15, DD_G11 = 16,
  DD_W = 22, DD_N = 23
};

/* ---------------------------------------------------------------- FRW --- */
/* utils/FRW.h:11-254, RT = double.  PI_L (FRW.h:4) is a long double literal,
 * so the K source term is formed in long double and rounded on accumulation. */
#define PI_L 3.141592653589793238462643383279502884L
#define OR_MAX_FLUIDS 8
typedef struct {
  double phi, K, alpha;
  int nf;
  double rho[OR_MAX_FLUIDS], w[OR_MAX_FLUIDS];
  double phi_K1, K_K1, phi_K2, K_K2, phi_K3, K_K3, phi_K4, K_K4;
  double fK1[OR_MAX_FLUIDS], fK2[OR_MAX_FLUIDS], fK3[OR_MAX_FLUIDS],
      fK4[OR_MAX_FLUIDS];
  double phi_get, K_get, rho_get, S_get;
} or_frw;

static void frw_sum_get(or_frw *F, const double *rhos) {
  F->rho_get = 0.0;
  F->S_get = 0.0;
  for (int n = 0; n < F->nf; ++n) {
    double rho = rhos[n];
    F->rho_get += rho;
    F->S_get += 3.0 * rho * F->w[n];
  }
}

/* FRW.h:93-129 (P1), 131-169 (P2), 171-209 (P3), 211-254 (RK_total) */
static void frw_stage(or_frw *F, int stage, double h) {
  double rh[OR_MAX_FLUIDS];
  double K_in, frac;
  const double *prevK;
  double *curK;
  double *phiK, *KK;
  switch (stage) {
  case 1: K_in = F->K; prevK = NULL; curK = F->fK1; phiK = &F->phi_K1; KK = &F->K_K1; frac = 0; break;
  case 2: K_in = F->K + (h / 2.0) * F->K_K1; prevK = F->fK1; curK = F->fK2; phiK = &F->phi_K2; KK = &F->K_K2; frac = h / 2.0; break;
  case 3: K_in = F->K + (h / 2.0) * F->K_K2; prevK = F->fK2; curK = F->fK3; phiK = &F->phi_K3; KK = &F->K_K3; frac = h / 2.0; break;
  default: K_in = F->K + h * F->K_K3; prevK = F->fK3; curK = F->fK4; phiK = &F->phi_K4; KK = &F->K_K4; frac = h; break;
  }
  *phiK = -1.0 * K_in / 6.0;
  double K_source = 0.0;
  for (int n = 0; n < F->nf; ++n) {
    double rho_in = prevK ? F->rho[n] + frac * prevK[n] : F->rho[n];
    rh[n] = rho_in;
    K_source += 4.0 * PI_L * rho_in * (1.0 + 3.0 * F->w[n]);
  }
  *KK = 1.0 / 3.0 * K_in * K_in + K_source;
  for (int n = 0; n < F->nf; ++n) curK[n] = K_in * rh[n] * (1.0 + F->w[n]);

  double rget[OR_MAX_FLUIDS];
  if (stage == 1 || stage == 2) {
    F->phi_get = F->phi + (h / 2.0) * *phiK;
    F->K_get = F->K + (h / 2.0) * *KK;
    for (int n = 0; n < F->nf; ++n) rget[n] = F->rho[n] + (h / 2.0) * curK[n];
  } else if (stage == 3) {
    F->phi_get = F->phi + h * *phiK;
    F->K_get = F->K + h * *KK;
    for (int n = 0; n < F->nf; ++n) rget[n] = F->rho[n] + h * curK[n];
  } else {
    F->phi += h / 6.0 * (F->phi_K1 + 2.0 * F->phi_K2 + 2.0 * F->phi_K3 + F->phi_K4);
    F->K += h / 6.0 * (F->K_K1 + 2.0 * F->K_K2 + 2.0 * F->K_K3 + F->K_K4);
    for (int n = 0; n < F->nf; ++n)
      F->rho[n] += h / 6.0 * (F->fK1[n] + 2.0 * F->fK2[n] + 2.0 * F->fK3[n] + F->fK4[n]);
    F->phi_get = F->phi;
    F->K_get = F->K;
    for (int n = 0; n < F->nf; ++n) rget[n] = F->rho[n];
  }
  frw_sum_get(F, rget);
}

/* ---------------------------------------------------------------- sim --- */
struct or_sim {
  or_params p;
  long npts;
  long *xo, *yo, *zo; /* wrapped offsets, index range [-HALO, n+HALO) */
  double *reg[OR_NF][4];
  double *xreg[OR_NX][4]; /* shift-build registers (n_extra of them allocated) */
  double *src[OR_NSRC];
  double *ricci, *AijAij, *K0;
  double *DIFFD;
  double *dust[OR_NDUST][4];
  double *dd[DD_N];
  or_frw frw;
  double K_min, K_avg, avg_vol, rho_avg;
};

static double *zalloc(long n) { return (double *)calloc((size_t)n, sizeof(double)); }

static long *mk_wrap(int n, long stride) {
  long *w = (long *)malloc(sizeof(long) * (size_t)(n + 2 * HALO));
  for (int t = -HALO; t < n + HALO; ++t) {
    /* INDEX macro, cosmo_macros.h:116: ((i+4*N)%N) */
    long m = ((long)t + 4L * n) % n;
    w[t + HALO] = m * stride;
  }
  return w + HALO;
}

or_sim *or_create(const or_params *p) {
  or_sim *s = (or_sim *)calloc(1, sizeof(or_sim));
  s->p = *p;
  s->npts = (long)p->nx * p->ny * p->nz;
  s->xo = mk_wrap(p->nx, (long)p->ny * p->nz);
  s->yo = mk_wrap(p->ny, p->nz);
  s->zo = mk_wrap(p->nz, 1);
  for (int f = 0; f < OR_NF; ++f)
    for (int r = 0; r < 4; ++r) s->reg[f][r] = zalloc(s->npts);
  for (int f = 0; f < n_extra(p); ++f)
    for (int r = 0; r < 4; ++r) s->xreg[f][r] = zalloc(s->npts);
  for (int f = 0; f < OR_NSRC; ++f) s->src[f] = zalloc(s->npts);
  s->ricci = zalloc(s->npts);
  s->AijAij = zalloc(s->npts);
  s->K0 = zalloc(s->npts);
  s->DIFFD = zalloc(s->npts);
  if (p->matter == OR_MATTER_DUST) {
    for (int f = 0; f < OR_NDUST; ++f)
      for (int r = 0; r < 4; ++r) s->dust[f][r] = zalloc(s->npts);
    for (int f = 0; f < DD_N; ++f) s->dd[f] = zalloc(s->npts);
  }
  /* FRW(0,0): FRW.h:37-45 */
  s->frw.alpha = 1.0;
  return s;
}

void or_destroy(or_sim *s) {
  if (!s) return;
  for (int f = 0; f < OR_NF; ++f)
    for (int r = 0; r < 4; ++r) free(s->reg[f][r]);
  for (int f = 0; f < OR_NX; ++f)
    for (int r = 0; r < 4; ++r) free(s->xreg[f][r]);
  for (int f = 0; f < OR_NSRC; ++f) free(s->src[f]);
  free(s->ricci); free(s->AijAij); free(s->K0); free(s->DIFFD);
  for (int f = 0; f < OR_NDUST; ++f)
    for (int r = 0; r < 4; ++r) free(s->dust[f][r]);
  for (int f = 0; f < DD_N; ++f) free(s->dd[f]);
  free(s->xo - HALO); free(s->yo - HALO); free(s->zo - HALO);
  free(s);
}

long or_points(const or_sim *s) { return s->npts; }

#ifdef _OPENMP
#include <omp.h>
#endif
int or_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Reference field-map keys (cosmo_macros.h:154-187): "<field>_<p|a|c|f>" for
 * RK4 registers, "<src>_a" for sources/extras, "DIFFD_a" for static dust.
 * The dust-fluid registers (dust.h:32-35) have no map in the reference; they
 * are addressed here as "dust_D_<r>", "dust_S1_<r>", ... */
double *or_field(or_sim *s, const char *name) {
  static const char REGS[4] = {'p', 'a', 'c', 'f'};
  size_t L = strlen(name);
  if (L < 3 || L > 60 || name[L - 2] != '_') return NULL;
  char base[64];
  memcpy(base, name, L - 2);
  base[L - 2] = 0;
  int r = -1;
  for (int t = 0; t < 4; ++t)
    if (REGS[t] == name[L - 1]) r = t;
  if (r < 0) return NULL;
  if (!strncmp(base, "dust_", 5)) {
    if (s->p.matter != OR_MATTER_DUST) return NULL;
    for (int f = 0; f < OR_NDUST; ++f)
      if (!strcmp(base + 5, DUST_NAMES[f])) return s->dust[f][r];
    return NULL;
  }
  for (int f = 0; f < OR_NF; ++f)
    if (!strcmp(base, FIELD_NAMES[f])) return s->reg[f][r];
  for (int f = 0; f < n_extra(&s->p); ++f)
    if (!strcmp(base, X_NAMES[f])) return s->xreg[f][r];
  if (r != OR_A) return NULL;
  for (int f = 0; f < OR_NSRC; ++f)
    if (!strcmp(base, SRC_NAMES[f])) return s->src[f];
  if (!strcmp(base, "ricci")) return s->ricci;
  if (!strcmp(base, "AijAij")) return s->AijAij;
  if (!strcmp(base, "K0")) return s->K0;
  if (!strcmp(base, "DIFFD")) return s->DIFFD;
  return NULL;
}

/* ----------------------------------------------------------- stencils --- */
/* utils/math.h; `/dx` applied last exactly as there. */
#define FV(f, a, b, c) ((f)[s->xo[(a)] + s->yo[(b)] + s->zo[(c)]])

/* sample f at offset t along direction d (1..3) */
static inline double at1(const or_sim *s, const double *f, long i, long j,
                         long k, int d, int t) {
  return d == 1 ? FV(f, i + t, j, k) : d == 2 ? FV(f, i, j + t, k) : FV(f, i, j, k + t);
}
/* sample f at offsets (t1 along d1, t2 along d2) */
static inline double at2(const or_sim *s, const double *f, long i, long j,
                         long k, int d1, int t1, int d2, int t2) {
  long a = i, b = j, c = k;
  if (d1 == 1) a += t1; else if (d1 == 2) b += t1; else c += t1;
  if (d2 == 1) a += t2; else if (d2 == 2) b += t2; else c += t2;
  return FV(f, a, b, c);
}

static double deriv(const or_sim *s, const double *f, long i, long j, long k, int d) {
  const double dx = s->p.dx;
#define P(t) at1(s, f, i, j, k, d, (t))
  switch (s->p.order) {
  case 2: /* math.h:74-100 */
    return ((-1.0 / 2.0 * P(-1) + 1.0 / 2.0 * P(1)) / dx);
  case 4: /* math.h:231-263 */
    return (+1.0 / 12.0 * P(-2) - 2.0 / 3.0 * P(-1) + 2.0 / 3.0 * P(1) - 1.0 / 12.0 * P(2)) / dx;
  case 6: /* math.h:412-450 */
    return (-1.0 / 60.0 * P(-3) + 3.0 / 20.0 * P(-2) - 3.0 / 4.0 * P(-1) + 3.0 / 4.0 * P(1) -
            3.0 / 20.0 * P(2) + 1.0 / 60.0 * P(3)) / dx;
  default: /* 8: math.h:554-586 */
    return ((1.0 / 280.0 * P(-4) - 1.0 / 280.0 * P(4)) - (4.0 / 105.0 * P(-3) - 4.0 / 105.0 * P(3)) +
            (1.0 / 5.0 * P(-2) - 1.0 / 5.0 * P(2)) - (4.0 / 5.0 * P(-1) - 4.0 / 5.0 * P(1))) / dx;
  }
#undef P
}

static double dderiv_pure(const or_sim *s, const double *f, long i, long j, long k, int d) {
  const double dx = s->p.dx;
#define P(t) at1(s, f, i, j, k, d, (t))
  switch (s->p.order) {
  case 2: /* math.h:873-901 */
    return (P(-1) - 2.0 * P(0) + P(1)) / dx / dx;
  case 4: /* math.h:958-993 */
    return (-1.0 / 12.0 * P(-2) + 4.0 / 3.0 * P(-1) - 5.0 / 2.0 * P(0) + 4.0 / 3.0 * P(1) -
            1.0 / 12.0 * P(2)) / dx / dx;
  case 6: /* math.h:995-1036 */
    return (1.0 / 90.0 * P(-3) - 3.0 / 20.0 * P(-2) + 3.0 / 2.0 * P(-1) - 49.0 / 18.0 * P(0) +
            3.0 / 2.0 * P(1) - 3.0 / 20.0 * P(2) + 1.0 / 90.0 * P(3)) / dx / dx;
  default: /* math.h:1038-1085 */
    return (-1.0 / 560.0 * P(-4) + 8.0 / 315.0 * P(-3) - 1.0 / 5.0 * P(-2) + 8.0 / 5.0 * P(-1) -
            205.0 / 72.0 * P(0) + 8.0 / 5.0 * P(1) - 1.0 / 5.0 * P(2) + 8.0 / 315.0 * P(3) -
            1.0 / 560.0 * P(4)) / dx / dx;
  }
#undef P
}

/* diagonal-only mixed stencils, math.h:676-855.  (a,b) = (lower, higher)
 * direction; the four-point group at distance m is
 * -f(+m,-m) + f(+m,+m) + f(-m,-m) - f(-m,+m). */
static double dderiv_mixed(const or_sim *s, const double *f, long i, long j, long k, int d1, int d2) {
  const double dx = s->p.dx;
  int a = d1 < d2 ? d1 : d2, b = d1 < d2 ? d2 : d1;
#define Q(m) (-at2(s, f, i, j, k, a, (m), b, -(m)) + at2(s, f, i, j, k, a, (m), b, (m)) + \
              at2(s, f, i, j, k, a, -(m), b, -(m)) - at2(s, f, i, j, k, a, -(m), b, (m)))
  switch (s->p.order) {
  case 2: return (Q(1)) / 4.0 / dx / dx;
  case 4: return ((Q(1)) - 1.0 / 16.0 * (Q(2))) / 3.0 / dx / dx;
  case 6: return (135.0 * (Q(1)) - 27.0 / 2.0 * (Q(2)) + (Q(3))) / 360.0 / dx / dx;
  default: return (2.0 / 5.0 * (Q(1)) - 1.0 / 20.0 * (Q(2)) + 2.0 / 315.0 * (Q(3)) - 1.0 / 2240.0 * (Q(4))) / dx / dx;
  }
#undef Q
}

/* math.h:1319-1330 */
static double dderiv(const or_sim *s, const double *f, long i, long j, long k, int d1, int d2) {
  if (d1 == d2) return dderiv_pure(s, f, i, j, k, d1);
  return dderiv_mixed(s, f, i, j, k, d1, d2);
}

/* KO_dissipation_Q, math.h:36-72: orders 2 and 8 only, 0 otherwise. */
static double ko_q(const or_sim *s, const double *f, long i, long j, long k) {
  const double c = s->p.ko_coeff, dx = s->p.dx;
  if (c == 0) return 0;
#define X(t) FV(f, i + (t), j, k)
#define Y(t) FV(f, i, j + (t), k)
#define Z(t) FV(f, i, j, k + (t))
  if (s->p.order == 2) {
    double st = (1.0 * X(-2) + 1.0 * Y(-2) + 1.0 * Z(-2) - 4.0 * X(-1) - 4.0 * Y(-1) - 4.0 * Z(-1) +
                 6.0 * X(0) + 6.0 * Y(0) + 6.0 * Z(0) - 4.0 * X(1) - 4.0 * Y(1) - 4.0 * Z(1) +
                 1.0 * X(2) + 1.0 * Y(2) + 1.0 * Z(2)) / pow(dx, 4.0);
    return c * pow(dx, 3.0) / 64.0 * st;
  }
  if (s->p.order == 8) {
    double st = (1.0 * X(-5) + 1.0 * Y(-5) + 1.0 * Z(-5) - 10.0 * X(-4) - 10.0 * Y(-4) - 10.0 * Z(-4) +
                 45.0 * X(-3) + 45.0 * Y(-3) + 45.0 * Z(-3) - 120.0 * X(-2) - 120.0 * Y(-2) - 120.0 * Z(-2) +
                 210.0 * X(-1) + 210.0 * Y(-1) + 210.0 * Z(-1) - 252.0 * X(0) - 252.0 * Y(0) - 252.0 * Z(0) +
                 210.0 * X(1) + 210.0 * Y(1) + 210.0 * Z(1) - 120.0 * X(2) - 120.0 * Y(2) - 120.0 * Z(2) +
                 45.0 * X(3) + 45.0 * Y(3) + 45.0 * Z(3) - 10.0 * X(4) - 10.0 * Y(4) - 10.0 * Z(4) +
                 1.0 * X(5) + 1.0 * Y(5) + 1.0 * Z(5)) / pow(dx, 10.0);
    return -c * pow(dx, 9.0) / 1024.0 * st;
  }
#undef X
#undef Y
#undef Z
  return 0.0;
}

/* lop_forward_derivative / lop_backward_derivative (math.h:1220-1232): the
 * lopsided first derivative of the stencil order plus its forward/backward
 * dissipation stencil.  Order 2: math.h:166-230 and 1087-1130 (the second
 * derivative stencils 888-956); order 4: math.h:265-337 (no dissipation,
 * 1132-1146); orders 6 and 8 are stubs that return 0 (math.h:539-549,
 * 661-673, 1148-1178). */
static double lop(const or_sim *s, const double *f, long i, long j, long k, int d, int fwd) {
  const double dx = s->p.dx;
#define P(t) at1(s, f, i, j, k, d, (t))
  if (s->p.order == 2) {
    if (fwd)
      return ((-1.0 / 2.0 * P(2) + 2.0 * P(1) - 3.0 / 2.0 * P(0)) / dx) +
             -1.0 / 2.0 * dx * ((2.0 * P(0) - 5.0 * P(1) + 4.0 * P(2) - 1.0 * P(3)) / dx / dx);
    return ((+1.0 / 2.0 * P(-2) - 2.0 * P(-1) + 3.0 / 2.0 * P(0)) / dx) +
           +1.0 / 2.0 * dx * ((2.0 * P(0) - 5.0 * P(-1) + 4.0 * P(-2) - 1.0 * P(-3)) / dx / dx);
  }
  if (s->p.order == 4) {
    if (fwd)
      return (+1.0 / 12.0 * P(3) - 1.0 / 2.0 * P(2) + 3.0 / 2.0 * P(1) - 5.0 / 6.0 * P(0) - 1.0 / 4.0 * P(-1)) / dx +
             0.0;
    return (-1.0 / 12.0 * P(-3) + 1.0 / 2.0 * P(-2) - 3.0 / 2.0 * P(-1) + 5.0 / 6.0 * P(0) + 1.0 / 4.0 * P(1)) / dx +
           0.0;
  }
  return 0.0 + 0.0;
#undef P
}

/* upwind_derivative, math.h:1247-1252 */
static double upwind(const or_sim *s, const double *f, long i, long j, long k, int d, double c) {
  if (c > 0) return c * lop(s, f, i, j, k, d, 1);
  return c * lop(s, f, i, j, k, d, 0);
}

void or_upwind(or_sim *s, const double *f, const double *c, double *out) {
  const long P = s->npts;
  for (int d = 1; d <= 3; ++d)
    for (long i = 0; i < s->p.nx; ++i)
      for (long j = 0; j < s->p.ny; ++j)
        for (long k = 0; k < s->p.nz; ++k) {
          long idx = (i * s->p.ny + j) * s->p.nz + k;
          out[(d - 1) * P + idx] = upwind(s, f, i, j, k, d, c[idx]);
        }
}

double or_stencil(or_sim *s, const double *f, int kind, int d1, int d2, long i, long j, long k) {
  switch (kind) {
  case 0: return deriv(s, f, i, j, k, d1);
  case 1: return dderiv(s, f, i, j, k, d1, d2);
  default: return ko_q(s, f, i, j, k);
  }
}

/* ------------------------------------------------------- point algebra --- */
/* Restates BSSN::set_bd_values (bssn.cc:361-421) after the stencil loads,
 * then the ev_* functions (bssn.cc:748-891).  theta = 0 (Z4c compiled out by
 * default, cosmo_macros.h:88-90).  beta = 0 unless p->shift selects the
 * USE_BSSN_SHIFT build (cosmo_macros.h:43-44): then the shift terms of
 * bssn_macros.h:292-385 and bssn.cc:786-828 are added, with the advection
 * terms (upwind derivatives) supplied in bd->adv, and the extra fields'
 * right-hand sides (bssn.cc:852-891) are formed in bd->rhsx. */
void or_point_compute(const or_params *p, or_point *bd) {
  const double *v = bd->v;
  const double theta = 0.0;
  /* bssn.cc:384-394 */
  bd->phi = v[F_PHI] + bd->phi_F;
  bd->K = v[F_K] + bd->K_F;
  for (int a = 0; a < 6; ++a) bd->g[a] = v[a] + ((a == 0 || a == 3 || a == 5) ? 1.0 : 0.0);
  bd->r = bd->src[0] + bd->rho_F;
  bd->S = bd->src[1] + bd->S_F;
  bd->alpha = v[F_ALPHA] + 1.0;
  /* set_gammai_values, bssn.cc:447-455 (unit-det inverse from DIFF vars) */
  {
    double g11 = v[0], g12 = v[1], g13 = v[2], g22 = v[3], g23 = v[4], g33 = v[5];
    bd->gi[0] = 1.0 + g22 + g33 - PW2(g23) + g22 * g33;
    bd->gi[3] = 1.0 + g11 + g33 - PW2(g13) + g11 * g33;
    bd->gi[5] = 1.0 + g11 + g22 - PW2(g12) + g11 * g22;
    bd->gi[1] = g13 * g23 - g12 * (1.0 + g33);
    bd->gi[2] = g12 * g23 - g13 * (1.0 + g22);
    bd->gi[4] = g12 * g13 - g23 * (1.0 + g11);
  }
  const double *gi = bd->gi, *g = bd->g;
#define GI(a, b) gi[SY[a][b]]
#define GG(a, b) g[SY[a][b]]
#define AA(a, b) v[F_A11 + SY[a][b]]
  /* calculate_Acont, bssn.cc:463-472, BSSN_CALCULATE_ACONT bssn_macros.h:233-237 */
  for (int I = 0; I < 3; ++I)
    for (int J = I; J < 3; ++J) {
      double acc = 0.0;
      int first = 1;
      for (int b = 0; b < 3; ++b)
        for (int a = 0; a < 3; ++a) {
          double t = GI(I, a) * GI(J, b) * AA(a, b);
          acc = first ? t : acc + t;
          first = 0;
        }
      bd->Acont[SY[I][J]] = acc;
    }
  bd->AijAij = bd->Acont[0] * v[F_A11 + 0] + bd->Acont[3] * v[F_A11 + 3] + bd->Acont[5] * v[F_A11 + 5] +
               2.0 * (bd->Acont[1] * v[F_A11 + 1] + bd->Acont[2] * v[F_A11 + 2] + bd->Acont[4] * v[F_A11 + 4]);

  /* calculate_conformal_christoffels, bssn.cc:577-590,
   * bssn_macros.h:221-229 */
#define DG(k, a, b) bd->dg[k][SY[a][b]]
  for (int I = 0; I < 3; ++I)
    for (int J = 0; J < 3; ++J)
      for (int K = J; K < 3; ++K) {
        double t0 = GI(I, 0) * (DG(J, K, 0) + DG(K, J, 0) - DG(0, J, K));
        double t1 = GI(I, 1) * (DG(J, K, 1) + DG(K, J, 1) - DG(1, J, K));
        double t2 = GI(I, 2) * (DG(J, K, 2) + DG(K, J, 2) - DG(2, J, K));
        bd->G[I][SY[J][K]] = 0.5 * (t0 + t1 + t2);
        bd->GL[I][SY[J][K]] = 0.5 * (DG(J, K, I) + DG(K, J, I) - DG(I, J, K));
      }
  for (int I = 0; I < 3; ++I) {
    const double *Gi = bd->G[I];
    bd->Gammad[I] = Gi[0] * gi[0] + Gi[3] * gi[3] + Gi[5] * gi[5] +
                    2.0 * (Gi[1] * gi[1] + Gi[2] * gi[2] + Gi[4] * gi[4]);
  }
#define GU(i, a, b) bd->G[i][SY[a][b]]
#define GLo(i, a, b) bd->GL[i][SY[a][b]]
  /* calculateDDphi, bssn.cc:592-602 */
  for (int I = 0; I < 3; ++I)
    for (int J = I; J < 3; ++J)
      bd->DDphi[SY[I][J]] = bd->ddphi[SY[I][J]] -
                            (GU(0, I, J) * bd->dphi[0] + GU(1, I, J) * bd->dphi[1] + GU(2, I, J) * bd->dphi[2]);

  /* calculateDDalphaTF, bssn.cc:604-626, BSSN_CALCULATE_DIDJALPHA
   * bssn_macros.h:241-245 */
  {
    double gild[3];
    for (int I = 0; I < 3; ++I)
      gild[I] = GI(I, 0) * bd->dphi[0] + GI(I, 1) * bd->dphi[1] + GI(I, 2) * bd->dphi[2];
    for (int I = 0; I < 3; ++I)
      for (int J = I; J < 3; ++J) {
        double acc = 0.0;
        for (int m = 0; m < 3; ++m) {
          double t = (GU(m, I, J) + 2.0 * ((double)(m == I) * bd->dphi[J] + (double)(m == J) * bd->dphi[I] -
                                           GG(I, J) * gild[m])) * bd->dalpha[m];
          acc = m == 0 ? t : acc + t;
        }
        bd->DDaTF[SY[I][J]] = bd->ddalpha[SY[I][J]] - (acc);
      }
    double *D = bd->DDaTF;
    bd->DDaTR = gi[0] * D[0] + gi[3] * D[3] + gi[5] * D[5] + 2.0 * (gi[1] * D[1] + gi[2] * D[2] + gi[4] * D[4]);
    for (int a = 0; a < 6; ++a) D[a] -= (1.0 / 3.0) * g[a] * bd->DDaTR;
    bd->DDaTR *= exp(-4.0 * bd->phi);
  }

  /* calculateRicciTF, bssn.cc:629-674; BSSN_CALCULATE_RICCI_UNITARY
   * bssn_macros.h:248-277 */
  for (int I = 0; I < 3; ++I)
    for (int J = I; J < 3; ++J) {
      double t1 = 0.0, t2a = 0.0, t2b = 0.0, t3a = 0.0, t3b = 0.0, t4 = 0.0;
      int first = 1;
      for (int K = 0; K < 3; ++K)
        for (int L = 0; L < 3; ++L) {
          double t = GI(K, L) * bd->ddg[SY[K][L]][SY[I][J]];
          t1 = first ? t : t1 + t;
          first = 0;
        }
      for (int K = 0; K < 3; ++K) {
        double ta = GG(K, I) * bd->dGam[J][K];
        double tb = GG(K, J) * bd->dGam[I][K];
        double tc = bd->Gammad[K] * GLo(I, J, K);
        double td = bd->Gammad[K] * GLo(J, I, K);
        t2a = K ? t2a + ta : ta;
        t2b = K ? t2b + tb : tb;
        t3a = K ? t3a + tc : tc;
        t3b = K ? t3b + td : td;
      }
      first = 1;
      for (int K = 0; K < 3; ++K)
        for (int L = 0; L < 3; ++L)
          for (int M = 0; M < 3; ++M) {
            double t = GI(L, M) * (GU(K, L, I) * GLo(J, K, M) + GU(K, L, J) * GLo(I, K, M) +
                                   0.5 * (GU(K, I, M) * GLo(K, L, J) + GU(K, J, M) * GLo(K, L, I)));
            t4 = first ? t : t4 + t;
            first = 0;
          }
      bd->Uricci[SY[I][J]] = (-0.5 * (t1) + 0.5 * (t2a + t2b) + 0.5 * (t3a + t3b) + t4);
    }
  {
    const double *dp = bd->dphi, *DD = bd->DDphi;
    double expression =
        (gi[0] * (DD[0] + 2.0 * dp[0] * dp[0]) + gi[3] * (DD[3] + 2.0 * dp[1] * dp[1]) +
         gi[5] * (DD[5] + 2.0 * dp[2] * dp[2]) +
         2.0 * (gi[1] * (DD[1] + 2.0 * dp[0] * dp[1]) + gi[2] * (DD[2] + 2.0 * dp[0] * dp[2]) +
                gi[4] * (DD[4] + 2.0 * dp[1] * dp[2])));
    for (int I = 0; I < 3; ++I)
      for (int J = I; J < 3; ++J) {
        int a = SY[I][J];
        bd->ricci_ij[a] = bd->Uricci[a] - 2.0 * (DD[a] - 2.0 * dp[I] * dp[J] + g[a] * (expression));
      }
    const double *R = bd->ricci_ij;
    bd->ricci = exp(-4.0 * bd->phi) *
                (R[0] * gi[0] + R[3] * gi[3] + R[5] * gi[5] + 2.0 * (R[1] * gi[1] + R[2] * gi[2] + R[4] * gi[4]));
    for (int a = 0; a < 6; ++a)
      bd->ricciTF[a] = R[a] - (1.0 / 3.0) * exp(4.0 * bd->phi) * g[a] * bd->ricci;
  }
  /* hamiltonianConstraintCalc, bssn.cc:1073-1084 */
  bd->H = -exp(5.0 * bd->phi) / 8.0 *
          (bd->ricci + 2.0 / 3.0 * PW2(bd->K + 2.0 * theta) - bd->AijAij - 16.0 * PI * bd->r);

  /* ---- ev_* : bssn.cc:748-831, bssn_macros.h:304-385 ---- */
  const double dt = p->dt, alpha = bd->alpha, H = bd->H;
  const int shift = p->shift != OR_SHIFT_NONE;
  const double *b = bd->xv; /* beta^1..3 */
#define DB(j, i) bd->dbeta[j][i]
  const double divb = DB(0, 0) + DB(1, 1) + DB(2, 2);
  /* BSSN_DT_DIFFGAMMAIJ, bssn_macros.h:292-307 */
  for (int I = 0; I < 3; ++I)
    for (int J = I; J < 3; ++J) {
      int a = SY[I][J];
      double dtg = -2.0 * alpha * v[F_A11 + a];
      if (shift)
        dtg = -2.0 * alpha * v[F_A11 + a] + bd->adv[a] + GG(I, 0) * DB(J, 0) + GG(I, 1) * DB(J, 1) +
              GG(I, 2) * DB(J, 2) + GG(J, 0) * DB(I, 0) + GG(J, 1) * DB(I, 1) + GG(J, 2) * DB(I, 2) -
              (2.0 / 3.0) * GG(I, J) * (divb);
      bd->rhs[a] = (dtg) + 0.5 * p->a_adj_amp * dt * H * v[a] - bd->ko[a];
    }
  for (int I = 0; I < 3; ++I)
    for (int J = I; J < 3; ++J) {
      int a = SY[I][J];
      double AA2 = 0.0;
      int first = 1;
      for (int m = 0; m < 3; ++m)
        for (int n = 0; n < 3; ++n) {
          double t = GI(m, n) * AA(m, I) * AA(n, J);
          AA2 = first ? t : AA2 + t;
          first = 0;
        }
      double dtA = (exp(-4.0 * bd->phi) * (alpha * (bd->ricciTF[a] - 8.0 * PI * bd->src[5 + a]) - bd->DDaTF[a]) +
                    alpha * ((bd->K + 2.0 * theta) * v[F_A11 + a] - 2.0 * (AA2)));
      if (shift) /* bssn_macros.h:319-333 */
        dtA = dtA + bd->adv[F_A11 + a] + AA(I, 0) * DB(J, 0) + AA(I, 1) * DB(J, 1) + AA(I, 2) * DB(J, 2) +
              AA(J, 0) * DB(I, 0) + AA(J, 1) * DB(I, 1) + AA(J, 2) * DB(I, 2) -
              (2.0 / 3.0) * AA(I, J) * (divb);
      bd->rhs[F_A11 + a] = dtA - 1.0 * p->a_adj_amp * dt * v[F_A11 + a] * H - bd->ko[F_A11 + a];
    }
  for (int I = 0; I < 3; ++I) {
    const double *Ac = bd->Acont;
    double second = GU(I, 0, 0) * Ac[0] + GU(I, 1, 1) * Ac[3] + GU(I, 2, 2) * Ac[5] +
                    2.0 * (GU(I, 0, 1) * Ac[1] + GU(I, 0, 2) * Ac[2] + GU(I, 1, 2) * Ac[4]) +
                    6.0 * (Ac[SY[I][0]] * bd->dphi[0] + Ac[SY[I][1]] * bd->dphi[1] + Ac[SY[I][2]] * bd->dphi[2]);
    double noshift =
        (-2.0 * (Ac[SY[I][0]] * bd->dalpha[0] + Ac[SY[I][1]] * bd->dalpha[1] + Ac[SY[I][2]] * bd->dalpha[2]) +
         2.0 * alpha *
             (+second -
              (1.0 / 3.0) * (2.0 * (GI(I, 0) * bd->dK[0] + GI(I, 1) * bd->dK[1] + GI(I, 2) * bd->dK[2]) + 0.0) -
              2.0 * alpha * 0.0 * (v[F_GAM1 + I] - bd->Gammad[I]) -
              8.0 * PI * (GI(I, 0) * bd->src[2] + GI(I, 1) * bd->src[3] + GI(I, 2) * bd->src[4])));
    double shiftG = 0.0;
    if (shift) { /* BSSN_DT_GAMMAI_SHIFT, bssn_macros.h:363-382 */
      const double(*ddb)[6] = bd->ddbeta;
#define DDB(k, l, m) ddb[m][SY[k][l]] /* d_k d_l beta^m */
      shiftG = bd->adv[F_GAM1 + I] - bd->Gammad[0] * DB(0, I) - bd->Gammad[1] * DB(1, I) -
               bd->Gammad[2] * DB(2, I) + (2.0 / 3.0) * bd->Gammad[I] * (divb) +
               (1.0 / 3.0) * (GI(I, 0) * DDB(0, 0, 0) + GI(I, 0) * DDB(1, 0, 1) + GI(I, 0) * DDB(2, 0, 2) +
                              GI(I, 1) * DDB(0, 1, 0) + GI(I, 1) * DDB(1, 1, 1) + GI(I, 1) * DDB(2, 1, 2) +
                              GI(I, 2) * DDB(0, 2, 0) + GI(I, 2) * DDB(1, 2, 1) + GI(I, 2) * DDB(2, 2, 2)) +
               (GI(0, 0) * DDB(0, 0, I) + GI(1, 1) * DDB(1, 1, I) + GI(2, 2) * DDB(2, 2, I) +
                2.0 * (GI(0, 1) * DDB(0, 1, I) + GI(0, 2) * DDB(0, 2, I) + GI(1, 2) * DDB(1, 2, I)));
#undef DDB
    }
    bd->rhs[F_GAM1 + I] = (noshift + shiftG) - bd->ko[F_GAM1 + I];
  }
  bd->rhs[F_K] = (-bd->DDaTR +
                  alpha * (1.0 / 3.0 * (v[F_K] + 2.0 * theta) * (v[F_K] + 2.0 * theta + 2.0 * bd->K_F) + bd->AijAij) +
                  4.0 * PI * alpha * (bd->src[0] + bd->src[1]) + 4.0 * PI * v[F_ALPHA] * (bd->rho_F + bd->S_F) +
                  bd->adv[F_K] - 1.0 * p->k_damping_amp * H * exp(-5.0 * bd->phi) + 0.0 * (1.0 - 0.0) * theta -
                  bd->ko[F_K]);
  bd->rhs[F_PHI] = (0.1 * p->a_adj_amp * dt * H -
                    1.0 / 6.0 * (alpha * (v[F_K] + 2.0 * theta) + v[F_ALPHA] * bd->K_F - (DB(0, 0) + DB(1, 1) + DB(2, 2))) +
                    bd->adv[F_PHI] - bd->ko[F_PHI]);
  /* gauge: BSSNGaugeHandler.cc:16-35 */
  double lapse = 0.0;
  switch (p->lapse) {
  case OR_LAPSE_HARMONIC: lapse = -1.0 * alpha * (bd->K); break;                        /* :32-35 */
  case OR_LAPSE_CONFORMAL_FLRW: lapse = -1.0 / 3.0 * alpha * bd->K_avg; break;          /* :24-27 */
  case OR_LAPSE_ANHARMONIC: lapse = 1.0 * PW2(alpha) * (bd->K - bd->K_avg); break;      /* :50-58 */
  case OR_LAPSE_ONE_PLUS_LOG:                                                           /* :64-68 */
    lapse = -2.0 * alpha * (bd->K - bd->K_avg) * p->gd_c + b[0] * bd->dalpha[0] + b[1] * bd->dalpha[1] +
            b[2] * bd->dalpha[2];
    break;
  case OR_LAPSE_CONFORMAL_SYNC: lapse = -1.0 / 3.0 * alpha * bd->K_F; break;            /* :84-87 */
  case OR_LAPSE_AWA_GAUGE_WAVE: lapse = -1.0 * PW2(alpha) * v[F_K]; break;              /* :209-212 */
  case OR_LAPSE_AWA_SHIFTED_WAVE: lapse = -1.0 * v[F_K]; break;                         /* :217-220 */
  case OR_LAPSE_DAMPED_WAVE:                                                            /* :151-155 */
    lapse = PW2(alpha) * (p->dw_mu_l * (12.0 * bd->phi * p->dw_p - log(alpha)) - bd->K) + b[0] * bd->dalpha[0] +
            b[1] * bd->dalpha[1] + b[2] * bd->dalpha[2];
    break;
  case OR_LAPSE_TEST_K_DRIVER: /* :93-102; ev_DIFFK is the complete right-hand side above */
    lapse = -1.0 * p->k_driver_coeff *
            ((bd->rhs[F_K] - (PW2(bd->K_avg) / 3.0 + 4.0 * PI * bd->rho_avg)) + (bd->K - bd->K_avg));
    break;
  default: lapse = 0.0; break;                                  /* Static, TestAijDriverLapse */
  }
  bd->rhs[F_ALPHA] = lapse + bd->adv[F_ALPHA] - bd->ko[F_ALPHA];
  if (shift) {
    /* ev_beta^i = shift gauge - KO (bssn.cc:852-869, BSSNGaugeHandler.cc:113-273) */
    for (int I = 0; I < 3; ++I) {
      double sh = 0.0;
      switch (p->shift) {
      case OR_SHIFT_GAMMA_DRIVER: sh = bd->xv[OR_X_AUXB1 + I]; break; /* :113-143 */
      case OR_SHIFT_DAMPED_WAVE: {                                     /* :157-207 */
        /* the y-component reads Gamma3 where Gamma2 is meant (:185), kept */
        const double Gq = v[F_GAM1 + (I == 0 ? 0 : 2)];
        sh = b[0] * DB(0, I) + b[1] * DB(1, I) + b[2] * DB(2, I) - p->dw_mu_s * alpha * b[I] +
             alpha * (-p->dw_mu_l * (12.0 * bd->phi * p->dw_p - log(alpha)) * b[I] +
                      exp(-4.0 * bd->phi) *
                          (-GI(I, 0) * bd->dalpha[0] - GI(I, 1) * bd->dalpha[1] - GI(I, 2) * bd->dalpha[2] +
                           alpha * (Gq - 2.0 * (GI(I, 0) * bd->dphi[0] + GI(I, 1) * bd->dphi[1] +
                                                GI(I, 2) * bd->dphi[2]))));
        break;
      }
      case OR_SHIFT_AWA_SHIFTED_WAVE: /* :225-253 */
        sh = (p->gauge_wave_dir == I + 1) ? -2.0 * bd->K * alpha : 0.0;
        break;
      case OR_SHIFT_AIJ_DRIVER: /* :262-273: |ev_A_Ij + A_Ij| over j */
        sh = -p->k_driver_coeff * (fabs(bd->rhs[F_A11 + SY[I][0]] + AA(I, 0)) +
                                   fabs(bd->rhs[F_A11 + SY[I][1]] + AA(I, 1)) +
                                   fabs(bd->rhs[F_A11 + SY[I][2]] + AA(I, 2)));
        break;
      default: sh = 0.0; break; /* Static */
      }
      bd->rhsx[OR_X_BETA1 + I] = sh - bd->kox[OR_X_BETA1 + I];
    }
    /* ev_expN, bssn.cc:871-875 */
    bd->rhsx[OR_X_EXPN] = b[0] * bd->dexpN[0] + b[1] * bd->dexpN[1] + b[2] * bd->dexpN[2] - alpha * bd->K / 3.0;
    /* ev_auxB^i, bssn.cc:878-891 */
    if (p->shift == OR_SHIFT_GAMMA_DRIVER)
      for (int I = 0; I < 3; ++I)
        bd->rhsx[OR_X_AUXB1 + I] = 0.75 * bd->rhs[F_GAM1 + I] - p->gd_eta * bd->xv[OR_X_AUXB1 + I];
  }
#undef DB
#undef GI
#undef GG
#undef AA
#undef DG
#undef GU
#undef GLo
}

/* BSSN_MI, bssn_macros.h:487-506 (momentum constraint, index lowered) */
void or_point_momentum(const or_point *bd, double M[3]) {
  const double *gi = bd->gi, *v = bd->v, *Ac = bd->Acont;
#define GI(a, b) gi[SY[a][b]]
#define AA(a, b) v[F_A11 + SY[a][b]]
#define DA(j, a, b) bd->dA[j][SY[a][b]]
  for (int I = 0; I < 3; ++I) {
    double t6 = 0.0, td = 0.0;
    int first = 1;
    for (int n = 0; n < 3; ++n)
      for (int m = 0; m < 3; ++m) {
        double t = GI(m, n) * AA(m, I) * bd->dphi[n];
        t6 = first ? t : t6 + t;
        first = 0;
      }
    first = 1;
    for (int m = 0; m < 3; ++m)
      for (int n = 0; n < 3; ++n) {
        double t = GI(m, n) * DA(n, m, I);
        td = first ? t : td + t;
        first = 0;
      }
    for (int m = 0; m < 3; ++m) td = td - bd->Gammad[m] * AA(m, I);
    for (int n = 0; n < 3; ++n)
      for (int m = 0; m < 3; ++m) td = td - bd->GL[m][SY[n][I]] * Ac[SY[m][n]];
    M[I] = exp(6.0 * bd->phi) *
           (-2.0 / 3.0 * bd->dK[I] - 8 * PI * (bd->src[2 + I]) - 2.0 / 3.0 * 2.0 * 0.0 + 6.0 * (t6) + (td));
  }
#undef GI
#undef AA
#undef DA
}

/* ------------------------------------------------------- grid driver --- */
/* set_bd_values (bssn.cc:361-421) at a grid point: local values, stencils,
 * then the algebra. */
static void fill_point(or_sim *s, long i, long j, long k, or_point *bd, int with_ko, int with_dA) {
  memset(bd, 0, sizeof(*bd));
  long idx = ((long)i * s->p.ny + j) * s->p.nz + k;
  for (int f = 0; f < OR_NF; ++f) bd->v[f] = s->reg[f][OR_A][idx];
  for (int f = 0; f < OR_NSRC; ++f) bd->src[f] = s->src[f][idx];
  bd->phi_F = s->frw.phi_get;
  bd->K_F = s->frw.K_get;
  bd->rho_F = s->frw.rho_get;
  bd->S_F = s->frw.S_get;
  bd->K_avg = s->K_avg;
  bd->avg_vol = s->avg_vol;
  bd->rho_avg = s->rho_avg;
  for (int a = 0; a < 6; ++a) {
    const double *f = s->reg[a][OR_A];
    for (int kk = 0; kk < 3; ++kk) bd->dg[kk][a] = deriv(s, f, i, j, k, kk + 1);
  }
  /* calculate_ddgamma: BSSN_APPLY_TO_IJ_PERMS over (I,J) then fields */
  for (int I = 0; I < 3; ++I)
    for (int J = I; J < 3; ++J)
      for (int a = 0; a < 6; ++a) bd->ddg[SY[I][J]][a] = dderiv(s, s->reg[a][OR_A], i, j, k, I + 1, J + 1);
  const double *phi = s->reg[F_PHI][OR_A], *al = s->reg[F_ALPHA][OR_A];
  for (int d = 0; d < 3; ++d) bd->dphi[d] = deriv(s, phi, i, j, k, d + 1);
  for (int I = 0; I < 3; ++I)
    for (int J = I; J < 3; ++J) {
      bd->ddphi[SY[I][J]] = dderiv(s, phi, i, j, k, I + 1, J + 1);
      bd->ddalpha[SY[I][J]] = dderiv(s, al, i, j, k, I + 1, J + 1);
    }
  for (int d = 0; d < 3; ++d) bd->dalpha[d] = deriv(s, al, i, j, k, d + 1);
  for (int d = 0; d < 3; ++d) bd->dK[d] = deriv(s, s->reg[F_K][OR_A], i, j, k, d + 1);
  for (int jj = 0; jj < 3; ++jj)
    for (int kk = 0; kk < 3; ++kk) bd->dGam[jj][kk] = deriv(s, s->reg[F_GAM1 + kk][OR_A], i, j, k, jj + 1);
  if (with_dA)
    for (int a = 0; a < 6; ++a)
      for (int jj = 0; jj < 3; ++jj) bd->dA[jj][a] = deriv(s, s->reg[F_A11 + a][OR_A], i, j, k, jj + 1);
  if (with_ko)
    for (int f = 0; f < OR_NF; ++f) bd->ko[f] = ko_q(s, s->reg[f][OR_A], i, j, k);
  if (s->p.shift != OR_SHIFT_NONE) {
    /* calculate_dbeta, calculate_dexpN (bssn.cc:547-566); the double
     * derivatives of beta that BSSN_DT_GAMMAI_SHIFT evaluates in place;
     * the upwind advection terms of every field (bssn_macros.h:296-370,
     * bssn.cc:786-828) */
    for (int f = 0; f < n_extra(&s->p); ++f) bd->xv[f] = s->xreg[f][OR_A][idx];
    for (int m = 0; m < 3; ++m) {
      const double *bm = s->xreg[OR_X_BETA1 + m][OR_A];
      for (int jj = 0; jj < 3; ++jj) bd->dbeta[jj][m] = deriv(s, bm, i, j, k, jj + 1);
      for (int I = 0; I < 3; ++I)
        for (int J = I; J < 3; ++J) bd->ddbeta[m][SY[I][J]] = dderiv(s, bm, i, j, k, I + 1, J + 1);
      if (with_ko) bd->kox[OR_X_BETA1 + m] = ko_q(s, bm, i, j, k);
    }
    for (int d = 0; d < 3; ++d) bd->dexpN[d] = deriv(s, s->xreg[OR_X_EXPN][OR_A], i, j, k, d + 1);
    for (int f = 0; f < OR_NF; ++f) {
      const double *fa = s->reg[f][OR_A];
      bd->adv[f] = upwind(s, fa, i, j, k, 1, bd->xv[0]) + upwind(s, fa, i, j, k, 2, bd->xv[1]) +
                   upwind(s, fa, i, j, k, 3, bd->xv[2]);
    }
  }
  or_point_compute(&s->p, bd);
  /* bssn.cc:469-471, 663 */
  s->AijAij[idx] = bd->AijAij;
  s->ricci[idx] = bd->ricci;
}

/* BSSN::RKEvolve, bssn.cc:231-261 */
void or_bssn_rk_evolve(or_sim *s) {
  const int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; ++i) {
    or_point bd;
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        fill_point(s, i, j, k, &bd, 1, 0);
        long idx = ((long)i * ny + j) * nz + k;
        for (int f = 0; f < OR_NF; ++f) s->reg[f][OR_C][idx] = bd.rhs[f];
        for (int f = 0; f < n_extra(&s->p); ++f) s->xreg[f][OR_C][idx] = bd.rhsx[f];
      }
  }
}

/* CosmoSim::prepBSSNOutput, sims/sim.cc:214-231 */
/* BSSN::set1DConstraintOutput, bssn.cc:903-938: H and the momentum
 * component M_axis at N_axis points; the running index is passed as i for
 * every axis (bssn.cc:921, 931), and the stencils wrap it periodically. */
void or_constraint_strip(or_sim *s, int axis, long n1, long n2, double *H, double *M) {
  const long ext[3] = {s->p.nx, s->p.ny, s->p.nz};
  const long nx = s->p.nx, ny = s->p.ny, nz = s->p.nz;
  const long j = ((n1 % ny) + ny) % ny, k = ((n2 % nz) + nz) % nz;
  or_point bd;
  for (long x = 0; x < ext[axis - 1]; ++x) {
    double Mv[3];
    fill_point(s, x % nx, j, k, &bd, 0, 1);
    or_point_momentum(&bd, Mv);
    H[x] = bd.H;
    M[x] = Mv[axis - 1];
  }
}

void or_bssn_prep_output(or_sim *s) {
  const int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; ++i) {
    or_point bd;
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) fill_point(s, i, j, k, &bd, 0, 0);
  }
}

/* ------------------------------------------------------- RK4 register --- */
/* utils/RK4Register.h:118-176 */
static void reg_step_init(double *r[4], long n) {
#pragma omp parallel for schedule(static)
  for (long x = 0; x < n; ++x) {
    r[OR_A][x] = r[OR_P][x];
    r[OR_F][x] = 0;
  }
}
static void reg_finalize(double *r[4], long n, int stage, double dt) {
  double *p = r[OR_P], *c = r[OR_C], *f = r[OR_F];
#pragma omp parallel for schedule(static)
  for (long x = 0; x < n; ++x) {
    switch (stage) {
    case 1: f[x] += dt * c[x] / 6.0; c[x] = p[x] + dt * c[x] / 2.0; break;
    case 2: f[x] += dt * c[x] / 3.0; c[x] = p[x] + dt * c[x] / 2.0; break;
    case 3: f[x] += dt * c[x] / 3.0; c[x] = p[x] + dt * c[x]; break;
    default:
      f[x] += dt * c[x] / 6.0 + p[x];
      c[x] = f[x];
      p[x] = f[x];
      break;
    }
  }
  double *t = r[OR_A]; /* swap_a_c, RK4Register.h:103-107 */
  r[OR_A] = r[OR_C];
  r[OR_C] = t;
}

/* ---------------------------------------------------------- reductions --- */
/* math.h:1382-1415 (volume_average, conformal_average), 1506-1518 (min);
 * summed in index order (the reference sums per OpenMP thread). */
static double conformal_average(const or_sim *s, const double *f, const double *dphi, double phiF) {
  double sum = 0.0, vsum = 0.0;
  for (long x = 0; x < s->npts; ++x) sum += exp(6.0 * (dphi[x] + phiF)) * f[x];
  for (long x = 0; x < s->npts; ++x) vsum += exp(6.0 * (dphi[x] + phiF));
  double vol = vsum / (double)s->npts;
  return sum / (double)s->npts / vol;
}

/* BSSN::setExtraFieldData, bssn.cc:165-170 */
void or_bssn_extra_field_data(or_sim *s) {
  const double *K = s->reg[F_K][OR_A], *ph = s->reg[F_PHI][OR_A];
  double mn = K[0];
  for (long x = 0; x < s->npts; ++x)
    if (K[x] < mn) mn = K[x];
  s->K_min = mn;
  s->K_avg = conformal_average(s, K, ph, s->frw.phi_get);
  s->avg_vol = exp(6.0 * conformal_average(s, ph, ph, s->frw.phi_get));
  s->rho_avg = conformal_average(s, s->src[0], ph, s->frw.phi_get);
}

void or_bssn_get_extra(const or_sim *s, double out[4]) {
  out[0] = s->K_min; out[1] = s->K_avg; out[2] = s->avg_vol; out[3] = s->rho_avg;
}

/* numNaNs, math.h:1524-1541: tests the float-rounded value */
long or_num_nans(or_sim *s, const char *name) {
  const double *f = or_field(s, name);
  if (!f) return -1;
  long n = 0;
  for (long x = 0; x < s->npts; ++x) {
    union { float v; uint32_t u; } u;
    u.v = (float)f[x];
    if ((u.u << 1) > 0xff000000u) n++;
  }
  return n;
}

/* conformal_standard_deviation, utils/math.h:1458-1484 */
static double conformal_stdev(const or_sim *s, const double *f, const double *dphi, double phiF) {
  const double avg = conformal_average(s, f, dphi, phiF);
  double sum = 0.0, vsum = 0.0;
  for (long x = 0; x < s->npts; ++x) sum += exp(6.0 * (dphi[x] + phiF)) * PW2(avg - f[x]);
  for (long x = 0; x < s->npts; ++x) vsum += exp(6.0 * (dphi[x] + phiF));
  double vol = vsum / (double)s->npts;
  return sqrt(sum / ((double)s->npts - 1.0) / vol);
}

/* reductions exported for the bitwise checks against utils/math.h:
 * 0 average (1359-1371), 1 volume_average (1382-1393), 2 conformal_average
 * (1403-1415), 3 min (1506-1518), 4 max (1491-1503), 5 numNaNs (1524-1541),
 * 6 conformal_standard_deviation (1458-1484) */
double or_reduction(or_sim *s, int kind, const double *f, const double *dphi, double phiF) {
  const long n = s->npts;
  double acc = 0.0;
  switch (kind) {
  case 0:
    for (long x = 0; x < n; ++x) acc += f[x];
    return acc / (double)n;
  case 1:
    for (long x = 0; x < n; ++x) acc += exp(6.0 * (dphi[x] + phiF));
    return acc / (double)n;
  case 2: return conformal_average(s, f, dphi, phiF);
  case 6: return conformal_stdev(s, f, dphi, phiF);
  case 3: {
    double m = f[0];
    for (long x = 0; x < n; ++x)
      if (f[x] < m) m = f[x];
    return m;
  }
  case 4: {
    double m = f[0];
    for (long x = 0; x < n; ++x)
      if (f[x] > m) m = f[x];
    return m;
  }
  default: {
    long c = 0;
    for (long x = 0; x < n; ++x) {
      union { float v; uint32_t u; } u;
      u.v = (float)f[x];
      if ((u.u << 1) > 0xff000000u) c++;
    }
    return (double)c;
  }
  }
}

/* RK4Register arithmetic on a standalone register set, x' = c0 cos x + c1 x
 * (the same ODE the reference driver integrates with utils/RK4Register.h);
 * with c0 = c1 = 0 it integrates x' = exp(-x), the ODE of tests/rk4.cc. */
void or_rk4_demo(const double *x0, long n, double h, double c0, double c1, int nsteps, double *traj,
                 double *last_stage_a) {
  double *r[4];
  for (int q = 0; q < 4; ++q) r[q] = zalloc(n);
  for (long x = 0; x < n; ++x) r[OR_P][x] = x0[x];
  for (int st = 0; st < nsteps; ++st) {
    reg_step_init(r, n);
    for (int k = 1; k <= 4; ++k) {
      for (long x = 0; x < n; ++x) {
        double a = r[OR_A][x];
        if (st == nsteps - 1) last_stage_a[(long)(k - 1) * n + x] = a;
        r[OR_C][x] = (c0 == 0.0 && c1 == 0.0) ? exp(-a) : cos(a) * c0 + c1 * a;
      }
      reg_finalize(r, n, k, h);
    }
    for (long x = 0; x < n; ++x) traj[(long)st * n + x] = r[OR_P][x];
  }
  for (int q = 0; q < 4; ++q) free(r[q]);
}

/* ---------------------------------------------------- BSSN methods --- */
/* set_DIFFgamma_Aij_norm, bssn.cc:61-119 (on _a only) */
void or_bssn_normalize(or_sim *s) {
  double *G[6], *A[6];
  for (int a = 0; a < 6; ++a) { G[a] = s->reg[a][OR_A]; A[a] = s->reg[F_A11 + a][OR_A]; }
#pragma omp parallel for schedule(static)
  for (long x = 0; x < s->npts; ++x) {
    double g11 = G[0][x], g12 = G[1][x], g13 = G[2][x], g22 = G[3][x], g23 = G[4][x], g33 = G[5][x];
    double omd = -1.0 * (g11 + g22 + g33 - PW2(g12) - PW2(g13) - PW2(g23) + g11 * g22 + g11 * g33 + g22 * g33 -
                         PW2(g23) * g11 - PW2(g13) * g22 - PW2(g12) * g33 + 2.0 * g12 * g13 * g23 + g11 * g22 * g33);
    double t = -1.0 * expm1(log1p(-1.0 * omd) / 3.0);
    g11 = (g11 + t) / (1.0 - t);
    g22 = (g22 + t) / (1.0 - t);
    g33 = (g33 + t) / (1.0 - t);
    g12 = (g12) / (1.0 - t);
    g13 = (g13) / (1.0 - t);
    g23 = (g23) / (1.0 - t);
    G[0][x] = g11; G[1][x] = g12; G[2][x] = g13; G[3][x] = g22; G[4][x] = g23; G[5][x] = g33;
    double gi11 = 1.0 + g22 + g33 - PW2(g23) + g22 * g33;
    double gi22 = 1.0 + g11 + g33 - PW2(g13) + g11 * g33;
    double gi33 = 1.0 + g11 + g22 - PW2(g12) + g11 * g22;
    double gi12 = g13 * g23 - g12 * (1.0 + g33);
    double gi13 = g12 * g23 - g13 * (1.0 + g22);
    double gi23 = g12 * g13 - g23 * (1.0 + g11);
    double trA = gi11 * A[0][x] + gi22 * A[3][x] + gi33 * A[5][x] +
                 2.0 * (gi12 * A[1][x] + gi13 * A[2][x] + gi23 * A[4][x]);
    A[0][x] = (A[0][x] - 1.0 / 3.0 * (1.0 + g11) * trA) / (1.0 - t);
    A[3][x] = (A[3][x] - 1.0 / 3.0 * (1.0 + g22) * trA) / (1.0 - t);
    A[5][x] = (A[5][x] - 1.0 / 3.0 * (1.0 + g33) * trA) / (1.0 - t);
    A[1][x] = (A[1][x] - 1.0 / 3.0 * g12 * trA) / (1.0 - t);
    A[2][x] = (A[2][x] - 1.0 / 3.0 * g13 * trA) / (1.0 - t);
    A[4][x] = (A[4][x] - 1.0 / 3.0 * g23 * trA) / (1.0 - t);
  }
}

/* BSSN::stepInit, bssn.cc:219-226 */
void or_bssn_step_init(or_sim *s) {
  for (int f = 0; f < OR_NF; ++f) reg_step_init(s->reg[f], s->npts);
  for (int f = 0; f < n_extra(&s->p); ++f) reg_step_init(s->xreg[f], s->npts);
  or_bssn_extra_field_data(s);
  if (s->p.normalize_metric) or_bssn_normalize(s);
}

/* BSSN::K{n}Finalize, bssn.cc:267-305: FRW stage first, then registers */
void or_bssn_k_finalize(or_sim *s, int n) {
  frw_stage(&s->frw, n, s->p.dt);
  for (int f = 0; f < OR_NF; ++f) reg_finalize(s->reg[f], s->npts, n, s->p.dt);
  for (int f = 0; f < n_extra(&s->p); ++f) reg_finalize(s->xreg[f], s->npts, n, s->p.dt);
  or_bssn_extra_field_data(s);
}

/* BSSN::clearSrc, bssn.cc:310-319 */
void or_bssn_clear_src(or_sim *s) {
  for (int f = 0; f < OR_NSRC; ++f) memset(s->src[f], 0, sizeof(double) * (size_t)s->npts);
}

/* Static::addBSSNSrc, components/static/static.cc:19-36 */
void or_static_add_src(or_sim *s) {
  double *r = s->src[0];
  const double *ph = s->reg[F_PHI][OR_A];
  const double phiF = s->frw.phi_get, rhoF = s->frw.rho_get;
  for (long x = 0; x < s->npts; ++x)
    r[x] += exp(-6.0 * (ph[x] + phiF)) * s->DIFFD[x] + rhoF * expm1(-6.0 * ph[x]);
}

/* Lambda::addBSSNSource, components/Lambda/lambda.cc:26-43 */
void or_lambda_add_src(or_sim *s) {
  double lam = s->p.lambda;
  for (long x = 0; x < s->npts; ++x) {
    s->src[0][x] += lam;
    s->src[1][x] += -3.0 * lam;
  }
}

/* ------------------------------------------------------- dust fluid --- */
/* Dust::populateDerivedFields + getDustData, dust.cc:109-158, 177-200 */
static void dust_populate(or_sim *s) {
  const int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; ++i) {
    or_point bd;
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        long idx = ((long)i * ny + j) * nz + k;
        fill_point(s, i, j, k, &bd, 0, 0);
        double D = s->dust[0][OR_A][idx], S1 = s->dust[1][OR_A][idx], S2 = s->dust[2][OR_A][idx],
               S3 = s->dust[3][OR_A][idx];
        const double *gi = bd.gi, *g = bd.g;
        double em4phi = exp(-4.0 * bd.phi), e6phi = exp(6.0 * bd.phi);
        double giSS = em4phi * (gi[0] * S1 * S1 + gi[3] * S2 * S2 + gi[5] * S3 * S3 +
                                2.0 * (gi[1] * S1 * S2 + gi[2] * S1 * S3 + gi[4] * S2 * S3));
        double W = sqrt(1.0 + giSS / D / D);
        double v1 = em4phi / e6phi * (gi[0] * S1 + gi[1] * S2 + gi[2] * S3);
        double v2 = em4phi / e6phi * (gi[1] * S1 + gi[3] * S2 + gi[4] * S3);
        double v3 = em4phi / e6phi * (gi[2] * S1 + gi[4] * S2 + gi[5] * S3);
        double vv[3] = {v1, v2, v3}, SS[3] = {S1, S2, S3};
        for (int d = 0; d < 3; ++d) s->dd[DD_ADV1 + d][idx] = bd.alpha * D * vv[d];
        for (int a = 0; a < 3; ++a)
          for (int d = 0; d < 3; ++d) s->dd[DD_AS1V1 + 3 * a + d][idx] = bd.alpha * SS[a] * vv[d];
        for (int d = 0; d < 3; ++d) {
          const double *dgd = bd.dg[d];
          double dp = bd.dphi[d];
          s->dd[DD_S1SRC + d][idx] =
              1.0 / 2.0 * bd.alpha * W * D * exp(4.0 * bd.phi) *
                  (v1 * v1 * (4.0 * g[0] * dp + dgd[0]) + v2 * v2 * (4.0 * g[3] * dp + dgd[3]) +
                   v3 * v3 * (4.0 * g[5] * dp + dgd[5]) +
                   2.0 * (v1 * v2 * (4.0 * g[1] * dp + dgd[1]) + v1 * v3 * (4.0 * g[2] * dp + dgd[2]) +
                          v2 * v3 * (4.0 * g[4] * dp + dgd[4]))) -
              W * D * bd.dalpha[d];
        }
        double e4phi = exp(4.0 * bd.phi);
        s->dd[DD_DETG][idx] = exp(6.0 * bd.phi);
        for (int a = 0; a < 6; ++a) s->dd[DD_G11 + a][idx] = e4phi * g[a];
        s->dd[DD_W][idx] = W;
      }
  }
}

void or_dust_step_init(or_sim *s) {
  for (int f = 0; f < OR_NDUST; ++f) reg_step_init(s->dust[f], s->npts);
  dust_populate(s);
}

/* Dust::RKEvolve, dust.cc:160-174, 203-224 */
void or_dust_rk_evolve(or_sim *s) {
  dust_populate(s);
  const int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        long idx = ((long)i * ny + j) * nz + k;
        s->dust[0][OR_C][idx] = deriv(s, s->dd[DD_ADV1], i, j, k, 1) + deriv(s, s->dd[DD_ADV1 + 1], i, j, k, 2) +
                                deriv(s, s->dd[DD_ADV1 + 2], i, j, k, 3);
        for (int a = 0; a < 3; ++a)
          s->dust[1 + a][OR_C][idx] = deriv(s, s->dd[DD_AS1V1 + 3 * a], i, j, k, 1) +
                                      deriv(s, s->dd[DD_AS1V1 + 3 * a + 1], i, j, k, 2) +
                                      deriv(s, s->dd[DD_AS1V1 + 3 * a + 2], i, j, k, 3) +
                                      s->dd[DD_S1SRC + a][idx];
      }
}

void or_dust_k_finalize(or_sim *s, int n) {
  for (int f = 0; f < OR_NDUST; ++f) reg_finalize(s->dust[f], s->npts, n, s->p.dt);
}

/* Dust::addBSSNSrc, dust.cc:226-264: W, detg, g_ij from the last
 * populateDerivedFields call, D and S_i from the current _a register. */
void or_dust_add_src(or_sim *s) {
  const double *D = s->dust[0][OR_A], *S1 = s->dust[1][OR_A], *S2 = s->dust[2][OR_A], *S3 = s->dust[3][OR_A];
  const double *W = s->dd[DD_W], *detg = s->dd[DD_DETG];
  const double *gg[6];
  for (int a = 0; a < 6; ++a) gg[a] = s->dd[DD_G11 + a];
  double **src = s->src;
  for (long x = 0; x < s->npts; ++x) {
    src[0][x] += W[x] * D[x] / detg[x];
    src[1][x] += D[x] / detg[x] * (W[x] * W[x] - 1.0) / W[x];
    src[2][x] += S1[x] / detg[x];
    src[3][x] += S2[x] / detg[x];
    src[4][x] += S3[x] / detg[x];
    const double SS[3] = {S1[x], S2[x], S3[x]};
    for (int I = 0; I < 3; ++I)
      for (int J = I; J < 3; ++J) {
        int a = SY[I][J];
        src[5 + a][x] += SS[I] * SS[J] / W[x] / D[x] / detg[x] - 1.0 / 3.0 * gg[a][x] * src[1][x];
      }
  }
}

/* ------------------------------------------------------------ drivers --- */
static int step_once(or_sim *s) {
  const int dust = s->p.matter == OR_MATTER_DUST;
  /* initStaticStep (sims/static.cc:74-109) / initDustStep (sims/dust.cc:53-92) */
  or_bssn_step_init(s);
  if (dust) or_dust_step_init(s);
  or_bssn_clear_src(s);
  if (s->p.matter == OR_MATTER_STATIC) or_static_add_src(s);
  if (dust) or_dust_add_src(s);
  or_lambda_add_src(s);
  /* runCommonStepTasks NaN guard, sims/sim.cc:179-183 */
  if (or_num_nans(s, "DIFFphi_a") > 0) return 10;
  /* runStaticStep / runDustStep */
  for (int n = 1; n <= 4; ++n) {
    if (n > 1) {
      or_bssn_clear_src(s);
      if (s->p.matter == OR_MATTER_STATIC) or_static_add_src(s);
      if (dust) or_dust_add_src(s);
      or_lambda_add_src(s);
    }
    or_bssn_rk_evolve(s);
    if (dust) or_dust_rk_evolve(s);
    or_bssn_k_finalize(s, n);
    if (dust) or_dust_k_finalize(s, n);
  }
  return 0;
}

int or_run_steps(or_sim *s, int nsteps) {
  for (int t = 0; t < nsteps; ++t) {
    int rc = step_once(s);
    if (rc) return rc;
  }
  return 0;
}

/* ---------------------------------------------------------- FRW API --- */
void or_frw_set(or_sim *s, double phi, double K) {
  s->frw.phi = s->frw.phi_get = phi;
  s->frw.K = s->frw.K_get = K;
}
void or_frw_add_fluid(or_sim *s, double rho, double w) {
  or_frw *F = &s->frw;
  if (F->nf >= OR_MAX_FLUIDS) return;
  F->rho[F->nf] = rho;
  F->w[F->nf] = w;
  F->nf++;
  frw_sum_get(F, F->rho);
}
void or_frw_get(const or_sim *s, double out[4]) {
  out[0] = s->frw.phi_get; out[1] = s->frw.K_get; out[2] = s->frw.rho_get; out[3] = s->frw.S_get;
}
void or_frw_step_stage(or_sim *s, int n, double h) { frw_stage(&s->frw, n, h); }

/* ------------------------------------------------------- constraints --- */
/* BSSN::setConstraintCalcs, bssn.cc:965-1071, bssn_macros.h:530-585 */
typedef struct {
  double val, scale, scaled;
} cstat;

static void point_constraints(const or_point *bd, cstat out[5]) {
  const double *gi = bd->gi, *v = bd->v;
  /* H: bssn.cc:1073-1100 */
  out[0].val = bd->H;
  out[0].scale = (exp(5.0 * bd->phi) / 8.0) *
                 sqrt(PW2(bd->ricci) + PW2(bd->AijAij) + PW2(2.0 / 3.0 * PW2(bd->K + 2.0 * 0.0)) + PW2(16.0 * PI * bd->r));
  /* M: BSSN_MI / BSSN_MI_SCALE */
  double M[3], Ms[3];
  or_point_momentum(bd, M);
  {
    const double *Ac = bd->Acont;
#define GI(a, b) gi[SY[a][b]]
#define AA(a, b) v[F_A11 + SY[a][b]]
    for (int I = 0; I < 3; ++I) {
      double t6 = 0.0, td = 0.0;
      int first = 1;
      for (int n = 0; n < 3; ++n)
        for (int m = 0; m < 3; ++m) {
          double t = GI(m, n) * AA(m, I) * bd->dphi[n];
          t6 = first ? t : t6 + t;
          first = 0;
        }
      first = 1;
      for (int m = 0; m < 3; ++m)
        for (int n = 0; n < 3; ++n) {
          double t = GI(m, n) * bd->dA[n][SY[m][I]];
          td = first ? t : td + t;
          first = 0;
        }
      for (int m = 0; m < 3; ++m) td = td - bd->Gammad[m] * AA(m, I);
      for (int n = 0; n < 3; ++n)
        for (int m = 0; m < 3; ++m) td = td - bd->GL[m][SY[n][I]] * Ac[SY[m][n]];
      Ms[I] = exp(6.0 * bd->phi) * (fabs(2.0 / 3.0 * bd->dK[I]) + fabs(8 * PI * (bd->src[2 + I])) +
                                    fabs(2.0 / 3.0 * 2.0 * 0.0) + 6.0 * fabs(t6) + fabs(td));
    }
  }
  out[1].val = sqrt(PW2(M[0]) + PW2(M[1]) + PW2(M[2]));
  out[1].scale = sqrt(PW2(Ms[0]) + PW2(Ms[1]) + PW2(Ms[2]));
  /* G: BSSN_GI_CALC / BSSN_GI_SCALE bssn_macros.h:530-536 */
  double Gc[3], Gs[3];
  for (int I = 0; I < 3; ++I) {
    const double *Gi = bd->G[I];
    Gc[I] = v[F_GAM1 + I] - gi[0] * Gi[0] - gi[3] * Gi[3] - gi[5] * Gi[5] -
            2.0 * (gi[1] * Gi[1] + gi[2] * Gi[2] + gi[4] * Gi[4]);
    Gs[I] = fabs(v[F_GAM1 + I]) + fabs(gi[0] * Gi[0] - gi[3] * Gi[3] - gi[5] * Gi[5] -
                                       2.0 * (gi[1] * Gi[1] + gi[2] * Gi[2] + gi[4] * Gi[4]));
  }
  out[2].val = sqrt(PW2(Gc[0]) + PW2(Gc[1]) + PW2(Gc[2]));
  out[2].scale = sqrt(PW2(Gs[0]) + PW2(Gs[1]) + PW2(Gs[2]));
  /* A: bssn.cc:1182-1192 */
  out[3].val = gi[0] * AA(0, 0) + gi[3] * AA(1, 1) + gi[5] * AA(2, 2) +
               2.0 * (gi[1] * AA(0, 1) + gi[2] * AA(0, 2) + gi[4] * AA(1, 2));
  out[3].scale = fabs(gi[0] * AA(0, 0)) + fabs(gi[3] * AA(1, 1)) + fabs(gi[5] * AA(2, 2)) +
                 2.0 * (fabs(gi[1] * AA(0, 1)) + fabs(gi[2] * AA(0, 2)) + fabs(gi[4] * AA(1, 2)));
#undef GI
#undef AA
  /* S: bssn.cc:1199-1219 */
  {
    double g11 = v[0], g12 = v[1], g13 = v[2], g22 = v[3], g23 = v[4], g33 = v[5];
    out[4].val = fabs(g11 + g22 + g33 - PW2(g12) - PW2(g13) - PW2(g23) + g11 * g22 + g11 * g33 + g22 * g33 -
                      PW2(g23) * g11 - PW2(g13) * g22 - PW2(g12) * g33 + 2.0 * g12 * g13 * g23 + g11 * g22 * g33);
    out[4].scale = (fabs(g11) + fabs(g22) + fabs(g33) + PW2(g12) + PW2(g13) + PW2(g23) + fabs(g11 * g22) +
                    fabs(g11 * g33) + fabs(g22 * g33) + PW2(g23) * g11 + PW2(g13) * g22 + PW2(g12) * g33 +
                    2.0 * fabs(g12 * g13 * g23) + fabs(g11 * g22 * g33));
  }
  for (int c = 0; c < 5; ++c) out[c].scaled = out[c].val / out[c].scale;
}

/* Test access to the per-point values setConstraintCalcs reduces
 * (bssn.cc:965-1071): {val, scale, scaled} of H, M, G, A, S, 15 doubles.
 * or_point_constraints: one point filled by or_point_compute (with dA);
 * or_constraint_points: every grid point of _a, out[15 * idx + q]. */
void or_point_constraints(const or_point *bd, double out[15]) {
  cstat c[5];
  point_constraints(bd, c);
  for (int q = 0; q < 5; ++q) {
    out[3 * q] = c[q].val;
    out[3 * q + 1] = c[q].scale;
    out[3 * q + 2] = c[q].scaled;
  }
}

void or_constraint_points(or_sim *s, double *out) {
  const int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; ++i) {
    or_point bd;
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        long idx = ((long)i * ny + j) * nz + k;
        fill_point(s, i, j, k, &bd, 0, 1);
        or_point_constraints(&bd, out + 15 * idx);
      }
  }
}

void or_bssn_constraints(or_sim *s, double H[8], double M[8], double G[7], double A[7], double S[7]) {
  const long n = s->npts;
  const int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz;
  cstat *all = (cstat *)malloc(sizeof(cstat) * 5 * (size_t)n);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nx; ++i) {
    or_point bd;
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        long idx = ((long)i * ny + j) * nz + k;
        fill_point(s, i, j, k, &bd, 0, 1);
        point_constraints(&bd, all + 5 * idx);
      }
  }
  double *outs[5] = {H, M, G, A, S};
  const double dx = s->p.dx;
  for (int c = 0; c < 5; ++c) {
    double mean = 0, mean_scale = 0, mean_scaled = 0, mx = 0, mx_scaled = 0, L2 = 0;
    for (long x = 0; x < n; ++x) {
      cstat t = all[5 * x + c];
      mean += t.val;
      mean_scale += t.scale;
      mean_scaled += t.scaled;
      L2 += t.val * t.val * dx * dx * dx;
      if (fabs(t.val) > mx) mx = fabs(t.val);
      if (fabs(t.scaled) > mx_scaled) mx_scaled = fabs(t.scaled);
    }
    mean /= (double)n;          /* BSSN_NORMALIZE_MEAN: mean_scale is left as a sum */
    mean_scaled /= (double)n;
    double sd = 0, sd_scaled = 0;
    for (long x = 0; x < n; ++x) {
      cstat t = all[5 * x + c];
      sd += PW2(t.val - mean);
      sd_scaled += PW2(t.scaled - mean_scaled);
    }
    sd = sqrt(sd / ((double)n - 1.0));
    sd_scaled = sqrt(sd_scaled / ((double)n - 1.0));
    double *o = outs[c];
    o[0] = mean; o[1] = sd; o[2] = mx; o[3] = mean_scale; o[4] = mean_scaled; o[5] = sd_scaled; o[6] = mx_scaled;
    if (c < 2) o[7] = sqrt(L2);
  }
  free(all);
}
