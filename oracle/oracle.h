/*
 * oracle.h -- CPU restatement of CosmoGRaPH's BSSN evolution hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  Nothing in cosmograph_amd/ links, loads or calls it.
 *
 * Every function restates one piece of the reference (file:line cited next
 * to it in oracle.c).  Layout follows the reference: one contiguous double
 * array per field per register, index (i*ny + j)*nz + k, periodic in all
 * three directions.  All arithmetic is FP64 (long double only where the
 * reference's FRW integrator promotes through its PI_L literal).
 *
 * Parity pins (see tests/):
 *   stencils, KO, reductions   -> bitwise vs the reference's utils/math.h
 *                                 built from source into oracle/_ref
 *   RK4 register arithmetic    -> bitwise vs utils/RK4Register.h (oracle/_ref)
 *   FRW integrator             -> bitwise vs utils/FRW.h (oracle/_ref)
 *   static-dust source         -> bitwise vs components/static/static.cc
 *   point algebra (RHS, H, M)  -> brute-force ADM identities (autodiff/sympy)
 *   grid composition/sequencing-> 4th-order convergence to the brute-force
 *                                 RHS, homogeneous-mode RK4 equivalence
 */
#ifndef COSMO_ORACLE_H
#define COSMO_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OR_NF 18     /* evolved BSSN fields, reference order bssn_macros.h:77-99 */
#define OR_NSRC 11   /* source arrays bssn_macros.h:101-112 */
#define OR_NDUST 4   /* D, S1, S2, S3 (components/dust_fluid/dust.h:32-35) */
#define OR_NX 7      /* shift build: beta1..3, expN (USE_BSSN_SHIFT), auxB1..3 (USE_GAMMA_DRIVER),
                        bssn_macros.h:18-47 */

enum { OR_P = 0, OR_A = 1, OR_C = 2, OR_F = 3 };
enum { OR_MATTER_NONE = 0, OR_MATTER_STATIC = 1, OR_MATTER_DUST = 2 };
/* lapse gauges of BSSNGaugeHandler (components/bssn/BSSNGaugeHandler.h:96-111)
 * that need no shift */
enum {
  OR_LAPSE_STATIC = 0, OR_LAPSE_HARMONIC = 1, OR_LAPSE_CONFORMAL_FLRW = 2, OR_LAPSE_ANHARMONIC = 3,
  OR_LAPSE_ONE_PLUS_LOG = 4, OR_LAPSE_CONFORMAL_SYNC = 5, OR_LAPSE_AWA_GAUGE_WAVE = 6,
  OR_LAPSE_AWA_SHIFTED_WAVE = 7, OR_LAPSE_DAMPED_WAVE = 8, OR_LAPSE_TEST_AIJ_DRIVER = 9,
  OR_LAPSE_TEST_K_DRIVER = 10
};
/* shift gauges (BSSNGaugeHandler.h:113-135).  0 = the default build without
 * USE_BSSN_SHIFT (beta = 0, no extra fields); any other value is the
 * USE_BSSN_SHIFT build with that shift, GammaDriver also with
 * USE_GAMMA_DRIVER (the auxB fields). */
enum {
  OR_SHIFT_NONE = 0, OR_SHIFT_STATIC = 1, OR_SHIFT_GAMMA_DRIVER = 2, OR_SHIFT_DAMPED_WAVE = 3,
  OR_SHIFT_AWA_SHIFTED_WAVE = 4, OR_SHIFT_AIJ_DRIVER = 5
};
enum { OR_X_BETA1 = 0, OR_X_EXPN = 3, OR_X_AUXB1 = 4 };

typedef struct {
  int nx, ny, nz;
  int order;              /* stencil order 2,4,6,8 */
  double dx, dt;
  double ko_coeff;        /* KO_damping_coefficient */
  double a_adj_amp;
  double k_damping_amp;
  int normalize_metric;
  int lapse;
  int matter;
  double lambda;          /* Lambda::lambda (rho_Lambda) */
  double gd_c;            /* gauge parameters (BSSNGaugeHandler.h:139-149) */
  double dw_mu_l, dw_p;
  int shift;              /* OR_SHIFT_* */
  int gauge_wave_dir;     /* AwA shifted wave direction 1..3 (BSSNGaugeHandler.h:71) */
  double gd_eta;          /* Gamma-driver damping (bssn.cc:18) */
  double dw_mu_s;         /* damped-wave shift parameter (BSSNGaugeHandler.h:67) */
  double k_driver_coeff;  /* TestKDriverLapse / AijDriverShift (BSSNGaugeHandler.h:85) */
} or_params;

/* Per-point data: the oracle's analogue of BSSNData.  Inputs first (filled
 * by the grid driver from stencils, or by a test from exact derivatives),
 * then everything the algebra derives.  Only doubles, so a ctypes mirror is
 * a flat array. */
typedef struct {
  /* ---- inputs ---- */
  double v[OR_NF];        /* DIFF field values (_a register) */
  double src[OR_NSRC];    /* DIFFr DIFFS S1 S2 S3 STF11 STF12 STF13 STF22 STF23 STF33 */
  double dg[3][6];        /* d_k DIFFgamma_ij  [k][sym ij] */
  double ddg[6][6];       /* d_k d_l DIFFgamma_ij [sym kl][sym ij] */
  double dphi[3], ddphi[6];
  double dalpha[3], ddalpha[6];
  double dK[3];
  double dGam[3][3];      /* d_j Gamma^k  [j][k] */
  double dA[3][6];        /* d_j A_kl [j][sym kl]; only for the momentum constraint */
  double ko[OR_NF];       /* KO_dissipation_Q per field */
  double phi_F, K_F, rho_F, S_F;
  double K_avg, rho_avg, avg_vol;
  /* ---- derived ---- */
  double phi, K, r, S, alpha;
  double g[6], gi[6];
  double Acont[6], AijAij;
  double G[3][6];         /* Gamma^i_jk  [i][sym jk] */
  double GL[3][6];        /* Gamma_ijk   [i][sym jk] */
  double Gammad[3];
  double DDphi[6];
  double DDaTF[6], DDaTR;
  double Uricci[6], ricci_ij[6], ricci, ricciTF[6];
  double H;
  double rhs[OR_NF];
  /* ---- shift build inputs (zero without USE_BSSN_SHIFT) ---- */
  double xv[OR_NX];       /* beta1..3, expN, auxB1..3 (_a register) */
  double dbeta[3][3];     /* d_j beta^i  [j][i] */
  double ddbeta[3][6];    /* d_k d_l beta^i  [i][sym kl] */
  double dexpN[3];
  double adv[OR_NF];      /* per field: sum over d of upwind_derivative(d, field, beta^d) */
  double kox[OR_NX];      /* KO_dissipation_Q of beta^i (the other extras have none) */
  /* ---- shift build outputs ---- */
  double rhsx[OR_NX];
} or_point;

typedef struct or_sim or_sim;

/* construction / state access */
or_sim *or_create(const or_params *p);
void or_destroy(or_sim *s);
double *or_field(or_sim *s, const char *name);   /* reference names, e.g. "DIFFphi_p", "DIFFD_a", "D_p" */
long or_points(const or_sim *s);
int or_num_threads(void);

/* FRW reference integrator (utils/FRW.h) */
void or_frw_set(or_sim *s, double phi, double K);
void or_frw_add_fluid(or_sim *s, double rho, double w);
void or_frw_get(const or_sim *s, double out[4]);       /* phi K rho S */
void or_frw_step_stage(or_sim *s, int n, double h);    /* n=1..4: P1,P2,P3,RK_total */

/* BSSN class methods (components/bssn/bssn.cc) */
void or_bssn_step_init(or_sim *s);
void or_bssn_rk_evolve(or_sim *s);
void or_bssn_k_finalize(or_sim *s, int n);
void or_bssn_clear_src(or_sim *s);
void or_bssn_normalize(or_sim *s);
void or_bssn_extra_field_data(or_sim *s);
void or_bssn_get_extra(const or_sim *s, double out[4]);   /* K_min K_avg avg_vol rho_avg */
void or_bssn_constraints(or_sim *s, double H[8], double M[8], double G[7], double A[7], double S[7]);
void or_bssn_prep_output(or_sim *s);   /* fills ricci_a / AijAij_a like prepBSSNOutput */
long or_num_nans(or_sim *s, const char *name);

/* matter (components/static, components/Lambda, components/dust_fluid) */
void or_static_add_src(or_sim *s);
void or_lambda_add_src(or_sim *s);
void or_dust_step_init(or_sim *s);
void or_dust_rk_evolve(or_sim *s);
void or_dust_k_finalize(or_sim *s, int n);
void or_dust_add_src(or_sim *s);

/* whole RK4 steps in the StaticSim / DustSim sequence (sims/static.cc:74-170,
 * sims/dust.cc:53-161) without output; returns 10 on a NaN (sims/sim.cc:179). */
int or_run_steps(or_sim *s, int nsteps);

/* reductions (utils/math.h) and RK4Register arithmetic, for bitwise checks */
double or_reduction(or_sim *s, int kind, const double *f, const double *dphi, double phiF);
void or_constraint_strip(or_sim *s, int axis, long n1, long n2, double *H, double *M);
void or_rk4_demo(const double *x0, long n, double h, double c0, double c1, int nsteps, double *traj,
                 double *last_stage_a);

/* point algebra on caller-filled inputs (for brute-force identity tests) */
void or_point_compute(const or_params *p, or_point *bd);
void or_point_momentum(const or_point *bd, double M[3]);
/* {val, scale, scaled} of H, M, G, A, S (setConstraintCalcs' per-point
 * values, bssn.cc:965-1071) at one point / at every point of _a */
void or_point_constraints(const or_point *bd, double out[15]);
void or_constraint_points(or_sim *s, double *out);
/* grid stencil probes (for bitwise comparison with utils/math.h) */
double or_stencil(or_sim *s, const double *f, int kind, int d1, int d2, long i, long j, long k);
/* upwind_derivative (utils/math.h:1247-1252) along d with speed c at every point:
 * out[(d-1)*P + idx], c per point */
void or_upwind(or_sim *s, const double *f, const double *c, double *out);

#ifdef __cplusplus
}
#endif
#endif
