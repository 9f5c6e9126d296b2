/*
 * CPU oracle for the CaPS round-trip-matrix hot path (FP64, plain C99).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against, and the CPU baseline timed beside it.  Only tests,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product library (caps_amd/libcaps_amd.so) never links or calls it.
 *
 * Each function restates one algorithm of the reference (CaPS v0.5) and cites
 * the reference file:line it follows.  Where the oracle deliberately differs
 * from the reference, the difference is stated at the function:
 *   - k-integrals use an adaptive Gauss-Legendre bisection (G10 vs G20 error
 *     estimate) instead of QUADPACK dqags/dqage/dqagi, run to a tighter
 *     tolerance (default 1e-12) than the reference's 1e-8;
 *   - log I_0 / log I_1 use the power series for every argument instead of
 *     the cephes Chebyshev tables (the arguments here stay below ~30);
 *   - the determinant is a dense LU with partial pivoting for every detalg;
 *     the HODLR mode keeps the reference's trace shortcuts around it.
 *
 * Parity status of every function is listed in oracle/README.md (what pins it).
 */
#ifndef CAPS_ORACLE_H
#define CAPS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#include <stddef.h>

typedef signed char orc_sign_t;

/* ---- logfac.c ---------------------------------------------------------- */
double orc_logi(unsigned n);
double orc_lfac(unsigned n);
double orc_lfac2(unsigned n);

/* ---- misc.c ------------------------------------------------------------ */
double orc_kahan_sum(const double *x, size_t n);
double orc_sqrtpm1(double x);
double orc_logadd(double a, double b);
double orc_logadd_ms(const double *v, const orc_sign_t *s, int n, orc_sign_t *sign);

/* ---- bessel.c ---------------------------------------------------------- */
double orc_bessel_logIn_series(double nu, double x);
double orc_bessel_logIn(int n, double x);
double orc_bessel_ratioI(double nu, double x);
double orc_bessel_logInu_asymp(double nu, double x);
double orc_bessel_logKnu_asymp(double nu, double x);
void   orc_bessel_logInKn_half(int n, double x, double *logI, double *logK);
double orc_bessel_logIn_half(int n, double x);
double orc_bessel_logKn_half(int n, double x);

/* ---- plm.c ------------------------------------------------------------- */
double orc_lnPl(int l, double x);
double orc_lnPlm(int l, int m, double x);
double orc_lnPlm_upwards(int l, int m, double x);
double orc_lnPlm_downwards(int l, int m, double x);
double orc_Plm_cf(long l, long m, double x);
double orc_dlnPlm(int l, int m, double x, double *d2);

/* ---- roundtrip.c ------------------------------------------------------- */
/* material: epsm1 = ε(iξ)-1 of plate and sphere at the frequency of the
 * block; +INFINITY means perfect reflector. */
typedef struct {
    double R, L, LbyR, calL, y;
    int ldim;
    double epsrel;          /* quadrature tolerance of the oracle */
    int detalg;             /* 0 HODLR-mode shortcuts, 1 LU-mode shortcuts */
    /* Drude parameters in rad/s (omegap = INFINITY: perfect reflector) */
    double omegap_plate, gamma_plate, omegap_sphere, gamma_sphere;
} orc_caps_t;

void   orc_caps_init(orc_caps_t *c, double R, double L);
double orc_epsm1_plate(const orc_caps_t *c, double xi_);
double orc_epsm1_sphere(const orc_caps_t *c, double xi_);
double orc_lnLambda(int l1, int l2, int m);
int    orc_estimate_lmin(const orc_caps_t *c, int m);
void   orc_fresnel(double epsm1, double xi_, double k_, double *rTE, double *rTM);
void   orc_mie(const orc_caps_t *c, double xi_, int l, double *lna, double *lnb);
double orc_K_estimate(int nu, int m, double alpha, double eps, double *a, double *b, double *approx);
double orc_integrate_plasma(int nu, double omega, double epsrel);   /* ξ = 0 plasma integral */

typedef struct orc_block orc_block_t;
orc_block_t *orc_block_new(const orc_caps_t *c, double xi_, int m);
void   orc_block_free(orc_block_t *b);
int    orc_block_lmin(const orc_block_t *b);
double orc_block_K(orc_block_t *b, int nu, int p, orc_sign_t *sign);
double orc_block_I(orc_block_t *b, int l1, int l2, int p, orc_sign_t *sign);
double orc_block_I_abs(orc_block_t *B, int l1, int l2, int p);   /* log Σ|terms| of I */
double orc_block_elem(orc_block_t *b, int l1, int l2, char p1, char p2);
void   orc_block_matrix(orc_block_t *b, double *M);   /* N x N row-major, N = 2 ldim */
long   orc_block_counters(const orc_block_t *b, long *n_K, long *n_Kevals, long *n_I, long *n_qterms);

/* ---- logdet.c ---------------------------------------------------------- */
double orc_logdet_lu(double *A, int n);              /* log|det A|, A overwritten */
double orc_logdet_1mM(const double *M, int n, int detalg);  /* log det(1-M) with shortcuts */
int    orc_dominant_1mM(const double *M, int n);              /* 1 - M column diagonally dominant */
double orc_logdet_1mM_delta(const double *M, int n);          /* log|det(1-M)|, delta form, no pivoting */
double orc_logdetD(const orc_caps_t *c, double xi_, int m);
void   orc_logdetD_many(const orc_caps_t *c, const double *xi_, const int *m, int n, double *out, int threads);
double orc_ht_drude(double LbyR);

/* ---- xi0.c ------------------------------------------------------------- */
void   orc_logdetD0(const orc_caps_t *c, int m, double omegap, double *EE, double *MM, double *MMp);
double orc_ht_perf(const orc_caps_t *c, double eps);
double orc_ht_plasma(const orc_caps_t *c, double omegap, double eps);

#ifdef __cplusplus
}
#endif

#endif
