/*
 * Oracle: Fresnel and Mie coefficients, the k-integrals K_ν and I_{l1 l2},
 * the A/B/C/D combinations and the round-trip matrix elements of one
 * (ξ, m) block.  Test infrastructure only (see oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define CAPS_C 299792458.

static int sgn(double v) { return (v > 0) - (v < 0); }

/* Geometry and defaults.  Reference: src/libcaps.c:183-214
 * (ldim = ceil(max(20, 5 R/L)), epsrel 1e-8, y = -log 2 - log1p(L/R)). */
void orc_caps_init(orc_caps_t *c, double R, double L)
{
    memset(c, 0, sizeof(*c));
    c->R = R;
    c->L = L;
    c->LbyR = L/R;
    c->calL = L + R;
    c->y = -log(2.) - log1p(c->LbyR);
    const double l = 5/c->LbyR;
    c->ldim = (int)ceil(l > 20 ? l : 20);
    c->epsrel = 1e-12;
    c->detalg = 0;
    c->omegap_plate = c->omegap_sphere = INFINITY;
    c->gamma_plate = c->gamma_sphere = 0;
}

/* ε(iξ)-1 of the Drude model ωp²/(ξ(ξ+γ)) at ξ = xi_ c/(L+R); +∞ for perfect
 * reflectors.  Reference: src/libcaps.c:104-120,129-132,150-156. */
static double drude_epsm1(double omegap, double gamma, double xi_, double calL)
{
    if(isinf(omegap))
        return INFINITY;
    const double xi = xi_*CAPS_C/calL;
    return omegap*omegap/(xi*(xi + gamma));
}

double orc_epsm1_plate(const orc_caps_t *c, double xi_)
{
    return drude_epsm1(c->omegap_plate, c->gamma_plate, xi_, c->calL);
}

double orc_epsm1_sphere(const orc_caps_t *c, double xi_)
{
    return drude_epsm1(c->omegap_sphere, c->gamma_sphere, xi_, c->calL);
}

/* log Λ_{l1 l2}^{(m)}.  Reference: src/libcaps.c:45-48. */
double orc_lnLambda(int l1, int l2, int m)
{
    return (orc_logi(2*l1 + 1) + orc_logi(2*l2 + 1) - orc_logi(l1) - orc_logi(l1 + 1)
            - orc_logi(l2) - orc_logi(l2 + 1) + orc_lfac(l1 - m) + orc_lfac(l2 - m)
            - orc_lfac(l1 + m) - orc_lfac(l2 + m))/2.0;
}

/* Lower end of the sliding ℓ-window [lmin, lmin+ldim).
 * Reference: src/libcaps.c:63-88. */
int orc_estimate_lmin(const orc_caps_t *c, int m)
{
    if(m == 0)
        return 1;
    const int l = (int)round(m/sqrt(2*c->LbyR));
    int lmin = l - c->ldim/2;
    if(lmin < m)
        lmin = m;
    return lmin;
}

/* Fresnel coefficients at k_ = k(L+R).  Reference: src/libcaps.c:616-645. */
void orc_fresnel(double epsm1, double xi_, double k_, double *rTE, double *rTM)
{
    if(isinf(epsm1)) {
        *rTM = 1.0;
        *rTE = -1.0;
        return;
    }
    const double x = xi_*xi_/(xi_*xi_ + k_*k_)*epsm1;
    const double bm1 = orc_sqrtpm1(x);
    *rTE = -bm1/(2 + bm1);
    *rTM = (epsm1 - bm1)/(epsm1 + 2 + bm1);
}

/* Mie coefficients log|a_l|, log|b_l| at χ = xi_/(1+L/R).
 * Reference: src/libcaps.c:485-520 (perfect reflector) and 557-602 (metal,
 * expressions of Canaguier-Durand et al., PRA 83, 039905 (2011)). */
void orc_mie(const orc_caps_t *c, double xi_, int l, double *lna, double *lnb)
{
    const double epsm1 = orc_epsm1_sphere(c, xi_);
    const double chi = xi_/(1 + c->LbyR);
    const double ln_chi = log(xi_) - log1p(c->LbyR);
    const double lnpi2 = log(M_PI) - log(2.);
    double lIp, lKp, lIm, lKm;
    orc_bessel_logInKn_half(l, chi, &lIp, &lKp);
    orc_bessel_logInKn_half(l - 1, chi, &lIm, &lKm);

    if(isinf(epsm1)) {
        *lnb = lnpi2 + lIp - lKp;
        const double ratio = orc_bessel_ratioI(l - 0.5, chi);
        const double num = lnpi2 + lIp + log(chi*ratio - l);
        const double den = orc_logadd(orc_logi(l) + lKp, ln_chi + lKm);
        *lna = num - den;
        return;
    }

    const double ln_l = orc_logi(l);
    const double ln_n = log1p(epsm1)/2;
    const double n = exp(ln_n);
    double lIp_n, lKp_n, lIm_n, lKm_n;
    orc_bessel_logInKn_half(l, n*chi, &lIp_n, &lKp_n);
    orc_bessel_logInKn_half(l - 1, n*chi, &lIm_n, &lKm_n);
    const double ratio = orc_bessel_ratioI(l - 0.5, chi);
    const double ratio_n = orc_bessel_ratioI(l - 0.5, n*chi);

    const double ln_gB = lIp_n + lIp + ln_chi + log(n*ratio_n - ratio);
    const double ln_gC = log(epsm1) + lIp_n + orc_logadd(ln_chi + lKm, ln_l + lKp);
    const double ln_gD = ln_chi + orc_logadd(lIp_n + lKm, ln_n + lKp + lIm_n);
    const double num = lIp_n + lIp + log(-l*epsm1 + n*chi*(n*ratio - ratio_n));
    *lnb = lnpi2 + ln_gB - ln_gD;
    *lna = lnpi2 + num - orc_logadd(ln_gC, ln_gD);
}

/* ------------------------------------------------------------------------ */
/* K-integral window: Laplace estimate of the peak of the integrand.        */

static double kf(double x, int nu, int m, double alpha)
{
    if(x == 1) {
        if(m != 1)
            return -INFINITY;
        return alpha - orc_logi(nu + 2) - orc_logi(nu + 1) - orc_logi(nu) - orc_logi(nu - 1) + log(8.);
    }
    if(m == 0)
        return alpha*x - orc_lnPlm(nu, 2, x);
    return alpha*x - orc_lnPlm(nu, 2*m, x) + log(x*x - 1);
}

/* Peak position x_max (Newton on f'), Laplace estimate of log K (approx) and
 * the window [a,b] where the integrand exceeds eps times its maximum.
 * Reference: src/integration.c:104-255.  Returns NAN on failure where the
 * reference terminates. */
double orc_K_estimate(int nu, int m, double alpha, double eps, double *a, double *b, double *approx)
{
    const int maxiter = 75;
    const int mpos = m > 0;
    const int m_ = m > 1 ? m : 1;
    double fxmax, fp, fpp = 0, xmax;
    int i;

    if(m == 1) {
        const double threshold = (nu - 2.)*(nu + 3.)/(6*alpha);
        if(threshold < 2) {
            *a = 1;
            *b = 1 - log(eps)/alpha;
            const double logt1 = 1.5*log(alpha) + log(2/M_PI)/2 + orc_bessel_logKn_half(nu, alpha);
            const double logt2 = -alpha + log(alpha + nu*(nu + 1)/2.);
            const double arg = -exp(logt2 - logt1);
            fxmax = alpha - log((nu - 1.)*nu*(nu + 1.)*(nu + 2)/8.);
            xmax = 1;
            if(fabs(arg) < 1)
                *approx = logt1 + log1p(arg);
            else
                *approx = -fxmax;
            goto bordercheck;
        }
    }

    if(nu == 2*m) {
        const int l = nu/2;
        const double ratio = (l - 1)/alpha;
        xmax = ratio + sqrt(1 + ratio*ratio);
    } else
        xmax = sqrt(1 + pow((nu + 0.5)/alpha, 2));

    for(i = 0; i < maxiter; i++) {
        const double xold = xmax, x2m1 = xmax*xmax - 1;
        double d2;
        const double d = orc_dlnPlm(nu, 2*m_, xmax, &d2);
        fp = alpha - d + mpos*2*xmax/x2m1;
        fpp = -d2 - mpos*2*(xmax*xmax + 1)/(x2m1*x2m1);
        xmax = xmax - fp/fpp;
        if(xmax <= 1)
            xmax = 1 + (xold - 1)/2;
        const double delta = fabs(xmax - xold);
        if(delta < 1e-13 || (xmax > 1.001 && delta < 1e-6))
            break;
    }
    if(!isfinite(xmax))
        return NAN;
    fxmax = kf(xmax, nu, m, alpha);
    if(!isfinite(fxmax) || !isfinite(fpp) || fpp < 0)
        return NAN;
    *approx = log(2*M_PI/fpp)/2 - fxmax;
    {
        const double width = -log(eps)/sqrt(fpp);
        *a = fmax(1, xmax - width);
        *b = xmax + width;
    }

bordercheck:
    if(*a > 1) {
        for(i = 0; i < maxiter; i++) {
            if(exp(fxmax - kf(*a, nu, m, alpha)) < eps)
                break;
            *a = 1 + 0.5*(*a - 1);
        }
        if(i == maxiter)
            return NAN;
    }
    for(i = 0; i < maxiter; i++) {
        if(exp(fxmax - kf(*b, nu, m, alpha)) < eps)
            break;
        *b = 1 + 2*(*b - 1);
    }
    if(i == maxiter)
        return NAN;
    return xmax;
}

/* ------------------------------------------------------------------------ */
/* Adaptive Gauss-Legendre quadrature (the oracle's stand-in for QUADPACK). */

#define GL_LO 10
#define GL_HI 20
static double gl_x_lo[GL_LO], gl_w_lo[GL_LO], gl_x_hi[GL_HI], gl_w_hi[GL_HI];
static int gl_ready = 0;

/* Gauss-Legendre nodes on [-1,1] by Newton iteration on P_n (Bonnet). */
static void gl_nodes(int n, double *x, double *w)
{
    for(int i = 0; i < n; i++) {
        double z = cos(M_PI*(i + 0.75)/(n + 0.5)), dp = 0;
        for(int it = 0; it < 100; it++) {
            double p0 = 1, p1 = z;
            for(int k = 2; k <= n; k++) {
                const double p2 = ((2*k - 1)*z*p1 - (k - 1)*p0)/k;
                p0 = p1; p1 = p2;
            }
            dp = n*(z*p1 - p0)/(z*z - 1);
            const double dz = p1/dp;
            z -= dz;
            if(fabs(dz) < 1e-17) break;
        }
        {
            double p0 = 1, p1 = z;
            for(int k = 2; k <= n; k++) {
                const double p2 = ((2*k - 1)*z*p1 - (k - 1)*p0)/k;
                p0 = p1; p1 = p2;
            }
            dp = n*(z*p1 - p0)/(z*z - 1);
        }
        x[i] = z;
        w[i] = 2/((1 - z*z)*dp*dp);
    }
}

static void gl_init(void)
{
    if(gl_ready) return;
    gl_nodes(GL_LO, gl_x_lo, gl_w_lo);
    gl_nodes(GL_HI, gl_x_hi, gl_w_hi);
    gl_ready = 1;
}

typedef double (*orc_fn)(double, void *);

static void gl_panel(orc_fn f, void *arg, double a, double b, double *val, double *err, long *nev)
{
    const double c = 0.5*(a + b), h = 0.5*(b - a);
    double lo = 0, hi = 0;
    for(int i = 0; i < GL_LO; i++) lo += gl_w_lo[i]*f(c + h*gl_x_lo[i], arg);
    for(int i = 0; i < GL_HI; i++) hi += gl_w_hi[i]*f(c + h*gl_x_hi[i], arg);
    *nev += GL_LO + GL_HI;
    *val = h*hi;
    *err = fabs(h*(hi - lo));
}

/* Globally adaptive bisection until Σerr <= max(epsabs, epsrel |I|). */
static double gl_adapt(orc_fn f, void *arg, double a, double b, double epsabs, double epsrel, long *nev)
{
    enum { LIMIT = 2000 };
    static __thread double pa[LIMIT], pb[LIMIT], pv[LIMIT], pe[LIMIT];
    int n = 1;
    pa[0] = a; pb[0] = b;
    gl_panel(f, arg, a, b, &pv[0], &pe[0], nev);
    for(;;) {
        double s = 0, e = 0;
        int worst = 0;
        for(int i = 0; i < n; i++) {
            s += pv[i]; e += pe[i];
            if(pe[i] > pe[worst]) worst = i;
        }
        if(e <= fmax(epsabs, epsrel*fabs(s)) || n + 1 >= LIMIT || pe[worst] <= 1e-300)
            return s;
        const double m = 0.5*(pa[worst] + pb[worst]);
        pa[n] = m; pb[n] = pb[worst];
        pb[worst] = m;
        gl_panel(f, arg, pa[worst], pb[worst], &pv[worst], &pe[worst], nev);
        gl_panel(f, arg, pa[n], pb[n], &pv[n], &pe[n], nev);
        n++;
    }
}

/* ------------------------------------------------------------------------ */
/* One (ξ, m) block.                                                        */

struct orc_block {
    orc_caps_t c;
    int m, lmin, ldim, is_pr;
    double xi_, alpha, epsm1_plate;
    int nu_cap;            /* capacity of the K caches */
    double *K[2];          /* log|K_ν|, NAN = not computed; index ν - 2m */
    signed char *Ks[2];
    int llo, ltab;         /* I table covers l in [llo, llo+ltab) */
    double *I[2];          /* log|I| for l1 >= l2, NAN = not computed */
    signed char *Is[2];
    double *al, *bl;       /* Mie, index l - lmin, NAN = not computed */
    long n_K, n_Kevals, n_I, n_qterms;
};

typedef struct {
    int nu, m, p;
    double alpha, lognorm, factor, epsm1, xi_;
} kint_t;

/* Integrand of K_ν, scaled by the Laplace estimate.
 * Reference: src/integration.c:257-286. */
static double K_integrand(double x, void *arg_)
{
    const kint_t *a = (const kint_t *)arg_;
    x *= a->factor;
    const double x2m1 = (x + 1)*(x - 1);
    double v;
    if(a->m)
        v = exp(-a->lognorm + orc_lnPlm(a->nu, 2*a->m, x) - a->alpha*x - log(x2m1));
    else
        v = exp(-a->lognorm + orc_lnPlm(a->nu, 2, x) - a->alpha*x);
    double rTE, rTM;
    orc_fresnel(a->epsm1, a->xi_, a->xi_*sqrt(x2m1), &rTE, &rTM);
    return (a->p == 0 ? rTE : rTM)*v;
}

/* The substituted tail: x = b + (1-t)/(α t), t in (0,1]. */
typedef struct { kint_t *k; double b; } ktail_t;
static double K_tail(double t, void *arg_)
{
    const ktail_t *a = (const ktail_t *)arg_;
    const double s = (1 - t)/t;
    return K_integrand(a->b + s/a->k->alpha, a->k)/(t*t*a->k->alpha);
}

/* log|K_ν^p| and its sign: K_ν = ∫_1^∞ r_p e^{-αx} P_ν^{2m}(x)/(x²-1) dx
 * (m > 0), ∫_1^∞ r_p e^{-αx} P_ν^2(x) dx (m = 0).  Reference:
 * src/integration.c:288-427: window [a,b] from the Laplace estimate, pieces
 * [1,a], [a,b], [b,∞).  The oracle integrates every piece with adaptive
 * Gauss-Legendre to the block's epsrel (the reference: dqags on [a,b] to
 * 1e-8, dqage GK7-15 on [1,a] only when (a-1) f(a) > I2 epsrel, dqagi on
 * the tail). */
static double K_compute(orc_block_t *B, int nu, int p, orc_sign_t *sign)
{
    gl_init();
    double a, b, lognorm;
    const int m = B->m;
    const double xmax = orc_K_estimate(nu, m, B->alpha, 1e-6, &a, &b, &lognorm);
    if(isnan(xmax)) {
        *sign = 0;
        return NAN;
    }
    if(a < 1.0001)
        a = 1;
    kint_t k = { nu, m, p, B->alpha, lognorm, 1, B->epsm1_plate, B->alpha/2 };
    const double eps = B->c.epsrel;
    long nev = 0;
    const double I2 = gl_adapt(K_integrand, &k, a, b, 0, eps, &nev);
    double I1 = 0;
    if(a > 1)
        I1 = gl_adapt(K_integrand, &k, 1, a, 0.01*eps*fabs(I2), eps, &nev);
    ktail_t tl = { &k, b };
    const double I3 = gl_adapt(K_tail, &tl, 0, 1, 0.01*eps*fabs(I2), eps, &nev);
    const double sum = I1 + I2 + I3;
    B->n_K++;
    B->n_Kevals += nev;
    *sign = sgn(sum);
    return log(fabs(sum)) + lognorm;
}

/* ξ = 0 plasma-model integral (1/ν!) ∫_0^∞ -r_TE(t) e^{-t} t^ν dt with
 * -r_TE = β₋/(2 + β₋), β₋ = sqrt(1 + (2ω/t)²) - 1, ω = ωp (L+R)/c.
 * Reference: src/integration.c:924-1019 (dqags on a window around the
 * peak, dqage GK7-15 left of it, dqagi on the tail, all to epsrel).  The
 * oracle: adaptive Gauss-Legendre on [0, ν], [ν, ν + 40√ν + 40] and the
 * tail t = b + s/(1-s), s in [0, 1), to epsrel. */
typedef struct { int nu; double omega, lp; } pint_t;
typedef struct { const pint_t *p; double b; } ptail_t;

static double plasma_integrand(double t, void *a_)
{
    const pint_t *a = (const pint_t *)a_;
    if(t <= 0)
        return 0;
    const double q = 2*a->omega/t;
    const double bm1 = orc_sqrtpm1(q*q);
    return bm1/(2 + bm1)*exp(a->lp - t + a->nu*log(t));
}

static double plasma_tail(double s, void *a_)
{
    const ptail_t *a = (const ptail_t *)a_;
    if(s >= 1)
        return 0;
    const double u = 1 - s;
    return plasma_integrand(a->b + s/u, (void *)a->p)/(u*u);
}

double orc_integrate_plasma(int nu, double omega, double epsrel)
{
    gl_init();
    pint_t p = { nu, omega, -orc_lfac((unsigned)nu) };
    const double a = nu, b = nu + 40*sqrt((double)nu) + 40;
    long nev = 0;
    const double I2 = gl_adapt(plasma_integrand, &p, a, b, 0, epsrel, &nev);
    const double I1 = gl_adapt(plasma_integrand, &p, 0, a, 0.01*epsrel*fabs(I2), epsrel, &nev);
    ptail_t tl = { &p, b };
    const double I3 = gl_adapt(plasma_tail, &tl, 0, 1, 0.01*epsrel*fabs(I2), epsrel, &nev);
    return I1 + I2 + I3;
}

/* Cached K.  Reference: src/integration.c:452-489 (sign preset: TE -, TM +). */
double orc_block_K(orc_block_t *B, int nu, int p, orc_sign_t *sign)
{
    const int idx = nu - 2*B->m;
    if(idx < 0 || idx >= B->nu_cap) {
        /* outside the block's table: compute without caching */
        return K_compute(B, nu, p, sign);
    }
    if(isnan(B->K[p][idx])) {
        orc_sign_t s;
        B->K[p][idx] = K_compute(B, nu, p, &s);
        B->Ks[p][idx] = s;
    }
    *sign = B->Ks[p][idx];
    return B->K[p][idx];
}

/* eq. (3) of the I recurrence. */
static double alpha_pnnu(double p, double n, double nu)
{
    return ((p*p - (n + nu + 1)*(n + nu + 1))*(p*p - (n - nu)*(n - nu)))/(4*p*p - 1);
}

/* I_{l1 l2}^p = e^{log a0} Σ_q a_q K_{l1+l2-2q} with the a_q recurrence.
 * Reference: src/integration.c:497-609 (eqs. (20),(24),(26)-(30),(35); the
 * 1e±100 rescaling of a_q; the early exit after three consecutive terms
 * below e^{-60} times the first). */
static double I_compute_abs(orc_block_t *B, int l1, int l2, int p, orc_sign_t *sign, double *log_abs);

static double I_compute(orc_block_t *B, int l1, int l2, int p, orc_sign_t *sign)
{
    return I_compute_abs(B, l1, l2, p, sign, NULL);
}

/* The same sum, and (log_abs != NULL) log of Σ_q |a_q K_{l1+l2-2q}| e^{log a0}:
 * the sum of the terms' magnitudes, the scale of its rounding and of the
 * K-integrals' errors (the condition of the sum is that over |I|). */
static double I_compute_abs(orc_block_t *B, int l1, int l2, int p, orc_sign_t *sign, double *log_abs)
{
    const int m_ = B->m > 0 ? B->m : 1;
    const double n = l1, nu = l2, m = m_;
    const double n4 = l1 + l2 - 2*m_;
    const int l12 = l1 + l2;
    int qmax = (l12 - 2*m_)/2;
    if(n < qmax) qmax = (int)n;
    if(nu < qmax) qmax = (int)nu;
    if(qmax < 0) {
        *sign = 0;
        return NAN;
    }
    const double Ap = -2*m*(n - nu)*(n + nu + 1);
    const double log_a0 = orc_lfac(2*l1) - orc_lfac(l1) + orc_lfac(2*l2) - orc_lfac(l2)
        + orc_lfac(l12) - orc_lfac(2*l12) + orc_lfac(l12 - 2*m_) - orc_lfac(l1 - m_) - orc_lfac(l2 - m_);

    double *aq = malloc((qmax + 1)*sizeof(double));
    double *tv = malloc((qmax + 1)*sizeof(double));
    orc_sign_t *ts = malloc((qmax + 1)*sizeof(orc_sign_t));
    orc_sign_t s;
    int q = 0;
    aq[0] = 1;
    tv[0] = orc_block_K(B, l12, p, &s);
    ts[0] = s;
    if(qmax == 0) goto done;

    q = 1;
    aq[1] = (n + nu - 1.5)*(1 - (2*n + 2*nu - 1)/(n4*(n4 - 1))*((m - n)*(m - n + 1)/(2*n - 1) + (m - nu)*(m - nu + 1)/(2*nu - 1)));
    tv[1] = orc_block_K(B, l12 - 2, p, &s) + log(fabs(aq[1]));
    ts[1] = sgn(aq[1])*s;
    if(qmax == 1) goto done;

    q = 2;
    aq[2] = (2*n + 2*nu - 1)*(2*n + 2*nu - 7)/4*((2*n + 2*nu - 3)/(n4*(n4 - 1))*((2*n + 2*nu - 5)/(2*(n4 - 2)*(n4 - 3))
            *((m - n)*(m - n + 1)*(m - n + 2)*(m - n + 3)/(2*n - 1)/(2*n - 3)
            + 2*(m - n)*(m - n + 1)*(m - nu)*(m - nu + 1)/((2*n - 1)*(2*nu - 1))
            + (m - nu)*(m - nu + 1)*(m - nu + 2)*(m - nu + 3)/(2*nu - 1)/(2*nu - 3)) - (m - n)*(m - n + 1)/(2*n - 1)
            - (m - nu)*(m - nu + 1)/(2*nu - 1)) + 0.5);
    tv[2] = orc_block_K(B, l12 - 4, p, &s) + log(fabs(aq[2]));
    ts[2] = sgn(aq[2])*s;
    if(qmax == 2) goto done;

    {
        double log_scaling = 0;
        int count = 0;
        for(q = 3; q <= qmax; q++) {
            const double pp = n + nu - 2*q, p1 = pp - 2*m, p2 = pp + 2*m;
            if(Ap != 0) {
                const double c0 = (pp + 2)*(pp + 3)*(p1 + 1)*(p1 + 2)*Ap*alpha_pnnu(pp + 1, n, nu);
                const double c1 = Ap*(Ap*Ap + (pp + 1)*(pp + 3)*(p1 + 2)*(p2 + 2)*alpha_pnnu(pp + 2, n, nu)
                                  + (pp + 2)*(pp + 4)*(p1 + 3)*(p2 + 3)*alpha_pnnu(pp + 3, n, nu));
                const double c2 = -(pp + 2)*(pp + 3)*(p2 + 3)*(p2 + 4)*Ap*alpha_pnnu(pp + 4, n, nu);
                aq[q] = (c1*aq[q - 1] + c2*aq[q - 2])/c0;
            } else
                aq[q] = (pp + 1)*(p2 + 2)*alpha_pnnu(pp + 2, n, nu)*aq[q - 1]/((pp + 2)*(p1 + 1)*alpha_pnnu(pp + 1, n, nu));

            if(fabs(aq[q]) > 1e100) {
                log_scaling += log(fabs(aq[q]));
                aq[q - 1] /= fabs(aq[q]);
                aq[q] = sgn(aq[q]);
            } else if(fabs(aq[q]) < 1e-100 && fabs(aq[q]) > 0) {
                log_scaling -= log(fabs(aq[q]));
                aq[q - 1] *= fabs(aq[q]);
                aq[q] = sgn(aq[q]);
            }
            tv[q] = log_scaling + orc_block_K(B, l12 - 2*q, p, &s) + log(fabs(aq[q]));
            ts[q] = sgn(aq[q])*s;
            if(tv[q] - tv[0] < -60) {
                if(++count >= 3)
                    break;
            } else
                count = 0;
        }
    }
done:
    B->n_I++;
    {
        const int nt = (q < qmax ? q : qmax) + 1;
        B->n_qterms += nt;
        const double r = log_a0 + orc_logadd_ms(tv, ts, nt, sign);
        if(log_abs) {
            double a = -INFINITY;
            for(int k = 0; k < nt; k++) a = orc_logadd(a, tv[k]);
            *log_abs = log_a0 + a;
        }
        free(aq); free(tv); free(ts);
        return r;
    }
}

/* log Σ_q |a_q K_{l1+l2-2q}| e^{log a0} of I_{l1 l2}^p (TM for perfect
 * reflectors, whose TE is -|TM|); l1 >= l2 >= m.  Test aid: the scale of
 * the sum's absolute error. */
double orc_block_I_abs(orc_block_t *B, int l1, int l2, int p)
{
    if(l1 < l2) { const int t = l1; l1 = l2; l2 = t; }
    if(l2 < B->m) return -INFINITY;
    if(B->is_pr) p = 1;
    orc_sign_t s;
    double a;
    I_compute_abs(B, l1, l2, p, &s, &a);
    return a;
}

/* Cached, symmetrised I.  Reference: src/integration.c:627-671 (zero for
 * l < m; perfect reflectors: I_TE = -|I_TM|; I symmetric in l1,l2). */
double orc_block_I(orc_block_t *B, int l1, int l2, int p, orc_sign_t *sign)
{
    const int m = B->m;
    if(l1 < m || l2 < m) {
        *sign = 0;
        return -INFINITY;
    }
    if(B->is_pr && p == 0) {
        const double v = orc_block_I(B, l1, l2, 1, sign);
        *sign = -1;
        return v;
    }
    if(l1 < l2) { const int t = l1; l1 = l2; l2 = t; }
    if(l2 < B->llo || l1 >= B->llo + B->ltab)
        return I_compute(B, l1, l2, p, sign);
    const size_t idx = (size_t)(l1 - B->llo)*B->ltab + (l2 - B->llo);
    if(isnan(B->I[p][idx])) {
        orc_sign_t s;
        B->I[p][idx] = I_compute(B, l1, l2, p, &s);
        B->Is[p][idx] = s;
    }
    *sign = B->Is[p][idx];
    return B->I[p][idx];
}

/* A, B, C, D.  Reference: src/integration.c:762-888. */
static double int_A(orc_block_t *B, int l1, int l2, int p, orc_sign_t *sign)
{
    if(B->m == 0) { *sign = 0; return -INFINITY; }
    return 2*orc_logi(B->m) + orc_block_I(B, l1, l2, p, sign);
}

static double int_B(orc_block_t *B, int l1, int l2, int p, orc_sign_t *sign)
{
    const int m = B->m;
    if(m == 0)
        return orc_block_I(B, l1, l2, p, sign);
    orc_sign_t s1, s2, s3, s4;
    const double I1 = orc_block_I(B, l1 - 1, l2 - 1, p, &s1);
    const double I2 = orc_block_I(B, l1 + 1, l2 - 1, p, &s2);
    const double I3 = orc_block_I(B, l1 - 1, l2 + 1, p, &s3);
    const double I4 = orc_block_I(B, l1 + 1, l2 + 1, p, &s4);
    const double den = (2*l1 + 1.)*(2*l2 + 1.);
    const double mx = fmax(fmax(I1, I2), fmax(I3, I4));
    double I = (l1 + 1.)*(l1 + m)*(l2 + 1.)*(l2 + m)/den*s1*exp(I1 - mx);
    I -= l1*(l1 - m + 1.)*(l2 + 1.)*(l2 + m)/den*s2*exp(I2 - mx);
    I -= (l1 + 1.)*(l1 + m)*l2*(l2 - m + 1.)/den*s3*exp(I3 - mx);
    I += l1*(l1 - m + 1.)*l2*(l2 - m + 1.)/den*s4*exp(I4 - mx);
    *sign = sgn(I);
    return mx + log(fabs(I));
}

static double int_C(orc_block_t *B, int l1, int l2, int p, orc_sign_t *sign)
{
    const int m = B->m;
    if(m == 0) { *sign = 0; return -INFINITY; }
    orc_sign_t s1, s2;
    const double I1 = orc_block_I(B, l1, l2 - 1, p, &s1);
    const double I2 = orc_block_I(B, l1, l2 + 1, p, &s2);
    const double den = 2*l2 + 1;
    double I = -(l2 + 1.)*(l2 + m)/den*s1*exp(I1 - I2);
    I += l2*(l2 - m + 1.)/den*s2;
    *sign = sgn(I);
    return orc_logi(m) + I2 + log(fabs(I));
}

static double int_D(orc_block_t *B, int l1, int l2, int p, orc_sign_t *sign)
{
    return int_C(B, l2, l1, p, sign);
}

orc_block_t *orc_block_new(const orc_caps_t *c, double xi_, int m)
{
    if(!(xi_ > 0) || m < 0)
        return NULL;
    orc_block_t *B = calloc(1, sizeof(*B));
    B->c = *c;
    B->m = m;
    B->ldim = c->ldim;
    B->lmin = orc_estimate_lmin(c, m);
    B->xi_ = xi_;
    B->alpha = 2*xi_;
    B->epsm1_plate = orc_epsm1_plate(c, xi_);
    B->is_pr = isinf(orc_epsm1_plate(c, INFINITY)) || isinf(c->omegap_plate);
    B->nu_cap = 2*(B->lmin + B->ldim + 2) - 2*m + 4;
    for(int p = 0; p < 2; p++) {
        B->K[p] = malloc(B->nu_cap*sizeof(double));
        B->Ks[p] = malloc(B->nu_cap);
        for(int i = 0; i < B->nu_cap; i++) B->K[p][i] = NAN;
    }
    B->llo = B->lmin > 0 ? B->lmin - 1 : 0;
    B->ltab = B->lmin + B->ldim + 2 - B->llo;
    const size_t nt = (size_t)B->ltab*B->ltab;
    for(int p = 0; p < 2; p++) {
        B->I[p] = malloc(nt*sizeof(double));
        B->Is[p] = malloc(nt);
        for(size_t i = 0; i < nt; i++) B->I[p][i] = NAN;
    }
    B->al = malloc(B->ldim*sizeof(double));
    B->bl = malloc(B->ldim*sizeof(double));
    for(int i = 0; i < B->ldim; i++) B->al[i] = B->bl[i] = NAN;
    return B;
}

void orc_block_free(orc_block_t *B)
{
    if(!B) return;
    for(int p = 0; p < 2; p++) {
        free(B->K[p]); free(B->Ks[p]); free(B->I[p]); free(B->Is[p]);
    }
    free(B->al); free(B->bl);
    free(B);
}

int orc_block_lmin(const orc_block_t *B) { return B->lmin; }

long orc_block_counters(const orc_block_t *B, long *n_K, long *n_Kevals, long *n_I, long *n_qterms)
{
    if(n_K) *n_K = B->n_K;
    if(n_Kevals) *n_Kevals = B->n_Kevals;
    if(n_I) *n_I = B->n_I;
    if(n_qterms) *n_qterms = B->n_qterms;
    return B->n_Kevals;
}

/* Matrix element M^{(m)}_{l1 l2}(p1,p2).  Reference: src/libcaps.c:761-831. */
double orc_block_elem(orc_block_t *B, int l1, int l2, char p1, char p2)
{
    const int lmin = B->lmin;
    if(isnan(B->al[l1 - lmin])) orc_mie(&B->c, B->xi_, l1, &B->al[l1 - lmin], &B->bl[l1 - lmin]);
    if(isnan(B->al[l2 - lmin])) orc_mie(&B->c, B->xi_, l2, &B->al[l2 - lmin], &B->bl[l2 - lmin]);
    const double lnL = orc_lnLambda(l1, l2, B->m);
    const double al1 = B->al[l1 - lmin], bl1 = B->bl[l1 - lmin];
    const double al2 = B->al[l2 - lmin], bl2 = B->bl[l2 - lmin];
    const int TE = 0, TM = 1;
    orc_sign_t s1, s2;
    if(p1 == p2) {
        if(p1 == 'E') {
            const double A = int_A(B, l1, l2, TE, &s1);
            const double Bv = int_B(B, l1, l2, TM, &s2);
            const double mie = (al1 + al2)/2;
            return exp(Bv + mie + lnL)*s2 - exp(A + mie + lnL)*s1;
        } else {
            const double A = int_A(B, l1, l2, TM, &s1);
            const double Bv = int_B(B, l1, l2, TE, &s2);
            const double mie = (bl1 + bl2)/2;
            return exp(A + mie + lnL)*s1 - exp(Bv + mie + lnL)*s2;
        }
    }
    if(B->m == 0)
        return 0;
    if(p1 == 'E') {
        const double C = int_C(B, l1, l2, TE, &s1);
        const double D = int_D(B, l1, l2, TM, &s2);
        const double mie = (al1 + bl2)/2;
        return exp(D + mie + lnL)*s2 - exp(C + mie + lnL)*s1;
    } else {
        const double C = int_C(B, l1, l2, TM, &s1);
        const double D = int_D(B, l1, l2, TE, &s2);
        const double mie = (bl1 + al2)/2;
        return exp(C + mie + lnL)*s1 - exp(D + mie + lnL)*s2;
    }
}

/* Whole matrix, interleaved 2x2 (E,M) index map: row i <-> (l = lmin + i/2,
 * p = i%2).  Reference: src/libcaps.c:717-743 (variant A). */
void orc_block_matrix(orc_block_t *B, double *M)
{
    const int N = 2*B->ldim;
    const char pol[2] = { 'E', 'M' };
    for(int i = 0; i < N; i++)
        for(int j = 0; j < N; j++)
            M[(size_t)i*N + j] = orc_block_elem(B, B->lmin + i/2, B->lmin + j/2, pol[i%2], pol[j%2]);
}
