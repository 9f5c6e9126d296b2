/*
 * Oracle: associated Legendre functions P_l^m(x), x > 1, in the log domain,
 * without the Condon-Shortley phase (all values positive).
 * Test infrastructure only (see oracle.h).
 */
#include <math.h>
#include "oracle.h"

/* Upward recurrence in the degree, DLMF 14.10.3, from
 * P_m^m = (2m)!/(2^m m!) (x²-1)^{m/2} (DLMF 14.7.15), rescaled by 1e±100.
 * Reference: src/plm.c:72-112. */
double orc_lnPlm_upwards(int l, int m, double x)
{
    double pref = orc_lfac(2*m) - m*log(2.) - orc_lfac(m) + m/2.*log((x + 1)*(x - 1));
    if(l == m)
        return pref;
    double pm = 1, p = x*(2*m + 1);
    if(p == 0)
        return -INFINITY;
    for(int ll = 2; ll < l + 1 - m; ll++) {
        const double k = (2.*m - 1.)/ll;
        const double pn = (2 + k)*x*p - (1 + k)*pm;
        pm = p;
        p = pn;
        const double a = fabs(p);
        if(a < 1e-100) {
            pref -= log(1e100);
            p *= 1e100;
            pm *= 1e100;
        } else if(a > 1e100) {
            pref += log(1e100);
            p *= 1e-100;
            pm *= 1e-100;
        }
    }
    if(isnan(p))
        return NAN;
    return pref + log(fabs(p));
}

/* log P_l(x) for (l+1) sinh ξ > 25, x = cosh ξ: the large-argument expansion of
 * Bogaert, Michiels and Fostier (SIAM J. Sci. Comput. 34 (2012), §3.2), 17
 * terms.  C_0 = Γ(l+1)/Γ(l+3/2) comes from its asymptotic series in 1/l.
 * Reference: src/plm.c:133-165. */
static double Pl_large(int l, double x, double sinhxi)
{
    const int M = 17;
    const double xi = acosh(x);
    const double expxi = x + sinhxi;
    const double k = 1./l;
    double C = (1 + k*(-3./8 + k*(25./128 + k*(-105./1024 + k*(1659./32768
               + k*(-6237./262144 + k*(50765./4194304)))))))/sqrt((double)l);
    double sum = 0, sinh_m = 1, exp_m = 1;
    for(int m = 0; m < M - 1; m++) {
        sum += C*(exp_m + exp(-(m + 2*l + 1)*xi))/sinh_m;
        sinh_m *= sinhxi;
        exp_m *= expxi;
        C *= (m + 0.5)*(m + 0.5)/((2*m + 2)*(l + 1.5 + m));
    }
    sum += C*(exp_m + exp(-(2*l + M)*xi))/sinh_m;
    return -(log(2.) + log(M_PI))/2 - log(sinhxi)/2 + log(sum) + (l + 0.5)*xi;
}

/* Coefficients f_n of the small-argument Bessel-series expansion of P_l,
 * Bogaert et al. eqs. (3.27)-(3.31), as combinations of h_k = (-y)^k I_k(y).
 * Reference: src/plm.c:168-189. */
static double bogaert_f(int n, const double h[13])
{
    switch(n) {
    case 0:  return h[0];
    case 2:  return h[1]/8. - h[2]/12.;
    case 4:  return 11./384*h[2] - 7./160*h[3] + 1./160*h[4];
    case 6:  return 173./15360*h[3] - 101./3584*h[4] + 671./80640*h[5] - 61./120960*h[6];
    case 8:  return 22931./3440640*h[4] - 90497./3870720*h[5] + 217./20480*h[6]
                  - 1261./967680*h[7] + 1261./29030400*h[8];
    case 10: return 1319183./247726080*h[5] - 10918993./454164480*h[6] + 1676287./113541120*h[7]
                  - 7034857./2554675200*h[8] + 1501./8110080*h[9] - 79./20275200*h[10];
    case 12: return 233526463./43599790080*h[6] - 1396004969./47233105920*h[7]
                  + 2323237523./101213798400*h[8] - 72836747./12651724800*h[9]
                  + 3135577./5367398400*h[10] - 1532789./61993451520*h[11]
                  + 66643./185980354560*h[12];
    default: return 0;
    }
}

/* log P_l(x) for (l+1) sinh ξ <= 25: Bogaert et al. §3.3, series in
 * 1/(l+1/2)² with Bessel-function coefficients.  Reference: src/plm.c:208-233. */
static double Pl_small(int l, double x)
{
    const double v = l + 0.5;
    const double xi = acosh(x);
    const double y = xi*v;
    double h[13], yn = 1;
    for(int n = 0; n < 13; n++) {
        h[n] = yn*exp(orc_bessel_logIn(n, y));
        yn *= -y;
    }
    double s = 0, vn = 1;
    for(int n = 0; n < 13; n += 2) {
        s += bogaert_f(n, h)/vn;
        vn *= v*v;
    }
    return log(s);
}

/* log P_l(x) by Bonnet's recurrence (n+1)P_{n+1} = (2n+1)x P_n - n P_{n-1}.
 * Reference: src/plm.c:251-274. */
static double Pl_recur(int l, double x)
{
    double pref = 0;
    if(l == 0) return 0;
    if(l == 1) return log(x);
    double v0 = 1, v1 = x;
    for(int n = 1; n < l; n++) {
        const double v = ((2*n + 1)*x*v1 - n*v0)/(n + 1);
        v0 = v1;
        v1 = v;
        if(v > 1e100) {
            pref += log(1e100);
            v0 *= 1e-100;
            v1 *= 1e-100;
        }
    }
    return pref + log(v1);
}

/* log P_l(x).  Reference: src/plm.c:299-311 (recurrence below l = 100,
 * expansions above, switching at (l+1) sinh ξ = 25). */
double orc_lnPl(int l, double x)
{
    if(l < 100)
        return Pl_recur(l, x);
    const double sinhxi = sqrt((x + 1)*(x - 1));
    if((l + 1)*sinhxi > 25)
        return Pl_large(l, x, sinhxi);
    return Pl_small(l, x);
}

/* P_l^{m-1}(x)/P_l^m(x) from the continued fraction of DLMF 14.14.1,
 * evaluated by the fundamental recurrences (DLMF 1.12.5-6) in sweeps of 16
 * and stopped when two successive convergents agree exactly.
 * Reference: src/plm.c:328-373 (gives up after 2048 sweeps). */
double orc_Plm_cf(long l, long m, double x)
{
    const double alpha = (1 - 1/(x*x))/4;
    double Am = 0, Bm = 1;
    double A = x*(l - m + 1)*(l + m)*alpha;
    double B = m;
    long j = 1 + m;
    for(int i = 1; i < 2048; i++) {
        const long k = j + 16;
        for(; j < k; j++) {
            const double aj = (l + 1 - j)*(l + j)*alpha;
            const double An = j*A + aj*Am;
            Am = A; A = An;
            const double Bn = j*B + aj*Bm;
            Bm = B; B = Bn;
        }
        if(A*Bm == Am*B)
            return (B*x*sqrt(1 - 1/(x*x)))/(2*A);
        const double inv = 1/B;
        Am *= inv; A *= inv; Bm *= inv; B = 1;
    }
    return NAN;
}

/* log P_l^m(x) from P_l(x), the continued fraction for P_l^{m-1}/P_l^m and
 * the downward recurrence in the order, DLMF 14.10.6.
 * Reference: src/plm.c:391-420. */
double orc_lnPlm_downwards(int l, int m, double x)
{
    double pref = orc_lnPl(l, x);
    if(m == 0)
        return pref;
    const double c = 2*x/sqrt((x - 1)*(x + 1));
    double vp = 1;
    double vm = orc_Plm_cf(l, m, x);
    for(int mm = m - 2; mm >= 0; mm--) {
        const double v = (vp + (mm + 1)*vm*c)/((l + 1. + mm)*(l - mm));
        vp = vm;
        vm = v;
        if(vm < 1e-100) {
            vm *= 1e100;
            vp *= 1e100;
            pref += log(1e100);
        }
    }
    return pref - log(fabs(vm));
}

/* Dispatch.  Reference: src/plm.c:47-55. */
double orc_lnPlm(int l, int m, double x)
{
    if(m == 0)
        return orc_lnPl(l, x);
    if(l - m <= 200)
        return orc_lnPlm_upwards(l, m, x);
    return orc_lnPlm_downwards(l, m, x);
}

/* First and second logarithmic derivatives of P_l^m(x) from the ratios
 * P_l^{m+1}/P_l^m and P_l^{m+2}/P_l^m (DLMF 14.10.5 rewritten).
 * Reference: src/plm.c:436-455. */
double orc_dlnPlm(int l, int m, double x, double *d2)
{
    double p = 0, q = 0;
    const double c2 = (x + 1)*(x - 1);
    const double c = sqrt(c2);
    if(m + 1 <= l)
        p = 1/orc_Plm_cf(l, m + 1, x);
    const double df = p/c + m*x/c2;
    if(d2) {
        if(m + 2 <= l)
            q = p/orc_Plm_cf(l, m + 2, x);
        *d2 = (q - p*p - m*(x*x + 1)/c2)/c2;
    }
    return df;
}
