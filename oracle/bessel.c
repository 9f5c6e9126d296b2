/*
 * Oracle: modified Bessel functions in the log domain.
 * Test infrastructure only (see oracle.h).
 */
#include <math.h>
#include "oracle.h"

/* log I_ν(x) from the power series Σ_k (x²/4)^k / (k! Γ(k+ν+1)) (DLMF 10.25.2).
 * Reference: src/bessel.c:744-759 (same series, same 1e-16 stop).  Returns
 * NAN if 2048 terms do not converge. */
double orc_bessel_logIn_series(double nu, double x)
{
    const double y = 0.25*x*x;
    double term = 1, sum = 0;
    for(int k = 1; k < 2048; k++) {
        term *= y/(k*(k + nu));
        sum += term;
        if(term < 1e-16*sum)
            return 0.5*nu*log(y) - lgamma(1 + nu) + log1p(sum);
    }
    return NAN;
}

/* log I_n(x), integer n.  Reference: src/bessel.c:583-615.  The reference
 * takes cephes Chebyshev fits for n = 0,1 at x >= 8 and Miller's downward
 * recurrence for large x; the oracle uses the power series throughout.  Its
 * only caller (the Legendre expansion for small arguments, plm.c) passes
 * x < 30, where the series of positive terms is accurate to a few ulp. */
double orc_bessel_logIn(int n, double x)
{
    if(n < 0) n = -n;
    if(x == 0) return n == 0 ? 0 : -INFINITY;
    if(n >= 100) return orc_bessel_logInu_asymp(n, x);
    return orc_bessel_logIn_series(n, x);
}

/* I_ν(x)/I_{ν+1}(x) by the continued fraction of DLMF 10.33.1, evaluated with
 * the modified Lentz recurrences until the ratio stops changing.
 * Reference: src/bessel.c:635-660 (iterate to an exact fixed point).  The
 * oracle caps the iteration count, which the reference does not. */
double orc_bessel_ratioI(double nu, double x)
{
    const double invx2 = 2/x;
    const double a1 = (nu + 1)*invx2;
    const double a2 = (nu + 2)*invx2;
    double num = a2 + 1/a1;
    double den = a2;
    double ratio = a1*num/den;
    double last = 0;
    for(long l = 3; l < 100000000L; l++) {
        const double an = invx2*(nu + l);
        num = an + 1/num;
        den = an + 1/den;
        ratio *= num/den;
        if(ratio == last)
            return ratio;
        last = ratio;
    }
    return NAN;
}

/* Debye polynomials u_k(p)/p^k of the uniform asymptotic expansion, written in
 * t = p² (DLMF 10.41.10, Abramowitz-Stegun 9.3.9-9.3.10), k = 1..5. */
static void debye_u(double t, double u[6])
{
    u[1] = (3 - 5*t)/24.;
    u[2] = (81 + t*(-462 + 385*t))/1152.;
    u[3] = (30375 + t*(-369603 + t*(765765 - 425425*t)))/414720.;
    u[4] = (4465125 + t*(-94121676 + t*(349922430 + t*(-446185740 + 185910725*t))))/39813120.;
    u[5] = (1519035525. + t*(-49286948607. + t*(284499769554. + t*(-614135872350.
           + t*(566098157625. - 188699385875.*t)))))/6688604160.;
}

/* log I_ν(x) by the uniform expansion for large order (DLMF 10.41.3).
 * Reference: src/bessel.c:673-694. */
double orc_bessel_logInu_asymp(double nu, double x)
{
    const double z = x/nu;
    const double a = sqrt(1 + z*z);
    const double t = 1/(1 + z*z);
    double u[6];
    debye_u(t, u);
    const double y = 1/(nu*a);
    const double sum = log1p(y*(u[1] + y*(u[2] + y*(u[3] + y*(u[4] + y*u[5])))));
    const double eta = a + log(z/(1 + a));
    return nu*eta - 0.5*log(2*M_PI*nu*a) + sum;
}

/* log K_ν(x) by the uniform expansion for large order (DLMF 10.41.4).
 * Reference: src/bessel.c:706-726. */
double orc_bessel_logKnu_asymp(double nu, double x)
{
    const double z = x/nu;
    const double a = sqrt(1 + z*z);
    const double t = 1/(1 + z*z);
    double u[6];
    debye_u(t, u);
    const double y = -1/(nu*a);
    const double sum = log1p(y*(u[1] + y*(u[2] + y*(u[3] + y*(u[4] + y*u[5])))));
    const double eta = a + log(z/(1 + a));
    return -nu*eta - 0.5*log(a) + 0.5*log(0.5*M_PI/nu) + sum;
}

/* log I_{n+1/2}(x) and log K_{n+1/2}(x).  Reference: src/bessel.c:781-833:
 * K by the upward recurrence K_{j+1/2} = K_{j-3/2} + (2j-1)/x K_{j-1/2}
 * starting from K_{1/2}, K_{3/2} (exact for half-integer order) with 1e±100
 * rescaling, for n < 100; the uniform expansion for n >= 100.  I follows from
 * the Wronskian I_ν K_{ν+1} + I_{ν+1} K_ν = 1/x and the ratio I_ν/I_{ν+1}. */
void orc_bessel_logInKn_half(int n, double x, double *logI, double *logK)
{
    const double logx = log(x);
    const double invx = 1/x;
    double Kn = 1, Knp = 1 + invx;
    double pref = -x + 0.5*(log(M_PI/2) - logx);
    double logKn, logKnp;

    if(n == 0) {
        logKn = pref + log(Kn);
        logKnp = pref + log(Knp);
    } else if(n >= 100) {
        logKnp = orc_bessel_logKnu_asymp(n + 1.5, x);
        logKn = orc_bessel_logKnu_asymp(n + 0.5, x);
    } else {
        for(int l = 2; l <= n + 1; l++) {
            const double k = (2*l - 1)*Knp*invx + Kn;
            Kn = Knp;
            Knp = k;
            if(Kn > 1e100) {
                Kn *= 1e-100;
                Knp *= 1e-100;
                pref += log(1e100);
            }
        }
        logKnp = pref + log(Knp);
        logKn = pref + log(Kn);
    }
    if(logK) *logK = logKn;
    if(logI) {
        if(n >= 100)
            *logI = orc_bessel_logInu_asymp(n + 0.5, x);
        else
            *logI = -logx - logKnp - log1p(exp(logKn - logKnp - log(orc_bessel_ratioI(n + 0.5, x))));
    }
}

double orc_bessel_logIn_half(int n, double x)
{
    double v;
    orc_bessel_logInKn_half(n, x, &v, NULL);
    return v;
}

double orc_bessel_logKn_half(int n, double x)
{
    double v;
    orc_bessel_logInKn_half(n, x, NULL, &v);
    return v;
}
