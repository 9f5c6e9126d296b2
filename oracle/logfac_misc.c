/*
 * Oracle: log n, log n!, log n!! and the log-domain helpers.
 * Test infrastructure only (see oracle.h).
 */
#include <math.h>
#include "oracle.h"

/* log n.  Reference: src/logfac.c:66597-66605 (a table of log n for
 * n < 65536, libm log above).  The table holds correctly rounded log n, so
 * libm log gives the same doubles. */
double orc_logi(unsigned n)
{
    return log((double)n);
}

/* log n!.  Reference: src/logfac.c:66615-66628.  Below 1024 the reference
 * reads a table of log Γ(n+1); from 1024 on it uses Stirling's series cut
 * after the 1/(12n) term.  The oracle keeps that cut exactly (its truncation
 * error, ~1/(360 n^3), is part of the reference's arithmetic). */
double orc_lfac(unsigned n)
{
    if(n < 1024)
        return lgamma((double)n + 1.0);
    const double C = 0.91893853320467274178; /* log(2π)/2 */
    return (n + 0.5)*log((double)n) - n + C + (1./12)/n;
}

/* log n!!.  Reference: src/logfac.c:66637-66651. */
double orc_lfac2(unsigned n)
{
    if(n % 2 == 0) {
        const unsigned k = n/2;
        return k*log(2.) + orc_lfac(k);
    }
    const unsigned k = (n + 1)/2;
    return orc_lfac(2*k) - k*log(2.) - orc_lfac(k);
}

/* Compensated (Kahan) summation.  Reference: src/misc.c:26-40. */
double orc_kahan_sum(const double *x, size_t n)
{
    double s = 0, c = 0;
    for(size_t i = 0; i < n; i++) {
        const double y = x[i] - c;
        const double t = s + y;
        c = (t - s) - y;
        s = t;
    }
    return s;
}

/* sqrt(1+x)-1 without cancellation.  Reference: src/misc.c:56-59. */
double orc_sqrtpm1(double x)
{
    return x/(sqrt(1 + x) + 1);
}

/* log(e^a + e^b) for non-negative numbers.  Reference: src/misc.c:70-81. */
double orc_logadd(double a, double b)
{
    if(isinf(a) && a < 0) return b;
    if(isinf(b) && b < 0) return a;
    if(a > b) return a + log1p(exp(b - a));
    return b + log1p(exp(a - b));
}

/* Signed log-sum: log|Σ s_i e^{v_i}| and its sign, by scaling with the
 * largest exponent.  Reference: src/misc.c:95-109. */
double orc_logadd_ms(const double *v, const orc_sign_t *s, int n, orc_sign_t *sign)
{
    double mx = v[0];
    for(int i = 1; i < n; i++)
        if(v[i] > mx) mx = v[i];
    double sum = 0;
    for(int i = 0; i < n; i++)
        sum += s[i]*exp(v[i] - mx);
    *sign = (sum > 0) - (sum < 0);
    return mx + log(fabs(sum));
}
