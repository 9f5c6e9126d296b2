/*
 * Oracle: the ξ = 0 blocks (perfect reflector / Drude EE and MM, plasma MM)
 * and the high-temperature free energies built from them.
 * Test infrastructure only (see oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include "oracle.h"

#define CAPS_C 299792458.0

/* log of the common ξ = 0 element factor
 * ((l1+l2+1) y + log (l1+l2)! - ½ Σ log (l ± m)!).  Reference:
 * src/libcaps.c:860-869 (EE), 925-941 (MM), 882-906 (plasma MM). */
static double m0_log(const orc_caps_t *c, int l1, int l2, int m)
{
    return (l1 + l2 + 1)*c->y + orc_lfac(l1 + l2)
         - 0.5*(orc_lfac(l1 + m) + orc_lfac(l1 - m) + orc_lfac(l2 + m) + orc_lfac(l2 - m));
}

/* caps_logdetD0: log det(1 - M) of the ldim x ldim blocks EE, MM and the
 * plasma-model MM (omegap in rad/s); NULL outputs are skipped.  Reference:
 * src/libcaps.c:985-1017 (window from caps_estimate_lminmax). */
void orc_logdetD0(const orc_caps_t *c, int m, double omegap, double *EE, double *MM, double *MMp)
{
    const int n = c->ldim, lmin = orc_estimate_lmin(c, m);
    double *M = malloc((size_t)n*n*sizeof(double));
    if(EE) {
        for(int i = 0; i < n; i++)
            for(int j = 0; j < n; j++)
                M[(size_t)i*n + j] = exp(m0_log(c, lmin + i, lmin + j, m));
        *EE = orc_logdet_1mM(M, n, c->detalg);
    }
    if(MM) {
        for(int i = 0; i < n; i++)
            for(int j = 0; j < n; j++) {
                const double l1 = lmin + i, l2 = lmin + j;
                M[(size_t)i*n + j] = exp(m0_log(c, lmin + i, lmin + j, m))*sqrt(l1*l2/((l1 + 1)*(l2 + 1)));
            }
        *MM = orc_logdet_1mM(M, n, c->detalg);
    }
    if(MMp) {
        /* plasma: ω = ωp (L+R)/c, α = ω/(1 + L/R); per l the factor
         * 1 - (2l+1)/(α I_{l-1/2}(α)/I_{l+1/2}(α)) (src/integration.c:897-915,
         * 966-980; src/libcaps.c:896-903) */
        const double omega = omegap*c->calL/CAPS_C, alpha = omega/(1 + c->LbyR);
        double *I = malloc((size_t)(2*n)*sizeof(double)), *f = malloc((size_t)n*sizeof(double));
        for(int k = 0; k < 2*n - 1; k++)
            I[k] = orc_integrate_plasma(2*lmin + k, omega, c->epsrel);
        for(int i = 0; i < n; i++) {
            const double l = lmin + i;
            const double fl = 1 - (2*l + 1)/(alpha*orc_bessel_ratioI(l - 0.5, alpha));
            f[i] = l*fl/(l + 1);
        }
        for(int i = 0; i < n; i++)
            for(int j = 0; j < n; j++)
                M[(size_t)i*n + j] = sqrt(f[i]*f[j])*exp(m0_log(c, lmin + i, lmin + j, m))*I[i + j];
        *MMp = orc_logdet_1mM(M, n, c->detalg);
        free(I);
        free(f);
    }
    free(M);
}

/* ξ = 0 free energy of perfect reflectors (units of k_B T): the Drude term
 * plus Σ_m MM, m = 0 by Bimonte's series, until |v/Σ| < eps.  Reference:
 * src/libcaps.c:1089-1116. */
double orc_ht_perf(const orc_caps_t *c, double eps)
{
    const double LbyR = c->LbyR;
    const double drude = orc_ht_drude(LbyR);
    double MM = 0;
    const double Z = 1/(1 + LbyR + sqrt(LbyR*(2 + LbyR)));
    for(int l = 0; ; l++) {
        const double v = 0.5*log1p(-pow(Z, 2*l + 3));
        MM += v;
        if(fabs(v/MM) < eps)
            break;
    }
    for(int m = 1; ; m++) {
        double v;
        orc_logdetD0(c, m, 0, NULL, &v, NULL);
        MM += v;
        if(fabs(v/MM) < eps)
            return drude + MM;
    }
}

/* ξ = 0 free energy of the plasma model (units of k_B T): the Drude term
 * plus Σ'_m MM_plasma (m = 0 halved) until |v/Σ| < eps.  Reference:
 * src/libcaps.c:1127-1145. */
double orc_ht_plasma(const orc_caps_t *c, double omegap, double eps)
{
    const double drude = orc_ht_drude(c->LbyR);
    double s = 0;
    for(int m = 0; ; m++) {
        double v;
        orc_logdetD0(c, m, omegap, NULL, NULL, &v);
        s += m == 0 ? v/2 : v;
        if(fabs(v/s) < eps)
            return drude + s;
    }
}
