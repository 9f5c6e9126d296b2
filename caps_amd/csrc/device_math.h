// FP64 device math for the round-trip-matrix kernels (gfx950).
//
// log n!, modified Bessel functions of half-integer order and the Mie
// coefficients, evaluated per lane.  Each function follows the CaPS v0.5
// algorithm it cites, so that the GPU values agree with the reference to a
// few ulp of the log-domain results.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace capsd {

constexpr double LOG2 = 0.69314718055994530942;
constexpr double LOGPI = 1.14472988584940017414;
constexpr double HALF_LOG_2PI = 0.91893853320467274178;
constexpr int LFAC_TABLE = 1024;

// log n! : table of log Γ(n+1) below 1024 (filled by the host from lgamma),
// Stirling's series cut after 1/(12n) above.  Reference: src/logfac.c:66615-66628.
__device__ __forceinline__ double lfac(const double* __restrict__ tab, int n)
{
    if (n < LFAC_TABLE) return tab[n];
    const double dn = (double)n;
    return (dn + 0.5) * log(dn) - dn + HALF_LOG_2PI + (1. / 12) / dn;
}

__device__ __forceinline__ double logi(int n) { return log((double)n); }

// log(e^a + e^b), a, b >= -inf.  Reference: src/misc.c:70-81.
__device__ __forceinline__ double logadd(double a, double b)
{
    if (isinf(a) && a < 0) return b;
    if (isinf(b) && b < 0) return a;
    return a > b ? a + log1p(exp(b - a)) : b + log1p(exp(a - b));
}

// Debye polynomials u_k(p)/p^k in t = p², k = 1..5 (DLMF 10.41.10).
__device__ __forceinline__ void debye_u(double t, double& u1, double& u2, double& u3, double& u4, double& u5)
{
    u1 = (3 - 5 * t) / 24.;
    u2 = (81 + t * (-462 + 385 * t)) / 1152.;
    u3 = (30375 + t * (-369603 + t * (765765 - 425425 * t))) / 414720.;
    u4 = (4465125 + t * (-94121676 + t * (349922430 + t * (-446185740 + 185910725 * t)))) / 39813120.;
    u5 = (1519035525. + t * (-49286948607. + t * (284499769554. + t * (-614135872350.
          + t * (566098157625. - 188699385875. * t))))) / 6688604160.;
}

// log I_ν(x), log K_ν(x) by the uniform large-order expansion (DLMF 10.41.3-4).
// Reference: src/bessel.c:673-726.
__device__ inline double logInu_asymp(double nu, double x)
{
    const double z = x / nu, a = sqrt(1 + z * z), t = 1 / (1 + z * z);
    double u1, u2, u3, u4, u5;
    debye_u(t, u1, u2, u3, u4, u5);
    const double y = 1 / (nu * a);
    const double s = log1p(y * (u1 + y * (u2 + y * (u3 + y * (u4 + y * u5)))));
    return nu * (a + log(z / (1 + a))) - 0.5 * log(2 * M_PI * nu * a) + s;
}

__device__ inline double logKnu_asymp(double nu, double x)
{
    const double z = x / nu, a = sqrt(1 + z * z), t = 1 / (1 + z * z);
    double u1, u2, u3, u4, u5;
    debye_u(t, u1, u2, u3, u4, u5);
    const double y = -1 / (nu * a);
    const double s = log1p(y * (u1 + y * (u2 + y * (u3 + y * (u4 + y * u5)))));
    return -nu * (a + log(z / (1 + a))) - 0.5 * log(a) + 0.5 * log(0.5 * M_PI / nu) + s;
}

// I_ν(x)/I_{ν+1}(x) from the continued fraction DLMF 10.33.1 iterated to an
// exact fixed point.  Reference: src/bessel.c:635-660.  The iteration count
// is bounded (the reference's is not): NAN after 2^24 steps.
__device__ inline double ratioI(double nu, double x)
{
    const double invx2 = 2 / x;
    const double a1 = (nu + 1) * invx2, a2 = (nu + 2) * invx2;
    double num = a2 + 1 / a1, den = a2;
    double ratio = a1 * num / den, last = 0;
    for (int l = 3; l < (1 << 24); l++) {
        const double an = invx2 * (nu + l);
        num = an + 1 / num;
        den = an + 1 / den;
        ratio *= num / den;
        if (ratio == last) return ratio;
        last = ratio;
    }
    return NAN;
}

// log I_{n+1/2}(x), log K_{n+1/2}(x).  Reference: src/bessel.c:781-833.
__device__ inline void logInKn_half(int n, double x, double& logI, double& logK)
{
    const double logx = log(x), invx = 1 / x;
    double Kn = 1, Knp = 1 + invx;
    double pref = -x + 0.5 * (log(M_PI / 2) - logx);
    double lKn, lKnp;
    if (n == 0) {
        lKn = pref;
        lKnp = pref + log(Knp);
    } else if (n >= 100) {
        lKnp = logKnu_asymp(n + 1.5, x);
        lKn = logKnu_asymp(n + 0.5, x);
    } else {
        for (int l = 2; l <= n + 1; l++) {
            const double k = (2 * l - 1) * Knp * invx + Kn;
            Kn = Knp;
            Knp = k;
            if (Kn > 1e100) {
                Kn *= 1e-100;
                Knp *= 1e-100;
                pref += log(1e100);
            }
        }
        lKnp = pref + log(Knp);
        lKn = pref + log(Kn);
    }
    logK = lKn;
    if (n >= 100)
        logI = logInu_asymp(n + 0.5, x);
    else
        logI = -logx - lKnp - log1p(exp(lKn - lKnp - log(ratioI(n + 0.5, x))));
}

// log|a_l|, log|b_l| at χ = xi_/(1+L/R); epsm1 = ε(iξ)-1 of the sphere
// (+inf: perfect reflector).  Reference: src/libcaps.c:485-520, 557-602.
__device__ inline void mie(double xi_, double LbyR, double epsm1, int l, double& lna, double& lnb)
{
    const double chi = xi_ / (1 + LbyR);
    const double ln_chi = log(xi_) - log1p(LbyR);
    const double lnpi2 = LOGPI - LOG2;
    double lIp, lKp, lIm, lKm;
    logInKn_half(l, chi, lIp, lKp);
    logInKn_half(l - 1, chi, lIm, lKm);
    if (isinf(epsm1)) {
        lnb = lnpi2 + lIp - lKp;
        const double ratio = ratioI(l - 0.5, chi);
        lna = lnpi2 + lIp + log(chi * ratio - l) - logadd(logi(l) + lKp, ln_chi + lKm);
        return;
    }
    const double ln_l = logi(l);
    const double ln_n = log1p(epsm1) / 2;
    const double n = exp(ln_n);
    double lIp_n, lKp_n, lIm_n, lKm_n;
    logInKn_half(l, n * chi, lIp_n, lKp_n);
    logInKn_half(l - 1, n * chi, lIm_n, lKm_n);
    const double ratio = ratioI(l - 0.5, chi);
    const double ratio_n = ratioI(l - 0.5, n * chi);
    const double ln_gB = lIp_n + lIp + ln_chi + log(n * ratio_n - ratio);
    const double ln_gC = log(epsm1) + lIp_n + logadd(ln_chi + lKm, ln_l + lKp);
    const double ln_gD = ln_chi + logadd(lIp_n + lKm, ln_n + lKp + lIm_n);
    const double num = lIp_n + lIp + log(-l * epsm1 + n * chi * (n * ratio - ratio_n));
    lnb = lnpi2 + ln_gB - ln_gD;
    lna = lnpi2 + num - logadd(ln_gC, ln_gD);
}

// Fresnel coefficients at k_ = xi_ sqrt(x²-1) (x = cosh of the integration
// variable).  Reference: src/libcaps.c:616-645.
__device__ __forceinline__ void fresnel(double epsm1, double xi_, double k_, double& rTE, double& rTM)
{
    if (isinf(epsm1)) {
        rTM = 1.0;
        rTE = -1.0;
        return;
    }
    const double x = xi_ * xi_ / (xi_ * xi_ + k_ * k_) * epsm1;
    const double bm1 = x / (sqrt(1 + x) + 1);
    rTE = -bm1 / (2 + bm1);
    rTM = (epsm1 - bm1) / (epsm1 + 2 + bm1);
}

}  // namespace capsd
