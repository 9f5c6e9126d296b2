// Round-trip matrix of (ξ, m) blocks on gfx950: Mie coefficients, the
// k-integrals K_ν, the integrals I_{l1 l2} and the matrix 1 - M.
//
// Every kernel is batched: blockIdx.y (or .z) selects the (ξ, m) block, the
// remaining grid dimensions cover the work of one block, lanes beyond a
// block's own extent idle.  All arithmetic is FP64.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "blocks.h"
#include "device_math.h"
#include "wave.h"

namespace capsd {

template <typename T>
__device__ __forceinline__ T* at(char* base, int64_t off) { return reinterpret_cast<T*>(base + off); }

// ---------------------------------------------------------------------------
// Mie coefficients: one lane per l of the block's window.
// Reference: src/libcaps.c:557-602 (caps_mie), cached per block as in
// src/libcaps.c:690-694 and 766-770.
//
// Besides log|a_l|, log|b_l| it stores the per-l halves of the element
// prefactors: log Λ_{l1 l2} = (g(l1) + g(l2))/2 with
// g(l) = log(2l+1) - log l - log(l+1) + log(l-m)! - log(l+m)! (the
// separable form of src/libcaps.c:45-48), so ea(l) = (log|a_l| + g(l))/2 and
// eb(l) = (log|b_l| + g(l))/2 give every element prefactor as a product of
// two factors e^{ea}, e^{eb}, stored as mantissa × 2^exponent.
__global__ void __launch_bounds__(256) k_mie(const BlockDesc* __restrict__ desc, char* __restrict__ arena,
                                             const double* __restrict__ lfac_tab)
{
    const BlockDesc& d = desc[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double* al = at<double>(arena, d.off_mie);   // [a | b | ea | eb], ldim each
    {
        // separable parts of log a0 of the I integrals:
        //   log a0(l1, l2) = g1(l1) + g1(l2) + g2(l1 + l2),
        //   g1(l) = log(2l)! - log l! - log(l-m)!,
        //   g2(s) = log s! - log(2s)! + log(s-2m)!      (m -> max(m, 1))
        const int m_ = d.m > 0 ? d.m : 1;
        double* g1 = al + 5 * d.ldim;
        double* g2 = g1 + d.nl;
        if (i < d.nl) {
            const int l = d.lo + i;
            g1[i] = l >= m_ ? lfac(lfac_tab, 2 * l) - lfac(lfac_tab, l) - lfac(lfac_tab, l - m_) : 0.0;
        }
        if (i < 2 * d.nl - 1) {
            const int sl = 2 * d.lo + i;
            g2[i] = sl >= 2 * m_ ? lfac(lfac_tab, sl) - lfac(lfac_tab, 2 * sl) + lfac(lfac_tab, sl - 2 * m_) : 0.0;
        }
    }
    if (i >= d.ldim) return;
    double lna, lnb;
    const int l = d.lmin + i, m = d.m;
    mie(d.xi_, d.LbyR, d.epsm1_sphere, l, lna, lnb);
    const double g = logi(2 * l + 1) - logi(l) - logi(l + 1) + lfac(lfac_tab, l - m) - lfac(lfac_tab, l + m);
    al[i] = lna;
    al[d.ldim + i] = lnb;
    int* ex = reinterpret_cast<int*>(al + 4 * d.ldim);
    // e^{ea}, e^{eb} as mantissa in [1, 2) times 2^exponent
    const double xa = (lna + g) / 2 * (1 / LOG2), xb = (lnb + g) / 2 * (1 / LOG2);
    const double fa = floor(xa), fb = floor(xb);
    al[2 * d.ldim + i] = exp2(xa - fa);
    al[3 * d.ldim + i] = exp2(xb - fb);
    ex[i] = (int)fa;
    ex[d.ldim + i] = (int)fb;
}

// Mie coefficients at arbitrary (ξ, l) points (scalar API caps_mie).
__global__ void k_mie_points(int n, const double* __restrict__ xi_, const int* __restrict__ l, double LbyR,
                             const double* __restrict__ epsm1, double* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a, b;
    mie(xi_[i], LbyR, epsm1[i], l[i], a, b);
    out[2 * i] = a;
    out[2 * i + 1] = b;
}

// Fresnel coefficients at arbitrary points (scalar API caps_fresnel).
__global__ void k_fresnel_points(int n, const double* __restrict__ xi_, const double* __restrict__ k_,
                                 const double* __restrict__ epsm1, double* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double te, tm;
    fresnel(epsm1[i], xi_[i], k_[i], te, tm);
    out[2 * i] = te;
    out[2 * i + 1] = tm;
}

// log I_{n+1/2}(x), log K_{n+1/2}(x) and I_ν(x)/I_{ν+1}(x) at arbitrary
// points (capsx_bessel_points): out[3i..3i+2].  order[i] < 0 skips the
// half-integer pair (NaN), nu[i] < 0 the ratio.
__global__ void k_bessel_points(int n, const int* __restrict__ order, const double* __restrict__ nu,
                                const double* __restrict__ x, double* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double lI = NAN, lK = NAN, r = NAN;
    if (order[i] >= 0) logInKn_half(order[i], x[i], lI, lK);
    if (nu[i] >= 0) r = ratioI(nu[i], x[i]);
    out[3 * i] = lI;
    out[3 * i + 1] = lK;
    out[3 * i + 2] = r;
}

// ---------------------------------------------------------------------------
// K_ν for all ν of a block.
//
// K_ν^p = ∫_1^∞ r_p(x) e^{-αx} P_ν^μ(x) w(x) dx, μ = 2m (m > 0) or 2 (m = 0),
// w = 1/(x²-1) (m > 0) or 1 (m = 0); the reference integrates each ν
// adaptively (src/integration.c:257-427).  Here all ν share one quadrature
// grid: the trapezoidal rule in v with x(v) = 1 + A² softplus(v)², which
// sends x = 1 to v = -∞ (integrand decays like e^{2v}) and makes the step a
// fixed fraction of the local peak width sqrt(x/α) for large x, so the rule
// converges exponentially for every ν at once.  At each node (one lane) the
// values P_ν^μ(x) for ν = μ, μ+1, ... follow from the upward recurrence in the
// degree (DLMF 14.10.3, the recurrence of src/plm.c:72-112 continued over
// all ν), kept as mantissa × 2^E.  For every ν the wave sums its 64 node
// terms after aligning them to their largest exponent (ldexp, exact): the
// terms of 16 ν at a time go through LDS, four lanes per ν add 16 nodes
// each and combine within their quad (no wave-wide reduction per ν); one
// KPart per (ν, wave) goes to HBM and k_K_combine adds the waves.

__device__ __forceinline__ double plm_step(double c1, double c2, double x, double p, double p_prev)
{
    return c1 * x * p - c2 * p_prev;
}

__global__ void __launch_bounds__(64) k_K_partial(const BlockDesc* __restrict__ desc, char* __restrict__ arena,
                                                  const double* __restrict__ lfac_tab, int nseg)
{
    const BlockDesc& d = desc[blockIdx.y];
    const int wave = blockIdx.x / nseg, seg = blockIdx.x % nseg;
    if (wave >= d.n_waves) return;  // uniform over the workgroup
    const int lane = threadIdx.x;
    const int k = wave * 64 + lane;
    const bool active = k < d.n_nodes;
    const int mu = d.mu;

    // node, weight and Fresnel factors (once per node)
    double x = 2, mant_te = 0, mant_tm = 0;
    int E = -1000000;
    if (active) {
        const double v = d.v0 + k * d.h;
        const double L = v > 40 ? v : log1p(exp(v));
        const double sg = 1 / (1 + exp(-v));
        const double xm1 = d.A * d.A * L * L;
        x = 1 + xm1;
        const double x2m1 = xm1 * (x + 1);
        const double dx = 2 * d.A * d.A * L * sg;
        const double lpmm = lfac(lfac_tab, 2 * mu) - mu * LOG2 - lfac(lfac_tab, mu) + 0.5 * mu * log(x2m1);
        double lw = log(d.h * dx) - d.alpha * x + lpmm;
        if (d.m > 0) lw -= log(x2m1);
        double rte, rtm;
        fresnel(d.epsm1_plate, d.xi_, d.xi_ * sqrt(x2m1), rte, rtm);
        const double lw2 = lw * (1 / LOG2);
        const double fl = floor(lw2);
        E = (int)fl;
        const double w = exp((lw2 - fl) * LOG2);
        mant_te = w * fabs(rte);
        mant_tm = w * fabs(rtm);
    }

    KPart* part = at<KPart>(arena, d.off_kpart);
    // per chunk of 16 ν: each node's term p_ν · 2^E as a mantissa in [1, 2)
    // and an exponent (rows ν, columns node; odd pitch), then per ν the sum
    // over the wave's 64 nodes: lane 4j + q adds nodes 16q .. 16q+15 of the
    // chunk's ν j (exponent-aligned to their largest), the quad combines
    // its four partial sums
    constexpr int CH = 16, TP = 65;
    __shared__ double c1s[64], c2s[64];
    __shared__ double tm_[CH * TP];
    __shared__ int te_[CH * TP];
    __shared__ double2 mant_s[64];   // per node: (|r_TE|, |r_TM|) times the weight mantissa
    mant_s[lane] = make_double2(mant_te, mant_tm);
    const int rj = lane >> 2, rq = lane & 3;   // reduction role: ν j of the chunk, node quarter q
    double p_prev = 0, p = 1;
    constexpr int NONE = -1000000;
    // this workgroup's ν range: segment seg of nseg of the 64-ν chunks (a
    // short grid splits the ν loop, whose steps are latency-bound, over
    // more waves); the recurrence is run alone up to the segment's start,
    // the same steps as the full loop, so the terms are the same bits
    const int nch = (d.nu_max - mu + 64) / 64;
    const int nu_beg = mu + 64 * (int)((int64_t)nch * seg / nseg);
    const int nu_end = min(d.nu_max, mu + 64 * (int)((int64_t)nch * (seg + 1) / nseg) - 1);
    for (int base = mu; base < nu_beg; base += 64) {
        {
            const int nu = base + lane;
            const double ll = (double)(nu + 1 - mu);
            c1s[lane] = (2.0 * nu + 1) / ll;
            c2s[lane] = (double)(nu + mu) / ll;
        }
        __syncthreads();
        for (int jj = 0; jj < 64; jj++) {
            const double pn = plm_step(c1s[jj], c2s[jj], x, p, p_prev);
            p_prev = p;
            p = pn;
            if (p > 0x1p600) {
                p *= 0x1p-600;
                p_prev *= 0x1p-600;
                E += 600;
            }
        }
        __syncthreads();
    }
    for (int base = nu_beg; base <= nu_end; base += 64) {
        // recurrence coefficients for the next 64 steps, one division per lane
        {
            const int nu = base + lane;
            const double ll = (double)(nu + 1 - mu);
            c1s[lane] = (2.0 * nu + 1) / ll;
            c2s[lane] = (double)(nu + mu) / ll;
        }
        __syncthreads();
        const int nend = min(base + 64, nu_end + 1);
        for (int c0 = base; c0 < nend; c0 += CH) {
            // this node's terms of ν = c0 .. c0 + CH - 1; the recurrence
            // gives p at ν+1, kept below 2^600 by moving powers of two into E
#pragma unroll
            for (int j = 0; j < CH; j++) {
                const int e = active ? E + ilogb(p) : NONE;
                tm_[j * TP + lane] = active ? ldexp(p, E - e) : 0.0;
                te_[j * TP + lane] = e;
                const int jj = c0 - base + j;
                if (jj < 64) {
                    const double pn = plm_step(c1s[jj], c2s[jj], x, p, p_prev);
                    p_prev = p;
                    p = pn;
                    if (p > 0x1p600) {
                        p *= 0x1p-600;
                        p_prev *= 0x1p-600;
                        E += 600;
                    }
                }
            }
            __syncthreads();
            // partial sums over the lane's 16 nodes
            const double* rowm = tm_ + rj * TP + 16 * rq;
            const int* rowe = te_ + rj * TP + 16 * rq;
            int F = NONE;
#pragma unroll
            for (int t = 0; t < 16; t++) F = max(F, rowe[t]);
            double ste = 0, stm = 0;
#pragma unroll
            for (int t = 0; t < 16; t++) {
                const double2 mt = mant_s[16 * rq + t];
                const double m = rowm[t];
                const int sh = rowe[t] - F;
                ste += ldexp(m * mt.x, sh);
                stm += ldexp(m * mt.y, sh);
            }
            // the quad's four partial sums, aligned to their largest exponent
            int Fq = max(F, dpp_i<DPP_QUAD_X1>(F));
            Fq = max(Fq, dpp_i<DPP_QUAD_X2>(Fq));
            ste = ldexp(ste, F - Fq);
            stm = ldexp(stm, F - Fq);
            ste += dpp_d<DPP_QUAD_X1>(ste);
            stm += dpp_d<DPP_QUAD_X1>(stm);
            ste += dpp_d<DPP_QUAD_X2>(ste);
            stm += dpp_d<DPP_QUAD_X2>(stm);
            const int nu = c0 + rj;
            if (rq == 0 && nu < nend) {
                KPart kp;
                kp.s_te = ste;
                kp.s_tm = stm;
                kp.e = Fq;
                kp.pad_ = 0;
                part[(int64_t)(nu - mu) * d.n_waves + wave] = kp;
            }
            __syncthreads();   // tm_/te_ are rewritten by the next chunk
        }
    }
}

// Sum the waves' partial sums: log|K_ν|, and K_ν = kmant · 2^kexp with
// |kmant| in [1, 2) carrying the sign (TE negative, TM positive).
__global__ void __launch_bounds__(256) k_K_combine(const BlockDesc* __restrict__ desc, char* __restrict__ arena,
                                                   int* __restrict__ status)
{
    const BlockDesc& d = desc[blockIdx.y];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x < K_GUARD) {
        // the zeroed guard in front of the table (read by k_I's last chunk)
        KEnt z;
        z.m = 0;
        z.e = 0;
        z.pad_ = 0;
        at<KEnt>(arena, d.off_k)[(int)threadIdx.x - K_GUARD] = z;
    }
    if (j >= d.nK) return;
    const KPart* part = at<KPart>(arena, d.off_kpart) + (int64_t)j * d.n_waves;
    int F = -2000000000;
    for (int w = 0; w < d.n_waves; w++)
        if (part[w].s_tm > 0 || part[w].s_te > 0) F = max(F, part[w].e);
    double ste = 0, stm = 0;
    for (int w = 0; w < d.n_waves; w++) {
        ste += ldexp(part[w].s_te, part[w].e - F);
        stm += ldexp(part[w].s_tm, part[w].e - F);
    }
    KEnt* kt = at<KEnt>(arena, d.off_k);                                     // [2][nK]
    double* lnk = at<double>(arena, d.off_k + 2 * sizeof(KEnt) * (int64_t)d.nK);  // [2][nK]
    const double s[2] = {ste, stm};
    for (int pol = 0; pol < 2; pol++) {
        const double v = s[pol];
        KEnt k;
        k.pad_ = 0;
        if (!(v > 0) || !isfinite(v)) {
            if (!(pol == 0 && d.is_pr)) atomicOr(status, 1);
            k.m = 0;
            k.e = 0;
            kt[pol * d.nK + j] = k;
            lnk[pol * d.nK + j] = -INFINITY;
            continue;
        }
        const int e = ilogb(v);
        const double mnt = ldexp(v, -e);
        k.m = pol == 0 ? -mnt : mnt;
        k.e = F + e;
        kt[pol * d.nK + j] = k;
        lnk[pol * d.nK + j] = log(v) + (double)F * LOG2;
    }
}

// ---------------------------------------------------------------------------
// I_{l1 l2}^p for all pairs of the block's table, both polarizations per
// lane (the a_q depend on l1, l2, m only).
//
// I = e^{log a0} Σ_q a_q K_{l1+l2-2q}: eqs. (20), (24), (26)-(30), (35) of the
// reference (src/integration.c:497-609), with its rescaling of a_q (here by
// powers of two) and its early exit after three consecutive terms below
// e^{-60} times the first one (per polarization).  Terms are summed in
// linear arithmetic after aligning the K exponents, as logadd_ms does
// (src/misc.c:95-109).  Symmetric in l1, l2; zero below l = max(m,1);
// perfect reflectors: I_TE = -|I_TM| (src/integration.c:627-671).

// x / y for the a_q recurrence, whose y are products of integers of
// moderate size (normal, far from overflow): the hardware reciprocal
// estimate refined by two Newton-Raphson steps, r <- r (2 - y r), then one
// residual correction of the quotient, q <- q + r (x - y q) (Markstein's
// scheme; within an ulp or two of the rounded quotient, 8 instructions
// against 11 of the IEEE division with its scaling and fix-up steps).
__device__ __forceinline__ double div_nr(double x, double y)
{
    double r = __builtin_amdgcn_rcp(y);
    r = fma(fma(-y, r, 1.0), r, r);
    r = fma(fma(-y, r, 1.0), r, r);
    const double q = x * r;
    return fma(fma(-y, q, x), r, q);
}

// α(p, n, ν) = Nα(p)/Dα(p) with Nα = (p² - (n+ν+1)²)(p² - (n-ν)²),
// Dα = 4p² - 1 (the reference's alpha, src/integration.c:485-488).  The
// recurrence keeps numerator and denominator apart so that one step of
// eq. (35) costs a single division.
struct AlphaND {
    double N, D;
};

__device__ __forceinline__ AlphaND alpha_nd(double p, double s2, double d2)
{
    const double p2 = p * p;
    return {(p2 - s2) * (p2 - d2), 4 * p2 - 1};
}

// One group: up to GM blocks that share (m, ℓ-window, polarizations) — in
// a free energy the blocks of one m at different frequencies — so the a_q of
// a pair are computed once and applied to every block's K table.
// band != 0: only the pairs with i1 - i2 <= 2, which the diagonal of M
// needs (the trace pass); use_skip != 0: blocks whose trace shortcut is
// already decided (out[b].skip) take no part.
// C: steps of the a_q recurrence per chunk (see below)
template <int NPOL, int GM, int C>
__global__ void __launch_bounds__(256) k_I(const BlockDesc* __restrict__ desc, char* __restrict__ arena,
                                           const IGroup* __restrict__ groups, const double* __restrict__ lfac_tab,
                                           int* __restrict__ status, const BlockOut* __restrict__ out, int band,
                                           int use_skip)
{
    static_assert(C >= 3, "the convergence test looks at a chunk's last three terms");
    static_assert(2 * (C - 1) <= K_GUARD, "a chunk reads at most 2 (C - 1) entries below the K table");
    const IGroup& grp = groups[blockIdx.y];
    const int G = grp.count;
    // the group's blocks that take part (uniform over the workgroup)
    bool act[GM];
    bool any = false;
#pragma unroll
    for (int g = 0; g < GM; g++) {
        act[g] = g < G && !(use_skip && out[grp.blk[g]].skip);
        any = any || act[g];
    }
    if (!any) return;
    const BlockDesc& d = desc[grp.blk[0]];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nl = d.nl;
    int i1, i2;
    if (band) {
        if (t >= 3 * (int64_t)nl) return;
        i1 = (int)(t / 3);
        i2 = i1 - (int)(t % 3);
        if (i2 < 0) return;
    } else {
        const int64_t ntri = (int64_t)nl * (nl + 1) / 2;
        if (t >= ntri) return;
        // triangular index -> (i1 >= i2)
        i1 = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
        while ((int64_t)i1 * (i1 + 1) / 2 > t) i1--;
        while ((int64_t)(i1 + 1) * (i1 + 2) / 2 <= t) i1++;
        i2 = (int)(t - (int64_t)i1 * (i1 + 1) / 2);
    }
    const int l1 = d.lo + i1, l2 = d.lo + i2;
    // the table keeps the lower triangle only (row i1 >= column i2, row
    // major): a wave's stores are contiguous, and readers take I(l1, l2) at
    // (max, min)
    const int64_t a = (int64_t)i1 * nl + i2;
    const int64_t nl2 = (int64_t)nl * nl;

    const int m_ = d.m > 0 ? d.m : 1;
    if (l2 < m_ || l2 < d.m) {
        for (int g = 0; g < G; g++) {
            if (!act[g]) continue;
            const BlockDesc& dg = desc[grp.blk[g]];
            double* Im = at<double>(arena, dg.off_i);
            int* Ie = at<int>(arena, dg.off_i + 16 * nl2);
            for (int pol = 0; pol < 2; pol++) {
                Im[pol * nl2 + a] = 0;
                Ie[pol * nl2 + a] = 0;
            }
        }
        return;
    }

    const int mu = d.mu;
    constexpr int POL0 = NPOL == 1 ? 1 : 0;   // pol index 1 = TM; perfect reflector: TM only
    const double n = l1, nu = l2, m = m_;
    const double n4 = l1 + l2 - 2 * m_;
    const int l12 = l1 + l2;
    int qmax = (l12 - 2 * m_) / 2;
    qmax = min(qmax, min(l1, l2));
    const double Ap = -2 * m * (n - nu) * (n + nu + 1);
    const double* g1 = at<double>(arena, d.off_mie) + 5 * d.ldim;   // separable parts of log a0 (k_mie)
    const double* g2 = g1 + nl;
    const double log_a0 = g1[i1] + g1[i2] + g2[i1 + i2];

    // per block and polarization: running sum (exponent-aligned to the q = 0
    // term) and its shift S - E0 (a_q = aq 2^S, E0 the exponent of the q = 0
    // term)
    // the group's K tables (same nK: same m and ℓ-window); polarization pol
    // of block g at kb[g] + pol nK
    const KEnt* kb[GM];
    const int nK = d.nK;
    double sum[GM][NPOL];
    int sh[GM][NPOL];
    // bit g NPOL + pp: that sum has converged (absent blocks count as such)
    unsigned done = 0;
#pragma unroll
    for (int g = 0; g < GM; g++) {
        kb[g] = act[g] ? at<KEnt>(arena, desc[grp.blk[g]].off_k) : nullptr;
#pragma unroll
        for (int pp = 0; pp < NPOL; pp++) {
            sh[g][pp] = 0;
            sum[g][pp] = 0;
            if (act[g]) {
                const KEnt k0 = kb[g][(POL0 + pp) * nK + l12 - mu];
                sh[g][pp] = -k0.e;   // S - E0, S = 0
                sum[g][pp] = k0.m;
            } else {
                done |= 1u << (g * NPOL + pp);
            }
        }
    }
    constexpr unsigned ALL = NPOL * GM >= 32 ? ~0u : (1u << (NPOL * GM)) - 1;

    // q = 1, 2: eqs. (26)-(30), added to every sum
    auto add_term = [&](int q, double aq) {
        const int j = l12 - 2 * q - mu;
#pragma unroll
        for (int g = 0; g < GM; g++)
#pragma unroll
            for (int pp = 0; pp < NPOL; pp++)
                if (act[g]) {
                    const KEnt kv = kb[g][(POL0 + pp) * nK + j];
                    sum[g][pp] += ldexp(aq * kv.m, kv.e + sh[g][pp]);
                }
    };
    double aqm2 = 1, aqm1 = 1, aq = 1;  // a_{q-2}, a_{q-1}, a_q
    double tn = 0, tv = 0, r3 = 0, rn1 = 0, rv1 = 0;
    if (qmax >= 1) {
        // the reciprocals shared between q = 1 and 2
        rn1 = 1 / (2 * n - 1);
        rv1 = 1 / (2 * nu - 1);
        r3 = 1 / (n4 * (n4 - 1));
        tn = (m - n) * (m - n + 1) * rn1;
        tv = (m - nu) * (m - nu + 1) * rv1;
        aq = (n + nu - 1.5) * (1 - (2 * n + 2 * nu - 1) * r3 * (tn + tv));
        add_term(1, aq);
    }
    if (qmax >= 2) {
        aqm2 = 1;
        aqm1 = aq;
        const double rn3 = 1 / (2 * n - 3), rv3 = 1 / (2 * nu - 3), r6 = 1 / (2 * (n4 - 2) * (n4 - 3));
        aq = (2 * n + 2 * nu - 1) * (2 * n + 2 * nu - 7) / 4 * ((2 * n + 2 * nu - 3) * r3
             * ((2 * n + 2 * nu - 5) * r6
             * (tn * (m - n + 2) * (m - n + 3) * rn3 + 2 * tn * tv + tv * (m - nu + 2) * (m - nu + 3) * rv3)
             - tn - tv) + 0.5);
        add_term(2, aq);
    }

    // q >= 3: eq. (35), in chunks of C steps.  A chunk first runs the
    // recurrence for its C steps (a_q of steps past qmax are zero), then
    // adds the C terms to every sum that has not converged, with the C K
    // loads of a sum issued together.  A sum has converged when the last
    // three terms of a chunk are each below 2^-56 of it (less than half an
    // ulp: they leave it unchanged); the lane stops when all have.  Later
    // terms of a converged sum are still added where other lanes of the
    // wave need the chunk for that sum (they are below the same bound).
    // The reference stops a sum after three consecutive terms below e^{-60}
    // times its q = 0 term; where that bound is the larger (a sum that
    // cancels to below 2^56 e^{-60} of its first term) this takes more
    // terms than the reference, never fewer significant ones.
    // α(p+3), α(p+4) of a step are α(p+1), α(p+2) of the step before
    // (p = l1 + l2 - 2q).  aq is renormalized to a power of two at chunk
    // starts, so that a chunk's steps share the scale 2^S.
    const double s2 = (n + nu + 1) * (n + nu + 1), d2 = (n - nu) * (n - nu);
    AlphaND al3 = alpha_nd(n + nu - 3, s2, d2), al4 = alpha_nd(n + nu - 2, s2, d2);
    int S = 0;   // a_q = aq 2^S
    for (int q0 = 3; q0 <= qmax && done != ALL; q0 += C) {
        {
            const double fa = fabs(aq);
            if (fa > 0x1p200 || (fa < 0x1p-200 && fa > 0)) {
                const int e = ilogb(aq);
                aq = ldexp(aq, -e);
                aqm1 = ldexp(aqm1, -e);
                S += e;
#pragma unroll
                for (int g = 0; g < GM; g++)
#pragma unroll
                    for (int pp = 0; pp < NPOL; pp++) sh[g][pp] += e;
            }
        }
        double ac[C];
#pragma unroll
        for (int t = 0; t < C; t++) {
            const int q = q0 + t;
            aqm2 = aqm1;
            aqm1 = aq;
            const double pq = n + nu - 2 * q, p1 = pq - 2 * m, p2 = pq + 2 * m;
            const AlphaND al1 = alpha_nd(pq + 1, s2, d2), al2 = alpha_nd(pq + 2, s2, d2);
            if (Ap != 0) {
                // a_q = (c1 a_{q-1} + c2 a_{q-2}) / c0 of eq. (35) with
                //   c0 = (p+2)(p+3)(p1+1)(p1+2) Ap α1,
                //   c1 = Ap [Ap² + (p+1)(p+3)(p1+2)(p2+2) α2 + (p+2)(p+4)(p1+3)(p2+3) α3],
                //   c2 = -(p+2)(p+3)(p2+3)(p2+4) Ap α4;
                // Ap cancels, and numerator and denominator are multiplied by
                // Dα1 Dα2 Dα3 Dα4
                const double k2 = (pq + 1) * (pq + 3) * (p1 + 2) * (p2 + 2);
                const double k3 = (pq + 2) * (pq + 4) * (p1 + 3) * (p2 + 3);
                const double k4 = (pq + 2) * (pq + 3) * (p2 + 3) * (p2 + 4);
                const double d23 = al2.D * al3.D;
                const double num1 = (Ap * Ap * d23 + k2 * al2.N * al3.D + k3 * al3.N * al2.D) * al4.D;
                const double num2 = k4 * al4.N * d23;
                const double den = (pq + 2) * (pq + 3) * (p1 + 1) * (p1 + 2) * al1.N * d23 * al4.D;
                aq = div_nr((num1 * aqm1 - num2 * aqm2) * al1.D, den);
            } else {
                aq = div_nr((pq + 1) * (p2 + 2) * al2.N * al1.D * aqm1, (pq + 2) * (p1 + 1) * al1.N * al2.D);
            }
            al3 = al1;
            al4 = al2;
            ac[t] = q <= qmax ? aq : 0.0;
        }
        const int jb = l12 - 2 * q0 - mu;   // K index of the chunk's first step
#pragma unroll
        for (int g = 0; g < GM; g++) {
            if (!act[g]) continue;
#pragma unroll
            for (int pp = 0; pp < NPOL; pp++) {
                const unsigned bit = 1u << (g * NPOL + pp);
                if (!__any(!(done & bit))) continue;   // uniform: converged in every lane
                // index jb - 2t; steps past qmax (coefficient 0) may reach
                // 2 (C - 1) entries below the table: the zeroed guard there
                // (polarization 0) or the other polarization's last entries
                const KEnt* kp = kb[g] + (POL0 + pp) * nK + jb;
                KEnt kv[C];
#pragma unroll
                for (int t = 0; t < C; t++) kv[t] = kp[-2 * t];
                double tm[C];
#pragma unroll
                for (int t = 0; t < C; t++) {
                    tm[t] = ldexp(ac[t] * kv[t].m, kv[t].e + sh[g][pp]);
                    sum[g][pp] += tm[t];
                }
                const double lim = fabs(sum[g][pp]) * 0x1p-56;
                if (fabs(tm[C - 1]) < lim && fabs(tm[C - 2]) < lim && fabs(tm[C - 3]) < lim) done |= bit;
            }
        }
    }

    // I = e^{log a0} · sum · 2^{E0} (E0 = S - sh), stored as a signed
    // mantissa times 2^exp at (i1, i2), i1 >= i2
    const double la2 = log_a0 * (1 / LOG2);
    const double fl = floor(la2);
    const double a0m = exp2(la2 - fl);
    const int a0e = (int)fl;
#pragma unroll
    for (int g = 0; g < GM; g++) {
        if (!act[g]) continue;
        const BlockDesc& dg = desc[grp.blk[g]];
        double* Im = at<double>(arena, dg.off_i);                 // [2][nl*nl]
        int* Ie = at<int>(arena, dg.off_i + 16 * nl2);            // [2][nl*nl]
#pragma unroll
        for (int pp = 0; pp < NPOL; pp++) {
            const int pol = POL0 + pp;
            const double sv = sum[g][pp];
            if (!(sv != 0) || !isfinite(sv)) atomicOr(status, 2);
            int e2;
            const double mt = frexp(sv * a0m, &e2);
            const int ex = e2 + a0e + S - sh[g][pp];
            // (plain stores: non-temporal ones leave the waves' partial lines
            // unmerged in L2, 10 % more bytes written to HBM)
            Im[pol * nl2 + a] = mt;
            Ie[pol * nl2 + a] = ex;
            if (NPOL == 1) {
                Im[a] = -fabs(mt);
                Ie[a] = ex;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Matrix assembly: one lane per 2x2 (E,M) tile (l1, l2), 16 x 16 tiles per
// workgroup.  Writes
// mode 0: 1 - M (the matrix the LU factors), mode 1: M, column-major with
// leading dimension lda; row i = 2(l1-lmin) + p1, column j = 2(l2-lmin) + p2
// (the interleaved index map of src/libcaps.c:717-743).  Elements follow
// src/libcaps.c:761-831 and A/B/C/D src/integration.c:762-888.

// Element contributions in linear form: with the prefactor e^p = pm · 2^pe
// (Mie and Λ factors), X·e^p is summed directly from the I values (stored as
// mantissa × 2^exponent),
//   A·e^p = m² I e^p,
//   B·e^p = Σ_k c_k I_k e^p   (4 terms; m = 0: I itself),
//   C·e^p = m [-(l2+1)(l2+m)/(2l2+1) I_1 e^p + l2(l2-m+1)/(2l2+1) I_2 e^p],
// the same quantities as src/integration.c:762-888 followed by the exp of
// src/libcaps.c:784-829, without the intermediate log: each term is a
// product of mantissas and one ldexp.  Terms below the smallest double
// vanish, as they do there.

// Coefficients of B and C, separable in l1 and l2:
//   u(l) = (l+1)(l+m)/(2l+1),  v(l) = l(l-m+1)/(2l+1).
struct UV {
    double u, v;
};

typedef double nt_double2 __attribute__((ext_vector_type(2)));   // 16-byte non-temporal stores

__device__ __forceinline__ UV uv(int l, int m)
{
    const double r = 1.0 / (2 * l + 1);
    return {(l + 1.) * (l + m) * r, l * (l - m + 1.) * r};
}

// flags bit 0: the trace pass: only the diagonal tiles and only the
// diagonal of M (from I pairs with |l1 - l2| <= 2); bit 1: leave out the
// blocks with out[b].skip set; bit 2: the symmetric delta-form LU follows,
// so a mirror tile outside the 64 x 64 diagonal blocks (tile pairs 2k, 2k+1)
// is not written, only its column partial sums (from the lower tile's rows).
// Full pass: workgroup blockIdx.x takes the tile pair (bx, by), bx >= by
// (linear index over the lower triangle of 16 x 16 tiles of (l1, l2)): the
// tile at rows of bx and columns of by and its mirror at rows of by and
// columns of bx, both from one I window (I is symmetric), so every I value
// is read about once.  Trace pass: blockIdx.x = bx = by.
// mode 0 (LU path) writes -M, i.e. 1 - M without its unit diagonal (the LU's
// delta form; k_dominance puts the unit diagonal back where the LU needs
// partial pivoting), and for every column of a tile's 32 rows the sum of
// |M_ij| over its rows i != j into off_cpart (at the tile's row tile);
// mode 1 writes M.
__global__ void __launch_bounds__(256) k_assemble(const BlockDesc* __restrict__ desc, char* __restrict__ arena,
                                                  const double* __restrict__ lfac_tab, int mode,
                                                  const BlockOut* __restrict__ out, int flags)
{
    const BlockDesc& d = desc[blockIdx.z];
    if ((flags & 2) && out[blockIdx.z].skip) return;
    const bool diag_only = flags & 1;
    int bx, by;
    if (diag_only) {
        bx = by = blockIdx.x;
    } else {
        const int64_t t = blockIdx.x;
        bx = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
        while ((int64_t)bx * (bx + 1) / 2 > t) bx--;
        while ((int64_t)(bx + 1) * (bx + 2) / 2 <= t) bx++;
        by = (int)(t - (int64_t)bx * (bx + 1) / 2);
    }
    const int L1 = d.lmin + bx * 16, L2 = d.lmin + by * 16;
    if (L1 >= d.lmin + d.ldim) return;   // uniform over the workgroup (L2 <= L1)
    const int m = d.m;
    // the I values of the workgroup's 16 x 16 tile and its ±1 halo, both
    // polarizations, loaded once into LDS; the window holds I(l1, l2) at
    // [pol][l2 - L2 + 1][l1 - L1 + 1] (rows padded to 19: conflict-free
    // column accesses).  The table keeps I(l1, l2) = I(l2, l1) once, at row
    // max(l1, l2), column min(l1, l2) (row major): below the diagonal
    // (L1 > L2) consecutive lanes take consecutive l2, so that the reads are
    // contiguous
    __shared__ double wm[2][18][19];
    __shared__ int we[2][18][19];
    __shared__ double csum[16][2][16];   // [column pair][parity][row pair]: |M| of the lane's 2 rows
    __shared__ double rsum[16][2][16];   // [row pair][parity][column pair]: |M| of the lane's 2 columns
    {
        const double* Im = at<double>(arena, d.off_i);
        const int* Ie = at<int>(arena, d.off_i + 16 * (int64_t)d.nl * d.nl);
        const int64_t nl2 = (int64_t)d.nl * d.nl;
        const bool below = L1 > L2;
        for (int t = threadIdx.x; t < 2 * 18 * 18; t += 256) {
            const int pol = t / 324, u = (t / 18) % 18, v = t % 18;
            const int r = below ? v : u, c = below ? u : v;
            const int l1 = L1 - 1 + c, l2 = L2 - 1 + r;
            const bool ok = l1 >= d.lo && l2 >= d.lo && l1 < d.lo + d.nl && l2 < d.lo + d.nl && l1 >= m && l2 >= m;
            const int hi = max(l1, l2), lw = min(l1, l2);
            const int64_t idx = pol * nl2 + (int64_t)(hi - d.lo) * d.nl + (lw - d.lo);
            wm[pol][r][c] = ok ? Im[idx] : 0.0;
            we[pol][r][c] = ok ? Ie[idx] : 0;
        }
    }
    __syncthreads();
    const int c0 = threadIdx.x & 15, r0 = threadIdx.x >> 4;
    const bool partials = mode == 0 && !diag_only && d.n_rtiles > 0;
    // flags bit 2: its columns' partial sums are the lower tile's row sums
    // (|M_ij| = |M_ji|)
    const bool skip_mirror = (flags & 4) && !diag_only && bx != by && (bx >> 1) != (by >> 1);
    const bool rows_too = partials && skip_mirror;
    const double* ma = at<double>(arena, d.off_mie) + 2 * d.ldim;   // mantissas of e^{ea}, e^{eb}
    const double* mb = ma + d.ldim;
    const int* xa = reinterpret_cast<const int*>(ma + 2 * d.ldim);  // their exponents
    const int* xb = xa + d.ldim;
    double* Am = at<double>(arena, d.off_mat);
    const int64_t lda = d.lda;
    const double sg = mode == 0 ? -1.0 : 1.0;

    // The 2 x 2 block of matrix rows 2 i2, 2 i2 + 1 and columns 2 j2,
    // 2 j2 + 1 (l1 = lmin + i2, l2 = lmin + j2), written, and for the
    // partial column sums its |M| over its two rows per column (i2 == j2:
    // without the diagonal) into csum[cp][.][rp].  I(l1 + a, l2 + b) is
    // wm[pol][wr + b][wc + a] (tr = 0) or wm[pol][wr + a][wc + b] (tr = 1,
    // the mirror tile: the window of (l2, l1)).
    auto block = [&](int i2, int j2, int wr, int wc, bool tr, int cp, int rp) {
        const int l1 = d.lmin + i2, l2 = d.lmin + j2;
        // I(l1 + a, l2 + b) of polarization pol times the prefactor pm 2^pe
        auto X = [&](int pol, int a, int b, double pm, int pe) {
            const int rr = tr ? wr + a : wr + b, cc = tr ? wc + b : wc + a;
            return ldexp(wm[pol][rr][cc] * pm, we[pol][rr][cc] + pe);
        };
        const double ma1 = ma[i2], mb1 = mb[i2], ma2 = ma[j2], mb2 = mb[j2];
        const int xa1 = xa[i2], xb1 = xb[i2], xa2 = xa[j2], xb2 = xb[j2];
        const int TE = 0, TM = 1;
        const UV c1 = uv(l1, m), c2 = uv(l2, m);
        // A = m² I;  B = u1 u2 I(-1,-1) - v1 u2 I(+1,-1) - u1 v2 I(-1,+1) + v1 v2 I(+1,+1)
        // (m = 0: B = I);  C(l1,l2) = m [-u2 I(0,-1) + v2 I(0,+1)];
        // D(l1,l2) = C(l2,l1) = m [-u1 I(-1,0) + v1 I(+1,0)]   (I symmetric)
        auto A = [&](int pol, double pm, int pe) { return m == 0 ? 0.0 : (double)m * m * X(pol, 0, 0, pm, pe); };
        auto B = [&](int pol, double pm, int pe) {
            if (m == 0) return X(pol, 0, 0, pm, pe);
            return c1.u * c2.u * X(pol, -1, -1, pm, pe) - c1.v * c2.u * X(pol, 1, -1, pm, pe)
                 - c1.u * c2.v * X(pol, -1, 1, pm, pe) + c1.v * c2.v * X(pol, 1, 1, pm, pe);
        };
        auto C = [&](int pol, double pm, int pe) {
            return m * (-c2.u * X(pol, 0, -1, pm, pe) + c2.v * X(pol, 0, 1, pm, pe));
        };
        auto D = [&](int pol, double pm, int pe) {
            return m * (-c1.u * X(pol, -1, 0, pm, pe) + c1.v * X(pol, 1, 0, pm, pe));
        };
        // EE: √(a_l1 a_l2) Λ (B_TM - A_TE);  MM: √(b_l1 b_l2) Λ (A_TM - B_TE)
        const double pEE = ma1 * ma2, pMM = mb1 * mb2;
        const int eEE = xa1 + xa2, eMM = xb1 + xb2;
        const double EE = B(TM, pEE, eEE) - A(TE, pEE, eEE);
        const double MM = A(TM, pMM, eMM) - B(TE, pMM, eMM);
        if (diag_only) {
            double* dg = at<double>(arena, d.off_diag);
            dg[2 * i2] = EE;
            dg[2 * i2 + 1] = MM;
            return;
        }
        double EM = 0, ME = 0;
        if (m != 0) {
            // EM: D_TM - C_TE with (a_l1 + b_l2)/2;  ME: C_TM - D_TE with (b_l1 + a_l2)/2
            const double pEM = ma1 * mb2, pME = mb1 * ma2;
            const int eEM = xa1 + xb2, eME = xb1 + xa2;
            EM = D(TM, pEM, eEM) - C(TE, pEM, eEM);
            ME = C(TM, pME, eME) - D(TE, pME, eME);
        }
        const int64_t r = 2 * i2, c = 2 * j2;
        // rows r, r+1 of a column are adjacent: one 16-byte store per column
        // (r and lda even, the matrix 256-byte aligned)
        // (non-temporal: the matrices of a batch are far larger than the
        // caches, and these lines would only evict the I windows)
        __builtin_nontemporal_store((nt_double2){sg * EE, sg * ME}, reinterpret_cast<nt_double2*>(Am + r + c * lda));
        __builtin_nontemporal_store((nt_double2){sg * EM, sg * MM},
                                    reinterpret_cast<nt_double2*>(Am + r + (c + 1) * lda));
        if (i2 == j2) {
            double* dg = at<double>(arena, d.off_diag);
            dg[r] = EE;
            dg[r + 1] = MM;
        }
        if (partials) {
            // column c: EE (row r, the diagonal when i2 = j2) and ME (row r+1);
            // column c+1: EM (row r) and MM (row r+1, the diagonal when i2 = j2)
            csum[cp][0][rp] = (i2 == j2 ? 0.0 : fabs(EE)) + fabs(ME);
            csum[cp][1][rp] = fabs(EM) + (i2 == j2 ? 0.0 : fabs(MM));
            if (rows_too) {
                // row r: EE, EM; row r+1: ME, MM (no diagonal in such a tile)
                rsum[rp][0][cp] = fabs(EE) + fabs(EM);
                rsum[rp][1][cp] = fabs(ME) + fabs(MM);
            }
        }
    };
    // the 32 columns of a tile at column tile ct, row tile rt: the sum over
    // its 32 rows in row order
    auto column_sums = [&](int rt, int ct) {
        if (threadIdx.x < 32) {
            const int jp = threadIdx.x >> 1, par = threadIdx.x & 1;
            const int col = 2 * (ct * 16 + jp) + par;
            if (col < d.N) {
                double sum = 0;
#pragma unroll
                for (int q = 0; q < 16; q++) sum += csum[jp][par][q];
                at<double>(arena, d.off_cpart)[(int64_t)rt * d.N + col] = sum;
            }
        }
    };

    // the tile (rows of bx, columns of by): lane (c0, r0) takes i2 = 16 bx + c0,
    // j2 = 16 by + r0
    {
        const int i2 = bx * 16 + c0, j2 = by * 16 + r0;
        const bool in = i2 < d.ldim && j2 < d.ldim && (!diag_only || i2 == j2);
        if (in) block(i2, j2, r0 + 1, c0 + 1, false, r0, c0);
        else if (partials) {
            csum[r0][0][c0] = 0.0;
            csum[r0][1][c0] = 0.0;
            if (rows_too) {
                rsum[c0][0][r0] = 0.0;
                rsum[c0][1][r0] = 0.0;
            }
        }
        if (diag_only) return;   // no barrier follows in the trace pass
        if (partials) {
            __syncthreads();
            column_sums(bx, by);
        }
    }
    if (bx == by) return;   // uniform: a diagonal tile is its own mirror
    if (skip_mirror) {
        // the mirror's 32 columns (rows of bx) over the 32 rows of by: the
        // lower tile's row sums in column order
        if (rows_too && threadIdx.x >= 32 && threadIdx.x < 64) {
            const int ip = (threadIdx.x - 32) >> 1, par = threadIdx.x & 1;
            const int col = 2 * (bx * 16 + ip) + par;
            if (col < d.N) {
                double sum = 0;
#pragma unroll
                for (int q = 0; q < 16; q++) sum += rsum[ip][par][q];
                at<double>(arena, d.off_cpart)[(int64_t)by * d.N + col] = sum;
            }
        }
        return;
    }
    // the mirror (rows of by, columns of bx): lane (c0, r0) takes
    // i2 = 16 by + c0, j2 = 16 bx + r0, whose I(l1 + a, l2 + b) is the
    // window's I(l2 + b, l1 + a)
    {
        if (partials) __syncthreads();   // csum is reused
        const int i2 = by * 16 + c0, j2 = bx * 16 + r0;
        const bool in = i2 < d.ldim && j2 < d.ldim;
        if (in) block(i2, j2, c0 + 1, r0 + 1, true, r0, c0);
        else if (partials) {
            csum[r0][0][c0] = 0.0;
            csum[r0][1][c0] = 0.0;
        }
        if (partials) {
            __syncthreads();
            column_sums(by, bx);
        }
    }
}

// Column diagonal dominance of 1 - M (the matrix holds -M, the column sums
// of |M_ij|, i != j, per row tile are in off_cpart): column j is dominant if
// Σ_{i≠j} |M_ij| < |1 - M_jj|.  A matrix whose columns are all strictly
// dominant keeps that property in every Schur complement, so LU with partial
// pivoting exchanges no rows and the elimination without pivot search is the
// same factorization (growth factor <= 2).  out[b].delta starts at 1 (k_trace)
// and any non-dominant column clears it (an AND, independent of the order);
// allow = 0 clears it for every block (the pivoted LU for all, a test hook).
__global__ void __launch_bounds__(256) k_dominance(const BlockDesc* __restrict__ desc, char* __restrict__ arena,
                                                   BlockOut* __restrict__ out, int allow)
{
    const int b = blockIdx.y;
    const BlockDesc& d = desc[b];
    if (out[b].skip || d.n_rtiles <= 0) return;
    const int j = blockIdx.x * 256 + threadIdx.x;
    bool ok = true;
    if (j < d.N) {
        const double* cp = at<double>(arena, d.off_cpart) + j;
        double s = 0;
        for (int t = 0; t < d.n_rtiles; t++) s += cp[(int64_t)t * d.N];
        const double djj = at<double>(arena, d.off_mat)[j + (int64_t)j * d.lda];   // -M_jj
        ok = allow && s < fabs(1.0 + djj);
    }
    if (__syncthreads_or(!ok) && threadIdx.x == 0) atomicAnd(&out[b].delta, 0);
}

// Blocks left without column dominance get their unit diagonal back: the
// matrix becomes 1 - M (the same values as writing 1 - M directly) for the LU
// with partial pivoting.
__global__ void __launch_bounds__(256) k_dominance_fix(const BlockDesc* __restrict__ desc, char* __restrict__ arena,
                                                       const BlockOut* __restrict__ out)
{
    const int b = blockIdx.y;
    const BlockDesc& d = desc[b];
    if (out[b].skip || d.n_rtiles <= 0 || out[b].delta) return;
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j < d.N) at<double>(arena, d.off_mat)[j + (int64_t)j * d.lda] += 1.0;
}

// ξ = 0 blocks (EE, MM, plasma MM; src/libcaps.c:860-941): elements
// M_ij = exp(h_i + h_j + T_{i+j}) with per-row terms h (n doubles) and
// per-(l1+l2) terms T (2n - 1 doubles), both at arena offset off_tab.
// Writes -M column-major (leading dimension lda; the LU's delta form, see
// k_assemble), the diagonal of M, and per 64-row tile (blockIdx.x) the
// column sums of |M_ij|, i != j.  Lanes run down the columns (coalesced
// stores), 64 x 4 per workgroup.
__global__ void __launch_bounds__(256) k_fill_xi0(const BlockDesc* __restrict__ desc, char* __restrict__ arena,
                                                  int64_t off_tab)
{
    const BlockDesc& d = desc[0];
    const int n = d.N;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63), j = blockIdx.y * 4 + (threadIdx.x >> 6);
    double v = 0;
    if (i < n && j < n) {
        const double* h = at<double>(arena, off_tab);
        const double* T = h + n;
        v = exp(h[i] + h[j] + T[i + j]);
        at<double>(arena, d.off_mat)[i + (int64_t)j * d.lda] = -v;
        if (i == j) at<double>(arena, d.off_diag)[i] = v;
    }
    // all 64 lanes of the wavefront take part in the reduction
    const double s = wave_sum(i == j ? 0.0 : fabs(v));
    if ((threadIdx.x & 63) == 0 && j < n) at<double>(arena, d.off_cpart)[(int64_t)blockIdx.x * n + j] = s;
}

// Trace of M (Kahan, in index order as src/matrix.c:63-68) and the trace
// shortcut flag |tr M| < 1e-8 (src/matrix.c:78-82).  One wavefront per block.
// Also starts the block's delta flag at 1 (cleared by k_dominance).
__global__ void __launch_bounds__(64) k_trace(const BlockDesc* __restrict__ desc, char* __restrict__ arena,
                                             BlockOut* __restrict__ out, int nblocks)
{
    const int b = blockIdx.x;   // one wavefront per block
    if (b >= nblocks) return;
    const BlockDesc& d = desc[b];
    const double s = wave_kahan_sum(at<double>(arena, d.off_diag), d.N);
    if (threadIdx.x == 0) {
        out[b].trace = s;
        // skip 1: trace shortcut; 3: non-finite trace (no LU, NaN result)
        out[b].skip = fabs(s) < 1e-8 ? 1 : (isfinite(s) ? 0 : 3);
        out[b].logdet = NAN;
        out[b].status = 0;
        out[b].delta = d.n_rtiles > 0 ? 1 : 0;
    }
}

// ||M||_F² and tr M² = Σ_ij M_ij M_ji of M (from the stored matrix, -M off
// the diagonal in both forms, and the exact diagonal of M) and the LU-mode
// Mercator shortcut of src/matrix.c:489-505: if ||M||_F < 1 and the bound on
// the neglected terms, |n²/2 + n + log(1-n)|, is below 1e-8 of
// -tr M - tr M²/2, the block's LU is skipped (skip = 2) and the host returns
// the two-term series.  lower != 0 (CHOLESKY): both sums over the matrix the
// reference builds for that mode, M with its upper triangle left at zero
// (src/matrix.c:93-102), i.e. ||M||_F² over i >= j and tr M² over the
// diagonal.  One workgroup per block, over pairs of 64 x 64 tiles (I, J),
// I <= J: both tiles read coalesced (one column segment per load) into LDS,
// where the transposed products m_ij m_ji are formed.
__global__ void __launch_bounds__(256) k_frobenius(const BlockDesc* __restrict__ desc, char* __restrict__ arena,
                                                   BlockOut* __restrict__ out, int lower)
{
    const BlockDesc& d = desc[blockIdx.x];
    if (out[blockIdx.x].skip == 1) return;   // trace shortcut already taken
    const double* A = at<double>(arena, d.off_mat);
    const double* dg = at<double>(arena, d.off_diag);
    const int N = d.N, nt = (N + 63) / 64;
    const int64_t lda = d.lda;
    __shared__ double T1[64 * 65], T2[64 * 65];   // tiles (I, J) and (J, I), [row][column]
    const int r = threadIdx.x & 63, w = threadIdx.x >> 6;
    auto m = [&](int i, int j) { return i >= N || j >= N ? 0.0 : i == j ? dg[i] : -A[i + (int64_t)j * lda]; };
    double f = 0, t2 = 0;
    for (int I = 0; I < nt; I++)
        for (int J = I; J < nt; J++) {
#pragma unroll 4
            for (int s2 = 0; s2 < 16; s2++) {
                const int c = w + 4 * s2;
                T1[r * 65 + c] = m(64 * I + r, 64 * J + c);
                if (J != I) T2[r * 65 + c] = m(64 * J + r, 64 * I + c);
            }
            __syncthreads();
#pragma unroll 4
            for (int s2 = 0; s2 < 16; s2++) {
                const int c = w + 4 * s2;   // element (64I + r, 64J + c) and its mirror
                const double a = T1[r * 65 + c];
                if (J == I) {
                    if (!lower) {
                        f += a * a;
                        t2 += a * T1[c * 65 + r];
                    } else if (r >= c) {
                        f += a * a;
                        if (r == c) t2 += a * a;
                    }
                } else {
                    const double b = T2[c * 65 + r];   // element (64J + c, 64I + r)
                    if (!lower) {
                        f += a * a + b * b;
                        t2 += 2 * (a * b);
                    } else {
                        f += b * b;   // below the diagonal: rows of tile J > columns of tile I
                    }
                }
            }
            __syncthreads();
        }
    __shared__ double sf[256], st[256];
    sf[threadIdx.x] = f;
    st[threadIdx.x] = t2;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            sf[threadIdx.x] += sf[threadIdx.x + o];
            st[threadIdx.x] += st[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        BlockOut& o = out[blockIdx.x];
        o.frob2 = sf[0];
        o.tr2 = st[0];
        const double norm = sqrt(sf[0]);
        if (!o.skip && norm < 1) {
            const double mercator = -o.trace - st[0] / 2;
            const double err = fabs(norm * norm / 2 + norm + log1p(-norm));
            if (fabs(err / mercator) < 1e-8) o.skip = 2;
        }
    }
}

}  // namespace capsd

// ---------------------------------------------------------------------------
// Launchers (C linkage, used by the host library).
using namespace capsd;

extern "C" {

void capsd_launch_mie(const BlockDesc* desc, char* arena, const double* lfac_tab, int nblocks, int max_ldim,
                      hipStream_t s)
{
    dim3 g((2 * (max_ldim + 2) + 255) / 256, nblocks);
    hipLaunchKernelGGL(k_mie, g, dim3(256), 0, s, desc, arena, lfac_tab);
}

void capsd_launch_K(const BlockDesc* desc, char* arena, const double* lfac_tab, int nblocks, int max_waves,
                    int max_nK, int nseg, int* status, hipStream_t s)
{
    hipLaunchKernelGGL(k_K_partial, dim3(max_waves * nseg, nblocks), dim3(64), 0, s, desc, arena, lfac_tab, nseg);
    hipLaunchKernelGGL(k_K_combine, dim3((max_nK + 255) / 256, nblocks), dim3(256), 0, s, desc, arena, status);
}

// I tables of the batch's groups, ordered by class: perfect reflectors
// (TM only) with at most 1, 2, 4, 8 blocks per group, then both
// polarizations likewise; counts[8] holds the groups per class.
// band: only the pairs the diagonal of M needs (max_ntri is then the
// largest nl); use_skip: leave out the blocks with out[b].skip set.
void capsd_launch_I(const BlockDesc* desc, char* arena, const void* groups, const int* counts,
                    const double* lfac_tab, int64_t max_ntri, int* status, const BlockOut* out, int band,
                    int use_skip, hipStream_t s)
{
    const IGroup* gr = (const IGroup*)groups;
    const unsigned gx = (unsigned)(((band ? 3 * max_ntri : max_ntri) + 255) / 256);
    // recurrence steps per chunk: 4 (default) or 8 (CAPS_AMD_I_CHUNK=8)
    static const int chunk = [] {
        const char* e = getenv("CAPS_AMD_I_CHUNK");
        return e && atoi(e) == 8 ? 8 : 4;
    }();
    int off = 0;
#define KI(NP, GMV, CLS)                                                                                       \
    if (counts[CLS] > 0) {                                                                                    \
        if (chunk == 8)                                                                                       \
            hipLaunchKernelGGL((k_I<NP, GMV, 8>), dim3(gx, counts[CLS]), dim3(256), 0, s, desc, arena,       \
                               gr + off, lfac_tab, status, out, band, use_skip);                              \
        else                                                                                                  \
            hipLaunchKernelGGL((k_I<NP, GMV, 4>), dim3(gx, counts[CLS]), dim3(256), 0, s, desc, arena,       \
                               gr + off, lfac_tab, status, out, band, use_skip);                              \
    }                                                                                                         \
    off += counts[CLS];
    KI(1, 1, 0) KI(1, 2, 1) KI(1, 4, 2) KI(1, 8, 3)
    KI(2, 1, 4) KI(2, 2, 5) KI(2, 4, 6) KI(2, 8, 7)
#undef KI
}

// flags: see k_assemble (1 = diagonal only, 2 = leave out skipped blocks)
void capsd_launch_assemble(const BlockDesc* desc, char* arena, const double* lfac_tab, int nblocks, int max_ldim,
                           int mode, const BlockOut* out, int flags, hipStream_t s)
{
    const unsigned T = (unsigned)((max_ldim + 15) / 16);
    dim3 g((flags & 1) ? T : T * (T + 1) / 2, 1, nblocks);
    hipLaunchKernelGGL(k_assemble, g, dim3(256), 0, s, desc, arena, lfac_tab, mode, out, flags);
}

void capsd_launch_frobenius(const BlockDesc* desc, char* arena, BlockOut* out, int nblocks, int lower, hipStream_t s)
{
    hipLaunchKernelGGL(k_frobenius, dim3(nblocks), dim3(256), 0, s, desc, arena, out, lower);
}

void capsd_launch_fill_xi0(const BlockDesc* desc, char* arena, int64_t off_tab, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_xi0, dim3((n + 63) / 64, (n + 3) / 4), dim3(256), 0, s, desc, arena, off_tab);
}

// Column dominance of the blocks' 1 - M (after their assembly): the delta
// flags, and the unit diagonal written back where a block is not dominant.
void capsd_launch_dominance(const BlockDesc* desc, char* arena, BlockOut* out, int nblocks, int maxN, int allow,
                            hipStream_t s)
{
    const dim3 g((maxN + 255) / 256, nblocks);
    hipLaunchKernelGGL(k_dominance, g, dim3(256), 0, s, desc, arena, out, allow);
    hipLaunchKernelGGL(k_dominance_fix, g, dim3(256), 0, s, desc, arena, out);
}

void capsd_launch_trace(const BlockDesc* desc, char* arena, BlockOut* out, int nblocks, int frobenius, hipStream_t s)
{
    hipLaunchKernelGGL(k_trace, dim3(nblocks), dim3(64), 0, s, desc, arena, out, nblocks);
    if (frobenius) hipLaunchKernelGGL(k_frobenius, dim3(nblocks), dim3(256), 0, s, desc, arena, out, frobenius == 2);
}

void capsd_launch_mie_points(int n, const double* xi_, const int* l, double LbyR, const double* epsm1, double* out,
                             hipStream_t s)
{
    hipLaunchKernelGGL(k_mie_points, dim3((n + 63) / 64), dim3(64), 0, s, n, xi_, l, LbyR, epsm1, out);
}

void capsd_launch_bessel_points(int n, const int* order, const double* nu, const double* x, double* out,
                                hipStream_t s)
{
    hipLaunchKernelGGL(k_bessel_points, dim3((n + 63) / 64), dim3(64), 0, s, n, order, nu, x, out);
}

void capsd_launch_fresnel_points(int n, const double* xi_, const double* k_, const double* epsm1, double* out,
                                 hipStream_t s)
{
    hipLaunchKernelGGL(k_fresnel_points, dim3((n + 63) / 64), dim3(64), 0, s, n, xi_, k_, epsm1, out);
}

}  // extern "C"
