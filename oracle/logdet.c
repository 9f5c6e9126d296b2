/*
 * Oracle: log det(1-M) of a dense block and the whole caps_logdetD path.
 * Test infrastructure only (see oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* log|det A| by LU with partial pivoting (row-major, right-looking), the
 * log|u_ii| summed with Kahan.  Reference: src/matrix.c:555-577 (dgetrf_) and
 * 462-471 (Kahan Σ log|U_ii|).  Returns NAN for a singular matrix. */
double orc_logdet_lu(double *A, int n)
{
    double *logs = malloc((size_t)n*sizeof(double));
    for(int k = 0; k < n; k++) {
        int piv = k;
        double best = fabs(A[(size_t)k*n + k]);
        for(int i = k + 1; i < n; i++) {
            const double v = fabs(A[(size_t)i*n + k]);
            if(v > best) { best = v; piv = i; }
        }
        if(best == 0) { free(logs); return NAN; }
        if(piv != k)
            for(int j = 0; j < n; j++) {
                const double t = A[(size_t)k*n + j];
                A[(size_t)k*n + j] = A[(size_t)piv*n + j];
                A[(size_t)piv*n + j] = t;
            }
        const double d = A[(size_t)k*n + k];
        logs[k] = log(fabs(d));
        const double inv = 1/d;
        for(int i = k + 1; i < n; i++) {
            double *ri = A + (size_t)i*n;
            const double l = ri[k]*inv;
            ri[k] = l;
            if(l == 0) continue;
            const double *rk = A + (size_t)k*n;
            for(int j = k + 1; j < n; j++)
                ri[j] -= l*rk[j];
        }
    }
    const double r = orc_kahan_sum(logs, n);
    free(logs);
    return r;
}

/* log|1 + d| keeping the relative accuracy of d (a pivot u = 1 + d near 1
 * stored as u itself loses d below one ulp of 1). */
static double log_abs_1p(double d) { return d > -0.5 ? log1p(d) : log(fabs(1 + d)); }

/* 1 if 1 - M (row-major M) is strictly column diagonally dominant:
 * sum_{i != j} |M_ij| < |1 - M_jj| for every column j.  Then every Schur
 * complement is too, and LU with partial pivoting exchanges no rows. */
int orc_dominant_1mM(const double *M, int n)
{
    for(int j = 0; j < n; j++) {
        double s = 0;
        for(int i = 0; i < n; i++)
            if(i != j) s += fabs(M[(size_t)i*n + j]);
        if(!(s < fabs(1 - M[(size_t)j*n + j])))
            return 0;
    }
    return 1;
}

/* log|det(1 - M)| for column diagonally dominant 1 - M, in delta form: the
 * elimination runs on D = -M (1 - M without its unit diagonal) with pivot
 * u_k = 1 + d_kk and no row exchanges (partial pivoting would make none),
 * every update being the same on D as on 1 - M; log|u_k| = log1p(d_kk),
 * Kahan-summed.  The reference factors 1 - M with LAPACK (src/matrix.c:
 * 555-577); the delta form gives the same determinant without the rounding
 * of 1 - M_kk to the nearest double near 1. */
double orc_logdet_1mM_delta(const double *M, int n)
{
    double *D = malloc((size_t)n*n*sizeof(double));
    double *logs = malloc((size_t)n*sizeof(double));
    for(size_t i = 0; i < (size_t)n*n; i++) D[i] = -M[i];
    for(int k = 0; k < n; k++) {
        const double d = D[(size_t)k*n + k];
        logs[k] = log_abs_1p(d);
        const double inv = 1/(1 + d);
        for(int i = k + 1; i < n; i++) {
            double *ri = D + (size_t)i*n;
            const double l = ri[k]*inv;
            ri[k] = l;
            if(l == 0) continue;
            const double *rk = D + (size_t)k*n;
            for(int j = k + 1; j < n; j++)
                ri[j] -= l*rk[j];
        }
    }
    const double r = orc_kahan_sum(logs, n);
    free(logs);
    free(D);
    return r;
}

/* log det(1-M) with the reference's shortcuts.  Reference: src/matrix.c:57-149
 * (|tr M| < 1e-8 -> -tr M; HODLR mode: -tr M if |tr M| > |logdet|) and
 * 489-544 (LU mode: Mercator -tr M - tr M²/2 if ||M||_F < 1 and its error bound
 * is below 1e-8 relative).  detalg 0 = HODLR mode (the reference default),
 * 1 = LU mode.  The determinant itself is always a dense LU here: the delta
 * form for column diagonally dominant 1 - M, else partial pivoting. */
double orc_logdet_1mM(const double *M, int n, int detalg)
{
    double *d = malloc((size_t)n*sizeof(double));
    for(int i = 0; i < n; i++) d[i] = M[(size_t)i*n + i];
    const double tr = orc_kahan_sum(d, n);
    free(d);
    if(fabs(tr) < 1e-8)
        return -tr;

    if(detalg == 1) {
        double norm = 0;
        for(size_t i = 0; i < (size_t)n*n; i++) norm += M[i]*M[i];
        norm = sqrt(norm);
        if(norm < 1) {
            double *row = malloc((size_t)n*sizeof(double));
            for(int i = 0; i < n; i++) {
                double s = 0;
                for(int j = 0; j < n; j++) s += M[(size_t)i*n + j]*M[(size_t)j*n + i];
                row[i] = s;
            }
            const double trA = -tr;
            const double trA2 = orc_kahan_sum(row, n);
            free(row);
            const double mercator = trA - trA2/2;
            const double err = fabs(norm*norm/2 + norm + log1p(-norm));
            if(fabs(err/mercator) < 1e-8)
                return mercator;
        }
    }

    double ld;
    if(orc_dominant_1mM(M, n))
        ld = orc_logdet_1mM_delta(M, n);
    else {
        double *A = malloc((size_t)n*n*sizeof(double));
        for(size_t i = 0; i < (size_t)n*n; i++) A[i] = -M[i];
        for(int i = 0; i < n; i++) A[(size_t)i*n + i] += 1;
        ld = orc_logdet_lu(A, n);
        free(A);
    }
    if(detalg == 0 && fabs(tr) > fabs(ld))
        return -tr;
    return ld;
}

/* caps_logdetD.  Reference: src/libcaps.c:950-962 (N = 2 ldim). */
double orc_logdetD(const orc_caps_t *c, double xi_, int m)
{
    orc_block_t *B = orc_block_new(c, xi_, m);
    if(!B) return NAN;
    const int N = 2*c->ldim;
    double *M = malloc((size_t)N*N*sizeof(double));
    orc_block_matrix(B, M);
    const double r = orc_logdet_1mM(M, N, c->detalg);
    free(M);
    orc_block_free(B);
    return r;
}

/* Many blocks, one per thread (mirrors the reference's one-block-per-process
 * MPI farm, src/caps.c:163-232,869-957). */
void orc_logdetD_many(const orc_caps_t *c, const double *xi_, const int *m, int n, double *out, int threads)
{
    (void)threads;
#ifdef _OPENMP
    #pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
#endif
    for(int i = 0; i < n; i++)
        out[i] = orc_logdetD(c, xi_[i], m[i]);
}

/* ξ = 0 Drude term (Bimonte-Emig, PRL 109, 160403 (2012), eq. (8)).
 * Reference: src/libcaps.c:1043-1065. */
double orc_ht_drude(double x)
{
    const double x2 = x*x;
    const double Z = (1 + x)*(1 - sqrt((x2 + 2*x)/(x2 + 2*x + 1)));
    double s1 = 0, s2 = 0;
    for(int l = 1; ; l++) {
        const double a = pow(Z, 2*l);
        const double b = a*Z;
        const double v1 = (2*l + 1)*log1p(-b);
        const double v2 = b*(1 - a)/(1 - b);
        s1 += v1;
        s2 += v2;
        if(fabs(v1/s1) < 1e-15 && fabs(v2/s2) < 1e-15)
            break;
    }
    return 0.5*(s1 + log1p(-(1 - Z*Z)*s2));
}
