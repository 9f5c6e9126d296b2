// log|det A| of a batch of FP64 matrices by blocked right-looking LU with
// partial pivoting (the dgetrf_ route of src/matrix.c:555-577 and the Kahan
// sum of log|u_ii| of src/matrix.c:462-471), on gfx950.
//
// Per panel of NB = 64 columns, three batched kernels; with two (three)
// levels, steps of 128 (256) columns whose panels are factored with updates
// restricted to the step's own columns, followed by one K = 128 (K = 256)
// update of the trailing matrix (schedules at capsd_launch_lu):
//   k_lu_panel      one workgroup per matrix factors the N-k x 64 panel:
//                   left-looking over leaves of 2..8 columns, each leaf held
//                   in registers (R rows per lane; 256, 512 or 1024 lanes by
//                   panel height), pivot search reduced over the wavefronts,
//                   rows swapped through a logical->physical row map in LDS
//                   (no row moves in HBM during the panel); writes the
//                   factored panel in pivoted order to a side buffer, the
//                   pivots (LAPACK ipiv convention) and log|u_ii|;
//   k_lu_swap_trsm  applies the row interchanges to the trailing columns
//                   and solves U12 = L11^{-1} A12 (one lane per column),
//                   writing U12 to the step's row-major U buffer;
//   k_lu_gemm_dma   A22 -= L21 U12 on FP64 MFMA (v_mfma_f64_16x16x4f64),
//                   128 x 128 tiles, operand stages moved straight from
//                   HBM into LDS (L from the panel buffers, U from the U
//                   buffer), double-buffered;
//   k_lu_gemm       the same product with register-staged operands, for
//                   128 x 64, 64 x 128, 64 x 64 and 256 x 128 tiles.
// Column-dominant blocks take the delta form instead (no pivot search, the
// matrix holding -M with u = 1 + d): k_lu_diag factors the 64 x 64 diagonal
// block and inverts its triangles, k_lu_l21 forms L21 = A21 U11^{-1} and
// k_lu_trsm_mfma U12 = L11^{-1} A12.  Round-trip blocks are symmetric: there
// the updates form U12 = D11 L21^T from the L21 panel buffers and the
// pivots (no TRSM, no U buffer), keep the lower triangle, and a square
// region launches only its lower triangle of tiles.
// k_lu_finalize sums log|u_ii| in double-double arithmetic.
// Matrices are column-major with leading dimension lda.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include <mutex>

#include "blocks.h"
#include "wave.h"

namespace capsd {

constexpr int NB = 64;       // panel width
#ifndef LU_KS
#define LU_KS 16
#endif

template <typename T>
__device__ __forceinline__ T* at(char* base, int64_t off) { return reinterpret_cast<T*>(base + off); }

// Side buffers of the LU of one block (inside the block's arena slice):
//   off_piv:              N ints (unused by the logdet; kept for the layout)
// and, appended by the host after the matrix:
struct LuSide {
    int64_t off_panel;   // NPANEL x lda x NB doubles: the factored panels of a
                         // step (panel q at column k + q NB), rows relative to
                         // the panel's first row, pivoted order
    int64_t off_logs;    // N doubles: log|u_ii|
    int64_t off_moved;   // NPANEL x NB ints: the row interchanges of each
                         // panel, LAPACK ipiv convention relative to the
                         // panel's first row (at column c rows c and ipiv[c]
                         // were exchanged)
    int64_t off_ut;      // NPANEL NB x lda doubles: the step's U rows, row
                         // k + i at i lda (columns at their matrix index);
                         // symmetric form: only panel q's pivots, at row q NB,
                         // columns 0 .. 63 (never U12 entries)
    int64_t off_inv;     // NPANEL x 2 NB x NB doubles (delta form): per panel
                         // L11^{-1} column-major, then U11^{-1} row-major
};
constexpr int NPANEL = 8;    // panels per step of the four-level schedule

__device__ __forceinline__ double* panel_buf(char* arena, const LuSide& sd, int64_t lda, int pb)
{
    return at<double>(arena, sd.off_panel) + (int64_t)pb * lda * NB;
}
__device__ __forceinline__ int* moved_buf(char* arena, const LuSide& sd, int pb)
{
    return at<int>(arena, sd.off_moved) + pb * NB;
}
__device__ __forceinline__ double* inv_buf(char* arena, const LuSide& sd, int pb)
{
    return at<double>(arena, sd.off_inv) + (int64_t)pb * 2 * NB * NB;
}

typedef double double4_t __attribute__((ext_vector_type(4)));

// log|1 + d| of a delta-form pivot u = 1 + d: log1p keeps the relative
// accuracy of d (a pivot near 1 stored as 1 + d loses d below one ulp of 1).
__device__ __forceinline__ double log_abs_1p(double d) { return d > -0.5 ? log1p(d) : log(fabs(1.0 + d)); }

// Largest v over the wavefront, lowest idx among equal values (a total
// order, so the result does not depend on the reduction tree), in every
// lane: within rows of 16 by DPP (quad exchanges 1 and 2, half-row and row
// mirrors), then across the four rows through scalar registers.
__device__ __forceinline__ void wave_argmax(double& v, int& idx)
{
    auto take = [&](double ov, int oi) {
        if (ov > v || (ov == v && oi < idx)) {
            v = ov;
            idx = oi;
        }
    };
    take(dpp_d<DPP_QUAD_X1>(v), dpp_i<DPP_QUAD_X1>(idx));
    take(dpp_d<DPP_QUAD_X2>(v), dpp_i<DPP_QUAD_X2>(idx));
    take(dpp_d<DPP_HALF_MIRROR>(v), dpp_i<DPP_HALF_MIRROR>(idx));
    take(dpp_d<DPP_MIRROR>(v), dpp_i<DPP_MIRROR>(idx));
    double bv = lane_d(v, 0);
    int bi = __builtin_amdgcn_readlane(idx, 0);
#pragma unroll
    for (int r = 16; r < 64; r += 16) {
        const double ov = lane_d(v, r);
        const int oi = __builtin_amdgcn_readlane(idx, r);
        if (ov > bv || (ov == bv && oi < bi)) {
            bv = ov;
            bi = oi;
        }
    }
    v = bv;
    idx = bi;
}

// Panel body: R rows per lane, leaves of LF columns, NT threads, the
// left-looking update in chunks of CW columns of L (CW divides LF); the
// logical->physical row map lives in dynamic LDS (R*NT ints).  Every leaf
// re-reads the L columns left of it (64 columns, leaves of 4: 480 L columns
// per panel; of 8: 224), while each column's pivot exchange and elimination
// touch all R x LF leaf registers: leaves of 8 are the fastest measured.
template <int R, int LF, int NT, int CW>
__device__ __forceinline__ void lu_panel_body(const BlockDesc* __restrict__ desc, char* __restrict__ arena,
                                              const LuSide* __restrict__ side, const BlockOut* __restrict__ out, int k,
                                              int pb)
{
    const int b = blockIdx.x;
    const BlockDesc& d = desc[b];
    if (out[b].skip || out[b].delta || k >= d.N) return;   // delta form: k_lu_diag / k_lu_l21
    const int N = d.N;
    const int64_t lda = d.lda;
    const int np = N - k;                 // panel rows
    const int nbk = min(NB, np);          // panel columns
    double* __restrict__ A = at<double>(arena, d.off_mat);
    double* __restrict__ P = panel_buf(arena, side[b], lda, pb);
    double* __restrict__ logs = at<double>(arena, side[b].off_logs);
    int* __restrict__ ipiv = moved_buf(arena, side[b], pb);
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;

    extern __shared__ int perm[];   // R * NT
    __shared__ double red_v2[2][NT / 64];
    __shared__ int red_i2[2][NT / 64];
    __shared__ double prow2[2][LF], xrow2[2][LF];
    __shared__ double Uc[2][CW * LF];   // solved U rows of one chunk, [row][col]
    __shared__ double pivv[NB];         // |pivot| per panel column

    for (int r = tid; r < np; r += NT) perm[r] = r;
    __syncthreads();

    double a[R][LF];
    for (int c0 = 0; c0 < nbk; c0 += LF) {
        const int w = min(LF, nbk - c0);
        // load the leaf (logical rows, through the row map)
#pragma unroll
        for (int i = 0; i < R; i++) {
            const int r = tid + i * NT;
#pragma unroll
            for (int j = 0; j < LF; j++)
                a[i][j] = (r < np && j < w) ? A[(k + perm[r]) + (k + c0 + j) * lda] : 0.0;
        }
        // left-looking update with the panel's earlier columns, CW columns
        // of L per chunk.  Rows [cc0, cc0+CW) of the leaf (lanes cc0.. of
        // wave 0, register row 0) become final U rows by a forward
        // substitution inside the wavefront; every row below then subtracts
        // the chunk's rank-CW product.
        for (int cc0 = 0; cc0 < c0; cc0 += CW) {
            const int buf = (cc0 / CW) & 1;
            double Lr[R][CW];
#pragma unroll
            for (int i = 0; i < R; i++) {
                const int r = tid + i * NT;
                const bool row_ok = r > cc0 && r < np;
                const int64_t pr = k + (row_ok ? perm[r] : 0);
#pragma unroll
                for (int jj = 0; jj < CW; jj++)
                    Lr[i][jj] = (row_ok && cc0 + jj < r) ? A[pr + (k + cc0 + jj) * lda] : 0.0;
            }
            if (wid == 0) {
#pragma unroll
                for (int t = 0; t < CW; t++) {
                    const int q = cc0 + t;
                    double uq[LF];
#pragma unroll
                    for (int j = 0; j < LF; j++) uq[j] = lane_d(a[0][j], q);
                    if (lane > q && lane < cc0 + CW) {
#pragma unroll
                        for (int j = 0; j < LF; j++) a[0][j] -= Lr[0][t] * uq[j];
                    }
                }
                if (lane >= cc0 && lane < cc0 + CW) {
#pragma unroll
                    for (int j = 0; j < LF; j++) Uc[buf][(lane - cc0) * LF + j] = a[0][j];
                }
            }
            __syncthreads();
#pragma unroll
            for (int t = 0; t < CW; t++) {
#pragma unroll
                for (int j = 0; j < LF; j++) {
                    const double u = Uc[buf][t * LF + j];
#pragma unroll
                    for (int i = 0; i < R; i++)
                        if (i > 0 || tid >= cc0 + CW) a[i][j] -= Lr[i][t] * u;
                }
            }
        }
        // factor the leaf with partial pivoting (unrolled: j indexes registers)
#pragma unroll
        for (int j = 0; j < LF; j++) {
            if (j >= w) break;
            const int c = c0 + j;
            // per-column scratch alternates between two buffers, so one
            // barrier after the reduction and one after the pivot row is
            // published suffice
            double* red_v = red_v2[c & 1];
            int* red_i = red_i2[c & 1];
            double* prow = prow2[c & 1];
            double* xrow = xrow2[c & 1];
            double bv = -1.0;
            int bi = 0x7fffffff;
#pragma unroll
            for (int i = 0; i < R; i++) {
                const int r = tid + i * NT;
                const double v = fabs(a[i][j]);
                if (r >= c && r < np && (v > bv || (v == bv && r < bi))) {
                    bv = v;
                    bi = r;
                }
            }
            wave_argmax(bv, bi);
            if (lane == 0) {
                red_v[wid] = bv;
                red_i[wid] = bi;
            }
            __syncthreads();
            double gv = red_v[0];
            int p = red_i[0];
            for (int q = 1; q < NT / 64; q++)
                if (red_v[q] > gv || (red_v[q] == gv && red_i[q] < p)) {
                    gv = red_v[q];
                    p = red_i[q];
                }
            // no finite candidate (a column of NaN): keep row c, so the row
            // map and ipiv stay in range; the NaN reaches log|u| and the
            // determinant
            if (p >= np) {
                p = c;
                gv = NAN;
            }
            // owners of rows p and c publish them
            const int own_p = p % NT, own_c = c % NT;
            if (tid == own_p) {
#pragma unroll
                for (int i = 0; i < R; i++)
                    if (tid + i * NT == p) {
#pragma unroll
                        for (int jj = 0; jj < LF; jj++) prow[jj] = a[i][jj];
                    }
            }
            if (tid == own_c && p != c) {
#pragma unroll
                for (int i = 0; i < R; i++)
                    if (tid + i * NT == c) {
#pragma unroll
                        for (int jj = 0; jj < LF; jj++) xrow[jj] = a[i][jj];
                    }
            }
            if (tid == 0) {
                if (p != c) {
                    const int t = perm[c];
                    perm[c] = perm[p];
                    perm[p] = t;
                }
                pivv[c] = gv;
                ipiv[c] = p;
            }
            __syncthreads();
            const double u = prow[j];
#pragma unroll
            for (int i = 0; i < R; i++) {
                const int r = tid + i * NT;
                if (r == c) {
#pragma unroll
                    for (int jj = 0; jj < LF; jj++) a[i][jj] = prow[jj];
                } else if (r == p) {
#pragma unroll
                    for (int jj = 0; jj < LF; jj++) a[i][jj] = xrow[jj];
                }
            }
            // l = a / u as a product with 1/u, as LAPACK's dgetf2 scales the
            // column (a division only below the smallest normal number); u is
            // the same in every lane, so the choice is uniform
            const int by_rcp = __builtin_amdgcn_readfirstlane((int)(fabs(u) >= 2.2250738585072014e-308));
            const double ru = 1.0 / u;
#pragma unroll
            for (int i = 0; i < R; i++) {
                const int r = tid + i * NT;
                if (r > c && r < np) {
                    const double l = by_rcp ? a[i][j] * ru : a[i][j] / u;
                    a[i][j] = l;
#pragma unroll
                    for (int jj = 0; jj < LF; jj++)
                        if (jj > j) a[i][jj] -= l * prow[jj];
                }
            }
        }
        // write the leaf back through the row map
#pragma unroll
        for (int i = 0; i < R; i++) {
            const int r = tid + i * NT;
            if (r < np) {
                const int64_t pr = k + perm[r];
#pragma unroll
                for (int j = 0; j < LF; j++)
                    if (j < w) A[pr + (k + c0 + j) * lda] = a[i][j];
            }
        }
        __syncthreads();
    }
    // log|u_ii| and the panel in pivoted (logical) order
    if (tid < nbk) logs[k + tid] = log(pivv[tid]);
    for (int r = tid; r < np; r += NT) {
        const int64_t pr = k + perm[r];
#pragma unroll
        for (int j0 = 0; j0 < NB; j0 += 16) {
            double v[16];
#pragma unroll
            for (int j = 0; j < 16; j++) v[j] = j0 + j < nbk ? A[pr + (k + j0 + j) * lda] : 0.0;
#pragma unroll
            for (int j = 0; j < 16; j++)
                if (j0 + j < nbk) P[r + (j0 + j) * lda] = v[j];
        }
    }
}


// ---- delta form: panels without pivot search --------------------------------
// For a column diagonally dominant 1 - M (the matrix holds -M, out[b].delta)
// partial pivoting exchanges no rows, so a panel is the unpivoted block
// factorization [A11; A21] = [L11; L21] U11: the 64 x 64 diagonal block is
// factored by one workgroup (k_lu_diag, which also forms L11^{-1} and
// U11^{-1}), then L21 = A21 U11^{-1} (k_lu_l21) and, for the step's U rows,
// U12 = L11^{-1} A12 (k_lu_trsm_mfma) are products on the FP64 MFMA units,
// spread over as many workgroups as the panel has row (column) tiles.  The
// diagonal blocks of a dominant matrix are well conditioned (|u_ii| close to
// 1), so products with the triangular inverses are as accurate as the
// substitutions they replace.

// The 64 x 64 diagonal block of panel pb at step k, by one workgroup of four
// wavefronts with the block in LDS, in compact loops (a fully unrolled
// 64-step form is larger than the instruction cache, and a lone block's LU
// then waits on instruction fetch).  Unpivoted blocked LU over four 16-column
// panels, pivots u = 1 + d (d the stored diagonal), log|u_ii| = log1p(d):
// (a) wavefront 0 factors the panel's rows c0..63 (one lane per row, the 16
//     columns in registers, pivot rows broadcast by readlane);
// (b) U12 of the panel's 16 rows: forward substitution with the panel's unit
//     lower block, one lane per column;
// (c) the trailing block minus L21 U12 on the FP64 MFMA units (16 x 16
//     tiles, K = 16, spread over the wavefronts).
// The factored block goes to the panel buffer's top rows.  Then the
// triangular inverses, blocked by 16: the diagonal blocks of L11^{-1} and
// U11^{-1} by substitution (lane = column of one of the four blocks), the
// blocks at distance 1, 2, 3 from the diagonal as MFMA products,
// X_ij = -X_ii sum_{k=j}^{i-1} L_ik X_kj and Y_ij = -Y_ii sum_{k=i+1}^{j}
// U_ik Y_kj; L11^{-1} is stored column-major, U11^{-1} row-major.  Rows and
// columns past the matrix (a last panel narrower than 64) act as identity.
__global__ void __launch_bounds__(256, 2) k_lu_diag(const BlockDesc* __restrict__ desc, char* __restrict__ arena,
                                                 const LuSide* __restrict__ side, const BlockOut* __restrict__ out,
                                                 int k, int pb)
{
    constexpr int WP = NB + 1;   // LDS pitch (doubles)
    // W: the block, after the LU L (strict lower) and U - I (upper);
    // Z: L11^{-1} below the diagonal, U11^{-1} on and above it
    __shared__ double W[NB * WP];
    __shared__ double Z[NB * WP];
    const int b = blockIdx.x;
    const BlockDesc& d = desc[b];
    if (out[b].skip || !out[b].delta) return;
    const int N = d.N, k0 = k + pb * NB;
    if (k0 >= N) return;
    const int nbk = min(NB, N - k0);
    const int64_t lda = d.lda;
    const double* __restrict__ A = at<double>(arena, d.off_mat);
    double* __restrict__ P = panel_buf(arena, side[b], lda, pb);
    double* __restrict__ inv = inv_buf(arena, side[b], pb);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;   // MFMA operand row / k index
#pragma unroll 4
    for (int s = 0; s < 16; s++) {
        const int j = wv + 4 * s;
        W[lane * WP + j] = (lane < nbk && j < nbk) ? A[(k0 + lane) + (int64_t)(k0 + j) * lda] : 0.0;
    }
    __syncthreads();
#pragma unroll 1
    for (int p = 0; p < 4; p++) {
        const int c0 = 16 * p;
        // (a) panel columns c0 .. c0+15, rows c0 .. 63
        if (wv == 0) {
            double a[16];
#pragma unroll
            for (int t = 0; t < 16; t++) a[t] = W[lane * WP + c0 + t];
#pragma unroll
            for (int t = 0; t < 16; t++) {
                const int pr = c0 + t;
                double u[16];
#pragma unroll
                for (int j = t; j < 16; j++) u[j] = lane_d(a[j], pr);
                const double ru = 1.0 / (1.0 + u[t]);
                if (lane > pr) {
                    const double l = a[t] * ru;
                    a[t] = l;
#pragma unroll
                    for (int j = t + 1; j < 16; j++) a[j] -= l * u[j];
                }
            }
            if (lane >= c0) {
#pragma unroll
                for (int t = 0; t < 16; t++) W[lane * WP + c0 + t] = a[t];
            }
        }
        __syncthreads();
        if (p == 3) break;
        // (b) rows c0 .. c0+15, columns c0+16 .. 63: y_t -= L_ts y_s, s < t
        if (wv == 0 && lane < 48 - c0) {
            const int j = c0 + 16 + lane;
            double y[16];
#pragma unroll
            for (int t = 0; t < 16; t++) y[t] = W[(c0 + t) * WP + j];
#pragma unroll
            for (int t = 1; t < 16; t++)
#pragma unroll
                for (int s2 = 0; s2 < t; s2++) y[t] -= W[(c0 + t) * WP + c0 + s2] * y[s2];
#pragma unroll
            for (int t = 1; t < 16; t++) W[(c0 + t) * WP + j] = y[t];
        }
        __syncthreads();
        // (c) trailing tiles: W -= L21 U12 (K = 16)
        const int nt = 3 - p;
#pragma unroll 1
        for (int tile = wv; tile < nt * nt; tile += 4) {
            const int i0 = c0 + 16 + 16 * (tile / nt), j0 = c0 + 16 + 16 * (tile % nt);
            double4_t acc;
#pragma unroll
            for (int q = 0; q < 4; q++) acc[q] = W[(i0 + lk + 4 * q) * WP + j0 + li];
#pragma unroll
            for (int kk = 0; kk < 16; kk += 4) {
                const double av = -W[(i0 + li) * WP + c0 + kk + lk];
                const double bv = W[(c0 + kk + lk) * WP + j0 + li];
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
            }
#pragma unroll
            for (int q = 0; q < 4; q++) W[(i0 + lk + 4 * q) * WP + j0 + li] = acc[q];
        }
        __syncthreads();
    }
    // the factored block, log|u_ii|, the (identity) row order
#pragma unroll 4
    for (int s = 0; s < 16; s++) {
        const int j = wv + 4 * s;
        if (lane < nbk && j < nbk) P[lane + (int64_t)j * lda] = W[lane * WP + j];
    }
    if (wv == 0 && lane < nbk) {
        at<double>(arena, side[b].off_logs)[k0 + lane] = log_abs_1p(W[lane * WP + lane]);
        moved_buf(arena, side[b], pb)[lane] = lane;
        // the pivots u = 1 + d for the symmetric form's updates, contiguous at
        // columns 0 .. 63 of the panel's first U-buffer row (never a U12
        // entry: those lie at columns k0 + NB and beyond)
        at<double>(arena, side[b].off_ut)[(int64_t)pb * NB * lda + lane] = 1.0 + W[lane * WP + lane];
    }
    // diagonal blocks of the inverses: wavefront 0 L^{-1}, wavefront 1 U^{-1};
    // lane = 16 * block + column
    {
        const int bb = 16 * lk, c = li;
        double x[16];
#pragma unroll
        for (int i = 0; i < 16; i++) x[i] = (i == c) ? 1.0 : 0.0;
        if (wv == 0) {
            // x = e_c; for q: x_i -= L_iq x_q, i > q
#pragma unroll
            for (int q = 0; q < 15; q++)
#pragma unroll
                for (int i = q + 1; i < 16; i++) x[i] -= W[(bb + i) * WP + bb + q] * x[q];
#pragma unroll
            for (int i = 0; i < 16; i++)
                if (i > c) Z[(bb + i) * WP + bb + c] = x[i];
        } else if (wv == 1) {
            // x = e_c; for q from the last: x_q /= u_qq, x_i -= U_iq x_q, i < q
            double ru[16];
#pragma unroll
            for (int q = 0; q < 16; q++) ru[q] = 1.0 / (1.0 + W[(bb + q) * WP + bb + q]);
#pragma unroll
            for (int q = 15; q >= 0; q--) {
                const double xq = x[q] * ru[q];
                x[q] = xq;
#pragma unroll
                for (int i = 0; i < q; i++) x[i] -= W[(bb + i) * WP + bb + q] * xq;
            }
#pragma unroll
            for (int i = 0; i < 16; i++)
                if (i <= c) Z[(bb + i) * WP + bb + c] = x[i];
        }
    }
    __syncthreads();
    // blocks at distance dd from the diagonal: tasks 0 .. 3-dd of L^{-1}
    // (block (dd + t, t)), then those of U^{-1} (block (t, dd + t))
#pragma unroll 1
    for (int dd = 1; dd < 4; dd++) {
        const int ntask = 4 - dd;
#pragma unroll 1
        for (int task = wv; task < 2 * ntask; task += 4) {
            const bool lower = task < ntask;
            const int t = lower ? task : task - ntask;
            const int bi = lower ? dd + t : t, bj = lower ? t : dd + t;   // block row, column
            const int kb0 = lower ? bj : bi + 1, kb1 = lower ? bi - 1 : bj;
            double4_t acc = (double4_t){0, 0, 0, 0};
#pragma unroll 1
            for (int kb = kb0; kb <= kb1; kb++) {
#pragma unroll
                for (int kk = 0; kk < 16; kk += 4) {
                    // W block (bi, kb) times Z block (kb, bj); a diagonal
                    // Z block is unit lower (L^{-1}) or upper (U^{-1})
                    const double av = W[(16 * bi + li) * WP + 16 * kb + kk + lk];
                    const int zr = kk + lk;
                    double bv = Z[(16 * kb + zr) * WP + 16 * bj + li];
                    if (kb == bj) bv = lower ? (zr > li ? bv : (zr == li ? 1.0 : 0.0)) : (zr <= li ? bv : 0.0);
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
                }
            }
            // times -(diagonal block bi of the inverse); the sum's register q
            // holds rows 4q .. 4q+3 of K: the B operand of step q
            double4_t res = (double4_t){0, 0, 0, 0};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int zc = 4 * q + lk;
                const double z = Z[(16 * bi + li) * WP + 16 * bi + zc];
                const double av = lower ? (li > zc ? -z : (li == zc ? -1.0 : 0.0)) : (li <= zc ? -z : 0.0);
                res = __builtin_amdgcn_mfma_f64_16x16x4f64(av, acc[q], res, 0, 0, 0);
            }
#pragma unroll
            for (int q = 0; q < 4; q++) Z[(16 * bi + lk + 4 * q) * WP + 16 * bj + li] = res[q];
        }
        __syncthreads();
    }
    // L11^{-1} column-major (unit diagonal, zero above), U11^{-1} row-major
    // (zero below)
#pragma unroll 4
    for (int s = 0; s < 16; s++) {
        const int c = wv + 4 * s;
        const int r = lane;
        inv[(int64_t)c * NB + r] = r > c ? Z[r * WP + c] : (r == c ? 1.0 : 0.0);
        inv[NB * NB + (int64_t)c * NB + r] = c <= r ? Z[c * WP + r] : 0.0;
    }
}

// L21 = A21 U11^{-1} for the rows below the diagonal block of panel pb, in
// MFMA tiles of 16 x 16 over K = 64.  A wavefront takes 64 rows and CT of
// the four 16-column tiles of the result: CT = 4 (workgroups of 256 rows,
// 256 MFMAs per wavefront) when the batch has rows enough to fill the GPU,
// CT = 1 (workgroups of 64 rows, the four wavefronts one column tile each,
// 64 MFMAs per wavefront, all operands loaded before the first product) for
// a short grid, whose time is the latency of one workgroup.  A21 read from
// the matrix (16-row column segments), U11^{-1} (row-major, L2-resident) as
// the B operand; the product, with swapped operands so each register holds
// whole 16-row column segments, goes to the panel buffer (rows relative to
// the panel's first row), the trailing updates' L operand (and, in the
// symmetric form, their U operand: U12 = D11 L21^T).
template <int CT>
__global__ void __launch_bounds__(256, 2) k_lu_l21(const BlockDesc* __restrict__ desc, char* __restrict__ arena,
                                               const LuSide* __restrict__ side, const BlockOut* __restrict__ out,
                                               int k, int pb)
{
    static_assert(CT == 1 || CT == 4, "column tiles per wavefront");
    constexpr int WPR = 4 / CT;   // wavefronts per 64-row group
    const int b = blockIdx.y;
    const BlockDesc& d = desc[b];
    if (out[b].skip || !out[b].delta) return;
    const int N = d.N, k0 = k + pb * NB;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int m0 = k0 + NB + blockIdx.x * (64 * CT) + (wv / WPR) * 64;   // this wave's first row
    if (m0 >= N) return;
    const int j0 = (wv % WPR) * CT;   // first column tile
    const int64_t lda = d.lda;
    const double* __restrict__ A = at<double>(arena, d.off_mat);
    double* __restrict__ P = panel_buf(arena, side[b], lda, pb);
    const double* __restrict__ Ui = inv_buf(arena, side[b], pb) + NB * NB;
    const int li = lane & 15, lk = lane >> 4;
    double4_t acc[4][CT];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < CT; j++) acc[i][j] = (double4_t){0, 0, 0, 0};
    if constexpr (CT == 4) {
#pragma unroll
        for (int kk = 0; kk < NB; kk += 4) {
            double av[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int row = m0 + 16 * i + li;
                av[i] = row < N ? A[row + (int64_t)(k0 + kk + lk) * lda] : 0.0;
            }
#pragma unroll
            for (int j = 0; j < 4; j++) bv[j] = Ui[(kk + lk) * NB + 16 * j + li];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j], av[i], acc[i][j], 0, 0, 0);
        }
    } else {
        double av[NB / 4][4], bv[NB / 4];
#pragma unroll
        for (int s = 0; s < NB / 4; s++) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int row = m0 + 16 * i + li;
                av[s][i] = row < N ? A[row + (int64_t)(k0 + 4 * s + lk) * lda] : 0.0;
            }
            bv[s] = Ui[(4 * s + lk) * NB + 16 * j0 + li];
        }
#pragma unroll
        for (int s = 0; s < NB / 4; s++)
#pragma unroll
            for (int i = 0; i < 4; i++)
                acc[i][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[s], av[s][i], acc[i][0], 0, 0, 0);
    }
    // lane l, register q of acc[i][j]: row m0 + 16i + l%16, column
    // 16 (j0 + j) + l/16 + 4q
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int row = m0 + 16 * i + li;
        if (row < N) {
#pragma unroll
            for (int j = 0; j < CT; j++)
#pragma unroll
                for (int q = 0; q < 4; q++)
                    P[(row - k0) + (int64_t)(16 * (j0 + j) + lk + 4 * q) * lda] = acc[i][j][q];
        }
    }
}

// U12 = L11^{-1} A12 for the 64 rows of panel pb over the columns
// [r0 + cs_off, min(N, r0 + cs_off + climit)), r0 = k + pb NB, written to the
// step's row-major U buffer (the trailing updates' B operand).  A workgroup
// takes TC columns: the 64 x TC tile of A12 is staged in LDS (one column's
// 64 rows per wavefront load, 512 contiguous bytes, all of a wavefront's
// loads issued before its LDS stores), each wavefront computes 16 rows x TC
// columns (TC/16 MFMA tiles) over K = 64 with L11^{-1} (column-major,
// L2-resident) as the A operand; the result's 16-column row segments go
// straight to the U buffer.  TC = 64 keeps four workgroups per CU (33 KB of
// LDS each), so one workgroup's loads run under another's MFMAs.
template <int TC>
__global__ void __launch_bounds__(256) k_lu_trsm_mfma(const BlockDesc* __restrict__ desc, char* __restrict__ arena,
                                                     const LuSide* __restrict__ side,
                                                     const BlockOut* __restrict__ out, int k, int pb, int cs_off,
                                                     int climit)
{
    constexpr int PITCH = TC + 1, CPW = TC / 4, NT = TC / 16;
    __shared__ double Bs[NB * PITCH];   // [k][column]
    const int b = blockIdx.y;
    const BlockDesc& d = desc[b];
    if (out[b].skip || !out[b].delta) return;
    const int N = d.N;
    const int r0 = k + pb * NB;
    if (r0 + cs_off >= N) return;
    const int cend = min(N, r0 + cs_off + climit);
    const int c0 = r0 + cs_off + blockIdx.x * TC;
    if (c0 >= cend) return;
    const int64_t lda = d.lda;
    const double* __restrict__ A = at<double>(arena, d.off_mat);
    const double* __restrict__ Li = inv_buf(arena, side[b], pb);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // stage A12[0..64) x [c0, c0+TC): wavefront wv loads columns wv, wv+4, ...
    {
        double v[CPW];
#pragma unroll
        for (int t = 0; t < CPW; t++) {
            const int col = c0 + wv + 4 * t;
            v[t] = col < cend ? A[(r0 + lane) + (int64_t)col * lda] : 0.0;
        }
#pragma unroll
        for (int t = 0; t < CPW; t++) Bs[lane * PITCH + wv + 4 * t] = v[t];
    }
    __syncthreads();
    const int li = lane & 15, lk = lane >> 4;
    double4_t acc[NT];
#pragma unroll
    for (int j = 0; j < NT; j++) acc[j] = (double4_t){0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < NB; kk += 4) {
        const double av = Li[(int64_t)(kk + lk) * NB + 16 * wv + li];   // L11^{-1}[16 wv + li][kk + lk]
#pragma unroll
        for (int j = 0; j < NT; j++)
            acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, Bs[(kk + lk) * PITCH + 16 * j + li], acc[j], 0, 0, 0);
    }
    // lane l, register q of acc[j]: row 16 wv + l/16 + 4q, column c0 + 16j + l%16
    double* Ut = at<double>(arena, side[b].off_ut) + (int64_t)pb * NB * lda;
#pragma unroll
    for (int j = 0; j < NT; j++) {
        const int col = c0 + 16 * j + li;
        if (col < cend) {
#pragma unroll
            for (int q = 0; q < 4; q++) Ut[(int64_t)(16 * wv + lk + 4 * q) * lda + col] = acc[j][q];
        }
    }
}

// Tall panels (more rows than one workgroup holds in registers): the panel
// is factored one column at a time by grid-wide kernels, 256 rows per
// workgroup; rows are exchanged in place within the panel's columns.  The
// panel buffer serves as scratch for the per-workgroup pivot candidates
// (value, row) until the panel is copied into it at the end.
constexpr int GP_ROWS = 256;

__device__ __forceinline__ void block_argmax_store(double v, int r, double* scratch, int slot)
{
    __shared__ double sv[GP_ROWS / 64];
    __shared__ int si[GP_ROWS / 64];
    wave_argmax(v, r);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        sv[wid] = v;
        si[wid] = r;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double bv = sv[0];
        int bi = si[0];
        for (int q = 1; q < GP_ROWS / 64; q++)
            if (sv[q] > bv || (sv[q] == bv && si[q] < bi)) {
                bv = sv[q];
                bi = si[q];
            }
        scratch[2 * slot] = bv;
        scratch[2 * slot + 1] = (double)bi;
    }
}

// pivot candidates of column 0 of the panel at k
__global__ void __launch_bounds__(GP_ROWS) k_gp_max(const BlockDesc* __restrict__ desc, char* __restrict__ arena,
                                                    const LuSide* __restrict__ side, const BlockOut* __restrict__ out,
                                                    int k, int pb)
{
    const int b = blockIdx.y;
    const BlockDesc& d = desc[b];
    const int N = d.N;
    if (out[b].skip || out[b].delta || k >= N || k + (int)blockIdx.x * GP_ROWS >= N) return;
    const int64_t lda = d.lda;
    const double* A = at<double>(arena, d.off_mat);
    const int r = k + blockIdx.x * GP_ROWS + threadIdx.x;
    const double v = r < N ? fabs(A[r + (int64_t)k * lda]) : -1.0;
    block_argmax_store(v, r < N ? r : 0x7fffffff, panel_buf(arena, side[b], lda, pb), blockIdx.x);
}

// pivot of column c: reduce the candidates, exchange rows k+c and p within
// the panel, record ipiv and log|u_cc|
__global__ void __launch_bounds__(64) k_gp_pivot(const BlockDesc* __restrict__ desc, char* __restrict__ arena,
                                                 const LuSide* __restrict__ side, const BlockOut* __restrict__ out,
                                                 int k, int pb, int c)
{
    const int b = blockIdx.x;
    const BlockDesc& d = desc[b];
    const int N = d.N;
    if (out[b].skip || out[b].delta || k >= N) return;
    const int nbk = min(NB, N - k);
    if (c >= nbk) return;
    const int64_t lda = d.lda;
    double* A = at<double>(arena, d.off_mat);
    const double* sc = panel_buf(arena, side[b], lda, pb);
    const int nch = (N - k + GP_ROWS - 1) / GP_ROWS;
    double v = -1.0;
    int p = 0x7fffffff;
    for (int q = threadIdx.x; q < nch; q += 64) {
        const double qv = sc[2 * q];
        const int qi = (int)sc[2 * q + 1];
        if (qv > v || (qv == v && qi < p)) {
            v = qv;
            p = qi;
        }
    }
    wave_argmax(v, p);
    // no finite candidate (a NaN column): keep row k+c; the NaN reaches
    // log|u| and the determinant
    if (p >= N) {
        p = k + c;
        v = NAN;
    }
    const int j = threadIdx.x;
    if (j < nbk && p != k + c) {
        double* col = A + (int64_t)(k + j) * lda;
        const double t = col[k + c];
        col[k + c] = col[p];
        col[p] = t;
    }
    if (j == 0) {
        moved_buf(arena, side[b], pb)[c] = p - k;
        at<double>(arena, side[b].off_logs)[k + c] = log(v);
    }
}

// column c: l = a/u below the pivot, rank-1 update of the panel columns
// c+1.., and the pivot candidates of column c+1
__global__ void __launch_bounds__(GP_ROWS) k_gp_update(const BlockDesc* __restrict__ desc, char* __restrict__ arena,
                                                       const LuSide* __restrict__ side,
                                                       const BlockOut* __restrict__ out, int k, int pb, int c)
{
    const int b = blockIdx.y;
    const BlockDesc& d = desc[b];
    const int N = d.N;
    if (out[b].skip || out[b].delta || k >= N || k + (int)blockIdx.x * GP_ROWS >= N) return;
    const int nbk = min(NB, N - k);
    if (c >= nbk) return;
    const int64_t lda = d.lda;
    double* A = at<double>(arena, d.off_mat);
    const int r = k + blockIdx.x * GP_ROWS + threadIdx.x;
    const int pr = k + c;
    double vnext = -1.0;
    if (r > pr && r < N) {
        const double u = A[pr + (int64_t)pr * lda];
        const double l = A[r + (int64_t)pr * lda] / u;
        A[r + (int64_t)pr * lda] = l;
        for (int j = c + 1; j < nbk; j++) {
            const int64_t cj = (int64_t)(k + j) * lda;
            const double x = A[r + cj] - l * A[pr + cj];
            A[r + cj] = x;
            if (j == c + 1) vnext = fabs(x);
        }
    }
    if (c + 1 < nbk)
        block_argmax_store(vnext, (r > pr && r < N) ? r : 0x7fffffff, panel_buf(arena, side[b], lda, pb),
                           blockIdx.x);
}

// the factored panel in (now physical = logical) row order into its buffer
__global__ void __launch_bounds__(GP_ROWS) k_gp_copy(const BlockDesc* __restrict__ desc, char* __restrict__ arena,
                                                     const LuSide* __restrict__ side, const BlockOut* __restrict__ out,
                                                     int k, int pb)
{
    const int b = blockIdx.y;
    const BlockDesc& d = desc[b];
    const int N = d.N;
    if (out[b].skip || out[b].delta || k >= N) return;
    const int nbk = min(NB, N - k);
    const int64_t lda = d.lda;
    const double* A = at<double>(arena, d.off_mat);
    double* P = panel_buf(arena, side[b], lda, pb);
    const int r = k + blockIdx.x * GP_ROWS + threadIdx.x;
    if (r >= N) return;
    for (int j = 0; j < nbk; j++) P[(r - k) + (int64_t)j * lda] = A[r + (int64_t)(k + j) * lda];
}

// R rows per lane, leaves of LF columns, NT lanes, chunks of CW L columns;
// OCC waves per SIMD (OCC NT/256 workgroups per CU share the registers that
// hold the leaf and one chunk of L).
template <int R, int LF, int NT, int OCC, int CW = LF>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(OCC, OCC)))
k_lu_panel(const BlockDesc* __restrict__ desc, char* __restrict__ arena, const LuSide* __restrict__ side,
           const BlockOut* __restrict__ out, int k, int pb)
{
    lu_panel_body<R, LF, NT, CW>(desc, arena, side, out, k, pb);
}

// Row interchanges of one column (lane), in place: the panel's pivots as
// nbk sequential exchanges of rows r0+c and r0+ipiv[c] (uniform over the
// wavefront).
__device__ __forceinline__ void swap_rows_ipiv(double* __restrict__ Ac, const int* __restrict__ ipiv, int nbk, int r0)
{
    for (int c = 0; c < nbk; c++) {
        const int p = ipiv[c];
        if (p != c) {
            const double t = Ac[r0 + c];
            Ac[r0 + c] = Ac[r0 + p];
            Ac[r0 + p] = t;
        }
    }
}

// Row interchanges and U12 = L11^{-1} A12 for the 64-row tile of rows
// [r0, r0+64), r0 = k + pb NB, columns [r0+cs_off, min(N, r0+cs_off+climit)),
// one lane per column, NTH columns per workgroup.  For each set bit q of
// swaps (in increasing q) the rows are exchanged by the pivots of panel q
// (relative to row k + q NB); then the column's 64 entries are solved in
// registers with the unit lower L11 of panel buffer pb, staged once per
// workgroup in LDS (every lane reads the same L11 entry at a time: one
// broadcast LDS read instead of a vector load per entry).
template <int NTH>
__global__ void __launch_bounds__(NTH) k_lu_swap_trsm(const BlockDesc* __restrict__ desc, char* __restrict__ arena,
                                                      const LuSide* __restrict__ side,
                                                      const BlockOut* __restrict__ out, int k, int pb, int cs_off,
                                                      int climit, int swaps)
{
    __shared__ double sL[NB * NB];   // L11, [column][row]
    const int b = blockIdx.y;
    const BlockDesc& d = desc[b];
    const int N = d.N;
    const int r0 = k + pb * NB;
    if (out[b].skip || out[b].delta || r0 + cs_off >= N) return;   // delta form: k_lu_trsm_mfma
    const int cend = min(N, r0 + cs_off + climit);
    const int64_t lda = d.lda;
    const double* P = panel_buf(arena, side[b], lda, pb);
    for (int idx = threadIdx.x; idx < NB * NB; idx += NTH) sL[idx] = P[(idx % NB) + (int64_t)(idx / NB) * lda];
    __syncthreads();
    const int col = r0 + cs_off + blockIdx.x * NTH + threadIdx.x;
    if (col >= cend) return;
    double* Ac = at<double>(arena, d.off_mat) + (int64_t)col * lda;
    // the pivots of the panel at row k + q NB cover min(NB, N - k - q NB)
    // columns
    for (int q = 0; q < NPANEL; q++)
        if ((swaps >> q) & 1) swap_rows_ipiv(Ac, moved_buf(arena, side[b], q), min(NB, N - k - q * NB), k + q * NB);
    double x[NB];
#pragma unroll
    for (int i = 0; i < NB; i++) x[i] = Ac[r0 + i];
#pragma unroll
    for (int i = 0; i < NB; i++) {
#pragma unroll
        for (int j = i + 1; j < NB; j++) x[j] -= sL[i * NB + j] * x[i];
    }
    // the solved U rows go to the step's row-major U buffer (coalesced:
    // lanes are columns), the B operand of every trailing update; nothing
    // reads them from the matrix again (the determinant needs only the
    // diagonal), so they are not written back there, which saved 64 stores
    // per lane that each touched 64 columns
    double* Ut = at<double>(arena, side[b].off_ut) + (int64_t)pb * NB * lda + col;
#pragma unroll
    for (int i = 0; i < NB; i++) Ut[i * lda] = x[i];
}

// Later panels' row interchanges applied to the L rows of earlier panels of
// the step: wavefront w handles buffer q = qbegin + w and applies the pivots
// of panels max(pfirst, q+1) .. plast in order to its rows below them (panel
// p's rows start (p - q) NB into buffer q).  One lane per column of the
// panel.
__global__ void __launch_bounds__(64 * (NPANEL - 1))
    k_lu_swap_panel(const BlockDesc* __restrict__ desc, char* __restrict__ arena, const LuSide* __restrict__ side,
                    const BlockOut* __restrict__ out, int k, int qbegin, int pfirst, int plast)
{
    const int b = blockIdx.x;
    const BlockDesc& d = desc[b];
    const int q = qbegin + (threadIdx.x >> 6);
    if (out[b].skip || out[b].delta) return;   // delta form: no interchanges
    double* Pq = panel_buf(arena, side[b], d.lda, q) + (int64_t)(threadIdx.x & 63) * d.lda;
    for (int p = max(pfirst, q + 1); p <= plast; p++)
        if (k + p * NB < d.N) swap_rows_ipiv(Pq, moved_buf(arena, side[b], p), min(NB, d.N - k - p * NB), (p - q) * NB);
}

// A22 -= L21 U12 with FP64 MFMA.  Workgroup tile 128x128 (or 64x64) of A22,
// waves of 64x64 (4x4 tiles of 16x16 per wave, 16 accumulators), K = 64 or
// 128 staged through LDS 16 (one-wave tiles: 8) at a time.

// WM x WN wavefronts of 64 x 64 each: workgroup tile (64 WM) x (64 WN).
// End of a trailing-update tile (64 x 64 per wavefront, swapped-operand
// MFMA layout: lane l, register q of acc[i][j] holds row r0 + 16i + rr,
// column c0 + 16j + cq + 4q, rr = l%16, cq = l/16): acc holds C - L U
// (pre) or -L U, to be added to C (loads of half a 16-row band issued
// before its stores).  NT: the pre tile's stores are non-temporal.
template <bool NT = false>
__device__ __forceinline__ void gemm_store_tile(double* __restrict__ A, int64_t lda, const double4_t (&acc)[4][4],
                                                int r0, int c0, int rr, int cq, int rend, int cend, bool pre)
{
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int row = r0 + 16 * i + rr;
        if (row >= rend) continue;
        if (pre) {
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int col = c0 + 16 * j + cq + 4 * q;
                    if (col < cend) {
                        if constexpr (NT)
                            __builtin_nontemporal_store(acc[i][j][q], &A[row + (int64_t)col * lda]);
                        else
                            A[row + (int64_t)col * lda] = acc[i][j][q];
                    }
                }
        } else {
            // half a band at a time (8 loads in flight: no spills beside
            // the accumulators at three workgroups per CU)
#pragma unroll
            for (int h = 0; h < 4; h += 2) {
                double cv[2][4];
#pragma unroll
                for (int j = 0; j < 2; j++)
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const int col = c0 + 16 * (h + j) + cq + 4 * q;
                        cv[j][q] = col < cend ? A[row + (int64_t)col * lda] : 0.0;
                    }
#pragma unroll
                for (int j = 0; j < 2; j++)
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const int col = c0 + 16 * (h + j) + cq + 4 * q;
                        if (col < cend) A[row + (int64_t)col * lda] = cv[j][q] + acc[i][h + j][q];
                    }
            }
        }
    }
}

// The pivots u_j = 1 + d_jj of the step's columns k + kk0 .. k + kk0 + KK
// into LDS (k_lu_diag keeps panel q's 64 at the start of U-buffer row
// q NB); the first barrier after it publishes them.
__device__ __forceinline__ void load_pivots(double* su, const double* Ut, int64_t lda, int kk0, int KK, int tid,
                                            int nthreads)
{
    for (int t = tid; t < KK; t += nthreads) {
        const int j = kk0 + t;
        su[t] = Ut[(int64_t)(j / NB) * NB * lda + j % NB];
    }
}

template <int WM, int WN>
__global__ void __launch_bounds__(64 * WM * WN) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_lu_gemm(const BlockDesc* __restrict__ desc, char* __restrict__ arena, const LuSide* __restrict__ side,
          const BlockOut* __restrict__ out, int k, int r_off, int r_lim, int c_off, int c_lim, int kk0, int KK,
          int gx, int gy, int nblocks, int xcd_order, int lower)
{
    constexpr int NT = 64 * WM * WN, GTM = 64 * WM, GTN = 64 * WN;
    constexpr int KS = NT >= 256 ? LU_KS : 8;   // K per LDS stage (prefetch registers)
    // padded LDS rows: GTPM keeps the A rows 128-byte aligned (its stores and
    // reads run along rows); the B stores run along rows too (U from the
    // row-major U buffer), GTPN odd as when they ran down columns
    constexpr int GTPM = GTM + 16, GTPN = GTN + 17;
    // XCD-aware order (xcd_order = 1): workgroup i runs on XCD i % 8, so all
    // tiles of block b go to XCD b % 8 and share its L2 for the L and U
    // operands; xcd_order = 0: block-major order
    const int wg = blockIdx.x, tiles = gx * gy;
    int b, tile;
    if (xcd_order) {
        const int xcd = wg & 7, local = wg >> 3;
        b = (local / tiles) * 8 + xcd;
        tile = local % tiles;
    } else {
        b = wg / tiles;
        tile = wg % tiles;
    }
    if (b >= nblocks) return;
    const BlockDesc& d = desc[b];
    if (out[b].skip) return;
    const int N = d.N;
    // the updated region: rows [k + r_off, k + r_off + r_lim) and columns
    // [k + c_off, k + c_off + c_lim), both cut at N; depth: the step's L
    // columns and U rows k + kk0 .. k + kk0 + KK (the caller keeps
    // kk0 + KK <= r_off, so they exist whenever the region does)
    const int r0 = k + r_off, rend = min(N, r0 + r_lim);
    const int c0 = k + c_off, cend = min(N, c0 + c_lim);
    const int m0 = r0 + (tile % gx) * GTM;   // rows of the update
    const int n0 = c0 + (tile / gx) * GTN;   // columns of the update
    if (m0 >= rend || n0 >= cend) return;
    if (lower && m0 + GTM <= n0) return;   // symmetric form: a tile wholly above the diagonal
    const int64_t lda = d.lda;
    double* A = at<double>(arena, d.off_mat);
    const double* Pb = panel_buf(arena, side[b], lda, 0);
    const double* Ut = at<double>(arena, side[b].off_ut);
    // one LDS buffer: the A and B stages
    constexpr int SMEM = KS * (GTPM + GTPN);
    __shared__ double smem[SMEM];
    // symmetric form: U row j = u_j L column j (U12 = D11 L21^T), with the
    // pivots u_j = 1 + the factored diagonal blocks' stored diagonals
    __shared__ double su[NPANEL * NB];
    if (lower) load_pivots(su, Ut, lda, kk0, KK, threadIdx.x, NT);
    double(*As)[GTPM] = reinterpret_cast<double(*)[GTPM]>(smem);             // [kk][row]
    double(*Bs)[GTPN] = reinterpret_cast<double(*)[GTPN]>(smem + KS * GTPM);  // [kk][col]
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int wr = (wv % WM) * 64, wc = (wv / WM) * 64;
    const int li = lane & 15, lk = lane >> 4;
    // With the operands swapped the MFMA yields the transposed tile: lane l,
    // register q of acc[i][j] holds C row wr + 16i + l%16, column
    // wc + 16j + l/16 + 4q, so each register's 64 lanes cover four whole
    // 128-byte column segments.  The accumulators take -L U.  Delta-form
    // blocks (entries of -M, small): they start as the C tile (loaded beside
    // the first stage) and the tile ends with stores only.  Pivoted blocks
    // (unit diagonal): they start at zero and C + acc is formed at the end,
    // one rounding at the scale of C per update instead of one per MFMA.
    const int rr = lane & 15, cq = lane >> 4;
    const bool pre = out[b].delta;
    double4_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int row = m0 + wr + 16 * i + rr;
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int col = n0 + wc + 16 * j + cq + 4 * q;
                acc[i][j][q] = (pre && row < rend && col < cend) ? A[row + (int64_t)col * lda] : 0.0;
            }
    }
    // the next stage's operands are fetched into registers while the
    // current stage runs on the MFMA units
    constexpr int PERA = (KS * GTM) / NT, PERB = (KS * GTN) / NT;
    double pa[PERA], pb[PERB];
    auto fetch = [&](int ks) {
        // L columns k + kk0 + ks .. +KS, all in panel q = (kk0 + ks) / NB,
        // whose buffer rows start at matrix row k + q NB
        const int j = kk0 + ks, q = j / NB;
        const double* Lp = Pb + (int64_t)q * lda * NB + (int64_t)(j % NB) * lda - (k + q * NB);
#pragma unroll
        for (int t = 0; t < PERA; t++) {
            const int idx = tid + NT * t;
            const int kk = idx / GTM, r = idx % GTM;
            const int row = m0 + r;
            pa[t] = row < rend ? Lp[row + (int64_t)kk * lda] : 0.0;
        }
        // U rows from the step's row-major U buffer, or (symmetric form) L
        // columns at rows n0 .. to be scaled by u_j (consecutive lanes read
        // consecutive columns)
#pragma unroll
        for (int t = 0; t < PERB; t++) {
            const int idx = tid + NT * t;
            const int kk2 = idx / GTN, c = idx % GTN;
            const int colg = n0 + c;
            pb[t] = colg >= cend ? 0.0
                  : lower        ? Lp[colg + (int64_t)kk2 * lda]
                                 : Ut[(int64_t)(kk0 + ks + kk2) * lda + colg];
        }
    };
    fetch(0);
    for (int ks = 0; ks < KK; ks += KS) {
        __syncthreads();
#pragma unroll
        for (int t = 0; t < PERA; t++) {
            const int idx = tid + NT * t;
            As[idx / GTM][idx % GTM] = pa[t];
        }
#pragma unroll
        for (int t = 0; t < PERB; t++) {
            const int idx = tid + NT * t;
            Bs[idx / GTN][idx % GTN] = lower ? pb[t] * su[ks + idx / GTN] : pb[t];
        }
        __syncthreads();
        if (ks + KS < KK) fetch(ks + KS);
#pragma unroll
        for (int kk = 0; kk < KS; kk += 4) {
            double av[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; i++) av[i] = -As[kk + lk][wr + 16 * i + li];
#pragma unroll
            for (int j = 0; j < 4; j++) bv[j] = Bs[kk + lk][wc + 16 * j + li];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j], av[i], acc[i][j], 0, 0, 0);
        }
    }
    gemm_store_tile(A, lda, acc, m0 + wr, n0 + wc, rr, cq, rend, cend, pre);
}

// A22 -= L21 U12 on 128x128 tiles (2x2 wavefronts of 64x64) with the
// operand stages moved straight from HBM into LDS (global_load_lds_dwordx4:
// no staging registers, no LDS store instructions) and double-buffered, so
// one barrier per stage suffices and the next stage's loads run under the
// current stage's MFMAs.  L comes from the step's panel buffers,
// U from the row-major U buffer, so both stages are rows of 128 contiguous
// doubles, one 1 KB row per wavefront instruction (lane l moves elements 2l
// and 2l+1).  The loads are issued from inline assembly: the compiler's
// wait-count tracking would otherwise make every LDS read wait for all
// outstanding LDS-bound loads, the next stage's included.  They are
// waited for explicitly (vmcnt(0) before the barrier); the compiler's own
// vmcnt waits stay correct, only stricter, because loads retire in order.
// Rows (columns) past the region load from inside the padded leading
// dimension (lda > N when N is odd) or not at all, and only feed rows
// (columns) of the product that are not stored.
typedef __attribute__((address_space(3))) void lds_void;
// KS = 16: 73.7 KB of LDS, two workgroups per CU; KS = 8: 36.9 KB, OCC
// workgroups per CU as the registers allow.  NT: the C tile is loaded and
// stored non-temporally (launches whose C region is larger than the caches)
template <int KS, int OCC, bool NT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC, OCC)))
k_lu_gemm_dma(const BlockDesc* __restrict__ desc, char* __restrict__ arena, const LuSide* __restrict__ side,
              const BlockOut* __restrict__ out, int k, int r_off, int r_lim, int c_off, int c_lim, int kk0, int KK,
              int gx, int gy, int nblocks, int xcd_order, int lower, int tri)
{
    constexpr int GT = 128, PITCH = GT + 16;
    constexpr int KW = KS / 4;   // kk rows of a stage moved by each wavefront
    // tri: a square region of the symmetric form (gx = gy), launched over
    // its lower triangle of tiles only (row tile >= column tile)
    const int wg = blockIdx.x, tiles = tri ? gx * (gx + 1) / 2 : gx * gy;
    int b, tile;
    if (xcd_order) {
        const int xcd = wg & 7, local = wg >> 3;
        b = (local / tiles) * 8 + xcd;
        tile = local % tiles;
    } else {
        b = wg / tiles;
        tile = wg % tiles;
    }
    if (b >= nblocks) return;
    const BlockDesc& d = desc[b];
    if (out[b].skip) return;
    const int N = d.N;
    const int r0 = k + r_off, rend = min(N, r0 + r_lim);
    const int c0 = k + c_off, cend = min(N, c0 + c_lim);
    int ti = tile % gx, tj = tile / gx;
    if (tri) {
        // column-major over the triangle: column tile j holds row tiles j .. gx - 1
        tj = 0;
        ti = tile;
        while (ti >= gx - tj) {
            ti -= gx - tj;
            tj++;
        }
        ti += tj;
    }
    const int m0 = r0 + ti * GT;
    const int n0 = c0 + tj * GT;
    if (m0 >= rend || n0 >= cend) return;
    if (lower && m0 + GT <= n0) return;   // symmetric form: a tile wholly above the diagonal
    const int64_t lda = d.lda;
    double* A = at<double>(arena, d.off_mat);
    const double* Pb = panel_buf(arena, side[b], lda, 0);
    const double* Ut = at<double>(arena, side[b].off_ut);
    __shared__ double sA[2][KS][PITCH];   // [stage][kk][row]
    __shared__ double sB[2][KS][PITCH];   // [stage][kk][col]
    // symmetric form: U row j = u_j L column j (U12 = D11 L21^T); the
    // stage holds the L column at rows n0 .., scaled when read
    __shared__ double su[NPANEL * NB];
    if (lower) load_pivots(su, Ut, lda, kk0, KK, threadIdx.x, 256);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wr = (wv & 1) * 64, wc = (wv >> 1) * 64;
    const int li = lane & 15, lk = lane >> 4;
    const bool arow_ok = m0 + 2 * lane < rend, bcol_ok = n0 + 2 * lane < cend;
    // stage of depth KS at kk0 + ks into buffer buf: wavefront wv moves kk
    // rows KW wv .. KW wv + KW - 1 of both operands
    auto issue = [&](int ks, int buf) {
        const int j = kk0 + ks, q = j / NB;
        const double* Lq = Pb + (int64_t)q * lda * NB + (int64_t)(j % NB) * lda - (k + q * NB);
        const double* Lp = Lq + m0 + 2 * lane;
        const double* Up = lower ? Lq + n0 + 2 * lane : Ut + (int64_t)j * lda + n0 + 2 * lane;
#pragma unroll
        for (int t = 0; t < KW; t++) {
            const int kk = KW * wv + t;
            const uint32_t la = __builtin_amdgcn_readfirstlane((uint32_t)(uintptr_t)(lds_void*)&sA[buf][kk][0]);
            const uint32_t lb = __builtin_amdgcn_readfirstlane((uint32_t)(uintptr_t)(lds_void*)&sB[buf][kk][0]);
            if (arow_ok)
                asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off"
                             :: "s"(la), "v"(Lp + (int64_t)kk * lda) : "memory", "m0");
            if (bcol_ok)
                asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off"
                             :: "s"(lb), "v"(Up + (int64_t)kk * lda) : "memory", "m0");
        }
    };
    // swapped operands: lane l, register q of acc[i][j] holds C row
    // wr + 16i + l%16, column wc + 16j + l/16 + 4q.  The accumulators take
    // -L U.  Delta-form blocks (entries of -M, small): they start as the C
    // tile (loaded beside the first stage, one latency for both) and the
    // tile ends with stores only.  Pivoted blocks (unit diagonal): they
    // start at zero and C + acc is formed at the end, one rounding at the
    // scale of C per update instead of one per MFMA.  The C tile is read
    // and written once per launch; non-temporally (NT) it does not evict the
    // L and U rows that the block's other tiles read again from L2.
    const int rr = lane & 15, cq = lane >> 4;
    const bool pre = out[b].delta;
    // symmetric form: a wavefront whose 64 x 64 quadrant lies above the
    // diagonal and outside the 64 x 64 diagonal blocks (the upper right one
    // of a diagonal tile; nothing reads those entries) computes and stores
    // nothing; it still moves its share of every stage
    const bool idle = lower && (m0 + wr + 63) / NB < (n0 + wc) / NB;
    double4_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int row = m0 + wr + 16 * i + rr;
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int col = n0 + wc + 16 * j + cq + 4 * q;
                acc[i][j][q] = 0.0;
                if (pre && !idle && row < rend && col < cend) {
                    const double* cp = A + row + (int64_t)col * lda;
                    acc[i][j][q] = NT ? __builtin_nontemporal_load(cp) : *cp;
                }
            }
    }
    issue(0, 0);
    const int S = KK / KS;
    for (int s = 0; s < S; s++) {
        const int buf = s & 1;
        __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0): this wavefront's stage-s loads are in LDS
        __syncthreads();                       // everyone's are, and stage s-1's buffer is free
        if (s + 1 < S) issue((s + 1) * KS, buf ^ 1);
        if (idle) continue;   // uniform over the wavefront
        // symmetric form: this lane's pivots of the stage (rows kk + lk)
        double us[KS / 4];
        if (lower) {
#pragma unroll
            for (int t = 0; t < KS / 4; t++) us[t] = su[s * KS + 4 * t + lk];
        }
#pragma unroll
        for (int kk = 0; kk < KS; kk += 4) {
            double av[4], bv[4];
#pragma unroll
            for (int j = 0; j < 4; j++) bv[j] = sB[buf][kk + lk][wc + 16 * j + li];
#pragma unroll
            for (int i = 0; i < 4; i++) av[i] = -sA[buf][kk + lk][wr + 16 * i + li];
            if (lower) {
#pragma unroll
                for (int j = 0; j < 4; j++) bv[j] *= us[kk / 4];
            }
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j], av[i], acc[i][j], 0, 0, 0);
        }
    }
    if (!idle) gemm_store_tile<NT>(A, lda, acc, m0 + wr, n0 + wc, rr, cq, rend, cend, pre);
}

// Sum of log|u_ii| (the reference sums them with Kahan compensation,
// src/matrix.c:462-471), in double-double arithmetic over the wavefront.
__global__ void __launch_bounds__(64) k_lu_finalize(const BlockDesc* __restrict__ desc, char* __restrict__ arena,
                                                   const LuSide* __restrict__ side, BlockOut* __restrict__ out,
                                                   int nblocks)
{
    const int b = blockIdx.x;   // one wavefront per block
    if (b >= nblocks || out[b].skip) return;
    const double s = wave_sum_dd(at<double>(arena, side[b].off_logs), desc[b].N);
    if (threadIdx.x == 0) out[b].logdet = s;
}

// Synthetic LU input (autotuning), X_ij uniform in [-1, 1) from a
// counter-based hash of (seed, block, i, j), column-major: A = 1 - 0.3 X /
// sqrt(N) (pivoted form), or -M with M = 0.3 X / N (delta form: 1 - M
// column diagonally dominant, column sums of |M| about 0.15).
__global__ void k_lu_fill(const BlockDesc* __restrict__ desc, char* __restrict__ arena, unsigned long long seed,
                          int delta)
{
    const BlockDesc& d = desc[blockIdx.y];
    const int N = d.N;
    double* A = at<double>(arena, d.off_mat);
    const double sc = delta ? 0.3 / N : 0.3 / sqrt((double)N);
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < (int64_t)N * N;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % N), j = (int)(e / N);
        unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (e + 1 + ((unsigned long long)blockIdx.y << 40));
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const double x = (double)(z >> 11) * 0x1p-53 * 2 - 1;
        A[i + (int64_t)j * d.lda] = (i == j && !delta ? 1.0 : 0.0) - sc * x;
    }
}

// ---- allocation-free LU (capsx_lu_logdet_batch_noalloc) ---------------------
// Its descriptors and the LU side buffers live in this fixed device array
// (each GPU holds its own copy until the process ends); host code never
// uses the array's name as an address, only the pointer that
// hipGetSymbolAddress returns for the device (noalloc_scratch_ptr).
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];

// Descriptors of n column-major dim x dim matrices at the caller's pointer
// (arena offset a_off, matrix b at a_off + b lda dim doubles): BlockDesc at
// scratch offset 0, LuSide at off_side, BlockOut at off_out (zeroed: no
// shortcut, explicit form, so the LU runs with partial pivoting as for any
// matrix), and per block the side buffers from off_buf on, `per` bytes per
// block in the layout of the library's lu_side_layout.
__global__ void k_lu_noalloc_setup(char* __restrict__ base, int64_t a_off, int dim, int lda, int n, int64_t off_side,
                                   int64_t off_out, int64_t off_buf, int64_t per, int64_t o_logs, int64_t o_moved,
                                   int64_t o_ut, int64_t o_inv)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    BlockDesc d = {};
    d.N = dim;
    d.ldim = (dim + 1) / 2;
    d.lda = lda;
    d.off_mat = a_off + (int64_t)b * lda * dim * 8;
    reinterpret_cast<BlockDesc*>(base)[b] = d;
    LuSide s;
    s.off_panel = off_buf + b * per;
    s.off_logs = s.off_panel + o_logs;
    s.off_moved = s.off_panel + o_moved;
    s.off_ut = s.off_panel + o_ut;
    s.off_inv = s.off_panel + o_inv;
    reinterpret_cast<LuSide*>(base + off_side)[b] = s;
    BlockOut o = {};
    o.logdet = NAN;
    reinterpret_cast<BlockOut*>(base + off_out)[b] = o;
}

__global__ void k_lu_noalloc_result(const BlockOut* __restrict__ out, double* __restrict__ logdet, int n)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < n) logdet[b] = out[b].logdet;
}

// MFMA operand-layout probe: D = A(16x4) B(4x16) for one wave.  A is
// row-major 16x4, B row-major 4x16, D row-major 16x16.
__global__ void k_mfma_probe(const double* __restrict__ A, const double* __restrict__ B, double* __restrict__ D)
{
    const int lane = threadIdx.x;
    const double a = A[(lane & 15) * 4 + (lane >> 4)];
    const double bb = B[(lane >> 4) * 16 + (lane & 15)];
    double4_t acc = (double4_t){0, 0, 0, 0};
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bb, acc, 0, 0, 0);
#pragma unroll
    for (int q = 0; q < 4; q++) D[((lane >> 4) + 4 * q) * 16 + (lane & 15)] = acc[q];
}

}  // namespace capsd

using namespace capsd;

extern "C" {

int capsd_lu_max_dim(void) { return 24 * 1024; }
int capsd_lu_nb(void) { return NB; }

// Panels taller than grid_panel_rows use the column-at-a-time grid kernels
// (CAPS_AMD_GRID_PANEL_ROWS, default 6144: above it the one-workgroup panel
// needs 24 rows per lane).
static int grid_panel_rows = 6144;
// one-workgroup panel variants (CAPS_AMD_PANEL_SET): 1 (default) = leaves
// of 8 columns with 256-lane workgroups up to 2560 rows, 0 = leaves of 4
// above 512 rows.  Both give bit-identical factors: every entry subtracts
// the same products in the same column order.
static int panel_set = 1;
// trailing-update kernel of the 128 x 128 tile (CAPS_AMD_GEMM_KERNEL):
// k_lu_gemm_dma (operand stages moved straight into LDS, double-buffered)
// with 8-deep stages and three workgroups per CU (2, the default) or 16-deep
// stages and two (1), or k_lu_gemm (stages through registers, 0)
static int gemm_kernel = 2;
// 128 x 64 and 64 x 128 tiles for regions one panel wide or high, with the
// 128 x 128 tile (CAPS_AMD_NARROW_TILES, default 1)
static int narrow_tiles = 1;
// columns per workgroup of the delta form's U12 kernel (CAPS_AMD_TRSM_COLS:
// 64, the default, or 128)
static int trsm_cols = 64;

static void launch_panel(int rows, const BlockDesc* desc, char* arena, const LuSide* side, const BlockOut* out, int nb,
                         int k, int pb, hipStream_t s)
{
#define PANEL(RR, LF, NT, OCC)                                                                              \
    hipLaunchKernelGGL((k_lu_panel<RR, LF, NT, OCC>), dim3(nb), dim3(NT), (RR) * (NT) * sizeof(int), s, desc, arena, \
                       side, out, k, pb)
#define PANELC(RR, LF, NT, OCC, CW)                                                                       \
    hipLaunchKernelGGL((k_lu_panel<RR, LF, NT, OCC, CW>), dim3(nb), dim3(NT), (RR) * (NT) * sizeof(int), s, desc, \
                       arena, side, out, k, pb)
    if (rows > grid_panel_rows) {
        const int nch = (rows + GP_ROWS - 1) / GP_ROWS;
        hipLaunchKernelGGL(k_gp_max, dim3(nch, nb), dim3(GP_ROWS), 0, s, desc, arena, side, out, k, pb);
        for (int c = 0; c < NB; c++) {
            hipLaunchKernelGGL(k_gp_pivot, dim3(nb), dim3(64), 0, s, desc, arena, side, out, k, pb, c);
            hipLaunchKernelGGL(k_gp_update, dim3(nch, nb), dim3(GP_ROWS), 0, s, desc, arena, side, out, k, pb, c);
        }
        hipLaunchKernelGGL(k_gp_copy, dim3(nch, nb), dim3(GP_ROWS), 0, s, desc, arena, side, out, k, pb);
    } else if (panel_set == 1) {
        // leaves of 8 left-looking in chunks of 4 (2 at 2560 rows), two
        // 256-lane workgroups per CU up to 2560 rows
        if (rows <= 2 * 256) PANEL(2, 8, 256, 2);
        else if (rows <= 4 * 256) PANELC(4, 8, 256, 2, 4);
        else if (rows <= 8 * 256) PANELC(8, 8, 256, 2, 4);
        else if (rows <= 10 * 256) PANELC(10, 8, 256, 2, 2);
        else if (rows <= 6 * 512) PANELC(6, 8, 512, 2, 4);
        else if (rows <= 12 * 512) PANELC(12, 8, 512, 2, 1);
        else PANEL(24, 2, 1024, 4);
    } else {
        // leaves of 4 above 512 rows
        if (rows <= 2 * 256) PANEL(2, 8, 256, 2);
        else if (rows <= 4 * 256) PANEL(4, 4, 256, 2);
        else if (rows <= 8 * 256) PANEL(8, 4, 256, 2);
        else if (rows <= 6 * 512) PANEL(6, 4, 512, 2);
        else if (rows <= 12 * 512) PANEL(12, 4, 512, 2);
        else PANEL(24, 2, 1024, 4);
    }
#undef PANEL
#undef PANELC
}

int capsd_noalloc_scratch(char** ptr, size_t* bytes);

// Library set-up on the current device: variant switches from the
// environment, the 1024-lane panel's dynamic LDS limit, and the
// allocation-free LU's scratch address for this device.
int capsd_lu_init(void)
{
    if (const char* e = getenv("CAPS_AMD_GRID_PANEL_ROWS")) grid_panel_rows = atoi(e);
    if (const char* e = getenv("CAPS_AMD_PANEL_SET")) panel_set = atoi(e);
    if (const char* e = getenv("CAPS_AMD_GEMM_KERNEL")) gemm_kernel = atoi(e);
    if (const char* e = getenv("CAPS_AMD_NARROW_TILES")) narrow_tiles = atoi(e);
    if (const char* e = getenv("CAPS_AMD_TRSM_COLS")) trsm_cols = atoi(e) == 128 ? 128 : 64;
    char* p;
    size_t bytes;
    return capsd_noalloc_scratch(&p, &bytes);
}

// Kernel variants (values < 0 keep the current one; prev receives the
// values before the call): the 128 x 128 trailing-update kernel (0, 1, 2),
// the panel variant set (0, 1), the narrow tiles (0, 1).  Returns 1 for
// invalid values.
int capsd_lu_kernels(int gk, int ps, int nt, int prev[3])
{
    if (prev) {
        prev[0] = gemm_kernel;
        prev[1] = panel_set;
        prev[2] = narrow_tiles;
    }
    if (gk > 2 || ps > 1 || nt > 1) return 1;
    if (gk >= 0) gemm_kernel = gk;
    if (ps >= 0) panel_set = ps;
    if (nt >= 0) narrow_tiles = nt;
    return 0;
}

// Panel height above which the grid-wide panel kernels run; returns the
// previous value (rows <= 0: query only).
int capsd_lu_grid_panel_rows(int rows)
{
    const int old = grid_panel_rows;
    if (rows > 0) grid_panel_rows = rows;
    return old;
}

// Upper bound of the number of trailing-update launches of one LU (the
// four-level schedule has 20 per 512 columns, the others fewer).
int capsd_lu_max_updates(int maxN) { return 20 * ((maxN + 8 * NB - 1) / (8 * NB)) + 2; }

// Panel buffers per block (LuSide::off_panel holds this many lda x NB panels).
int capsd_lu_npanels(void) { return NPANEL; }

// Whole LU of a batch (host loop over steps; all launches on stream s).
//   levels 1: NB-column steps: panel, swaps + TRSM, one K = NB update.
//   levels 2: 2NB-column steps: panel 0, its swaps + TRSM of the second
//     panel's columns and their update, panel 1, panel 1's pivots into
//     panel 0's L rows, swaps of both panels + TRSM of the first 64 U rows
//     of the trailing columns, a K = NB update of the next 64, their TRSM,
//     one K = 2NB update of the trailing matrix.
//   levels 3: 4NB-column steps: the levels-2 step on the step's own 256
//     columns (panels 0, 1 and the K = 2NB update of columns 128..256), the
//     same for panels 2 and 3 on columns 128..256, the later panels' pivots
//     into the earlier panels' L rows; for the trailing columns all four
//     panels' swaps and the 256-row TRSM as four 64-row TRSMs with three
//     strip updates between them; one K = 4NB update of the trailing matrix
//     (half the read-modify-write of C per flop of levels 2).
//   levels 4: 8NB-column steps: the levels-3 step on panels 0..3 applied to
//     the step's 512 columns, the same on panels 4..7 applied to their own,
//     the pivots of panels 4..7 into the L rows of panels 0..3; for the
//     trailing columns all eight panels' swaps and the 512-row TRSM (eight
//     64-row TRSMs with the updates from the panels above between them);
//     one K = 8NB update of the trailing matrix.
// tile: 128 (update workgroups of 128 x 128, 4 wavefronts), 256 (256 x 128,
// 8 wavefronts) or 64 (64 x 64, one wavefront); xcd_order: tile order of the
// update (k_lu_gemm).
// gemm_ev: NULL, or 2 events per update launch recorded around it;
// gemm_info receives (k, r_off, r_lim, c_off, c_lim, KK, dma) of each update
// launch (dma = 1 for k_lu_gemm_dma, 0 for the k_lu_gemm tiles).  kinds: bit 0 set if the batch may hold blocks for the pivoted LU
// (out[b].delta = 0), bit 1 if it may hold delta-form blocks; kernels for a
// kind the caller rules out are not launched; bit 2 (with bit 1 alone): the
// delta-form blocks are symmetric (1 - M of a round trip): the updates take
// U12 = D11 L21^T from the L21 panel buffers and the pivots (no U buffer),
// no TRSM or update of U rows runs, and the trailing updates skip the tiles
// wholly above the diagonal (the diagonal blocks are kept whole for
// k_lu_diag).  Returns the number of update launches.
int capsd_launch_lu(const BlockDesc* desc, char* arena, const void* side, BlockOut* out, int nblocks, int maxN,
                    hipStream_t s, int levels, int xcd_order, int tile, hipEvent_t* gemm_ev, int* gemm_info, int kinds)
{
    const bool piv = kinds & 1, dlt = kinds & 2;
    // symmetric delta form (bit 2, delta-form blocks only): U12 = D11 L21^T
    // is formed by the updates from k_lu_l21's L21, so no TRSM and no update
    // of U rows runs, and the trailing updates keep the lower triangle
    const int sym = (kinds & 4) && dlt && !piv ? 1 : 0;
    const LuSide* sd = (const LuSide*)side;
    const int nb8 = xcd_order ? (nblocks + 7) & ~7 : nblocks;
    constexpr int ALL = 1 << 30;   // region limit: up to N
    int p = 0;
    // A[k+r_off .., k+c_off ..] -= L[.., k+kk0 ..] U[k+kk0 .., ..], depth KK
    auto update = [&](int k, int r_off, int r_lim, int c_off, int c_lim, int kk0, int KK) {
        const int rows = min(maxN, k + r_off + min(r_lim, maxN)) - (k + r_off);
        const int cols = min(maxN, k + c_off + min(c_lim, maxN)) - (k + c_off);
        if (rows <= 0 || cols <= 0) return;
        if (sym && r_off + rows <= c_off) return;   // wholly above the diagonal: U rows
        if (gemm_ev) (void)hipEventRecord(gemm_ev[2 * p], s);
        int dma = 0;   // 1: this launch is k_lu_gemm_dma
        if (tile == 64) {
            const int gx = (rows + 63) / 64, gy = (cols + 63) / 64;
            hipLaunchKernelGGL((k_lu_gemm<1, 1>), dim3(gx * gy * nb8), dim3(64), 0, s, desc, arena, sd, out, k, r_off,
                               r_lim, c_off, c_lim, kk0, KK, gx, gy, nblocks, xcd_order, sym);
        } else if (tile == 256) {
            const int gx = (rows + 255) / 256, gy = (cols + 127) / 128;
            hipLaunchKernelGGL((k_lu_gemm<4, 2>), dim3(gx * gy * nb8), dim3(512), 0, s, desc, arena, sd, out, k,
                               r_off, r_lim, c_off, c_lim, kk0, KK, gx, gy, nblocks, xcd_order, sym);
        } else if (narrow_tiles && cols <= 64 && rows > 64) {
            // regions one panel wide: 128 x 64 tiles (a 128 x 128 tile would
            // compute half of its product for nothing)
            const int gx = (rows + 127) / 128, gy = (cols + 63) / 64;
            hipLaunchKernelGGL((k_lu_gemm<2, 1>), dim3(gx * gy * nb8), dim3(128), 0, s, desc, arena, sd, out, k,
                               r_off, r_lim, c_off, c_lim, kk0, KK, gx, gy, nblocks, xcd_order, sym);
        } else if (narrow_tiles && rows <= 64 && cols > 64) {
            const int gx = (rows + 63) / 64, gy = (cols + 127) / 128;
            hipLaunchKernelGGL((k_lu_gemm<1, 2>), dim3(gx * gy * nb8), dim3(128), 0, s, desc, arena, sd, out, k,
                               r_off, r_lim, c_off, c_lim, kk0, KK, gx, gy, nblocks, xcd_order, sym);
        } else if (gemm_kernel == 0) {
            const int gx = (rows + 127) / 128, gy = (cols + 127) / 128;
            hipLaunchKernelGGL((k_lu_gemm<2, 2>), dim3(gx * gy * nb8), dim3(256), 0, s, desc, arena, sd, out, k,
                               r_off, r_lim, c_off, c_lim, kk0, KK, gx, gy, nblocks, xcd_order, sym);
        } else {
            const int gx = (rows + 127) / 128, gy = (cols + 127) / 128;
            dma = 1;
            // non-temporal C tiles when the launch's C region (the blocks'
            // trailing matrices, the lower half in the symmetric form) is
            // larger than the 256 MB memory-side cache: then nothing keeps
            // it for the next panel's kernels anyway, and cached it would
            // evict the L rows the tiles share
            const bool nt = (double)rows * cols * 8.0 * nblocks * (sym ? 0.5 : 1.0) > 256e6;
            // a square region of the symmetric form: only its lower
            // triangle of tiles is launched (the others would exit at once)
            const int tri = sym && r_off == c_off && rows == cols ? 1 : 0;
            const unsigned ntile = tri ? (unsigned)(gx * (gx + 1) / 2) : (unsigned)(gx * gy);
#define GEMM_DMA(KSV, OCCV, NTV)                                                                                    \
    hipLaunchKernelGGL((k_lu_gemm_dma<KSV, OCCV, NTV>), dim3(ntile * nb8), dim3(256), 0, s, desc, arena, sd, out, k, \
                       r_off, r_lim, c_off, c_lim, kk0, KK, gx, gy, nblocks, xcd_order, sym, tri)
            if (gemm_kernel == 2 && nt)
                GEMM_DMA(8, 3, true);
            else if (gemm_kernel == 2)
                GEMM_DMA(8, 3, false);
            else if (nt)
                GEMM_DMA(16, 2, true);
            else
                GEMM_DMA(16, 2, false);
#undef GEMM_DMA
        }
        if (gemm_ev) (void)hipEventRecord(gemm_ev[2 * p + 1], s);
        if (gemm_info) {
            int* g = gemm_info + 7 * p;
            g[0] = k, g[1] = r_off, g[2] = r_lim, g[3] = c_off, g[4] = c_lim, g[5] = KK, g[6] = dma;
        }
        p++;
    };
    // swaps (bit q: panel q's pivots) and TRSM of rows k + pb NB .. +NB in
    // the columns [k + pb NB + cs_off, +cols)
    // pivoted blocks: k_lu_swap_trsm; delta-form blocks: k_lu_trsm_mfma
    // (each kernel leaves the other kind of block alone)
    auto trsm = [&](int k, int pb, int cs_off, int cols, int swaps) {
        const int c0 = k + pb * NB + cs_off;
        const int n = min(maxN - c0, cols);
        if (n <= 0) return;
        if (piv && n > 64)
            hipLaunchKernelGGL(k_lu_swap_trsm<256>, dim3((n + 255) / 256, nblocks), dim3(256), 0, s, desc, arena, sd,
                               out, k, pb, cs_off, cols, swaps);
        else if (piv)
            hipLaunchKernelGGL(k_lu_swap_trsm<64>, dim3((n + 63) / 64, nblocks), dim3(64), 0, s, desc, arena, sd, out,
                               k, pb, cs_off, cols, swaps);
        if (sym) return;   // the U rows come from k_lu_l21
        if (dlt && trsm_cols == 128)
            hipLaunchKernelGGL(k_lu_trsm_mfma<128>, dim3((n + 127) / 128, nblocks), dim3(256), 0, s, desc, arena, sd,
                               out, k, pb, cs_off, cols);
        else if (dlt)
            hipLaunchKernelGGL(k_lu_trsm_mfma<64>, dim3((n + 63) / 64, nblocks), dim3(256), 0, s, desc, arena, sd,
                               out, k, pb, cs_off, cols);
    };
    // pivoted blocks: the one-workgroup (or grid-wide) panel; delta-form
    // blocks: the diagonal block, then L21 = A21 U11^{-1}
    auto panel = [&](int k, int pb) {
        const int k0 = k + pb * NB;
        if (k0 >= maxN) return;
        if (piv) launch_panel(maxN - k0, desc, arena, sd, out, nblocks, k0, pb, s);
        if (!dlt) return;
        hipLaunchKernelGGL(k_lu_diag, dim3(nblocks), dim3(256), 0, s, desc, arena, sd, out, k, pb);
        const int below = maxN - k0 - NB;
        // whole 64 x 64 tiles per wavefront while they give two workgroups
        // per CU or more, else one column tile per wavefront
        if (below > 0) {
            if ((int64_t)((below + 255) / 256) * nblocks >= 512)
                hipLaunchKernelGGL(k_lu_l21<4>, dim3((below + 255) / 256, nblocks), dim3(256), 0, s, desc, arena, sd,
                                   out, k, pb);
            else
                hipLaunchKernelGGL(k_lu_l21<1>, dim3((below + 63) / 64, nblocks), dim3(256), 0, s, desc, arena, sd,
                                   out, k, pb);
        }
    };
    // buffers [qbegin, qend) take the pivots of the later panels pfirst..plast
    auto swap_panels = [&](int k, int qbegin, int qend, int pfirst, int plast) {
        if (piv && k + pfirst * NB < maxN)
            hipLaunchKernelGGL(k_lu_swap_panel, dim3(nblocks), dim3(64 * (qend - qbegin)), 0, s, desc, arena, sd, out,
                               k, qbegin, pfirst, plast);
    };
    // the two-panel step at column k0 = k + h NB (panels h, h+1) applied to
    // the columns up to k0 + 2NB + tail: its own second panel and `tail`
    // columns beyond it
    auto step2 = [&](int k, int h, int tail) {
        const int o = h * NB;
        panel(k, h);
        trsm(k, h, NB, NB, 1 << h);
        update(k, o + NB, ALL, o + NB, NB, o, NB);
        panel(k, h + 1);
        swap_panels(k, h, h + 1, h + 1, h + 1);
        trsm(k, h, 2 * NB, tail, 3 << h);
        update(k, o + NB, NB, o + 2 * NB, tail, o, NB);
        trsm(k, h + 1, NB, tail, 0);
        update(k, o + 2 * NB, ALL, o + 2 * NB, tail, o, 2 * NB);
    };
    // the four-panel step on panels h..h+3 (columns from k0 = k + h NB)
    // applied to the columns up to k0 + 4NB + tail: panels h, h+1 with the
    // two-panel step (applied to the step's columns), panels h+2, h+3, their
    // pivots into the first two panels' L rows; for the columns beyond the
    // step's own the four panels' swaps and U rows (64-row TRSMs with the
    // updates from the panels above between them) and one K = 4NB update
    auto step4 = [&](int k, int h, int tail) {
        const int o = h * NB;
        step2(k, h, 2 * NB);
        step2(k, h + 2, 0);
        swap_panels(k, h, h + 2, h + 2, h + 3);
        trsm(k, h, 4 * NB, tail, 15 << h);
        update(k, o + NB, NB, o + 4 * NB, tail, o, NB);
        trsm(k, h + 1, 3 * NB, tail, 0);
        update(k, o + 2 * NB, 2 * NB, o + 4 * NB, tail, o, 2 * NB);
        trsm(k, h + 2, 2 * NB, tail, 0);
        update(k, o + 3 * NB, NB, o + 4 * NB, tail, o + 2 * NB, NB);
        trsm(k, h + 3, NB, tail, 0);
        update(k, o + 4 * NB, ALL, o + 4 * NB, tail, o, 4 * NB);
    };
    if (levels <= 1) {
        for (int k = 0; k < maxN; k += NB) {
            panel(k, 0);
            trsm(k, 0, NB, ALL, 1);
            update(k, NB, ALL, NB, ALL, 0, NB);
        }
    } else if (levels == 2) {
        for (int k = 0; k < maxN; k += 2 * NB) step2(k, 0, ALL);
    } else if (levels == 3) {
        for (int k = 0; k < maxN; k += 4 * NB) step4(k, 0, ALL);
    } else {
        // 8NB-column steps: the four-panel step on panels 0..3 applied to
        // the step's own columns, then on panels 4..7, the later panels'
        // pivots into the earlier panels' L rows; for the trailing columns
        // the eight panels' swaps and their 512 U rows (forward
        // substitution in 64-row TRSMs and updates from the panels above);
        // one K = 8NB update of the trailing matrix (half the C traffic per
        // flop of levels 3)
        for (int k = 0; k < maxN; k += 8 * NB) {
            step4(k, 0, 4 * NB);
            step4(k, 4, 0);
            swap_panels(k, 0, 4, 4, 7);
            trsm(k, 0, 8 * NB, ALL, 255);
            update(k, NB, NB, 8 * NB, ALL, 0, NB);
            trsm(k, 1, 7 * NB, ALL, 0);
            update(k, 2 * NB, 2 * NB, 8 * NB, ALL, 0, 2 * NB);
            trsm(k, 2, 6 * NB, ALL, 0);
            update(k, 3 * NB, NB, 8 * NB, ALL, 2 * NB, NB);
            trsm(k, 3, 5 * NB, ALL, 0);
            update(k, 4 * NB, 4 * NB, 8 * NB, ALL, 0, 4 * NB);
            trsm(k, 4, 4 * NB, ALL, 0);
            update(k, 5 * NB, NB, 8 * NB, ALL, 4 * NB, NB);
            trsm(k, 5, 3 * NB, ALL, 0);
            update(k, 6 * NB, 2 * NB, 8 * NB, ALL, 4 * NB, 2 * NB);
            trsm(k, 6, 2 * NB, ALL, 0);
            update(k, 7 * NB, NB, 8 * NB, ALL, 6 * NB, NB);
            trsm(k, 7, NB, ALL, 0);
            update(k, 8 * NB, ALL, 8 * NB, ALL, 0, 8 * NB);
        }
    }
    hipLaunchKernelGGL(k_lu_finalize, dim3(nblocks), dim3(64), 0, s, desc, arena, sd, out, nblocks);
    return p;
}

// Device address of the allocation-free LU's scratch array on the current
// device, looked up once per device (hipGetSymbolAddress with that device
// current) and kept; its size in *bytes.  Returns 0, or the HIP error of
// the lookup.
static constexpr int NOALLOC_MAX_DEVICES = 64;
static char* noalloc_ptr[NOALLOC_MAX_DEVICES];
static std::mutex noalloc_mu;

int capsd_noalloc_scratch(char** ptr, size_t* bytes)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return (int)hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (!noalloc_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        // the 1024-lane panel's dynamic LDS, as capsd_lu_init sets it
        e = hipFuncSetAttribute((const void*)k_lu_panel<24, 2, 1024, 4>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                24 * 1024 * sizeof(int));
        if (e != hipSuccess) return (int)e;
        noalloc_ptr[dev] = (char*)p;
    }
    *ptr = noalloc_ptr[dev];
    *bytes = NOALLOC_SCRATCH_BYTES;
    return 0;
}

void capsd_launch_noalloc_setup(char* base, int64_t a_off, int dim, int lda, int n, int64_t off_side, int64_t off_out,
                                int64_t off_buf, int64_t per, int64_t o_logs, int64_t o_moved, int64_t o_ut,
                                int64_t o_inv, hipStream_t s)
{
    hipLaunchKernelGGL(k_lu_noalloc_setup, dim3((n + 63) / 64), dim3(64), 0, s, base, a_off, dim, lda, n, off_side,
                       off_out, off_buf, per, o_logs, o_moved, o_ut, o_inv);
}

void capsd_launch_noalloc_result(const void* out, double* logdet, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_lu_noalloc_result, dim3((n + 63) / 64), dim3(64), 0, s, (const BlockOut*)out, logdet, n);
}

void capsd_launch_lu_fill(const BlockDesc* desc, char* arena, int nblocks, unsigned long long seed, int delta,
                          hipStream_t s)
{
    hipLaunchKernelGGL(k_lu_fill, dim3(256, nblocks), dim3(256), 0, s, desc, arena, seed, delta);
}

void capsd_launch_mfma_probe(const double* A, const double* B, double* D, hipStream_t s)
{
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, s, A, B, D);
}

}  // extern "C"
