// log|det A| of a batch of FP64 matrices by blocked right-looking LU with
// partial pivoting (the dgetrf_ route of src/matrix.c:555-577 and the Kahan
// sum of log|u_ii| of src/matrix.c:462-471), on gfx950.
//
// Per panel of NB = 64 columns, three batched kernels:
//   k_lu_panel      one workgroup per matrix factors the N-k x 64 panel:
//                   left-looking over leaves of 8 columns, each leaf held in
//                   registers (R rows per lane), pivot search reduced over the
//                   wavefronts, rows swapped through a logical->physical row
//                   map in LDS (no row moves in HBM during the panel); writes
//                   the factored panel in pivoted order to a side buffer, the
//                   list of moved rows and log|u_ii|;
//   k_lu_swap_trsm  applies the row permutation to the trailing columns and
//                   solves U12 = L11^{-1} A12 (one lane per column);
//   k_lu_gemm       A22 -= L21 U12 on FP64 MFMA (v_mfma_f64_16x16x4f64).
// k_lu_finalize sums log|u_ii| with Kahan compensation.
// Matrices are column-major with leading dimension lda.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "blocks.h"

namespace capsd {

constexpr int NB = 64;       // panel width
constexpr int LEAF = 8;      // leaf width inside the panel
constexpr int PT = 256;      // threads of the panel kernel
constexpr int MAXMOVE = 2 * NB;

template <typename T>
__device__ __forceinline__ T* at(char* base, int64_t off) { return reinterpret_cast<T*>(base + off); }

// Side buffers of the LU of one block (inside the block's arena slice):
//   off_piv:              N ints (unused by the logdet; kept for the layout)
// and, appended by the host after the matrix:
struct LuSide {
    int64_t off_panel;   // lda x NB doubles: factored panel, pivoted order
    int64_t off_logs;    // N doubles: log|u_ii|
    int64_t off_moved;   // 1 + 2*MAXMOVE ints: count, (dst, src) pairs
};

__device__ __forceinline__ void wave_argmax(double& v, int& idx)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(v, o, 64);
        const int oi = __shfl_xor(idx, o, 64);
        if (ov > v || (ov == v && oi < idx)) {
            v = ov;
            idx = oi;
        }
    }
}

// Panel body: R rows per lane, leaves of LF columns, NT threads; the
// logical->physical row map lives in dynamic LDS (R*NT ints).
template <int R, int LF, int NT>
__device__ __forceinline__ void lu_panel_body(const BlockDesc* __restrict__ desc, char* __restrict__ arena,
                                              const LuSide* __restrict__ side, const BlockOut* __restrict__ out, int k)
{
    const int b = blockIdx.x;
    const BlockDesc& d = desc[b];
    if (out[b].skip || k >= d.N) return;
    const int N = d.N;
    const int64_t lda = d.lda;
    const int np = N - k;                 // panel rows
    const int nbk = min(NB, np);          // panel columns
    double* A = at<double>(arena, d.off_mat);
    double* P = at<double>(arena, side[b].off_panel);
    double* logs = at<double>(arena, side[b].off_logs);
    int* moved = at<int>(arena, side[b].off_moved);
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;

    extern __shared__ int perm[];   // R * NT
    __shared__ double red_v2[2][NT / 64];
    __shared__ int red_i2[2][NT / 64];
    __shared__ double prow2[2][LF], xrow2[2][LF];
    __shared__ double U[NB * LF];       // U part of the current leaf, [row][col]
    __shared__ double Lt[NB * NB];        // L11 of the panel so far, [row][col]
    __shared__ int nmoved;

    for (int r = tid; r < np; r += NT) perm[r] = r;
    if (tid == 0) nmoved = 0;
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
        if (c0 > 0) {
            // U part: solve L11[0:c0,0:c0] U = leaf[0:c0] (unit lower), one lane per column
            for (int idx = tid; idx < c0 * c0; idx += NT) {
                const int rr = idx / c0, cc = idx % c0;
                Lt[rr * NB + cc] = A[(k + perm[rr]) + (k + cc) * lda];
            }
            for (int idx = tid; idx < c0 * LF; idx += NT) {
                const int rr = idx / LF, cc = idx % LF;
                U[rr * LF + cc] = cc < w ? A[(k + perm[rr]) + (k + c0 + cc) * lda] : 0.0;
            }
            __syncthreads();
            if (tid < LF) {
                for (int i = 0; i < c0; i++) {
                    const double xi = U[i * LF + tid];
                    for (int jj = i + 1; jj < c0; jj++) U[jj * LF + tid] -= Lt[jj * NB + i] * xi;
                }
            }
            __syncthreads();
            // rows < c0 take the solved U; rows >= c0 get the left-looking
            // update, one earlier leaf (8 columns of L) per step with all of
            // the step's loads issued before their first use
#pragma unroll
            for (int i = 0; i < R; i++) {
                const int r = tid + i * NT;
                if (r < c0) {
#pragma unroll
                    for (int j = 0; j < LF; j++) a[i][j] = U[r * LF + j];
                }
            }
            for (int cc0 = 0; cc0 < c0; cc0 += LF) {
                double Lr[R][LF];
#pragma unroll
                for (int i = 0; i < R; i++) {
                    const int r = tid + i * NT;
                    const bool row_ok = r >= c0 && r < np;
                    const int64_t pr = k + (row_ok ? perm[r] : 0);
#pragma unroll
                    for (int jj = 0; jj < LF; jj++) Lr[i][jj] = row_ok ? A[pr + (k + cc0 + jj) * lda] : 0.0;
                }
#pragma unroll
                for (int jj = 0; jj < LF; jj++) {
                    double u[LF];
#pragma unroll
                    for (int j = 0; j < LF; j++) u[j] = U[(cc0 + jj) * LF + j];
#pragma unroll
                    for (int i = 0; i < R; i++)
#pragma unroll
                        for (int j = 0; j < LF; j++) a[i][j] -= Lr[i][jj] * u[j];
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
                logs[k + c] = log(fabs(gv));
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
#pragma unroll
            for (int i = 0; i < R; i++) {
                const int r = tid + i * NT;
                if (r > c && r < np) {
                    const double l = a[i][j] / u;
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
    // panel in pivoted (logical) order, and the list of moved rows
    for (int r = tid; r < np; r += NT) {
        const int64_t pr = k + perm[r];
        for (int j = 0; j < nbk; j++) P[r + j * lda] = A[pr + (k + j) * lda];
        if (perm[r] != r) {
            const int slot = atomicAdd(&nmoved, 1);
            if (slot < MAXMOVE) {
                moved[1 + 2 * slot] = r;
                moved[2 + 2 * slot] = perm[r];
            }
        }
    }
    __syncthreads();
    if (tid == 0) moved[0] = min(nmoved, MAXMOVE);
}


// Up to 12 rows per lane: leaves of 8 columns; 16 or 24 rows: leaves of 4
// (register budget of one wave per SIMD).
template <int R>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
k_lu_panel(const BlockDesc* __restrict__ desc, char* __restrict__ arena, const LuSide* __restrict__ side,
           const BlockOut* __restrict__ out, int k)
{
    lu_panel_body<R, (R <= 12 ? 8 : 4), 256>(desc, arena, side, out, k);
}

// Large matrices (N up to 24 x 1024 rows): 1024 lanes, leaves of 2 columns.
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(4, 4)))
k_lu_panel_big(const BlockDesc* __restrict__ desc, char* __restrict__ arena, const LuSide* __restrict__ side,
               const BlockOut* __restrict__ out, int k)
{
    lu_panel_body<24, 2, 1024>(desc, arena, side, out, k);
}

// Row permutation of the trailing columns and U12 = L11^{-1} A12.  One lane
// per trailing column, 64 columns per workgroup.  The moved rows of all 64
// columns are gathered into LDS before any is written (dynamic LDS:
// MAXMOVE x 64 doubles).
__global__ void __launch_bounds__(64) k_lu_swap_trsm(const BlockDesc* __restrict__ desc, char* __restrict__ arena,
                                                     const LuSide* __restrict__ side,
                                                     const BlockOut* __restrict__ out, int k)
{
    const int b = blockIdx.y;
    const BlockDesc& d = desc[b];
    if (out[b].skip || k + NB >= d.N) return;
    const int N = d.N;
    const int64_t lda = d.lda;
    if (blockIdx.x * 64 + k + NB >= N) return;
    const int col = k + NB + blockIdx.x * 64 + threadIdx.x;
    double* A = at<double>(arena, d.off_mat);
    const double* P = at<double>(arena, side[b].off_panel);
    const int* moved = at<int>(arena, side[b].off_moved);
    extern __shared__ double tmpS[];   // [MAXMOVE][64]
    if (col >= N) return;
    double* Ac = A + (int64_t)col * lda;
    const int nm = moved[0];
    for (int t = 0; t < nm; t++) tmpS[t * 64 + threadIdx.x] = Ac[k + moved[2 + 2 * t]];
    for (int t = 0; t < nm; t++) Ac[k + moved[1 + 2 * t]] = tmpS[t * 64 + threadIdx.x];
    // forward substitution with the unit lower L11 (uniform loads of P)
    double x[NB];
#pragma unroll
    for (int i = 0; i < NB; i++) x[i] = Ac[k + i];
#pragma unroll
    for (int i = 0; i < NB; i++) {
#pragma unroll
        for (int j = i + 1; j < NB; j++) x[j] -= P[j + (int64_t)i * lda] * x[i];
    }
#pragma unroll
    for (int i = 0; i < NB; i++) Ac[k + i] = x[i];
}

// A22 -= L21 U12 with FP64 MFMA.  Workgroup tile 128x128 of A22, 4 waves
// of 64x64 (4x4 tiles of 16x16 per wave, 16 accumulators), K = NB = 64
// staged through LDS 16 at a time.  LDS rows are padded to 144 doubles so
// the four k-rows a wavefront reads in one instruction fall on disjoint bank
// halves.
typedef double double4_t __attribute__((ext_vector_type(4)));
constexpr int GT = 128;          // workgroup tile
constexpr int KS = 16;           // K per LDS stage
constexpr int GTP = GT + 16;     // padded LDS row

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_lu_gemm(const BlockDesc* __restrict__ desc, char* __restrict__ arena,
                                                 const LuSide* __restrict__ side, const BlockOut* __restrict__ out,
                                                 int k)
{
    const int b = blockIdx.z;
    const BlockDesc& d = desc[b];
    if (out[b].skip) return;
    const int N = d.N;
    const int m0 = k + NB + blockIdx.x * GT;   // rows of A22
    const int n0 = k + NB + blockIdx.y * GT;   // columns of A22
    if (m0 >= N || n0 >= N) return;
    const int64_t lda = d.lda;
    double* A = at<double>(arena, d.off_mat);
    const double* P = at<double>(arena, side[b].off_panel);
    // one LDS buffer: the A and B stages, reused for the C strips at the end
    __shared__ double smem[2 * KS * GTP];
    double(*As)[GTP] = reinterpret_cast<double(*)[GTP]>(smem);            // [kk][row]
    double(*Bs)[GTP] = reinterpret_cast<double(*)[GTP]>(smem + KS * GTP);  // [kk][col]
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int wr = (wv & 1) * 64, wc = (wv >> 1) * 64;
    double4_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = (double4_t){0, 0, 0, 0};
    const int li = lane & 15, lk = lane >> 4;
    // the next stage's operands are fetched into registers while the
    // current stage runs on the MFMA units
    constexpr int PER = (KS * GT) / 256;
    double pa[PER], pb[PER];
    auto fetch = [&](int ks) {
#pragma unroll
        for (int t = 0; t < PER; t++) {
            const int idx = tid + 256 * t;
            const int kk = idx / GT, r = idx % GT;
            const int row = m0 + r;
            pa[t] = row < N ? P[(row - k) + (int64_t)(ks + kk) * lda] : 0.0;
            const int kk2 = idx % KS, c = idx / KS;
            const int colg = n0 + c;
            pb[t] = colg < N ? A[(k + ks + kk2) + (int64_t)colg * lda] : 0.0;
        }
    };
    fetch(0);
    for (int ks = 0; ks < NB; ks += KS) {
        __syncthreads();
#pragma unroll
        for (int t = 0; t < PER; t++) {
            const int idx = tid + 256 * t;
            As[idx / GT][idx % GT] = pa[t];
            Bs[idx % KS][idx / KS] = pb[t];
        }
        __syncthreads();
        if (ks + KS < NB) fetch(ks + KS);
#pragma unroll
        for (int kk = 0; kk < KS; kk += 4) {
            double av[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; i++) av[i] = As[kk + lk][wr + 16 * i + li];
#pragma unroll
            for (int j = 0; j < 4; j++) bv[j] = Bs[kk + lk][wc + 16 * j + li];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[i][j], 0, 0, 0);
        }
    }
    // D layout (measured with k_mfma_probe): lane l, register q holds row
    // l/16 + 4q, column l%16 of the 16x16 tile.  Each wave passes its
    // 64x64 result through LDS one 16-row strip at a time so that the
    // read-modify-write of C touches whole 128-byte column segments: lane
    // (r = l%16, c = l/16) updates rows r of columns c, c+4, ..., c+60, all
    // loads of a strip issued before its stores.
    double* Ts = smem + wv * (16 * 64 + 4);        // per-wave strip [col][row], 16 x 64
    __syncthreads();                                // As/Bs are reused
    const int rr = lane & 15, cc = lane >> 4;
#pragma unroll
    for (int i = 0; i < 4; i++) {
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) Ts[(16 * j + li) * 16 + (lk + 4 * q)] = acc[i][j][q];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int row = m0 + wr + 16 * i + rr;
        double cv[16];
#pragma unroll
        for (int t = 0; t < 16; t++) {
            const int col = n0 + wc + cc + 4 * t;
            cv[t] = (row < N && col < N) ? A[row + (int64_t)col * lda] : 0.0;
        }
#pragma unroll
        for (int t = 0; t < 16; t++) {
            const int col = n0 + wc + cc + 4 * t;
            if (row < N && col < N) A[row + (int64_t)col * lda] = cv[t] - Ts[(cc + 4 * t) * 16 + rr];
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// Kahan sum of log|u_ii| (src/matrix.c:462-471).
__global__ void k_lu_finalize(const BlockDesc* __restrict__ desc, char* __restrict__ arena,
                              const LuSide* __restrict__ side, BlockOut* __restrict__ out, int nblocks)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblocks) return;
    if (out[b].skip) return;
    const BlockDesc& d = desc[b];
    const double* logs = at<double>(arena, side[b].off_logs);
    double s = 0, c = 0;
    for (int i = 0; i < d.N; i++) {
        const double y = logs[i] - c;
        const double t = s + y;
        c = (t - s) - y;
        s = t;
    }
    out[b].logdet = s;
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

// Rows per lane the panel kernel needs for a matrix of dimension N.
int capsd_lu_rows_per_lane(int N) { return (N + PT - 1) / PT; }
int capsd_lu_max_dim(void) { return 24 * 1024; }
int capsd_lu_nb(void) { return NB; }
int capsd_lu_maxmove(void) { return MAXMOVE; }

static void launch_panel(int R, const BlockDesc* desc, char* arena, const LuSide* side, const BlockOut* out, int nb,
                         int k, hipStream_t s)
{
#define PANEL(RR) \
    hipLaunchKernelGGL(k_lu_panel<RR>, dim3(nb), dim3(256), (RR) * 256 * sizeof(int), s, desc, arena, side, out, k)
    if (R <= 2) PANEL(2);
    else if (R <= 4) PANEL(4);
    else if (R <= 8) PANEL(8);
    else if (R <= 12) PANEL(12);
    else if (R <= 16) PANEL(16);
    else if (R <= 24) PANEL(24);
    else hipLaunchKernelGGL(k_lu_panel_big, dim3(nb), dim3(1024), 24 * 1024 * sizeof(int), s, desc, arena, side, out, k);
#undef PANEL
}

// Whole LU of a batch (host loop over panels; all launches on stream s).
int capsd_lu_init(void)
{
    hipError_t e = hipFuncSetAttribute((const void*)k_lu_swap_trsm, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       MAXMOVE * 64 * sizeof(double));
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)k_lu_panel_big, hipFuncAttributeMaxDynamicSharedMemorySize,
                                24 * 1024 * sizeof(int));
    return (int)e;
}

// gemm_ev: NULL, or 2 events per panel recorded around each k_lu_gemm
// launch (per-kernel timing of the trailing update).
void capsd_launch_lu(const BlockDesc* desc, char* arena, const void* side, BlockOut* out, int nblocks, int maxN,
                     hipStream_t s, hipEvent_t* gemm_ev)
{
    const LuSide* sd = (const LuSide*)side;
    int p = 0;
    for (int k = 0; k < maxN; k += NB, p++) {
        const int R = (maxN - k + PT - 1) / PT;
        launch_panel(R, desc, arena, sd, out, nblocks, k, s);
        const int rest = maxN - k - NB;
        if (rest > 0) {
            hipLaunchKernelGGL(k_lu_swap_trsm, dim3((rest + 63) / 64, nblocks), dim3(64),
                               MAXMOVE * 64 * sizeof(double), s, desc, arena, sd, out, k);
            if (gemm_ev) (void)hipEventRecord(gemm_ev[2 * p], s);
            hipLaunchKernelGGL(k_lu_gemm, dim3((rest + GT - 1) / GT, (rest + GT - 1) / GT, nblocks), dim3(256), 0, s,
                               desc, arena, sd, out, k);
            if (gemm_ev) (void)hipEventRecord(gemm_ev[2 * p + 1], s);
        }
    }
    hipLaunchKernelGGL(k_lu_finalize, dim3((nblocks + 63) / 64), dim3(64), 0, s, desc, arena, sd, out, nblocks);
}

void capsd_launch_mfma_probe(const double* A, const double* B, double* D, hipStream_t s)
{
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, s, A, B, D);
}

}  // extern "C"
