// Surrogate for the planner: an MLP 5 -> 64 -> 64 -> 1 (tanh, FP32) that maps
// (log R/L, log ξ(L+R)/c, m/ldim, log ωp, log γ) to log|logdetD|, trained
// on blocks the GPU path computes, with hand-written forward, backward and
// Adam kernels.  It has no counterpart in the reference; its only use is to
// choose, before any block is computed, which (ξ, m) blocks to batch.  The
// free energy never depends on it: every planned term is verified against
// the reference's cutoff rules and extended when the plan falls short.
//
// Parameter layout (floats): W1[64][5], b1[64], W2[64][64], b2[64], W3[64],
// b3.  Inputs and target are standardised on the host.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

namespace capsm {

constexpr int NIN = 5, H = 64;
constexpr int OW1 = 0, OB1 = OW1 + H * NIN, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H, OB3 = OW3 + H;
constexpr int NPAR = OB3 + 1;   // 4609
constexpr int TILE = 64;        // samples per workgroup in the gradient kernel

// Forward pass, one lane per sample; the parameters are staged in LDS.
__global__ void __launch_bounds__(256) k_mlp_fwd(const float* __restrict__ P, const float* __restrict__ X,
                                                 float* __restrict__ Y, int n)
{
    __shared__ float w[NPAR];
    for (int i = threadIdx.x; i < NPAR; i += blockDim.x) w[i] = P[i];
    __syncthreads();
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float x[NIN];
#pragma unroll
    for (int i = 0; i < NIN; i++) x[i] = X[s * NIN + i];
    float h1[H];
#pragma unroll
    for (int j = 0; j < H; j++) {
        float a = w[OB1 + j];
#pragma unroll
        for (int i = 0; i < NIN; i++) a += w[OW1 + j * NIN + i] * x[i];
        h1[j] = tanhf(a);
    }
    float y = w[OB3];
    for (int k = 0; k < H; k++) {
        float a = w[OB2 + k];
#pragma unroll
        for (int j = 0; j < H; j++) a += w[OW2 + k * H + j] * h1[j];
        y += w[OW3 + k] * tanhf(a);
    }
    Y[s] = y;
}

// Gradient of the mean squared error (1/n) Σ (y - t)² over samples
// [s0, s0 + n) of the (index-permuted) data set; one workgroup of 64 lanes
// per tile of 64 samples writes its partial gradient G[tile][NPAR].
// Dynamic LDS: parameters + per-sample H1, H2 (then dH2), dH1, x, dy.
__global__ void __launch_bounds__(TILE) k_mlp_grad(const float* __restrict__ P, const float* __restrict__ X,
                                                   const float* __restrict__ T, const int* __restrict__ idx, int s0,
                                                   int n, float inv_n, float* __restrict__ G)
{
    extern __shared__ float sh[];
    float* w = sh;                       // NPAR
    float* H1 = w + NPAR + 3;            // [TILE][H]
    float* H2 = H1 + TILE * H;           // [TILE][H], later dH2
    float* D1 = H2 + TILE * H;           // [TILE][H]
    float* XS = D1 + TILE * H;           // [TILE][NIN]
    float* DY = XS + TILE * NIN;         // [TILE]
    const int t = threadIdx.x;
    for (int i = t; i < NPAR; i += TILE) w[i] = P[i];
    __syncthreads();
    const int si = blockIdx.x * TILE + t;     // sample within the batch
    const bool ok = si < n;
    const int s = ok ? (idx ? idx[s0 + si] : s0 + si) : 0;
    // forward
#pragma unroll
    for (int i = 0; i < NIN; i++) XS[t * NIN + i] = ok ? X[s * NIN + i] : 0.f;
    for (int j = 0; j < H; j++) {
        float a = w[OB1 + j];
#pragma unroll
        for (int i = 0; i < NIN; i++) a += w[OW1 + j * NIN + i] * XS[t * NIN + i];
        H1[t * H + j] = tanhf(a);
    }
    float y = w[OB3];
    for (int k = 0; k < H; k++) {
        float a = w[OB2 + k];
        for (int j = 0; j < H; j++) a += w[OW2 + k * H + j] * H1[t * H + j];
        const float h2 = tanhf(a);
        H2[t * H + k] = h2;
        y += w[OW3 + k] * h2;
    }
    const float dy = ok ? 2.f * (y - T[s]) * inv_n : 0.f;
    DY[t] = dy;
    __syncthreads();
    float* g = G + (int64_t)blockIdx.x * NPAR;
    // dW3, b3 (lane k)
    {
        float a = 0;
        for (int q = 0; q < TILE; q++) a += DY[q] * H2[q * H + t];
        g[OW3 + t] = a;
        if (t == 0) {
            float b = 0;
            for (int q = 0; q < TILE; q++) b += DY[q];
            g[OB3] = b;
        }
    }
    __syncthreads();
    // dH2 in place (lane = sample), then dH1
    for (int k = 0; k < H; k++) {
        const float h2 = H2[t * H + k];
        H2[t * H + k] = dy * w[OW3 + k] * (1.f - h2 * h2);
    }
    for (int j = 0; j < H; j++) {
        float a = 0;
        for (int k = 0; k < H; k++) a += w[OW2 + k * H + j] * H2[t * H + k];
        const float h1 = H1[t * H + j];
        D1[t * H + j] = a * (1.f - h1 * h1);
    }
    __syncthreads();
    // dW2[k][j] = Σ_q dH2[q][k] H1[q][j]; lane t covers rows k = t
    for (int j = 0; j < H; j++) {
        float a = 0;
        for (int q = 0; q < TILE; q++) a += H2[q * H + t] * H1[q * H + j];
        g[OW2 + t * H + j] = a;
    }
    {
        float b2 = 0, b1 = 0;
        for (int q = 0; q < TILE; q++) {
            b2 += H2[q * H + t];
            b1 += D1[q * H + t];
        }
        g[OB2 + t] = b2;
        g[OB1 + t] = b1;
    }
#pragma unroll
    for (int i = 0; i < NIN; i++) {
        float a = 0;
        for (int q = 0; q < TILE; q++) a += D1[q * H + t] * XS[q * NIN + i];
        g[OW1 + t * NIN + i] = a;
    }
}

// Sum of the tiles' partial gradients, in tile order (deterministic).
__global__ void k_mlp_reduce(const float* __restrict__ G, int ntiles, float* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPAR) return;
    float a = 0;
    for (int b = 0; b < ntiles; b++) a += G[(int64_t)b * NPAR + i];
    out[i] = a;
}

// Adam (Kingma & Ba 2015) with bias correction.
__global__ void k_adam(float* __restrict__ P, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPAR) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1 - b1) * gi;
    const float vi = b2 * v[i] + (1 - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    P[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Mean squared error of the current parameters over the first n samples.
__global__ void k_mse(const float* __restrict__ Y, const float* __restrict__ T, int n, float* __restrict__ out)
{
    __shared__ float sh[256];
    float a = 0;
    for (int i = threadIdx.x; i < n; i += 256) {
        const float d = Y[i] - T[i];
        a += d * d;
    }
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0] / n;
}

constexpr size_t GRAD_LDS = (NPAR + 3 + 3 * TILE * H + TILE * NIN + TILE) * sizeof(float);

}  // namespace capsm

using namespace capsm;

#define MCHECK(x)                                                                                 \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) {                                                                   \
            fprintf(stderr, "Fatal error: HIP error %s at %s (%s:%d)\n", hipGetErrorString(e_), #x, \
                    __FILE__, __LINE__);                                                          \
            abort();                                                                              \
        }                                                                                         \
    } while (0)

struct capsx_mlp {
    float* P = nullptr;   // device parameters
    float* M = nullptr;   // Adam moments
    float* V = nullptr;
    long step = 0;
    float xmean[NIN], xstd[NIN], ymean = 0, ystd = 1;
    hipStream_t s = nullptr;
};

namespace {

uint64_t splitmix(uint64_t& st)
{
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double urand(uint64_t& st) { return (splitmix(st) >> 11) * (1.0 / 9007199254740992.0); }

void normalise(const capsx_mlp* h, const float* X, int n, std::vector<float>& out)
{
    out.resize((size_t)n * NIN);
    for (int s = 0; s < n; s++)
        for (int i = 0; i < NIN; i++) out[s * NIN + i] = (X[s * NIN + i] - h->xmean[i]) / h->xstd[i];
}

}  // namespace

extern "C" {

int capsx_mlp_nparams(void) { return NPAR; }

// New network, Glorot-uniform weights from SplitMix64(seed), zero biases.
capsx_mlp* capsx_mlp_new(unsigned long long seed)
{
    capsx_mlp* h = new capsx_mlp();
    std::vector<float> p(NPAR, 0.f);
    uint64_t st = seed;
    auto fill = [&](int off, int fan_out, int fan_in) {
        const double a = sqrt(6.0 / (fan_in + fan_out));
        for (int i = 0; i < fan_out * fan_in; i++) p[off + i] = (float)((2 * urand(st) - 1) * a);
    };
    fill(OW1, H, NIN);
    fill(OW2, H, H);
    fill(OW3, 1, H);
    for (int i = 0; i < NIN; i++) {
        h->xmean[i] = 0;
        h->xstd[i] = 1;
    }
    MCHECK(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking));
    MCHECK(hipMalloc(&h->P, NPAR * sizeof(float)));
    MCHECK(hipMalloc(&h->M, NPAR * sizeof(float)));
    MCHECK(hipMalloc(&h->V, NPAR * sizeof(float)));
    MCHECK(hipMemcpy(h->P, p.data(), NPAR * sizeof(float), hipMemcpyHostToDevice));
    MCHECK(hipMemset(h->M, 0, NPAR * sizeof(float)));
    MCHECK(hipMemset(h->V, 0, NPAR * sizeof(float)));
    MCHECK(hipFuncSetAttribute((const void*)k_mlp_grad, hipFuncAttributeMaxDynamicSharedMemorySize, GRAD_LDS));
    return h;
}

void capsx_mlp_free(capsx_mlp* h)
{
    if (!h) return;
    (void)hipStreamSynchronize(h->s);
    (void)hipFree(h->P);
    (void)hipFree(h->M);
    (void)hipFree(h->V);
    (void)hipStreamDestroy(h->s);
    delete h;
}

void capsx_mlp_get_params(capsx_mlp* h, float* out)
{
    MCHECK(hipMemcpy(out, h->P, NPAR * sizeof(float), hipMemcpyDeviceToHost));
}

void capsx_mlp_set_params(capsx_mlp* h, const float* in)
{
    MCHECK(hipMemcpy(h->P, in, NPAR * sizeof(float), hipMemcpyHostToDevice));
}

// Standardisation: x' = (x - mean)/std per feature, y' = (y - ymean)/ystd.
void capsx_mlp_set_norm(capsx_mlp* h, const float* xmean, const float* xstd, float ymean, float ystd)
{
    for (int i = 0; i < NIN; i++) {
        h->xmean[i] = xmean[i];
        h->xstd[i] = xstd[i] > 0 ? xstd[i] : 1;
    }
    h->ymean = ymean;
    h->ystd = ystd > 0 ? ystd : 1;
}

void capsx_mlp_get_norm(capsx_mlp* h, float* xmean, float* xstd, float* ymean, float* ystd)
{
    for (int i = 0; i < NIN; i++) {
        xmean[i] = h->xmean[i];
        xstd[i] = h->xstd[i];
    }
    *ymean = h->ymean;
    *ystd = h->ystd;
}

// Predictions in raw units (log|logdetD|) for raw features X[n][5].
void capsx_mlp_predict(capsx_mlp* h, const float* X, int n, float* Y)
{
    if (n <= 0) return;
    std::vector<float> xn;
    normalise(h, X, n, xn);
    float *dX, *dY;
    MCHECK(hipMalloc(&dX, (size_t)n * NIN * sizeof(float)));
    MCHECK(hipMalloc(&dY, (size_t)n * sizeof(float)));
    MCHECK(hipMemcpyAsync(dX, xn.data(), (size_t)n * NIN * sizeof(float), hipMemcpyHostToDevice, h->s));
    hipLaunchKernelGGL(k_mlp_fwd, dim3((n + 255) / 256), dim3(256), 0, h->s, h->P, dX, dY, n);
    MCHECK(hipMemcpyAsync(Y, dY, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, h->s));
    MCHECK(hipStreamSynchronize(h->s));
    MCHECK(hipFree(dX));
    MCHECK(hipFree(dY));
    for (int i = 0; i < n; i++) Y[i] = Y[i] * h->ystd + h->ymean;
}

// Gradient of the standardised MSE over all n samples (test hook for the
// CPU check of the backward kernel).  Xn, Tn already standardised.
void capsx_mlp_grad(capsx_mlp* h, const float* Xn, const float* Tn, int n, float* grad)
{
    float *dX, *dT, *dG, *dg;
    const int ntiles = (n + TILE - 1) / TILE;
    MCHECK(hipMalloc(&dX, (size_t)n * NIN * sizeof(float)));
    MCHECK(hipMalloc(&dT, (size_t)n * sizeof(float)));
    MCHECK(hipMalloc(&dG, (size_t)ntiles * NPAR * sizeof(float)));
    MCHECK(hipMalloc(&dg, NPAR * sizeof(float)));
    MCHECK(hipMemcpy(dX, Xn, (size_t)n * NIN * sizeof(float), hipMemcpyHostToDevice));
    MCHECK(hipMemcpy(dT, Tn, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mlp_grad, dim3(ntiles), dim3(TILE), GRAD_LDS, h->s, h->P, dX, dT, (const int*)nullptr, 0, n,
                       1.0f / n, dG);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((NPAR + 255) / 256), dim3(256), 0, h->s, dG, ntiles, dg);
    MCHECK(hipStreamSynchronize(h->s));
    MCHECK(hipMemcpy(grad, dg, NPAR * sizeof(float), hipMemcpyDeviceToHost));
    MCHECK(hipFree(dX));
    MCHECK(hipFree(dT));
    MCHECK(hipFree(dG));
    MCHECK(hipFree(dg));
}

// Train on raw (X, Y) with Adam, minibatches of `batch` samples, the
// sample order shuffled each epoch (SplitMix64(seed)), the step size
// cosine-annealed from lr in the first epoch to lr/20 in the last.  Sets the
// standardisation from the data.  Returns the final standardised MSE over
// the whole set; loss_curve (may be NULL) receives one value per epoch.
float capsx_mlp_train(capsx_mlp* h, const float* X, const float* Y, int n, int epochs, int batch, float lr,
                      unsigned long long seed, float* loss_curve)
{
    if (n <= 0) return NAN;
    double mu[NIN] = {0}, sd[NIN] = {0}, ym = 0, ys = 0;
    for (int s = 0; s < n; s++) {
        for (int i = 0; i < NIN; i++) mu[i] += X[s * NIN + i];
        ym += Y[s];
    }
    for (int i = 0; i < NIN; i++) mu[i] /= n;
    ym /= n;
    for (int s = 0; s < n; s++) {
        for (int i = 0; i < NIN; i++) sd[i] += (X[s * NIN + i] - mu[i]) * (X[s * NIN + i] - mu[i]);
        ys += (Y[s] - ym) * (Y[s] - ym);
    }
    float xm[NIN], xs[NIN];
    for (int i = 0; i < NIN; i++) {
        xm[i] = (float)mu[i];
        xs[i] = (float)sqrt(sd[i] / n);
    }
    capsx_mlp_set_norm(h, xm, xs, (float)ym, (float)sqrt(ys / n));
    std::vector<float> xn, tn(n);
    normalise(h, X, n, xn);
    for (int s = 0; s < n; s++) tn[s] = (Y[s] - h->ymean) / h->ystd;
    batch = std::max(TILE, std::min(batch, n));
    const int ntiles = (batch + TILE - 1) / TILE;
    float *dX, *dT, *dG, *dg, *dY, *dl;
    int* dI;
    MCHECK(hipMalloc(&dX, (size_t)n * NIN * sizeof(float)));
    MCHECK(hipMalloc(&dT, (size_t)n * sizeof(float)));
    MCHECK(hipMalloc(&dY, (size_t)n * sizeof(float)));
    MCHECK(hipMalloc(&dI, (size_t)n * sizeof(int)));
    MCHECK(hipMalloc(&dG, (size_t)ntiles * NPAR * sizeof(float)));
    MCHECK(hipMalloc(&dg, NPAR * sizeof(float)));
    MCHECK(hipMalloc(&dl, sizeof(float)));
    MCHECK(hipMemcpy(dX, xn.data(), (size_t)n * NIN * sizeof(float), hipMemcpyHostToDevice));
    MCHECK(hipMemcpy(dT, tn.data(), (size_t)n * sizeof(float), hipMemcpyHostToDevice));
    std::vector<int> perm(n);
    for (int i = 0; i < n; i++) perm[i] = i;
    uint64_t st = seed;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float loss = NAN;
    for (int ep = 0; ep < epochs; ep++) {
        for (int i = n - 1; i > 0; i--) {
            const int j = (int)(splitmix(st) % (uint64_t)(i + 1));
            std::swap(perm[i], perm[j]);
        }
        MCHECK(hipMemcpyAsync(dI, perm.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, h->s));
        const float lr_ep =
            lr / 20 + (lr - lr / 20) * 0.5f * (1 + cosf(3.14159265f * ep / std::max(1, epochs - 1)));
        for (int s0 = 0; s0 + batch <= n || s0 == 0; s0 += batch) {
            const int nb = std::min(batch, n - s0);
            const int nt = (nb + TILE - 1) / TILE;
            hipLaunchKernelGGL(k_mlp_grad, dim3(nt), dim3(TILE), GRAD_LDS, h->s, h->P, dX, dT, dI, s0, nb,
                               1.0f / nb, dG);
            hipLaunchKernelGGL(k_mlp_reduce, dim3((NPAR + 255) / 256), dim3(256), 0, h->s, dG, nt, dg);
            h->step++;
            const float c1 = 1 - powf(b1, (float)h->step), c2 = 1 - powf(b2, (float)h->step);
            hipLaunchKernelGGL(k_adam, dim3((NPAR + 255) / 256), dim3(256), 0, h->s, h->P, dg, h->M, h->V, lr_ep, b1,
                               b2, eps, c1, c2);
            if (s0 + batch >= n) break;
        }
        if (loss_curve || ep == epochs - 1) {
            hipLaunchKernelGGL(k_mlp_fwd, dim3((n + 255) / 256), dim3(256), 0, h->s, h->P, dX, dY, n);
            hipLaunchKernelGGL(k_mse, dim3(1), dim3(256), 0, h->s, dY, dT, n, dl);
            MCHECK(hipMemcpyAsync(&loss, dl, sizeof(float), hipMemcpyDeviceToHost, h->s));
            MCHECK(hipStreamSynchronize(h->s));
            if (loss_curve) loss_curve[ep] = loss;
        }
    }
    MCHECK(hipStreamSynchronize(h->s));
    MCHECK(hipFree(dX));
    MCHECK(hipFree(dT));
    MCHECK(hipFree(dY));
    MCHECK(hipFree(dI));
    MCHECK(hipFree(dG));
    MCHECK(hipFree(dg));
    MCHECK(hipFree(dl));
    return loss;
}

}  // extern "C"
