// k_mlp.hip -- the MLP that warm-starts the discrete-Maxwellian Newton solve,
// trained on the device by its own forward and backward kernels (no
// counterpart in the reference).
//
// Network: 5 -> 64 (tanh) -> 64 (tanh) -> 5, FP32 master parameters, Adam.
// Inputs and targets are normalised with statistics of the recorded rows.
// Loss over a batch of B rows with weights w (0 for rows that must not
// count): L = sum_b w_b |y_b - t_b|^2 / (5 B), normalised units.
//
// Training step (all on the context's stream, deterministic):
//   k_mlp_sample   row indices of the minibatch (splitmix64 of seed, step, b)
//   k_mlp_fwdbwd   one thread per sample: forward, loss, backward to the
//                  pre-activation gradients; activations to a work buffer
//   k_mlp_wgrad    parameter gradients, the batch split over kSplit blocks
//   k_mlp_gsum     sum of the kSplit partial gradients (fixed order)
//   k_mlp_adam     Adam update of the FP32 parameters
// After training the weights are rounded to BF16 for the warm start.
#include "ab_device.hpp"
#include "ab_dm.hpp"
#include "ab_internal.hpp"

#include <cmath>
#include <cstring>
#include <random>
#include <vector>

namespace ab {

namespace {

constexpr int kWs = 2 * kMlpHid + 2 * kMlpHid + kMlpIn + kMlpOut;   // work floats per sample
constexpr int kOffH1 = 0, kOffH2 = kMlpHid, kOffDz2 = 2 * kMlpHid, kOffDz1 = 3 * kMlpHid;
constexpr int kOffXn = 4 * kMlpHid, kOffDy = 4 * kMlpHid + kMlpIn;
constexpr int kSplit = 32;          // batch slices of the gradient reduction

__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__global__ void k_mlp_sample(int *idx, int B, unsigned long long seed, long step, int nrows)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const unsigned long long h = splitmix64(seed ^ splitmix64((unsigned long long)step * 1000003ull + b));
    idx[b] = (int)(h % (unsigned long long)nrows);
}

// Forward and backward of one sample per thread. idx == nullptr: row b.
__global__ __launch_bounds__(128) void k_mlp_fwdbwd(const float *__restrict__ P,
                                                    const float *__restrict__ rows,
                                                    const int *__restrict__ idx, int B,
                                                    float *__restrict__ work,
                                                    float *__restrict__ lossb)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const float *row = rows + (size_t)(idx ? idx[b] : b) * kMlpRow;
    float *wk = work + (size_t)b * kWs;
    float xn[kMlpIn], tn[kMlpOut];
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) xn[i] = (row[i] - P[kMlpXm + i]) / P[kMlpXs + i];
#pragma unroll
    for (int o = 0; o < kMlpOut; ++o) tn[o] = (row[kMlpIn + o] - P[kMlpYm + o]) / P[kMlpYs + o];
    const float w = row[kMlpIn + kMlpOut];
    float h1[kMlpHid];
#pragma unroll
    for (int j = 0; j < kMlpHid; ++j)
    {
        float z = P[kMlpB1 + j];
#pragma unroll
        for (int i = 0; i < kMlpIn; ++i) z += P[kMlpW1 + j * kMlpIn + i] * xn[i];
        h1[j] = tanhf(z);
        wk[kOffH1 + j] = h1[j];
    }
    float y[kMlpOut];
#pragma unroll
    for (int o = 0; o < kMlpOut; ++o) y[o] = P[kMlpB3 + o];
#pragma unroll 4
    for (int j = 0; j < kMlpHid; ++j)
    {
        float z = P[kMlpB2 + j];
#pragma unroll
        for (int k = 0; k < kMlpHid; ++k) z += P[kMlpW2 + j * kMlpHid + k] * h1[k];
        const float h2 = tanhf(z);
        wk[kOffH2 + j] = h2;
#pragma unroll
        for (int o = 0; o < kMlpOut; ++o) y[o] += P[kMlpW3 + o * kMlpHid + j] * h2;
    }
    const float scale = 1.0f / (float)(kMlpOut * B);
    float dy[kMlpOut], l = 0.0f;
#pragma unroll
    for (int o = 0; o < kMlpOut; ++o)
    {
        const float r = y[o] - tn[o];
        l += r * r;
        dy[o] = 2.0f * w * r * scale;
        wk[kOffDy + o] = dy[o];
    }
    lossb[b] = w * l * scale;
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) wk[kOffXn + i] = xn[i];
    float dh1[kMlpHid];
#pragma unroll
    for (int k = 0; k < kMlpHid; ++k) dh1[k] = 0.0f;
#pragma unroll 4
    for (int j = 0; j < kMlpHid; ++j)
    {
        const float h2 = wk[kOffH2 + j];
        float dh2 = 0.0f;
#pragma unroll
        for (int o = 0; o < kMlpOut; ++o) dh2 += P[kMlpW3 + o * kMlpHid + j] * dy[o];
        const float dz2 = dh2 * (1.0f - h2 * h2);
        wk[kOffDz2 + j] = dz2;
#pragma unroll
        for (int k = 0; k < kMlpHid; ++k) dh1[k] += P[kMlpW2 + j * kMlpHid + k] * dz2;
    }
#pragma unroll
    for (int k = 0; k < kMlpHid; ++k) wk[kOffDz1 + k] = dh1[k] * (1.0f - h1[k] * h1[k]);
}

// Partial gradient of parameter p over the batch slice blockIdx.y.
__global__ __launch_bounds__(256) void k_mlp_wgrad(const float *__restrict__ work, int B,
                                                   float *__restrict__ part)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpTrain) return;
    int a, bo;               // work offsets of the two factors; bo < 0: bias
    if (p < kMlpB1) { a = kOffDz1 + p / kMlpIn; bo = kOffXn + p % kMlpIn; }
    else if (p < kMlpW2) { a = kOffDz1 + (p - kMlpB1); bo = -1; }
    else if (p < kMlpB2) { const int q = p - kMlpW2; a = kOffDz2 + q / kMlpHid; bo = kOffH1 + q % kMlpHid; }
    else if (p < kMlpW3) { a = kOffDz2 + (p - kMlpB2); bo = -1; }
    else if (p < kMlpB3) { const int q = p - kMlpW3; a = kOffDy + q / kMlpHid; bo = kOffH2 + q % kMlpHid; }
    else { a = kOffDy + (p - kMlpB3); bo = -1; }
    const int per = (B + kSplit - 1) / kSplit;
    const int b0 = blockIdx.y * per, b1 = min(B, b0 + per);
    float s = 0.0f;
    for (int b = b0; b < b1; ++b)
    {
        const float *wk = work + (size_t)b * kWs;
        s += (bo >= 0) ? wk[a] * wk[bo] : wk[a];
    }
    part[(size_t)blockIdx.y * kMlpTrain + p] = s;
}

__global__ void k_mlp_gsum(const float *__restrict__ part, float *__restrict__ grad)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpTrain) return;
    float s = 0.0f;
    for (int k = 0; k < kSplit; ++k) s += part[(size_t)k * kMlpTrain + p];
    grad[p] = s;
}

__global__ void k_mlp_adam(float *__restrict__ P, const float *__restrict__ g, float *m, float *v,
                           float lr, float c1, float c2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpTrain) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float mi = b1 * m[p] + (1.0f - b1) * g[p];
    const float vi = b2 * v[p] + (1.0f - b2) * g[p] * g[p];
    m[p] = mi;
    v[p] = vi;
    P[p] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Sum of n floats by one block, fixed order.
__global__ __launch_bounds__(256) void k_mlp_sumf(const float *__restrict__ a, int n, float *out)
{
    __shared__ double sh[256 / kWave];
    double v[1] = {0.0};
    for (int i = threadIdx.x; i < n; i += 256) v[0] += a[i];
    block_sum<1>(v, sh);
    if (threadIdx.x == 0) *out = (float)v[0];
}

// Input/target normalisation from the rows with weight, FP64 sums.
__global__ __launch_bounds__(256) void k_mlp_norm(const float *__restrict__ rows, int n, float *P)
{
    __shared__ double sh[21 * (256 / kWave)];
    double v[21];
#pragma unroll
    for (int q = 0; q < 21; ++q) v[q] = 0.0;
    for (int r = threadIdx.x; r < n; r += 256)
    {
        const float *row = rows + (size_t)r * kMlpRow;
        if (!(row[kMlpIn + kMlpOut] > 0.0f)) continue;
#pragma unroll
        for (int q = 0; q < 10; ++q) { const double x = row[q]; v[q] += x; v[10 + q] += x * x; }
        v[20] += 1.0;
    }
    block_sum<21>(v, sh);
    if (threadIdx.x < 10)
    {
        const int q = threadIdx.x;
        const double cnt = v[20] > 0.0 ? v[20] : 1.0;
        const double mean = v[q] / cnt;
        const double var = fmax(v[10 + q] / cnt - mean * mean, 0.0);
        double sd = sqrt(var);
        if (!(sd > 1e-12 * fabs(mean) + 1e-30)) sd = 1.0;     // constant feature
        const int base = q < kMlpIn ? kMlpXm + q : kMlpYm + (q - kMlpIn);
        const int sbase = q < kMlpIn ? kMlpXs + q : kMlpYs + (q - kMlpIn);
        P[base] = (float)mean;
        P[sbase] = (float)sd;
    }
}

__global__ void k_mlp_bf16(const float *__restrict__ P, unsigned short *__restrict__ wb)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpTrain) return;
    const unsigned u = __float_as_uint(P[p]);
    const unsigned r = u + 0x7FFFu + ((u >> 16) & 1u);      // round to nearest even
    wb[p] = (unsigned short)(r >> 16);
}

// Forward with the FP32 parameters, de-normalised outputs; x rows of kMlpRow.
__global__ __launch_bounds__(128) void k_mlp_forward(const float *__restrict__ P,
                                                     const float *__restrict__ rows, int n,
                                                     float *__restrict__ y)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    const float *row = rows + (size_t)b * kMlpRow;
    float xn[kMlpIn];
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) xn[i] = (row[i] - P[kMlpXm + i]) / P[kMlpXs + i];
    float h1[kMlpHid];
#pragma unroll
    for (int j = 0; j < kMlpHid; ++j)
    {
        float z = P[kMlpB1 + j];
#pragma unroll
        for (int i = 0; i < kMlpIn; ++i) z += P[kMlpW1 + j * kMlpIn + i] * xn[i];
        h1[j] = tanhf(z);
    }
    float out[kMlpOut];
#pragma unroll
    for (int o = 0; o < kMlpOut; ++o) out[o] = P[kMlpB3 + o];
#pragma unroll 4
    for (int j = 0; j < kMlpHid; ++j)
    {
        float z = P[kMlpB2 + j];
#pragma unroll
        for (int k = 0; k < kMlpHid; ++k) z += P[kMlpW2 + j * kMlpHid + k] * h1[k];
        const float h2 = tanhf(z);
#pragma unroll
        for (int o = 0; o < kMlpOut; ++o) out[o] += P[kMlpW3 + o * kMlpHid + j] * h2;
    }
#pragma unroll
    for (int o = 0; o < kMlpOut; ++o) y[(size_t)b * kMlpOut + o] = out[o] * P[kMlpYs + o] + P[kMlpYm + o];
}

// The warm start's inference as the solver runs it: BF16 weights on MFMA,
// FP32 accumulation, one wave per 64 rows (rows past n compute on zeros and
// are not written).
__global__ __launch_bounds__(256) void k_mlp_infer(const unsigned short *__restrict__ wb,
                                                   const float *__restrict__ pf,
                                                   const float *__restrict__ x, int n,
                                                   float *__restrict__ y)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    float xi[kMlpIn], out[kMlpOut];
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) xi[i] = b < n ? x[(size_t)b * kMlpIn + i] : 0.0f;
    mlp_wave_mfma(wb, pf, xi, out);
    if (b < n)
#pragma unroll
        for (int o = 0; o < kMlpOut; ++o) y[(size_t)b * kMlpOut + o] = out[o];
}

#define MLPCHK(ctx, expr)                                                        \
    do {                                                                         \
        hipError_t e_ = (expr);                                                  \
        if (e_ != hipSuccess) {                                                  \
            (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);      \
            return AB_ERR_HIP;                                                   \
        }                                                                        \
    } while (0)

int mlp_alloc(ab_ctx *c)
{
    if (c->mlp_p) return AB_OK;
    bool ok = hipMalloc((void **)&c->mlp_p, sizeof(float) * kMlpParams) == hipSuccess &&
              hipMalloc((void **)&c->mlp_wb, sizeof(unsigned short) * kMlpTrain) == hipSuccess &&
              hipMalloc((void **)&c->mlp_m, sizeof(float) * kMlpTrain) == hipSuccess &&
              hipMalloc((void **)&c->mlp_v, sizeof(float) * kMlpTrain) == hipSuccess &&
              hipMalloc((void **)&c->mlp_g, sizeof(float) * kMlpTrain * (kSplit + 1)) == hipSuccess;
    if (!ok) { c->err = "MLP: device allocation failed"; return AB_ERR_NOMEM; }
    return AB_OK;
}

// Work buffers for a batch of B samples (activations, per-sample loss,
// indices).
int mlp_work(ab_ctx *c, int B)
{
    if (B <= c->mlp_work_cap) return AB_OK;
    if (c->mlp_work) (void)hipFree(c->mlp_work);
    if (c->mlp_idx) (void)hipFree(c->mlp_idx);
    c->mlp_work = nullptr; c->mlp_idx = nullptr; c->mlp_work_cap = 0;
    if (hipMalloc((void **)&c->mlp_work, sizeof(float) * ((size_t)B * kWs + B + 1)) != hipSuccess ||
        hipMalloc((void **)&c->mlp_idx, sizeof(int) * (size_t)B) != hipSuccess)
    { c->err = "MLP: work allocation failed"; return AB_ERR_NOMEM; }
    c->mlp_work_cap = B;
    return AB_OK;
}

int mlp_rows(ab_ctx *c, int n)
{
    if (n <= c->mlp_rows_cap) return AB_OK;
    if (c->mlp_rows) (void)hipFree(c->mlp_rows);
    c->mlp_rows = nullptr; c->mlp_rows_cap = 0;
    if (hipMalloc((void **)&c->mlp_rows, sizeof(float) * (size_t)n * (kMlpRow + kMlpOut)) != hipSuccess)
    { c->err = "MLP: row allocation failed"; return AB_ERR_NOMEM; }
    c->mlp_rows_cap = n;
    return AB_OK;
}

// One forward/backward over B samples (rows or rows[idx]) into mlp_g;
// the summed loss into *loss_dev when given.
void mlp_grad(ab_ctx *c, const float *rows, const int *idx, int B, float *loss_dev)
{
    hipStream_t s = c->stream;
    float *lossb = c->mlp_work + (size_t)B * kWs;
    k_mlp_fwdbwd<<<(B + 127) / 128, 128, 0, s>>>(c->mlp_p, rows, idx, B, c->mlp_work, lossb);
    dim3 gw((kMlpTrain + 255) / 256, kSplit);
    float *part = c->mlp_g + kMlpTrain;
    k_mlp_wgrad<<<gw, 256, 0, s>>>(c->mlp_work, B, part);
    k_mlp_gsum<<<(kMlpTrain + 255) / 256, 256, 0, s>>>(part, c->mlp_g);
    if (loss_dev) k_mlp_sumf<<<1, 256, 0, s>>>(lossb, B, loss_dev);
}

void mlp_to_bf16(ab_ctx *c)
{
    k_mlp_bf16<<<(kMlpTrain + 255) / 256, 256, 0, c->stream>>>(c->mlp_p, c->mlp_wb);
}

// Host-side Xavier-uniform initialisation, biases zero, identity
// normalisation.
std::vector<float> mlp_random(unsigned long long seed)
{
    std::vector<float> p(kMlpParams, 0.0f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int fan_out, int fan_in) {
        const double a = std::sqrt(6.0 / (fan_in + fan_out));
        std::uniform_real_distribution<double> u(-a, a);
        for (int k = 0; k < fan_out * fan_in; ++k) p[off + k] = (float)u(rng);
    };
    fill(kMlpW1, kMlpHid, kMlpIn);
    fill(kMlpW2, kMlpHid, kMlpHid);
    fill(kMlpW3, kMlpOut, kMlpHid);
    for (int i = 0; i < kMlpIn; ++i) p[kMlpXs + i] = 1.0f;
    for (int o = 0; o < kMlpOut; ++o) p[kMlpYs + o] = 1.0f;
    return p;
}

}  // namespace

void mlp_free(ab_ctx *c)
{
    void *ptrs[] = {c->mlp_p, c->mlp_wb, c->mlp_m, c->mlp_v, c->mlp_g, c->mlp_work, c->mlp_idx,
                    c->mlp_rows};
    for (void *p : ptrs) if (p) (void)hipFree(p);
    c->mlp_p = nullptr; c->mlp_wb = nullptr; c->mlp_m = c->mlp_v = c->mlp_g = nullptr;
    c->mlp_work = nullptr; c->mlp_idx = nullptr; c->mlp_rows = nullptr;
    c->mlp_work_cap = c->mlp_rows_cap = 0;
    c->mlp_ready = c->mlp_on = false;
}

}  // namespace ab

using namespace ab;

extern "C" {

int ab_mlp_init(ab_ctx *c, unsigned long long seed)
{
    if (!c) return AB_ERR_ARG;
    int rc;
    if ((rc = mlp_alloc(c))) return rc;
    std::vector<float> p = mlp_random(seed);
    hipStream_t s = c->stream;
    MLPCHK(c, hipMemcpyAsync(c->mlp_p, p.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice, s));
    std::vector<float> z(kMlpTrain, 0.0f);
    MLPCHK(c, hipMemcpyAsync(c->mlp_m, z.data(), sizeof(float) * kMlpTrain, hipMemcpyHostToDevice, s));
    MLPCHK(c, hipMemcpyAsync(c->mlp_v, z.data(), sizeof(float) * kMlpTrain, hipMemcpyHostToDevice, s));
    const int nrows = c->dm_rec_filled * c->N;
    if (nrows > 0) k_mlp_norm<<<1, 256, 0, s>>>(c->dm_rec, nrows, c->mlp_p);
    mlp_to_bf16(c);
    MLPCHK(c, hipStreamSynchronize(s));
    c->mlp_step = 0;
    c->mlp_ready = true;
    return AB_OK;
}

int ab_mlp_train(ab_ctx *c, int iters, int batch, float lr, unsigned long long seed,
                 float *loss_first, float *loss_last)
{
    if (!c || iters < 0 || batch <= 0 || !(lr > 0.0f)) return AB_ERR_ARG;
    if (!c->mlp_ready) { c->err = "ab_mlp_train before ab_mlp_init"; return AB_ERR_STATE; }
    const int nrows = c->dm_rec_filled * c->N;
    if (nrows <= 0) { c->err = "ab_mlp_train: no recorded rows (ab_dm_record)"; return AB_ERR_STATE; }
    int rc;
    if ((rc = mlp_work(c, batch))) return rc;
    hipStream_t s = c->stream;
    float *loss_dev = c->mlp_work + (size_t)batch * kWs + batch;
    float l0 = 0.0f, l1 = 0.0f;
    for (int it = 0; it < iters; ++it)
    {
        k_mlp_sample<<<(batch + 255) / 256, 256, 0, s>>>(c->mlp_idx, batch, seed, it, nrows);
        const bool want = (it == 0 || it == iters - 1);
        mlp_grad(c, c->dm_rec, c->mlp_idx, batch, want ? loss_dev : nullptr);
        if (want)
        {
            float l;
            MLPCHK(c, hipMemcpyAsync(&l, loss_dev, sizeof l, hipMemcpyDeviceToHost, s));
            MLPCHK(c, hipStreamSynchronize(s));
            if (it == 0) l0 = l;
            l1 = l;
        }
        ++c->mlp_step;
        const float c1 = 1.0f - std::pow(0.9f, (float)c->mlp_step);
        const float c2 = 1.0f - std::pow(0.999f, (float)c->mlp_step);
        k_mlp_adam<<<(kMlpTrain + 255) / 256, 256, 0, s>>>(c->mlp_p, c->mlp_g, c->mlp_m, c->mlp_v,
                                                           lr, c1, c2);
    }
    mlp_to_bf16(c);
    MLPCHK(c, hipGetLastError());
    MLPCHK(c, hipStreamSynchronize(s));
    if (loss_first) *loss_first = l0;
    if (loss_last) *loss_last = l1;
    return AB_OK;
}

int ab_mlp_enable(ab_ctx *c, int on)
{
    if (!c) return AB_ERR_ARG;
    if (on && !c->mlp_ready) { c->err = "ab_mlp_enable before ab_mlp_init"; return AB_ERR_STATE; }
    c->mlp_on = on != 0;
    return AB_OK;
}

int ab_mlp_get(ab_ctx *c, float *params)
{
    if (!c || !params) return AB_ERR_ARG;
    if (!c->mlp_ready) { c->err = "MLP not initialised"; return AB_ERR_STATE; }
    MLPCHK(c, hipMemcpyAsync(params, c->mlp_p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost,
                             c->stream));
    MLPCHK(c, hipStreamSynchronize(c->stream));
    return AB_OK;
}

int ab_mlp_set(ab_ctx *c, const float *params)
{
    if (!c || !params) return AB_ERR_ARG;
    int rc;
    if ((rc = mlp_alloc(c))) return rc;
    for (int q = 0; q < kMlpIn; ++q)
        if (!(params[kMlpXs + q] > 0.0f)) { c->err = "ab_mlp_set: input scale must be > 0"; return AB_ERR_ARG; }
    for (int q = 0; q < kMlpOut; ++q)
        if (!(params[kMlpYs + q] > 0.0f)) { c->err = "ab_mlp_set: target scale must be > 0"; return AB_ERR_ARG; }
    hipStream_t s = c->stream;
    MLPCHK(c, hipMemcpyAsync(c->mlp_p, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, s));
    std::vector<float> z(kMlpTrain, 0.0f);
    MLPCHK(c, hipMemcpyAsync(c->mlp_m, z.data(), sizeof(float) * kMlpTrain, hipMemcpyHostToDevice, s));
    MLPCHK(c, hipMemcpyAsync(c->mlp_v, z.data(), sizeof(float) * kMlpTrain, hipMemcpyHostToDevice, s));
    mlp_to_bf16(c);
    MLPCHK(c, hipStreamSynchronize(s));
    c->mlp_step = 0;
    c->mlp_ready = true;
    return AB_OK;
}

int ab_mlp_forward(ab_ctx *c, const float *x, int n, float *y)
{
    if (!c || !x || !y || n < 0) return AB_ERR_ARG;
    if (!c->mlp_ready) { c->err = "MLP not initialised"; return AB_ERR_STATE; }
    if (n == 0) return AB_OK;
    int rc;
    if ((rc = mlp_rows(c, n))) return rc;
    std::vector<float> rows((size_t)n * kMlpRow, 0.0f);
    for (int b = 0; b < n; ++b)
        for (int i = 0; i < kMlpIn; ++i) rows[(size_t)b * kMlpRow + i] = x[(size_t)b * kMlpIn + i];
    hipStream_t s = c->stream;
    float *ydev = c->mlp_rows + (size_t)n * kMlpRow;
    MLPCHK(c, hipMemcpyAsync(c->mlp_rows, rows.data(), sizeof(float) * rows.size(), hipMemcpyHostToDevice, s));
    k_mlp_forward<<<(n + 127) / 128, 128, 0, s>>>(c->mlp_p, c->mlp_rows, n, ydev);
    MLPCHK(c, hipMemcpyAsync(y, ydev, sizeof(float) * (size_t)n * kMlpOut, hipMemcpyDeviceToHost, s));
    MLPCHK(c, hipStreamSynchronize(s));
    return AB_OK;
}

int ab_mlp_infer(ab_ctx *c, const float *x, int n, float *y)
{
    if (!c || !x || !y || n < 0) return AB_ERR_ARG;
    if (!c->mlp_ready) { c->err = "MLP not initialised"; return AB_ERR_STATE; }
    if (n == 0) return AB_OK;
    int rc;
    if ((rc = mlp_rows(c, n))) return rc;
    hipStream_t s = c->stream;
    float *ydev = c->mlp_rows + (size_t)n * kMlpRow;
    MLPCHK(c, hipMemcpyAsync(c->mlp_rows, x, sizeof(float) * (size_t)n * kMlpIn, hipMemcpyHostToDevice, s));
    k_mlp_infer<<<(n + 255) / 256, 256, 0, s>>>(c->mlp_wb, c->mlp_p, c->mlp_rows, n, ydev);
    MLPCHK(c, hipMemcpyAsync(y, ydev, sizeof(float) * (size_t)n * kMlpOut, hipMemcpyDeviceToHost, s));
    MLPCHK(c, hipStreamSynchronize(s));
    return AB_OK;
}

int ab_mlp_loss_grad(ab_ctx *c, const float *x, const float *t, const float *w, int n,
                     float *loss, float *grad)
{
    if (!c || !x || !t || !w || n <= 0 || !loss || !grad) return AB_ERR_ARG;
    if (!c->mlp_ready) { c->err = "MLP not initialised"; return AB_ERR_STATE; }
    int rc;
    if ((rc = mlp_rows(c, n)) || (rc = mlp_work(c, n))) return rc;
    std::vector<float> rows((size_t)n * kMlpRow);
    for (int b = 0; b < n; ++b)
    {
        float *r = rows.data() + (size_t)b * kMlpRow;
        for (int i = 0; i < kMlpIn; ++i) r[i] = x[(size_t)b * kMlpIn + i];
        for (int o = 0; o < kMlpOut; ++o) r[kMlpIn + o] = t[(size_t)b * kMlpOut + o];
        r[kMlpIn + kMlpOut] = w[b];
    }
    hipStream_t s = c->stream;
    MLPCHK(c, hipMemcpyAsync(c->mlp_rows, rows.data(), sizeof(float) * rows.size(), hipMemcpyHostToDevice, s));
    float *loss_dev = c->mlp_work + (size_t)n * kWs + n;
    mlp_grad(c, c->mlp_rows, nullptr, n, loss_dev);
    MLPCHK(c, hipMemcpyAsync(grad, c->mlp_g, sizeof(float) * kMlpTrain, hipMemcpyDeviceToHost, s));
    MLPCHK(c, hipMemcpyAsync(loss, loss_dev, sizeof(float), hipMemcpyDeviceToHost, s));
    MLPCHK(c, hipStreamSynchronize(s));
    return AB_OK;
}

}  // extern "C"
