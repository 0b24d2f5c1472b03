// ab_dm.hpp -- discrete-Maxwellian collision target and its learned warm
// start (device helpers). Not in the reference: its collision relaxes to the
// continuous Maxwellian of (rho, u, T) (kCollisionNode3D,
// Solver/CudaBackend.cu:1181-1197), whose grid moments differ from those of g
// by the quadrature error (1e-5 .. 1e-4 at Nv = 10), so the collision does not
// conserve momentum and energy on the grid. The discrete target
//     M_d(v) = exp(a . psi(v)),   psi = (1, xi_x, xi_y, xi_z, |xi|^2),  xi = v / cs,
// has exactly the grid moments of g (discrete-velocity BGK, Mieussens 2000):
// Newton in FP64 on the five a from the continuous Maxwellian's parameters,
// or from those plus a small MLP's prediction of the difference.
#pragma once
#include "ab_device.hpp"

namespace ab {

constexpr int kMlpIn = 5;      // u_x/cs, u_y/cs, u_z/cs, R T / cs^2, dv / cs
constexpr int kMlpHid = 64;    // two hidden layers of one wave's width
constexpr int kMlpOut = 5;     // a_discrete - a_continuous
// FP32 parameter block: W1[64][5] b1[64] W2[64][64] b2[64] W3[5][64] b3[5],
// then the input and output normalisation xm[5] xs[5] ym[5] ys[5].
constexpr int kMlpW1 = 0;
constexpr int kMlpB1 = kMlpW1 + kMlpHid * kMlpIn;
constexpr int kMlpW2 = kMlpB1 + kMlpHid;
constexpr int kMlpB2 = kMlpW2 + kMlpHid * kMlpHid;
constexpr int kMlpW3 = kMlpB2 + kMlpHid;
constexpr int kMlpB3 = kMlpW3 + kMlpOut * kMlpHid;
constexpr int kMlpTrain = kMlpB3 + kMlpOut;              // trainable count (4869)
constexpr int kMlpXm = kMlpTrain;
constexpr int kMlpXs = kMlpXm + kMlpIn;
constexpr int kMlpYm = kMlpXs + kMlpIn;
constexpr int kMlpYs = kMlpYm + kMlpOut;
constexpr int kMlpParams = kMlpYs + kMlpOut;             // 4889
// Recorded training row: 5 features, 5 targets, weight (1 valid / 0 not).
constexpr int kMlpRow = kMlpIn + kMlpOut + 1;

// Continuous Maxwellian rho (2 pi R T)^-3/2 exp(-|v-u|^2 / 2RT) in the basis.
__device__ __forceinline__ void dm_guess(double rho, double ux, double uy, double uz, double T,
                                         double R, double cs, double a[5])
{
    const double RT = R * T;
    a[0] = log(rho) - 1.5 * log(2.0 * AB_PI * RT) - (ux * ux + uy * uy + uz * uz) / (2.0 * RT);
    a[1] = cs * ux / RT;
    a[2] = cs * uy / RT;
    a[3] = cs * uz / RT;
    a[4] = -cs * cs / (2.0 * RT);
}

// The 15 distinct entries of J = sum e psi psi^T dv^3 (upper triangle, row
// by row); entries 0..4 (row 0) are the moments of M_d.
__device__ __forceinline__ void dm_accum(double S[15], double e, double x, double y, double z)
{
    const double q = x * x + y * y + z * z;
    const double ex = e * x, ey = e * y, ez = e * z, eq = e * q;
    S[0] += e;  S[1] += ex; S[2] += ey; S[3] += ez; S[4] += eq;
    S[5] += ex * x; S[6] += ex * y; S[7] += ex * z; S[8] += ex * q;
    S[9] += ey * y; S[10] += ey * z; S[11] += ey * q;
    S[12] += ez * z; S[13] += ez * q;
    S[14] += eq * q;
}

__device__ __forceinline__ int dm_uidx(int p, int q)
{
    // packed upper-triangle index of (min, max)
    const int i = p < q ? p : q, j = p < q ? q : p;
    return i * 5 - (i * (i - 1)) / 2 + (j - i);
}

// Solve J d = F by Cholesky (J symmetric positive definite from the packed
// sums). Returns false when a pivot is not positive.
__device__ __forceinline__ bool dm_solve5(const double S[15], const double F[5], double d[5])
{
    double L[5][5];
    double y[5];
#pragma unroll
    for (int i = 0; i < 5; ++i)
    {
#pragma unroll
        for (int j = 0; j <= i; ++j)
        {
            double s = S[dm_uidx(i, j)];
#pragma unroll
            for (int k = 0; k < j; ++k) s -= L[i][k] * L[j][k];
            if (i == j)
            {
                if (!(s > 0.0)) return false;
                L[i][i] = sqrt(s);
            }
            else L[i][j] = s / L[j][j];
        }
    }
#pragma unroll
    for (int i = 0; i < 5; ++i)
    {
        double s = F[i];
#pragma unroll
        for (int k = 0; k < i; ++k) s -= L[i][k] * y[k];
        y[i] = s / L[i][i];
    }
#pragma unroll
    for (int i = 4; i >= 0; --i)
    {
        double s = y[i];
#pragma unroll
        for (int k = i + 1; k < 5; ++k) s -= L[k][i] * d[k];
        d[i] = s / L[i][i];
    }
    return true;
}

__device__ __forceinline__ float bf16_to_float(unsigned short h)
{
    return __uint_as_float((unsigned)h << 16);
}

// MLP forward for one sample by one wave (lane j = hidden unit j), BF16
// weights (W1, W2, W3 packed in that order), FP32 biases, normalisation and
// accumulation. Every lane returns the five outputs (de-normalised).
__device__ __forceinline__ void mlp_wave_forward(const unsigned short *__restrict__ wb,
                                                 const float *__restrict__ pf,
                                                 const float x[kMlpIn], float out[kMlpOut])
{
    const int lane = threadIdx.x & (kWave - 1);
    float xn[kMlpIn];
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) xn[i] = (x[i] - pf[kMlpXm + i]) / pf[kMlpXs + i];
    float h1 = pf[kMlpB1 + lane];
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) h1 += bf16_to_float(wb[kMlpW1 + lane * kMlpIn + i]) * xn[i];
    h1 = tanhf(h1);
    float h2 = pf[kMlpB2 + lane];
    const unsigned short *w2 = wb + kMlpW2 + lane * kMlpHid;
#pragma unroll
    for (int k = 0; k < kMlpHid; ++k)
    {
        const float hk = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, h1), k));
        h2 += bf16_to_float(w2[k]) * hk;
    }
    h2 = tanhf(h2);
#pragma unroll
    for (int o = 0; o < kMlpOut; ++o)
    {
        const float y = wave_sum(bf16_to_float(wb[kMlpW3 + o * kMlpHid + lane]) * h2) + pf[kMlpB3 + o];
        out[o] = y * pf[kMlpYs + o] + pf[kMlpYm + o];
    }
}

}  // namespace ab
