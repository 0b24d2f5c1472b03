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

// The same 15 sums over the tensor grid xi_i x xi_j x xi_k without visiting
// its Nv^3 nodes. exp(a . psi) = e^a0 E_x(xi_i) E_y(xi_j) E_z(xi_k) with
// E_x(t) = exp(a1 t + a4 t^2) (and a2, a3 for y, z), so every sum of
// e x^p y^q z^r is e^a0 X_p Y_q Z_r, X_p = sum_i E_x(xi_i) xi_i^p. One
// particle per thread: 3 Nv exponentials instead of Nv^3.
__device__ __forceinline__ void dm_grid_sums(const double a[5], int Nv, double vlo, double dv,
                                             double ics, double dv3, double S[15])
{
    double X[5] = {0, 0, 0, 0, 0}, Y[5] = {0, 0, 0, 0, 0}, Z[5] = {0, 0, 0, 0, 0};
    for (int l = 0; l < Nv; ++l)
    {
        const double xi = vnode(vlo, dv, l) * ics, t2 = a[4] * xi * xi;
        double ex = exp(a[1] * xi + t2), ey = exp(a[2] * xi + t2), ez = exp(a[3] * xi + t2);
#pragma unroll
        for (int p = 0; p < 5; ++p)
        {
            X[p] += ex; Y[p] += ey; Z[p] += ez;
            ex *= xi; ey *= xi; ez *= xi;
        }
    }
    const double e0 = exp(a[0]) * dv3;
    const double x0 = e0 * X[0];
    const double y0z0 = Y[0] * Z[0], y1z0 = Y[1] * Z[0], y0z1 = Y[0] * Z[1];
    const double y2z0 = Y[2] * Z[0], y0z2 = Y[0] * Z[2], y1z1 = Y[1] * Z[1];
    S[0] = x0 * y0z0;
    S[1] = e0 * X[1] * y0z0;
    S[2] = x0 * y1z0;
    S[3] = x0 * y0z1;
    S[4] = e0 * (X[2] * y0z0 + X[0] * (y2z0 + y0z2));
    S[5] = e0 * X[2] * y0z0;
    S[6] = e0 * X[1] * y1z0;
    S[7] = e0 * X[1] * y0z1;
    S[8] = e0 * (X[3] * y0z0 + X[1] * (y2z0 + y0z2));
    S[9] = x0 * y2z0;
    S[10] = x0 * y1z1;
    S[11] = e0 * (X[2] * y1z0 + X[0] * (Y[3] * Z[0] + Y[1] * Z[2]));
    S[12] = x0 * y0z2;
    S[13] = e0 * (X[2] * y0z1 + X[0] * (Y[2] * Z[1] + Y[0] * Z[3]));
    S[14] = e0 * (X[4] * y0z0 + X[0] * (Y[4] * Z[0] + Y[0] * Z[4] + 2.0 * Y[2] * Z[2])
                  + 2.0 * X[2] * (y2z0 + y0z2));
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
    double L[5][5], inv[5];
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
                inv[i] = 1.0 / L[i][i];
            }
            else L[i][j] = s * inv[j];
        }
    }
#pragma unroll
    for (int i = 0; i < 5; ++i)
    {
        double s = F[i];
#pragma unroll
        for (int k = 0; k < i; ++k) s -= L[i][k] * y[k];
        y[i] = s * inv[i];
    }
#pragma unroll
    for (int i = 4; i >= 0; --i)
    {
        double s = y[i];
#pragma unroll
        for (int k = i + 1; k < 5; ++k) s -= L[k][i] * d[k];
        d[i] = s * inv[i];
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

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// tanh(x) = 1 - 2 / (1 + e^{2x}) with the hardware base-2 exponential and
// reciprocal (about 1e-7 absolute; the warm start needs no more). Saturates
// to +-1 for large |x| (e^{2x} -> inf or 0).
__device__ __forceinline__ float tanh_fast(float x)
{
    const float e = __builtin_amdgcn_exp2f(x * 2.8853900817779268f);     // 2 log2(e)
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(1.0f + e);
}

// MLP forward for the wave's 64 samples (lane l holds sample l; every lane
// of the wave must call it), BF16 weights on MFMA with FP32 accumulation.
// Layer 2 is computed transposed, Z2^T (n x p) = W2 (n x k) . H1^T (k x p),
// with v_mfma_f32_32x32x16_bf16 (A = W2 rows straight from memory, B = the
// layer-1 activations rounded to BF16, produced by each lane in the B layout:
// column p = l % 32 (+ 32), k = 16 ks + 8 (l / 32) + e). The result layout
// leaves each lane half of the 64 hidden units of its two columns, so layer 3
// is a lane-local dot product plus one exchange with lane l ^ 32.
__device__ __forceinline__ void mlp_wave_mfma(const unsigned short *__restrict__ wb,
                                              const float *__restrict__ pf,
                                              const float x[kMlpIn], float out[kMlpOut])
{
    const int lane = threadIdx.x & (kWave - 1), col = lane & 31, half = lane >> 5;
    // normalised inputs of the lane's columns p = col (pt 0) and col + 32 (pt 1)
    float xn[2][kMlpIn];
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i)
    {
        const float v = (x[i] - pf[kMlpXm + i]) / pf[kMlpXs + i];
        const float o = __shfl_xor(v, 32, kWave);
        xn[0][i] = half ? o : v;
        xn[1][i] = half ? v : o;
    }
    // layer 1: H1^T in the B layout, rounded to BF16
    bf16x8 B[2][4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
#pragma unroll
        for (int e = 0; e < 8; ++e)
        {
            const int k = 16 * ks + 8 * half + e;
            float w[kMlpIn];
#pragma unroll
            for (int i = 0; i < kMlpIn; ++i) w[i] = bf16_to_float(wb[kMlpW1 + k * kMlpIn + i]);
            const float b = pf[kMlpB1 + k];
#pragma unroll
            for (int pt = 0; pt < 2; ++pt)
            {
                float z = b;
#pragma unroll
                for (int i = 0; i < kMlpIn; ++i) z += w[i] * xn[pt][i];
                B[pt][ks][e] = (__bf16)tanh_fast(z);
            }
        }
    // layer 2: acc[nt][pt] = W2 rows 32 nt .. 32 nt + 31 times the columns of pt
    f32x16 acc[2][2];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int pt = 0; pt < 2; ++pt) acc[nt][pt] = f32x16{};
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
        {
            const bf16x8 A = *reinterpret_cast<const bf16x8 *>(
                wb + kMlpW2 + (32 * nt + col) * kMlpHid + 16 * ks + 8 * half);
#pragma unroll
            for (int pt = 0; pt < 2; ++pt)
                acc[nt][pt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, B[pt][ks], acc[nt][pt], 0, 0, 0);
        }
    // h2 = tanh(z2 + b2); the lane holds rows n = 32 nt + 8 (r / 4) + 4 half + r % 4
    float y[2][kMlpOut];
#pragma unroll
    for (int pt = 0; pt < 2; ++pt)
#pragma unroll
        for (int o = 0; o < kMlpOut; ++o) y[pt][o] = 0.0f;
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int r = 0; r < 16; ++r)
        {
            const int n = 32 * nt + 8 * (r >> 2) + 4 * half + (r & 3);
            const float b = pf[kMlpB2 + n];
            float w[kMlpOut];
#pragma unroll
            for (int o = 0; o < kMlpOut; ++o) w[o] = bf16_to_float(wb[kMlpW3 + o * kMlpHid + n]);
#pragma unroll
            for (int pt = 0; pt < 2; ++pt)
            {
                const float h = tanh_fast(acc[nt][pt][r] + b);
#pragma unroll
                for (int o = 0; o < kMlpOut; ++o) y[pt][o] += w[o] * h;
            }
        }
    // the other 32 hidden units of each column are in lane l ^ 32
#pragma unroll
    for (int pt = 0; pt < 2; ++pt)
#pragma unroll
        for (int o = 0; o < kMlpOut; ++o) y[pt][o] += __shfl_xor(y[pt][o], 32, kWave);
#pragma unroll
    for (int o = 0; o < kMlpOut; ++o)
        out[o] = ((half ? y[1][o] : y[0][o]) + pf[kMlpB3 + o]) * pf[kMlpYs + o] + pf[kMlpYm + o];
}

}  // namespace ab
