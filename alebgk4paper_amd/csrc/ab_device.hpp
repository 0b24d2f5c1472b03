// ab_device.hpp -- device helpers shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#ifndef AB_PI
#define AB_PI 3.14159265358979323846
#endif

namespace ab {

constexpr int kWave = 64;          // CDNA wavefront
constexpr int kBlock = 256;        // 4 waves per workgroup for the per-particle kernels
constexpr int kMaxNv = 48;         // velocity nodes per axis (ab_create refuses more)
// Lattice planes one z block of the 3-D table transport reads (its largest
// block, 12 outputs, plus the 4-plane stencil). Its buffer descriptors take
// 32-bit byte offsets, so such a span of the stored g must stay below 2 GiB.
constexpr int kMaxZBlockPlanes = 12 + 4;

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on
// XCD b % 8). Give every XCD a contiguous range of logical block ids so the
// particles a XCD processes are spatially close and share its L2.
__device__ __forceinline__ int xcd_remap(int b, int nblocks)
{
    int xcd = b & 7, slot = b >> 3;
    int per = nblocks >> 3, rem = nblocks & 7;
    return xcd * per + (xcd < rem ? xcd : rem) + slot;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// A value every lane of the wave holds (a wave reduction's result) moved to
// scalar registers: per-particle quantities then cost no VGPRs.
__device__ __forceinline__ double uniform(double x)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Sums eight values over the wave with 10 exchanges instead of 48. At each
// of the first three stages a lane keeps half of its values and adds its
// partner's copy of that half (its lane bit picks the half), then a plain
// butterfly over lane bits 2..0; value q ends in lanes 8q .. 8q + 7 and is
// read from lane 8q into scalar registers. Fixed order: deterministic.
__device__ __forceinline__ void wave_sum8(double (&v)[8])
{
    const int lane = threadIdx.x & (kWave - 1);
    const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8;
    double a[4], b[2];
#pragma unroll
    for (int j = 0; j < 4; ++j)
    {
        const double keep = b5 ? v[j + 4] : v[j], send = b5 ? v[j] : v[j + 4];
        a[j] = keep + __shfl_xor(send, 32, kWave);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j)
    {
        const double keep = b4 ? a[j + 2] : a[j], send = b4 ? a[j] : a[j + 2];
        b[j] = keep + __shfl_xor(send, 16, kWave);
    }
    double c = (b3 ? b[1] : b[0]) + __shfl_xor(b3 ? b[0] : b[1], 8, kWave);
    c += __shfl_xor(c, 4, kWave);
    c += __shfl_xor(c, 2, kWave);
    c += __shfl_xor(c, 1, kWave);
    const unsigned long long u = __builtin_bit_cast(unsigned long long, c);
#pragma unroll
    for (int q = 0; q < 8; ++q)
    {
        const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, 8 * q);
        const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), 8 * q);
        v[q] = __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
    }
}

// Lane l's value of v in scalar registers (l wave-uniform): a per-wave table
// held one entry per lane is read without memory.
__device__ __forceinline__ double lane_value(double v, int l)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)b, l);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(b >> 32), l);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Orders a wave's LDS writes before its later LDS reads of other lanes'
// entries (a wave-private table; no workgroup barrier, so waves that left
// the kernel early do not matter).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T>
__device__ __forceinline__ T wave_min(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, kWave));
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}

// Sum over a kBlock-thread block in a fixed order (wave tree, then the four
// wave totals in wave order). Every thread gets the result. `sh` needs
// kBlock/kWave slots per value; NV values are reduced together.
template <int NV, typename T>
__device__ __forceinline__ void block_sum(T (&v)[NV], T *sh)
{
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
    __syncthreads();
    if (lane == 0)
    {
#pragma unroll
        for (int k = 0; k < NV; ++k) sh[k * (kBlock / kWave) + w] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k)
    {
        T s = sh[k * (kBlock / kWave)];
#pragma unroll
        for (int q = 1; q < kBlock / kWave; ++q) s += sh[k * (kBlock / kWave) + q];
        v[k] = s;
    }
    __syncthreads();
}

// Sums over one wave (every lane gets them): the reductions of kernels that
// give each particle one wave.
template <int NV, typename T>
__device__ __forceinline__ void wave_sums(T (&v)[NV])
{
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
}

template <typename T>
__device__ __forceinline__ void block_minmax(T &mn, T &mx, T *sh)
{
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    mn = wave_min(mn);
    mx = wave_max(mx);
    __syncthreads();
    if (lane == 0) { sh[w] = mn; sh[kBlock / kWave + w] = mx; }
    __syncthreads();
    T a = sh[0], b = sh[kBlock / kWave];
#pragma unroll
    for (int q = 1; q < kBlock / kWave; ++q)
    {
        a = fmin(a, sh[q]);
        b = fmax(b, sh[kBlock / kWave + q]);
    }
    mn = a; mx = b;
    __syncthreads();
}

// Velocity node vlo + idx*dv, rounded as the reference's CPU code rounds it
// (product, then sum). A fused multiply-add differs in the last bit for some
// nodes, and on a corner particle (n = (1,1)/sqrt2) that bit decides whether
// the node with v_x = -v_y counts as incoming (c_n < 0) or not.
__host__ __device__ __forceinline__ double vnode(double vlo, double dv, int idx)
{
#pragma clang fp contract(off)
    return vlo + idx * dv;
}

// c_n = (v - U_w) . n, the sign that gates transport and splits the diffuse
// wall's fluxes (CudaBackend.cu:715-719, 974-986). Evaluated without FMA
// contraction, as the reference writes it: on an edge particle the two
// products are exact negatives for the nodes with v_x = v_y and c_n is then
// exactly 0 (neither incoming nor outgoing); a fused multiply-add leaves a
// +-1e-13 residue that flips those nodes.
__device__ __forceinline__ double wall_cn3(double vx, double vy, double vz, double Ux,
                                           double Uy, double Uz, double nx, double ny,
                                           double nz)
{
#pragma clang fp contract(off)
    return (vx - Ux) * nx + (vy - Uy) * ny + (vz - Uz) * nz;
}

__device__ __forceinline__ double wall_cn2(double vx, double vy, double Ux, double Uy,
                                           double nx, double ny)
{
#pragma clang fp contract(off)
    return (vx - Ux) * nx + (vy - Uy) * ny;
}

// maxwellian3D (MLSTransport/MLSBasis.hpp:112-122), FP64.
__device__ __forceinline__ double maxwellian3d(double vx, double vy, double vz,
                                               double ux, double uy, double uz,
                                               double T, double rho, double R)
{
    double twoRT = 2.0 * R * T;
    double e = -((vx - ux) * (vx - ux) + (vy - uy) * (vy - uy) + (vz - uz) * (vz - uz)) / twoRT;
    double norm = pow(AB_PI * twoRT, 1.5);
    return rho / norm * exp(e);
}

// maxwellianG1 (MLSTransport/MLSBasis.hpp:87-94), FP64.
__device__ __forceinline__ double maxwellian_g1(double vx, double vy, double ux, double uy,
                                                double T, double rho, double R)
{
    double twoRT = 2.0 * R * T;
    double e = -((vx - ux) * (vx - ux) + (vy - uy) * (vy - uy)) / twoRT;
    return rho / (AB_PI * twoRT) * exp(e);
}

}  // namespace ab
