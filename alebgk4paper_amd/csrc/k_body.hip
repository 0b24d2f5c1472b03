// k_body.hip -- raw velocity moments of body-surface particles for the
// surface-stress loads (computeBodyLoads, RigidBody/BodyDynamics.cpp:17-97).
//
// The reference integrates P_ab = sum f (v_a - U_a)(v_b - U_b) dv^3 for every
// trial wall velocity U, which the Heun coupling evaluates twice per step on
// the same g (BodyDynamics.cpp:353, :399). The ten raw moments
//   m0 = sum f dv^3,  m_a = sum f v_a dv^3,  m_ab = sum f v_a v_b dv^3
// give P(U) = m_ab - U_a m_b - U_b m_a + U_a U_b m0 in closed form, so one
// pass over the surface particles' g serves every U. One wave per surface
// particle, FP64 accumulation.
#include "ab_device.hpp"
#include "ab_internal.hpp"

namespace ab {

namespace {

// Table storage holds g = L (M0 + d); m0 == nullptr means g is stored as is.
template <typename T>
__global__ void k_body_moments(DevCloud c, const T *__restrict__ g, int stride, const double *lazy,
                               const double *__restrict__ m0, const int *__restrict__ idx, int n,
                               double *__restrict__ mom)
{
    const int k = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    if (k >= n) return;
    const int i = idx[k];
    const double scale = lazy ? *lazy : 1.0;
    const T *p = g + (size_t)i * stride;
    const int Nv = c.Nv;
    const double dvd = (c.dim == 3) ? c.dv * c.dv * c.dv : c.dv * c.dv;
    double m[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int q = lane; q < c.nodes; q += kWave)
    {
        const double vx = vnode(c.vlo, c.dv, q % Nv);
        const double vy = vnode(c.vlo, c.dv, (q / Nv) % Nv);
        const double vz = (c.dim == 3) ? vnode(c.vlo, c.dv, q / (Nv * Nv)) : 0.0;
        const double f = ((double)p[q] + (m0 ? m0[q] : 0.0)) * scale * dvd;
        m[0] += f;
        m[1] += f * vx; m[2] += f * vy; m[3] += f * vz;
        m[4] += f * vx * vx; m[5] += f * vy * vy; m[6] += f * vz * vz;
        m[7] += f * vx * vy; m[8] += f * vx * vz; m[9] += f * vy * vz;
    }
#pragma unroll
    for (int q = 0; q < 10; ++q) m[q] = wave_sum(m[q]);
    if (lane < 10) mom[(size_t)k * 10 + lane] = m[lane];
}

}  // namespace

void launch_body_moments(const DevCloud &c, const double *g_fp64, const void *g_table,
                         int table_fp32, int S, const double *lazy, const double *m0,
                         const int *idx, int n, double *mom, hipStream_t s)
{
    if (n <= 0) return;
    const int nb = (n * kWave + 255) / 256;
    if (g_fp64)
        k_body_moments<double><<<nb, 256, 0, s>>>(c, g_fp64, c.gstride, nullptr, nullptr, idx, n,
                                                  mom);
    else if (table_fp32)
        k_body_moments<float><<<nb, 256, 0, s>>>(c, (const float *)g_table, S, lazy, m0, idx, n,
                                                 mom);
    else
        k_body_moments<double><<<nb, 256, 0, s>>>(c, (const double *)g_table, S, lazy, m0, idx,
                                                  n, mom);
}

}  // namespace ab
