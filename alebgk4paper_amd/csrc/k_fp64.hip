// k_fp64.hip -- the reference formulation of the ALE-BGK step in FP64 on gfx950.
//
// Transport: one thread per (particle, velocity node), coalesced over the
// node index; a workgroup covers 256 nodes of one particle and stages that
// particle's neighbour positions in LDS (every lane reads the same entry: an
// LDS broadcast). Interior and boundary particles are launched from compact
// lists, each with its compile-time basis size, instead of sweeping the whole
// grid twice with early exits as the reference does (CudaBackend.cu:1667-1677).
//
// Per-particle stages: one 256-thread workgroup per particle, coalesced over
// the node index, with fixed-order workgroup reductions. They fuse reference
// kernels that touch only their own particle:
//   limit_commit  kPositivity3D + kNAD3D + kCommit  (+ mass partial)
//   wall_collide  kScaleG + kWallBC3D + kMoments3D + kCollisionNode3D
//   final         kScaleG + kMoments3D
// and the global sums stay on the device (no host round trip per rescale).
#include "ab_device.hpp"
#include "ab_dm.hpp"
#include "ab_internal.hpp"

namespace ab {

namespace {

__device__ __forceinline__ void node_velocity(const DevCloud &c, int lin, double &vx,
                                              double &vy, double &vz)
{
    int Nv = c.Nv;
    int iv = lin % Nv, jv = (lin / Nv) % Nv, kv = lin / (Nv * Nv);
    vx = vnode(c.vlo, c.dv, iv);
    vy = vnode(c.vlo, c.dv, jv);
    vz = (c.dim == 3) ? vnode(c.vlo, c.dv, kv) : 0.0;
}

// kCopyGt + kTransportNode3DT<P, BND> (Solver/CudaBackend.cu:191-197, 699-869).
template <int P, bool BND>
__global__ __launch_bounds__(kBlock) void k_transport3d(DevCloud c, const double *__restrict__ g,
                                                        double *__restrict__ gt, double dt,
                                                        double h, double tol, int clamp,
                                                        int *mlsflag, double *mlsovr,
                                                        const int *__restrict__ list,
                                                        int nchunk, const int *__restrict__ only)
{
    __shared__ double sx[AB_MAX_NEIGH], sy[AB_MAX_NEIGH], sz[AB_MAX_NEIGH];
    __shared__ int sj[AB_MAX_NEIGH];
    __shared__ double sred[2 * (kBlock / kWave)];
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int i = list[b / nchunk];
    if (only && !only[i]) return;            // whole block: one particle
    const int lin = (b % nchunk) * kBlock + threadIdx.x;
    const int nn = c.nneigh[i];
    for (int t = threadIdx.x; t < nn; t += kBlock)
    {
        int j = c.neigh[(size_t)i * AB_MAX_NEIGH + t];
        sj[t] = j; sx[t] = c.x[j]; sy[t] = c.y[j]; sz[t] = c.z[j];
    }
    __syncthreads();

    const size_t gs = (size_t)c.gstride;
    double excess = 0.0;
    int flag = 0;
    if (lin < c.nodes)
    {
        double vx, vy, vz;
        node_velocity(c, lin, vx, vy, vz);
        const double g_self = g[(size_t)i * gs + lin];
        double out = g_self;
        bool gated = false;
        if (BND)
        {
            double cn = wall_cn3(vx, vy, vz, c.wux[i], c.wuy[i], c.wuz[i], c.wnx[i], c.wny[i],
                                 c.wnz[i]);
            gated = (cn >= 0.0);
        }
        if (!gated)
        {
            const double xf = c.x[i] - (vx - c.mvx[i]) * dt;
            const double yf = c.y[i] - (vy - c.mvy[i]) * dt;
            const double zf = c.z[i] - (vz - c.mvz[i]) * dt;
            const double hh = h * h;
            double A[P * P], b1[P];
#pragma unroll
            for (int q = 0; q < P * P; ++q) A[q] = 0.0;
#pragma unroll
            for (int q = 0; q < P; ++q) b1[q] = 0.0;
            bool has = false;
            double sumW = 0.0, sumWG = 0.0, gmin = 1e300, gmax = -1e300;
            for (int m = 0; m < nn; ++m)
            {
                const double dxn = sx[m] - xf, dyn = sy[m] - yf, dzn = sz[m] - zf;
                const double r2 = dxn * dxn + dyn * dyn + dzn * dzn;
                const double q2 = r2 / hh;                  // mlsWeightSq
                double w = 0.0;
                if (q2 < 1.0) { double t1 = 1.0 - q2; w = t1 * t1; }
                if (w < 1e-30) continue;
                double basis[P];
                const double ssx = dxn / h, ssy = dyn / h, ssz = dzn / h;
                basis[0] = 1.0; basis[1] = ssx; basis[2] = ssy; basis[3] = ssz;
                if constexpr (P >= 10)
                {
                    basis[4] = ssx * ssx; basis[5] = ssy * ssy; basis[6] = ssz * ssz;
                    basis[7] = ssx * ssy; basis[8] = ssx * ssz; basis[9] = ssy * ssz;
                }
                const double gj = g[(size_t)sj[m] * gs + lin];
#pragma unroll
                for (int p1 = 0; p1 < P; ++p1)
                {
                    const double wb = w * basis[p1];
#pragma unroll
                    for (int p2 = p1; p2 < P; ++p2) A[p1 * P + p2] += wb * basis[p2];
                    b1[p1] += wb * gj;
                }
                sumW += w;
                sumWG += w * gj;
                gmin = fmin(gmin, gj);
                gmax = fmax(gmax, gj);
                has = true;
            }
            if (has)
            {
                double dmax = 0.0;
#pragma unroll
                for (int k = 0; k < P; ++k) dmax = fmax(dmax, A[k * P + k]);
                bool degenerate = (dmax <= 0.0);
#pragma unroll
                for (int k = 0; k < P; ++k)
                {
                    double d = A[k * P + k];
#pragma unroll
                    for (int j = 0; j < k; ++j) { double l = A[j * P + k]; d -= l * l; }
                    if (!(d > tol * dmax)) degenerate = true;
                    d = degenerate ? 1.0 : sqrt(d);
                    A[k * P + k] = d;
#pragma unroll
                    for (int m2 = k + 1; m2 < P; ++m2)
                    {
                        double v = A[k * P + m2];
#pragma unroll
                        for (int j = 0; j < k; ++j) v -= A[j * P + k] * A[j * P + m2];
                        A[k * P + m2] = v / d;
                    }
                }
#pragma unroll
                for (int k = 0; k < P; ++k)
                {
                    double v = b1[k];
#pragma unroll
                    for (int j = 0; j < k; ++j) v -= A[j * P + k] * b1[j];
                    b1[k] = v / A[k * P + k];
                }
#pragma unroll
                for (int k = P - 1; k >= 0; --k)
                {
                    double v = b1[k];
#pragma unroll
                    for (int j = k + 1; j < P; ++j) v -= A[k * P + j] * b1[j];
                    b1[k] = v / A[k * P + k];
                }
                out = b1[0];
                if (degenerate || !isfinite(out)) out = (sumW > 0.0) ? (sumWG / sumW) : g_self;
                if (degenerate || !isfinite(b1[0])) flag |= 1;
                excess = fmax(fmax(out - gmax, gmin - out), 0.0);
                bool overshoot = (out < gmin) || (out > gmax);
                if (overshoot) flag |= 2;
                if (overshoot && clamp) out = fmin(fmax(out, gmin), gmax);
            }
        }
        gt[(size_t)i * gs + lin] = out;
    }
    // One flag word and one overshoot sum per workgroup.
    double red[1] = {excess};
    block_sum<1>(red, sred);
    int f = flag;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) f |= __shfl_xor(f, o, kWave);
    __shared__ int sflag[kBlock / kWave];
    if ((threadIdx.x & (kWave - 1)) == 0) sflag[threadIdx.x / kWave] = f;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        int fa = 0;
        for (int q = 0; q < kBlock / kWave; ++q) fa |= sflag[q];
        if (fa && mlsflag) atomicOr(&mlsflag[i], fa);
        if (red[0] > 0.0 && mlsovr)
        {
            double dv3 = c.dv * c.dv * c.dv;
            atomicAdd(&mlsovr[i], red[0] * dv3);
        }
    }
}

// kCopyGt + kTransportNode, 2-D Chu layers (Solver/CudaBackend.cu:199-300):
// weight through sqrt + mlsWeight, Gauss with partial pivoting on both layers.
// Row swaps compare unrolled row indices against the pivot so A stays in
// registers. A is summed in its upper triangle and mirrored: the reference
// sums every entry, so the lower triangle differs from its by the rounding
// of w phi_a phi_b against w phi_b phi_a.
template <int P, bool BND>
__global__ __launch_bounds__(kBlock) void k_transport2d(DevCloud c, const double *__restrict__ g,
                                                        double *__restrict__ gt, double dt,
                                                        double h, const int *__restrict__ list,
                                                        int nchunk, const int *__restrict__ only)
{
    __shared__ double sx[AB_MAX_NEIGH], sy[AB_MAX_NEIGH];
    __shared__ int sj[AB_MAX_NEIGH];
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int i = list[b / nchunk];
    if (only && !only[i]) return;            // whole block: one particle
    const int lin = (b % nchunk) * kBlock + threadIdx.x;
    const int nn = c.nneigh[i];
    if (c.inst) h = c.ih[c.inst[i]];
    for (int t = threadIdx.x; t < nn; t += kBlock)
    {
        int j = c.neigh[(size_t)i * AB_MAX_NEIGH + t];
        sj[t] = j; sx[t] = c.x[j]; sy[t] = c.y[j];
    }
    __syncthreads();
    if (lin >= c.nodes) return;
    const size_t gs = (size_t)c.gstride;
    const int Nv2 = c.nodes;
    const double vx = vnode(c.vlo, c.dv, lin % c.Nv);
    const double vy = vnode(c.vlo, c.dv, lin / c.Nv);
    double o1 = g[(size_t)i * gs + lin], o2 = g[(size_t)i * gs + lin + Nv2];
    bool gated = false;
    if (BND)
    {
        double cn = wall_cn2(vx, vy, c.wux[i], c.wuy[i], c.wnx[i], c.wny[i]);
        gated = (cn >= 0.0);
    }
    if (!gated)
    {
        const double xf = c.x[i] - (vx - c.mvx[i]) * dt;
        const double yf = c.y[i] - (vy - c.mvy[i]) * dt;
        double A[P * P], bb1[P], bb2[P];
#pragma unroll
        for (int q = 0; q < P * P; ++q) A[q] = 0.0;
#pragma unroll
        for (int q = 0; q < P; ++q) { bb1[q] = 0.0; bb2[q] = 0.0; }
        bool has = false;
        for (int m = 0; m < nn; ++m)
        {
            const double dxn = sx[m] - xf, dyn = sy[m] - yf;
            const double r = sqrt(dxn * dxn + dyn * dyn);
            double w = 0.0;                                   // mlsWeight
            if (r < h) { double qq = r / h; double t1 = 1.0 - qq * qq; w = t1 * t1; }
            if (w < 1e-30) continue;
            double basis[P];
            const double ssx = dxn / h, ssy = dyn / h;
            basis[0] = 1.0; basis[1] = ssx; basis[2] = ssy;
            if constexpr (P >= 6) { basis[3] = ssx * ssx; basis[4] = ssx * ssy; basis[5] = ssy * ssy; }
            // A is symmetric: the upper triangle here, mirrored after the sum
#pragma unroll
            for (int p1 = 0; p1 < P; ++p1)
#pragma unroll
                for (int p2 = p1; p2 < P; ++p2) A[p1 * P + p2] += w * basis[p1] * basis[p2];
            const double *gj = g + (size_t)sj[m] * gs;
            const double gn1 = gj[lin], gn2 = gj[lin + Nv2];
#pragma unroll
            for (int q = 0; q < P; ++q)
            {
                bb1[q] += w * basis[q] * gn1;
                bb2[q] += w * basis[q] * gn2;
            }
            has = true;
        }
        if (has)
        {
#pragma unroll
            for (int p1 = 1; p1 < P; ++p1)
#pragma unroll
                for (int p2 = 0; p2 < p1; ++p2) A[p1 * P + p2] = A[p2 * P + p1];
#pragma unroll
            for (int col = 0; col < P; ++col)
            {
                int piv = col;
                double mx = fabs(A[col * P + col]);
#pragma unroll
                for (int row = col + 1; row < P; ++row)
                    if (fabs(A[row * P + col]) > mx) { mx = fabs(A[row * P + col]); piv = row; }
#pragma unroll
                for (int row = col + 1; row < P; ++row)
                {
                    if (row == piv)
                    {
#pragma unroll
                        for (int q = 0; q < P; ++q)
                        {
                            double t1 = A[col * P + q];
                            A[col * P + q] = A[row * P + q];
                            A[row * P + q] = t1;
                        }
                        double tb = bb1[col]; bb1[col] = bb1[row]; bb1[row] = tb;
                        tb = bb2[col]; bb2[col] = bb2[row]; bb2[row] = tb;
                    }
                }
                const double d = A[col * P + col];
                if (fabs(d) < 1e-30) continue;
#pragma unroll
                for (int row = col + 1; row < P; ++row)
                {
                    const double f = A[row * P + col] / d;
#pragma unroll
                    for (int q = col; q < P; ++q) A[row * P + q] -= f * A[col * P + q];
                    bb1[row] -= f * bb1[col];
                    bb2[row] -= f * bb2[col];
                }
            }
            double s1[P], s2[P];
#pragma unroll
            for (int row = P - 1; row >= 0; --row)
            {
                double v1 = bb1[row], v2 = bb2[row];
#pragma unroll
                for (int q = row + 1; q < P; ++q)
                {
                    v1 -= A[row * P + q] * s1[q];
                    v2 -= A[row * P + q] * s2[q];
                }
                const double d = A[row * P + row];
                s1[row] = fabs(d) > 1e-30 ? v1 / d : 0.0;
                s2[row] = fabs(d) > 1e-30 ? v2 / d : 0.0;
            }
            o1 = s1[0];
            o2 = s2[0];
        }
    }
    gt[(size_t)i * gs + lin] = o1;
    gt[(size_t)i * gs + lin + Nv2] = o2;
}

// Per-particle mass partial: sum of g*dv^d over the mass-carrying layer
// (kSumMass3D / kSumMass per-thread part, CudaBackend.cu:391-436).
__global__ __launch_bounds__(kBlock) void k_mass_partials(DevCloud c, const double *__restrict__ g,
                                                          double *__restrict__ pmass)
{
    __shared__ double sh[kBlock / kWave];
    const int i = xcd_remap(blockIdx.x, gridDim.x);
    const double dvd = (c.dim == 3) ? c.dv * c.dv * c.dv : c.dv * c.dv;
    const double *p = g + (size_t)i * c.gstride;
    double m[1] = {0.0};
    for (int k = threadIdx.x; k < c.nodes; k += kBlock) m[0] += p[k] * dvd;
    block_sum<1>(m, sh);
    if (threadIdx.x == 0) pmass[i] = m[0];
}

// Fixed-order sum of the per-particle partials and the rescale factor
// (sumMassHost + the scalar logic of cudaStep :1627, :1706-1710, :1773-1780),
// kept on the device so no host synchronisation is needed.
// With seg, block k sums instance k's particles seg[k] .. seg[k+1]-1 into
// its scalars (S_COUNT apart) in the same order as a cloud of its own.
__global__ __launch_bounds__(1024) void k_reduce_mass(int N, const double *__restrict__ pmass,
                                                      double *scal, int mode, int nofix,
                                                      const int *__restrict__ seg,
                                                      const double *__restrict__ pnew)
{
    __shared__ double sh[1024 / kWave], shn[1024 / kWave];
    int lo = 0, hi = N;
    if (seg)
    {
        lo = seg[blockIdx.x];
        hi = seg[blockIdx.x + 1];
        scal += (size_t)blockIdx.x * S_COUNT;
    }
    double s = 0.0, sn = 0.0;
    for (int i = lo + threadIdx.x; i < hi; i += 1024)
    {
        s += pmass[i];
        if (pnew) sn += pnew[i];
    }
    s = wave_sum(s);
    sn = wave_sum(sn);
    if ((threadIdx.x & (kWave - 1)) == 0) { sh[threadIdx.x / kWave] = s; shn[threadIdx.x / kWave] = sn; }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double t = 0.0, tn = 0.0;
        for (int q = 0; q < 1024 / kWave; ++q) { t += sh[q]; tn += shn[q]; }
        if (mode == R_MASS_FUSED)
        {
            // massOld / massNew, the collided mass after that rescale, and the
            // one end-of-step factor that the unscaled collided rows need:
            // massOld / massColl * massOld / massNew = massOld / t
            const double s1 = (tn > 1e-30) ? scal[S_MASS_OLD] / tn : 1.0;
            scal[S_MASS_NEW] = tn;
            scal[S_SCALE1] = s1;
            scal[S_MASS_COLL] = s1 * t;
            scal[S_SCALE2] = (!nofix && s1 * t > 1e-30) ? scal[S_MASS_OLD] / t : s1;
        }
        else if (mode == R_MASS_OLD)
            scal[S_MASS_OLD] = t;
        else if (mode == R_MASS_NEW)
        {
            scal[S_MASS_NEW] = t;
            scal[S_SCALE1] = (t > 1e-30) ? scal[S_MASS_OLD] / t : 1.0;
        }
        else
        {
            scal[S_MASS_COLL] = t;
            scal[S_SCALE2] = (!nofix && t > 1e-30) ? scal[S_MASS_OLD] / t : 1.0;
        }
    }
}

// kPositivity3D + kNAD3D + kCommit (CudaBackend.cu:882-954, 328-336) for 3-D;
// kPositivity + kCommit (:304-336) for 2-D. Writes the committed g, the
// limiter bits and the particle's mass.
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_limit_commit(DevCloud c, double *__restrict__ g,
                                                         const double *__restrict__ gt,
                                                         const double *__restrict__ rho,
                                                         double *__restrict__ limflag,
                                                         double *__restrict__ pmass)
{
    const int i = xcd_remap(blockIdx.x, gridDim.x) * (kBlock / kWave) + (int)(threadIdx.x / kWave);
    if (i >= c.N) return;                              // whole wave: one particle
    const int lane = threadIdx.x & (kWave - 1);
    const size_t gs = (size_t)c.gstride;
    double *gp = g + (size_t)i * gs;
    const double *tp = gt + (size_t)i * gs;
    const int nodes = c.nodes;
    double flag = 0.0;
    double sc = 1.0, sc2 = 1.0;
    bool fires = false, nad = false;

    if constexpr (DIM == 3)
    {
        const double dv3 = c.dv * c.dv * c.dv;
        double r[3] = {0.0, 0.0, 0.0};             // mPos, mNeg, mFallback
        for (int k = lane; k < nodes; k += kWave)
        {
            double v = tp[k];
            if (v >= 0) r[0] += v;
            else { r[1] += v; r[2] += gp[k]; }
        }
        wave_sums<3>(r);
        fires = (r[1] < -1e-30 && (r[0] + r[1]) > 0);
        if (fires)
        {
            flag += 1.0;
            double target = r[0] + r[1], replaced = r[0] + r[2];
            sc = (replaced > 1e-30) ? target / replaced : 0.0;
        }
        if (!c.bnd[i])
        {
            double rn[1] = {0.0};
            for (int k = lane; k < nodes; k += kWave)
            {
                double v = tp[k];
                if (fires) v = ((v < 0.0) ? gp[k] : v) * sc;
                rn[0] += v;
            }
            wave_sums<1>(rn);
            const double rho_new = rn[0] * dv3;
            const int nn = c.nneigh[i];
            double rmin = 1e300, rmax = -1e300;
            for (int k = lane; k < nn; k += kWave)
            {
                double q = rho[c.neigh[(size_t)i * AB_MAX_NEIGH + k]];
                rmin = fmin(rmin, q);
                rmax = fmax(rmax, q);
            }
            rmin = wave_min(rmin);
            rmax = wave_max(rmax);
            if (!(rmin > rmax))
            {
                double slack = 0.1 * fmax(rmax - rmin, 1e-12) + 1e-9;
                double lo = rmin - slack, hi = rmax + slack;
                if (rho_new > hi && rho_new > 1e-30) { nad = true; sc2 = hi / rho_new; }
                else if (rho_new < lo && rho_new > 1e-30) { nad = true; sc2 = lo / rho_new; }
                if (nad) flag += 2.0;
            }
        }
        double m[1] = {0.0};
        for (int k = lane; k < nodes; k += kWave)
        {
            double v = tp[k];
            if (fires) v = ((v < 0.0) ? gp[k] : v) * sc;
            if (nad) v *= sc2;
            gp[k] = v;
            m[0] += v * dv3;
        }
        wave_sums<1>(m);
        if (lane == 0) { pmass[i] = m[0]; limflag[i] = flag; }
    }
    else
    {
        double r[2] = {0.0, 0.0};                  // mPos, mNeg of g1
        for (int k = lane; k < nodes; k += kWave)
        {
            double v = tp[k];
            if (v >= 0) r[0] += v; else r[1] += v;
        }
        wave_sums<2>(r);
        fires = (r[1] < -1e-30 && r[0] > 1e-30 && (r[0] + r[1]) > 0);
        if (fires) { flag += 1.0; sc = (r[0] + r[1]) / r[0]; }
        const double dv2 = c.dv * c.dv;
        double m[1] = {0.0};
        for (int k = lane; k < nodes; k += kWave)
        {
            double v1 = tp[k], v2 = tp[k + nodes];
            if (fires)
            {
                if (v1 < 0) { v1 = 0; v2 = 0; }
                else { v1 *= sc; v2 *= sc; }
            }
            gp[k] = v1;
            gp[k + nodes] = v2;
            m[0] += v1 * dv2;
        }
        wave_sums<1>(m);
        if (lane == 0) { pmass[i] = m[0]; limflag[i] = flag; }
    }
}

// kScaleG (massOld/massNew) + kWallBC3D / kMoments3D + kCollisionNode3D
// (CudaBackend.cu:478-485, 956-1030, 532-569, 1181-1197) and their 2-D
// twins (:338-389, 124-158, 1165-1179). The scale is applied as the value is
// read, which rounds exactly as the stored product would.
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_wall_collide(DevCloud c, DmArgs dm, double *__restrict__ g,
                                                         const double *__restrict__ scal,
                                                         double dt, double tau, double R,
                                                         double *rho, double *ux, double *uy,
                                                         double *uz, double *T,
                                                         double *__restrict__ pmass,
                                                         const int *__restrict__ guard)
{
    const int i = xcd_remap(blockIdx.x, gridDim.x) * (kBlock / kWave) + (int)(threadIdx.x / kWave);
    if (i >= c.N) return;                              // whole wave: one particle
    const int lane = threadIdx.x & (kWave - 1);
    if (guard && (!*guard || c.bnd[i])) return;        // whole wave: one particle
    if (c.inst) scal += (size_t)c.inst[i] * S_COUNT;
    const double s1 = scal[S_SCALE1];
    const bool doScale = !guard && scal[S_MASS_NEW] > 1e-30;
    const size_t gs = (size_t)c.gstride;
    double *gp = g + (size_t)i * gs;
    const int nodes = c.nodes, Nv = c.Nv;
    const double dvd = (DIM == 3) ? c.dv * c.dv * c.dv : c.dv * c.dv;
    double m[1] = {0.0};

    if (c.bnd[i])
    {
        const double nx = c.wnx[i], ny = c.wny[i], nz = (DIM == 3) ? c.wnz[i] : 0.0;
        const double Uwx = c.wux[i], Uwy = c.wuy[i], Uwz = (DIM == 3) ? c.wuz[i] : 0.0;
        const double Tw = c.wt[i];
        const double RT2w = 2.0 * R * Tw;
        const double norm = (DIM == 3) ? pow(AB_PI * RT2w, 1.5) : (AB_PI * RT2w);
        double r[2] = {0.0, 0.0};                  // intM, intP
        for (int k = lane; k < nodes; k += kWave)
        {
            double vx = vnode(c.vlo, c.dv, k % Nv);
            double vy = vnode(c.vlo, c.dv, (k / Nv) % Nv);
            double vz = (DIM == 3) ? vnode(c.vlo, c.dv, k / (Nv * Nv)) : 0.0;
            double v = doScale ? gp[k] * s1 : gp[k];
            double cn = (DIM == 3) ? wall_cn3(vx, vy, vz, Uwx, Uwy, Uwz, nx, ny, nz)
                                   : wall_cn2(vx, vy, Uwx, Uwy, nx, ny);
            if (cn < 0.0) r[0] += cn * v * dvd;
            else if (cn > 0.0)
            {
                double e2 = (DIM == 3) ? (vx - Uwx) * (vx - Uwx) + (vy - Uwy) * (vy - Uwy)
                                             + (vz - Uwz) * (vz - Uwz)
                                       : (vx - Uwx) * (vx - Uwx) + (vy - Uwy) * (vy - Uwy);
                double Mw = exp(-(e2) / RT2w) / norm;
                r[1] += cn * Mw * dvd;
            }
        }
        wave_sums<2>(r);
        const double rho_w = (r[1] > 1e-30) ? (-r[0] / r[1]) : 1.0;
        for (int k = lane; k < nodes; k += kWave)
        {
            double vx = vnode(c.vlo, c.dv, k % Nv);
            double vy = vnode(c.vlo, c.dv, (k / Nv) % Nv);
            double vz = (DIM == 3) ? vnode(c.vlo, c.dv, k / (Nv * Nv)) : 0.0;
            double v = doScale ? gp[k] * s1 : gp[k];
            double v2 = 0.0;
            if constexpr (DIM == 2) v2 = doScale ? gp[k + nodes] * s1 : gp[k + nodes];
            double cn = (DIM == 3) ? wall_cn3(vx, vy, vz, Uwx, Uwy, Uwz, nx, ny, nz)
                                   : wall_cn2(vx, vy, Uwx, Uwy, nx, ny);
            if (cn > 0.0)
            {
                double e2 = (DIM == 3) ? (vx - Uwx) * (vx - Uwx) + (vy - Uwy) * (vy - Uwy)
                                             + (vz - Uwz) * (vz - Uwz)
                                       : (vx - Uwx) * (vx - Uwx) + (vy - Uwy) * (vy - Uwy);
                v = rho_w * exp(-(e2) / RT2w) / norm;
                if constexpr (DIM == 2) v2 = 0.5 * R * Tw * v;
            }
            gp[k] = v;
            if constexpr (DIM == 2) gp[k + nodes] = v2;
            m[0] += v * dvd;
        }
        if (lane == 0)
        {
            rho[i] = rho_w; ux[i] = Uwx; uy[i] = Uwy; T[i] = Tw;
            if (DIM == 3) uz[i] = Uwz;
        }
    }
    else
    {
        double r[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int k = lane; k < nodes; k += kWave)
        {
            double vx = vnode(c.vlo, c.dv, k % Nv);
            double vy = vnode(c.vlo, c.dv, (k / Nv) % Nv);
            double f = doScale ? gp[k] * s1 : gp[k];
            if constexpr (DIM == 3)
            {
                double vz = vnode(c.vlo, c.dv, k / (Nv * Nv));
                r[0] += f * dvd;
                r[1] += vx * f * dvd;
                r[2] += vy * f * dvd;
                r[3] += vz * f * dvd;
                r[4] += (vx * vx + vy * vy + vz * vz) * f * dvd;
            }
            else
            {
                double f2 = doScale ? gp[k + nodes] * s1 : gp[k + nodes];
                r[0] += f * dvd;
                r[1] += vx * f * dvd;
                r[2] += vy * f * dvd;
                r[4] += (vx * vx + vy * vy) * f * dvd + 2.0 * f2 * dvd;
            }
        }
        wave_sums<5>(r);
        double mrho = rho[i], mux = ux[i], muy = uy[i], muz = (DIM == 3) ? uz[i] : 0.0, mT = T[i];
        if (r[0] > 1e-30)
        {
            mrho = r[0];
            mux = r[1] / r[0];
            muy = r[2] / r[0];
            double u2;
            if constexpr (DIM == 3) { muz = r[3] / r[0]; u2 = mux * mux + muy * muy + muz * muz; }
            else u2 = mux * mux + muy * muy;
            double t1 = (r[4] / r[0] - u2) / (3.0 * R);
            mT = t1 > 1e-6 ? t1 : 1e-6;
        }
        const double lam = dt / tau, coeff = 1.0 / (1.0 + lam);
        // discrete-Maxwellian target: Newton on a from the continuous guess
        double a[5] = {0, 0, 0, 0, 0};
        bool useDm = false;
        if constexpr (DIM == 3)
        {
            if (dm.on && r[0] > 1e-30)
            {
                useDm = true;
                const double cs = dm.cs;
                double ac[5];
                dm_guess(mrho, mux, muy, muz, mT, R, cs, ac);
                const float feat[kMlpIn] = {float(mux / cs), float(muy / cs), float(muz / cs),
                                            float(R * mT / (cs * cs)), float(c.dv / cs)};
                float da[kMlpOut] = {0, 0, 0, 0, 0};
                if (dm.mlp) mlp_wave_forward(dm.wb, dm.pf, feat, da);
#pragma unroll
                for (int q = 0; q < 5; ++q) a[q] = ac[q] + (double)da[q];
                const double ms[5] = {r[0], r[1] / cs, r[2] / cs, r[3] / cs, r[4] / (cs * cs)};
                int pass = 0;
                double res = 0.0;
                bool ok = true;
                for (;;)
                {
                    ++pass;
                    double S[15];
#pragma unroll
                    for (int q = 0; q < 15; ++q) S[q] = 0.0;
                    for (int k = lane; k < nodes; k += kWave)
                    {
                        const double xx = vnode(c.vlo, c.dv, k % Nv) / cs;
                        const double yy = vnode(c.vlo, c.dv, (k / Nv) % Nv) / cs;
                        const double zz = vnode(c.vlo, c.dv, k / (Nv * Nv)) / cs;
                        const double e = exp(a[0] + a[1] * xx + a[2] * yy + a[3] * zz
                                             + a[4] * (xx * xx + yy * yy + zz * zz)) * dvd;
                        dm_accum(S, e, xx, yy, zz);
                    }
                    wave_sums<15>(S);
                    double F[5];
                    res = 0.0;
#pragma unroll
                    for (int q = 0; q < 5; ++q) { F[q] = S[q] - ms[q]; res = fmax(res, fabs(F[q])); }
                    res /= fabs(ms[0]);
                    if (res <= dm.tol || pass > dm.maxit) break;
                    double d[5];
                    if (!dm_solve5(S, F, d)) { ok = false; break; }
#pragma unroll
                    for (int q = 0; q < 5; ++q) a[q] -= d[q];
                }
                bool conv = ok && res <= dm.tol;
#pragma unroll
                for (int q = 0; q < 5; ++q) conv = conv && isfinite(a[q]);
                if (!conv)
#pragma unroll
                    for (int q = 0; q < 5; ++q) a[q] = ac[q];       // continuous fallback
                if (lane == 0)
                {
#pragma unroll
                    for (int q = 0; q < 5; ++q) dm.alpha[(size_t)i * 5 + q] = a[q];
                    dm.passes[i] = conv ? pass : dm.maxit + 2;
                    dm.res[i] = res;
                    if (dm.rec)
                    {
                        float *row = dm.rec + (size_t)i * kMlpRow;
#pragma unroll
                        for (int q = 0; q < kMlpIn; ++q) row[q] = feat[q];
#pragma unroll
                        for (int q = 0; q < kMlpOut; ++q) row[kMlpIn + q] = float(a[q] - ac[q]);
                        row[kMlpIn + kMlpOut] = conv ? 1.0f : 0.0f;
                    }
                }
            }
        }
        for (int k = lane; k < nodes; k += kWave)
        {
            double vx = vnode(c.vlo, c.dv, k % Nv);
            double vy = vnode(c.vlo, c.dv, (k / Nv) % Nv);
            double f = doScale ? gp[k] * s1 : gp[k];
            if constexpr (DIM == 3)
            {
                double vz = vnode(c.vlo, c.dv, k / (Nv * Nv));
                double M;
                if (useDm)
                {
                    const double xx = vx / dm.cs, yy = vy / dm.cs, zz = vz / dm.cs;
                    M = exp(a[0] + a[1] * xx + a[2] * yy + a[3] * zz
                            + a[4] * (xx * xx + yy * yy + zz * zz));
                }
                else M = maxwellian3d(vx, vy, vz, mux, muy, muz, mT, mrho, R);
                f = coeff * (f + lam * M);
            }
            else
            {
                double f2 = doScale ? gp[k + nodes] * s1 : gp[k + nodes];
                double Mg1 = maxwellian_g1(vx, vy, mux, muy, mT, mrho, R);
                f = coeff * (f + lam * Mg1);
                f2 = coeff * (f2 + lam * 0.5 * R * mT * Mg1);
                gp[k + nodes] = f2;
            }
            gp[k] = f;
            m[0] += f * dvd;
        }
        if (lane == 0)
        {
            rho[i] = mrho; ux[i] = mux; uy[i] = muy; T[i] = mT;
            if (DIM == 3) uz[i] = muz;
        }
    }
    wave_sums<1>(m);
    if (lane == 0) pmass[i] = m[0];
}

// kPositivity + kCommit + kScaleG + kWallBC / kMoments + kCollisionNode of
// the 2-D step (CudaBackend.cu:304-336, 478-485, 338-389, 124-158,
// 1165-1179) in one pass per particle for the 2-D table formulation: gt is
// read (the row stays in cache between the passes over it) and the collided
// g written once, where k_limit_commit<2> and k_wall_collide<2> read and
// write g twice. The rescale between commit and collision is left out: the
// wall density, the density moment and the Maxwellians scale with g, u and
// T do not, so the collided row of the unscaled commit is the collided row
// of the rescaled one divided by massOld / massNew. The reduction then takes
// the committed masses (pnew) and the collided masses (pmass) and gives
// k_final the one factor that the rows need; the wall densities get the
// first factor there. Results agree with the separate kernels to rounding.
// U > 0: rows of at most U * 64 nodes (Nv <= 16 for U = 4), which each lane
// loads into registers once (U values of each layer); U == 0: any row, read
// again from memory in every pass. Both sum in the same order, so they give
// identical results.
template <int U>
__global__ __launch_bounds__(kBlock) void k_commit_collide2(DevCloud c, double *__restrict__ g,
                                                            const double *__restrict__ gt,
                                                            double dt, double tau, double R,
                                                            double *rho, double *ux, double *uy,
                                                            double *T, double *__restrict__ limflag,
                                                            double *__restrict__ pnew,
                                                            double *__restrict__ pmass)
{
    const int i = xcd_remap(blockIdx.x, gridDim.x) * (kBlock / kWave) + (int)(threadIdx.x / kWave);
    if (i >= c.N) return;                              // whole wave: one particle
    const int lane = threadIdx.x & (kWave - 1);
    const size_t gs = (size_t)c.gstride;
    double *gp = g + (size_t)i * gs;
    const double *tp = gt + (size_t)i * gs;
    const int nodes = c.nodes, Nv = c.Nv;
    const double dvd = c.dv * c.dv;
    const int nu = U > 0 ? U : (nodes + kWave - 1) / kWave;   // passes of a wave over the row

    double t1[U > 0 ? U : 1], t2[U > 0 ? U : 1];
    if constexpr (U > 0)
    {
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            const int k = lane + u * kWave;
            t1[u] = k < nodes ? tp[k] : 0.0;
            t2[u] = k < nodes ? tp[k + nodes] : 0.0;
        }
    }
    auto raw1 = [&](int u, int k) { if constexpr (U > 0) return t1[u]; else return tp[k]; };
    auto raw2 = [&](int u, int k) { if constexpr (U > 0) return t2[u]; else return tp[k + nodes]; };

    double rp[2] = {0.0, 0.0};                     // mPos, mNeg of g1
#pragma unroll
    for (int u = 0; u < nu; ++u)
    {
        const int k = lane + u * kWave;
        if (k >= nodes) continue;
        const double v = raw1(u, k);
        if (v >= 0) rp[0] += v; else rp[1] += v;
    }
    wave_sums<2>(rp);
    const bool fires = (rp[1] < -1e-30 && rp[0] > 1e-30 && (rp[0] + rp[1]) > 0);
    const double sc = fires ? (rp[0] + rp[1]) / rp[0] : 1.0;
    auto committed = [&](int u, int k, double &v1, double &v2) {
        v1 = raw1(u, k);
        v2 = raw2(u, k);
        if (fires)
        {
            if (v1 < 0) { v1 = 0; v2 = 0; }
            else { v1 *= sc; v2 *= sc; }
        }
    };
    double m[1] = {0.0};
    double mnew = 0.0;

    if (c.bnd[i])
    {
        const double nx = c.wnx[i], ny = c.wny[i];
        const double Uwx = c.wux[i], Uwy = c.wuy[i];
        const double Tw = c.wt[i];
        const double RT2w = 2.0 * R * Tw;
        const double norm = AB_PI * RT2w;
        double r[3] = {0.0, 0.0, 0.0};             // intM, intP, committed mass
#pragma unroll
        for (int u = 0; u < nu; ++u)
        {
            const int k = lane + u * kWave;
            if (k >= nodes) continue;
            const double vx = vnode(c.vlo, c.dv, k % Nv);
            const double vy = vnode(c.vlo, c.dv, (k / Nv) % Nv);
            double v, v2;
            committed(u, k, v, v2);
            const double cn = wall_cn2(vx, vy, Uwx, Uwy, nx, ny);
            if (cn < 0.0) r[0] += cn * v * dvd;
            else if (cn > 0.0)
            {
                const double e2 = (vx - Uwx) * (vx - Uwx) + (vy - Uwy) * (vy - Uwy);
                const double Mw = exp(-(e2) / RT2w) / norm;
                r[1] += cn * Mw * dvd;
            }
            r[2] += v * dvd;
        }
        wave_sums<3>(r);
        const double rho_w = (r[1] > 1e-30) ? (-r[0] / r[1]) : 1.0;
        mnew = r[2];
#pragma unroll
        for (int u = 0; u < nu; ++u)
        {
            const int k = lane + u * kWave;
            if (k >= nodes) continue;
            const double vx = vnode(c.vlo, c.dv, k % Nv);
            const double vy = vnode(c.vlo, c.dv, (k / Nv) % Nv);
            double v, v2;
            committed(u, k, v, v2);
            const double cn = wall_cn2(vx, vy, Uwx, Uwy, nx, ny);
            if (cn > 0.0)
            {
                const double e2 = (vx - Uwx) * (vx - Uwx) + (vy - Uwy) * (vy - Uwy);
                v = rho_w * exp(-(e2) / RT2w) / norm;
                v2 = 0.5 * R * Tw * v;
            }
            gp[k] = v;
            gp[k + nodes] = v2;
            m[0] += v * dvd;
        }
        if (lane == 0) { rho[i] = rho_w; ux[i] = Uwx; uy[i] = Uwy; T[i] = Tw; }
    }
    else
    {
        double r[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int u = 0; u < nu; ++u)
        {
            const int k = lane + u * kWave;
            if (k >= nodes) continue;
            const double vx = vnode(c.vlo, c.dv, k % Nv);
            const double vy = vnode(c.vlo, c.dv, (k / Nv) % Nv);
            double f, f2;
            committed(u, k, f, f2);
            r[0] += f * dvd;
            r[1] += vx * f * dvd;
            r[2] += vy * f * dvd;
            r[4] += (vx * vx + vy * vy) * f * dvd + 2.0 * f2 * dvd;
        }
        wave_sums<5>(r);
        mnew = r[0];
        double mrho = rho[i], mux = ux[i], muy = uy[i], mT = T[i];
        if (r[0] > 1e-30)
        {
            mrho = r[0];
            mux = r[1] / r[0];
            muy = r[2] / r[0];
            const double t1m = (r[4] / r[0] - (mux * mux + muy * muy)) / (3.0 * R);
            mT = t1m > 1e-6 ? t1m : 1e-6;
        }
        const double lam = dt / tau, coeff = 1.0 / (1.0 + lam);
#pragma unroll
        for (int u = 0; u < nu; ++u)
        {
            const int k = lane + u * kWave;
            if (k >= nodes) continue;
            const double vx = vnode(c.vlo, c.dv, k % Nv);
            const double vy = vnode(c.vlo, c.dv, (k / Nv) % Nv);
            double f, f2;
            committed(u, k, f, f2);
            const double Mg1 = maxwellian_g1(vx, vy, mux, muy, mT, mrho, R);
            f = coeff * (f + lam * Mg1);
            f2 = coeff * (f2 + lam * 0.5 * R * mT * Mg1);
            gp[k] = f;
            gp[k + nodes] = f2;
            m[0] += f * dvd;
        }
        if (lane == 0) { rho[i] = mrho; ux[i] = mux; uy[i] = muy; T[i] = mT; }
    }
    wave_sums<1>(m);
    if (lane == 0)
    {
        pmass[i] = m[0];
        pnew[i] = mnew;
        limflag[i] = fires ? 1.0 : 0.0;
    }
}

// kScaleG (massOld/massColl) + kMoments3D / kMoments (CudaBackend.cu:1773-1788).
// One wave per particle, four per workgroup: each lane loads four nodes of
// both layers before it adds any, and the sums are wave reductions, so a
// row's loads are in flight together and no workgroup barrier is needed.
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_final(DevCloud c, double *__restrict__ g,
                                                  const double *__restrict__ scal, double R,
                                                  double *rho, double *ux, double *uy,
                                                  double *uz, double *T,
                                                  double *__restrict__ pmass,
                                                  const int *__restrict__ guard,
                                                  bool fold_scale1)
{
    constexpr int kPer = kBlock / kWave;
    const int i = xcd_remap(blockIdx.x, gridDim.x) * kPer + (int)(threadIdx.x / kWave);
    if (i >= c.N) return;                          // whole wave
    if (guard && !*guard) return;
    const int lane = threadIdx.x & (kWave - 1);
    if (c.inst) scal += (size_t)c.inst[i] * S_COUNT;
    const double s2 = scal[S_SCALE2];
    const double s1 = scal[S_SCALE1];
    const bool doScale = (s2 != 1.0);
    const size_t gs = (size_t)c.gstride;
    double *gp = g + (size_t)i * gs;
    const int nodes = c.nodes, Nv = c.Nv;
    const double dvd = (DIM == 3) ? c.dv * c.dv * c.dv : c.dv * c.dv;
    const bool inner = !c.bnd[i];
    double r[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};   // rho, rux, ruy, ruz, E, mass
    auto node = [&](int k, double f, double f2) {
        if (doScale)
        {
            f *= s2; gp[k] = f;
            if constexpr (DIM == 2) { f2 *= s2; gp[k + nodes] = f2; }
        }
        r[5] += f * dvd;
        if (inner)
        {
            double vx = vnode(c.vlo, c.dv, k % Nv);
            double vy = vnode(c.vlo, c.dv, (k / Nv) % Nv);
            r[0] += f * dvd;
            r[1] += vx * f * dvd;
            r[2] += vy * f * dvd;
            if constexpr (DIM == 3)
            {
                double vz = vnode(c.vlo, c.dv, k / (Nv * Nv));
                r[3] += vz * f * dvd;
                r[4] += (vx * vx + vy * vy + vz * vz) * f * dvd;
            }
            else
                r[4] += (vx * vx + vy * vy) * f * dvd + 2.0 * f2 * dvd;
        }
    };
    constexpr int U = 4;
    int k0 = 0;
    for (; k0 + U * kWave <= nodes; k0 += U * kWave)
    {
        double f[U], f2[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            f[u] = gp[k0 + u * kWave + lane];
            f2[u] = (DIM == 2) ? gp[k0 + u * kWave + lane + nodes] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) node(k0 + u * kWave + lane, f[u], f2[u]);
    }
    for (int k = k0 + lane; k < nodes; k += kWave)
        node(k, gp[k], (DIM == 2) ? gp[k + nodes] : 0.0);
#pragma unroll
    for (int q = 0; q < 6; ++q) r[q] = wave_sum(r[q]);
    if (lane == 0)
    {
        pmass[i] = r[5];
        if (!inner && fold_scale1) rho[i] *= s1;
        if (inner && r[0] > 1e-30)
        {
            double mux = r[1] / r[0], muy = r[2] / r[0], muz = 0.0, u2;
            if constexpr (DIM == 3) { muz = r[3] / r[0]; u2 = mux * mux + muy * muy + muz * muz; }
            else u2 = mux * mux + muy * muy;
            double t1 = (r[4] / r[0] - u2) / (3.0 * R);
            rho[i] = r[0]; ux[i] = mux; uy[i] = muy; T[i] = t1 > 1e-6 ? t1 : 1e-6;
            if (DIM == 3) uz[i] = muz;
        }
    }
}

// ---- the reference's CPU-step limiter (MOOD) ---------------------------------

// The positivity repair inside the CPU transport (MLSTransport/MLSTransport.cpp
// :189-200 in 3-D, :287-307 in 2-D) on one particle's row: negatives to
// zero, the rest scaled to keep the mass (2-D: on g1, g2 following). Returns
// whether it fired (every thread).
template <int DIM>
__device__ __forceinline__ bool positivity_zero(const DevCloud &c, double *p, double *sh)
{
    const int n = c.nodes;
    double r[2] = {0.0, 0.0};                  // mPos, mNeg
    for (int k = threadIdx.x; k < n; k += kBlock)
    {
        const double v = p[k];
        if (v >= 0.0) r[0] += v; else r[1] += v;
    }
    block_sum<2>(r, sh);
    if (!(r[1] < -1e-30 && r[0] > 1e-30 && (r[0] + r[1]) > 0.0)) return false;
    const double sc = (r[0] + r[1]) / r[0];
    for (int k = threadIdx.x; k < n; k += kBlock)
    {
        const bool neg = p[k] < 0.0;
        p[k] = neg ? 0.0 : p[k] * sc;
        if constexpr (DIM == 2) p[k + n] = neg ? 0.0 : p[k + n] * sc;
    }
    __syncthreads();
    return true;
}

// Positivity on every particle's transported row, then checkTransportPAD
// (MLSTransport/MLSTransport.cpp:324-379) on the interior ones: flagged when
// the row's mass is not finite or not positive, or its density leaves the
// neighbours' current range widened by 0.1 of the range (at least 1e-12) plus
// 1e-9. mood[i] = flagged; limflag: +1 positivity, +2 flagged.
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_mood_check(DevCloud c, double *__restrict__ gt,
                                                       const double *__restrict__ rho,
                                                       int *__restrict__ mood,
                                                       double *__restrict__ limflag)
{
    __shared__ double sh[2 * (kBlock / kWave)];
    const int i = xcd_remap(blockIdx.x, gridDim.x);
    double *p = gt + (size_t)i * c.gstride;
    const bool fired = positivity_zero<DIM>(c, p, sh);
    bool flag = false;
    if (!c.bnd[i])
    {
        double m[1] = {0.0};
        for (int k = threadIdx.x; k < c.nodes; k += kBlock) m[0] += p[k];
        block_sum<1>(m, sh);
        const double dvN = (DIM == 3) ? c.dv * c.dv * c.dv : c.dv * c.dv;
        flag = !isfinite(m[0]) || m[0] <= 0.0;     // a non-finite value makes the sum non-finite
        const int nn = c.nneigh[i];
        double lo = 1e300, hi = -1e300;
        for (int k = threadIdx.x; k < nn; k += kBlock)
        {
            const double q = rho[c.neigh[(size_t)i * AB_MAX_NEIGH + k]];
            lo = fmin(lo, q); hi = fmax(hi, q);
        }
        block_minmax(lo, hi, sh);
        if (!flag && nn > 0)
        {
            const double rn = m[0] * dvN;
            const double slack = 0.1 * fmax(hi - lo, 1e-12) + 1e-9;
            flag = (rn < lo - slack || rn > hi + slack);
        }
    }
    if (threadIdx.x == 0)
    {
        mood[i] = flag ? 1 : 0;
        limflag[i] = (fired ? 1.0 : 0.0) + (flag ? 2.0 : 0.0);
    }
}

// The flagged particles' rows came from the linear re-transport: positivity
// once more on those (the repair runs inside every transport pass); then the
// commit gt -> g for every particle and its mass partial.
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_mood_commit(DevCloud c, double *__restrict__ g,
                                                        double *__restrict__ gt,
                                                        const int *__restrict__ mood,
                                                        double *__restrict__ limflag,
                                                        double *__restrict__ pmass)
{
    __shared__ double sh[2 * (kBlock / kWave)];
    const int i = xcd_remap(blockIdx.x, gridDim.x);
    const size_t gs = (size_t)c.gstride;
    double *tp = gt + (size_t)i * gs;
    double *gp = g + (size_t)i * gs;
    if (mood[i])
    {
        const bool fired = positivity_zero<DIM>(c, tp, sh);
        if (fired && threadIdx.x == 0 && !((int)limflag[i] & 1)) limflag[i] += 1.0;
    }
    const double dvd = (DIM == 3) ? c.dv * c.dv * c.dv : c.dv * c.dv;
    double m[1] = {0.0};
    for (int k = threadIdx.x; k < (int)gs; k += kBlock)
    {
        const double v = tp[k];
        gp[k] = v;
        if (k < c.nodes) m[0] += v * dvd;
    }
    block_sum<1>(m, sh);
    if (threadIdx.x == 0) pmass[i] = m[0];
}

// moodCheck (MomentCalculations/Conservation.cpp:16-38) with padCheck
// (MLSTransport/MLSBasis.hpp:208-222) on the interior particles after the
// collision: rho < 1e-10, T < 1e-6 or a mass-carrying value below -1e-14.
// limflag +4 and *fail set for a failing particle.
__global__ __launch_bounds__(kBlock) void k_pad_check(DevCloud c, const double *__restrict__ g,
                                                      const double *__restrict__ rho,
                                                      const double *__restrict__ T,
                                                      double *__restrict__ limflag,
                                                      int *__restrict__ fail)
{
    __shared__ int sbad[kBlock / kWave];
    const int i = xcd_remap(blockIdx.x, gridDim.x);
    if (c.bnd[i]) return;
    const double *p = g + (size_t)i * c.gstride;
    int bad = 0;
    for (int k = threadIdx.x; k < c.nodes; k += kBlock) bad |= (p[k] < -1.0e-14) ? 1 : 0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) bad |= __shfl_xor(bad, o, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) sbad[threadIdx.x / kWave] = bad;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        int b = (rho[i] < 1.0e-10 || T[i] < 1.0e-6) ? 1 : 0;
        for (int q = 0; q < kBlock / kWave; ++q) b |= sbad[q];
        if (b)
        {
            limflag[i] += 4.0;
            atomicOr(fail, 1);
        }
    }
}

}  // namespace

void launch_transport_fp64(const DevCloud &c, const double *g, double *gt, double dt, double h,
                           int order, double tol, int clamp, int *mlsflag, double *mlsovr,
                           const int *list, int nlist, bool boundary, hipStream_t s,
                           const int *only)
{
    if (nlist <= 0) return;
    const int nchunk = (c.nodes + kBlock - 1) / kBlock;
    const int nb = nlist * nchunk;
    if (c.dim == 3)
    {
        if (boundary)
            k_transport3d<4, true><<<nb, kBlock, 0, s>>>(c, g, gt, dt, h, tol, clamp, mlsflag,
                                                         mlsovr, list, nchunk, only);
        else if (order >= 2)
            k_transport3d<10, false><<<nb, kBlock, 0, s>>>(c, g, gt, dt, h, tol, clamp, mlsflag,
                                                           mlsovr, list, nchunk, only);
        else
            k_transport3d<4, false><<<nb, kBlock, 0, s>>>(c, g, gt, dt, h, tol, clamp, mlsflag,
                                                          mlsovr, list, nchunk, only);
    }
    else
    {
        if (boundary)
            k_transport2d<3, true><<<nb, kBlock, 0, s>>>(c, g, gt, dt, h, list, nchunk, only);
        else if (order >= 2)
            k_transport2d<6, false><<<nb, kBlock, 0, s>>>(c, g, gt, dt, h, list, nchunk, only);
        else
            k_transport2d<3, false><<<nb, kBlock, 0, s>>>(c, g, gt, dt, h, list, nchunk, only);
    }
}

void launch_mass_partials(const DevCloud &c, const double *g, double *pmass, hipStream_t s)
{
    k_mass_partials<<<c.N, kBlock, 0, s>>>(c, g, pmass);
}

void launch_reduce_mass(int N, const double *pmass, double *scal, int mode, int nofix,
                        hipStream_t s, const int *seg, int nseg, const double *pnew)
{
    k_reduce_mass<<<seg ? nseg : 1, 1024, 0, s>>>(N, pmass, scal, mode, nofix, seg, pnew);
}

void launch_limit_commit(const DevCloud &c, double *g, const double *gt, const double *rho,
                         double *limflag, double *pmass, hipStream_t s)
{
    const int nb = (c.N + kBlock / kWave - 1) / (kBlock / kWave);
    if (c.dim == 3) k_limit_commit<3><<<nb, kBlock, 0, s>>>(c, g, gt, rho, limflag, pmass);
    else k_limit_commit<2><<<nb, kBlock, 0, s>>>(c, g, gt, rho, limflag, pmass);
}

void launch_wall_collide(const DevCloud &c, const DmArgs &dm, double *g, const double *scal,
                         double dt, double tau, double R, double *rho, double *ux, double *uy,
                         double *uz, double *T, double *pmass, hipStream_t s, const int *guard)
{
    const int nb = (c.N + kBlock / kWave - 1) / (kBlock / kWave);
    if (c.dim == 3)
        k_wall_collide<3><<<nb, kBlock, 0, s>>>(c, dm, g, scal, dt, tau, R, rho, ux, uy, uz, T,
                                                pmass, guard);
    else
        k_wall_collide<2><<<nb, kBlock, 0, s>>>(c, dm, g, scal, dt, tau, R, rho, ux, uy, uz, T,
                                                pmass, guard);
}

void launch_final(const DevCloud &c, double *g, const double *scal, double R, double *rho,
                  double *ux, double *uy, double *uz, double *T, double *pmass, hipStream_t s,
                  const int *guard, bool fold_scale1)
{
    const int nb = (c.N + kBlock / kWave - 1) / (kBlock / kWave);
    if (c.dim == 3)
        k_final<3><<<nb, kBlock, 0, s>>>(c, g, scal, R, rho, ux, uy, uz, T, pmass, guard,
                                         fold_scale1);
    else
        k_final<2><<<nb, kBlock, 0, s>>>(c, g, scal, R, rho, ux, uy, uz, T, pmass, guard,
                                         fold_scale1);
}

void launch_commit_collide2(const DevCloud &c, double *g, const double *gt, double dt,
                            double tau, double R, double *rho, double *ux, double *uy,
                            double *T, double *limflag, double *pnew, double *pmass,
                            hipStream_t s)
{
    const int nb = (c.N + kBlock / kWave - 1) / (kBlock / kWave);
    // rows of up to 256 nodes (Nv <= 16) stay in registers
    const int u = (c.nodes + kWave - 1) / kWave;
    if (u == 1)
        k_commit_collide2<1><<<nb, kBlock, 0, s>>>(c, g, gt, dt, tau, R, rho, ux, uy, T, limflag,
                                                   pnew, pmass);
    else if (u == 2)
        k_commit_collide2<2><<<nb, kBlock, 0, s>>>(c, g, gt, dt, tau, R, rho, ux, uy, T, limflag,
                                                   pnew, pmass);
    else if (u <= 4)
        k_commit_collide2<4><<<nb, kBlock, 0, s>>>(c, g, gt, dt, tau, R, rho, ux, uy, T, limflag,
                                                   pnew, pmass);
    else
        k_commit_collide2<0><<<nb, kBlock, 0, s>>>(c, g, gt, dt, tau, R, rho, ux, uy, T, limflag,
                                                   pnew, pmass);
}

void launch_mood_check(const DevCloud &c, double *gt, const double *rho, int *mood,
                       double *limflag, hipStream_t s)
{
    if (c.dim == 3) k_mood_check<3><<<c.N, kBlock, 0, s>>>(c, gt, rho, mood, limflag);
    else k_mood_check<2><<<c.N, kBlock, 0, s>>>(c, gt, rho, mood, limflag);
}

void launch_mood_commit(const DevCloud &c, double *g, double *gt, const int *mood,
                        double *limflag, double *pmass, hipStream_t s)
{
    if (c.dim == 3) k_mood_commit<3><<<c.N, kBlock, 0, s>>>(c, g, gt, mood, limflag, pmass);
    else k_mood_commit<2><<<c.N, kBlock, 0, s>>>(c, g, gt, mood, limflag, pmass);
}

void launch_pad_check(const DevCloud &c, const double *g, const double *rho, const double *T,
                      double *limflag, int *fail, hipStream_t s)
{
    k_pad_check<<<c.N, kBlock, 0, s>>>(c, g, rho, T, limflag, fail);
}

}  // namespace ab
