// k_body.hip -- rigid bodies on the device: the surface-stress loads
// (computeBodyLoads, RigidBody/BodyDynamics.cpp:17-97) and the moving-body
// block of a step (moveRigidBodies, :143-711, then cloud management,
// Solver/Solver.cpp:348-444).
//
// Loads. The reference integrates P_ab = sum f (v_a - U_a)(v_b - U_b) dv^3
// for every trial wall velocity U, which the Heun coupling evaluates twice
// per step on the same g (BodyDynamics.cpp:353, :399). The ten raw moments
//   m0 = sum f dv^3,  m_a = sum f v_a dv^3,  m_ab = sum f v_a v_b dv^3
// give P(U) = m_ab - U_a m_b - U_b m_a + U_a U_b m0 in closed form, so one
// pass over the surface particles' g serves every U: one wave per surface
// particle, FP64 accumulation. Each surface particle's traction then gives
// its force and torque contributions, and one workgroup per body sums them
// in a fixed order; only 6 values per body leave the device.
//
// Motion. The body ODE is a handful of FP64 operations per body and runs on
// the host between the device passes, operation for operation as the
// reference writes it; the per-particle work (rigid motion of the surface,
// the contact shift, the refreshed normals and wall velocities) runs on the
// device with the per-body constants (cos, sin, axis) computed on the host,
// so positions round exactly as the reference's.
#include "ab_device.hpp"
#include "ab_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace ab {

namespace {

// Per-body parameters of the load kernels: centre (the torque arm's
// origin), surface area per particle, and the rigid velocity field
// U = v + w x r of umode 2.
struct LoadBody
{
    double c[3];
    double area;
    double v[3];
    double w[3];
};

// Per-body constants of the rigid surface motion (BodyDynamics.cpp:405-440):
// 3-D: Rodrigues rotation by |w| dt about w / |w| (rot = 0 when |w| is
// negligible); 2-D: planar rotation by angularVel dt.
struct MoveBody
{
    double c[3], v[3];
    int rot;
    double a[3], cs, sn, omc;
};

// The state the surface wall state is refreshed from (:615-638).
struct ShapeBody
{
    int shape;
    double c[3], size, o[2], rm[9];
    double v[3], om[3], angv;
};

constexpr int kParDoubles = (int)(sizeof(ShapeBody) / sizeof(double)) + 1;
static_assert(sizeof(LoadBody) <= kParDoubles * sizeof(double), "parameter slot");
static_assert(sizeof(MoveBody) <= kParDoubles * sizeof(double), "parameter slot");

// Stable list of the body-surface particles (boundary, body id >= 0) in
// particle order; one workgroup of 1024 threads over contiguous ranges.
// With idx == nullptr only the counts are written: count[0] the surface
// particles, count[1] those whose body id is nb or more (no such body).
__global__ __launch_bounds__(1024) void k_surface_list(int n, int nb,
                                                       const unsigned char *__restrict__ bnd,
                                                       const int *__restrict__ bid,
                                                       int *__restrict__ idx,
                                                       int *__restrict__ body,
                                                       int *__restrict__ count)
{
    __shared__ int sh[1024];
    __shared__ int bad;
    const int t = threadIdx.x;
    if (t == 0) bad = 0;
    const int per = (n + 1023) / 1024, lo = min(t * per, n), hi = min(lo + per, n);
    int cnt = 0, nbad = 0;
    for (int i = lo; i < hi; ++i)
    {
        cnt += (bnd[i] && bid[i] >= 0);
        nbad += (bnd[i] && bid[i] >= nb);
    }
    __syncthreads();
    if (nbad) atomicAdd(&bad, nbad);
    sh[t] = cnt;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1)
    {
        const int v = t >= o ? sh[t - o] : 0;
        __syncthreads();
        sh[t] += v;
        __syncthreads();
    }
    if (t == 1023) { count[0] = sh[1023]; count[1] = bad; }
    if (!idx) return;
    int pos = sh[t] - cnt;
    for (int i = lo; i < hi; ++i)
        if (bnd[i] && bid[i] >= 0)
        {
            idx[pos] = i;
            body[pos] = bid[i];
            ++pos;
        }
}

// Raw moments of the listed particles. Table storage holds g = L (M0 + d);
// m0 == nullptr means g is stored as is.
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

// Force and torque contributions of each listed surface particle
// (BodyDynamics.cpp:41-94) for a wall velocity U taken from the cloud's wall
// state (umode 0), from wu (umode 1), or from the rigid motion of its body
// (umode 2, setSurfaceVelocity :103-125).
__global__ void k_surface_loads(DevCloud c, int n, const int *__restrict__ idx,
                                const int *__restrict__ body, const double *__restrict__ mom,
                                const double *__restrict__ wu, int umode,
                                const LoadBody *__restrict__ lb, double *__restrict__ con)
{
#pragma clang fp contract(off)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = idx[k];
    const LoadBody B = lb[body[k]];
    const double rx = c.x[i] - B.c[0], ry = c.y[i] - B.c[1], rz = c.z[i] - B.c[2];
    double Ux, Uy, Uz;
    if (umode == 0) { Ux = c.wux[i]; Uy = c.wuy[i]; Uz = c.wuz[i]; }
    else if (umode == 1) { Ux = wu[3 * k]; Uy = wu[3 * k + 1]; Uz = wu[3 * k + 2]; }
    else
    {
        Ux = B.v[0] + B.w[1] * rz - B.w[2] * ry;
        Uy = B.v[1] + B.w[2] * rx - B.w[0] * rz;
        Uz = B.v[2] + B.w[0] * ry - B.w[1] * rx;
    }
    const double *q = mom + (size_t)k * 10;
    // P_ab = m_ab - U_a m_b - U_b m_a + U_a U_b m0
    const double pxx = q[4] - 2.0 * Ux * q[1] + Ux * Ux * q[0];
    const double pyy = q[5] - 2.0 * Uy * q[2] + Uy * Uy * q[0];
    const double pxy = q[7] - Ux * q[2] - Uy * q[1] + Ux * Uy * q[0];
    const double nx = c.wnx[i], ny = c.wny[i], nz = c.wnz[i];
    const double A = B.area;
    double *o = con + (size_t)k * 6;
    if (c.dim == 3)
    {
        const double pzz = q[6] - 2.0 * Uz * q[3] + Uz * Uz * q[0];
        const double pxz = q[8] - Ux * q[3] - Uz * q[1] + Ux * Uz * q[0];
        const double pyz = q[9] - Uy * q[3] - Uz * q[2] + Uy * Uz * q[0];
        const double tx = -(nx * pxx + ny * pxy + nz * pxz);
        const double ty = -(nx * pxy + ny * pyy + nz * pyz);
        const double tz = -(nx * pxz + ny * pyz + nz * pzz);
        o[0] = A * tx; o[1] = A * ty; o[2] = A * tz;
        o[3] = A * (ry * tz - rz * ty);
        o[4] = A * (rz * tx - rx * tz);
        o[5] = A * (rx * ty - ry * tx);
    }
    else
    {
        const double vx = -nx * pxx - ny * pxy;
        const double vy = -nx * pxy - ny * pyy;
        o[0] = A * vx; o[1] = A * vy; o[2] = 0.0;
        o[3] = 0.0; o[4] = 0.0;
        o[5] = A * (rx * vy - ry * vx);
    }
}

// One workgroup per body: the sum of its surface particles' contributions,
// each thread over a strided share of the list, then a fixed tree.
__global__ __launch_bounds__(256) void k_body_sum(int n, const int *__restrict__ body,
                                                  const double *__restrict__ con,
                                                  double *__restrict__ out)
{
    __shared__ double sh[6][256];
    const int b = blockIdx.x, t = threadIdx.x;
    double a[6] = {0, 0, 0, 0, 0, 0};
    for (int k = t; k < n; k += 256)
        if (body[k] == b)
#pragma unroll
            for (int q = 0; q < 6; ++q) a[q] += con[(size_t)k * 6 + q];
#pragma unroll
    for (int q = 0; q < 6; ++q) sh[q][t] = a[q];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1)
    {
        if (t < o)
#pragma unroll
            for (int q = 0; q < 6; ++q) sh[q][t] += sh[q][t + o];
        __syncthreads();
    }
    if (t < 6) out[6 * b + t] = sh[t][0];
}

// Rigid motion of the surface with the body's state before the update
// (BodyDynamics.cpp:405-440).
__global__ void k_surface_move(int dim, int n, const int *__restrict__ idx,
                               const int *__restrict__ body, double dt,
                               const MoveBody *__restrict__ mb, double *x, double *y, double *z)
{
#pragma clang fp contract(off)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = idx[k];
    const MoveBody B = mb[body[k]];
    const double rx = x[i] - B.c[0], ry = y[i] - B.c[1], rz = z[i] - B.c[2];
    if (dim == 3)
    {
        double nxr = rx, nyr = ry, nzr = rz;
        if (B.rot)
        {
            const double ax = B.a[0], ay = B.a[1], az = B.a[2];
            const double adotr = ax * rx + ay * ry + az * rz;
            const double crx = ay * rz - az * ry;
            const double cry = az * rx - ax * rz;
            const double crz = ax * ry - ay * rx;
            nxr = rx * B.cs + crx * B.sn + ax * adotr * B.omc;
            nyr = ry * B.cs + cry * B.sn + ay * adotr * B.omc;
            nzr = rz * B.cs + crz * B.sn + az * adotr * B.omc;
        }
        x[i] = B.c[0] + dt * B.v[0] + nxr;
        y[i] = B.c[1] + dt * B.v[1] + nyr;
        z[i] = B.c[2] + dt * B.v[2] + nzr;
    }
    else
    {
        x[i] = B.c[0] + dt * B.v[0] + B.cs * rx - B.sn * ry;
        y[i] = B.c[1] + dt * B.v[1] + B.sn * rx + B.cs * ry;
    }
}

// Extent of each body's surface (BodyDynamics.cpp:560-573): one workgroup
// per body; out = xlo, xhi, ylo, yhi, zlo, zhi, count.
__global__ __launch_bounds__(256) void k_body_extent(int n, const int *__restrict__ idx,
                                                     const int *__restrict__ body,
                                                     const double *__restrict__ x,
                                                     const double *__restrict__ y,
                                                     const double *__restrict__ z,
                                                     double *__restrict__ out)
{
    __shared__ double sh[7][256];
    const int b = blockIdx.x, t = threadIdx.x;
    double e[7] = {1e300, -1e300, 1e300, -1e300, 1e300, -1e300, 0.0};
    for (int k = t; k < n; k += 256)
        if (body[k] == b)
        {
            const int i = idx[k];
            e[0] = fmin(e[0], x[i]); e[1] = fmax(e[1], x[i]);
            e[2] = fmin(e[2], y[i]); e[3] = fmax(e[3], y[i]);
            e[4] = fmin(e[4], z[i]); e[5] = fmax(e[5], z[i]);
            e[6] += 1.0;
        }
#pragma unroll
    for (int q = 0; q < 7; ++q) sh[q][t] = e[q];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1)
    {
        if (t < o)
        {
            for (int q = 0; q < 6; q += 2)
            {
                sh[q][t] = fmin(sh[q][t], sh[q][t + o]);
                sh[q + 1][t] = fmax(sh[q + 1][t], sh[q + 1][t + o]);
            }
            sh[6][t] += sh[6][t + o];
        }
        __syncthreads();
    }
    if (t < 7) out[7 * b + t] = sh[t][0];
}

// The contact's rigid shift of a body's surface (BodyDynamics.cpp:588-594);
// shift = 3 per body, applied only to bodies flagged in shift[3 nb + b].
__global__ void k_surface_shift(int n, int nb, const int *__restrict__ idx,
                                const int *__restrict__ body, const double *__restrict__ shift,
                                double *x, double *y, double *z)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int b = body[k];
    if (shift[3 * nb + b] == 0.0) return;
    const int i = idx[k];
    x[i] += shift[3 * b];
    y[i] += shift[3 * b + 1];
    z[i] += shift[3 * b + 2];
}

// getNormal / getNormal3 of the four shapes (Geometry.hpp:116-122,
// :148-164, :225-231, :298-310).
__device__ void shape_normal(const ShapeBody &b, double x, double y, double z, double n[3])
{
#pragma clang fp contract(off)
    switch (b.shape)
    {
    case AB_BODY_SQUARE:
    {
        const double ox = b.o[0], oy = b.o[1];
        const double xp = x - b.c[0], yp = y - b.c[1];
        const double nn = sqrt(xp * xp + yp * yp);
        const double d[4] = {(xp * ox + yp * oy) / nn, -(xp * ox + yp * oy) / nn,
                             (xp * oy - yp * ox) / nn, -(xp * oy - yp * ox) / nn};
        const double vx[4] = {ox, -ox, oy, -oy}, vy[4] = {oy, -oy, -ox, ox};
        int best = 0;
        for (int q = 1; q < 4; ++q)
            if (d[q] > d[best]) best = q;
        n[0] = vx[best]; n[1] = vy[best]; n[2] = 0.0;
        return;
    }
    case AB_BODY_CIRCLE:
    {
        const double dx = x - b.c[0], dy = y - b.c[1];
        const double nn = sqrt(dx * dx + dy * dy);
        n[0] = dx / nn; n[1] = dy / nn; n[2] = 0.0;
        return;
    }
    case AB_BODY_SPHERE:
    {
        const double dx = x - b.c[0], dy = y - b.c[1], dz = z - b.c[2];
        const double nn = sqrt(dx * dx + dy * dy + dz * dz);
        n[0] = dx / nn; n[1] = dy / nn; n[2] = dz / nn;
        return;
    }
    default:
    {
        const double dx = x - b.c[0], dy = y - b.c[1], dz = z - b.c[2];
        const double *R = b.rm;
        const double o[3] = {R[0] * dx + R[3] * dy + R[6] * dz, R[1] * dx + R[4] * dy + R[7] * dz,
                             R[2] * dx + R[5] * dy + R[8] * dz};
        int f = 0;
        for (int q = 1; q < 3; ++q)
            if (fabs(o[q]) > fabs(o[f])) f = q;
        const double s = (o[f] >= 0.0) ? 1.0 : -1.0;
        n[0] = s * R[f]; n[1] = s * R[3 + f]; n[2] = s * R[6 + f];
        return;
    }
    }
}

// The surface wall state from the new body state (BodyDynamics.cpp:615-638):
// normal, wall velocity v + w x r, and the mesh velocity that moves the
// transport's foot points with the surface.
__global__ void k_surface_refresh(DevCloud c, int n, const int *__restrict__ idx,
                                  const int *__restrict__ body, const ShapeBody *__restrict__ sb,
                                  double *wnx, double *wny, double *wnz, double *wux, double *wuy,
                                  double *wuz, double *mvx, double *mvy, double *mvz)
{
#pragma clang fp contract(off)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i = idx[k];
    const ShapeBody &B = sb[body[k]];
    const double px = c.x[i], py = c.y[i], pz = c.z[i];
    const double rx = px - B.c[0], ry = py - B.c[1], rz = pz - B.c[2];
    double nn[3];
    shape_normal(B, px, py, pz, nn);
    if (c.dim == 3)
    {
        wnx[i] = nn[0]; wny[i] = nn[1]; wnz[i] = nn[2];
        const double ux = B.v[0] + B.om[1] * rz - B.om[2] * ry;
        const double uy = B.v[1] + B.om[2] * rx - B.om[0] * rz;
        const double uz = B.v[2] + B.om[0] * ry - B.om[1] * rx;
        wux[i] = ux; wuy[i] = uy; wuz[i] = uz;
        mvx[i] = ux; mvy[i] = uy; mvz[i] = uz;
    }
    else
    {
        wnx[i] = nn[0]; wny[i] = nn[1];
        const double ux = B.v[0] - B.angv * ry;
        const double uy = B.v[1] + B.angv * rx;
        wux[i] = ux; wuy[i] = uy;
        mvx[i] = ux; mvy[i] = uy;
    }
}

#define BCHK(c, expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) {                                                         \
            (c)->err = std::string("body: ") + #expr + ": " + hipGetErrorString(e_);    \
            return AB_ERR_HIP;                                                          \
        }                                                                               \
    } while (0)

// Scratch of the body passes for ns surface particles (grown, kept).
int body_scratch(ab_ctx *c, int ns)
{
    if (!c->bpar)
    {
        if (hipMalloc((void **)&c->bpar, sizeof(double) * kParDoubles * kMaxBodies) != hipSuccess ||
            hipMalloc((void **)&c->bout, sizeof(double) * 8 * kMaxBodies) != hipSuccess ||
            hipMalloc((void **)&c->bcnt, 2 * sizeof(int)) != hipSuccess)
        { c->err = "body scratch: device allocation failed"; return AB_ERR_NOMEM; }
    }
    if (ns <= c->bmom_cap) return AB_OK;
    void *old[] = {c->bidx, c->bbody, c->bmom, c->bcon, c->bwu};
    for (void *p : old) if (p) (void)hipFree(p);
    c->bidx = c->bbody = nullptr;
    c->bmom = c->bcon = c->bwu = nullptr;
    c->bmom_cap = 0;
    const size_t n = (size_t)std::max(ns, 64);
    if (hipMalloc((void **)&c->bidx, sizeof(int) * n) != hipSuccess ||
        hipMalloc((void **)&c->bbody, sizeof(int) * n) != hipSuccess ||
        hipMalloc((void **)&c->bmom, sizeof(double) * 10 * n) != hipSuccess ||
        hipMalloc((void **)&c->bcon, sizeof(double) * 6 * n) != hipSuccess ||
        hipMalloc((void **)&c->bwu, sizeof(double) * 3 * n) != hipSuccess)
    { c->err = "body scratch: device allocation failed"; return AB_ERR_NOMEM; }
    c->bmom_cap = (int)n;
    return AB_OK;
}

void launch_moments(ab_ctx *c, const int *idx, int n, double *mom)
{
    if (n <= 0) return;
    const int nb = (n * kWave + 255) / 256;
    DevCloud v = c->view();
    const bool tab = c->cfg.precision != AB_PREC_FP64 && c->dim == 3;
    if (!tab)
        k_body_moments<double><<<nb, 256, 0, c->stream>>>(v, c->g, c->gstride, nullptr, nullptr,
                                                          idx, n, mom);
    else if (c->cfg.precision == AB_PREC_MIXED)
        k_body_moments<float><<<nb, 256, 0, c->stream>>>(v, (const float *)c->gs, c->S,
                                                         c->scal + S_LAZY, c->m0, idx, n, mom);
    else
        k_body_moments<double><<<nb, 256, 0, c->stream>>>(v, (const double *)c->gs, c->S,
                                                          c->scal + S_LAZY, c->m0, idx, n, mom);
}

// Loads of nb bodies from the moments of n listed surface particles:
// contributions, per-body sums, and the 6 nb results copied to host out.
int loads(ab_ctx *c, const int *idx, const int *body, int n, const double *mom,
          const double *wu, int umode, const LoadBody *lb, int nb, double *out, double *con,
          double *par, double *dout)
{
    BCHK(c, hipMemcpyAsync(par, lb, sizeof(LoadBody) * nb, hipMemcpyHostToDevice, c->stream));
    if (n > 0)
        k_surface_loads<<<(n + 255) / 256, 256, 0, c->stream>>>(c->view(), n, idx, body, mom, wu,
                                                                umode, (const LoadBody *)par, con);
    k_body_sum<<<nb, 256, 0, c->stream>>>(n, body, con, dout);
    BCHK(c, hipGetLastError());
    BCHK(c, hipMemcpyAsync(out, dout, sizeof(double) * 6 * nb, hipMemcpyDeviceToHost, c->stream));
    BCHK(c, hipStreamSynchronize(c->stream));
    return AB_OK;
}

// RigidBody::getMomentOfInertia and getSurfaceArea of the four shapes
// (Geometry.hpp:95-97, :136-137, :216-217, :271-273).
double inertia(const ab_rigid_body &b)
{
    switch (b.shape)
    {
    case AB_BODY_SPHERE: return 0.4 * b.mass * b.size * b.size;
    case AB_BODY_CIRCLE: return b.mass * b.size * b.size / 2.0;
    default: return b.mass * b.size * b.size / 6.0;     // square, cube
    }
}

double area(const ab_rigid_body &b)
{
    switch (b.shape)
    {
    case AB_BODY_SPHERE: return 4.0 * AB_PI * b.size * b.size / b.npts;
    case AB_BODY_SQUARE: return 4.0 * b.size / b.npts;
    case AB_BODY_CIRCLE: return 2.0 * AB_PI * b.size / b.npts;
    default: return 6.0 * b.size * b.size / b.npts;
    }
}

// Cube::rotate3 (Geometry.hpp:335-360): Rm <- D(w dt) Rm, re-orthonormalised.
void rotate3(double *Rm, double wx, double wy, double wz, double dt)
{
    const double w = std::sqrt(wx * wx + wy * wy + wz * wz);
    if (w < 1e-300) return;
    const double ax = wx / w, ay = wy / w, az = wz / w;
    const double th = w * dt, c = std::cos(th), s = std::sin(th), C = 1.0 - c;
    const double D[9] = {c + ax * ax * C,      ax * ay * C - az * s, ax * az * C + ay * s,
                         ay * ax * C + az * s, c + ay * ay * C,      ay * az * C - ax * s,
                         az * ax * C - ay * s, az * ay * C + ax * s, c + az * az * C};
    double Rn[9];
    for (int r = 0; r < 3; ++r)
        for (int col = 0; col < 3; ++col)
            Rn[3 * r + col] = D[3 * r + 0] * Rm[0 + col] + D[3 * r + 1] * Rm[3 + col] +
                              D[3 * r + 2] * Rm[6 + col];
    auto nrm = [&](int r) {
        const double m = std::sqrt(Rn[3 * r] * Rn[3 * r] + Rn[3 * r + 1] * Rn[3 * r + 1] +
                                   Rn[3 * r + 2] * Rn[3 * r + 2]);
        for (int k = 0; k < 3; ++k) Rn[3 * r + k] /= m;
    };
    nrm(0);
    const double d = Rn[0] * Rn[3] + Rn[1] * Rn[4] + Rn[2] * Rn[5];
    for (int k = 0; k < 3; ++k) Rn[3 + k] -= d * Rn[k];
    nrm(1);
    Rn[6] = Rn[1] * Rn[5] - Rn[2] * Rn[4];
    Rn[7] = Rn[2] * Rn[3] - Rn[0] * Rn[5];
    Rn[8] = Rn[0] * Rn[4] - Rn[1] * Rn[3];
    for (int k = 0; k < 9; ++k) Rm[k] = Rn[k];
}

}  // namespace

}  // namespace ab

using namespace ab;

namespace ab {
namespace {

// Device buffers of one ab_body_loads call.
struct LoadBufs
{
    int *idx, *body;
    double *wu, *mom, *con, *par, *out;
};

int check_loads_args(ab_ctx *c, int nsurf, const int *surf_idx, const int *surf_body,
                     int nbodies, const double *body_c, const double *body_area, double *out,
                     const char *who)
{
    if (!c || nsurf < 0 || nbodies <= 0 || nbodies > kMaxBodies || !body_c || !body_area ||
        !out || (nsurf > 0 && (!surf_idx || !surf_body)))
    {
        if (c) c->err = std::string(who) + ": invalid argument (at most 16 bodies)";
        return AB_ERR_ARG;
    }
    if (!c->uploaded) { c->err = "no cloud uploaded"; return AB_ERR_STATE; }
    for (int k = 0; k < nsurf; ++k)
        if (surf_idx[k] < 0 || surf_idx[k] >= c->N || surf_body[k] < 0 || surf_body[k] >= nbodies)
        {
            c->err = std::string(who) + ": surface index or body id out of range";
            return AB_ERR_ARG;
        }
    return AB_OK;
}

// The loads of ab_body_loads on the given buffers: the surface list (and
// the wall velocities) copied in, moments, contributions, per-body sums.
// Every buffer is written before it is read.
int body_loads_on(ab_ctx *c, const LoadBufs &B, int nsurf, const int *surf_idx,
                  const int *surf_body, int nbodies, const double *body_c,
                  const double *body_area, const double *wu, double *out)
{
    std::memset(out, 0, sizeof(double) * 6 * (size_t)nbodies);
    hipStream_t s = c->stream;
    if (nsurf > 0)
    {
        BCHK(c, hipMemcpyAsync(B.idx, surf_idx, sizeof(int) * nsurf, hipMemcpyHostToDevice, s));
        BCHK(c, hipMemcpyAsync(B.body, surf_body, sizeof(int) * nsurf, hipMemcpyHostToDevice, s));
        if (wu)
            BCHK(c, hipMemcpyAsync(B.wu, wu, sizeof(double) * 3 * nsurf, hipMemcpyHostToDevice, s));
    }
    launch_moments(c, B.idx, nsurf, B.mom);
    LoadBody lb[kMaxBodies];
    std::memset(lb, 0, sizeof lb);
    for (int b = 0; b < nbodies; ++b)
    {
        for (int a = 0; a < 3; ++a) lb[b].c[a] = body_c[3 * b + a];
        lb[b].area = body_area[b];
    }
    return loads(c, B.idx, B.body, nsurf, B.mom, wu ? B.wu : nullptr, wu ? 1 : 0, lb, nbodies,
                 out, B.con, B.par, B.out);
}

// ---- scratch of the no-allocation variant ----------------------------------
// One array per device, for up to kStaticSurf surface particles: indices,
// body ids, wall velocities (3), moments (10), contributions (6), per-body
// parameters and results. The library never allocates it; each device
// holds its own copy, reached through hipGetSymbolAddress.
constexpr int kStaticSurf = 8192;
constexpr size_t kOffIdx = 0;
constexpr size_t kOffBody = kOffIdx + sizeof(int) * kStaticSurf;
constexpr size_t kOffWu = kOffBody + sizeof(int) * kStaticSurf;
constexpr size_t kOffMom = kOffWu + sizeof(double) * 3 * kStaticSurf;
constexpr size_t kOffCon = kOffMom + sizeof(double) * 10 * kStaticSurf;
constexpr size_t kOffPar = kOffCon + sizeof(double) * 6 * kStaticSurf;
constexpr size_t kOffOut = kOffPar + sizeof(double) * kParDoubles * kMaxBodies;
constexpr size_t kScratchBytes = kOffOut + sizeof(double) * 8 * kMaxBodies;

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch[kMaxDevices] = {};

}  // namespace
}  // namespace ab

static __device__ __align__(256) unsigned char scratch[ab::kScratchBytes];

namespace ab {
namespace {

// The device address of this device's scratch, looked up once per device
// with that device current (the array's name on the host is no device
// address).
int static_scratch(ab_ctx *c, unsigned char **out)
{
    if (c->device < 0 || c->device >= kMaxDevices)
    { c->err = "static scratch: device index out of range"; return AB_ERR_UNSUPPORTED; }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[c->device])
    {
        int cur = -1;
        BCHK(c, hipGetDevice(&cur));
        if (cur != c->device) BCHK(c, hipSetDevice(c->device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != c->device) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
        {
            c->err = std::string("static scratch: hipGetSymbolAddress: ") + hipGetErrorString(e);
            return AB_ERR_HIP;
        }
        g_scratch[c->device] = p;
    }
    *out = (unsigned char *)g_scratch[c->device];
    return AB_OK;
}

}  // namespace
}  // namespace ab

extern "C" int ab_body_loads(ab_ctx *c, int nsurf, const int *surf_idx, const int *surf_body,
                             int nbodies, const double *body_c, const double *body_area,
                             const double *wu, double *out)
{
    int rc = check_loads_args(c, nsurf, surf_idx, surf_body, nbodies, body_c, body_area, out,
                              "ab_body_loads");
    if (rc) return rc;
    if ((rc = body_scratch(c, nsurf))) return rc;
    const LoadBufs B{c->bidx, c->bbody, c->bwu, c->bmom, c->bcon, c->bpar, c->bout};
    return body_loads_on(c, B, nsurf, surf_idx, surf_body, nbodies, body_c, body_area, wu, out);
}

extern "C" int ab_body_loads_static(ab_ctx *c, int nsurf, const int *surf_idx,
                                    const int *surf_body, int nbodies, const double *body_c,
                                    const double *body_area, const double *wu, double *out)
{
    int rc = check_loads_args(c, nsurf, surf_idx, surf_body, nbodies, body_c, body_area, out,
                              "ab_body_loads_static");
    if (rc) return rc;
    if (nsurf > kStaticSurf)
    {
        c->err = "ab_body_loads_static: " + std::to_string(nsurf) +
                 " surface particles exceed its scratch of " + std::to_string(kStaticSurf) +
                 " (use ab_body_loads)";
        return AB_ERR_ARG;
    }
    unsigned char *p = nullptr;
    if ((rc = static_scratch(c, &p))) return rc;
    const LoadBufs B{(int *)(p + kOffIdx), (int *)(p + kOffBody), (double *)(p + kOffWu),
                     (double *)(p + kOffMom), (double *)(p + kOffCon), (double *)(p + kOffPar),
                     (double *)(p + kOffOut)};
    return body_loads_on(c, B, nsurf, surf_idx, surf_body, nbodies, body_c, body_area, wu, out);
}

extern "C" int ab_body_step(ab_ctx *c, double dt, int nb, ab_rigid_body *bodies, double dx,
                            double min_dist, const double dom[6], double *loads_out, int *n_new,
                            int counts[3])
{
    if (!c) return AB_ERR_ARG;
    if (counts) counts[0] = counts[1] = counts[2] = 0;
    if (n_new) *n_new = c->N;
    if (nb < 0 || nb > kMaxBodies || (nb && !bodies) || !dom || !(dt > 0.0) ||
        !std::isfinite(dt) || !(dx != 0.0) || !std::isfinite(dx))
    { c->err = "ab_body_step: invalid argument (at most 16 bodies)"; return AB_ERR_ARG; }
    if (!c->uploaded) { c->err = "ab_body_step before ab_upload"; return AB_ERR_STATE; }
    if (c->cfg.precision != AB_PREC_FP64 || c->nbatch)
    { c->err = "ab_body_step: moving bodies run on a single AB_PREC_FP64 cloud"; return AB_ERR_UNSUPPORTED; }
    // the moving-body block runs only with bodies (Solver.cpp:349, :705)
    if (nb == 0) return AB_OK;
    for (int b = 0; b < nb; ++b)
        if (bodies[b].shape < AB_BODY_SQUARE || bodies[b].shape > AB_BODY_CUBE ||
            bodies[b].npts <= 0 || !(bodies[b].mass > 0.0) || !(bodies[b].size > 0.0))
        { c->err = "ab_body_step: invalid body"; return AB_ERR_ARG; }
    hipStream_t s = c->stream;
    const bool d3 = c->dim == 3;
    int rc;
    // host sources of asynchronous copies: alive until the calls below have
    // synchronised the stream
    MoveBody mb[kMaxBodies];
    ShapeBody sb[kMaxBodies];
    double sh[4 * kMaxBodies];
    if ((rc = body_scratch(c, 0))) return rc;
    // the surface list, in particle order
    k_surface_list<<<1, 1024, 0, s>>>(c->N, nb, c->bnd, c->bid, nullptr, nullptr, c->bcnt);
    int cnt2[2] = {0, 0};
    BCHK(c, hipMemcpyAsync(cnt2, c->bcnt, sizeof cnt2, hipMemcpyDeviceToHost, s));
    BCHK(c, hipStreamSynchronize(s));
    const int ns = cnt2[0];
    if (cnt2[1])
    {
        c->err = "ab_body_step: " + std::to_string(cnt2[1]) +
                 " surface particles carry a body id with no body given";
        return AB_ERR_ARG;
    }
    if ((rc = body_scratch(c, ns))) return rc;
    k_surface_list<<<1, 1024, 0, s>>>(c->N, nb, c->bnd, c->bid, c->bidx, c->bbody, c->bcnt);
    launch_moments(c, c->bidx, ns, c->bmom);
    BCHK(c, hipGetLastError());

    // 1. loads at the current wall state (BodyDynamics.cpp:353)
    LoadBody lb[kMaxBodies];
    std::memset(lb, 0, sizeof lb);
    for (int b = 0; b < nb; ++b)
    {
        for (int a = 0; a < 3; ++a) lb[b].c[a] = bodies[b].c[a];
        lb[b].area = area(bodies[b]);
    }
    double L0[6 * kMaxBodies], L1[6 * kMaxBodies];
    if ((rc = loads(c, c->bidx, c->bbody, ns, c->bmom, nullptr, 0, lb, nb, L0, c->bcon, c->bpar,
                    c->bout)))
        return rc;

    // 2. predictor (:367-395); 2-D spin lives in the z component
    for (int b = 0; b < nb; ++b)
    {
        const ab_rigid_body &B = bodies[b];
        const double I = inertia(B);
        const double *F = L0 + 6 * b, *T = F + 3;
        LoadBody &l = lb[b];
        l.v[0] = B.vel[0] + F[0] * dt / B.mass;
        l.v[1] = B.vel[1] + F[1] * dt / B.mass;
        l.v[2] = d3 ? B.vel[2] + F[2] * dt / B.mass : 0.0;
        if (B.rotates)
        {
            if (d3)
            {
                l.w[0] = B.om[0] + T[0] * dt / I;
                l.w[1] = B.om[1] + T[1] * dt / I;
                l.w[2] = B.om[2] + T[2] * dt / I;
            }
            else
            {
                l.w[0] = 0.0; l.w[1] = 0.0;
                l.w[2] = B.angular_vel + T[2] * dt / I;
            }
        }
        else
        {
            l.w[0] = d3 ? B.om[0] : 0.0;
            l.w[1] = d3 ? B.om[1] : 0.0;
            l.w[2] = d3 ? B.om[2] : B.angular_vel;
        }
    }
    // 3. corrector loads at the predicted wall velocities, same g (:398-399)
    if ((rc = loads(c, c->bidx, c->bbody, ns, c->bmom, nullptr, 2, lb, nb, L1, c->bcon, c->bpar,
                    c->bout)))
        return rc;
    if (loads_out)
        for (int b = 0; b < nb; ++b)
        {
            std::memcpy(loads_out + 12 * b, L0 + 6 * b, 6 * sizeof(double));
            std::memcpy(loads_out + 12 * b + 6, L1 + 6 * b, 6 * sizeof(double));
        }

    // 4. rigid motion of the surface with the old state (:405-440)
    {
        std::memset(mb, 0, sizeof mb);
        for (int b = 0; b < nb; ++b)
        {
            const ab_rigid_body &B = bodies[b];
            MoveBody &m = mb[b];
            for (int a = 0; a < 3; ++a) { m.c[a] = B.c[a]; m.v[a] = B.vel[a]; }
            if (d3)
            {
                const double wmag = std::sqrt(B.om[0] * B.om[0] + B.om[1] * B.om[1] + B.om[2] * B.om[2]);
                m.rot = wmag > 1e-300;
                if (m.rot)
                {
                    for (int a = 0; a < 3; ++a) m.a[a] = B.om[a] / wmag;
                    const double th = wmag * dt;
                    m.cs = std::cos(th); m.sn = std::sin(th);
                    m.omc = 1.0 - m.cs;
                }
            }
            else
            {
                const double th = B.angular_vel * dt;
                m.cs = std::cos(th); m.sn = std::sin(th);
            }
        }
        BCHK(c, hipMemcpyAsync(c->bpar, mb, sizeof(MoveBody) * nb, hipMemcpyHostToDevice, s));
        if (ns > 0)
            k_surface_move<<<(ns + 255) / 256, 256, 0, s>>>(c->dim, ns, c->bidx, c->bbody, dt,
                                                            (const MoveBody *)c->bpar, c->x, c->y,
                                                            c->z);
        BCHK(c, hipGetLastError());
    }

    // 5. Newton-Euler with the averaged loads (:451-519)
    for (int b = 0; b < nb; ++b)
    {
        ab_rigid_body &B = bodies[b];
        const double I = inertia(B);
        const double *F0 = L0 + 6 * b, *T0 = F0 + 3, *F1 = L1 + 6 * b, *T1 = F1 + 3;
        B.c[0] += dt * B.vel[0];
        B.c[1] += dt * B.vel[1];
        if (d3) B.c[2] += dt * B.vel[2];
        if (d3)
        {
            // only the cube carries an orientation matrix (Sphere's rotate3
            // is the base class's no-op)
            if (B.shape == AB_BODY_CUBE) rotate3(B.rm, B.om[0], B.om[1], B.om[2], dt);
        }
        else
        {
            const double th = B.angular_vel * dt;
            const double ox2 = B.o[0], oy2 = B.o[1];
            B.o[0] = std::cos(th) * ox2 - std::sin(th) * oy2;
            B.o[1] = std::sin(th) * ox2 + std::cos(th) * oy2;
        }
        for (int a = 0; a < 3; ++a)
        {
            B.force[a] = 0.5 * (F0[a] + F1[a]);
            B.tq[a] = 0.5 * (T0[a] + T1[a]);
        }
        B.vel[0] += B.force[0] * dt / B.mass;
        B.vel[1] += B.force[1] * dt / B.mass;
        if (d3) B.vel[2] += B.force[2] * dt / B.mass;
        if (B.rotates)
        {
            if (d3)
            {
                B.om[0] += B.tq[0] * dt / I;
                B.om[1] += B.tq[1] * dt / I;
                B.om[2] += B.tq[2] * dt / I;
            }
            else
                B.angular_vel += B.tq[2] * dt / I;
        }
        if (d3) B.angular_vel = B.om[2];
    }

    // 5b. box-wall contact (:550-610)
    {
        double ext[7 * kMaxBodies];
        k_body_extent<<<nb, 256, 0, s>>>(ns, c->bidx, c->bbody, c->x, c->y, c->z, c->bout);
        BCHK(c, hipGetLastError());
        BCHK(c, hipMemcpyAsync(ext, c->bout, sizeof(double) * 7 * nb, hipMemcpyDeviceToHost, s));
        BCHK(c, hipStreamSynchronize(s));
        // dom = xleft, xright, ybottom, ytop, zfront, zback
        const double gap = 0.4 * std::fabs(dx);
        const double xlo = std::min(dom[0], dom[1]), xhi = std::max(dom[0], dom[1]);
        const double ylo = std::min(dom[2], dom[3]), yhi = std::max(dom[2], dom[3]);
        const double zlo = std::min(dom[5], dom[4]), zhi = std::max(dom[5], dom[4]);
        std::memset(sh, 0, sizeof sh);
        bool any = false;
        for (int b = 0; b < nb; ++b)
        {
            const double *e = ext + 7 * b;
            if (e[6] == 0.0) continue;
            const double sx = std::max(0.0, (xlo + gap) - e[0]) - std::max(0.0, e[1] - (xhi - gap));
            const double sy = std::max(0.0, (ylo + gap) - e[2]) - std::max(0.0, e[3] - (yhi - gap));
            const double sz = d3 ? std::max(0.0, (zlo + gap) - e[4]) - std::max(0.0, e[5] - (zhi - gap))
                                 : 0.0;
            if (sx == 0.0 && sy == 0.0 && sz == 0.0) continue;
            ab_rigid_body &B = bodies[b];
            B.c[0] += sx; B.c[1] += sy; B.c[2] += sz;
            sh[3 * b] = sx; sh[3 * b + 1] = sy; sh[3 * b + 2] = sz;
            sh[3 * nb + b] = 1.0;
            any = true;
            if (sx > 0.0 && B.vel[0] < 0.0) B.vel[0] = 0.0;
            if (sx < 0.0 && B.vel[0] > 0.0) B.vel[0] = 0.0;
            if (sy > 0.0 && B.vel[1] < 0.0) B.vel[1] = 0.0;
            if (sy < 0.0 && B.vel[1] > 0.0) B.vel[1] = 0.0;
            if (d3)
            {
                if (sz > 0.0 && B.vel[2] < 0.0) B.vel[2] = 0.0;
                if (sz < 0.0 && B.vel[2] > 0.0) B.vel[2] = 0.0;
            }
        }
        if (any)
        {
            BCHK(c, hipMemcpyAsync(c->bout, sh, sizeof(double) * 4 * nb, hipMemcpyHostToDevice, s));
            k_surface_shift<<<(ns + 255) / 256, 256, 0, s>>>(ns, nb, c->bidx, c->bbody, c->bout,
                                                             c->x, c->y, c->z);
            BCHK(c, hipGetLastError());
        }
    }

    // 6. surface wall state from the new positions and state (:615-638)
    {
        std::memset(sb, 0, sizeof sb);
        for (int b = 0; b < nb; ++b)
        {
            const ab_rigid_body &B = bodies[b];
            ShapeBody &q = sb[b];
            q.shape = B.shape;
            q.size = B.size;
            q.angv = B.angular_vel;
            for (int a = 0; a < 3; ++a) { q.c[a] = B.c[a]; q.v[a] = B.vel[a]; q.om[a] = B.om[a]; }
            q.o[0] = B.o[0]; q.o[1] = B.o[1];
            for (int a = 0; a < 9; ++a) q.rm[a] = B.rm[a];
        }
        BCHK(c, hipMemcpyAsync(c->bpar, sb, sizeof(ShapeBody) * nb, hipMemcpyHostToDevice, s));
        if (ns > 0)
            k_surface_refresh<<<(ns + 255) / 256, 256, 0, s>>>(
                c->view(), ns, c->bidx, c->bbody, (const ShapeBody *)c->bpar, c->wnx, c->wny,
                c->wnz, c->wux, c->wuy, c->wuz, c->mvx, c->mvy, c->mvz);
        BCHK(c, hipGetLastError());
        BCHK(c, hipStreamSynchronize(s));
    }

    // 7. cloud management around the moved bodies (Solver.cpp:379-398)
    ab_body cb[kMaxBodies];
    for (int b = 0; b < nb; ++b)
    {
        const ab_rigid_body &B = bodies[b];
        std::memset(&cb[b], 0, sizeof cb[b]);
        cb[b].shape = B.shape;
        for (int a = 0; a < 3; ++a) cb[b].c[a] = B.c[a];
        cb[b].size = B.size;
        cb[b].o[0] = B.o[0]; cb[b].o[1] = B.o[1];
        for (int a = 0; a < 9; ++a) cb[b].rm[a] = B.rm[a];
    }
    return ab_cloud_manage(c, nb, cb, nullptr, dx, min_dist, dom, n_new, counts);
}

// Several clouds' moving-body blocks in one call, one context after another
// on the calling thread. (One host thread per context overlapped the
// instances' synchronisations, but crashed inside the runtime's kernel
// launch when a profiler traced the 16 threads; it is not used.)
extern "C" int ab_body_step_many(int n, ab_ctx *const *ctxs, const ab_body_task *tasks)
{
    if (n <= 0 || n > 64 || !ctxs || !tasks) return AB_ERR_ARG;
    for (int k = 0; k < n; ++k)
    {
        if (!ctxs[k]) return AB_ERR_ARG;
        for (int q = 0; q < k; ++q)
            if (ctxs[q] == ctxs[k]) { ctxs[k]->err = "ab_body_step_many: a context appears twice"; return AB_ERR_ARG; }
    }
    std::vector<int> rcs(n, AB_OK);
    int cur = 0;
    const bool have = hipGetDevice(&cur) == hipSuccess;
    for (int k = 0; k < n; ++k)
    {
        ab_ctx *c = ctxs[k];
        const ab_body_task &t = tasks[k];
        if (hipSetDevice(c->device) != hipSuccess)
        { c->err = "ab_body_step_many: hipSetDevice failed"; rcs[k] = AB_ERR_HIP; continue; }
        rcs[k] = ab_body_step(c, t.dt, t.nbodies, t.bodies, t.dx, t.min_dist, t.dom, t.loads,
                              t.n_new, t.counts);
    }
    if (have) (void)hipSetDevice(cur);               // the caller's device as it was
    for (int k = 0; k < n; ++k)
        if (rcs[k]) return rcs[k];
    return AB_OK;
}
