/*
 * oracle.c -- CPU FP64 restatement of the reference ALE-BGK time step.
 *
 * TEST INFRASTRUCTURE (see oracle.h): the checker, never the product.
 *
 * Each function names the reference lines it restates. Loops run in the
 * reference's order and expressions keep its association, so with
 * -ffp-contract=off the setup and the per-particle kernels agree with the
 * reference CPU build bit for bit wherever the two reference backends share
 * the arithmetic (lattice, neighbour lists, wall state, moments, collision,
 * diffuse wall).
 *
 * Parity pins (tests/test_oracle.py):
 *   orc_lattice, orc_neighbours, orc_wall_state, orc_initial_state,
 *   orc_moments*, orc_collision*, orc_wallbc*   -> reference CPU goldens, exact
 *   orc_transport3d/2d -> reference CPU goldens (Gauss vs Cholesky, 1e-12);
 *                         affine / quadratic reproduction; brute-force LS
 *   orc_positivity*, orc_nad3d -> hand-built cases against closed forms
 *   orc_sum_mass       -> golden total mass; independent fsum
 *   orc_step           -> uniform-equilibrium fixed point; mass conservation
 *   orc_body_loads     -> reference CPU golden (computeBodyLoads)
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static int imax(int a, int b) { return a > b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }

/* ======================================================================== */
/* Setup                                                                     */
/* ======================================================================== */

/* generateParticles for a non-periodic cavity
 * (Initialization/ParticleGeneration.cpp:35-166). geo->dim, Nx, Ny, Nz and
 * the domain are inputs; the rest of geo is filled. Particle order: k (z)
 * outer, j, i inner. Returns N. */
int orc_lattice(orc_geom *geo, double *x, double *y, double *z,
                unsigned char *boundary, int *voxel)
{
    int ndim = geo->dim;
    int Nx = geo->Nx, Ny = geo->Ny, Nz = (ndim == 2) ? 1 : geo->Nz;
    geo->Nz = Nz;
    double Lx = geo->xright - geo->xleft;
    double Ly = geo->ytop - geo->ybottom;
    double Lz = (ndim == 3) ? (geo->zback - geo->zfront) : 1.0;
    double dx = Lx / (Nx - 1);
    double dy = Ly / (Ny - 1);
    double dz = (ndim == 3 && Nz > 1) ? Lz / (Nz - 1) : 1.0;
    geo->dx = dx; geo->dy = dy; geo->dz = dz;
    {
        double a = fabs(dx), b = fabs(dy), c = (ndim == 3 ? fabs(dz) : fabs(dx));
        double m = a; if (b > m) m = b; if (c > m) m = c;
        geo->radius = 2.5 * m;
        double c2 = (ndim == 3 ? fabs(dz) : fabs(dy));
        double n = a; if (b < n) n = b; if (c2 < n) n = c2;
        geo->minDist = 0.5 * n;
    }
    double hb = 3.0 * dx, vb = 3.0 * dy, wb;
    geo->nbhBox = (int)ceil(Lx / hb) + 1;
    geo->nbvBox = (int)ceil(Ly / vb) + 1;
    if (ndim == 3) { wb = 3.0 * dz; geo->nbwBox = (int)ceil(Lz / wb) + 1; }
    else { wb = 1.0; geo->nbwBox = 1; }
    geo->hBox = hb; geo->vBox = vb; geo->wBox = wb;
    geo->nvox = geo->nbhBox * geo->nbvBox * geo->nbwBox;

    int n = 0;
    for (int k = 0; k < Nz; k++) {
        double zz = (ndim == 3) ? geo->zfront + k * dz : 0.0;
        for (int j = 0; j < Ny; j++) {
            double yy = geo->ybottom + j * dy;
            for (int i = 0; i < Nx; i++) {
                double xx = geo->xleft + i * dx;
                x[n] = xx; y[n] = yy; z[n] = zz;
                int bx = (i == 0 || i == Nx - 1);
                int by = (j == 0 || j == Ny - 1);
                int bz = (ndim == 3) && (k == 0 || k == Nz - 1);
                boundary[n] = (unsigned char)(bx || by || bz);
                int vi = imax(0, imin(geo->nbhBox - 1, (int)floor((xx - geo->xleft) / hb)));
                int vj = imax(0, imin(geo->nbvBox - 1, (int)floor((yy - geo->ybottom) / vb)));
                int vk = (ndim == 3)
                       ? imax(0, imin(geo->nbwBox - 1, (int)floor((zz - geo->zfront) / wb)))
                       : 0;
                voxel[n] = vi + geo->nbhBox * (vj + geo->nbvBox * vk);
                n++;
            }
        }
    }
    return n;
}

/* updateVoxelsAndNeighbours, non-periodic (Initialization/NeighbourSearch.cpp
 * :17-99): bin by voxel (ascending particle index within a voxel), scan the
 * 27 (9 in 2-D) voxels dk, dj, di, keep j != i with r^2 < radius^2 (strict),
 * stop at ORC_MAX_NEIGH. voxel[] is recomputed from the positions. Unused
 * neigh slots are set to -1. Returns 0, or -1 on allocation failure. */
int orc_neighbours(const orc_geom *geo, int N, const double *x,
                   const double *y, const double *z, int *voxel,
                   int *neigh, int *nneigh)
{
    int ndim = geo->dim, nvox = geo->nvox;
    int *count = (int *)calloc((size_t)nvox + 1, sizeof(int));
    int *start = (int *)calloc((size_t)nvox + 1, sizeof(int));
    int *list = (int *)malloc(sizeof(int) * (size_t)(N > 0 ? N : 1));
    if (!count || !start || !list) { free(count); free(start); free(list); return -1; }
    for (int i = 0; i < N; i++) {
        int vi = imax(0, imin(geo->nbhBox - 1, (int)floor((x[i] - geo->xleft) / geo->hBox)));
        int vj = imax(0, imin(geo->nbvBox - 1, (int)floor((y[i] - geo->ybottom) / geo->vBox)));
        int vk = (ndim == 3)
               ? imax(0, imin(geo->nbwBox - 1, (int)floor((z[i] - geo->zfront) / geo->wBox)))
               : 0;
        voxel[i] = vi + geo->nbhBox * (vj + geo->nbvBox * vk);
        if (voxel[i] < nvox) count[voxel[i]]++;
    }
    for (int v = 0; v < nvox; v++) start[v + 1] = start[v] + count[v];
    memset(count, 0, sizeof(int) * (size_t)nvox);
    for (int i = 0; i < N; i++)
        if (voxel[i] < nvox) list[start[voxel[i]] + count[voxel[i]]++] = i;

    double r2max = geo->radius * geo->radius;
    int dkmax = (ndim == 3) ? 1 : 0;
    for (int i = 0; i < N; i++) {
        int *nb = neigh + (size_t)i * ORC_MAX_NEIGH;
        int nn = 0;
        int vi = voxel[i] % geo->nbhBox;
        int tmp = voxel[i] / geo->nbhBox;
        int vj = tmp % geo->nbvBox;
        int vk = tmp / geo->nbvBox;
        int done = 0;
        for (int dk = -dkmax; dk <= dkmax && !done; dk++)
            for (int dj = -1; dj <= 1 && !done; dj++)
                for (int di = -1; di <= 1 && !done; di++) {
                    int ni = vi + di, nj = vj + dj, nk = vk + dk;
                    if (ni < 0 || ni >= geo->nbhBox) continue;
                    if (nj < 0 || nj >= geo->nbvBox) continue;
                    if (nk < 0 || nk >= geo->nbwBox) continue;
                    int vox = ni + geo->nbhBox * (nj + geo->nbvBox * nk);
                    for (int q = start[vox]; q < start[vox + 1]; q++) {
                        int j = list[q];
                        if (j == i) continue;
                        double ddx = x[j] - x[i];
                        double ddy = y[j] - y[i];
                        double ddz = (ndim == 3) ? z[j] - z[i] : 0.0;
                        if (ddx * ddx + ddy * ddy + ddz * ddz < r2max) {
                            nb[nn++] = j;
                            if (nn >= ORC_MAX_NEIGH) { done = 1; break; }
                        }
                    }
                }
        nneigh[i] = nn;
        for (int m = nn; m < ORC_MAX_NEIGH; m++) nb[m] = -1;
    }
    free(count); free(start); free(list);
    return 0;
}

/* getWallNormal (MLSTransport/MLSBasis.hpp:228-287). out7 = n(3), Uw(3), Tw. */
void orc_wall_normal(const orc_geom *geo, double x, double y, double z,
                     double Tw, double TwTop, const double *U, double *o)
{
    const double eps = 1.0e-13;
    double nx = 0, ny = 0, nz = 0, ux = 0, uy = 0, uz = 0, tws = 0;
    int nh = 0;
#define HIT(ax, ay, az, f, over)                                  \
    do { nx += ax; ny += ay; nz += az;                           \
         ux += U[3 * (f)]; uy += U[3 * (f) + 1]; uz += U[3 * (f) + 2]; \
         tws += ((over) > 0.0) ? (over) : Tw; ++nh; } while (0)
    if (fabs(x - geo->xleft) < eps)   HIT(1.0, 0.0, 0.0, 0, -1.0);
    if (fabs(x - geo->xright) < eps)  HIT(-1.0, 0.0, 0.0, 1, -1.0);
    if (fabs(y - geo->ytop) < eps)    HIT(0.0, -1.0, 0.0, 2, TwTop);
    if (fabs(y - geo->ybottom) < eps) HIT(0.0, 1.0, 0.0, 3, -1.0);
    if (fabs(z - geo->zfront) < eps)  HIT(0.0, 0.0, -1.0, 4, -1.0);
    if (fabs(z - geo->zback) < eps)   HIT(0.0, 0.0, 1.0, 5, -1.0);
#undef HIT
    double T;
    if (nh == 0) T = Tw;
    else if (nh == 1) T = tws;
    else {
        double nl = sqrt(nx * nx + ny * ny + nz * nz);
        if (nl > eps) { nx /= nl; ny /= nl; nz /= nl; }
        ux /= nh; uy /= nh; uz /= nh;
        T = tws / nh;
    }
    o[0] = nx; o[1] = ny; o[2] = nz; o[3] = ux; o[4] = uy; o[5] = uz; o[6] = T;
}

/* initWallData (BoundaryConditions/WallState.cpp:16-26); body surface
 * particles (bodyid >= 0) keep their state. The reference passes the 2-D
 * lattice z = 0 into getWallNormal, which then also matches zback = 0 of the
 * 2-D domain {.., zfront=0, zback=0}: that hit is reproduced. */
void orc_wall_state(const orc_geom *geo, int N, const double *x,
                    const double *y, const double *z,
                    const unsigned char *boundary, const int *bodyid,
                    double Tw, const double *Uface,
                    double *wnx, double *wny, double *wnz, double *wUx,
                    double *wUy, double *wUz, double *wT)
{
    for (int i = 0; i < N; i++) {
        if (!boundary[i] || (bodyid && bodyid[i] >= 0)) continue;
        double o[7];
        orc_wall_normal(geo, x[i], y[i], z[i], Tw, -1.0, Uface, o);
        wnx[i] = o[0]; wny[i] = o[1]; wnz[i] = o[2];
        wUx[i] = o[3]; wUy[i] = o[4]; wUz[i] = o[5]; wT[i] = o[6];
    }
}

/* applyLidProfile, parabolic lid (TestCases/TestCases.cpp:240-248). */
void orc_lid_parabolic(const orc_geom *geo, int N, const double *x,
                       const double *y, const unsigned char *boundary,
                       const int *bodyid, double peak, double *wUx)
{
    const double L = geo->xright - geo->xleft;
    for (int i = 0; i < N; i++) {
        if (!boundary[i] || (bodyid && bodyid[i] >= 0)) continue;
        if (fabs(y[i] - geo->ytop) > 1e-13) continue;
        const double s = x[i] / L;
        wUx[i] = peak * (2.0 * s) * (2.0 * s) * (2.0 - 2.0 * s) * (2.0 - 2.0 * s);
    }
}

/* maxwellian3D (MLSTransport/MLSBasis.hpp:112-122). */
double orc_maxwellian3d(double vx, double vy, double vz, double ux,
                        double uy, double uz, double T, double rho, double R)
{
    double twoRT = 2.0 * R * T;
    double exponent = -((vx - ux) * (vx - ux) + (vy - uy) * (vy - uy)
                        + (vz - uz) * (vz - uz)) / twoRT;
    double norm = pow(M_PI * twoRT, 1.5);
    return rho / norm * exp(exponent);
}

/* maxwellianG1 (MLSTransport/MLSBasis.hpp:87-94). */
double orc_maxwellian_g1(double vx, double vy, double ux, double uy,
                         double T, double rho, double R)
{
    double twoRT = 2.0 * R * T;
    double exponent = -((vx - ux) * (vx - ux) + (vy - uy) * (vy - uy)) / twoRT;
    return rho / (M_PI * twoRT) * exp(exponent);
}

/* applyIC: gas at rest at T0 (Initialization/InitialConditions.cpp:18-81). */
void orc_initial_state(int dim, int N, int Nv, double vlo, double dv,
                       double rho0, double T0, double R, double *g)
{
    int Nv2 = Nv * Nv;
    size_t gs = (dim == 3) ? (size_t)Nv * Nv2 : (size_t)2 * Nv2;
    for (int i = 0; i < N; i++) {
        double *p = g + (size_t)i * gs;
        if (dim == 3) {
            for (int kv = 0; kv < Nv; ++kv) {
                double vz = vlo + kv * dv;
                for (int jv = 0; jv < Nv; ++jv) {
                    double vy = vlo + jv * dv;
                    for (int iv = 0; iv < Nv; ++iv) {
                        double vx = vlo + iv * dv;
                        p[iv + Nv * (jv + Nv * kv)] =
                            orc_maxwellian3d(vx, vy, vz, 0.0, 0.0, 0.0, T0, rho0, R);
                    }
                }
            }
        } else {
            for (int jv = 0; jv < Nv; ++jv) {
                double vy = vlo + jv * dv;
                for (int iv = 0; iv < Nv; ++iv) {
                    double vx = vlo + iv * dv;
                    int lin = iv + Nv * jv;
                    double Mg1 = orc_maxwellian_g1(vx, vy, 0.0, 0.0, T0, rho0, R);
                    p[lin] = Mg1;
                    p[lin + Nv2] = 0.5 * R * T0 * Mg1;
                }
            }
        }
    }
}

/* ======================================================================== */
/* Transport                                                                 */
/* ======================================================================== */

static double mvel(const double *a, int i) { return a ? a[i] : 0.0; }

/* kCopyGt then kTransportNode3DT<POLY,false> and <4,true>
 * (Solver/CudaBackend.cu:191-197, 699-869; launch :1660-1677). POLY is 10
 * for mlsOrder >= 2 and 4 otherwise on interior particles, 4 on the
 * boundary. mlsflag / mlsovr may be NULL. The reference accumulates mlsOvr
 * with atomicAdd in nondeterministic order; here it is summed in lin order. */
void orc_transport3d(const orc_cloud *c, const double *g, double *gt,
                     double dt, double h, int mlsOrder, double tol,
                     int clamp, int *mlsflag, double *mlsovr)
{
    int Nv = c->Nv;
    size_t gs = (size_t)c->gstride;
    double dv3 = c->dv * c->dv * c->dv;
    memcpy(gt, g, sizeof(double) * gs * (size_t)c->N);
    for (int i = 0; i < c->N; i++) {
        int bnd = c->boundary[i] != 0;
        int P = bnd ? 4 : (mlsOrder >= 2 ? 10 : 4);
        int nn = c->nneigh[i];
        const int *nb = c->neigh + (size_t)i * ORC_MAX_NEIGH;
        for (int lin = 0; lin < (int)gs; lin++) {
            int iv = lin % Nv, jv = (lin / Nv) % Nv, kv = lin / (Nv * Nv);
            double vx = c->vlo + iv * c->dv;
            double vy = c->vlo + jv * c->dv;
            double vz = c->vlo + kv * c->dv;
            if (bnd) {
                double cn = (vx - c->wUx[i]) * c->wnx[i]
                          + (vy - c->wUy[i]) * c->wny[i]
                          + (vz - c->wUz[i]) * c->wnz[i];
                if (cn >= 0.0) continue;
            }
            double xf = c->x[i] - (vx - mvel(c->mvx, i)) * dt;
            double yf = c->y[i] - (vy - mvel(c->mvy, i)) * dt;
            double zf = c->z[i] - (vz - mvel(c->mvz, i)) * dt;
            double A[100], b1[10], basis[10];
            for (int q = 0; q < P * P; ++q) A[q] = 0.0;
            for (int q = 0; q < P; ++q) b1[q] = 0.0;
            int has = 0;
            double sumW = 0.0, sumWG = 0.0, gmin = 1e300, gmax = -1e300;
            for (int m = 0; m < nn; ++m) {
                int j = nb[m];
                double dxn = c->x[j] - xf;
                double dyn = c->y[j] - yf;
                double dzn = c->z[j] - zf;
                double r2 = dxn * dxn + dyn * dyn + dzn * dzn;
                /* mlsWeightSq (MLSBasis.hpp:142-148) */
                double q2 = r2 / (h * h);
                double w;
                if (q2 >= 1.0) w = 0.0;
                else { double t = 1.0 - q2; w = t * t; }
                if (w < 1e-30) continue;
                /* fillBasis 3-D (MLSBasis.hpp:158-195) */
                double sx = dxn / h, sy = dyn / h, sz = dzn / h;
                basis[0] = 1.0; basis[1] = sx; basis[2] = sy; basis[3] = sz;
                if (P >= 10) {
                    basis[4] = sx * sx; basis[5] = sy * sy; basis[6] = sz * sz;
                    basis[7] = sx * sy; basis[8] = sx * sz; basis[9] = sy * sz;
                }
                double gj = g[(size_t)j * gs + lin];
                for (int p1 = 0; p1 < P; ++p1) {
                    double wb = w * basis[p1];
                    for (int p2 = p1; p2 < P; ++p2) A[p1 * P + p2] += wb * basis[p2];
                    b1[p1] += wb * gj;
                }
                sumW += w;
                sumWG += w * gj;
                gmin = fmin(gmin, gj);
                gmax = fmax(gmax, gj);
                has = 1;
            }
            if (!has) continue;
            double dmax = 0.0;
            for (int k = 0; k < P; ++k) dmax = fmax(dmax, A[k * P + k]);
            int degenerate = (dmax <= 0.0);
            for (int k = 0; k < P; ++k) {
                double d = A[k * P + k];
                for (int j = 0; j < k; ++j) { double l = A[j * P + k]; d -= l * l; }
                if (!(d > tol * dmax)) degenerate = 1;
                d = degenerate ? 1.0 : sqrt(d);
                A[k * P + k] = d;
                for (int m2 = k + 1; m2 < P; ++m2) {
                    double v = A[k * P + m2];
                    for (int j = 0; j < k; ++j) v -= A[j * P + k] * A[j * P + m2];
                    A[k * P + m2] = v / d;
                }
            }
            for (int k = 0; k < P; ++k) {
                double v = b1[k];
                for (int j = 0; j < k; ++j) v -= A[j * P + k] * b1[j];
                b1[k] = v / A[k * P + k];
            }
            for (int k = P - 1; k >= 0; --k) {
                double v = b1[k];
                for (int j = k + 1; j < P; ++j) v -= A[k * P + j] * b1[j];
                b1[k] = v / A[k * P + k];
            }
            double out = b1[0];
            if (degenerate || !isfinite(out))
                out = (sumW > 0.0) ? (sumWG / sumW) : g[(size_t)i * gs + lin];
            if (mlsflag && (degenerate || !isfinite(b1[0]))) mlsflag[i] |= 1;
            double excess = fmax(fmax(out - gmax, gmin - out), 0.0);
            int overshoot = (out < gmin) || (out > gmax);
            if (overshoot && mlsflag) mlsflag[i] |= 2;
            if (excess > 0.0 && mlsovr) mlsovr[i] += excess * dv3;
            if (overshoot && clamp) out = fmin(fmax(out, gmin), gmax);
            gt[(size_t)i * gs + lin] = out;
        }
    }
}

/* kCopyGt then kTransportNode, 2-D Chu layers (Solver/CudaBackend.cu:199-300):
 * weight through sqrt + mlsWeight, pivoted Gauss on both layers. */
void orc_transport2d(const orc_cloud *c, const double *g, double *gt,
                     double dt, double h, int mlsOrder)
{
    int Nv = c->Nv, Nv2 = Nv * Nv;
    size_t gs = (size_t)c->gstride;
    memcpy(gt, g, sizeof(double) * gs * (size_t)c->N);
    for (int i = 0; i < c->N; i++) {
        int bnd = c->boundary[i] != 0;
        int poly = (mlsOrder >= 2 && !bnd) ? 6 : 3;
        int nn = c->nneigh[i];
        const int *nb = c->neigh + (size_t)i * ORC_MAX_NEIGH;
        for (int lin = 0; lin < Nv2; lin++) {
            int iv = lin % Nv, jv = lin / Nv;
            double vx = c->vlo + iv * c->dv;
            double vy = c->vlo + jv * c->dv;
            if (bnd) {
                double cn = (vx - c->wUx[i]) * c->wnx[i] + (vy - c->wUy[i]) * c->wny[i];
                if (cn >= 0.0) continue;
            }
            double xf = c->x[i] - (vx - mvel(c->mvx, i)) * dt;
            double yf = c->y[i] - (vy - mvel(c->mvy, i)) * dt;
            double A[36], bb1[6], bb2[6], s1[6], s2[6], basis[6];
            for (int q = 0; q < poly * poly; ++q) A[q] = 0;
            for (int q = 0; q < poly; ++q) { bb1[q] = 0; bb2[q] = 0; }
            int has = 0;
            for (int m = 0; m < nn; ++m) {
                int j = nb[m];
                double dxn = c->x[j] - xf, dyn = c->y[j] - yf;
                double r = sqrt(dxn * dxn + dyn * dyn);
                double w;                     /* mlsWeight (MLSBasis.hpp:127-133) */
                if (r >= h) w = 0.0;
                else { double q = r / h; double t = 1.0 - q * q; w = t * t; }
                if (w < 1e-30) continue;
                double sx = dxn / h, sy = dyn / h;
                basis[0] = 1.0; basis[1] = sx; basis[2] = sy;
                if (poly >= 6) { basis[3] = sx * sx; basis[4] = sx * sy; basis[5] = sy * sy; }
                for (int p1 = 0; p1 < poly; ++p1)
                    for (int p2 = 0; p2 < poly; ++p2)
                        A[p1 * poly + p2] += w * basis[p1] * basis[p2];
                const double *gj = g + (size_t)j * gs;
                double gn1 = gj[lin], gn2 = gj[lin + Nv2];
                for (int q = 0; q < poly; ++q) {
                    bb1[q] += w * basis[q] * gn1;
                    bb2[q] += w * basis[q] * gn2;
                }
                has = 1;
            }
            if (!has) continue;
            for (int col = 0; col < poly; ++col) {
                int piv = col;
                double mx = fabs(A[col * poly + col]);
                for (int row = col + 1; row < poly; ++row)
                    if (fabs(A[row * poly + col]) > mx) { mx = fabs(A[row * poly + col]); piv = row; }
                if (piv != col) {
                    for (int q = 0; q < poly; ++q) {
                        double t = A[col * poly + q];
                        A[col * poly + q] = A[piv * poly + q];
                        A[piv * poly + q] = t;
                    }
                    double tb = bb1[col]; bb1[col] = bb1[piv]; bb1[piv] = tb;
                    tb = bb2[col]; bb2[col] = bb2[piv]; bb2[piv] = tb;
                }
                double d = A[col * poly + col];
                if (fabs(d) < 1e-30) continue;
                for (int row = col + 1; row < poly; ++row) {
                    double f = A[row * poly + col] / d;
                    for (int q = col; q < poly; ++q) A[row * poly + q] -= f * A[col * poly + q];
                    bb1[row] -= f * bb1[col];
                    bb2[row] -= f * bb2[col];
                }
            }
            for (int row = poly - 1; row >= 0; --row) {
                double v1 = bb1[row], v2 = bb2[row];
                for (int q = row + 1; q < poly; ++q) {
                    v1 -= A[row * poly + q] * s1[q];
                    v2 -= A[row * poly + q] * s2[q];
                }
                double d = A[row * poly + row];
                s1[row] = fabs(d) > 1e-30 ? v1 / d : 0.0;
                s2[row] = fabs(d) > 1e-30 ? v2 / d : 0.0;
            }
            double *gtp = gt + (size_t)i * gs;
            gtp[lin] = s1[0];
            gtp[lin + Nv2] = s2[0];
        }
    }
}

/* ======================================================================== */
/* Limiters, commit, conservation                                            */
/* ======================================================================== */

/* kPositivity3D (Solver/CudaBackend.cu:882-907). */
void orc_positivity3d(const orc_cloud *c, const double *g, double *gt,
                      double *limflag)
{
    size_t gs = (size_t)c->gstride;
    for (int i = 0; i < c->N; i++) {
        double *t = gt + (size_t)i * gs;
        const double *g0 = g + (size_t)i * gs;
        double mPos = 0, mNeg = 0, mFallback = 0;
        for (size_t k = 0; k < gs; ++k) {
            double v = t[k];
            if (v >= 0) mPos += v;
            else { mNeg += v; mFallback += g0[k]; }
        }
        if (mNeg < -1e-30 && (mPos + mNeg) > 0) {
            limflag[i] += 1.0;
            double target = mPos + mNeg;
            double replaced = mPos + mFallback;
            double sc = (replaced > 1e-30) ? target / replaced : 0.0;
            for (size_t k = 0; k < gs; ++k) {
                double v = (t[k] < 0.0) ? g0[k] : t[k];
                t[k] = v * sc;
            }
        }
    }
}

/* kPositivity, 2-D g1 layer (Solver/CudaBackend.cu:304-326). */
void orc_positivity2d(const orc_cloud *c, double *gt, double *limflag)
{
    int Nv2 = c->Nv * c->Nv;
    size_t gs = (size_t)c->gstride;
    for (int i = 0; i < c->N; i++) {
        double *t = gt + (size_t)i * gs;
        double mPos = 0, mNeg = 0;
        for (int k = 0; k < Nv2; ++k) {
            double v = t[k];
            if (v >= 0) mPos += v; else mNeg += v;
        }
        if (mNeg < -1e-30 && mPos > 1e-30 && (mPos + mNeg) > 0) {
            limflag[i] += 1.0;
            double sc = (mPos + mNeg) / mPos;
            for (int k = 0; k < Nv2; ++k) {
                if (t[k] < 0) { t[k] = 0; t[k + Nv2] = 0; }
                else { t[k] *= sc; t[k + Nv2] *= sc; }
            }
        }
    }
}

/* kNAD3D (Solver/CudaBackend.cu:920-954): interior particles only. */
void orc_nad3d(const orc_cloud *c, double *gt, const double *rho,
               double *limflag)
{
    size_t gs = (size_t)c->gstride;
    double dv3 = c->dv * c->dv * c->dv;
    for (int i = 0; i < c->N; i++) {
        if (c->boundary[i]) continue;
        double *t = gt + (size_t)i * gs;
        double rho_new = 0.0;
        for (size_t k = 0; k < gs; ++k) rho_new += t[k];
        rho_new *= dv3;
        int nn = c->nneigh[i];
        const int *nb = c->neigh + (size_t)i * ORC_MAX_NEIGH;
        double rmin = 1e300, rmax = -1e300;
        for (int k = 0; k < nn; ++k) {
            double rn = rho[nb[k]];
            rmin = fmin(rmin, rn);
            rmax = fmax(rmax, rn);
        }
        if (rmin > rmax) continue;
        double slack = 0.1 * fmax(rmax - rmin, 1e-12) + 1e-9;
        double lo = rmin - slack, hi = rmax + slack;
        if (rho_new > hi && rho_new > 1e-30) {
            limflag[i] += 2.0;
            double sc = hi / rho_new;
            for (size_t k = 0; k < gs; ++k) t[k] *= sc;
        } else if (rho_new < lo && rho_new > 1e-30) {
            limflag[i] += 2.0;
            double sc = lo / rho_new;
            for (size_t k = 0; k < gs; ++k) t[k] *= sc;
        }
    }
}

/* sumMassHost with kSumMass3D / kSumMass (Solver/CudaBackend.cu:391-436,
 * 1384-1395): per particle a serial sum of g*dv^d (3-D: every node; 2-D: the
 * g1 layer), a shared-memory tree within each block of 128 particles, then a
 * serial host sum over blocks. */
double orc_sum_mass(const orc_cloud *c, const double *g)
{
    size_t gs = (size_t)c->gstride;
    int Nv2 = c->Nv * c->Nv;
    double dvd = (c->dim == 3) ? c->dv * c->dv * c->dv : c->dv * c->dv;
    size_t nk = (c->dim == 3) ? gs : (size_t)Nv2;
    int nb = (c->N + 127) / 128;
    double total = 0.0;
    for (int b = 0; b < nb; b++) {
        double sh[128];
        for (int t = 0; t < 128; t++) {
            int i = b * 128 + t;
            double m = 0;
            if (i < c->N) {
                const double *p = g + (size_t)i * gs;
                for (size_t k = 0; k < nk; ++k) m += p[k] * dvd;
            }
            sh[t] = m;
        }
        for (int sft = 64; sft > 0; sft >>= 1)
            for (int t = 0; t < sft; t++) sh[t] += sh[t + sft];
        total += sh[0];
    }
    return total;
}

/* kScaleG (Solver/CudaBackend.cu:478-485). */
void orc_scale(const orc_cloud *c, double *g, double s)
{
    size_t n = (size_t)c->gstride * (size_t)c->N;
    for (size_t k = 0; k < n; ++k) g[k] *= s;
}

/* ======================================================================== */
/* Diffuse wall, moments, collision                                          */
/* ======================================================================== */

/* kWallBC3D (Solver/CudaBackend.cu:956-1030); also applyBoundaryBC 3-D
 * (BoundaryConditions/DiffuseWall.cpp:44-98). */
void orc_wallbc3d(const orc_cloud *c, double *g, double R, double *rho,
                  double *ux, double *uy, double *uz, double *T)
{
    int Nv = c->Nv;
    size_t gs = (size_t)c->gstride;
    double dv3 = c->dv * c->dv * c->dv;
    for (int i = 0; i < c->N; i++) {
        if (!c->boundary[i]) continue;
        double *p = g + (size_t)i * gs;
        double nx = c->wnx[i], ny = c->wny[i], nz = c->wnz[i];
        double Uwx = c->wUx[i], Uwy = c->wUy[i], Uwz = c->wUz[i], Tw = c->wT[i];
        double RT2w = 2.0 * R * Tw;
        double norm = pow(M_PI * RT2w, 1.5);
        double intM = 0, intP = 0;
        for (int kv = 0; kv < Nv; ++kv) {
            double vz = c->vlo + kv * c->dv;
            for (int jv = 0; jv < Nv; ++jv) {
                double vy = c->vlo + jv * c->dv;
                for (int iv = 0; iv < Nv; ++iv) {
                    double vx = c->vlo + iv * c->dv;
                    double cn = (vx - Uwx) * nx + (vy - Uwy) * ny + (vz - Uwz) * nz;
                    int lin = iv + Nv * (jv + Nv * kv);
                    if (cn < 0.0) intM += cn * p[lin] * dv3;
                    else if (cn > 0.0) {
                        double Mw = exp(-((vx - Uwx) * (vx - Uwx) + (vy - Uwy) * (vy - Uwy)
                                          + (vz - Uwz) * (vz - Uwz)) / RT2w) / norm;
                        intP += cn * Mw * dv3;
                    }
                }
            }
        }
        double rho_w = (intP > 1e-30) ? (-intM / intP) : 1.0;
        for (int kv = 0; kv < Nv; ++kv) {
            double vz = c->vlo + kv * c->dv;
            for (int jv = 0; jv < Nv; ++jv) {
                double vy = c->vlo + jv * c->dv;
                for (int iv = 0; iv < Nv; ++iv) {
                    double vx = c->vlo + iv * c->dv;
                    double cn = (vx - Uwx) * nx + (vy - Uwy) * ny + (vz - Uwz) * nz;
                    if (cn > 0.0) {
                        int lin = iv + Nv * (jv + Nv * kv);
                        p[lin] = rho_w * exp(-((vx - Uwx) * (vx - Uwx) + (vy - Uwy) * (vy - Uwy)
                                               + (vz - Uwz) * (vz - Uwz)) / RT2w) / norm;
                    }
                }
            }
        }
        rho[i] = rho_w; ux[i] = Uwx; uy[i] = Uwy; uz[i] = Uwz; T[i] = Tw;
    }
}

/* kWallBC, 2-D (Solver/CudaBackend.cu:338-389). */
void orc_wallbc2d(const orc_cloud *c, double *g, double R, double *rho,
                  double *ux, double *uy, double *T)
{
    int Nv = c->Nv, Nv2 = Nv * Nv;
    size_t gs = (size_t)c->gstride;
    double dv2 = c->dv * c->dv;
    for (int i = 0; i < c->N; i++) {
        if (!c->boundary[i]) continue;
        double *p = g + (size_t)i * gs;
        double nx = c->wnx[i], ny = c->wny[i];
        double Uwx = c->wUx[i], Uwy = c->wUy[i], Tw = c->wT[i];
        double RT2w = 2.0 * R * Tw;
        double intM = 0, intP = 0;
        for (int jv = 0; jv < Nv; ++jv) {
            double vy = c->vlo + jv * c->dv;
            for (int iv = 0; iv < Nv; ++iv) {
                double vx = c->vlo + iv * c->dv;
                double cn = (vx - Uwx) * nx + (vy - Uwy) * ny;
                int lin = iv + Nv * jv;
                if (cn < 0) intM += cn * p[lin] * dv2;
                if (cn > 0) {
                    double Mw = exp(-((vx - Uwx) * (vx - Uwx) + (vy - Uwy) * (vy - Uwy)) / RT2w)
                              / (M_PI * RT2w);
                    intP += cn * Mw * dv2;
                }
            }
        }
        double rho_w = (intP > 1e-30) ? (-intM / intP) : 1.0;
        for (int jv = 0; jv < Nv; ++jv) {
            double vy = c->vlo + jv * c->dv;
            for (int iv = 0; iv < Nv; ++iv) {
                double vx = c->vlo + iv * c->dv;
                double cn = (vx - Uwx) * nx + (vy - Uwy) * ny;
                int lin = iv + Nv * jv;
                if (cn > 0) {
                    double Mw = rho_w
                              * exp(-((vx - Uwx) * (vx - Uwx) + (vy - Uwy) * (vy - Uwy)) / RT2w)
                              / (M_PI * RT2w);
                    p[lin] = Mw;
                    p[lin + Nv2] = 0.5 * R * Tw * Mw;
                }
            }
        }
        rho[i] = rho_w; ux[i] = Uwx; uy[i] = Uwy; T[i] = Tw;
    }
}

/* kMoments3D (Solver/CudaBackend.cu:532-569); computeMoments 3-D
 * (MomentCalculations/Moments.cpp:43-95). Interior particles only. */
void orc_moments3d(const orc_cloud *c, const double *g, double R,
                   double *rho, double *ux, double *uy, double *uz,
                   double *T)
{
    int Nv = c->Nv;
    size_t gs = (size_t)c->gstride;
    double dv3 = c->dv * c->dv * c->dv;
    for (int i = 0; i < c->N; i++) {
        if (c->boundary[i]) continue;
        const double *p = g + (size_t)i * gs;
        double r = 0, rux = 0, ruy = 0, ruz = 0, E = 0;
        for (int kv = 0; kv < Nv; ++kv) {
            double vz = c->vlo + kv * c->dv;
            for (int jv = 0; jv < Nv; ++jv) {
                double vy = c->vlo + jv * c->dv;
                for (int iv = 0; iv < Nv; ++iv) {
                    double vx = c->vlo + iv * c->dv;
                    double f = p[iv + Nv * (jv + Nv * kv)];
                    r += f * dv3;
                    rux += vx * f * dv3;
                    ruy += vy * f * dv3;
                    ruz += vz * f * dv3;
                    E += (vx * vx + vy * vy + vz * vz) * f * dv3;
                }
            }
        }
        if (r > 1e-30) {
            rho[i] = r;
            ux[i] = rux / r;
            uy[i] = ruy / r;
            uz[i] = ruz / r;
            double u2 = ux[i] * ux[i] + uy[i] * uy[i] + uz[i] * uz[i];
            double t = (E / r - u2) / (3.0 * R);
            T[i] = t > 1e-6 ? t : 1e-6;
        }
    }
}

/* kMoments, 2-D (Solver/CudaBackend.cu:124-158). */
void orc_moments2d(const orc_cloud *c, const double *g, double R,
                   double *rho, double *ux, double *uy, double *T)
{
    int Nv = c->Nv, Nv2 = Nv * Nv;
    size_t gs = (size_t)c->gstride;
    double dv2 = c->dv * c->dv;
    for (int i = 0; i < c->N; i++) {
        if (c->boundary[i]) continue;
        const double *p = g + (size_t)i * gs;
        double r = 0, rux = 0, ruy = 0, E = 0;
        for (int jv = 0; jv < Nv; ++jv) {
            double vy = c->vlo + jv * c->dv;
            for (int iv = 0; iv < Nv; ++iv) {
                double vx = c->vlo + iv * c->dv;
                double g1 = p[iv + Nv * jv];
                double g2 = p[iv + Nv * jv + Nv2];
                r += g1 * dv2;
                rux += vx * g1 * dv2;
                ruy += vy * g1 * dv2;
                E += (vx * vx + vy * vy) * g1 * dv2 + 2.0 * g2 * dv2;
            }
        }
        if (r > 1e-30) {
            rho[i] = r;
            ux[i] = rux / r;
            uy[i] = ruy / r;
            double u2 = ux[i] * ux[i] + uy[i] * uy[i];
            double t = (E / r - u2) / (3.0 * R);
            T[i] = t > 1e-6 ? t : 1e-6;
        }
    }
}

/* kCollisionNode3D (Solver/CudaBackend.cu:1181-1197); bgkCollision 3-D
 * (MomentCalculations/Collision.cpp:204-221). Interior particles only. */
void orc_collision3d(const orc_cloud *c, double *g, double dt, double tau,
                     double R, const double *rho, const double *ux,
                     const double *uy, const double *uz, const double *T)
{
    int Nv = c->Nv;
    size_t gs = (size_t)c->gstride;
    double lam = dt / tau, coeff = 1.0 / (1.0 + lam);
    for (int i = 0; i < c->N; i++) {
        if (c->boundary[i]) continue;
        double *p = g + (size_t)i * gs;
        for (int lin = 0; lin < (int)gs; ++lin) {
            int iv = lin % Nv, jv = (lin / Nv) % Nv, kv = lin / (Nv * Nv);
            double vx = c->vlo + iv * c->dv, vy = c->vlo + jv * c->dv,
                   vz = c->vlo + kv * c->dv;
            double M = orc_maxwellian3d(vx, vy, vz, ux[i], uy[i], uz[i], T[i], rho[i], R);
            p[lin] = coeff * (p[lin] + lam * M);
        }
    }
}

/* kCollisionNode, 2-D (Solver/CudaBackend.cu:1165-1179). */
void orc_collision2d(const orc_cloud *c, double *g, double dt, double tau,
                     double R, const double *rho, const double *ux,
                     const double *uy, const double *T)
{
    int Nv = c->Nv, Nv2 = Nv * Nv;
    size_t gs = (size_t)c->gstride;
    double lam = dt / tau, coeff = 1.0 / (1.0 + lam);
    for (int i = 0; i < c->N; i++) {
        if (c->boundary[i]) continue;
        double *p = g + (size_t)i * gs;
        for (int lin = 0; lin < Nv2; ++lin) {
            int iv = lin % Nv, jv = lin / Nv;
            double vx = c->vlo + iv * c->dv, vy = c->vlo + jv * c->dv;
            double Mg1 = orc_maxwellian_g1(vx, vy, ux[i], uy[i], T[i], rho[i], R);
            p[lin] = coeff * (p[lin] + lam * Mg1);
            p[lin + Nv2] = coeff * (p[lin + Nv2] + lam * 0.5 * R * T[i] * Mg1);
        }
    }
}

/* ======================================================================== */
/* Discrete-Maxwellian collision (not in the reference)                     */
/* ======================================================================== */

double orc_dm_scale(int Nv, double vlo, double dv)
{
    double vhi = vlo + (Nv - 1) * dv;
    return fmax(fabs(vlo), fabs(vhi));
}

/* -|v-u|^2/2RT + ln(rho/(2 pi R T)^1.5) in the scaled basis. */
void orc_dm_guess(double rho, double ux, double uy, double uz, double T,
                  double R, double cs, double a[5])
{
    double RT = R * T;
    a[0] = log(rho) - 1.5 * log(2.0 * M_PI * RT) - (ux * ux + uy * uy + uz * uz) / (2.0 * RT);
    a[1] = cs * ux / RT;
    a[2] = cs * uy / RT;
    a[3] = cs * uz / RT;
    a[4] = -cs * cs / (2.0 * RT);
}

/* Solve J x = F for symmetric positive definite J (5x5, full storage) by
 * Cholesky; returns 0 when a pivot is not positive. */
static int chol5(const double J[25], const double F[5], double x[5])
{
    double L[25] = {0}, y[5];
    for (int i = 0; i < 5; ++i) {
        for (int j = 0; j <= i; ++j) {
            double s = J[i * 5 + j];
            for (int k = 0; k < j; ++k) s -= L[i * 5 + k] * L[j * 5 + k];
            if (i == j) {
                if (!(s > 0.0)) return 0;
                L[i * 5 + i] = sqrt(s);
            } else {
                L[i * 5 + j] = s / L[j * 5 + j];
            }
        }
    }
    for (int i = 0; i < 5; ++i) {
        double s = F[i];
        for (int k = 0; k < i; ++k) s -= L[i * 5 + k] * y[k];
        y[i] = s / L[i * 5 + i];
    }
    for (int i = 4; i >= 0; --i) {
        double s = y[i];
        for (int k = i + 1; k < 5; ++k) s -= L[k * 5 + i] * x[k];
        x[i] = s / L[i * 5 + i];
    }
    return 1;
}

int orc_dm_solve(int Nv, double vlo, double dv, double cs, const double m[5],
                 double a[5], double tol, int maxit, double *res)
{
    double dv3 = dv * dv * dv;
    double r = INFINITY;
    int pass;
    for (pass = 1; pass <= maxit + 1; ++pass) {
        double J[25] = {0};
        for (int kv = 0; kv < Nv; ++kv) {
            double zz = (vlo + kv * dv) / cs;
            for (int jv = 0; jv < Nv; ++jv) {
                double yy = (vlo + jv * dv) / cs;
                for (int iv = 0; iv < Nv; ++iv) {
                    double xx = (vlo + iv * dv) / cs;
                    double psi[5] = {1.0, xx, yy, zz, xx * xx + yy * yy + zz * zz};
                    double e = exp(a[0] + a[1] * xx + a[2] * yy + a[3] * zz + a[4] * psi[4]) * dv3;
                    for (int p = 0; p < 5; ++p)
                        for (int q = p; q < 5; ++q) J[p * 5 + q] += e * psi[p] * psi[q];
                }
            }
        }
        for (int p = 0; p < 5; ++p)
            for (int q = 0; q < p; ++q) J[p * 5 + q] = J[q * 5 + p];
        double F[5];
        r = 0.0;
        for (int p = 0; p < 5; ++p) {
            F[p] = J[p] - m[p];          /* row 0 of J holds the moments of M_d */
            r = fmax(r, fabs(F[p]));
        }
        r /= fabs(m[0]);
        if (r <= tol || pass == maxit + 1) break;
        double d[5];
        if (!chol5(J, F, d)) { pass = maxit + 1; break; }
        for (int p = 0; p < 5; ++p) a[p] -= d[p];
    }
    if (res) *res = r;
    return (r <= tol) ? pass : maxit + 2;
}

void orc_collision3d_dm(const orc_cloud *c, double *g, double dt, double tau,
                        double R, double tol, int maxit, const double *a_init,
                        double *a_out, int *passes, double *res)
{
    int Nv = c->Nv;
    size_t gs = (size_t)c->gstride;
    double dv3 = c->dv * c->dv * c->dv;
    double cs = orc_dm_scale(Nv, c->vlo, c->dv);
    double lam = dt / tau, coeff = 1.0 / (1.0 + lam);
    for (int i = 0; i < c->N; i++) {
        if (c->boundary[i]) continue;
        double *p = g + (size_t)i * gs;
        double m[5] = {0, 0, 0, 0, 0};
        for (int kv = 0; kv < Nv; ++kv) {
            double vz = c->vlo + kv * c->dv;
            for (int jv = 0; jv < Nv; ++jv) {
                double vy = c->vlo + jv * c->dv;
                for (int iv = 0; iv < Nv; ++iv) {
                    double vx = c->vlo + iv * c->dv;
                    double f = p[iv + Nv * (jv + Nv * kv)] * dv3;
                    m[0] += f;
                    m[1] += vx * f;
                    m[2] += vy * f;
                    m[3] += vz * f;
                    m[4] += (vx * vx + vy * vy + vz * vz) * f;
                }
            }
        }
        if (!(m[0] > 1e-30)) continue;
        double rho = m[0], ux = m[1] / rho, uy = m[2] / rho, uz = m[3] / rho;
        double t = (m[4] / rho - (ux * ux + uy * uy + uz * uz)) / (3.0 * R);
        double T = t > 1e-6 ? t : 1e-6;
        double ms[5] = {m[0], m[1] / cs, m[2] / cs, m[3] / cs, m[4] / (cs * cs)};
        double a[5], a0[5];
        orc_dm_guess(rho, ux, uy, uz, T, R, cs, a0);
        for (int q = 0; q < 5; ++q) a[q] = a_init ? a_init[(size_t)i * 5 + q] : a0[q];
        double r;
        int np = orc_dm_solve(Nv, c->vlo, c->dv, cs, ms, a, tol, maxit, &r);
        int ok = np <= maxit + 1;
        for (int q = 0; q < 5 && ok; ++q) ok = isfinite(a[q]);
        if (!ok) for (int q = 0; q < 5; ++q) a[q] = a0[q];   /* continuous fallback */
        if (a_out) for (int q = 0; q < 5; ++q) a_out[(size_t)i * 5 + q] = a[q];
        if (passes) passes[i] = np;
        if (res) res[i] = r;
        for (int kv = 0; kv < Nv; ++kv) {
            double zz = (c->vlo + kv * c->dv) / cs;
            for (int jv = 0; jv < Nv; ++jv) {
                double yy = (c->vlo + jv * c->dv) / cs;
                for (int iv = 0; iv < Nv; ++iv) {
                    double xx = (c->vlo + iv * c->dv) / cs;
                    double M = exp(a[0] + a[1] * xx + a[2] * yy + a[3] * zz
                                   + a[4] * (xx * xx + yy * yy + zz * zz));
                    int lin = iv + Nv * (jv + Nv * kv);
                    p[lin] = coeff * (p[lin] + lam * M);
                }
            }
        }
    }
}

/* ======================================================================== */
/* One time step                                                             */
/* ======================================================================== */

/* cudaStep (Solver/CudaBackend.cu:1505-1870) without its diagnostics. */
static double step_impl(const orc_cloud *c, orc_state *s, double dt, double tau,
                        double R, double h, int mlsOrder, double tol, int nofix,
                        int dm, double dm_tol, int dm_maxit)
{
    int N = c->N;
    for (int i = 0; i < N; i++) {
        s->limflag[i] = 0.0;
        if (s->mlsflag) s->mlsflag[i] = 0;
        if (s->mlsovr) s->mlsovr[i] = 0.0;
    }
    double massOld = orc_sum_mass(c, s->g);
    if (c->dim == 3) {
        orc_transport3d(c, s->g, s->gt, dt, h, mlsOrder, tol, 0, s->mlsflag, s->mlsovr);
        orc_positivity3d(c, s->g, s->gt, s->limflag);
        orc_nad3d(c, s->gt, s->rho, s->limflag);
    } else {
        orc_transport2d(c, s->g, s->gt, dt, h, mlsOrder);
        orc_positivity2d(c, s->gt, s->limflag);
    }
    memcpy(s->g, s->gt, sizeof(double) * (size_t)c->gstride * (size_t)N);   /* kCommit */
    double massNew = orc_sum_mass(c, s->g);
    if (massNew > 1e-30) orc_scale(c, s->g, massOld / massNew);
    if (c->dim == 3) {
        orc_wallbc3d(c, s->g, R, s->rho, s->ux, s->uy, s->uz, s->T);
        orc_moments3d(c, s->g, R, s->rho, s->ux, s->uy, s->uz, s->T);
        if (dm) orc_collision3d_dm(c, s->g, dt, tau, R, dm_tol, dm_maxit, NULL, NULL, NULL, NULL);
        else orc_collision3d(c, s->g, dt, tau, R, s->rho, s->ux, s->uy, s->uz, s->T);
    } else {
        orc_wallbc2d(c, s->g, R, s->rho, s->ux, s->uy, s->T);
        orc_moments2d(c, s->g, R, s->rho, s->ux, s->uy, s->T);
        orc_collision2d(c, s->g, dt, tau, R, s->rho, s->ux, s->uy, s->T);
    }
    if (!nofix) {
        double mColl = orc_sum_mass(c, s->g);
        if (mColl > 1e-30) orc_scale(c, s->g, massOld / mColl);
    }
    if (c->dim == 3) orc_moments3d(c, s->g, R, s->rho, s->ux, s->uy, s->uz, s->T);
    else orc_moments2d(c, s->g, R, s->rho, s->ux, s->uy, s->T);
    return massOld;
}

double orc_step(const orc_cloud *c, orc_state *s, double dt, double tau,
                double R, double h, int mlsOrder, double tol, int nofix)
{
    return step_impl(c, s, dt, tau, R, h, mlsOrder, tol, nofix, 0, 0.0, 0);
}

double orc_step_dm(const orc_cloud *c, orc_state *s, double dt, double tau,
                   double R, double h, int mlsOrder, double tol, int nofix,
                   double dm_tol, int dm_maxit)
{
    return step_impl(c, s, dt, tau, R, h, mlsOrder, tol, nofix, 1, dm_tol, dm_maxit);
}

/* ======================================================================== */
/* The reference's CPU step with its MOOD limiter                           */
/* ======================================================================== */

/* The positivity repair inside the CPU transport (MLSTransport/MLSTransport.cpp
 * :189-200 in 3-D, :287-307 in 2-D): negative values to zero, the rest scaled
 * so that the mass is kept; in 2-D on the g1 layer, with g2 following g1.
 * Returns 1 when it fired. */
static int positivity_zero(const orc_cloud *c, double *p)
{
    const int n = (c->dim == 3) ? c->gstride : c->Nv * c->Nv;
    double mPos = 0.0, mNeg = 0.0;
    for (int k = 0; k < n; k++) {
        if (p[k] >= 0.0) mPos += p[k];
        else mNeg += p[k];
    }
    if (!(mNeg < -1.0e-30 && mPos > 1.0e-30 && (mPos + mNeg) > 0.0)) return 0;
    const double sc = (mPos + mNeg) / mPos;
    for (int k = 0; k < n; k++) {
        if (c->dim == 3) p[k] = (p[k] < 0.0) ? 0.0 : p[k] * sc;
        else if (p[k] < 0.0) { p[k] = 0.0; p[k + n] = 0.0; }
        else { p[k] *= sc; p[k + n] *= sc; }
    }
    return 1;
}

/* checkTransportPAD (MLSTransport/MLSTransport.cpp:324-379) for one interior
 * particle: the transported row is inadmissible when its mass is not finite
 * or not positive, or when its density leaves the neighbours' current range
 * widened by 0.1 of the range (at least 1e-12) plus 1e-9. */
static int transport_flag(const orc_cloud *c, int i, const double *p, const double *rho)
{
    const int n = (c->dim == 3) ? c->Nv * c->Nv * c->Nv : c->Nv * c->Nv;
    const double dvN = (c->dim == 3) ? c->dv * c->dv * c->dv : c->dv * c->dv;
    double m = 0.0;
    for (int k = 0; k < n; k++) {
        if (!isfinite(p[k])) return 1;
        m += p[k];
    }
    if (m <= 0.0) return 1;
    const double rn = m * dvN;
    double lo = INFINITY, hi = -INFINITY;
    for (int t = 0; t < c->nneigh[i]; t++) {
        const double q = rho[c->neigh[(size_t)i * ORC_MAX_NEIGH + t]];
        lo = fmin(lo, q);
        hi = fmax(hi, q);
    }
    if (!(isfinite(lo) && isfinite(hi))) return 0;
    const double slack = 0.1 * fmax(hi - lo, 1.0e-12) + 1.0e-9;
    return (rn < lo - slack || rn > hi + slack);
}

/* padCheck (MLSTransport/MLSBasis.hpp:208-222) of an interior particle. */
static int pad_fails(const orc_cloud *c, const double *p, double rho, double T)
{
    if (rho < 1.0e-10 || T < 1.0e-6) return 1;
    const int n = (c->dim == 3) ? c->Nv * c->Nv * c->Nv : c->Nv * c->Nv;
    for (int k = 0; k < n; k++)
        if (p[k] < -1.0e-14) return 1;
    return 0;
}

static void moments_of(const orc_cloud *c, orc_state *s, double R)
{
    if (c->dim == 3) orc_moments3d(c, s->g, R, s->rho, s->ux, s->uy, s->uz, s->T);
    else orc_moments2d(c, s->g, R, s->rho, s->ux, s->uy, s->T);
}

static void collide(const orc_cloud *c, orc_state *s, double dt, double tau, double R)
{
    if (c->dim == 3) orc_collision3d(c, s->g, dt, tau, R, s->rho, s->ux, s->uy, s->uz, s->T);
    else orc_collision2d(c, s->g, dt, tau, R, s->rho, s->ux, s->uy, s->T);
}

/* The time loop body of runSolver with the MOOD limiter (Solver/Solver.cpp
 * :641-693): transport with the positivity repair; checkTransportPAD, and
 * the flagged particles transported again with the linear basis; commit; one
 * mass rescale; wall, moments, collision, moments; moodCheck
 * (MomentCalculations/Conservation.cpp:16-38) and, if any interior particle
 * fails padCheck, a second collision of every interior particle and the
 * moments again. The MLS solve is this oracle's (the GPU path's Cholesky);
 * the CPU code solves by Gauss elimination, equal to round-off on these
 * stencils. limflag: +1 positivity repair, +2 transported again, +4 failed
 * the post-collision check. */
double orc_step_mood(const orc_cloud *c, orc_state *s, double dt, double tau, double R,
                     double h, int mlsOrder, double tol)
{
    const int N = c->N;
    const size_t gs = (size_t)c->gstride;
    for (int i = 0; i < N; i++) {
        s->limflag[i] = 0.0;
        if (s->mlsflag) s->mlsflag[i] = 0;
        if (s->mlsovr) s->mlsovr[i] = 0.0;
    }
    const double massOld = orc_sum_mass(c, s->g);
    if (c->dim == 3) orc_transport3d(c, s->g, s->gt, dt, h, mlsOrder, tol, 0, NULL, NULL);
    else orc_transport2d(c, s->g, s->gt, dt, h, mlsOrder);
    for (int i = 0; i < N; i++)
        if (positivity_zero(c, s->gt + (size_t)i * gs)) s->limflag[i] = 1.0;
    int nf = 0;
    for (int i = 0; i < N; i++)
        if (!c->boundary[i] && transport_flag(c, i, s->gt + (size_t)i * gs, s->rho)) {
            s->limflag[i] += 2.0;
            nf++;
        }
    if (nf > 0) {
        double *gl = (double *)malloc(sizeof(double) * gs * (size_t)N);
        if (c->dim == 3) orc_transport3d(c, s->g, gl, dt, h, 1, tol, 0, NULL, NULL);
        else orc_transport2d(c, s->g, gl, dt, h, 1);
        for (int i = 0; i < N; i++) {
            if (!((int)s->limflag[i] & 2)) continue;
            double *p = s->gt + (size_t)i * gs;
            memcpy(p, gl + (size_t)i * gs, sizeof(double) * gs);
            if (positivity_zero(c, p) && !((int)s->limflag[i] & 1)) s->limflag[i] += 1.0;
        }
        free(gl);
    }
    memcpy(s->g, s->gt, sizeof(double) * gs * (size_t)N);          /* mlsTransportCommit */
    const double massNew = orc_sum_mass(c, s->g);
    if (massNew > 1e-30) orc_scale(c, s->g, massOld / massNew);
    if (c->dim == 3) orc_wallbc3d(c, s->g, R, s->rho, s->ux, s->uy, s->uz, s->T);
    else orc_wallbc2d(c, s->g, R, s->rho, s->ux, s->uy, s->T);
    moments_of(c, s, R);
    collide(c, s, dt, tau, R);
    moments_of(c, s, R);
    int fail = 0;
    for (int i = 0; i < N; i++)
        if (!c->boundary[i] && pad_fails(c, s->g + (size_t)i * gs, s->rho[i], s->T[i])) {
            s->limflag[i] += 4.0;
            fail = 1;
        }
    if (fail) {
        collide(c, s, dt, tau, R);
        moments_of(c, s, R);
    }
    return massOld;
}

/* ======================================================================== */
/* Body loads                                                                */
/* ======================================================================== */

/* computeBodyLoads (RigidBody/BodyDynamics.cpp:17-97) for one body. */
void orc_body_loads(const orc_cloud *c, const double *g, const int *bodyid,
                    int body, double bcx, double bcy, double bcz, double area,
                    double *out)
{
    int Nv = c->Nv;
    size_t gs = (size_t)c->gstride;
    double Fx = 0, Fy = 0, Fz = 0, Tx = 0, Ty = 0, Tz = 0;
    for (int i = 0; i < c->N; i++) {
        if (!c->boundary[i] || bodyid[i] != body) continue;
        const double *p = g + (size_t)i * gs;
        double cdx = c->x[i] - bcx, cdy = c->y[i] - bcy, cdz = c->z[i] - bcz;
        if (c->dim == 3) {
            double dv3 = c->dv * c->dv * c->dv;
            double pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0;
            for (int kv = 0; kv < Nv; ++kv) {
                double cz = (c->vlo + kv * c->dv) - c->wUz[i];
                for (int jv = 0; jv < Nv; ++jv) {
                    double cy = (c->vlo + jv * c->dv) - c->wUy[i];
                    for (int iv = 0; iv < Nv; ++iv) {
                        double cx = (c->vlo + iv * c->dv) - c->wUx[i];
                        double f = p[iv + Nv * (jv + Nv * kv)] * dv3;
                        pxx += f * cx * cx; pyy += f * cy * cy;
                        pzz += f * cz * cz;
                        pxy += f * cx * cy; pxz += f * cx * cz;
                        pyz += f * cy * cz;
                    }
                }
            }
            double tx = -(c->wnx[i] * pxx + c->wny[i] * pxy + c->wnz[i] * pxz);
            double ty = -(c->wnx[i] * pxy + c->wny[i] * pyy + c->wnz[i] * pyz);
            double tz = -(c->wnx[i] * pxz + c->wny[i] * pyz + c->wnz[i] * pzz);
            Fx += area * tx; Fy += area * ty; Fz += area * tz;
            Tx += area * (cdy * tz - cdz * ty);
            Ty += area * (cdz * tx - cdx * tz);
            Tz += area * (cdx * ty - cdy * tx);
        } else {
            double dv2 = c->dv * c->dv;
            double phiX = 0.0, phiY = 0.0, phiXY = 0.0;
            for (int jv = 0; jv < Nv; ++jv) {
                double cy = (c->vlo + jv * c->dv) - c->wUy[i];
                for (int iv = 0; iv < Nv; ++iv) {
                    double cx = (c->vlo + iv * c->dv) - c->wUx[i];
                    double g1 = p[iv + Nv * jv];
                    phiX += g1 * dv2 * cx * cx;
                    phiY += g1 * dv2 * cy * cy;
                    phiXY += g1 * dv2 * cx * cy;
                }
            }
            double valX = -c->wnx[i] * phiX - c->wny[i] * phiXY;
            double valY = -c->wnx[i] * phiXY - c->wny[i] * phiY;
            Fx += area * valX;
            Fy += area * valY;
            Tz += area * (cdx * valY - cdy * valX);
        }
    }
    out[0] = Fx; out[1] = Fy; out[2] = Fz; out[3] = Tx; out[4] = Ty; out[5] = Tz;
}

/* ---- Cloud management around moving bodies --------------------------------- */

/* isInObject / isInObject3 (RigidBody/Geometry.hpp): Square :179-186 and
 * Circle :233-237 test (x, y) only; Sphere :105-109 and Cube :287-293 take
 * z = cz in 2-D (their isInObject). */
int orc_in_body(const orc_body *b, double x, double y, double z, int d3)
{
    switch (b->shape)
    {
    case 1:
    {
        double xC = x - b->c[0], yC = y - b->c[1];
        double D = b->o[0] * b->o[0] + b->o[1] * b->o[1];
        double xp = (b->o[0] * xC + b->o[1] * yC) / D;
        double yp = (b->o[1] * xC - b->o[0] * yC) / D;
        return (fabs(xp) <= b->size / 2) && (fabs(yp) <= b->size / 2);
    }
    case 2:
    {
        double dx = x - b->c[0], dy = y - b->c[1];
        return dx * dx + dy * dy <= b->size * b->size;
    }
    case 3:
    {
        double zz = d3 ? z : b->c[2];
        double dx = x - b->c[0], dy = y - b->c[1], dz = zz - b->c[2];
        return dx * dx + dy * dy + dz * dz <= b->size * b->size;
    }
    case 4:
    {
        double zz = d3 ? z : b->c[2];
        double dx = x - b->c[0], dy = y - b->c[1], dz = zz - b->c[2];
        const double *R = b->rm;
        double bx = R[0] * dx + R[3] * dy + R[6] * dz;   /* R^T (world -> body) */
        double by = R[1] * dx + R[4] * dy + R[7] * dz;
        double bz = R[2] * dx + R[5] * dy + R[8] * dz;
        return fabs(bx) <= b->size / 2 && fabs(by) <= b->size / 2 && fabs(bz) <= b->size / 2;
    }
    default:
        return 0;
    }
}

static int in_any_body(int nb, const orc_body *b, double x, double y, double z, int d3)
{
    for (int k = 0; k < nb; ++k)
        if (orc_in_body(&b[k], x, y, z, d3)) return 1;
    return 0;
}

/* |p - q|^2 as CloudManagement.cpp writes it (dz = 0 in 2-D) */
static double cm_d2(double px, double py, double pz, double qx, double qy, double qz, int d3)
{
    double dx = px - qx, dy = py - qy;
    double dz = d3 ? pz - qz : 0.0;
    return dx * dx + dy * dy + dz * dz;
}

/* manageRemove (RigidBody/CloudManagement.cpp:17-69) */
int orc_cloud_remove(int dim, int N, const double *x, const double *y, const double *z,
                     const unsigned char *bnd, const int *bodyid, int nb, const orc_body *b,
                     double dx, unsigned char *keep)
{
    const int d3 = dim == 3;
    const double hug2 = pow(0.4 * fabs(dx), 2);
    int removed = 0;
    for (int i = 0; i < N; ++i)
    {
        int drop = 0;
        if (!bnd[i] && nb > 0)
        {
            drop = in_any_body(nb, b, x[i], y[i], z[i], d3);
            for (int q = 0; q < N && !drop; ++q)      /* the surface points, in order */
                if (bnd[q] && bodyid[q] >= 0)
                    drop = cm_d2(x[i], y[i], z[i], x[q], y[q], z[q], d3) < hug2;
        }
        keep[i] = drop ? 0 : 1;
        removed += drop;
    }
    return removed;
}

/* mergeCloseParticles (RigidBody/CloudManagement.cpp:80-119) */
int orc_cloud_merge(int dim, int N, double *x, double *y, double *z, const unsigned char *bnd,
                    const int *neigh, const int *nneigh, double min_dist, int gstride,
                    double *g, unsigned char *dead)
{
    const double min2 = min_dist * min_dist;
    memset(dead, 0, (size_t)N);
    if (min2 <= 0.0) return 0;
    int merged = 0;
    for (int i = 0; i < N; ++i)
    {
        if (dead[i] || bnd[i]) continue;
        for (int k = 0; k < nneigh[i]; ++k)
        {
            int j = neigh[(size_t)i * ORC_MAX_NEIGH + k];
            if (j <= i || j >= N || dead[j]) continue;
            if (bnd[j]) continue;
            if (cm_d2(x[i], y[i], z[i], x[j], y[j], z[j], dim == 3) < min2)
            {
                x[i] = 0.5 * (x[i] + x[j]);
                y[i] = 0.5 * (y[i] + y[j]);
                if (dim == 3) z[i] = 0.5 * (z[i] + z[j]);
                for (int l = 0; l < gstride; ++l)
                    g[(size_t)i * gstride + l] =
                        0.5 * (g[(size_t)i * gstride + l] + g[(size_t)j * gstride + l]);
                dead[j] = 1;
                merged++;
                break;
            }
        }
    }
    return merged;
}

/* manageRefill (RigidBody/CloudManagement.cpp:121-275) */
int orc_cloud_refill(int dim, int N, const double *x, const double *y, const double *z,
                     const unsigned char *bnd, const int *bodyid, const double *mom, int nb,
                     const orc_body *b, double dx, const double *dom6, double radius, int Nv,
                     double vlo, double dv, double R, int gstride, int cap, double *out,
                     double *gout)
{
    if (nb <= 0) return 0;
    const int d3 = dim == 3;
    const double dxl = fabs(dx);
    const double hole2 = pow(0.8 * dxl, 2), clr2 = pow(0.6 * dxl, 2);
    double bx0 = 1e300, bx1 = -1e300, by0 = 1e300, by1 = -1e300, bz0 = 1e300, bz1 = -1e300;
    int any = 0;
    for (int q = 0; q < N; ++q)
    {
        if (!bnd[q] || bodyid[q] < 0) continue;
        bx0 = fmin(bx0, x[q]); bx1 = fmax(bx1, x[q]);
        by0 = fmin(by0, y[q]); by1 = fmax(by1, y[q]);
        if (d3) { bz0 = fmin(bz0, z[q]); bz1 = fmax(bz1, z[q]); }
        any = 1;
    }
    if (!any) return 0;
    const double margin = 3.5 * dxl;
    bx0 -= margin; bx1 += margin; by0 -= margin; by1 += margin;
    if (d3) { bz0 -= margin; bz1 += margin; }
    const int nxs = (int)llround((dom6[1] - dom6[0]) / dxl);
    const int nys = (int)llround((dom6[3] - dom6[2]) / dxl);
    const int nzs = d3 ? (int)llround((dom6[5] - dom6[4]) / dxl) : 0;
    int i0 = (int)ceil((bx0 - dom6[0]) / dxl), i1 = (int)floor((bx1 - dom6[0]) / dxl);
    int j0 = (int)ceil((by0 - dom6[2]) / dxl), j1 = (int)floor((by1 - dom6[2]) / dxl);
    int k0 = d3 ? (int)ceil((bz0 - dom6[4]) / dxl) : 0, k1 = d3 ? (int)floor((bz1 - dom6[4]) / dxl) : 0;
    if (i0 < 1) i0 = 1;
    if (i1 > nxs - 1) i1 = nxs - 1;
    if (j0 < 1) j0 = 1;
    if (j1 > nys - 1) j1 = nys - 1;
    if (d3 && k0 < 1) k0 = 1;
    if (d3 && k1 > nzs - 1) k1 = nzs - 1;
    const double r2 = radius * radius;
    int added = 0;
    for (int vk = k0; vk <= k1; ++vk)
        for (int vj = j0; vj <= j1; ++vj)
            for (int vi = i0; vi <= i1; ++vi)
            {
                double xm = dom6[0] + vi * dxl, ym = dom6[2] + vj * dxl;
                double zm = d3 ? dom6[4] + vk * dxl : 0.0;
                /* the cloud, then the particles added so far (P grows) */
                int hole = 1;
                for (int q = 0; q < N + added && hole; ++q)
                {
                    double qx = q < N ? x[q] : out[q - N], qy = q < N ? y[q] : out[cap + q - N];
                    double qz = q < N ? z[q] : out[2 * cap + q - N];
                    if (cm_d2(qx, qy, qz, xm, ym, zm, d3) < hole2) hole = 0;
                }
                if (!hole) continue;
                if (in_any_body(nb, b, xm, ym, zm, d3)) continue;
                int hugs = 0;
                for (int q = 0; q < N && !hugs; ++q)
                    if (bnd[q] && bodyid[q] >= 0)
                        hugs = cm_d2(x[q], y[q], z[q], xm, ym, zm, d3) < clr2;
                if (hugs) continue;
                double wsum = 0, rho = 0, ux = 0, uy = 0, uz = 0, T = 0;
                for (int q = 0; q < N + added; ++q)
                {
                    const int a = q >= N, k = q - N;
                    double qx = a ? out[k] : x[q], qy = a ? out[cap + k] : y[q];
                    double qz = a ? out[2 * cap + k] : z[q];
                    double d2 = cm_d2(qx, qy, qz, xm, ym, zm, d3);
                    if (d2 >= r2) continue;
                    double q2 = d2 / (radius * radius);     /* mlsWeightSq, MLSBasis.hpp:142 */
                    double w = q2 >= 1.0 ? 0.0 : (1.0 - q2) * (1.0 - q2);
                    wsum += w;
                    rho += w * (a ? out[3 * cap + k] : mom[5 * q + 0]);
                    ux += w * (a ? out[4 * cap + k] : mom[5 * q + 1]);
                    uy += w * (a ? out[5 * cap + k] : mom[5 * q + 2]);
                    uz += w * (a ? out[6 * cap + k] : mom[5 * q + 3]);
                    T += w * (a ? out[7 * cap + k] : mom[5 * q + 4]);
                }
                if (wsum < 1e-30) continue;
                rho /= wsum; ux /= wsum; uy /= wsum; uz /= wsum; T /= wsum;
                if (added >= cap) return -1;
                const int k = added;
                out[k] = xm; out[cap + k] = ym; out[2 * cap + k] = zm;
                out[3 * cap + k] = rho; out[4 * cap + k] = ux; out[5 * cap + k] = uy;
                out[6 * cap + k] = d3 ? uz : 0.0; out[7 * cap + k] = T;
                double *gr = gout + (size_t)k * gstride;
                if (d3)
                {
                    for (int kv = 0; kv < Nv; ++kv)
                        for (int jv = 0; jv < Nv; ++jv)
                            for (int iv = 0; iv < Nv; ++iv)
                                gr[iv + Nv * (jv + Nv * kv)] = orc_maxwellian3d(
                                    vlo + iv * dv, vlo + jv * dv, vlo + kv * dv, ux, uy, uz, T,
                                    rho, R);
                }
                else
                {
                    for (int jv = 0; jv < Nv; ++jv)
                        for (int iv = 0; iv < Nv; ++iv)
                        {
                            double g1 = orc_maxwellian_g1(vlo + iv * dv, vlo + jv * dv, ux, uy, T,
                                                          rho, R);
                            gr[iv + Nv * jv] = g1;
                            gr[iv + Nv * jv + Nv * Nv] = 0.5 * R * T * g1;
                        }
                }
                added++;
            }
    return added;
}
