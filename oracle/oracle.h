/*
 * oracle.h -- CPU FP64 restatement of the reference ALE-BGK time step.
 *
 * TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load liboracle.so, and only as the checker. The
 * product path (alebgk4paper_amd) never links or calls it.
 *
 * Semantics follow the reference's GPU path, Solver/CudaBackend.cu
 * (cudaStep :1505-1870), which is the component the HIP library replaces:
 * Cholesky with a relative pivot tolerance and a Shepard fallback in 3-D,
 * pivoted Gauss in 2-D, positivity by fallback to the pre-transport value,
 * the NAD density clamp, and two global mass rescales per step. Setup
 * (lattice, neighbour lists, wall state, initial condition) follows the
 * reference CPU sources it shares with the GPU path.
 *
 * Compiled with -ffp-contract=off so a*b+c is never fused: the expressions
 * evaluate in the order the reference writes them.
 *
 * Layout: structure of arrays; g is particle-major, g[i*gstride + lin] with
 * lin = iv + Nv*(jv + Nv*kv) in 3-D; in 2-D g1 at lin = iv + Nv*jv and g2 at
 * lin + Nv*Nv.
 */
#ifndef ALEBGK_ORACLE_H
#define ALEBGK_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_NEIGH 128

/* Voxel and lattice geometry (the CalcParameters fields generateParticles
 * fills, Initialization/ParticleGeneration.cpp:47-114). */
typedef struct {
    int dim;
    int Nx, Ny, Nz;
    double xleft, xright, ybottom, ytop, zfront, zback;
    double dx, dy, dz;
    double radius, minDist;
    double hBox, vBox, wBox;
    int nbhBox, nbvBox, nbwBox, nvox;
} orc_geom;

/* Read-only view of a cloud. */
typedef struct {
    int dim, N, Nv, gstride;
    double vlo, dv;
    const double *x, *y, *z;
    const unsigned char *boundary;
    const double *wnx, *wny, *wnz, *wUx, *wUy, *wUz, *wT;
    const double *mvx, *mvy, *mvz;        /* ALE mesh velocity, may be NULL */
    const int *neigh;                     /* N * ORC_MAX_NEIGH */
    const int *nneigh;
} orc_cloud;

/* Mutable per-step state. */
typedef struct {
    double *g, *gt;                       /* N * gstride each */
    double *rho, *ux, *uy, *uz, *T;       /* N each */
    double *limflag;                      /* N: +1 positivity, +2 NAD */
    int *mlsflag;                         /* N: |1 degenerate, |2 extrapolation */
    double *mlsovr;                       /* N: overshoot L1, density units */
} orc_state;

/* ---- setup --------------------------------------------------------------- */
int  orc_lattice(orc_geom *geo, double *x, double *y, double *z,
                 unsigned char *boundary, int *voxel);
int  orc_neighbours(const orc_geom *geo, int N, const double *x,
                    const double *y, const double *z, int *voxel,
                    int *neigh, int *nneigh);
void orc_wall_normal(const orc_geom *geo, double x, double y, double z,
                     double Tw, double TwTop, const double *Uface /* 6x3:
                     left,right,top,bottom,front,back */, double *out7);
void orc_wall_state(const orc_geom *geo, int N, const double *x,
                    const double *y, const double *z,
                    const unsigned char *boundary, const int *bodyid,
                    double Tw, const double *Uface,
                    double *wnx, double *wny, double *wnz, double *wUx,
                    double *wUy, double *wUz, double *wT);
void orc_lid_parabolic(const orc_geom *geo, int N, const double *x,
                       const double *y, const unsigned char *boundary,
                       const int *bodyid, double peak, double *wUx);
void orc_initial_state(int dim, int N, int Nv, double vlo, double dv,
                       double rho0, double T0, double R, double *g);

/* ---- kernels (one reference kernel each) --------------------------------- */
double orc_maxwellian3d(double vx, double vy, double vz, double ux,
                        double uy, double uz, double T, double rho, double R);
double orc_maxwellian_g1(double vx, double vy, double ux, double uy,
                         double T, double rho, double R);

void orc_transport3d(const orc_cloud *c, const double *g, double *gt,
                     double dt, double h, int mlsOrder, double tol,
                     int clamp, int *mlsflag, double *mlsovr);
void orc_transport2d(const orc_cloud *c, const double *g, double *gt,
                     double dt, double h, int mlsOrder);
void orc_positivity3d(const orc_cloud *c, const double *g, double *gt,
                      double *limflag);
void orc_positivity2d(const orc_cloud *c, double *gt, double *limflag);
void orc_nad3d(const orc_cloud *c, double *gt, const double *rho,
               double *limflag);
double orc_sum_mass(const orc_cloud *c, const double *g);
void orc_scale(const orc_cloud *c, double *g, double s);
void orc_wallbc3d(const orc_cloud *c, double *g, double R, double *rho,
                  double *ux, double *uy, double *uz, double *T);
void orc_wallbc2d(const orc_cloud *c, double *g, double R, double *rho,
                  double *ux, double *uy, double *T);
void orc_moments3d(const orc_cloud *c, const double *g, double R,
                   double *rho, double *ux, double *uy, double *uz,
                   double *T);
void orc_moments2d(const orc_cloud *c, const double *g, double R,
                   double *rho, double *ux, double *uy, double *T);
void orc_collision3d(const orc_cloud *c, double *g, double dt, double tau,
                     double R, const double *rho, const double *ux,
                     const double *uy, const double *uz, const double *T);
void orc_collision2d(const orc_cloud *c, double *g, double dt, double tau,
                     double R, const double *rho, const double *ux,
                     const double *uy, const double *T);

/* Discrete-Maxwellian collision (not in the reference). The BGK target is
 * M_d(v) = exp(a0 + a1 xi_x + a2 xi_y + a3 xi_z + a4 |xi|^2), xi = v / cs,
 * whose grid moments sum_k M_d(v_k) psi_k dv^3, psi = (1, xi, |xi|^2), equal
 * those of g, so the collision conserves mass, momentum and energy on the
 * velocity grid exactly (discrete-velocity BGK, Mieussens 2000). Newton in
 * FP64 on the 5 unknowns from the continuous Maxwellian of (rho, u, T).
 * cs: velocity scale (largest |v| of the grid). */
double orc_dm_scale(int Nv, double vlo, double dv);
void orc_dm_guess(double rho, double ux, double uy, double uz, double T,
                  double R, double cs, double a[5]);
/* m: target moments in scaled form (sum g psi dv^3). Each pass over the
 * grid evaluates the residual F = sum M_d psi dv^3 - m and the Jacobian at
 * the current a; a pass with max|F_j| <= tol * m0 ends the solve. Returns
 * the number of passes (> maxit when it did not converge), *res the last
 * relative residual. */
int orc_dm_solve(int Nv, double vlo, double dv, double cs, const double m[5],
                 double a[5], double tol, int maxit, double *res);
/* Collision with M_d for every interior particle. a_init (N x 5, may be
 * NULL: continuous guess) warm-starts Newton; a_out (N x 5), passes (N) and
 * res (N) may be NULL. */
void orc_collision3d_dm(const orc_cloud *c, double *g, double dt, double tau,
                        double R, double tol, int maxit, const double *a_init,
                        double *a_out, int *passes, double *res);

/* The whole cudaStep sequence. Returns massOld. nofix disables the second
 * mass restore (ALEBGK_NO_MASS_FIX). */
double orc_step(const orc_cloud *c, orc_state *s, double dt, double tau,
                double R, double h, int mlsOrder, double tol, int nofix);
/* The same step with the discrete-Maxwellian collision (3-D). */
double orc_step_dm(const orc_cloud *c, orc_state *s, double dt, double tau,
                   double R, double h, int mlsOrder, double tol, int nofix,
                   double dm_tol, int dm_maxit);

/* The reference's CPU step with its default limiter, MOOD (Solver/Solver.cpp
 * :641-693): positivity to zero, checkTransportPAD with a linear
 * re-transport of the flagged particles, one mass rescale, and a second
 * collision when moodCheck fails. Returns massOld; limflag +1 positivity,
 * +2 transported again, +4 failed the post-collision check. */
double orc_step_mood(const orc_cloud *c, orc_state *s, double dt, double tau, double R,
                     double h, int mlsOrder, double tol);

/* Surface-stress loads of one body, computeBodyLoads
 * (RigidBody/BodyDynamics.cpp:17-97). out6 = Fx,Fy,Fz,Tx,Ty,Tz. */
void orc_body_loads(const orc_cloud *c, const double *g, const int *bodyid,
                    int body, double cx, double cy, double cz, double area,
                    double *out6);

/* ---- Cloud management around moving bodies ---------------------------------
 * RigidBody/CloudManagement.cpp. A body (RigidBody/Geometry.hpp): shape 1
 * square, 2 circle, 3 sphere, 4 cube; centre c, size (side or radius), the
 * square's orientation o, the cube's body-to-world rotation rm (row-major). */
typedef struct {
    int shape;
    double c[3];
    double size;
    double o[2];
    double rm[9];
} orc_body;

/* isInObject / isInObject3 (Geometry.hpp:105-109, 179-186, 233-237, 287-293). */
int orc_in_body(const orc_body *b, double x, double y, double z, int d3);

/* manageRemove (:17-69): keep[i] = 0 for a removed particle; returns the
 * number removed. */
int orc_cloud_remove(int dim, int N, const double *x, const double *y, const double *z,
                     const unsigned char *bnd, const int *bodyid, int nb, const orc_body *b,
                     double dx, unsigned char *keep);

/* mergeCloseParticles (:80-119) on the current neighbour lists: moves the
 * merging particles (x, y, z in place) and averages their g; dead[j] = 1 for
 * a particle merged away. Returns the number of merges. */
int orc_cloud_merge(int dim, int N, double *x, double *y, double *z, const unsigned char *bnd,
                    const int *neigh, const int *nneigh, double min_dist, int gstride,
                    double *g, unsigned char *dead);

/* manageRefill (:121-275): the particles it appends, at most cap. mom holds
 * the cloud's rho ux uy uz T (5 per particle). out: 8 x cap values (x, y, z,
 * rho, ux, uy, uz, T of each added particle, field-major), gout: cap rows of
 * gstride. Returns the number added, -1 when cap is too small. */
int orc_cloud_refill(int dim, int N, const double *x, const double *y, const double *z,
                     const unsigned char *bnd, const int *bodyid, const double *mom, int nb,
                     const orc_body *b, double dx, const double *dom6, double radius, int Nv,
                     double vlo, double dv, double R, int gstride, int cap, double *out,
                     double *gout);

#ifdef __cplusplus
}
#endif
#endif
