/*
 * alebgk_host.h -- C API of the host library (libalebgk_host.so): the
 * reference's case table, lattice generation, wall state and initial state,
 * producing the arrays ab_upload takes.
 *
 * Mirrors TestCases/TestCases.{hpp,cpp} (parseCaseName, configureCase,
 * applyLidProfile), Initialization/ParticleGeneration.cpp (generateParticles),
 * BoundaryConditions/WallState.cpp (initWallData) and
 * Initialization/InitialConditions.cpp (applyIC).
 */
#ifndef ALEBGK_HOST_H
#define ALEBGK_HOST_H

#include "alebgk_dev.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ALH_LID_UNIFORM_TOP 0     /* uniform Ux on y = L (cavity2d) */
#define ALH_LID_UNIFORM_FRONT 1   /* uniform Ux on z = zfront (cavity3d) */
#define ALH_LID_PARABOLIC_TOP 2   /* parabolic Ux on y = L (body cases) */

/* One case of TestCases.cpp:56-228 after command-line overrides. */
typedef struct alh_case {
    int id;                 /* 1..6 (TestCase) */
    char name[32];
    int dim;
    int has_body;
    int lid;                /* ALH_LID_* */
    double lid_speed;       /* uniform value or parabolic peak, m/s */
    int nx, ny, nz, nv;
    double vmin, vmax;
    double dt, tfinal, tau;
    double rho0, Tw, R, d_hs;
    int save_every;
    int mls_order;          /* SimParameters::mlsOrder (reference default 1) */
    /* DomainBoundary: xleft, xright, ytop, ybottom, zfront, zback */
    double dom[6];
} alh_case;

/* parseCaseName + configureCase (TestCases.cpp:91-228): token is the case
 * number or its name. nx/nv <= 0 keep the case defaults (--Nx / --Nv). */
int alh_configure(const char *token, int nx, int nv, alh_case *out);

/* Particle count of the case's lattice (no body). */
int alh_lattice_size(const alh_case *cs, int *n);

/* generateParticles + initWallData + applyLidProfile + applyIC for a plain
 * (body-free) case. Fills the ab_config (lattice, voxel grid, radius, gas)
 * and the arrays, each sized for alh_lattice_size particles:
 *   x, y, z, boundary              N
 *   wall                           7N  (wnx, wny, wnz, wUx, wUy, wUz, wT)
 *   mom                            5N  (rho, ux, uy, uz, T)
 *   g                              N * gstride, particle-major
 * precision goes into cfg->precision. */
int alh_build_lattice(const alh_case *cs, int precision, ab_config *cfg, double *x, double *y,
                      double *z, unsigned char *boundary, double *wall, double *mom, double *g);

#ifdef __cplusplus
}
#endif
#endif
