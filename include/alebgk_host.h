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

/* ---- Body cases -------------------------------------------------------------
 * makeBodies (TestCases/TestCases.cpp:251-294): the case's body (one for the
 * body cases, none otherwise) at rest, scaled by body_scale (--body-scale). */
int alh_make_bodies(const alh_case *cs, double body_scale, ab_rigid_body *bodies,
                    int *nbodies);
/* Particles of a body case: the lattice minus the particles inside a body or
 * within 0.4 dx of its surface, plus the surface particles
 * (embedRigidBodies, Initialization/BodyEmbedding.cpp:18-104). */
int alh_body_case_size(const alh_case *cs, double body_scale, int *n);
/* generateParticles + embedRigidBodies + initWallData + applyLidProfile +
 * applyIC for a body case (main.cpp:331-345, Solver.cpp applyIC), arrays
 * sized by alh_body_case_size: as alh_build_lattice, plus body_id (N) and the
 * bodies (up to 1; *nbodies set). geo (may be NULL) = dx, dy, dz, minDist
 * (CalcParameters). cfg->nx/ny/nz are 0: the cloud is not a lattice. */
int alh_build_body_case(const alh_case *cs, double body_scale, int precision, ab_config *cfg,
                        double *x, double *y, double *z, unsigned char *boundary, double *wall,
                        double *mom, double *g, int *body_id, ab_rigid_body *bodies,
                        int *nbodies, double geo[4]);

/* ---- Run outputs (Core/Output.hpp:17-115, Solver/Solver.cpp:96-121) ------
 * Written with the reference's stream formatting: byte-identical files for
 * the same values. */
/* computeDiagnostics: means over the interior (non-boundary) particles. */
typedef struct alh_diag {
    double rho_avg, KE, T_avg, p_avg;
    int mood_count, n;
} alh_diag;
int alh_diagnostics(int n, const unsigned char *boundary, const double *rho, const double *ux,
                    const double *uy, const double *uz, const double *T, const double *p,
                    const unsigned char *mood, alh_diag *d);
/* writeVTK: out_dir/output_<step, 6 digits>.vtk, ASCII unstructured grid of
 * the points with rho, T, p, Nv_local, moodFlag, boundary and velocity. */
int alh_write_vtk(const char *out_dir, int step, double t, int n, const double *x,
                  const double *y, const double *z, const double *rho, const double *T,
                  const double *p, int nv_local, const unsigned char *mood,
                  const unsigned char *boundary, const double *ux, const double *uy,
                  const double *uz);
/* writeTimeseries: one CSV row; first != 0 truncates and writes the header. */
int alh_write_timeseries(const char *path, int first, double t, const alh_diag *d);
/* printStep's line (no newline) into out. */
int alh_format_step(int step, double t, double dt, const alh_diag *d, char *out, int cap);
/* writeBodyLog: one bodies.csv row per body; first != 0 writes the header. */
int alh_write_bodies(const char *path, int first, double t, int dim, int nbodies,
                     const ab_rigid_body *bodies);

/* ---- Stability guard (Core/Stability.hpp:46-123) -------------------------
 * Stops a run whose diagnostics leave physical bounds: mean T outside
 * [0.60, 1.30] Tw, mean rho outside [0.20, 5.00] rho0, KE above 100 x its
 * first positive value, or anything non-finite. ALEBGK_GUARD_TMIN / _TMAX /
 * _RHOMIN / _RHOMAX / _KEMAX override the factors; ALEBGK_NO_GUARD=1 turns
 * it off (read at alh_guard_init). */
typedef struct alh_guard {
    double Tw, rho0, ke_ref;
    int have_ke, enabled;
    double t_lo, t_hi, rho_lo, rho_hi, ke_hi;
} alh_guard;
int alh_guard_init(alh_guard *g, double wall_temp, double rho0);
/* 1: diverged (the reason in why), 0: within bounds. */
int alh_guard_check(alh_guard *g, int step, double t, const alh_diag *d, char *why, int cap);

/* ---- Restart files (Core/Checkpoint.hpp:37-234, format v1) -------------
 * The reference's checkpoint, read and written byte for byte: a 56-byte
 * header, N particle records of 200 bytes, N x gstride FP64 g, nbodies body
 * records of 232 bytes, little-endian. */
typedef struct alh_ckpt_info {
    int dim;
    long long n, gstride, step;
    double t;
    int nbodies;
} alh_ckpt_info;

/* The record fields as arrays (N entries per field unless noted). */
typedef struct alh_ckpt_fields {
    double *x, *y, *z;
    double *mom;           /* N x 6: rho, ux, uy, uz, T, p */
    double *mesh;          /* N x 3: meshVx, meshVy, meshVz */
    double *wall;          /* N x 7: wnx, wny, wnz, wUx, wUy, wUz, wT */
    double *grid;          /* N x 3: vlo, vhi, dv_local */
    int *ints;             /* N x 4: Nv_local, dim, bodyId, voxel */
    unsigned char *flags;  /* N x 3: boundary, validg, moodFlag */
    double *g;             /* N x gstride */
    double *bodies;        /* nbodies x 30: cx cy cz velx vely velz angularVel
                              omx omy omz forcex forcey forcez torque tqx tqy
                              tqz ox oy, Rm[9], dim, rotates */
} alh_ckpt_fields;

/* The header only (AB_ERR_ARG: unreadable or not a v1 file). */
int alh_ckpt_info_read(const char *path, alh_ckpt_info *info);
/* ckpt::restore (Checkpoint.hpp:161): a file whose N, gstride or body count
 * differ from *expect is refused with AB_ERR_UNSUPPORTED; a short file is
 * AB_ERR_ARG. */
int alh_ckpt_read(const char *path, const alh_ckpt_info *expect, alh_ckpt_fields *out);
/* ckpt::save (Checkpoint.hpp:78): written to path.tmp, then renamed. */
int alh_ckpt_write(const char *path, const alh_ckpt_info *info, const alh_ckpt_fields *in);

#ifdef __cplusplus
}
#endif
#endif
