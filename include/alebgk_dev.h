/*
 * alebgk_dev.h -- C ABI of the MI355X (gfx950) ALE-BGK time-step library
 * (libalebgk_hip.so).
 *
 * Replaces the reference's GPU driver boundary, the free functions declared
 * at Solver/Solver.cpp:67-89 and defined in Solver/CudaBackend.cu, plus the
 * host neighbour search (Initialization/NeighbourSearch.cpp:17) and the
 * body-load reduction (RigidBody/BodyDynamics.cpp:17).
 *
 * Conventions
 *   - Every call returns an int status: AB_OK (0) or an AB_ERR_* code; the
 *     message is ab_last_error(ctx). Nothing calls exit().
 *   - Host arrays are borrowed for the duration of the call only; device
 *     buffers are owned by the context.
 *   - One context = one cloud = one HIP stream; calls on one context must be
 *     serialised by the caller. There are no hidden globals.
 *   - FP64 at the interface whatever the internal precision.
 *   - g is particle-major: g[i*gstride + lin], lin = iv + Nv*(jv + Nv*kv) in
 *     3-D (gstride = Nv^3); in 2-D g1 at lin = iv + Nv*jv, g2 at lin + Nv^2
 *     (gstride = 2*Nv^2). Neighbour lists are N x AB_MAX_NEIGH int32.
 */
#ifndef ALEBGK_DEV_H
#define ALEBGK_DEV_H

#ifdef __cplusplus
extern "C" {
#endif

#define AB_MAX_NEIGH 128          /* Core/Types.hpp:22 MAX_NEIGH */

#define AB_OK 0
#define AB_ERR_ARG 1              /* invalid argument */
#define AB_ERR_HIP 2              /* HIP runtime error (message has details) */
#define AB_ERR_STATE 3            /* call out of order (e.g. step before upload) */
#define AB_ERR_UNSUPPORTED 4      /* configuration this build does not run */
#define AB_ERR_NOMEM 5            /* device allocation failed */

/* Arithmetic of the step. */
#define AB_PREC_FP64 0        /* the reference formulation in FP64: per
                                 (particle, node) normal equations and
                                 Cholesky every step (kTransportNode3DT) */
#define AB_PREC_MIXED 1       /* MLS coefficients factorised once in FP64,
                                 transport contraction, collision and storage
                                 in FP32; conservation sums, wall scaling and
                                 body loads in FP64 */
#define AB_PREC_FP64_TABLE 2  /* 3-D: the MIXED pipeline with FP64 storage
                                 and arithmetic. 2-D: the FP64 step with
                                 the MLS transport as per-class coefficient
                                 tables, and positivity, commit, wall and
                                 collision in one pass with the two mass
                                 rescales applied as one (equal to the
                                 FP64 step to rounding; lattice clouds;
                                 MIXED is 3-D only) */

typedef struct ab_ctx ab_ctx;

typedef struct ab_config {
    int dim;                  /* 2 or 3 */
    int n_particles;          /* N */
    int nv;                   /* velocity nodes per axis (SimParameters::Nv) */
    double vmin, vmax;        /* velocity grid, dv = (vmax-vmin)/(nv-1) */
    double R;                 /* gas constant (GasConstants::R) */
    double tau;               /* BGK relaxation time (SimParameters::tao) */
    double radius;            /* MLS support = neighbour radius (cp.radius) */
    int mls_order;            /* 1 linear, 2 quadratic (SimParameters::mlsOrder) */
    double mls_tol;           /* relative Cholesky pivot tolerance; <= 0 -> 1e-12
                                 (ALEBGK_MLS_TOL, CudaBackend.cu:680) */
    int mls_clamp;            /* clamp reconstructions to the neighbour range
                                 (ALEBGK_MLS_CLAMP, CudaBackend.cu:691) */
    int no_mass_fix;          /* skip the end-of-step mass restore
                                 (ALEBGK_NO_MASS_FIX, CudaBackend.cu:1768) */
    int precision;            /* AB_PREC_* */
    /* Lattice the cloud was generated on, particle p = ix + nx*(iy + ny*iz)
     * (Initialization/ParticleGeneration.cpp:119-185). Required by the table
     * pipelines (MIXED, FP64_TABLE); 0 when the cloud is not a pure lattice. */
    int nx, ny, nz;
    /* Voxel grid of the neighbour search (CalcParameters hBox.. nbwBox). */
    double xleft, ybottom, zfront;
    double hbox, vbox, wbox;
    int nbhbox, nbvbox, nbwbox;
    /* Collision target (3-D). AB_COLLISION_CONTINUOUS is the reference's
     * continuous Maxwellian of (rho, u, T) (kCollisionNode3D,
     * CudaBackend.cu:1181-1197); AB_COLLISION_DISCRETE relaxes to the
     * discrete Maxwellian with the grid moments of g (no counterpart in the
     * reference). dm_tol <= 0 -> 1e-12, dm_maxit <= 0 -> 8. */
    int collision;
    double dm_tol;
    int dm_maxit;
    /* Limiter semantics. AB_LIMITER_GPU is the reference's GPU step
     * (cudaStep: positivity falls back to the pre-transport value, NAD
     * rescales into the neighbours' range, two mass rescales). AB_LIMITER_MOOD
     * is its CPU step with the default limiter, MOOD (runSolver,
     * Solver/Solver.cpp:641-693: positivity to zero, checkTransportPAD with a
     * linear re-transport of the flagged particles, one mass rescale, a
     * second collision when moodCheck fails); FP64 with the continuous
     * collision only. */
    int limiter;
} ab_config;

#define AB_COLLISION_CONTINUOUS 0
#define AB_COLLISION_DISCRETE 1
#define AB_LIMITER_GPU 0
#define AB_LIMITER_MOOD 1

/* Borrowed host view of a cloud (the SoA fields cudaUploadCloud gathers,
 * Solver/CudaBackend.cu:1223-1381). Any of mvx/mvy/mvz may be NULL (zero
 * mesh velocity). neigh/nneigh may be NULL: then ab_upload rebuilds the lists
 * on the device. */
typedef struct ab_cloud_view {
    const double *x, *y, *z;
    const unsigned char *boundary;
    const double *wnx, *wny, *wnz;
    const double *wux, *wuy, *wuz;
    const double *wt;
    const double *mvx, *mvy, *mvz;
    const double *rho, *ux, *uy, *uz, *t;
    const double *g;
    const int *neigh;
    const int *nneigh;
    const int *body_id;      /* Particle::bodyId (N), or NULL: no bodies */
} ab_cloud_view;

/* Per-step report (cudaLimiterCounts / cudaMlsCounts / cudaMlsOvershoot /
 * cudaMassBudget, Solver/CudaBackend.cu:1986-2030). Filling it costs one
 * host synchronisation. */
typedef struct ab_step_stats {
    double mass_old;               /* total mass at the start of the step */
    double mass_after_transport;   /* after limiters and commit, before rescale */
    double mass_after_collision;   /* after wall + collision, before rescale */
    int n_positivity;              /* particles the positivity repair touched */
    int n_nad;                     /* particles the NAD clamp touched */
    int n_mls_degenerate;          /* particles with a collapsed MLS pivot */
    int n_mls_extrapolate;         /* particles with an overshooting value */
    double mls_ovr_sum, mls_ovr_max;
    /* discrete-Maxwellian collision: particles solved, those that did not
     * converge (continuous fallback), grid passes summed over particles,
     * largest final relative residual */
    int dm_particles;
    int dm_unconverged;
    long long dm_passes;
    double dm_res_max;
    /* AB_LIMITER_MOOD: n_positivity counts the positivity repairs, n_nad the
     * particles the transport check sent to the linear re-transport, and
     * n_pad_fail those that failed the post-collision check (any one makes
     * the step collide every interior particle again). */
    int n_pad_fail;
} ab_step_stats;

/* Lifecycle. Replaces the file-static DeviceCloud (CudaBackend.cu:115) with
 * an explicit handle. */
int ab_create(const ab_config *cfg, ab_ctx **out);
/* The checks ab_create makes, on the host only (no device is touched):
 * AB_OK, or the code ab_create would return and a message in msg (cap
 * bytes, may be NULL). The reference refuses such configurations at upload
 * with a message (Solver/Solver.cpp:260-267). Besides argument ranges it
 * refuses a 3-D table pipeline whose z block of stored g would exceed the
 * 2 GiB its 32-bit buffer offsets reach (for example cavity3d 60^3, Nv=20
 * with AB_PREC_FP64_TABLE; AB_PREC_MIXED fits). */
int ab_check_config(const ab_config *cfg, char *msg, int cap);
int ab_destroy(ab_ctx *ctx);
const char *ab_last_error(const ab_ctx *ctx);

/* cudaUploadCloud (Solver/CudaBackend.cu:1223). Re-upload with the same N
 * reuses the device buffers. For the table pipelines this also builds the
 * MLS coefficient tables (the FP64 factorisation). */
int ab_upload(ab_ctx *ctx, const ab_cloud_view *view);

/* updateVoxelsAndNeighbours (Initialization/NeighbourSearch.cpp:17) on the
 * device: integer counting sort into voxels, then the 27-voxel scan in the
 * reference order; lists are identical to the reference's. Asynchronous on
 * the context's stream, except in the table pipelines, which wait for the
 * lists to check the lattice classes again. */
int ab_rebuild_neighbours(ab_ctx *ctx);
int ab_download_neighbours(ab_ctx *ctx, int *neigh, int *nneigh);
/* The voxel of every particle from the last device rebuild (Particle::voxel,
 * which restart files record). */
int ab_download_voxels(ab_ctx *ctx, int *voxel);

/* cudaStep (Solver/CudaBackend.cu:1505). stats may be NULL (no host sync). */
int ab_step(ab_ctx *ctx, double dt, ab_step_stats *stats);
/* nsteps calls of ab_step without stats, enqueued back to back. */
int ab_steps(ab_ctx *ctx, double dt, int nsteps);

/* The transport stage alone (kCopyGt + kTransportNode3DT / kTransportNode,
 * CudaBackend.cu:191-300, 699-869): gt from the current g, g untouched.
 * AB_PREC_FP64 only. */
int ab_transport_only(ab_ctx *ctx, double dt);

/* cudaDownloadFull / cudaDownloadMoments (CudaBackend.cu:2056, 1917).
 * ab_download_moments: any pointer may be NULL; limflag gets the per-particle
 * limiter bits of the last step (1 positivity, 2 NAD). */
int ab_download_g(ab_ctx *ctx, double *g);
int ab_download_gt(ab_ctx *ctx, double *gt);
int ab_download_moments(ab_ctx *ctx, double *rho, double *ux, double *uy,
                        double *uz, double *t, double *limflag);
/* The particle fields of cudaDownloadFull (CudaBackend.cu:2056) that cloud
 * management changes: positions, the boundary flag and the wall state
 * (wall: 7 per particle, wn xyz, wU xyz, wT). Any pointer may be NULL. */
int ab_download_particles(ab_ctx *ctx, double *x, double *y, double *z,
                          unsigned char *boundary, double *wall);

/* cudaTotalMass (CudaBackend.cu:2018). */
int ab_total_mass(ab_ctx *ctx, double *mass);

/* Surface-stress loads on rigid bodies, computeBodyLoads
 * (RigidBody/BodyDynamics.cpp:17-97), reduced on the device in FP64.
 * surf_idx[k] is a particle index, surf_body[k] its body id (< nbodies);
 * body_c is 3*nbodies centres, body_area the per-particle surface area of
 * each body; wu (3*nsurf) is the wall velocity of each surface particle, or
 * NULL to use the uploaded wall velocity. out = 6*nbodies
 * (Fx, Fy, Fz, Tx, Ty, Tz per body; 2-D: Fx, Fy, 0, 0, 0, Tz). Moments,
 * tractions and the per-body sums (a fixed-order reduction) stay on the
 * device; only the 6*nbodies results come back. */
int ab_body_loads(ab_ctx *ctx, int nsurf, const int *surf_idx,
                  const int *surf_body, int nbodies, const double *body_c,
                  const double *body_area, const double *wu, double *out);
/* ab_body_loads without any device allocation. No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). The
 * scratch lives in a file-scope device array of the library, one per GPU,
 * for at most AB_STATIC_SURF surface particles (about 1.3 MB per GPU, held
 * until the process ends); larger calls are refused with AB_ERR_ARG. The
 * array is found once per device (hipGetSymbolAddress) on the first call;
 * every call on a GPU uses the same array and overwrites what it reads, so
 * calls that overlap on one GPU (other threads, other streams) race. Same
 * arguments and results as ab_body_loads. */
#define AB_STATIC_SURF 8192
int ab_body_loads_static(ab_ctx *ctx, int nsurf, const int *surf_idx,
                         const int *surf_body, int nbodies, const double *body_c,
                         const double *body_area, const double *wu, double *out);

/* Update wall state and mesh velocity of selected particles after a body
 * moved (the uploads of cudaSyncBodies, CudaBackend.cu:1884-1915). Any array
 * may be NULL. */
int ab_set_wall_state(ab_ctx *ctx, int n, const int *idx, const double *x,
                      const double *y, const double *z, const double *wn,
                      const double *wu, const double *mv);

int ab_synchronize(ab_ctx *ctx);

/* ---- Cloud management around moving bodies --------------------------------
 * manageRemove, mergeCloseParticles and manageRefill
 * (RigidBody/CloudManagement.cpp:17-275) with the neighbour rebuilds the
 * moving-body step puts between them (Solver/Solver.cpp:379-398), on the
 * device: the cloud is re-laid out in place, with no download or upload of
 * g (the reference's cudaDownloadFull / cudaUploadCloud round trip,
 * Solver.cpp:373,401). AB_PREC_FP64 only (a moving body leaves the lattice).
 *
 * A body is one of the reference's rigid bodies (RigidBody/Geometry.hpp):
 * shape AB_BODY_SQUARE (side size, orientation o), AB_BODY_CIRCLE (radius
 * size), AB_BODY_SPHERE (radius size), AB_BODY_CUBE (side size, body-to-world
 * rotation rm, row-major as Cube::Rm). body_id holds Particle::bodyId of the
 * current cloud (N entries, -1 for none), or is NULL for the body ids the
 * device cloud carries (ab_cloud_view.body_id; given ones replace them): the
 * body surface particles are the boundary particles with body_id >= 0. The
 * device's body ids follow the new cloud. dx is CalcParameters::dx, min_dist
 * CalcParameters::minDist (<= 0: no merging), dom the box (xleft, xright,
 * ybottom, ytop, zfront, zback); the MLS radius and the velocity grid are the
 * configuration's. On return the cloud holds *n_new particles and counts (may
 * be NULL) holds {removed, merged, added}. Particles keep their order,
 * refilled ones are appended in site order, as in the reference.
 * ab_cloud_source_map then gives, for each particle of the new cloud, the
 * index it had before the call, or -1 for a refilled one (bodyId -1): src
 * holds *n_new entries. */
#define AB_BODY_SQUARE 1
#define AB_BODY_CIRCLE 2
#define AB_BODY_SPHERE 3
#define AB_BODY_CUBE 4
typedef struct ab_body {
    int shape;
    double c[3];
    double size;
    double o[2];
    double rm[9];
} ab_body;
int ab_cloud_manage(ab_ctx *ctx, int nbodies, const ab_body *bodies, const int *body_id,
                    double dx, double min_dist, const double dom[6], int *n_new,
                    int counts[3]);
int ab_cloud_source_map(ab_ctx *ctx, int *src);
/* Particle::bodyId of every particle of the device cloud (uploaded with
 * ab_cloud_view.body_id, carried through cloud management; -1 for none). */
int ab_download_body_ids(ab_ctx *ctx, int *body_id);

/* ---- Moving rigid bodies ---------------------------------------------------
 * A rigid body with its dynamic state (RigidBody and its shapes,
 * RigidBody/Geometry.hpp:19-362): shape and size as ab_body, npts the surface
 * particles it was sampled with (the surface area per particle is the
 * shape's area over npts), mass, surface temperature T, whether it spins
 * (rotates), centre c, 2-D orientation o, cube rotation rm (body to world,
 * row-major), velocity vel, 2-D spin angular_vel, 3-D angular velocity om,
 * and the Heun-averaged force and torque of the last step (2-D: tq[2] is
 * RigidBody::torque). */
typedef struct ab_rigid_body {
    int shape;
    int npts;
    int rotates;
    double size;
    double mass, T;
    double c[3];
    double o[2];
    double rm[9];
    double vel[3];
    double angular_vel;
    double om[3];
    double force[3];
    double tq[3];
} ab_rigid_body;

/* The moving-body block of one time step, after ab_step: moveRigidBodies
 * (RigidBody/BodyDynamics.cpp:143-711) and the cloud management that follows
 * it (manageRemove, mergeCloseParticles, manageRefill with their neighbour
 * rebuilds, Solver/Solver.cpp:348-444 and :705-714), all on the device's
 * cloud with no download or upload of g:
 *   loads L0 from the surface particles' wall velocities; the predicted
 *   body velocities; loads L1 at those velocities on the same g (one pass of
 *   raw surface moments serves both); the surface particles moved rigidly
 *   with the old state; Newton-Euler in FP64 with the averaged loads; the
 *   box-wall contact (a rigid shift by the surface extent, 0.4 dx clear,
 *   inelastic); the surface wall state (normal, wall and mesh velocity)
 *   from the new state; then cloud management around the moved bodies.
 * bodies (nbodies, in/out) carry the state; the body surface particles are
 * the boundary particles whose body id (ab_download_body_ids) is >= 0. dx,
 * min_dist and dom as for ab_cloud_manage. loads (may be NULL) receives
 * 12 per body: the L0 force and torque, then L1's. n_new and counts as
 * for ab_cloud_manage. AB_PREC_FP64 only (the cloud leaves the lattice). The
 * load sums are reduced on the device in a fixed order (the reference sums
 * particle by particle: equal to rounding); the body arithmetic is the
 * reference's, operation for operation. */
int ab_body_step(ab_ctx *ctx, double dt, int nbodies, ab_rigid_body *bodies, double dx,
                 double min_dist, const double dom[6], double *loads, int *n_new,
                 int counts[3]);

/* ab_body_step of several independent clouds in one call (no counterpart in
 * the reference, which has one cloud): task k applies to ctxs[k], the
 * contexts one after another on the calling thread (whose current device
 * is restored). Each cloud ends exactly as ab_body_step alone leaves it.
 * Returns the first failing task's status (its message in that context's
 * ab_last_error). n <= 64. */
typedef struct ab_body_task {
    double dt;
    int nbodies;
    ab_rigid_body *bodies;
    double dx, min_dist;
    double dom[6];
    double *loads;           /* may be NULL */
    int *n_new;              /* may be NULL */
    int *counts;             /* 3 ints, may be NULL */
} ab_body_task;
int ab_body_step_many(int n, ab_ctx *const *ctxs, const ab_body_task *tasks);

/* ---- Autotuner of the transport launch shape -----------------------------
 * No counterpart in the reference. Times every launch shape of the table
 * pipelines' transport contraction on the context's current state, `reps`
 * launches each after a warm-up (the transport writes only its scratch
 * output, the state is not advanced), and keeps the fastest for the
 * following steps. Shapes: one lattice column per wave with z blocks of K
 * outputs ("K9/W3": bit-identical results for every K), or TX x TY blocks of
 * interior columns per wave ("B2x4/K6/S": the same sum in another order,
 * equal to FP32 rounding; "/S": items of one z block, "/S2" of two, edge
 * columns first).
 * names/ms (cap entries each, may be NULL) receive every
 * candidate's name and best time in ms, "n/a" and -1 for a shape that does
 * not tile the lattice; *n the number of candidates, *chosen the index kept.
 * ab_set_transport_shape forces an index (-1: the default for the lattice).
 * On a 2-D table context the candidates are the apply's run lengths ("R4",
 * "R8", "R12": outputs per work item along a lattice line, bit-identical
 * results; -1 gives R8). */
int ab_autotune(ab_ctx *ctx, double dt, int reps, int cap, char (*names)[32], float *ms,
                int *n, int *chosen);
int ab_set_transport_shape(ab_ctx *ctx, int shape);

/* Measurement helper (no counterpart in the reference): `reps` launches of a
 * plain copy kernel over `bytes` bytes (read once, written once), so a
 * profiler's HBM counters (FETCH_SIZE, WRITE_SIZE) can be scaled to bytes.
 * Allocates and frees its own buffers. */
int ab_copy_calibration(unsigned long long bytes, int reps);

/* ---- Several contexts in one HIP graph ----------------------------------
 * No counterpart in the reference (its time loop launches every kernel from
 * the host). ab_graph_create captures nsteps steps of each of the n
 * contexts (uploaded, distinct) into one graph in which every context is an
 * independent branch; ab_graph_launch enqueues one replay (asynchronous, on
 * the graph's own stream, ordered after the work already enqueued on every
 * member context's stream; work enqueued on those streams afterwards waits
 * for the replay, as after eager steps) and ab_graph_synchronize waits for
 * it. A replay gives exactly the eager steps' results. Refused
 * (AB_ERR_STATE) while a context profiles or records training rows; table
 * coefficients are built for dt before the capture, and again before a
 * replay where an eager step at another dt or a wall-state change left them
 * stale (on the context's stream; the launch stays asynchronous). A graph
 * holds the clouds' device addresses: once any of its contexts is
 * re-uploaded or re-laid out (ab_upload, ab_cloud_manage, ab_body_step), or
 * a 2-D table context's transport is set up again (ab_rebuild_neighbours,
 * ab_set_wall_state, ab_autotune, ab_set_transport_shape), ab_graph_launch
 * refuses with AB_ERR_STATE; create the graph again. */
typedef struct ab_graph ab_graph;
int ab_graph_create(ab_ctx *const *ctxs, int n, double dt, int nsteps, ab_graph **out);
int ab_graph_launch(ab_graph *graph);
int ab_graph_synchronize(ab_graph *graph);
int ab_graph_destroy(ab_graph *graph);

/* ---- Several 2-D clouds in one context ----------------------------------
 * No counterpart in the reference (one cloud per run). ab_batch_create
 * copies the uploaded 2-D contexts' clouds (any sizes; the same velocity
 * grid, gas, MLS order and precision, AB_PREC_FP64 or AB_PREC_FP64_TABLE,
 * the GPU limiter and the continuous collision) into one new context, one
 * after another, and steps them together: one launch per stage covers every
 * cloud, while the mass sums, rescales and MLS radii stay per cloud, so each
 * cloud steps exactly as it would alone. Downloads give the concatenated
 * arrays; ab_total_mass the sum over the clouds. ab_batch_scatter copies g
 * and the moments back into the member contexts, which must outlive the
 * batch, and keep their clouds: ab_batch_scatter refuses (AB_ERR_STATE) a
 * member re-uploaded or re-laid out since the batch was formed. A batch
 * refuses ab_upload, ab_rebuild_neighbours and
 * ab_cloud_manage (work on the members, then form the batch again);
 * ab_destroy frees it and leaves the members. */
int ab_batch_create(ab_ctx *const *ctxs, int n, ab_ctx **out);
int ab_batch_scatter(ab_ctx *batch);

/* ---- Discrete-Maxwellian collision and its learned warm start ----------
 * No counterpart in the reference. With cfg.collision = AB_COLLISION_DISCRETE
 * every interior particle solves, in FP64, for a (5) such that
 * M_d(v) = exp(a . (1, v/cs, |v|^2/cs^2)) has the grid moments (1, v, |v|^2)
 * of g; cs is the largest |v| of the grid. Newton starts from the continuous
 * Maxwellian's parameters, or from those plus a small MLP's prediction of
 * the difference (ab_mlp_enable); Newton always runs to dm_tol, so the MLP
 * only changes how many grid passes that takes. */

/* The last step's solution: alpha N x 5, passes N (> dm_maxit + 1: did not
 * converge, continuous fallback), residual N; any pointer may be NULL. */
int ab_dm_download(ab_ctx *ctx, double *alpha, int *passes, double *residual);

/* Record MLP training rows during the next `steps` steps (a ring of up to 16
 * steps of N rows: 5 features u_x/cs, u_y/cs, u_z/cs, R T/cs^2, dv/cs;
 * 5 targets a_discrete - a_continuous; a weight, 0 for boundary particles
 * and unconverged solves). ab_dm_rows reports the rows held and how many
 * carry weight. */
int ab_dm_record(ab_ctx *ctx, int steps);
int ab_dm_rows(ab_ctx *ctx, int *rows, int *valid);

/* MLP 5 -> 64 -> 64 -> 5 (tanh), FP32 master parameters, BF16 weights in the
 * warm start. Parameter block of AB_MLP_PARAMS floats: W1[64][5] b1[64]
 * W2[64][64] b2[64] W3[5][64] b3[5] (AB_MLP_TRAINABLE), then input and
 * target normalisation xm[5] xs[5] ym[5] ys[5]. */
#define AB_MLP_TRAINABLE 4869
#define AB_MLP_PARAMS 4889
/* Random initialisation (seeded) and normalisation from the recorded rows. */
int ab_mlp_init(ab_ctx *ctx, unsigned long long seed);
/* Adam on the recorded rows, on the device: iters minibatches of `batch`
 * rows drawn with the seed; loss = sum w |y - t|^2 / (5 batch) in
 * normalised units. loss_first / loss_last may be NULL. */
int ab_mlp_train(ab_ctx *ctx, int iters, int batch, float lr, unsigned long long seed,
                 float *loss_first, float *loss_last);
int ab_mlp_enable(ab_ctx *ctx, int on);
int ab_mlp_get(ab_ctx *ctx, float *params);
int ab_mlp_set(ab_ctx *ctx, const float *params);
/* Forward pass with the FP32 parameters: x n x 5 -> y n x 5 (de-normalised). */
int ab_mlp_forward(ab_ctx *ctx, const float *x, int n, float *y);
/* The warm start's own inference, as the step runs it: BF16 weights on
 * MFMA, layer-1 activations rounded to BF16, FP32 accumulation. */
int ab_mlp_infer(ab_ctx *ctx, const float *x, int n, float *y);
/* Loss and its gradient (AB_MLP_TRAINABLE floats) over rows x (n x 5),
 * targets t (n x 5, de-normalised), weights w (n), FP32 parameters. */
int ab_mlp_loss_grad(ab_ctx *ctx, const float *x, const float *t, const float *w, int n,
                     float *loss, float *grad);

/* Per-stage device timing with HIP events on the context's stream (the
 * [GPU-STEADY] table of CudaBackend.cu:1518-1623). After ab_profile(ctx, 1),
 * every step records events around its stages; ab_profile_read returns the
 * accumulated milliseconds and launch counts per stage, in the order of
 * ab_stage_name(k) for k < *nstages. */
int ab_profile(ab_ctx *ctx, int enable);
int ab_profile_read(ab_ctx *ctx, double *ms, long *counts, int *nstages);
const char *ab_stage_name(int k);

#ifdef __cplusplus
}
#endif
#endif
