// ab_internal.hpp -- the context behind the C ABI and the launchers each
// kernel translation unit exports to ab_api.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/alebgk_dev.h"
#include "ab_dm.hpp"

namespace ab {

constexpr int kMaxBodies = 16;     // rigid bodies per cloud

// Device view of one cloud: flat SoA arrays, mirroring the fields of the
// reference's DeviceCloud (Solver/CudaBackend.cu:69-113).
struct DevCloud
{
    int dim, N, Nv;
    int nodes;        // velocity nodes of the mass-carrying layer: Nv^3 or Nv^2
    int gstride;      // values per particle: Nv^3 or 2*Nv^2
    double vlo, dv;
    const double *x, *y, *z;
    const unsigned char *bnd;
    const double *wnx, *wny, *wnz, *wux, *wuy, *wuz, *wt;
    const double *mvx, *mvy, *mvz;
    const int *neigh, *nneigh;
    // several independent clouds in one context (ab_batch_create): the
    // instance of each particle and each instance's MLS support radius;
    // null for a single cloud
    const int *inst;
    const double *ih;
};

// Discrete-Maxwellian collision settings and outputs (3-D interior
// particles). Pointers are per particle; rec/wb/pf may be null.
struct DmArgs
{
    int on;                   // relax to the discrete Maxwellian
    int mlp;                  // warm-start Newton from the MLP's prediction
    double cs, tol;           // velocity scale; relative residual to stop at
    int maxit;                // Newton updates at most
    double *alpha;            // N x 5 solution
    int *passes;              // N: grid passes (> maxit + 1: not converged)
    double *res;              // N: final relative residual
    float *rec;               // N x kMlpRow training rows, or null
    const unsigned short *wb; // MLP weights, BF16
    const float *pf;          // MLP biases and normalisation, FP32
};

// Device scalars of one step (slots of Ctx::scal).
enum ScalarSlot
{
    S_MASS_OLD = 0,     // total mass at the start of the step
    S_MASS_NEW,         // after transport, limiters and commit
    S_SCALE1,           // massOld / massNew (1 when skipped)
    S_MASS_COLL,        // after wall and collision
    S_SCALE2,           // massOld / massColl (1 when skipped)
    S_LAZY,             // table pipelines: factor the stored g still carries
    S_MASS_START,       // mass at the start of the last completed step
    S_COUNT = 8
};

// R_MASS_FUSED: the committed masses (second array) and the collided
// masses of the unscaled rows (first array) of the 2-D table step's fused
// commit and collision, giving both scales at once.
enum ReduceMode { R_MASS_OLD = 0, R_MASS_NEW = 1, R_MASS_COLL = 2, R_MASS_FUSED = 3 };

// Stages timed by ab_profile, in report order.
enum Stage
{
    ST_TRANSPORT = 0,   // MLS transport (the dominant kernel)
    ST_LIMIT_COMMIT,    // positivity + NAD + commit
    ST_WALL_COLLIDE,    // mass rescale + diffuse wall + moments + collision
    ST_FINAL,           // end-of-step rescale + moments
    ST_REDUCE,          // global mass sums
    ST_NEIGH,           // neighbour rebuild
    ST_COUNT
};

}  // namespace ab

struct ab_ctx
{
    ab_config cfg{};
    int dim = 3, N = 0, Nv = 0, nodes = 0, gstride = 0;
    double vlo = 0, dv = 0;
    hipStream_t stream = nullptr;
    int device = 0;                // the device current at ab_create
    std::string err;
    bool uploaded = false;

    // cloud (device)
    double *x = nullptr, *y = nullptr, *z = nullptr;
    unsigned char *bnd = nullptr;
    double *wnx = nullptr, *wny = nullptr, *wnz = nullptr;
    double *wux = nullptr, *wuy = nullptr, *wuz = nullptr, *wt = nullptr;
    double *mvx = nullptr, *mvy = nullptr, *mvz = nullptr;
    double *rho = nullptr, *ux = nullptr, *uy = nullptr, *uz = nullptr, *T = nullptr;
    int *neigh = nullptr, *nneigh = nullptr;
    int *bid = nullptr;            // Particle::bodyId (-1: none)
    int *ilist = nullptr, *blist = nullptr;   // interior / boundary particle lists
    int n_int = 0, n_bnd = 0;

    // FP64 reference-formulation state
    double *g = nullptr, *gt = nullptr;
    double *limflag = nullptr;
    int *mlsflag = nullptr;
    int *mood = nullptr;           // N: flagged by the MOOD transport check (AB_LIMITER_MOOD)
    int *moodfail = nullptr;       // 1: some particle failed the post-collision check
    double *mlsovr = nullptr;
    double *pmass = nullptr;       // per-particle partial sums
    double *scal = nullptr;        // ab::ScalarSlot
    bool mass_valid = false;       // pmass holds the current g's masses
    // Bumped whenever the device arrays are freed, re-allocated or refilled
    // (upload, re-layout, destroy): graphs and batches record it when they
    // are made and refuse to run on a cloud that has changed since.
    unsigned long long gen = 0;
    // Bumped whenever the 2-D table transport's device arrays are allocated
    // again (its setup after a neighbour rebuild or a wall-state change, a
    // new run length): graphs record it and refuse to replay stale pointers.
    unsigned long long tgen = 0;

    // neighbour rebuild scratch (vlist: 2 N ints, the sorted list and the
    // scatter before the sort)
    int *vox = nullptr, *vcount = nullptr, *vstart = nullptr, *vlist = nullptr;
    int nvox = 0;

    // table pipelines (AB_PREC_MIXED / AB_PREC_FP64_TABLE)
    int S = 0;                     // padded per-particle stride of the tables
    void *gs = nullptr;            // stored g, [N][S], FP32 or FP64
    void *gts = nullptr;           // transported g, [N][S]
    void *coef = nullptr;          // [ncls][81][S] MLS coefficients
    int *cls = nullptr;            // representative particle of each class
    int *lat = nullptr;            // lattice index (ix, iy, iz) per particle, 3*N
    int ncls = 0;
    double coef_dt = -1.0;         // dt the tables were built for
    double *m0 = nullptr;          // equilibrium the stored deviation is taken from, S
    double *m0x = nullptr;         // its 1-D factors e_l (kMaxNv); M0 = ((m0c e_i) e_j) e_k
    double *m0xy = nullptr;        // its pair factors (m0c e_i) e_j (kMaxNv^2)
    double m0sum = 0.0;            // sum of M0 over the nodes
    double m0c = 0.0;
    double m0mom[5] = {0};         // its moments: sum M0 dv^3, sum v M0 dv^3, sum |v|^2 M0 dv^3
    double m0rho = 1.0, m0T = 270.0;
    double *part = nullptr;        // per-particle partials of the table pipeline
    double *red = nullptr;         // block sums of the partials (2 per reduction block)
    double *pst = nullptr;         // per-particle state between the discrete-collision kernels
    int *bad = nullptr;            // stencil check result
    int apply_shape = -1;          // transport launch shape (-1: default for nz)

    // 2-D table transport (AB_PREC_FP64_TABLE with dim 2)
    double *t2_coef = nullptr;     // [instances x 25 classes][21 offsets][Nv^2] MLS coefficients
    int *t2_cls = nullptr;         // class of each particle, N
    int *t2_rep = nullptr;         // representative particle of each class
    int *t2_lat = nullptr;         // (ix, iy, nx, ny) of each particle, 4*N
    int t2_ncls = 0;
    int *t2_items = nullptr;       // apply work items, 8 ints each: column runs, row runs, corners
    int t2_nv = 0, t2_nh = 0, t2_nit = 0;   // column runs, row runs, all items
    int t2_run = -1;               // run length index of the apply (-1: the default, 8 outputs)
    double t2_dt = -1.0;           // dt the coefficients were built for
    double *t2_pnew = nullptr;     // committed mass per particle of the fused commit and collision

    // discrete-Maxwellian collision (cfg.collision) and its MLP warm start
    double *dm_alpha = nullptr;    // N x 5 solution of the last step
    int *dm_pass = nullptr;        // N grid passes of the last step
    double *dm_res = nullptr;      // N final relative residuals
    float *dm_rec = nullptr;       // recorded training rows, dm_rec_slots x N x kMlpRow
    int dm_rec_slots = 0;          // steps the ring holds
    int dm_rec_next = 0;           // slot the next recorded step writes
    int dm_rec_filled = 0;         // slots holding rows
    int dm_rec_left = 0;           // steps still to record
    float *mlp_p = nullptr;        // FP32 parameters (kMlpParams)
    unsigned short *mlp_wb = nullptr;   // BF16 copy of the trainable block
    float *mlp_m = nullptr, *mlp_v = nullptr, *mlp_g = nullptr;   // Adam moments, gradient
    float *mlp_work = nullptr;     // per-sample activations of a training batch
    int mlp_work_cap = 0;          // samples the work buffer holds
    int *mlp_idx = nullptr;        // rows of a batch
    float *mlp_rows = nullptr;     // caller rows (forward / gradient check)
    int mlp_rows_cap = 0;
    long mlp_step = 0;             // Adam step count
    bool mlp_ready = false;        // parameters initialised

    // a batch of independent 2-D clouds (ab_batch_create): per instance its
    // first particle (nbatch + 1 entries), lattice size and support radius;
    // empty for a single cloud
    int nbatch = 0;
    int *inst = nullptr;           // instance of each particle, N
    double *ih = nullptr;          // support radius of each instance
    int *seg = nullptr;            // first particle of each instance, nbatch + 1
    std::vector<int> bt_base, bt_nx, bt_ny;
    std::vector<double> bt_h;
    std::vector<ab_ctx *> members;
    std::vector<unsigned long long> member_gen;   // members' generations when batched

    // cloud management (ab_cloud_manage): the last call's source map, and
    // its device scratch kept between calls (slot k: the k-th buffer a call
    // asks for, grown when a call needs more)
    std::vector<int> cm_src;
    std::vector<void *> cm_pool;
    std::vector<size_t> cm_pool_bytes;
    bool mlp_on = false;           // warm start enabled

    // body loads scratch, grown to the surface-particle count: indices and
    // body of each surface particle, its raw moments (10), its load
    // contributions (6), given wall velocities (3); per-body parameters and
    // results; a counter
    int bmom_cap = 0;
    int *bidx = nullptr, *bbody = nullptr, *bcnt = nullptr;
    double *bmom = nullptr, *bcon = nullptr, *bwu = nullptr;
    double *bpar = nullptr, *bout = nullptr;

    // profiling: event pairs recorded around stages on the context's stream,
    // resolved (synchronised and accumulated) only when the pool fills or the
    // caller reads the totals, so the timed work is not serialised.
    bool prof = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> evpool;   // 2 per recorded stage
    std::vector<int> evstage;
    int nrec = 0;
    double prof_ms[ab::ST_COUNT] = {0};
    long prof_n[ab::ST_COUNT] = {0};

    ab::DevCloud view() const
    {
        ab::DevCloud c;
        c.dim = dim; c.N = N; c.Nv = Nv; c.nodes = nodes; c.gstride = gstride;
        c.vlo = vlo; c.dv = dv;
        c.x = x; c.y = y; c.z = z; c.bnd = bnd;
        c.wnx = wnx; c.wny = wny; c.wnz = wnz; c.wux = wux; c.wuy = wuy; c.wuz = wuz;
        c.wt = wt; c.mvx = mvx; c.mvy = mvy; c.mvz = mvz;
        c.neigh = neigh; c.nneigh = nneigh;
        c.inst = inst; c.ih = ih;
        return c;
    }
};

namespace ab {

// ---- ab_api.hip: stage timing ------------------------------------------------
void prof_begin(ab_ctx *c);
void prof_end(ab_ctx *c, int stage);
void prof_flush(ab_ctx *c);

// ---- k_fp64.hip ------------------------------------------------------------
void launch_transport_fp64(const DevCloud &c, const double *g, double *gt, double dt,
                           double h, int order, double tol, int clamp, int *mlsflag,
                           double *mlsovr, const int *list, int nlist, bool boundary,
                           hipStream_t s, const int *only = nullptr);
// The reference's CPU-step limiter (MOOD): positivity to zero and the
// transport check (flags), the commit after the flagged particles'
// re-transport, and the post-collision admissibility check.
void launch_mood_check(const DevCloud &c, double *gt, const double *rho, int *mood,
                       double *limflag, hipStream_t s);
void launch_mood_commit(const DevCloud &c, double *g, double *gt, const int *mood,
                        double *limflag, double *pmass, hipStream_t s);
void launch_pad_check(const DevCloud &c, const double *g, const double *rho, const double *T,
                      double *limflag, int *fail, hipStream_t s);
void launch_mass_partials(const DevCloud &c, const double *g, double *pmass, hipStream_t s);
// One sum per instance when seg is given (nseg instances, first particles
// seg[0..nseg], scalars S_COUNT apart); else one over all N.
void launch_reduce_mass(int N, const double *pmass, double *scal, int mode, int nofix,
                        hipStream_t s, const int *seg = nullptr, int nseg = 1,
                        const double *pnew = nullptr);
void launch_limit_commit(const DevCloud &c, double *g, const double *gt, const double *rho,
                         double *limflag, double *pmass, hipStream_t s);
// guard (may be null): when given, the launch acts only if *guard is set,
// and the wall-collide launch then collides the interior particles again
// without rescaling (the MOOD limiter's second collision).
void launch_wall_collide(const DevCloud &c, const DmArgs &dm, double *g, const double *scal, double dt,
                         double tau, double R, double *rho, double *ux, double *uy,
                         double *uz, double *T, double *pmass, hipStream_t s,
                         const int *guard = nullptr);
// fold_scale1: after launch_commit_collide2, whose wall densities still
// lack the first rescale, the boundary particles' rho is multiplied by it.
void launch_final(const DevCloud &c, double *g, const double *scal, double R, double *rho,
                  double *ux, double *uy, double *uz, double *T, double *pmass,
                  hipStream_t s, const int *guard = nullptr, bool fold_scale1 = false);
// The 2-D step's positivity, commit, wall and collision in one pass (the
// first rescale folded into the end-of-step one): g from gt, the committed
// masses into pnew, the collided masses into pmass.
void launch_commit_collide2(const DevCloud &c, double *g, const double *gt, double dt,
                            double tau, double R, double *rho, double *ux, double *uy,
                            double *T, double *limflag, double *pnew, double *pmass,
                            hipStream_t s);

// ---- k_neigh.hip -------------------------------------------------------------
int launch_neighbours(const DevCloud &c, const ab_config &cfg, int *vox, int *vcount,
                      int *vstart, int *vlist, int nvox, int *neigh, int *nneigh,
                      hipStream_t s);

// ---- k_table.hip -------------------------------------------------------------
int table_setup(ab_ctx *ctx);
int table_upload_g(ab_ctx *ctx, const double *g_host);
int table_download_g(ab_ctx *ctx, double *g_host);
int table_step(ab_ctx *ctx, double dt);
int table_prepare(ab_ctx *ctx, double dt);   // coefficients for dt
int table_autotune(ab_ctx *ctx, double dt, int reps, int cap, char (*names)[32], float *ms,
                   int *n, int *chosen);
int table_set_shape(ab_ctx *ctx, int shape);
// Re-runs the lattice class check (after a wall-state change).
int table_recheck(ab_ctx *ctx);

// ---- k_table2d.hip -------------------------------------------------------------
// The 2-D table transport: classes and their check (after the upload and
// every change of the cloud's geometry or wall state), gt from g for dt
// (the coefficients built when dt changes), the coefficients for dt alone.
int table2d_setup(ab_ctx *ctx);
int table2d_transport(ab_ctx *ctx, double dt);
int table2d_prepare(ab_ctx *ctx, double dt);
void table2d_free(ab_ctx *ctx);
// The apply's run lengths (outputs per work item along a lattice line):
// timed, or one forced (-1: the default); every run length gives
// bit-identical results.
int table2d_autotune(ab_ctx *ctx, double dt, int reps, int cap, char (*names)[32], float *ms,
                     int *n, int *chosen);
int table2d_set_run(ab_ctx *ctx, int run);

// Discrete-Maxwellian settings of the next step (records training rows when
// asked); defined in ab_api.hip.
DmArgs dm_args(ab_ctx *ctx);
void dm_step_done(ab_ctx *ctx);
void mlp_free(ab_ctx *ctx);   // k_mlp.hip
void table_free(ab_ctx *ctx);

// Cloud re-layout (ab_api.hip): particle i of the new cloud takes old
// particle src[i] (device array), zero where src[i] < 0; the interior and
// boundary lists from the boundary flags.
int relayout_cloud(ab_ctx *ctx, const int *src, int n);
int rebuild_lists(ab_ctx *ctx);

}  // namespace ab
