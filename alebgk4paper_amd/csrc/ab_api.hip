// ab_api.hip -- the extern "C" entry points of libalebgk_hip.so
// (include/alebgk_dev.h): context, upload/download, the step driver.
//
// The step keeps every global sum on the device: the reference reads the mass
// back to the host three times per step (sumMassHost, CudaBackend.cu:1384),
// here the rescale factors are computed and consumed on the device and a step
// enqueues without a single host synchronisation.
#include "ab_device.hpp"
#include "ab_internal.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

using namespace ab;

namespace {

#define HIPCHK(ctx, expr)                                                        \
    do {                                                                         \
        hipError_t e_ = (expr);                                                  \
        if (e_ != hipSuccess) {                                                  \
            (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);      \
            return AB_ERR_HIP;                                                   \
        }                                                                        \
    } while (0)

// The 2-D cloud with AB_PREC_FP64_TABLE: the FP64 step with the table
// transport (k_table2d.hip).
bool table2d(const ab_ctx *c) { return c->dim == 2 && c->cfg.precision == AB_PREC_FP64_TABLE; }
// The 3-D table pipelines (and the 2-D AB_PREC_MIXED request, which their
// setup refuses).
bool table_mode(const ab_ctx *c) { return c->cfg.precision != AB_PREC_FP64 && !table2d(c); }

template <typename T>
int dalloc(ab_ctx *ctx, T **p, size_t n)
{
    if (hipMalloc((void **)p, n * sizeof(T)) != hipSuccess)
    {
        ctx->err = "device allocation of " + std::to_string(n * sizeof(T)) + " bytes failed";
        *p = nullptr;
        return AB_ERR_NOMEM;
    }
    return AB_OK;
}

void free_cloud(ab_ctx *c)
{
    c->gen++;
    void *ptrs[] = {c->x, c->y, c->z, c->bnd, c->wnx, c->wny, c->wnz, c->wux, c->wuy, c->wuz,
                    c->wt, c->mvx, c->mvy, c->mvz, c->rho, c->ux, c->uy, c->uz, c->T,
                    c->neigh, c->nneigh, c->ilist, c->blist, c->g, c->gt, c->limflag,
                    c->mlsflag, c->mlsovr, c->pmass, c->vox, c->vcount, c->vstart, c->vlist,
                    c->bmom, c->bidx, c->bwu, c->bbody, c->bcnt, c->bcon, c->bpar, c->bout,
                    c->bid, c->dm_alpha, c->dm_pass, c->dm_res, c->dm_rec,
                    c->mood, c->moodfail, c->inst, c->ih, c->seg};
    for (void *p : ptrs) if (p) (void)hipFree(p);
    c->inst = c->seg = nullptr; c->ih = nullptr;
    c->dm_alpha = c->dm_res = nullptr; c->dm_pass = nullptr; c->dm_rec = nullptr;
    c->mood = c->moodfail = nullptr;
    c->dm_rec_slots = c->dm_rec_next = c->dm_rec_filled = c->dm_rec_left = 0;
    c->x = c->y = c->z = nullptr; c->bnd = nullptr;
    c->wnx = c->wny = c->wnz = c->wux = c->wuy = c->wuz = c->wt = nullptr;
    c->mvx = c->mvy = c->mvz = nullptr;
    c->rho = c->ux = c->uy = c->uz = c->T = nullptr;
    c->neigh = c->nneigh = c->ilist = c->blist = nullptr;
    c->g = c->gt = c->limflag = c->mlsovr = c->pmass = nullptr;
    c->mlsflag = nullptr;
    c->vox = c->vcount = c->vstart = c->vlist = nullptr;
    c->bmom = c->bcon = c->bwu = c->bpar = c->bout = nullptr;
    c->bidx = c->bbody = c->bcnt = nullptr;
    c->bmom_cap = 0;
    c->bid = nullptr;
    table_free(c);
    table2d_free(c);
    c->coef_dt = -1.0;
}

int alloc_cloud(ab_ctx *c)
{
    const size_t N = (size_t)c->N;
    int rc = 0;
    double **dd[] = {&c->x, &c->y, &c->z, &c->wnx, &c->wny, &c->wnz, &c->wux, &c->wuy, &c->wuz,
                     &c->wt, &c->mvx, &c->mvy, &c->mvz, &c->rho, &c->ux, &c->uy, &c->uz, &c->T,
                     &c->limflag, &c->mlsovr, &c->pmass};
    for (double **p : dd) if ((rc = dalloc(c, p, N))) return rc;
    if ((rc = dalloc(c, &c->bnd, N))) return rc;
    if ((rc = dalloc(c, &c->neigh, N * AB_MAX_NEIGH))) return rc;
    if ((rc = dalloc(c, &c->nneigh, N))) return rc;
    if ((rc = dalloc(c, &c->bid, N))) return rc;
    if ((rc = dalloc(c, &c->ilist, N))) return rc;
    if ((rc = dalloc(c, &c->blist, N))) return rc;
    if ((rc = dalloc(c, &c->mlsflag, N))) return rc;
    if ((rc = dalloc(c, &c->mood, N)) || (rc = dalloc(c, &c->moodfail, 1))) return rc;
    if ((rc = dalloc(c, &c->g, N * c->gstride))) return rc;
    if (!table_mode(c) && (rc = dalloc(c, &c->gt, N * c->gstride))) return rc;
    if (c->cfg.collision == AB_COLLISION_DISCRETE)
    {
        if ((rc = dalloc(c, &c->dm_alpha, N * 5)) || (rc = dalloc(c, &c->dm_pass, N)) ||
            (rc = dalloc(c, &c->dm_res, N)))
            return rc;
        if (hipMemset(c->dm_pass, 0, sizeof(int) * N) != hipSuccess ||
            hipMemset(c->dm_res, 0, sizeof(double) * N) != hipSuccess ||
            hipMemset(c->dm_alpha, 0, sizeof(double) * N * 5) != hipSuccess)
        { c->err = "DM state: memset failed"; return AB_ERR_HIP; }
    }
    return AB_OK;
}

__global__ void k_fill_int(int *p, int n, int v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
// Shifts the neighbour indices of n particles by add (a cloud moved into a
// batch at particle add).
__global__ void k_offset_neigh(int *neigh, const int *__restrict__ nneigh, int n, int add)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * AB_MAX_NEIGH) return;
    const int i = (int)(e / AB_MAX_NEIGH), m = (int)(e % AB_MAX_NEIGH);
    if (m < nneigh[i]) neigh[e] += add;
}

__global__ void k_fill_dbl(double *p, int n, double v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_scatter_wall(int n, const int *__restrict__ idx, const double *__restrict__ src,
                               int has_xyz, int has_wn, int has_wu, int has_mv, double *x,
                               double *y, double *z, double *wnx, double *wny, double *wnz,
                               double *wux, double *wuy, double *wuz, double *mvx, double *mvy,
                               double *mvz)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int i = idx[k];
    const double *s = src + (size_t)k * 12;
    if (has_xyz) { x[i] = s[0]; y[i] = s[1]; z[i] = s[2]; }
    if (has_wn) { wnx[i] = s[3]; wny[i] = s[4]; wnz[i] = s[5]; }
    if (has_wu) { wux[i] = s[6]; wuy[i] = s[7]; wuz[i] = s[8]; }
    if (has_mv) { mvx[i] = s[9]; mvy[i] = s[10]; mvz[i] = s[11]; }
}

// The per-step flags cleared in one launch (a fill kernel, not a memset, so
// a graph capture holds no memset node).
__global__ void k_zero_flags(int n, int *mlsflag, double *mlsovr, double *limflag, int *moodfail)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    mlsflag[i] = 0;
    mlsovr[i] = 0.0;
    limflag[i] = 0.0;
    if (i == 0) moodfail[0] = 0;
}

int zero_step_flags(ab_ctx *c)
{
    const int nb = (c->N + 255) / 256;
    k_zero_flags<<<nb, 256, 0, c->stream>>>(c->N, c->mlsflag, c->mlsovr, c->limflag, c->moodfail);
    return hipGetLastError() == hipSuccess ? AB_OK : AB_ERR_HIP;
}


struct Timer
{
    ab_ctx *c;
    explicit Timer(ab_ctx *ctx) : c(ctx) {}
    void tic() { prof_begin(c); }
    void toc(int st) { prof_end(c, st); }
};

int fp64_transport(ab_ctx *c, double dt)
{
    if (table2d(c)) return table2d_transport(c, dt);
    const double tol = c->cfg.mls_tol > 0 ? c->cfg.mls_tol : 1e-12;
    DevCloud v = c->view();
    launch_transport_fp64(v, c->g, c->gt, dt, c->cfg.radius, c->cfg.mls_order, tol,
                          c->cfg.mls_clamp, c->mlsflag, c->mlsovr, c->ilist, c->n_int, false,
                          c->stream);
    launch_transport_fp64(v, c->g, c->gt, dt, c->cfg.radius, c->cfg.mls_order, tol,
                          c->cfg.mls_clamp, c->mlsflag, c->mlsovr, c->blist, c->n_bnd, true,
                          c->stream);
    return hipGetLastError() == hipSuccess ? AB_OK : AB_ERR_HIP;
}

int fp64_step(ab_ctx *c, double dt)
{
    DevCloud v = c->view();
    hipStream_t s = c->stream;
    Timer tm(c);
    int rc;
    tm.tic();
    if (!c->mass_valid) launch_mass_partials(v, c->g, c->pmass, s);
    launch_reduce_mass(c->N, c->pmass, c->scal, R_MASS_OLD, 0, s, c->seg, c->nbatch);
    if ((rc = zero_step_flags(c))) return rc;
    tm.toc(ST_REDUCE);
    tm.tic();
    if ((rc = fp64_transport(c, dt))) return rc;
    tm.toc(ST_TRANSPORT);
    const bool fused = table2d(c);
    if (fused)
    {
        // the 2-D table formulation: positivity, commit, wall and collision
        // in one pass, both rescales in the end-of-step one
        tm.tic();
        launch_commit_collide2(v, c->g, c->gt, dt, c->cfg.tau, c->cfg.R, c->rho, c->ux, c->uy,
                               c->T, c->limflag, c->t2_pnew, c->pmass, s);
        tm.toc(ST_WALL_COLLIDE);
        tm.tic();
        launch_reduce_mass(c->N, c->pmass, c->scal, R_MASS_FUSED, c->cfg.no_mass_fix, s, c->seg,
                           c->nbatch, c->t2_pnew);
        tm.toc(ST_REDUCE);
    }
    else
    {
        tm.tic();
        launch_limit_commit(v, c->g, c->gt, c->rho, c->limflag, c->pmass, s);
        tm.toc(ST_LIMIT_COMMIT);
        tm.tic();
        launch_reduce_mass(c->N, c->pmass, c->scal, R_MASS_NEW, 0, s, c->seg, c->nbatch);
        tm.toc(ST_REDUCE);
        tm.tic();
        launch_wall_collide(v, dm_args(c), c->g, c->scal, dt, c->cfg.tau, c->cfg.R, c->rho, c->ux,
                            c->uy, c->uz, c->T, c->pmass, s);
        tm.toc(ST_WALL_COLLIDE);
        dm_step_done(c);
        tm.tic();
        launch_reduce_mass(c->N, c->pmass, c->scal, R_MASS_COLL, c->cfg.no_mass_fix, s, c->seg,
                           c->nbatch);
        tm.toc(ST_REDUCE);
    }
    tm.tic();
    launch_final(v, c->g, c->scal, c->cfg.R, c->rho, c->ux, c->uy, c->uz, c->T, c->pmass, s,
                 nullptr, fused);
    tm.toc(ST_FINAL);
    c->mass_valid = true;
    return hipGetLastError() == hipSuccess ? AB_OK : AB_ERR_HIP;
}

// The reference's CPU step with its MOOD limiter (Solver/Solver.cpp:641-693):
// transport with the positivity repair; the transport check and the flagged
// particles' linear re-transport; commit; one mass rescale (M_old / M_new);
// wall, moments, collision, moments; the post-collision check and, if any
// interior particle fails it, a second collision of every interior particle
// and the moments again. No host synchronisation: the re-transport and the
// second collision are launched always and act only where flagged.
int fp64_step_mood(ab_ctx *c, double dt)
{
    DevCloud v = c->view();
    hipStream_t s = c->stream;
    Timer tm(c);
    int rc;
    tm.tic();
    if (!c->mass_valid) launch_mass_partials(v, c->g, c->pmass, s);
    launch_reduce_mass(c->N, c->pmass, c->scal, R_MASS_OLD, 0, s, c->seg, c->nbatch);
    if ((rc = zero_step_flags(c))) return rc;
    tm.toc(ST_REDUCE);
    tm.tic();
    if ((rc = fp64_transport(c, dt))) return rc;
    tm.toc(ST_TRANSPORT);
    tm.tic();
    launch_mood_check(v, c->gt, c->rho, c->mood, c->limflag, s);
    const double tol = c->cfg.mls_tol > 0 ? c->cfg.mls_tol : 1e-12;
    launch_transport_fp64(v, c->g, c->gt, dt, c->cfg.radius, 1, tol, c->cfg.mls_clamp,
                          c->mlsflag, c->mlsovr, c->ilist, c->n_int, false, s, c->mood);
    launch_mood_commit(v, c->g, c->gt, c->mood, c->limflag, c->pmass, s);
    tm.toc(ST_LIMIT_COMMIT);
    tm.tic();
    launch_reduce_mass(c->N, c->pmass, c->scal, R_MASS_NEW, 0, s, c->seg, c->nbatch);
    tm.toc(ST_REDUCE);
    tm.tic();
    launch_wall_collide(v, dm_args(c), c->g, c->scal, dt, c->cfg.tau, c->cfg.R, c->rho, c->ux,
                        c->uy, c->uz, c->T, c->pmass, s);
    tm.toc(ST_WALL_COLLIDE);
    tm.tic();
    launch_reduce_mass(c->N, c->pmass, c->scal, R_MASS_COLL, 1, s, c->seg, c->nbatch);
    tm.toc(ST_REDUCE);
    tm.tic();
    launch_final(v, c->g, c->scal, c->cfg.R, c->rho, c->ux, c->uy, c->uz, c->T, c->pmass, s);
    launch_pad_check(v, c->g, c->rho, c->T, c->limflag, c->moodfail, s);
    launch_wall_collide(v, dm_args(c), c->g, c->scal, dt, c->cfg.tau, c->cfg.R, c->rho, c->ux,
                        c->uy, c->uz, c->T, c->pmass, s, c->moodfail);
    launch_final(v, c->g, c->scal, c->cfg.R, c->rho, c->ux, c->uy, c->uz, c->T, c->pmass, s,
                 c->moodfail);
    tm.toc(ST_FINAL);
    c->mass_valid = true;
    return hipGetLastError() == hipSuccess ? AB_OK : AB_ERR_HIP;
}

int do_step(ab_ctx *c, double dt)
{
    if (!c->uploaded) { c->err = "ab_step before ab_upload"; return AB_ERR_STATE; }
    if (!(dt > 0.0) || !std::isfinite(dt)) { c->err = "dt must be positive and finite"; return AB_ERR_ARG; }
    if (table_mode(c))
    {
        int rc = table_step(c, dt);
        if (rc && c->err.empty()) c->err = "table step failed";
        return rc;
    }
    if (c->cfg.limiter == AB_LIMITER_MOOD) return fp64_step_mood(c, dt);
    return fp64_step(c, dt);
}

}  // namespace

namespace ab {

int graph_coefficients(ab_ctx *c, double dt);

// The step without host work, for graph capture; graph_prepare refuses what
// a capture cannot hold and builds the table coefficients for dt eagerly.
int graph_prepare(ab_ctx *c, double dt)
{
    if (!c->uploaded) { c->err = "graph capture before ab_upload"; return AB_ERR_STATE; }
    if (c->prof) { c->err = "graph capture with profiling on"; return AB_ERR_STATE; }
    if (c->dm_rec_left > 0) { c->err = "graph capture while recording training rows"; return AB_ERR_STATE; }
    // The FP64 steps sum the start-of-step masses from pmass, which every
    // step leaves current; after an upload it is computed first. Do that
    // here, eagerly, so the capture (which executes nothing) neither marks
    // pmass current without computing it nor holds a launch that a replay
    // after an eager step would not need.
    if (!table_mode(c) && !c->mass_valid)
    {
        launch_mass_partials(c->view(), c->g, c->pmass, c->stream);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->mass_valid = true;
    }
    return graph_coefficients(c, dt);
}

// The table pipelines' coefficients for dt, built on the context's stream
// when they are not current (a graph replays the apply without building
// them: an eager step at another dt, or a wall-state change, leaves them
// stale).
int graph_coefficients(ab_ctx *c, double dt)
{
    if (table_mode(c) && c->coef_dt != dt)
    {
        int rc = table_prepare(c, dt);
        if (rc) return rc;
    }
    if (table2d(c) && c->t2_dt != dt)
    {
        int rc = table2d_prepare(c, dt);
        if (rc) return rc;
    }
    return AB_OK;
}

int graph_step(ab_ctx *c, double dt) { return do_step(c, dt); }

template <typename T>
__global__ void k_gather(const T *__restrict__ src, T *__restrict__ dst,
                         const int *__restrict__ map, int n, int stride)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * stride) return;
    const int i = (int)(e / stride), k = (int)(e - (size_t)i * stride);
    const int j = map[i];
    dst[e] = j >= 0 ? src[(size_t)j * stride + k] : T(0);
}

__global__ void k_gather_bid(const int *__restrict__ src, int *__restrict__ dst,
                             const int *__restrict__ map, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = map[i] >= 0 ? src[map[i]] : -1;
}

// Re-lays the cloud out for n particles: particle i takes the state of old
// particle src[i] (src: n entries on the device), or zero where src[i] < 0
// (a particle the caller fills in). The per-step scratch starts at zero;
// the neighbour-search scratch and the training-row ring, sized by N, are
// dropped. Neighbour lists and the interior/boundary lists are left to the
// caller.
int relayout_cloud(ab_ctx *c, const int *src, int n)
{
    hipStream_t s = c->stream;
    double **carried[] = {&c->x, &c->y, &c->z, &c->wnx, &c->wny, &c->wnz, &c->wux, &c->wuy,
                          &c->wuz, &c->wt, &c->mvx, &c->mvy, &c->mvz, &c->rho, &c->ux, &c->uy,
                          &c->uz, &c->T};
    constexpr int nc = sizeof(carried) / sizeof(carried[0]);
    double *od[nc];
    for (int k = 0; k < nc; ++k) { od[k] = *carried[k]; *carried[k] = nullptr; }
    double *og = c->g, *ogt = c->gt, *olim = c->limflag, *oovr = c->mlsovr, *opm = c->pmass;
    double *oalpha = c->dm_alpha, *ores = c->dm_res;
    int *opass = c->dm_pass, *omls = c->mlsflag, *omood = c->mood, *omf = c->moodfail;
    int *onb = c->neigh, *onn = c->nneigh, *oil = c->ilist, *obl = c->blist, *obid = c->bid;
    unsigned char *obnd = c->bnd;
    c->g = c->gt = c->limflag = c->mlsovr = c->pmass = c->dm_alpha = c->dm_res = nullptr;
    c->dm_pass = c->mlsflag = c->mood = c->moodfail = nullptr;
    c->neigh = c->nneigh = c->ilist = c->blist = c->bid = nullptr;
    c->bnd = nullptr;
    const int oldN = c->N;
    c->N = n;
    c->cfg.n_particles = n;
    int rc = alloc_cloud(c);
    auto gather = [&](auto *from, auto *to, int stride) {
        const size_t tot = (size_t)n * stride;
        if (!from || !to || !tot) return;
        k_gather<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(from, to, src, n, stride);
    };
    if (!rc)
    {
        for (int k = 0; k < nc; ++k) gather(od[k], *carried[k], 1);
        gather(obnd, c->bnd, 1);
        gather(og, c->g, c->gstride);
        gather(oalpha, c->dm_alpha, 5);
        gather(ores, c->dm_res, 1);
        gather(opass, c->dm_pass, 1);
        const int nb = (n + 255) / 256;
        // a particle filled in by the caller belongs to no body
        if (n) k_gather_bid<<<nb, 256, 0, s>>>(obid, c->bid, src, n);
        if (n)
        {
            k_fill_dbl<<<nb, 256, 0, s>>>(c->limflag, n, 0.0);
            k_fill_dbl<<<nb, 256, 0, s>>>(c->mlsovr, n, 0.0);
            k_fill_dbl<<<nb, 256, 0, s>>>(c->pmass, n, 0.0);
            k_fill_int<<<nb, 256, 0, s>>>(c->mlsflag, n, 0);
            k_fill_int<<<nb, 256, 0, s>>>(c->mood, n, 0);
            k_fill_int<<<nb, 256, 0, s>>>(c->nneigh, n, 0);
        }
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
        { c->err = "cloud re-layout: gather failed"; rc = AB_ERR_HIP; }
    }
    void *old[] = {og, ogt, olim, oovr, opm, oalpha, ores, opass, omls, omood, omf,
                   onb, onn, oil, obl, obnd, obid, c->vox, c->vlist, c->dm_rec};
    for (void *p : old) if (p) (void)hipFree(p);
    for (int k = 0; k < nc; ++k) if (od[k]) (void)hipFree(od[k]);
    c->vox = c->vlist = nullptr;
    c->dm_rec = nullptr;
    c->dm_rec_slots = c->dm_rec_next = c->dm_rec_filled = c->dm_rec_left = 0;
    c->mass_valid = false;
    c->gen++;
    if (rc) { free_cloud(c); c->uploaded = false; c->N = oldN; }
    return rc;
}

// Rebuilds the interior and boundary lists from the device boundary flags.
int rebuild_lists(ab_ctx *c)
{
    const int N = c->N;
    std::vector<unsigned char> b(N);
    if (N && hipMemcpyAsync(b.data(), c->bnd, N, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
    { c->err = "list rebuild: copy failed"; return AB_ERR_HIP; }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<int> il, bl;
    il.reserve(N); bl.reserve(N);
    for (int i = 0; i < N; ++i) (b[i] ? bl : il).push_back(i);
    c->n_int = (int)il.size(); c->n_bnd = (int)bl.size();
    if (c->n_int) HIPCHK(c, hipMemcpyAsync(c->ilist, il.data(), 4 * il.size(), hipMemcpyHostToDevice, c->stream));
    if (c->n_bnd) HIPCHK(c, hipMemcpyAsync(c->blist, bl.data(), 4 * bl.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return AB_OK;
}

constexpr int kRecSlots = 16;     // steps the training-row ring holds

__global__ void k_fill_f(float *p, size_t n, float v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

bool dm_on(const ab_ctx *c) { return c->cfg.collision == AB_COLLISION_DISCRETE && c->dim == 3; }

double dm_scale(const ab_ctx *c)
{
    const double vhi = c->vlo + (c->Nv - 1) * c->dv;
    return std::fmax(std::fabs(c->vlo), std::fabs(vhi));
}

// Per-step settings of the discrete-Maxwellian collision. When rows are
// being recorded, the slot the step writes is cleared first (weight 0 for
// every particle the kernels do not fill).
DmArgs dm_args(ab_ctx *c)
{
    DmArgs d{};
    if (!dm_on(c) || !c->dm_alpha) return d;
    d.on = 1;
    d.cs = dm_scale(c);
    d.tol = c->cfg.dm_tol > 0 ? c->cfg.dm_tol : 1e-12;
    d.maxit = c->cfg.dm_maxit > 0 ? c->cfg.dm_maxit : 8;
    d.alpha = c->dm_alpha;
    d.passes = c->dm_pass;
    d.res = c->dm_res;
    d.mlp = (c->mlp_on && c->mlp_ready) ? 1 : 0;
    d.wb = c->mlp_wb;
    d.pf = c->mlp_p;
    if (c->dm_rec_left > 0 && c->dm_rec)
    {
        float *slot = c->dm_rec + (size_t)c->dm_rec_next * c->N * kMlpRow;
        const size_t n = (size_t)c->N * kMlpRow;
        k_fill_f<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(slot, n, 0.0f);
        d.rec = slot;
    }
    return d;
}

void dm_step_done(ab_ctx *c)
{
    if (!dm_on(c) || c->dm_rec_left <= 0 || !c->dm_rec) return;
    c->dm_rec_next = (c->dm_rec_next + 1) % c->dm_rec_slots;
    if (c->dm_rec_filled < c->dm_rec_slots) c->dm_rec_filled++;
    c->dm_rec_left--;
}

constexpr int kEvPairs = 2048;

void prof_flush(ab_ctx *c)
{
    if (c->nrec == 0) return;
    (void)hipStreamSynchronize(c->stream);
    for (int k = 0; k < c->nrec; ++k)
    {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->evpool[2 * k], c->evpool[2 * k + 1]) == hipSuccess)
        {
            c->prof_ms[c->evstage[k]] += ms;
            c->prof_n[c->evstage[k]] += 1;
        }
    }
    c->nrec = 0;
}

void prof_begin(ab_ctx *c)
{
    if (!c->prof) return;
    if (c->evpool.empty())
    {
        c->evpool.resize(2 * kEvPairs, nullptr);
        c->evstage.resize(kEvPairs, 0);
        for (auto &e : c->evpool)
            if (hipEventCreate(&e) != hipSuccess) { c->prof = false; return; }
    }
    if (c->nrec >= kEvPairs) prof_flush(c);
    (void)hipEventRecord(c->evpool[2 * c->nrec], c->stream);
}

void prof_end(ab_ctx *c, int stage)
{
    if (!c->prof || c->evpool.empty()) return;
    (void)hipEventRecord(c->evpool[2 * c->nrec + 1], c->stream);
    c->evstage[c->nrec] = stage;
    c->nrec++;
}

}  // namespace ab

extern "C" {

int ab_check_config(const ab_config *cfg, char *msg, int cap)
{
    auto fail = [&](int code, const char *m) {
        if (msg && cap > 0) std::snprintf(msg, (size_t)cap, "%s", m);
        return code;
    };
    if (msg && cap > 0) msg[0] = 0;
    if (!cfg) return fail(AB_ERR_ARG, "null configuration");
    if ((cfg->dim != 2 && cfg->dim != 3) || cfg->n_particles <= 0 || cfg->nv < 2 ||
        cfg->nv > ab::kMaxNv || !(cfg->vmax > cfg->vmin) || !(cfg->R > 0) || !(cfg->tau > 0) ||
        !(cfg->radius > 0) || (cfg->mls_order != 1 && cfg->mls_order != 2) ||
        cfg->precision < AB_PREC_FP64 || cfg->precision > AB_PREC_FP64_TABLE ||
        (cfg->collision != AB_COLLISION_CONTINUOUS && cfg->collision != AB_COLLISION_DISCRETE))
        return fail(AB_ERR_ARG, "invalid configuration (dim, N, nv <= 48, velocity grid, gas, "
                                "radius, MLS order, precision or collision)");
    if (cfg->collision == AB_COLLISION_DISCRETE && cfg->dim != 3)
        return fail(AB_ERR_UNSUPPORTED, "the discrete-Maxwellian collision is 3-D only");
    if (cfg->limiter != AB_LIMITER_GPU && cfg->limiter != AB_LIMITER_MOOD)
        return fail(AB_ERR_ARG, "unknown limiter");
    if (cfg->limiter == AB_LIMITER_MOOD &&
        (cfg->precision != AB_PREC_FP64 || cfg->collision != AB_COLLISION_CONTINUOUS))
        return fail(AB_ERR_UNSUPPORTED, "the MOOD limiter runs with AB_PREC_FP64 and the "
                                        "continuous collision");
    if (cfg->dim == 3 && cfg->precision != AB_PREC_FP64 && cfg->nx > 0 && cfg->ny > 0)
    {
        // one z block of the table transport spans kMaxZBlockPlanes lattice
        // planes of the stored g, addressed with 32-bit byte offsets
        const long long nodes = (long long)cfg->nv * cfg->nv * cfg->nv;
        const long long S = (nodes + ab::kWave - 1) / ab::kWave * ab::kWave;
        const long long es = cfg->precision == AB_PREC_MIXED ? 4 : 8;
        const double span = (double)ab::kMaxZBlockPlanes * cfg->nx * cfg->ny * (double)S * es;
        if (span >= 2147483648.0)
            return fail(AB_ERR_UNSUPPORTED,
                        cfg->precision == AB_PREC_MIXED
                            ? "table transport: one z block of this lattice spans more than "
                              "2 GiB of g (nx * ny * Nv^3 too large)"
                            : "table transport: one z block of this lattice spans more than "
                              "2 GiB of FP64 g; run it with AB_PREC_MIXED or AB_PREC_FP64");
    }
    return AB_OK;
}

int ab_create(const ab_config *cfg, ab_ctx **out)
{
    if (!cfg || !out) return AB_ERR_ARG;
    *out = nullptr;
    int crc = ab_check_config(cfg, nullptr, 0);
    if (crc) return crc;
    ab_ctx *c = new (std::nothrow) ab_ctx;
    if (!c) return AB_ERR_NOMEM;
    c->cfg = *cfg;
    c->dim = cfg->dim;
    c->Nv = cfg->nv;
    c->nodes = (cfg->dim == 3) ? cfg->nv * cfg->nv * cfg->nv : cfg->nv * cfg->nv;
    c->gstride = (cfg->dim == 3) ? c->nodes : 2 * c->nodes;
    c->vlo = cfg->vmin;
    c->dv = (cfg->vmax - cfg->vmin) / (cfg->nv - 1);
    if (hipGetDevice(&c->device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipMalloc((void **)&c->scal, sizeof(double) * S_COUNT) != hipSuccess ||
        hipMemset(c->scal, 0, sizeof(double) * S_COUNT) != hipSuccess)
    {
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return AB_ERR_HIP;
    }
    *out = c;
    return AB_OK;
}

int ab_destroy(ab_ctx *c)
{
    if (!c) return AB_ERR_ARG;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_cloud(c);
    mlp_free(c);
    for (void *p : c->cm_pool) if (p) (void)hipFree(p);
    if (c->scal) (void)hipFree(c->scal);
    for (auto e : c->evpool) if (e) (void)hipEventDestroy(e);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return AB_OK;
}

const char *ab_last_error(const ab_ctx *c) { return c ? c->err.c_str() : "null context"; }

int ab_upload(ab_ctx *c, const ab_cloud_view *v)
{
    if (!c || !v) return AB_ERR_ARG;
    if (c->nbatch) { c->err = "a batch takes its clouds from its members (ab_batch_create)"; return AB_ERR_STATE; }
    if (!v->x || !v->y || !v->z || !v->boundary || !v->wnx || !v->wny || !v->wnz || !v->wux ||
        !v->wuy || !v->wuz || !v->wt || !v->rho || !v->ux || !v->uy || !v->uz || !v->t || !v->g)
    { c->err = "ab_upload: a required array is NULL"; return AB_ERR_ARG; }
    const int N = c->cfg.n_particles;
    if (c->N != N) { free_cloud(c); c->N = N; }
    c->N = N;
    int rc;
    if (!c->x && (rc = alloc_cloud(c))) { free_cloud(c); return rc; }
    hipStream_t s = c->stream;
    const size_t n8 = sizeof(double) * (size_t)N;
    auto up = [&](double *dst, const double *src) -> bool {
        if (src) return hipMemcpyAsync(dst, src, n8, hipMemcpyHostToDevice, s) == hipSuccess;
        return hipMemsetAsync(dst, 0, n8, s) == hipSuccess;
    };
    bool ok = up(c->x, v->x) && up(c->y, v->y) && up(c->z, v->z) && up(c->wnx, v->wnx) &&
              up(c->wny, v->wny) && up(c->wnz, v->wnz) && up(c->wux, v->wux) &&
              up(c->wuy, v->wuy) && up(c->wuz, v->wuz) && up(c->wt, v->wt) &&
              up(c->mvx, v->mvx) && up(c->mvy, v->mvy) && up(c->mvz, v->mvz) &&
              up(c->rho, v->rho) && up(c->ux, v->ux) && up(c->uy, v->uy) && up(c->uz, v->uz) &&
              up(c->T, v->t);
    if (!ok) { c->err = "ab_upload: host-to-device copy failed"; return AB_ERR_HIP; }
    HIPCHK(c, hipMemcpyAsync(c->bnd, v->boundary, N, hipMemcpyHostToDevice, s));
    if (v->body_id)
        HIPCHK(c, hipMemcpyAsync(c->bid, v->body_id, sizeof(int) * (size_t)N, hipMemcpyHostToDevice, s));
    else
        k_fill_int<<<(N + 255) / 256, 256, 0, s>>>(c->bid, N, -1);
    // interior / boundary lists
    std::vector<int> il, bl;
    il.reserve(N); bl.reserve(N);
    for (int i = 0; i < N; ++i) (v->boundary[i] ? bl : il).push_back(i);
    c->n_int = (int)il.size(); c->n_bnd = (int)bl.size();
    if (c->n_int) HIPCHK(c, hipMemcpyAsync(c->ilist, il.data(), 4 * il.size(), hipMemcpyHostToDevice, s));
    if (c->n_bnd) HIPCHK(c, hipMemcpyAsync(c->blist, bl.data(), 4 * bl.size(), hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(c->g, v->g, n8 * c->gstride, hipMemcpyHostToDevice, s));
    if (v->neigh && v->nneigh)
    {
        HIPCHK(c, hipMemcpyAsync(c->neigh, v->neigh, sizeof(int) * (size_t)N * AB_MAX_NEIGH,
                                 hipMemcpyHostToDevice, s));
        HIPCHK(c, hipMemcpyAsync(c->nneigh, v->nneigh, sizeof(int) * (size_t)N,
                                 hipMemcpyHostToDevice, s));
        // host lists are borrowed: finish the copies before returning
        HIPCHK(c, hipStreamSynchronize(s));
    }
    else
    {
        HIPCHK(c, hipStreamSynchronize(s));
        c->uploaded = true;
        if ((rc = ab_rebuild_neighbours(c))) { c->uploaded = false; return rc; }
    }
    c->uploaded = true;
    c->mass_valid = false;
    c->gen++;
    if (table_mode(c))
    {
        if ((rc = table_setup(c))) { c->uploaded = false; return rc; }
        c->coef_dt = -1.0;
        if ((rc = table_upload_g(c, c->g))) { c->err = "table upload failed"; return rc; }
    }
    if (table2d(c) && (rc = table2d_setup(c))) { c->uploaded = false; return rc; }
    HIPCHK(c, hipStreamSynchronize(s));
    return AB_OK;
}

int ab_rebuild_neighbours(ab_ctx *c)
{
    if (!c) return AB_ERR_ARG;
    if (!c->uploaded) { c->err = "ab_rebuild_neighbours before ab_upload"; return AB_ERR_STATE; }
    if (c->nbatch) { c->err = "a batch keeps its members' neighbour lists; rebuild them in the members"; return AB_ERR_UNSUPPORTED; }
    const ab_config &f = c->cfg;
    const int nbw = (c->dim == 3) ? f.nbwbox : 1;
    if (f.nbhbox <= 0 || f.nbvbox <= 0 || nbw <= 0 || f.hbox == 0 || f.vbox == 0 ||
        (c->dim == 3 && f.wbox == 0))
    { c->err = "neighbour rebuild needs the voxel grid in ab_config"; return AB_ERR_ARG; }
    const int nvox = f.nbhbox * f.nbvbox * nbw;
    if (nvox != c->nvox)
    {
        if (c->vcount) (void)hipFree(c->vcount);
        if (c->vstart) (void)hipFree(c->vstart);
        c->vcount = c->vstart = nullptr;
        int rc;
        if ((rc = dalloc(c, &c->vcount, (size_t)nvox)) || (rc = dalloc(c, &c->vstart, (size_t)nvox + 1)))
            return rc;
        c->nvox = nvox;
    }
    if (!c->vox)
    {
        int rc;
        if ((rc = dalloc(c, &c->vox, (size_t)c->N)) || (rc = dalloc(c, &c->vlist, 2 * (size_t)c->N)))
            return rc;
    }
    Timer tm(c);
    tm.tic();
    if (int e = launch_neighbours(c->view(), f, c->vox, c->vcount, c->vstart, c->vlist, nvox,
                                  c->neigh, c->nneigh, c->stream))
    {
        c->err = std::string("neighbour rebuild launch failed: ") +
                 hipGetErrorString((hipError_t)e);
        return AB_ERR_HIP;
    }
    tm.toc(ST_NEIGH);
    // the table transports take their stencils from the lists: check the
    // classes again (on the host) and rebuild the coefficients at the next
    // step; other clouds need no host synchronisation (later work on the
    // stream, downloads included, runs after the rebuild)
    int rc = AB_OK;
    if (table_mode(c) && c->gs)
    {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        rc = table_recheck(c);
        c->coef_dt = -1.0;
    }
    else if (table2d(c) && c->t2_coef)
    {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        rc = table2d_setup(c);
    }
    if (rc) c->uploaded = false;
    return rc;
}

int ab_download_neighbours(ab_ctx *c, int *neigh, int *nneigh)
{
    if (!c) return AB_ERR_ARG;
    if (!c->uploaded) { c->err = "no cloud uploaded"; return AB_ERR_STATE; }
    if (neigh) HIPCHK(c, hipMemcpyAsync(neigh, c->neigh, sizeof(int) * (size_t)c->N * AB_MAX_NEIGH,
                                        hipMemcpyDeviceToHost, c->stream));
    if (nneigh) HIPCHK(c, hipMemcpyAsync(nneigh, c->nneigh, sizeof(int) * (size_t)c->N,
                                         hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return AB_OK;
}

int ab_download_voxels(ab_ctx *c, int *voxel)
{
    if (!c || !voxel) return AB_ERR_ARG;
    if (!c->uploaded || !c->vox) { c->err = "no device neighbour rebuild has run"; return AB_ERR_STATE; }
    HIPCHK(c, hipMemcpyAsync(voxel, c->vox, sizeof(int) * (size_t)c->N, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return AB_OK;
}

int ab_step(ab_ctx *c, double dt, ab_step_stats *st)
{
    if (!c) return AB_ERR_ARG;
    int rc = do_step(c, dt);
    if (rc || !st) return rc;
    double sc[S_COUNT];
    HIPCHK(c, hipMemcpyAsync(sc, c->scal, sizeof sc, hipMemcpyDeviceToHost, c->stream));
    std::vector<double> lim(c->N), ovr(c->N);
    std::vector<int> fl(c->N);
    HIPCHK(c, hipMemcpyAsync(lim.data(), c->limflag, 8 * (size_t)c->N, hipMemcpyDeviceToHost, c->stream));
    if (!table_mode(c))
    {
        HIPCHK(c, hipMemcpyAsync(ovr.data(), c->mlsovr, 8 * (size_t)c->N, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(fl.data(), c->mlsflag, 4 * (size_t)c->N, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::memset(st, 0, sizeof *st);
    st->mass_old = table_mode(c) ? sc[S_MASS_START] : sc[S_MASS_OLD];
    st->mass_after_transport = sc[S_MASS_NEW];
    st->mass_after_collision = sc[S_MASS_COLL];
    for (int i = 0; i < c->N; ++i)
    {
        int f = (int)(lim[i] + 0.5);
        if (f & 1) st->n_positivity++;
        if (f & 2) st->n_nad++;
        if (f & 4) st->n_pad_fail++;
    }
    if (table_mode(c))
    {
        // the tables fold degeneracy into the coefficients and do not track
        // per-value extrapolation
        st->n_mls_degenerate = st->n_mls_extrapolate = -1;
        st->mls_ovr_sum = st->mls_ovr_max = -1.0;
    }
    else
        for (int i = 0; i < c->N; ++i)
        {
            if (fl[i] & 1) st->n_mls_degenerate++;
            if (fl[i] & 2) st->n_mls_extrapolate++;
            st->mls_ovr_sum += ovr[i];
            if (ovr[i] > st->mls_ovr_max) st->mls_ovr_max = ovr[i];
        }
    if (dm_on(c) && c->dm_pass)
    {
        std::vector<int> pass(c->N);
        std::vector<double> res(c->N);
        HIPCHK(c, hipMemcpyAsync(pass.data(), c->dm_pass, 4 * (size_t)c->N, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(res.data(), c->dm_res, 8 * (size_t)c->N, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        const int maxit = c->cfg.dm_maxit > 0 ? c->cfg.dm_maxit : 8;
        for (int i = 0; i < c->N; ++i)
        {
            if (pass[i] <= 0) continue;
            st->dm_particles++;
            st->dm_passes += pass[i];
            if (pass[i] > maxit + 1) st->dm_unconverged++;
            if (res[i] > st->dm_res_max) st->dm_res_max = res[i];
        }
    }
    return AB_OK;
}

__global__ void k_copy_calibration(const float4 *__restrict__ a, float4 *__restrict__ b, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

int ab_copy_calibration(unsigned long long bytes, int reps)
{
    if (bytes < 16 || reps < 1) return AB_ERR_ARG;
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc((void **)&a, bytes) != hipSuccess) return AB_ERR_NOMEM;
    if (hipMalloc((void **)&b, bytes) != hipSuccess) { (void)hipFree(a); return AB_ERR_NOMEM; }
    bool ok = hipMemset(a, 0, bytes) == hipSuccess && hipMemset(b, 0, bytes) == hipSuccess;
    const size_t n = bytes / 16;
    for (int r = 0; r < reps && ok; ++r)
    {
        k_copy_calibration<<<4096, 256>>>(a, b, n);
        ok = hipGetLastError() == hipSuccess;
    }
    ok = ok && hipDeviceSynchronize() == hipSuccess;
    (void)hipFree(a);
    (void)hipFree(b);
    return ok ? AB_OK : AB_ERR_HIP;
}

int ab_autotune(ab_ctx *c, double dt, int reps, int cap, char (*names)[32], float *ms, int *n,
                int *chosen)
{
    if (!c || reps < 1 || cap < 0 || !(dt > 0.0)) return AB_ERR_ARG;
    if (!c->uploaded) { c->err = "ab_autotune before ab_upload"; return AB_ERR_STATE; }
    if (table2d(c)) return table2d_autotune(c, dt, reps, cap, names, ms, n, chosen);
    if (!table_mode(c)) { c->err = "ab_autotune tunes the table pipelines' transport"; return AB_ERR_UNSUPPORTED; }
    return table_autotune(c, dt, reps, cap, names, ms, n, chosen);
}

int ab_set_transport_shape(ab_ctx *c, int shape)
{
    if (!c) return AB_ERR_ARG;
    if (table2d(c)) return table2d_set_run(c, shape);
    if (!table_mode(c)) { c->err = "transport shapes exist in the table pipelines"; return AB_ERR_UNSUPPORTED; }
    return table_set_shape(c, shape);
}

int ab_dm_download(ab_ctx *c, double *alpha, int *passes, double *residual)
{
    if (!c) return AB_ERR_ARG;
    if (!dm_on(c) || !c->dm_alpha) { c->err = "the context does not run the discrete-Maxwellian collision"; return AB_ERR_STATE; }
    const size_t N = (size_t)c->N;
    if (alpha) HIPCHK(c, hipMemcpyAsync(alpha, c->dm_alpha, 8 * 5 * N, hipMemcpyDeviceToHost, c->stream));
    if (passes) HIPCHK(c, hipMemcpyAsync(passes, c->dm_pass, 4 * N, hipMemcpyDeviceToHost, c->stream));
    if (residual) HIPCHK(c, hipMemcpyAsync(residual, c->dm_res, 8 * N, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return AB_OK;
}

int ab_dm_record(ab_ctx *c, int steps)
{
    if (!c || steps < 0) return AB_ERR_ARG;
    if (!dm_on(c) || !c->uploaded) { c->err = "ab_dm_record needs an uploaded discrete-Maxwellian context"; return AB_ERR_STATE; }
    if (!c->dm_rec)
    {
        int rc;
        if ((rc = dalloc(c, &c->dm_rec, (size_t)kRecSlots * c->N * kMlpRow))) return rc;
        c->dm_rec_slots = kRecSlots;
        c->dm_rec_next = c->dm_rec_filled = 0;
    }
    c->dm_rec_left = steps;
    return AB_OK;
}

int ab_dm_rows(ab_ctx *c, int *rows, int *valid)
{
    if (!c) return AB_ERR_ARG;
    const int n = c->dm_rec_filled * c->N;
    if (rows) *rows = n;
    if (valid)
    {
        *valid = 0;
        if (n > 0)
        {
            std::vector<float> r((size_t)n * kMlpRow);
            HIPCHK(c, hipMemcpyAsync(r.data(), c->dm_rec, sizeof(float) * r.size(), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            for (int k = 0; k < n; ++k) *valid += r[(size_t)k * kMlpRow + kMlpIn + kMlpOut] > 0.0f;
        }
    }
    return AB_OK;
}

int ab_steps(ab_ctx *c, double dt, int nsteps)
{
    if (!c || nsteps < 0) return AB_ERR_ARG;
    for (int k = 0; k < nsteps; ++k)
    {
        int rc = do_step(c, dt);
        if (rc) return rc;
    }
    return AB_OK;
}

int ab_transport_only(ab_ctx *c, double dt)
{
    if (!c) return AB_ERR_ARG;
    if (!c->uploaded) { c->err = "no cloud uploaded"; return AB_ERR_STATE; }
    if (table_mode(c)) { c->err = "ab_transport_only runs the AB_PREC_FP64 formulation"; return AB_ERR_UNSUPPORTED; }
    int rc;
    if ((rc = zero_step_flags(c))) return rc;
    if ((rc = fp64_transport(c, dt))) { c->err = "transport launch failed"; return rc; }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return AB_OK;
}

int ab_download_g(ab_ctx *c, double *g)
{
    if (!c || !g) return AB_ERR_ARG;
    if (!c->uploaded) { c->err = "no cloud uploaded"; return AB_ERR_STATE; }
    if (table_mode(c))
    {
        int rc = table_download_g(c, c->g);
        if (rc) { c->err = "table download failed"; return rc; }
    }
    HIPCHK(c, hipMemcpyAsync(g, c->g, sizeof(double) * (size_t)c->N * c->gstride,
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return AB_OK;
}

int ab_download_gt(ab_ctx *c, double *gt)
{
    if (!c || !gt) return AB_ERR_ARG;
    if (!c->uploaded || table_mode(c)) { c->err = "gt is kept by the AB_PREC_FP64 formulation only"; return AB_ERR_STATE; }
    HIPCHK(c, hipMemcpyAsync(gt, c->gt, sizeof(double) * (size_t)c->N * c->gstride,
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return AB_OK;
}

int ab_download_moments(ab_ctx *c, double *rho, double *ux, double *uy, double *uz, double *t,
                        double *limflag)
{
    if (!c) return AB_ERR_ARG;
    if (!c->uploaded) { c->err = "no cloud uploaded"; return AB_ERR_STATE; }
    const size_t n8 = sizeof(double) * (size_t)c->N;
    double *dst[] = {rho, ux, uy, uz, t, limflag};
    const double *src[] = {c->rho, c->ux, c->uy, c->uz, c->T, c->limflag};
    for (int k = 0; k < 6; ++k)
        if (dst[k]) HIPCHK(c, hipMemcpyAsync(dst[k], src[k], n8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return AB_OK;
}

int ab_download_particles(ab_ctx *c, double *x, double *y, double *z, unsigned char *boundary,
                          double *wall)
{
    if (!c) return AB_ERR_ARG;
    if (!c->uploaded) { c->err = "no cloud uploaded"; return AB_ERR_STATE; }
    const size_t N = (size_t)c->N, n8 = sizeof(double) * N;
    double *dst[] = {x, y, z};
    const double *src[] = {c->x, c->y, c->z};
    for (int k = 0; k < 3; ++k)
        if (dst[k]) HIPCHK(c, hipMemcpyAsync(dst[k], src[k], n8, hipMemcpyDeviceToHost, c->stream));
    if (boundary) HIPCHK(c, hipMemcpyAsync(boundary, c->bnd, N, hipMemcpyDeviceToHost, c->stream));
    std::vector<double> w;
    if (wall)
    {
        w.resize(7 * N);
        const double *ws[] = {c->wnx, c->wny, c->wnz, c->wux, c->wuy, c->wuz, c->wt};
        for (int k = 0; k < 7; ++k)
            HIPCHK(c, hipMemcpyAsync(w.data() + k * N, ws[k], n8, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (wall)
        for (size_t i = 0; i < N; ++i)
            for (int k = 0; k < 7; ++k) wall[7 * i + k] = w[k * N + i];
    return AB_OK;
}

int ab_total_mass(ab_ctx *c, double *mass)
{
    if (!c || !mass) return AB_ERR_ARG;
    if (!c->uploaded) { c->err = "no cloud uploaded"; return AB_ERR_STATE; }
    double sc[S_COUNT];
    if (table_mode(c))
    {
        // the step already leaves the current mass in S_MASS_OLD
        HIPCHK(c, hipMemcpyAsync(sc, c->scal, sizeof sc, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        *mass = sc[S_MASS_OLD];
        return AB_OK;
    }
    launch_mass_partials(c->view(), c->g, c->pmass, c->stream);
    launch_reduce_mass(c->N, c->pmass, c->scal, R_MASS_OLD, 0, c->stream);
    c->mass_valid = true;
    HIPCHK(c, hipMemcpyAsync(sc, c->scal, sizeof sc, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *mass = sc[S_MASS_OLD];
    return AB_OK;
}

int ab_set_wall_state(ab_ctx *c, int n, const int *idx, const double *x, const double *y,
                      const double *z, const double *wn, const double *wu, const double *mv)
{
    if (!c || n < 0 || (n > 0 && !idx)) return AB_ERR_ARG;
    if (!c->uploaded) { c->err = "no cloud uploaded"; return AB_ERR_STATE; }
    if (n == 0) return AB_OK;
    const bool hx = x && y && z;
    std::vector<double> buf((size_t)n * 12, 0.0);
    for (int k = 0; k < n; ++k)
    {
        if (idx[k] < 0 || idx[k] >= c->N) { c->err = "ab_set_wall_state: index out of range"; return AB_ERR_ARG; }
        double *s = &buf[(size_t)k * 12];
        if (hx) { s[0] = x[k]; s[1] = y[k]; s[2] = z[k]; }
        for (int a = 0; a < 3; ++a)
        {
            if (wn) s[3 + a] = wn[3 * k + a];
            if (wu) s[6 + a] = wu[3 * k + a];
            if (mv) s[9 + a] = mv[3 * k + a];
        }
    }
    int *di = nullptr;
    double *ds = nullptr;
    int rc;
    if ((rc = dalloc(c, &di, (size_t)n)) || (rc = dalloc(c, &ds, buf.size())))
    { if (di) (void)hipFree(di); return rc; }
    hipError_t e = hipMemcpyAsync(di, idx, 4 * (size_t)n, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(ds, buf.data(), 8 * buf.size(), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
    {
        k_scatter_wall<<<(n + 255) / 256, 256, 0, c->stream>>>(
            n, di, ds, hx, wn != nullptr, wu != nullptr, mv != nullptr, c->x, c->y, c->z, c->wnx,
            c->wny, c->wnz, c->wux, c->wuy, c->wuz, c->mvx, c->mvy, c->mvz);
        e = hipStreamSynchronize(c->stream);
    }
    (void)hipFree(di);
    (void)hipFree(ds);
    if (e != hipSuccess) { c->err = std::string("ab_set_wall_state: ") + hipGetErrorString(e); return AB_ERR_HIP; }
    // the table transports hold one stencil per lattice class: check the
    // classes again, and rebuild the coefficients at the next step
    if (table_mode(c) || table2d(c))
    {
        rc = table_mode(c) ? table_recheck(c) : table2d_setup(c);
        c->coef_dt = -1.0;
        if (rc) { c->uploaded = false; return rc; }
    }
    return AB_OK;
}

int ab_batch_create(ab_ctx *const *m, int n, ab_ctx **out)
{
    if (!m || !out || n <= 0) return AB_ERR_ARG;
    *out = nullptr;
    for (int k = 0; k < n; ++k)
    {
        if (!m[k]) return AB_ERR_ARG;
        if (!m[k]->uploaded || m[k]->nbatch) return AB_ERR_STATE;
    }
    const ab_config &f = m[0]->cfg;
    long total = 0;
    for (int k = 0; k < n; ++k)
    {
        const ab_config &g = m[k]->cfg;
        if (g.dim != 2 || g.nv != f.nv || g.vmin != f.vmin || g.vmax != f.vmax || g.R != f.R ||
            g.tau != f.tau || g.mls_order != f.mls_order || g.mls_tol != f.mls_tol ||
            g.mls_clamp != f.mls_clamp || g.no_mass_fix != f.no_mass_fix ||
            g.precision != f.precision || g.limiter != AB_LIMITER_GPU ||
            g.collision != AB_COLLISION_CONTINUOUS)
            return AB_ERR_UNSUPPORTED;
        total += m[k]->N;
    }
    if (total > (1L << 30)) return AB_ERR_ARG;
    ab_config cfg = f;
    cfg.n_particles = (int)total;
    cfg.nx = cfg.ny = 0;
    cfg.nz = 1;
    ab_ctx *b = nullptr;
    int rc = ab_create(&cfg, &b);
    if (rc) return rc;
    auto fail = [&](int code) { ab_destroy(b); return code; };
    b->N = (int)total;
    if ((rc = alloc_cloud(b))) return fail(rc);
    (void)hipFree(b->scal);
    b->scal = nullptr;
    if ((rc = dalloc(b, &b->scal, (size_t)n * S_COUNT)) || (rc = dalloc(b, &b->inst, (size_t)total)) ||
        (rc = dalloc(b, &b->ih, (size_t)n)) || (rc = dalloc(b, &b->seg, (size_t)n + 1)))
        return fail(rc);
    hipStream_t s = b->stream;
    b->bt_base.assign(1, 0);
    for (int k = 0; k < n; ++k)
    {
        ab_ctx *a = m[k];
        if (hipStreamSynchronize(a->stream) != hipSuccess) return fail(AB_ERR_HIP);
        const size_t o = (size_t)b->bt_base[k], nk = (size_t)a->N;
        const double *src[] = {a->x, a->y, a->z, a->wnx, a->wny, a->wnz, a->wux, a->wuy, a->wuz,
                               a->wt, a->mvx, a->mvy, a->mvz, a->rho, a->ux, a->uy, a->uz, a->T};
        double *dst[] = {b->x, b->y, b->z, b->wnx, b->wny, b->wnz, b->wux, b->wuy, b->wuz,
                         b->wt, b->mvx, b->mvy, b->mvz, b->rho, b->ux, b->uy, b->uz, b->T};
        bool ok = true;
        for (int q = 0; q < 18 && ok; ++q)
            ok = hipMemcpyAsync(dst[q] + o, src[q], 8 * nk, hipMemcpyDeviceToDevice, s) == hipSuccess;
        ok = ok && hipMemcpyAsync(b->bnd + o, a->bnd, nk, hipMemcpyDeviceToDevice, s) == hipSuccess &&
             hipMemcpyAsync(b->g + o * b->gstride, a->g, 8 * nk * b->gstride,
                            hipMemcpyDeviceToDevice, s) == hipSuccess &&
             hipMemcpyAsync(b->neigh + o * AB_MAX_NEIGH, a->neigh, 4 * nk * AB_MAX_NEIGH,
                            hipMemcpyDeviceToDevice, s) == hipSuccess &&
             hipMemcpyAsync(b->nneigh + o, a->nneigh, 4 * nk, hipMemcpyDeviceToDevice, s) == hipSuccess &&
             hipMemcpyAsync(b->bid + o, a->bid, 4 * nk, hipMemcpyDeviceToDevice, s) == hipSuccess;
        if (!ok) { b->err = "ab_batch_create: device copy failed"; return fail(AB_ERR_HIP); }
        const size_t ne = nk * AB_MAX_NEIGH;
        k_offset_neigh<<<(unsigned)((ne + 255) / 256), 256, 0, s>>>(b->neigh + o * AB_MAX_NEIGH,
                                                                    b->nneigh + o, (int)nk, (int)o);
        k_fill_int<<<(unsigned)((nk + 255) / 256), 256, 0, s>>>(b->inst + o, (int)nk, k);
        b->bt_base.push_back((int)(o + nk));
        b->bt_nx.push_back(a->cfg.nx);
        b->bt_ny.push_back(a->cfg.ny);
        b->bt_h.push_back(a->cfg.radius);
        b->members.push_back(a);
        b->member_gen.push_back(a->gen);
    }
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(b->ih, b->bt_h.data(), 8 * (size_t)n, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(b->seg, b->bt_base.data(), 4 * ((size_t)n + 1), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemsetAsync(b->scal, 0, sizeof(double) * S_COUNT * (size_t)n, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
    { b->err = "ab_batch_create: device copy failed"; return fail(AB_ERR_HIP); }
    b->nbatch = n;
    if ((rc = rebuild_lists(b))) return fail(rc);
    b->uploaded = true;
    b->mass_valid = false;
    if (table2d(b) && (rc = table2d_setup(b))) return fail(rc);
    *out = b;
    return AB_OK;
}

int ab_batch_scatter(ab_ctx *b)
{
    if (!b) return AB_ERR_ARG;
    if (!b->nbatch) { b->err = "ab_batch_scatter: not a batch"; return AB_ERR_STATE; }
    for (int k = 0; k < b->nbatch; ++k)
        if (b->members[k]->gen != b->member_gen[k] ||
            b->members[k]->N != b->bt_base[k + 1] - b->bt_base[k])
        {
            b->err = "ab_batch_scatter: member " + std::to_string(k) +
                     " was re-uploaded or re-laid out after the batch was formed";
            return AB_ERR_STATE;
        }
    HIPCHK(b, hipStreamSynchronize(b->stream));
    for (int k = 0; k < b->nbatch; ++k)
    {
        ab_ctx *a = b->members[k];
        const size_t o = (size_t)b->bt_base[k], nk = (size_t)a->N;
        const double *src[] = {b->rho, b->ux, b->uy, b->uz, b->T};
        double *dst[] = {a->rho, a->ux, a->uy, a->uz, a->T};
        HIPCHK(b, hipStreamSynchronize(a->stream));
        for (int q = 0; q < 5; ++q)
            HIPCHK(b, hipMemcpyAsync(dst[q], src[q] + o, 8 * nk, hipMemcpyDeviceToDevice, b->stream));
        HIPCHK(b, hipMemcpyAsync(a->g, b->g + o * b->gstride, 8 * nk * b->gstride,
                                 hipMemcpyDeviceToDevice, b->stream));
        // the member's per-particle masses follow its new g, so its next
        // step (eager or a graph replay) starts from the right sum
        launch_mass_partials(a->view(), a->g, a->pmass, b->stream);
        a->mass_valid = true;
    }
    HIPCHK(b, hipStreamSynchronize(b->stream));
    return AB_OK;
}

int ab_synchronize(ab_ctx *c)
{
    if (!c) return AB_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return AB_OK;
}

int ab_profile(ab_ctx *c, int enable)
{
    if (!c) return AB_ERR_ARG;
    prof_flush(c);
    c->prof = enable != 0;
    for (int k = 0; k < ST_COUNT; ++k) { c->prof_ms[k] = 0.0; c->prof_n[k] = 0; }
    return AB_OK;
}

int ab_profile_read(ab_ctx *c, double *ms, long *counts, int *nstages)
{
    if (!c) return AB_ERR_ARG;
    prof_flush(c);
    if (nstages) *nstages = ST_COUNT;
    for (int k = 0; k < ST_COUNT; ++k)
    {
        if (ms) ms[k] = c->prof_ms[k];
        if (counts) counts[k] = c->prof_n[k];
    }
    return AB_OK;
}

const char *ab_stage_name(int k)
{
    static const char *names[ST_COUNT] = {"transport", "limit_commit", "wall_collide",
                                          "final", "reduce", "neighbours"};
    return (k >= 0 && k < ST_COUNT) ? names[k] : "";
}

}  // extern "C"
