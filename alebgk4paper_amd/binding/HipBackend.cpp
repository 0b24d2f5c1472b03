// HipBackend.cpp -- the reference's GPU backend functions, declared at
// Solver/Solver.cpp:67-89 and defined in Solver/CudaBackend.cu, implemented on
// the C ABI of libalebgk_hip.so (include/alebgk_dev.h). A maintainer compiles
// this file in place of CudaBackend.cu (with -DALEBGK_WITH_CUDA, which keeps
// runSolver's GPU branch) and links libalebgk_hip.so; runSolver and the rest
// of the reference stay as they are.
//
// One ab_ctx replaces the file-static DeviceCloud (CudaBackend.cu:115).
// cudaUploadCloud receives no CalcParameters (the MLS radius is
// CalcParameters::radius), so the upload is staged on the host and handed to
// the device at the next call that has them (cudaStep, cudaSyncBodies);
// queries in between answer from the staged cloud. Errors print and exit(1),
// as the reference's CUCHECK does (CudaBackend.cu:58-66).
//
// Precision: ALEBGK_HIP_PRECISION = fp64 (default; the reference's
// formulation), mixed or fp64-table (coefficient tables; lattice clouds only,
// refused otherwise at the first step).
#include "Config.hpp"
#include "Types.hpp"
#include "Geometry.hpp"

#include "alebgk_dev.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

void moveRigidBodies(std::vector<Particle> &P, std::vector<std::unique_ptr<RigidBody>> &bodies,
                     double dt, const DomainBoundary &dom, const CalcParameters &cp);

namespace {

ab_ctx *g_ctx = nullptr;
ab_step_stats g_last{};          // report of the most recent step
int g_N = 0, g_gs = 0, g_dim = 3;
double g_dvd = 0.0;              // dv^dim: the mass quadrature weight
double g_R = GasConstants{}.R;   // the gas constant cudaStep was last given

// the staged upload (cudaUploadCloud before the parameters are known)
struct Staged
{
    bool pending = false;
    SimParameters sp;
    std::vector<double> x, y, z, wnx, wny, wnz, wux, wuy, wuz, wt, mvx, mvy, mvz;
    std::vector<double> rho, ux, uy, uz, T, g;
    std::vector<unsigned char> bnd;
    std::vector<int> nb, nn, bid;
} g_st;

[[noreturn]] void die(const char *what, const char *msg)
{
    std::fprintf(stderr, "[hip] %s failed: %s\n", what, msg);
    std::exit(1);
}

void check(int rc, const char *what)
{
    if (rc != AB_OK) die(what, g_ctx ? ab_last_error(g_ctx) : "no context");
}

int precision_from_env()
{
    const char *e = std::getenv("ALEBGK_HIP_PRECISION");
    if (!e || !*e || !std::strcmp(e, "fp64")) return AB_PREC_FP64;
    if (!std::strcmp(e, "mixed")) return AB_PREC_MIXED;
    if (!std::strcmp(e, "fp64-table")) return AB_PREC_FP64_TABLE;
    die("ALEBGK_HIP_PRECISION", "expected fp64, mixed or fp64-table");
}

// Hands the staged cloud to the device, creating the context the first
// time or when N or the velocity grid changed.
void flush(const CalcParameters &cp)
{
    if (!g_st.pending) return;
    const SimParameters &sp = g_st.sp;
    const int N = g_N;
    ab_config c{};
    c.dim = g_dim;
    c.n_particles = N;
    c.nv = sp.Nv;
    c.vmin = sp.VMin;
    c.vmax = sp.VMax;
    c.R = g_R;
    c.tau = sp.tao;
    c.radius = cp.radius;
    c.mls_order = sp.mlsOrder;
    c.mls_tol = 1e-12;
    c.precision = precision_from_env();
    const long lat = (long)sp.Nx * sp.Ny * (g_dim == 3 ? sp.Nz : 1);
    if (lat == N)
    {
        c.nx = sp.Nx;
        c.ny = sp.Ny;
        c.nz = (g_dim == 3) ? sp.Nz : 1;
    }
    c.hbox = cp.hBox; c.vbox = cp.vBox; c.wbox = cp.wBox;
    c.nbhbox = cp.nbhBox; c.nbvbox = cp.nbvBox; c.nbwbox = cp.nbwBox;
    if (sp.limiter == LimiterMode::MOOD && c.precision == AB_PREC_FP64 &&
        std::getenv("ALEBGK_HIP_CPU_SEMANTICS"))
        c.limiter = AB_LIMITER_MOOD;
    if (g_ctx)
    {
        ab_destroy(g_ctx);
        g_ctx = nullptr;
    }
    char msg[256];
    if (int rc = ab_check_config(&c, msg, sizeof msg)) { (void)rc; die("ab_create", msg); }
    check(ab_create(&c, &g_ctx), "ab_create");
    Staged &s = g_st;
    ab_cloud_view v{s.x.data(), s.y.data(), s.z.data(), s.bnd.data(), s.wnx.data(),
                    s.wny.data(), s.wnz.data(), s.wux.data(), s.wuy.data(), s.wuz.data(),
                    s.wt.data(), s.mvx.data(), s.mvy.data(), s.mvz.data(), s.rho.data(),
                    s.ux.data(), s.uy.data(), s.uz.data(), s.T.data(), s.g.data(),
                    s.nb.data(), s.nn.data(), s.bid.data()};
    check(ab_upload(g_ctx, &v), "ab_upload");
    s = Staged{};
}

}  // namespace

// cudaUploadCloud (CudaBackend.cu:1223): the SoA gather of the particles.
void cudaUploadCloud(const std::vector<Particle> &P, const SimParameters &sp)
{
    const int N = (int)P.size();
    if (N == 0) die("cudaUploadCloud", "empty cloud");
    const int dim = P[0].dim, Nv = sp.Nv;
    g_N = N;
    g_dim = dim;
    g_gs = (dim == 3) ? Nv * Nv * Nv : 2 * Nv * Nv;
    const double dv = (sp.VMax - sp.VMin) / (Nv - 1);
    g_dvd = (dim == 3) ? dv * dv * dv : dv * dv;
    Staged &s = g_st;
    s = Staged{};
    s.sp = sp;
    for (auto *a : {&s.x, &s.y, &s.z, &s.wnx, &s.wny, &s.wnz, &s.wux, &s.wuy, &s.wuz, &s.wt,
                    &s.mvx, &s.mvy, &s.mvz, &s.rho, &s.ux, &s.uy, &s.uz, &s.T})
        a->resize(N);
    s.g.resize((size_t)N * g_gs);
    s.bnd.resize(N);
    s.nb.assign((size_t)N * AB_MAX_NEIGH, 0);
    s.nn.resize(N);
    s.bid.resize(N);
    for (int i = 0; i < N; ++i)
    {
        const Particle &p = P[i];
        s.x[i] = p.x; s.y[i] = p.y; s.z[i] = p.z; s.bnd[i] = p.boundary ? 1 : 0;
        s.wnx[i] = p.wnx; s.wny[i] = p.wny; s.wnz[i] = p.wnz;
        s.wux[i] = p.wUx; s.wuy[i] = p.wUy; s.wuz[i] = p.wUz; s.wt[i] = p.wT;
        s.mvx[i] = p.meshVx; s.mvy[i] = p.meshVy; s.mvz[i] = p.meshVz;
        s.rho[i] = p.rho; s.ux[i] = p.ux; s.uy[i] = p.uy; s.uz[i] = p.uz; s.T[i] = p.T;
        s.bid[i] = p.bodyId;
        if ((int)p.g.size() < g_gs) die("cudaUploadCloud", "a particle's g is shorter than Nv^d");
        std::copy(p.g.begin(), p.g.begin() + g_gs, s.g.begin() + (size_t)i * g_gs);
        s.nn[i] = (int)std::min<size_t>(p.neighindex.size(), AB_MAX_NEIGH);
        std::copy(p.neighindex.begin(), p.neighindex.begin() + s.nn[i],
                  s.nb.begin() + (size_t)i * AB_MAX_NEIGH);
    }
    s.pending = true;
}

// cudaStep (CudaBackend.cu:1505).
void cudaStep(const SimParameters &sp, const CalcParameters &cp, double dt, double R)
{
    (void)sp;
    if (g_ctx && R != g_R && !g_st.pending)
        die("cudaStep", "the gas constant changed after the upload");
    g_R = R;
    flush(cp);
    check(ab_step(g_ctx, dt, &g_last), "ab_step");
}

// cudaDownloadMoments (CudaBackend.cu:1917): rho, u, T, the limiter flags,
// and p = rho T as the reference's GPU download writes it (:1971).
void cudaDownloadMoments(std::vector<Particle> &P)
{
    const int N = g_N;
    if ((int)P.size() != N) die("cudaDownloadMoments", "particle count differs from the upload");
    if (g_st.pending) return;     // nothing stepped since the upload: P holds the state
    std::vector<double> rho(N), ux(N), uy(N), uz(N), T(N), lim(N);
    check(ab_download_moments(g_ctx, rho.data(), ux.data(), uy.data(), uz.data(), T.data(),
                              lim.data()),
          "ab_download_moments");
    for (int i = 0; i < N; ++i)
    {
        Particle &p = P[i];
        p.rho = rho[i]; p.ux = ux[i]; p.uy = uy[i];
        if (g_dim == 3) p.uz = uz[i];
        p.T = T[i];
        p.moodFlag = ((int)(lim[i] + 0.5)) != 0;
        p.p = p.rho * p.T;
    }
}

// cudaDownloadFull (CudaBackend.cu:2056): the moments and g.
void cudaDownloadFull(std::vector<Particle> &P)
{
    cudaDownloadMoments(P);
    if (g_st.pending) return;
    std::vector<double> g((size_t)g_N * g_gs);
    check(ab_download_g(g_ctx, g.data()), "ab_download_g");
    for (int i = 0; i < g_N; ++i)
        std::copy(g.begin() + (size_t)i * g_gs, g.begin() + (size_t)(i + 1) * g_gs,
                  P[i].g.begin());
}

// cudaSyncBodies (CudaBackend.cu:1884): the surface particles' g to the host,
// the reference's moveRigidBodies there, and the moved surface's positions,
// wall state and mesh velocity back to the device.
void cudaSyncBodies(std::vector<Particle> &P, std::vector<std::unique_ptr<RigidBody>> &bodies,
                    double dt, const DomainBoundary &dom, const CalcParameters &cp)
{
    flush(cp);
    const int N = g_N;
    std::vector<double> g((size_t)N * g_gs);
    check(ab_download_g(g_ctx, g.data()), "ab_download_g");
    std::vector<int> idx;
    for (int i = 0; i < N; ++i)
        if (P[i].boundary && P[i].bodyId >= 0)
        {
            std::copy(g.begin() + (size_t)i * g_gs, g.begin() + (size_t)(i + 1) * g_gs,
                      P[i].g.begin());
            idx.push_back(i);
        }
    moveRigidBodies(P, bodies, dt, dom, cp);
    const int n = (int)idx.size();
    std::vector<double> x(n), y(n), z(n), wn(3 * (size_t)n), wu(3 * (size_t)n),
        mv(3 * (size_t)n);
    for (int k = 0; k < n; ++k)
    {
        const Particle &p = P[idx[k]];
        x[k] = p.x; y[k] = p.y; z[k] = p.z;
        const double a[3] = {p.wnx, p.wny, p.wnz}, b[3] = {p.wUx, p.wUy, p.wUz},
                     c[3] = {p.meshVx, p.meshVy, p.meshVz};
        std::memcpy(&wn[3 * k], a, sizeof a);
        std::memcpy(&wu[3 * k], b, sizeof b);
        std::memcpy(&mv[3 * k], c, sizeof c);
    }
    check(ab_set_wall_state(g_ctx, n, idx.data(), x.data(), y.data(), z.data(), wn.data(),
                            wu.data(), mv.data()),
          "ab_set_wall_state");
}

// The step reports (CudaBackend.cu:1986-2030), from the last step's stats.
void cudaLimiterCounts(int *nPositivity, int *nNAD)
{
    if (nPositivity) *nPositivity = g_last.n_positivity;
    if (nNAD) *nNAD = g_last.n_nad;
}

void cudaMlsCounts(int *nDegenerate, int *nExtrapolate)
{
    if (nDegenerate) *nDegenerate = g_last.n_mls_degenerate;
    if (nExtrapolate) *nExtrapolate = g_last.n_mls_extrapolate;
}

void cudaMlsOvershoot(double *sum, double *worst)
{
    if (sum) *sum = g_last.mls_ovr_sum;
    if (worst) *worst = g_last.mls_ovr_max;
}

void cudaMlsOvershootNow(double *sum, double *worst) { cudaMlsOvershoot(sum, worst); }

double cudaTotalMass()
{
    if (g_st.pending)
    {
        double m = 0.0;
        const int nodes = (g_dim == 3) ? g_gs : g_gs / 2;
        for (int i = 0; i < g_N; ++i)
        {
            double s = 0.0;
            for (int k = 0; k < nodes; ++k) s += g_st.g[(size_t)i * g_gs + k];
            m += s * g_dvd;
        }
        return m;
    }
    double m = 0.0;
    check(ab_total_mass(g_ctx, &m), "ab_total_mass");
    return m;
}

void cudaMassBudget(double *afterCorrection, double *afterCollision)
{
    if (afterCorrection) *afterCorrection = g_last.mass_after_transport;
    if (afterCollision) *afterCollision = g_last.mass_after_collision;
}

// cudaWallMomentum (CudaBackend.cu:1977) reads the per-particle x-momentum
// change of the wall step, a diagnostic this library does not record: the
// reference returns zeros when that array is absent, and so does this.
void cudaWallMomentum(std::vector<double> &out)
{
    static bool told = false;
    if (!told)
    {
        std::fprintf(stderr, "[hip] cudaWallMomentum: the wall momentum diagnostic is not "
                             "recorded by libalebgk_hip; reporting zeros\n");
        told = true;
    }
    out.assign(g_N, 0.0);
}
