// main.cpp -- alebgk_amd: the reference's program on the MI355X library.
//
// The command line of TestCases/main.cpp:114-366 (case by name or number,
// --fixed, --Nx, --Nv, --dt, --tfinal, --body-scale, --estimate, --tag; --slab
// is refused: this build has no MPI) and the GPU time loop of runSolver
// (Solver/Solver.cpp:256-590): the cloud is built on the host as main() builds
// it, uploaded once, stepped on the device (with the moving-body block every
// step for a free body), and every saveEvery steps the moments are downloaded
// for the step line, the stability guard, the VTK file and timeseries.csv;
// bodies.csv logs the body state; ALEBGK_CHECKPOINT_EVERY / ALEBGK_RESTART
// write and resume restart.ckpt (format v1).
//
// Options beyond the reference's: --precision fp64|mixed|fp64-table (the
// reference's formulation in FP64 by default), --mls-order 1|2 (default 1,
// SimParameters::mlsOrder), --collision continuous|discrete, --limiter gpu|mood,
// --out DIR (the output root, default "output").
#include "alebgk_host.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <vector>

namespace {

struct Options
{
    std::string token, tag, out_root = "output";
    bool fixed = false;
    double nx = -1, nv = -1, dt = -1, tfinal = -1, body_scale = 1.0, estimate = -1;
    int precision = AB_PREC_FP64, mls_order = 1, collision = AB_COLLISION_CONTINUOUS;
    int limiter = AB_LIMITER_GPU;
};

void list_cases()
{
    std::printf("\nAvailable cases:\n\n");
    const char *names[6] = {"cavity2d", "cavity3d", "cavity2d-square", "cavity2d-circle",
                            "cavity3d-sphere", "cavity3d-cube"};
    const char *desc[6] = {"2-D lid-driven cavity", "3-D lid-driven cavity",
                           "2-D lid-driven cavity with a square body",
                           "2-D lid-driven cavity with a circular body",
                           "3-D lid-driven cavity with a spherical body",
                           "3-D lid-driven cavity with a cubic body"};
    for (int i = 0; i < 6; ++i) std::printf("  %d  %-17s %s\n", i + 1, names[i], desc[i]);
    std::printf("\nOptions:\n"
                "  --fixed          hold the body at its initial pose (body cases)\n"
                "  --Nx N           particles per spatial direction\n"
                "  --Nv N           velocity-grid points per direction\n"
                "  --dt T           time step, s\n"
                "  --tfinal T       final time, s\n"
                "  --body-scale S   scale the body's linear size (1.0 = reference)\n"
                "  --estimate N     project the whole run from step 2; 0 = stop there\n"
                "  --tag NAME       suffix for the output directory\n"
                "  --precision P    fp64 (default), mixed or fp64-table\n"
                "  --mls-order K    1 (default, linear) or 2 (quadratic)\n"
                "  --collision C    continuous (default) or discrete\n"
                "  --limiter L      gpu (default: the reference's GPU step) or mood\n"
                "  --out DIR        output root (default output)\n"
                "\nExample:  alebgk_amd cavity3d-cube --Nx 50 --tfinal 1.8e-7\n\n");
}

bool next_value(int argc, char **argv, int &i, double *out)
{
    if (i + 1 >= argc) return false;
    char *end = nullptr;
    const double v = std::strtod(argv[++i], &end);
    if (end == argv[i]) return false;
    *out = v;
    return true;
}

// 0 ok, 1 usage error (message printed)
int parse(int argc, char **argv, Options &o)
{
    o.token = argv[1];
    for (int i = 2; i < argc; ++i)
    {
        const std::string a = argv[i];
        bool ok = true;
        auto word = [&](std::string &dst) {
            if (i + 1 >= argc) return false;
            dst = argv[++i];
            return true;
        };
        std::string w;
        if (a == "--fixed") o.fixed = true;
        else if (a == "--Nx") ok = next_value(argc, argv, i, &o.nx);
        else if (a == "--Nv") ok = next_value(argc, argv, i, &o.nv);
        else if (a == "--dt") ok = next_value(argc, argv, i, &o.dt);
        else if (a == "--tfinal") ok = next_value(argc, argv, i, &o.tfinal);
        else if (a == "--body-scale") ok = next_value(argc, argv, i, &o.body_scale);
        else if (a == "--estimate") ok = next_value(argc, argv, i, &o.estimate);
        else if (a == "--tag") ok = word(o.tag);
        else if (a == "--out") ok = word(o.out_root);
        else if (a == "--slab")
        {
            std::printf("--slab asks for the MPI slab decomposition; this build runs one "
                        "cloud on one GPU.\n");
            return 1;
        }
        else if (a == "--precision")
        {
            ok = word(w);
            if (ok && w == "fp64") o.precision = AB_PREC_FP64;
            else if (ok && w == "mixed") o.precision = AB_PREC_MIXED;
            else if (ok && w == "fp64-table") o.precision = AB_PREC_FP64_TABLE;
            else ok = false;
        }
        else if (a == "--mls-order")
        {
            double v = 0;
            ok = next_value(argc, argv, i, &v) && (v == 1 || v == 2);
            o.mls_order = (int)v;
        }
        else if (a == "--collision")
        {
            ok = word(w);
            if (ok && w == "continuous") o.collision = AB_COLLISION_CONTINUOUS;
            else if (ok && w == "discrete") o.collision = AB_COLLISION_DISCRETE;
            else ok = false;
        }
        else if (a == "--limiter")
        {
            ok = word(w);
            if (ok && w == "gpu") o.limiter = AB_LIMITER_GPU;
            else if (ok && w == "mood") o.limiter = AB_LIMITER_MOOD;
            else ok = false;
        }
        else
        {
            std::printf("Unknown option '%s'.\n", a.c_str());
            list_cases();
            return 1;
        }
        if (!ok)
        {
            std::printf("Option '%s' needs a valid value.\n", a.c_str());
            return 1;
        }
    }
    return 0;
}

void make_dirs(const std::string &path)
{
    std::string cur;
    for (size_t k = 0; k <= path.size(); ++k)
    {
        if (k == path.size() || path[k] == '/')
        {
            if (!cur.empty()) mkdir(cur.c_str(), 0755);
        }
        if (k < path.size()) cur += path[k];
    }
}

// The host copy of the cloud, with the fields the outputs and restart files
// need.
struct Cloud
{
    ab_config cfg{};
    int N = 0, gs = 0;
    std::vector<double> x, y, z, wall, mom, g, mesh;
    std::vector<unsigned char> bnd;
    std::vector<int> bid;
    std::vector<ab_rigid_body> bodies;
    double geo[4] = {0, 0, 0, 0};

    void resize(int n)
    {
        N = n;
        x.resize(n); y.resize(n); z.resize(n); bnd.resize(n); bid.assign(n, -1);
        wall.resize(7 * (size_t)n); mom.resize(5 * (size_t)n); mesh.assign(3 * (size_t)n, 0.0);
        g.resize((size_t)n * gs);
    }
};

[[noreturn]] void fail(ab_ctx *c, const char *what, int rc)
{
    std::fprintf(stderr, "[alebgk_amd] %s failed (%d): %s\n", what, rc,
                 c ? ab_last_error(c) : "");
    std::exit(1);
}

// The device moments and particles as the outputs take them (the GPU
// download of the reference sets p = rho T, CudaBackend.cu:1971).
struct Snapshot
{
    std::vector<double> x, y, z, rho, ux, uy, uz, T, p, lim, wall;
    std::vector<unsigned char> bnd, mood;
};

void snapshot(ab_ctx *c, int N, Snapshot &s)
{
    for (auto *v : {&s.x, &s.y, &s.z, &s.rho, &s.ux, &s.uy, &s.uz, &s.T, &s.p, &s.lim})
        v->resize(N);
    s.wall.resize(7 * (size_t)N);
    s.bnd.resize(N);
    s.mood.resize(N);
    int rc = ab_download_moments(c, s.rho.data(), s.ux.data(), s.uy.data(), s.uz.data(),
                                 s.T.data(), s.lim.data());
    if (rc) fail(c, "ab_download_moments", rc);
    rc = ab_download_particles(c, s.x.data(), s.y.data(), s.z.data(), s.bnd.data(), s.wall.data());
    if (rc) fail(c, "ab_download_particles", rc);
    for (int i = 0; i < N; ++i)
    {
        s.p[i] = s.rho[i] * s.T[i];
        s.mood[i] = ((int)(s.lim[i] + 0.5)) != 0;
    }
}

void write_outputs(const Snapshot &s, int N, int nv, const std::string &dir, int step, double t,
                   alh_diag &d)
{
    alh_diagnostics(N, s.bnd.data(), s.rho.data(), s.ux.data(), s.uy.data(), s.uz.data(),
                    s.T.data(), s.p.data(), s.mood.data(), &d);
    alh_write_vtk(dir.c_str(), step, t, N, s.x.data(), s.y.data(), s.z.data(), s.rho.data(),
                  s.T.data(), s.p.data(), nv, s.mood.data(), s.bnd.data(), s.ux.data(),
                  s.uy.data(), s.uz.data());
}

// ckpt::save of the device state (Core/Checkpoint.hpp:78): g, moments
// (p = rho R T), positions, wall state, mesh velocity (the surface's wall
// velocity on a moving body, zero elsewhere), voxels, body ids and bodies.
int save_restart(ab_ctx *c, const Cloud &cl, const std::string &path, long long step,
                 double t, double R)
{
    const int N = c ? cl.N : 0;
    Snapshot s;
    snapshot(c, N, s);
    std::vector<double> g((size_t)N * cl.gs), mom(6 * (size_t)N), mesh(3 * (size_t)N, 0.0),
        grid(3 * (size_t)N);
    std::vector<int> ints(4 * (size_t)N), vox(N), bid(N);
    std::vector<unsigned char> flags(3 * (size_t)N);
    int rc = ab_download_g(c, g.data());
    if (!rc) rc = ab_download_voxels(c, vox.data());
    if (!rc) rc = ab_download_body_ids(c, bid.data());
    if (rc) return rc;
    const double dv = (cl.cfg.vmax - cl.cfg.vmin) / (cl.cfg.nv - 1);
    for (int i = 0; i < N; ++i)
    {
        const double m[6] = {s.rho[i], s.ux[i], s.uy[i], s.uz[i], s.T[i], s.rho[i] * R * s.T[i]};
        std::memcpy(&mom[6 * (size_t)i], m, sizeof m);
        if (s.bnd[i] && bid[i] >= 0)
        {
            mesh[3 * (size_t)i] = s.wall[7 * (size_t)i + 3];
            mesh[3 * (size_t)i + 1] = s.wall[7 * (size_t)i + 4];
            if (cl.cfg.dim == 3) mesh[3 * (size_t)i + 2] = s.wall[7 * (size_t)i + 5];
        }
        grid[3 * (size_t)i] = cl.cfg.vmin; grid[3 * (size_t)i + 1] = cl.cfg.vmax;
        grid[3 * (size_t)i + 2] = dv;
        ints[4 * (size_t)i] = cl.cfg.nv; ints[4 * (size_t)i + 1] = cl.cfg.dim;
        ints[4 * (size_t)i + 2] = bid[i]; ints[4 * (size_t)i + 3] = vox[i];
        flags[3 * (size_t)i] = s.bnd[i]; flags[3 * (size_t)i + 1] = 1;
        flags[3 * (size_t)i + 2] = s.mood[i];
    }
    std::vector<double> brec(30 * cl.bodies.size(), 0.0);
    for (size_t b = 0; b < cl.bodies.size(); ++b)
    {
        const ab_rigid_body &r = cl.bodies[b];
        double *q = &brec[30 * b];
        const double v[19] = {r.c[0], r.c[1], r.c[2], r.vel[0], r.vel[1], r.vel[2],
                              r.angular_vel, r.om[0], r.om[1], r.om[2], r.force[0],
                              r.force[1], r.force[2], r.tq[2], r.tq[0], r.tq[1], r.tq[2],
                              r.o[0], r.o[1]};
        std::memcpy(q, v, sizeof v);
        std::memcpy(q + 19, r.rm, 9 * sizeof(double));
        q[28] = cl.cfg.dim;
        q[29] = r.rotates;
    }
    alh_ckpt_info info{};
    info.dim = cl.cfg.dim;
    info.n = N;
    info.gstride = cl.gs;
    info.step = step;
    info.t = t;
    info.nbodies = (int)cl.bodies.size();
    alh_ckpt_fields f{s.x.data(), s.y.data(), s.z.data(), mom.data(), mesh.data(), s.wall.data(),
                      grid.data(), ints.data(), flags.data(), g.data(), brec.data()};
    return alh_ckpt_write(path.c_str(), &info, &f);
}

// ckpt::restore (Core/Checkpoint.hpp:161) into the host cloud: every field
// the file holds, the bodies' state onto the case's bodies.
int load_restart(const std::string &path, Cloud &cl, long long *step, double *t)
{
    alh_ckpt_info info{};
    int rc = alh_ckpt_info_read(path.c_str(), &info);
    if (rc) return rc;
    if (info.dim != cl.cfg.dim || info.gstride != cl.gs || info.nbodies != (int)cl.bodies.size())
        return AB_ERR_UNSUPPORTED;
    cl.resize((int)info.n);
    cl.cfg.n_particles = (int)info.n;
    std::vector<double> mom(6 * (size_t)cl.N), grid(3 * (size_t)cl.N),
        brec(30 * (size_t)std::max(info.nbodies, 1));
    std::vector<int> ints(4 * (size_t)cl.N);
    std::vector<unsigned char> flags(3 * (size_t)cl.N);
    alh_ckpt_fields f{cl.x.data(), cl.y.data(), cl.z.data(), mom.data(), cl.mesh.data(),
                      cl.wall.data(), grid.data(), ints.data(), flags.data(), cl.g.data(),
                      brec.data()};
    if ((rc = alh_ckpt_read(path.c_str(), &info, &f))) return rc;
    for (int i = 0; i < cl.N; ++i)
    {
        for (int k = 0; k < 5; ++k) cl.mom[5 * (size_t)i + k] = mom[6 * (size_t)i + k];
        cl.bnd[i] = flags[3 * (size_t)i];
        cl.bid[i] = ints[4 * (size_t)i + 2];
    }
    for (size_t b = 0; b < cl.bodies.size(); ++b)
    {
        ab_rigid_body &r = cl.bodies[b];
        const double *q = &brec[30 * b];
        r.c[0] = q[0]; r.c[1] = q[1]; r.c[2] = q[2];
        r.vel[0] = q[3]; r.vel[1] = q[4]; r.vel[2] = q[5];
        r.angular_vel = q[6];
        r.om[0] = q[7]; r.om[1] = q[8]; r.om[2] = q[9];
        r.force[0] = q[10]; r.force[1] = q[11]; r.force[2] = q[12];
        r.tq[0] = q[14]; r.tq[1] = q[15]; r.tq[2] = q[16];
        r.o[0] = q[17]; r.o[1] = q[18];
        std::memcpy(r.rm, q + 19, 9 * sizeof(double));
        r.rotates = q[29] != 0.0;
    }
    *step = info.step;
    *t = info.t;
    return AB_OK;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 2)
    {
        std::printf("Usage: alebgk_amd <case> [options]\n");
        list_cases();
        return 1;
    }
    Options o;
    if (parse(argc, argv, o)) return 1;
    alh_case cs{};
    if (alh_configure(o.token.c_str(), o.nx > 0 ? (int)o.nx : 0, o.nv > 0 ? (int)o.nv : 0, &cs))
    {
        std::printf("Unknown case '%s'.\n", o.token.c_str());
        list_cases();
        return 1;
    }
    if (o.dt > 0) cs.dt = o.dt;
    if (o.tfinal > 0) cs.tfinal = o.tfinal;
    cs.mls_order = o.mls_order;
    const bool moving = cs.has_body && !o.fixed;
    std::string out_dir = o.out_root + "/" + cs.name + "_Nx" + std::to_string(cs.nx) + "_Nv" +
                          std::to_string(cs.nv);
    if (o.fixed && cs.has_body) out_dir += "_fixed";
    if (!o.tag.empty()) out_dir += "_" + o.tag;

    std::printf("\n==============================================\n");
    std::printf("  ALE-BGK meshfree kinetic solver (MI355X)\n");
    std::printf("  Case %d: %s\n", cs.id, cs.name);
    if (cs.has_body)
        std::printf("  Body:   %s\n", moving ? "free to translate and rotate" : "fixed");
    std::printf("  Grid:   %d^%d particles, Nv = %d\n", cs.nx, cs.dim, cs.nv);
    std::printf("  Time:   dt = %.3e s, tfinal = %.3e s (%ld steps)\n", cs.dt, cs.tfinal,
                long(cs.tfinal / cs.dt + 0.5));
    std::printf("  Output: %s\n", out_dir.c_str());
    std::printf("==============================================\n\n");
    std::fflush(stdout);
    const auto t_start = std::chrono::steady_clock::now();

    // the cloud, as main() builds it
    Cloud cl;
    cl.gs = (cs.dim == 3) ? cs.nv * cs.nv * cs.nv : 2 * cs.nv * cs.nv;
    int N = 0, rc;
    if (cs.has_body)
    {
        if ((rc = alh_body_case_size(&cs, o.body_scale, &N))) fail(nullptr, "alh_body_case_size", rc);
        cl.resize(N);
        cl.bodies.resize(4);
        int nb = 0;
        rc = alh_build_body_case(&cs, o.body_scale, o.precision, &cl.cfg, cl.x.data(), cl.y.data(),
                                 cl.z.data(), cl.bnd.data(), cl.wall.data(), cl.mom.data(),
                                 cl.g.data(), cl.bid.data(), cl.bodies.data(), &nb, cl.geo);
        if (rc) fail(nullptr, "alh_build_body_case", rc);
        cl.bodies.resize(nb);
    }
    else
    {
        alh_lattice_size(&cs, &N);
        cl.resize(N);
        rc = alh_build_lattice(&cs, o.precision, &cl.cfg, cl.x.data(), cl.y.data(), cl.z.data(),
                               cl.bnd.data(), cl.wall.data(), cl.mom.data(), cl.g.data());
        if (rc) fail(nullptr, "alh_build_lattice", rc);
        const double dx = 1e-6 / (cs.nx - 1);
        cl.geo[0] = dx; cl.geo[3] = 0.5 * dx;
    }
    cl.cfg.precision = o.precision;
    cl.cfg.collision = o.collision;
    cl.cfg.limiter = o.limiter;
    cl.cfg.mls_order = o.mls_order;
    make_dirs(out_dir);

    long long step = 0;
    double t = 0.0;
    const std::string ck_path = out_dir + "/restart.ckpt";
    if (const char *rf = std::getenv("ALEBGK_RESTART"))
        if (rf[0])
        {
            if ((rc = load_restart(rf, cl, &step, &t)))
            {
                std::printf("[ckpt] restore FAILED -- refusing to start from a half-known state\n");
                return 1;
            }
            cl.cfg.n_particles = cl.N;
            if (cl.N != cs.nx * cs.ny * (cs.dim == 3 ? cs.nz : 1)) cl.cfg.nx = cl.cfg.ny = cl.cfg.nz = 0;
            std::printf("[ckpt] restored step %lld, t = %.6e from %s\n", step, t, rf);
        }
    char msg[256];
    if ((rc = ab_check_config(&cl.cfg, msg, sizeof msg)))
    {
        std::printf("Configuration refused: %s\n", msg);
        return 1;
    }
    ab_ctx *c = nullptr;
    if ((rc = ab_create(&cl.cfg, &c))) fail(nullptr, "ab_create (is a GPU present?)", rc);
    std::vector<double> cols(12 * (size_t)cl.N);
    double *col[12];
    for (int k = 0; k < 12; ++k) col[k] = &cols[(size_t)k * cl.N];
    for (int i = 0; i < cl.N; ++i)
    {
        for (int k = 0; k < 7; ++k) col[k][i] = cl.wall[7 * (size_t)i + k];
        for (int k = 0; k < 5; ++k) col[7 + k][i] = cl.mom[5 * (size_t)i + k];
    }
    std::vector<double> mv(3 * (size_t)cl.N);
    for (int i = 0; i < cl.N; ++i)
        for (int k = 0; k < 3; ++k) mv[(size_t)k * cl.N + i] = cl.mesh[3 * (size_t)i + k];
    ab_cloud_view v{cl.x.data(), cl.y.data(), cl.z.data(), cl.bnd.data(), col[0], col[1], col[2],
                    col[3], col[4], col[5], col[6], &mv[0], &mv[cl.N], &mv[2 * (size_t)cl.N],
                    col[7], col[8], col[9], col[10], col[11], cl.g.data(), nullptr, nullptr,
                    cl.bid.data()};
    if ((rc = ab_upload(c, &v))) fail(c, "ab_upload", rc);
    std::printf("Particles: %d\n\n", cl.N);
    std::printf("=== ALE-BGK Solver: N=%d  dim=%d  ranks=1 ===\n", cl.N, cs.dim);

    // foot-point check (Solver.cpp:178-221)
    double dt = cs.dt;
    {
        const double vmax = std::max(std::fabs(cs.vmax), std::fabs(cs.vmin));
        const double dxmin = std::fabs(cl.geo[0]);
        const double cfl = vmax * dt / dxmin;
        std::printf("[FOOT-POINT CHECK] dt= %.3g, v_max*dt/dx = %.3f (max allowed 1.5)\n", dt, cfl);
        if (cfl > 1.5)
        {
            const double safe = 0.75 * dxmin / vmax;
            std::printf("[FOOT-POINT WARNING] dt= %.3g, v_max*dt/dx = %.3f > 1.5 (outside MLS "
                        "stencil).\n              Clamping dt from %.3e to %.3e (CFL 0.75).\n",
                        dt, cfl, dt, safe);
            dt = safe;
        }
        else
            std::printf("[FOOT-POINT OK] v_max*dt/dx = %.3f (within MLS stencil)\n", cfl);
    }
    alh_guard guard;
    alh_guard_init(&guard, cs.Tw, cs.rho0);
    Snapshot snap;
    alh_diag d{};
    if (step == 0)
    {
        // the initial state is written from the host cloud, as the
        // reference does before its first step (Solver.cpp:293): p = rho R T
        snapshot(c, cl.N, snap);
        for (int i = 0; i < cl.N; ++i) snap.p[i] = snap.rho[i] * cs.R * snap.T[i];
        write_outputs(snap, cl.N, cs.nv, out_dir, 0, 0.0, d);
    }
    int ck_every = 0;
    if (const char *e = std::getenv("ALEBGK_CHECKPOINT_EVERY"))
        if (e[0]) ck_every = std::atoi(e);
    const std::string ts_path = out_dir + "/timeseries.csv", body_path = out_dir + "/bodies.csv";
    bool ts_first = true, body_first = true, diverged = false;
    const long est_total = (long)std::ceil(cs.tfinal / cs.dt);
    auto loop_start = std::chrono::steady_clock::now();
    double est_t1 = 0.0;
    while (t < cs.tfinal)
    {
        dt = std::min(dt, cs.tfinal - t);
        if ((rc = ab_step(c, dt, nullptr))) fail(c, "ab_step", rc);
        if (moving && !cl.bodies.empty())
        {
            int n_new = 0, counts[3];
            rc = ab_body_step(c, dt, (int)cl.bodies.size(), cl.bodies.data(), cl.geo[0],
                              cl.geo[3], std::vector<double>{cs.dom[0], cs.dom[1], cs.dom[3],
                                                             cs.dom[2], cs.dom[4], cs.dom[5]}
                                             .data(),
                              nullptr, &n_new, counts);
            if (rc) fail(c, "ab_body_step", rc);
            cl.N = n_new;
            if ((counts[0] || counts[2]) && step % cs.save_every == 0)
                std::printf("  [bodies] removed %d, added %d, N=%d\n", counts[0], counts[2], n_new);
            if (step % cs.save_every == 0)
            {
                alh_write_bodies(body_path.c_str(), body_first, t + dt, cs.dim,
                                 (int)cl.bodies.size(), cl.bodies.data());
                body_first = false;
            }
        }
        t += dt;
        step++;
        if (o.estimate >= 0 && step <= 2)
        {
            ab_synchronize(c);
            const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() -
                                                            loop_start).count();
            if (step == 1) est_t1 = el;
            else
            {
                const double per = el - est_t1;
                std::printf("[ESTIMATE] step 2 took %.3f s -> projected full run for time = %.3g"
                            "(%ld steps): %.3g h\n",
                            per, cs.tfinal, est_total, per * est_total / 3600.0);
                std::fflush(stdout);
                if (o.estimate == 0) break;
            }
        }
        if (step % cs.save_every == 0)
        {
            snapshot(c, cl.N, snap);
            write_outputs(snap, cl.N, cs.nv, out_dir, (int)step, t, d);
            alh_format_step((int)step, t, dt, &d, msg, sizeof msg);
            std::printf("%s\n", msg);
            std::fflush(stdout);
            char why[256];
            if (alh_guard_check(&guard, (int)step, t, &d, why, sizeof why))
            {
                std::printf("\n*** STOPPING: the solution has left physical bounds.\n"
                            "***   step %lld, t = %.4e s\n***   %s\n"
                            "***   rho = %.6g (rho0 = %.6g), T = %.6g K (Tw = %.6g K), KE = %.6g\n***\n",
                            step, t, why, d.rho_avg, cs.rho0, d.T_avg, cs.Tw, d.KE);
                diverged = true;
            }
            alh_write_timeseries(ts_path.c_str(), ts_first, t, &d);
            ts_first = false;
        }
        if (ck_every > 0 && step % ck_every == 0)
        {
            if ((rc = save_restart(c, cl, ck_path, step, t, cs.R)))
                std::printf("[ckpt] write of %s failed (%d)\n", ck_path.c_str(), rc);
            else
                std::printf("[ckpt] wrote %s at step %lld\n", ck_path.c_str(), step);
        }
        if (diverged) break;
    }
    // the end of the GPU loop (Solver.cpp:570-577): the moments and the step
    // line; the files were written on the saveEvery steps
    snapshot(c, cl.N, snap);
    alh_diagnostics(cl.N, snap.bnd.data(), snap.rho.data(), snap.ux.data(), snap.uy.data(),
                    snap.uz.data(), snap.T.data(), snap.p.data(), snap.mood.data(), &d);
    alh_format_step((int)step, t, dt, &d, msg, sizeof msg);
    std::printf("%s\n", msg);
    const double elapsed =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    std::printf("\nDone. Steps=%lld  t=%.4e  Wall=%.2f s\n", step, t, elapsed);
    ab_destroy(c);
    return diverged ? 2 : 0;
}
