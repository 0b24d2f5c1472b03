// ref_harness.cpp -- drives the reference CPU functions in place, to write
// golden fixtures and to time the reference CPU step.
//
// This file is test infrastructure. It is compiled by oracle/Makefile against
// the reference sources where they lie (REF=/root/reference by default) and
// the binary goes to oracle/_ref/. Nothing in the product loads it.
//
// Modes
//   ref_harness case3d  <outdir> <Nx> <Nv>         cavity3d fixtures
//   ref_harness case2d  <outdir> <Nx> <Nv>         cavity2d fixtures
//   ref_harness cube    <outdir> <Nx> <Nv> <scale> cavity3d-cube body loads
//   ref_harness cpustep <Nx> <Nv> <order> <nsteps> time the reference CPU step
//   ref_harness ckpt    <outdir> <case> <Nx> <Nv>   a restart file + its fields
//   ref_harness mood    <outdir> <case> <Nx> <Nv> <order> <nsteps> <a> <b>
//                                                  CPU steps with the MOOD limiter
//   ref_harness ckptread <file> <outdir> <case> <Nx> <Nv>  restore a file, dump fields
//   ref_harness pos2d   <outdir> <Nx> <Nv> <a> <T> <cfl>   2-D transport where positivity fires
//   ref_harness bodycase <outdir> <case> <Nx> <Nv> <scale>   a body case's cloud and body
//   ref_harness writers <outdir> <case> <Nx> <Nv>   the VTK / CSV writers on a cloud
//   ref_harness bodysteps <outdir> <case> <Nx> <Nv> <scale> <nsteps> <a> <order>
//                                                  CPU steps with a free body
//   ref_harness cloud   <outdir> <case> <Nx> <Nv> <scale> <shift> <jitter>
//                                                  cloud management after a body move
//
// Functions called (reference file:line):
//   configureCase            TestCases/TestCases.cpp:113
//   makeBodies               TestCases/TestCases.cpp:251
//   generateParticles        Initialization/ParticleGeneration.cpp:30
//   embedRigidBodies         Initialization/BodyEmbedding.cpp:18
//   updateVoxelsAndNeighbours Initialization/NeighbourSearch.cpp:17
//   initWallData             BoundaryConditions/WallState.cpp:16
//   applyLidProfile          TestCases/TestCases.cpp:230
//   applyIC                  Initialization/InitialConditions.cpp:18
//   moveRigidBodies          RigidBody/BodyDynamics.cpp:143
//   mlsTransportCompute      MLSTransport/MLSTransport.cpp:88
//   computeMoments           MomentCalculations/Moments.cpp:17
//   bgkCollision             MomentCalculations/Collision.cpp:70
//   applyBoundaryBC          BoundaryConditions/DiffuseWall.cpp:19
//   sumMass / scaleG         MomentCalculations/Conservation.cpp:48,92
//   checkTransportPAD, mlsTransportCommit  MLSTransport/MLSTransport.cpp:324, 384
//   moodCheck                MomentCalculations/Conservation.cpp:16
//   computeBodyLoads (static) RigidBody/BodyDynamics.cpp:17 (TU included below)
//   ckpt::save / ckpt::restore Core/Checkpoint.hpp:78, 161
//   manageRemove, mergeCloseParticles, manageRefill  RigidBody/CloudManagement.cpp:17, 80, 121
// =============================================================================
#include "Config.hpp"
#include "Types.hpp"
#include "MLSBasis.hpp"
#include "Geometry.hpp"
#include "Initialization.hpp"
#include "BoundaryConditions.hpp"
#include "MLSTransport.hpp"
#include "Moments.hpp"
#include "TestCases.hpp"
#include "Checkpoint.hpp"
#include "Output.hpp"
#include "RigidBody.hpp"

#include <chrono>
#include <random>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

// computeBodyLoads is file-static in the reference; compile that translation
// unit here (in place, unmodified) so the harness can call it.
#include "BodyDynamics.cpp"

namespace {

// ---- minimal .npy writer (little-endian, C order) -------------------------
void writeNpy(const std::string &path, const char *descr, const void *data,
              size_t elemSize, const std::vector<size_t> &shape)
{
    std::string dict = "{'descr': '";
    dict += descr;
    dict += "', 'fortran_order': False, 'shape': (";
    size_t count = 1;
    for (size_t k = 0; k < shape.size(); ++k)
    {
        dict += std::to_string(shape[k]);
        dict += (shape.size() == 1) ? "," : (k + 1 < shape.size() ? ", " : "");
        count *= shape[k];
    }
    dict += "), }";
    size_t hdr = 10 + dict.size() + 1;
    size_t pad = (64 - hdr % 64) % 64;
    dict.append(pad, ' ');
    dict += '\n';
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) { std::fprintf(stderr, "cannot write %s\n", path.c_str()); std::exit(2); }
    const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
    std::fwrite(magic, 1, 8, f);
    uint16_t hl = static_cast<uint16_t>(dict.size());
    std::fwrite(&hl, 2, 1, f);
    std::fwrite(dict.data(), 1, dict.size(), f);
    std::fwrite(data, elemSize, count, f);
    std::fclose(f);
}
void npyF8(const std::string &p, const std::vector<double> &v,
           std::vector<size_t> shape) { writeNpy(p, "<f8", v.data(), 8, shape); }
void npyI4(const std::string &p, const std::vector<int32_t> &v,
           std::vector<size_t> shape) { writeNpy(p, "<i4", v.data(), 4, shape); }
void npyU1(const std::string &p, const std::vector<uint8_t> &v,
           std::vector<size_t> shape) { writeNpy(p, "|u1", v.data(), 1, shape); }

struct Setup
{
    TestCase tc;
    SimParameters sp;
    DomainBoundary dom;
    BoundaryConditions BC;
    GasConstants gc;
    CalcParameters cp{};
    std::vector<std::unique_ptr<RigidBody>> bodies;
    std::vector<Particle> P;
};

// The setup order of main() (TestCases/main.cpp:193-345).
void build(Setup &S, TestCase tc, int Nx, int Nv, double bodyScale)
{
    S.tc = tc;
    configureCase(tc, BodyMotion::Free, S.sp, S.dom, S.BC, S.gc);
    S.sp.problem = ProblemID::DrivenCavity;
    const CaseInfo &info = caseInfo(tc);
    S.sp.Nx = S.sp.Ny = Nx;
    if (info.dim == 3) S.sp.Nz = Nx;
    S.sp.Nv = S.sp.NvMin = S.sp.NvMax = Nv;
    S.bodies = makeBodies(tc, BodyMotion::Free, S.BC.rho, bodyScale);
    generateParticles(S.P, S.sp, S.dom, S.cp);
    embedRigidBodies(S.P, S.bodies, S.sp, S.cp);
    updateVoxelsAndNeighbours(S.P, S.cp, S.dom);
    initWallData(S.P, S.dom, S.BC);
    applyLidProfile(tc, S.P, S.dom);
    double Lx = S.dom.xright - S.dom.xleft, Ly = S.dom.ytop - S.dom.ybottom;
    double Lz = (info.dim == 3) ? (S.dom.zback - S.dom.zfront) : 1.0;
    applyIC(S.P, S.BC.rho, S.gc.R, Lx, Ly, Lz);
}

// A smooth, positive, non-equilibrium distribution: a local Maxwellian whose
// density, velocity and temperature vary across the box. Deterministic.
void perturb(Setup &S)
{
    const double L = 1e-6, R = S.gc.R, twoPi = 2.0 * M_PI;
    for (auto &p : S.P)
    {
        double sx = p.x / L, sy = p.y / L, sz = p.z / L;
        double rho = S.BC.rho * (1.0 + 0.2 * std::sin(twoPi * sx) * std::cos(M_PI * sy));
        double ux = 60.0 * std::sin(M_PI * sy);
        double uy = -40.0 * std::sin(M_PI * sx);
        double uz = 30.0 * std::cos(M_PI * sz);
        double T = 270.0 * (1.0 + 0.1 * std::sin(M_PI * sz + 0.3));
        int Nv = p.Nv_local;
        for (int kv = 0; kv < (p.dim == 3 ? Nv : 1); ++kv)
            for (int jv = 0; jv < Nv; ++jv)
                for (int iv = 0; iv < Nv; ++iv)
                {
                    double vx = p.vlo + iv * p.dv_local;
                    double vy = p.vlo + jv * p.dv_local;
                    double vz = p.vlo + kv * p.dv_local;
                    if (p.dim == 3)
                        p.g[iv + Nv * (jv + Nv * kv)] =
                            maxwellian3D(vx, vy, vz, ux, uy, uz, T, rho, R);
                    else
                    {
                        double g1 = maxwellianG1(vx, vy, ux, uy, T, rho, R);
                        p.g[iv + Nv * jv] = g1;
                        p.g[iv + Nv * jv + Nv * Nv] = 0.5 * R * T * g1 * (1.0 + 0.05 * sx);
                    }
                }
    }
}

std::vector<double> gatherG(const std::vector<Particle> &P, bool gt)
{
    size_t gs = P[0].g.size();
    std::vector<double> out(P.size() * gs);
    for (size_t i = 0; i < P.size(); ++i)
        std::memcpy(&out[i * gs], (gt ? P[i].gt : P[i].g).data(), gs * 8);
    return out;
}
void scatterG(std::vector<Particle> &P, const std::vector<double> &in)
{
    size_t gs = P[0].g.size();
    for (size_t i = 0; i < P.size(); ++i)
        std::memcpy(P[i].g.data(), &in[i * gs], gs * 8);
}
std::vector<double> moments(const std::vector<Particle> &P)
{
    std::vector<double> m(P.size() * 5);
    for (size_t i = 0; i < P.size(); ++i)
    {
        m[5 * i + 0] = P[i].rho; m[5 * i + 1] = P[i].ux; m[5 * i + 2] = P[i].uy;
        m[5 * i + 3] = P[i].uz;  m[5 * i + 4] = P[i].T;
    }
    return m;
}
void setMoments(std::vector<Particle> &P, const std::vector<double> &m)
{
    for (size_t i = 0; i < P.size(); ++i)
    {
        P[i].rho = m[5 * i + 0]; P[i].ux = m[5 * i + 1]; P[i].uy = m[5 * i + 2];
        P[i].uz = m[5 * i + 3];  P[i].T = m[5 * i + 4];
    }
}

void writeMeta(const std::string &dir, const Setup &S, double dtT)
{
    const CalcParameters &cp = S.cp;
    fs::create_directories(dir);
    FILE *f = std::fopen((dir + "/meta.json").c_str(), "w");
    if (!f) { std::fprintf(stderr, "cannot write %s/meta.json\n", dir.c_str()); std::exit(2); }
    std::fprintf(f, "{\n");
    std::fprintf(f, "  \"case\": \"%s\",\n", caseInfo(S.tc).name);
    std::fprintf(f, "  \"dim\": %d, \"N\": %zu, \"Nx\": %d, \"Ny\": %d, \"Nz\": %d, \"Nv\": %d,\n",
                 cp.dim, S.P.size(), S.sp.Nx, S.sp.Ny, S.sp.Nz, S.sp.Nv);
    std::fprintf(f, "  \"gStride\": %zu,\n", S.P[0].g.size());
    std::fprintf(f, "  \"vmin\": %.17g, \"vmax\": %.17g, \"dv\": %.17g,\n",
                 S.sp.VMin, S.sp.VMax, cp.dv);
    std::fprintf(f, "  \"dx\": %.17g, \"dy\": %.17g, \"dz\": %.17g,\n", cp.dx, cp.dy, cp.dz);
    std::fprintf(f, "  \"radius\": %.17g, \"minDist\": %.17g,\n", cp.radius, cp.minDist);
    std::fprintf(f, "  \"hBox\": %.17g, \"vBox\": %.17g, \"wBox\": %.17g,\n", cp.hBox, cp.vBox, cp.wBox);
    std::fprintf(f, "  \"nbhBox\": %d, \"nbvBox\": %d, \"nbwBox\": %d, \"nvox\": %d,\n",
                 cp.nbhBox, cp.nbvBox, cp.nbwBox, cp.nvox);
    std::fprintf(f, "  \"xleft\": %.17g, \"xright\": %.17g, \"ybottom\": %.17g, \"ytop\": %.17g, "
                 "\"zfront\": %.17g, \"zback\": %.17g,\n",
                 S.dom.xleft, S.dom.xright, S.dom.ybottom, S.dom.ytop, S.dom.zfront, S.dom.zback);
    std::fprintf(f, "  \"R\": %.17g, \"tau\": %.17g, \"dt_case\": %.17g, \"dt_transport\": %.17g,\n",
                 S.gc.R, S.sp.tao, S.sp.dt, dtT);
    std::fprintf(f, "  \"rho0\": %.17g, \"Tw\": %.17g, \"mass_in\": 0\n", S.BC.rho, S.BC.Tw);
    std::fprintf(f, "}\n");
    std::fclose(f);
}

void writeCloud(const std::string &dir, const Setup &S)
{
    const auto &P = S.P;
    size_t N = P.size();
    std::vector<double> x(N), y(N), z(N), wall(N * 7);
    std::vector<uint8_t> bnd(N);
    std::vector<int32_t> vox(N), body(N), nn(N), flat;
    for (size_t i = 0; i < N; ++i)
    {
        x[i] = P[i].x; y[i] = P[i].y; z[i] = P[i].z;
        bnd[i] = P[i].boundary ? 1 : 0; vox[i] = P[i].voxel; body[i] = P[i].bodyId;
        double w[7] = {P[i].wnx, P[i].wny, P[i].wnz, P[i].wUx, P[i].wUy, P[i].wUz, P[i].wT};
        std::memcpy(&wall[7 * i], w, sizeof w);
        nn[i] = static_cast<int32_t>(P[i].neighindex.size());
        for (int j : P[i].neighindex) flat.push_back(j);
    }
    npyF8(dir + "/x.npy", x, {N});
    npyF8(dir + "/y.npy", y, {N});
    npyF8(dir + "/z.npy", z, {N});
    npyU1(dir + "/boundary.npy", bnd, {N});
    npyI4(dir + "/voxel.npy", vox, {N});
    npyI4(dir + "/bodyid.npy", body, {N});
    npyF8(dir + "/wall.npy", wall, {N, 7});
    npyI4(dir + "/nneigh.npy", nn, {N});
    npyI4(dir + "/neigh_flat.npy", flat, {flat.size()});
}

// Particles whose CPU positivity repair fired during mlsTransportCompute: the
// repair zeroes the negatives, so an exact zero in gt marks them
// (MLSTransport/MLSTransport.cpp:189-200, :283-309).
std::vector<uint8_t> positivityMask(const std::vector<Particle> &P)
{
    std::vector<uint8_t> m(P.size(), 0);
    for (size_t i = 0; i < P.size(); ++i)
        for (double v : P[i].gt)
            if (v == 0.0) { m[i] = 1; break; }
    return m;
}

int modeCase(const std::string &dir, TestCase tc, int Nx, int Nv)
{
    Setup S;
    build(S, tc, Nx, Nv, 1.0);
    size_t N = S.P.size(), gs = S.P[0].g.size();
    double dxmin = std::fabs(S.cp.dx);
    if (std::fabs(S.cp.dy) < dxmin) dxmin = std::fabs(S.cp.dy);
    if (S.cp.dim == 3 && std::fabs(S.cp.dz) < dxmin) dxmin = std::fabs(S.cp.dz);
    double dtT = 0.5 * dxmin / S.sp.VMax;        // foot-point CFL 0.5
    writeMeta(dir, S, dtT);
    writeCloud(dir, S);
    npyF8(dir + "/g_ic.npy", gatherG(S.P, false), {N, gs});

    perturb(S);
    std::vector<double> gin = gatherG(S.P, false);
    npyF8(dir + "/g_in.npy", gin, {N, gs});
    std::vector<double> mass{sumMass(S.P)};
    npyF8(dir + "/mass_in.npy", mass, {1});

    for (int order = 1; order <= 2; ++order)
    {
        scatterG(S.P, gin);
        mlsTransportCompute(S.P, dtT, S.cp, S.dom, S.BC, order);
        std::string sfx = "_o" + std::to_string(order);
        npyF8(dir + "/gt" + sfx + ".npy", gatherG(S.P, true), {N, gs});
        npyU1(dir + "/posfired" + sfx + ".npy", positivityMask(S.P), {N});
    }

    // Moments of the perturbed field (interior particles; the boundary keeps
    // the state applyIC gave it).
    scatterG(S.P, gin);
    std::vector<double> m0 = moments(S.P);
    computeMoments(S.P, S.gc.R);
    std::vector<double> m1 = moments(S.P);
    npyF8(dir + "/mom_in.npy", m1, {N, 5});

    // BGK collision with those moments.
    bgkCollision(S.P, S.sp.dt, S.sp.tao, S.gc.R, false, S.sp.d);
    npyF8(dir + "/g_coll.npy", gatherG(S.P, false), {N, gs});

    // Diffuse wall on the perturbed field.
    scatterG(S.P, gin);
    setMoments(S.P, m0);
    applyBoundaryBC(S.P, S.gc.R, S.dom, S.BC);
    npyF8(dir + "/g_wall.npy", gatherG(S.P, false), {N, gs});
    npyF8(dir + "/mom_wall.npy", moments(S.P), {N, 5});
    std::printf("wrote %s: N=%zu gStride=%zu\n", dir.c_str(), N, gs);
    return 0;
}

int modeCube(const std::string &dir, int Nx, int Nv, double scale)
{
    Setup S;
    build(S, TestCase::Cavity3D_Cube, Nx, Nv, scale);
    size_t N = S.P.size(), gs = S.P[0].g.size();
    writeMeta(dir, S, 0.0);
    writeCloud(dir, S);
    perturb(S);
    // Body loads are a pure function of the surface particles' g and wall
    // state; store only those particles' g.
    std::vector<int32_t> surf;
    for (size_t i = 0; i < N; ++i)
        if (S.P[i].boundary && S.P[i].bodyId >= 0) surf.push_back(static_cast<int32_t>(i));
    std::vector<double> gs_surf(surf.size() * gs);
    for (size_t k = 0; k < surf.size(); ++k)
        std::memcpy(&gs_surf[k * gs], S.P[surf[k]].g.data(), gs * 8);
    npyI4(dir + "/surf_idx.npy", surf, {surf.size()});
    npyF8(dir + "/g_surf.npy", gs_surf, {surf.size(), gs});
    // Give the surface a non-zero wall velocity so the load's dependence on
    // wU is exercised: a rigid rotation plus translation.
    std::vector<double> wall(surf.size() * 3);
    for (size_t k = 0; k < surf.size(); ++k)
    {
        Particle &p = S.P[surf[k]];
        const RigidBody &b = *S.bodies[p.bodyId];
        double rx = p.x - b.cx, ry = p.y - b.cy, rz = p.z - b.cz;
        double vx = 0.7, vy = -0.4, vz = 0.2, ox = 1e6, oy = -2e6, oz = 3e6;
        p.wUx = vx + oy * rz - oz * ry;
        p.wUy = vy + oz * rx - ox * rz;
        p.wUz = vz + ox * ry - oy * rx;
        wall[3 * k] = p.wUx; wall[3 * k + 1] = p.wUy; wall[3 * k + 2] = p.wUz;
    }
    npyF8(dir + "/surf_wU.npy", wall, {surf.size(), 3});
    size_t nb = S.bodies.size();
    std::vector<double> Fx(nb), Fy(nb), Fz(nb), Tx(nb), Ty(nb), Tz(nb);
    computeBodyLoads(S.P, S.bodies, Fx, Fy, Fz, Tx, Ty, Tz);
    std::vector<double> loads{Fx[0], Fy[0], Fz[0], Tx[0], Ty[0], Tz[0]};
    npyF8(dir + "/loads.npy", loads, {6});
    const RigidBody &b = *S.bodies[0];
    std::vector<double> bstate{b.cx, b.cy, b.cz, b.getSurfaceArea(), b.mass,
                               b.getMomentOfInertia()};
    npyF8(dir + "/body.npy", bstate, {6});
    std::printf("wrote %s: N=%zu surface=%zu\n", dir.c_str(), N, surf.size());
    return 0;
}

// The CPU time loop body of runSolver (Solver/Solver.cpp:632-705), run for
// nsteps fixed steps; prints one JSON line with the per-step wall times.
int modeCpuStep(int Nx, int Nv, int order, int nsteps)
{
    Setup S;
    build(S, TestCase::Cavity3D, Nx, Nv, 1.0);
    S.sp.mlsOrder = order;
    auto &P = S.P;
    const SimParameters &sp = S.sp;
    double dt = sp.dt, R = S.gc.R;
    std::vector<double> sec;
    for (int s = 0; s < nsteps; ++s)
    {
        auto t0 = std::chrono::steady_clock::now();
        double massOld = sumMass(P);
        mlsTransportCompute(P, dt, S.cp, S.dom, S.BC, sp.mlsOrder);
        if (sp.limiter == LimiterMode::MOOD)
        {
            int nf = checkTransportPAD(P);
            if (nf > 0) mlsTransportCompute(P, dt, S.cp, S.dom, S.BC, 1, true);
        }
        mlsTransportCommit(P);
        double massNew = sumMass(P);
        if (massNew > 1.0e-30) scaleG(P, massOld / massNew);
        applyGravity(P, sp.gravityY, dt);
        applyBoundaryBC(P, R, S.dom, S.BC);
        computeMoments(P, R);
        bgkCollision(P, dt, sp.tao, R, sp.tauVariable, sp.d);
        computeMoments(P, R);
        if (sp.limiter == LimiterMode::MOOD && moodCheck(P))
        {
            bgkCollision(P, dt, sp.tao, R, sp.tauVariable, sp.d);
            computeMoments(P, R);
        }
        sec.push_back(std::chrono::duration<double>(
            std::chrono::steady_clock::now() - t0).count());
        std::fprintf(stderr, "[cpustep] step %d %.3f s\n", s + 1, sec.back());
    }
    int threads = 1;
#ifdef _OPENMP
    threads = omp_get_max_threads();
#endif
    double rho = 0, T = 0; int n = 0;
    for (auto &p : P) if (!p.boundary) { rho += p.rho; T += p.T; ++n; }
    std::printf("{\"N\": %zu, \"Nv\": %d, \"order\": %d, \"threads\": %d, \"step_seconds\": [",
                P.size(), Nv, order, threads);
    for (size_t k = 0; k < sec.size(); ++k)
        std::printf("%s%.6f", k ? ", " : "", sec[k]);
    std::printf("], \"rho_avg\": %.10f, \"T_avg\": %.10f}\n", rho / n, T / n);
    return 0;
}

}  // namespace

// Every field a restart record carries, as arrays (for the format tests).
void writeRecordFields(const std::string &dir, const Setup &S)
{
    const auto &P = S.P;
    size_t N = P.size(), gs = P[0].g.size();
    writeCloud(dir, S);
    std::vector<double> m(N * 6), mesh(N * 3), grid(N * 3);
    std::vector<int32_t> ints(N * 2);
    std::vector<uint8_t> flags(N * 2);
    for (size_t i = 0; i < N; ++i)
    {
        const Particle &q = P[i];
        double mm[6] = {q.rho, q.ux, q.uy, q.uz, q.T, q.p};
        std::memcpy(&m[6 * i], mm, sizeof mm);
        mesh[3 * i] = q.meshVx; mesh[3 * i + 1] = q.meshVy; mesh[3 * i + 2] = q.meshVz;
        grid[3 * i] = q.vlo; grid[3 * i + 1] = q.vhi; grid[3 * i + 2] = q.dv_local;
        ints[2 * i] = q.Nv_local; ints[2 * i + 1] = q.dim;
        flags[2 * i] = q.validg ? 1 : 0; flags[2 * i + 1] = q.moodFlag ? 1 : 0;
    }
    npyF8(dir + "/rec_mom.npy", m, {N, 6});
    npyF8(dir + "/rec_mesh.npy", mesh, {N, 3});
    npyF8(dir + "/rec_grid.npy", grid, {N, 3});
    npyI4(dir + "/rec_ints.npy", ints, {N, 2});
    npyU1(dir + "/rec_flags.npy", flags, {N, 2});
    npyF8(dir + "/g.npy", gatherG(P, false), {N, gs});
    std::vector<double> b;
    for (const auto &rb : S.bodies)
    {
        double v[19] = {rb->cx, rb->cy, rb->cz, rb->velx, rb->vely, rb->velz, rb->angularVel,
                        rb->omx, rb->omy, rb->omz, rb->forcex, rb->forcey, rb->forcez,
                        rb->torque, rb->tqx, rb->tqy, rb->tqz, rb->ox, rb->oy};
        b.insert(b.end(), v, v + 19);
        const double *Rm = rb->orientation();
        double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        b.insert(b.end(), Rm ? Rm : I, (Rm ? Rm : I) + 9);
        b.push_back(rb->dim);
        b.push_back(rb->rotates ? 1.0 : 0.0);
    }
    npyF8(dir + "/bodies.npy", b, {S.bodies.size(), 30});
}

TestCase caseOf(const std::string &name)
{
    if (name == "cube") return TestCase::Cavity3D_Cube;
    if (name == "sphere") return TestCase::Cavity3D_Sphere;
    if (name == "circle") return TestCase::Cavity2D_Circle;
    if (name == "square") return TestCase::Cavity2D_Square;
    if (name == "cavity2d") return TestCase::Cavity2D;
    return TestCase::Cavity3D;
}

// A restart file of a perturbed state with moments, mesh velocity and a
// moving body, written by the reference's ckpt::save.
int modeCkpt(const std::string &dir, const std::string &cname, int Nx, int Nv)
{
    Setup S;
    build(S, caseOf(cname), Nx, Nv, 1.0);
    perturb(S);
    computeMoments(S.P, S.gc.R);
    for (auto &q : S.P)
    {
        q.p = q.rho * S.gc.R * q.T;
        q.meshVx = 1e-3 * q.x / 1e-6; q.meshVy = -2e-3; q.meshVz = 0.5e-3 * q.z / 1e-6;
        q.moodFlag = (q.voxel % 3) == 0;
    }
    for (auto &rb : S.bodies)
    {
        rb->velx = 0.25; rb->vely = -0.5; rb->velz = 0.125;
        rb->omx = 1e5; rb->omy = -2e5; rb->omz = 3e5; rb->angularVel = 4e5;
        rb->forcex = 1e-9; rb->tqz = 2e-15;
    }
    fs::create_directories(dir);
    writeMeta(dir, S, 0.0);
    writeRecordFields(dir, S);
    if (!ckpt::save(dir + "/restart.ckpt", 7, 1.5e-10, S.P, S.bodies)) return 3;
    return 0;
}

int modeCkptRead(const std::string &file, const std::string &dir, const std::string &cname,
                 int Nx, int Nv)
{
    Setup S;
    build(S, caseOf(cname), Nx, Nv, 1.0);
    long step = 0;
    double t = 0.0;
    if (!ckpt::restore(file, &step, &t, S.P, S.bodies)) return 3;
    fs::create_directories(dir);
    writeRecordFields(dir, S);
    std::vector<double> st{(double)step, t};
    npyF8(dir + "/step_t.npy", st, {2});
    return 0;
}

// The CPU time loop of runSolver with its default limiter, MOOD
// (Solver/Solver.cpp:641-693), for nsteps fixed steps from a start that
// makes the limiters act: two opposing streams, the Maxwellian of the case's
// density and wall temperature moving at u_x = +a left of the domain's
// middle and -a right of it (the density piles up where they meet, a new
// extremum for the NAD check), with every value scaled by (1 + b u), u
// uniform in [-1, 1] (mt19937_64, seed 20261202); the moments are
// recomputed. Writes the start (g0, mom0), the end (g1, mom1), the particles
// the transport check flagged at each step (flags, nsteps x N) and whether
// the post-collision check failed (pad).
int modeMood(const std::string &dir, const std::string &cname, int Nx, int Nv, int order,
             int nsteps, double a, double b)
{
    Setup S;
    build(S, caseOf(cname), Nx, Nv, 1.0);
    S.sp.mlsOrder = order;
    auto &P = S.P;
    const SimParameters &sp = S.sp;
    const double dt = sp.dt, R = S.gc.R;
    std::mt19937_64 rng(20261202);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    const double xmid = 0.5 * (S.dom.xleft + S.dom.xright);
    for (auto &p : P)
    {
        const double ux = (p.x < xmid) ? a : -a, rho = S.BC.rho, T = S.BC.Tw;
        const int Nv = p.Nv_local, Nv2 = Nv * Nv;
        for (int kv = 0; kv < (p.dim == 3 ? Nv : 1); ++kv)
            for (int jv = 0; jv < Nv; ++jv)
                for (int iv = 0; iv < Nv; ++iv)
                {
                    const double vx = p.vlo + iv * p.dv_local, vy = p.vlo + jv * p.dv_local;
                    const double vz = p.vlo + kv * p.dv_local;
                    if (p.dim == 3)
                        p.g[iv + Nv * (jv + Nv * kv)] = maxwellian3D(vx, vy, vz, ux, 0.0, 0.0, T, rho, R);
                    else
                    {
                        const double g1 = maxwellianG1(vx, vy, ux, 0.0, T, rho, R);
                        p.g[iv + Nv * jv] = g1;
                        p.g[iv + Nv * jv + Nv2] = 0.5 * R * T * g1;
                    }
                }
        for (auto &v : p.g) v *= 1.0 + b * U(rng);
    }
    computeMoments(P, R);
    const size_t N = P.size();
    auto moments = [&]() {
        std::vector<double> m(N * 5);
        for (size_t i = 0; i < N; ++i)
        {
            double q[5] = {P[i].rho, P[i].ux, P[i].uy, P[i].uz, P[i].T};
            std::memcpy(&m[5 * i], q, sizeof q);
        }
        return m;
    };
    writeMeta(dir, S, dt);
    const size_t gs = P[0].g.size();
    npyF8(dir + "/g0.npy", gatherG(P, false), {N, gs});
    npyF8(dir + "/mom0.npy", moments(), {N, 5});
    std::vector<uint8_t> flags;
    std::vector<uint8_t> pad;
    for (int st = 0; st < nsteps; ++st)
    {
        double massOld = sumMass(P);
        mlsTransportCompute(P, dt, S.cp, S.dom, S.BC, sp.mlsOrder);
        int nf = checkTransportPAD(P);
        for (const auto &p : P) flags.push_back(p.moodFlag ? 1 : 0);
        if (nf > 0) mlsTransportCompute(P, dt, S.cp, S.dom, S.BC, 1, true);
        mlsTransportCommit(P);
        double massNew = sumMass(P);
        if (massNew > 1.0e-30) scaleG(P, massOld / massNew);
        applyGravity(P, sp.gravityY, dt);
        applyBoundaryBC(P, R, S.dom, S.BC);
        computeMoments(P, R);
        bgkCollision(P, dt, sp.tao, R, sp.tauVariable, sp.d);
        computeMoments(P, R);
        const bool fail = moodCheck(P);
        pad.push_back(fail ? 1 : 0);
        if (fail)
        {
            bgkCollision(P, dt, sp.tao, R, sp.tauVariable, sp.d);
            computeMoments(P, R);
        }
        std::printf("step %d: %d flagged, post-collision check %s\n", st + 1, nf,
                    fail ? "failed" : "passed");
    }
    npyF8(dir + "/g1.npy", gatherG(P, false), {N, gs});
    npyF8(dir + "/mom1.npy", moments(), {N, 5});
    npyU1(dir + "/flags.npy", flags, {(size_t)nsteps, N});
    npyU1(dir + "/pad.npy", pad, {(size_t)nsteps});
    std::vector<double> par{(double)order, (double)nsteps, a, b};
    npyF8(dir + "/params.npy", par, {4});
    std::printf("wrote %s: N=%zu\n", dir.c_str(), N);
    return 0;
}

// Cloud management after a body displacement: the moving-body block of the
// GPU step without the body dynamics (Solver/Solver.cpp:379-398) --
// manageRemove, updateVoxelsAndNeighbours, mergeCloseParticles (+ rebuild),
// manageRefill (+ rebuild). The interior particles are jittered first
// (uniform, +-jitter dx per coordinate, mt19937_64 seed 20261203) so that
// some pairs come closer than minDist; the body and its surface particles
// are then translated by (shift, 0.6 shift, 0.3 shift) dx (z in 3-D only).
// A cube body is also turned 0.3 rad about z (its Rm; the surface points
// keep their positions: the tests only use the inside test).
int modeCloud(const std::string &dir, const std::string &cname, int Nx, int Nv, double scale,
              double shift, double jitter)
{
    Setup S;
    build(S, caseOf(cname), Nx, Nv, scale);
    perturb(S);
    computeMoments(S.P, S.gc.R);
    auto &P = S.P;
    const bool d3 = S.cp.dim == 3;
    const double dx = std::fabs(S.cp.dx);
    std::mt19937_64 rng(20261203);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    for (auto &p : P)
        if (!p.boundary)
        {
            p.x += jitter * dx * U(rng);
            p.y += jitter * dx * U(rng);
            if (d3) p.z += jitter * dx * U(rng);
        }
    RigidBody &b = *S.bodies[0];
    const double sx = shift * dx, sy = 0.6 * shift * dx, sz = d3 ? 0.3 * shift * dx : 0.0;
    b.cx += sx; b.cy += sy; b.cz += sz;
    for (auto &p : P)
        if (p.boundary && p.bodyId >= 0) { p.x += sx; p.y += sy; p.z += sz; }
    // shape, centre, size, orientation, rotation (AB_BODY_* numbering)
    std::vector<double> body(16, 0.0);
    body[1] = b.cx; body[2] = b.cy; body[3] = b.cz; body[5] = b.ox; body[6] = b.oy;
    body[7] = body[11] = body[15] = 1.0;
    if (auto *q = dynamic_cast<Square *>(&b)) { body[0] = 1; body[4] = q->L; }
    else if (auto *q = dynamic_cast<Circle *>(&b)) { body[0] = 2; body[4] = q->R; }
    else if (auto *q = dynamic_cast<Sphere *>(&b)) { body[0] = 3; body[4] = q->R; }
    else if (auto *q = dynamic_cast<Cube *>(&b))
    {
        body[0] = 4; body[4] = q->L;
        const double a = 0.3, c = std::cos(a), s_ = std::sin(a);
        const double Rm[9] = {c, -s_, 0.0, s_, c, 0.0, 0.0, 0.0, 1.0};
        q->setOrientation(Rm);
        for (int k = 0; k < 9; ++k) body[7 + k] = Rm[k];
    }
    auto dump = [&](const std::string &tag) {
        const size_t N = P.size(), gs = P[0].g.size();
        std::vector<double> x(N), y(N), z(N);
        std::vector<uint8_t> bnd(N);
        std::vector<int32_t> bid(N);
        for (size_t i = 0; i < N; ++i)
        {
            x[i] = P[i].x; y[i] = P[i].y; z[i] = P[i].z;
            bnd[i] = P[i].boundary ? 1 : 0; bid[i] = P[i].bodyId;
        }
        npyF8(dir + "/x_" + tag + ".npy", x, {N});
        npyF8(dir + "/y_" + tag + ".npy", y, {N});
        npyF8(dir + "/z_" + tag + ".npy", z, {N});
        npyU1(dir + "/boundary_" + tag + ".npy", bnd, {N});
        npyI4(dir + "/bodyid_" + tag + ".npy", bid, {N});
        npyF8(dir + "/g_" + tag + ".npy", gatherG(P, false), {N, gs});
        npyF8(dir + "/mom_" + tag + ".npy", moments(P), {N, 5});
    };
    writeMeta(dir, S, 0.0);
    {
        const size_t N = P.size();
        std::vector<double> wall(N * 7);
        for (size_t i = 0; i < N; ++i)
        {
            const double w[7] = {P[i].wnx, P[i].wny, P[i].wnz, P[i].wUx, P[i].wUy, P[i].wUz, P[i].wT};
            std::memcpy(&wall[7 * i], w, sizeof w);
        }
        npyF8(dir + "/wall_before.npy", wall, {N, 7});
    }
    dump("before");
    npyF8(dir + "/body.npy", body, {16});
    const int removed = manageRemove(P, S.bodies, S.cp);
    updateVoxelsAndNeighbours(P, S.cp, S.dom);
    const int merged = mergeCloseParticles(P, S.cp);
    if (merged > 0) updateVoxelsAndNeighbours(P, S.cp, S.dom);
    const int added = manageRefill(P, S.bodies, S.sp, S.cp, S.dom, S.gc.R);
    if (added > 0) updateVoxelsAndNeighbours(P, S.cp, S.dom);
    dump("after");
    std::vector<int32_t> cnt{removed, merged, added};
    npyI4(dir + "/counts.npy", cnt, {3});
    std::vector<double> par{scale, shift, jitter};
    npyF8(dir + "/params.npy", par, {3});
    std::printf("wrote %s: removed %d merged %d added %d, N %zu\n", dir.c_str(), removed, merged,
                added, P.size());
    return 0;
}

// The 2-D transport with its positivity repair on a field that trips it:
// opposing streams, u_x = +a left of the middle and -a right of it, at
// temperature T, so the transported g1 of the particles where the streams
// meet overshoots below zero. Writes the cloud, g_in, and per MLS order the
// transported g (gt_o1, gt_o2: the repair already applied, as
// mlsTransportCompute does) and the particles it repaired (posfired_o*).
int modePos2d(const std::string &dir, int Nx, int Nv, double a, double T, double cfl)
{
    Setup S;
    build(S, TestCase::Cavity2D, Nx, Nv, 1.0);
    const double R = S.gc.R, xmid = 0.5 * (S.dom.xleft + S.dom.xright);
    for (auto &p : S.P)
    {
        const double ux = (p.x < xmid) ? a : -a;
        const int n = p.Nv_local, n2 = n * n;
        for (int jv = 0; jv < n; ++jv)
            for (int iv = 0; iv < n; ++iv)
            {
                const double g1 = maxwellianG1(p.vlo + iv * p.dv_local, p.vlo + jv * p.dv_local,
                                               ux, 0.0, T, S.BC.rho, R);
                p.g[iv + n * jv] = g1;
                p.g[iv + n * jv + n2] = 0.5 * R * T * g1;
            }
    }
    const double dtT = cfl * std::fabs(S.cp.dx) / S.sp.VMax;
    writeMeta(dir, S, dtT);
    writeCloud(dir, S);
    const size_t N = S.P.size(), gs = S.P[0].g.size();
    std::vector<double> gin = gatherG(S.P, false);
    npyF8(dir + "/g_in.npy", gin, {N, gs});
    int fired[3] = {0, 0, 0};
    for (int order = 1; order <= 2; ++order)
    {
        scatterG(S.P, gin);
        mlsTransportCompute(S.P, dtT, S.cp, S.dom, S.BC, order);
        std::string sfx = "_o" + std::to_string(order);
        npyF8(dir + "/gt" + sfx + ".npy", gatherG(S.P, true), {N, gs});
        std::vector<uint8_t> m = positivityMask(S.P);
        for (uint8_t v : m) fired[order] += v;
        npyU1(dir + "/posfired" + sfx + ".npy", m, {N});
    }
    std::printf("wrote %s: N=%zu, positivity fired on %d (linear) / %d (quadratic) particles\n",
                dir.c_str(), N, fired[1], fired[2]);
    return 0;
}

// Body record of the harness fixtures: shape (AB_BODY_* numbering), Npts,
// mass, T, size, centre (3), orientation (2), Rm (9), velocity (3),
// angularVel, omega (3), force (3), torque vector (3), moment of inertia,
// surface area per particle: 36 values.
std::vector<double> bodyRecord(const RigidBody &b)
{
    std::vector<double> r(36, 0.0);
    const double *Rm = b.orientation();
    if (auto *q = dynamic_cast<const Square *>(&b)) { r[0] = 1; r[4] = q->L; }
    else if (auto *q = dynamic_cast<const Circle *>(&b)) { r[0] = 2; r[4] = q->R; }
    else if (auto *q = dynamic_cast<const Sphere *>(&b)) { r[0] = 3; r[4] = q->R; }
    else if (auto *q = dynamic_cast<const Cube *>(&b)) { r[0] = 4; r[4] = q->L; }
    r[1] = b.Npts; r[2] = b.mass; r[3] = b.T;
    r[5] = b.cx; r[6] = b.cy; r[7] = b.cz; r[8] = b.ox; r[9] = b.oy;
    for (int k = 0; k < 9; ++k) r[10 + k] = Rm ? Rm[k] : (k % 4 == 0 ? 1.0 : 0.0);
    r[19] = b.velx; r[20] = b.vely; r[21] = b.velz; r[22] = b.angularVel;
    r[23] = b.omx; r[24] = b.omy; r[25] = b.omz;
    r[26] = b.forcex; r[27] = b.forcey; r[28] = b.forcez;
    r[29] = b.tqx; r[30] = b.tqy; r[31] = b.tqz;
    r[32] = b.getMomentOfInertia(); r[33] = b.getSurfaceArea();
    return r;
}

// The cloud of a body case as main() builds it (embedRigidBodies, wall
// state, parabolic lid, applyIC) and its body.
int modeBodyCase(const std::string &dir, const std::string &cname, int Nx, int Nv, double scale)
{
    Setup S;
    build(S, caseOf(cname), Nx, Nv, scale);
    const size_t N = S.P.size(), gs = S.P[0].g.size();
    writeMeta(dir, S, 0.0);
    writeCloud(dir, S);
    // applyIC gives every particle the same row: keep the first lattice
    // particle's and the last (body-surface) particle's
    std::vector<double> rows(S.P[0].g);
    rows.insert(rows.end(), S.P[N - 1].g.begin(), S.P[N - 1].g.end());
    npyF8(dir + "/g_ic_rows.npy", rows, {2, gs});
    npyF8(dir + "/mom_ic.npy", moments(S.P), {N, 5});
    npyF8(dir + "/body.npy", bodyRecord(*S.bodies[0]), {36});
    std::vector<double> par{scale};
    npyF8(dir + "/params.npy", par, {1});
    std::printf("wrote %s: N=%zu\n", dir.c_str(), N);
    return 0;
}

// nsteps of the reference's CPU time loop with a free body
// (Solver/Solver.cpp:632-714: transport, MOOD, commit, mass rescale, wall,
// moments, collision, moments, post-collision check; moveRigidBodies;
// manageRemove, mergeCloseParticles, manageRefill with the neighbour
// rebuilds) from opposing streams u_x = +-a (left / right of the middle)
// laid over the gas, so the body feels a shear from the first step. Writes
// the start (g0, mom0, wall0), the body record after every step (steps x 36)
// with the cloud counts (steps x 4: N, removed, merged, added), and the end
// cloud (x1, y1, z1, boundary1, bodyid1, wall1, mesh1, g1, mom1).
int modeBodySteps(const std::string &dir, const std::string &cname, int Nx, int Nv,
                  double scale, int nsteps, double a, int order)
{
    Setup S;
    build(S, caseOf(cname), Nx, Nv, scale);
    S.sp.mlsOrder = order;
    auto &P = S.P;
    const SimParameters &sp = S.sp;
    const double dt = sp.dt, R = S.gc.R;
    const double xmid = 0.5 * (S.dom.xleft + S.dom.xright);
    for (auto &p : P)
    {
        const double ux = (p.x < xmid) ? a : -a, rho = S.BC.rho, T = S.BC.Tw;
        const int n = p.Nv_local, n2 = n * n;
        for (int kv = 0; kv < (p.dim == 3 ? n : 1); ++kv)
            for (int jv = 0; jv < n; ++jv)
                for (int iv = 0; iv < n; ++iv)
                {
                    const double vx = p.vlo + iv * p.dv_local, vy = p.vlo + jv * p.dv_local;
                    const double vz = p.vlo + kv * p.dv_local;
                    if (p.dim == 3)
                        p.g[iv + n * (jv + n * kv)] = maxwellian3D(vx, vy, vz, ux, 0.0, 0.0, T, rho, R);
                    else
                    {
                        const double g1 = maxwellianG1(vx, vy, ux, 0.0, T, rho, R);
                        p.g[iv + n * jv] = g1;
                        p.g[iv + n * jv + n2] = 0.5 * R * T * g1;
                    }
                }
    }
    computeMoments(P, R);
    auto wallOf = [&]() {
        std::vector<double> w(P.size() * 7);
        for (size_t i = 0; i < P.size(); ++i)
        {
            const double q[7] = {P[i].wnx, P[i].wny, P[i].wnz, P[i].wUx, P[i].wUy, P[i].wUz, P[i].wT};
            std::memcpy(&w[7 * i], q, sizeof q);
        }
        return w;
    };
    writeMeta(dir, S, dt);
    writeCloud(dir, S);
    {
        const size_t N = P.size(), gs = P[0].g.size();
        npyF8(dir + "/g0.npy", gatherG(P, false), {N, gs});
        npyF8(dir + "/mom0.npy", moments(P), {N, 5});
        npyF8(dir + "/wall0.npy", wallOf(), {N, 7});
        npyF8(dir + "/body0.npy", bodyRecord(*S.bodies[0]), {36});
    }
    std::vector<double> hist;
    std::vector<int32_t> counts;
    for (int st = 0; st < nsteps; ++st)
    {
        double massOld = sumMass(P);
        mlsTransportCompute(P, dt, S.cp, S.dom, S.BC, sp.mlsOrder);
        if (sp.limiter == LimiterMode::MOOD)
        {
            int nf = checkTransportPAD(P);
            if (nf > 0) mlsTransportCompute(P, dt, S.cp, S.dom, S.BC, 1, true);
        }
        mlsTransportCommit(P);
        double massNew = sumMass(P);
        if (massNew > 1.0e-30) scaleG(P, massOld / massNew);
        applyGravity(P, sp.gravityY, dt);
        applyBoundaryBC(P, R, S.dom, S.BC);
        computeMoments(P, R);
        bgkCollision(P, dt, sp.tao, R, sp.tauVariable, sp.d);
        computeMoments(P, R);
        if (sp.limiter == LimiterMode::MOOD && moodCheck(P))
        {
            bgkCollision(P, dt, sp.tao, R, sp.tauVariable, sp.d);
            computeMoments(P, R);
        }
        moveRigidBodies(P, S.bodies, dt, S.dom, S.cp);
        const int removed = manageRemove(P, S.bodies, S.cp);
        updateVoxelsAndNeighbours(P, S.cp, S.dom);
        const int merged = mergeCloseParticles(P, S.cp);
        if (merged > 0) updateVoxelsAndNeighbours(P, S.cp, S.dom);
        const int added = manageRefill(P, S.bodies, S.sp, S.cp, S.dom, R);
        if (added > 0) updateVoxelsAndNeighbours(P, S.cp, S.dom);
        S.cp.N = static_cast<int>(P.size());
        const std::vector<double> r = bodyRecord(*S.bodies[0]);
        hist.insert(hist.end(), r.begin(), r.end());
        counts.insert(counts.end(), {(int32_t)P.size(), removed, merged, added});
    }
    npyF8(dir + "/body_hist.npy", hist, {(size_t)nsteps, 36});
    npyI4(dir + "/counts.npy", counts, {(size_t)nsteps, 4});
    const size_t N = P.size(), gs = P[0].g.size();
    std::vector<double> x(N), y(N), z(N), mesh(3 * N);
    std::vector<uint8_t> bnd(N);
    std::vector<int32_t> bid(N);
    for (size_t i = 0; i < N; ++i)
    {
        x[i] = P[i].x; y[i] = P[i].y; z[i] = P[i].z;
        bnd[i] = P[i].boundary ? 1 : 0; bid[i] = P[i].bodyId;
        mesh[3 * i] = P[i].meshVx; mesh[3 * i + 1] = P[i].meshVy; mesh[3 * i + 2] = P[i].meshVz;
    }
    npyF8(dir + "/x1.npy", x, {N});
    npyF8(dir + "/y1.npy", y, {N});
    npyF8(dir + "/z1.npy", z, {N});
    npyU1(dir + "/boundary1.npy", bnd, {N});
    npyI4(dir + "/bodyid1.npy", bid, {N});
    npyF8(dir + "/wall1.npy", wallOf(), {N, 7});
    npyF8(dir + "/mesh1.npy", mesh, {N, 3});
    npyF8(dir + "/g1.npy", gatherG(P, false), {N, gs});
    npyF8(dir + "/mom1.npy", moments(P), {N, 5});
    std::vector<double> par{scale, a, (double)order, (double)nsteps};
    npyF8(dir + "/params.npy", par, {4});
    const std::vector<double> r = bodyRecord(*S.bodies[0]);
    std::printf("wrote %s: N=%zu, body at (%.6e, %.6e, %.6e), v (%.4e, %.4e, %.4e), "
                "w (%.4e, %.4e, %.4e / %.4e)\n", dir.c_str(), N, r[5], r[6], r[7], r[19], r[20],
                r[21], r[23], r[24], r[25], r[22]);
    return 0;
}

// The reference's output writers on a perturbed cloud: its fields (x, y,
// z, rho, T, p, ux, uy, uz, boundary, moodFlag, Nv_local) and what
// writeVTK, computeDiagnostics and writeTimeseries (Core/Output.hpp) make of
// them: output_000007.vtk, timeseries.csv (two rows, t = 1e-10 and 2e-10)
// and the diagnostics. Some particles carry moodFlag so the count shows.
int modeWriters(const std::string &dir, const std::string &cname, int Nx, int Nv)
{
    Setup S;
    build(S, caseOf(cname), Nx, Nv, 1.0);
    perturb(S);
    computeMoments(S.P, S.gc.R);
    auto &P = S.P;
    const size_t N = P.size();
    for (size_t i = 0; i < N; ++i) P[i].moodFlag = (i % 7 == 3);
    writeMeta(dir, S, 0.0);
    std::vector<double> f(N * 9);
    std::vector<uint8_t> b(N), m(N);
    for (size_t i = 0; i < N; ++i)
    {
        const Particle &q = P[i];
        const double v[9] = {q.x, q.y, q.z, q.rho, q.T, q.p, q.ux, q.uy, q.uz};
        std::memcpy(&f[9 * i], v, sizeof v);
        b[i] = q.boundary ? 1 : 0;
        m[i] = q.moodFlag ? 1 : 0;
    }
    npyF8(dir + "/fields.npy", f, {N, 9});
    npyU1(dir + "/boundary.npy", b, {N});
    npyU1(dir + "/mood.npy", m, {N});
    writeVTK(P, dir, 7, 7e-11);
    const Diagnostics d = computeDiagnostics(P);
    std::vector<double> dv{d.rho_avg, d.KE, d.T_avg, d.p_avg, (double)d.moodCount, (double)d.n};
    npyF8(dir + "/diag.npy", dv, {6});
    writeTimeseries(dir + "/timeseries.csv", 1e-10, d);
    writeTimeseries(dir + "/timeseries.csv", 2e-10, d);
    std::printf("wrote %s: N=%zu\n", dir.c_str(), N);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 2) { std::fprintf(stderr, "usage: see header\n"); return 1; }
    std::string mode = argv[1];
    if (mode == "case3d" && argc == 5)
        return modeCase(argv[2], TestCase::Cavity3D, std::atoi(argv[3]), std::atoi(argv[4]));
    if (mode == "case2d" && argc == 5)
        return modeCase(argv[2], TestCase::Cavity2D, std::atoi(argv[3]), std::atoi(argv[4]));
    if (mode == "cube" && argc == 6)
        return modeCube(argv[2], std::atoi(argv[3]), std::atoi(argv[4]), std::atof(argv[5]));
    if (mode == "cpustep" && argc == 6)
        return modeCpuStep(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]),
                           std::atoi(argv[5]));
    if (mode == "ckpt" && argc == 6)
        return modeCkpt(argv[2], argv[3], std::atoi(argv[4]), std::atoi(argv[5]));
    if (mode == "mood" && argc == 10)
        return modeMood(argv[2], argv[3], std::atoi(argv[4]), std::atoi(argv[5]),
                        std::atoi(argv[6]), std::atoi(argv[7]), std::atof(argv[8]),
                        std::atof(argv[9]));
    if (mode == "cloud" && argc == 9)
        return modeCloud(argv[2], argv[3], std::atoi(argv[4]), std::atoi(argv[5]),
                         std::atof(argv[6]), std::atof(argv[7]), std::atof(argv[8]));
    if (mode == "bodycase" && argc == 7)
        return modeBodyCase(argv[2], argv[3], std::atoi(argv[4]), std::atoi(argv[5]),
                            std::atof(argv[6]));
    if (mode == "bodysteps" && argc == 10)
        return modeBodySteps(argv[2], argv[3], std::atoi(argv[4]), std::atoi(argv[5]),
                             std::atof(argv[6]), std::atoi(argv[7]), std::atof(argv[8]),
                             std::atoi(argv[9]));
    if (mode == "writers" && argc == 6)
        return modeWriters(argv[2], argv[3], std::atoi(argv[4]), std::atoi(argv[5]));
    if (mode == "pos2d" && argc == 8)
        return modePos2d(argv[2], std::atoi(argv[3]), std::atoi(argv[4]), std::atof(argv[5]),
                         std::atof(argv[6]), std::atof(argv[7]));
    if (mode == "ckptread" && argc == 7)
        return modeCkptRead(argv[2], argv[3], argv[4], std::atoi(argv[5]), std::atoi(argv[6]));
    std::fprintf(stderr, "bad arguments\n");
    return 1;
}
