// cases.cpp -- the six cavity cases (TestCases/TestCases.cpp:56-228) and the
// plain-cavity cloud: lattice, wall state, lid, initial state.
//
// Built with -ffp-contract=off (and no -march): positions, voxel indices and
// wall normals then round exactly as the reference's host code does, which
// keeps the neighbour lists identical to the reference's.
#include "alebgk_host.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;

struct CaseRow
{
    int id;
    const char *name;
    int dim;
    int has_body;
    int lid;
    double lid_speed;
};

// TestCases.cpp:56-75
const CaseRow kCases[6] = {
    {1, "cavity2d", 2, 0, ALH_LID_UNIFORM_TOP, 1.0},
    {2, "cavity3d", 3, 0, ALH_LID_UNIFORM_FRONT, 1.0},
    {3, "cavity2d-square", 2, 1, ALH_LID_PARABOLIC_TOP, 10.0},
    {4, "cavity2d-circle", 2, 1, ALH_LID_PARABOLIC_TOP, 10.0},
    {5, "cavity3d-sphere", 3, 1, ALH_LID_PARABOLIC_TOP, 10.0},
    {6, "cavity3d-cube", 3, 1, ALH_LID_PARABOLIC_TOP, 10.0},
};

int parse_case(const char *tok)
{
    if (!tok || !*tok) return -1;
    std::string t(tok);
    if (t.find_first_not_of("0123456789") == std::string::npos)
    {
        int n = std::atoi(tok);
        return (n >= 1 && n <= 6) ? n : -1;
    }
    for (const auto &c : kCases)
        if (t == c.name) return c.id;
    return -1;
}

double maxwellian3d(double vx, double vy, double vz, double ux, double uy, double uz, double T,
                    double rho, double R)
{
    double twoRT = 2.0 * R * T;
    double e = -((vx - ux) * (vx - ux) + (vy - uy) * (vy - uy) + (vz - uz) * (vz - uz)) / twoRT;
    double norm = std::pow(kPi * twoRT, 1.5);
    return rho / norm * std::exp(e);
}

double maxwellian_g1(double vx, double vy, double ux, double uy, double T, double rho, double R)
{
    double twoRT = 2.0 * R * T;
    double e = -((vx - ux) * (vx - ux) + (vy - uy) * (vy - uy)) / twoRT;
    return rho / (kPi * twoRT) * std::exp(e);
}

}  // namespace

extern "C" {

int alh_configure(const char *token, int nx, int nv, alh_case *o)
{
    if (!o) return AB_ERR_ARG;
    int id = parse_case(token);
    if (id < 0) return AB_ERR_ARG;
    std::memset(o, 0, sizeof *o);
    const CaseRow &r = kCases[id - 1];
    o->id = id;
    std::snprintf(o->name, sizeof o->name, "%s", r.name);
    o->dim = r.dim;
    o->has_body = r.has_body;
    o->lid = r.lid;
    o->lid_speed = r.lid_speed;
    o->Tw = 270.0;                 // kWallTemp
    o->R = 208.0;                  // GasConstants::R
    o->d_hs = 0.368e-9;            // kHSDiam
    o->mls_order = 1;              // SimParameters::mlsOrder default (Types.hpp:148)
    switch (id)
    {
    case 1:  // TestCases.cpp:151-160
        o->rho0 = 1.0; o->tau = 3.7142e-10; o->nx = o->ny = 100; o->nz = 1; o->nv = 44;
        o->vmax = 1303; o->vmin = -1303; o->dt = 1e-13; o->tfinal = 5e-9; o->save_every = 500;
        break;
    case 2:  // :164-173
        o->rho0 = 1.0; o->tau = 3.7142e-10; o->nx = o->ny = o->nz = 31; o->nv = 20;
        o->vmax = 1303; o->vmin = -1303; o->dt = 1e-11; o->tfinal = 5e-9; o->save_every = 500;
        break;
    case 3:
    case 4:  // :185-195
        o->rho0 = 11.0; o->tau = 3.37e-11; o->nx = o->ny = 60; o->nz = 1; o->nv = 30;
        o->vmax = 1500; o->vmin = -1500; o->dt = 1.5e-11; o->tfinal = 1e-6; o->save_every = 200;
        break;
    default:  // 5, 6: :214-224
        o->rho0 = 4.4; o->tau = 8.425e-11; o->nx = o->ny = o->nz = 40; o->nv = 10;
        o->vmax = 1303; o->vmin = -1303; o->dt = 1.5e-11; o->tfinal = 1.8e-7; o->save_every = 500;
        break;
    }
    // main.cpp:198-203 overrides
    if (nx > 0) { o->nx = o->ny = nx; if (r.dim == 3) o->nz = nx; }
    if (nv > 0) o->nv = nv;
    // DomainBoundary {xleft, xright, ytop, ybottom, zfront, zback} (:141-142)
    double d2[6] = {0, 1e-6, 1e-6, 0, 0, 0}, d3[6] = {0, 1e-6, 1e-6, 0, 1e-6, 0};
    std::memcpy(o->dom, r.dim == 2 ? d2 : d3, sizeof o->dom);
    return AB_OK;
}

int alh_lattice_size(const alh_case *cs, int *n)
{
    if (!cs || !n) return AB_ERR_ARG;
    *n = cs->nx * cs->ny * (cs->dim == 3 ? cs->nz : 1);
    return AB_OK;
}

}  // extern "C"

namespace alh {

// generateParticles (ParticleGeneration.cpp:35-114, non-periodic lattice),
// initWallData (WallState.cpp:16-26: box-wall particles only), the lid
// (TestCases.cpp:141-144 uniform, :230-249 parabolic) and the macro fields of
// applyIC (InitialConditions.cpp:18-81) for the lattice of a case. g is
// left to fill_ic. geo receives dx, dy, dz, minDist.
int lattice_core(const alh_case *cs, int precision, ab_config *cfg, double *x, double *y,
                 double *z, unsigned char *boundary, double *wall, double *mom, double geo[4])
{
    if (!cs || !cfg || !x || !y || !z || !boundary || !wall || !mom) return AB_ERR_ARG;
    const int ndim = cs->dim;
    const int Nx = cs->nx, Ny = cs->ny, Nz = (ndim == 3) ? cs->nz : 1;
    if (Nx < 2 || Ny < 2 || (ndim == 3 && Nz < 2) || cs->nv < 2) return AB_ERR_ARG;
    const double xleft = cs->dom[0], xright = cs->dom[1], ytop = cs->dom[2],
                 ybottom = cs->dom[3], zfront = cs->dom[4], zback = cs->dom[5];
    // generateParticles (ParticleGeneration.cpp:35-114), non-periodic
    const double Lx = xright - xleft, Ly = ytop - ybottom;
    const double Lz = (ndim == 3) ? (zback - zfront) : 1.0;
    const double dx = Lx / (Nx - 1), dy = Ly / (Ny - 1);
    const double dz = (ndim == 3 && Nz > 1) ? Lz / (Nz - 1) : 1.0;
    const double radius = 2.5 * std::max({std::fabs(dx), std::fabs(dy),
                                          (ndim == 3 ? std::fabs(dz) : std::fabs(dx))});
    if (geo)
    {
        geo[0] = dx; geo[1] = dy; geo[2] = dz;
        geo[3] = 0.5 * std::min({std::fabs(dx), std::fabs(dy),
                                 (ndim == 3 ? std::fabs(dz) : std::fabs(dy))});
    }
    const double hb = 3.0 * dx, vb = 3.0 * dy;
    const int nbh = (int)std::ceil(Lx / hb) + 1, nbv = (int)std::ceil(Ly / vb) + 1;
    double wb = 1.0;
    int nbw = 1;
    if (ndim == 3) { wb = 3.0 * dz; nbw = (int)std::ceil(Lz / wb) + 1; }

    std::memset(cfg, 0, sizeof *cfg);
    cfg->dim = ndim;
    cfg->n_particles = Nx * Ny * Nz;
    cfg->nv = cs->nv;
    cfg->vmin = cs->vmin; cfg->vmax = cs->vmax;
    cfg->R = cs->R; cfg->tau = cs->tau;
    cfg->radius = radius;
    cfg->mls_order = cs->mls_order;
    cfg->mls_tol = 1e-12;
    cfg->precision = precision;
    cfg->nx = Nx; cfg->ny = Ny; cfg->nz = (ndim == 3) ? Nz : 1;
    cfg->xleft = xleft; cfg->ybottom = ybottom; cfg->zfront = zfront;
    cfg->hbox = hb; cfg->vbox = vb; cfg->wbox = wb;
    cfg->nbhbox = nbh; cfg->nbvbox = nbv; cfg->nbwbox = nbw;

    // Face wall velocities (TestCases.cpp:143-144): left, right, top,
    // bottom, front, back.
    double U[6][3] = {{0}};
    if (cs->lid == ALH_LID_UNIFORM_TOP) U[2][0] = cs->lid_speed;
    if (cs->lid == ALH_LID_UNIFORM_FRONT) U[4][0] = cs->lid_speed;

    int n = 0;
    for (int k = 0; k < Nz; k++)
    {
        const double zz = (ndim == 3) ? zfront + k * dz : 0.0;
        for (int j = 0; j < Ny; j++)
        {
            const double yy = ybottom + j * dy;
            for (int i = 0; i < Nx; i++, n++)
            {
                const double xx = xleft + i * dx;
                x[n] = xx; y[n] = yy; z[n] = zz;
                const bool bx = (i == 0 || i == Nx - 1), by = (j == 0 || j == Ny - 1);
                const bool bz = (ndim == 3) && (k == 0 || k == Nz - 1);
                boundary[n] = (bx || by || bz) ? 1 : 0;
                double *w = wall + 7 * (size_t)n;
                for (int q = 0; q < 7; ++q) w[q] = 0.0;
                w[6] = 0.0;
                if (boundary[n])
                {
                    // getWallNormal (MLSBasis.hpp:228-287). In 2-D the
                    // lattice z = 0 matches both zfront and zback.
                    const double eps = 1.0e-13;
                    double nx = 0, ny = 0, nz = 0, ux = 0, uy = 0, uz = 0, tws = 0;
                    int nh = 0;
                    auto hit = [&](double ax, double ay, double az, int f) {
                        nx += ax; ny += ay; nz += az;
                        ux += U[f][0]; uy += U[f][1]; uz += U[f][2];
                        tws += cs->Tw; ++nh;
                    };
                    if (std::fabs(xx - xleft) < eps) hit(1.0, 0.0, 0.0, 0);
                    if (std::fabs(xx - xright) < eps) hit(-1.0, 0.0, 0.0, 1);
                    if (std::fabs(yy - ytop) < eps) hit(0.0, -1.0, 0.0, 2);
                    if (std::fabs(yy - ybottom) < eps) hit(0.0, 1.0, 0.0, 3);
                    if (std::fabs(zz - zfront) < eps) hit(0.0, 0.0, -1.0, 4);
                    if (std::fabs(zz - zback) < eps) hit(0.0, 0.0, 1.0, 5);
                    double T;
                    if (nh == 0) T = cs->Tw;
                    else if (nh == 1) T = tws;
                    else
                    {
                        double nl = std::sqrt(nx * nx + ny * ny + nz * nz);
                        if (nl > eps) { nx /= nl; ny /= nl; nz /= nl; }
                        ux /= nh; uy /= nh; uz /= nh;
                        T = tws / nh;
                    }
                    w[0] = nx; w[1] = ny; w[2] = nz; w[3] = ux; w[4] = uy; w[5] = uz; w[6] = T;
                    // applyLidProfile (TestCases.cpp:230-249), parabolic lid
                    if (cs->lid == ALH_LID_PARABOLIC_TOP && std::fabs(yy - ytop) <= 1e-13)
                    {
                        const double s = xx / (xright - xleft);
                        w[3] = cs->lid_speed * (2.0 * s) * (2.0 * s) * (2.0 - 2.0 * s) *
                               (2.0 - 2.0 * s);
                    }
                }
                // applyIC (InitialConditions.cpp:18-81): rest, T = 270 K
                double *m = mom + 5 * (size_t)n;
                m[0] = cs->rho0; m[1] = m[2] = m[3] = 0.0; m[4] = 270.0;
            }
        }
    }
    return AB_OK;
}

// applyIC's distribution: every particle starts from the same Maxwellian at
// rest (rho0, 270 K), one row evaluated node by node as applyIC does and
// copied to all n particles.
void fill_ic(const alh_case *cs, double *g, size_t n)
{
    const int Nv = cs->nv, ndim = cs->dim;
    const double dv = (cs->vmax - cs->vmin) / (Nv - 1), vlo = cs->vmin;
    const size_t gs = (ndim == 3) ? (size_t)Nv * Nv * Nv : (size_t)2 * Nv * Nv;
    std::vector<double> row(gs);
    if (ndim == 3)
    {
        for (int kv = 0; kv < Nv; ++kv)
            for (int jv = 0; jv < Nv; ++jv)
                for (int iv = 0; iv < Nv; ++iv)
                    row[iv + Nv * (jv + Nv * kv)] =
                        maxwellian3d(vlo + iv * dv, vlo + jv * dv, vlo + kv * dv, 0.0, 0.0, 0.0,
                                     270.0, cs->rho0, cs->R);
    }
    else
    {
        for (int jv = 0; jv < Nv; ++jv)
            for (int iv = 0; iv < Nv; ++iv)
            {
                const int lin = iv + Nv * jv;
                const double Mg1 =
                    maxwellian_g1(vlo + iv * dv, vlo + jv * dv, 0.0, 0.0, 270.0, cs->rho0, cs->R);
                row[lin] = Mg1;
                row[lin + Nv * Nv] = 0.5 * cs->R * 270.0 * Mg1;
            }
    }
    for (size_t p = 0; p < n; ++p) std::memcpy(g + p * gs, row.data(), gs * sizeof(double));
}

}  // namespace alh

extern "C" {

int alh_build_lattice(const alh_case *cs, int precision, ab_config *cfg, double *x, double *y,
                      double *z, unsigned char *boundary, double *wall, double *mom, double *g)
{
    if (!cs || !g) return AB_ERR_ARG;
    if (cs->has_body) return AB_ERR_UNSUPPORTED;
    int rc = alh::lattice_core(cs, precision, cfg, x, y, z, boundary, wall, mom, nullptr);
    if (rc) return rc;
    alh::fill_ic(cs, g, (size_t)cfg->n_particles);
    return AB_OK;
}

}  // extern "C"
