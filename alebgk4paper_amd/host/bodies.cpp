// bodies.cpp -- the body cases on the host: the reference's four rigid bodies
// (RigidBody/Geometry.hpp:19-362), makeBodies (TestCases/TestCases.cpp:251-294)
// and embedRigidBodies (Initialization/BodyEmbedding.cpp:18-104), producing
// the cloud of a body case and its bodies' initial state.
//
// Built like cases.cpp with -ffp-contract=off and no -march: the surface
// points, the inside tests and the normals round as the reference's host code
// does, so the embedded cloud is identical to the reference's.
#include "alebgk_host.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <vector>

namespace alh {
int lattice_core(const alh_case *cs, int precision, ab_config *cfg, double *x, double *y,
                 double *z, unsigned char *boundary, double *wall, double *mom, double geo[4]);
void fill_ic(const alh_case *cs, double *g, size_t n);
}  // namespace alh

namespace {

constexpr double kPi = 3.14159265358979323846;

// TestCases.cpp:36-54: body density ratio, sizes, surface samples, centre
constexpr double kBodyDensityRatio = 10.0;
constexpr double kSquareSide = 7.5e-8;
constexpr double kCircleRadius = kSquareSide / 2.0;
constexpr double kCubeSide = 1.5e-7;
constexpr double kSphereRadius = kCubeSide / 2.0;
constexpr int kSquarePts = 16;
constexpr int kCirclePts = 32;
constexpr int kSpherePts = 48;
constexpr int kCubePts = 96;
constexpr double kCx = 6e-7, kCy = 7e-7, kCz = 5e-7;
constexpr double kWallT = 270.0;

// Cube::toBody (Geometry.hpp:279-285): R^T (x - c).
void to_body(const ab_rigid_body &b, double x, double y, double z, double o[3])
{
    const double dx = x - b.c[0], dy = y - b.c[1], dz = z - b.c[2];
    const double *R = b.rm;
    o[0] = R[0] * dx + R[3] * dy + R[6] * dz;
    o[1] = R[1] * dx + R[4] * dy + R[7] * dz;
    o[2] = R[2] * dx + R[5] * dy + R[8] * dz;
}

// isInObject (2-D bodies: Square :178-185, Circle :233-237) and isInObject3
// (Sphere :105-109, Cube :287-293).
bool inside(const ab_rigid_body &b, double x, double y, double z, bool d3)
{
    switch (b.shape)
    {
    case AB_BODY_SQUARE:
    {
        const double xC = x - b.c[0], yC = y - b.c[1];
        const double D = b.o[0] * b.o[0] + b.o[1] * b.o[1];
        const double xp = (b.o[0] * xC + b.o[1] * yC) / D;
        const double yp = (b.o[1] * xC - b.o[0] * yC) / D;
        return (std::fabs(xp) <= b.size / 2) && (std::fabs(yp) <= b.size / 2);
    }
    case AB_BODY_CIRCLE:
    {
        const double dx = x - b.c[0], dy = y - b.c[1];
        return dx * dx + dy * dy <= b.size * b.size;
    }
    case AB_BODY_SPHERE:
    {
        // the 2-D test takes z = cz (Geometry.hpp:102-103)
        const double dx = x - b.c[0], dy = y - b.c[1], dz = (d3 ? z : b.c[2]) - b.c[2];
        return dx * dx + dy * dy + dz * dz <= b.size * b.size;
    }
    default:
    {
        double o[3];
        to_body(b, x, y, d3 ? z : b.c[2], o);
        return std::fabs(o[0]) <= b.size / 2 && std::fabs(o[1]) <= b.size / 2 &&
               std::fabs(o[2]) <= b.size / 2;
    }
    }
}

// getNormal / getNormal3: Sphere :116-122, Square :148-164, Circle :225-231,
// Cube :298-310 (the face whose body-frame coordinate is largest).
void normal(const ab_rigid_body &b, double x, double y, double z, double n[3])
{
    switch (b.shape)
    {
    case AB_BODY_SQUARE:
    {
        const double ox = b.o[0], oy = b.o[1];
        const double xp = x - b.c[0], yp = y - b.c[1];
        const double nn = std::sqrt(xp * xp + yp * yp);
        const double dots[4][3] = {
            {(xp * ox + yp * oy) / nn, ox, oy},
            {-(xp * ox + yp * oy) / nn, -ox, -oy},
            {(xp * oy - yp * ox) / nn, oy, -ox},
            {-(xp * oy - yp * ox) / nn, -oy, ox},
        };
        int best = 0;
        for (int i = 1; i < 4; ++i)
            if (dots[i][0] > dots[best][0]) best = i;
        n[0] = dots[best][1]; n[1] = dots[best][2]; n[2] = 0.0;
        return;
    }
    case AB_BODY_CIRCLE:
    {
        const double dx = x - b.c[0], dy = y - b.c[1];
        const double nn = std::sqrt(dx * dx + dy * dy);
        n[0] = dx / nn; n[1] = dy / nn; n[2] = 0.0;
        return;
    }
    case AB_BODY_SPHERE:
    {
        const double dx = x - b.c[0], dy = y - b.c[1], dz = z - b.c[2];
        const double nn = std::sqrt(dx * dx + dy * dy + dz * dz);
        n[0] = dx / nn; n[1] = dy / nn; n[2] = dz / nn;
        return;
    }
    default:
    {
        double o[3];
        to_body(b, x, y, z, o);
        int f = 0;
        for (int i = 1; i < 3; ++i)
            if (std::fabs(o[i]) > std::fabs(o[f])) f = i;
        const double s = (o[f] >= 0.0) ? 1.0 : -1.0;
        n[0] = s * b.rm[0 + f];
        n[1] = s * b.rm[3 + f];
        n[2] = s * b.rm[6 + f];
        return;
    }
    }
}

// generateBoundaryParticles(3): Sphere :130-143 (Fibonacci lattice), Square
// :187-210 (Npts/4 per side), Circle :239-246, Cube :315-333 (n x n per face).
std::vector<std::array<double, 3>> surface(const ab_rigid_body &b)
{
    std::vector<std::array<double, 3>> v;
    const double cx = b.c[0], cy = b.c[1], cz = b.c[2];
    switch (b.shape)
    {
    case AB_BODY_SPHERE:
    {
        const double R = b.size;
        const double ga = kPi * (3.0 - std::sqrt(5.0));
        for (int i = 0; i < b.npts; ++i)
        {
            const double zf = 1.0 - 2.0 * (i + 0.5) / b.npts;
            const double rf = std::sqrt(1.0 - zf * zf);
            const double th = ga * i;
            v.push_back({cx + R * rf * std::cos(th), cy + R * rf * std::sin(th), cz + R * zf});
        }
        break;
    }
    case AB_BODY_SQUARE:
    {
        const double L = b.size;
        const double t = std::atan2(b.o[1], b.o[0]);
        const double c = std::cos(t), s = std::sin(t);
        const int nSide = b.npts / 4;
        const double dl = L / nSide;
        const double tlx = cx + L * (c - s) / 2, tly = cy + L * (s + c) / 2;
        const double brx = cx + L * (s - c) / 2, bry = cy - L * (s + c) / 2;
        for (int i = 0; i < nSide; ++i) v.push_back({brx + dl * (i + 0.5) * c, bry + dl * (i + 0.5) * s, 0.0});
        for (int i = 0; i < nSide; ++i) v.push_back({tlx - dl * (i + 0.5) * c, tly - dl * (i + 0.5) * s, 0.0});
        for (int i = 0; i < nSide; ++i) v.push_back({brx - dl * (i + 0.5) * s, bry + dl * (i + 0.5) * c, 0.0});
        for (int i = 0; i < nSide; ++i) v.push_back({tlx + dl * (i + 0.5) * s, tly - dl * (i + 0.5) * c, 0.0});
        break;
    }
    case AB_BODY_CIRCLE:
    {
        const double R = b.size, dth = 2.0 * kPi / b.npts;
        for (int i = 0; i < b.npts; ++i)
            v.push_back({R * std::cos(i * dth) + cx, R * std::sin(i * dth) + cy, 0.0});
        break;
    }
    default:
    {
        const double L = b.size;
        const int n = std::max(1, (int)std::lround(std::sqrt(b.npts / 6.0)));
        const double dl = L / n;
        const double *R = b.rm;
        for (int f = 0; f < 3; ++f)
            for (int sgn = -1; sgn <= 1; sgn += 2)
                for (int i = 0; i < n; ++i)
                    for (int j = 0; j < n; ++j)
                    {
                        double q[3];
                        q[f] = sgn * L / 2.0;
                        q[(f + 1) % 3] = (i + 0.5) * dl - L / 2.0;
                        q[(f + 2) % 3] = (j + 0.5) * dl - L / 2.0;
                        v.push_back({cx + R[0] * q[0] + R[1] * q[1] + R[2] * q[2],
                                     cy + R[3] * q[0] + R[4] * q[1] + R[5] * q[2],
                                     cz + R[6] * q[0] + R[7] * q[1] + R[8] * q[2]});
                    }
        break;
    }
    }
    return v;
}

// The embedding of a body case: lattice particles kept, then each body's
// surface particles (BodyEmbedding.cpp:35-100). keep[p] for every lattice
// particle; surf[b] the surface points of body b.
struct Embedding
{
    std::vector<unsigned char> keep;
    std::vector<std::vector<std::array<double, 3>>> surf;
    int n_keep = 0, n_surf = 0;
};

int embed(const alh_case *cs, const ab_rigid_body *bodies, int nb, const double *x,
          const double *y, const double *z, int nlat, double dx, Embedding &E)
{
    const bool d3 = cs->dim == 3;
    E.surf.assign(nb, {});
    E.n_surf = 0;
    for (int b = 0; b < nb; ++b)
    {
        E.surf[b] = surface(bodies[b]);
        E.n_surf += (int)E.surf[b].size();
    }
    const double hug2 = std::pow(0.4 * std::fabs(dx), 2);
    E.keep.assign(nlat, 1);
    E.n_keep = 0;
    for (int p = 0; p < nlat; ++p)
    {
        bool drop = false;
        for (int b = 0; b < nb && !drop; ++b)
        {
            if (inside(bodies[b], x[p], y[p], z[p], d3)) { drop = true; break; }
            for (const auto &q : E.surf[b])
            {
                const double ddx = x[p] - q[0], ddy = y[p] - q[1];
                const double d2 = d3 ? ddx * ddx + ddy * ddy + (z[p] - q[2]) * (z[p] - q[2])
                                     : ddx * ddx + ddy * ddy;
                if (d2 < hug2) { drop = true; break; }
            }
        }
        E.keep[p] = drop ? 0 : 1;
        E.n_keep += drop ? 0 : 1;
    }
    return AB_OK;
}

// Lattice of the case (no g) into temporary arrays.
struct Lattice
{
    ab_config cfg{};
    std::vector<double> x, y, z, wall, mom;
    std::vector<unsigned char> bnd;
    double geo[4] = {0, 0, 0, 0};
    int n = 0;
};

int lattice(const alh_case *cs, int precision, Lattice &L)
{
    alh_case plain = *cs;
    int n = 0;
    alh_lattice_size(&plain, &n);
    L.n = n;
    L.x.resize(n); L.y.resize(n); L.z.resize(n);
    L.bnd.resize(n);
    L.wall.resize(7 * (size_t)n);
    L.mom.resize(5 * (size_t)n);
    return alh::lattice_core(&plain, precision, &L.cfg, L.x.data(), L.y.data(), L.z.data(),
                             L.bnd.data(), L.wall.data(), L.mom.data(), L.geo);
}

}  // namespace

extern "C" {

int alh_make_bodies(const alh_case *cs, double body_scale, ab_rigid_body *bodies, int *nbodies)
{
    if (!cs || !nbodies) return AB_ERR_ARG;
    *nbodies = 0;
    if (!cs->has_body) return AB_OK;
    if (!bodies || !(body_scale > 0.0)) return AB_ERR_ARG;
    ab_rigid_body &b = bodies[0];
    std::memset(&b, 0, sizeof b);
    const double rho = cs->rho0 * kBodyDensityRatio;
    b.T = kWallT;
    b.rotates = 1;
    b.c[0] = kCx; b.c[1] = kCy;
    b.o[0] = 1.0; b.o[1] = 0.0;                 // RigidBody defaults (Geometry.hpp:37-38)
    for (int k = 0; k < 9; ++k) b.rm[k] = (k % 4 == 0) ? 1.0 : 0.0;
    switch (cs->id)
    {
    case 3:   // Square (Geometry.hpp:145-150): orientation (0, 1), normalised
    {
        b.shape = AB_BODY_SQUARE;
        b.size = kSquareSide * body_scale;
        b.npts = kSquarePts;
        b.mass = rho * b.size * b.size;
        double ox = 0.0, oy = 1.0, n = std::sqrt(ox * ox + oy * oy);
        b.o[0] = ox / n; b.o[1] = oy / n;
        break;
    }
    case 4:   // Circle (:219-221)
        b.shape = AB_BODY_CIRCLE;
        b.size = kCircleRadius * body_scale;
        b.npts = kCirclePts;
        b.mass = rho * kPi * b.size * b.size;
        break;
    case 5:   // Sphere (:93-100)
        b.shape = AB_BODY_SPHERE;
        b.size = kSphereRadius * body_scale;
        b.npts = kSpherePts;
        b.c[2] = kCz;
        b.mass = rho * 4.0 / 3.0 * kPi * b.size * b.size * b.size;
        break;
    case 6:   // Cube (:262-268)
        b.shape = AB_BODY_CUBE;
        b.size = kCubeSide * body_scale;
        b.npts = kCubePts;
        b.c[2] = kCz;
        b.mass = rho * b.size * b.size * b.size;
        break;
    default:
        return AB_ERR_ARG;
    }
    *nbodies = 1;
    return AB_OK;
}

int alh_body_case_size(const alh_case *cs, double body_scale, int *n)
{
    if (!cs || !n) return AB_ERR_ARG;
    Lattice L;
    int rc = lattice(cs, AB_PREC_FP64, L);
    if (rc) return rc;
    ab_rigid_body bodies[1];
    int nb = 0;
    if ((rc = alh_make_bodies(cs, body_scale, bodies, &nb))) return rc;
    Embedding E;
    embed(cs, bodies, nb, L.x.data(), L.y.data(), L.z.data(), L.n, L.geo[0], E);
    *n = E.n_keep + E.n_surf;
    return AB_OK;
}

int alh_build_body_case(const alh_case *cs, double body_scale, int precision, ab_config *cfg,
                        double *x, double *y, double *z, unsigned char *boundary, double *wall,
                        double *mom, double *g, int *body_id, ab_rigid_body *bodies,
                        int *nbodies, double geo[4])
{
    if (!cs || !cfg || !x || !y || !z || !boundary || !wall || !mom || !g || !body_id ||
        !nbodies)
        return AB_ERR_ARG;
    Lattice L;
    int rc = lattice(cs, precision, L);
    if (rc) return rc;
    ab_rigid_body local[1];
    ab_rigid_body *bb = bodies ? bodies : local;
    int nb = 0;
    if ((rc = alh_make_bodies(cs, body_scale, bb, &nb))) return rc;
    *nbodies = nb;
    Embedding E;
    embed(cs, bb, nb, L.x.data(), L.y.data(), L.z.data(), L.n, L.geo[0], E);
    *cfg = L.cfg;
    int n = 0;
    for (int p = 0; p < L.n; ++p)
    {
        if (!E.keep[p]) continue;
        x[n] = L.x[p]; y[n] = L.y[p]; z[n] = L.z[p];
        boundary[n] = L.bnd[p];
        std::memcpy(wall + 7 * (size_t)n, &L.wall[7 * (size_t)p], 7 * sizeof(double));
        std::memcpy(mom + 5 * (size_t)n, &L.mom[5 * (size_t)p], 5 * sizeof(double));
        body_id[n] = -1;
        ++n;
    }
    const bool d3 = cs->dim == 3;
    for (int b = 0; b < nb; ++b)
        for (const auto &q : E.surf[b])
        {
            x[n] = q[0]; y[n] = q[1]; z[n] = d3 ? q[2] : 0.0;
            boundary[n] = 1;
            double nn[3];
            normal(bb[b], x[n], y[n], z[n], nn);
            double *w = wall + 7 * (size_t)n;
            w[0] = nn[0]; w[1] = nn[1]; w[2] = d3 ? nn[2] : 0.0;
            w[3] = bb[b].vel[0]; w[4] = bb[b].vel[1]; w[5] = bb[b].vel[2];
            w[6] = bb[b].T;
            // applyIC then sets every particle to the gas at rest
            double *m = mom + 5 * (size_t)n;
            m[0] = cs->rho0; m[1] = m[2] = m[3] = 0.0; m[4] = 270.0;
            body_id[n] = b;
            ++n;
        }
    cfg->n_particles = n;
    // a body cloud is no lattice: the table pipelines refuse it
    cfg->nx = cfg->ny = cfg->nz = 0;
    alh::fill_ic(cs, g, (size_t)n);
    if (geo) std::memcpy(geo, L.geo, sizeof L.geo);
    return AB_OK;
}

}  // extern "C"
