// output.cpp -- the reference's run outputs from downloaded arrays: the ASCII
// VTK file, the diagnostics and their CSV row, the one-line step summary
// (Core/Output.hpp:17-115), the body log (Solver/Solver.cpp:96-121) and the
// stability guard (Core/Stability.hpp:46-123).
//
// The files are written with the same stream formatting as the reference
// (scientific after the VTK header, 6 digits), so they are byte-identical to
// the reference's for the same values; the guard reads the same environment
// variables and stops on the same bounds.
#include "alebgk_host.h"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>

namespace {

double env_fac(const char *name, double fallback)
{
    const char *e = std::getenv(name);
    if (!e || !*e) return fallback;
    const double v = std::atof(e);
    return (v > 0.0) ? v : fallback;
}

}  // namespace

extern "C" {

int alh_write_vtk(const char *out_dir, int step, double t, int n, const double *x,
                  const double *y, const double *z, const double *rho, const double *T,
                  const double *p, int nv_local, const unsigned char *mood,
                  const unsigned char *boundary, const double *ux, const double *uy,
                  const double *uz)
{
    if (!out_dir || n < 0 || (n > 0 && (!x || !y || !z || !rho || !T || !p || !mood ||
                                        !boundary || !ux || !uy || !uz)))
        return AB_ERR_ARG;
    std::ostringstream fname;
    fname << out_dir << "/output_" << std::setw(6) << std::setfill('0') << step << ".vtk";
    std::ofstream f(fname.str());
    if (!f) return AB_ERR_ARG;
    f << "# vtk DataFile Version 3.0\n";
    f << "ALEBGK t=" << std::scientific << t << "\n";
    f << "ASCII\nDATASET UNSTRUCTURED_GRID\n";
    f << "POINTS " << n << " double\n";
    for (int i = 0; i < n; ++i) f << x[i] << " " << y[i] << " " << z[i] << "\n";
    f << "CELLS 0 0\nCELL_TYPES 0\nPOINT_DATA " << n << "\n";
    auto scalar = [&](const char *name, auto fn) {
        f << "SCALARS " << name << " double 1\nLOOKUP_TABLE default\n";
        for (int i = 0; i < n; ++i) f << fn(i) << "\n";
    };
    scalar("rho", [&](int i) { return rho[i]; });
    scalar("T", [&](int i) { return T[i]; });
    scalar("p", [&](int i) { return p[i]; });
    scalar("Nv_local", [&](int) { return static_cast<double>(nv_local); });
    scalar("moodFlag", [&](int i) { return static_cast<double>(mood[i] != 0); });
    scalar("boundary", [&](int i) { return static_cast<double>(boundary[i] != 0); });
    f << "VECTORS velocity double\n";
    for (int i = 0; i < n; ++i) f << ux[i] << " " << uy[i] << " " << uz[i] << "\n";
    f.close();
    return f ? AB_OK : AB_ERR_ARG;
}

int alh_diagnostics(int n, const unsigned char *boundary, const double *rho, const double *ux,
                    const double *uy, const double *uz, const double *T, const double *p,
                    const unsigned char *mood, alh_diag *d)
{
    if (!d || n < 0 || (n > 0 && (!boundary || !rho || !ux || !uy || !uz || !T || !p || !mood)))
        return AB_ERR_ARG;
    std::memset(d, 0, sizeof *d);
    for (int i = 0; i < n; ++i)
    {
        if (boundary[i]) continue;
        d->rho_avg += rho[i];
        d->KE += 0.5 * rho[i] * (ux[i] * ux[i] + uy[i] * uy[i] + uz[i] * uz[i]);
        d->T_avg += T[i];
        d->p_avg += p[i];
        d->mood_count += (mood[i] != 0);
        d->n++;
    }
    if (d->n > 0)
    {
        const double m = d->n;
        d->rho_avg /= m;
        d->KE /= m;
        d->T_avg /= m;
        d->p_avg /= m;
    }
    return AB_OK;
}

int alh_write_timeseries(const char *path, int first, double t, const alh_diag *d)
{
    if (!path || !d) return AB_ERR_ARG;
    std::ofstream f(path, first ? std::ios::out : std::ios::app);
    if (!f) return AB_ERR_ARG;
    if (first) f << "t,rho_avg,KE,T_avg,p_avg,mood_count\n";
    f << std::scientific << t << "," << d->rho_avg << "," << d->KE << "," << d->T_avg << ","
      << d->p_avg << "," << d->mood_count << "\n";
    return f ? AB_OK : AB_ERR_ARG;
}

int alh_format_step(int step, double t, double dt, const alh_diag *d, char *out, int cap)
{
    if (!d || !out || cap <= 0) return AB_ERR_ARG;
    std::snprintf(out, (size_t)cap, "Step%6d|t=%9.3e|dt=%7.2e|rho=%7.4f|T=%7.4f|KE=%7.3e|MOOD=%3d",
                  step, t, dt, d->rho_avg, d->T_avg, d->KE, d->mood_count);
    return AB_OK;
}

int alh_write_bodies(const char *path, int first, double t, int dim, int nbodies,
                     const ab_rigid_body *bodies)
{
    if (!path || nbodies < 0 || (nbodies > 0 && !bodies)) return AB_ERR_ARG;
    std::ofstream f(path, first ? std::ios::out : std::ios::app);
    if (!f) return AB_ERR_ARG;
    if (first) f << "t,body,cx,cy,velx,vely,angVel,Fx,Fy,torque,cz,velz,omx,omy,Fz,tqx,tqy\n";
    for (int b = 0; b < nbodies; ++b)
    {
        const ab_rigid_body &r = bodies[b];
        const double wz = (dim == 3) ? r.om[2] : r.angular_vel;
        f << std::scientific << t << "," << b << "," << r.c[0] << "," << r.c[1] << ","
          << r.vel[0] << "," << r.vel[1] << "," << wz << "," << r.force[0] << ","
          << r.force[1] << "," << r.tq[2] << "," << r.c[2] << "," << r.vel[2] << ","
          << r.om[0] << "," << r.om[1] << "," << r.force[2] << "," << r.tq[0] << ","
          << r.tq[1] << "\n";
    }
    return f ? AB_OK : AB_ERR_ARG;
}

int alh_guard_init(alh_guard *g, double wall_temp, double rho0)
{
    if (!g) return AB_ERR_ARG;
    std::memset(g, 0, sizeof *g);
    g->Tw = wall_temp;
    g->rho0 = rho0;
    g->enabled = std::getenv("ALEBGK_NO_GUARD") == nullptr;
    g->t_lo = env_fac("ALEBGK_GUARD_TMIN", 0.60);
    g->t_hi = env_fac("ALEBGK_GUARD_TMAX", 1.30);
    g->rho_lo = env_fac("ALEBGK_GUARD_RHOMIN", 0.20);
    g->rho_hi = env_fac("ALEBGK_GUARD_RHOMAX", 5.00);
    g->ke_hi = env_fac("ALEBGK_GUARD_KEMAX", 100.0);
    return AB_OK;
}

int alh_guard_check(alh_guard *g, int step, double t, const alh_diag *d, char *why, int cap)
{
    if (!g || !d) return AB_ERR_ARG;
    (void)step;
    (void)t;
    if (why && cap > 0) why[0] = 0;
    if (!g->enabled) return 0;
    auto fail = [&](const char *fmt, ...) {
        if (why && cap > 0)
        {
            va_list ap;
            va_start(ap, fmt);
            std::vsnprintf(why, (size_t)cap, fmt, ap);
            va_end(ap);
        }
        return 1;
    };
    if (!std::isfinite(d->rho_avg) || !std::isfinite(d->T_avg) || !std::isfinite(d->KE))
        return fail("a diagnostic is NaN or infinite");
    if (d->T_avg < g->t_lo * g->Tw)
        return fail("mean temperature below %.4g K (%.2f x Tw)", g->t_lo * g->Tw, g->t_lo);
    if (d->T_avg > g->t_hi * g->Tw)
        return fail("mean temperature above %.4g K (%.2f x Tw)", g->t_hi * g->Tw, g->t_hi);
    if (g->rho0 > 0.0)
    {
        if (d->rho_avg < g->rho_lo * g->rho0)
            return fail("mean density below %.4g (%.2f x rho0)", g->rho_lo * g->rho0, g->rho_lo);
        if (d->rho_avg > g->rho_hi * g->rho0)
            return fail("mean density above %.4g (%.2f x rho0)", g->rho_hi * g->rho0, g->rho_hi);
    }
    // the kinetic-energy reference is the first diagnostic with a usable
    // value: the gas and the body start at rest
    if (!g->have_ke && d->KE > 0.0)
    {
        g->ke_ref = d->KE;
        g->have_ke = 1;
    }
    if (g->have_ke && d->KE > g->ke_hi * g->ke_ref)
        return fail("kinetic energy %.3g x its early value (ceiling %.0f x)", d->KE / g->ke_ref,
                    g->ke_hi);
    return 0;
}

}  // extern "C"
