// k_cloud.hip -- cloud management around moving rigid bodies, on the device:
// manageRemove, mergeCloseParticles and manageRefill
// (RigidBody/CloudManagement.cpp:17-275) with the neighbour rebuilds the
// moving-body step puts between them (Solver/Solver.cpp:379-398).
//
// The reference runs this block on the host after downloading the whole
// cloud and uploads it again afterwards (Solver.cpp:373,401); here the cloud
// stays on the device. Each stage turns into a keep/source map over the
// particles and one re-layout of the device arrays (relayout_cloud), so the
// order of the particles is the reference's: kept particles in their order,
// refilled ones appended in site order.
//
//   remove   one thread per particle: inside a body, or within 0.4 dx of a
//            body-surface particle (CloudManagement.cpp:17-69)
//   merge    one thread per particle finds the candidates (an interior
//            neighbour j > i closer than minDist); one thread then runs the
//            reference's greedy pass over the candidates in order (the
//            decisions depend on the earlier ones through the merged marks
//            only), and one block per merged pair averages g
//            (CloudManagement.cpp:71-119)
//   refill   one wave per lattice site of the margin box: the hole test over
//            the cloud, the body test and the surface clearance; then one
//            block adds the accepted sites in order, each with the
//            weight-averaged moments of the particles in its radius, the
//            previously added ones included (CloudManagement.cpp:121-275).
//            A site added earlier never blocks a later one: sites are dx
//            apart and the hole test's radius is 0.8 dx.
// The inside tests, the distances and the site positions are evaluated as
// the reference writes them, without contraction; the refill's weighted
// sums are a fixed-order block reduction (the reference sums in particle
// order: the averages agree to rounding).
#include "ab_device.hpp"
#include "ab_internal.hpp"

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

using namespace ab;

namespace {

struct DevBody
{
    int shape;
    double cx, cy, cz, size, ox, oy;
    double rm[9];
};

struct Bodies
{
    int n;
    DevBody b[kMaxBodies];
};

// isInObject / isInObject3 of the reference's bodies (RigidBody/Geometry.hpp):
// Square :178-185 and Circle :233-237 are 2-D (their 3-D test ignores z),
// Sphere :105-109 and Cube :287-293 are 3-D (their 2-D test takes z = cz).
__host__ __device__ inline bool in_body(const DevBody &b, double x, double y, double z, bool d3)
{
#pragma clang fp contract(off)
    switch (b.shape)
    {
    case AB_BODY_SQUARE:
    {
        const double xC = x - b.cx, yC = y - b.cy;
        const double D = b.ox * b.ox + b.oy * b.oy;
        const double xp = (b.ox * xC + b.oy * yC) / D;
        const double yp = (b.oy * xC - b.ox * yC) / D;
        return (fabs(xp) <= b.size / 2) && (fabs(yp) <= b.size / 2);
    }
    case AB_BODY_CIRCLE:
    {
        const double dx = x - b.cx, dy = y - b.cy;
        return dx * dx + dy * dy <= b.size * b.size;
    }
    case AB_BODY_SPHERE:
    {
        const double zz = d3 ? z : b.cz;
        const double dx = x - b.cx, dy = y - b.cy, dz = zz - b.cz;
        return dx * dx + dy * dy + dz * dz <= b.size * b.size;
    }
    case AB_BODY_CUBE:
    {
        const double zz = d3 ? z : b.cz;
        const double dx = x - b.cx, dy = y - b.cy, dz = zz - b.cz;
        const double bx = b.rm[0] * dx + b.rm[3] * dy + b.rm[6] * dz;
        const double by = b.rm[1] * dx + b.rm[4] * dy + b.rm[7] * dz;
        const double bz = b.rm[2] * dx + b.rm[5] * dy + b.rm[8] * dz;
        return fabs(bx) <= b.size / 2 && fabs(by) <= b.size / 2 && fabs(bz) <= b.size / 2;
    }
    default:
        return false;
    }
}

__device__ inline bool in_any(const Bodies &B, double x, double y, double z, bool d3)
{
    for (int k = 0; k < B.n; ++k)
        if (in_body(B.b[k], x, y, z, d3)) return true;
    return false;
}

// squared distance of (ax, ay, az) - (bx, by, bz), z ignored in 2-D, summed
// as the reference writes it
__device__ inline double dist2(double ax, double ay, double az, double bx, double by, double bz,
                               bool d3)
{
#pragma clang fp contract(off)
    const double dx = ax - bx, dy = ay - by;
    const double dz = d3 ? az - bz : 0.0;
    return dx * dx + dy * dy + dz * dz;
}

// Stable compaction of the indices i < n with flag[i] == want into out (one
// block of 1024 threads, each a contiguous range); count[0] = their number.
__global__ __launch_bounds__(1024) void k_compact(const unsigned char *__restrict__ flag, int want,
                                                  int n, int *__restrict__ out,
                                                  int *__restrict__ count)
{
    __shared__ int sh[1024];
    const int t = threadIdx.x;
    const int per = (n + 1023) / 1024, lo = min(t * per, n), hi = min(lo + per, n);
    int c = 0;
    for (int i = lo; i < hi; ++i) c += (flag[i] == want);
    sh[t] = c;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1)
    {
        const int v = t >= o ? sh[t - o] : 0;
        __syncthreads();
        sh[t] += v;
        __syncthreads();
    }
    int pos = sh[t] - c;
    for (int i = lo; i < hi; ++i)
        if (flag[i] == want) out[pos++] = i;
    if (t == 1023) count[0] = sh[1023];
}

__global__ void k_gather_int(const int *__restrict__ src, int *__restrict__ dst,
                             const int *__restrict__ map, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = map[i] >= 0 ? src[map[i]] : -1;
}

__global__ void k_not(const unsigned char *__restrict__ a, unsigned char *__restrict__ b, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i] ? 0 : 1;
}

__global__ void k_surface_flags(int n, const unsigned char *__restrict__ bnd,
                                const int *__restrict__ body, unsigned char *__restrict__ flag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = (bnd[i] && body[i] >= 0) ? 1 : 0;
}

// manageRemove (CloudManagement.cpp:17-69): keep[i] = 0 for an interior
// particle inside a body or within 0.4 dx of a body-surface particle.
__global__ void k_remove(int n, const double *__restrict__ x, const double *__restrict__ y,
                         const double *__restrict__ z, const unsigned char *__restrict__ bnd,
                         Bodies B, int d3, const int *__restrict__ surf, int nsurf, double hug2,
                         unsigned char *__restrict__ keep)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    bool drop = false;
    if (!bnd[i])
    {
        const double px = x[i], py = y[i], pz = z[i];
        drop = in_any(B, px, py, pz, d3);
        for (int k = 0; !drop && k < nsurf; ++k)
        {
            const int q = surf[k];
            drop = dist2(px, py, pz, x[q], y[q], z[q], d3) < hug2;
        }
    }
    keep[i] = drop ? 0 : 1;
}

// mergeCloseParticles, the candidates (CloudManagement.cpp:87-99): an
// interior particle with an interior neighbour j > i closer than minDist.
__global__ void k_merge_cand(int n, const double *__restrict__ x, const double *__restrict__ y,
                             const double *__restrict__ z, const unsigned char *__restrict__ bnd,
                             const int *__restrict__ neigh, const int *__restrict__ nneigh,
                             int d3, double min2, unsigned char *__restrict__ cand)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    bool c = false;
    if (!bnd[i])
    {
        const int *nb = neigh + (size_t)i * AB_MAX_NEIGH;
        for (int k = 0; k < nneigh[i] && !c; ++k)
        {
            const int j = nb[k];
            if (j <= i || j >= n || bnd[j]) continue;
            c = dist2(x[i], y[i], z[i], x[j], y[j], z[j], d3) < min2;
        }
    }
    cand[i] = c ? 1 : 0;
}

// mergeCloseParticles, the greedy pass (CloudManagement.cpp:87-111) over the
// candidates in order: particle i takes the mean position of itself and its
// first live interior neighbour j > i closer than minDist, and j is merged
// away. One thread: the decisions are sequential.
__global__ void k_merge_greedy(int n, const int *__restrict__ cand, const int *__restrict__ ncand,
                               double *__restrict__ x, double *__restrict__ y,
                               double *__restrict__ z, const unsigned char *__restrict__ bnd,
                               const int *__restrict__ neigh, const int *__restrict__ nneigh,
                               int d3, double min2, unsigned char *__restrict__ dead,
                               int *__restrict__ pairs, int *__restrict__ npairs)
{
#pragma clang fp contract(off)
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int np = 0;
    const int nc = ncand[0];
    for (int c = 0; c < nc; ++c)
    {
        const int i = cand[c];
        if (dead[i]) continue;
        const int *nb = neigh + (size_t)i * AB_MAX_NEIGH;
        for (int k = 0; k < nneigh[i]; ++k)
        {
            const int j = nb[k];
            if (j <= i || j >= n || dead[j] || bnd[j]) continue;
            if (dist2(x[i], y[i], z[i], x[j], y[j], z[j], d3) < min2)
            {
                x[i] = 0.5 * (x[i] + x[j]);
                y[i] = 0.5 * (y[i] + y[j]);
                if (d3) z[i] = 0.5 * (z[i] + z[j]);
                pairs[2 * np] = i;
                pairs[2 * np + 1] = j;
                ++np;
                dead[j] = 1;
                break;
            }
        }
    }
    npairs[0] = np;
}

// g of a merged pair: the average (one block per pair; the pairs are
// disjoint, a merged-away particle never merges again).
__global__ void k_merge_g(const int *__restrict__ pairs, const int *__restrict__ npairs,
                          double *__restrict__ g, int gstride)
{
    const int p = blockIdx.x;
    if (p >= npairs[0]) return;
    const size_t i = pairs[2 * p], j = pairs[2 * p + 1];
    for (int k = threadIdx.x; k < gstride; k += blockDim.x)
        g[i * gstride + k] = 0.5 * (g[i * gstride + k] + g[j * gstride + k]);
}

// Surface bounding box (CloudManagement.cpp:144-155): min x, max x, min y,
// max y, min z, max z over the surface particles.
__global__ __launch_bounds__(256) void k_bbox(const int *__restrict__ surf, int nsurf,
                                              const double *__restrict__ x,
                                              const double *__restrict__ y,
                                              const double *__restrict__ z,
                                              double *__restrict__ box)
{
    __shared__ double sh[6][256];
    double v[6] = {1e300, -1e300, 1e300, -1e300, 1e300, -1e300};
    for (int k = threadIdx.x; k < nsurf; k += 256)
    {
        const int q = surf[k];
        v[0] = fmin(v[0], x[q]); v[1] = fmax(v[1], x[q]);
        v[2] = fmin(v[2], y[q]); v[3] = fmax(v[3], y[q]);
        v[4] = fmin(v[4], z[q]); v[5] = fmax(v[5], z[q]);
    }
    for (int c = 0; c < 6; ++c) sh[c][threadIdx.x] = v[c];
    __syncthreads();
    if (threadIdx.x < 6)
    {
        const int c = threadIdx.x;
        double r = sh[c][0];
        for (int k = 1; k < 256; ++k) r = (c & 1) ? fmax(r, sh[c][k]) : fmin(r, sh[c][k]);
        box[c] = r;
    }
}

// The lattice sites of the margin box, site s = (vk - k0) ny nx + (vj - j0) nx
// + (vi - i0): the reference's loop order.
struct Sites
{
    int i0, j0, k0, nx, ny, n;
    double xleft, ybottom, zfront, dxl;
};

__device__ inline void site_pos(const Sites &S, int s, int d3, double &xm, double &ym, double &zm)
{
#pragma clang fp contract(off)
    const int vi = S.i0 + s % S.nx, vj = S.j0 + (s / S.nx) % S.ny, vk = S.k0 + s / (S.nx * S.ny);
    xm = S.xleft + vi * S.dxl;
    ym = S.ybottom + vj * S.dxl;
    zm = d3 ? S.zfront + vk * S.dxl : 0.0;
}

// manageRefill's site tests (CloudManagement.cpp:177-208), one wave per
// site: no particle within 0.8 dx, outside every body, no body-surface
// particle within 0.6 dx.
__global__ __launch_bounds__(256) void k_sites(Sites S, int n, const double *__restrict__ x,
                                               const double *__restrict__ y,
                                               const double *__restrict__ z, Bodies B, int d3,
                                               const int *__restrict__ surf, int nsurf,
                                               double hole2, double clr2,
                                               unsigned char *__restrict__ ok)
{
    const int s = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    if (s >= S.n) return;
    double xm, ym, zm;
    site_pos(S, s, d3, xm, ym, zm);
    bool near = false;
    for (int q = lane; q < n && !near; q += kWave)
        near = dist2(x[q], y[q], z[q], xm, ym, zm, d3) < hole2;
    bool hug = false;
    for (int k = lane; k < nsurf && !hug; k += kWave)
    {
        const int q = surf[k];
        hug = dist2(x[q], y[q], z[q], xm, ym, zm, d3) < clr2;
    }
    const bool bad = __ballot(near || hug) != 0;
    if (lane == 0) ok[s] = (!bad && !in_any(B, xm, ym, zm, d3)) ? 1 : 0;
}

// manageRefill's additions (CloudManagement.cpp:210-243), one block, the
// accepted sites in order: the mlsWeightSq-weighted means of rho, u and T
// over the particles within the radius (the earlier additions included); a
// site with no weight is skipped. add: 8 x nc (x y z rho ux uy uz T).
__global__ __launch_bounds__(1024) void k_fill(Sites S, const int *__restrict__ cand,
                                               const int *__restrict__ ncand, int n,
                                               const double *__restrict__ x,
                                               const double *__restrict__ y,
                                               const double *__restrict__ z,
                                               const double *__restrict__ rho,
                                               const double *__restrict__ ux,
                                               const double *__restrict__ uy,
                                               const double *__restrict__ uz,
                                               const double *__restrict__ T, int d3,
                                               double radius, double *__restrict__ add,
                                               int *__restrict__ nadd)
{
#pragma clang fp contract(off)
    __shared__ double part[6][1024 / kWave];
    __shared__ int na;
    const int t = threadIdx.x, w = t / kWave, lane = t & (kWave - 1);
    const int nc = ncand[0];
    if (t == 0) na = 0;
    __syncthreads();
    const double r2 = radius * radius;
    for (int c = 0; c < nc; ++c)
    {
        double xm, ym, zm;
        site_pos(S, cand[c], d3, xm, ym, zm);
        double v[6] = {0, 0, 0, 0, 0, 0};
        const int tot = n + na;
        for (int q = t; q < tot; q += 1024)
        {
            const bool a = q >= n;
            const int k = q - n;
            const double qx = a ? add[k] : x[q], qy = a ? add[nc + k] : y[q];
            const double qz = a ? add[2 * nc + k] : z[q];
            const double d2 = dist2(qx, qy, qz, xm, ym, zm, d3);
            if (d2 >= r2) continue;
            const double q2 = d2 / (radius * radius);       // mlsWeightSq, MLSBasis.hpp:142-148
            if (q2 >= 1.0) continue;
            const double wq = (1.0 - q2) * (1.0 - q2);
            v[0] += wq;
            v[1] += wq * (a ? add[3 * nc + k] : rho[q]);
            v[2] += wq * (a ? add[4 * nc + k] : ux[q]);
            v[3] += wq * (a ? add[5 * nc + k] : uy[q]);
            v[4] += wq * (a ? add[6 * nc + k] : uz[q]);
            v[5] += wq * (a ? add[7 * nc + k] : T[q]);
        }
        for (int m = 0; m < 6; ++m)
        {
            double r = wave_sum(v[m]);
            if (lane == 0) part[m][w] = r;
        }
        __syncthreads();
        if (t == 0)
        {
            double s[6];
            for (int m = 0; m < 6; ++m)
            {
                s[m] = 0.0;
                for (int k = 0; k < 1024 / kWave; ++k) s[m] += part[m][k];
            }
            if (!(s[0] < 1e-30))
            {
                const int k = na;
                add[k] = xm; add[nc + k] = ym; add[2 * nc + k] = zm;
                add[3 * nc + k] = s[1] / s[0];
                add[4 * nc + k] = s[2] / s[0];
                add[5 * nc + k] = s[3] / s[0];
                add[6 * nc + k] = d3 ? s[4] / s[0] : 0.0;
                add[7 * nc + k] = s[5] / s[0];
                na = k + 1;
            }
        }
        __syncthreads();
    }
    if (t == 0) nadd[0] = na;
}

// The added particles' state (CloudManagement.cpp:245-271): position, the
// averaged moments and the local Maxwellian (3-D: maxwellian3D; 2-D: g1 =
// maxwellianG1, g2 = 0.5 R T g1). One block per added particle.
__global__ void k_fill_rows(int n0, int na, int nc, const double *__restrict__ add, int d3,
                            int Nv, double vlo, double dv, double R, int gstride,
                            double *__restrict__ x, double *__restrict__ y,
                            double *__restrict__ z, double *__restrict__ rho,
                            double *__restrict__ ux, double *__restrict__ uy,
                            double *__restrict__ uz, double *__restrict__ T,
                            double *__restrict__ g)
{
    const int k = blockIdx.x;
    if (k >= na) return;
    const size_t i = (size_t)n0 + k;
    const double r = add[3 * nc + k], u = add[4 * nc + k], v = add[5 * nc + k];
    const double w = add[6 * nc + k], t = add[7 * nc + k];
    if (threadIdx.x == 0)
    {
        x[i] = add[k]; y[i] = add[nc + k]; z[i] = add[2 * nc + k];
        rho[i] = r; ux[i] = u; uy[i] = v; uz[i] = w; T[i] = t;
    }
    double *gi = g + i * gstride;
    if (d3)
    {
        for (int l = threadIdx.x; l < Nv * Nv * Nv; l += blockDim.x)
        {
            const int iv = l % Nv, jv = (l / Nv) % Nv, kv = l / (Nv * Nv);
            gi[l] = maxwellian3d(vnode(vlo, dv, iv), vnode(vlo, dv, jv), vnode(vlo, dv, kv), u,
                                 v, w, t, r, R);
        }
    }
    else
    {
#pragma clang fp contract(off)
        for (int l = threadIdx.x; l < Nv * Nv; l += blockDim.x)
        {
            const int iv = l % Nv, jv = l / Nv;
            const double g1 = maxwellian_g1(vnode(vlo, dv, iv), vnode(vlo, dv, jv), u, v, t, r, R);
            gi[l] = g1;
            gi[l + Nv * Nv] = 0.5 * R * t * g1;
        }
    }
}

// Device scratch of one call: the k-th request takes the context's pool
// slot k, allocated (or grown) only when it is smaller than asked and kept
// for the next call, so steady calls allocate and free nothing (a free
// synchronises the whole device). The requests of one call never share a
// slot, and a call's stream is synchronised before it returns.
struct Scratch
{
    size_t next = 0;
    ab_ctx *owner = nullptr;
    // an early return may leave kernels using the slots in flight
    ~Scratch() { if (owner) (void)hipStreamSynchronize(owner->stream); }
    template <typename T>
    int get(ab_ctx *c, T **p, size_t n)
    {
        owner = c;
        const size_t bytes = std::max<size_t>(n, 1) * sizeof(T), k = next++;
        if (k >= c->cm_pool.size())
        {
            c->cm_pool.push_back(nullptr);
            c->cm_pool_bytes.push_back(0);
        }
        if (c->cm_pool_bytes[k] < bytes)
        {
            if (c->cm_pool[k]) (void)hipFree(c->cm_pool[k]);
            c->cm_pool[k] = nullptr;
            c->cm_pool_bytes[k] = 0;
            void *q = nullptr;
            if (hipMalloc(&q, bytes) != hipSuccess)
            {
                c->err = "cloud management: device allocation failed";
                *p = nullptr;
                return AB_ERR_NOMEM;
            }
            c->cm_pool[k] = q;
            c->cm_pool_bytes[k] = bytes;
        }
        *p = static_cast<T *>(c->cm_pool[k]);
        return AB_OK;
    }
};

#define CMCHK(c, expr)                                                                  \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) {                                                         \
            (c)->err = std::string("cloud management: ") + #expr + ": " + hipGetErrorString(e_); \
            return AB_ERR_HIP;                                                          \
        }                                                                               \
    } while (0)

int read_int(ab_ctx *c, const int *dev, int *host)
{
    CMCHK(c, hipMemcpyAsync(host, dev, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    CMCHK(c, hipStreamSynchronize(c->stream));
    return AB_OK;
}

// Re-lays the cloud out by the keep flags (stable) and carries the body ids
// and the source map along; rebuilds neighbour and particle lists.
int apply_keep(ab_ctx *c, Scratch &sc, const unsigned char *keep, int **body)
{
    const int N = c->N;
    int *map = nullptr, *cnt = nullptr, *nb = nullptr;
    int rc;
    if ((rc = sc.get(c, &map, N)) || (rc = sc.get(c, &cnt, 1))) return rc;
    k_compact<<<1, 1024, 0, c->stream>>>(keep, 1, N, map, cnt);
    int n1 = 0;
    if ((rc = read_int(c, cnt, &n1))) return rc;
    if (n1 == N) return AB_OK;
    std::vector<int> hmap(n1);
    if (n1) CMCHK(c, hipMemcpyAsync(hmap.data(), map, sizeof(int) * n1, hipMemcpyDeviceToHost, c->stream));
    if ((rc = sc.get(c, &nb, n1))) return rc;
    if (n1) k_gather_int<<<(n1 + 255) / 256, 256, 0, c->stream>>>(*body, nb, map, n1);
    CMCHK(c, hipStreamSynchronize(c->stream));
    *body = nb;
    std::vector<int> src(n1);
    for (int i = 0; i < n1; ++i) src[i] = c->cm_src[hmap[i]];
    c->cm_src.swap(src);
    if ((rc = relayout_cloud(c, map, n1))) return rc;
    if ((rc = rebuild_lists(c))) return rc;
    return ab_rebuild_neighbours(c);
}

}  // namespace

extern "C" int ab_cloud_manage(ab_ctx *c, int nbodies, const ab_body *bodies, const int *body_id,
                               double dx, double min_dist, const double dom[6], int *n_new,
                               int counts[3])
{
    if (!c) return AB_ERR_ARG;
    if (!c->uploaded) { c->err = "ab_cloud_manage before ab_upload"; return AB_ERR_STATE; }
    if (c->cfg.precision != AB_PREC_FP64)
    { c->err = "ab_cloud_manage: the cloud changes, which the lattice tables cannot follow (AB_PREC_FP64 only)"; return AB_ERR_UNSUPPORTED; }
    if (c->nbatch)
    { c->err = "ab_cloud_manage: manage the members' clouds, then form the batch again"; return AB_ERR_UNSUPPORTED; }
    if (nbodies < 0 || nbodies > kMaxBodies || (nbodies && !bodies) || !dom ||
        !(dx != 0.0) || !std::isfinite(dx))
    { c->err = "ab_cloud_manage: invalid argument"; return AB_ERR_ARG; }
    Bodies B{};
    B.n = nbodies;
    for (int k = 0; k < nbodies; ++k)
    {
        const ab_body &b = bodies[k];
        if (b.shape < AB_BODY_SQUARE || b.shape > AB_BODY_CUBE)
        { c->err = "ab_cloud_manage: unknown body shape"; return AB_ERR_ARG; }
        DevBody &d = B.b[k];
        d.shape = b.shape; d.cx = b.c[0]; d.cy = b.c[1]; d.cz = b.c[2]; d.size = b.size;
        d.ox = b.o[0]; d.oy = b.o[1];
        for (int m = 0; m < 9; ++m) d.rm[m] = b.rm[m];
    }
    const int N0 = c->N;
    c->cm_src.resize(N0);
    for (int i = 0; i < N0; ++i) c->cm_src[i] = i;
    int removed = 0, merged = 0, added = 0;
    if (counts) counts[0] = counts[1] = counts[2] = 0;
    if (n_new) *n_new = N0;
    // the moving-body block runs only with bodies (Solver.cpp:349)
    if (nbodies == 0)
    {
        if (body_id && N0)
        {
            CMCHK(c, hipMemcpyAsync(c->bid, body_id, sizeof(int) * N0, hipMemcpyHostToDevice, c->stream));
            CMCHK(c, hipStreamSynchronize(c->stream));
        }
        return AB_OK;
    }
    hipStream_t s = c->stream;
    const bool d3 = c->dim == 3;
    const double dxl = std::fabs(dx);
    Scratch sc;
    int rc;
    int *body = nullptr, *surf = nullptr, *nsurf_d = nullptr;
    unsigned char *flag = nullptr;
    if ((rc = sc.get(c, &body, N0)) || (rc = sc.get(c, &nsurf_d, 1))) return rc;
    // the body ids given replace the device cloud's; without them the
    // device's are used (the relayouts below carry c->bid along)
    if (body_id && N0)
        CMCHK(c, hipMemcpyAsync(c->bid, body_id, sizeof(int) * N0, hipMemcpyHostToDevice, s));
    if (N0) CMCHK(c, hipMemcpyAsync(body, c->bid, sizeof(int) * N0, hipMemcpyDeviceToDevice, s));
    auto surface = [&](int &nsurf) -> int {
        const int N = c->N;
        int r;
        if ((r = sc.get(c, &flag, N)) || (r = sc.get(c, &surf, N))) return r;
        if (N) k_surface_flags<<<(N + 255) / 256, 256, 0, s>>>(N, c->bnd, body, flag);
        k_compact<<<1, 1024, 0, s>>>(flag, 1, N, surf, nsurf_d);
        return read_int(c, nsurf_d, &nsurf);
    };

    // 1. manageRemove, then the neighbour rebuild (Solver.cpp:379-387)
    {
        int nsurf = 0;
        if ((rc = surface(nsurf))) return rc;
        const double hug2 = std::pow(0.4 * dxl, 2);
        unsigned char *keep = nullptr;
        if ((rc = sc.get(c, &keep, N0))) return rc;
        if (N0)
            k_remove<<<(N0 + 255) / 256, 256, 0, s>>>(N0, c->x, c->y, c->z, c->bnd, B, d3, surf,
                                                      nsurf, hug2, keep);
        CMCHK(c, hipGetLastError());
        if ((rc = apply_keep(c, sc, keep, &body))) return rc;
        removed = N0 - c->N;
        if (removed == 0 && (rc = ab_rebuild_neighbours(c))) return rc;
    }
    // 2. mergeCloseParticles, and a rebuild when it merged (Solver.cpp:389-395)
    const double min2 = min_dist * min_dist;
    if (min2 > 0.0 && c->N > 0)
    {
        const int N = c->N;
        unsigned char *cand = nullptr, *dead = nullptr;
        int *clist = nullptr, *ncand = nullptr, *pairs = nullptr, *npairs = nullptr;
        if ((rc = sc.get(c, &cand, N)) || (rc = sc.get(c, &dead, N)) || (rc = sc.get(c, &clist, N)) ||
            (rc = sc.get(c, &ncand, 1)) || (rc = sc.get(c, &pairs, 2 * (size_t)N)) ||
            (rc = sc.get(c, &npairs, 1)))
            return rc;
        k_merge_cand<<<(N + 255) / 256, 256, 0, s>>>(N, c->x, c->y, c->z, c->bnd, c->neigh,
                                                     c->nneigh, d3, min2, cand);
        k_compact<<<1, 1024, 0, s>>>(cand, 1, N, clist, ncand);
        int nc = 0;
        if ((rc = read_int(c, ncand, &nc))) return rc;
        if (nc > 0)
        {
            CMCHK(c, hipMemsetAsync(dead, 0, N, s));
            k_merge_greedy<<<1, 1, 0, s>>>(N, clist, ncand, c->x, c->y, c->z, c->bnd, c->neigh,
                                           c->nneigh, d3, min2, dead, pairs, npairs);
            int np = 0;
            if ((rc = read_int(c, npairs, &np))) return rc;
            if (np > 0)
            {
                k_merge_g<<<np, 256, 0, s>>>(pairs, npairs, c->g, c->gstride);
                unsigned char *keep = cand;
                k_not<<<(N + 255) / 256, 256, 0, s>>>(dead, keep, N);
                if ((rc = apply_keep(c, sc, keep, &body))) return rc;
                merged = np;
            }
        }
    }
    // 3. manageRefill, and a rebuild when it added (Solver.cpp:394-396)
    {
        int nsurf = 0;
        if ((rc = surface(nsurf))) return rc;
        if (nsurf > 0)
        {
            double *box = nullptr;
            if ((rc = sc.get(c, &box, 6))) return rc;
            k_bbox<<<1, 256, 0, s>>>(surf, nsurf, c->x, c->y, c->z, box);
            double hb[6];
            CMCHK(c, hipMemcpyAsync(hb, box, sizeof hb, hipMemcpyDeviceToHost, s));
            CMCHK(c, hipStreamSynchronize(s));
            // the site range, as CloudManagement.cpp:156-171 computes it
            const double margin = 3.5 * dxl;
            double bx0 = hb[0] - margin, bx1 = hb[1] + margin;
            double by0 = hb[2] - margin, by1 = hb[3] + margin;
            double bz0 = hb[4] - margin, bz1 = hb[5] + margin;
            const int nxs = (int)std::llround((dom[1] - dom[0]) / dxl);
            const int nys = (int)std::llround((dom[3] - dom[2]) / dxl);
            const int nzs = d3 ? (int)std::llround((dom[5] - dom[4]) / dxl) : 0;
            const int i0 = std::max(1, (int)std::ceil((bx0 - dom[0]) / dxl));
            const int i1 = std::min(nxs - 1, (int)std::floor((bx1 - dom[0]) / dxl));
            const int j0 = std::max(1, (int)std::ceil((by0 - dom[2]) / dxl));
            const int j1 = std::min(nys - 1, (int)std::floor((by1 - dom[2]) / dxl));
            const int k0 = d3 ? std::max(1, (int)std::ceil((bz0 - dom[4]) / dxl)) : 0;
            const int k1 = d3 ? std::min(nzs - 1, (int)std::floor((bz1 - dom[4]) / dxl)) : 0;
            Sites S{};
            S.i0 = i0; S.j0 = j0; S.k0 = k0;
            S.nx = std::max(0, i1 - i0 + 1);
            S.ny = std::max(0, j1 - j0 + 1);
            const int nz = std::max(0, k1 - k0 + 1);
            S.n = S.nx * S.ny * nz;
            S.xleft = dom[0]; S.ybottom = dom[2]; S.zfront = dom[4]; S.dxl = dxl;
            if (S.n > 0)
            {
                const int N = c->N;
                unsigned char *ok = nullptr;
                int *clist = nullptr, *ncand = nullptr, *nadd = nullptr;
                double *add = nullptr;
                if ((rc = sc.get(c, &ok, S.n)) || (rc = sc.get(c, &clist, S.n)) ||
                    (rc = sc.get(c, &ncand, 1)) || (rc = sc.get(c, &nadd, 1)))
                    return rc;
                const double hole2 = std::pow(0.8 * dxl, 2), clr2 = std::pow(0.6 * dxl, 2);
                k_sites<<<(S.n + 3) / 4, 256, 0, s>>>(S, N, c->x, c->y, c->z, B, d3, surf, nsurf,
                                                      hole2, clr2, ok);
                k_compact<<<1, 1024, 0, s>>>(ok, 1, S.n, clist, ncand);
                int nc = 0;
                if ((rc = read_int(c, ncand, &nc))) return rc;
                if (nc > 0)
                {
                    if ((rc = sc.get(c, &add, (size_t)8 * nc))) return rc;
                    k_fill<<<1, 1024, 0, s>>>(S, clist, ncand, N, c->x, c->y, c->z, c->rho, c->ux,
                                              c->uy, c->uz, c->T, d3, c->cfg.radius, add, nadd);
                    int na = 0;
                    if ((rc = read_int(c, nadd, &na))) return rc;
                    if (na > 0)
                    {
                        const int n3 = N + na;
                        std::vector<int> hmap(n3);
                        for (int i = 0; i < n3; ++i) hmap[i] = i < N ? i : -1;
                        int *map = nullptr, *nb = nullptr;
                        if ((rc = sc.get(c, &map, n3)) || (rc = sc.get(c, &nb, n3))) return rc;
                        CMCHK(c, hipMemcpyAsync(map, hmap.data(), sizeof(int) * n3, hipMemcpyHostToDevice, s));
                        // refilled particles carry bodyId -1
                        k_gather_int<<<(n3 + 255) / 256, 256, 0, s>>>(body, nb, map, n3);
                        body = nb;
                        if ((rc = relayout_cloud(c, map, n3))) return rc;
                        k_fill_rows<<<na, 256, 0, s>>>(N, na, nc, add, d3, c->Nv, c->vlo, c->dv,
                                                       c->cfg.R, c->gstride, c->x, c->y, c->z,
                                                       c->rho, c->ux, c->uy, c->uz, c->T, c->g);
                        CMCHK(c, hipGetLastError());
                        c->cm_src.resize(n3, -1);
                        if ((rc = rebuild_lists(c)) || (rc = ab_rebuild_neighbours(c))) return rc;
                        added = na;
                    }
                }
            }
        }
    }
    CMCHK(c, hipStreamSynchronize(s));
    if (n_new) *n_new = c->N;
    if (counts) { counts[0] = removed; counts[1] = merged; counts[2] = added; }
    return AB_OK;
}

extern "C" int ab_download_body_ids(ab_ctx *c, int *body_id)
{
    if (!c || !body_id) return AB_ERR_ARG;
    if (!c->uploaded) { c->err = "no cloud uploaded"; return AB_ERR_STATE; }
    if (c->N)
    {
        CMCHK(c, hipMemcpyAsync(body_id, c->bid, sizeof(int) * (size_t)c->N, hipMemcpyDeviceToHost, c->stream));
        CMCHK(c, hipStreamSynchronize(c->stream));
    }
    return AB_OK;
}

extern "C" int ab_cloud_source_map(ab_ctx *c, int *src)
{
    if (!c || !src) return AB_ERR_ARG;
    if ((int)c->cm_src.size() != c->N)
    { c->err = "ab_cloud_source_map: no cloud management since the last upload"; return AB_ERR_STATE; }
    std::copy(c->cm_src.begin(), c->cm_src.end(), src);
    return AB_OK;
}
