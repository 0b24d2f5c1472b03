// k_table.hip -- the table pipeline of the ALE-BGK step (AB_PREC_MIXED with
// FP32 storage and arithmetic, AB_PREC_FP64_TABLE with FP64).
//
// MLS transport is linear in g: for every (particle i, node v)
//     gt[i,v] = sum_o C[i,o,v] g[i+o,v]
// with C fixed by the cloud geometry (foot point x_i - v dt, the weights and
// the basis). The reference re-assembles and factorises the normal equations
// every step (Solver/CudaBackend.cu:699-869). On a lattice every particle's
// stencil is one of 5^3 = 125 classes (per axis: first, second, interior,
// second-to-last, last lattice plane), so the setup factorises each class once
// in FP64 (q = A^-1 e0, C_o = w_o phi_o^T q, the Shepard weights where the
// pivot collapses, identity on gated boundary nodes) and the step applies it.
//
// Stencil: the 81 lattice offsets with ox^2+oy^2+oz^2 <= 6 (radius 2.5 dx
// includes exactly these; self included for the gated nodes), grouped in 21
// (ox, oy) columns of 5 or 3 z-offsets.
//
// Step (four kernels, six with the discrete collision; no host synchronisation):
//   k_apply      dt = sum_o C d  (stored deviation, below) one wave = 64 nodes of a
//                lattice column (ix, iy); marches along z in blocks of K
//                outputs with the class coefficients in registers and each
//                loaded value reused by up to 5 outputs.
//   k_post       one wave per particle: positivity (fallback to the
//                pre-transport value), NAD, then the diffuse wall (boundary)
//                or moments + BGK collision (interior), written back in place
//                of g unscaled; FP64 partial sums per particle. With the
//                discrete collision it stops after the moments, k_dm_solve
//                (one thread per particle) finds the discrete Maxwellian and
//                k_collide (one wave per particle) applies the collision.
//   k_reduce     mass sums -> the two global rescale factors, folded into one
//                lazy factor the stored g carries into the next step.
//   k_finalize   per-particle rho, u, T of the step's end.
// Homogeneity makes the lazy factor exact: wall, moments (u, T) and
// collision are all homogeneous of degree 1 in a uniform rescale of g.
#include "ab_device.hpp"
#include "ab_internal.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <utility>
#include <vector>

namespace ab {

namespace {

constexpr int kOff = 81;      // stencil offsets
constexpr int kCols = 21;     // (ox, oy) columns
constexpr int kNcls = 125;    // 5 categories per axis

struct Col { int ox, oy, d, o0; };

// Columns in canonical order; o0 is the index of the column's oz = -d offset
// (the running sum of 2 d + 1). Tables, not loops: the kernels' templates
// evaluate these per (union column, output) at compile time.
__host__ __device__ constexpr Col col_at(int c)
{
    constexpr int t[kCols][4] = {
        {-1, -1, 2, 0}, {0, -1, 2, 5}, {1, -1, 2, 10}, {-1, 0, 2, 15}, {0, 0, 2, 20},
        {1, 0, 2, 25}, {-1, 1, 2, 30}, {0, 1, 2, 35}, {1, 1, 2, 40},
        {-2, -1, 1, 45}, {-2, 0, 1, 48}, {-2, 1, 1, 51}, {2, -1, 1, 54}, {2, 0, 1, 57},
        {2, 1, 1, 60}, {-1, -2, 1, 63}, {0, -2, 1, 66}, {1, -2, 1, 69}, {-1, 2, 1, 72},
        {0, 2, 1, 75}, {1, 2, 1, 78}};
    return Col{t[c][0], t[c][1], t[c][2], t[c][3]};
}

// Stencil column of offset (ox, oy), or -1.
__host__ __device__ constexpr int col_of(int ox, int oy)
{
    constexpr int t[5][5] = {{-1, 9, 10, 11, -1}, {15, 0, 3, 6, 18}, {16, 1, 4, 7, 19},
                             {17, 2, 5, 8, 20}, {-1, 12, 13, 14, -1}};
    return (ox < -2 || ox > 2 || oy < -2 || oy > 2) ? -1 : t[ox + 2][oy + 2];
}

// The tables agree with each other and with the stencil |o|^2 < 6.25.
constexpr bool col_tables_agree()
{
    int o0 = 0, n = 0;
    for (int c = 0; c < kCols; ++c)
    {
        const Col k = col_at(c);
        if (k.o0 != o0 || col_of(k.ox, k.oy) != c) return false;
        o0 += 2 * k.d + 1;
    }
    for (int ox = -2; ox <= 2; ++ox)
        for (int oy = -2; oy <= 2; ++oy)
        {
            int d = -1;
            for (int oz = 0; oz <= 2; ++oz)
                if (4 * (ox * ox + oy * oy + oz * oz) < 25) d = oz;
            const int c = col_of(ox, oy);
            if ((c < 0) != (d < 0) || (c >= 0 && col_at(c).d != d)) return false;
            n += (c >= 0);
        }
    return o0 == kOff && n == kCols;
}
static_assert(col_tables_agree(), "stencil column tables");

// Canonical index of lattice offset (ox, oy, oz), or -1 outside the stencil.
__host__ __device__ inline int canon_index(int ox, int oy, int oz)
{
    for (int c = 0; c < kCols; ++c)
    {
        Col k = col_at(c);
        if (k.ox == ox && k.oy == oy)
            return (oz >= -k.d && oz <= k.d) ? k.o0 + oz + k.d : -1;
    }
    return -1;
}

__host__ __device__ inline int category(int i, int n)
{
    return i == 0 ? 0 : (i == 1 ? 1 : (i == n - 1 ? 4 : (i == n - 2 ? 3 : 2)));
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// ---- coefficient tables (FP64 factorisation, once per cloud) ---------------
// Thread per (class, node): the assembly of kTransportNode3DT for the class's
// representative particle, then q = A^-1 e0 by the same Cholesky and
// substitutions, then C_m = w_m * phi_m . q for each neighbour m.
template <typename T, int P, bool BND>
__device__ void build_class(const DevCloud &c, int i, int lin, const int *lat, double dt,
                            double h, double tol, T *out, int S)
{
    double vx = vnode(c.vlo, c.dv, lin % c.Nv);
    double vy = vnode(c.vlo, c.dv, (lin / c.Nv) % c.Nv);
    double vz = vnode(c.vlo, c.dv, lin / (c.Nv * c.Nv));
    const int self = canon_index(0, 0, 0);
    if (BND)
    {
        double cn = wall_cn3(vx, vy, vz, c.wux[i], c.wuy[i], c.wuz[i], c.wnx[i], c.wny[i],
                             c.wnz[i]);
        if (cn >= 0.0) { out[(size_t)self * S] = T(1); return; }
    }
    const double xf = c.x[i] - (vx - c.mvx[i]) * dt;
    const double yf = c.y[i] - (vy - c.mvy[i]) * dt;
    const double zf = c.z[i] - (vz - c.mvz[i]) * dt;
    const double hh = h * h;
    const int nn = c.nneigh[i];
    const int *nb = c.neigh + (size_t)i * AB_MAX_NEIGH;
    double A[P * P];
#pragma unroll
    for (int q = 0; q < P * P; ++q) A[q] = 0.0;
    bool has = false;
    double sumW = 0.0;
    auto weight_basis = [&](int j, double *basis) -> double {
        const double dxn = c.x[j] - xf, dyn = c.y[j] - yf, dzn = c.z[j] - zf;
        const double r2 = dxn * dxn + dyn * dyn + dzn * dzn;
        const double q2 = r2 / hh;
        double w = 0.0;
        if (q2 < 1.0) { double t1 = 1.0 - q2; w = t1 * t1; }
        const double sx = dxn / h, sy = dyn / h, sz = dzn / h;
        basis[0] = 1.0; basis[1] = sx; basis[2] = sy; basis[3] = sz;
        if constexpr (P >= 10)
        {
            basis[4] = sx * sx; basis[5] = sy * sy; basis[6] = sz * sz;
            basis[7] = sx * sy; basis[8] = sx * sz; basis[9] = sy * sz;
        }
        return w;
    };
    for (int m = 0; m < nn; ++m)
    {
        double basis[P];
        double w = weight_basis(nb[m], basis);
        if (w < 1e-30) continue;
#pragma unroll
        for (int p1 = 0; p1 < P; ++p1)
        {
            const double wb = w * basis[p1];
#pragma unroll
            for (int p2 = p1; p2 < P; ++p2) A[p1 * P + p2] += wb * basis[p2];
        }
        sumW += w;
        has = true;
    }
    if (!has) { out[(size_t)self * S] = T(1); return; }
    double dmax = 0.0;
#pragma unroll
    for (int k = 0; k < P; ++k) dmax = fmax(dmax, A[k * P + k]);
    bool degenerate = (dmax <= 0.0);
#pragma unroll
    for (int k = 0; k < P; ++k)
    {
        double d = A[k * P + k];
#pragma unroll
        for (int j = 0; j < k; ++j) { double l = A[j * P + k]; d -= l * l; }
        if (!(d > tol * dmax)) degenerate = true;
        d = degenerate ? 1.0 : sqrt(d);
        A[k * P + k] = d;
#pragma unroll
        for (int m2 = k + 1; m2 < P; ++m2)
        {
            double v = A[k * P + m2];
#pragma unroll
            for (int j = 0; j < k; ++j) v -= A[j * P + k] * A[j * P + m2];
            A[k * P + m2] = v / d;
        }
    }
    double q[P];
#pragma unroll
    for (int k = 0; k < P; ++k)
    {
        double v = (k == 0) ? 1.0 : 0.0;
#pragma unroll
        for (int j = 0; j < k; ++j) v -= A[j * P + k] * q[j];
        q[k] = v / A[k * P + k];
    }
#pragma unroll
    for (int k = P - 1; k >= 0; --k)
    {
        double v = q[k];
#pragma unroll
        for (int j = k + 1; j < P; ++j) v -= A[k * P + j] * q[j];
        q[k] = v / A[k * P + k];
    }
    const int ix = lat[3 * i], iy = lat[3 * i + 1], iz = lat[3 * i + 2];
    for (int m = 0; m < nn; ++m)
    {
        int j = nb[m];
        double basis[P];
        double w = weight_basis(j, basis);
        if (w < 1e-30) continue;
        double cm;
        if (degenerate) cm = w / sumW;
        else
        {
            cm = 0.0;
#pragma unroll
            for (int p = 0; p < P; ++p) cm += basis[p] * q[p];
            cm *= w;
        }
        int o = canon_index(lat[3 * j] - ix, lat[3 * j + 1] - iy, lat[3 * j + 2] - iz);
        if (o >= 0) out[(size_t)o * S] = T(cm);
    }
}

template <typename T>
__global__ void k_build(DevCloud c, const int *__restrict__ rep, const int *__restrict__ lat,
                        double dt, double h, double tol, int order, T *__restrict__ coef, int S)
{
    const int cls = blockIdx.y;
    const int lin = blockIdx.x * blockDim.x + threadIdx.x;
    if (lin >= S) return;
    // layout [class][64-node chunk][81 offsets][64 nodes]: a wave's 81
    // coefficient rows are one contiguous 81 x 64 block (constant offsets)
    T *out = coef + ((size_t)cls * (S / kWave) + lin / kWave) * kOff * kWave + (lin % kWave);
    for (int o = 0; o < kOff; ++o) out[(size_t)o * kWave] = T(0);
    if (lin >= c.nodes) return;
    const int i = rep[cls];
    if (c.bnd[i]) build_class<T, 4, true>(c, i, lin, lat, dt, h, tol, out, kWave);
    else if (order >= 2) build_class<T, 10, false>(c, i, lin, lat, dt, h, tol, out, kWave);
    else build_class<T, 4, false>(c, i, lin, lat, dt, h, tol, out, kWave);
}

// ---- transport apply ---------------------------------------------------------
struct Lat { int nx, ny, nz, S, nchunk; };

// Addresses: every pointer below is wave-uniform and only the node index
// varies across lanes. Loads and stores go through raw buffer instructions
// whose descriptor is built in SGPRs from the uniform row pointer, with the
// lane's byte offset as the only VGPR operand (a plain global access would
// carry a 64-bit VGPR address per value in flight).
template <typename T>
__device__ __forceinline__ T bload(const T *row, int voff)
{
    __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc((void *)row, (short)0,
                                                                 0x7fffffff, 0x00020000);
    if constexpr (sizeof(T) == 4)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, 0, 0));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, 0, 0));
}

// A descriptor built once for a base and reused with a wave-uniform byte
// offset (SGPR soffset) per row: one descriptor per column and z block
// instead of one per load. Offsets stay below 2^31 bytes (a z block spans at
// most 13 lattice planes).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t brsrc(const void *base)
{
    return __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, 0x7fffffff, 0x00020000);
}

template <typename T>
__device__ __forceinline__ T bload_so(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    if constexpr (sizeof(T) == 4)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

template <typename T>
__device__ __forceinline__ void bstore(T *row, int voff, T v)
{
    __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc((void *)row, (short)0,
                                                                 0x7fffffff, 0x00020000);
    if constexpr (sizeof(T) == 4)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, 0, 0);
    else
    {
        typedef unsigned u2 __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, voff, 0, 0);
    }
}

template <typename T>
__device__ __forceinline__ void bstore_so(__amdgpu_buffer_rsrc_t r, int voff, int soff, T v)
{
    if constexpr (sizeof(T) == 4)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
    else
    {
        typedef unsigned u2 __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, voff, soff, 0);
    }
}

// The loads of column CC+1 are issued before the FMAs of column CC (two
// register buffers), so every column's memory latency overlaps the previous
// column's arithmetic. Each step ends in a register pin and a scheduling
// barrier: without them the optimiser sinks all the arithmetic below all
// the loads and spills the values.

// End planes in pairs: planes 0 and 1 (and nz-2, nz-1) read the same four
// lattice planes z_lo .. z_lo+3 of each stencil column, so one load of the
// column's values serves both outputs; the two classes' coefficients are
// loaded per offset. TOP selects the upper pair (z_lo = nz - 4).
template <bool TOP>
__host__ __device__ constexpr int pair_idx(int which, int dz)
{
    // index into the column's four planes for output `which` (0: lower
    // plane of the pair) and offset dz, with the lattice clamping
    if (!TOP)
    {
        const int z = which + dz;                  // planes 0, 1
        return z < 0 ? 0 : z;
    }
    const int z = 2 + which + dz;                  // relative to nz - 4
    return z > 3 ? 3 : z;
}

template <typename T, bool TOP, int CC>
__device__ __forceinline__ void pair_load(const T *__restrict__ g, __amdgpu_buffer_rsrc_t ra,
                                          __amdgpu_buffer_rsrc_t rb, const Lat &L, int ix, int iy,
                                          int zlo, int lin, T (&gv)[4], T (&ca)[5], T (&cb)[5])
{
    constexpr Col k = col_at(CC);
    const int xx = clampi(ix + k.ox, 0, L.nx - 1), yy = clampi(iy + k.oy, 0, L.ny - 1);
    const __amdgpu_buffer_rsrc_t rg =
        brsrc(g + ((size_t)xx + (size_t)L.nx * (yy + (size_t)L.ny * zlo)) * L.S);
    const int pb = L.nx * L.ny * L.S * (int)sizeof(T);
    // only the planes this column's offsets reach
    constexpr int lo = TOP ? 2 - k.d : 0, hi = TOP ? 3 : 1 + k.d;
#pragma clang loop unroll(full)
    for (int j = lo; j <= hi; ++j) gv[j] = bload_so<T>(rg, lin, j * pb);
#pragma clang loop unroll(full)
    for (int j = 0; j <= 2 * k.d; ++j)
    {
        ca[j] = bload_so<T>(ra, lin, (k.o0 + j) * kWave * (int)sizeof(T));
        cb[j] = bload_so<T>(rb, lin, (k.o0 + j) * kWave * (int)sizeof(T));
    }
}

template <typename T, bool TOP, int CC>
__device__ __forceinline__ void pair_fma(const T (&gv)[4], const T (&ca)[5], const T (&cb)[5],
                                         T &acc0, T &acc1)
{
    constexpr Col k = col_at(CC);
#pragma clang loop unroll(full)
    for (int j = 0; j <= 2 * k.d; ++j)
    {
        acc0 += ca[j] * gv[pair_idx<TOP>(0, j - k.d)];
        acc1 += cb[j] * gv[pair_idx<TOP>(1, j - k.d)];
    }
}

template <typename T, bool TOP, int CC>
__device__ __forceinline__ void pair_step(const T *__restrict__ g, __amdgpu_buffer_rsrc_t ra,
                                          __amdgpu_buffer_rsrc_t rb, const Lat &L, int ix, int iy,
                                          int zlo, int lin, T (&g0)[4], T (&a0)[5], T (&b0)[5],
                                          T (&g1)[4], T (&a1)[5], T (&b1)[5], T &acc0, T &acc1)
{
    if constexpr (CC + 1 < kCols)
    {
        if constexpr ((CC & 1) == 0) pair_load<T, TOP, CC + 1>(g, ra, rb, L, ix, iy, zlo, lin, g1, a1, b1);
        else pair_load<T, TOP, CC + 1>(g, ra, rb, L, ix, iy, zlo, lin, g0, a0, b0);
    }
    if constexpr ((CC & 1) == 0) pair_fma<T, TOP, CC>(g0, a0, b0, acc0, acc1);
    else pair_fma<T, TOP, CC>(g1, a1, b1, acc0, acc1);
    asm volatile("" : "+v"(acc0), "+v"(acc1)::"memory");
    __builtin_amdgcn_sched_barrier(0);
}

// Outputs of the pair's two planes (classes cfa, cfb).
template <typename T, bool TOP, int... CC>
__device__ __forceinline__ void apply_pair(const T *__restrict__ g, const T *__restrict__ cfa,
                                           const T *__restrict__ cfb, const Lat &L, int ix, int iy,
                                           int lin, T &acc0, T &acc1,
                                           std::integer_sequence<int, CC...>)
{
    const int zlo = TOP ? L.nz - 4 : 0;
    const __amdgpu_buffer_rsrc_t ra = brsrc(cfa), rb = brsrc(cfb);
    T g0[4], a0[5], b0[5], g1[4], a1[5], b1[5];
    acc0 = T(0);
    acc1 = T(0);
    pair_load<T, TOP, 0>(g, ra, rb, L, ix, iy, zlo, lin, g0, a0, b0);
    (pair_step<T, TOP, CC>(g, ra, rb, L, ix, iy, zlo, lin, g0, a0, b0, g1, a1, b1, acc0, acc1), ...);
}

// Interior: the values of column CC for outputs iz0 .. iz0+K-1.
template <typename T, int K, int CC>
__device__ __forceinline__ void col_load(const T *__restrict__ g, const Lat &L, int ix, int iy,
                                         int iz0, int lin, size_t zstride, T (&v)[K + 4])
{
    constexpr Col kc = col_at(CC);
    constexpr int NV = K + 2 * kc.d;
    const int xx = clampi(ix + kc.ox, 0, L.nx - 1), yy = clampi(iy + kc.oy, 0, L.ny - 1);
    // the column from plane iz0 - d (>= 0 in the interior march)
    const int z0 = iz0 - kc.d;
    const __amdgpu_buffer_rsrc_t rg =
        brsrc(g + ((size_t)xx + (size_t)L.nx * yy) * L.S + (size_t)z0 * zstride);
    const int pb = (int)(zstride * sizeof(T));
    const int zmax = L.nz - 1 - z0;
#pragma clang loop unroll(full)
    for (int j = 0; j < NV; ++j) v[j] = bload_so<T>(rg, lin, (j < zmax ? j : zmax) * pb);
}

template <typename T, int K, int CC, typename CS>
__device__ __forceinline__ void col_step(const T *__restrict__ g, const CS &cr,
                                         const Lat &L, int ix, int iy, int iz0, int lin,
                                         size_t zstride, T (&va)[K + 4], T (&vb)[K + 4],
                                         T (&acc)[K])
{
    constexpr Col kc = col_at(CC);
    if constexpr (CC + 1 < kCols)
    {
        if constexpr ((CC & 1) == 0) col_load<T, K, CC + 1>(g, L, ix, iy, iz0, lin, zstride, vb);
        else col_load<T, K, CC + 1>(g, L, ix, iy, iz0, lin, zstride, va);
    }
#pragma clang loop unroll(full)
    for (int k = 0; k < K; ++k)
#pragma clang loop unroll(full)
        for (int dz = 0; dz <= 2 * kc.d; ++dz)
        {
            if constexpr ((CC & 1) == 0) acc[k] += cr[kc.o0 + dz] * va[k + dz];
            else acc[k] += cr[kc.o0 + dz] * vb[k + dz];
        }
#pragma clang loop unroll(full)
    for (int k = 0; k < K; ++k) asm volatile("" : "+v"(acc[k])::"memory");
    __builtin_amdgcn_sched_barrier(0);
}

template <typename T, int K, typename CS, int... CC>
__device__ __forceinline__ void apply_columns(const T *__restrict__ g, const CS &cr,
                                              const Lat &L, int ix, int iy, int iz0, int lin,
                                              size_t zstride, T (&acc)[K],
                                              std::integer_sequence<int, CC...>)
{
    T va[K + 4], vb[K + 4];
    col_load<T, K, 0>(g, L, ix, iy, iz0, lin, zstride, va);
    (col_step<T, K, CC>(g, cr, L, ix, iy, iz0, lin, zstride, va, vb, acc), ...);
}

// The lattice column (ix, iy) of a wave in an "all columns" or "edge
// columns only" enumeration (edges: ix or iy in {0, 1, n-2, n-1}).
__device__ __forceinline__ void column_of(const Lat &L, int e, bool edges, int &ix, int &iy)
{
    if (!edges) { ix = e % L.nx; iy = e / L.nx; }
    else if (e < 4 * L.nx)
    {
        const int r = e / L.nx;                  // rows 0, 1, ny-2, ny-1
        iy = r < 2 ? r : L.ny - 4 + r;
        ix = e % L.nx;
    }
    else
    {
        const int q = e - 4 * L.nx;              // columns 0, 1, nx-2, nx-1 of rows 2 .. ny-3
        iy = 2 + q / 4;
        ix = (q % 4) < 2 ? q % 4 : L.nx - 4 + q % 4;
    }
}

// One wave = lattice column (ix, iy) x one 64-node chunk: the end-plane
// pairs, then the interior march in z blocks of K outputs.
template <typename T, int K>
__device__ __forceinline__ void column_wave(const Lat &L, const T *__restrict__ g, T *__restrict__ gt,
                                            const T *__restrict__ coef, int chunk, int ix, int iy,
                                            int voff)
{
    const int col = ix + L.nx * iy;
    const int cx = category(ix, L.nx), cy = category(iy, L.ny);
    const size_t plane = (size_t)L.nx * L.ny;
    // this wave's chunk of nodes: offset every base pointer once
    const size_t c0 = (size_t)chunk * kWave;
    const T *gw = g + c0;
    T *gtw = gt + c0;

    // The four end planes, in pairs (0, 1) and (nz-2, nz-1): their own
    // classes, coefficients loaded per offset, column values shared.
    auto end_pair = [&](auto top) {
        constexpr bool TOP = decltype(top)::value;
        const int iz = TOP ? L.nz - 2 : 0;
        const T *cfa = coef + ((size_t)(cx + 5 * (cy + 5 * category(iz, L.nz))) * L.nchunk + chunk) * kOff * kWave;
        const T *cfb = coef + ((size_t)(cx + 5 * (cy + 5 * category(iz + 1, L.nz))) * L.nchunk + chunk) * kOff * kWave;
        T a0, a1;
        apply_pair<T, TOP>(gw, cfa, cfb, L, ix, iy, voff, a0, a1,
                           std::make_integer_sequence<int, kCols>{});
        bstore(gtw + ((size_t)col + plane * iz) * L.S, voff, a0);
        bstore(gtw + ((size_t)col + plane * (iz + 1)) * L.S, voff, a1);
        __builtin_amdgcn_sched_barrier(0);
    };
    end_pair(std::false_type{});
    end_pair(std::true_type{});

    // Interior planes iz = 2 .. nz-3 share one class: coefficients in
    // registers
    const T *cf = coef + ((size_t)(cx + 5 * (cy + 5 * 2)) * L.nchunk + chunk) * kOff * kWave;
    auto march = [&](const auto &cr) {
        for (int iz0 = 2; iz0 <= L.nz - 3; iz0 += K)
        {
            T acc[K];
#pragma unroll
            for (int k = 0; k < K; ++k) acc[k] = T(0);
            apply_columns<T, K>(gw, cr, L, ix, iy, iz0, voff, plane * L.S, acc,
                                std::make_integer_sequence<int, kCols>{});
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (iz0 + k <= L.nz - 3)
                    bstore(gtw + ((size_t)col + plane * (iz0 + k)) * L.S, voff, acc[k]);
        }
    };
    T cr[kOff];
    const __amdgpu_buffer_rsrc_t rc = brsrc(cf);
#pragma unroll
    for (int o = 0; o < kOff; ++o) cr[o] = bload_so<T>(rc, voff, o * kWave * (int)sizeof(T));
    march(cr);
}

template <typename T, int K, int W>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W)))
void k_apply(Lat L, const T *__restrict__ g, T *__restrict__ gt, const T *__restrict__ coef,
             const double *__restrict__ scal)
{
    (void)scal;
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    // wave-uniform: keep it (and every address part derived from it) in SGPRs
    const int wv = __builtin_amdgcn_readfirstlane(b * (int)(blockDim.x / kWave) + (int)threadIdx.x / kWave);
    const int voff = (threadIdx.x & (kWave - 1)) * (int)sizeof(T);   // the lane's byte offset
    const int ncol = L.nx * L.ny;
    if (wv >= ncol * L.nchunk) return;
    // chunk-major: the waves of one XCD (a contiguous range after the remap)
    // share one or two 64-node chunks, so the chunk's coefficient rows and
    // the neighbouring columns' values stay in that XCD's L2
    int ix, iy;
    column_of(L, wv % ncol, false, ix, iy);
    column_wave<T, K>(L, g, gt, coef, wv / ncol, ix, iy, voff);
}

// ---- blocks of interior columns ------------------------------------------
// Columns with ix, iy in [2, n-3] all belong to the classes (2, 2, cz): one
// wave can compute a TX x TY block of them with one set of coefficients.
// Stencil columns of neighbouring outputs overlap, so the block loads each
// column of the union of its outputs' stencils once per z block and feeds
// every output whose stencil contains it: for 2 x 2 outputs, 32 union
// columns instead of 4 x 21.


struct UCol { int ox, oy, d; };

template <int TX, int TY>
struct Blk
{
    static constexpr int NT = TX * TY;
    // union columns relative to output (0, 0), row-major; d: the deepest
    // z reach any output needs from it
    struct Tab { UCol u[(TX + 4) * (TY + 4)]; int n; };
    static constexpr Tab make()
    {
        Tab r{};
        int n = 0;
        for (int oy = -2; oy <= TY + 1; ++oy)
            for (int ox = -2; ox <= TX + 1; ++ox)
            {
                int d = 0;
                for (int ty = 0; ty < TY; ++ty)
                    for (int tx = 0; tx < TX; ++tx)
                    {
                        const int c = col_of(ox - tx, oy - ty);
                        if (c >= 0 && col_at(c).d > d) d = col_at(c).d;
                    }
                if (d) r.u[n++] = UCol{ox, oy, d};
            }
        r.n = n;
        return r;
    }
    static constexpr Tab tab = make();
    static constexpr int NU = tab.n;
    static constexpr UCol ucol(int u) { return tab.u[u]; }
};

// Addressing of a block's z block: one buffer descriptor at the lowest union
// column (bx - 2, by - 2) and plane iz0 - 2; the lane's node at the vector
// offset, and in the scalar offset union column U (cofs(U), cheap to
// recompute) plus plane iz0 - 2 + zz (zof[zz], clamped to the top plane,
// per z block): no descriptor per column, no per-load vector arithmetic.
template <typename T, int K>
struct BlkAddr
{
    __amdgpu_buffer_rsrc_t r;
    int lin;                  // the lane's byte offset
    int col, row;             // bytes between neighbouring columns in x and in y
    int zof[K + 4];
    int bx, by, nx, ny;       // block origin and lattice extent (edge blocks clamp)
    template <int U, int TX, int TY, bool CLAMP = false>
    __device__ __forceinline__ int cofs() const
    {
        constexpr UCol uc = Blk<TX, TY>::ucol(U);
        if constexpr (!CLAMP) return (uc.ox + 2) * col + (uc.oy + 2) * row;
        // a union column outside the lattice reads the lattice's last column,
        // as the stencil of an edge particle does
        return (clampi(bx + uc.ox, 0, nx - 1) - bx + 2) * col +
               (clampi(by + uc.oy, 0, ny - 1) - by + 2) * row;
    }
};

// Values of union column U for outputs iz0 .. iz0+K-1 (no x/y clamping: the
// block lies two columns inside the lattice).
template <typename T, int TX, int TY, int K, int U, bool CLAMP = false>
__device__ __forceinline__ void blk_load(const BlkAddr<T, K> &a, T (&v)[K + 4])
{
    constexpr UCol uc = Blk<TX, TY>::ucol(U);
    constexpr int NV = K + 2 * uc.d;
    const int so = a.template cofs<U, TX, TY, CLAMP>();
#pragma clang loop unroll(full)
    for (int j = 0; j < NV; ++j) v[j] = bload_so<T>(a.r, a.lin, so + a.zof[j + 2 - uc.d]);
}

template <typename T, int TX, int TY, int K, int U, typename CS, int... TI>
__device__ __forceinline__ void blk_fma(const CS &cr, const T (&v)[K + 4],
                                        T (&acc)[TX * TY][K], std::integer_sequence<int, TI...>)
{
    constexpr UCol uc = Blk<TX, TY>::ucol(U);
    auto one = [&](auto tic) {
        constexpr int t = decltype(tic)::value;
        constexpr int c = col_of(uc.ox - t % TX, uc.oy - t / TX);
        if constexpr (c >= 0)
        {
            constexpr int dc = col_at(c).d, o0 = col_at(c).o0;
#pragma clang loop unroll(full)
            for (int k = 0; k < K; ++k)
#pragma clang loop unroll(full)
                for (int dz = -dc; dz <= dc; ++dz)
                    acc[t][k] += cr[o0 + dz + dc] * v[k + dz + uc.d];
        }
    };
    (one(std::integral_constant<int, TI>{}), ...);
}

// R register buffers of union-column values: the loads run R - 1 columns
// ahead of the arithmetic (R = 2: one column ahead).
template <typename T, int TX, int TY, int K, bool CLAMP, int R, int U, typename CS>
__device__ __forceinline__ void blk_step(const BlkAddr<T, K> &a, const CS &cr, T (&v)[R][K + 4],
                                         T (&acc)[TX * TY][K])
{
    constexpr int ahead = U + R - 1;
    if constexpr (ahead < Blk<TX, TY>::NU) blk_load<T, TX, TY, K, ahead, CLAMP>(a, v[ahead % R]);
    blk_fma<T, TX, TY, K, U>(cr, v[U % R], acc, std::make_integer_sequence<int, TX * TY>{});
#pragma clang loop unroll(full)
    for (int t = 0; t < TX * TY; ++t)
#pragma clang loop unroll(full)
        for (int k = 0; k < K; ++k) asm volatile("" : "+v"(acc[t][k])::"memory");
    __builtin_amdgcn_sched_barrier(0);
}

template <typename T, int TX, int TY, int K, bool CLAMP, int R, int... U>
__device__ __forceinline__ void blk_prime(const BlkAddr<T, K> &a, T (&v)[R][K + 4],
                                          std::integer_sequence<int, U...>)
{
    auto one = [&](auto uic) {
        constexpr int u = decltype(uic)::value;
        if constexpr (u < Blk<TX, TY>::NU) blk_load<T, TX, TY, K, u, CLAMP>(a, v[u]);
    };
    (one(std::integral_constant<int, U>{}), ...);
}

template <typename T, int TX, int TY, int K, bool CLAMP, int R, typename CS, int... U>
__device__ __forceinline__ void blk_columns(const T *__restrict__ g, const CS &cr,
                                            const Lat &L, int bx, int by, int iz0, int lin,
                                            size_t zstride, T (&acc)[TX * TY][K],
                                            std::integer_sequence<int, U...>)
{
    BlkAddr<T, K> a;
    const int zlo = iz0 - 2;
    // the descriptor's base (bx - 2, by - 2) may lie before the array for an
    // edge block; every address it serves is inside (offsets are >= 0)
    a.r = brsrc(g + ((ptrdiff_t)(bx - 2) + (ptrdiff_t)L.nx * (by - 2)) * L.S + (size_t)zlo * zstride);
    a.lin = lin;
    a.col = L.S * (int)sizeof(T);
    a.row = L.nx * L.S * (int)sizeof(T);
    a.bx = bx; a.by = by; a.nx = L.nx; a.ny = L.ny;
    const int pb = (int)(zstride * sizeof(T)), ztop = L.nz - 1 - zlo;
#pragma clang loop unroll(full)
    for (int zz = 0; zz < K + 4; ++zz) a.zof[zz] = (zz < ztop ? zz : ztop) * pb;
    T v[R][K + 4];
    blk_prime<T, TX, TY, K, CLAMP, R>(a, v, std::make_integer_sequence<int, R - 1>{});
    (blk_step<T, TX, TY, K, CLAMP, R, U>(a, cr, v, acc), ...);
}

// End planes of a block: planes 0 and 1 (TOP: nz-1 and nz-2) have classes
// of their own. The coefficients of their taps that stay inside the lattice
// (the table holds zeros for the others) are loaded into registers in one
// batch, 51 for the plane on the wall (A) and 72 for the next plane (B);
// then the union columns are walked as in the z-march, each column's four
// end planes loaded one column ahead and added into every output whose
// stencil holds the column.
template <bool TOP>
struct EndTaps
{
    static constexpr int lo_a(int d) { return TOP ? -d : 0; }
    static constexpr int hi_a(int d) { return TOP ? 0 : d; }
    static constexpr int lo_b(int d) { return TOP ? -d : -1; }
    static constexpr int hi_b(int d) { return TOP ? 1 : d; }
    // window index (the end's four planes, ascending) of tap dz
    static constexpr int va(int dz) { return TOP ? 3 + dz : dz; }
    static constexpr int vb(int dz) { return TOP ? 2 + dz : 1 + dz; }
    static constexpr int base_a(int c)
    {
        int n = 0;
        for (int k = 0; k < c; ++k) n += hi_a(col_at(k).d) - lo_a(col_at(k).d) + 1;
        return n;
    }
    static constexpr int base_b(int c)
    {
        int n = 0;
        for (int k = 0; k < c; ++k) n += hi_b(col_at(k).d) - lo_b(col_at(k).d) + 1;
        return n;
    }
    static constexpr int NA = base_a(kCols), NB = base_b(kCols);
};

template <typename T, bool TOP, int... C>
__device__ __forceinline__ void end_taps(__amdgpu_buffer_rsrc_t rw, __amdgpu_buffer_rsrc_t rn,
                                         int lin, T (&ca)[EndTaps<TOP>::NA],
                                         T (&cb)[EndTaps<TOP>::NB], std::integer_sequence<int, C...>)
{
    using E = EndTaps<TOP>;
    auto one = [&](auto cic) {
        constexpr int c = decltype(cic)::value;
        constexpr Col k = col_at(c);
#pragma clang loop unroll(full)
        for (int dz = E::lo_a(k.d); dz <= E::hi_a(k.d); ++dz)
            ca[E::base_a(c) + dz - E::lo_a(k.d)] =
                bload_so<T>(rw, lin, (k.o0 + dz + k.d) * kWave * (int)sizeof(T));
#pragma clang loop unroll(full)
        for (int dz = E::lo_b(k.d); dz <= E::hi_b(k.d); ++dz)
            cb[E::base_b(c) + dz - E::lo_b(k.d)] =
                bload_so<T>(rn, lin, (k.o0 + dz + k.d) * kWave * (int)sizeof(T));
    };
    (one(std::integral_constant<int, C>{}), ...);
}

// The end planes of union column U that any output's taps reach.
template <typename T, int TX, int TY, bool TOP, bool CLAMP, int U>
__device__ __forceinline__ void end_load(const BlkAddr<T, 1> &a, int pb, T (&v)[4])
{
    constexpr UCol uc = Blk<TX, TY>::ucol(U);
    constexpr int lo = TOP ? 2 - uc.d : 0, hi = TOP ? 3 : 1 + uc.d;
    const int so = a.template cofs<U, TX, TY, CLAMP>();
#pragma clang loop unroll(full)
    for (int k = lo; k <= hi; ++k) v[k] = bload_so<T>(a.r, a.lin, so + k * pb);
}

template <typename T, int TX, int TY, bool TOP, int U, int... TI>
__device__ __forceinline__ void end_fma(const T (&ca)[EndTaps<TOP>::NA],
                                        const T (&cb)[EndTaps<TOP>::NB], const T (&v)[4],
                                        T (&aw)[TX * TY], T (&an)[TX * TY],
                                        std::integer_sequence<int, TI...>)
{
    using E = EndTaps<TOP>;
    constexpr UCol uc = Blk<TX, TY>::ucol(U);
    auto one = [&](auto tic) {
        constexpr int t = decltype(tic)::value;
        constexpr int c = col_of(uc.ox - t % TX, uc.oy - t / TX);
        if constexpr (c >= 0)
        {
            constexpr int d = col_at(c).d;
#pragma clang loop unroll(full)
            for (int dz = E::lo_a(d); dz <= E::hi_a(d); ++dz)
                aw[t] += ca[E::base_a(c) + dz - E::lo_a(d)] * v[E::va(dz)];
#pragma clang loop unroll(full)
            for (int dz = E::lo_b(d); dz <= E::hi_b(d); ++dz)
                an[t] += cb[E::base_b(c) + dz - E::lo_b(d)] * v[E::vb(dz)];
        }
    };
    (one(std::integral_constant<int, TI>{}), ...);
}

// Ring of column buffers: the loads run ring - 1 columns ahead of the
// arithmetic, which is only a few FMAs per column at the ends (blocks of six
// or more outputs do more per column and have fewer registers to spare).
template <int NT>
constexpr int end_ring() { return NT >= 6 ? 4 : 7; }

template <typename T, int TX, int TY, bool TOP, bool CLAMP, int U>
__device__ __forceinline__ void end_step(const BlkAddr<T, 1> &a, int pb,
                                         const T (&ca)[EndTaps<TOP>::NA],
                                         const T (&cb)[EndTaps<TOP>::NB],
                                         T (&v)[end_ring<TX * TY>()][4], T (&aw)[TX * TY],
                                         T (&an)[TX * TY])
{
    constexpr int RG = end_ring<TX * TY>(), ahead = U + RG - 1;
    if constexpr (ahead < Blk<TX, TY>::NU)
        end_load<T, TX, TY, TOP, CLAMP, ahead>(a, pb, v[ahead % RG]);
    end_fma<T, TX, TY, TOP, U>(ca, cb, v[U % RG], aw, an,
                               std::make_integer_sequence<int, TX * TY>{});
#pragma clang loop unroll(full)
    for (int t = 0; t < TX * TY; ++t) asm volatile("" : "+v"(aw[t]), "+v"(an[t])::"memory");
    __builtin_amdgcn_sched_barrier(0);
}

template <typename T, int TX, int TY, bool TOP, bool CLAMP, int... U>
__device__ __forceinline__ void end_prime(const BlkAddr<T, 1> &a, int pb,
                                          T (&v)[end_ring<TX * TY>()][4],
                                          std::integer_sequence<int, U...>)
{
    auto one = [&](auto uic) {
        constexpr int u = decltype(uic)::value;
        if constexpr (u < Blk<TX, TY>::NU) end_load<T, TX, TY, TOP, CLAMP, u>(a, pb, v[u]);
    };
    (one(std::integral_constant<int, U>{}), ...);
}

// aw: the plane on the wall (0, or nz-1 for TOP), an: the next one.
template <typename T, int TX, int TY, bool TOP, bool CLAMP, int... U>
__device__ __forceinline__ void blk_ends(const T *__restrict__ g, const T *__restrict__ cfw,
                                         const T *__restrict__ cfn, const Lat &L, int bx, int by,
                                         int lin, T (&aw)[TX * TY], T (&an)[TX * TY],
                                         std::integer_sequence<int, U...>)
{
    T ca[EndTaps<TOP>::NA], cb[EndTaps<TOP>::NB];
    end_taps<T, TOP>(brsrc(cfw), brsrc(cfn), lin, ca, cb, std::make_integer_sequence<int, kCols>{});
    const int zlo = TOP ? L.nz - 4 : 0;
    const int pb = L.nx * L.ny * L.S * (int)sizeof(T);
    BlkAddr<T, 1> a;
    a.r = brsrc(g + ((ptrdiff_t)(bx - 2) + (ptrdiff_t)L.nx * (by - 2)) * L.S +
                (size_t)zlo * L.nx * L.ny * L.S);
    a.lin = lin;
    a.col = L.S * (int)sizeof(T);
    a.row = L.nx * L.S * (int)sizeof(T);
    a.bx = bx; a.by = by; a.nx = L.nx; a.ny = L.ny;
#pragma clang loop unroll(full)
    for (int t = 0; t < TX * TY; ++t) { aw[t] = T(0); an[t] = T(0); }
    constexpr int RG = end_ring<TX * TY>();
    T v[RG][4];
    end_prime<T, TX, TY, TOP, CLAMP>(a, pb, v, std::make_integer_sequence<int, RG - 1>{});
    (end_step<T, TX, TY, TOP, CLAMP, U>(a, pb, ca, cb, v, aw, an), ...);
}

// One wave = a TX x TY block of columns with origin (bx, by), all of the
// x/y class (cx, cy), x one 64-node chunk. Interior blocks (cx = cy = 2)
// lie two columns inside the lattice; edge blocks (CLAMP) run along an edge
// of the lattice, their stencil columns clamped to it.
template <typename T, int TX, int TY, int K, bool CLAMP, int R = 2, int ZB = 1>
__device__ __forceinline__ void block_at(const Lat &L, const T *__restrict__ g, T *__restrict__ gt,
                                         const T *__restrict__ coef, int chunk, int bx, int by,
                                         int cx, int cy, int slab, int voff)
{
    const size_t plane = (size_t)L.nx * L.ny;
    const size_t c0 = (size_t)chunk * kWave;
    const T *gw = g + c0;
    T *gtw = gt + c0;
    auto cls_rows = [&](int cz) {
        return coef + ((size_t)(cx + 5 * (cy + 5 * cz)) * L.nchunk + chunk) * kOff * kWave;
    };
    auto store = [&](int t, int iz, T v) {
        bstore(gtw + ((size_t)(bx + t % TX) + (size_t)L.nx * (by + t / TX) + plane * iz) * L.S, voff, v);
    };
    auto end_pair = [&](auto top) {
        constexpr bool TOP = decltype(top)::value;
        const int iw = TOP ? L.nz - 1 : 0, in = TOP ? L.nz - 2 : 1;
        T aw[TX * TY], an[TX * TY];
        blk_ends<T, TX, TY, TOP, CLAMP>(gw, cls_rows(category(iw, L.nz)),
                                        cls_rows(category(in, L.nz)), L, bx, by, voff, aw, an,
                                        std::make_integer_sequence<int, Blk<TX, TY>::NU>{});
#pragma unroll
        for (int t = 0; t < TX * TY; ++t) { store(t, iw, aw[t]); store(t, in, an[t]); }
        __builtin_amdgcn_sched_barrier(0);
    };
    if (slab < 0)
    {
        end_pair(std::false_type{});
        end_pair(std::true_type{});
        if (slab == -1) return;
    }

    // slab k >= 0: the ZB z blocks from plane 2 + k ZB K
    const int zb = slab < 0 ? 2 : 2 + slab * ZB * K;
    const int ze = slab < 0 ? L.nz - 3 : min(zb + ZB * K - 1, L.nz - 3);
    auto march = [&](const auto &cr) {
        for (int iz0 = zb; iz0 <= ze; iz0 += K)
        {
            T acc[TX * TY][K];
#pragma unroll
            for (int t = 0; t < TX * TY; ++t)
#pragma unroll
                for (int k = 0; k < K; ++k) acc[t][k] = T(0);
            blk_columns<T, TX, TY, K, CLAMP, R>(gw, cr, L, bx, by, iz0, voff, plane * L.S, acc,
                                             std::make_integer_sequence<int, Blk<TX, TY>::NU>{});
#pragma unroll
            for (int t = 0; t < TX * TY; ++t)
#pragma unroll
                for (int k = 0; k < K; ++k)
                    if (iz0 + k <= L.nz - 3) store(t, iz0 + k, acc[t][k]);
        }
    };
    T cr[kOff];
    const __amdgpu_buffer_rsrc_t rc = brsrc(cls_rows(2));
#pragma unroll
    for (int o = 0; o < kOff; ++o) cr[o] = bload_so<T>(rc, voff, o * kWave * (int)sizeof(T));
    march(cr);
}

// Interior block blk (row-major over the interior). Requires (nx - 4) % TX
// == 0 and (ny - 4) % TY == 0.
template <typename T, int TX, int TY, int K, int R = 2, int ZB = 1>
__device__ __forceinline__ void block_wave(const Lat &L, const T *__restrict__ g, T *__restrict__ gt,
                                           const T *__restrict__ coef, int chunk, int blk, int slab,
                                           int voff)
{
    const int nbx = (L.nx - 4) / TX;
    block_at<T, TX, TY, K, false, R, ZB>(L, g, gt, coef, chunk, 2 + (blk % nbx) * TX,
                                  2 + (blk / nbx) * TY, 2, 2, slab, voff);
}

// The whole transport in one launch: per chunk, the edge columns (z blocks
// of KE), then the interior blocks (TX x TY columns per wave). One
// launch keeps the edge waves from running as a second, short and
// latency-bound grid after the blocks. Where the interior is a multiple of 4
// columns wide, the edge columns along an edge share a class and run as
// blocks too: 1 x 4 along the x edges (ix = 0, 1, nx-2, nx-1), 4 x 1 along
// the y edges, and only the 16 corner columns one per wave.
//
// SLAB: a block wave covers one z block of K planes instead of the whole
// column, and the waves of a chunk run slab by slab (first every block's end
// planes, then every block's first z block, ...), so the waves in flight on
// an XCD share one slab of the chunk's columns in L2 instead of spanning all
// of its planes.
template <int K>
__host__ __device__ constexpr int zblocks(int nz) { return (nz - 4 + K - 1) / K; }

__host__ __device__ inline bool edge_blocks(int nx, int ny)
{
    return ((nx - 4) & 3) == 0 && ((ny - 4) & 3) == 0;
}

// edge waves per chunk: blocks along the four x edges and four y edges plus
// the corners, or one wave per edge column
__host__ __device__ inline int edge_items(int nx, int ny, int slabs = 1)
{
    return edge_blocks(nx, ny) ? ((ny - 4) + (nx - 4)) * slabs + 16 : nx * ny - (nx - 4) * (ny - 4);
}

// Edge blocks march in z blocks of at most kEdgeK (12 would raise the
// kernel's register count above three waves per SIMD).
constexpr int kEdgeK = 8;

// Items per edge block: the whole column, or in the SLAB order its end
// planes and each of its z blocks (of kEdgeK planes) as items of their own.
__host__ __device__ inline int edge_slabs(bool slab, int nz)
{
    return slab ? zblocks<kEdgeK>(nz) + 1 : 1;
}

// R: the register buffers of the z-march (loads R - 1 union columns ahead).
// SLAB: 0 whole columns; n > 0 items of n z blocks of the interior blocks.
template <typename T, int TX, int TY, int K, int KE, int SLAB, int R = 2>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void k_apply_mix(Lat L, const T *__restrict__ g, T *__restrict__ gt, const T *__restrict__ coef)
{
    const int nblk = ((L.nx - 4) / TX) * ((L.ny - 4) / TY);
    const int nitem = SLAB ? nblk * (zblocks<SLAB * K>(L.nz) + 1) : nblk;
    const int eslabs = edge_slabs(SLAB != 0, L.nz);
    const int per = nitem + edge_items(L.nx, L.ny, eslabs);
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int wv = __builtin_amdgcn_readfirstlane(b * (int)(blockDim.x / kWave) + (int)threadIdx.x / kWave);
    const int voff = (threadIdx.x & (kWave - 1)) * (int)sizeof(T);
    if (wv >= per * L.nchunk) return;
    const int chunk = wv / per;                             // chunk-major
    // within a chunk the edge items first: they march whole columns and are
    // the longest, so started last they would set the launch's tail
    const int ne = per - nitem, r0 = wv % per;
    const int r = r0 < ne ? nitem + r0 : r0 - ne;
    if (r < nitem)
    {
        if constexpr (SLAB) block_wave<T, TX, TY, K, R, SLAB>(L, g, gt, coef, chunk, r % nblk, r / nblk - 1, voff);
        else block_wave<T, TX, TY, K, R>(L, g, gt, coef, chunk, r, -2, voff);
        return;
    }
    int e = r - nitem;
    constexpr int KB = KE > kEdgeK ? kEdgeK : KE;
    static_assert(!SLAB || KB == kEdgeK, "edge slabs are counted in z blocks of kEdgeK");
    if (edge_blocks(L.nx, L.ny))
    {
        const int qy = (L.ny - 4) / 4, qx = (L.nx - 4) / 4;
        const int nblke = 4 * (qx + qy);
        if (e < nblke * eslabs)
        {
            // edge block eb, slab sl (-2: the whole column; -1: the end
            // planes; k >= 0: z block k)
            const int eb = e / eslabs, sl = SLAB ? e % eslabs - 1 : -2;
            if (eb < 4 * qy)                                 // x edges: 1 x 4 along y
            {
                const int side = eb / qy, ix = side < 2 ? side : L.nx - 4 + side;
                block_at<T, 1, 4, KB, true, R>(L, g, gt, coef, chunk, ix, 2 + 4 * (eb % qy),
                                                  category(ix, L.nx), 2, sl, voff);
                return;
            }
            const int f = eb - 4 * qy;                       // y edges: 4 x 1 along x
            const int side = f / qx, iy = side < 2 ? side : L.ny - 4 + side;
            block_at<T, 4, 1, KB, true, R>(L, g, gt, coef, chunk, 2 + 4 * (f % qx), iy, 2,
                                              category(iy, L.ny), sl, voff);
            return;
        }
        e -= nblke * eslabs;                                 // the corners
        const int a = e & 3, c = e >> 2;
        const int ix = a < 2 ? a : L.nx - 4 + a, iy = c < 2 ? c : L.ny - 4 + c;
        column_wave<T, KE>(L, g, gt, coef, chunk, ix, iy, voff);
        return;
    }
    int ix, iy;
    column_of(L, e, true, ix, iy);
    column_wave<T, KE>(L, g, gt, coef, chunk, ix, iy, voff);
}

// ---- per-particle limiters, wall, moments, collision -------------------------
// Storage: g = L * (M0 + d), d in the stored type, M0 the equilibrium
// Maxwellian of the reference state (rho_ref, u = 0, T_ref) in FP64 and L the
// step's pending rescale. A low-Mach flow is a small deviation from M0, so
// FP32 keeps ~7 digits of the deviation itself instead of ~7 digits of g
// (the velocity moment of a 1 m/s flow is 3e-3 of the thermal moments).
// MLS rows sum to 1, so the transport of M0 is M0 and only d is applied.
// Everything below works in L-units (values / L); partials are written in
// true units.
//
// Partial sums written per particle (FP64, true units):
//   0  mass after limiters and commit (before the first rescale)
//   1  mass after wall / collision, before the rescales
//   2..4  sum v_a h dv^3 (interior)   | 2 = rho_w before the rescales (boundary)
//   5  sum |v|^2 h dv^3 (interior)
//   6  limiter bits (1 positivity, 2 NAD)
constexpr int kPart = 8;

struct Ref
{
    const double *m0;       // M0 per node, S entries
    const double *m0xy;     // its (v_x, v_y) factor (m0c e_iv) e_jv per pair p = iv + Nv jv
    const double *m0x;      // its 1-D factors e_l = exp(-v_l^2 / 2RT), Nv entries:
    double m0c;             //   M0(iv, jv, kv) = ((m0c e_iv) e_jv) e_kv exactly
    double m0sum;           // sum over the nodes of M0
    double mom[5];          // sum M0 dv^3, sum v_a M0 dv^3, sum |v|^2 M0 dv^3
    double rho, T;          // reference state of M0
};

// A particle's row visited pair by pair: the lane takes two (iv, jv) pairs
// of the velocity plane at a time (slots s = 0, 1: p = 128 c + 64 s + lane
// for chunk pair c) and, for each, the Nv nodes k = p + P kv (P = Nv^2).
// v_x, v_y (and anything that depends only on them) are then fixed per pair
// and v_z is the same on every lane, so a moment sum_k psi(v) d costs three
// accumulators per node (sum d, sum kv d, sum kv^2 d), combined with v_x,
// v_y once per pair; what depends on v_z alone is computed once per plane
// for both slots (the `plane` callback).
// Accesses are buffer operations with the lane's pair in the vector offset
// and the node's plane kv P in the scalar offset (no per-node address
// arithmetic). An item is kB planes of both slots: all its loads are in
// flight together (for Nv <= 10 the whole row), the first item's issued
// before `setup`, whose own memory latency (per-particle constants) then
// overlaps them. A slot past the last pair gets an out-of-range offset
// (its loads return 0, its stores are dropped) and an M0 factor of 0, so
// it adds nothing to any sum; planes past the last reload the last one
// and are skipped.
// M0 of a node is the pair's factor (m0c e_iv) e_jv (a table built with M0)
// times the plane's e_kv, rounded as the table of M0 was built. Per-wave
// tables over v_z (e_kv here, the callers' separable factors) are held one
// entry per lane (lane l: node l) and read into scalar registers with
// lane_value: no memory access per node.
//   setup()  begin(s, iv, jv, vx, vy, ok)  plane(kv)  node(s, p, kv, t, m0)  end(s)
template <typename T>
struct Row
{
    static constexpr int kB = 10;
    static constexpr int kOut = (int)0x80000000;   // vector offset past any buffer
    __amdgpu_buffer_rsrc_t rg;    // stored row (pre-transport values; written last)
    __amdgpu_buffer_rsrc_t rt;    // transported row
    const double *m0xy;
    double ezl;                   // lane l: e_l (0 past the last node)
    int lane, Nv, P;
    double vlo, dv;
    __device__ Row(T *g, T *t, const Ref &r, int ln, int nv, double lo, double d)
        : rg(brsrc(g)), rt(brsrc(t)), m0xy(r.m0xy), ezl(ln < nv ? r.m0x[ln] : 0.0), lane(ln),
          Nv(nv), P(nv * nv), vlo(lo), dv(d)
    {
    }
    // (the plane offset is wave-uniform; readfirstlane keeps it in a scalar
    // register where the compiler would otherwise loop over the lanes)
    __device__ __forceinline__ int soff(int kv) const
    { return __builtin_amdgcn_readfirstlane(kv * P * (int)sizeof(T)); }
    // the byte offset of pair p's value in a plane (kOut for a slot past the end)
    __device__ __forceinline__ int voff(int p) const
    { return p < P ? p * (int)sizeof(T) : kOut; }
    __device__ __forceinline__ T g(int p, int kv) const { return bload_so<T>(rg, voff(p), soff(kv)); }
    __device__ __forceinline__ void set_g(int p, int kv, T v) const
    { bstore_so<T>(rg, voff(p), soff(kv), v); }
    __device__ __forceinline__ void set_t(int p, int kv, T v) const
    { bstore_so<T>(rt, voff(p), soff(kv), v); }
    template <typename FS, typename FB, typename FP, typename FN, typename FE>
    __device__ __forceinline__ void pairs(FS &&setup, FB &&begin, FP &&plane, FN &&node,
                                          FE &&end) const
    {
        const int ncp = (P + 2 * kWave - 1) / (2 * kWave);
        const int nkb = (Nv + kB - 1) / kB;
        for (int c = 0; c < ncp; ++c)
        {
            int p[2];
#pragma unroll
            for (int s = 0; s < 2; ++s) p[s] = (2 * c + s) * kWave + lane;
            for (int b = 0; b < nkb; ++b)
            {
                const int k0 = __builtin_amdgcn_readfirstlane(b * kB);
                T v[2][kB];
                double mxy[2];
#pragma unroll
                for (int s = 0; s < 2; ++s)
                {
                    mxy[s] = p[s] < P ? m0xy[p[s]] : 0.0;
#pragma unroll
                    for (int j = 0; j < kB; ++j)
                        v[s][j] = bload_so<T>(rt, voff(p[s]), soff(min(k0 + j, Nv - 1)));
                }
                if (c == 0 && b == 0) setup();
                if (b == 0)
                {
#pragma unroll
                    for (int s = 0; s < 2; ++s)
                    {
                        const int q = min(p[s], P - 1), jv = q / Nv, iv = q - jv * Nv;
                        begin(s, iv, jv, vnode(vlo, dv, iv), vnode(vlo, dv, jv), p[s] < P);
                    }
                }
#pragma unroll
                for (int j = 0; j < kB; ++j)
                {
                    const int kv = k0 + j;
                    if (kv < Nv)
                    {
                        const double ez = lane_value(ezl, kv);
                        plane(kv);
#pragma unroll
                        for (int s = 0; s < 2; ++s) node(s, p[s], kv, (double)v[s][j], mxy[s] * ez);
                    }
                    // planes one after the other: hoisting the later planes'
                    // conversions would hold them all in registers
                    __builtin_amdgcn_sched_barrier(0);
                }
                if (k0 + kB >= Nv)
                {
#pragma unroll
                    for (int s = 0; s < 2; ++s) end(s);
                }
            }
        }
    }
};

// Boundary particle (one wave): positivity, then the diffuse wall in FP64
// (kWallBC3D, CudaBackend.cu:956-1030), written back as a deviation. Two
// passes over the row. The wall's sums need the limited row, which depends
// on the positivity outcome only through the rescale factor sc and linearly:
// with sel = the fallback value where m + t < 0, else t, the limited row is
// d = sc (m + sel) - m when the repair fires and t when not. The first pass
// therefore sums both variants alongside the positivity sums, and the
// outcome picks one: sum d = sc S_sel - S_m or S_t, and the incoming flux
// sum_{c_n<0} c_n (m + d) = sc I_sel or I_t.
template <typename T>
__device__ __forceinline__ void post_boundary(const DevCloud &c, int i, const Row<T> &row,
                                              const Ref &ref, double Lo, double R,
                                              double out[kPart])
{
    const int lane = row.lane, Nv = row.Nv;
    const double dv3 = c.dv * c.dv * c.dv, vlo = c.vlo, dv = c.dv;
    auto nop = [] {};
    const double nx = c.wnx[i], ny = c.wny[i], nz = c.wnz[i];
    const double Uwx = c.wux[i], Uwy = c.wuy[i], Uwz = c.wuz[i], Tw = c.wt[i];
    double RT2w = 0.0, inorm = 0.0;
    // the wall Maxwellian exp(-|v - U_w|^2 / 2RT_w) / norm: the v_z factor
    // tabulated by the wave (lane l < Nv: node l) while the row's first loads
    // are in flight, the (v_x, v_y) factor per pair (0 for a slot past the
    // last pair)
    double wz = 0.0;                                   // lane l: node l's v_z factor
    auto wall_tab = [&] {
        RT2w = 2.0 * R * Tw;
        inorm = 1.0 / pow(AB_PI * RT2w, 1.5);
        if (lane < Nv)
        {
            const double t = vnode(vlo, dv, lane);
            wz = exp(-((t - Uwz) * (t - Uwz)) / RT2w);
        }
    };
    // c_n = ((v_x - U_x) n_x + (v_y - U_y) n_y) + (v_z - U_z) n_z, summed in
    // that order without contraction: the pair's part first, as written
    double cnxy[2], wxy[2], cz = 0.0, wzk = 0.0;
    auto pair_wall = [&](int s, int, int, double vx, double vy, bool ok) {
#pragma clang fp contract(off)
        cnxy[s] = (vx - Uwx) * nx + (vy - Uwy) * ny;
        wxy[s] = ok ? exp(-((vx - Uwx) * (vx - Uwx) + (vy - Uwy) * (vy - Uwy)) / RT2w) * inorm
                    : 0.0;
    };
    auto plane_wall = [&](int kv) {
        cz = (vnode(vlo, dv, kv) - Uwz) * nz;
        wzk = lane_value(wz, kv);
    };
    auto cn_of = [&](int s) -> double {
#pragma clang fp contract(off)
        return cnxy[s] + cz;
    };
    // pass 1: positivity sums; S_t, S_sel, S_m; I_t, I_sel; the outgoing flux
    double mPos = 0.0, mNeg = 0.0, mFb = 0.0;
    double St = 0.0, Ssel = 0.0, Sm = 0.0, It = 0.0, Isel = 0.0, intP = 0.0;
    row.pairs(wall_tab, pair_wall, plane_wall, [&](int s, int p, int kv, double t, double m) {
        // (selected values, not a selected accumulator: the latter becomes a
        // private-memory array indexed per node)
        const double v = m + t;
        const bool neg = v < 0.0;
        double vs = v;                                  // m + sel
        if (neg) vs = m + (double)row.g(p, kv);
        mPos += neg ? 0.0 : v;
        mNeg += neg ? v : 0.0;
        mFb += neg ? vs : 0.0;
        St += t; Ssel += vs; Sm += m;
        const double cn = cn_of(s);
        if (cn < 0.0) { It += cn * v; Isel += cn * vs; }
        else if (cn > 0.0) intP += cn * (wxy[s] * wzk);
    }, [](int) {});
    double r[9] = {mPos, mNeg, mFb, St, Ssel, Sm, It, Isel, intP};
    wave_sum8(*reinterpret_cast<double(*)[8]>(r));
    r[8] = uniform(wave_sum(r[8]));
    mPos = r[0]; mNeg = r[1]; mFb = r[2];
    const bool fires = (Lo * mNeg < -1e-30 && (mPos + mNeg) > 0);
    double sc = 1.0;
    if (fires)
    {
        const double replaced = mPos + mFb;
        sc = (Lo * replaced > 1e-30) ? (mPos + mNeg) / replaced : 0.0;
    }
    const double m0s = fires ? sc * r[4] - r[5] : r[3];
    const double intM = (fires ? sc * r[7] : r[6]) * dv3;
    intP = r[8] * dv3;
    // intM is in L-units; the reference's fallback rho_w = 1 is in true units
    const double rhoWL = (intP > 1e-30) ? (-intM / intP) : 1.0 / Lo;
    out[0] = Lo * (ref.mom[0] + m0s * dv3);
    out[2] = Lo * rhoWL;
    // pass 2: the wall's re-emission where c_n > 0, the limited row elsewhere
    const T Lt = T(Lo);
    double h0 = 0;
    row.pairs(nop, pair_wall, plane_wall, [&](int s, int p, int kv, double t, double m) {
        double dh;
        if (cn_of(s) > 0.0) dh = rhoWL * (wxy[s] * wzk) - m;
        else if (!fires) dh = t;
        else
        {
            const double sel = (m + t < 0.0) ? (double)row.g(p, kv) : t;
            dh = sc * sel + (sc - 1.0) * m;
        }
        const T dhT = T(dh);
        row.set_g(p, kv, Lt * dhT + T((Lo - 1.0) * m));
        h0 += (double)dhT;
    }, [](int) {});
    out[1] = Lo * (ref.mom[0] + wave_sum(h0) * dv3);
    out[6] = fires ? 1.0 : 0.0;
}

// An interior particle after positivity and NAD (L-units): the NAD factor,
// the limiter bits and the moments before the NAD factor.
struct Raw
{
    double sc2, flag;
    double rho, px, py, pz, E;    // sum (1, v, |v|^2) (M0 + d) dv^3
};

// The committed state derived from Raw (scalar arithmetic of one particle:
// a wave evaluates it uniformly, or one thread per particle).
struct Committed
{
    double sc2;               // NAD factor
    double flag;              // limiter bits
    double mr, px, py, pz, E; // committed moments: sc2 (rho, rho u, sum |v|^2 g)
    double ux, uy, uz, T;     // velocity and temperature (ratios: unaffected by sc2)
    bool okm;                 // mass positive: u and T defined
};

// Positivity (fallback to the pre-transport value, then the rescale to the
// transported mass), the moments, and NAD against the neighbours' current
// densities. The raw moments are summed in the positivity pass and redone
// only when the repair fires (it then rewrites the transported row). The
// pass sums the negative nodes only (mNeg, and their fallback values mFb);
// the positive part is the row's total sum M0 + sum d less mNeg.
template <typename T>
__device__ __forceinline__ Raw commit_interior(const DevCloud &c, int i, const Row<T> &row,
                                               const Ref &ref, const double *rho, double Lo)
{
    const int lane = row.lane;
    const double dv3 = c.dv * c.dv * c.dv;
    auto nop = [] {};
    // NAD bounds: the neighbours' current densities, two per lane at most
    // (lanes past the list repeat its last entry). The list is read before
    // the row's loads are issued and the densities while they are in
    // flight; min and max are taken after the pass.
    static_assert(AB_MAX_NEIGH <= 2 * kWave, "two neighbours per lane");
    const int nn = c.nneigh[i];
    const int *nb = c.neigh + (size_t)i * AB_MAX_NEIGH;
    int j0 = 0, j1 = 0;
    if (nn > 0) { j0 = nb[min(lane, nn - 1)]; j1 = nb[min(lane + kWave, nn - 1)]; }
    double r0 = 0.0, r1 = 0.0;
    auto gather = [&] { if (nn > 0) { r0 = rho[j0]; r1 = rho[j1]; } };
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};      // mPos mNeg mFb r0..r4
    // per pair: A = sum d, K1 = sum kv d, K2 = sum kv^2 d, then the pair's
    // share of the moments (v_z = vlo + kv dv)
    const double vlo = c.vlo, dv = c.dv;
    double A[2], K1[2], K2[2], px[2], py[2], q = 0.0, q2 = 0.0;
    auto begin = [&](int sl, int, int, double vx, double vy, bool) {
        A[sl] = K1[sl] = K2[sl] = 0.0; px[sl] = vx; py[sl] = vy;
    };
    auto plane = [&](int kv) { q = (double)kv; q2 = q * q; };
    auto end = [&](int sl) {
        s[3] += A[sl]; s[4] += px[sl] * A[sl]; s[5] += py[sl] * A[sl];
        s[6] += vlo * A[sl] + dv * K1[sl];
        s[7] += (px[sl] * px[sl] + py[sl] * py[sl] + vlo * vlo) * A[sl]
              + dv * (2.0 * vlo * K1[sl] + dv * K2[sl]);
    };
    auto accum = [&](int sl, double d) { A[sl] += d; K1[sl] += q * d; K2[sl] += q2 * d; };
    row.pairs(gather, begin, plane, [&](int sl, int p, int kv, double d, double m) {
        const double v = m + d;
        if (v < 0.0) { s[1] += v; s[2] += m + (double)row.g(p, kv); }
        accum(sl, d);
    }, end);
    wave_sum8(s);
    s[0] = (ref.m0sum + s[3]) - s[1];
    Raw r;
    r.flag = 0.0;
    if (Lo * s[1] < -1e-30 && (s[0] + s[1]) > 0)
    {
        r.flag = 1.0;
        const double replaced = s[0] + s[2];
        const double sc = (Lo * replaced > 1e-30) ? (s[0] + s[1]) / replaced : 0.0;
#pragma unroll
        for (int k = 3; k < 8; ++k) s[k] = 0.0;
        row.pairs(nop, begin, plane, [&](int sl, int p, int kv, double dg, double m) {
            const double sel = (m + dg < 0.0) ? (double)row.g(p, kv) : dg;
            const T lim = T(sc * sel + (sc - 1.0) * m);
            row.set_t(p, kv, lim);
            accum(sl, (double)lim);
        }, end);
        s[0] = s[1] = s[2] = 0.0;
        wave_sum8(s);
    }
    const double rhoL = ref.mom[0] + s[3] * dv3;
    const double pxL = ref.mom[1] + s[4] * dv3;
    const double pyL = ref.mom[2] + s[5] * dv3;
    const double pzL = ref.mom[3] + s[6] * dv3;
    const double EL = ref.mom[4] + s[7] * dv3;
    double rmin = 1e300, rmax = -1e300;
    if (nn > 0)
    {
        rmin = uniform(wave_min(fmin(r0, r1)));
        rmax = uniform(wave_max(fmax(r0, r1)));
    }
    const double rho_new = Lo * rhoL;
    double sc2 = 1.0;
    if (!(rmin > rmax))
    {
        const double slack = 0.1 * fmax(rmax - rmin, 1e-12) + 1e-9;
        const double lo = rmin - slack, hi = rmax + slack;
        bool nad = false;
        if (rho_new > hi && rho_new > 1e-30) { sc2 = hi / rho_new; nad = true; }
        else if (rho_new < lo && rho_new > 1e-30) { sc2 = lo / rho_new; nad = true; }
        if (nad) r.flag += 2.0;
    }
    r.sc2 = sc2;
    r.rho = rhoL; r.px = pxL; r.py = pyL; r.pz = pzL; r.E = EL;
    return r;
}

__device__ __forceinline__ Committed finish_commit(const Raw &w, const Ref &ref, double Lo,
                                                  double R)
{
    Committed r;
    r.sc2 = w.sc2;
    r.flag = w.flag;
    r.mr = w.sc2 * w.rho;
    r.px = w.sc2 * w.px; r.py = w.sc2 * w.py; r.pz = w.sc2 * w.pz; r.E = w.sc2 * w.E;
    r.ux = 0.0; r.uy = 0.0; r.uz = 0.0; r.T = ref.T;
    r.okm = Lo * r.mr > 1e-30;
    if (r.okm)
    {
        const double ir = 1.0 / w.rho;
        r.ux = w.px * ir; r.uy = w.py * ir; r.uz = w.pz * ir;
        const double u2 = r.ux * r.ux + r.uy * r.uy + r.uz * r.uz;
        const double t1 = (w.E * ir - u2) / (3.0 * R);
        r.T = t1 > 1e-6 ? t1 : 1e-6;
    }
    return r;
}

// Collision target relative to M0 as an exponent D(v) = A0 + Bq |v|^2 + C . v
// (M = M0 exp(D)); the continuous Maxwellian of the committed moments:
// D = ln(rho/rho_ref) - 1.5 ln(T/T_ref) - |u|^2/2RT + |v|^2 (1/2RT_ref - 1/2RT) + v.u/RT
__device__ __forceinline__ void continuous_target(const Committed &m, const Ref &ref, double R,
                                                  double D[5])
{
    D[0] = m.okm ? log(m.mr / ref.rho) - 1.5 * log(m.T / ref.T)
                       - (m.ux * m.ux + m.uy * m.uy + m.uz * m.uz) / (2.0 * R * m.T)
                 : 0.0;
    D[1] = 1.0 / (2.0 * R * ref.T) - 1.0 / (2.0 * R * m.T);
    D[2] = m.ux / (R * m.T); D[3] = m.uy / (R * m.T); D[4] = m.uz / (R * m.T);
}

// BGK collision of one row: h = coeff (g'' + lam M)  =>  dh = coeff (d'' + lam (M - M0)),
// d'' = sc2 d + (sc2 - 1) M0 the committed deviation, M - M0 = M0 (exp(D) - 1)
// in FP64 (target 0 when !okm). exp(D) = G_xy(v_x, v_y) G_z(v_z) with
// G_xy = exp(A0 + Bq (v_x^2 + v_y^2) + Cx v_x + Cy v_y) once per pair and
// G_z(t) = exp(Bq t^2 + Cz t) tabulated by the wave (lane l < Nv: node l):
// no exponential per node. Per node the update is evaluated as
// dh = c1 d + M0 (c2 + c3 (E - 1)) with c1 = coeff sc2, c2 = coeff (sc2 - 1),
// c3 = coeff lam. Writes the stored row (folding the pending rescale Lo into
// it) and returns the grid moments of dh (wave-uniform).
template <typename T>
__device__ __forceinline__ void collide_row(const Row<T> &row, double sc2, bool okm,
                                            const double D[5], double lam, double Lo, double h[5])
{
    const int lane = row.lane;
    const double coeff = 1.0 / (1.0 + lam);
    const double c1 = coeff * sc2, c2 = coeff * (sc2 - 1.0), c3 = coeff * lam;
    // G_z (lane l: node l), tabulated while the row's first loads are in flight
    double gz = 0.0;
    auto gz_tab = [&] {
        if (lane < row.Nv)
        {
            const double t = vnode(row.vlo, row.dv, lane);
            gz = exp(D[1] * t * t + D[4] * t);
        }
    };
    const T Lt = T(Lo);
    const double Lo1 = Lo - 1.0;
#pragma unroll
    for (int k = 0; k < 5; ++k) h[k] = 0.0;
    // per pair: A = sum dh, K1 = sum kv dh, K2 = sum kv^2 dh (v_z = vlo + kv dv)
    const double vlo = row.vlo, dv = row.dv;
    double A[2], K1[2], K2[2], px[2], py[2], Exy[2], gzk = 0.0, q = 0.0, q2 = 0.0;
    row.pairs(gz_tab, [&](int s, int, int, double vx, double vy, bool) {
        A[s] = K1[s] = K2[s] = 0.0; px[s] = vx; py[s] = vy;
        Exy[s] = okm ? exp(D[0] + D[1] * (vx * vx + vy * vy) + D[2] * vx + D[3] * vy) : 0.0;
    }, [&](int kv) {
        gzk = lane_value(gz, kv);
        q = (double)kv; q2 = q * q;
    }, [&](int s, int p, int kv, double t, double m) {
        const double Em1 = fma(Exy[s], gzk, -1.0);
        const T dh = T(fma(c1, t, m * fma(c3, Em1, c2)));
        row.set_g(p, kv, Lt * dh + T(Lo1 * m));
        const double hd = (double)dh;
        A[s] += hd; K1[s] += q * hd; K2[s] += q2 * hd;
    }, [&](int s) {
        h[0] += A[s]; h[1] += px[s] * A[s]; h[2] += py[s] * A[s];
        h[3] += vlo * A[s] + dv * K1[s];
        h[4] += (px[s] * px[s] + py[s] * py[s] + vlo * vlo) * A[s]
              + dv * (2.0 * vlo * K1[s] + dv * K2[s]);
    });
    double t[8] = {h[0], h[1], h[2], h[3], h[4], 0.0, 0.0, 0.0};
    wave_sum8(t);
#pragma unroll
    for (int k = 0; k < 5; ++k) h[k] = t[k];
}

// Per-particle state between the discrete-Maxwellian kernels, structure of
// arrays (slot q of particle i at q N + i).
// P_RAW: the moments before the NAD factor (Raw), P_D: the collision target
// exponent (continuous, or discrete where Newton converged).
enum PostSlot { P_SC2, P_OKM, P_RAW, P_D = P_RAW + 5, P_COUNT = P_D + 5 };

// The particle of a wave of a kernel over a particle list, in a scalar
// register (the row's buffer descriptors are uniform); -1 past the list.
// Consecutive blocks go to different XCDs (no remap): a wave reads only its
// own row.
__device__ __forceinline__ int list_particle(const int *__restrict__ list, int n)
{
    const int w = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
    return w < n ? __builtin_amdgcn_readfirstlane(list[w]) : -1;
}

// The same list from its end: a pass that follows another over the same
// rows starts with the rows the other read last, which the memory-side cache
// still holds.
__device__ __forceinline__ int list_particle_rev(const int *__restrict__ list, int n)
{
    const int w = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
    return w < n ? __builtin_amdgcn_readfirstlane(list[n - 1 - w]) : -1;
}

// One wave per boundary particle (the boundary list): positivity and the
// diffuse wall. Partials: [0] committed mass, [1] mass after the wall, [2]
// the wall density, [6] limiter bits.
template <typename T>
__global__ __launch_bounds__(256) void k_wall(DevCloud c, Lat L, T *__restrict__ g,
                                              T *__restrict__ gt, Ref ref,
                                              const double *__restrict__ scal, double R,
                                              double *__restrict__ part,
                                              const int *__restrict__ blist, int nb)
{
    const int i = list_particle(blist, nb);
    const int lane = threadIdx.x & (kWave - 1);
    if (i < 0) return;
    const double Lo = scal[S_LAZY];
    const Row<T> row(g + (size_t)i * L.S, gt + (size_t)i * L.S, ref, lane, c.Nv, c.vlo, c.dv);
    double out[kPart] = {0, 0, 0, 0, 0, 0, 0, 0};
    post_boundary<T>(c, i, row, ref, Lo, R, out);
    if (lane < kPart - 1) part[(size_t)i * kPart + lane] = out[lane];
}

// One wave per interior particle (the interior list): positivity, NAD and
// moments; then (continuous collision) the BGK collision in the same kernel,
// or (DM) the raw moments and the NAD factor are handed to k_dm_solve, which
// derives the committed state one thread per particle, and k_collide.
// Partials: [0] committed mass, [1..5] moments after the collision, [6]
// limiter bits.
template <typename T, bool DM>
__global__ __launch_bounds__(256) void k_post(DevCloud c, Lat L, T *__restrict__ g,
                                              T *__restrict__ gt, Ref ref,
                                              const double *__restrict__ rho,
                                              const double *__restrict__ scal, double dt,
                                              double tau, double R, double *__restrict__ part,
                                              double *__restrict__ pst,
                                              const int *__restrict__ ilist, int ni)
{
    const int i = list_particle(ilist, ni);
    const int lane = threadIdx.x & (kWave - 1);
    if (i < 0) return;
    const double Lo = scal[S_LAZY];
    const Row<T> row(g + (size_t)i * L.S, gt + (size_t)i * L.S, ref, lane, c.Nv, c.vlo, c.dv);
    double *pp = part + (size_t)i * kPart;
    const Raw w = commit_interior<T>(c, i, row, ref, rho, Lo);
    if constexpr (DM)
    {
        if (lane == 0)
        {
            const size_t N = c.N;
            const double raw[5] = {w.rho, w.px, w.py, w.pz, w.E};
            pst[P_SC2 * N + i] = w.sc2;
#pragma unroll
            for (int q = 0; q < 5; ++q) pst[(P_RAW + q) * N + i] = raw[q];
            pp[0] = Lo * (w.sc2 * w.rho);
            pp[6] = w.flag;
        }
    }
    else
    {
        const Committed m = finish_commit(w, ref, Lo, R);
        double D[5];
        continuous_target(m, ref, R, D);
        double h[5];
        collide_row<T>(row, m.sc2, m.okm, D, dt / tau, Lo, h);
        const double dv3 = c.dv * c.dv * c.dv;
        double out = 0.0;
        if (lane == 0) out = Lo * m.mr;
#pragma unroll
        for (int q = 0; q < 5; ++q) if (lane == 1 + q) out = Lo * (ref.mom[q] + h[q] * dv3);
        if (lane == 6) out = m.flag;
        if (lane < kPart - 1) pp[lane] = out;
    }
}

// Discrete-Maxwellian Newton, one thread per interior particle (64 particles
// share each instruction of the 5x5 FP64 solves): the continuous Maxwellian's
// parameters, plus the MLP's prediction of the difference when enabled (the
// wave's 64 particles on MFMA; every lane takes part), then Newton on the
// separable grid sums until the relative residual passes. A converged
// solution replaces the collision target exponent.
__global__ __launch_bounds__(256) void k_dm_solve(int N, const unsigned char *__restrict__ bnd,
                                                  DmArgs dm, Ref ref, const double *__restrict__ scal,
                                                  double R, int Nv, double vlo, double dv,
                                                  double *__restrict__ pst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = N;
    const bool inner = i < N && !bnd[i];
    // the committed state and the continuous target (one thread per particle)
    Raw w{1.0, 0.0, ref.rho, 0.0, 0.0, 0.0, 0.0};
    if (inner)
    {
        w.sc2 = pst[P_SC2 * n + i];
        w.rho = pst[(P_RAW + 0) * n + i]; w.px = pst[(P_RAW + 1) * n + i];
        w.py = pst[(P_RAW + 2) * n + i]; w.pz = pst[(P_RAW + 3) * n + i];
        w.E = pst[(P_RAW + 4) * n + i];
    }
    const Committed m = finish_commit(w, ref, scal[S_LAZY], R);
    double D[5];
    continuous_target(m, ref, R, D);
    if (inner)
    {
        pst[P_OKM * n + i] = m.okm ? 1.0 : 0.0;
#pragma unroll
        for (int q = 0; q < 5; ++q) pst[(P_D + q) * n + i] = D[q];
    }
    const bool act = inner && m.okm;
    const double cs = dm.cs, ics = 1.0 / cs, dv3 = dv * dv * dv;
    const double ux = m.ux, uy = m.uy, uz = m.uz, T = m.T;
    double ms[5] = {m.mr, m.px, m.py, m.pz, m.E};
    const float feat[kMlpIn] = {float(ux * ics), float(uy * ics), float(uz * ics),
                                float(R * T * ics * ics), float(dv * ics)};
    float da[kMlpOut] = {0, 0, 0, 0, 0};
    if (dm.mlp) mlp_wave_mfma(dm.wb, dm.pf, feat, da);
    if (!act) return;
    double ac[5], aref[5], a[5];
    dm_guess(ms[0], ux, uy, uz, T, R, cs, ac);
    dm_guess(ref.rho, 0.0, 0.0, 0.0, ref.T, R, cs, aref);
#pragma unroll
    for (int q = 0; q < 5; ++q) a[q] = ac[q] + (double)da[q];
    // the moments in the scaled basis psi = (1, v/cs, |v|^2/cs^2)
    ms[1] *= ics; ms[2] *= ics; ms[3] *= ics; ms[4] *= ics * ics;
    bool ok = true;
    int pass = 0;
    double res = 0.0;
    for (;;)
    {
        ++pass;
        double S[15], F[5];
        dm_grid_sums(a, Nv, vlo, dv, ics, dv3, S);
        res = 0.0;
#pragma unroll
        for (int q = 0; q < 5; ++q) { F[q] = S[q] - ms[q]; res = fmax(res, fabs(F[q])); }
        res /= fabs(ms[0]);
        if (res <= dm.tol || pass > dm.maxit) break;
        double d[5];
        if (!dm_solve5(S, F, d)) { ok = false; break; }
#pragma unroll
        for (int q = 0; q < 5; ++q) a[q] -= d[q];
    }
    bool conv = ok && res <= dm.tol;
#pragma unroll
    for (int q = 0; q < 5; ++q) conv = conv && isfinite(a[q]);
    if (!conv)
    {
#pragma unroll
        for (int q = 0; q < 5; ++q) a[q] = ac[q];               // continuous target kept
    }
    else
    {
        // D = (a - a_ref) . psi in unscaled velocities, a_ref the parameters of M0
        pst[(P_D + 0) * n + i] = a[0] - aref[0];
        pst[(P_D + 1) * n + i] = (a[4] - aref[4]) * ics * ics;
        pst[(P_D + 2) * n + i] = a[1] * ics;
        pst[(P_D + 3) * n + i] = a[2] * ics;
        pst[(P_D + 4) * n + i] = a[3] * ics;
    }
#pragma unroll
    for (int q = 0; q < 5; ++q) dm.alpha[(size_t)i * 5 + q] = a[q];
    dm.passes[i] = conv ? pass : 2 * dm.maxit + 2;
    dm.res[i] = res;
    if (dm.rec)
    {
        float *rw = dm.rec + (size_t)i * kMlpRow;
#pragma unroll
        for (int q = 0; q < kMlpIn; ++q) rw[q] = feat[q];
#pragma unroll
        for (int q = 0; q < kMlpOut; ++q) rw[kMlpIn + q] = float(a[q] - ac[q]);
        rw[kMlpIn + kMlpOut] = conv ? 1.0f : 0.0f;
    }
}

// The collision of the interior particles towards the target k_dm_solve left
// (one wave per particle, the list in reverse: k_post read it forwards);
// partials [1..5].
template <typename T>
__global__ __launch_bounds__(256) void k_collide(DevCloud c, Lat L, T *__restrict__ g,
                                                 const T *__restrict__ gt, Ref ref,
                                                 const double *__restrict__ scal, double dt,
                                                 double tau, const double *__restrict__ pst,
                                                 double *__restrict__ part,
                                                 const int *__restrict__ ilist, int ni)
{
    const int i = list_particle_rev(ilist, ni);
    const int lane = threadIdx.x & (kWave - 1);
    if (i < 0) return;
    const double Lo = scal[S_LAZY];
    const size_t N = c.N;
    const Row<T> row(g + (size_t)i * L.S, const_cast<T *>(gt) + (size_t)i * L.S, ref, lane,
                     c.Nv, c.vlo, c.dv);
    double D[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) D[q] = pst[(P_D + q) * N + i];
    double h[5];
    collide_row<T>(row, pst[P_SC2 * N + i], pst[P_OKM * N + i] != 0.0, D, dt / tau, Lo, h);
    const double dv3 = c.dv * c.dv * c.dv;
#pragma unroll
    for (int q = 0; q < 5; ++q)
        if (lane == q) part[(size_t)i * kPart + 1 + q] = Lo * (ref.mom[q] + h[q] * dv3);
}

// Global sums of the partials -> the two rescale factors of cudaStep
// (:1706-1710, :1773-1780), folded into the lazy factor. Two stages with a
// fixed partition (deterministic): kRedBlocks blocks sum contiguous particle
// ranges, then one block combines them.
constexpr int kRedBlocks = 256;

__global__ __launch_bounds__(256) void k_reduce_stage(int N, const double *__restrict__ part,
                                                      double *__restrict__ red)
{
    __shared__ double sh[2 * (256 / kWave)];
    const int per = (N + kRedBlocks - 1) / kRedBlocks;
    const int i0 = blockIdx.x * per, i1 = min(N, i0 + per);
    double v[2] = {0.0, 0.0};
    for (int i = i0 + (int)threadIdx.x; i < i1; i += 256)
    {
        v[0] += part[(size_t)i * kPart];
        v[1] += part[(size_t)i * kPart + 1];
    }
    block_sum<2>(v, sh);
    if (threadIdx.x == 0) { red[2 * blockIdx.x] = v[0]; red[2 * blockIdx.x + 1] = v[1]; }
}

__global__ __launch_bounds__(256) void k_reduce(const double *__restrict__ red, double *scal,
                                                int nofix)
{
    __shared__ double sh[2 * (256 / kWave)];
    double v[2] = {0.0, 0.0};
    for (int b = threadIdx.x; b < kRedBlocks; b += 256) { v[0] += red[2 * b]; v[1] += red[2 * b + 1]; }
    block_sum<2>(v, sh);
    if (threadIdx.x == 0)
    {
        const double mNew = v[0], hsum = v[1];
        const double mOld = scal[S_MASS_OLD];
        const double s1 = (mNew > 1e-30) ? mOld / mNew : 1.0;
        const double mColl = s1 * hsum;
        const double s2 = (!nofix && mColl > 1e-30) ? mOld / mColl : 1.0;
        scal[S_MASS_START] = mOld;
        scal[S_MASS_NEW] = mNew;
        scal[S_SCALE1] = s1;
        scal[S_MASS_COLL] = mColl;
        scal[S_SCALE2] = s2;
        scal[S_LAZY] = s1 * s2;
        scal[S_MASS_OLD] = s2 * mColl;          // the next step's starting mass
    }
}

__global__ void k_finalize(DevCloud c, const double *__restrict__ part,
                           const double *__restrict__ scal, double R, double *rho, double *ux,
                           double *uy, double *uz, double *T, double *limflag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.N) return;
    const double *p = part + (size_t)i * kPart;
    limflag[i] = p[6];
    if (c.bnd[i])
    {
        rho[i] = scal[S_SCALE1] * p[2];
        ux[i] = c.wux[i]; uy[i] = c.wuy[i]; uz[i] = c.wuz[i]; T[i] = c.wt[i];
        return;
    }
    const double r = scal[S_LAZY] * p[1];
    if (r > 1e-30)
    {
        const double mux = p[2] / p[1], muy = p[3] / p[1], muz = p[4] / p[1];
        const double u2 = mux * mux + muy * muy + muz * muz;
        const double t1 = (p[5] / p[1] - u2) / (3.0 * R);
        rho[i] = r; ux[i] = mux; uy[i] = muy; uz[i] = muz; T[i] = t1 > 1e-6 ? t1 : 1e-6;
    }
}

// Mass of the stored field, sum L (M0 + d) dv^3 (L applied by k_mass_from_part).
template <typename T>
__global__ void k_table_mass(int N, int nodes, int S, const T *__restrict__ g,
                             const double *__restrict__ m0, double dv3,
                             double *__restrict__ part)
{
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    if (i >= N) return;
    double m = 0.0;
    for (int k = lane; k < nodes; k += kWave) m += m0[k] + (double)g[(size_t)i * S + k];
    m = wave_sum(m);
    if (lane == 0) part[(size_t)i * kPart + 7] = m * dv3;
}

__global__ __launch_bounds__(1024) void k_mass_from_part(int N, const double *__restrict__ part,
                                                         double *scal)
{
    __shared__ double sh[1024 / kWave];
    double a = 0.0;
    for (int i = threadIdx.x; i < N; i += 1024) a += part[(size_t)i * kPart + 7];
    a = wave_sum(a);
    if ((threadIdx.x & (kWave - 1)) == 0) sh[threadIdx.x / kWave] = a;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double t = 0.0;
        for (int q = 0; q < 1024 / kWave; ++q) t += sh[q];
        scal[S_MASS_OLD] = t * scal[S_LAZY];
    }
}

// FP64 g -> stored deviation d = g - M0 (L = 1), padded rows.
template <typename T>
__global__ void k_pack(int N, int nodes, int gstride, int S, const double *__restrict__ src,
                       const double *__restrict__ m0, T *__restrict__ dst)
{
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (size_t)N * S) return;
    size_t i = k / S, lin = k % S;
    dst[k] = (lin < (size_t)nodes) ? T(src[i * gstride + lin] - m0[lin]) : T(0);
}

template <typename T>
__global__ void k_unpack(int N, int gstride, int S, const T *__restrict__ src,
                         const double *__restrict__ m0, const double *__restrict__ scal,
                         double *__restrict__ dst)
{
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (size_t)N * gstride) return;
    size_t i = k / gstride, lin = k % gstride;
    dst[k] = (m0[lin] + (double)src[i * S + lin]) * scal[S_LAZY];
}

__global__ void k_set_scalar(double *scal, int slot, double v) { scal[slot] = v; }

bool is_fp32(const ab_ctx *ctx) { return ctx->cfg.precision == AB_PREC_MIXED; }

// Launch shapes of the transport contraction the autotuner chooses from:
// K outputs per z block (values reused across K + 4 planes), W the waves per
// SIMD the register allocation targets. Every shape does the same products
// in the same order per output, so the results do not depend on the choice.
// TX x TY > 1: interior blocks + edge columns; S > 0: the slab order, items of S z blocks
struct ApplyShape { int K, W, TX, TY, S; };
constexpr ApplyShape kShapes32[] = {{12, 3, 1, 1, 0}, {9, 3, 1, 1, 0}, {12, 2, 2, 2, 0},
                                    {12, 2, 2, 2, 1}, {8, 2, 2, 3, 1}, {6, 2, 2, 4, 1},
                                    {6, 2, 2, 4, 2}};
constexpr ApplyShape kShapes64[] = {{9, 2, 1, 1, 0}, {8, 2, 1, 1, 0}, {7, 2, 1, 1, 0},
                                    {6, 2, 1, 1, 0}};
constexpr int kMaxBlockPlanes = kMaxZBlockPlanes;

template <typename T>
constexpr int nshapes() { return sizeof(T) == 4 ? 7 : 4; }

template <typename T>
ApplyShape shape_at(int i) { return sizeof(T) == 4 ? kShapes32[i] : kShapes64[i]; }

// Launch `shape` over the lattice (a block shape: k_apply_mix, blocks and the
// edge columns in one grid).
template <typename T>
void launch_shape(int i, Lat L, const T *g, T *gt, const T *cf, const double *scal,
                  hipStream_t s)
{
    auto waves = [&](int ncol) { return (ncol * L.nchunk + 3) / 4; };
    const int all = L.nx * L.ny, inner = (L.nx - 4) * (L.ny - 4);
    if constexpr (sizeof(T) == 4)
    {
        const ApplyShape sh = kShapes32[i];
        if (sh.TX > 1)
        {
            const int nblk = inner / (sh.TX * sh.TY);
            const int nw = (sh.S ? nblk * ((L.nz - 4 + sh.S * sh.K - 1) / (sh.S * sh.K) + 1) : nblk) +
                           edge_items(L.nx, L.ny, edge_slabs(sh.S != 0, L.nz));
            switch (i)
            {
            case 3: k_apply_mix<T, 2, 2, 12, 12, true><<<waves(nw), 256, 0, s>>>(L, g, gt, cf); break;
            case 4: k_apply_mix<T, 2, 3, 8, 8, true><<<waves(nw), 256, 0, s>>>(L, g, gt, cf); break;
            case 5: k_apply_mix<T, 2, 4, 6, 8, 1><<<waves(nw), 256, 0, s>>>(L, g, gt, cf); break;
            case 6: k_apply_mix<T, 2, 4, 6, 8, 2><<<waves(nw), 256, 0, s>>>(L, g, gt, cf); break;
            default: k_apply_mix<T, 2, 2, 12, 12, false><<<waves(nw), 256, 0, s>>>(L, g, gt, cf); break;
            }
            return;
        }
        const int nb = waves(all);
        if (i == 1) k_apply<T, 9, 3><<<nb, 256, 0, s>>>(L, g, gt, cf, scal);
        else k_apply<T, 12, 3><<<nb, 256, 0, s>>>(L, g, gt, cf, scal);
    }
    else
    {
        const int nb = waves(all);
        switch (i)
        {
        case 0: k_apply<T, 9, 2><<<nb, 256, 0, s>>>(L, g, gt, cf, scal); break;
        case 1: k_apply<T, 8, 2><<<nb, 256, 0, s>>>(L, g, gt, cf, scal); break;
        case 2: k_apply<T, 7, 2><<<nb, 256, 0, s>>>(L, g, gt, cf, scal); break;
        default: k_apply<T, 6, 2><<<nb, 256, 0, s>>>(L, g, gt, cf, scal); break;
        }
    }
}

// A shape applies to this lattice (blocks must tile the interior columns).
template <typename T>
bool shape_fits(int i, int nx, int ny)
{
    const ApplyShape sh = shape_at<T>(i);
    return sh.TX == 1 || ((nx - 4) % sh.TX == 0 && (ny - 4) % sh.TY == 0 && nx > 4 && ny > 4);
}

// Default shape before tuning: in FP32 the 2 x 4 column blocks with K = 6
// in the slab order where they tile the interior (the fastest shape measured,
// 267 us at 40^3 / Nv 10 and 860 us at 60^3 / Nv 10, against 612 and 1934
// us for single columns), else the 2 x 2 blocks, else single columns; in
// FP64 the single-column z block that divides the interior run.
template <typename T>
int default_shape(int nx, int ny, int nz)
{
    if (sizeof(T) == 4)
        return shape_fits<T>(5, nx, ny) ? 5 : (shape_fits<T>(3, nx, ny) ? 3 : 0);
    const int run = nz - 4;
    return run % 9 == 0 ? 0 : (run % 7 == 0 ? 2 : 1);
}

template <typename T>
int apply_launch(ab_ctx *ctx, const Lat &L, int shape)
{
    // buffer offsets within a z block must stay below 2^31
    if ((double)kMaxBlockPlanes * ctx->cfg.nx * ctx->cfg.ny * (double)ctx->S * sizeof(T) >= 2147483648.0)
    {
        ctx->err = "table transport: a z block of the lattice exceeds 2 GiB";
        return AB_ERR_UNSUPPORTED;
    }
    if (shape < 0 || shape >= nshapes<T>() || !shape_fits<T>(shape, L.nx, L.ny))
        shape = default_shape<T>(ctx->cfg.nx, ctx->cfg.ny, ctx->cfg.nz);
    launch_shape<T>(shape, L, (const T *)ctx->gs, (T *)ctx->gts, (const T *)ctx->coef, ctx->scal,
                    ctx->stream);
    return hipGetLastError() == hipSuccess ? AB_OK : AB_ERR_HIP;
}

}  // namespace

void table_free(ab_ctx *ctx)
{
    void *ptrs[] = {ctx->gs, ctx->gts, ctx->coef, ctx->cls, ctx->lat, ctx->part, ctx->bad,
                    ctx->m0, ctx->m0x, ctx->m0xy, ctx->red, ctx->pst};
    for (void *p : ptrs) if (p) (void)hipFree(p);
    ctx->gs = ctx->gts = ctx->coef = nullptr;
    ctx->m0 = ctx->m0x = ctx->m0xy = ctx->red = ctx->pst = nullptr;
    ctx->cls = ctx->lat = ctx->bad = nullptr;
    ctx->part = nullptr;
}

// Validate that the cloud is the lattice the config names (every particle's
// neighbour set is its class's stencil and its wall state is its class's),
// then build the tables. Host-side check of the uploaded arrays.
// The lattice classes of the uploaded cloud and their check: every
// particle has its class representative's stencil (neighbours at the same
// lattice offsets and displacements, to 1e-12 of the support radius; the
// list order may differ) and wall state. lat: (ix, iy, iz) per particle; rep:
// the representative of each class.
static int table_check(ab_ctx *ctx, std::vector<int> &lat, std::vector<int> &rep)
{
    const ab_config &cfg = ctx->cfg;
    const int N = ctx->N;
    if (cfg.dim != 3)
    { ctx->err = "table pipelines need dim 3 (2-D runs with AB_PREC_FP64)"; return AB_ERR_UNSUPPORTED; }
    if (cfg.nx < 5 || cfg.ny < 5 || cfg.nz < 5 || (long)cfg.nx * cfg.ny * cfg.nz != N)
    { ctx->err = "table pipelines need a lattice cloud: nx, ny, nz >= 5 and N = nx*ny*nz"; return AB_ERR_UNSUPPORTED; }
    if (cfg.mls_clamp)
    { ctx->err = "mls_clamp is nonlinear in g; run it with AB_PREC_FP64"; return AB_ERR_UNSUPPORTED; }

    std::vector<double> x(N), y(N), z(N), w(10 * (size_t)N);
    std::vector<unsigned char> b(N);
    std::vector<int> nb((size_t)N * AB_MAX_NEIGH), nn(N);
    hipStream_t s = ctx->stream;
    auto dl = [&](void *dst, const void *src, size_t n) {
        return hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, s) == hipSuccess;
    };
    bool ok = dl(x.data(), ctx->x, 8 * (size_t)N) && dl(y.data(), ctx->y, 8 * (size_t)N)
           && dl(z.data(), ctx->z, 8 * (size_t)N) && dl(b.data(), ctx->bnd, N)
           && dl(nb.data(), ctx->neigh, 4 * (size_t)N * AB_MAX_NEIGH)
           && dl(nn.data(), ctx->nneigh, 4 * (size_t)N);
    const double *wsrc[10] = {ctx->wnx, ctx->wny, ctx->wnz, ctx->wux, ctx->wuy, ctx->wuz,
                              ctx->wt, ctx->mvx, ctx->mvy, ctx->mvz};
    for (int k = 0; k < 10 && ok; ++k) ok = dl(&w[(size_t)k * N], wsrc[k], 8 * (size_t)N);
    if (!ok || hipStreamSynchronize(s) != hipSuccess)
    { ctx->err = "table setup: device-to-host copy failed"; return AB_ERR_HIP; }

    const int nx = cfg.nx, ny = cfg.ny, nz = cfg.nz;
    lat.assign(3 * (size_t)N, 0);
    rep.assign(kNcls, -1);
    std::vector<int> cls(N);
    std::vector<unsigned long long> mlo(N, 0), mhi(N, 0);
    for (int p = 0; p < N; ++p)
    {
        int ix = p % nx, iy = (p / nx) % ny, iz = p / (nx * ny);
        lat[3 * p] = ix; lat[3 * p + 1] = iy; lat[3 * p + 2] = iz;
        cls[p] = category(ix, nx) + 5 * (category(iy, ny) + 5 * category(iz, nz));
        if (rep[cls[p]] < 0) rep[cls[p]] = p;
    }
    for (int p = 0; p < N; ++p)
    {
        for (int m = 0; m < nn[p]; ++m)
        {
            int j = nb[(size_t)p * AB_MAX_NEIGH + m];
            int jx = j % nx, jy = (j / nx) % ny, jz = j / (nx * ny);
            int o = canon_index(jx - lat[3 * p], jy - lat[3 * p + 1], jz - lat[3 * p + 2]);
            if (o < 0)
            {
                ctx->err = "table setup: particle " + std::to_string(p) +
                           " has a neighbour outside the lattice stencil";
                return AB_ERR_UNSUPPORTED;
            }
            if (o < 64) mlo[p] |= 1ull << o; else mhi[p] |= 1ull << (o - 64);
        }
    }
    const double tol = 1e-12 * cfg.radius;
    for (int p = 0; p < N; ++p)
    {
        int r = rep[cls[p]];
        bool same = (mlo[p] == mlo[r]) && (mhi[p] == mhi[r]) && (nn[p] == nn[r]) && (b[p] == b[r]);
        for (int k = 0; k < 10 && same; ++k) same = (w[(size_t)k * N + p] == w[(size_t)k * N + r]);
        // the neighbours at the same displacements (the representative has
        // a neighbour at every lattice offset p has: the masks are equal)
        for (int m = 0; m < nn[p] && same; ++m)
        {
            const int jp = nb[(size_t)p * AB_MAX_NEIGH + m], jr = r + (jp - p);
            same = std::fabs((x[jp] - x[p]) - (x[jr] - x[r])) <= tol &&
                   std::fabs((y[jp] - y[p]) - (y[jr] - y[r])) <= tol &&
                   std::fabs((z[jp] - z[p]) - (z[jr] - z[r])) <= tol;
        }
        if (!same)
        {
            ctx->err = "table setup: particle " + std::to_string(p) +
                       " differs from its lattice class (stencil, geometry or wall state)";
            return AB_ERR_UNSUPPORTED;
        }
    }
    return AB_OK;
}

int table_recheck(ab_ctx *ctx)
{
    std::vector<int> lat, rep;
    return table_check(ctx, lat, rep);
}

int table_setup(ab_ctx *ctx)
{
    const int N = ctx->N;
    std::vector<int> lat, rep;
    int rc = table_check(ctx, lat, rep);
    if (rc) return rc;
    std::vector<double> hr(N), hT(N);
    hipStream_t s = ctx->stream;
    if (hipMemcpyAsync(hr.data(), ctx->rho, 8 * (size_t)N, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(hT.data(), ctx->T, 8 * (size_t)N, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
    { ctx->err = "table setup: device-to-host copy failed"; return AB_ERR_HIP; }
    const ab_config &cfg = ctx->cfg;

    const bool f32 = is_fp32(ctx);
    const size_t es = f32 ? sizeof(float) : sizeof(double);
    ctx->S = ((ctx->nodes + kWave - 1) / kWave) * kWave;
    ctx->ncls = kNcls;
    const size_t gsz = (size_t)N * ctx->S * es;
    if (!ctx->gs)
    {
        bool a = hipMalloc(&ctx->gs, gsz) == hipSuccess && hipMalloc(&ctx->gts, gsz) == hipSuccess
              && hipMalloc(&ctx->coef, (size_t)kNcls * kOff * ctx->S * es) == hipSuccess
              && hipMalloc((void **)&ctx->lat, 3 * sizeof(int) * (size_t)N) == hipSuccess
              && hipMalloc((void **)&ctx->cls, sizeof(int) * kNcls) == hipSuccess
              && hipMalloc((void **)&ctx->part, sizeof(double) * kPart * (size_t)N) == hipSuccess
              && hipMalloc((void **)&ctx->m0, sizeof(double) * ctx->S) == hipSuccess
              && hipMalloc((void **)&ctx->m0x, sizeof(double) * kMaxNv) == hipSuccess
              && hipMalloc((void **)&ctx->m0xy, sizeof(double) * kMaxNv * kMaxNv) == hipSuccess
              && hipMalloc((void **)&ctx->red, sizeof(double) * 2 * kRedBlocks) == hipSuccess
              && (ctx->cfg.collision != AB_COLLISION_DISCRETE
                  || hipMalloc((void **)&ctx->pst, sizeof(double) * P_COUNT * (size_t)N) == hipSuccess);
        if (!a) { table_free(ctx); ctx->err = "table setup: device allocation failed"; return AB_ERR_NOMEM; }
        if (hipMemsetAsync(ctx->gts, 0, gsz, s) != hipSuccess)
        { ctx->err = "table setup: memset failed"; return AB_ERR_HIP; }
    }
    // The equilibrium the deviation is stored against: the mean state of the
    // uploaded cloud at rest (any reference is exact; a close one keeps the
    // stored deviation small).
    double rs = 0.0, ts = 0.0;
    for (int p = 0; p < N; ++p) { rs += hr[p]; ts += hT[p]; }
    ctx->m0rho = rs / N;
    ctx->m0T = ts / N;
    {
        const int Nv = ctx->Nv;
        const double R = cfg.R, twoRT = 2.0 * R * ctx->m0T;
        const double norm = std::pow(AB_PI * twoRT, 1.5);
        const double dv3 = ctx->dv * ctx->dv * ctx->dv;
        // M0 as the product of its 1-D factors, in the order the kernels
        // recompute it: ((m0c e_iv) e_jv) e_kv
        std::vector<double> m0(ctx->S, 0.0), ex(kMaxNv, 0.0);
        for (int l = 0; l < Nv; ++l)
        {
            const double v = vnode(ctx->vlo, ctx->dv, l);
            ex[l] = std::exp(-(v * v) / twoRT);
        }
        ctx->m0c = ctx->m0rho / norm;
        std::vector<double> mxy((size_t)kMaxNv * kMaxNv, 0.0);
        for (int p = 0; p < Nv * Nv; ++p) mxy[p] = (ctx->m0c * ex[p % Nv]) * ex[p / Nv];
        ctx->m0sum = 0.0;
        for (int q = 0; q < 5; ++q) ctx->m0mom[q] = 0.0;
        for (int k = 0; k < ctx->nodes; ++k)
        {
            const int iv = k % Nv, jv = (k / Nv) % Nv, kv = k / (Nv * Nv);
            const double vx = vnode(ctx->vlo, ctx->dv, iv);
            const double vy = vnode(ctx->vlo, ctx->dv, jv);
            const double vz = vnode(ctx->vlo, ctx->dv, kv);
            m0[k] = ((ctx->m0c * ex[iv]) * ex[jv]) * ex[kv];
            ctx->m0sum += m0[k];
            ctx->m0mom[0] += m0[k] * dv3;
            ctx->m0mom[1] += vx * m0[k] * dv3;
            ctx->m0mom[2] += vy * m0[k] * dv3;
            ctx->m0mom[3] += vz * m0[k] * dv3;
            ctx->m0mom[4] += (vx * vx + vy * vy + vz * vz) * m0[k] * dv3;
        }
        if (hipMemcpyAsync(ctx->m0, m0.data(), sizeof(double) * ctx->S, hipMemcpyHostToDevice, s)
                != hipSuccess
            || hipMemcpyAsync(ctx->m0x, ex.data(), sizeof(double) * kMaxNv, hipMemcpyHostToDevice, s)
               != hipSuccess
            || hipMemcpyAsync(ctx->m0xy, mxy.data(), sizeof(double) * mxy.size(), hipMemcpyHostToDevice, s)
                != hipSuccess
            || hipStreamSynchronize(s) != hipSuccess)
        { ctx->err = "table setup: host-to-device copy failed"; return AB_ERR_HIP; }
    }
    if (hipMemcpyAsync(ctx->lat, lat.data(), 3 * sizeof(int) * (size_t)N, hipMemcpyHostToDevice, s) != hipSuccess
        || hipMemcpyAsync(ctx->cls, rep.data(), sizeof(int) * kNcls, hipMemcpyHostToDevice, s) != hipSuccess)
    { ctx->err = "table setup: host-to-device copy failed"; return AB_ERR_HIP; }
    return AB_OK;
}

// Pack g (FP64, particle-major) into the padded table layout and (re)build
// the coefficient tables for time step dt.
static int table_build(ab_ctx *ctx, double dt)
{
    const double tol = ctx->cfg.mls_tol > 0 ? ctx->cfg.mls_tol : 1e-12;
    dim3 grid((ctx->S + 255) / 256, kNcls);
    ab::DevCloud c = ctx->view();
    if (is_fp32(ctx))
        k_build<float><<<grid, 256, 0, ctx->stream>>>(c, ctx->cls, ctx->lat, dt, ctx->cfg.radius,
                                                      tol, ctx->cfg.mls_order, (float *)ctx->coef,
                                                      ctx->S);
    else
        k_build<double><<<grid, 256, 0, ctx->stream>>>(c, ctx->cls, ctx->lat, dt,
                                                       ctx->cfg.radius, tol, ctx->cfg.mls_order,
                                                       (double *)ctx->coef, ctx->S);
    return hipGetLastError() == hipSuccess ? AB_OK : AB_ERR_HIP;
}

int table_upload_g(ab_ctx *ctx, const double *g_dev)
{
    const size_t n = (size_t)ctx->N * ctx->S;
    const int bs = 256;
    const unsigned nb = (unsigned)((n + bs - 1) / bs);
    k_set_scalar<<<1, 1, 0, ctx->stream>>>(ctx->scal, S_LAZY, 1.0);
    if (is_fp32(ctx))
        k_pack<float><<<nb, bs, 0, ctx->stream>>>(ctx->N, ctx->nodes, ctx->gstride, ctx->S, g_dev, ctx->m0,
                                                  (float *)ctx->gs);
    else
        k_pack<double><<<nb, bs, 0, ctx->stream>>>(ctx->N, ctx->nodes, ctx->gstride, ctx->S, g_dev, ctx->m0,
                                                   (double *)ctx->gs);
    const int nbm = (ctx->N * kWave + 255) / 256;
    const double dv3 = ctx->dv * ctx->dv * ctx->dv;
    if (is_fp32(ctx))
        k_table_mass<float><<<nbm, 256, 0, ctx->stream>>>(ctx->N, ctx->nodes, ctx->S,
                                                          (const float *)ctx->gs, ctx->m0, dv3,
                                                          ctx->part);
    else
        k_table_mass<double><<<nbm, 256, 0, ctx->stream>>>(ctx->N, ctx->nodes, ctx->S,
                                                           (const double *)ctx->gs, ctx->m0, dv3,
                                                           ctx->part);
    k_mass_from_part<<<1, 1024, 0, ctx->stream>>>(ctx->N, ctx->part, ctx->scal);
    return hipGetLastError() == hipSuccess ? AB_OK : AB_ERR_HIP;
}

int table_download_g(ab_ctx *ctx, double *g_dev)
{
    const size_t n = (size_t)ctx->N * ctx->gstride;
    const int bs = 256;
    const unsigned nb = (unsigned)((n + bs - 1) / bs);
    if (is_fp32(ctx))
        k_unpack<float><<<nb, bs, 0, ctx->stream>>>(ctx->N, ctx->gstride, ctx->S,
                                                    (const float *)ctx->gs, ctx->m0, ctx->scal,
                                                    g_dev);
    else
        k_unpack<double><<<nb, bs, 0, ctx->stream>>>(ctx->N, ctx->gstride, ctx->S,
                                                     (const double *)ctx->gs, ctx->m0, ctx->scal,
                                                     g_dev);
    return hipGetLastError() == hipSuccess ? AB_OK : AB_ERR_HIP;
}

int table_prepare(ab_ctx *ctx, double dt)
{
    int rc = table_build(ctx, dt);
    if (rc) { ctx->err = "table build failed"; return rc; }
    ctx->coef_dt = dt;
    return AB_OK;
}

int table_step(ab_ctx *ctx, double dt)
{
    // The coefficients depend on dt through the foot points: rebuild when it
    // changes (the last step of a run is shortened to land on tfinal).
    if (ctx->coef_dt != dt)
    {
        int rc = table_prepare(ctx, dt);
        if (rc) return rc;
    }
    Lat L{ctx->cfg.nx, ctx->cfg.ny, ctx->cfg.nz, ctx->S, ctx->S / kWave};
    ab::DevCloud c = ctx->view();
    hipStream_t s = ctx->stream;
    auto tic = [&]() { prof_begin(ctx); };
    auto toc = [&](int st) { prof_end(ctx, st); };
    Ref ref;
    ref.m0 = ctx->m0;
    ref.m0xy = ctx->m0xy;
    ref.m0x = ctx->m0x;
    ref.m0c = ctx->m0c;
    ref.m0sum = ctx->m0sum;
    for (int q = 0; q < 5; ++q) ref.mom[q] = ctx->m0mom[q];
    ref.rho = ctx->m0rho;
    ref.T = ctx->m0T;
    tic();
    const int arc = is_fp32(ctx) ? apply_launch<float>(ctx, L, ctx->apply_shape)
                                 : apply_launch<double>(ctx, L, ctx->apply_shape);
    if (arc) return arc;
    toc(ST_TRANSPORT);
    tic();
    const DmArgs dm = dm_args(ctx);
    const int nbs = (ctx->N + 255) / 256;
    // one wave per particle of each list (four per block)
    const int ni = ctx->n_int, nb = ctx->n_bnd;
    const int nbi = (ni + 3) / 4, nbb = (nb + 3) / 4;
    auto post = [&](auto *gs) {
        using T = std::remove_pointer_t<decltype(gs)>;
        T *gts = (T *)ctx->gts;
        if (nb)
            k_wall<T><<<nbb, 256, 0, s>>>(c, L, gs, gts, ref, ctx->scal, ctx->cfg.R, ctx->part,
                                          ctx->blist, nb);
        if (!ni) return;
        if (!dm.on)
        {
            k_post<T, false><<<nbi, 256, 0, s>>>(c, L, gs, gts, ref, ctx->rho, ctx->scal, dt,
                                                 ctx->cfg.tau, ctx->cfg.R, ctx->part, nullptr,
                                                 ctx->ilist, ni);
            return;
        }
        k_post<T, true><<<nbi, 256, 0, s>>>(c, L, gs, gts, ref, ctx->rho, ctx->scal, dt,
                                            ctx->cfg.tau, ctx->cfg.R, ctx->part, ctx->pst,
                                            ctx->ilist, ni);
        k_dm_solve<<<nbs, 256, 0, s>>>(ctx->N, c.bnd, dm, ref, ctx->scal, ctx->cfg.R, ctx->Nv,
                                       ctx->vlo, ctx->dv, ctx->pst);
        k_collide<T><<<nbi, 256, 0, s>>>(c, L, gs, gts, ref, ctx->scal, dt, ctx->cfg.tau,
                                         ctx->pst, ctx->part, ctx->ilist, ni);
    };
    if (is_fp32(ctx)) post((float *)ctx->gs);
    else post((double *)ctx->gs);
    toc(ST_WALL_COLLIDE);
    dm_step_done(ctx);
    tic();
    k_reduce_stage<<<kRedBlocks, 256, 0, s>>>(ctx->N, ctx->part, ctx->red);
    k_reduce<<<1, 256, 0, s>>>(ctx->red, ctx->scal, ctx->cfg.no_mass_fix);
    k_finalize<<<(ctx->N + 255) / 256, 256, 0, s>>>(c, ctx->part, ctx->scal, ctx->cfg.R,
                                                    ctx->rho, ctx->ux, ctx->uy, ctx->uz,
                                                    ctx->T, ctx->limflag);
    toc(ST_FINAL);
    return hipGetLastError() == hipSuccess ? AB_OK : AB_ERR_HIP;
}

// Times every launch shape of the transport on the current state (the
// transport writes only its scratch output gt) and keeps the fastest.
int table_autotune(ab_ctx *ctx, double dt, int reps, int cap, char (*names)[32], float *ms,
                   int *n, int *chosen)
{
    if (ctx->coef_dt != dt)
    {
        int rc = table_prepare(ctx, dt);
        if (rc) return rc;
    }
    Lat L{ctx->cfg.nx, ctx->cfg.ny, ctx->cfg.nz, ctx->S, ctx->S / kWave};
    const bool f32 = is_fp32(ctx);
    const int ns = f32 ? nshapes<float>() : nshapes<double>();
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    { ctx->err = "autotune: event creation failed"; return AB_ERR_HIP; }
    int best = -1;
    float tbest = 0.0f;
    int rc = AB_OK;
    for (int i = 0; i < ns && !rc; ++i)
    {
        const bool fits = f32 ? shape_fits<float>(i, L.nx, L.ny) : shape_fits<double>(i, L.nx, L.ny);
        if (!fits)
        {
            if (ms && i < cap) ms[i] = -1.0f;
            if (names && i < cap) snprintf(names[i], 32, "n/a");
            continue;
        }
        float tmin = 0.0f;
        for (int r = 0; r <= reps && !rc; ++r)           // r == 0 warms up
        {
            (void)hipEventRecord(e0, ctx->stream);
            rc = f32 ? apply_launch<float>(ctx, L, i) : apply_launch<double>(ctx, L, i);
            (void)hipEventRecord(e1, ctx->stream);
            float t = 0.0f;
            if (!rc && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess))
            { ctx->err = "autotune: timing failed"; rc = AB_ERR_HIP; }
            if (r == 1 || (r > 1 && t < tmin)) tmin = t;
        }
        if (rc) break;
        if (ms && i < cap) ms[i] = tmin;
        if (names && i < cap)
        {
            const ApplyShape sh = f32 ? shape_at<float>(i) : shape_at<double>(i);
            if (sh.TX > 1)
                snprintf(names[i], 32, "B%dx%d/K%d%s", sh.TX, sh.TY, sh.K,
                         sh.S == 2 ? "/S2" : (sh.S ? "/S" : ""));
            else snprintf(names[i], 32, "K%d/W%d", sh.K, sh.W);
        }
        if (best < 0 || tmin < tbest) { best = i; tbest = tmin; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    ctx->apply_shape = best;
    if (n) *n = ns;
    if (chosen) *chosen = best;
    return AB_OK;
}

int table_set_shape(ab_ctx *ctx, int shape)
{
    const int ns = is_fp32(ctx) ? nshapes<float>() : nshapes<double>();
    if (shape < -1 || shape >= ns) { ctx->err = "transport shape index out of range"; return AB_ERR_ARG; }
    const bool fits = shape < 0 || (is_fp32(ctx) ? shape_fits<float>(shape, ctx->cfg.nx, ctx->cfg.ny)
                                                 : shape_fits<double>(shape, ctx->cfg.nx, ctx->cfg.ny));
    if (!fits) { ctx->err = "transport shape does not tile this lattice"; return AB_ERR_UNSUPPORTED; }
    ctx->apply_shape = shape;
    return AB_OK;
}

}  // namespace ab
