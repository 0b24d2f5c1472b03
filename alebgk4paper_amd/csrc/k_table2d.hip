// k_table2d.hip -- table-driven transport of the 2-D Chu path
// (AB_PREC_FP64_TABLE with dim 2).
//
// The 2-D MLS transport (kTransportNode2D, Solver/CudaBackend.cu:199-300) is
// linear in g: with the normal equations A s = b, b = sum_m w_m phi_m g_m,
// solved by Gauss elimination with partial pivoting and gt = s[0], the
// elimination's steps depend on A only, so s[0] = sum_k b_k s0(e_k) and
//     gt[i, v] = sum_m C[i, m, v] g[m, v],  C = w_m sum_k phi_m[k] s0(e_k).
// On a lattice every particle's stencil is one of 5 x 5 classes (per axis:
// first, second, interior, second-to-last, last column), so the setup runs the
// reference's assembly and elimination once per (class, velocity node) on the
// class's representative particle, with the P unit vectors as right-hand
// sides, and the step applies the 21-offset stencil (the lattice offsets with
// ox^2 + oy^2 <= 6) to both layers g1 and g2. Gated boundary nodes
// (c_n >= 0) and particles without a weighted neighbour keep their value, as
// in the reference. The rest of the step is the FP64 formulation's.
#include "ab_device.hpp"
#include "ab_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

using namespace ab;

namespace {

constexpr int kOff2 = 21;
constexpr int kNcls2 = 25;

struct Off2 { int ox, oy; };

__host__ __device__ constexpr Off2 off2_at(int o)
{
    int n = 0;
    for (int oy = -2; oy <= 2; ++oy)
        for (int ox = -2; ox <= 2; ++ox)
            if (ox * ox + oy * oy <= 6)
            {
                if (n == o) return Off2{ox, oy};
                ++n;
            }
    return Off2{0, 0};
}

__host__ __device__ constexpr int off2_index(int ox, int oy)
{
    int n = 0;
    for (int y = -2; y <= 2; ++y)
        for (int x = -2; x <= 2; ++x)
            if (x * x + y * y <= 6)
            {
                if (x == ox && y == oy) return n;
                ++n;
            }
    return -1;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__host__ __device__ inline int category2(int i, int n)
{
    return i == 0 ? 0 : i == 1 ? 1 : i == n - 2 ? 3 : i == n - 1 ? 4 : 2;
}

// One (class, node): the reference's assembly for the representative
// particle i (its neighbour list, in order), then the pivoted elimination
// with the P unit vectors as right-hand sides. out: kOff2 coefficients at
// stride `stride`.
template <int P, bool BND>
__device__ void build2(const DevCloud &c, int i, int lin, double dt, double h,
                       const int *__restrict__ lat, double *out, size_t stride)
{
    const double vx = vnode(c.vlo, c.dv, lin % c.Nv);
    const double vy = vnode(c.vlo, c.dv, lin / c.Nv);
    const int self = off2_index(0, 0);
    if (BND && wall_cn2(vx, vy, c.wux[i], c.wuy[i], c.wnx[i], c.wny[i]) >= 0.0)
    {
        out[self * stride] = 1.0;
        return;
    }
    const double xf = c.x[i] - (vx - c.mvx[i]) * dt;
    const double yf = c.y[i] - (vy - c.mvy[i]) * dt;
    const int nn = c.nneigh[i];
    const int *nb = c.neigh + (size_t)i * AB_MAX_NEIGH;
    auto weight_basis = [&](int j, double *basis) -> double {
        const double dxn = c.x[j] - xf, dyn = c.y[j] - yf;
        const double r = sqrt(dxn * dxn + dyn * dyn);
        double w = 0.0;                                   // mlsWeight
        if (r < h) { const double qq = r / h; const double t1 = 1.0 - qq * qq; w = t1 * t1; }
        const double ssx = dxn / h, ssy = dyn / h;
        basis[0] = 1.0; basis[1] = ssx; basis[2] = ssy;
        if constexpr (P >= 6) { basis[3] = ssx * ssx; basis[4] = ssx * ssy; basis[5] = ssy * ssy; }
        return w;
    };
    double A[P * P], B[P * P];                            // B: the right-hand sides, B[row][k]
#pragma unroll
    for (int q = 0; q < P * P; ++q) { A[q] = 0.0; B[q] = 0.0; }
#pragma unroll
    for (int k = 0; k < P; ++k) B[k * P + k] = 1.0;
    bool has = false;
    for (int m = 0; m < nn; ++m)
    {
        double basis[P];
        const double w = weight_basis(nb[m], basis);
        if (w < 1e-30) continue;
#pragma unroll
        for (int p1 = 0; p1 < P; ++p1)
#pragma unroll
            for (int p2 = 0; p2 < P; ++p2) A[p1 * P + p2] += w * basis[p1] * basis[p2];
        has = true;
    }
    if (!has)
    {
        out[self * stride] = 1.0;
        return;
    }
#pragma unroll
    for (int col = 0; col < P; ++col)
    {
        int piv = col;
        double mx = fabs(A[col * P + col]);
#pragma unroll
        for (int row = col + 1; row < P; ++row)
            if (fabs(A[row * P + col]) > mx) { mx = fabs(A[row * P + col]); piv = row; }
#pragma unroll
        for (int row = col + 1; row < P; ++row)
            if (row == piv)
            {
#pragma unroll
                for (int q = 0; q < P; ++q)
                {
                    double t1 = A[col * P + q]; A[col * P + q] = A[row * P + q]; A[row * P + q] = t1;
                    t1 = B[col * P + q]; B[col * P + q] = B[row * P + q]; B[row * P + q] = t1;
                }
            }
        const double d = A[col * P + col];
        if (fabs(d) < 1e-30) continue;
#pragma unroll
        for (int row = col + 1; row < P; ++row)
        {
            const double f = A[row * P + col] / d;
#pragma unroll
            for (int q = col; q < P; ++q) A[row * P + q] -= f * A[col * P + q];
#pragma unroll
            for (int k = 0; k < P; ++k) B[row * P + k] -= f * B[col * P + k];
        }
    }
    // back substitution, every right-hand side; only s[0] is kept
    double s[P * P];
#pragma unroll
    for (int row = P - 1; row >= 0; --row)
    {
        const double d = A[row * P + row];
#pragma unroll
        for (int k = 0; k < P; ++k)
        {
            double v = B[row * P + k];
#pragma unroll
            for (int q = row + 1; q < P; ++q) v -= A[row * P + q] * s[q * P + k];
            s[row * P + k] = fabs(d) > 1e-30 ? v / d : 0.0;
        }
    }
    for (int m = 0; m < nn; ++m)
    {
        const int j = nb[m];
        double basis[P];
        const double w = weight_basis(j, basis);
        if (w < 1e-30) continue;
        double cm = 0.0;
#pragma unroll
        for (int k = 0; k < P; ++k) cm += basis[k] * s[k];   // s[0 * P + k] = s0(e_k)
        const int o = off2_index(lat[4 * j] - lat[4 * i], lat[4 * j + 1] - lat[4 * i + 1]);
        if (o >= 0) out[o * stride] += w * cm;
    }
}

// Coefficients [class][offset][node]; classes of several instances use
// their instance's support radius.
__global__ void k_build2(DevCloud c, const int *__restrict__ rep, const int *__restrict__ lat,
                         double dt, double h, int order, double *__restrict__ coef)
{
    const int cls = blockIdx.y;
    const int lin = blockIdx.x * blockDim.x + threadIdx.x;
    if (lin >= c.nodes) return;
    const size_t stride = c.nodes;
    double *out = coef + (size_t)cls * kOff2 * stride + lin;
    for (int o = 0; o < kOff2; ++o) out[o * stride] = 0.0;
    const int i = rep[cls];
    if (i < 0) return;
    if (c.inst) h = c.ih[c.inst[i]];
    if (c.bnd[i]) build2<3, true>(c, i, lin, dt, h, lat, out, stride);
    else if (order >= 2) build2<6, false>(c, i, lin, dt, h, lat, out, stride);
    else build2<3, false>(c, i, lin, dt, h, lat, out, stride);
}

// Offset index of (ox, oy), -1 outside the stencil, as a table the unrolled
// loops below fold to constants.
struct OffTab
{
    int idx[5][5];
};
__host__ __device__ constexpr OffTab make_offtab()
{
    OffTab t{};
    for (int oy = -2; oy <= 2; ++oy)
        for (int ox = -2; ox <= 2; ++ox) t.idx[oy + 2][ox + 2] = off2_index(ox, oy);
    return t;
}
constexpr OffTab kOffTab = make_offtab();

// Run lengths of the apply (outputs per work item along a lattice line):
// longer runs load fewer input rows per output ((RUN + 4) / RUN), shorter
// ones hold fewer accumulators and give more items. Each output's sum takes
// the same terms in the same order for every run length.
constexpr int kRuns[] = {4, 8, 12};
constexpr int kNRuns = 3;
constexpr int kRunDefault = 1;   // 8 outputs

// The row offset `ro` passes through an empty asm that takes every
// accumulator of the run, so the loads of a row issue after the additions
// of the rows before it.
template <int RUN>
__device__ __forceinline__ void pin_after(int &ro, const double *a1, const double *a2)
{
    if constexpr (RUN == 4)
        asm volatile("" : "+v"(ro)
                     : "v"(a1[0]), "v"(a1[1]), "v"(a1[2]), "v"(a1[3]), "v"(a2[0]), "v"(a2[1]),
                       "v"(a2[2]), "v"(a2[3]));
    else if constexpr (RUN == 8)
        asm volatile("" : "+v"(ro)
                     : "v"(a1[0]), "v"(a1[1]), "v"(a1[2]), "v"(a1[3]), "v"(a1[4]), "v"(a1[5]),
                       "v"(a1[6]), "v"(a1[7]), "v"(a2[0]), "v"(a2[1]), "v"(a2[2]), "v"(a2[3]),
                       "v"(a2[4]), "v"(a2[5]), "v"(a2[6]), "v"(a2[7]));
    else
    {
        static_assert(RUN == 12, "pins exist for runs of 4, 8 and 12 outputs");
        asm volatile("" : "+v"(ro)
                     : "v"(a1[0]), "v"(a1[1]), "v"(a1[2]), "v"(a1[3]), "v"(a1[4]), "v"(a1[5]),
                       "v"(a1[6]), "v"(a1[7]), "v"(a1[8]), "v"(a1[9]), "v"(a1[10]), "v"(a1[11]));
        asm volatile("" : "+v"(ro)
                     : "v"(a2[0]), "v"(a2[1]), "v"(a2[2]), "v"(a2[3]), "v"(a2[4]), "v"(a2[5]),
                       "v"(a2[6]), "v"(a2[7]), "v"(a2[8]), "v"(a2[9]), "v"(a2[10]), "v"(a2[11]));
    }
}

// A run of up to RUN outputs of one class along a lattice line (HOR: along
// x, else along y), every output at least two sites from the line's ends,
// so the inputs along the line stay inside it (an input past the last site is
// clamped: it feeds only outputs that are not stored). Across the line the
// five inputs are clamped into the lattice; their coefficients are 0 there.
// Each input row is loaded once and added into the outputs whose stencil
// holds it, with the class's 21 coefficients in registers. Column runs sum
// each output in stencil order (oy outer, ox inner), as the per-particle
// path does.
// Item: p0, along stride, count, last along offset, across position, across
// size, across stride.
template <bool HOR, int RUN>
__device__ __forceinline__ void run2(const int *__restrict__ it, int lin, int nodes, int gstride,
                                     const int *__restrict__ cls, const double *__restrict__ coef,
                                     const double *__restrict__ g, double *__restrict__ gt)
{
    const int p0 = it[0], as = it[1], cnt = it[2], amax = it[3], pos = it[4], nac = it[5],
              cs = it[6];
    const double *cf = coef + (size_t)cls[p0] * kOff2 * nodes + lin;
    double w[kOff2];
#pragma unroll
    for (int o = 0; o < kOff2; ++o) w[o] = cf[(size_t)o * nodes];
    // g through a buffer resource with 32-bit byte offsets (the setup
    // checks that g spans less than 2 GiB); the second layer via soffset
    const __amdgpu_buffer_rsrc_t rg =
        __builtin_amdgcn_make_buffer_rsrc((void *)g, (short)0, 0x7fffffff, 0x00020000);
    const int l2 = nodes * 8;
    int cof[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) cof[c] = (clampi(pos + c - 2, 0, nac - 1) - pos) * cs * gstride * 8;
    double a1[RUN], a2[RUN];
#pragma unroll
    for (int r = 0; r < RUN; ++r) { a1[r] = 0.0; a2[r] = 0.0; }
    // One input row in flight ahead of the row being added: a row's offset
    // is pinned after the additions of the rows before the current one
    // (without it the compiler hoists all rows' loads and runs at one wave
    // per SIMD).
    double x1[5], x2[5], y1[5], y2[5];
    auto load = [&](int v, double *z1, double *z2) {
        const int d = v - 2 < amax ? v - 2 : amax;
        int ro = ((p0 + d * as) * gstride + lin) * 8;
        pin_after<RUN>(ro, a1, a2);
#pragma unroll
        for (int c = 0; c < 5; ++c)
        {
            z1[c] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rg, ro + cof[c], 0, 0));
            z2[c] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rg, ro + cof[c], l2, 0));
        }
    };
    load(0, x1, x2);
#pragma unroll
    for (int v = 0; v < RUN + 4; ++v)
    {
        if (v + 1 < RUN + 4) load(v + 1, y1, y2);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int r = 0; r < RUN; ++r)
        {
            const int dv = v - 2 - r;
            if (dv < -2 || dv > 2) continue;
#pragma unroll
            for (int c = 0; c < 5; ++c)
            {
                const int o = HOR ? kOffTab.idx[c][dv + 2] : kOffTab.idx[dv + 2][c];
                if (o < 0) continue;
                a1[r] += w[o] * x1[c];
                a2[r] += w[o] * x2[c];
            }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int c = 0; c < 5; ++c) { x1[c] = y1[c]; x2[c] = y2[c]; }
    }
#pragma unroll
    for (int r = 0; r < RUN; ++r)
        if (r < cnt)
        {
            double *out = gt + ((long)p0 + (long)r * as) * gstride + lin;
            out[0] = a1[r];
            out[nodes] = a2[r];
        }
}

// One particle (the lattice corners): its class stencil with clamped
// offsets, in stencil order.
__device__ __forceinline__ void point2(int i, int lin, int nodes, int gstride,
                                       const int *__restrict__ lat, const int *__restrict__ cls,
                                       const double *__restrict__ coef,
                                       const double *__restrict__ g, double *__restrict__ gt)
{
    const int4 l = reinterpret_cast<const int4 *>(lat)[i];   // ix, iy, nx, ny
    const int ix = l.x, iy = l.y, nx = l.z, ny = l.w;
    const size_t base = (size_t)i - ix - (size_t)nx * iy;
    const double *cf = coef + (size_t)cls[i] * kOff2 * nodes + lin;
    double a1 = 0.0, a2 = 0.0;
#pragma unroll
    for (int o = 0; o < kOff2; ++o)
    {
        const Off2 d = off2_at(o);
        const int jx = clampi(ix + d.ox, 0, nx - 1), jy = clampi(iy + d.oy, 0, ny - 1);
        const double *gj = g + (base + jx + (size_t)nx * jy) * gstride + lin;
        const double w = cf[(size_t)o * nodes];
        a1 += w * gj[0];
        a2 += w * gj[nodes];
    }
    gt[(size_t)i * gstride + lin] = a1;
    gt[(size_t)i * gstride + lin + nodes] = a2;
}

// gt = the class stencils applied to both layers. One thread per (item,
// node), S threads per item (the nodes rounded up to whole waves, so each
// wave takes one kind of item): column runs, then row runs, then corners.
template <int RUN>
__global__ __launch_bounds__(256) void k_apply2(int nv, int nh, int nit, int S, int nodes,
                                                int gstride, const int *__restrict__ items,
                                                const int *__restrict__ lat,
                                                const int *__restrict__ cls,
                                                const double *__restrict__ coef,
                                                const double *__restrict__ g,
                                                double *__restrict__ gt)
{
    // each XCD takes a contiguous range of items, so the runs of
    // neighbouring columns that share input rows share its L2
    const long e = (long)xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int item = (int)(e / S), lin = (int)(e - (long)item * S);
    if (item >= nit || lin >= nodes) return;
    const int *it = items + 8 * (size_t)item;
    if (item < nv) run2<false, RUN>(it, lin, nodes, gstride, cls, coef, g, gt);
    else if (item < nv + nh) run2<true, RUN>(it, lin, nodes, gstride, cls, coef, g, gt);
    else point2(it[0], lin, nodes, gstride, lat, cls, coef, g, gt);
}

// The lattice instances of a 2-D context: particle ranges and sizes.
struct Inst2
{
    std::vector<int> base, nx, ny;
    std::vector<double> h;
};

Inst2 instances2(const ab_ctx *c)
{
    if (c->nbatch > 0) return Inst2{c->bt_base, c->bt_nx, c->bt_ny, c->bt_h};
    return Inst2{{0, c->N}, {c->cfg.nx}, {c->cfg.ny}, {c->cfg.radius}};
}

// The apply's work items for runs of `run` outputs: column runs over rows
// 2 .. ny-3 of every column, row runs over columns 2 .. nx-3 of rows 0, 1,
// ny-2, ny-1, and the 16 corners of each instance.
std::vector<int> apply_items2(const Inst2 &L, int run, int &nv, int &nh)
{
    std::vector<int> itv, ith, itc;
    for (size_t k = 0; k < L.nx.size(); ++k)
    {
        const int nx = L.nx[k], ny = L.ny[k], b0 = L.base[k];
        for (int ix = 0; ix < nx; ++ix)
            for (int iy = 2; iy <= ny - 3; iy += run)
            {
                const int it[8] = {b0 + ix + nx * iy, nx, std::min(run, ny - 2 - iy), ny - 1 - iy,
                                   ix, nx, 1, 0};
                itv.insert(itv.end(), it, it + 8);
            }
        for (int iy : {0, 1, ny - 2, ny - 1})
        {
            for (int ix = 2; ix <= nx - 3; ix += run)
            {
                const int it[8] = {b0 + ix + nx * iy, 1, std::min(run, nx - 2 - ix), nx - 1 - ix,
                                   iy, ny, nx, 0};
                ith.insert(ith.end(), it, it + 8);
            }
            for (int ix : {0, 1, nx - 2, nx - 1})
            {
                const int it[8] = {b0 + ix + nx * iy, 0, 0, 0, 0, 0, 0, 0};
                itc.insert(itc.end(), it, it + 8);
            }
        }
    }
    nv = (int)itv.size() / 8;
    nh = (int)ith.size() / 8;
    std::vector<int> items(itv);
    items.insert(items.end(), ith.begin(), ith.end());
    items.insert(items.end(), itc.begin(), itc.end());
    return items;
}

// Builds and uploads the items of run length index `run` (-1: the default)
// into a fresh device array (the old one is freed after the stream drains).
int upload_items2(ab_ctx *c, int run)
{
    int nv = 0, nh = 0;
    const std::vector<int> items =
        apply_items2(instances2(c), kRuns[run < 0 ? kRunDefault : run], nv, nh);
    int *d = nullptr;
    if (hipMalloc((void **)&d, sizeof(int) * items.size()) != hipSuccess)
    { c->err = "2-D table setup: device allocation failed"; return AB_ERR_NOMEM; }
    if (hipMemcpyAsync(d, items.data(), sizeof(int) * items.size(), hipMemcpyHostToDevice,
                       c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
    {
        (void)hipFree(d);
        c->err = "2-D table setup: host-to-device copy failed";
        return AB_ERR_HIP;
    }
    if (c->t2_items) (void)hipFree(c->t2_items);
    c->t2_items = d;
    c->t2_nv = nv;
    c->t2_nh = nh;
    c->t2_nit = (int)items.size() / 8;
    c->t2_run = run;
    c->tgen++;
    return AB_OK;
}

template <int RUN>
int apply2_launch(ab_ctx *c)
{
    const int S = (c->nodes + 63) / 64 * 64;
    const long n = (long)c->t2_nit * S;
    k_apply2<RUN><<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(
        c->t2_nv, c->t2_nh, c->t2_nit, S, c->nodes, c->gstride, c->t2_items, c->t2_lat, c->t2_cls,
        c->t2_coef, c->g, c->gt);
    if (hipGetLastError() != hipSuccess) { c->err = "2-D table apply launch failed"; return AB_ERR_HIP; }
    return AB_OK;
}

int apply2(ab_ctx *c)
{
    switch (c->t2_run < 0 ? kRunDefault : c->t2_run)
    {
    case 0: return apply2_launch<kRuns[0]>(c);
    case 2: return apply2_launch<kRuns[2]>(c);
    default: return apply2_launch<kRuns[1]>(c);
    }
}

}  // namespace

namespace ab {

void table2d_free(ab_ctx *c)
{
    c->tgen++;
    void *ptrs[] = {c->t2_coef, c->t2_cls, c->t2_rep, c->t2_lat, c->t2_items, c->t2_pnew};
    for (void *p : ptrs) if (p) (void)hipFree(p);
    c->t2_coef = nullptr; c->t2_cls = nullptr; c->t2_rep = nullptr; c->t2_lat = nullptr;
    c->t2_items = nullptr; c->t2_pnew = nullptr;
    c->t2_ncls = 0; c->t2_nv = c->t2_nh = c->t2_nit = 0;
    c->t2_dt = -1.0;
}

// Classes of the 2-D lattice clouds (one, or each instance of a batch) and
// their checks: every particle has its class representative's neighbour
// offsets and displacements (to 1e-12 of the support radius; the list order
// may differ) and wall state. Run after the upload and after every change of
// the cloud's geometry, lists or wall state.
int table2d_setup(ab_ctx *c)
{
    const ab_config &cfg = c->cfg;
    const int N = c->N;
    const Inst2 inst = instances2(c);
    const std::vector<int> &base = inst.base, &nxs = inst.nx, &nys = inst.ny;
    const std::vector<double> &hs = inst.h;
    const int ninst = (int)nxs.size();
    table2d_free(c);
    for (int k = 0; k < ninst; ++k)
        if (nxs[k] < 5 || nys[k] < 5 || (long)nxs[k] * nys[k] != base[k + 1] - base[k])
        { c->err = "the 2-D table transport needs lattice clouds: nx, ny >= 5 and N = nx*ny"; return AB_ERR_UNSUPPORTED; }
    if (cfg.mls_clamp)
    { c->err = "mls_clamp is nonlinear in g; run it with AB_PREC_FP64"; return AB_ERR_UNSUPPORTED; }
    if ((long long)N * c->gstride * 8 >= (1LL << 31))
    { c->err = "the 2-D table transport addresses g with 32-bit offsets: N * 2 Nv^2 * 8 bytes must stay below 2 GiB"; return AB_ERR_UNSUPPORTED; }
    std::vector<unsigned char> b(N);
    std::vector<int> nb((size_t)N * AB_MAX_NEIGH), nn(N);
    std::vector<double> w(7 * (size_t)N), x(N), y(N);
    hipStream_t s = c->stream;
    auto dl = [&](void *dst, const void *src, size_t n) {
        return hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, s) == hipSuccess;
    };
    bool ok = dl(b.data(), c->bnd, N) && dl(nb.data(), c->neigh, 4 * (size_t)N * AB_MAX_NEIGH) &&
              dl(nn.data(), c->nneigh, 4 * (size_t)N) && dl(x.data(), c->x, 8 * (size_t)N) &&
              dl(y.data(), c->y, 8 * (size_t)N);
    const double *ws[7] = {c->wnx, c->wny, c->wux, c->wuy, c->wt, c->mvx, c->mvy};
    for (int k = 0; k < 7 && ok; ++k) ok = dl(&w[(size_t)k * N], ws[k], 8 * (size_t)N);
    if (!ok || hipStreamSynchronize(s) != hipSuccess)
    { c->err = "2-D table setup: device-to-host copy failed"; return AB_ERR_HIP; }
    const int ncls = kNcls2 * ninst;
    std::vector<int> cls(N), rep(ncls, -1), lat(4 * (size_t)N);
    std::vector<unsigned> mask(N, 0);
    for (int k = 0; k < ninst; ++k)
    {
        const int nx = nxs[k], ny = nys[k];
        for (int p = base[k]; p < base[k + 1]; ++p)
        {
            const int q = p - base[k], ix = q % nx, iy = q / nx;
            lat[4 * p] = ix; lat[4 * p + 1] = iy; lat[4 * p + 2] = nx; lat[4 * p + 3] = ny;
            cls[p] = kNcls2 * k + category2(ix, nx) + 5 * category2(iy, ny);
            if (rep[cls[p]] < 0) rep[cls[p]] = p;
        }
    }
    for (int k = 0; k < ninst; ++k)
        for (int p = base[k]; p < base[k + 1]; ++p)
            for (int m = 0; m < nn[p]; ++m)
            {
                const int j = nb[(size_t)p * AB_MAX_NEIGH + m];
                const int o = (j < base[k] || j >= base[k + 1])
                                  ? -1 : off2_index(lat[4 * j] - lat[4 * p], lat[4 * j + 1] - lat[4 * p + 1]);
                if (o < 0)
                {
                    c->err = "2-D table setup: particle " + std::to_string(p) +
                             " has a neighbour outside the lattice stencil";
                    return AB_ERR_UNSUPPORTED;
                }
                mask[p] |= 1u << o;
            }
    for (int k = 0; k < ninst; ++k)
    {
        const double tol = 1e-12 * hs[k];
        for (int p = base[k]; p < base[k + 1]; ++p)
        {
            const int r = rep[cls[p]];
            bool same = mask[p] == mask[r] && nn[p] == nn[r] && b[p] == b[r];
            for (int q = 0; q < 7 && same; ++q) same = w[(size_t)q * N + p] == w[(size_t)q * N + r];
            // the neighbours at the same displacements (the representative
            // has a neighbour at every lattice offset p has: the masks are equal)
            for (int m = 0; m < nn[p] && same; ++m)
            {
                const int jp = nb[(size_t)p * AB_MAX_NEIGH + m], jr = r + (jp - p);
                same = std::fabs((x[jp] - x[p]) - (x[jr] - x[r])) <= tol &&
                       std::fabs((y[jp] - y[p]) - (y[jr] - y[r])) <= tol;
            }
            if (!same)
            {
                c->err = "2-D table setup: particle " + std::to_string(p) +
                         " differs from its lattice class (stencil, geometry or wall state)";
                return AB_ERR_UNSUPPORTED;
            }
        }
    }
    if (hipMalloc((void **)&c->t2_coef, sizeof(double) * ncls * kOff2 * (size_t)c->nodes) != hipSuccess ||
        hipMalloc((void **)&c->t2_cls, sizeof(int) * (size_t)N) != hipSuccess ||
        hipMalloc((void **)&c->t2_rep, sizeof(int) * ncls) != hipSuccess ||
        hipMalloc((void **)&c->t2_lat, sizeof(int) * 4 * (size_t)N) != hipSuccess ||
        hipMalloc((void **)&c->t2_pnew, sizeof(double) * (size_t)N) != hipSuccess)
    { table2d_free(c); c->err = "2-D table setup: device allocation failed"; return AB_ERR_NOMEM; }
    c->t2_ncls = ncls;
    if (hipMemcpyAsync(c->t2_cls, cls.data(), sizeof(int) * N, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(c->t2_rep, rep.data(), sizeof(int) * ncls, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(c->t2_lat, lat.data(), sizeof(int) * 4 * (size_t)N, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
    { table2d_free(c); c->err = "2-D table setup: host-to-device copy failed"; return AB_ERR_HIP; }
    const int rc = upload_items2(c, c->t2_run);
    if (rc) table2d_free(c);
    return rc;
}

int table2d_prepare(ab_ctx *c, double dt)
{
    if (!c->t2_coef) { c->err = "2-D table transport before its setup"; return AB_ERR_STATE; }
    dim3 gb((c->nodes + 63) / 64, c->t2_ncls);
    k_build2<<<gb, 64, 0, c->stream>>>(c->view(), c->t2_rep, c->t2_lat, dt, c->cfg.radius,
                                       c->cfg.mls_order, c->t2_coef);
    if (hipGetLastError() != hipSuccess) { c->err = "2-D table build launch failed"; return AB_ERR_HIP; }
    c->t2_dt = dt;
    return AB_OK;
}

// gt from g: the coefficients for dt (built when dt changes), then the apply.
int table2d_transport(ab_ctx *c, double dt)
{
    int rc;
    if (c->t2_dt != dt && (rc = table2d_prepare(c, dt))) return rc;
    return apply2(c);
}

// Times the apply at every run length (`reps` launches after a warm-up; the
// apply writes only gt) and keeps the fastest. New items invalidate the
// graphs captured on the context before.
int table2d_autotune(ab_ctx *c, double dt, int reps, int cap, char (*names)[32], float *ms,
                     int *n, int *chosen)
{
    int rc;
    if (!c->t2_coef) { c->err = "2-D table transport before its setup"; return AB_ERR_STATE; }
    if (c->t2_dt != dt && (rc = table2d_prepare(c, dt))) return rc;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) { c->err = "autotune: event creation failed"; return AB_ERR_HIP; }
    if (hipEventCreate(&e1) != hipSuccess)
    {
        (void)hipEventDestroy(e0);
        c->err = "autotune: event creation failed";
        return AB_ERR_HIP;
    }
    int best = -1;
    float tbest = 0.0f;
    rc = AB_OK;
    for (int i = 0; i < kNRuns && !rc; ++i)
    {
        rc = upload_items2(c, i);
        float tmin = 0.0f;
        for (int r = 0; r <= reps && !rc; ++r)           // r == 0 warms up
        {
            (void)hipEventRecord(e0, c->stream);
            rc = apply2(c);
            (void)hipEventRecord(e1, c->stream);
            float t = 0.0f;
            if (!rc && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess))
            { c->err = "autotune: timing failed"; rc = AB_ERR_HIP; }
            if (r == 1 || (r > 1 && t < tmin)) tmin = t;
        }
        if (rc) break;
        if (ms && i < cap) ms[i] = tmin;
        if (names && i < cap) snprintf(names[i], 32, "R%d", kRuns[i]);
        if (best < 0 || tmin < tbest) { best = i; tbest = tmin; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!rc) rc = upload_items2(c, best);
    if (rc) return rc;
    if (n) *n = kNRuns;
    if (chosen) *chosen = best;
    return AB_OK;
}

int table2d_set_run(ab_ctx *c, int run)
{
    if (run < -1 || run >= kNRuns) { c->err = "transport shape index out of range"; return AB_ERR_ARG; }
    if (!c->t2_coef) { c->t2_run = run; return AB_OK; }   // taken at the setup
    if (run == c->t2_run) return AB_OK;
    return upload_items2(c, run);
}

}  // namespace ab
