// k_neigh.hip -- device rebuild of the MLS neighbour lists, identical to
// updateVoxelsAndNeighbours (Initialization/NeighbourSearch.cpp:17-99).
//
// Integer counting sort of particles into the reference's voxels: count,
// exclusive scan, scatter, then each voxel's segment ranked by particle index
// (one wave per voxel), so that the order within a voxel is ascending, as the
// reference's push_back in index order gives. The search then runs one
// workgroup per voxel: every particle of a voxel has the same 27 (9 in 2-D)
// candidate voxels, and for fixed (dk, dj) the three voxels di = -1, 0, 1 are
// consecutive in the sorted list, so the candidates form at most 9 contiguous
// runs in the reference's scan order. The workgroup stages them (ids and
// positions) in LDS once; each wave then takes one particle of the voxel and
// tests 64 candidates per pass, and a ballot appends the hits in candidate
// order, which keeps j != i with r^2 < radius^2 in the reference's order,
// capped at AB_MAX_NEIGH. A voxel whose candidates exceed the LDS stage reads
// them from memory in the same order.
//
// The FP64 voxel and distance arithmetic must round exactly as the reference
// host build does (no FMA contraction), so these kernels switch contraction
// off.
#include "ab_device.hpp"
#include "ab_internal.hpp"

namespace ab {

namespace {

struct VoxGrid
{
    int dim;
    double xleft, ybottom, zfront, hb, vb, wb;
    int nbh, nbv, nbw;
    double radius;
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__global__ void k_voxel(VoxGrid vg, int N, const double *__restrict__ x,
                        const double *__restrict__ y, const double *__restrict__ z,
                        int *__restrict__ vox, int *__restrict__ count)
{
#pragma clang fp contract(off)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int vi = clampi((int)floor((x[i] - vg.xleft) / vg.hb), 0, vg.nbh - 1);
    int vj = clampi((int)floor((y[i] - vg.ybottom) / vg.vb), 0, vg.nbv - 1);
    int vk = (vg.dim == 3) ? clampi((int)floor((z[i] - vg.zfront) / vg.wb), 0, vg.nbw - 1) : 0;
    int v = vi + vg.nbh * (vj + vg.nbv * vk);
    vox[i] = v;
    atomicAdd(&count[v], 1);
}

// Exclusive scan of the voxel counts in one workgroup; clears the counts so
// they can serve as fill cursors for the scatter.
__global__ __launch_bounds__(1024) void k_scan(int nvox, int *__restrict__ count,
                                               int *__restrict__ start)
{
    __shared__ int sh[1024];
    const int t = threadIdx.x;
    const int per = (nvox + 1023) / 1024;
    const int lo = t * per, hi = min(lo + per, nvox);
    int s = 0;
    for (int v = lo; v < hi; ++v) s += count[v];
    sh[t] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1)
    {
        int add = (t >= o) ? sh[t - o] : 0;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    int run = (t == 0) ? 0 : sh[t - 1];
    for (int v = lo; v < hi; ++v)
    {
        start[v] = run;
        run += count[v];
        count[v] = 0;
    }
    if (t == 1023) start[nvox] = sh[1023];
}

__global__ void k_scatter(int N, const int *__restrict__ vox, int *__restrict__ cursor,
                          const int *__restrict__ start, int *__restrict__ tmp)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int v = vox[i];
    int slot = atomicAdd(&cursor[v], 1);
    tmp[start[v] + slot] = i;
}

// One wave per voxel: each id of the scattered segment goes to its rank
// among the segment's ids (all distinct), which sorts the segment ascending.
__global__ __launch_bounds__(256) void k_rank_segments(int nvox, const int *__restrict__ start,
                                                       const int *__restrict__ tmp,
                                                       int *__restrict__ list)
{
    const int v = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (v >= nvox) return;                         // whole wave
    const int lo = start[v], n = start[v + 1] - lo;
    if (n <= 64)
    {
        const int key = (lane < n) ? tmp[lo + lane] : 0x7fffffff;
        int rank = 0;
        for (int b = 0; b < n; ++b) rank += (__shfl(key, b) < key);
        if (lane < n) list[lo + rank] = key;
        return;
    }
    for (int a = lane; a < n; a += 64)
    {
        const int key = tmp[lo + a];
        int rank = 0;
        for (int b = 0; b < n; ++b) rank += (tmp[lo + b] < key);
        list[lo + rank] = key;
    }
}

constexpr int kStage = 1536;                       // candidates staged in LDS

// One workgroup (4 waves) per voxel; see the file comment.
__global__ __launch_bounds__(256) void k_search(VoxGrid vg, const double *__restrict__ x,
                                                const double *__restrict__ y,
                                                const double *__restrict__ z,
                                                const int *__restrict__ start,
                                                const int *__restrict__ list,
                                                int *__restrict__ neigh,
                                                int *__restrict__ nneigh)
{
#pragma clang fp contract(off)
    __shared__ double sx[kStage], sy[kStage], sz[kStage];
    __shared__ int sj[kStage];
    __shared__ int rlo[9], roff[10], nrow_s;
    const int v = blockIdx.x;
    const int plo = start[v], phi = start[v + 1];
    if (plo == phi) return;                        // whole workgroup
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0)
    {
        const int vi = v % vg.nbh, t = v / vg.nbh;
        const int vj = t % vg.nbv, vk = t / vg.nbv;
        const int dkmax = (vg.dim == 3) ? 1 : 0;
        const int ilo = max(vi - 1, 0), ihi = min(vi + 1, vg.nbh - 1);
        int nr = 0, off = 0;
        for (int dk = -dkmax; dk <= dkmax; dk++)
            for (int dj = -1; dj <= 1; dj++)
            {
                const int nj = vj + dj, nk = vk + dk;
                if (nj < 0 || nj >= vg.nbv || nk < 0 || nk >= vg.nbw) continue;
                const int row = vg.nbh * (nj + vg.nbv * nk);
                rlo[nr] = start[ilo + row];
                roff[nr] = off;
                off += start[ihi + row + 1] - start[ilo + row];
                nr++;
            }
        roff[nr] = off;
        nrow_s = nr;
    }
    __syncthreads();
    const int nr = nrow_s, total = roff[nr];
    const bool staged = total <= kStage;
    // candidate k of the voxel's scan order: its position in the sorted list
    auto at = [&](int k) {
        int r = 0;
        while (r + 1 < nr && roff[r + 1] <= k) ++r;
        return rlo[r] + (k - roff[r]);
    };
    if (staged)
    {
        for (int k = tid; k < total; k += 256)
        {
            const int j = list[at(k)];
            sj[k] = j;
            sx[k] = x[j];
            sy[k] = y[j];
            sz[k] = (vg.dim == 3) ? z[j] : 0.0;
        }
        __syncthreads();
    }
    const double r2max = vg.radius * vg.radius;
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int p = plo + wave; p < phi; p += 4)
    {
        const int i = list[p];
        const double xi = x[i], yi = y[i], zi = (vg.dim == 3) ? z[i] : 0.0;
        int *nb = neigh + (size_t)i * AB_MAX_NEIGH;
        int nn = 0;
        for (int k0 = 0; k0 < total && nn < AB_MAX_NEIGH; k0 += 64)
        {
            const int k = k0 + lane;
            bool hit = false;
            int j = -1;
            if (k < total)
            {
                double xj, yj, zj;
                if (staged) { j = sj[k]; xj = sx[k]; yj = sy[k]; zj = sz[k]; }
                else
                {
                    j = list[at(k)];
                    xj = x[j];
                    yj = y[j];
                    zj = (vg.dim == 3) ? z[j] : 0.0;
                }
                const double ddx = xj - xi;
                const double ddy = yj - yi;
                const double ddz = (vg.dim == 3) ? zj - zi : 0.0;
                hit = (j != i) && (ddx * ddx + ddy * ddy + ddz * ddz < r2max);
            }
            const unsigned long long m = __ballot(hit);
            const int pos = nn + __popcll(m & below);
            if (hit && pos < AB_MAX_NEIGH) nb[pos] = j;
            nn = min(nn + __popcll(m), AB_MAX_NEIGH);
        }
        if (lane == 0) nneigh[i] = nn;
        for (int q = nn + lane; q < AB_MAX_NEIGH; q += 64) nb[q] = -1;
    }
}

}  // namespace

int launch_neighbours(const DevCloud &c, const ab_config &cfg, int *vox, int *vcount,
                      int *vstart, int *vlist, int nvox, int *neigh, int *nneigh,
                      hipStream_t s)
{
    VoxGrid vg;
    vg.dim = c.dim;
    vg.xleft = cfg.xleft; vg.ybottom = cfg.ybottom; vg.zfront = cfg.zfront;
    vg.hb = cfg.hbox; vg.vb = cfg.vbox; vg.wb = cfg.wbox;
    vg.nbh = cfg.nbhbox; vg.nbv = cfg.nbvbox; vg.nbw = (c.dim == 3) ? cfg.nbwbox : 1;
    vg.radius = cfg.radius;
    const int N = c.N, bs = 256;
    hipError_t e = hipMemsetAsync(vcount, 0, sizeof(int) * (size_t)nvox, s);
    if (e != hipSuccess) return (int)e;
    k_voxel<<<(N + bs - 1) / bs, bs, 0, s>>>(vg, N, c.x, c.y, c.z, vox, vcount);
    k_scan<<<1, 1024, 0, s>>>(nvox, vcount, vstart);
    int *tmp = vlist + N;                          // vlist holds 2 N ints
    k_scatter<<<(N + bs - 1) / bs, bs, 0, s>>>(N, vox, vcount, vstart, tmp);
    k_rank_segments<<<(nvox + 3) / 4, 256, 0, s>>>(nvox, vstart, tmp, vlist);
    k_search<<<nvox, 256, 0, s>>>(vg, c.x, c.y, c.z, vstart, vlist, neigh, nneigh);
    return (int)hipGetLastError();
}

}  // namespace ab
