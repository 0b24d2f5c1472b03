// k_neigh.hip -- device rebuild of the MLS neighbour lists, identical to
// updateVoxelsAndNeighbours (Initialization/NeighbourSearch.cpp:17-99).
//
// Integer counting sort of particles into the reference's voxels (count,
// exclusive scan, scatter, then each voxel's short segment sorted by particle
// index so that the order within a voxel is ascending, as the reference's
// push_back in index order gives), then one thread per particle scans the 27
// (9 in 2-D) voxels in the reference order dk, dj, di and keeps j != i with
// r^2 < radius^2, capped at AB_MAX_NEIGH.
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
                          const int *__restrict__ start, int *__restrict__ list)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int v = vox[i];
    int slot = atomicAdd(&cursor[v], 1);
    list[start[v] + slot] = i;
}

__global__ void k_sort_segments(int nvox, const int *__restrict__ start, int *__restrict__ list)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    int lo = start[v], hi = start[v + 1];
    for (int a = lo + 1; a < hi; ++a)
    {
        int key = list[a], b = a - 1;
        while (b >= lo && list[b] > key) { list[b + 1] = list[b]; --b; }
        list[b + 1] = key;
    }
}

__global__ void k_search(VoxGrid vg, int N, const double *__restrict__ x,
                         const double *__restrict__ y, const double *__restrict__ z,
                         const int *__restrict__ vox, const int *__restrict__ start,
                         const int *__restrict__ list, int *__restrict__ neigh,
                         int *__restrict__ nneigh)
{
#pragma clang fp contract(off)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double r2max = vg.radius * vg.radius;
    const int dkmax = (vg.dim == 3) ? 1 : 0;
    const int v = vox[i];
    const int vi = v % vg.nbh, tmp = v / vg.nbh;
    const int vj = tmp % vg.nbv, vk = tmp / vg.nbv;
    const double xi = x[i], yi = y[i], zi = z[i];
    int *nb = neigh + (size_t)i * AB_MAX_NEIGH;
    int nn = 0;
    bool done = false;
    for (int dk = -dkmax; dk <= dkmax && !done; dk++)
        for (int dj = -1; dj <= 1 && !done; dj++)
            for (int di = -1; di <= 1 && !done; di++)
            {
                int ni = vi + di, nj = vj + dj, nk = vk + dk;
                if (ni < 0 || ni >= vg.nbh) continue;
                if (nj < 0 || nj >= vg.nbv) continue;
                if (nk < 0 || nk >= vg.nbw) continue;
                int w = ni + vg.nbh * (nj + vg.nbv * nk);
                for (int q = start[w]; q < start[w + 1]; ++q)
                {
                    int j = list[q];
                    if (j == i) continue;
                    double ddx = x[j] - xi;
                    double ddy = y[j] - yi;
                    double ddz = (vg.dim == 3) ? z[j] - zi : 0.0;
                    if (ddx * ddx + ddy * ddy + ddz * ddz < r2max)
                    {
                        nb[nn++] = j;
                        if (nn >= AB_MAX_NEIGH) { done = true; break; }
                    }
                }
            }
    nneigh[i] = nn;
    for (int m = nn; m < AB_MAX_NEIGH; ++m) nb[m] = -1;
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
    k_scatter<<<(N + bs - 1) / bs, bs, 0, s>>>(N, vox, vcount, vstart, vlist);
    k_sort_segments<<<(nvox + bs - 1) / bs, bs, 0, s>>>(nvox, vstart, vlist);
    k_search<<<(N + bs - 1) / bs, bs, 0, s>>>(vg, N, c.x, c.y, c.z, vox, vstart, vlist, neigh,
                                              nneigh);
    return (int)hipGetLastError();
}

}  // namespace ab
