// libpd_hip: the neighbour list of a point cloud on the GPU.
//
// Same result as pd_neighbour_list (the host cell-list search): node j is a
// neighbour of i (j != i) when ((dx*dx + dy*dy) + dz*dz) <= horizon^2,
// d = x_j - x_i, each product and sum rounded (no contraction); rows sorted
// ascending by j and padded with pad_value.  The host bins the nodes into
// cells of edge `horizon` (a counting sort, O(N)); the GPU then counts each
// node's neighbours over its 27 cells, fills the rows, and sorts every row
// in LDS.  This replaces the KD-tree query of model.py:536-597 for setup.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "pd_dev.h"

using pdsetup::Dev;

namespace {

struct Grid {
    double lo[3];
    double h, h2;
    int dim[3];
};

__device__ __forceinline__ int cell_coord(double x, double lo, double h,
                                          int dim)
{
    const double c = floor((x - lo) / h);
    return (int)fmin(fmax(c, 0.0), (double)(dim - 1));
}

// Visit every other node j within the horizon of node i, in cell order and,
// within a cell, ascending j.
template <typename F>
__device__ __forceinline__ void for_neighbours(const Grid &g,
                                               const double *x,
                                               const long long *cstart,
                                               const int *members, int i,
                                               F &&fn)
{
    const double xi = x[3 * (size_t)i], yi = x[3 * (size_t)i + 1],
                 zi = x[3 * (size_t)i + 2];
    const int c0 = cell_coord(xi, g.lo[0], g.h, g.dim[0]);
    const int c1 = cell_coord(yi, g.lo[1], g.h, g.dim[1]);
    const int c2 = cell_coord(zi, g.lo[2], g.h, g.dim[2]);
    for (int a = max(0, c0 - 1); a <= min(g.dim[0] - 1, c0 + 1); ++a)
        for (int b = max(0, c1 - 1); b <= min(g.dim[1] - 1, c1 + 1); ++b)
            for (int e = max(0, c2 - 1); e <= min(g.dim[2] - 1, c2 + 1);
                 ++e) {
                const long long c =
                    ((long long)a * g.dim[1] + b) * g.dim[2] + e;
                for (long long m = cstart[c]; m < cstart[c + 1]; ++m) {
                    const int j = members[m];
                    if (j == i)
                        continue;
                    double d0 = x[3 * (size_t)j] - xi;
                    double d1 = x[3 * (size_t)j + 1] - yi;
                    double d2 = x[3 * (size_t)j + 2] - zi;
                    d0 = d0 * d0;
                    d1 = d1 * d1;
                    d2 = d2 * d2;
                    if ((d0 + d1) + d2 <= g.h2)
                        fn(j);
                }
            }
}

__global__ void k_nl_count(int n, Grid g, const double *x,
                           const long long *cstart, const int *members,
                           int *family)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    int c = 0;
    for_neighbours(g, x, cstart, members, i, [&](int) { ++c; });
    family[i] = c;
}

// Row i of the (n, K) list: its neighbours sorted ascending (insertion sort
// in the thread's own LDS slice, stride K+1 words so that the lanes' slices
// start in different banks), then pad_value.  family[i] <= K was checked.
__global__ void k_nl_fill(int n, Grid g, const double *x,
                          const long long *cstart, const int *members,
                          const int *family, int K, int pad, int *nlist)
{
    extern __shared__ int rows[];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    int *r = rows + (size_t)threadIdx.x * (K + 1);
    int f = 0;
    for_neighbours(g, x, cstart, members, i, [&](int j) {
        int p = f++;
        while (p > 0 && r[p - 1] > j) {
            r[p] = r[p - 1];
            --p;
        }
        r[p] = j;
    });
    int *dst = nlist + (size_t)i * K;
    for (int p = 0; p < K; ++p)
        dst[p] = p < f ? r[p] : pad;
    (void)family;
}

}  // namespace


extern "C" int pd_neighbour_list_device(int device, int n, const double *x,
                                        double horizon, int *family,
                                        int *nlist, int max_neighbours,
                                        int pad_value)
{
    if (n <= 0 || !x || !family)
        return pd_fail(PD_E_ARG, "nnodes, coords and family are required");
    if (!(horizon > 0.0))
        return pd_fail(PD_E_ARG, "horizon must be > 0");
    if (nlist && (max_neighbours <= 0 || max_neighbours > 4096))
        return pd_fail(PD_E_LIMIT, "max_neighbours must be in 1..4096");
    if (int rc = pdsetup::use_device(device))
        return rc;

    // cells of edge `horizon` over the bounding box, as the host search
    Grid g;
    double hi[3];
    for (int d = 0; d < 3; ++d)
        g.lo[d] = hi[d] = x[d];
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            g.lo[d] = std::min(g.lo[d], x[3 * (size_t)i + d]);
            hi[d] = std::max(hi[d], x[3 * (size_t)i + d]);
        }
    long long total = 1;
    for (int d = 0; d < 3; ++d) {
        long dim = std::max(1L, (long)std::floor((hi[d] - g.lo[d]) / horizon)
                                    + 1);
        dim = std::min(dim, 1L << 20);
        g.dim[d] = (int)dim;
        total *= dim;
        if (total > (1LL << 28))
            return pd_fail(PD_E_LIMIT, "cell grid too large for this horizon");
    }
    g.h = horizon;
    g.h2 = horizon * horizon;
    auto cell_of = [&](int i, int d) {
        long c = (long)std::floor((x[3 * (size_t)i + d] - g.lo[d]) / horizon);
        return std::min(std::max(c, 0L), (long)g.dim[d] - 1);
    };
    std::vector<long long> cstart(total + 1, 0);
    std::vector<long long> cid(n);
    for (int i = 0; i < n; ++i) {
        cid[i] = (cell_of(i, 0) * g.dim[1] + cell_of(i, 1)) * g.dim[2] +
                 cell_of(i, 2);
        cstart[cid[i] + 1]++;
    }
    for (long long c = 0; c < total; ++c)
        cstart[c + 1] += cstart[c];
    std::vector<int> members(n);
    {
        std::vector<long long> fillp(cstart.begin(), cstart.end() - 1);
        for (int i = 0; i < n; ++i)
            members[fillp[cid[i]]++] = i;
    }

    Dev<double> dx;
    Dev<long long> dcs;
    Dev<int> dm, dfam;
    PD_SETUP_CHK(dx.alloc(3 * (size_t)n));
    PD_SETUP_CHK(dcs.alloc(cstart.size()));
    PD_SETUP_CHK(dm.alloc(n));
    PD_SETUP_CHK(dfam.alloc(n));
    PD_SETUP_CHK(hipMemcpy(dx.p, x, 3 * (size_t)n * sizeof(double),
                    hipMemcpyHostToDevice));
    PD_SETUP_CHK(hipMemcpy(dcs.p, cstart.data(), cstart.size() * sizeof(long long),
                    hipMemcpyHostToDevice));
    PD_SETUP_CHK(hipMemcpy(dm.p, members.data(), (size_t)n * sizeof(int),
                    hipMemcpyHostToDevice));
    const int blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_nl_count, dim3(blocks), dim3(256), 0, 0, n, g, dx.p,
                       dcs.p, dm.p, dfam.p);
    PD_SETUP_CHK(hipGetLastError());
    PD_SETUP_CHK(hipMemcpy(family, dfam.p, (size_t)n * sizeof(int),
                    hipMemcpyDeviceToHost));
    if (!nlist)
        return PD_OK;
    const int fmax = *std::max_element(family, family + n);
    if (fmax > max_neighbours)
        return pd_fail(PD_E_LIMIT, "a node has more than max_neighbours bonds");
    const int K = max_neighbours;
    // LDS per block: threads x (K+1) ints, at most 64 KiB
    const int tpb = std::max(1, std::min(64, 16384 / (K + 1)));
    Dev<int> dnl;
    PD_SETUP_CHK(dnl.alloc((size_t)n * K));
    hipLaunchKernelGGL(k_nl_fill, dim3((n + tpb - 1) / tpb), dim3(tpb),
                       (size_t)tpb * (K + 1) * sizeof(int), 0, n, g, dx.p,
                       dcs.p, dm.p, dfam.p, K, pad_value, dnl.p);
    PD_SETUP_CHK(hipGetLastError());
    PD_SETUP_CHK(hipMemcpy(nlist, dnl.p, (size_t)n * K * sizeof(int),
                    hipMemcpyDeviceToHost));
    return PD_OK;
}
