// libpd_hip: the initial crack.
//
// pd_line_crack_pairs finds the bonds that cross a straight crack, on the
// GPU (one thread per node row) or on the host: the bulk form of the
// pairwise predicates PeriPy's initial_crack_helper (model.py:1578-1612)
// calls once per bond in Python, e.g. peripy/test/conftest.py:43-63.  The
// pairs come out in the helper's order: i ascending, then i's row order,
// i < j.  pd_create_crack removes them from the neighbour list with
// create_crack.pyx's swap-with-last rule (create_crack.pyx:1-43); that pass
// is O(pairs * row length) and runs on the host.
#include <hip/hip_runtime.h>

#include <vector>

#include "pd_dev.h"

using pdsetup::Dev;

namespace {

// A straight crack in the (axis, cross) plane: a bond (x, y) crosses it when
// its end with the smaller `axis` coordinate, p1, lies below `split` and the
// other end, p2, above it (strictly), and the line through p1 and p2 meets
// coordinate[axis] = at at a cross coordinate strictly between lo and hi.
// The arithmetic is the predicate's, rounded step by step:
//   m = (p2_c - p1_c) / (p2_a - p1_a);  b = p1_c - m * p1_a;
//   height = m * at + b.
struct LineCrack {
    int axis, cross;
    double split, at, lo, hi;
};

__host__ __device__ inline bool crosses(const LineCrack &c, const double *x,
                                        const double *y)
{
    const double *p1 = x, *p2 = y;
    if (x[c.axis] > y[c.axis]) {
        p1 = y;
        p2 = x;
    }
    if (!(p1[c.axis] < c.split && p2[c.axis] > c.split))
        return false;
    const double m = (p2[c.cross] - p1[c.cross]) / (p2[c.axis] - p1[c.axis]);
    const double b = p1[c.cross] - m * p1[c.axis];
    const double height = m * c.at + b;
    return height > c.lo && height < c.hi;
}

// Crack pairs (i, j), i < j, of row i in row order: counted when `out` is
// null, else written to out[0..].
__host__ __device__ inline int row_pairs(const LineCrack &c, int n, int K,
                                         const double *x, const int *nlist,
                                         const int *n_neigh, int i, int *out)
{
    const int *row = nlist + (size_t)i * K;
    const double *xi = x + 3 * (size_t)i;
    int cnt = 0;
    for (int s = 0; s < n_neigh[i]; ++s) {
        const int j = row[s];
        if (j <= i || j >= n)
            continue;
        if (crosses(c, xi, x + 3 * (size_t)j)) {
            if (out) {
                out[2 * cnt] = i;
                out[2 * cnt + 1] = j;
            }
            ++cnt;
        }
    }
    return cnt;
}

__global__ void __launch_bounds__(256)
    k_crack_count(LineCrack c, int n, int K, const double *x,
                  const int *nlist, const int *n_neigh, int *count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        count[i] = row_pairs(c, n, K, x, nlist, n_neigh, i, nullptr);
}

__global__ void __launch_bounds__(256)
    k_crack_fill(LineCrack c, int n, int K, const double *x,
                 const int *nlist, const int *n_neigh, const long long *start,
                 int *pairs)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && start[i + 1] > start[i])
        row_pairs(c, n, K, x, nlist, n_neigh, i, pairs + 2 * start[i]);
}

int check_graph(int n, int K, const int *nlist, const int *n_neigh)
{
    if (n <= 0 || K <= 0 || !nlist || !n_neigh)
        return pd_fail(PD_E_ARG,
                       "nnodes, max_neighbours, nlist and n_neigh are "
                       "required");
    for (int i = 0; i < n; ++i)
        if (n_neigh[i] < 0 || n_neigh[i] > K)
            return pd_fail(PD_E_ARG, "n_neigh must lie in [0, max_neighbours]");
    return PD_OK;
}

}  // namespace

extern "C" int pd_line_crack_pairs(int device, int nnodes, int max_neighbours,
                                   const double *coords, const int *nlist,
                                   const int *n_neigh, int axis, int cross_axis,
                                   double split, double at, double lo,
                                   double hi, int *pairs, long long capacity,
                                   long long *npairs)
{
    if (int rc = check_graph(nnodes, max_neighbours, nlist, n_neigh))
        return rc;
    if (!coords || !npairs)
        return pd_fail(PD_E_ARG, "coords and npairs are required");
    if (axis < 0 || axis > 2 || cross_axis < 0 || cross_axis > 2 ||
        axis == cross_axis)
        return pd_fail(PD_E_ARG, "axis and cross_axis must be two of 0, 1, 2");
    const LineCrack c{axis, cross_axis, split, at, lo, hi};
    const int n = nnodes, K = max_neighbours;
    std::vector<long long> start(n + 1, 0);

    if (device < 0) {
        for (int i = 0; i < n; ++i)
            start[i + 1] = start[i] +
                           row_pairs(c, n, K, coords, nlist, n_neigh, i,
                                     nullptr);
        *npairs = start[n];
        if (!pairs)
            return PD_OK;
        if (capacity < start[n])
            return pd_fail(PD_E_LIMIT, "pairs holds fewer than npairs pairs");
        for (int i = 0; i < n; ++i)
            row_pairs(c, n, K, coords, nlist, n_neigh, i,
                      pairs + 2 * start[i]);
        return PD_OK;
    }

    if (int rc = pdsetup::use_device(device))
        return rc;
    const size_t slots = (size_t)n * K;
    Dev<double> dx;
    Dev<int> dnl, dnn, dcount, dpairs;
    Dev<long long> dstart;
    PD_SETUP_CHK(dx.alloc(3 * (size_t)n));
    PD_SETUP_CHK(dnl.alloc(slots));
    PD_SETUP_CHK(dnn.alloc(n));
    PD_SETUP_CHK(dcount.alloc(n));
    PD_SETUP_CHK(hipMemcpy(dx.p, coords, 3 * (size_t)n * sizeof(double),
                           hipMemcpyHostToDevice));
    PD_SETUP_CHK(hipMemcpy(dnl.p, nlist, slots * sizeof(int),
                           hipMemcpyHostToDevice));
    PD_SETUP_CHK(hipMemcpy(dnn.p, n_neigh, (size_t)n * sizeof(int),
                           hipMemcpyHostToDevice));
    const dim3 grid((n + 255) / 256);
    hipLaunchKernelGGL(k_crack_count, grid, dim3(256), 0, 0, c, n, K, dx.p,
                       dnl.p, dnn.p, dcount.p);
    PD_SETUP_CHK(hipGetLastError());
    std::vector<int> count(n);
    PD_SETUP_CHK(hipMemcpy(count.data(), dcount.p, (size_t)n * sizeof(int),
                           hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i)
        start[i + 1] = start[i] + count[i];
    *npairs = start[n];
    if (!pairs)
        return PD_OK;
    if (capacity < start[n])
        return pd_fail(PD_E_LIMIT, "pairs holds fewer than npairs pairs");
    if (start[n] == 0)
        return PD_OK;
    PD_SETUP_CHK(dstart.alloc(n + 1));
    PD_SETUP_CHK(dpairs.alloc(2 * (size_t)start[n]));
    PD_SETUP_CHK(hipMemcpy(dstart.p, start.data(),
                           (size_t)(n + 1) * sizeof(long long),
                           hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_crack_fill, grid, dim3(256), 0, 0, c, n, K, dx.p,
                       dnl.p, dnn.p, dstart.p, dpairs.p);
    PD_SETUP_CHK(hipGetLastError());
    PD_SETUP_CHK(hipMemcpy(pairs, dpairs.p,
                           2 * (size_t)start[n] * sizeof(int),
                           hipMemcpyDeviceToHost));
    return PD_OK;
}

// create_crack.pyx:13-43: for each pair (i, j) in order, remove j from i's
// row, then i from j's row: the first slot (within n_neigh) holding it takes
// the row's last live entry, the last live slot becomes -1, n_neigh drops by
// one; nothing happens if the row does not hold it.
extern "C" int pd_create_crack(int nnodes, int max_neighbours,
                               long long ncrack, const int *crack, int *nlist,
                               int *n_neigh)
{
    if (int rc = check_graph(nnodes, max_neighbours, nlist, n_neigh))
        return rc;
    if (ncrack < 0 || (ncrack > 0 && !crack))
        return pd_fail(PD_E_ARG, "crack (ncrack, 2) is required");
    for (long long q = 0; q < 2 * ncrack; ++q)
        if (crack[q] < 0 || crack[q] >= nnodes)
            return pd_fail(PD_E_ARG, "a crack pair names a node outside "
                                     "[0, nnodes)");
    const size_t K = (size_t)max_neighbours;
    for (long long q = 0; q < ncrack; ++q) {
        const int ends[2][2] = {{crack[2 * q], crack[2 * q + 1]},
                                {crack[2 * q + 1], crack[2 * q]}};
        for (const auto &e : ends) {
            int *row = nlist + (size_t)e[0] * K;
            const int last = n_neigh[e[0]] - 1;
            for (int s = 0; s <= last; ++s) {
                if (row[s] == e[1]) {
                    row[s] = row[last];
                    row[last] = -1;
                    n_neigh[e[0]] = last;
                    break;
                }
            }
        }
    }
    return PD_OK;
}
