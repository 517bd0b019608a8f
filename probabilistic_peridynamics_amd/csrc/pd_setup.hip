// libpd_hip: set-up arrays from built-in predicates.
//
// PeriPy builds a model's bond types and per-node arrays (boundary
// conditions, densities, tips) by calling Python predicates once per bond
// or node (model.py:657-780, 1093-1212): 1e8 calls on a million-node mesh.
// Two common families of predicates are evaluated here instead, on the GPU
// or on the host (device -1), with the same arithmetic either way:
//
//   pd_plane_bond_types  a body split by the plane x[axis] = position: a bond
//                        whose ends both lie below it (x < position) has type
//                        `below`, both at or above it type `above`, else
//                        `across` (the two-material composite's bond types,
//                        an interface type for bonds that cross);
//   pd_first_box         for every node, the index of the first of a list of
//                        axis-aligned boxes that contains it (-1 if none):
//                        from it the caller looks its boundary condition,
//                        density or tip up in a table.
#include <hip/hip_runtime.h>

#include <vector>

#include "pd_dev.h"

using pdsetup::Dev;

namespace {

struct Plane {
    int axis;
    double position;
    int below, above, across;
};

__host__ __device__ inline int plane_type(const Plane &p, double xi,
                                          double xj)
{
    const bool bi = xi < p.position, bj = xj < p.position;
    if (bi && bj)
        return p.below;
    if (!bi && !bj)
        return p.above;
    return p.across;
}

// Row i: the type of each live slot (the first n_neigh[i]), 0 elsewhere (the
// reference's array starts as zeros, model.py:752-753).
__host__ __device__ inline void plane_row(const Plane &p, int n, int K,
                                          const double *x, const int *nlist,
                                          const int *n_neigh, int i, int *out)
{
    const int *row = nlist + (size_t)i * K;
    int *o = out + (size_t)i * K;
    const int c = n_neigh[i];
    const double xi = x[3 * (size_t)i + p.axis];
    for (int s = 0; s < K; ++s) {
        int t = 0;
        if (s < c) {
            const int j = row[s];
            if (j >= 0 && j < n)
                t = plane_type(p, xi, x[3 * (size_t)j + p.axis]);
        }
        o[s] = t;
    }
}

__global__ void k_plane_types(Plane p, int n, int K, const double *x,
                              const int *nlist, const int *n_neigh, int *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        plane_row(p, n, K, x, nlist, n_neigh, i, out);
}

// Box b holds x when lo_b <= x < hi_b in every coordinate (bounds may be
// infinite); the first such box in the list decides.
__host__ __device__ inline int first_box(int nb, const double *lo,
                                         const double *hi, const double *x)
{
    for (int b = 0; b < nb; ++b) {
        bool in = true;
        for (int d = 0; d < 3; ++d)
            in = in && lo[3 * b + d] <= x[d] && x[d] < hi[3 * b + d];
        if (in)
            return b;
    }
    return -1;
}

__global__ void k_first_box(int n, int nb, const double *lo,
                            const double *hi, const double *x, int *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = first_box(nb, lo, hi, x + 3 * (size_t)i);
}

}  // namespace

extern "C" {

int pd_plane_bond_types(int device, int nnodes, int max_neighbours,
                        const double *coords, const int *nlist,
                        const int *n_neigh, int axis, double position,
                        int below, int above, int across, int *bond_types)
{
    if (nnodes <= 0 || max_neighbours <= 0 || !coords || !nlist ||
        !n_neigh || !bond_types)
        return pd_fail(PD_E_ARG, "pd_plane_bond_types: bad arguments");
    if (axis < 0 || axis > 2)
        return pd_fail(PD_E_ARG, "axis must be 0, 1 or 2");
    for (int i = 0; i < nnodes; ++i)
        if (n_neigh[i] < 0 || n_neigh[i] > max_neighbours)
            return pd_fail(PD_E_ARG, "n_neigh out of range");
    const Plane p{axis, position, below, above, across};
    const size_t N = (size_t)nnodes, NK = N * max_neighbours;
    if (device < 0) {
        for (int i = 0; i < nnodes; ++i)
            plane_row(p, nnodes, max_neighbours, coords, nlist, n_neigh, i,
                      bond_types);
        return PD_OK;
    }
    int rc = pdsetup::use_device(device);
    if (rc)
        return rc;
    Dev<double> x;
    Dev<int> nl, nn, out;
    PD_SETUP_CHK(x.alloc(3 * N));
    PD_SETUP_CHK(nl.alloc(NK));
    PD_SETUP_CHK(nn.alloc(N));
    PD_SETUP_CHK(out.alloc(NK));
    PD_SETUP_CHK(hipMemcpy(x.p, coords, 3 * N * sizeof(double),
                           hipMemcpyHostToDevice));
    PD_SETUP_CHK(hipMemcpy(nl.p, nlist, NK * sizeof(int),
                           hipMemcpyHostToDevice));
    PD_SETUP_CHK(hipMemcpy(nn.p, n_neigh, N * sizeof(int),
                           hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_plane_types, dim3((nnodes + 255) / 256), dim3(256),
                       0, 0, p, nnodes, max_neighbours, x.p, nl.p, nn.p,
                       out.p);
    PD_SETUP_CHK(hipGetLastError());
    PD_SETUP_CHK(hipMemcpy(bond_types, out.p, NK * sizeof(int),
                           hipMemcpyDeviceToHost));
    return PD_OK;
}

int pd_first_box(int device, int nnodes, const double *coords, int nboxes,
                 const double *lo, const double *hi, int *box)
{
    if (nnodes < 0 || nboxes < 0 || (nnodes > 0 && (!coords || !box)) ||
        (nboxes > 0 && (!lo || !hi)))
        return pd_fail(PD_E_ARG, "pd_first_box: bad arguments");
    if (nnodes == 0)
        return PD_OK;
    const size_t N = (size_t)nnodes, B = 3 * (size_t)nboxes;
    if (device < 0) {
        for (int i = 0; i < nnodes; ++i)
            box[i] = first_box(nboxes, lo, hi, coords + 3 * (size_t)i);
        return PD_OK;
    }
    int rc = pdsetup::use_device(device);
    if (rc)
        return rc;
    Dev<double> x, l, h;
    Dev<int> out;
    PD_SETUP_CHK(x.alloc(3 * N));
    PD_SETUP_CHK(l.alloc(B));
    PD_SETUP_CHK(h.alloc(B));
    PD_SETUP_CHK(out.alloc(N));
    PD_SETUP_CHK(hipMemcpy(x.p, coords, 3 * N * sizeof(double),
                           hipMemcpyHostToDevice));
    if (B) {
        PD_SETUP_CHK(hipMemcpy(l.p, lo, B * sizeof(double),
                               hipMemcpyHostToDevice));
        PD_SETUP_CHK(hipMemcpy(h.p, hi, B * sizeof(double),
                               hipMemcpyHostToDevice));
    }
    hipLaunchKernelGGL(k_first_box, dim3((nnodes + 255) / 256), dim3(256), 0,
                       0, nnodes, nboxes, l.p, h.p, x.p, out.p);
    PD_SETUP_CHK(hipGetLastError());
    PD_SETUP_CHK(hipMemcpy(box, out.p, N * sizeof(int),
                           hipMemcpyDeviceToHost));
    return PD_OK;
}

}  // extern "C"
