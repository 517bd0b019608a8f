// libpd_hip: the stiffness-correction factors of peripy/correction.pyx
// (micromodulus function, partial-volume correction, surface correction by
// the volume method), computed on the GPU -- or on the host with the same
// code when device < 0.
//
// The four functions of correction.pyx share one sequential loop (:32-56,
// :97-120, :155-183, :236-262): rows i ascending, live slots in order; for a
// neighbour j > i the pair's factor f multiplies slot (i, s) and the FIRST
// slot of row j (within n_neigh[j]) that holds i.  Every slot is written by
// exactly one pair, so the loop is parallel over slots:
//   j > i   corr[i][s] *= f, once;
//   j < i   if s is the first slot of row i that holds j, corr[i][s] is
//           multiplied by f once for every slot of row j that holds i (m
//           times in a row; m = 1 in a symmetric list without repeats);
//           any later slot holding j is left alone;
//   j == i, or j outside [0, nnodes): left alone.
// f is computed with the smaller index first, as the sequential loop does,
// and is a function of the unordered pair anyway: |r_b - r_a| squares the
// same differences either way round and fv_a + fv_b commutes.  Products and
// sums are rounded one by one (-ffp-contract=off), sqrt and division are the
// correctly rounded ones, so every factor is bitwise the reference's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "pd_dev.h"

using pdsetup::Dev;

namespace {

enum Kind { KIND_CONICAL = 0, KIND_PA_HHB = 1, KIND_VOLUME_METHOD = 2 };

struct CorrArgs {
    int n, K;
    int kind;
    int flag;             // the reference's algorithm flag: 0 = the formula,
                          // anything else = factor 1.00
    const int *nlist;     // (n, K)
    const int *n_neigh;   // (n,)
    const double *r0;     // (n, 3), geometric kinds
    const double *fam;    // (n,) family volumes, or null: n_neigh * average
    double horizon, radius, average, bulk;
    double *corr;         // (n, K), multiplied in place
};

// spatial.pyx:20-31 (ceuclid)
__host__ __device__ inline double euclid(const double *a, const double *b)
{
    double d0 = b[0] - a[0];
    double d1 = b[1] - a[1];
    double d2 = b[2] - a[2];
    d0 = d0 * d0;
    d1 = d1 * d1;
    d2 = d2 * d2;
    return sqrt((d0 + d1) + d2);
}

// Factor of the pair (a, b), a < b.
__host__ __device__ inline double pair_factor(const CorrArgs &c, int a, int b)
{
    if (c.kind == KIND_VOLUME_METHOD) {
        // correction.pyx:42-45, :107-109
        double va, vb;
        if (c.fam) {
            va = c.fam[a];
            vb = c.fam[b];
        } else {
            va = c.n_neigh[a] * c.average;
            vb = c.n_neigh[b] * c.average;
        }
        return 2. * c.bulk / (va + vb);
    }
    if (c.flag != 0)
        return 1.00;  // correction.pyx:171-172, :250-251
    const double l0 = euclid(c.r0 + 3 * (size_t)a, c.r0 + 3 * (size_t)b);
    const double h = c.horizon;
    if (c.kind == KIND_CONICAL) {
        // correction.pyx:265-277 (cmicromodulus_connical)
        if (l0 <= h)
            return (h - l0) / h;
        return 0.00;
    }
    // correction.pyx:186-205 (cPA_HHB)
    const double r = c.radius;
    if (l0 <= h - r)
        return 1.00;
    if (l0 <= h + r)
        return (h + r - l0) / (2.0 * r);
    return 0.00;
}

__host__ __device__ inline void correct_slot(const CorrArgs &c, int i, int s)
{
    const size_t K = (size_t)c.K;
    if (s >= c.n_neigh[i])
        return;
    const int *row = c.nlist + (size_t)i * K;
    const int j = row[s];
    if (j < 0 || j >= c.n || j == i)
        return;
    int reps = 1;
    if (j < i) {
        for (int t = 0; t < s; ++t)
            if (row[t] == j)
                return;
        const int *rj = c.nlist + (size_t)j * K;
        const int nj = c.n_neigh[j];
        reps = 0;
        for (int t = 0; t < nj; ++t)
            reps += rj[t] == i;
        if (reps == 0)
            return;
    }
    const double f = pair_factor(c, min(i, j), max(i, j));
    double v = c.corr[(size_t)i * K + s];
    for (int r = 0; r < reps; ++r)
        v *= f;
    c.corr[(size_t)i * K + s] = v;
}

// correction.pyx:87-95: family volume of i = in-order sum of the volumes of
// its live neighbours, from 0.0.
__host__ __device__ inline double family_volume(int K, const int *nlist,
                                                const int *n_neigh,
                                                const double *vol, int n,
                                                int i)
{
    const int *row = nlist + (size_t)i * K;
    double t = 0.0;
    for (int s = 0; s < n_neigh[i]; ++s) {
        const int j = row[s];
        if (j >= 0 && j < n)
            t += vol[j];
    }
    return t;
}

// One thread per slot; a wave covers consecutive slots of one or two rows.
__global__ void __launch_bounds__(256) k_correct(CorrArgs c)
{
    const size_t total = (size_t)c.n * c.K;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
         q += stride)
        correct_slot(c, (int)(q / c.K), (int)(q % c.K));
}

__global__ void __launch_bounds__(256)
    k_family_volume(int n, int K, const int *nlist, const int *n_neigh,
                    const double *vol, double *fam)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        fam[i] = family_volume(K, nlist, n_neigh, vol, n, i);
}

// Runs one correction over the caller's host arrays.  `vol` is given for the
// precise surface correction only.
int run(int device, CorrArgs c, const double *vol)
{
    if (c.n <= 0 || c.K <= 0 || !c.nlist || !c.n_neigh || !c.corr)
        return pd_fail(PD_E_ARG,
                       "nnodes, max_neighbours, the correction array, nlist "
                       "and n_neigh are required");
    if (c.kind != KIND_VOLUME_METHOD && c.flag == 0 && !c.r0)
        return pd_fail(PD_E_ARG, "r0 is required");
    for (int i = 0; i < c.n; ++i)
        if (c.n_neigh[i] < 0 || c.n_neigh[i] > c.K)
            return pd_fail(PD_E_ARG, "n_neigh must lie in [0, max_neighbours]");
    const size_t slots = (size_t)c.n * c.K;

    if (device < 0) {
        std::vector<double> fam;
        if (vol) {
            fam.resize(c.n);
            for (int i = 0; i < c.n; ++i)
                fam[i] = family_volume(c.K, c.nlist, c.n_neigh, vol, c.n, i);
            c.fam = fam.data();
        }
        for (int i = 0; i < c.n; ++i)
            for (int s = 0; s < c.K; ++s)
                correct_slot(c, i, s);
        return PD_OK;
    }

    if (int rc = pdsetup::use_device(device))
        return rc;
    Dev<int> dnl, dnn;
    Dev<double> dr0, dvol, dfam, dcorr;
    PD_SETUP_CHK(dnl.alloc(slots));
    PD_SETUP_CHK(dnn.alloc(c.n));
    PD_SETUP_CHK(dcorr.alloc(slots));
    PD_SETUP_CHK(hipMemcpy(dnl.p, c.nlist, slots * sizeof(int),
                           hipMemcpyHostToDevice));
    PD_SETUP_CHK(hipMemcpy(dnn.p, c.n_neigh, (size_t)c.n * sizeof(int),
                           hipMemcpyHostToDevice));
    PD_SETUP_CHK(hipMemcpy(dcorr.p, c.corr, slots * sizeof(double),
                           hipMemcpyHostToDevice));
    CorrArgs d = c;
    d.nlist = dnl.p;
    d.n_neigh = dnn.p;
    d.corr = dcorr.p;
    d.r0 = nullptr;
    if (c.r0) {
        PD_SETUP_CHK(dr0.alloc(3 * (size_t)c.n));
        PD_SETUP_CHK(hipMemcpy(dr0.p, c.r0, 3 * (size_t)c.n * sizeof(double),
                               hipMemcpyHostToDevice));
        d.r0 = dr0.p;
    }
    if (vol) {
        PD_SETUP_CHK(dvol.alloc(c.n));
        PD_SETUP_CHK(dfam.alloc(c.n));
        PD_SETUP_CHK(hipMemcpy(dvol.p, vol, (size_t)c.n * sizeof(double),
                               hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_family_volume, dim3((c.n + 255) / 256),
                           dim3(256), 0, 0, c.n, c.K, d.nlist, d.n_neigh,
                           dvol.p, dfam.p);
        PD_SETUP_CHK(hipGetLastError());
        d.fam = dfam.p;
    }
    const size_t blocks = std::min<size_t>((slots + 255) / 256, 1u << 20);
    hipLaunchKernelGGL(k_correct, dim3((unsigned)blocks), dim3(256), 0, 0, d);
    PD_SETUP_CHK(hipGetLastError());
    PD_SETUP_CHK(hipMemcpy(c.corr, dcorr.p, slots * sizeof(double),
                           hipMemcpyDeviceToHost));
    return PD_OK;
}

CorrArgs args(int n, int K, int kind, int flag, double *corr,
              const double *r0, const int *nlist, const int *n_neigh)
{
    CorrArgs c{};
    c.n = n;
    c.K = K;
    c.kind = kind;
    c.flag = flag;
    c.corr = corr;
    c.r0 = r0;
    c.nlist = nlist;
    c.n_neigh = n_neigh;
    return c;
}

}  // namespace

extern "C" int pd_set_micromodulus_function(int device, int nnodes,
                                            int max_neighbours,
                                            double *micromodulus_values,
                                            const double *r0,
                                            const int *nlist,
                                            const int *n_neigh,
                                            double horizon,
                                            int micromodulus_function)
{
    CorrArgs c = args(nnodes, max_neighbours, KIND_CONICAL,
                      micromodulus_function, micromodulus_values, r0, nlist,
                      n_neigh);
    c.horizon = horizon;
    return run(device, c, nullptr);
}

extern "C" int pd_set_volume_correction(int device, int nnodes,
                                        int max_neighbours,
                                        double *volume_corrections,
                                        const double *r0, const int *nlist,
                                        const int *n_neigh, double horizon,
                                        double node_radius,
                                        int volume_correction)
{
    CorrArgs c = args(nnodes, max_neighbours, KIND_PA_HHB, volume_correction,
                      volume_corrections, r0, nlist, n_neigh);
    c.horizon = horizon;
    c.radius = node_radius;
    return run(device, c, nullptr);
}

extern "C" int pd_set_imprecise_surface_correction(
    int device, int nnodes, int max_neighbours, double *stiffness_corrections,
    const int *nlist, const int *n_neigh, double average_volume,
    double family_volume_bulk)
{
    CorrArgs c = args(nnodes, max_neighbours, KIND_VOLUME_METHOD, 0,
                      stiffness_corrections, nullptr, nlist, n_neigh);
    c.average = average_volume;
    c.bulk = family_volume_bulk;
    return run(device, c, nullptr);
}

extern "C" int pd_set_precise_surface_correction(
    int device, int nnodes, int max_neighbours, double *stiffness_corrections,
    const int *nlist, const int *n_neigh, const double *volume,
    double family_volume_bulk)
{
    if (!volume)
        return pd_fail(PD_E_ARG, "volume is required");
    CorrArgs c = args(nnodes, max_neighbours, KIND_VOLUME_METHOD, 0,
                      stiffness_corrections, nullptr, nlist, n_neigh);
    c.bulk = family_volume_bulk;
    return run(device, c, volume);
}
