// libpd_hip: bond-based peridynamics time stepping for MI355X (gfx950).
//
// One fused kernel per time step (k_step): for every node i, every live bond
// (i, j) is tested for breaking and its force gathered into i; then the force
// boundary condition is added and the integrator update (Euler,
// Euler-Cromer or Velocity-Verlet) writes u (and ud, udd) of node i.  The
// positions r = r0 + u that neighbours gather are ping-ponged between two
// record arrays, so every bond of a step sees the positions at the start of
// the step (the break-before-force order of peripy/integrators.py:339-349 and
// the single bond_force pass of cl/peridynamics.cl both read u_n only).
//
// Data layout in HBM (N nodes, Np = N rounded up to 64, Kc = longest row,
// C = ceil(Kc/4) + 8 chunk rows):
//   rec[2][N]  NodeRec, 64 bytes: (r0x, r0y, r0z, V | rx, ry, rz, c) with
//              r = r0 + u (ping-pong; the r0/V half is static in both) and
//              c the per-node stiffness (or 0).  A neighbour gather reads one
//              64-byte-aligned record.
//   nbr[C*Np*4] int32, neighbour-major in chunks of 4: bond k of node i is at
//              ((k/4)*Np + i)*4 + k%4, so one 16-byte load gives a lane four
//              consecutive bonds and a wavefront reads 1 KiB contiguously.
//              Rows are sorted ascending by j (the k-th neighbours of
//              consecutive nodes are then close in memory).  Bit 31 marks a
//              broken bond (DEAD), bit 30 one broken in the current step
//              (NEW, compact layout only); bits 0-29 hold j.  The 8 extra
//              chunk rows are DEAD padding for the one-chunk-ahead prefetch.
//   cnt[N]     number of bond slots of node i
//   u, ud, udd, force, body_force (N,3) f64 in the reference's layout
//   flags[N]   bit d: displacement BC on DOF d; bit 3+d: force BC on DOF d
//   bcv, fbcv, dens (N,3) f64, read only for flagged DOFs / dynamic
//              integrators
//   corr (f64), state (u16: bond type << 8 | regime), pos (u16: reference
//              column of the slot): per-slot, in nbr's order
//   nlist_ref (N,K) i32 + n_neigh: the reference-layout list of the compact
//              (Cython) layout, maintained by in-kernel compaction.
//
// Stretch arithmetic: exactly spatial.pyx's (dr = r_j - r_i,
// l = sqrt((dr0^2 + dr1^2) + dr2^2), s = (l - l0)/l0), built without
// floating-point contraction, so l and l0 are bitwise those of the Cython
// code.  The square roots use the correctly rounded rsq-Goldschmidt sequence
// that the compiler emits for sqrt(double), without its range scaling (taken
// only for q outside [2^-767, inf), where the library sqrt is called).  Its
// by-product h ~ 0.5/sqrt(q) (relative error ~1e-13) replaces the two
// divisions; a stretch within a relative 1e-10 of any threshold it is
// compared with is recomputed with the IEEE division, so every break and
// regime decision equals the one made with the correctly rounded s.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pd_hip.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(call)                                                        \
    do {                                                                    \
        hipError_t e_ = (call);                                             \
        if (e_ != hipSuccess)                                               \
            return fail(PD_E_HIP, std::string(#call) + ": " +               \
                                      hipGetErrorString(e_));               \
    } while (0)

}  // namespace

namespace pdk {

constexpr uint32_t DEAD = 0x80000000u;
constexpr uint32_t NEWB = 0x40000000u;
constexpr uint32_t KEYM = 0x3FFFFFFFu;
constexpr int MAX_NODES = 1 << 26;  // gathers use a 32-bit byte offset j*64
constexpr double GUARD = 1e-10;
constexpr int MAX_TAB = 256;  // nbond_types * nregimes
constexpr int PAD_CHUNKS = 8;

// damage models of the step kernel
constexpr int DM_PMB = 0;    // scalar stiffness
constexpr int DM_NLIN = 1;   // n-linear / composite tables
constexpr int DM_FIELD = 2;  // PMB with k_ij = (c_i + c_j)/2

struct alignas(64) NodeRec {
    double4 a;  // r0x, r0y, r0z, V
    double4 b;  // rx, ry, rz, c
};

struct StepArgs {
    int n, np, K;
    int integrator, layout, has_corr;
    unsigned absmask;  // 0x7fffffff: break on |s|; 0xffffffff: on s
    int nregimes, ntab;
    double dt, damping, s0h, s0h_guard, k4;
    const NodeRec *rec_old;
    NodeRec *rec_new;
    double *u, *ud, *udd, *force, *body_force;
    const int *flags;
    const double *bcv, *fbcv, *dens;
    const int *cnt;
    int *nbr;
    const double *corr;
    unsigned short *state;
    const double *k_tab, *s_tab, *pc_tab;
    int *nlist_ref, *n_neigh;
};

__host__ __device__ __forceinline__ size_t sidx(int k, uint32_t i, int np)
{
    return ((size_t)(k >> 2) * (size_t)np + i) * 4 + (k & 3);
}

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------

// sqrt(q) rounded to nearest for q in [2^-767, inf): the rsq seed, one
// Goldschmidt step for (g ~ sqrt q, h ~ 0.5/sqrt q) and two residual
// corrections of g -- the sequence LLVM emits for llvm.sqrt.f64 on gfx9,
// minus the input scaling and the zero/infinity selects it needs outside
// that range.
__device__ __forceinline__ void sqrt_core(double q, double &g, double &h)
{
    const double y = __builtin_amdgcn_rsq(q);
    g = q * y;
    h = y * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, q);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, q);
    g = __builtin_fma(d, h, g);
}

__device__ __forceinline__ bool core_range(double q)
{
    return (q >= 0x1p-767) & (q < __builtin_inf());
}

// |x| (mask 0x7fffffff) or x (mask 0xffffffff): clears the sign bit or not
__device__ __forceinline__ double sign_mask(double x, unsigned mask)
{
    return __hiloint2double((int)((unsigned)__double2hiint(x) & mask),
                            __double2loint(x));
}

__device__ __forceinline__ uint32_t ld_nbr(const int *p)
{
    return (uint32_t)__hip_atomic_load(p, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
}

// Is j a bond of node i that broke in this step?  Binary search of the
// (ascending) keys of i's slots.
__device__ bool broke_now(const int *nbr, int np, uint32_t i, int cnt, int j)
{
    int lo = 0, hi = cnt - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const uint32_t e = ld_nbr(&nbr[sidx(mid, i, np)]);
        const int key = (int)(e & KEYM);
        if (key == j)
            return (e & NEWB) != 0;
        if (key < j)
            lo = mid + 1;
        else
            hi = mid - 1;
    }
    return false;
}

// Reproduces, for row i of the reference-layout list, what
// peripy/peridynamics.pyx:124-156 (break_bonds) does to that row in one pass:
// first every node k < i whose bond to i broke removes itself from i's row by
// swap-with-last, in ascending k (the outer loop order of break_bonds), then
// i's own sweep replaces each broken j > i by the last live entry without
// advancing.  All of it depends only on row i and the set of bonds of i that
// broke this step, which are the NEW slots of node i.
__device__ __noinline__ void compact_row(int *nbr, int np, int *nlist_ref,
                                         int K, int *n_neigh, uint32_t i,
                                         int cnt)
{
    __threadfence();
    int *R = nlist_ref + (size_t)i * K;
    int n = n_neigh[i];
    for (int k = 0; k < cnt; ++k) {
        const uint32_t e = ld_nbr(&nbr[sidx(k, i, np)]);
        const int j = (int)(e & KEYM);
        if ((e & NEWB) && j < (int)i) {
            for (int p = 0; p < n; ++p) {
                if (R[p] == j) {
                    R[p] = R[n - 1];
                    --n;
                    break;
                }
            }
        }
    }
    int p = 0;
    while (p < n) {
        const int j = R[p];
        if (j > (int)i && broke_now(nbr, np, i, cnt, j)) {
            R[p] = R[n - 1];
            --n;
        } else {
            ++p;
        }
    }
    for (int k = 0; k < cnt; ++k) {
        int *q = &nbr[sidx(k, i, np)];
        const uint32_t e = ld_nbr(q);
        if (e & NEWB)
            *q = (int)(e & ~NEWB);
    }
    n_neigh[i] = n;
}

// ---------------------------------------------------------------------------
// the step kernel
// ---------------------------------------------------------------------------

// One bond (i, j = e) in slot `slot` (an index into nbr / corr / state).
template <int DM>
__device__ __forceinline__ void bond(const StepArgs &a, const char *rb,
                                     const double4 &r0i, const double4 &ri,
                                     int e, size_t slot, const double *ts,
                                     const double *tk, const double *tp,
                                     double &fx, double &fy, double &fz,
                                     int &nbroken)
{
    if (e < 0)
        return;
    const NodeRec *pj =
        reinterpret_cast<const NodeRec *>(rb + ((uint32_t)e << 6));
    const double4 r0j = pj->a;
    const double4 rj = pj->b;

    // spatial.pyx:20-31 with r = r0 + u: dr = r_j - r_i
    const double dx = rj.x - ri.x;
    const double dy = rj.y - ri.y;
    const double dz = rj.z - ri.z;
    const double q = (dx * dx + dy * dy) + dz * dz;
    const double ex = r0j.x - r0i.x;
    const double ey = r0j.y - r0i.y;
    const double ez = r0j.z - r0i.z;
    const double q0 = (ex * ex + ey * ey) + ez * ez;

    double l, hl, l0, h0;
    sqrt_core(q, l, hl);
    sqrt_core(q0, l0, h0);
    double dl = l - l0;
    double sh = dl * h0;  // s/2, from h0 ~ 0.5/l0
    const bool fast = core_range(q) & core_range(q0);

    double g;
    if (DM != DM_NLIN) {
        double as = sign_mask(sh, a.absmask);
        if (!fast || fabs(as - a.s0h) <= a.s0h_guard) {
            // rare: outside the core's range, or near s0
            if (!fast) {
                l = sqrt(q);
                l0 = sqrt(q0);
                hl = 0.5 / l;
                dl = l - l0;
            }
            sh = (dl / l0) * 0.5;
            as = sign_mask(sh, a.absmask);
        }
        if (!(as < a.s0h)) {
            a.nbr[slot] = (int)((uint32_t)e | DEAD |
                                (a.layout == PD_LAYOUT_COMPACT ? NEWB : 0u));
            ++nbroken;
            return;
        }
        // s*k/l*V_j = (s/2)(0.5/l)(4k) V_j
        const double k4 = DM == DM_FIELD ? (ri.w + rj.w) * 2.0 : a.k4;
        g = sh * hl * k4 * r0j.w;
    } else {
        // cl/peridynamics.cl:326-378 (bond_force3/4) with thresholds halved
        const uint32_t st = a.state[slot];
        const int t = (int)(st >> 8);
        const int reg = (int)(st & 0xffu);
        const int base = t * a.nregimes;
        const double c_hi = ts[base + reg];
        const double c_lo = reg > 0 ? ts[base + reg - 1] : -__builtin_inf();
        if (!fast || fabs(sh - c_hi) <= GUARD * fabs(c_hi) ||
            fabs(sh - c_lo) <= GUARD * fabs(c_lo)) {
            if (!fast) {
                l = sqrt(q);
                l0 = sqrt(q0);
                hl = 0.5 / l;
                dl = l - l0;
            }
            sh = (dl / l0) * 0.5;
        }
        int nreg = reg;
        if (sh < c_hi) {
            if (reg > 0 && sh < c_lo)
                nreg = reg - 1;
        } else {
            nreg = reg + 1;
        }
        if (nreg != reg)
            a.state[slot] = (unsigned short)((t << 8) | nreg);
        if (nreg >= a.nregimes) {
            a.nbr[slot] = (int)((uint32_t)e | DEAD);
            ++nbroken;
            return;
        }
        const double fn = __builtin_fma(sh + sh, tk[base + nreg], tp[base + nreg]);
        g = fn * (hl + hl) * r0j.w;
    }
    if (a.has_corr)
        g *= a.corr[slot];
    fx = __builtin_fma(g, dx, fx);
    fy = __builtin_fma(g, dy, fy);
    fz = __builtin_fma(g, dz, fz);
}

// DM: DM_PMB, DM_NLIN or DM_FIELD.  LPN lanes share a node, taking its
// 4-bond chunks round-robin, and reduce with cross-lane shuffles.
template <int DM, int LPN>
__global__ __launch_bounds__(256) void k_step(StepArgs a, double dmag,
                                              double fmag)
{
    __shared__ double tk[DM == DM_NLIN ? MAX_TAB : 1],
        ts[DM == DM_NLIN ? MAX_TAB : 1], tp[DM == DM_NLIN ? MAX_TAB : 1];
    if (DM == DM_NLIN) {
        for (int t = threadIdx.x; t < a.ntab; t += blockDim.x) {
            tk[t] = a.k_tab[t];
            ts[t] = a.s_tab[t] * 0.5;
            tp[t] = a.pc_tab[t];
        }
        __syncthreads();
    }

    const uint32_t gt = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t i = gt / LPN;
    const int sub = (int)(gt % LPN);
    const bool valid = i < (uint32_t)a.n;

    double fx = 0.0, fy = 0.0, fz = 0.0;
    int nbroken = 0;
    int cnt = 0;
    double4 r0i = make_double4(0.0, 0.0, 0.0, 0.0), ri = r0i;
    if (valid) {
        const NodeRec *pi = a.rec_old + i;
        r0i = pi->a;
        ri = pi->b;
        cnt = a.cnt[i];
        const char *rb = reinterpret_cast<const char *>(a.rec_old);
        const int nch = (cnt + 3) >> 2;
        const int4 *nb4 = reinterpret_cast<const int4 *>(a.nbr);
        int4 cur = nb4[(size_t)sub * a.np + i];
        for (int c = sub; c < nch; c += LPN) {
            // prefetch the next chunk (in bounds: PAD_CHUNKS rows of DEAD)
            const int4 nxt = nb4[(size_t)(c + LPN) * a.np + i];
            const size_t s0 = ((size_t)c * a.np + i) * 4;
            bond<DM>(a, rb, r0i, ri, cur.x, s0 + 0, ts, tk, tp, fx, fy, fz, nbroken);
            bond<DM>(a, rb, r0i, ri, cur.y, s0 + 1, ts, tk, tp, fx, fy, fz, nbroken);
            bond<DM>(a, rb, r0i, ri, cur.z, s0 + 2, ts, tk, tp, fx, fy, fz, nbroken);
            bond<DM>(a, rb, r0i, ri, cur.w, s0 + 3, ts, tk, tp, fx, fy, fz, nbroken);
            cur = nxt;
        }
    }

#pragma unroll
    for (int off = LPN / 2; off > 0; off >>= 1) {
        fx += __shfl_xor(fx, off, LPN);
        fy += __shfl_xor(fy, off, LPN);
        fz += __shfl_xor(fz, off, LPN);
        nbroken += __shfl_xor(nbroken, off, LPN);
    }
    if (!valid || sub != 0)
        return;

    if (a.layout == PD_LAYOUT_COMPACT && nbroken > 0)
        compact_row(a.nbr, a.np, a.nlist_ref, a.K, a.n_neigh, i, cnt);

    // force boundary condition after the bond sum (peridynamics.pyx:93-97,
    // cl/peridynamics.cl:117-129)
    const int fl = a.flags[i];
    const size_t b = 3 * (size_t)i;
    const double bf[3] = {fx, fy, fz};
    double F[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        F[d] = (fl >> (3 + d)) & 1 ? bf[d] + fmag * a.fbcv[b + d] : bf[d];
        a.body_force[b + d] = bf[d];
        a.force[b + d] = F[d];
    }

    double u[3] = {a.u[b], a.u[b + 1], a.u[b + 2]};
    if (a.integrator == PD_EULER) {
        // peridynamics.pyx:179-185, cl/euler.cl:21-23
#pragma unroll
        for (int d = 0; d < 3; ++d)
            u[d] = (fl >> d) & 1 ? dmag * a.bcv[b + d] : u[d] + a.dt * F[d];
    } else if (a.integrator == PD_EULER_CROMER) {
        // cl/euler_cromer.cl:29-34
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double ud = a.ud[b + d];
            const double uddi = (F[d] - a.damping * ud) / a.dens[b + d];
            ud = ud + uddi * a.dt;
            a.udd[b + d] = uddi;
            a.ud[b + d] = ud;
            u[d] = (fl >> d) & 1 ? dmag * a.bcv[b + d] : u[d] + a.dt * ud;
        }
    } else {
        // cl/velocity_verlet.cl:29-35
        const double hdt = a.dt / 2;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double ud1 = a.ud[b + d] + hdt * a.udd[b + d];
            const double udd1 = (F[d] - a.damping * ud1) / a.dens[b + d];
            const double ud = ud1 + hdt * udd1;
            a.ud[b + d] = ud;
            a.udd[b + d] = udd1;
            u[d] = (fl >> d) & 1 ? dmag * a.bcv[b + d]
                                 : u[d] + a.dt * (ud + hdt * udd1);
        }
    }
    a.u[b] = u[0];
    a.u[b + 1] = u[1];
    a.u[b + 2] = u[2];
    a.rec_new[i].b =
        make_double4(r0i.x + u[0], r0i.y + u[1], r0i.z + u[2], ri.w);
}

// The stretch pipeline of bond<DM_PMB>, one bond per thread, for checking:
// l, l0 (bitwise those of spatial.pyx), the working s (2*sh), the exact
// s = (l - l0)/l0 and the break decision.
__global__ void k_bond_check(int n, const double *ri, const double *rj,
                             const double *r0i, const double *r0j,
                             double s0, unsigned absmask, double *l_out,
                             double *l0_out, double *s_out, double *s_exact,
                             int *broken)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    const size_t b = 3 * (size_t)t;
    const double dx = rj[b] - ri[b];
    const double dy = rj[b + 1] - ri[b + 1];
    const double dz = rj[b + 2] - ri[b + 2];
    const double q = (dx * dx + dy * dy) + dz * dz;
    const double ex = r0j[b] - r0i[b];
    const double ey = r0j[b + 1] - r0i[b + 1];
    const double ez = r0j[b + 2] - r0i[b + 2];
    const double q0 = (ex * ex + ey * ey) + ez * ez;
    double l, hl, l0, h0;
    sqrt_core(q, l, hl);
    sqrt_core(q0, l0, h0);
    double dl = l - l0;
    double sh = dl * h0;
    const bool fast = core_range(q) & core_range(q0);
    const double s0h = s0 * 0.5;
    double as = sign_mask(sh, absmask);
    if (!fast || fabs(as - s0h) <= GUARD * s0h) {
        if (!fast) {
            l = sqrt(q);
            l0 = sqrt(q0);
            dl = l - l0;
        }
        sh = (dl / l0) * 0.5;
        as = sign_mask(sh, absmask);
    }
    l_out[t] = l;
    l0_out[t] = l0;
    s_out[t] = sh + sh;
    s_exact[t] = dl / l0;
    broken[t] = !(as < s0h);
}

// Records of both ping-pong buffers from r0, V, u and c: a = (r0, V),
// b = (r0 + u, c).
__global__ void k_records(int n, const double *r0, const double *vol,
                          const double *u, const double *c, NodeRec *rec0,
                          NodeRec *rec1)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const size_t b = 3 * (size_t)i;
    NodeRec r;
    r.a = make_double4(r0[b], r0[b + 1], r0[b + 2], vol[i]);
    r.b = make_double4(r0[b] + u[b], r0[b + 1] + u[b + 1], r0[b + 2] + u[b + 2],
                       c ? c[i] : 0.0);
    rec0[i] = r;
    rec1[i] = r;
}

// Live-slot count per node, the damage (family - n)/family
// (peridynamics.pyx:29), and, for the slot layout, the reference-layout
// neighbour list: -1 everywhere but at the live slots' columns.
__global__ void k_export(int n, int np, int K, int layout, const int *cnt,
                         const int *nbr, const unsigned short *pos,
                         const int *family, const int *n_neigh_c,
                         int *nlist_out, int *n_neigh_out, double *damage)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    int live;
    if (layout == PD_LAYOUT_SLOT) {
        int *row = nlist_out + (size_t)i * K;
        for (int p = 0; p < K; ++p)
            row[p] = -1;
        live = 0;
        const int c = cnt[i];
        for (int k = 0; k < c; ++k) {
            const size_t s = sidx(k, (uint32_t)i, np);
            const int e = nbr[s];
            if (e >= 0) {
                row[pos[s]] = e;
                ++live;
            }
        }
    } else {
        live = n_neigh_c[i];
    }
    n_neigh_out[i] = live;
    const double f = (double)family[i];
    damage[i] = (f - (double)live) / f;
}

__global__ void k_export_regimes(int n, int np, int K, const int *cnt,
                                 const unsigned short *pos,
                                 const unsigned short *state, int *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int c = cnt[i];
    for (int k = 0; k < c; ++k) {
        const size_t s = sidx(k, (uint32_t)i, np);
        out[(size_t)i * K + pos[s]] = state[s] & 0xff;
    }
}

__global__ void k_live(int n, int np, const int *cnt, const int *nbr,
                       unsigned long long *total)
{
    __shared__ unsigned long long part[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long live = 0;
    if (i < n) {
        const int c = cnt[i];
        for (int k = 0; k < c; ++k)
            live += nbr[sidx(k, (uint32_t)i, np)] >= 0;
    }
    part[threadIdx.x] = live;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        atomicAdd(total, part[0]);
}

}  // namespace pdk

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

namespace {

using namespace pdk;

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t count)
    {
        release();
        n = count;
        if (count == 0)
            return hipSuccess;
        return hipMalloc(reinterpret_cast<void **>(&p), count * sizeof(T));
    }
};

}  // namespace

struct pd_sim {
    int device = 0;
    int integrator = PD_EULER, layout = PD_LAYOUT_SLOT;
    double dt = 0.0, damping = 0.0;
    int n = 0, np = 0, K = 0, kc = 0;
    int tension_only = 0;
    int lpn = 1, block = 256;
    bool has_buffers = false, has_corr = false, has_btypes = false,
         has_cnode = false;
    int nregimes = 1, nbond_types = 1;
    double k = 0.0, s0 = 0.0;
    int cur = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    std::vector<double> h_corr;  // (N,K) as given to build
    std::vector<int> h_btypes;   // (N,K)

    DevBuf<NodeRec> rec[2];
    DevBuf<double> r0, vol, u, ud, udd, force, body_force, bcv, fbcv, dens,
        damage, cnode, corr, k_tab, s_tab, pc_tab;
    DevBuf<int> flags, family, cnt, nbr, nlist_ref, n_neigh, out_nlist,
        out_nneigh;
    DevBuf<unsigned short> pos, state;
    DevBuf<unsigned long long> counter;
};

namespace {

void run_parallel(int n, const std::function<void(int, int)> &fn)
{
    const unsigned hw = std::thread::hardware_concurrency();
    int nt = (int)std::min<unsigned>(hw ? hw : 1, 16);
    if (n < 4096)
        nt = 1;
    std::vector<std::thread> th;
    const int chunk = (n + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        const int lo = t * chunk, hi = std::min(n, lo + chunk);
        if (lo >= hi)
            break;
        th.emplace_back(fn, lo, hi);
    }
    for (auto &x : th)
        x.join();
}

template <typename T>
hipError_t upload(DevBuf<T> &d, const T *h, size_t count)
{
    hipError_t e = d.alloc(count);
    if (e != hipSuccess || count == 0)
        return e;
    return hipMemcpy(d.p, h, count * sizeof(T), hipMemcpyHostToDevice);
}

int check_sim(pd_sim *s)
{
    if (!s)
        return fail(PD_E_ARG, "null simulation handle");
    HIPCHK(hipSetDevice(s->device));
    return PD_OK;
}

int dm_of(const pd_sim *s)
{
    if (s->nregimes > 1 || s->nbond_types > 1)
        return DM_NLIN;
    return s->has_cnode ? DM_FIELD : DM_PMB;
}

StepArgs make_args(pd_sim *s)
{
    StepArgs a;
    a.n = s->n;
    a.np = s->np;
    a.K = s->K;
    a.integrator = s->integrator;
    a.layout = s->layout;
    a.has_corr = s->has_corr ? 1 : 0;
    a.absmask = s->tension_only ? 0xffffffffu : 0x7fffffffu;
    a.nregimes = s->nregimes;
    a.ntab = s->nregimes * s->nbond_types;
    a.dt = s->dt;
    a.damping = s->damping;
    a.s0h = s->s0 * 0.5;
    a.s0h_guard = GUARD * a.s0h;
    a.k4 = s->k * 4.0;
    a.rec_old = s->rec[s->cur].p;
    a.rec_new = s->rec[1 - s->cur].p;
    a.u = s->u.p;
    a.ud = s->ud.p;
    a.udd = s->udd.p;
    a.force = s->force.p;
    a.body_force = s->body_force.p;
    a.flags = s->flags.p;
    a.bcv = s->bcv.p;
    a.fbcv = s->fbcv.p;
    a.dens = s->dens.p;
    a.cnt = s->cnt.p;
    a.nbr = s->nbr.p;
    a.corr = s->corr.p;
    a.state = s->state.p;
    a.k_tab = s->k_tab.p;
    a.s_tab = s->s_tab.p;
    a.pc_tab = s->pc_tab.p;
    a.nlist_ref = s->nlist_ref.p;
    a.n_neigh = s->n_neigh.p;
    return a;
}

template <int DM>
hipError_t launch_dm(pd_sim *s, const StepArgs &a, double dmag, double fmag)
{
    const unsigned threads = (unsigned)s->n * (unsigned)s->lpn;
    const unsigned grid = (threads + s->block - 1) / s->block;
    switch (s->lpn) {
    case 1:
        hipLaunchKernelGGL((k_step<DM, 1>), dim3(grid), dim3(s->block), 0,
                           s->stream, a, dmag, fmag);
        break;
    case 2:
        hipLaunchKernelGGL((k_step<DM, 2>), dim3(grid), dim3(s->block), 0,
                           s->stream, a, dmag, fmag);
        break;
    case 4:
        hipLaunchKernelGGL((k_step<DM, 4>), dim3(grid), dim3(s->block), 0,
                           s->stream, a, dmag, fmag);
        break;
    default:
        hipLaunchKernelGGL((k_step<DM, 8>), dim3(grid), dim3(s->block), 0,
                           s->stream, a, dmag, fmag);
        break;
    }
    return hipGetLastError();
}

int launch_step(pd_sim *s, double dmag, double fmag)
{
    const StepArgs a = make_args(s);
    hipError_t e;
    switch (dm_of(s)) {
    case DM_NLIN:
        e = launch_dm<DM_NLIN>(s, a, dmag, fmag);
        break;
    case DM_FIELD:
        e = launch_dm<DM_FIELD>(s, a, dmag, fmag);
        break;
    default:
        e = launch_dm<DM_PMB>(s, a, dmag, fmag);
        break;
    }
    if (e != hipSuccess)
        return fail(PD_E_HIP, std::string("step kernel launch: ") +
                                  hipGetErrorString(e));
    s->cur = 1 - s->cur;
    return PD_OK;
}

// (Re)write both record buffers from r0, V, the device u and c.
int refresh_records(pd_sim *s)
{
    const int blocks = (s->n + 255) / 256;
    hipLaunchKernelGGL(k_records, dim3(blocks), dim3(256), 0, s->stream, s->n,
                       s->r0.p, s->vol.p, s->u.p,
                       s->has_cnode ? s->cnode.p : nullptr, s->rec[0].p,
                       s->rec[1].p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    s->cur = 0;
    return PD_OK;
}

}  // namespace

extern "C" {

const char *pd_last_error(void) { return g_err.c_str(); }

int pd_device_count(int *count)
{
    if (!count)
        return fail(PD_E_ARG, "count is NULL");
    int c = 0;
    const hipError_t e = hipGetDeviceCount(&c);
    *count = (e == hipSuccess) ? c : 0;
    return PD_OK;
}

int pd_device_name(int device, char *name, int name_len, char *arch,
                   int arch_len)
{
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, device));
    if (name && name_len > 0)
        snprintf(name, name_len, "%s", p.name);
    if (arch && arch_len > 0)
        snprintf(arch, arch_len, "%s", p.gcnArchName);
    return PD_OK;
}

int pd_build(pd_sim **out, int device, int integrator, int layout, double dt,
             double damping, int nnodes, int max_neighbours,
             const double *coords, const double *volume, const int *family,
             const int *bc_types, const double *bc_values,
             const int *force_bc_types, const double *force_bc_values,
             const double *stiffness_corrections, const int *bond_types,
             const double *densities)
{
    if (!out)
        return fail(PD_E_ARG, "sim is NULL");
    *out = nullptr;
    if (integrator < PD_EULER || integrator > PD_VELOCITY_VERLET)
        return fail(PD_E_ARG, "integrator must be 0, 1 or 2");
    if (layout != PD_LAYOUT_COMPACT && layout != PD_LAYOUT_SLOT)
        return fail(PD_E_ARG, "layout must be 0 (compact) or 1 (slot)");
    if (nnodes <= 0 || nnodes > MAX_NODES)
        return fail(PD_E_LIMIT, "nnodes must be in 1..2^26");
    if (max_neighbours <= 0 || max_neighbours > 65535)
        return fail(PD_E_LIMIT, "max_neighbours must be in 1..65535");
    if (!coords || !volume || !family || !bc_types || !bc_values ||
        !force_bc_types || !force_bc_values)
        return fail(PD_E_ARG, "a required build array is NULL");
    if (integrator == PD_EULER && densities)
        return fail(PD_E_ARG,
                    "densities are not supported by this integrator: it "
                    "neglects inertial effects");
    if (integrator != PD_EULER && !densities)
        return fail(PD_E_ARG,
                    "densities must be supplied for a dynamic integrator");
    if (layout == PD_LAYOUT_COMPACT && (stiffness_corrections || bond_types))
        return fail(PD_E_ARG,
                    "stiffness_corrections and bond_types need the slot "
                    "(-1 marking) neighbour-list layout");
    for (int i = 0; i < nnodes; ++i)
        if (family[i] <= 0)
            return fail(PD_E_ARG, "family must be > 0 for every node");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(PD_E_NODEV, "no HIP device is available");
    if (device < 0 || device >= ndev)
        return fail(PD_E_ARG, "device index out of range");

    HIPCHK(hipSetDevice(device));
    pd_sim *s = new pd_sim();
    s->device = device;
    s->integrator = integrator;
    s->layout = layout;
    s->dt = dt;
    s->damping = damping;
    s->n = nnodes;
    s->np = (nnodes + 63) & ~63;
    s->K = max_neighbours;
    const size_t N = (size_t)nnodes;

    std::vector<int> flags(N);
    for (size_t i = 0; i < N; ++i) {
        int f = 0;
        for (int d = 0; d < 3; ++d) {
            if (bc_types[3 * i + d] != 0)
                f |= 1 << d;
            if (force_bc_types[3 * i + d] != 0)
                f |= 1 << (3 + d);
        }
        flags[i] = f;
    }
    int rc = PD_OK;
    auto chk = [&](hipError_t e, const char *what) {
        if (e != hipSuccess && rc == PD_OK)
            rc = fail(PD_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    };
    chk(upload(s->r0, coords, 3 * N), "coords");
    chk(upload(s->vol, volume, N), "volume");
    chk(upload(s->flags, flags.data(), N), "flags");
    chk(upload(s->family, family, N), "family");
    chk(upload(s->bcv, bc_values, 3 * N), "bc_values");
    chk(upload(s->fbcv, force_bc_values, 3 * N), "force_bc_values");
    if (densities)
        chk(upload(s->dens, densities, 3 * N), "densities");
    chk(s->rec[0].alloc(N), "rec");
    chk(s->rec[1].alloc(N), "rec");
    chk(s->u.alloc(3 * N), "u");
    chk(s->ud.alloc(3 * N), "ud");
    chk(s->udd.alloc(3 * N), "udd");
    chk(s->force.alloc(3 * N), "force");
    chk(s->body_force.alloc(3 * N), "body_force");
    chk(s->damage.alloc(N), "damage");
    chk(s->out_nneigh.alloc(N), "n_neigh");
    chk(s->counter.alloc(1), "counter");
    chk(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "stream");
    chk(hipEventCreate(&s->ev0), "event");
    chk(hipEventCreate(&s->ev1), "event");
    if (rc == PD_OK) {
        chk(hipMemset(s->force.p, 0, 3 * N * sizeof(double)), "memset");
        chk(hipMemset(s->body_force.p, 0, 3 * N * sizeof(double)), "memset");
    }
    if (stiffness_corrections) {
        s->h_corr.assign(stiffness_corrections,
                         stiffness_corrections + N * max_neighbours);
        s->has_corr = true;
    }
    if (bond_types) {
        s->h_btypes.assign(bond_types, bond_types + N * max_neighbours);
        s->has_btypes = true;
    }
    if (rc != PD_OK) {
        pd_destroy(s);
        return rc;
    }
    *out = s;
    return PD_OK;
}

int pd_create_buffers(pd_sim *s, const int *nlist, const int *n_neigh,
                      const double *bond_stiffness,
                      const double *critical_stretch, const double *plus_cs,
                      int nregimes, int nbond_types, const double *u,
                      const double *ud, const double *udd,
                      const int *regimes)
{
    int rc = check_sim(s);
    if (rc)
        return rc;
    if (!nlist || !bond_stiffness || !critical_stretch)
        return fail(PD_E_ARG, "nlist, bond_stiffness and critical_stretch "
                              "are required");
    if (s->layout == PD_LAYOUT_COMPACT && !n_neigh)
        return fail(PD_E_ARG, "n_neigh is required for the compact layout");
    if (nregimes < 1 || nbond_types < 1)
        return fail(PD_E_ARG, "nregimes and nbond_types must be >= 1");
    const bool nlin = nregimes > 1 || nbond_types > 1;
    if (nlin) {
        if (s->layout == PD_LAYOUT_COMPACT)
            return fail(PD_E_ARG,
                        "n-linear damage models are not supported by this "
                        "integrator (compact layout)");
        if (!s->has_btypes)
            return fail(PD_E_ARG, "bond_types must be given to build for an "
                                  "n-linear or composite damage model");
        if (!plus_cs)
            return fail(PD_E_ARG, "plus_cs is required with nregimes > 1");
        if (nregimes * nbond_types > MAX_TAB || nregimes > 254 ||
            nbond_types > 255)
            return fail(PD_E_LIMIT, "damage model table too large");
    }
    const int N = s->n, K = s->K, NP = s->np;
    if (nlin) {
        for (size_t q = 0; q < (size_t)N * K; ++q) {
            if (nlist[q] == -1)
                continue;
            const int t = s->h_btypes[q];
            if (t < 0 || t >= nbond_types)
                return fail(PD_E_ARG, "bond_types value out of range");
            if (regimes && (regimes[q] < 0 || regimes[q] >= nregimes))
                return fail(PD_E_ARG, "regimes value out of range for a "
                                      "live bond");
        }
    }

    // compute rows: live entries, sorted ascending by j
    std::vector<int> cnt(N);
    std::atomic<int> bad{0};
    run_parallel(N, [&](int lo, int hi) {
        for (int i = lo; i < hi; ++i) {
            const int *row = nlist + (size_t)i * K;
            int c = 0;
            if (s->layout == PD_LAYOUT_COMPACT) {
                c = n_neigh[i];
                if (c < 0 || c > K) {
                    bad = 1;
                    c = 0;
                }
                for (int p = 0; p < c; ++p)
                    if (row[p] < 0 || row[p] >= N)
                        bad = 2;
            } else {
                for (int p = 0; p < K; ++p) {
                    if (row[p] == -1)
                        continue;
                    if (row[p] < 0 || row[p] >= N)
                        bad = 2;
                    ++c;
                }
            }
            cnt[i] = c;
        }
    });
    if (bad == 1)
        return fail(PD_E_ARG, "n_neigh out of range [0, max_neighbours]");
    if (bad == 2)
        return fail(PD_E_ARG, "nlist entry out of range [0, nnodes)");
    int kc = 0;
    for (int i = 0; i < N; ++i)
        kc = std::max(kc, cnt[i]);
    s->kc = kc;
    const size_t S = ((size_t)((kc + 3) / 4) + PAD_CHUNKS) * NP * 4;
    std::vector<int> nbr(S, (int)DEAD);
    std::vector<unsigned short> pos(S, 0);
    std::vector<double> corr;
    std::vector<unsigned short> state;
    if (s->has_corr)
        corr.assign(S, 0.0);
    if (nlin)
        state.assign(S, 0);
    run_parallel(N, [&](int lo, int hi) {
        std::vector<std::pair<int, int>> ent;
        for (int i = lo; i < hi; ++i) {
            const int *row = nlist + (size_t)i * K;
            ent.clear();
            if (s->layout == PD_LAYOUT_COMPACT) {
                for (int p = 0; p < cnt[i]; ++p)
                    ent.emplace_back(row[p], p);
            } else {
                for (int p = 0; p < K; ++p)
                    if (row[p] != -1)
                        ent.emplace_back(row[p], p);
            }
            std::stable_sort(ent.begin(), ent.end(),
                             [](const std::pair<int, int> &x,
                                const std::pair<int, int> &y) {
                                 return x.first < y.first;
                             });
            for (int k = 0; k < (int)ent.size(); ++k) {
                const size_t sl = sidx(k, (uint32_t)i, NP);
                const size_t src = (size_t)i * K + ent[k].second;
                nbr[sl] = ent[k].first;
                pos[sl] = (unsigned short)ent[k].second;
                if (s->has_corr)
                    corr[sl] = s->h_corr[src];
                if (nlin)
                    state[sl] = (unsigned short)((s->h_btypes[src] << 8) |
                                                 (regimes ? regimes[src] : 0));
            }
        }
    });

    HIPCHK(hipStreamSynchronize(s->stream));
    auto chk = [&](hipError_t e, const char *what) {
        if (e != hipSuccess && rc == PD_OK)
            rc = fail(PD_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    };
    chk(upload(s->cnt, cnt.data(), (size_t)N), "cnt");
    chk(upload(s->nbr, nbr.data(), nbr.size()), "nbr");
    chk(upload(s->pos, pos.data(), pos.size()), "pos");
    if (s->has_corr)
        chk(upload(s->corr, corr.data(), corr.size()), "corr");
    else
        s->corr.release();
    if (nlin)
        chk(upload(s->state, state.data(), state.size()), "state");
    else
        s->state.release();
    if (s->layout == PD_LAYOUT_COMPACT) {
        chk(upload(s->nlist_ref, nlist, (size_t)N * K), "nlist");
        chk(upload(s->n_neigh, n_neigh, (size_t)N), "n_neigh");
    } else {
        s->nlist_ref.release();
        s->n_neigh.release();
    }
    chk(s->out_nlist.alloc(s->layout == PD_LAYOUT_SLOT ? (size_t)N * K : 0),
        "nlist export");
    const size_t B = 3 * (size_t)N * sizeof(double);
    if (u)
        chk(hipMemcpy(s->u.p, u, B, hipMemcpyHostToDevice), "u");
    else
        chk(hipMemset(s->u.p, 0, B), "u");
    if (ud)
        chk(hipMemcpy(s->ud.p, ud, B, hipMemcpyHostToDevice), "ud");
    else
        chk(hipMemset(s->ud.p, 0, B), "ud");
    if (udd)
        chk(hipMemcpy(s->udd.p, udd, B, hipMemcpyHostToDevice), "udd");
    else
        chk(hipMemset(s->udd.p, 0, B), "udd");
    s->nregimes = nregimes;
    s->nbond_types = nbond_types;
    if (nlin) {
        const size_t T = (size_t)nregimes * nbond_types;
        chk(upload(s->k_tab, bond_stiffness, T), "bond_stiffness");
        chk(upload(s->s_tab, critical_stretch, T), "critical_stretch");
        chk(upload(s->pc_tab, plus_cs, T), "plus_cs");
        s->k = 0.0;
        s->s0 = 0.0;
    } else {
        s->k_tab.release();
        s->s_tab.release();
        s->pc_tab.release();
        s->k = bond_stiffness[0];
        s->s0 = critical_stretch[0];
    }
    if (rc)
        return rc;
    rc = refresh_records(s);
    if (rc)
        return rc;
    s->has_buffers = true;
    return PD_OK;
}

int pd_set_node_stiffness(pd_sim *s, const double *c)
{
    int rc = check_sim(s);
    if (rc)
        return rc;
    HIPCHK(hipStreamSynchronize(s->stream));
    if (!c) {
        s->cnode.release();
        s->has_cnode = false;
    } else {
        HIPCHK(upload(s->cnode, c, (size_t)s->n));
        s->has_cnode = true;
    }
    if (s->has_buffers) {
        // the records carry c: rewrite both from the device u
        rc = refresh_records(s);
        if (rc)
            return rc;
    }
    return PD_OK;
}

int pd_set_break_mode(pd_sim *s, int tension_only)
{
    int rc = check_sim(s);
    if (rc)
        return rc;
    s->tension_only = tension_only ? 1 : 0;
    return PD_OK;
}

int pd_set_launch(pd_sim *s, int lanes_per_node, int block_size)
{
    int rc = check_sim(s);
    if (rc)
        return rc;
    if (lanes_per_node != 1 && lanes_per_node != 2 && lanes_per_node != 4 &&
        lanes_per_node != 8)
        return fail(PD_E_ARG, "lanes_per_node must be 1, 2, 4 or 8");
    if (block_size < 64 || block_size > 256 || block_size % 64)
        return fail(PD_E_ARG, "block_size must be 64, 128, 192 or 256");
    s->lpn = lanes_per_node;
    s->block = block_size;
    return PD_OK;
}

int pd_step(pd_sim *s, double dmag, double fmag)
{
    int rc = check_sim(s);
    if (rc)
        return rc;
    if (!s->has_buffers)
        return fail(PD_E_STATE, "pd_create_buffers has not been called");
    return launch_step(s, dmag, fmag);
}

int pd_run(pd_sim *s, int nsteps, const double *disp_mags,
           const double *force_mags)
{
    int rc = check_sim(s);
    if (rc)
        return rc;
    if (!s->has_buffers)
        return fail(PD_E_STATE, "pd_create_buffers has not been called");
    if (nsteps < 0)
        return fail(PD_E_ARG, "nsteps < 0");
    for (int t = 0; t < nsteps; ++t) {
        rc = launch_step(s, disp_mags ? disp_mags[t] : 0.0,
                         force_mags ? force_mags[t] : 0.0);
        if (rc)
            return rc;
    }
    return PD_OK;
}

int pd_run_timed(pd_sim *s, int nsteps, const double *disp_mags,
                 const double *force_mags, float *elapsed_ms)
{
    int rc = check_sim(s);
    if (rc)
        return rc;
    if (!elapsed_ms)
        return fail(PD_E_ARG, "elapsed_ms is NULL");
    HIPCHK(hipEventRecord(s->ev0, s->stream));
    rc = pd_run(s, nsteps, disp_mags, force_mags);
    if (rc)
        return rc;
    HIPCHK(hipEventRecord(s->ev1, s->stream));
    HIPCHK(hipEventSynchronize(s->ev1));
    HIPCHK(hipEventElapsedTime(elapsed_ms, s->ev0, s->ev1));
    return PD_OK;
}

int pd_write(pd_sim *s, double *u, double *ud, double *udd, double *force,
             double *body_force, double *damage, int *nlist, int *n_neigh)
{
    int rc = check_sim(s);
    if (rc)
        return rc;
    if (!s->has_buffers)
        return fail(PD_E_STATE, "pd_create_buffers has not been called");
    const int N = s->n;
    const int blocks = (N + 255) / 256;
    hipLaunchKernelGGL(k_export, dim3(blocks), dim3(256), 0, s->stream, N,
                       s->np, s->K, s->layout, s->cnt.p, s->nbr.p, s->pos.p,
                       s->family.p, s->n_neigh.p, s->out_nlist.p,
                       s->out_nneigh.p, s->damage.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    const size_t B = 3 * (size_t)N * sizeof(double);
    if (u)
        HIPCHK(hipMemcpy(u, s->u.p, B, hipMemcpyDeviceToHost));
    if (ud)
        HIPCHK(hipMemcpy(ud, s->ud.p, B, hipMemcpyDeviceToHost));
    if (udd)
        HIPCHK(hipMemcpy(udd, s->udd.p, B, hipMemcpyDeviceToHost));
    if (force)
        HIPCHK(hipMemcpy(force, s->force.p, B, hipMemcpyDeviceToHost));
    if (body_force)
        HIPCHK(hipMemcpy(body_force, s->body_force.p, B,
                         hipMemcpyDeviceToHost));
    if (damage)
        HIPCHK(hipMemcpy(damage, s->damage.p, (size_t)N * sizeof(double),
                         hipMemcpyDeviceToHost));
    if (nlist) {
        const size_t NB = (size_t)N * s->K * sizeof(int);
        if (s->layout == PD_LAYOUT_SLOT)
            HIPCHK(hipMemcpy(nlist, s->out_nlist.p, NB, hipMemcpyDeviceToHost));
        else
            HIPCHK(hipMemcpy(nlist, s->nlist_ref.p, NB, hipMemcpyDeviceToHost));
    }
    if (n_neigh)
        HIPCHK(hipMemcpy(n_neigh, s->out_nneigh.p, (size_t)N * sizeof(int),
                         hipMemcpyDeviceToHost));
    return PD_OK;
}

int pd_get_regimes(pd_sim *s, int *regimes)
{
    int rc = check_sim(s);
    if (rc)
        return rc;
    if (!s->has_buffers)
        return fail(PD_E_STATE, "pd_create_buffers has not been called");
    if (!regimes)
        return fail(PD_E_ARG, "regimes is NULL");
    const size_t NK = (size_t)s->n * s->K;
    if (!s->state.p) {
        std::memset(regimes, 0, NK * sizeof(int));
        return PD_OK;
    }
    DevBuf<int> out;
    HIPCHK(out.alloc(NK));
    HIPCHK(hipMemsetAsync(out.p, 0, NK * sizeof(int), s->stream));
    const int blocks = (s->n + 255) / 256;
    hipLaunchKernelGGL(k_export_regimes, dim3(blocks), dim3(256), 0,
                       s->stream, s->n, s->np, s->K, s->cnt.p, s->pos.p,
                       s->state.p, out.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipMemcpy(regimes, out.p, NK * sizeof(int), hipMemcpyDeviceToHost));
    return PD_OK;
}

int pd_live_bonds(pd_sim *s, long long *count)
{
    int rc = check_sim(s);
    if (rc)
        return rc;
    if (!count)
        return fail(PD_E_ARG, "count is NULL");
    if (!s->has_buffers)
        return fail(PD_E_STATE, "pd_create_buffers has not been called");
    HIPCHK(hipMemsetAsync(s->counter.p, 0, sizeof(unsigned long long),
                          s->stream));
    const int blocks = (s->n + 255) / 256;
    hipLaunchKernelGGL(k_live, dim3(blocks), dim3(256), 0, s->stream, s->n,
                       s->np, s->cnt.p, s->nbr.p, s->counter.p);
    HIPCHK(hipGetLastError());
    unsigned long long h = 0;
    HIPCHK(hipMemcpyAsync(&h, s->counter.p, sizeof(h), hipMemcpyDeviceToHost,
                          s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    *count = (long long)h;
    return PD_OK;
}

int pd_synchronize(pd_sim *s)
{
    int rc = check_sim(s);
    if (rc)
        return rc;
    HIPCHK(hipStreamSynchronize(s->stream));
    return PD_OK;
}

void pd_destroy(pd_sim *s)
{
    if (!s)
        return;
    (void)hipSetDevice(s->device);
    if (s->stream)
        (void)hipStreamSynchronize(s->stream);
    if (s->ev0)
        (void)hipEventDestroy(s->ev0);
    if (s->ev1)
        (void)hipEventDestroy(s->ev1);
    if (s->stream)
        (void)hipStreamDestroy(s->stream);
    delete s;
}

int pd_bond_check(int device, int n, const double *ri, const double *rj,
                  const double *r0i, const double *r0j, double s0,
                  int tension_only, double *l, double *l0, double *s,
                  double *s_exact, int *broken)
{
    if (n <= 0 || !ri || !rj || !r0i || !r0j || !l || !l0 || !s ||
        !s_exact || !broken)
        return fail(PD_E_ARG, "pd_bond_check: bad arguments");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(PD_E_NODEV, "no HIP device is available");
    HIPCHK(hipSetDevice(device));
    const size_t N = (size_t)n;
    DevBuf<double> d_ri, d_rj, d_r0i, d_r0j, d_l, d_l0, d_s, d_se;
    DevBuf<int> d_b;
    HIPCHK(upload(d_ri, ri, 3 * N));
    HIPCHK(upload(d_rj, rj, 3 * N));
    HIPCHK(upload(d_r0i, r0i, 3 * N));
    HIPCHK(upload(d_r0j, r0j, 3 * N));
    HIPCHK(d_l.alloc(N));
    HIPCHK(d_l0.alloc(N));
    HIPCHK(d_s.alloc(N));
    HIPCHK(d_se.alloc(N));
    HIPCHK(d_b.alloc(N));
    hipLaunchKernelGGL(k_bond_check, dim3((n + 255) / 256), dim3(256), 0, 0,
                       n, d_ri.p, d_rj.p, d_r0i.p, d_r0j.p, s0,
                       tension_only ? 0xffffffffu : 0x7fffffffu, d_l.p,
                       d_l0.p, d_s.p, d_se.p, d_b.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(l, d_l.p, N * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(l0, d_l0.p, N * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(s, d_s.p, N * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(s_exact, d_se.p, N * sizeof(double),
                     hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(broken, d_b.p, N * sizeof(int), hipMemcpyDeviceToHost));
    return PD_OK;
}

int pd_neighbour_list(int n, const double *x, double horizon, int *family,
                      int *nlist, int max_neighbours, int pad_value)
{
    if (n <= 0 || !x || !family)
        return fail(PD_E_ARG, "nnodes, coords and family are required");
    if (!(horizon > 0.0))
        return fail(PD_E_ARG, "horizon must be > 0");
    double lo[3] = {x[0], x[1], x[2]}, hi[3] = {x[0], x[1], x[2]};
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], x[3 * (size_t)i + d]);
            hi[d] = std::max(hi[d], x[3 * (size_t)i + d]);
        }
    long dim[3];
    long total = 1;
    for (int d = 0; d < 3; ++d) {
        dim[d] = std::max(1L, (long)std::floor((hi[d] - lo[d]) / horizon) + 1);
        dim[d] = std::min(dim[d], 1L << 20);
        total *= dim[d];
        if (total > (1L << 28))
            return fail(PD_E_LIMIT, "cell grid too large for this horizon");
    }
    auto cell_of = [&](int i, int d) {
        long c = (long)std::floor((x[3 * (size_t)i + d] - lo[d]) / horizon);
        return std::min(std::max(c, 0L), dim[d] - 1);
    };
    std::vector<long> cstart(total + 1, 0);
    std::vector<long> cid(n);
    for (int i = 0; i < n; ++i) {
        cid[i] = (cell_of(i, 0) * dim[1] + cell_of(i, 1)) * dim[2] + cell_of(i, 2);
        cstart[cid[i] + 1]++;
    }
    for (long c = 0; c < total; ++c)
        cstart[c + 1] += cstart[c];
    std::vector<int> members(n);
    {
        std::vector<long> fillp(cstart.begin(), cstart.end() - 1);
        for (int i = 0; i < n; ++i)
            members[fillp[cid[i]]++] = i;
    }
    const double h2 = horizon * horizon;
    std::atomic<int> overflow{0};
    run_parallel(n, [&](int a, int b) {
        std::vector<int> row;
        for (int i = a; i < b; ++i) {
            row.clear();
            const long ci[3] = {cell_of(i, 0), cell_of(i, 1), cell_of(i, 2)};
            for (long c0 = std::max(0L, ci[0] - 1); c0 <= std::min(dim[0] - 1, ci[0] + 1); ++c0)
                for (long c1 = std::max(0L, ci[1] - 1); c1 <= std::min(dim[1] - 1, ci[1] + 1); ++c1)
                    for (long c2 = std::max(0L, ci[2] - 1); c2 <= std::min(dim[2] - 1, ci[2] + 1); ++c2) {
                        const long c = (c0 * dim[1] + c1) * dim[2] + c2;
                        for (long m = cstart[c]; m < cstart[c + 1]; ++m) {
                            const int j = members[m];
                            if (j == i)
                                continue;
                            double d0 = x[3 * (size_t)j] - x[3 * (size_t)i];
                            double d1 = x[3 * (size_t)j + 1] - x[3 * (size_t)i + 1];
                            double d2 = x[3 * (size_t)j + 2] - x[3 * (size_t)i + 2];
                            d0 = d0 * d0;
                            d1 = d1 * d1;
                            d2 = d2 * d2;
                            if ((d0 + d1) + d2 <= h2)
                                row.push_back(j);
                        }
                    }
            std::sort(row.begin(), row.end());
            family[i] = (int)row.size();
            if (nlist) {
                if ((int)row.size() > max_neighbours) {
                    overflow = 1;
                    continue;
                }
                int *dst = nlist + (size_t)i * max_neighbours;
                std::copy(row.begin(), row.end(), dst);
                for (int p = (int)row.size(); p < max_neighbours; ++p)
                    dst[p] = pad_value;
            }
        }
    });
    if (overflow)
        return fail(PD_E_LIMIT, "a node has more than max_neighbours bonds");
    return PD_OK;
}

}  // extern "C"
