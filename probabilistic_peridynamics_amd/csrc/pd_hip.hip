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
// Tiles.  The nodes are cut into tiles: runs of up to `tile_nodes` nodes in
// Morton order of their initial coordinates, grown while the tile's halo --
// the tile's nodes and all their neighbours -- fits in LDS (1280 entries,
// 80 KiB, so two workgroups share a CU and one stages its halo while the
// other computes).  Within a tile the nodes are ordered by descending bond
// count, so a wavefront's nodes run loops of nearly equal length.  One
// workgroup runs one tile: it first copies the halo's records into LDS with
// coalesced loads (each record once), then every bond gathers its neighbour
// from LDS.
// A direct global gather would fetch each 64-byte record with four 16-byte
// loads from 64 different records per instruction, and the vector L1 then
// handles one 64-byte sector per clock: that, not HBM, was the bound.
// Tiles are mapped to workgroups so that each XCD (blockIdx % 8) runs a
// contiguous range of tiles, whose halos overlap in that XCD's L2.
//
// Data layout in HBM (N nodes, S bond slots).  Every per-node array is kept
// in tile order: internal node p is node tnode[p] of the caller's numbering,
// a tile's nodes are a contiguous range of p, and a halo is a few contiguous
// runs, so per-node loads and stores coalesce.  Inputs are permuted into
// this order on upload and outputs back on download.
//   rec[2][N]  NodeRec, 64 bytes: (r0x, r0y, r0z, w | rx, ry, rz, v) with
//              r = r0 + u (compact layout) or r = u (slot layout, whose
//              stretch is formed from u), the bond loop's weights w (4kV for the scalar
//              model, 2V for a stiffness field c, V for n-linear models) and
//              v (2cV for a field, else 0); ping-pong, with r0, w and v
//              static in both.
//   tiles[T]   TileInfo: node range [node_off, node_off + nodes) of p, halo
//              range in halo[], first chunk of the tile's bond slots, chunk
//              rows.
//   tnode[N]   caller's node id of internal node p; halo: ascending internal
//              ids per tile.
//   bond slots, in chunks of 4: chunk (c, l) of a tile with n nodes is
//              chunk_off + c*n + l and holds bonds 4c..4c+3 of its l-th node,
//              rows sorted ascending by j; a wavefront reads its nodes' c-th
//              chunks contiguously.  A tile has as many chunk rows as its
//              longest row needs (at least one); the kernel's look-ahead
//              chunk loads are clamped to the last of them.
//     loc   u16  1 + halo-local index of j; 0: broken or empty (hot loop;
//                entry 0 of a staged halo is a zero record)
//     gidx  i32  j, with bit 31 DEAD and bit 30 NEW (broken in this step,
//                compact layout) -- read only on breaks and exports
//     pos   u16  reference column of the slot (exports)
//     corr  f64, state u16 (bond type << 8 | regime): optional
//   chunk0[N], stride[N]: node p's chunk 0 and its tile's node count, so
//              slot(p, k) = (chunk0 + (k/4)*stride)*4 + k%4
//   u, ud, udd, force, body_force (N,3) f64 in the reference's layout
//   flags[N]   bit d: displacement BC on DOF d; bit 3+d: force BC on DOF d
//   bcv, fbcv, dens (N,3) f64, read only for flagged DOFs / dynamic
//              integrators
//   rslot (N,K) u16 + n_neigh: the reference-layout list of the compact
//              (Cython) layout, in the caller's numbering, as the slot index
//              of each row position, maintained by in-kernel compaction;
//              nlist_ref (N,K) i32 holds the caller's rows, rebuilt from
//              rslot at export.
//
// Stretch arithmetic.  Compact layout (the Cython path): exactly
// spatial.pyx's (dr = r_j - r_i with r = r0 + u, l = sqrt((dr0^2 + dr1^2) +
// dr2^2), s = (l - l0)/l0), built without floating-point contraction, so l
// and l0 are bitwise those of the Cython code.  Slot layout (the OpenCL
// integrators): cl/peridynamics.cl:65-75's (xi = r0_j - r0_i, xi_eta =
// (u_j - u_i) + xi), the squared lengths contracted as
// fma(z, z, fma(x, x, y*y)) and the updates contracted, as the OpenCL
// compiler that wrote the reference's Euler-Cromer and Velocity-Verlet
// fixtures evaluated them.  The square roots use the correctly rounded
// rsq-Goldschmidt sequence that the compiler emits for sqrt(double), without
// its range scaling (taken only for q outside [2^-767, inf), where the
// library sqrt is called).  Its by-product h ~ 0.5/sqrt(q) (relative error
// ~1e-13) replaces the two divisions; a stretch within a relative 1e-10 of
// any threshold it is compared with is recomputed with the IEEE division, so
// every break and regime decision equals the one made with the correctly
// rounded s.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <mutex>
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

// Error reporting for the library's other translation units.
int pd_fail(int code, const char *msg) { return fail(code, msg); }

namespace pdk {

constexpr uint32_t DEAD = 0x80000000u;
constexpr uint32_t NEWB = 0x40000000u;
constexpr uint32_t KEYM = 0x3FFFFFFFu;
constexpr unsigned short LOC_DEAD = 0;  // loc of a broken or empty slot
constexpr int MAX_NODES = 1 << 26;
constexpr double GUARD = 1e-10;
constexpr int MAX_TAB = 256;  // nbond_types * nregimes
constexpr int MAX_LPN = 8;
constexpr int XCDS = 8;
// Halo capacity of a tile: entries of 4 x 16 bytes, stored piece-major (piece
// q of entry e at byte q*16*cap + 16*e), so that one address register and
// immediate offsets reach all four pieces and random entries spread over all
// LDS banks.  4*16*cap bytes (plus the n-linear tables) stay within 80 KiB,
// which lets two workgroups share a CU's 160 KiB.
constexpr int LDS_CAP_PMB = 1280;
constexpr int LDS_CAP_NLIN = 1152;
constexpr unsigned LDS_BYTES = 81920;

// damage models of the step kernel
constexpr int DM_PMB = 0;    // scalar stiffness
constexpr int DM_NLIN = 1;   // n-linear / composite tables
constexpr int DM_FIELD = 2;  // PMB with k_ij = (c_i + c_j)/2

struct alignas(64) NodeRec {
    double4 a;  // r0x, r0y, r0z, V
    double4 b;  // rx, ry, rz, c
};

struct TileInfo {
    int node_off, nodes, halo_off, halo;
    long long chunk_off;
    int nch, pad;
};

struct StepArgs {
    int n, K, hstride, ntiles, tiles_per_xcd;
    int integrator, layout, has_corr;
    unsigned absmask;  // 0x7fffffff: break on |s|; 0xffffffff: on s
    int nregimes, ntab;
    double dt, damping, s0h, s0h_guard;
    const NodeRec *rec_old;
    NodeRec *rec_new;
    const TileInfo *tiles;
    const int *tnode, *halo;
    const double *u;  // u_n
    double *u_out;    // u_{n+1}: u itself, or the next slot of a replay
    double *ud, *udd, *force, *body_force;
    const int *flags;
    const double *bcv, *fbcv, *dens;
    const double *dens1;  // per-node density when all three DOFs share it
    const int *cnt;
    unsigned short *loc;
    int *gidx;
    const double *corr;
    unsigned short *state;
    const double *k_tab, *s_tab, *pc_tab;
    const long long *chunk0;
    const int *stride;
    unsigned short *rslot;  // compact layout: row position -> slot
    int *n_neigh;
    const double *cnode;  // node stiffness c (DM_FIELD)
    int *brk;  // if set: the step at which each slot's bond broke
    int step;  // index of this step (for brk)
    // exact-order step: lower-slot counts, upper-list lengths and the upper
    // lists (slot index << 16 | halo loc), the scalar stiffness
    const int *nlow;
    int *nup;
    unsigned *uord;
    double kx;
};

__host__ __device__ __forceinline__ size_t slot_of(long long chunk0,
                                                   int stride, int k)
{
    return (size_t)(chunk0 + (long long)(k >> 2) * stride) * 4 + (k & 3);
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

__device__ __forceinline__ uint32_t ld_slot(const int *p)
{
    return (uint32_t)__hip_atomic_load(p, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
}

// --- compaction of the compact (Cython) layout's rows -----------------------
//
// rslot (N, K) u16, rows in the caller's numbering: position p < n_neigh[m]
// of node m's reference-layout row holds the index of the slot (of m's
// internal node) whose bond it lists, so the row's entries are those slots'
// keys.  Positions from n_neigh[m] up to the row's count at set-up keep the
// values that swap-with-last left there, as the reference's rows do; the
// caller's row is rebuilt from rslot and gidx at export (k_export_rows).
//
// What peridynamics.pyx:124-156 (break_bonds) does to row i in one pass:
// first every node k < i whose bond to i broke removes itself from i's row by
// swap-with-last, in ascending k (the outer loop order of break_bonds), then
// i's own sweep replaces each broken j > i by the last entry without
// advancing.  All of it depends only on row i and on which of node i's bonds
// broke this step (its NEW slots).  Slots are sorted by key, so the lower
// neighbours are slots [0, nlow) and the upper ones [nhi, cnt).

__device__ __forceinline__ int slot_key(const int *gidx, long long c0,
                                        int stride, int k)
{
    return (int)(ld_slot(&gidx[slot_of(c0, stride, k)]) & KEYM);
}

// One node's row rs[0, n) in LDS, fl[k] = 1 for its slots that broke this
// step; i is the node's caller id (slot keys are caller ids).  Applies
// break_bonds' removals and returns the new count; nhi is set to the first
// slot whose key exceeds i.
__device__ __forceinline__ int compact_rows_seq(const int *gidx, long long c0,
                                             int stride, int cnt, int nlow,
                                             uint32_t i, unsigned short *rs,
                                             const unsigned char *fl, int n,
                                             int &nhi_out)
{
    int nhi = nlow;
    while (nhi < cnt && slot_key(gidx, c0, stride, nhi) == (int)i)
        ++nhi;
    nhi_out = nhi;
    for (int k = 0; k < nlow; ++k) {
        if (!fl[k])
            continue;
        // slots [kl, kh] repeat k's key; the reference removes the first
        // position holding that key
        const int j = slot_key(gidx, c0, stride, k);
        int kl = k, kh = k;
        while (kl > 0 && slot_key(gidx, c0, stride, kl - 1) == j)
            --kl;
        while (kh + 1 < nlow && slot_key(gidx, c0, stride, kh + 1) == j)
            ++kh;
        int p = 0;
        while (p < n && (rs[p] < kl || rs[p] > kh))
            ++p;
        if (p == n)
            continue;
        if (rs[p] != k) {
            // a repeat: the position that held slot k keeps the key, with
            // the slot that stays live
            int p2 = p + 1;
            while (p2 < n && rs[p2] != k)
                ++p2;
            if (p2 < n)
                rs[p2] = rs[p];
        }
        rs[p] = rs[n - 1];
        --n;
    }
    int rem = 0;
    for (int k = nhi; k < cnt; ++k)
        rem += fl[k];
    int p = 0;
    while (rem > 0 && p < n) {
        const int k = rs[p];
        if (k >= nhi && fl[k]) {
            rs[p] = rs[n - 1];
            --n;
            --rem;
        } else {
            ++p;
        }
    }
    return n;
}

// LDS one node's row takes in compact_tile (u16 positions, u8 break flags);
// the scratch also holds 64 bytes of counts.
__host__ __device__ constexpr unsigned compact_row_bytes(int K)
{
    return (3u * (unsigned)K + 15u) & ~15u;
}

// Every thread of the block calls this after its bond loop, with nbroken the
// node's count of bonds broken in this step (the same in all its lanes).  The
// rows of the nodes that lost bonds are staged in `scratch` (`bytes` of LDS
// the block no longer reads once every thread is here), compacted by one
// lane each, and written back; with EXACT, the node's upper list (uord, nup)
// is rebuilt in the new row order.  NEW flags are cleared.  Nodes take
// scratch regions in groups when they do not all fit.  The first 64 bytes
// hold each wave's count of such nodes, from which every node's region
// follows without atomics.
template <int LPN, bool EXACT>
__device__ void compact_tile(const StepArgs &a, const TileInfo &ti,
                             char *scratch, unsigned bytes, bool valid,
                             int l, int sub, uint32_t i, int cnt, int nbroken)
{
    const bool mine = valid && nbroken > 0;
    __syncthreads();  // every bond loop is done: the scratch is free
    const unsigned long long bm = __ballot(mine && sub == 0);
    const int lane = (int)(threadIdx.x & 63u), w = (int)(threadIdx.x >> 6);
    int *ctr = reinterpret_cast<int *>(scratch);
    if (lane == 0)
        ctr[w] = __popcll(bm);
    __syncthreads();
    int total = 0, rank = 0;
    for (int v = 0; v < (int)(blockDim.x >> 6); ++v) {
        const int c = ctr[v];
        total += c;
        rank += v < w ? c : 0;
    }
    if (total == 0)
        return;  // the same in every thread
    rank += __popcll(bm & ((1ull << lane) - 1ull));
    rank = __shfl(rank, 0, LPN);  // the node's first lane holds its rank
    const int K = a.K;
    const unsigned nb = compact_row_bytes(K);
    const int G = (int)((bytes - 64u) / nb);
    if (G < 1)
        return;  // rows too wide: rejected by the host before any launch
    const long long c0 = ti.chunk_off + l;
    const uint32_t m = mine ? (uint32_t)a.tnode[i] : 0u;
    unsigned short *R = a.rslot + (size_t)m * K;
    for (int g0 = 0; g0 < total; g0 += G) {
        const bool act = mine && rank >= g0 && rank < g0 + G;
        unsigned short *rs = reinterpret_cast<unsigned short *>(
            scratch + 64u + (unsigned)(act ? rank - g0 : 0) * nb);
        unsigned char *fl = reinterpret_cast<unsigned char *>(rs + K);
        int n = 0;
        if (act) {
            n = a.n_neigh[m];
            for (int p = sub; p < n; p += LPN)
                rs[p] = R[p];
            for (int k = sub; k < cnt; k += LPN) {
                int *q = &a.gidx[slot_of(c0, ti.nodes, k)];
                const uint32_t e = ld_slot(q);
                fl[k] = (e & NEWB) ? 1 : 0;
                if (e & NEWB)
                    *q = (int)(e & ~NEWB);
            }
        }
        __syncthreads();
        if (act && sub == 0) {
            int nhi = 0;
            // keys are caller ids: the node's own is m
            const int n1 = compact_rows_seq(a.gidx, c0, ti.nodes, cnt,
                                            a.nlow[i], m, rs, fl, n, nhi);
            a.n_neigh[m] = n1;
            if (EXACT) {
                int c = 0;
                for (int p = 0; p < n1; ++p) {
                    const int k = rs[p];
                    if (k >= nhi)
                        a.uord[slot_of(c0, ti.nodes, c++)] =
                            ((unsigned)k << 16) |
                            a.loc[slot_of(c0, ti.nodes, k)];
                }
                a.nup[i] = c;
            }
        }
        __syncthreads();
        if (act)
            for (int p = sub; p < n; p += LPN)
                R[p] = rs[p];
        __syncthreads();  // the next group reuses the scratch
    }
}

// ---------------------------------------------------------------------------
// the step kernel
// ---------------------------------------------------------------------------

template <int DM>
constexpr unsigned piece_bytes()
{
    return 16u * (DM == DM_NLIN ? LDS_CAP_NLIN : LDS_CAP_PMB);
}

// Squared length of (x, y, z): spatial.pyx's ((x*x + y*y) + z*z), uncontracted
// (CLA false), or cl/peridynamics.cl:72-73's as contracted,
// fma(z, z, fma(x, x, y*y)) (CLA true).
template <bool CLA>
__device__ __forceinline__ double sq3(double x, double y, double z)
{
    if (CLA)
        return __builtin_fma(z, z, __builtin_fma(x, x, y * y));
    return (x * x + y * y) + z * z;
}

// W bonds of node i at slots s0 .. s0+W-1: ev their halo indices (LOC_DEAD:
// broken or empty), st their (bond type << 8 | regime) for DM_NLIN, ci node
// i's stiffness for DM_FIELD.  Compile-time switches, so that the common
// runs spend no instruction on what they do not use: CORR, the per-slot
// stiffness corrections multiply the force (bond_force2/4); TENS, a PMB bond
// breaks on !(s < s0) (cl/peridynamics.cl:78) instead of !(|s| < s0)
// (peridynamics.pyx:133); CLA, the OpenCL stretch (records hold u; the bond
// vector is (u_j - u_i) + xi) instead of the Cython one (records hold
// r0 + u; the bond vector is r_j - r_i).
template <int DM, int W, bool CORR, bool TENS, bool CLA>
__device__ __forceinline__ void bonds(const StepArgs &a, const char *ldsb,
                                      const double4 &r0i, const double4 &ri,
                                      double ci,
                                      const unsigned *ev,
                                      const unsigned short *st, size_t s0,
                                      const double *ts, const double *tk,
                                      const double *tp, double &fx,
                                      double &fy, double &fz, int &nbroken)
{
    constexpr unsigned PB = piece_bytes<DM>();
    // per bond, what the force needs: dr, 0.5/l, s/2 and the weight
    double dx[W], dy[W], dz[W], hl[W], sh[W], w[W];
    bool live[W], slow[W];
#pragma unroll
    for (int q = 0; q < W; ++q) {
        live[q] = ev[q] != LOC_DEAD;
        // a dead slot reads the zero record at entry 0 (weight 0)
        const char *pe = ldsb + 16u * ev[q];
        const double2 p0 = *reinterpret_cast<const double2 *>(pe);
        const double2 p1 = *reinterpret_cast<const double2 *>(pe + PB);
        const double2 p2 = *reinterpret_cast<const double2 *>(pe + 2 * PB);
        // c_j is read only by the node-field model
        double2 p3;
        if (DM == DM_FIELD) {
            p3 = *reinterpret_cast<const double2 *>(pe + 3 * PB);
        } else {
            p3.x = *reinterpret_cast<const double *>(pe + 3 * PB);
            p3.y = 0.0;
        }
        const double ex = p0.x - r0i.x;
        const double ey = p0.y - r0i.y;
        const double ez = p1.x - r0i.z;
        if (CLA) {
            // cl/peridynamics.cl:65-71: xi_eta = (u_j - u_i) + xi
            dx[q] = (p2.x - ri.x) + ex;
            dy[q] = (p2.y - ri.y) + ey;
            dz[q] = (p3.x - ri.z) + ez;
        } else {
            // spatial.pyx:20-31 with r = r0 + u: dr = r_j - r_i
            dx[q] = p2.x - ri.x;
            dy[q] = p2.y - ri.y;
            dz[q] = p3.x - ri.z;
        }
        const double qq = sq3<CLA>(dx[q], dy[q], dz[q]);
        const double q0 = sq3<CLA>(ex, ey, ez);
        double l, l0, h0;
        sqrt_core(qq, l, hl[q]);
        sqrt_core(q0, l0, h0);
        sh[q] = (l - l0) * h0;  // s/2, from h0 ~ 0.5/l0
        // k = (c_i + c_j)/2 for a field: s k/l V_j = (s/2)(0.5/l) w with
        // w = 2(c_i+c_j)V_j = c_i (2 V_j) + (2 c_j V_j), the staged pair
        w[q] = DM == DM_FIELD ? __builtin_fma(ci, p1.y, p3.y) : p1.y;
        // sqrt_core is exact on [2^-767, inf).  With !(|s| < s0) as the
        // break test, one compare covers both lengths: with q >= 2^-700, a
        // q0 below 2^-767 (l0 < 2^-383.5, zero included) gives s > 2^33 or
        // NaN and an infinite or NaN q0 gives NaN, so the bond breaks here
        // as in the reference (unless s0 > 2^33); an infinite q gives NaN
        // here and an infinite s there, both breaking.  q below 2^-700, or
        // NaN, takes the exact path.  The tension test and the regime walk
        // check both lengths.
        bool fast;
        if constexpr (DM != DM_NLIN && !TENS)
            fast = qq >= 0x1p-700;
        else
            fast = core_range(qq) & core_range(q0);
        bool near;
        if (DM != DM_NLIN) {
            near = fabs((TENS ? sh[q] : fabs(sh[q])) - a.s0h) <= a.s0h_guard;
        } else {
            // (lower, upper) thresholds of the bond's regime; regime 0's
            // lower one is -DBL_MAX, which no stretch is near
            const double2 c = reinterpret_cast<const double2 *>(
                ts)[(st[q] >> 8) * a.nregimes + (st[q] & 0xff)];
            near = fabs(sh[q] - c.y) <= GUARD * fabs(c.y) ||
                   fabs(sh[q] - c.x) <= GUARD * fabs(c.x);
        }
        slow[q] = live[q] & (!fast | near);
    }
#pragma unroll
    for (int q = 0; q < W; ++q) {
        if (slow[q]) {
            // rare: near a threshold, or a square root outside the core's
            // range: IEEE square roots and division (spatial.pyx:87-97,
            // cl/peridynamics.cl:72-74)
            const char *pe = ldsb + 16u * ev[q];
            const double2 p0 = *reinterpret_cast<const double2 *>(pe);
            const double2 p1 = *reinterpret_cast<const double2 *>(pe + PB);
            const double qq = sq3<CLA>(dx[q], dy[q], dz[q]);
            const double ex = p0.x - r0i.x;
            const double ey = p0.y - r0i.y;
            const double ez = p1.x - r0i.z;
            const double q0 = sq3<CLA>(ex, ey, ez);
            const double l = sqrt(qq), l0 = sqrt(q0);
            if (!(core_range(qq) & core_range(q0)))
                hl[q] = 0.5 / l;
            sh[q] = ((l - l0) / l0) * 0.5;
        }
    }
    double cr[W];
#pragma unroll
    for (int q = 0; q < W; ++q)
        cr[q] = CORR ? a.corr[s0 + q] : 1.0;
#pragma unroll
    for (int q = 0; q < W; ++q) {
        if (!live[q])
            continue;
        double g;
        if (DM != DM_NLIN) {
            if (!((TENS ? sh[q] : fabs(sh[q])) < a.s0h)) {
                a.loc[s0 + q] = LOC_DEAD;
                if (a.brk)
                    a.brk[s0 + q] = a.step;
                a.gidx[s0 + q] |=
                    (int)(DEAD | (a.layout == PD_LAYOUT_COMPACT ? NEWB : 0u));
                ++nbroken;
                continue;
            }
            // s k/l V_j = (s/2)(0.5/l)(4 k V_j)
            g = sh[q] * hl[q] * w[q];
        } else {
            // cl/peridynamics.cl:326-378 (bond_force3/4), thresholds halved
            const int t = st[q] >> 8;
            const int reg = st[q] & 0xff;
            const int base = t * a.nregimes;
            const double2 c =
                reinterpret_cast<const double2 *>(ts)[base + reg];
            const double c_hi = c.y, c_lo = c.x;
            int nreg = reg;
            if (sh[q] < c_hi) {
                if (reg > 0 && sh[q] < c_lo)
                    nreg = reg - 1;
            } else {
                nreg = reg + 1;
            }
            if (nreg != reg)
                a.state[s0 + q] = (unsigned short)((t << 8) | nreg);
            if (nreg >= a.nregimes) {
                a.loc[s0 + q] = LOC_DEAD;
                if (a.brk)
                    a.brk[s0 + q] = a.step;
                a.gidx[s0 + q] |= (int)DEAD;
                ++nbroken;
                continue;
            }
            // (bond stiffness, plus_cs) of the new regime
            const double2 kp =
                reinterpret_cast<const double2 *>(tk)[base + nreg];
            const double fn = __builtin_fma(sh[q] + sh[q], kp.x, kp.y);
            g = fn * (hl[q] + hl[q]) * w[q];
        }
        if (CORR)
            g *= cr[q];
        fx = __builtin_fma(g, dx[q], fx);
        fy = __builtin_fma(g, dy[q], fy);
        fz = __builtin_fma(g, dz[q], fz);
    }
}

// A chunk's bond data: the halo indices and, for DM_NLIN, the bond states.
struct Chunk {
    ushort4 loc, st;
};

// A chunk of four bonds of node i, slots s0 .. s0+3.  The pieces of halo
// entry h are (r0x, r0y) (r0z, w) (rx, ry) (rz, v), as the records: w =
// 4 k V for DM_PMB, 2 V for DM_FIELD and V for DM_NLIN; v = 2 c V for
// DM_FIELD.  The bonds go two at a time: their reads are
// issued first and the arithmetic is straight-line code (a dead slot
// computes on the zero record and is masked), so the two chains interleave.
// (Four at a time is no faster and needs 40 more registers, which the
// field and n-linear models would spill.)  Only the exact-division path -- a stretch
// near a threshold, or a square root outside the core's range -- and
// breaking branch.
template <int DM, bool CORR, bool TENS, bool CLA>
__device__ __forceinline__ void chunk4(const StepArgs &a, const char *ldsb,
                                       const double4 &r0i, const double4 &ri,
                                       double ci,
                                       const Chunk &ch, size_t s0,
                                       const double *ts, const double *tk,
                                       const double *tp, double &fx,
                                       double &fy, double &fz, int &nbroken)
{
    const unsigned ev[4] = {ch.loc.x, ch.loc.y, ch.loc.z, ch.loc.w};
    const unsigned short st[4] = {ch.st.x, ch.st.y, ch.st.z, ch.st.w};
    bonds<DM, 2, CORR, TENS, CLA>(a, ldsb, r0i, ri, ci, ev, st, s0, ts, tk,
                                  tp, fx, fy, fz, nbroken);
    bonds<DM, 2, CORR, TENS, CLA>(a, ldsb, r0i, ri, ci, ev + 2, st + 2,
                                  s0 + 2, ts, tk, tp, fx, fy, fz, nbroken);
}

// DM: DM_PMB, DM_NLIN or DM_FIELD.  One workgroup per tile, LPN lanes per
// node taking its 4-bond chunks round-robin and reducing with cross-lane
// shuffles.  LDS (static, LDS_BYTES): the halo, four pieces of
// piece_bytes<DM>() each, then for DM_NLIN the three damage-model tables.
// CLA: the slot layout's OpenCL arithmetic (records hold u, updates
// contracted); otherwise the Cython path's (records hold r0 + u).
template <int DM, int LPN, bool CORR, bool TENS, bool CLA>
__global__ __launch_bounds__(1024) void k_step(StepArgs a, double dmag,
                                               double fmag)
{
    __shared__ double2 lds[LDS_BYTES / sizeof(double2)];
    const int b = blockIdx.x;
    const int tile = (b % XCDS) * a.tiles_per_xcd + b / XCDS;
    if (tile >= a.ntiles)
        return;
    const TileInfo ti = a.tiles[tile];
    char *ldsb = reinterpret_cast<char *>(lds);

    // node state and the first DEPTH chunks of halo indices are requested
    // before the halo is staged, so their latency overlaps the staging; the
    // bond loop then keeps DEPTH chunk loads in flight.  Chunk rows are
    // clamped to the tile's longest row (a node's surplus loads are unused).
    constexpr int DEPTH = 4;
    const int l = threadIdx.x / LPN;
    const int sub = threadIdx.x % LPN;
    const bool valid = l < ti.nodes;
    const ushort4 *L4 = reinterpret_cast<const ushort4 *>(a.loc);
    const long long c0 = ti.chunk_off + (valid ? l : 0);
    const int last_row = max(ti.nch, 1) - 1;
    const ushort4 *S4 = reinterpret_cast<const ushort4 *>(a.state);
    auto chunk_at = [&](int c) {
        const long long k = c0 + (long long)min(c, last_row) * ti.nodes;
        Chunk ch;
        ch.loc = L4[k];
        ch.st = DM == DM_NLIN ? S4[k] : make_ushort4(0, 0, 0, 0);
        return ch;
    };
    Chunk q0 = chunk_at(sub), q1 = chunk_at(sub + LPN),
          q2 = chunk_at(sub + 2 * LPN), q3 = chunk_at(sub + 3 * LPN);
    int cnt = 0;
    const uint32_t i = (uint32_t)(ti.node_off + (valid ? l : 0));
    double4 r0i = make_double4(0.0, 0.0, 0.0, 0.0), ri = r0i;
    if (valid) {
        const NodeRec *pi = a.rec_old + i;
        r0i = pi->a;
        ri = pi->b;
        cnt = a.cnt[i];
    }
    const double ci = DM == DM_FIELD && valid ? a.cnode[i] : 0.0;
    // with 4 or more lanes per node, lane d < 3 updates component d: its
    // inputs are requested now too (the prescribed values, needed only on
    // flagged DOFs, are read at the update)
    constexpr bool PRE = LPN >= 4;
    int pfl = 0;
    double pu = 0.0, pud = 0.0, pudd = 0.0, pdens = 1.0;
    if (PRE && valid && sub < 3) {
        const size_t k = 3 * (size_t)i + sub;
        pfl = a.flags[i];
        pu = a.u[k];
        if (a.integrator != PD_EULER) {
            pud = a.ud[k];
            pdens = a.dens1 ? a.dens1[i] : a.dens[k];
            if (a.integrator == PD_VELOCITY_VERLET)
                pudd = a.udd[k];
        }
    }

    // stage the halo: piece x copies 16-byte piece x%4 of record x/4 into
    // entry x/4 + 1, SP pieces per thread in flight at once (loads clamped in
    // bounds, stores predicated)
    constexpr int SP = 10;
    constexpr unsigned PB = piece_bytes<DM>();
    const int np = 4 * ti.halo;
    for (int x0 = threadIdx.x; x0 < np; x0 += SP * blockDim.x) {
        double2 v[SP];
#pragma unroll
        for (int k = 0; k < SP; ++k) {
            const int x = min(x0 + k * (int)blockDim.x, np - 1);
            const int g = a.halo[ti.halo_off + (x >> 2)];
            v[k] = reinterpret_cast<const double2 *>(a.rec_old + g)[x & 3];
        }
#pragma unroll
        for (int k = 0; k < SP; ++k) {
            const int x = x0 + k * (int)blockDim.x;
            if (x < np) {
                const double2 w = v[k];
                *reinterpret_cast<double2 *>(
                    ldsb + (x & 3) * PB + 16u * (unsigned)((x >> 2) + 1)) = w;
            }
        }
    }
    if (threadIdx.x < 4)  // entry 0: the zero record dead slots read
        *reinterpret_cast<double2 *>(ldsb + threadIdx.x * PB) =
            make_double2(0.0, 0.0);
    // n-linear tables, one 16-byte read each: (lower, upper) halved
    // thresholds of regime r of type t, and (stiffness, plus_cs)
    double2 *tsp = reinterpret_cast<double2 *>(ldsb + 4 * PB);
    double2 *tkp = tsp + MAX_TAB;
    const double *ts = reinterpret_cast<const double *>(tsp);
    const double *tk = reinterpret_cast<const double *>(tkp);
    const double *tp = nullptr;
    if (DM == DM_NLIN) {
        for (int t = threadIdx.x; t < a.ntab; t += blockDim.x) {
            const bool first = t % a.nregimes == 0;
            tsp[t] = make_double2(first ? -__DBL_MAX__ : a.s_tab[t - 1] * 0.5,
                                  a.s_tab[t] * 0.5);
            tkp[t] = make_double2(a.k_tab[t], a.pc_tab[t]);
        }
    }
    __syncthreads();

    double fx = 0.0, fy = 0.0, fz = 0.0;
    int nbroken = 0;
    if (valid) {
        const int nch = (cnt + 3) >> 2;
        for (int c = sub; c < nch; c += LPN) {
            const Chunk nx = chunk_at(c + DEPTH * LPN);
            const size_t s0 = (size_t)(c0 + (long long)c * ti.nodes) * 4;
            chunk4<DM, CORR, TENS, CLA>(a, ldsb, r0i, ri, ci, q0, s0, ts, tk,
                                        tp, fx, fy, fz, nbroken);
            q0 = q1;
            q1 = q2;
            q2 = q3;
            q3 = nx;
        }
    }

#pragma unroll
    for (int off = LPN / 2; off > 0; off >>= 1) {
        fx += __shfl_xor(fx, off, LPN);
        fy += __shfl_xor(fy, off, LPN);
        fz += __shfl_xor(fz, off, LPN);
        nbroken += __shfl_xor(nbroken, off, LPN);
    }
    // the compact layout (the Cython path, launched with !CLA only): once
    // the halo is no longer read, its LDS holds the rows being compacted
    if constexpr (!CLA) {
        if (a.layout == PD_LAYOUT_COMPACT)
            compact_tile<LPN, false>(a, ti, ldsb, LDS_BYTES, valid, l, sub,
                                     i, cnt, nbroken);
    }
    if (!valid)
        return;

    // the node's lanes share its three components (after the xor reduction
    // every lane holds the sums), so the per-node loads and stores coalesce
    const int fl = PRE ? pfl : a.flags[i];
    const size_t b3 = 3 * (size_t)i;
    for (int d = sub; d < 3; d += LPN) {
        const size_t k = b3 + d;
        const double bf = d == 0 ? fx : (d == 1 ? fy : fz);
        // force boundary condition after the bond sum
        // (peridynamics.pyx:93-97, cl/peridynamics.cl:117-129)
        // body_force differs from force only on a force-BC DOF, so only
        // those store it; reads take force elsewhere
        const bool fbc = (fl >> (3 + d)) & 1;
        double F = bf;
        if (fbc)
            F = CLA ? __builtin_fma(fmag, a.fbcv[k], bf)
                    : bf + fmag * a.fbcv[k];
        if (fbc)
            a.body_force[k] = bf;
        a.force[k] = F;
        const bool fixed = (fl >> d) & 1;
        double u = PRE ? pu : a.u[k];
        if (a.integrator == PD_EULER) {
            // peridynamics.pyx:179-185; cl/euler.cl:21-23 (contracted)
            const double un = CLA ? __builtin_fma(a.dt, F, u) : u + a.dt * F;
            u = fixed ? dmag * a.bcv[k] : un;
        } else if (a.integrator == PD_EULER_CROMER) {
            // cl/euler_cromer.cl:29-34 (contracted)
            double ud = PRE ? pud : a.ud[k];
            const double dens =
                PRE ? pdens : (a.dens1 ? a.dens1[i] : a.dens[k]);
            const double udd = __builtin_fma(-a.damping, ud, F) / dens;
            ud = __builtin_fma(udd, a.dt, ud);
            a.udd[k] = udd;
            a.ud[k] = ud;
            u = fixed ? dmag * a.bcv[k] : __builtin_fma(a.dt, ud, u);
        } else {
            // cl/velocity_verlet.cl:29-35 (contracted)
            const double hdt = a.dt / 2;
            const double ud1 = __builtin_fma(hdt, PRE ? pudd : a.udd[k],
                                             PRE ? pud : a.ud[k]);
            const double dens =
                PRE ? pdens : (a.dens1 ? a.dens1[i] : a.dens[k]);
            const double udd1 = __builtin_fma(-a.damping, ud1, F) / dens;
            const double ud = __builtin_fma(hdt, udd1, ud1);
            a.ud[k] = ud;
            a.udd[k] = udd1;
            u = fixed ? dmag * a.bcv[k]
                      : __builtin_fma(a.dt, __builtin_fma(hdt, udd1, ud), u);
        }
        a.u_out[k] = u;
        // the bond loop's position for the next step: r0 + u (compact) or
        // u (slot); the record's c (w) is static
        const double r0d = d == 0 ? r0i.x : (d == 1 ? r0i.y : r0i.z);
        reinterpret_cast<double *>(&a.rec_new[i].b)[d] = CLA ? u : r0d + u;
    }
}

// The stretch pipeline of bond<DM_PMB>, one bond per thread, for checking:
// l, l0 (bitwise those of spatial.pyx, or of cl/peridynamics.cl:65-75 as
// contracted with cla), the working s (2*sh), the exact s = (l - l0)/l0 and
// the break decision.  With cla, ri and rj are the displacements u_i, u_j
// (the OpenCL kernels form the bond from u); otherwise positions r0 + u.
template <bool CLA>
__device__ void bond_check_one(const double *ri, const double *rj,
                               const double *r0i, const double *r0j,
                               double s0, unsigned absmask, double &l,
                               double &l0, double &s, double &se, int &brk)
{
    const double ex = r0j[0] - r0i[0];
    const double ey = r0j[1] - r0i[1];
    const double ez = r0j[2] - r0i[2];
    double dx = rj[0] - ri[0], dy = rj[1] - ri[1], dz = rj[2] - ri[2];
    if (CLA) {
        dx = dx + ex;
        dy = dy + ey;
        dz = dz + ez;
    }
    const double q = sq3<CLA>(dx, dy, dz);
    const double q0 = sq3<CLA>(ex, ey, ez);
    double hl, h0;
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
    s = sh + sh;
    se = dl / l0;
    brk = !(as < s0h);
}

__global__ void k_bond_check(int n, int cla, const double *ri,
                             const double *rj, const double *r0i,
                             const double *r0j, double s0, unsigned absmask,
                             double *l_out, double *l0_out, double *s_out,
                             double *s_exact, int *broken)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    const size_t b = 3 * (size_t)t;
    double l, l0, s, se;
    int brk;
    if (cla)
        bond_check_one<true>(ri + b, rj + b, r0i + b, r0j + b, s0, absmask,
                             l, l0, s, se, brk);
    else
        bond_check_one<false>(ri + b, rj + b, r0i + b, r0j + b, s0, absmask,
                              l, l0, s, se, brk);
    l_out[t] = l;
    l0_out[t] = l0;
    s_out[t] = s;
    s_exact[t] = se;
    broken[t] = brk;
}

// Records of both ping-pong buffers from r0, V, u and c: a = (r0, wscale V),
// b = (r0 + u, v) -- (u, v) with ucla, the slot layout's OpenCL stretch --
// with v = 2 c V for a field c, else 0.  The weights are the factors the
// bond loop multiplies by: wscale = 4k for the scalar model, 2 for the field
// model, 1 for the n-linear model.
__global__ void k_records(int n, double wscale, int ucla, const double *r0,
                          const double *vol, const double *u, const double *c,
                          NodeRec *rec0, NodeRec *rec1)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const size_t b = 3 * (size_t)i;
    NodeRec r;
    r.a = make_double4(r0[b], r0[b + 1], r0[b + 2], vol[i] * wscale);
    const double v = c ? 2.0 * c[i] * vol[i] : 0.0;
    r.b = ucla ? make_double4(u[b], u[b + 1], u[b + 2], v)
               : make_double4(r0[b] + u[b], r0[b + 1] + u[b + 1],
                              r0[b + 2] + u[b + 2], v);
    rec0[i] = r;
    rec1[i] = r;
}

// Live-slot count per node, the damage (family - n)/family
// (peridynamics.pyx:29), and, for the slot layout, the reference-layout
// neighbour list: -1 everywhere but at the live slots' columns.  Thread p
// handles internal node p and writes row tnode[p] of the outputs.
// rows == 0 counts the live bonds without rebuilding the slot-layout rows.
__global__ void k_export(int n, int K, int layout, int rows, const int *tnode,
                         const int *cnt, const long long *chunk0,
                         const int *stride, const int *gidx,
                         const unsigned short *pos, const int *family,
                         const int *n_neigh_c, int *nlist_out,
                         int *n_neigh_out, double *damage)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n)
        return;
    const int i = tnode[p];
    int live;
    if (layout == PD_LAYOUT_SLOT) {
        int *row = nlist_out + (size_t)i * K;
        if (rows)
            for (int q = 0; q < K; ++q)
                row[q] = -1;
        live = 0;
        const int c = cnt[p];
        for (int k = 0; k < c; ++k) {
            const size_t s = slot_of(chunk0[p], stride[p], k);
            const int e = gidx[s];
            if (e >= 0) {
                if (rows)
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

// The compact layout's reference-layout rows in the caller's numbering: the
// positions below the row's count at set-up list their slots' keys (live
// ones first, then what swap-with-last left behind); the others keep the
// caller's values (uploaded into nlist_ref at set-up).
__global__ void k_export_rows(int n, int K, const int *tnode, const int *cnt,
                              const long long *chunk0, const int *stride,
                              const int *gidx, const unsigned short *rslot,
                              int *nlist_ref)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n)
        return;
    const size_t row = (size_t)tnode[p] * K;
    const int c = cnt[p];
    for (int q = 0; q < c; ++q)
        nlist_ref[row + q] =
            (int)((uint32_t)gidx[slot_of(chunk0[p], stride[p],
                                         rslot[row + q])] & KEYM);
}

// Monitored sums of a write (model.py:1320-1360).  Pass 1: block b sums a
// fixed strided share of the DOFs -- u, ud, udd, force*V, body_force*V --
// and of the nodes' damage, then reduces its 256 lanes by a fixed tree, so
// the result does not depend on scheduling.
constexpr int MON_BLOCKS = 512;
constexpr int MON_SUMS = 6;

__global__ void __launch_bounds__(256)
    k_monitor_part(int n, const double *u, const double *ud,
                   const double *udd, const double *force,
                   const double *body, const int *flags, const double *vol,
                   const double *damage, double *part)
{
    __shared__ double sh[MON_SUMS][256];
    const int t = threadIdx.x;
    const size_t first = (size_t)blockIdx.x * 256 + t;
    const size_t step = (size_t)MON_BLOCKS * 256;
    double a[MON_SUMS] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (size_t q = first; q < 3 * (size_t)n; q += step) {
        const size_t p = q / 3;
        const double v = vol[p];
        const bool fbc = (flags[p] >> (3 + (int)(q - 3 * p))) & 1;
        a[0] += u[q];
        a[1] += ud[q];
        a[2] += udd[q];
        a[3] += force[q] * v;
        a[4] += (fbc ? body[q] : force[q]) * v;
    }
    for (size_t i = first; i < (size_t)n; i += step)
        a[5] += damage[i];
    for (int k = 0; k < MON_SUMS; ++k)
        sh[k][t] = a[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
            for (int k = 0; k < MON_SUMS; ++k)
                sh[k][t] += sh[k][t + w];
        __syncthreads();
    }
    if (t < MON_SUMS)
        part[(size_t)blockIdx.x * MON_SUMS + t] = sh[t][0];
}

// Pass 2 (one block): the block sums in block order, and each tip group's
// sums accumulated entry by entry from 0.0 in the order the entries were
// given -- the reference's loop over the group's (node, dof) list
// (model.py:1320-1336), so these are its values bitwise.
__global__ void __launch_bounds__(256)
    k_monitor_final(const double *part, int ngroups, const int *tip_off,
                    const int *tip_q, const double *u, const double *ud,
                    const double *udd, const double *force,
                    const double *body, const int *flags, const double *vol,
                    double *out)
{
    const int t = threadIdx.x;
    if (t < MON_SUMS) {
        double acc = 0.0;
        for (int b = 0; b < MON_BLOCKS; ++b)
            acc += part[b * MON_SUMS + t];
        out[t] = acc;
    }
    for (int g = t; g < ngroups; g += blockDim.x) {
        double a[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int e = tip_off[g]; e < tip_off[g + 1]; ++e) {
            const int q = tip_q[e];
            const int p = q / 3;
            const double v = vol[p];
            const bool fbc = (flags[p] >> (3 + q - 3 * p)) & 1;
            a[0] += u[q];
            a[1] += ud[q];
            a[2] += udd[q];
            a[3] += force[q] * v;
            a[4] += (fbc ? body[q] : force[q]) * v;
        }
        for (int k = 0; k < 5; ++k)
            out[MON_SUMS + 5 * g + k] = a[k];
    }
}

__global__ void k_export_regimes(int n, int K, const int *tnode,
                                 const int *cnt, const long long *chunk0,
                                 const int *stride,
                                 const unsigned short *pos,
                                 const unsigned short *state, int *out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n)
        return;
    const size_t row = (size_t)tnode[p] * K;
    const int c = cnt[p];
    for (int k = 0; k < c; ++k) {
        const size_t s = slot_of(chunk0[p], stride[p], k);
        out[row + pos[s]] = state[s] & 0xff;
    }
}

__global__ void k_live(int n, const int *cnt, const long long *chunk0,
                       const int *stride, const int *gidx,
                       unsigned long long *total)
{
    __shared__ unsigned long long part[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long live = 0;
    if (i < n) {
        const int c = cnt[i];
        for (int k = 0; k < c; ++k)
            live += gidx[slot_of(chunk0[i], stride[i], k)] >= 0;
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

// --- exact-order Euler step of the Cython path ----------------------------
//
// peridynamics.pyx:34-99 (bond_force) visits i ascending and, for each live
// j > i of i's row in row order, adds f V_j to F_i and subtracts f V_i from
// F_j.  So F_m is the left fold, from +0.0, of first the bonds to live
// k < m in ascending k, then the bonds to the live j > m of m's current row
// in row order, then the force boundary condition.  Every term is
// (fn (r_j - r_m)) V_j with fn = ((l - l0)/l0 k)/l, l = |r_j - r_m|,
// l0 = |r0_j - r0_m| (spatial.pyx:20-31, 87-97): for a lower neighbour the
// reference subtracts (fn (r_m - r_k)) V_k, which is the same double,
// negated, as r_m - r_k is the exact negation of r_k - r_m.  The fold order
// depends only on node m's own row, so each tile folds its own nodes.
//
// A node's fold list is its lower slots in ascending order (slot order:
// rows are sorted by caller id; a dead slot adds nothing), then its upper
// neighbours in current row order (uord: slot index << 16 | halo loc,
// rebuilt by compact_tile after every break of the node).  The node's four
// lanes take four consecutive positions per round: each evaluates its bond
// with the reference's operations -- bitwise square roots, IEEE divisions,
// no contraction -- decides the break from that exact stretch
// (peridynamics.pyx:133), and writes its three force terms to the node's
// 96-byte LDS window; lanes 0-2 then add the four terms of their component
// in order.  Breaking is applied before the force, as in Euler.__call__: a
// bond that breaks adds nothing, and when any bond of the node breaks the
// row is compacted (peridynamics.pyx:102-156) and the upper part, whose
// order the compaction may change, is folded again from the sum of the
// lower part.  Assumes symmetric neighbour lists without repeats, as the
// model builds them.

// LDS of the exact step: the halo (pieces of EXACT_CAP entries), then the
// fold windows (12 doubles per node, tiles of up to EXACT_NODES nodes).
constexpr int EXACT_NODES = 128;
constexpr int EXACT_CAP = (int)((LDS_BYTES - EXACT_NODES * 96u) / 64u);

// Rows of K slots fit compact_tile's groups in the exact step's fold
// windows (its 64 bytes of counts, then at least one row).
__host__ __device__ constexpr bool exact_rows_fit(int K)
{
    return compact_row_bytes(K) + 64u <= EXACT_NODES * 96u;
}

// The exact bond evaluation: the term's components and the stretch.
struct ExactTerm {
    double t0, t1, t2, s;
};

// a / b rounded to nearest by the compiler's f64 division sequence (rcp, two
// Newton steps, the quotient and one residual correction) without its
// v_div_scale / v_div_fixup steps.  Those change nothing -- the result is
// the IEEE quotient -- when a is zero, or when a, b and a/b are normal, a >=
// 2^-969, 1/b is normal and a's exponent exceeds b's by less than 768.
__device__ __forceinline__ double div_core(double a, double b)
{
    double y = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    const double q = a * y;
    const double r = __builtin_fma(-b, q, a);
    return __builtin_fma(r, y, q);
}

// The exact bond evaluation.  With q and q0 at least 2^-700, l and l0 come
// from sqrt_core (exact there, and l, l0 lie in [2^-350, 2^512]).  With the
// stiffness k in [2^-100, 2^100] and s0 below 2^300 (SAFE, checked by the
// host) both quotients come from div_core: s = (l - l0)/l0 is 0 or at least
// 2^-54 in magnitude (l - l0 is 0 or at least an ulp of the shorter), and for a
// bond that does not break it is below s0, so every operand and quotient is
// normal, numerators exceed 2^-969 and no exponent gap reaches 768; a bond
// that breaks only needs its stretch on the right side of s0, which a huge
// (or NaN: l0 from a q0 outside the range) value is, as the reference's
// infinite or NaN one is.  The rest take the library's square roots and
// divisions.
template <bool SAFE>
__device__ __forceinline__ ExactTerm exact_term(const char *ldsb, unsigned e,
                                                const double4 &r0m,
                                                const double4 &rm, double k)
{
    constexpr unsigned PB = 16u * EXACT_CAP;
    const char *pe = ldsb + 16u * e;
    const double2 p0 = *reinterpret_cast<const double2 *>(pe);
    const double2 p1 = *reinterpret_cast<const double2 *>(pe + PB);
    const double2 p2 = *reinterpret_cast<const double2 *>(pe + 2 * PB);
    const double p3 = *reinterpret_cast<const double *>(pe + 3 * PB);
    // spatial.pyx:20-31: squares of r_j - r_m summed left to right
    const double dx = p2.x - rm.x, dy = p2.y - rm.y, dz = p3 - rm.z;
    const double q = (dx * dx + dy * dy) + dz * dz;
    const double ex = p0.x - r0m.x, ey = p0.y - r0m.y, ez = p1.x - r0m.z;
    const double q0 = (ex * ex + ey * ey) + ez * ez;
    const bool fast = (q >= 0x1p-700) & (q0 >= 0x1p-700);
    double l, l0, h;
    sqrt_core(q, l, h);
    sqrt_core(q0, l0, h);
    ExactTerm t;
    double fn;
    // spatial.pyx:87-97, peridynamics.pyx:75-89
    if (SAFE && fast) {
        t.s = div_core(l - l0, l0);
        fn = div_core(t.s * k, l);
    } else {
        if (!fast) {
            l = sqrt(q);
            l0 = sqrt(q0);
        }
        t.s = (l - l0) / l0;
        fn = t.s * k;
        fn = fn / l;
    }
    const double v = p1.y;  // V_j
    t.t0 = (fn * dx) * v;
    t.t1 = (fn * dy) * v;
    t.t2 = (fn * dz) * v;
    return t;
}

// The fold window of a node: [component][position of the round].
__device__ __forceinline__ void window_put(double *win, int sub,
                                          const ExactTerm &t)
{
    win[sub] = t.t0;
    win[4 + sub] = t.t1;
    win[8 + sub] = t.t2;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void window_done()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One workgroup per tile of at most EXACT_NODES nodes, four lanes per node
// (block 4 * tile_nodes), on k_step's tiles and halo staging.  Records hold
// (r0, V | r0 + u, 0).
template <bool SAFE>
__global__ __launch_bounds__(512) void k_exact_step(StepArgs a, double dmag,
                                                    double fmag)
{
    __shared__ double2 lds[LDS_BYTES / sizeof(double2)];
    constexpr int LPN = 4;
    constexpr unsigned PB = 16u * EXACT_CAP;
    const int b = blockIdx.x;
    const int tile = (b % XCDS) * a.tiles_per_xcd + b / XCDS;
    if (tile >= a.ntiles)
        return;
    const TileInfo ti = a.tiles[tile];
    char *ldsb = reinterpret_cast<char *>(lds);
    const int l = threadIdx.x / LPN;
    const int sub = threadIdx.x % LPN;
    const bool valid = l < ti.nodes;
    const uint32_t i = (uint32_t)(ti.node_off + (valid ? l : 0));
    const long long c0 = ti.chunk_off + (valid ? l : 0);
    double4 r0m = make_double4(0.0, 0.0, 0.0, 0.0), rm = r0m;
    int cnt = 0, nlow = 0, nup = 0, fl = 0;
    double pu = 0.0;
    if (valid) {
        r0m = a.rec_old[i].a;
        rm = a.rec_old[i].b;
        cnt = a.cnt[i];
        nlow = a.nlow[i];
        nup = a.nup[i];
        if (sub < 3) {
            fl = a.flags[i];
            pu = a.u[3 * (size_t)i + sub];
        }
    }
    // stage the halo (as k_step)
    constexpr int SP = 10;
    const int np = 4 * ti.halo;
    for (int x0 = threadIdx.x; x0 < np; x0 += SP * blockDim.x) {
        double2 v[SP];
#pragma unroll
        for (int k = 0; k < SP; ++k) {
            const int x = min(x0 + k * (int)blockDim.x, np - 1);
            const int g = a.halo[ti.halo_off + (x >> 2)];
            v[k] = reinterpret_cast<const double2 *>(a.rec_old + g)[x & 3];
        }
#pragma unroll
        for (int k = 0; k < SP; ++k) {
            const int x = x0 + k * (int)blockDim.x;
            if (x < np)
                *reinterpret_cast<double2 *>(
                    ldsb + (x & 3) * PB + 16u * (unsigned)((x >> 2) + 1)) =
                    v[k];
        }
    }
    if (threadIdx.x < 4)
        *reinterpret_cast<double2 *>(ldsb + threadIdx.x * PB) =
            make_double2(0.0, 0.0);
    __syncthreads();

    double *win = reinterpret_cast<double *>(ldsb + 4 * PB) + 12 * l;
    const double s0 = 2.0 * a.s0h;
    double F = 0.0, Flow = 0.0;  // lane d < 3: component d
    int nbroken = 0;
    if (valid) {
        const int total = nlow + nup;
        // fold position -> slot index << 16 | halo loc (lower positions are
        // the slots themselves, upper ones come from uord), requested two
        // rounds ahead of their use
        auto entry = [&](int pos) -> unsigned {
            if (pos < nlow)
                return ((unsigned)pos << 16) |
                       a.loc[slot_of(c0, ti.nodes, pos)];
            if (pos < total)
                return a.uord[slot_of(c0, ti.nodes, pos - nlow)];
            return LOC_DEAD;
        };
        unsigned ahead0 = entry(sub), ahead1 = entry(sub + LPN);
        for (int r = 0; r < total; r += LPN) {
            const unsigned cur = ahead0;
            ahead0 = ahead1;
            ahead1 = entry(r + 2 * LPN + sub);
            const unsigned e = cur & 0xffffu;
            const size_t slot = slot_of(c0, ti.nodes, (int)(cur >> 16));
            ExactTerm t = {0.0, 0.0, 0.0, 0.0};
            if (e != LOC_DEAD) {
                t = exact_term<SAFE>(ldsb, e, r0m, rm, a.kx);
                const double as = sign_mask(t.s, a.absmask);
                if (!(as < s0)) {  // peridynamics.pyx:133
                    a.loc[slot] = LOC_DEAD;
                    a.gidx[slot] |= (int)(DEAD | NEWB);
                    ++nbroken;
                    t.t0 = t.t1 = t.t2 = 0.0;
                }
            }
            window_put(win, sub, t);
            if (sub < 3) {
#pragma unroll
                for (int q = 0; q < LPN; ++q) {
                    if (r + q == nlow)
                        Flow = F;
                    F = F + win[4 * sub + q];
                }
            }
            window_done();
        }
        if (nlow >= total)
            Flow = F;
    }
#pragma unroll
    for (int off = LPN / 2; off > 0; off >>= 1)
        nbroken += __shfl_xor(nbroken, off, LPN);
    // the fold windows are free: they hold the rows being compacted, and
    // the upper lists of those nodes are rebuilt in the new row order
    compact_tile<LPN, true>(a, ti, ldsb + 4 * PB, EXACT_NODES * 96u, valid, l,
                            sub, i, cnt, nbroken);
    if (!valid)
        return;
    if (nbroken > 0) {
        // the upper part again, in the new row order
        const int nup2 = ld_slot(&a.nup[i]);
        F = Flow;
        for (int r = 0; r < nup2; r += LPN) {
            const int pos = r + sub;
            ExactTerm t = {0.0, 0.0, 0.0, 0.0};
            if (pos < nup2) {
                const unsigned u = ld_slot(reinterpret_cast<const int *>(
                    &a.uord[slot_of(c0, ti.nodes, pos)]));
                t = exact_term<SAFE>(ldsb, u & 0xffffu, r0m, rm, a.kx);
            }
            window_put(win, sub, t);
            if (sub < 3) {
#pragma unroll
                for (int q = 0; q < LPN; ++q)
                    F = F + win[4 * sub + q];
            }
            window_done();
        }
    }
    if (sub < 3) {
        // peridynamics.pyx:93-97, then :179-185
        const size_t k = 3 * (size_t)i + sub;
        const bool fbc = (fl >> (3 + sub)) & 1;
        double Fd = F;
        if (fbc) {
            Fd = F + fmag * a.fbcv[k];
            a.body_force[k] = F;
        }
        a.force[k] = Fd;
        const double u = (fl >> sub) & 1 ? dmag * a.bcv[k] : pu + a.dt * Fd;
        a.u_out[k] = u;
        const double r0d = sub == 0 ? r0m.x : (sub == 1 ? r0m.y : r0m.z);
        reinterpret_cast<double *>(&a.rec_new[i].b)[sub] = r0d + u;
    }
}

// --- adjoint of the Euler step with a recorded break mask -------------------
//
// Forward (PD_EULER, node stiffness field c): u_{n+1} = u_n + dt F(u_n) on
// free DOFs (prescribed DOFs are set), with
//   F_m = sum_j k_mj V_j phi(d_mj),  phi(d) = (1/l0 - 1/l) d = (s/l) d,
//   d_mj = r_j - r_m, k_mj = (c_m + c_j)/2,
// over the bonds live in step n (broken at a step > n, from brk).  For a
// loss L(u_N) and mu = P_free lambda_{n+1}:
//   lambda_n = mu + dt J_n^T mu,
//   (J^T mu)_m = sum_j k_mj A_mj (V_m mu_j - V_j mu_m),
//   A = dphi/dd = (1/l0 - 1/l) I + d d^T / l^3  (symmetric, even in d),
//   dL/dc_m += dt/2 sum_j (1/l0 - 1/l)_mj d_mj . (V_j mu_m - V_m mu_j).

// Bond state at the start of step n0: alive iff it broke at a step >= n0.
__global__ void k_adj_restore(int n, int n0, const int *cnt,
                              const long long *chunk0, const int *stride,
                              const int *brk, const unsigned short *loc0,
                              unsigned short *loc, int *gidx)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n)
        return;
    const int c = cnt[p];
    for (int k = 0; k < c; ++k) {
        const size_t s = slot_of(chunk0[p], stride[p], k);
        const bool alive = brk[s] >= n0;
        loc[s] = alive ? loc0[s] : LOC_DEAD;
        const uint32_t e = (uint32_t)gidx[s] & ~(DEAD | NEWB);
        gidx[s] = (int)(alive ? e : (e | DEAD));
    }
}

// A per-node array in the caller's node order: dst[tnode[p]] = src[p], w
// values per node.
__global__ void k_unperm(int n, int w, const int *tnode, const double *src,
                         double *dst)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n)
        return;
    const size_t c = (size_t)tnode[p] * w, q = (size_t)p * w;
    for (int d = 0; d < w; ++d)
        dst[c + d] = src[q + d];
}

// A per-node array in the library's node order from the caller's:
// dst[p] = src[tnode[p]], w values per node.
__global__ void k_perm(int n, int w, const int *tnode, const double *src,
                       double *dst)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n)
        return;
    const size_t c = (size_t)tnode[p] * w, q = (size_t)p * w;
    for (int d = 0; d < w; ++d)
        dst[q + d] = src[c + d];
}

// brk at the start of a recording: -1 for bonds already broken, else "never".
__global__ void k_adj_init(int n, const int *cnt, const long long *chunk0,
                           const int *stride, const unsigned short *loc,
                           int *brk)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n)
        return;
    const int c = cnt[p];
    for (int k = 0; k < c; ++k) {
        const size_t s = slot_of(chunk0[p], stride[p], k);
        brk[s] = loc[s] == LOC_DEAD ? -1 : 0x7fffffff;
    }
}

// 1/sqrt(q) for q in [2^-767, 2^767]: the hardware estimate and two Newton
// steps y <- y (3/2 - q y^2 / 2).
__device__ __forceinline__ double rsqrt_newton(double q)
{
    double y = __builtin_amdgcn_rsq(q);
    const double hq = 0.5 * q;
    y = y * __builtin_fma(-hq * y, y, 1.5);
    y = y * __builtin_fma(-hq * y, y, 1.5);
    return y;
}

// 1/l0 of every slot's bond, once per recording (inv: caller id -> internal
// id); the slots past a node's count keep the 0 they were cleared to.  The
// same reciprocal square root of the same squared length as the backward
// kernel's 1/l, so that a bond with l == l0 (bitwise, e.g. both ends at
// rest) gets 1/l0 - 1/l = 0 exactly, as its stretch is 0 in the forward
// step: a one-ulp difference there, times k dt, is not small.
__global__ void k_adj_il0(int n, const int *cnt, const long long *chunk0,
                          const int *stride, const int *gidx, const int *inv,
                          const double *r0, double *il0)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n)
        return;
    const int c = cnt[p];
    const size_t b = 3 * (size_t)p;
    for (int k = 0; k < c; ++k) {
        const size_t s = slot_of(chunk0[p], stride[p], k);
        const size_t bj = 3 * (size_t)inv[(uint32_t)gidx[s] & KEYM];
        const double ex = r0[bj] - r0[b];
        const double ey = r0[bj + 1] - r0[b + 1];
        const double ez = r0[bj + 2] - r0[b + 2];
        il0[s] = rsqrt_newton((ex * ex + ey * ey) + ez * ez);
    }
}

// The per-node inputs of one backward step as 64-byte records, from which
// the tiles stage their halos: a = (rx, ry, rz, V), b = (mux, muy, muz, c)
// with r = r0 + u_n and mu = lambda_{n+1} with prescribed DOFs zeroed.
__global__ void k_adj_pack(int n, const double *r0, const double *u,
                           const double *vol, const double *c,
                           const int *flags, const double *lam, NodeRec *rec)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n)
        return;
    const size_t b = 3 * (size_t)p;
    const int f = flags[p];
    NodeRec o;
    o.a = make_double4(r0[b] + u[b], r0[b + 1] + u[b + 1],
                       r0[b + 2] + u[b + 2], vol[p]);
    o.b = make_double4(f & 1 ? 0.0 : lam[b], f & 2 ? 0.0 : lam[b + 1],
                       f & 4 ? 0.0 : lam[b + 2], c[p]);
    rec[p] = o;
}

struct AdjArgs {
    int ntiles, tiles_per_xcd, step;
    double dt;
    const TileInfo *tiles;
    const int *halo, *cnt, *brk;
    const NodeRec *rec;  // the records of this step
    const unsigned short *loc0;  // loc at the start of the recording
    const double *il0;
    double *lam_out, *gradc;
    // if set, the records of the step before (k_adj_pack's, from u_prev =
    // u_{n-1} and lambda_n) are written too
    NodeRec *rec_next;
    const double *u_prev, *r0;
    const int *flags;
};

// One backward step on the step kernel's tiles: the tile's halo records
// (k_adj_pack) are staged into LDS piece-major, as in k_step, and every bond
// of a node reads its neighbour from there.  LPN lanes per node take its
// 4-bond chunks round-robin; a slot is live in step n if it held a bond at
// the start of the recording (loc0) that broke at a step > n (brk).
template <int LPN>
__global__ __launch_bounds__(1024) void k_adj_tile(AdjArgs a)
{
    __shared__ double2 lds[LDS_BYTES / sizeof(double2)];
    const int tile = (blockIdx.x % XCDS) * a.tiles_per_xcd + blockIdx.x / XCDS;
    if (tile >= a.ntiles)
        return;
    const TileInfo ti = a.tiles[tile];
    char *ldsb = reinterpret_cast<char *>(lds);
    const int l = threadIdx.x / LPN;
    const int sub = threadIdx.x % LPN;
    const bool valid = l < ti.nodes;
    const uint32_t i = (uint32_t)(ti.node_off + (valid ? l : 0));
    const NodeRec own = a.rec[i];
    const int cnt = valid ? a.cnt[i] : 0;

    constexpr int SP = 10;
    constexpr unsigned PB = 16u * LDS_CAP_PMB;
    const int np = 4 * ti.halo;
    for (int x0 = threadIdx.x; x0 < np; x0 += SP * blockDim.x) {
        double2 v[SP];
#pragma unroll
        for (int k = 0; k < SP; ++k) {
            const int x = min(x0 + k * (int)blockDim.x, np - 1);
            const int g = a.halo[ti.halo_off + (x >> 2)];
            v[k] = reinterpret_cast<const double2 *>(a.rec + g)[x & 3];
        }
#pragma unroll
        for (int k = 0; k < SP; ++k) {
            const int x = x0 + k * (int)blockDim.x;
            if (x < np)
                *reinterpret_cast<double2 *>(
                    ldsb + (x & 3) * PB + 16u * (unsigned)((x >> 2) + 1)) =
                    v[k];
        }
    }
    __syncthreads();

    const double rm[3] = {own.a.x, own.a.y, own.a.z};
    const double mum[3] = {own.b.x, own.b.y, own.b.z};
    const double vm = own.a.w, cm = own.b.w;
    double jt[3] = {0.0, 0.0, 0.0};
    double gc = 0.0;
    const ushort4 *L4 = reinterpret_cast<const ushort4 *>(a.loc0);
    const int4 *B4 = reinterpret_cast<const int4 *>(a.brk);
    const double2 *I2 = reinterpret_cast<const double2 *>(a.il0);
    const long long c0 = ti.chunk_off + l;
    const int nch = (cnt + 3) >> 2;
    for (int c = sub; c < nch; c += LPN) {
        const long long k = c0 + (long long)c * ti.nodes;
        const ushort4 lc = L4[k];
        const int4 bk = B4[k];
        const double2 ia = I2[2 * k], ib = I2[2 * k + 1];
        const unsigned ev[4] = {lc.x, lc.y, lc.z, lc.w};
        const int br[4] = {bk.x, bk.y, bk.z, bk.w};
        const double il0[4] = {ia.x, ia.y, ib.x, ib.y};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (ev[q] == LOC_DEAD || br[q] <= a.step)
                continue;  // empty, or broken before or in this step
            const char *pe = ldsb + 16u * ev[q];
            const double2 p0 = *reinterpret_cast<const double2 *>(pe);
            const double2 p1 = *reinterpret_cast<const double2 *>(pe + PB);
            const double2 p2 =
                *reinterpret_cast<const double2 *>(pe + 2 * PB);
            const double2 p3 =
                *reinterpret_cast<const double2 *>(pe + 3 * PB);
            const double dd[3] = {p0.x - rm[0], p0.y - rm[1], p1.x - rm[2]};
            const double muj[3] = {p2.x, p2.y, p3.x};
            const double vj = p1.y, cj = p3.y;
            double w[3];
#pragma unroll
            for (int d = 0; d < 3; ++d)
                w[d] = vm * muj[d] - vj * mum[d];
            // 1/l from the reciprocal square root of the squared length,
            // two Newton steps (relative error ~1e-16): no square root or
            // division in the loop
            const double il =
                rsqrt_newton((dd[0] * dd[0] + dd[1] * dd[1]) + dd[2] * dd[2]);
            const double psi = il0[q] - il;
            const double kk = (cm + cj) * 0.5;
            const double dw = dd[0] * w[0] + dd[1] * w[1] + dd[2] * w[2];
            const double il3 = il * il * il;
#pragma unroll
            for (int d = 0; d < 3; ++d)
                jt[d] += kk * (psi * w[d] + dd[d] * dw * il3);
            gc -= 0.5 * psi * dw;
        }
    }
#pragma unroll
    for (int off = LPN / 2; off > 0; off >>= 1) {
        jt[0] += __shfl_xor(jt[0], off, LPN);
        jt[1] += __shfl_xor(jt[1], off, LPN);
        jt[2] += __shfl_xor(jt[2], off, LPN);
        gc += __shfl_xor(gc, off, LPN);
    }
    if (!valid)
        return;
    const size_t b = 3 * (size_t)i;
    const int fl = a.rec_next ? a.flags[i] : 0;
    for (int d = sub; d < 4; d += LPN) {
        if (d < 3) {
            const double lam = mum[d] + a.dt * jt[d];
            a.lam_out[b + d] = lam;
            if (a.rec_next) {
                // as k_adj_pack: r = r0 + u_{n-1}, mu = lambda_n masked
                double *ra = reinterpret_cast<double *>(&a.rec_next[i].a);
                double *rb = reinterpret_cast<double *>(&a.rec_next[i].b);
                ra[d] = a.r0[b + d] + a.u_prev[b + d];
                rb[d] = (fl >> d) & 1 ? 0.0 : lam;
            }
        } else if (a.rec_next) {
            reinterpret_cast<double *>(&a.rec_next[i].a)[3] = vm;
            reinterpret_cast<double *>(&a.rec_next[i].b)[3] = cm;
        }
    }
    if (sub == 0)
        a.gradc[i] += a.dt * gc;
}

}  // namespace pdk

// Device scratch of pd_monitor_noalloc: the per-block partial sums of
// k_monitor_part, then the staged outputs of k_monitor_final (6 model sums
// and 5 per tip group), for at most NOALLOC_MAX_GROUPS tip groups.  File
// scope: every device the library runs on holds its own copy (64 KiB) until
// the process ends.  Host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device.
constexpr int NOALLOC_MAX_GROUPS = 1024;
constexpr size_t NOALLOC_DOUBLES = (size_t)pdk::MON_BLOCKS * pdk::MON_SUMS +
                                   pdk::MON_SUMS + 5 * NOALLOC_MAX_GROUPS;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

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
    int n = 0, K = 0;
    int tension_only = 0;
    int lpn = 4, tile_nodes = 128;
    int hmax = 0;           // halo capacity of a tile (LDS budget)
    int hstride = 0;        // largest halo of the current tiles
    size_t lds_budget = 0;  // bytes of LDS a workgroup may use
    int ntiles = 0;
    bool retile_pending = false;  // tile_nodes changed since the tiling
    bool has_buffers = false, has_corr = false, has_btypes = false,
         has_cnode = false;
    int nregimes = 1, nbond_types = 1;
    double k = 0.0, s0 = 0.0;
    int cur = 0;
    bool exact = false;  // exact-order Euler step (k_exact_step)
    // a replay's u_n and u_{n+1} (slots of the segment buffer); null: u
    const double *u_in = nullptr;
    double *u_out = nullptr;
    DevBuf<int> nlow, nup;  // exact step: lower slots, upper-list length
    DevBuf<unsigned> uord;  // exact step: upper lists in row order
    DevBuf<double> wbuf;  // a per-node array reordered for download
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    // build inputs in the caller's node order; uploaded in tile order by
    // pd_create_buffers, which fixes that order
    std::vector<double> h_coords, h_vol, h_bcv, h_fbcv, h_dens, h_cnode;
    std::vector<int> h_flags;
    std::vector<double> h_corr;    // (N,K) as given to build
    std::vector<int> h_btypes;     // (N,K)
    std::vector<int> perm;         // internal node p -> caller's node id
    std::vector<int> h_group;      // tiling groups (empty: one group)

    DevBuf<NodeRec> rec[2];
    DevBuf<TileInfo> tiles;
    DevBuf<double> dens1;
    DevBuf<double> r0, vol, u, ud, udd, force, body_force, bcv, fbcv, dens,
        damage, cnode, corr, k_tab, s_tab, pc_tab;
    DevBuf<double> cnode_caller;  // a field given on the device (caller order)
    DevBuf<int> flags, family, cnt, tnode, halo, gidx, stride, nlist_ref,
        n_neigh, out_nlist, out_nneigh;
    DevBuf<long long> chunk0;
    DevBuf<unsigned short> loc, pos, state, rslot;
    DevBuf<unsigned long long> counter;

    // adjoint recording (Euler, node field): break steps, checkpoints
    bool adj_recording = false, adj_ready = false, adj_replay = false;
    int adj_nsteps = 0, adj_every = 0, step_counter = 0;
    std::vector<double> adj_dmag, adj_fmag;
    DevBuf<int> brk;
    DevBuf<unsigned short> loc0;
    DevBuf<double> il0;    // 1/l0 per slot
    DevBuf<NodeRec> arec[2];  // records of a backward step, ping-pong
    DevBuf<double> ckpt, seg, lam[2], gradc, fin_u, fin_force, fin_bforce;

    // monitored sums: tip entries (caller's node, dof) grouped, in order
    std::vector<int> h_tip_node, h_tip_dof, h_tip_off;
    DevBuf<int> tip_q, tip_off;  // internal DOF index 3p + d; group offsets
    DevBuf<double> mon_part, mon_out;
};

namespace {

void run_parallel(int n, const std::function<void(int, int)> &fn,
                  int min_per_thread = 4096)
{
    const unsigned hw = std::thread::hardware_concurrency();
    int nt = (int)std::min<unsigned>(hw ? hw : 1, 16);
    if (n < min_per_thread)
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

// dst[p] = src[perm[p]] for rows of w values (caller order -> tile order)
template <typename T>
std::vector<T> to_internal(const std::vector<int> &perm, const T *src, int w)
{
    std::vector<T> dst(perm.size() * (size_t)w);
    for (size_t p = 0; p < perm.size(); ++p)
        for (int d = 0; d < w; ++d)
            dst[p * w + d] = src[(size_t)perm[p] * w + d];
    return dst;
}

// Upload rows of w values given in the caller's order into tile order.
template <typename T>
hipError_t upload_perm(DevBuf<T> &d, const std::vector<int> &perm,
                       const T *src, int w)
{
    const std::vector<T> v = to_internal(perm, src, w);
    return upload(d, v.data(), v.size());
}

// Download rows of w values in tile order into the caller's order.
// n*w doubles of an internal-order device array into dst in the caller's
// node order: reordered on the device, then one copy.
int download_caller(pd_sim *s, double *dst, const double *src, int w)
{
    const size_t count = (size_t)s->n * w;
    if (count == 0)
        return PD_OK;
    if (s->wbuf.n < count)
        HIPCHK(s->wbuf.alloc(count));
    hipLaunchKernelGGL(k_unperm, dim3((s->n + 255) / 256), dim3(256), 0,
                       s->stream, s->n, w, s->tnode.p, src, s->wbuf.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(dst, s->wbuf.p, count * sizeof(double),
                          hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return PD_OK;
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
    a.K = s->K;
    a.hstride = s->hstride;
    a.ntiles = s->ntiles;
    a.tiles_per_xcd = (s->ntiles + XCDS - 1) / XCDS;
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
    a.rec_old = s->rec[s->cur].p;
    a.rec_new = s->rec[1 - s->cur].p;
    a.tiles = s->tiles.p;
    a.tnode = s->tnode.p;
    a.halo = s->halo.p;
    a.u = s->u_in ? s->u_in : s->u.p;
    a.u_out = s->u_out ? s->u_out : s->u.p;
    a.ud = s->ud.p;
    a.udd = s->udd.p;
    a.force = s->force.p;
    a.body_force = s->body_force.p;
    a.flags = s->flags.p;
    a.bcv = s->bcv.p;
    a.fbcv = s->fbcv.p;
    a.dens = s->dens.p;
    a.dens1 = s->dens1.p;
    a.cnt = s->cnt.p;
    a.loc = s->loc.p;
    a.gidx = s->gidx.p;
    a.corr = s->corr.p;
    a.state = s->state.p;
    a.k_tab = s->k_tab.p;
    a.s_tab = s->s_tab.p;
    a.pc_tab = s->pc_tab.p;
    a.chunk0 = s->chunk0.p;
    a.stride = s->stride.p;
    a.rslot = s->rslot.p;
    a.n_neigh = s->n_neigh.p;
    a.cnode = s->cnode.p;
    a.brk = s->adj_recording ? s->brk.p : nullptr;
    a.step = s->step_counter;
    a.nlow = s->nlow.p;
    a.nup = s->nup.p;
    a.uord = s->uord.p;
    a.kx = s->k;
    return a;
}

template <int DM, int LPN, bool CLA>
hipError_t launch_cla(pd_sim *s, const StepArgs &a, double dmag, double fmag)
{
    const unsigned grid = (unsigned)XCDS * a.tiles_per_xcd;
    const unsigned block = (unsigned)(s->tile_nodes * LPN);
    // the n-linear model breaks by regimes, not by the sign mode
    const bool tens = DM != DM_NLIN && a.absmask == 0xffffffffu;
    if (a.has_corr && tens)
        hipLaunchKernelGGL((k_step<DM, LPN, true, DM != DM_NLIN, CLA>),
                           dim3(grid), dim3(block), 0, s->stream, a, dmag,
                           fmag);
    else if (a.has_corr)
        hipLaunchKernelGGL((k_step<DM, LPN, true, false, CLA>), dim3(grid),
                           dim3(block), 0, s->stream, a, dmag, fmag);
    else if (tens)
        hipLaunchKernelGGL((k_step<DM, LPN, false, DM != DM_NLIN, CLA>),
                           dim3(grid), dim3(block), 0, s->stream, a, dmag,
                           fmag);
    else
        hipLaunchKernelGGL((k_step<DM, LPN, false, false, CLA>), dim3(grid),
                           dim3(block), 0, s->stream, a, dmag, fmag);
    return hipGetLastError();
}

// The slot layout steps with the OpenCL arithmetic; the compact layout (the
// Cython path: scalar stiffness, no corrections) with the Cython one.
template <int DM, int LPN>
hipError_t launch_one(pd_sim *s, const StepArgs &a, double dmag, double fmag)
{
    if (a.layout == PD_LAYOUT_SLOT)
        return launch_cla<DM, LPN, true>(s, a, dmag, fmag);
    const unsigned grid = (unsigned)XCDS * a.tiles_per_xcd;
    const unsigned block = (unsigned)(s->tile_nodes * LPN);
    if (DM != DM_PMB || a.has_corr)
        return hipErrorInvalidValue;  // rejected before launch
    if (a.absmask == 0xffffffffu)
        hipLaunchKernelGGL((k_step<DM_PMB, LPN, false, true, false>),
                           dim3(grid), dim3(block), 0, s->stream, a, dmag,
                           fmag);
    else
        hipLaunchKernelGGL((k_step<DM_PMB, LPN, false, false, false>),
                           dim3(grid), dim3(block), 0, s->stream, a, dmag,
                           fmag);
    return hipGetLastError();
}

template <int DM>
hipError_t launch_dm(pd_sim *s, const StepArgs &a, double dmag, double fmag)
{
    switch (s->lpn) {
    case 1:
        return launch_one<DM, 1>(s, a, dmag, fmag);
    case 2:
        return launch_one<DM, 2>(s, a, dmag, fmag);
    case 4:
        return launch_one<DM, 4>(s, a, dmag, fmag);
    default:
        return launch_one<DM, 8>(s, a, dmag, fmag);
    }
}


// One backward step's k_adj_tile, with the step kernel's grid and block.
hipError_t launch_adjoint(pd_sim *s, const AdjArgs &a)
{
    const unsigned grid = (unsigned)XCDS * a.tiles_per_xcd;
    switch (s->lpn) {
    case 1:
        hipLaunchKernelGGL(k_adj_tile<1>, dim3(grid), dim3(s->tile_nodes), 0,
                           s->stream, a);
        break;
    case 2:
        hipLaunchKernelGGL(k_adj_tile<2>, dim3(grid),
                           dim3(2 * s->tile_nodes), 0, s->stream, a);
        break;
    case 4:
        hipLaunchKernelGGL(k_adj_tile<4>, dim3(grid),
                           dim3(4 * s->tile_nodes), 0, s->stream, a);
        break;
    default:
        hipLaunchKernelGGL(k_adj_tile<8>, dim3(grid),
                           dim3(8 * s->tile_nodes), 0, s->stream, a);
        break;
    }
    return hipGetLastError();
}

void adjoint_release(pd_sim *s)
{
    s->adj_recording = s->adj_ready = s->adj_replay = false;
    s->adj_nsteps = s->adj_every = s->step_counter = 0;
    s->adj_dmag.clear();
    s->adj_fmag.clear();
    s->brk.release();
    s->loc0.release();
    s->il0.release();
    s->arec[0].release();
    s->arec[1].release();
    s->ckpt.release();
    s->seg.release();
    s->lam[0].release();
    s->lam[1].release();
    s->gradc.release();
    s->fin_u.release();
    s->fin_force.release();
    s->fin_bforce.release();
}

// Bookkeeping of an adjoint recording after each forward step: the step's
// magnitudes, and a checkpoint of u every adj_every steps.
int after_step(pd_sim *s, double dmag, double fmag)
{
    if (s->adj_replay)
        return PD_OK;
    if (!s->adj_recording) {
        s->adj_ready = false;  // stepped past a finished recording
        return PD_OK;
    }
    s->adj_dmag.push_back(dmag);
    s->adj_fmag.push_back(fmag);
    ++s->step_counter;
    const size_t B = 3 * (size_t)s->n;
    if (s->step_counter % s->adj_every == 0 &&
        s->step_counter < s->adj_nsteps)
        HIPCHK(hipMemcpyAsync(s->ckpt.p + (s->step_counter / s->adj_every) * B,
                              s->u.p, B * sizeof(double),
                              hipMemcpyDeviceToDevice, s->stream));
    if (s->step_counter == s->adj_nsteps) {
        s->adj_recording = false;
        s->adj_ready = true;
    }
    return PD_OK;
}

int launch_step(pd_sim *s, double dmag, double fmag)
{
    if (s->retile_pending)
        return fail(PD_E_STATE, "tile_nodes changed: call pd_create_buffers "
                                "to re-tile before stepping");
    if (s->ntiles == 0) {
        s->cur = 1 - s->cur;
        return PD_OK;
    }
    if (s->layout == PD_LAYOUT_COMPACT && dm_of(s) != DM_PMB)
        return fail(PD_E_STATE, "the compact (Cython) layout steps the "
                                "scalar stiffness model only");
    if (s->exact && (s->layout != PD_LAYOUT_COMPACT ||
                     s->integrator != PD_EULER || dm_of(s) != DM_PMB ||
                     s->has_corr || !s->uord.p))
        return fail(PD_E_STATE, "the exact-order step needs the compact "
                                "layout, the Euler integrator and the scalar "
                                "model without corrections");
    if (s->exact && (s->lpn != 4 || s->tile_nodes > EXACT_NODES))
        return fail(PD_E_STATE, "the exact-order step runs 4 lanes per node "
                                "and tiles of at most 128 nodes");
    if (s->exact && !exact_rows_fit(s->K))
        return fail(PD_E_LIMIT, "the exact-order step takes at most 4074 "
                                "neighbours per node");
    const StepArgs a = make_args(s);
    hipError_t e;
    if (s->exact) {
        // the stiffness range in which div_core is the IEEE division
        const double ak = std::fabs(s->k);
        if (ak >= 0x1p-100 && ak <= 0x1p100 && s->s0 < 0x1p300)
            hipLaunchKernelGGL(k_exact_step<true>,
                               dim3(XCDS * a.tiles_per_xcd),
                               dim3(4 * s->tile_nodes), 0, s->stream, a, dmag,
                               fmag);
        else
            hipLaunchKernelGGL(k_exact_step<false>,
                               dim3(XCDS * a.tiles_per_xcd),
                               dim3(4 * s->tile_nodes), 0, s->stream, a, dmag,
                               fmag);
        e = hipGetLastError();
    } else {
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
    }
    if (e != hipSuccess)
        return fail(PD_E_HIP, std::string("step kernel launch: ") +
                                  hipGetErrorString(e));
    s->cur = 1 - s->cur;
    return after_step(s, dmag, fmag);
}

// The tip entries as internal DOF indices 3p + d for the current tiling.
int upload_tips(pd_sim *s)
{
    if (s->h_tip_off.empty() || s->perm.empty())
        return PD_OK;
    std::vector<int> inv(s->n);
    for (int p = 0; p < s->n; ++p)
        inv[s->perm[p]] = p;
    std::vector<int> q(s->h_tip_node.size());
    for (size_t e = 0; e < q.size(); ++e)
        q[e] = 3 * inv[s->h_tip_node[e]] + s->h_tip_dof[e];
    HIPCHK(upload(s->tip_q, q.data(), q.size()));
    HIPCHK(upload(s->tip_off, s->h_tip_off.data(), s->h_tip_off.size()));
    return PD_OK;
}

// (Re)write both record buffers from r0, V, the device u and c.
int refresh_records(pd_sim *s)
{
    const int blocks = (s->n + 255) / 256;
    const int dm = dm_of(s);
    // the exact-order step reads V itself
    const double wscale =
        s->exact ? 1.0
                 : (dm == DM_PMB ? 4.0 * s->k : (dm == DM_FIELD ? 2.0 : 1.0));
    hipLaunchKernelGGL(k_records, dim3(blocks), dim3(256), 0, s->stream, s->n,
                       wscale, s->layout == PD_LAYOUT_SLOT ? 1 : 0, s->r0.p,
                       s->vol.p, s->u.p,
                       s->has_cnode ? s->cnode.p : nullptr, s->rec[0].p,
                       s->rec[1].p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    s->cur = 0;
    return PD_OK;
}

uint64_t spread3(uint64_t v)
{
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

// LDS entries of a tile's halo, chosen so that the records the bond loop
// reads together fall into different banks.  One ds_read_b128 over 16
// consecutive lanes reads a 16-byte piece of each lane's neighbour record;
// records whose entries are equal modulo HALO_BANK_GROUPS share banks, and
// distinct records in one bank are read one after another.  rows[l] lists
// the tile's node l's neighbours as halo indices in slot order; lane
// t = l * lpn + sub takes the chunks of four slots sub, sub + lpn, ...  Read
// groups are (16 lanes, round, bond of the chunk).  Greedy: records in the
// most groups first, each to the residue least loaded over its groups among
// the entries left (entry 0 is the zero record).  Returns entry - 1 for each
// halo index; the order changes no result.
constexpr int HALO_BANK_GROUPS = 16;

std::vector<int> spread_halo(const std::vector<std::vector<int>> &rows,
                             int lpn, int H)
{
    constexpr int NB = HALO_BANK_GROUPS;
    std::vector<std::vector<int>> memb(H);
    std::vector<int> seen(H, -1);
    const int lanes = (int)rows.size() * lpn;
    int ngroups = 0;
    for (int q0 = 0; q0 < lanes; q0 += 16) {
        const int q1 = std::min(q0 + 16, lanes);
        int rounds = 0;
        for (int t = q0; t < q1; ++t) {
            const int nch = ((int)rows[t / lpn].size() + 3) / 4;
            rounds = std::max(rounds, (nch - t % lpn + lpn - 1) / lpn);
        }
        for (int k = 0; k < rounds; ++k)
            for (int b = 0; b < 4; ++b) {
                const int g = ngroups++;
                for (int t = q0; t < q1; ++t) {
                    const std::vector<int> &row = rows[t / lpn];
                    const int p = 4 * (t % lpn + lpn * k) + b;
                    if (p < (int)row.size() && seen[row[p]] != g) {
                        seen[row[p]] = g;
                        memb[row[p]].push_back(g);
                    }
                }
            }
    }
    std::vector<int> avail(NB, 0);
    for (int e = 1; e <= H; ++e)
        ++avail[e % NB];
    std::vector<int> order(H);
    for (int h = 0; h < H; ++h)
        order[h] = h;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return memb[x].size() > memb[y].size();
    });
    std::vector<unsigned short> load((size_t)ngroups * NB, 0);
    std::vector<int> res(H);
    int score[NB];
    for (int h : order) {
        for (int r = 0; r < NB; ++r)
            score[r] = 0;
        for (int g : memb[h])
            for (int r = 0; r < NB; ++r)
                score[r] += load[(size_t)g * NB + r];
        int best = -1;
        for (int r = 0; r < NB; ++r)
            if (avail[r] > 0 && (best < 0 || score[r] < score[best]))
                best = r;
        res[h] = best;
        --avail[best];
        for (int g : memb[h])
            ++load[(size_t)g * NB + best];
    }
    int next[NB];
    for (int r = 0; r < NB; ++r)
        next[r] = r == 0 ? NB : r;
    std::vector<int> pos(H);
    for (int h = 0; h < H; ++h) {
        pos[h] = next[res[h]] - 1;
        next[res[h]] += NB;
    }
    return pos;
}

// Node order for tiling: Morton order of the initial coordinates on a grid
// whose cell holds about one node.
// Node order for tiling: by group (when given), then by Morton code of the
// coordinates on a grid of about one node per cell.
std::vector<int> morton_order(int n, const double *x, const int *group)
{
    double lo[3] = {x[0], x[1], x[2]}, hi[3] = {x[0], x[1], x[2]};
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], x[3 * (size_t)i + d]);
            hi[d] = std::max(hi[d], x[3 * (size_t)i + d]);
        }
    double vol = 1.0;
    int dims = 0;
    for (int d = 0; d < 3; ++d)
        if (hi[d] > lo[d]) {
            vol *= hi[d] - lo[d];
            ++dims;
        }
    double h = dims ? std::pow(vol / n, 1.0 / dims) : 1.0;
    if (!(h > 0.0))
        h = 1.0;
    std::vector<uint64_t> code(n);
    for (int i = 0; i < n; ++i) {
        uint64_t c = 0;
        for (int d = 0; d < 3; ++d) {
            double v = std::floor((x[3 * (size_t)i + d] - lo[d]) / h);
            v = std::min(std::max(v, 0.0), 2097151.0);
            c |= spread3((uint64_t)v) << d;
        }
        code[i] = c;
    }
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i)
        order[i] = i;
    if (group)
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            return group[a] != group[b] ? group[a] < group[b]
                                        : code[a] < code[b];
        });
    else
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return code[a] < code[b]; });
    return order;
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
    if (name && name_len > 0) {
        if (p.name[0])
            snprintf(name, name_len, "%s", p.name);
        else  // the marketing name is not always populated
            snprintf(name, name_len, "HIP device %d, %d CUs", device,
                     p.multiProcessorCount);
    }
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
    if (layout == PD_LAYOUT_COMPACT &&
        compact_row_bytes(max_neighbours) + 64u > LDS_BYTES)
        return fail(PD_E_LIMIT, "the compact layout steps at most 27285 "
                                "neighbours per node");
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
    int lds_max = 0;
    HIPCHK(hipDeviceGetAttribute(&lds_max,
                                 hipDeviceAttributeMaxSharedMemoryPerBlock,
                                 device));
    pd_sim *s = new pd_sim();
    s->device = device;
    s->integrator = integrator;
    s->layout = layout;
    s->dt = dt;
    s->damping = damping;
    s->n = nnodes;
    s->K = max_neighbours;
    s->lds_budget = (size_t)lds_max;
    if ((size_t)lds_max < LDS_BYTES)
        return fail(PD_E_NODEV, "device has less than 80 KiB of LDS per "
                                "workgroup");
    s->hmax = LDS_CAP_PMB - 1;
    const size_t N = (size_t)nnodes;
    s->h_coords.assign(coords, coords + 3 * N);

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
    s->h_vol.assign(volume, volume + N);
    s->h_flags = std::move(flags);
    s->h_bcv.assign(bc_values, bc_values + 3 * N);
    s->h_fbcv.assign(force_bc_values, force_bc_values + 3 * N);
    if (densities)
        s->h_dens.assign(densities, densities + 3 * N);
    chk(upload(s->family, family, N), "family");
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
    const int N = s->n, K = s->K;
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

    // bond rows: live entries (j, reference column), sorted ascending by j
    std::vector<int> cnt(N);
    std::vector<long long> rstart(N + 1, 0);
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
    for (int i = 0; i < N; ++i)
        rstart[i + 1] = rstart[i] + cnt[i];
    std::vector<int> rj(rstart[N]);
    std::vector<unsigned short> rp(rstart[N]);
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
            for (size_t k = 0; k < ent.size(); ++k) {
                rj[rstart[i] + k] = ent[k].first;
                rp[rstart[i] + k] = (unsigned short)ent[k].second;
            }
        }
    });

    // tiles: greedy over the Morton order, capped by tile_nodes and by the
    // LDS halo capacity
    // entry 0 of a staged halo is the zero record of dead slots
    // the compact layout's tiles leave room for the exact step's windows
    s->hmax = (nlin ? LDS_CAP_NLIN
                    : (s->layout == PD_LAYOUT_COMPACT ? EXACT_CAP
                                                      : LDS_CAP_PMB)) -
              1;
    const std::vector<int> order = morton_order(
        N, s->h_coords.data(), s->h_group.empty() ? nullptr : s->h_group.data());
    std::vector<int> stamp(N, -1), probe(N, -1);
    std::vector<TileInfo> tiles;
    std::vector<int> tnode;
    tnode.reserve(N);
    {
        int cur_nodes = 0, cur_halo = 0, tid = 0, nprobe = 0;
        // distinct nodes of i's neighbourhood not yet in tile t (a row may
        // list a neighbour more than once)
        auto count_new = [&](int i, int t) {
            ++nprobe;
            int c = 0;
            auto visit = [&](int j) {
                if (stamp[j] != t && probe[j] != nprobe) {
                    probe[j] = nprobe;
                    ++c;
                }
            };
            visit(i);
            for (long long q = rstart[i]; q < rstart[i + 1]; ++q)
                visit(rj[q]);
            return c;
        };
        for (int q = 0; q < N; ++q) {
            const int i = order[q];
            if (1 + cnt[i] > s->hmax)
                return fail(PD_E_LIMIT, "a node's neighbourhood exceeds the "
                                        "LDS halo capacity");
            int add = count_new(i, tid);
            const bool new_group =
                cur_nodes > 0 && !s->h_group.empty() &&
                s->h_group[i] != s->h_group[tnode.back()];
            if (cur_nodes > 0 && (cur_nodes == s->tile_nodes || new_group ||
                                  cur_halo + add > s->hmax)) {
                TileInfo t;
                t.node_off = (int)tnode.size() - cur_nodes;
                t.nodes = cur_nodes;
                t.halo = cur_halo;
                tiles.push_back(t);
                ++tid;
                cur_nodes = 0;
                cur_halo = 0;
                add = count_new(i, tid);
            }
            stamp[i] = tid;
            for (long long r = rstart[i]; r < rstart[i + 1]; ++r)
                stamp[rj[r]] = tid;
            cur_halo += add;
            ++cur_nodes;
            tnode.push_back(i);
        }
        if (cur_nodes > 0) {
            TileInfo t;
            t.node_off = (int)tnode.size() - cur_nodes;
            t.nodes = cur_nodes;
            t.halo = cur_halo;
            tiles.push_back(t);
        }
    }
    // within a tile, nodes by descending bond count: the nodes sharing a
    // wavefront then run loops of nearly equal length
    for (const auto &t : tiles)
        std::stable_sort(tnode.begin() + t.node_off,
                         tnode.begin() + t.node_off + t.nodes,
                         [&](int x, int y) { return cnt[x] > cnt[y]; });
    std::vector<int> inv(N);  // caller's node id -> internal id
    for (int q = 0; q < N; ++q)
        inv[tnode[q]] = q;
    const int ntiles = (int)tiles.size();
    long long halo_total = 0, chunk_total = 0;
    for (auto &t : tiles) {
        t.halo_off = (int)halo_total;
        halo_total += t.halo;
        int nch = 0;
        for (int l = 0; l < t.nodes; ++l)
            nch = std::max(nch, (cnt[tnode[t.node_off + l]] + 3) >> 2);
        t.nch = nch;
        t.pad = 0;
        t.chunk_off = chunk_total;
        // at least one row, which the kernel's clamped loads may read
        chunk_total += (long long)std::max(nch, 1) * t.nodes;
    }
    if (halo_total > 0x7fffffffLL)
        return fail(PD_E_LIMIT, "total halo size exceeds 2^31");
    const size_t S = (size_t)chunk_total * 4;
    std::vector<int> halo(halo_total);
    std::vector<unsigned short> loc(S, LOC_DEAD);
    std::vector<int> gidx(S, (int)DEAD);
    std::vector<unsigned short> pos(S, 0);
    std::vector<double> corr;
    std::vector<unsigned short> state;
    if (s->has_corr)
        corr.assign(S, 0.0);
    if (nlin)
        state.assign(S, 0);
    std::vector<long long> chunk0(N);
    std::vector<int> stride(N);
    const bool compact = s->layout == PD_LAYOUT_COMPACT;
    std::vector<int> nlow_h(compact ? N : 0), nup_h(compact ? N : 0);
    std::vector<unsigned> uord_h(compact ? S : 0, 0u);
    std::vector<unsigned short> rslot_h(compact ? (size_t)N * K : 0, 0);
    std::atomic<int> halo_mismatch{0};
    run_parallel(ntiles, [&](int lo, int hi) {
        std::vector<int> lmap(N, -1);  // internal id -> halo position
        std::vector<int> hl;
        std::vector<std::vector<int>> rows;
        for (int t = lo; t < hi; ++t) {
            const TileInfo &ti = tiles[t];
            hl.clear();
            for (int l = 0; l < ti.nodes; ++l) {
                const int i = tnode[ti.node_off + l];
                hl.push_back(ti.node_off + l);
                for (long long r = rstart[i]; r < rstart[i + 1]; ++r)
                    hl.push_back(inv[rj[r]]);
            }
            std::sort(hl.begin(), hl.end());
            hl.erase(std::unique(hl.begin(), hl.end()), hl.end());
            if ((int)hl.size() != ti.halo) {
                halo_mismatch = 1;
                continue;
            }
            for (size_t h = 0; h < hl.size(); ++h)
                lmap[hl[h]] = (int)h;
            rows.resize(ti.nodes);
            for (int l = 0; l < ti.nodes; ++l) {
                const int i = tnode[ti.node_off + l];
                rows[l].clear();
                for (int k = 0; k < cnt[i]; ++k)
                    rows[l].push_back(lmap[inv[rj[rstart[i] + k]]]);
            }
            const std::vector<int> hpos =
                spread_halo(rows, s->lpn, (int)hl.size());
            for (size_t h = 0; h < hl.size(); ++h) {
                lmap[hl[h]] = hpos[h];
                halo[ti.halo_off + hpos[h]] = hl[h];
            }
            for (int l = 0; l < ti.nodes; ++l) {
                const int i = tnode[ti.node_off + l];
                chunk0[ti.node_off + l] = ti.chunk_off + l;
                stride[ti.node_off + l] = ti.nodes;
                for (int k = 0; k < cnt[i]; ++k) {
                    const int j = rj[rstart[i] + k];
                    const int p = rp[rstart[i] + k];
                    const size_t sl = slot_of(ti.chunk_off + l, ti.nodes, k);
                    loc[sl] = (unsigned short)(lmap[inv[j]] + 1);
                    gidx[sl] = j;
                    pos[sl] = (unsigned short)p;
                    const size_t src = (size_t)i * K + p;
                    if (s->has_corr)
                        corr[sl] = s->h_corr[src];
                    if (nlin)
                        state[sl] = (unsigned short)(
                            (s->h_btypes[src] << 8) | (regimes ? regimes[src] : 0));
                }
                if (compact) {
                    // row position -> slot, then the exact step's lists:
                    // lower slots, and the upper neighbours in row order
                    for (int k = 0; k < cnt[i]; ++k)
                        rslot_h[(size_t)i * K + rp[rstart[i] + k]] =
                            (unsigned short)k;
                    const int p = ti.node_off + l;
                    const int *rb = rj.data() + rstart[i];
                    nlow_h[p] = (int)(std::lower_bound(rb, rb + cnt[i], i) - rb);
                    const int *row = nlist + (size_t)i * K;
                    int c = 0;
                    for (int q = 0; q < cnt[i]; ++q) {
                        const int j = row[q];
                        if (j <= i)
                            continue;
                        const int k =
                            (int)(std::lower_bound(rb, rb + cnt[i], j) - rb);
                        uord_h[slot_of(ti.chunk_off + l, ti.nodes, c++)] =
                            ((unsigned)k << 16) |
                            loc[slot_of(ti.chunk_off + l, ti.nodes, k)];
                    }
                    nup_h[p] = c;
                }
            }
            for (int h : hl)
                lmap[h] = -1;
        }
    }, 16);
    if (halo_mismatch)
        return fail(PD_E_STATE, "internal: tile halo count mismatch");
    for (const auto &t : tiles)
        if (t.halo > s->hmax)
            return fail(PD_E_LIMIT, "internal: tile halo over capacity");

    HIPCHK(hipStreamSynchronize(s->stream));
    // from here the device state is replaced: until it is whole again no
    // step may run (a failure below leaves the simulation without buffers)
    s->has_buffers = false;
    adjoint_release(s);
    auto chk = [&](hipError_t e, const char *what) {
        if (e != hipSuccess && rc == PD_OK)
            rc = fail(PD_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    };
    s->ntiles = ntiles;
    s->hstride = 1;
    for (const auto &t : tiles)
        s->hstride = std::max(s->hstride, t.halo);
    chk(upload(s->tiles, tiles.data(), tiles.size()), "tiles");
    chk(upload(s->tnode, tnode.data(), tnode.size()), "tnode");
    chk(upload(s->halo, halo.data(), halo.size()), "halo");
    s->perm = tnode;
    chk(upload_perm(s->cnt, s->perm, cnt.data(), 1), "cnt");
    chk(upload_perm(s->r0, s->perm, s->h_coords.data(), 3), "coords");
    chk(upload_perm(s->vol, s->perm, s->h_vol.data(), 1), "volume");
    chk(upload_perm(s->flags, s->perm, s->h_flags.data(), 1), "flags");
    chk(upload_perm(s->bcv, s->perm, s->h_bcv.data(), 3), "bc_values");
    chk(upload_perm(s->fbcv, s->perm, s->h_fbcv.data(), 3),
        "force_bc_values");
    s->dens1.release();
    if (!s->h_dens.empty()) {
        chk(upload_perm(s->dens, s->perm, s->h_dens.data(), 3), "densities");
        // one density per node (the reference's tiled layout, model.py:
        // 691-702) is read as one value per node
        bool per_node = true;
        for (size_t i = 0; per_node && i < N; ++i)
            per_node = s->h_dens[3 * i] == s->h_dens[3 * i + 1] &&
                       s->h_dens[3 * i] == s->h_dens[3 * i + 2];
        if (per_node) {
            std::vector<double> d1(N);
            for (size_t i = 0; i < N; ++i)
                d1[i] = s->h_dens[3 * i];
            chk(upload_perm(s->dens1, s->perm, d1.data(), 1), "densities");
        }
    }
    if (s->has_cnode && !s->h_cnode.empty()) {
        chk(upload_perm(s->cnode, s->perm, s->h_cnode.data(), 1),
            "node stiffness");
    } else if (s->has_cnode) {
        // given on the device: into the new node order there
        chk(s->cnode.alloc(N), "node stiffness");
        if (rc == PD_OK)
            hipLaunchKernelGGL(k_perm, dim3((N + 255) / 256), dim3(256), 0,
                               s->stream, (int)N, 1, s->tnode.p,
                               s->cnode_caller.p, s->cnode.p);
        chk(hipGetLastError(), "node stiffness");
    }
    chk(upload(s->chunk0, chunk0.data(), (size_t)N), "chunk0");
    chk(upload(s->stride, stride.data(), (size_t)N), "stride");
    chk(upload(s->loc, loc.data(), loc.size()), "loc");
    chk(upload(s->gidx, gidx.data(), gidx.size()), "gidx");
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
        chk(upload(s->rslot, rslot_h.data(), rslot_h.size()), "nlist");
        chk(upload(s->n_neigh, n_neigh, (size_t)N), "n_neigh");
        chk(upload(s->nlow, nlow_h.data(), nlow_h.size()), "exact lists");
        chk(upload(s->nup, nup_h.data(), nup_h.size()), "exact lists");
        chk(upload(s->uord, uord_h.data(), uord_h.size()), "exact lists");
    } else {
        s->nlist_ref.release();
        s->rslot.release();
        s->n_neigh.release();
        s->nlow.release();
        s->nup.release();
        s->uord.release();
    }
    chk(s->out_nlist.alloc(s->layout == PD_LAYOUT_SLOT ? (size_t)N * K : 0),
        "nlist export");
    const size_t B = 3 * (size_t)N * sizeof(double);
    if (u)
        chk(upload_perm(s->u, s->perm, u, 3), "u");
    else
        chk(hipMemset(s->u.p, 0, B), "u");
    if (ud)
        chk(upload_perm(s->ud, s->perm, ud, 3), "ud");
    else
        chk(hipMemset(s->ud.p, 0, B), "ud");
    if (udd)
        chk(upload_perm(s->udd, s->perm, udd, 3), "udd");
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
    rc = upload_tips(s);
    if (rc)
        return rc;
    s->has_buffers = true;
    s->retile_pending = false;
    adjoint_release(s);  // slot arrays were rebuilt
    return PD_OK;
}

int pd_set_tips(pd_sim *s, int nentries, const int *node, const int *dof,
                const int *group, int ngroups)
{
    int rc = check_sim(s);
    if (rc)
        return rc;
    if (nentries < 0 || ngroups < 0 ||
        (nentries > 0 && (!node || !dof || !group)))
        return fail(PD_E_ARG, "node, dof and group (nentries) are required");
    std::vector<int> off(ngroups + 1, 0);
    for (int e = 0; e < nentries; ++e) {
        if (node[e] < 0 || node[e] >= s->n || dof[e] < 0 || dof[e] > 2 ||
            group[e] < 0 || group[e] >= ngroups)
            return fail(PD_E_ARG, "a tip entry is out of range");
        off[group[e] + 1]++;
    }
    for (int g = 0; g < ngroups; ++g)
        off[g + 1] += off[g];
    // group the entries, keeping their order inside a group
    std::vector<int> fill(off.begin(), off.end() - 1);
    s->h_tip_node.assign(nentries, 0);
    s->h_tip_dof.assign(nentries, 0);
    for (int e = 0; e < nentries; ++e) {
        const int at = fill[group[e]]++;
        s->h_tip_node[at] = node[e];
        s->h_tip_dof[at] = dof[e];
    }
    s->h_tip_off = off;
    s->tip_q.release();
    s->tip_off.release();
    if (s->has_buffers)
        return upload_tips(s);
    return PD_OK;
}

int pd_monitor(pd_sim *s, double *model, double *tips)
{
    int rc = check_sim(s);
    if (rc)
        return rc;
    if (!s->has_buffers)
        return fail(PD_E_STATE, "pd_create_buffers has not been called");
    if (!model)
        return fail(PD_E_ARG, "model (6 doubles) is required");
    const int ngroups =
        s->h_tip_off.empty() ? 0 : (int)s->h_tip_off.size() - 1;
    if (ngroups > 0 && !tips)
        return fail(PD_E_ARG, "tips (5 doubles per group) is required");
    const int N = s->n;
    if (!s->mon_part.p) {
        HIPCHK(s->mon_part.alloc((size_t)MON_BLOCKS * MON_SUMS));
    }
    const size_t nout = MON_SUMS + 5 * (size_t)ngroups;
    if (s->mon_out.n < nout) {
        HIPCHK(s->mon_out.alloc(nout));
    }
    hipLaunchKernelGGL(k_export, dim3((N + 255) / 256), dim3(256), 0,
                       s->stream, N, s->K, s->layout, 0, s->tnode.p,
                       s->cnt.p, s->chunk0.p, s->stride.p, s->gidx.p,
                       s->pos.p, s->family.p, s->n_neigh.p, s->out_nlist.p,
                       s->out_nneigh.p, s->damage.p);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_monitor_part, dim3(MON_BLOCKS), dim3(256), 0,
                       s->stream, N, s->u.p, s->ud.p, s->udd.p, s->force.p,
                       s->body_force.p, s->flags.p, s->vol.p, s->damage.p,
                       s->mon_part.p);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_monitor_final, dim3(1), dim3(256), 0, s->stream,
                       s->mon_part.p, ngroups, s->tip_off.p, s->tip_q.p,
                       s->u.p, s->ud.p, s->udd.p, s->force.p,
                       s->body_force.p, s->flags.p, s->vol.p, s->mon_out.p);
    HIPCHK(hipGetLastError());
    std::vector<double> h(nout);
    HIPCHK(hipMemcpyAsync(h.data(), s->mon_out.p, nout * sizeof(double),
                          hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    std::copy(h.begin(), h.begin() + MON_SUMS, model);
    if (ngroups > 0)
        std::copy(h.begin() + MON_SUMS, h.end(), tips);
    return PD_OK;
}

namespace {

// The device address of `scratch` for each device, looked up once per
// device with that device current.
constexpr int NOALLOC_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch[NOALLOC_DEVICES] = {};

int scratch_address(int device, double **out)
{
    if (device < 0 || device >= NOALLOC_DEVICES)
        return fail(PD_E_LIMIT, "pd_monitor_noalloc: device index beyond 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(PD_E_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                      hipGetErrorString(e));
        g_scratch[device] = static_cast<double *>(p);
    }
    *out = g_scratch[device];
    return PD_OK;
}

}  // namespace

int pd_monitor_noalloc(pd_sim *s, double *model, double *tips)
{
    int rc = check_sim(s);  // makes the simulation's device current
    if (rc)
        return rc;
    if (!s->has_buffers)
        return fail(PD_E_STATE, "pd_create_buffers has not been called");
    if (!model)
        return fail(PD_E_ARG, "model (6 doubles) is required");
    const int ngroups =
        s->h_tip_off.empty() ? 0 : (int)s->h_tip_off.size() - 1;
    if (ngroups > NOALLOC_MAX_GROUPS)
        return fail(PD_E_LIMIT, "pd_monitor_noalloc: more than 1024 tip "
                                "groups (use pd_monitor)");
    if (ngroups > 0 && !tips)
        return fail(PD_E_ARG, "tips (5 doubles per group) is required");
    double *sp = nullptr;
    rc = scratch_address(s->device, &sp);
    if (rc)
        return rc;
    // every entry of both regions is written by the kernels below before it
    // is read: nothing needs clearing between calls
    double *part = sp;
    double *staged = sp + (size_t)MON_BLOCKS * MON_SUMS;
    const int N = s->n;
    hipLaunchKernelGGL(k_export, dim3((N + 255) / 256), dim3(256), 0,
                       s->stream, N, s->K, s->layout, 0, s->tnode.p,
                       s->cnt.p, s->chunk0.p, s->stride.p, s->gidx.p,
                       s->pos.p, s->family.p, s->n_neigh.p, s->out_nlist.p,
                       s->out_nneigh.p, s->damage.p);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_monitor_part, dim3(MON_BLOCKS), dim3(256), 0,
                       s->stream, N, s->u.p, s->ud.p, s->udd.p, s->force.p,
                       s->body_force.p, s->flags.p, s->vol.p, s->damage.p,
                       part);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_monitor_final, dim3(1), dim3(256), 0, s->stream,
                       part, ngroups, s->tip_off.p, s->tip_q.p, s->u.p,
                       s->ud.p, s->udd.p, s->force.p, s->body_force.p,
                       s->flags.p, s->vol.p, staged);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(model, staged, MON_SUMS * sizeof(double),
                          hipMemcpyDeviceToHost, s->stream));
    if (ngroups > 0)
        HIPCHK(hipMemcpyAsync(tips, staged + MON_SUMS,
                              5 * (size_t)ngroups * sizeof(double),
                              hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return PD_OK;
}

int pd_set_node_stiffness(pd_sim *s, const double *c)
{
    int rc = check_sim(s);
    if (rc)
        return rc;
    HIPCHK(hipStreamSynchronize(s->stream));
    s->cnode_caller.release();
    if (!c) {
        s->cnode.release();
        s->h_cnode.clear();
        s->has_cnode = false;
    } else {
        s->h_cnode.assign(c, c + s->n);
        s->has_cnode = true;
        // uploaded in tile order, which pd_create_buffers fixes
        if (s->has_buffers)
            HIPCHK(upload_perm(s->cnode, s->perm, c, 1));
    }
    if (s->has_buffers) {
        // the records carry c: rewrite both from the device u
        rc = refresh_records(s);
        if (rc)
            return rc;
    }
    return PD_OK;
}

int pd_set_node_stiffness_device(pd_sim *s, const double *c)
{
    int rc = check_sim(s);
    if (rc)
        return rc;
    if (!c)
        return fail(PD_E_ARG, "c is NULL (pd_set_node_stiffness clears)");
    HIPCHK(hipStreamSynchronize(s->stream));
    const size_t N = (size_t)s->n;
    HIPCHK(s->cnode_caller.alloc(N));
    HIPCHK(hipMemcpyAsync(s->cnode_caller.p, c, N * sizeof(double),
                          hipMemcpyDeviceToDevice, s->stream));
    s->h_cnode.clear();
    s->has_cnode = true;
    if (s->has_buffers) {
        HIPCHK(s->cnode.alloc(N));
        hipLaunchKernelGGL(k_perm, dim3((s->n + 255) / 256), dim3(256), 0,
                           s->stream, s->n, 1, s->tnode.p, s->cnode_caller.p,
                           s->cnode.p);
        HIPCHK(hipGetLastError());
        return refresh_records(s);  // synchronizes the stream
    }
    HIPCHK(hipStreamSynchronize(s->stream));
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

int pd_set_exact_order(pd_sim *s, int on)
{
    int rc = check_sim(s);
    if (rc)
        return rc;
    if (on && (s->layout != PD_LAYOUT_COMPACT || s->integrator != PD_EULER))
        return fail(PD_E_ARG, "the exact-order step is the Cython path's: "
                              "PD_LAYOUT_COMPACT and PD_EULER only");
    if (on && !exact_rows_fit(s->K))
        return fail(PD_E_LIMIT, "the exact-order step takes at most 4074 "
                                "neighbours per node");
    const bool was = s->exact;
    s->exact = on != 0;
    // the records carry V (exact) or 4kV (default): rewrite them
    if (s->has_buffers && was != s->exact)
        return refresh_records(s);
    return PD_OK;
}

int pd_set_launch(pd_sim *s, int lanes_per_node, int tile_nodes)
{
    int rc = check_sim(s);
    if (rc)
        return rc;
    if (lanes_per_node != 1 && lanes_per_node != 2 && lanes_per_node != 4 &&
        lanes_per_node != 8)
        return fail(PD_E_ARG, "lanes_per_node must be 1, 2, 4 or 8");
    if (tile_nodes < 64 || tile_nodes % 64 ||
        tile_nodes * lanes_per_node > 1024)
        return fail(PD_E_ARG, "tile_nodes must be a multiple of 64 with "
                              "tile_nodes * lanes_per_node <= 1024");
    // a new tile size takes effect when pd_create_buffers re-tiles; until
    // then the current tiles (of the old size) cannot be stepped
    if (s->has_buffers && tile_nodes != s->tile_nodes)
        s->retile_pending = true;
    s->lpn = lanes_per_node;
    s->tile_nodes = tile_nodes;
    return PD_OK;
}

int pd_set_groups(pd_sim *s, const int *group)
{
    int rc = check_sim(s);
    if (rc)
        return rc;
    if (s->has_buffers)
        return fail(PD_E_STATE, "groups must be set before pd_create_buffers");
    if (!group) {
        s->h_group.clear();
        return PD_OK;
    }
    s->h_group.assign(group, group + s->n);
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

int pd_run_graph(pd_sim *s, int nsteps, const double *disp_mags,
                 const double *force_mags, float *elapsed_ms)
{
    int rc = check_sim(s);
    if (rc)
        return rc;
    if (!s->has_buffers)
        return fail(PD_E_STATE, "pd_create_buffers has not been called");
    if (nsteps < 0)
        return fail(PD_E_ARG, "nsteps < 0");
    if (nsteps == 0) {
        if (elapsed_ms)
            *elapsed_ms = 0.0f;
        return PD_OK;
    }
    if (s->adj_recording)
        return fail(PD_E_STATE, "an adjoint recording runs step by step "
                                "(pd_run), not as a graph");
    HIPCHK(hipStreamSynchronize(s->stream));
    const int cur0 = s->cur;
    HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    for (int t = 0; t < nsteps && rc == PD_OK; ++t)
        rc = launch_step(s, disp_mags ? disp_mags[t] : 0.0,
                         force_mags ? force_mags[t] : 0.0);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(s->stream, &graph);
    if (rc != PD_OK || ec != hipSuccess) {
        if (graph)
            (void)hipGraphDestroy(graph);
        s->cur = cur0;
        return rc != PD_OK ? rc
                           : fail(PD_E_HIP, std::string("graph capture: ") +
                                                hipGetErrorString(ec));
    }
    hipGraphExec_t exec = nullptr;
    hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e == hipSuccess && elapsed_ms)
        e = hipEventRecord(s->ev0, s->stream);
    if (e == hipSuccess)
        e = hipGraphLaunch(exec, s->stream);
    if (e == hipSuccess && elapsed_ms)
        e = hipEventRecord(s->ev1, s->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s->stream);
    if (e == hipSuccess && elapsed_ms)
        e = hipEventElapsedTime(elapsed_ms, s->ev0, s->ev1);
    if (exec)
        (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        s->cur = cur0;
        return fail(PD_E_HIP, std::string("graph run: ") +
                                  hipGetErrorString(e));
    }
    return PD_OK;
}

int pd_adjoint_start(pd_sim *s, int nsteps, int checkpoint_every)
{
    int rc = check_sim(s);
    if (rc)
        return rc;
    if (!s->has_buffers || s->retile_pending)
        return fail(PD_E_STATE, "pd_create_buffers has not been called");
    if (s->integrator != PD_EULER || s->layout != PD_LAYOUT_SLOT ||
        !s->has_cnode || s->nregimes > 1 || s->nbond_types > 1)
        return fail(PD_E_ARG, "the adjoint needs the Euler integrator, the "
                              "slot layout and a node stiffness field");
    if (nsteps < 1 || checkpoint_every < 1)
        return fail(PD_E_ARG, "nsteps and checkpoint_every must be >= 1");
    adjoint_release(s);
    const size_t S = s->loc.n, N = (size_t)s->n, B = 3 * N;
    const int nck = (nsteps + checkpoint_every - 1) / checkpoint_every;
    std::vector<int> inv(N);
    for (size_t q = 0; q < N; ++q)
        inv[s->perm[q]] = (int)q;
    DevBuf<int> inv_d;
    HIPCHK(upload(inv_d, inv.data(), N));
    HIPCHK(s->brk.alloc(S));
    HIPCHK(s->il0.alloc(S));
    HIPCHK(s->arec[0].alloc(N));
    HIPCHK(s->arec[1].alloc(N));
    HIPCHK(s->loc0.alloc(S));
    HIPCHK(s->ckpt.alloc((size_t)nck * B));
    HIPCHK(s->seg.alloc((size_t)checkpoint_every * B));
    HIPCHK(s->lam[0].alloc(B));
    HIPCHK(s->lam[1].alloc(B));
    HIPCHK(s->gradc.alloc(N));
    HIPCHK(s->fin_u.alloc(B));
    HIPCHK(s->fin_force.alloc(B));
    HIPCHK(s->fin_bforce.alloc(B));
    HIPCHK(hipMemcpyAsync(s->loc0.p, s->loc.p, S * sizeof(unsigned short),
                          hipMemcpyDeviceToDevice, s->stream));
    // slots past a node's count: brk -1 (all bytes set), 1/l0 zero
    HIPCHK(hipMemsetAsync(s->brk.p, 0xff, S * sizeof(int), s->stream));
    HIPCHK(hipMemsetAsync(s->il0.p, 0, S * sizeof(double), s->stream));
    const int blocks = (s->n + 255) / 256;
    hipLaunchKernelGGL(k_adj_init, dim3(blocks), dim3(256), 0, s->stream,
                       s->n, s->cnt.p, s->chunk0.p, s->stride.p, s->loc.p,
                       s->brk.p);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_adj_il0, dim3(blocks), dim3(256), 0, s->stream,
                       s->n, s->cnt.p, s->chunk0.p, s->stride.p, s->gidx.p,
                       inv_d.p, s->r0.p, s->il0.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(s->ckpt.p, s->u.p, B * sizeof(double),
                          hipMemcpyDeviceToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    s->adj_nsteps = nsteps;
    s->adj_every = checkpoint_every;
    s->step_counter = 0;
    s->adj_recording = true;
    return PD_OK;
}

namespace {

// Put the state at the start of step n0 from u (already in s->u): bond
// states from brk, records from u.
int adjoint_restore(pd_sim *s, int n0)
{
    const int blocks = (s->n + 255) / 256;
    hipLaunchKernelGGL(k_adj_restore, dim3(blocks), dim3(256), 0, s->stream,
                       s->n, n0, s->cnt.p, s->chunk0.p, s->stride.p,
                       s->brk.p, s->loc0.p, s->loc.p, s->gidx.p);
    HIPCHK(hipGetLastError());
    return refresh_records(s);
}

}  // namespace

namespace {

// pd_adjoint_backward with dL_du and dL_dc in host memory (device false) or
// in device memory (device true), both in the caller's node order.
int adjoint_backward(pd_sim *s, const double *dL_du, double *dL_dc,
                     bool device)
{
    int rc = check_sim(s);
    if (rc)
        return rc;
    if (!s->adj_ready)
        return fail(PD_E_STATE, "no finished adjoint recording: call "
                                "pd_adjoint_start and run its steps");
    if (!dL_du || !dL_dc)
        return fail(PD_E_ARG, "dL_du and dL_dc are required");
    const size_t N = (size_t)s->n, B = 3 * N, BB = B * sizeof(double);
    const int every = s->adj_every, nsteps = s->adj_nsteps;
    const int nck = (nsteps + every - 1) / every;
    hipStream_t st = s->stream;
    HIPCHK(hipStreamSynchronize(st));
    HIPCHK(hipMemcpyAsync(s->fin_u.p, s->u.p, BB, hipMemcpyDeviceToDevice,
                          st));
    HIPCHK(hipMemcpyAsync(s->fin_force.p, s->force.p, BB,
                          hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(s->fin_bforce.p, s->body_force.p, BB,
                          hipMemcpyDeviceToDevice, st));
    if (device) {
        hipLaunchKernelGGL(k_perm, dim3((s->n + 255) / 256), dim3(256), 0, st,
                           s->n, 3, s->tnode.p, dL_du, s->lam[0].p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemsetAsync(s->gradc.p, 0, N * sizeof(double), st));
    } else {
        const std::vector<double> v = to_internal(s->perm, dL_du, 3);
        HIPCHK(hipMemcpyAsync(s->lam[0].p, v.data(), BB,
                              hipMemcpyHostToDevice, st));
        HIPCHK(hipMemsetAsync(s->gradc.p, 0, N * sizeof(double), st));
        HIPCHK(hipStreamSynchronize(st));
    }
    int cur = 0;
    const int pblocks = (int)((N + 255) / 256);
    AdjArgs aa{};
    aa.ntiles = s->ntiles;
    aa.tiles_per_xcd = (s->ntiles + XCDS - 1) / XCDS;
    aa.dt = s->dt;
    aa.tiles = s->tiles.p;
    aa.halo = s->halo.p;
    aa.cnt = s->cnt.p;
    aa.brk = s->brk.p;
    aa.r0 = s->r0.p;
    aa.flags = s->flags.p;
    aa.loc0 = s->loc0.p;
    aa.il0 = s->il0.p;
    aa.gradc = s->gradc.p;
    struct ReplayGuard {  // steps below are replays, also on early returns
        pd_sim *s;
        ~ReplayGuard()
        {
            s->adj_replay = false;
            s->u_in = nullptr;
            s->u_out = nullptr;
        }
    } guard{s};
    s->adj_replay = true;
    for (int k = nck - 1; k >= 0 && rc == PD_OK; --k) {
        const int n0 = k * every, n1 = std::min(n0 + every, nsteps);
        HIPCHK(hipMemcpyAsync(s->u.p, s->ckpt.p + (size_t)k * B, BB,
                              hipMemcpyDeviceToDevice, st));
        rc = adjoint_restore(s, n0);
        if (rc)
            break;
        HIPCHK(hipMemcpyAsync(s->seg.p, s->u.p, BB, hipMemcpyDeviceToDevice,
                              st));
        // replay the segment: the same kernel from the same state reproduces
        // the recorded steps (and their breaks) bitwise; step n reads u_n
        // from slot n - n0 of the segment and writes u_{n+1} to the next
        for (int n = n0; n < n1 - 1 && rc == PD_OK; ++n) {
            s->step_counter = n;
            s->u_in = s->seg.p + (size_t)(n - n0) * B;
            s->u_out = s->seg.p + (size_t)(n - n0 + 1) * B;
            rc = launch_step(s, s->adj_dmag[n], s->adj_fmag[n]);
        }
        s->u_in = nullptr;
        s->u_out = nullptr;
        if (rc)
            break;
        // the records of the segment's last step; each backward step then
        // writes those of the step before it, within the segment
        int ra = 0;
        hipLaunchKernelGGL(k_adj_pack, dim3(pblocks), dim3(256), 0, st, s->n,
                           s->r0.p, s->seg.p + (size_t)(n1 - 1 - n0) * B,
                           s->vol.p, s->cnode.p, s->flags.p, s->lam[cur].p,
                           s->arec[ra].p);
        HIPCHK(hipGetLastError());
        for (int n = n1 - 1; n >= n0; --n) {
            aa.step = n;
            aa.rec = s->arec[ra].p;
            aa.lam_out = s->lam[1 - cur].p;
            aa.rec_next = n > n0 ? s->arec[1 - ra].p : nullptr;
            aa.u_prev = n > n0 ? s->seg.p + (size_t)(n - 1 - n0) * B
                               : nullptr;
            if (s->ntiles > 0)
                HIPCHK(launch_adjoint(s, aa));
            cur = 1 - cur;
            ra = 1 - ra;
        }
    }
    s->adj_replay = false;
    if (rc)
        return rc;
    // back to the end of the recording
    HIPCHK(hipMemcpyAsync(s->u.p, s->fin_u.p, BB, hipMemcpyDeviceToDevice,
                          st));
    HIPCHK(hipMemcpyAsync(s->force.p, s->fin_force.p, BB,
                          hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(s->body_force.p, s->fin_bforce.p, BB,
                          hipMemcpyDeviceToDevice, st));
    rc = adjoint_restore(s, nsteps);
    if (rc)
        return rc;
    s->step_counter = nsteps;
    if (device) {
        hipLaunchKernelGGL(k_unperm, dim3((s->n + 255) / 256), dim3(256), 0,
                           st, s->n, 1, s->tnode.p, s->gradc.p, dL_dc);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
        return PD_OK;
    }
    return download_caller(s, dL_dc, s->gradc.p, 1);
}

}  // namespace

int pd_adjoint_backward(pd_sim *s, const double *dL_du, double *dL_dc)
{
    return adjoint_backward(s, dL_du, dL_dc, false);
}

int pd_adjoint_backward_device(pd_sim *s, const double *dL_du, double *dL_dc)
{
    return adjoint_backward(s, dL_du, dL_dc, true);
}

int pd_write_device(pd_sim *s, double *u, double *damage)
{
    int rc = check_sim(s);
    if (rc)
        return rc;
    if (!s->has_buffers)
        return fail(PD_E_STATE, "pd_create_buffers has not been called");
    const int N = s->n;
    const int blocks = (N + 255) / 256;
    if (damage) {
        hipLaunchKernelGGL(k_export, dim3(blocks), dim3(256), 0, s->stream,
                           N, s->K, s->layout, 0, s->tnode.p, s->cnt.p,
                           s->chunk0.p, s->stride.p, s->gidx.p, s->pos.p,
                           s->family.p, s->n_neigh.p, s->out_nlist.p,
                           s->out_nneigh.p, s->damage.p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(damage, s->damage.p, (size_t)N * sizeof(double),
                              hipMemcpyDeviceToDevice, s->stream));
    }
    if (u) {
        hipLaunchKernelGGL(k_unperm, dim3(blocks), dim3(256), 0, s->stream, N,
                           3, s->tnode.p, s->u.p, u);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    return PD_OK;
}

int pd_adjoint_stop(pd_sim *s)
{
    int rc = check_sim(s);
    if (rc)
        return rc;
    HIPCHK(hipStreamSynchronize(s->stream));
    adjoint_release(s);
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
                       s->K, s->layout, nlist ? 1 : 0, s->tnode.p, s->cnt.p,
                       s->chunk0.p,
                       s->stride.p, s->gidx.p, s->pos.p, s->family.p,
                       s->n_neigh.p, s->out_nlist.p, s->out_nneigh.p,
                       s->damage.p);
    HIPCHK(hipGetLastError());
    if (nlist && s->layout == PD_LAYOUT_COMPACT) {
        hipLaunchKernelGGL(k_export_rows, dim3(blocks), dim3(256), 0,
                           s->stream, N, s->K, s->tnode.p, s->cnt.p,
                           s->chunk0.p, s->stride.p, s->gidx.p, s->rslot.p,
                           s->nlist_ref.p);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    const struct {
        double *dst;
        const double *src;
    } fields[] = {{u, s->u.p}, {ud, s->ud.p}, {udd, s->udd.p},
                  {force, s->force.p}, {body_force, s->body_force.p}};
    for (const auto &f : fields) {
        if (f.dst) {
            rc = download_caller(s, f.dst, f.src, 3);
            if (rc)
                return rc;
        }
    }
    if (body_force) {
        // stored only on force-BC DOFs; equal to force everywhere else
        std::vector<double> tmp;
        const double *f = force;
        if (!f) {
            tmp.resize(3 * (size_t)N);
            rc = download_caller(s, tmp.data(), s->force.p, 3);
            if (rc)
                return rc;
            f = tmp.data();
        }
        for (size_t q = 0; q < 3 * (size_t)N; ++q)
            if (!((s->h_flags[q / 3] >> (3 + (int)(q % 3))) & 1))
                body_force[q] = f[q];
    }
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
                       s->stream, s->n, s->K, s->tnode.p, s->cnt.p,
                       s->chunk0.p, s->stride.p, s->pos.p, s->state.p, out.p);
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
                       s->cnt.p, s->chunk0.p, s->stride.p, s->gidx.p,
                       s->counter.p);
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

int pd_tile_stats(pd_sim *s, int *ntiles, int *hmax, int *hstride,
                  long long *halo_total, long long *slots)
{
    int rc = check_sim(s);
    if (rc)
        return rc;
    if (ntiles)
        *ntiles = s->ntiles;
    if (hmax)
        *hmax = s->hmax;
    if (hstride)
        *hstride = s->hstride;
    if (halo_total)
        *halo_total = (long long)s->halo.n;
    if (slots)
        *slots = (long long)s->loc.n;
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
                  int flags, double *l, double *l0, double *s,
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
                       n, (flags & PD_CHECK_OPENCL) ? 1 : 0, d_ri.p, d_rj.p,
                       d_r0i.p, d_r0j.p, s0,
                       (flags & PD_CHECK_TENSION) ? 0xffffffffu : 0x7fffffffu,
                       d_l.p,
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
