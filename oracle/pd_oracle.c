/*
 * Oracle: a CPU restatement of PeriPy's time-stepping kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP path: it is
 * compiled, loaded and called only by tests/, by __graft_entry__.smoke() and by
 * the cpu_baseline leg of bench.py.  The product (probabilistic_peridynamics_amd)
 * never links, loads or calls it.
 *
 * Build: gcc -O3 -ffp-contract=off -fPIC -shared (no -march), matching the
 * reference's own compile flags (setup.py:12: extra_compile_args=['-O3'], plain
 * x86-64, so no FMA instructions are available to the compiler) and therefore
 * the reference's IEEE-double rounding, operation for operation.
 *
 * Two families of functions live here.
 *
 *  (1) Exact restatements of the Cython oracle (peripy/peridynamics.pyx,
 *      peripy/spatial.pyx).  Every floating-point operation is written in the
 *      order of the Cython source, so results are bit-identical to it.
 *
 *  (2) The semantics the reference only has in OpenCL (per-slot "-1" bond
 *      breaking, stiffness corrections, the n-linear / composite regime state
 *      machine, the Euler-Cromer and Velocity-Verlet updates), restated from
 *      the peripy/cl kernels, but with the stretch computed by the Cython arithmetic
 *      of (1) (r = r0 + u per node, then spatial.pyx's ceuclid/cstrain).  This
 *      is the definition the HIP library implements for those integrators.
 *
 * Pinning (see tests/test_oracle_*.py): the reference's own known-answer
 * tests (test_peridynamics.py, test_spatial.py, test_cl_euler*.py,
 * test_cl_velocity_verlet.py, test_cl_peridynamics.py), its 10-step
 * regression fixtures (expected_displacements/force/damage/connectivity), and
 * closed forms / invariants / brute force for what those leave open.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Threads of the per-node loops below (slot bond force, updates,
 * positions).  Every node's arithmetic is the same at any thread count;
 * the Cython break_bonds / bond_force scatter stays single-threaded. */
void oracle_set_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

int oracle_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---------------------------------------------------------------------- */
/* (1) spatial.pyx                                                         */
/* ---------------------------------------------------------------------- */

/* spatial.pyx:20-31 (ceuclid): dr_i = r2_i - r1_i; dr_i = dr_i*dr_i;
 * return sqrt(dr_0 + dr_1 + dr_2)  -- C evaluates the sum left to right. */
double oracle_euclid(const double *r1, const double *r2)
{
    double d0 = r2[0] - r1[0];
    double d1 = r2[1] - r1[1];
    double d2 = r2[2] - r1[2];
    d0 = d0 * d0;
    d1 = d1 * d1;
    d2 = d2 * d2;
    return sqrt((d0 + d1) + d2);
}

/* spatial.pyx:54-66 (cstrain): l = |r2-r1|, l0 = |r20-r10|, (l - l0)/l0 */
double oracle_strain(const double *r1, const double *r2,
                     const double *r10, const double *r20)
{
    double l = oracle_euclid(r1, r2);
    double l0 = oracle_euclid(r10, r20);
    double dl = l - l0;
    return dl / l0;
}

/* spatial.pyx:87-97 (cstrain2): given l, (l - |r20-r10|)/|r20-r10| */
double oracle_strain2(double l, const double *r10, const double *r20)
{
    double l0 = oracle_euclid(r10, r20);
    double dl = l - l0;
    return dl / l0;
}

/* ---------------------------------------------------------------------- */
/* (1) peridynamics.pyx                                                    */
/* ---------------------------------------------------------------------- */

/* peridynamics.pyx:5-31 (damage): ifamily = family[i] (as double);
 * (ifamily - n_neigh[i]) / ifamily */
void oracle_damage(int nnodes, const int *n_neigh, const int *family,
                   double *out)
{
    for (int i = 0; i < nnodes; ++i) {
        double ifamily = (double)family[i];
        out[i] = (ifamily - (double)n_neigh[i]) / ifamily;
    }
}

/* peridynamics.pyx:34-99 (bond_force).  Node i, for each of its current
 * neighbours j > i (in nlist order): fn = strain*k; fn = fn/l;
 * f_d = fn*(r_j,d - r_i,d); F_i += f_d*V_j; F_j -= f_d*V_i.  After node i's
 * own loop its force boundary condition is added (peridynamics.pyx:93-97).
 * `force` is overwritten (the Cython function returns a new zero array). */
void oracle_bond_force(int nnodes, int max_neigh, const double *r,
                       const double *r0, const int *nlist, const int *n_neigh,
                       const double *volume, double bond_stiffness,
                       const double *force_bc_values,
                       const int *force_bc_types, double force_bc_scale,
                       double *force)
{
    for (int i = 0; i < 3 * nnodes; ++i)
        force[i] = 0.0;
    for (int i = 0; i < nnodes; ++i) {
        int i_n_neigh = n_neigh[i];
        for (int neigh = 0; neigh < i_n_neigh; ++neigh) {
            int j = nlist[(size_t)i * max_neigh + neigh];
            if (i < j) {
                double l = oracle_euclid(&r[3 * (size_t)i], &r[3 * (size_t)j]);
                double strain = oracle_strain2(l, &r0[3 * (size_t)i],
                                               &r0[3 * (size_t)j]);
                double force_norm = strain * bond_stiffness;
                force_norm = force_norm / l;
                double f[3];
                for (int d = 0; d < 3; ++d)
                    f[d] = force_norm * (r[3 * (size_t)j + d] - r[3 * (size_t)i + d]);
                for (int d = 0; d < 3; ++d) {
                    force[3 * (size_t)i + d] = force[3 * (size_t)i + d] + f[d] * volume[j];
                    force[3 * (size_t)j + d] = force[3 * (size_t)j + d] - f[d] * volume[i];
                }
            }
        }
        for (int d = 0; d < 3; ++d) {
            if (force_bc_types[3 * (size_t)i + d] != 0)
                force[3 * (size_t)i + d] = force[3 * (size_t)i + d]
                    + force_bc_scale * force_bc_values[3 * (size_t)i + d];
        }
    }
}

/* peridynamics.pyx:102-156 (break_bonds).  Sequential over i; a bond j > i
 * breaks when !(|strain| < s0): it is replaced by the last live entry of i's
 * row (index not advanced), and i is removed from j's row by the same
 * swap-with-last.  Stale values beyond n_neigh are left untouched. */
void oracle_break_bonds(int nnodes, int max_neigh, const double *r,
                        const double *r0, int *nlist, int *n_neigh,
                        double critical_strain)
{
    for (int i = 0; i < nnodes; ++i) {
        int *row = &nlist[(size_t)i * max_neigh];
        int i_n_neigh = n_neigh[i];
        int neigh = 0;
        while (neigh < i_n_neigh) {
            int j = row[neigh];
            if (i < j) {
                double s = oracle_strain(&r[3 * (size_t)i], &r[3 * (size_t)j],
                                         &r0[3 * (size_t)i], &r0[3 * (size_t)j]);
                if (fabs(s) < critical_strain) {
                    neigh += 1;
                } else {
                    row[neigh] = row[i_n_neigh - 1];
                    i_n_neigh -= 1;
                    int *jrow = &nlist[(size_t)j * max_neigh];
                    int j_n_neigh = n_neigh[j];
                    for (int jneigh = 0; jneigh < j_n_neigh; ++jneigh) {
                        if (jrow[jneigh] == i) {
                            jrow[jneigh] = jrow[j_n_neigh - 1];
                            n_neigh[j] = n_neigh[j] - 1;
                            break;
                        }
                    }
                }
            } else {
                neigh += 1;
            }
        }
        n_neigh[i] = i_n_neigh;
    }
}

/* peridynamics.pyx:159-185 (update_displacement, Euler):
 * u = bc_types==0 ? u + dt*F : bc_scale*bc_values  (per DOF) */
void oracle_update_displacement(int nnodes, double *u,
                                const double *bc_values, const int *bc_types,
                                const double *force, double bc_scale,
                                double dt)
{
#pragma omp parallel for schedule(static) if (nnodes > 4096)
    for (size_t i = 0; i < 3 * (size_t)nnodes; ++i) {
        if (bc_types[i] == 0)
            u[i] = u[i] + dt * force[i];
        else
            u[i] = bc_scale * bc_values[i];
    }
}

/* create_crack.pyx:1-43: for each pair (i, j) remove j from i's row and i
 * from j's row: swap with the last live entry, write -1 into the vacated last
 * slot, decrement n_neigh. */
void oracle_create_crack(int ncrack, const int *crack, int max_neigh,
                         int *nlist, int *n_neigh)
{
    for (int c = 0; c < ncrack; ++c) {
        int i = crack[2 * c], j = crack[2 * c + 1];
        int pair[2][2] = {{i, j}, {j, i}};
        for (int side = 0; side < 2; ++side) {
            int a = pair[side][0], b = pair[side][1];
            int *row = &nlist[(size_t)a * max_neigh];
            for (int neigh = 0; neigh < n_neigh[a]; ++neigh) {
                if (row[neigh] == b) {
                    row[neigh] = row[n_neigh[a] - 1];
                    row[n_neigh[a] - 1] = -1;
                    n_neigh[a] = n_neigh[a] - 1;
                    break;
                }
            }
        }
    }
}

/* ---------------------------------------------------------------------- */
/* (2) OpenCL-only semantics restated with the Cython stretch arithmetic  */
/* ---------------------------------------------------------------------- */

/* Per-slot bond force with in-place breaking, the structure of
 * cl/peridynamics.cl bond_force1..4 (:4-131, :134-262, :265-414, :417-565):
 * every node i visits every slot p of its row; nlist[i][p] == -1 is a broken
 * (or padding) slot.  Differences from the OpenCL text, by design:
 *   - the stretch is spatial.pyx's: l = |r_j - r_i| with r = r0 + u formed per
 *     node by the caller, l0 = |r0_j - r0_i|, s = (l - l0)/l0;
 *   - the force is written in peridynamics.pyx:77-89's order:
 *     fn = s*k [*corr]; fn = fn/l; f_d = fn*(r_j,d - r_i,d); F_i,d += f_d*V_j
 *     (n-linear: fn = s*k[t,r] + plus_cs[t,r] [*corr] before the division);
 *   - PMB breaking uses peridynamics.pyx:133's !(|s| < s0) unless
 *     tension_only, which gives cl/peridynamics.cl:78's !(s < s0);
 *   - with a per-node stiffness field cnode, k_ij = (c_i + c_j)*0.5.
 * The n-linear state machine is cl/peridynamics.cl:326-367 verbatim in
 * meaning (at most one regime change per step, break when regime >= nregimes),
 * except that the table is only read for a live slot.
 * Tables are (nbond_types, nregimes) flattened t*nregimes + r
 * (model.py:1042-1070).  body_force = bond sum; force = bond sum + BC term
 * (cl/peridynamics.cl:117-129). */
void oracle_slot_bond_force(
    int nnodes, int max_neigh, const double *r, const double *r0,
    int *nlist, const double *volume, double bond_stiffness,
    double critical_stretch, int tension_only, const double *cnode,
    const double *corr, const int *bond_types, int *regimes,
    const double *k_tab, const double *s_tab, const double *pc_tab,
    int nregimes, const int *force_bc_types, const double *force_bc_values,
    double force_bc_scale, double *force, double *body_force)
{
#pragma omp parallel for schedule(static) if (nnodes > 4096)
    for (int i = 0; i < nnodes; ++i) {
        double acc[3] = {0.0, 0.0, 0.0};
        for (int p = 0; p < max_neigh; ++p) {
            size_t slot = (size_t)i * max_neigh + p;
            int j = nlist[slot];
            if (j == -1)
                continue;
            double l = oracle_euclid(&r[3 * (size_t)i], &r[3 * (size_t)j]);
            double s = oracle_strain2(l, &r0[3 * (size_t)i], &r0[3 * (size_t)j]);
            double fn;
            if (k_tab == NULL) {
                int broken = tension_only ? !(s < critical_stretch)
                                          : !(fabs(s) < critical_stretch);
                if (broken) {
                    nlist[slot] = -1;
                    continue;
                }
                double k = bond_stiffness;
                if (cnode != NULL)
                    k = (cnode[i] + cnode[j]) * 0.5;
                fn = s * k;
            } else {
                int t = bond_types[slot];
                int reg = regimes[slot];
                if (s < s_tab[t * nregimes + reg]) {
                    if (reg > 0 && s < s_tab[t * nregimes + reg - 1]) {
                        reg -= 1;
                        regimes[slot] = reg;
                    }
                } else {
                    reg += 1;
                    regimes[slot] = reg;
                }
                if (reg >= nregimes) {
                    nlist[slot] = -1;
                    continue;
                }
                fn = s * k_tab[t * nregimes + reg] + pc_tab[t * nregimes + reg];
            }
            if (corr != NULL)
                fn = fn * corr[slot];
            fn = fn / l;
            for (int d = 0; d < 3; ++d) {
                double f = fn * (r[3 * (size_t)j + d] - r[3 * (size_t)i + d]);
                acc[d] = acc[d] + f * volume[j];
            }
        }
        for (int d = 0; d < 3; ++d) {
            size_t q = 3 * (size_t)i + d;
            body_force[q] = acc[d];
            force[q] = force_bc_types[q] == 0
                           ? acc[d]
                           : acc[d] + force_bc_scale * force_bc_values[q];
        }
    }
}

/* Count of live (!= -1) slots per row and the damage from it, in
 * peridynamics.pyx:29's form (family - n)/family (cl/peridynamics.cl:568-610
 * counts the same slots). */
void oracle_slot_damage(int nnodes, int max_neigh, const int *nlist,
                        const int *family, int *n_neigh, double *damage)
{
    for (int i = 0; i < nnodes; ++i) {
        int n = 0;
        for (int p = 0; p < max_neigh; ++p)
            n += nlist[(size_t)i * max_neigh + p] != -1;
        n_neigh[i] = n;
        double ifamily = (double)family[i];
        damage[i] = (ifamily - (double)n) / ifamily;
    }
}

/* cl/euler_cromer.cl:29-34, per DOF q:
 * udd = (F - eta*ud)/rho; ud += udd*dt; u = bc ? scale*val : u + dt*ud */
void oracle_update_euler_cromer(int nnodes, const double *force, double *u,
                                double *ud, double *udd, const int *bc_types,
                                const double *bc_values,
                                const double *densities, double bc_scale,
                                double damping, double dt)
{
#pragma omp parallel for schedule(static) if (nnodes > 4096)
    for (size_t q = 0; q < 3 * (size_t)nnodes; ++q) {
        double uddi = (force[q] - damping * ud[q]) / densities[q];
        udd[q] = uddi;
        ud[q] = ud[q] + uddi * dt;
        u[q] = bc_types[q] == 0 ? (u[q] + dt * ud[q]) : (bc_scale * bc_values[q]);
    }
}

/* cl/velocity_verlet.cl:29-35, per DOF q:
 * ud1 = ud + (dt/2)*udd; udd1 = (F - eta*ud1)/rho; ud = ud1 + (dt/2)*udd1;
 * udd = udd1; u = bc ? scale*val : u + dt*(ud + (dt/2)*udd1) */
void oracle_update_velocity_verlet(int nnodes, const double *force, double *u,
                                   double *ud, double *udd,
                                   const int *bc_types, const double *bc_values,
                                   const double *densities, double bc_scale,
                                   double damping, double dt)
{
#pragma omp parallel for schedule(static) if (nnodes > 4096)
    for (size_t q = 0; q < 3 * (size_t)nnodes; ++q) {
        double ud1 = ud[q] + (dt / 2) * udd[q];
        double udd1 = (force[q] - damping * ud1) / densities[q];
        ud[q] = ud1 + (dt / 2) * udd1;
        udd[q] = udd1;
        u[q] = bc_types[q] == 0 ? (u[q] + dt * (ud[q] + (dt / 2) * udd1))
                                : (bc_scale * bc_values[q]);
    }
}

/* r = r0 + u, elementwise (the numpy `self.coords + u` of
 * integrators.py:440,450). */
void oracle_positions(int nnodes, const double *r0, const double *u,
                      double *r)
{
#pragma omp parallel for schedule(static) if (nnodes > 4096)
    for (size_t q = 0; q < 3 * (size_t)nnodes; ++q)
        r[q] = r0[q] + u[q];
}

/* ---------------------------------------------------------------------- */
/* (1) correction.pyx: stiffness-correction factors (set-up)               */
/* ---------------------------------------------------------------------- */

/* The loop every function of correction.pyx shares (:32-56, :97-120,
 * :155-183, :236-262): rows i ascending, live slots in order; for a
 * neighbour j > i the pair's factor f multiplies slot (i, neigh) and the
 * FIRST slot of row j (within n_neigh[j]) that holds i.  Slots whose
 * neighbour is out of range are skipped (the Cython reads out of bounds). */
typedef double (*oracle_pair_fn)(int i, int j, const void *ctx);

static void oracle_pairwise(int nnodes, int max_neigh, double *corr,
                            const int *nlist, const int *n_neigh,
                            oracle_pair_fn fn, const void *ctx)
{
    for (int i = 0; i < nnodes; ++i) {
        for (int neigh = 0; neigh < n_neigh[i]; ++neigh) {
            int j = nlist[(size_t)i * max_neigh + neigh];
            if (j < 0 || j >= nnodes || !(i < j))
                continue;
            double f = fn(i, j, ctx);
            corr[(size_t)i * max_neigh + neigh] *= f;
            for (int jn = 0; jn < n_neigh[j]; ++jn) {
                if (nlist[(size_t)j * max_neigh + jn] == i) {
                    corr[(size_t)j * max_neigh + jn] *= f;
                    break;
                }
            }
        }
    }
}

struct oracle_geom {
    const double *r0;
    double horizon, node_radius;
    int flag;
};

/* correction.pyx:265-277 (cmicromodulus_connical):
 * l0 <= h ? (h - l0)/h : 0; any flag but 0 gives 1.00 (:246-251). */
static double oracle_conical(int i, int j, const void *ctx)
{
    const struct oracle_geom *g = (const struct oracle_geom *)ctx;
    if (g->flag != 0)
        return 1.00;
    double l0 = oracle_euclid(&g->r0[3 * (size_t)i], &g->r0[3 * (size_t)j]);
    if (l0 <= g->horizon)
        return (g->horizon - l0) / g->horizon;
    return 0.00;
}

/* correction.pyx:186-205 (cPA_HHB): l0 <= h - r -> 1; l0 <= h + r ->
 * (h + r - l0)/(2r); else 0.  Any flag but 0 gives 1.00 (:166-172). */
static double oracle_pa_hhb(int i, int j, const void *ctx)
{
    const struct oracle_geom *g = (const struct oracle_geom *)ctx;
    if (g->flag != 0)
        return 1.00;
    double l0 = oracle_euclid(&g->r0[3 * (size_t)i], &g->r0[3 * (size_t)j]);
    double h = g->horizon, r = g->node_radius;
    if (l0 <= h - r)
        return 1.00;
    if (l0 <= h + r)
        return (h + r - l0) / (2.0 * r);
    return 0.00;
}

struct oracle_surface {
    const double *family_volume; /* NULL: n_neigh[i] * average */
    const int *n_neigh;
    double average, bulk;
};

/* correction.pyx:42-45, :107-109: 2*V_bulk / (V_fam,i + V_fam,j) */
static double oracle_volume_method(int i, int j, const void *ctx)
{
    const struct oracle_surface *s = (const struct oracle_surface *)ctx;
    double vi, vj;
    if (s->family_volume) {
        vi = s->family_volume[i];
        vj = s->family_volume[j];
    } else {
        vi = s->n_neigh[i] * s->average;
        vj = s->n_neigh[j] * s->average;
    }
    return 2. * s->bulk / (vi + vj);
}

/* correction.pyx:208-262 (set_micromodulus_function) */
void oracle_set_micromodulus_function(int nnodes, int max_neigh,
                                      double *values, const double *r0,
                                      const int *nlist, const int *n_neigh,
                                      double horizon, int flag)
{
    struct oracle_geom g = {r0, horizon, 0.0, flag};
    oracle_pairwise(nnodes, max_neigh, values, nlist, n_neigh,
                    oracle_conical, &g);
}

/* correction.pyx:123-183 (set_volume_correction) */
void oracle_set_volume_correction(int nnodes, int max_neigh, double *values,
                                  const double *r0, const int *nlist,
                                  const int *n_neigh, double horizon,
                                  double node_radius, int flag)
{
    struct oracle_geom g = {r0, horizon, node_radius, flag};
    oracle_pairwise(nnodes, max_neigh, values, nlist, n_neigh,
                    oracle_pa_hhb, &g);
}

/* correction.pyx:5-56 (set_imprecise_surface_correction) */
void oracle_set_imprecise_surface_correction(int nnodes, int max_neigh,
                                             double *values,
                                             const int *nlist,
                                             const int *n_neigh,
                                             double average_volume,
                                             double family_volume_bulk)
{
    struct oracle_surface s = {NULL, n_neigh, average_volume,
                               family_volume_bulk};
    oracle_pairwise(nnodes, max_neigh, values, nlist, n_neigh,
                    oracle_volume_method, &s);
}

/* correction.pyx:59-120 (set_precise_surface_correction): the family
 * volume of i is the in-order sum of its live neighbours' volumes
 * (:87-95), starting from 0.0.  `family_volume` is caller scratch (nnodes). */
void oracle_set_precise_surface_correction(int nnodes, int max_neigh,
                                           double *values, const int *nlist,
                                           const int *n_neigh,
                                           const double *volume,
                                           double family_volume_bulk,
                                           double *family_volume)
{
    for (int i = 0; i < nnodes; ++i) {
        double tmp = 0.0;
        for (int neigh = 0; neigh < n_neigh[i]; ++neigh) {
            int j = nlist[(size_t)i * max_neigh + neigh];
            if (j >= 0 && j < nnodes)
                tmp += volume[j];
        }
        family_volume[i] = tmp;
    }
    struct oracle_surface s = {family_volume, n_neigh, 0.0,
                               family_volume_bulk};
    oracle_pairwise(nnodes, max_neigh, values, nlist, n_neigh,
                    oracle_volume_method, &s);
}
