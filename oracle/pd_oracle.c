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
 *      the peripy/cl kernels with their own arithmetic: xi = r0_j - r0_i,
 *      xi_eta = (u_j - u_i) + xi, the squared lengths contracted into fused
 *      multiply-adds as the OpenCL compiler that wrote the reference's
 *      Euler-Cromer and Velocity-Verlet fixtures evaluated them (those
 *      fixtures are met to ~6e-19 in u with this arithmetic, ~2e-16 without
 *      the contraction and ~7e-15 with the Cython stretch), and the
 *      integrator updates contracted likewise.  An `arith` switch keeps the
 *      Cython stretch available for comparison.  This is the definition the
 *      HIP library implements for the slot-layout integrators.
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

/* Squared length of (x, y, z) as cl/peridynamics.cl:72-73 evaluates
 * x*x + y*y + z*z once contracted: fma(z, z, fma(x, x, y*y)). */
static double cl_sq(double x, double y, double z)
{
    return fma(z, z, fma(x, x, y * y));
}

/* The stretch of one bond with cl/peridynamics.cl:65-75's arithmetic as
 * contracted: xi = r0_j - r0_i, xe = (u_j - u_i) + xi, l0 = sqrt(|xi|^2),
 * l = sqrt(|xe|^2); returns (l - l0)/l0 and stores l, l0. */
double oracle_cl_strain(const double *ui, const double *uj, const double *r0i,
                        const double *r0j, double *l_out, double *l0_out)
{
    double xi[3], xe[3];
    for (int d = 0; d < 3; ++d) {
        xi[d] = r0j[d] - r0i[d];
        xe[d] = (uj[d] - ui[d]) + xi[d];
    }
    double l0 = sqrt(cl_sq(xi[0], xi[1], xi[2]));
    double l = sqrt(cl_sq(xe[0], xe[1], xe[2]));
    *l_out = l;
    *l0_out = l0;
    return (l - l0) / l0;
}

/* Per-slot bond force with in-place breaking, cl/peridynamics.cl
 * bond_force1..4 (:4-131, :134-262, :265-414, :417-565): every node i visits
 * every slot p of its row; nlist[i][p] == -1 is a broken (or padding) slot.
 *   arith 1 (the OpenCL kernels, :65-75 and :206-216):
 *     xi = r0_j - r0_i; xe = (u_j - u_i) + xi; l0 = sqrt(|xi|^2),
 *     y = sqrt(|xe|^2) (squared lengths contracted, cl_sq); s = (y - l0)/l0;
 *     f = s*k*V_j (bond_force1), s*k*corr*V_j (bond_force2),
 *     fma(s, k[t,r], plus_cs[t,r])*V_j (bond_force3; times corr before V_j
 *     in bond_force4); F_i,d += f*(xe_d/y).
 *   arith 0 (the Cython stretch, spatial.pyx:20-31, :87-97): r = r0 + u per
 *     node, l = |r_j - r_i|, l0 = |r0_j - r0_i|, s = (l - l0)/l0, and the
 *     force in peridynamics.pyx:77-89's order: fn = s*k [*corr]; fn = fn/l;
 *     F_i,d += (fn*(r_j,d - r_i,d))*V_j.
 * PMB breaking is peridynamics.pyx:133's !(|s| < s0) unless tension_only,
 * which gives cl/peridynamics.cl:78's !(s < s0).  With a per-node
 * stiffness field cnode, k_ij = (c_i + c_j)*0.5.  The n-linear state machine
 * is cl/peridynamics.cl:326-367 (at most one regime change per step, break
 * when regime >= nregimes), except that the table is only read for a live
 * slot.  Tables are (nbond_types, nregimes) flattened t*nregimes + r
 * (model.py:1042-1070).  Slots are summed in row order from 0.0 (the OpenCL
 * kernel sums them by a tree).  body_force = bond sum; force = bond sum + BC
 * term (cl/peridynamics.cl:117-129; contracted with arith 1). */
void oracle_slot_bond_force(
    int nnodes, int max_neigh, const double *u, const double *r0,
    int *nlist, const double *volume, double bond_stiffness,
    double critical_stretch, int tension_only, const double *cnode,
    const double *corr, const int *bond_types, int *regimes,
    const double *k_tab, const double *s_tab, const double *pc_tab,
    int nregimes, const int *force_bc_types, const double *force_bc_values,
    double force_bc_scale, double *force, double *body_force, int arith)
{
#pragma omp parallel for schedule(static) if (nnodes > 4096)
    for (int i = 0; i < nnodes; ++i) {
        double acc[3] = {0.0, 0.0, 0.0};
        const double *r0i = &r0[3 * (size_t)i], *ui = &u[3 * (size_t)i];
        double ri[3];
        for (int d = 0; d < 3; ++d)
            ri[d] = r0i[d] + ui[d];
        for (int p = 0; p < max_neigh; ++p) {
            size_t slot = (size_t)i * max_neigh + p;
            int j = nlist[slot];
            if (j == -1)
                continue;
            const double *r0j = &r0[3 * (size_t)j], *uj = &u[3 * (size_t)j];
            double dir[3], l, s;
            if (arith) {
                double xi[3];
                for (int d = 0; d < 3; ++d) {
                    xi[d] = r0j[d] - r0i[d];
                    dir[d] = (uj[d] - ui[d]) + xi[d];
                }
                double l0 = sqrt(cl_sq(xi[0], xi[1], xi[2]));
                l = sqrt(cl_sq(dir[0], dir[1], dir[2]));
                s = (l - l0) / l0;
            } else {
                double rj[3];
                for (int d = 0; d < 3; ++d) {
                    rj[d] = r0j[d] + uj[d];
                    dir[d] = rj[d] - ri[d];
                }
                l = oracle_euclid(ri, rj);
                s = oracle_strain2(l, r0i, r0j);
            }
            double k;
            if (k_tab == NULL) {
                int broken = tension_only ? !(s < critical_stretch)
                                          : !(fabs(s) < critical_stretch);
                if (broken) {
                    nlist[slot] = -1;
                    continue;
                }
                k = bond_stiffness;
                if (cnode != NULL)
                    k = (cnode[i] + cnode[j]) * 0.5;
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
                k = k_tab[t * nregimes + reg];
            }
            if (arith) {
                double f = k_tab == NULL
                               ? s * k
                               : fma(s, k, pc_tab[bond_types[slot] * nregimes +
                                                  regimes[slot]]);
                if (corr != NULL)
                    f = f * corr[slot];
                f = f * volume[j];
                for (int d = 0; d < 3; ++d)
                    acc[d] = acc[d] + f * (dir[d] / l);
            } else {
                double fn = k_tab == NULL
                                ? s * k
                                : s * k + pc_tab[bond_types[slot] * nregimes +
                                                 regimes[slot]];
                if (corr != NULL)
                    fn = fn * corr[slot];
                fn = fn / l;
                for (int d = 0; d < 3; ++d)
                    acc[d] = acc[d] + (fn * dir[d]) * volume[j];
            }
        }
        for (int d = 0; d < 3; ++d) {
            size_t q = 3 * (size_t)i + d;
            body_force[q] = acc[d];
            if (force_bc_types[q] == 0)
                force[q] = acc[d];
            else if (arith)
                force[q] = fma(force_bc_scale, force_bc_values[q], acc[d]);
            else
                force[q] = acc[d] + force_bc_scale * force_bc_values[q];
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

/* cl/euler.cl:21-23 as contracted: u = bc ? scale*val : fma(dt, F, u) */
void oracle_update_euler_cl(int nnodes, double *u, const double *bc_values,
                            const int *bc_types, const double *force,
                            double bc_scale, double dt)
{
#pragma omp parallel for schedule(static) if (nnodes > 4096)
    for (size_t q = 0; q < 3 * (size_t)nnodes; ++q)
        u[q] = bc_types[q] == 0 ? fma(dt, force[q], u[q])
                                : bc_scale * bc_values[q];
}

/* cl/euler_cromer.cl:29-34 as contracted, per DOF q:
 * udd = fma(-eta, ud, F)/rho; ud = fma(udd, dt, ud);
 * u = bc ? scale*val : fma(dt, ud, u) */
void oracle_update_euler_cromer(int nnodes, const double *force, double *u,
                                double *ud, double *udd, const int *bc_types,
                                const double *bc_values,
                                const double *densities, double bc_scale,
                                double damping, double dt)
{
#pragma omp parallel for schedule(static) if (nnodes > 4096)
    for (size_t q = 0; q < 3 * (size_t)nnodes; ++q) {
        double uddi = fma(-damping, ud[q], force[q]) / densities[q];
        udd[q] = uddi;
        ud[q] = fma(uddi, dt, ud[q]);
        u[q] = bc_types[q] == 0 ? fma(dt, ud[q], u[q])
                                : (bc_scale * bc_values[q]);
    }
}

/* cl/velocity_verlet.cl:29-35 as contracted, per DOF q:
 * ud1 = fma(dt/2, udd, ud); udd1 = fma(-eta, ud1, F)/rho;
 * ud = fma(dt/2, udd1, ud1); udd = udd1;
 * u = bc ? scale*val : fma(dt, fma(dt/2, udd1, ud), u) */
void oracle_update_velocity_verlet(int nnodes, const double *force, double *u,
                                   double *ud, double *udd,
                                   const int *bc_types, const double *bc_values,
                                   const double *densities, double bc_scale,
                                   double damping, double dt)
{
#pragma omp parallel for schedule(static) if (nnodes > 4096)
    for (size_t q = 0; q < 3 * (size_t)nnodes; ++q) {
        double ud1 = fma(dt / 2, udd[q], ud[q]);
        double udd1 = fma(-damping, ud1, force[q]) / densities[q];
        ud[q] = fma(dt / 2, udd1, ud1);
        udd[q] = udd1;
        u[q] = bc_types[q] == 0 ? fma(dt, fma(dt / 2, udd1, ud[q]), u[q])
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
