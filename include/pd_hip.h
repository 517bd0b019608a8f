/*
 * pd_hip.h -- C ABI of libpd_hip.so, the MI355X (gfx950) bond-based
 * peridynamics time-stepping library.
 *
 * Every entry point takes plain pointers and sizes, copies what it reads
 * before returning, and returns 0 on success or a negative PD_E_* code; the
 * message of the last failure on the calling thread is pd_last_error().
 * Host arrays are C-contiguous, float64 ("double") or int32 ("int"), in the
 * reference's shapes: (nnodes, 3) for per-node vectors, (nnodes,
 * max_neighbours) for per-bond arrays.
 *
 * The entry points mirror the reference's Integrator interface
 * (peripy/integrators.py), which is what a binding of this path replaces:
 *
 *   pd_device_count      <- peripy/cl/utilities.py:28-44 get_context
 *   pd_build             <- Integrator.build, integrators.py:84-188
 *   pd_create_buffers    <- Integrator.create_buffers, integrators.py:190-254
 *   pd_step              <- Integrator.__call__, integrators.py:330-349,
 *                           484-503, 586-606, 700-720
 *   pd_run               <- the step loop of Model.simulate, model.py:1295-1301
 *   pd_write             <- Integrator.write, integrators.py:285-300
 *   pd_destroy           <- (Python garbage collection of the integrator)
 *
 * and the set-up the reference runs before the first step:
 *
 *   pd_neighbour_list(_device)  <- Model._set_neighbour_list, model.py:536-597
 *   pd_line_crack_pairs         <- initial_crack_helper, model.py:1578-1612
 *   pd_create_crack             <- create_crack.pyx:1-43
 *   pd_set_*_correction,
 *   pd_set_micromodulus_function <- correction.pyx:5-277
 *   pd_set_tips, pd_monitor     <- the per-write sums of model.py:1303-1360
 *
 * Kernel semantics: see the per-function comments and
 * probabilistic_peridynamics_amd/csrc/pd_hip.hip.
 */
#ifndef PD_HIP_H
#define PD_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pd_sim pd_sim;

/* integrator kinds */
#define PD_EULER 0            /* peridynamics.pyx:159-185, cl/euler.cl:3-24 */
#define PD_EULER_CROMER 1     /* cl/euler_cromer.cl:3-35 */
#define PD_VELOCITY_VERLET 2  /* cl/velocity_verlet.cl:3-36 */

/* neighbour-list layouts (model.py:575-585 chooses by integrator.context) */
#define PD_LAYOUT_COMPACT 0  /* Cython: width family.max(), broken bonds are
                                removed by swap-with-last compaction
                                (peridynamics.pyx:102-156) */
#define PD_LAYOUT_SLOT 1     /* OpenCL: width a power of two, -1 marks a broken
                                or padding slot (cl/peridynamics.cl:91) */

/* error codes */
#define PD_OK 0
#define PD_E_ARG -1     /* invalid argument (shape, value, missing array) */
#define PD_E_HIP -2     /* a HIP runtime call failed */
#define PD_E_NODEV -3   /* no usable GPU */
#define PD_E_STATE -4   /* call out of order (e.g. step before buffers) */
#define PD_E_LIMIT -5   /* a size limit of this library was exceeded */

/* Last error message of the calling thread ("" if none). */
const char *pd_last_error(void);

/* Number of visible HIP devices; 0 (and PD_OK) when none.  Replaces
 * get_context (peripy/cl/utilities.py:28-44): the caller raises its
 * ContextError when *count == 0. */
int pd_device_count(int *count);

/* Name and architecture of a device (e.g. "gfx950"), NUL-terminated. */
int pd_device_name(int device, char *name, int name_len, char *arch,
                   int arch_len);

/* Integrator.build (integrators.py:84-188): static per-node arrays.
 *   coords (N,3) f64, volume (N) f64, family (N) i32, bc_types (N,3) i32,
 *   bc_values (N,3) f64, force_bc_types (N,3) i32, force_bc_values (N,3) f64,
 *   stiffness_corrections (N,K) f64 or NULL (bond_force2/4),
 *   bond_types (N,K) i32 or NULL (bond_force3/4),
 *   densities (N,3) f64: required for EULER_CROMER / VELOCITY_VERLET and
 *   rejected for EULER (integrators.py:415-423, 511-519, 610-620).
 * max_neighbours is K, the width of every (N,K) array: at most 65535, and
 * at most 27285 for PD_LAYOUT_COMPACT, whose rows are compacted in LDS
 * (PD_E_LIMIT beyond). */
int pd_build(pd_sim **sim, int device, int integrator, int layout, double dt,
             double damping, int nnodes, int max_neighbours,
             const double *coords, const double *volume, const int *family,
             const int *bc_types, const double *bc_values,
             const int *force_bc_types, const double *force_bc_values,
             const double *stiffness_corrections, const int *bond_types,
             const double *densities);

/* Integrator.create_buffers (integrators.py:190-254): the simulation state.
 *   nlist (N,K) i32, n_neigh (N) i32: for PD_LAYOUT_COMPACT the live bonds of
 *   node i are nlist[i, :n_neigh[i]]; for PD_LAYOUT_SLOT every entry != -1.
 *   nregimes == nbond_types == 1: bond_stiffness / critical_stretch point at
 *   one double each (the PMB model) and plus_cs / regimes are ignored.
 *   Otherwise they are (nbond_types, nregimes) tables, row-major, with
 *   plus_cs from model.py:941-1091, and regimes (N,K) i32 or NULL (zeros);
 *   bond_types must then have been given to pd_build (all zeros allowed).
 *   u, ud, udd: (N,3) f64 or NULL (zeros).
 * May be called again to restart from new state (Model.simulate does). */
int pd_create_buffers(pd_sim *sim, const int *nlist, const int *n_neigh,
                      const double *bond_stiffness,
                      const double *critical_stretch, const double *plus_cs,
                      int nregimes, int nbond_types, const double *u,
                      const double *ud, const double *udd,
                      const int *regimes);

/* Per-node stiffness field c (N) f64, or NULL to clear it.  When set, the PMB
 * bond stiffness of bond (i,j) is (c_i + c_j) * 0.5 instead of the scalar.
 * No counterpart in the reference's kernels: it is the stiffness model of
 * the probabilistic (random-field) samples, as PeriPy users build them by
 * re-running simulate() with new bond_stiffness (model.py:1523-1542). */
int pd_set_node_stiffness(pd_sim *sim, const double *c);

/* pd_set_node_stiffness with c (N) f64 in device memory of the simulation's
 * device (for callers whose field lives on the GPU, e.g. a torch tensor).
 * The caller's writes to c must be complete (its stream synchronized)
 * before the call; c is copied, so it may be reused on return. */
int pd_set_node_stiffness_device(pd_sim *sim, const double *c);

/* Breaking criterion of the PMB model: 0 (default) breaks when
 * !(|s| < s0) (peridynamics.pyx:133); 1 breaks when !(s < s0)
 * (cl/peridynamics.cl:78). */
int pd_set_break_mode(pd_sim *sim, int tension_only);

/* Exact-order step of the Cython path (PD_LAYOUT_COMPACT, PD_EULER, scalar
 * stiffness, no corrections): 1 sums each node's bond forces in the order
 * of peridynamics.pyx:34-99 (bond_force) -- the live neighbours k below the
 * node in ascending k, then those above it in the node's current row order,
 * then the force boundary condition -- with the reference's operations in
 * IEEE double without contraction, so that force and u
 * (peridynamics.pyx:159-185) equal that arithmetic bitwise, not only to
 * rounding.  The step runs on the default step's tiles and LDS halo, four
 * lanes per node folding consecutive list positions, at about twice the
 * default step's time; integrators.Euler turns it on, the library's default
 * is 0 (off).  Rows of at most 4074 slots (PD_E_LIMIT beyond).  Assumes
 * symmetric neighbour lists without repeats, as the model builds them.  No
 * counterpart in the reference (it replaces the whole of Euler.__call__,
 * integrators.py:330-349, for the Cython path). */
int pd_set_exact_order(pd_sim *sim, int on);

/* Launch shape of the step kernel: lanes per node (1, 2, 4 or 8) and the
 * largest number of nodes per tile (a multiple of 64, with
 * tile_nodes * lanes_per_node <= 1024; one workgroup runs one tile).
 * Defaults: 4 and 128.  lanes_per_node takes effect at the next step;
 * tile_nodes at the next pd_create_buffers, which builds the tiles (stepping
 * in between fails with PD_E_STATE).  A tile's
 * halo (its nodes and all their neighbours) is capped at 1280 entries (1152
 * for n-linear models), so a node may have at most 1279 (1151) bonds. */
int pd_set_launch(pd_sim *sim, int lanes_per_node, int tile_nodes);

/* Tiling of the current state: number of tiles, halo capacity per tile
 * (from the device's LDS per workgroup), largest halo of any tile (sets the
 * LDS each workgroup requests), total halo entries, bond slots. */
int pd_tile_stats(pd_sim *sim, int *ntiles, int *hmax, int *hstride,
                  long long *halo_total, long long *slots);

/* Tiling groups: group (N) i32, or NULL for one group.  Tiles never span two
 * groups, so the disjoint meshes of a batch of samples concatenated into one
 * simulation are tiled one sample at a time.  No counterpart in the
 * reference: PeriPy runs its probabilistic samples one simulate() call each
 * (model.py:1523-1542).  Must be set before pd_create_buffers. */
int pd_set_groups(pd_sim *sim, const int *group);

/* Integrator.__call__: one step (break, bond force, boundary conditions,
 * integration).  Asynchronous with respect to the host. */
int pd_step(pd_sim *sim, double displacement_bc_magnitude,
            double force_bc_magnitude);

/* nsteps calls of pd_step with magnitudes disp_mags[s], force_mags[s]
 * (either may be NULL for zeros). */
int pd_run(pd_sim *sim, int nsteps, const double *disp_mags,
           const double *force_mags);

/* pd_run bracketed by HIP events recorded on the library's own stream, the
 * stream every step kernel is launched on; *elapsed_ms is the time between
 * them (synchronous). */
int pd_run_timed(pd_sim *sim, int nsteps, const double *disp_mags,
                 const double *force_mags, float *elapsed_ms);

/* pd_run captured as one hipGraph and launched once (synchronous); if
 * elapsed_ms is not NULL it receives the graph's run time, measured with
 * events on the library's stream.  Replaces the step loop of Model.simulate
 * (model.py:1295-1301) when the magnitudes of all steps are known. */
int pd_run_graph(pd_sim *sim, int nsteps, const double *disp_mags,
                 const double *force_mags, float *elapsed_ms);

/* Adjoint of the Euler step with a per-node stiffness field, for gradients
 * of a loss of the final displacements with respect to the field.  No
 * counterpart in the reference.  Usage: pd_adjoint_start, then nsteps steps
 * by pd_step / pd_run (not pd_run_graph), then pd_adjoint_backward (any
 * number of times).  The recording keeps the step at which each bond broke
 * (the break mask is held fixed in the backward pass) and u every
 * checkpoint_every steps; the backward pass replays each segment from its
 * checkpoint with the step kernel (bitwise the recorded trajectory) and
 * runs the adjoint kernel (on the step kernel's tiles) back through it.
 * Device memory: 14 bytes per bond slot, and per node 256 bytes plus 24
 * per checkpoint and per step of a segment.  Requires PD_EULER,
 * PD_LAYOUT_SLOT and pd_set_node_stiffness.  pd_create_buffers or
 * pd_adjoint_stop ends a recording. */
int pd_adjoint_start(pd_sim *sim, int nsteps, int checkpoint_every);

/* dL_du (N,3) f64: dL/du after the recorded steps.  dL_dc (N) f64 receives
 * dL/dc for the node stiffness field c (k_ij = (c_i + c_j)/2).  Leaves the
 * state as it was at the end of the recording. */
int pd_adjoint_backward(pd_sim *sim, const double *dL_du, double *dL_dc);

/* pd_adjoint_backward with dL_du and dL_dc in device memory of the
 * simulation's device, caller's node order; nothing crosses PCIe.  dL_du
 * must be complete before the call; dL_dc is complete on return. */
int pd_adjoint_backward_device(pd_sim *sim, const double *dL_du,
                               double *dL_dc);

/* Free the recording. */
int pd_adjoint_stop(pd_sim *sim);

/* Tip groups for pd_monitor: entry e is DOF dof[e] (0..2) of node node[e]
 * and belongs to group group[e] (0..ngroups-1); entries keep their order
 * inside a group.  These are the tip lists Model builds from is_tip
 * (model.py:1093-1212).  May be called before or after pd_create_buffers;
 * nentries == 0 clears them. */
int pd_set_tips(pd_sim *sim, int nentries, const int *node, const int *dof,
                const int *group, int ngroups);

/* The sums one periodic write of Model.simulate records (model.py:1303-1360),
 * computed on the device instead of after a full pd_write: model[6] = sum of
 * u, ud, udd, force*V and body_force*V over all DOFs, and sum of damage;
 * tips[5*g .. 5*g+4] = the same five sums over tip group g, accumulated from
 * 0.0 in entry order as the reference's loop does (its values bitwise).  The
 * model sums use a fixed reduction order (not numpy's).  Synchronizes. */
int pd_monitor(pd_sim *sim, double *model, double *tips);

/* pd_monitor without any allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Its partial sums
 * and staged outputs live in a file-scope device array of the library
 * (64 KiB per device, held until the process ends) instead of buffers
 * allocated on first use; results equal pd_monitor's bitwise.  At most 1024
 * tip groups (PD_E_LIMIT beyond).  Every call on one GPU shares that GPU's
 * array, so calls that overlap on one GPU (from several threads or
 * simulations at once) race; calls in sequence do not. */
int pd_monitor_noalloc(pd_sim *sim, double *model, double *tips);

/* Integrator.write (integrators.py:285-300): damage, then device -> host
 * copies.  Any pointer may be NULL to skip that array.  nlist (N,K) and
 * n_neigh (N) come back in the layout given to pd_build, reflecting every
 * broken bond; damage is (family - n_neigh)/family (peridynamics.pyx:29). */
int pd_write(pd_sim *sim, double *u, double *ud, double *udd, double *force,
             double *body_force, double *damage, int *nlist, int *n_neigh);

/* u (N,3) and damage (N) into device memory of the simulation's device, in
 * the caller's node order (either may be NULL); complete on return. */
int pd_write_device(pd_sim *sim, double *u, double *damage);

/* Bond regimes (N,K) i32 of the n-linear model (zeros for PMB). */
int pd_get_regimes(pd_sim *sim, int *regimes);

/* Sum over nodes of the live bond count at this moment (directed bonds). */
int pd_live_bonds(pd_sim *sim, long long *count);

/* Block the host until all work of this simulation has finished. */
int pd_synchronize(pd_sim *sim);

void pd_destroy(pd_sim *sim);

/* Flags of pd_bond_check. */
#define PD_CHECK_TENSION 1 /* break on !(s < s0) instead of !(|s| < s0) */
#define PD_CHECK_OPENCL 2  /* the slot layout's OpenCL stretch */

/* Verification entry point: the step kernel's stretch pipeline on n bonds
 * (r_i, r_j, r0_i, r0_j each (n,3) f64), run on `device`.  Outputs per bond:
 * l = |r_j - r_i| and l0 = |r0_j - r0_i| (spatial.pyx:20-31, bitwise), the
 * kernel's working stretch s, the IEEE-division stretch (l - l0)/l0
 * (spatial.pyx:87-97) and the break decision !(|s| < s0)
 * (peridynamics.pyx:133; !(s < s0) with PD_CHECK_TENSION).  With
 * PD_CHECK_OPENCL, r_i and r_j are the displacements u_i, u_j and the
 * lengths are cl/peridynamics.cl:65-75's, l = |(u_j - u_i) + xi| and
 * l0 = |xi| with xi = r0_j - r0_i, their squared lengths contracted as
 * fma(z, z, fma(x, x, y*y)). */
int pd_bond_check(int device, int n, const double *ri, const double *rj,
                  const double *r0i, const double *r0j, double s0,
                  int flags, double *l, double *l0, double *s,
                  double *s_exact, int *broken);

/* Bond types of a body split by the plane coords[axis] = position, the
 * bulk form of a per-bond is_bond_type predicate (model.py:709-780 calls it
 * once per bond in Python): slot s < n_neigh[i] of row i gets `below` when
 * both ends lie below the plane (x < position), `above` when both lie at or
 * above it, `across` otherwise; the other slots get 0.  coords (N,3) f64,
 * nlist (N,K) and n_neigh (N) i32, bond_types (N,K) i32 out.  device -1
 * runs the same code on the host. */
int pd_plane_bond_types(int device, int nnodes, int max_neighbours,
                        const double *coords, const int *nlist,
                        const int *n_neigh, int axis, double position,
                        int below, int above, int across, int *bond_types);

/* For every node, the index of the first of nboxes axis-aligned boxes
 * [lo_b, hi_b) (lo (B,3), hi (B,3) f64; lo <= x < hi in every coordinate,
 * bounds may be infinite) that contains it, -1 if none: the bulk form of
 * per-node boundary-condition, density and tip predicates built from
 * boxes (model.py:657-707, 1093-1212 call them once per node).  box (N)
 * i32 out.  device -1 runs the same code on the host. */
int pd_first_box(int device, int nnodes, const double *coords, int nboxes,
                 const double *lo, const double *hi, int *box);

/* Neighbour list of a point cloud (no counterpart kernel in the reference:
 * model.py:536-597 builds it on the host with a KD-tree).  Node j is a
 * neighbour of i (j != i) when ((dx*dx + dy*dy) + dz*dz) <= horizon*horizon,
 * d = x_j - x_i.  Rows are sorted ascending by j.
 * Pass 1: nlist == NULL fills family (N) and returns.  Pass 2: nlist (N,K)
 * is filled with pad_value after each row's family[i] entries.  Runs on the
 * host (no GPU needed). */
int pd_neighbour_list(int nnodes, const double *coords, double horizon,
                      int *family, int *nlist, int max_neighbours,
                      int pad_value);

/* pd_neighbour_list computed on a GPU: the same family, rows and padding
 * (the host bins the nodes into cells; the GPU searches, fills and sorts the
 * rows).  max_neighbours at most 4096.  Replaces, like pd_neighbour_list,
 * the KD-tree query of model.py:536-597. */
int pd_neighbour_list_device(int device, int nnodes, const double *coords,
                             double horizon, int *family, int *nlist,
                             int max_neighbours, int pad_value);

/* Stiffness-correction factors, the four functions of peripy/correction.pyx
 * with their argument meaning and order, after a leading `device` (a GPU
 * index, or -1 to run the same code on the host).  Each multiplies the
 * (nnodes, max_neighbours) float64 array it is given, in place, for every
 * live bond (slot s < n_neigh[i]) exactly as the reference's loop does: a
 * pair i < j multiplies slot (i, s) and the first slot of row j that holds
 * i; factors are bitwise the reference's.  Slots naming a node outside
 * [0, nnodes) or i itself are left alone.  n_neigh must lie in
 * [0, max_neighbours].  r0 is (nnodes, 3). */

/* correction.pyx:208-277 set_micromodulus_function: conical
 * (h - l0)/h for l0 <= h, else 0; factor 1 when micromodulus_function != 0 */
int pd_set_micromodulus_function(int device, int nnodes, int max_neighbours,
                                 double *micromodulus_values,
                                 const double *r0, const int *nlist,
                                 const int *n_neigh, double horizon,
                                 int micromodulus_function);

/* correction.pyx:123-205 set_volume_correction (cPA_HHB): 1 for
 * l0 <= h - r, (h + r - l0)/(2r) for l0 <= h + r, else 0; factor 1 when
 * volume_correction != 0 */
int pd_set_volume_correction(int device, int nnodes, int max_neighbours,
                             double *volume_corrections, const double *r0,
                             const int *nlist, const int *n_neigh,
                             double horizon, double node_radius,
                             int volume_correction);

/* correction.pyx:5-56 set_imprecise_surface_correction:
 * 2*V_bulk / (n_neigh[i]*V_avg + n_neigh[j]*V_avg) */
int pd_set_imprecise_surface_correction(int device, int nnodes,
                                        int max_neighbours,
                                        double *stiffness_corrections,
                                        const int *nlist, const int *n_neigh,
                                        double average_volume,
                                        double family_volume_bulk);

/* correction.pyx:59-120 set_precise_surface_correction:
 * 2*V_bulk / (V_fam,i + V_fam,j), V_fam,i the in-order sum of the volumes
 * of i's live neighbours */
int pd_set_precise_surface_correction(int device, int nnodes,
                                      int max_neighbours,
                                      double *stiffness_corrections,
                                      const int *nlist, const int *n_neigh,
                                      const double *volume,
                                      double family_volume_bulk);

/* Initial crack: the bonds (i, j), i < j, that cross a straight crack, in
 * the order initial_crack_helper lists them (model.py:1578-1612: i
 * ascending, then i's row order), found on `device` (-1: the host).  The
 * crack is the bulk form of the reference's pairwise predicates
 * (peripy/test/conftest.py:43-63): with p1 the end of smaller coordinate
 * `axis` and p2 the other, a bond crosses when p1[axis] < split < p2[axis]
 * and the line through p1 and p2 meets coordinate `axis` = at at a
 * `cross_axis` coordinate strictly between lo and hi, computed as
 * m = (p2[c]-p1[c])/(p2[a]-p1[a]); b = p1[c] - m*p1[a]; m*at + b.
 * *npairs receives the number of pairs; pairs (capacity rows of 2 ints) is
 * filled when given and large enough (PD_E_LIMIT otherwise). */
int pd_line_crack_pairs(int device, int nnodes, int max_neighbours,
                        const double *coords, const int *nlist,
                        const int *n_neigh, int axis, int cross_axis,
                        double split, double at, double lo, double hi,
                        int *pairs, long long capacity, long long *npairs);

/* create_crack.pyx:1-43, on the host: for each pair (i, j) in order, remove
 * j from row i and then i from row j -- the first slot holding it takes the
 * row's last live entry, that slot becomes -1, n_neigh decreases.  Every
 * pair must name nodes in [0, nnodes) (checked before anything changes). */
int pd_create_crack(int nnodes, int max_neighbours, long long ncrack,
                    const int *crack, int *nlist, int *n_neigh);

#ifdef __cplusplus
}
#endif

#endif /* PD_HIP_H */
