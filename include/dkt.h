/*
 * dkt.h -- C ABI of the MI355X-native adaptive k-tree FEM matvec.
 *
 * Plain C: pointers, sizes and int status codes (0 = success, nonzero = error,
 * message in dkt_last_error()).  No torch or C++ types cross this boundary.
 * Every handle owns its memory; calls on different handles are independent
 * (no global mutable state besides the per-thread error string), and device
 * work is ordered on the caller's HIP stream.
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the Dendro-KT source tree).
 */
#ifndef DKT_H
#define DKT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Space-filling curve of the tree order (reference: CMake option
 * HILBERT_ORDERING, CMakeLists.txt:38,214-216; tables src/KDhcurvedata.cpp:25-60). */
#define DKT_CURVE_MORTON 0
#define DKT_CURVE_HILBERT 1

/* Elemental operators with device implementations.
 * DKT_OP_STIFFNESS: Poisson stiffness K_e = h^(dim-2) K_ref (equal to
 *   HeatEq::HeatMat<3>::elementalMatVec, FEM/examples/src/heatMat.cpp:46-117, in 3D).
 * DKT_OP_MASS: mass matrix h^dim M_ref (the operator of HeatVec's RHS,
 *   FEM/examples/src/heatVec.cpp:31-78).
 * DKT_OP_IDENTITY: identity elemental op (test/testMatvec.cpp:187-198).
 * DKT_OP_CAPTURED: dense per-level elemental matrices supplied by the caller
 *   (captured from a user elementalMatVec, FEM/include/feMatrix.h:56),
 *   op_data = [maxdepth + 1][nPe][nPe] row-major (out = K in).  With the
 *   reference's traversal semantics they run on the per-leaf/per-node kernels;
 *   with DKT_PLAN_CONFORMING on the tiled conforming kernel. */
#define DKT_OP_STIFFNESS 0
#define DKT_OP_MASS 1
#define DKT_OP_IDENTITY 2
#define DKT_OP_CAPTURED 3

/* dkt_plan_create flags.  DKT_PLAN_DRYRUN builds the host-side plan and its
 * statistics (dkt_plan_info) without touching the device; such a plan cannot
 * run a matvec. */
#define DKT_PLAN_SIMPLE 1
#define DKT_PLAN_DRYRUN 2
/* The conforming operator sum_e C_e^T K_e C_e (symmetric; see
 * dkt_da_conforming_terms) instead of the reference's traversal semantics;
 * runs on tiles of mesh elements (default 256, halved while the LDS
 * footprint does not fit; at most 2048 constraint terms per tile), or on the
 * per-leaf/per-node kernels with DKT_PLAN_SIMPLE. */
#define DKT_PLAN_CONFORMING 4
/* Leaves per tile of the tiled path: flags |= log2(leaves) << DKT_PLAN_TILE_SHIFT
 * with leaves in {64, 128, 256}; 0 = 128, halved while the LDS footprint does
 * not fit.  Tiles also end early at 32 stale-buffer readers. */
#define DKT_PLAN_TILE_SHIFT 8

typedef struct dkt_octs dkt_octs; /* host list of cells/points */
typedef struct dkt_da dkt_da;     /* host mesh: tree + CG nodes + traversal */
typedef struct dkt_plan dkt_plan; /* device-resident matvec plan */

/* Last error message of the calling thread ("" if none). */
const char *dkt_last_error(void);

/* ---- Octant lists ------------------------------------------------------- */
/* A list of cells or points of a dim-dimensional tree with integer
 * coordinates in [0, 2^maxdepth].  Layout of `cells`: uint32 [n][dim+1] =
 * (anchor coords..., level); a point is a cell of level maxdepth. */
int dkt_octs_create(int dim, int maxdepth, int curve, const uint32_t *cells, size_t n, dkt_octs **out);
int dkt_octs_destroy(dkt_octs *o);
size_t dkt_octs_size(const dkt_octs *o);
int dkt_octs_copy(const dkt_octs *o, uint32_t *cells_out);

/* Sorts in SFC order, ancestors first, removing duplicates.
 * Replaces SFC_Tree::distTreeSort + locRemoveDuplicates (src/tsort.cpp:201-224,718-741). */
int dkt_tree_sort(dkt_octs *o);

/* Complete tree whose leaves contain at most max_pts of the given cells/points.
 * Replaces SFC_Tree::distTreeConstruction (src/tsort.cpp:648-715). */
int dkt_tree_construct(const dkt_octs *pts, size_t max_pts, dkt_octs **tree_out);

/* 2:1 balanced complete tree refining construct(pts, max_pts).
 * Replaces SFC_Tree::distTreeBalancing (src/tsort.cpp:860-878). */
int dkt_tree_balance(const dkt_octs *pts, size_t max_pts, dkt_octs **tree_out);

/* Synthetic inputs: tree refined around a sphere surface (level < base_level
 * always split; level < max_level split when the cell meets the sphere), and
 * Gaussian-mixture points (Irwin-Hall normals, integer exact). */
int dkt_tree_sphere(int dim, int maxdepth, int curve, int base_level, int max_level, const double *center,
                    double radius, dkt_octs **out);
int dkt_points_gauss(int dim, int maxdepth, int curve, size_t n, int ncomp, double sigma, uint64_t seed,
                     dkt_octs **out);
/* U[-1,1) values keyed by node coordinates: x[i] = f(seed, nodes[i]). */
int dkt_keyed_uniform(uint64_t seed, int dim, const uint32_t *nodes, size_t n, double *x_out);

/* ---- Mesh (DA) ---------------------------------------------------------- */
/* Builds the CG node list of a sorted, 2:1 balanced complete tree in the
 * reference's node order.  The matvec traversal over them is flattened on
 * first use (a reference-semantics plan, counts 3-7, dkt_da_traversal), the
 * conforming element records likewise (a conforming plan, counts 8-9).
 * Replaces ot::DA<dim>::DA(tree, nEle, comm, order) (src/oda.cpp:45-151). */
int dkt_da_create(const dkt_octs *tree, int order, dkt_da **out);
int dkt_da_destroy(dkt_da *da);
/* Counts: 0 elements, 1 nodes (DA::getLocalNodalSz), 2 boundary nodes,
 * 3 traversal leaves, 4 hanging leaves, 5 nodes per element, 6 hanging
 * leaves that read a stale parent-buffer entry, 7 stale-term entries,
 * 8 conforming-constraint entries, 9 elements with constrained ranks. */
size_t dkt_da_count(const dkt_da *da, int what);
/* uint32 [N][dim+1] node coords + level (DA::getTNCoords, oda.h:252). */
int dkt_da_nodes(const dkt_da *da, uint32_t *out);
/* uint32 [E][dim+1] (the tree). */
int dkt_da_elements(const dkt_da *da, uint32_t *out);
/* Boundary node ids (DA::getBoundaryNodeIndices, oda.h:264). */
int dkt_da_boundary(const dkt_da *da, uint32_t *out);
/* Flattened traversal: leaf cells uint32 [L][dim+1], child index uint8 [L],
 * slots int32 [L][nPe] (node id or -1), hang_of int32 [L] (-1 or row of
 * pslot), pslot int32 [H][nPe] (parent-level node per parent rank or -1).
 * Any output pointer may be NULL. */
int dkt_da_traversal(const dkt_da *da, uint32_t *leaf_cells, uint8_t *child, int32_t *slot, int32_t *hang_of,
                     int32_t *pslot);

/* Missing nodes of a hanging leaf are interpolated from its parent-level
 * nodes (pslot) as the reference does in its static parent buffer
 * (FEM/include/matvec.h:403-457).  Where a parent rank has no node at the
 * parent level the reference reads the value a previous hanging leaf left in
 * that buffer entry; these stale terms are returned here: for hanging leaf h
 * and k in [off[h], off[h+1]), missing rank rank[k] adds w[k] * x[node[k]].
 * off: int64 [H+1]; rank: uint8, node: int32, w: double [count 7]. */
int dkt_da_stale_terms(const dkt_da *da, int64_t *off, uint8_t *rank, int32_t *node, double *w);

/* Conforming hanging-node constraints (this build's operator for solves at
 * scale; the reference has none -- its static-buffer interpolation above is
 * neither symmetric nor conforming, and its traversal leaves part of the
 * domain unvisited).  Over the mesh elements in order: the h-th element that
 * has ranks at points which are not nodes owns k in [off[h], off[h+1]), where
 * rank rank[k] takes w[k] * x[node[k]]; the value at such a point is the
 * parent cell's interpolant there, recursively through coarser cells.
 * off: int64 [count 9 + 1]; rank, node, w: [count 8]. */
int dkt_da_conforming_terms(const dkt_da *da, int64_t *off, uint8_t *rank, int32_t *node, double *w);

/* Writes the mesh as an ASCII VTK unstructured grid (.vtu) for dim 2 (quads)
 * and 3 (hexahedra), one cell per element with its own corner points in
 * [0,1]^dim, cell data "level", and, when values (double [count 1], host) is
 * not NULL, point data `name` at the corners: the node's value, or at a
 * hanging corner the conforming interpolant.  Replaces the reference's
 * oct2vtu (IO/vtk/include/oct2vtk.h:272-273); 4D meshes are refused. */
int dkt_da_write_vtu(const dkt_da *da, const char *path, const double *values, const char *name);

/* ---- Device plan and matvec -------------------------------------------- */
/* Uploads the flattened traversal of `da` to the current HIP device.
 * op: DKT_OP_*; op_data: for DKT_OP_CAPTURED, double [(maxdepth+1)][nPe][nPe]
 * elemental matrices indexed by leaf level (row-major, out = K in), else NULL.
 * flags: 0 = tiled kernels (default); DKT_PLAN_SIMPLE = per-leaf/per-node
 * kernels through a global contribution buffer (an independent check path;
 * always used for DKT_OP_CAPTURED).
 * Replaces the per-call setup of fem::matvec (FEM/include/matvec.h:231-240):
 * the traversal is flattened once per mesh instead of per call. */
int dkt_plan_create(const dkt_da *da, int op, const double *op_data, int flags, dkt_plan **out);
int dkt_plan_destroy(dkt_plan *plan);

/* y = A x on device, ordered on `stream` (a hipStream_t; NULL = default).
 * dtype 0: double, 1: float (float accumulation).  x, y: device arrays
 * [N][nrhs] (node-major, the reference's dof-interleaved layout), not
 * aliased.  masked = 0: fem::matvec (FEM/include/matvec.h:231-240);
 * masked = 1: feMatrix::matVec with HeatMat's Dirichlet hooks
 * (FEM/include/feMatrix.h:189-259, FEM/examples/src/heatMat.cpp:120-139):
 * boundary entries of x are read as zero and boundary entries of y are zero;
 * masked = 2: feVector::computeVec with HeatVec's hooks
 * (FEM/include/feVector.h:122-168, FEM/examples/src/heatVec.cpp:78-97):
 * boundary entries of y are zero (x is read as given). */
int dkt_plan_matvec(dkt_plan *plan, int dtype, const void *x, void *y, int nrhs, int masked, void *stream);

/* dkt_plan_matvec that allocates no device memory, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same arguments and
 * result, bit for bit, as dkt_plan_matvec.  The matvec's scratch (the partial
 * sums of nodes shared by several tiles; the contribution buffer of simple
 * plans) lives in one fixed device array of DKT_NOALLOC_SCRATCH_BYTES per GPU,
 * held until the process ends; a matvec that needs more (partial slots x
 * columns per launch x element size) is refused with an error and nothing
 * runs.  Every call on one GPU shares that array: calls that overlap on one
 * GPU (other streams or threads) race and must be ordered by the caller. */
#define DKT_NOALLOC_SCRATCH_BYTES (32u << 20)
int dkt_plan_matvec_noalloc(dkt_plan *plan, int dtype, const void *x, void *y, int nrhs, int masked, void *stream);

/* Replaces the plan's Dirichlet node set (the nodes the masked matvecs and
 * the solvers' masked operators zero; dkt_plan_create takes the DA's boundary
 * nodes, DA::getBoundaryNodeIndices, as HeatMat's hooks do) with ids[0..n).
 * For operators whose essential boundary is part of the domain boundary, such
 * as the space-time heat form (spatial boundary and initial time only).  Waits
 * for the device to finish work that may read the old set.  No counterpart in
 * the reference (its hooks are the leaf class's preMatVec/postMatVec,
 * FEM/include/feMatrix.h:91-121). */
int dkt_plan_set_dirichlet(dkt_plan *plan, const uint32_t *ids, size_t n);

/* ---- Solver ------------------------------------------------------------ */
/* Conjugate gradients on A x = b with the plan's operator, as
 * HeatMat::cgSolve (FEM/examples/src/heatMat.cpp:165-325): unpreconditioned,
 * one matvec per iteration, stop when ||b - A x||_inf / ||b||_inf <= tol
 * (||b||_inf = 1 if b = 0) or after max_iter iterations.  masked = 1 uses the
 * Dirichlet-masked operator of feMatrix::matVec (as cgSolve does through
 * HeatMat's pre/post hooks); then b should vanish on boundary nodes.
 * mode DKT_CG_F64: all FP64.  The mixed modes run CG with FP32 vectors and
 * matvecs (FP32 accumulation) and FP64 dot products; the stop test uses the
 * FP64 residual.  DKT_CG_MIXED_RELIABLE: reliable updates, the FP32 update of
 * x is folded into x in FP64 and the residual replaced by the FP64 b - A x
 * whenever the FP32 residual has dropped tenfold, keeping the search
 * direction.  DKT_CG_MIXED_RESTART: iterative refinement, each FP32 inner
 * solve restarted from zero on the FP64 residual.  DKT_CG_MIXED: reliable
 * updates on conforming plans (symmetric operator), restarts otherwise (the
 * reference operator is not symmetric at hanging nodes).
 * b, x: device double [N]; x holds the initial guess on entry.  Reductions
 * run in a fixed order (bitwise reproducible).  stats (may be NULL): [0]
 * CG iterations (FP32 iterations for the mixed modes), [1] final relative
 * residual, [2] residual replacements / refinement steps, [3] matvecs,
 * [4] stop reason: DKT_CG_CONVERGED, DKT_CG_MAX_ITER, DKT_CG_BREAKDOWN
 * (p.Ap <= 0 or a non-finite step: the operator is not positive definite on
 * the search direction; the iteration stops before x changes) or
 * DKT_CG_STALLED (the mixed modes' refinement made no progress).
 * Returns 0 when the solve converged, reached max_iter or stalled (compare
 * stats[1] with tol), DKT_ERR_BREAKDOWN on a breakdown (stats are filled and
 * x holds the last iterate), 1 on invalid arguments or a device error. */
#define DKT_CG_CONVERGED 1
#define DKT_CG_MAX_ITER 2
#define DKT_CG_BREAKDOWN 3
#define DKT_CG_STALLED 4
#define DKT_ERR_BREAKDOWN 2
#define DKT_CG_F64 0
#define DKT_CG_MIXED 1
#define DKT_CG_MIXED_RESTART 2
#define DKT_CG_MIXED_RELIABLE 3
int dkt_cg_solve(dkt_plan *plan, int mode, const double *b, double *x, int max_iter, double tol, int masked,
                 double *stats, void *stream);

/* BiCGStab (van der Vorst) on A x = b for nonsymmetric operators, FP64, e.g.
 * the stabilised space-time heat form of a DKT_OP_CAPTURED plan.  Same
 * arguments, stop rule (relative L_inf residual), masks, stats layout and
 * status codes as dkt_cg_solve (stats[3]: 1 + 2 matvecs per iteration);
 * breakdown (r^.v = 0, rho = 0, omega = 0) returns DKT_ERR_BREAKDOWN.
 * Reductions run in a fixed order (bitwise reproducible).  No counterpart in
 * the reference library (its heat driver hands the operator to PETSc's KSP,
 * FEM/examples/src/heatEq.cpp:153-165). */
int dkt_bicgstab_solve(dkt_plan *plan, const double *b, double *x, int max_iter, double tol, int masked,
                       double *stats, void *stream);

/* Batched FP64 conjugate gradients: solves A_k x_k = b_k for nplans tiled
 * plans (any mix of (dim, p), reference or conforming; no simple or
 * per-level plans)
 * together on one stream, each as dkt_cg_solve with DKT_CG_F64 from x_k = 0
 * (x_k is overwritten).  Per iteration one matvec launch per (dim, p) group
 * and one launch per vector step serve every solve; a converged solve's
 * kernels return at once.  Each solve's arithmetic and reduction order are
 * those of its own dkt_cg_solve: x_k is bitwise what dkt_cg_solve(plans[k],
 * DKT_CG_F64, b_k, x_k = 0, ...) returns.  A plan may appear several times
 * (each solve has its own workspace).  stats (may be NULL): 3 * nplans doubles,
 * [3k] iterations, [3k + 1] final relative residual and [3k + 2] stop reason
 * (DKT_CG_CONVERGED, DKT_CG_MAX_ITER or DKT_CG_BREAKDOWN) of solve k.
 * No counterpart in the reference (HeatMat::cgSolve solves one system per
 * call, FEM/examples/src/heatMat.cpp:165); for servers batching requests. */
int dkt_cg_solve_batch(dkt_plan *const *plans, int nplans, const double *const *b, double *const *x, int max_iter,
                       double tol, int masked, double *stats, void *stream);

/* Continuous batching of CG solves (dkt_cg_solve_batch runs on it): solves
 * join with dkt_cg_engine_add(_mode) and leave when they stop, between chunks
 * of 8 iterations, so a server keeps the device busy with whatever is running
 * instead of waiting for the slowest solve of a batch.  Each solve is bitwise
 * its own dkt_cg_solve from x = 0 in its mode (x is overwritten), with its own
 * tol and max_iter; masked is per engine.  No counterpart in the reference.
 * dkt_cg_engine_create: at most max_slots (1..65535) solves at a time.
 * dkt_cg_engine_add: an FP64 solve (DKT_CG_F64) on a tiled plan (reference or
 *   conforming; several solves may share one), device b and x of its N nodes
 *   (they must stay valid until the solve stops); *slot gets the solve's
 *   slot, whose workspace the engine keeps for later solves.  Fails when
 *   every slot is taken.
 * dkt_cg_engine_add_mode: the same with mode DKT_CG_F64, or
 *   DKT_CG_MIXED / DKT_CG_MIXED_RELIABLE: mixed precision with reliable
 *   updates (FP32 iteration and matvecs, FP64 residual replacements, stop on
 *   the FP64 residual), bitwise dkt_cg_solve(DKT_CG_MIXED_RELIABLE) from
 *   x = 0.  The replacements of every mixed solve that is due run together
 *   at the start of a step.
 * dkt_cg_engine_step: starts the solves added since the last step, runs
 *   chunks * 8 iterations of every running solve on `stream`, synchronizes,
 *   and reports the solves that stopped (converged, max_iter reached or
 *   broken down): finished[i] = slot, stats[3i] = iterations, stats[3i + 1]
 *   = relative residual, stats[3i + 2] = stop reason (DKT_CG_*; stats may be
 *   NULL), *nfinished = their count; finished has max_slots entries, stats
 *   3 * max_slots.  Their slots are free again.  One engine, one thread. */
typedef struct dkt_cg_engine dkt_cg_engine;
int dkt_cg_engine_create(int max_slots, int masked, dkt_cg_engine **out);
int dkt_cg_engine_add(dkt_cg_engine *engine, dkt_plan *plan, const double *b, double *x, int max_iter, double tol,
                      int *slot);
int dkt_cg_engine_add_mode(dkt_cg_engine *engine, dkt_plan *plan, int mode, const double *b, double *x, int max_iter,
                           double tol, int *slot);
int dkt_cg_engine_step(dkt_cg_engine *engine, int chunks, void *stream, int *finished, double *stats,
                       int *nfinished);
int dkt_cg_engine_destroy(dkt_cg_engine *engine);

/* ---- Batches ----------------------------------------------------------- */
/* One launch per (dim, p) over the tiles of many plans (independent meshes),
 * each with its own device vectors x[k], y[k] of [N_k][nrhs] (dtype 0 f64,
 * 1 f32), plus one fix-up launch for the shared nodes of every plan.  flags
 * DKT_BATCH_GRAPH: the launches are captured into a hipGraph on the first
 * dkt_batch_run (the (dim, p) launches on parallel branches, joined before the
 * fix-up) and replayed after (the stream must not be the default stream;
 * the branches use streams of default priority).  Plans and vectors are bound at
 * creation and must outlive the batch.  Every y[k] equals dkt_plan_matvec of
 * plan k bit for bit.  No counterpart in the reference (one fem::matvec per
 * call, FEM/include/matvec.h:231-240). */
typedef struct dkt_batch dkt_batch;
#define DKT_BATCH_GRAPH 1
/* DKT_BATCH_PERSISTENT: one persistent launch per run for the whole batch,
 * any mix of (dim, p): the tiles of every plan (work claimed from eight
 * SFC-contiguous queues, one per XCD) and then the fix-up of their shared
 * nodes, each fix-up block waiting on its plan's tile-completion counter; the
 * kernel resets its counters on exit, so with DKT_BATCH_GRAPH a replay is one
 * graph node.  Tiled reference-semantics plans only.  Switched off:
 * dkt_batch_create refuses the flag (its launch did not finish on the device
 * and the cause is not found). */
#define DKT_BATCH_PERSISTENT 2
/* Columns per tile record for nrhs > 1: flags |= (log2(columns) + 1) <<
 * DKT_BATCH_COLUMNS_SHIFT with columns in {1, 2, 4, ..., 64}; 0 = as many
 * as the tiles' LDS allows.  The columns are cut into chunks of that many,
 * each chunk a record of the same launch (one tile launch and one fix-up
 * launch per (dim, p) group).  One column per record runs the single-column
 * tile kernels (fewer registers, more waves per SIMD) over the same
 * [N][nrhs] vectors; the autotuner chooses it per (dim, p).  Each column's
 * result is bitwise the same either way. */
#define DKT_BATCH_COLUMNS_SHIFT 8
/* x[k], y[k] column-major, [nrhs][N_k] (column q at q * N_k), instead of
 * [N_k][nrhs]: every (column, plan) pair is a single-column record (the
 * single-column kernels, whose reads of x are then unit-stride), and all
 * columns of a (dim, p) group share one tile launch and one fix-up launch.
 * Bitwise the results of the node-major layout, column by column. */
#define DKT_BATCH_COLUMN_MAJOR 4
int dkt_batch_create(dkt_plan *const *plans, int nplans, int dtype, int nrhs, const void *const *x, void *const *y,
                     int masked, int flags, dkt_batch **out);
int dkt_batch_run(dkt_batch *batch, void *stream);
/* Persistent batches: *err = 1 when a fix-up block gave up waiting for its
 * plan's tiles (a bounded spin; the results of that run are not valid), else 0.
 * Synchronizes with the device. */
int dkt_batch_status(dkt_batch *batch, int *err);
int dkt_batch_destroy(dkt_batch *batch);

/* info[0..8] = nodes, leaves, hanging leaves, nodes per element, device map
 * bytes, tiles, shared nodes, partial slots, max nodes per tile;
 * info[9..18] = tile statistics of the tiled path: mean and max stale-buffer
 * readers per tile, mean and max tile-start buffer terms per tile, mean
 * sibling groups, mean nodes per tile, hanging-transpose entries, mean cone
 * length per reader, mean and max longest replay per tile with readers;
 * info[19] = leaves per tile; info[20] = max constraint terms per tile of a
 * tiled conforming plan; info[21] = max hanging leaves per tile; info[22] =
 * LDS bytes per workgroup of the tiled kernel (FP64, one column). */
int dkt_plan_info(const dkt_plan *plan, double *info, int n);

/* Event timing of subsequent tiled matvecs on their stream (no counterpart in
 * the reference; replaces the bench timers of matvec_bench_adaptive.cpp:199-216).
 * profile_read synchronizes the recorded events and returns the summed tile
 * kernel time and whole-matvec time in ms over ncalls launches, then clears. */
int dkt_plan_profile(dkt_plan *plan, int enable);
int dkt_plan_profile_read(dkt_plan *plan, double *tile_ms, double *total_ms, int *ncalls);

#ifdef __cplusplus
}
#endif
#endif /* DKT_H */
