/*
 * statics_hip: an MI355X (gfx950) backend for the statics crate's solve:
 * the forces in a voxel structure under load, all nodes at once, by
 * conjugate gradients preconditioned with an octree multigrid V-cycle.
 *
 * The model is the reference's face-node/beam lattice (statics/src/model.rs,
 * statics/src/lattice.rs): every free cell's six faces are nodes with six
 * degrees of freedom (translation, rotation); the cell joins each pair of its
 * faces by a beam of four springs ("bonds"). A face of a held cell is held.
 * Cells are stored on the device in 8^3 bricks aligned to multiples of 8 of
 * the lattice (octree level 3 above the lattice's level).
 *
 * Every entry point returns an StStatus. Output buffers are owned by the
 * caller. One context owns one device stream and its buffers; calls on one
 * context are serialised by the caller (as a Rust &mut would be).
 *
 * Node order. Nodes are the faces of the free cells. Each is named by its
 * face: orientation o (0 x, 1 y, 2 z) and the lattice coordinates of the
 * cell on its + side (the greater of the two cells it lies between, which is
 * the reference's around[v][1]). st_node_faces lists them in the device's
 * order; st_lookup_faces maps faces to device indices, so a caller can keep
 * the reference's own node order (statics/src/lattice.rs:124-163).
 */
#ifndef STATICS_HIP_H
#define STATICS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  ST_OK = 0,
  ST_BAD_INPUT = 1,     /* inconsistent or out-of-range arguments */
  ST_NO_CONVERGENCE = 2,/* the iteration cap met before the tolerance (x holds the last iterate) */
  ST_NOT_SPD = 3,       /* a non-positive curvature p.Ap or a singular block */
  ST_NAN = 4,           /* a NaN or Inf in the iteration */
  ST_OOM = 5,           /* a device allocation failed */
  ST_DEVICE = 6,        /* any other HIP error */
  ST_NO_STRUCTURE = 7   /* an entry point needing uploaded cells called before st_upload */
} StStatus;

typedef struct StCtx StCtx;

/* One structure's cells, as the reference's `solids` after hold_bulk
 * (statics/src/lattice.rs:50-58, 87-119): each solid cell, its site in
 * lattice units (the reference's site >> level), whether it is held, and for
 * free cells the stiffness of its material along (e) and across (g) a bond,
 * in pascals. beams: per cell a 15-bit mask of its beams that are alive, bit
 * q for the q-th face pair (a, b), a < b in VOXEL_FACES order
 * (lattice/src/lib.rs:64); NULL means all alive. Cells may come in any order;
 * a site may appear once. */
typedef struct {
  int64_t n_cells;
  const int32_t* sites; /* 3 per cell: x, y, z */
  const uint8_t* held;  /* 1 per cell */
  const double* e;      /* 1 per cell (ignored for held cells) */
  const double* g;      /* 1 per cell */
  const uint16_t* beams;/* 1 per cell, or NULL */
  double spacing;       /* metres per cell: UNIT * 2^level (statics/src/lattice.rs:34-36) */
  const double* section;/* 15 per cell, or NULL: beam q's section over SECTION s^2, each beam's
                           own section as gather::measured gives it (statics/src/gather.rs:
                           386-390): its force terms scale with it, its couple with its square
                           (K4 = A / d, R2 = A^2 / (4 pi d)). NULL: 1 for every beam, the
                           sections lattice::build gives (lattice.rs:151-158) */
} StCells;

/* ST_FP64: everything in FP64. ST_MIXED: the CG (vectors, K p, inner products,
 * step lengths) stays in FP64 and the V-cycle preconditioner runs in FP32 on the
 * fine level and the first four coarse levels (their operators, transfers and
 * smoothers), FP64 above and in the dense top. ST_MIXED32: the V-cycle as ST_MIXED, and the
 * CG's vectors and its K p in FP32 too, with FP64 kept for the inner products, the step
 * lengths, the solution (the FP32 steps are added into it in FP64) and the residual: each
 * time the FP32 recursive residual has fallen a hundredfold, and when it meets the
 * tolerance, the FP64 true residual f - K x replaces it (iterative refinement by residual
 * replacement). All recompute the FP64 true residual when the recursive one meets the
 * tolerance, and go on from it if it does not. */
typedef enum { ST_FP64 = 0, ST_MIXED = 1, ST_MIXED32 = 2 } StPrecision;

typedef struct {
  int32_t precision;   /* StPrecision */
  double tol;          /* stop when ||f - K x||_2 <= tol * ||f||_2 (statics/src/bodies.rs:350-352) */
  int32_t max_iter;    /* cap on CG iterations (the reference: 100,000) */
  int32_t refresh;     /* 0: the FP64 true residual must meet tol (the solve restarts from
                          it when the recursive residual does, and stops when it stalls);
                          < 0: the reference's rule (statics/src/bodies.rs:350-352), the
                          recursive residual decides; rel_residual reports the true one */
  int32_t warm_start;  /* nonzero: start from x_out's contents instead of 0 */
  int32_t release_hierarchy; /* nonzero: free the multigrid hierarchy when the solve ends,
                                so the next solve (st_backward's adjoint included) recomputes
                                it instead of reading a stored one; 0: keep it for the next
                                solve of the same operator in the same precision mode */
} StSolveOpts;

typedef struct {
  int32_t iters;          /* CG iterations taken */
  int32_t levels;         /* multigrid levels, the finest included */
  int32_t refreshes;      /* FP64 true-residual recomputations */
  int32_t status;         /* StStatus of the solve */
  double rel_residual;    /* final ||f - K x|| / ||f||, recomputed in FP64 */
  double setup_ms;        /* numbering, smoother, hierarchy (bookkeeping only when kept) */
  double solve_ms;        /* the iterations */
  int64_t nodes;          /* face nodes (held and gone included) */
  int64_t dof_nodes;      /* free nodes: not held, not gone */
  int64_t cells;          /* free cells */
  int64_t bricks;         /* bricks on the device */
  int32_t retries;        /* restarts, 0 to 2: under ST_MIXED32, the FP64 CG taking over from
                             an FP32 CG that stopped short (from its solution); and the
                             solve repeated from zero with the robust cycle settings (two
                             cycles on coarse levels 3-6, FP64 from level 3) after the fast
                             ones failed */
  int32_t first_status;   /* with retries > 0: the StStatus the first attempt ended with */
  int64_t mem_peak;       /* the most device memory (bytes) the library's buffers held in
                             this process during the call */
} StStats;

/* The layouts a binding must reproduce (x86-64 and aarch64 C ABI; statics_hip checks
 * them at compile time): StSolveOpts is 32 bytes, precision at 0, tol at 8, max_iter at
 * 16, refresh at 20, warm_start at 24, release_hierarchy at 28. StStats is 88 bytes:
 * iters 0, levels 4, refreshes 8, status 12, rel_residual 16, setup_ms 24, solve_ms 32,
 * nodes 40, dof_nodes 48, cells 56, bricks 64, retries 72, first_status 76, mem_peak 80. */

/* ---- context ---------------------------------------------------------- */
int st_create(StCtx** ctx, int device);
int st_destroy(StCtx* ctx);
/* The last error's text for this context (static storage, overwritten by the next call). */
const char* st_last_error(const StCtx* ctx);

/* ---- structure -------------------------------------------------------- */
/* Replaces: lattice::build (statics/src/lattice.rs:124-163). Uploads the cells
 * into device bricks and numbers the face nodes; no lattice of bonds is ever
 * materialised. */
int st_upload(StCtx* ctx, const StCells* cells);
/* A world tile as the reference's brick grid, lattice::Grid
 * (lattice/src/lib.rs:182-205), as its caller holds it: */
typedef struct {
  int32_t min[3];         /* the box's least site (Grid.min) */
  int32_t n[3];           /* the box's size in sites (Grid.n) */
  const int32_t* bricks;  /* Grid.bricks: per 8^3 place, z-major, then y, then x, where its
                             sites start in Grid.cells in units of 512; -1 if empty */
  int64_t nbricks;        /* length of bricks: ceil(n/8) over the three axes */
  const uint32_t* solid;  /* Grid.solid: bit i % 32 of word i / 32 for site i of Grid.cells */
  int64_t nsites;         /* length of Grid.cells (a multiple of 512) */
  /* Or a worldgen::grid::VoxelGrid (worldgen/src/grid.rs:69-81, the same brick layout
   * and site indexing, worldgen/src/grid.rs:530-537): with solid_bytes not NULL, a
   * site is a cell when VoxelGrid.solid (how much of ALPHA = 255 of the site stops a
   * walking body) is at least solid_min, and `solid` is not read; nsites is then the
   * length of VoxelGrid.solid. */
  const uint8_t* solid_bytes;
  uint8_t solid_min;
} StGrid;

/* Replaces: statics::lattice::solids + hold_bulk (statics/src/lattice.rs:87-119)
 * for a tile held as a lattice::Grid (or a VoxelGrid, above). The cells are the
 * sites whose Grid.solid bit is set (Grid::solid, lattice/src/lib.rs:332-335) or
 * whose VoxelGrid.solid byte reaches solid_min, in solids' order (x outermost,
 * z innermost); each is held if every site within hold_radius metres (at spacing
 * UNIT * 2^level) is a cell, as hold_bulk decides it. Both run on the device.
 * Writes the count to *count_out; with sites_out (3 int32 per cell, grid units),
 * held_out (1 per cell) and index_out (the cell's index in Grid.cells, for its
 * material) not NULL, also the cells, each array sized for the count (a first
 * call with them NULL gives it). The result is st_upload's input, with e and g
 * per cell from the caller's materials. */
/* A lattice as the reference holds it (statics/src/lattice.rs:61-71, model.rs:92-111): per
 * node its position, weight, material, held flag and the two sites it lies between
 * (Lattice.around: lesser first, level-0 units); per bond its nodes, point, direction,
 * share of the section, beam length and material. A material is six values, rho, e, g,
 * ft, fc, fs (model.rs:43-51); two are the same material when all six are equal (the
 * reference compares names, one to each set of values among its materials). */
typedef struct {
  uint32_t level;
  int64_t nnodes;
  const double* x;         /* 3 per node */
  const double* m;
  const double* node_mat;  /* 6 per node */
  const uint8_t* held;
  const int64_t* around;   /* 6 per node */
  int64_t nbonds;
  const int64_t* ij;       /* 2 per bond */
  const double* p;         /* 3 per bond */
  const double* n;         /* 3 per bond */
  const double* a;
  const double* d;
  const double* bond_mat;  /* 6 per bond */
} StLattice;

/* Replaces: statics::gather::gathered (statics/src/gather.rs:193-195; against_full = 1)
 * and measured (:228-395; against_full = 0): the next level up, each of its cells' beams
 * measured from the finer matter in it. The finer lattice can be one built at level 0 or
 * one gathered before; a coarse cell holds at most 12 free finer nodes, 6 faces and 480
 * bonds (a lattice the reference builds or gathers never has more; more is ST_BAD_INPUT).
 * The result is kept in the context until the next st_gather: its sizes go to
 * nnodes_out and nbonds_out, st_gather_read copies it out in StLattice's layout, in the
 * reference's order (nodes by face, bonds cell by cell, PER_BEAM to a beam). */
int st_gather(StCtx* ctx, const StLattice* fine, int against_full, int64_t* nnodes_out, int64_t* nbonds_out);

/* The cells a lattice was built from, for st_upload. */
typedef struct {
  int64_t n_cells;        /* out: cells (those with beams first, then held ones) */
  double spacing;         /* out: UNIT * 2^level */
  int32_t* sites;         /* 3 per cell, level units; NULL on the sizing call */
  uint8_t* held;          /* 1 per cell, or NULL */
  double* e;              /* 1 per cell, or NULL (0 for held cells) */
  double* g;              /* 1 per cell, or NULL */
  double* rho;            /* 1 per cell, or NULL */
  uint16_t* beams;        /* 1 per cell, or NULL: the beams the lattice holds (bit q, pair q) */
  double* strength;       /* 45 per cell, or NULL: ft, fc, fs of each beam q (0 where absent) */
  int64_t* bond_index;    /* 1 per bond, or NULL: the bond's place 60 * cell + 4 * q + k in the
                             per-bond outputs (st_bond_used; st_bond_forces at 3 times it), or
                             -1 minus that place for a beam that runs from the greater face to
                             the lesser (gather::measured makes such): the force on its j side
                             is then the negated st_bond_forces value, its used share the same */
  double* section;        /* 15 per cell, or NULL: each beam's section over SECTION s^2 (1 where
                             absent), for StCells.section */
} StLatticeCells;

/* Replaces: the caller's side of lattice::build (statics/src/lattice.rs:124-163) run
 * backwards, for a caller that holds a Lattice (statics/src/lattice.rs:61-71) rather than
 * its cells: each beam's cell is the site its end nodes share (as statics/src/gather.rs:245
 * finds it), its pair the two faces of that cell its nodes are; a held node's other cell is
 * held (lattice.rs:138). Host only; no context. Takes a lattice of full cells at its level
 * as build makes them (strengths per beam allowed, bodies::as_generated,
 * statics/src/bodies.rs:611-625), or as gather::measured makes the next level's (beams of
 * their own sections and strengths, statics/src/gather.rs:386-390; out->section then
 * receives the sections, and a lattice with any section other than SECTION s^2 is refused
 * when out->section is NULL): every beam four bonds in lattice::beam's order and geometry
 * (lattice.rs:76-84) between face centres, one e and g to a cell, one material to a beam.
 * Anything else is ST_BAD_INPUT, with the reason in message. Call once with out->sites
 * NULL for n_cells, then with arrays of that size. */
int st_lattice_cells(const StLattice* lat, StLatticeCells* out, char* message, int64_t message_len);
int st_gather_read(StCtx* ctx, uint32_t* level, double* x, double* m, double* node_mat, uint8_t* held,
                   int64_t* around, int64_t* ij, double* p, double* n, double* a, double* d, double* bond_mat);

int st_grid_cells(StCtx* ctx, const StGrid* grid, int32_t level, double hold_radius, int64_t* count_out,
                  int32_t* sites_out, uint8_t* held_out, int64_t* index_out);

/* Counts of what st_upload built. */
int st_sizes(StCtx* ctx, int64_t* nodes, int64_t* cells, int64_t* bricks);
/* Each node's face, 4 int32 per node: o, x, y, z (the + side cell), in device order. */
int st_node_faces(StCtx* ctx, int32_t* faces_out);
/* Device index of each face (4 int32 each, as st_node_faces); -1 where no node. */
int st_lookup_faces(StCtx* ctx, int64_t n, const int32_t* faces, int64_t* index_out);
/* Per node: 1 if held (a face of a held cell; statics/src/lattice.rs:138). */
int st_node_held(StCtx* ctx, uint8_t* held_out);

/* ---- classification --------------------------------------------------- */
/* Replaces: bodies::joined_to_held (statics/src/bodies.rs:557-574). Per node,
 * 1 if a chain of alive beams joins it to a held node. beams_alive: per cell,
 * in st_upload's cell order, the 15-bit alive mask (NULL: the uploaded masks).
 * The result also becomes the context's `gone` (gone = !joined) unless
 * keep_gone is nonzero. */
int st_joined_to_held(StCtx* ctx, const uint16_t* beams_alive, uint8_t* joined_out, int keep_gone);
/* Replaces: bodies::joined_to_held (statics/src/bodies.rs:557-574) with the reference's own
 * arguments, a lattice's nodes and bonds (no upload needed): per node, 1 if a chain of alive
 * bonds joins it to a held node. held: 1 per node; ij: 2 per bond (the bond's nodes i, j);
 * alive: 1 per bond, or NULL (all alive). Bit-exact (reachability is order-free). */
int st_joined_edges(StCtx* ctx, int64_t nnodes, const uint8_t* held, int64_t nbonds, const int64_t* ij,
                    const uint8_t* alive, uint8_t* joined_out);
/* Sets the nodes taken out of the solve (the `gone` argument of motions,
 * statics/src/bodies.rs:429); NULL clears it. Beams touching a gone node are
 * not live (statics/src/bodies.rs:434). */
int st_set_gone(StCtx* ctx, const uint8_t* gone);
/* Replaces the per-cell alive masks (settle's broken beams, statics/src/bodies.rs:525-533). */
int st_set_beams(StCtx* ctx, const uint16_t* beams_alive);

/* New stiffnesses e, g for the uploaded cells (upload order; held cells ignored),
 * keeping the lattice: what an optimisation loop over per-cell stiffness changes
 * between solves (the reference rebuilds its Lattice, statics/src/lattice.rs:124-163).
 * The next solve rebuilds the hierarchy. */
int st_set_stiffness(StCtx* ctx, const double* e, const double* g);

/* ---- operator --------------------------------------------------------- */
/* Replaces: Level::apply at level 0 (statics/src/bodies.rs:193-206): y = K x
 * over the live beams, 6 doubles per node in device order; rows of held and
 * gone nodes are 0 and their x is ignored. */
int st_apply(StCtx* ctx, const double* x, double* y);

/* ---- solve ------------------------------------------------------------ */
/* Replaces: bodies::motions (statics/src/bodies.rs:429-468). loads: 6 doubles
 * per node in device order (force, moment), read for free nodes only. x_out:
 * 6 doubles per node; 0 for held and gone nodes.
 * The multigrid hierarchy built by a solve is kept for the next solve of the same
 * operator in the same precision mode (st_solve, st_solve_resident, st_backward's
 * adjoint): st_upload, st_set_beams, st_set_gone, st_joined_to_held, st_classify
 * and st_set_stiffness drop it, and the solve after them rebuilds it. A solve on a
 * kept hierarchy reports only its per-solve bookkeeping as setup_ms. */
int st_solve(StCtx* ctx, const double* loads, const StSolveOpts* opts, double* x_out, StStats* stats);

/* The same solve in three steps, for callers that keep the loads and the
 * solution in device memory between calls (the benchmark's timed region holds
 * no host transfer): st_set_loads copies the loads to the device once;
 * st_solve_resident solves with them (warm_start: from the last solution);
 * st_get_solution copies the last solution back. */
int st_set_loads(StCtx* ctx, const double* loads);
int st_solve_resident(StCtx* ctx, const StSolveOpts* opts, StStats* stats);
int st_get_solution(StCtx* ctx, double* x_out);

/* No counterpart in the reference: st_joined_to_held (keep_gone = 1, the context's
 * current beam masks) for callers that cannot allocate device memory, such as code under
 * HIP graph capture or a tight serving loop. It allocates nothing, copies nothing and does
 * not wait: it only enqueues kernels on `stream` (a hipStream_t; NULL: the context's own
 * stream). joined_out: device memory, one byte per node (device order), 1 where a chain of
 * alive beams joins the node to a held node; count_out: device memory for one int32, the
 * number joined, or NULL. Its scratch (union-find parents and held-root flags) is a fixed
 * device array of the library, 32 MiB on each GPU it runs on, held until the process
 * ends, so structures of more than 4,194,304 nodes are refused (ST_BAD_INPUT). Every call
 * on one GPU shares that array: calls that overlap on one GPU (from several contexts or
 * streams) race, and the caller must order them. */
int st_joined_to_held_static(StCtx* ctx, void* stream, uint8_t* joined_out, int32_t* count_out);

/* Replaces: bodies::joined_to_held (statics/src/bodies.rs:557-574) when the
 * caller only needs `gone`: classifies on the device with the current beam
 * masks and sets gone = !joined without a host round trip. Writes the number
 * of nodes joined to a held node to *joined_count if not NULL. */
int st_classify(StCtx* ctx, int64_t* joined_count);

/* Device time of the launches of one kernel, for measuring it against its
 * roofline: enable != 0 starts recording (and clears the totals) for kernel
 * `which` (0: the fine operator K p of the CG with its fused p.Ap: FP64 under ST_FP64
 * and ST_MIXED, FP32 under ST_MIXED32); enable == 0 stops. st_kernel_timing reads the totals so far:
 * launches, summed milliseconds, and the algorithmic bytes of one launch
 * (6 values per free node read and written, plus e, g, the beam mask and the
 * six face indices per free cell). */
int st_kernel_timing_enable(StCtx* ctx, int which, int enable);
int st_kernel_timing(StCtx* ctx, int64_t* launches, double* total_ms, double* bytes_per_launch);

/* No counterpart in the reference: the schedule the autotuner chose for the fine
 * operator in FP64 (fp64 != 0) or FP32: its index (the value of the
 * STATICS_HIP_SCHEDULE environment variable that forces it), the name of its
 * kernel and a short description (static strings), and the number of kernel
 * dispatches one application takes (the cell-centric schedule launches its
 * small- and large-brick classes separately). Tunes first if the upload is not
 * tuned yet. Any output pointer may be NULL. */
int st_apply_schedule(StCtx* ctx, int fp64, int* index, const char** kernel, const char** description,
                      int* dispatches);

/* Replaces: bodies::moved (statics/src/bodies.rs:420-423) for every live bond:
 * per cell (in upload order) 15 beams x 4 bonds x 3 doubles, the force on
 * the j side of each bond, bonds in the reference's order
 * (statics/src/lattice.rs:76-84); 0 for dead beams. x: 6 per node. */
int st_bond_forces(StCtx* ctx, const double* x, double* forces_out);

/* Replaces: bodies::used (statics/src/bodies.rs:544-546) over the bonds of
 * bodies::moved, with Bond::used and Mat::used (statics/src/model.rs:123-129,
 * 66-68): per cell (in upload order) 15 beams x 4 bonds, the share of the
 * cell's strengths that the bond's force under motions x uses (> 1: the bond
 * breaks); 0 for dead beams and held cells. strength: per_beam = 0, 3 doubles per
 * cell, the material's ft, fc, fs (model.rs:43-58); per_beam != 0, 45 per cell, each
 * beam's own (beam q at 3 q), as lattices whose beams were made stronger one by one
 * hold them (bodies::as_generated, statics/src/bodies.rs:611-625; st_lattice_cells
 * gives them). x: 6 per node. */
int st_bond_used(StCtx* ctx, const double* strength, int per_beam, const double* x, double* used_out);

/* Replaces: bodies::settle (statics/src/bodies.rs:511-541) on the device. From
 * the context's beam masks (all alive after st_upload with beams NULL, as the
 * reference starts), rounds of: gone = !joined_to_held (bodies.rs:518-519); the
 * solve under `loads` (the reference's forces: gravity, bodies.rs:407-409,
 * given here as 6 doubles per node in device order); every live beam one of
 * whose bonds has used > 1 (strength and per_beam as st_bond_used) dies whole
 * (bodies.rs:523-533); until a round breaks nothing. opts->warm_start != 0 starts
 * each round after the first from the previous round's solution (the reference
 * starts from 0; the converged solution is the same). The reference settles at
 * tol 1e-6 (bodies.rs:432): rounds after breaks can leave systems whose FP64 true
 * residual floors above 1e-8. Outputs (each may be NULL):
 * gone_out per node (device order), beams_out the final 15-bit masks per cell
 * (upload order; 0 for held cells), rounds_out, and stats of the last round with
 * iters, setup_ms and solve_ms summed over the rounds. The fallen share
 * (bodies.rs:535-540) is the mass of the gone nodes over the free mass, which
 * the caller holds. Afterwards the context keeps the final masks and gone set,
 * and the last round's solution. stats->retries counts the rounds whose solve was
 * repeated with the robust cycle settings; stats->first_status is ST_NO_CONVERGENCE
 * when some round under the reference's rule (refresh < 0) ended with a true residual
 * above tol (its rel_residual is the worst such round's), else ST_OK. */
int st_settle(StCtx* ctx, const double* strength, int per_beam, const double* loads, const StSolveOpts* opts, uint8_t* gone_out,
              uint16_t* beams_out, int32_t* rounds_out, StStats* stats);

/* ---- gradients (no counterpart in the reference) ---------------------- */
typedef enum { ST_COMPLIANCE = 0, ST_SOFTMAX_STRESS = 1, ST_VALUE_ONLY = 256 } StObjective;
/* After st_solve or st_solve_resident: J and dJ/dtheta per cell (upload order; 0 for held
 * cells), where theta is the cell's log-stiffness (e and g both scale by
 * exp(theta)). ST_COMPLIANCE: J = f.x. ST_SOFTMAX_STRESS: J = (1/beta) log
 * sum_b exp(beta sigma_b) over live bonds b, sigma_b = (f_b . n_b)/a_b, the
 * bond's normal stress. The adjoint solve K lambda = dJ/dx reuses the operator (K is
 * symmetric) and runs with the precision and tolerance of opts. Its multigrid hierarchy
 * is the one the forward solve kept if that solve ran with release_hierarchy = 0 in the
 * same precision mode; otherwise (release_hierarchy = 1 on the forward solve, the
 * recompute-instead-of-store mode) the adjoint builds it anew and, with
 * opts->release_hierarchy, frees it again when done.
 * dJ_dtheta_out may be NULL (the gradient then stays on the device).
 * objective | ST_VALUE_ONLY: J alone, no adjoint solve, no gradient. */
int st_backward(StCtx* ctx, int objective, double beta, const StSolveOpts* opts, double* J_out,
                double* dJ_dtheta_out, StStats* stats);

/* ---- batches ---------------------------------------------------------- */
/* Many independent structures in one set of launches: every kernel covers
 * all structures' bricks; CG scalars are per structure and a structure that
 * has converged stops updating. Each structure is classified first (what is
 * not joined to a held node is gone), as forces does before its solve. cells[s] and loads[s] as for st_upload and
 * st_solve (loads may be NULL: gravity along -z of weight 1 per free cell,
 * a sixth on each face). x_out[s] may be NULL. stats: one per structure. */
int st_solve_batch(StCtx* ctx, int32_t n, const StCells* const* cells, const double* const* loads,
                   const StSolveOpts* opts, double* const* x_out, StStats* stats);

/* ---- synthetic structures (the workloads the benchmarks are quoted on) -- */
typedef struct {
  int32_t nx, ny, nz;      /* lattice extent; z is up */
  uint64_t seed;
  double pillar_spacing;   /* Poisson-disk radius between pillar centres, cells (8) */
  int32_t pillar_min, pillar_max; /* pillar width range, cells (2, 6) */
  int32_t slab_min, slab_max;     /* gap between slabs, cells (8, 16) */
  double hole_p;           /* probability a 2x2 tile of a slab is a hole (0.2) */
  double sigma;            /* log-stiffness standard deviation (0: unit stiffness) */
} StGenParams;
/* Fills occ (nx*ny*nz bytes, index (x*ny + y)*nz + z): 0 empty, 1 free, 2
 * held; theta (same size, may be NULL): log-stiffness per cell, N(0, sigma^2). */
int st_generate(const StGenParams* p, uint8_t* occ, float* theta);

#ifdef __cplusplus
}
#endif
#endif
