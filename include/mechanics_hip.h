/* mechanics_hip: the reference's mechanics::solve::Solver on one MI355X (gfx950).
 *
 * A symmetric system over the corner nodes of co-rotated cubes, three unknowns to a node
 * and a 3x3 block to each pair joined (mechanics/src/solve.rs:45-72), solved by conjugate
 * gradients preconditioned by V-cycles over the octree: a level up, a part is the finer
 * parts of one octree cell that are joined, moving as a rigid body (six unknowns), its
 * blocks the Galerkin lifts of its members' (solve.rs:1-13, 253-301), until a few parts are
 * solved directly (solve.rs:343-372). Setup and iterations run on the device in FP64; the
 * host does the bookkeeping of the levels (which parts, which pairs) and factors the top.
 *
 * Status codes are statics_hip's (StStatus). A solver belongs to one host thread at a
 * time. */
#ifndef MECHANICS_HIP_H
#define MECHANICS_HIP_H
#include <stdint.h>

#include "statics_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct MeSolver MeSolver;

/* mechanics::solve::System (solve.rs:49-59) as flat arrays. */
typedef struct {
  int64_t n;              /* unknowns (< 2^31) */
  const int64_t* sites;   /* 3 per unknown: its octree key, the node's place over the cells' size
                             (mechanics/src/lib.rs:640); read by me_solver_new only */
  const double* at;       /* 3 per unknown: where it is, metres; read by me_solver_new only */
  const double* diag;     /* 9 per unknown: its block, row-major */
  int64_t npairs;
  const uint32_t* pairs;  /* 2 per pair: i, j (i != j, both < n) */
  const double* blocks;   /* 9 per pair: the block in row i, column j, row-major; its transpose
                             is the block in row j, column i; a pair twice is the sum */
} MeSystem;

/* Replaces: Solver::new (mechanics/src/solve.rs:387-420): the inverses of the diagonal
 * blocks, the smoothing step (solve.rs:309-327), and the levels up to the direct one, for
 * cells `size` metres across, on device `device`. */
int me_solver_new(MeSolver** out, const MeSystem* system, double size, int device);

/* Replaces: Solver::solve (solve.rs:508-537): solves `system` (the one the solver was made
 * for, or one close to it: the same unknowns, other values or other pairs) for x, from
 * what x holds, till r^T P r <= tolerance^2 b^T P b with P the V-cycle (at most 300
 * rounds). *rounds: the rounds taken, or -1 where the reference returns None (the levels
 * do not fit the system: a negative or NaN measure), x then no farther than it was.
 * b, x: 3 per unknown, host memory. */
int me_solve(MeSolver* s, const MeSystem* system, const double* b, double* x, double tolerance, int32_t* rounds);

int me_solver_free(MeSolver* s);
/* The last error's text (static storage, overwritten by the next call on this solver). */
const char* me_last_error(const MeSolver* s);

/* No counterpart in the reference: the levels made (count; per level up to cap: parts,
 * pairs, smoothing step, whether solved directly), the finest level's step, and the
 * wall-clock time of the last setup and solve, milliseconds (host bookkeeping and
 * copies included). Any pointer may be NULL. */
int me_solver_info(const MeSolver* s, int32_t cap, int32_t* count, int64_t* parts, int64_t* pairs, double* steps,
                   int32_t* direct, double* step0, double* setup_ms, double* solve_ms);

/* No counterpart in the reference: device time of the CG's K p (the level-0 product of
 * me_solve, one launch per round) for measuring it against its roofline. enable 1 starts
 * recording (clearing the totals), 0 stops, -1 only reads: launches and summed
 * milliseconds so far, and the algorithmic bytes of one launch (the diagonal blocks, row
 * pointers, column indices and pair blocks read once, x read once, the product written). */
int me_kernel_timing(MeSolver* s, int enable, int64_t* launches, double* total_ms, double* bytes_per_launch);

#ifdef __cplusplus
}
#endif
#endif
