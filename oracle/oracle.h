/*
 * oracle.h - CPU restatement of the reference's Bacchus hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP path is
 * compared against: only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it.  The product never routes through it.
 *
 * Restated from (reference file:line):
 *   oracle_matvec             src/Minerva/scheduler/scheduler.c:80-145,
 *                             194-418 (dispatch, diag/off-diag, per-thread
 *                             output replicas src/Minerva/vector_block/
 *                             vector_block.c:47-67,146-155) and the six
 *                             kernels src/Minerva/matrix_vector_multiplication/
 *                             matrix_vector_multiplication.c:7-462
 *   oracle_lanczos            src/Bacchus/lanczos/lanczos.c:60-253 with the
 *                             block-wise vector ops of src/Bacchus/vector/
 *                             vector.c:240-476 and src/Bacchus/math_tools/
 *                             math_tools.c:13-57
 *   oracle_tridiagonal_eigenvalues
 *                             the eigenvalues dsteqr_ returns in
 *                             src/Bacchus/diagonalization/diagonalization.c:
 *                             16-68 (LAPACK is not vendored; restated as
 *                             Sturm-sequence bisection, see oracle/README)
 */
#ifndef ORACLE_H
#define ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* instructions: 8 int32 per row {type, rank, block_in, block_out,
 * neutron_list, proton_list, matrix_block, pad}; type 1 N, 2 P, 3 NP.
 * matrix_force[v] in {1,2,3} selects scales[force-1] (NULL = 1.0).
 * y += H x, like a Bacchus matvec into pre-zeroed files. */
int oracle_matvec(int64_t num_blocks, const int64_t *block_np,
		  const int64_t *block_nn, const int64_t *list_offsets,
		  const int32_t *list_entries, const int64_t *matrix_offsets,
		  const int64_t *matrix_dn, const int32_t *matrix_force,
		  const double *scales, const double *matrix_elements,
		  int64_t num_instructions, const int32_t *instructions,
		  const double *x, double *y, int num_threads);

/* Eigenvalues of the symmetric tridiagonal matrix (diag[0..n-1],
 * off[0..n-2]), ascending, into out[0..count-1] (the lowest count). */
void oracle_tridiagonal_eigenvalues(int64_t n, const double *diag,
				    const double *off, int64_t count,
				    double *out);

enum { ORACLE_START_E0 = 0, ORACLE_START_GIVEN = 2 };
enum { ORACLE_REORTH_REFERENCE = 0, ORACLE_REORTH_FULL = 1,
       ORACLE_REORTH_NONE = 2 };
enum { ORACLE_CONVERGE_EIGENVALUES = 0, ORACLE_CONVERGE_NONE = 2 };

/* Lanczos with a callback matvec y = H x (y zeroed by the caller's
 * callback contract: oracle_lanczos passes a zero y, as
 * basis_append_vector creates a zero vector, src/Bacchus/basis/basis.c:
 * 39-53).  Returns the Krylov dimension K; alpha/beta get K values (beta
 * K-1 meaningful).  eigenvalues gets min(K, num_eigenvalues) values. */
typedef void (*oracle_matvec_fn)(void *ctx, const double *x, double *y);
int64_t oracle_lanczos(int64_t dimension, oracle_matvec_fn matvec, void *ctx,
		       int64_t max_num_iterations, int start,
		       const double *start_vector, int reorthogonalization,
		       int convergence, double tolerance,
		       int64_t num_eigenvalues, double *eigenvalues,
		       double *alpha, double *beta);

#ifdef __cplusplus
}
#endif
#endif
