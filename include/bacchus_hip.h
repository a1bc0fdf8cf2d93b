/*
 * bacchus_hip.h - C ABI of the MI355X (gfx950) Bacchus hot path: the
 * index-list driven block Hamiltonian x vector product (Minerva) and the
 * Lanczos recurrence around it.
 *
 * Plain pointers and sizes only.  Every function returns 0 on success and a
 * nonzero code on failure; bacchus_hip_last_error() then holds the message
 * ("Error in <function> @ <file>:<line>: ..." like the reference's error()
 * macro, src/Utilities/error/error.h:7-13).  The file-based entry points
 * keep the reference's inputs unchanged: comb.txt, the BLOCK: evaluation
 * order, index_list_<id>, <id>_matrix_elements and vec_<id> directories.
 */
#ifndef BACCHUS_HIP_H
#define BACCHUS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Storage precision of the matrix elements on the device. */
enum bacchus_hip_precision {
	BACCHUS_HIP_FP64 = 0, /* FP64 elements, FP64 accumulation            */
	BACCHUS_HIP_FP32 = 1, /* FP32 elements, FP32 products and accumulation */
	BACCHUS_HIP_BF16 = 2  /* BF16 elements, FP32 products and accumulation */
};

/* One evaluation-order instruction with dense 0-based indices
 * (src/Minerva/evaluation_order/evaluation_order.h:22-31).
 * type: 1 = neutron, 2 = proton, 3 = neutron-proton block
 * (evaluation_order.h:13-20); -1 marks an unused list. */
typedef struct bacchus_hip_instruction {
	int32_t type;
	int32_t num_particles;
	int32_t block_in;
	int32_t block_out;
	int32_t neutron_list;
	int32_t proton_list;
	int32_t matrix_block;
	int32_t pad;
} bacchus_hip_instruction_t;

/* Host-side description of an operator, the in-memory form of the files a
 * scheduler reads (src/Minerva/memory_manager/memory_manager.c:455-509). */
typedef struct bacchus_hip_problem {
	int64_t num_blocks;
	const int64_t *block_np;         /* proton states per basis block   */
	const int64_t *block_nn;         /* neutron states per basis block  */
	int64_t num_lists;
	const int64_t *list_offsets;     /* num_lists+1, in triples         */
	const int32_t *list_entries;     /* {in, out, matrix} triples       */
	int64_t num_matrix_blocks;
	const int64_t *matrix_offsets;   /* num_matrix_blocks+1, elements   */
	const int64_t *matrix_dn;        /* neutron matrix dimension        */
	const int32_t *matrix_force;     /* 1, 2 or 3 body (scale class)    */
	const double *matrix_elements;
	int64_t num_instructions;
	const bacchus_hip_instruction_t *instructions;
} bacchus_hip_problem_t;

typedef struct bacchus_hip_operator bacchus_hip_operator_t;

/* Validate the problem on the host (every index in bounds), build the
 * device tables and upload everything to HBM on the current device.  The
 * tables are built on up to 16 host threads that end before the call
 * returns; the problem's arrays are only read.  matrix_force may be NULL
 * (every block then counts as 2-body for the scales). */
int bacchus_hip_operator_create(const bacchus_hip_problem_t *problem,
				int precision, bacchus_hip_operator_t **op);

/* File-based constructor: replaces new_combination_table
 * (src/Utilities/combination_table/combination_table.c:83-137) +
 * read_evaluation_order (src/Minerva/evaluation_order/evaluation_order.c:
 * 43-112) + new_scheduler (src/Minerva/scheduler/scheduler.c:62-78).
 * Reads comb.txt, the evaluation order, index_list_<id> and
 * <id>_matrix_elements files and uploads them once. */
int bacchus_hip_operator_load(const char *combination_table_file,
			      int num_protons, int num_neutrons,
			      const char *evaluation_order_file,
			      const char *index_lists_base_directory,
			      const char *matrix_file_base_directory,
			      int precision, bacchus_hip_operator_t **op);

/* One evaluation-order instruction as the reference's parser hands it over
 * (evaluation_instruction_t, src/Minerva/evaluation_order/
 * evaluation_order.h:22-31): basis blocks, index lists and the element
 * block by their comb.txt ids; a list id of 0 means "none".  type: 1 N,
 * 2 P, 3 NP.  num_particles: 1, 2 or 3 (the force class of the element
 * block for the scales), or 0 when unknown (2-body). */
typedef struct bacchus_hip_file_instruction {
	int32_t type;
	int32_t num_particles;
	int64_t block_in;
	int64_t block_out;
	int64_t neutron_list;
	int64_t proton_list;
	int64_t matrix_block;
} bacchus_hip_file_instruction_t;

/* Constructor from already-parsed tables: replaces new_scheduler
 * (src/Minerva/scheduler/scheduler.c:62-78, scheduler.h:10-14) for a caller
 * that holds the reference's combination table and evaluation order.  The
 * basis blocks are given in vector order (comb.txt order) with their ids;
 * index_list_<id> and <id>_matrix_elements are read from the two
 * directories.  maximum_loaded_memory (bytes; 0 = none) is checked as the
 * reference's memory manager checks it: an instruction whose vector
 * blocks, element block and lists exceed it is an error
 * (src/Minerva/memory_manager/memory_manager.c:354-365); the operator is
 * kept resident in device memory, so one whose device tables exceed it is
 * an error too (there is no out-of-core mode). */
int bacchus_hip_operator_load_blocks(int64_t num_blocks, const int64_t *block_ids,
				     const int64_t *block_np, const int64_t *block_nn,
				     int64_t num_instructions,
				     const bacchus_hip_file_instruction_t *instructions,
				     const char *index_lists_base_directory,
				     const char *matrix_file_base_directory,
				     size_t maximum_loaded_memory, int precision,
				     bacchus_hip_operator_t **op);

void bacchus_hip_operator_free(bacchus_hip_operator_t *op);

int64_t bacchus_hip_operator_dimension(const bacchus_hip_operator_t *op);

/* Evaluation-order generator: replaces the Mars program
 * (src/Mars/programs/mars.c:64-199).  Reads the CALCBLOCK rows of comb.txt
 * and writes the BLOCK: rows in Mars's order (1-body blocks dropped, and
 * the 3-body ones when two_nucleon_forces_only != 0, as its
 * --only-two-nucleon-forces option). */
int bacchus_hip_write_evaluation_order(const char *combination_table_file,
				       int two_nucleon_forces_only,
				       const char *output_path);

/* Scale factors per force class (1-, 2-, 3-body); default 1, 1, 1.
 * H(c1, c2) = H1 + c1*V2 + c2*V3 without rewriting element blocks
 * (the linear combination src/Vulcan/programs/vulcan.c:114-207 writes). */
int bacchus_hip_operator_set_scales(bacchus_hip_operator_t *op, double s1,
				    double s2, double s3);

/* y = H x (accumulate = 0) or y += H x (accumulate = 1) on device memory,
 * D doubles each, on the given hipStream_t (NULL = default stream).
 * Device-resident form of run_matrix_vector_multiplication
 * (src/Minerva/scheduler/scheduler.c:80-145). */
int bacchus_hip_operator_matvec(bacchus_hip_operator_t *op, const double *x,
				double *y, int accumulate, void *stream);

/* Algorithmic traffic and work of one matvec: bytes = s_V * sum|V| +
 * 12 * sum|L| + 16 * D, flops = 2 * multiplications (off-diagonal
 * instructions count twice). */
int bacchus_hip_operator_stats(const bacchus_hip_operator_t *op,
			       double *algorithmic_bytes, double *flops,
			       int64_t *num_work_items);

/* Drop-in for run_matrix_vector_multiplication
 * (src/Minerva/scheduler/scheduler.c:80-145): reads <in>/vec_<id>, adds
 * H x to the existing <out>/vec_<id> files (the reference loads the output
 * files and accumulates, src/Minerva/vector_block/vector_block.c:47-67)
 * and writes them back. */
int bacchus_hip_run_matrix_vector_multiplication(
	const char *output_vector_base_directory,
	const char *input_vector_base_directory, bacchus_hip_operator_t *op);

/* Lanczos on the device (src/Bacchus/lanczos/lanczos.c:60-253). */
enum bacchus_hip_start { BACCHUS_HIP_START_E0 = 0,     /* lanczos.c:60-72 */
			 BACCHUS_HIP_START_RANDOM = 1, /* seeded, normalized */
			 BACCHUS_HIP_START_GIVEN = 2 };
enum bacchus_hip_reorth { BACCHUS_HIP_REORTH_REFERENCE = 0, /* MGS vs
							      u_0..u_{k-3},
							      lanczos.c:155-169 */
			  BACCHUS_HIP_REORTH_FULL = 1,     /* CGS2 vs all */
			  BACCHUS_HIP_REORTH_NONE = 2 };
enum bacchus_hip_convergence { BACCHUS_HIP_CONVERGE_EIGENVALUES = 0,
			       /* |a - a'| of the target eigenvector's Krylov
				  amplitudes between iterations (lanczos.c:
				  276-287 difference_eigenvectors) */
			       BACCHUS_HIP_CONVERGE_EIGENVECTORS = 1,
			       BACCHUS_HIP_CONVERGE_NONE = 2,
			       BACCHUS_HIP_CONVERGE_TARGETS = 3,
			       /* block solver: |r_k|^2 / gap_k <= tol |theta_k|
				  for the lowest num_eigenvalues Ritz pairs, with
				  gap_k = min_{i != k} |theta_i - theta_k| - |r_i|
				  (or |r_k| <= tol |theta_k| where that gap is
				  not positive) */
			       BACCHUS_HIP_CONVERGE_RESIDUALS = 4 };

typedef struct bacchus_hip_lanczos_settings {
	int64_t max_num_iterations;
	int32_t start;            /* enum bacchus_hip_start            */
	int32_t reorthogonalization; /* enum bacchus_hip_reorth         */
	int32_t convergence;      /* enum bacchus_hip_convergence      */
	int32_t num_eigenvalues;  /* lowest eigenvalues reported       */
	double eigenvalue_tolerance;
	uint64_t seed;            /* random start                       */
	const double *start_vector; /* device pointer, START_GIVEN     */
	const double *target_eigenvalues; /* CONVERGE_TARGETS: stop when the
					     lowest num_eigenvalues are within
					     eigenvalue_tolerance (relative) */
	int32_t check_every;      /* iterations between eigenvalue checks */
	int32_t target_eigenvalue; /* index (ascending) of the eigenvalue the
				      EIGENVALUES / EIGENVECTORS criteria
				      watch (lanczos.h:23, bacchus.c:55) */
	int32_t num_ritz_vectors; /* Ritz vectors written to ritz_vector
				     (0 = 1)                             */
	int32_t pad;
	double *krylov_basis;     /* NULL, or host memory for
				     min(D, max_num_iterations) x D doubles
				     receiving the Krylov vectors u_0 ..
				     u_{k-1} of the solve (the basis
				     get_eigensystem combines, lanczos.c:
				     255-265)                            */
} bacchus_hip_lanczos_settings_t;

typedef struct bacchus_hip_lanczos_result {
	int64_t iterations;       /* Krylov dimension reached            */
	double matvec_seconds;    /* device time in the matvec            */
	double total_seconds;     /* wall time of the solve               */
	int32_t converged;        /* 1: the convergence criterion stopped
				     the solve; 0: the iteration (block)
				     limit did                            */
	int32_t pad;
} bacchus_hip_lanczos_result_t;

/* Runs Lanczos; writes the lowest num_eigenvalues Ritz values to
 * eigenvalues (host), alpha/beta (host, max_num_iterations each, may be
 * NULL), and, if ritz_vector is non-NULL (device, num_ritz_vectors x D
 * doubles), the Ritz vectors of the lowest num_ritz_vectors eigenvalues
 * (ascending, unit norm, Krylov amplitudes signed so that the first is
 * non-negative as diagonalization.c:55-58 does). */
int bacchus_hip_lanczos(bacchus_hip_operator_t *op,
			const bacchus_hip_lanczos_settings_t *settings,
			double *eigenvalues, double *alpha, double *beta,
			double *ritz_vector,
			bacchus_hip_lanczos_result_t *result);

/* Batched form (no counterpart in the reference, where a parameter sweep is
 * one Bacchus job per Hamiltonian after a Vulcan linear combination,
 * src/Vulcan/programs/vulcan.c:114-207).  A batch holds several operators
 * (basis sizes may differ) and B_i vectors per operator; vector b of
 * operator i is multiplied by H_i(s) = s1 H1 + s2 V2 + s3 V3 with its own
 * scales, the element blocks being shared by the whole batch.  One
 * persistent launch covers every operator.  Operators must be FP32 or BF16
 * (FP32 vectors, FP32 accumulation); they must outlive the batch.
 * Vectors are device FP32 arrays [D_i][B_i], the vector index fastest. */
typedef struct bacchus_hip_batch bacchus_hip_batch_t;

typedef struct bacchus_hip_batch_tuning {
	int32_t tiles_per_cu;     /* tile size: total work / (CUs * this); 0 = 16 */
	int32_t max_tile_vectors; /* vectors per tile, power of two <= 64; 0 = 64 */
	int32_t workgroups_per_cu; /* persistent grid; 0 = 8                    */
	int32_t tile_words;       /* FP64 LDS sums per tile, <= 8192; 0 = 4096   */
	int32_t waves_per_simd;   /* 4: kernel held to 128 VGPRs (four waves per
				     SIMD); 0 = the compiler's register count */
} bacchus_hip_batch_tuning_t;

int bacchus_hip_batch_create(bacchus_hip_operator_t *const *ops,
			     const int32_t *num_vectors, int32_t num_ops,
			     const bacchus_hip_batch_tuning_t *tuning,
			     bacchus_hip_batch_t **batch);
void bacchus_hip_batch_free(bacchus_hip_batch_t *batch);

/* scales: num_vectors[op_index] x 3 doubles (s1, s2, s3 per vector) */
int bacchus_hip_batch_set_scales(bacchus_hip_batch_t *batch, int32_t op_index,
				 const double *scales);

/* y_i = H_i x_i for every operator (x, y: num_ops device pointers; y is
 * overwritten).  Pointers that change between calls are re-uploaded. */
int bacchus_hip_batch_matvec(bacchus_hip_batch_t *batch, const float *const *x,
			     float *const *y, void *stream);

/* Times candidate tilings (repetitions matvecs each, on internal buffers),
 * then the best of them with the other waves_per_simd, and keeps the
 * fastest; writes it and its time per batched matvec. */
int bacchus_hip_batch_autotune(bacchus_hip_batch_t *batch, int32_t repetitions,
			       bacchus_hip_batch_tuning_t *best, double *best_ms);

/* bytes = sum_i (s_V |V_i| + 12 |L_i| + 8 D_i B_i); flops = 2 sum_i
 * mults_i B_i; num_items = work items of the persistent launch */
int bacchus_hip_batch_stats(const bacchus_hip_batch_t *batch, double *bytes,
			    double *flops, int64_t *num_items);

/* The matrix-core form of the batch's P-form instructions (P, and N in the
 * transposed [neutron][proton] space; no counterpart in the reference):
 * each is the product of its list, as a matrix with the summed signed
 * elements, with a dense panel of (states x vectors) columns, stored as
 * 16 x 4 tiles and run as v_mfma_f32_16x16x4f32 after the tile launch.
 * Off when BACCHUS_HIP_BATCH_MFMA=0 is set at batch creation; instructions
 * with a nonzero element below 2e-9 in magnitude, and operators while a
 * scale is nonzero and below 1e-3 in magnitude, stay on the vector path, so
 * the 1e-12 element skip of matrix_vector_multiplication.c:50-53 and
 * :105-108 never applies to a matrix-core product.  kgroups = 16 x 4 tiles;
 * mfma_per_matvec = MFMA instructions of one batched matvec with every
 * operator on; fill = list entries / (64 kgroups); operators_on =
 * operators whose P-form instructions run on the matrix cores now. */
int bacchus_hip_batch_mfma_stats(const bacchus_hip_batch_t *batch, int64_t *kgroups,
				 double *mfma_per_matvec, double *fill,
				 int32_t *operators_on);

/* One Lanczos recurrence per (operator, vector) in lockstep: full CGS2
 * reorthogonalization and all dot products in FP64, the matvec batched.
 * Each iteration is one captured hipGraph when use_graph != 0. */
typedef struct bacchus_hip_batch_lanczos_settings {
	int64_t max_num_iterations;
	int32_t num_eigenvalues;   /* lowest eigenvalues per instance        */
	int32_t use_graph;
	double eigenvalue_tolerance; /* 0: run max_num_iterations; else stop
					when every instance's lowest
					num_eigenvalues moved less than this
					(relative) between checks */
	int32_t check_every;
	int32_t num_ritz_vectors;  /* Ritz vectors per instance (0 = 1)      */
	uint64_t seed;             /* random starts                          */
	const double *const *start_vectors; /* NULL or per operator NULL/device
					       FP64 [D_i][B_i]             */
} bacchus_hip_batch_lanczos_settings_t;

/* eigenvalues: host, sum_i B_i * num_eigenvalues, operator-major then
 * vector then ascending eigenvalue; ritz_vectors: NULL or per operator
 * NULL/device FP64 [num_ritz_vectors][D_i][B_i] receiving each instance's
 * lowest Ritz vectors (unit norm, ascending eigenvalue). */
int bacchus_hip_batch_lanczos(bacchus_hip_batch_t *batch,
			      const bacchus_hip_batch_lanczos_settings_t *settings,
			      double *eigenvalues, double *const *ritz_vectors,
			      bacchus_hip_lanczos_result_t *result);

/* Block Lanczos over a batch of one operator (no counterpart in the
 * reference, whose Lanczos is single-vector, src/Bacchus/lanczos/
 * lanczos.c:74-253).  The batch's b vectors form the block (1 <= b <= 16);
 * every vector of the batch must carry the same scales.  One batched
 * launch per block step; Krylov blocks stored FP32 [D][b], inner products
 * in FP64, classical block Gram-Schmidt against all blocks twice (once in
 * the step that converges), block
 * tridiagonal eigensystem in FP64 on the host.  Krylov dimension
 * b * max_num_blocks <= 256.  The batch keeps the solver's device workspace
 * between calls (freed with the batch); calls on one batch must not
 * overlap. */
typedef struct bacchus_hip_block_lanczos_settings {
	int32_t max_num_blocks;
	int32_t num_eigenvalues;  /* 1..b                                      */
	int32_t convergence;      /* TARGETS, RESIDUALS or NONE                */
	int32_t pad;
	double tolerance;         /* relative                                  */
	const double *target_eigenvalues; /* CONVERGE_TARGETS                  */
	const float *start_block; /* device FP32 [D][b]; NULL = seeded random  */
	uint64_t seed;
} bacchus_hip_block_lanczos_settings_t;

/* eigenvalues, residuals: host, num_eigenvalues each (residual = |H y - theta
 * y| of the Ritz pair, from the recurrence); ritz_vectors: NULL or device
 * FP64 [num_eigenvalues][D]; result->iterations = block steps taken. */
int bacchus_hip_block_lanczos(bacchus_hip_batch_t *batch,
			      const bacchus_hip_block_lanczos_settings_t *settings,
			      double *eigenvalues, double *residuals,
			      double *ritz_vectors,
			      bacchus_hip_lanczos_result_t *result);

/* Dense symmetric eigensystem on the host (cyclic Jacobi, FP64), the
 * solver of the block tridiagonal matrix: A n x n row-major (symmetrized),
 * w ascending, V (may be NULL) eigenvector k in column k.  No GPU needed. */
int bacchus_hip_symmetric_eigen(int64_t n, const double *A, double *w, double *V);

/* Ceres's projections onto n stored vectors, in FP64 on the device.  V
 * holds the vectors as device FP64 rows ([n][D], row i = v_i); the result is
 * a host n x n row-major matrix, element [i][j] for j >= i mirrored into
 * [j][i] as the reference fills it; on `stream` (NULL: the default stream),
 * returning once the matrix is on the host.
 *   norm_matrix replaces create_norm_matrix
 *     (src/Ceres/norm_matrix/norm_matrix.c:4-28): N[i][j] = v_i . v_j
 *   subspace_operator replaces create_subspace_operator
 *     (src/Ceres/subspace_operator/subspace_operator.c:7-61):
 *     M[i][j] = v_j . (H v_i), one matvec of op per vector (any storage
 *     precision; the scales set on op apply) */
int bacchus_hip_norm_matrix(const double *V, int64_t n, int64_t D, double *N, void *stream);
int bacchus_hip_subspace_operator(bacchus_hip_operator_t *op, const double *V, int64_t n,
				  double *M, void *stream);

/* Stored-basis helpers of the learned start (no counterpart in the
 * reference).  U holds n FP32 vectors of dimension D ([n][D]).
 *   gram:      G[i][j] = U_i . U_j                  (host, n x n, FP64)
 *   project:   out[i][b] = U_i . V[:, b]            (host, n x B, FP64;
 *              V device FP64 [D][B])
 *   combine:   out = sum_i c_i U_i / |sum_i c_i U_i| (c device FP32, n;
 *              out device FP64, D) */
int bacchus_hip_basis_gram(const float *U, int64_t n, int64_t D, double *G);
int bacchus_hip_basis_project(const float *U, int64_t n, int64_t D,
			      const double *V, int64_t B, double *out);
/* The same projection with no device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * graph capture, serving loops, harnesses that forbid hipMalloc).  Its
 * partial sums live in a static device array of 2^18 doubles per GPU,
 * held until the process ends, so n * B <= 2^18 / 65 = 4032 (larger sizes
 * are rejected); every call on a GPU uses that GPU's array, so calls that
 * overlap on one GPU race.  Results equal bacchus_hip_basis_project's. */
int bacchus_hip_basis_project_static(const float *U, int64_t n, int64_t D,
				     const double *V, int64_t B, double *out);
int bacchus_hip_basis_combine(const float *U, int64_t n, int64_t D,
			      const float *c, double *out, void *stream);
/*   transform: out[j] = sum_i C[j][i] U_i for j < r (C host FP64 r x n;
 *              out device FP32 [r][D]) - e.g. an orthonormalized basis; C
 *              goes through a device buffer the library keeps per device,
 *              and calls are serialized; synchronous */
int bacchus_hip_basis_transform(const float *U, int64_t n, int64_t D,
				const double *C, int64_t r, float *out);

/* Learned Lanczos start (no counterpart in the reference): an MLP
 * in -> hidden (x depth, GELU) -> out with BF16 weights on MFMA, FP32
 * accumulation and FP32 master weights trained with Adam on the device.
 * Parameters are all weight matrices (out x in, row-major, layer order)
 * followed by all biases.  x, t, y are device FP32 arrays, row-major
 * (M rows). */
typedef struct bacchus_hip_mlp bacchus_hip_mlp_t;
int bacchus_hip_mlp_create(int in, int hidden, int depth, int out,
			   uint64_t seed, bacchus_hip_mlp_t **mlp);
void bacchus_hip_mlp_free(bacchus_hip_mlp_t *mlp);
int64_t bacchus_hip_mlp_num_params(const bacchus_hip_mlp_t *mlp);
int bacchus_hip_mlp_get_params(const bacchus_hip_mlp_t *mlp, float *host);
int bacchus_hip_mlp_set_params(bacchus_hip_mlp_t *mlp, const float *host);
int bacchus_hip_mlp_forward(bacchus_hip_mlp_t *mlp, const float *x, int M,
			    float *y, void *stream);
/* gradients of the mean squared error against t (host copy) */
int bacchus_hip_mlp_gradients(bacchus_hip_mlp_t *mlp, const float *x,
			      const float *t, int M, float *grads_host,
			      float *loss_host, void *stream);
int bacchus_hip_mlp_train_step(bacchus_hip_mlp_t *mlp, const float *x,
			       const float *t, int M, float lr,
			       float *loss_host, void *stream);

/* Precision of the MLP's GEMMs (default 0: every operand rounded to BF16).
 * BACCHUS_HIP_MLP_SPLIT_ACTIVATIONS: inputs, activations and backward
 * operands other than the weights are carried as two BF16 parts
 * (v = hi + lo, 16 significant bits; one more MFMA per tile), the weights
 * staying BF16.  BACCHUS_HIP_MLP_FP32_LAST: the output layer's FP32 master
 * weights are used, as two BF16 parts. */
#define BACCHUS_HIP_MLP_SPLIT_ACTIVATIONS 1
#define BACCHUS_HIP_MLP_FP32_LAST 2
int bacchus_hip_mlp_set_precision(bacchus_hip_mlp_t *mlp, int32_t flags);

const char *bacchus_hip_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
