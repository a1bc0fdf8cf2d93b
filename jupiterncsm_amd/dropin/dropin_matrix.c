/*
 * dropin_matrix.c - matrix.h (src/Bacchus/matrix/matrix.h) on the MI355X:
 * replaces src/Bacchus/matrix/matrix.c.
 *
 * The same two kinds of matrix: explicit (dense, the reference's test
 * backend, delegated unchanged to src/Bacchus/explicit_matrix/
 * explicit_matrix.c) and generative (new_generative_matrix -> the device
 * scheduler of dropin_scheduler.c).  The one addition is the device
 * operator that dropin_lanczos.c runs the recurrence on: the scheduler's,
 * or, for an explicit matrix, its elements uploaded once as a one-block
 * operator (1 x D states, one neutron-proton instruction whose element
 * block is the dense matrix, so every element counts as the reference's
 * explicit product counts it, with no 1e-12 skip).
 */
#include <error/error.h>
#include <explicit_matrix/explicit_matrix.h>
#include <matrix/matrix.h>
#include <scheduler/scheduler.h>

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "dropin.h"

typedef enum
{
	DROPIN_EXPLICIT_MATRIX,
	DROPIN_GENERATIVE_MATRIX
} dropin_matrix_type_t;

struct _matrix_
{
	dropin_matrix_type_t type;
	explicit_matrix_t explicit_matrix;
	scheduler_t scheduler;
	bacchus_hip_operator_t *dense; /* explicit matrices, on first use */
};

static matrix_t new_explicit(explicit_matrix_t e)
{
	matrix_t matrix = (matrix_t)calloc(1, sizeof(struct _matrix_));
	if (!matrix)
		error("out of host memory\n");
	matrix->type = DROPIN_EXPLICIT_MATRIX;
	matrix->explicit_matrix = e;
	return matrix;
}

matrix_t new_zero_matrix(size_t num_rows, size_t num_columns)
{
	return new_explicit(new_zero_explicit_matrix(num_rows, num_columns));
}

matrix_t new_random_symmetric_matrix(size_t side_length)
{
	return new_explicit(new_random_symmetric_explicit_matrix(side_length));
}

matrix_t new_matrix_from_numpy(FILE *matrix_file)
{
	return new_explicit(new_explicit_matrix_from_numpy(matrix_file));
}

matrix_t new_generative_matrix(evaluation_order_t evaluation_order,
			       combination_table_t combination_table,
			       const char *index_lists_base_directory,
			       const char *matrix_file_base_directory,
			       size_t maximum_loaded_memory)
{
	matrix_t matrix = (matrix_t)calloc(1, sizeof(struct _matrix_));
	if (!matrix)
		error("out of host memory\n");
	matrix->type = DROPIN_GENERATIVE_MATRIX;
	matrix->scheduler = new_scheduler(evaluation_order, combination_table,
					  index_lists_base_directory,
					  matrix_file_base_directory, maximum_loaded_memory);
	return matrix;
}

void matrix_vector_multiplication(vector_t result_vector, const matrix_t matrix,
				  const vector_t vector)
{
	printf("Matrix vector multiplication start:\n");
	struct timespec t_start, t_end;
	clock_gettime(CLOCK_REALTIME, &t_start);
	switch (matrix->type) {
	case DROPIN_EXPLICIT_MATRIX:
		explicit_matrix_vector_multiplication(result_vector, matrix->explicit_matrix,
						      vector);
		break;
	case DROPIN_GENERATIVE_MATRIX:
		run_matrix_vector_multiplication(get_vector_path(result_vector),
						 get_vector_path(vector), matrix->scheduler);
		break;
	}
	clock_gettime(CLOCK_REALTIME, &t_end);
	printf("Matrix vector multiplication end after %lg µs\n",
	       (t_end.tv_sec - t_start.tv_sec) * 1e6 + (t_end.tv_nsec - t_start.tv_nsec) * 1e-3);
}

static explicit_matrix_t explicit_of(matrix_t matrix, const char *what)
{
	if (matrix->type != DROPIN_EXPLICIT_MATRIX)
		error("%s needs an explicit matrix\n", what);
	return matrix->explicit_matrix;
}

size_t get_num_rows(matrix_t matrix)
{
	return get_explicit_matrix_num_rows(explicit_of(matrix, "get_num_rows"));
}

size_t get_num_columns(matrix_t matrix)
{
	return get_explicit_matrix_num_columns(explicit_of(matrix, "get_num_columns"));
}

/* a changed explicit matrix is uploaded again on the next device use */
static void drop_dense(matrix_t matrix)
{
	bacchus_hip_operator_free(matrix->dense);
	matrix->dense = NULL;
}

void set_matrix_elements(matrix_t matrix, const double *elements)
{
	set_explicit_matrix_elements(explicit_of(matrix, "set_matrix_elements"), elements);
	drop_dense(matrix);
}

void set_matrix_element(matrix_t matrix, size_t i, size_t j, double element)
{
	set_explicit_matrix_element(explicit_of(matrix, "set_matrix_element"), i, j, element);
	drop_dense(matrix);
}

double *get_matrix_elements(matrix_t matrix)
{
	/* the caller may write through the pointer */
	drop_dense(matrix);
	return get_explicit_matrix_elements(explicit_of(matrix, "get_matrix_elements"));
}

void save_numpy_matrix(FILE *file, const matrix_t matrix)
{
	save_numpy_explicit_matrix(file, explicit_of(matrix, "save_numpy_matrix"));
}

equality_status_t compare_matrices(const matrix_t first_matrix, const matrix_t second_matrix)
{
	if (first_matrix->type != second_matrix->type)
		return NOT_EQUAL;
	if (first_matrix->type == DROPIN_EXPLICIT_MATRIX)
		return compare_explicit_matrices(first_matrix->explicit_matrix,
						 second_matrix->explicit_matrix);
	return first_matrix->scheduler == second_matrix->scheduler ? EQUAL : NOT_EQUAL;
}

void free_matrix(matrix_t matrix)
{
	if (!matrix)
		return;
	if (matrix->type == DROPIN_EXPLICIT_MATRIX)
		free_explicit_matrix(matrix->explicit_matrix);
	else
		free_scheduler(matrix->scheduler);
	bacchus_hip_operator_free(matrix->dense);
	free(matrix);
}

bacchus_hip_operator_t *dropin_matrix_operator(matrix_t matrix)
{
	if (matrix->type == DROPIN_GENERATIVE_MATRIX)
		return dropin_scheduler_operator(matrix->scheduler);
	if (matrix->dense)
		return matrix->dense;
	const size_t rows = get_explicit_matrix_num_rows(matrix->explicit_matrix);
	const size_t cols = get_explicit_matrix_num_columns(matrix->explicit_matrix);
	if (rows != cols || rows == 0)
		error("Lanczos needs a square matrix (%zu x %zu)\n", rows, cols);
	const int64_t D = (int64_t)rows;
	/* element i D + j = M[i][j] couples input state j to output state i
	 * (explicit_matrix.c:76-90); the neutron list holds every (j, i) pair
	 * and the single proton entry (0, 0, 0) */
	const size_t n = (size_t)D * (size_t)D;
	int32_t *entries = (int32_t *)malloc((3 * n + 3) * sizeof(int32_t));
	if (!entries)
		error("out of host memory\n");
	for (int64_t i = 0; i < D; i++)
		for (int64_t j = 0; j < D; j++) {
			int32_t *e = entries + 3 * (size_t)(i * D + j);
			e[0] = (int32_t)j;
			e[1] = (int32_t)i;
			e[2] = (int32_t)(i * D + j);
		}
	entries[3 * n] = entries[3 * n + 1] = entries[3 * n + 2] = 0;
	const int64_t block_np = 1, block_nn = D;
	const int64_t list_offsets[3] = {0, (int64_t)n, (int64_t)n + 1};
	const int64_t matrix_offsets[2] = {0, (int64_t)n};
	const int64_t matrix_dn = (int64_t)n;
	const int32_t force = 2;
	bacchus_hip_instruction_t instruction = {3, 2, 0, 0, 0, 1, 0, 0};
	bacchus_hip_problem_t p = {0};
	p.num_blocks = 1;
	p.block_np = &block_np;
	p.block_nn = &block_nn;
	p.num_lists = 2;
	p.list_offsets = list_offsets;
	p.list_entries = entries;
	p.num_matrix_blocks = 1;
	p.matrix_offsets = matrix_offsets;
	p.matrix_dn = &matrix_dn;
	p.matrix_force = &force;
	p.matrix_elements = get_explicit_matrix_elements(matrix->explicit_matrix);
	p.num_instructions = 1;
	p.instructions = &instruction;
	if (bacchus_hip_operator_create(&p, BACCHUS_HIP_FP64, &matrix->dense) != 0)
		error("%s\n", bacchus_hip_last_error());
	free(entries);
	return matrix->dense;
}
