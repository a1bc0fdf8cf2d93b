/*
 * dropin_lanczos.c - lanczos.h (src/Bacchus/lanczos/lanczos.h:30-36) on the
 * MI355X: replaces src/Bacchus/lanczos/lanczos.c.
 *
 * diagonalize runs the reference's recurrence on the device through
 * bacchus_hip_lanczos: the e_0 start (lanczos.c:60-72), the three-term step
 * (:74-153), MGS of each new vector against u_0..u_{k-3} with no
 * renormalization (:155-169), and the stopping rule of :171-253 on the
 * eigenvalue index settings.target_eigenvalue (eigenvalues or the
 * eigenvector-amplitude criterion, or none), at most
 * min(dimension, max_num_iterations) iterations.  The matrix is the
 * device operator behind settings.matrix (dropin_matrix.c).  Afterwards the
 * Krylov vectors u_0..u_{k-1} are written to the Krylov directory as the
 * reference keeps them (<dir>/vector_<k>/vec_<b>, basis.c:37-52), so
 * get_eigensystem returns what the reference's does: the eigensystem of
 * the tridiagonal matrix from diagonalization.c (dsteqr_) over that
 * basis, and get_eigenvector / bacchus.c:70-87 combine the eigenvectors
 * from it unchanged.
 */
#include <basis/basis.h>
#include <diagonalization/diagonalization.h>
#include <directory_tools/directory_tools.h>
#include <eigensystem/eigensystem.h>
#include <error/error.h>
#include <lanczos/lanczos.h>

#include <errno.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "dropin.h"

struct _lanczos_environment_
{
	lanczos_settings_t settings;
	basis_t krylow_basis;
	double *diagonal_elements;
	double *off_diagonal_elements;
	size_t dimension_krylow_space;
};

lanczos_environment_t new_lanczos_environment(lanczos_settings_t settings)
{
	lanczos_environment_t environment =
		(lanczos_environment_t)calloc(1, sizeof(struct _lanczos_environment_));
	const size_t n = settings.max_num_iterations ? settings.max_num_iterations : 1;
	if (!environment)
		error("out of host memory\n");
	environment->settings = settings;
	environment->krylow_basis = new_basis_empty(settings.vector_settings,
						    settings.krylow_vectors_directory_name,
						    settings.max_num_iterations + 1);
	environment->diagonal_elements = (double *)calloc(n, sizeof(double));
	environment->off_diagonal_elements = (double *)calloc(n, sizeof(double));
	if (!environment->diagonal_elements || !environment->off_diagonal_elements)
		error("out of host memory\n");
	if (!directory_exists(settings.krylow_vectors_directory_name) &&
	    create_directory(settings.krylow_vectors_directory_name) != 0)
		error("Could not create krylow vector directory \"%s\". %s\n",
		      settings.krylow_vectors_directory_name, strerror(errno));
	return environment;
}

/* u_k of the host copy into the files of Krylov vector k, block by block
 * in the vector settings' order (vec_<i+1>, vector.c:98-128) */
static void write_krylow_vector(lanczos_environment_t environment, size_t k,
				const double *u)
{
	const vector_settings_t *vs = &environment->settings.vector_settings;
	basis_append_vector(environment->krylow_basis);
	size_t offset = 0;
	for (size_t b = 0; b < vs->num_blocks; b++) {
		char path[4096];
		snprintf(path, sizeof path, "%s/vector_%zu/vec_%zu",
			 environment->settings.krylow_vectors_directory_name, k, b + 1);
		FILE *f = fopen(path, "wb");
		if (!f)
			error("Could not open vector block file %s. %s\n", path, strerror(errno));
		const size_t len = vs->block_sizes[b];
		if (fwrite(u + offset, sizeof(double), len, f) != len)
			error("Could not write vector block file %s\n", path);
		fclose(f);
		offset += len;
	}
}

void diagonalize(lanczos_environment_t environment)
{
	struct timespec t_start, t_end;
	printf("Diagonalzation start:\n");
	clock_gettime(CLOCK_REALTIME, &t_start);
	const lanczos_settings_t *s = &environment->settings;
	bacchus_hip_operator_t *op = dropin_matrix_operator(s->matrix);
	const int64_t D = bacchus_hip_operator_dimension(op);
	size_t blocks_total = 0;
	for (size_t b = 0; b < s->vector_settings.num_blocks; b++)
		blocks_total += s->vector_settings.block_sizes[b];
	if ((size_t)D != s->dimension || blocks_total != s->dimension)
		error("dimension %zu, vector blocks %zu and matrix %lld disagree\n", s->dimension,
		      blocks_total, (long long)D);
	const size_t K = s->dimension < s->max_num_iterations ? s->dimension
							       : s->max_num_iterations;
	if (K == 0)
		error("max_num_iterations must be positive\n");
	bacchus_hip_lanczos_settings_t st;
	memset(&st, 0, sizeof st);
	st.max_num_iterations = (int64_t)s->max_num_iterations;
	st.start = BACCHUS_HIP_START_E0;
	st.reorthogonalization = BACCHUS_HIP_REORTH_REFERENCE;
	switch (s->convergence_critera) {
	case converge_eigenvalues:
		st.convergence = BACCHUS_HIP_CONVERGE_EIGENVALUES;
		break;
	case converge_eigenvectors:
		st.convergence = BACCHUS_HIP_CONVERGE_EIGENVECTORS;
		break;
	default:
		st.convergence = BACCHUS_HIP_CONVERGE_NONE;
		break;
	}
	st.num_eigenvalues = 1;
	st.eigenvalue_tolerance = s->eigenvalue_tolerance;
	st.check_every = 1;
	st.target_eigenvalue = (int32_t)s->target_eigenvalue;
	st.krylov_basis = (double *)malloc(K * (size_t)D * sizeof(double));
	if (!st.krylov_basis)
		error("out of host memory for %zu Krylov vectors\n", K);
	double eigenvalue = 0;
	bacchus_hip_lanczos_result_t result;
	memset(&result, 0, sizeof result);
	if (bacchus_hip_lanczos(op, &st, &eigenvalue, environment->diagonal_elements,
				environment->off_diagonal_elements, NULL, &result) != 0)
		error("%s\n", bacchus_hip_last_error());
	environment->dimension_krylow_space = (size_t)result.iterations;
	for (size_t k = 0; k < (size_t)result.iterations; k++)
		write_krylow_vector(environment, k, st.krylov_basis + k * (size_t)D);
	free(st.krylov_basis);
	clock_gettime(CLOCK_REALTIME, &t_end);
	printf("Lanczos on the GPU: %lld iterations, %s, matvec %lg s\n",
	       (long long)result.iterations, result.converged ? "converged" : "iteration limit",
	       result.matvec_seconds);
	printf("Diagonalization end after %lg µs\n",
	       (t_end.tv_sec - t_start.tv_sec) * 1e6 + (t_end.tv_nsec - t_start.tv_nsec) * 1e-3);
}

eigensystem_t get_eigensystem(lanczos_environment_t environment)
{
	eigensystem_t system = diagonalize_tridiagonal_matrix(
		environment->diagonal_elements, environment->off_diagonal_elements,
		basis_get_dimension(environment->krylow_basis));
	set_basis(system, environment->krylow_basis);
	return system;
}

void free_lanczos_environment(lanczos_environment_t environment)
{
	if (!environment)
		return;
	free_basis(environment->krylow_basis);
	free(environment->diagonal_elements);
	free(environment->off_diagonal_elements);
	free(environment);
}
