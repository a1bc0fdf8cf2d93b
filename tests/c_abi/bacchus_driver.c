/*
 * bacchus_driver.c - a plain C caller of libbacchus_hip, linked the way a
 * Bacchus build links the library in place of the Minerva objects: loads
 * the operator from the reference's files (comb.txt, evaluation order,
 * index_list_<id>, <id>_matrix_elements), runs the file-based matvec
 * (vec_<id> of in_dir added into out_dir, as run_matrix_vector_multiplication
 * does, src/Minerva/scheduler/scheduler.c:80-145), then the reference-mode
 * Lanczos (e_0 start, MGS against u_0..u_{k-3}, |lambda_0 - previous| < 1e-7,
 * src/Bacchus/lanczos/lanczos.c:60-253), and prints the Krylov dimension and
 * the eigenvalues.  Exit status 1 with the library's message on failure,
 * 2 on a usage error.
 */
#include <stdio.h>
#include <stdlib.h>

#include "bacchus_hip.h"

static int fail(const char *what)
{
	fprintf(stderr, "%s: %s\n", what, bacchus_hip_last_error());
	return 1;
}

int main(int argc, char **argv)
{
	if (argc != 10) {
		fprintf(stderr,
			"usage: %s comb.txt num_protons num_neutrons evaluation_order "
			"index_lists_dir matrix_elements_dir in_vector_dir out_vector_dir "
			"max_iterations\n",
			argv[0]);
		return 2;
	}
	const int max_iterations = atoi(argv[9]);
	if (max_iterations < 1)
		return 2;
	bacchus_hip_operator_t *op = NULL;
	if (bacchus_hip_operator_load(argv[1], atoi(argv[2]), atoi(argv[3]), argv[4], argv[5],
				      argv[6], BACCHUS_HIP_FP64, &op))
		return fail("bacchus_hip_operator_load");
	if (bacchus_hip_run_matrix_vector_multiplication(argv[8], argv[7], op))
		return fail("bacchus_hip_run_matrix_vector_multiplication");

	bacchus_hip_lanczos_settings_t st = {0};
	st.max_num_iterations = max_iterations;
	st.start = BACCHUS_HIP_START_E0;
	st.reorthogonalization = BACCHUS_HIP_REORTH_REFERENCE;
	st.convergence = BACCHUS_HIP_CONVERGE_EIGENVALUES;
	st.num_eigenvalues = max_iterations;
	st.eigenvalue_tolerance = 1e-7;
	st.check_every = 1;
	double *eigenvalues = calloc((size_t)max_iterations, sizeof(double));
	if (!eigenvalues)
		return 1;
	bacchus_hip_lanczos_result_t res;
	if (bacchus_hip_lanczos(op, &st, eigenvalues, NULL, NULL, NULL, &res))
		return fail("bacchus_hip_lanczos");
	printf("iterations %lld\n", (long long)res.iterations);
	for (long long k = 0; k < res.iterations && k < max_iterations; k++)
		printf("%.17g\n", eigenvalues[k]);
	free(eigenvalues);
	bacchus_hip_operator_free(op);
	return 0;
}
