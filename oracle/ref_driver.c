/*
 * ref_driver.c - TEST INFRASTRUCTURE ONLY.  A main() around the reference's
 * own Bacchus/Minerva library code (compiled from the sources under the
 * reference tree by oracle/Makefile into oracle/_ref/), bypassing the
 * libconfig settings parser exactly as the reference's own tests do
 * (src/Bacchus/matrix_builder/matrix_builder.c:230-255,
 * src/Bacchus/lanczos/lanczos.c:295-352).
 *
 *   bacchus_ref matvec  <comb> <Z> <N> <evaluation_order> <index_lists_dir>
 *                       <matrix_dir> <in_dir> <out_dir>
 *   bacchus_ref lanczos <comb> <Z> <N> <evaluation_order> <index_lists_dir>
 *                       <matrix_dir> <krylow_dir> <max_iterations>
 *                       <tolerance> <criterion 0|1|2> <result_file>
 *                       [<eigenvector_dir> <target_eigenvector>]
 *
 *   bacchus_ref explicit <matrix_file> <dimension> <krylow_dir>
 *                        <max_iterations> <tolerance> <criterion> <result_file>
 *
 * explicit: a dense D x D matrix (raw row-major doubles) through the
 * EXPLICIT backend (new_zero_matrix + set_matrix_elements, as the
 * reference's own Lanczos tests build theirs, lanczos.c:295-352).
 *
 * The optional eigenvector arguments write eigenvector_<i+1>/vec_<b> for
 * i = 0..target_eigenvector the way the bacchus program does
 * (src/Bacchus/programs/bacchus.c:70-87).
 *
 * The same driver is linked twice by oracle/Makefile: with the reference's
 * own matrix.c, scheduler.c and lanczos.c (_ref/bacchus_ref), and with the
 * three drop-in files of jupiterncsm_amd/dropin/ in their place, calling
 * the GPU library (_ref/bacchus_dropin).
 *
 * LAPACK is not part of the reference; dsteqr_/dsyev_ forward to the
 * LAPACK build that ships with scipy (exported as scipy_dsteqr_ and
 * scipy_dsyev_).
 */
#include <combination_table/combination_table.h>
#include <evaluation_order/evaluation_order.h>
#include <scheduler/scheduler.h>
#include <matrix/matrix.h>
#include <lanczos/lanczos.h>
#include <vector/vector.h>
#include <eigensystem/eigensystem.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

extern void scipy_dsteqr_(char *, int *, double *, double *, double *, int *,
			  double *, int *);
extern void scipy_dsyev_(char *, char *, int *, double *, int *, double *,
			 double *, int *, int *);

void dsteqr_(char *compz, int *n, double *d, double *e, double *z, int *ldz,
	     double *work, int *info)
{
	scipy_dsteqr_(compz, n, d, e, z, ldz, work, info);
}

void dsyev_(char *jobz, char *uplo, int *n, double *a, int *lda, double *w,
	    double *work, int *lwork, int *info)
{
	scipy_dsyev_(jobz, uplo, n, a, lda, w, work, lwork, info);
}

int main(int argc, char **argv)
{
	if (argc >= 9 && strcmp(argv[1], "matvec") == 0) {
		combination_table_t table =
			new_combination_table(argv[2], atoi(argv[3]), atoi(argv[4]));
		evaluation_order_t order = read_evaluation_order(argv[5], table);
		scheduler_t scheduler = new_scheduler(order, table, argv[6], argv[7],
						      (size_t)16 << 30);
		run_matrix_vector_multiplication(argv[9], argv[8], scheduler);
		free_scheduler(scheduler);
		free_evaluation_order(order);
		free_combination_table(table);
		return EXIT_SUCCESS;
	}
	if (argc >= 13 && strcmp(argv[1], "lanczos") == 0) {
		combination_table_t table =
			new_combination_table(argv[2], atoi(argv[3]), atoi(argv[4]));
		evaluation_order_t order = read_evaluation_order(argv[5], table);
		lanczos_settings_t settings = {
			.dimension = get_full_dimension(table),
			.vector_settings = setup_vector_settings(table),
			.krylow_vectors_directory_name = argv[8],
			.max_num_iterations = (size_t)atol(argv[9]),
			.eigenvalue_tolerance = atof(argv[10]),
			.convergence_critera = (convergence_critera_t)atoi(argv[11]),
			.target_eigenvalue = 0,
			.matrix = new_generative_matrix(order, table, argv[6], argv[7],
							(size_t)16 << 30)};
		lanczos_environment_t env = new_lanczos_environment(settings);
		diagonalize(env);
		eigensystem_t system = get_eigensystem(env);
		FILE *out = fopen(argv[12], "w");
		if (!out)
			return EXIT_FAILURE;
		for (size_t i = 0; i < get_num_eigenvalues(system); i++)
			fprintf(out, "%.17g\n", get_eigenvalue(system, i));
		fclose(out);
		if (argc >= 15) {
			/* bacchus.c:70-87 */
			const char *dir = argv[13];
			const size_t target = (size_t)atol(argv[14]);
			for (size_t i = 0; i <= target; i++) {
				vector_settings_t vs = settings.vector_settings;
				vs.directory_name = (char *)calloc(strlen(dir) + 256, sizeof(char));
				sprintf(vs.directory_name, "%s/eigenvector_%lu", dir, i + 1);
				vector_t eigenvector = new_zero_vector(vs);
				get_eigenvector(eigenvector, system, i);
				save_vector(eigenvector);
				free_vector(eigenvector);
				free(vs.directory_name);
			}
		}
		free_eigensystem(system);
		free_lanczos_environment(env);
		free_matrix(settings.matrix);
		free_evaluation_order(order);
		free_combination_table(table);
		return EXIT_SUCCESS;
	}
	if (argc >= 9 && strcmp(argv[1], "explicit") == 0) {
		const size_t D = (size_t)atol(argv[3]);
		double *elements = (double *)malloc(D * D * sizeof(double));
		FILE *in = fopen(argv[2], "rb");
		if (!elements || !in || fread(elements, sizeof(double), D * D, in) != D * D)
			return EXIT_FAILURE;
		fclose(in);
		matrix_t matrix = new_zero_matrix(D, D);
		set_matrix_elements(matrix, elements);
		size_t block_size = D;
		vector_settings_t vs = {.directory_name = NULL, .num_blocks = 1,
					.block_sizes = &block_size};
		lanczos_settings_t settings = {
			.dimension = D,
			.vector_settings = vs,
			.krylow_vectors_directory_name = argv[4],
			.max_num_iterations = (size_t)atol(argv[5]),
			.eigenvalue_tolerance = atof(argv[6]),
			.convergence_critera = (convergence_critera_t)atoi(argv[7]),
			.target_eigenvalue = 0,
			.matrix = matrix};
		lanczos_environment_t env = new_lanczos_environment(settings);
		diagonalize(env);
		eigensystem_t system = get_eigensystem(env);
		FILE *out = fopen(argv[8], "w");
		if (!out)
			return EXIT_FAILURE;
		for (size_t i = 0; i < get_num_eigenvalues(system); i++)
			fprintf(out, "%.17g\n", get_eigenvalue(system, i));
		fclose(out);
		free_eigensystem(system);
		free_lanczos_environment(env);
		free_matrix(matrix);
		free(elements);
		return EXIT_SUCCESS;
	}
	fprintf(stderr, "usage: %s matvec|lanczos|explicit ...\n", argv[0]);
	return EXIT_FAILURE;
}
