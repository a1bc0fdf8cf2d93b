/*
 * dropin_scheduler.c - scheduler.h (src/Minerva/scheduler/scheduler.h:10-20)
 * on the MI355X: replaces src/Minerva/scheduler/scheduler.c.
 *
 * new_scheduler reads the basis blocks of the combination table and the
 * instructions of the evaluation order through their public iterators,
 * loads index_list_<id> and <id>_matrix_elements once and keeps the operator
 * resident in device memory; run_matrix_vector_multiplication adds H x to
 * the output vector's vec_<id> files, as the reference's does
 * (scheduler.c:80-145, the output blocks loaded and accumulated into).
 * Errors print the library's message through the reference's error()
 * (src/Utilities/error/error.h:7-13) and exit with EXIT_FAILURE.
 */
#include <combination_table/combination_table.h>
#include <error/error.h>
#include <evaluation_order/evaluation_order.h>
#include <scheduler/scheduler.h>
#include <string_tools/string_tools.h>

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "dropin.h"

struct _scheduler_
{
	bacchus_hip_operator_t *op;
	/* copied, as the reference copies them (scheduler.c:62-78) */
	char *index_lists_base_directory;
	char *matrix_file_base_directory;
};

/* force class (1, 2, 3 body) of every element block id, from the
 * table's 1NF/2NF/3NF iterators; 0 for ids it does not list */
static int *element_block_ranks(combination_table_t table, size_t *num_ids)
{
	const size_t n = get_num_arrays(table) + 1;
	int *rank = (int *)calloc(n, sizeof(int));
	if (!rank)
		error("out of host memory\n");
	reset_1nf_block_iterator(table);
	while (has_next_1nf_block(table)) {
		const matrix_block_setting_t s = next_1nf_block_iterator(table);
		if (s.matrix_block_id < n)
			rank[s.matrix_block_id] = 1;
	}
	reset_2nf_block_iterator(table);
	while (has_next_2nf_block(table)) {
		const matrix_block_setting_t s = next_2nf_block_iterator(table);
		if (s.matrix_block_id < n)
			rank[s.matrix_block_id] = 2;
	}
	reset_3nf_block_iterator(table);
	while (has_next_3nf_block(table)) {
		const matrix_block_setting_t s = next_3nf_block_iterator(table);
		if (s.matrix_block_id < n)
			rank[s.matrix_block_id] = 3;
	}
	*num_ids = n;
	return rank;
}

scheduler_t new_scheduler(evaluation_order_t evaluation_order,
			  combination_table_t combination_table,
			  const char *index_lists_base_directory,
			  const char *matrix_file_base_directory,
			  size_t maximum_loaded_memory)
{
	const size_t nb = get_num_basis_blocks(combination_table);
	int64_t *ids = (int64_t *)calloc(nb ? nb : 1, sizeof(int64_t));
	int64_t *np = (int64_t *)calloc(nb ? nb : 1, sizeof(int64_t));
	int64_t *nn = (int64_t *)calloc(nb ? nb : 1, sizeof(int64_t));
	const size_t ni = get_num_instructions(evaluation_order);
	bacchus_hip_file_instruction_t *ins = (bacchus_hip_file_instruction_t *)calloc(
		ni ? ni : 1, sizeof(bacchus_hip_file_instruction_t));
	if (!ids || !np || !nn || !ins)
		error("out of host memory\n");
	/* basis blocks in vector order (the order setup_vector_settings gives
	 * the vec_<i+1> files, vector.c:68-96) */
	size_t k = 0;
	reset_basis_block_iterator(combination_table);
	while (has_next_basis_block(combination_table) && k < nb) {
		const basis_block_t b = next_basis_block(combination_table);
		ids[k] = (int64_t)b.block_id;
		np[k] = (int64_t)b.num_proton_states;
		nn[k] = (int64_t)b.num_neutron_states;
		k++;
	}
	size_t num_ids = 0;
	int *rank = element_block_ranks(combination_table, &num_ids);
	/* instructions: the evaluation order's rows (UNLOAD_ rows are not
	 * kept by read_evaluation_order, evaluation_order.c:175-190) */
	size_t m = 0;
	evaluation_order_iterator_t it = get_evaluation_order_iterator(evaluation_order);
	while (has_next_instruction(it)) {
		const evaluation_instruction_t I = next_instruction(it);
		if (I.type != neutron_block && I.type != proton_block &&
		    I.type != neutron_proton_block)
			continue;
		if (m >= ni)
			break;
		bacchus_hip_file_instruction_t *f = &ins[m++];
		f->type = (int32_t)I.type;
		f->block_in = (int64_t)I.vector_block_in;
		f->block_out = (int64_t)I.vector_block_out;
		f->neutron_list = I.type != proton_block ? (int64_t)I.neutron_index : 0;
		f->proton_list = I.type != neutron_block ? (int64_t)I.proton_index : 0;
		f->matrix_block = (int64_t)I.matrix_element_file;
		f->num_particles = I.matrix_element_file < num_ids ? rank[I.matrix_element_file] : 0;
	}
	free_evaluation_order_iterator(it);
	scheduler_t scheduler = (scheduler_t)calloc(1, sizeof(struct _scheduler_));
	if (!scheduler)
		error("out of host memory\n");
	scheduler->index_lists_base_directory = copy_string(index_lists_base_directory);
	scheduler->matrix_file_base_directory = copy_string(matrix_file_base_directory);
	if (bacchus_hip_operator_load_blocks((int64_t)k, ids, np, nn, (int64_t)m, ins,
					     index_lists_base_directory,
					     matrix_file_base_directory, maximum_loaded_memory,
					     BACCHUS_HIP_FP64, &scheduler->op) != 0)
		error("%s\n", bacchus_hip_last_error());
	free(rank);
	free(ins);
	free(nn);
	free(np);
	free(ids);
	return scheduler;
}

void run_matrix_vector_multiplication(const char *output_vector_base_directory,
				      const char *input_vector_base_directory,
				      scheduler_t scheduler)
{
	if (bacchus_hip_run_matrix_vector_multiplication(output_vector_base_directory,
							 input_vector_base_directory,
							 scheduler->op) != 0)
		error("%s\n", bacchus_hip_last_error());
}

void free_scheduler(scheduler_t scheduler)
{
	if (!scheduler)
		return;
	bacchus_hip_operator_free(scheduler->op);
	free(scheduler->index_lists_base_directory);
	free(scheduler->matrix_file_base_directory);
	free(scheduler);
}

bacchus_hip_operator_t *dropin_scheduler_operator(scheduler_t scheduler)
{
	return scheduler ? scheduler->op : NULL;
}
