/*
 * ncsm_inputs.h - synthetic M-scheme inputs for the Bacchus hot path.
 *
 * The reference gets its inputs from external programs (Anicre for the basis
 * and index lists, Mercury for the matrix-element blocks, Mars for the
 * evaluation order).  None of them can run offline, so this library builds
 * the same artefacts from scratch for a seeded synthetic interaction and can
 * write them in the reference's file formats:
 *
 *   comb.txt                    parsed by src/Utilities/combination_table/
 *                               combination_table.c:83-137,419-640 and
 *                               src/Utilities/calculation_blocks/
 *                               calculation_blocks.c:31-84
 *   evaluation_order            "BLOCK: in out L [L2] V" rows, parsed by
 *                               src/Minerva/evaluation_order/
 *                               evaluation_order.c:175-243 (written like
 *                               src/Mars/programs/mars.c:172-199)
 *   index_lists/index_list_<id> raw {int in, out, matrix} triples, bit 31 of
 *                               matrix = minus sign (src/Minerva/index_list/
 *                               index_list.c:69-97, src/Minerva/
 *                               matrix_vector_multiplication/
 *                               matrix_vector_multiplication.c:7-18)
 *   matrix_elements/<id>_matrix_elements
 *                               size_t Dn, size_t Dp, double[] (src/Minerva/
 *                               matrix_block/matrix_block.c:22-62; element
 *                               order of src/Mercury/connection_list/
 *                               connection_list.c:61-180,377-446)
 *
 * All ids are global and 1-based: basis blocks 1..B, then index lists, then
 * matrix-element blocks.  Everything is also available in memory through the
 * accessors below, with dense 0-based indices (id - first id of the kind).
 */
#ifndef NCSM_INPUTS_H
#define NCSM_INPUTS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ncsm_config {
	int num_protons;      /* particles of species p (<= 4)                 */
	int num_neutrons;     /* particles of species n (<= 4)                 */
	int nmax;             /* HO quanta allowed above the lowest filling    */
	int twice_m;          /* total 2*M of the many-body basis (0)          */
	int max_body;         /* 1, 2 or 3: highest operator rank generated    */
	int include_one_body; /* 1: emit the diagonal 1-body blocks            */
	uint64_t seed;        /* seed of the counter-based element generator   */
	double hw;            /* one-body energy hw*(N+1.5) per orbital        */
	double sigma2;        /* std. dev. of the 2-body elements              */
	double sigma3;        /* std. dev. of the 3-body elements              */
} ncsm_config_t;

typedef struct ncsm_inputs ncsm_inputs_t;

/* Instruction record shared with the matvec library.  type: 1 = neutron
 * (N/NN/NNN), 2 = proton (P/PP/PPP), 3 = neutron-proton (NP/NNP/NPP).
 * All indices are dense and 0-based; -1 marks an unused list. */
typedef struct ncsm_instruction {
	int32_t type;
	int32_t num_particles;  /* operator rank 1, 2 or 3 (CALCBLOCK:k:)     */
	int32_t block_in;
	int32_t block_out;
	int32_t neutron_list;
	int32_t proton_list;
	int32_t matrix_block;
	int32_t pad;
} ncsm_instruction_t;

/* Build the basis, index lists, matrix-element blocks and instructions.
 * Returns NULL and prints a message on invalid configurations. */
ncsm_inputs_t *ncsm_inputs_build(const ncsm_config_t *config);
/* The same inputs with the two bulk stages on the GPU (current HIP device):
 * the index-list entries (connection_list.c:377-446 contract; one device
 * thread per input determinant and annihilated subset walks the creation
 * candidates) and the matrix-element blocks (one thread per element, the
 * same keyed Gaussians).  Index lists, instructions and shapes are
 * identical to ncsm_inputs_build's; elements agree to a few units in the
 * last place (the device's log and cos).  Returns NULL and prints a
 * message on invalid configurations or device errors. */
ncsm_inputs_t *ncsm_inputs_build_gpu(const ncsm_config_t *config);
void ncsm_inputs_free(ncsm_inputs_t *inputs);

/* Write comb.txt, evaluation_order, index_lists/ and matrix_elements/ under
 * directory (created if missing).  Returns 0 on success. */
int ncsm_inputs_write(const ncsm_inputs_t *inputs, const char *directory);

int64_t ncsm_dimension(const ncsm_inputs_t *inputs);
int64_t ncsm_num_blocks(const ncsm_inputs_t *inputs);
/* per block: Ep, 2Mp, En, 2Mn, Np, Nn (6 int64 per block) */
const int64_t *ncsm_blocks(const ncsm_inputs_t *inputs);

int64_t ncsm_num_lists(const ncsm_inputs_t *inputs);
/* offsets in triples, num_lists+1 entries */
const int64_t *ncsm_list_offsets(const ncsm_inputs_t *inputs);
/* {in,out,matrix} triples, 3 int32 each */
const int32_t *ncsm_list_entries(const ncsm_inputs_t *inputs);
/* per list: species (0 = proton, 1 = neutron), k, E_bra, 2M_bra, E_ket,
 * 2M_ket, depth (7 int32 per list) */
const int32_t *ncsm_list_settings(const ncsm_inputs_t *inputs);

int64_t ncsm_num_matrix_blocks(const ncsm_inputs_t *inputs);
/* offsets in elements, num_matrix_blocks+1 entries */
const int64_t *ncsm_matrix_offsets(const ncsm_inputs_t *inputs);
const double *ncsm_matrix_elements(const ncsm_inputs_t *inputs);
/* per block: Dn, Dp, force class (1,2,3) (3 int64 per block) */
const int64_t *ncsm_matrix_shapes(const ncsm_inputs_t *inputs);

int64_t ncsm_num_instructions(const ncsm_inputs_t *inputs);
const ncsm_instruction_t *ncsm_instructions(const ncsm_inputs_t *inputs);

/* Number of orbitals per species and their quantum numbers (e, l, 2j, 2m),
 * 4 int32 per orbital, in the Antoine order of src/Neptune/bases/
 * sp_states.c:7-35 (e, l, j ascending, m descending). */
int32_t ncsm_num_orbitals(const ncsm_inputs_t *inputs);
const int32_t *ncsm_orbitals(const ncsm_inputs_t *inputs);

/* Per block and species, the Slater determinants of the block's proton or
 * neutron sub-basis in storage order: species 0 = proton, 1 = neutron.
 * Writes count*k orbital indices to out (if non-NULL) and returns count. */
int64_t ncsm_subbasis_states(const ncsm_inputs_t *inputs, int64_t block,
			     int species, int32_t *out);

/* Seeded matrix element of the synthetic interaction between the created
 * (bra) and annihilated (ket) orbital tuples.  Orbitals are global sp
 * indices: proton orbital s -> 2s, neutron orbital s -> 2s+1, protons
 * first, each species ascending.  Returns 0 for rank mismatches, M or
 * parity violation.  Symmetric under bra <-> ket. */
double ncsm_element_value(const ncsm_inputs_t *inputs, int rank,
			  const int32_t *bra, const int32_t *ket);

#ifdef __cplusplus
}
#endif
#endif
