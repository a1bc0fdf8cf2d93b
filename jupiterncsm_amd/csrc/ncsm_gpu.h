// ncsm_gpu.h - the device stages of the synthetic input generator (internal
// to libncsm_inputs, not part of its C ABI).
//
// The host keeps the combinatorics that are small (orbitals, k-particle
// state sets, sub-bases, connection-table offsets) and hands the two bulk
// stages to the GPU as flat tables:
//   * index-list entries: one work item per (list group, input determinant,
//     annihilated k'-subset); the device walks the item's creation
//     candidates (Pauli check, sorted output determinant, binary search in
//     the output sub-basis, fermionic phase) in a count pass and a fill
//     pass, so the entries come out in the host generator's order;
//   * matrix elements: one thread per element of every block.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace ncsm {

// A k-particle determinant or k'-tuple: four int16 orbital indices, -1
// padded, packed little-endian (entry i in bits 16 i .. 16 i + 15).
using Packed = uint64_t;

struct GenItem {
	Packed det;         // input determinant (k orbitals, ascending)
	int64_t cre_first;  // first creation candidate in the state table
	int64_t m_base;     // connection index of (C, first candidate)
	int64_t out_first;  // first key of the output sub-basis
	int32_t out_count;  // its determinants
	int32_t cre_count;  // creation candidates
	int32_t din;        // index of det in the input sub-basis
	uint8_t sub;        // bit mask of the annihilated positions of det
	uint8_t k;          // particles per determinant
	uint8_t kk;         // operator rank k'
	uint8_t pad;
};

// Entries {din, out, matrix | phase bit 31} of every item, in item order;
// counts[i] = entries of item i.  states: packed k'-tuples of every state
// set; keys: sorted big-endian keys of every sub-basis (see key_be).
int gpu_list_entries(const std::vector<GenItem> &items, const std::vector<Packed> &states,
		     const std::vector<uint64_t> &keys, std::vector<int32_t> &entries,
		     std::vector<int32_t> &counts, std::string &error);

// One matrix-element block: pairs [pair_first, pair_first + np) of the
// proton part and [pair_first + np, ... + nn) of the neutron part (a part
// of rank 0 has one empty pair); element (mp, mn) at offset + mn + nn mp.
struct ElemBlock {
	int64_t offset;
	int64_t pair_first;
	int32_t np, nn;
	int32_t kp, kn;
	double sigma;
};

// Every element of the blocks: rank 1 the diagonal one-body energy
// hw (e + 1.5), else sigma times the keyed Gaussian of (seed, rank, bra,
// ket) of the host generator.  pairs: {annihilated, created} packed tuples;
// orbital_e: HO energy of every orbital.
int gpu_elements(const std::vector<ElemBlock> &blocks, const std::vector<Packed> &pairs,
		 const std::vector<int32_t> &orbital_e, uint64_t seed, double hw,
		 double *out, int64_t total, std::string &error);

// Big-endian key of an ascending k-tuple: lexicographic order of tuples of
// one length is numeric order of keys.
inline uint64_t key_be(const int16_t *t, int k)
{
	uint64_t key = 0;
	for (int i = 0; i < k; i++)
		key |= uint64_t(uint16_t(t[i] + 1)) << (16 * (3 - i));
	return key;
}

}  // namespace ncsm
