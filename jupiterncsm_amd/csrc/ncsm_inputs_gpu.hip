// ncsm_inputs_gpu.hip - the bulk stages of the synthetic input generator on
// gfx950: index-list entries and matrix-element blocks (see ncsm_gpu.h).
//
// Index lists follow the contract the reference's consumers read
// (src/Mercury/connection_list/connection_list.c:377-446: a connection
// annihilates the k'-tuple C of an input determinant and creates A, the
// element index is the connection's position in the annihilated-major,
// created-minor enumeration, the phase of a+_A a_C goes to bit 31,
// src/Minerva/matrix_vector_multiplication/matrix_vector_multiplication.c:
// 7-18).  The host generator (ncsm_inputs.cpp, build_list_group) is the
// same algorithm, one determinant at a time; here a thread takes one
// (determinant, annihilated subset) and walks its creation candidates.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>

#include "ncsm_gpu.h"

namespace ncsm {
namespace {

__device__ __forceinline__ int16_t unpack(Packed p, int i)
{
	return (int16_t)((p >> (16 * i)) & 0xFFFFu);
}

// phase of a+_{A1}..a+_{Ak'} a_{Ck'}..a_{C1} on the ascending determinant
// det: annihilate C1 first (sign of its position), then create Ak' first
// (sign of its insertion position)
__device__ int operator_phase(const int16_t *det, int k, const int16_t *C, const int16_t *A,
			      int kk)
{
	int16_t occ[8];
	int n = k;
	for (int i = 0; i < k; i++)
		occ[i] = det[i];
	int sign = 1;
	for (int c = 0; c < kk; c++) {
		int pos = 0;
		while (occ[pos] != C[c])
			pos++;
		if (pos & 1)
			sign = -sign;
		for (int i = pos; i < n - 1; i++)
			occ[i] = occ[i + 1];
		n--;
	}
	for (int a = kk - 1; a >= 0; a--) {
		int pos = 0;
		while (pos < n && occ[pos] < A[a])
			pos++;
		if (pos & 1)
			sign = -sign;
		for (int i = n; i > pos; i--)
			occ[i] = occ[i - 1];
		occ[pos] = A[a];
		n++;
	}
	return sign;
}

// count pass (out == nullptr): counts[i] = entries of item i; fill pass:
// entries of item i from triple offsets[i] on
__global__ void list_entries_kernel(const GenItem *__restrict__ items, int64_t n,
				    const Packed *__restrict__ states,
				    const uint64_t *__restrict__ keys, int32_t *__restrict__ counts,
				    const int64_t *__restrict__ offsets, int32_t *__restrict__ out)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const GenItem it = items[i];
	const int k = it.k, kk = it.kk;
	int16_t det[4], C[4], R[4];
	int nc = 0, nr = 0;
	for (int j = 0; j < k; j++) {
		det[j] = unpack(it.det, j);
		if ((it.sub >> j) & 1)
			C[nc++] = det[j];
		else
			R[nr++] = det[j];
	}
	const uint64_t *ok = keys + it.out_first;
	int64_t pos = out ? offsets[i] : 0;
	int32_t count = 0;
	for (int a = 0; a < it.cre_count; a++) {
		const Packed ap = states[it.cre_first + a];
		int16_t A[4];
		bool clash = false;
		for (int x = 0; x < kk; x++) {
			A[x] = unpack(ap, x);
			for (int y = 0; y < nr; y++)
				clash |= A[x] == R[y];
		}
		if (clash)
			continue;
		// output determinant: the merge of the ascending R and A
		int16_t O[4];
		for (int x = 0, y = 0, o = 0; o < k; o++)
			O[o] = (y >= kk || (x < nr && R[x] < A[y])) ? R[x++] : A[y++];
		uint64_t key = 0;
		for (int o = 0; o < k; o++)
			key |= uint64_t(uint16_t(O[o] + 1)) << (16 * (3 - o));
		int lo = 0, hi = it.out_count;
		while (lo < hi) {
			const int mid = (lo + hi) >> 1;
			if (ok[mid] < key)
				lo = mid + 1;
			else
				hi = mid;
		}
		if (lo >= it.out_count || ok[lo] != key)
			continue;
		if (out) {
			const int sign = operator_phase(det, k, C, A, kk);
			uint32_t mi = (uint32_t)(it.m_base + a);
			if (sign < 0)
				mi |= 0x80000000u;
			out[3 * pos + 0] = it.din;
			out[3 * pos + 1] = lo;
			out[3 * pos + 2] = (int32_t)mi;
			pos++;
		}
		count++;
	}
	if (!out)
		counts[i] = count;
}

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
	z += 0x9e3779b97f4a7c15ULL;
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
	return z ^ (z >> 31);
}

// the host generator's keyed standard normal variate: splitmix64 over
// (seed, rank, the lexicographically smaller of bra/ket, then the other),
// Box-Muller on the two uniform variates it yields
__device__ double keyed_gaussian(uint64_t seed, int rank, const int32_t *bra,
				 const int32_t *ket, int n)
{
	bool swap = false;
	for (int i = 0; i < n; i++)
		if (ket[i] != bra[i]) {
			swap = ket[i] < bra[i];
			break;
		}
	const int32_t *lo = swap ? ket : bra, *hi = swap ? bra : ket;
	uint64_t h = mix64(seed ^ (uint64_t(rank) * 0x100000001b3ULL));
	for (int i = 0; i < n; i++)
		h = mix64(h ^ uint64_t(uint32_t(lo[i])));
	h = mix64(h ^ 0x5bd1e995ULL);
	for (int i = 0; i < n; i++)
		h = mix64(h ^ uint64_t(uint32_t(hi[i])));
	const double u1 = double((h >> 11) + 1) * 0x1.0p-53;
	const double u2 = double(mix64(h) >> 11) * 0x1.0p-53;
	return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925286766559 * u2);
}

// element e of the chunk (blocks[b].offset rebased to the chunk)
__global__ void elements_kernel(const ElemBlock *__restrict__ blocks, int nblocks, int64_t base,
				const Packed *__restrict__ pairs, int64_t pair_base,
				const int32_t *__restrict__ orbital_e, uint64_t seed, double hw,
				double *__restrict__ out, int64_t n)
{
	const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= n)
		return;
	int lo = 0, hi = nblocks - 1;
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if (blocks[mid].offset - base <= e)
			lo = mid;
		else
			hi = mid - 1;
	}
	const ElemBlock b = blocks[lo];
	const int64_t local = e - (b.offset - base);
	const int64_t mp = local / b.nn, mn = local - mp * b.nn;
	const Packed *pp = pairs + 2 * (b.pair_first - pair_base + mp);
	const Packed *pn = pairs + 2 * (b.pair_first - pair_base + b.np + mn);
	// bra: created orbitals, ket: annihilated; protons 2s first, then
	// neutrons 2s + 1
	int32_t bra[6], ket[6];
	int c = 0;
	for (int i = 0; i < b.kp; i++, c++) {
		bra[c] = 2 * unpack(pp[1], i);
		ket[c] = 2 * unpack(pp[0], i);
	}
	for (int i = 0; i < b.kn; i++, c++) {
		bra[c] = 2 * unpack(pn[1], i) + 1;
		ket[c] = 2 * unpack(pn[0], i) + 1;
	}
	const int rank = b.kp + b.kn;
	double v;
	if (rank == 1)
		v = bra[0] == ket[0] ? hw * (orbital_e[bra[0] >> 1] + 1.5) : 0.0;
	else
		v = b.sigma * keyed_gaussian(seed, rank, bra, ket, c);
	out[e] = v;
}

struct DeviceBuffer {
	void *p = nullptr;
	~DeviceBuffer() { (void)hipFree(p); }
};

#define NCSM_HIP(call)                                                                  \
	do {                                                                           \
		hipError_t e_ = (call);                                                \
		if (e_ != hipSuccess) {                                                \
			error = std::string(#call) + ": " + hipGetErrorString(e_);     \
			return 1;                                                      \
		}                                                                      \
	} while (0)

template <typename T> int upload(DeviceBuffer &b, const T *src, size_t n, std::string &error)
{
	NCSM_HIP(hipMalloc(&b.p, std::max<size_t>(n, 1) * sizeof(T)));
	if (n)
		NCSM_HIP(hipMemcpy(b.p, src, n * sizeof(T), hipMemcpyHostToDevice));
	return 0;
}

constexpr int64_t kItemChunk = 1 << 21;      // work items per device pass
constexpr int64_t kElementChunk = 1 << 26;   // elements per device pass

}  // namespace

int gpu_list_entries(const std::vector<GenItem> &items, const std::vector<Packed> &states,
		     const std::vector<uint64_t> &keys, std::vector<int32_t> &entries,
		     std::vector<int32_t> &counts, std::string &error)
{
	DeviceBuffer d_states, d_keys;
	if (upload(d_states, states.data(), states.size(), error) ||
	    upload(d_keys, keys.data(), keys.size(), error))
		return 1;
	counts.assign(items.size(), 0);
	entries.clear();
	for (int64_t i0 = 0; i0 < (int64_t)items.size(); i0 += kItemChunk) {
		const int64_t n = std::min<int64_t>(kItemChunk, (int64_t)items.size() - i0);
		DeviceBuffer d_items, d_counts, d_offsets, d_out;
		if (upload(d_items, items.data() + i0, (size_t)n, error))
			return 1;
		NCSM_HIP(hipMalloc(&d_counts.p, n * sizeof(int32_t)));
		const unsigned grid = (unsigned)((n + 255) / 256);
		hipLaunchKernelGGL(list_entries_kernel, dim3(grid), dim3(256), 0, nullptr,
				   (const GenItem *)d_items.p, n, (const Packed *)d_states.p,
				   (const uint64_t *)d_keys.p, (int32_t *)d_counts.p,
				   (const int64_t *)nullptr, (int32_t *)nullptr);
		NCSM_HIP(hipGetLastError());
		NCSM_HIP(hipMemcpy(counts.data() + i0, d_counts.p, n * sizeof(int32_t),
				   hipMemcpyDeviceToHost));
		std::vector<int64_t> offsets(n);
		int64_t total = 0;
		for (int64_t j = 0; j < n; j++) {
			offsets[j] = total;
			total += counts[i0 + j];
		}
		if (upload(d_offsets, offsets.data(), offsets.size(), error))
			return 1;
		NCSM_HIP(hipMalloc(&d_out.p, std::max<int64_t>(total, 1) * 3 * sizeof(int32_t)));
		hipLaunchKernelGGL(list_entries_kernel, dim3(grid), dim3(256), 0, nullptr,
				   (const GenItem *)d_items.p, n, (const Packed *)d_states.p,
				   (const uint64_t *)d_keys.p, (int32_t *)nullptr,
				   (const int64_t *)d_offsets.p, (int32_t *)d_out.p);
		NCSM_HIP(hipGetLastError());
		const size_t at = entries.size();
		entries.resize(at + 3 * (size_t)total);
		if (total)
			NCSM_HIP(hipMemcpy(entries.data() + at, d_out.p, 3 * total * sizeof(int32_t),
					   hipMemcpyDeviceToHost));
	}
	return 0;
}

int gpu_elements(const std::vector<ElemBlock> &blocks, const std::vector<Packed> &pairs,
		 const std::vector<int32_t> &orbital_e, uint64_t seed, double hw, double *out,
		 int64_t total, std::string &error)
{
	DeviceBuffer d_orb;
	if (upload(d_orb, orbital_e.data(), orbital_e.size(), error))
		return 1;
	// chunks of whole blocks (a block larger than a chunk is a chunk)
	size_t b0 = 0;
	while (b0 < blocks.size()) {
		const int64_t base = blocks[b0].offset;
		size_t b1 = b0 + 1;
		while (b1 < blocks.size() && blocks[b1].offset - base < kElementChunk)
			b1++;
		const int64_t end = b1 < blocks.size() ? blocks[b1].offset : total;
		const int64_t n = end - base;
		const int64_t p0 = blocks[b0].pair_first;
		const int64_t p1 = b1 < blocks.size() ? blocks[b1].pair_first
						      : (int64_t)pairs.size() / 2;
		if (n > 0) {
			DeviceBuffer d_blocks, d_pairs, d_out;
			if (upload(d_blocks, blocks.data() + b0, b1 - b0, error) ||
			    upload(d_pairs, pairs.data() + 2 * p0, (size_t)(2 * (p1 - p0)), error))
				return 1;
			NCSM_HIP(hipMalloc(&d_out.p, n * sizeof(double)));
			hipLaunchKernelGGL(elements_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256),
					   0, nullptr, (const ElemBlock *)d_blocks.p, (int)(b1 - b0), base,
					   (const Packed *)d_pairs.p, p0, (const int32_t *)d_orb.p, seed,
					   hw, (double *)d_out.p, n);
			NCSM_HIP(hipGetLastError());
			NCSM_HIP(hipMemcpy(out + base, d_out.p, n * sizeof(double),
					   hipMemcpyDeviceToHost));
		}
		b0 = b1;
	}
	return 0;
}

}  // namespace ncsm
