// np.hip - the neutron-proton instructions of the Bacchus matvec
// (multiplication_neutrons_protons and its off-diagonal form,
// src/Minerva/matrix_vector_multiplication/matrix_vector_multiplication.c:
// 134-210, 364-462), element-row-major.
//
// The reference loops over every (neutron entry, proton entry) pair and
// loads V[m_n + Dn m_p] for each.  An element W[a][b] of an instruction is
// used by every pair of entries with compact indices (a, b); here the
// row-list entries are grouped by b, so one element value loaded into a
// register is applied to all row-list entries with that b, each adding into
// its own output row.  The lane list is sliced by output state (SELL: q
// lanes per state, 64 lanes per slice, entries of a state dealt round-robin
// to its lanes, states sorted by entry count), and the element rows are
// stored expanded into that order per b with the lane entry's phase folded
// in: a wave reads each W_b with coalesced 8-byte loads and the input
// indices with 4-byte ones, sums a state's products in registers (then
// over its q lanes) and adds one value per (row entry, state) into the
// workgroup's LDS copy of the output block.  A workgroup owns a whole output
// block, so every row-list entry of an instruction direction, whatever its
// output row, is in the same workgroup as the other entries of its b: each
// expanded element row is streamed from HBM once per matvec.
#include "bacchus_internal.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <exception>
#include <map>
#include <mutex>
#include <numeric>
#include <thread>
#include <tuple>
#include <type_traits>

namespace bacchus {
namespace {

constexpr uint32_t kMask = 0x7FFFFFFFu;
constexpr uint32_t kPad = 0xFFFFFFFFu;

template <typename VT> struct NpElem;
template <> struct NpElem<double> {
	using acc_t = double;
	using x_t = double;
	__device__ static double load(const double *p, size_t i) { return p[i]; }
};
template <> struct NpElem<float> {
	using acc_t = float;
	using x_t = float;
	__device__ static float load(const float *p, size_t i) { return p[i]; }
};
template <> struct NpElem<uint16_t> {
	using acc_t = float;
	using x_t = float;
	__device__ static float load(const uint16_t *p, size_t i)
	{
		return __uint_as_float(uint32_t(p[i]) << 16);
	}
};

// One workgroup per tile: the output block in LDS (FP64 sums) followed by
// one x staging area of kNpStage words per wave; waves take the tile's
// units in turn.  A unit is one instruction direction, a range of row-list
// compact indices b and a range of lane-list slices.  For instructions
// whose x rows fit the staging area (I.stage rows per pass), a pass copies
// its rows there once and every slice reads x from LDS; the others gather
// x from global memory, kNpChunk rows per pass.
template <typename VT>
__global__ __launch_bounds__(kNpThreads) void np_kernel(
	const NpTile *__restrict__ tiles, const NpUnit *__restrict__ units,
	const NpInstr *__restrict__ ins, const NpSlice *__restrict__ slices,
	const uint32_t *__restrict__ sell_in, const int32_t *__restrict__ sell_row,
	const uint2 *__restrict__ jl, const uint32_t *__restrict__ bptr,
	const VT *__restrict__ W, const typename NpElem<VT>::x_t *__restrict__ x_normal,
	const typename NpElem<VT>::x_t *__restrict__ x_transposed, double *__restrict__ y_normal,
	double *__restrict__ y_transposed, int32_t block_words, double s1, double s2, double s3)
{
	using acc_t = typename NpElem<VT>::acc_t;
	using x_t = typename NpElem<VT>::x_t;
	extern __shared__ double blk[];
	const NpTile T = tiles[blockIdx.x];
	const int words = T.rows * T.cols;
	for (int t = threadIdx.x; t < words; t += blockDim.x)
		blk[t] = 0.0;
	__syncthreads();
	const x_t *__restrict__ xs = T.space ? x_transposed : x_normal;
	const int lane = threadIdx.x & 63;
	const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const int nwaves = blockDim.x >> 6;
	// this wave's staging area (x values in their own type)
	x_t *__restrict__ stage =
		reinterpret_cast<x_t *>(blk + block_words) + (size_t)wave * kNpStage;
	for (int32_t ui = T.u_begin + wave; ui < T.u_end; ui += nwaves) {
		const NpUnit U = units[ui];
		const NpInstr I = ins[U.vi];
		const acc_t sc = acc_t(I.force == 1 ? s1 : (I.force == 2 ? s2 : s3));
		const x_t *__restrict__ xb = xs + I.x_off;
		const uint32_t nl = (uint32_t)I.nl_in;
		const int rows_per_pass = I.stage ? I.stage : kNpChunk;
		// the unit's metadata in lane registers (the builder keeps every
		// unit within 64 slices, 63 compact indices and 64 row-list
		// entries), read back with readlane: no dependent loads per slice
		// or pass
		const uint32_t nsl = U.k1 - U.k0, nb = U.b1 - U.b0;
		uint4 sl = make_uint4(0, 0, 0, 0);
		if ((uint32_t)lane < nsl) {
			const NpSlice q = slices[I.slice0 + U.k0 + lane];
			sl = make_uint4(q.off, q.steps, q.row_base, q.shape);
		}
		const uint32_t bp = (uint32_t)lane <= nb ? bptr[I.bbase + U.b0 + lane] : 0u;
		uint32_t jfirst = __builtin_amdgcn_readfirstlane(bp);
		uint32_t jend = __builtin_amdgcn_readlane(bp, nb);
		if (U.jsub) {
			// one compact index with more than 64 entries: this unit's
			// part of them
			jfirst += U.jsub >> 16;
			jend = jfirst + (U.jsub & 0xFFFFu);
		}
		const uint2 je = (uint32_t)lane < jend - jfirst ? jl[jfirst + lane] : make_uint2(0u, 0u);
		for (uint32_t bi = 0; bi < nb; bi++) {
			const uint32_t b = U.b0 + bi;
			const uint32_t j0 = U.jsub ? 0u : __builtin_amdgcn_readlane(bp, bi) - jfirst;
			const uint32_t j1 = U.jsub ? jend - jfirst
						   : __builtin_amdgcn_readlane(bp, bi + 1) - jfirst;
			const VT *__restrict__ wb = W + I.w_off + (size_t)b * I.sell_len;
			for (uint32_t jc = j0; jc < j1; jc += (uint32_t)rows_per_pass) {
				const int n = (int)min((uint32_t)rows_per_pass, j1 - jc);
				// row-list entries of this pass (wave-uniform)
				uint32_t xrow[kNpChunk], orow[kNpChunk];
				acc_t sg[kNpChunk];
#pragma unroll
				for (int c = 0; c < kNpChunk; c++) {
					const uint32_t at = jc + (uint32_t)(c < n ? c : 0);
					xrow[c] = nl * __builtin_amdgcn_readlane(je.x, at);
					const uint32_t so = __builtin_amdgcn_readlane(je.y, at);
					orow[c] = (so & kMask) * (uint32_t)T.cols;
					sg[c] = (so & 0x80000000u) ? -sc : sc;
				}
				if (I.stage) {
					// the pass's x rows into the staging area; the wave's
					// earlier reads of it are complete (their values fed
					// the sums already flushed)
					__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
					// all loads first (one wait), then the LDS writes:
					// position p of the area is row p / nl, column p % nl
					{
						const uint32_t total = (uint32_t)n * nl;
						x_t tmp[kNpStage / 64];
#pragma unroll
						for (int k = 0; k < kNpStage / 64; k++) {
							const uint32_t p = (uint32_t)lane + 64u * (uint32_t)k;
							tmp[k] = x_t(0);
							if (p < total) {
								const uint32_t c = p / nl, i = p - c * nl;
								const uint32_t r = c == 0 ? xrow[0] : c == 1 ? xrow[1]
										 : c == 2 ? xrow[2] : xrow[3];
								tmp[k] = xb[r + i];
							}
						}
#pragma unroll
						for (int k = 0; k < kNpStage / 64; k++) {
							const uint32_t p = (uint32_t)lane + 64u * (uint32_t)k;
							if (p < total)
								stage[p] = tmp[k];
						}
					}
					__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
					// rows past n read row 0 (their sums are never used)
#pragma unroll
					for (int c = 0; c < kNpChunk; c++)
						xrow[c] = (uint32_t)(c < n ? c : 0) * nl;
				}
				// the pass's slices as one pipeline of rounds (kNpSteps
				// steps of one slice each): the index, element and row
				// loads of the next round are issued before the products
				// of the current one, across slice boundaries
				const uint32_t *__restrict__ ein0 = sell_in + I.sell_base;
				const x_t *__restrict__ xsrc = I.stage ? stage : xb;
				auto issue = [&](uint32_t k, uint32_t t, uint32_t (&e)[kNpSteps],
						 acc_t (&v)[kNpSteps]) {
					const uint32_t off = __builtin_amdgcn_readlane(sl.x, k);
					const uint32_t steps = __builtin_amdgcn_readlane(sl.y, k);
					const uint32_t width = __builtin_amdgcn_readlane(sl.w, k) & 0xFFFFu;
#pragma unroll
					for (int u = 0; u < kNpSteps; u++) {
						const bool ok = (uint32_t)lane < width && t + (uint32_t)u < steps;
						const uint32_t p = off + (t + (uint32_t)u) * width + (uint32_t)lane;
						e[u] = ok ? ein0[p] : 0u;
						v[u] = ok ? NpElem<VT>::load(wb, p) : acc_t(0);
					}
				};
				auto row_of = [&](uint32_t k) {
					return sell_row[__builtin_amdgcn_readlane(sl.z, k) + (uint32_t)lane];
				};
				uint32_t ec[kNpSteps], en[kNpSteps];
				acc_t vc[kNpSteps], vn[kNpSteps];
				acc_t acc[kNpChunk];
#pragma unroll
				for (int c = 0; c < kNpChunk; c++)
					acc[c] = 0;
				uint32_t k = 0, t = 0;
				issue(0, 0, ec, vc);
				int32_t row_c = row_of(0), row_n = -1;
				for (;;) {
					const uint32_t steps = __builtin_amdgcn_readlane(sl.y, k);
					const bool same = t + kNpSteps < steps;
					const uint32_t kn = same ? k : k + 1, tn = same ? t + kNpSteps : 0u;
					const bool more = kn < nsl;
					if (more) {
						issue(kn, tn, en, vn);
						if (!same)
							row_n = row_of(kn);
					}
					// products of the current round (missing steps have zero
					// elements and read input 0)
#pragma unroll
					for (int u = 0; u < kNpSteps; u++) {
						acc_t xv[kNpChunk];
#pragma unroll
						for (int c = 0; c < kNpChunk; c++)
							xv[c] = c < n ? acc_t(xsrc[xrow[c] + ec[u]]) : acc_t(0);
#pragma unroll
						for (int c = 0; c < kNpChunk; c++)
							acc[c] += vc[u] * xv[c];
					}
					if (!same) {
						// slice k done: a state's q lanes hold parts of its
						// sum (width is a multiple of q, so a group is all in
						// or all out)
						const int q = 1 << (__builtin_amdgcn_readlane(sl.w, k) >> 16);
						for (int o = q >> 1; o > 0; o >>= 1)
#pragma unroll
							for (int c = 0; c < kNpChunk; c++)
								acc[c] += __shfl_xor(acc[c], o, 64);
						if ((lane & (q - 1)) == 0 && row_c >= 0) {
#pragma unroll
							for (int c = 0; c < kNpChunk; c++)
								if (c < n)
									atomicAdd(&blk[orow[c] + (uint32_t)row_c],
										  (double)(sg[c] * acc[c]));
						}
#pragma unroll
						for (int c = 0; c < kNpChunk; c++)
							acc[c] = 0;
						row_c = row_n;
					}
					if (!more)
						break;
					k = kn;
					t = tn;
#pragma unroll
					for (int u = 0; u < kNpSteps; u++) {
						ec[u] = en[u];
						vc[u] = vn[u];
					}
				}
			}
		}
	}
	__syncthreads();
	double *__restrict__ y = (T.space ? y_transposed : y_normal) + T.y_off;
	for (int t = threadIdx.x; t < words; t += blockDim.x) {
		const double v = blk[t];
		if (v != 0.0)
			atomicAdd(y + t, v);
	}
}

template <typename F> void for_each_parallel(size_t n, F &&fn)
{
	const unsigned hw = std::thread::hardware_concurrency();
	const size_t nthreads = std::min<size_t>(n, std::max(1u, std::min(16u, hw ? hw : 1u)));
	std::atomic<size_t> next{0};
	std::exception_ptr failure;
	std::mutex lock;
	auto worker = [&]() {
		try {
			for (size_t q; (q = next.fetch_add(1)) < n;)
				fn(q);
		} catch (...) {
			std::lock_guard<std::mutex> g(lock);
			if (!failure)
				failure = std::current_exception();
			next.store(n);
		}
	};
	std::vector<std::thread> pool;
	for (size_t t = 1; t < nthreads; t++) {
		try {
			pool.emplace_back(worker);
		} catch (...) {
			break;
		}
	}
	worker();
	for (auto &t : pool)
		t.join();
	if (failure)
		std::rethrow_exception(failure);
}

uint16_t bf16_of(double v)
{
	float f = (float)v;
	uint32_t u;
	memcpy(&u, &f, 4);
	if ((u & 0x7f800000u) == 0x7f800000u)
		return (uint16_t)(u >> 16);
	u += 0x7fffu + ((u >> 16) & 1u);
	return (uint16_t)(u >> 16);
}

// The SELL form of one lane list in one direction (host side).
struct Sell {
	int64_t base = 0;             // first position in the global arrays
	uint32_t len = 0;             // padded length (positions)
	uint32_t slice0 = 0, nslices = 0;
	std::vector<uint32_t> comp;   // compact | sign per position, kPad at
				      // padding (host only, for the elements)
};

// The row-list entries of one list in one direction grouped by compact
// index.
struct JGroup {
	uint32_t bbase = 0;  // index of bptr[0]
	std::vector<uint32_t> count;
};

}  // namespace

int np_build(bacchus_hip_operator *op, const NpContext &c)
{
	const bacchus_hip_problem_t *p = c.p;
	const auto &mvals = *c.mvals;
	const auto &swapped = *c.swapped;
	const auto &rowptr = *c.rowptr;
	const auto &entries = *c.entries;
	const int64_t ni = p->num_instructions;
	auto list_len = [&](int32_t l) -> int64_t {
		return l >= 0 ? p->list_offsets[l + 1] - p->list_offsets[l] : 0;
	};
	// every NP virtual instruction: (instruction, direction)
	struct Virt {
		int64_t i;
		int dir;
		int32_t L, S;      // lane list, row list
		int64_t bin, bout; // input and output basis block
		int space;
	};
	std::vector<Virt> virts;
	for (int64_t i = 0; i < ni; i++) {
		const bacchus_hip_instruction_t &I = p->instructions[i];
		if (I.type != 3 || list_len(I.neutron_list) == 0 || list_len(I.proton_list) == 0)
			continue;
		for (int dir = 0; dir < (I.block_in == I.block_out ? 1 : 2); dir++) {
			Virt v;
			v.i = i;
			v.dir = dir;
			v.space = swapped[i] ? 1 : 0;
			v.L = swapped[i] ? I.proton_list : I.neutron_list;
			v.S = swapped[i] ? I.neutron_list : I.proton_list;
			v.bin = dir == 0 ? I.block_in : I.block_out;
			v.bout = dir == 0 ? I.block_out : I.block_in;
			virts.push_back(v);
		}
	}
	op->np_path = true;
	if (virts.empty())
		return 0;

	// --- SELL copies of the lane lists, per direction ---
	std::map<std::pair<int32_t, int>, size_t> sell_of;
	std::vector<Sell> sells;
	for (const Virt &v : virts)
		if (!sell_of.count({v.L, v.dir})) {
			sell_of[{v.L, v.dir}] = sells.size();
			sells.emplace_back();
		}
	std::vector<std::pair<int32_t, int>> sell_keys(sells.size());
	for (const auto &kv : sell_of)
		sell_keys[kv.second] = kv.first;
	// slices per lane list, built in parallel
	struct SliceBuild {
		std::vector<NpSlice> slices;    // off relative to the list's base
		std::vector<int32_t> rows;      // 64 per slice
		std::vector<uint32_t> in;       // positions
	};
	std::vector<SliceBuild> sb(sells.size());
	for_each_parallel(sells.size(), [&](size_t q) {
		const int32_t l = sell_keys[q].first;
		const int dir = sell_keys[q].second;
		const int64_t r0 = (dir == 0 ? *c.rows_by_out : *c.rows_by_in)[l];
		const int64_t nrows = (dir == 0 ? *c.n_out : *c.n_in)[l];
		std::vector<std::pair<uint32_t, int32_t>> rows;  // (count, state)
		uint32_t maxlen = 0;
		for (int64_t r = 0; r < nrows; r++) {
			const uint32_t cnt = rowptr[r0 + r + 1] - rowptr[r0 + r];
			if (cnt) {
				rows.push_back({cnt, (int32_t)r});
				maxlen = std::max(maxlen, cnt);
			}
		}
		(void)maxlen;
		std::stable_sort(rows.begin(), rows.end(),
				 [](const auto &a, const auto &b) { return a.first > b.first; });
		Sell &S = sells[q];
		SliceBuild &B = sb[q];
		uint32_t pos = 0;
		size_t k0 = 0;
		while (k0 < rows.size()) {
			// q = 2^lq lanes per state for the next 64 / q states (the
			// longest first): the q that wastes the smallest share of
			// positions, more states on a tie
			int lq = 0;
			double best = 1e300;
			for (int l = 0; l <= 6; l++) {
				const uint32_t qn = 1u << l;
				const size_t k1 = std::min(rows.size(), k0 + (size_t)(64u >> l));
				uint64_t used = 0;
				for (size_t k = k0; k < k1; k++)
					used += rows[k].first;
				const uint64_t steps = (rows[k0].first + qn - 1) / qn;
				const double waste = double(steps * (k1 - k0) * qn) / double(used);
				if (waste < best - 1e-12) {
					best = waste;
					lq = l;
				}
			}
			const uint32_t qn = 1u << lq;
			const size_t k1 = std::min(rows.size(), k0 + (size_t)(64u >> lq));
			const uint32_t width = (uint32_t)(k1 - k0) * qn;
			const uint32_t steps = (rows[k0].first + qn - 1) / qn;
			NpSlice s{};
			s.off = pos;
			s.steps = steps;
			s.shape = width | ((uint32_t)lq << 16);
			B.slices.push_back(s);
			for (uint32_t ln = 0; ln < 64; ln++) {
				const size_t k = k0 + ln / qn;
				B.rows.push_back(ln < width ? rows[k].second : -1);
			}
			B.in.resize(pos + (size_t)steps * width, 0u);
			S.comp.resize(pos + (size_t)steps * width, kPad);
			for (size_t k = k0; k < k1; k++) {
				const uint32_t ln0 = (uint32_t)(k - k0) * qn;
				const uint32_t a = rowptr[r0 + rows[k].second];
				// the state's entries, ordered by input state so a lane's
				// x loads ascend
				std::vector<uint2> es(entries.begin() + a,
						      entries.begin() + a + rows[k].first);
				std::stable_sort(es.begin(), es.end(),
						 [](const uint2 &u, const uint2 &w) { return u.x < w.x; });
				for (uint32_t e = 0; e < rows[k].first; e++) {
					const uint32_t t = e / qn, u = e % qn;
					const size_t at = pos + (size_t)t * width + ln0 + u;
					B.in[at] = es[e].x;
					S.comp[at] = es[e].y;
				}
			}
			pos += steps * width;
			k0 = k1;
		}
		S.len = pos;
		S.nslices = (uint32_t)B.slices.size();
	});
	std::vector<NpSlice> h_slices;
	std::vector<int32_t> h_rows;
	std::vector<uint32_t> h_in;
	{
		size_t total = 0;
		for (size_t q = 0; q < sells.size(); q++)
			total += sb[q].in.size();
		if (total > 0xFFFFFFFFull)
			return BACCHUS_FAIL("sliced lane lists exceed 2^32 positions");
		h_in.reserve(total);
		for (size_t q = 0; q < sells.size(); q++) {
			Sell &S = sells[q];
			S.base = (int64_t)h_in.size();
			S.slice0 = (uint32_t)h_slices.size();
			const uint32_t row0 = (uint32_t)h_rows.size();
			for (uint32_t k = 0; k < S.nslices; k++) {
				NpSlice sl = sb[q].slices[k];
				sl.row_base = row0 + 64u * k;
				h_slices.push_back(sl);
			}
			h_rows.insert(h_rows.end(), sb[q].rows.begin(), sb[q].rows.end());
			h_in.insert(h_in.end(), sb[q].in.begin(), sb[q].in.end());
			SliceBuild().slices.swap(sb[q].slices);
			std::vector<int32_t>().swap(sb[q].rows);
			std::vector<uint32_t>().swap(sb[q].in);
		}
	}

	// --- row-list entries grouped by compact index, per direction ---
	std::map<std::pair<int32_t, int>, size_t> jg_of;
	std::vector<std::pair<int32_t, int>> jg_keys;
	for (const Virt &v : virts)
		if (!jg_of.count({v.S, v.dir})) {
			jg_of[{v.S, v.dir}] = jg_keys.size();
			jg_keys.push_back({v.S, v.dir});
		}
	std::vector<JGroup> jgs(jg_keys.size());
	std::vector<uint2> h_j;
	std::vector<uint32_t> h_bptr;
	for (size_t q = 0; q < jg_keys.size(); q++) {
		const int32_t l = jg_keys[q].first;
		const int dir = jg_keys[q].second;
		const int64_t r0 = (dir == 0 ? *c.rows_by_out : *c.rows_by_in)[l];
		const int64_t nrows = (dir == 0 ? *c.n_out : *c.n_in)[l];
		const size_t cs = mvals[l].size();
		std::vector<uint32_t> count(cs + 1, 0);
		const uint32_t a = rowptr[r0], b = rowptr[r0 + nrows];
		for (uint32_t k = a; k < b; k++)
			count[(entries[k].y & kMask) + 1]++;
		for (size_t s = 0; s < cs; s++)
			count[s + 1] += count[s];
		const uint32_t jbase = (uint32_t)h_j.size();
		jgs[q].bbase = (uint32_t)h_bptr.size();
		for (size_t s = 0; s <= cs; s++)
			h_bptr.push_back(jbase + count[s]);
		jgs[q].count.assign(cs, 0);
		for (size_t s = 0; s < cs; s++)
			jgs[q].count[s] = count[s + 1] - count[s];
		h_j.resize(jbase + (b - a));
		std::vector<uint32_t> fill(count.begin(), count.end() - 1);
		// CSR rows are the direction's output states, entries' .x its
		// input states
		for (int64_t r = 0; r < nrows; r++)
			for (uint32_t k = rowptr[r0 + r]; k < rowptr[r0 + r + 1]; k++) {
				const uint32_t cb = entries[k].y & kMask;
				h_j[jbase + fill[cb]++] =
					make_uint2(entries[k].x, (uint32_t)r | (entries[k].y & 0x80000000u));
			}
	}

	// --- expanded element rows per (element sub-block, direction) ---
	struct WJob {
		int64_t off;
		size_t sell;
		const Virt *v;
	};
	std::map<std::tuple<int32_t, int32_t, int32_t, int>, int64_t> w_of;
	std::vector<WJob> wjobs;
	int64_t w_total = 0;
	std::vector<int64_t> virt_w(virts.size());
	for (size_t k = 0; k < virts.size(); k++) {
		const Virt &v = virts[k];
		const bacchus_hip_instruction_t &I = p->instructions[v.i];
		const auto key = std::make_tuple(I.matrix_block, I.neutron_list, I.proton_list, v.dir);
		auto it = w_of.find(key);
		if (it != w_of.end()) {
			virt_w[k] = it->second;
			continue;
		}
		const size_t sq = sell_of[{v.L, v.dir}];
		const int64_t size = (int64_t)mvals[v.S].size() * sells[sq].len;
		w_of[key] = w_total;
		virt_w[k] = w_total;
		wjobs.push_back(WJob{w_total, sq, &v});
		w_total += size;
	}
	op->np_elements = w_total;
	const int prec = op->precision;
	const size_t esz = prec == BACCHUS_HIP_FP64 ? 8 : prec == BACCHUS_HIP_FP32 ? 4 : 2;
	std::vector<char> h_w((size_t)w_total * esz);
	for_each_parallel(wjobs.size(), [&](size_t q) {
		const WJob &j = wjobs[q];
		const Virt &v = *j.v;
		const bacchus_hip_instruction_t &I = p->instructions[v.i];
		const Sell &S = sells[j.sell];
		const auto &rowvals = mvals[v.S];
		const auto &lanevals = mvals[v.L];
		const double *Vb = p->matrix_elements + p->matrix_offsets[I.matrix_block];
		const int64_t Dn = p->matrix_dn[I.matrix_block];
		const bool sw = swapped[v.i] != 0;
		for (size_t b = 0; b < rowvals.size(); b++)
			for (uint32_t pos = 0; pos < S.len; pos++) {
				const uint32_t cm = S.comp[pos];
				double val = 0.0;
				if (cm != kPad) {
					const uint32_t a = lanevals[cm & kMask];
					// element index m_n + Dn m_p
					const double e = sw ? Vb[rowvals[b] + Dn * (int64_t)a]
							    : Vb[a + Dn * (int64_t)rowvals[b]];
					val = (cm & 0x80000000u) ? -e : e;
				}
				const size_t at = (size_t)j.off + b * S.len + pos;
				if (prec == BACCHUS_HIP_FP64)
					reinterpret_cast<double *>(h_w.data())[at] = val;
				else if (prec == BACCHUS_HIP_FP32)
					reinterpret_cast<float *>(h_w.data())[at] = (float)val;
				else
					reinterpret_cast<uint16_t *>(h_w.data())[at] = bf16_of(val);
			}
	});
	for (Sell &S : sells)
		std::vector<uint32_t>().swap(S.comp);

	// --- instructions, units, tiles ---
	// work of one slice for one row-list entry: its lane-steps, a full
	// wave's worth per step whatever the width, plus the per-slice flush
	auto slice_cost = [](const NpSlice &sl) { return 64.0 * sl.steps + 64.0; };
	std::vector<NpInstr> h_ins(virts.size());
	std::vector<std::vector<NpUnit>> slot_units(2 * op->block_np.size());
	const double unit_target = [] {
		const char *e = getenv("BACCHUS_HIP_NP_UNIT");
		return e ? std::max(1024.0, atof(e)) : 65536.0;
	}();
	double total_cost = 0;
	// x rows staged in LDS (BACCHUS_HIP_NP_STAGE=0: gathered from global
	// memory through the caches)
	const bool stage_on = [] {
		const char *e = getenv("BACCHUS_HIP_NP_STAGE");
		return !e || atoi(e) != 0;
	}();
	// row-list entries of a unit at most (lane registers; smaller values
	// exercise the split of a compact index's entries)
	const uint32_t jmax = [] {
		const char *e = getenv("BACCHUS_HIP_NP_JMAX");
		const int v = e ? atoi(e) : 64;
		return (uint32_t)std::min(64, std::max(1, v));
	}();
	for (size_t k = 0; k < virts.size(); k++) {
		const Virt &v = virts[k];
		const bacchus_hip_instruction_t &I = p->instructions[v.i];
		const Sell &S = sells[sell_of[{v.L, v.dir}]];
		const JGroup &J = jgs[jg_of[{v.S, v.dir}]];
		NpInstr &n = h_ins[k];
		n.x_off = op->block_off[v.bin];
		n.w_off = virt_w[k];
		n.sell_len = S.len;
		n.sell_base = (uint32_t)S.base;
		n.slice0 = S.slice0;
		n.nslices = S.nslices;
		n.bbase = J.bbase;
		n.nl_in = (int32_t)(v.space ? op->block_np[v.bin] : op->block_nn[v.bin]);
		const int force = p->matrix_force ? p->matrix_force[I.matrix_block] : 2;
		n.force = (int16_t)std::min(3, std::max(1, force));
		// x rows staged in LDS when they fit and the lane list is long
		// enough to pay for copying them
		{
			const int fit = kNpStage / std::max(1, n.nl_in);
			n.stage = (int16_t)((stage_on && fit >= 1 && S.len >= (uint32_t)n.nl_in / 2)
						    ? std::min(kNpChunk, fit) : 0);
		}
		// cost of one b: its entries x the lane-steps of every slice
		double steps = 0;
		for (uint32_t s = 0; s < S.nslices; s++)
			steps += slice_cost(h_slices[S.slice0 + s]);
		auto &units = slot_units[(size_t)v.space * op->block_np.size() + v.bout];
		uint32_t b = 0;
		const uint32_t cs = (uint32_t)J.count.size();
		while (b < cs) {
			if (!J.count[b]) {
				b++;
				continue;
			}
			const double cb = J.count[b] * steps;
			if (cb > 2 * unit_target || S.nslices > kNpUnitSlices || J.count[b] > jmax) {
				// one b over several units: its entries in groups of at
				// most 64 (jsub = first << 16 | count), its slices in
				// groups of about unit_target
				for (uint32_t j0 = 0; j0 < J.count[b]; j0 += jmax) {
					const uint32_t jn = std::min(jmax, J.count[b] - j0);
					const uint32_t jsub = J.count[b] > jmax ? (j0 << 16 | jn) : 0u;
					uint32_t s0 = 0;
					double acc = 0;
					for (uint32_t s = 0; s < S.nslices; s++) {
						acc += jn * slice_cost(h_slices[S.slice0 + s]);
						if (acc >= unit_target || s + 1 == S.nslices ||
						    s + 1 - s0 == kNpUnitSlices) {
							units.push_back(NpUnit{(int32_t)k, b, b + 1, s0, s + 1, jsub});
							s0 = s + 1;
							acc = 0;
						}
					}
				}
				if (J.count[b] > 0xFFFFu)
					return BACCHUS_FAIL("a row-list compact index with %u entries",
							    J.count[b]);
				total_cost += cb;
				b++;
				continue;
			}
			uint32_t e = b;
			double acc = 0;
			uint32_t entries_in = 0;
			while (e < cs && acc + J.count[e] * steps <= std::max(unit_target, cb) &&
			       e - b < kNpUnitSlices - 1 && entries_in + J.count[e] <= jmax) {
				if (J.count[e] * steps > 2 * unit_target)
					break;
				acc += J.count[e] * steps;
				entries_in += J.count[e];
				e++;
			}
			units.push_back(NpUnit{(int32_t)k, b, e, 0, S.nslices, 0});
			total_cost += acc;
			b = e;
		}
	}
	// tiles: the units of one output block in groups of about tile_target
	std::vector<NpTile> h_tiles;
	std::vector<NpUnit> h_units;
	std::vector<double> tile_cost;
	const double tile_target = [&] {
		const char *e = getenv("BACCHUS_HIP_NP_TILES_PER_CU");
		const double per = e ? std::max(0.25, atof(e)) : 4.0;
		return std::max(4 * unit_target, total_cost / (op->num_cu * per));
	}();
	const int64_t nb = (int64_t)op->block_np.size();
	int64_t max_words = 1;
	std::vector<std::vector<NpUnit>> tile_units;
	for (int64_t slot = 0; slot < 2 * nb; slot++) {
		auto &us = slot_units[slot];
		if (us.empty())
			continue;
		const int space = slot >= nb;
		const int64_t bo = slot % nb;
		NpTile t{};
		t.y_off = op->block_off[bo];
		t.rows = (int32_t)(space ? op->block_nn[bo] : op->block_np[bo]);
		t.cols = (int32_t)(space ? op->block_np[bo] : op->block_nn[bo]);
		t.space = space;
		max_words = std::max<int64_t>(max_words, (int64_t)t.rows * t.cols);
		double acc = 0;
		std::vector<NpUnit> cur;
		auto cost_of = [&](const NpUnit &u) {
			const NpInstr &n = h_ins[u.vi];
			double steps = 0;
			for (uint32_t s = u.k0; s < u.k1; s++)
				steps += slice_cost(h_slices[n.slice0 + s]);
			return steps * (h_bptr[n.bbase + u.b1] - h_bptr[n.bbase + u.b0]);
		};
		for (size_t k = 0; k < us.size(); k++) {
			cur.push_back(us[k]);
			acc += cost_of(us[k]);
			if (acc >= tile_target || k + 1 == us.size()) {
				h_tiles.push_back(t);
				tile_cost.push_back(acc);
				tile_units.push_back(cur);
				cur.clear();
				acc = 0;
			}
		}
	}
	// heaviest first
	std::vector<size_t> order(h_tiles.size());
	std::iota(order.begin(), order.end(), 0);
	std::stable_sort(order.begin(), order.end(),
			 [&](size_t a, size_t b) { return tile_cost[a] > tile_cost[b]; });
	std::vector<NpTile> tiles;
	for (size_t k : order) {
		NpTile t = h_tiles[k];
		t.u_begin = (int32_t)h_units.size();
		// longest units first inside a tile too
		auto us = tile_units[k];
		h_units.insert(h_units.end(), us.begin(), us.end());
		t.u_end = (int32_t)h_units.size();
		tiles.push_back(t);
	}
	// LDS: the largest output block (FP64 sums), then the waves' x staging
	const size_t xsz = prec == BACCHUS_HIP_FP64 ? 8 : 4;
	const size_t lds = (size_t)max_words * 8 + (size_t)(kNpThreads / 64) * kNpStage * xsz;
	if (lds > 160 * 1024)
		return BACCHUS_FAIL("an output block of %lld states does not fit the LDS",
				    (long long)max_words);
	op->np_num_tiles = (int64_t)tiles.size();
	op->np_lds_bytes = lds;
	op->np_block_words = (int32_t)max_words;
	if (getenv("BACCHUS_HIP_VERBOSE"))
		fprintf(stderr, "np: %zu instructions, %zu units, %zu tiles, %lld elements, "
			"%zu positions, LDS %zu B\n", virts.size(), h_units.size(), tiles.size(),
			(long long)w_total, h_in.size(), op->np_lds_bytes);

	// host tables to files, for checking the layout without a device
	if (const char *dump = getenv("BACCHUS_HIP_NP_DUMP")) {
		auto put = [&](const char *name, const void *data, size_t bytes) {
			const std::string path = std::string(dump) + "/" + name;
			if (FILE *f = fopen(path.c_str(), "wb")) {
				fwrite(data, 1, bytes, f);
				fclose(f);
			}
		};
		put("w.bin", h_w.data(), h_w.size());
		put("in.bin", h_in.data(), h_in.size() * 4);
		put("row.bin", h_rows.data(), h_rows.size() * 4);
		put("slices.bin", h_slices.data(), h_slices.size() * sizeof(NpSlice));
		put("j.bin", h_j.data(), h_j.size() * sizeof(uint2));
		put("bptr.bin", h_bptr.data(), h_bptr.size() * 4);
		put("instr.bin", h_ins.data(), h_ins.size() * sizeof(NpInstr));
		put("units.bin", h_units.data(), h_units.size() * sizeof(NpUnit));
		put("tiles.bin", tiles.data(), tiles.size() * sizeof(NpTile));
	}

	// --- upload ---
	auto upload = [&](void **dst, const void *src, size_t bytes) -> int {
		if (bytes == 0)
			bytes = 16;
		BACCHUS_HIP_CHECK(hipMalloc(dst, bytes));
		op->device_bytes += bytes;
		if (src)
			BACCHUS_HIP_CHECK(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
		return 0;
	};
	int rc;
	if ((rc = upload(&op->d_np_w, h_w.data(), h_w.size())) ||
	    (rc = upload((void **)&op->d_np_in, h_in.data(), h_in.size() * 4)) ||
	    (rc = upload((void **)&op->d_np_row, h_rows.data(), h_rows.size() * 4)) ||
	    (rc = upload((void **)&op->d_np_slices, h_slices.data(), h_slices.size() * sizeof(NpSlice))) ||
	    (rc = upload((void **)&op->d_np_j, h_j.data(), h_j.size() * sizeof(uint2))) ||
	    (rc = upload((void **)&op->d_np_bptr, h_bptr.data(), h_bptr.size() * 4)) ||
	    (rc = upload((void **)&op->d_np_instr, h_ins.data(), h_ins.size() * sizeof(NpInstr))) ||
	    (rc = upload((void **)&op->d_np_units, h_units.data(), h_units.size() * sizeof(NpUnit))) ||
	    (rc = upload((void **)&op->d_np_tiles, tiles.data(), tiles.size() * sizeof(NpTile))))
		return rc;
	if (op->np_lds_bytes > 64 * 1024) {
		hipError_t e = hipSuccess;
		if (prec == BACCHUS_HIP_FP64)
			e = hipFuncSetAttribute((const void *)np_kernel<double>,
						hipFuncAttributeMaxDynamicSharedMemorySize,
						(int)op->np_lds_bytes);
		else if (prec == BACCHUS_HIP_FP32)
			e = hipFuncSetAttribute((const void *)np_kernel<float>,
						hipFuncAttributeMaxDynamicSharedMemorySize,
						(int)op->np_lds_bytes);
		else
			e = hipFuncSetAttribute((const void *)np_kernel<uint16_t>,
						hipFuncAttributeMaxDynamicSharedMemorySize,
						(int)op->np_lds_bytes);
		if (e != hipSuccess)
			return BACCHUS_FAIL("LDS of %zu B per workgroup refused: %s", op->np_lds_bytes,
					    hipGetErrorString(e));
	}
	return 0;
}

int np_launch(const bacchus_hip_operator *op, const void *x, const void *xT, double *y,
	      double s1, double s2, double s3, hipStream_t s)
{
	if (!op->np_num_tiles)
		return 0;
	{
		const dim3 grid((unsigned)op->np_num_tiles), block(kNpThreads);
		const NpTile *tiles = op->d_np_tiles;
		const size_t lds = op->np_lds_bytes;
		const int32_t bw = op->np_block_words;
		if (op->precision == BACCHUS_HIP_FP64)
			hipLaunchKernelGGL(np_kernel<double>, grid, block, lds, s, tiles, op->d_np_units,
					   op->d_np_instr, op->d_np_slices, op->d_np_in, op->d_np_row,
					   op->d_np_j, op->d_np_bptr, (const double *)op->d_np_w,
					   (const double *)x, (const double *)xT, y, op->d_yT, bw, s1, s2, s3);
		else if (op->precision == BACCHUS_HIP_FP32)
			hipLaunchKernelGGL(np_kernel<float>, grid, block, lds, s, tiles, op->d_np_units,
					   op->d_np_instr, op->d_np_slices, op->d_np_in, op->d_np_row,
					   op->d_np_j, op->d_np_bptr, (const float *)op->d_np_w,
					   (const float *)x, (const float *)xT, y, op->d_yT, bw, s1, s2, s3);
		else
			hipLaunchKernelGGL(np_kernel<uint16_t>, grid, block, lds, s, tiles,
					   op->d_np_units, op->d_np_instr, op->d_np_slices, op->d_np_in,
					   op->d_np_row, op->d_np_j, op->d_np_bptr,
					   (const uint16_t *)op->d_np_w, (const float *)x, (const float *)xT,
					   y, op->d_yT, bw, s1, s2, s3);
		BACCHUS_HIP_CHECK(hipGetLastError());
	}
	return 0;
}

void np_free(bacchus_hip_operator *op)
{
	(void)hipFree(op->d_np_w);
	(void)hipFree(op->d_np_in);
	(void)hipFree(op->d_np_row);
	(void)hipFree(op->d_np_slices);
	(void)hipFree(op->d_np_j);
	(void)hipFree(op->d_np_bptr);
	(void)hipFree(op->d_np_instr);
	(void)hipFree(op->d_np_units);
	(void)hipFree(op->d_np_tiles);
}

}  // namespace bacchus
