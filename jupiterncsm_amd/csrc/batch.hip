// batch.hip - the batched form of the Bacchus matvec and Lanczos: several
// operators, several vectors per operator, one persistent launch.
//
// A batch multiplies vector b of operator i by H_i(s_b) = s1 H1 + s2 V2 +
// s3 V3 (the linear combination src/Vulcan/programs/vulcan.c:114-207
// writes into new element files, applied here as per-vector scales of the
// shared element blocks).  The work units are the owner-computes tiles of
// the single-vector kernel (operator.hip) widened by bt vectors: the LDS
// tile holds rows x neutron states x bt FP32 sums, lanes are (entry slot,
// vector) pairs, so every element and list entry a lane group loads is used
// for bt vectors and the input loads are bt contiguous floats.  Tiles of all
// operators, times their vector chunks, form one list ordered heaviest
// first that a persistent grid walks with a static stride.
#include "bacchus_internal.h"

#include <type_traits>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

namespace bacchus {
namespace {

constexpr uint32_t kMask = 0x7FFFFFFFu;
constexpr int64_t kBatchTileWords = 8192;  // FP64 sums of LDS per tile (64 KiB)
constexpr int kMaxVectors = 256;
constexpr int64_t kMinTileRows = 8;

struct BatchOp {
	const TileWork *tiles;
	const Task *tasks;
	const TInstr *ins;
	const uint32_t *rowptr;
	const uint2 *ent;
	const uint32_t *keys;
	const void *V;
	const float *x;
	float *y;
	const float *xT;      // per-block transposed copies ([neutron][proton])
	float *yT;
	const float *scales;  // [4][B]: row f = scale of force class f
	int32_t B;            // vectors, padded to a multiple of 4 (row stride)
	int32_t precision;
	// P-form virtual instructions applied by batch_mfma_kernel (nonzero
	// entry per instruction) and whether that path is on for this launch
	const uint32_t *mfma_vi;
	int32_t mfma;
	int32_t pad;
};

// --- the matrix-core form of P-form instructions ---
//
// A P instruction (and an N instruction in the transposed space) with list
// entries (p_in -> p_out, element m, sign) is, over the flattened columns
// (state n, vector b) of the batched vectors [state][vector],
//   Y[p_out][(n, b)] += s_b A[p_out][p_in] X[p_in][(n, b)],
//   A[p_out][p_in] = sum of sign V[m] over the entries connecting them,
// a product of a sparse matrix with a dense panel of n_states x B columns.
// A is stored as 16 x 4 tiles: for every 16-row group of an output block,
// the input rows its entries read, four at a time ("K-groups"), with the
// 64 values in v_mfma_f32_16x16x4f32 operand order (lane l holds
// A[l % 16][l / 16]).  The scale s_b of the instruction's force class
// multiplies the X operand.
struct MfmaUnit {
	int64_t y_off;    // block offset in y (yT in the transposed space), states
	int32_t op;       // operator of the batch
	int32_t space;
	int32_t nrows;    // rows of the block in its space
	int32_t ncols;    // states per row
	int32_t r0;       // first row of the group
	int32_t g[4];     // K-groups of force class f: [g[f - 1], g[f])
};
struct MfmaKGroup {
	int64_t x_off;    // input block offset, states
	int32_t rows[4];  // input rows of A's four columns
};
// one wave's work: K-groups [g0, g1) of a unit over one column chunk
struct MfmaItem {
	int32_t unit;
	int32_t chunk;
	int32_t g0, g1;
};
constexpr int kMfmaSplit = 32;  // K-groups per item at most

struct BatchItem {
	int32_t op;
	int32_t tile;
	int32_t b0;
	int32_t lbt;  // log2 of the vectors of the tile (>= 2)
};

__device__ __forceinline__ float sign_of(float v, uint32_t m)
{
	return __int_as_float(__float_as_int(v) ^ (int)(m & 0x80000000u));
}

// The operator arrays reach the kernel through BatchOp descriptors in
// memory, so the compiler sees generic pointers and would emit flat loads
// (which also count against the LDS counter: every wait for an LDS add or
// a scalar load would then wait for all loads in flight).  Loads go through
// global-address-space pointers instead.
#if defined(__HIP_DEVICE_COMPILE__)
#define BACCHUS_GLOBAL __attribute__((address_space(1)))
#else
#define BACCHUS_GLOBAL
#endif
template <typename T> __device__ __forceinline__ const T BACCHUS_GLOBAL *global_of(const T *p)
{
	return (const T BACCHUS_GLOBAL *)p;
}

__device__ __forceinline__ float4 ld4(const float *p)
{
	return *reinterpret_cast<const float4 BACCHUS_GLOBAL *>(global_of(p));
}

__device__ __forceinline__ void fma4(float4 &a, float v, const float4 &x)
{
	a.x += v * x.x;
	a.y += v * x.y;
	a.z += v * x.z;
	a.w += v * x.w;
}

// a += v x for the vectors whose scaled element |s v| is not below 1e-12:
// N and P instructions skip small elements per vector, on the element
// scaled by that vector's force-class scale, as the reference does on the
// (Vulcan-combined) element it reads (matrix_vector_multiplication.c:
// 50-53, :105-108)
__device__ __forceinline__ void fma4_kept(float4 &a, float v, const float4 &x, const float4 &s)
{
	const float av = fabsf(v);
	a.x += (av * fabsf(s.x) < 1e-12f) ? 0.f : v * x.x;
	a.y += (av * fabsf(s.y) < 1e-12f) ? 0.f : v * x.y;
	a.z += (av * fabsf(s.z) < 1e-12f) ? 0.f : v * x.z;
	a.w += (av * fabsf(s.w) < 1e-12f) ? 0.f : v * x.w;
}

// the LDS tile holds FP64 sums (ds_add_f64)
__device__ __forceinline__ void add4_lds(double *t, const float4 &s, const float4 &v)
{
	atomicAdd(t + 0, (double)(s.x * v.x));
	atomicAdd(t + 1, (double)(s.y * v.y));
	atomicAdd(t + 2, (double)(s.z * v.z));
	atomicAdd(t + 3, (double)(s.w * v.w));
}

// Task and virtual-instruction descriptors travel as dwords in the lanes
// of one VGPR (lane i holds dword i): a wave loads the next task's ahead of
// time with one vector load and reads the fields out with readlane when it
// gets there, so no dependent descriptor load sits between two tasks.
static_assert(sizeof(Task) == 8 * sizeof(uint32_t), "Task is 8 dwords");
static_assert(sizeof(TInstr) == 12 * sizeof(uint32_t), "TInstr is 12 dwords");
__device__ __forceinline__ uint32_t task_words(const Task *tasks, int32_t ti, int lane)
{
	return lane < 8 ? global_of(reinterpret_cast<const uint32_t *>(tasks + ti))[lane] : 0u;
}
__device__ __forceinline__ uint32_t ins_words(const TInstr *ins, uint32_t vi, int lane)
{
	return lane < 12 ? global_of(reinterpret_cast<const uint32_t *>(ins + vi))[lane] : 0u;
}
template <typename S> __device__ __forceinline__ S unpack_words(uint32_t w)
{
	uint32_t d[sizeof(S) / sizeof(uint32_t)];
#pragma unroll
	for (int i = 0; i < (int)(sizeof(S) / sizeof(uint32_t)); i++)
		d[i] = __builtin_amdgcn_readlane(w, i);
	S s;
	__builtin_memcpy(&s, d, sizeof(S));
	return s;
}

template <typename VT> __device__ __forceinline__ float elem_at(__amdgpu_buffer_rsrc_t r, uint32_t off);
template <> __device__ __forceinline__ float elem_at<float>(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
	return buf_f32(r, off);
}
template <> __device__ __forceinline__ float elem_at<uint16_t>(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
	return buf_bf16(r, off);
}

template <typename VT> struct Load;
template <> struct Load<float> {
	__device__ static float at(const float *p, size_t i) { return global_of(p)[i]; }
};
template <> struct Load<uint16_t> {
	__device__ static float at(const uint16_t *p, size_t i)
	{
		return __uint_as_float(uint32_t(global_of(p)[i]) << 16);
	}
};

// One work item: tile w of operator o for vectors [b0, b0 + bt).  Every
// lane carries 4 consecutive vectors (one 16-byte load of x per product
// term, each element value used for 4 FMAs); the lanes of a tile form
// (entry slot, vector group) pairs: bt / 4 lanes per slot, 64 / (bt / 4)
// slots per wave.
//   NP (matrix_vector_multiplication.c:134-210): slots over lane-side
//     entries; row-side entries of the tile's rows are broadcast with
//     readlane; sums go to the LDS tile when the row changes.
//   P (:79-132): slots over columns, same broadcast; for narrow rows,
//     slots of column groups each walking its own entries.
//   N (:20-77) runs in the transposed space as P there.
// N and P skip elements with |s V| < 1e-12 per vector (s its force-class
// scale), as the reference does on the element it reads (:50-53, :105-108).
template <typename VT, int kAhead>
__device__ void batch_tile(const BatchOp &o, const TileWork &w, int b0, int lbt,
			   double *tile)
{
	const int bt = 1 << lbt;
	const int lq = lbt - 2;       // log2(lanes per slot)
	const int nrow = w.po_end - w.po_begin;
	const int nn = w.nn_out;
	// an output element's bt sums at a pitch of bt + 1 doubles, so the
	// lanes' LDS adds to different outputs fall into different banks
	const int pitch = (1 << lbt) + 1;
	const int twords = nrow * nn * pitch;
	float *ys = w.space ? o.yT : o.y;
	for (int t = threadIdx.x; t < twords; t += blockDim.x)
		tile[t] = 0.0;
	__syncthreads();
	const int lane = threadIdx.x & 63;
	const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const int nwaves = blockDim.x >> 6;
	const int slot = lane >> lq;
	const int G = 64 >> lq;       // entry slots per pass
	const int q4 = (lane & ((1 << lq) - 1)) * 4;  // first vector of the lane in the tile
	const int B = o.B;
	const int b = b0 + q4;
	const bool bvalid = b < B;    // B is a multiple of 4: all 4 or none
	const VT *V = (const VT *)o.V;
	// uniform, so the input-row addresses of the NP loop stay scalar
	const float *xs = reinterpret_cast<const float *>(
		uniform_ptr(reinterpret_cast<const char *>(w.space ? o.xT : o.x)));
	// descriptor words of the current task, its instruction, and the next
	// task (the next task's instruction is loaded when this one starts)
	int32_t ti = w.t_begin + wave;
	uint32_t tw_cur = 0, iw_cur = 0, tw_nxt = 0;
	if (ti < w.t_end) {
		tw_cur = task_words(o.tasks, ti, lane);
		iw_cur = ins_words(o.ins, __builtin_amdgcn_readlane(tw_cur, 0), lane);
		if (ti + nwaves < w.t_end)
			tw_nxt = task_words(o.tasks, ti + nwaves, lane);
	}
	for (; ti < w.t_end; ti += nwaves) {
		const Task T = unpack_words<Task>(tw_cur);
		const TInstr I = unpack_words<TInstr>(iw_cur);
		uint32_t iw_nxt = 0, tw_nn = 0;
		if (ti + nwaves < w.t_end) {
			iw_nxt = ins_words(o.ins, __builtin_amdgcn_readlane(tw_nxt, 0), lane);
			if (ti + 2 * nwaves < w.t_end)
				tw_nn = task_words(o.tasks, ti + 2 * nwaves, lane);
		}
		const float4 sc = bvalid ? ld4(o.scales + (size_t)I.force * B + b)
					 : make_float4(0.f, 0.f, 0.f, 0.f);
		// lane's vectors of the input block: element e at xb + e * B
		const float *xb = xs + (size_t)I.x_off * B + (bvalid ? b : 0);
		const uint32_t nn_in = (uint32_t)I.nn_in;
		if (I.type == 3) {
			// every lane holds kLaneEntries lane-side entries (slot
			// chunks), so a broadcast row-side entry puts 2 kLaneEntries
			// independent loads in flight; the loads of row-side entries
			// t + 1 .. t + kAhead are issued before the products of entry
			// t, so kAhead + 1 entries' loads overlap (the element rows
			// stream from HBM)
			const VT *vblk = V + I.v_off;
			const uint32_t dn = (uint32_t)I.dn;
			// list position of the lane's entry in chunk c of the pass at
			// e0 (the task's first entry where the chunk runs past its end)
			auto lane_k = [&](uint32_t e0, int c) {
				const uint32_t e = e0 + (uint32_t)(c * G + slot);
				return I.n_base + (e < T.e1 ? e : T.e0);
			};
			// with two row entries ahead, the lane-side entries of the
			// next pass are loaded while this pass runs its rows
			uint2 en_nx[kLaneEntries];
			uint32_t key_nx[kLaneEntries];
			if constexpr (kAhead == 2) {
#pragma unroll
				for (int c = 0; c < kLaneEntries; c++) {
					const uint32_t k = lane_k(T.e0, c);
					en_nx[c] = global_of(o.ent)[k];
					key_nx[c] = global_of(o.keys)[k];
				}
			}
			for (uint32_t e0 = T.e0; e0 < T.e1; e0 += kLaneEntries * G) {
				// slot chunks that hold any entry (wave-uniform); the
				// loads of empty chunks are skipped
				const int nc = (int)min((uint32_t)kLaneEntries, (T.e1 - e0 + G - 1) / (uint32_t)G);
				bool vl[kLaneEntries];
				// byte offsets of the lane's element (from the element
				// row) and of its 4 input values (from the input row); the
				// lane's tile offset within a row (output column x pitch +
				// first vector) and its phase bit
				uint32_t colb[kLaneEntries], xob[kLaneEntries], outc[kLaneEntries];
				int lidx[kLaneEntries];
				uint32_t lsb[kLaneEntries];
				float4 acc[kLaneEntries];
				const uint32_t bl = bvalid ? (uint32_t)b : 0u;
				uint2 en_c[kLaneEntries];
#pragma unroll
				for (int c = 0; c < kLaneEntries; c++) {
					if constexpr (kAhead == 2) {
						en_c[c] = en_nx[c];
						outc[c] = key_nx[c];
					} else {
						const uint32_t k = lane_k(e0, c);
						en_c[c] = global_of(o.ent)[k];
						outc[c] = global_of(o.keys)[k];
					}
				}
				if constexpr (kAhead == 2) {
					const uint32_t en0 = e0 + kLaneEntries * G;
					if (en0 < T.e1) {
#pragma unroll
						for (int c = 0; c < kLaneEntries; c++) {
							const uint32_t k = lane_k(en0, c);
							en_nx[c] = global_of(o.ent)[k];
							key_nx[c] = global_of(o.keys)[k];
						}
					}
				}
#pragma unroll
				for (int c = 0; c < kLaneEntries; c++) {
					const uint32_t e = e0 + (uint32_t)(c * G + slot);
					vl[c] = e < T.e1 && bvalid;
					const uint32_t k = lane_k(e0, c);
					const uint2 en = en_c[c];
					// element column: compact index, or lane-list
					// position for expanded element rows
					colb[c] = (I.expanded ? k - I.n_base : en.y & kMask) * (uint32_t)sizeof(VT);
					xob[c] = (en.x * (uint32_t)B + bl) * 4u;
					lidx[c] = (int)outc[c] * pitch + q4;
					lsb[c] = en.y & 0x80000000u;
					acc[c] = make_float4(0.f, 0.f, 0.f, 0.f);
				}
				auto rows_pass = [&](auto full) {
				constexpr bool kFull = decltype(full)::value;
				float va[kAhead + 1][kLaneEntries];
				float4 xa[kAhead + 1][kLaneEntries];
				uint2 pl;
				// loads of row-side entry t (lane t of pl) into buffer buf
				auto issue = [&](uint32_t t, auto buf) {
					constexpr int k = decltype(buf)::value;
					const uint32_t px = __builtin_amdgcn_readlane(pl.x, t);
					const uint32_t py = __builtin_amdgcn_readlane(pl.y, t);
					// element row and input row: uniform bases
					// a lane's column (compact index or lane-list
					// position) is below the lane list's length
					const __amdgpu_buffer_rsrc_t vr = rsrc_of(
						reinterpret_cast<const char *>(vblk + (size_t)dn * (py & kMask)),
						(uint32_t)I.n_len * (uint32_t)sizeof(VT));
					const __amdgpu_buffer_rsrc_t xr = rsrc_of(
						reinterpret_cast<const char *>(
							xs + ((size_t)I.x_off + (size_t)nn_in * px) * B),
						nn_in * (uint32_t)B * 4u);
#pragma unroll
					for (int c = 0; c < kLaneEntries; c++) {
						if (kFull || c < nc) {
							va[k][c] = elem_at<VT>(vr, colb[c]);
							xa[k][c] = buf_f32x4(xr, xob[c]);
						} else {
							va[k][c] = 0.f;
							xa[k][c] = make_float4(0.f, 0.f, 0.f, 0.f);
						}
					}
				};
				int cur = -1;
				// the sums carry both phases (row entry and lane entry), so
				// a flush scales them by the vectors' force scales only
				auto flush = [&]() {
					const int rb = (cur - w.po_begin) * nn * pitch;
#pragma unroll
					for (int c = 0; c < kLaneEntries; c++)
						if (vl[c])
							add4_lds(&tile[rb + lidx[c]], sc, acc[c]);
				};
				// products of entry t from buffer buf, after a flush when
				// its row differs from the current one
				auto consume = [&](uint32_t t, uint32_t prow, auto buf) {
					constexpr int k = decltype(buf)::value;
					const int row = (int)__builtin_amdgcn_readlane(prow, t);
					if (row != cur) {
						if (cur >= 0)
							flush();
#pragma unroll
						for (int c = 0; c < kLaneEntries; c++)
							acc[c] = make_float4(0.f, 0.f, 0.f, 0.f);
						cur = row;
					}
					const uint32_t ps = __builtin_amdgcn_readlane(pl.y, t) & 0x80000000u;
#pragma unroll
					for (int c = 0; c < kLaneEntries; c++)
						fma4(acc[c], __int_as_float(__float_as_int(va[k][c]) ^ (int)(ps ^ lsb[c])),
						     xa[k][c]);
				};
				using B0 = std::integral_constant<int, 0>;
				using B1 = std::integral_constant<int, 1>;
				using B2 = std::integral_constant<int, kAhead == 2 ? 2 : 0>;
				for (uint32_t j0 = T.jb; j0 < T.je; j0 += 64) {
					const uint32_t cnt = min(64u, T.je - j0);
					const uint32_t jl = j0 + min((uint32_t)lane, cnt - 1);
					pl = global_of(o.ent)[jl];
					const uint32_t prow = global_of(o.keys)[jl];
					issue(0u, B0{});
					if constexpr (kAhead == 1) {
						for (uint32_t t = 0; t < cnt; t += 2) {
							if (t + 1 < cnt)
								issue(t + 1, B1{});
							consume(t, prow, B0{});
							if (t + 1 < cnt) {
								if (t + 2 < cnt)
									issue(t + 2, B0{});
								consume(t + 1, prow, B1{});
							}
						}
					} else {
						// the loads of entries t + 1 and t + 2 in flight
						// while the products of t are formed
						if (1 < cnt)
							issue(1u, B1{});
						for (uint32_t t = 0; t < cnt; t += 3) {
							if (t + 2 < cnt)
								issue(t + 2, B2{});
							consume(t, prow, B0{});
							if (t + 1 < cnt) {
								if (t + 3 < cnt)
									issue(t + 3, B0{});
								consume(t + 1, prow, B1{});
							}
							if (t + 2 < cnt) {
								if (t + 4 < cnt)
									issue(t + 4, B1{});
								consume(t + 2, prow, B2{});
							}
						}
					}
				}
				if (cur >= 0)
					flush();
				};
				if (nc == kLaneEntries)
					rows_pass(std::true_type{});
				else
					rows_pass(std::false_type{});
			}
		} else if (I.type == 2 && o.mfma && global_of(o.mfma_vi)[T.vi]) {
			// applied on the matrix cores by batch_mfma_kernel
		} else if (I.type == 2 && 2 * (T.e1 - T.e0) <= (uint32_t)G) {
			// narrow rows: slots of C >= (columns) lanes groups, C a power
			// of two, G / C of them; slot g walks entries g, g + G / C, ...
			// of [jb, je), so its x loads are one contiguous run of
			// columns x vectors; sums go to the tile on row changes
			const uint32_t ncol = T.e1 - T.e0;
			const int lc = ncol <= 1 ? 0 : 32 - __builtin_clz(ncol - 1);
			const uint32_t Gs = (uint32_t)(G >> lc);
			const uint32_t g = (uint32_t)(slot >> lc);
			const uint32_t c = (uint32_t)slot & ((1u << lc) - 1u);
			const bool valid = c < ncol && bvalid;
			const uint32_t nl = T.e0 + (c < ncol ? c : 0u);
			float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
			int cur = -1;
			for (uint32_t j0 = T.jb + g; j0 < T.je; j0 += 2u * Gs) {
				uint2 en[2];
				int rw[2];
				float e[2];
				float4 xv[2];
#pragma unroll
				for (int u = 0; u < 2; u++) {
					const uint32_t j = j0 + (uint32_t)u * Gs;
					const bool ok = j < T.je;
					en[u] = global_of(o.ent)[ok ? j : T.jb];
					rw[u] = ok ? (int)global_of(o.keys)[j] : -1;
				}
#pragma unroll
				for (int u = 0; u < 2; u++) {
					e[u] = Load<VT>::at(V, (size_t)I.v_off + (en[u].y & kMask));
					xv[u] = ld4(xb + ((size_t)nn_in * en[u].x + nl) * B);
				}
#pragma unroll
				for (int u = 0; u < 2; u++) {
					if (rw[u] < 0)
						break;
					if (rw[u] != cur) {
						if (cur >= 0 && valid)
							add4_lds(&tile[((cur - w.po_begin) * nn + (int)nl) * pitch + q4], sc, a);
						a = make_float4(0.f, 0.f, 0.f, 0.f);
						cur = rw[u];
					}
					fma4_kept(a, sign_of(e[u], en[u].y), xv[u], sc);
				}
			}
			if (cur >= 0 && valid)
				add4_lds(&tile[((cur - w.po_begin) * nn + (int)nl) * pitch + q4], sc, a);
		} else if (I.type == 2) {
			for (int n0 = (int)T.e0; n0 < (int)T.e1; n0 += G) {
				const int n = n0 + slot;
				const bool valid = n < (int)T.e1 && bvalid;
				const uint32_t nl = n < (int)T.e1 ? (uint32_t)n : T.e0;
				float4 a0 = make_float4(0.f, 0.f, 0.f, 0.f);
				int cur = -1;
				for (uint32_t j0 = T.jb; j0 < T.je; j0 += 64) {
					const uint32_t cnt = min(64u, T.je - j0);
					const uint32_t jl = j0 + min((uint32_t)lane, cnt - 1);
					const uint2 pl = global_of(o.ent)[jl];
					const uint32_t prow = global_of(o.keys)[jl];
					for (uint32_t t = 0; t < cnt; t++) {
						const int row = (int)__builtin_amdgcn_readlane(prow, t);
						if (row != cur) {
							if (cur >= 0 && valid)
								add4_lds(&tile[((cur - w.po_begin) * nn + n) * pitch + q4], sc, a0);
							a0 = make_float4(0.f, 0.f, 0.f, 0.f);
							cur = row;
						}
						const uint32_t px = __builtin_amdgcn_readlane(pl.x, t);
						const uint32_t py = __builtin_amdgcn_readlane(pl.y, t);
						const float e = Load<VT>::at(V, (size_t)I.v_off + (py & kMask));
						fma4_kept(a0, sign_of(e, py), ld4(xb + ((size_t)nn_in * px + nl) * B), sc);
					}
				}
				if (cur >= 0 && valid)
					add4_lds(&tile[((cur - w.po_begin) * nn + n) * pitch + q4], sc, a0);
			}
		}
		tw_cur = tw_nxt;
		iw_cur = iw_nxt;
		tw_nxt = tw_nn;
	}
	__syncthreads();
	float *yblk = ys + ((size_t)w.y_off + (size_t)nn * w.po_begin) * B + b0;
	for (int t = threadIdx.x; t < twords; t += blockDim.x) {
		const int e = t / pitch, c = t - e * pitch;
		if (c >= bt || b0 + c >= B)
			continue;
		const float v = (float)tile[t];
		float *dst = yblk + (size_t)e * B + c;
		if (w.shared) {
			if (v != 0.f)
				atomicAdd(dst, v);
		} else
			*dst += v;
	}
}

// Persistent grid over nq work queues (items qbegin[q]..qbegin[q + 1] of
// queue q, heaviest first): a workgroup takes items from the queue of its
// XCD (workgroup i runs on XCD i mod kXcds) until it is exhausted, then
// from the others in turn.  The queue heads next[q] are cleared by
// reset_queue before every launch.
template <int kAhead>
__device__ __forceinline__ void batch_body(const BatchOp *__restrict__ ops,
					   const BatchItem *__restrict__ items,
					   const int *__restrict__ qbegin, int nq, int *__restrict__ next)
{
	extern __shared__ double tile[];
	__shared__ int s_item;
	const int home = (int)(blockIdx.x % kXcds) % nq;
	for (int qi = 0; qi < nq; qi++) {
		const int q = (home + qi) % nq;
		const int base = qbegin[q], count = qbegin[q + 1] - base;
		for (;;) {
			if (threadIdx.x == 0)
				s_item = atomicAdd(next + q, 1);
			__syncthreads();
			const int it = __builtin_amdgcn_readfirstlane(s_item);
			__syncthreads();
			if (it >= count)
				break;
			const BatchItem bi = items[base + it];
			const BatchOp o = ops[bi.op];
			const TileWork w = o.tiles[bi.tile];
			if (o.precision == BACCHUS_HIP_BF16)
				batch_tile<uint16_t, kAhead>(o, w, bi.b0, bi.lbt, tile);
			else
				batch_tile<float, kAhead>(o, w, bi.b0, bi.lbt, tile);
			__syncthreads();
		}
	}
}

// The same kernel compiled twice: with the registers the compiler chooses
// and the NP row loop loading two row entries ahead (152 VGPRs, three waves
// per SIMD; one entry ahead: 6.64 ms for the C3 8-vector block against
// 6.38 ms), and held to 128 VGPRs for a fourth wave with one entry ahead;
// the autotuner keeps the faster for the batch.
__global__ __launch_bounds__(kTileThreads) void batch_kernel(
	const BatchOp *__restrict__ ops, const BatchItem *__restrict__ items,
	const int *__restrict__ qbegin, int nq, int *__restrict__ next)
{
	batch_body<2>(ops, items, qbegin, nq, next);
}
__global__ __launch_bounds__(kTileThreads, 4) void batch_kernel_w4(
	const BatchOp *__restrict__ ops, const BatchItem *__restrict__ items,
	const int *__restrict__ qbegin, int nq, int *__restrict__ next)
{
	batch_body<1>(ops, items, qbegin, nq, next);
}

typedef float mfma_f32x4 __attribute__((ext_vector_type(4)));

// One wave of batch_mfma_kernel: K-groups [g0, g1) of unit u over the
// column chunk `chunk` of 64 NQ flattened columns (n, b) of its 16 rows,
// with v_mfma_f32_16x16x4f32 (lane l: A[l % 16][l / 16], X[l / 16][l % 16];
// result rows 4 (l / 16) + j, column l % 16).  The 4 NQ panels of a chunk
// are interleaved: column j of panel (q, e) is column 64 q + 4 j + e of the
// chunk, so lane l loads its four panels' X values of quad q as one float4
// (columns 64 q + 4 (l % 16) + 0..3) and adds their results to y as one
// float4.  The loads of K-group g + 1 are issued before the products of g.
// Units whose K-groups are split over several items add atomically.
template <int NQ>
__device__ void mfma_item(const BatchOp &o, const MfmaUnit &u, int chunk, int g0, int g1,
			  bool split, const MfmaKGroup *__restrict__ kgroups,
			  const float *__restrict__ A)
{
	const int lane = threadIdx.x & 63;
	const int Bp = o.B;
	const int64_t width = (int64_t)u.ncols * Bp;  // floats per row
	const float *xs = u.space ? o.xT : o.x;
	float *ys = u.space ? o.yT : o.y;
	const int kq = lane >> 4;
	// Bp is a multiple of 4: a quad of columns is all inside the row or
	// all outside, and its 4 vectors are consecutive
	int64_t c0[NQ];
	bool cv[NQ];
	mfma_f32x4 acc[4 * NQ];
#pragma unroll
	for (int q = 0; q < NQ; q++) {
		c0[q] = (int64_t)chunk * 64 * NQ + 64 * q + 4 * (lane & 15);
		cv[q] = c0[q] < width;
	}
#pragma unroll
	for (int k = 0; k < 4 * NQ; k++)
		acc[k] = mfma_f32x4{0.f, 0.f, 0.f, 0.f};
	auto xrow_of = [&](int g) {
		const MfmaKGroup K = kgroups[g];
		const int row = kq == 0 ? K.rows[0] : kq == 1 ? K.rows[1] : kq == 2 ? K.rows[2] : K.rows[3];
		return xs + (size_t)(K.x_off + (int64_t)row * u.ncols) * Bp;
	};
	// K-groups by force class, the class's scales on the X operand
	for (int f = 1; f <= 3; f++) {
		int g = max(g0, u.g[f - 1]);
		const int gend = min(g1, u.g[f]);
		if (g >= gend)
			continue;
		float4 sc[NQ];
#pragma unroll
		for (int q = 0; q < NQ; q++)
			sc[q] = ld4(o.scales + (size_t)f * Bp + (int)(c0[q] % Bp));
		float a = A[(size_t)g * 64 + lane];
		float4 xv[NQ];
		{
			const float *xr = xrow_of(g);
#pragma unroll
			for (int q = 0; q < NQ; q++)
				xv[q] = cv[q] ? ld4(xr + c0[q]) : make_float4(0.f, 0.f, 0.f, 0.f);
		}
		for (; g < gend; g++) {
			float an = 0.f;
			float4 xn[NQ];
			if (g + 1 < gend) {
				an = A[(size_t)(g + 1) * 64 + lane];
				const float *xr = xrow_of(g + 1);
#pragma unroll
				for (int q = 0; q < NQ; q++)
					xn[q] = cv[q] ? ld4(xr + c0[q]) : make_float4(0.f, 0.f, 0.f, 0.f);
			}
#pragma unroll
			for (int q = 0; q < NQ; q++) {
				acc[4 * q + 0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, xv[q].x * sc[q].x, acc[4 * q + 0], 0, 0, 0);
				acc[4 * q + 1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, xv[q].y * sc[q].y, acc[4 * q + 1], 0, 0, 0);
				acc[4 * q + 2] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, xv[q].z * sc[q].z, acc[4 * q + 2], 0, 0, 0);
				acc[4 * q + 3] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, xv[q].w * sc[q].w, acc[4 * q + 3], 0, 0, 0);
			}
			if (g + 1 < gend) {
				a = an;
#pragma unroll
				for (int q = 0; q < NQ; q++)
					xv[q] = xn[q];
			}
		}
	}
#pragma unroll
	for (int q = 0; q < NQ; q++) {
		if (!cv[q])
			continue;
#pragma unroll
		for (int j = 0; j < 4; j++) {
			const int row = u.r0 + 4 * kq + j;
			if (row >= u.nrows)
				continue;
			float *dst = ys + (size_t)(u.y_off + (int64_t)row * u.ncols) * Bp + c0[q];
			if (split) {
				atomicAdd(dst + 0, acc[4 * q + 0][j]);
				atomicAdd(dst + 1, acc[4 * q + 1][j]);
				atomicAdd(dst + 2, acc[4 * q + 2][j]);
				atomicAdd(dst + 3, acc[4 * q + 3][j]);
			} else {
				float4 v = *reinterpret_cast<float4 *>(dst);
				v.x += acc[4 * q + 0][j];
				v.y += acc[4 * q + 1][j];
				v.z += acc[4 * q + 2][j];
				v.w += acc[4 * q + 3][j];
				*reinterpret_cast<float4 *>(dst) = v;
			}
		}
	}
}

// The P-form instructions of the operators with the matrix-core path on,
// after batch_kernel (which skips them) and before the transposed blocks
// are added back; one wave per item, chunks of 128 columns for rows wider
// than 64 columns, of 64 otherwise.
__global__ __launch_bounds__(256) void batch_mfma_kernel(
	const BatchOp *__restrict__ ops, const MfmaUnit *__restrict__ units,
	const MfmaKGroup *__restrict__ kgroups, const float *__restrict__ A,
	const MfmaItem *__restrict__ items, int nitems)
{
	const int it = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
	if (it >= nitems)
		return;
	const MfmaItem item = items[it];
	const MfmaUnit u = units[item.unit];
	const BatchOp o = ops[u.op];
	if (!o.mfma)
		return;
	const bool split = item.g0 > u.g[0] || item.g1 < u.g[3];
	if ((int64_t)u.ncols * o.B > 64)
		mfma_item<2>(o, u, item.chunk, item.g0, item.g1, split, kgroups, A);
	else
		mfma_item<1>(o, u, item.chunk, item.g0, item.g1, split, kgroups, A);
}

// dst[r][b] = src[r][b] (b < B), 0 for the padding vectors b in [B, Bp)
__global__ void pad_cols(const float *__restrict__ src, int64_t D, int B, int Bp,
			 float *__restrict__ dst)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= D * Bp)
		return;
	const int64_t r = i / Bp;
	const int b = (int)(i - r * Bp);
	dst[i] = b < B ? src[r * B + b] : 0.f;
}

__global__ void unpad_cols(const float *__restrict__ src, int64_t D, int B, int Bp,
			   float *__restrict__ dst)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= D * B)
		return;
	const int64_t r = i / B;
	dst[i] = src[r * Bp + (i - r * B)];
}

__global__ void reset_queue(int *next, int nq)
{
	if (blockIdx.x == 0 && (int)threadIdx.x < nq)
		next[threadIdx.x] = 0;
}

__global__ void fill_float(float *p, int64_t n, float v)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		p[i] = v;
}

int blocks_of(int64_t n) { return (int)((n + 255) / 256); }

int pow2_floor(int v)
{
	int p = 1;
	while (p * 2 <= v)
		p *= 2;
	return p;
}

}  // namespace
}  // namespace bacchus

using namespace bacchus;

struct bacchus_hip_batch {
	std::vector<bacchus_hip_operator_t *> ops;
	std::vector<int32_t> B;   // vectors per operator
	std::vector<int32_t> Bp;  // padded to a multiple of 4 (row stride)
	bacchus_hip_batch_tuning_t tuning{};
	std::vector<float *> d_scales;
	std::vector<TileWork *> d_tiles;
	std::vector<Task *> d_tasks;
	std::vector<BatchOp> h_ops;
	BatchOp *d_ops = nullptr;
	BatchItem *d_items = nullptr;
	int *d_next = nullptr;    // work-queue heads (kXcds)
	int *d_qbegin = nullptr;  // first item of every queue (nq + 1)
	int nq = 1;
	int64_t num_items = 0;
	size_t lds_bytes = 0;
	int grid = 0;
	// internal [D][Bp] vector buffers (padding, autotuning, Lanczos)
	std::vector<float *> d_x, d_y;
	// per-operator transposed copies of x and y (operators with
	// transposed-space instructions)
	std::vector<float *> d_xT, d_yT;
	double bytes = 0, flops = 0;
	// matrix-core path of the P-form instructions (batch_mfma_kernel):
	// off with BACCHUS_HIP_BATCH_MFMA=0 at creation; per operator, on
	// while its scales allow it (mfma_scales_ok)
	bool mfma_enabled = false;
	std::vector<uint32_t *> d_mfma_vi;
	std::vector<char> mfma_has, mfma_scales_ok;
	MfmaUnit *d_mfma_units = nullptr;
	MfmaKGroup *d_mfma_kg = nullptr;
	float *d_mfma_a = nullptr;
	MfmaItem *d_mfma_items = nullptr;
	int64_t mfma_items = 0;
	int64_t mfma_kgroups = 0;    // K-groups (one MFMA per 16 columns each)
	double mfma_entries = 0;     // list entries they hold
	double mfma_columns = 0;     // sum over units of K-groups x columns / 16

	void release_plan()
	{
		for (auto *p : d_tiles)
			(void)hipFree(p);
		for (auto *p : d_tasks)
			(void)hipFree(p);
		(void)hipFree(d_items);
		(void)hipFree(d_qbegin);
		d_tiles.clear();
		d_tasks.clear();
		d_items = nullptr;
		d_qbegin = nullptr;
	}
	~bacchus_hip_batch()
	{
		release_plan();
		for (auto *p : d_scales)
			(void)hipFree(p);
		for (auto *p : d_x)
			(void)hipFree(p);
		for (auto *p : d_y)
			(void)hipFree(p);
		for (auto *p : d_xT)
			(void)hipFree(p);
		for (auto *p : d_yT)
			(void)hipFree(p);
		(void)hipFree(d_ops);
		(void)hipFree(d_next);
		(void)hipFree(workspace);
		for (auto *p : d_mfma_vi)
			(void)hipFree(p);
		(void)hipFree(d_mfma_units);
		(void)hipFree(d_mfma_kg);
		(void)hipFree(d_mfma_a);
		(void)hipFree(d_mfma_items);
	}
	// device workspace of the block solver, kept between calls (grown on
	// demand); calls on one batch must not overlap
	void *workspace = nullptr;
	size_t workspace_bytes = 0;
};

namespace bacchus {
namespace {

int upload_ops(bacchus_hip_batch_t *bt)
{
	BACCHUS_HIP_CHECK(hipMemcpy(bt->d_ops, bt->h_ops.data(),
				    bt->h_ops.size() * sizeof(BatchOp), hipMemcpyHostToDevice));
	return 0;
}

// Scales with which a P-form instruction may run on the matrix cores: the
// reference skips a product when the (Vulcan-combined) element s V has
// |s V| < 1e-12 (matrix_vector_multiplication.c:50-53, :105-108), which a
// shared A cannot do per vector.  The matrix-core path is therefore used
// only where no product can be skipped: instructions whose nonzero elements
// all have |V| >= kMfmaMinElement (build_mfma), under scales that are 0
// or have |s| >= kMfmaMinScale (the products of a zero scale are zero
// either way).
constexpr float kMfmaMinElement = 2e-9f;
constexpr float kMfmaMinScale = 1e-3f;

bool mfma_scales_allowed(const std::vector<float> &sc)
{
	for (float s : sc)
		if (s != 0.f && !(std::fabs(s) >= kMfmaMinScale))
			return false;
	return true;
}

// The matrix-core tables of every operator's P-form virtual instructions,
// built on the host from the operator's CSR row pointers and its entries
// and elements read back from the device.  Instructions with an element
// below kMfmaMinElement stay in batch_kernel.
int build_mfma(bacchus_hip_batch_t *bt)
{
	const char *env = getenv("BACCHUS_HIP_BATCH_MFMA");
	bt->mfma_enabled = !(env && atoi(env) == 0);
	const int nops = (int)bt->ops.size();
	bt->d_mfma_vi.assign(nops, nullptr);
	bt->mfma_has.assign(nops, 0);
	bt->mfma_scales_ok.assign(nops, 1);
	std::vector<MfmaUnit> units;
	std::vector<MfmaKGroup> kgs;
	std::vector<double> A;
	double entries_used = 0;
	for (int i = 0; i < nops; i++) {
		const bacchus_hip_operator_t *op = bt->ops[i];
		const int64_t nb = (int64_t)op->block_nn.size();
		std::vector<uint32_t> elig(std::max<size_t>(1, op->h_tinstr.size()), 0);
		for (int64_t slot = 0; bt->mfma_enabled && slot < 2 * nb; slot++) {
			const int64_t b = slot % nb;
			const int space = slot >= nb;
			const int64_t nrows = space ? op->block_nn[b] : op->block_np[b];
			const int64_t ncols = space ? op->block_np[b] : op->block_nn[b];
			struct PInstr {
				int32_t v;
				std::vector<uint2> ent;
				std::vector<float> val;
			};
			std::vector<PInstr> pis;
			for (int32_t v = op->blk_vi[slot]; v < op->blk_vi[slot + 1]; v++) {
				const TInstr &t = op->h_tinstr[v];
				if (t.type != 2 || t.nn_in != ncols)
					continue;
				const uint32_t e0 = op->h_rowptr[t.p_rows];
				const uint32_t e1 = op->h_rowptr[t.p_rows + nrows];
				if (e1 <= e0)
					continue;
				PInstr pi;
				pi.v = v;
				pi.ent.resize(e1 - e0);
				BACCHUS_HIP_CHECK(hipMemcpy(pi.ent.data(), op->d_entries + e0,
							    (e1 - e0) * sizeof(uint2), hipMemcpyDeviceToHost));
				uint32_t mmax = 0;
				for (const uint2 &en : pi.ent)
					mmax = std::max(mmax, en.y & kMask);
				pi.val.resize((size_t)mmax + 1);
				if (op->precision == BACCHUS_HIP_FP32) {
					BACCHUS_HIP_CHECK(hipMemcpy(pi.val.data(), (const float *)op->d_v + t.v_off,
								    pi.val.size() * sizeof(float),
								    hipMemcpyDeviceToHost));
				} else {
					std::vector<uint16_t> h(pi.val.size());
					BACCHUS_HIP_CHECK(hipMemcpy(h.data(), (const uint16_t *)op->d_v + t.v_off,
								    h.size() * sizeof(uint16_t),
								    hipMemcpyDeviceToHost));
					for (size_t k = 0; k < h.size(); k++) {
						const uint32_t u = uint32_t(h[k]) << 16;
						memcpy(&pi.val[k], &u, 4);
					}
				}
				bool ok = true;
				for (const uint2 &en : pi.ent) {
					const float a = std::fabs(pi.val[en.y & kMask]);
					if (a != 0.f && !(a >= kMfmaMinElement))
						ok = false;
				}
				if (!ok)
					continue;
				elig[v] = 1;
				pis.push_back(std::move(pi));
			}
			std::stable_sort(pis.begin(), pis.end(), [&](const PInstr &x, const PInstr &y) {
				return op->h_tinstr[x.v].force < op->h_tinstr[y.v].force;
			});
			for (int64_t r0 = 0; r0 < nrows && !pis.empty(); r0 += 16) {
				const int64_t r1 = std::min<int64_t>(nrows, r0 + 16);
				MfmaUnit un{};
				un.y_off = op->block_off[b];
				un.op = i;
				un.space = space;
				un.nrows = (int32_t)nrows;
				un.ncols = (int32_t)ncols;
				un.r0 = (int32_t)r0;
				un.g[0] = (int32_t)kgs.size();
				int fcur = 1;
				for (const PInstr &pi : pis) {
					const TInstr &t = op->h_tinstr[pi.v];
					const uint32_t base = op->h_rowptr[t.p_rows];
					std::vector<uint32_t> cols;
					for (int64_t r = r0; r < r1; r++)
						for (uint32_t e = op->h_rowptr[t.p_rows + r];
						     e < op->h_rowptr[t.p_rows + r + 1]; e++)
							cols.push_back(pi.ent[e - base].x);
					if (cols.empty())
						continue;
					std::sort(cols.begin(), cols.end());
					cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
					// pis are ordered by force class
					for (; fcur < t.force; fcur++)
						un.g[fcur] = (int32_t)kgs.size();
					const size_t first = kgs.size();
					for (size_t k0 = 0; k0 < cols.size(); k0 += 4) {
						MfmaKGroup kg{};
						kg.x_off = t.x_off;
						for (size_t q = 0; q < 4; q++)
							kg.rows[q] = (int32_t)cols[std::min(k0 + q, cols.size() - 1)];
						kgs.push_back(kg);
					}
					A.resize(kgs.size() * 64, 0.0);
					for (int64_t r = r0; r < r1; r++)
						for (uint32_t e = op->h_rowptr[t.p_rows + r];
						     e < op->h_rowptr[t.p_rows + r + 1]; e++) {
							const uint2 en = pi.ent[e - base];
							const size_t k = (size_t)(std::lower_bound(cols.begin(), cols.end(), en.x) -
										  cols.begin());
							const double v = pi.val[en.y & kMask];
							A[(first + k / 4) * 64 + (size_t)(r - r0) + 16 * (k % 4)] +=
								(en.y & 0x80000000u) ? -v : v;
							entries_used += 1;
						}
				}
				for (; fcur <= 3; fcur++)
					un.g[fcur] = (int32_t)kgs.size();
				if (un.g[3] > un.g[0])
					units.push_back(un);
			}
		}
		BACCHUS_HIP_CHECK(hipMalloc(&bt->d_mfma_vi[i], elig.size() * sizeof(uint32_t)));
		BACCHUS_HIP_CHECK(hipMemcpy(bt->d_mfma_vi[i], elig.data(), elig.size() * sizeof(uint32_t),
					    hipMemcpyHostToDevice));
		bt->mfma_has[i] = std::any_of(elig.begin(), elig.end(), [](uint32_t e) { return e != 0; });
		bt->h_ops[i].mfma_vi = bt->d_mfma_vi[i];
		bt->h_ops[i].mfma = bt->mfma_has[i] && bt->mfma_scales_ok[i];
	}
	if (kgs.size() > (size_t)INT32_MAX)
		return BACCHUS_FAIL("more than 2^31 matrix-core K-groups");
	// items: every unit in column chunks (128 columns for rows wider than
	// 64 columns, 64 otherwise) and runs of at most kMfmaSplit K-groups,
	// the longest runs first
	std::vector<std::pair<int32_t, MfmaItem>> items;
	double mfmas = 0;
	for (size_t k = 0; k < units.size(); k++) {
		const MfmaUnit &u = units[k];
		const int64_t width = (int64_t)u.ncols * bt->Bp[u.op];
		const int nq = width > 64 ? 2 : 1;
		const int64_t chunks = (width + 64 * nq - 1) / (64 * nq);
		mfmas += double(u.g[3] - u.g[0]) * double(chunks * 4 * nq);
		for (int64_t c = 0; c < chunks; c++)
			for (int32_t g = u.g[0]; g < u.g[3]; g += kMfmaSplit) {
				const int32_t g1 = std::min(u.g[3], g + kMfmaSplit);
				items.push_back({(g1 - g) * nq, MfmaItem{(int32_t)k, (int32_t)c, g, g1}});
			}
	}
	std::stable_sort(items.begin(), items.end(),
			 [](const std::pair<int32_t, MfmaItem> &a, const std::pair<int32_t, MfmaItem> &b) {
				 return a.first > b.first;
			 });
	std::vector<MfmaItem> it(items.size());
	for (size_t k = 0; k < items.size(); k++)
		it[k] = items[k].second;
	if (getenv("BACCHUS_HIP_VERBOSE") && !units.empty()) {
		int64_t kmax = 0, narrow = 0;
		for (const MfmaUnit &u : units) {
			kmax = std::max<int64_t>(kmax, u.g[3] - u.g[0]);
			narrow += (int64_t)u.ncols * bt->Bp[u.op] <= 64;
		}
		fprintf(stderr, "[bacchus] matrix-core P-form path: %zu units (%lld with rows of <= 64 "
			"columns), %zu K-groups (at most %lld per unit), %zu items, fill %.3f, %.0f MFMA "
			"per matvec\n", units.size(), (long long)narrow, kgs.size(), (long long)kmax,
			it.size(), kgs.empty() ? 0.0 : entries_used / (64.0 * kgs.size()), mfmas);
	}
	std::vector<float> Af(A.begin(), A.end());
	bt->mfma_items = (int64_t)it.size();
	bt->mfma_kgroups = (int64_t)kgs.size();
	bt->mfma_entries = entries_used;
	bt->mfma_columns = mfmas;
	if (!it.empty()) {
		BACCHUS_HIP_CHECK(hipMalloc(&bt->d_mfma_units, units.size() * sizeof(MfmaUnit)));
		BACCHUS_HIP_CHECK(hipMalloc(&bt->d_mfma_kg, kgs.size() * sizeof(MfmaKGroup)));
		BACCHUS_HIP_CHECK(hipMalloc(&bt->d_mfma_a, Af.size() * sizeof(float)));
		BACCHUS_HIP_CHECK(hipMalloc(&bt->d_mfma_items, it.size() * sizeof(MfmaItem)));
		BACCHUS_HIP_CHECK(hipMemcpy(bt->d_mfma_units, units.data(), units.size() * sizeof(MfmaUnit),
					    hipMemcpyHostToDevice));
		BACCHUS_HIP_CHECK(hipMemcpy(bt->d_mfma_kg, kgs.data(), kgs.size() * sizeof(MfmaKGroup),
					    hipMemcpyHostToDevice));
		BACCHUS_HIP_CHECK(hipMemcpy(bt->d_mfma_a, Af.data(), Af.size() * sizeof(float),
					    hipMemcpyHostToDevice));
		BACCHUS_HIP_CHECK(hipMemcpy(bt->d_mfma_items, it.data(), it.size() * sizeof(MfmaItem),
					    hipMemcpyHostToDevice));
	}
	return 0;
}

// Builds the tiles of every operator for the batch's tuning and the merged,
// heaviest-first work list.
int build_plan(bacchus_hip_batch_t *bt)
{
	bt->release_plan();
	const int nops = (int)bt->ops.size();
	const bacchus_hip_batch_tuning_t &tu = bt->tuning;
	const int tiles_per_cu = tu.tiles_per_cu > 0 ? tu.tiles_per_cu : 64;
	const int64_t words = tu.tile_words > 0 ? std::min<int64_t>(kBatchTileWords, tu.tile_words) : 4096;
	const int max_vec = tu.max_tile_vectors > 0 ? pow2_floor(std::min(64, tu.max_tile_vectors)) : 64;
	const int num_cu = bt->ops[0]->num_cu;
	double work = 0;
	for (int i = 0; i < nops; i++)
		work += bt->ops[i]->mults * bt->Bp[i];
	const double target = std::max(2.0e4, work / (double(num_cu) * tiles_per_cu));
	struct Cand {
		double cost;
		BatchItem item;
		int space;
		int64_t y_off;
		int po_begin;
	};
	// input groups (balanced over all operators), one work queue per XCD;
	// with more groups than XCDs (BACCHUS_HIP_BATCH_GROUPS), group g goes to
	// queue g mod 8 (measured: 16 groups C3 block 6.53 ms, C4 2.13 ms; 8
	// groups 6.60 and 0.76 ms)
	int groups = kXcds;
	if (const char *ge = getenv("BACCHUS_HIP_BATCH_GROUPS"))
		groups = std::max(1, std::min(256, atoi(ge)));
	const int nqueues = std::min(groups, kXcds);
	std::vector<double> group_load(groups, 0.0);
	std::vector<std::vector<Cand>> queues(nqueues);
	int64_t max_words = 1;
	for (int i = 0; i < nops; i++) {
		const bacchus_hip_operator_t *op = bt->ops[i];
		const int64_t nb = (int64_t)op->block_nn.size();
		const int Bp = bt->Bp[i];
		const int vec = std::max(4, std::min(max_vec, pow2_floor(Bp) * (pow2_floor(Bp) < Bp ? 2 : 1)));
		PlanParams pp;
		pp.target = target;
		pp.in_groups = groups;
		pp.group_load = &group_load;
		// row widths: neutron states (normal space), proton states
		// (transposed space)
		int64_t widest = 1;
		for (int64_t b = 0; b < nb; b++)
			widest = std::max({widest, op->block_nn[b], op->block_np[b]});
		// the LDS tile holds rows x nn x (bt + 1) doubles (the pitch of
		// batch_tile): the plan's rows x nn x bt at most 4/5 of the words
		pp.tile_elems = std::max<int64_t>(words * 4 / 5, 4 * widest);
		pp.bt.resize(2 * nb);
		for (int64_t slot = 0; slot < 2 * nb; slot++) {
			// widest vector chunk that still leaves kMinTileRows rows in
			// the LDS tile: row-side NP entries are few per row, and a pass
			// over the lane-side entries should serve several of them
			const int64_t width = slot < nb ? op->block_nn[slot] : op->block_np[slot - nb];
			const int64_t rows = slot < nb ? op->block_np[slot] : op->block_nn[slot - nb];
			const int64_t want = std::min<int64_t>(kMinTileRows, rows);
			int v = vec;
			while (v > 4 && width * v * want > pp.tile_elems)
				v >>= 1;
			pp.bt[slot] = v;
		}
		Plan plan;
		if (int rc = make_plan(op, pp, plan))
			return rc;
		TileWork *dt = nullptr;
		Task *dk = nullptr;
		BACCHUS_HIP_CHECK(hipMalloc(&dt, std::max<size_t>(16, plan.tiles.size() * sizeof(TileWork))));
		bt->d_tiles.push_back(dt);
		BACCHUS_HIP_CHECK(hipMalloc(&dk, std::max<size_t>(16, plan.tasks.size() * sizeof(Task))));
		bt->d_tasks.push_back(dk);
		if (!plan.tiles.empty())
			BACCHUS_HIP_CHECK(hipMemcpy(dt, plan.tiles.data(), plan.tiles.size() * sizeof(TileWork),
						    hipMemcpyHostToDevice));
		if (!plan.tasks.empty())
			BACCHUS_HIP_CHECK(hipMemcpy(dk, plan.tasks.data(), plan.tasks.size() * sizeof(Task),
						    hipMemcpyHostToDevice));
		bt->h_ops[i].tiles = dt;
		bt->h_ops[i].tasks = dk;
		for (size_t t = 0; t < plan.tiles.size(); t++) {
			const TileWork &tw = plan.tiles[t];
			max_words = std::max<int64_t>(max_words, int64_t(tw.po_end - tw.po_begin) *
									 tw.nn_out * (plan.bt[t] + 1));
		}
		for (size_t t = 0; t < plan.tiles.size(); t++) {
			const int v = plan.bt[t];
			int lbt = 0;
			while ((1 << lbt) < v)
				lbt++;
			for (int b0 = 0; b0 < Bp; b0 += v)
				queues[plan.tiles[t].group % nqueues].push_back(
					Cand{plan.cost[t], BatchItem{i, (int32_t)t, b0, lbt}, plan.tiles[t].space,
					     plan.tiles[t].y_off, plan.tiles[t].po_begin});
		}
	}
	std::vector<BatchItem> items;
	std::vector<int> qbegin(1, 0);
	// within a queue: every item heaviest first; BACCHUS_HIP_BATCH_ORDER=1
	// puts the items of one output block one after another instead, blocks
	// heaviest first (measured slower through the dynamic queues: C3 block
	// 6.84 against 6.61 ms, C4 0.773 against 0.751 ms)
	const char *ord = getenv("BACCHUS_HIP_BATCH_ORDER");
	const bool by_block = ord && atoi(ord) == 1;
	for (auto &q : queues) {
		if (by_block) {
			std::map<std::tuple<int, int, int64_t>, double> bc;
			for (const Cand &c : q)
				bc[std::make_tuple(c.item.op, c.space, c.y_off)] += c.cost;
			std::stable_sort(q.begin(), q.end(), [&](const Cand &a, const Cand &b) {
				const double ca = bc[std::make_tuple(a.item.op, a.space, a.y_off)];
				const double cb = bc[std::make_tuple(b.item.op, b.space, b.y_off)];
				if (ca != cb)
					return ca > cb;
				return std::make_tuple(a.item.op, a.space, a.y_off, a.item.b0, a.po_begin) <
				       std::make_tuple(b.item.op, b.space, b.y_off, b.item.b0, b.po_begin);
			});
		} else
			std::stable_sort(q.begin(), q.end(),
					 [](const Cand &a, const Cand &b) { return a.cost > b.cost; });
		for (const Cand &c : q)
			items.push_back(c.item);
		if (items.size() > 0x7FFFFFFFull)
			return BACCHUS_FAIL("more than 2^31 batch work items");
		qbegin.push_back((int)items.size());
	}
	bt->num_items = (int64_t)items.size();
	bt->nq = nqueues;
	BACCHUS_HIP_CHECK(hipMalloc(&bt->d_items, std::max<size_t>(16, items.size() * sizeof(BatchItem))));
	if (!items.empty())
		BACCHUS_HIP_CHECK(hipMemcpy(bt->d_items, items.data(), items.size() * sizeof(BatchItem),
					    hipMemcpyHostToDevice));
	BACCHUS_HIP_CHECK(hipMalloc(&bt->d_qbegin, qbegin.size() * sizeof(int)));
	BACCHUS_HIP_CHECK(hipMemcpy(bt->d_qbegin, qbegin.data(), qbegin.size() * sizeof(int),
				    hipMemcpyHostToDevice));
	if (max_words > kBatchTileWords)
		return BACCHUS_FAIL("a batched tile needs %lld doubles of LDS (%lld at most)",
				    (long long)max_words, (long long)kBatchTileWords);
	bt->lds_bytes = (size_t)max_words * sizeof(double);
	const int wg_per_cu = tu.workgroups_per_cu > 0 ? tu.workgroups_per_cu : 8;
	// a multiple of kXcds, so every XCD has the same number of workgroups
	const int64_t want = std::min<int64_t>(bt->num_items, (int64_t)num_cu * wg_per_cu);
	bt->grid = (int)std::max<int64_t>(kXcds, (want + kXcds - 1) / kXcds * kXcds);
	return upload_ops(bt);
}

// the launch in three parts: operator i's inputs (output cleared, input
// transposed), the batched kernels of all operators, operator i's outputs
// (transposed-space sums added); launch_batch runs them on one stream
void launch_batch_in(bacchus_hip_batch_t *bt, size_t i, hipStream_t s)
{
	const bacchus_hip_operator_t *op = bt->ops[i];
	const int64_t n = op->dimension * bt->Bp[i];
	hipLaunchKernelGGL(fill_float, dim3(blocks_of(n)), dim3(256), 0, s, bt->h_ops[i].y, n, 0.f);
	if (op->has_transposed) {
		hipLaunchKernelGGL(fill_float, dim3(blocks_of(n)), dim3(256), 0, s, bt->d_yT[i], n, 0.f);
		hipLaunchKernelGGL(transpose_blocks<float>,
				   dim3((unsigned)op->block_np.size(), kTransposeSplits), dim3(256), 0, s,
				   bt->h_ops[i].x, bt->d_xT[i], op->d_blk_off, op->d_blk_np,
				   op->d_blk_nn, bt->Bp[i]);
	}
}

void launch_batch_core(bacchus_hip_batch_t *bt, hipStream_t s)
{
	if (bt->num_items) {
		hipLaunchKernelGGL(reset_queue, dim3(1), dim3(64), 0, s, bt->d_next, bt->nq);
		hipLaunchKernelGGL(bt->tuning.waves_per_simd == 4 ? batch_kernel_w4 : batch_kernel,
				   dim3(bt->grid), dim3(kTileThreads), bt->lds_bytes, s,
				   bt->d_ops, bt->d_items, bt->d_qbegin, bt->nq, bt->d_next);
	}
	if (bt->mfma_items)
		hipLaunchKernelGGL(batch_mfma_kernel, dim3((unsigned)((bt->mfma_items + 3) / 4)), dim3(256),
				   0, s, bt->d_ops, bt->d_mfma_units, bt->d_mfma_kg, bt->d_mfma_a,
				   bt->d_mfma_items, (int)bt->mfma_items);
}

void launch_batch_out(bacchus_hip_batch_t *bt, size_t i, hipStream_t s)
{
	const bacchus_hip_operator_t *op = bt->ops[i];
	if (op->has_transposed)
		hipLaunchKernelGGL(add_transposed_blocks<float>,
				   dim3((unsigned)op->block_np.size(), kTransposeSplits), dim3(256), 0, s,
				   bt->h_ops[i].y, bt->d_yT[i], op->d_blk_off, op->d_blk_np,
				   op->d_blk_nn, bt->Bp[i]);
}

int launch_batch(bacchus_hip_batch_t *bt, hipStream_t s)
{
	for (size_t i = 0; i < bt->ops.size(); i++)
		launch_batch_in(bt, i, s);
	launch_batch_core(bt, s);
	for (size_t i = 0; i < bt->ops.size(); i++)
		launch_batch_out(bt, i, s);
	BACCHUS_HIP_CHECK(hipGetLastError());
	return 0;
}

int ensure_buffers(bacchus_hip_batch_t *bt)
{
	for (size_t i = 0; i < bt->ops.size(); i++) {
		const size_t n = (size_t)bt->ops[i]->dimension * bt->Bp[i];
		if (!bt->d_x[i]) {
			BACCHUS_HIP_CHECK(hipMalloc(&bt->d_x[i], n * sizeof(float)));
			BACCHUS_HIP_CHECK(hipMemset(bt->d_x[i], 0, n * sizeof(float)));
		}
		if (!bt->d_y[i])
			BACCHUS_HIP_CHECK(hipMalloc(&bt->d_y[i], n * sizeof(float)));
	}
	return 0;
}

// points every operator's descriptor at the internal buffers
int use_internal(bacchus_hip_batch_t *bt)
{
	if (int rc = ensure_buffers(bt))
		return rc;
	for (size_t i = 0; i < bt->ops.size(); i++) {
		bt->h_ops[i].x = bt->d_x[i];
		bt->h_ops[i].y = bt->d_y[i];
	}
	return upload_ops(bt);
}

}  // namespace

int batch_workspace(bacchus_hip_batch_t *bt, size_t bytes, void **ptr)
{
	if (bytes > bt->workspace_bytes) {
		(void)hipFree(bt->workspace);
		bt->workspace = nullptr;
		bt->workspace_bytes = 0;
		BACCHUS_HIP_CHECK(hipMalloc(&bt->workspace, bytes));
		bt->workspace_bytes = bytes;
	}
	*ptr = bt->workspace;
	return 0;
}

int batch_shape(const bacchus_hip_batch_t *bt, int32_t *num_ops, int32_t *num_vectors,
		int64_t *dimension)
{
	if (!bt || bt->ops.empty())
		return BACCHUS_FAIL("empty batch");
	*num_ops = (int32_t)bt->ops.size();
	*num_vectors = bt->B[0];
	*dimension = bt->ops[0]->dimension;
	return 0;
}

}  // namespace bacchus

extern "C" {

int bacchus_hip_batch_create(bacchus_hip_operator_t *const *ops, const int32_t *num_vectors,
			     int32_t num_ops, const bacchus_hip_batch_tuning_t *tuning,
			     bacchus_hip_batch_t **out)
{
	if (!ops || !num_vectors || !out || num_ops <= 0)
		return BACCHUS_FAIL("NULL argument or no operators");
	*out = nullptr;
	for (int i = 0; i < num_ops; i++) {
		if (!ops[i])
			return BACCHUS_FAIL("operator %d is NULL", i);
		if (ops[i]->precision == BACCHUS_HIP_FP64)
			return BACCHUS_FAIL("operator %d stores FP64 elements; the batched "
					    "form takes FP32 or BF16 operators", i);
		if (num_vectors[i] < 1 || num_vectors[i] > kMaxVectors)
			return BACCHUS_FAIL("operator %d: %d vectors (1..%d allowed)", i,
					    num_vectors[i], kMaxVectors);
		if (ops[i]->dimension * (int64_t)num_vectors[i] > (int64_t)INT32_MAX)
			return BACCHUS_FAIL("operator %d: D x B = %lld exceeds 2^31", i,
					    (long long)(ops[i]->dimension * num_vectors[i]));
		for (size_t b = 0; b < ops[i]->block_nn.size(); b++)
			if (std::max(ops[i]->block_nn[b], ops[i]->block_np[b]) > kBatchTileWords / 5)
				return BACCHUS_FAIL("operator %d has a block of %lld x %lld states; the "
						    "batched tile holds rows of at most %lld", i,
						    (long long)ops[i]->block_np[b], (long long)ops[i]->block_nn[b],
						    (long long)(kBatchTileWords / 5));
	}
	auto *bt = new bacchus_hip_batch;
	if (tuning)
		bt->tuning = *tuning;
	bt->ops.assign(ops, ops + num_ops);
	bt->B.assign(num_vectors, num_vectors + num_ops);
	for (int i = 0; i < num_ops; i++)
		bt->Bp.push_back((num_vectors[i] + 3) / 4 * 4);
	bt->d_x.assign(num_ops, nullptr);
	bt->d_y.assign(num_ops, nullptr);
	bt->d_xT.assign(num_ops, nullptr);
	bt->d_yT.assign(num_ops, nullptr);
	bt->h_ops.resize(num_ops);
	auto fail = [&](int rc) {
		delete bt;
		return rc;
	};
	for (int i = 0; i < num_ops; i++) {
		const bacchus_hip_operator_t *op = ops[i];
		const int B = num_vectors[i], Bp = bt->Bp[i];
		// [4][Bp]: scale of force class f for vector b at f * Bp + b;
		// padding vectors scaled by 0
		std::vector<float> sc((size_t)Bp * 4, 0.0f);
		for (int f = 1; f < 4; f++)
			for (int b = 0; b < B; b++)
				sc[(size_t)f * Bp + b] = 1.0f;
		float *ds = nullptr;
		if (hipMalloc(&ds, sc.size() * sizeof(float)) != hipSuccess)
			return fail(BACCHUS_FAIL("hipMalloc of the scales failed"));
		bt->d_scales.push_back(ds);
		if (hipMemcpy(ds, sc.data(), sc.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
			return fail(BACCHUS_FAIL("upload of the scales failed"));
		if (op->has_transposed) {
			const size_t n = (size_t)op->dimension * Bp * sizeof(float);
			if (hipMalloc(&bt->d_xT[i], n) != hipSuccess ||
			    hipMalloc(&bt->d_yT[i], n) != hipSuccess)
				return fail(BACCHUS_FAIL("hipMalloc of the transposed vectors failed"));
		}
		BatchOp &o = bt->h_ops[i];
		o = BatchOp{};
		o.ins = op->d_tinstr;
		o.rowptr = op->d_rowptr;
		o.xT = bt->d_xT[i];
		o.yT = bt->d_yT[i];
		o.ent = op->d_entries;
		o.keys = op->d_keys;
		o.V = op->d_v;
		o.scales = ds;
		o.B = Bp;
		o.precision = op->precision;
		const double sv = op->precision == BACCHUS_HIP_FP32 ? 4 : 2;
		bt->bytes += sv * double(op->element_count) + 12.0 * double(op->list_entry_count) +
			     8.0 * double(op->dimension) * B;
		bt->flops += 2.0 * op->mults * B;
	}
	if (hipMalloc(&bt->d_ops, num_ops * sizeof(BatchOp)) != hipSuccess ||
	    hipMalloc(&bt->d_next, kXcds * sizeof(int)) != hipSuccess)
		return fail(BACCHUS_FAIL("hipMalloc of the descriptors failed"));
	if (int rc = build_mfma(bt))
		return fail(rc);
	if (int rc = build_plan(bt))
		return fail(rc);
	if (int rc = use_internal(bt))
		return fail(rc);
	*out = bt;
	return 0;
}

void bacchus_hip_batch_free(bacchus_hip_batch_t *batch) { delete batch; }

int bacchus_hip_batch_set_scales(bacchus_hip_batch_t *bt, int32_t i, const double *scales)
{
	if (!bt || !scales)
		return BACCHUS_FAIL("NULL argument");
	if (i < 0 || i >= (int32_t)bt->ops.size())
		return BACCHUS_FAIL("operator index %d out of range", i);
	const int B = bt->B[i], Bp = bt->Bp[i];
	std::vector<float> sc((size_t)Bp * 4, 0.0f);
	for (int b = 0; b < B; b++)
		for (int f = 0; f < 3; f++)
			sc[(size_t)(f + 1) * Bp + b] = (float)scales[(size_t)b * 3 + f];
	BACCHUS_HIP_CHECK(hipMemcpy(bt->d_scales[i], sc.data(), sc.size() * sizeof(float),
				    hipMemcpyHostToDevice));
	bt->mfma_scales_ok[i] = mfma_scales_allowed(sc);
	const int32_t on = bt->mfma_has[i] && bt->mfma_scales_ok[i];
	if (on != bt->h_ops[i].mfma) {
		bt->h_ops[i].mfma = on;
		return upload_ops(bt);
	}
	return 0;
}

int bacchus_hip_batch_matvec(bacchus_hip_batch_t *bt, const float *const *x, float *const *y,
			     void *stream)
{
	if (!bt || !x || !y)
		return BACCHUS_FAIL("NULL argument");
	hipStream_t s = (hipStream_t)stream;
	bool changed = false;
	std::vector<char> direct(bt->ops.size());
	for (size_t i = 0; i < bt->ops.size(); i++) {
		if (!x[i] || !y[i])
			return BACCHUS_FAIL("NULL vector for operator %zu", i);
		// caller arrays are used in place when their rows are whole
		// 16-byte groups; otherwise they go through the padded buffers
		direct[i] = bt->B[i] == bt->Bp[i] && ((uintptr_t)x[i] & 15) == 0 &&
			    ((uintptr_t)y[i] & 15) == 0;
		const float *xi = direct[i] ? x[i] : bt->d_x[i];
		float *yi = direct[i] ? y[i] : bt->d_y[i];
		if (bt->h_ops[i].x != xi || bt->h_ops[i].y != yi) {
			bt->h_ops[i].x = xi;
			bt->h_ops[i].y = yi;
			changed = true;
		}
	}
	if (changed) {
		BACCHUS_HIP_CHECK(hipStreamSynchronize(s));
		if (int rc = upload_ops(bt))
			return rc;
	}
	for (size_t i = 0; i < bt->ops.size(); i++)
		if (!direct[i]) {
			const int64_t D = bt->ops[i]->dimension;
			hipLaunchKernelGGL(pad_cols, dim3(blocks_of(D * bt->Bp[i])), dim3(256), 0, s, x[i], D,
					   bt->B[i], bt->Bp[i], bt->d_x[i]);
		}
	if (int rc = launch_batch(bt, s))
		return rc;
	for (size_t i = 0; i < bt->ops.size(); i++)
		if (!direct[i]) {
			const int64_t D = bt->ops[i]->dimension;
			hipLaunchKernelGGL(unpad_cols, dim3(blocks_of(D * bt->B[i])), dim3(256), 0, s,
					   bt->d_y[i], D, bt->B[i], bt->Bp[i], y[i]);
		}
	BACCHUS_HIP_CHECK(hipGetLastError());
	return 0;
}

int bacchus_hip_batch_stats(const bacchus_hip_batch_t *bt, double *bytes, double *flops,
			    int64_t *num_items)
{
	if (!bt)
		return BACCHUS_FAIL("NULL argument");
	if (bytes)
		*bytes = bt->bytes;
	if (flops)
		*flops = bt->flops;
	if (num_items)
		*num_items = bt->num_items;
	return 0;
}

int bacchus_hip_batch_mfma_stats(const bacchus_hip_batch_t *bt, int64_t *kgroups,
				 double *mfma_per_matvec, double *fill, int32_t *operators_on)
{
	if (!bt)
		return BACCHUS_FAIL("NULL argument");
	if (kgroups)
		*kgroups = bt->mfma_kgroups;
	if (mfma_per_matvec)
		*mfma_per_matvec = bt->mfma_columns;
	if (fill)
		*fill = bt->mfma_kgroups ? bt->mfma_entries / (64.0 * double(bt->mfma_kgroups)) : 0.0;
	if (operators_on) {
		int32_t on = 0;
		for (const BatchOp &o : bt->h_ops)
			on += o.mfma != 0;
		*operators_on = on;
	}
	return 0;
}

int bacchus_hip_batch_autotune(bacchus_hip_batch_t *bt, int32_t reps,
			       bacchus_hip_batch_tuning_t *best, double *best_ms)
{
	if (!bt)
		return BACCHUS_FAIL("NULL argument");
	if (reps < 1)
		reps = 3;
	const std::vector<BatchOp> saved = bt->h_ops;
	if (int rc = use_internal(bt))
		return rc;
	hipEvent_t e0, e1;
	BACCHUS_HIP_CHECK(hipEventCreate(&e0));
	BACCHUS_HIP_CHECK(hipEventCreate(&e1));
	// candidates (C3 8-vector block, measured over a wider grid: 8 tiles per
	// CU about 1 % faster than 16, 64 slower; 6144 words slightly ahead)
	const int tpc[] = {8, 16};
	const int vec[] = {4, 16, 64};
	const int wpc[] = {4, 8};
	const int wds[] = {2048, 4096, 6144};
	bacchus_hip_batch_tuning_t best_t = bt->tuning;
	double best_v = 1e300;
	int rc = 0;
	for (int a : tpc)
		for (int v : vec)
			for (int g : wpc)
				for (int wd : wds) {
					bacchus_hip_batch_tuning_t t{a, v, g, wd, bt->tuning.waves_per_simd};
					bt->tuning = t;
					if ((rc = build_plan(bt)))
						goto done;
					if ((rc = launch_batch(bt, nullptr)))
						goto done;
					(void)hipEventRecord(e0, nullptr);
					for (int r = 0; r < reps; r++)
						if ((rc = launch_batch(bt, nullptr)))
							goto done;
					(void)hipEventRecord(e1, nullptr);
					if (hipEventSynchronize(e1) != hipSuccess) {
						rc = BACCHUS_FAIL("autotune run failed");
						goto done;
					}
					float ms = 0;
					(void)hipEventElapsedTime(&ms, e0, e1);
					if (ms / reps < best_v) {
						best_v = ms / reps;
						best_t = t;
					}
				}
	// the best tiling again with the kernel held to four waves per SIMD
	{
		bacchus_hip_batch_tuning_t t = best_t;
		t.waves_per_simd = best_t.waves_per_simd == 4 ? 0 : 4;
		bt->tuning = t;
		if ((rc = build_plan(bt)))
			goto done;
		if ((rc = launch_batch(bt, nullptr)))
			goto done;
		(void)hipEventRecord(e0, nullptr);
		for (int r = 0; r < reps; r++)
			if ((rc = launch_batch(bt, nullptr)))
				goto done;
		(void)hipEventRecord(e1, nullptr);
		if (hipEventSynchronize(e1) != hipSuccess) {
			rc = BACCHUS_FAIL("autotune run failed");
			goto done;
		}
		float ms = 0;
		(void)hipEventElapsedTime(&ms, e0, e1);
		if (ms / reps < best_v) {
			best_v = ms / reps;
			best_t = t;
		}
	}
	bt->tuning = best_t;
	rc = build_plan(bt);
done:
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	for (size_t i = 0; i < saved.size(); i++) {
		bt->h_ops[i].x = saved[i].x;
		bt->h_ops[i].y = saved[i].y;
	}
	if (!rc)
		rc = upload_ops(bt);
	if (rc)
		return rc;
	if (best)
		*best = best_t;
	if (best_ms)
		*best_ms = best_v;
	return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------
// Batched Lanczos (src/Bacchus/lanczos/lanczos.c:74-253 per instance, with
// full reorthogonalization instead of lanczos.c:155-169).  Vectors of an
// operator are stored [D][B] in FP64; the Krylov basis Q[j] for j <= K is
// kept in HBM.  Every kernel reads the iteration number k from device
// memory, so one iteration is a fixed kernel sequence that is captured
// once as a hipGraph and replayed.
// ---------------------------------------------------------------------
namespace bacchus {
namespace {

constexpr int kDotChunks = 128;

struct LState {
	int64_t D = 0;
	int B = 0;
	double *Q = nullptr;      // [K + 1][D][B]
	double *w = nullptr;      // [D][B]
	double *alpha = nullptr;  // [K][B]
	double *beta = nullptr;   // [K][B]
	double *h = nullptr;      // [K + 1][B]
	double *partial = nullptr;  // [K + 1][chunks][B]
	double *ss = nullptr;     // [B]
	int64_t *k = nullptr;     // iteration
};

// xf[r][b] = Q_k[r][b] into the padded [D][Bp] matvec input
__global__ void to_float_cols(const double *__restrict__ Q, const int64_t *kp, int64_t D, int B,
			      int Bp, float *__restrict__ xf)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= D * Bp)
		return;
	const int64_t r = i / Bp;
	const int b = (int)(i - r * Bp);
	xf[i] = b < B ? (float)Q[*kp * D * B + r * B + b] : 0.f;
}

// w[r][b] = yf[r][b] from the padded matvec output
__global__ void to_double_cols(const float *__restrict__ yf, int64_t D, int B, int Bp,
			       double *__restrict__ w)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= D * B)
		return;
	const int64_t r = i / B;
	w[i] = (double)yf[r * Bp + (i - r * B)];
}

// partial[j][c][b] = sum over rows of chunk c of U_j[r][b] * w[r][b],
// j = 0..k (grid.y = K + 1; blocks with j > k return).  If self != 0 the
// products are w*w (j = 0 only).  Threads: b = tid % B, row lane = tid / B.
__global__ __launch_bounds__(256) void col_dot_partial(const double *__restrict__ U,
						      const double *__restrict__ w, int64_t D,
						      int B, const int64_t *kp, int self,
						      double *__restrict__ partial)
{
	__shared__ double sh[256];
	const int j = blockIdx.y;
	if (self ? j > 0 : j > *kp)
		return;
	const int R = 256 / B;
	const int tid = threadIdx.x;
	const int b = tid % B, rl = tid / B;
	const int64_t per = (D + gridDim.x - 1) / gridDim.x;
	const int64_t r0 = per * blockIdx.x, r1 = min(D, r0 + per);
	double s = 0;
	if (rl < R) {
		const double *u = self ? w : U + (int64_t)j * D * B;
		for (int64_t r = r0 + rl; r < r1; r += R) {
			const int64_t i = r * B + b;
			s += u[i] * w[i];
		}
	}
	sh[tid] = s;
	__syncthreads();
	if (tid < B) {
		double t = 0;
		for (int q = 0; q < R; q++)
			t += sh[tid + q * B];
		partial[((int64_t)j * gridDim.x + blockIdx.x) * B + tid] = t;
	}
}

// out[j][b] = sum_c partial[j][c][b] in a fixed order: block j, thread
// (g, b) of the 256 / B chunk groups adds chunks g, g + 256 / B, ..., then
// the groups' sums are added in LDS (B <= kMaxVectors = 256).  alpha (may
// be null): the block j = k also stores (add = 0) or adds (add = 1) its
// sums into alpha[k][b], the Lanczos coefficient of q_k
__global__ __launch_bounds__(256) void col_dot_reduce(const double *__restrict__ partial, int chunks,
						     int B, const int64_t *kp, int self,
						     double *__restrict__ out,
						     double *__restrict__ alpha = nullptr, int add = 0)
{
	__shared__ double sh[256];
	const int j = blockIdx.x;
	if (self ? j > 0 : j > *kp)
		return;
	const int G = 256 / B;
	const int t = threadIdx.x;
	const int g = t / B, b = t - g * B;
	double v = 0;
	if (g < G)
		for (int c = g; c < chunks; c += G)
			v += partial[((int64_t)j * chunks + c) * B + b];
	sh[t] = v;
	__syncthreads();
	if (t < B) {
		double s = 0;
		for (int q = 0; q < G; q++)
			s += sh[q * B + t];
		out[(int64_t)j * B + t] = s;
		if (alpha && !self && j == *kp) {
			double *a = alpha + (int64_t)j * B + t;
			*a = add ? *a + s : s;
		}
	}
}

// w[r][b] -= sum_{j <= k} h[j][b] Q_j[r][b]
__global__ void col_subtract(double *__restrict__ w, const double *__restrict__ Q,
			     const double *__restrict__ h, int64_t n, int B, const int64_t *kp)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const int b = (int)(i % B);
	const int64_t k = *kp;
	double v = w[i];
	for (int64_t j = 0; j <= k; j++)
		v -= h[j * B + b] * Q[j * n + i];
	w[i] = v;
}

// beta[k][b] = |w_b|, Q[k+1] = w / beta (zero after a breakdown), k += 1
__global__ void finish_iteration(double *__restrict__ Q, const double *__restrict__ w,
				 const double *__restrict__ ss, double *__restrict__ beta,
				 int64_t n, int B, const int64_t *kp)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const int b = (int)(i % B);
	const int64_t k = *kp;
	const double bb = sqrt(ss[b]);
	if (i < B)
		beta[k * B + i] = bb;
	Q[(k + 1) * n + i] = bb > 1e-300 ? w[i] / bb : 0.0;
}

__global__ void advance(int64_t *kp)
{
	if (threadIdx.x == 0 && blockIdx.x == 0)
		*kp += 1;
}

__global__ void set_counter(int64_t *kp, int64_t v)
{
	if (threadIdx.x == 0 && blockIdx.x == 0)
		*kp = v;
}

__device__ __forceinline__ uint64_t mix(uint64_t z)
{
	z += 0x9e3779b97f4a7c15ULL;
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
	return z ^ (z >> 31);
}

// uniform [-1, 1) keyed by (seed, operator, flat index)
__global__ void random_cols(double *u, int64_t n, uint64_t seed)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		u[i] = 2.0 * double(mix(seed ^ mix((uint64_t)i)) >> 11) * 0x1.0p-53 - 1.0;
}

__global__ void copy_cols(double *dst, const double *src, int64_t n)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		dst[i] = src[i];
}

// u[r][b] /= sqrt(ss[b])
__global__ void normalize_cols(double *u, const double *ss, int64_t n, int B)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) {
		const double s = sqrt(ss[i % B]);
		u[i] = s > 0 ? u[i] / s : 0.0;
	}
}

// out[r][b] = sum_j z[j][b] Q_j[r][b]
__global__ void combine_cols(double *__restrict__ out, const double *__restrict__ Q,
			     const double *__restrict__ z, int64_t n, int B, int64_t k)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const int b = (int)(i % B);
	double v = 0;
	for (int64_t j = 0; j < k; j++)
		v += z[j * B + b] * Q[j * n + i];
	out[i] = v;
}

struct BatchLanczos {
	std::vector<LState> st;
	hipStream_t s = nullptr;
	int64_t K = 0;
	// after the batched matvec each operator's vector work (independent of
	// the others') runs on a stream of its own, forked from and joined back
	// to s by events, so the operators' chains of small launches overlap
	// (parallel branches when captured); all streams at the default priority
	std::vector<hipStream_t> os;
	// ev[0]: iteration start; ev[1]: matvec done; ev[2 + i], ev[2 + n + i]:
	// operator i's inputs ready, its iteration done (n operators)
	std::vector<hipEvent_t> ev;

	~BatchLanczos()
	{
		for (auto e : ev)
			(void)hipEventDestroy(e);
		for (auto x : os)
			(void)hipStreamDestroy(x);
		for (auto &l : st) {
			(void)hipFree(l.Q);
			(void)hipFree(l.w);
			(void)hipFree(l.alpha);
			(void)hipFree(l.beta);
			(void)hipFree(l.h);
			(void)hipFree(l.partial);
			(void)hipFree(l.ss);
			(void)hipFree(l.k);
		}
	}

	void norm_sq(LState &l, const double *u, hipStream_t s)
	{
		hipLaunchKernelGGL(col_dot_partial, dim3(kDotChunks, 1), dim3(256), 0, s, nullptr, u,
				   l.D, l.B, l.k, 1, l.partial);
		hipLaunchKernelGGL(col_dot_reduce, dim3(1), dim3(256), 0, s, l.partial, kDotChunks, l.B,
				   l.k, 1, l.ss);
	}
	void project(LState &l, int add, hipStream_t s)
	{
		const int64_t n = l.D * l.B;
		hipLaunchKernelGGL(col_dot_partial, dim3(kDotChunks, (unsigned)(K + 1)), dim3(256), 0, s,
				   l.Q, l.w, l.D, l.B, l.k, 0, l.partial);
		hipLaunchKernelGGL(col_dot_reduce, dim3((unsigned)(K + 1)), dim3(256), 0, s, l.partial,
				   kDotChunks, l.B, l.k, 0, l.h, l.alpha, add);
		hipLaunchKernelGGL(col_subtract, dim3(blocks_of(n)), dim3(256), 0, s, l.w, l.Q, l.h, n,
				   l.B, l.k);
	}
	// one iteration of every instance: w = H q_k, CGS2 against q_0..q_k,
	// beta, q_{k+1}
	int iteration(bacchus_hip_batch_t *bt)
	{
		const bool fork = !os.empty();
		const size_t n_ops = st.size();
		if (fork)
			BACCHUS_HIP_CHECK(hipEventRecord(ev[0], s));
		for (size_t i = 0; i < n_ops; i++) {
			LState &l = st[i];
			const int Bp = bt->Bp[i];
			hipStream_t x = s;
			if (fork) {
				BACCHUS_HIP_CHECK(hipStreamWaitEvent(os[i], ev[0], 0));
				x = os[i];
			}
			hipLaunchKernelGGL(to_float_cols, dim3(blocks_of(l.D * Bp)), dim3(256), 0, x, l.Q, l.k,
					   l.D, l.B, Bp, bt->d_x[i]);
			launch_batch_in(bt, i, x);
			if (fork) {
				BACCHUS_HIP_CHECK(hipEventRecord(ev[2 + i], x));
				BACCHUS_HIP_CHECK(hipStreamWaitEvent(s, ev[2 + i], 0));
			}
		}
		launch_batch_core(bt, s);
		if (fork)
			BACCHUS_HIP_CHECK(hipEventRecord(ev[1], s));
		for (size_t i = 0; i < n_ops; i++) {
			LState &l = st[i];
			const int64_t n = l.D * l.B;
			hipStream_t x = s;
			if (fork) {
				BACCHUS_HIP_CHECK(hipStreamWaitEvent(os[i], ev[1], 0));
				x = os[i];
			}
			launch_batch_out(bt, i, x);
			hipLaunchKernelGGL(to_double_cols, dim3(blocks_of(n)), dim3(256), 0, x, bt->d_y[i], l.D,
					   l.B, bt->Bp[i], l.w);
			project(l, 0, x);
			project(l, 1, x);
			norm_sq(l, l.w, x);
			hipLaunchKernelGGL(finish_iteration, dim3(blocks_of(n)), dim3(256), 0, x, l.Q, l.w,
					   l.ss, l.beta, n, l.B, l.k);
			hipLaunchKernelGGL(advance, dim3(1), dim3(64), 0, x, l.k);
			if (fork) {
				BACCHUS_HIP_CHECK(hipEventRecord(ev[2 + n_ops + i], x));
				BACCHUS_HIP_CHECK(hipStreamWaitEvent(s, ev[2 + n_ops + i], 0));
			}
		}
		BACCHUS_HIP_CHECK(hipGetLastError());
		return 0;
	}
};

// lowest `count` eigenvalues of instance b's k x k tridiagonal, Krylov
// space cut at the first breakdown
std::vector<double> instance_eigenvalues(const std::vector<double> &alpha,
					 const std::vector<double> &beta, int B, int b,
					 int64_t k, int count, int64_t *used)
{
	std::vector<double> d, e;
	double scale = 0;
	for (int64_t j = 0; j < k; j++)
		scale = std::max(scale, std::fabs(alpha[j * B + b]));
	for (int64_t j = 0; j < k; j++) {
		d.push_back(alpha[j * B + b]);
		if (j + 1 < k) {
			const double bb = beta[j * B + b];
			if (bb <= 1e-12 * scale)
				break;
			e.push_back(bb);
		}
	}
	if (used)
		*used = (int64_t)d.size();
	e.resize(d.size() > 0 ? d.size() - 1 : 0);
	e.push_back(0.0);
	return tridiagonal_lowest_host(d.data(), e.data(), (int64_t)d.size(), count);
}

}  // namespace
}  // namespace bacchus

extern "C" int bacchus_hip_batch_lanczos(bacchus_hip_batch_t *bt,
					 const bacchus_hip_batch_lanczos_settings_t *set,
					 double *eigenvalues, double *const *ritz,
					 bacchus_hip_lanczos_result_t *result)
{
	if (!bt || !set || !eigenvalues)
		return BACCHUS_FAIL("NULL argument");
	const auto t_start = std::chrono::steady_clock::now();
	const int nops = (int)bt->ops.size();
	int64_t Dmin = INT64_MAX;
	for (auto *op : bt->ops)
		Dmin = std::min(Dmin, op->dimension);
	const int64_t K = std::min<int64_t>(set->max_num_iterations, Dmin);
	const int nev = set->num_eigenvalues;
	if (K < 1 || nev < 1)
		return BACCHUS_FAIL("need at least one iteration and one eigenvalue");
	BatchLanczos L;
	L.K = K;
	if (int rc = use_internal(bt))
		return rc;
	BACCHUS_HIP_CHECK(hipStreamCreate(&L.s));
	struct StreamGuard {
		hipStream_t s;
		~StreamGuard() { (void)hipStreamDestroy(s); }
	} guard{L.s};
	L.st.resize(nops);
	for (int i = 0; i < nops; i++) {
		LState &l = L.st[i];
		l.D = bt->ops[i]->dimension;
		l.B = bt->B[i];
		const size_t n = (size_t)l.D * l.B;
		BACCHUS_HIP_CHECK(hipMalloc(&l.Q, (size_t)(K + 1) * n * sizeof(double)));
		BACCHUS_HIP_CHECK(hipMalloc(&l.w, n * sizeof(double)));
		BACCHUS_HIP_CHECK(hipMalloc(&l.alpha, (size_t)K * l.B * sizeof(double)));
		BACCHUS_HIP_CHECK(hipMalloc(&l.beta, (size_t)K * l.B * sizeof(double)));
		BACCHUS_HIP_CHECK(hipMalloc(&l.h, (size_t)(K + 1) * l.B * sizeof(double)));
		BACCHUS_HIP_CHECK(hipMalloc(&l.partial, (size_t)(K + 1) * kDotChunks * l.B * sizeof(double)));
		BACCHUS_HIP_CHECK(hipMalloc(&l.ss, (size_t)l.B * sizeof(double)));
		BACCHUS_HIP_CHECK(hipMalloc(&l.k, sizeof(int64_t)));
		hipLaunchKernelGGL(set_counter, dim3(1), dim3(64), 0, L.s, l.k, (int64_t)0);
		const double *start = set->start_vectors ? set->start_vectors[i] : nullptr;
		if (start)
			hipLaunchKernelGGL(copy_cols, dim3(blocks_of(n)), dim3(256), 0, L.s, l.Q, start, (int64_t)n);
		else
			hipLaunchKernelGGL(random_cols, dim3(blocks_of(n)), dim3(256), 0, L.s, l.Q, (int64_t)n,
					   set->seed ^ (0x5bd1e995ULL * (uint64_t)(i + 1)));
		L.norm_sq(l, l.Q, L.s);
		hipLaunchKernelGGL(normalize_cols, dim3(blocks_of(n)), dim3(256), 0, L.s, l.Q, l.ss,
				   (int64_t)n, l.B);
	}
	BACCHUS_HIP_CHECK(hipGetLastError());
	// BACCHUS_HIP_LANCZOS_STREAMS=0 keeps every operator's vector work on
	// one stream, one operator after the other
	const char *fenv = getenv("BACCHUS_HIP_LANCZOS_STREAMS");
	if (nops > 1 && !(fenv && atoi(fenv) == 0)) {
		L.ev.resize(2 + 2 * (size_t)nops, nullptr);
		for (auto &e : L.ev)
			BACCHUS_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
		L.os.resize(nops, nullptr);
		for (auto &x : L.os)
			BACCHUS_HIP_CHECK(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
	}
	hipGraphExec_t exec = nullptr;
	hipGraph_t graph = nullptr;
	if (set->use_graph) {
		BACCHUS_HIP_CHECK(hipStreamBeginCapture(L.s, hipStreamCaptureModeThreadLocal));
		const int rc = L.iteration(bt);
		hipGraph_t g = nullptr;
		const hipError_t ec = hipStreamEndCapture(L.s, &g);
		if (rc)
			return rc;
		if (ec != hipSuccess)
			return BACCHUS_FAIL("graph capture failed: %s", hipGetErrorString(ec));
		graph = g;
		BACCHUS_HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
	}
	struct GraphGuard {
		hipGraph_t g;
		hipGraphExec_t e;
		~GraphGuard()
		{
			if (e)
				(void)hipGraphExecDestroy(e);
			if (g)
				(void)hipGraphDestroy(g);
		}
	} gguard{graph, exec};
	const int check = set->check_every > 0 ? set->check_every : 5;
	const bool converge = set->eigenvalue_tolerance > 0;
	std::vector<std::vector<double>> prev(nops);
	int64_t k = 0;
	bool converged = false;
	std::vector<std::vector<double>> ha(nops), hb(nops);
	auto read_ab = [&](int64_t kk) -> int {
		for (int i = 0; i < nops; i++) {
			const LState &l = L.st[i];
			ha[i].resize((size_t)kk * l.B);
			hb[i].resize((size_t)kk * l.B);
			BACCHUS_HIP_CHECK(hipMemcpy(ha[i].data(), l.alpha, ha[i].size() * sizeof(double),
						    hipMemcpyDeviceToHost));
			BACCHUS_HIP_CHECK(hipMemcpy(hb[i].data(), l.beta, hb[i].size() * sizeof(double),
						    hipMemcpyDeviceToHost));
		}
		return 0;
	};
	while (k < K) {
		const int64_t steps = converge ? std::min<int64_t>(check, K - k) : K - k;
		for (int64_t r = 0; r < steps; r++) {
			if (exec)
				BACCHUS_HIP_CHECK(hipGraphLaunch(exec, L.s));
			else if (int rc = L.iteration(bt))
				return rc;
		}
		k += steps;
		BACCHUS_HIP_CHECK(hipStreamSynchronize(L.s));
		if (!converge || k >= K)
			continue;
		if (int rc = read_ab(k))
			return rc;
		bool all = true;
		for (int i = 0; i < nops; i++) {
			const int B = L.st[i].B;
			std::vector<double> cur;
			for (int b = 0; b < B; b++) {
				auto ev = instance_eigenvalues(ha[i], hb[i], B, b, k, nev, nullptr);
				ev.resize(nev, NAN);
				cur.insert(cur.end(), ev.begin(), ev.end());
			}
			if (prev[i].size() != cur.size())
				all = false;
			else
				for (size_t q = 0; q < cur.size(); q++)
					if (!(std::fabs(cur[q] - prev[i][q]) <=
					      set->eigenvalue_tolerance * std::fabs(cur[q])))
						all = false;
			prev[i] = cur;
		}
		if (all) {
			converged = true;
			break;
		}
	}
	if (int rc = read_ab(k))
		return rc;
	const int nritz = std::max(1, set->num_ritz_vectors);
	size_t pos = 0;
	for (int i = 0; i < nops; i++) {
		LState &l = L.st[i];
		const int B = l.B;
		const int64_t n = l.D * B;
		// z[r][j][b]: Krylov coefficients of Ritz vector r of instance b
		std::vector<double> z((size_t)nritz * k * B, 0.0);
		for (int b = 0; b < B; b++) {
			int64_t used = 0;
			auto ev = instance_eigenvalues(ha[i], hb[i], B, b, k, std::max(nev, nritz), &used);
			for (int q = 0; q < nev; q++)
				eigenvalues[pos++] = q < (int)ev.size() ? ev[q] : NAN;
			if (!ritz || !ritz[i])
				continue;
			std::vector<double> d(used), e(used > 0 ? used - 1 : 0);
			for (int64_t j = 0; j < used; j++)
				d[j] = ha[i][j * B + b];
			for (int64_t j = 0; j + 1 < used; j++)
				e[j] = hb[i][j * B + b];
			for (int r = 0; r < nritz && r < (int)ev.size(); r++) {
				const auto zz = tridiagonal_vector(d, e, used, ev[r]);
				for (int64_t j = 0; j < used; j++)
					z[((size_t)r * k + j) * B + b] = zz[j];
			}
		}
		if (ritz && ritz[i]) {
			double *dz = nullptr;
			BACCHUS_HIP_CHECK(hipMalloc(&dz, std::max<size_t>(8, z.size() * sizeof(double))));
			BACCHUS_HIP_CHECK(hipMemcpy(dz, z.data(), z.size() * sizeof(double), hipMemcpyHostToDevice));
			for (int r = 0; r < nritz; r++)
				hipLaunchKernelGGL(combine_cols, dim3(blocks_of(n)), dim3(256), 0, L.s,
						   ritz[i] + (size_t)r * n, l.Q, dz + (size_t)r * k * B, n, B, k);
			const hipError_t e = hipStreamSynchronize(L.s);
			(void)hipFree(dz);
			if (e != hipSuccess)
				return BACCHUS_FAIL("Ritz vectors: %s", hipGetErrorString(e));
		}
	}
	BACCHUS_HIP_CHECK(hipStreamSynchronize(L.s));
	if (result) {
		result->iterations = k;
		result->converged = converged ? 1 : 0;
		result->pad = 0;
		result->matvec_seconds = 0;
		result->total_seconds =
			std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
	}
	return 0;
}
