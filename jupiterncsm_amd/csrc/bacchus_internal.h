// bacchus_internal.h - shared host/device definitions of the gfx950 Bacchus
// library (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "bacchus_hip.h"

namespace bacchus {

// Records "Error in <func> @ <file>:<line>: <msg>" for
// bacchus_hip_last_error() (the message shape of
// src/Utilities/error/error.h:7-13) and returns a nonzero code.
int set_error(const char *func, const char *file, int line, const char *fmt,
	      ...) __attribute__((format(printf, 4, 5)));
#define BACCHUS_FAIL(...) \
	::bacchus::set_error(__func__, __FILE__, __LINE__, __VA_ARGS__)
#define BACCHUS_HIP_CHECK(call)                                              \
	do {                                                                 \
		hipError_t err_ = (call);                                    \
		if (err_ != hipSuccess)                                      \
			return BACCHUS_FAIL("%s: %s", #call,                 \
					    hipGetErrorString(err_));        \
	} while (0)

// One direction of one instruction, as the tile kernel applies it:
// forward (H_out,in x_in) or, for off-diagonal instructions, transposed
// (H_out,in^T x_out into y_in) using the list copies sorted by input index.
struct TInstr {
	int64_t x_off;     // offset of the input block in x
	int64_t v_off;     // first matrix element
	int64_t dn;        // neutron matrix dimension (NP row stride of V)
	uint32_t n_base;   // NP: first entry of the element-ordered neutron
			   // copy; N: row pointer of neutron row 0
	uint32_t n_len;    // NP: neutron entries
	uint32_t p_rows;   // NP, P: row pointer of proton row 0
	int32_t nn_in;     // neutron states of the input block
	int16_t type;      // 1 N, 2 P, 3 NP
	int16_t force;     // 1, 2, 3 body
	int32_t expanded;  // NP: 1 = element rows expanded into lane-list
			   // order (column = entry position, row stride n_len);
			   // 2 = the same with lanes paired by input (FP64)
};

// One output tile: rows [po_begin, po_end) of a basis block (all columns)
// and the range of virtual instructions summed into it.  In the normal
// space a block is [proton][neutron] (rows = proton states, columns =
// neutron states); in the transposed space (space = 1) it is
// [neutron][proton], read from and written to the transposed copies of x
// and y, and proton and neutron lists exchange their roles.
struct TileWork {
	int64_t y_off;     // offset of the block in y
	int32_t nn_out;    // columns of the block (neutron states; protons in
			   // the transposed space)
	int32_t po_begin;
	int32_t po_end;
	int32_t vi_begin;
	int32_t vi_end;
	int32_t shared;    // 1: other tiles cover the same rows, flush
			   // atomically
	int32_t t_begin;   // tasks [t_begin, t_end)
	int32_t t_end;
	int32_t space;     // 0 normal, 1 transposed
	int32_t group;     // input group of its instructions (0 without
			   // input groups)
};

// One unit of a tile's work: row-side entries [jb, je) (rows of the tile,
// in row order) of a P/NP virtual instruction with the range [e0, e1) of
// lane-side entries (NP) or columns (P) it covers.  (N instructions run as
// P instructions of the transposed space.)
struct Task {
	int32_t vi;
	int32_t r;
	uint32_t jb;
	uint32_t je;
	uint32_t e0;
	uint32_t e1;
	uint32_t pad0;
	uint32_t pad1;
};

constexpr int kTileThreads = 256;
constexpr int kTileMaxElems = 2048;  // doubles of LDS per tile (16 KiB)
constexpr uint64_t kTaskMults = 65536;     // multiplications per P/NP task
// lane-side NP entries held by one lane of the single-vector kernel
constexpr int kLaneEntries = 4;
// XCDs of the device: workgroup i of a launch runs on XCD i mod kXcds
constexpr int kXcds = 8;

// Tiling of one operator's output: the tiles (rows of a block x all
// neutron states x bt vectors) with their tasks, heaviest first.
struct PlanParams {
	double target = 0;          // multiplications per tile (x bt)
	std::vector<int> bt;        // vectors per tile, per basis block
	int64_t tile_elems = kTileMaxElems;  // LDS words per tile (rows*nn*bt)
	uint64_t task_mults = kTaskMults;    // multiplications x bt per task
	int xcds = 0;               // > 1: group output blocks per XCD
	bool skip_np = false;       // leave NP instructions out (np_kernel
				    // applies them)
	// > 1: input groups.  Every basis block is given one of in_groups
	// groups (balanced by the work that reads it, added to *group_load
	// when that is set, so several operators share one balance), every
	// tile holds instructions of one group only, and with xcds > 1 the
	// tiles of group g go to XCD g: an XCD's L2 then holds the input
	// vectors of its own blocks only.  Output rows reached from several
	// groups are flushed atomically.
	int in_groups = 0;
	std::vector<double> *group_load = nullptr;
};
struct Plan {
	std::vector<TileWork> tiles;
	std::vector<Task> tasks;
	std::vector<double> cost;   // multiplications x bt of every tile
	std::vector<int> bt;        // vectors of every tile
	// virtual instructions of tile t: vi_order[vi_begin, vi_end)
	std::vector<int32_t> vi_order;
	int64_t max_tile_elems = 1;
};

// --- the neutron-proton (NP) path: element-row-major units ---
//
// An NP virtual instruction (one direction of an NP instruction, in the
// space where its lane list is the longer list) is
//   y[s_out(j)][l_out(i)] += sign_j sign_i W[a_i][b_j] x[s_in(j)][l_in(i)]
// over the row-list entries j (compact index b_j) and lane-list entries i
// (compact index a_i).  Its lane list is stored sliced ("SELL"): the output
// states with entries, sorted by entry count, q lanes per state, 64 lanes
// per slice, entry t q + u of a state at [slice][t][lane]; the element rows
// are expanded into that order per row-list compact index b (W_b[pos] =
// sign_i W[a_i][b], zero at the padding), so a wave streams W_b and the
// input indices with 16- and 4-byte coalesced loads.  The row-list entries
// are grouped by b: one W_b value in a register serves every j with
// compact index b, each j adding into its own output row.
struct NpInstr {
	int64_t x_off;      // input block offset
	int64_t w_off;      // first element of W_0 (row stride sell_len)
	uint32_t sell_len;  // padded SELL length of the lane list
	uint32_t slice0;    // first slice (global index into the slice table)
	uint32_t nslices;
	uint32_t bbase;     // row-list entries of compact index b:
			    // [bptr[bbase + b], bptr[bbase + b + 1])
	int32_t nl_in;      // lane-side states of the input block (x row stride)
	int16_t force;
	int16_t stage;      // x rows staged in LDS per pass (0: x gathered
			    // from global memory, kNpChunk rows per pass)
	uint32_t sell_base; // first position of the lane list's SELL copy
};
// A slice: `width` lanes (its states x q lanes each, q = 1 << lq chosen
// per slice to waste the fewest positions), `steps` positions per lane,
// position t width + lane.
struct NpSlice {
	uint32_t off;       // first SELL position of the slice, from the
			    // lane list's sell_base (= its element column)
	uint32_t steps;     // entries per lane
	uint32_t row_base;  // row table entry of lane 0 (64 per slice)
	uint32_t shape;     // width | lq << 16
};
// one wave's work: instruction vi, compact indices [b0, b1), slices
// [k0, k1) of its lane list
struct NpUnit {
	int32_t vi;
	uint32_t b0, b1;
	uint32_t k0, k1;
	uint32_t jsub;  // 0: every entry of [b0, b1); else b1 = b0 + 1 and the
			// entries first .. first + count of b0 (first << 16 |
			// count, count <= 64)
};
// one workgroup: a whole output block (rows x cols of its space) in LDS
// and the units that write it
struct NpTile {
	int64_t y_off;
	int32_t rows;
	int32_t cols;
	int32_t space;
	int32_t u_begin;
	int32_t u_end;
	int32_t pad;
};
constexpr int kNpThreads = 512;
constexpr int kNpChunk = 4;  // row-list entries (same b) per pass
constexpr int kNpSteps = 8;  // lane steps per round of loads
constexpr uint32_t kNpUnitSlices = 64;  // slices (and compact indices + 1)
					// of a unit at most
constexpr int kNpStage = 1024;  // x values of a wave's LDS staging area

// Shape of a batch: operators, vectors of operator 0, its dimension.
int batch_shape(const bacchus_hip_batch_t *batch, int32_t *num_ops, int32_t *num_vectors,
		int64_t *dimension);
// Device workspace of at least bytes owned by the batch (kept between calls,
// freed with the batch).
int batch_workspace(bacchus_hip_batch_t *batch, size_t bytes, void **ptr);

// Host tridiagonal helpers (diag d, off-diagonal e, order k): lowest
// `count` eigenvalues by Sturm bisection, and the eigenvector of one
// eigenvalue by inverse iteration (component 0 >= 0).
std::vector<double> tridiagonal_lowest_host(const double *d, const double *e,
					    int64_t k, int count);
std::vector<double> tridiagonal_vector(const std::vector<double> &d,
				       const std::vector<double> &e, int64_t k,
				       double lambda);

}  // namespace bacchus

struct bacchus_hip_operator {
	int precision = BACCHUS_HIP_FP64;
	int64_t dimension = 0;
	std::vector<int64_t> block_np, block_nn, block_off;
	std::vector<int64_t> block_ids;  // vec_<id> names (comb.txt ids)
	double scales[3] = {1.0, 1.0, 1.0};
	// device tables
	void *d_v = nullptr;            // elements in storage precision
	uint32_t *d_rowptr = nullptr;   // CSR row pointers (absolute)
	uint2 *d_entries = nullptr;     // {other index, matrix index|sign}
	uint32_t *d_keys = nullptr;     // row of every entry
	bacchus::TInstr *d_tinstr = nullptr;
	bacchus::TileWork *d_tiles = nullptr;
	bacchus::Task *d_tasks = nullptr;
	int64_t num_tasks = 0;
	float *d_xf = nullptr;          // FP32 copy of x (FP32/BF16 modes)
	// transposed space: per-block [neutron][proton] copies of x and y
	bool has_transposed = false;
	void *d_xT = nullptr;           // storage-precision vector type
	double *d_yT = nullptr;
	int64_t *d_blk_off = nullptr;   // block offsets, proton and neutron
	int32_t *d_blk_np = nullptr;    // counts (device copies)
	int32_t *d_blk_nn = nullptr;
	int64_t num_tiles = 0;
	int64_t num_tinstr = 0;
	size_t tile_lds_bytes = 0;
	// host copies the planner needs: CSR row pointers and the virtual
	// instructions grouped by output block ([blk_vi[b], blk_vi[b+1]))
	std::vector<uint32_t> h_rowptr;
	std::vector<bacchus::TInstr> h_tinstr;
	std::vector<int32_t> blk_vi;
	std::vector<int32_t> blk_np;    // first NP virtual instruction of each
					// slot (they come last in a slot)
	double mults = 0;
	int num_cu = 256;
	// NP path (np.hip)
	bool np_path = false;           // NP instructions run in np_kernel
	void *d_np_w = nullptr;         // expanded element rows (storage type)
	uint32_t *d_np_in = nullptr;    // SELL input states (all lane lists)
	int32_t *d_np_row = nullptr;    // SELL output state of every lane
	bacchus::NpSlice *d_np_slices = nullptr;
	uint2 *d_np_j = nullptr;        // row-list entries {s_in, s_out|sign}
	uint32_t *d_np_bptr = nullptr;  // per (row list, direction): cS + 1
	bacchus::NpInstr *d_np_instr = nullptr;
	bacchus::NpUnit *d_np_units = nullptr;
	bacchus::NpTile *d_np_tiles = nullptr;
	int64_t np_num_tiles = 0;
	size_t np_lds_bytes = 0;
	int32_t np_block_words = 0;     // LDS words of the largest block
	int64_t np_elements = 0;        // stored expanded elements
	// statistics
	size_t device_bytes = 0;        // device memory the operator holds
	double algorithmic_bytes = 0;
	double flops = 0;
	int64_t element_count = 0;
	int64_t compact_elements = 0;  // elements of the dense sub-blocks
	int64_t list_entry_count = 0;
};

namespace bacchus {
int make_plan(const bacchus_hip_operator *op, const PlanParams &pp, Plan &out);

// What the NP tables are built from (bacchus_hip_operator_create's host
// tables): the problem, every list's compact element indices, which
// instructions run in the transposed space, and the CSR copies of every
// used list (entries {other index, compact | sign}, rows by output at
// rows_by_out[l], by input at rows_by_in[l], both absolute in rowptr).
struct NpContext {
	const bacchus_hip_problem_t *p;
	const std::vector<std::vector<uint32_t>> *mvals;
	const std::vector<char> *swapped;
	const std::vector<uint32_t> *rowptr;
	const std::vector<uint2> *entries;
	const std::vector<int64_t> *rows_by_out, *rows_by_in;
	const std::vector<int64_t> *n_in, *n_out;  // sub-basis sizes per list
};
// Builds and uploads the NP tables; sets op->np_path.
int np_build(bacchus_hip_operator *op, const NpContext &c);
// Applies every NP instruction: y (and yT) += H_NP x, atomically.
int np_launch(const bacchus_hip_operator *op, const void *x, const void *xT, double *y,
	      double s1, double s2, double s3, hipStream_t s);
// Frees the NP tables.
void np_free(bacchus_hip_operator *op);

// Loads at a uniform base plus a 32-bit per-lane byte offset, as raw
// buffer loads: the base and the range (bytes) form a buffer resource in
// SGPRs, a lane needs one VGPR of offset per load (not a 64-bit address),
// and an offset beyond the range reads 0.
__device__ __forceinline__ const char *uniform_ptr(const char *p)
{
	const uint64_t v = (uint64_t)p;
	const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
	const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
	return (const char *)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc_of(const void *base, uint32_t bytes)
{
	// third dword: the raw-buffer format bits of gfx9 targets
	return __builtin_amdgcn_make_buffer_rsrc(
		(void *)uniform_ptr(static_cast<const char *>(base)), (short)0,
		(int)__builtin_amdgcn_readfirstlane(bytes), 0x00020000);
}
__device__ __forceinline__ float buf_f32(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
	return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
__device__ __forceinline__ float buf_bf16(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
	return __uint_as_float(uint32_t(__builtin_amdgcn_raw_buffer_load_b16(r, (int)off, 0, 0)) << 16);
}
__device__ __forceinline__ double buf_f64(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
	typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
	const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0);
	return __longlong_as_double((long long)(((uint64_t)v[1] << 32) | v[0]));
}
__device__ __forceinline__ void buf_f64x2(__amdgpu_buffer_rsrc_t r, uint32_t off, double &a,
					  double &b)
{
	typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
	const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
	a = __longlong_as_double((long long)(((uint64_t)v[1] << 32) | v[0]));
	b = __longlong_as_double((long long)(((uint64_t)v[3] << 32) | v[2]));
}
__device__ __forceinline__ float4 buf_f32x4(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
	typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
	const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
	return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]),
			   __uint_as_float(v[3]));
}

constexpr int kTransposeSplits = 8;

// xT[b][n][p] = x[b][p][n] for every block (B interleaved vectors; one
// workgroup row per block, grid.y splits a block)
template <typename T>
__global__ void transpose_blocks(const T *__restrict__ x, T *__restrict__ xT,
				 const int64_t *__restrict__ off, const int32_t *__restrict__ np,
				 const int32_t *__restrict__ nn, int B)
{
	const int blk = blockIdx.x;
	const int64_t o = off[blk];
	const int P = np[blk], N = nn[blk];
	const int64_t total = (int64_t)P * N * B;
	for (int64_t t = (int64_t)blockIdx.y * blockDim.x + threadIdx.x; t < total;
	     t += (int64_t)blockDim.x * gridDim.y) {
		const int64_t e = t / B;
		const int b = (int)(t - e * B);
		const int p = (int)(e / N), n = (int)(e - (int64_t)p * N);
		xT[(o + (int64_t)n * P + p) * B + b] = x[(o + e) * B + b];
	}
}

// y[b][p][n] += yT[b][n][p] for every block
template <typename T>
__global__ void add_transposed_blocks(T *__restrict__ y, const T *__restrict__ yT,
				      const int64_t *__restrict__ off, const int32_t *__restrict__ np,
				      const int32_t *__restrict__ nn, int B)
{
	const int blk = blockIdx.x;
	const int64_t o = off[blk];
	const int P = np[blk], N = nn[blk];
	const int64_t total = (int64_t)P * N * B;
	for (int64_t t = (int64_t)blockIdx.y * blockDim.x + threadIdx.x; t < total;
	     t += (int64_t)blockDim.x * gridDim.y) {
		const int64_t e = t / B;
		const int b = (int)(t - e * B);
		const int p = (int)(e / N), n = (int)(e - (int64_t)p * N);
		y[(o + e) * B + b] += yT[(o + (int64_t)n * P + p) * B + b];
	}
}

}
