// bacchus_internal.h - shared host/device definitions of the gfx950 Bacchus
// library (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

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
};
struct Plan {
	std::vector<TileWork> tiles;
	std::vector<Task> tasks;
	std::vector<double> cost;   // multiplications x bt of every tile
	std::vector<int> bt;        // vectors of every tile
	int64_t max_tile_elems = 1;
};

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
	double mults = 0;
	int num_cu = 256;
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
