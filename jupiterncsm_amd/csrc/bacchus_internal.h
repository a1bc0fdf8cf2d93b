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
	int32_t pad;
};

// One output tile: rows [po_begin, po_end) of a basis block (all neutron
// states) and the range of virtual instructions summed into it.
struct TileWork {
	int64_t y_off;     // offset of the block in y
	int32_t nn_out;    // neutron states of the block
	int32_t po_begin;
	int32_t po_end;
	int32_t vi_begin;
	int32_t vi_end;
	int32_t shared;    // other tiles cover the same rows: flush atomically
	int32_t t_begin;   // tasks [t_begin, t_end)
	int32_t t_end;
	int32_t pad;
};

// One unit of a tile's work: entries [jb, je) of an N instruction over all
// rows (r = -1), or proton entries [jb, je) (rows of the tile, in row
// order) of a P/NP instruction with the range [e0, e1) of neutron entries
// (NP) or neutron states (P) it covers.
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
constexpr uint32_t kNTaskEntries = 2048;  // list entries per N task
constexpr uint64_t kTaskMults = 65536;     // multiplications per P/NP task

// Tiling of one operator's output: the tiles (rows of a block x all
// neutron states x bt vectors) with their tasks, heaviest first.
struct PlanParams {
	double target = 0;          // multiplications per tile (x bt)
	std::vector<int> bt;        // vectors per tile, per basis block
	int64_t tile_elems = kTileMaxElems;  // LDS words per tile (rows*nn*bt)
	uint64_t task_mults = kTaskMults;    // multiplications x bt per task
};
struct Plan {
	std::vector<TileWork> tiles;
	std::vector<Task> tasks;
	std::vector<double> cost;   // multiplications x bt of every tile
	std::vector<int> bt;        // vectors of every tile
	int64_t max_tile_elems = 1;
};

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
	double algorithmic_bytes = 0;
	double flops = 0;
	int64_t element_count = 0;
	int64_t compact_elements = 0;  // elements of the dense sub-blocks
	int64_t list_entry_count = 0;
};

namespace bacchus {
int make_plan(const bacchus_hip_operator *op, const PlanParams &pp, Plan &out);
}
