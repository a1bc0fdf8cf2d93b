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

// One output-element chunk of one basis block and the range of virtual
// instructions it sums over.
struct WorkItem {
	int64_t y_off;     // offset of the chunk's first element in y
	int32_t count;     // elements in the chunk (<= kChunk)
	int32_t nn_out;    // neutron states of the output block
	int32_t elem0;     // block-local index of the chunk's first element
	int32_t vi_begin;  // virtual instruction range
	int32_t vi_end;
	int32_t pad;
};

// An instruction applied in one direction: forward (H_out,in x_in) or, for
// off-diagonal instructions, transposed (H_out,in^T x_out into y_in).  Row
// pointers select the CSR copy of each list sorted by this direction's
// output index.
struct VInstr {
	int64_t x_off;     // offset of the input block in x
	int64_t v_off;     // first matrix element
	int64_t dn;        // neutron matrix dimension (NP stride)
	uint32_t n_rows;   // first row pointer of the neutron list copy
	uint32_t p_rows;   // first row pointer of the proton list copy
	int32_t nn_in;     // neutron states of the input block
	int16_t type;      // 1 N, 2 P, 3 NP
	int16_t force;     // 1, 2, 3 body
};

// One neutron-entry chunk of an NP instruction (one direction) against a
// range of proton output rows.
struct NPWork {
	int64_t y_off;     // output block offset in y
	int64_t x_off;     // input block offset in x
	int64_t v_off;     // first matrix element of the block
	int64_t dn;        // neutron matrix dimension (row stride of V)
	uint32_t n_begin;  // first neutron entry of the chunk (absolute)
	uint32_t p_rows;   // row pointer of proton output row 0
	int32_t n_count;   // neutron entries in the chunk (<= kChunk)
	int32_t nn_in;     // neutron states of the input block
	int32_t nn_out;    // neutron states of the output block
	int32_t po_begin;  // proton output rows [po_begin, po_end)
	int32_t po_end;
	int32_t force;     // 1, 2, 3 body
};

constexpr int kChunk = 256;
constexpr uint32_t kProtonRange = 4096;
constexpr int kXcds = 8;

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
	bacchus::VInstr *d_vinstr = nullptr;
	bacchus::WorkItem *d_work = nullptr;
	uint32_t *d_keys = nullptr;     // row of every CSR entry
	bacchus::NPWork *d_npwork = nullptr;
	int64_t num_npwork = 0;
	float *d_xf = nullptr;          // FP32 copy of x (FP32/BF16 modes)
	int64_t num_work = 0;
	int64_t num_vinstr = 0;
	// statistics
	double algorithmic_bytes = 0;
	double flops = 0;
	int64_t element_count = 0;
	int64_t list_entry_count = 0;
};
