// operator.hip - device tables and the gfx950 matvec kernel of the Bacchus
// hot path.
//
// The reference runs the six scalar scatter loops of
// src/Minerva/matrix_vector_multiplication/matrix_vector_multiplication.c:
// 20-462 from an OpenMP instruction loop (src/Minerva/scheduler/
// scheduler.c:97-137) with per-thread output replicas.  Here the product is
// computed owner-side instead: every output element is summed by one lane
// from CSR copies of the index lists sorted by output index, so there are no
// scatter conflicts inside an instruction.  Off-diagonal instructions are
// applied twice, forward and transposed (scheduler.c:310-418), the
// transposed direction reading the list copy sorted by input index.
#include "bacchus_internal.h"

#include <hip/hip_bf16.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>

namespace bacchus {

static thread_local std::string g_last_error;

int set_error(const char *func, const char *file, int line, const char *fmt,
	      ...)
{
	char msg[1024];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(msg, sizeof msg, fmt, ap);
	va_end(ap);
	const char *base = strrchr(file, '/');
	char full[1400];
	snprintf(full, sizeof full, "Error in %s @ %s:%d: %s", func,
		 base ? base + 1 : file, line, msg);
	g_last_error = full;
	return 1;
}

namespace {

constexpr uint32_t kIndexMask = 0x7FFFFFFFu;

__device__ __forceinline__ double apply_sign(double v, uint32_t m)
{
	return __longlong_as_double(__double_as_longlong(v) ^
				    ((long long)(m & 0x80000000u) << 32));
}
__device__ __forceinline__ float apply_sign(float v, uint32_t m)
{
	return __int_as_float(__float_as_int(v) ^ (int)(m & 0x80000000u));
}

template <typename VT> struct Elem;
template <> struct Elem<double> {
	using acc_t = double;
	using x_t = double;
	__device__ static double load(const double *p, int64_t i) { return p[i]; }
};
template <> struct Elem<float> {
	using acc_t = float;
	using x_t = float;
	__device__ static float load(const float *p, int64_t i) { return p[i]; }
};
template <> struct Elem<uint16_t> {
	using acc_t = float;
	using x_t = float;
	__device__ static float load(const uint16_t *p, int64_t i)
	{
		return __uint_as_float(uint32_t(p[i]) << 16);
	}
};

// One lane per output element.  For element (n_o, p_o) of the output block:
//   N  : sum_{i in Ln(n_o)}  s_i V[m_i] x[n_i + Nn_in p_o]       (c.f. :44-76)
//   P  : sum_{j in Lp(p_o)}  s_j V[m_j] x[n_o + Nn_in p_j]       (c.f. :99-131)
//   NP : sum_{j in Lp(p_o)} s_j sum_{i in Ln(n_o)} s_i V[m_i + Dn m_j]
//                                      x[n_i + Nn_in p_j]        (c.f. :164-209)
// The N and P forms skip |V| < 1e-12 as the reference does (:52, :107).
template <typename VT>
__global__ __launch_bounds__(kChunk) void matvec_kernel(
	const WorkItem *__restrict__ work, const VInstr *__restrict__ vins,
	const uint32_t *__restrict__ rowptr, const uint2 *__restrict__ ent,
	const VT *__restrict__ V, const typename Elem<VT>::x_t *__restrict__ x,
	double *__restrict__ y, double s1, double s2, double s3)
{
	using acc_t = typename Elem<VT>::acc_t;
	const WorkItem w = work[blockIdx.x];
	const int t = threadIdx.x;
	if (t >= w.count)
		return;
	const int e = w.elem0 + t;
	const int no = e % w.nn_out;
	const int po = e / w.nn_out;
	acc_t acc = 0;
	for (int vi = w.vi_begin; vi < w.vi_end; vi++) {
		const VInstr I = vins[vi];
		const acc_t sc = acc_t(I.force == 1 ? s1 : (I.force == 2 ? s2 : s3));
		acc_t part = 0;
		if (I.type == 1) {
			const uint32_t b = rowptr[I.n_rows + no];
			const uint32_t f = rowptr[I.n_rows + no + 1];
			const auto *xc = x + I.x_off + (int64_t)I.nn_in * po;
			for (uint32_t k = b; k < f; k++) {
				const uint2 en = ent[k];
				const acc_t v = Elem<VT>::load(V, I.v_off + (en.y & kIndexMask));
				if (fabs(v) < acc_t(1e-12))
					continue;
				part += apply_sign(v, en.y) * acc_t(xc[en.x]);
			}
		} else if (I.type == 2) {
			const uint32_t b = rowptr[I.p_rows + po];
			const uint32_t f = rowptr[I.p_rows + po + 1];
			const auto *xc = x + I.x_off + no;
			for (uint32_t k = b; k < f; k++) {
				const uint2 en = ent[k];
				const acc_t v = Elem<VT>::load(V, I.v_off + (en.y & kIndexMask));
				if (fabs(v) < acc_t(1e-12))
					continue;
				part += apply_sign(v, en.y) *
					acc_t(xc[(int64_t)I.nn_in * en.x]);
			}
		} else {
			const uint32_t pb = rowptr[I.p_rows + po];
			const uint32_t pf = rowptr[I.p_rows + po + 1];
			const uint32_t nb = rowptr[I.n_rows + no];
			const uint32_t nf = rowptr[I.n_rows + no + 1];
			for (uint32_t j = pb; j < pf; j++) {
				const uint2 pe = ent[j];
				const int64_t vrow = I.v_off + I.dn * (pe.y & kIndexMask);
				const auto *xc = x + I.x_off + (int64_t)I.nn_in * pe.x;
				acc_t inner = 0;
				for (uint32_t k = nb; k < nf; k++) {
					const uint2 ne = ent[k];
					const acc_t v = Elem<VT>::load(V, vrow + (ne.y & kIndexMask));
					inner += apply_sign(v, ne.y) * acc_t(xc[ne.x]);
				}
				part += apply_sign(inner, pe.y);
			}
		}
		acc += sc * part;
	}
	atomicAdd(&y[w.y_off + t], (double)acc);
}

// XCD-aware order: the hardware deals workgroups round-robin over the
// XCDs, so logical item (xcd * per + k) is given to the k-th block of that
// XCD and neighbouring items (same element block) share one L2.
__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t n)
{
	const int64_t per = n / kXcds;
	return (b % kXcds) * per + b / kXcds;
}

// NP product (c.f. matrix_vector_multiplication.c:134-210 and the
// off-diagonal form :364-462).  Lane i holds neutron entry i of a chunk of
// the list copy sorted by output row (then element index); the proton
// entries of one output row p_o are wave-uniform (scalar loads):
//   acc_i = s_i * sum_{j in Lp(p_o)} s_j V[m_i + Dn m_j] x[n_i + Nn_in p_j]
// then lanes with equal output row n_o are summed by a segmented
// reduction and the first lane of each run adds the sum to y[n_o + Nn_out p_o].
template <typename VT>
__global__ __launch_bounds__(kChunk) void np_kernel(
	const NPWork *__restrict__ work, int64_t nwork,
	const uint32_t *__restrict__ rowptr, const uint2 *__restrict__ ent,
	const uint32_t *__restrict__ keys, const VT *__restrict__ V,
	const typename Elem<VT>::x_t *__restrict__ x, double *__restrict__ y,
	double s1, double s2, double s3)
{
	using acc_t = typename Elem<VT>::acc_t;
	const NPWork w = work[xcd_remap(blockIdx.x, nwork)];
	const int t = threadIdx.x;
	const int lane = t & 63;
	const bool valid = t < w.n_count;
	const uint32_t k = w.n_begin + (valid ? t : 0);
	const uint2 ne = ent[k];
	const uint32_t key = valid ? keys[k] : 0xFFFFFFFFu;
	// step s of the segmented reduction adds lane + 2^s when it is in the
	// same run of equal keys (runs are contiguous: the copy is sorted)
	uint32_t add = 0;
#pragma unroll
	for (int s = 0; s < 6; s++) {
		const uint32_t other = __shfl_down(key, 1 << s, 64);
		if (lane + (1 << s) < 64 && other == key)
			add |= 1u << s;
	}
	const uint32_t prev = __shfl_up(key, 1, 64);
	const bool head = valid && (lane == 0 || prev != key);
	const acc_t sc = apply_sign(acc_t(w.force == 1 ? s1 : (w.force == 2 ? s2 : s3)), ne.y);
	const VT *vrow0 = V + w.v_off + (ne.y & kIndexMask);
	const auto *xcol0 = x + w.x_off + ne.x;
	double *ycol0 = y + w.y_off + key;
	for (int po = w.po_begin; po < w.po_end; po++) {
		const uint32_t jb = rowptr[w.p_rows + po];
		const uint32_t je = rowptr[w.p_rows + po + 1];
		if (jb == je)
			continue;
		acc_t acc = 0;
		for (uint32_t j = jb; j < je; j++) {
			const uint2 pe = ent[j];
			const acc_t v = Elem<VT>::load(vrow0, w.dn * (int64_t)(pe.y & kIndexMask));
			const acc_t xv = acc_t(xcol0[(int64_t)w.nn_in * pe.x]);
			acc += apply_sign(v, pe.y) * xv;
		}
		acc *= sc;
#pragma unroll
		for (int s = 0; s < 6; s++) {
			const acc_t o = __shfl_down(acc, 1 << s, 64);
			if (add & (1u << s))
				acc += o;
		}
		if (head)
			atomicAdd(ycol0 + (int64_t)w.nn_out * po, (double)acc);
	}
}

__global__ void scale_kernel(double *y, int64_t n, double s)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		y[i] = s == 0.0 ? 0.0 : y[i] * s;
}

__global__ void to_float_kernel(const double *x, float *xf, int64_t n)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		xf[i] = (float)x[i];
}

template <typename VT>
void launch(const bacchus_hip_operator *op, const VT *V, const typename Elem<VT>::x_t *x,
	    double *y, double s1, double s2, double s3, hipStream_t s)
{
	if (op->num_npwork)
		hipLaunchKernelGGL(np_kernel<VT>, dim3(op->num_npwork), dim3(kChunk), 0, s,
				   op->d_npwork, op->num_npwork, op->d_rowptr, op->d_entries,
				   op->d_keys, V, x, y, s1, s2, s3);
	if (op->num_work)
		hipLaunchKernelGGL(matvec_kernel<VT>, dim3(op->num_work), dim3(kChunk), 0, s,
				   op->d_work, op->d_vinstr, op->d_rowptr, op->d_entries, V, x, y,
				   s1, s2, s3);
}

uint16_t to_bf16(double v)
{
	float f = (float)v;
	uint32_t u;
	memcpy(&u, &f, 4);
	if ((u & 0x7f800000u) == 0x7f800000u)
		return (uint16_t)(u >> 16);
	u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
	return (uint16_t)(u >> 16);
}

struct ListUse {
	int64_t n_in = -1, n_out = -1;
	bool used = false;
};

int check_list_use(ListUse &u, int64_t n_in, int64_t n_out, int64_t list)
{
	if (!u.used) {
		u = {n_in, n_out, true};
		return 0;
	}
	if (u.n_in != n_in || u.n_out != n_out)
		return BACCHUS_FAIL("index list %lld is used between sub-bases of "
				    "different sizes (%lld->%lld and %lld->%lld)",
				    (long long)list, (long long)u.n_in,
				    (long long)u.n_out, (long long)n_in,
				    (long long)n_out);
	return 0;
}

}  // namespace
}  // namespace bacchus

using namespace bacchus;

extern "C" {

const char *bacchus_hip_last_error(void) { return g_last_error.c_str(); }

int64_t bacchus_hip_operator_dimension(const bacchus_hip_operator_t *op)
{
	return op ? op->dimension : -1;
}

void bacchus_hip_operator_free(bacchus_hip_operator_t *op)
{
	if (!op)
		return;
	(void)hipFree(op->d_v);
	(void)hipFree(op->d_rowptr);
	(void)hipFree(op->d_entries);
	(void)hipFree(op->d_vinstr);
	(void)hipFree(op->d_work);
	(void)hipFree(op->d_keys);
	(void)hipFree(op->d_npwork);
	(void)hipFree(op->d_xf);
	delete op;
}

int bacchus_hip_operator_set_scales(bacchus_hip_operator_t *op, double s1,
				    double s2, double s3)
{
	if (!op)
		return BACCHUS_FAIL("operator is NULL");
	op->scales[0] = s1;
	op->scales[1] = s2;
	op->scales[2] = s3;
	return 0;
}

int bacchus_hip_operator_stats(const bacchus_hip_operator_t *op,
			       double *algorithmic_bytes, double *flops,
			       int64_t *num_work_items)
{
	if (!op)
		return BACCHUS_FAIL("operator is NULL");
	if (algorithmic_bytes)
		*algorithmic_bytes = op->algorithmic_bytes;
	if (flops)
		*flops = op->flops;
	if (num_work_items)
		*num_work_items = op->num_work;
	return 0;
}

int bacchus_hip_operator_create(const bacchus_hip_problem_t *p,
				int precision, bacchus_hip_operator_t **out)
{
	if (!p || !out)
		return BACCHUS_FAIL("NULL argument");
	*out = nullptr;
	if (precision < BACCHUS_HIP_FP64 || precision > BACCHUS_HIP_BF16)
		return BACCHUS_FAIL("unknown precision %d", precision);
	if (p->num_blocks <= 0)
		return BACCHUS_FAIL("no basis blocks");
	const int64_t nb = p->num_blocks, nl = p->num_lists,
		      nv = p->num_matrix_blocks, ni = p->num_instructions;
	auto *op = new bacchus_hip_operator;
	op->precision = precision;
	op->block_np.assign(p->block_np, p->block_np + nb);
	op->block_nn.assign(p->block_nn, p->block_nn + nb);
	op->block_off.resize(nb + 1, 0);
	op->block_ids.resize(nb);
	for (int64_t b = 0; b < nb; b++)
		op->block_ids[b] = b + 1;
	for (int64_t b = 0; b < nb; b++) {
		if (op->block_np[b] <= 0 || op->block_nn[b] <= 0 ||
		    op->block_np[b] * op->block_nn[b] > INT32_MAX) {
			delete op;
			return BACCHUS_FAIL("basis block %lld has invalid size %lldx%lld",
					    (long long)b, (long long)op->block_np[b],
					    (long long)op->block_nn[b]);
		}
		op->block_off[b + 1] = op->block_off[b] + op->block_np[b] * op->block_nn[b];
	}
	op->dimension = op->block_off[nb];
	auto fail = [&](int code) {
		delete op;
		return code;
	};

	// --- host validation: every index a kernel will touch is in bounds ---
	std::vector<ListUse> use(nl);
	std::vector<int64_t> vlen(nv), vdn(nv);
	for (int64_t v = 0; v < nv; v++) {
		vlen[v] = p->matrix_offsets[v + 1] - p->matrix_offsets[v];
		vdn[v] = p->matrix_dn[v];
		if (vlen[v] < 0)
			return fail(BACCHUS_FAIL("matrix block %lld has negative size",
						 (long long)v));
	}
	std::vector<uint32_t> list_max_m(nl, 0);
	for (int64_t l = 0; l < nl; l++)
		for (int64_t k = p->list_offsets[l]; k < p->list_offsets[l + 1]; k++)
			list_max_m[l] = std::max(list_max_m[l],
						 (uint32_t)p->list_entries[3 * k + 2] & 0x7FFFFFFFu);
	std::vector<int64_t> list_max_in(nl, -1), list_max_out(nl, -1);
	for (int64_t l = 0; l < nl; l++)
		for (int64_t k = p->list_offsets[l]; k < p->list_offsets[l + 1]; k++) {
			const int32_t a = p->list_entries[3 * k], b = p->list_entries[3 * k + 1];
			if (a < 0 || b < 0)
				return fail(BACCHUS_FAIL("index list %lld has a negative index",
							 (long long)l));
			list_max_in[l] = std::max<int64_t>(list_max_in[l], a);
			list_max_out[l] = std::max<int64_t>(list_max_out[l], b);
		}
	double mults = 0;
	for (int64_t i = 0; i < ni; i++) {
		const bacchus_hip_instruction_t &I = p->instructions[i];
		if (I.block_in < 0 || I.block_in >= nb || I.block_out < 0 ||
		    I.block_out >= nb || I.matrix_block < 0 || I.matrix_block >= nv)
			return fail(BACCHUS_FAIL("instruction %lld references a missing "
						 "block", (long long)i));
		const int64_t bi = I.block_in, bo = I.block_out, v = I.matrix_block;
		const bool need_n = I.type == 1 || I.type == 3;
		const bool need_p = I.type == 2 || I.type == 3;
		if (I.type < 1 || I.type > 3)
			return fail(BACCHUS_FAIL("instruction %lld has unknown type %d",
						 (long long)i, I.type));
		if ((need_n && (I.neutron_list < 0 || I.neutron_list >= nl)) ||
		    (need_p && (I.proton_list < 0 || I.proton_list >= nl)))
			return fail(BACCHUS_FAIL("instruction %lld references a missing "
						 "index list", (long long)i));
		if (I.type == 1 && op->block_np[bi] != op->block_np[bo])
			return fail(BACCHUS_FAIL("neutron instruction %lld joins blocks with "
						 "different proton dimensions", (long long)i));
		if (I.type == 2 && op->block_nn[bi] != op->block_nn[bo])
			return fail(BACCHUS_FAIL("proton instruction %lld joins blocks with "
						 "different neutron dimensions", (long long)i));
		if (need_n && check_list_use(use[I.neutron_list], op->block_nn[bi],
					     op->block_nn[bo], I.neutron_list))
			return fail(1);
		if (need_p && check_list_use(use[I.proton_list], op->block_np[bi],
					     op->block_np[bo], I.proton_list))
			return fail(1);
		int64_t Dn, Dp;
		if (I.type == 1) {
			Dn = vlen[v];
			Dp = 1;
		} else if (I.type == 2) {
			Dn = 1;
			Dp = vlen[v];
		} else {
			Dn = vdn[v];
			if (Dn <= 0 || vlen[v] % Dn != 0)
				return fail(BACCHUS_FAIL("matrix block %lld: %lld elements do not "
							 "split into neutron dimension %lld",
							 (long long)v, (long long)vlen[v],
							 (long long)Dn));
			Dp = vlen[v] / Dn;
		}
		if (need_n && p->list_offsets[I.neutron_list + 1] > p->list_offsets[I.neutron_list] &&
		    (int64_t)list_max_m[I.neutron_list] >= Dn)
			return fail(BACCHUS_FAIL("index list %d addresses element %u of a "
						 "neutron dimension %lld", I.neutron_list,
						 list_max_m[I.neutron_list], (long long)Dn));
		if (need_p && p->list_offsets[I.proton_list + 1] > p->list_offsets[I.proton_list] &&
		    (int64_t)list_max_m[I.proton_list] >= Dp)
			return fail(BACCHUS_FAIL("index list %d addresses element %u of a "
						 "proton dimension %lld", I.proton_list,
						 list_max_m[I.proton_list], (long long)Dp));
		const int64_t ln = need_n ? p->list_offsets[I.neutron_list + 1] -
						    p->list_offsets[I.neutron_list] : 0;
		const int64_t lp = need_p ? p->list_offsets[I.proton_list + 1] -
						    p->list_offsets[I.proton_list] : 0;
		const double m = I.type == 1 ? double(ln) * op->block_np[bi]
				 : I.type == 2 ? double(lp) * op->block_nn[bi]
					       : double(ln) * lp;
		mults += (bi == bo) ? m : 2 * m;
	}
	for (int64_t l = 0; l < nl; l++) {
		if (!use[l].used)
			continue;
		if (list_max_in[l] >= use[l].n_in || list_max_out[l] >= use[l].n_out)
			return fail(BACCHUS_FAIL("index list %lld indexes beyond its "
						 "sub-bases (%lld/%lld, %lld/%lld)",
						 (long long)l, (long long)list_max_in[l],
						 (long long)use[l].n_in,
						 (long long)list_max_out[l],
						 (long long)use[l].n_out));
	}

	// --- CSR copies of every used list: by output and by input index,
	// each row ordered by matrix index (neighbouring lanes then read
	// neighbouring elements); keys[k] = row of entry k ---
	std::vector<uint32_t> rowptr;
	std::vector<uint2> entries;
	std::vector<uint32_t> keys;
	std::vector<int64_t> rows_by_out(nl, -1), rows_by_in(nl, -1);
	int64_t used_entries = 0;
	for (int64_t l = 0; l < nl; l++) {
		if (!use[l].used)
			continue;
		const int64_t a = p->list_offsets[l], b = p->list_offsets[l + 1];
		used_entries += b - a;
		for (int dir = 0; dir < 2; dir++) {
			const int64_t nrows = dir == 0 ? use[l].n_out : use[l].n_in;
			const int key = dir == 0 ? 1 : 0;  // sort field: out or in
			(dir == 0 ? rows_by_out : rows_by_in)[l] = (int64_t)rowptr.size();
			std::vector<uint32_t> count(nrows + 1, 0);
			for (int64_t k = a; k < b; k++)
				count[p->list_entries[3 * k + key] + 1]++;
			const uint32_t base = (uint32_t)entries.size();
			for (int64_t r = 0; r < nrows; r++)
				count[r + 1] += count[r];
			for (int64_t r = 0; r <= nrows; r++)
				rowptr.push_back(base + count[r]);
			entries.resize(entries.size() + (b - a));
			keys.resize(keys.size() + (b - a));
			std::vector<uint32_t> fill(count.begin(), count.end() - 1);
			for (int64_t k = a; k < b; k++) {
				const int32_t *e3 = &p->list_entries[3 * k];
				const uint32_t r = (uint32_t)e3[key];
				uint2 en;
				en.x = (uint32_t)e3[1 - key];
				en.y = (uint32_t)e3[2];
				keys[base + fill[r]] = r;
				entries[base + fill[r]++] = en;
			}
			for (int64_t r = 0; r < nrows; r++)
				std::sort(entries.begin() + base + count[r],
					  entries.begin() + base + count[r + 1],
					  [](const uint2 &u, const uint2 &w) {
						  return (u.y & 0x7FFFFFFFu) < (w.y & 0x7FFFFFFFu);
					  });
			if (entries.size() > 0xFFFFFFFFull || rowptr.size() > 0xFFFFFFFFull)
				return fail(BACCHUS_FAIL("index lists exceed 2^32 entries"));
		}
	}
	op->list_entry_count = used_entries;
	// Optional ordering of the neutron lists of NP instructions by element
	// index alone (coalesced V rows, shorter reduction runs); rows stay in
	// keys[] so the NP kernel's segmented reduction is unchanged.
	const char *order_env = getenv("BACCHUS_HIP_NP_ORDER");
	const bool element_order = order_env && strcmp(order_env, "element") == 0;
	std::vector<int64_t> elem_by_out(nl, -1), elem_by_in(nl, -1);
	if (element_order) {
		for (int64_t i = 0; i < ni; i++) {
			const bacchus_hip_instruction_t &I = p->instructions[i];
			if (I.type != 3 || elem_by_out[I.neutron_list] >= 0)
				continue;
			const int64_t l = I.neutron_list;
			for (int dir = 0; dir < 2; dir++) {
				const int64_t r0 = (dir == 0 ? rows_by_out : rows_by_in)[l];
				const int64_t nrows = dir == 0 ? use[l].n_out : use[l].n_in;
				const uint32_t a = rowptr[r0], b = rowptr[r0 + nrows];
				std::vector<uint32_t> perm(b - a);
				std::iota(perm.begin(), perm.end(), a);
				std::stable_sort(perm.begin(), perm.end(), [&](uint32_t u, uint32_t w) {
					return (entries[u].y & 0x7FFFFFFFu) < (entries[w].y & 0x7FFFFFFFu);
				});
				(dir == 0 ? elem_by_out : elem_by_in)[l] = (int64_t)entries.size();
				for (uint32_t k : perm) {
					entries.push_back(entries[k]);
					keys.push_back(keys[k]);
				}
			}
		}
		if (entries.size() > 0xFFFFFFFFull)
			return fail(BACCHUS_FAIL("index lists exceed 2^32 entries"));
	}

	// --- virtual instructions: N and P grouped by output block; NP split
	// into (neutron-entry chunk x proton output-row range) work items ---
	std::vector<std::vector<VInstr>> per_block(nb);
	std::vector<std::vector<double>> per_block_cost(nb);
	struct NPV {
		NPWork proto;
		int64_t n_base, n_len, p_rows, n_pout;
	};
	std::vector<NPV> npv;
	for (int64_t i = 0; i < ni; i++) {
		const bacchus_hip_instruction_t &I = p->instructions[i];
		const int force = p->matrix_force ? p->matrix_force[I.matrix_block] : 2;
		const int64_t ln = I.neutron_list >= 0 ? p->list_offsets[I.neutron_list + 1] -
								 p->list_offsets[I.neutron_list] : 0;
		const int64_t lp = I.proton_list >= 0 ? p->list_offsets[I.proton_list + 1] -
								p->list_offsets[I.proton_list] : 0;
		for (int dir = 0; dir < (I.block_in == I.block_out ? 1 : 2); dir++) {
			const int64_t bin = dir == 0 ? I.block_in : I.block_out;
			const int64_t bout = dir == 0 ? I.block_out : I.block_in;
			const auto &rows = dir == 0 ? rows_by_out : rows_by_in;
			const int f = std::min(3, std::max(1, force));
			if (I.type == 3) {
				if (ln == 0 || lp == 0)
					continue;
				NPV q{};
				q.proto.y_off = op->block_off[bout];
				q.proto.x_off = op->block_off[bin];
				q.proto.v_off = p->matrix_offsets[I.matrix_block];
				q.proto.dn = vdn[I.matrix_block];
				q.proto.nn_in = (int32_t)op->block_nn[bin];
				q.proto.nn_out = (int32_t)op->block_nn[bout];
				q.proto.force = f;
				const int64_t nr = rows[I.neutron_list];
				q.n_base = element_order
					? (dir == 0 ? elem_by_out : elem_by_in)[I.neutron_list]
					: (int64_t)rowptr[nr];
				q.n_len = ln;
				q.p_rows = rows[I.proton_list];
				q.n_pout = op->block_np[bout];
				npv.push_back(q);
				continue;
			}
			VInstr v{};
			v.x_off = op->block_off[bin];
			v.v_off = p->matrix_offsets[I.matrix_block];
			v.dn = 0;
			v.nn_in = (int32_t)op->block_nn[bin];
			v.type = (int16_t)I.type;
			v.force = (int16_t)f;
			if (I.type == 1)
				v.n_rows = (uint32_t)rows[I.neutron_list];
			else
				v.p_rows = (uint32_t)rows[I.proton_list];
			const double m = I.type == 1 ? double(ln) * op->block_np[bin]
						     : double(lp) * op->block_nn[bin];
			per_block[bout].push_back(v);
			per_block_cost[bout].push_back(m);
		}
	}
	// --- N/P work items: 256-element chunks x instruction ranges ---
	std::vector<VInstr> vinstr;
	std::vector<WorkItem> work;
	int num_cu = 256;
	{
		int dev = 0;
		hipDeviceProp_t prop;
		if (hipGetDevice(&dev) == hipSuccess &&
		    hipGetDeviceProperties(&prop, dev) == hipSuccess)
			num_cu = prop.multiProcessorCount;
	}
	double np_mults = 0;
	for (auto &c : per_block_cost)
		for (double v : c)
			np_mults += v;
	const double target = std::max(2.0e4, np_mults / (num_cu * 32.0));
	for (int64_t b = 0; b < nb; b++) {
		auto &vs = per_block[b];
		auto &cs = per_block_cost[b];
		if (vs.empty())
			continue;
		std::vector<int64_t> order(vs.size());
		std::iota(order.begin(), order.end(), 0);
		std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t c) {
			return vs[a].v_off < vs[c].v_off;
		});
		const int32_t first = (int32_t)vinstr.size();
		std::vector<double> cost;
		for (int64_t k : order) {
			vinstr.push_back(vs[k]);
			cost.push_back(cs[k]);
		}
		const int64_t size = op->block_np[b] * op->block_nn[b];
		for (int64_t e0 = 0; e0 < size; e0 += kChunk) {
			const int32_t cnt = (int32_t)std::min<int64_t>(kChunk, size - e0);
			const double frac = double(cnt) / double(size);
			int32_t begin = 0;
			double acc = 0;
			for (int32_t k = 0; k < (int32_t)cost.size(); k++) {
				acc += cost[k] * frac + 64.0;
				if (acc >= target || k + 1 == (int32_t)cost.size()) {
					WorkItem w{};
					w.y_off = op->block_off[b] + e0;
					w.count = cnt;
					w.nn_out = (int32_t)op->block_nn[b];
					w.elem0 = (int32_t)e0;
					w.vi_begin = first + begin;
					w.vi_end = first + k + 1;
					work.push_back(w);
					begin = k + 1;
					acc = 0;
				}
			}
		}
	}
	op->num_work = (int64_t)work.size();
	op->num_vinstr = (int64_t)vinstr.size();

	// --- NP work items, grouped by element block (V locality), then
	// proton output-row range, then neutron chunk ---
	std::stable_sort(npv.begin(), npv.end(), [](const NPV &a, const NPV &b) {
		return a.proto.v_off < b.proto.v_off;
	});
	std::vector<NPWork> npwork;
	for (const NPV &q : npv) {
		// proton output rows split so a work item sees <= kProtonRange
		// proton entries (a single long row stays whole)
		int32_t po0 = 0;
		while (po0 < q.n_pout) {
			int32_t po1 = po0;
			uint32_t taken = 0;
			while (po1 < q.n_pout) {
				const uint32_t len = rowptr[q.p_rows + po1 + 1] - rowptr[q.p_rows + po1];
				if (po1 > po0 && taken + len > kProtonRange)
					break;
				taken += len;
				po1++;
			}
			if (taken > 0) {
				for (int64_t c = 0; c < q.n_len; c += kChunk) {
					NPWork w = q.proto;
					w.n_begin = (uint32_t)(q.n_base + c);
					w.n_count = (int32_t)std::min<int64_t>(kChunk, q.n_len - c);
					w.p_rows = (uint32_t)q.p_rows;
					w.po_begin = po0;
					w.po_end = po1;
					npwork.push_back(w);
				}
			}
			po0 = po1;
		}
	}
	// pad to a multiple of the XCD count so the XCD-aware block remap
	// is a bijection (empty items do nothing)
	while (npwork.size() % kXcds)
		npwork.push_back(NPWork{});
	op->num_npwork = (int64_t)npwork.size();

	// --- statistics: compulsory bytes of referenced arrays ---
	std::vector<char> vref(nv, 0);
	for (int64_t i = 0; i < ni; i++)
		vref[p->instructions[i].matrix_block] = 1;
	int64_t elems = 0;
	for (int64_t v = 0; v < nv; v++)
		if (vref[v])
			elems += vlen[v];
	op->element_count = elems;
	const double sv = precision == BACCHUS_HIP_FP64 ? 8 : precision == BACCHUS_HIP_FP32 ? 4 : 2;
	op->algorithmic_bytes = sv * double(elems) + 12.0 * double(used_entries) +
				16.0 * double(op->dimension);
	op->flops = 2.0 * mults;

	// --- upload ---
	const int64_t total_v = nv ? p->matrix_offsets[nv] : 0;
	auto upload = [&](void **dst, const void *src, size_t bytes) -> int {
		if (bytes == 0)
			bytes = 16;
		BACCHUS_HIP_CHECK(hipMalloc(dst, bytes));
		if (src)
			BACCHUS_HIP_CHECK(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
		return 0;
	};
	int rc = 0;
	if (precision == BACCHUS_HIP_FP64) {
		rc = upload(&op->d_v, p->matrix_elements, total_v * sizeof(double));
	} else if (precision == BACCHUS_HIP_FP32) {
		std::vector<float> tmp(total_v);
		for (int64_t k = 0; k < total_v; k++)
			tmp[k] = (float)p->matrix_elements[k];
		rc = upload(&op->d_v, tmp.data(), total_v * sizeof(float));
	} else {
		std::vector<uint16_t> tmp(total_v);
		for (int64_t k = 0; k < total_v; k++)
			tmp[k] = to_bf16(p->matrix_elements[k]);
		rc = upload(&op->d_v, tmp.data(), total_v * sizeof(uint16_t));
	}
	if (rc)
		return fail(rc);
	if ((rc = upload((void **)&op->d_rowptr, rowptr.data(), rowptr.size() * 4)) ||
	    (rc = upload((void **)&op->d_entries, entries.data(), entries.size() * 8)) ||
	    (rc = upload((void **)&op->d_keys, keys.data(), keys.size() * 4)) ||
	    (rc = upload((void **)&op->d_npwork, npwork.data(), npwork.size() * sizeof(NPWork))) ||
	    (rc = upload((void **)&op->d_vinstr, vinstr.data(), vinstr.size() * sizeof(VInstr))) ||
	    (rc = upload((void **)&op->d_work, work.data(), work.size() * sizeof(WorkItem))))
		return fail(rc);
	if (precision != BACCHUS_HIP_FP64 &&
	    (rc = upload((void **)&op->d_xf, nullptr, op->dimension * sizeof(float))))
		return fail(rc);
	*out = op;
	return 0;
}

int bacchus_hip_operator_matvec(bacchus_hip_operator_t *op, const double *x,
				double *y, int accumulate, void *stream)
{
	if (!op || !x || !y)
		return BACCHUS_FAIL("NULL argument");
	hipStream_t s = (hipStream_t)stream;
	const int64_t n = op->dimension;
	const int threads = 256;
	const int64_t blocks = (n + threads - 1) / threads;
	if (!accumulate)
		hipLaunchKernelGGL(scale_kernel, dim3(blocks), dim3(threads), 0, s, y, n, 0.0);
	const double s1 = op->scales[0], s2 = op->scales[1], s3 = op->scales[2];
	if (op->precision == BACCHUS_HIP_FP64) {
		launch<double>(op, (const double *)op->d_v, x, y, s1, s2, s3, s);
	} else {
		hipLaunchKernelGGL(to_float_kernel, dim3(blocks), dim3(threads), 0, s, x, op->d_xf, n);
		if (op->precision == BACCHUS_HIP_FP32)
			launch<float>(op, (const float *)op->d_v, op->d_xf, y, s1, s2, s3, s);
		else
			launch<uint16_t>(op, (const uint16_t *)op->d_v, op->d_xf, y, s1, s2, s3, s);
	}
	BACCHUS_HIP_CHECK(hipGetLastError());
	return 0;
}

}  // extern "C"
