// mlp.hip - the learned Lanczos start: a small MLP (in -> hidden x depth ->
// out, GELU) with BF16 weights feeding MFMA and FP32 accumulation, trained on
// the GPU by hand-written forward/backward kernels and Adam on FP32 master
// weights.  Not a reference component: it predicts coefficients of the
// starting vector over stored ground states for bacchus_hip_lanczos.
#include "bacchus_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace bacchus {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short to_bf16_bits(float f)
{
	uint32_t u = __float_as_uint(f);
	if ((u & 0x7f800000u) != 0x7f800000u)
		u += 0x7fffu + ((u >> 16) & 1u);
	return (short)(u >> 16);
}

__device__ __forceinline__ float bf16_to_float(uint16_t b)
{
	return __uint_as_float(uint32_t(b) << 16);
}

// Operand of a GEMM: element (r, c) at base[r * rs + c * cs], FP32 or BF16.
struct Operand {
	const void *p;
	int64_t rs, cs;
	int bf16;
};

__device__ __forceinline__ float load_op(const Operand &o, int64_t r, int64_t c)
{
	const int64_t i = r * o.rs + c * o.cs;
	return o.bf16 ? bf16_to_float(((const uint16_t *)o.p)[i]) : ((const float *)o.p)[i];
}

__device__ __forceinline__ float gelu(float z)
{
	return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}

__device__ __forceinline__ float gelu_grad(float z)
{
	return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
	       z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// C[M x N] = A[M x K] B[K x N] (+ bias[n]) with BF16 operands on
// v_mfma_f32_16x16x16_bf16 and FP32 accumulation; one wave per 16 x 16
// tile.  Lane l holds A[l%16][4(l/16)+j], B[4(l/16)+j][l%16] and
// D[4(l/16)+j][l%16], j = 0..3.  mode: 0 store, 1 GELU (Z to z_out, GELU(Z)
// to C), 2 accumulate into C.  split: an operand with its bit set (1 = A,
// 2 = B) is carried as two BF16 parts v = hi + lo, hi = bf16(v), lo =
// bf16(v - hi) (16 significant bits), and the product takes the extra
// MFMAs lo_A hi_B and hi_A lo_B.
__global__ __launch_bounds__(64) void gemm_bf16(Operand A, Operand B, int M, int N, int K,
						const float *bias, float *C, int64_t ldc,
						float *z_out, int mode, int split)
{
	const int tn = (N + 15) / 16;
	const int tile = blockIdx.x;
	const int m0 = (tile / tn) * 16, n0 = (tile % tn) * 16;
	const int l = threadIdx.x;
	const int row = l % 16, kg = l / 16;
	f32x4 acc = {0, 0, 0, 0};
	for (int k0 = 0; k0 < K; k0 += 16) {
		bf16x4 a, b, alo, blo;
		for (int j = 0; j < 4; j++) {
			const int k = k0 + 4 * kg + j;
			const int am = m0 + row, bn = n0 + row;
			const float av = (am < M && k < K) ? load_op(A, am, k) : 0.0f;
			const float bv = (bn < N && k < K) ? load_op(B, k, bn) : 0.0f;
			a[j] = to_bf16_bits(av);
			b[j] = to_bf16_bits(bv);
			alo[j] = to_bf16_bits(av - bf16_to_float((uint16_t)a[j]));
			blo[j] = to_bf16_bits(bv - bf16_to_float((uint16_t)b[j]));
		}
		acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
		if (split & 1)
			acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(alo, b, acc, 0, 0, 0);
		if (split & 2)
			acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, blo, acc, 0, 0, 0);
	}
	for (int j = 0; j < 4; j++) {
		const int m = m0 + 4 * kg + j, n = n0 + row;
		if (m >= M || n >= N)
			continue;
		float v = acc[j] + (bias ? bias[n] : 0.0f);
		float *dst = C + (int64_t)m * ldc + n;
		if (mode == 1) {
			z_out[(int64_t)m * ldc + n] = v;
			*dst = gelu(v);
		} else if (mode == 2)
			*dst += v;
		else
			*dst = v;
	}
}

// dZ = dA * GELU'(Z), elementwise
__global__ void gelu_backward(const float *dA, const float *Z, float *dZ, int64_t n)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		dZ[i] = dA[i] * gelu_grad(Z[i]);
}

// db[n] = sum_m dZ[m, n]
__global__ void column_sum(const float *dZ, int M, int N, float *db)
{
	const int n = blockIdx.x * blockDim.x + threadIdx.x;
	if (n >= N)
		return;
	float s = 0;
	for (int m = 0; m < M; m++)
		s += dZ[(int64_t)m * N + n];
	db[n] = s;
}

// mean squared error and its gradient dY = 2 (Y - T) / (M * N)
__global__ void mse_grad(const float *Y, const float *T, int64_t n, float *dY, float *loss)
{
	__shared__ float sh[256];
	float s = 0;
	for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
		const float d = Y[i] - T[i];
		dY[i] = 2.0f * d / (float)n;
		s += d * d;
	}
	sh[threadIdx.x] = s;
	__syncthreads();
	for (int o = blockDim.x / 2; o > 0; o >>= 1) {
		if (threadIdx.x < o)
			sh[threadIdx.x] += sh[threadIdx.x + o];
		__syncthreads();
	}
	if (threadIdx.x == 0)
		*loss = sh[0] / (float)n;
}

// Adam on FP32 master weights, refreshing the BF16 copy
__global__ void adam_step(float *w, uint16_t *w_bf16, const float *g, float *m, float *v,
			  int64_t n, float lr, float b1, float b2, float eps, float c1, float c2)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const float gi = g[i];
	const float mi = b1 * m[i] + (1 - b1) * gi;
	const float vi = b2 * v[i] + (1 - b2) * gi * gi;
	m[i] = mi;
	v[i] = vi;
	const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
	w[i] = wi;
	if (w_bf16)
		w_bf16[i] = (uint16_t)to_bf16_bits(wi);
}

__global__ void to_bf16_kernel(const float *w, uint16_t *b, int64_t n)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		b[i] = (uint16_t)to_bf16_bits(w[i]);
}

int grid(int64_t n) { return (int)((n + 255) / 256); }

}  // namespace
}  // namespace bacchus

using namespace bacchus;

struct bacchus_hip_mlp {
	int in = 0, hidden = 0, depth = 0, out = 0;
	int nlayers = 0;
	std::vector<int> din, dout;            // per layer
	std::vector<int64_t> w_off, b_off;     // offsets into params
	int64_t nparams = 0;
	float *params = nullptr;               // FP32 master: all W then b
	uint16_t *params_bf16 = nullptr;       // BF16 copy (weights used)
	float *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
	int64_t step = 0;
	int precision = 0;                     // BACCHUS_HIP_MLP_* flags
	// activations for a batch of up to cap rows
	int cap = 0;
	std::vector<float *> Z, H;             // per layer pre-/post-activation
	float *dA = nullptr, *dZ = nullptr, *dY = nullptr, *loss = nullptr;
};

namespace {

void free_acts(bacchus_hip_mlp *m)
{
	for (float *p : m->Z)
		(void)hipFree(p);
	for (float *p : m->H)
		(void)hipFree(p);
	m->Z.clear();
	m->H.clear();
	(void)hipFree(m->dA);
	(void)hipFree(m->dZ);
	(void)hipFree(m->dY);
	m->dA = m->dZ = m->dY = nullptr;
	m->cap = 0;
}

int ensure_capacity(bacchus_hip_mlp *m, int M)
{
	if (M <= m->cap)
		return 0;
	free_acts(m);
	const int widest = std::max(m->hidden, m->out);
	m->Z.assign(m->nlayers, nullptr);
	m->H.assign(m->nlayers, nullptr);
	for (int l = 0; l < m->nlayers; l++) {
		BACCHUS_HIP_CHECK(hipMalloc(&m->Z[l], (size_t)M * m->dout[l] * 4));
		BACCHUS_HIP_CHECK(hipMalloc(&m->H[l], (size_t)M * m->dout[l] * 4));
	}
	BACCHUS_HIP_CHECK(hipMalloc(&m->dA, (size_t)M * widest * 4));
	BACCHUS_HIP_CHECK(hipMalloc(&m->dZ, (size_t)M * widest * 4));
	BACCHUS_HIP_CHECK(hipMalloc(&m->dY, (size_t)M * m->out * 4));
	m->cap = M;
	return 0;
}

void gemm(hipStream_t s, Operand A, Operand B, int M, int N, int K, const float *bias,
	  float *C, int64_t ldc, float *z, int mode, int split)
{
	const int tiles = ((M + 15) / 16) * ((N + 15) / 16);
	hipLaunchKernelGGL(gemm_bf16, dim3(tiles), dim3(64), 0, s, A, B, M, N, K, bias, C, ldc,
			   z, mode, split);
}

// The weights of layer l as the GEMM operand W[n, k] at (k, n) (strides
// as given): the BF16 copy, or for the last layer under
// BACCHUS_HIP_MLP_FP32_LAST the FP32 master weights, carried as two BF16
// parts (split flag 2 of gemm_bf16).
Operand weights(const bacchus_hip_mlp *m, int l, int64_t rs, int64_t cs, int *split_b)
{
	if (l + 1 == m->nlayers && (m->precision & BACCHUS_HIP_MLP_FP32_LAST)) {
		*split_b = 2;
		return Operand{m->params + m->w_off[l], rs, cs, 0};
	}
	*split_b = 0;
	return Operand{m->params_bf16 + m->w_off[l], rs, cs, 1};
}

// forward pass; the last layer is linear
void forward(bacchus_hip_mlp *m, const float *x, int M, hipStream_t s)
{
	const int sa = (m->precision & BACCHUS_HIP_MLP_SPLIT_ACTIVATIONS) ? 1 : 0;
	for (int l = 0; l < m->nlayers; l++) {
		const int K = m->din[l], N = m->dout[l];
		Operand A = l == 0 ? Operand{x, K, 1, 0} : Operand{m->H[l - 1], K, 1, 0};
		// B(k, n) = W[n, k] (W is N x K row-major)
		int sb;
		Operand B = weights(m, l, 1, K, &sb);
		const bool last = l + 1 == m->nlayers;
		gemm(s, A, B, M, N, K, m->params + m->b_off[l], m->H[l], N, m->Z[l], last ? 0 : 1,
		     sa | sb);
	}
}

// gradients of the mean squared error into m->grads
void backward(bacchus_hip_mlp *m, const float *x, const float *t, int M, hipStream_t s)
{
	const int L = m->nlayers;
	const int64_t ny = (int64_t)M * m->out;
	hipLaunchKernelGGL(mse_grad, dim3(1), dim3(256), 0, s, m->H[L - 1], t, ny, m->dY, m->loss);
	const float *dz = m->dY;
	// with split activations every backward operand but the weights is
	// carried as two BF16 parts too
	const int sa = (m->precision & BACCHUS_HIP_MLP_SPLIT_ACTIVATIONS) ? 1 : 0;
	for (int l = L - 1; l >= 0; l--) {
		const int K = m->din[l], N = m->dout[l];
		const float *a = l == 0 ? x : m->H[l - 1];
		// dW[n, k] = sum_m dZ[m, n] a[m, k]
		gemm(s, Operand{dz, 1, N, 0}, Operand{a, K, 1, 0}, N, K, M, nullptr,
		     m->grads + m->w_off[l], K, nullptr, 0, sa ? 3 : 0);
		hipLaunchKernelGGL(column_sum, dim3(grid(N)), dim3(256), 0, s, dz, M, N,
				   m->grads + m->b_off[l]);
		if (l == 0)
			break;
		// dA[m, k] = sum_n dZ[m, n] W[n, k]
		int sb;
		Operand W = weights(m, l, K, 1, &sb);
		gemm(s, Operand{dz, N, 1, 0}, W, M, K, N, nullptr, m->dA, K, nullptr, 0, sa | sb);
		hipLaunchKernelGGL(gelu_backward, dim3(grid((int64_t)M * K)), dim3(256), 0, s, m->dA,
				   m->Z[l - 1], m->dZ, (int64_t)M * K);
		dz = m->dZ;
		// dZ is reused as the next dz; dA is free again for the next layer
	}
}

}  // namespace

extern "C" {

int bacchus_hip_mlp_create(int in, int hidden, int depth, int out, uint64_t seed,
			   bacchus_hip_mlp_t **mlp)
{
	if (!mlp || in < 1 || hidden < 1 || depth < 1 || out < 1)
		return BACCHUS_FAIL("invalid MLP shape");
	auto *m = new bacchus_hip_mlp;
	m->in = in;
	m->hidden = hidden;
	m->depth = depth;
	m->out = out;
	m->nlayers = depth + 1;
	for (int l = 0; l < m->nlayers; l++) {
		m->din.push_back(l == 0 ? in : hidden);
		m->dout.push_back(l == depth ? out : hidden);
	}
	for (int l = 0; l < m->nlayers; l++) {
		m->w_off.push_back(m->nparams);
		m->nparams += (int64_t)m->din[l] * m->dout[l];
	}
	for (int l = 0; l < m->nlayers; l++) {
		m->b_off.push_back(m->nparams);
		m->nparams += m->dout[l];
	}
	// He-style uniform init from a splitmix64 stream
	std::vector<float> p(m->nparams, 0.0f);
	uint64_t st = seed;
	auto next = [&]() {
		uint64_t z = (st += 0x9e3779b97f4a7c15ULL);
		z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
		z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
		return double((z ^ (z >> 31)) >> 11) * 0x1.0p-53;
	};
	for (int l = 0; l < m->nlayers; l++) {
		const float bound = std::sqrt(6.0f / m->din[l]);
		for (int64_t i = 0; i < (int64_t)m->din[l] * m->dout[l]; i++)
			p[m->w_off[l] + i] = (float)((2 * next() - 1) * bound);
	}
	const size_t bytes = m->nparams * 4;
	if (hipMalloc(&m->params, bytes) != hipSuccess ||
	    hipMalloc(&m->params_bf16, m->nparams * 2) != hipSuccess ||
	    hipMalloc(&m->grads, bytes) != hipSuccess || hipMalloc(&m->adam_m, bytes) != hipSuccess ||
	    hipMalloc(&m->adam_v, bytes) != hipSuccess || hipMalloc(&m->loss, 4) != hipSuccess ||
	    hipMemcpy(m->params, p.data(), bytes, hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemset(m->adam_m, 0, bytes) != hipSuccess ||
	    hipMemset(m->adam_v, 0, bytes) != hipSuccess) {
		bacchus_hip_mlp_free(m);
		return BACCHUS_FAIL("MLP device allocation failed");
	}
	hipLaunchKernelGGL(to_bf16_kernel, dim3(grid(m->nparams)), dim3(256), 0, nullptr, m->params,
			   m->params_bf16, m->nparams);
	*mlp = m;
	return 0;
}

void bacchus_hip_mlp_free(bacchus_hip_mlp_t *m)
{
	if (!m)
		return;
	free_acts(m);
	(void)hipFree(m->params);
	(void)hipFree(m->params_bf16);
	(void)hipFree(m->grads);
	(void)hipFree(m->adam_m);
	(void)hipFree(m->adam_v);
	(void)hipFree(m->loss);
	delete m;
}

int64_t bacchus_hip_mlp_num_params(const bacchus_hip_mlp_t *m) { return m ? m->nparams : -1; }

int bacchus_hip_mlp_set_precision(bacchus_hip_mlp_t *m, int32_t flags)
{
	if (!m)
		return BACCHUS_FAIL("NULL argument");
	if (flags & ~(BACCHUS_HIP_MLP_SPLIT_ACTIVATIONS | BACCHUS_HIP_MLP_FP32_LAST))
		return BACCHUS_FAIL("unknown MLP precision flags %d", (int)flags);
	m->precision = flags;
	return 0;
}

int bacchus_hip_mlp_get_params(const bacchus_hip_mlp_t *m, float *host)
{
	if (!m || !host)
		return BACCHUS_FAIL("NULL argument");
	BACCHUS_HIP_CHECK(hipMemcpy(host, m->params, m->nparams * 4, hipMemcpyDeviceToHost));
	return 0;
}

int bacchus_hip_mlp_set_params(bacchus_hip_mlp_t *m, const float *host)
{
	if (!m || !host)
		return BACCHUS_FAIL("NULL argument");
	BACCHUS_HIP_CHECK(hipMemcpy(m->params, host, m->nparams * 4, hipMemcpyHostToDevice));
	hipLaunchKernelGGL(to_bf16_kernel, dim3(grid(m->nparams)), dim3(256), 0, nullptr, m->params,
			   m->params_bf16, m->nparams);
	BACCHUS_HIP_CHECK(hipDeviceSynchronize());
	return 0;
}

int bacchus_hip_mlp_forward(bacchus_hip_mlp_t *m, const float *x, int M, float *y, void *stream)
{
	if (!m || !x || !y || M < 1)
		return BACCHUS_FAIL("invalid argument");
	if (ensure_capacity(m, M))
		return 1;
	hipStream_t s = (hipStream_t)stream;
	forward(m, x, M, s);
	BACCHUS_HIP_CHECK(hipMemcpyAsync(y, m->H[m->nlayers - 1], (size_t)M * m->out * 4,
					 hipMemcpyDeviceToDevice, s));
	BACCHUS_HIP_CHECK(hipGetLastError());
	return 0;
}

int bacchus_hip_mlp_gradients(bacchus_hip_mlp_t *m, const float *x, const float *t, int M,
			      float *grads_host, float *loss_host, void *stream)
{
	if (!m || !x || !t || M < 1)
		return BACCHUS_FAIL("invalid argument");
	if (ensure_capacity(m, M))
		return 1;
	hipStream_t s = (hipStream_t)stream;
	forward(m, x, M, s);
	backward(m, x, t, M, s);
	if (grads_host)
		BACCHUS_HIP_CHECK(hipMemcpyAsync(grads_host, m->grads, m->nparams * 4,
						 hipMemcpyDeviceToHost, s));
	if (loss_host)
		BACCHUS_HIP_CHECK(hipMemcpyAsync(loss_host, m->loss, 4, hipMemcpyDeviceToHost, s));
	BACCHUS_HIP_CHECK(hipStreamSynchronize(s));
	return 0;
}

int bacchus_hip_mlp_train_step(bacchus_hip_mlp_t *m, const float *x, const float *t, int M,
			       float lr, float *loss_host, void *stream)
{
	if (!m || !x || !t || M < 1)
		return BACCHUS_FAIL("invalid argument");
	if (ensure_capacity(m, M))
		return 1;
	hipStream_t s = (hipStream_t)stream;
	forward(m, x, M, s);
	backward(m, x, t, M, s);
	m->step++;
	const float b1 = 0.9f, b2 = 0.999f;
	const float c1 = 1.0f - std::pow(b1, (float)m->step);
	const float c2 = 1.0f - std::pow(b2, (float)m->step);
	hipLaunchKernelGGL(adam_step, dim3(grid(m->nparams)), dim3(256), 0, s, m->params,
			   m->params_bf16, m->grads, m->adam_m, m->adam_v, m->nparams, lr, b1, b2,
			   1e-8f, c1, c2);
	if (loss_host) {
		BACCHUS_HIP_CHECK(hipMemcpyAsync(loss_host, m->loss, 4, hipMemcpyDeviceToHost, s));
		BACCHUS_HIP_CHECK(hipStreamSynchronize(s));
	}
	BACCHUS_HIP_CHECK(hipGetLastError());
	return 0;
}

}  // extern "C"
