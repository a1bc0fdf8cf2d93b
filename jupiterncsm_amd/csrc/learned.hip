// learned.hip - the stored-basis side of the learned Lanczos start: Gram
// matrix and projections that turn solved eigenvectors into training
// targets, and the combination that turns the network's coefficients into
// a start vector on the device.  FP32 storage, FP64 accumulation.
#include "bacchus_internal.h"

#include <mutex>
#include <vector>

namespace bacchus {
namespace {

constexpr int kChunks = 64;
constexpr int kMaxDevices = 64;

__device__ double block_reduce(double v, double *sh)
{
	for (int o = 32; o > 0; o >>= 1)
		v += __shfl_xor(v, o, 64);
	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
	if (lane == 0)
		sh[wid] = v;
	__syncthreads();
	double t = 0;
	if (threadIdx.x == 0)
		for (int w = 0; w < (int)(blockDim.x >> 6); w++)
			t += sh[w];
	__syncthreads();
	return t;
}

// partial[(i * n + j) * chunks + c] = U_i . U_j over chunk c (j <= i)
__global__ __launch_bounds__(256) void gram_partial(const float *__restrict__ U, int n, int64_t D,
						   double *__restrict__ partial)
{
	__shared__ double sh[4];
	const int i = blockIdx.y / n, j = blockIdx.y % n;
	if (j > i)
		return;
	const int64_t per = (D + gridDim.x - 1) / gridDim.x;
	const int64_t r0 = per * blockIdx.x, r1 = min(D, r0 + per);
	const float *a = U + (int64_t)i * D, *b = U + (int64_t)j * D;
	double s = 0;
	for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x)
		s += (double)a[r] * (double)b[r];
	s = block_reduce(s, sh);
	if (threadIdx.x == 0)
		partial[((int64_t)i * n + j) * gridDim.x + blockIdx.x] = s;
}

// partial[(i * chunks + c) * B + b] = U_i . V[:, b] over chunk c
__global__ __launch_bounds__(256) void project_partial(const float *__restrict__ U, int64_t D,
						      const double *__restrict__ V, int B,
						      double *__restrict__ partial)
{
	__shared__ double sh[256];
	const int i = blockIdx.y;
	const int R = 256 / B;
	const int b = threadIdx.x % B, rl = threadIdx.x / B;
	const int64_t per = (D + gridDim.x - 1) / gridDim.x;
	const int64_t r0 = per * blockIdx.x, r1 = min(D, r0 + per);
	const float *u = U + (int64_t)i * D;
	double s = 0;
	if (rl < R)
		for (int64_t r = r0 + rl; r < r1; r += R)
			s += (double)u[r] * V[r * B + b];
	sh[threadIdx.x] = s;
	__syncthreads();
	if ((int)threadIdx.x < B) {
		double t = 0;
		for (int q = 0; q < R; q++)
			t += sh[threadIdx.x + q * B];
		partial[((int64_t)i * gridDim.x + blockIdx.x) * B + threadIdx.x] = t;
	}
}

// out[i * B + b] = sum_c partial[(i * chunks + c) * B + b], the chunks in
// order (one thread per output)
__global__ void reduce_project(const double *__restrict__ partial, int n, int B, int chunks,
			       double *__restrict__ out)
{
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= n * B)
		return;
	const int i = t / B, b = t - i * B;
	double s = 0;
	for (int c = 0; c < chunks; c++)
		s += partial[((int64_t)i * chunks + c) * B + b];
	out[t] = s;
}

// Scratch of bacchus_hip_basis_project_static: the partial sums and the
// reduced projections, kProjectScratch doubles in every GPU's memory (held
// until the process ends).  Host code uses the address hipGetSymbolAddress
// returns for the current device, never this name.
constexpr int64_t kProjectScratch = 1 << 18;
static __device__ __align__(256) double scratch[kProjectScratch];

// projections of n basis vectors onto B device vectors into partial
// (n * kChunks * B doubles) and their sums into finals (n * B doubles)
void project_into(const float *U, int64_t n, int64_t D, const double *V, int64_t B,
		  double *partial, double *finals)
{
	hipLaunchKernelGGL(project_partial, dim3(kChunks, (unsigned)n), dim3(256), 0, nullptr, U, D, V,
			   (int)B, partial);
	hipLaunchKernelGGL(reduce_project, dim3((unsigned)((n * B + 255) / 256)), dim3(256), 0,
			   nullptr, partial, (int)n, (int)B, kChunks, finals);
}

// out[r] = sum_i c_i U_i[r] (FP64 accumulation)
__global__ void combine_basis(const float *__restrict__ U, int n, int64_t D,
			      const float *__restrict__ c, double *__restrict__ out)
{
	const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= D)
		return;
	double v = 0;
	for (int i = 0; i < n; i++)
		v += (double)c[i] * (double)U[(int64_t)i * D + r];
	out[r] = v;
}

// out[j][r] = sum_i C[j][i] U[i][r] (FP64 accumulation), grid.y = j
__global__ void transform_basis(const float *__restrict__ U, int n, int64_t D,
				const double *__restrict__ C, float *__restrict__ out)
{
	const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= D)
		return;
	const double *c = C + (int64_t)blockIdx.y * n;
	double v = 0;
	for (int i = 0; i < n; i++)
		v += c[i] * (double)U[(int64_t)i * D + r];
	out[(int64_t)blockIdx.y * D + r] = (float)v;
}

// the same for at most kFewOutputs outputs, one thread per component
// reading U once for all of them (the same sums, in the same order)
constexpr int kFewOutputs = 16;
__global__ __launch_bounds__(256) void transform_basis_few(const float *__restrict__ U, int n,
							   int64_t D, const double *__restrict__ C,
							   int m, float *__restrict__ out)
{
	const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= D)
		return;
	double acc[kFewOutputs];
#pragma unroll
	for (int q = 0; q < kFewOutputs; q++)
		acc[q] = 0.0;
#pragma unroll 8
	for (int i = 0; i < n; i++) {
		const double u = (double)U[(int64_t)i * D + r];
#pragma unroll
		for (int q = 0; q < kFewOutputs; q++)
			if (q < m)
				acc[q] += C[(int64_t)q * n + i] * u;
	}
#pragma unroll
	for (int q = 0; q < kFewOutputs; q++)
		if (q < m)
			out[(int64_t)q * D + r] = (float)acc[q];
}

// one block: ss = |out|^2 in a fixed order, then out /= sqrt(ss)
__global__ __launch_bounds__(256) void normalize_vector(double *out, int64_t D)
{
	__shared__ double sh[4];
	__shared__ double total;
	double s = 0;
	for (int64_t r = threadIdx.x; r < D; r += blockDim.x)
		s += out[r] * out[r];
	s = block_reduce(s, sh);
	if (threadIdx.x == 0)
		total = s;
	__syncthreads();
	const double inv = total > 0 ? 1.0 / sqrt(total) : 0.0;
	for (int64_t r = threadIdx.x; r < D; r += blockDim.x)
		out[r] *= inv;
}

}  // namespace
}  // namespace bacchus

using namespace bacchus;

extern "C" {

int bacchus_hip_basis_gram(const float *U, int64_t n, int64_t D, double *G)
{
	if (!U || !G || n <= 0 || D <= 0 || n > 255)
		return BACCHUS_FAIL("bad arguments (n = %lld, D = %lld)", (long long)n, (long long)D);
	double *dp = nullptr;
	const size_t np = (size_t)n * n * kChunks;
	BACCHUS_HIP_CHECK(hipMalloc(&dp, np * sizeof(double)));
	hipLaunchKernelGGL(gram_partial, dim3(kChunks, (unsigned)(n * n)), dim3(256), 0, nullptr, U,
			   (int)n, D, dp);
	std::vector<double> h(np);
	const hipError_t e = hipMemcpy(h.data(), dp, np * sizeof(double), hipMemcpyDeviceToHost);
	(void)hipFree(dp);
	if (e != hipSuccess)
		return BACCHUS_FAIL("gram: %s", hipGetErrorString(e));
	for (int64_t i = 0; i < n; i++)
		for (int64_t j = 0; j <= i; j++) {
			double s = 0;
			for (int c = 0; c < kChunks; c++)
				s += h[(i * n + j) * kChunks + c];
			G[i * n + j] = G[j * n + i] = s;
		}
	return 0;
}

int bacchus_hip_basis_project(const float *U, int64_t n, int64_t D, const double *V, int64_t B,
			      double *out)
{
	if (!U || !V || !out || n <= 0 || D <= 0 || B <= 0 || B > 256)
		return BACCHUS_FAIL("bad arguments (n = %lld, D = %lld, B = %lld)", (long long)n,
				    (long long)D, (long long)B);
	double *dp = nullptr;
	const size_t np = (size_t)n * kChunks * B;
	BACCHUS_HIP_CHECK(hipMalloc(&dp, (np + (size_t)n * B) * sizeof(double)));
	project_into(U, n, D, V, B, dp, dp + np);
	const hipError_t e = hipMemcpy(out, dp + np, (size_t)n * B * sizeof(double),
				       hipMemcpyDeviceToHost);
	(void)hipFree(dp);
	if (e != hipSuccess)
		return BACCHUS_FAIL("project: %s", hipGetErrorString(e));
	return 0;
}

int bacchus_hip_basis_project_static(const float *U, int64_t n, int64_t D, const double *V,
				     int64_t B, double *out)
{
	if (!U || !V || !out || n <= 0 || D <= 0 || B <= 0 || B > 256)
		return BACCHUS_FAIL("bad arguments (n = %lld, D = %lld, B = %lld)", (long long)n,
				    (long long)D, (long long)B);
	const int64_t need = n * B * (kChunks + 1);
	if (need > kProjectScratch)
		return BACCHUS_FAIL("n B = %lld needs %lld scratch doubles, more than the %lld of "
				    "the static scratch (n B <= %lld)", (long long)(n * B),
				    (long long)need, (long long)kProjectScratch,
				    (long long)(kProjectScratch / (kChunks + 1)));
	// the scratch array's address on the current device, looked up once per
	// device (each GPU holds its own copy)
	static std::mutex lock;
	static double *address[kMaxDevices] = {};
	int dev = 0;
	BACCHUS_HIP_CHECK(hipGetDevice(&dev));
	if (dev < 0 || dev >= kMaxDevices)
		return BACCHUS_FAIL("device %d beyond the %d this entry point serves", dev, kMaxDevices);
	double *base = nullptr;
	{
		std::lock_guard<std::mutex> g(lock);
		if (!address[dev]) {
			void *p = nullptr;
			BACCHUS_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
			address[dev] = static_cast<double *>(p);
		}
		base = address[dev];
	}
	const int64_t np = n * kChunks * B;
	project_into(U, n, D, V, B, base, base + np);
	BACCHUS_HIP_CHECK(hipMemcpy(out, base + np, (size_t)(n * B) * sizeof(double),
				    hipMemcpyDeviceToHost));
	return 0;
}

int bacchus_hip_basis_combine(const float *U, int64_t n, int64_t D, const float *c, double *out,
			      void *stream)
{
	if (!U || !c || !out || n <= 0 || D <= 0)
		return BACCHUS_FAIL("bad arguments");
	hipStream_t s = (hipStream_t)stream;
	hipLaunchKernelGGL(combine_basis, dim3((unsigned)((D + 255) / 256)), dim3(256), 0, s, U, (int)n,
			   D, c, out);
	hipLaunchKernelGGL(normalize_vector, dim3(1), dim3(256), 0, s, out, D);
	BACCHUS_HIP_CHECK(hipGetLastError());
	return 0;
}

int bacchus_hip_basis_transform(const float *U, int64_t n, int64_t D, const double *C, int64_t r,
				float *out)
{
	if (!U || !C || !out || n <= 0 || D <= 0 || r <= 0 || r > 65535)
		return BACCHUS_FAIL("bad arguments");
	// the coefficients go to a device buffer kept per device (grown when a
	// call needs more); calls hold the lock until their kernel has ended
	static std::mutex mu;
	static double *buf[kMaxDevices] = {};
	static size_t cap[kMaxDevices] = {};
	std::lock_guard<std::mutex> lock(mu);
	int dev = 0;
	BACCHUS_HIP_CHECK(hipGetDevice(&dev));
	if (dev < 0 || dev >= kMaxDevices)
		return BACCHUS_FAIL("device %d beyond %d", dev, kMaxDevices);
	const size_t need = (size_t)r * n * sizeof(double);
	if (cap[dev] < need) {
		if (buf[dev])
			BACCHUS_HIP_CHECK(hipFree(buf[dev]));
		buf[dev] = nullptr;
		cap[dev] = 0;
		BACCHUS_HIP_CHECK(hipMalloc(&buf[dev], need));
		cap[dev] = need;
	}
	double *dc = buf[dev];
	BACCHUS_HIP_CHECK(hipMemcpy(dc, C, need, hipMemcpyHostToDevice));
	if (r <= kFewOutputs)
		hipLaunchKernelGGL(transform_basis_few, dim3((unsigned)((D + 255) / 256)), dim3(256), 0,
				   nullptr, U, (int)n, D, dc, (int)r, out);
	else
		hipLaunchKernelGGL(transform_basis, dim3((unsigned)((D + 255) / 256), (unsigned)r), dim3(256),
				   0, nullptr, U, (int)n, D, dc, out);
	const hipError_t e = hipDeviceSynchronize();
	if (e != hipSuccess)
		return BACCHUS_FAIL("transform: %s", hipGetErrorString(e));
	return 0;
}

}  // extern "C"
