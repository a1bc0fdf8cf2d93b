// subspace.hip - Ceres's projections in FP64 on the device: the norm
// matrix of stored vectors (src/Ceres/norm_matrix/norm_matrix.c:4-28) and
// the operator projected onto them (src/Ceres/subspace_operator/
// subspace_operator.c:7-61).
//
// The reference multiplies every stored vector by H with one Bacchus matvec
// (disk-backed vectors) and forms the n(n+1)/2 dot products on the host.
// Here the vectors are device-resident [n][D] FP64 rows, each product is
// one FP64 single-vector matvec, and the n dot products of a product with
// all rows run as one two-stage deterministic reduction (fixed chunks of D
// per workgroup, then the chunks in order), so results do not depend on
// scheduling.  The matrices come back to the host row-major, the upper
// triangle mirrored as the reference fills it.
#include "bacchus_internal.h"

#include <vector>

namespace bacchus {
namespace {

constexpr int kDotChunks = 128;   // workgroups per row
constexpr int kDotThreads = 256;

// partial[j][c] = sum over chunk c of V[j][k] y[k]
__global__ __launch_bounds__(kDotThreads) void rows_dot_partial(const double *__restrict__ V,
								int64_t D, const double *__restrict__ y,
								double *__restrict__ partial)
{
	const int64_t j = blockIdx.y;
	const int c = blockIdx.x;
	const int64_t per = (D + kDotChunks - 1) / kDotChunks;
	const int64_t k0 = (int64_t)c * per, k1 = k0 + per < D ? k0 + per : D;
	const double *row = V + j * D;
	double s = 0.0;
	for (int64_t k = k0 + threadIdx.x; k < k1; k += kDotThreads)
		s += row[k] * y[k];
	__shared__ double red[kDotThreads];
	red[threadIdx.x] = s;
	__syncthreads();
	for (int w = kDotThreads / 2; w > 0; w >>= 1) {
		if ((int)threadIdx.x < w)
			red[threadIdx.x] += red[threadIdx.x + w];
		__syncthreads();
	}
	if (threadIdx.x == 0)
		partial[j * kDotChunks + c] = red[0];
}

// out[j] = sum_c partial[j][c], chunks in order
__global__ void rows_dot_sum(const double *__restrict__ partial, int64_t n, double *__restrict__ out)
{
	const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= n)
		return;
	double s = 0.0;
	for (int c = 0; c < kDotChunks; c++)
		s += partial[j * kDotChunks + c];
	out[j] = s;
}

// out[j] = V[j] . y for the n rows of V, on stream s
void rows_dot(const double *V, int64_t n, int64_t D, const double *y, double *partial, double *out,
	      hipStream_t s)
{
	hipLaunchKernelGGL(rows_dot_partial, dim3(kDotChunks, (unsigned)n), dim3(kDotThreads), 0, s, V,
			   D, y, partial);
	hipLaunchKernelGGL(rows_dot_sum, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, partial, n,
			   out);
}

// the n x n matrix of dots[i][j] (device, row i from vector i) to the host,
// dots[i][j] for j >= i mirrored into [j][i]
int symmetric_to_host(const double *dots, int64_t n, double *M, hipStream_t s)
{
	std::vector<double> h((size_t)n * n);
	BACCHUS_HIP_CHECK(hipMemcpyAsync(h.data(), dots, h.size() * sizeof(double),
					 hipMemcpyDeviceToHost, s));
	BACCHUS_HIP_CHECK(hipStreamSynchronize(s));
	for (int64_t i = 0; i < n; i++)
		for (int64_t j = i; j < n; j++)
			M[i * n + j] = M[j * n + i] = h[(size_t)(i * n + j)];
	return 0;
}

struct DeviceBuffers {
	double *p = nullptr;
	~DeviceBuffers() { (void)hipFree(p); }
};

}  // namespace
}  // namespace bacchus

using namespace bacchus;

extern "C" {

int bacchus_hip_norm_matrix(const double *V, int64_t n, int64_t D, double *N, void *stream)
{
	if (!V || !N || n <= 0 || D <= 0 || n > 65535)
		return BACCHUS_FAIL("bad arguments (n = %lld, D = %lld)", (long long)n, (long long)D);
	hipStream_t s = (hipStream_t)stream;
	DeviceBuffers buf;
	BACCHUS_HIP_CHECK(hipMalloc(&buf.p, (size_t)n * (kDotChunks + n) * sizeof(double)));
	double *partial = buf.p, *dots = buf.p + (size_t)n * kDotChunks;
	for (int64_t i = 0; i < n; i++)
		rows_dot(V, n, D, V + i * D, partial, dots + i * n, s);
	BACCHUS_HIP_CHECK(hipGetLastError());
	return symmetric_to_host(dots, n, N, s);
}

int bacchus_hip_subspace_operator(bacchus_hip_operator_t *op, const double *V, int64_t n, double *M,
				  void *stream)
{
	if (!op || !V || !M || n <= 0 || n > 65535)
		return BACCHUS_FAIL("bad arguments (n = %lld)", (long long)n);
	const int64_t D = bacchus_hip_operator_dimension(op);
	if (D <= 0)
		return BACCHUS_FAIL("operator has dimension %lld", (long long)D);
	hipStream_t s = (hipStream_t)stream;
	DeviceBuffers buf;
	BACCHUS_HIP_CHECK(hipMalloc(&buf.p, ((size_t)D + (size_t)n * (kDotChunks + n)) * sizeof(double)));
	double *y = buf.p, *partial = buf.p + D, *dots = partial + (size_t)n * kDotChunks;
	for (int64_t i = 0; i < n; i++) {
		// H v_i, then v_j . (H v_i) for every j (the reference keeps j >= i)
		if (int rc = bacchus_hip_operator_matvec(op, V + i * D, y, 0, s))
			return rc;
		rows_dot(V, n, D, y, partial, dots + i * n, s);
	}
	BACCHUS_HIP_CHECK(hipGetLastError());
	return symmetric_to_host(dots, n, M, s);
}

}  // extern "C"
