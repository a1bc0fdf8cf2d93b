// lanczos.hip - the Lanczos recurrence of src/Bacchus/lanczos/lanczos.c:
// 98-291 on the device: the Krylov basis stays in HBM, alpha/beta, the
// reorthogonalization coefficients and the tridiagonal eigenvalues are
// computed in FP64 on the GPU, and the host only reads a few eigenvalues
// back at convergence checks.
#include "bacchus_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

namespace bacchus {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxPartials = 1024;

__device__ __forceinline__ double wave_sum(double v)
{
	for (int o = 32; o > 0; o >>= 1)
		v += __shfl_xor(v, o, 64);
	return v;
}

// block-wide sum; every thread gets the result
__device__ double block_sum(double v, double *sh)
{
	v = wave_sum(v);
	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
	if (lane == 0)
		sh[wid] = v;
	__syncthreads();
	double t = 0;
	if (threadIdx.x < 64) {
		t = threadIdx.x < (blockDim.x >> 6) ? sh[threadIdx.x] : 0.0;
		t = wave_sum(t);
		if (threadIdx.x == 0)
			sh[0] = t;
	}
	__syncthreads();
	t = sh[0];
	__syncthreads();
	return t;
}

// partial[j * nblk + blk] = sum over this block's slice of U_j . w,
// j = 0..nvec-1 (nvec = 1 gives a plain dot product).  Grid (nblk, vector
// groups of kDotVecs): a thread reads w[i] once for the group's vectors and
// keeps their sums in registers; each sum is reduced over the block in a
// fixed order (wave shuffles, then the waves in order).
constexpr int kDotVecs = 8;
__global__ __launch_bounds__(kThreads) void multi_dot_partial(
	const double *__restrict__ U, int64_t ld, int nvec,
	const double *__restrict__ w, int64_t n, double *__restrict__ partial)
{
	__shared__ double sh[kDotVecs][kThreads / 64];
	const int j0 = blockIdx.y * kDotVecs;
	const int nv = min(kDotVecs, nvec - j0);
	const int64_t per = (n + gridDim.x - 1) / gridDim.x;
	const int64_t b = per * blockIdx.x, e = min(n, b + per);
	double s[kDotVecs];
#pragma unroll
	for (int q = 0; q < kDotVecs; q++)
		s[q] = 0.0;
	for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
		const double wi = w[i];
#pragma unroll
		for (int q = 0; q < kDotVecs; q++)
			if (q < nv)
				s[q] += U[ld * (j0 + q) + i] * wi;
	}
	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
	for (int q = 0; q < kDotVecs; q++) {
		const double v = wave_sum(s[q]);
		if (lane == 0)
			sh[q][wid] = v;
	}
	__syncthreads();
	if ((int)threadIdx.x < nv) {
		double t = 0;
		for (int k = 0; k < (int)(blockDim.x >> 6); k++)
			t += sh[threadIdx.x][k];
		partial[(int64_t)(j0 + threadIdx.x) * gridDim.x + blockIdx.x] = t;
	}
}

// out[j] = scale_j * sum_b partial[j * nblk + b] in a fixed order (one block
// per j, deterministic)
__global__ __launch_bounds__(kThreads) void reduce_partials(
	const double *__restrict__ partial, int nblk, double *__restrict__ out)
{
	__shared__ double sh[kThreads / 64];
	const int j = blockIdx.x;
	double s = 0;
	for (int b = threadIdx.x; b < nblk; b += blockDim.x)
		s += partial[(int64_t)j * nblk + b];
	s = block_sum(s, sh);
	if (threadIdx.x == 0)
		out[j] = s;
}

// w -= alpha_k u_k (+ beta_{k-1} u_{k-1}) with alpha, beta read on device
// (lanczos.c:113-135)
__global__ void three_term(double *__restrict__ w, const double *__restrict__ uk,
			   const double *__restrict__ ukm1, const double *alpha_k,
			   const double *beta_km1, int64_t n)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	double v = w[i] - (*alpha_k) * uk[i];
	if (ukm1)
		v -= (*beta_km1) * ukm1[i];
	w[i] = v;
}

// w -= sum_j c_j U_j  (classical Gram-Schmidt update)
__global__ void subtract_combination(double *__restrict__ w,
				     const double *__restrict__ U, int64_t ld,
				     int nvec, const double *__restrict__ c,
				     int64_t n)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	double v = w[i];
#pragma unroll 4
	for (int j = 0; j < nvec; j++)
		v -= c[j] * U[ld * j + i];
	w[i] = v;
}

// w -= c_0 u (single modified Gram-Schmidt step, c read on device)
__global__ void subtract_one(double *__restrict__ w, const double *__restrict__ u,
			     const double *c, int64_t n)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		w[i] -= (*c) * u[i];
}

// beta = sqrt(ss); w *= 1/beta (lanczos.c:137-141)
__global__ void finish_beta(const double *ss, double *beta_k)
{
	*beta_k = sqrt(*ss);
}
__global__ void scale_by_inverse(double *__restrict__ w, const double *beta_k,
				 int64_t n)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		w[i] *= 1.0 / (*beta_k);
}

__global__ void combine_vectors(double *__restrict__ out,
				const double *__restrict__ U, int64_t ld, int nvec,
				const double *__restrict__ z, int64_t n)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	double v = 0;
	for (int j = 0; j < nvec; j++)
		v += z[j] * U[ld * j + i];
	out[i] = v;
}

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
	z += 0x9e3779b97f4a7c15ULL;
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
	return z ^ (z >> 31);
}

// uniform in [-1, 1) keyed by (seed, index), like new_random_vector's
// 2*drand48()-1 (src/Bacchus/vector/vector.c:130-137) but reproducible
__global__ void random_fill(double *u, int64_t n, uint64_t seed)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		u[i] = 2.0 * double(mix64(seed ^ mix64((uint64_t)i)) >> 11) * 0x1.0p-53 - 1.0;
}

__global__ void set_e0(double *u, int64_t n)
{
	const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		u[i] = i == 0 ? 1.0 : 0.0;
}

// Lowest `count` eigenvalues of the k x k tridiagonal T (diag alpha, off
// beta) by Sturm-sequence bisection, one thread per eigenvalue.
__global__ void tridiagonal_lowest(const double *__restrict__ d,
				   const double *__restrict__ e, int k, int count,
				   double *__restrict__ out)
{
	__shared__ double lo_sh[kThreads / 64], hi_sh[kThreads / 64];
	double lo = 1e300, hi = -1e300;
	for (int i = threadIdx.x; i < k; i += blockDim.x) {
		const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < k ? fabs(e[i]) : 0.0);
		lo = fmin(lo, d[i] - r);
		hi = fmax(hi, d[i] + r);
	}
	for (int o = 32; o > 0; o >>= 1) {
		lo = fmin(lo, __shfl_xor(lo, o, 64));
		hi = fmax(hi, __shfl_xor(hi, o, 64));
	}
	if ((threadIdx.x & 63) == 0) {
		lo_sh[threadIdx.x >> 6] = lo;
		hi_sh[threadIdx.x >> 6] = hi;
	}
	__syncthreads();
	lo = lo_sh[0];
	hi = hi_sh[0];
	for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
		lo = fmin(lo, lo_sh[w]);
		hi = fmax(hi, hi_sh[w]);
	}
	const double span = fmax(fabs(lo), fabs(hi));
	lo -= 1e-12 * span + 1e-300;
	hi += 1e-12 * span + 1e-300;
	for (int idx = threadIdx.x; idx < count && idx < k; idx += blockDim.x) {
		double a = lo, b = hi;
		for (int it = 0; it < 200; it++) {
			const double mid = 0.5 * (a + b);
			if (mid <= a || mid >= b)
				break;
			int cnt = 0;
			double q = d[0] - mid;
			if (q < 0)
				cnt++;
			for (int i = 1; i < k; i++) {
				if (q == 0.0)
					q = 1e-300;
				q = d[i] - mid - e[i - 1] * e[i - 1] / q;
				if (q < 0)
					cnt++;
			}
			if (cnt > idx)
				b = mid;
			else
				a = mid;
		}
		out[idx] = 0.5 * (a + b);
	}
}

int blocks_for(int64_t n) { return (int)((n + kThreads - 1) / kThreads); }

struct Lanczos {
	hipStream_t s = nullptr;
	int64_t D = 0;
	int nblk = 0;
	double *partial = nullptr, *scal = nullptr;

	int dot(const double *a, const double *b, double *out)
	{
		hipLaunchKernelGGL(multi_dot_partial, dim3(nblk, 1), dim3(kThreads), 0, s, a, D, 1,
				   b, D, partial);
		hipLaunchKernelGGL(reduce_partials, dim3(1), dim3(kThreads), 0, s, partial,
				   nblk, out);
		return 0;
	}
	int multi_dot(const double *U, int nvec, const double *w, double *out)
	{
		hipLaunchKernelGGL(multi_dot_partial,
				   dim3(nblk, (unsigned)((nvec + kDotVecs - 1) / kDotVecs)), dim3(kThreads),
				   0, s, U, D, nvec, w, D, partial);
		hipLaunchKernelGGL(reduce_partials, dim3(nvec), dim3(kThreads), 0, s, partial,
				   nblk, out);
		return 0;
	}
};

}  // namespace

// eigenvector of the tridiagonal T for eigenvalue lambda by inverse
// iteration, normalized with component 0 >= 0 (diagonalization.c:55-58)
std::vector<double> tridiagonal_vector(const std::vector<double> &d,
				       const std::vector<double> &e, int64_t k,
				       double lambda)
{
	std::vector<double> z(k, 1.0), c(k), dd(k), rhs(k);
	for (int sweep = 0; sweep < 4; sweep++) {
		// solve (T - lambda I) z_new = z with Gaussian elimination
		// (Thomas algorithm, tiny pivots nudged)
		double scale = 0;
		for (int64_t i = 0; i < k; i++)
			scale = std::max(scale, std::fabs(d[i]) + (i ? std::fabs(e[i - 1]) : 0));
		const double tiny = 1e-14 * (scale > 0 ? scale : 1.0);
		for (int64_t i = 0; i < k; i++) {
			dd[i] = d[i] - lambda;
			rhs[i] = z[i];
		}
		for (int64_t i = 0; i < k; i++) {
			if (i > 0) {
				const double m = e[i - 1] / dd[i - 1];
				dd[i] -= m * c[i - 1];
				rhs[i] -= m * rhs[i - 1];
			}
			if (std::fabs(dd[i]) < tiny)
				dd[i] = dd[i] < 0 ? -tiny : tiny;
			c[i] = i + 1 < k ? e[i] : 0.0;
		}
		for (int64_t i = k - 1; i >= 0; i--)
			z[i] = (rhs[i] - (i + 1 < k ? c[i] * z[i + 1] : 0.0)) / dd[i];
		double nrm = 0;
		for (double v : z)
			nrm += v * v;
		nrm = std::sqrt(nrm);
		for (double &v : z)
			v /= nrm;
	}
	if (z[0] < 0)
		for (double &v : z)
			v = -v;
	return z;
}


std::vector<double> tridiagonal_lowest_host(const double *d, const double *e,
					    int64_t k, int count)
{
	double lo = 1e300, hi = -1e300;
	for (int64_t i = 0; i < k; i++) {
		const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < k ? std::fabs(e[i]) : 0.0);
		lo = std::min(lo, d[i] - r);
		hi = std::max(hi, d[i] + r);
	}
	const double span = std::max(std::fabs(lo), std::fabs(hi));
	lo -= 1e-12 * span + 1e-300;
	hi += 1e-12 * span + 1e-300;
	std::vector<double> out;
	for (int idx = 0; idx < count && idx < k; idx++) {
		double a = lo, b = hi;
		for (int it = 0; it < 200; it++) {
			const double mid = 0.5 * (a + b);
			if (mid <= a || mid >= b)
				break;
			int64_t cnt = 0;
			double q = d[0] - mid;
			if (q < 0)
				cnt++;
			for (int64_t i = 1; i < k; i++) {
				if (q == 0.0)
					q = 1e-300;
				q = d[i] - mid - e[i - 1] * e[i - 1] / q;
				if (q < 0)
					cnt++;
			}
			if (cnt > idx)
				b = mid;
			else
				a = mid;
		}
		out.push_back(0.5 * (a + b));
	}
	return out;
}

}  // namespace bacchus

using namespace bacchus;

extern "C" int bacchus_hip_lanczos(bacchus_hip_operator_t *op,
				   const bacchus_hip_lanczos_settings_t *st,
				   double *eigenvalues, double *alpha_out,
				   double *beta_out, double *ritz_vector,
				   bacchus_hip_lanczos_result_t *result)
{
	if (!op || !st || !eigenvalues)
		return BACCHUS_FAIL("NULL argument");
	const auto t0 = std::chrono::steady_clock::now();
	const int64_t D = op->dimension;
	const int64_t K = std::min<int64_t>(D, st->max_num_iterations);
	const int nritz = std::max(1, st->num_ritz_vectors);
	const int target = st->target_eigenvalue;
	// eigenvalues computed per check: the reported ones, the Ritz vectors'
	// and the target's
	const int m = std::max({1, st->num_eigenvalues, nritz, target + 1});
	if (K < 1)
		return BACCHUS_FAIL("max_num_iterations must be positive");
	if (target < 0)
		return BACCHUS_FAIL("target_eigenvalue %d is negative", target);
	if (st->convergence != BACCHUS_HIP_CONVERGE_EIGENVALUES &&
	    st->convergence != BACCHUS_HIP_CONVERGE_EIGENVECTORS &&
	    st->convergence != BACCHUS_HIP_CONVERGE_NONE &&
	    st->convergence != BACCHUS_HIP_CONVERGE_TARGETS)
		return BACCHUS_FAIL("convergence %d not supported by the Lanczos solver",
				    st->convergence);
	if (st->start == BACCHUS_HIP_START_GIVEN && !st->start_vector)
		return BACCHUS_FAIL("start vector missing");
	if (st->convergence == BACCHUS_HIP_CONVERGE_TARGETS && !st->target_eigenvalues)
		return BACCHUS_FAIL("target eigenvalues missing");
	Lanczos L;
	L.D = D;
	L.nblk = (int)std::min<int64_t>(kMaxPartials, std::max<int64_t>(1, D / (4 * kThreads)));
	double *U = nullptr, *dev = nullptr;
	// device scalars: alpha[K], beta[K], coeff[K+1], eig[m], ss[1]
	const int64_t nscal = 2 * K + (K + 1) + m + 8;
	BACCHUS_HIP_CHECK(hipMalloc(&U, (size_t)(K + 1) * D * sizeof(double)));
	if (hipMalloc(&dev, nscal * sizeof(double)) != hipSuccess ||
	    hipMalloc(&L.partial, (size_t)(K + 1) * L.nblk * sizeof(double)) != hipSuccess) {
		(void)hipFree(U);
		(void)hipFree(dev);
		return BACCHUS_FAIL("device allocation failed");
	}
	double *d_alpha = dev, *d_beta = dev + K, *d_coef = dev + 2 * K;
	double *d_eig = d_coef + (K + 1), *d_ss = d_eig + m;
	// one event pair per matvec, read once at the end so the host never
	// waits on a matvec between convergence checks
	std::vector<hipEvent_t> ev(2 * K, nullptr);
	for (auto &e : ev)
		(void)hipEventCreate(&e);
	int rc = 0;
	const int gb = blocks_for(D);
	// u_0
	if (st->start == BACCHUS_HIP_START_GIVEN)
		rc = hipMemcpyAsync(U, st->start_vector, D * sizeof(double),
				    hipMemcpyDeviceToDevice, L.s) != hipSuccess;
	else if (st->start == BACCHUS_HIP_START_RANDOM) {
		hipLaunchKernelGGL(random_fill, dim3(gb), dim3(kThreads), 0, L.s, U, D, st->seed);
		L.dot(U, U, d_ss);
		hipLaunchKernelGGL(finish_beta, dim3(1), dim3(1), 0, L.s, d_ss, d_ss + 1);
		hipLaunchKernelGGL(scale_by_inverse, dim3(gb), dim3(kThreads), 0, L.s, U, d_ss + 1, D);
	} else
		hipLaunchKernelGGL(set_e0, dim3(gb), dim3(kThreads), 0, L.s, U, D);
	double matvec_ms = 0;
	int64_t it = 0;
	double previous = 0;
	// EIGENVECTORS: the target eigenvector's amplitudes of the previous
	// check, zero beyond its length (lanczos.c:183-184, 238-240)
	std::vector<double> previous_amplitudes(K + 1, 0.0);
	std::vector<double> h_eig(m), h_alpha(K), h_beta(K);
	const int check_every = std::max(1, st->check_every);
	bool converged = false;
	for (it = 0; it < K && !rc; it++) {
		double *u = U + it * D, *w = U + (it + 1) * D;
		(void)hipEventRecord(ev[2 * it], L.s);
		rc = bacchus_hip_operator_matvec(op, u, w, 0, L.s);
		(void)hipEventRecord(ev[2 * it + 1], L.s);
		if (rc)
			break;
		L.dot(u, w, d_alpha + it);
		hipLaunchKernelGGL(three_term, dim3(gb), dim3(kThreads), 0, L.s, w, u,
				   it > 0 ? U + (it - 1) * D : nullptr, d_alpha + it,
				   it > 0 ? d_beta + it - 1 : nullptr, D);
		if (st->reorthogonalization == BACCHUS_HIP_REORTH_FULL) {
			for (int pass = 0; pass < 2; pass++) {
				L.multi_dot(U, (int)(it + 1), w, d_coef);
				hipLaunchKernelGGL(subtract_combination, dim3(gb), dim3(kThreads), 0,
						   L.s, w, U, D, (int)(it + 1), d_coef, D);
			}
		}
		L.dot(w, w, d_ss);
		hipLaunchKernelGGL(finish_beta, dim3(1), dim3(1), 0, L.s, d_ss, d_beta + it);
		hipLaunchKernelGGL(scale_by_inverse, dim3(gb), dim3(kThreads), 0, L.s, w,
				   d_beta + it, D);
		if (st->reorthogonalization == BACCHUS_HIP_REORTH_REFERENCE && it > 2) {
			// sequential MGS against u_0..u_{it-3} (vector.c:461-476)
			for (int64_t j = 0; j < it - 2; j++) {
				L.dot(w, U + j * D, d_coef);
				hipLaunchKernelGGL(subtract_one, dim3(gb), dim3(kThreads), 0, L.s, w,
						   U + j * D, d_coef, D);
			}
		}
		const bool last = it + 1 == K;
		const bool check =
			st->convergence != BACCHUS_HIP_CONVERGE_NONE && ((it + 1) % check_every == 0);
		if (check && !last) {
			const int cnt = std::min<int64_t>(m, it + 1);
			hipLaunchKernelGGL(tridiagonal_lowest, dim3(1), dim3(kThreads), 0, L.s,
					   d_alpha, d_beta, (int)(it + 1), cnt, d_eig);
			const bool amplitudes = st->convergence == BACCHUS_HIP_CONVERGE_EIGENVECTORS;
			if (hipMemcpyAsync(h_eig.data(), d_eig, cnt * sizeof(double),
					   hipMemcpyDeviceToHost, L.s) != hipSuccess ||
			    (amplitudes &&
			     (hipMemcpyAsync(h_alpha.data(), d_alpha, (it + 1) * sizeof(double),
					     hipMemcpyDeviceToHost, L.s) != hipSuccess ||
			      hipMemcpyAsync(h_beta.data(), d_beta, (it + 1) * sizeof(double),
					     hipMemcpyDeviceToHost, L.s) != hipSuccess)) ||
			    hipStreamSynchronize(L.s) != hipSuccess) {
				rc = BACCHUS_FAIL("eigenvalue readback failed");
				break;
			}
			bool done = false;
			// the reference watches eigenvalue index target_eigenvalue of
			// T; while T is smaller, the highest it has (get_eigenvalue
			// asserts the index, eigensystem.c:60-66)
			const int tk = std::min<int>(target, cnt - 1);
			if (st->convergence == BACCHUS_HIP_CONVERGE_EIGENVALUES) {
				done = std::fabs(h_eig[tk] - previous) < st->eigenvalue_tolerance;
				previous = h_eig[tk];
			} else if (amplitudes) {
				// difference_eigenvectors (lanczos.c:276-287): the first
				// n - 1 amplitudes against the previous ones, the new last
				// one against zero
				const int64_t n = it + 1;
				const std::vector<double> a =
					tridiagonal_vector(h_alpha, h_beta, n, h_eig[tk]);
				double d2 = 0;
				for (int64_t i = 0; i + 1 < n; i++)
					d2 += (a[i] - previous_amplitudes[i]) * (a[i] - previous_amplitudes[i]);
				d2 += a[n - 1] * a[n - 1];
				done = std::sqrt(d2) < st->eigenvalue_tolerance;
				std::copy(a.begin(), a.end(), previous_amplitudes.begin());
			} else if (cnt == m) {
				done = true;
				for (int k = 0; k < m; k++)
					if (std::fabs(h_eig[k] - st->target_eigenvalues[k]) >
					    st->eigenvalue_tolerance * std::fabs(st->target_eigenvalues[k]))
						done = false;
			}
			if (done) {
				converged = true;
				it++;
				break;
			}
		}
	}
	const int64_t dim = it;
	if (!rc) {
		const int cnt = (int)std::min<int64_t>(m, dim);
		hipLaunchKernelGGL(tridiagonal_lowest, dim3(1), dim3(kThreads), 0, L.s, d_alpha,
				   d_beta, (int)dim, cnt, d_eig);
		if (hipMemcpyAsync(h_eig.data(), d_eig, cnt * sizeof(double), hipMemcpyDeviceToHost,
				   L.s) != hipSuccess ||
		    hipMemcpyAsync(h_alpha.data(), d_alpha, dim * sizeof(double),
				   hipMemcpyDeviceToHost, L.s) != hipSuccess ||
		    hipMemcpyAsync(h_beta.data(), d_beta, dim * sizeof(double),
				   hipMemcpyDeviceToHost, L.s) != hipSuccess ||
		    hipStreamSynchronize(L.s) != hipSuccess)
			rc = BACCHUS_FAIL("result readback failed");
		// the caller's array holds num_eigenvalues values (m may be larger
		// for the Ritz vectors or the target)
		for (int k = 0; k < cnt && k < std::max(1, st->num_eigenvalues) && !rc; k++)
			eigenvalues[k] = h_eig[k];
		if (!rc && alpha_out)
			memcpy(alpha_out, h_alpha.data(), dim * sizeof(double));
		if (!rc && beta_out)
			memcpy(beta_out, h_beta.data(), dim * sizeof(double));
		// Ritz vector r = sum_j z_j u_j, z the eigenvector of T for the
		// r-th lowest eigenvalue (the amplitudes get_eigenvector combines,
		// eigensystem.c:85-95)
		for (int r = 0; r < nritz && !rc && ritz_vector; r++) {
			if (r >= cnt) {
				rc = BACCHUS_FAIL("%d Ritz vectors asked for, the Krylov space has "
						  "dimension %lld", nritz, (long long)dim);
				break;
			}
			std::vector<double> z = tridiagonal_vector(h_alpha, h_beta, dim, h_eig[r]);
			if (hipMemcpy(d_coef, z.data(), dim * sizeof(double), hipMemcpyHostToDevice) !=
			    hipSuccess)
				rc = BACCHUS_FAIL("ritz coefficients upload failed");
			else {
				hipLaunchKernelGGL(combine_vectors, dim3(gb), dim3(kThreads), 0, L.s,
						   ritz_vector + (size_t)r * D, U, D, (int)dim, d_coef, D);
				if (hipStreamSynchronize(L.s) != hipSuccess)
					rc = BACCHUS_FAIL("ritz vector failed");
			}
		}
	}
	if (!rc && st->krylov_basis && dim > 0) {
		if (hipMemcpyAsync(st->krylov_basis, U, (size_t)dim * D * sizeof(double),
				   hipMemcpyDeviceToHost, L.s) != hipSuccess ||
		    hipStreamSynchronize(L.s) != hipSuccess)
			rc = BACCHUS_FAIL("Krylov basis copy failed");
	}
	for (int64_t k = 0; k < dim && !rc; k++) {
		float ms = 0;
		if (hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]) == hipSuccess)
			matvec_ms += ms;
	}
	for (auto &e : ev)
		(void)hipEventDestroy(e);
	(void)hipFree(U);
	(void)hipFree(dev);
	(void)hipFree(L.partial);
	if (result) {
		result->iterations = dim;
		result->converged = converged ? 1 : 0;
		result->pad = 0;
		result->matvec_seconds = matvec_ms * 1e-3;
		result->total_seconds =
			std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
	}
	return rc;
}
