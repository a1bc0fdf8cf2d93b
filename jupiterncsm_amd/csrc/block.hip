// block.hip - block Lanczos over the batched matvec, started from a given
// block of vectors (no counterpart in the reference, whose Lanczos is the
// single-vector recurrence of src/Bacchus/lanczos/lanczos.c:74-253).
//
// The b vectors of a block go through one batched launch per step (FP32 or
// BF16 elements, FP32 vectors).  Krylov blocks are stored in FP32 as
// [D][b] (vector index fastest, the batched layout); every inner product
// accumulates in FP64 and the small dense algebra (block tridiagonal
// matrix, its eigensystem, Gram matrices) runs in FP64 on the host.
// Orthogonalization is classical block Gram-Schmidt against every stored
// block followed by an orthonormalization of the block through the
// eigensystem of its Gram matrix (SVQB), done twice for every block the
// solve goes on with (once for the block of the step that converges, whose
// successor is never used); columns without a direction of their own are
// refilled with seeded random vectors.
#include "bacchus_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>
#include <vector>

namespace bacchus {
namespace {

constexpr int kInnerChunks = 512;
constexpr int kMaxBlock = 16;
constexpr int kMaxKrylov = 256;

__device__ __forceinline__ uint64_t splitmix(uint64_t z)
{
	z += 0x9e3779b97f4a7c15ULL;
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
	return z ^ (z >> 31);
}

// partial[(j * chunks + c) * b * b + p * b + q] = sum over the rows r of
// chunk c of Q_j[r][p] * W[r][q]; Q holds nq blocks [D][b], and a grid row
// j = nq (grid.y = nq + 1) takes W itself in place of Q_j.  Thread
// (g, q) of the block's 256 / b row groups walks rows g, g + 256 / b, ...
// of the chunk with the b sums of column q in registers (its W loads and
// the row's Q values are contiguous across the wave), then the row groups'
// sums are added in LDS in a fixed order.
__global__ __launch_bounds__(256) void block_inner(const float *__restrict__ Q,
						   const float *__restrict__ W, int64_t D, int b, int nq,
						   double *__restrict__ partial)
{
	__shared__ double sh[256 * kMaxBlock];
	const int j = blockIdx.y;
	const int bb = b * b;
	const int G = 256 / b;
	const int t = threadIdx.x;
	const int g = t / b, q = t - g * b;
	const int64_t per = (D + gridDim.x - 1) / gridDim.x;
	const int64_t r0 = per * blockIdx.x, r1 = min(D, r0 + per);
	const float *Qj = j < nq ? Q + (int64_t)j * D * b : W;
	double s[kMaxBlock];
#pragma unroll
	for (int p = 0; p < kMaxBlock; p++)
		s[p] = 0.0;
	if (g < G)
		for (int64_t r = r0 + g; r < r1; r += G) {
			const double w = (double)W[r * b + q];
			const float *row = Qj + r * b;
#pragma unroll
			for (int p = 0; p < kMaxBlock; p++)
				if (p < b)
					s[p] += (double)row[p] * w;
		}
	if (g < G)
#pragma unroll
		for (int p = 0; p < kMaxBlock; p++)
			if (p < b)
				sh[(g * b + p) * b + q] = s[p];
	__syncthreads();
	if (t < bb) {
		double v = 0;
		for (int k = 0; k < G; k++)
			v += sh[k * bb + t];
		partial[((int64_t)j * gridDim.x + blockIdx.x) * bb + t] = v;
	}
}

// out[e] = sum over the chunks of partial[(j * chunks + c) * bb + e'] for
// e = j * bb + e': one block per output, thread t adding chunks t, t + 256,
// ..., then a fixed shuffle and LDS tree (the same order on every call)
__global__ __launch_bounds__(256) void block_inner_reduce(const double *__restrict__ partial,
							  int chunks, int bb, double *__restrict__ out)
{
	__shared__ double sh[4];
	const int e = blockIdx.x;
	const int j = e / bb, k = e - j * bb;
	double v = 0;
	for (int c = threadIdx.x; c < chunks; c += blockDim.x)
		v += partial[((int64_t)j * chunks + c) * bb + k];
	for (int o = 32; o > 0; o >>= 1)
		v += __shfl_xor(v, o, 64);
	if ((threadIdx.x & 63) == 0)
		sh[threadIdx.x >> 6] = v;
	__syncthreads();
	if (threadIdx.x == 0)
		out[e] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// out[r][q] = in[r][q] (or 0) + sum_{j, p} Q_j[r][p] C[(j b + p) m + q] for
// q < m; in may alias out (each thread reads its row before writing it).
// row_major: out is [m][D] instead of [D][m].
template <typename OT>
__global__ __launch_bounds__(256) void block_combine(const float *__restrict__ Q, int nblk, int b,
						     int64_t D, const double *__restrict__ C, int m,
						     const float *in, OT *out, int row_major)
{
	const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= D)
		return;
	double acc[kMaxBlock];
#pragma unroll
	for (int q = 0; q < kMaxBlock; q++)
		acc[q] = (in && q < m) ? (double)in[r * m + q] : 0.0;
	for (int j = 0; j < nblk; j++) {
		const float *row = Q + ((int64_t)j * D + r) * b;
		for (int p = 0; p < b; p++) {
			const double v = (double)row[p];
			const double *c = C + ((int64_t)j * b + p) * m;
#pragma unroll
			for (int q = 0; q < kMaxBlock; q++)
				if (q < m)
					acc[q] += v * c[q];
		}
	}
#pragma unroll
	for (int q = 0; q < kMaxBlock; q++)
		if (q < m) {
			if (row_major)
				out[(int64_t)q * D + r] = (OT)acc[q];
			else
				out[r * m + q] = (OT)acc[q];
		}
}

// column col of W [D][b] <- uniform [-1, 1) keyed by (seed, row)
__global__ void block_random_column(float *W, int64_t D, int b, int col, uint64_t seed)
{
	const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (r < D)
		W[r * b + col] =
			(float)(2.0 * double(splitmix(seed ^ splitmix((uint64_t)r)) >> 11) * 0x1.0p-53 - 1.0);
}

unsigned grid_of(int64_t n) { return (unsigned)((n + 255) / 256); }

// Cyclic Jacobi (two-sided rotations chosen as in Golub & Van Loan,
// Matrix Computations, Alg. 8.5.2): A = V diag(w) V^T, w ascending,
// eigenvector k in column k of V (row-major n x n).
void jacobi_eigen(int n, std::vector<double> A, std::vector<double> &w, std::vector<double> &V)
{
	std::vector<double> Z((size_t)n * n, 0.0);
	for (int i = 0; i < n; i++)
		Z[(size_t)i * n + i] = 1.0;
	auto a = [&](int i, int j) -> double & { return A[(size_t)i * n + j]; };
	double norm2 = 0;
	for (double v : A)
		norm2 += v * v;
	for (int sweep = 0; sweep < 64; sweep++) {
		double off = 0;
		for (int i = 0; i < n; i++)
			for (int j = i + 1; j < n; j++)
				off += a(i, j) * a(i, j);
		if (off <= 1e-32 * norm2 || off == 0.0)
			break;
		for (int p = 0; p < n; p++)
			for (int q = p + 1; q < n; q++) {
				const double apq = a(p, q);
				if (apq == 0.0)
					continue;
				const double tau = (a(q, q) - a(p, p)) / (2.0 * apq);
				const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
				const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
				// A <- J^T A J with J the rotation in the (p, q) plane
				for (int k = 0; k < n; k++) {
					const double akp = a(k, p), akq = a(k, q);
					a(k, p) = c * akp - s * akq;
					a(k, q) = s * akp + c * akq;
				}
				for (int k = 0; k < n; k++) {
					const double apk = a(p, k), aqk = a(q, k);
					a(p, k) = c * apk - s * aqk;
					a(q, k) = s * apk + c * aqk;
				}
				for (int k = 0; k < n; k++) {
					double *z = &Z[(size_t)k * n];
					const double zp = z[p], zq = z[q];
					z[p] = c * zp - s * zq;
					z[q] = s * zp + c * zq;
				}
			}
	}
	std::vector<int> order(n);
	std::iota(order.begin(), order.end(), 0);
	std::stable_sort(order.begin(), order.end(), [&](int i, int j) { return a(i, i) < a(j, j); });
	w.assign(n, 0.0);
	V.assign((size_t)n * n, 0.0);
	for (int k = 0; k < n; k++) {
		w[k] = a(order[k], order[k]);
		for (int i = 0; i < n; i++)
			V[(size_t)i * n + k] = Z[(size_t)i * n + order[k]];
	}
}

struct Block {
	int64_t D = 0;
	int b = 0;
	int max_blocks = 0;
	float *Q = nullptr;      // max_blocks + 1 blocks [D][b]
	float *X = nullptr;      // matvec input  [D][b]
	float *W = nullptr;      // matvec output [D][b]
	float *Y = nullptr;      // scratch       [D][b]
	double *dC = nullptr;    // combination coefficients
	double *dpart = nullptr; // inner-product partial sums
	double *dsum = nullptr;  // their sums over the chunks
	double matvec_ms = 0;

	float *blk(int j) const { return Q + (int64_t)j * D * b; }

	// every buffer carved from the batch's workspace (256-byte aligned)
	int alloc(bacchus_hip_batch_t *batch)
	{
		auto up = [](size_t n) { return (n + 255) & ~size_t(255); };
		const size_t vec = up((size_t)D * b * sizeof(float));
		const size_t nq = vec * (max_blocks + 1);
		const size_t nc = up(sizeof(double) * (size_t)(max_blocks + 1) * b * kMaxBlock);
		const size_t np = up(sizeof(double) * (size_t)(max_blocks + 1) * kInnerChunks * b * b);
		const size_t ns = up(sizeof(double) * (size_t)(max_blocks + 1) * b * b);
		void *ws = nullptr;
		if (int rc = batch_workspace(batch, nq + 3 * vec + nc + np + ns, &ws))
			return rc;
		char *p = static_cast<char *>(ws);
		Q = reinterpret_cast<float *>(p);
		X = reinterpret_cast<float *>(p + nq);
		W = reinterpret_cast<float *>(p + nq + vec);
		Y = reinterpret_cast<float *>(p + nq + 2 * vec);
		dC = reinterpret_cast<double *>(p + nq + 3 * vec);
		dpart = reinterpret_cast<double *>(p + nq + 3 * vec + nc);
		dsum = reinterpret_cast<double *>(p + nq + 3 * vec + nc + np);
		return 0;
	}

	// C[(j b + p) b + q] = Q_j[:, p] . V[:, q] for the blocks j0 <= j < j0 + nblk
	// with_self: one more block of results, V^T V, after the nblk ones
	int inner(const float *Qs, int nblk, const float *V, std::vector<double> &C, bool with_self = false)
	{
		const int bb = b * b;
		const int rows = nblk + (with_self ? 1 : 0);
		hipLaunchKernelGGL(block_inner, dim3(kInnerChunks, (unsigned)rows), dim3(256), 0, nullptr, Qs, V,
				   D, b, nblk, dpart);
		const int n = rows * bb;
		hipLaunchKernelGGL(block_inner_reduce, dim3((unsigned)n), dim3(256), 0, nullptr, dpart,
				   kInnerChunks, bb, dsum);
		C.resize((size_t)n);
		BACCHUS_HIP_CHECK(hipMemcpy(C.data(), dsum, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
		return 0;
	}

	// out = in + Qs C ((nblk b) x m coefficients)
	template <typename OT>
	int combine(const float *Qs, int nblk, const std::vector<double> &C, int m, const float *in, OT *out,
		    int row_major)
	{
		BACCHUS_HIP_CHECK(hipMemcpy(dC, C.data(), C.size() * sizeof(double), hipMemcpyHostToDevice));
		hipLaunchKernelGGL(block_combine<OT>, dim3(grid_of(D)), dim3(256), 0, nullptr, Qs, nblk, b, D, dC,
				   m, in, out, row_major);
		BACCHUS_HIP_CHECK(hipGetLastError());
		return 0;
	}

	// Projects V off the first nblk stored blocks: V <- V - Q C, C added
	// into coef ((nblk b) x b, may be null)
	int project_out(int nblk, float *V, std::vector<double> *coef)
	{
		if (nblk == 0)
			return 0;
		std::vector<double> C;
		if (int rc = inner(Q, nblk, V, C))
			return rc;
		if (coef)
			for (size_t i = 0; i < C.size(); i++)
				(*coef)[i] += C[i];
		for (double &c : C)
			c = -c;
		return combine<float>(Q, nblk, C, b, V, V, 0);
	}

	// V (orthogonal to the stored blocks) = out R, out orthonormal: SVQB
	// (the Gram matrix's eigensystem); directions whose Gram eigenvalue is
	// below floor get seeded random columns, orthogonalized, with zero
	// rows in R.  V, out and Y are distinct buffers; gram, when given, is
	// V's Gram matrix.
	int svqb(int nblk, float *V, float *out, std::vector<double> &R, double floor, uint64_t seed,
		 double *cond = nullptr, const std::vector<double> *gram = nullptr)
	{
		std::vector<double> G, lam, U;
		if (gram)
			G = *gram;
		else if (int rc = inner(V, 1, V, G))
			return rc;
		jacobi_eigen(b, G, lam, U);
		if (cond)
			*cond = lam[0] > floor && lam[0] > 0 ? lam[b - 1] / lam[0] : INFINITY;
		// out = V U diag(lam^-1/2); R = diag(lam^1/2) U^T
		std::vector<double> C((size_t)b * b, 0.0);
		R.assign((size_t)b * b, 0.0);
		std::vector<int> dead;
		for (int k = 0; k < b; k++) {
			if (!(lam[k] > floor)) {
				dead.push_back(k);
				continue;
			}
			const double s = std::sqrt(lam[k]);
			for (int p = 0; p < b; p++) {
				C[(size_t)p * b + k] = U[(size_t)p * b + k] / s;
				R[(size_t)k * b + p] = s * U[(size_t)p * b + k];
			}
		}
		if (int rc = combine<float>(V, 1, C, b, nullptr, out, 0))
			return rc;
		for (size_t i = 0; i < dead.size(); i++) {
			// a random direction orthogonal to the stored blocks and to
			// the block's live columns
			hipLaunchKernelGGL(block_random_column, dim3(grid_of(D)), dim3(256), 0, nullptr, out,
					   D, b, dead[i], seed + 7919 * (i + 1));
		}
		if (!dead.empty()) {
			for (int pass = 0; pass < 2; pass++) {
				if (int rc = project_out(nblk, out, nullptr))
					return rc;
				// Gram-Schmidt of the dead columns against the block
				std::vector<double> Gb;
				if (int rc = inner(out, 1, out, Gb))
					return rc;
				std::vector<double> M((size_t)b * b, 0.0);
				for (int k = 0; k < b; k++)
					M[(size_t)k * b + k] = 1.0;
				std::vector<char> is_dead(b, 0);
				for (int k : dead)
					is_dead[k] = 1;
				for (int k : dead)
					for (int p = 0; p < b; p++)
						if (!is_dead[p])
							M[(size_t)p * b + k] = -Gb[(size_t)p * b + k];
				if (int rc = combine<float>(out, 1, M, b, nullptr, Y, 0))
					return rc;
				BACCHUS_HIP_CHECK(hipMemcpy(out, Y, (size_t)D * b * sizeof(float),
							    hipMemcpyDeviceToDevice));
			}
			// normalize every column once more (dead columns are not
			// yet orthonormal among themselves); a direction with
			// nothing left (the Krylov space fills the whole space) stays
			// zero, which only happens for the block after the last
			std::vector<double> Gb, l2, U2;
			if (int rc = inner(out, 1, out, Gb))
				return rc;
			jacobi_eigen(b, Gb, l2, U2);
			const double lmax = *std::max_element(l2.begin(), l2.end());
			std::vector<double> C2((size_t)b * b, 0.0), R2((size_t)b * b, 0.0);
			for (int k = 0; k < b; k++) {
				if (!(l2[k] > 1e-12 * lmax))
					continue;
				const double s = std::sqrt(l2[k]);
				for (int p = 0; p < b; p++) {
					C2[(size_t)p * b + k] = U2[(size_t)p * b + k] / s;
					R2[(size_t)k * b + p] = s * U2[(size_t)p * b + k];
				}
			}
			if (int rc = combine<float>(out, 1, C2, b, nullptr, Y, 0))
				return rc;
			BACCHUS_HIP_CHECK(hipMemcpy(out, Y, (size_t)D * b * sizeof(float), hipMemcpyDeviceToDevice));
			// V = out_old R = out_new (R2 R)
			std::vector<double> RR((size_t)b * b, 0.0);
			for (int i = 0; i < b; i++)
				for (int k = 0; k < b; k++)
					for (int j = 0; j < b; j++)
						RR[(size_t)i * b + j] += R2[(size_t)i * b + k] * R[(size_t)k * b + j];
			R.swap(RR);
		}
		return 0;
	}

	// V (b columns) -> out orthonormal and orthogonal to the first nblk
	// blocks, V = Q coef + out R, in up to two rounds of (project out,
	// SVQB).  The first round writes out, coef and R and the condition
	// number of V's Gram matrix after the projection; V is overwritten.
	int first_round(int nblk, float *V, float *out, std::vector<double> &coef,
			std::vector<double> &R, uint64_t seed, double *cond)
	{
		// Q^T V for the stored blocks and V^T V in one launch
		std::vector<double> C;
		if (int rc = inner(Q, nblk, V, C, true))
			return rc;
		const size_t bb = (size_t)b * b;
		const std::vector<double> G0(C.end() - bb, C.end());
		C.resize(C.size() - bb);
		double scale = 0;
		for (int k = 0; k < b; k++)
			scale = std::max(scale, G0[(size_t)k * b + k]);
		const double floor = 1e-24 * scale;
		coef = C;
		if (nblk > 0) {
			for (double &c : C)
				c = -c;
			if (int rc = combine<float>(Q, nblk, C, b, V, V, 0))
				return rc;
		}
		// with no stored blocks V is unchanged and G0 is its Gram matrix
		return svqb(nblk, V, out, R, floor, seed, cond, nblk == 0 ? &G0 : nullptr);
	}

	// The second round on out (X is its scratch output): out = Q c2 +
	// out2 R2, so V = Q (coef + c2 R) + out2 (R2 R); coef and R updated.
	int second_round(int nblk, float *out, std::vector<double> &coef, std::vector<double> &R,
			 uint64_t seed)
	{
		std::vector<double> c2((size_t)nblk * b * b, 0.0);
		if (int rc = project_out(nblk, out, &c2))
			return rc;
		std::vector<double> R2;
		if (int rc = svqb(nblk, out, X, R2, 0.0, seed ^ 0x9e3779b9ULL))
			return rc;
		BACCHUS_HIP_CHECK(hipMemcpy(out, X, (size_t)D * b * sizeof(float), hipMemcpyDeviceToDevice));
		for (int i = 0; i < nblk * b; i++)
			for (int j = 0; j < b; j++) {
				double s = 0;
				for (int k = 0; k < b; k++)
					s += c2[(size_t)i * b + k] * R[(size_t)k * b + j];
				coef[(size_t)i * b + j] += s;
			}
		std::vector<double> RR((size_t)b * b, 0.0);
		for (int i = 0; i < b; i++)
			for (int k = 0; k < b; k++)
				for (int j = 0; j < b; j++)
					RR[(size_t)i * b + j] += R2[(size_t)i * b + k] * R[(size_t)k * b + j];
		R.swap(RR);
		return 0;
	}
};

}  // namespace
}  // namespace bacchus

using namespace bacchus;

extern "C" {

int bacchus_hip_symmetric_eigen(int64_t n, const double *A, double *w, double *V)
{
	if (n <= 0 || n > 4096 || !A || !w)
		return BACCHUS_FAIL("bad arguments (n = %lld)", (long long)n);
	std::vector<double> M(A, A + n * n), lam, Z;
	for (int64_t i = 0; i < n; i++)
		for (int64_t j = 0; j < i; j++) {
			const double s = 0.5 * (M[i * n + j] + M[j * n + i]);
			M[i * n + j] = M[j * n + i] = s;
		}
	jacobi_eigen((int)n, M, lam, Z);
	std::copy(lam.begin(), lam.end(), w);
	if (V)
		std::copy(Z.begin(), Z.end(), V);
	return 0;
}

int bacchus_hip_block_lanczos(bacchus_hip_batch_t *batch,
			      const bacchus_hip_block_lanczos_settings_t *set, double *eigenvalues,
			      double *residuals, double *ritz_vectors,
			      bacchus_hip_lanczos_result_t *result)
{
	const auto t_start = std::chrono::steady_clock::now();
	if (!batch || !set || !eigenvalues)
		return BACCHUS_FAIL("NULL argument");
	int32_t num_ops = 0, b = 0;
	int64_t D = 0;
	if (int rc = batch_shape(batch, &num_ops, &b, &D))
		return rc;
	if (num_ops != 1)
		return BACCHUS_FAIL("the batch must hold one operator (has %d)", num_ops);
	if (b < 1 || b > kMaxBlock || b > D)
		return BACCHUS_FAIL("block size %d outside 1..%d (and the dimension %lld)", b, kMaxBlock,
				    (long long)D);
	const int nev = set->num_eigenvalues;
	if (nev < 1 || nev > b)
		return BACCHUS_FAIL("num_eigenvalues %d outside 1..%d (the block size)", nev, b);
	int J = set->max_num_blocks;
	if (J < 1 || (int64_t)J * b > kMaxKrylov)
		return BACCHUS_FAIL("max_num_blocks %d: Krylov dimension must stay within %d", J, kMaxKrylov);
	// every multiplied block is a full set of new directions
	J = (int)std::min<int64_t>(J, D / b);
	const int conv = set->convergence;
	if (conv != BACCHUS_HIP_CONVERGE_TARGETS && conv != BACCHUS_HIP_CONVERGE_RESIDUALS &&
	    conv != BACCHUS_HIP_CONVERGE_NONE)
		return BACCHUS_FAIL("convergence %d not supported by the block solver", conv);
	if (conv == BACCHUS_HIP_CONVERGE_TARGETS && !set->target_eigenvalues)
		return BACCHUS_FAIL("CONVERGE_TARGETS needs target_eigenvalues");

	Block L;
	L.D = D;
	L.b = b;
	L.max_blocks = J;
	if (int rc = L.alloc(batch))
		return rc;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	BACCHUS_HIP_CHECK(hipEventCreate(&e0));
	BACCHUS_HIP_CHECK(hipEventCreate(&e1));
	struct Events {
		hipEvent_t a, b;
		~Events()
		{
			(void)hipEventDestroy(a);
			(void)hipEventDestroy(b);
		}
	} ev_guard{e0, e1};

	// start block: given or seeded random, orthonormalized
	if (set->start_block)
		BACCHUS_HIP_CHECK(hipMemcpy(L.W, set->start_block, (size_t)D * b * sizeof(float),
					    hipMemcpyDeviceToDevice));
	else
		for (int k = 0; k < b; k++)
			hipLaunchKernelGGL(block_random_column, dim3(grid_of(D)), dim3(256), 0, nullptr, L.W, D,
					   b, k, set->seed * 0x100000001b3ULL + (uint64_t)k);
	std::vector<double> coef, R;
	double cond = INFINITY;
	if (int rc = L.first_round(0, L.W, L.blk(0), coef, R, set->seed + 1, &cond))
		return rc;
	// a well-conditioned start block is orthonormal to storage precision
	// after one round
	if (!(cond < 4.0))
		if (int rc = L.second_round(0, L.blk(0), coef, R, set->seed + 1))
			return rc;

	const int K = J * b;
	std::vector<double> T((size_t)(K + b) * (K + b), 0.0);
	std::vector<double> theta, S, res(nev, 0.0);
	int done = 0;
	bool converged = false;
	for (int j = 0; j < J; j++) {
		// W = H Q_j
		BACCHUS_HIP_CHECK(hipMemcpy(L.X, L.blk(j), (size_t)D * b * sizeof(float),
					    hipMemcpyDeviceToDevice));
		BACCHUS_HIP_CHECK(hipEventRecord(e0, nullptr));
		const float *xs[1] = {L.X};
		float *ys[1] = {L.W};
		if (int rc = bacchus_hip_batch_matvec(batch, xs, ys, nullptr))
			return rc;
		BACCHUS_HIP_CHECK(hipEventRecord(e1, nullptr));
		BACCHUS_HIP_CHECK(hipEventSynchronize(e1));
		float ms = 0;
		BACCHUS_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
		L.matvec_ms += ms;
		// W = Q_{0..j} coef + Q_{j+1} R: the first round decides the
		// eigenvalues of this step; the second one only matters for the
		// blocks after it and runs when the solve goes on
		const uint64_t seed_j = set->seed + 0x51ed27ULL * (uint64_t)(j + 2);
		if (int rc = L.first_round(j + 1, L.W, L.blk(j + 1), coef, R, seed_j, nullptr))
			return rc;
		const int m = (j + 1) * b, ld = K + b;
		auto fill_block = [&]() {
			for (int p = 0; p < b; p++)
				for (int q = 0; q < b; q++) {
					// diagonal block: Q_j^T H Q_j, symmetrized
					const double a = 0.5 * (coef[((size_t)j * b + p) * b + q] +
								coef[((size_t)j * b + q) * b + p]);
					T[(size_t)(j * b + p) * ld + j * b + q] = a;
					// H Q_j = ... + Q_{j+1} R: T[j+1][j] = R, T[j][j+1] = R^T
					T[(size_t)((j + 1) * b + p) * ld + j * b + q] = R[(size_t)p * b + q];
					T[(size_t)(j * b + q) * ld + (j + 1) * b + p] = R[(size_t)p * b + q];
				}
		};
		fill_block();
		std::vector<double> Tm((size_t)m * m);
		for (int p = 0; p < m; p++)
			for (int q = 0; q < m; q++)
				Tm[(size_t)p * m + q] = T[(size_t)p * ld + q];
		jacobi_eigen(m, Tm, theta, S);
		// residual of Ritz pair k: |R S[j b .. (j+1) b, k]| (every pair:
		// the gap of the bound below subtracts the others' residuals)
		res.assign(m, 0.0);
		for (int k = 0; k < m; k++) {
			double s2 = 0;
			for (int p = 0; p < b; p++) {
				double s = 0;
				for (int q = 0; q < b; q++)
					s += R[(size_t)p * b + q] * S[(size_t)(j * b + q) * m + k];
				s2 += s * s;
			}
			res[k] = std::sqrt(s2);
		}
		done = j + 1;
		bool ok = true;
		for (int k = 0; k < nev; k++) {
			if (conv == BACCHUS_HIP_CONVERGE_TARGETS) {
				const double t = set->target_eigenvalues[k];
				ok = ok && std::fabs(theta[k] - t) <= set->tolerance * std::fabs(t);
			} else if (conv == BACCHUS_HIP_CONVERGE_RESIDUALS) {
				// |theta_k - lambda| <= |r_k|^2 / gap_k, gap_k a lower bound
				// of the distance from theta_k to the rest of the spectrum:
				// every other Ritz value theta_i has an eigenvalue within
				// |r_i| of it, so gap_k = min_i (|theta_i - theta_k| - |r_i|).
				// Where that is not positive, the plain residual bound
				// |theta_k - lambda| <= |r_k| decides.
				double gap = INFINITY;
				for (int i = 0; i < m; i++)
					if (i != k)
						gap = std::min(gap, std::fabs(theta[i] - theta[k]) - res[i]);
				const double bound = std::isfinite(gap) && gap > 0 ? res[k] * res[k] / gap : res[k];
				ok = ok && bound <= set->tolerance * std::fabs(theta[k]);
			} else {
				ok = false;
			}
		}
		converged = ok;
		if (ok || j + 1 == J)
			break;
		if (int rc = L.second_round(j + 1, L.blk(j + 1), coef, R, seed_j))
			return rc;
		fill_block();
	}
	const int m = done * b;
	for (int k = 0; k < nev; k++) {
		eigenvalues[k] = theta[k];
		if (residuals)
			residuals[k] = res[k];
	}
	if (ritz_vectors) {
		// Ritz vector k = sum_j Q_j S[j b .. (j+1) b, k], written [nev][D]
		std::vector<double> C((size_t)m * nev);
		for (int i = 0; i < m; i++)
			for (int k = 0; k < nev; k++)
				C[(size_t)i * nev + k] = S[(size_t)i * m + k];
		if (int rc = L.combine<double>(L.Q, done, C, nev, nullptr, ritz_vectors, 1))
			return rc;
		BACCHUS_HIP_CHECK(hipDeviceSynchronize());
	}
	if (result) {
		result->iterations = done;
		result->converged = converged ? 1 : 0;
		result->pad = 0;
		result->matvec_seconds = L.matvec_ms * 1e-3;
		result->total_seconds =
			std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
	}
	return 0;
}

}  // extern "C"
