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
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>
#include <type_traits>
#include <tuple>

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

// fn(k) for every k in [0, n) on up to 16 host threads, indices taken from
// a shared counter; the calling thread works too (and does everything if no
// thread can be started).  The first exception a call of fn throws (e.g.
// std::bad_alloc while a table grows) stops the remaining work and is
// rethrown in the calling thread after every worker has joined.
template <typename F> void parallel_for(size_t n, F &&fn)
{
	const unsigned hw = std::thread::hardware_concurrency();
	const size_t nthreads = std::min<size_t>(n, std::max(1u, std::min(16u, hw ? hw : 1u)));
	std::atomic<size_t> next{0};
	std::exception_ptr failure;
	std::mutex failure_lock;
	auto worker = [&]() {
		try {
			for (size_t q; (q = next.fetch_add(1)) < n;)
				fn(q);
		} catch (...) {
			std::lock_guard<std::mutex> g(failure_lock);
			if (!failure)
				failure = std::current_exception();
			next.store(n);
		}
	};
	std::vector<std::thread> pool;
	for (size_t t = 1; t < nthreads; t++) {
		try {
			pool.emplace_back(worker);
		} catch (...) {
			break;
		}
	}
	worker();
	for (auto &t : pool)
		t.join();
	if (failure)
		std::rethrow_exception(failure);
}

// host stage times of operator creation on stderr under BACCHUS_HIP_VERBOSE
struct StageClock {
	bool on = getenv("BACCHUS_HIP_VERBOSE") != nullptr;
	std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
	void mark(const char *what)
	{
		const auto now = std::chrono::steady_clock::now();
		if (on)
			fprintf(stderr, "[bacchus] operator create: %s %.3f s\n", what,
				std::chrono::duration<double>(now - t).count());
		t = now;
	}
};

// allocator whose resize() leaves new elements uninitialized: every range
// the element table opens is written in full afterwards, so zeroing it
// first would only add a pass over gigabytes
template <typename T> struct NoInitAlloc : std::allocator<T> {
	template <typename U> struct rebind {
		using other = NoInitAlloc<U>;
	};
	NoInitAlloc() = default;
	template <typename U> NoInitAlloc(const NoInitAlloc<U> &) noexcept {}
	template <typename U> void construct(U *q) noexcept { ::new ((void *)q) U; }
	template <typename U, typename... A> void construct(U *q, A &&...a)
	{
		::new ((void *)q) U(std::forward<A>(a)...);
	}
};

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
// load: one element; load2 (FP64): elements i and i + 1 (i even) in one
// 16-byte load
template <> struct Elem<double> {
	using acc_t = double;
	using x_t = double;
	__device__ static double load(const double *p, int64_t i) { return p[i]; }
	__device__ static void load2(const double *p, int64_t i, double &a, double &b)
	{
		const double2 v = *reinterpret_cast<const double2 *>(p + i);
		a = v.x;
		b = v.y;
	}
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

// Owner-computes tile kernel.  A workgroup owns an output tile: the rows
// p_o in [po_begin, po_end) of one basis block, all columns, held in LDS.
// It applies every virtual instruction that writes the block (the
// instruction itself, or for off-diagonal instructions its transpose,
// scheduler.c:310-418) and writes the tile to y once.  Tiles of the
// transposed space read xT and write yT (blocks stored [neutron][proton]),
// with the proton and neutron lists in exchanged roles.
//   NP (c.f. matrix_vector_multiplication.c:134-210): for each row p_o,
//     threads over lane-side entries i (ordered by element index):
//     tile[p_o][n_o(i)] += s_i sum_{j in Lp(p_o)} s_j W[c_i + |C_n| c_j] x[n_i + Nn_in p_j]
//     with W the instruction's dense element sub-block and c the compact
//     element indices (see bacchus_hip_operator_create)
//   P  (:79-132): tile[p_o][n] += sum_{j in Lp(p_o)} s_j V[m_j] x[n + Nn_in p_j]
//     over the tile's rows p_o and all columns n: lanes over columns,
//     entries broadcast (rows of more than 32 columns), or slots of lanes
//     each walking its own entries (narrow rows).
//   N  (:20-77) runs in the transposed space, where it has the form of P
//     with the neutron list over the rows and the protons as columns.
// N and P skip elements with |s V| < 1e-12, s the force-class scale, as the
// reference does on the (Vulcan-combined) element it reads (:50-53,
// :105-108).
template <typename VT>
__global__ __launch_bounds__(kTileThreads) void tile_kernel(
	const TileWork *__restrict__ work, int64_t nwork, const Task *__restrict__ tasks,
	const TInstr *__restrict__ ins,
	const uint32_t *__restrict__ rowptr, const uint2 *__restrict__ ent,
	const uint32_t *__restrict__ keys, const VT *__restrict__ V,
	const typename Elem<VT>::x_t *__restrict__ x_normal,
	const typename Elem<VT>::x_t *__restrict__ x_transposed, double *__restrict__ y_normal,
	double *__restrict__ y_transposed, double s1, double s2, double s3)
{
	using acc_t = typename Elem<VT>::acc_t;
	extern __shared__ double tile[];
	// tiles are sorted heaviest first and dispatched in that order
	const TileWork w = work[blockIdx.x];
	(void)nwork;
	const auto *__restrict__ x = w.space ? x_transposed : x_normal;
	double *__restrict__ y = w.space ? y_transposed : y_normal;
	const int nrow = w.po_end - w.po_begin;
	const int nn = w.nn_out;
	const int telems = nrow * nn;
	for (int t = threadIdx.x; t < telems; t += blockDim.x)
		tile[t] = 0.0;
	__syncthreads();
	// tasks are taken by waves independently; all accumulation into the
	// tile is by LDS atomics
	const int lane = threadIdx.x & 63;
	// wave-uniform by construction; readfirstlane tells the compiler, so
	// the task, the instruction and all address bases live in SGPRs
	const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const int nwaves = kTileThreads >> 6;
	for (int32_t ti = w.t_begin + wave; ti < w.t_end; ti += nwaves) {
		const Task T = tasks[ti];
		const TInstr I = ins[T.vi];
		const acc_t sc = acc_t(I.force == 1 ? s1 : (I.force == 2 ? s2 : s3));
		if (sizeof(VT) == sizeof(double) && I.type == 3 && I.expanded == 2) {
			if constexpr (sizeof(VT) == sizeof(double)) {
			// expanded element rows with paired lanes: lane entries 2k and
			// 2k + 1 share their input index (runs padded with zero
			// elements), so every lane takes pairs: one 2-element load of
			// the row and one load of x serve two products.  Otherwise as
			// below.
			const VT *vblk = V + I.v_off;
			const auto *xblk = x + I.x_off;
			const uint32_t dn = (uint32_t)I.dn, nn_in = (uint32_t)I.nn_in;
			constexpr int KP = kLaneEntries / 2;  // pairs per lane
			for (uint32_t e0 = T.e0; e0 < T.e1; e0 += 128 * KP) {
				const int nc = (int)min((uint32_t)KP, (T.e1 - e0 + 127u) / 128u);
				bool vl[KP];
				uint32_t col[KP], xin[KP], outa[KP], outb[KP];
				acc_t sga[KP], sgb[KP], acca[KP], accb[KP];
#pragma unroll
				for (int c = 0; c < KP; c++) {
					// e0, e1 and n_len are even: a pair never straddles
					// the task's end
					const uint32_t e = e0 + 128 * c + 2 * lane;
					vl[c] = e < T.e1;
					const uint32_t k = I.n_base + (vl[c] ? e : T.e0);
					// k is even (paired copies start at even entries)
					const uint4 eab = *reinterpret_cast<const uint4 *>(ent + k);
					const uint2 kab = *reinterpret_cast<const uint2 *>(keys + k);
					col[c] = k - I.n_base;
					xin[c] = eab.x;
					outa[c] = kab.x;
					outb[c] = kab.y;
					sga[c] = apply_sign(sc, eab.y);
					sgb[c] = apply_sign(sc, eab.w);
					acca[c] = accb[c] = 0;
				}
				auto pairs_pass = [&](auto live) {
				constexpr int NC = decltype(live)::value;
				int cur = -1;
				for (uint32_t j0 = T.jb; j0 < T.je; j0 += 64) {
					const uint32_t cnt = min(64u, T.je - j0);
					const uint32_t jl = j0 + min((uint32_t)lane, cnt - 1);
					const uint2 pl = ent[jl];
					const uint32_t prow = keys[jl];
					for (uint32_t t = 0; t < cnt; t++) {
						const int row = (int)__builtin_amdgcn_readlane(prow, t);
						if (row != cur) {
							if (cur >= 0) {
								const int r = cur - w.po_begin;
#pragma unroll
								for (int c = 0; c < NC; c++)
									if (vl[c]) {
										atomicAdd(&tile[r * nn + outa[c]], (double)(sga[c] * acca[c]));
										atomicAdd(&tile[r * nn + outb[c]], (double)(sgb[c] * accb[c]));
									}
							}
#pragma unroll
							for (int c = 0; c < NC; c++)
								acca[c] = accb[c] = 0;
							cur = row;
						}
						const uint32_t px = __builtin_amdgcn_readlane(pl.x, t);
						const uint32_t py = __builtin_amdgcn_readlane(pl.y, t);
						const VT *vrow = vblk + (size_t)(dn * (py & kIndexMask));
						const auto *xcol = xblk + (size_t)(nn_in * px);
						acc_t va[KP], vb[KP], xa[KP];
#pragma unroll
						for (int c = 0; c < NC; c++) {
							Elem<VT>::load2(vrow, col[c], va[c], vb[c]);
							xa[c] = acc_t(xcol[xin[c]]);
						}
#pragma unroll
						for (int c = 0; c < NC; c++) {
							const acc_t xs = apply_sign(xa[c], py);
							acca[c] += va[c] * xs;
							accb[c] += vb[c] * xs;
						}
					}
				}
				if (cur >= 0) {
					const int r = cur - w.po_begin;
#pragma unroll
					for (int c = 0; c < NC; c++)
						if (vl[c]) {
							atomicAdd(&tile[r * nn + outa[c]], (double)(sga[c] * acca[c]));
							atomicAdd(&tile[r * nn + outb[c]], (double)(sgb[c] * accb[c]));
						}
				}
				};
				static_assert(KP == 2, "one pass per count of live pair chunks");
				if (nc == KP)
					pairs_pass(std::integral_constant<int, 2>{});
				else
					pairs_pass(std::integral_constant<int, 1>{});
			}
			}
		} else if (I.type == 3) {
			// row-side entries [jb, je) of the tile's rows (contiguous in
			// the copy ordered by row; keys[] gives each entry's row):
			// every lane holds kLaneEntries neighbouring lane-side entries
			// (chunks of 64), so each broadcast row-side entry puts
			// 2 kLaneEntries independent loads in flight.  Row-side entries
			// arrive with one coalesced load per 64 and are broadcast with
			// readlane; a lane's partial sums go to the tile whenever the
			// (uniform) row changes.
			const VT *vblk = V + I.v_off;
			const auto *xblk = x + I.x_off;
			const uint32_t dn = (uint32_t)I.dn, nn_in = (uint32_t)I.nn_in;
			for (uint32_t e0 = T.e0; e0 < T.e1; e0 += 64 * kLaneEntries) {
				// chunks of 64 lane entries that hold any (wave-uniform):
				// the loads of empty chunks are skipped
				const int nc = (int)min((uint32_t)kLaneEntries, (T.e1 - e0 + 63u) / 64u);
				bool vl[kLaneEntries];
				uint32_t col[kLaneEntries], xin[kLaneEntries], outc[kLaneEntries];
				acc_t sg[kLaneEntries], acc[kLaneEntries];
#pragma unroll
				for (int c = 0; c < kLaneEntries; c++) {
					const uint32_t e = e0 + 64 * c + lane;
					vl[c] = e < T.e1;
					const uint32_t k = I.n_base + (vl[c] ? e : T.e0);
					const uint2 en = ent[k];
					// element column: the lane entry's compact index, or
					// its position in the lane list for expanded rows
					col[c] = I.expanded ? k - I.n_base : en.y & kIndexMask;
					xin[c] = en.x;
					outc[c] = keys[k];
					sg[c] = apply_sign(sc, en.y);
					acc[c] = 0;
				}
				// the row loop is compiled for a fixed count NC of chunks
				// that hold entries (empty chunks then cost nothing), or
				// with NC = 0 for a count checked at run time
				auto rows_pass = [&](auto live) {
				constexpr int NC = decltype(live)::value;
				constexpr int NL = NC ? NC : kLaneEntries;
				int cur = -1;
				for (uint32_t j0 = T.jb; j0 < T.je; j0 += 64) {
					const uint32_t cnt = min(64u, T.je - j0);
					const uint32_t jl = j0 + min((uint32_t)lane, cnt - 1);
					const uint2 pl = ent[jl];
					const uint32_t prow = keys[jl];
					for (uint32_t t = 0; t < cnt; t++) {
						const int row = (int)__builtin_amdgcn_readlane(prow, t);
						if (row != cur) {
							if (cur >= 0) {
								const int r = cur - w.po_begin;
#pragma unroll
								for (int c = 0; c < NL; c++)
									if (vl[c])
										atomicAdd(&tile[r * nn + outc[c]], (double)(sg[c] * acc[c]));
							}
#pragma unroll
							for (int c = 0; c < NL; c++)
								acc[c] = 0;
							cur = row;
						}
						const uint32_t px = __builtin_amdgcn_readlane(pl.x, t);
						const uint32_t py = __builtin_amdgcn_readlane(pl.y, t);
						const VT *vrow = vblk + (size_t)(dn * (py & kIndexMask));
						const auto *xcol = xblk + (size_t)(nn_in * px);
						acc_t va[kLaneEntries], xa[kLaneEntries];
#pragma unroll
						for (int c = 0; c < NL; c++) {
							if (NC || c < nc) {
								va[c] = Elem<VT>::load(vrow, col[c]);
								xa[c] = acc_t(xcol[xin[c]]);
							} else {
								va[c] = 0;
								xa[c] = 0;
							}
						}
#pragma unroll
						for (int c = 0; c < NL; c++)
							acc[c] += apply_sign(va[c], py) * xa[c];
					}
				}
				if (cur >= 0) {
					const int r = cur - w.po_begin;
#pragma unroll
					for (int c = 0; c < NL; c++)
						if (vl[c])
							atomicAdd(&tile[r * nn + outc[c]], (double)(sg[c] * acc[c]));
				}
				};
				static_assert(kLaneEntries == 4, "one row loop per count of live chunks");
				if (nc == kLaneEntries)
					rows_pass(std::integral_constant<int, 4>{});
				else if constexpr (sizeof(VT) == sizeof(double)) {
					// FP64 elements: a copy per count (measured faster
					// there; the FP32/BF16 kernels lose with the extra
					// copies)
					if (nc == 3)
						rows_pass(std::integral_constant<int, 3>{});
					else if (nc == 2)
						rows_pass(std::integral_constant<int, 2>{});
					else
						rows_pass(std::integral_constant<int, 1>{});
				} else
					rows_pass(std::integral_constant<int, 0>{});
			}
		} else if (I.type == 2 && T.e1 - T.e0 <= 32u) {
			// narrow rows (<= 32 columns): 64 / S entry slots of S lanes,
			// one column per lane; slot g walks entries g, g + G, ... of
			// [jb, je) (in row order), so a slot's x loads are one
			// contiguous row segment and its entry and element loads are
			// shared by its lanes; sums go to the tile on row changes
			const auto *xblk = x + I.x_off;
			const uint32_t nn_in = (uint32_t)I.nn_in;
			const uint32_t ncol = T.e1 - T.e0;
			const int ls = ncol <= 1 ? 0 : 32 - __builtin_clz(ncol - 1);
			const int G = 64 >> ls;
			const int g = lane >> ls;
			const uint32_t c = (uint32_t)lane & ((1u << ls) - 1u);
			const bool valid = c < ncol;
			const uint32_t nl = T.e0 + (valid ? c : 0u);
			acc_t a = 0;
			int cur = -1;
			for (uint32_t j0 = T.jb + (uint32_t)g; j0 < T.je; j0 += 4u * (uint32_t)G) {
				uint2 en[4];
				int rw[4];
				acc_t e[4], xv[4];
#pragma unroll
				for (int u = 0; u < 4; u++) {
					const uint32_t j = j0 + (uint32_t)(u * G);
					const bool ok = j < T.je;
					en[u] = ent[ok ? j : T.jb];
					rw[u] = ok ? (int)keys[j] : -1;
				}
#pragma unroll
				for (int u = 0; u < 4; u++) {
					e[u] = Elem<VT>::load(V, I.v_off + (en[u].y & kIndexMask));
					xv[u] = acc_t((xblk + (size_t)nn_in * en[u].x)[nl]);
				}
#pragma unroll
				for (int u = 0; u < 4; u++) {
					if (rw[u] < 0)
						break;
					if (rw[u] != cur) {
						if (cur >= 0 && valid && a != acc_t(0))
							atomicAdd(&tile[(cur - w.po_begin) * nn + nl], (double)(sc * a));
						a = 0;
						cur = rw[u];
					}
					// the reference skips on the scaled element (a
					// Vulcan-combined file holds s V)
					if (!(fabs(sc * e[u]) < acc_t(1e-12)))
						a += apply_sign(e[u], en[u].y) * xv[u];
				}
			}
			if (cur >= 0 && valid && a != acc_t(0))
				atomicAdd(&tile[(cur - w.po_begin) * nn + nl], (double)(sc * a));
		} else if (I.type == 2) {
			// proton entries [jb, je) of the tile's rows: lanes over
			// neighbouring neutron states n (coalesced x); entries broadcast
			// with readlane; partial sums go to the tile on row changes
			const auto *xblk = x + I.x_off;
			const uint32_t nn_in = (uint32_t)I.nn_in;
			for (int n0 = (int)T.e0; n0 < (int)T.e1; n0 += 64) {
				const int n = n0 + lane;
				const bool valid = n < (int)T.e1;
				const uint32_t nl = valid ? (uint32_t)n : T.e0;
				acc_t a0 = 0, a1 = 0;
				int cur = -1;
				for (uint32_t j0 = T.jb; j0 < T.je; j0 += 64) {
					const uint32_t cnt = min(64u, T.je - j0);
					const uint32_t jl = j0 + min((uint32_t)lane, cnt - 1);
					const uint2 pl = ent[jl];
					const uint32_t prow = keys[jl];
					for (uint32_t t = 0; t < cnt; t++) {
						const int row = (int)__builtin_amdgcn_readlane(prow, t);
						if (row != cur) {
							if (cur >= 0 && valid && (a0 + a1) != acc_t(0))
								atomicAdd(&tile[(cur - w.po_begin) * nn + n], (double)(sc * (a0 + a1)));
							a0 = a1 = 0;
							cur = row;
						}
						const uint32_t px = __builtin_amdgcn_readlane(pl.x, t);
						const uint32_t py = __builtin_amdgcn_readlane(pl.y, t);
						const acc_t e = Elem<VT>::load(V, I.v_off + (py & kIndexMask));
						const acc_t xv = acc_t((xblk + (size_t)(nn_in * px))[nl]);
						if (!(fabs(sc * e) < acc_t(1e-12))) {
							if (t & 1)
								a1 += apply_sign(e, py) * xv;
							else
								a0 += apply_sign(e, py) * xv;
						}
					}
				}
				if (cur >= 0 && valid && (a0 + a1) != acc_t(0))
					atomicAdd(&tile[(cur - w.po_begin) * nn + n], (double)(sc * (a0 + a1)));
			}
		}
	}
	__syncthreads();
	double *yblk = y + w.y_off + (int64_t)nn * w.po_begin;
	for (int t = threadIdx.x; t < telems; t += blockDim.x) {
		const double v = tile[t];
		if (w.shared) {
			if (v != 0.0)
				atomicAdd(yblk + t, v);
		} else
			yblk[t] += v;
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
	    const typename Elem<VT>::x_t *xT, double *y, double s1, double s2, double s3,
	    hipStream_t s)
{
	if (op->num_tiles)
		hipLaunchKernelGGL(tile_kernel<VT>, dim3(op->num_tiles), dim3(kTileThreads),
				   op->tile_lds_bytes, s, op->d_tiles, op->num_tiles, op->d_tasks, op->d_tinstr,
				   op->d_rowptr, op->d_entries, op->d_keys, V, x, xT, y, op->d_yT, s1, s2, s3);
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

namespace bacchus {

// Rows of every block are grouped into tiles of about pp.target
// multiplications (times the tile's vector count) within the LDS budget;
// a single overloaded group is split over instruction ranges (those tiles
// flush with atomics).  Tiles are ordered heaviest first so the long ones
// do not form the tail; every tile gets tasks: N instructions in entry
// ranges over all rows, P and NP instructions as pieces of the contiguous
// proton segment of the tile's rows times ranges of neutron entries (NP)
// or neutron states (P).  With input groups the instructions of a block
// are split by the group of their input block first, and every group gets
// its own tiles.
int make_plan(const bacchus_hip_operator *op, const PlanParams &pp, Plan &out)
{
	const std::vector<uint32_t> &rowptr = op->h_rowptr;
	const std::vector<TInstr> &tinstr = op->h_tinstr;
	const int64_t nb = (int64_t)op->block_np.size();
	out = Plan();
	out.vi_order.resize(tinstr.size());
	std::iota(out.vi_order.begin(), out.vi_order.end(), 0);
	auto slot_end = [&](int64_t slot) {
		return pp.skip_np ? op->blk_np[slot] : op->blk_vi[slot + 1];
	};
	auto dims = [&](int64_t slot, int64_t &nn, int64_t &np) {
		const int64_t b = slot % nb;
		nn = slot >= nb ? op->block_np[b] : op->block_nn[b];
		np = slot >= nb ? op->block_nn[b] : op->block_np[b];
	};
	// multiplications of virtual instruction v over rows [po0, po1)
	auto icost = [&](const TInstr &t, int64_t nn, int64_t po0, int64_t po1) {
		const double e = double(rowptr[t.p_rows + po1] - rowptr[t.p_rows + po0]);
		const double w = t.type == 3 ? double(t.n_len) : double(nn);
		return w * e + w * 0.25 * double(po1 - po0);
	};
	// input group of every virtual instruction
	std::vector<int32_t> vgroup(tinstr.size(), 0);
	const int G = std::max(1, pp.in_groups);
	if (G > 1) {
		std::map<int64_t, int64_t> block_at;
		for (int64_t b = 0; b < nb; b++)
			block_at[op->block_off[b]] = b;
		std::vector<int64_t> vblock(tinstr.size(), 0);
		std::vector<double> bcost(nb, 0.0);
		for (int64_t slot = 0; slot < 2 * nb; slot++) {
			int64_t nn, np;
			dims(slot, nn, np);
			const int bt = pp.bt.empty() ? 1 : pp.bt[slot];
			for (int32_t v = op->blk_vi[slot]; v < slot_end(slot); v++) {
				const auto it = block_at.find(tinstr[v].x_off);
				if (it == block_at.end())
					return BACCHUS_FAIL("instruction input at %lld is no block start",
							    (long long)tinstr[v].x_off);
				vblock[v] = it->second;
				bcost[it->second] += icost(tinstr[v], nn, 0, np) * bt;
			}
		}
		std::vector<double> own(G, 0.0);
		std::vector<double> &load = pp.group_load ? *pp.group_load : own;
		load.resize(G, 0.0);
		std::vector<int64_t> bo(nb);
		std::iota(bo.begin(), bo.end(), 0);
		std::stable_sort(bo.begin(), bo.end(),
				 [&](int64_t a, int64_t b) { return bcost[a] > bcost[b]; });
		std::vector<int32_t> bgroup(nb, 0);
		for (int64_t b : bo) {
			if (bcost[b] == 0)
				continue;
			const int g = (int)(std::min_element(load.begin(), load.end()) - load.begin());
			bgroup[b] = g;
			load[g] += bcost[b];
		}
		for (size_t v = 0; v < tinstr.size(); v++)
			vgroup[v] = bgroup[vblock[v]];
	}
	std::vector<int> tile_bt;
	for (int64_t slot = 0; slot < 2 * nb; slot++) {
		const int32_t first = op->blk_vi[slot];
		const int32_t last = slot_end(slot);
		if (first == last)
			continue;
		const int64_t b = slot % nb;
		const int space = slot >= nb;
		// rows x columns of the block in its space
		int64_t nn, np;
		dims(slot, nn, np);
		const int bt = pp.bt.empty() ? 1 : pp.bt[slot];
		if (nn * bt > pp.tile_elems)
			return BACCHUS_FAIL("basis block %lld has rows of %lld states; a tile "
					    "of %d vectors holds at most %lld", (long long)b,
					    (long long)nn, bt, (long long)(pp.tile_elems / bt));
		// the slot's instructions by group (stable: instruction order
		// within a group)
		int32_t *ord = out.vi_order.data();
		std::stable_sort(ord + first, ord + last,
				 [&](int32_t u, int32_t w) { return vgroup[u] < vgroup[w]; });
		int ngroups = 0;
		for (int32_t k = first; k < last; k++)
			ngroups += k == first || vgroup[ord[k]] != vgroup[ord[k - 1]];
		for (int32_t gfirst = first; gfirst < last;) {
			const int32_t g = vgroup[ord[gfirst]];
			int32_t glast = gfirst;
			while (glast < last && vgroup[ord[glast]] == g)
				glast++;
			// cost of every (instruction, row) and of every row
			const size_t nv = (size_t)(glast - gfirst);
			std::vector<double> cost(nv * np, 0.0), row_cost(np, 0.0);
			for (size_t k = 0; k < nv; k++) {
				const TInstr &t = tinstr[ord[gfirst + k]];
				for (int64_t po = 0; po < np; po++) {
					const double c = icost(t, nn, po, po + 1) * bt;
					cost[k * np + po] = c;
					row_cost[po] += c;
				}
			}
			const int64_t rows_max = std::max<int64_t>(1, pp.tile_elems / (nn * bt));
			int64_t po0 = 0;
			while (po0 < np) {
				int64_t po1 = po0 + 1;
				double c = row_cost[po0];
				while (po1 < np && po1 - po0 < rows_max && c + row_cost[po1] <= pp.target) {
					c += row_cost[po1];
					po1++;
				}
				out.max_tile_elems = std::max(out.max_tile_elems, (po1 - po0) * nn * bt);
				TileWork w{};
				w.y_off = op->block_off[b];
				w.nn_out = (int32_t)nn;
				w.po_begin = (int32_t)po0;
				w.po_end = (int32_t)po1;
				w.space = space;
				w.group = g;
				const int pieces = c > 2.0 * pp.target ? (int)std::ceil(c / pp.target) : 1;
				if (pieces == 1) {
					w.vi_begin = gfirst;
					w.vi_end = glast;
					w.shared = ngroups > 1;
					out.tiles.push_back(w);
					out.cost.push_back(c);
					tile_bt.push_back(bt);
				} else {
					const double per = c / pieces;
					int32_t begin = 0;
					double acc = 0;
					for (size_t k = 0; k < nv; k++) {
						for (int64_t po = po0; po < po1; po++)
							acc += cost[k * np + po];
						if (acc >= per || k + 1 == nv) {
							w.vi_begin = gfirst + begin;
							w.vi_end = gfirst + (int32_t)k + 1;
							w.shared = 1;
							out.tiles.push_back(w);
							out.cost.push_back(acc);
							tile_bt.push_back(bt);
							begin = (int32_t)k + 1;
							acc = 0;
						}
					}
				}
				po0 = po1;
			}
			gfirst = glast;
		}
	}
	// heaviest first
	std::vector<size_t> order(out.tiles.size());
	std::iota(order.begin(), order.end(), 0);
	std::stable_sort(order.begin(), order.end(),
			 [&](size_t a, size_t b) { return out.cost[a] > out.cost[b]; });
	std::vector<TileWork> tiles;
	std::vector<double> costs;
	std::vector<int> bts;
	for (size_t k : order) {
		tiles.push_back(out.tiles[k]);
		costs.push_back(out.cost[k]);
		bts.push_back(tile_bt[k]);
	}
	out.tiles.swap(tiles);
	out.cost.swap(costs);
	out.bt = bts;
	if (pp.xcds > 1 && !out.tiles.empty()) {
		// XCD-aware order: workgroup i runs on XCD i mod xcds, so every
		// output block (its element sub-blocks are walked by all its tiles
		// in the same instruction order) goes to one XCD and its tiles are
		// dispatched together there, sharing that XCD's L2.  Blocks are
		// balanced over the XCDs by cost; holes are padded with empty
		// tiles.  With input groups, group g takes the place of the
		// output blocks: its tiles go to one XCD.
		std::map<std::pair<int, int64_t>, std::vector<size_t>> groups;
		for (size_t k = 0; k < out.tiles.size(); k++)
			if (G > 1)
				groups[{0, out.tiles[k].group}].push_back(k);
			else
				groups[{out.tiles[k].space, out.tiles[k].y_off}].push_back(k);
		std::vector<std::pair<double, std::vector<size_t>>> gl;
		for (auto &g : groups) {
			double c = 0;
			for (size_t k : g.second)
				c += out.cost[k];
			gl.push_back({c, g.second});
		}
		std::stable_sort(gl.begin(), gl.end(),
				 [](const auto &a, const auto &b) { return a.first > b.first; });
		const int X = pp.xcds;
		std::vector<double> load(X, 0.0);
		std::vector<std::vector<size_t>> bins(X);
		for (auto &g : gl) {
			const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
			load[k] += g.first;
			bins[k].insert(bins[k].end(), g.second.begin(), g.second.end());
		}
		// within an XCD: the tiles of one output block (they read the same
		// element sub-blocks, split by rows) one after another, so their
		// shared element rows are still in the XCD's L2; blocks heaviest
		// first (BACCHUS_HIP_TILE_ORDER=0: every tile heaviest first)
		const char *ord = getenv("BACCHUS_HIP_TILE_ORDER");
		if (!(ord && atoi(ord) == 0))
			for (auto &b : bins) {
				std::map<std::pair<int, int64_t>, double> bc;
				for (size_t i : b)
					bc[{out.tiles[i].space, out.tiles[i].y_off}] += out.cost[i];
				std::stable_sort(b.begin(), b.end(), [&](size_t u, size_t v) {
					const TileWork &a = out.tiles[u], &c = out.tiles[v];
					const double ca = bc[{a.space, a.y_off}], cc = bc[{c.space, c.y_off}];
					if (ca != cc)
						return ca > cc;
					return std::make_tuple(a.space, a.y_off, a.vi_begin, a.po_begin) <
					       std::make_tuple(c.space, c.y_off, c.vi_begin, c.po_begin);
				});
			}
		size_t len = 0;
		for (auto &b : bins)
			len = std::max(len, b.size());
		TileWork empty{};
		empty.space = 0;
		std::vector<TileWork> t2;
		std::vector<double> c2;
		std::vector<int> b2;
		for (size_t r = 0; r < len; r++)
			for (int k = 0; k < X; k++) {
				if (r < bins[k].size()) {
					const size_t i = bins[k][r];
					t2.push_back(out.tiles[i]);
					c2.push_back(out.cost[i]);
					b2.push_back(out.bt[i]);
				} else {
					t2.push_back(empty);
					c2.push_back(0.0);
					b2.push_back(1);
				}
			}
		out.tiles.swap(t2);
		out.cost.swap(c2);
		out.bt.swap(b2);
	}
	for (size_t ti = 0; ti < out.tiles.size(); ti++) {
		TileWork &t = out.tiles[ti];
		const uint64_t task_mults = std::max<uint64_t>(4096, pp.task_mults / out.bt[ti]);
		t.t_begin = (int32_t)out.tasks.size();
		for (int32_t k = t.vi_begin; k < t.vi_end; k++) {
			const int32_t v = out.vi_order[k];
			const TInstr &I = tinstr[v];
			const uint32_t jb = rowptr[I.p_rows + t.po_begin], je = rowptr[I.p_rows + t.po_end];
			if (je <= jb)
				continue;
			const uint32_t span = I.type == 3 ? I.n_len : (uint32_t)t.nn_out;
			const uint32_t unit = I.type == 3 ? 64u * kLaneEntries : 64u;
			const uint32_t e_step = std::min<uint32_t>((span + unit - 1) / unit * unit, 1024u);
			const uint32_t j_step = std::max<uint32_t>(64u, (uint32_t)(task_mults / e_step));
			for (uint32_t jj = jb; jj < je; jj += j_step)
				for (uint32_t e = 0; e < span; e += e_step)
					out.tasks.push_back(Task{v, 0, jj, std::min(je, jj + j_step), e,
								 std::min(span, e + e_step), 0, 0});
		}
		t.t_end = (int32_t)out.tasks.size();
		if (out.tasks.size() > 0x7FFFFFFFull)
			return BACCHUS_FAIL("more than 2^31 tile tasks");
	}
	return 0;
}

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
	(void)hipFree(op->d_tinstr);
	(void)hipFree(op->d_tiles);
	(void)hipFree(op->d_tasks);
	(void)hipFree(op->d_keys);
	(void)hipFree(op->d_xf);
	(void)hipFree(op->d_xT);
	(void)hipFree(op->d_yT);
	(void)hipFree(op->d_blk_off);
	(void)hipFree(op->d_blk_np);
	(void)hipFree(op->d_blk_nn);
	np_free(op);
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
		*num_work_items = op->num_tiles;
	return 0;
}

static int operator_create(const bacchus_hip_problem_t *p, int precision,
			   bacchus_hip_operator_t **out);

// the tables of a large operator take gigabytes of host memory: a failed
// allocation while building them comes back as an error code, not as an
// exception through the C ABI
int bacchus_hip_operator_create(const bacchus_hip_problem_t *p,
				int precision, bacchus_hip_operator_t **out)
{
	try {
		return operator_create(p, precision, out);
	} catch (const std::bad_alloc &) {
		return BACCHUS_FAIL("out of host memory while building the operator tables");
	} catch (const std::exception &e) {
		return BACCHUS_FAIL("operator tables: %s", e.what());
	}
}

static int operator_create(const bacchus_hip_problem_t *p, int precision,
			   bacchus_hip_operator_t **out)
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
	// frees the operator (host tables and device memory) on every early
	// return and exception; released on success
	std::unique_ptr<bacchus_hip_operator, void (*)(bacchus_hip_operator *)> guard(
		op, bacchus_hip_operator_free);
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
			return BACCHUS_FAIL("basis block %lld has invalid size %lldx%lld",
					    (long long)b, (long long)op->block_np[b],
					    (long long)op->block_nn[b]);
		}
		op->block_off[b + 1] = op->block_off[b] + op->block_np[b] * op->block_nn[b];
	}
	op->dimension = op->block_off[nb];
	auto fail = [](int code) { return code; };

	StageClock clk;
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

	clk.mark("validation");
	// --- compact element indices: every used list gets the sorted set of
	// element indices it references; entries carry the position in that
	// set (phase bit kept).  Each (element block, neutron list, proton list)
	// then gets a dense sub-block W: N/P = the referenced elements, NP =
	// rows over the proton set, columns over the neutron set.  An
	// instruction's element loads are then neighbours and every loaded line
	// is used. ---
	std::vector<std::vector<uint32_t>> mvals(nl);
	parallel_for((size_t)nl, [&](size_t l) {
		if (!use[l].used)
			return;
		auto &mv = mvals[l];
		mv.reserve(p->list_offsets[l + 1] - p->list_offsets[l]);
		for (int64_t k = p->list_offsets[l]; k < p->list_offsets[l + 1]; k++)
			mv.push_back((uint32_t)p->list_entries[3 * k + 2] & 0x7FFFFFFFu);
		std::sort(mv.begin(), mv.end());
		mv.erase(std::unique(mv.begin(), mv.end()), mv.end());
	});
	auto compact = [&](int64_t l, int32_t m) -> uint32_t {
		const uint32_t v = (uint32_t)m & 0x7FFFFFFFu;
		const auto &mv = mvals[l];
		const uint32_t c = (uint32_t)(std::lower_bound(mv.begin(), mv.end(), v) - mv.begin());
		return c | ((uint32_t)m & 0x80000000u);
	};
	clk.mark("compact element indices");
	// --- which instructions run in the transposed space (every block read
	// and written as [neutron][proton], protons fastest): NP instructions
	// whose proton list is the longer one (the kernel's lanes then walk the
	// long list, as for the others) and P instructions of blocks with few
	// neutron states (they become N instructions there, whose lanes walk
	// the list instead of the neutron states) ---
	auto list_len = [&](int32_t l) -> int64_t {
		return l >= 0 ? p->list_offsets[l + 1] - p->list_offsets[l] : 0;
	};
	std::vector<char> swapped(ni, 0);
	for (int64_t i = 0; i < ni; i++) {
		const bacchus_hip_instruction_t &I = p->instructions[i];
		if (I.type == 3)
			swapped[i] = list_len(I.proton_list) > list_len(I.neutron_list);
		else
			swapped[i] = I.type == 1;
	}
	// lane-side lists whose elements are used at most about twice, or
	// three times for 3-body instructions (whose element sub-blocks are too
	// large to stay in cache, so contiguous rows pay for more copies; a
	// 2-body sub-block is small and its gathers hit the caches): their
	// instructions store the element rows expanded into lane-entry order,
	// so lane loads of elements are contiguous and the lane list can be
	// ordered by input index (clustered x loads)
	std::vector<char> expanded(nl, 0), lane_of_3body(nl, 0);
	for (int64_t i = 0; i < ni; i++) {
		const bacchus_hip_instruction_t &I = p->instructions[i];
		if (I.type == 3 && p->matrix_force && p->matrix_force[I.matrix_block] == 3)
			lane_of_3body[swapped[i] ? I.proton_list : I.neutron_list] = 1;
	}
	// uses per element up to which a lane list is expanded: 2 for 2-body
	// lists, 3 for 3-body ones (BACCHUS_HIP_EXPAND=<2-body>:<3-body> to
	// measure others; 0 expands nothing)
	int64_t expand2 = 2, expand3 = 3;
	if (const char *ex = getenv("BACCHUS_HIP_EXPAND")) {
		long a = 2, b = 3;
		if (sscanf(ex, "%ld:%ld", &a, &b) >= 1) {
			expand2 = a;
			expand3 = strchr(ex, ':') ? b : a;
		}
	}
	for (int64_t i = 0; i < ni; i++) {
		const bacchus_hip_instruction_t &I = p->instructions[i];
		if (I.type != 3)
			continue;
		const int32_t l = swapped[i] ? I.proton_list : I.neutron_list;
		expanded[l] = list_len(l) <= (lane_of_3body[l] ? expand3 : expand2) * (int64_t)mvals[l].size();
	}
	auto lane_list = [&](const bacchus_hip_instruction_t &I, int64_t i) -> int32_t {
		return swapped[i] ? I.proton_list : I.neutron_list;
	};
	std::map<std::tuple<int32_t, int32_t, int32_t>, int64_t> sub_of;
	std::vector<int64_t> sub_off(ni, -1);
	std::vector<double, NoInitAlloc<double>> W;
	// places in instruction order first, then every sub-block filled by
	// host threads into its own range
	std::vector<std::pair<int64_t, int64_t>> sub_jobs;  // (offset, instruction)
	int64_t sub_total = 0;
	for (int64_t i = 0; i < ni; i++) {
		const bacchus_hip_instruction_t &I = p->instructions[i];
		if (I.type == 3 && expanded[lane_list(I, i)])
			continue;  // built per direction below
		const auto key = std::make_tuple(I.matrix_block, I.type != 2 ? I.neutron_list : -1,
						 I.type != 1 ? I.proton_list : -1);
		auto it = sub_of.find(key);
		if (it != sub_of.end()) {
			sub_off[i] = it->second;
			continue;
		}
		const int64_t size = I.type == 1 ? (int64_t)mvals[I.neutron_list].size()
				   : I.type == 2 ? (int64_t)mvals[I.proton_list].size()
						 : (int64_t)mvals[I.neutron_list].size() *
							   (int64_t)mvals[I.proton_list].size();
		sub_jobs.push_back({sub_total, i});
		sub_of[key] = sub_total;
		sub_off[i] = sub_total;
		sub_total += size;
	}
	W.resize(sub_total);
	parallel_for(sub_jobs.size(), [&](size_t q) {
		const int64_t i = sub_jobs[q].second;
		const bacchus_hip_instruction_t &I = p->instructions[i];
		const double *Vb = p->matrix_elements + p->matrix_offsets[I.matrix_block];
		double *dst = W.data() + sub_jobs[q].first;
		if (I.type == 1 || I.type == 2) {
			for (uint32_t m : mvals[I.type == 1 ? I.neutron_list : I.proton_list])
				*dst++ = Vb[m];
		} else {
			// normal: rows over the proton set, columns over the neutron
			// set; transposed space: the other way round.  The swap
			// decision depends on the two lists only, so every
			// instruction sharing the sub-block agrees with it.
			const auto &mn = mvals[I.neutron_list], &mp = mvals[I.proton_list];
			const int64_t Dn = vdn[I.matrix_block];
			for (size_t b = 0; b < mp.size(); b++)
				for (size_t a2 = 0; a2 < mn.size(); a2++) {
					const double v = Vb[mn[a2] + Dn * (int64_t)mp[b]];
					if (swapped[i])
						dst[a2 * mp.size() + b] = v;
					else
						dst[b * mn.size() + a2] = v;
				}
		}
	});

	clk.mark("transposed space, element sub-blocks");
	// --- CSR copies of every used list, by output and by input index,
	// each row ordered by element index; keys[k] = row of entry k.  Every
	// (list, direction) gets its place first; the copies are then built by
	// host threads into their own ranges ---
	std::vector<uint32_t> rowptr;
	std::vector<uint2> entries;
	std::vector<uint32_t> keys;
	std::vector<int64_t> rows_by_out(nl, -1), rows_by_in(nl, -1);
	struct CsrJob {
		int64_t l;
		int dir;
		int64_t rp_off;
		int64_t base;
	};
	std::vector<CsrJob> csr_jobs;
	int64_t used_entries = 0, rp_total = 0, ent_total = 0;
	for (int64_t l = 0; l < nl; l++) {
		if (!use[l].used)
			continue;
		const int64_t a = p->list_offsets[l], b = p->list_offsets[l + 1];
		used_entries += b - a;
		for (int dir = 0; dir < 2; dir++) {
			const int64_t nrows = dir == 0 ? use[l].n_out : use[l].n_in;
			(dir == 0 ? rows_by_out : rows_by_in)[l] = rp_total;
			csr_jobs.push_back(CsrJob{l, dir, rp_total, ent_total});
			rp_total += nrows + 1;
			ent_total += b - a;
		}
	}
	if (ent_total > 0xFFFFFFFFll || rp_total > 0xFFFFFFFFll)
		return fail(BACCHUS_FAIL("index lists exceed 2^32 entries"));
	rowptr.resize(rp_total);
	entries.resize(ent_total);
	keys.resize(ent_total);
	parallel_for(csr_jobs.size(), [&](size_t q) {
		const CsrJob &j = csr_jobs[q];
		const int64_t l = j.l;
		const int64_t a = p->list_offsets[l], b = p->list_offsets[l + 1];
		const int64_t nrows = j.dir == 0 ? use[l].n_out : use[l].n_in;
		const int key = j.dir == 0 ? 1 : 0;  // sort field: out or in
		std::vector<uint32_t> count(nrows + 1, 0);
		for (int64_t k = a; k < b; k++)
			count[p->list_entries[3 * k + key] + 1]++;
		const uint32_t base = (uint32_t)j.base;
		for (int64_t r = 0; r < nrows; r++)
			count[r + 1] += count[r];
		for (int64_t r = 0; r <= nrows; r++)
			rowptr[j.rp_off + r] = base + count[r];
		std::vector<uint32_t> fill(count.begin(), count.end() - 1);
		for (int64_t k = a; k < b; k++) {
			const int32_t *e3 = &p->list_entries[3 * k];
			const uint32_t r = (uint32_t)e3[key];
			uint2 en;
			en.x = (uint32_t)e3[1 - key];
			en.y = compact(l, e3[2]);
			keys[base + fill[r]] = r;
			entries[base + fill[r]++] = en;
		}
		for (int64_t r = 0; r < nrows; r++)
			std::sort(entries.begin() + base + count[r], entries.begin() + base + count[r + 1],
				  [](const uint2 &u, const uint2 &w) {
					  return (u.y & 0x7FFFFFFFu) < (w.y & 0x7FFFFFFFu);
				  });
	});
	op->list_entry_count = used_entries;
	// the lane-side list of every NP instruction (neutron list, or proton
	// list in the transposed space): a copy ordered by element index alone
	// (rows kept in keys[]), or for expanded lists by (input, element).
	// FP64 operators pad every input's run of an expanded list to an even
	// length with an entry whose element is zero, so entries 2k and 2k + 1
	// share their input and the kernel takes them as a pair (measured
	// faster in FP64 only; FP32/BF16 lose with pairs).  The copies are
	// sorted by host threads and appended in instruction order.
	const bool pair_lanes = op->precision == BACCHUS_HIP_FP64;
	std::vector<int64_t> elem_by_out(nl, -1), elem_by_in(nl, -1);
	std::vector<int64_t> elen_by_out(nl, 0), elen_by_in(nl, 0);
	std::vector<char> pad_entry(entries.size(), 0);
	struct LaneJob {
		int64_t l;
		int dir;
		std::vector<uint2> e;
		std::vector<uint32_t> k;
		std::vector<char> pad;
	};
	std::vector<LaneJob> lane_jobs;
	{
		std::vector<char> seen(nl, 0);
		for (int64_t i = 0; i < ni; i++) {
			const bacchus_hip_instruction_t &I = p->instructions[i];
			if (I.type != 3)
				continue;
			const int64_t l = swapped[i] ? I.proton_list : I.neutron_list;
			if (seen[l])
				continue;
			seen[l] = 1;
			for (int dir = 0; dir < 2; dir++) {
				lane_jobs.emplace_back();
				lane_jobs.back().l = l;
				lane_jobs.back().dir = dir;
			}
		}
	}
	parallel_for(lane_jobs.size(), [&](size_t q) {
		LaneJob &j = lane_jobs[q];
		const int64_t l = j.l;
		const int64_t r0 = (j.dir == 0 ? rows_by_out : rows_by_in)[l];
		const int64_t nrows = j.dir == 0 ? use[l].n_out : use[l].n_in;
		const uint32_t a = rowptr[r0], b = rowptr[r0 + nrows];
		std::vector<uint32_t> perm(b - a);
		std::iota(perm.begin(), perm.end(), a);
		if (expanded[l])
			std::stable_sort(perm.begin(), perm.end(), [&](uint32_t u, uint32_t w) {
				return std::make_pair(entries[u].x, entries[u].y & 0x7FFFFFFFu) <
				       std::make_pair(entries[w].x, entries[w].y & 0x7FFFFFFFu);
			});
		else
			std::stable_sort(perm.begin(), perm.end(), [&](uint32_t u, uint32_t w) {
				return (entries[u].y & 0x7FFFFFFFu) < (entries[w].y & 0x7FFFFFFFu);
			});
		j.e.reserve(perm.size());
		j.k.reserve(perm.size());
		j.pad.reserve(perm.size());
		for (size_t t = 0; t < perm.size(); t++) {
			const uint2 e = entries[perm[t]];
			const uint32_t r = keys[perm[t]];
			j.e.push_back(e);
			j.k.push_back(r);
			j.pad.push_back(0);
			const bool run_ends = t + 1 == perm.size() || entries[perm[t + 1]].x != e.x;
			if (pair_lanes && expanded[l] && run_ends && j.e.size() % 2) {
				j.e.push_back(make_uint2(e.x, 0u));
				j.k.push_back(r);
				j.pad.push_back(1);
			}
		}
	});
	{
		size_t total = entries.size();
		for (const LaneJob &j : lane_jobs)
			total += j.e.size() + 1;
		if (total > 0xFFFFFFFFull)
			return fail(BACCHUS_FAIL("index lists exceed 2^32 entries"));
		entries.reserve(total);
		keys.reserve(total);
		pad_entry.reserve(total);
		for (LaneJob &j : lane_jobs) {
			// paired copies start at an even entry, so a pair's two
			// entries are one aligned 16-byte load and their keys one
			// 8-byte load
			if (pair_lanes && expanded[j.l] && entries.size() % 2) {
				entries.push_back(make_uint2(0u, 0u));
				keys.push_back(0u);
				pad_entry.push_back(1);
			}
			const int64_t start = (int64_t)entries.size();
			(j.dir == 0 ? elem_by_out : elem_by_in)[j.l] = start;
			(j.dir == 0 ? elen_by_out : elen_by_in)[j.l] = (int64_t)j.e.size();
			entries.insert(entries.end(), j.e.begin(), j.e.end());
			keys.insert(keys.end(), j.k.begin(), j.k.end());
			pad_entry.insert(pad_entry.end(), j.pad.begin(), j.pad.end());
			std::vector<uint2>().swap(j.e);
			std::vector<uint32_t>().swap(j.k);
			std::vector<char>().swap(j.pad);
		}
	}

	// expanded element rows of an NP instruction in one direction: row =
	// compact index of the row-side list, column = position in the lane
	// copy of that direction (zero at the padding entries); with paired
	// lanes, rows start at even element offsets, so an entry pair is one
	// aligned 2-element load
	std::map<std::tuple<int32_t, int32_t, int32_t, int>, int64_t> exp_of;
	// (space is reserved here, in instruction order; the rows are filled
	// afterwards, in parallel, see fill_expanded)
	struct ExpJob {
		int64_t off;
		int64_t i;
		int dir;
	};
	std::vector<ExpJob> exp_jobs;
	auto expanded_rows = [&](const bacchus_hip_instruction_t &I, int64_t i, int dir) -> int64_t {
		const auto key = std::make_tuple(I.matrix_block, I.neutron_list, I.proton_list, dir);
		auto it = exp_of.find(key);
		if (it != exp_of.end())
			return it->second;
		const int32_t l = lane_list(I, i);
		const int32_t rl = swapped[i] ? I.neutron_list : I.proton_list;
		const int64_t len = (dir == 0 ? elen_by_out : elen_by_in)[l];
		if (pair_lanes && W.size() % 2)
			W.push_back(0.0);
		const int64_t off = (int64_t)W.size();
		W.resize(off + (int64_t)mvals[rl].size() * len);
		exp_jobs.push_back(ExpJob{off, i, dir});
		exp_of[key] = off;
		return off;
	};
	auto fill_expanded = [&](const ExpJob &job) {
		const bacchus_hip_instruction_t &I = p->instructions[job.i];
		const int32_t l = lane_list(I, job.i);
		const int32_t rl = swapped[job.i] ? I.neutron_list : I.proton_list;
		const int64_t base = (job.dir == 0 ? elem_by_out : elem_by_in)[l];
		const int64_t len = (job.dir == 0 ? elen_by_out : elen_by_in)[l];
		const auto &rows = mvals[rl];
		const auto &cols = mvals[l];
		const double *Vb = p->matrix_elements + p->matrix_offsets[I.matrix_block];
		const int64_t Dn = vdn[I.matrix_block];
		double *dst = W.data() + job.off;
		for (size_t r = 0; r < rows.size(); r++)
			for (int64_t k = 0; k < len; k++) {
				if (pad_entry[base + k]) {
					dst[r * len + k] = 0.0;
					continue;
				}
				const uint32_t m = cols[entries[base + k].y & 0x7FFFFFFFu];
				// element index m_n + Dn m_p
				dst[r * len + k] = swapped[job.i] ? Vb[rows[r] + Dn * (int64_t)m]
								  : Vb[m + Dn * (int64_t)rows[r]];
			}
	};

	// room for every expanded block at once (no regrowth copies of W)
	{
		std::map<std::tuple<int32_t, int32_t, int32_t, int>, int64_t> sizes;
		for (int64_t i = 0; i < ni; i++) {
			const bacchus_hip_instruction_t &I = p->instructions[i];
			if (I.type != 3 || list_len(I.neutron_list) == 0 || list_len(I.proton_list) == 0)
				continue;
			const int32_t l = lane_list(I, i);
			if (!expanded[l])
				continue;
			const int32_t rl = swapped[i] ? I.neutron_list : I.proton_list;
			for (int dir = 0; dir < (I.block_in == I.block_out ? 1 : 2); dir++)
				sizes[std::make_tuple(I.matrix_block, I.neutron_list, I.proton_list, dir)] =
					(int64_t)mvals[rl].size() * (dir == 0 ? elen_by_out : elen_by_in)[l] + 1;
		}
		int64_t extra = 0;
		for (const auto &kv : sizes)
			extra += kv.second;
		W.reserve(W.size() + (size_t)extra);
	}

	clk.mark("CSR and lane copies");
	// --- virtual instructions (each direction) grouped by output block;
	// slots [0, nb) are the normal space, [nb, 2 nb) the transposed one ---
	std::vector<std::vector<TInstr>> per_block(2 * nb);
	for (int64_t i = 0; i < ni; i++) {
		const bacchus_hip_instruction_t &I = p->instructions[i];
		const int force = p->matrix_force ? p->matrix_force[I.matrix_block] : 2;
		const int64_t ln = I.neutron_list >= 0 ? p->list_offsets[I.neutron_list + 1] -
								 p->list_offsets[I.neutron_list] : 0;
		const int64_t lp = I.proton_list >= 0 ? p->list_offsets[I.proton_list + 1] -
								p->list_offsets[I.proton_list] : 0;
		if ((I.type != 2 && ln == 0) || (I.type != 1 && lp == 0))
			continue;
		for (int dir = 0; dir < (I.block_in == I.block_out ? 1 : 2); dir++) {
			const int64_t bin = dir == 0 ? I.block_in : I.block_out;
			const int64_t bout = dir == 0 ? I.block_out : I.block_in;
			const auto &rows = dir == 0 ? rows_by_out : rows_by_in;
			TInstr t{};
			t.x_off = op->block_off[bin];
			t.v_off = sub_off[i];
			t.force = (int16_t)std::min(3, std::max(1, force));
			const auto &elem = dir == 0 ? elem_by_out : elem_by_in;
			const bool expd = I.type == 3 && expanded[lane_list(I, i)];
			if (expd) {
				t.v_off = expanded_rows(I, i, dir);
				t.expanded = pair_lanes ? 2 : 1;
			}
			if (!swapped[i]) {
				t.dn = I.type == 3 ? (int64_t)mvals[I.neutron_list].size() : 0;
				t.nn_in = (int32_t)op->block_nn[bin];
				t.type = (int16_t)I.type;
				if (I.type == 3) {
					t.n_base = (uint32_t)elem[I.neutron_list];
					t.n_len = (uint32_t)(dir == 0 ? elen_by_out : elen_by_in)[I.neutron_list];
					if (expd)
						t.dn = (int64_t)t.n_len;
				} else if (I.type == 1)
					t.n_base = (uint32_t)rows[I.neutron_list];
				if (I.type != 1)
					t.p_rows = (uint32_t)rows[I.proton_list];
				per_block[bout].push_back(t);
			} else {
				// transposed space: protons fastest, the roles of the
				// two lists exchanged
				t.nn_in = (int32_t)op->block_np[bin];
				if (I.type == 3) {
					t.type = 3;
					t.n_base = (uint32_t)elem[I.proton_list];
					t.n_len = (uint32_t)(dir == 0 ? elen_by_out : elen_by_in)[I.proton_list];
					t.dn = expd ? (int64_t)t.n_len : (int64_t)mvals[I.proton_list].size();
					t.p_rows = (uint32_t)rows[I.neutron_list];
				} else {
					// an N instruction there walks its list over the
					// rows with the protons as contiguous columns, the
					// form of a P instruction in the normal space
					t.type = 2;
					t.p_rows = (uint32_t)rows[I.neutron_list];
				}
				per_block[nb + bout].push_back(t);
			}
		}
	}

	// expanded element rows: disjoint ranges of W, filled by host threads
	// (largest first, taken from a shared counter) from read-only tables
	{
		std::vector<int64_t> job_size(exp_jobs.size());
		std::vector<size_t> order(exp_jobs.size());
		for (size_t k = 0; k < exp_jobs.size(); k++) {
			const ExpJob &j = exp_jobs[k];
			const bacchus_hip_instruction_t &I = p->instructions[j.i];
			const int32_t l = lane_list(I, j.i);
			const int32_t rl = swapped[j.i] ? I.neutron_list : I.proton_list;
			job_size[k] = (int64_t)mvals[rl].size() *
				      (j.dir == 0 ? elen_by_out : elen_by_in)[l];
			order[k] = k;
		}
		std::sort(order.begin(), order.end(),
			  [&](size_t a, size_t b) { return job_size[a] > job_size[b]; });
		parallel_for(order.size(), [&](size_t q) { fill_expanded(exp_jobs[order[q]]); });
	}
	op->compact_elements = (int64_t)W.size();

	clk.mark("virtual instructions, expanded rows");
	// --- NP virtual instructions of one output block that share their
	// lane copy, force class and input block are merged into one: their
	// row-side entries, by output row, in a new CSR whose entries carry the
	// element row directly (the instruction's first element + compact
	// index x row stride, with the entry's phase in bit 31).  A lane's sum
	// for an output then runs over the row entries of every merged
	// instruction before it goes to the tile, which takes that many fewer
	// LDS adds.  Merged instructions have v_off 0 and dn 1, so the kernels
	// read them unchanged ---
	{
		for (int64_t slot = 0; slot < 2 * nb; slot++) {
			auto &vs = per_block[slot];
			const int64_t nrows = slot < nb ? op->block_np[slot] : op->block_nn[slot - nb];
			std::map<std::tuple<uint32_t, uint32_t, int, int64_t, int32_t, int32_t>,
				 std::vector<size_t>> groups;
			for (size_t k = 0; k < vs.size(); k++) {
				const TInstr &t = vs[k];
				if (t.type != 3)
					continue;
				groups[std::make_tuple(t.n_base, t.n_len, (int)t.force, t.x_off, t.nn_in,
						       t.expanded)].push_back(k);
			}
			std::vector<char> drop(vs.size(), 0);
			std::vector<TInstr> merged;
			for (const auto &g : groups) {
				const std::vector<size_t> &ks = g.second;
				if (ks.size() < 2)
					continue;
				bool fits = true;
				for (size_t k : ks) {
					const TInstr &t = vs[k];
					const uint32_t a = rowptr[t.p_rows], b = rowptr[t.p_rows + nrows];
					uint32_t bmax = 0;
					for (uint32_t e = a; e < b; e++)
						bmax = std::max(bmax, entries[e].y & kIndexMask);
					if (t.v_off + (int64_t)bmax * t.dn >= (int64_t)kIndexMask)
						fits = false;
				}
				if (!fits)
					continue;
				const uint32_t base = (uint32_t)rowptr.size();
				if (rowptr.size() + nrows + 1 > 0xFFFFFFFFull)
					return fail(BACCHUS_FAIL("index lists exceed 2^32 entries"));
				for (int64_t r = 0; r < nrows; r++) {
					rowptr.push_back((uint32_t)entries.size());
					for (size_t k : ks) {
						const TInstr &t = vs[k];
						for (uint32_t e = rowptr[t.p_rows + r]; e < rowptr[t.p_rows + r + 1]; e++) {
							const uint2 en = entries[e];
							const uint32_t row = (uint32_t)(t.v_off + (int64_t)(en.y & kIndexMask) * t.dn);
							entries.push_back(make_uint2(en.x, row | (en.y & ~kIndexMask)));
							keys.push_back((uint32_t)r);
							pad_entry.push_back(0);
						}
					}
				}
				rowptr.push_back((uint32_t)entries.size());
				if (entries.size() > 0xFFFFFFFFull)
					return fail(BACCHUS_FAIL("index lists exceed 2^32 entries"));
				TInstr m = vs[ks[0]];
				m.v_off = 0;
				m.dn = 1;
				m.p_rows = base;
				merged.push_back(m);
				for (size_t k : ks)
					drop[k] = 1;
			}
			if (merged.empty())
				continue;
			std::vector<TInstr> kept;
			for (size_t k = 0; k < vs.size(); k++)
				if (!drop[k])
					kept.push_back(vs[k]);
			kept.insert(kept.end(), merged.begin(), merged.end());
			vs.swap(kept);
		}
	}

	clk.mark("merged row-side lists");
	// --- host copies for the planner; virtual instructions of a block
	// sorted by element sub-block so neighbouring tasks share elements ---
	op->blk_vi.assign(2 * nb + 1, 0);
	op->blk_np.assign(2 * nb, 0);
	for (int64_t b = 0; b < 2 * nb; b++) {
		auto &vs = per_block[b];
		// N and P first, NP last (np_kernel can take those), each by
		// element sub-block
		std::stable_sort(vs.begin(), vs.end(), [](const TInstr &u, const TInstr &w) {
			return std::make_pair(u.type == 3, u.v_off) < std::make_pair(w.type == 3, w.v_off);
		});
		op->blk_np[b] = (int32_t)(op->h_tinstr.size() +
					  std::count_if(vs.begin(), vs.end(),
							[](const TInstr &t) { return t.type != 3; }));
		op->h_tinstr.insert(op->h_tinstr.end(), vs.begin(), vs.end());
		op->blk_vi[b + 1] = (int32_t)op->h_tinstr.size();
		if (b >= nb && !vs.empty())
			op->has_transposed = true;
	}
	op->h_rowptr = rowptr;
	op->mults = mults;
	{
		int dev = 0;
		hipDeviceProp_t prop;
		if (hipGetDevice(&dev) == hipSuccess &&
		    hipGetDeviceProperties(&prop, dev) == hipSuccess)
			op->num_cu = prop.multiProcessorCount;
	}
	// NP instructions in np_kernel (element-row-major units over whole
	// output blocks) with BACCHUS_HIP_NP=1; by default they stay in the tile
	// kernel, which measured faster (4.7 against 15.9 ms on the 3+3 Nmax 6
	// FP64 matvec: the element-row-major form leaves most lane slots empty,
	// each element row serving 1-4 row entries)
	{
		const char *e = getenv("BACCHUS_HIP_NP");
		if (e && atoi(e) != 0) {
			std::vector<int64_t> n_in(nl, 0), n_out(nl, 0);
			for (int64_t l = 0; l < nl; l++) {
				n_in[l] = use[l].n_in;
				n_out[l] = use[l].n_out;
			}
			NpContext nc{p, &mvals, &swapped, &rowptr, &entries, &rows_by_out, &rows_by_in,
				     &n_in, &n_out};
			if (int rc = np_build(op, nc))
				return fail(rc);
		}
	}
	// ~64 tiles per CU: small tiles keep every CU busy to the end
	PlanParams pp;
	pp.target = std::max(2.0e4, mults / (op->num_cu * 64.0));
	if (const char *tm = getenv("BACCHUS_HIP_TILE_MULTS"))
		pp.target = std::max(1.0e3, atof(tm));
	pp.bt.assign(2 * nb, 1);
	pp.xcds = kXcds;
	pp.skip_np = op->np_path;
	if (const char *xe = getenv("BACCHUS_HIP_XCDS"))
		pp.xcds = std::max(1, std::min(64, atoi(xe)));
	// tiles grouped by the input blocks they read, eight groups per XCD
	// (3+3 Nmax 6 FP64 matvec: 4.3 ms grouped by output block, 3.80 ms
	// with one input group per XCD, 3.39 ms with 64; 16, 32 and 48 groups
	// 3.62, 3.48, 3.45 ms; 96 and more 3.51 ms)
	pp.in_groups = 8 * kXcds;
	if (const char *ge = getenv("BACCHUS_HIP_IN_GROUPS"))
		pp.in_groups = std::max(1, std::min(4096, atoi(ge)));
	// LDS doubles per tile (measurement; the kernel's LDS follows the plan)
	if (const char *te = getenv("BACCHUS_HIP_TILE_ELEMS"))
		pp.tile_elems = std::max<int64_t>(1024, std::min<int64_t>(16384, atoll(te)));
	Plan plan;
	if (int rc = make_plan(op, pp, plan))
		return fail(rc);
	std::vector<TileWork> &tiles = plan.tiles;
	std::vector<Task> &tasks = plan.tasks;
	const std::vector<TInstr> &tinstr = op->h_tinstr;
	op->num_tasks = (int64_t)tasks.size();
	if (getenv("BACCHUS_HIP_VERBOSE")) {
		std::vector<double> tc = plan.cost;
		std::sort(tc.begin(), tc.end());
		double sum = 0;
		for (double c : tc)
			sum += c;
		if (!tc.empty())
			fprintf(stderr, "tiles %zu tasks %zu target %.3g sum %.4g mean %.3g p50 %.3g p90 %.3g p99 %.3g max %.3g\n",
				tc.size(), tasks.size(), pp.target, sum, sum / tc.size(), tc[tc.size() / 2],
				tc[tc.size() * 9 / 10], tc[tc.size() * 99 / 100], tc.back());
	}
	op->num_tiles = (int64_t)tiles.size();
	op->num_tinstr = (int64_t)tinstr.size();
	op->tile_lds_bytes = (size_t)plan.max_tile_elems * sizeof(double);

	clk.mark("plan");
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

	clk.mark("statistics");
	// --- upload ---
	auto upload = [&](void **dst, const void *src, size_t bytes) -> int {
		if (bytes == 0)
			bytes = 16;
		BACCHUS_HIP_CHECK(hipMalloc(dst, bytes));
		op->device_bytes += bytes;
		if (src)
			BACCHUS_HIP_CHECK(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
		return 0;
	};
	const int64_t total_v = (int64_t)W.size();
	int rc = 0;
	if (precision == BACCHUS_HIP_FP64) {
		rc = upload(&op->d_v, W.data(), total_v * sizeof(double));
	} else if (precision == BACCHUS_HIP_FP32) {
		std::vector<float, NoInitAlloc<float>> tmp(total_v);
		const int64_t chunk = 1 << 22;
		parallel_for((size_t)((total_v + chunk - 1) / chunk), [&](size_t c) {
			const int64_t k1 = std::min(total_v, ((int64_t)c + 1) * chunk);
			for (int64_t k = (int64_t)c * chunk; k < k1; k++)
				tmp[k] = (float)W[k];
		});
		rc = upload(&op->d_v, tmp.data(), total_v * sizeof(float));
	} else {
		std::vector<uint16_t, NoInitAlloc<uint16_t>> tmp(total_v);
		const int64_t chunk = 1 << 22;
		parallel_for((size_t)((total_v + chunk - 1) / chunk), [&](size_t c) {
			const int64_t k1 = std::min(total_v, ((int64_t)c + 1) * chunk);
			for (int64_t k = (int64_t)c * chunk; k < k1; k++)
				tmp[k] = to_bf16(W[k]);
		});
		rc = upload(&op->d_v, tmp.data(), total_v * sizeof(uint16_t));
	}
	if (rc)
		return fail(rc);
	if ((rc = upload((void **)&op->d_rowptr, rowptr.data(), rowptr.size() * 4)) ||
	    (rc = upload((void **)&op->d_entries, entries.data(), entries.size() * 8)) ||
	    (rc = upload((void **)&op->d_keys, keys.data(), keys.size() * 4)) ||
	    (rc = upload((void **)&op->d_tiles, tiles.data(), tiles.size() * sizeof(TileWork))) ||
	    (rc = upload((void **)&op->d_tasks, tasks.data(), tasks.size() * sizeof(Task))) ||
	    (rc = upload((void **)&op->d_tinstr, tinstr.data(), tinstr.size() * sizeof(TInstr))))
		return fail(rc);
	if (precision != BACCHUS_HIP_FP64 &&
	    (rc = upload((void **)&op->d_xf, nullptr, op->dimension * sizeof(float))))
		return fail(rc);
	{
		std::vector<int32_t> bnp(nb), bnn(nb);
		for (int64_t b = 0; b < nb; b++) {
			bnp[b] = (int32_t)op->block_np[b];
			bnn[b] = (int32_t)op->block_nn[b];
		}
		const size_t xs = precision == BACCHUS_HIP_FP64 ? sizeof(double) : sizeof(float);
		if ((rc = upload(&op->d_xT, nullptr, op->dimension * xs)) ||
		    (rc = upload((void **)&op->d_yT, nullptr, op->dimension * sizeof(double))) ||
		    (rc = upload((void **)&op->d_blk_off, op->block_off.data(), nb * sizeof(int64_t))) ||
		    (rc = upload((void **)&op->d_blk_np, bnp.data(), nb * sizeof(int32_t))) ||
		    (rc = upload((void **)&op->d_blk_nn, bnn.data(), nb * sizeof(int32_t))))
			return fail(rc);
	}
	clk.mark("upload");
	*out = guard.release();
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
	const dim3 tgrid((unsigned)op->block_np.size(), kTransposeSplits);
	if (op->has_transposed)
		hipLaunchKernelGGL(scale_kernel, dim3(blocks), dim3(threads), 0, s, op->d_yT, n, 0.0);
	if (op->precision == BACCHUS_HIP_FP64) {
		double *xT = (double *)op->d_xT;
		if (op->has_transposed)
			hipLaunchKernelGGL(transpose_blocks<double>, tgrid, dim3(threads), 0, s, x, xT,
					   op->d_blk_off, op->d_blk_np, op->d_blk_nn, 1);
		launch<double>(op, (const double *)op->d_v, x, xT, y, s1, s2, s3, s);
		if (int rc = np_launch(op, x, xT, y, s1, s2, s3, s))
			return rc;
	} else {
		float *xT = (float *)op->d_xT;
		hipLaunchKernelGGL(to_float_kernel, dim3(blocks), dim3(threads), 0, s, x, op->d_xf, n);
		if (op->has_transposed)
			hipLaunchKernelGGL(transpose_blocks<float>, tgrid, dim3(threads), 0, s, op->d_xf, xT,
					   op->d_blk_off, op->d_blk_np, op->d_blk_nn, 1);
		if (op->precision == BACCHUS_HIP_FP32)
			launch<float>(op, (const float *)op->d_v, op->d_xf, xT, y, s1, s2, s3, s);
		else
			launch<uint16_t>(op, (const uint16_t *)op->d_v, op->d_xf, xT, y, s1, s2, s3, s);
		if (int rc = np_launch(op, op->d_xf, xT, y, s1, s2, s3, s))
			return rc;
	}
	if (op->has_transposed)
		hipLaunchKernelGGL(add_transposed_blocks<double>, tgrid, dim3(threads), 0, s, y,
				   op->d_yT, op->d_blk_off, op->d_blk_np, op->d_blk_nn, 1);
	BACCHUS_HIP_CHECK(hipGetLastError());
	return 0;
}

}  // extern "C"
