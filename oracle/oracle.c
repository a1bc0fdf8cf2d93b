/*
 * oracle.c - CPU restatement of the Bacchus hot path (Minerva matvec and
 * Bacchus Lanczos).  TEST INFRASTRUCTURE ONLY: the checker for the HIP
 * path, never part of it.  See oracle.h for the reference lines followed.
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* matrix_vector_multiplication.c:7-13 */
static int retrive_phase(int32_t value)
{
	return (value & (int32_t)0x80000000) ? -1 : 1;
}

/* matrix_vector_multiplication.c:15-18 */
static int32_t strip_phase(int32_t value)
{
	return value & 0x7FFFFFFF;
}

/* multiplication_neutrons, matrix_vector_multiplication.c:20-77 */
static void mult_n(double *out, const double *in, int64_t Np, int64_t Nn_in,
		   int64_t Nn_out, const double *V, double scale,
		   const int32_t *L, int64_t len)
{
	for (int64_t i = 0; i < len; i++) {
		const int64_t n_in = L[3 * i], n_out = L[3 * i + 1];
		const int sign = retrive_phase(L[3 * i + 2]);
		const double me = V[strip_phase(L[3 * i + 2])] * scale;
		if (fabs(me) < 1e-12)
			continue;
		for (int64_t p = 0; p < Np; p++)
			out[n_out + Nn_out * p] += sign * me * in[n_in + Nn_in * p];
	}
}

/* multiplication_protons, matrix_vector_multiplication.c:79-132 */
static void mult_p(double *out, const double *in, int64_t Nn,
		   const double *V, double scale, const int32_t *L,
		   int64_t len)
{
	for (int64_t i = 0; i < len; i++) {
		const int64_t p_in = L[3 * i], p_out = L[3 * i + 1];
		const int sign = retrive_phase(L[3 * i + 2]);
		const double me = V[strip_phase(L[3 * i + 2])] * scale;
		if (fabs(me) < 1e-12)
			continue;
		for (int64_t n = 0; n < Nn; n++)
			out[Nn * p_out + n] += sign * me * in[Nn * p_in + n];
	}
}

/* multiplication_neutrons_protons, matrix_vector_multiplication.c:134-210
 * (no 1e-12 skip here, as in the reference) */
static void mult_np(double *out, const double *in, int64_t Nn_in,
		    int64_t Nn_out, const double *V, int64_t Dn, double scale,
		    const int32_t *Ln, int64_t len_n, const int32_t *Lp,
		    int64_t len_p)
{
	for (int64_t i = 0; i < len_n; i++) {
		const int64_t n_in = Ln[3 * i], n_out = Ln[3 * i + 1];
		const int32_t mn = Ln[3 * i + 2];
		for (int64_t j = 0; j < len_p; j++) {
			const int64_t p_in = Lp[3 * j], p_out = Lp[3 * j + 1];
			const int32_t mp = Lp[3 * j + 2];
			const int64_t in_index = n_in + Nn_in * p_in;
			const int64_t out_index = n_out + Nn_out * p_out;
			const int64_t m = strip_phase(mn) + Dn * strip_phase(mp);
			const int sign = retrive_phase(mn) * retrive_phase(mp);
			out[out_index] += sign * (V[m] * scale) * in[in_index];
		}
	}
}

/* The *_off_diag kernels (matrix_vector_multiplication.c:212-462) apply the
 * same element to (out <- in) and to the transpose (in <- out) in one loop;
 * left = y[out] += H x[in], right = y[in] += H^T x[out]
 * (scheduler.c:310-418). */
static void mult_n_off(double *out_l, double *out_r, const double *in_l,
		       const double *in_r, int64_t Np, int64_t Nn_in,
		       int64_t Nn_out, const double *V, double scale,
		       const int32_t *L, int64_t len)
{
	for (int64_t i = 0; i < len; i++) {
		const int64_t n_in = L[3 * i], n_out = L[3 * i + 1];
		const int sign = retrive_phase(L[3 * i + 2]);
		const double me = V[strip_phase(L[3 * i + 2])] * scale;
		if (fabs(me) < 1e-12)
			continue;
		for (int64_t p = 0; p < Np; p++) {
			const int64_t o = n_out + Nn_out * p, q = n_in + Nn_in * p;
			out_l[o] += sign * me * in_l[q];
			out_r[q] += sign * me * in_r[o];
		}
	}
}

static void mult_p_off(double *out_l, double *out_r, const double *in_l,
		       const double *in_r, int64_t Nn, const double *V,
		       double scale, const int32_t *L, int64_t len)
{
	for (int64_t i = 0; i < len; i++) {
		const int64_t p_in = L[3 * i], p_out = L[3 * i + 1];
		const int sign = retrive_phase(L[3 * i + 2]);
		const double me = V[strip_phase(L[3 * i + 2])] * scale;
		if (fabs(me) < 1e-12)
			continue;
		for (int64_t n = 0; n < Nn; n++) {
			const int64_t o = Nn * p_out + n, q = Nn * p_in + n;
			out_l[o] += sign * me * in_l[q];
			out_r[q] += sign * me * in_r[o];
		}
	}
}

static void mult_np_off(double *out_l, double *out_r, const double *in_l,
			const double *in_r, int64_t Nn_in, int64_t Nn_out,
			const double *V, int64_t Dn, double scale,
			const int32_t *Ln, int64_t len_n, const int32_t *Lp,
			int64_t len_p)
{
	for (int64_t i = 0; i < len_n; i++) {
		const int64_t n_in = Ln[3 * i], n_out = Ln[3 * i + 1];
		const int32_t mn = Ln[3 * i + 2];
		for (int64_t j = 0; j < len_p; j++) {
			const int64_t p_in = Lp[3 * j], p_out = Lp[3 * j + 1];
			const int32_t mp = Lp[3 * j + 2];
			const int64_t q = n_in + Nn_in * p_in;
			const int64_t o = n_out + Nn_out * p_out;
			const int64_t m = strip_phase(mn) + Dn * strip_phase(mp);
			const int sign = retrive_phase(mn) * retrive_phase(mp);
			const double me = V[m] * scale;
			out_l[o] += sign * me * in_l[q];
			out_r[q] += sign * me * in_r[o];
		}
	}
}

int oracle_matvec(int64_t num_blocks, const int64_t *block_np,
		  const int64_t *block_nn, const int64_t *list_offsets,
		  const int32_t *list_entries, const int64_t *matrix_offsets,
		  const int64_t *matrix_dn, const int32_t *matrix_force,
		  const double *scales, const double *matrix_elements,
		  int64_t num_instructions, const int32_t *instructions,
		  const double *x, double *y, int num_threads)
{
	int64_t *off = (int64_t *)malloc((num_blocks + 1) * sizeof(int64_t));
	if (!off)
		return 1;
	off[0] = 0;
	for (int64_t b = 0; b < num_blocks; b++)
		off[b + 1] = off[b] + block_np[b] * block_nn[b];
	const int64_t D = off[num_blocks];
	if (num_threads < 1)
		num_threads = 1;
	/* one output replica per thread (vector_block.c:47-67), reduced
	 * into y at the end (vector_block.c:146-155) */
	double *rep = (double *)calloc((size_t)num_threads * (size_t)D, sizeof(double));
	if (!rep) {
		free(off);
		return 1;
	}
#pragma omp parallel num_threads(num_threads)
	{
		double *yt = rep + (size_t)omp_get_thread_num() * (size_t)D;
#pragma omp for schedule(dynamic, 1)
		for (int64_t k = 0; k < num_instructions; k++) {
			const int32_t *I = instructions + 8 * k;
			const int32_t type = I[0], bi = I[2], bo = I[3];
			const int32_t ln = I[4], lp = I[5], v = I[6];
			const double scale =
				scales ? scales[(matrix_force ? matrix_force[v] : 2) - 1] : 1.0;
			const double *V = matrix_elements + matrix_offsets[v];
			const int32_t *Ln = ln >= 0 ? list_entries + 3 * list_offsets[ln] : NULL;
			const int32_t *Lp = lp >= 0 ? list_entries + 3 * list_offsets[lp] : NULL;
			const int64_t len_n = ln >= 0 ? list_offsets[ln + 1] - list_offsets[ln] : 0;
			const int64_t len_p = lp >= 0 ? list_offsets[lp + 1] - list_offsets[lp] : 0;
			if (bi == bo) {
				if (type == 1)
					mult_n(yt + off[bo], x + off[bi], block_np[bi],
					       block_nn[bi], block_nn[bo], V, scale, Ln, len_n);
				else if (type == 2)
					mult_p(yt + off[bo], x + off[bi], block_nn[bi], V,
					       scale, Lp, len_p);
				else
					mult_np(yt + off[bo], x + off[bi], block_nn[bi],
						block_nn[bo], V, matrix_dn[v], scale, Ln, len_n,
						Lp, len_p);
			} else {
				if (type == 1)
					mult_n_off(yt + off[bo], yt + off[bi], x + off[bi],
						   x + off[bo], block_np[bi], block_nn[bi],
						   block_nn[bo], V, scale, Ln, len_n);
				else if (type == 2)
					mult_p_off(yt + off[bo], yt + off[bi], x + off[bi],
						   x + off[bo], block_nn[bi], V, scale, Lp, len_p);
				else
					mult_np_off(yt + off[bo], yt + off[bi], x + off[bi],
						    x + off[bo], block_nn[bi], block_nn[bo], V,
						    matrix_dn[v], scale, Ln, len_n, Lp, len_p);
			}
		}
	}
	for (int t = 0; t < num_threads; t++)
		for (int64_t i = 0; i < D; i++)
			y[i] += rep[(size_t)t * (size_t)D + i];
	free(rep);
	free(off);
	return 0;
}

/* Number of eigenvalues of T below x (Sturm sequence of the LDL^T
 * factorization of T - x I). */
static int64_t sturm_count(int64_t n, const double *d, const double *e,
			   double x)
{
	int64_t count = 0;
	double q = d[0] - x;
	if (q < 0)
		count++;
	for (int64_t i = 1; i < n; i++) {
		if (q == 0.0)
			q = 1e-300;
		q = d[i] - x - e[i - 1] * e[i - 1] / q;
		if (q < 0)
			count++;
	}
	return count;
}

void oracle_tridiagonal_eigenvalues(int64_t n, const double *d,
				    const double *e, int64_t count,
				    double *out)
{
	double lo = d[0], hi = d[0];
	for (int64_t i = 0; i < n; i++) {
		const double r = (i > 0 ? fabs(e[i - 1]) : 0) + (i + 1 < n ? fabs(e[i]) : 0);
		if (d[i] - r < lo)
			lo = d[i] - r;
		if (d[i] + r > hi)
			hi = d[i] + r;
	}
	const double span = fmax(fabs(lo), fabs(hi));
	lo -= 1e-12 * span + 1e-300;
	hi += 1e-12 * span + 1e-300;
	if (count > n)
		count = n;
	for (int64_t k = 0; k < count; k++) {
		double a = lo, b = hi;
		for (int it = 0; it < 200; it++) {
			const double mid = 0.5 * (a + b);
			if (mid <= a || mid >= b)
				break;
			if (sturm_count(n, d, e, mid) > k)
				b = mid;
			else
				a = mid;
		}
		out[k] = 0.5 * (a + b);
	}
}

static double dot(const double *a, const double *b, int64_t n)
{
	double s = 0.0;
	for (int64_t i = 0; i < n; i++)
		s += a[i] * b[i];
	return s;
}

/* lanczos.c:60-253; vector ops vector.c:240-476, math_tools.c:13-57 */
int64_t oracle_lanczos(int64_t D, oracle_matvec_fn matvec, void *ctx,
		       int64_t max_num_iterations, int start,
		       const double *start_vector, int reorth, int convergence,
		       double tolerance, int64_t num_eigenvalues,
		       double *eigenvalues, double *alpha, double *beta)
{
	const int64_t K = D < max_num_iterations ? D : max_num_iterations;
	double *U = (double *)calloc((size_t)(K + 1) * (size_t)D, sizeof(double));
	double *c = (double *)calloc((size_t)(K + 1), sizeof(double));
	double *ev = (double *)calloc((size_t)(K + 1), sizeof(double));
	if (!U || !c || !ev) {
		free(U);
		free(c);
		free(ev);
		return -1;
	}
	/* lanczos.c:60-72: u_0 = e_0 */
	if (start == ORACLE_START_GIVEN && start_vector)
		memcpy(U, start_vector, (size_t)D * sizeof(double));
	else
		U[0] = 1.0;
	double previous = 0.0;
	int64_t it;
	for (it = 0; it < K; it++) {
		double *u = U + (size_t)it * D, *w = U + (size_t)(it + 1) * D;
		matvec(ctx, u, w); /* lanczos.c:98-100, w starts at zero */
		const double a = dot(u, w, D); /* :108-111 */
		alpha[it] = a;
		if (it == 0) { /* :117-121 */
			for (int64_t i = 0; i < D; i++)
				w[i] -= a * u[i];
		} else { /* :122-135, vector.c:324-380 */
			const double bp = beta[it - 1];
			const double *up = U + (size_t)(it - 1) * D;
			for (int64_t i = 0; i < D; i++) {
				w[i] -= a * u[i];
				w[i] -= bp * up[i];
			}
		}
		if (reorth == ORACLE_REORTH_FULL) {
			/* classical Gram-Schmidt twice against u_0..u_it */
			for (int pass = 0; pass < 2; pass++) {
				for (int64_t j = 0; j <= it; j++)
					c[j] = dot(U + (size_t)j * D, w, D);
				for (int64_t j = 0; j <= it; j++) {
					const double *uj = U + (size_t)j * D;
					for (int64_t i = 0; i < D; i++)
						w[i] -= c[j] * uj[i];
				}
			}
		}
		const double b = sqrt(dot(w, w, D)); /* :137-138, vector.c:382-404 */
		beta[it] = b;
		const double inv = 1.0 / b; /* :141, vector.c:437-459 */
		for (int64_t i = 0; i < D; i++)
			w[i] *= inv;
		/* lanczos.c:155-169, vector.c:461-476: MGS of u_{it+1} against
		 * u_0..u_{it-3}, no renormalization */
		if (reorth == ORACLE_REORTH_REFERENCE && it > 2) {
			for (int64_t j = 0; j < it - 2; j++) {
				const double *uj = U + (size_t)j * D;
				const double p = dot(w, uj, D);
				for (int64_t i = 0; i < D; i++)
					w[i] -= p * uj[i];
			}
		}
		/* :192-243, target eigenvalue 0 */
		oracle_tridiagonal_eigenvalues(it + 1, alpha, beta, 1, ev);
		const double diff = convergence == ORACLE_CONVERGE_EIGENVALUES
					    ? fabs(ev[0] - previous)
					    : 2 * tolerance;
		if (diff < tolerance) {
			it++;
			break;
		}
		previous = ev[0];
	}
	const int64_t dim = it;
	const int64_t n = num_eigenvalues < dim ? num_eigenvalues : dim;
	oracle_tridiagonal_eigenvalues(dim, alpha, beta, n, eigenvalues);
	free(U);
	free(c);
	free(ev);
	return dim;
}
