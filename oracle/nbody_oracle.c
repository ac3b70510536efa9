/*
 * nbody_oracle.c -- CPU restatement of the drons/nbody force path (test
 * infrastructure; see nbody_oracle.h for who may use it and how it is pinned).
 *
 * Compiled with -O2 -ffp-contract=off so that every operation below rounds
 * exactly where it is written.  Each function cites the reference lines it
 * restates (paths relative to the reference repository root).
 */
#include "nbody_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* force law: nbody/nbody_data.cpp:35-44 (G = 1, nbody_data.cpp:15)          */
/* ------------------------------------------------------------------------ */
void oracle_force(const double v1[3], const double v2[3], double m1, double m2, double out[3])
{
	double dx = v1[0] - v2[0];
	double dy = v1[1] - v2[1];
	double dz = v1[2] - v2[2];
	double r2 = dx * dx + dy * dy + dz * dz;      /* vertex3::norm, vertex.h:151-154 */
	if(r2 < ORACLE_MIN_DISTANCE)
	{
		r2 = ORACLE_MIN_DISTANCE;
	}
	double coeff = ((-1.0 * m1) * m2) / (r2 * sqrt(r2));
	out[0] = dx * coeff;
	out[1] = dy * coeff;
	out[2] = dz * coeff;
}

/* nbody/nbody_data.cpp:46-55: pairs closer than sqrt(MinDistance) contribute 0 */
double oracle_potential_energy(const double v1[3], const double v2[3], double m1, double m2)
{
	double dx = v1[0] - v2[0];
	double dy = v1[1] - v2[1];
	double dz = v1[2] - v2[2];
	double r2 = dx * dx + dy * dy + dz * dz;
	if(r2 < ORACLE_MIN_DISTANCE)
	{
		return 0.0;
	}
	return -((1.0 * m1) * m2) / sqrt(r2);
}

/* ------------------------------------------------------------------------ */
/* direct fcompute                                                           */
/* ------------------------------------------------------------------------ */
static void accel_one(size_t n, const double* y, const double* mass, size_t body1, double acc[3])
{
	const double* rx = y;
	const double* ry = y + n;
	const double* rz = y + 2 * n;
	double v1[3] = { rx[body1], ry[body1], rz[body1] };
	double tot[3] = { 0.0, 0.0, 0.0 };
	for(size_t body2 = 0; body2 != n; ++body2)
	{
		if(body1 == body2)           /* nbody_engine_simple.cpp:158-161 */
		{
			continue;
		}
		double v2[3] = { rx[body2], ry[body2], rz[body2] };
		double fr[3];
		oracle_force(v1, v2, mass[body1], mass[body2], fr);
		tot[0] += fr[0];
		tot[1] += fr[1];
		tot[2] += fr[2];
	}
	acc[0] = tot[0] / mass[body1];   /* nbody_engine_simple.cpp:169-171 */
	acc[1] = tot[1] / mass[body1];
	acc[2] = tot[2] / mass[body1];
}

/* nbody/nbody_engine_simple.cpp:116-173 */
void oracle_fcompute_simple(size_t n, const double* y, const double* mass, double* f)
{
	for(size_t i = 0; i < n; ++i)
	{
		double acc[3];
		accel_one(n, y, mass, i, acc);
		f[i] = y[3 * n + i];
		f[n + i] = y[4 * n + i];
		f[2 * n + i] = y[5 * n + i];
		f[3 * n + i] = acc[0];
		f[4 * n + i] = acc[1];
		f[5 * n + i] = acc[2];
	}
}

/* nbody/nbody_engine_openmp.cpp:19-77 */
void oracle_fcompute_openmp(size_t n, const double* y, const double* mass, double* f, int threads)
{
	long long nn = (long long)n;
#ifdef _OPENMP
	#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : omp_get_max_threads())
#else
	(void)threads;
#endif
	for(long long ii = 0; ii < nn; ++ii)
	{
		size_t i = (size_t)ii;
		double acc[3];
		accel_one(n, y, mass, i, acc);
		f[i] = y[3 * n + i];
		f[n + i] = y[4 * n + i];
		f[2 * n + i] = y[5 * n + i];
		f[3 * n + i] = acc[0];
		f[4 * n + i] = acc[1];
		f[5 * n + i] = acc[2];
	}
}

void oracle_accel_targets(size_t n, const double* y, const double* mass,
                          const size_t* idx, size_t ntgt, double* acc, int threads)
{
	long long nt = (long long)ntgt;
#ifdef _OPENMP
	#pragma omp parallel for schedule(dynamic, 16) num_threads(threads > 0 ? threads : omp_get_max_threads())
#else
	(void)threads;
#endif
	for(long long t = 0; t < nt; ++t)
	{
		accel_one(n, y, mass, idx[t], acc + 3 * t);
	}
}

/* nbody/nbody_engine_block.cpp:14-126: targets and sources in tiles of 64
   bodies; per pair dr = r1 - r2, coeff = m2 / (sqrt(r2) * r2) on the clamped r2,
   total -= dr * coeff; f[3n:6n] = total (no division by m1).  The self pair is
   summed: its dr is +0, so it subtracts +0 and changes nothing.  The reference
   reads past the end unless n % 64 == 0; here every n works and n % 64 == 0
   gives the reference's sums (sources in index order for every target). */
void oracle_fcompute_block(size_t n, const double* y, const double* mass, double* f, int threads)
{
	const double* rx = y;
	const double* ry = y + n;
	const double* rz = y + 2 * n;
	long long nn = (long long)n;
#ifdef _OPENMP
	#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : omp_get_max_threads())
#else
	(void)threads;
#endif
	for(long long ii = 0; ii < nn; ++ii)
	{
		size_t i = (size_t)ii;
		double tx = 0.0, ty = 0.0, tz = 0.0;
		for(size_t j = 0; j < n; ++j)
		{
			double dx = rx[i] - rx[j];
			double dy = ry[i] - ry[j];
			double dz = rz[i] - rz[j];
			double r2 = dx * dx + dy * dy + dz * dz;
			if(r2 < ORACLE_MIN_DISTANCE)
			{
				r2 = ORACLE_MIN_DISTANCE;
			}
			double r = sqrt(r2);
			double coeff = mass[j] / (r * r2);
			tx -= dx * coeff;
			ty -= dy * coeff;
			tz -= dz * coeff;
		}
		f[i] = y[3 * n + i];
		f[n + i] = y[4 * n + i];
		f[2 * n + i] = y[5 * n + i];
		f[3 * n + i] = tx;
		f[4 * n + i] = ty;
		f[5 * n + i] = tz;
	}
}

/* ------------------------------------------------------------------------ */
/* Ahmad-Cohen neighbour scheme: nbody/nbody_engine_ah.cpp:16-160            */
/* ------------------------------------------------------------------------ */
/* Full pass (nbody_engine_ah.cpp:44-114): every pair as the simple engine;
   body2 is a neighbour of body1 when |r1 - r2|^2 < max_dist_sqr or
   |F|^2 > min_force_sqr (vertex3::norm is the squared length, vertex.h:151-154),
   and the other bodies' forces are summed into the frozen far force univ.
   Neighbour lists are stored in index order as CSR: offs[n + 1], list. */
int oracle_ah_full(oracle_ah* h, const double* y, const double* mass, double* f)
{
	const size_t n = h->n;
	const double* rx = y;
	const double* ry = y + n;
	const double* rz = y + 2 * n;
	size_t total = 0;
	for(int pass = 0; pass < 2; ++pass)
	{
		if(pass == 1)
		{
			free(h->list);
			h->list = (unsigned*)malloc((total > 0 ? total : 1) * sizeof(unsigned));
			if(h->list == NULL)
			{
				return -1;
			}
		}
		total = 0;
		for(size_t b1 = 0; b1 < n; ++b1)
		{
			const double v1[3] = { rx[b1], ry[b1], rz[b1] };
			double tu[3] = { 0.0, 0.0, 0.0 };
			double tf[3] = { 0.0, 0.0, 0.0 };
			h->offs[b1] = total;
			for(size_t b2 = 0; b2 != n; ++b2)
			{
				if(b1 == b2)
				{
					continue;
				}
				const double v2[3] = { rx[b2], ry[b2], rz[b2] };
				double fr[3];
				oracle_force(v1, v2, mass[b1], mass[b2], fr);
				const double dx = v1[0] - v2[0], dy = v1[1] - v2[1], dz = v1[2] - v2[2];
				const double d2 = dx * dx + dy * dy + dz * dz;
				const double f2 = fr[0] * fr[0] + fr[1] * fr[1] + fr[2] * fr[2];
				if(d2 < h->max_dist_sqr || f2 > h->min_force_sqr)
				{
					if(pass == 1)
					{
						h->list[total] = (unsigned)b2;
					}
					++total;
				}
				else
				{
					tu[0] += fr[0];
					tu[1] += fr[1];
					tu[2] += fr[2];
				}
				tf[0] += fr[0];
				tf[1] += fr[1];
				tf[2] += fr[2];
			}
			if(pass == 1)
			{
				f[b1] = y[3 * n + b1];
				f[n + b1] = y[4 * n + b1];
				f[2 * n + b1] = y[5 * n + b1];
				f[3 * n + b1] = tf[0] / mass[b1];
				f[4 * n + b1] = tf[1] / mass[b1];
				f[5 * n + b1] = tf[2] / mass[b1];
				h->univ[3 * b1] = tu[0];
				h->univ[3 * b1 + 1] = tu[1];
				h->univ[3 * b1 + 2] = tu[2];
			}
		}
		h->offs[n] = total;
	}
	h->total = total;
	return 0;
}

/* Sparse pass (nbody_engine_ah.cpp:116-160): the frozen far force plus the
   neighbours' forces at the current positions, in list order. */
void oracle_ah_sparse(const oracle_ah* h, const double* y, const double* mass, double* f)
{
	const size_t n = h->n;
	const double* rx = y;
	const double* ry = y + n;
	const double* rz = y + 2 * n;
	for(size_t b1 = 0; b1 < n; ++b1)
	{
		const double v1[3] = { rx[b1], ry[b1], rz[b1] };
		double tf[3] = { h->univ[3 * b1], h->univ[3 * b1 + 1], h->univ[3 * b1 + 2] };
		for(size_t k = h->offs[b1]; k < h->offs[b1 + 1]; ++k)
		{
			const size_t b2 = h->list[k];
			const double v2[3] = { rx[b2], ry[b2], rz[b2] };
			double fr[3];
			oracle_force(v1, v2, mass[b1], mass[b2], fr);
			tf[0] += fr[0];
			tf[1] += fr[1];
			tf[2] += fr[2];
		}
		f[b1] = y[3 * n + b1];
		f[n + b1] = y[4 * n + b1];
		f[2 * n + b1] = y[5 * n + b1];
		f[3 * n + b1] = tf[0] / mass[b1];
		f[4 * n + b1] = tf[1] / mass[b1];
		f[5 * n + b1] = tf[2] / mass[b1];
	}
}

int oracle_ah_alloc(oracle_ah* h, size_t n, double max_dist, double min_force)
{
	memset(h, 0, sizeof(*h));
	h->n = n;
	h->max_dist_sqr = max_dist * max_dist;     /* nbody_engine_ah.cpp:3-9 */
	h->min_force_sqr = min_force * min_force;
	h->offs = (size_t*)calloc(n + 1, sizeof(size_t));
	h->univ = (double*)calloc(3 * (n > 0 ? n : 1), sizeof(double));
	if(h->offs == NULL || h->univ == NULL)
	{
		oracle_ah_free(h);
		return -1;
	}
	return 0;
}

void oracle_ah_free(oracle_ah* h)
{
	free(h->offs);
	free(h->univ);
	free(h->list);
	memset(h, 0, sizeof(*h));
}

/* ------------------------------------------------------------------------ */
/* vector ops                                                                */
/* ------------------------------------------------------------------------ */
/* nbody_engine_simple.cpp:259-278 */
void oracle_fill(double* a, double v, size_t count)
{
	for(size_t i = 0; i < count; ++i)
	{
		a[i] = v;
	}
}

/* nbody_engine_simple.cpp:280-306: a[i] += b[i]*c */
void oracle_fmadd_inplace(double* a, const double* b, double c, size_t count)
{
	for(size_t i = 0; i < count; ++i)
	{
		a[i] = fma(b[i], c, a[i]);
	}
}

/* nbody_engine_simple.cpp:308-339: a[i] = b[i] + c[i]*d */
void oracle_fmadd(double* a, const double* b, const double* c, double d, size_t count)
{
	for(size_t i = 0; i < count; ++i)
	{
		a[i] = fma(c[i], d, b[i]);
	}
}

/* nbody/nbody_engine.cpp:47-66 */
void oracle_fmaddn_inplace(double* a, const double* const* b, const double* c, size_t csize, size_t count)
{
	if(c == NULL)
	{
		return;
	}
	for(size_t k = 0; k != csize; ++k)
	{
		if(c[k] == 0.0)
		{
			continue;
		}
		oracle_fmadd_inplace(a, b[k], c[k], count);
	}
}

/* nbody/nbody_engine.cpp:74-108 */
void oracle_fmaddn(double* a, const double* b, const double* const* c, const double* d,
                   size_t dsize, size_t count)
{
	if(d == NULL)
	{
		return;
	}
	int initiated = 0;
	if(b == NULL)
	{
		oracle_fill(a, 0.0, count);
		initiated = 1;
	}
	for(size_t k = 0; k != dsize; ++k)
	{
		if(d[k] == 0.0)
		{
			continue;
		}
		if(initiated)
		{
			oracle_fmadd(a, a, c[k], d[k], count);
		}
		else
		{
			oracle_fmadd(a, b, c[k], d[k], count);
		}
		initiated = 1;
	}
}

/* nbody/nbody_engine_simple.cpp:341-394 with summation_k (nbody/summation.h:7-14) */
void oracle_fmaddn_corr(double* a, double* corr, const double* const* b, const double* c,
                        size_t csize, size_t count)
{
	if(c == NULL)
	{
		return;
	}
	for(size_t i = 0; i < count; ++i)
	{
		for(size_t k = 0; k < csize; ++k)
		{
			if(c[k] == 0.0)
			{
				continue;
			}
			volatile double term = b[k][i] * c[k];
			double corrected = term - corr[i];
			double new_sum = a[i] + corrected;
			corr[i] = (new_sum - a[i]) - corrected;
			a[i] = new_sum;
		}
	}
}

/* nbody_engine_simple.cpp:396-419 */
double oracle_fmaxabs(const double* a, size_t count)
{
	if(count == 0)
	{
		return 0.0;
	}
	double result = fabs(a[0]);
	for(size_t i = 0; i < count; ++i)
	{
		double v = fabs(a[i]);
		if(v > result)
		{
			result = v;
		}
	}
	return result;
}

/* nbody_engine_simple.cpp:175-199 */
void oracle_clamp(double* y, size_t nbodies, double b)
{
	double diam = 2.0 * b;
	for(int d = 0; d < 3; ++d)
	{
		double* r = y + (size_t)d * nbodies;
		for(size_t i = 0; i < nbodies; ++i)
		{
			if(r[i] > +b) { r[i] -= diam; }
			if(r[i] < -b) { r[i] += diam; }
		}
	}
}

/* ------------------------------------------------------------------------ */
/* heap index algebra: nbody/nbody_space_heap_func_priv.h:4-72              */
/* ------------------------------------------------------------------------ */
size_t oracle_heap_left(size_t i) { return i << 1; }
size_t oracle_heap_right(size_t i) { return (i << 1) + 1; }
size_t oracle_heap_parent(size_t i) { return i >> 1; }

/* Strip the trailing run of 1 bits (the chain of "right child" steps), then
   step to the right sibling.  Written with a loop, not ctz, so it is an
   independent statement of the same map. */
size_t oracle_heap_next_down(size_t i)
{
	while(i & 1u)
	{
		i >>= 1;
	}
	return i + 1;
}

size_t oracle_heap_skip_idx(size_t i) { return oracle_heap_next_down(i); }

size_t oracle_heap_next_up(size_t i, size_t tree_size)
{
	size_t l = oracle_heap_left(i);
	if(l < tree_size)
	{
		return l;
	}
	return oracle_heap_next_down(i);
}

/* ------------------------------------------------------------------------ */
/* heap kd-tree                                                              */
/* ------------------------------------------------------------------------ */
int oracle_heap_alloc(oracle_heap* h, size_t count)
{
	memset(h, 0, sizeof(*h));
	if(count == 0 || (count & (count - 1)) != 0)
	{
		return -1;
	}
	h->count = count;
	h->size = 2 * count;    /* 2*count - 1 + NBODY_HEAP_ROOT_INDEX, nbody_space_heap.cpp:23 */
	h->cm = (double*)calloc(h->size * 3, sizeof(double));
	h->mass = (double*)calloc(h->size, sizeof(double));
	h->crit_r2 = (double*)calloc(h->size, sizeof(double));
	h->bmin = (double*)calloc(h->size * 3, sizeof(double));
	h->bmax = (double*)calloc(h->size * 3, sizeof(double));
	h->body_n = (size_t*)malloc(h->size * sizeof(size_t));
	if(!h->cm || !h->mass || !h->crit_r2 || !h->bmin || !h->bmax || !h->body_n)
	{
		oracle_heap_free(h);
		return -1;
	}
	return 0;
}

void oracle_heap_free(oracle_heap* h)
{
	free(h->cm);
	free(h->mass);
	free(h->crit_r2);
	free(h->bmin);
	free(h->bmax);
	free(h->body_n);
	memset(h, 0, sizeof(*h));
}

/* std::min / std::max as the reference calls them (left operand wins ties) */
static double min_ref(double a, double b) { return (b < a) ? b : a; }
static double max_ref(double a, double b) { return (a < b) ? b : a; }

static double len3(double x, double y, double z)
{
	return sqrt(x * x + y * y + z * z);   /* vertex3::length, vertex.h:155-158 */
}

/* nbody/nbody_space_heap.cpp:119-134 */
static void heap_update(oracle_heap* h, size_t idx, size_t l, size_t r)
{
	double ml = h->mass[l];
	double mr = h->mass[r];
	double m = ml + mr;
	h->mass[idx] = m;
	for(int d = 0; d < 3; ++d)
	{
		h->cm[3 * idx + d] = (h->cm[3 * l + d] * ml + h->cm[3 * r + d] * mr) / m;
		h->bmin[3 * idx + d] = min_ref(h->bmin[3 * l + d], h->bmin[3 * r + d]);
		h->bmax[3 * idx + d] = max_ref(h->bmax[3 * l + d], h->bmax[3 * r + d]);
	}
	const double* lo = h->bmin + 3 * idx;
	const double* hi = h->bmax + 3 * idx;
	const double* c = h->cm + 3 * idx;
	double rad = len3(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]) * 0.5 +
	             len3((hi[0] + lo[0]) / 2.0 - c[0],
	                  (hi[1] + lo[1]) / 2.0 - c[1],
	                  (hi[2] + lo[2]) / 2.0 - c[2]);
	h->crit_r2[idx] = (rad * rad) * h->ratio_sqr;
}

/* Total order used by the median split: coordinate first, body index on ties.
   For distinct coordinates this is exactly std::nth_element's partition by
   `r[a] < r[b]` (nbody_space_heap.cpp:157-159); ties are made deterministic. */
static int key_less(const double* key, size_t a, size_t b)
{
	double ka = key[a];
	double kb = key[b];
	if(ka < kb) { return 1; }
	if(kb < ka) { return 0; }
	return a < b;
}

/* Rearrange ind[0..count) so that ind[0..k) hold the k smallest (by key_less). */
static void select_k(size_t* ind, size_t count, size_t k, const double* key)
{
	size_t lo = 0;
	size_t hi = count;            /* half-open working range */
	while(hi - lo > 1)
	{
		/* median of three pivot */
		size_t mid = lo + (hi - lo) / 2;
		size_t a = ind[lo], b = ind[mid], c = ind[hi - 1];
		size_t piv;
		if(key_less(key, a, b))
		{
			piv = key_less(key, b, c) ? b : (key_less(key, a, c) ? c : a);
		}
		else
		{
			piv = key_less(key, a, c) ? a : (key_less(key, b, c) ? c : b);
		}
		/* three-way split: [lo,lt) < piv, [lt,gt) == piv, [gt,hi) > piv */
		size_t lt = lo, i = lo, gt = hi;
		while(i < gt)
		{
			size_t v = ind[i];
			if(key_less(key, v, piv))
			{
				ind[i] = ind[lt]; ind[lt] = v; ++lt; ++i;
			}
			else if(key_less(key, piv, v))
			{
				--gt; ind[i] = ind[gt]; ind[gt] = v;
			}
			else
			{
				++i;
			}
		}
		if(k < lt) { hi = lt; }
		else if(k >= gt) { lo = gt; }
		else { return; }
	}
}

/* nbody/nbody_space_heap.cpp:136-191 (serial; the OpenMP tasks there do not
   change the result) */
static void heap_build_p(oracle_heap* h, size_t count, size_t* ind, const double* r[3],
                         const double* mass, size_t idx, int dim)
{
	if(count == 1)
	{
		size_t b = *ind;
		for(int d = 0; d < 3; ++d)
		{
			h->cm[3 * idx + d] = r[d][b];
			h->bmin[3 * idx + d] = r[d][b];
			h->bmax[3 * idx + d] = r[d][b];
		}
		h->mass[idx] = mass[b];
		h->body_n[idx] = b;
		return;
	}
	size_t left_size = count / 2;
	select_k(ind, count, left_size, r[dim]);
	int next = (dim + 1) % 3;
	heap_build_p(h, left_size, ind, r, mass, oracle_heap_left(idx), next);
	heap_build_p(h, count - left_size, ind + left_size, r, mass, oracle_heap_right(idx), next);
	heap_update(h, idx, oracle_heap_left(idx), oracle_heap_right(idx));
}

/* heap_build_p with the two halves of every segment above 64 bodies as OpenMP
   tasks, as nbody/nbody_space_heap.cpp:175-183 does; the subtrees are disjoint,
   so the result equals the serial build's.  Used to time the CPU baseline. */
static void heap_build_task(oracle_heap* h, size_t count, size_t* ind, const double* r[3],
                            const double* mass, size_t idx, int dim)
{
	if(count <= 64)
	{
		heap_build_p(h, count, ind, r, mass, idx, dim);
		return;
	}
	size_t left_size = count / 2;
	select_k(ind, count, left_size, r[dim]);
	int next = (dim + 1) % 3;
	#pragma omp task
	heap_build_task(h, left_size, ind, r, mass, oracle_heap_left(idx), next);
	#pragma omp task
	heap_build_task(h, count - left_size, ind + left_size, r, mass, oracle_heap_right(idx), next);
	#pragma omp taskwait
	heap_update(h, idx, oracle_heap_left(idx), oracle_heap_right(idx));
}

int oracle_heap_build_omp(oracle_heap* h, const double* rx, const double* ry, const double* rz,
                          const double* mass, double ratio, int threads)
{
	size_t n = h->count;
	size_t* ind = (size_t*)malloc(n * sizeof(size_t));
	if(ind == NULL)
	{
		return -1;
	}
	for(size_t i = 0; i < n; ++i)
	{
		ind[i] = i;
	}
	for(size_t i = 0; i < h->size; ++i)
	{
		h->body_n[i] = ORACLE_NO_BODY;
	}
	h->ratio_sqr = ratio * ratio;
	const double* r[3] = { rx, ry, rz };
	#pragma omp parallel num_threads(threads > 0 ? threads : omp_get_max_threads())
	#pragma omp single
	heap_build_task(h, n, ind, r, mass, ORACLE_HEAP_ROOT, 0);
	free(ind);
	return 0;
}

/* The Barnes-Hut accelerations (force / m) of the bodies idx[0..m) with the
   layout's walk (1 = heap stack walk, 2 = stackless), OpenMP over targets:
   the per-target work of oracle_bh_fcompute on a sample. */
void oracle_bh_accel_targets(const oracle_heap* h, int layout, size_t n, const double* y, const double* mass,
                             const size_t* idx, size_t m, double* out, int threads)
{
	long long mm = (long long)m;
	#pragma omp parallel for schedule(dynamic, 4) num_threads(threads > 0 ? threads : omp_get_max_threads())
	for(long long t = 0; t < mm; ++t)
	{
		size_t b = idx[t];
		double v1[3] = { y[b], y[n + b], y[2 * n + b] };
		double tot[3];
		if(layout == 2)
		{
			oracle_heap_traverse_stackless(h, v1, mass[b], tot);
		}
		else
		{
			oracle_heap_traverse(h, v1, mass[b], tot);
		}
		for(int d = 0; d < 3; ++d)
		{
			out[3 * t + d] = tot[d] / mass[b];
		}
	}
}

/* nbody/nbody_space_heap.cpp:13-37 */
int oracle_heap_build(oracle_heap* h, const double* rx, const double* ry, const double* rz,
                      const double* mass, double ratio)
{
	size_t n = h->count;
	size_t* ind = (size_t*)malloc(n * sizeof(size_t));
	if(ind == NULL)
	{
		return -1;
	}
	for(size_t i = 0; i < n; ++i)
	{
		ind[i] = i;
	}
	for(size_t i = 0; i < h->size; ++i)
	{
		h->body_n[i] = ORACLE_NO_BODY;
	}
	h->ratio_sqr = ratio * ratio;
	const double* r[3] = { rx, ry, rz };
	heap_build_p(h, n, ind, r, mass, ORACLE_HEAP_ROOT, 0);
	free(ind);
	return 0;
}

/* nbody/nbody_space_heap.cpp:39-62 (the reference's last pass also writes the
   unused slot 0; it is never read, so that write is not restated) */
void oracle_heap_rebuild(oracle_heap* h, const double* rx, const double* ry, const double* rz,
                         double ratio)
{
	size_t n = h->count;
	h->ratio_sqr = ratio * ratio;
	for(size_t idx = n; idx < 2 * n; ++idx)
	{
		size_t b = h->body_n[idx];
		double p[3] = { rx[b], ry[b], rz[b] };
		for(int d = 0; d < 3; ++d)
		{
			h->cm[3 * idx + d] = p[d];
			h->bmin[3 * idx + d] = p[d];
			h->bmax[3 * idx + d] = p[d];
		}
	}
	for(size_t level = n; level > 1; level /= 2)
	{
		for(size_t idx = level / 2; idx < level; ++idx)
		{
			heap_update(h, idx, oracle_heap_left(idx), oracle_heap_right(idx));
		}
	}
}

#define ORACLE_MAX_STACK 64   /* MAX_STACK_SIZE, nbody/nbody_space_tree.h:10 */

static double dist2_to(const oracle_heap* h, const double v1[3], size_t node)
{
	double dx = v1[0] - h->cm[3 * node + 0];
	double dy = v1[1] - h->cm[3 * node + 1];
	double dz = v1[2] - h->cm[3 * node + 2];
	return dx * dx + dy * dy + dz * dz;
}

/* nbody/nbody_space_heap.cpp:64-97 */
void oracle_heap_traverse(const oracle_heap* h, const double v1[3], double mass1, double out[3])
{
	double tot[3] = { 0.0, 0.0, 0.0 };
	size_t stack[ORACLE_MAX_STACK];
	size_t sp = 0;
	stack[sp++] = ORACLE_HEAP_ROOT;
	while(sp != 0)
	{
		size_t curr = stack[--sp];
		if(dist2_to(h, v1, curr) > h->crit_r2[curr])
		{
			double fr[3];
			oracle_force(v1, h->cm + 3 * curr, mass1, h->mass[curr], fr);
			tot[0] += fr[0];
			tot[1] += fr[1];
			tot[2] += fr[2];
		}
		else
		{
			size_t l = oracle_heap_left(curr);
			size_t r = oracle_heap_right(curr);
			if(r < h->size) { stack[sp++] = r; }
			if(l < h->size) { stack[sp++] = l; }
		}
	}
	out[0] = tot[0];
	out[1] = tot[1];
	out[2] = tot[2];
}

/* nbody/nbody_space_heap_stackless.cpp:3-26 */
void oracle_heap_traverse_stackless(const oracle_heap* h, const double v1[3], double mass1, double out[3])
{
	double tot[3] = { 0.0, 0.0, 0.0 };
	size_t curr = ORACLE_HEAP_ROOT;
	do
	{
		if(dist2_to(h, v1, curr) > h->crit_r2[curr])
		{
			double fr[3];
			oracle_force(v1, h->cm + 3 * curr, mass1, h->mass[curr], fr);
			tot[0] += fr[0];
			tot[1] += fr[1];
			tot[2] += fr[2];
			curr = oracle_heap_skip_idx(curr);
		}
		else
		{
			curr = oracle_heap_next_up(curr, h->size);
		}
	}
	while(curr != ORACLE_HEAP_ROOT);
	out[0] = tot[0];
	out[1] = tot[1];
	out[2] = tot[2];
}

/* nbody/nbody_engine_simple_bh.cpp:62-92 */
void oracle_bh_fcompute(const oracle_heap* h, int layout, int traverse, size_t n,
                        const double* y, const double* mass, double* f, int threads)
{
	long long nn = (long long)n;
#ifdef _OPENMP
	#pragma omp parallel for schedule(dynamic, 4) num_threads(threads > 0 ? threads : omp_get_max_threads())
#else
	(void)threads;
#endif
	for(long long t = 0; t < nn; ++t)
	{
		size_t body1;
		double v1[3];
		double m1;
		if(traverse == 0)        /* ett_cycle */
		{
			body1 = (size_t)t;
			v1[0] = y[body1];
			v1[1] = y[n + body1];
			v1[2] = y[2 * n + body1];
			m1 = mass[body1];
		}
		else                     /* ett_nested_tree: leaves in heap order */
		{
			size_t idx = n + (size_t)t;
			body1 = h->body_n[idx];
			v1[0] = h->cm[3 * idx + 0];
			v1[1] = h->cm[3 * idx + 1];
			v1[2] = h->cm[3 * idx + 2];
			m1 = h->mass[idx];
		}
		double tot[3];
		if(layout == 2)
		{
			oracle_heap_traverse_stackless(h, v1, m1, tot);
		}
		else
		{
			oracle_heap_traverse(h, v1, m1, tot);
		}
		f[body1] = y[3 * n + body1];
		f[n + body1] = y[4 * n + body1];
		f[2 * n + body1] = y[5 * n + body1];
		f[3 * n + body1] = tot[0] / mass[body1];
		f[4 * n + body1] = tot[1] / mass[body1];
		f[5 * n + body1] = tot[2] / mass[body1];
	}
}

/* ------------------------------------------------------------------------ */
/* conservation checks                                                       */
/* ------------------------------------------------------------------------ */
/* summation_k on a scalar (nbody/summation.h:7-14) */
static double kahan_step(double a, double b, double* corr)
{
	double corrected = b - *corr;
	double new_sum = a + corrected;
	*corr = (new_sum - a) - corrected;
	return new_sum;
}

/* nbody/nbody_data.cpp:57-102 with summation() over the proxies of
   nbody/summation_proxy.h (first element seeds the sum, summation.h:19-37) */
void oracle_stats(size_t n, const double* pos, const double* vel, const double* mass,
                  int want_energy, double out[12])
{
	memset(out, 0, 12 * sizeof(double));
	if(n == 0)
	{
		return;
	}
	/* P: impulce_proxy = vel * m */
	{
		double s[3], c[3] = { 0, 0, 0 };
		for(int d = 0; d < 3; ++d) { s[d] = vel[d] * mass[0]; }
		for(size_t i = 1; i < n; ++i)
			for(int d = 0; d < 3; ++d) { s[d] = kahan_step(s[d], vel[3 * i + d] * mass[i], &c[d]); }
		out[0] = s[0]; out[1] = s[1]; out[2] = s[2];
	}
	/* L: impulce_moment_proxy = pos ^ (vel * m) */
	{
		double s[3] = { 0, 0, 0 }, c[3] = { 0, 0, 0 };
		for(size_t i = 0; i < n; ++i)
		{
			const double* p = pos + 3 * i;
			double pv[3] = { vel[3 * i] * mass[i], vel[3 * i + 1] * mass[i], vel[3 * i + 2] * mass[i] };
			double l[3] = { p[1] * pv[2] - p[2] * pv[1],
			                p[2] * pv[0] - p[0] * pv[2],
			                p[0] * pv[1] - p[1] * pv[0] };
			for(int d = 0; d < 3; ++d)
			{
				s[d] = (i == 0) ? l[d] : kahan_step(s[d], l[d], &c[d]);
			}
		}
		out[3] = s[0]; out[4] = s[1]; out[5] = s[2];
	}
	if(want_energy)
	{
		/* KE: kinetic_energy_proxy = |v|^2 * m */
		double ke = 0.0, kc = 0.0;
		for(size_t i = 0; i < n; ++i)
		{
			const double* v = vel + 3 * i;
			double t = (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * mass[i];
			ke = (i == 0) ? t : kahan_step(ke, t, &kc);
		}
		/* PE: potential_energy_proxy over the flat n*n index (summation_proxy.h:68-87) */
		double pe = 0.0, pc = 0.0;
		for(size_t n1 = 0; n1 < n; ++n1)
		{
			for(size_t n2 = 0; n2 < n; ++n2)
			{
				double t = (n1 == n2) ? 0.0 :
				           oracle_potential_energy(pos + 3 * n1, pos + 3 * n2, mass[n1], mass[n2]);
				pe = (n1 == 0 && n2 == 0) ? t : kahan_step(pe, t, &pc);
			}
		}
		out[6] = ke / 2.0;
		out[7] = pe / 2.0;
	}
	/* V: total mass and mass_center_proxy = pos * m */
	{
		double tm = mass[0], tc = 0.0;
		for(size_t i = 1; i < n; ++i) { tm = kahan_step(tm, mass[i], &tc); }
		double s[3], c[3] = { 0, 0, 0 };
		for(int d = 0; d < 3; ++d) { s[d] = pos[d] * mass[0]; }
		for(size_t i = 1; i < n; ++i)
			for(int d = 0; d < 3; ++d) { s[d] = kahan_step(s[d], pos[3 * i + d] * mass[i], &c[d]); }
		out[8] = s[0] / tm; out[9] = s[1] / tm; out[10] = s[2] / tm;
		out[11] = tm;
	}
}

/* ------------------------------------------------------------------------
 * Morton radix tree (nbh_tree_build_radix's definition; no reference
 * counterpart).  Karras, "Maximizing parallelism in the construction of BVHs,
 * octrees, and k-d trees" (HPG 2012), restated sequentially.
 * ------------------------------------------------------------------------ */
static unsigned long long radix_spread(unsigned long long v)
{
	v &= 0x1fffffull;
	v = (v | (v << 32)) & 0x1f00000000ffffull;
	v = (v | (v << 16)) & 0x1f0000ff0000ffull;
	v = (v | (v << 8)) & 0x100f00f00f00f00full;
	v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
	v = (v | (v << 2)) & 0x1249249249249249ull;
	return v;
}

static unsigned long long radix_cell(double x, double lo, double scale)
{
	double c = floor((x - lo) * scale);
	if(c < 0.0) { c = 0.0; }
	if(c > 2097151.0) { c = 2097151.0; }
	return (unsigned long long)c;
}

typedef struct radix_key {
	unsigned long long code;
	unsigned body;
} radix_key;

static int radix_key_cmp(const void* a, const void* b)
{
	const radix_key* x = (const radix_key*)a;
	const radix_key* y = (const radix_key*)b;
	if(x->code != y->code) { return x->code < y->code ? -1 : 1; }
	return x->body < y->body ? -1 : (x->body > y->body ? 1 : 0);
}

/* common prefix length of sorted keys i and j, the index as tie-breaking suffix */
static int radix_delta(const radix_key* k, long long n, long long i, long long j)
{
	if(j < 0 || j >= n) { return -1; }
	if(k[i].code == k[j].code)
	{
		return 64 + __builtin_clzll((unsigned long long)(i ^ j));
	}
	return __builtin_clzll(k[i].code ^ k[j].code);
}

static double radix_min(double a, double b) { return (b < a) ? b : a; }
static double radix_max(double a, double b) { return (a < b) ? b : a; }

/* internal id `id` from its children (the heap tree's update of a node) */
static void radix_node(size_t id, size_t l, size_t r, double ratio_sqr, double* cm, double* nmass, double* crit,
                       double* bmin, double* bmax, double* quad)
{
	const double ml = nmass[l], mr = nmass[r];
	const double m = ml + mr;
	double c[3], lo[3], hi[3];
	for(int d = 0; d < 3; ++d)
	{
		c[d] = (cm[3 * l + d] * ml + cm[3 * r + d] * mr) / m;
		lo[d] = radix_min(bmin[3 * l + d], bmin[3 * r + d]);
		hi[d] = radix_max(bmax[3 * l + d], bmax[3 * r + d]);
		bmin[3 * id + d] = lo[d];
		bmax[3 * id + d] = hi[d];
	}
	const double ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
	const double gx = (hi[0] + lo[0]) / 2.0 - c[0];
	const double gy = (hi[1] + lo[1]) / 2.0 - c[1];
	const double gz = (hi[2] + lo[2]) / 2.0 - c[2];
	const double rad = sqrt(ex * ex + ey * ey + ez * ez) * 0.5 + sqrt(gx * gx + gy * gy + gz * gz);
	nmass[id] = m;
	for(int d = 0; d < 3; ++d) { cm[3 * id + d] = c[d]; }
	crit[id] = (rad * rad) * ratio_sqr;
	double q[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
	for(int side = 0; side < 2; ++side)
	{
		const size_t ch = side == 0 ? l : r;
		const double mc = nmass[ch];
		const double dx = cm[3 * ch] - c[0], dy = cm[3 * ch + 1] - c[1], dz = cm[3 * ch + 2] - c[2];
		q[0] += quad[6 * ch + 0] + mc * dx * dx;
		q[1] += quad[6 * ch + 1] + mc * dy * dy;
		q[2] += quad[6 * ch + 2] + mc * dz * dz;
		q[3] += quad[6 * ch + 3] + mc * dx * dy;
		q[4] += quad[6 * ch + 4] + mc * dx * dz;
		q[5] += quad[6 * ch + 5] + mc * dy * dz;
	}
	for(int k = 0; k < 6; ++k) { quad[6 * id + k] = q[k]; }
}

int oracle_radix_build(size_t n, const double* rx, const double* ry, const double* rz, const double* mass,
                       double ratio, unsigned* order, int* child, int* parent, double* cm, double* nmass,
                       double* crit, double* bmin, double* bmax, double* quad)
{
	if(n < 2) { return -1; }
	radix_key* k = (radix_key*)malloc(n * sizeof(radix_key));
	size_t* stack = (size_t*)malloc(2 * n * sizeof(size_t));
	unsigned char* seen = (unsigned char*)calloc(2 * n, 1);
	if(k == NULL || stack == NULL || seen == NULL)
	{
		free(k); free(stack); free(seen);
		return -1;
	}
	double lo[3] = {rx[0], ry[0], rz[0]}, hi[3] = {rx[0], ry[0], rz[0]};
	for(size_t i = 1; i < n; ++i)
	{
		const double v[3] = {rx[i], ry[i], rz[i]};
		for(int d = 0; d < 3; ++d)
		{
			lo[d] = fmin(lo[d], v[d]);
			hi[d] = fmax(hi[d], v[d]);
		}
	}
	const double side = fmax(fmax(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
	const double scale = side > 0.0 ? 2097152.0 / side : 0.0;
	for(size_t i = 0; i < n; ++i)
	{
		k[i].code = (radix_spread(radix_cell(rx[i], lo[0], scale)) << 2) |
		            (radix_spread(radix_cell(ry[i], lo[1], scale)) << 1) | radix_spread(radix_cell(rz[i], lo[2], scale));
		k[i].body = (unsigned)i;
	}
	qsort(k, n, sizeof(radix_key), radix_key_cmp);
	memset(child, 0, 2 * n * sizeof(int));
	memset(parent, 0, 2 * n * sizeof(int));
	const long long nn = (long long)n;
	for(long long i = 0; i < nn - 1; ++i)
	{
		const int dir = radix_delta(k, nn, i, i + 1) - radix_delta(k, nn, i, i - 1) > 0 ? 1 : -1;
		const int dmin = radix_delta(k, nn, i, i - dir);
		long long lmax = 2;
		while(radix_delta(k, nn, i, i + lmax * dir) > dmin) { lmax *= 2; }
		long long l = 0;
		for(long long t = lmax / 2; t >= 1; t /= 2)
		{
			if(radix_delta(k, nn, i, i + (l + t) * dir) > dmin) { l += t; }
		}
		const long long j = i + l * dir;
		const int dnode = radix_delta(k, nn, i, j);
		/* binary search of the split: the last position whose key shares more
		   than dnode bits with key i */
		long long s = 0;
		for(long long t = (l + 1) / 2;; t = (t + 1) / 2)
		{
			if(s + t <= l && radix_delta(k, nn, i, i + (s + t) * dir) > dnode) { s += t; }
			if(t <= 1) { break; }
		}
		const long long gamma = i + s * dir + (dir < 0 ? -1 : 0);
		const long long first = i < j ? i : j, last = i < j ? j : i;
		const int left = first == gamma ? (int)(nn + gamma) : (int)(gamma + 1);
		const int right = last == gamma + 1 ? (int)(nn + gamma + 1) : (int)(gamma + 2);
		child[2 * (i + 1)] = left;
		child[2 * (i + 1) + 1] = right;
		parent[left] = (int)(i + 1);
		parent[right] = (int)(i + 1);
	}
	memset(quad, 0, 12 * n * sizeof(double));
	for(size_t j = 0; j < n; ++j)
	{
		const size_t id = n + j;
		const unsigned b = k[j].body;
		order[j] = b;
		const double v[3] = {rx[b], ry[b], rz[b]};
		for(int d = 0; d < 3; ++d)
		{
			cm[3 * id + d] = v[d];
			bmin[3 * id + d] = v[d];
			bmax[3 * id + d] = v[d];
		}
		nmass[id] = mass[b];
		crit[id] = 0.0;
	}
	/* post-order from the root: a node is computed once both children are */
	const double ratio_sqr = ratio * ratio;
	size_t sp = 0;
	stack[sp++] = 1;
	while(sp > 0)
	{
		const size_t id = stack[sp - 1];
		const size_t l = (size_t)child[2 * id], r = (size_t)child[2 * id + 1];
		if(!seen[id])
		{
			seen[id] = 1;
			if(r < n) { stack[sp++] = r; }
			if(l < n) { stack[sp++] = l; }
			continue;
		}
		--sp;
		radix_node(id, l, r, ratio_sqr, cm, nmass, crit, bmin, bmax, quad);
	}
	free(k);
	free(stack);
	free(seen);
	return 0;
}
