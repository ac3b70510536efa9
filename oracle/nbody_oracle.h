/*
 * nbody_oracle.h -- CPU restatement of the drons/nbody force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engines:
 * only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
 * load it.  The product (libnbody_hip.so, libnbody.so) never links or calls it.
 *
 * Arithmetic conventions (shared bit-for-bit with the FP64 HIP kernels):
 *   - everything in IEEE binary64, compiled with -ffp-contract=off;
 *   - the vector ops (fmadd / fmaddn chains) use an explicit fma() for each
 *     `a + b*c` term, which is what the reference's `-O3 -march=native` build
 *     contracts those statements to on an FMA host;
 *   - the force law, the kd-tree moments and the Kahan steps use plain, separately
 *     rounded operations in the order written in the reference sources.
 *
 * Pinning (what checks each function against something other than itself):
 *   oracle_force / oracle_fcompute_simple   -> 20 solver goldens at 1e-12
 *                                              (test/data/<name>.txt), closed-form
 *                                              2-body and symmetric cases,
 *                                              long-double brute force
 *   oracle_heap_*  (index algebra)          -> reference KATs
 *                                              (test/engine/test_nbody_engine.cpp:1107-1147),
 *                                              exhaustive pointer-DFS comparison
 *   oracle_heap_build                       -> definition by full sorting (tests/),
 *                                              subtree invariants, mass conservation
 *   oracle_heap_traverse*                   -> ratio 1e8 == direct sum in tree order,
 *                                              heap == stackless bit-exact
 *   oracle vector ops                       -> element-wise host expressions and the
 *                                              eps/2 Kahan known-answer test
 *   oracle_stats                            -> closed forms (2-body circular orbit)
 */
#ifndef NBODY_ORACLE_H
#define NBODY_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_MIN_DISTANCE 1e-8   /* nbody/nbtype.h:79, applied to r^2 */
#define ORACLE_HEAP_ROOT 1         /* nbody/nbody_space_heap_func.h:8 */
#define ORACLE_NO_BODY ((size_t)-1)

/* ---- force law (nbody/nbody_data.cpp:35-55) ---- */
void   oracle_force(const double v1[3], const double v2[3], double m1, double m2, double out[3]);
double oracle_potential_energy(const double v1[3], const double v2[3], double m1, double m2);

/* ---- direct fcompute, state y = [rx ry rz vx vy vz] of 6n doubles ---- */
/* nbody/nbody_engine_simple.cpp:116-173 (single thread, body2 in index order) */
void oracle_fcompute_simple(size_t n, const double* y, const double* mass, double* f);
/* nbody/nbody_engine_openmp.cpp:19-77 (same arithmetic, OpenMP over targets) */
void oracle_fcompute_openmp(size_t n, const double* y, const double* mass, double* f, int threads);
/* Same sum restricted to the targets listed in idx[0..ntgt): writes acc[3*t..3*t+2]
   (acceleration, i.e. F/m1) for target idx[t].  Used to price sampled targets. */
void oracle_accel_targets(size_t n, const double* y, const double* mass,
                          const size_t* idx, size_t ntgt, double* acc, int threads);

/* nbody/nbody_engine_block.cpp:14-126: a_i = -sum_j (r_i - r_j) m_j / (r r2),
   self pair included (it adds 0), no division by m_i. */
void oracle_fcompute_block(size_t n, const double* y, const double* mass, double* f, int threads);

/* ---- Ahmad-Cohen neighbour scheme (nbody/nbody_engine_ah.cpp:16-160) ---- */
typedef struct oracle_ah {
	size_t    n;
	double    max_dist_sqr;
	double    min_force_sqr;
	size_t*   offs;     /* [n + 1] CSR offsets of the neighbour lists */
	unsigned* list;     /* [total] neighbours of each body in index order */
	double*   univ;     /* [3n] frozen far force (F, not F/m) */
	size_t    total;
} oracle_ah;

int  oracle_ah_alloc(oracle_ah* h, size_t n, double max_dist, double min_force);
void oracle_ah_free(oracle_ah* h);
/* full recompute (fcompute_full, :44-114): f, neighbour lists and far forces */
int  oracle_ah_full(oracle_ah* h, const double* y, const double* mass, double* f);
/* between full recomputes (fcompute_sparse, :116-160) */
void oracle_ah_sparse(const oracle_ah* h, const double* y, const double* mass, double* f);

/* ---- engine vector ops over `count` doubles ---- */
void   oracle_fill(double* a, double v, size_t count);
void   oracle_fmadd_inplace(double* a, const double* b, double c, size_t count);        /* a += b*c */
void   oracle_fmadd(double* a, const double* b, const double* c, double d, size_t count); /* a = b + c*d */
/* Base-class decompositions (nbody/nbody_engine.cpp:47-113): zero coefficients skipped;
   fmaddn with b == NULL fills a with 0 first; with b != NULL and every d == 0, a is untouched. */
void   oracle_fmaddn_inplace(double* a, const double* const* b, const double* c, size_t csize, size_t count);
void   oracle_fmaddn(double* a, const double* b, const double* const* c, const double* d,
                     size_t dsize, size_t count);
/* Kahan per term, element-major (nbody/nbody_engine_simple.cpp:341-394, summation.h:7-14) */
void   oracle_fmaddn_corr(double* a, double* corr, const double* const* b, const double* c,
                          size_t csize, size_t count);
double oracle_fmaxabs(const double* a, size_t count);                     /* nbody_engine_simple.cpp:396-419 */
void   oracle_clamp(double* y, size_t nbodies, double b);                 /* nbody_engine_simple.cpp:175-199 */

/* ---- heap index algebra (nbody/nbody_space_heap_func_priv.h:4-72) ---- */
size_t oracle_heap_left(size_t i);
size_t oracle_heap_right(size_t i);
size_t oracle_heap_parent(size_t i);
size_t oracle_heap_next_down(size_t i);
size_t oracle_heap_skip_idx(size_t i);
size_t oracle_heap_next_up(size_t i, size_t tree_size);

/* ---- heap kd-tree (nbody/nbody_space_heap.cpp) ---- */
typedef struct oracle_heap {
	size_t  count;       /* bodies; must be a power of two */
	size_t  size;        /* heap_size = 2*count (index 0 unused) */
	double* cm;          /* [size*3] mass centre, xyz interleaved */
	double* mass;        /* [size] */
	double* crit_r2;     /* [size] radius_sqr; 0 on leaves */
	double* bmin;        /* [size*3] */
	double* bmax;        /* [size*3] */
	size_t* body_n;      /* [size]; ORACLE_NO_BODY on internal nodes */
	double  ratio_sqr;
} oracle_heap;

int  oracle_heap_alloc(oracle_heap* h, size_t count);
void oracle_heap_free(oracle_heap* h);
/* nbody_space_heap.cpp:13-37 + build_p :136-191 + update :119-134 */
int  oracle_heap_build(oracle_heap* h, const double* rx, const double* ry, const double* rz,
                       const double* mass, double ratio);
/* The same build with OpenMP tasks for the segments above 64 bodies
   (nbody/nbody_space_heap.cpp:175-183); identical result.  For timing. */
int  oracle_heap_build_omp(oracle_heap* h, const double* rx, const double* ry, const double* rz,
                           const double* mass, double ratio, int threads);
/* nbody_space_heap.cpp:39-62 (refit with the existing body_n) */
void oracle_heap_rebuild(oracle_heap* h, const double* rx, const double* ry, const double* rz,
                         double ratio);
/* Stack traversal, right pushed first => left subtree first (nbody_space_heap.cpp:64-97).
   Returns the total force (not divided by mass1). */
void oracle_heap_traverse(const oracle_heap* h, const double v1[3], double mass1, double out[3]);
/* Stackless preorder (nbody_space_heap_stackless.cpp:3-26). */
void oracle_heap_traverse_stackless(const oracle_heap* h, const double v1[3], double mass1, double out[3]);

/* Barnes-Hut fcompute (nbody/nbody_engine_simple_bh.cpp:16-93).
   layout: 1 = heap, 2 = heap_stackless.  traverse: 0 = cycle, 1 = nested_tree.
   The tree must already be built / refitted for y. */
void oracle_bh_fcompute(const oracle_heap* h, int layout, int traverse, size_t n,
                        const double* y, const double* mass, double* f, int threads);

/* Accelerations of the bodies idx[0..m) by the walk of `layout` (the
   per-target work of oracle_bh_fcompute), out[3m]; for sampled checks and the
   CPU baseline. */
void oracle_bh_accel_targets(const oracle_heap* h, int layout, size_t n, const double* y, const double* mass,
                             const size_t* idx, size_t m, double* out, int threads);

/* ---- Morton radix tree (no reference counterpart; the fast tree of the FP32
        group walks, nbh_tree_build_radix) ----
   Restated from its definition, sequentially: 63-bit Morton codes of the
   positions on a 2^21-cell grid per axis over the bounding cube (cell =
   floor((x - lo) * 2^21 / side), clamped to [0, 2^21 - 1]; side = the longest
   extent of the bounding box), bodies sorted by (code, body number), internal
   nodes by Karras' binary radix tree construction (node i covers a key range
   with i at one end, split at the highest differing bit), then moments bottom-up
   in post-order with the heap tree's formulas (children left then right).
   Ids as nbh_tree_build_radix: root 1, internal i -> id i + 1, sorted leaf j ->
   id n + j.  Arrays sized for ids 0..2n-1 (id 0 unused): order[n] sorted body
   numbers, child[2n] (child[2 id], child[2 id + 1] for internal ids),
   parent[2n] (0 for the root and id 0), cm[6n], mass[2n], crit[2n], bmin[6n],
   bmax[6n], quad[12n] (second central moments xx yy zz xy xz yz).  Returns 0,
   or -1 when n < 2 or memory runs out. */
int oracle_radix_build(size_t n, const double* rx, const double* ry, const double* rz, const double* mass,
                       double ratio, unsigned* order, int* child, int* parent, double* cm, double* nmass,
                       double* crit, double* bmin, double* bmax, double* quad);

/* ---- conservation checks (nbody/nbody_data.cpp:57-145, summation.h:19-45,
        summation_proxy.h:6-87) ----
   Bodies given in AoS form exactly as nbody_data holds them (pos[3n], vel[3n]).
   out: [0..2] P, [3..5] L, [6] KE (already /2), [7] PE (already /2),
        [8..10] mass centre (divided by total mass), [11] total mass. */
void oracle_stats(size_t n, const double* pos, const double* vel, const double* mass,
                  int want_energy, double out[12]);

#ifdef __cplusplus
}
#endif

#endif /* NBODY_ORACLE_H */
