/*
 * nbody_hip.h -- C ABI of the MI355X (gfx950) N-body kernels (libnbody_hip.so).
 *
 * Plain pointers and sizes only.  Device pointers come from nbh_malloc.  Every
 * call returns an nbh_status; on failure nbh_last_error() holds a message for
 * the calling thread.  All work of one context is ordered on that context's
 * single HIP stream; calls that hand a result to host memory (nbh_memcpy_d2h,
 * nbh_maxabs, nbh_stats, nbh_bh_get_tree) return only after it is complete.
 *
 * Each entry point names the reference interface it stands in for (paths are
 * relative to the drons/nbody repository root).  The reference's de-facto C
 * layer is the set of host launchers in nbody/nbody_engine_cuda_impl.h:19-115.
 */
#ifndef NBODY_HIP_H
#define NBODY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum nbh_status {
	NBH_OK = 0,
	NBH_ERR_INVALID = 1,      /* bad argument: NULL pointer, size, count, ... */
	NBH_ERR_HIP = 2,          /* a HIP runtime call failed */
	NBH_ERR_NOMEM = 3,        /* device allocation failed */
	NBH_ERR_UNSUPPORTED = 4   /* valid request this build does not support */
} nbh_status;

/* Arithmetic of the force kernels. */
typedef enum nbh_precision {
	NBH_FP64 = 0,   /* IEEE binary64, same operation order as the reference CPU engines */
	NBH_FP32 = 1,   /* FP32 interaction math on relative coordinates, compensated sums,
	                   FP64 state in and out */
	NBH_FP32_CORR = 2, /* direct sum only: NBH_FP32 plus an FP64 correction pass -- for each
	                   target, the sources within 32 places of it in the Morton order of the
	                   FP32 tiles (its nearest neighbours along the curve) are recomputed in
	                   FP64 and their FP32 terms, recomputed the tiles' way, taken out */
	NBH_FP64_BLOCK = 3 /* direct sum only: IEEE binary64 in the operation order of the
	                   reference's `block` engine (nbody/nbody_engine_block.cpp:14-126):
	                   a_i = -sum_j (r_i - r_j) m_j / (sqrt(r2) r2), self pair included,
	                   no division by m_i */
} nbh_precision;

/* Tree layouts of the Barnes-Hut engine (nbody/nbody_engine_simple_bh.h:13-20). */
typedef enum nbh_tree_layout {
	NBH_LAYOUT_HEAP = 1,            /* stack traversal, left subtree first */
	NBH_LAYOUT_HEAP_STACKLESS = 2   /* skip/next_up preorder */
} nbh_tree_layout;

typedef struct nbh_ctx nbh_ctx;     /* one device + one stream */
typedef struct nbh_tree nbh_tree;   /* device-resident heap kd-tree */
typedef struct nbh_mlp nbh_mlp;     /* correction network (parameters + Adam state) */
typedef struct nbh_snapshot nbh_snapshot;   /* staging for asynchronous state copies */
typedef struct nbh_ah nbh_ah;       /* Ahmad-Cohen neighbour lists + frozen far forces */

/* ---- context, errors ---------------------------------------------------- */
const char* nbh_last_error(void);
int nbh_version(void);
/* Number of visible HIP devices (0 when there is no GPU). */
nbh_status nbh_device_count(int* count);
/* Replaces nbody_engine_cuda::select_devices (nbody/nbody_engine_cuda.cpp:56-89)
   for a single device id. */
nbh_status nbh_ctx_create(int device, nbh_ctx** ctx);
nbh_status nbh_ctx_destroy(nbh_ctx* ctx);
nbh_status nbh_ctx_synchronize(nbh_ctx* ctx);
/* Device name and compute-unit count, for reports. */
nbh_status nbh_ctx_info(nbh_ctx* ctx, char* name, size_t name_len, int* cu_count);

/* ---- memory: nbody_engine::create_buffer/free_buffer/read_buffer/write_buffer/
        copy_buffer (nbody/nbody_engine.h:54-63; CUDA nbody_engine_cuda.cpp:267-355) ---- */
nbh_status nbh_malloc(nbh_ctx* ctx, size_t bytes, void** dptr);
nbh_status nbh_free(nbh_ctx* ctx, void* dptr);
nbh_status nbh_memcpy_h2d(nbh_ctx* ctx, void* dst, const void* src, size_t bytes);
nbh_status nbh_memcpy_d2h(nbh_ctx* ctx, void* dst, const void* src, size_t bytes);
nbh_status nbh_memcpy_d2d(nbh_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ---- asynchronous state snapshots for frame dumps ------------------------
   The reference dumps a frame by copying the state to the host before the next
   step (nbody/nbody_solver.cpp:93-103 through nbody_engine_cuda::get_data,
   nbody_engine_cuda.cpp:141-165).  A snapshot copies `bytes` from a device
   buffer to a device staging slot on the context's stream (so it sees the
   state of the work queued so far, and later work may overwrite the source),
   then from staging to pinned host memory on a second stream of default
   priority, overlapping what the context's stream runs next.  slots: 1 or 2
   independent slots.  nbh_snapshot_wait blocks until the host copy of `slot`
   is complete and returns its host address (valid until the next
   nbh_snapshot_begin of that slot or nbh_snapshot_destroy); it may be called
   from another host thread than nbh_snapshot_begin, but not concurrently with
   calls on the same slot. */
nbh_status nbh_snapshot_create(nbh_ctx* ctx, size_t bytes, int slots, nbh_snapshot** snap);
nbh_status nbh_snapshot_destroy(nbh_snapshot* snap);
nbh_status nbh_snapshot_begin(nbh_ctx* ctx, nbh_snapshot* snap, int slot, const void* src);
nbh_status nbh_snapshot_wait(nbh_ctx* ctx, nbh_snapshot* snap, int slot, const void** host);

/* ---- FP64 vector ops over n doubles ------------------------------------- */
/* fill_buffer (nbody_engine_cuda_impl.cu:716-721) */
nbh_status nbh_fill(nbh_ctx* ctx, double* a, double value, size_t n);
/* Fused Butcher-table combination, one pass over memory:
     acc = (init ? init[i] : 0);  for t in 0..k-1: acc = fma(src[t][i], coef[t], acc);  a[i] = acc
   Terms are applied in the order given; the caller drops zero coefficients.
   init may equal a; src[t] may equal a.  k may be 0 (then a = init or 0).
   Replaces the fmadd / fmadd_inplace chains of nbody_engine::fmaddn and
   fmaddn_inplace (nbody/nbody_engine.cpp:47-113; kernels kfmadd,
   kfmadd_inplace, nbody_engine_cuda_impl.cu:750-805). */
nbh_status nbh_fmaddn(nbh_ctx* ctx, double* a, const double* init,
                      const double* const* src, const double* coef, size_t k, size_t n);
/* Kahan-compensated a[i] += sum_t src[t][i]*coef[t] with per-element
   correction corr[i], terms in order (summation_k, nbody/summation.h:7-14).
   Replaces the per-term kfmadd_inplace_corr launches
   (nbody/nbody_engine_cuda.cpp:402-459, nbody_engine_cuda_impl.cu:766-787). */
nbh_status nbh_fmaddn_corr(nbh_ctx* ctx, double* a, double* corr,
                           const double* const* src, const double* coef, size_t k, size_t n);
/* max |a[i]| returned to host (fmaxabs, nbody_engine_cuda_impl.cu:808-817). n == 0 gives 0. */
nbh_status nbh_maxabs(nbh_ctx* ctx, const double* a, size_t n, double* result);
/* Periodic wrap of the first 3*nbodies entries of y by +-2b
   (kclamp_coord, nbody_engine_cuda_impl.cu:724-747). */
nbh_status nbh_clamp(nbh_ctx* ctx, double* y, size_t nbodies, double b);

/* ---- direct all-pairs force ---------------------------------------------
   y: 6n doubles [rx ry rz vx vy vz]; mass: n doubles; f: 6n doubles, written
   as [vx vy vz ax ay az].  Replaces fcompute_block + fcompute_xyz
   (nbody_engine_cuda_impl.cu:10-124; host nbody/nbody_engine_cuda.cpp:202-247).
   NBH_FP64 reproduces nbody_engine_simple::fcompute (nbody_engine_simple.cpp:116-173)
   operation for operation.  block_size: work-group size of the FP64 kernel
   (64..1024, multiple of 64), 0 = default. */
nbh_status nbh_fcompute_direct(nbh_ctx* ctx, const double* y, const double* mass, double* f,
                               size_t n, nbh_precision precision, int block_size);

/* nbh_fcompute_direct without any device allocation: no counterpart in the
   reference.  For callers that cannot allocate (code under HIP graph capture,
   tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
   the same result bits as nbh_fcompute_direct; the FP32 paths keep their
   device scratch (sort keys, Morton-ordered records, split partial sums) in
   one fixed 64 MiB array per device instead of the context's grown buffer, so
   n is limited to 131072 bodies (NBH_ERR_INVALID beyond, or when the scratch
   the device's CU count asks for exceeds the array).  The array is shared by
   every context on a device and keeps the last call's contents: calls on one
   device whose executions overlap (different streams or threads) race. */
nbh_status nbh_fcompute_direct_noalloc(nbh_ctx* ctx, const double* y, const double* mass, double* f,
                                       size_t n, nbh_precision precision, int block_size);

/* FP64 accelerations of the m bodies listed in idx (device array of indices < n):
   acc[3t..3t+2] for body idx[t], same operations and source order as
   nbody_engine_simple::fcompute.  Used to measure the error of the FP32 and
   Barnes-Hut paths on sampled targets where a full FP64 sum is too costly. */
nbh_status nbh_direct_sample(nbh_ctx* ctx, const double* y, const double* mass, size_t n,
                             const uint32_t* idx, size_t m, double* acc);

/* ---- Ahmad-Cohen neighbour scheme (nbody/nbody_engine_ah.cpp:16-160) -------
   State of one engine: per body the frozen far force and the neighbour list.
   The reference keeps them in std::vector members of nbody_engine_ah
   (nbody/nbody_engine_ah.h:13-17); here they live in HBM. */
nbh_status nbh_ah_create(nbh_ctx* ctx, size_t n, nbh_ah** ah);
nbh_status nbh_ah_destroy(nbh_ah* ah);
/* Full recompute, replacing nbody_engine_ah::fcompute_full (nbody_engine_ah.cpp:44-114):
   f = [v ; F/m] over all pairs in FP64 (nbody_data::force, body2 in index order),
   body2 listed as a neighbour of body1 when |r1 - r2|^2 < max_dist^2 or
   |F12|^2 > min_force^2, and the far force (sum of the other pairs' F) frozen.
   Reads one 8-byte total back to the host to size the lists (grown as needed). */
nbh_status nbh_ah_fcompute_full(nbh_ctx* ctx, nbh_ah* ah, const double* y, const double* mass, double* f,
                                double max_dist, double min_force);
/* Between full recomputes, replacing nbody_engine_ah::fcompute_sparse
   (nbody_engine_ah.cpp:116-160): f = [v ; (far + sum over neighbours of F)/m]. */
nbh_status nbh_ah_fcompute_sparse(nbh_ctx* ctx, nbh_ah* ah, const double* y, const double* mass, double* f);
/* The lists of the last full recompute, copied to host memory: *total entries;
   offs (n + 1 offsets), list (*total body numbers) and univ (far forces [3][n])
   are written when not NULL. */
nbh_status nbh_ah_lists(nbh_ctx* ctx, const nbh_ah* ah, size_t* total, uint64_t* offs, uint32_t* list,
                        double* univ);

/* ---- Barnes-Hut heap kd-tree --------------------------------------------
   nbh_tree_build needs n to be a power of two (the heap layout of
   nbody/nbody_space_heap.cpp:23 puts the leaves at [n, 2n)).  The tree is built on the GPU with the same
   median splits as nbody_space_heap::build (nbody_space_heap.cpp:13-37, 136-191). */
nbh_status nbh_tree_create(nbh_ctx* ctx, size_t n, nbh_tree** tree);
nbh_status nbh_tree_destroy(nbh_tree* tree);
/* Full build: body order per node, moments, boxes, crit_r2 = (ratio*r)^2. */
nbh_status nbh_tree_build(nbh_ctx* ctx, nbh_tree* tree, const double* y, const double* mass,
                          double ratio);
/* Refit with the existing body order (nbody_space_heap::rebuild, :39-62;
   kupdate_leaf_bh / kupdate_node_bh, nbody_engine_cuda_impl.cu:477-641); a
   Morton tree keeps its structure and recomputes its moments. */
nbh_status nbh_tree_refit(nbh_ctx* ctx, nbh_tree* tree, const double* y, double ratio);
/* Morton radix tree, no counterpart in the reference: the bodies sorted by
   63-bit Morton codes over their bounding cube (21 bits per axis, ties by body
   number), internal nodes by the binary radix tree construction of Karras
   (2012), moments, boxes and opening radii as nbh_tree_build computes them.
   Ids: root 1, internal nodes 1..n-1, the body at sorted position j is leaf
   n + j.  Its cells are compact where the heap's count-balanced nodes of a
   centrally concentrated state span from the core to the outskirts, so a walk
   takes far fewer of them at the same ratio.  Only the group walks
   (nbh_bh_fcompute_group, _quad, _local, _local2) run on it: their walk groups
   are the maximal subtrees of at most 128 bodies (one wave, two bodies per
   lane), the two-level walk's super-groups the maximal subtrees of at most 2048
   bodies that hold more than one walk group (a walk group in none walks in one
   level).  A walk step whose children would not fit on the walk stack takes
   the nodes as monopoles (counted in nbh_bh_walk_counters out[2]).  Any n >= 2
   builds and walks.  The level counts are read back to the host once per build
   (through pinned memory), to size the launches that follow. */
nbh_status nbh_tree_build_radix(nbh_ctx* ctx, nbh_tree* tree, const double* y, const double* mass,
                                double ratio);
/* Lean Morton builds (lean != 0; no counterpart in the reference): the next
   nbh_tree_build_radix / nbh_tree_refit of this tree write the walk records of
   every node but its FP64 node arrays (mass centres, masses, boxes, second
   moments) only for the nodes the build reads back: the roots of its 64-body
   build groups, the nodes above them and one-body groups.  The group walks give
   the same results bit for bit; nbh_tree_get then returns valid entries for
   those nodes only.  About 0.86 GB less written per build at 4M bodies.
   Default off. */
nbh_status nbh_tree_set_radix_lean(nbh_ctx* ctx, nbh_tree* tree, int lean);
/* Children of internal ids (child[2k], child[2k + 1] for id k = 1..n-1; entry 0
   unused) and parents (parent[id] for ids 2..2n-1) of the last Morton tree;
   either pointer may be NULL. */
nbh_status nbh_tree_get_children(nbh_ctx* ctx, const nbh_tree* tree, int32_t* child, int32_t* parent);
/* Traversal: f[6n] = [v ; F/m] for every body, targets taken in heap leaf
   order.  FP64 reproduces nbody_space_heap::traverse (layout HEAP) and
   nbody_space_heap_stackless::traverse (layout HEAP_STACKLESS) with the
   nested_tree target order of nbody_engine_simple_bh.cpp:78-92.
   Replaces kfcompute_heap_bh / kfcompute_heap_bh_stackless (+ _tex),
   nbody_engine_cuda_impl.cu:130-451. */
nbh_status nbh_bh_fcompute(nbh_ctx* ctx, const nbh_tree* tree, const double* y,
                           const double* mass, double* f, nbh_tree_layout layout,
                           nbh_precision precision);
/* Group walk (FP32 interaction math, FP64 state): the 64 bodies under each heap
   node of level log2(n/64) share one walk; a cell is taken when the distance
   from the group's box to its mass centre exceeds its opening radius, so each
   body sees a refinement of the cells its own walk would take (error <= that
   of nbh_bh_fcompute at the same ratio).  Replaces the per-thread walk of
   kfcompute_heap_bh_stackless (nbody_engine_cuda_impl.cu:372-451) on the fast
   path; n >= 64 on the heap tree (on a Morton tree the groups are its walk
   groups, nbh_tree_build_radix). */
nbh_status nbh_bh_fcompute_group(nbh_ctx* ctx, const nbh_tree* tree, const double* y, double* f);
/* nbh_bh_fcompute_group plus the quadrupole term of every internal cell taken as
   a monopole (second central moments of the cell about its mass centre, kept
   per node by nbh_tree_build / nbh_tree_refit); FP32.  A cell whose mass centre
   lies within near_k half-diagonals of the group's box centre gets the term per
   body; a farther one gets its quadrupole field and Jacobian at the box centre,
   once per group, applied to each body to first order (near_k = INFINITY: all
   per body).  Not in the reference, whose walks are monopole only: an option
   that lowers the error at a given opening ratio. */
nbh_status nbh_bh_fcompute_group_quad(nbh_ctx* ctx, const nbh_tree* tree, const double* y, double* f,
                                      float near_k);
/* nbh_bh_fcompute_group_quad with the far cells' monopole field expanded too:
   a cell whose mass centre lies at least near_k (>= 1) half-diagonals from the
   group's box centre contributes its monopole field and first to third
   derivatives at the centre plus its quadrupole field and Jacobian, once per
   group; each body applies the sum (monopole to third order, quadrupole to
   first) in its offset from the centre.  Nearer cells and leaves are swept per body as in
   nbh_bh_fcompute_group_quad.  Not in the reference: a local (Taylor) expansion
   in the manner of fast multipole methods, replacing most of the per-body work
   of a group walk. */
nbh_status nbh_bh_fcompute_group_local(nbh_ctx* ctx, const nbh_tree* tree, const double* y, double* f,
                                       float near_k);
/* nbh_bh_fcompute_group_local in two levels: the tree is walked once per
   super-group (heap tree: 1024 bodies, 16 groups; Morton tree: see
   nbh_tree_build_radix) with the super-group's box; accepted cells at least
   max(near_k, 5) half-diagonals of that box from its centre form one expansion
   about it (monopole to third order, quadrupole to first), and the other
   accepted nodes a list that each group takes instead of walking (leaves and
   nearer cells per body, farther ones into the group's own expansion).  A node
   accepted for the super-group's box is accepted for every group box inside
   it, so the groups keep the acceptance rule; on the Morton tree the
   super-group walk also opens every rejected node wider than half its box
   rather than listing it.  Lists longer than 32768
   entries fall back to the one-level walk for that super-group (counted in
   walk counter out[2]).  The lists (n * 128 bytes at the default capacity) are
   allocated by nbh_tree_set_list_cap, or by the first call when that was not
   called; a failed allocation leaves none of them allocated.
   Not in the reference. */
nbh_status nbh_bh_fcompute_group_local2(nbh_ctx* ctx, nbh_tree* tree, const double* y, double* f, float near_k);
/* nbh_bh_fcompute_group_local2 with a learned correction of the super-groups'
   shared far fields.  The super-group walk also lists each super-group's far
   cells (up to 8192; further ones keep the plain expansion), and one wave per
   super-group evaluates the network (nbh_mlp_*, BF16 on the matrix cores) on
   them, 16 cells per tile: from the invariants of (super box centre p, cell)
   it predicts the cell's field remainder after the monopole and quadrupole
   terms -- R(p) and J = dR/dp through the basis of nbh_mlp_forward, with
   R / w and kJac J d / w as outputs, w = M rho^3 / d^2 -- and adds them to the
   super-group's expansion (field and Jacobian), which every body of the
   super-group applies to first order in its offset from p.  The far-cell lists
   take n / 1024 * 8192 * 4 bytes, allocated on first use.  Not in the reference. */
nbh_status nbh_bh_fcompute_group_local2_mlp(nbh_ctx* ctx, nbh_tree* tree, const nbh_mlp* net, const double* y,
                                            double* f, float near_k);
/* Training samples of nbh_bh_fcompute_group_local2_mlp: runs the super-group
   walk (near_k as for the walk) and, for the super-groups listed in the host
   array supers[n_supers], writes every listed far cell's features x[16],
   basis[16] and exact target t[9] (FP64 field of the cell's bodies at the
   super box centre minus its monopole and quadrupole terms, and the same for
   the Jacobian, scaled as above, rounded to FP32) into device arrays of
   `capacity` rows; cells (device, may be NULL) receives the (super-group,
   heap node) pair of each row; *count = rows found (rows beyond capacity are
   not written).  Synchronous.  Not in the reference. */
nbh_status nbh_bh_super_samples(nbh_ctx* ctx, nbh_tree* tree, float near_k, const uint32_t* supers,
                                size_t n_supers, float* x, float* basis, float* target, uint32_t* cells,
                                size_t capacity, size_t* count);
/* List capacity per super-group of nbh_bh_fcompute_group_local2 (default 32768
   entries; lists take n / 1024 * cap * 4 bytes).  Super-groups with longer lists
   walk in one level.  Frees the lists already allocated and allocates them at
   the new capacity (all of them or, on failure, none: NBH_ERR_NOMEM); cap in
   [64, 2^24].  No counterpart in the reference: it bounds the two-level walk's
   memory and lets a caller allocate it outside the force evaluation. */
nbh_status nbh_tree_set_list_cap(nbh_ctx* ctx, nbh_tree* tree, int cap);
/* Device bytes held by the tree: its build buffers and node records, plus the
   two-level walk's lists when they are allocated.  No counterpart in the
   reference (memory accounting). */
nbh_status nbh_tree_bytes(nbh_ctx* ctx, const nbh_tree* tree, size_t* bytes);
/* Copy the tree to host arrays of tree size 2n (index 0 unused); any pointer may
   be NULL.  cm/bmin/bmax are xyz-interleaved [2n*3]; body_n holds -1 on
   internal nodes.  Used to check the layout against the reference heap
   (nbody_space_heap::get_body_n / get_mass_center / get_radius_sqr). */
nbh_status nbh_tree_get(nbh_ctx* ctx, const nbh_tree* tree, int32_t* body_n, double* cm,
                        double* mass, double* crit_r2, double* bmin, double* bmax);
/* Number of accepted node interactions and of node visits of the last
   nbh_bh_fcompute (summed over targets); either pointer may be NULL. */
nbh_status nbh_bh_counters(nbh_ctx* ctx, const nbh_tree* tree, uint64_t* interactions,
                           uint64_t* visits);

/* Counters of the last traversal: out[0] pair interactions, out[1] node visits,
   out[2] nodes taken as monopoles because the walk stack was full, out[3] far
   cells (corrected by the network, or expanded about the group centre by
   nbh_bh_fcompute_group_quad) and out[4] near cells given their quadrupole term
   per body. */
nbh_status nbh_bh_walk_counters(nbh_ctx* ctx, const nbh_tree* tree, uint64_t out[5]);

/* ---- learned correction of the group walk ---------------------------------
   No counterpart in the reference.  The internal cells that the group walk
   takes as monopoles are split by the distance dc from the group's box centre
   p to their mass centre, in units of the box's half-diagonal h:
     near: dc < near_k * h.  The cell's second central moments Q are applied to
           every body of the group (quadrupole term, FP32).
     far:  otherwise, when the cell's bounding radius r exceeds rho_min * dc.  A
           16-64-64-9 tanh MLP (BF16 operands on the matrix cores, FP32
           accumulate) maps 16 rotation invariants of (p, cell: mass, box,
           Q) to 9 coefficients of a basis built from the direction uh to the
           cell, w = q uh (q = Q / (M r^2)), the cell's box-centre offset o and
           q (sample layout of basis[16]: uh, w, o, q as xx yy zz xy xz yz, 0):
             R / scale = c0 uh + c1 w + c2 o,
             J dc / (4 scale) = c3 q + c4 (uh w' + w uh') + c5 I + c6 uh uh'
                                + c7 (uh o' + o uh') + c8 o o',
           R(p) = exact field of the cell's bodies - monopole field, J = dR/dp,
           scale = M r^2 / dc^4 (target layout t[9]: R (3), J as xx yy zz xy
           xz yz).  Each body of the group gets the sum over far cells of
           R(p) + J (x - p).
   nbh_mlp_* keep an FP32 master copy of the 6288 parameters (row-major
   W1[16][64] b1[64] W2[64][64] b2[64] W3[64][16] b3[16], output columns 9..15
   unused) and its BF16 matrix-core packing. */
nbh_status nbh_mlp_create(nbh_ctx* ctx, uint64_t seed, nbh_mlp** net);
nbh_status nbh_mlp_destroy(nbh_mlp* net);
nbh_status nbh_mlp_param_count(int* count);
/* Host arrays of nbh_mlp_param_count floats; set also resets the Adam state. */
nbh_status nbh_mlp_set_params(nbh_ctx* ctx, nbh_mlp* net, const float* params);
nbh_status nbh_mlp_get_params(nbh_ctx* ctx, const nbh_mlp* net, float* params);
/* Device x[count*16] -> device out[count*9], the basis coefficients (asynchronous). */
nbh_status nbh_mlp_forward(nbh_ctx* ctx, const nbh_mlp* net, const float* x, size_t count,
                           float* out);
/* Gradient of L = sum |B(basis) out - t|^2 / count over device x[count*16],
   basis[count*16], t[count*9] into host grad[param_count]; *loss = L (may be
   NULL).  Synchronous. */
nbh_status nbh_mlp_gradient(nbh_ctx* ctx, nbh_mlp* net, const float* x, const float* basis,
                            const float* t, size_t count, float* grad, float* loss);
/* Adam (beta 0.9/0.999, eps 1e-8) with a cosine learning-rate decay from lr to 0
   over epochs * ceil(count/batch) steps; batch b of an epoch holds the samples
   b, b + nb, b + 2nb, ...  epoch_loss[epochs] (may be NULL) receives the mean
   loss of the last batch of each epoch.  Synchronous. */
nbh_status nbh_mlp_train(nbh_ctx* ctx, nbh_mlp* net, const float* x, const float* basis,
                         const float* t, size_t count, int epochs, size_t batch, float lr,
                         float* epoch_loss);
/* Training samples from a built tree: for the groups listed in the host array
   groups[n_groups] (group g = the 64 bodies under heap node n/64 + g), every far
   cell's features x[16], basis[16] and exact target t[9] (FP64 sums over the
   cell's bodies, rounded to FP32) into device arrays of `capacity` rows; cells (device, may be
   NULL) receives the (group, heap node) pair of each row.  *count = number of
   far cells found (rows beyond capacity are not written; their order depends
   on the schedule, their set does not).  Synchronous. */
nbh_status nbh_bh_band_samples(nbh_ctx* ctx, const nbh_tree* tree, const uint32_t* groups,
                               size_t n_groups, float rho_min, float near_k, float* x,
                               float* basis, float* target, uint32_t* cells, size_t capacity,
                               size_t* count);
/* nbh_bh_fcompute_group plus the near-cell quadrupole terms and the network's
   correction of the far cells.  near_k = 0 and an all-zero network give
   nbh_bh_fcompute_group's result exactly. */
nbh_status nbh_bh_fcompute_group_mlp(nbh_ctx* ctx, const nbh_tree* tree, const nbh_mlp* net,
                                     const double* y, double* f, float rho_min, float near_k);

/* ---- kernel timers -------------------------------------------------------
   When enabled, the main kernel of each entry point is bracketed by HIP events
   on the context's stream.  nbh_timing_read synchronises the stream and returns
   the summed device time (ms) and launch count of one kernel class since the
   last reset.  Classes: "direct_fp64", "direct_fp32", "direct_fp64_block",
   "tree_build", "tree_refit", "bh_traverse" (a whole walk), "bh_group_walk"
   (the group walks of the two-level walk alone), "ah_full", "ah_sparse",
   "fmaddn", "stats".  No counterpart in the
   reference (its timing is host wall clock, test/bench/bench.h:105-115). */
nbh_status nbh_timing_enable(nbh_ctx* ctx, int enable);
nbh_status nbh_timing_reset(nbh_ctx* ctx);
nbh_status nbh_timing_read(nbh_ctx* ctx, const char* kernel_class, double* total_ms, uint64_t* launches);

/* ---- conservation checks ------------------------------------------------
   FP64 compensated reductions of the P/L/E/V quantities of
   nbody_data::print_statistics (nbody/nbody_data.cpp:57-102, proxies
   nbody/summation_proxy.h:6-87) over the device state y[6n], mass[n].
   out[12]: P(3) L(3) KE PE CM(3, divided by total mass) total_mass.
   with_energy = 0 skips the O(n^2) potential energy (out[7] = 0). */
nbh_status nbh_stats(nbh_ctx* ctx, const double* y, const double* mass, size_t n,
                     int with_energy, double out[12]);

#ifdef __cplusplus
}
#endif

#endif /* NBODY_HIP_H */
