// nbody_engine_hip.h -- the `hip` (direct all-pairs) and `hip_bh` (Barnes-Hut)
// engines: nbody_engine implementations whose buffers live in MI355X HBM and
// whose every operation is one or a few launches of the gfx950 kernels behind
// the C ABI of include/nbody_hip.h.
//
// Reference counterparts: nbody_engine_cuda (nbody/nbody_engine_cuda.{h,cpp})
// and nbody_engine_cuda_bh / nbody_engine_cuda_bh_tex
// (nbody/nbody_engine_cuda_bh*.{h,cpp}).  Differences by design: one device
// per engine, the multi-term ops are single fused launches, the Barnes-Hut
// tree is built on the GPU, and HIP failures are logged and latched in
// has_error() instead of terminating the process.
#pragma once

#include <string>
#include <vector>

#include "nbody_engine.h"
#include "nbody_hip.h"

enum e_traverse_type
{
	ett_cycle,
	ett_nested_tree
};

enum e_tree_layout
{
	etl_tree,
	etl_heap,
	etl_heap_stackless,
	etl_unknown = 0xffff
};

const char* tree_layout_name(e_tree_layout tree_layout);
e_tree_layout tree_layout_from_str(const std::string& name);

class nbody_engine_hip : public nbody_engine
{
public:
	class smemory : public memory
	{
		nbh_ctx* m_ctx;
		void*    m_data;
		size_t   m_size;
	public:
		smemory(nbh_ctx* ctx, size_t size);
		~smemory() override;
		bool valid() const { return m_size == 0 || m_data != nullptr; }
		double* data() { return static_cast<double*>(m_data); }
		const double* data() const { return static_cast<const double*>(m_data); }
		size_t size() const override { return m_size; }
		size_t count() const { return m_size / sizeof(double); }
	};

protected:
	nbh_ctx*      m_ctx;
	int           m_device;
	int           m_block_size;
	nbh_precision m_precision;
	nbody_data*   m_data;
	smemory*      m_mass;
	smemory*      m_y;
	nbh_snapshot* m_snap;       // frame snapshots (created on the first dump)
	bool          m_error;

	bool check(nbh_status s, const char* what);
	bool check_f(const memory* y, const memory* f, const smemory** ys, smemory** fs);
public:
	nbody_engine_hip();
	~nbody_engine_hip() override;

	//! Select the HIP device; returns 0 on success (nbody_engine_cuda::select_devices).
	int select_device(int device);
	void set_block_size(int block_size);
	void set_precision(nbh_precision p);
	nbh_precision get_precision() const { return m_precision; }
	//! True once any HIP call of this engine has failed.
	bool has_error() const { return m_error; }
	nbh_ctx* context() { return m_ctx; }
	const smemory* mass_buffer() const { return m_mass; }

	const char* type_name() const override;
	bool init(nbody_data* data) override;
	void get_data(nbody_data* data) override;
	size_t problem_size() const override;
	memory* get_y() override;
	void advise_time(const nbcoord_t& dt) override;
	nbcoord_t get_time() const override;
	void set_time(nbcoord_t t) override;
	size_t get_step() const override;
	void set_step(size_t s) override;
	void fcompute(const nbcoord_t& t, const memory* y, memory* f) override;
	void clamp(memory* y, nbcoord_t b) override;
	//! Device-side snapshots: a staging copy on the engine's stream, then a
	//! copy to pinned host memory that overlaps the following steps.
	bool snapshot_begin(int slot) override;
	const nbcoord_t* snapshot_state(int slot) override;

	smemory* create_buffer(size_t bytes) override;
	void free_buffer(memory* m) override;
	void read_buffer(void* dst, const memory* src) override;
	void write_buffer(memory* dst, const void* src) override;
	void copy_buffer(memory* a, const memory* b) override;
	void fill_buffer(memory* a, const nbcoord_t& value) override;
	void fmadd_inplace(memory* a, const memory* b, const nbcoord_t& c) override;
	void fmadd(memory* a, const memory* b, const memory* c, const nbcoord_t& d) override;
	void fmaddn_inplace(memory* a, const memory_array& b, const nbcoord_t* c, size_t csize) override;
	void fmaddn_corr(memory* a, memory* corr, const memory_array& b, const nbcoord_t* c, size_t csize) override;
	void fmaddn(memory* a, const memory* b, const memory_array& c, const nbcoord_t* d, size_t dsize) override;
	void fmaxabs(const memory* a, nbcoord_t& result) override;
	void print_info() const override;
	bool compute_statistics(bool with_energy, double out[12]) override;
};

// The `block` engine's arithmetic on the GPU (nbody/nbody_engine_block.{h,cpp}):
// FP64 source tiles, a = -sum m_j (r_i - r_j) / (sqrt(r2) r2) with the self pair
// included and no division by m_i.  Everything else as nbody_engine_hip.
class nbody_engine_hip_block : public nbody_engine_hip
{
public:
	nbody_engine_hip_block();
	const char* type_name() const override;
};

// Ahmad-Cohen neighbour scheme (nbody/nbody_engine_ah.{h,cpp}) on the GPU: a
// full FP64 recompute every full_recompute_rate-th step builds the neighbour
// lists (|r1 - r2| < max_dist or |F12| > min_force) and freezes the far force;
// the other steps sum the far force and the neighbours only.
class nbody_engine_hip_ah : public nbody_engine_hip
{
	size_t  m_rate;
	double  m_max_dist;
	double  m_min_force;
	nbh_ah* m_ah;
	bool    m_have_lists;
public:
	nbody_engine_hip_ah(size_t full_recompute_rate, double max_dist, double min_force);
	~nbody_engine_hip_ah() override;
	const char* type_name() const override;
	bool init(nbody_data* data) override;
	void fcompute(const nbcoord_t& t, const memory* y, memory* f) override;
	void print_info() const override;
	const nbh_ah* neighbours() const { return m_ah; }
};

class nbody_engine_hip_bh : public nbody_engine_hip
{
	nbcoord_t       m_ratio;
	e_traverse_type m_traverse_type;
	size_t          m_tree_build_rate;
	e_tree_layout   m_tree_layout;
	nbh_tree*       m_tree;
	bool            m_tree_built;
	bool            m_group_walk;
	bool            m_quadrupole;
	float           m_quad_near;
	bool            m_local_far;
	bool            m_morton;          // group walks on the Morton radix tree
	std::vector<float> m_net_params;   // learned correction (empty: none)
	float           m_net_rho;
	float           m_net_near;
	std::string     m_net_path;
	bool            m_net_two_level;   // the network corrects the two-level walk's super-group far fields
	nbh_mlp*        m_net;
public:
	nbody_engine_hip_bh(nbcoord_t distance_to_node_radius_ratio, e_traverse_type tt,
	                    size_t tree_build_rate, e_tree_layout tl);
	//! Use the shared 64-body group walk (FP32 only) instead of one walk per body.
	void set_group_walk(bool g) { m_group_walk = g; }
	//! Add the quadrupole term of taken cells to the group walk.
	void set_quadrupole(bool q) { m_quadrupole = q; }
	//! Cells within near_k group half-diagonals get the quadrupole per body,
	//! farther ones through an expansion about the group centre.
	void set_quad_near(float near_k) { m_quad_near = near_k; }
	//! Expand the far cells' monopole field about the group centre too (second
	//! order) instead of sweeping it per body.
	void set_local_far_field(bool l) { m_local_far = l; }
	//! Build the Morton radix tree (nbh_tree_build_radix) instead of the heap
	//! kd-tree; only the group walks run on it.
	void set_morton_tree(bool m) { m_morton = m; }
	//! Read a correction network file ("NBHMLP01", int32 parameter count,
	//! float32 rho_min, float32 near_k, float32 parameters); rho >= 0 and
	//! near_k >= 0 override the thresholds stored in it.  False if the file is
	//! unreadable or malformed.
	bool load_correction(const std::string& path, double rho, double near_k);
	const nbh_mlp* correction() const { return m_net; }
	~nbody_engine_hip_bh() override;
	const char* type_name() const override;
	bool init(nbody_data* data) override;
	void fcompute(const nbcoord_t& t, const memory* y, memory* f) override;
	void print_info() const override;
	const nbh_tree* tree() const { return m_tree; }
	nbcoord_t ratio() const { return m_ratio; }
	e_tree_layout layout() const { return m_tree_layout; }
};
