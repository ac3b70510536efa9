// nbody_engine_mi355x.h -- the `hip` and `hip_bh` engines as classes of the
// drons/nbody tree itself (Qt build), forwarding every nbody_engine method to
// the C ABI of include/nbody_hip.h.  This file is compiled inside the
// reference tree (qmake, with pri/hip.pri as INTEGRATION.md shows), not by this
// repository's build; tests/test_integration_adapter.py compiles it against the
// reference's own nbody_engine.h and nbody_data.h to keep it in step with them.
//
// Reference counterparts: nbody_engine_cuda / nbody_engine_cuda_bh_tex
// (nbody/nbody_engine_cuda.h, nbody/nbody_engine_cuda_bh_tex.h).  Misuse is
// logged with qDebug and the call returns, as in nbody/nbody_engine_simple.cpp:
// 122-131; a failed nbh_* call is logged with nbh_last_error() (the CUDA
// engine exits the process instead, nbody/nbody_engine_cuda.cpp:8-15).
#ifndef NBODY_ENGINE_MI355X_H
#define NBODY_ENGINE_MI355X_H

#include <QVariantMap>

#include "nbody_engine.h"
#include "nbody_hip.h"

class NBODY_DLL nbody_engine_mi355x : public nbody_engine
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
		size_t size() const override;
		double* data() const;
		size_t count() const;
	};
protected:
	nbh_ctx*      m_ctx;
	nbh_precision m_precision;
	nbody_data*   m_data;
	smemory*      m_mass;
	smemory*      m_y;
	bool ok(nbh_status s, const char* what) const;
	bool check_f(const memory* y, const memory* f, const smemory** ys, smemory** fs) const;
	//! the non-zero terms of b[0..csize) with their coefficients
	bool gather(const memory_array& b, const nbcoord_t* c, size_t csize, size_t n,
	            std::vector<const double*>& src, std::vector<double>& coef) const;
public:
	nbody_engine_mi355x();
	~nbody_engine_mi355x() override;
	//! one HIP device per engine (nbody_engine_cuda::select_devices for one id)
	int select_device(int device);
	void set_precision(nbh_precision p);

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
	memory* create_buffer(size_t s) override;
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
};

class NBODY_DLL nbody_engine_mi355x_bh : public nbody_engine_mi355x
{
	nbh_tree*       m_tree;
	nbcoord_t       m_ratio;
	size_t          m_tree_build_rate;
	nbh_tree_layout m_layout;
	bool            m_built;
public:
	nbody_engine_mi355x_bh(nbcoord_t distance_to_node_radius_ratio, size_t tree_build_rate,
	                       nbh_tree_layout layout);
	~nbody_engine_mi355x_bh() override;
	const char* type_name() const override;
	bool init(nbody_data* data) override;
	void fcompute(const nbcoord_t& t, const memory* y, memory* f) override;
	void print_info() const override;
};

//! The factory branch of the two engines: "hip" or "hip_bh" with the keys
//! device, force_precision (fp64 | fp32 | fp32_corr for hip), and for hip_bh
//! distance_to_node_radius_ratio, tree_build_rate, tree_layout
//! (heap | heap_stackless), traverse_type (cycle | nested_tree: same numbers);
//! NULL for anything else, as nbody_create_engine (nbody/nbody_engines.cpp:21-190).
NBODY_DLL nbody_engine* nbody_create_engine_mi355x(const QVariantMap& param);

#endif // NBODY_ENGINE_MI355X_H
