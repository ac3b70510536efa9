// nbody_solver.h -- ODE solvers that drive an nbody_engine.
//
// Same classes, defaults and per-step engine-op sequences as the reference
// solvers (nbody/nbody_solver*.{h,cpp}, nbody/nbody_extrapolator.{h,cpp},
// factory nbody/nbody_solvers.cpp:17-110).  They call the engine only through
// the nbody_engine interface, so the `hip` and `hip_bh` engines run under
// them unchanged.
//
// Licence: this file restates code of drons/nbody (the reference files cited
// in it), which is distributed under the GNU Lesser General Public License,
// version 3; this file is distributed under the same licence.
#pragma once

#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "nbody_butcher_table.h"
#include "nbody_data_stream.h"
#include "nbody_engine.h"
#include "nbody_params.h"

class nbody_solver
{
	nbody_engine* m_engine;
	nbcoord_t     m_min_step;
	nbcoord_t     m_max_step;
	bool          m_clamp_to_box;
	std::vector<std::function<void(nbody_data*)>> m_check_visitors;
	nbody_solver(const nbody_solver&) = delete;
	nbody_solver& operator=(const nbody_solver&) = delete;
public:
	nbody_solver();
	virtual ~nbody_solver();
	void set_engine(nbody_engine*);
	nbody_engine* engine();
	void set_time_step(nbcoord_t min_step, nbcoord_t max_step);
	nbcoord_t get_min_step() const;
	nbcoord_t get_max_step() const;
	void set_clamp_to_box(bool clamp);
	void add_check_visitor(std::function<void(nbody_data*)> v);
	//! Integrate to max_time in steps of max_step (nbody/nbody_solver.cpp:55-105):
	//! statistics every check_dt; with a stream and dump_dt > 0 a frame at the
	//! start (when the time is <= 0) and every dump_dt.  Frames are written by a
	//! second thread from snapshots the engine takes between steps, so the
	//! steps go on while a frame is copied and written; on return every frame
	//! is written and `data` holds the last one.  -1 when a frame could not be
	//! taken or written.
	int run(nbody_data* data, nbody_data_stream* stream, nbcoord_t max_time, nbcoord_t dump_dt,
	        nbcoord_t check_dt);

	virtual const char* type_name() const = 0;
	virtual void advise(nbcoord_t dt) = 0;
	virtual void print_info() const;
	virtual e_ode_order get_ode_order() const;
	virtual void reset();
};

class nbody_solver_euler : public nbody_solver
{
	nbody_engine::memory* m_dy;
public:
	nbody_solver_euler();
	~nbody_solver_euler() override;
	const char* type_name() const override;
	void advise(nbcoord_t dt) override;
};

class nbody_solver_midpoint : public nbody_solver
{
	nbody_engine::memory* m_k1;
	nbody_engine::memory* m_k2;
	nbody_engine::memory* m_tmp;
public:
	nbody_solver_midpoint();
	~nbody_solver_midpoint() override;
	const char* type_name() const override;
	void advise(nbcoord_t dt) override;
};

class nbody_solver_midpoint_stetter : public nbody_solver
{
	nbody_engine::memory_array m_uv;
	nbody_engine::memory_array m_fu;
	nbody_engine::memory*      m_tmp;
	nbody_engine::memory*      m_du;
	bool                       m_uv_initiated;
public:
	nbody_solver_midpoint_stetter();
	~nbody_solver_midpoint_stetter() override;
	const char* type_name() const override;
	void advise(nbcoord_t dt) override;
	void reset() override;
};

class nbody_solver_rk4 : public nbody_solver
{
	nbody_engine::memory_array m_k;
	nbody_engine::memory*      m_tmp;
public:
	nbody_solver_rk4();
	~nbody_solver_rk4() override;
	const char* type_name() const override;
	void advise(nbcoord_t dt) override;
};

class nbody_solver_trapeze : public nbody_solver
{
	nbody_engine::memory_array m_f;
	nbody_engine::memory*      m_predictor;
	size_t                     m_refine_steps_count;
public:
	nbody_solver_trapeze();
	~nbody_solver_trapeze() override;
	const char* type_name() const override;
	void set_refine_steps_count(size_t v);
	void advise(nbcoord_t dt) override;
	void print_info() const override;
};

class nbody_solver_adams : public nbody_solver
{
	nbody_solver*              m_starter;
	nbody_engine::memory_array m_f;
	nbody_engine::memory*      m_corr_data;
	size_t                     m_rank;
	size_t                     m_step;
	bool                       m_correction;
public:
	nbody_solver_adams(nbody_solver* starter, size_t rank, bool corr);
	~nbody_solver_adams() override;
	const char* type_name() const override;
	void advise(nbcoord_t dt) override;
	void print_info() const override;
	void reset() override;
};

class nbody_solver_rk_butcher : public nbody_solver
{
	nbody_butcher_table*       m_bt;
	std::string                m_type_name;
	nbody_engine::memory_array m_k;
	nbody_engine::memory*      m_t;
	nbody_engine::memory*      m_tmpk;
	nbody_engine::memory_array m_y_stack;
	nbody_engine::memory*      m_ycorr_data;
	nbody_engine::memory*      m_tcorr_data;
	size_t                     m_max_recursion;
	size_t                     m_substep_subdivisions;
	nbcoord_t                  m_error_threshold;
	size_t                     m_refine_steps_count;
	bool                       m_correction;
	size_t                     m_substeps_done;
public:
	explicit nbody_solver_rk_butcher(nbody_butcher_table* t);
	~nbody_solver_rk_butcher() override;
	const char* type_name() const override;
	void set_max_recursion(size_t v);
	void set_substep_subdivisions(size_t v);
	void set_error_threshold(nbcoord_t v);
	void set_refine_steps_count(size_t v);
	void set_correction(bool corr);
	void advise(nbcoord_t dt) override;
	void print_info() const override;
	void reset() override;
	const nbody_butcher_table* table() const;
	//! Number of (sub)steps applied to y since creation (1 per accepted step).
	size_t substeps_done() const { return m_substeps_done; }
private:
	void sub_step(size_t substeps_count, nbcoord_t t, nbcoord_t dt, nbody_engine::memory* y,
	              size_t recursion_level);
	void sub_step_implicit(size_t steps, const nbcoord_t** a, nbcoord_t* coeff,
	                       const nbody_engine::memory* y, bool need_first_approach_k,
	                       const nbcoord_t* c, nbcoord_t t, nbcoord_t dt);
	void sub_step_explicit(size_t steps, const nbcoord_t** a, nbcoord_t* coeff,
	                       const nbody_engine::memory* y, const nbcoord_t* c, nbcoord_t t, nbcoord_t dt);
};

//! Extrapolation of a solution to substep_count -> infinity
//! (nbody/nbody_extrapolator.h:9-27).
class nbody_extrapolator
{
	nbody_engine*       m_e;
	std::vector<size_t> m_sub_steps_count;
public:
	nbody_extrapolator(nbody_engine* engine, size_t order, const std::vector<size_t>& substeps_count);
	virtual ~nbody_extrapolator();
	nbody_engine* engine() const;
	size_t sub_steps_count(size_t level) const;
	virtual void update_table(size_t level, nbody_engine::memory* y) = 0;
	virtual nbcoord_t estimate_error(size_t level) const = 0;
	virtual void extrapolate(size_t level, nbody_engine::memory* ext_y) const = 0;
};

//! Berrut rational extrapolation (nbody/nbody_extrapolator.cpp:34-103)
class nbody_extrapolator_berrut : public nbody_extrapolator
{
	nbody_engine::memory_array m_table;
	nbody_engine::memory*      m_diff;
public:
	nbody_extrapolator_berrut(nbody_engine* engine, size_t order, const std::vector<size_t>& substeps_count);
	~nbody_extrapolator_berrut() override;
	void update_table(size_t level, nbody_engine::memory* y) override;
	nbcoord_t estimate_error(size_t level) const override;
	void extrapolate(size_t level, nbody_engine::memory* ext_y) const override;
private:
	std::vector<nbcoord_t> weights(size_t level) const;
};

//! Neville extrapolation, Numerical Recipes (17.3.8) (nbody/nbody_extrapolator.cpp:105-181)
class nbody_extrapolator_neville : public nbody_extrapolator
{
	nbody_engine::memory_array m_table;
	nbody_engine::memory_array m_table_prev;
	nbody_engine::memory*      m_diff;
public:
	nbody_extrapolator_neville(nbody_engine* engine, size_t order, const std::vector<size_t>& substeps_count);
	~nbody_extrapolator_neville() override;
	void update_table(size_t level, nbody_engine::memory* y) override;
	nbcoord_t estimate_error(size_t level) const override;
	void extrapolate(size_t level, nbody_engine::memory* ext_y) const override;
};

nbody_extrapolator* nbody_create_extrapolator(const std::string& type, nbody_engine* engine, size_t order,
                                              const std::vector<size_t>& substeps_count);

class nbody_solver_bulirsch_stoer : public nbody_solver
{
	nbody_solver_midpoint_stetter* m_internal;
	size_t                         m_max_level;
	nbcoord_t                      m_error_threshold;
	std::vector<size_t>            m_sub_steps_count;
	nbody_engine::memory*          m_y0;
	nbody_extrapolator*            m_extrapolator;
	size_t                         m_last_level;
public:
	nbody_solver_bulirsch_stoer(size_t max_level, nbcoord_t error_threshold);
	~nbody_solver_bulirsch_stoer() override;
	const char* type_name() const override;
	void advise(nbcoord_t dt) override;
	void print_info() const override;
	//! Level at which the last step stopped.
	size_t last_level() const { return m_last_level; }
private:
	void compute_substep(size_t level, nbcoord_t dt, nbcoord_t t0);
};

//! Factory with the keys and defaults of nbody/nbody_solvers.cpp:17-110.
//! Returns NULL for an unknown "solver" (or adams "starter_solver").
nbody_solver* nbody_create_solver(const nbody_params& param);
