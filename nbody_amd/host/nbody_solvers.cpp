// nbody_solvers.cpp -- solver implementations (see nbody_solver.h).  Each
// advise() issues the same engine calls, with the same coefficients, as the
// reference method cited beside it (paths relative to the drons/nbody root).
//
// Licence: this file restates code of drons/nbody (the reference files cited
// in it), which is distributed under the GNU Lesser General Public License,
// version 3; this file is distributed under the same licence.
#include "nbody_solver.h"

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <limits>
#include <mutex>
#include <numeric>
#include <thread>

#include "nbody_log.h"

// ---------------------------------------------------------------------------
// base: nbody/nbody_solver.cpp
// ---------------------------------------------------------------------------
nbody_solver::nbody_solver() : m_engine(nullptr), m_min_step(0), m_max_step(0), m_clamp_to_box(false) {}
nbody_solver::~nbody_solver() {}

void nbody_solver::set_engine(nbody_engine* e)
{
	m_engine = e;
	m_engine->set_ode_order(get_ode_order());
}

nbody_engine* nbody_solver::engine() { return m_engine; }

void nbody_solver::set_time_step(nbcoord_t min_step, nbcoord_t max_step)
{
	m_min_step = min_step;
	m_max_step = max_step;
}

nbcoord_t nbody_solver::get_min_step() const { return m_min_step; }
nbcoord_t nbody_solver::get_max_step() const { return m_max_step; }
void nbody_solver::set_clamp_to_box(bool clamp) { m_clamp_to_box = clamp; }
void nbody_solver::add_check_visitor(std::function<void(nbody_data*)> v) { m_check_visitors.push_back(v); }

// nbody/nbody_solver.cpp:55-105 (stream dumps are outside this build)
namespace {

// Frames of a run's dumps, written by a second thread: the run loop has the
// engine take a snapshot of the state (nbody_engine::snapshot_begin) and queues
// it; the writer waits for the copy, turns the SoA state into the AoS frame of
// the stream format and appends it, while the loop goes on with the next steps.
// The two snapshot slots alternate; a dump waits until the frame that last
// used its slot is written.
class frame_writer
{
	struct job
	{
		int       slot;
		size_t    step;
		nbcoord_t time;
	};
	nbody_engine*           m_engine;
	nbody_data_stream*      m_stream;
	const nbody_data*       m_data;
	std::mutex              m_mu;
	std::condition_variable m_cv;
	std::deque<job>         m_queue;
	bool                    m_busy[2] = {false, false};
	bool                    m_stop = false;
	bool                    m_failed = false;
	int                     m_next_slot = 0;
	bool                    m_written = false;
	std::vector<nbvertex_t> m_pos, m_vel;       // the last frame written
	std::thread             m_thread;

	void loop()
	{
		const size_t n = m_data->get_count();
		for(;;)
		{
			job j;
			{
				std::unique_lock<std::mutex> lock(m_mu);
				m_cv.wait(lock, [this] { return m_stop || !m_queue.empty(); });
				if(m_queue.empty())
				{
					return;
				}
				j = m_queue.front();
				m_queue.pop_front();
			}
			bool ok = false;
			const nbcoord_t* y = m_engine->snapshot_state(j.slot);
			if(y != nullptr)
			{
				m_pos.resize(n);
				m_vel.resize(n);
				for(size_t i = 0; i < n; ++i)
				{
					m_pos[i] = nbvertex_t(y[i], y[n + i], y[2 * n + i]);
					m_vel[i] = nbvertex_t(y[3 * n + i], y[4 * n + i], y[5 * n + i]);
				}
				ok = m_stream->write_frame(m_data, j.step, j.time, m_pos.data(), m_vel.data()) == 0;
			}
			std::lock_guard<std::mutex> lock(m_mu);
			m_busy[j.slot] = false;
			m_failed = m_failed || !ok;
			m_written = m_written || ok;
			m_cv.notify_all();
		}
	}

public:
	frame_writer(nbody_engine* e, nbody_data_stream* stream, const nbody_data* data) :
		m_engine(e), m_stream(stream), m_data(data)
	{
		m_thread = std::thread([this] { loop(); });
	}

	~frame_writer()
	{
		stop();
	}

	void stop()
	{
		{
			std::lock_guard<std::mutex> lock(m_mu);
			m_stop = true;
		}
		m_cv.notify_all();
		if(m_thread.joinable())
		{
			m_thread.join();
		}
	}

	//! Queue a frame of the current state; -1 if an earlier frame failed or the
	//! snapshot cannot be taken.
	int dump(size_t step, nbcoord_t time)
	{
		const int slot = m_next_slot;
		m_next_slot ^= 1;
		{
			std::unique_lock<std::mutex> lock(m_mu);
			m_cv.wait(lock, [this, slot] { return m_failed || !m_busy[slot]; });
			if(m_failed)
			{
				return -1;
			}
		}
		if(!m_engine->snapshot_begin(slot))
		{
			return -1;
		}
		std::lock_guard<std::mutex> lock(m_mu);
		m_busy[slot] = true;
		m_queue.push_back(job{slot, step, time});
		m_cv.notify_all();
		return 0;
	}

	//! Write what is queued, stop the thread and put the last frame into data.
	int finish(nbody_data* data)
	{
		stop();
		if(m_written)
		{
			std::copy(m_pos.begin(), m_pos.end(), data->get_vertites());
			std::copy(m_vel.begin(), m_vel.end(), data->get_velosites());
		}
		return m_failed ? -1 : 0;
	}
};

}  // namespace

int nbody_solver::run(nbody_data* data, nbody_data_stream* stream, nbcoord_t max_time, nbcoord_t dump_dt,
                      nbcoord_t check_dt)
{
	nbcoord_t dt = get_max_step();
	nbcoord_t last_check = data->get_time();
	nbcoord_t last_dump = last_check;
	std::unique_ptr<frame_writer> writer;
	if(stream != nullptr && dump_dt > 0)
	{
		writer.reset(new frame_writer(m_engine, stream, data));
		if(last_dump <= 0)
		{
			if(writer->dump(data->get_step(), data->get_time()) != 0)
			{
				writer->finish(data);
				nbody_log("run: can't write the first frame");
				return -1;
			}
			last_dump = data->get_time();
		}
	}
	while(data->get_time() < max_time)
	{
		if(m_clamp_to_box)
		{
			m_engine->clamp(m_engine->get_y(), static_cast<nbcoord_t>(data->get_box_size()));
		}
		advise(dt);
		nbcoord_t t = data->get_time();
		if(check_dt > 0 && t >= last_check + check_dt - dt * 0.1)
		{
			data->print_statistics(m_engine);
			for(auto& v : m_check_visitors)
			{
				v(data);
			}
			last_check = t;
		}
		if(writer && t >= last_dump + dump_dt - dt * 0.1)
		{
			if(writer->dump(data->get_step(), t) != 0)
			{
				writer->finish(data);
				nbody_log("run: can't write the frame of step %zu", data->get_step());
				return -1;
			}
			last_dump = t;
		}
	}
	if(writer && writer->finish(data) != 0)
	{
		nbody_log("run: a frame could not be written");
		return -1;
	}
	return 0;
}

void nbody_solver::print_info() const
{
	nbody_print("\tmin_step: %g", m_min_step);
	nbody_print("\tmax_step: %g", m_max_step);
	nbody_print("\tODE order: %d", static_cast<int>(get_ode_order()));
	nbody_print("\tClamp: %d", m_clamp_to_box ? 1 : 0);
}

e_ode_order nbody_solver::get_ode_order() const { return eode_first_order; }
void nbody_solver::reset() {}

static size_t state_bytes(nbody_engine* e)
{
	return sizeof(nbcoord_t) * e->problem_size();
}

// ---------------------------------------------------------------------------
// euler: nbody/nbody_solver_euler.cpp:19-32
// ---------------------------------------------------------------------------
nbody_solver_euler::nbody_solver_euler() : m_dy(nullptr) {}

nbody_solver_euler::~nbody_solver_euler()
{
	if(engine() != nullptr) { engine()->free_buffer(m_dy); }
}

const char* nbody_solver_euler::type_name() const { return "nbody_solver_euler"; }

void nbody_solver_euler::advise(nbcoord_t dt)
{
	nbody_engine::memory* y = engine()->get_y();
	if(m_dy == nullptr)
	{
		m_dy = engine()->create_buffer(state_bytes(engine()));
	}
	engine()->fcompute(engine()->get_time(), y, m_dy);
	engine()->fmadd_inplace(y, m_dy, dt);
	engine()->advise_time(dt);
}

// ---------------------------------------------------------------------------
// midpoint: nbody/nbody_solver_midpoint.cpp:23-40
// ---------------------------------------------------------------------------
nbody_solver_midpoint::nbody_solver_midpoint() : m_k1(nullptr), m_k2(nullptr), m_tmp(nullptr) {}

nbody_solver_midpoint::~nbody_solver_midpoint()
{
	if(engine() != nullptr)
	{
		engine()->free_buffer(m_k1);
		engine()->free_buffer(m_k2);
		engine()->free_buffer(m_tmp);
	}
}

const char* nbody_solver_midpoint::type_name() const { return "nbody_solver_midpoint"; }

void nbody_solver_midpoint::advise(nbcoord_t dt)
{
	if(m_k1 == nullptr)
	{
		m_k1 = engine()->create_buffer(state_bytes(engine()));
		m_k2 = engine()->create_buffer(state_bytes(engine()));
		m_tmp = engine()->create_buffer(state_bytes(engine()));
	}
	nbody_engine::memory* y = engine()->get_y();
	nbcoord_t t = engine()->get_time();
	engine()->fcompute(t, y, m_k1);
	engine()->fmadd(m_tmp, y, m_k1, dt / 2.0);
	engine()->fcompute(t + dt / 2.0, m_tmp, m_k2);
	engine()->fmadd_inplace(y, m_k2, dt);
	engine()->advise_time(dt);
}

// ---------------------------------------------------------------------------
// midpoint-st (Stetter): nbody/nbody_solver_midpoint_stetter.cpp:23-76
// ---------------------------------------------------------------------------
nbody_solver_midpoint_stetter::nbody_solver_midpoint_stetter() : m_tmp(nullptr), m_du(nullptr), m_uv_initiated(false) {}

nbody_solver_midpoint_stetter::~nbody_solver_midpoint_stetter()
{
	if(engine() != nullptr)
	{
		engine()->free_buffers(m_fu);
		engine()->free_buffers(m_uv);
		engine()->free_buffer(m_tmp);
		engine()->free_buffer(m_du);
	}
}

const char* nbody_solver_midpoint_stetter::type_name() const { return "nbody_solver_midpoint_stetter"; }

void nbody_solver_midpoint_stetter::advise(nbcoord_t dt)
{
	nbody_engine::memory* y = engine()->get_y();
	nbcoord_t t = engine()->get_time();
	bool first_run(m_fu.empty());
	if(first_run)
	{
		size_t size = state_bytes(engine());
		m_tmp = engine()->create_buffer(size);
		m_du = engine()->create_buffer(size);
		m_fu = engine()->create_buffers(size, 2);
		m_uv = engine()->create_buffers(size, 2);
	}
	if(!m_uv_initiated || first_run)
	{
		engine()->copy_buffer(m_uv[0], y);
		engine()->copy_buffer(m_uv[1], y);
		first_run = true;
		m_uv_initiated = true;
	}
	// u_{k+1}
	if(first_run)
	{
		engine()->fcompute(t, m_uv[0], m_fu[0]);
	}
	engine()->fmadd(m_tmp, m_uv[1], m_fu[0], dt / 2.0);
	engine()->fcompute(t + dt / 2.0, m_tmp, m_du);
	engine()->fmadd_inplace(m_uv[0], m_du, dt);
	// v_{k+1}
	{
		nbcoord_t v_coeff[] = { dt / 2.0, dt / 2.0 };
		engine()->fcompute(t + dt, m_uv[0], m_fu[1]);
		engine()->fmaddn_inplace(m_uv[1], m_fu, v_coeff, 2);
	}
	// y = (u + v) / 2
	nbcoord_t y_coeff[] = { 1.0 / 2.0, 1.0 / 2.0 };
	engine()->fmaddn(y, nullptr, m_uv, y_coeff, m_uv.size());
	std::swap(m_fu[1], m_fu[0]);
	engine()->advise_time(dt);
}

void nbody_solver_midpoint_stetter::reset()
{
	m_uv_initiated = false;
}

// ---------------------------------------------------------------------------
// rk4: nbody/nbody_solver_rk4.cpp:19-48
// ---------------------------------------------------------------------------
nbody_solver_rk4::nbody_solver_rk4() : m_tmp(nullptr) {}

nbody_solver_rk4::~nbody_solver_rk4()
{
	if(engine() != nullptr)
	{
		engine()->free_buffers(m_k);
		engine()->free_buffer(m_tmp);
	}
}

const char* nbody_solver_rk4::type_name() const { return "nbody_solver_rk4"; }

void nbody_solver_rk4::advise(nbcoord_t dt)
{
	nbody_engine::memory* y = engine()->get_y();
	nbcoord_t t = engine()->get_time();
	if(m_k.empty())
	{
		m_k = engine()->create_buffers(state_bytes(engine()), 4);
		m_tmp = engine()->create_buffer(state_bytes(engine()));
	}
	engine()->fcompute(t, y, m_k[0]);
	engine()->fmadd(m_tmp, y, m_k[0], dt / 2.0);
	engine()->fcompute(t + dt / 2.0, m_tmp, m_k[1]);
	engine()->fmadd(m_tmp, y, m_k[1], dt / 2.0);
	engine()->fcompute(t + dt / 2.0, m_tmp, m_k[2]);
	engine()->fmadd(m_tmp, y, m_k[2], dt);
	engine()->fcompute(t + dt, m_tmp, m_k[3]);
	const nbcoord_t coeff[] = { dt / 6.0, dt / 3.0, dt / 3.0, dt / 6.0 };
	engine()->fmaddn_inplace(y, m_k, coeff, 4);
	engine()->advise_time(dt);
}

// ---------------------------------------------------------------------------
// trapeze: nbody/nbody_solver_trapeze.cpp:26-55
// ---------------------------------------------------------------------------
nbody_solver_trapeze::nbody_solver_trapeze() : m_predictor(nullptr), m_refine_steps_count(1) {}

nbody_solver_trapeze::~nbody_solver_trapeze()
{
	if(engine() != nullptr)
	{
		engine()->free_buffers(m_f);
		engine()->free_buffer(m_predictor);
	}
}

const char* nbody_solver_trapeze::type_name() const { return "nbody_solver_trapeze"; }
void nbody_solver_trapeze::set_refine_steps_count(size_t v) { m_refine_steps_count = v; }

void nbody_solver_trapeze::advise(nbcoord_t dt)
{
	nbody_engine::memory* y = engine()->get_y();
	nbcoord_t t = engine()->get_time();
	if(m_f.empty())
	{
		m_f = engine()->create_buffers(state_bytes(engine()), 2);
		m_predictor = engine()->create_buffer(state_bytes(engine()));
	}
	engine()->fcompute(t, y, m_f[0]);
	engine()->fmadd(m_predictor, y, m_f[0], dt);
	for(size_t s = 0; s <= m_refine_steps_count; ++s)
	{
		const nbcoord_t coeff[] = { dt / 2.0, dt / 2.0 };
		engine()->fcompute(t, m_predictor, m_f[1]);
		if(s == m_refine_steps_count)
		{
			engine()->fmaddn_inplace(y, m_f, coeff, 2);
		}
		else
		{
			engine()->fmaddn(m_predictor, y, m_f, coeff, 2);
		}
	}
	engine()->advise_time(dt);
}

void nbody_solver_trapeze::print_info() const
{
	nbody_solver::print_info();
	nbody_print("\trefine_steps_count %zu", m_refine_steps_count);
}

// ---------------------------------------------------------------------------
// adams (Adams-Bashforth, ranks 1..5): nbody/nbody_solver_adams.cpp:30-101
// ---------------------------------------------------------------------------
nbody_solver_adams::nbody_solver_adams(nbody_solver* starter, size_t rank, bool corr) :
	m_starter(starter), m_corr_data(nullptr), m_rank(rank), m_step(0), m_correction(corr)
{
}

nbody_solver_adams::~nbody_solver_adams()
{
	delete m_starter;
	if(engine() != nullptr)
	{
		engine()->free_buffers(m_f);
		engine()->free_buffer(m_corr_data);
	}
}

const char* nbody_solver_adams::type_name() const { return "nbody_solver_adams"; }

void nbody_solver_adams::advise(nbcoord_t dt)
{
	static const nbcoord_t a1[1] = { 1.0 };
	static const nbcoord_t a2[2] = { 3.0 / 2.0, -1.0 / 2.0 };
	static const nbcoord_t a3[3] = { 23.0 / 12.0, -4.0 / 3.0, 5.0 / 12.0 };
	static const nbcoord_t a4[4] = { 55.0 / 24.0, -59.0 / 24.0, 37.0 / 24.0, -3.0 / 8.0 };
	static const nbcoord_t a5[5] = { 1901.0 / 720.0, -1387.0 / 360.0, 109.0 / 30.0, -637.0 / 360.0, 251.0 / 720.0 };
	static const nbcoord_t* ar[] = { nullptr, a1, a2, a3, a4, a5 };
	const nbcoord_t* a = ar[m_rank];

	nbody_engine::memory* y = engine()->get_y();
	nbcoord_t t = engine()->get_time();
	size_t fnum = m_step % m_rank;
	if(m_f.empty())
	{
		m_starter->set_engine(engine());
		m_f = engine()->create_buffers(state_bytes(engine()), m_rank);
		if(m_correction)
		{
			m_corr_data = engine()->create_buffer(state_bytes(engine()));
			engine()->fill_buffer(m_corr_data, 0);
		}
	}
	if(m_step > m_rank)
	{
		std::vector<nbcoord_t> coeff(m_rank);
		engine()->fcompute(t, y, m_f[fnum]);
		for(size_t n = 0; n < m_rank; ++n)
		{
			coeff[(m_rank + fnum - n) % m_rank] = a[n] * dt;
		}
		if(m_correction)
		{
			engine()->fmaddn_corr(y, m_corr_data, m_f, coeff.data(), m_rank);
		}
		else
		{
			engine()->fmaddn_inplace(y, m_f, coeff.data(), m_rank);
		}
		engine()->advise_time(dt);
	}
	else
	{
		engine()->fcompute(t, y, m_f[fnum]);
		m_starter->advise(dt);
	}
	++m_step;
}

void nbody_solver_adams::print_info() const
{
	nbody_solver::print_info();
	nbody_print("\trank %zu", m_rank);
	nbody_print("\tcorrection %d", m_correction ? 1 : 0);
	nbody_print("\tstarter %s", m_starter->type_name());
}

void nbody_solver_adams::reset()
{
	if(m_corr_data != nullptr)
	{
		engine()->fill_buffer(m_corr_data, 0);
	}
	m_step = 0;
}

// ---------------------------------------------------------------------------
// generic Runge-Kutta on a Butcher table: nbody/nbody_solver_rk_butcher.cpp
// ---------------------------------------------------------------------------
nbody_solver_rk_butcher::nbody_solver_rk_butcher(nbody_butcher_table* t) :
	m_bt(t),
	m_type_name(std::string("nbody_solver_") + t->name()),
	m_t(nullptr),
	m_tmpk(nullptr),
	m_ycorr_data(nullptr),
	m_tcorr_data(nullptr),
	m_max_recursion(8),
	m_substep_subdivisions(8),
	m_error_threshold(1e-4),
	m_refine_steps_count(1),
	m_correction(false),
	m_substeps_done(0)
{
}

nbody_solver_rk_butcher::~nbody_solver_rk_butcher()
{
	delete m_bt;
	if(engine() != nullptr)
	{
		engine()->free_buffers(m_k);
		engine()->free_buffer(m_t);
		engine()->free_buffer(m_tmpk);
		engine()->free_buffer(m_ycorr_data);
		engine()->free_buffer(m_tcorr_data);
		engine()->free_buffers(m_y_stack);
	}
}

const char* nbody_solver_rk_butcher::type_name() const { return m_type_name.c_str(); }
void nbody_solver_rk_butcher::set_max_recursion(size_t v) { m_max_recursion = v; }
void nbody_solver_rk_butcher::set_substep_subdivisions(size_t v) { m_substep_subdivisions = v; }
void nbody_solver_rk_butcher::set_error_threshold(nbcoord_t v) { m_error_threshold = v; }
void nbody_solver_rk_butcher::set_refine_steps_count(size_t v) { m_refine_steps_count = v; }
void nbody_solver_rk_butcher::set_correction(bool corr) { m_correction = corr; }
const nbody_butcher_table* nbody_solver_rk_butcher::table() const { return m_bt; }

// :60-68
void nbody_solver_rk_butcher::advise(nbcoord_t dt)
{
	nbody_engine::memory* y = engine()->get_y();
	nbcoord_t t = engine()->get_time();
	sub_step(1, t, dt, y, 0);
	engine()->advise_time(dt);
}

void nbody_solver_rk_butcher::print_info() const
{
	nbody_solver::print_info();
	nbody_print("\tmax_recursion %zu", m_max_recursion);
	nbody_print("\tsubstep_subdivisions %zu", m_substep_subdivisions);
	nbody_print("\terror_threshold %g", m_error_threshold);
	nbody_print("\trefine_steps_count %zu", m_refine_steps_count);
	nbody_print("\tcorrection %d", m_correction ? 1 : 0);
}

void nbody_solver_rk_butcher::reset()
{
	if(m_ycorr_data != nullptr) { engine()->fill_buffer(m_ycorr_data, 0); }
	if(m_tcorr_data != nullptr) { engine()->fill_buffer(m_tcorr_data, 0); }
}

// :97-128
void nbody_solver_rk_butcher::sub_step_implicit(size_t steps, const nbcoord_t** a, nbcoord_t* coeff,
                                                const nbody_engine::memory* y, bool need_first_approach_k,
                                                const nbcoord_t* c, nbcoord_t t, nbcoord_t dt)
{
	if(need_first_approach_k)
	{
		engine()->fcompute(t, y, m_tmpk);
		for(size_t i = 0; i != steps; ++i)
		{
			engine()->fmadd(m_t, y, m_tmpk, dt * c[i]);
			engine()->fcompute(t + c[i] * dt, m_t, m_k[i]);
		}
	}
	for(size_t iter = 0; iter != m_refine_steps_count; ++iter)
	{
		for(size_t i = 0; i != steps; ++i)
		{
			for(size_t n = 0; n != steps; ++n)
			{
				coeff[n] = dt * a[i][n];
			}
			engine()->fmaddn(m_t, y, m_k, coeff, m_k.size());
			engine()->fcompute(t + c[i] * dt, m_t, m_k[i]);
		}
	}
}

// :130-161
void nbody_solver_rk_butcher::sub_step_explicit(size_t steps, const nbcoord_t** a, nbcoord_t* coeff,
                                                const nbody_engine::memory* y, const nbcoord_t* c,
                                                nbcoord_t t, nbcoord_t dt)
{
	for(size_t i = 0; i < steps; ++i)
	{
		if(i == 0)
		{
			engine()->fcompute(t + c[i] * dt, y, m_k[i]);
		}
		else
		{
			for(size_t n = 0; n != i; ++n)
			{
				coeff[n] = dt * a[i][n];
			}
			if(m_correction)
			{
				engine()->copy_buffer(m_tcorr_data, m_ycorr_data);
				engine()->copy_buffer(m_t, y);
				engine()->fmaddn_corr(m_t, m_tcorr_data, m_k, coeff, i);
			}
			else
			{
				engine()->fmaddn(m_t, y, m_k, coeff, i);
			}
			engine()->fcompute(t + c[i] * dt, m_t, m_k[i]);
		}
	}
}

// :163-247 -- embedded tables split a step whose error estimate
// max|sum((b2 - b1) k)| exceeds error_threshold into substep_subdivisions
// substeps, recursively up to max_recursion levels.
void nbody_solver_rk_butcher::sub_step(size_t substeps_count, nbcoord_t t, nbcoord_t dt,
                                       nbody_engine::memory* y, size_t recursion_level)
{
	const size_t steps = m_bt->get_steps();
	const nbcoord_t** a = m_bt->get_a();
	const nbcoord_t* b1 = m_bt->get_b1();
	const nbcoord_t* b2 = m_bt->get_b2();
	const nbcoord_t* c = m_bt->get_c();
	size_t ps = engine()->problem_size();
	bool need_first_approach_k = false;
	std::vector<nbcoord_t> coeff(steps + 1);

	if(m_k.empty())
	{
		need_first_approach_k = true;
		m_k = engine()->create_buffers(sizeof(nbcoord_t) * ps, steps);
		m_t = engine()->create_buffer(sizeof(nbcoord_t) * ps);
		m_tmpk = engine()->create_buffer(sizeof(nbcoord_t) * ps);
		m_y_stack = engine()->create_buffers(sizeof(nbcoord_t) * ps, m_max_recursion);
		if(m_correction)
		{
			m_ycorr_data = engine()->create_buffer(sizeof(nbcoord_t) * ps);
			m_tcorr_data = engine()->create_buffer(sizeof(nbcoord_t) * ps);
			engine()->fill_buffer(m_ycorr_data, 0);
			engine()->fill_buffer(m_tcorr_data, 0);
		}
	}

	for(size_t sub_n = 0; sub_n != substeps_count; ++sub_n, t += dt)
	{
		if(m_bt->is_implicit())
		{
			sub_step_implicit(steps, a, coeff.data(), y, need_first_approach_k, c, t, dt);
		}
		else
		{
			sub_step_explicit(steps, a, coeff.data(), y, c, t, dt);
		}

		nbcoord_t max_error = 0;
		if(m_bt->is_embedded())
		{
			for(size_t n = 0; n != steps; ++n)
			{
				coeff[n] = (b2[n] - b1[n]);
			}
			engine()->fmaddn(m_t, nullptr, m_k, coeff.data(), steps);
			engine()->fmaxabs(m_t, max_error);
		}

		bool can_subdivide = (m_bt->is_embedded() && recursion_level < m_max_recursion) && dt > get_min_step();
		bool need_subdivide = max_error > m_error_threshold;

		if(can_subdivide && need_subdivide)
		{
			nbcoord_t new_dt = dt / static_cast<nbcoord_t>(m_substep_subdivisions);
			nbody_engine::memory* curr_y = m_y_stack[recursion_level];
			engine()->copy_buffer(curr_y, y);
			sub_step(m_substep_subdivisions, t, new_dt, curr_y, recursion_level + 1);
			engine()->copy_buffer(y, curr_y);
		}
		else
		{
			for(size_t n = 0; n != steps; ++n)
			{
				coeff[n] = b2[n] * dt;
			}
			if(m_correction)
			{
				engine()->fmaddn_corr(y, m_ycorr_data, m_k, coeff.data(), m_k.size());
			}
			else
			{
				engine()->fmaddn_inplace(y, m_k, coeff.data(), m_k.size());
			}
			++m_substeps_done;
		}
	}
}

// ---------------------------------------------------------------------------
// extrapolators: nbody/nbody_extrapolator.cpp
// ---------------------------------------------------------------------------
nbody_extrapolator::nbody_extrapolator(nbody_engine* e, size_t order, const std::vector<size_t>& substeps_count) :
	m_e(e), m_sub_steps_count(substeps_count)
{
	if(order != 1)
	{
		for(size_t& s : m_sub_steps_count)
		{
			s = static_cast<size_t>(std::pow(s, order));
		}
	}
}

nbody_extrapolator::~nbody_extrapolator() {}
nbody_engine* nbody_extrapolator::engine() const { return m_e; }
size_t nbody_extrapolator::sub_steps_count(size_t level) const { return m_sub_steps_count[level]; }

nbody_extrapolator_berrut::nbody_extrapolator_berrut(nbody_engine* e, size_t order,
                                                     const std::vector<size_t>& substeps_count) :
	nbody_extrapolator(e, order, substeps_count)
{
	size_t size = sizeof(nbcoord_t) * e->problem_size();
	m_table = e->create_buffers(size, substeps_count.size());
	m_diff = e->create_buffer(size);
}

nbody_extrapolator_berrut::~nbody_extrapolator_berrut()
{
	engine()->free_buffers(m_table);
	engine()->free_buffer(m_diff);
}

void nbody_extrapolator_berrut::update_table(size_t level, nbody_engine::memory* y)
{
	engine()->copy_buffer(m_table[level], y);
}

nbcoord_t nbody_extrapolator_berrut::estimate_error(size_t level) const
{
	if(level == 0)
	{
		return std::numeric_limits<nbcoord_t>::max();
	}
	nbcoord_t maxdiff = 0;
	const std::vector<nbcoord_t> w0(weights(level - 1));
	std::vector<nbcoord_t> w1(weights(level));
	for(size_t n = 0; n != w0.size(); ++n)
	{
		w1[n] -= w0[n];
	}
	engine()->fill_buffer(m_diff, 0);
	engine()->fmaddn_inplace(m_diff, m_table, w1.data(), level);
	engine()->fmaxabs(m_diff, maxdiff);
	return maxdiff;
}

void nbody_extrapolator_berrut::extrapolate(size_t level, nbody_engine::memory* ext_y) const
{
	const std::vector<nbcoord_t> w(weights(level));
	engine()->fill_buffer(ext_y, 0);
	if(level == 0)
	{
		return;
	}
	engine()->fmaddn_inplace(ext_y, m_table, w.data(), level);
}

std::vector<nbcoord_t> nbody_extrapolator_berrut::weights(size_t level) const
{
	std::vector<nbcoord_t> w(level);
	for(size_t n = 0; n != level; ++n)
	{
		w[n] = std::pow(-1, n + 1) * static_cast<nbcoord_t>(sub_steps_count(n));
	}
	nbcoord_t sum = std::accumulate(w.begin(), w.end(), 0.0);
	for(size_t n = 0; n != level; ++n)
	{
		w[n] /= sum;
	}
	return w;
}

nbody_extrapolator_neville::nbody_extrapolator_neville(nbody_engine* e, size_t order,
                                                       const std::vector<size_t>& substeps_count) :
	nbody_extrapolator(e, order, substeps_count)
{
	size_t size = sizeof(nbcoord_t) * e->problem_size();
	m_table = e->create_buffers(size, substeps_count.size());
	m_table_prev = e->create_buffers(size, substeps_count.size());
	m_diff = e->create_buffer(size);
}

nbody_extrapolator_neville::~nbody_extrapolator_neville()
{
	engine()->free_buffers(m_table);
	engine()->free_buffers(m_table_prev);
	engine()->free_buffer(m_diff);
}

// T_{k,j+1} = (T_{k,j} n_k - T_{k-1,j} n_{k-j-1}) / (n_k - n_{k-j-1}) with the
// squared substep counts n (order 2), row k = y.
void nbody_extrapolator_neville::update_table(size_t k, nbody_engine::memory* y)
{
	std::swap(m_table, m_table_prev);
	engine()->copy_buffer(m_table[0], y);
	if(k == 0)
	{
		return;
	}
	for(size_t j = 0; j != k; ++j)
	{
		nbcoord_t step1(static_cast<nbcoord_t>(sub_steps_count(k)));
		nbcoord_t step2(static_cast<nbcoord_t>(sub_steps_count(k - j - 1)));
		nbcoord_t step_diff = step1 - step2;
		nbcoord_t d[2] = { step1 / step_diff, -step2 / step_diff };
		nbody_engine::memory_array c = { m_table[j], m_table_prev[j] };
		engine()->fmaddn(m_table[j + 1], nullptr, c, d, 2);
	}
}

nbcoord_t nbody_extrapolator_neville::estimate_error(size_t level) const
{
	if(level == 0)
	{
		return std::numeric_limits<nbcoord_t>::max();
	}
	nbcoord_t maxdiff = 0;
	engine()->fmadd(m_diff, m_table[level], m_table[level - 1], -1.0);
	engine()->fmaxabs(m_diff, maxdiff);
	return maxdiff;
}

void nbody_extrapolator_neville::extrapolate(size_t level, nbody_engine::memory* ext_y) const
{
	engine()->copy_buffer(ext_y, m_table[level]);
}

nbody_extrapolator* nbody_create_extrapolator(const std::string& type, nbody_engine* engine, size_t order,
                                              const std::vector<size_t>& substeps_count)
{
	if(type == "berrut")
	{
		return new nbody_extrapolator_berrut(engine, order, substeps_count);
	}
	if(type == "neville")
	{
		return new nbody_extrapolator_neville(engine, order, substeps_count);
	}
	return nullptr;
}

// ---------------------------------------------------------------------------
// Bulirsch-Stoer: nbody/nbody_solver_bulirsch_stoer.cpp
// ---------------------------------------------------------------------------
namespace {
// Substep counts 2, 4, 6, 8, 12, 16, 24, 32, ... (:7-38)
std::vector<size_t> bs_steps(size_t n)
{
	std::vector<size_t> steps;
	if(n == 0) { return steps; }
	steps.push_back(2);
	if(n > 1) { steps.push_back(4); }
	if(n > 2) { steps.push_back(6); }
	for(size_t j = 3; j < n; ++j)
	{
		steps.push_back(2 * steps[j - 2]);
	}
	return steps;
}
}  // namespace

nbody_solver_bulirsch_stoer::nbody_solver_bulirsch_stoer(size_t max_level, nbcoord_t error_threshold) :
	m_internal(nullptr),
	m_max_level(max_level),
	m_error_threshold(error_threshold),
	m_sub_steps_count(bs_steps(max_level)),
	m_y0(nullptr),
	m_extrapolator(nullptr),
	m_last_level(0)
{
}

nbody_solver_bulirsch_stoer::~nbody_solver_bulirsch_stoer()
{
	if(engine() != nullptr)
	{
		engine()->free_buffer(m_y0);
	}
	delete m_extrapolator;
	delete m_internal;
}

const char* nbody_solver_bulirsch_stoer::type_name() const { return "nbody_solver_bulirsch_stoer"; }

// :65-83
void nbody_solver_bulirsch_stoer::compute_substep(size_t level, nbcoord_t dt, nbcoord_t t0)
{
	nbody_engine::memory* y = engine()->get_y();
	engine()->copy_buffer(y, m_y0);
	engine()->set_time(t0);
	size_t substeps_count = m_sub_steps_count[level];
	nbcoord_t substep_dt = dt / static_cast<nbcoord_t>(substeps_count);
	m_internal->reset();
	for(size_t substep_n = 0; substep_n != substeps_count; ++substep_n)
	{
		m_internal->advise(substep_dt);
	}
	m_extrapolator->update_table(level, y);
}

// :85-122
void nbody_solver_bulirsch_stoer::advise(nbcoord_t dt)
{
	if(m_y0 == nullptr)
	{
		m_y0 = engine()->create_buffer(state_bytes(engine()));
		m_extrapolator = nbody_create_extrapolator("neville", engine(), 2, m_sub_steps_count);
		m_internal = new nbody_solver_midpoint_stetter();
		m_internal->set_engine(engine());
	}
	nbody_engine::memory* y = engine()->get_y();
	nbcoord_t t0(engine()->get_time());
	engine()->copy_buffer(m_y0, y);
	for(size_t level = 0; level != m_max_level; ++level)
	{
		compute_substep(level, dt, t0);
		if(level < 2)
		{
			continue;
		}
		nbcoord_t maxdiff = m_extrapolator->estimate_error(level);
		if(maxdiff < m_error_threshold || level == m_max_level - 1)
		{
			m_extrapolator->extrapolate(level, y);
			m_last_level = level;
			break;
		}
	}
	engine()->set_time(t0);
	engine()->advise_time(dt);
}

void nbody_solver_bulirsch_stoer::print_info() const
{
	nbody_solver::print_info();
	nbody_print("\tmax_level %zu", m_max_level);
	nbody_print("\terror_threshold %g", m_error_threshold);
}

// ---------------------------------------------------------------------------
// factory: nbody/nbody_solvers.cpp:3-110
// ---------------------------------------------------------------------------
static nbody_solver* create_butcher_solver(const std::string& table, const nbody_params& param)
{
	double error_threshold = 0;
	long long max_recursion = 0, refine = 0, subdiv = 0;
	bool correction = false;
	if(!param.get_double("error_threshold", 1e-4, &error_threshold) ||
	   !param.get_long("max_recursion", 8, &max_recursion) ||
	   !param.get_long("refine_steps_count", 1, &refine) ||
	   !param.get_long("substep_subdivisions", 8, &subdiv) ||
	   !param.get_bool("correction", false, &correction) ||
	   max_recursion < 0 || refine < 0 || subdiv < 1)
	{
		nbody_log("invalid Runge-Kutta solver parameters");
		return nullptr;
	}
	nbody_butcher_table* bt = nbody_create_butcher_table(table);
	if(bt == nullptr)
	{
		return nullptr;
	}
	nbody_solver_rk_butcher* solver = new nbody_solver_rk_butcher(bt);
	solver->set_error_threshold(error_threshold);
	solver->set_max_recursion(static_cast<size_t>(max_recursion));
	solver->set_refine_steps_count(static_cast<size_t>(refine));
	solver->set_substep_subdivisions(static_cast<size_t>(subdiv));
	solver->set_correction(correction);
	return solver;
}

nbody_solver* nbody_create_solver(const nbody_params& param)
{
	const std::string type(param.value("solver", ""));
	nbody_solver* solver = nullptr;
	if(type == "adams")
	{
		nbody_params starter_param(param);
		starter_param.set("solver", param.value("starter_solver", "euler"));
		long long rank = 0;
		bool corr = false;
		if(!param.get_long("rank", 1, &rank) || rank < 1 || rank > 5 ||
		   !param.get_bool("correction", false, &corr))
		{
			nbody_log("invalid adams parameters (rank must be 1..5)");
			return nullptr;
		}
		nbody_solver* starter = nbody_create_solver(starter_param);
		if(starter == nullptr)
		{
			return nullptr;
		}
		solver = new nbody_solver_adams(starter, static_cast<size_t>(rank), corr);
	}
	else if(type == "bs")
	{
		long long max_level = 0;
		double thr = 0;
		if(!param.get_long("max_level", 8, &max_level) || max_level < 1 ||
		   !param.get_double("error_threshold", 1e-11, &thr))
		{
			nbody_log("invalid bs parameters");
			return nullptr;
		}
		solver = new nbody_solver_bulirsch_stoer(static_cast<size_t>(max_level), thr);
	}
	else if(type == "euler") { solver = new nbody_solver_euler(); }
	else if(type == "midpoint") { solver = new nbody_solver_midpoint(); }
	else if(type == "midpoint-st") { solver = new nbody_solver_midpoint_stetter(); }
	else if(type == "rk4") { solver = new nbody_solver_rk4(); }
	else if(type == "rkck" || type == "rkdp" || type == "rkdverk" || type == "rkf" ||
	        type == "rkfeagin10" || type == "rkfeagin12" || type == "rkfeagin14" ||
	        type == "rkgl" || type == "rklc")
	{
		solver = create_butcher_solver(type, param);
		if(solver == nullptr) { return nullptr; }
	}
	else if(type == "trapeze")
	{
		long long refine = 0;
		if(!param.get_long("refine_steps_count", 1, &refine) || refine < 0)
		{
			nbody_log("invalid trapeze parameters");
			return nullptr;
		}
		nbody_solver_trapeze* trapeze = new nbody_solver_trapeze();
		trapeze->set_refine_steps_count(static_cast<size_t>(refine));
		solver = trapeze;
	}
	else
	{
		return nullptr;
	}
	double min_step = 0, max_step = 0;
	bool clamp = false;
	if(!param.get_double("min_step", 1e-9, &min_step) || !param.get_double("max_step", 1e-2, &max_step) ||
	   !param.get_bool("clamp", false, &clamp))
	{
		nbody_log("invalid step parameters");
		delete solver;
		return nullptr;
	}
	solver->set_time_step(min_step, max_step);
	solver->set_clamp_to_box(clamp);
	return solver;
}
