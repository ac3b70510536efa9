// nbody_engine.h -- the engine interface the solvers drive.
//
// Same virtual methods, argument meaning and error behaviour as the
// reference's nbody_engine (nbody/nbody_engine.h:16-96): solvers see only
// opaque byte-sized `memory` buffers; a wrong buffer type, a null coefficient
// array or csize > b.size() is logged and the call returns without effect.
//
// Licence: this file restates code of drons/nbody (the reference files cited
// in it), which is distributed under the GNU Lesser General Public License,
// version 3; this file is distributed under the same licence.
#pragma once

#include <vector>

#include "nbody_data.h"

enum e_ode_order
{
	eode_first_order = 1,
	eode_second_order = 1
};

class nbody_engine
{
	size_t      m_compute_count;
	e_ode_order m_ode_order;
	std::vector<nbcoord_t> m_snapshot[2];
	nbody_engine(const nbody_engine&) = delete;
	nbody_engine& operator=(const nbody_engine&) = delete;
public:
	class memory
	{
	public:
		memory();
		//! Memory size in bytes
		virtual size_t size() const = 0;
		virtual ~memory();
	};
	typedef std::vector<memory*> memory_array;

public:
	nbody_engine();
	virtual ~nbody_engine();
	virtual const char* type_name() const = 0;
	virtual bool init(nbody_data* data) = 0;
	virtual void get_data(nbody_data* data) = 0;
	virtual size_t problem_size() const = 0;
	virtual memory* get_y() = 0;
	virtual void advise_time(const nbcoord_t& dt) = 0;
	virtual nbcoord_t get_time() const = 0;
	virtual void set_time(nbcoord_t t) = 0;
	virtual size_t get_step() const = 0;
	virtual void set_step(size_t s) = 0;
	//! f = dy/dt at (t, y)
	virtual void fcompute(const nbcoord_t& t, const memory* y, memory* f) = 0;
	virtual void clamp(memory* y, nbcoord_t b) = 0;

	virtual memory* create_buffer(size_t bytes) = 0;
	virtual void free_buffer(memory*) = 0;
	virtual memory_array create_buffers(size_t size, size_t count);
	virtual void free_buffers(memory_array&);
	virtual void read_buffer(void* dst, const memory* src) = 0;
	virtual void write_buffer(memory* dst, const void* src) = 0;
	//! a[i] = b[i]
	virtual void copy_buffer(memory* a, const memory* b) = 0;
	//! a[i] = value
	virtual void fill_buffer(memory* a, const nbcoord_t& value) = 0;
	//! a[i] += b[i]*c
	virtual void fmadd_inplace(memory* a, const memory* b, const nbcoord_t& c) = 0;
	//! a[i] = b[i] + c[i]*d
	virtual void fmadd(memory* a, const memory* b, const memory* c, const nbcoord_t& d) = 0;
	//! a[i] += sum( b[k][i]*c[k], k=[0...csize) )
	virtual void fmaddn_inplace(memory* a, const memory_array& b, const nbcoord_t* c, size_t csize);
	//! a[i] += sum( b[k][i]*c[k], k=[0...csize) ) with Kahan correction
	virtual void fmaddn_corr(memory* a, memory* corr, const memory_array& b, const nbcoord_t* c, size_t csize);
	//! a[i] = b[i] + sum( c[k][i]*d[k], k=[0...dsize) ), b may be NULL
	virtual void fmaddn(memory* a, const memory* b, const memory_array& c, const nbcoord_t* d, size_t dsize);
	//! result = max( fabs(a[k]) )
	virtual void fmaxabs(const memory* a, nbcoord_t& result) = 0;
	virtual void print_info() const;

	//! Optional device-side P/L/E/V: fill out[12] (P, L, KE, PE, CM, M) for
	//! the current state and return true, or return false to let the caller
	//! compute them on the host.  No counterpart in the reference interface.
	virtual bool compute_statistics(bool with_energy, double out[12]);

	//! Frame snapshots for the run loop's dumps (no counterpart in the
	//! reference interface, whose run loop copies with get_data):
	//! snapshot_begin(slot) takes a copy of the state y for host slot 0 or 1;
	//! snapshot_state(slot) returns it -- problem_size() doubles
	//! [rx ry rz vx vy vz] -- once the copy is complete, or NULL on failure.
	//! The defaults copy at once with read_buffer; a device engine lets the
	//! copy to the host overlap the steps that follow.  The two calls of one
	//! slot may come from different threads (snapshot_state from a writer).
	virtual bool snapshot_begin(int slot);
	virtual const nbcoord_t* snapshot_state(int slot);

	void advise_compute_count();
	size_t get_compute_count() const;
	void set_ode_order(e_ode_order);
	e_ode_order get_ode_order() const;
};
