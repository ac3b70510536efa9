// nbody_engine.cpp -- default buffer-array helpers and the default
// decompositions of the multi-term ops into single-term ops, with the
// semantics of nbody/nbody_engine.cpp:21-113 (zero coefficients skipped;
// fmaddn with b == NULL starts from 0; fmaddn_corr ignores corr by default).
//
// Licence: this file restates code of drons/nbody (the reference files cited
// in it), which is distributed under the GNU Lesser General Public License,
// version 3; this file is distributed under the same licence.
#include "nbody_engine.h"

#include "nbody_log.h"

nbody_engine::memory::memory() {}
nbody_engine::memory::~memory() {}

nbody_engine::nbody_engine() : m_compute_count(0), m_ode_order(eode_first_order) {}
nbody_engine::~nbody_engine() {}

nbody_engine::memory_array nbody_engine::create_buffers(size_t size, size_t count)
{
	memory_array mema;
	mema.reserve(count);
	for(size_t n = 0; n != count; ++n)
	{
		memory* m = create_buffer(size);
		if(m == nullptr)
		{
			free_buffers(mema);
			return memory_array();
		}
		mema.push_back(m);
	}
	return mema;
}

void nbody_engine::free_buffers(memory_array& mema)
{
	for(memory* m : mema)
	{
		free_buffer(m);
	}
	mema.clear();
}

void nbody_engine::fmaddn_inplace(memory* a, const memory_array& b, const nbcoord_t* c, size_t csize)
{
	if(c == nullptr)
	{
		return;
	}
	if(csize > b.size())
	{
		nbody_log("csize > b.size()");
		return;
	}
	for(size_t i = 0; i != csize; ++i)
	{
		if(c[i] == 0.0)
		{
			continue;
		}
		fmadd_inplace(a, b[i], c[i]);
	}
}

void nbody_engine::fmaddn_corr(memory* a, memory* corr, const memory_array& b, const nbcoord_t* c, size_t csize)
{
	(void)corr;
	fmaddn_inplace(a, b, c, csize);
}

void nbody_engine::fmaddn(memory* a, const memory* b, const memory_array& c, const nbcoord_t* d, size_t dsize)
{
	if(d == nullptr)
	{
		nbody_log("d == NULL");
		return;
	}
	if(dsize > c.size())
	{
		nbody_log("dsize > c.size()");
		return;
	}
	bool a_initiated = false;
	if(b == nullptr)
	{
		fill_buffer(a, 0);
		a_initiated = true;
	}
	for(size_t i = 0; i != dsize; ++i)
	{
		if(d[i] == 0.0)
		{
			continue;
		}
		if(a_initiated)
		{
			fmadd(a, a, c[i], d[i]);
		}
		else
		{
			fmadd(a, b, c[i], d[i]);
		}
		a_initiated = true;
	}
}

void nbody_engine::print_info() const {}

bool nbody_engine::compute_statistics(bool, double*)
{
	return false;
}

bool nbody_engine::snapshot_begin(int slot)
{
	memory* y = get_y();
	if(slot < 0 || slot > 1 || y == nullptr)
	{
		nbody_log("snapshot_begin: bad slot %d or engine not initialised", slot);
		return false;
	}
	m_snapshot[slot].resize(y->size() / sizeof(nbcoord_t));
	read_buffer(m_snapshot[slot].data(), y);
	return true;
}

const nbcoord_t* nbody_engine::snapshot_state(int slot)
{
	if(slot < 0 || slot > 1 || m_snapshot[slot].empty())
	{
		nbody_log("snapshot_state: slot %d holds no snapshot", slot);
		return nullptr;
	}
	return m_snapshot[slot].data();
}

void nbody_engine::advise_compute_count()
{
	m_compute_count++;
}

size_t nbody_engine::get_compute_count() const
{
	return m_compute_count;
}

void nbody_engine::set_ode_order(e_ode_order order)
{
	m_ode_order = order;
}

e_ode_order nbody_engine::get_ode_order() const
{
	return m_ode_order;
}
