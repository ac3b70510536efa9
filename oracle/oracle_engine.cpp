// oracle_engine.cpp -- TEST INFRASTRUCTURE.  The CPU restatement of the
// reference's `simple`, `openmp`, `block`, `ah` and `simple_bh` engines
// (nbody_oracle.c) behind the
// product's nbody_engine interface, so the product's solvers can be run on the
// oracle and compared with the HIP engines.  Loaded only by tests/ (and by the
// cpu_baseline leg of bench.py); nothing in the product links it.
//
// Buffer ops follow nbody/nbody_engine_simple.cpp:201-419 (problem_size()
// elements, except fill_buffer which covers the whole buffer); fmaddn and
// fmaddn_inplace use the base-class decomposition (nbody/nbody_engine.cpp:47-113).
#include <cstdlib>
#include <cstring>

#include "nbody_engine.h"
#include "nbody_host.h"
#include "nbody_log.h"
#include "nbody_oracle.h"

namespace {

class oracle_engine : public nbody_engine
{
public:
	class smemory : public memory
	{
		void*  m_data;
		size_t m_size;
	public:
		explicit smemory(size_t s) : m_data(std::malloc(s == 0 ? 1 : s)), m_size(s) {}
		~smemory() override { std::free(m_data); }
		double* data() { return static_cast<double*>(m_data); }
		const double* data() const { return static_cast<const double*>(m_data); }
		size_t size() const override { return m_size; }
	};

private:
	nbody_data* m_data;
	smemory*    m_mass;
	smemory*    m_y;
	bool        m_bh;
	double      m_ratio;
	int         m_layout;     // 1 heap, 2 heap_stackless
	int         m_traverse;   // 0 cycle, 1 nested_tree
	size_t      m_rate;
	oracle_heap m_heap;
	bool        m_built;
	int         m_threads;
	int         m_direct;     // direct sums: 0 simple/openmp, 1 block, 2 ah
	oracle_ah   m_ah;
	double      m_ah_dist, m_ah_force;

public:
	oracle_engine(bool bh, double ratio, int layout, int traverse, size_t rate, int threads, int direct = 0,
	              double ah_dist = 0, double ah_force = 0) :
		m_data(nullptr), m_mass(nullptr), m_y(nullptr), m_bh(bh), m_ratio(ratio), m_layout(layout),
		m_traverse(traverse), m_rate(rate), m_built(false), m_threads(threads), m_direct(direct),
		m_ah_dist(ah_dist), m_ah_force(ah_force)
	{
		std::memset(&m_heap, 0, sizeof(m_heap));
		std::memset(&m_ah, 0, sizeof(m_ah));
	}

	~oracle_engine() override
	{
		delete m_mass;
		delete m_y;
		oracle_heap_free(&m_heap);
		oracle_ah_free(&m_ah);
	}

	const oracle_ah* ah() const { return &m_ah; }

	const oracle_heap* heap() const { return &m_heap; }

	const char* type_name() const override
	{
		if(m_bh) { return "oracle_engine_simple_bh"; }
		return m_direct == 1 ? "oracle_engine_block" : m_direct == 2 ? "oracle_engine_ah" : "oracle_engine_simple";
	}

	bool init(nbody_data* data) override
	{
		m_data = data;
		size_t n = data->get_count();
		delete m_mass;
		delete m_y;
		m_mass = new smemory(sizeof(double) * n);
		m_y = new smemory(sizeof(double) * 6 * n);
		double* y = m_y->data();
		for(size_t i = 0; i < n; ++i)
		{
			y[i] = data->get_vertites()[i].x;
			y[n + i] = data->get_vertites()[i].y;
			y[2 * n + i] = data->get_vertites()[i].z;
			y[3 * n + i] = data->get_velosites()[i].x;
			y[4 * n + i] = data->get_velosites()[i].y;
			y[5 * n + i] = data->get_velosites()[i].z;
			m_mass->data()[i] = data->get_mass()[i];
		}
		if(m_direct == 2)
		{
			oracle_ah_free(&m_ah);
			if(oracle_ah_alloc(&m_ah, n, m_ah_dist, m_ah_force) != 0)
			{
				nbody_log("oracle: out of memory");
				return false;
			}
		}
		if(m_bh)
		{
			oracle_heap_free(&m_heap);
			if(oracle_heap_alloc(&m_heap, n) != 0)
			{
				nbody_log("oracle: body count must be a power of two");
				return false;
			}
			m_built = false;
		}
		return true;
	}

	void get_data(nbody_data* data) override
	{
		size_t n = m_data->get_count();
		const double* y = m_y->data();
		for(size_t i = 0; i < n; ++i)
		{
			data->get_vertites()[i] = nbvertex_t(y[i], y[n + i], y[2 * n + i]);
			data->get_velosites()[i] = nbvertex_t(y[3 * n + i], y[4 * n + i], y[5 * n + i]);
		}
	}

	size_t problem_size() const override { return 6 * m_data->get_count(); }
	memory* get_y() override { return m_y; }
	void advise_time(const nbcoord_t& dt) override { m_data->advise_time(dt); }
	nbcoord_t get_time() const override { return m_data->get_time(); }
	void set_time(nbcoord_t t) override { m_data->set_time(t); }
	size_t get_step() const override { return m_data->get_step(); }
	void set_step(size_t s) override { m_data->set_step(s); }

	void fcompute(const nbcoord_t&, const memory* _y, memory* _f) override
	{
		const smemory* y = dynamic_cast<const smemory*>(_y);
		smemory* f = dynamic_cast<smemory*>(_f);
		if(y == nullptr) { nbody_log("y is not smemory"); return; }
		if(f == nullptr) { nbody_log("f is not smemory"); return; }
		advise_compute_count();
		size_t n = m_data->get_count();
		if(m_direct == 1)
		{
			oracle_fcompute_block(n, y->data(), m_mass->data(), f->data(), m_threads);
			return;
		}
		if(m_direct == 2)
		{
			// nbody/nbody_engine_ah.cpp:16-42: full recompute every rate-th step
			if(m_data->get_step() % m_rate == 0)
			{
				if(oracle_ah_full(&m_ah, y->data(), m_mass->data(), f->data()) != 0)
				{
					nbody_log("oracle: out of memory");
				}
			}
			else
			{
				oracle_ah_sparse(&m_ah, y->data(), m_mass->data(), f->data());
			}
			return;
		}
		if(!m_bh)
		{
			if(m_threads == 1)
			{
				oracle_fcompute_simple(n, y->data(), m_mass->data(), f->data());
			}
			else
			{
				oracle_fcompute_openmp(n, y->data(), m_mass->data(), f->data(), m_threads);
			}
			return;
		}
		const double* yy = y->data();
		if(m_rate == 0 || !m_built || (m_data->get_step() % m_rate) == 0)
		{
			oracle_heap_build(&m_heap, yy, yy + n, yy + 2 * n, m_mass->data(), m_ratio);
			m_built = true;
		}
		else
		{
			oracle_heap_rebuild(&m_heap, yy, yy + n, yy + 2 * n, m_ratio);
		}
		oracle_bh_fcompute(&m_heap, m_layout, m_traverse, n, yy, m_mass->data(), f->data(), m_threads);
	}

	void clamp(memory* _y, nbcoord_t b) override
	{
		smemory* y = dynamic_cast<smemory*>(_y);
		if(y == nullptr) { nbody_log("y is not smemory"); return; }
		oracle_clamp(y->data(), m_data->get_count(), b);
	}

	smemory* create_buffer(size_t s) override { return new smemory(s); }
	void free_buffer(memory* m) override { delete m; }

	void read_buffer(void* dst, const memory* _src) override
	{
		const smemory* src = dynamic_cast<const smemory*>(_src);
		if(src == nullptr) { nbody_log("src is not smemory"); return; }
		std::memcpy(dst, src->data(), src->size());
	}

	void write_buffer(memory* _dst, const void* src) override
	{
		smemory* dst = dynamic_cast<smemory*>(_dst);
		if(dst == nullptr) { nbody_log("dst is not smemory"); return; }
		std::memcpy(dst->data(), src, dst->size());
	}

	void copy_buffer(memory* _a, const memory* _b) override
	{
		smemory* a = dynamic_cast<smemory*>(_a);
		const smemory* b = dynamic_cast<const smemory*>(_b);
		if(a == nullptr) { nbody_log("a is not smemory"); return; }
		if(b == nullptr) { nbody_log("b is not smemory"); return; }
		std::memcpy(a->data(), b->data(), sizeof(double) * problem_size());
	}

	void fill_buffer(memory* _a, const nbcoord_t& value) override
	{
		smemory* a = dynamic_cast<smemory*>(_a);
		if(a == nullptr) { nbody_log("a is not smemory"); return; }
		oracle_fill(a->data(), value, a->size() / sizeof(double));
	}

	void fmadd_inplace(memory* _a, const memory* _b, const nbcoord_t& c) override
	{
		smemory* a = dynamic_cast<smemory*>(_a);
		const smemory* b = dynamic_cast<const smemory*>(_b);
		if(a == nullptr) { nbody_log("a is not smemory"); return; }
		if(b == nullptr) { nbody_log("b is not smemory"); return; }
		oracle_fmadd_inplace(a->data(), b->data(), c, problem_size());
	}

	void fmadd(memory* _a, const memory* _b, const memory* _c, const nbcoord_t& d) override
	{
		smemory* a = dynamic_cast<smemory*>(_a);
		const smemory* b = dynamic_cast<const smemory*>(_b);
		const smemory* c = dynamic_cast<const smemory*>(_c);
		if(a == nullptr) { nbody_log("a is not smemory"); return; }
		if(b == nullptr) { nbody_log("b is not smemory"); return; }
		if(c == nullptr) { nbody_log("c is not smemory"); return; }
		oracle_fmadd(a->data(), b->data(), c->data(), d, problem_size());
	}

	void fmaddn_corr(memory* _a, memory* _corr, const memory_array& _b, const nbcoord_t* c, size_t csize) override
	{
		smemory* a = dynamic_cast<smemory*>(_a);
		smemory* corr = dynamic_cast<smemory*>(_corr);
		if(a == nullptr) { nbody_log("a is not smemory"); return; }
		if(corr == nullptr) { nbody_log("corr is not smemory"); return; }
		if(c == nullptr) { nbody_log("c is not smemory"); return; }
		if(csize > _b.size()) { nbody_log("csize > b.size()"); return; }
		std::vector<const double*> b;
		for(auto i : _b)
		{
			const smemory* m = dynamic_cast<const smemory*>(i);
			if(m == nullptr) { nbody_log("b is not smemory"); return; }
			b.push_back(m->data());
		}
		oracle_fmaddn_corr(a->data(), corr->data(), b.data(), c, csize, problem_size());
	}

	void fmaxabs(const memory* _a, nbcoord_t& result) override
	{
		const smemory* a = dynamic_cast<const smemory*>(_a);
		if(a == nullptr) { nbody_log("a is not smemory"); return; }
		result = oracle_fmaxabs(a->data(), problem_size());
	}
};

}  // namespace

extern "C" {

nbs_engine* oracle_engine_create_simple(int threads)
{
	return reinterpret_cast<nbs_engine*>(static_cast<nbody_engine*>(
		new oracle_engine(false, 0, 0, 0, 0, threads)));
}

nbs_engine* oracle_engine_create_simple_bh(double ratio, int layout, int traverse, size_t rate, int threads)
{
	return reinterpret_cast<nbs_engine*>(static_cast<nbody_engine*>(
		new oracle_engine(true, ratio, layout, traverse, rate, threads)));
}

// nbody/nbody_engines.cpp:16-19 (`block`, OpenMP over target tiles)
nbs_engine* oracle_engine_create_block(int threads)
{
	return reinterpret_cast<nbs_engine*>(static_cast<nbody_engine*>(
		new oracle_engine(false, 0, 0, 0, 0, threads, 1)));
}

// nbody/nbody_engines.cpp:8-15 (`ah`: full_recompute_rate, max_dist, min_force)
nbs_engine* oracle_engine_create_ah(size_t rate, double max_dist, double min_force)
{
	if(rate == 0)
	{
		return nullptr;
	}
	return reinterpret_cast<nbs_engine*>(static_cast<nbody_engine*>(
		new oracle_engine(false, 0, 0, 0, rate, 1, 2, max_dist, min_force)));
}

// Neighbour lists of the last full recompute of an `ah` oracle engine: the
// total count, and with non-null arrays the CSR offsets [n + 1] and the list.
size_t oracle_engine_ah_lists(nbs_engine* e, size_t* offs, unsigned* list)
{
	const oracle_engine* o = dynamic_cast<const oracle_engine*>(reinterpret_cast<nbody_engine*>(e));
	if(o == nullptr || o->ah()->offs == nullptr)
	{
		return 0;
	}
	const oracle_ah* h = o->ah();
	if(offs != nullptr) { std::memcpy(offs, h->offs, (h->n + 1) * sizeof(size_t)); }
	if(list != nullptr && h->list != nullptr) { std::memcpy(list, h->list, h->total * sizeof(unsigned)); }
	return h->total;
}

// Copy the last built heap (tree size 2n, xyz interleaved); body_n as int64 (-1 internal).
int oracle_engine_heap(nbs_engine* e, int64_t* body_n, double* cm, double* mass, double* crit_r2,
                       double* bmin, double* bmax)
{
	oracle_engine* o = dynamic_cast<oracle_engine*>(reinterpret_cast<nbody_engine*>(e));
	if(o == nullptr || o->heap()->size == 0)
	{
		return 0;
	}
	const oracle_heap* h = o->heap();
	for(size_t i = 0; i < h->size; ++i)
	{
		if(body_n) { body_n[i] = h->body_n[i] == ORACLE_NO_BODY ? -1 : static_cast<int64_t>(h->body_n[i]); }
		if(mass) { mass[i] = h->mass[i]; }
		if(crit_r2) { crit_r2[i] = h->crit_r2[i]; }
		for(int d = 0; d < 3; ++d)
		{
			if(cm) { cm[3 * i + d] = h->cm[3 * i + d]; }
			if(bmin) { bmin[3 * i + d] = h->bmin[3 * i + d]; }
			if(bmax) { bmax[3 * i + d] = h->bmax[3 * i + d]; }
		}
	}
	return 1;
}

}  // extern "C"
