// nbody_engine_hip.cpp -- see nbody_engine_hip.h.
#include "nbody_engine_hip.h"

#include <cstdio>
#include <cstring>
#include <vector>

#include "nbody_log.h"

// ---------------------------------------------------------------------------
// buffers
// ---------------------------------------------------------------------------
nbody_engine_hip::smemory::smemory(nbh_ctx* ctx, size_t size) : m_ctx(ctx), m_data(nullptr), m_size(size)
{
	if(size != 0 && nbh_malloc(ctx, size, &m_data) != NBH_OK)
	{
		nbody_log("hip: allocation of %zu bytes failed: %s", size, nbh_last_error());
		m_data = nullptr;
	}
}

nbody_engine_hip::smemory::~smemory()
{
	if(m_data != nullptr && nbh_free(m_ctx, m_data) != NBH_OK)
	{
		nbody_log("hip: free failed: %s", nbh_last_error());
	}
}

// ---------------------------------------------------------------------------
// engine
// ---------------------------------------------------------------------------
nbody_engine_hip::nbody_engine_hip() :
	m_ctx(nullptr),
	m_device(-1),
	m_block_size(NBODY_DATA_BLOCK_SIZE),
	m_precision(NBH_FP64),
	m_data(nullptr),
	m_mass(nullptr),
	m_y(nullptr),
	m_snap(nullptr),
	m_error(false)
{
}

nbody_engine_hip::~nbody_engine_hip()
{
	nbh_snapshot_destroy(m_snap);
	delete m_mass;
	delete m_y;
	if(m_ctx != nullptr)
	{
		nbh_ctx_destroy(m_ctx);
	}
}

bool nbody_engine_hip::check(nbh_status s, const char* what)
{
	if(s == NBH_OK)
	{
		return true;
	}
	if(s != NBH_ERR_INVALID)
	{
		m_error = true;
	}
	nbody_log("hip: %s: %s", what, nbh_last_error());
	return false;
}

int nbody_engine_hip::select_device(int device)
{
	if(m_ctx != nullptr)
	{
		nbh_ctx_destroy(m_ctx);
		m_ctx = nullptr;
	}
	nbh_status s = nbh_ctx_create(device, &m_ctx);
	if(s != NBH_OK)
	{
		nbody_log("hip: can't use device %d: %s", device, nbh_last_error());
		m_ctx = nullptr;
		return -1;
	}
	m_device = device;
	return 0;
}

void nbody_engine_hip::set_block_size(int block_size)
{
	m_block_size = block_size;
}

void nbody_engine_hip::set_precision(nbh_precision p)
{
	m_precision = p;
}

const char* nbody_engine_hip::type_name() const
{
	return "nbody_engine_hip";
}

// nbody/nbody_engine_cuda.cpp:91-139: AoS host bodies -> SoA device state.
bool nbody_engine_hip::init(nbody_data* data)
{
	if(m_ctx == nullptr)
	{
		nbody_log("hip: no device selected");
		return false;
	}
	m_data = data;
	size_t count = data->get_count();
	nbh_snapshot_destroy(m_snap);
	m_snap = nullptr;
	delete m_mass;
	delete m_y;
	m_mass = create_buffer(sizeof(nbcoord_t) * count);
	m_y = create_buffer(sizeof(nbcoord_t) * problem_size());
	if(m_mass == nullptr || m_y == nullptr)
	{
		return false;
	}
	std::vector<nbcoord_t> ybuf(6 * count);
	const nbvertex_t* vrt = data->get_vertites();
	const nbvertex_t* vel = data->get_velosites();
	for(size_t i = 0; i != count; ++i)
	{
		ybuf[i] = vrt[i].x;
		ybuf[count + i] = vrt[i].y;
		ybuf[2 * count + i] = vrt[i].z;
		ybuf[3 * count + i] = vel[i].x;
		ybuf[4 * count + i] = vel[i].y;
		ybuf[5 * count + i] = vel[i].z;
	}
	write_buffer(m_y, ybuf.data());
	write_buffer(m_mass, data->get_mass());
	return !m_error;
}

// nbody/nbody_engine_cuda.cpp:141-165
void nbody_engine_hip::get_data(nbody_data* data)
{
	if(m_y == nullptr)
	{
		nbody_log("hip: engine not initialised");
		return;
	}
	size_t count = m_data->get_count();
	std::vector<nbcoord_t> ybuf(6 * count);
	read_buffer(ybuf.data(), m_y);
	nbvertex_t* vrt = data->get_vertites();
	nbvertex_t* vel = data->get_velosites();
	for(size_t i = 0; i != count; ++i)
	{
		vrt[i] = nbvertex_t(ybuf[i], ybuf[count + i], ybuf[2 * count + i]);
		vel[i] = nbvertex_t(ybuf[3 * count + i], ybuf[4 * count + i], ybuf[5 * count + i]);
	}
}

size_t nbody_engine_hip::problem_size() const
{
	return m_data == nullptr ? 0 : 6 * m_data->get_count();
}

nbody_engine::memory* nbody_engine_hip::get_y()
{
	return m_y;
}

bool nbody_engine_hip::snapshot_begin(int slot)
{
	if(m_y == nullptr || slot < 0 || slot > 1)
	{
		nbody_log("hip: snapshot_begin: bad slot %d or engine not initialised", slot);
		return false;
	}
	if(m_snap == nullptr && !check(nbh_snapshot_create(m_ctx, m_y->size(), 2, &m_snap), "snapshot allocation"))
	{
		return false;
	}
	return check(nbh_snapshot_begin(m_ctx, m_snap, slot, m_y->data()), "snapshot");
}

// Called from the run loop's writer thread: no error latching here, the
// failure reaches the run loop as a NULL state.
const nbcoord_t* nbody_engine_hip::snapshot_state(int slot)
{
	const void* host = nullptr;
	if(m_snap == nullptr || nbh_snapshot_wait(m_ctx, m_snap, slot, &host) != NBH_OK)
	{
		nbody_log("hip: snapshot of slot %d not available: %s", slot, nbh_last_error());
		return nullptr;
	}
	return static_cast<const nbcoord_t*>(host);
}

void nbody_engine_hip::advise_time(const nbcoord_t& dt)
{
	m_data->advise_time(dt);
}

nbcoord_t nbody_engine_hip::get_time() const
{
	return m_data->get_time();
}

void nbody_engine_hip::set_time(nbcoord_t t)
{
	m_data->set_time(t);
}

size_t nbody_engine_hip::get_step() const
{
	return m_data->get_step();
}

void nbody_engine_hip::set_step(size_t s)
{
	m_data->set_step(s);
}

bool nbody_engine_hip::check_f(const memory* _y, const memory* _f, const smemory** ys, smemory** fs)
{
	const smemory* y = dynamic_cast<const smemory*>(_y);
	smemory* f = dynamic_cast<smemory*>(const_cast<memory*>(_f));
	if(y == nullptr)
	{
		nbody_log("y is not smemory");
		return false;
	}
	if(f == nullptr)
	{
		nbody_log("f is not smemory");
		return false;
	}
	if(m_data == nullptr || m_mass == nullptr)
	{
		nbody_log("hip: engine not initialised");
		return false;
	}
	size_t need = sizeof(nbcoord_t) * problem_size();
	if(y->size() < need || f->size() < need)
	{
		nbody_log("fcompute: buffers smaller than problem size");
		return false;
	}
	if(y == f)
	{
		nbody_log("fcompute: y and f must be different buffers");
		return false;
	}
	*ys = y;
	*fs = f;
	return true;
}

// nbody/nbody_engine_cuda.cpp:202-247
void nbody_engine_hip::fcompute(const nbcoord_t& t, const memory* _y, memory* _f)
{
	(void)t;
	const smemory* y = nullptr;
	smemory* f = nullptr;
	if(!check_f(_y, _f, &y, &f))
	{
		return;
	}
	advise_compute_count();
	check(nbh_fcompute_direct(m_ctx, y->data(), m_mass->data(), f->data(), m_data->get_count(),
	                          m_precision, m_block_size >= 64 && m_block_size % 64 == 0 &&
	                          m_block_size <= 1024 ? m_block_size : 0),
	      "fcompute");
}

void nbody_engine_hip::clamp(memory* _y, nbcoord_t b)
{
	smemory* y = dynamic_cast<smemory*>(_y);
	if(y == nullptr)
	{
		nbody_log("y is not smemory");
		return;
	}
	size_t count = m_data->get_count();
	if(y->count() < 3 * count)
	{
		nbody_log("clamp: buffer smaller than 3*count");
		return;
	}
	check(nbh_clamp(m_ctx, y->data(), count, b), "clamp");
}

nbody_engine_hip::smemory* nbody_engine_hip::create_buffer(size_t bytes)
{
	if(m_ctx == nullptr)
	{
		nbody_log("hip: no device selected");
		return nullptr;
	}
	smemory* m = new smemory(m_ctx, bytes);
	if(!m->valid())
	{
		delete m;
		return nullptr;
	}
	return m;
}

void nbody_engine_hip::free_buffer(memory* m)
{
	delete m;
}

void nbody_engine_hip::read_buffer(void* dst, const memory* _src)
{
	const smemory* src = dynamic_cast<const smemory*>(_src);
	if(src == nullptr)
	{
		nbody_log("src is not smemory");
		return;
	}
	if(dst == nullptr)
	{
		nbody_log("dst is NULL");
		return;
	}
	check(nbh_memcpy_d2h(m_ctx, dst, src->data(), src->size()), "read_buffer");
}

void nbody_engine_hip::write_buffer(memory* _dst, const void* src)
{
	smemory* dst = dynamic_cast<smemory*>(_dst);
	if(dst == nullptr)
	{
		nbody_log("dst is not smemory");
		return;
	}
	if(src == nullptr)
	{
		nbody_log("src is NULL");
		return;
	}
	check(nbh_memcpy_h2d(m_ctx, dst->data(), src, dst->size()), "write_buffer");
}

void nbody_engine_hip::copy_buffer(memory* _a, const memory* _b)
{
	smemory* a = dynamic_cast<smemory*>(_a);
	const smemory* b = dynamic_cast<const smemory*>(_b);
	if(a == nullptr)
	{
		nbody_log("a is not smemory");
		return;
	}
	if(b == nullptr)
	{
		nbody_log("b is not smemory");
		return;
	}
	if(a->size() != b->size())
	{
		nbody_log("Size does not match");
		return;
	}
	check(nbh_memcpy_d2d(m_ctx, a->data(), b->data(), a->size()), "copy_buffer");
}

void nbody_engine_hip::fill_buffer(memory* _a, const nbcoord_t& value)
{
	smemory* a = dynamic_cast<smemory*>(_a);
	if(a == nullptr)
	{
		nbody_log("a is not smemory");
		return;
	}
	check(nbh_fill(m_ctx, a->data(), value, a->count()), "fill_buffer");
}

// a[i] += b[i]*c
void nbody_engine_hip::fmadd_inplace(memory* _a, const memory* _b, const nbcoord_t& c)
{
	smemory* a = dynamic_cast<smemory*>(_a);
	const smemory* b = dynamic_cast<const smemory*>(_b);
	if(a == nullptr)
	{
		nbody_log("a is not smemory");
		return;
	}
	if(b == nullptr)
	{
		nbody_log("b is not smemory");
		return;
	}
	if(b->count() < a->count())
	{
		nbody_log("b is smaller than a");
		return;
	}
	const double* src[1] = { b->data() };
	double coef[1] = { c };
	check(nbh_fmaddn(m_ctx, a->data(), a->data(), src, coef, 1, a->count()), "fmadd_inplace");
}

// a[i] = b[i] + c[i]*d
void nbody_engine_hip::fmadd(memory* _a, const memory* _b, const memory* _c, const nbcoord_t& d)
{
	smemory* a = dynamic_cast<smemory*>(_a);
	const smemory* b = dynamic_cast<const smemory*>(_b);
	const smemory* c = dynamic_cast<const smemory*>(_c);
	if(a == nullptr)
	{
		nbody_log("a is not smemory");
		return;
	}
	if(b == nullptr)
	{
		nbody_log("b is not smemory");
		return;
	}
	if(c == nullptr)
	{
		nbody_log("c is not smemory");
		return;
	}
	if(b->count() < a->count() || c->count() < a->count())
	{
		nbody_log("b or c is smaller than a");
		return;
	}
	const double* src[1] = { c->data() };
	double coef[1] = { d };
	check(nbh_fmaddn(m_ctx, a->data(), b->data(), src, coef, 1, a->count()), "fmadd");
}

namespace {

// Gather the terms with non-zero coefficients (the only ones the reference's
// decomposition touches, nbody/nbody_engine.cpp:59-65, 94-107).  Any such
// term that is not a device buffer of at least `count` doubles rejects the
// whole call.
bool gather_terms(const nbody_engine::memory_array& b, const nbcoord_t* c, size_t csize,
                  size_t count, std::vector<const double*>* src, std::vector<double>* coef)
{
	for(size_t k = 0; k != csize; ++k)
	{
		if(c[k] == 0.0)
		{
			continue;
		}
		const nbody_engine_hip::smemory* m = dynamic_cast<const nbody_engine_hip::smemory*>(b[k]);
		if(m == nullptr)
		{
			nbody_log("b is not smemory");
			return false;
		}
		if(m->count() < count)
		{
			nbody_log("b[%zu] is smaller than a", k);
			return false;
		}
		src->push_back(m->data());
		coef->push_back(c[k]);
	}
	return true;
}

}  // namespace

// a[i] += sum(b[k][i]*c[k]) in one pass (base class: one fmadd_inplace per term)
void nbody_engine_hip::fmaddn_inplace(memory* _a, const memory_array& b, const nbcoord_t* c, size_t csize)
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
	smemory* a = dynamic_cast<smemory*>(_a);
	if(a == nullptr)
	{
		nbody_log("a is not smemory");
		return;
	}
	std::vector<const double*> src;
	std::vector<double> coef;
	if(!gather_terms(b, c, csize, a->count(), &src, &coef))
	{
		return;
	}
	if(src.empty())
	{
		return;
	}
	check(nbh_fmaddn(m_ctx, a->data(), a->data(), src.data(), coef.data(), src.size(), a->count()),
	      "fmaddn_inplace");
}

// nbody/nbody_engine_simple.cpp:341-394 semantics, all terms in one pass.
void nbody_engine_hip::fmaddn_corr(memory* _a, memory* _corr, const memory_array& b,
                                   const nbcoord_t* c, size_t csize)
{
	smemory* a = dynamic_cast<smemory*>(_a);
	smemory* corr = dynamic_cast<smemory*>(_corr);
	if(a == nullptr)
	{
		nbody_log("a is not smemory");
		return;
	}
	if(corr == nullptr)
	{
		nbody_log("corr is not smemory");
		return;
	}
	if(c == nullptr)
	{
		nbody_log("c is NULL");
		return;
	}
	if(csize > b.size())
	{
		nbody_log("csize > b.size()");
		return;
	}
	for(size_t k = 0; k < csize; ++k)
	{
		if(dynamic_cast<const smemory*>(b[k]) == nullptr)
		{
			nbody_log("b is not smemory");
			return;
		}
	}
	if(corr->count() < a->count())
	{
		nbody_log("corr is smaller than a");
		return;
	}
	std::vector<const double*> src;
	std::vector<double> coef;
	if(!gather_terms(b, c, csize, a->count(), &src, &coef))
	{
		return;
	}
	if(src.empty())
	{
		return;
	}
	check(nbh_fmaddn_corr(m_ctx, a->data(), corr->data(), src.data(), coef.data(), src.size(), a->count()),
	      "fmaddn_corr");
}

// a[i] = b[i] + sum(c[k][i]*d[k]) in one pass, b may be NULL (then 0).
// As in the base class, with b != NULL and no non-zero d the call leaves a as is.
void nbody_engine_hip::fmaddn(memory* _a, const memory* _b, const memory_array& c, const nbcoord_t* d, size_t dsize)
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
	smemory* a = dynamic_cast<smemory*>(_a);
	if(a == nullptr)
	{
		nbody_log("a is not smemory");
		return;
	}
	const smemory* b = nullptr;
	if(_b != nullptr)
	{
		b = dynamic_cast<const smemory*>(_b);
		if(b == nullptr)
		{
			nbody_log("b is not smemory");
			return;
		}
		if(b->count() < a->count())
		{
			nbody_log("b is smaller than a");
			return;
		}
	}
	std::vector<const double*> src;
	std::vector<double> coef;
	if(!gather_terms(c, d, dsize, a->count(), &src, &coef))
	{
		return;
	}
	if(b != nullptr && src.empty())
	{
		return;
	}
	check(nbh_fmaddn(m_ctx, a->data(), b == nullptr ? nullptr : b->data(), src.data(), coef.data(),
	                 src.size(), a->count()),
	      "fmaddn");
}

void nbody_engine_hip::fmaxabs(const memory* _a, nbcoord_t& result)
{
	const smemory* a = dynamic_cast<const smemory*>(_a);
	if(a == nullptr)
	{
		nbody_log("a is not smemory");
		return;
	}
	double r = 0;
	if(check(nbh_maxabs(m_ctx, a->data(), a->count(), &r), "fmaxabs"))
	{
		result = r;
	}
}

void nbody_engine_hip::print_info() const
{
	char name[128] = {0};
	int cus = 0;
	if(m_ctx != nullptr)
	{
		nbh_ctx_info(m_ctx, name, sizeof(name), &cus);
	}
	nbody_print("\tdevice: %d (%s, %d CUs)", m_device, name, cus);
	nbody_print("\tblock_size: %d", m_block_size);
	nbody_print("\tforce_precision: %s",
	            m_precision == NBH_FP64 ? "fp64" : m_precision == NBH_FP32 ? "fp32" :
	            m_precision == NBH_FP32_CORR ? "fp32_corr" : "fp64 (block operation order)");
}

// ---------------------------------------------------------------------------
// block (nbody/nbody_engine_block.cpp:14-126)
// ---------------------------------------------------------------------------
nbody_engine_hip_block::nbody_engine_hip_block()
{
	m_precision = NBH_FP64_BLOCK;
}

const char* nbody_engine_hip_block::type_name() const
{
	return "nbody_engine_hip_block";
}

// ---------------------------------------------------------------------------
// Ahmad-Cohen (nbody/nbody_engine_ah.cpp:3-160)
// ---------------------------------------------------------------------------
nbody_engine_hip_ah::nbody_engine_hip_ah(size_t full_recompute_rate, double max_dist, double min_force) :
	m_rate(full_recompute_rate),
	m_max_dist(max_dist),
	m_min_force(min_force),
	m_ah(nullptr),
	m_have_lists(false)
{
}

nbody_engine_hip_ah::~nbody_engine_hip_ah()
{
	nbh_ah_destroy(m_ah);
}

const char* nbody_engine_hip_ah::type_name() const
{
	return "nbody_engine_hip_ah";
}

bool nbody_engine_hip_ah::init(nbody_data* data)
{
	if(!nbody_engine_hip::init(data))
	{
		return false;
	}
	nbh_ah_destroy(m_ah);
	m_ah = nullptr;
	m_have_lists = false;
	return check(nbh_ah_create(m_ctx, data->get_count(), &m_ah), "neighbour-list allocation");
}

// nbody/nbody_engine_ah.cpp:16-42.  A full recompute also runs when no lists
// exist yet (the first call of a resumed run at a step that is not a multiple
// of the rate), where the reference would read empty lists.
void nbody_engine_hip_ah::fcompute(const nbcoord_t& t, const memory* _y, memory* _f)
{
	(void)t;
	const smemory* y = nullptr;
	smemory* f = nullptr;
	if(!check_f(_y, _f, &y, &f))
	{
		return;
	}
	if(m_ah == nullptr)
	{
		nbody_log("hip_ah: engine not initialised");
		return;
	}
	advise_compute_count();
	if(!m_have_lists || m_data->get_step() % m_rate == 0)
	{
		m_have_lists = check(nbh_ah_fcompute_full(m_ctx, m_ah, y->data(), m_mass->data(), f->data(), m_max_dist,
		                                          m_min_force),
		                     "fcompute (full recompute)");
	}
	else
	{
		check(nbh_ah_fcompute_sparse(m_ctx, m_ah, y->data(), m_mass->data(), f->data()), "fcompute (neighbours)");
	}
}

void nbody_engine_hip_ah::print_info() const
{
	nbody_engine_hip::print_info();
	nbody_print("\tfull_recompute_rate: %zu", m_rate);
	nbody_print("\tmax_dist: %g", m_max_dist);
	nbody_print("\tmin_force: %g", m_min_force);
}

bool nbody_engine_hip::compute_statistics(bool with_energy, double out[12])
{
	if(m_y == nullptr || m_mass == nullptr)
	{
		return false;
	}
	return check(nbh_stats(m_ctx, m_y->data(), m_mass->data(), m_data->get_count(), with_energy ? 1 : 0, out),
	             "statistics");
}

// ---------------------------------------------------------------------------
// Barnes-Hut
// ---------------------------------------------------------------------------
const char* tree_layout_name(e_tree_layout tree_layout)
{
	switch(tree_layout)
	{
	case etl_tree:
		return "tree";
	case etl_heap:
		return "heap";
	case etl_heap_stackless:
		return "heap_stackless";
	default:
		return "";
	}
}

e_tree_layout tree_layout_from_str(const std::string& name)
{
	if(name == "tree") { return etl_tree; }
	if(name == "heap") { return etl_heap; }
	if(name == "heap_stackless") { return etl_heap_stackless; }
	return etl_unknown;
}

nbody_engine_hip_bh::nbody_engine_hip_bh(nbcoord_t ratio, e_traverse_type tt, size_t rate, e_tree_layout tl) :
	m_ratio(ratio),
	m_traverse_type(tt),
	m_tree_build_rate(rate),
	m_tree_layout(tl),
	m_tree(nullptr),
	m_tree_built(false),
	m_group_walk(false),
	m_quadrupole(false),
	m_quad_near(3.0f),
	m_local_far(false),
	m_morton(false),
	m_net_rho(0.0f),
	m_net_near(0.0f),
	m_net_two_level(false),
	m_net(nullptr)
{
}

nbody_engine_hip_bh::~nbody_engine_hip_bh()
{
	if(m_tree != nullptr)
	{
		nbh_tree_destroy(m_tree);
	}
	if(m_net != nullptr)
	{
		nbh_mlp_destroy(m_net);
	}
}

// Correction file: 8 bytes of magic, int32 parameter count, float32 rho_min,
// float32 near_k, then the float32 parameters (little-endian), as written by
// nbody_amd.correction.save.  "NBHMLP01": a network of the monopole group walk
// (far cells of each 64-body group); "NBHMLP2L": a network of the two-level
// local walk (far cells of each super-group, remainder after the quadrupole),
// which needs bh_multipole=quadrupole and bh_far_field=local.
bool nbody_engine_hip_bh::load_correction(const std::string& path, double rho, double near_k)
{
	FILE* fp = std::fopen(path.c_str(), "rb");
	if(fp == nullptr)
	{
		nbody_log("learned_correction: cannot open '%s'", path.c_str());
		return false;
	}
	char magic[8] = {0};
	int32_t count = 0;
	float file_rho = 0.0f;
	float file_near = 0.0f;
	int expect = 0;
	nbh_mlp_param_count(&expect);
	const bool read_magic = std::fread(magic, 1, 8, fp) == 8;
	const bool group_net = read_magic && std::memcmp(magic, "NBHMLP01", 8) == 0;
	const bool two_level_net = read_magic && std::memcmp(magic, "NBHMLP2L", 8) == 0;
	bool ok = (group_net || two_level_net) &&
	          std::fread(&count, sizeof(count), 1, fp) == 1 && count == expect &&
	          std::fread(&file_rho, sizeof(file_rho), 1, fp) == 1 && file_rho >= 0.0f &&
	          std::fread(&file_near, sizeof(file_near), 1, fp) == 1 && file_near >= 0.0f;
	std::vector<float> params;
	if(ok)
	{
		params.resize(static_cast<size_t>(count));
		ok = std::fread(params.data(), sizeof(float), params.size(), fp) == params.size() && std::fgetc(fp) == EOF;
	}
	std::fclose(fp);
	if(!ok)
	{
		nbody_log("learned_correction: '%s' is not a correction network file", path.c_str());
		return false;
	}
	if(two_level_net && !(m_quadrupole && m_local_far))
	{
		nbody_log("learned_correction: '%s' corrects the two-level walk: it needs bh_multipole=quadrupole and "
		          "bh_far_field=local", path.c_str());
		return false;
	}
	if(group_net && m_quadrupole)
	{
		nbody_log("learned_correction: '%s' corrects the monopole group walk: it needs bh_multipole=monopole",
		          path.c_str());
		return false;
	}
	m_net_two_level = two_level_net;
	m_net_params.swap(params);
	m_net_rho = rho >= 0.0 ? static_cast<float>(rho) : file_rho;
	m_net_near = near_k >= 0.0 ? static_cast<float>(near_k) : file_near;
	m_net_path = path;
	return true;
}

const char* nbody_engine_hip_bh::type_name() const
{
	return "nbody_engine_hip_bh";
}

bool nbody_engine_hip_bh::init(nbody_data* data)
{
	size_t n = data->get_count();
	if(m_morton ? n < 2 : (n == 0 || (n & (n - 1)) != 0))
	{
		if(m_morton)
		{
			nbody_log("hip_bh: body count %zu: the Morton tree needs at least 2 bodies", n);
		}
		else
		{
			nbody_log("hip_bh: body count %zu is not a power of two (heap layout needs 2^k leaves)", n);
		}
		return false;
	}
	if(!nbody_engine_hip::init(data))
	{
		return false;
	}
	if(m_tree != nullptr)
	{
		nbh_tree_destroy(m_tree);
		m_tree = nullptr;
	}
	m_tree_built = false;
	if(!check(nbh_tree_create(m_ctx, n, &m_tree), "tree allocation"))
	{
		return false;
	}
	// the Morton tree's group walks read its walk records only: its builds skip
	// the FP64 node arrays of the nodes inside the 64-body build groups
	if(m_morton && !check(nbh_tree_set_radix_lean(m_ctx, m_tree, 1), "tree setup"))
	{
		return false;
	}
	// the two-level walk's lists (default capacity, 32768 entries per 1024 bodies)
	// are allocated here, not inside the first force evaluation
	if(m_group_walk && m_quadrupole && m_local_far && m_precision == NBH_FP32 &&
	   !check(nbh_tree_set_list_cap(m_ctx, m_tree, 32768), "two-level walk list allocation"))
	{
		return false;
	}
	if(!m_net_params.empty() && m_net == nullptr)
	{
		if(!check(nbh_mlp_create(m_ctx, 0, &m_net), "correction network allocation") ||
		   !check(nbh_mlp_set_params(m_ctx, m_net, m_net_params.data()), "correction network upload"))
		{
			return false;
		}
	}
	return true;
}

// Build or refit, then traverse (nbody/nbody_engine_simple_bh.cpp:16-93,
// nbody/nbody_engine_cuda_bh_tex.cpp:46-201).  Targets are always walked in
// heap leaf order; `cycle` gives the same numbers (same positions and masses).
void nbody_engine_hip_bh::fcompute(const nbcoord_t& t, const memory* _y, memory* _f)
{
	(void)t;
	const smemory* y = nullptr;
	smemory* f = nullptr;
	if(!check_f(_y, _f, &y, &f))
	{
		return;
	}
	if(m_tree == nullptr)
	{
		nbody_log("hip_bh: engine not initialised");
		return;
	}
	advise_compute_count();
	// tree.is_empty() || rate == 0 || step % rate == 0 -> full build
	bool build = !m_tree_built || m_tree_build_rate == 0 || (m_data->get_step() % m_tree_build_rate) == 0;
	if(build)
	{
		const nbh_status st = m_morton ? nbh_tree_build_radix(m_ctx, m_tree, y->data(), m_mass->data(), m_ratio)
		                               : nbh_tree_build(m_ctx, m_tree, y->data(), m_mass->data(), m_ratio);
		if(!check(st, "tree build"))
		{
			return;
		}
		m_tree_built = true;
	}
	else if(!check(nbh_tree_refit(m_ctx, m_tree, y->data(), m_ratio), "tree refit"))
	{
		return;
	}
	if(m_group_walk && m_precision == NBH_FP32 && m_data->get_count() >= 64)
	{
		if(m_net != nullptr && m_net_two_level)
		{
			check(nbh_bh_fcompute_group_local2_mlp(m_ctx, m_tree, m_net, y->data(), f->data(), m_quad_near),
			      "bh fcompute (two-level walk, learned correction)");
			return;
		}
		if(m_net != nullptr)
		{
			check(nbh_bh_fcompute_group_mlp(m_ctx, m_tree, m_net, y->data(), f->data(), m_net_rho, m_net_near),
			      "bh fcompute (group walk, learned correction)");
			return;
		}
		if(m_quadrupole && m_local_far)
		{
			check(nbh_bh_fcompute_group_local2(m_ctx, m_tree, y->data(), f->data(), m_quad_near),
			      "bh fcompute (quadrupole walk, local far field)");
			return;
		}
		if(m_quadrupole)
		{
			check(nbh_bh_fcompute_group_quad(m_ctx, m_tree, y->data(), f->data(), m_quad_near),
			      "bh fcompute (quadrupole walk)");
			return;
		}
		check(nbh_bh_fcompute_group(m_ctx, m_tree, y->data(), f->data()), "bh fcompute (group walk)");
		return;
	}
	check(nbh_bh_fcompute(m_ctx, m_tree, y->data(), m_mass->data(), f->data(),
	                      m_tree_layout == etl_heap ? NBH_LAYOUT_HEAP : NBH_LAYOUT_HEAP_STACKLESS, m_precision),
	      "bh fcompute");
}

void nbody_engine_hip_bh::print_info() const
{
	nbody_engine_hip::print_info();
	nbody_print("\tdistance_to_node_radius_ratio: %g", m_ratio);
	nbody_print("\ttraverse_type: %s", m_traverse_type == ett_cycle ? "cycle" : "nested_tree");
	nbody_print("\ttree_layout: %s", tree_layout_name(m_tree_layout));
	nbody_print("\ttree_build_rate: %zu", m_tree_build_rate);
	nbody_print("\tbh_walk: %s", m_group_walk ? "group" : "body");
	nbody_print("\tbh_tree: %s", m_morton ? "morton" : "heap");
	nbody_print("\tbh_multipole: %s", m_quadrupole ? "quadrupole" : "monopole");
	if(m_quadrupole)
	{
		nbody_print("\tbh_quad_near: %g", static_cast<double>(m_quad_near));
		nbody_print("\tbh_far_field: %s", m_local_far ? "local" : "per_body");
	}
	nbody_print("\tlearned_correction: %s", m_net_params.empty() ? "none" : m_net_path.c_str());
	if(!m_net_params.empty())
	{
		nbody_print("\tcorrection_rho: %g", static_cast<double>(m_net_rho));
		nbody_print("\tcorrection_near: %g", static_cast<double>(m_net_near));
	}
}
