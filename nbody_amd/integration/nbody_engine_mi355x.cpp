// nbody_engine_mi355x.cpp -- see nbody_engine_mi355x.h.  Compiled inside the
// drons/nbody tree; each method is one or a few nbh_* calls.
#include "nbody_engine_mi355x.h"

#include <QDebug>
#include <vector>

// ---- buffers -----------------------------------------------------------------
nbody_engine_mi355x::smemory::smemory(nbh_ctx* ctx, size_t size) : m_ctx(ctx), m_data(nullptr), m_size(size)
{
	if(size != 0 && nbh_malloc(ctx, size, &m_data) != NBH_OK)
	{
		qDebug() << "hip: allocation of" << size << "bytes failed:" << nbh_last_error();
		m_data = nullptr;
		m_size = 0;
	}
}

nbody_engine_mi355x::smemory::~smemory()
{
	if(m_data != nullptr)
	{
		nbh_free(m_ctx, m_data);
	}
}

size_t nbody_engine_mi355x::smemory::size() const { return m_size; }
double* nbody_engine_mi355x::smemory::data() const { return static_cast<double*>(m_data); }
size_t nbody_engine_mi355x::smemory::count() const { return m_size / sizeof(double); }

// ---- direct engine -------------------------------------------------------------
nbody_engine_mi355x::nbody_engine_mi355x() :
	m_ctx(nullptr), m_precision(NBH_FP64), m_data(nullptr), m_mass(nullptr), m_y(nullptr)
{
}

nbody_engine_mi355x::~nbody_engine_mi355x()
{
	delete m_mass;
	delete m_y;
	if(m_ctx != nullptr)
	{
		nbh_ctx_destroy(m_ctx);
	}
}

bool nbody_engine_mi355x::ok(nbh_status s, const char* what) const
{
	if(s != NBH_OK)
	{
		qDebug() << "hip:" << what << "failed:" << nbh_last_error();
		return false;
	}
	return true;
}

int nbody_engine_mi355x::select_device(int device)
{
	if(m_ctx != nullptr)
	{
		nbh_ctx_destroy(m_ctx);
		m_ctx = nullptr;
	}
	return ok(nbh_ctx_create(device, &m_ctx), "nbh_ctx_create") ? 0 : -1;
}

void nbody_engine_mi355x::set_precision(nbh_precision p) { m_precision = p; }
const char* nbody_engine_mi355x::type_name() const { return "nbody_engine_hip"; }

// AoS bodies -> SoA device state (nbody/nbody_engine_cuda.cpp:91-139).
bool nbody_engine_mi355x::init(nbody_data* data)
{
	if(m_ctx == nullptr)
	{
		qDebug() << "hip: no device selected";
		return false;
	}
	m_data = data;
	const size_t n = data->get_count();
	delete m_mass;
	delete m_y;
	m_mass = static_cast<smemory*>(create_buffer(sizeof(nbcoord_t) * n));
	m_y = static_cast<smemory*>(create_buffer(sizeof(nbcoord_t) * problem_size()));
	if(m_mass == nullptr || m_y == nullptr)
	{
		return false;
	}
	std::vector<nbcoord_t> y(6 * n);
	const nbvertex_t* r = data->get_vertites();
	const nbvertex_t* v = data->get_velosites();
	for(size_t i = 0; i != n; ++i)
	{
		y[i] = r[i].x;
		y[n + i] = r[i].y;
		y[2 * n + i] = r[i].z;
		y[3 * n + i] = v[i].x;
		y[4 * n + i] = v[i].y;
		y[5 * n + i] = v[i].z;
	}
	return ok(nbh_memcpy_h2d(m_ctx, m_y->data(), y.data(), m_y->size()), "upload") &&
	       ok(nbh_memcpy_h2d(m_ctx, m_mass->data(), data->get_mass(), m_mass->size()), "upload");
}

// SoA device state -> AoS bodies (nbody/nbody_engine_cuda.cpp:141-165).
void nbody_engine_mi355x::get_data(nbody_data* data)
{
	if(m_y == nullptr)
	{
		qDebug() << "hip: engine not initialised";
		return;
	}
	const size_t n = m_data->get_count();
	std::vector<nbcoord_t> y(6 * n);
	if(!ok(nbh_memcpy_d2h(m_ctx, y.data(), m_y->data(), m_y->size()), "download"))
	{
		return;
	}
	nbvertex_t* r = data->get_vertites();
	nbvertex_t* v = data->get_velosites();
	for(size_t i = 0; i != n; ++i)
	{
		r[i] = nbvertex_t(y[i], y[n + i], y[2 * n + i]);
		v[i] = nbvertex_t(y[3 * n + i], y[4 * n + i], y[5 * n + i]);
	}
}

size_t nbody_engine_mi355x::problem_size() const { return m_data == nullptr ? 0 : 6 * m_data->get_count(); }
nbody_engine::memory* nbody_engine_mi355x::get_y() { return m_y; }
void nbody_engine_mi355x::advise_time(const nbcoord_t& dt) { m_data->advise_time(dt); }
nbcoord_t nbody_engine_mi355x::get_time() const { return m_data->get_time(); }
void nbody_engine_mi355x::set_time(nbcoord_t t) { m_data->set_time(t); }
size_t nbody_engine_mi355x::get_step() const { return m_data->get_step(); }
void nbody_engine_mi355x::set_step(size_t s) { m_data->set_step(s); }

bool nbody_engine_mi355x::check_f(const memory* _y, const memory* _f, const smemory** ys, smemory** fs) const
{
	*ys = dynamic_cast<const smemory*>(_y);
	*fs = dynamic_cast<smemory*>(const_cast<memory*>(_f));
	if(*ys == nullptr || *fs == nullptr)
	{
		qDebug() << "y or f is not smemory";
		return false;
	}
	if(m_mass == nullptr || (*ys)->count() < problem_size() || (*fs)->count() < problem_size())
	{
		qDebug() << "fcompute: engine not initialised or buffers smaller than the problem";
		return false;
	}
	return true;
}

void nbody_engine_mi355x::fcompute(const nbcoord_t& t, const memory* _y, memory* _f)
{
	Q_UNUSED(t);
	const smemory* y = nullptr;
	smemory* f = nullptr;
	if(!check_f(_y, _f, &y, &f))
	{
		return;
	}
	advise_compute_count();
	ok(nbh_fcompute_direct(m_ctx, y->data(), m_mass->data(), f->data(), m_data->get_count(), m_precision, 0),
	   "fcompute");
}

void nbody_engine_mi355x::clamp(memory* _y, nbcoord_t b)
{
	smemory* y = dynamic_cast<smemory*>(_y);
	if(y == nullptr)
	{
		qDebug() << "y is not smemory";
		return;
	}
	ok(nbh_clamp(m_ctx, y->data(), m_data->get_count(), b), "clamp");
}

nbody_engine::memory* nbody_engine_mi355x::create_buffer(size_t s)
{
	smemory* m = new smemory(m_ctx, s);
	if(s != 0 && m->size() == 0)
	{
		delete m;
		return nullptr;
	}
	return m;
}

void nbody_engine_mi355x::free_buffer(memory* m) { delete m; }

void nbody_engine_mi355x::read_buffer(void* dst, const memory* _src)
{
	const smemory* src = dynamic_cast<const smemory*>(_src);
	if(src == nullptr)
	{
		qDebug() << "src is not smemory";
		return;
	}
	ok(nbh_memcpy_d2h(m_ctx, dst, src->data(), src->size()), "read_buffer");
}

void nbody_engine_mi355x::write_buffer(memory* _dst, const void* src)
{
	smemory* dst = dynamic_cast<smemory*>(_dst);
	if(dst == nullptr)
	{
		qDebug() << "dst is not smemory";
		return;
	}
	ok(nbh_memcpy_h2d(m_ctx, dst->data(), src, dst->size()), "write_buffer");
}

void nbody_engine_mi355x::copy_buffer(memory* _a, const memory* _b)
{
	smemory* a = dynamic_cast<smemory*>(_a);
	const smemory* b = dynamic_cast<const smemory*>(_b);
	if(a == nullptr || b == nullptr || a->size() != b->size())
	{
		qDebug() << "copy_buffer: not smemory or sizes differ";
		return;
	}
	ok(nbh_memcpy_d2d(m_ctx, a->data(), b->data(), a->size()), "copy_buffer");
}

void nbody_engine_mi355x::fill_buffer(memory* _a, const nbcoord_t& value)
{
	smemory* a = dynamic_cast<smemory*>(_a);
	if(a == nullptr)
	{
		qDebug() << "a is not smemory";
		return;
	}
	ok(nbh_fill(m_ctx, a->data(), value, a->count()), "fill_buffer");
}

void nbody_engine_mi355x::fmadd_inplace(memory* _a, const memory* _b, const nbcoord_t& c)
{
	smemory* a = dynamic_cast<smemory*>(_a);
	const smemory* b = dynamic_cast<const smemory*>(_b);
	if(a == nullptr || b == nullptr || b->count() < a->count())
	{
		qDebug() << "fmadd_inplace: not smemory or b shorter than a";
		return;
	}
	const double* src[1] = {b->data()};
	ok(nbh_fmaddn(m_ctx, a->data(), a->data(), src, &c, 1, a->count()), "fmadd_inplace");
}

void nbody_engine_mi355x::fmadd(memory* _a, const memory* _b, const memory* _c, const nbcoord_t& d)
{
	smemory* a = dynamic_cast<smemory*>(_a);
	const smemory* b = dynamic_cast<const smemory*>(_b);
	const smemory* c = dynamic_cast<const smemory*>(_c);
	if(a == nullptr || b == nullptr || c == nullptr || b->count() < a->count() || c->count() < a->count())
	{
		qDebug() << "fmadd: not smemory or an operand shorter than a";
		return;
	}
	const double* src[1] = {c->data()};
	ok(nbh_fmaddn(m_ctx, a->data(), b->data(), src, &d, 1, a->count()), "fmadd");
}

// The terms of b with a non-zero coefficient, in order (nbody/nbody_engine.cpp:47-113
// applies exactly those, one fmadd at a time).
bool nbody_engine_mi355x::gather(const memory_array& b, const nbcoord_t* c, size_t csize, size_t n,
                                 std::vector<const double*>& src, std::vector<double>& coef) const
{
	if(c == nullptr || csize > b.size())
	{
		qDebug() << "fmaddn: c == NULL or csize > b.size()";
		return false;
	}
	for(size_t k = 0; k != csize; ++k)
	{
		if(c[k] == 0)
		{
			continue;
		}
		const smemory* bk = dynamic_cast<const smemory*>(b[k]);
		if(bk == nullptr || bk->count() < n)
		{
			qDebug() << "fmaddn: term" << k << "is not smemory or shorter than a";
			return false;
		}
		src.push_back(bk->data());
		coef.push_back(c[k]);
	}
	return true;
}

void nbody_engine_mi355x::fmaddn_inplace(memory* _a, const memory_array& b, const nbcoord_t* c, size_t csize)
{
	smemory* a = dynamic_cast<smemory*>(_a);
	std::vector<const double*> src;
	std::vector<double> coef;
	if(a == nullptr || !gather(b, c, csize, a->count(), src, coef) || src.empty())
	{
		return;
	}
	ok(nbh_fmaddn(m_ctx, a->data(), a->data(), src.data(), coef.data(), src.size(), a->count()), "fmaddn_inplace");
}

void nbody_engine_mi355x::fmaddn_corr(memory* _a, memory* _corr, const memory_array& b, const nbcoord_t* c,
                                      size_t csize)
{
	smemory* a = dynamic_cast<smemory*>(_a);
	smemory* corr = dynamic_cast<smemory*>(_corr);
	std::vector<const double*> src;
	std::vector<double> coef;
	if(a == nullptr || corr == nullptr || corr->count() < a->count() || !gather(b, c, csize, a->count(), src, coef) ||
	   src.empty())
	{
		return;
	}
	ok(nbh_fmaddn_corr(m_ctx, a->data(), corr->data(), src.data(), coef.data(), src.size(), a->count()),
	   "fmaddn_corr");
}

void nbody_engine_mi355x::fmaddn(memory* _a, const memory* _b, const memory_array& c, const nbcoord_t* d,
                                 size_t dsize)
{
	smemory* a = dynamic_cast<smemory*>(_a);
	const smemory* b = _b == nullptr ? nullptr : dynamic_cast<const smemory*>(_b);
	std::vector<const double*> src;
	std::vector<double> coef;
	if(a == nullptr || (_b != nullptr && (b == nullptr || b->count() < a->count())) ||
	   !gather(c, d, dsize, a->count(), src, coef))
	{
		return;
	}
	if(b != nullptr && src.empty())
	{
		return;     // a left as it is, as nbody_engine::fmaddn does with no non-zero term
	}
	ok(nbh_fmaddn(m_ctx, a->data(), b == nullptr ? nullptr : b->data(), src.data(), coef.data(), src.size(),
	              a->count()), "fmaddn");
}

void nbody_engine_mi355x::fmaxabs(const memory* _a, nbcoord_t& result)
{
	const smemory* a = dynamic_cast<const smemory*>(_a);
	if(a == nullptr)
	{
		qDebug() << "a is not smemory";
		return;
	}
	double r = 0.0;
	if(ok(nbh_maxabs(m_ctx, a->data(), a->count(), &r), "fmaxabs"))
	{
		result = r;
	}
}

void nbody_engine_mi355x::print_info() const
{
	char name[128] = {0};
	int cus = 0;
	if(m_ctx != nullptr && nbh_ctx_info(m_ctx, name, sizeof(name), &cus) == NBH_OK)
	{
		qDebug() << "\tdevice:" << name << cus << "CUs";
	}
	qDebug() << "\tforce_precision:" << (m_precision == NBH_FP64 ? "fp64" : m_precision == NBH_FP32 ? "fp32" : "fp32_corr");
}

// ---- Barnes-Hut engine -----------------------------------------------------------
nbody_engine_mi355x_bh::nbody_engine_mi355x_bh(nbcoord_t ratio, size_t rate, nbh_tree_layout layout) :
	m_tree(nullptr), m_ratio(ratio), m_tree_build_rate(rate), m_layout(layout), m_built(false)
{
}

nbody_engine_mi355x_bh::~nbody_engine_mi355x_bh()
{
	nbh_tree_destroy(m_tree);
}

const char* nbody_engine_mi355x_bh::type_name() const { return "nbody_engine_hip_bh"; }

bool nbody_engine_mi355x_bh::init(nbody_data* data)
{
	const size_t n = data->get_count();
	if(n == 0 || (n & (n - 1)) != 0)
	{
		qDebug() << "hip_bh: body count" << n << "is not a power of two";
		return false;
	}
	if(!nbody_engine_mi355x::init(data))
	{
		return false;
	}
	nbh_tree_destroy(m_tree);
	m_tree = nullptr;
	m_built = false;
	return ok(nbh_tree_create(m_ctx, n, &m_tree), "nbh_tree_create");
}

// Build or refit, then walk (nbody/nbody_engine_simple_bh.cpp:16-93).
void nbody_engine_mi355x_bh::fcompute(const nbcoord_t& t, const memory* _y, memory* _f)
{
	Q_UNUSED(t);
	const smemory* y = nullptr;
	smemory* f = nullptr;
	if(!check_f(_y, _f, &y, &f) || m_tree == nullptr)
	{
		return;
	}
	advise_compute_count();
	const bool build = !m_built || m_tree_build_rate == 0 || m_data->get_step() % m_tree_build_rate == 0;
	if(build)
	{
		if(!ok(nbh_tree_build(m_ctx, m_tree, y->data(), m_mass->data(), m_ratio), "nbh_tree_build"))
		{
			return;
		}
		m_built = true;
	}
	else if(!ok(nbh_tree_refit(m_ctx, m_tree, y->data(), m_ratio), "nbh_tree_refit"))
	{
		return;
	}
	ok(nbh_bh_fcompute(m_ctx, m_tree, y->data(), m_mass->data(), f->data(), m_layout, m_precision),
	   "nbh_bh_fcompute");
}

void nbody_engine_mi355x_bh::print_info() const
{
	nbody_engine_mi355x::print_info();
	qDebug() << "\tdistance_to_node_radius_ratio:" << m_ratio;
	qDebug() << "\ttree_build_rate:" << m_tree_build_rate;
	qDebug() << "\ttree_layout:" << (m_layout == NBH_LAYOUT_HEAP ? "heap" : "heap_stackless");
}

// ---- factory branch ------------------------------------------------------------------
nbody_engine* nbody_create_engine_mi355x(const QVariantMap& param)
{
	const QString type(param.value("engine").toString());
	if(type != "hip" && type != "hip_bh")
	{
		return nullptr;
	}
	bool ok_dev = false;
	const int device = param.value("device", "0").toString().toInt(&ok_dev);
	const QString prec(param.value("force_precision", "fp64").toString());
	nbh_precision p = NBH_FP64;
	if(prec == "fp32")
	{
		p = NBH_FP32;
	}
	else if(prec == "fp32_corr" && type == "hip")
	{
		p = NBH_FP32_CORR;
	}
	else if(prec != "fp64")
	{
		qDebug() << "Invalid force_precision" << prec;
		return nullptr;
	}
	if(!ok_dev)
	{
		qDebug() << "Invalid device (one id per engine)";
		return nullptr;
	}
	nbody_engine_mi355x* engine = nullptr;
	if(type == "hip")
	{
		engine = new nbody_engine_mi355x();
	}
	else
	{
		const QString layout(param.value("tree_layout", "heap").toString());
		const QString traverse(param.value("traverse_type", "cycle").toString());
		if((layout != "heap" && layout != "heap_stackless") || (traverse != "cycle" && traverse != "nested_tree"))
		{
			qDebug() << "Invalid tree_layout or traverse_type";
			return nullptr;
		}
		engine = new nbody_engine_mi355x_bh(param.value("distance_to_node_radius_ratio", 10).toDouble(),
		                                    param.value("tree_build_rate", 0).toULongLong(),
		                                    layout == "heap" ? NBH_LAYOUT_HEAP : NBH_LAYOUT_HEAP_STACKLESS);
	}
	if(engine->select_device(device) != 0)
	{
		delete engine;
		return nullptr;
	}
	engine->set_precision(p);
	return engine;
}
