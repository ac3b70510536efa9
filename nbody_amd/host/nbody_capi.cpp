// nbody_capi.cpp -- C ABI of the host layer (include/nbody_host.h).
#include "nbody_host.h"

#include <cstring>
#include <vector>

#include "nbody_arg_parser.h"
#include "nbody_engine_hip.h"
#include "nbody_engines.h"
#include "nbody_generators.h"
#include "nbody_log.h"
#include "nbody_solver.h"

namespace {

nbody_data* D(nbs_data* d) { return reinterpret_cast<nbody_data*>(d); }
const nbody_data* D(const nbs_data* d) { return reinterpret_cast<const nbody_data*>(d); }
nbody_engine* E(nbs_engine* e) { return reinterpret_cast<nbody_engine*>(e); }
const nbody_engine* E(const nbs_engine* e) { return reinterpret_cast<const nbody_engine*>(e); }
nbody_engine::memory* M(nbs_memory* m) { return reinterpret_cast<nbody_engine::memory*>(m); }
const nbody_engine::memory* M(const nbs_memory* m) { return reinterpret_cast<const nbody_engine::memory*>(m); }
nbody_solver* S(nbs_solver* s) { return reinterpret_cast<nbody_solver*>(s); }
const nbody_solver* S(const nbs_solver* s) { return reinterpret_cast<const nbody_solver*>(s); }

nbody_engine::memory_array MA(nbs_memory* const* b, size_t nb)
{
	nbody_engine::memory_array a(nb);
	for(size_t i = 0; i < nb; ++i)
	{
		a[i] = reinterpret_cast<nbody_engine::memory*>(b[i]);
	}
	return a;
}

// A buffer type no engine knows: the negative-branch tests pass it to every op
// (test/engine/test_nbody_engine.cpp:748-759).
class fake_memory : public nbody_engine::memory
{
	size_t m_size;
public:
	explicit fake_memory(size_t s) : m_size(s) {}
	size_t size() const override { return m_size; }
};

}  // namespace

extern "C" {

size_t nbs_log_count(void) { return nbody_log_count(); }
const char* nbs_log_last(void) { return nbody_log_last(); }
void nbs_log_quiet(int quiet) { nbody_log_set_quiet(quiet != 0); }

nbs_data* nbs_data_create(void) { return reinterpret_cast<nbs_data*>(new nbody_data()); }
void nbs_data_destroy(nbs_data* d) { delete D(d); }

int nbs_data_load(nbs_data* d, const char* path, const char* type)
{
	return D(d)->load_initial(path, type == nullptr ? "G1" : type) ? 1 : 0;
}

int nbs_data_save(const nbs_data* d, const char* path)
{
	return D(d)->save(path) ? 1 : 0;
}

void nbs_data_make_universe(nbs_data* d, size_t star_count, double sx, double sy, double sz)
{
	D(d)->make_universe(star_count, sx, sy, sz);
}

int nbs_data_generate(nbs_data* d, const char* kind, size_t n, uint64_t seed)
{
	return nbody_generate(D(d), kind, n, seed) ? 1 : 0;
}

void nbs_data_set(nbs_data* d, size_t n, const double* pos, const double* vel, const double* mass,
                  const double* radius)
{
	nbody_data* x = D(d);
	x->clear();
	for(size_t i = 0; i < n; ++i)
	{
		x->add_body(nbvertex_t(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]),
		            nbvertex_t(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]), mass[i],
		            radius == nullptr ? 0.0 : radius[i]);
	}
}

size_t nbs_data_count(const nbs_data* d) { return D(d)->get_count(); }

void nbs_data_get(const nbs_data* d, double* pos, double* vel, double* mass, double* radius)
{
	const nbody_data* x = D(d);
	for(size_t i = 0; i < x->get_count(); ++i)
	{
		if(pos != nullptr)
		{
			pos[3 * i] = x->get_vertites()[i].x;
			pos[3 * i + 1] = x->get_vertites()[i].y;
			pos[3 * i + 2] = x->get_vertites()[i].z;
		}
		if(vel != nullptr)
		{
			vel[3 * i] = x->get_velosites()[i].x;
			vel[3 * i + 1] = x->get_velosites()[i].y;
			vel[3 * i + 2] = x->get_velosites()[i].z;
		}
		if(mass != nullptr) { mass[i] = x->get_mass()[i]; }
		if(radius != nullptr) { radius[i] = x->get_radius()[i]; }
	}
}

double nbs_data_time(const nbs_data* d) { return D(d)->get_time(); }
size_t nbs_data_step(const nbs_data* d) { return D(d)->get_step(); }
void nbs_data_set_time(nbs_data* d, double t) { D(d)->set_time(t); }
void nbs_data_set_step(nbs_data* d, size_t s) { D(d)->set_step(s); }
void nbs_data_advise_time(nbs_data* d, double dt) { D(d)->advise_time(dt); }
size_t nbs_data_box_size(const nbs_data* d) { return D(d)->get_box_size(); }
void nbs_data_set_check_list(nbs_data* d, const char* list) { D(d)->set_check_list(list); }
void nbs_data_set_quiet(nbs_data* d, int quiet) { D(d)->set_quiet(quiet != 0); }
void nbs_data_print_statistics(nbs_data* d, nbs_engine* e) { D(d)->print_statistics(E(e)); }
void nbs_data_statistics(nbs_data* d, nbs_engine* e, double out[12]) { D(d)->compute_statistics(E(e), out); }

void nbs_data_errors(const nbs_data* d, double out[4])
{
	const nbody_data* x = D(d);
	out[0] = x->get_impulce_err();
	out[1] = x->get_impulce_moment_err();
	out[2] = x->get_energy_err();
	nbvertex_t vcm = (x->get_last_mass_center() - x->get_initial_mass_center()) / x->get_time();
	out[3] = vcm.length();
}

int nbs_data_is_equal(const nbs_data* a, const nbs_data* b, double eps)
{
	return D(a)->is_equal(*D(b), eps) ? 1 : 0;
}

void nbs_data_resize(nbs_data* d, size_t n) { D(d)->resize(n); }

nbs_engine* nbs_engine_create(const char* params)
{
	nbody_params p = nbody_params::parse(params == nullptr ? "" : params);
	return reinterpret_cast<nbs_engine*>(nbody_create_engine(p));
}

void nbs_engine_destroy(nbs_engine* e) { delete E(e); }
const char* nbs_engine_type_name(const nbs_engine* e) { return E(e)->type_name(); }
int nbs_engine_init(nbs_engine* e, nbs_data* d) { return E(e)->init(D(d)) ? 1 : 0; }
void nbs_engine_get_data(nbs_engine* e, nbs_data* d) { E(e)->get_data(D(d)); }
size_t nbs_engine_problem_size(const nbs_engine* e) { return E(e)->problem_size(); }
nbs_memory* nbs_engine_get_y(nbs_engine* e) { return reinterpret_cast<nbs_memory*>(E(e)->get_y()); }
void nbs_engine_advise_time(nbs_engine* e, double dt) { E(e)->advise_time(dt); }
double nbs_engine_get_time(const nbs_engine* e) { return E(e)->get_time(); }
void nbs_engine_set_time(nbs_engine* e, double t) { E(e)->set_time(t); }
size_t nbs_engine_get_step(const nbs_engine* e) { return E(e)->get_step(); }
void nbs_engine_set_step(nbs_engine* e, size_t s) { E(e)->set_step(s); }

void nbs_engine_fcompute(nbs_engine* e, double t, const nbs_memory* y, nbs_memory* f)
{
	E(e)->fcompute(t, M(y), M(f));
}

void nbs_engine_clamp(nbs_engine* e, nbs_memory* y, double b) { E(e)->clamp(M(y), b); }

nbs_memory* nbs_engine_create_buffer(nbs_engine* e, size_t bytes)
{
	return reinterpret_cast<nbs_memory*>(E(e)->create_buffer(bytes));
}

void nbs_engine_free_buffer(nbs_engine* e, nbs_memory* m) { E(e)->free_buffer(M(m)); }
void nbs_engine_read_buffer(nbs_engine* e, void* dst, const nbs_memory* src) { E(e)->read_buffer(dst, M(src)); }
void nbs_engine_write_buffer(nbs_engine* e, nbs_memory* dst, const void* src) { E(e)->write_buffer(M(dst), src); }
void nbs_engine_copy_buffer(nbs_engine* e, nbs_memory* a, const nbs_memory* b) { E(e)->copy_buffer(M(a), M(b)); }
void nbs_engine_fill_buffer(nbs_engine* e, nbs_memory* a, double value) { E(e)->fill_buffer(M(a), value); }

void nbs_engine_fmadd_inplace(nbs_engine* e, nbs_memory* a, const nbs_memory* b, double c)
{
	E(e)->fmadd_inplace(M(a), M(b), c);
}

void nbs_engine_fmadd(nbs_engine* e, nbs_memory* a, const nbs_memory* b, const nbs_memory* c, double d)
{
	E(e)->fmadd(M(a), M(b), M(c), d);
}

void nbs_engine_fmaddn_inplace(nbs_engine* e, nbs_memory* a, nbs_memory* const* b, size_t nb,
                               const double* c, size_t csize)
{
	E(e)->fmaddn_inplace(M(a), MA(b, nb), c, csize);
}

void nbs_engine_fmaddn_corr(nbs_engine* e, nbs_memory* a, nbs_memory* corr, nbs_memory* const* b, size_t nb,
                            const double* c, size_t csize)
{
	E(e)->fmaddn_corr(M(a), M(corr), MA(b, nb), c, csize);
}

void nbs_engine_fmaddn(nbs_engine* e, nbs_memory* a, const nbs_memory* b, nbs_memory* const* c, size_t nc,
                       const double* d, size_t dsize)
{
	E(e)->fmaddn(M(a), M(b), MA(c, nc), d, dsize);
}

void nbs_engine_fmaxabs(nbs_engine* e, const nbs_memory* a, double* result)
{
	nbcoord_t r = *result;
	E(e)->fmaxabs(M(a), r);
	*result = r;
}

size_t nbs_engine_compute_count(const nbs_engine* e) { return E(e)->get_compute_count(); }
void nbs_engine_print_info(const nbs_engine* e) { E(e)->print_info(); }

int nbs_engine_has_error(const nbs_engine* e)
{
	const nbody_engine_hip* h = dynamic_cast<const nbody_engine_hip*>(E(e));
	return (h != nullptr && h->has_error()) ? 1 : 0;
}

void* nbs_engine_hip_context(nbs_engine* e)
{
	nbody_engine_hip* h = dynamic_cast<nbody_engine_hip*>(E(e));
	return h == nullptr ? nullptr : h->context();
}

const void* nbs_engine_hip_tree(nbs_engine* e)
{
	nbody_engine_hip_bh* h = dynamic_cast<nbody_engine_hip_bh*>(E(e));
	return h == nullptr ? nullptr : h->tree();
}

const void* nbs_engine_hip_neighbours(nbs_engine* e)
{
	nbody_engine_hip_ah* h = dynamic_cast<nbody_engine_hip_ah*>(E(e));
	return h == nullptr ? nullptr : h->neighbours();
}

void* nbs_memory_device_ptr(nbs_memory* m)
{
	nbody_engine_hip::smemory* s = dynamic_cast<nbody_engine_hip::smemory*>(M(m));
	return s == nullptr ? nullptr : s->data();
}

size_t nbs_memory_size(const nbs_memory* m) { return M(m)->size(); }

nbs_memory* nbs_memory_fake_create(size_t bytes)
{
	return reinterpret_cast<nbs_memory*>(static_cast<nbody_engine::memory*>(new fake_memory(bytes)));
}

void nbs_memory_fake_destroy(nbs_memory* m) { delete M(m); }

nbs_solver* nbs_solver_create(const char* params)
{
	nbody_params p = nbody_params::parse(params == nullptr ? "" : params);
	return reinterpret_cast<nbs_solver*>(nbody_create_solver(p));
}

void nbs_solver_destroy(nbs_solver* s) { delete S(s); }
const char* nbs_solver_type_name(const nbs_solver* s) { return S(s)->type_name(); }
void nbs_solver_set_engine(nbs_solver* s, nbs_engine* e) { S(s)->set_engine(E(e)); }
void nbs_solver_set_time_step(nbs_solver* s, double min_step, double max_step) { S(s)->set_time_step(min_step, max_step); }

int nbs_solver_run(nbs_solver* s, nbs_data* d, nbs_stream* stream, double max_time, double dump_dt, double check_dt)
{
	return S(s)->run(D(d), reinterpret_cast<nbody_data_stream*>(stream), max_time, dump_dt, check_dt);
}

size_t nbs_parse_arguments(int argc, const char* const* argv, char* out, size_t out_len)
{
	const std::string text = nbody_parse_arguments(argc, argv).str();
	if(out != nullptr && out_len > 0)
	{
		const size_t k = text.size() < out_len - 1 ? text.size() : out_len - 1;
		std::memcpy(out, text.data(), k);
		out[k] = '\0';
	}
	return text.size();
}

nbs_stream* nbs_stream_create(void) { return reinterpret_cast<nbs_stream*>(new nbody_data_stream()); }
void nbs_stream_destroy(nbs_stream* s) { delete reinterpret_cast<nbody_data_stream*>(s); }
int nbs_stream_open(nbs_stream* s, const char* base, int64_t max_part_size, const nbs_stream_reader* append_to)
{
	return reinterpret_cast<nbody_data_stream*>(s)->open(base != nullptr ? base : "", max_part_size,
	                                                     reinterpret_cast<const nbody_data_stream_reader*>(append_to));
}
int nbs_stream_write(nbs_stream* s, const nbs_data* d)
{
	return reinterpret_cast<nbody_data_stream*>(s)->write(reinterpret_cast<const nbody_data*>(d));
}
void nbs_stream_close(nbs_stream* s) { reinterpret_cast<nbody_data_stream*>(s)->close(); }

static nbody_data_stream_reader* R(nbs_stream_reader* r) { return reinterpret_cast<nbody_data_stream_reader*>(r); }
static const nbody_data_stream_reader* RC(const nbs_stream_reader* r)
{
	return reinterpret_cast<const nbody_data_stream_reader*>(r);
}
nbs_stream_reader* nbs_stream_reader_create(void)
{
	return reinterpret_cast<nbs_stream_reader*>(new nbody_data_stream_reader());
}
void nbs_stream_reader_destroy(nbs_stream_reader* r) { delete R(r); }
int nbs_stream_reader_load(nbs_stream_reader* r, const char* base) { return R(r)->load(base != nullptr ? base : ""); }
void nbs_stream_reader_close(nbs_stream_reader* r) { R(r)->close(); }
size_t nbs_stream_reader_frame_count(const nbs_stream_reader* r) { return RC(r)->get_frame_count(); }
size_t nbs_stream_reader_steps_count(const nbs_stream_reader* r) { return RC(r)->get_steps_count(); }
double nbs_stream_reader_max_time(const nbs_stream_reader* r) { return RC(r)->get_max_time(); }
int nbs_stream_reader_seek(nbs_stream_reader* r, size_t frame) { return R(r)->seek(frame); }
size_t nbs_stream_reader_current_frame(const nbs_stream_reader* r) { return RC(r)->get_current_frame(); }
size_t nbs_stream_reader_body_count(const nbs_stream_reader* r) { return RC(r)->get_body_count(); }
size_t nbs_stream_reader_coord_size(const nbs_stream_reader* r) { return RC(r)->get_coord_size(); }
double nbs_stream_reader_box_size(const nbs_stream_reader* r) { return RC(r)->get_box_size(); }
size_t nbs_stream_reader_last_file_n(const nbs_stream_reader* r) { return RC(r)->get_last_file_n(); }
int nbs_stream_reader_read(nbs_stream_reader* r, nbs_data* d) { return R(r)->read(reinterpret_cast<nbody_data*>(d)); }

void nbs_data_get_color(const nbs_data* d, float* rgba)
{
	const nbody_data* x = reinterpret_cast<const nbody_data*>(d);
	const nbcolor_t* c = x->get_color();
	for(size_t i = 0; i < x->get_count(); ++i)
	{
		rgba[4 * i] = c[i].r;
		rgba[4 * i + 1] = c[i].g;
		rgba[4 * i + 2] = c[i].b;
		rgba[4 * i + 3] = c[i].a;
	}
}

void nbs_solver_advise(nbs_solver* s, double dt) { S(s)->advise(dt); }
void nbs_solver_print_info(const nbs_solver* s) { S(s)->print_info(); }

int nbs_solver_rk_configure(nbs_solver* s, size_t max_recursion, size_t substep_subdivisions,
                            size_t refine_steps_count, double error_threshold)
{
	nbody_solver_rk_butcher* b = dynamic_cast<nbody_solver_rk_butcher*>(S(s));
	if(b == nullptr)
	{
		return 0;
	}
	b->set_max_recursion(max_recursion);
	b->set_substep_subdivisions(substep_subdivisions);
	b->set_refine_steps_count(refine_steps_count);
	b->set_error_threshold(error_threshold);
	return 1;
}

int nbs_solver_rk_table(const nbs_solver* s, size_t* steps, int* implicit, int* embedded)
{
	const nbody_solver_rk_butcher* b = dynamic_cast<const nbody_solver_rk_butcher*>(S(s));
	if(b == nullptr)
	{
		return 0;
	}
	*steps = b->table()->get_steps();
	*implicit = b->table()->is_implicit() ? 1 : 0;
	*embedded = b->table()->is_embedded() ? 1 : 0;
	return 1;
}

int nbs_solver_rk_coeffs(const nbs_solver* s, double* a, double* b1, double* b2, double* c)
{
	const nbody_solver_rk_butcher* b = dynamic_cast<const nbody_solver_rk_butcher*>(S(s));
	if(b == nullptr)
	{
		return 0;
	}
	const nbody_butcher_table* t = b->table();
	size_t n = t->get_steps();
	for(size_t i = 0; i < n; ++i)
	{
		size_t jmax = t->is_implicit() ? n : i;
		for(size_t j = 0; j < n; ++j)
		{
			a[i * n + j] = j < jmax ? t->get_a()[i][j] : 0.0;
		}
		b1[i] = t->get_b1()[i];
		b2[i] = t->get_b2()[i];
		c[i] = t->get_c()[i];
	}
	return 1;
}

size_t nbs_solver_rk_substeps(const nbs_solver* s)
{
	const nbody_solver_rk_butcher* b = dynamic_cast<const nbody_solver_rk_butcher*>(S(s));
	return b == nullptr ? 0 : b->substeps_done();
}

nbs_solver* nbs_solver_create_rk_butcher(const char* table)
{
	nbody_butcher_table* t = nbody_create_butcher_table(table == nullptr ? "" : table);
	if(t == nullptr)
	{
		return nullptr;
	}
	return reinterpret_cast<nbs_solver*>(static_cast<nbody_solver*>(new nbody_solver_rk_butcher(t)));
}

nbs_solver* nbs_solver_create_adams(size_t rank, int correction)
{
	if(rank < 1 || rank > 5)
	{
		return nullptr;
	}
	return reinterpret_cast<nbs_solver*>(static_cast<nbody_solver*>(
		new nbody_solver_adams(new nbody_solver_euler(), rank, correction != 0)));
}

nbs_extrapolator* nbs_extrapolator_create(const char* type, nbs_engine* e, size_t order,
                                          const size_t* substeps_count, size_t levels)
{
	std::vector<size_t> sc(substeps_count, substeps_count + levels);
	return reinterpret_cast<nbs_extrapolator*>(nbody_create_extrapolator(type == nullptr ? "" : type, E(e), order, sc));
}

void nbs_extrapolator_destroy(nbs_extrapolator* x) { delete reinterpret_cast<nbody_extrapolator*>(x); }

void nbs_extrapolator_update_table(nbs_extrapolator* x, size_t level, nbs_memory* y)
{
	reinterpret_cast<nbody_extrapolator*>(x)->update_table(level, M(y));
}

double nbs_extrapolator_estimate_error(const nbs_extrapolator* x, size_t level)
{
	return reinterpret_cast<const nbody_extrapolator*>(x)->estimate_error(level);
}

void nbs_extrapolator_extrapolate(const nbs_extrapolator* x, size_t level, nbs_memory* y)
{
	reinterpret_cast<const nbody_extrapolator*>(x)->extrapolate(level, M(y));
}

}  // extern "C"
