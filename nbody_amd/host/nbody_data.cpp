// nbody_data.cpp -- see nbody_data.h.  Reference counterparts are cited per
// function (paths relative to the drons/nbody repository root).
//
// Licence: this file restates code of drons/nbody (the reference files cited
// in it), which is distributed under the GNU Lesser General Public License,
// version 3; this file is distributed under the same licence.
#include "nbody_data.h"

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>

#include "nbody_engine.h"
#include "nbody_log.h"

namespace {

constexpr nbcoord_t GravityConst = 1.0;   // nbody/nbody_data.cpp:15

double wall_seconds()
{
	using clk = std::chrono::steady_clock;
	return std::chrono::duration<double>(clk::now().time_since_epoch()).count();
}

bool contains(const std::string& s, char c)
{
	return s.find(c) != std::string::npos;
}

// Kahan sum of container[0..size) seeded with container[0]
// (nbody/summation.h:19-45).
template<class T, class F>
T kahan_range(F item, size_t size)
{
	if(size == 0)
	{
		return T();
	}
	T correction = T();
	T sum = item(0);
	for(size_t i = 1; i < size; ++i)
	{
		sum = summation_k(sum, item(i), correction);
	}
	return sum;
}

}  // namespace

nbody_data::nbody_data() :
	m_count(0),
	m_time(0),
	m_step(0),
	m_box_size(0),
	m_check_list("PLV"),
	m_initial_kinetic_energy(0),
	m_initial_potential_energy(0),
	m_last_total_kinetic_energy(0),
	m_last_total_potential_energy(0),
	m_initial_values_computed(false),
	m_prev_compute_count(0),
	m_timer_start(wall_seconds()),
	m_timer_step(0),
	m_quiet(false)
{
}

// nbody/nbody_data.cpp:35-44
nbvertex_t nbody_data::force(const nbvertex_t& v1, const nbvertex_t& v2, nbcoord_t mass1, nbcoord_t mass2) const
{
	nbvertex_t dr(v1 - v2);
	nbcoord_t r2(dr.norm());
	if(r2 < nbody::MinDistance)
	{
		r2 = nbody::MinDistance;
	}
	return dr * ((-GravityConst * mass1 * mass2) / (r2 * std::sqrt(r2)));
}

// nbody/nbody_data.cpp:46-55
nbcoord_t nbody_data::potential_energy(const nbvertex_t* vertites, size_t body1, size_t body2) const
{
	nbvertex_t dr(vertites[body1] - vertites[body2]);
	nbcoord_t r2(dr.norm());
	if(r2 < nbody::MinDistance)
	{
		return 0;
	}
	return -(GravityConst * m_mass[body1] * m_mass[body2]) / std::sqrt(r2);
}

// The P/L/E/V sums of nbody/nbody_data.cpp:70-102 (proxies of
// nbody/summation_proxy.h:6-87), on the device when the engine offers it.
void nbody_data::compute_statistics(nbody_engine* engine, double out[12])
{
	const bool want_e = contains(m_check_list, 'E');
	if(engine != nullptr && engine->compute_statistics(want_e, out))
	{
		if(!want_e)
		{
			out[6] = 0.0;
			out[7] = 0.0;
		}
		return;
	}
	std::memset(out, 0, 12 * sizeof(double));
	const size_t n = m_count;
	const nbvertex_t* pos = m_vertites.data();
	const nbvertex_t* vel = m_velosites.data();
	const nbcoord_t* m = m_mass.data();
	nbvertex_t p = kahan_range<nbvertex_t>([&](size_t i) { return vel[i] * m[i]; }, n);
	nbvertex_t l = kahan_range<nbvertex_t>([&](size_t i) { return pos[i] ^ (vel[i] * m[i]); }, n);
	out[0] = p.x; out[1] = p.y; out[2] = p.z;
	out[3] = l.x; out[4] = l.y; out[5] = l.z;
	if(want_e)
	{
		nbcoord_t ke = kahan_range<nbcoord_t>([&](size_t i) { return vel[i].norm() * m[i]; }, n);
		nbcoord_t pe = kahan_range<nbcoord_t>([&](size_t k)
		{
			size_t n1 = k / n;
			size_t n2 = k % n;
			return n1 == n2 ? 0.0 : potential_energy(pos, n1, n2);
		}, n * n);
		out[6] = ke / 2;
		out[7] = pe / 2;
	}
	nbcoord_t total_mass = kahan_range<nbcoord_t>([&](size_t i) { return m[i]; }, n);
	nbvertex_t cm = kahan_range<nbvertex_t>([&](size_t i) { return pos[i] * m[i]; }, n);
	cm /= total_mass;
	out[8] = cm.x; out[9] = cm.y; out[10] = cm.z;
	out[11] = total_mass;
}

// nbody/nbody_data.cpp:57-145
void nbody_data::print_statistics(nbody_engine* engine)
{
	size_t compute_count = 0;
	if(engine != nullptr)
	{
		engine->get_data(this);
		compute_count = engine->get_compute_count();
	}
	double timer_end = wall_seconds();
	double st[12];
	compute_statistics(engine, st);

	nbvertex_t total_impulce(st[0], st[1], st[2]);
	nbvertex_t total_impulce_moment(st[3], st[4], st[5]);
	nbcoord_t total_kinetic_energy = st[6];
	nbcoord_t total_potential_energy = st[7];
	nbvertex_t mass_center(st[8], st[9], st[10]);
	if(!contains(m_check_list, 'P')) { total_impulce = nbvertex_t(); }
	if(!contains(m_check_list, 'L')) { total_impulce_moment = nbvertex_t(); }
	if(!contains(m_check_list, 'V')) { mass_center = nbvertex_t(); }

	if(!m_initial_values_computed)
	{
		m_initial_impulce = total_impulce;
		m_initial_impulce_moment = total_impulce_moment;
		m_initial_mass_center = mass_center;
		m_initial_kinetic_energy = total_kinetic_energy;
		m_initial_potential_energy = total_potential_energy;
		m_initial_values_computed = true;
	}

	char line[512];
	int len = std::snprintf(line, sizeof(line), "# %08zu t %6.6f CC %08zu SC %8zu",
	                        m_step, m_time, compute_count, compute_count - m_prev_compute_count);
	m_prev_compute_count = compute_count;
	if(contains(m_check_list, 'P'))
	{
		m_last_total_impulce = total_impulce;
		len += std::snprintf(line + len, sizeof(line) - len, " dP %4.3e", get_impulce_err());
	}
	if(contains(m_check_list, 'L'))
	{
		m_last_total_impulce_moment = total_impulce_moment;
		len += std::snprintf(line + len, sizeof(line) - len, " dL %4.3e", get_impulce_moment_err());
	}
	if(contains(m_check_list, 'V'))
	{
		m_last_mass_center = mass_center;
		nbvertex_t mass_center_vel((get_last_mass_center() - get_initial_mass_center()) / m_time);
		len += std::snprintf(line + len, sizeof(line) - len, " Vcm %4.3e", mass_center_vel.length());
	}
	if(contains(m_check_list, 'E'))
	{
		m_last_total_kinetic_energy = total_kinetic_energy;
		m_last_total_potential_energy = total_potential_energy;
		len += std::snprintf(line + len, sizeof(line) - len, " dE %4.3e", get_energy_err());
	}
	double dt = static_cast<double>(m_step - m_timer_step);
	std::snprintf(line + len, sizeof(line) - len, " St %g Wt %g",
	              (timer_end - m_timer_start) / dt, (wall_seconds() - m_timer_start) / dt);
	if(!m_quiet)
	{
		nbody_print("%s", line);
	}
	m_timer_start = wall_seconds();
	m_timer_step = m_step;
}

bool nbody_data::resize(size_t s)
{
	m_vertites.resize(s);
	m_velosites.resize(s);
	m_mass.resize(s);
	m_radius.resize(s);
	m_color.resize(s);
	m_count = s;
	return true;
}

void nbody_data::add_body(const nbvertex_t& r, const nbvertex_t& v, const nbcoord_t& m, const nbcoord_t& radius,
                          const nbcolor_t& color)
{
	m_vertites.push_back(r);
	m_velosites.push_back(v);
	m_mass.push_back(m);
	m_radius.push_back(radius);
	m_color.push_back(color);
	++m_count;
}

void nbody_data::advise_time(nbcoord_t dt)
{
	m_time += dt;
	++m_step;
}

nbcoord_t nbody_data::get_time() const { return m_time; }
void nbody_data::set_time(nbcoord_t t) { m_time = t; }
size_t nbody_data::get_step() const { return m_step; }
void nbody_data::set_step(size_t s) { m_step = s; }
nbvertex_t* nbody_data::get_vertites() { return m_vertites.data(); }
const nbvertex_t* nbody_data::get_vertites() const { return m_vertites.data(); }
nbvertex_t* nbody_data::get_velosites() { return m_velosites.data(); }
const nbvertex_t* nbody_data::get_velosites() const { return m_velosites.data(); }
nbcoord_t* nbody_data::get_mass() { return m_mass.data(); }
const nbcoord_t* nbody_data::get_mass() const { return m_mass.data(); }
nbcoord_t* nbody_data::get_radius() { return m_radius.data(); }
const nbcoord_t* nbody_data::get_radius() const { return m_radius.data(); }
nbcolor_t* nbody_data::get_color() { return m_color.data(); }
const nbcolor_t* nbody_data::get_color() const { return m_color.data(); }
size_t nbody_data::get_box_size() const { return m_box_size; }
void nbody_data::set_box_size(size_t s) { m_box_size = s; }
size_t nbody_data::get_count() const { return m_count; }
void nbody_data::set_check_list(const std::string& check_list) { m_check_list = check_list; }
const std::string& nbody_data::get_check_list() const { return m_check_list; }
void nbody_data::set_quiet(bool quiet) { m_quiet = quiet; }

static size_t round_up(size_t n, size_t factor)
{
	return factor * (1 + (n - 1) / factor);
}

// nbody/nbody_data.cpp:263-303: a disk galaxy around a 99.9% central mass,
// drawn with the default-seeded std::mt19937_64.
void nbody_data::add_galaxy(const nbvertex_t& center, const nbvertex_t& velosity, nbcoord_t radius,
                            nbcoord_t total_mass, size_t count, const nbcolor_t& color)
{
	std::mt19937_64 rng;
	std::uniform_real_distribution<double> uniform(-radius, radius);
	count = round_up(count, NBODY_DATA_BLOCK_SIZE) - 1;

	nbvertex_t up(0, 0, 1);
	nbvertex_t right(1, 0, 0);
	nbcoord_t black_hole_mass_ratio = 0.999;
	nbcoord_t black_hole_mass = total_mass * black_hole_mass_ratio;
	nbcoord_t star_mass = (total_mass - black_hole_mass) / static_cast<nbcoord_t>(count);
	nbcoord_t all_stars_mass = static_cast<nbcoord_t>(count) * star_mass;

	add_body(center, velosity, black_hole_mass, 0, nbcolor_t(0, 1, 0, 1));

	for(size_t n = 0; n != count; ++n)
	{
		nbvertex_t r(static_cast<nbcoord_t>(uniform(rng)),
		             static_cast<nbcoord_t>(uniform(rng)),
		             static_cast<nbcoord_t>(uniform(rng)));
		nbcoord_t rlen(r.length());
		if(rlen > radius)
		{
			--n;
			continue;
		}
		r.z *= 0.3;
		r += center;
		nbvertex_t v((r - center) ^ up);
		if(v.length() < 1e-7)
		{
			v = (r - center) ^ right;
		}
		v.normalize();
		nbcoord_t effective_mass = std::pow(rlen / radius, 3.0) * all_stars_mass + black_hole_mass;
		v *= std::sqrt(force(r, center, star_mass, effective_mass).length() * (r - center).length() / star_mass);
		add_body(r, v + velosity, star_mass, 0, color);
	}
}

// nbody/nbody_data.cpp:305-324
void nbody_data::make_universe(size_t star_count, nbcoord_t sx, nbcoord_t sy, nbcoord_t sz)
{
	clear();
	nbcoord_t radius = sx * 0.5;
	nbcoord_t galaxy_mass = 1000;
	nbvertex_t center(sx * 0.5, sy * 0.5, sz * 0.5);
	nbvertex_t base(radius, 0, 0);
	nbvertex_t velosity(0, std::sqrt(force(nbvertex_t(), base, galaxy_mass, galaxy_mass).length() *
	                                 (base).length() / (2 * galaxy_mass)), 0);
	const float intensity = 0.8f;
	add_galaxy(center - base, velosity / 3, radius, galaxy_mass, star_count,
	           nbcolor_t(intensity * 0.5f, intensity * 0.5f, intensity, 1.0f));
	add_galaxy(center + base, -velosity / 3, radius, galaxy_mass, star_count,
	           nbcolor_t(intensity, intensity * 0.5f, intensity * 0.5f, 1.0f));
	m_box_size = static_cast<size_t>(std::max(std::max(static_cast<nbcoord_t>(1), sx), std::max(sy, sz)));
}

nbvertex_t nbody_data::get_initial_impulce() const { return m_initial_impulce; }
nbvertex_t nbody_data::get_initial_impulce_moment() const { return m_initial_impulce_moment; }
nbvertex_t nbody_data::get_initial_mass_center() const { return m_initial_mass_center; }
nbcoord_t nbody_data::get_initial_energy() const { return m_initial_kinetic_energy + m_initial_potential_energy; }
nbvertex_t nbody_data::get_last_total_impulce() const { return m_last_total_impulce; }
nbvertex_t nbody_data::get_last_total_impulce_moment() const { return m_last_total_impulce_moment; }
nbvertex_t nbody_data::get_last_mass_center() const { return m_last_mass_center; }
nbcoord_t nbody_data::get_last_total_energy() const
{
	return m_last_total_kinetic_energy + m_last_total_potential_energy;
}

// nbody/nbody_data.cpp:366-381 (percent of the initial magnitude)
nbcoord_t nbody_data::get_impulce_err() const
{
	return std::fabs(100.0 * (get_last_total_impulce() - get_initial_impulce()).length() /
	                 get_initial_impulce().length());
}

nbcoord_t nbody_data::get_impulce_moment_err() const
{
	return std::fabs(100.0 * (get_last_total_impulce_moment() - get_initial_impulce_moment()).length() /
	                 get_initial_impulce_moment().length());
}

nbcoord_t nbody_data::get_energy_err() const
{
	return std::fabs(100.0 * (get_last_total_energy() - get_initial_energy()) / get_initial_energy());
}

// nbody/nbody_data.cpp:383-405
bool nbody_data::is_equal(const nbody_data& other, nbcoord_t eps) const
{
	if(m_count != other.m_count)
	{
		return false;
	}
	for(size_t n = 0; n != m_count; ++n)
	{
		if(std::fabs(m_mass[n] - other.m_mass[n]) > eps ||
		   std::fabs(m_radius[n] - other.m_radius[n]) > eps ||
		   std::fabs(m_vertites[n].x - other.m_vertites[n].x) > eps ||
		   std::fabs(m_vertites[n].y - other.m_vertites[n].y) > eps ||
		   std::fabs(m_vertites[n].z - other.m_vertites[n].z) > eps ||
		   std::fabs(m_velosites[n].x - other.m_velosites[n].x) > eps ||
		   std::fabs(m_velosites[n].y - other.m_velosites[n].y) > eps ||
		   std::fabs(m_velosites[n].z - other.m_velosites[n].z) > eps)
		{
			return false;
		}
	}
	return true;
}

void nbody_data::clear()
{
	m_vertites.clear();
	m_velosites.clear();
	m_mass.clear();
	m_radius.clear();
	m_color.clear();
	m_count = 0;
}

// nbody/nbody_data.cpp:407-425: 17 significant digits, forced sign, so the
// text round-trips the FP64 values exactly.
bool nbody_data::save(const std::string& fn) const
{
	FILE* f = std::fopen(fn.c_str(), "w");
	if(f == nullptr)
	{
		nbody_log("Can't open file %s", fn.c_str());
		return false;
	}
	for(size_t n = 0; n != m_count; ++n)
	{
		std::fprintf(f, "%+.16e %+.16e %+.16e %+.16e %+.16e %+.16e %+.16e %+.16e\n",
		             m_vertites[n].x, m_vertites[n].y, m_vertites[n].z,
		             m_velosites[n].x, m_velosites[n].y, m_velosites[n].z,
		             m_mass[n], m_radius[n]);
	}
	std::fclose(f);
	return true;
}

static bool parse_double(const std::string& s, double* out)
{
	if(s.empty())
	{
		return false;
	}
	const char* b = s.c_str();
	char* e = nullptr;
	double v = std::strtod(b, &e);
	if(e == b || *e != '\0')
	{
		return false;
	}
	*out = v;
	return true;
}

static std::vector<std::string> split_ws(const std::string& line)
{
	std::vector<std::string> out;
	std::istringstream is(line);
	std::string tok;
	while(is >> tok)
	{
		out.push_back(tok);
	}
	return out;
}

// nbody/nbody_data.cpp:427-486
bool nbody_data::load(const std::string& fn, e_units_type unit_type)
{
	std::ifstream file(fn);
	if(!file)
	{
		nbody_log("Can't open file %s", fn.c_str());
		return false;
	}
	clear();
	nbcoord_t mass_factor = get_mass_factor(unit_type);
	std::string line;
	for(size_t line_n = 0; std::getline(file, line); ++line_n)
	{
		size_t comment = line.find("//");
		if(comment != std::string::npos)
		{
			line = line.substr(0, comment);
		}
		std::vector<std::string> columns = split_ws(line);
		if(columns.empty())
		{
			continue;
		}
		if(columns.size() < 7)
		{
			nbody_log("Failed to parse line %zu <%s>: format is 'X Y Z Vx Vy Vz Mass <Radius>'",
			          line_n, line.c_str());
			return false;
		}
		double rvm[7];
		for(int n = 0; n != 7; ++n)
		{
			if(!parse_double(columns[n], &rvm[n]))
			{
				nbody_log("Can't convert column %d to double at line %zu", n, line_n);
				return false;
			}
		}
		double radius = 0;
		if(columns.size() > 7 && !parse_double(columns[7], &radius))
		{
			nbody_log("Can't convert 'radius' column to double at line %zu", line_n);
			return false;
		}
		add_body(nbvertex_t(rvm[0], rvm[1], rvm[2]), nbvertex_t(rvm[3], rvm[4], rvm[5]),
		         rvm[6] * mass_factor, radius);
	}
	return true;
}

// nbody/nbody_data.cpp:488-573
bool nbody_data::load_zeno_ascii(const std::string& fn)
{
	std::ifstream file(fn);
	if(!file)
	{
		nbody_log("Can't open file %s", fn.c_str());
		return false;
	}
	clear();
	std::string line;
	unsigned long long count = 0;
	unsigned long long dim = 0;
	if(!std::getline(file, line) || std::sscanf(line.c_str(), "%llu", &count) != 1)
	{
		nbody_log("Can't read body count from %s", fn.c_str());
		return false;
	}
	if(!std::getline(file, line) || std::sscanf(line.c_str(), "%llu", &dim) != 1)
	{
		nbody_log("Can't read body dimensions from %s", fn.c_str());
		return false;
	}
	if(dim != 3)
	{
		nbody_log("Invalid dimension count %llu in %s", dim, fn.c_str());
		return false;
	}
	if(!std::getline(file, line) || line.empty())
	{
		return false;
	}
	resize(count);
	for(size_t i = 0; i != count; ++i)
	{
		if(!std::getline(file, line))
		{
			nbody_log("Unexpected data stream end");
			return false;
		}
		m_mass[i] = std::strtod(line.c_str(), nullptr);
	}
	for(int part = 0; part < 2; ++part)
	{
		for(size_t i = 0; i != count; ++i)
		{
			if(!std::getline(file, line))
			{
				nbody_log("Unexpected data stream end");
				return false;
			}
			std::vector<std::string> cols = split_ws(line);
			if(cols.size() != 3)
			{
				nbody_log("Invalid vector dimensions");
				return false;
			}
			nbvertex_t v(std::strtod(cols[0].c_str(), nullptr), std::strtod(cols[1].c_str(), nullptr),
			             std::strtod(cols[2].c_str(), nullptr));
			(part == 0 ? m_vertites : m_velosites)[i] = v;
		}
	}
	m_box_size = 1;
	return true;
}

// nbody/nbody_data.cpp:575-616
bool nbody_data::load_initial(const std::string& fn, const std::string& type)
{
	std::string t(type);
	std::transform(t.begin(), t.end(), t.begin(), [](unsigned char c) { return std::tolower(c); });
	bool ret = false;
	if(t == "zeno") { ret = load_zeno_ascii(fn); }
	else if(t == "g1") { ret = load(fn, eut_G1); }
	else if(t == "si") { ret = load(fn, eut_SI); }
	else if(t == "adk") { ret = load(fn, eut_au_day_kg); }
	else
	{
		nbody_log("Unknown initial state type. Possible values are: Zeno, G1, SI, ADK.");
		return false;
	}
	if(!ret)
	{
		return false;
	}
	nbvertex_t r(1, 1, 1);
	for(const auto& v : m_vertites)
	{
		r.x = std::max(v.x, r.x);
		r.y = std::max(v.y, r.y);
		r.z = std::max(v.z, r.z);
	}
	m_box_size = static_cast<size_t>(std::max(r.x, std::max(r.y, r.z)));
	return ret;
}

nbcoord_t nbody_data::get_mass_factor(e_units_type unit_type)
{
	switch(unit_type)
	{
	case eut_G1:
		return 1.0;
	case eut_SI:
		return nbody::GravityConstSI;
	case eut_au_day_kg:
		return nbody::GravityConstAuDayKg;
	}
	return 0.0;
}
