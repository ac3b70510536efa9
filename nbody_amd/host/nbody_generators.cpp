// nbody_generators.cpp -- see nbody_generators.h.
#include "nbody_generators.h"

#include <cmath>
#include <random>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;

struct rng_t
{
	std::mt19937_64 g;
	explicit rng_t(uint64_t seed) : g(seed) {}
	// uniform in [0, 1) with 53 random bits
	double uniform() { return static_cast<double>(g() >> 11) * (1.0 / 9007199254740992.0); }
	// uniform in (0, 1)
	double uniform_open()
	{
		double u;
		do { u = uniform(); } while(u == 0.0);
		return u;
	}
	double normal()
	{
		double u1 = uniform_open();
		double u2 = uniform();
		return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
	}
	nbvertex_t isotropic(double len)
	{
		double z = (1.0 - 2.0 * uniform()) * len;
		double rxy = std::sqrt(std::fmax(len * len - z * z, 0.0));
		double phi = 2.0 * kPi * uniform();
		return nbvertex_t(rxy * std::cos(phi), rxy * std::sin(phi), z);
	}
};

}  // namespace

void nbody_gen_ball(nbody_data* d, size_t n, uint64_t seed, double sigma)
{
	d->clear();
	rng_t r(seed);
	double m = 1.0 / static_cast<double>(n);
	for(size_t i = 0; i < n; ++i)
	{
		nbvertex_t p;
		do
		{
			p = nbvertex_t(2.0 * r.uniform() - 1.0, 2.0 * r.uniform() - 1.0, 2.0 * r.uniform() - 1.0);
		}
		while(p.norm() > 1.0);
		nbvertex_t v(sigma * r.normal(), sigma * r.normal(), sigma * r.normal());
		d->add_body(p, v, m);
	}
	d->set_box_size(1);
}

void nbody_gen_plummer(nbody_data* d, size_t n, uint64_t seed, double a, double M, double rmax,
                       const nbvertex_t& center, const nbvertex_t& velocity)
{
	rng_t r(seed);
	double m = M / static_cast<double>(n);
	for(size_t i = 0; i < n; ++i)
	{
		double rad;
		do
		{
			double x = r.uniform_open();
			rad = a / std::sqrt(std::pow(x, -2.0 / 3.0) - 1.0);
		}
		while(!(rad <= rmax * a));
		// speed in units of the local escape speed: q with density q^2 (1 - q^2)^(7/2)
		double q, g;
		do
		{
			q = r.uniform();
			g = 0.1 * r.uniform();
		}
		while(g > q * q * std::pow(1.0 - q * q, 3.5));
		double vesc = std::sqrt(2.0 * M / a) * std::pow(1.0 + (rad * rad) / (a * a), -0.25);
		nbvertex_t p = r.isotropic(rad) + center;
		nbvertex_t v = r.isotropic(q * vesc) + velocity;
		d->add_body(p, v, m);
	}
}

void nbody_gen_two_plummer(nbody_data* d, size_t n, uint64_t seed, double sep, double vrel)
{
	d->clear();
	size_t h = n / 2;
	nbody_gen_plummer(d, h, seed * 2 + 1, 1.0, 1.0, 10.0, nbvertex_t(-0.5 * sep, 0, 0), nbvertex_t(0.5 * vrel, 0, 0));
	nbody_gen_plummer(d, n - h, seed * 2 + 2, 1.0, 1.0, 10.0, nbvertex_t(0.5 * sep, 0, 0), nbvertex_t(-0.5 * vrel, 0, 0));
	d->set_box_size(static_cast<size_t>(sep + 20));
}

void nbody_gen_cube_lognormal(nbody_data* d, size_t n, uint64_t seed, double h, double sigma)
{
	d->clear();
	rng_t r(seed);
	std::vector<double> m(n);
	std::vector<nbvertex_t> p(n);
	double total = 0.0;
	for(size_t i = 0; i < n; ++i)
	{
		p[i] = nbvertex_t(h * (2.0 * r.uniform() - 1.0), h * (2.0 * r.uniform() - 1.0), h * (2.0 * r.uniform() - 1.0));
		m[i] = std::exp(sigma * r.normal());
		total += m[i];
	}
	for(size_t i = 0; i < n; ++i)
	{
		d->add_body(p[i], nbvertex_t(), m[i] / total);
	}
	d->set_box_size(static_cast<size_t>(std::ceil(h)));
}

bool nbody_generate(nbody_data* d, const std::string& kind, size_t n, uint64_t seed)
{
	if(kind == "ball")
	{
		nbody_gen_ball(d, n, seed, 0.3);
	}
	else if(kind == "plummer")
	{
		d->clear();
		nbody_gen_plummer(d, n, seed, 1.0, 1.0, 10.0, nbvertex_t(), nbvertex_t());
		d->set_box_size(10);
	}
	else if(kind == "two_plummer")
	{
		nbody_gen_two_plummer(d, n, seed, 10.0, 0.5);
	}
	else if(kind == "cube")
	{
		nbody_gen_cube_lognormal(d, n, seed, 1.0, 0.5);
	}
	else
	{
		return false;
	}
	d->set_time(0);
	d->set_step(0);
	return true;
}
