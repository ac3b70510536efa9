// nbody_data.h -- bodies, force law and conservation checks of the host layer.
//
// Mirrors the parts of the reference's nbody_data (nbody/nbody_data.h:10-93)
// that the engines, solvers and tests use: state arrays, time/step counters,
// the force law, the P/L/E/V checks, the G1/SI/ADK/Zeno loaders and the G1
// writer, and make_universe.  Qt types are replaced by the standard library.
//
// Licence: this file restates code of drons/nbody (the reference files cited
// in it), which is distributed under the GNU Lesser General Public License,
// version 3; this file is distributed under the same licence.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "nbtype.h"

class nbody_engine;

class nbody_data
{
	size_t                  m_count;
	nbcoord_t               m_time;
	size_t                  m_step;
	size_t                  m_box_size;
	std::string             m_check_list;
	std::vector<nbvertex_t> m_vertites;
	std::vector<nbvertex_t> m_velosites;
	std::vector<nbcoord_t>  m_mass;
	std::vector<nbcoord_t>  m_radius;
	std::vector<nbcolor_t>  m_color;

	nbvertex_t  m_initial_impulce;
	nbvertex_t  m_initial_impulce_moment;
	nbvertex_t  m_initial_mass_center;
	nbcoord_t   m_initial_kinetic_energy;
	nbcoord_t   m_initial_potential_energy;
	nbvertex_t  m_last_total_impulce;
	nbvertex_t  m_last_total_impulce_moment;
	nbvertex_t  m_last_mass_center;
	nbcoord_t   m_last_total_kinetic_energy;
	nbcoord_t   m_last_total_potential_energy;
	bool        m_initial_values_computed;
	size_t      m_prev_compute_count;
	double      m_timer_start;
	size_t      m_timer_step;
	bool        m_quiet;
public:
	nbody_data();

	//! F = -m1 m2 (r1 - r2) / max(|r1 - r2|^2, MinDistance)^(3/2), G = 1
	nbvertex_t force(const nbvertex_t& v1, const nbvertex_t& v2, nbcoord_t mass1, nbcoord_t mass2) const;
	//! -m1 m2 / |r1 - r2|, or 0 when |r1 - r2|^2 < MinDistance
	nbcoord_t potential_energy(const nbvertex_t* vertites, size_t body1, size_t body2) const;

	//! Compute P/L/E/V (the quantities of check_list) for the engine's current
	//! state and print the check line; the first call records initial values.
	void print_statistics(nbody_engine* engine);
	//! The same quantities without printing (P, L, KE, PE, CM, total mass).
	void compute_statistics(nbody_engine* engine, double out[12]);

	bool resize(size_t s);
	void add_body(const nbvertex_t& r, const nbvertex_t& v, const nbcoord_t& m, const nbcoord_t& radius = 0,
	              const nbcolor_t& color = nbcolor_t());
	void advise_time(nbcoord_t dt);
	nbcoord_t get_time() const;
	void set_time(nbcoord_t t);
	size_t get_step() const;
	void set_step(size_t s);
	nbvertex_t* get_vertites();
	const nbvertex_t* get_vertites() const;
	nbvertex_t* get_velosites();
	const nbvertex_t* get_velosites() const;
	nbcoord_t* get_mass();
	const nbcoord_t* get_mass() const;
	nbcoord_t* get_radius();
	const nbcoord_t* get_radius() const;
	nbcolor_t* get_color();
	const nbcolor_t* get_color() const;
	size_t get_box_size() const;
	void set_box_size(size_t s);
	size_t get_count() const;
	void set_check_list(const std::string& check_list);
	const std::string& get_check_list() const;
	void set_quiet(bool quiet);

	void add_galaxy(const nbvertex_t& center, const nbvertex_t& velosity, nbcoord_t radius,
	                nbcoord_t total_mass, size_t count, const nbcolor_t& color = nbcolor_t());
	void make_universe(size_t star_count, nbcoord_t sx, nbcoord_t sy, nbcoord_t sz);

	nbvertex_t get_initial_impulce() const;
	nbvertex_t get_initial_impulce_moment() const;
	nbvertex_t get_initial_mass_center() const;
	nbcoord_t get_initial_energy() const;
	nbvertex_t get_last_total_impulce() const;
	nbvertex_t get_last_total_impulce_moment() const;
	nbvertex_t get_last_mass_center() const;
	nbcoord_t get_last_total_energy() const;
	nbcoord_t get_impulce_err() const;
	nbcoord_t get_impulce_moment_err() const;
	nbcoord_t get_energy_err() const;
	bool is_equal(const nbody_data& other, nbcoord_t eps) const;
	void clear();

	//! G1-style text writer: "X Y Z Vx Vy Vz Mass Radius", %+.16e.
	bool save(const std::string& fn) const;
	//! G1/SI/ADK text reader ("//" comments, >= 7 columns, optional radius).
	bool load(const std::string& fn, e_units_type unit_type = eut_G1);
	bool load_zeno_ascii(const std::string& fn);
	//! type: "Zeno", "G1", "SI" or "ADK" (case-insensitive).
	bool load_initial(const std::string& fn, const std::string& type);
	static nbcoord_t get_mass_factor(e_units_type unit_type);
};
