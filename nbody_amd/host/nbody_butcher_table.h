// nbody_butcher_table.h -- Runge-Kutta tableau (a, b1, b2, c) with the
// accessors of the reference's nbody_butcher_table (nbody/nbody_butcher_table.h:7-22).
//
// Licence: this file restates code of drons/nbody (the reference files cited
// in it), which is distributed under the GNU Lesser General Public License,
// version 3; this file is distributed under the same licence.
#pragma once

#include <string>
#include <vector>

#include "nbtype.h"

class nbody_butcher_table
{
	std::string                   m_name;
	size_t                        m_steps;
	bool                          m_implicit;
	bool                          m_embedded;
	std::vector<const nbcoord_t*> m_rows;
	const nbcoord_t*              m_b1;
	const nbcoord_t*              m_b2;
	const nbcoord_t*              m_c;
public:
	nbody_butcher_table(const char* name, size_t steps, bool implicit, bool embedded,
	                    const nbcoord_t* a, const nbcoord_t* b1, const nbcoord_t* b2, const nbcoord_t* c);
	const std::string& name() const { return m_name; }
	size_t get_steps() const { return m_steps; }
	//! Row i holds a[i][0..i) (explicit) or a[i][0..steps) (implicit).
	const nbcoord_t** get_a() const { return const_cast<const nbcoord_t**>(m_rows.data()); }
	const nbcoord_t* get_b1() const { return m_b1; }
	const nbcoord_t* get_b2() const { return m_b2; }
	const nbcoord_t* get_c() const { return m_c; }
	bool is_implicit() const { return m_implicit; }
	bool is_embedded() const { return m_embedded; }
};

//! "rk4", "rkck", "rkdp", "rkdverk", "rkf", "rkfeagin10", "rkfeagin12",
//! "rkfeagin14", "rkgl", "rklc"; NULL for anything else.
nbody_butcher_table* nbody_create_butcher_table(const std::string& name);
