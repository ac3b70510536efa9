// nbody_generators.h -- seeded synthetic initial states of the benchmark
// workloads (G = 1 units).  Random numbers come from std::mt19937_64(seed)
// with explicit conversions (53-bit uniforms, Box-Muller normals), so the
// bodies are identical on every platform with the same seed.
#pragma once

#include <cstdint>
#include <string>

#include "nbody_data.h"

//! Uniform in the unit ball (rejection from [-1,1]^3), velocities N(0, sigma^2)
//! per component, masses 1/n.
void nbody_gen_ball(nbody_data* d, size_t n, uint64_t seed, double sigma);
//! Plummer sphere, scale radius a, total mass M, radii truncated at rmax*a;
//! positions by inversion of the cumulative mass, speeds by the rejection
//! method of Aarseth, Henon & Wielen (1974), isotropic directions.  Bodies are
//! shifted by `center` and `velocity`.
void nbody_gen_plummer(nbody_data* d, size_t n, uint64_t seed, double a, double M, double rmax,
                       const nbvertex_t& center, const nbvertex_t& velocity);
//! Two Plummer spheres (a = 1, M = 1 each) of n/2 bodies with independent
//! seeds, centred at (-sep/2,0,0) and (+sep/2,0,0), approaching along x with
//! relative speed vrel.
void nbody_gen_two_plummer(nbody_data* d, size_t n, uint64_t seed, double sep, double vrel);
//! Cold uniform cube [-h, h]^3, zero velocities, masses lognormal(0, sigma)
//! normalised to a total of 1.
void nbody_gen_cube_lognormal(nbody_data* d, size_t n, uint64_t seed, double h, double sigma);
//! Dispatch by name ("ball", "plummer", "two_plummer", "cube"); false if unknown.
bool nbody_generate(nbody_data* d, const std::string& kind, size_t n, uint64_t seed);
