// nbody_arg_parser.h -- command-line parameters of the front ends.
//
// nbody_parse_arguments (nbody/nbody_arg_parser.cpp:5-28): every argument of
// the form --key=value becomes one parameter; arguments without the "--"
// prefix are skipped, and one with "--" but not exactly one '=' is logged and
// skipped.  The keys are those of the engine and solver factories plus the
// front end's own (max_time, dump_step, output, ...).
#pragma once

#include "nbody_params.h"

nbody_params nbody_parse_arguments(int argc, const char* const* argv);
