// nbody_arg_parser.cpp -- see nbody_arg_parser.h.
#include "nbody_arg_parser.h"

#include <string>

#include "nbody_log.h"

nbody_params nbody_parse_arguments(int argc, const char* const* argv)
{
	nbody_params param;
	const std::string prefix("--");
	for(int k = 1; k < argc; ++k)
	{
		const std::string arg(argv[k] != nullptr ? argv[k] : "");
		if(arg.compare(0, prefix.size(), prefix) != 0)
		{
			continue;
		}
		const std::string body = arg.substr(prefix.size());
		const size_t eq = body.find('=');
		if(eq == std::string::npos || body.find('=', eq + 1) != std::string::npos)
		{
			nbody_log("invalid argument format %s", arg.c_str());
			continue;
		}
		param.set(body.substr(0, eq), body.substr(eq + 1));
	}
	return param;
}
