// nbody-simulation -- command-line front end of the engines and solvers, with
// the arguments of the reference's nbody-simulation (test/main/main.cpp:12-123):
//
//   --engine=hip|hip_bh|...  --solver=rk4|...  and every other factory key
//   --stars_count=64 --box_size=100   make_universe when neither of the next two is given
//   --initial_state=FILE --initial_type=Zeno|G1|SI|ADK
//   --resume=BASE                     continue the stream BASE from its last frame
//   --output=/tmp/nbody/main-stream   stream base name (frames every dump_step)
//   --max_part_size=1073741824 --max_time=1 --dump_step=0.01 --check_step=0.1
//   --check_list=PLV --verbose=0
//
// Exit status: 0 on success, 255 (-1) when the engine, solver, state or
// stream cannot be set up or a frame cannot be written.
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>

#include "nbody_arg_parser.h"
#include "nbody_data.h"
#include "nbody_data_stream.h"
#include "nbody_engines.h"
#include "nbody_log.h"
#include "nbody_solver.h"

namespace {

double number(const nbody_params& p, const char* key, double def)
{
	double v = def;
	if(!p.get_double(key, def, &v))
	{
		nbody_log("argument --%s=%s is not a number; using %g", key, p.value(key, "").c_str(), def);
		return def;
	}
	return v;
}

}  // namespace

int main(int argc, char** argv)
{
	const nbody_params param = nbody_parse_arguments(argc, argv);
	const double box_size = number(param, "box_size", 100);
	const double max_part_size = number(param, "max_part_size", 1 << 30);
	const double max_time = number(param, "max_time", 1);
	const double dump_step = number(param, "dump_step", 1e-2);
	const double check_step = number(param, "check_step", 1e-1);
	const std::string check_list = param.value("check_list", "PLV");
	std::string output = param.value("output", "/tmp/nbody/main-stream");
	const std::string resume = param.value("resume", "");
	const std::string initial_state = param.value("initial_state", "");
	const std::string initial_type = param.value("initial_type", "Zeno");

	nbody_data data;
	std::unique_ptr<nbody_data_stream> stream(new nbody_data_stream());
	std::unique_ptr<nbody_data_stream_reader> resume_stream;

	if(!resume.empty())
	{
		resume_stream.reset(new nbody_data_stream_reader());
		if(resume_stream->load(resume) != 0)
		{
			nbody_log("can't open stream to resume %s", resume.c_str());
			return -1;
		}
		if(resume_stream->seek(resume_stream->get_frame_count() - 1) != 0)
		{
			nbody_log("can't seek to the end of stream %s", resume.c_str());
			return -1;
		}
		data.resize(resume_stream->get_body_count());
		if(resume_stream->read(&data) != 0)
		{
			nbody_log("can't read the last frame from stream %s", resume.c_str());
			return -1;
		}
		data.set_box_size(static_cast<size_t>(resume_stream->get_box_size()));
		output = resume;
	}
	else if(!initial_state.empty())
	{
		if(!data.load_initial(initial_state, initial_type))
		{
			nbody_log("can't load initial state %s", initial_state.c_str());
			return -1;
		}
	}
	else
	{
		double stars = number(param, "stars_count", 64);
		if(stars < 1)
		{
			stars = 64;
		}
		data.make_universe(static_cast<size_t>(stars), box_size, box_size, box_size);
	}

	std::unique_ptr<nbody_engine> engine(nbody_create_engine(param));
	if(engine == nullptr)
	{
		nbody_log("can't create engine %s", param.value("engine", "").c_str());
		return -1;
	}
	std::unique_ptr<nbody_solver> solver(nbody_create_solver(param));
	if(solver == nullptr)
	{
		nbody_log("can't create solver %s", param.value("solver", "").c_str());
		return -1;
	}
	if(!engine->init(&data))
	{
		nbody_log("can't init engine %s", engine->type_name());
		return -1;
	}
	solver->set_engine(engine.get());
	data.set_check_list(check_list);

	if(number(param, "verbose", 0) != 0)
	{
		nbody_print("General:");
		nbody_print("\tStars count: %zu", data.get_count());
		nbody_print("\tBox size: %zu", data.get_box_size());
		nbody_print("\toutput: %s", output.c_str());
		if(!resume.empty())
		{
			nbody_print("\tresume: %s", resume.c_str());
		}
		nbody_print("\tinitial_state: %s", initial_state.c_str());
		nbody_print("\tmax_part_size: %.0f", max_part_size);
		nbody_print("\tmax_time: %g", max_time);
		nbody_print("\tdump_step: %g", dump_step);
		nbody_print("\tcheck_step: %g", check_step);
		nbody_print("\tcheck_list: %s", check_list.c_str());
		nbody_print("Solver: %s", solver->type_name());
		solver->print_info();
		nbody_print("Engine: %s", engine->type_name());
		engine->print_info();
	}

	if(stream->open(output, static_cast<int64_t>(max_part_size), resume_stream.get()) != 0)
	{
		nbody_log("fail to open stream %s", output.c_str());
		return -1;
	}
	if(resume_stream != nullptr)
	{
		resume_stream->close();
	}
	const int rc = solver->run(&data, stream.get(), max_time, dump_step, check_step);
	stream->close();
	solver.reset();
	engine.reset();
	return rc;
}
