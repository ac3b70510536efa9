// nbody-bench -- the reference's benchmark front end (test/bench/main.cpp:36-216,
// harness test/bench/bench.h:60-120 and 308-360) for the engines of this build:
// each parameter set is run with euler to max_time for each value of the swept
// key, and the table gives the wall time per step (the reference's "time"),
// in the reference's txt layout.
//
//   --bench=gpu             hip (fp64, fp32) and hip_bh (body walk, group walk,
//                           two-level walk) over stars_count 8K..256K
//   --bench=gpu_tree_ratio  hip_bh variants over distance_to_node_radius_ratio
//                           0.1 .. 1024 (--ratio_count=k keeps the first k)
//   --bench=gpu_tree_block  hip fp64 over its work-group sizes 64 .. 1024
//   --stars_count=N (tree benches, default 131072), --max_time=T (default as
//   the reference: 1 for gpu, 0.1 for the tree benches), --values=a,b,... to
//   replace the swept values.
#include <chrono>
#include <cstdio>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "nbody_arg_parser.h"
#include "nbody_data.h"
#include "nbody_engines.h"
#include "nbody_log.h"
#include "nbody_solver.h"

namespace {

struct entry
{
	std::string                                      name;
	std::vector<std::pair<std::string, std::string>> keys;
};

nbody_params make(const entry& e, const std::string& field, const std::string& value)
{
	nbody_params p;
	for(const auto& kv : e.keys)
	{
		p.set(kv.first, kv.second);
	}
	p.set("solver", "euler");
	p.set(field, value);
	return p;
}

// Seconds per step of one run (nbody bench.h run(): time / engine step), or -1.
double run_one(const nbody_params& p, double max_time)
{
	double stars = 1024;
	p.get_double("stars_count", 1024, &stars);
	nbody_data data;
	data.set_quiet(true);
	data.make_universe(static_cast<size_t>(stars) / 2, 100, 100, 100);
	std::unique_ptr<nbody_engine> engine(nbody_create_engine(p));
	std::unique_ptr<nbody_solver> solver(nbody_create_solver(p));
	if(engine == nullptr || solver == nullptr || !engine->init(&data))
	{
		nbody_log("bench: can't set up %s", p.str().c_str());
		return -1;
	}
	solver->set_engine(engine.get());
	const auto t0 = std::chrono::steady_clock::now();
	if(solver->run(&data, nullptr, max_time, 0, max_time) != 0 || engine->get_step() == 0)
	{
		return -1;
	}
	const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
	return wall / static_cast<double>(engine->get_step());
}

std::vector<std::string> split(const std::string& s)
{
	std::vector<std::string> out;
	std::stringstream ss(s);
	std::string item;
	while(std::getline(ss, item, ','))
	{
		if(!item.empty()) { out.push_back(item); }
	}
	return out;
}

}  // namespace

int main(int argc, char** argv)
{
	const nbody_params param = nbody_parse_arguments(argc, argv);
	const std::string bench = param.value("bench", "gpu");
	const std::string stars = param.value("stars_count", "131072");
	std::vector<entry> sets;
	std::vector<std::string> values;
	std::string field;
	double max_time = 0.1;
	if(bench == "gpu")
	{
		sets = {{"hip", {{"engine", "hip"}}},
		        {"hip+fp32", {{"engine", "hip"}, {"force_precision", "fp32"}}},
		        {"hip_bh+heap_stackless", {{"engine", "hip_bh"}, {"tree_layout", "heap_stackless"},
		                                   {"traverse_type", "nested_tree"}}},
		        {"hip_bh+group", {{"engine", "hip_bh"}, {"force_precision", "fp32"}, {"bh_walk", "group"}}},
		        {"hip_bh+two_level", {{"engine", "hip_bh"}, {"force_precision", "fp32"}, {"bh_walk", "group"},
		                              {"bh_multipole", "quadrupole"}, {"bh_far_field", "local"}}}};
		values = {"8192", "16384", "32768", "65536", "131072", "262144"};
		field = "stars_count";
		max_time = 1;
	}
	else if(bench == "gpu_tree_ratio")
	{
		const std::vector<std::pair<std::string, std::string>> common = {{"stars_count", stars}, {"max_step", "0.01"}};
		auto with = [&](std::vector<std::pair<std::string, std::string>> k) {
			k.insert(k.end(), common.begin(), common.end());
			return k;
		};
		sets = {{"hip_bh+heap+nested", with({{"engine", "hip_bh"}, {"tree_layout", "heap"},
		                                     {"traverse_type", "nested_tree"}})},
		        {"hip_bh+heap_stackless", with({{"engine", "hip_bh"}, {"tree_layout", "heap_stackless"},
		                                        {"traverse_type", "nested_tree"}})},
		        {"hip_bh+group", with({{"engine", "hip_bh"}, {"force_precision", "fp32"}, {"bh_walk", "group"}})},
		        {"hip_bh+two_level", with({{"engine", "hip_bh"}, {"force_precision", "fp32"}, {"bh_walk", "group"},
		                                   {"bh_multipole", "quadrupole"}, {"bh_far_field", "local"}})}};
		values = {"0.1", "0.5", "1", "2", "4", "16", "64", "256", "1024"};
		double rc = 0;
		if(param.get_double("ratio_count", 0, &rc) && rc > 0 && rc < values.size())
		{
			values.resize(static_cast<size_t>(rc));
		}
		field = "distance_to_node_radius_ratio";
	}
	else if(bench == "gpu_tree_block")
	{
		sets = {{"hip", {{"engine", "hip"}, {"stars_count", stars}, {"max_step", "0.01"}}}};
		values = {"64", "128", "256", "512", "1024"};
		field = "block_size";
	}
	else
	{
		nbody_log("unknown bench '%s' (gpu, gpu_tree_ratio, gpu_tree_block)", bench.c_str());
		return -1;
	}
	if(param.contains("values"))
	{
		values = split(param.value("values", ""));
	}
	param.get_double("max_time", max_time, &max_time);

	std::printf("%%%% %s = {", field.c_str());
	for(size_t j = 0; j < values.size(); ++j)
	{
		std::printf("%s%s", j ? ", " : "", values[j].c_str());
	}
	std::printf("};\n%%%% Parameters:\n");
	for(size_t i = 0; i < sets.size(); ++i)
	{
		std::printf("%%%%\tset #%zu{\n", i);
		for(const auto& kv : sets[i].keys)
		{
			std::printf("%%%%\t\t%s=%s,\n", kv.first.c_str(), kv.second.c_str());
		}
		std::printf("%%%%\t}\n");
	}
	std::vector<std::vector<double>> result(sets.size(), std::vector<double>(values.size(), -1));
	int failures = 0;
	for(size_t i = 0; i < sets.size(); ++i)
	{
		for(size_t j = 0; j < values.size(); ++j)
		{
			result[i][j] = run_one(make(sets[i], field, values[j]), max_time);
			failures += result[i][j] < 0 ? 1 : 0;
			std::fprintf(stderr, "%s %s=%s: %.6g s/step\n", sets[i].name.c_str(), field.c_str(), values[j].c_str(),
			             result[i][j]);
		}
	}
	std::printf("time\n%s ", bench == "gpu" ? "engine" : "name");
	for(const auto& v : values)
	{
		std::printf("%s ", v.c_str());
	}
	std::printf("\n");
	for(size_t i = 0; i < sets.size(); ++i)
	{
		std::printf("%s ", sets[i].name.c_str());
		for(size_t j = 0; j < values.size(); ++j)
		{
			std::printf("%.8g ", result[i][j]);
		}
		std::printf("\n");
	}
	return failures == 0 ? 0 : -1;
}
