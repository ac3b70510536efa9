// nbody_engines.cpp -- see nbody_engines.h.
#include "nbody_engines.h"

#include <cstdlib>

#include "nbody_engine_hip.h"
#include "nbody_log.h"

namespace {

bool parse_device(const std::string& s, int* device)
{
	if(s.empty() || s.find_first_not_of("0123456789") != std::string::npos)
	{
		return false;
	}
	*device = std::atoi(s.c_str());
	return true;
}

bool common_hip_params(const nbody_params& param, nbody_engine_hip* engine)
{
	int device = 0;
	std::string dev = param.value("device", "0");
	if(!parse_device(dev, &device))
	{
		nbody_log("Invalid device '%s': one device id is expected", dev.c_str());
		return false;
	}
	long long block_size = 0;
	if(!param.get_long("block_size", NBODY_DATA_BLOCK_SIZE, &block_size) || block_size <= 0)
	{
		nbody_log("Invalid block_size");
		return false;
	}
	std::string prec = param.value("force_precision", "fp64");
	if(prec != "fp64" && prec != "fp32" && prec != "fp32_corr")
	{
		nbody_log("Invalid force_precision. Allowed values are 'fp64', 'fp32' or 'fp32_corr'");
		return false;
	}
	if(prec == "fp32_corr" && param.value("engine", "") != "hip")
	{
		nbody_log("force_precision=fp32_corr is a direct-sum mode (engine hip)");
		return false;
	}
	const std::string type = param.value("engine", "");
	if(prec != "fp64" && (type == "hip_block" || type == "hip_ah"))
	{
		nbody_log("engine %s computes in FP64 only", type.c_str());
		return false;
	}
	if(engine->select_device(device) != 0)
	{
		return false;
	}
	engine->set_block_size(static_cast<int>(block_size));
	if(type != "hip_block")
	{
		engine->set_precision(prec == "fp64" ? NBH_FP64 : prec == "fp32" ? NBH_FP32 : NBH_FP32_CORR);
	}
	return true;
}

}  // namespace

nbody_engine* nbody_create_engine(const nbody_params& param)
{
	const std::string type(param.value("engine", ""));
	if(type == "hip")
	{
		nbody_engine_hip* engine = new nbody_engine_hip();
		if(!common_hip_params(param, engine))
		{
			delete engine;
			return nullptr;
		}
		return engine;
	}
	// nbody/nbody_engines.cpp:16-19 (`block`) on the GPU
	if(type == "hip_block")
	{
		nbody_engine_hip_block* engine = new nbody_engine_hip_block();
		if(!common_hip_params(param, engine))
		{
			delete engine;
			return nullptr;
		}
		return engine;
	}
	// nbody/nbody_engines.cpp:8-15 (`ah`) on the GPU; the same keys and defaults.
	// full_recompute_rate 0 (a division by zero in the reference) returns NULL.
	if(type == "hip_ah")
	{
		long long rate = 0;
		double max_dist = 0, min_force = 0;
		if(!param.get_long("full_recompute_rate", 1000, &rate) || rate <= 0)
		{
			nbody_log("Invalid full_recompute_rate: must be a positive integer");
			return nullptr;
		}
		if(!param.get_double("max_dist", 10, &max_dist) || !(max_dist >= 0))
		{
			nbody_log("Invalid max_dist");
			return nullptr;
		}
		if(!param.get_double("min_force", 1e-4, &min_force) || !(min_force >= 0))
		{
			nbody_log("Invalid min_force");
			return nullptr;
		}
		nbody_engine_hip_ah* engine = new nbody_engine_hip_ah(static_cast<size_t>(rate), max_dist, min_force);
		if(!common_hip_params(param, engine))
		{
			delete engine;
			return nullptr;
		}
		return engine;
	}
	if(type == "hip_bh")
	{
		double ratio = 0;
		long long rate = 0;
		if(!param.get_double("distance_to_node_radius_ratio", 10, &ratio) || !(ratio > 0))
		{
			nbody_log("Invalid distance_to_node_radius_ratio");
			return nullptr;
		}
		if(!param.get_long("tree_build_rate", 0, &rate) || rate < 0)
		{
			nbody_log("Invalid tree_build_rate");
			return nullptr;
		}
		e_tree_layout tl = tree_layout_from_str(param.value("tree_layout", "heap"));
		if(tl != etl_heap && tl != etl_heap_stackless)
		{
			nbody_log("Invalid tree_layout. Allowed values are 'heap' or 'heap_stackless'");
			return nullptr;
		}
		std::string strtt = param.value("traverse_type", "cycle");
		e_traverse_type tt;
		if(strtt == "cycle")
		{
			tt = ett_cycle;
		}
		else if(strtt == "nested_tree")
		{
			tt = ett_nested_tree;
		}
		else
		{
			nbody_log("Invalid traverse_type. Allowed values are 'cycle' or 'nested_tree'");
			return nullptr;
		}
		std::string walk = param.value("bh_walk", "body");
		if(walk != "body" && walk != "group")
		{
			nbody_log("Invalid bh_walk. Allowed values are 'body' or 'group'");
			return nullptr;
		}
		if(walk == "group" && param.value("force_precision", "fp64") != "fp32")
		{
			nbody_log("bh_walk=group needs force_precision=fp32");
			return nullptr;
		}
		std::string multipole = param.value("bh_multipole", "monopole");
		if(multipole != "monopole" && multipole != "quadrupole")
		{
			nbody_log("Invalid bh_multipole. Allowed values are 'monopole' or 'quadrupole'");
			return nullptr;
		}
		if(multipole == "quadrupole" && walk != "group")
		{
			nbody_log("bh_multipole=quadrupole needs bh_walk=group");
			return nullptr;
		}
		double quad_near = 3.0;
		if(!param.get_double("bh_quad_near", 3.0, &quad_near) || !(quad_near >= 0.0))
		{
			nbody_log("Invalid bh_quad_near");
			return nullptr;
		}
		std::string far_field = param.value("bh_far_field", "per_body");
		if(far_field != "per_body" && far_field != "local")
		{
			nbody_log("Invalid bh_far_field. Allowed values are 'per_body' or 'local'");
			return nullptr;
		}
		if(far_field == "local" && (multipole != "quadrupole" || !(quad_near >= 1.0)))
		{
			nbody_log("bh_far_field=local needs bh_multipole=quadrupole and bh_quad_near >= 1");
			return nullptr;
		}
		// learned correction of the group walk (no reference counterpart)
		std::string net = param.value("learned_correction", "none");
		double rho = -1.0;
		if(!param.get_double("correction_rho", -1.0, &rho) || !(rho == -1.0 || rho >= 0.0))
		{
			nbody_log("Invalid correction_rho");
			return nullptr;
		}
		double near_k = -1.0;
		if(!param.get_double("correction_near", -1.0, &near_k) || !(near_k == -1.0 || near_k >= 0.0))
		{
			nbody_log("Invalid correction_near");
			return nullptr;
		}
		if(net != "none" && walk != "group")
		{
			nbody_log("learned_correction needs bh_walk=group");
			return nullptr;
		}
		// the tree the group walks run on (no reference counterpart): the heap
		// kd-tree of the reference, or a Morton radix tree
		const std::string tree = param.value("bh_tree", "heap");
		if(tree != "heap" && tree != "morton")
		{
			nbody_log("Invalid bh_tree. Allowed values are 'heap' or 'morton'");
			return nullptr;
		}
		if(tree == "morton" && (walk != "group" || net != "none"))
		{
			nbody_log("bh_tree=morton needs bh_walk=group and no learned_correction");
			return nullptr;
		}
		nbody_engine_hip_bh* engine = new nbody_engine_hip_bh(ratio, tt, static_cast<size_t>(rate), tl);
		engine->set_group_walk(walk == "group");
		engine->set_quadrupole(multipole == "quadrupole");
		engine->set_quad_near(static_cast<float>(quad_near));
		engine->set_local_far_field(far_field == "local");
		engine->set_morton_tree(tree == "morton");
		if(net != "none" && !engine->load_correction(net, rho, near_k))
		{
			delete engine;
			return nullptr;
		}
		if(!common_hip_params(param, engine))
		{
			delete engine;
			return nullptr;
		}
		return engine;
	}
	return nullptr;
}
