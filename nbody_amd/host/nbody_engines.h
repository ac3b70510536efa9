// nbody_engines.h -- engine factory (counterpart of nbody/nbody_engines.cpp:4-193,
// with the two engines of this build registered as "hip" and "hip_bh").
//
// Keys (all optional):
//   engine                          "hip" | "hip_bh"
//   device                          one HIP device id, default "0" (a list such
//                                   as "0,0" is rejected: one GPU per engine)
//   block_size                      work-group size of the FP64 direct kernel
//                                   (default 64; any other positive value is
//                                   accepted and mapped to a valid size)
//   force_precision                 "fp64" (default) | "fp32" | "fp32_corr" (hip only)
//   distance_to_node_radius_ratio   lambda = 1/theta (hip_bh), default 10
//   tree_build_rate                 0 = rebuild every fcompute (default); k > 0
//                                   = rebuild when step % k == 0, refit otherwise
//   tree_layout                     "heap" (default) | "heap_stackless"
//   traverse_type                   "cycle" (default) | "nested_tree"
//   bh_walk                         "body" (default: one walk per body, the
//                                   reference traversal) | "group" (one walk per
//                                   64 bodies, fp32 only)
// Any unknown engine name or invalid value returns NULL.
#pragma once

#include "nbody_engine.h"
#include "nbody_params.h"

nbody_engine* nbody_create_engine(const nbody_params& param);
