// nbh_tree.hip -- Barnes-Hut heap kd-tree (build, refit, traversal).
#include "nbh_common.h"

struct nbh_tree {
	nbh_ctx* ctx = nullptr;
	size_t   n = 0;
};

extern "C" {

nbh_status nbh_tree_create(nbh_ctx* ctx, size_t n, nbh_tree** tree)
{
	(void)ctx; (void)n; (void)tree;
	nbh::set_error("Barnes-Hut tree not available in this build");
	return NBH_ERR_UNSUPPORTED;
}

nbh_status nbh_tree_destroy(nbh_tree* tree)
{
	delete tree;
	return NBH_OK;
}

nbh_status nbh_tree_build(nbh_ctx*, nbh_tree*, const double*, const double*, double)
{
	nbh::set_error("Barnes-Hut tree not available in this build");
	return NBH_ERR_UNSUPPORTED;
}

nbh_status nbh_tree_refit(nbh_ctx*, nbh_tree*, const double*, double)
{
	nbh::set_error("Barnes-Hut tree not available in this build");
	return NBH_ERR_UNSUPPORTED;
}

nbh_status nbh_bh_fcompute(nbh_ctx*, const nbh_tree*, const double*, const double*, double*,
                           nbh_tree_layout, nbh_precision)
{
	nbh::set_error("Barnes-Hut tree not available in this build");
	return NBH_ERR_UNSUPPORTED;
}

nbh_status nbh_tree_get(nbh_ctx*, const nbh_tree*, int32_t*, double*, double*, double*, double*, double*)
{
	nbh::set_error("Barnes-Hut tree not available in this build");
	return NBH_ERR_UNSUPPORTED;
}

nbh_status nbh_bh_counters(nbh_ctx*, const nbh_tree*, uint64_t* a, uint64_t* b)
{
	if(a) { *a = 0; }
	if(b) { *b = 0; }
	return NBH_OK;
}

}  // extern "C"
