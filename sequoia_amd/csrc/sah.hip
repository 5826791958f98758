// sah.hip — binned-SAH and grid builders (placeholder until the GPU SAH lands).
#include "internal.h"

namespace sq {

int sah_build(sq_ctx* c, const sq_triangle* tris, uint32_t n, const sq_build_params* p, sq_builder kind,
              sq_node* d_nodes, uint32_t* root, uint32_t* depth) {
    (void)c; (void)tris; (void)n; (void)p; (void)kind; (void)d_nodes; (void)root; (void)depth;
    SQ_RET(SQ_ERR_UNSUPPORTED, "SAH/grid builders are not available in this build");
}

int refit_tree(sq_ctx* c, sq_bvh* b) {
    (void)c; (void)b;
    SQ_RET(SQ_ERR_UNSUPPORTED, "refit is not available in this build");
}

}  // namespace sq
