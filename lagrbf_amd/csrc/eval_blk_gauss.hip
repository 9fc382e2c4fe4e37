// Cell-table evaluation kernels of the Gaussian, with the grid's shape parameter or each patch's (every launch shape and field count);
// the kernels themselves are in eval_blk_impl.hpp.
#include "eval_blk_impl.hpp"

namespace lrb {
LRB_EVAL_BLK_INST(K_GAUSS)
LRB_EVAL_BLK_INST(K_GAUSS_LOCAL)
}  // namespace lrb
