// Cell-table evaluation kernels of the linear, Wendland C0, bump, multiquadric and inverse multiquadric kernels (every launch shape and field count);
// the kernels themselves are in eval_blk_impl.hpp.
#include "eval_blk_impl.hpp"

namespace lrb {
LRB_EVAL_BLK_INST(K_LINEAR)
LRB_EVAL_BLK_INST(K_WC0)
LRB_EVAL_BLK_INST(K_BUMP)
LRB_EVAL_BLK_INST(K_MQ)
LRB_EVAL_BLK_INST(K_IMQ)
}  // namespace lrb
