// Cell-table evaluation kernels of the Wendland C2 kernel (every launch shape and field count);
// the kernels themselves are in eval_blk_impl.hpp.
#include "eval_blk_impl.hpp"

namespace lrb {
LRB_EVAL_BLK_INST(K_WC2)
}  // namespace lrb
