// Cell-table evaluation kernels of the r^3 kernel (every launch shape and field count);
// the kernels themselves are in eval_blk_impl.hpp.
#include "eval_blk_impl.hpp"

namespace lrb {
LRB_EVAL_BLK_INST(K_PHS3)
}  // namespace lrb
