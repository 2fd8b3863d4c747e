// Line-parallel kernels for CDF3.7 bc 'n' (see line_parallel.h).
#include "line_parallel.h"

namespace whhd {
template bool lp_launch<W_Cdf37N, 0>(int, bool, bool, float*, const float*, const lp::LpGeom&,
                                   int, int,
                                   hipStream_t);
}  // namespace whhd
