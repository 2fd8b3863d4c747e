// Line-parallel kernels for CDF4.6 bc 'd' (see line_parallel.h).
#include "line_parallel.h"

namespace whhd {
template bool lp_launch<W_Cdf46D, 1>(int, bool, bool, float*, const float*, const lp::LpGeom&,
                                   int, int,
                                   hipStream_t);
}  // namespace whhd
