// trace_uni_map.hip -- trace kernels for uniaxial stones under the HDR lat-long map.
#include "trace_launch.h"

GR_INSTANTIATE_TRACE_CLASSES(1, 4)
