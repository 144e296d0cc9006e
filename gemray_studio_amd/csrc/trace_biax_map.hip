// trace_biax_map.hip -- trace kernels for biaxial stones under the HDR lat-long map.
#include "trace_launch.h"

GR_INSTANTIATE_TRACE_CLASSES(17, 4)
