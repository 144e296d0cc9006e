// trace_iso_map.hip -- trace kernels for isotropic stones under the HDR lat-long map.
#include "trace_launch.h"

GR_INSTANTIATE_TRACE_CLASSES(0, 4)
