// trace_iso_studio.hip -- trace kernels for isotropic stones under the studio rig.
#include "trace_launch.h"

GR_INSTANTIATE_TRACE_CLASSES(0, 0)
