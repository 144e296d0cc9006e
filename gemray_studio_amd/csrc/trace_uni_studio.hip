// trace_uni_studio.hip -- trace kernels for uniaxial stones under the studio rig.
#include "trace_launch.h"

GR_INSTANTIATE_TRACE_CLASSES(1, 0)
