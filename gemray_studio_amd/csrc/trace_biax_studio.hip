// trace_biax_studio.hip -- trace kernels for biaxial stones under the studio rig.
#include "trace_launch.h"

GR_INSTANTIATE_TRACE_CLASSES(17, 0)
