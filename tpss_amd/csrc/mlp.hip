// mlp.hip -- learned per-patch damping (placeholder translation unit).
#include <hip/hip_runtime.h>
