// mlp.hip -- learned preconditioner kernels (filled in below)
#include <hip/hip_runtime.h>
