// MLP warm-start kernels (placeholder translation unit; filled in with the
// learned GJK guess network).
#include <hip/hip_runtime.h>
