// adjoint.hip -- tangent-linear and adjoint kernels of one RK3 substep (placeholder).
#include "kernels.h"
