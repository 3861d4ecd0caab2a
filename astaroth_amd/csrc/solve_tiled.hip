// solve_tiled.hip -- LDS-tiled 2.5D solve kernel (placeholder until the tiled kernel lands:
// no candidate shapes, so the device layer uses the direct kernel).
#include "kernels.h"

namespace acamd {

int tiled_num_shapes() { return 0; }
TiledShape tiled_shape(int) { return TiledShape{0, 0, 0}; }
bool tiled_shape_fits(TiledShape, int, int, int) { return false; }

template <typename T>
bool launch_solve_tiled(const SolveArgs<T>&, bool, bool, int, TiledShape, hipStream_t)
{
    return false;
}

template bool launch_solve_tiled<AcRealT>(const SolveArgs<AcRealT>&, bool, bool, int, TiledShape,
                                          hipStream_t);

}  // namespace acamd
