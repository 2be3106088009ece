// sweep_f32_dd.hip -- instantiates the brick sweep kernels for float precision and
// the DD scheme (one translation unit per pair keeps compile times parallel).
#include "sweep_impl.h"

DOMA_SWEEP_INSTANTIATE(float, DOMA_SCHEME_DD)
