// sweep_f32_tw.hip -- instantiates the brick sweep kernels for float precision and
// the TW_DD scheme (one translation unit per pair keeps compile times parallel).
#include "sweep_impl.h"

DOMA_SWEEP_INSTANTIATE(float, DOMA_SCHEME_TW_DD)
