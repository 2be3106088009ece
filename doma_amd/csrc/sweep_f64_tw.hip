// sweep_f64_tw.hip -- instantiates the brick sweep kernels for double precision and
// the TW_DD scheme (one translation unit per pair keeps compile times parallel).
#include "sweep_impl.h"

DOMA_SWEEP_INSTANTIATE(double, DOMA_SCHEME_TW_DD)
