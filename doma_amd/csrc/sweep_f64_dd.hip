// sweep_f64_dd.hip -- instantiates the brick sweep kernels for double precision and
// the DD scheme (one translation unit per pair keeps compile times parallel).
#include "sweep_impl.h"

DOMA_SWEEP_INSTANTIATE(double, DOMA_SCHEME_DD)
