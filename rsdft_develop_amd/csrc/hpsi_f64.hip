// hpsi_f64.hip -- FP64 instantiation of the hpsi kernels, plus the launch-
// shape query and the autotuner (which time both precisions).
#define RSG_HPSI_TYPE double
#define RSG_HPSI_TYPE_IS_DOUBLE 1
#include "hpsi_impl.hpp"
