// hpsi_f32.hip -- FP32 instantiation of the hpsi kernels.
#define RSG_HPSI_TYPE float
#define RSG_HPSI_TYPE_IS_DOUBLE 0
#include "hpsi_impl.hpp"
