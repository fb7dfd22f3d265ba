// Instantiation unit: Hessian products and the device TN loop, FP64.
#include "ofdft_core.h"

OFDFT_TU_HVP(, double)
