// Instantiation unit: evaluation (E and V), FP64.
#include "ofdft_core.h"

OFDFT_TU_EVAL(, double)
