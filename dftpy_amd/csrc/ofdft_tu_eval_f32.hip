// Instantiation unit: evaluation (E and V), FP32.
#include "ofdft_core.h"

OFDFT_TU_EVAL(, float)
