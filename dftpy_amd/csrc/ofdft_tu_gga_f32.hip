// Instantiation unit: GGA kinetic functionals, FP32.
#include "ofdft_core.h"

OFDFT_TU_GGA(, float)
