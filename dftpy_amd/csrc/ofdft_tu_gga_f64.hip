// Instantiation unit: GGA kinetic functionals, FP64.
#include "ofdft_core.h"

OFDFT_TU_GGA(, double)
