// Instantiation unit: LWT / LMGP kinetic functionals, FP32.
#include "ofdft_core.h"

OFDFT_TU_LWT(, float)
