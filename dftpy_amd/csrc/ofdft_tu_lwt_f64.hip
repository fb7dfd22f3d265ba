// Instantiation unit: LWT / LMGP kinetic functionals, FP64.
#include "ofdft_core.h"

OFDFT_TU_LWT(, double)
