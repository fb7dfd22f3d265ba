// Instantiation unit: the persistent batched TN loop, FP64.
#include "tn_batch.h"

OFDFT_TU_BATCH(, double)
