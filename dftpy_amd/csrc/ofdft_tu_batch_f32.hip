// Instantiation unit: the persistent batched TN loop, FP32.
#include "tn_batch.h"

OFDFT_TU_BATCH(, float)
