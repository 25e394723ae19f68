// Double-precision instantiation of the batched-ensemble pass kernels.
#include "sdns_batch.hpp"
SDNS_BATCH_INSTANTIATE(double)
