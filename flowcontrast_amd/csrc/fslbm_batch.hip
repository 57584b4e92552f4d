// Batched multi-domain step (placeholder until written).
#include "fslbm_internal.h"
