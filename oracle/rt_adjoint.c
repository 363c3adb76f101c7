/* Placeholder translation unit for the FP64 CPU adjoint (filled in later). */
#include "rt_oracle.h"
