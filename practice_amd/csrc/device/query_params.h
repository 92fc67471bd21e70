// Device-side name of the kernel argument block (include/cq_query_params.h).
#include "cq_query_params.h"
