// bloom_filter.h -- the reference's header name (sdbf/bloom_filter.h) for callers that include it:
// the declarations are those of sdbf_amd.h.
#ifndef SDBFX_BLOOM_FILTER_H
#define SDBFX_BLOOM_FILTER_H
#include "sdbf_defines.h"
#endif
