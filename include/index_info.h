// index_info.h -- the reference's header name (sdbf/index_info.h) for callers that include it:
// the declarations are those of sdbf_amd.h.
#ifndef SDBFX_INDEX_INFO_H
#define SDBFX_INDEX_INFO_H
#include "sdbf_defines.h"
#endif
