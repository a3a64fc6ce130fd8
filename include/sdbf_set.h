// sdbf_set.h -- the reference's header name (sdbf/sdbf_set.h) for callers that include it:
// the declarations are those of sdbf_amd.h.
#ifndef SDBFX_SDBF_SET_H
#define SDBFX_SDBF_SET_H
#include "sdbf_defines.h"
#endif
