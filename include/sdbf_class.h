// sdbf_class.h -- the reference's header name (sdbf/sdbf_class.h) for callers that include it:
// the declarations are those of sdbf_amd.h.
#ifndef SDBFX_SDBF_CLASS_H
#define SDBFX_SDBF_CLASS_H
#include "sdbf_defines.h"
#endif
