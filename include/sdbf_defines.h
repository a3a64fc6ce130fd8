// sdbf_defines.h -- the reference's definitions header (sdbf/sdbf_defines.h,
// with sdbf/util.h's size units), for callers that include it: the class
// declarations of this engine (sdbf_amd.h) and the constants of the digest
// path with the reference's names and values.
#ifndef SDBFX_SDBF_DEFINES_H
#define SDBFX_SDBF_DEFINES_H

#include "sdbf_amd.h"

#define _MAX_ELEM_COUNT 160      // stream-mode elements per filter
#define _MAX_ELEM_COUNT_DD 192   // block-mode elements per filter
#define BF_SIZE 256
#define ENTR_POWER 10
#define ENTR_SCALE (1000 * (1 << ENTR_POWER))
#define MIN_FILE_SIZE 512
#define MIN_ELEM_COUNT 16
#define POP_WIN_SIZE 64
#define BIGFILTER 16384
#define BIGFILTER_ELEM 8738

#ifndef KB
#define KB 1024
#define MB (KB * KB)
#define GB (MB * KB)
#endif

#endif
