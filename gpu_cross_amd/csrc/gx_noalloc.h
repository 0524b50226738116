// gx_noalloc.h -- the scratch-array FP64 path behind gx_execute_noalloc (gx_noalloc.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "gx_common.h"
#include "gx_internal.h"

namespace gx {

// Look up the device address of the library's scratch array on `device` (which must
// be current) once; later calls return at once.  Called by gx_create.
int noalloc_setup(int device);

// Enqueue line prep, block ranges and FP64 accumulation of one point into intens with
// every intermediate buffer in the scratch array of `device`.  Calls no allocator.
int execute_noalloc(int device, const PointConsts& pc, const GridView& gv, const LineStore& ls, int nlines,
                    double* intens, hipStream_t s);

}  // namespace gx
