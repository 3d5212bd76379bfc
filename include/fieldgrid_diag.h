/* fieldgrid measurement diagnostics: entry points of libfieldgrid.so that calibrate the
 * measurements and are no part of the drop-in interface (include/fieldgrid.h). They have no
 * counterpart in the reference, and a caller of the FDTD path never needs them.
 */
#ifndef FIELDGRID_DIAG_H
#define FIELDGRID_DIAG_H

#include "fieldgrid.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Bandwidth probe: copies `bytes` between two device buffers `iterations` times and returns the
   achieved read+write GB/s. It calibrates the HBM roofline and the profiler's FETCH_SIZE /
   WRITE_SIZE counters against a known traffic. */
int fg_stream_copy(fg_device *dev, uint64_t bytes, uint32_t iterations, double *gbps);

#ifdef __cplusplus
}
#endif

#endif /* FIELDGRID_DIAG_H */
