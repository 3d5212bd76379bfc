/* fieldgrid far field: the near-to-far-field transform of the NTFF accumulators on the device.
 *
 * Replaces `Far` (crates/sims-field/src/fdtd/ntff.rs:66-137): `Far::load` (76-88) takes one
 * frequency's J and M accumulators per surface patch, times a complex scale; `Far::field`
 * (90-115) sums them over every patch with the phase e^{i k d.pos} times the patch area for a
 * direction d; `intensity` (117-119) and `radiated` (121-136) follow. The sums over patches run
 * on the GPU in f64 (the reference's f64 arithmetic; the order of the sum differs, so results
 * agree to f64 rounding). A caller such as sims-antenna's solve.rs:416-440 (91 x 121
 * directions over ~10^5 patches) moves from O(dirs x patches) on one host core to one launch.
 */
#ifndef FIELDGRID_FAR_H
#define FIELDGRID_FAR_H

#include "fieldgrid.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fg_far fg_far;

/* Far::load (ntff.rs:76-88): `count` patches with centres pos[3*q..] (m) and areas area[q]
   (m^2), accumulators acc laid out as fg_record.ntff ([(q*nf + f)*12 + {J re/im xyz, M re/im
   xyz}]), frequency slot k of nf at freq Hz, each value times (scale_re + i scale_im). The
   arrays are copied. */
int fg_far_new(fg_device *dev, uint64_t count, const double *pos, const double *area, const float *acc, uint32_t nf,
               uint32_t k, double freq, double scale_re, double scale_im, fg_far **out);
/* Far::field (ntff.rs:90-115) for n unit directions dirs[3*d..]: out[6*d + 2*c + {0,1}] =
   {re, im} of component c. */
int fg_far_field(fg_far *far, uint64_t n, const double *dirs, double *out);
/* Far::intensity (ntff.rs:117-119): |field|^2 / (2 eta0) per direction. */
int fg_far_intensity(fg_far *far, uint64_t n, const double *dirs, double *out);
/* Far::radiated (ntff.rs:121-136): total radiated power and peak intensity over a steps x
   2 steps grid of directions (the intensities on the device, the sums in the reference's
   order on the host). */
int fg_far_radiated(fg_far *far, uint32_t steps, double *total, double *peak);
void fg_far_free(fg_far *far);

#ifdef __cplusplus
}
#endif

#endif /* FIELDGRID_FAR_H */
