/* Measurement helpers for the roofline of the gather-bound sweeps.
 *
 * No counterpart in the reference and not part of the drop-in ABI
 * (gt_hip.h): libgt_probe.so times the memory system's rate for the access
 * pattern that bounds the pull sweeps -- 4-byte gathers at random labels --
 * so a sweep's time can be compared with the time its own gathers would
 * take at that rate.  bench.py reports the comparison.
 */
#ifndef GT_PROBE_H
#define GT_PROBE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Times `reps` launches (after two untimed ones) of a kernel laid out like
 * the hot-slice sweep: one workgroup of eight wavefronts per compute unit,
 * each wavefront walking its share of `tiles` tiles of 1,024 4-byte labels
 * (16 per lane, loaded coalesced), then gathering src[label] for the first
 * `active` entries of each tile (the other lanes masked off) and summing.
 * Labels are uniform in [0, window) (a fixed hash of the entry's position),
 * src is a device array of `window` floats.  Writes the mean launch time in
 * microseconds and the device's compute-unit count.  Returns 0, or -1 with a
 * message in gt_probe_error(); allocates and frees its own buffers. */
int gt_probe_gather(uint32_t window, uint32_t tiles, uint32_t active, uint32_t reps,
                    double *us_per_launch, uint32_t *compute_units);

/* The last error message of this thread ("" when none). */
const char *gt_probe_error(void);

#ifdef __cplusplus
}
#endif

#endif /* GT_PROBE_H */
