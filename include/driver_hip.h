/* The series driver: what the reference's find_spots does between reading
 * chunks and writing the table (src/find_spots.cc:460-700), on the device.
 *
 * Frames are submitted by number as a reader yields them, each as the chunk
 * HDF5 stored (or NULL: a frame the writer never wrote). They are held until
 * `chunk` of them wait, then taken in number order: decoded on the device,
 * thresholded there, and grouped in order with z = number + z_offset
 * (src/find_spots.cc:539-600). A frame whose z is not above the last grouped
 * one is dropped and counted out of order; a frame never written is counted
 * and leaves a gap that no spot crosses; a frame that fails to decode or to
 * threshold is logged ("frame N: message") and counted as a failure. The
 * table does not depend on `chunk`. sf_hip_driver_finish ends the series
 * and logs dials.find_spots' summary lines, worded as the reference words
 * them (src/find_spots.cc:642-682); sf_hip_driver_write writes the .refl
 * and logs its line (src/find_spots.cc:684-713).
 */
#ifndef SF_DRIVER_HIP_H
#define SF_DRIVER_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "decode_hip.h"
#include "dext_hip.h"
#include "refl_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* src/find_spots.cc:70-90 (Options): --chunk, -z, the grouping options. */
typedef struct sf_driver_options {
  int32_t chunk;        /* frames held before a batch; 0: 64 */
  int64_t z_offset;     /* z = frame number + z_offset */
  sf_options grouping;  /* min/max spot size, max separation, two_d */
} sf_driver_options;

/* What the summary counts (src/find_spots.cc:465-470, 541-548). */
typedef struct sf_driver_stats {
  uint64_t dispatched;   /* frames submitted */
  uint64_t read;         /* decoded and thresholded */
  uint64_t missing;      /* never written */
  uint64_t failed;       /* decode, threshold or grouping failures */
  uint64_t grouped;      /* frames given to the grouping */
  uint64_t out_of_order; /* dropped: not above the last grouped frame */
} sf_driver_stats;

typedef struct sf_driver sf_driver;

/* A driver for frames of height x width elements of bit_depth bits, on the
 * current device. options NULL: chunk 64, z_offset 0, the dials.find_spots
 * grouping defaults. NULL on error (sf_hip_driver_last_error()). */
sf_driver *sf_hip_driver_create(unsigned bit_depth, size_t height,
                                size_t width, const sf_driver_options *options);

/* Frame `number` as stored: `data` (size bytes, algorithm "bslz4", "lz4" or
 * "") is copied; data NULL means never written. Runs a batch once `chunk`
 * frames wait. SF_OK, or SF_BAD_ARGUMENT / SF_DEVICE_ERROR (frame failures
 * are logged and counted, not returned). */
int sf_hip_driver_submit(sf_driver *driver, int64_t number,
                         const uint8_t *data, size_t size,
                         const char *algorithm);

/* The detector mask (the bit array of sf_hip_stamp_mask, device memory the
 * caller keeps; NULL: none) stamped into every decoded frame before the
 * threshold, which the reference leaves out (README.md:227-234). 16- and
 * 32-bit frames. height x width is the mask's shape: SF_BAD_ARGUMENT unless
 * it is the driver's frame shape (nothing is launched then). The driver
 * keeps the pointer: the mask must stay allocated until
 * sf_hip_driver_finish() returns or another mask (or NULL) is set. */
int sf_hip_driver_set_mask(sf_driver *driver, const uint32_t *bits,
                           size_t height, size_t width);

/* Runs what waits, ends the grouping and logs the summary. */
int sf_hip_driver_finish(sf_driver *driver);

/* Writes the .refl (refl::write's options) and logs its line. SF_IO_ERROR
 * when the writer fails (the message is logged too). */
int sf_hip_driver_write(sf_driver *driver, const char *path,
                        const sf_refl_options *options);

/* The log so far: failures as they happened, then the summary. Copies at
 * most cap - 1 bytes and a terminator; returns the full length. */
size_t sf_hip_driver_log(sf_driver *driver, char *text, size_t cap);
int sf_hip_driver_stats(sf_driver *driver, sf_driver_stats *stats);
/* The driver's labeller: spots, pixels and counts after finish. */
sf_labeller *sf_hip_driver_labeller(sf_driver *driver);
void sf_hip_driver_destroy(sf_driver *driver);
const char *sf_hip_driver_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
