/* Writing a DIALS reflection table (.refl) from the spot and pixel tables the
 * labeller returns: the reference's refl::write (src/refl.hh:21-78), as a C
 * entry point. Host code only: it reads host tables and writes a file.
 *
 * The file is one msgpack document (src/refl.hh:3-15):
 *   [ "dials::af::reflection_table", 2,
 *     { "identifiers": { id: identifier } or {},
 *       "nrows": N,
 *       "data": { name: [ type, [ N, bin(raw little-endian column) ] ] } } ]
 * with the columns dials.find_spots writes, in name order (src/refl.hh:22-33):
 * bbox, flags, id, intensity.sum.value, intensity.sum.variance, n_signal,
 * panel, shoebox (optional), xyzobs.px.value, xyzobs.px.variance. Every
 * header takes msgpack's shortest form, so the bytes equal the reference's.
 */
#ifndef SF_REFL_HIP_H
#define SF_REFL_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "dext_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* src/refl.hh:52-63 (refl::Options). */
typedef struct sf_refl_options {
  uint64_t panel;         /* 0: one panel */
  int32_t id;             /* 0: the experiment every spot belongs to */
  int32_t shoeboxes;      /* 1: write the shoebox column (dense boxes) */
  const char *identifier; /* the .expt identifier; NULL or "" writes none */
} sf_refl_options;

/* src/refl.hh:65-74 (refl::write). `spots` and `pixels` as the labeller
 * returns them (spot s owns pixels[first .. first + n_signal - 1]); `width`
 * is the frame width the labeller was given. Returns SF_OK; SF_BAD_ARGUMENT
 * for a null path or table, a zero width, or a spot whose pixels lie outside
 * the pixel table or its box (where the reference would read past its
 * vectors); SF_IO_ERROR when the file cannot be written, or a column would
 * pass msgpack's 4 GB binary limit (where refl::write throws). The message is
 * in sf_hip_refl_last_error(). */
int sf_hip_refl_write(const char *path, const sf_spot *spots, size_t n_spots,
                      const sf_pixel *pixels, size_t n_pixels, size_t width,
                      const sf_refl_options *options);

/* The calling thread's last sf_hip_refl_write message. */
const char *sf_hip_refl_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SF_REFL_HIP_H */
