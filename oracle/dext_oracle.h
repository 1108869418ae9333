/*
 * CPU oracle for the spot-finder hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's CPU path, used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
 * Nothing in the product library links or calls this.
 *
 *   oracle_dext_*          src/dext.cc:8-225 (summed-area-table form)
 *   oracle_dext_direct_*   the same windows summed directly, clipped at the
 *                          frame edge (the identity src/dext_cuda.cu:110-115
 *                          asserts between its direct and tile kernels)
 *   oracle_by_index        src/signal_order.cc:19-72 (order and duplicate
 *                          contract of src/signal_order.hh:51-62)
 *   oracle_labeller_*      src/dials_spots.cc:11-360, restated as a batch:
 *                          all frames are labelled at once and components are
 *                          emitted in the order the streaming Labeller closes
 *                          them, (last frame, first (frame, index) member).
 *
 * Pinning (what checks each function, beyond retyping its formula):
 *   dext          golden vectors written by the compiled reference
 *                 (tests/golden), the direct form (independent summation),
 *                 the published 12-frame fixture answer (8928 pixels, 48
 *                 spots; README.md:95-107) and invariants.
 *   by_index      equality with a stable sort (tests/test_signal_order.cc).
 *   labeller      the hand-derived known answers of tests/test_dials_spots.cc,
 *                 brute-force flood fill on tiny inputs, and golden spot
 *                 tables written by the compiled reference Labeller.
 */
#ifndef DEXT_ORACLE_H
#define DEXT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/signal_pixel.hh:20-28 -- 16 bytes. */
typedef struct {
  uint32_t index;
  uint32_t value;
  float background;
  uint16_t population;
  uint16_t reserved;
} oracle_signal_pixel;

/* src/dials_spots.hh:48-52 -- 12 bytes. */
typedef struct {
  int32_t frame;
  uint32_t index;
  uint32_t value;
} oracle_pixel;

/* src/dials_spots.hh:55-79 -- 104 bytes with the same field offsets. */
typedef struct {
  int32_t bbox[6];
  double position[3];
  double variance[3];
  double intensity;
  double intensity_variance;
  uint32_t n_signal;
  uint64_t first;
} oracle_spot;

/* src/dials_spots.hh:99-106 */
typedef struct {
  uint64_t signal_pixels, groups, too_small, too_large, separated, accepted;
} oracle_counts;

/* Threshold one frame. Writes at most `capacity` entries, ascending by index,
 * and stores the true count in *found. Returns 0, or -1 exactly where
 * dext() does (src/dext.cc:14-27). */
int oracle_dext_u16(const uint16_t *image, size_t height, size_t width,
                    oracle_signal_pixel *out, size_t capacity, size_t *found);
int oracle_dext_u32(const uint32_t *image, size_t height, size_t width,
                    oracle_signal_pixel *out, size_t capacity, size_t *found);
int oracle_dext_direct_u16(const uint16_t *image, size_t height, size_t width,
                           oracle_signal_pixel *out, size_t capacity,
                           size_t *found);
int oracle_dext_direct_u32(const uint32_t *image, size_t height, size_t width,
                           oracle_signal_pixel *out, size_t capacity,
                           size_t *found);

/* Threshold `frames` frames of one size on `threads` host threads, one frame
 * per thread at a time with thread-local scratch. Returns total signal pixels
 * or a negative error. Used for timing. */
long long oracle_dext_many_u16(const uint16_t *frames, size_t count,
                               size_t height, size_t width, int threads);

void oracle_by_index(oracle_signal_pixel *pixels, size_t n);

typedef struct oracle_labeller oracle_labeller;

oracle_labeller *oracle_labeller_new(size_t height, size_t width,
                                     size_t min_spot_size,
                                     size_t max_spot_size,
                                     double max_separation, int two_d);
void oracle_labeller_free(oracle_labeller *labeller);
/* 0 ok; -1 frame out of order or repeated; -2 indices not ascending;
 * -3 index off the frame (the three throws of src/dials_spots.cc:58-91). */
int oracle_labeller_add(oracle_labeller *labeller, int64_t frame,
                        const oracle_signal_pixel *signal, size_t n);
int oracle_labeller_finish(oracle_labeller *labeller);
size_t oracle_labeller_spot_count(const oracle_labeller *labeller);
size_t oracle_labeller_pixel_count(const oracle_labeller *labeller);
void oracle_labeller_results(const oracle_labeller *labeller,
                             oracle_spot *spots, oracle_pixel *pixels,
                             oracle_counts *counts);

/* The reference's shared test-frame generator (tests/synthetic_frame.hh,
 * make_frame): fills height * width pixels. */
void oracle_synthetic_frame(uint16_t *pixels, size_t height, size_t width,
                            uint64_t seed, double background, int with_mask);

#ifdef __cplusplus
}
#endif

#endif
