/*
 * dext_hip -- the C ABI of the MI355X (gfx950) spot-finder hot path.
 *
 * Plain C: pointers, sizes and status codes; no exception crosses it and no
 * HIP or torch type appears in a signature. Each entry point names the
 * reference interface it stands in for (paths relative to the reference
 * repository graeme-winter/dials-metal-spotfinder).
 *
 * Status codes mirror the reference's return codes and add three:
 *   SF_OK (0), SF_BAD_ARGUMENT (-1, dext_gpu.hh:141 / dext_cuda.cu:588-593),
 *   SF_OVERFLOW (-2, more signal pixels than the list holds,
 *   dext_gpu.hh:148-151), SF_DEVICE_ERROR (-3, where gpu::find throws
 *   std::runtime_error, dext_gpu.hh:153-155), SF_ORDER_ERROR (-4, where
 *   dials_spots::Labeller::add throws, dials_spots.cc:58-91), SF_IO_ERROR
 *   (-5, where refl::write throws, refl.cc:36-80 and 160-164).
 * sf_hip_last_error() returns the calling thread's last message.
 */
#ifndef DEXT_HIP_H
#define DEXT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SF_OK 0
#define SF_BAD_ARGUMENT (-1)
#define SF_OVERFLOW (-2)
#define SF_DEVICE_ERROR (-3)
#define SF_ORDER_ERROR (-4)
#define SF_IO_ERROR (-5)
#define SF_BAD_DATA (-6) /* a chunk that does not decode to its frame */

/* src/signal_pixel.hh:20-28: 16 bytes, identical layout to SignalPixel. */
typedef struct sf_signal_pixel {
  uint32_t index;
  uint32_t value;
  float background;
  uint16_t population;
  uint16_t reserved;
} sf_signal_pixel;

/* src/dials_spots.hh:48-52: 12 bytes, identical layout to Pixel. */
typedef struct sf_pixel {
  int32_t frame;
  uint32_t index;
  uint32_t value;
} sf_pixel;

/* src/dials_spots.hh:55-79: 104 bytes, identical layout to Spot. */
typedef struct sf_spot {
  int32_t bbox[6];
  double position[3];
  double variance[3];
  double intensity;
  double intensity_variance;
  uint32_t n_signal;
  uint64_t first;
} sf_spot;

/* src/dials_spots.hh:99-106 */
typedef struct sf_counts {
  uint64_t signal_pixels, groups, too_small, too_large, separated, accepted;
} sf_counts;

/* src/dials_spots.hh:81-97 */
typedef struct sf_options {
  uint64_t min_spot_size;  /* 3 */
  uint64_t max_spot_size;  /* 1000 */
  double max_separation;   /* 2.0; <= 0 turns the filter off */
  int32_t two_d;           /* 0 */
} sf_options;

/* ---- device and host memory (dext_gpu.hh:125-139) ---------------------- */

/* dext_gpu.hh:28 backend(): "HIP". */
const char *sf_hip_backend(void);
/* dext_gpu.hh:126 available(). */
int sf_hip_available(void);
/* dext_gpu.hh:131 memory_free(). */
size_t sf_hip_memory_free(void);
/* dext_gpu.hh:138-139 host_alloc / host_free: page-locked (hipHostMalloc). */
void *sf_hip_host_alloc(size_t bytes);
void sf_hip_host_free(void *pointer);
/* The calling thread's last error message ("" if none). */
const char *sf_hip_last_error(void);

/* Device buffers and copies, for callers that keep frames resident in HBM
 * (no reference counterpart: the reference always uploads per frame). */
void *sf_hip_device_alloc(size_t bytes);
void sf_hip_device_free(void *pointer);
int sf_hip_copy_to_device(void *device, const void *host, size_t bytes);
int sf_hip_copy_to_host(void *host, const void *device, size_t bytes);
int sf_hip_synchronize(void);

/* ---- window choice and profiling (dext_gpu.hh:49-123) -------------------
 * Window 0 = Direct, 1 = Tile. On this backend the pair selects the form:
 * (Tile, Tile) runs the fused single-pass kernel; any Direct runs the plain
 * three-kernel form whose stages sum their windows directly. Defaults are
 * Tile/Tile, overridable by SPOTFINDER_GPU_STAGE0 / SPOTFINDER_GPU_STAGE2. */
int sf_hip_stage0_window(void);
void sf_hip_set_stage0_window(int window);
int sf_hip_stage2_window(void);
void sf_hip_set_stage2_window(int window);
/* dext_gpu.hh:112-116: upload, stage0, stage1, stage2, copy, sort (ms).
 * For the fused form stage0 holds the fused kernel, stage1 zero and stage2
 * the ordering kernel. */
void sf_hip_profile_stages(int on);
int sf_hip_profiling(void);
void sf_hip_last_stage_times(double times[6]);
/* dext_gpu.hh:123 report_windows(). */
void sf_hip_report_windows(int on);

/* ---- the drop-in threshold (dext_gpu.hh:156-158 gpu::find<T>) -----------
 * Host frame in, ascending list out, synchronous. Writes at most `capacity`
 * entries to `out`, stores the number found in *found. Returns SF_OK,
 * SF_BAD_ARGUMENT, SF_OVERFLOW when found exceeds height*width*sizeof(T)/16
 * (the device list's size, dext_cuda.cu:614-616), or SF_DEVICE_ERROR.
 * Uses a per-thread stream and workspace (dext_cuda.cu:468-542). */
int sf_hip_find_u16(const uint16_t *image, size_t height, size_t width,
                    sf_signal_pixel *out, size_t capacity, size_t *found);
int sf_hip_find_u32(const uint32_t *image, size_t height, size_t width,
                    sf_signal_pixel *out, size_t capacity, size_t *found);
/* When the last find() on this thread found more than `capacity`, copy its
 * first n entries (n <= found) into out. */
int sf_hip_take_found(sf_signal_pixel *out, size_t n);

/* ---- the threshold with no allocation --------------------------------------
 * No counterpart in the reference: this is for callers that cannot allocate
 * device memory -- code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc. It is sf_hip_find_u16 / sf_hip_find_u32
 * (same kernels, same list, same count) on a frame already in device memory,
 * with every buffer beyond its inputs and outputs in one static device
 * array of SF_FIND_STATIC_SCRATCH_BYTES per device, held until the process
 * ends. It calls no hipMalloc, hipFree or other allocating routine.
 *   pixel_bytes 2 or 4; frame: height x width device pixels; out: device
 *   list of `capacity` entries; found: a device uint32 that receives the
 *   number of signal pixels found. Asynchronous on `stream` (a hipStream_t;
 *   NULL: the null stream). The list is written, ascending by index, when
 *   the count is at most min(capacity, height * width * pixel_bytes / 16);
 *   a larger count is the SF_OVERFLOW of find() and leaves the list
 *   unwritten.
 * Returns SF_OK once enqueued, SF_BAD_ARGUMENT (message in
 * sf_hip_last_error) for a null pointer, a zero size or capacity, another
 * pixel type, more than INT32_MAX pixels, or a frame whose tables exceed the
 * array (a 4362 x 4148 frame fits at either pixel type), SF_DEVICE_ERROR
 * when a launch fails. The first call on a device looks the array up and
 * uploads two small tables (synchronous work: make it outside a capture).
 * Calls on one device share its array: calls that overlap there (other
 * streams, other threads) race, and must be serialised by the caller. */
#define SF_FIND_STATIC_SCRATCH_BYTES ((size_t)96 << 20)
int sf_hip_find_static(int pixel_bytes, const void *frame, size_t height,
                       size_t width, sf_signal_pixel *out, size_t capacity,
                       uint32_t *found, void *stream);

/* ---- device-resident batches (no reference counterpart: the reference
 * launches one frame at a time, dext_cuda.cu:603-656) ----------------------
 * A plan fixes a list of frame shapes of one pixel type (2 or 4 bytes) and
 * owns the staging, cell tables and ordered output lists for them. run()
 * thresholds every frame of the plan in one sequence of launches (threshold,
 * then ordering), asynchronously on the plan's stream.
 * form 0 = fused (default: the sparse passes -- a probe per 5x5 cell, then
 * erosion, Poisson test and ordering only around active cells -- for 16-bit
 * frames and for 32-bit frames below 2^29 pixels; the register-sweep kernel
 * for larger 32-bit ones), 1 = plain three-kernel form, one frame per launch
 * sequence, 2 = the register-sweep kernel. Any other form is refused with
 * SF_BAD_ARGUMENT. */
typedef struct sf_plan sf_plan;
sf_plan *sf_hip_plan_create(int pixel_bytes, int n_frames,
                            const uint32_t *heights, const uint32_t *widths,
                            int band_rows);
void sf_hip_plan_destroy(sf_plan *plan);
int sf_hip_plan_set_form(sf_plan *plan, int form);
/* Persistent grid: 0 = one block per work item, >0 = that many blocks
 * striding over the items. */
int sf_hip_plan_set_grid(sf_plan *plan, int blocks);
/* The autotuner: times the candidate launch shapes of the plan's kernel on
 * these frames (each `trials` times, after a warm-up run) and keeps the
 * fastest. For the sweep kernel the candidates are band heights (the tile's
 * rows: 44, 108, 236, 492) crossed with grids (0 and 4 / 8 / 16 resident
 * blocks a CU); the sparse passes' launch shape follows the frames, so they
 * have nothing to time (band_rows reports their 80-row scout band). Reports
 * the chosen band rows and grid through band_rows / blocks (either may be
 * NULL). Drops a captured graph: capture again. */
int sf_hip_plan_autotune(sf_plan *plan, const void *const *device_frames,
                         int trials, int *band_rows, int *blocks);
int sf_hip_plan_run(sf_plan *plan, const void *const *device_frames);
/* Point the plan at these device frames without launching (the descriptor
 * upload that run() does first). A captured graph reads the descriptors at
 * execution, so bind() then replay() thresholds new frames. */
int sf_hip_plan_bind(sf_plan *plan, const void *const *device_frames);
/* Capture run()'s launches for the bound frames into a hipGraph; replay()
 * launches it. The graph clears its counters with a kernel, not a memset. */
int sf_hip_plan_capture(sf_plan *plan, const void *const *device_frames);
int sf_hip_plan_replay(sf_plan *plan);
int sf_hip_plan_wait(sf_plan *plan);
/* After wait(): per-frame counts (may exceed capacity -> that frame is -2). */
int sf_hip_plan_counts(sf_plan *plan, uint32_t *counts);
/* Device pointer to frame f's ordered list, and its capacity. */
const sf_signal_pixel *sf_hip_plan_output(sf_plan *plan, int frame);
uint32_t sf_hip_plan_capacity(sf_plan *plan, int frame);
/* Kernel time of the last run (ms): fused/plain kernels, ordering. Needs
 * sf_hip_plan_set_timing(plan, 1) before the run. */
int sf_hip_plan_set_timing(sf_plan *plan, int on);
int sf_hip_plan_last_times(sf_plan *plan, double *threshold_ms,
                           double *order_ms);
/* The last run's threshold split by pass (ms): the first kernel (the sparse
 * form's scout, or the single threshold kernel of the other fused forms),
 * the threshold's remaining kernels (the sparse form's erosion and Poisson
 * passes), and the ordering. Any output may be NULL. */
int sf_hip_plan_last_pass_times(sf_plan *plan, double *first_ms,
                                double *rest_ms, double *order_ms);
/* The plan's stream, as an opaque handle (hipStream_t). */
void *sf_hip_plan_stream(sf_plan *plan);

/* Pipelined steps over a series of chunks (16-bit plans of form 0; no
 * reference counterpart). The sparse threshold's first pass (the scout)
 * streams every pixel from HBM; its later passes (erosion, Poisson test)
 * read little and keep the vector units busy. A step finishes the chunk
 * bound to `current` -- scouted by the step before -- and scouts the chunk
 * bound to `next`, in two launches that each mix half of next's scout with
 * one of current's passes, then orders current's lists. Afterwards
 * current's counts and lists are those of run(), and next awaits the step
 * that finishes it. Either plan may be NULL: (NULL, first) only scouts,
 * (last, NULL) only finishes. SF_BAD_ARGUMENT if current was not scouted by
 * a step since its frames were bound, if next has no frames bound, or for a
 * plan of another pixel type or form. The step runs on current's stream
 * (next's when current is NULL), after next's stream has finished its work;
 * next's stream then waits for the step. */
int sf_hip_plan_pipeline(sf_plan *current, sf_plan *next);
/* The same step captured into a hipGraph that `current` (or `next` when
 * current is NULL) owns; replay_pipeline launches it for the same pair,
 * reading whatever frames are bound to the two plans at that time, with the
 * same checks as sf_hip_plan_pipeline (a capture runs nothing and needs no
 * frames). Capture again after either plan changes form. */
int sf_hip_plan_capture_pipeline(sf_plan *current, sf_plan *next);
int sf_hip_plan_replay_pipeline(sf_plan *current, sf_plan *next);
/* Deferred ordering (no reference counterpart). With `deferred` set, a
 * pipelined step that finishes this plan (as `current`, eager or captured
 * after the setting) stops after its Poisson test and leaves the ordering of
 * its lists to sf_hip_plan_order, which enqueues it on the plan's stream.
 * The next step, on the next plan's stream, does not wait for it: with three
 * plans in turn, a chunk's ordering runs beside the next chunk's step. Until
 * sf_hip_plan_order the plan's lists and counts are not yet those of the
 * step, and the plan cannot be the `next` of a step (SF_BAD_ARGUMENT);
 * sf_hip_plan_order with no ordering pending, or a setting change while one
 * is, is SF_BAD_ARGUMENT. A run or replay orders its own lists.
 * sf_hip_labeller_add_plan takes a plan whose ordering is pending: it
 * enqueues the ordering's places on its own stream and takes the survivors
 * in list order by them; the lists are then written only by a later
 * sf_hip_plan_order (before the plan is scouted again), and the next step
 * that scouts the plan waits for that reader. */
int sf_hip_plan_set_deferred_order(sf_plan *plan, int deferred);
int sf_hip_plan_order(sf_plan *plan);

/* ---- 3-D grouping and centroids on the device
 * (dials_spots::Labeller, src/dials_spots.hh:108-158) -----------------------
 * Same options, same output: the Spot and Pixel tables and Counts come out
 * identical, row for row, to the reference Labeller's after finish(). Frames
 * are labelled in chunks on the device; a component is emitted once a frame
 * arrives that it has no pixel on, ordered as the reference orders them. */
typedef struct sf_labeller sf_labeller;
/* Labeller(height, width, options); NULL where the constructor throws
 * (src/dials_spots.cc:22-30). options NULL = the dials.find_spots defaults. */
sf_labeller *sf_hip_labeller_create(size_t height, size_t width,
                                    const sf_options *options);
void sf_hip_labeller_destroy(sf_labeller *labeller);
/* Labeller::add (src/dials_spots.cc:57-145) with a host list. SF_ORDER_ERROR
 * where add() throws: a frame not above the last, indices not strictly
 * ascending, an index off the frame (the frame is then not added). */
int sf_hip_labeller_add(sf_labeller *labeller, int64_t frame,
                        const sf_signal_pixel *signal, size_t n);
/* Add every frame of a plan's last run, as frames first_frame,
 * first_frame + 1, ... straight from the device lists (no host copy). A
 * frame whose list overflowed (-2) is added empty. Returns without waiting
 * for the plan: the survivors are gathered into the labeller's node arrays
 * on a stream of its own, after the plan's work enqueued so far (the plan's
 * next write of its lists waits for the gather, so the plan may be rebound
 * and run again at once), and a worker thread labels them after the gather.
 * A plan whose ordering is pending is taken by the ordering's places (see
 * sf_hip_plan_set_deferred_order). A failure of that labelling is returned
 * by the next call on the labeller (add, add_plan, flush, finish, wait or a
 * query). Holds two chunks' nodes in device memory for the gathers. */
int sf_hip_labeller_add_plan(sf_labeller *labeller, sf_plan *plan,
                             int64_t first_frame);
/* Add n frames whose ordered lists are already in device memory, as frame
 * numbers frames[0] < frames[1] < ... (above the last added; gaps close
 * components as src/dials_spots.cc:68-69 does). For frame k: lists[k] and
 * count_ptrs[k] are device pointers, capacities[k] its list's entries,
 * counts[k] the count as the host knows it (a count above the capacity adds
 * the frame empty). The lists are read on `stream` (NULL: the labeller's
 * own), so the caller may reuse them once later work on that stream runs.
 * No counterpart in the reference, whose Labeller takes host vectors: for
 * callers whose lists are on the device already (the serving loop). */
int sf_hip_labeller_add_lists(sf_labeller *labeller, int n,
                              const int64_t *frames,
                              const sf_signal_pixel *const *lists,
                              const uint32_t *const *count_ptrs,
                              const uint32_t *capacities,
                              const uint32_t *counts, void *stream);
/* Label what has been added so far (queries do this themselves). */
int sf_hip_labeller_flush(sf_labeller *labeller);
/* Labeller::finish (src/dials_spots.cc:147-151). */
int sf_hip_labeller_finish(sf_labeller *labeller);
int sf_hip_labeller_wait(sf_labeller *labeller);
/* Labeller::live(): pixels held for components that may still grow. */
size_t sf_hip_labeller_live(sf_labeller *labeller);
/* Sizes of spots() and pixels() so far. */
int sf_hip_labeller_sizes(sf_labeller *labeller, size_t *spots,
                          size_t *pixels);
size_t sf_hip_labeller_spot_count(sf_labeller *labeller);
size_t sf_hip_labeller_pixel_count(sf_labeller *labeller);
/* Room for `spots` more spot rows and `pixels` more pixel rows in the
 * device result tables, allocated now (after waiting for what was added), so
 * that the next chunks allocate nothing. No counterpart in the reference,
 * whose tables are host vectors: for callers that know their series'
 * length, since a large device allocation can stall for hundreds of
 * milliseconds or more while a series runs. */
int sf_hip_labeller_reserve(sf_labeller *labeller, size_t spots,
                            size_t pixels);
/* spots(), pixels() and counts() (any pointer may be NULL). */
int sf_hip_labeller_results(sf_labeller *labeller, sf_spot *spots,
                            sf_pixel *pixels, sf_counts *counts);
const char *sf_hip_labeller_last_error(void);

/* ---- the serving loop ----------------------------------------------------
 * No counterpart in the reference, whose driver runs one series at a time
 * (src/find_spots.cc:560-640): this takes device-resident frames of several
 * concurrent series as they arrive, batches whatever is admitted into shared
 * launches (continuous batching over a few streams), keeps each series'
 * union-find frontier in its own device labeller, and admits frames by an
 * HBM budget for their per-frame buffers (slots). A series' lists and
 * tables are the same alone and in any mix. Calls are thread-safe. */
typedef struct sf_server sf_server;
typedef struct sf_server_stats {
  uint64_t batches;               /* launches */
  uint64_t batched_frames;        /* frames over all launches */
  uint64_t admission_waits;       /* times a frame found no slot in budget */
  uint64_t max_batches_in_flight;
  uint64_t bytes_in_use;          /* slot bytes allocated */
  uint64_t peak_bytes;
  uint64_t budget;
  /* frames with more signal pixels than their list holds (the -2 of
   * dext_gpu.hh:148-151): their count is reported as found (above the
   * slot's capacity) and the labeller takes them as empty frames, as the
   * driver does not group a frame that failed */
  uint64_t overflowed;
  /* bytes of host-fed frames (sf_hip_server_open_host) copied to the device */
  uint64_t ingested_bytes;
} sf_server_stats;
/* hbm_budget: bytes of slots the server may hold; max_batch_frames: frames
 * per launch; keep_lists: keep every frame's ordered list on the host (for
 * checking; costs a copy per frame). */
sf_server *sf_hip_server_create(size_t hbm_budget, int max_batch_frames,
                                int keep_lists);
/* Drains, then frees everything (labellers included). */
void sf_hip_server_destroy(sf_server *server);
/* Opens a series; returns its id (>= 0) or a negative status. The labeller
 * options are as sf_hip_labeller_create's. */
int sf_hip_server_open(sf_server *server, int pixel_bytes, size_t height,
                       size_t width, const sf_options *options);
/* Opens a series whose frames arrive in host memory: sf_hip_server_submit
 * then takes a host pointer (pinned -- sf_hip_host_alloc -- for the copy to
 * overlap; it must stay unchanged until the frame completes). Each frame is
 * copied into its slot (which holds a frame more) on the serving loop's
 * batch stream just before its batch's kernels, so its latencies include
 * the copy. No counterpart in the reference, whose driver reads frames
 * into host memory and copies each in on its worker's stream
 * (src/find_spots.cc:205-221, src/dext_cuda.cu:633). */
int sf_hip_server_open_host(sf_server *server, int pixel_bytes,
                            size_t height, size_t width,
                            const sf_options *options);
/* Hands over one frame (device memory, or pinned host memory for a series
 * opened with sf_hip_server_open_host; valid until the frame completes);
 * frame numbers ascend per series. Returns at once. _at stamps the arrival
 * at a time of sf_hip_server_clock() (seconds) instead of now. */
int sf_hip_server_submit(sf_server *server, int series, int64_t frame,
                         const void *device_pixels);
int sf_hip_server_submit_at(sf_server *server, int series, int64_t frame,
                            const void *device_pixels, double arrival_s);
double sf_hip_server_clock(void);
/* Waits until every submitted frame is thresholded and labelled. */
int sf_hip_server_drain(sf_server *server);
/* Drains and finishes the series' labeller; its results are then final. */
int sf_hip_server_close(sf_server *server, int series);
/* The series' labeller, for sf_hip_labeller_results and friends. */
sf_labeller *sf_hip_server_labeller(sf_server *server, int series);
/* Per completed frame of a series, in frame order: number, signal count
 * (above sf_hip_server_slot_bytes' list capacity, height * width *
 * pixel_bytes / 16, when the frame overflowed and was labelled empty),
 * arrival-to-ordered-list latency and arrival-to-labelled latency (ms). Any
 * output may be NULL; *n is the number of frames completed. */
int sf_hip_server_frames(sf_server *server, int series, int64_t *frames,
                         uint32_t *counts, double *mask_ms, double *table_ms,
                         size_t capacity, size_t *n);
/* The k-th completed frame's list (keep_lists servers only). */
int sf_hip_server_frame_list(sf_server *server, int series, size_t k,
                             sf_signal_pixel *out, size_t capacity,
                             size_t *n);
int sf_hip_server_stats(sf_server *server, sf_server_stats *stats);
/* Bytes of budget one frame's slot of this shape takes. */
size_t sf_hip_server_slot_bytes(int pixel_bytes, size_t height, size_t width);
/* Plays an arrival schedule and drains: frame f of series i (f < n_frames[i])
 * arrives at t0 + f / rate_hz[i] and is numbered first_frame + f; its pixels
 * are rings[sum(ring_lengths[:i]) + f % ring_lengths[i]]. Latencies are
 * measured from the scheduled arrival. */
int sf_hip_server_play(sf_server *server, int n_series, const int *series,
                       const double *rate_hz, const int64_t *n_frames,
                       const void *const *rings, const int *ring_lengths,
                       int64_t first_frame);
const char *sf_hip_server_last_error(void);

/* ---- synthetic frames on the device (benchmark and test input) ----------
 * Fills `frames` device frames of height x width (pixel_bytes 2 or 4) at
 * `device` with a Poisson background of mean `background`, masked module
 * gaps (0xFFFF.. at the module pitch when gaps != 0), hot pixels at
 * max-1 with probability hot_fraction, and the given spots: each spot is
 * (x, y, sigma, total counts, first frame, frame span) as doubles, six per
 * spot, drawn with a Gaussian rocking curve of sigma_z frames. */
int sf_hip_synth_frames(void *device, int pixel_bytes, uint32_t height,
                        uint32_t width, uint32_t frames, uint64_t seed,
                        double background, int gaps, double hot_fraction,
                        const double *spots, uint32_t n_spots,
                        double sigma_z);

/* --- The detector's pixel mask ------------------------------------------
 * No counterpart in the reference, which names it as its one omission
 * (README.md:227-234): DIALS masks the pixels of
 * /entry/instrument/detector/pixel_mask, the reference only the sentinels at
 * or above max - 1 (src/dext.cc:40). Stamping max(T) into a frame before the
 * threshold closes the gap: the threshold then treats a masked pixel as it
 * treats a module gap. The mask is a device bit array: bit (p & 31) of
 * bits[p >> 5] for pixel p of a row-major height x width frame.
 */

/* Writes max(T) (0xFFFF or 0xFFFFFFFF) at every masked pixel of each of the
 * n_frames device frames (pixel_bytes 2 or 4), and nothing else; a clear
 * mask byte (8 pixels) costs no frame traffic. Asynchronous on `stream` (a hipStream_t;
 * NULL: the null stream). Returns SF_OK, SF_BAD_ARGUMENT or SF_DEVICE_ERROR
 * (message in sf_hip_mask_last_error()). */
int sf_hip_stamp_mask(int pixel_bytes, void *const *frames, int n_frames,
                      size_t height, size_t width, const uint32_t *bits,
                      void *stream);

/* Packs a device byte mask (height*width bytes, nonzero = masked, as NXmx's
 * pixel_mask) into the bit array above (ceil(height*width / 32) words).
 * Asynchronous on `stream`. Returns SF_OK, SF_BAD_ARGUMENT or
 * SF_DEVICE_ERROR. */
int sf_hip_pack_mask(const uint8_t *mask, size_t height, size_t width,
                     uint32_t *bits, void *stream);

const char *sf_hip_mask_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* DEXT_HIP_H */
