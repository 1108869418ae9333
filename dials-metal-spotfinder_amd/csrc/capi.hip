// Host runtime behind include/dext_hip.h: plans (device buffers for a set of
// frame shapes), the per-thread workspace of the drop-in find(), window and
// profiling policy, graphs, errors.
//
// Reference counterparts: src/dext_gpu.cc (window policy, announce, stage
// times), src/dext_cuda.cu:442-728 (workspace, stream guard, find). The
// structure here is this build's own: a plan owns every buffer a launch
// needs, a frame is thresholded by one fused kernel and one ordering kernel,
// and the list comes back already ascending, so there is no host sort.
#include "sf_device.hh"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace sf {
void upload_sqrt_tables();
hipError_t launch_zero(const FrameDesc *, int, hipStream_t);
hipError_t launch_fused(int, bool, bool, const FrameDesc *, int, uint32_t, int,
                        hipStream_t);
hipError_t launch_order(const FrameDesc *, int, uint32_t, hipStream_t);
hipError_t launch_plain(int, const FrameDesc *, int, uint32_t, uint32_t,
                        uint8_t *, uint8_t *, hipStream_t, hipEvent_t,
                        hipEvent_t);
void sparse_layout(uint32_t, uint32_t, FrameDesc &);
size_t sparse_bytes(uint32_t, uint32_t, uint32_t);
void sparse_carve(uint8_t *, FrameDesc &);
hipError_t launch_sparse32_threshold(const FrameDesc *, int, uint32_t,
                                     hipStream_t, hipEvent_t *);
hipError_t launch_sparse_threshold(const FrameDesc *, int, uint32_t,
                                   hipStream_t, hipEvent_t *);
hipError_t launch_sparse_order(const FrameDesc *, int, uint32_t, hipStream_t);
hipError_t launch_sparse_places(const FrameDesc *, int, uint32_t, hipStream_t);
hipError_t launch_sparse_copy(const FrameDesc *, int, hipStream_t);
hipError_t launch_sparse(const FrameDesc *, int, uint32_t, uint32_t,
                         hipStream_t);
hipError_t launch_sparse32_pipeline(const FrameDesc *, int, uint32_t,
                                    uint32_t, const FrameDesc *, int, uint32_t,
                                    hipStream_t, hipEvent_t *, bool);
hipError_t launch_sparse_pipeline(const FrameDesc *, int, uint32_t, uint32_t,
                                  const FrameDesc *, int, uint32_t,
                                  hipStream_t, hipEvent_t *, bool);
} // namespace sf

namespace {

thread_local std::string g_last_error;

int fail(const char *what, hipError_t err) {
  g_last_error = std::string(what) + ": " + hipGetErrorString(err);
  return SF_DEVICE_ERROR;
}

int bad(const std::string &what) {
  g_last_error = what;
  return SF_BAD_ARGUMENT;
}

#define SF_CHECK(call, what)                                                   \
  do {                                                                         \
    const hipError_t e_ = (call);                                              \
    if (e_ != hipSuccess)                                                      \
      return fail(what, e_);                                                   \
  } while (0)

// The square-root table, once per device.
std::once_flag g_tables_once;
void ensure_tables() { std::call_once(g_tables_once, sf::upload_sqrt_tables); }

// ---------------------------------------------------------------------------
// Window policy (src/dext_gpu.cc:39-135): environment default, setters,
// announce on change.
// ---------------------------------------------------------------------------
std::atomic<int> g_window0{1}, g_window2{1};
std::once_flag g_windows_once;
std::atomic<bool> g_profiling{false};
std::atomic<bool> g_reporting{true};
std::atomic<int> g_announced{-1};

void resolve_one(const char *suffix, std::atomic<int> &into) {
  char name[64];
  std::snprintf(name, sizeof(name), "SPOTFINDER_GPU_%s", suffix);
  const char *asked = std::getenv(name);
  if (!asked)
    return;
  if (std::strcmp(asked, "direct") == 0)
    into = 0;
  else if (std::strcmp(asked, "tile") == 0)
    into = 1;
  else
    std::fprintf(stderr,
                 "warning: %s=%s is not 'tile' or 'direct'; leaving it alone\n",
                 name, asked);
}

void resolve_windows() {
  std::call_once(g_windows_once, [] {
    resolve_one("STAGE0", g_window0);
    resolve_one("STAGE2", g_window2);
  });
}

const char *window_name(int w) { return w ? "tile" : "direct"; }

void announce(int w0, int w2) {
  if (!g_reporting.load(std::memory_order_relaxed))
    return;
  const int now = w0 * 2 + w2;
  if (g_announced.exchange(now, std::memory_order_relaxed) == now)
    return;
  std::fprintf(stderr, "HIP: stage0 %s window, stage2 %s window (%s form)\n",
               window_name(w0), window_name(w2),
               (w0 && w2) ? "fused" : "plain");
}

thread_local double g_stage_times[6] = {0, 0, 0, 0, 0, 0};

} // namespace

// ---------------------------------------------------------------------------
// Plans.
// ---------------------------------------------------------------------------
struct sf_plan {
  int pixel_bytes = 2;
  int n = 0;
  std::vector<uint32_t> heights, widths;
  std::vector<uint32_t> capacity;
  int band_rows = 236;
  int form = 0;   // 0 fused (sparse passes; the sweep for 32-bit frames
                  // of 2^29 pixels or more), 1 plain, 2 sweep
  int blocks = 0; // 0: one block per item
  hipStream_t stream = nullptr;
  bool timing = false;

  std::vector<sf::FrameDesc> desc; // host copy
  sf::FrameDesc *h_desc = nullptr; // pinned staging for the upload
  sf::FrameDesc *d_desc = nullptr;
  sf_signal_pixel *d_staging = nullptr, *d_out = nullptr;
  uint8_t *d_cell_count = nullptr;
  uint32_t *d_cell_base = nullptr, *d_band_total = nullptr;
  uint32_t max_bands = 0, max_cell_rows = 0;
  size_t band_capacity = 0; // entries of d_band_total
  uint8_t *d_sparse = nullptr; // the sparse form's tables, every frame
  uint32_t *d_counters = nullptr, *d_found = nullptr;
  uint8_t *d_mask0 = nullptr, *d_mask1 = nullptr;
  size_t max_pixels = 0;
  uint32_t *h_found = nullptr;
  uint32_t items = 0;
  bool any_vec = false, any_scalar = false;
  hipEvent_t ev[6] = {};
  hipEvent_t uploaded = nullptr;
  bool upload_pending = false;
  std::vector<const void *> bound;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_form = -1;
  // pipelined steps: whether the bound frames have been scouted by a step
  // (and await their erosion, Poisson pass and ordering), the captured step
  // this plan owns and the plans it was captured for, and an event for the
  // waits between the two plans' streams
  bool scouted = false;
  hipGraph_t pipe_graph = nullptr;
  hipGraphExec_t pipe_exec = nullptr;
  sf_plan *pipe_cur = nullptr, *pipe_next = nullptr;
  hipEvent_t handoff = nullptr;
  // deferred ordering: a pipelined step that finishes this plan leaves its
  // ordering to sf_hip_plan_order (order_pending until then); whether the
  // captured step was captured so
  bool defer_order = false, order_pending = false, pipe_deferred = false;
  // a reader of a pending ordering took its places (plan_take_places): only
  // the copy into the lists is left, which sf_hip_plan_order does
  bool copy_pending = false;
  // a reader of the lists on another stream (the labeller's snapshot): the
  // plan's next write of its lists waits for it
  hipEvent_t read_done = nullptr;
  bool reader_pending = false;
};

namespace {

// The kernel a form runs for this plan's pixel type.
enum Kernel { kSweep, kPlain, kSparse };

Kernel kernel_for(const sf_plan *p, int form) {
  if (form == 1)
    return kPlain;
  // the sparse form at 32 bits addresses a frame's bytes with 32-bit
  // offsets: frames below 2^29 pixels (the sweep takes larger ones)
  if (form == 0 && (p->pixel_bytes == 2 || p->max_pixels < (1u << 29)))
    return kSparse;
  return kSweep;
}

uint32_t strips_for(const sf_plan *p, uint32_t width, int form) {
  const Kernel k = kernel_for(p, form);
  if (k == kPlain)
    return (width + 63) / 64;
  const uint32_t ow = p->pixel_bytes == 2 ? sf::strip_width<uint16_t>()
                                          : sf::strip_width<uint32_t>();
  return (width + ow - 1) / ow;
}

void drop_pipe_graph(sf_plan *p) {
  if (p->pipe_exec)
    (void)hipGraphExecDestroy(p->pipe_exec);
  if (p->pipe_graph)
    (void)hipGraphDestroy(p->pipe_graph);
  p->pipe_exec = nullptr;
  p->pipe_graph = nullptr;
  p->pipe_cur = p->pipe_next = nullptr;
}

void plan_free(sf_plan *p) {
  if (p->stream)
    (void)hipStreamSynchronize(p->stream);
  if (p->reader_pending)
    (void)hipEventSynchronize(p->read_done);
  drop_pipe_graph(p);
  if (p->handoff)
    (void)hipEventDestroy(p->handoff);
  if (p->read_done)
    (void)hipEventDestroy(p->read_done);
  if (p->exec)
    (void)hipGraphExecDestroy(p->exec);
  if (p->graph)
    (void)hipGraphDestroy(p->graph);
  for (hipEvent_t &e : p->ev)
    if (e)
      (void)hipEventDestroy(e);
  if (p->uploaded)
    (void)hipEventDestroy(p->uploaded);
  (void)hipFree(p->d_desc);
  (void)hipFree(p->d_staging);
  (void)hipFree(p->d_out);
  (void)hipFree(p->d_cell_count);
  (void)hipFree(p->d_cell_base);
  (void)hipFree(p->d_band_total);
  (void)hipFree(p->d_counters);
  (void)hipFree(p->d_found);
  (void)hipFree(p->d_mask0);
  (void)hipFree(p->d_mask1);
  (void)hipFree(p->d_sparse);
  if (p->h_desc)
    (void)hipHostFree(p->h_desc);
  if (p->h_found)
    (void)hipHostFree(p->h_found);
  if (p->stream)
    (void)hipStreamDestroy(p->stream);
}

// Lay out descriptors for the current form (strip width differs by form).
void layout(sf_plan *p) {
  uint32_t item = 0;
  size_t staging = 0, cells = 0, bands = 0, sparse = 0;
  p->any_vec = p->any_scalar = false;
  p->max_bands = p->max_cell_rows = 0;
  const Kernel kernel = kernel_for(p, p->form);
  for (int f = 0; f < p->n; f++) {
    sf::FrameDesc &d = p->desc[f];
    d.height = p->heights[f];
    d.width = p->widths[f];
    d.capacity = p->capacity[f];
    if (kernel == kSparse) {
      sf::sparse_layout(d.height, d.width, d);
      sf::sparse_carve(p->d_sparse + sparse, d);
      sparse += sf::sparse_bytes(d.height, d.width, d.capacity);
      p->max_cell_rows = std::max(p->max_cell_rows, d.n_cell_rows);
    } else {
      d.n_strips = strips_for(p, d.width, p->form);
      d.band_rows = (uint32_t)p->band_rows;
      d.n_bands = (d.height + d.band_rows - 1) / d.band_rows;
    }
    d.item_begin = item;
    item += d.n_strips * d.n_bands;
    d.staging = p->d_staging + staging;
    d.out = p->d_out + staging;
    staging += d.capacity;
    d.cell_count = p->d_cell_count + cells;
    d.cell_base = p->d_cell_base + cells;
    cells += (size_t)d.height * ((d.width + 63) / 64);
    d.band_total = p->d_band_total + bands;
    bands += d.n_bands;
    p->max_bands = d.n_bands > p->max_bands ? d.n_bands : p->max_bands;
    d.counter = p->d_counters + f;
    d.out_count = p->d_found + f;
  }
  p->items = item;
}

// Work enqueued on `s` that writes the plan's lists waits for their last
// reader on another stream.
int wait_reader(sf_plan *p, hipStream_t s) {
  if (p && p->reader_pending) {
    SF_CHECK(hipStreamWaitEvent(s, p->read_done, 0), "hipStreamWaitEvent");
    p->reader_pending = false;
  }
  return SF_OK;
}

int upload_desc(sf_plan *p, const void *const *frames) {
  const int k = p->pixel_bytes == 2 ? 4 : 2;
  bool changed = p->bound.size() != (size_t)p->n;
  if (!changed)
    for (int f = 0; f < p->n; f++)
      changed |= p->bound[f] != frames[f];
  // Row loads of several pixels: the sparse scout reads 16 bytes from an
  // even column, so 4-byte aligned frames of even width suffice; the other
  // forms read K pixels a lane and need 8-byte alignment and K | width.
  const bool sparse = kernel_for(p, p->form) == kSparse;
  p->any_vec = p->any_scalar = false;
  for (int f = 0; f < p->n; f++) {
    sf::FrameDesc &d = p->desc[f];
    d.pixels = frames[f];
    const uintptr_t at = reinterpret_cast<uintptr_t>(frames[f]);
    d.vec = sparse ? (at % 4 == 0 && (p->pixel_bytes == 4 || d.width % 2 == 0)
                          ? 1u
                          : 0u)
                   : (at % 8 == 0 && d.width % (uint32_t)k == 0 ? 1u : 0u);
    p->any_vec |= d.vec != 0;
    p->any_scalar |= d.vec == 0;
  }
  if (!changed)
    return SF_OK;
  // The pinned staging may still be read by an earlier upload.
  if (p->upload_pending)
    SF_CHECK(hipEventSynchronize(p->uploaded), "waiting for an upload");
  std::memcpy(p->h_desc, p->desc.data(), sizeof(sf::FrameDesc) * p->n);
  SF_CHECK(hipMemcpyAsync(p->d_desc, p->h_desc, sizeof(sf::FrameDesc) * p->n,
                          hipMemcpyHostToDevice, p->stream),
           "uploading frame descriptors");
  SF_CHECK(hipEventRecord(p->uploaded, p->stream), "hipEventRecord");
  p->upload_pending = true;
  p->bound.assign(frames, frames + p->n);
  return SF_OK;
}

int enqueue(sf_plan *p, bool timing) {
  const Kernel kernel = kernel_for(p, p->form);
  if (kernel == kSparse) { // clears its own counters and orders its lists
    // ev[0] after the clearing kernel, ev[1] after the scout, ev[2] = ev[3]
    // after the Poisson pass, ev[4] after the ordering
    if (p->pixel_bytes == 4)
      SF_CHECK(sf::launch_sparse32_threshold(p->d_desc, p->n, p->items,
                                             p->stream,
                                             timing ? p->ev : nullptr),
               "launching the sparse threshold");
    else
      SF_CHECK(sf::launch_sparse_threshold(p->d_desc, p->n, p->items,
                                           p->stream,
                                           timing ? p->ev : nullptr),
               "launching the sparse threshold");
    if (timing)
      for (int e = 2; e <= 3; e++)
        SF_CHECK(hipEventRecord(p->ev[e], p->stream), "hipEventRecord");
    SF_CHECK(sf::launch_sparse_order(p->d_desc, p->n, p->max_cell_rows,
                                     p->stream),
             "launching the sparse ordering");
    if (timing)
      SF_CHECK(hipEventRecord(p->ev[4], p->stream), "hipEventRecord");
    return SF_OK;
  }
  SF_CHECK(sf::launch_zero(p->d_desc, p->n, p->stream), "zeroing counters");
  if (timing)
    SF_CHECK(hipEventRecord(p->ev[0], p->stream), "hipEventRecord");
  if (kernel == kSweep) {
    SF_CHECK(sf::launch_fused(p->pixel_bytes, p->any_vec, p->any_scalar,
                              p->d_desc, p->n, p->items, p->blocks,
                              p->stream),
             "launching the fused threshold");
    if (timing) {
      SF_CHECK(hipEventRecord(p->ev[1], p->stream), "hipEventRecord");
      SF_CHECK(hipEventRecord(p->ev[2], p->stream), "hipEventRecord");
    }
  } else {
    for (int f = 0; f < p->n; f++) {
      const bool first = f == 0 && timing;
      SF_CHECK(sf::launch_plain(p->pixel_bytes, p->d_desc, f, p->heights[f],
                                p->widths[f], p->d_mask0, p->d_mask1,
                                p->stream, first ? p->ev[1] : nullptr,
                                first ? p->ev[2] : nullptr),
               "launching the plain threshold");
    }
  }
  if (timing)
    SF_CHECK(hipEventRecord(p->ev[3], p->stream), "hipEventRecord");
  SF_CHECK(sf::launch_order(p->d_desc, p->n, p->max_bands, p->stream),
           "launching the ordering kernel");
  if (timing)
    SF_CHECK(hipEventRecord(p->ev[4], p->stream), "hipEventRecord");
  return SF_OK;
}

} // namespace

extern "C" {

sf_plan *sf_hip_plan_create(int pixel_bytes, int n_frames,
                            const uint32_t *heights, const uint32_t *widths,
                            int band_rows) {
  if ((pixel_bytes != 2 && pixel_bytes != 4) || n_frames <= 0 || !heights ||
      !widths) {
    bad("sf_hip_plan_create: bad argument");
    return nullptr;
  }
  for (int f = 0; f < n_frames; f++) {
    if (!heights[f] || !widths[f] || heights[f] > 0x7fffffffu / widths[f]) {
      bad("sf_hip_plan_create: a frame has no pixels or too many");
      return nullptr;
    }
  }
  ensure_tables();
  sf_plan *p = new sf_plan;
  p->pixel_bytes = pixel_bytes;
  p->n = n_frames;
  p->heights.assign(heights, heights + n_frames);
  p->widths.assign(widths, widths + n_frames);
  if (band_rows > 0)
    p->band_rows = ((band_rows + 20 + 15) & ~15) - 20;
  size_t staging = 0, cells = 0, bands = 0, sparse = 0;
  for (int f = 0; f < n_frames; f++) {
    const size_t px = (size_t)heights[f] * widths[f];
    // The list holds one entry per 16 bytes of frame, as the reference's
    // device buffer does (src/dext_cuda.cu:614-616).
    p->capacity.push_back((uint32_t)(px * (size_t)pixel_bytes / 16));
    staging += p->capacity.back();
    sparse += sf::sparse_bytes(heights[f], widths[f], p->capacity.back());
    cells += (size_t)heights[f] * ((widths[f] + 63) / 64);
    // band totals for the plan's band (the autotuner grows them for
    // shorter candidates)
    bands += (heights[f] + (uint32_t)p->band_rows - 1) / (uint32_t)p->band_rows;
    p->max_pixels = px > p->max_pixels ? px : p->max_pixels;
  }
  p->desc.resize(n_frames);
  bool ok = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) ==
            hipSuccess;
  ok = ok && hipMalloc(&p->d_desc, sizeof(sf::FrameDesc) * n_frames) ==
                 hipSuccess;
  ok = ok && hipHostMalloc(&p->h_desc, sizeof(sf::FrameDesc) * n_frames) ==
                 hipSuccess;
  ok = ok && hipMalloc(&p->d_staging,
                       sizeof(sf_signal_pixel) * (staging ? staging : 1)) ==
                 hipSuccess;
  ok = ok && hipMalloc(&p->d_out, sizeof(sf_signal_pixel) *
                                      (staging ? staging : 1)) == hipSuccess;
  // the sparse form's tables (u32 frames of 2^29 pixels and more never
  // take that form)
  if (pixel_bytes == 4 && p->max_pixels >= (1u << 29))
    sparse = 0;
  if (sparse)
    ok = ok && hipMalloc(&p->d_sparse, sparse) == hipSuccess;
  ok = ok && hipMalloc(&p->d_cell_count, cells) == hipSuccess;
  ok = ok && hipMalloc(&p->d_cell_base, sizeof(uint32_t) * cells) ==
                 hipSuccess;
  ok = ok && hipMalloc(&p->d_band_total, sizeof(uint32_t) * bands) ==
                 hipSuccess;
  p->band_capacity = bands;
  ok = ok && hipMalloc(&p->d_counters, sizeof(uint32_t) * n_frames) ==
                 hipSuccess;
  ok = ok && hipMalloc(&p->d_found, sizeof(uint32_t) * n_frames) ==
                 hipSuccess;
  ok = ok && hipHostMalloc(&p->h_found, sizeof(uint32_t) * n_frames) ==
                 hipSuccess;
  for (hipEvent_t &e : p->ev)
    ok = ok && hipEventCreate(&e) == hipSuccess;
  ok = ok && hipEventCreateWithFlags(&p->uploaded, hipEventDisableTiming) ==
                 hipSuccess;
  ok = ok && hipEventCreateWithFlags(&p->handoff, hipEventDisableTiming) ==
                 hipSuccess;
  ok = ok && hipEventCreateWithFlags(&p->read_done, hipEventDisableTiming) ==
                 hipSuccess;
  if (!ok) {
    g_last_error = "sf_hip_plan_create: device allocation failed";
    plan_free(p);
    delete p;
    return nullptr;
  }
  layout(p);
  return p;
}

void sf_hip_plan_destroy(sf_plan *plan) {
  if (!plan)
    return;
  plan_free(plan);
  delete plan;
}

int sf_hip_plan_set_form(sf_plan *p, int form) {
  if (!p || form < 0 || form > 2)
    return bad("sf_hip_plan_set_form: bad argument");
  if (form == p->form)
    return SF_OK;
  if (kernel_for(p, form) == kPlain && !p->d_mask0) {
    SF_CHECK(hipMalloc(&p->d_mask0, p->max_pixels), "allocating mask0");
    SF_CHECK(hipMalloc(&p->d_mask1, p->max_pixels), "allocating mask1");
  }
  SF_CHECK(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
  p->form = form;
  layout(p);
  p->bound.clear(); // force a descriptor upload
  p->scouted = false;
  drop_pipe_graph(p);
  return SF_OK;
}

int sf_hip_plan_set_grid(sf_plan *p, int blocks) {
  if (!p || blocks < 0)
    return bad("sf_hip_plan_set_grid: bad argument");
  p->blocks = blocks;
  p->graph_form = -1;
  p->scouted = false;
  drop_pipe_graph(p);
  return SF_OK;
}

int sf_hip_plan_autotune(sf_plan *p, const void *const *frames, int trials,
                         int *band_rows, int *blocks) {
  if (!p || !frames || trials <= 0)
    return bad("sf_hip_plan_autotune: bad argument");
  const Kernel kernel = kernel_for(p, p->form);
  if (kernel == kPlain)
    return bad("sf_hip_plan_autotune: the plain form has nothing to tune");
  if (kernel == kSparse) { // its launch shape follows the frames
    if (band_rows)
      *band_rows = (int)p->desc[0].band_rows;
    if (blocks)
      *blocks = 0;
    return SF_OK;
  }
  int cus = 256;
  {
    int device = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&device) == hipSuccess &&
        hipGetDeviceProperties(&prop, device) == hipSuccess)
      cus = prop.multiProcessorCount;
  }
  // candidate band heights (the rows a sweep tile covers)
  const int bands[] = {44, 108, 236, 492};
  const int grids[] = {0, 4 * cus, 8 * cus, 16 * cus};
  // band totals for the shortest candidate band
  size_t need_bands = 0;
  for (int f = 0; f < p->n; f++)
    need_bands += (p->heights[f] + 43) / 44;
  SF_CHECK(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
  if (p->exec) {
    (void)hipGraphExecDestroy(p->exec);
    p->exec = nullptr;
  }
  if (p->graph) {
    (void)hipGraphDestroy(p->graph);
    p->graph = nullptr;
  }
  p->graph_form = -1;
  p->scouted = false;
  drop_pipe_graph(p);
  if (need_bands > p->band_capacity) {
    uint32_t *grown = nullptr;
    SF_CHECK(hipMalloc(&grown, sizeof(uint32_t) * need_bands),
             "allocating band totals");
    (void)hipFree(p->d_band_total);
    p->d_band_total = grown;
    p->band_capacity = need_bands;
  }
  const bool was_timing = p->timing;
  double best = 1e30;
  int best_band = p->band_rows, best_grid = p->blocks;
  for (const int band : bands) {
    p->band_rows = band;
    layout(p);
    p->bound.clear(); // force a descriptor upload
    for (const int grid : grids) {
      p->blocks = grid;
      int status = sf_hip_plan_run(p, frames); // warm-up
      if (status != SF_OK)
        return status;
      std::vector<float> times;
      for (int t = 0; t < trials; t++) {
        status = enqueue(p, true);
        if (status != SF_OK)
          return status;
        SF_CHECK(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
        float ms = 0.0f;
        SF_CHECK(hipEventElapsedTime(&ms, p->ev[0], p->ev[4]), "event time");
        times.push_back(ms);
      }
      std::sort(times.begin(), times.end());
      const double median = times[times.size() / 2];
      if (median < best) {
        best = median;
        best_band = band;
        best_grid = grid;
      }
    }
  }
  p->timing = was_timing;
  p->band_rows = best_band;
  p->blocks = best_grid;
  layout(p);
  p->bound.clear();
  SF_CHECK(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
  p->upload_pending = false;
  if (band_rows)
    *band_rows = best_band;
  if (blocks)
    *blocks = best_grid;
  return SF_OK;
}


int sf_hip_plan_set_timing(sf_plan *p, int on) {
  if (!p)
    return bad("sf_hip_plan_set_timing: no plan");
  p->timing = on != 0;
  return SF_OK;
}

int sf_hip_plan_run(sf_plan *p, const void *const *frames) {
  if (!p || !frames)
    return bad("sf_hip_plan_run: bad argument");
  for (int f = 0; f < p->n; f++)
    if (!frames[f])
      return bad("sf_hip_plan_run: a frame pointer is null");
  const int up = upload_desc(p, frames);
  if (up != SF_OK)
    return up;
  p->scouted = false;
  p->order_pending = false; // a run orders its own lists
  p->copy_pending = false;
  const int waited = wait_reader(p, p->stream);
  if (waited != SF_OK)
    return waited;
  return enqueue(p, p->timing);
}

int sf_hip_plan_bind(sf_plan *p, const void *const *frames) {
  if (!p || !frames)
    return bad("sf_hip_plan_bind: bad argument");
  for (int f = 0; f < p->n; f++)
    if (!frames[f])
      return bad("sf_hip_plan_bind: a frame pointer is null");
  p->scouted = false; // the tables no longer describe the bound frames
  return upload_desc(p, frames);
}

int sf_hip_plan_capture(sf_plan *p, const void *const *frames) {
  if (!p || !frames)
    return bad("sf_hip_plan_capture: bad argument");
  const int up = upload_desc(p, frames);
  if (up != SF_OK)
    return up;
  SF_CHECK(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
  if (p->exec) {
    (void)hipGraphExecDestroy(p->exec);
    p->exec = nullptr;
  }
  if (p->graph) {
    (void)hipGraphDestroy(p->graph);
    p->graph = nullptr;
  }
  SF_CHECK(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal),
           "hipStreamBeginCapture");
  const int status = enqueue(p, false);
  hipGraph_t graph = nullptr;
  const hipError_t end = hipStreamEndCapture(p->stream, &graph);
  if (status != SF_OK)
    return status;
  SF_CHECK(end, "hipStreamEndCapture");
  p->graph = graph;
  SF_CHECK(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0),
           "hipGraphInstantiate");
  p->graph_form = p->form;
  return SF_OK;
}

int sf_hip_plan_replay(sf_plan *p) {
  if (!p || !p->exec)
    return bad("sf_hip_plan_replay: nothing captured");
  p->order_pending = false; // a replayed run orders its own lists
  p->copy_pending = false;
  const int waited = wait_reader(p, p->stream);
  if (waited != SF_OK)
    return waited;
  SF_CHECK(hipGraphLaunch(p->exec, p->stream), "hipGraphLaunch");
  return SF_OK;
}

int sf_hip_plan_wait(sf_plan *p) {
  if (!p)
    return bad("sf_hip_plan_wait: no plan");
  SF_CHECK(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
  p->upload_pending = false;
  return SF_OK;
}

int sf_hip_plan_counts(sf_plan *p, uint32_t *counts) {
  if (!p || !counts)
    return bad("sf_hip_plan_counts: bad argument");
  SF_CHECK(hipMemcpyAsync(p->h_found, p->d_found, sizeof(uint32_t) * p->n,
                          hipMemcpyDeviceToHost, p->stream),
           "reading counts");
  SF_CHECK(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
  std::memcpy(counts, p->h_found, sizeof(uint32_t) * p->n);
  return SF_OK;
}

const sf_signal_pixel *sf_hip_plan_output(sf_plan *p, int frame) {
  if (!p || frame < 0 || frame >= p->n)
    return nullptr;
  return p->desc[frame].out;
}

uint32_t sf_hip_plan_capacity(sf_plan *p, int frame) {
  if (!p || frame < 0 || frame >= p->n)
    return 0;
  return p->capacity[frame];
}

int sf_hip_plan_last_times(sf_plan *p, double *threshold_ms,
                           double *order_ms) {
  if (!p || !p->timing)
    return bad("sf_hip_plan_last_times: timing is off");
  float a = 0.0f, b = 0.0f;
  SF_CHECK(hipEventElapsedTime(&a, p->ev[0], p->ev[3]), "event time");
  SF_CHECK(hipEventElapsedTime(&b, p->ev[3], p->ev[4]), "event time");
  if (threshold_ms)
    *threshold_ms = a;
  if (order_ms)
    *order_ms = b;
  return SF_OK;
}

int sf_hip_plan_last_pass_times(sf_plan *p, double *first_ms,
                                double *rest_ms, double *order_ms) {
  if (!p || !p->timing)
    return bad("sf_hip_plan_last_pass_times: timing is off");
  float a = 0.0f, b = 0.0f, c = 0.0f;
  SF_CHECK(hipEventElapsedTime(&a, p->ev[0], p->ev[1]), "event time");
  SF_CHECK(hipEventElapsedTime(&b, p->ev[1], p->ev[3]), "event time");
  SF_CHECK(hipEventElapsedTime(&c, p->ev[3], p->ev[4]), "event time");
  if (first_ms)
    *first_ms = a;
  if (rest_ms)
    *rest_ms = b;
  if (order_ms)
    *order_ms = c;
  return SF_OK;
}

void *sf_hip_plan_stream(sf_plan *p) { return p ? (void *)p->stream : nullptr; }

} // extern "C"

namespace {

// Checks the pair of a pipelined step and returns the stream it runs on:
// the current plan's (the next plan's when there is none).
int pipeline_check(sf_plan *cur, sf_plan *next, const char *what,
                   hipStream_t *stream, bool running = true) {
  if ((!cur && !next) || cur == next)
    return bad(std::string(what) + ": needs one plan or two different ones");
  for (sf_plan *p : {cur, next}) {
    if (p && kernel_for(p, p->form) != kSparse)
      return bad(std::string(what) +
                 ": pipelined steps take plans of the sparse form (form 0)");
  }
  if (cur && next && cur->pixel_bytes != next->pixel_bytes)
    return bad(std::string(what) +
               ": the two plans' pixels differ in size");
  // a capture runs nothing: what it needs is checked when it is replayed
  if (running && cur && !cur->scouted)
    return bad(std::string(what) +
               ": the current plan's frames were not scouted by a step");
  // scouting clears the tables that a pending ordering reads
  if (running && next && next->order_pending)
    return bad(std::string(what) +
               ": the next plan's ordering is pending (sf_hip_plan_order)");
  if (running && next && next->bound.size() != (size_t)next->n)
    return bad(std::string(what) + ": the next plan has no frames bound");
  *stream = cur ? cur->stream : next->stream;
  return SF_OK;
}

// The other plan's stream must be done with what the step overwrites (its
// tables and descriptors) before the step, and its own later work must
// follow the step.
int pipeline_wait_other(sf_plan *next, hipStream_t s) {
  if (next && next->stream != s) {
    SF_CHECK(hipEventRecord(next->handoff, next->stream), "hipEventRecord");
    SF_CHECK(hipStreamWaitEvent(s, next->handoff, 0), "hipStreamWaitEvent");
  }
  return SF_OK;
}

int pipeline_after(sf_plan *cur, sf_plan *next, hipStream_t s) {
  sf_plan *owner = cur ? cur : next;
  if (next && next->stream != s) {
    SF_CHECK(hipEventRecord(owner->handoff, s), "hipEventRecord");
    SF_CHECK(hipStreamWaitEvent(next->stream, owner->handoff, 0),
             "hipStreamWaitEvent");
  }
  if (cur) {
    cur->scouted = false;
    cur->order_pending = cur->defer_order;
    cur->copy_pending = false;
  }
  if (next) {
    next->scouted = true;
    next->copy_pending = false; // the scout rewrote what the copy reads
  }
  return SF_OK;
}

int pipeline_enqueue(sf_plan *cur, sf_plan *next, hipStream_t s,
                     bool timing, bool order) {
  sf_plan *owner = cur ? cur : next;
  // events as the plan's pass times read them: first mixed launch, second,
  // ordering
  hipEvent_t marks[4] = {owner->ev[0], owner->ev[1], owner->ev[3],
                         owner->ev[4]};
  const auto launch = owner->pixel_bytes == 4 ? sf::launch_sparse32_pipeline
                                               : sf::launch_sparse_pipeline;
  SF_CHECK(launch(cur ? cur->d_desc : nullptr, cur ? cur->n : 0,
                  cur ? cur->items : 0, cur ? cur->max_cell_rows : 0,
                  next ? next->d_desc : nullptr, next ? next->n : 0,
                  next ? next->items : 0, s, timing ? marks : nullptr,
                  order),
           "launching a pipelined step");
  return SF_OK;
}

} // namespace

extern "C" {

int sf_hip_plan_pipeline(sf_plan *cur, sf_plan *next) {
  hipStream_t s = nullptr;
  int status = pipeline_check(cur, next, "sf_hip_plan_pipeline", &s);
  if (status != SF_OK)
    return status;
  if ((status = pipeline_wait_other(next, s)) != SF_OK)
    return status;
  // the next plan's clearing and scout rewrite the tables that a reader
  // which took its last ordering's places may still use (plan_take_places);
  // a reader of its lists waits only for their next write, by the ordering
  if (next && next->copy_pending && (status = wait_reader(next, s)) != SF_OK)
    return status;
  const bool order = !(cur && cur->defer_order);
  if (order && (status = wait_reader(cur, s)) != SF_OK)
    return status;
  sf_plan *owner = cur ? cur : next;
  if ((status = pipeline_enqueue(cur, next, s, owner->timing, order)) !=
      SF_OK)
    return status;
  return pipeline_after(cur, next, s);
}

int sf_hip_plan_capture_pipeline(sf_plan *cur, sf_plan *next) {
  hipStream_t s = nullptr;
  const int status = pipeline_check(cur, next, "sf_hip_plan_capture_pipeline",
                                    &s, false);
  if (status != SF_OK)
    return status;
  sf_plan *owner = cur ? cur : next;
  SF_CHECK(hipStreamSynchronize(s), "hipStreamSynchronize");
  drop_pipe_graph(owner);
  SF_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal),
           "hipStreamBeginCapture");
  const bool order = !(cur && cur->defer_order);
  const int enqueued = pipeline_enqueue(cur, next, s, false, order);
  hipGraph_t graph = nullptr;
  const hipError_t end = hipStreamEndCapture(s, &graph);
  if (enqueued != SF_OK)
    return enqueued;
  SF_CHECK(end, "hipStreamEndCapture");
  owner->pipe_graph = graph;
  SF_CHECK(hipGraphInstantiate(&owner->pipe_exec, graph, nullptr, nullptr, 0),
           "hipGraphInstantiate");
  owner->pipe_cur = cur;
  owner->pipe_next = next;
  owner->pipe_deferred = !order;
  return SF_OK;
}

int sf_hip_plan_replay_pipeline(sf_plan *cur, sf_plan *next) {
  hipStream_t s = nullptr;
  int status = pipeline_check(cur, next, "sf_hip_plan_replay_pipeline", &s);
  if (status != SF_OK)
    return status;
  sf_plan *owner = cur ? cur : next;
  if (!owner->pipe_exec || owner->pipe_cur != cur || owner->pipe_next != next)
    return bad("sf_hip_plan_replay_pipeline: no step captured for these "
               "plans");
  if (cur && owner->pipe_deferred != cur->defer_order)
    return bad("sf_hip_plan_replay_pipeline: the step was captured with "
               "the other ordering setting; capture it again");
  if ((status = pipeline_wait_other(next, s)) != SF_OK)
    return status;
  if (next && next->copy_pending && (status = wait_reader(next, s)) != SF_OK)
    return status;
  if (!owner->pipe_deferred && (status = wait_reader(cur, s)) != SF_OK)
    return status;
  SF_CHECK(hipGraphLaunch(owner->pipe_exec, s), "hipGraphLaunch");
  return pipeline_after(cur, next, s);
}

int sf_hip_plan_set_deferred_order(sf_plan *p, int deferred) {
  if (!p)
    return bad("sf_hip_plan_set_deferred_order: no plan");
  if (p->order_pending)
    return bad("sf_hip_plan_set_deferred_order: an ordering is pending");
  p->defer_order = deferred != 0;
  return SF_OK;
}

int sf_hip_plan_order(sf_plan *p) {
  if (!p)
    return bad("sf_hip_plan_order: no plan");
  if (!p->order_pending && !p->copy_pending)
    return bad("sf_hip_plan_order: no pipelined step left this plan's "
               "ordering pending");
  // after the step (enqueued on this plan's stream), and after the last
  // reader of the lists it rewrites (or of the places it copies by)
  const int waited = wait_reader(p, p->stream);
  if (waited != SF_OK)
    return waited;
  if (p->order_pending)
    SF_CHECK(sf::launch_sparse_order(p->d_desc, p->n, p->max_cell_rows,
                                     p->stream),
             "launching the ordering");
  else
    SF_CHECK(sf::launch_sparse_copy(p->d_desc, p->n, p->stream),
             "launching the copy into the lists");
  p->order_pending = false;
  p->copy_pending = false;
  return SF_OK;
}

// ---------------------------------------------------------------------------
// Device, memory, policy.
// ---------------------------------------------------------------------------
const char *sf_hip_backend(void) { return "HIP"; }

int sf_hip_available(void) {
  int devices = 0;
  return hipGetDeviceCount(&devices) == hipSuccess && devices > 0;
}

size_t sf_hip_memory_free(void) {
  size_t free_bytes = 0, total = 0;
  if (hipMemGetInfo(&free_bytes, &total) != hipSuccess)
    return 0;
  return free_bytes;
}

void *sf_hip_host_alloc(size_t bytes) {
  void *p = nullptr;
  if (hipHostMalloc(&p, bytes) != hipSuccess)
    return nullptr;
  return p;
}

void sf_hip_host_free(void *p) {
  if (p)
    (void)hipHostFree(p);
}

const char *sf_hip_last_error(void) { return g_last_error.c_str(); }

void *sf_hip_device_alloc(size_t bytes) {
  void *p = nullptr;
  const hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    fail("hipMalloc", e);
    return nullptr;
  }
  return p;
}

void sf_hip_device_free(void *p) {
  if (p)
    (void)hipFree(p);
}

int sf_hip_copy_to_device(void *device, const void *host, size_t bytes) {
  SF_CHECK(hipMemcpy(device, host, bytes, hipMemcpyHostToDevice), "hipMemcpy");
  return SF_OK;
}

int sf_hip_copy_to_host(void *host, const void *device, size_t bytes) {
  SF_CHECK(hipMemcpy(host, device, bytes, hipMemcpyDeviceToHost), "hipMemcpy");
  return SF_OK;
}

int sf_hip_synchronize(void) {
  SF_CHECK(hipDeviceSynchronize(), "hipDeviceSynchronize");
  return SF_OK;
}

int sf_hip_stage0_window(void) {
  resolve_windows();
  return g_window0.load();
}
void sf_hip_set_stage0_window(int w) {
  resolve_windows();
  g_window0 = w ? 1 : 0;
}
int sf_hip_stage2_window(void) {
  resolve_windows();
  return g_window2.load();
}
void sf_hip_set_stage2_window(int w) {
  resolve_windows();
  g_window2 = w ? 1 : 0;
}
void sf_hip_profile_stages(int on) { g_profiling = on != 0; }
int sf_hip_profiling(void) { return g_profiling.load() ? 1 : 0; }
void sf_hip_last_stage_times(double times[6]) {
  for (int i = 0; i < 6; i++)
    times[i] = g_stage_times[i];
}
void sf_hip_report_windows(int on) { g_reporting = on != 0; }

} // extern "C"

// ---------------------------------------------------------------------------
// The drop-in find(): per-thread workspace and stream, synchronous.
// ---------------------------------------------------------------------------
namespace {

struct Workspace {
  sf_plan *plan = nullptr;
  int pixel_bytes = 0;
  uint32_t height = 0, width = 0;
  void *frame = nullptr;
  size_t frame_bytes = 0;
  hipEvent_t before_upload = nullptr;
  ~Workspace() {
    if (plan)
      sf_hip_plan_destroy(plan);
    if (frame)
      (void)hipFree(frame);
    if (before_upload)
      (void)hipEventDestroy(before_upload);
  }
};

Workspace &workspace() {
  static thread_local Workspace w;
  return w;
}

// Leaves the stream idle on every exit path (the reference's StreamGuard,
// src/dext_cuda.cu:442-464).
struct Idle {
  hipStream_t s;
  ~Idle() {
    if (s)
      (void)hipStreamSynchronize(s);
  }
};

template <typename T>
int find(const T *image, size_t height, size_t width, sf_signal_pixel *out,
         size_t capacity, size_t *found) {
  if (found)
    *found = 0;
  if (!image || height == 0 || width == 0 || !found)
    return bad("find: null image or zero size");
  if (height > 0x7fffffffu / width)
    return bad("find: frame larger than INT32_MAX pixels");
  Workspace &w = workspace();
  const int pb = (int)sizeof(T);
  const int w0 = sf_hip_stage0_window(), w2 = sf_hip_stage2_window();
  announce(w0, w2);
  const int form = (w0 && w2) ? 0 : 1;
  if (!w.plan || w.pixel_bytes != pb || w.height != height ||
      w.width != width) {
    if (w.plan)
      sf_hip_plan_destroy(w.plan);
    const uint32_t h32 = (uint32_t)height, w32 = (uint32_t)width;
    w.plan = sf_hip_plan_create(pb, 1, &h32, &w32, 0);
    if (!w.plan)
      return SF_DEVICE_ERROR;
    w.pixel_bytes = pb;
    w.height = (uint32_t)height;
    w.width = (uint32_t)width;
  }
  int status = sf_hip_plan_set_form(w.plan, form);
  if (status != SF_OK)
    return status;
  const size_t bytes = height * width * sizeof(T);
  if (bytes > w.frame_bytes) {
    (void)hipStreamSynchronize(w.plan->stream);
    if (w.frame)
      (void)hipFree(w.frame);
    w.frame = nullptr;
    w.frame_bytes = 0;
    SF_CHECK(hipMalloc(&w.frame, bytes), "allocating the frame");
    w.frame_bytes = bytes;
  }
  if (!w.before_upload)
    SF_CHECK(hipEventCreate(&w.before_upload), "hipEventCreate");
  sf_plan *p = w.plan;
  const Idle idle{p->stream};
  const bool timing = sf_hip_profiling() != 0;
  p->timing = timing;
  if (timing)
    SF_CHECK(hipEventRecord(w.before_upload, p->stream), "hipEventRecord");
  SF_CHECK(hipMemcpyAsync(w.frame, image, bytes, hipMemcpyHostToDevice,
                          p->stream),
           "copying a frame to the device");
  if (timing)
    SF_CHECK(hipEventRecord(p->ev[5], p->stream), "hipEventRecord");
  const void *frames[1] = {w.frame};
  status = sf_hip_plan_run(p, frames);
  if (status != SF_OK)
    return status;
  uint32_t n = 0;
  status = sf_hip_plan_counts(p, &n);
  if (status != SF_OK)
    return status;
  p->upload_pending = false;
  double times[6] = {0, 0, 0, 0, 0, 0};
  if (timing) {
    float ms = 0.0f;
    SF_CHECK(hipEventElapsedTime(&ms, w.before_upload, p->ev[5]), "time");
    times[0] = ms;
    if (form == 0) {
      SF_CHECK(hipEventElapsedTime(&ms, p->ev[0], p->ev[1]), "time");
      times[1] = ms;
      SF_CHECK(hipEventElapsedTime(&ms, p->ev[3], p->ev[4]), "time");
      times[3] = ms;
    } else {
      SF_CHECK(hipEventElapsedTime(&ms, p->ev[0], p->ev[1]), "time");
      times[1] = ms;
      SF_CHECK(hipEventElapsedTime(&ms, p->ev[1], p->ev[2]), "time");
      times[2] = ms;
      SF_CHECK(hipEventElapsedTime(&ms, p->ev[2], p->ev[4]), "time");
      times[3] = ms;
    }
  }
  *found = n;
  if (n > p->capacity[0]) {
    if (timing)
      std::memcpy(g_stage_times, times, sizeof(times));
    g_last_error = "more signal pixels than the device list holds";
    return SF_OVERFLOW;
  }
  const auto copy_started = std::chrono::steady_clock::now();
  const size_t take = n < capacity ? n : capacity;
  if (take && out) {
    SF_CHECK(hipMemcpyAsync(out, p->desc[0].out, take * sizeof(sf_signal_pixel),
                            hipMemcpyDeviceToHost, p->stream),
             "copying the signal pixels back");
    SF_CHECK(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
  }
  if (timing) {
    times[4] = std::chrono::duration<double, std::milli>(
                   std::chrono::steady_clock::now() - copy_started)
                   .count();
    times[5] = 0.0; // the list arrives ordered
    std::memcpy(g_stage_times, times, sizeof(times));
  }
  return SF_OK;
}

} // namespace

extern "C" {

int sf_hip_find_u16(const uint16_t *image, size_t height, size_t width,
                    sf_signal_pixel *out, size_t capacity, size_t *found) {
  return find<uint16_t>(image, height, width, out, capacity, found);
}

int sf_hip_find_u32(const uint32_t *image, size_t height, size_t width,
                    sf_signal_pixel *out, size_t capacity, size_t *found) {
  return find<uint32_t>(image, height, width, out, capacity, found);
}

int sf_hip_take_found(sf_signal_pixel *out, size_t n) {
  Workspace &w = workspace();
  if (!w.plan || !out)
    return bad("sf_hip_take_found: nothing found yet");
  if (n == 0)
    return SF_OK;
  SF_CHECK(hipMemcpyAsync(out, w.plan->desc[0].out,
                          n * sizeof(sf_signal_pixel), hipMemcpyDeviceToHost,
                          w.plan->stream),
           "copying the signal pixels back");
  SF_CHECK(hipStreamSynchronize(w.plan->stream), "hipStreamSynchronize");
  return SF_OK;
}

} // extern "C"

namespace sf {
int plan_lists(sf_plan *p, std::vector<const sf_signal_pixel *> &lists,
               std::vector<const uint32_t *> &counts,
               std::vector<uint32_t> &capacity, hipStream_t *stream) {
  if (!p)
    return SF_BAD_ARGUMENT;
  lists.resize(p->n);
  counts.resize(p->n);
  capacity.resize(p->n);
  for (int f = 0; f < p->n; f++) {
    lists[f] = p->desc[f].out;
    counts[f] = p->desc[f].out_count;
    capacity[f] = p->capacity[f];
  }
  *stream = p->stream;
  return SF_OK;
}

// For a reader of a plan whose last pipelined step left its ordering
// pending: the ordering's places (row prefixes, each survivor's place in
// the list, each frame's count) enqueued on `reader`, and per frame where
// the survivors sit (staging) and their places (dstof), so that the reader
// takes the survivors in list order without the lists being written; they
// are written only if sf_hip_plan_order is called. `taken` is false, and
// nothing is enqueued, when no ordering is pending (the lists are ordered).
int plan_take_places(sf_plan *p, hipStream_t reader,
                     std::vector<const sf_signal_pixel *> &staging,
                     std::vector<const uint32_t *> &places, bool &taken) {
  taken = false;
  if (!p)
    return SF_BAD_ARGUMENT;
  if (!p->order_pending)
    return SF_OK;
  SF_CHECK(sf::launch_sparse_places(p->d_desc, p->n, p->max_cell_rows,
                                    reader),
           "launching the places");
  staging.resize(p->n);
  places.resize(p->n);
  for (int f = 0; f < p->n; f++) {
    staging[f] = p->desc[f].staging;
    places[f] = p->desc[f].dstof;
  }
  p->order_pending = false;
  p->copy_pending = true;
  taken = true;
  return SF_OK;
}

// A reader of the plan's lists enqueued on `reader` (another stream): the
// plan's next write of its lists waits for what is enqueued there now.
int plan_add_reader(sf_plan *p, hipStream_t reader) {
  if (!p)
    return SF_BAD_ARGUMENT;
  SF_CHECK(hipEventRecord(p->read_done, reader), "hipEventRecord");
  p->reader_pending = true;
  return SF_OK;
}
} // namespace sf

namespace sf {
// The fused forms' layout of one frame (form 0: the sparse passes at 16
// bits, the sweep at 32): cell lattice or strips and bands, and the list
// capacity. Pointers (the sparse tables too), item offsets and the vector
// flag are the caller's to fill.
void fused_layout(int pixel_bytes, uint32_t height, uint32_t width,
                  FrameDesc &d) {
  d.height = height;
  d.width = width;
  d.capacity = (uint32_t)((size_t)height * width * (size_t)pixel_bytes / 16);
  if (pixel_bytes == 2) {
    sparse_layout(height, width, d);
    return;
  }
  // Short bands: a frame or two per launch still fills the chip (at 16
  // frames 108-row bands are 5% faster, at one frame 44-row bands 35%).
  const uint32_t ow = strip_width<uint32_t>();
  d.n_strips = (width + ow - 1) / ow;
  d.band_rows = 44;
  d.n_bands = (height + d.band_rows - 1) / d.band_rows;
}

// Cells of the per-row survivor table a frame needs, whatever the form.
size_t cell_entries(uint32_t height, uint32_t width) {
  return (size_t)height * ((width + 63) / 64);
}

// Zero the counters, threshold every frame of `d` (device descriptors), and
// lay each list out in index order: one batch on `stream`.
hipError_t launch_batch(int pixel_bytes, bool any_vec, bool any_scalar,
                        const FrameDesc *d, int n, uint32_t items,
                        uint32_t max_bands, uint32_t max_cell_rows,
                        hipStream_t stream) {
  if (pixel_bytes == 2)
    return launch_sparse(d, n, items, max_cell_rows, stream);
  hipError_t e = launch_zero(d, n, stream);
  if (e != hipSuccess)
    return e;
  e = launch_fused(pixel_bytes, any_vec, any_scalar, d, n, items, 0, stream);
  if (e != hipSuccess)
    return e;
  return launch_order(d, n, max_bands, stream);
}

void ensure_device_tables() { ensure_tables(); }
void set_last_error(const std::string &message) { g_last_error = message; }
} // namespace sf
