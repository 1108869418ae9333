// The serving loop: frames of several concurrent series thresholded,
// ordered and grouped as they arrive, with continuous batching.
//
// The reference has no counterpart; its driver (src/find_spots.cc:560-640)
// runs one series through a thread pool and one Labeller. Here:
//
//   submit    a caller hands over a device-resident frame of an open series
//             (frame numbers ascending per series) and returns at once;
//   admit     a frame needs a slot: the device buffers of one frame's list,
//             staging and cell table, sized by its series' shape. Slots come
//             from a per-series free list, or are allocated while the bytes
//             in use stay within the HBM budget the server was created
//             with; a frame that finds neither waits (its series' later
//             frames wait behind it) until a slot is released;
//   batch     the scheduler thread gathers every admitted frame of one pixel
//             type, whatever its series or shape, into one launch (zero
//             counters, threshold, ordering) on one of a few batch streams,
//             taking the pixel types in turn, oldest waiting frame first, so
//             batches overlap on the device and frames arriving while others
//             run join the next launch;
//   watch     a watcher thread waits on each batch's event in launch order
//             and stamps frame-to-mask latency the moment it completes;
//   label     a dispatcher hands each series' frames of the batch to that
//             series' own labelling thread, which adds them to the series'
//             device labeller (its union-find frontier stays on the device)
//             and stamps chunk-to-table latency when the labeller is done;
//             the last of them releases the batch and its slots.
//
// Every frame is computed by the same kernels as a plan, and each series has
// its own labeller, so a series' results are identical alone and in a mix.
#include "sf_device.hh"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <iterator>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace sf {
void fused_layout(int pixel_bytes, uint32_t height, uint32_t width,
                  FrameDesc &d);
size_t cell_entries(uint32_t height, uint32_t width);
hipError_t launch_batch(int pixel_bytes, bool any_vec, bool any_scalar,
                        const FrameDesc *d, int n, uint32_t items,
                        uint32_t max_bands, uint32_t max_cell_rows,
                        hipStream_t stream);
size_t sparse_bytes(uint32_t height, uint32_t width, uint32_t capacity);
void sparse_carve(uint8_t *base, FrameDesc &d);
void ensure_device_tables();
int labeller_add_lists(sf_labeller *l, int n, const int64_t *numbers,
                       const sf_signal_pixel *const *lists,
                       const uint32_t *const *count_ptrs,
                       const uint32_t *capacities, const uint32_t *counts,
                       hipStream_t producer);
hipStream_t labeller_stream(sf_labeller *l);
} // namespace sf

namespace {

using Clock = std::chrono::steady_clock;
thread_local std::string g_serve_error;

constexpr int kBatchStreams = 8;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Series;

struct Slot {
  Series *series = nullptr;
  void *base = nullptr;
  void *frame = nullptr; // a host-fed series' frame, copied in per batch
  sf::FrameDesc desc{}; // buffers filled; pixels, counters, items per batch
};

struct Request {
  int series = 0;
  int64_t frame = 0;
  const void *pixels = nullptr;
  Clock::time_point arrival;
};

struct Batch;

// One series' frames of one batch, for that series' labelling worker.
struct Job {
  Batch *batch = nullptr;
  int series = 0;
  std::vector<int> frames; // indices into the batch, in frame order
};

struct Series {
  int id = 0;
  int pixel_bytes = 2;
  uint32_t height = 0, width = 0;
  bool host = false; // frames arrive in (pinned) host memory
  sf_labeller *labeller = nullptr;
  size_t slot_bytes = 0;
  std::vector<Slot *> slots, free_slots;
  int64_t last_submitted = 0;
  bool any_submitted = false;
  bool closed = false;
  // per completed frame, in frame order
  std::vector<int64_t> frames;
  std::vector<uint32_t> counts;
  std::vector<double> mask_ms, table_ms;
  std::vector<std::vector<sf_signal_pixel>> lists;
  // the labelling worker
  std::thread worker;
  std::deque<Job> jobs;
  std::condition_variable cv;
  bool stop = false;
};

struct Batch {
  int pixel_bytes = 2;
  int capacity = 0;
  std::vector<Request> requests;
  std::vector<Slot *> slots;
  // descriptors and found counts pass through host-mapped memory, moved by
  // one-block kernels: no copy of a batch waits in the copy engines behind
  // another batch's frames
  sf::FrameDesc *h_desc = nullptr, *d_desc = nullptr, *hd_desc = nullptr;
  uint32_t *d_counters = nullptr, *d_found = nullptr, *h_found = nullptr,
           *hd_found = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t masked = nullptr, released = nullptr;
  Clock::time_point masked_at; // when the watcher saw `masked` complete
  hipError_t mask_error = hipSuccess;
  uint64_t ingested = 0;        // bytes copied in from host memory
  int jobs_left = 0;            // labelling jobs still reading the lists
  std::vector<std::vector<sf_signal_pixel>> lists; // keep_lists copies
};

} // namespace

struct sf_server {
  size_t budget = 0, used = 0, peak = 0;
  int max_batch = 16;
  bool keep_lists = false;
  std::mutex m;
  std::condition_variable cv_sched, cv_done, cv_label, cv_drain;
  std::deque<Request> pending;
  bool dirty = false;
  std::vector<Series *> series;
  std::vector<Batch *> batches, free_batches;
  std::deque<Batch *> inflight; // launched, list not yet ready
  std::deque<Batch *> to_label; // lists ready, waiting for the labellers
  std::thread scheduler, watcher, dispatcher;
  bool stop = false, scheduler_done = false, watcher_done = false;
  uint64_t submitted = 0, completed = 0;
  int status = SF_OK;
  std::string error;
  // statistics
  uint64_t n_batches = 0, batched_frames = 0, admission_waits = 0;
  uint64_t overflowed = 0; // frames whose list overflowed (labelled empty)
  uint64_t ingested = 0;   // bytes of host-fed frames copied to the device
  int max_inflight = 0;
};

namespace {

void set_failure(sf_server *s, int status, const std::string &what) {
  // called with s->m held
  if (s->status == SF_OK) {
    s->status = status;
    s->error = what;
  }
}

std::string hip_message(const char *what, hipError_t e) {
  return std::string(what) + ": " + hipGetErrorString(e);
}

void free_slot(Slot *slot) {
  if (slot->base)
    (void)hipFree(slot->base);
  delete slot;
}

// Bytes of a 16-bit frame's sparse tables (0 at 32 bits).
size_t sparse_part(int pixel_bytes, const sf::FrameDesc &d) {
  return pixel_bytes == 2 ? sf::sparse_bytes(d.height, d.width, d.capacity)
                          : 0;
}

// One frame's buffers for a series: staging and ordered list (capacity
// entries each), cell counts and offsets, band totals, and at 16 bits the
// sparse form's tables.
hipError_t make_slot(Series *se, Slot **out) {
  Slot *slot = new Slot;
  slot->series = se;
  sf::FrameDesc &d = slot->desc;
  std::memset(&d, 0, sizeof(d));
  sf::fused_layout(se->pixel_bytes, se->height, se->width, d);
  const size_t cap = std::max<size_t>(d.capacity, 1);
  const size_t cells = sf::cell_entries(se->height, se->width);
  size_t at = 0;
  const size_t staging_at = at;
  at += cap * sizeof(sf_signal_pixel);
  const size_t out_at = at;
  at += cap * sizeof(sf_signal_pixel);
  const size_t base_at = at;
  at += cells * sizeof(uint32_t);
  const size_t band_at = at;
  at += ((size_t)d.n_bands + 1) * sizeof(uint32_t);
  const size_t count_at = at;
  at += cells;
  at = (at + 255) & ~(size_t)255;
  const size_t sparse_at = at;
  at += sparse_part(se->pixel_bytes, d);
  at = (at + 255) & ~(size_t)255;
  const size_t frame_at = at;
  if (se->host)
    at += (size_t)se->pixel_bytes * se->height * se->width;
  const hipError_t e = hipMalloc(&slot->base, at);
  if (e != hipSuccess) {
    delete slot;
    return e;
  }
  char *b = static_cast<char *>(slot->base);
  d.staging = reinterpret_cast<sf_signal_pixel *>(b + staging_at);
  d.out = reinterpret_cast<sf_signal_pixel *>(b + out_at);
  d.cell_base = reinterpret_cast<uint32_t *>(b + base_at);
  d.band_total = reinterpret_cast<uint32_t *>(b + band_at);
  d.cell_count = reinterpret_cast<uint8_t *>(b + count_at);
  if (se->pixel_bytes == 2)
    sf::sparse_carve(reinterpret_cast<uint8_t *>(b + sparse_at), d);
  if (se->host)
    slot->frame = b + frame_at;
  *out = slot;
  return hipSuccess;
}

size_t slot_bytes(int pixel_bytes, uint32_t height, uint32_t width,
                  bool host) {
  sf::FrameDesc d{};
  sf::fused_layout(pixel_bytes, height, width, d);
  const size_t cap = std::max<size_t>(d.capacity, 1);
  const size_t cells = sf::cell_entries(height, width);
  return 2 * cap * sizeof(sf_signal_pixel) + cells * sizeof(uint32_t) +
         (((size_t)d.n_bands + 1) * sizeof(uint32_t)) +
         cells + 255 + sparse_part(pixel_bytes, d) + 255 +
         (host ? (size_t)pixel_bytes * height * width : 0);
}

hipError_t make_batch(int capacity, Batch **out) {
  Batch *b = new Batch;
  b->capacity = capacity;
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess)
    e = hipMalloc(&b->d_desc, sizeof(sf::FrameDesc) * capacity);
  if (e == hipSuccess)
    e = hipHostMalloc(&b->h_desc, sizeof(sf::FrameDesc) * capacity,
                      hipHostMallocMapped);
  if (e == hipSuccess)
    e = hipHostGetDevicePointer(reinterpret_cast<void **>(&b->hd_desc),
                                b->h_desc, 0);
  if (e == hipSuccess)
    e = hipMalloc(&b->d_counters, sizeof(uint32_t) * 2 * capacity);
  if (e == hipSuccess) {
    b->d_found = b->d_counters + capacity;
    e = hipHostMalloc(&b->h_found, sizeof(uint32_t) * capacity,
                      hipHostMallocMapped);
  }
  if (e == hipSuccess)
    e = hipHostGetDevicePointer(reinterpret_cast<void **>(&b->hd_found),
                                b->h_found, 0);
  if (e == hipSuccess)
    e = hipEventCreateWithFlags(&b->masked, hipEventDisableTiming);
  if (e == hipSuccess)
    e = hipEventCreateWithFlags(&b->released, hipEventDisableTiming);
  *out = b;
  return e;
}

void free_batch(Batch *b) {
  if (b->stream)
    (void)hipStreamSynchronize(b->stream);
  if (b->masked)
    (void)hipEventDestroy(b->masked);
  if (b->released)
    (void)hipEventDestroy(b->released);
  (void)hipFree(b->d_desc);
  (void)hipFree(b->d_counters);
  if (b->h_desc)
    (void)hipHostFree(b->h_desc);
  if (b->h_found)
    (void)hipHostFree(b->h_found);
  if (b->stream)
    (void)hipStreamDestroy(b->stream);
  delete b;
}

// n words from `from` to `to` by one block (either side host-mapped).
__global__ __launch_bounds__(256) void move_words(const uint32_t *from,
                                                  uint32_t *to, uint32_t n) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
    to[i] = from[i];
  if (threadIdx.x == 0)
    __threadfence_system();
}

// Enqueue a formed batch on its stream (no lock held).
hipError_t launch(Batch *b) {
  const int n = (int)b->requests.size();
  uint32_t items = 0, max_bands = 0, max_cell_rows = 0;
  bool any_vec = false, any_scalar = false;
  const uint32_t k = b->pixel_bytes == 2 ? 4 : 2;
  // host-fed frames first come in, on the batch's own stream
  b->ingested = 0;
  for (int f = 0; f < n; f++) {
    const Slot *slot = b->slots[f];
    if (!slot->series->host)
      continue;
    const size_t bytes = (size_t)slot->series->pixel_bytes *
                         slot->series->height * slot->series->width;
    const hipError_t e =
        hipMemcpyAsync(slot->frame, b->requests[f].pixels, bytes,
                       hipMemcpyHostToDevice, b->stream);
    if (e != hipSuccess)
      return e;
    b->ingested += bytes;
  }
  for (int f = 0; f < n; f++) {
    sf::FrameDesc d = b->slots[f]->desc;
    d.pixels = b->slots[f]->series->host ? b->slots[f]->frame
                                         : b->requests[f].pixels;
    d.counter = b->d_counters + f;
    d.out_count = b->d_found + f;
    d.item_begin = items;
    items += d.n_strips * d.n_bands;
    max_bands = std::max(max_bands, d.n_bands);
    max_cell_rows = std::max(max_cell_rows, d.n_cell_rows);
    // u16 frames go to the sparse scout (16-byte row loads from an even
    // column: 4-byte aligned frames of even width), u32 frames to the sweep
    // (K pixels a lane: 8-byte aligned, K | width)
    const uintptr_t at = reinterpret_cast<uintptr_t>(d.pixels);
    d.vec = b->pixel_bytes == 2
                ? (at % 4 == 0 && d.width % 2 == 0 ? 1u : 0u)
                : (at % 8 == 0 && d.width % k == 0 ? 1u : 0u);
    any_vec |= d.vec != 0;
    any_scalar |= d.vec == 0;
    b->h_desc[f] = d;
  }
  static_assert(sizeof(sf::FrameDesc) % 4 == 0, "descriptors move as words");
  move_words<<<1, 256, 0, b->stream>>>(
      reinterpret_cast<const uint32_t *>(b->hd_desc),
      reinterpret_cast<uint32_t *>(b->d_desc),
      (uint32_t)(sizeof(sf::FrameDesc) / 4 * n));
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = sf::launch_batch(b->pixel_bytes, any_vec, any_scalar, b->d_desc, n,
                         items, max_bands, max_cell_rows, b->stream);
  if (e == hipSuccess) {
    move_words<<<1, 256, 0, b->stream>>>(b->d_found, b->hd_found,
                                         (uint32_t)n);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipEventRecord(b->masked, b->stream);
  return e;
}

// Give idle slots of other series back to the budget until `need` bytes fit
// (a series that stopped sending must not starve one that still sends).
// Called with the lock held.
void reclaim(sf_server *s, Series *keep, size_t need) {
  for (Series *other : s->series) {
    if (other == keep)
      continue;
    while (!other->free_slots.empty() && s->used + need > s->budget) {
      Slot *slot = other->free_slots.back();
      other->free_slots.pop_back();
      other->slots.erase(
          std::find(other->slots.begin(), other->slots.end(), slot));
      free_slot(slot);
      s->used -= other->slot_bytes;
    }
    if (s->used + need <= s->budget)
      return;
  }
}

// Form one batch of `pixel_bytes` frames from the pending queue, in arrival
// order, skipping every later frame of a series that could not be admitted.
// Called with the lock held; returns null when nothing can be admitted.
Batch *form_batch(sf_server *s, int pixel_bytes) {
  if (s->free_batches.empty()) {
    if ((int)s->batches.size() >= kBatchStreams)
      return nullptr;
    Batch *b = nullptr;
    const hipError_t e = make_batch(s->max_batch, &b);
    if (e != hipSuccess) {
      if (b)
        free_batch(b);
      set_failure(s, SF_DEVICE_ERROR, hip_message("creating a batch", e));
      return nullptr;
    }
    s->batches.push_back(b);
    s->free_batches.push_back(b);
  }
  Batch *b = s->free_batches.back();
  b->pixel_bytes = pixel_bytes;
  b->requests.clear();
  b->slots.clear();
  std::vector<int> blocked;
  std::deque<Request> rest;
  while (!s->pending.empty()) {
    Request r = s->pending.front();
    s->pending.pop_front();
    Series *se = s->series[r.series];
    const bool skip =
        se->pixel_bytes != pixel_bytes ||
        (int)b->requests.size() >= s->max_batch ||
        std::find(blocked.begin(), blocked.end(), r.series) != blocked.end();
    if (skip) {
      rest.push_back(r);
      continue;
    }
    Slot *slot = nullptr;
    if (se->free_slots.empty() && s->used + se->slot_bytes > s->budget)
      reclaim(s, se, se->slot_bytes);
    if (!se->free_slots.empty()) {
      slot = se->free_slots.back();
      se->free_slots.pop_back();
    } else if (s->used + se->slot_bytes <= s->budget) {
      const hipError_t e = make_slot(se, &slot);
      if (e != hipSuccess) {
        set_failure(s, SF_DEVICE_ERROR, hip_message("allocating a slot", e));
        rest.push_back(r);
        blocked.push_back(r.series);
        continue;
      }
      se->slots.push_back(slot);
      s->used += se->slot_bytes;
      s->peak = std::max(s->peak, s->used);
    } else {
      s->admission_waits++;
      rest.push_back(r);
      blocked.push_back(r.series);
      continue;
    }
    b->requests.push_back(r);
    b->slots.push_back(slot);
  }
  s->pending.swap(rest);
  if (b->requests.empty())
    return nullptr;
  s->free_batches.pop_back();
  return b;
}

void scheduler_loop(sf_server *s) {
  std::unique_lock<std::mutex> lk(s->m);
  for (;;) {
    s->cv_sched.wait(lk, [s] {
      return s->stop || (!s->pending.empty() && s->dirty);
    });
    if (s->stop && (s->pending.empty() || s->status != SF_OK))
      break;
    s->dirty = false;
    if (s->status != SF_OK) {
      // drop what cannot run; drain() reports the failure
      s->completed += s->pending.size();
      s->pending.clear();
      s->cv_drain.notify_all();
      continue;
    }
    // One batch per pixel type in turn, the type of the oldest waiting
    // frame first, so that neither type can take every batch stream.
    const int first =
        s->pending.empty() ? 2 : s->series[s->pending.front().series]->pixel_bytes;
    for (bool progress = true; progress && !s->pending.empty();) {
      progress = false;
      for (int pixel_bytes : {first, 6 - first}) {
        Batch *b = form_batch(s, pixel_bytes);
        if (!b)
          continue;
        progress = true;
        lk.unlock();
        const hipError_t e = launch(b);
        lk.lock();
        if (e != hipSuccess)
          set_failure(s, SF_DEVICE_ERROR, hip_message("launching a batch", e));
        s->n_batches++;
        s->batched_frames += b->requests.size();
        s->ingested += b->ingested;
        s->inflight.push_back(b);
        s->max_inflight = std::max(
            s->max_inflight, (int)(s->batches.size() - s->free_batches.size()));
        s->cv_done.notify_one();
      }
    }
  }
  s->scheduler_done = true;
  s->cv_done.notify_all();
}

// Stamps each batch's frame-to-mask time as its event completes, in launch
// order, on a thread of its own so that labelling never delays the stamp.
void watch_loop(sf_server *s) {
  std::unique_lock<std::mutex> lk(s->m);
  for (;;) {
    s->cv_done.wait(lk, [s] {
      return !s->inflight.empty() || s->scheduler_done;
    });
    if (s->inflight.empty() && s->scheduler_done)
      break;
    Batch *b = s->inflight.front();
    s->inflight.pop_front();
    lk.unlock();
    b->mask_error = hipEventSynchronize(b->masked);
    b->masked_at = Clock::now();
    lk.lock();
    s->to_label.push_back(b);
    s->cv_label.notify_one();
  }
  s->watcher_done = true;
  s->cv_label.notify_all();
}

// A batch's lists are ready: copy them out if asked, then hand each series'
// frames of the batch to that series' labelling worker. The batch is
// released by the last worker to finish with it.
void release_batch(sf_server *s, Batch *b) {
  // called with s->m held
  for (Slot *slot : b->slots)
    slot->series->free_slots.push_back(slot);
  s->completed += b->requests.size();
  b->requests.clear();
  b->slots.clear();
  s->free_batches.push_back(b);
  s->dirty = true;
  s->cv_sched.notify_one();
  s->cv_drain.notify_all();
}

void dispatch(sf_server *s, Batch *b) {
  const int n = (int)b->requests.size();
  int status = SF_OK;
  std::string what;
  {
    std::lock_guard<std::mutex> lk(s->m);
    if (s->status != SF_OK)
      status = s->status;
  }
  if (status == SF_OK && b->mask_error != hipSuccess) {
    status = SF_DEVICE_ERROR;
    what = hip_message("a batch failed", b->mask_error);
  }
  b->lists.assign(s->keep_lists && status == SF_OK ? n : 0, {});
  for (int f = 0; f < (int)b->lists.size() && status == SF_OK; f++) {
    const uint32_t c = std::min(b->h_found[f], b->h_desc[f].capacity);
    b->lists[f].resize(c);
    const hipError_t e =
        c ? hipMemcpy(b->lists[f].data(), b->h_desc[f].out,
                      c * sizeof(sf_signal_pixel), hipMemcpyDeviceToHost)
          : hipSuccess;
    if (e != hipSuccess) {
      status = SF_DEVICE_ERROR;
      what = hip_message("reading a list", e);
    }
  }
  std::lock_guard<std::mutex> lk(s->m);
  if (status != SF_OK) {
    if (!what.empty())
      set_failure(s, status, what);
    (void)hipStreamSynchronize(b->stream);
    release_batch(s, b);
    return;
  }
  // one job per series, its frames in batch order (which is frame order)
  std::vector<Job> jobs;
  for (int f = 0; f < n; f++) {
    const int sid = b->requests[f].series;
    auto it = std::find_if(jobs.begin(), jobs.end(),
                           [sid](const Job &j) { return j.series == sid; });
    if (it == jobs.end()) {
      jobs.push_back(Job{b, sid, {}});
      it = jobs.end() - 1;
    }
    it->frames.push_back(f);
  }
  b->jobs_left = (int)jobs.size();
  for (Job &job : jobs) {
    Series *se = s->series[job.series];
    se->jobs.push_back(std::move(job));
    se->cv.notify_one();
  }
}

void dispatch_loop(sf_server *s) {
  std::unique_lock<std::mutex> lk(s->m);
  for (;;) {
    s->cv_label.wait(lk, [s] {
      return !s->to_label.empty() || s->watcher_done;
    });
    if (s->to_label.empty() && s->watcher_done)
      break;
    Batch *b = s->to_label.front();
    s->to_label.pop_front();
    lk.unlock();
    dispatch(s, b);
    lk.lock();
  }
}

// Label a series' queued jobs -- its frames of one or more batches, oldest
// first -- as one chunk on the series' own labeller. Every list is complete
// (the watcher saw its batch's event) and stays untouched until the job's
// batch is released after this returns, so the gather needs no producer
// stream.
int label_jobs(Series *se, const std::vector<Job> &jobs,
               Clock::time_point *tabled) {
  std::vector<int64_t> numbers;
  std::vector<const sf_signal_pixel *> ptrs;
  std::vector<const uint32_t *> count_ptrs;
  std::vector<uint32_t> caps, counts;
  for (const Job &job : jobs) {
    const Batch *b = job.batch;
    for (int f : job.frames) {
      numbers.push_back(b->requests[f].frame);
      ptrs.push_back(b->h_desc[f].out);
      count_ptrs.push_back(b->d_found + f);
      caps.push_back(b->h_desc[f].capacity);
      counts.push_back(b->h_found[f]);
    }
  }
  int status = sf::labeller_add_lists(
      se->labeller, (int)numbers.size(), numbers.data(), ptrs.data(),
      count_ptrs.data(), caps.data(), counts.data(), nullptr);
  if (status != SF_OK) {
    g_serve_error = std::string("labelling: ") + sf_hip_labeller_last_error();
    return status;
  }
  const hipError_t e = hipStreamSynchronize(sf::labeller_stream(se->labeller));
  *tabled = Clock::now();
  if (e != hipSuccess) {
    g_serve_error = hip_message("labelling", e);
    return SF_DEVICE_ERROR;
  }
  return SF_OK;
}

void series_loop(sf_server *s, Series *se) {
  std::unique_lock<std::mutex> lk(s->m);
  for (;;) {
    se->cv.wait(lk, [se] { return !se->jobs.empty() || se->stop; });
    if (se->jobs.empty() && se->stop)
      break;
    // everything queued for this series so far becomes one chunk
    std::vector<Job> jobs(std::make_move_iterator(se->jobs.begin()),
                          std::make_move_iterator(se->jobs.end()));
    se->jobs.clear();
    const bool failed = s->status != SF_OK;
    lk.unlock();
    Clock::time_point tabled = Clock::now();
    int status = SF_OK;
    if (!failed)
      status = label_jobs(se, jobs, &tabled);
    lk.lock();
    if (status != SF_OK)
      set_failure(s, status, g_serve_error);
    for (Job &job : jobs) {
      Batch *b = job.batch;
      if (!failed && status == SF_OK) {
        for (int f : job.frames) {
          const Request &r = b->requests[f];
          // more signal pixels than the slot's list holds (the -2 of
          // src/dext_gpu.hh:148-151): the labeller took the frame as empty
          if (b->h_found[f] > b->h_desc[f].capacity)
            s->overflowed++;
          se->frames.push_back(r.frame);
          se->counts.push_back(b->h_found[f]);
          se->mask_ms.push_back(ms_between(r.arrival, b->masked_at));
          se->table_ms.push_back(ms_between(r.arrival, tabled));
          if (s->keep_lists)
            se->lists.push_back(std::move(b->lists[f]));
        }
      }
      if (--b->jobs_left == 0)
        release_batch(s, b);
    }
  }
}

int fail_arg(const char *what) {
  g_serve_error = what;
  return SF_BAD_ARGUMENT;
}

Series *series_of(sf_server *s, int id) {
  if (!s || id < 0 || id >= (int)s->series.size())
    return nullptr;
  return s->series[id];
}

int drain_locked(sf_server *s, std::unique_lock<std::mutex> &lk) {
  s->cv_drain.wait(lk, [s] { return s->completed >= s->submitted; });
  if (s->status != SF_OK) {
    g_serve_error = s->error;
    return s->status;
  }
  return SF_OK;
}

} // namespace

extern "C" {

sf_server *sf_hip_server_create(size_t hbm_budget, int max_batch_frames,
                                int keep_lists) {
  if (hbm_budget == 0 || max_batch_frames <= 0 || max_batch_frames > 4096) {
    fail_arg("sf_hip_server_create: bad argument");
    return nullptr;
  }
  sf::ensure_device_tables();
  sf_server *s = new sf_server;
  s->budget = hbm_budget;
  s->max_batch = max_batch_frames;
  s->keep_lists = keep_lists != 0;
  s->scheduler = std::thread(scheduler_loop, s);
  s->watcher = std::thread(watch_loop, s);
  s->dispatcher = std::thread(dispatch_loop, s);
  return s;
}

void sf_hip_server_destroy(sf_server *s) {
  if (!s)
    return;
  {
    std::unique_lock<std::mutex> lk(s->m);
    (void)drain_locked(s, lk);
    s->stop = true;
    s->cv_sched.notify_all();
  }
  s->scheduler.join();
  s->watcher.join();
  s->dispatcher.join();
  for (Series *se : s->series) {
    {
      std::lock_guard<std::mutex> lk(s->m);
      se->stop = true;
      se->cv.notify_all();
    }
    se->worker.join();
  }
  for (Batch *b : s->batches)
    free_batch(b);
  for (Series *se : s->series) {
    for (Slot *slot : se->slots)
      free_slot(slot);
    sf_hip_labeller_destroy(se->labeller);
    delete se;
  }
  delete s;
}

} // extern "C"

namespace {

int open_series(sf_server *s, int pixel_bytes, size_t height, size_t width,
                const sf_options *options, bool host) {
  if (!s || (pixel_bytes != 2 && pixel_bytes != 4) || height == 0 ||
      width == 0 || height > 0x7fffffffu / width)
    return fail_arg("sf_hip_server_open: bad argument");
  const size_t bytes =
      slot_bytes(pixel_bytes, (uint32_t)height, (uint32_t)width, host);
  if (bytes > s->budget)
    return fail_arg("sf_hip_server_open: one frame's buffers exceed the "
                    "HBM budget");
  sf_labeller *lab = sf_hip_labeller_create(height, width, options);
  if (!lab) {
    g_serve_error = std::string("sf_hip_server_open: ") +
                    sf_hip_labeller_last_error();
    return SF_BAD_ARGUMENT;
  }
  Series *se = new Series;
  se->pixel_bytes = pixel_bytes;
  se->height = (uint32_t)height;
  se->width = (uint32_t)width;
  se->host = host;
  se->labeller = lab;
  se->slot_bytes = bytes;
  std::lock_guard<std::mutex> lk(s->m);
  se->id = (int)s->series.size();
  s->series.push_back(se);
  se->worker = std::thread(series_loop, s, se);
  return se->id;
}

} // namespace

extern "C" {

int sf_hip_server_open(sf_server *s, int pixel_bytes, size_t height,
                       size_t width, const sf_options *options) {
  return open_series(s, pixel_bytes, height, width, options, false);
}

int sf_hip_server_open_host(sf_server *s, int pixel_bytes, size_t height,
                            size_t width, const sf_options *options) {
  return open_series(s, pixel_bytes, height, width, options, true);
}

int sf_hip_server_submit_at(sf_server *s, int series, int64_t frame,
                            const void *device_pixels, double arrival_s) {
  if (!s || !device_pixels)
    return fail_arg("sf_hip_server_submit: bad argument");
  std::lock_guard<std::mutex> lk(s->m);
  Series *se = series_of(s, series);
  if (!se || se->closed)
    return fail_arg("sf_hip_server_submit: no such open series");
  if (s->status != SF_OK) {
    g_serve_error = s->error;
    return s->status;
  }
  if (se->any_submitted && frame <= se->last_submitted) {
    g_serve_error = "frames must be submitted in ascending order";
    return SF_ORDER_ERROR;
  }
  se->any_submitted = true;
  se->last_submitted = frame;
  Request r;
  r.series = series;
  r.frame = frame;
  r.pixels = device_pixels;
  r.arrival = arrival_s < 0.0
                  ? Clock::now()
                  : Clock::time_point(std::chrono::duration_cast<
                                      Clock::duration>(
                        std::chrono::duration<double>(arrival_s)));
  s->pending.push_back(r);
  s->submitted++;
  s->dirty = true;
  s->cv_sched.notify_one();
  return SF_OK;
}

int sf_hip_server_submit(sf_server *s, int series, int64_t frame,
                         const void *device_pixels) {
  return sf_hip_server_submit_at(s, series, frame, device_pixels, -1.0);
}

int sf_hip_server_drain(sf_server *s) {
  if (!s)
    return fail_arg("sf_hip_server_drain: no server");
  std::unique_lock<std::mutex> lk(s->m);
  return drain_locked(s, lk);
}

int sf_hip_server_close(sf_server *s, int series) {
  if (!s)
    return fail_arg("sf_hip_server_close: no server");
  std::unique_lock<std::mutex> lk(s->m);
  Series *se = series_of(s, series);
  if (!se)
    return fail_arg("sf_hip_server_close: no such series");
  const int status = drain_locked(s, lk);
  if (status != SF_OK)
    return status;
  if (se->closed)
    return SF_OK;
  se->closed = true; // no more submissions; its worker is idle
  lk.unlock();
  const int done = sf_hip_labeller_finish(se->labeller);
  if (done != SF_OK)
    g_serve_error = sf_hip_labeller_last_error();
  return done;
}

sf_labeller *sf_hip_server_labeller(sf_server *s, int series) {
  if (!s)
    return nullptr;
  std::lock_guard<std::mutex> lk(s->m);
  Series *se = series_of(s, series);
  return se ? se->labeller : nullptr;
}

double sf_hip_server_clock(void) {
  return std::chrono::duration<double>(Clock::now().time_since_epoch())
      .count();
}

int sf_hip_server_frames(sf_server *s, int series, int64_t *frames,
                         uint32_t *counts, double *mask_ms, double *table_ms,
                         size_t capacity, size_t *n) {
  if (!s || !n)
    return fail_arg("sf_hip_server_frames: bad argument");
  std::lock_guard<std::mutex> lk(s->m);
  Series *se = series_of(s, series);
  if (!se)
    return fail_arg("sf_hip_server_frames: no such series");
  *n = se->frames.size();
  const size_t take = std::min(capacity, se->frames.size());
  for (size_t k = 0; k < take; k++) {
    if (frames)
      frames[k] = se->frames[k];
    if (counts)
      counts[k] = se->counts[k];
    if (mask_ms)
      mask_ms[k] = se->mask_ms[k];
    if (table_ms)
      table_ms[k] = se->table_ms[k];
  }
  return SF_OK;
}

int sf_hip_server_frame_list(sf_server *s, int series, size_t k,
                             sf_signal_pixel *out, size_t capacity,
                             size_t *n) {
  if (!s || !n)
    return fail_arg("sf_hip_server_frame_list: bad argument");
  std::lock_guard<std::mutex> lk(s->m);
  Series *se = series_of(s, series);
  if (!se || !s->keep_lists || k >= se->lists.size())
    return fail_arg("sf_hip_server_frame_list: no such list kept");
  const auto &list = se->lists[k];
  *n = list.size();
  const size_t take = std::min(capacity, list.size());
  if (take && out)
    std::memcpy(out, list.data(), take * sizeof(sf_signal_pixel));
  return SF_OK;
}

int sf_hip_server_stats(sf_server *s, sf_server_stats *stats) {
  if (!s || !stats)
    return fail_arg("sf_hip_server_stats: bad argument");
  std::lock_guard<std::mutex> lk(s->m);
  stats->batches = s->n_batches;
  stats->batched_frames = s->batched_frames;
  stats->admission_waits = s->admission_waits;
  stats->max_batches_in_flight = (uint64_t)s->max_inflight;
  stats->bytes_in_use = s->used;
  stats->peak_bytes = s->peak;
  stats->budget = s->budget;
  stats->overflowed = s->overflowed;
  stats->ingested_bytes = s->ingested;
  return SF_OK;
}

size_t sf_hip_server_slot_bytes(int pixel_bytes, size_t height,
                                size_t width) {
  if ((pixel_bytes != 2 && pixel_bytes != 4) || !height || !width ||
      height > 0x7fffffffu / width)
    return 0;
  return slot_bytes(pixel_bytes, (uint32_t)height, (uint32_t)width, false);
}

int sf_hip_server_play(sf_server *s, int n_series, const int *series,
                       const double *rate_hz, const int64_t *n_frames,
                       const void *const *rings, const int *ring_lengths,
                       int64_t first_frame) {
  if (!s || n_series <= 0 || !series || !rate_hz || !n_frames || !rings ||
      !ring_lengths)
    return fail_arg("sf_hip_server_play: bad argument");
  std::vector<size_t> ring_at(n_series);
  size_t at = 0;
  for (int i = 0; i < n_series; i++) {
    if (rate_hz[i] <= 0.0 || ring_lengths[i] <= 0 || n_frames[i] < 0)
      return fail_arg("sf_hip_server_play: bad rate, ring or count");
    ring_at[i] = at;
    at += (size_t)ring_lengths[i];
  }
  std::vector<int64_t> next(n_series, 0);
  const Clock::time_point t0 = Clock::now() + std::chrono::milliseconds(2);
  for (;;) {
    // the earliest due frame over all series
    int pick = -1;
    double due = 0.0;
    for (int i = 0; i < n_series; i++) {
      if (next[i] >= n_frames[i])
        continue;
      const double t = (double)next[i] / rate_hz[i];
      if (pick < 0 || t < due) {
        pick = i;
        due = t;
      }
    }
    if (pick < 0)
      break;
    const Clock::time_point when =
        t0 + std::chrono::duration_cast<Clock::duration>(
                 std::chrono::duration<double>(due));
    std::this_thread::sleep_until(when);
    const int64_t f = next[pick]++;
    const void *px = rings[ring_at[pick] + (size_t)(f % ring_lengths[pick])];
    const double arrival =
        std::chrono::duration<double>(when.time_since_epoch()).count();
    const int status = sf_hip_server_submit_at(s, series[pick],
                                               first_frame + f, px, arrival);
    if (status != SF_OK)
      return status;
  }
  return sf_hip_server_drain(s);
}

const char *sf_hip_server_last_error(void) { return g_serve_error.c_str(); }

} // extern "C"
