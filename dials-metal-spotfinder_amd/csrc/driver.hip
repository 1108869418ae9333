// The series driver on the device: src/find_spots.cc:460-713 (dispatch in
// chunks, collect in number order, group with z = number + z_offset, count
// what was never written, failed or came late, print dials.find_spots'
// summary, write the table), with the reference's worker threads replaced
// by batches: a chunk of frames is decoded in one launch, thresholded by a
// plan per run of consecutive frame numbers, and handed to the device
// labeller straight from the plan's lists. Host code only: it drives the
// library's own entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/driver_hip.h"

namespace {

thread_local std::string g_driver_error;

int fail(int status, std::string what) {
  g_driver_error = std::move(what);
  return status;
}

constexpr int kDefaultChunk = 64;

// A submitted frame. Its chunk's bytes are in the driver's pinned staging
// at [at, at + size), and on their way to the same place in the device
// staging.
struct Pending {
  int64_t number;
  bool written;
  std::string algorithm;
  size_t at = 0, size = 0;
};

std::string format(const char *fmt, ...) {
  va_list args;
  va_start(args, fmt);
  char small[512];
  const int n = vsnprintf(small, sizeof small, fmt, args);
  va_end(args);
  if (n < (int)sizeof small)
    return std::string(small, n > 0 ? n : 0);
  std::string big(n + 1, '\0');
  va_start(args, fmt);
  vsnprintf(&big[0], big.size(), fmt, args);
  va_end(args);
  big.resize(n);
  return big;
}

// "chunk f: message" from a batch of the decoder -> the message
std::string strip_chunk(const std::string &message) {
  if (message.rfind("chunk ", 0) == 0) {
    const size_t colon = message.find(": ");
    if (colon != std::string::npos)
      return message.substr(colon + 2);
  }
  return message;
}

} // namespace

struct sf_driver {
  unsigned bit_depth = 16;
  size_t height = 0, width = 0, frame_bytes = 0;
  int chunk = kDefaultChunk;
  int64_t z_offset = 0;
  sf_options grouping{};
  sf_decoder *dec = nullptr;
  sf_labeller *lab = nullptr;
  std::map<int, sf_plan *> plans; // by frame count
  uint8_t *d_frames = nullptr;    // `chunk` frame slots
  // The submitted chunks, 256-byte aligned, in pinned host memory and at the
  // same offsets in device memory: a submit copies its chunk in and starts
  // its copy to the device on the decoder's stream, so that the copies run
  // while later frames are submitted. Both are reused from the start once
  // no frame is waiting.
  uint8_t *h_stage = nullptr;
  uint8_t *d_stage = nullptr;
  size_t stage_cap = 0, stage_used = 0;
  std::vector<Pending> waiting;
  const uint32_t *mask_bits = nullptr; // the detector mask, device bits
  std::string log;
  sf_driver_stats stats{};
  int64_t highest = 0;
  bool any = false;
  bool finished = false;
  std::chrono::steady_clock::time_point began;
};

namespace {

void note(sf_driver *d, const std::string &line) { d->log += line + "\n"; }

// Pinned host memory takes long to allocate (its pages are locked), so the
// drivers of a process share the staging buffers they have finished with:
// a driver takes the smallest free one that is large enough, and gives it
// back when it grows or is destroyed. The pool keeps the two largest.
std::mutex g_pinned_mu;
std::vector<std::pair<uint8_t *, size_t>> g_pinned;

uint8_t *pinned_take(size_t need, size_t &cap) {
  {
    std::lock_guard<std::mutex> lock(g_pinned_mu);
    size_t best = g_pinned.size();
    for (size_t k = 0; k < g_pinned.size(); k++)
      if (g_pinned[k].second >= need &&
          (best == g_pinned.size() || g_pinned[k].second < g_pinned[best].second))
        best = k;
    if (best < g_pinned.size()) {
      uint8_t *p = g_pinned[best].first;
      cap = g_pinned[best].second;
      g_pinned.erase(g_pinned.begin() + best);
      return p;
    }
  }
  uint8_t *p = nullptr;
  if (hipHostMalloc(reinterpret_cast<void **>(&p), need) != hipSuccess)
    return nullptr;
  cap = need;
  return p;
}

void pinned_give(uint8_t *p, size_t cap) {
  if (!p)
    return;
  std::lock_guard<std::mutex> lock(g_pinned_mu);
  g_pinned.push_back({p, cap});
  if (g_pinned.size() > 2) {
    size_t small = 0;
    for (size_t k = 1; k < g_pinned.size(); k++)
      if (g_pinned[k].second < g_pinned[small].second)
        small = k;
    (void)hipHostFree(g_pinned[small].first);
    g_pinned.erase(g_pinned.begin() + small);
  }
}

hipStream_t decoder_stream(sf_driver *d) {
  return static_cast<hipStream_t>(sf_hip_decoder_stream(d->dec));
}

const uint8_t *chunk_bytes(const sf_driver *d, const Pending &p) {
  return p.size ? d->h_stage + p.at : reinterpret_cast<const uint8_t *>("");
}

// Room for `end` bytes of staging, `hint` bytes at least when it grows (a
// batch's worth of chunks like the one being staged). Copies to the device in
// flight read the old buffers, so the decoder's stream is drained first; the
// chunks staged so far move to the new buffers.
int stage_room(sf_driver *d, size_t end, size_t hint) {
  if (end <= d->stage_cap)
    return SF_OK;
  size_t cap = std::max({end, hint, 2 * d->stage_cap, (size_t)64 << 20});
  int status = sf_hip_decoder_wait(d->dec);
  if (status != SF_OK)
    return fail(status, sf_hip_decode_last_error());
  uint8_t *dev = nullptr;
  uint8_t *host = pinned_take(cap, cap);
  hipError_t e = host ? hipSuccess : hipErrorOutOfMemory;
  if (e == hipSuccess)
    e = hipMalloc(reinterpret_cast<void **>(&dev), cap);
  if (e == hipSuccess && d->stage_used) {
    std::memcpy(host, d->h_stage, d->stage_used);
    e = hipMemcpy(dev, d->d_stage, d->stage_used, hipMemcpyDeviceToDevice);
  }
  if (e != hipSuccess) {
    pinned_give(host, cap);
    if (dev)
      (void)hipFree(dev);
    return fail(SF_DEVICE_ERROR,
                std::string("staging chunks: ") + hipGetErrorString(e));
  }
  pinned_give(d->h_stage, d->stage_cap);
  if (d->d_stage)
    (void)hipFree(d->d_stage);
  d->h_stage = host;
  d->d_stage = dev;
  d->stage_cap = cap;
  return SF_OK;
}

// A submitted chunk into the staging, and its copy to the device started.
int stage_chunk(sf_driver *d, const uint8_t *data, size_t size, Pending &p) {
  p.at = (d->stage_used + 255) / 256 * 256;
  p.size = size;
  if (size == 0)
    return SF_OK;
  const size_t hint = ((size + 255) / 256 * 256) * ((size_t)d->chunk + 1);
  const int status = stage_room(d, p.at + size, hint);
  if (status != SF_OK)
    return status;
  // a large chunk is copied in by four threads
  uint8_t *to = d->h_stage + p.at;
  if (size >= ((size_t)2 << 20)) {
    const size_t part = (size / 4 + 4095) / 4096 * 4096;
    std::thread helpers[3];
    for (int t = 1; t < 4; t++) {
      const size_t a = std::min(size, t * part), b = std::min(size, a + part);
      helpers[t - 1] = std::thread([=] { std::memcpy(to + a, data + a, b - a); });
    }
    std::memcpy(to, data, std::min(size, part));
    for (auto &h : helpers)
      h.join();
  } else {
    std::memcpy(to, data, size);
  }
  const hipError_t e =
      hipMemcpyAsync(d->d_stage + p.at, d->h_stage + p.at, size,
                     hipMemcpyHostToDevice, decoder_stream(d));
  if (e != hipSuccess)
    return fail(SF_DEVICE_ERROR,
                std::string("hipMemcpyAsync: ") + hipGetErrorString(e));
  d->stage_used = p.at + size;
  return SF_OK;
}

void frame_failed(sf_driver *d, int64_t number, const std::string &message) {
  // src/find_spots.cc:528-533: "key: what()"
  note(d, format("frame %lld: %s", (long long)number, message.c_str()));
  d->stats.failed++;
}

sf_plan *plan_for(sf_driver *d, int n) {
  auto it = d->plans.find(n);
  if (it != d->plans.end())
    return it->second;
  std::vector<uint32_t> h(n, (uint32_t)d->height), w(n, (uint32_t)d->width);
  sf_plan *p = sf_hip_plan_create((int)(d->bit_depth / 8), n, h.data(),
                                  w.data(), 0);
  if (p)
    d->plans[n] = p;
  return p;
}

// Thresholds a run of frames with consecutive numbers (slots into d_frames)
// and groups them. A frame with more signal than the plan's list holds fails
// as in src/find_spots.cc:253-257 and splits the run.
int threshold_run(sf_driver *d, const std::vector<int> &slots,
                  const std::vector<int64_t> &numbers) {
  const int n = (int)slots.size();
  if (n == 0)
    return SF_OK;
  sf_plan *plan = plan_for(d, n);
  if (!plan)
    return fail(SF_DEVICE_ERROR,
                std::string("plan: ") + sf_hip_last_error());
  std::vector<const void *> ptrs(n);
  for (int k = 0; k < n; k++)
    ptrs[k] = d->d_frames + (size_t)slots[k] * d->frame_bytes;
  int status = sf_hip_plan_run(plan, ptrs.data());
  if (status == SF_OK)
    status = sf_hip_plan_wait(plan);
  if (status != SF_OK && status != SF_OVERFLOW)
    return fail(SF_DEVICE_ERROR,
                std::string("threshold: ") + sf_hip_last_error());
  std::vector<uint32_t> counts(n);
  sf_hip_plan_counts(plan, counts.data());
  std::vector<int> over;
  for (int k = 0; k < n; k++)
    if (counts[k] > sf_hip_plan_capacity(plan, k))
      over.push_back(k);
  if (!over.empty()) {
    std::vector<int> s;
    std::vector<int64_t> z;
    size_t next = 0;
    for (int k = 0; k <= n; k++) {
      const bool cut = k == n || (next < over.size() && over[next] == k);
      if (cut) {
        status = threshold_run(d, s, z);
        if (status != SF_OK)
          return status;
        s.clear();
        z.clear();
        if (k < n) {
          frame_failed(d, numbers[k],
                       "more signal pixels than the device buffer holds; a "
                       "frame that dense is not a frame of spots");
          next++;
        }
      } else {
        s.push_back(slots[k]);
        z.push_back(numbers[k]);
      }
    }
    return SF_OK;
  }
  d->stats.read += n;
  status = sf_hip_labeller_add_plan(d->lab, plan, numbers[0] + d->z_offset);
  if (status == SF_OK)
    status = sf_hip_labeller_wait(d->lab); // the plan is reused next
  if (status != SF_OK) {
    for (int k = 0; k < n; k++) // src/find_spots.cc:580-586
      note(d, format("grouping frame %lld: %s", (long long)numbers[k],
                     sf_hip_labeller_last_error()));
    d->stats.failed += n;
    return SF_OK;
  }
  d->highest = numbers[n - 1] + d->z_offset;
  d->any = true;
  d->stats.grouped += n;
  return SF_OK;
}

// Decodes the written frames of one algorithm into their slots; frames that
// fail are logged, and their slots dropped from `good`.
int decode_group(sf_driver *d, std::vector<const Pending *> &frames,
                 std::vector<int> &slots, const std::string &algorithm) {
  for (;;) {
    if (frames.empty())
      return SF_OK;
    std::vector<const uint8_t *> host(frames.size());
    std::vector<size_t> sizes(frames.size());
    for (size_t f = 0; f < frames.size(); f++) {
      host[f] = chunk_bytes(d, *frames[f]);
      sizes[f] = frames[f]->size;
    }
    const int status = sf_hip_decoder_index(d->dec, host.data(), sizes.data(),
                                            (int)frames.size(),
                                            algorithm.c_str());
    if (status == SF_OK)
      break;
    if (status == SF_DEVICE_ERROR)
      return fail(status, sf_hip_decode_last_error());
    // the walk stops at the first bad chunk: log it, go on without it
    const std::string message = sf_hip_decode_last_error();
    size_t bad = 0;
    if (frames.size() > 1 && message.rfind("chunk ", 0) == 0)
      bad = (size_t)std::atoll(message.c_str() + 6);
    if (bad >= frames.size())
      bad = 0;
    frame_failed(d, frames[bad]->number, strip_chunk(message));
    frames.erase(frames.begin() + bad);
    slots.erase(slots.begin() + bad);
  }
  // the chunks are in the device staging (copied on the decoder's stream
  // when submitted, so the decode follows their copies)
  std::vector<const uint8_t *> chunks(frames.size());
  std::vector<void *> out(frames.size());
  for (size_t f = 0; f < frames.size(); f++) {
    chunks[f] = d->d_stage ? d->d_stage + frames[f]->at : nullptr;
    out[f] = d->d_frames + (size_t)slots[f] * d->frame_bytes;
  }
  int status = sf_hip_decoder_decode(d->dec, chunks.data(), out.data());
  if (status == SF_OK)
    status = sf_hip_decoder_wait(d->dec);
  if (status == SF_BAD_DATA) {
    // a block that does not decode: find which, one chunk at a time
    std::vector<const Pending *> kept;
    std::vector<int> kept_slots;
    for (size_t f = 0; f < frames.size(); f++) {
      const int one =
          sf_hip_decoder_image(d->dec, chunk_bytes(d, *frames[f]),
                               frames[f]->size, algorithm.c_str(), out[f]);
      if (one == SF_OK) {
        kept.push_back(frames[f]);
        kept_slots.push_back(slots[f]);
      } else if (one == SF_BAD_DATA) {
        frame_failed(d, frames[f]->number, sf_hip_decode_last_error());
      } else {
        return fail(one, sf_hip_decode_last_error());
      }
    }
    frames.swap(kept);
    slots.swap(kept_slots);
    return SF_OK;
  }
  if (status != SF_OK)
    return fail(status, sf_hip_decode_last_error());
  return SF_OK;
}

// One batch: the first n waiting frames in number order
// (src/find_spots.cc:550-600).
int run_batch(sf_driver *d, size_t n) {
  std::stable_sort(d->waiting.begin(), d->waiting.end(),
                   [](const Pending &a, const Pending &b) {
                     return a.number < b.number;
                   });
  n = std::min(n, d->waiting.size());
  std::vector<Pending> batch(std::make_move_iterator(d->waiting.begin()),
                             std::make_move_iterator(d->waiting.begin() + n));
  d->waiting.erase(d->waiting.begin(), d->waiting.begin() + n);
  // decode by algorithm, each written frame in a slot of its own
  std::map<std::string, std::pair<std::vector<const Pending *>,
                                  std::vector<int>>> groups;
  for (size_t k = 0; k < batch.size(); k++) {
    if (!batch[k].written) {
      d->stats.missing++;
      continue;
    }
    groups[batch[k].algorithm].first.push_back(&batch[k]);
    groups[batch[k].algorithm].second.push_back((int)k);
  }
  std::vector<std::pair<int64_t, int>> decoded; // number, slot
  for (auto &g : groups) {
    int status = decode_group(d, g.second.first, g.second.second, g.first);
    if (status != SF_OK)
      return status;
    for (size_t f = 0; f < g.second.first.size(); f++)
      decoded.push_back({g.second.first[f]->number, g.second.second[f]});
  }
  // the batch's chunks are decoded (the decoder waited for its stream):
  // with no frame waiting, the staging starts over
  if (d->waiting.empty())
    d->stage_used = 0;
  std::sort(decoded.begin(), decoded.end());
  if (d->mask_bits && !decoded.empty()) {
    // the detector mask into the decoded frames, before the threshold
    std::vector<void *> ptrs;
    for (auto &f : decoded)
      ptrs.push_back(d->d_frames + (size_t)f.second * d->frame_bytes);
    // on the decoder's stream, which wrote the frames; waiting for that
    // stream alone leaves every other stream of the process running
    int status = sf_hip_stamp_mask((int)(d->bit_depth / 8), ptrs.data(),
                                   (int)ptrs.size(), d->height, d->width,
                                   d->mask_bits,
                                   sf_hip_decoder_stream(d->dec));
    if (status != SF_OK)
      return fail(SF_DEVICE_ERROR,
                  std::string("mask: ") + sf_hip_mask_last_error());
    status = sf_hip_decoder_wait(d->dec);
    if (status != SF_OK)
      return fail(SF_DEVICE_ERROR,
                  std::string("mask: ") + sf_hip_decode_last_error());
  }
  if (d->bit_depth != 16 && d->bit_depth != 32) {
    for (auto &f : decoded) // src/find_spots.cc:515-519
      frame_failed(d, f.first,
                   format("the threshold does not support %u-bit data",
                          d->bit_depth));
    return SF_OK;
  }
  // late frames: not above the last grouped z (src/find_spots.cc:571-578)
  std::vector<int> slots;
  std::vector<int64_t> numbers;
  auto flush = [&]() {
    const int status = threshold_run(d, slots, numbers);
    slots.clear();
    numbers.clear();
    return status;
  };
  for (auto &f : decoded) {
    // not above the last grouped frame, or the last of the run being built
    // (a number that came twice)
    const int64_t z = f.first + d->z_offset;
    const bool late = !numbers.empty() ? z <= numbers.back() + d->z_offset
                                       : d->any && z <= d->highest;
    if (late) {
      d->stats.read++;
      d->stats.out_of_order++;
      continue;
    }
    if (!numbers.empty() && f.first != numbers.back() + 1) {
      const int status = flush();
      if (status != SF_OK)
        return status;
    }
    slots.push_back(f.second);
    numbers.push_back(f.first);
  }
  return flush();
}

} // namespace

extern "C" sf_driver *sf_hip_driver_create(unsigned bit_depth, size_t height,
                                           size_t width,
                                           const sf_driver_options *options) {
  if (height == 0 || width == 0 || (bit_depth != 8 && bit_depth != 16 &&
                                    bit_depth != 32 && bit_depth != 64)) {
    fail(SF_BAD_ARGUMENT, "sf_hip_driver_create: bad argument");
    return nullptr;
  }
  auto *d = new sf_driver;
  d->bit_depth = bit_depth;
  d->height = height;
  d->width = width;
  d->frame_bytes = height * width * bit_depth / 8;
  d->grouping = sf_options{3, 1000, 2.0, 0}; // dials.find_spots defaults
  if (options) {
    d->chunk = options->chunk > 0 ? options->chunk : kDefaultChunk;
    d->z_offset = options->z_offset;
    d->grouping = options->grouping;
  }
  d->lab = sf_hip_labeller_create(height, width, &d->grouping);
  if (!d->lab) {
    fail(SF_BAD_ARGUMENT, std::string("sf_hip_driver_create: ") +
                              sf_hip_labeller_last_error());
    sf_hip_driver_destroy(d);
    return nullptr;
  }
  int status = sf_hip_decoder_create(bit_depth, height, width, d->chunk,
                                     &d->dec);
  if (status != SF_OK) {
    fail(status, std::string("sf_hip_driver_create: ") +
                     sf_hip_decode_last_error());
    sf_hip_driver_destroy(d);
    return nullptr;
  }
  const hipError_t e = hipMalloc(reinterpret_cast<void **>(&d->d_frames),
                                 (size_t)d->chunk * d->frame_bytes);
  if (e != hipSuccess) {
    fail(SF_DEVICE_ERROR,
         std::string("sf_hip_driver_create: ") + hipGetErrorString(e));
    sf_hip_driver_destroy(d);
    return nullptr;
  }
  d->began = std::chrono::steady_clock::now();
  return d;
}

extern "C" int sf_hip_driver_submit(sf_driver *d, int64_t number,
                                    const uint8_t *data, size_t size,
                                    const char *algorithm) {
  if (!d || (data && !algorithm))
    return fail(SF_BAD_ARGUMENT, "sf_hip_driver_submit: bad argument");
  if (d->finished)
    return fail(SF_BAD_ARGUMENT, "sf_hip_driver_submit: after finish");
  Pending p;
  p.number = number;
  p.written = data != nullptr;
  if (data) {
    p.algorithm = algorithm;
    const int status = stage_chunk(d, data, size, p);
    if (status != SF_OK)
      return status;
  }
  d->waiting.push_back(std::move(p));
  d->stats.dispatched++;
  if (d->waiting.size() >= (size_t)d->chunk)
    return run_batch(d, (size_t)d->chunk);
  return SF_OK;
}

extern "C" int sf_hip_driver_set_mask(sf_driver *d, const uint32_t *bits,
                                      size_t height, size_t width) {
  if (!d)
    return fail(SF_BAD_ARGUMENT, "sf_hip_driver_set_mask: bad argument");
  if (d->bit_depth != 16 && d->bit_depth != 32)
    return fail(SF_BAD_ARGUMENT,
                "sf_hip_driver_set_mask: 16- or 32-bit frames only");
  // the stamp reads ceil(height * width / 32) words: a mask of another
  // shape would be read past its end, or mask the wrong pixels
  if (bits && (height != d->height || width != d->width))
    return fail(SF_BAD_ARGUMENT,
                format("sf_hip_driver_set_mask: a %zu x %zu mask for %zu x "
                       "%zu frames",
                       height, width, d->height, d->width));
  d->mask_bits = bits;
  return SF_OK;
}

extern "C" int sf_hip_driver_finish(sf_driver *d) {
  if (!d)
    return fail(SF_BAD_ARGUMENT, "sf_hip_driver_finish: bad argument");
  if (d->finished)
    return SF_OK;
  while (!d->waiting.empty()) {
    const int status = run_batch(d, (size_t)d->chunk);
    if (status != SF_OK)
      return status;
  }
  int status = sf_hip_labeller_finish(d->lab);
  if (status != SF_OK)
    return fail(status, sf_hip_labeller_last_error());
  d->finished = true;
  const double seconds = std::chrono::duration<double>(
                             std::chrono::steady_clock::now() - d->began)
                             .count();
  const sf_driver_stats &s = d->stats;
  // src/find_spots.cc:642-682, wording included
  note(d, format("Thresholded %llu of %llu images in %.1f s (%.1f images/s), "
                 "%llu never written, %llu failures",
                 (unsigned long long)s.read, (unsigned long long)s.dispatched,
                 seconds, seconds > 0 ? s.read / seconds : 0.0,
                 (unsigned long long)s.missing, (unsigned long long)s.failed));
  if (s.out_of_order > 0)
    note(d, format("warning: %llu frames arrived after a later one and were "
                   "dropped",
                   (unsigned long long)s.out_of_order));
  sf_counts c{};
  status = sf_hip_labeller_results(d->lab, nullptr, nullptr, &c);
  if (status != SF_OK)
    return fail(status, sf_hip_labeller_last_error());
  const unsigned long long sized = c.groups - c.too_small - c.too_large;
  note(d, format("Found %llu signal pixels on %llu frames",
                 (unsigned long long)c.signal_pixels,
                 (unsigned long long)s.grouped));
  note(d, format("Extracted %llu spots", (unsigned long long)c.groups));
  note(d, format("Removed %llu spots with size < %zu pixels",
                 (unsigned long long)c.too_small,
                 (size_t)d->grouping.min_spot_size));
  note(d, format("Removed %llu spots with size > %zu pixels",
                 (unsigned long long)c.too_large,
                 (size_t)d->grouping.max_spot_size));
  note(d, format("Calculated %llu spot centroids", sized));
  note(d, format("Calculated %llu spot intensities", sized));
  if (d->grouping.max_separation > 0.0)
    note(d, format("Filtered %llu of %llu spots by peak-centroid distance",
                   (unsigned long long)c.accepted, sized));
  return SF_OK;
}

extern "C" int sf_hip_driver_write(sf_driver *d, const char *path,
                                   const sf_refl_options *options) {
  if (!d || !path)
    return fail(SF_BAD_ARGUMENT, "sf_hip_driver_write: bad argument");
  if (!d->finished)
    return fail(SF_BAD_ARGUMENT, "sf_hip_driver_write: before finish");
  size_t n_spots = 0, n_pixels = 0;
  int status = sf_hip_labeller_sizes(d->lab, &n_spots, &n_pixels);
  if (status != SF_OK)
    return fail(status, sf_hip_labeller_last_error());
  std::vector<sf_spot> spots(n_spots);
  std::vector<sf_pixel> pixels(n_pixels);
  status = sf_hip_labeller_results(d->lab, spots.data(), pixels.data(),
                                   nullptr);
  if (status != SF_OK)
    return fail(status, sf_hip_labeller_last_error());
  sf_refl_options defaults{0, 0, 1, ""};
  const sf_refl_options &o = options ? *options : defaults;
  status = sf_hip_refl_write(path, spots.data(), n_spots, pixels.data(),
                             n_pixels, d->width, &o);
  if (status != SF_OK) {
    note(d, sf_hip_refl_last_error());
    return fail(status, sf_hip_refl_last_error());
  }
  if (n_spots == 0) {
    note(d, format("No reflections found; %s is a well formed table with no "
                   "rows in it, which dials.index will refuse",
                   path));
  } else {
    double bytes = 0;
    if (FILE *f = std::fopen(path, "rb")) {
      std::fseek(f, 0, SEEK_END);
      bytes = (double)std::ftell(f);
      std::fclose(f);
    }
    note(d, format("Wrote %zu reflections to %s (%.1f MB%s)", n_spots, path,
                   bytes / 1e6, o.shoeboxes ? ", most of it shoeboxes" : ""));
  }
  return SF_OK;
}

extern "C" size_t sf_hip_driver_log(sf_driver *d, char *text, size_t cap) {
  if (!d)
    return 0;
  if (text && cap) {
    const size_t n = std::min(cap - 1, d->log.size());
    memcpy(text, d->log.data(), n);
    text[n] = '\0';
  }
  return d->log.size();
}

extern "C" int sf_hip_driver_stats(sf_driver *d, sf_driver_stats *stats) {
  if (!d || !stats)
    return fail(SF_BAD_ARGUMENT, "sf_hip_driver_stats: bad argument");
  *stats = d->stats;
  return SF_OK;
}

extern "C" sf_labeller *sf_hip_driver_labeller(sf_driver *d) {
  return d ? d->lab : nullptr;
}

extern "C" void sf_hip_driver_destroy(sf_driver *d) {
  if (!d)
    return;
  for (auto &p : d->plans)
    sf_hip_plan_destroy(p.second);
  if (d->dec) {
    (void)sf_hip_decoder_wait(d->dec); // copies from the staging in flight
    sf_hip_decoder_destroy(d->dec);
  }
  if (d->lab)
    sf_hip_labeller_destroy(d->lab);
  if (d->d_frames)
    (void)hipFree(d->d_frames);
  if (d->d_stage)
    (void)hipFree(d->d_stage);
  pinned_give(d->h_stage, d->stage_cap);
  delete d;
}

extern "C" const char *sf_hip_driver_last_error(void) {
  return g_driver_error.c_str();
}
