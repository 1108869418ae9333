// C entry points over the reference's own CPU path -- TEST INFRASTRUCTURE.
//
// Compiled together with the reference's src/dext.cc, src/dials_spots.cc,
// src/signal_order.cc and src/refl.cc where they lie (oracle/Makefile); the result goes to
// oracle/_ref/ and is never committed. Tests use it to pin the oracle and the
// device path against the real reference, and bench.py times it as the CPU
// baseline. Everything here is glue: the arithmetic is the reference's.
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

#include "dext.hh"
#include "dials_spots.hh"
#include "refl.hh"
#include "signal_order.hh"

static_assert(sizeof(dials_spots::Spot) == 104, "Spot layout");
static_assert(sizeof(dials_spots::Pixel) == 12, "Pixel layout");

namespace {

template <typename T>
int threshold(const T *image, std::size_t height, std::size_t width,
              SignalPixel *out, std::size_t capacity, std::size_t *found) {
  static thread_local dext_scratch<T> scratch;
  static thread_local std::vector<SignalPixel> signal;
  const int status = dext<T>(image, signal, height, width, scratch);
  *found = signal.size();
  const std::size_t n = signal.size() < capacity ? signal.size() : capacity;
  if (n)
    std::memcpy(out, signal.data(), n * sizeof(SignalPixel));
  return status;
}

template <typename T>
long long threshold_many(const T *frames, std::size_t count,
                         std::size_t height, std::size_t width, int threads) {
  std::atomic<std::size_t> next{0};
  std::atomic<long long> total{0};
  std::atomic<int> failed{0};
  const std::size_t n = height * width;
  auto work = [&] {
    dext_scratch<T> scratch;
    std::vector<SignalPixel> signal;
    for (;;) {
      const std::size_t f = next.fetch_add(1);
      if (f >= count)
        break;
      if (dext<T>(frames + f * n, signal, height, width, scratch) != 0)
        failed = 1;
      total += static_cast<long long>(signal.size());
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < threads; t++)
    pool.emplace_back(work);
  work();
  for (auto &t : pool)
    t.join();
  return failed ? -1 : total.load();
}

// `jobs` thresholds over `count` frames (frame j % count), on `threads`
// threads, each with its own scratch as find_spots' workers hold it
// (src/find_spots.cc:254): the CPU throughput at any thread count from a
// bounded sample of frames.
template <typename T>
long long threshold_cycle(const T *frames, std::size_t count,
                          std::size_t height, std::size_t width,
                          std::size_t jobs, int threads) {
  std::atomic<std::size_t> next{0};
  std::atomic<long long> total{0};
  std::atomic<int> failed{0};
  const std::size_t n = height * width;
  auto work = [&] {
    dext_scratch<T> scratch;
    std::vector<SignalPixel> signal;
    for (;;) {
      const std::size_t j = next.fetch_add(1);
      if (j >= jobs)
        break;
      if (dext<T>(frames + (j % count) * n, signal, height, width,
                  scratch) != 0)
        failed = 1;
      total += static_cast<long long>(signal.size());
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < threads; t++)
    pool.emplace_back(work);
  work();
  for (auto &t : pool)
    t.join();
  return failed ? -1 : total.load();
}

struct RefLabeller {
  dials_spots::Labeller labeller;
  std::vector<SignalPixel> buffer;
  RefLabeller(std::size_t h, std::size_t w, const dials_spots::Options &o)
      : labeller(h, w, o) {}
};

} // namespace

extern "C" {

int ref_dext_u16(const std::uint16_t *image, std::size_t height,
                 std::size_t width, SignalPixel *out, std::size_t capacity,
                 std::size_t *found) {
  return threshold<std::uint16_t>(image, height, width, out, capacity, found);
}

int ref_dext_u32(const std::uint32_t *image, std::size_t height,
                 std::size_t width, SignalPixel *out, std::size_t capacity,
                 std::size_t *found) {
  return threshold<std::uint32_t>(image, height, width, out, capacity, found);
}

long long ref_dext_many_u16(const std::uint16_t *frames, std::size_t count,
                            std::size_t height, std::size_t width,
                            int threads) {
  return threshold_many<std::uint16_t>(frames, count, height, width, threads);
}

long long ref_dext_many_u32(const std::uint32_t *frames, std::size_t count,
                            std::size_t height, std::size_t width,
                            int threads) {
  return threshold_many<std::uint32_t>(frames, count, height, width, threads);
}

void ref_by_index(SignalPixel *pixels, std::size_t n) {
  std::vector<SignalPixel> v(pixels, pixels + n);
  signal_order::by_index(v);
  if (n)
    std::memcpy(pixels, v.data(), n * sizeof(SignalPixel));
}

void *ref_labeller_new(std::size_t height, std::size_t width,
                       std::size_t min_spot_size, std::size_t max_spot_size,
                       double max_separation, int two_d) {
  dials_spots::Options options;
  options.min_spot_size = min_spot_size;
  options.max_spot_size = max_spot_size;
  options.max_separation = max_separation;
  options.two_d = two_d != 0;
  try {
    return new RefLabeller(height, width, options);
  } catch (const std::exception &) {
    return nullptr;
  }
}

void ref_labeller_free(void *handle) {
  delete static_cast<RefLabeller *>(handle);
}

int ref_labeller_add(void *handle, std::int64_t frame,
                     const SignalPixel *signal, std::size_t n) {
  RefLabeller *l = static_cast<RefLabeller *>(handle);
  l->buffer.assign(signal, signal + n);
  try {
    l->labeller.add(frame, l->buffer);
  } catch (const std::exception &) {
    return -1;
  }
  return 0;
}

int ref_labeller_finish(void *handle) {
  static_cast<RefLabeller *>(handle)->labeller.finish();
  return 0;
}

std::size_t ref_labeller_live(void *handle) {
  return static_cast<RefLabeller *>(handle)->labeller.live();
}

std::size_t ref_labeller_spot_count(void *handle) {
  return static_cast<RefLabeller *>(handle)->labeller.spots().size();
}

std::size_t ref_labeller_pixel_count(void *handle) {
  return static_cast<RefLabeller *>(handle)->labeller.pixels().size();
}

void ref_labeller_results(void *handle, dials_spots::Spot *spots,
                          dials_spots::Pixel *pixels,
                          dials_spots::Counts *counts) {
  const dials_spots::Labeller &l =
      static_cast<RefLabeller *>(handle)->labeller;
  if (spots && !l.spots().empty())
    std::memcpy(spots, l.spots().data(),
                l.spots().size() * sizeof(dials_spots::Spot));
  if (pixels && !l.pixels().empty())
    std::memcpy(pixels, l.pixels().data(),
                l.pixels().size() * sizeof(dials_spots::Pixel));
  if (counts)
    *counts = l.counts();
}

long long ref_dext_cycle_u16(const std::uint16_t *frames, std::size_t count,
                             std::size_t height, std::size_t width,
                             std::size_t jobs, int threads) {
  return threshold_cycle<std::uint16_t>(frames, count, height, width, jobs,
                                        threads);
}

long long ref_dext_cycle_u32(const std::uint32_t *frames, std::size_t count,
                             std::size_t height, std::size_t width,
                             std::size_t jobs, int threads) {
  return threshold_cycle<std::uint32_t>(frames, count, height, width, jobs,
                                        threads);
}

// The whole CPU path of dials.find_spots over a stack of u16 frames, as
// src/find_spots.cc:460-713 runs it: `threads` workers threshold frames
// (each with its own scratch), the calling thread groups them in frame
// order with the default options as they complete, then finishes and
// writes the table with shoeboxes to `path`. Returns 0, -1 if a threshold
// failed, -5 if the write threw; the signal pixel, spot and pixel counts
// come back through the pointers.
int ref_pipeline_u16(const std::uint16_t *frames, std::size_t count,
                     std::size_t height, std::size_t width, int threads,
                     const char *path, std::size_t *n_signal,
                     std::size_t *n_spots, std::size_t *n_pixels) {
  const std::size_t n = height * width;
  std::vector<std::vector<SignalPixel>> lists(count);
  std::vector<char> done(count, 0);
  std::mutex m;
  std::condition_variable cv;
  std::atomic<std::size_t> next{0};
  std::atomic<int> failed{0};
  auto work = [&] {
    dext_scratch<std::uint16_t> scratch;
    for (;;) {
      const std::size_t f = next.fetch_add(1);
      if (f >= count)
        break;
      std::vector<SignalPixel> signal;
      if (dext<std::uint16_t>(frames + f * n, signal, height, width,
                              scratch) != 0)
        failed = 1;
      {
        std::lock_guard<std::mutex> lock(m);
        lists[f].swap(signal);
        done[f] = 1;
      }
      cv.notify_all();
    }
  };
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; t++)
    pool.emplace_back(work);
  dials_spots::Labeller labeller(height, width, dials_spots::Options{});
  std::size_t signal_total = 0;
  for (std::size_t f = 0; f < count; f++) {
    std::vector<SignalPixel> signal;
    {
      std::unique_lock<std::mutex> lock(m);
      cv.wait(lock, [&] { return done[f] != 0; });
      signal.swap(lists[f]);
    }
    signal_total += signal.size();
    labeller.add(static_cast<std::int64_t>(f), signal);
  }
  for (auto &t : pool)
    t.join();
  labeller.finish();
  *n_signal = signal_total;
  *n_spots = labeller.spots().size();
  *n_pixels = labeller.pixels().size();
  if (failed)
    return -1;
  try {
    refl::Options options;
    refl::write(path, labeller.spots(), labeller.pixels(), width, options);
  } catch (const std::exception &) {
    return -5;
  }
  return 0;
}

// refl::write over C tables: 0, or -5 where it throws.
int ref_refl_write(const char *path, const dials_spots::Spot *spots,
                   std::size_t n_spots, const dials_spots::Pixel *pixels,
                   std::size_t n_pixels, std::size_t width, std::size_t panel,
                   int id, const char *identifier, int shoeboxes) {
  try {
    refl::Options options;
    options.panel = panel;
    options.id = id;
    options.identifier = identifier ? identifier : "";
    options.shoeboxes = shoeboxes != 0;
    refl::write(path, std::vector<dials_spots::Spot>(spots, spots + n_spots),
                std::vector<dials_spots::Pixel>(pixels, pixels + n_pixels),
                width, options);
    return 0;
  } catch (const std::exception &) {
    return -5;
  }
}

} // extern "C"
