// The reference's device interface (src/dext_gpu.hh, namespace gpu), re-made
// over the C ABI of libdext_hip.so so that a caller written against it --
// src/find_spots.cc's threshold<T>() and -gpu check, tests/test_dext_gpu.cc,
// tests/bench_dext_gpu.cc -- compiles and links unchanged.
//
// Header-only: nothing here includes a HIP header. Link with -ldext_hip.
//
// Semantics follow src/dext_gpu.hh:
//   find<T>      clears `signal_out` and refills it ascending by index;
//                returns 0, -1 (null image, zero size, height*width beyond
//                INT32_MAX) or -2 (more signal pixels than the device list
//                holds, height*width*sizeof(T)/16); throws
//                std::runtime_error on a device failure. No host sort: the
//                list comes back ordered.
//   windows      Tile/Tile (the default here) runs the fused single-pass
//                form, any Direct the plain three-kernel form; both give
//                identical results. SPOTFINDER_GPU_STAGE0/STAGE2 override
//                the defaults as in src/dext_gpu.cc:54-82.
//   StageTimes   upload, stage0 = threshold kernel, stage1 (plain form's
//                erode), stage2 = ordering (and the plain form's last stage),
//                copy, sort (always 0).
#ifndef SPOTFINDER_DEXT_GPU_HH
#define SPOTFINDER_DEXT_GPU_HH

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "dext_hip.h"

#if __has_include("signal_pixel.hh")
#include "signal_pixel.hh"
#else
// src/signal_pixel.hh:20-28: the record both sides of the boundary share.
struct SignalPixel {
  std::uint32_t index;
  std::uint32_t value;
  float background;
  std::uint16_t population;
  std::uint16_t reserved;
};
#endif

static_assert(sizeof(SignalPixel) == sizeof(sf_signal_pixel),
              "SignalPixel and sf_signal_pixel must share their layout");

namespace gpu {

inline const char *backend() { return sf_hip_backend(); }

enum class Window { Direct, Tile };

inline const char *name(Window window) {
  return window == Window::Tile ? "tile" : "direct";
}

inline Window stage0_window() {
  return sf_hip_stage0_window() ? Window::Tile : Window::Direct;
}
inline void stage0_window(Window window) {
  sf_hip_set_stage0_window(window == Window::Tile ? 1 : 0);
}
inline Window stage2_window() {
  return sf_hip_stage2_window() ? Window::Tile : Window::Direct;
}
inline void stage2_window(Window window) {
  sf_hip_set_stage2_window(window == Window::Tile ? 1 : 0);
}

struct StageTimes {
  double upload = 0.0;
  double stage0 = 0.0;
  double stage1 = 0.0;
  double stage2 = 0.0;
  double copy = 0.0;
  double sort = 0.0;
};

inline void profile_stages(bool on) { sf_hip_profile_stages(on ? 1 : 0); }
inline bool profile_stages() { return sf_hip_profiling() != 0; }

inline StageTimes last_stage_times() {
  double t[6];
  sf_hip_last_stage_times(t);
  StageTimes times;
  times.upload = t[0];
  times.stage0 = t[1];
  times.stage1 = t[2];
  times.stage2 = t[3];
  times.copy = t[4];
  times.sort = t[5];
  return times;
}

inline void report_windows(bool on) { sf_hip_report_windows(on ? 1 : 0); }
inline bool available() { return sf_hip_available() != 0; }
inline std::size_t memory_free() { return sf_hip_memory_free(); }
inline void *host_alloc(std::size_t bytes) { return sf_hip_host_alloc(bytes); }
inline void host_free(void *pointer) { sf_hip_host_free(pointer); }

namespace detail {
template <typename T> struct entry;
template <> struct entry<std::uint16_t> {
  static int call(const std::uint16_t *image, std::size_t h, std::size_t w,
                  sf_signal_pixel *out, std::size_t capacity,
                  std::size_t *found) {
    return sf_hip_find_u16(image, h, w, out, capacity, found);
  }
};
template <> struct entry<std::uint32_t> {
  static int call(const std::uint32_t *image, std::size_t h, std::size_t w,
                  sf_signal_pixel *out, std::size_t capacity,
                  std::size_t *found) {
    return sf_hip_find_u32(image, h, w, out, capacity, found);
  }
};
} // namespace detail

template <typename T>
int find(const T *image_in, std::vector<SignalPixel> &signal_out,
         std::size_t height, std::size_t width) {
  signal_out.clear();
  // Fill into whatever capacity the vector kept from earlier frames; fetch
  // the rest from the device only when this frame found more.
  const std::size_t capacity = signal_out.capacity();
  signal_out.resize(capacity);
  std::size_t found = 0;
  const int status = detail::entry<T>::call(
      image_in, height, width,
      reinterpret_cast<sf_signal_pixel *>(signal_out.data()), capacity,
      &found);
  if (status == SF_DEVICE_ERROR) {
    signal_out.clear();
    throw std::runtime_error(std::string("gpu::find: ") + sf_hip_last_error());
  }
  if (status != SF_OK) {
    signal_out.clear();
    return status;
  }
  signal_out.resize(found);
  if (found > capacity) {
    if (sf_hip_take_found(reinterpret_cast<sf_signal_pixel *>(
                              signal_out.data()),
                          found) != SF_OK) {
      signal_out.clear();
      throw std::runtime_error(std::string("gpu::find: ") +
                               sf_hip_last_error());
    }
  }
  return 0;
}

} // namespace gpu

#endif // SPOTFINDER_DEXT_GPU_HH
