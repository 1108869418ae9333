// A gpu:: backend for the reference's own source tree, over libdext_hip.so.
//
// This translation unit takes the place of src/dext_cuda.cu: it defines what
// that file defines (src/dext_cuda.cu:546-735) -- backend(), available(),
// memory_free(), host_alloc(), host_free(), find<uint16_t>, find<uint32_t>
// and the backend's default windows -- against the reference's unchanged
// src/dext_gpu.hh and src/dext_gpu_internal.hh. The window policy, the
// SPOTFINDER_GPU_STAGE0/STAGE2 overrides, the announcement and the per-thread
// stage times stay in the reference's src/dext_gpu.cc. Compile it with
// -I <reference>/src -I <this repository>/include and link -ldext_hip.
//
// Both windows name the same computation here: Tile/Tile runs the fused
// single-pass kernel, any Direct the plain three-kernel form. Results are
// identical either way, so the windows only choose a speed.
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "dext_gpu.hh"
#include "dext_gpu_internal.hh"
#include "dext_hip.h"

static_assert(sizeof(SignalPixel) == sizeof(sf_signal_pixel),
              "SignalPixel and sf_signal_pixel must share their layout");

namespace gpu {

namespace {

int entry(const std::uint16_t *image, std::size_t height, std::size_t width,
          sf_signal_pixel *out, std::size_t capacity, std::size_t *found) {
  return sf_hip_find_u16(image, height, width, out, capacity, found);
}

int entry(const std::uint32_t *image, std::size_t height, std::size_t width,
          sf_signal_pixel *out, std::size_t capacity, std::size_t *found) {
  return sf_hip_find_u32(image, height, width, out, capacity, found);
}

[[noreturn]] void fail(const char *what) {
  throw std::runtime_error(std::string(what) + ": " + sf_hip_last_error());
}

} // namespace

const char *backend() { return sf_hip_backend(); }

namespace internal {

// The fused kernel is the fast form on this device, by a factor of about ten
// over the plain one (measured at 4362 x 4148 u16), so Tile for both.
Window default_stage0_window() { return Window::Tile; }
Window default_stage2_window() { return Window::Tile; }

} // namespace internal

bool available() { return sf_hip_available() != 0; }

std::size_t memory_free() { return sf_hip_memory_free(); }

void *host_alloc(std::size_t bytes) { return sf_hip_host_alloc(bytes); }

void host_free(void *pointer) { sf_hip_host_free(pointer); }

template <typename T>
int find(const T *image_in, std::vector<SignalPixel> &signal_out,
         std::size_t height, std::size_t width) {
  signal_out.clear();
  if (image_in == nullptr || height == 0 || width == 0)
    return -1;
  if (height > 0x7fffffffu / width)
    return -1;

  // src/dext_gpu.cc announces the windows; the library stays quiet.
  static const bool quiet = (sf_hip_report_windows(0), true);
  (void)quiet;
  const Window window0 = stage0_window();
  const Window window2 = stage2_window();
  internal::announce_windows(window0, window2);
  sf_hip_set_stage0_window(window0 == Window::Tile ? 1 : 0);
  sf_hip_set_stage2_window(window2 == Window::Tile ? 1 : 0);
  const bool timing = profile_stages();
  sf_hip_profile_stages(timing ? 1 : 0);

  // Fill whatever capacity the vector kept from earlier frames; fetch the
  // rest only when this frame found more. The list arrives ordered by index,
  // so there is no host sort.
  const std::size_t capacity = signal_out.capacity();
  signal_out.resize(capacity);
  std::size_t found = 0;
  const int status =
      entry(image_in, height, width,
            reinterpret_cast<sf_signal_pixel *>(signal_out.data()), capacity,
            &found);
  if (status == SF_DEVICE_ERROR) {
    signal_out.clear();
    fail("gpu::find");
  }
  if (status != SF_OK) {
    signal_out.clear();
    return status;
  }
  signal_out.resize(found);
  if (found > capacity &&
      sf_hip_take_found(reinterpret_cast<sf_signal_pixel *>(signal_out.data()),
                        found) != SF_OK) {
    signal_out.clear();
    fail("gpu::find");
  }
  if (timing) {
    double t[6];
    sf_hip_last_stage_times(t);
    StageTimes times;
    times.upload = t[0];
    times.stage0 = t[1];
    times.stage1 = t[2];
    times.stage2 = t[3];
    times.copy = t[4];
    times.sort = t[5];
    internal::set_stage_times(times);
  }
  return 0;
}

template int find<std::uint16_t>(const std::uint16_t *,
                                 std::vector<SignalPixel> &, std::size_t,
                                 std::size_t);
template int find<std::uint32_t>(const std::uint32_t *,
                                 std::vector<SignalPixel> &, std::size_t,
                                 std::size_t);

} // namespace gpu
