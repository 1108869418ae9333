// The detector's pixel mask stamped into frames before the threshold: the
// step the reference names as its one omission (README.md:227-234). DIALS
// masks the pixels of /entry/instrument/detector/pixel_mask; the reference
// recognises only the sentinel values at or above max - 1 (src/dext.cc:40),
// so a bad pixel recorded with a plausible value is masked there and not
// here, and closing the gap needs the sentinel written into the frame
// before the threshold, not a kernel change. A masked pixel becomes max(T)
// (0xFFFF, 0xFFFFFFFF), which every threshold form treats as the reference
// treats a module gap.
//
// The mask is a device bit array: bit (p & 31) of word p >> 5 is pixel p of
// a row-major frame. Each thread takes 8 pixels (one byte of the array) and
// writes the sentinel at its set bits in every frame of the launch, so a
// clear byte costs one load and no frame traffic: a detector mask (gaps, a
// few bad pixels) costs its bits plus the masked pixels' writes, not a pass
// over the frames.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>

#include "../../include/dext_hip.h"

namespace {

thread_local std::string g_mask_error;

// A launch's frames, passed by value (no pointer table to upload).
constexpr int kMaskFrames = 64;
struct FramePtrs {
  void *p[kMaskFrames];
};

// One lane per 8 adjacent pixels (one byte of the bit array), looping over
// the launch's frames: neighbouring lanes write neighbouring 16-byte pieces of
// a row, so a module gap is written by whole-wave coalesced stores, and a
// fully masked byte on a 16-byte aligned frame is one or two 16-byte stores.
template <typename T>
__global__ __launch_bounds__(256) void stamp_kernel(FramePtrs frames,
                                                    int n_frames,
                                                    const uint32_t *bits,
                                                    uint64_t groups,
                                                    uint64_t pixels) {
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
       g < groups; g += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t m = (bits[g >> 2] >> (8 * (g & 3))) & 0xFFu;
    if (8 * g + 8 > pixels) // bits past the frame's end
      m &= (1u << (pixels - 8 * g)) - 1u;
    if (!m)
      continue;
    for (int f = 0; f < n_frames; f++) {
      T *px = static_cast<T *>(frames.p[f]) + 8 * g;
      if (m == 0xFFu && (reinterpret_cast<uintptr_t>(frames.p[f]) & 15) == 0) {
        uint4 *v = reinterpret_cast<uint4 *>(px);
        const uint4 ones = make_uint4(~0u, ~0u, ~0u, ~0u);
        v[0] = ones;
        if (sizeof(T) == 4)
          v[1] = ones;
      } else {
        for (uint32_t b = m; b; b &= b - 1)
          px[__builtin_ctz(b)] = (T)~(T)0;
      }
    }
  }
}

// Bit p of the array = (mask[p] != 0): one ballot per 64 pixels, so each
// wave writes two words.
__global__ __launch_bounds__(256) void pack_kernel(const uint8_t *mask,
                                                   uint64_t pixels,
                                                   uint32_t *bits) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
       p < ((pixels + 63) & ~63ull); p += stride) {
    const bool set = p < pixels && mask[p] != 0;
    const uint64_t word = __ballot(set);
    if ((threadIdx.x & 63) == 0) {
      const uint64_t at = p >> 5; // the wave's 64 pixels start at p
      if (32 * at < pixels)
        bits[at] = (uint32_t)word;
      if (32 * (at + 1) < pixels)
        bits[at + 1] = (uint32_t)(word >> 32);
    }
  }
}

int fail(int status, const std::string &what) {
  g_mask_error = what;
  return status;
}

uint32_t grid_for(uint64_t n) {
  return (uint32_t)std::min<uint64_t>(std::max<uint64_t>((n + 255) / 256, 1),
                                      65535);
}

} // namespace

extern "C" int sf_hip_stamp_mask(int pixel_bytes, void *const *frames,
                                 int n_frames, size_t height, size_t width,
                                 const uint32_t *bits, void *stream) {
  if ((pixel_bytes != 2 && pixel_bytes != 4) || n_frames < 0 ||
      (n_frames && (!frames || !bits)) || height == 0 || width == 0)
    return fail(SF_BAD_ARGUMENT, "sf_hip_stamp_mask: bad argument");
  for (int f = 0; f < n_frames; f++)
    if (!frames[f])
      return fail(SF_BAD_ARGUMENT, "sf_hip_stamp_mask: a null frame");
  const uint64_t pixels = (uint64_t)height * width;
  const uint64_t groups = (pixels + 7) / 8;
  hipStream_t s = static_cast<hipStream_t>(stream);
  for (int f0 = 0; f0 < n_frames; f0 += kMaskFrames) {
    FramePtrs batch{};
    const int n = std::min(kMaskFrames, n_frames - f0);
    for (int f = 0; f < n; f++)
      batch.p[f] = frames[f0 + f];
    if (pixel_bytes == 2)
      stamp_kernel<uint16_t><<<grid_for(groups), 256, 0, s>>>(
          batch, n, bits, groups, pixels);
    else
      stamp_kernel<uint32_t><<<grid_for(groups), 256, 0, s>>>(
          batch, n, bits, groups, pixels);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return fail(SF_DEVICE_ERROR,
                std::string("sf_hip_stamp_mask: ") + hipGetErrorString(e));
  return SF_OK;
}

extern "C" int sf_hip_pack_mask(const uint8_t *mask, size_t height,
                                size_t width, uint32_t *bits, void *stream) {
  if (!mask || !bits || height == 0 || width == 0)
    return fail(SF_BAD_ARGUMENT, "sf_hip_pack_mask: bad argument");
  const uint64_t pixels = (uint64_t)height * width;
  pack_kernel<<<grid_for(pixels), 256, 0, static_cast<hipStream_t>(stream)>>>(
      mask, pixels, bits);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return fail(SF_DEVICE_ERROR,
                std::string("sf_hip_pack_mask: ") + hipGetErrorString(e));
  return SF_OK;
}

extern "C" const char *sf_hip_mask_last_error(void) {
  return g_mask_error.c_str();
}
