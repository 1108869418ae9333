// The plain per-frame form: three kernels, one pixel per thread, every window
// summed directly. It is the comparison point for the fused kernel (the
// "one frame per launch" form) and an independent second device
// implementation of src/dext.cc:8-225 for the parity tests.
//
// stage0: 7x7 clipped window -> mask0 (1 byte a pixel), src/dext.cc:69-113
// stage1: 5x5 erode -> mask1, src/dext.cc:117-142
// stage2: 11x11 window of pixels neither masked nor mask1 at candidates,
//         Poisson test, emit, src/dext.cc:151-222. A wave is 64 consecutive
//         pixels of one row, so a (row, 64-column block) cell is written by
//         one wave and the same ordering kernel as the fused form applies.
#include "sf_device.hh"

namespace sf {

template <typename T>
__global__ __launch_bounds__(256) void plain_stage0(const FrameDesc *frames,
                                                    uint8_t *mask0) {
  using Tr = Traits<T>;
  using W = typename Tr::W;
  using R = typename Tr::R;
  const FrameDesc &fd = frames[blockIdx.z];
  const int width = (int)fd.width, height = (int)fd.height;
  const int j = blockIdx.x * 64 + (threadIdx.x & 63);
  const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= height || j >= width)
    return;
  const T *pix = static_cast<const T *>(fd.pixels);
  W n = 0, s = 0, q = 0;
  for (int a = max(i - 3, 0); a <= min(i + 3, height - 1); a++) {
    const T *row = pix + (size_t)a * width;
    for (int b = max(j - 3, 0); b <= min(j + 3, width - 1); b++) {
      const W v = row[b];
      if (v < Tr::kMasked) {
        n += 1;
        s += v;
        q += v * v;
      }
    }
  }
  bool hit = false;
  if (n >= 2) {
    const R m = (R)n, rs = (R)s, rq = (R)q;
    hit = (m * rq - rs * rs - rs * (m - (R)1)) >
          (rs * (R)6 * sqrt((R)2 * (m - (R)1)));
  }
  const size_t k = (size_t)i * width + j;
  mask0[k] = (hit || pix[k] >= Tr::kMasked) ? 1 : 0;
}

template <typename T>
__global__ __launch_bounds__(256) void plain_stage1(const FrameDesc *frames,
                                                    const uint8_t *mask0,
                                                    uint8_t *mask1) {
  using Tr = Traits<T>;
  const FrameDesc &fd = frames[blockIdx.z];
  const int width = (int)fd.width, height = (int)fd.height;
  const int j = blockIdx.x * 64 + (threadIdx.x & 63);
  const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= height || j >= width)
    return;
  const T *pix = static_cast<const T *>(fd.pixels);
  const size_t k = (size_t)i * width + j;
  uint8_t m = mask0[k];
  if (pix[k] >= Tr::kMasked) {
    m = 1;
  } else if (m) {
    for (int a = max(i - 2, 0); a <= min(i + 2, height - 1); a++)
      for (int b = max(j - 2, 0); b <= min(j + 2, width - 1); b++)
        m &= mask0[(size_t)a * width + b];
  }
  mask1[k] = m;
}

template <typename T>
__global__ __launch_bounds__(256) void plain_stage2(const FrameDesc *frames,
                                                    const uint8_t *mask1) {
  using Tr = Traits<T>;
  using W = typename Tr::W;
  using R = typename Tr::R;
  const FrameDesc &fd = frames[blockIdx.z];
  const int width = (int)fd.width, height = (int)fd.height;
  const int j = blockIdx.x * 64 + (threadIdx.x & 63);
  const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= height) // whole wave: the row is uniform across it
    return;
  const T *pix = static_cast<const T *>(fd.pixels);
  bool hit = false;
  float background = 0.0f;
  uint32_t population = 0;
  uint32_t p = 0;
  if (j < width) {
    const size_t k = (size_t)i * width + j;
    const uint32_t v = pix[k];
    p = v >= Tr::kMasked ? 0u : v;
    if (p > 0 && mask1[k]) {
      W n = 0, s = 0;
      for (int a = max(i - 5, 0); a <= min(i + 5, height - 1); a++)
        for (int b = max(j - 5, 0); b <= min(j + 5, width - 1); b++) {
          const size_t kk = (size_t)a * width + b;
          const W u = pix[kk];
          if (u < Tr::kMasked && !mask1[kk]) {
            n += 1;
            s += u;
          }
        }
      const R mean = n >= 2 ? (R)s / (R)n : (R)0;
      if ((R)p >= mean + (R)3 * sqrt(mean)) {
        hit = true;
        background = (float)mean;
        population = (uint32_t)n;
      }
    }
  }
  const uint64_t votes = ballot(hit);
  const uint32_t total = (uint32_t)__builtin_popcountll(votes);
  uint32_t base = 0;
  if (total != 0) {
    if (lane_id() == 0)
      base = atomicAdd(fd.counter, total);
    base = __builtin_amdgcn_readfirstlane(base);
    if (hit) {
      const uint32_t at = base + count_below(votes);
      if (at < fd.capacity) {
        sf_signal_pixel e;
        e.index = (uint32_t)((size_t)i * width + j);
        e.value = p;
        e.background = background;
        e.population = (uint16_t)population;
        e.reserved = 0;
        fd.staging[at] = e;
      }
    }
  }
  if (lane_id() == 0) {
    const size_t c = (size_t)i * fd.n_strips + blockIdx.x;
    fd.cell_count[c] = (uint8_t)total;
    if (total) {
      fd.cell_base[c] = base;
      atomicAdd(&fd.band_total[i / fd.band_rows], total);
    }
  }
}

hipError_t launch_plain(int pixel_bytes, const FrameDesc *frames,
                        int frame_index, uint32_t height, uint32_t width,
                        uint8_t *mask0, uint8_t *mask1, hipStream_t stream,
                        hipEvent_t after0, hipEvent_t after1) {
  const dim3 grid((width + 63) / 64, (height + 3) / 4, 1);
  const FrameDesc *one = frames + frame_index;
  if (pixel_bytes == 2) {
    plain_stage0<uint16_t><<<grid, 256, 0, stream>>>(one, mask0);
    if (after0)
      (void)hipEventRecord(after0, stream);
    plain_stage1<uint16_t><<<grid, 256, 0, stream>>>(one, mask0, mask1);
    if (after1)
      (void)hipEventRecord(after1, stream);
    plain_stage2<uint16_t><<<grid, 256, 0, stream>>>(one, mask1);
  } else {
    plain_stage0<uint32_t><<<grid, 256, 0, stream>>>(one, mask0);
    if (after0)
      (void)hipEventRecord(after0, stream);
    plain_stage1<uint32_t><<<grid, 256, 0, stream>>>(one, mask0, mask1);
    if (after1)
      (void)hipEventRecord(after1, stream);
    plain_stage2<uint32_t><<<grid, 256, 0, stream>>>(one, mask1);
  }
  return hipGetLastError();
}

} // namespace sf
