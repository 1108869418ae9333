// Synthetic diffraction-like frames, generated on the device so that a
// benchmark series of thousands of full-size frames can be made resident in
// HBM without a host round trip. Not part of the threshold: input only.
//
// Background: Poisson(mean) by inversion of a host-built CDF table, from a
// counter-based hash of (seed, frame, pixel). Module gaps at the Eiger pitch
// (514 x 1030 modules, 38-row and 10-column gaps) are set to max(T); hot
// pixels to max(T) - 1. Spots are 2-D Gaussians with a Gaussian rocking
// curve over their frame span, each pixel's expected count drawn as a
// Poisson deviate (Knuth below 30, normal approximation above), added with a
// saturating compare-and-swap that never creates a sentinel.
#include "sf_device.hh"

#include <cmath>
#include <vector>

namespace sf {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ double unit(uint64_t h) {
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

template <typename T>
__global__ void synth_background(T *frames, uint32_t height, uint32_t width,
                                 uint32_t n_frames, uint64_t seed,
                                 const uint32_t *cdf, int cdf_len, int gaps,
                                 uint32_t hot_threshold) {
  const size_t n = (size_t)height * width;
  const size_t total = n * n_frames;
  const T top = (T)~(T)0;
  for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
       g += (size_t)gridDim.x * blockDim.x) {
    const size_t k = g % n;
    const uint32_t row = (uint32_t)(k / width);
    const uint32_t col = (uint32_t)(k % width);
    const uint64_t h = mix64(seed ^ mix64(g * 2 + 1));
    T v;
    if (gaps && ((row % 552u) >= 514u || (col % 1040u) >= 1030u)) {
      v = top;
    } else if ((uint32_t)(h >> 32) < hot_threshold) {
      v = (T)(top - 1);
    } else {
      const uint32_t u = (uint32_t)h;
      int c = 0;
      while (c < cdf_len - 1 && u >= cdf[c])
        c++;
      v = (T)c;
    }
    frames[g] = v;
  }
}

// Saturating add into one pixel of a T frame, leaving sentinels alone.
template <typename T>
__device__ void add_counts(T *frames, size_t at, uint32_t add);

template <>
__device__ void add_counts<uint16_t>(uint16_t *frames, size_t at,
                                     uint32_t add) {
  uint32_t *word = reinterpret_cast<uint32_t *>(
      reinterpret_cast<uintptr_t>(frames + at) & ~(uintptr_t)3);
  const int shift =
      (reinterpret_cast<uintptr_t>(frames + at) & 2) ? 16 : 0;
  uint32_t old = *word, assumed;
  do {
    assumed = old;
    const uint32_t cur = (assumed >> shift) & 0xFFFFu;
    if (cur >= 0xFFFEu)
      return;
    uint32_t next = cur + add;
    next = next > 0xFFFDu ? 0xFFFDu : next;
    const uint32_t merged =
        (assumed & ~(0xFFFFu << shift)) | (next << shift);
    old = atomicCAS(word, assumed, merged);
  } while (old != assumed);
}

template <>
__device__ void add_counts<uint32_t>(uint32_t *frames, size_t at,
                                     uint32_t add) {
  uint32_t *word = frames + at;
  uint32_t old = *word, assumed;
  do {
    assumed = old;
    if (assumed >= 0xFFFFFFFEu)
      return;
    const uint64_t next = (uint64_t)assumed + add;
    const uint32_t clipped =
        next > 0xFFFFFFFDull ? 0xFFFFFFFDu : (uint32_t)next;
    old = atomicCAS(word, assumed, clipped);
  } while (old != assumed);
}

struct SpotInstance {
  float x, y, sigma;
  float expected; // counts of this spot on this frame
  uint32_t frame;
  uint32_t id;
};

template <typename T>
__global__ void synth_spots(T *frames, uint32_t height, uint32_t width,
                            uint32_t n_frames, uint64_t seed,
                            const SpotInstance *spots) {
  const SpotInstance s = spots[blockIdx.x];
  if (s.frame >= n_frames)
    return;
  const int radius = min(7, (int)ceilf(4.0f * s.sigma));
  const int side = 2 * radius + 1;
  const int cx = (int)floorf(s.x), cy = (int)floorf(s.y);
  const float norm = 1.0f / (6.2831853f * s.sigma * s.sigma);
  for (int t = threadIdx.x; t < side * side; t += blockDim.x) {
    const int px = cx - radius + t % side;
    const int py = cy - radius + t / side;
    if (px < 0 || py < 0 || px >= (int)width || py >= (int)height)
      continue;
    const float dx = (float)px + 0.5f - s.x, dy = (float)py + 0.5f - s.y;
    const double e = (double)(s.expected * norm *
                              expf(-(dx * dx + dy * dy) /
                                   (2.0f * s.sigma * s.sigma)));
    if (e < 0.05)
      continue;
    uint64_t h = mix64(seed ^ mix64(((uint64_t)s.id << 24) ^
                                    ((uint64_t)s.frame << 8) ^ (uint64_t)t ^
                                    0x5bd1e995ull));
    uint32_t draw;
    if (e < 30.0) {
      const double limit = exp(-e);
      double prod = unit(h);
      draw = 0;
      while (prod > limit && draw < 200) {
        h = mix64(h);
        prod *= unit(h);
        draw++;
      }
    } else {
      const uint64_t h2 = mix64(h);
      const double u1 = unit(h) + 1e-300, u2 = unit(h2);
      const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
      const double v = e + sqrt(e) * z;
      draw = v <= 0.0 ? 0u : (uint32_t)fmin(v + 0.5, 4.0e9);
    }
    if (draw)
      add_counts<T>(frames, (size_t)s.frame * height * width +
                                (size_t)py * width + px,
                    draw);
  }
}

} // namespace sf

extern "C" int sf_hip_synth_frames(void *device, int pixel_bytes,
                                   uint32_t height, uint32_t width,
                                   uint32_t frames, uint64_t seed,
                                   double background, int gaps,
                                   double hot_fraction, const double *spots,
                                   uint32_t n_spots, double sigma_z) {
  using namespace sf;
  if (!device || (pixel_bytes != 2 && pixel_bytes != 4) || !height ||
      !width || !frames)
    return SF_BAD_ARGUMENT;
  // Poisson CDF as 32-bit thresholds.
  std::vector<uint32_t> cdf;
  {
    double p = std::exp(-background), acc = 0.0;
    for (int k = 0; k < 4096; k++) {
      acc += p;
      const double t = acc * 4294967296.0;
      cdf.push_back(t >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)t);
      if (t >= 4294967295.0 || (k > background && p < 1e-12))
        break;
      p *= background / (double)(k + 1);
    }
    cdf.back() = 0xFFFFFFFFu;
  }
  std::vector<SpotInstance> inst;
  for (uint32_t s = 0; s < n_spots; s++) {
    const double *sp = spots + 6 * (size_t)s;
    const int first = (int)sp[4], span = (int)sp[5] < 1 ? 1 : (int)sp[5];
    const double centre = first + 0.5 * (span - 1);
    double wsum = 0.0;
    for (int f = 0; f < span; f++)
      wsum += std::exp(-(f + first - centre) * (f + first - centre) /
                       (2.0 * sigma_z * sigma_z));
    for (int f = 0; f < span; f++) {
      const double w = std::exp(-(f + first - centre) * (f + first - centre) /
                                (2.0 * sigma_z * sigma_z)) /
                       wsum;
      if (first + f < 0 || (uint32_t)(first + f) >= frames)
        continue;
      SpotInstance i;
      i.x = (float)sp[0];
      i.y = (float)sp[1];
      i.sigma = (float)sp[2];
      i.expected = (float)(sp[3] * w);
      i.frame = (uint32_t)(first + f);
      i.id = s;
      inst.push_back(i);
    }
  }
  uint32_t *d_cdf = nullptr;
  SpotInstance *d_inst = nullptr;
  if (hipMalloc(&d_cdf, cdf.size() * sizeof(uint32_t)) != hipSuccess)
    return SF_DEVICE_ERROR;
  hipError_t err = hipMemcpy(d_cdf, cdf.data(), cdf.size() * sizeof(uint32_t),
                             hipMemcpyHostToDevice);
  if (err == hipSuccess && !inst.empty()) {
    err = hipMalloc(&d_inst, inst.size() * sizeof(SpotInstance));
    if (err == hipSuccess)
      err = hipMemcpy(d_inst, inst.data(), inst.size() * sizeof(SpotInstance),
                      hipMemcpyHostToDevice);
  }
  const uint32_t hot = (uint32_t)(hot_fraction * 4294967296.0);
  if (err == hipSuccess) {
    if (pixel_bytes == 2) {
      synth_background<uint16_t><<<4096, 256>>>(
          (uint16_t *)device, height, width, frames, seed, d_cdf,
          (int)cdf.size(), gaps, hot);
      if (!inst.empty())
        synth_spots<uint16_t><<<(uint32_t)inst.size(), 256>>>(
            (uint16_t *)device, height, width, frames, seed, d_inst);
    } else {
      synth_background<uint32_t><<<4096, 256>>>(
          (uint32_t *)device, height, width, frames, seed, d_cdf,
          (int)cdf.size(), gaps, hot);
      if (!inst.empty())
        synth_spots<uint32_t><<<(uint32_t)inst.size(), 256>>>(
            (uint32_t *)device, height, width, frames, seed, d_inst);
    }
    err = hipGetLastError();
    if (err == hipSuccess)
      err = hipDeviceSynchronize();
  }
  (void)hipFree(d_cdf);
  if (d_inst)
    (void)hipFree(d_inst);
  return err == hipSuccess ? SF_OK : SF_DEVICE_ERROR;
}
