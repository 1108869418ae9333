// The fused, single-pass extended-dispersion threshold for gfx950.
//
// Reference semantics: src/dext.cc:8-225 (the CPU path is the spec) and the
// device contract of src/dext_gpu.hh:141-158. This is not a translation of
// src/dext_cuda.cu's three kernels: one wave sweeps a vertical strip of the
// frame top to bottom and carries every window as a running sum in its own
// registers, so the frame is read from HBM once and nothing but the signal
// list and a small cell table is written back.
//
// Work item = (frame, strip, band). A strip is 64 lanes x K columns
// (K = 4 at 16 bits, 2 at 32 bits) of which the middle strip_width() columns
// are output and HALO columns on each side feed the windows. A band is
// band_rows output rows; the sweep starts 10 rows above it and ends 10 below,
// which is the total reach of the three chained windows (3 + 2 + 5).
//
// Per input row y a lane:
//   1. takes its K pixels of row y (prefetched 5 rows ahead into a 16-row
//      register ring; off-frame pixels read as the masked sentinel, which is
//      exactly the clipped-window and erode-edge rule of src/dext.cc:71-95,
//      117-142: they add nothing to a sum and count as set for the erode);
//   2. adds row y and drops row y-7 from three 7-row column sums (count and
//      sum packed in one word, and sum of squares), then sums 7 columns
//      across lanes with DPP moves -> the 7x7 window of row y-3 -> the
//      dispersion test in R (src/dext.cc:99-111) -> mask0 bits of row y-3;
//   3. erodes mask0 5x5 with bit operations (DPP for the neighbouring
//      lanes' bits, a shift register for the rows) -> the erode of row y-5;
//   4. builds the 11x11 pair (count, sum) of pixels neither masked nor
//      eroded-set for row y-5 and slides it like (2) -> window of row y-10;
//   5. where row y-10 has candidates (p > 0 and eroded-set), the Poisson
//      test in R (src/dext.cc:193-207) and the emit.
// Survivors of a (row, strip) cell are written contiguously to a staging
// area at an atomically reserved offset, and the cell's {count, offset} is
// recorded; order_kernel then lays every cell out in row-major order, which
// is ascending index order, so no host sort is needed (replacing
// src/signal_order.cc on the host).
#include "sf_device.hh"

namespace sf {

// sqrt(2 (n - 1)) for window counts n = 2..49, correctly rounded, which is
// what the reference evaluates per pixel (src/dext.cc:107-108). n = 0, 1 map
// to +inf, so s * 6 * inf is inf or NaN and "lhs > rhs" is false there: the
// n >= 2 guard of src/dext.cc:102 without a branch.
__constant__ float kSqrtLutF[64];
__constant__ double kSqrtLutD[64];

template <typename R> __device__ __forceinline__ const R *sqrt_lut();
template <> __device__ __forceinline__ const float *sqrt_lut<float>() {
  return kSqrtLutF;
}
template <> __device__ __forceinline__ const double *sqrt_lut<double>() {
  return kSqrtLutD;
}

// Horizontal window sum of radius RAD at each of a lane's K columns, from the
// lane's own column sums and its neighbours' (DPP, several hops when RAD > K).
template <int K, int RAD, typename W>
__device__ __forceinline__ void window_across(const W (&a)[K], W (&out)[K]) {
  constexpr int HOPS = (RAD + K - 1) / K;
  W ext[K + 2 * RAD];
#pragma unroll
  for (int i = 0; i < K; i++)
    ext[RAD + i] = a[i];
  W lo[K], hi[K];
#pragma unroll
  for (int i = 0; i < K; i++) {
    lo[i] = a[i];
    hi[i] = a[i];
  }
#pragma unroll
  for (int h = 1; h <= HOPS; h++) {
#pragma unroll
    for (int i = 0; i < K; i++) {
      const int dl = -h * K + i;  // column offset reached from below
      const int dh = h * K + i;   // and from above
      if (dl >= -RAD) {
        lo[i] = from_below(lo[i]);
        ext[RAD + dl] = lo[i];
      }
      if (dh <= K - 1 + RAD) {
        hi[i] = from_above(hi[i]);
        ext[RAD + dh] = hi[i];
      }
    }
  }
  W s = 0;
#pragma unroll
  for (int d = 0; d <= 2 * RAD; d++)
    s += ext[d];
  out[0] = s;
#pragma unroll
  for (int i = 1; i < K; i++) {
    s += ext[i + 2 * RAD] - ext[i - 1];
    out[i] = s;
  }
}

template <typename T> struct RowSignal {
  float background[Traits<T>::K];
  uint32_t population[Traits<T>::K];
};

// The Poisson test at a row's candidates and the emit of the survivors.
// Wave-uniform: called only when some lane of the wave has a candidate.
// Returns {count, staging offset} of the cell.
template <typename T>
__device__ __noinline__ uint2 poisson_and_emit(
    const FrameDesc &fd, uint32_t cand, uint32_t row, int c_lane,
    const typename Traits<T>::W (&win)[Traits<T>::K],
    const uint32_t (&p)[Traits<T>::K]) {
  using Tr = Traits<T>;
  using W = typename Tr::W;
  using R = typename Tr::R;
  constexpr int K = Tr::K;
  constexpr W kLow = ((W)1 << Tr::kNs2Shift) - 1;
  uint32_t sig = 0;
  float background[K];
  uint32_t population[K];
#pragma unroll
  for (int i = 0; i < K; i++) {
    background[i] = 0.0f;
    population[i] = 0;
    if ((cand >> i) & 1u) {
      const uint32_t n = (uint32_t)(win[i] >> Tr::kNs2Shift);
      const W s = win[i] & kLow;
      // src/dext.cc:203-207: mean over at least two pixels, else zero.
      const R mean = n >= 2 ? (R)s / (R)n : (R)0;
      const bool hit = (R)p[i] >= mean + (R)3 * sqrt(mean);
      if (hit) {
        sig |= 1u << i;
        background[i] = (float)mean;
        population[i] = n;
      }
    }
  }
  const uint32_t cnt = __builtin_popcount(sig);
  const uint64_t b0 = ballot(cnt & 1u), b1 = ballot(cnt & 2u),
                 b2 = ballot(cnt & 4u);
  const uint32_t below =
      count_below(b0) + 2 * count_below(b1) + 4 * count_below(b2);
  const uint32_t total = (uint32_t)(__builtin_popcountll(b0) +
                                    2 * __builtin_popcountll(b1) +
                                    4 * __builtin_popcountll(b2));
  uint32_t base = 0;
  if (total != 0) {
    if (lane_id() == 0)
      base = atomicAdd(fd.counter, total);
    base = __builtin_amdgcn_readfirstlane(base);
    uint32_t at = base + below;
    const uint32_t row_index = row * fd.width;
#pragma unroll
    for (int i = 0; i < K; i++) {
      if ((sig >> i) & 1u) {
        if (at < fd.capacity) {
          sf_signal_pixel e;
          e.index = row_index + (uint32_t)(c_lane + i);
          e.value = p[i];
          e.background = background[i];
          e.population = (uint16_t)population[i];
          e.reserved = 0;
          fd.staging[at] = e;
        }
        at++;
      }
    }
  }
  return make_uint2(total, base);
}

template <typename T, bool VEC>
__device__ __forceinline__ void load_row(const T *pix, int y, int height,
                                         int width, int c_lane,
                                         uint32_t (&v)[Traits<T>::K]) {
  constexpr int K = Traits<T>::K;
  if (y < 0 || y >= height) {
#pragma unroll
    for (int i = 0; i < K; i++)
      v[i] = Traits<T>::kOffFrame;
    return;
  }
  const T *row = pix + (size_t)y * (size_t)width;
  if (VEC) {
    if (c_lane >= 0 && c_lane + K <= width) {
      const uint2 d = *reinterpret_cast<const uint2 *>(row + c_lane);
      if (K == 4) {
        v[0] = d.x & 0xFFFFu;
        v[1] = d.x >> 16;
        v[2 % K] = d.y & 0xFFFFu;
        v[3 % K] = d.y >> 16;
      } else {
        v[0] = d.x;
        v[1 % K] = d.y;
      }
    } else {
#pragma unroll
      for (int i = 0; i < K; i++)
        v[i] = Traits<T>::kOffFrame;
    }
  } else {
#pragma unroll
    for (int i = 0; i < K; i++) {
      const int c = c_lane + i;
      v[i] = (c >= 0 && c < width) ? (uint32_t)row[c] : Traits<T>::kOffFrame;
    }
  }
}

template <typename T, bool VEC>
__device__ __forceinline__ void threshold_item(
    const FrameDesc &fd, uint32_t strip, uint32_t band,
    const typename Traits<T>::R *lut) {
  using Tr = Traits<T>;
  using W = typename Tr::W;
  using R = typename Tr::R;
  constexpr int K = Tr::K;
  constexpr uint32_t KMASK = (1u << K) - 1u;
  constexpr int HALO = halo<T>();
  constexpr int OW = strip_width<T>();
  constexpr W kOne = (W)1 << Tr::kNsShift;
  constexpr W kLow = kOne - 1;
  constexpr W kOne2 = (W)1 << Tr::kNs2Shift;

  const T *pix = static_cast<const T *>(fd.pixels);
  const int height = (int)fd.height, width = (int)fd.width;
  const uint32_t lane = lane_id();
  const int c_out0 = (int)strip * OW;
  const int c_lane = c_out0 - HALO + (int)lane * K;
  const int r0 = (int)(band * fd.band_rows);
  const int r1 = min(r0 + (int)fd.band_rows, height);
  const int steps = ((r1 - r0) + 20 + 15) & ~15;
  uint32_t band_sum = 0;

  // Output columns of this lane: inside the strip's output span and the
  // frame.
  uint32_t colmask = 0;
#pragma unroll
  for (int i = 0; i < K; i++) {
    const int c = c_lane + i;
    if (c >= c_out0 && c < c_out0 + OW && c < width)
      colmask |= 1u << i;
  }

  uint32_t raw[16][K];
  W ring0n[8][K], ring0q[8][K], ring2[16][K];
  W vn[K], vq[K], v2[K];
#pragma unroll
  for (int i = 0; i < K; i++) {
    vn[i] = 0;
    vq[i] = 0;
    v2[i] = 0;
#pragma unroll
    for (int r = 0; r < 8; r++) {
      ring0n[r][i] = 0;
      ring0q[r][i] = 0;
    }
#pragma unroll
    for (int r = 0; r < 16; r++)
      ring2[r][i] = 0;
  }
  // Bit histories, newest row in the low K bits: masked pixels (rows
  // y .. y-5), horizontally eroded mask0 (rows y-3 .. y-7), candidates
  // (rows y-5 .. y-10).
  uint32_t masked_h = 0, erode_h = 0, cand_h = 0;

  const int y0 = r0 - 10;
#pragma unroll
  for (int j = 0; j < 5; j++)
    load_row<T, VEC>(pix, y0 + j, height, width, c_lane, raw[j]);

  for (int t0 = 0; t0 < steps; t0 += 16) {
#pragma unroll
    for (int j = 0; j < 16; j++) {
      const int t = t0 + j;
      const int y = y0 + t;
      load_row<T, VEC>(pix, y + 5, height, width, c_lane, raw[(j + 5) & 15]);

      // (1)-(2): the 7x7 sums of row y-3.
      uint32_t mbits = 0;
#pragma unroll
      for (int i = 0; i < K; i++) {
        const uint32_t v = raw[j & 15][i];
        const bool ok = v < Tr::kMasked;
        mbits |= (ok ? 0u : 1u) << i;
        const W ns = ok ? (W)v + kOne : (W)0;
        const W q = ok ? (W)v * (W)v : (W)0;
        vn[i] += ns - ring0n[(j + 1) & 7][i];
        vq[i] += q - ring0q[(j + 1) & 7][i];
        ring0n[j & 7][i] = ns;
        ring0q[j & 7][i] = q;
      }
      masked_h = (masked_h << K) | mbits;

      W hn[K], hq[K];
      window_across<K, 3>(vn, hn);
      window_across<K, 3>(vq, hq);

      // Dispersion test, evaluated left to right in R exactly as
      // src/dext.cc:103-108 writes it.
      uint32_t mask0 = (masked_h >> (3 * K)) & KMASK;
#pragma unroll
      for (int i = 0; i < K; i++) {
        const uint32_t n = (uint32_t)(hn[i] >> Tr::kNsShift);
        const R m = (R)n;
        const R s = (R)(hn[i] & kLow);
        const R s2 = (R)hq[i];
        const R lhs = m * s2 - s * s - s * (m - (R)1);
        const R rhs = s * (R)6 * lut[n];
        mask0 |= (lhs > rhs ? 1u : 0u) << i;
      }

      // (3): 5x5 erode; off-frame columns and rows are masked, so they read
      // as set, as src/dext.cc:123-137 treats them.
      const uint32_t x =
          from_below(mask0) | (mask0 << K) | (from_above(mask0) << (2 * K));
      const uint32_t across =
          ((x & (x >> 1) & (x >> 2) & (x << 1) & (x << 2)) >> K) & KMASK;
      erode_h = (erode_h << K) | across;
      const uint32_t eroded = erode_h & (erode_h >> K) & (erode_h >> (2 * K)) &
                              (erode_h >> (3 * K)) & (erode_h >> (4 * K)) &
                              KMASK;

      // (4): pixels of row y-5 that are neither masked nor eroded-set feed
      // the 11x11 window (src/dext.cc:151-161); set ones with p > 0 are
      // candidates (src/dext.cc:196).
      const uint32_t masked5 = (masked_h >> (5 * K)) & KMASK;
      uint32_t positive = 0;
#pragma unroll
      for (int i = 0; i < K; i++) {
        const uint32_t p = raw[(j + 11) & 15][i];
        const bool use = !(((masked5 | eroded) >> i) & 1u);
        const W ns2 = use ? (W)p + kOne2 : (W)0;
        positive |= (p != 0 ? 1u : 0u) << i;
        v2[i] += ns2 - ring2[(j + 5) & 15][i];
        ring2[j & 15][i] = ns2;
      }
      cand_h = (cand_h << K) | (eroded & positive & ~masked5);

      // (5): row y-10 is an output row of this band once t >= 20. Its
      // 11x11 sums are formed across lanes only when some lane of the wave
      // has a candidate there (about one wave-row in ten on diffraction
      // frames); the column sums above are kept up to date every row.
      const int row = y - 10;
      if (t >= 20 && row < r1) {
        const uint32_t cand = (cand_h >> (5 * K)) & KMASK & colmask;
        uint2 cell = make_uint2(0u, 0u);
        if (ballot(cand != 0) != 0) {
          W h2[K];
          window_across<K, 5>(v2, h2);
          uint32_t p10[K];
#pragma unroll
          for (int i = 0; i < K; i++)
            p10[i] = raw[(j + 6) & 15][i];
          cell = poisson_and_emit<T>(fd, cand, (uint32_t)row, c_lane, h2, p10);
        }
        band_sum += cell.x;
        if (lane == 0) {
          const size_t c = (size_t)row * fd.n_strips + strip;
          fd.cell_count[c] = (uint8_t)cell.x;
          if (cell.x)
            fd.cell_base[c] = cell.y;
        }
      }
    }
  }
  if (lane == 0 && band_sum)
    atomicAdd(&fd.band_total[band], band_sum);
}

// One wave per block. Items are (frame, band, strip), strip fastest; with
// `persistent` the grid strides over them, otherwise one block per item.
template <typename T, bool VEC>
__global__ __launch_bounds__(64) void fused_kernel(const FrameDesc *frames,
                                                   int n_frames,
                                                   uint32_t items) {
  using R = typename Traits<T>::R;
  // The square-root table in LDS: a per-lane lookup there costs one LDS
  // read instead of a dependent load through the vector memory path.
  __shared__ R lut[64];
  lut[threadIdx.x] = sqrt_lut<R>()[threadIdx.x];
  __syncthreads();
  const uint32_t blocks = gridDim.x;
  for (uint32_t b = blockIdx.x; b < ((items + 7) & ~7u); b += blocks) {
    const uint32_t item =
        blocks >= ((items + 7) & ~7u) ? xcd_local_item(b, blocks) : b;
    if (item >= items)
      continue;
    const int f = frame_of_item(frames, n_frames, item);
    const FrameDesc &fd = frames[f];
    if (VEC != (fd.vec != 0))
      continue;
    const uint32_t local = item - fd.item_begin;
    threshold_item<T, VEC>(fd, local % fd.n_strips, local / fd.n_strips, lut);
  }
}

// Exclusive prefix over a block of 256 threads (4 waves); *total gets the
// block's sum.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v,
                                                         uint32_t *total,
                                                         uint32_t *lds) {
  const uint32_t lane = lane_id();
  const uint32_t wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, d, 64);
    if (lane >= (uint32_t)d)
      x += y;
  }
  if (lane == 63)
    lds[wave] = x;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < 4; w++) {
    before += (uint32_t)w < wave ? lds[w] : 0u;
    all += lds[w];
  }
  *total = all;
  __syncthreads();
  return before + x - v;
}

// Lays one band of one frame out in row-major cell order, which is ascending
// index order: the band starts after every earlier band's survivors, and
// each cell's run after the cells before it in the band. Grid (bands,
// frames).
__global__ __launch_bounds__(256) void order_kernel(const FrameDesc *frames) {
  __shared__ uint32_t lds[8];
  const FrameDesc &fd = frames[blockIdx.y];
  const uint32_t band = blockIdx.x;
  if (band >= fd.n_bands)
    return;
  const uint32_t found = *fd.counter;
  if (band == 0 && threadIdx.x == 0)
    *fd.out_count = found;
  if (found > fd.capacity)
    return;
  uint32_t mine = 0;
  for (uint32_t b = threadIdx.x; b < band; b += blockDim.x)
    mine += fd.band_total[b];
  uint32_t before = 0;
  (void)block_exclusive_scan(mine, &before, lds);
  const size_t c0 = (size_t)band * fd.band_rows * fd.n_strips;
  const size_t r1 = min((size_t)(band + 1) * fd.band_rows, (size_t)fd.height);
  const size_t c1 = r1 * fd.n_strips;
  const size_t per = (c1 - c0 + blockDim.x - 1) / blockDim.x;
  const size_t my0 = min(c0 + threadIdx.x * per, c1);
  const size_t my1 = min(my0 + per, c1);
  uint32_t sum = 0;
  for (size_t c = my0; c < my1; c++)
    sum += fd.cell_count[c];
  uint32_t band_sum = 0;
  uint32_t dest = before + block_exclusive_scan(sum, &band_sum, lds);
  for (size_t c = my0; c < my1; c++) {
    const uint32_t k = fd.cell_count[c];
    if (!k)
      continue;
    const uint32_t base = fd.cell_base[c];
    // The runs lie inside the reserved staging and the counts sum to
    // `found`, so neither bound binds; they are kept as guards.
    for (uint32_t e = 0; e < k; e++)
      if (dest + e < fd.capacity && base + e < fd.capacity)
        fd.out[dest + e] = fd.staging[base + e];
    dest += k;
  }
}

__global__ void zero_counters(const FrameDesc *frames) {
  const FrameDesc &fd = frames[blockIdx.x];
  if (threadIdx.x == 0)
    *fd.counter = 0;
  for (uint32_t b = threadIdx.x; b < fd.n_bands; b += blockDim.x)
    fd.band_total[b] = 0;
}

} // namespace sf

// ---------------------------------------------------------------------------
// Host-side launchers (called from the plan in capi.cpp).
// ---------------------------------------------------------------------------
namespace sf {

void upload_sqrt_tables() {
  float f[64];
  double d[64];
  for (int n = 0; n < 64; n++) {
    if (n < 2) {
      f[n] = __builtin_inff();
      d[n] = __builtin_inf();
    } else {
      f[n] = __builtin_sqrtf(2.0f * (float)(n - 1));
      d[n] = __builtin_sqrt(2.0 * (double)(n - 1));
    }
  }
  (void)hipMemcpyToSymbol(HIP_SYMBOL(kSqrtLutF), f, sizeof(f));
  (void)hipMemcpyToSymbol(HIP_SYMBOL(kSqrtLutD), d, sizeof(d));
}

hipError_t launch_zero(const FrameDesc *frames, int n_frames,
                       hipStream_t stream) {
  zero_counters<<<n_frames, 64, 0, stream>>>(frames);
  return hipGetLastError();
}

hipError_t launch_fused(int pixel_bytes, bool any_vec, bool any_scalar,
                        const FrameDesc *frames, int n_frames, uint32_t items,
                        int blocks, hipStream_t stream) {
  const uint32_t padded = (items + 7) & ~7u;
  const uint32_t grid =
      blocks > 0 ? (uint32_t)blocks < padded ? (uint32_t)blocks : padded
                 : padded;
  if (pixel_bytes == 2) {
    if (any_vec)
      fused_kernel<uint16_t, true><<<grid, 64, 0, stream>>>(frames, n_frames,
                                                            items);
    if (any_scalar)
      fused_kernel<uint16_t, false><<<grid, 64, 0, stream>>>(frames, n_frames,
                                                             items);
  } else {
    if (any_vec)
      fused_kernel<uint32_t, true><<<grid, 64, 0, stream>>>(frames, n_frames,
                                                            items);
    if (any_scalar)
      fused_kernel<uint32_t, false><<<grid, 64, 0, stream>>>(frames, n_frames,
                                                             items);
  }
  return hipGetLastError();
}

hipError_t launch_order(const FrameDesc *frames, int n_frames,
                        uint32_t max_bands, hipStream_t stream) {
  order_kernel<<<dim3(max_bands, n_frames), 256, 0, stream>>>(frames);
  return hipGetLastError();
}

} // namespace sf
