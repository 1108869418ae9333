// Six-connected 3-D grouping and the centroid/intensity pass on the device.
//
// Reference semantics: dials_spots::Labeller (src/dials_spots.cc:11-360).
// The reference labels one frame at a time on the host and emits a component
// as soon as a frame arrives that it has no pixel on. Here a chunk of frames
// is labelled at once on the device:
//
//   nodes    the pixels still open from earlier chunks (the frontier), then
//            every signal pixel of the chunk, in (frame, index) order;
//   unite    each node hooks to the pixel on its left (same row and frame),
//            the pixel above, and the same pixel on the previous frame when
//            the frame numbers are consecutive and not two_d
//            (src/dials_spots.cc:102-137); roots always point to the lower
//            node, so a component's root is its first (frame, index) pixel;
//   emit     a component is closed when it has no pixel on the chunk's last
//            frame (or at finish, or always in two_d), and closed components
//            are emitted ordered by (last frame, first pixel) -- the order
//            the streaming Labeller flushes them (src/dials_spots.cc:153-194);
//   finalise one thread per component walks its members in (frame, index)
//            order and repeats the reference's FP64 arithmetic operation for
//            operation (src/dials_spots.cc:221-360), so sums, centroid,
//            variance and the separation test are bit-identical;
//   carry    the open components' pixels, still in (frame, index) order,
//            become the next chunk's frontier, each pointing at its
//            component's first pixel.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "sf_device.hh"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

namespace sf {


struct ChunkDesc {
  const sf_signal_pixel *list; // device list of this frame
  const uint32_t *count;       // device count (may exceed capacity)
  uint32_t capacity;
  uint32_t offset;             // first node of this frame
  int32_t number;              // frame number
  int32_t linked;              // consecutive with the frame before it
};

struct LabelArgs {
  uint32_t n_nodes, n_carry, carry_tail; // carry_tail: first carried node on
                                         // the last carried frame
  uint32_t n_frames;                     // frames in this chunk
  uint32_t width, pixels;
  int32_t closing;                       // close everything (finish, two_d)
  int32_t first_linked;                  // chunk frame 0 follows the carry
  int32_t two_d;
  uint64_t min_size, max_size;
  double max_separation;
};

__device__ __forceinline__ uint32_t find_root(uint32_t *parent, uint32_t i) {
  uint32_t p = parent[i];
  while (p != i) {
    const uint32_t gp = parent[p];
    if (gp != p)
      parent[i] = gp; // path halving; racing writes only shorten paths
    i = p;
    p = gp;
  }
  return i;
}

// Hook the larger root under the smaller, lock-free.
__device__ void unite(uint32_t *parent, uint32_t a, uint32_t b) {
  for (;;) {
    a = find_root(parent, a);
    b = find_root(parent, b);
    if (a == b)
      return;
    if (a > b) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    const uint32_t old = atomicCAS(&parent[b], b, a);
    if (old == b)
      return;
    b = old;
  }
}

__device__ __forceinline__ int chunk_frame_of(const ChunkDesc *frames,
                                              uint32_t n_frames, uint32_t g) {
  int lo = 0, hi = (int)n_frames - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (frames[mid].offset <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Position of `index` among nodes [from, to), or -1.
__device__ __forceinline__ long find_index(const uint32_t *index,
                                           uint32_t from, uint32_t to,
                                           uint32_t want) {
  uint32_t lo = from, hi = to;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (index[mid] < want)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < to && index[lo] == want) ? (long)lo : -1;
}

__global__ void gather_nodes(LabelArgs a, const ChunkDesc *frames,
                             const int32_t *carry_frame,
                             const uint32_t *carry_index,
                             const uint32_t *carry_value,
                             const uint32_t *carry_root, int32_t *node_frame,
                             uint32_t *node_index, uint32_t *node_value,
                             uint32_t *node_slot, uint32_t *parent,
                             uint32_t *last_slot) {
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < a.n_nodes;
       g += gridDim.x * blockDim.x) {
    last_slot[g] = 0;
    if (g < a.n_carry) {
      node_frame[g] = carry_frame[g];
      node_index[g] = carry_index[g];
      node_value[g] = carry_value[g];
      node_slot[g] = 0;
      parent[g] = carry_root[g];
    } else {
      const int k = chunk_frame_of(frames, a.n_frames, g);
      const sf_signal_pixel e = frames[k].list[g - frames[k].offset];
      node_frame[g] = frames[k].number;
      node_index[g] = e.index;
      node_value[g] = e.value;
      node_slot[g] = (uint32_t)k + 1;
      parent[g] = g;
    }
  }
}

// row_start[slot * (height + 1) + y] = the first node of `slot` whose index
// is at least y * width (slot 0: the carried nodes of the last carried
// frame; slot k + 1: chunk frame k). One thread per (slot, row).
__global__ void fill_rows(LabelArgs a, const ChunkDesc *frames,
                          const uint32_t *node_index, uint32_t *row_start) {
  const uint32_t height = a.pixels / a.width;
  const uint32_t total = (a.n_frames + 1) * (height + 1);
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += gridDim.x * blockDim.x) {
    const uint32_t slot = t / (height + 1), y = t % (height + 1);
    uint32_t from, to;
    if (slot == 0) {
      from = a.carry_tail;
      to = a.n_carry;
    } else {
      from = frames[slot - 1].offset;
      to = slot < a.n_frames ? frames[slot].offset : a.n_nodes;
    }
    const uint32_t want = y * a.width; // y == height: past every index
    uint32_t lo = from, hi = to;
    if (y == height) {
      lo = to;
    } else {
      while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (node_index[mid] < want)
          lo = mid + 1;
        else
          hi = mid;
      }
    }
    row_start[t] = lo;
  }
}

// Horizontal runs without atomics: every chunk node points at the first
// node of its run (consecutive indices on one row of one frame), which is
// lower, so roots still point down.
__global__ void link_runs(LabelArgs a, const ChunkDesc *frames,
                          const uint32_t *node_index, const uint32_t *slot,
                          uint32_t *parent, uint32_t *run) {
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < a.n_nodes;
       g += gridDim.x * blockDim.x) {
    if (g < a.n_carry) { // carried nodes already point at their roots
      run[g] = g;
      continue;
    }
    const int k = (int)slot[g] - 1; // chunk frame, from the gather
    const uint32_t from = frames[k].offset;
    uint32_t r = g, idx = node_index[g];
    while (r > from && idx % a.width != 0 && node_index[r - 1] == idx - 1) {
      r--;
      idx--;
    }
    parent[g] = r;
    run[g] = r;
  }
}

// The up and previous-frame edges (src/dials_spots.cc:106-137). An edge is
// skipped when the left neighbour has the same edge one column over: the two
// targets are then neighbours on their own row and already joined, as are
// the node and its left neighbour.
__global__ void unite_edges(LabelArgs a, const ChunkDesc *frames,
                            const uint32_t *node_index, const uint32_t *slot,
                            const uint32_t *row_start, uint32_t *parent) {
  const uint32_t stride = a.pixels / a.width + 1;
  for (uint32_t g = a.n_carry + blockIdx.x * blockDim.x + threadIdx.x;
       g < a.n_nodes; g += gridDim.x * blockDim.x) {
    const int k = (int)slot[g] - 1; // chunk frame, from the gather
    const uint32_t from = frames[k].offset;
    const uint32_t idx = node_index[g];
    const uint32_t y = idx / a.width;
    const uint32_t *rows = row_start + (size_t)(k + 1) * stride;
    const bool left =
        g > from && idx % a.width != 0 && node_index[g - 1] == idx - 1;
    if (y > 0) {
      const uint32_t lo = rows[y - 1];
      const long up = find_index(node_index, lo, rows[y], idx - a.width);
      if (up >= 0 && !(left && (uint32_t)up > lo &&
                       node_index[up - 1] == idx - a.width - 1))
        unite(parent, g, (uint32_t)up);
    }
    if (!a.two_d && frames[k].linked) {
      const uint32_t *prev = row_start + (size_t)k * stride;
      const uint32_t lo = prev[y];
      const long back = find_index(node_index, lo, prev[y + 1], idx);
      if (back >= 0 &&
          !(left && (uint32_t)back > lo && node_index[back - 1] == idx - 1))
        unite(parent, g, (uint32_t)back);
    }
  }
}

// Every union is done: resolve each node's root with a read-only walk into
// its own array. (Writing roots back into `parent` while other threads still
// walk it would let a path-halving write replace a finished root with an
// intermediate ancestor.)
__global__ void resolve_roots(LabelArgs a, const uint32_t *parent,
                              const uint32_t *run, const uint32_t *slot,
                              uint32_t *root, uint32_t *last_slot) {
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < a.n_nodes;
       g += gridDim.x * blockDim.x) {
    if (run[g] != g)
      continue; // resolved from its run's first node below
    uint32_t r = g;
    while (parent[r] != r)
      r = parent[r];
    root[g] = r;
    if (slot[g]) // a run lies on one frame: one update per run
      atomicMax(&last_slot[r], slot[g]);
  }
}

__global__ void resolve_runs(LabelArgs a, const uint32_t *run,
                             uint32_t *root) {
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < a.n_nodes;
       g += gridDim.x * blockDim.x) {
    const uint32_t r = run[g];
    if (r != g)
      root[g] = root[r];
  }
}

// Sort keys: closed components by (last slot, root), the slot shifted left
// by `shift`; open ones last (all bits set). Also the flags of the nodes that
// stay open.
template <typename K>
__global__ void make_keys(LabelArgs a, const uint32_t *root,
                          const uint32_t *last_slot, K *keys,
                          uint32_t *values, uint32_t *open_flag, int shift) {
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < a.n_nodes;
       g += gridDim.x * blockDim.x) {
    const uint32_t r = root[g];
    const uint32_t last = last_slot[r];
    const bool open = !a.closing && !a.two_d && a.n_frames > 0 &&
                      last == a.n_frames;
    keys[g] = open ? ~(K)0 : (K)(((K)last << shift) | r);
    values[g] = g;
    open_flag[g] = open ? 1u : 0u;
  }
}

// start_flag[i] = sorted entry i begins a closed component.
template <typename K>
__global__ void mark_starts(LabelArgs a, const K *keys, K open_key,
                            uint32_t *start_flag) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n_nodes;
       i += gridDim.x * blockDim.x)
    start_flag[i] =
        (keys[i] != open_key && (i == 0 || keys[i] != keys[i - 1])) ? 1u : 0u;
}

template <typename K>
__global__ void list_starts(LabelArgs a, const uint32_t *start_flag,
                            const uint32_t *start_rank, uint32_t *starts,
                            const K *keys, K open_key,
                            uint32_t *n_closed_nodes) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n_nodes;
       i += gridDim.x * blockDim.x) {
    if (start_flag[i])
      starts[start_rank[i]] = i;
    const bool closed = keys[i] != open_key;
    const bool next_open = i + 1 == a.n_nodes || keys[i + 1] == open_key;
    if (closed && next_open)
      *n_closed_nodes = i + 1;
  }
}

struct SpotOut {
  sf_spot spot;
  uint32_t accepted; // 1 if the spot is kept
  uint32_t members;  // its pixel count (when kept)
};

enum Emitted { kAccepted = 0, kTooSmall, kTooLarge, kSeparated };

// The part of the reference's emit() (src/dials_spots.cc:297-360) that
// follows the sums: centroid, variance, the single-frame z rule and the
// separation test. `delta_sq` is filled by `variance_sums`, which must add
// the terms in member order.
template <class VarianceSums>
__device__ int finish_spot(const LabelArgs &a, sf_spot &spot, double sum,
                           double sum_sq, const double (&weighted)[3],
                           const double (&peak)[3], VarianceSums variance_sums) {
  spot.intensity = sum;
  spot.intensity_variance = sum;
  if (sum > 0.0) {
    for (int axis = 0; axis < 3; axis++)
      spot.position[axis] = weighted[axis] / sum;
    double delta_sq[3] = {0.0, 0.0, 0.0};
    variance_sums(spot.position, delta_sq);
    const double twelfth = 1.0 / 12.0;
    if (sum * sum > sum_sq) {
      for (int axis = 0; axis < 3; axis++) {
        const double variance = delta_sq[axis] * sum / (sum * sum - sum_sq);
        spot.variance[axis] = variance / sum + twelfth;
      }
    } else {
      for (int axis = 0; axis < 3; axis++)
        spot.variance[axis] = twelfth;
    }
    if (spot.bbox[5] == spot.bbox[4] + 1)
      spot.position[2] = (double)spot.bbox[4] + 0.5;
  } else {
    spot.position[0] = (spot.bbox[1] + spot.bbox[0]) / 2.0;
    spot.position[1] = (spot.bbox[3] + spot.bbox[2]) / 2.0;
    spot.position[2] = (spot.bbox[5] + spot.bbox[4]) / 2.0;
  }
  if (a.max_separation > 0.0) {
    const double dx = peak[0] - spot.position[0];
    const double dy = peak[1] - spot.position[1];
    const double dz = peak[2] - spot.position[2];
    if (sqrt(dx * dx + dy * dy + dz * dz) > a.max_separation)
      return kSeparated;
  }
  return kAccepted;
}

// One component by one thread: the reference's emit() (src/dials_spots.cc:
// 221-360) operation for operation, every sum in member order.
__device__ int emit_sequential(const LabelArgs &a, uint32_t s0, uint32_t n,
                               const int32_t *mem_frame,
                               const uint32_t *mem_index,
                               const uint32_t *mem_value, sf_spot &spot) {
  spot.n_signal = n;
  double sum = 0.0, sum_sq = 0.0;
  double weighted[3] = {0.0, 0.0, 0.0};
  int32_t x0 = 0, x1 = 0, y0 = 0, y1 = 0, z0 = 0, z1 = 0;
  uint32_t peak_value = 0;
  double peak[3] = {0.0, 0.0, 0.0};
  for (uint32_t k = 0; k < n; k++) {
    const uint32_t g = s0 + k;
    const uint32_t idx = mem_index[g];
    const int32_t x = (int32_t)(idx % a.width);
    const int32_t y = (int32_t)(idx / a.width);
    const int32_t z = mem_frame[g];
    if (k == 0) {
      x0 = x;
      x1 = x + 1;
      y0 = y;
      y1 = y + 1;
      z0 = z;
      z1 = z + 1;
    } else {
      x0 = x < x0 ? x : x0;
      x1 = x + 1 > x1 ? x + 1 : x1;
      y0 = y < y0 ? y : y0;
      y1 = y + 1 > y1 ? y + 1 : y1;
      z0 = z < z0 ? z : z0;
      z1 = z + 1 > z1 ? z + 1 : z1;
    }
    const double value = (double)mem_value[g];
    const double position[3] = {(double)x + 0.5, (double)y + 0.5,
                                (double)z + 0.5};
    sum += value;
    sum_sq += value * value;
    for (int axis = 0; axis < 3; axis++)
      weighted[axis] += value * position[axis];
    if (mem_value[g] > peak_value) {
      peak_value = mem_value[g];
      peak[0] = position[0];
      peak[1] = position[1];
      peak[2] = position[2];
    }
  }
  spot.bbox[0] = x0;
  spot.bbox[1] = x1;
  spot.bbox[2] = y0;
  spot.bbox[3] = y1;
  spot.bbox[4] = z0;
  spot.bbox[5] = z1;
  return finish_spot(a, spot, sum, sum_sq, weighted, peak,
                     [&](const double (&centre)[3], double (&delta_sq)[3]) {
                       for (uint32_t k = 0; k < n; k++) {
                         const uint32_t g = s0 + k;
                         const uint32_t idx = mem_index[g];
                         const double value = (double)mem_value[g];
                         const double position[3] = {
                             (double)(idx % a.width) + 0.5,
                             (double)(idx / a.width) + 0.5,
                             (double)mem_frame[g] + 0.5};
                         for (int axis = 0; axis < 3; axis++) {
                           const double delta = position[axis] - centre[axis];
                           delta_sq[axis] += value * delta * delta;
                         }
                       }
                     });
}

template <typename T> __device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1)
    v += __shfl_xor(v, d, 64);
  return v;
}
template <typename T> __device__ __forceinline__ T wave_min(T v) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const T o = __shfl_xor(v, d, 64);
    v = o < v ? o : v;
  }
  return v;
}
template <typename T> __device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const T o = __shfl_xor(v, d, 64);
    v = o > v ? o : v;
  }
  return v;
}

// The closed members' frame, index and value in emission order (the sorted
// order), gathered once: the finalise and pixel-table passes then read them
// in order instead of through `order`.
__global__ void gather_members(const uint32_t *n_closed, const uint32_t *order,
                               const int32_t *node_frame,
                               const uint32_t *node_index,
                               const uint32_t *node_value, int32_t *mem_frame,
                               uint32_t *mem_index, uint32_t *mem_value) {
  const uint32_t end = *n_closed;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < end;
       i += gridDim.x * blockDim.x) {
    const uint32_t g = order[i];
    mem_frame[i] = node_frame[g];
    mem_index[i] = node_index[g];
    mem_value[i] = node_value[g];
  }
}

// One wave per closed component, in emission order, with the result the
// sequential emit() gives bit for bit:
//  * bbox, peak (first strict maximum in member order) and the sums are
//    taken across the wave. The reference sums in double in member order;
//    when every term and partial sum is an integer (or half-integer) below
//    2^53 -- always so for 16-bit frames -- those double sums are exact and
//    equal the integer sums computed here in any order. Otherwise one lane
//    runs the sequential emit();
//  * the variance terms v (p - c)^2 are formed in parallel, then added in
//    member order, as src/dials_spots.cc:304-316 adds them.
__global__ __launch_bounds__(256) void finalise(
    LabelArgs a, const uint32_t *starts, const uint32_t *n_starts,
    const uint32_t *n_closed, const int32_t *mem_frame,
    const uint32_t *mem_index, const uint32_t *mem_value, SpotOut *out,
    unsigned long long *counts) {
  __shared__ double terms[4][3 * 64]; // a wave's variance terms, by axis
  const uint32_t comps = *n_starts;
  const uint32_t end_all = *n_closed;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  uint32_t tally[4] = {0, 0, 0, 0};
  for (uint32_t c = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
       c < comps; c += waves) {
    const uint32_t s0 = starts[c];
    const uint32_t s1 = c + 1 < comps ? starts[c + 1] : end_all;
    const uint32_t n = s1 - s0;
    SpotOut r;
    memset(&r, 0, sizeof(r));
    int kind = kAccepted;
    if (n < a.min_size) {
      kind = kTooSmall;
    } else if (n > a.max_size) {
      kind = kTooLarge;
    } else {
      // Members in registers: lane holds members lane + 64 j, j < KMAX.
      constexpr int KMAX = 16;
      const bool held = n <= 64u * KMAX;
      uint32_t m_idx[KMAX], m_val[KMAX];
      int32_t m_z[KMAX];
#pragma unroll
      for (int j = 0; j < KMAX; j++) {
        const uint32_t k = lane + 64u * j;
        m_idx[j] = 0;
        m_val[j] = 0;
        m_z[j] = 0;
        if (held && k < n) {
          m_idx[j] = mem_index[s0 + k];
          m_val[j] = mem_value[s0 + k];
          m_z[j] = mem_frame[s0 + k];
        }
      }
      // pass 1: bbox, peak, integer sums, bounds
      int32_t x0 = INT32_MAX, x1 = INT32_MIN, y0 = INT32_MAX, y1 = INT32_MIN,
              z0 = INT32_MAX, z1 = INT32_MIN;
      uint64_t sv = 0, svv = 0, sx = 0, sy = 0;
      int64_t sz = 0;
      uint32_t vmax = 0, kmax = 0xFFFFFFFFu;
      int32_t zabs = 0;
      auto visit = [&](uint32_t k, uint32_t idx, uint32_t v, int32_t z) {
        const int32_t x = (int32_t)(idx % a.width);
        const int32_t y = (int32_t)(idx / a.width);
        x0 = min(x0, x);
        x1 = max(x1, x + 1);
        y0 = min(y0, y);
        y1 = max(y1, y + 1);
        z0 = min(z0, z);
        z1 = max(z1, z + 1);
        sv += v;
        svv += (uint64_t)v * v; // exact while v < 2^32
        sx += (uint64_t)v * (uint64_t)(2 * x + 1);
        sy += (uint64_t)v * (uint64_t)(2 * y + 1);
        sz += (int64_t)v * (int64_t)(2 * (int64_t)z + 1);
        zabs = max(zabs, z < 0 ? -z : z);
        if (v > vmax) { // lanes see members in order, so first per lane
          vmax = v;
          kmax = k;
        }
      };
      if (held) {
#pragma unroll
        for (int j = 0; j < KMAX; j++) {
          const uint32_t k = lane + 64u * j;
          if (k < n)
            visit(k, m_idx[j], m_val[j], m_z[j]);
        }
      } else {
        for (uint32_t k = lane; k < n; k += 64)
          visit(k, mem_index[s0 + k], mem_value[s0 + k], mem_frame[s0 + k]);
      }
      x0 = wave_min(x0);
      x1 = wave_max(x1);
      y0 = wave_min(y0);
      y1 = wave_max(y1);
      z0 = wave_min(z0);
      z1 = wave_max(z1);
      zabs = wave_max(zabs);
      const uint32_t vtop = wave_max(vmax);
      const uint32_t ktop = wave_min(vmax == vtop ? kmax : 0xFFFFFFFFu);
      sf_spot &spot = r.spot;
      // Exactness of the reference's double sums: every term and partial
      // sum below 2^52 (with margin) in magnitude.
      const double bound = 4503599627370496.0; // 2^52
      const double top = (double)vtop, cnt = (double)n;
      const double reach = (double)max(max(a.width, a.pixels / a.width),
                                       (uint32_t)zabs) * 2.0 + 1.0;
      const bool exact = cnt * top * top < bound && cnt * top * reach < bound;
      if (!exact) {
        if (lane == 0)
          kind = emit_sequential(a, s0, n, mem_frame, mem_index, mem_value,
                                 spot);
        kind = __shfl(kind, 0, 64);
      } else {
        sv = wave_sum(sv);
        svv = wave_sum(svv);
        sx = wave_sum(sx);
        sy = wave_sum(sy);
        sz = wave_sum(sz);
        spot.n_signal = n;
        spot.bbox[0] = x0;
        spot.bbox[1] = x1;
        spot.bbox[2] = y0;
        spot.bbox[3] = y1;
        spot.bbox[4] = z0;
        spot.bbox[5] = z1;
        const double sum = (double)sv, sum_sq = (double)svv;
        const double weighted[3] = {(double)sx * 0.5, (double)sy * 0.5,
                                    (double)sz * 0.5};
        double peak[3] = {0.0, 0.0, 0.0};
        if (vtop > 0) {
          const uint32_t idx = mem_index[s0 + ktop];
          peak[0] = (double)(idx % a.width) + 0.5;
          peak[1] = (double)(idx / a.width) + 0.5;
          peak[2] = (double)mem_frame[s0 + ktop] + 0.5;
        }
        // the variance terms of one 64-member chunk, added in member order
        auto chunk_terms = [&](uint32_t k0, uint32_t idx, uint32_t v,
                               int32_t z, const double (&centre)[3],
                               double (&delta_sq)[3]) {
          double t[3] = {0.0, 0.0, 0.0};
          if (k0 + lane < n) {
            const double value = (double)v;
            const double position[3] = {(double)(idx % a.width) + 0.5,
                                        (double)(idx / a.width) + 0.5,
                                        (double)z + 0.5};
            for (int axis = 0; axis < 3; axis++) {
              const double delta = position[axis] - centre[axis];
              t[axis] = value * delta * delta;
            }
          }
          // terms through LDS, then every lane adds them in member order
          // (the same sums in every lane; the reads are broadcasts)
          double *mine = terms[threadIdx.x >> 6];
          for (int axis = 0; axis < 3; axis++)
            mine[axis * 64 + lane] = t[axis];
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
          const uint32_t m = min(64u, n - k0);
          for (uint32_t i = 0; i < m; i++)
            for (int axis = 0; axis < 3; axis++)
              delta_sq[axis] += mine[axis * 64 + i];
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        };
        kind = finish_spot(
            a, spot, sum, sum_sq, weighted, peak,
            [&](const double (&centre)[3], double (&delta_sq)[3]) {
              if (held) {
#pragma unroll
                for (int j = 0; j < KMAX; j++)
                  if (64u * j < n)
                    chunk_terms(64u * j, m_idx[j], m_val[j], m_z[j], centre,
                                delta_sq);
              } else {
                for (uint32_t k0 = 0; k0 < n; k0 += 64) {
                  uint32_t idx = 0, v = 0;
                  int32_t z = 0;
                  if (k0 + lane < n) {
                    idx = mem_index[s0 + k0 + lane];
                    v = mem_value[s0 + k0 + lane];
                    z = mem_frame[s0 + k0 + lane];
                  }
                  chunk_terms(k0, idx, v, z, centre, delta_sq);
                }
              }
            });
      }
    }
    if (kind == kAccepted) {
      r.accepted = 1;
      r.members = n;
    }
    if (lane == 0) {
      out[c] = r;
      tally[kind]++;
    }
  }
  // counts: groups, too_small, too_large, separated, accepted -- summed
  // over the block first, one atomic per counter and block
  __shared__ uint32_t block_tally[4];
  if (threadIdx.x < 4)
    block_tally[threadIdx.x] = 0;
  __syncthreads();
  if (lane == 0)
    for (int t = 0; t < 4; t++)
      if (tally[t])
        atomicAdd(&block_tally[t], tally[t]);
  __syncthreads();
  if (threadIdx.x == 0) {
    const int slot[4] = {5, 2, 3, 4}; // kAccepted, kTooSmall, ...
    for (int t = 0; t < 4; t++)
      if (block_tally[t])
        atomicAdd(&counts[slot[t]], (unsigned long long)block_tally[t]);
    if (blockIdx.x == 0)
      atomicAdd(&counts[1], (unsigned long long)comps);
  }
}

// Per component: accepted << 32 | kept members, so that one exclusive scan
// gives both the spot rank (high half) and the pixel rank (low half; the
// members of a chunk stay below 2^32).
__global__ void spot_flags(const SpotOut *res, const uint32_t *n_comps,
                           unsigned long long *packed) {
  const uint32_t comps = *n_comps;
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < comps;
       c += gridDim.x * blockDim.x)
    packed[c] = ((unsigned long long)res[c].accepted << 32) | res[c].members;
}

// Appends the kept spots after what earlier chunks wrote.
__global__ void write_spots(const SpotOut *res, const uint32_t *n_comps,
                            const unsigned long long *rank,
                            const unsigned long long *totals,
                            sf_spot *spots) {
  const uint32_t comps = *n_comps;
  const unsigned long long spot_base = totals[0], pixel_base = totals[1];
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < comps;
       c += gridDim.x * blockDim.x) {
    if (!res[c].accepted)
      continue;
    sf_spot s = res[c].spot;
    s.first = pixel_base + (uint32_t)rank[c];
    spots[spot_base + (rank[c] >> 32)] = s;
  }
}

// Appends the kept spots' pixels, one thread per member: entry i of the
// emission order belongs to component start_rank[i] + start_flag[i] - 1 and
// is member i - starts[c] of it.
__global__ void write_pixels(const SpotOut *res, const uint32_t *n_closed,
                             const uint32_t *start_flag,
                             const uint32_t *start_rank,
                             const uint32_t *starts,
                             const unsigned long long *rank,
                             const int32_t *mem_frame,
                             const uint32_t *mem_index,
                             const uint32_t *mem_value,
                             const unsigned long long *totals,
                             sf_pixel *pixels) {
  const uint32_t end = *n_closed;
  const unsigned long long pixel_base = totals[1];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < end;
       i += gridDim.x * blockDim.x) {
    const uint32_t c = start_rank[i] + start_flag[i] - 1;
    if (!res[c].accepted)
      continue;
    sf_pixel p;
    p.frame = mem_frame[i];
    p.index = mem_index[i];
    p.value = mem_value[i];
    pixels[pixel_base + (uint32_t)rank[c] + (i - starts[c])] = p;
  }
}

__global__ void advance_totals(const uint32_t *n_comps,
                               const unsigned long long *rank,
                               const SpotOut *res,
                               unsigned long long *totals) {
  const uint32_t comps = *n_comps;
  if (comps) {
    const uint32_t last = comps - 1;
    totals[0] += (rank[last] >> 32) + res[last].accepted;
    totals[1] += (uint32_t)rank[last] + res[last].members;
  }
}

// The open nodes, in node order, become the next frontier; each points at
// its component's first open node (the root, which is open with it).
__global__ void carry_out(LabelArgs a, const uint32_t *open_flag,
                          const uint32_t *open_rank, const uint32_t *root,
                          const int32_t *node_frame,
                          const uint32_t *node_index,
                          const uint32_t *node_value, int32_t *carry_frame,
                          uint32_t *carry_index, uint32_t *carry_value,
                          uint32_t *carry_root, uint32_t *n_carry_out) {
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < a.n_nodes;
       g += gridDim.x * blockDim.x) {
    if (g + 1 == a.n_nodes)
      *n_carry_out = open_rank[g] + open_flag[g];
    if (!open_flag[g])
      continue;
    const uint32_t at = open_rank[g];
    carry_frame[at] = node_frame[g];
    carry_index[at] = node_index[g];
    carry_value[at] = node_value[g];
    carry_root[at] = open_rank[root[g]];
  }
}

__global__ void count_components(const uint32_t *rank, const uint32_t *flag,
                                 uint32_t n, uint32_t *out) {
  *out = rank[n - 1] + flag[n - 1];
}

__global__ void add_signal_count(const ChunkDesc *frames, uint32_t n_frames,
                                 unsigned long long *counts) {
  unsigned long long total = 0;
  for (uint32_t k = 0; k < n_frames; k++) {
    const uint32_t c = *frames[k].count;
    total += c <= frames[k].capacity ? c : 0;
  }
  counts[0] += total;
}

} // namespace sf

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
namespace {
thread_local std::string g_label_error;
} // namespace

namespace sf {

template <typename T> struct DevArray {
  T *ptr = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t n) {
    if (n <= cap)
      return hipSuccess;
    size_t want = cap ? cap : 1024;
    while (want < n)
      want *= 2;
    T *fresh = nullptr;
    const hipError_t e = hipMalloc(&fresh, want * sizeof(T));
    if (e != hipSuccess)
      return e;
    if (ptr)
      (void)hipFree(ptr);
    ptr = fresh;
    cap = want;
    return hipSuccess;
  }
  // grow keeping the first `keep` elements (stream must be idle)
  hipError_t grow_keep(size_t n, size_t keep) {
    if (n <= cap)
      return hipSuccess;
    size_t want = cap ? cap : 1024;
    while (want < n)
      want *= 2;
    T *fresh = nullptr;
    hipError_t e = hipMalloc(&fresh, want * sizeof(T));
    if (e != hipSuccess)
      return e;
    if (ptr && keep)
      e = hipMemcpy(fresh, ptr, keep * sizeof(T), hipMemcpyDeviceToDevice);
    if (ptr)
      (void)hipFree(ptr);
    ptr = fresh;
    cap = want;
    return e;
  }
  void release() {
    if (ptr)
      (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
  }
};

} // namespace sf

struct sf_labeller {
  size_t height = 0, width = 0;
  sf_options options{};
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool started = false;
  int64_t last_frame = 0;
  // frontier (double-buffered)
  sf::DevArray<int32_t> carry_frame[2];
  sf::DevArray<uint32_t> carry_index[2], carry_value[2], carry_root[2];
  int carry_side = 0;
  uint32_t n_carry = 0;      // host copy, valid after a sync
  uint32_t carry_tail = 0;   // first carried node on the last frame
  uint32_t last_count = 0;   // pixels of the last labelled frame
  uint32_t *d_n_carry = nullptr;
  uint32_t *h_scalars = nullptr; // pinned: n_carry
  // node arrays
  sf::DevArray<int32_t> node_frame;
  sf::DevArray<uint32_t> node_index, node_value, node_slot, parent, root,
      last_slot;
  sf::DevArray<uint64_t> keys, keys_sorted;
  sf::DevArray<uint32_t> row_start; // per (slot, row): first node
  sf::DevArray<uint32_t> values, order, open_flag, open_rank, start_flag,
      start_rank, starts;
  sf::DevArray<sf::SpotOut> results;
  sf::DevArray<unsigned long long> spot_pack, spot_rank;
  sf::DevArray<uint8_t> temp;
  sf::DevArray<sf::ChunkDesc> chunk;
  uint32_t *d_scalars = nullptr; // n_starts, n_closed
  // results
  sf::DevArray<sf_spot> spots;
  sf::DevArray<sf_pixel> pixels;
  unsigned long long *d_counts = nullptr; // 6
  unsigned long long *d_totals = nullptr; // spots, pixels written
  uint64_t bound_spots = 0, bound_pixels = 0;
  // host-list frames waiting for the device
  std::vector<sf_signal_pixel> pending;
  std::vector<int64_t> pending_frames;
  std::vector<uint32_t> pending_counts;
  sf::DevArray<sf_signal_pixel> pending_dev;
  sf::DevArray<uint32_t> pending_count_dev;
  std::vector<sf::ChunkDesc> host_chunk;
};

namespace {

int lab_fail(const char *what, hipError_t e) {
  g_label_error = std::string(what) + ": " + hipGetErrorString(e);
  return SF_DEVICE_ERROR;
}

#define LAB_CHECK(call, what)                                                  \
  do {                                                                         \
    const hipError_t e_ = (call);                                              \
    if (e_ != hipSuccess)                                                      \
      return lab_fail(what, e_);                                               \
  } while (0)

uint32_t grid_for(size_t n) {
  const size_t blocks = (n + 255) / 256;
  return (uint32_t)std::min<size_t>(std::max<size_t>(blocks, 1), 65535);
}

// The component sort on 32-bit keys: keys of 25 to 30 bits go in three
// 10-bit passes instead of four 8-bit ones.
using Sort10 = rocprim::radix_sort_config<
    rocprim::default_config, rocprim::default_config,
    rocprim::radix_sort_onesweep_config<
        rocprim::kernel_config<1024, 16>, rocprim::kernel_config<1024, 16>, 10,
        rocprim::block_radix_rank_algorithm::match>>;

hipError_t sort32(void *temp, size_t &bytes, const uint32_t *keys_in,
                  uint32_t *keys_out, const uint32_t *values_in,
                  uint32_t *values_out, uint32_t n, int bits,
                  hipStream_t s) {
  if (bits > 24 && bits <= 30)
    return rocprim::radix_sort_pairs<Sort10>(temp, bytes, keys_in, keys_out,
                                             values_in, values_out, n, 0, bits,
                                             s);
  return rocprim::radix_sort_pairs(temp, bytes, keys_in, keys_out, values_in,
                                   values_out, n, 0, bits, s);
}

// Label one chunk whose frame lists are device-resident. `frames` holds
// list/count/capacity/number for each frame; offsets are filled here from
// the host counts.
int run_chunk(sf_labeller *l, std::vector<sf::ChunkDesc> &frames,
              const std::vector<uint32_t> &counts, bool closing,
              hipStream_t producer) {
  hipStream_t s = l->stream;
  // the frontier and the scratch of the previous chunk must be done with
  LAB_CHECK(hipStreamSynchronize(s), "waiting for the previous chunk");
  l->n_carry = *l->h_scalars;
  // Every pixel of the last labelled frame is in an open component, so the
  // frontier ends with exactly that frame's pixels.
  l->carry_tail = l->n_carry >= l->last_count ? l->n_carry - l->last_count : 0;
  const uint32_t n_frames = (uint32_t)frames.size();
  uint64_t total = l->n_carry;
  for (uint32_t k = 0; k < n_frames; k++) {
    frames[k].offset = (uint32_t)total;
    const uint32_t c = counts[k] <= frames[k].capacity ? counts[k] : 0;
    total += c;
  }
  if (total > 0xF0000000ull) {
    g_label_error = "more than 4 billion pixels held open at once";
    return SF_BAD_ARGUMENT;
  }
  const uint32_t n = (uint32_t)total;
  sf::LabelArgs a{};
  a.n_nodes = n;
  a.n_carry = l->n_carry;
  a.carry_tail = l->carry_tail;
  a.n_frames = n_frames;
  a.width = (uint32_t)l->width;
  a.pixels = (uint32_t)(l->height * l->width);
  a.closing = closing ? 1 : 0;
  a.first_linked = n_frames && frames[0].linked;
  a.two_d = l->options.two_d;
  a.min_size = l->options.min_spot_size;
  a.max_size = l->options.max_spot_size;
  a.max_separation = l->options.max_separation;
  // outputs: at worst every node is kept
  l->bound_pixels += n - l->n_carry;
  l->bound_spots += (n - l->n_carry) / std::max<uint64_t>(1, a.min_size) + 1;
  {
    unsigned long long totals[2] = {0, 0};
    // on the labeller's own stream (idle here), so that no null-stream copy
    // waits for other streams' work
    LAB_CHECK(hipMemcpyAsync(totals, l->d_totals, sizeof(totals),
                             hipMemcpyDeviceToHost, s),
              "reading totals");
    LAB_CHECK(hipStreamSynchronize(s), "reading totals");
    LAB_CHECK(l->spots.grow_keep(totals[0] + n / std::max<uint64_t>(
                                                  1, a.min_size) + 1,
                                 totals[0]),
              "growing the spot table");
    LAB_CHECK(l->pixels.grow_keep(totals[1] + n + 1, totals[1]),
              "growing the pixel table");
  }
  l->last_count = 0;
  if (n_frames && !closing && !l->options.two_d)
    l->last_count = counts[n_frames - 1] <= frames[n_frames - 1].capacity
                        ? counts[n_frames - 1]
                        : 0;
  if (n == 0) {
    *l->h_scalars = 0;
    if (n_frames) {
      LAB_CHECK(l->chunk.ensure(n_frames), "chunk table");
      LAB_CHECK(hipMemcpyAsync(l->chunk.ptr, frames.data(),
                               n_frames * sizeof(sf::ChunkDesc),
                               hipMemcpyHostToDevice, s),
                "chunk table");
      sf::add_signal_count<<<1, 1, 0, s>>>(l->chunk.ptr, n_frames,
                                           l->d_counts);
    }
    LAB_CHECK(hipMemsetAsync(l->d_n_carry, 0, sizeof(uint32_t), s), "memset");
    return SF_OK;
  }
  LAB_CHECK(l->node_frame.ensure(n), "alloc");
  LAB_CHECK(l->node_index.ensure(n), "alloc");
  LAB_CHECK(l->node_value.ensure(n), "alloc");
  LAB_CHECK(l->node_slot.ensure(n), "alloc");
  LAB_CHECK(l->parent.ensure(n), "alloc");
  LAB_CHECK(l->root.ensure(n), "alloc");
  LAB_CHECK(l->last_slot.ensure(n), "alloc");
  LAB_CHECK(l->keys.ensure(n), "alloc");
  LAB_CHECK(l->keys_sorted.ensure(n), "alloc");
  LAB_CHECK(l->values.ensure(n), "alloc");
  LAB_CHECK(l->order.ensure(n), "alloc");
  LAB_CHECK(l->open_flag.ensure(n), "alloc");
  LAB_CHECK(l->open_rank.ensure(n), "alloc");
  LAB_CHECK(l->start_flag.ensure(n), "alloc");
  LAB_CHECK(l->start_rank.ensure(n), "alloc");
  LAB_CHECK(l->starts.ensure(n), "alloc");
  LAB_CHECK(l->results.ensure(n), "alloc");
  LAB_CHECK(l->spot_pack.ensure(n), "alloc");
  LAB_CHECK(l->spot_rank.ensure(n), "alloc");
  const int in = l->carry_side, outside = 1 - l->carry_side;
  LAB_CHECK(l->carry_frame[outside].ensure(n), "alloc");
  LAB_CHECK(l->carry_index[outside].ensure(n), "alloc");
  LAB_CHECK(l->carry_value[outside].ensure(n), "alloc");
  LAB_CHECK(l->carry_root[outside].ensure(n), "alloc");
  LAB_CHECK(l->chunk.ensure(std::max<uint32_t>(n_frames, 1)), "alloc");
  // temp storage: the largest of the rocPRIM calls
  size_t need = 0, bytes = 0;
  // Sort keys: 32 bits when (last slot, root) fit, else 64.
  int root_bits = 1, slot_bits = 1;
  while (root_bits < 32 && (1ull << root_bits) < (uint64_t)n)
    root_bits++;
  while ((1ull << slot_bits) < (uint64_t)n_frames + 2)
    slot_bits++;
  const bool narrow = root_bits + slot_bits <= 32;
  uint32_t *keys32 = reinterpret_cast<uint32_t *>(l->keys.ptr);
  uint32_t *keys32_sorted = reinterpret_cast<uint32_t *>(l->keys_sorted.ptr);
  if (narrow)
    LAB_CHECK(sort32(nullptr, bytes, keys32, keys32_sorted, l->values.ptr,
                     l->order.ptr, n, root_bits + slot_bits, s),
              "sort size");
  else
    LAB_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, l->keys.ptr,
                                        l->keys_sorted.ptr, l->values.ptr,
                                        l->order.ptr, n, 0, 64, s),
              "sort size");
  need = std::max(need, bytes);
  LAB_CHECK(rocprim::exclusive_scan(nullptr, bytes, l->open_flag.ptr,
                                    l->open_rank.ptr, 0u, (size_t)n,
                                    rocprim::plus<uint32_t>(), s),
            "scan size");
  need = std::max(need, bytes);
  LAB_CHECK(rocprim::exclusive_scan(nullptr, bytes, l->spot_pack.ptr,
                                    l->spot_rank.ptr, 0ull, (size_t)n,
                                    rocprim::plus<unsigned long long>(), s),
            "scan size");
  need = std::max(need, bytes);
  LAB_CHECK(l->temp.ensure(need + 256), "alloc temp");

  if (n_frames)
    LAB_CHECK(hipMemcpyAsync(l->chunk.ptr, frames.data(),
                             n_frames * sizeof(sf::ChunkDesc),
                             hipMemcpyHostToDevice, s),
              "chunk table");
  // The gather reads the producer's lists: run it on the producer's stream so
  // that the producer cannot overwrite them first, and have this stream wait
  // for it.
  hipStream_t gs = producer ? producer : s;
  if (producer) {
    LAB_CHECK(hipEventRecord(l->done, s), "event");
    LAB_CHECK(hipStreamWaitEvent(gs, l->done, 0), "event wait");
  }
  const uint32_t grid = grid_for(n);
  sf::gather_nodes<<<grid, 256, 0, gs>>>(
      a, l->chunk.ptr, l->carry_frame[in].ptr, l->carry_index[in].ptr,
      l->carry_value[in].ptr, l->carry_root[in].ptr, l->node_frame.ptr,
      l->node_index.ptr, l->node_value.ptr, l->node_slot.ptr, l->parent.ptr,
      l->last_slot.ptr);
  if (n_frames)
    sf::add_signal_count<<<1, 1, 0, gs>>>(l->chunk.ptr, n_frames,
                                          l->d_counts);
  LAB_CHECK(hipGetLastError(), "gather");
  if (producer) {
    LAB_CHECK(hipEventRecord(l->done, gs), "event");
    LAB_CHECK(hipStreamWaitEvent(s, l->done, 0), "event wait");
  }
  {
    const size_t rows = (size_t)(n_frames + 1) * (l->height + 1);
    LAB_CHECK(l->row_start.ensure(rows), "alloc");
    sf::fill_rows<<<grid_for(rows), 256, 0, s>>>(a, l->chunk.ptr,
                                                 l->node_index.ptr,
                                                 l->row_start.ptr);
  }
  // `values` holds each node's run start until make_keys reuses it
  sf::link_runs<<<grid, 256, 0, s>>>(a, l->chunk.ptr, l->node_index.ptr,
                                     l->node_slot.ptr, l->parent.ptr,
                                     l->values.ptr);
  sf::unite_edges<<<grid, 256, 0, s>>>(a, l->chunk.ptr, l->node_index.ptr,
                                       l->node_slot.ptr, l->row_start.ptr,
                                       l->parent.ptr);
  sf::resolve_roots<<<grid, 256, 0, s>>>(a, l->parent.ptr, l->values.ptr,
                                         l->node_slot.ptr, l->root.ptr,
                                         l->last_slot.ptr);
  sf::resolve_runs<<<grid, 256, 0, s>>>(a, l->values.ptr, l->root.ptr);
  if (narrow)
    sf::make_keys<uint32_t><<<grid, 256, 0, s>>>(
        a, l->root.ptr, l->last_slot.ptr, keys32, l->values.ptr,
        l->open_flag.ptr, root_bits);
  else
    sf::make_keys<uint64_t><<<grid, 256, 0, s>>>(
        a, l->root.ptr, l->last_slot.ptr, l->keys.ptr, l->values.ptr,
        l->open_flag.ptr, 32);
  LAB_CHECK(hipGetLastError(), "label kernels");
  bytes = l->temp.cap;
  if (narrow) {
    LAB_CHECK(sort32(l->temp.ptr, bytes, keys32, keys32_sorted,
                     l->values.ptr, l->order.ptr, n, root_bits + slot_bits,
                     s),
              "sorting components");
    // closed keys stay below 2^(root_bits + slot_bits) - 1, so the open key
    // (all ones) sorts last on those bits too
    sf::mark_starts<uint32_t><<<grid, 256, 0, s>>>(a, keys32_sorted,
                                                   ~0u, l->start_flag.ptr);
  } else {
    LAB_CHECK(rocprim::radix_sort_pairs(l->temp.ptr, bytes, l->keys.ptr,
                                        l->keys_sorted.ptr, l->values.ptr,
                                        l->order.ptr, n, 0, 64, s),
              "sorting components");
    sf::mark_starts<uint64_t><<<grid, 256, 0, s>>>(a, l->keys_sorted.ptr,
                                                   ~0ull, l->start_flag.ptr);
  }
  bytes = l->temp.cap;
  LAB_CHECK(rocprim::exclusive_scan(l->temp.ptr, bytes, l->start_flag.ptr,
                                    l->start_rank.ptr, 0u, (size_t)n,
                                    rocprim::plus<uint32_t>(), s),
            "ranking components");
  LAB_CHECK(hipMemsetAsync(l->d_scalars, 0, 2 * sizeof(uint32_t), s),
            "memset");
  if (narrow)
    sf::list_starts<uint32_t><<<grid, 256, 0, s>>>(
        a, l->start_flag.ptr, l->start_rank.ptr, l->starts.ptr, keys32_sorted,
        ~0u, l->d_scalars + 1);
  else
    sf::list_starts<uint64_t><<<grid, 256, 0, s>>>(
        a, l->start_flag.ptr, l->start_rank.ptr, l->starts.ptr,
        l->keys_sorted.ptr, ~0ull, l->d_scalars + 1);
  sf::count_components<<<1, 1, 0, s>>>(l->start_rank.ptr, l->start_flag.ptr,
                                       n, l->d_scalars);
  // the members in emission order, in the sort's input arrays (dead now)
  int32_t *mem_frame = reinterpret_cast<int32_t *>(l->values.ptr);
  uint32_t *mem_index = reinterpret_cast<uint32_t *>(l->keys.ptr);
  uint32_t *mem_value = mem_index + n;
  sf::gather_members<<<grid, 256, 0, s>>>(l->d_scalars + 1, l->order.ptr,
                                          l->node_frame.ptr,
                                          l->node_index.ptr,
                                          l->node_value.ptr, mem_frame,
                                          mem_index, mem_value);
  // a bounded grid: each wave loops over components, each block adds its
  // tallies once
  sf::finalise<<<std::min<uint32_t>(grid, 1024), 256, 0, s>>>(
      a, l->starts.ptr, l->d_scalars, l->d_scalars + 1, mem_frame, mem_index,
      mem_value, l->results.ptr, l->d_counts);
  sf::spot_flags<<<grid, 256, 0, s>>>(l->results.ptr, l->d_scalars,
                                      l->spot_pack.ptr);
  bytes = l->temp.cap;
  LAB_CHECK(rocprim::exclusive_scan(l->temp.ptr, bytes, l->spot_pack.ptr,
                                    l->spot_rank.ptr, 0ull, (size_t)n,
                                    rocprim::plus<unsigned long long>(), s),
            "ranking spots and their pixels");
  sf::write_spots<<<grid, 256, 0, s>>>(l->results.ptr, l->d_scalars,
                                       l->spot_rank.ptr, l->d_totals,
                                       l->spots.ptr);
  sf::write_pixels<<<grid, 256, 0, s>>>(
      l->results.ptr, l->d_scalars + 1, l->start_flag.ptr, l->start_rank.ptr,
      l->starts.ptr, l->spot_rank.ptr, mem_frame, mem_index, mem_value,
      l->d_totals, l->pixels.ptr);
  sf::advance_totals<<<1, 1, 0, s>>>(l->d_scalars, l->spot_rank.ptr,
                                     l->results.ptr, l->d_totals);
  bytes = l->temp.cap;
  LAB_CHECK(rocprim::exclusive_scan(l->temp.ptr, bytes, l->open_flag.ptr,
                                    l->open_rank.ptr, 0u, (size_t)n,
                                    rocprim::plus<uint32_t>(), s),
            "ranking the frontier");
  sf::carry_out<<<grid, 256, 0, s>>>(
      a, l->open_flag.ptr, l->open_rank.ptr, l->root.ptr, l->node_frame.ptr,
      l->node_index.ptr, l->node_value.ptr, l->carry_frame[outside].ptr,
      l->carry_index[outside].ptr, l->carry_value[outside].ptr,
      l->carry_root[outside].ptr, l->d_n_carry);
  LAB_CHECK(hipGetLastError(), "finalise kernels");
  LAB_CHECK(hipMemcpyAsync(l->h_scalars, l->d_n_carry, sizeof(uint32_t),
                           hipMemcpyDeviceToHost, s),
            "reading the frontier size");
  l->carry_side = outside;
  return SF_OK;
}

} // namespace

namespace sf {
// Defined with the plans: the device lists, counts and capacities of a
// plan's frames, and the stream their last run was enqueued on.
int plan_lists(sf_plan *plan, std::vector<const sf_signal_pixel *> &lists,
               std::vector<const uint32_t *> &counts,
               std::vector<uint32_t> &capacity, hipStream_t *stream);
} // namespace sf

namespace {

// Processes the frames added from the host since the last flush.
int flush_pending(sf_labeller *l) {
  if (l->pending_frames.empty())
    return SF_OK;
  LAB_CHECK(hipStreamSynchronize(l->stream), "waiting for the labeller");
  const size_t total = l->pending.size();
  LAB_CHECK(l->pending_dev.ensure(std::max<size_t>(total, 1)), "alloc");
  LAB_CHECK(l->pending_count_dev.ensure(l->pending_counts.size()), "alloc");
  if (total)
    LAB_CHECK(hipMemcpy(l->pending_dev.ptr, l->pending.data(),
                        total * sizeof(sf_signal_pixel),
                        hipMemcpyHostToDevice),
              "uploading signal lists");
  LAB_CHECK(hipMemcpy(l->pending_count_dev.ptr, l->pending_counts.data(),
                      l->pending_counts.size() * sizeof(uint32_t),
                      hipMemcpyHostToDevice),
            "uploading counts");
  std::vector<sf::ChunkDesc> frames(l->pending_frames.size());
  size_t at = 0;
  for (size_t k = 0; k < frames.size(); k++) {
    frames[k].list = l->pending_dev.ptr + at;
    frames[k].count = l->pending_count_dev.ptr + k;
    frames[k].capacity = l->pending_counts[k];
    frames[k].number = (int32_t)l->pending_frames[k];
    frames[k].linked = 0;
    at += l->pending_counts[k];
  }
  // link flags: consecutive frame numbers; frame 0 against the frontier
  for (size_t k = 0; k < frames.size(); k++) {
    const int64_t before = k ? l->pending_frames[k - 1] : l->last_frame;
    const bool have_before = k > 0 || l->started;
    frames[k].linked = have_before && l->pending_frames[k] == before + 1;
  }
  const std::vector<uint32_t> counts = l->pending_counts;
  const int64_t last = l->pending_frames.back();
  const int status = run_chunk(l, frames, counts, false, nullptr);
  l->pending.clear();
  l->pending_frames.clear();
  l->pending_counts.clear();
  if (status != SF_OK)
    return status;
  l->started = true;
  l->last_frame = last;
  return SF_OK;
}

} // namespace

namespace sf {
// Label frames whose ordered lists are already on the device, any frame
// numbers in ascending order (gaps close components as in
// src/dials_spots.cc:68-69). `producer` is the stream the lists were written
// on; the gather runs there, so the producer may reuse the lists once work
// it enqueues afterwards runs. The serving loop's path into the labeller.
int labeller_add_lists(sf_labeller *l, int n, const int64_t *numbers,
                       const sf_signal_pixel *const *lists,
                       const uint32_t *const *count_ptrs,
                       const uint32_t *capacities, const uint32_t *counts,
                       hipStream_t producer) {
  if (!l || n <= 0 || !numbers || !lists || !count_ptrs || !capacities ||
      !counts) {
    g_label_error = "labeller_add_lists: bad argument";
    return SF_BAD_ARGUMENT;
  }
  int status = flush_pending(l);
  if (status != SF_OK)
    return status;
  for (int k = 0; k < n; k++) {
    const int64_t before = k ? numbers[k - 1] : l->last_frame;
    if ((k > 0 || l->started) && numbers[k] <= before) {
      g_label_error = "frames must be added in ascending order";
      return SF_ORDER_ERROR;
    }
  }
  std::vector<sf::ChunkDesc> frames(n);
  std::vector<uint32_t> host_counts(counts, counts + n);
  for (int k = 0; k < n; k++) {
    frames[k].list = lists[k];
    frames[k].count = count_ptrs[k];
    frames[k].capacity = capacities[k];
    frames[k].number = (int32_t)numbers[k];
    const int64_t before = k ? numbers[k - 1] : l->last_frame;
    frames[k].linked = (k > 0 || l->started) && numbers[k] == before + 1;
  }
  status = run_chunk(l, frames, host_counts, false, producer);
  if (status != SF_OK)
    return status;
  l->started = true;
  l->last_frame = numbers[n - 1];
  return SF_OK;
}

hipStream_t labeller_stream(sf_labeller *l) { return l ? l->stream : nullptr; }

} // namespace sf

extern "C" {

sf_labeller *sf_hip_labeller_create(size_t height, size_t width,
                                    const sf_options *options) {
  sf_options o;
  if (options) {
    o = *options;
  } else {
    o.min_spot_size = 3;
    o.max_spot_size = 1000;
    o.max_separation = 2.0;
    o.two_d = 0;
  }
  // the constructor's three refusals (src/dials_spots.cc:22-30)
  if (height == 0 || width == 0 || o.max_spot_size < o.min_spot_size ||
      o.min_spot_size == 0 || height > 0x7fffffffu / width) {
    g_label_error = "labeller options rejected";
    return nullptr;
  }
  sf_labeller *l = new sf_labeller;
  l->height = height;
  l->width = width;
  l->options = o;
  bool ok =
      hipStreamCreateWithFlags(&l->stream, hipStreamNonBlocking) == hipSuccess;
  ok = ok && hipEventCreateWithFlags(&l->done, hipEventDisableTiming) ==
                 hipSuccess;
  ok = ok && hipMalloc(&l->d_n_carry, sizeof(uint32_t)) == hipSuccess;
  ok = ok && hipMalloc(&l->d_scalars, 4 * sizeof(uint32_t)) == hipSuccess;
  ok = ok && hipMalloc(&l->d_counts, 6 * sizeof(unsigned long long)) ==
                 hipSuccess;
  ok = ok && hipMalloc(&l->d_totals, 2 * sizeof(unsigned long long)) ==
                 hipSuccess;
  ok = ok && hipHostMalloc(&l->h_scalars, 4 * sizeof(uint32_t)) == hipSuccess;
  if (ok) {
    ok = hipMemset(l->d_counts, 0, 6 * sizeof(unsigned long long)) ==
             hipSuccess &&
         hipMemset(l->d_totals, 0, 2 * sizeof(unsigned long long)) ==
             hipSuccess &&
         hipMemset(l->d_n_carry, 0, sizeof(uint32_t)) == hipSuccess;
    l->h_scalars[0] = 0;
  }
  if (!ok) {
    g_label_error = "labeller: device allocation failed";
    sf_hip_labeller_destroy(l);
    return nullptr;
  }
  return l;
}

void sf_hip_labeller_destroy(sf_labeller *l) {
  if (!l)
    return;
  if (l->stream)
    (void)hipStreamSynchronize(l->stream);
  for (int i = 0; i < 2; i++) {
    l->carry_frame[i].release();
    l->carry_index[i].release();
    l->carry_value[i].release();
    l->carry_root[i].release();
  }
  l->node_frame.release();
  l->node_index.release();
  l->node_value.release();
  l->node_slot.release();
  l->parent.release();
  l->root.release();
  l->last_slot.release();
  l->keys.release();
  l->row_start.release();
  l->keys_sorted.release();
  l->values.release();
  l->order.release();
  l->open_flag.release();
  l->open_rank.release();
  l->start_flag.release();
  l->start_rank.release();
  l->starts.release();
  l->results.release();
  l->spot_pack.release();
  l->spot_rank.release();
  l->temp.release();
  l->chunk.release();
  l->spots.release();
  l->pixels.release();
  l->pending_dev.release();
  l->pending_count_dev.release();
  if (l->d_n_carry)
    (void)hipFree(l->d_n_carry);
  if (l->d_scalars)
    (void)hipFree(l->d_scalars);
  if (l->d_counts)
    (void)hipFree(l->d_counts);
  if (l->d_totals)
    (void)hipFree(l->d_totals);
  if (l->h_scalars)
    (void)hipHostFree(l->h_scalars);
  if (l->done)
    (void)hipEventDestroy(l->done);
  if (l->stream)
    (void)hipStreamDestroy(l->stream);
  delete l;
}

int sf_hip_labeller_add(sf_labeller *l, int64_t frame,
                        const sf_signal_pixel *signal, size_t n) {
  if (!l || (n && !signal)) {
    g_label_error = "sf_hip_labeller_add: bad argument";
    return SF_BAD_ARGUMENT;
  }
  const int64_t last = !l->pending_frames.empty() ? l->pending_frames.back()
                                                  : l->last_frame;
  const bool any = l->started || !l->pending_frames.empty();
  if (any && frame <= last) {
    g_label_error = "frames must be added in ascending order";
    return SF_ORDER_ERROR;
  }
  const size_t pixels = l->height * l->width;
  for (size_t i = 0; i < n; i++) {
    if (signal[i].index >= pixels) {
      g_label_error = "a signal pixel is off the end of the frame";
      return SF_ORDER_ERROR;
    }
    if (i && signal[i].index <= signal[i - 1].index) {
      g_label_error = "signal pixels are not ascending by index";
      return SF_ORDER_ERROR;
    }
  }
  l->pending.insert(l->pending.end(), signal, signal + n);
  l->pending_frames.push_back(frame);
  l->pending_counts.push_back((uint32_t)n);
  if (l->pending.size() > (4u << 20) || l->pending_frames.size() >= 256)
    return flush_pending(l);
  return SF_OK;
}

int sf_hip_labeller_add_plan(sf_labeller *l, sf_plan *plan,
                             int64_t first_frame) {
  if (!l || !plan) {
    g_label_error = "sf_hip_labeller_add_plan: bad argument";
    return SF_BAD_ARGUMENT;
  }
  int status = flush_pending(l);
  if (status != SF_OK)
    return status;
  if (l->started && first_frame <= l->last_frame) {
    g_label_error = "frames must be added in ascending order";
    return SF_ORDER_ERROR;
  }
  std::vector<const sf_signal_pixel *> lists;
  std::vector<const uint32_t *> count_ptrs;
  std::vector<uint32_t> caps;
  hipStream_t producer = nullptr;
  status = sf::plan_lists(plan, lists, count_ptrs, caps, &producer);
  if (status != SF_OK)
    return status;
  std::vector<uint32_t> counts(lists.size());
  status = sf_hip_plan_counts(plan, counts.data());
  if (status != SF_OK)
    return status;
  std::vector<sf::ChunkDesc> frames(lists.size());
  for (size_t k = 0; k < frames.size(); k++) {
    frames[k].list = lists[k];
    frames[k].count = count_ptrs[k];
    frames[k].capacity = caps[k];
    frames[k].number = (int32_t)(first_frame + (int64_t)k);
    frames[k].linked = k > 0 || (l->started && first_frame == l->last_frame + 1);
  }
  status = run_chunk(l, frames, counts, false, producer);
  if (status != SF_OK)
    return status;
  l->started = true;
  l->last_frame = first_frame + (int64_t)frames.size() - 1;
  return SF_OK;
}

int sf_hip_labeller_flush(sf_labeller *l) {
  if (!l)
    return SF_BAD_ARGUMENT;
  return flush_pending(l);
}

int sf_hip_labeller_finish(sf_labeller *l) {
  if (!l)
    return SF_BAD_ARGUMENT;
  int status = flush_pending(l);
  if (status != SF_OK)
    return status;
  if (!l->started)
    return SF_OK;
  std::vector<sf::ChunkDesc> none;
  std::vector<uint32_t> no_counts;
  status = run_chunk(l, none, no_counts, true, nullptr);
  if (status != SF_OK)
    return status;
  LAB_CHECK(hipStreamSynchronize(l->stream), "hipStreamSynchronize");
  l->n_carry = *l->h_scalars;
  return SF_OK;
}

int sf_hip_labeller_wait(sf_labeller *l) {
  if (!l)
    return SF_BAD_ARGUMENT;
  LAB_CHECK(hipStreamSynchronize(l->stream), "hipStreamSynchronize");
  l->n_carry = *l->h_scalars;
  return SF_OK;
}

size_t sf_hip_labeller_live(sf_labeller *l) {
  if (!l || flush_pending(l) != SF_OK || sf_hip_labeller_wait(l) != SF_OK)
    return 0;
  return l->n_carry;
}

int sf_hip_labeller_sizes(sf_labeller *l, size_t *spots, size_t *pixels) {
  if (!l)
    return SF_BAD_ARGUMENT;
  int status = flush_pending(l);
  if (status == SF_OK)
    status = sf_hip_labeller_wait(l);
  if (status != SF_OK)
    return status;
  unsigned long long totals[2];
  LAB_CHECK(hipMemcpy(totals, l->d_totals, sizeof(totals),
                      hipMemcpyDeviceToHost),
            "reading totals");
  if (spots)
    *spots = totals[0];
  if (pixels)
    *pixels = totals[1];
  return SF_OK;
}

size_t sf_hip_labeller_spot_count(sf_labeller *l) {
  size_t s = 0;
  return sf_hip_labeller_sizes(l, &s, nullptr) == SF_OK ? s : 0;
}

size_t sf_hip_labeller_pixel_count(sf_labeller *l) {
  size_t p = 0;
  return sf_hip_labeller_sizes(l, nullptr, &p) == SF_OK ? p : 0;
}

int sf_hip_labeller_results(sf_labeller *l, sf_spot *spots, sf_pixel *pixels,
                            sf_counts *counts) {
  size_t n_spots = 0, n_pixels = 0;
  const int status = sf_hip_labeller_sizes(l, &n_spots, &n_pixels);
  if (status != SF_OK)
    return status;
  if (spots && n_spots)
    LAB_CHECK(hipMemcpy(spots, l->spots.ptr, n_spots * sizeof(sf_spot),
                        hipMemcpyDeviceToHost),
              "reading spots");
  if (pixels && n_pixels)
    LAB_CHECK(hipMemcpy(pixels, l->pixels.ptr, n_pixels * sizeof(sf_pixel),
                        hipMemcpyDeviceToHost),
              "reading pixels");
  if (counts) {
    unsigned long long c[6];
    LAB_CHECK(hipMemcpy(c, l->d_counts, sizeof(c), hipMemcpyDeviceToHost),
              "reading counts");
    counts->signal_pixels = c[0];
    counts->groups = c[1];
    counts->too_small = c[2];
    counts->too_large = c[3];
    counts->separated = c[4];
    counts->accepted = c[5];
  }
  return SF_OK;
}

const char *sf_hip_labeller_last_error(void) { return g_label_error.c_str(); }

} // extern "C"
