// Six-connected 3-D grouping and the centroid/intensity pass on the device.
//
// Reference semantics: dials_spots::Labeller (src/dials_spots.cc:11-360).
// The reference labels one frame at a time on the host and emits a component
// as soon as a frame arrives that it has no pixel on. Here a chunk of frames
// is labelled at once on the device:
//
//   nodes    the pixels still open from earlier chunks (the frontier), then
//            every signal pixel of the chunk, in (frame, index) order;
//   runs     horizontal runs of nodes (consecutive indices on one row of one
//            frame; each carried pixel is a run of its own), numbered in
//            node order. The grouping works on runs: a chunk of 16-bit frames
//            has about seven pixels to a run;
//   unite    each run joins the runs above it on its frame and the runs on
//            the same row of the previous frame when the frame numbers are
//            consecutive and not two_d (src/dials_spots.cc:102-137); roots
//            always point to the lower run, so a component's root is its
//            first run and that run's first pixel its first pixel;
//   emit     a component is closed when it has no pixel on the chunk's last
//            frame (or at finish, or always in two_d), and closed components
//            are emitted ordered by (last frame, first pixel) -- the order
//            the streaming Labeller flushes them (src/dials_spots.cc:153-194):
//            a radix sort of the runs on (last frame, root run) keeps each
//            component's runs, and so its pixels, in (frame, index) order;
//   finalise the reference's emit() arithmetic (src/dials_spots.cc:221-360)
//            per component, bit for bit: small components one to a lane,
//            larger ones one to a wave with the variance terms added in
//            member order;
//   carry    the open components' pixels, still in (frame, index) order,
//            become the next chunk's frontier, each pointing at its
//            component's first pixel.
#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "sf_device.hh"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

namespace sf {


struct ChunkDesc {
  const sf_signal_pixel *list; // device list of this frame (or its
                               // survivors in staging order, with `place`)
  const uint32_t *place;       // null, or each survivor's place in the list
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
  double inv_width;                      // 1 / width, rounded
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

// Row of pixel `idx`. The product with the rounded reciprocal is within
// 2^-20 of idx / width (idx < 2^32), so for widths below 2^20 its truncation
// is the quotient or one less, and one comparison settles it.
__device__ __forceinline__ uint32_t row_of(const LabelArgs &a, uint32_t idx) {
  if (a.width >= (1u << 20))
    return idx / a.width;
  uint32_t q = (uint32_t)((double)idx * a.inv_width);
  if (idx - q * a.width >= a.width)
    q++;
  return q;
}

// Whole-wave reductions through DPP, with all 64 lanes active: a scan within
// rows of 16 lanes (row_shr 1, 2, 4, 8), then the last lane of row 0 and of
// row 2 broadcast into rows 1 and 3 (row_bcast 15), and the last lane of
// row 1 into rows 2 and 3 (row_bcast 31); lane 63 ends with the result.
// Lanes with no source read the identity.
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_take(uint32_t x, uint32_t identity) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)identity, (int)x, CTRL,
                                               ROWS, 0xF, false);
}
template <class Op>
__device__ __forceinline__ uint32_t wave_reduce(uint32_t x, uint32_t identity,
                                                Op op) {
  x = op(x, dpp_take<0x111, 0xF>(x, identity));
  x = op(x, dpp_take<0x112, 0xF>(x, identity));
  x = op(x, dpp_take<0x114, 0xF>(x, identity));
  x = op(x, dpp_take<0x118, 0xF>(x, identity));
  x = op(x, dpp_take<0x142, 0xA>(x, identity));
  x = op(x, dpp_take<0x143, 0xC>(x, identity));
  return (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
}
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t x) {
  return wave_reduce(x, 0xFFFFFFFFu,
                     [](uint32_t p, uint32_t q) { return p < q ? p : q; });
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t x) {
  return wave_reduce(x, 0u,
                     [](uint32_t p, uint32_t q) { return p > q ? p : q; });
}
__device__ __forceinline__ int32_t wave_min_i32(int32_t x) {
  return (int32_t)wave_reduce((uint32_t)x, 0x7FFFFFFFu, [](uint32_t p,
                                                           uint32_t q) {
    return (int32_t)p < (int32_t)q ? p : q;
  });
}
__device__ __forceinline__ int32_t wave_max_i32(int32_t x) {
  return (int32_t)wave_reduce((uint32_t)x, 0x80000000u, [](uint32_t p,
                                                           uint32_t q) {
    return (int32_t)p > (int32_t)q ? p : q;
  });
}
__device__ __forceinline__ double readlane_f64(double x, int l) {
  const uint64_t b = (uint64_t)__double_as_longlong(x);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
  const uint32_t hi =
      (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_take_f64(double x) {
  const uint64_t b = (uint64_t)__double_as_longlong(x);
  const uint32_t lo = dpp_take<CTRL, ROWS>((uint32_t)b, 0u);
  const uint32_t hi = dpp_take<CTRL, ROWS>((uint32_t)(b >> 32), 0u);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
// The sum of integer-valued doubles: exact in any order while every partial
// sum stays below 2^53.
__device__ __forceinline__ double wave_sum_f64(double x) {
  x += dpp_take_f64<0x111, 0xF>(x);
  x += dpp_take_f64<0x112, 0xF>(x);
  x += dpp_take_f64<0x114, 0xF>(x);
  x += dpp_take_f64<0x118, 0xF>(x);
  x += dpp_take_f64<0x142, 0xA>(x);
  x += dpp_take_f64<0x143, 0xC>(x);
  return readlane_f64(x, 63);
}

// Nodes per tile of gather_chunk and number_runs, threads per tile of
// number_runs (two nodes a thread), and each wave's share of a tile.
constexpr uint32_t kTile = 2048;
constexpr uint32_t kTileThreads = 1024;
constexpr uint32_t kTileWaves = kTileThreads / 64;
constexpr uint32_t kWaveNodes = kTile / kTileWaves;

// A chunk's own nodes (its signal pixels, without the carried frontier) at
// chunk-local places j: frame k's from first[k] on, in (frame, index) order.
// index, value and slot (k + 1) per node; bit (j & 63) of heads[j >> 6]
// marks node j a run head; tile_heads[t] counts the heads of nodes
// [t, t + 1) * kTile.
struct ChunkNodes {
  uint32_t *index, *value, *slot;
  uint64_t *heads;
  uint32_t *tile_heads;
};

// The chunk's nodes out of its frames' lists: node j of frame k is entry
// j - first[k] of frame k's list (first[n_frames]: the total). A node
// begins a run when it is its frame's first, the first column of a row, or
// not the index after its list predecessor's. Blocks take tiles of kTile
// nodes in turn (the grid may be an upper bound: the total is on the
// device).
__global__ __launch_bounds__(256) void gather_chunk(LabelArgs a,
                                                    const ChunkDesc *frames,
                                                    uint32_t n_frames,
                                                    const uint32_t *first,
                                                    ChunkNodes out) {
  __shared__ int tile_frame; // the frame of the tile's first node
  __shared__ uint32_t listed;
  if (n_frames == 0)
    return;
  const uint32_t total = first[n_frames];
  const uint32_t lane = lane_id();
  for (uint32_t t = blockIdx.x; t < (total + kTile - 1) / kTile;
       t += gridDim.x) {
    const uint32_t tile = t * kTile;
    if (threadIdx.x == 0) {
      listed = 0;
      int lo = 0, hi = (int)n_frames - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= tile)
          lo = mid;
        else
          hi = mid - 1;
      }
      tile_frame = lo;
    }
    __syncthreads();
    int k = tile_frame; // stepped forward to each node's frame
    uint32_t mine = 0;
    const uint32_t end = min(tile + kTile, total);
    for (uint32_t j = tile + threadIdx.x; j < end; j += blockDim.x) {
      while (k + 1 < (int)n_frames && first[k + 1] <= j)
        k++;
      const sf_signal_pixel *list = frames[k].list;
      const uint32_t at = j - first[k];
      const sf_signal_pixel e = list[at];
      out.index[j] = e.index;
      out.value[j] = e.value;
      out.slot[j] = (uint32_t)k + 1;
      const bool head = at == 0 || e.index - row_of(a, e.index) * a.width == 0 ||
                        list[at - 1].index != e.index - 1;
      const uint64_t mask = ballot(head);
      if (lane == 0)
        out.heads[j >> 6] = mask; // the wave's 64 nodes start at j
      mine += head;
    }
    if (mine)
      atomicAdd(&listed, mine);
    __syncthreads();
    if (threadIdx.x == 0)
      out.tile_heads[t] = listed;
    __syncthreads();
  }
}

// The chunk's nodes from survivors in staging order with their places in
// the list (frame k's survivor e is node first[k] + place[e]); a list over
// its capacity counts 0. The run heads then come from heads_chunk. Grid
// (blocks, frames in turn).
__global__ __launch_bounds__(256) void scatter_chunk(const ChunkDesc *frames,
                                                     uint32_t n_frames,
                                                     const uint32_t *first,
                                                     ChunkNodes out) {
  for (uint32_t k = blockIdx.y; k < n_frames; k += gridDim.y) {
    const ChunkDesc f = frames[k];
    const uint32_t c = *f.count;
    if (c > f.capacity)
      continue;
    const uint32_t base = first[k];
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < c;
         e += gridDim.x * blockDim.x) {
      const uint32_t j = base + f.place[e];
      const sf_signal_pixel sp = f.list[e];
      out.index[j] = sp.index;
      out.value[j] = sp.value;
      out.slot[j] = k + 1;
    }
  }
}

// The run heads of a chunk's nodes already in place (scatter_chunk), as
// gather_chunk marks them: its frame's first, the first column of a row, or
// not the index after the node before it. Blocks take tiles in turn.
__global__ __launch_bounds__(256) void heads_chunk(LabelArgs a,
                                                   uint32_t n_frames,
                                                   const uint32_t *first,
                                                   ChunkNodes out) {
  __shared__ int tile_frame;
  __shared__ uint32_t listed;
  if (n_frames == 0)
    return;
  const uint32_t total = first[n_frames];
  const uint32_t lane = lane_id();
  for (uint32_t t = blockIdx.x; t < (total + kTile - 1) / kTile;
       t += gridDim.x) {
    const uint32_t tile = t * kTile;
    if (threadIdx.x == 0) {
      listed = 0;
      int lo = 0, hi = (int)n_frames - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= tile)
          lo = mid;
        else
          hi = mid - 1;
      }
      tile_frame = lo;
    }
    __syncthreads();
    int k = tile_frame;
    uint32_t mine = 0;
    const uint32_t end = min(tile + kTile, total);
    for (uint32_t j = tile + threadIdx.x; j < end; j += blockDim.x) {
      while (k + 1 < (int)n_frames && first[k + 1] <= j)
        k++;
      const uint32_t idx = out.index[j];
      const bool head = j == first[k] ||
                        idx - row_of(a, idx) * a.width == 0 ||
                        out.index[j - 1] != idx - 1;
      const uint64_t mask = ballot(head);
      if (lane == 0)
        out.heads[j >> 6] = mask;
      mine += head;
    }
    if (mine)
      atomicAdd(&listed, mine);
    __syncthreads();
    if (threadIdx.x == 0)
      out.tile_heads[t] = listed;
    __syncthreads();
  }
}

// Runs numbered in node order, over the carried nodes g < n_carry (blocks
// chunk_tiles on, kTile nodes a block: each a run of its own, numbered as
// its node, whose parent is its carried root) and the chunk's own nodes
// (block t: chunk tile t; node j is node g = n_carry + j, and run r is the
// r-th head). A chunk tile's first run is n_carry plus the heads of the
// chunk tiles before it, which each block adds up itself; each wave takes
// kWaveNodes consecutive nodes of its tile. Writes run_of per node; per run
// its head node, first and last pixel index, slot and initial parent (a
// chunk run's is itself), and a zero last_slot. The last chunk tile's block
// (the first block when there is none) writes the run count, here and to
// the host's mapped copy; the first block adds the chunk's own nodes to the
// signal-pixel count.
__global__ __launch_bounds__(kTileThreads) void number_runs(
    LabelArgs a, const uint32_t *cn_index, const uint32_t *cn_slot,
    const uint64_t *cn_heads, const uint32_t *cn_tile_heads,
    uint32_t chunk_tiles, const uint32_t *carry_index,
    const uint32_t *carry_root, uint32_t *run_of, uint32_t *run_head,
    uint32_t *run_idx, uint32_t *run_last, uint32_t *run_slot,
    uint32_t *parent, uint32_t *last_slot, uint32_t *n_runs,
    uint32_t *host_runs, unsigned long long *counts) {
  const uint32_t n_chunk = a.n_nodes - a.n_carry;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    counts[0] += n_chunk;
    if (chunk_tiles == 0) {
      *n_runs = a.n_carry;
      *host_runs = a.n_carry; // host-mapped
      __threadfence_system();
    }
  }
  if (blockIdx.x >= chunk_tiles) {
    const uint32_t base = (blockIdx.x - chunk_tiles) * kTile;
    const uint32_t end = min(base + kTile, a.n_carry);
    for (uint32_t g = base + threadIdx.x; g < end; g += blockDim.x) {
      const uint32_t idx = carry_index[g];
      run_of[g] = g;
      run_head[g] = g;
      run_idx[g] = idx;
      run_last[g] = idx;
      run_slot[g] = 0;
      parent[g] = carry_root[g];
      last_slot[g] = 0;
    }
    return;
  }
  // part[w]: wave w's share of the earlier tiles' heads; part[kTileWaves +
  // w]: the heads among wave w's own nodes
  __shared__ uint32_t part[2 * kTileWaves];
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  const uint32_t tile = blockIdx.x * kTile;
  uint32_t before = 0;
  for (uint32_t i = threadIdx.x; i < blockIdx.x; i += blockDim.x)
    before += cn_tile_heads[i];
  before = wave_reduce(before, 0u,
                       [](uint32_t p, uint32_t q) { return p + q; });
  const uint32_t first = tile + kWaveNodes * wave;
  uint32_t own = 0;
  if (lane < kWaveNodes / 64 && first + 64 * lane < n_chunk)
    own = (uint32_t)__builtin_popcountll(cn_heads[(first >> 6) + lane]);
  own = wave_reduce(own, 0u, [](uint32_t p, uint32_t q) { return p + q; });
  if (lane == 0) {
    part[wave] = before;
    part[kTileWaves + wave] = own;
  }
  __syncthreads();
  uint32_t tiles_before = a.n_carry, tile_runs = 0;
  for (uint32_t w = 0; w < kTileWaves; w++) {
    tiles_before += part[w];
    tile_runs += part[kTileWaves + w];
  }
  uint32_t run = tiles_before; // runs before this wave's nodes
  for (uint32_t w = 0; w < wave; w++)
    run += part[kTileWaves + w];
  if (blockIdx.x + 1 == chunk_tiles && threadIdx.x == 0) {
    *n_runs = tiles_before + tile_runs;
    *host_runs = *n_runs; // host-mapped
    __threadfence_system();
  }
  const uint64_t upto = ~0ull >> (63 - lane);
  for (uint32_t q = 0; q < kWaveNodes / 64; q++) {
    const uint32_t j = first + 64 * q + lane;
    if (first + 64 * q >= n_chunk)
      break;
    const uint64_t mask = cn_heads[(first >> 6) + q];
    // the next group's heads, for lane 63's tail test
    const uint64_t next = first + 64 * (q + 1) < n_chunk
                              ? cn_heads[(first >> 6) + q + 1]
                              : 1ull;
    if (j < n_chunk) {
      const uint32_t g = a.n_carry + j;
      const uint32_t r = run + (uint32_t)__builtin_popcountll(mask & upto) - 1;
      run_of[g] = r;
      const bool head = (mask >> lane) & 1;
      const bool tail = j + 1 == n_chunk ||
                        (lane < 63 ? (mask >> (lane + 1)) & 1 : next & 1);
      if (head || tail) {
        const uint32_t idx = cn_index[j];
        if (head) {
          const uint32_t slot = cn_slot[j];
          run_head[r] = g;
          run_idx[r] = idx;
          run_slot[r] = slot;
          parent[r] = r;
          // the last frame (slot) this run alone reaches, raised to its
          // component's by resolve_runs
          last_slot[r] = slot;
        }
        if (tail)
          run_last[r] = idx;
      }
    }
    run += (uint32_t)__builtin_popcountll(mask);
  }
}

// A node's frame, index and value: carried nodes g < n_carry from the
// frontier, the chunk's own from its chunk-local arrays (the frame by its
// slot in the chunk table).
struct NodeSrc {
  const int32_t *carry_frame;
  const uint32_t *carry_index, *carry_value;
  const uint32_t *index, *value, *slot;
  const ChunkDesc *frames;
};

__device__ __forceinline__ void node_at(const LabelArgs &a, const NodeSrc &s,
                                        uint32_t g, int32_t &frame,
                                        uint32_t &index, uint32_t &value) {
  if (g < a.n_carry) {
    frame = s.carry_frame[g];
    index = s.carry_index[g];
    value = s.carry_value[g];
  } else {
    const uint32_t j = g - a.n_carry;
    index = s.index[j];
    value = s.value[j];
    frame = s.frames[s.slot[j] - 1].number;
  }
}

// run_rows[slot * (height + 1) + y] = the first run of `slot` whose first
// index is at least y * width (slot 0: the carried runs -- single carried
// pixels -- of the last carried frame; slot k + 1: chunk frame k, whose runs
// start at its first node's run). One thread per (slot, row).
__global__ void fill_run_rows(LabelArgs a, const ChunkDesc *frames,
                              const uint32_t *run_of, const uint32_t *run_idx,
                              const uint32_t *n_runs, uint32_t *run_rows) {
  const uint32_t height = a.pixels / a.width;
  const uint32_t total = (a.n_frames + 1) * (height + 1);
  const uint32_t runs = *n_runs;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += gridDim.x * blockDim.x) {
    const uint32_t slot = t / (height + 1), y = t % (height + 1);
    uint32_t from, to;
    if (slot == 0) {
      from = a.carry_tail;
      to = a.n_carry;
    } else {
      const uint32_t f = frames[slot - 1].offset;
      const uint32_t e =
          slot < a.n_frames ? frames[slot].offset : a.n_nodes;
      from = f < a.n_nodes ? run_of[f] : runs;
      to = e < a.n_nodes ? run_of[e] : runs;
    }
    const uint32_t want = y * a.width; // y == height: past every index
    uint32_t lo = from, hi = to;
    if (y == height) {
      lo = to;
    } else {
      while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (run_idx[mid] < want)
          lo = mid + 1;
        else
          hi = mid;
      }
    }
    run_rows[t] = lo;
  }
}

// First entry in [lo, hi) of the sorted `key` that is at least `want` (hi
// if none). Four independent probes a step cut the range five ways.
__device__ __forceinline__ uint32_t lower_bound(const uint32_t *key,
                                                uint32_t lo, uint32_t hi,
                                                uint32_t want) {
  while (hi - lo > 4) {
    const uint32_t step = (hi - lo) / 5;
    const uint32_t p0 = lo + step, p1 = p0 + step, p2 = p1 + step,
                   p3 = p2 + step;
    const uint32_t v0 = key[p0], v1 = key[p1], v2 = key[p2], v3 = key[p3];
    if (v3 < want) {
      lo = p3 + 1;
    } else if (v2 < want) {
      lo = p2 + 1;
      hi = p3;
    } else if (v1 < want) {
      lo = p1 + 1;
      hi = p2;
    } else if (v0 < want) {
      lo = p0 + 1;
      hi = p1;
    } else {
      hi = p0;
    }
  }
  uint32_t below = 0;
#pragma unroll
  for (uint32_t q = 0; q < 4; q++)
    below += lo + q < hi && key[lo + q] < want;
  return lo + below;
}

// Joins run `r` to every run of [from, to) -- the runs of one row, in index
// order, disjoint -- that covers one of indices lo .. hi. Long rows start
// with a search on the runs' last indices; runs are then read four at a
// time.
__device__ __forceinline__ void unite_row(uint32_t *parent,
                                          const uint32_t *run_idx,
                                          const uint32_t *run_last,
                                          uint32_t r, uint32_t from,
                                          uint32_t to, uint32_t lo,
                                          uint32_t hi) {
  if (to - from > 8)
    from = lower_bound(run_last, from, to, lo);
  for (uint32_t u = from; u < to; u += 4) {
    uint32_t first[4], last[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const bool in = u + q < to;
      first[q] = in ? run_idx[u + q] : 0xFFFFFFFFu;
      last[q] = in ? run_last[u + q] : 0u;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      if (u + q >= to || first[q] > hi)
        return;
      if (last[q] >= lo)
        unite(parent, r, u + q);
    }
  }
}

// The up and previous-frame edges are src/dials_spots.cc:106-137's: a run
// joins each run above it on its frame, and each run on the same row of the
// frame before when the frame numbers are consecutive and not two_d.
// ---------------------------------------------------------------------------
// The same unions in two passes, with few global atomics. Cut the chunk into
// tiles of rows x frames: an up edge stays on its frame and a
// previous-frame edge on its row, so most edges join two runs of one tile.
// A workgroup takes a tile, loads its runs' columns and rows into LDS and
// unites them there; then one pass over all runs does, on the global
// forest, the edges the tiles could not: up edges from a tile's first row,
// previous-frame edges from a tile's first frame (to the carried frontier
// too), and every edge of a tile whose runs overflow LDS. Roots still always
// point to the lower run: local ids follow run order (slot, then row, then
// column), and the global pass hooks the larger root under the smaller.
// ---------------------------------------------------------------------------
constexpr uint32_t kTileRuns = 5632;   // runs a tile unites in LDS
constexpr uint32_t kTileFrames = 32;   // chunk frames a tile spans at most
constexpr uint32_t kTileMaxRows = 120; // rows a tile spans at most
constexpr uint32_t kTileMinRows = 8;

// The chunk's tiling, the same in both passes: frame groups of
// kTileFrames (or fewer), and rows per tile such that a tile is expected to
// hold three quarters of kTileRuns.
struct Tiling {
  uint32_t frames, rows, bands, groups;
};
__device__ __forceinline__ Tiling tiling_for(const LabelArgs &a,
                                             uint32_t runs) {
  Tiling t;
  const uint32_t height = a.pixels / a.width;
  const uint32_t nf = a.n_frames ? a.n_frames : 1u;
  t.frames = min(nf, kTileFrames);
  const uint64_t per_row = (uint64_t)(runs ? runs : 1u) * t.frames / nf;
  const uint64_t want = (uint64_t)kTileRuns * 3 / 4 * height /
                        (per_row ? per_row : 1u);
  t.rows = (uint32_t)min<uint64_t>(
      min<uint64_t>(max<uint64_t>(want, kTileMinRows), kTileMaxRows), height);
  t.bands = (height + t.rows - 1) / t.rows;
  t.groups = (nf + t.frames - 1) / t.frames;
  return t;
}

__device__ __forceinline__ uint32_t find_local(uint32_t *lp, uint32_t i) {
  uint32_t p = lp[i];
  while (p != i) {
    const uint32_t gp = lp[p];
    if (gp != p)
      lp[i] = gp;
    i = p;
    p = gp;
  }
  return i;
}

__device__ void unite_local(uint32_t *lp, uint32_t a, uint32_t b) {
  for (;;) {
    a = find_local(lp, a);
    b = find_local(lp, b);
    if (a == b)
      return;
    if (a > b) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    const uint32_t old = atomicCAS(&lp[b], b, a);
    if (old == b)
      return;
    b = old;
  }
}

// Inclusive sum over the wave's lanes.
__device__ __forceinline__ uint32_t wave_inclusive(uint32_t x) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)x, d, 64);
    if (lane >= (uint32_t)d)
      x += y;
  }
  return x;
}

// One tile per block: block b is band b % bands (rows [band * rows, ..))
// of frame group b / bands (slots 1 + group * frames .. ). Its runs united in
// LDS, then each run's parent set to its local root. A tile that cannot be
// held (more runs than kTileRuns, a frame wider than 65,536) is marked in
// spill and left to the global pass. The grid is an upper bound on the
// tiles; blocks past them return.
__global__ __launch_bounds__(1024) void unite_tiles(
    LabelArgs a, const ChunkDesc *frames, const uint32_t *run_idx,
    const uint32_t *run_last, const uint32_t *run_rows,
    const uint32_t *n_runs, uint32_t *parent, uint32_t *last_slot,
    uint32_t *spill) {
  __shared__ uint32_t lp[kTileRuns];
  __shared__ uint32_t cols[kTileRuns];  // first column | last column << 16
  __shared__ uint8_t row_of_run[kTileRuns];
  __shared__ uint16_t lrows[kTileFrames * (kTileMaxRows + 1)];
  __shared__ uint32_t slot_first[kTileFrames + 1], slot_off[kTileFrames + 1];
  __shared__ uint32_t linked[kTileFrames];
  __shared__ uint32_t total;
  const uint32_t height = a.pixels / a.width, stride = height + 1;
  const Tiling t = tiling_for(a, *n_runs);
  if (blockIdx.x >= t.bands * t.groups)
    return;
  const uint32_t band = blockIdx.x % t.bands, group = blockIdx.x / t.bands;
  const uint32_t y0 = band * t.rows, y1 = min(y0 + t.rows, height);
  const uint32_t s_lo = 1 + group * t.frames;
  const uint32_t nslots = min(t.frames, a.n_frames + 1 - s_lo);
  if (threadIdx.x < 64) { // each slot's runs in the tile, as local ranges
    const uint32_t lane = lane_id();
    uint32_t first = 0, count = 0;
    if (lane < nslots) {
      const uint32_t s = s_lo + lane;
      first = run_rows[(size_t)s * stride + y0];
      count = run_rows[(size_t)s * stride + y1] - first;
      linked[lane] = frames[s - 1].linked;
    }
    const uint32_t incl = wave_inclusive(count);
    if (lane < nslots) {
      slot_first[lane] = first;
      slot_off[lane] = incl - count;
    }
    if (lane == 63)
      total = incl;
  }
  __syncthreads();
  const uint32_t n = total;
  if (n > kTileRuns || a.width > 65536u) {
    if (threadIdx.x == 0)
      spill[blockIdx.x] = 1;
    return;
  }
  if (threadIdx.x == 0)
    spill[blockIdx.x] = 0;
  // the local slot of local run i: the last slot whose offset is at most i
  const auto slot_of = [&](uint32_t i) {
    uint32_t lo = 0, hi = nslots - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if (slot_off[mid] <= i)
        lo = mid;
      else
        hi = mid - 1;
    }
    return lo;
  };
  const uint32_t trows = y1 - y0;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    const uint32_t ls = slot_of(i);
    const uint32_t r = slot_first[ls] + (i - slot_off[ls]);
    const uint32_t idx = run_idx[r], last = run_last[r];
    const uint32_t y = row_of(a, idx);
    const uint32_t base = y * a.width;
    lp[i] = i;
    cols[i] = (idx - base) | ((last - base) << 16);
    row_of_run[i] = (uint8_t)(y - y0);
  }
  for (uint32_t e = threadIdx.x; e < nslots * (trows + 1); e += blockDim.x) {
    const uint32_t ls = e / (trows + 1), ry = e % (trows + 1);
    lrows[ls * (kTileMaxRows + 1) + ry] =
        (uint16_t)(slot_off[ls] +
                   (run_rows[(size_t)(s_lo + ls) * stride + y0 + ry] -
                    slot_first[ls]));
  }
  __syncthreads();
  // The neighbours of local run i (columns lo .. hi) among the runs
  // [from, to) of one row -- disjoint, in column order -- are those that
  // share a column with it: a range [first, end) of that row.
  const auto span = [&](uint32_t from, uint32_t to, uint32_t lo, uint32_t hi,
                        uint32_t &end) {
    if (to - from > 8) { // skip the runs that end left of lo
      uint32_t l = from, h = to;
      while (l < h) {
        const uint32_t mid = (l + h) >> 1;
        if ((cols[mid] >> 16) < lo)
          l = mid + 1;
        else
          h = mid;
      }
      from = l;
    }
    while (from < to && (cols[from] >> 16) < lo)
      from++;
    end = from;
    while (end < to && (cols[end] & 0xFFFFu) <= hi)
      end++;
    return from;
  };
  // Two passes, so that runs of one large component do not all race to
  // hook one root: first each run points at its lowest neighbour (every
  // neighbour comes earlier in local order), a plain store; then each
  // further edge is united, and most of them already share a root.
  const auto neighbours = [&](uint32_t i, uint32_t &up0, uint32_t &up1,
                              uint32_t &pv0, uint32_t &pv1) {
    const uint32_t ls = slot_of(i), ry = row_of_run[i];
    const uint32_t c = cols[i], lo = c & 0xFFFFu, hi = c >> 16;
    const uint16_t *rows = lrows + ls * (kTileMaxRows + 1);
    up0 = up1 = pv0 = pv1 = 0;
    if (ry > 0)
      up0 = span(rows[ry - 1], rows[ry], lo, hi, up1);
    if (!a.two_d && ls > 0 && linked[ls]) {
      const uint16_t *prev = lrows + (ls - 1) * (kTileMaxRows + 1);
      pv0 = span(prev[ry], prev[ry + 1], lo, hi, pv1);
    }
  };
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    uint32_t up0, up1, pv0, pv1;
    neighbours(i, up0, up1, pv0, pv1);
    uint32_t low = i;
    if (up1 > up0)
      low = up0;
    if (pv1 > pv0 && pv0 < low)
      low = pv0;
    lp[i] = low;
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    uint32_t up0, up1, pv0, pv1;
    neighbours(i, up0, up1, pv0, pv1);
    const uint32_t low = (up1 > up0 && !(pv1 > pv0 && pv0 < up0)) ? up0
                         : (pv1 > pv0)                            ? pv0
                                                                  : i;
    for (uint32_t u = up0; u < up1; u++)
      if (u != low)
        unite_local(lp, i, u);
    for (uint32_t u = pv0; u < pv1; u++)
      if (u != low)
        unite_local(lp, i, u);
  }
  __syncthreads();
  // local roots (every path compressed), the last slot each local component
  // reaches (in `cols`, free now), then each run's parent and, for local
  // roots, that last slot for resolve_runs to raise the component's by
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    uint32_t root = i;
    while (lp[root] != root)
      root = lp[root];
    lp[i] = root;
    cols[i] = 0;
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
    atomicMax(&cols[lp[i]], s_lo + slot_of(i));
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    const uint32_t root = lp[i];
    const uint32_t s = slot_of(i), q = slot_of(root);
    const uint32_t g = slot_first[s] + (i - slot_off[s]);
    parent[g] = slot_first[q] + (root - slot_off[q]);
    last_slot[g] = root == i ? cols[i] : 0u;
  }
}

// The edges the tiles left, on the global forest (a run at a time): up edges from a tile's first row, previous-frame edges from a
// tile's first frame (which includes those to the carried frontier), and
// every edge of a spilled tile.
__global__ __launch_bounds__(256) void unite_cross(
    LabelArgs a, const ChunkDesc *frames, const uint32_t *run_idx,
    const uint32_t *run_last, const uint32_t *run_slot,
    const uint32_t *run_rows, const uint32_t *n_runs, const uint32_t *spill,
    uint32_t *parent) {
  const uint32_t stride = a.pixels / a.width + 1;
  const uint32_t runs = *n_runs;
  const Tiling t = tiling_for(a, runs);
  for (uint32_t r = xcd_local_item(blockIdx.x, gridDim.x) * blockDim.x +
                    threadIdx.x + a.n_carry;
       r < runs; r += gridDim.x * blockDim.x) {
    const int k = (int)run_slot[r] - 1; // chunk frame
    const uint32_t idx = run_idx[r], last = run_last[r];
    const uint32_t y = row_of(a, idx);
    const uint32_t group = (uint32_t)k / t.frames;
    const bool spilled = spill[group * t.bands + y / t.rows] != 0;
    const uint32_t *mine = run_rows + (size_t)(k + 1) * stride;
    if (y > 0 && (spilled || y % t.rows == 0))
      unite_row(parent, run_idx, run_last, r, mine[y - 1], mine[y],
                idx - a.width, last - a.width);
    if (!a.two_d && frames[k].linked &&
        (spilled || (uint32_t)k % t.frames == 0)) {
      const uint32_t *prev = run_rows + (size_t)k * stride;
      unite_row(parent, run_idx, run_last, r, prev[y], prev[y + 1], idx,
                last);
    }
  }
}

// Every union is done: resolve each run's root (its component's first run,
// whose first pixel is the component's first pixel) with a read-only walk.
// (Writing roots back into `parent` while other threads still walk it would
// let a path-halving write replace a finished root with an intermediate
// ancestor.) Each run then raises its root's last slot by what it carries
// (one atomic per tile-local component, not per run).
__global__ void resolve_runs(const uint32_t *parent, const uint32_t *run_slot,
                             const uint32_t *n_runs, uint32_t *root,
                             uint32_t *last_slot) {
  const uint32_t runs = *n_runs;
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < runs;
       r += gridDim.x * blockDim.x) {
    uint32_t p = r;
    while (parent[p] != p)
      p = parent[p];
    root[r] = p;
    // last_slot holds what run r contributes: its own slot, or (for a root
    // of a component united in a tile) that component's last slot, or 0;
    // only roots are raised, so a non-root's entry is read undisturbed
    const uint32_t last = last_slot[r];
    if (p != r && last)
      atomicMax(&last_slot[p], last);
  }
}

// Sort keys per run: closed components by (last slot, root run), the slot
// shifted left by `shift`; open ones last (all bits set). Also each run's
// pixels if it stays open (else 0), for the frontier.
template <typename K>
__global__ void make_keys(LabelArgs a, const uint32_t *root,
                          const uint32_t *last_slot, const uint32_t *run_head,
                          const uint32_t *n_runs, K *keys, uint32_t *values,
                          uint32_t *open_len, int shift) {
  const uint32_t runs = *n_runs;
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < runs;
       r += gridDim.x * blockDim.x) {
    const uint32_t p = root[r];
    const uint32_t last = last_slot[p];
    const bool open = !a.closing && !a.two_d && a.n_frames > 0 &&
                      last == a.n_frames;
    keys[r] = open ? ~(K)0 : (K)(((K)last << shift) | p);
    values[r] = r;
    open_len[r] =
        open ? (r + 1 < runs ? run_head[r + 1] : a.n_nodes) - run_head[r] : 0;
  }
}

// Per sorted run i: start << 32 | pixels, where start = 1 if the run begins
// a closed component and pixels = its pixel count if closed (else 0). One
// exclusive scan of these gives each closed run's component (high half) and
// its first place among the closed members (low half). Also each run's
// sorted place (pos_of) and the number of closed runs (d_closed_runs).
template <typename K>
__global__ void mark_runs(LabelArgs a, const K *keys, K open_key,
                          const uint32_t *order, const uint32_t *run_head,
                          uint32_t runs, unsigned long long *pack,
                          uint32_t *pos_of, uint32_t *d_closed_runs) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < runs;
       i += gridDim.x * blockDim.x) {
    const K key = keys[i];
    const bool closed = key != open_key;
    const bool start = closed && (i == 0 || key != keys[i - 1]);
    const uint32_t r = order[i];
    pos_of[r] = i;
    uint32_t len = 0;
    if (closed)
      len = (r + 1 < runs ? run_head[r + 1] : a.n_nodes) - run_head[r];
    pack[i] = ((unsigned long long)start << 32) | len;
    if (closed && (i + 1 == runs || keys[i + 1] == open_key))
      *d_closed_runs = i + 1;
  }
}

// Component c's first run (starts) and first member (comp_mem); the number
// of components (d_scalars[0]) and of closed members (d_scalars[1]).
__global__ void list_starts(const unsigned long long *pack,
                            const unsigned long long *rank, uint32_t runs,
                            const uint32_t *d_closed_runs, uint32_t *starts,
                            uint32_t *comp_mem, uint32_t *d_scalars) {
  const uint32_t closed = *d_closed_runs;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < closed;
       i += gridDim.x * blockDim.x) {
    const unsigned long long p = pack[i], q = rank[i];
    if (p >> 32) {
      starts[q >> 32] = i;
      comp_mem[q >> 32] = (uint32_t)q;
    }
    if (i + 1 == closed) {
      d_scalars[0] = (uint32_t)(q >> 32) + (uint32_t)(p >> 32);
      d_scalars[1] = (uint32_t)q + (uint32_t)p;
    }
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

// Calls f(k, index, value, frame) for members k = 0 .. n - 1 in order; the
// members are read eight at a time, ahead of the arithmetic that uses them.
template <class F>
__device__ __forceinline__ void for_members(uint32_t s0, uint32_t n,
                                            const int32_t *mem_frame,
                                            const uint32_t *mem_index,
                                            const uint32_t *mem_value, F f) {
  for (uint32_t k0 = 0; k0 < n; k0 += 8) {
    uint32_t idx[8], val[8];
    int32_t z[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const bool in = k0 + q < n;
      idx[q] = in ? mem_index[s0 + k0 + q] : 0u;
      val[q] = in ? mem_value[s0 + k0 + q] : 0u;
      z[q] = in ? mem_frame[s0 + k0 + q] : 0;
    }
#pragma unroll
    for (int q = 0; q < 8; q++)
      if (k0 + q < n)
        f(k0 + q, idx[q], val[q], z[q]);
  }
}

// The closed members' frame, index and value in emission order, and each
// member's component, gathered once, one thread per node: node g of run r,
// at sorted place i, is member (i's first member) + (g - r's head) of
// component (i's scan rank, high half) + i's start bit - 1.
__global__ void gather_members(LabelArgs a, const uint32_t *d_closed_runs,
                               const uint32_t *run_of, const uint32_t *pos_of,
                               const uint32_t *run_head,
                               const unsigned long long *pack,
                               const unsigned long long *rank,
                               NodeSrc nodes, int32_t *mem_frame,
                               uint32_t *mem_index, uint32_t *mem_value,
                               uint32_t *mem_comp) {
  const uint32_t closed = *d_closed_runs;
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < a.n_nodes;
       g += gridDim.x * blockDim.x) {
    const uint32_t r = run_of[g];
    const uint32_t i = pos_of[r];
    if (i >= closed)
      continue;
    const unsigned long long q = rank[i];
    const uint32_t at = (uint32_t)q + (g - run_head[r]);
    node_at(a, nodes, g, mem_frame[at], mem_index[at], mem_value[at]);
    mem_comp[at] = (uint32_t)(q >> 32) + (uint32_t)(pack[i] >> 32) - 1;
  }
}

// A block's tallies of emitted kinds (kAccepted, kTooSmall, kTooLarge,
// kSeparated), summed in LDS and stored in the block's own row of `rows`
// (four counters a row); close_chunk adds the rows up. (One atomic per
// counter and block on the shared counters serialised over thousands of
// blocks.)
__device__ void store_tallies(const uint32_t (&tally)[4], uint32_t *rows) {
  __shared__ uint32_t block_tally[4];
  if (threadIdx.x < 4)
    block_tally[threadIdx.x] = 0;
  __syncthreads();
  for (int t = 0; t < 4; t++)
    if (tally[t])
      atomicAdd(&block_tally[t], tally[t]);
  __syncthreads();
  if (threadIdx.x < 4)
    rows[4 * blockIdx.x + threadIdx.x] = block_tally[threadIdx.x];
}

// What the centroid pass knows of a component before its last arithmetic:
// the sums over its members in member order (src/dials_spots.cc:233-296),
// its variance sums (src/dials_spots.cc:304-316), its bounding box, peak
// and size, and whether it was rejected for its size.
struct CompSums {
  double sum, sum_sq;
  double weighted[3];
  double peak[3];
  double delta_sq[3];
  int32_t bbox[6];
  uint32_t n;
  int32_t kind; // kTooSmall or kTooLarge, or -1 to be finished
};

// The sums of the reference's emit() (src/dials_spots.cc:233-296) for one
// lane, operation for operation: sums, bounding box and peak in member
// order.
__device__ void sequential_sums(const LabelArgs &a, uint32_t s0, uint32_t n,
                                const int32_t *mem_frame,
                                const uint32_t *mem_index,
                                const uint32_t *mem_value, CompSums &cs) {
  double sum = 0.0, sum_sq = 0.0;
  double weighted[3] = {0.0, 0.0, 0.0};
  int32_t x0 = 0, x1 = 0, y0 = 0, y1 = 0, z0 = 0, z1 = 0;
  uint32_t peak_value = 0;
  double peak[3] = {0.0, 0.0, 0.0};
  for_members(s0, n, mem_frame, mem_index, mem_value,
              [&](uint32_t k, uint32_t idx, uint32_t v, int32_t z) {
                const uint32_t row = row_of(a, idx);
                const int32_t x = (int32_t)(idx - row * a.width);
                const int32_t y = (int32_t)row;
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
                const double value = (double)v;
                const double position[3] = {(double)x + 0.5, (double)y + 0.5,
                                            (double)z + 0.5};
                sum += value;
                sum_sq += value * value;
                for (int axis = 0; axis < 3; axis++)
                  weighted[axis] += value * position[axis];
                if (v > peak_value) {
                  peak_value = v;
                  peak[0] = position[0];
                  peak[1] = position[1];
                  peak[2] = position[2];
                }
              });
  cs.sum = sum;
  cs.sum_sq = sum_sq;
  for (int axis = 0; axis < 3; axis++) {
    cs.weighted[axis] = weighted[axis];
    cs.peak[axis] = peak[axis];
  }
  cs.bbox[0] = x0;
  cs.bbox[1] = x1;
  cs.bbox[2] = y0;
  cs.bbox[3] = y1;
  cs.bbox[4] = z0;
  cs.bbox[5] = z1;
}

// Variance terms staged per round, and the stride of each axis's row (two
// doubles over, so the three chains read distinct banks).
constexpr uint32_t kTermRound = 256;
constexpr uint32_t kTermStride = kTermRound + 2;
static_assert(kTermRound % 16 == 0, "the variance sums take sixteen terms at a time");

// One wave per closed component, in emission order: everything of the
// reference's emit() (src/dials_spots.cc:221-330) that reads the members,
// bit for bit:
//  * bbox, peak (first strict maximum in member order) and the sums are
//    taken across the wave. The reference sums in double in member order;
//    when every term and partial sum is an integer (or half-integer) below
//    2^53 -- always so for 16-bit frames -- those double sums are exact and
//    equal the sums of the same integers taken in any order, which is what
//    the lanes and the DPP reduction add here. Otherwise one lane takes the
//    sums in member order;
//  * the variance terms v (p - c)^2 are formed in parallel and staged in
//    LDS, then lanes 0, 1 and 2 each add one axis's terms in member order,
//    as src/dials_spots.cc:304-316 adds them, the next terms read ahead of
//    the adds.
// The arithmetic that follows the sums is finish_components'.
__global__ __launch_bounds__(256) void component_sums(
    LabelArgs a, const uint32_t *starts, const uint32_t *n_starts,
    const uint32_t *n_closed, const int32_t *mem_frame,
    const uint32_t *mem_index, const uint32_t *mem_value, CompSums *sums) {
  __shared__ double terms[4][3 * kTermStride];
  const uint32_t comps = *n_starts;
  const uint32_t end_all = *n_closed;
  const uint32_t lane = lane_id();
  double *mine = terms[threadIdx.x >> 6];
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  uint32_t c = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  // each component's bounds are read one component ahead
  uint32_t next_s0 = c < comps ? starts[c] : 0u;
  uint32_t next_end = c + 1 < comps ? starts[c + 1] : end_all;
  for (; c < comps; c += waves) {
    const uint32_t s0 = next_s0;
    const uint32_t n = next_end - s0;
    if (c + waves < comps) {
      next_s0 = starts[c + waves];
      next_end = c + waves + 1 < comps ? starts[c + waves + 1] : end_all;
    }
    // fields are stored as soon as they are known (by lane 0, or lane
    // `axis` for the variance sums); a rejected size needs only n and kind
    CompSums *out = sums + c;
    if (n < a.min_size || n > a.max_size) {
      if (lane == 0) {
        out->n = n;
        out->kind = n < a.min_size ? kTooSmall : kTooLarge;
      }
      continue;
    }
    // the first round's members in registers: lane + 64 j, j < 4
    uint32_t m_idx[4], m_val[4];
    int32_t m_z[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const uint32_t k = lane + 64u * j;
      m_idx[j] = k < n ? mem_index[s0 + k] : 0;
      m_val[j] = k < n ? mem_value[s0 + k] : 0;
      m_z[j] = k < n ? mem_frame[s0 + k] : 0;
    }
    // bbox, peak, sums, and the bounds of the sums' exactness
    uint32_t x0 = 0xFFFFFFFFu, x1 = 0, y0 = 0xFFFFFFFFu, y1 = 0;
    int32_t z0 = INT32_MAX, z1 = INT32_MIN;
    double sv = 0.0, svv = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
    uint32_t vmax = 0, kmax = 0xFFFFFFFFu, zabs = 0, imax = 0;
    int32_t zmax = 0;
    auto visit = [&](uint32_t k, uint32_t idx, uint32_t v, int32_t z) {
      const uint32_t y = row_of(a, idx);
      const uint32_t x = idx - y * a.width;
      x0 = min(x0, x);
      x1 = max(x1, x + 1);
      y0 = min(y0, y);
      y1 = max(y1, y + 1);
      z0 = min(z0, z);
      z1 = max(z1, z + 1);
      const double value = (double)v;
      sv += value;
      svv += value * value;
      sx += value * (double)(2 * x + 1);
      sy += value * (double)(2 * y + 1);
      sz += value * (double)(2 * (int64_t)z + 1);
      zabs = max(zabs, (uint32_t)(z < 0 ? -(int64_t)z : (int64_t)z));
      if (v > vmax) { // lanes see members in order, so first per lane
        vmax = v;
        kmax = k;
        imax = idx;
        zmax = z;
      }
    };
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const uint32_t k = lane + 64u * j;
      if (k < n)
        visit(k, m_idx[j], m_val[j], m_z[j]);
    }
    for (uint32_t k = lane + kTermRound; k < n; k += 64)
      visit(k, mem_index[s0 + k], mem_value[s0 + k], mem_frame[s0 + k]);
    x0 = wave_min_u32(x0);
    x1 = wave_max_u32(x1);
    y0 = wave_min_u32(y0);
    y1 = wave_max_u32(y1);
    z0 = wave_min_i32(z0);
    z1 = wave_max_i32(z1);
    zabs = wave_max_u32(zabs);
    const uint32_t vtop = wave_max_u32(vmax);
    const uint32_t ktop = wave_min_u32(vmax == vtop ? kmax : 0xFFFFFFFFu);
    // Exactness of the reference's double sums: every term and partial
    // sum below 2^52 (with margin) in magnitude.
    const double bound = 4503599627370496.0; // 2^52
    const double top = (double)vtop, cnt = (double)n;
    const double reach =
        (double)max(max(a.width, a.pixels / a.width), zabs) * 2.0 + 1.0;
    const bool exact = cnt * top * top < bound && cnt * top * reach < bound;
    double sum, weighted[3];
    if (exact) {
      sum = wave_sum_f64(sv);
      const double sum_sq = wave_sum_f64(svv);
      weighted[0] = wave_sum_f64(sx) * 0.5;
      weighted[1] = wave_sum_f64(sy) * 0.5;
      weighted[2] = wave_sum_f64(sz) * 0.5;
      double peak[3] = {0.0, 0.0, 0.0};
      if (vtop > 0) { // member ktop was lane (ktop mod 64)'s
        const int owner = (int)(ktop & 63u);
        const uint32_t idx =
            (uint32_t)__builtin_amdgcn_readlane((int)imax, owner);
        const int32_t z = __builtin_amdgcn_readlane(zmax, owner);
        const uint32_t y = row_of(a, idx);
        peak[0] = (double)(idx - y * a.width) + 0.5;
        peak[1] = (double)y + 0.5;
        peak[2] = (double)z + 0.5;
      }
      if (lane == 0) {
        out->sum = sum;
        out->sum_sq = sum_sq;
        for (int axis = 0; axis < 3; axis++) {
          out->weighted[axis] = weighted[axis];
          out->peak[axis] = peak[axis];
        }
        out->bbox[0] = (int32_t)x0;
        out->bbox[1] = (int32_t)x1;
        out->bbox[2] = (int32_t)y0;
        out->bbox[3] = (int32_t)y1;
        out->bbox[4] = z0;
        out->bbox[5] = z1;
        out->n = n;
        out->kind = -1;
      }
    } else {
      CompSums cs; // lane 0's
      if (lane == 0) {
        memset(&cs, 0, sizeof(cs));
        sequential_sums(a, s0, n, mem_frame, mem_index, mem_value, cs);
        cs.n = n;
        cs.kind = -1;
        *out = cs;
      }
      sum = readlane_f64(cs.sum, 0);
      for (int axis = 0; axis < 3; axis++)
        weighted[axis] = readlane_f64(cs.weighted[axis], 0);
    }
    double acc = 0.0; // lane a < 3: axis a's variance sum, in member order
    if (sum > 0.0) {
      // the centroid as finish_spot forms it; then the terms
      double centre[3];
      for (int axis = 0; axis < 3; axis++)
        centre[axis] = weighted[axis] / sum;
      for (uint32_t k0 = 0; k0 < n; k0 += kTermRound) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const uint32_t k = k0 + lane + 64u * j;
          uint32_t idx = m_idx[j], v = m_val[j];
          int32_t z = m_z[j];
          if (k0 > 0) {
            idx = k < n ? mem_index[s0 + k] : 0;
            v = k < n ? mem_value[s0 + k] : 0;
            z = k < n ? mem_frame[s0 + k] : 0;
          }
          const uint32_t y = row_of(a, idx);
          const double value = (double)v;
          const double position[3] = {(double)(idx - y * a.width) + 0.5,
                                      (double)y + 0.5, (double)z + 0.5};
          for (int axis = 0; axis < 3; axis++) {
            const double delta = position[axis] - centre[axis];
            mine[axis * kTermStride + lane + 64u * j] = value * delta * delta;
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane < 3) {
          // Sixteen terms at a time, two sets of eight: each set is read
          // while the other is added. The round's slots past the last
          // member hold +0.0 (a zero value times a finite square), and
          // adding +0.0 to the sum, which is never -0.0, leaves it
          // unchanged, so the last sixteen need no tail.
          const uint32_t m = (min(kTermRound, n - k0) + 15u) & ~15u;
          const double *row = mine + lane * kTermStride;
          double p[8], q[8];
#pragma unroll
          for (int t = 0; t < 8; t++)
            p[t] = row[t];
          for (uint32_t i = 0; i < m; i += 16) {
#pragma unroll
            for (int t = 0; t < 8; t++)
              q[t] = row[i + 8 + t];
#pragma unroll
            for (int t = 0; t < 8; t++)
              acc += p[t];
            if (i + 16 < m) {
#pragma unroll
              for (int t = 0; t < 8; t++)
                p[t] = row[i + 16 + t];
            }
#pragma unroll
            for (int t = 0; t < 8; t++)
              acc += q[t];
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    }
    if (lane < 3)
      out->delta_sq[lane] = acc;
  }
}

// The rest of emit() (src/dials_spots.cc:297-360) one component a lane:
// centroid, variance, the single-frame z rule, the separation test; the
// spot row and the tallies.
__global__ __launch_bounds__(256) void finish_components(
    LabelArgs a, const uint32_t *n_starts, const CompSums *sums,
    SpotOut *out, unsigned long long *spot_pack, uint32_t *tallies) {
  const uint32_t comps = *n_starts;
  uint32_t tally[4] = {0, 0, 0, 0};
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < comps;
       c += gridDim.x * blockDim.x) {
    const CompSums cs = sums[c];
    SpotOut r;
    memset(&r, 0, sizeof(r));
    int kind = cs.kind;
    if (kind < 0) {
      sf_spot &spot = r.spot;
      spot.n_signal = cs.n;
      for (int q = 0; q < 6; q++)
        spot.bbox[q] = cs.bbox[q];
      kind = finish_spot(a, spot, cs.sum, cs.sum_sq, cs.weighted, cs.peak,
                         [&](const double (&)[3], double (&delta_sq)[3]) {
                           for (int axis = 0; axis < 3; axis++)
                             delta_sq[axis] = cs.delta_sq[axis];
                         });
    }
    if (kind == kAccepted) {
      r.accepted = 1;
      r.members = cs.n;
    }
    out[c] = r;
    spot_pack[c] = ((unsigned long long)r.accepted << 32) | r.members;
    tally[kind]++;
  }
  store_tallies(tally, tallies);
}

// Appends the kept spots and their pixels (a member a thread): member m of
// component c is pixel row (c's first pixel row) + (m - c's first member).
__global__ void write_pixels(const SpotOut *res, const uint32_t *n_comps,
                             const uint32_t *n_members,
                             const uint32_t *comp_mem,
                             const unsigned long long *spot_rank,
                             const int32_t *mem_frame,
                             const uint32_t *mem_index,
                             const uint32_t *mem_value,
                             const uint32_t *mem_comp,
                             const unsigned long long *totals,
                             sf_spot *spots, unsigned long long spots_from,
                             sf_pixel *pixels,
                             unsigned long long pixels_from) {
  // rows are numbered over the whole table; `spots` and `pixels` hold it
  // from row spots_from and pixels_from on
  const unsigned long long spot_base = totals[0], pixel_base = totals[1];
  sf_spot *spot_rows = spots + (spot_base - spots_from);
  sf_pixel *pixel_rows = pixels + (pixel_base - pixels_from);
  // the kept spots' rows, after what earlier chunks wrote
  const uint32_t comps = *n_comps;
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < comps;
       c += gridDim.x * blockDim.x) {
    if (!res[c].accepted)
      continue;
    sf_spot sp = res[c].spot;
    sp.first = pixel_base + (uint32_t)spot_rank[c];
    spot_rows[spot_rank[c] >> 32] = sp;
  }
  const uint32_t members = *n_members;
  for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < members;
       m += gridDim.x * blockDim.x) {
    const uint32_t c = mem_comp[m];
    if (!res[c].accepted)
      continue;
    sf_pixel px;
    px.frame = mem_frame[m];
    px.index = mem_index[m];
    px.value = mem_value[m];
    pixel_rows[(uint32_t)spot_rank[c] + (m - comp_mem[c])] = px;
  }
}

// The chunk's close, one block: counts (groups, too_small, too_large,
// separated, accepted, at 1..5) += the centroid passes' tally rows and the
// chunk's components; the spots and pixels written so far advance, also in
// the host's mapped copy.
__global__ __launch_bounds__(256) void close_chunk(
    const uint32_t *rows, uint32_t n_rows, const uint32_t *n_comps,
    const unsigned long long *rank, const SpotOut *res,
    unsigned long long *counts, unsigned long long *totals,
    unsigned long long *host_totals) {
  __shared__ unsigned long long part[4][4];
  unsigned long long t4[4] = {0, 0, 0, 0};
  for (uint32_t r = threadIdx.x; r < n_rows; r += blockDim.x)
    for (int t = 0; t < 4; t++)
      t4[t] += rows[4 * r + t];
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  for (int t = 0; t < 4; t++) {
    // the row counts stay below 2^32 in a chunk
    const uint32_t w = wave_reduce((uint32_t)t4[t], 0u,
                                   [](uint32_t p, uint32_t q) { return p + q; });
    if (lane == 0)
      part[wave][t] = w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int slot[4] = {5, 2, 3, 4}; // kAccepted, kTooSmall, ...
    for (int t = 0; t < 4; t++)
      counts[slot[t]] += part[0][t] + part[1][t] + part[2][t] + part[3][t];
    const uint32_t comps = *n_comps;
    counts[1] += comps;
    if (comps) {
      const uint32_t last = comps - 1;
      totals[0] += (rank[last] >> 32) + res[last].accepted;
      totals[1] += (uint32_t)rank[last] + res[last].members;
      host_totals[0] = totals[0];
      host_totals[1] = totals[1];
      __threadfence_system();
    }
  }
}

// The open runs' pixels, in node order, become the next frontier; each
// points at its component's first pixel -- the first pixel of its root run,
// which is open with it. One thread per run: the open runs (few) copy their
// nodes, the rest return.
__global__ void carry_out(LabelArgs a, const uint32_t *open_len,
                          const uint32_t *open_at, const uint32_t *root,
                          const uint32_t *run_head, const uint32_t *n_runs,
                          NodeSrc nodes, int32_t *carry_frame,
                          uint32_t *carry_index, uint32_t *carry_value,
                          uint32_t *carry_root, uint32_t *n_carry_out) {
  const uint32_t runs = *n_runs;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const uint32_t last = runs - 1;
    *n_carry_out = open_at[last] + open_len[last]; // host-mapped
    __threadfence_system();
  }
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < runs;
       r += gridDim.x * blockDim.x) {
    const uint32_t len = open_len[r];
    if (!len)
      continue;
    const uint32_t at = open_at[r], head = run_head[r];
    const uint32_t to_root = open_at[root[r]];
    for (uint32_t k = 0; k < len; k++) {
      node_at(a, nodes, head + k, carry_frame[at + k], carry_index[at + k],
              carry_value[at + k]);
      carry_root[at + k] = to_root;
    }
  }
}

// Where each frame's nodes start among a chunk's own nodes, from the counts
// on the device: first[k] = the counts of frames before k, where a list
// over its capacity counts 0; first[n_frames] their total. Each frame's
// count as found also goes to host_counts (host-mapped). One block. This
// is add_plan's, which copies nothing else: the plan's lists are gathered
// into the chunk's node arrays (gather_chunk) on the snapshot stream.
__global__ __launch_bounds__(256) void chunk_offsets(const ChunkDesc *frames,
                                                     uint32_t n_frames,
                                                     uint32_t *first,
                                                     uint32_t *host_counts) {
  __shared__ uint32_t part[4];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0)
    carry = 0;
  __syncthreads();
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  for (uint32_t k0 = 0; k0 < n_frames; k0 += 256) {
    const uint32_t k = k0 + threadIdx.x;
    uint32_t c = 0;
    if (k < n_frames) {
      const uint32_t found = *frames[k].count;
      host_counts[k] = found;
      c = found <= frames[k].capacity ? found : 0u;
    }
    const uint32_t incl = wave_inclusive(c);
    if (lane == 63)
      part[wave] = incl;
    __syncthreads();
    uint32_t before = carry;
    for (uint32_t w = 0; w < wave; w++)
      before += part[w];
    if (k < n_frames)
      first[k] = before + incl - c;
    __syncthreads();
    if (threadIdx.x == 255)
      carry = before + incl;
    __syncthreads();
  }
  if (threadIdx.x == 0)
    first[n_frames] = carry;
  __threadfence_system();
}

// ---------------------------------------------------------------------------
// A stable LSD radix sort of (key, value) pairs, for the component sort's
// 32-bit keys: three passes of 9-bit digits. A pass is a histogram per tile
// of 1024 pairs, a wave per digit turning the histograms into each tile's
// first place within the digit, and a scatter that ranks each tile's pairs
// per digit in input order (a wave's 64 pairs by ballot matches of their
// digits, the waves of a tile and the rounds of a wave in order).
// ---------------------------------------------------------------------------
constexpr uint32_t kSortBits = 9;
constexpr uint32_t kSortBuckets = 1u << kSortBits;
constexpr uint32_t kSortTile = 1024; // pairs a block takes: 4 waves x 256

__global__ __launch_bounds__(256) void sort_histogram(const uint32_t *keys,
                                                      uint32_t n, int shift,
                                                      uint32_t *hist,
                                                      uint32_t tiles) {
  __shared__ uint32_t counts[kSortBuckets];
  for (uint32_t d = threadIdx.x; d < kSortBuckets; d += blockDim.x)
    counts[d] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * kSortTile;
  for (uint32_t k = threadIdx.x; k < kSortTile; k += blockDim.x)
    if (base + k < n)
      atomicAdd(&counts[(keys[base + k] >> shift) & (kSortBuckets - 1)], 1u);
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < kSortBuckets; d += blockDim.x)
    hist[d * tiles + blockIdx.x] = counts[d];
}

// Each digit's row of hist (digit-major, tiles per digit) becomes each
// tile's first place within the digit, and the digit's total goes to
// totals[d]: a wave per digit.
__global__ __launch_bounds__(256) void sort_offsets(uint32_t *hist,
                                                    uint32_t tiles,
                                                    uint32_t *totals) {
  const uint32_t lane = lane_id();
  const uint32_t d = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (d >= kSortBuckets)
    return;
  uint32_t run = 0;
  for (uint32_t t0 = 0; t0 < tiles; t0 += 64) {
    const uint32_t t = t0 + lane;
    const uint32_t v = t < tiles ? hist[d * tiles + t] : 0u;
    const uint32_t incl = wave_inclusive(v);
    if (t < tiles)
      hist[d * tiles + t] = run + incl - v;
    run += (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
  }
  if (lane == 0)
    totals[d] = run;
}

// The lanes of this round with the same digit as this lane (all lanes
// active; `live` marks the lanes that hold a pair).
__device__ __forceinline__ uint64_t digit_peers(uint32_t digit,
                                                uint64_t live) {
  uint64_t mask = live;
#pragma unroll
  for (uint32_t b = 0; b < kSortBits; b++) {
    const uint64_t set = ballot((digit >> b) & 1u);
    mask &= ((digit >> b) & 1u) ? set : ~set;
  }
  return mask;
}

__global__ __launch_bounds__(256) void sort_scatter(
    const uint32_t *keys_in, const uint32_t *values_in, uint32_t n,
    int shift, const uint32_t *offsets, const uint32_t *totals,
    uint32_t tiles, uint32_t *keys_out, uint32_t *values_out) {
  __shared__ uint32_t start[4][kSortBuckets]; // per wave and digit
  __shared__ uint32_t base[kSortBuckets];     // each digit's first place
  __shared__ uint32_t wave_sum[4];
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  { // the digits' first places: an exclusive prefix of their totals, two
    // digits a thread
    const uint32_t v0 = totals[2 * threadIdx.x], v1 = totals[2 * threadIdx.x + 1];
    const uint32_t incl = wave_inclusive(v0 + v1);
    if (lane == 63)
      wave_sum[wave] = incl;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t w = 0; w < wave; w++)
      before += wave_sum[w];
    base[2 * threadIdx.x] = before + incl - (v0 + v1);
    base[2 * threadIdx.x + 1] = before + incl - v1;
  }
  const uint64_t below = (1ull << lane) - 1ull;
  for (uint32_t d = lane; d < kSortBuckets; d += 64)
    start[wave][d] = 0;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const uint32_t first = blockIdx.x * kSortTile + wave * (kSortTile / 4);
  // this wave's pairs per digit
  for (uint32_t r = 0; r < kSortTile / 4; r += 64) {
    const uint32_t i = first + r + lane;
    const bool live = i < n;
    const uint32_t digit =
        live ? (keys_in[i] >> shift) & (kSortBuckets - 1) : 0u;
    const uint64_t peers = digit_peers(digit, ballot(live));
    if (live && (peers & below) == 0)
      start[wave][digit] += (uint32_t)__builtin_popcountll(peers);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  __syncthreads();
  // each wave's first place per digit: the tile's, after earlier waves'
  for (uint32_t d = threadIdx.x; d < kSortBuckets; d += blockDim.x) {
    uint32_t run = base[d] + offsets[d * tiles + blockIdx.x];
    for (uint32_t w = 0; w < 4; w++) {
      const uint32_t c = start[w][d];
      start[w][d] = run;
      run += c;
    }
  }
  __syncthreads();
  for (uint32_t r = 0; r < kSortTile / 4; r += 64) {
    const uint32_t i = first + r + lane;
    const bool live = i < n;
    const uint32_t key = live ? keys_in[i] : 0u;
    const uint32_t value = live ? values_in[i] : 0u;
    const uint32_t digit = (key >> shift) & (kSortBuckets - 1);
    const uint64_t peers = digit_peers(digit, ballot(live));
    const uint32_t at =
        start[wave][digit] + (uint32_t)__builtin_popcountll(peers & below);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (live) {
      keys_out[at] = key;
      values_out[at] = value;
      if ((peers & below) == 0)
        start[wave][digit] += (uint32_t)__builtin_popcountll(peers);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
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
  void release() {
    if (ptr)
      (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
  }
};

// A result table that only grows: a list of device segments, each holding
// the rows from its `base` (the table's row count when it was started) on.
// When a chunk's rows may not fit in the last segment, a new one (twice as
// large, or larger) starts; nothing is copied or freed until release(), so
// growing never waits for the device. A large hipMalloc can itself take a
// long time (from hundreds of milliseconds to seconds has been seen, and
// other threads' allocations wait meanwhile), so once the last segment has
// room for fewer than three more chunks like the current one the next is
// allocated ahead, on a helper thread, and taken when it is needed; and a
// caller that knows its series' length reserves the rows up front.
template <typename T> struct Segments {
  struct Seg {
    T *ptr;
    size_t base, cap;
  };
  std::vector<Seg> segs;
  int device = 0;
  std::thread ahead;     // allocating `spare`
  T *spare = nullptr;    // written by `ahead`, read after joining it
  size_t spare_cap = 0;
  // room for `need` rows after the `total` already written
  hipError_t room(size_t total, size_t need, const Seg *&out) {
    if (!segs.empty()) {
      const Seg &last = segs.back();
      if (total >= last.base && total - last.base + need <= last.cap) {
        out = &last;
        // the chunks after the next few may not fit: start the next segment
        if (last.cap - (total - last.base + need) < 3 * need && !spare &&
            !ahead.joinable())
          start_ahead(2 * last.cap);
        return hipSuccess;
      }
    }
    size_t cap = segs.empty() ? 4 * need : 2 * segs.back().cap;
    cap = std::max<size_t>(cap, 1024);
    while (cap < need)
      cap *= 2;
    if (ahead.joinable())
      ahead.join();
    T *fresh = nullptr;
    if (spare && spare_cap >= need) {
      fresh = spare;
      cap = spare_cap;
    } else {
      if (spare)
        (void)hipFree(spare);
      const hipError_t e = hipMalloc(&fresh, cap * sizeof(T));
      if (e != hipSuccess) {
        spare = nullptr;
        spare_cap = 0;
        return e;
      }
    }
    spare = nullptr;
    spare_cap = 0;
    segs.push_back(Seg{fresh, total, cap});
    out = &segs.back();
    return hipSuccess;
  }
  // room for `rows` more rows after `total` without allocating again: a
  // new last segment (base `total`) unless the last one has it
  hipError_t reserve(size_t total, size_t rows) {
    if (!segs.empty()) {
      const Seg &last = segs.back();
      if (total >= last.base && total - last.base + rows <= last.cap)
        return hipSuccess;
    }
    if (ahead.joinable())
      ahead.join();
    const size_t cap = std::max<size_t>(rows, 1024);
    T *fresh = nullptr;
    if (spare && spare_cap >= cap) {
      fresh = spare;
    } else {
      const hipError_t e = hipMalloc(&fresh, cap * sizeof(T));
      if (e != hipSuccess)
        return e;
      if (spare)
        (void)hipFree(spare);
    }
    segs.push_back(Seg{fresh, total, spare == fresh ? spare_cap : cap});
    spare = nullptr;
    spare_cap = 0;
    return hipSuccess;
  }
  void start_ahead(size_t cap) {
    ahead = std::thread([this, cap] {
      T *p = nullptr;
      if (hipSetDevice(device) == hipSuccess &&
          hipMalloc(&p, cap * sizeof(T)) == hipSuccess) {
        spare = p;
        spare_cap = cap;
      }
    });
  }
  // the first `total` rows, in order, into host memory
  hipError_t read(T *host, size_t total) const {
    for (size_t k = 0; k < segs.size(); k++) {
      const size_t lo = segs[k].base;
      const size_t hi = std::min(total, k + 1 < segs.size() ? segs[k + 1].base
                                                           : total);
      if (hi > lo) {
        const hipError_t e = hipMemcpy(host + lo, segs[k].ptr,
                                       (hi - lo) * sizeof(T),
                                       hipMemcpyDeviceToHost);
        if (e != hipSuccess)
          return e;
      }
    }
    return hipSuccess;
  }
  ~Segments() {
    if (ahead.joinable())
      ahead.join();
  }
  void release() {
    if (ahead.joinable())
      ahead.join();
    if (spare)
      (void)hipFree(spare);
    spare = nullptr;
    spare_cap = 0;
    for (auto &g : segs)
      (void)hipFree(g.ptr);
    segs.clear();
  }
};

} // namespace sf

struct sf_labeller {
  size_t height = 0, width = 0;
  sf_options options{};
  int device = 0;
  hipStream_t stream = nullptr;
  // add_plan's chunks: the plan's lists are gathered on the snapshot stream
  // into one of two sides' node sets (chunk_offsets, gather_chunk), and a
  // worker thread labels the chunk once the gather is done, so add_plan
  // returns without waiting for the plan or the labelling. A chunk is
  // handed over while the worker may still hold the one before it
  // (jobs_given - jobs_done <= 2).
  struct Job {
    std::vector<sf::ChunkDesc> frames;
    int side = 0;
  };
  std::thread worker;
  std::mutex mu;
  std::condition_variable cv;
  bool job_ready = false, busy = false, stopping = false;
  uint64_t jobs_given = 0, jobs_done = 0;
  Job job;
  int deferred_status = 0; // the first failure of a handed-over chunk
  std::string deferred_error;
  // A chunk's own nodes (sf::ChunkNodes) and its frames' starts among them:
  // two sets for add_plan's sides, gathered on the snapshot stream, and one
  // for the chunks that run_chunk gathers itself.
  struct NodeSet {
    sf::DevArray<uint32_t> index, value, slot, tile_heads, first;
    sf::DevArray<uint64_t> heads;
    hipError_t ensure(size_t nodes, size_t frames) {
      const size_t n = std::max<size_t>(nodes, 1);
      hipError_t e;
      if ((e = index.ensure(n)) != hipSuccess ||
          (e = value.ensure(n)) != hipSuccess ||
          (e = slot.ensure(n)) != hipSuccess ||
          (e = heads.ensure(n / 64 + 1)) != hipSuccess ||
          (e = tile_heads.ensure(n / sf::kTile + 1)) != hipSuccess)
        return e;
      return first.ensure(frames + 1);
    }
    sf::ChunkNodes view() const {
      return {index.ptr, value.ptr, slot.ptr, heads.ptr, tile_heads.ptr};
    }
    void release() {
      index.release();
      value.release();
      slot.release();
      tile_heads.release();
      first.release();
      heads.release();
    }
  };
  NodeSet snap_nodes[2], own_nodes;
  uint32_t *h_first = nullptr; // pinned staging of own_nodes.first
  size_t h_first_cap = 0;
  sf::DevArray<sf::ChunkDesc> snap_table[2];
  sf::ChunkDesc *h_snap_table[2] = {nullptr, nullptr}; // pinned staging
  uint32_t *h_snap_counts[2] = {nullptr, nullptr};     // host-mapped
  uint32_t *hd_snap_counts[2] = {nullptr, nullptr};
  size_t snap_frames_cap[2] = {0, 0};
  hipEvent_t snap_ready[2] = {nullptr, nullptr}, snap_free[2] = {nullptr, nullptr};
  // add_plan's snapshots run on a stream of their own, after the plan's run
  // that wrote the lists (snap_after), so that the plan's later work does not
  // wait for them
  hipStream_t snap_stream = nullptr;
  hipEvent_t snap_after[2] = {nullptr, nullptr};
  int snap_side = 0;
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
  // pinned and mapped, written by the kernels: the frontier size of the
  // last chunk, the current chunk's run count; the spots and pixels written
  uint32_t *h_scalars = nullptr, *hd_scalars = nullptr;
  unsigned long long *h_totals = nullptr, *hd_totals = nullptr;
  hipEvent_t runs_ready = nullptr;
  hipEvent_t synced = nullptr; // recorded and polled to wait for the stream
  sf::ChunkDesc *h_chunk = nullptr; // pinned staging of the chunk table
  uint32_t h_chunk_cap = 0;
  // per node
  sf::DevArray<uint32_t> run_of;
  // per run (at most one per node)
  sf::DevArray<uint32_t> run_head, run_idx, run_last, run_slot, parent, root,
      last_slot, pos_of;
  sf::DevArray<uint32_t> run_rows; // per (slot, row): first run
  sf::DevArray<uint32_t> tile_spill; // per tile of unite_tiles: left over
  sf::DevArray<uint64_t> keys, keys_sorted;
  sf::DevArray<uint32_t> values, order, open_len, open_at;
  sf::DevArray<unsigned long long> run_pack, run_rank;
  // per component
  sf::DevArray<uint32_t> starts, comp_mem;
  sf::DevArray<uint32_t> tallies; // per centroid-pass block: kinds emitted
  sf::DevArray<sf::SpotOut> results;
  sf::DevArray<sf::CompSums> comp_sums;
  sf::DevArray<uint32_t> sort_hist; // per digit and tile of the run sort
  sf::DevArray<unsigned long long> spot_pack, spot_rank;
  sf::DevArray<uint8_t> temp;
  sf::DevArray<sf::ChunkDesc> chunk;
  uint32_t *d_scalars = nullptr; // 8 device counters of a chunk
  // results
  sf::Segments<sf_spot> spots;
  sf::Segments<sf_pixel> pixels;
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

// The component sort. rocPRIM sorts up to a million items by merge sort
// unless told otherwise (about twenty passes over a chunk's runs); above 64K
// items these configurations take the radix (onesweep) path, and keys of 25
// to 30 bits go in three 10-bit passes instead of four 8-bit ones.
constexpr size_t kMergeSortLimit = 64 * 1024;
using Sort10 = rocprim::radix_sort_config<
    rocprim::default_config, rocprim::default_config,
    rocprim::radix_sort_onesweep_config<
        rocprim::kernel_config<1024, 16>, rocprim::kernel_config<1024, 16>, 10,
        rocprim::block_radix_rank_algorithm::match>,
    kMergeSortLimit>;
using Sort8 = rocprim::radix_sort_config<rocprim::default_config,
                                         rocprim::default_config,
                                         rocprim::default_config,
                                         kMergeSortLimit>;

hipError_t sort32(void *temp, size_t &bytes, const uint32_t *keys_in,
                  uint32_t *keys_out, const uint32_t *values_in,
                  uint32_t *values_out, uint32_t n, int bits,
                  hipStream_t s) {
  if (bits > 24 && bits <= 30)
    return rocprim::radix_sort_pairs<Sort10>(temp, bytes, keys_in, keys_out,
                                             values_in, values_out, n, 0, bits,
                                             s);
  return rocprim::radix_sort_pairs<Sort8>(temp, bytes, keys_in, keys_out,
                                          values_in, values_out, n, 0, bits,
                                          s);
}

hipError_t sort64(void *temp, size_t &bytes, const uint64_t *keys_in,
                  uint64_t *keys_out, const uint32_t *values_in,
                  uint32_t *values_out, uint32_t n, hipStream_t s) {
  return rocprim::radix_sort_pairs<Sort8>(temp, bytes, keys_in, keys_out,
                                          values_in, values_out, n, 0, 64, s);
}

// Waits for an event by polling it: a short spin, then a yield between
// queries. A blocking wait sleeps until the device's interrupt, whose
// wake-up has taken from tens of microseconds to milliseconds, and the
// labeller's waits sit between launches that keep the device busy.
hipError_t poll_event(hipEvent_t e) {
  for (uint32_t k = 0;; k++) {
    const hipError_t q = hipEventQuery(e);
    if (q != hipErrorNotReady)
      return q;
    if (k < 64) {
      for (int i = 0; i < 32; i++)
        __builtin_ia32_pause();
    } else {
      std::this_thread::yield();
    }
  }
}

// Everything enqueued on the labeller's stream so far, done.
hipError_t poll_stream(sf_labeller *l) {
  const hipError_t e = hipEventRecord(l->synced, l->stream);
  return e == hipSuccess ? poll_event(l->synced) : e;
}

// Label one chunk whose frame lists are device-resident. `frames` holds
// list/count/capacity/number for each frame; offsets are filled here from
// the host counts. `gathered`: the chunk's own nodes, already gathered from
// the lists (add_plan's side); otherwise they are gathered here, on
// `producer` when given.
int run_chunk(sf_labeller *l, std::vector<sf::ChunkDesc> &frames,
              const std::vector<uint32_t> &counts, bool closing,
              hipStream_t producer,
              const sf_labeller::NodeSet *gathered = nullptr) {
  hipStream_t s = l->stream;
  // the frontier and the scratch of the previous chunk must be done with
  LAB_CHECK(poll_stream(l), "waiting for the previous chunk");
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
  a.inv_width = 1.0 / (double)l->width;
  // outputs: at worst every node is kept
  l->bound_pixels += n - l->n_carry;
  l->bound_spots += (n - l->n_carry) / std::max<uint64_t>(1, a.min_size) + 1;
  // the segments the chunk's spots and pixels go to: the totals as the
  // last chunk's kernels left them in mapped memory (the stream is idle)
  const sf::Segments<sf_spot>::Seg *spot_seg = nullptr;
  const sf::Segments<sf_pixel>::Seg *pixel_seg = nullptr;
  {
    const unsigned long long totals[2] = {l->h_totals[0], l->h_totals[1]};
    LAB_CHECK(l->spots.room(totals[0],
                            n / std::max<uint64_t>(1, a.min_size) + 1,
                            spot_seg),
              "growing the spot table");
    LAB_CHECK(l->pixels.room(totals[1], (size_t)n + 1, pixel_seg),
              "growing the pixel table");
  }
  // the chunk table, through pinned staging (the stream is idle, so the
  // last chunk's copy out of it is done)
  if (n_frames > l->h_chunk_cap) {
    if (l->h_chunk)
      (void)hipHostFree(l->h_chunk);
    l->h_chunk = nullptr;
    l->h_chunk_cap = 0;
    LAB_CHECK(hipHostMalloc(&l->h_chunk, n_frames * sizeof(sf::ChunkDesc)),
              "chunk staging");
    l->h_chunk_cap = n_frames;
  }
  if (n_frames)
    std::memcpy(l->h_chunk, frames.data(), n_frames * sizeof(sf::ChunkDesc));
  l->last_count = 0;
  if (n_frames && !closing && !l->options.two_d)
    l->last_count = counts[n_frames - 1] <= frames[n_frames - 1].capacity
                        ? counts[n_frames - 1]
                        : 0;
  if (n == 0) {
    // nothing carried and every frame empty (or over its capacity): no
    // signal pixel to count
    *l->h_scalars = 0;
    return SF_OK;
  }
  const uint32_t n_chunk = n - l->n_carry;
  const uint32_t chunk_tiles = (n_chunk + sf::kTile - 1) / sf::kTile;
  const uint32_t carry_tiles = (l->n_carry + sf::kTile - 1) / sf::kTile;
  if (!gathered) {
    LAB_CHECK(l->own_nodes.ensure(n_chunk, n_frames), "alloc");
    if (n_frames + 1 > l->h_first_cap) {
      if (l->h_first)
        (void)hipHostFree(l->h_first);
      l->h_first = nullptr;
      l->h_first_cap = 0;
      LAB_CHECK(hipHostMalloc(&l->h_first, (n_frames + 1) * sizeof(uint32_t)),
                "chunk staging");
      l->h_first_cap = n_frames + 1;
    }
    for (uint32_t k = 0; k < n_frames; k++)
      l->h_first[k] = frames[k].offset - l->n_carry;
    l->h_first[n_frames] = n_chunk;
  }
  const sf_labeller::NodeSet &nodes = gathered ? *gathered : l->own_nodes;
  LAB_CHECK(l->run_of.ensure(n), "alloc");
  LAB_CHECK(l->run_head.ensure(n), "alloc");
  LAB_CHECK(l->run_idx.ensure(n), "alloc");
  LAB_CHECK(l->run_last.ensure(n), "alloc");
  LAB_CHECK(l->pos_of.ensure(n), "alloc");
  LAB_CHECK(l->run_slot.ensure(n), "alloc");
  LAB_CHECK(l->parent.ensure(n), "alloc");
  LAB_CHECK(l->root.ensure(n), "alloc");
  LAB_CHECK(l->last_slot.ensure(n), "alloc");
  LAB_CHECK(l->keys.ensure(n), "alloc");
  LAB_CHECK(l->keys_sorted.ensure(n), "alloc");
  LAB_CHECK(l->values.ensure(n), "alloc");
  LAB_CHECK(l->order.ensure(n), "alloc");
  LAB_CHECK(l->open_len.ensure(n), "alloc");
  LAB_CHECK(l->open_at.ensure(n), "alloc");
  LAB_CHECK(l->run_pack.ensure(n), "alloc");
  LAB_CHECK(l->run_rank.ensure(n), "alloc");
  LAB_CHECK(l->starts.ensure(n), "alloc");
  LAB_CHECK(l->comp_mem.ensure(n), "alloc");
  LAB_CHECK(l->results.ensure(n), "alloc");
  LAB_CHECK(l->spot_pack.ensure(n), "alloc");
  LAB_CHECK(l->spot_rank.ensure(n), "alloc");
  const int in = l->carry_side, outside = 1 - l->carry_side;
  LAB_CHECK(l->carry_frame[outside].ensure(n), "alloc");
  LAB_CHECK(l->carry_index[outside].ensure(n), "alloc");
  LAB_CHECK(l->carry_value[outside].ensure(n), "alloc");
  LAB_CHECK(l->carry_root[outside].ensure(n), "alloc");
  LAB_CHECK(l->chunk.ensure(std::max<uint32_t>(n_frames, 1)), "alloc");
  uint32_t *keys32 = reinterpret_cast<uint32_t *>(l->keys.ptr);
  uint32_t *keys32_sorted = reinterpret_cast<uint32_t *>(l->keys_sorted.ptr);
  // d_scalars: components, closed members, closed runs, (unused), runs
  uint32_t *d_comps = l->d_scalars, *d_closed_mem = l->d_scalars + 1,
           *d_closed_runs = l->d_scalars + 2,
           *n_runs = l->d_scalars + 4;
  LAB_CHECK(hipMemsetAsync(l->d_scalars, 0, 8 * sizeof(uint32_t), s),
            "memset");
  if (n_frames)
    LAB_CHECK(hipMemcpyAsync(l->chunk.ptr, l->h_chunk,
                             n_frames * sizeof(sf::ChunkDesc),
                             hipMemcpyHostToDevice, s),
              "chunk table");
  if (!gathered && n_chunk) {
    LAB_CHECK(hipMemcpyAsync(l->own_nodes.first.ptr, l->h_first,
                             (n_frames + 1) * sizeof(uint32_t),
                             hipMemcpyHostToDevice, s),
              "chunk table");
    // The gather reads the producer's lists: run it on the producer's
    // stream so that the producer cannot overwrite them first, and have this
    // stream wait for it.
    hipStream_t gs = producer ? producer : s;
    if (producer) {
      LAB_CHECK(hipEventRecord(l->done, s), "event");
      LAB_CHECK(hipStreamWaitEvent(gs, l->done, 0), "event wait");
    }
    sf::gather_chunk<<<std::min<uint32_t>(chunk_tiles, 4096), 256, 0, gs>>>(
        a, l->chunk.ptr, n_frames, l->own_nodes.first.ptr,
        l->own_nodes.view());
    LAB_CHECK(hipGetLastError(), "gather");
    if (producer) {
      LAB_CHECK(hipEventRecord(l->done, gs), "event");
      LAB_CHECK(hipStreamWaitEvent(s, l->done, 0), "event wait");
    }
  }
  const uint32_t grid = grid_for(n);
  sf::number_runs<<<chunk_tiles + carry_tiles, sf::kTileThreads, 0, s>>>(
      a, nodes.index.ptr, nodes.slot.ptr, nodes.heads.ptr,
      nodes.tile_heads.ptr, chunk_tiles, l->carry_index[in].ptr,
      l->carry_root[in].ptr, l->run_of.ptr, l->run_head.ptr, l->run_idx.ptr,
      l->run_last.ptr, l->run_slot.ptr, l->parent.ptr, l->last_slot.ptr,
      n_runs, l->hd_scalars + 1, l->d_counts);
  // where the later kernels read each node's frame, index and value
  const sf::NodeSrc node_src{l->carry_frame[in].ptr, l->carry_index[in].ptr,
                             l->carry_value[in].ptr, nodes.index.ptr,
                             nodes.value.ptr,        nodes.slot.ptr,
                             l->chunk.ptr};
  // the host needs the run count for the sort; the kernels up to the keys
  // are queued before it waits
  LAB_CHECK(hipEventRecord(l->runs_ready, s), "event");
  {
    const size_t rows = (size_t)(n_frames + 1) * (l->height + 1);
    LAB_CHECK(l->run_rows.ensure(rows), "alloc");
    sf::fill_run_rows<<<grid_for(rows), 256, 0, s>>>(
        a, l->chunk.ptr, l->run_of.ptr, l->run_idx.ptr, n_runs,
        l->run_rows.ptr);
  }
  // the run count is on the device: bounded grids that stride (a multiple of
  // 8 blocks for the XCD mapping of unite_cross)
  const uint32_t run_grid = std::min<uint32_t>((grid + 7) & ~7u, 4096);
  {
    // tiles: at most ceil(height / kTileMinRows) bands of every frame group
    const uint32_t height = (uint32_t)l->height;
    const uint32_t bands =
        (height + sf::kTileMinRows - 1) / sf::kTileMinRows;
    const uint32_t groups =
        std::max<uint32_t>(1, (n_frames + sf::kTileFrames - 1) /
                                  sf::kTileFrames);
    LAB_CHECK(l->tile_spill.ensure((size_t)bands * groups), "alloc");
    sf::unite_tiles<<<bands * groups, 1024, 0, s>>>(
        a, l->chunk.ptr, l->run_idx.ptr, l->run_last.ptr, l->run_rows.ptr,
        n_runs, l->parent.ptr, l->last_slot.ptr, l->tile_spill.ptr);
    sf::unite_cross<<<run_grid, 256, 0, s>>>(
        a, l->chunk.ptr, l->run_idx.ptr, l->run_last.ptr, l->run_slot.ptr,
        l->run_rows.ptr, n_runs, l->tile_spill.ptr, l->parent.ptr);
  }
  sf::resolve_runs<<<run_grid, 256, 0, s>>>(l->parent.ptr, l->run_slot.ptr,
                                            n_runs, l->root.ptr,
                                            l->last_slot.ptr);
  LAB_CHECK(hipGetLastError(), "label kernels");
  // the keys, the sort and the scans take their sizes from the host
  LAB_CHECK(poll_event(l->runs_ready), "waiting for the run count");
  const uint32_t runs = l->h_scalars[1];
  // Sort keys: 32 bits when (last slot, root run) fit, else 64.
  int root_bits = 1, slot_bits = 1;
  while (root_bits < 32 && (1ull << root_bits) < (uint64_t)runs)
    root_bits++;
  while ((1ull << slot_bits) < (uint64_t)n_frames + 2)
    slot_bits++;
  const bool narrow = root_bits + slot_bits <= 32;
  if (narrow)
    sf::make_keys<uint32_t><<<run_grid, 256, 0, s>>>(
        a, l->root.ptr, l->last_slot.ptr, l->run_head.ptr, n_runs, keys32,
        l->values.ptr, l->open_len.ptr, root_bits);
  else
    sf::make_keys<uint64_t><<<run_grid, 256, 0, s>>>(
        a, l->root.ptr, l->last_slot.ptr, l->run_head.ptr, n_runs,
        l->keys.ptr, l->values.ptr, l->open_len.ptr, 32);
  // temp storage: the largest of the rocPRIM calls
  size_t need = 0, bytes = 0;
  if (narrow)
    LAB_CHECK(sort32(nullptr, bytes, keys32, keys32_sorted, l->values.ptr,
                     l->order.ptr, runs, root_bits + slot_bits, s),
              "sort size");
  else
    LAB_CHECK(sort64(nullptr, bytes, l->keys.ptr, l->keys_sorted.ptr,
                     l->values.ptr, l->order.ptr, runs, s),
              "sort size");
  need = std::max(need, bytes);
  LAB_CHECK(rocprim::exclusive_scan(nullptr, bytes, l->open_len.ptr,
                                    l->open_at.ptr, 0u, (size_t)runs,
                                    rocprim::plus<uint32_t>(), s),
            "scan size");
  need = std::max(need, bytes);
  LAB_CHECK(rocprim::exclusive_scan(nullptr, bytes, l->run_pack.ptr,
                                    l->run_rank.ptr, 0ull, (size_t)runs,
                                    rocprim::plus<unsigned long long>(), s),
            "scan size");
  need = std::max(need, bytes);
  LAB_CHECK(l->temp.ensure(need + 256), "alloc temp");
  const uint32_t rgrid = grid_for(runs);
  bytes = l->temp.cap;
  if (narrow) {
    if (root_bits + slot_bits <= 3 * (int)sf::kSortBits) {
      // three 9-bit passes: keys32 -> keys32_sorted -> keys32 ->
      // keys32_sorted, the values likewise (the inputs are dead after)
      const uint32_t sort_tiles = (runs + sf::kSortTile - 1) / sf::kSortTile;
      LAB_CHECK(l->sort_hist.ensure((size_t)sf::kSortBuckets *
                                    (std::max<uint32_t>(sort_tiles, 1) + 1)),
                "alloc");
      uint32_t *kin = keys32, *kout = keys32_sorted;
      uint32_t *vin = l->values.ptr, *vout = l->order.ptr;
      for (int pass = 0; pass < 3; pass++) {
        const int shift = pass * (int)sf::kSortBits;
        if (sort_tiles) {
          sf::sort_histogram<<<sort_tiles, 256, 0, s>>>(
              kin, runs, shift, l->sort_hist.ptr, sort_tiles);
          uint32_t *totals =
              l->sort_hist.ptr + (size_t)sf::kSortBuckets * sort_tiles;
          sf::sort_offsets<<<sf::kSortBuckets / 4, 256, 0, s>>>(
              l->sort_hist.ptr, sort_tiles, totals);
          sf::sort_scatter<<<sort_tiles, 256, 0, s>>>(
              kin, vin, runs, shift, l->sort_hist.ptr, totals, sort_tiles,
              kout, vout);
        }
        std::swap(kin, kout);
        std::swap(vin, vout);
      }
    } else {
      LAB_CHECK(sort32(l->temp.ptr, bytes, keys32, keys32_sorted,
                       l->values.ptr, l->order.ptr, runs,
                       root_bits + slot_bits, s),
                "sorting components");
    }
    // closed keys stay below 2^(root_bits + slot_bits) - 1, so the open key
    // (all ones) sorts last on those bits too
    sf::mark_runs<uint32_t><<<rgrid, 256, 0, s>>>(
        a, keys32_sorted, ~0u, l->order.ptr, l->run_head.ptr, runs,
        l->run_pack.ptr, l->pos_of.ptr, d_closed_runs);
  } else {
    LAB_CHECK(sort64(l->temp.ptr, bytes, l->keys.ptr, l->keys_sorted.ptr,
                     l->values.ptr, l->order.ptr, runs, s),
              "sorting components");
    sf::mark_runs<uint64_t><<<rgrid, 256, 0, s>>>(
        a, l->keys_sorted.ptr, ~0ull, l->order.ptr, l->run_head.ptr, runs,
        l->run_pack.ptr, l->pos_of.ptr, d_closed_runs);
  }
  bytes = l->temp.cap;
  LAB_CHECK(rocprim::exclusive_scan(l->temp.ptr, bytes, l->run_pack.ptr,
                                    l->run_rank.ptr, 0ull, (size_t)runs,
                                    rocprim::plus<unsigned long long>(), s),
            "ranking components and their members");
  sf::list_starts<<<rgrid, 256, 0, s>>>(l->run_pack.ptr, l->run_rank.ptr,
                                        runs, d_closed_runs, l->starts.ptr,
                                        l->comp_mem.ptr, l->d_scalars);
  // the members in emission order, in the sort's input arrays (dead now)
  int32_t *mem_frame = reinterpret_cast<int32_t *>(l->values.ptr);
  uint32_t *mem_index = reinterpret_cast<uint32_t *>(l->keys.ptr);
  uint32_t *mem_value = mem_index + n;
  uint32_t *mem_comp = keys32_sorted; // the sort's output, dead now
  sf::gather_members<<<grid, 256, 0, s>>>(
      a, d_closed_runs, l->run_of.ptr, l->pos_of.ptr, l->run_head.ptr,
      l->run_pack.ptr, l->run_rank.ptr, node_src, mem_frame, mem_index,
      mem_value, mem_comp);
  // the centroid pass: a wave per component for what reads the members,
  // then a lane per component for the rest; each loops over components on a
  // bounded grid
  const uint32_t sums_grid = std::min<uint32_t>(rgrid, 4096);
  const uint32_t finish_grid = std::min<uint32_t>(rgrid, 1024);
  LAB_CHECK(l->comp_sums.ensure(n), "alloc");
  LAB_CHECK(l->tallies.ensure(4 * finish_grid), "alloc");
  sf::component_sums<<<sums_grid, 256, 0, s>>>(
      a, l->comp_mem.ptr, d_comps, d_closed_mem, mem_frame, mem_index,
      mem_value, l->comp_sums.ptr);
  sf::finish_components<<<finish_grid, 256, 0, s>>>(
      a, d_comps, l->comp_sums.ptr, l->results.ptr, l->spot_pack.ptr,
      l->tallies.ptr);
  bytes = l->temp.cap;
  // components are at most the runs
  LAB_CHECK(rocprim::exclusive_scan(l->temp.ptr, bytes, l->spot_pack.ptr,
                                    l->spot_rank.ptr, 0ull, (size_t)runs,
                                    rocprim::plus<unsigned long long>(), s),
            "ranking spots and their pixels");
  sf::write_pixels<<<grid, 256, 0, s>>>(
      l->results.ptr, d_comps, d_closed_mem, l->comp_mem.ptr,
      l->spot_rank.ptr, mem_frame, mem_index, mem_value, mem_comp,
      l->d_totals, spot_seg->ptr, spot_seg->base, pixel_seg->ptr,
      pixel_seg->base);
  sf::close_chunk<<<1, 256, 0, s>>>(l->tallies.ptr, finish_grid,
                                    d_comps, l->spot_rank.ptr, l->results.ptr,
                                    l->d_counts, l->d_totals, l->hd_totals);
  bytes = l->temp.cap;
  LAB_CHECK(rocprim::exclusive_scan(l->temp.ptr, bytes, l->open_len.ptr,
                                    l->open_at.ptr, 0u, (size_t)runs,
                                    rocprim::plus<uint32_t>(), s),
            "ranking the frontier");
  sf::carry_out<<<rgrid, 256, 0, s>>>(
      a, l->open_len.ptr, l->open_at.ptr, l->root.ptr, l->run_head.ptr,
      n_runs, node_src, l->carry_frame[outside].ptr,
      l->carry_index[outside].ptr, l->carry_value[outside].ptr,
      l->carry_root[outside].ptr, l->hd_scalars);
  LAB_CHECK(hipGetLastError(), "finalise kernels");
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
int plan_add_reader(sf_plan *plan, hipStream_t reader);
int plan_take_places(sf_plan *plan, hipStream_t reader,
                     std::vector<const sf_signal_pixel *> &staging,
                     std::vector<const uint32_t *> &places, bool &taken);
} // namespace sf

namespace {

// Waits until the worker has labelled (enqueued) every handed-over chunk
// but `ahead`, and the hand-over slot is free; reports the first failure
// among them, once, as this thread's error.
int drain(sf_labeller *l, uint64_t ahead = 0) {
  std::unique_lock<std::mutex> lock(l->mu);
  l->cv.wait(lock, [&] {
    return !l->job_ready && l->jobs_done + ahead >= l->jobs_given;
  });
  const int status = l->deferred_status;
  if (status != SF_OK) {
    g_label_error = l->deferred_error;
    l->deferred_status = SF_OK;
  }
  return status;
}

// The worker's part of a handed-over chunk: once the snapshot is copied,
// the chunk is labelled from it like any other, and the side is marked
// free for the snapshot after next once the labelling has read it.
int label_job(sf_labeller *l, sf_labeller::Job &job) {
  const int side = job.side;
  LAB_CHECK(poll_event(l->snap_ready[side]), "waiting for the plan's lists");
  std::vector<uint32_t> counts(job.frames.size());
  for (size_t k = 0; k < counts.size(); k++)
    counts[k] = l->h_snap_counts[side][k];
  const int status =
      run_chunk(l, job.frames, counts, false, nullptr, &l->snap_nodes[side]);
  LAB_CHECK(hipEventRecord(l->snap_free[side], l->stream), "event");
  return status;
}

void worker_loop(sf_labeller *l) {
  (void)hipSetDevice(l->device);
  std::unique_lock<std::mutex> lock(l->mu);
  for (;;) {
    l->cv.wait(lock, [&] { return l->job_ready || l->stopping; });
    if (!l->job_ready)
      return; // stopping, and nothing handed over
    sf_labeller::Job job = std::move(l->job);
    l->job_ready = false;
    l->busy = true;
    lock.unlock();
    const int status = label_job(l, job);
    lock.lock();
    if (status != SF_OK && l->deferred_status == SF_OK) {
      l->deferred_status = status;
      l->deferred_error = g_label_error;
    }
    l->busy = false;
    l->jobs_done++;
    l->cv.notify_all();
  }
}

void stop_worker(sf_labeller *l) {
  if (!l->worker.joinable())
    return;
  {
    std::lock_guard<std::mutex> lock(l->mu);
    l->stopping = true;
  }
  l->cv.notify_all();
  l->worker.join();
}

// Processes the frames added from the host since the last flush (after
// every handed-over chunk).
int flush_pending(sf_labeller *l) {
  const int drained = drain(l);
  if (drained != SF_OK)
    return drained;
  if (l->pending_frames.empty())
    return SF_OK;
  LAB_CHECK(poll_stream(l), "waiting for the labeller");
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
  ok = ok && hipStreamCreateWithFlags(&l->snap_stream, hipStreamNonBlocking) ==
                 hipSuccess;
  ok = ok && hipEventCreateWithFlags(&l->done, hipEventDisableTiming) ==
                 hipSuccess;
  ok = ok && hipEventCreateWithFlags(&l->runs_ready, hipEventDisableTiming) ==
                 hipSuccess;
  ok = ok && hipEventCreateWithFlags(&l->synced, hipEventDisableTiming) ==
                 hipSuccess;
  ok = ok && hipGetDevice(&l->device) == hipSuccess;
  l->spots.device = l->pixels.device = l->device;
  for (int i = 0; i < 2; i++) {
    ok = ok && hipEventCreateWithFlags(&l->snap_ready[i],
                                       hipEventDisableTiming) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&l->snap_free[i],
                                       hipEventDisableTiming) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&l->snap_after[i],
                                       hipEventDisableTiming) == hipSuccess;
  }
  ok = ok && hipMalloc(&l->d_scalars, 8 * sizeof(uint32_t)) == hipSuccess;
  ok = ok && hipMalloc(&l->d_counts, 6 * sizeof(unsigned long long)) ==
                 hipSuccess;
  ok = ok && hipMalloc(&l->d_totals, 2 * sizeof(unsigned long long)) ==
                 hipSuccess;
  ok = ok && hipHostMalloc(&l->h_scalars, 4 * sizeof(uint32_t),
                           hipHostMallocMapped) == hipSuccess;
  ok = ok && hipHostGetDevicePointer(reinterpret_cast<void **>(&l->hd_scalars),
                                     l->h_scalars, 0) == hipSuccess;
  ok = ok && hipHostMalloc(&l->h_totals, 2 * sizeof(unsigned long long),
                           hipHostMallocMapped) == hipSuccess;
  ok = ok && hipHostGetDevicePointer(reinterpret_cast<void **>(&l->hd_totals),
                                     l->h_totals, 0) == hipSuccess;
  if (ok) {
    ok = hipMemset(l->d_counts, 0, 6 * sizeof(unsigned long long)) ==
             hipSuccess &&
         hipMemset(l->d_totals, 0, 2 * sizeof(unsigned long long)) ==
             hipSuccess;
    std::memset(l->h_scalars, 0, 4 * sizeof(uint32_t));
    std::memset(l->h_totals, 0, 2 * sizeof(unsigned long long));
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
  stop_worker(l); // labels what was handed over first
  if (l->stream)
    (void)hipStreamSynchronize(l->stream);
  if (l->snap_stream) // a gather may still be writing a side
    (void)hipStreamSynchronize(l->snap_stream);
  for (int i = 0; i < 2; i++) {
    if (l->snap_ready[i]) {
      (void)hipEventSynchronize(l->snap_ready[i]);
      (void)hipEventDestroy(l->snap_ready[i]);
    }
    if (l->snap_free[i])
      (void)hipEventDestroy(l->snap_free[i]);
    if (l->snap_after[i])
      (void)hipEventDestroy(l->snap_after[i]);
    l->snap_nodes[i].release();
    l->snap_table[i].release();
    if (l->h_snap_table[i])
      (void)hipHostFree(l->h_snap_table[i]);
    if (l->h_snap_counts[i])
      (void)hipHostFree(l->h_snap_counts[i]);
  }
  for (int i = 0; i < 2; i++) {
    l->carry_frame[i].release();
    l->carry_index[i].release();
    l->carry_value[i].release();
    l->carry_root[i].release();
  }
  l->own_nodes.release();
  if (l->h_first)
    (void)hipHostFree(l->h_first);
  l->run_of.release();
  l->run_rows.release();
  l->run_head.release();
  l->run_idx.release();
  l->run_last.release();
  l->pos_of.release();
  l->run_slot.release();
  l->parent.release();
  l->root.release();
  l->last_slot.release();
  l->keys.release();
  l->keys_sorted.release();
  l->values.release();
  l->order.release();
  l->open_len.release();
  l->open_at.release();
  l->run_pack.release();
  l->run_rank.release();
  l->starts.release();
  l->comp_mem.release();
  l->tallies.release();
  l->results.release();
  l->comp_sums.release();
  l->sort_hist.release();
  l->spot_pack.release();
  l->spot_rank.release();
  l->temp.release();
  l->chunk.release();
  l->spots.release();
  l->pixels.release();
  l->pending_dev.release();
  l->pending_count_dev.release();
  if (l->runs_ready)
    (void)hipEventDestroy(l->runs_ready);
  if (l->synced)
    (void)hipEventDestroy(l->synced);
  if (l->h_totals)
    (void)hipHostFree(l->h_totals);
  if (l->h_chunk)
    (void)hipHostFree(l->h_chunk);
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
  if (l->snap_stream)
    (void)hipStreamDestroy(l->snap_stream);
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
  // Frames added from the host go first, after every handed-over chunk.
  // Otherwise the worker may keep the last chunk while this one is handed
  // over: its snapshot side's last reader, the chunk before that, is done.
  int status = l->pending_frames.empty() ? drain(l, 1) : flush_pending(l);
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
  const uint32_t n_frames = (uint32_t)lists.size();
  if (n_frames == 0)
    return SF_OK;
  // the snapshot side: its last reader (the chunk before last) is enqueued
  const int side = l->snap_side;
  l->snap_side ^= 1;
  uint64_t total = 0;
  for (uint32_t k = 0; k < n_frames; k++)
    total += caps[k];
  if (total > 0xF0000000ull) {
    g_label_error = "a plan whose lists hold more than 4 billion pixels";
    return SF_BAD_ARGUMENT;
  }
  LAB_CHECK(l->snap_nodes[side].ensure(total, n_frames), "snapshot alloc");
  LAB_CHECK(l->snap_table[side].ensure(n_frames), "snapshot alloc");
  if (n_frames > l->snap_frames_cap[side]) {
    if (l->h_snap_table[side])
      (void)hipHostFree(l->h_snap_table[side]);
    if (l->h_snap_counts[side])
      (void)hipHostFree(l->h_snap_counts[side]);
    l->h_snap_table[side] = nullptr;
    l->h_snap_counts[side] = nullptr;
    l->snap_frames_cap[side] = 0;
    LAB_CHECK(hipHostMalloc(&l->h_snap_table[side],
                            n_frames * sizeof(sf::ChunkDesc)),
              "snapshot staging");
    LAB_CHECK(hipHostMalloc(&l->h_snap_counts[side],
                            n_frames * sizeof(uint32_t), hipHostMallocMapped),
              "snapshot counts");
    LAB_CHECK(hipHostGetDevicePointer(
                  reinterpret_cast<void **>(&l->hd_snap_counts[side]),
                  l->h_snap_counts[side], 0),
              "snapshot counts");
    l->snap_frames_cap[side] = n_frames;
  }
  // on the snapshot stream: after the plan's run that wrote the lists (what
  // is enqueued on the plan's stream now) and after the side's last reader,
  // the chunk's nodes gathered from the lists into the side's node set; the
  // plan's next write of its lists waits for the gather, and nothing else
  // the plan's stream does later waits for it. A plan whose pipelined step
  // left its ordering pending has no lists yet: its ordering's places are
  // enqueued here instead, and the nodes are taken from its staging by
  // them (the lists are written only if the plan is ordered later).
  LAB_CHECK(hipEventRecord(l->snap_after[side], producer), "event");
  LAB_CHECK(hipStreamWaitEvent(l->snap_stream, l->snap_after[side], 0),
            "event wait");
  LAB_CHECK(hipStreamWaitEvent(l->snap_stream, l->snap_free[side], 0),
            "event wait");
  std::vector<const sf_signal_pixel *> staging;
  std::vector<const uint32_t *> places;
  bool placed = false;
  status = sf::plan_take_places(plan, l->snap_stream, staging, places, placed);
  if (status != SF_OK)
    return status;
  std::vector<sf::ChunkDesc> frames(n_frames);
  for (uint32_t k = 0; k < n_frames; k++) {
    sf::ChunkDesc &f = l->h_snap_table[side][k];
    f = sf::ChunkDesc{};
    f.list = placed ? staging[k] : lists[k];
    f.place = placed ? places[k] : nullptr;
    f.count = count_ptrs[k];
    f.capacity = caps[k];
    frames[k].capacity = caps[k];
    frames[k].number = (int32_t)(first_frame + (int64_t)k);
    frames[k].linked =
        k > 0 || (l->started && first_frame == l->last_frame + 1);
  }
  LAB_CHECK(hipMemcpyAsync(l->snap_table[side].ptr, l->h_snap_table[side],
                           n_frames * sizeof(sf::ChunkDesc),
                           hipMemcpyHostToDevice, l->snap_stream),
            "snapshot table");
  sf::chunk_offsets<<<1, 256, 0, l->snap_stream>>>(
      l->snap_table[side].ptr, n_frames, l->snap_nodes[side].first.ptr,
      l->hd_snap_counts[side]);
  sf::LabelArgs ga{};
  ga.width = (uint32_t)l->width;
  ga.pixels = (uint32_t)(l->height * l->width);
  ga.inv_width = 1.0 / (double)l->width;
  const uint64_t tiles = (total + sf::kTile - 1) / sf::kTile;
  const uint32_t tile_grid =
      (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(tiles, 4096));
  if (placed) {
    sf::scatter_chunk<<<dim3(64, std::min<uint32_t>(n_frames, 65535)), 256,
                        0, l->snap_stream>>>(
        l->snap_table[side].ptr, n_frames, l->snap_nodes[side].first.ptr,
        l->snap_nodes[side].view());
    sf::heads_chunk<<<tile_grid, 256, 0, l->snap_stream>>>(
        ga, n_frames, l->snap_nodes[side].first.ptr,
        l->snap_nodes[side].view());
  } else {
    sf::gather_chunk<<<tile_grid, 256, 0, l->snap_stream>>>(
        ga, l->snap_table[side].ptr, n_frames, l->snap_nodes[side].first.ptr,
        l->snap_nodes[side].view());
  }
  LAB_CHECK(hipGetLastError(), "snapshot");
  LAB_CHECK(hipEventRecord(l->snap_ready[side], l->snap_stream), "event");
  status = sf::plan_add_reader(plan, l->snap_stream);
  if (status != SF_OK)
    return status;
  {
    std::lock_guard<std::mutex> lock(l->mu);
    l->job.frames = std::move(frames);
    l->job.side = side;
    l->job_ready = true;
    l->jobs_given++;
    if (!l->worker.joinable())
      l->worker = std::thread(worker_loop, l);
  }
  l->cv.notify_all();
  l->started = true;
  l->last_frame = first_frame + (int64_t)n_frames - 1;
  return SF_OK;
}

int sf_hip_labeller_add_lists(sf_labeller *l, int n, const int64_t *frames,
                              const sf_signal_pixel *const *lists,
                              const uint32_t *const *count_ptrs,
                              const uint32_t *capacities,
                              const uint32_t *counts, void *stream) {
  return sf::labeller_add_lists(l, n, frames, lists, count_ptrs, capacities,
                                counts, static_cast<hipStream_t>(stream));
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
  LAB_CHECK(poll_stream(l), "waiting for the labeller");
  l->n_carry = *l->h_scalars;
  return SF_OK;
}

int sf_hip_labeller_wait(sf_labeller *l) {
  if (!l)
    return SF_BAD_ARGUMENT;
  const int drained = drain(l);
  if (drained != SF_OK)
    return drained;
  LAB_CHECK(poll_stream(l), "waiting for the labeller");
  l->n_carry = *l->h_scalars;
  return SF_OK;
}

int sf_hip_labeller_reserve(sf_labeller *l, size_t spots, size_t pixels) {
  if (!l) {
    g_label_error = "sf_hip_labeller_reserve: bad argument";
    return SF_BAD_ARGUMENT;
  }
  int status = flush_pending(l);
  if (status == SF_OK)
    status = sf_hip_labeller_wait(l);
  if (status != SF_OK)
    return status;
  const unsigned long long totals[2] = {l->h_totals[0], l->h_totals[1]};
  LAB_CHECK(l->spots.reserve(totals[0], spots), "reserving spot rows");
  LAB_CHECK(l->pixels.reserve(totals[1], pixels), "reserving pixel rows");
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
    LAB_CHECK(l->spots.read(spots, n_spots), "reading spots");
  if (pixels && n_pixels)
    LAB_CHECK(l->pixels.read(pixels, n_pixels), "reading pixels");
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
