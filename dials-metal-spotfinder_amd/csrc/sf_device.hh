// Device-side types and helpers shared by the spot-finder kernels.
//
// Everything here is gfx950 (CDNA4, wave64) HIP. Compiled with
// -ffp-contract=off: the threshold's float and double tests must round every
// product separately, as the reference's baseline-x86-64 build does
// (src/dext.cc:99-109, 203-207), so no multiply-add may be fused.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/dext_hip.h"

namespace sf {

// ---------------------------------------------------------------------------
// Pixel-type policy (src/dext.hh:18-38): T pixel, W wrapping window
// accumulator, R real type of the two tests. Masked means v >= max(T) - 1
// (src/dext.cc:40).
//
// The 7x7 count n and sum s are carried packed in one W: the count sits above
// bit NS_SHIFT, which no window sum of s reaches (u16: 49 * 65533 < 2^22 <=
// 2^26; u32: 49 * (2^32 - 3) < 2^38 <= 2^40). The 11x11 pair likewise above
// NS2_SHIFT (u16: 121 * 65533 < 2^23; u32: 121 * 2^32 < 2^40). Running sums
// that add and subtract packed values wrap modulo 2^bits and unpack exactly,
// because each field's true window value fits its field.
// ---------------------------------------------------------------------------
template <typename T> struct Traits;

template <> struct Traits<uint16_t> {
  using W = uint32_t;
  using R = float;
  static constexpr int K = 4;               // columns per lane
  static constexpr uint32_t kMasked = 0xFFFEu;
  static constexpr uint32_t kOffFrame = 0xFFFFu; // reads as masked
  static constexpr int kNsShift = 26;
  static constexpr int kNs2Shift = 23;
};

template <> struct Traits<uint32_t> {
  using W = uint64_t;
  using R = double;
  static constexpr int K = 2;
  static constexpr uint32_t kMasked = 0xFFFFFFFEu;
  static constexpr uint32_t kOffFrame = 0xFFFFFFFFu;
  static constexpr int kNsShift = 40;
  static constexpr int kNs2Shift = 40;
};

// Columns a lane-strip reaches beyond its output on each side: 3 (7x7) + 2
// (5x5 erode) + 5 (11x11) = 10, rounded up to whole lane groups so that a
// lane's K columns are either all in the frame or all outside it whenever the
// width is a multiple of K.
template <typename T> constexpr int halo() {
  return ((10 + Traits<T>::K - 1) / Traits<T>::K) * Traits<T>::K;
}
template <typename T> constexpr int strip_width() {
  return 64 * Traits<T>::K - 2 * halo<T>();
}

// One frame of a launch. Device-resident, built by the host plan.
struct FrameDesc {
  const void *pixels;          // T[height * width], row-major
  sf_signal_pixel *staging;    // capacity entries, unordered cell runs
  uint8_t *cell_count;         // height * n_strips survivors per cell
  uint32_t *cell_base;         // their staging offset (where count > 0)
  uint32_t *band_total;        // n_bands: survivors per band of rows
  uint32_t *counter;           // staging allocation counter
  sf_signal_pixel *out;        // ordered list, capacity entries
  uint32_t *out_count;         // found (may exceed capacity: -2)
  uint32_t height, width;
  uint32_t n_strips, n_bands;
  uint32_t band_rows;
  uint32_t capacity;
  uint32_t item_begin;         // first work item of this frame
  uint32_t vec;                // 1 if rows may be read K pixels at a time
  // The sparse u16 form's per-frame tables (dext_sparse.hip): 5x5 cells on
  // a global lattice, n_cell_rows x n_cell_cols of them, 64 cell columns to
  // a word of `act`. Per-cell tables are indexed row * n_cell_cols + column
  // and hold a value only where the cell's `act` bit is set.
  // A scout item (16 cell rows x 64 cell columns, n_strips x n_bands of them)
  // lists its active cells in its own segment of 1024 entries.
  uint64_t *act;               // n_cell_rows * n_strips: active cells
  uint32_t *icount;            // per item: active cells it listed
  uint32_t *alist;             // per item segment: cells, row << 16 | column
  uint32_t *ccand;             // beside alist: each cell's candidates
  uint32_t *ecell;             // per cell: its eroded pixels (25 bits)
  uint32_t *scell;             // per cell: its surviving pixels
  uint32_t *sbase;             // per cell: staging offset of its survivors
  uint32_t *rowsurv;           // per pixel row: survivors, then first place
  uint32_t *dstof;             // per staging entry: its place in `out`
  uint32_t n_cell_rows, n_cell_cols;
};

// ---------------------------------------------------------------------------
// Cross-lane moves. wave_shr:1 gives each lane the value of the lane below it
// (lane - 1) and wave_shl:1 the lane above (lane + 1); bound_ctrl makes the
// lane with no source read zero. Both are plain VALU moves that the compiler
// can fold into the add that consumes them.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t from_below(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x138, 0xF, 0xF, true);
}
__device__ __forceinline__ uint32_t from_above(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x130, 0xF, 0xF, true);
}
__device__ __forceinline__ uint64_t from_below(uint64_t x) {
  const uint32_t lo = from_below((uint32_t)x);
  const uint32_t hi = from_below((uint32_t)(x >> 32));
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t from_above(uint64_t x) {
  const uint32_t lo = from_above((uint32_t)x);
  const uint32_t hi = from_above((uint32_t)(x >> 32));
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Number of set bits of `mask` in lanes below this one.
__device__ __forceinline__ uint32_t count_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) {
  return __ballot(p);
}

// Item i of a launch -> its frame, by binary search over item_begin.
__device__ __forceinline__ int frame_of_item(const FrameDesc *frames,
                                             int n_frames, uint32_t item) {
  int lo = 0, hi = n_frames - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (frames[mid].item_begin <= item)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Blocks are dealt to the eight XCDs round-robin. Remap so that a run of
// consecutive work items -- neighbouring strips, which share halo columns --
// lands on one XCD and so in one L2.
__device__ __forceinline__ uint32_t xcd_local_item(uint32_t block,
                                                   uint32_t blocks) {
  const uint32_t per = blocks / 8; // the launch pads blocks to a multiple of 8
  return (block % 8) * per + block / 8;
}

// The sparse forms' scout items (dext_sparse.hip, dext_sparse32.hip): 16
// cell rows by 64 cell columns, each listing its active cells in its own
// segment of the frame's list.
constexpr int kBandCells = 16;
constexpr int kSegment = 64 * kBandCells;

// Work item `item` of a launch (4 a block, XCD-local runs of blocks): its
// frame, copied into registers (the passes store through pointers the
// compiler cannot tell apart from the descriptor), and its index within the
// frame; false past the end.
__device__ __forceinline__ bool wave_item(const FrameDesc *frames,
                                          int n_frames, uint32_t first,
                                          uint32_t end, uint32_t block,
                                          FrameDesc &fd, uint32_t &local) {
  // the wave's index, as a scalar: everything derived from it is uniform
  const uint32_t item = (uint32_t)__builtin_amdgcn_readfirstlane(
      (int)(first + 4 * block + (threadIdx.x >> 6)));
  if (item >= end)
    return false;
  fd = frames[frame_of_item(frames, n_frames, item)];
  local = item - fd.item_begin;
  return true;
}

// One chunk's part of a mixed launch of the sparse forms (a scout of one
// chunk beside the erosion or Poisson pass of another).
struct Side {
  const FrameDesc *frames; // descriptors of the side's chunk
  int n_frames;
  uint32_t first, end;     // its work items in this launch
  uint32_t blocks;         // blocks of four items
};

// Block `block` of a mixed launch: 1 for a scout block, 0 for a sparse
// block, with its index among its kind; -1 for the grid's padding. Blocks
// are dealt to the eight XCDs round-robin, and each XCD draws HBM through
// its own share of the fabric: so each XCD takes
// an eighth of each kind (contiguous runs: neighbouring strips share an
// L2), the two kinds spread through its blocks in proportion. The grid is
// mixed_grid() blocks.
__device__ __forceinline__ int mixed_role(uint32_t block,
                                          uint32_t scout_blocks,
                                          uint32_t sparse_blocks,
                                          uint32_t &index) {
  const uint32_t x = block % 8, k = block / 8;
  const uint32_t s0 = scout_blocks * x / 8, s1 = scout_blocks * (x + 1) / 8;
  const uint32_t e0 = sparse_blocks * x / 8, e1 = sparse_blocks * (x + 1) / 8;
  const uint32_t ns = s1 - s0, n = ns + (e1 - e0);
  index = 0;
  if (k >= n)
    return -1;
  // scout blocks among this XCD's first k and first k + 1
  const uint32_t before = (uint32_t)((uint64_t)k * ns / n);
  const bool scout = (uint32_t)((uint64_t)(k + 1) * ns / n) > before;
  index = scout ? s0 + before : e0 + (k - before);
  return scout ? 1 : 0;
}

} // namespace sf
