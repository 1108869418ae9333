// The u32 threshold as sparse passes over the u16 form's 5x5 cell lattice
// (dext_sparse.hip, which also argues why one probe per cell suffices): a
// dense scout that tests each cell's probe pixel, erosion and the Poisson
// test only at the cells it marks, then the u16 form's ordering. The
// tables, work items and ordering are shared with the u16 form; the
// arithmetic is the reference's for 32-bit pixels (src/dext.hh:18-38):
// sums in wrapping 64-bit integers, both tests in double, a pixel masked at
// 0xFFFFFFFE and above (src/dext.cc:40).
//
//   scout   one wave per 64 cell columns x 16 cell rows: each lane reads
//           eight columns of every pixel row around its probes (two
//           16-byte loads), sums its probe's 7x7 window (the two rows that
//           neighbouring windows share carried), tests it, and the wave
//           lists its active cells.
//   erode   sixteen lanes per active cell: mask0 on the 9x9 pixels around
//           the cell from running column sums and 7-wide lane windows, then
//           E by bit operations on ballots; cells with candidates listed.
//   poisson sixteen lanes per candidate cell: 11x11 sums of the pixels
//           neither masked nor eroded, the test at the candidates,
//           survivors to staging.
#include "sf_device.hh"

namespace sf {

hipError_t launch_sparse_zero(const FrameDesc *frames, int n_frames,
                              hipStream_t stream);
hipError_t launch_sparse_order(const FrameDesc *frames, int n_frames,
                               uint32_t max_cell_rows, hipStream_t stream);
uint32_t mixed_grid(uint32_t scout_blocks, uint32_t sparse_blocks);

namespace {

constexpr uint32_t kMasked = 0xFFFFFFFEu; // masked at and above

// sqrt(2 (n - 1)) for n = 2..49 as the reference takes it in double
// (src/dext.cc:108); +inf for n = 0, 1, so that lhs > rhs is false there
// (0 * inf is NaN), as the reference tests only windows of two or more.
__device__ __forceinline__ void fill_lut(double *lut) {
  if (threadIdx.x < 64)
    lut[threadIdx.x] = threadIdx.x < 2
                           ? __builtin_inf()
                           : sqrt(2.0 * (double)(threadIdx.x - 1));
}

// The dispersion test (src/dext.cc:97-109) in double, in the reference's
// order, from the window's count n and its wrapped sums s, q.
__device__ __forceinline__ bool dispersed(uint64_t s, uint64_t q, uint32_t n,
                                          const double *lut) {
  const double rm = (double)n, rs = (double)s, rq = (double)q;
  const double lhs = rm * rq - rs * rs - rs * (rm - 1.0);
  const double rhs = rs * 6.0 * lut[n];
  return lhs > rhs;
}

// The frame as a buffer resource: reads past its end return 0, never fault
// (the plan keeps u32 frames of this form below 2^29 pixels).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t frame_buffer(
    const FrameDesc &fd) {
  return __builtin_amdgcn_make_buffer_rsrc(
      const_cast<void *>(fd.pixels), 0, (int)(4u * fd.height * fd.width),
      0x00020000);
}

// One pixel row of a lane's probe window: eight columns; bit k of `win`
// (of `cell`) says column k is in the probe's window (the cell).
__device__ __forceinline__ void scout_row(const uint32_t (&d)[8],
                                          uint32_t win, uint32_t cell,
                                          uint64_t &s, uint64_t &q,
                                          uint32_t &n7, uint32_t &n5) {
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const bool ok = d[k] < kMasked;
    const bool in7 = ok && ((win >> k) & 1u);
    const uint32_t p = in7 ? d[k] : 0u;
    s += p;
    q += (uint64_t)p * p;
    n7 += in7 ? 1u : 0u;
    n5 += ok && ((cell >> k) & 1u) ? 1u : 0u;
  }
}

// Scout one item: cell columns 64 * strip + lane, cell rows
// [16 * band, 16 * band + 16). VEC: the frame is at least 8 pixels wide
// and each lane reads its 8 columns as two 16-byte buffer loads. INTERIOR
// (with VEC): every lane's window lies inside the frame's columns, so the
// eight columns start one left of the cell: window columns 0..6, cell
// columns 1..5, probe column 3, all known here.
template <bool VEC, bool INTERIOR>
__device__ __forceinline__ void scout_item(const FrameDesc &fd, uint32_t strip,
                                           uint32_t band, const double *lut,
                                           uint64_t *masks) {
  const int height = (int)fd.height, width = (int)fd.width;
  const int ncr = (int)fd.n_cell_rows, ncc = (int)fd.n_cell_cols;
  const int lane = (int)lane_id();
  const int j = 64 * (int)strip + lane;
  const bool live = j < ncc;
  const int pc = min(5 * j + 2, width - 1);
  const int wlo = max(pc - 3, 0), whi = min(pc + 3, width - 1);
  const int clo = 5 * j, chi = min(5 * j + 4, width - 1);
  const int ls = INTERIOR ? pc - 3
                 : VEC    ? min(max(pc - 3, 0), width - 8)
                          : pc - 3;
  uint32_t win = 0, cell = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int c = ls + k;
    if (live && c >= wlo && c <= whi)
      win |= 1u << k;
    if (live && c >= clo && c <= chi)
      cell |= 1u << k;
  }
  if (INTERIOR) {
    win = 0x7Fu;
    cell = 0x3Eu;
  }
  const int probe_at = INTERIOR ? 3 : pc - ls; // 0..7
  const __amdgpu_buffer_rsrc_t rsrc = frame_buffer(fd);
  const uint32_t lane_at = 4u * (uint32_t)ls;

  // a pixel row's eight columns; rows off the frame read as masked (CHECK:
  // the item reaches the frame's top or bottom)
  const auto load_row = [&]<bool CHECK>(int y, uint32_t (&d)[8]) {
    if (CHECK && (y < 0 || y >= height)) { // uniform
#pragma unroll
      for (int k = 0; k < 8; k++)
        d[k] = 0xFFFFFFFFu;
      return;
    }
    if (VEC) {
      typedef int i32x4 __attribute__((ext_vector_type(4)));
      const int row = (int)(4u * (uint32_t)y * (uint32_t)width);
      const i32x4 a =
          __builtin_amdgcn_raw_buffer_load_b128(rsrc, (int)lane_at, row, 0);
      const i32x4 b = __builtin_amdgcn_raw_buffer_load_b128(
          rsrc, (int)lane_at + 16, row, 0);
      d[0] = (uint32_t)a.x;
      d[1] = (uint32_t)a.y;
      d[2] = (uint32_t)a.z;
      d[3] = (uint32_t)a.w;
      d[4] = (uint32_t)b.x;
      d[5] = (uint32_t)b.y;
      d[6] = (uint32_t)b.z;
      d[7] = (uint32_t)b.w;
    } else { // one pixel a load; columns off the frame read as masked
      const int row = (int)(4u * (uint32_t)y * (uint32_t)width);
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const int c = ls + k;
        const uint32_t v = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(
            rsrc, 4 * min(max(c, 0), width - 1), row, 0);
        d[k] = c >= 0 && c < width ? v : 0xFFFFFFFFu;
      }
    }
  };

  // Whether full cell row i's cell (lane's column j), whose probe lies in a
  // gap, holds no eroded pixel (src/dext.cc:117-142): E(x) = 1 needs mask0
  // on every in-frame pixel of x's 5x5 window, so one unmasked pixel z with
  // mask0(z) = 0 clears E within two rows and columns of it. The points
  // tried lie on the probe's column pc (every column of the cell is within
  // two of it): row 5i+1 (else 5i) and row 5i+3 (else 5i+4), where
  // unmasked; their 7x7 windows have the probe window's columns, so the
  // lane's eight loaded columns serve. The cell is free when each of its
  // rows that holds an unmasked pixel is within two rows of a point whose
  // test is negative. Rows 5i-3 .. 5i+7 are read again, once.
  const auto eroded_free = [&](int i) -> bool {
    uint64_t ws[4] = {0, 0, 0, 0}, wq[4] = {0, 0, 0, 0};
    uint32_t wn[4] = {0, 0, 0, 0};
    uint32_t um = 0, zc = 0; // per cell row: unmasked pixel / at column pc
#pragma unroll 1
    for (int o = -3; o <= 7; o++) {
      uint32_t d[8];
      load_row.template operator()<true>(5 * i + o, d);
      uint64_t s_ = 0, q_ = 0;
      uint32_t n_ = 0, c_ = 0;
      scout_row(d, win, cell, s_, q_, n_, c_);
      // windows of the points at cell rows 0, 1, 3, 4 (slots 0..3)
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const int r = t < 2 ? t : t + 1;
        if (o >= r - 3 && o <= r + 3) {
          ws[t] += s_;
          wq[t] += q_;
          wn[t] += n_;
        }
      }
      if (o >= 0 && o <= 4) {
        uint32_t w = d[0];
#pragma unroll
        for (int k = 1; k < 8; k++)
          w = probe_at == k ? d[k] : w;
        um |= (c_ != 0 ? 1u : 0u) << o;
        zc |= (w < kMasked ? 1u : 0u) << o;
      }
    }
    uint32_t covered = 0;
    // top point: row 1, else row 0; bottom point: row 3, else row 4
    const int top = (zc >> 1) & 1u ? 1 : (zc & 1u) ? 0 : -1;
    const int bottom = (zc >> 3) & 1u ? 2 : (zc >> 4) & 1u ? 3 : -1;
    const uint32_t rows_near[4] = {0x07u, 0x0Fu, 0x1Eu, 0x1Cu};
#pragma unroll
    for (int t = 0; t < 4; t++)
      if ((t == top || t == bottom) && !dispersed(ws[t], wq[t], wn[t], lut))
        covered |= rows_near[t];
    return (um & ~covered) == 0;
  };

  const int i0 = (int)band * kBandCells, i1 = min(i0 + kBandCells, ncr);
  // rows 5 i0 - 1 and 5 i0: the first window's top, carried like a tail
  uint64_t ts = 0, tq = 0;
  uint32_t tn = 0, top5 = 0, unused = 0;
  {
    uint32_t d[8];
    load_row.template operator()<true>(5 * i0 - 1, d);
    scout_row(d, win, cell, ts, tq, tn, unused);
    load_row.template operator()<true>(5 * i0, d);
    scout_row(d, win, cell, ts, tq, tn, top5);
  }
  const auto load_rows = [&]<bool CHECK>(int y, uint32_t (&rows)[5][8]) {
#pragma unroll
    for (int r = 0; r < 5; r++)
      load_row.template operator()<CHECK>(y + r, rows[r]);
  };
  // cell row i from its rows 5i+1 .. 5i+5 and the carried pair 5i-1, 5i
  const auto cell_row = [&](int i, const uint32_t (&rows)[5][8]) {
    uint64_t s = ts, q = tq, s2 = 0, q2 = 0;
    uint32_t n7 = tn, n5 = top5, n2 = 0, n5n = 0;
    scout_row(rows[0], win, cell, s, q, n7, n5);
    scout_row(rows[1], win, cell, s, q, n7, n5);
    scout_row(rows[2], win, cell, s, q, n7, n5);
    scout_row(rows[3], win, cell, s2, q2, n2, n5);
    scout_row(rows[4], win, cell, s2, q2, n2, n5n);
    s += s2;
    q += q2;
    n7 += n2;
    ts = s2;
    tq = q2;
    tn = n2;
    top5 = n5n;
    // the probe's own value, from row 5i+2
    uint32_t probe_row[8];
#pragma unroll
    for (int k = 0; k < 8; k++)
      probe_row[k] = rows[1][k];
    if (5 * i + 2 >= height) { // uniform: a last cell row of one or two rows
      // probe on the last row instead, window rows height-4 .. height-1
      s = q = 0;
      n7 = 0;
      uint32_t d[8];
#pragma unroll
      for (int r = 4; r >= 1; r--) {
        load_row.template operator()<true>(height - r, d);
        scout_row(d, win, cell, s, q, n7, unused);
      }
#pragma unroll
      for (int k = 0; k < 8; k++)
        probe_row[k] = d[k];
    }
    uint32_t pv = probe_row[0];
#pragma unroll
    for (int k = 1; k < 8; k++)
      pv = probe_at == k ? probe_row[k] : pv;
    const bool active =
        live && (pv >= kMasked || dispersed(s, q, n7, lut)) && n5 > 0;
    // A full cell whose probe lies in a gap (0xFFFFFFFF: a module gap that
    // cuts the cell) usually holds no eroded pixel: such cells are checked
    // after the sweep, when its registers are free. (Cells whose probe is a
    // hot pixel, 0xFFFFFFFE, are scattered one to a wave and stay.)
    const bool full = 5 * i + 4 < height && 5 * j + 4 < width;
    const uint64_t bits = ballot(active);
    const uint64_t check = ballot(active && pv == 0xFFFFFFFFu && full);
    if (lane == 0) {
      masks[i - i0] = bits;
      masks[kBandCells + i - i0] = check;
    }
  };
  if (5 * i1 + 5 <= height) { // rows 5 i0 + 1 .. 5 i1 + 4 all in the frame
    for (int i = i0; i < i1; i++) {
      uint32_t rows[5][8];
      load_rows.template operator()<false>(5 * i + 1, rows);
      cell_row(i, rows);
    }
  } else {
    for (int i = i0; i < i1; i++) {
      uint32_t rows[5][8];
      load_rows.template operator()<true>(5 * i + 1, rows);
      cell_row(i, rows);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // drop the checked cells that hold no eroded pixel
  for (int i = i0; i < i1; i++) {
    const uint64_t check = masks[kBandCells + i - i0];
    if (check) { // uniform, and rare
      const bool mine = (check >> lane) & 1u;
      const uint64_t bits = masks[i - i0] & ~ballot(eroded_free(i) && mine);
      if (lane == 0)
        masks[i - i0] = bits;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
  // publish the bits, and the item's active cells in its own segment
  uint32_t total = 0;
  const uint32_t local = band * fd.n_strips + strip;
  uint32_t *list = fd.alist + (size_t)local * kSegment;
  for (int i = i0; i < i1; i++) {
    const uint64_t bits = masks[i - i0];
    if (lane == 0)
      fd.act[(size_t)i * fd.n_strips + strip] = bits;
    if ((bits >> lane) & 1u)
      list[total + count_below(bits)] = ((uint32_t)i << 16) | (uint32_t)j;
    total += (uint32_t)__builtin_popcountll(bits);
  }
  if (lane == 0)
    fd.icount[local] = total;
}

__device__ __forceinline__ bool block_item(const FrameDesc *frames,
                                           int n_frames, uint32_t items,
                                           FrameDesc &fd, uint32_t &local) {
  return wave_item(frames, n_frames, 0, items,
                   xcd_local_item(blockIdx.x, gridDim.x), fd, local);
}

// A scout work item on this wave: its frame's lattice, strip and band.
__device__ __forceinline__ void scout_wave(const FrameDesc &fd, uint32_t local,
                                           const double *lut,
                                           uint64_t *masks) {
  const uint32_t strip = local % fd.n_strips, band = local / fd.n_strips;
  // interior: every lane's cell column j has columns 5j-1 .. 5j+6 in frame
  const int j = 64 * (int)strip + (int)lane_id();
  const bool inside = j >= 1 && j < (int)fd.n_cell_cols &&
                      5 * j + 6 <= (int)fd.width - 1;
  if (fd.vec && fd.width >= 8 && !ballot(!inside))
    scout_item<true, true>(fd, strip, band, lut, masks);
  else if (fd.vec && fd.width >= 8)
    scout_item<true, false>(fd, strip, band, lut, masks);
  else
    scout_item<false, false>(fd, strip, band, lut, masks);
}

__global__ __launch_bounds__(256, 4) void scout32_kernel(
    const FrameDesc *frames, int n_frames, uint32_t items) {
  __shared__ double lut[64];
  // per wave: each cell row's active cells, then the cells to check
  __shared__ uint64_t masks[4][2 * kBandCells];
  fill_lut(lut);
  __syncthreads();
  FrameDesc fd;
  uint32_t local;
  if (block_item(frames, n_frames, items, fd, local))
    scout_wave(fd, local, lut, masks[threadIdx.x >> 6]);
}

// ---------------------------------------------------------------------------
// Sixteen lanes to a cell, four cells to a wave, as in the u16 form: lane c
// of a cell's group holds pixel column 5j - 5 + c (c = 15 idles: no result
// reads it) and the cell's rows 5i - 5 .. 5i + 9 in registers. Column sums
// run down the rows; window sums move within the group's DPP row (16
// lanes): right<N> reads lane c + N, and lanes past the row's end read 0.
// ---------------------------------------------------------------------------
template <int N> __device__ __forceinline__ uint32_t right(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x100 + N, 0xF, 0xF, true);
}
template <int N> __device__ __forceinline__ uint64_t right64(uint64_t x) {
  const uint32_t lo = right<N>((uint32_t)x), hi = right<N>((uint32_t)(x >> 32));
  return ((uint64_t)hi << 32) | lo;
}
// x over lanes c .. c+6: the 7-wide window centred on lane c+3 (valid at
// c = 0..8); additions only.
template <typename T, T (*R1)(T), T (*R2)(T), T (*R4)(T), T (*R6)(T)>
__device__ __forceinline__ T window7_of(T x) {
  const T a = x + R1(x); // c .. c+1
  const T b = a + R2(a); // c .. c+3
  return (b + R4(a)) + R6(x);
}
__device__ __forceinline__ uint32_t window7(uint32_t x) {
  return window7_of<uint32_t, right<1>, right<2>, right<4>, right<6>>(x);
}
__device__ __forceinline__ uint64_t window7(uint64_t x) {
  return window7_of<uint64_t, right64<1>, right64<2>, right64<4>,
                    right64<6>>(x);
}
// x over lanes c .. c+10: the 11-wide window centred on lane c+5 (valid at
// c = 0..4).
template <typename T, T (*R1)(T), T (*R2)(T), T (*R4)(T), T (*R8)(T),
          T (*R10)(T)>
__device__ __forceinline__ T window11_of(T x) {
  const T a = x + R1(x); // c .. c+1
  const T b = a + R2(a); // c .. c+3
  const T d = b + R4(b); // c .. c+7
  return (d + R8(a)) + R10(x);
}
__device__ __forceinline__ uint32_t window11(uint32_t x) {
  return window11_of<uint32_t, right<1>, right<2>, right<4>, right<8>,
                     right<10>>(x);
}
__device__ __forceinline__ uint64_t window11(uint64_t x) {
  return window11_of<uint64_t, right64<1>, right64<2>, right64<4>,
                     right64<8>, right64<10>>(x);
}

constexpr uint64_t kCellLanes = 0x03E003E003E003E0ull; // lanes 5..9 a group

// The 15 rows of a lane's column around `cell`; off the frame, or on an
// invalid group, pixels read as masked.
__device__ __forceinline__ void load_column(const FrameDesc &fd,
                                            __amdgpu_buffer_rsrc_t rsrc,
                                            bool valid, uint32_t cell, int c,
                                            uint32_t (&v)[15]) {
  const int height = (int)fd.height, width = (int)fd.width;
  const int i = (int)(cell >> 16), j = (int)(cell & 0xFFFFu);
  const int x = 5 * j - 5 + c, y0 = 5 * i - 5;
  // byte offset of row y0 (wraps below the frame: then out of range)
  const uint32_t at = 4u * (uint32_t)(y0 * width + x);
  const uint32_t pitch = 4u * (uint32_t)width;
#pragma unroll
  for (int r = 0; r < 15; r++)
    v[r] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(
        rsrc, (int)(at + (uint32_t)r * pitch), 0, 0);
  // the frame's edges and invalid groups (uniform test): mask what is off
  const bool edge = !valid || (c < 15 && (x < 0 || x >= width)) || y0 < 0 ||
                    y0 + 15 > height;
  if (ballot(edge)) {
    const bool col = valid && x >= 0 && x < width;
#pragma unroll
    for (int r = 0; r < 15; r++)
      if (!col || (uint32_t)(y0 + r) >= (uint32_t)height)
        v[r] = 0xFFFFFFFFu;
  }
}

// mask0 on the 9x9 pixels around a cell, then E on the cell and its
// candidates (E, unmasked, 0 < p): src/dext.cc:65-142. Writes the cell's
// E and clears its survivors; returns its candidates.
__device__ uint32_t erode_group(const FrameDesc &fd, bool valid, uint32_t cell,
                                const uint32_t (&v)[15], const double *lut) {
  const uint32_t lane = lane_id();
  // the lane's column sums over 7 rows: count, value, value^2 (wrapping)
  uint32_t cn = 0;
  uint64_t cs = 0, cq = 0;
  const auto take = [&](uint32_t x, bool in) {
    const bool ok = x < kMasked;
    const uint32_t p = ok ? x : 0u;
    const uint64_t p2 = (uint64_t)p * p;
    if (in) {
      cn += ok ? 1u : 0u;
      cs += p;
      cq += p2;
    } else {
      cn -= ok ? 1u : 0u;
      cs -= p;
      cq -= p2;
    }
  };
#pragma unroll
  for (int r = 0; r < 6; r++)
    take(v[r], true);
  // mask0 or masked ("set") at pixel rows 5i-2+o, columns 3..11: 7x7
  // windows (rows o .. o+6), tested on lane c for column c+3
  uint64_t h5[9];
#pragma unroll
  for (int o = 0; o < 9; o++) {
    take(v[o + 6], true);
    if (o > 0)
      take(v[o - 1], false);
    const uint32_t n = window7(cn);
    const uint64_t s = window7(cs), q = window7(cq);
    const uint64_t b = (ballot(dispersed(s, q, n, lut)) << 3) |
                       ballot(v[o + 3] >= kMasked);
    // set on all five columns c-2 .. c+2 (bit c: column c of its group)
    h5[o] = b & (b << 1) & (b << 2) & (b >> 1) & (b >> 2);
  }
  const uint32_t at = 16u * (lane >> 4) + 5u;
  uint32_t e25 = 0, c25 = 0;
#pragma unroll
  for (int t = 0; t < 5; t++) {
    const uint32_t p = v[t + 5];
    const uint64_t unmasked = ballot(p < kMasked);
    const uint64_t positive = ballot(p - 1u < kMasked - 1u); // 0 < p < kMasked
    const uint64_t e = h5[t] & h5[t + 1] & h5[t + 2] & h5[t + 3] & h5[t + 4] &
                       unmasked & kCellLanes;
    e25 |= (uint32_t)((e >> at) & 31u) << (5 * t);
    c25 |= (uint32_t)(((e & positive) >> at) & 31u) << (5 * t);
  }
  if (valid && (lane & 15u) == 0) {
    const uint32_t ci = (cell >> 16) * fd.n_cell_cols + (cell & 0xFFFFu);
    fd.ecell[ci] = e25;
    fd.scell[ci] = 0;
  }
  return c25;
}

// One item's active cells, 64 at a time, four to a wave-step; the next four
// cells' pixels are fetched while the current four are tested.
__device__ __forceinline__ void erode_item(const FrameDesc &fd, uint32_t local,
                                           const double *lut) {
  const uint32_t n = fd.icount[local];
  const uint32_t lane = lane_id(), grp = lane >> 4;
  const int c = (int)(lane & 15u);
  const uint32_t *cells = fd.alist + (size_t)local * kSegment;
  uint32_t *cand = fd.ccand + (size_t)local * kSegment;
  const __amdgpu_buffer_rsrc_t rsrc = frame_buffer(fd);
  for (uint32_t c0 = 0; c0 < n; c0 += 64) {
    const uint32_t m = min(64u, n - c0);
    const uint32_t mine = lane < m ? cells[c0 + lane] : 0u;
    const uint32_t steps = (m + 3) / 4;
    bool valid = grp < m;
    uint32_t cell = (uint32_t)__shfl((int)mine, (int)grp, 64);
    uint32_t v[15];
    load_column(fd, rsrc, valid, cell, c, v);
    for (uint32_t g = 0; g < steps; g++) {
      const uint32_t next = 4 * (g + 1) + grp;
      const bool next_valid = next < m;
      const uint32_t next_cell =
          (uint32_t)__shfl((int)mine, (int)(next & 63u), 64);
      uint32_t vn[15];
      if (g + 1 < steps)
        load_column(fd, rsrc, next_valid, next_cell, c, vn);
      const uint32_t c25 = erode_group(fd, valid, cell, v, lut);
      if (valid && c == 0)
        cand[c0 + 4 * g + grp] = c25;
#pragma unroll
      for (int r = 0; r < 15; r++)
        v[r] = vn[r];
      cell = next_cell;
      valid = next_valid;
    }
  }
}

__global__ __launch_bounds__(256) void erode32_kernel(const FrameDesc *frames,
                                                      int n_frames,
                                                      uint32_t items) {
  __shared__ double lut[64];
  fill_lut(lut);
  __syncthreads();
  FrameDesc fd;
  uint32_t local;
  if (block_item(frames, n_frames, items, fd, local))
    erode_item(fd, local, lut);
}

// What a lane needs for one cell's Poisson sums: its column's 15 pixels and
// the E words of the cells above, at and below in its cell column.
struct PoissonIn {
  uint32_t v[15];
  uint64_t word[3];
  uint32_t eroded[3];
};

__device__ __forceinline__ void poisson_load(const FrameDesc &fd,
                                             __amdgpu_buffer_rsrc_t rsrc,
                                             bool valid, uint32_t cell,
                                             PoissonIn &in) {
  const int c = (int)(lane_id() & 15u);
  const int i = (int)(cell >> 16), j = (int)(cell & 0xFFFFu);
  const int x = 5 * j - 5 + c, ncr = (int)fd.n_cell_rows;
  const bool col = valid && c < 15 && x >= 0 && x < (int)fd.width;
  const uint32_t jj = col ? (uint32_t)x / 5u : 0u;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const uint32_t ii = (uint32_t)min(max(i - 1 + a, 0), ncr - 1);
    in.word[a] = fd.act[ii * fd.n_strips + (jj >> 6)];
    in.eroded[a] = fd.ecell[ii * fd.n_cell_cols + jj];
  }
  load_column(fd, rsrc, valid, cell, c, in.v);
}

struct Compact;
__device__ __forceinline__ void put_candidate(Compact &cp, uint32_t mine,
                                              uint32_t first, int t,
                                              uint32_t n, uint64_t sum,
                                              uint32_t p);

// The 11x11 sums of a cell's pixels neither masked nor eroded
// (src/dext.cc:146-189), for the lane's column c of the cell (lanes c = 0..4
// of the group), row by row of the cell: each candidate's count, sum and
// value go to the step's compacted list as they are formed. `mine`: the
// cell's candidates; `first`: the list entry of its first.
__device__ __forceinline__ void poisson_sums(const FrameDesc &fd, bool valid,
                                             uint32_t cell, const PoissonIn &in,
                                             uint32_t mine, uint32_t first,
                                             Compact &cp) {
  const int c = (int)(lane_id() & 15u);
  const int i = (int)(cell >> 16), j = (int)(cell & 0xFFFFu);
  const int x = 5 * j - 5 + c, ncr = (int)fd.n_cell_rows;
  const bool col = valid && c < 15 && x >= 0 && x < (int)fd.width;
  const uint32_t jj = col ? (uint32_t)x / 5u : 0u;
  const uint32_t cc = col ? (uint32_t)x - 5u * jj : 0u;
  // E where the cell is active (E holds only in active cells) and in frame
  uint32_t e[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const int ii = i - 1 + a;
    e[a] = (col && ii >= 0 && ii < ncr && ((in.word[a] >> (jj & 63u)) & 1u))
               ? in.eroded[a]
               : 0u;
  }
  // the rows' pixels that count: neither masked nor eroded
  uint32_t counted = 0;
#pragma unroll
  for (int r = 0; r < 15; r++) {
    const uint32_t out = (e[r / 5] >> (5u * (r % 5) + cc)) & 1u;
    counted |= (in.v[r] < kMasked && !out ? 1u : 0u) << r;
  }
  // the lane's column sums over 11 rows: count and value
  uint32_t cn = 0;
  uint64_t cs = 0;
#pragma unroll
  for (int r = 0; r < 10; r++) {
    const bool ok = (counted >> r) & 1u;
    cn += ok ? 1u : 0u;
    cs += ok ? in.v[r] : 0u;
  }
#pragma unroll
  for (int t = 0; t < 5; t++) {
    {
      const bool ok = (counted >> (t + 10)) & 1u;
      cn += ok ? 1u : 0u;
      cs += ok ? in.v[t + 10] : 0u;
    }
    if (t > 0) {
      const bool ok = (counted >> (t - 1)) & 1u;
      cn -= ok ? 1u : 0u;
      cs -= ok ? in.v[t - 1] : 0u;
    }
    const uint32_t n = window11(cn);
    const uint64_t sum = window11(cs);
    const uint32_t p = right<5>(in.v[t + 5]);
    put_candidate(cp, mine, first, t, n, sum, p);
  }
}

// The Poisson test (src/dext.cc:201-206) in double, in the reference's order:
// the 11x11 window's count n and sum, the candidate's value p; the mean goes
// out through `mean`.
__device__ __forceinline__ bool poisson_signal(uint32_t n, uint64_t sum,
                                               uint32_t p, double &mean) {
  const double m = n >= 2 ? (double)sum / (double)n : 0.0;
  mean = m;
  return (double)p >= m + 3.0 * sqrt(m);
}

// Orders one wave's LDS stores before its later LDS loads (lanes read what
// other lanes wrote).
__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A wave's survivors wait here until a whole run of them takes its place in
// staging under one atomic: each cell's run in its pixel order.
constexpr int kStash = 384;
constexpr int kStashCells = 64;
struct Stash {
  float mean[kStash];
  uint32_t count[kStash];
  uint32_t value[kStash];
  uint32_t index[kStash]; // pixel index in the frame
  uint32_t cell[kStashCells];
  uint32_t s25[kStashCells];
  uint32_t at[kStashCells]; // the cell's run in the stash
};

// A wave-step's candidates, compacted in (cell, row, column) order: at most
// four cells of 25 pixels.
constexpr int kCompact = 100;
struct Compact {
  uint64_t sum[kCompact];  // window sum
  uint32_t value[kCompact];
  uint32_t info[kCompact]; // count | (5 row + column) << 8 | group << 16
  uint32_t s25[4];         // each group's survivors
};

// Lane c (0..4) of a group: its pixel on the cell's row t, if a candidate,
// to its place in the step's compacted list.
__device__ __forceinline__ void put_candidate(Compact &cp, uint32_t mine,
                                              uint32_t first, int t,
                                              uint32_t n, uint64_t sum,
                                              uint32_t p) {
  const uint32_t lane = lane_id(), c = lane & 15u;
  const uint32_t bit = 5u * (uint32_t)t + c;
  if (c <= 4 && ((mine >> bit) & 1u)) {
    const uint32_t at =
        first + (uint32_t)__builtin_popcount(mine & ((1u << bit) - 1u));
    cp.sum[at] = sum;
    cp.value[at] = p;
    cp.info[at] = n | (bit << 8) | ((lane >> 4) << 16);
  }
}

__device__ __forceinline__ void flush(const FrameDesc &fd, Stash &st,
                                      uint32_t &n, uint32_t &cells) {
  const uint32_t lane = lane_id();
  if (n) {
    uint32_t base = 0;
    if (lane == 0)
      base = atomicAdd(fd.counter, n);
    base = (uint32_t)__builtin_amdgcn_readfirstlane((int)base);
    for (uint32_t k = lane; k < cells; k += 64) {
      const uint32_t cell = st.cell[k], s25 = st.s25[k];
      const uint32_t i = cell >> 16;
      const uint32_t ci = i * fd.n_cell_cols + (cell & 0xFFFFu);
      fd.scell[ci] = s25;
      fd.sbase[ci] = base + st.at[k];
#pragma unroll
      for (int t = 0; t < 5; t++) {
        const uint32_t row = (s25 >> (5 * t)) & 31u;
        if (row)
          atomicAdd(&fd.rowsurv[5 * i + t], (uint32_t)__builtin_popcount(row));
      }
    }
    for (uint32_t e = lane; e < n; e += 64) {
      const uint32_t dst = base + e;
      if (dst < fd.capacity) {
        sf_signal_pixel px;
        px.index = st.index[e];
        px.value = st.value[e];
        px.background = st.mean[e];
        px.population = (uint16_t)st.count[e];
        px.reserved = 0;
        fd.staging[dst] = px;
      }
    }
  }
  n = 0;
  cells = 0;
}

// One item's candidate cells: each 64 of its active cells are compacted to
// their candidate cells (in LDS) and summed four at a time, the next four
// cells' inputs fetched while the current four are summed. The step's
// candidate pixels are then compacted in (cell, row, column) order, a lane
// each, and tested there, so that the double division and square root run
// once a candidate; survivors gather in the stash in that order, which is
// each cell's pixel order.
__device__ __forceinline__ void poisson_item(const FrameDesc &fd,
                                             uint32_t local, uint64_t *slots,
                                             Stash &st, Compact &cp) {
  const uint32_t n = fd.icount[local];
  const uint32_t lane = lane_id(), grp = lane >> 4;
  const uint32_t *cells = fd.alist + (size_t)local * kSegment;
  const uint32_t *cand = fd.ccand + (size_t)local * kSegment;
  const __amdgpu_buffer_rsrc_t rsrc = frame_buffer(fd);
  const uint32_t width = fd.width;
  uint32_t stashed = 0, stashed_cells = 0;
  for (uint32_t c0 = 0; c0 < n; c0 += 64) {
    const bool here = c0 + lane < n;
    const uint32_t cell = here ? cells[c0 + lane] : 0u;
    const uint32_t c25 = here ? cand[c0 + lane] : 0u;
    const uint64_t has = ballot(c25 != 0);
    if (c25)
      slots[count_below(has)] = (uint64_t)cell | ((uint64_t)c25 << 32);
    wave_sync_lds();
    const uint32_t nc = (uint32_t)__builtin_popcountll(has);
    const uint32_t steps = (nc + 3) / 4;
    if (!steps)
      continue;
    bool valid = grp < nc;
    uint64_t entry = valid ? slots[grp] : 0ull;
    PoissonIn in;
    poisson_load(fd, rsrc, valid, (uint32_t)entry, in);
    for (uint32_t g = 0; g < steps; g++) {
      const uint32_t next = 4 * (g + 1) + grp;
      const bool next_valid = next < nc;
      const uint64_t next_entry = next_valid ? slots[next] : 0ull;
      PoissonIn ahead;
      if (g + 1 < steps)
        poisson_load(fd, rsrc, next_valid, (uint32_t)next_entry, ahead);
      // the step's candidates: counts of the four cells, in group order
      const uint32_t mine = (uint32_t)(entry >> 32); // 0 on invalid groups
      const uint32_t k = (uint32_t)__builtin_popcount(mine);
      const uint32_t k0 = (uint32_t)__builtin_amdgcn_readlane((int)k, 0);
      const uint32_t k1 = (uint32_t)__builtin_amdgcn_readlane((int)k, 16);
      const uint32_t k2 = (uint32_t)__builtin_amdgcn_readlane((int)k, 32);
      const uint32_t k3 = (uint32_t)__builtin_amdgcn_readlane((int)k, 48);
      const uint32_t total = k0 + k1 + k2 + k3;
      if (stashed + total > (uint32_t)kStash ||
          stashed_cells + 4 > (uint32_t)kStashCells)
        flush(fd, st, stashed, stashed_cells);
      const uint32_t first =
          (grp > 0 ? k0 : 0u) + (grp > 1 ? k1 : 0u) + (grp > 2 ? k2 : 0u);
      if (lane < 4)
        cp.s25[lane] = 0;
      poisson_sums(fd, valid, (uint32_t)entry, in, mine, first, cp);
      wave_sync_lds();
      // a lane a candidate; survivors to the stash in candidate order
      uint32_t kept = 0;
      for (uint32_t b = 0; b < total; b += 64) {
        const uint32_t at = b + lane;
        bool ok = false;
        double mean = 0.0;
        uint32_t info = 0, value = 0;
        if (at < total) {
          info = cp.info[at];
          value = cp.value[at];
          ok = poisson_signal(info & 0xFFu, cp.sum[at], value, mean);
        }
        const uint64_t okb = ballot(ok);
        if (ok) {
          const uint32_t e = stashed + kept + count_below(okb);
          const uint32_t gi = info >> 16, bit = (info >> 8) & 0xFFu;
          const uint32_t ce = (uint32_t)slots[4 * g + gi];
          const uint32_t t = bit / 5u, cc = bit - 5u * t;
          st.mean[e] = (float)mean;
          st.count[e] = info & 0xFFu;
          st.value[e] = value;
          st.index[e] =
              (5u * (ce >> 16) + t) * width + 5u * (ce & 0xFFFFu) + cc;
          atomicOr(&cp.s25[gi], 1u << bit);
        }
        kept += (uint32_t)__builtin_popcountll(okb);
      }
      if (kept) {
        wave_sync_lds();
        // each cell with survivors: its record, by the lane of its group
        const uint32_t s25 = lane < 4 ? cp.s25[lane] : 0u;
        const uint32_t ks = (uint32_t)__builtin_popcount(s25);
        const uint64_t lead = ballot(s25 != 0);
        uint32_t run = 0; // survivors of the groups before this lane's
#pragma unroll
        for (int q = 0; q < 3; q++) {
          const uint32_t kq = (uint32_t)__builtin_amdgcn_readlane((int)ks, q);
          run += lane > (uint32_t)q ? kq : 0u;
        }
        if (s25) {
          const uint32_t slot = stashed_cells + count_below(lead);
          st.cell[slot] = (uint32_t)slots[4 * g + lane];
          st.s25[slot] = s25;
          st.at[slot] = stashed + run;
        }
        stashed += kept;
        stashed_cells += (uint32_t)__builtin_popcountll(lead);
      }
      wave_sync_lds();
      in = ahead;
      entry = next_entry;
      valid = next_valid;
    }
  }
  flush(fd, st, stashed, stashed_cells);
}

__global__ __launch_bounds__(256) void poisson32_kernel(const FrameDesc *frames,
                                                        int n_frames,
                                                        uint32_t items) {
  __shared__ uint64_t slots[4][64];
  __shared__ Stash stash[4];
  __shared__ Compact compact[4];
  FrameDesc fd;
  uint32_t local;
  const uint32_t wave = threadIdx.x >> 6;
  if (block_item(frames, n_frames, items, fd, local))
    poisson_item(fd, local, slots[wave], stash[wave], compact[wave]);
}

// Pipelined launches, as the u16 form's (dext_sparse.hip): a chunk's scout
// beside the previous chunk's erosion (one launch) and beside its Poisson
// pass (the next), half the scout items in each, the two kinds spread
// evenly through each XCD's blocks. (Four blocks a CU: 128 VGPRs at most.)
__global__ __launch_bounds__(256, 4) void scout_erode32_kernel(Side scout,
                                                               Side sparse) {
  __shared__ double lut[64];
  __shared__ uint64_t masks[4][2 * kBandCells];
  fill_lut(lut);
  __syncthreads();
  uint32_t index;
  const int role = mixed_role(blockIdx.x, scout.blocks, sparse.blocks, index);
  FrameDesc fd;
  uint32_t local;
  if (role < 0)
    return;
  if (role == 1) {
    if (wave_item(scout.frames, scout.n_frames, scout.first, scout.end,
                  index, fd, local))
      scout_wave(fd, local, lut, masks[threadIdx.x >> 6]);
  } else if (wave_item(sparse.frames, sparse.n_frames, sparse.first,
                       sparse.end, index, fd, local)) {
    erode_item(fd, local, lut);
  }
}

__global__ __launch_bounds__(256, 4) void scout_poisson32_kernel(Side scout,
                                                                 Side sparse) {
  __shared__ double lut[64];
  __shared__ uint64_t masks[4][2 * kBandCells];
  __shared__ uint64_t slots[4][64];
  __shared__ Stash stash[4];
  __shared__ Compact compact[4];
  fill_lut(lut);
  __syncthreads();
  const uint32_t wave = threadIdx.x >> 6;
  uint32_t index;
  const int role = mixed_role(blockIdx.x, scout.blocks, sparse.blocks, index);
  FrameDesc fd;
  uint32_t local;
  if (role < 0)
    return;
  if (role == 1) {
    if (wave_item(scout.frames, scout.n_frames, scout.first, scout.end,
                  index, fd, local))
      scout_wave(fd, local, lut, masks[wave]);
  } else if (wave_item(sparse.frames, sparse.n_frames, sparse.first,
                       sparse.end, index, fd, local)) {
    poisson_item(fd, local, slots[wave], stash[wave], compact[wave]);
  }
}

} // namespace

// The u32 sparse form's threshold on `stream` (clear, scout, erode,
// Poisson); `items`: the launch's scout items. `marks` (optional, two
// events): recorded after the clearing kernel and after the scout.
hipError_t launch_sparse32_threshold(const FrameDesc *frames, int n_frames,
                                     uint32_t items, hipStream_t stream,
                                     hipEvent_t *marks) {
  if (n_frames <= 0)
    return hipSuccess;
  hipError_t e = launch_sparse_zero(frames, n_frames, stream);
  if (e != hipSuccess)
    return e;
  if (marks && (e = hipEventRecord(marks[0], stream)) != hipSuccess)
    return e;
  const uint32_t blocks = ((items + 3) / 4 + 7) & ~7u;
  if (items)
    scout32_kernel<<<blocks, 256, 0, stream>>>(frames, n_frames, items);
  if (marks && (e = hipEventRecord(marks[1], stream)) != hipSuccess)
    return e;
  if (items) {
    erode32_kernel<<<blocks, 256, 0, stream>>>(frames, n_frames, items);
    poisson32_kernel<<<blocks, 256, 0, stream>>>(frames, n_frames, items);
  }
  return hipGetLastError();
}

// One pipelined step of the u32 form on `stream`, as
// launch_sparse_pipeline for the u16 form: the threshold of `cur` (scouted
// by the previous step) finished -- erosion, Poisson, ordering -- while
// `next` is cleared and scouted. Either side may be empty. `marks`
// (optional, four events): before the first mixed launch, after it, after
// the second, after the ordering.
hipError_t launch_sparse32_pipeline(const FrameDesc *cur, int n_cur,
                                    uint32_t cur_items,
                                    uint32_t cur_max_cell_rows,
                                    const FrameDesc *next, int n_next,
                                    uint32_t next_items, hipStream_t stream,
                                    hipEvent_t *marks, bool order) {
  hipError_t e;
  if (n_next > 0 && (e = launch_sparse_zero(next, n_next, stream)) !=
                        hipSuccess)
    return e;
  if (n_next <= 0)
    next_items = 0;
  if (n_cur <= 0)
    cur_items = 0;
  // the scout's items in two halves of whole blocks
  const uint32_t half = ((next_items / 2) + 3) & ~3u;
  const uint32_t split = half < next_items ? half : next_items;
  const uint32_t sparse_blocks = (cur_items + 3) / 4;
  const Side a_scout{next, n_next, 0, split, (split + 3) / 4};
  const Side b_scout{next, n_next, split, next_items,
                     (next_items - split + 3) / 4};
  const Side sparse{cur, n_cur, 0, cur_items, sparse_blocks};
  if (marks && (e = hipEventRecord(marks[0], stream)) != hipSuccess)
    return e;
  const uint32_t grid_a = a_scout.blocks + sparse_blocks
                              ? mixed_grid(a_scout.blocks, sparse_blocks)
                              : 0;
  if (grid_a)
    scout_erode32_kernel<<<grid_a, 256, 0, stream>>>(a_scout, sparse);
  if (marks && (e = hipEventRecord(marks[1], stream)) != hipSuccess)
    return e;
  const uint32_t grid_b = b_scout.blocks + sparse_blocks
                              ? mixed_grid(b_scout.blocks, sparse_blocks)
                              : 0;
  if (grid_b)
    scout_poisson32_kernel<<<grid_b, 256, 0, stream>>>(b_scout, sparse);
  if (marks && (e = hipEventRecord(marks[2], stream)) != hipSuccess)
    return e;
  if (n_cur > 0 && order) {
    e = launch_sparse_order(cur, n_cur, cur_max_cell_rows, stream);
    if (e != hipSuccess)
      return e;
  }
  if (marks && (e = hipEventRecord(marks[3], stream)) != hipSuccess)
    return e;
  return hipGetLastError();
}

hipError_t launch_sparse32(const FrameDesc *frames, int n_frames,
                           uint32_t items, uint32_t max_cell_rows,
                           hipStream_t stream) {
  const hipError_t e =
      launch_sparse32_threshold(frames, n_frames, items, stream, nullptr);
  if (e != hipSuccess)
    return e;
  return launch_sparse_order(frames, n_frames, max_cell_rows, stream);
}

} // namespace sf
