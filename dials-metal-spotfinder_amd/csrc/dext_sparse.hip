// The u16 threshold as four sparse passes: a dense scout that tests one probe
// pixel per 5x5 cell, then the erosion, the Poisson test and the ordering,
// each only where the scout found something.
//
// Reference semantics: src/dext.cc:8-225, exactly (the parity tests compare
// every SignalPixel field with the reference's own output).
//
// Why probes suffice. The output depends on the first mask (mask0: masked or
// dispersion-positive) only through its 5x5 erosion E (src/dext.cc:117-142):
// E(x) = 1 iff mask0 = 1 on every in-frame pixel of the 5x5 window N5(x),
// and E matters only at unmasked pixels (a masked pixel is excluded from the
// 11x11 sums and is never a candidate whatever E says). Cut the frame into
// 5x5 cells on a fixed global lattice; cell (i, j) is rows 5i..5i+4 and
// columns 5j..5j+4 clipped to the frame, and its probe is (5i+2, 5j+2), or
// the frame's last row or column where that is off the frame (a cell cut to
// one or two rows or columns is still within two of it). Every N5(x) of the
// cell contains the probe, so where the probe has mask0 = 0, E = 0 on the
// whole cell. A cell is active when its probe is masked or
// dispersion-positive and the cell has an unmasked pixel; nothing else can
// hold a candidate or an eroded pixel.
//
// The passes, each a kernel over every frame of a launch:
//
//   scout   one wave per 64 cell columns x 16 cell rows: each lane sums its
//           probe's 7x7 window from one 16-byte load per pixel row (shared
//           rows carried between windows), tests it, and the wave appends
//           its active cells to the frame's list and writes their bits.
//   erode   sixteen lanes per active cell, four cells a wave: mask0 on the
//           9x9 pixels around the cell by separable window sums (columns in
//           lanes, rows in registers), then E by bit operations on ballots.
//           Cells with candidates (E, unmasked, p > 0) are listed.
//   poisson sixteen lanes per candidate cell: the 11x11 sums of pixels
//           neither masked nor eroded (E read from the neighbouring cells'
//           tables), the test at the candidates, survivors to staging.
//   order   one wave per cell row: survivors laid out in index order.
#include "sf_device.hh"

namespace sf {

namespace {

typedef unsigned short us2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef u32x4 u32x4_a4 __attribute__((aligned(4))); // 16 bytes, 4-aligned
typedef __attribute__((address_space(1))) const u32x4_a4 *global_q4;
typedef __attribute__((address_space(1))) const uint16_t *global_u16;

__device__ __forceinline__ uint32_t udot2(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_udot2(__builtin_bit_cast(us2, a),
                                __builtin_bit_cast(us2, b), c, false);
}


// sqrt(2 (n - 1)) for n = 2..49 as the reference rounds it (src/dext.cc:108);
// n = 0, 1 give +inf so that "lhs > rhs" is false there (0 * inf is NaN).
__device__ __forceinline__ void fill_lut(float *lut) {
  if (threadIdx.x < 64)
    lut[threadIdx.x] = threadIdx.x < 2 ? __builtin_inff()
                                       : sqrtf(2.0f * (float)(threadIdx.x - 1));
}

// The frame as a buffer resource: reads past its end return 0, never fault.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t frame_buffer(
    const FrameDesc &fd) {
  return __builtin_amdgcn_make_buffer_rsrc(
      const_cast<void *>(fd.pixels), 0,
      (int)(2u * fd.height * fd.width), 0x00020000);
}

// One pixel row of a lane's probe window. The lane holds eight columns from
// an even column `ls`; m[k] keeps the window's columns of word k (0xFFFF a
// half) and w5[k] weighs the cell's own columns. With p = v + 1 at unmasked
// pixels and 0 at masked ones (v + 2 wraps the two sentinels to 0 and 1):
// s += sum p, q += sum p^2, n7 += window count, n5 += cell count. `one` is
// {1, 1} held in a register, so that min(p, 1) stays one packed minimum.
__device__ __forceinline__ void scout_row(const uint32_t (&d)[4],
                                          const uint32_t (&m)[4],
                                          const uint32_t (&w5)[4], us2 one,
                                          uint32_t &s, uint32_t &q,
                                          uint32_t &n7, uint32_t &n5) {
  const us2 two = {2, 2};
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const us2 a = __builtin_bit_cast(us2, d[k]) + two;
    const uint32_t p =
        __builtin_bit_cast(uint32_t, __builtin_elementwise_sub_sat(a, one)) &
        m[k];
    const uint32_t f = __builtin_bit_cast(
        uint32_t, __builtin_elementwise_min(__builtin_bit_cast(us2, p), one));
    s = udot2(p, 0x00010001u, s);
    q = udot2(p, p, q);
    n7 = udot2(f, 0x00010001u, n7);
    n5 = udot2(f, w5[k], n5);
  }
}

// The dispersion test (src/dext.cc:97-109) from window sums of p = v + 1
// over the window's unmasked pixels (s = sum p, q = sum p^2, n7 their
// count): the sums of v follow exactly, modulo 2^32 as the reference's
// 32-bit accumulator wraps. In float, in the reference's order.
__device__ __forceinline__ bool dispersed(uint32_t s, uint32_t q, uint32_t n7,
                                          const float *lut) {
  const uint32_t sv = s - n7;
  const uint32_t qv = q - 2u * sv - n7;
  const float fm = (float)n7, fs = (float)sv, fq = (float)qv;
  const float lhs = fm * fq - fs * fs - fs * (fm - 1.0f);
  const float rhs = fs * 6.0f * lut[n7];
  return lhs > rhs;
}

// Scout one item: cell columns 64 * strip + lane, cell rows
// [16 * band, 16 * band + 16). VEC: the frame's rows are 4-byte aligned and
// at least 8 pixels wide, so every lane reads its 8 columns as one 16-byte
// load (a buffer load: the row's offset is a scalar).
template <bool VEC>
__device__ void scout_item(const FrameDesc &fd, uint32_t strip, uint32_t band,
                           const float *lut, uint64_t *masks) {
  const int height = (int)fd.height, width = (int)fd.width;
  const int ncr = (int)fd.n_cell_rows, ncc = (int)fd.n_cell_cols;
  const int lane = (int)lane_id();
  const int j = 64 * (int)strip + lane;
  const bool live = j < ncc;
  const int pc = min(5 * j + 2, width - 1);
  // window columns [pc-3, pc+3] and the cell's columns, in the frame
  const int wlo = max(pc - 3, 0), whi = min(pc + 3, width - 1);
  const int clo = 5 * j, chi = min(5 * j + 4, width - 1);
  int ls = (pc - 3) & ~1;
  if (VEC)
    ls = min(max(ls, 0), width - 8);
  uint32_t m[4], w5[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    m[k] = 0;
    w5[k] = 0;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int c = ls + 2 * k + h;
      if (live && c >= wlo && c <= whi)
        m[k] |= 0xFFFFu << (16 * h);
      if (live && c >= clo && c <= chi)
        w5[k] |= 1u << (16 * h);
    }
  }
  us2 one = {1, 1};
  __asm__ volatile("" : "+v"(one));
  const int probe_at = pc - ls; // 0..7 when the probe is in the frame
  const uint16_t *pix = static_cast<const uint16_t *>(fd.pixels);
  const __amdgpu_buffer_rsrc_t rsrc = frame_buffer(fd);
  const uint32_t lane_at = 2u * (uint32_t)ls;

  // a pixel row's eight columns; rows off the frame read as masked (CHECK:
  // the item reaches the frame's top or bottom)
  const auto load_row = [&]<bool CHECK>(int y, uint32_t (&d)[4]) {
    if (CHECK && (y < 0 || y >= height)) { // uniform
#pragma unroll
      for (int k = 0; k < 4; k++)
        d[k] = 0xFFFFFFFFu;
      return;
    }
    if (VEC) {
      typedef int i32x4 __attribute__((ext_vector_type(4)));
      const i32x4 v = __builtin_amdgcn_raw_buffer_load_b128(
          rsrc, (int)lane_at, (int)(2u * (uint32_t)y * (uint32_t)width), 0);
      d[0] = (uint32_t)v.x;
      d[1] = (uint32_t)v.y;
      d[2] = (uint32_t)v.z;
      d[3] = (uint32_t)v.w;
    } else {
      const uint16_t *row = pix + (size_t)y * (size_t)width;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        uint32_t lo = 0xFFFFu, hi = 0xFFFFu;
        if (m[k] & 0xFFFFu)
          lo = *(global_u16)(row + ls + 2 * k);
        if (m[k] >> 16)
          hi = *(global_u16)(row + ls + 2 * k + 1);
        d[k] = lo | (hi << 16);
      }
    }
  };
  const auto load_rows = [&]<bool CHECK>(int y, uint32_t (&rows)[5][4]) {
#pragma unroll
    for (int r = 0; r < 5; r++)
      load_row.template operator()<CHECK>(y + r, rows[r]);
  };

  // Whether full cell row i's cell (lane's column j), whose probe is masked,
  // holds no eroded pixel (src/dext.cc:117-142): E(x) = 1 needs mask0 on
  // every in-frame pixel of x's 5x5 window, so one unmasked pixel z with
  // mask0(z) = 0 clears E within two rows and columns of it. The points
  // tried lie on the probe's column pc (every column of the cell is within
  // two of it): row 5i+1 (else 5i) and row 5i+3 (else 5i+4), where
  // unmasked; their 7x7 windows have the probe window's columns, so the
  // lane's eight loaded columns serve. The cell is free when each of its
  // rows that holds an unmasked pixel is within two rows of a point whose
  // test is negative. Rows 5i-3 .. 5i+7 are read again, once.
  const auto eroded_free = [&](int i) -> bool {
    uint32_t ws[4] = {0, 0, 0, 0}, wq[4] = {0, 0, 0, 0}, wn[4] = {0, 0, 0, 0};
    uint32_t um = 0, zc = 0; // per cell row: unmasked pixel / at column pc
    const int pk = probe_at >> 1, half = 16 * (probe_at & 1);
#pragma unroll 1
    for (int o = -3; o <= 7; o++) {
      uint32_t d[4];
      load_row.template operator()<true>(5 * i + o, d);
      uint32_t s_ = 0, q_ = 0, n_ = 0, c_ = 0;
      scout_row(d, m, w5, one, s_, q_, n_, c_);
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
        const uint32_t w = pk == 0   ? d[0]
                           : pk == 1 ? d[1]
                           : pk == 2 ? d[2]
                                     : d[3];
        um |= (c_ != 0 ? 1u : 0u) << o;
        zc |= (((w >> half) & 0xFFFFu) < 0xFFFEu ? 1u : 0u) << o;
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
  uint32_t ts = 0, tq = 0, tn = 0, top5 = 0, unused = 0;
  {
    uint32_t d[4];
    load_row.template operator()<true>(5 * i0 - 1, d);
    scout_row(d, m, w5, one, ts, tq, tn, unused);
    load_row.template operator()<true>(5 * i0, d);
    scout_row(d, m, w5, one, ts, tq, tn, top5);
  }
  // cell row i from its rows 5i+1 .. 5i+5 and the carried pair 5i-1, 5i
  const auto cell_row = [&](int i, const uint32_t (&rows)[5][4]) {
    uint32_t s = ts, q = tq, n7 = tn, n5 = top5;
    uint32_t s2 = 0, q2 = 0, n2 = 0, n5n = 0;
    scout_row(rows[0], m, w5, one, s, q, n7, n5);
    scout_row(rows[1], m, w5, one, s, q, n7, n5);
    scout_row(rows[2], m, w5, one, s, q, n7, n5);
    scout_row(rows[3], m, w5, one, s2, q2, n2, n5);
    scout_row(rows[4], m, w5, one, s2, q2, n2, n5n);
    s += s2;
    q += q2;
    n7 += n2;
    ts = s2;
    tq = q2;
    tn = n2;
    top5 = n5n;
    // the probe's own value, from row 5i+2
    uint32_t probe_row[4] = {rows[1][0], rows[1][1], rows[1][2], rows[1][3]};
    if (5 * i + 2 >= height) { // uniform: a last cell row of one or two rows
      // probe on the last row instead, window rows height-4 .. height-1
      s = q = n7 = 0;
      uint32_t d[4];
#pragma unroll
      for (int r = 4; r >= 1; r--) {
        load_row.template operator()<true>(height - r, d);
        scout_row(d, m, w5, one, s, q, n7, unused);
      }
#pragma unroll
      for (int k = 0; k < 4; k++)
        probe_row[k] = d[k];
    }
    const int pk = probe_at >> 1;
    const uint32_t word = pk == 0   ? probe_row[0]
                          : pk == 1 ? probe_row[1]
                          : pk == 2 ? probe_row[2]
                                    : probe_row[3];
    const uint32_t pv = (word >> (16 * (probe_at & 1))) & 0xFFFFu;
    const bool masked = pv >= 0xFFFEu;
    const bool active = live && (masked || dispersed(s, q, n7, lut)) && n5 > 0;
    // A full cell whose probe lies in a gap (0xFFFF: a module gap that cuts
    // the cell) usually holds no eroded pixel: such cells are checked after
    // the sweep, when its registers are free. (Cells whose probe is a hot
    // pixel, 0xFFFE, are scattered one to a wave: re-reading their windows
    // would cost more than their erosion, so they stay.)
    const bool full = 5 * i + 4 < height && 5 * j + 4 < width;
    const uint64_t bits = ballot(active);
    const uint64_t check = ballot(active && pv == 0xFFFFu && full);
    if (lane == 0) {
      masks[i - i0] = bits;
      masks[kBandCells + i - i0] = check;
    }
  };
  // two row sets in turn: one loads while the other is summed
  const auto sweep = [&]<bool CHECK>() {
    uint32_t ra[5][4], rb[5][4];
    load_rows.template operator()<CHECK>(5 * i0 + 1, ra);
    for (int i = i0; i < i1; i += 2) {
      if (i + 1 < i1)
        load_rows.template operator()<CHECK>(5 * i + 6, rb);
      cell_row(i, ra);
      if (i + 1 >= i1)
        break;
      if (i + 2 < i1)
        load_rows.template operator()<CHECK>(5 * i + 11, ra);
      cell_row(i + 1, rb);
    }
  };
  if (5 * i1 + 5 <= height) // rows 5 i0 + 1 .. 5 i1 + 4 all in the frame
    sweep.template operator()<false>();
  else
    sweep.template operator()<true>();
  // drop the checked cells that hold no eroded pixel, then publish the bits
  for (int i = i0; i < i1; i++) {
    uint64_t bits = masks[i - i0];
    const uint64_t check = masks[kBandCells + i - i0];
    if (check) { // uniform, and rare
      const bool mine = (check >> lane) & 1u;
      bits &= ~ballot(eroded_free(i) && mine);
      if (lane == 0)
        masks[i - i0] = bits;
    }
    if (lane == 0)
      fd.act[(size_t)i * fd.n_strips + strip] = bits;
  }
  // the item's active cells, in its own segment of the list
  uint32_t total = 0;
  const uint32_t local = band * fd.n_strips + strip;
  uint32_t *list = fd.alist + (size_t)local * kSegment;
  for (int i = i0; i < i1; i++) {
    const uint64_t bits = masks[i - i0];
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
                                           const float *lut, uint64_t *masks) {
  const uint32_t strip = local % fd.n_strips, band = local / fd.n_strips;
  if (fd.vec && fd.width >= 8)
    scout_item<true>(fd, strip, band, lut, masks);
  else
    scout_item<false>(fd, strip, band, lut, masks);
}

__global__ __launch_bounds__(256) void scout_kernel(const FrameDesc *frames,
                                                    int n_frames,
                                                    uint32_t items) {
  __shared__ float lut[64];
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
// Sixteen lanes to a cell, four cells to a wave. Lane c of a cell's group
// holds pixel column 5j - 5 + c (c = 15 idles: no result reads it) and the
// cell's rows 5i - 5 .. 5i + 9 in registers. Row sums are prefix
// differences; column sums move within the group's DPP row (16 lanes):
// right<N> reads lane c + N, left<N> lane c - N, and lanes past the row's end
// read 0.
// ---------------------------------------------------------------------------
template <int N> __device__ __forceinline__ uint32_t right(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x100 + N, 0xF, 0xF, true);
}
template <int N> __device__ __forceinline__ uint32_t left(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x110 + N, 0xF, 0xF, true);
}
// x over lanes c .. c+6: the 7-wide window centred on lane c+3 (valid at
// c = 0..8). Every step folds its lane move into an add (additions only: a
// subtraction of a moved operand, v_subrev_u32_dpp, has been seen to take
// the move from the other operand on this hardware).
__device__ __forceinline__ uint32_t window7(uint32_t x) {
  const uint32_t a = x + right<1>(x); // c .. c+1
  const uint32_t b = a + right<2>(a); // c .. c+3
  return (b + right<4>(a)) + right<6>(x);
}
// x over lanes c .. c+10: the 11-wide window centred on lane c+5 (valid at
// c = 0..4).
__device__ __forceinline__ uint32_t window11(uint32_t x) {
  const uint32_t a = x + right<1>(x); // c .. c+1
  const uint32_t b = a + right<2>(a); // c .. c+3
  const uint32_t d = b + right<4>(b); // c .. c+7
  return (d + right<8>(a)) + right<10>(x);
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
  const uint32_t at = 2u * (uint32_t)(y0 * width + x);
  const uint32_t pitch = 2u * (uint32_t)width;
#pragma unroll
  for (int r = 0; r < 15; r++)
    v[r] = (uint32_t)(uint16_t)__builtin_amdgcn_raw_buffer_load_b16(
        rsrc, (int)(at + (uint32_t)r * pitch), 0, 0);
  // the frame's edges and invalid groups (uniform test): mask what is off
  const bool edge = !valid || (c < 15 && (x < 0 || x >= width)) || y0 < 0 ||
                    y0 + 15 > height;
  if (ballot(edge)) {
    const bool col = valid && x >= 0 && x < width;
#pragma unroll
    for (int r = 0; r < 15; r++)
      if (!col || (uint32_t)(y0 + r) >= (uint32_t)height)
        v[r] = 0xFFFFu;
  }
}

// mask0 on the 9x9 pixels around a cell, then E on the cell and its
// candidates (E, unmasked, 0 < p): src/dext.cc:65-142. Writes the cell's
// E and clears its survivors; returns its candidates.
__device__ uint32_t erode_group(const FrameDesc &fd, bool valid, uint32_t cell,
                                const uint32_t (&v)[15], const float *lut) {
  const uint32_t lane = lane_id();
  // count above bit 26 and sum, packed; squares; as prefix sums over rows
  uint32_t pn[16], pq[16];
  pn[0] = pq[0] = 0;
#pragma unroll
  for (int r = 0; r < 15; r++) {
    const uint32_t x = v[r] < 0xFFFEu ? (v[r] | (1u << 26)) : 0u;
    pn[r + 1] = pn[r] + x;
    pq[r + 1] = pq[r] + __umul24(x, x); // low 24 bits: v
  }
  // mask0 or masked ("set") at pixel rows 5i-2+o, columns 3..11: 7x7
  // windows (src/dext.cc:69-111), tested on lane c for column c+3
  uint64_t h5[9];
#pragma unroll
  for (int o = 0; o < 9; o++) {
    const uint32_t ns = window7(pn[o + 7] - pn[o]);
    const uint32_t sq = window7(pq[o + 7] - pq[o]);
    const uint32_t n = ns >> 26, sum = ns & 0x3FFFFFFu;
    const float fm = (float)n, fs = (float)sum, fq = (float)sq;
    const float lhs = fm * fq - fs * fs - fs * (fm - 1.0f);
    const float rhs = fs * 6.0f * lut[n];
    const uint64_t b =
        (ballot(lhs > rhs) << 3) | ballot(v[o + 3] >= 0xFFFEu);
    // set on all five columns c-2 .. c+2 (bit c: column c of its group)
    h5[o] = b & (b << 1) & (b << 2) & (b >> 1) & (b >> 2);
  }
  const uint32_t at = 16u * (lane >> 4) + 5u;
  uint32_t e25 = 0, c25 = 0;
#pragma unroll
  for (int t = 0; t < 5; t++) {
    const uint32_t p = v[t + 5];
    const uint64_t unmasked = ballot(p < 0xFFFEu);
    const uint64_t positive = ballot(p - 1u < 0xFFFDu); // 0 < p < 0xFFFE
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
                           const float *lut) {
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

__global__ __launch_bounds__(256) void erode_kernel(const FrameDesc *frames,
                                                    int n_frames,
                                                    uint32_t items) {
  __shared__ float lut[64];
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
                                              uint32_t ns, uint32_t p);

// The 11x11 sums of a cell's pixels neither masked nor eroded
// (src/dext.cc:146-189), for the lane's column c of the cell (lanes c = 0..4
// of the group), row by row of the cell: each candidate's packed count and
// sum (count << 23 | sum) and its value go to the step's compacted list as
// they are formed. `mine`: the cell's candidates; `first`: the list entry of
// its first.
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
  // count above bit 23 and sum, packed, as prefix sums over rows
  uint32_t pn[16];
  pn[0] = 0;
#pragma unroll
  for (int r = 0; r < 15; r++) {
    const uint32_t out = (e[r / 5] >> (5u * (r % 5) + cc)) & 1u;
    const uint32_t xr =
        (in.v[r] < 0xFFFEu && !out) ? (in.v[r] | (1u << 23)) : 0u;
    pn[r + 1] = pn[r] + xr;
  }
#pragma unroll
  for (int t = 0; t < 5; t++)
    put_candidate(cp, mine, first, t, window11(pn[t + 11] - pn[t]),
                  right<5>(in.v[t + 5]));
}

// The Poisson test (src/dext.cc:201-206) in the reference's order and
// rounding, from the 11x11 window's count and sum and the candidate's value
// p; the mean goes out through `mean`.
__device__ __forceinline__ bool poisson_signal(uint32_t ns, uint32_t p,
                                               float &mean) {
  const uint32_t n = ns >> 23, sum = ns & 0x7FFFFFu;
  const float m = n >= 2 ? (float)sum / (float)n : 0.0f;
  mean = m;
  return (float)p >= m + 3.0f * sqrtf(m);
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
constexpr int kStash = 512;
constexpr int kStashCells = 64;
struct Stash {
  float mean[kStash];
  uint32_t pop[kStash];   // count | value << 16
  uint32_t index[kStash]; // pixel index in the frame
  uint32_t cell[kStashCells];
  uint32_t s25[kStashCells];
  uint32_t at[kStashCells]; // the cell's run in the stash
};

// A wave-step's candidates, compacted in (cell, row, column) order: at most
// four cells of 25 pixels.
constexpr int kCompact = 100;
struct Compact {
  uint32_t ns[kCompact];  // window count << 23 | sum
  uint32_t pix[kCompact]; // value | (5 row + column) << 16 | group << 24
  uint32_t s25[4];        // each group's survivors
};

// Lane c (0..4) of a group: its pixel on the cell's row t, if a candidate,
// to its place in the step's compacted list.
__device__ __forceinline__ void put_candidate(Compact &cp, uint32_t mine,
                                              uint32_t first, int t,
                                              uint32_t ns, uint32_t p) {
  const uint32_t lane = lane_id(), c = lane & 15u;
  const uint32_t bit = 5u * (uint32_t)t + c;
  if (c <= 4 && ((mine >> bit) & 1u)) {
    const uint32_t at =
        first + (uint32_t)__builtin_popcount(mine & ((1u << bit) - 1u));
    cp.ns[at] = ns;
    cp.pix[at] = p | (bit << 16) | ((lane >> 4) << 24);
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
        px.value = st.pop[e] >> 16;
        px.background = st.mean[e];
        px.population = (uint16_t)(st.pop[e] & 0xFFFFu);
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
// each, and tested there, so that the division and square root run once a
// candidate; survivors gather in the stash in that order, which is each
// cell's pixel order.
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
    const uint32_t nc = (uint32_t)__builtin_popcountll(has);
    const uint32_t steps = (nc + 3) / 4;
    if (!steps)
      continue;
    wave_sync_lds();
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
        float mean = 0.0f;
        uint32_t ns_at = 0, pix = 0;
        if (at < total) {
          ns_at = cp.ns[at];
          pix = cp.pix[at];
          ok = poisson_signal(ns_at, pix & 0xFFFFu, mean);
        }
        const uint64_t okb = ballot(ok);
        if (ok) {
          const uint32_t e = stashed + kept + count_below(okb);
          const uint32_t gi = pix >> 24, bit = (pix >> 16) & 0xFFu;
          const uint32_t ce = (uint32_t)slots[4 * g + gi];
          const uint32_t t = bit / 5u, cc = bit - 5u * t;
          st.mean[e] = mean;
          st.pop[e] = (ns_at >> 23) | ((pix & 0xFFFFu) << 16);
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

__global__ __launch_bounds__(256) void poisson_kernel(const FrameDesc *frames,
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

// ---------------------------------------------------------------------------
// Pipelined launches. The scout streams every pixel from HBM with VALU to
// spare; erosion and Poisson read little but keep the VALU busy. A chunk's
// scout therefore runs beside the previous chunk's erosion (one launch) and
// beside its Poisson pass (the next launch): half the scout items in each.
// Blocks of the two kinds alternate in the launch, so that every CU holds
// both. No block of a launch waits for another: each launch's kinds work on
// different chunks, whose tables are apart.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void scout_erode_kernel(Side scout,
                                                          Side sparse) {
  __shared__ float lut[64];
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

// (Four blocks a CU, the scout's occupancy: 128 VGPRs at most.)
__global__ __launch_bounds__(256, 4) void scout_poisson_kernel(Side scout,
                                                            Side sparse) {
  __shared__ float lut[64];
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

// Inclusive sum of x over the lanes up to this one (DPP: within rows of 16,
// then row 0's and row 1's totals broadcast upwards).
__device__ __forceinline__ uint32_t wave_prefix(uint32_t x) {
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, false);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, false);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, false);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, false);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false);
  return x;
}

// rowsurv, per pixel row, becomes its exclusive prefix (the row's first
// place in the list), with the frame's total after the last row. One block a
// frame.
__global__ __launch_bounds__(1024) void rowscan_kernel(const FrameDesc *frames) {
  __shared__ uint32_t sums[16];
  __shared__ uint32_t carry;
  const FrameDesc fd = frames[blockIdx.x];
  const uint32_t n = 5 * fd.n_cell_rows;
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  uint32_t run = 0; // the carry, as thread 1023 knows it
  if (threadIdx.x == 0)
    carry = 0;
  __syncthreads();
  for (uint32_t b = 0; b < n; b += 4096) {
    const uint32_t at = b + 4 * threadIdx.x;
    uint32_t v[4], local = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      v[q] = at + q < n ? fd.rowsurv[at + q] : 0u;
      local += v[q];
    }
    const uint32_t incl = wave_prefix(local);
    if (lane == 63)
      sums[wave] = incl;
    __syncthreads();
    uint32_t before = carry;
    for (uint32_t w = 0; w < wave; w++)
      before += sums[w];
    uint32_t x = before + incl - local;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      if (at + q < n)
        fd.rowsurv[at + q] = x;
      x += v[q];
    }
    run = x;
    __syncthreads();
    if (threadIdx.x == 1023)
      carry = run;
    __syncthreads();
  }
  if (threadIdx.x == 0)
    fd.rowsurv[n] = carry;
}

constexpr int kTagWords = 8; // act words a cell row's pass reads at once

// Places in index order for one cell row's survivors: pixel row 5i+t before
// 5i+t+1, within a pixel row by cell, then by column. The row's cells that
// hold survivors are staged in LDS in column order, up to 64 at a time (one
// per lane), from kTagWords act words at a time; for each 64 staged cells,
// one prefix over them per pixel row gives each cell its places. Each of
// the row's staging entries gets its place; no entry is read.
__device__ __forceinline__ void tag_row(const FrameDesc &fd, uint32_t i,
                                        uint32_t *cell_s, uint32_t *cell_from) {
  const uint32_t lane = lane_id();
  const uint32_t *first = fd.rowsurv + 5 * i; // places of the five rows
  if (first[5] == first[0])
    return; // no survivor on the cell row
  uint32_t start[5];
#pragma unroll
  for (int t = 0; t < 5; t++)
    start[t] = first[t];
  // the staged cells' places, and then the stage is empty again
  uint32_t staged = 0;
  auto place = [&]() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint32_t cs = lane < staged ? cell_s[lane] : 0u;
    const uint32_t cf = lane < staged ? cell_from[lane] : 0u;
#pragma unroll
    for (int t = 0; t < 5; t++) {
      const uint32_t k = (uint32_t)__builtin_popcount((cs >> (5 * t)) & 31u);
      const uint32_t incl = wave_prefix(k);
      const uint32_t dst = start[t] + incl - k;
      const uint32_t src =
          cf + (uint32_t)__builtin_popcount(cs & ((1u << (5 * t)) - 1u));
#pragma unroll
      for (int q = 0; q < 5; q++)
        if ((uint32_t)q < k && src + q < fd.capacity)
          fd.dstof[src + q] = dst + q;
      start[t] += (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    staged = 0;
  };
  const uint64_t *act = fd.act + (size_t)i * fd.n_strips;
  for (uint32_t w0 = 0; w0 < fd.n_strips; w0 += kTagWords) {
    uint64_t bits[kTagWords];
    uint32_t s[kTagWords], from[kTagWords];
#pragma unroll
    for (int u = 0; u < kTagWords; u++)
      bits[u] = w0 + u < fd.n_strips ? act[w0 + u] : 0ull;
#pragma unroll
    for (int u = 0; u < kTagWords; u++) {
      const uint32_t ci = i * fd.n_cell_cols + 64 * (w0 + u) + lane;
      const bool on = (bits[u] >> lane) & 1u;
      s[u] = on ? fd.scell[ci] : 0u;
      from[u] = on ? fd.sbase[ci] : 0u;
    }
#pragma unroll
    for (int u = 0; u < kTagWords; u++) {
      const uint64_t has = ballot(s[u] != 0);
      const uint32_t nh = (uint32_t)__builtin_popcountll(has);
      if (staged + nh > 64)
        place();
      if (s[u]) {
        const uint32_t at = staged + count_below(has);
        cell_s[at] = s[u];
        cell_from[at] = from[u];
      }
      staged += nh;
    }
  }
  if (staged)
    place();
}

// Grid (cell-row quads, frames): a wave a cell row. Places are given only
// when the list fits (the capacity contract of src/dext_gpu.cc).
__global__ __launch_bounds__(256) void tag_kernel(const FrameDesc *frames) {
  // per wave: the staged cells, one per lane
  __shared__ uint32_t cells[4][2][64];
  const FrameDesc fd = frames[blockIdx.y];
  const uint32_t wave = threadIdx.x >> 6;
  const uint32_t i = (uint32_t)__builtin_amdgcn_readfirstlane(
      (int)(4 * blockIdx.x + wave));
  if (i >= fd.n_cell_rows || *fd.counter > fd.capacity)
    return;
  tag_row(fd, i, cells[wave][0], cells[wave][1]);
}

// Every staging entry to its place; the count goes out whether or not the
// list fits. Grid (blocks, frames).
__global__ __launch_bounds__(256) void place_kernel(const FrameDesc *frames) {
  const FrameDesc fd = frames[blockIdx.y];
  const uint32_t found = *fd.counter;
  if (blockIdx.x == 0 && threadIdx.x == 0)
    *fd.out_count = found;
  if (found > fd.capacity)
    return;
  const uint4 *from = reinterpret_cast<const uint4 *>(fd.staging);
  uint4 *to = reinterpret_cast<uint4 *>(fd.out);
  for (uint32_t e = blockIdx.x * 256 + threadIdx.x; e < found;
       e += gridDim.x * 256) {
    const uint32_t at = fd.dstof[e];
    if (at < found)
      to[at] = from[e];
  }
}

__global__ void zero_kernel(const FrameDesc *frames) {
  const FrameDesc &fd = frames[blockIdx.x];
  if (threadIdx.x == 0)
    *fd.counter = 0;
  for (uint32_t r = threadIdx.x; r < 5 * fd.n_cell_rows; r += blockDim.x)
    fd.rowsurv[r] = 0;
}

} // namespace

// The clearing kernel of the sparse forms on its own (the u32 form's
// threshold starts with it).
hipError_t launch_sparse_zero(const FrameDesc *frames, int n_frames,
                              hipStream_t stream) {
  if (n_frames > 0)
    zero_kernel<<<n_frames, 256, 0, stream>>>(frames);
  return hipGetLastError();
}

// Items of a frame for the scout: 64-cell strips x 16-cell-row bands.
void sparse_layout(uint32_t height, uint32_t width, FrameDesc &d) {
  d.n_cell_rows = (height + 4) / 5;
  d.n_cell_cols = (width + 4) / 5;
  d.n_strips = (d.n_cell_cols + 63) / 64;
  d.band_rows = 5 * kBandCells;
  d.n_bands = (d.n_cell_rows + kBandCells - 1) / kBandCells;
}

// Bytes of the sparse tables of one frame, and their layout from `base`
// (256-byte aligned pieces); `capacity`: the frame's list entries.
static size_t sparse_piece(size_t b) { return (b + 255) & ~(size_t)255; }

size_t sparse_bytes(uint32_t height, uint32_t width, uint32_t capacity) {
  FrameDesc d{};
  sparse_layout(height, width, d);
  const size_t cells = (size_t)d.n_cell_rows * d.n_cell_cols;
  const size_t items = (size_t)d.n_strips * d.n_bands;
  return sparse_piece(8 * (size_t)d.n_cell_rows * d.n_strips) +
         sparse_piece(4 * items) + 2 * sparse_piece(4 * items * kSegment) +
         3 * sparse_piece(4 * cells) +
         sparse_piece(20 * (size_t)d.n_cell_rows + 4) +
         sparse_piece(4 * (size_t)capacity);
}

void sparse_carve(uint8_t *base, FrameDesc &d) {
  const size_t cells = (size_t)d.n_cell_rows * d.n_cell_cols;
  const size_t items = (size_t)d.n_strips * d.n_bands;
  size_t at = 0;
  const auto take = [&](size_t bytes) {
    uint8_t *p = base + at;
    at += sparse_piece(bytes);
    return p;
  };
  d.act = reinterpret_cast<uint64_t *>(take(8 * (size_t)d.n_cell_rows * d.n_strips));
  d.icount = reinterpret_cast<uint32_t *>(take(4 * items));
  d.alist = reinterpret_cast<uint32_t *>(take(4 * items * kSegment));
  d.ccand = reinterpret_cast<uint32_t *>(take(4 * items * kSegment));
  d.ecell = reinterpret_cast<uint32_t *>(take(4 * cells));
  d.scell = reinterpret_cast<uint32_t *>(take(4 * cells));
  d.sbase = reinterpret_cast<uint32_t *>(take(4 * cells));
  d.rowsurv =
      reinterpret_cast<uint32_t *>(take(20 * (size_t)d.n_cell_rows + 4));
  d.dstof = reinterpret_cast<uint32_t *>(take(4 * (size_t)d.capacity));
}

// The sparse form on `stream` in two parts: the threshold (clear, scout,
// erode, Poisson) and the ordering (row prefix, places, copy). `items`: the
// launch's scout items; `max_cell_rows`: the most cell rows of any frame.
// `marks` (optional, two events): recorded after the clearing kernel and
// after the scout, to time the scout on its own.
hipError_t launch_sparse_threshold(const FrameDesc *frames, int n_frames,
                                   uint32_t items, hipStream_t stream,
                                   hipEvent_t *marks) {
  if (n_frames <= 0)
    return hipSuccess;
  zero_kernel<<<n_frames, 256, 0, stream>>>(frames);
  hipError_t e;
  if (marks && (e = hipEventRecord(marks[0], stream)) != hipSuccess)
    return e;
  const uint32_t blocks = ((items + 3) / 4 + 7) & ~7u;
  if (items)
    scout_kernel<<<blocks, 256, 0, stream>>>(frames, n_frames, items);
  if (marks && (e = hipEventRecord(marks[1], stream)) != hipSuccess)
    return e;
  if (items) {
    erode_kernel<<<blocks, 256, 0, stream>>>(frames, n_frames, items);
    poisson_kernel<<<blocks, 256, 0, stream>>>(frames, n_frames, items);
  }
  return hipGetLastError();
}

hipError_t launch_sparse_order(const FrameDesc *frames, int n_frames,
                               uint32_t max_cell_rows, hipStream_t stream) {
  if (n_frames <= 0)
    return hipSuccess;
  rowscan_kernel<<<n_frames, 1024, 0, stream>>>(frames);
  if (max_cell_rows)
    tag_kernel<<<dim3((max_cell_rows + 3) / 4, n_frames), 256, 0, stream>>>(
        frames);
  place_kernel<<<dim3(64, n_frames), 256, 0, stream>>>(frames);
  return hipGetLastError();
}

// The ordering in two parts, for a reader that takes the survivors from
// staging by their places (dstof) itself: the places (row prefix and tags;
// each frame's count to out_count too), then, should the lists be wanted
// later, the copy into them.
namespace {
__global__ void found_kernel(const FrameDesc *frames) {
  const FrameDesc &fd = frames[blockIdx.x];
  if (threadIdx.x == 0)
    *fd.out_count = *fd.counter;
}
} // namespace

hipError_t launch_sparse_places(const FrameDesc *frames, int n_frames,
                                uint32_t max_cell_rows, hipStream_t stream) {
  if (n_frames <= 0)
    return hipSuccess;
  rowscan_kernel<<<n_frames, 1024, 0, stream>>>(frames);
  if (max_cell_rows)
    tag_kernel<<<dim3((max_cell_rows + 3) / 4, n_frames), 256, 0, stream>>>(
        frames);
  found_kernel<<<n_frames, 64, 0, stream>>>(frames);
  return hipGetLastError();
}

hipError_t launch_sparse_copy(const FrameDesc *frames, int n_frames,
                              hipStream_t stream) {
  if (n_frames <= 0)
    return hipSuccess;
  place_kernel<<<dim3(64, n_frames), 256, 0, stream>>>(frames);
  return hipGetLastError();
}

// Blocks of a mixed launch: on each XCD its eighth of both kinds.
uint32_t mixed_grid(uint32_t scout_blocks, uint32_t sparse_blocks) {
  return 8 * ((scout_blocks + 7) / 8 + (sparse_blocks + 7) / 8);
}

// One pipelined step on `stream`: the threshold of `cur` (n_cur frames,
// scouted by the previous step) finished -- erosion, Poisson, and the
// ordering unless `order` is false (then launch_sparse_order runs it later)
// -- while `next` (n_next frames) is cleared and scouted. Either side may
// be empty. `marks` (optional, four events): before the first mixed launch,
// after it, after the second, after the ordering.
hipError_t launch_sparse_pipeline(const FrameDesc *cur, int n_cur,
                                  uint32_t cur_items,
                                  uint32_t cur_max_cell_rows,
                                  const FrameDesc *next, int n_next,
                                  uint32_t next_items, hipStream_t stream,
                                  hipEvent_t *marks, bool order) {
  hipError_t e;
  if (n_next > 0)
    zero_kernel<<<n_next, 256, 0, stream>>>(next);
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
    scout_erode_kernel<<<grid_a, 256, 0, stream>>>(a_scout, sparse);
  if (marks && (e = hipEventRecord(marks[1], stream)) != hipSuccess)
    return e;
  const uint32_t grid_b = b_scout.blocks + sparse_blocks
                              ? mixed_grid(b_scout.blocks, sparse_blocks)
                              : 0;
  if (grid_b)
    scout_poisson_kernel<<<grid_b, 256, 0, stream>>>(b_scout, sparse);
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

hipError_t launch_sparse(const FrameDesc *frames, int n_frames, uint32_t items,
                         uint32_t max_cell_rows, hipStream_t stream) {
  const hipError_t e =
      launch_sparse_threshold(frames, n_frames, items, stream, nullptr);
  if (e != hipSuccess)
    return e;
  return launch_sparse_order(frames, n_frames, max_cell_rows, stream);
}

} // namespace sf
