// The u16 threshold as a sparse tile kernel: test the 7x7 dispersion window at
// one probe pixel per 5x5 cell, and everywhere else only where a probe says
// it can matter.
//
// Reference semantics: src/dext.cc:8-225, exactly (the parity tests compare
// every SignalPixel field with the reference's own output).
//
// Why probes suffice. The output depends on the first mask (mask0: masked or
// dispersion-positive) only through its 5x5 erosion E (src/dext.cc:117-142):
// E(x) = 1 iff mask0 = 1 on every in-frame pixel of the 5x5 window N5(x),
// and E matters only at unmasked pixels (a masked pixel is excluded from the
// 11x11 sums and is never a candidate whatever E says). Cut the frame into
// 5x5 cells on a fixed global lattice and call each cell's centre its probe:
// every N5(x) contains the probe of x's own cell. So where that probe has
// mask0 = 0, E = 0 on the whole cell, and the cell costs one dispersion test
// instead of 25. Only cells whose probe is positive are refined: mask0 of
// their unmasked pixels, then of the unmasked neighbours of those that come
// out positive, then E by bitwise erosion. The 11x11 Poisson test runs only
// at candidates (E, unmasked, p > 0). On diffraction frames about 1.6% of
// cells need refinement and 0.4% of pixels are candidates.
//
// One block of NT threads per output tile of TH x TW pixels. The tile plus a
// 10-row / 12-column halo is staged in LDS as u16 values with masked pixels
// zeroed, plus a bitmap of masked (and off-frame) pixels. Window sums use
// v_dot2_u32_u16 (two pixels' values and squares per instruction, wrapping
// mod 2^32 as the reference's u32 accumulator does) and v_bcnt for counts.
// Survivors are emitted in row-major order per tile row run, with the same
// (row, tile) cell table and ordering kernel as the other forms.
#include "sf_device.hh"

namespace sf {

typedef unsigned short ushort2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t udot2(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_udot2(__builtin_bit_cast(ushort2v, a),
                                __builtin_bit_cast(ushort2v, b), c, false);
}

__device__ __forceinline__ int floor_div5(int v) {
  return v >= 0 ? v / 5 : -((-v + 4) / 5);
}

template <int TH_, int TW_, int NT_, int LIST_ = 512> struct Probe {
  static constexpr int TH = TH_, TW = TW_;
  static constexpr int NT = NT_;            // threads a block
  static constexpr int ROWS = TH + 20;      // global rows r0-10 .. r0+TH+9
  static constexpr int COLS = TW + 24;      // global cols c0-12 .. c0+TW+11
  static constexpr int WORDS = COLS / 32;   // bitmap words per row
  static constexpr int HALF = COLS / 2;     // u16 pairs per row
  static constexpr int PITCH = HALF + 2;    // LDS row pitch in words: rows
                                            // start on different banks
  static constexpr int NCR = (TH + 9) / 5 + 2;
  static constexpr int NCC = (TW + 9) / 5 + 2;
  static constexpr int LIST = LIST_;        // candidates per round
  static constexpr int OUTWORDS = TH * WORDS;
  static_assert(COLS % 64 == 0, "rows load as whole waves of 4-pixel groups");
  static_assert(LIST >= TW, "a round holds at least one full row");
  static_assert(NT % 64 == 0 && NT <= 1024 && LIST % NT == 0,
                "whole waves; the list spreads evenly over the block");
};

// Bits [a, a + len) of a bitmap row (len <= 25), as the low bits of a word.
__device__ __forceinline__ uint32_t row_bits(const uint32_t *row, int a,
                                             int len) {
  const int j = a >> 5;
  const uint32_t lo = row[j], hi = row[j + 1];
  return __builtin_amdgcn_alignbit(hi, lo, (uint32_t)(a & 31)) &
         ((1u << len) - 1u);
}

// Word j of a bitmap row moved by d columns: bit b of the result is column
// 32j + b + d of the row (d in -2..2; outside the row reads zero).
template <int WORDS>
__device__ __forceinline__ uint32_t moved(const uint32_t *row, int j, int d) {
  const uint32_t w = row[j];
  if (d == 0)
    return w;
  if (d > 0) {
    const uint32_t next = j + 1 < WORDS ? row[j + 1] : 0u;
    return (w >> d) | (next << (32 - d));
  }
  const uint32_t prev = j > 0 ? row[j - 1] : 0u;
  return (w << -d) | (prev >> (32 + d));
}

// mask0 of an unmasked pixel: the dispersion test over its clipped 7x7 window
// (src/dext.cc:69-111), from LDS. Off-frame pixels are in the masked bitmap
// with value 0, so they count and sum as nothing, as the clipped window does.
template <class G>
__device__ __forceinline__ bool dispersion(const uint32_t *px32,
                                           const uint32_t *mbits, int lr,
                                           int lc, const float *lut) {
  const int a = lc - 3;
  const int w = a >> 1;
  const uint32_t drop = (a & 1) ? 0xFFFF0000u : 0xFFFFFFFFu;
  const uint32_t keep = (a & 1) ? 0xFFFFFFFFu : 0x0000FFFFu;
  uint32_t masked = 0, s = 0, q = 0;
#pragma unroll
  for (int r = lr - 3; r <= lr + 3; r++) {
    masked += __builtin_popcount(row_bits(mbits + r * G::WORDS, a, 7));
    const uint32_t *p = px32 + r * G::PITCH + w;
    const uint32_t d0 = p[0] & drop, d1 = p[1], d2 = p[2], d3 = p[3] & keep;
    s = udot2(d0, 0x00010001u, s);
    s = udot2(d1, 0x00010001u, s);
    s = udot2(d2, 0x00010001u, s);
    s = udot2(d3, 0x00010001u, s);
    q = udot2(d0, d0, q);
    q = udot2(d1, d1, q);
    q = udot2(d2, d2, q);
    q = udot2(d3, d3, q);
  }
  const uint32_t n = 49u - masked;
  const float m = (float)n, fs = (float)s, fq = (float)q;
  const float lhs = m * fq - fs * fs - fs * (m - 1.0f);
  const float rhs = fs * 6.0f * lut[n];
  return lhs > rhs;
}

// Exclusive scan over the block (NT threads); *total = the block's sum.
template <int NT>
__device__ __forceinline__ uint32_t scan_block(uint32_t v, uint32_t *total,
                                               uint32_t *tmp) {
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(x, d, 64);
    if (lane >= (uint32_t)d)
      x += y;
  }
  if (lane == 63)
    tmp[wave] = x;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (int k = 0; k < NT / 64; k++) {
    before += (uint32_t)k < wave ? tmp[k] : 0u;
    all += tmp[k];
  }
  *total = all;
  // No closing barrier: every caller passes another one before it scans
  // again, so `tmp` is not overwritten while a wave still reads it.
  return before + x - v;
}

template <class G> struct ProbeShared {
  uint32_t px[G::ROWS * G::PITCH];
  uint32_t mbits[G::ROWS * G::WORDS + 1]; // masked or off-frame
  uint32_t need[G::ROWS * G::WORDS + 1];  // unmasked pixels of positive cells
  uint32_t m0[G::ROWS * G::WORDS + 1];    // mask0, where computed
  uint32_t work[G::ROWS * G::WORDS + 1];  // second-round pixels, candidates
  uint32_t ebits[G::ROWS * G::WORDS + 1]; // E at `need` pixels
  uint32_t cellmask[G::NCR * G::WORDS];   // columns of positive cells
  uint32_t word_off[G::ROWS * G::WORDS + 1];
  uint32_t list[G::LIST];
  float bg[G::LIST];
  uint32_t pop[G::LIST];
  uint32_t rank[G::LIST];
  uint32_t rowok[G::TH];
  uint32_t rowbase[G::TH];
  uint32_t rounds[G::TH + 2];
  uint32_t scalars[8];
  uint32_t tmp[16];
  float lut[64];
};

// Ranks of the set bits of bitmap rows [ra, rb) in row-major order:
// word_off[k] = bits before word k of the range; returns the total.
template <class G>
__device__ uint32_t rank_bits(ProbeShared<G> &sh, const uint32_t *bits,
                              int ra, int rb) {
  constexpr int PER = (G::ROWS * G::WORDS + G::NT - 1) / G::NT;
  const int k0 = ra * G::WORDS, n = (rb - ra) * G::WORDS;
  uint32_t mine[PER];
  uint32_t sum = 0;
#pragma unroll
  for (int i = 0; i < PER; i++) {
    const int k = (int)threadIdx.x * PER + i;
    mine[i] = k < n ? (uint32_t)__builtin_popcount(bits[k0 + k]) : 0u;
    sum += mine[i];
  }
  uint32_t total = 0;
  uint32_t at = scan_block<G::NT>(sum, &total, sh.tmp);
#pragma unroll
  for (int i = 0; i < PER; i++) {
    const int k = (int)threadIdx.x * PER + i;
    if (k < n)
      sh.word_off[k] = at;
    at += mine[i];
  }
  __syncthreads();
  return total;
}

// Lists the set bits ranked [first, first + count) as (row << 16 | column).
template <class G>
__device__ void list_bits(ProbeShared<G> &sh, const uint32_t *bits, int ra,
                          int rb, uint32_t first, uint32_t count) {
  const int n = (rb - ra) * G::WORDS;
  for (int k = threadIdx.x; k < n; k += G::NT) {
    const uint32_t off = sh.word_off[k];
    uint32_t b = bits[ra * G::WORDS + k];
    if (!b || off >= first + count ||
        off + (uint32_t)__builtin_popcount(b) <= first)
      continue;
    const int lr = ra + k / G::WORDS, j = k % G::WORDS;
    uint32_t at = off;
    while (b) {
      const int bit = __builtin_ctz(b);
      b &= b - 1;
      if (at >= first && at < first + count)
        sh.list[at - first] = ((uint32_t)lr << 16) | (uint32_t)(32 * j + bit);
      at++;
    }
  }
  __syncthreads();
}

// mask0 at every set bit of `bits` (rows [ra, rb)), or-ed into sh.m0. The
// pixels are spread evenly over the block, LIST at a time.
template <class G>
__device__ void refine(ProbeShared<G> &sh, const uint32_t *bits, int ra,
                       int rb) {
  const uint32_t total = rank_bits<G>(sh, bits, ra, rb);
  for (uint32_t first = 0; first < total; first += G::LIST) {
    const uint32_t count = min((uint32_t)G::LIST, total - first);
    list_bits<G>(sh, bits, ra, rb, first, count);
    for (uint32_t i = threadIdx.x; i < count; i += G::NT) {
      const int lr = (int)(sh.list[i] >> 16), lc = (int)(sh.list[i] & 0xFFFFu);
      if (dispersion<G>(sh.px, sh.mbits, lr, lc, sh.lut))
        atomicOr(&sh.m0[lr * G::WORDS + (lc >> 5)], 1u << (lc & 31));
    }
    __syncthreads();
  }
}

template <class G>
__device__ void probe_tile(const FrameDesc &fd, uint32_t tile_x,
                           uint32_t tile_y, ProbeShared<G> &sh) {
  typedef unsigned short us2 __attribute__((ext_vector_type(2)));
  const int height = (int)fd.height, width = (int)fd.width;
  const int r0 = (int)tile_y * G::TH, c0 = (int)tile_x * G::TW;
  const uint16_t *pix = static_cast<const uint16_t *>(fd.pixels);
  const uint32_t tid = threadIdx.x;
  constexpr int NW = G::ROWS * G::WORDS;

  // -- clear the bitmaps ----------------------------------------------------
  // (need, work and ebits are written whole by the phases that make them)
  for (int k = tid; k < NW + 1; k += G::NT)
    sh.m0[k] = 0;
  for (int k = tid; k < G::NCR * G::WORDS; k += G::NT)
    sh.cellmask[k] = 0;
  if (tid == 0)
    sh.mbits[NW] = 0;
  for (int r = tid; r < G::TH; r += G::NT)
    sh.rowok[r] = 0;

  // -- stage the tile: values with masked pixels zeroed, masked bitmap -----
  // A thread keeps one 4-pixel column group and steps down STEP rows at a
  // time; every load of the tile is issued before any is used.
  {
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    typedef __attribute__((address_space(1))) const u32x2 *global_u2;
    typedef __attribute__((address_space(1))) const uint16_t *global_u16;
    constexpr int GROUPS = G::COLS / 4;         // 32 groups a row
    constexpr int STEP = G::NT / GROUPS;        // rows per iteration
    constexpr int ITER = (G::ROWS + STEP - 1) / STEP;
    const int grp = (int)tid % GROUPS, sub = (int)tid / GROUPS;
    const int c = c0 - 12 + 4 * grp;
    uint32_t raw[ITER][2];
    const bool interior = fd.vec != 0 && r0 >= 10 &&
                          r0 + G::TH + 10 <= height && c0 >= 12 &&
                          c0 + G::TW + 12 <= width;
    if (interior) {
      // every staged row in the frame, every group 8-byte aligned
      const uint16_t *at = pix + ((size_t)(r0 - 10 + sub) * (size_t)width + c);
      const size_t stride = (size_t)STEP * (size_t)width;
#pragma unroll
      for (int it = 0; it < ITER; it++) {
        raw[it][0] = raw[it][1] = 0xFFFFFFFFu;
        if (it * STEP + sub < G::ROWS) {
          const u32x2 d = *(global_u2)(at + (size_t)it * stride);
          raw[it][0] = d.x;
          raw[it][1] = d.y;
        }
      }
    } else {
      const bool vec = fd.vec != 0 && c >= 0 && c + 4 <= width;
#pragma unroll
      for (int it = 0; it < ITER; it++) {
        const int lr = it * STEP + sub;
        const int y = r0 - 10 + lr;
        raw[it][0] = 0xFFFFFFFFu;
        raw[it][1] = 0xFFFFFFFFu;
        if (lr < G::ROWS && y >= 0 && y < height) {
          const uint16_t *row = pix + (size_t)y * (size_t)width;
          if (vec) {
            const u32x2 d = *(global_u2)(row + c);
            raw[it][0] = d.x;
            raw[it][1] = d.y;
          } else {
            uint32_t v[4] = {0xFFFFu, 0xFFFFu, 0xFFFFu, 0xFFFFu};
#pragma unroll
            for (int i = 0; i < 4; i++)
              if (c + i >= 0 && c + i < width)
                v[i] = *(global_u16)(row + c + i);
            raw[it][0] = v[0] | (v[1] << 16);
            raw[it][1] = v[2] | (v[3] << 16);
          }
        }
      }
    }
    const us2 two = {2, 2}, one = {1, 1};
    const uint32_t place = 4u * (uint32_t)(grp & 7); // nibble slot in a word
#pragma unroll
    for (int it = 0; it < ITER; it++) {
      const int lr = it * STEP + sub;
      // v + 2 wraps to 0 or 1 exactly for the two sentinels 0xFFFE, 0xFFFF
      const us2 a = __builtin_bit_cast(us2, raw[it][0]) + two;
      const us2 b = __builtin_bit_cast(us2, raw[it][1]) + two;
      const us2 pa = __builtin_elementwise_sub_sat(a, two);
      const us2 pb = __builtin_elementwise_sub_sat(b, two);
      // masked flags 1 / 0 a pixel: min(2 - a, 1) saturating
      const us2 ma = __builtin_elementwise_min(
          __builtin_elementwise_sub_sat(two, a), one);
      const us2 mb = __builtin_elementwise_min(
          __builtin_elementwise_sub_sat(two, b), one);
      // the four flags as bits 0..3 (weights 1, 2 and 4, 8)
      const uint32_t nib =
          udot2(__builtin_bit_cast(uint32_t, mb), 0x00080004u,
                udot2(__builtin_bit_cast(uint32_t, ma), 0x00020001u, 0u));
      if (lr < G::ROWS) {
        uint2 out;
        out.x = __builtin_bit_cast(uint32_t, pa);
        out.y = __builtin_bit_cast(uint32_t, pb);
        *reinterpret_cast<uint2 *>(&sh.px[lr * G::PITCH + 2 * grp]) = out;
      }
      // or eight lanes' nibbles, each in its slot, into the first lane
      uint32_t x = nib << place;
      x |= (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x101, 0xF, 0xF, true);
      x |= (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x102, 0xF, 0xF, true);
      x |= (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x104, 0xF, 0xF, true);
      if (lr < G::ROWS && (grp & 7) == 0)
        sh.mbits[lr * G::WORDS + (grp >> 3)] = x;
    }
  }
  __syncthreads();

  // -- probes: one dispersion test per 5x5 cell ------------------------------
  const int ci0 = floor_div5(r0 - 5), ci1 = floor_div5(r0 + G::TH + 4);
  const int cj0 = floor_div5(c0 - 5), cj1 = floor_div5(c0 + G::TW + 4);
  const int ncr = ci1 - ci0 + 1, ncc = cj1 - cj0 + 1;
  for (int k = tid; k < ncr * ncc; k += G::NT) {
    const int ci = ci0 + k / ncc, cj = cj0 + k % ncc;
    const int plr = 5 * ci + 2 - r0 + 10, plc = 5 * cj + 2 - c0 + 12;
    const bool masked =
        (sh.mbits[plr * G::WORDS + (plc >> 5)] >> (plc & 31)) & 1u;
    const bool pos =
        masked || dispersion<G>(sh.px, sh.mbits, plr, plc, sh.lut);
    if (pos) {
      // the cell's five columns, in its cell row's column mask
      const int lc0 = 5 * cj - c0 + 12; // 3 <= lc0 and lc0 + 4 < COLS
      uint32_t *row = sh.cellmask + (ci - ci0) * G::WORDS;
      const int j = lc0 >> 5, off = lc0 & 31;
      atomicOr(&row[j], 0x1Fu << off);
      if (off > 27)
        atomicOr(&row[j + 1], 0x1Fu >> (32 - off));
    }
  }
  __syncthreads();

  // -- the unmasked pixels of positive cells, inside the E region ----------
  // E region: rows [5, TH+15), columns [7, TW+17) of the tile's LDS frame.
  int any_need = 0;
  for (int k = tid; k < NW; k += G::NT) {
    const int lr = k / G::WORDS, j = k % G::WORDS;
    if (lr < 5 || lr >= G::TH + 15) {
      sh.need[k] = 0;
      continue;
    }
    const int ci = floor_div5(r0 - 10 + lr);
    const int lo = max(7 - 32 * j, 0), hi = min(G::TW + 17 - 32 * j, 32);
    const uint32_t cols =
        hi <= lo ? 0u
                 : ((hi - lo == 32 ? 0xFFFFFFFFu : ((1u << (hi - lo)) - 1u))
                    << lo);
    sh.need[k] = sh.cellmask[(ci - ci0) * G::WORDS + j] & cols & ~sh.mbits[k];
    any_need |= sh.need[k] != 0;
  }
  // No unmasked pixel of a positive cell: E is 0 at every unmasked pixel, so
  // nothing in the tile can be signal (about half the tiles of a diffraction
  // frame). Record empty rows and stop.
  const auto no_signal = [&] {
    const int rows = min(G::TH, height - r0);
    for (int r = tid; r < rows; r += G::NT)
      fd.cell_count[(size_t)(r0 + r) * fd.n_strips + tile_x] = 0;
    __syncthreads();
  };
  if (!__syncthreads_or(any_need)) {
    no_signal();
    return;
  }

  // -- refine, round one: mask0 at those pixels ------------------------------
  refine<G>(sh, sh.need, 5, G::TH + 15);
  int any_m0 = 0;
  for (int k = tid; k < NW; k += G::NT)
    any_m0 |= sh.m0[k] != 0;
  // Every refined pixel negative: E(x) needs mask0(x) = 1 at the unmasked x
  // itself, so E is 0 on the tile (typical beside module gaps).
  if (!__syncthreads_or(any_m0)) {
    no_signal();
    return;
  }

  // -- round two: the unmasked neighbours (5x5) of positive ones ------------
  // Separable dilation of mask0: 5 columns wide into ebits (scratch until the
  // erosion), then 5 rows tall over the mask0 region, rows [3, TH+17).
  for (int k = tid; k < NW; k += G::NT) {
    const int lr = k / G::WORDS, j = k % G::WORDS;
    uint32_t h = 0;
    if (lr >= 1 && lr < G::TH + 19) {
      const uint32_t *row = sh.m0 + lr * G::WORDS;
      h = moved<G::WORDS>(row, j, -2) | moved<G::WORDS>(row, j, -1) | row[j] |
          moved<G::WORDS>(row, j, 1) | moved<G::WORDS>(row, j, 2);
    }
    sh.ebits[k] = h;
  }
  __syncthreads();
  for (int k = tid; k < NW; k += G::NT) {
    const int lr = k / G::WORDS;
    uint32_t grown = 0;
    if (lr >= 3 && lr < G::TH + 17) {
#pragma unroll
      for (int r = -2; r <= 2; r++)
        grown |= sh.ebits[k + r * G::WORDS];
      grown &= ~sh.mbits[k] & ~sh.need[k];
    }
    sh.work[k] = grown;
  }
  __syncthreads();
  refine<G>(sh, sh.work, 3, G::TH + 17);

  // -- E: 5x5 erosion of (mask0 or masked), kept at the refined pixels ------
  // Separable: 5 columns wide into work over rows [3, TH+17), then 5 rows
  // tall over the E region, rows [5, TH+15). Off-frame pixels are masked, so
  // they count as set, the reference's erode edge (src/dext.cc:117-142).
  for (int k = tid; k < NW; k += G::NT) {
    const int lr = k / G::WORDS, j = k % G::WORDS;
    uint32_t h = 0;
    if (lr >= 3 && lr < G::TH + 17) {
      const uint32_t *m0 = sh.m0 + lr * G::WORDS;
      const uint32_t *mb = sh.mbits + lr * G::WORDS;
      h = 0xFFFFFFFFu;
#pragma unroll
      for (int d = -2; d <= 2; d++)
        h &= moved<G::WORDS>(m0, j, d) | moved<G::WORDS>(mb, j, d);
    }
    sh.work[k] = h;
  }
  __syncthreads();
  for (int k = tid; k < NW; k += G::NT) {
    const int lr = k / G::WORDS;
    uint32_t e = 0;
    const uint32_t need = sh.need[k];
    if (need && lr >= 5 && lr < G::TH + 15) {
      e = need;
#pragma unroll
      for (int r = -2; r <= 2; r++)
        e &= sh.work[k + r * G::WORDS];
    }
    sh.ebits[k] = e;
  }
  __syncthreads();

  // -- candidates: E, unmasked, p > 0, inside the output tile ---------------
  // output: rows [10, TH+10), columns [12, TW+12). The same pass zeroes the
  // eroded-set pixels in LDS: the 11x11 sums leave out masked and
  // eroded-set pixels, masked ones are already 0, so each window sum becomes
  // a plain sum (candidates' own values are read from the frame). A word's
  // candidates are tested before its pixels are zeroed, by the same thread.
  {
    uint16_t *px16 = reinterpret_cast<uint16_t *>(sh.px);
    for (int k = tid; k < NW; k += G::NT) {
      const int lr = k / G::WORDS, j = k % G::WORDS;
      uint32_t bits = sh.ebits[k]; // set only in the E region
      if (lr >= 10 && lr < G::TH + 10) {
        uint32_t cand = 0, x = bits;
        while (x) {
          const int b = __builtin_ctz(x);
          x &= x - 1;
          const int lc = 32 * j + b;
          if (lc < 12 || lc >= G::TW + 12)
            continue;
          if (px16[lr * 2 * G::PITCH + lc])
            cand |= 1u << b;
        }
        sh.work[k] = cand;
      }
      while (bits) {
        const int b = __builtin_ctz(bits);
        bits &= bits - 1;
        px16[lr * 2 * G::PITCH + 32 * j + b] = 0;
      }
    }
  }
  __syncthreads();

  // Row-major ranks of the candidates, and rounds of whole rows each at most
  // LIST candidates.
  const uint32_t total = rank_bits<G>(sh, sh.work, 10, 10 + G::TH);
  if (total == 0) { // uniform: no candidate in the tile
    no_signal();
    return;
  }
  if (tid == 0) {
    int n = 0;
    sh.rounds[n++] = 0;
    if (total > (uint32_t)G::LIST) { // rare: split into rounds of whole rows
      uint32_t start = 0;
      for (int r = 0; r < G::TH; r++) {
        const uint32_t end =
            r + 1 < G::TH ? sh.word_off[(r + 1) * G::WORDS] : total;
        if (end - start > (uint32_t)G::LIST) {
          sh.rounds[n++] = (uint32_t)r;
          start = sh.word_off[r * G::WORDS];
        }
      }
    }
    sh.rounds[n] = G::TH;
    sh.scalars[0] = (uint32_t)n; // number of rounds
    sh.scalars[1] = 0;           // survivors so far
  }
  __syncthreads();

  const uint32_t n_rounds = sh.scalars[0];
  const int out_rows = min(G::TH, height - r0);
  for (uint32_t round = 0; round < n_rounds; round++) {
    const int ra = (int)sh.rounds[round], rb = (int)sh.rounds[round + 1];
    const uint32_t first = sh.word_off[ra * G::WORDS];
    const uint32_t last = rb < G::TH ? sh.word_off[rb * G::WORDS] : total;
    const uint32_t count = last - first;
    list_bits<G>(sh, sh.work + 10 * G::WORDS, 0, G::TH, first, count);
    // the Poisson test over the clipped 11x11 window of pixels neither
    // masked nor eroded-set (src/dext.cc:165-207); eight lanes a candidate,
    // lane g summing rows g and g + 8 of the window, then added across the
    // eight (integer sums, so the order is immaterial)
    const uint32_t g = tid & 7u;
    typedef __attribute__((address_space(1))) const uint16_t *global_u16;
    const uint16_t *pix = static_cast<const uint16_t *>(fd.pixels);
    for (uint32_t at0 = 0; at0 < count; at0 += G::NT / 8) {
      const uint32_t i = at0 + (tid >> 3);
      const bool active = i < count;
      uint32_t n = 0, s2 = 0, p = 0;
      if (active) {
        const int lr = 10 + (int)(sh.list[i] >> 16);
        const int lc = (int)(sh.list[i] & 0xFFFFu);
        if (g == 0) // the candidate's own value: eroded-set, so 0 in LDS
          p = *(global_u16)(pix + (size_t)(r0 + lr - 10) * (size_t)width +
                            (size_t)(c0 + lc - 12));
        const int a = lc - 5;
        const int w = a >> 1;
        const uint32_t drop = (a & 1) ? 0xFFFF0000u : 0xFFFFFFFFu;
        const uint32_t keep = (a & 1) ? 0xFFFFFFFFu : 0x0000FFFFu;
#pragma unroll
        for (int t = 0; t < 2; t++) {
          const int row = (int)g + 8 * t;
          if (row < 11) {
            const int r = lr - 5 + row;
            n += 11u - __builtin_popcount(
                           row_bits(sh.mbits + r * G::WORDS, a, 11) |
                           row_bits(sh.ebits + r * G::WORDS, a, 11));
            const uint32_t *q = sh.px + r * G::PITCH + w;
            s2 = udot2(q[0] & drop, 0x00010001u, s2);
            s2 = udot2(q[1], 0x00010001u, s2);
            s2 = udot2(q[2], 0x00010001u, s2);
            s2 = udot2(q[3], 0x00010001u, s2);
            s2 = udot2(q[4], 0x00010001u, s2);
            s2 = udot2(q[5] & keep, 0x00010001u, s2);
          }
        }
      }
#pragma unroll
      for (int d = 1; d < 8; d <<= 1) {
        n += __shfl_xor(n, d, 64);
        s2 += __shfl_xor(s2, d, 64);
      }
      if (active && g == 0) {
        const float mean = n >= 2 ? (float)s2 / (float)n : 0.0f;
        const bool ok = (float)p >= mean + 3.0f * sqrtf(mean);
        sh.rank[i] = ok ? 1u : 0u;
        sh.bg[i] = mean;
        sh.pop[i] = n | (p << 16);
      }
    }
    __syncthreads();
    // survivors in list order
    constexpr int LPER = G::LIST / G::NT;
    uint32_t flags[LPER];
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < LPER; i++) {
      const uint32_t e = tid * LPER + i;
      flags[i] = e < count ? sh.rank[e] : 0u;
      sum += flags[i];
    }
    uint32_t survivors = 0;
    uint32_t at = scan_block<G::NT>(sum, &survivors, sh.tmp);
    if (tid == 0)
      sh.scalars[2] = survivors ? atomicAdd(fd.counter, survivors) : 0u;
    // each entry's exclusive survivor rank, for the row offsets below
#pragma unroll
    for (int i = 0, r = (int)at; i < LPER; i++) {
      const uint32_t e = tid * LPER + i;
      if (e < count)
        sh.rank[e] = (uint32_t)r;
      r += (int)flags[i];
    }
    __syncthreads();
    const uint32_t base = sh.scalars[2];
#pragma unroll
    for (int i = 0; i < LPER; i++) {
      const uint32_t e = tid * LPER + i;
      if (!flags[i])
        continue;
      const int lr = 10 + (int)(sh.list[e] >> 16);
      const int lc = (int)(sh.list[e] & 0xFFFFu);
      atomicAdd(&sh.rowok[lr - 10], 1u);
      const uint32_t dst = base + at;
      at++;
      if (dst < fd.capacity) {
        sf_signal_pixel px;
        px.index = (uint32_t)(r0 + lr - 10) * (uint32_t)width +
                   (uint32_t)(c0 + lc - 12);
        px.value = sh.pop[e] >> 16;
        px.background = sh.bg[e];
        px.population = (uint16_t)(sh.pop[e] & 0xFFFFu);
        px.reserved = 0;
        fd.staging[dst] = px;
      }
    }
    // a row's survivors start at the survivor rank of its first candidate
    for (int r = ra + (int)tid; r < rb; r += G::NT) {
      const uint32_t fi = sh.word_off[r * G::WORDS] - first;
      if (fi < count)
        sh.rowbase[r] = base + sh.rank[fi];
    }
    if (tid == 0)
      sh.scalars[1] += survivors;
    __syncthreads();
  }
  for (int r = tid; r < out_rows; r += G::NT) {
    const size_t c = (size_t)(r0 + r) * fd.n_strips + tile_x;
    fd.cell_count[c] = (uint8_t)sh.rowok[r];
    if (sh.rowok[r])
      fd.cell_base[c] = sh.rowbase[r];
  }
  if (tid == 0 && sh.scalars[1])
    atomicAdd(&fd.band_total[tile_y], sh.scalars[1]);
  __syncthreads();
}

template <class G>
__global__ __launch_bounds__(G::NT, 4) void probe_kernel(
    const FrameDesc *frames, int n_frames, uint32_t items) {
  __shared__ ProbeShared<G> sh;
  if (threadIdx.x < 64)
    sh.lut[threadIdx.x] = threadIdx.x < 2
                              ? __builtin_inff()
                              : sqrtf(2.0f * (float)(threadIdx.x - 1));
  __syncthreads();
  const uint32_t padded = (items + 7) & ~7u;
  const uint32_t blocks = gridDim.x;
  for (uint32_t b = blockIdx.x; b < padded; b += blocks) {
    // each XCD takes a contiguous run of tiles (blocks and padded are
    // multiples of 8, so a block stays on its XCD's run)
    const uint32_t item = xcd_local_item(b, padded);
    if (item >= items)
      continue;
    const int f = frame_of_item(frames, n_frames, item);
    const FrameDesc &fd = frames[f];
    const uint32_t local = item - fd.item_begin;
    probe_tile<G>(fd, local % fd.n_strips, local / fd.n_strips, sh);
  }
}

// The tile shapes the autotuner chooses from: 64 x 104 with 256 threads
// (four blocks a CU), 128 x 104 or 64 x 232 with 512 threads (two blocks a
// CU, about 12% less halo to stage), and 32 x 104 with 128 threads (seven
// blocks a CU, more halo). On 4362 x 4148 frames the first is fastest.
using ProbeShape0 = Probe<64, 104, 256>;
using ProbeShape1 = Probe<128, 104, 512>;
using ProbeShape2 = Probe<64, 232, 512>;
using ProbeShape3 = Probe<32, 104, 128, 128>;

int probe_shapes() { return 4; }
int probe_tile_width(int shape) {
  return shape == 1   ? ProbeShape1::TW
         : shape == 2 ? ProbeShape2::TW
         : shape == 3 ? ProbeShape3::TW
                      : ProbeShape0::TW;
}
int probe_tile_height(int shape) {
  return shape == 1   ? ProbeShape1::TH
         : shape == 2 ? ProbeShape2::TH
         : shape == 3 ? ProbeShape3::TH
                      : ProbeShape0::TH;
}

hipError_t launch_probe(int shape, const FrameDesc *frames, int n_frames,
                        uint32_t items, int blocks, hipStream_t stream) {
  const uint32_t padded = (items + 7) & ~7u;
  const uint32_t asked = ((uint32_t)blocks + 7) & ~7u; // whole XCD rounds
  const uint32_t grid = blocks > 0 ? (asked < padded ? asked : padded) : padded;
  if (shape == 1)
    probe_kernel<ProbeShape1><<<grid, ProbeShape1::NT, 0, stream>>>(
        frames, n_frames, items);
  else if (shape == 2)
    probe_kernel<ProbeShape2><<<grid, ProbeShape2::NT, 0, stream>>>(
        frames, n_frames, items);
  else if (shape == 3)
    probe_kernel<ProbeShape3><<<grid, ProbeShape3::NT, 0, stream>>>(
        frames, n_frames, items);
  else
    probe_kernel<ProbeShape0><<<grid, ProbeShape0::NT, 0, stream>>>(
        frames, n_frames, items);
  return hipGetLastError();
}

} // namespace sf
