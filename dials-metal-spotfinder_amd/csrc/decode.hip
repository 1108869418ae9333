// HDF5 chunks decoded into device frames: decompress::image
// (src/decompress.cc:76-136) for a batch of chunks in one launch.
//
// The host walks each chunk's block lengths (a dependent chain of 4-byte
// reads that is cheap on the host and serial on a device) and writes a block
// table: where each block's LZ4 bytes sit in its chunk and where its decoded
// bytes go in its frame. The device decodes in two passes. The parse pass
// gives each bitshuffle block a thread, which walks its LZ4 sequences and
// records them (a wave per block would spend that walk on the scalar unit,
// which then bounds the decode). The copy pass gives each block a wave: it
// copies the block's LZ4 bytes into LDS with 16-byte loads, replays the
// recorded sequences there (literal and match bytes copied 64 lanes at a
// time), and writes the frame's elements from the bit planes, each lane 8
// elements at a time by 8x8 bit transposes in registers. The bit planes
// never reach HBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/decode_hip.h"

namespace {

thread_local std::string g_decode_error;

int fail(int status, std::string what) {
  g_decode_error = std::move(what);
  return status;
}

constexpr int kWave = 64;
constexpr size_t kHeader = 12;       // src/decompress.cc:14-16
constexpr uint32_t kMaxShuffled = 16384; // LDS holds a block and its LZ4 bytes
constexpr int kMinMatch = 4, kLastLiterals = 5, kMfLimit = 12;

enum Kind : uint32_t { kShuffled = 0, kLz4 = 1, kRaw = 2 };

// One block of the batch. `src` is the byte offset of its LZ4 bytes (after
// the 4-byte length) or raw bytes in chunk `chunk`; `dst` the byte offset of
// its decoded bytes in frame `chunk`.
struct Block {
  uint64_t src;
  uint64_t dst;
  uint32_t comp;  // LZ4 bytes (0 for a raw copy)
  uint32_t bytes; // decoded bytes
  uint32_t chunk;
  uint32_t room;  // plain LZ4: the output room the block is decoded with
                  // (its bytes may be fewer); otherwise equal to bytes
};
static_assert(sizeof(Block) == 32, "block table entry");

uint32_t be32(const uint8_t *p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) |
         ((uint32_t)p[2] << 8) | p[3];
}

size_t lz4_bound(size_t n) { return n + n / 255 + 16; }

// The bytes one LZ4 block decodes to with `room` bytes of output, under the
// rules wave_lz4 decodes by (the published block format, with
// LZ4_decompress_safe's end-of-block conditions): its sequences' lengths
// walked on the host, nothing copied; -1 when the block is malformed.
long long lz4_length(const uint8_t *src, size_t n, size_t room) {
  size_t i = 0, o = 0;
  for (;;) {
    if (i >= n)
      return -1;
    const unsigned token = src[i++];
    size_t lit = token >> 4;
    if (lit == 15) {
      unsigned b;
      do {
        if (i >= n)
          return -1;
        b = src[i++];
        lit += b;
      } while (b == 255);
    }
    if (o + lit + kMfLimit > room || i + lit + 8 > n) // the last sequence
      return i + lit == n && o + lit <= room ? (long long)(o + lit) : -1;
    i += lit;
    o += lit;
    const size_t off = (size_t)src[i] | ((size_t)src[i + 1] << 8);
    i += 2;
    if (off == 0 || off > o)
      return -1;
    size_t len = token & 15;
    if (len == 15) {
      unsigned b;
      do {
        if (i >= n)
          return -1;
        b = src[i++];
        len += b;
      } while (b == 255);
    }
    len += kMinMatch;
    if (o + len + kLastLiterals > room)
      return -1;
    o += len;
  }
}

// bitshuffle's default block: 8192 bytes of elements, a multiple of 8, at
// least 128 elements.
size_t default_block(size_t elem_size) {
  const size_t block = 8192 / elem_size / 8 * 8;
  return std::max<size_t>(block, 128);
}

// ---------------------------------------------------------------- device

__device__ __forceinline__ uint32_t uniform(uint32_t v) {
  return __builtin_amdgcn_readfirstlane(v);
}

// The 8x8 bit matrix in x (byte r, bit c) transposed: bit 8r + c moves to
// bit 8c + r. Byte k of the input holds bit plane k for 8 elements; byte i
// of the output is element i's byte.
__device__ __forceinline__ uint64_t transpose8(uint64_t x) {
  uint64_t t = (x ^ (x >> 7)) & 0x00AA00AA00AA00AAull;
  x ^= t ^ (t << 7);
  t = (x ^ (x >> 14)) & 0x0000CCCC0000CCCCull;
  x ^= t ^ (t << 14);
  t = (x ^ (x >> 28)) & 0x00000000F0F0F0F0ull;
  x ^= t ^ (t << 28);
  return x;
}

// len bytes from src to dst, 64 lanes at a time (four rounds' loads issued
// before their stores while 256 bytes remain). No byte read is one this
// copy writes.
template <typename I>
__device__ __forceinline__ void wave_copy(uint8_t *dst, const uint8_t *src,
                                          I len, int lane) {
  I base = 0;
  for (; base + 4 * kWave <= len; base += 4 * kWave) {
    uint8_t v[4];
#pragma unroll
    for (int u = 0; u < 4; u++)
      v[u] = src[base + u * kWave + lane];
#pragma unroll
    for (int u = 0; u < 4; u++)
      dst[base + u * kWave + lane] = v[u];
  }
  for (; base < len; base += kWave) {
    const I k = base + lane;
    if (k < len)
      dst[k] = src[k];
  }
}

// A match of len bytes at distance off < len: the pattern of the off bytes
// before dst repeats, dst[k] = dst[k % off - off]. Every byte read lies
// before dst, so no byte read is one this copy writes.
template <typename I>
__device__ __forceinline__ void wave_repeat(uint8_t *dst, I off, I len,
                                            int lane) {
  const uint8_t *pattern = dst - off;
  if (off == 1) {
    const uint8_t v = pattern[0];
    for (I k = lane; k < len; k += kWave)
      dst[k] = v;
    return;
  }
  I r = (I)lane % off;
  const I step = (I)kWave % off;
  for (I base = 0; base < len; base += kWave) {
    const I k = base + lane;
    if (k < len)
      dst[k] = pattern[r];
    r += step;
    if (r >= off)
      r -= off;
  }
}

// Byte q of src: from the window (byte w0 + lane of src in each lane) when
// it holds it, else read by every lane from src.
template <typename I>
__device__ __forceinline__ uint32_t byte_at(const uint8_t *src, uint32_t win,
                                            I w0, I q) {
  if (q - w0 < kWave)
    return __builtin_amdgcn_readlane(win, (int)(q - w0));
  return uniform(src[q]);
}

// One LZ4 block (the published block format, with LZ4_decompress_safe's
// end-of-block rules) from src[0, n) into dst[0, cap) by the whole wave:
// the bytes produced, or -1 when the block is malformed. Each sequence
// starts with one read of the 64 bytes at its token, a byte a lane: the
// token, lengths and offset come from it by readlane (uniform, so the parse
// runs on the scalar unit), and literals that lie in it are written from
// it. A match waits for the wave's stores only when it reads bytes written
// since the last wait.
template <typename I>
__device__ I wave_lz4(const uint8_t *src, I n, uint8_t *dst, I cap,
                      int lane) {
  I i = 0, o = 0, synced = 0;
  for (;;) {
    if (i >= n)
      return -1;
    const I w0 = i;
    const uint32_t win = w0 + lane < n ? (uint32_t)src[w0 + lane] : 0u;
    const uint32_t token = __builtin_amdgcn_readlane(win, 0);
    i++;
    I lit = token >> 4;
    if (lit == 15) {
      uint32_t b;
      do {
        if (i >= n)
          return -1;
        b = byte_at<I>(src, win, w0, i);
        i++;
        lit += b;
      } while (b == 255);
    }
    const bool last = o + lit + kMfLimit > cap || i + lit + 8 > n;
    if (last && (i + lit != n || o + lit > cap))
      return -1;
    if (i + lit <= w0 + kWave) {
      const I k = (I)lane - (i - w0); // window lanes [i - w0, +lit)
      if (k >= 0 && k < lit)
        dst[o + k] = (uint8_t)win;
    } else {
      wave_copy<I>(dst + o, src + i, lit, lane);
    }
    i += lit;
    o += lit;
    if (last)
      return o;
    const I off = byte_at<I>(src, win, w0, i) |
                  (byte_at<I>(src, win, w0, i + 1) << 8);
    i += 2;
    if (off == 0 || off > o)
      return -1;
    I len = token & 15;
    if (len == 15) {
      uint32_t b;
      do {
        if (i >= n)
          return -1;
        b = byte_at<I>(src, win, w0, i);
        i++;
        len += b;
      } while (b == 255);
    }
    len += kMinMatch;
    if (o + len + kLastLiterals > cap)
      return -1;
    if (o - off + (off < len ? off : len) > synced) {
      __syncthreads(); // the match reads bytes the lanes have just written
      synced = o;
    }
    if (off >= len)
      wave_copy<I>(dst + o, dst + o - off, len, lane);
    else
      wave_repeat<I>(dst + o, off, len, lane);
    o += len;
  }
}

// A wave's LDS operations execute in order: once the compiler keeps them in
// order too, its lanes see each other's LDS stores.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// k mod m for 0 <= k, m < 2^16, m > 0, on the vector unit: the float
// quotient is within one of the true one, and one correction fixes it.
__device__ __forceinline__ int mod_small(int k, int m, float inv) {
  int r = k - (int)((float)k * inv) * m;
  r = r < 0 ? r + m : r;
  return r >= m ? r - m : r;
}

// len bytes from src to dst within LDS, no byte read being one this copy
// writes. Idle lanes store to their own byte of `trash` rather than leave
// the wave's exec mask (no scalar mask work); reads may run up to 63 bytes
// past src + len.
__device__ __forceinline__ void lds_copy(uint8_t *dst, const uint8_t *src,
                                         int len, int lane, uint8_t *trash) {
  int base = 0;
  for (; base + 4 * kWave <= len; base += 4 * kWave) {
    uint8_t v[4];
#pragma unroll
    for (int u = 0; u < 4; u++)
      v[u] = src[base + u * kWave + lane];
#pragma unroll
    for (int u = 0; u < 4; u++)
      dst[base + u * kWave + lane] = v[u];
  }
  for (; base < len; base += kWave) {
    const int k = base + lane;
    const uint8_t v = src[k];
    *(k < len ? dst + k : trash) = v;
  }
}

// One LZ4 block held in LDS (src[0, n), readable to n + 63) into LDS
// (dst[0, cap), with 64 bytes after it for idle lanes): the same parse and
// checks as wave_lz4, shaped for the scalar unit, which bounds this kernel:
// no guard on the window read, no exec-mask branches around copies, k mod
// off on the vector unit.
__device__ int lds_lz4(const uint8_t *src, int n, uint8_t *dst, int cap,
                       int lane) {
  uint8_t *trash = dst + cap + lane;
  int i = 0, o = 0;
  for (;;) {
    const uint32_t win = src[i + lane];
    const uint32_t token = __builtin_amdgcn_readlane(win, 0);
    int p = i + 1;
    int lit = token >> 4;
    if (lit == 15) {
      uint32_t b;
      do {
        if (p >= n)
          return -1;
        b = byte_at<int>(src, win, i, p);
        p++;
        lit += b;
      } while (b == 255);
    }
    const int rem_out = cap - o;
    if (lit + kMfLimit > rem_out || lit + 8 > n - p) {
      if (lit != n - p || lit > rem_out)
        return -1;
      lds_copy(dst + o, src + p, lit, lane, trash);
      return o + lit;
    }
    if (p + lit <= i + kWave) {
      const int k = lane - (p - i); // the literals are window lanes
      *((unsigned)k < (unsigned)lit ? dst + o + k : trash) = (uint8_t)win;
    } else {
      lds_copy(dst + o, src + p, lit, lane, trash);
    }
    o += lit;
    p += lit;
    const int off = (int)(byte_at<int>(src, win, i, p) |
                          (byte_at<int>(src, win, i, p + 1) << 8));
    p += 2;
    int len = token & 15;
    if (len == 15) {
      uint32_t b;
      do {
        if (p >= n)
          return -1;
        b = byte_at<int>(src, win, i, p);
        p++;
        len += b;
      } while (b == 255);
    }
    len += kMinMatch;
    if (off == 0 || off > o || len + kLastLiterals > cap - o)
      return -1;
    wave_sync(); // the match may read bytes the lanes have just written
    uint8_t *at = dst + o;
    if (off >= len) {
      lds_copy(at, at - off, len, lane, trash);
    } else if (off == 1) {
      const uint8_t v = at[-1];
      for (int base = 0; base < len; base += kWave) {
        const int k = base + lane;
        *(k < len ? at + k : trash) = v;
      }
    } else {
      const float inv = 1.0f / (float)off;
      for (int base = 0; base < len; base += kWave) {
        const int k = base + lane;
        const uint8_t v = at[mod_small(k, off, inv) - off];
        *(k < len ? at + k : trash) = v;
      }
    }
    o += len;
    i = p;
  }
}

// Elements of ES bytes from the bit planes of one block (groups = elements
// / 8, plane p at planes + p * groups), each lane 8 elements at a time.
template <int ES>
__device__ __forceinline__ void unshuffle(const uint8_t *planes, int groups,
                                          uint8_t *out, bool aligned,
                                          int lane) {
  for (int g = lane; g < groups; g += kWave) {
    uint64_t x[ES];
#pragma unroll
    for (int j = 0; j < ES; j++) {
      uint64_t v = 0;
#pragma unroll
      for (int k = 0; k < 8; k++)
        v |= (uint64_t)planes[(8 * j + k) * groups + g] << (8 * k);
      x[j] = transpose8(v);
    }
    // element i's byte j is byte i of x[j]
    uint8_t *at = out + (size_t)g * 8 * ES;
    if (ES == 1) {
      if (aligned)
        *reinterpret_cast<uint64_t *>(at) = x[0];
      else
        for (int i = 0; i < 8; i++)
          at[i] = (uint8_t)(x[0] >> (8 * i));
    } else {
      uint32_t w[2 * ES]; // 8 elements of ES bytes as 32-bit words
#pragma unroll
      for (int q = 0; q < 2 * ES; q++) {
        uint32_t word = 0;
#pragma unroll
        for (int b = 0; b < 4; b++) {
          const int byte = 4 * q + b; // output byte: element byte / ES
          const int i = byte / ES, j = byte % ES;
          word |= (uint32_t)((x[j] >> (8 * i)) & 0xFF) << (8 * b);
        }
        w[q] = word;
      }
      if (aligned) {
        uint4 *v = reinterpret_cast<uint4 *>(at);
#pragma unroll
        for (int q = 0; q < 2 * ES; q += 4)
          v[q / 4] = make_uint4(w[q], w[q + 1], w[q + 2], w[q + 3]);
      } else {
        for (int q = 0; q < 2 * ES; q++)
          for (int b = 0; b < 4; b++)
            at[4 * q + b] = (uint8_t)(w[q] >> (8 * b));
      }
    }
  }
}

__global__ __launch_bounds__(64) void zero_status(int32_t *status, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x)
    status[i] = 0;
}

// Sequences recorded a block by the parse pass; a block with more is
// parsed again by its wave in LDS.
constexpr int kMaxSeq = 256;

// One LZ4 sequence: where its literals start in the block's LZ4 bytes, how
// many, the match offset and length (0: the last sequence, literals only).
__device__ __forceinline__ uint64_t pack_seq(int at, int lit, int off,
                                             int len) {
  return (uint64_t)(uint32_t)at | ((uint64_t)(uint32_t)lit << 16) |
         ((uint64_t)(uint32_t)off << 32) | ((uint64_t)(uint32_t)len << 48);
}

// The parse pass: a thread a bitshuffle block walks its LZ4 sequences on
// the vector unit, with the checks of lds_lz4, and records them. n_seq[b]
// is the count, -1 for a block that does not decode (its chunk's status
// -91), -2 for one with more than kMaxSeq sequences. Walking 64 blocks at
// once spreads the parse's per-sequence work over a wave's lanes, where a
// wave per block spends it on the scalar unit. One wave a workgroup, so
// that the few waves spread evenly over the CUs.
__global__ __launch_bounds__(64) void parse_kernel(
    const Block *blocks, uint32_t n_blocks, const uint8_t *const *chunks,
    uint64_t *seqs, int32_t *n_seq, int32_t *status) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n_blocks)
    return;
  const Block blk = blocks[b];
  const uint8_t *src = chunks[blk.chunk] + blk.src;
  const int n = (int)blk.comp, cap = (int)blk.bytes;
  uint64_t *out = seqs + (size_t)b * kMaxSeq;
  int i = 0, o = 0, k = 0, result = -1;
  while (i < n) {
    const uint32_t token = src[i];
    int p = i + 1;
    int lit = token >> 4;
    bool bad = false;
    if (lit == 15) {
      uint32_t v = 255;
      while (v == 255) {
        if (p >= n) {
          bad = true;
          break;
        }
        v = src[p++];
        lit += v;
      }
    }
    if (bad)
      break;
    if (lit + kMfLimit > cap - o || lit + 8 > n - p) {
      if (lit == n - p && lit <= cap - o) {
        if (k < kMaxSeq)
          out[k] = pack_seq(p, lit, 0, 0);
        k++;
        result = o + lit;
      }
      break;
    }
    const int at = p;
    p += lit;
    o += lit;
    const int off = (int)src[p] | ((int)src[p + 1] << 8);
    p += 2;
    int len = token & 15;
    if (len == 15) {
      uint32_t v = 255;
      while (v == 255) {
        if (p >= n) {
          bad = true;
          break;
        }
        v = src[p++];
        len += v;
      }
    }
    len += kMinMatch;
    if (bad || off == 0 || off > o || len + kLastLiterals > cap - o)
      break;
    if (k < kMaxSeq)
      out[k] = pack_seq(at, lit, off, len);
    k++;
    o += len;
    i = p;
  }
  if (result != cap) {
    n_seq[b] = -1;
    status[blk.chunk] = -91;
  } else {
    n_seq[b] = k <= kMaxSeq ? k : -2;
  }
}

// The copies of one recorded sequence list into LDS: literals from the
// staged LZ4 bytes, matches from the bytes before them.
__device__ __forceinline__ void replay_seq(uint64_t e, const uint8_t *src,
                                           uint8_t *dst, int &o,
                                           uint8_t *trash, int lane) {
  const int at = (int)(e & 0xFFFF), lit = (int)((e >> 16) & 0xFFFF);
  const int off = (int)((e >> 32) & 0xFFFF), len = (int)(e >> 48);
  if (lit <= kWave) {
    const uint8_t v = src[at + lane];
    *(lane < lit ? dst + o + lane : trash) = v;
  } else {
    lds_copy(dst + o, src + at, lit, lane, trash);
  }
  o += lit;
  if (len == 0)
    return;
  wave_sync(); // the match may read bytes the lanes have just written
  uint8_t *to = dst + o;
  if (off >= len) {
    if (len <= kWave) {
      const uint8_t v = to[lane - off];
      *(lane < len ? to + lane : trash) = v;
    } else {
      lds_copy(to, to - off, len, lane, trash);
    }
  } else if (off == 1) {
    const uint8_t v = to[-1];
    for (int base = 0; base < len; base += kWave) {
      const int k = base + lane;
      *(k < len ? to + k : trash) = v;
    }
  } else {
    const float inv = 1.0f / (float)off;
    for (int base = 0; base < len; base += kWave) {
      const int k = base + lane;
      const uint8_t v = to[mod_small(k, off, inv) - off];
      *(k < len ? to + k : trash) = v;
    }
  }
  o += len;
}

// One bitshuffle block a workgroup of one wave. LDS: `stage` bytes for the
// LZ4 bytes (from the 16-byte boundary below them, with 64 bytes to spare
// for the window reads), then the bit planes and 64 bytes for idle lanes.
template <int ES>
__global__ __launch_bounds__(64) void shuffled_kernel(
    const Block *blocks, const uint8_t *const *chunks, uint8_t *const *frames,
    uint32_t stage, const uint64_t *seqs, const int32_t *n_seq,
    int32_t *status) {
  extern __shared__ __align__(16) uint8_t lds[];
  const int lane = threadIdx.x;
  const int ns = n_seq[blockIdx.x];
  if (ns == -1)
    return; // the parse pass found it does not decode
  const Block b = blocks[blockIdx.x];
  const uint8_t *src = chunks[b.chunk] + b.src;
  // 16-byte loads from the boundary below src: each holds a byte of the
  // block, so none reaches past the page of a byte the chunk owns.
  const uintptr_t base = reinterpret_cast<uintptr_t>(src) & ~uintptr_t(15);
  const uint32_t head = (uint32_t)(reinterpret_cast<uintptr_t>(src) - base);
  const uint32_t span = head + b.comp;
  for (uint32_t k = 16 * lane; k < span; k += 16 * kWave)
    *reinterpret_cast<uint4 *>(lds + k) =
        *reinterpret_cast<const uint4 *>(base + k);
  __syncthreads();
  uint8_t *planes = lds + stage;
  if (ns >= 0) {
    // the recorded sequences, four a lane, replayed in order
    const uint64_t *mine = seqs + (size_t)blockIdx.x * kMaxSeq;
    uint64_t t[kMaxSeq / kWave];
#pragma unroll
    for (int q = 0; q < kMaxSeq / kWave; q++)
      t[q] = q * kWave + lane < ns ? mine[q * kWave + lane] : 0;
    uint8_t *trash = planes + b.bytes + lane;
    int o = 0;
#pragma unroll
    for (int q = 0; q < kMaxSeq / kWave; q++) {
      const int end = ns - q * kWave < kWave ? ns - q * kWave : kWave;
      for (int l = 0; l < end; l++) {
        const uint64_t e =
            (uint64_t)__builtin_amdgcn_readlane((uint32_t)t[q], l) |
            ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(t[q] >> 32), l)
             << 32);
        replay_seq(e, lds + head, planes, o, trash, lane);
      }
    }
  } else {
    const int got = lds_lz4(lds + head, (int)b.comp, planes, (int)b.bytes,
                            lane);
    if (got != (int)b.bytes) {
      if (lane == 0)
        status[b.chunk] = -91;
      return;
    }
  }
  __syncthreads();
  uint8_t *out = frames[b.chunk] + b.dst;
  const bool aligned = (reinterpret_cast<uintptr_t>(frames[b.chunk]) & 15) == 0;
  unshuffle<ES>(planes, (int)(b.bytes / ES / 8), out, aligned, lane);
}

// One plain-LZ4 block a wave, decoded from the chunk straight into the
// frame (a block may be the whole frame).
__global__ __launch_bounds__(64) void lz4_kernel(const Block *blocks,
                                                 const uint8_t *const *chunks,
                                                 uint8_t *const *frames,
                                                 int32_t *status) {
  const Block b = blocks[blockIdx.x];
  const long long got =
      wave_lz4<long long>(chunks[b.chunk] + b.src, (long long)b.comp,
                          frames[b.chunk] + b.dst, (long long)b.room,
                          threadIdx.x);
  if (got != (long long)b.bytes && threadIdx.x == 0)
    status[b.chunk] = -1;
}

// The bytes a bitshuffle chunk stores raw after its blocks.
__global__ __launch_bounds__(64) void raw_kernel(const Block *blocks,
                                                 const uint8_t *const *chunks,
                                                 uint8_t *const *frames) {
  const Block b = blocks[blockIdx.x];
  for (uint32_t k = threadIdx.x; k < b.bytes; k += kWave)
    frames[b.chunk][b.dst + k] = chunks[b.chunk][b.src + k];
}

// ------------------------------------------------------------ host encoder

size_t lz4_encode(const uint8_t *in, size_t n, uint8_t *out, size_t cap) {
  uint32_t table[1u << 12] = {}; // position + 1, by a hash of 4 bytes
  size_t o = 0;
  auto put_length = [&](size_t v) {
    while (v >= 255) {
      if (o >= cap)
        return false;
      out[o++] = 255;
      v -= 255;
    }
    if (o >= cap)
      return false;
    out[o++] = (uint8_t)v;
    return true;
  };
  // A sequence: literals in[from, from + lit), then a match (mlen 0: none).
  auto sequence = [&](size_t from, size_t lit, size_t off, size_t mlen) {
    if (o >= cap)
      return false;
    const size_t token_at = o++;
    uint8_t token = (uint8_t)(std::min<size_t>(lit, 15) << 4);
    if (lit >= 15 && !put_length(lit - 15))
      return false;
    if (lit > cap - o)
      return false;
    memcpy(out + o, in + from, lit);
    o += lit;
    if (mlen) {
      if (cap - o < 2)
        return false;
      out[o++] = (uint8_t)off;
      out[o++] = (uint8_t)(off >> 8);
      token |= (uint8_t)std::min<size_t>(mlen - kMinMatch, 15);
      if (mlen - kMinMatch >= 15 && !put_length(mlen - kMinMatch - 15))
        return false;
    }
    out[token_at] = token;
    return true;
  };
  size_t anchor = 0, misses = 0;
  // Matches start at least 12 bytes before the end and end at least 5
  // before it, as the block format requires. A run of misses takes longer
  // steps (incompressible bit planes).
  for (size_t i = 0; i + kMfLimit <= n;) {
    uint32_t word, seen;
    memcpy(&word, in + i, 4);
    const uint32_t h = (word * 2654435761u) >> 20;
    const size_t cand = table[h];
    table[h] = (uint32_t)(i + 1);
    if (cand && i - (cand - 1) <= 65535 &&
        (memcpy(&seen, in + cand - 1, 4), seen == word)) {
      const size_t from = cand - 1;
      size_t len = kMinMatch;
      while (i + len < n - kLastLiterals && in[from + len] == in[i + len])
        len++;
      if (!sequence(anchor, i - anchor, i - from, len))
        return 0;
      i += len;
      anchor = i;
      misses = 0;
    } else {
      i += 1 + (misses++ >> 6);
    }
  }
  if (!sequence(anchor, n - anchor, 0, 0))
    return 0;
  return o;
}

// bitshuffle's forward transpose of one block of n elements (n % 8 == 0):
// bit k of byte j of element i goes to bit i % 8 of byte i / 8 of plane
// 8 j + k.
void bitshuffle_block(const uint8_t *in, size_t n, size_t es, uint8_t *planes) {
  const size_t groups = n / 8;
  for (size_t g = 0; g < groups; g++)
    for (size_t j = 0; j < es; j++) {
      uint64_t x = 0;
      for (int i = 0; i < 8; i++)
        x |= (uint64_t)in[(8 * g + i) * es + j] << (8 * i);
      // transpose: bit 8i + k -> bit 8k + i
      uint64_t t = (x ^ (x >> 7)) & 0x00AA00AA00AA00AAull;
      x ^= t ^ (t << 7);
      t = (x ^ (x >> 14)) & 0x0000CCCC0000CCCCull;
      x ^= t ^ (t << 14);
      t = (x ^ (x >> 28)) & 0x00000000F0F0F0F0ull;
      x ^= t ^ (t << 28);
      for (int k = 0; k < 8; k++)
        planes[(8 * j + k) * groups + g] = (uint8_t)(x >> (8 * k));
    }
}

void put_be(uint8_t *p, uint64_t v, int bytes) {
  for (int i = bytes - 1; i >= 0; i--, v >>= 8)
    p[i] = (uint8_t)v;
}

} // namespace

// ------------------------------------------------------------- decoder

// One chunk's walk: its blocks by kind, its largest block, and its status
// and message when it does not walk.
struct ChunkWalk {
  std::vector<Block> kinds[3];
  uint32_t max_comp = 0, max_bytes = 0;
  int status = SF_OK;
  std::string message;
};

struct sf_decoder {
  int device = 0;
  hipStream_t stream = nullptr;
  size_t element = 0, expected = 0;
  int max_frames = 0;
  // the indexed batch
  int n = 0;
  bool stored = false;
  std::string algorithm;
  std::vector<Block> kinds[3];
  std::vector<ChunkWalk> walks; // per chunk of the batch, reused
  uint32_t first[3] = {0, 0, 0}, count[3] = {0, 0, 0};
  uint32_t max_comp = 0, max_bytes = 0;
  Block *h_blocks = nullptr, *d_blocks = nullptr;
  size_t block_cap = 0;
  // chunk pointers then frame pointers, pinned and on the device
  const void **h_ptrs = nullptr;
  void **d_ptrs = nullptr;
  bool ptrs_valid = false;
  int32_t *h_status = nullptr, *d_status = nullptr;
  // the parse pass's sequence lists (kMaxSeq a block) and counts
  uint64_t *d_seqs = nullptr;
  int32_t *d_n_seq = nullptr;
  size_t seq_blocks = 0;
  uint8_t *d_stage = nullptr;
  size_t stage_cap = 0;
  // HIP events around the decode launches, when timing is on
  bool timing = false;
  hipEvent_t events[2] = {nullptr, nullptr};
};

namespace {

std::string prefix(int n, int f) {
  return n > 1 ? "chunk " + std::to_string(f) + ": " : std::string();
}

// The checks of decompress::image (src/decompress.cc:83-135) on one chunk,
// and its blocks into `w`. Reads the decoder's settings only, so chunks can
// be walked on several threads at once.
int walk_chunk(const sf_decoder *d, const uint8_t *data, size_t size, int f,
               int n, ChunkWalk &w) {
  const std::string at = prefix(n, f);
  const size_t expected = d->expected, es = d->element;
  if (d->stored) {
    if (size != expected)
      return fail(SF_BAD_DATA, at + "uncompressed frame is " +
                                   std::to_string(size) + " bytes, expected " +
                                   std::to_string(expected));
    return SF_OK;
  }
  if (size < kHeader)
    return fail(SF_BAD_DATA, at + "too short for a filter header");
  uint64_t total = 0;
  for (size_t i = 0; i < 8; i++)
    total = (total << 8) | data[i];
  const uint32_t block = be32(data + 8);
  if (total != expected)
    return fail(SF_BAD_DATA, at + "filter header says " +
                                 std::to_string(total) +
                                 " bytes uncompressed, expected " +
                                 std::to_string(expected));
  size_t pos = kHeader;
  if (d->algorithm == "bslz4") {
    const std::string bad = at + "bitshuffle decompression failed with code ";
    const size_t elems = expected / es;
    size_t bs = block / es;
    if (bs == 0)
      bs = default_block(es);
    if (bs % 8)
      return fail(SF_BAD_DATA, bad + "-81");
    if (bs * es > kMaxShuffled)
      return fail(SF_BAD_ARGUMENT,
                  at + "bitshuffle blocks of " + std::to_string(bs * es) +
                      " bytes exceed the device decoder's " +
                      std::to_string(kMaxShuffled));
    const size_t full = elems / bs;
    const size_t last = elems % bs - elems % bs % 8;
    size_t done = 0;
    for (size_t k = 0; k < full + (last ? 1 : 0); k++) {
      const size_t count = k < full ? bs : last;
      if (size - pos < 4)
        return fail(SF_BAD_DATA, bad + "-91");
      const uint32_t comp = be32(data + pos);
      pos += 4;
      if (comp > size - pos || comp > lz4_bound(count * es))
        return fail(SF_BAD_DATA, bad + "-91");
      w.kinds[kShuffled].push_back(Block{pos, done * es, comp,
                                          (uint32_t)(count * es), (uint32_t)f,
                                          (uint32_t)(count * es)});
      w.max_comp = std::max(w.max_comp, comp);
      w.max_bytes = std::max(w.max_bytes, (uint32_t)(count * es));
      pos += comp;
      done += count;
    }
    const size_t leftover = (elems - done) * es;
    if (leftover > size - pos)
      return fail(SF_BAD_DATA, bad + "-91");
    if (leftover)
      w.kinds[kRaw].push_back(Block{pos, done * es, 0, (uint32_t)leftover,
                                     (uint32_t)f, (uint32_t)leftover});
    return SF_OK;
  }
  // "lz4": the HDF5 LZ4 filter's framing (src/decompress.cc:42-70). Each
  // block is decoded with room for min(block size, what is left) and the
  // next one goes where its output ended, so a block's length is walked
  // here (its sequences' lengths, no copies) to place the next one.
  const size_t block_size = block > 0 ? block : expected;
  size_t written = 0;
  while (written < expected) {
    if (pos + 4 > size)
      return fail(SF_BAD_DATA, at + "lz4 data truncated");
    const uint32_t comp = be32(data + pos);
    pos += 4;
    if (comp == 0 || pos + comp > size)
      return fail(SF_BAD_DATA, at + "lz4 block size out of range");
    const size_t want = std::min(expected - written, block_size);
    if (want > UINT32_MAX)
      return fail(SF_BAD_ARGUMENT, at + "an lz4 block over 4 GiB");
    const long long got = lz4_length(data + pos, comp, want);
    if (got < 0)
      return fail(SF_BAD_DATA, at + "lz4 decompression failed with code -1");
    if (got > 0)
      w.kinds[kLz4].push_back(Block{pos, written, comp, (uint32_t)got,
                                     (uint32_t)f, (uint32_t)want});
    pos += comp;
    written += (size_t)got;
  }
  return SF_OK;
}

// Walks the chunks of a batch, on up to 16 host threads when there are
// several: the block lengths form a dependent chain through each chunk, a
// cache and TLB miss a block, and chunks are independent. Their blocks join
// the tables in chunk order; the first chunk in order that does not walk
// gives the status and message.
int walk_chunks(sf_decoder *d, const uint8_t *const *chunks,
                const size_t *sizes, int n) {
  if ((int)d->walks.size() < n)
    d->walks.resize(n);
  for (int f = 0; f < n; f++) {
    ChunkWalk &w = d->walks[f];
    for (auto &k : w.kinds)
      k.clear();
    w.max_comp = w.max_bytes = 0;
    w.status = SF_OK;
    w.message.clear();
  }
  const auto walk = [&](int f) {
    ChunkWalk &w = d->walks[f];
    w.status = walk_chunk(d, chunks[f], sizes[f], f, n, w);
    if (w.status != SF_OK)
      w.message = g_decode_error;
  };
  const int threads = std::min<int>(
      {n, 16, std::max<int>(1, (int)std::thread::hardware_concurrency())});
  if (threads <= 1 || d->stored) {
    for (int f = 0; f < n; f++)
      walk(f);
  } else {
    std::vector<std::thread> pool;
    pool.reserve(threads - 1);
    for (int t = 1; t < threads; t++)
      pool.emplace_back([&, t] {
        for (int f = t; f < n; f += threads)
          walk(f);
      });
    for (int f = 0; f < n; f += threads)
      walk(f);
    for (auto &th : pool)
      th.join();
  }
  for (int f = 0; f < n; f++) {
    const ChunkWalk &w = d->walks[f];
    if (w.status != SF_OK)
      return fail(w.status, w.message);
    for (int k = 0; k < 3; k++)
      d->kinds[k].insert(d->kinds[k].end(), w.kinds[k].begin(),
                         w.kinds[k].end());
    d->max_comp = std::max(d->max_comp, w.max_comp);
    d->max_bytes = std::max(d->max_bytes, w.max_bytes);
  }
  return SF_OK;
}

int device_fail(const char *what, hipError_t e) {
  return fail(SF_DEVICE_ERROR,
              std::string(what) + ": " + hipGetErrorString(e));
}

} // namespace

extern "C" int sf_hip_decoder_create(unsigned bit_depth, size_t height,
                                     size_t width, int max_frames,
                                     sf_decoder **out) {
  if (!out || (bit_depth != 8 && bit_depth != 16 && bit_depth != 32 &&
               bit_depth != 64) ||
      height == 0 || width == 0 || max_frames <= 0)
    return fail(SF_BAD_ARGUMENT, "sf_hip_decoder_create: bad argument");
  *out = nullptr;
  auto *d = new sf_decoder;
  d->element = bit_depth / 8;
  d->expected = height * width * bit_depth / 8; // src/decompress.cc:71-74
  d->max_frames = max_frames;
  hipError_t e = hipGetDevice(&d->device);
  if (e == hipSuccess)
    e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
  if (e == hipSuccess)
    e = hipHostMalloc(reinterpret_cast<void **>(&d->h_ptrs),
                      2 * max_frames * sizeof(void *), hipHostMallocDefault);
  if (e == hipSuccess)
    e = hipMalloc(reinterpret_cast<void **>(&d->d_ptrs),
                  2 * max_frames * sizeof(void *));
  if (e == hipSuccess)
    e = hipHostMalloc(reinterpret_cast<void **>(&d->h_status),
                      max_frames * sizeof(int32_t), hipHostMallocDefault);
  if (e == hipSuccess)
    e = hipMalloc(reinterpret_cast<void **>(&d->d_status),
                  max_frames * sizeof(int32_t));
  for (int k = 0; k < 2 && e == hipSuccess; k++)
    e = hipEventCreate(&d->events[k]);
  if (e != hipSuccess) {
    sf_hip_decoder_destroy(d);
    return device_fail("sf_hip_decoder_create", e);
  }
  *out = d;
  return SF_OK;
}

extern "C" int sf_hip_decoder_index(sf_decoder *d,
                                    const uint8_t *const *chunks,
                                    const size_t *sizes, int n,
                                    const char *algorithm) {
  if (!d || !chunks || !sizes || !algorithm || n <= 0 || n > d->max_frames)
    return fail(SF_BAD_ARGUMENT, "sf_hip_decoder_index: bad argument");
  for (int f = 0; f < n; f++)
    if (!chunks[f])
      return fail(SF_BAD_ARGUMENT, "sf_hip_decoder_index: a null chunk");
  // The pinned table may still be on its way to the device.
  hipError_t e = hipStreamSynchronize(d->stream);
  if (e != hipSuccess)
    return device_fail("sf_hip_decoder_index", e);
  d->n = 0;
  d->algorithm = algorithm;
  d->stored = d->algorithm.empty();
  if (!d->stored && d->algorithm != "bslz4" && d->algorithm != "lz4")
    return fail(SF_BAD_DATA, "unknown compression algorithm " + d->algorithm);
  for (auto &k : d->kinds)
    k.clear();
  d->max_comp = 0;
  d->max_bytes = 0;
  const int walked = walk_chunks(d, chunks, sizes, n);
  if (walked != SF_OK)
    return walked;
  size_t total = 0;
  for (int k = 0; k < 3; k++) {
    if (d->kinds[k].size() > (size_t)INT32_MAX)
      return fail(SF_BAD_ARGUMENT, "sf_hip_decoder_index: too many blocks");
    d->first[k] = (uint32_t)total;
    d->count[k] = (uint32_t)d->kinds[k].size();
    total += d->kinds[k].size();
  }
  if (total > d->block_cap) {
    if (d->h_blocks)
      (void)hipHostFree(d->h_blocks);
    if (d->d_blocks)
      (void)hipFree(d->d_blocks);
    d->h_blocks = nullptr;
    d->d_blocks = nullptr;
    d->block_cap = 0;
    const size_t cap = std::max<size_t>(total, 1024);
    e = hipHostMalloc(reinterpret_cast<void **>(&d->h_blocks),
                      cap * sizeof(Block), hipHostMallocDefault);
    if (e == hipSuccess)
      e = hipMalloc(reinterpret_cast<void **>(&d->d_blocks),
                    cap * sizeof(Block));
    if (e != hipSuccess)
      return device_fail("sf_hip_decoder_index", e);
    d->block_cap = cap;
  }
  if (d->count[kShuffled] > d->seq_blocks) {
    if (d->d_seqs)
      (void)hipFree(d->d_seqs);
    if (d->d_n_seq)
      (void)hipFree(d->d_n_seq);
    d->d_seqs = nullptr;
    d->d_n_seq = nullptr;
    d->seq_blocks = 0;
    const size_t blocks = d->count[kShuffled];
    e = hipMalloc(reinterpret_cast<void **>(&d->d_seqs),
                  blocks * kMaxSeq * sizeof(uint64_t));
    if (e == hipSuccess)
      e = hipMalloc(reinterpret_cast<void **>(&d->d_n_seq),
                    blocks * sizeof(int32_t));
    if (e != hipSuccess)
      return device_fail("sf_hip_decoder_index", e);
    d->seq_blocks = blocks;
  }
  for (int k = 0; k < 3; k++)
    std::copy(d->kinds[k].begin(), d->kinds[k].end(),
              d->h_blocks + d->first[k]);
  if (total) {
    e = hipMemcpyAsync(d->d_blocks, d->h_blocks, total * sizeof(Block),
                       hipMemcpyHostToDevice, d->stream);
    if (e != hipSuccess)
      return device_fail("sf_hip_decoder_index", e);
  }
  d->n = n;
  d->ptrs_valid = false;
  return SF_OK;
}

extern "C" int sf_hip_decoder_decode(sf_decoder *d,
                                     const uint8_t *const *device_chunks,
                                     void *const *device_frames) {
  if (!d || !device_chunks || !device_frames)
    return fail(SF_BAD_ARGUMENT, "sf_hip_decoder_decode: bad argument");
  if (d->n == 0)
    return fail(SF_BAD_ARGUMENT, "sf_hip_decoder_decode: nothing indexed");
  const int n = d->n;
  for (int f = 0; f < n; f++)
    if (!device_chunks[f] || !device_frames[f])
      return fail(SF_BAD_ARGUMENT, "sf_hip_decoder_decode: a null pointer");
  hipError_t e;
  if (d->stored) {
    for (int f = 0; f < n; f++) {
      e = hipMemcpyAsync(device_frames[f], device_chunks[f], d->expected,
                         hipMemcpyDeviceToDevice, d->stream);
      if (e != hipSuccess)
        return device_fail("sf_hip_decoder_decode", e);
    }
    return SF_OK;
  }
  bool same = d->ptrs_valid;
  for (int f = 0; same && f < n; f++)
    same = d->h_ptrs[f] == device_chunks[f] &&
           d->h_ptrs[d->max_frames + f] == device_frames[f];
  if (!same) {
    // the pinned pointers may still be on their way to the device
    e = hipStreamSynchronize(d->stream);
    if (e != hipSuccess)
      return device_fail("sf_hip_decoder_decode", e);
    for (int f = 0; f < n; f++) {
      d->h_ptrs[f] = device_chunks[f];
      d->h_ptrs[d->max_frames + f] = device_frames[f];
    }
    e = hipMemcpyAsync(d->d_ptrs, d->h_ptrs,
                       2 * d->max_frames * sizeof(void *),
                       hipMemcpyHostToDevice, d->stream);
    if (e != hipSuccess)
      return device_fail("sf_hip_decoder_decode", e);
    d->ptrs_valid = true;
  }
  const auto *chunks = reinterpret_cast<const uint8_t *const *>(d->d_ptrs);
  auto *frames = reinterpret_cast<uint8_t *const *>(d->d_ptrs + d->max_frames);
  if (d->timing)
    (void)hipEventRecord(d->events[0], d->stream);
  zero_status<<<1, 64, 0, d->stream>>>(d->d_status, n);
  if (d->count[kShuffled]) {
    // LZ4 bytes from the 16-byte boundary below them and 64 bytes that
    // window reads may touch; the planes and 64 bytes for idle lanes
    const uint32_t stage = (d->max_comp + 15 + 64 + 15) / 16 * 16;
    const uint32_t lds = stage + (d->max_bytes + 15) / 16 * 16 + 64;
    const Block *b = d->d_blocks + d->first[kShuffled];
    const dim3 grid(d->count[kShuffled]);
    parse_kernel<<<(d->count[kShuffled] + 63) / 64, 64, 0, d->stream>>>(
        b, d->count[kShuffled], chunks, d->d_seqs, d->d_n_seq, d->d_status);
    const uint64_t *seqs = d->d_seqs;
    const int32_t *n_seq = d->d_n_seq;
    switch (d->element) {
    case 1:
      shuffled_kernel<1><<<grid, 64, lds, d->stream>>>(
          b, chunks, frames, stage, seqs, n_seq, d->d_status);
      break;
    case 2:
      shuffled_kernel<2><<<grid, 64, lds, d->stream>>>(
          b, chunks, frames, stage, seqs, n_seq, d->d_status);
      break;
    case 4:
      shuffled_kernel<4><<<grid, 64, lds, d->stream>>>(
          b, chunks, frames, stage, seqs, n_seq, d->d_status);
      break;
    default:
      shuffled_kernel<8><<<grid, 64, lds, d->stream>>>(
          b, chunks, frames, stage, seqs, n_seq, d->d_status);
      break;
    }
  }
  if (d->count[kLz4])
    lz4_kernel<<<d->count[kLz4], 64, 0, d->stream>>>(
        d->d_blocks + d->first[kLz4], chunks, frames, d->d_status);
  if (d->count[kRaw])
    raw_kernel<<<d->count[kRaw], 64, 0, d->stream>>>(
        d->d_blocks + d->first[kRaw], chunks, frames);
  if (d->timing)
    (void)hipEventRecord(d->events[1], d->stream);
  e = hipGetLastError();
  if (e != hipSuccess)
    return device_fail("sf_hip_decoder_decode", e);
  return SF_OK;
}

extern "C" int sf_hip_decoder_wait(sf_decoder *d) {
  if (!d)
    return fail(SF_BAD_ARGUMENT, "sf_hip_decoder_wait: bad argument");
  if (d->n == 0 || d->stored) {
    const hipError_t e = hipStreamSynchronize(d->stream);
    return e == hipSuccess ? SF_OK : device_fail("sf_hip_decoder_wait", e);
  }
  hipError_t e = hipMemcpyAsync(d->h_status, d->d_status,
                                d->n * sizeof(int32_t),
                                hipMemcpyDeviceToHost, d->stream);
  if (e == hipSuccess)
    e = hipStreamSynchronize(d->stream);
  if (e != hipSuccess)
    return device_fail("sf_hip_decoder_wait", e);
  for (int f = 0; f < d->n; f++)
    if (d->h_status[f] != 0) {
      const bool shuffled = d->algorithm == "bslz4";
      return fail(SF_BAD_DATA,
                  prefix(d->n, f) +
                      (shuffled ? "bitshuffle decompression failed with code "
                                : "lz4 decompression failed with code ") +
                      std::to_string(d->h_status[f]));
    }
  return SF_OK;
}

extern "C" int sf_hip_decoder_image(sf_decoder *d, const uint8_t *data,
                                    size_t size, const char *algorithm,
                                    void *device_frame) {
  if (!d || !data || !algorithm || !device_frame)
    return fail(SF_BAD_ARGUMENT, "sf_hip_decoder_image: bad argument");
  int status = sf_hip_decoder_index(d, &data, &size, 1, algorithm);
  if (status != SF_OK)
    return status;
  if (size > d->stage_cap) {
    if (d->d_stage)
      (void)hipFree(d->d_stage);
    d->d_stage = nullptr;
    d->stage_cap = 0;
    const hipError_t e = hipMalloc(reinterpret_cast<void **>(&d->d_stage),
                                   std::max<size_t>(size, 1 << 20));
    if (e != hipSuccess)
      return device_fail("sf_hip_decoder_image", e);
    d->stage_cap = std::max<size_t>(size, 1 << 20);
  }
  hipError_t e = hipMemcpyAsync(d->d_stage, data, size, hipMemcpyHostToDevice,
                                d->stream);
  if (e != hipSuccess)
    return device_fail("sf_hip_decoder_image", e);
  const uint8_t *staged = d->d_stage;
  status = sf_hip_decoder_decode(d, &staged, &device_frame);
  if (status != SF_OK)
    return status;
  return sf_hip_decoder_wait(d);
}

extern "C" int sf_hip_decoder_set_timing(sf_decoder *d, int on) {
  if (!d)
    return fail(SF_BAD_ARGUMENT, "sf_hip_decoder_set_timing: bad argument");
  d->timing = on != 0;
  return SF_OK;
}

extern "C" double sf_hip_decoder_last_ms(sf_decoder *d) {
  float ms = -1.0f;
  if (!d || !d->timing ||
      hipEventSynchronize(d->events[1]) != hipSuccess ||
      hipEventElapsedTime(&ms, d->events[0], d->events[1]) != hipSuccess)
    return -1.0;
  return ms;
}

extern "C" void *sf_hip_decoder_stream(sf_decoder *d) {
  return d ? static_cast<void *>(d->stream) : nullptr;
}

extern "C" void sf_hip_decoder_destroy(sf_decoder *d) {
  if (!d)
    return;
  if (d->stream)
    (void)hipStreamSynchronize(d->stream);
  if (d->h_blocks)
    (void)hipHostFree(d->h_blocks);
  if (d->d_blocks)
    (void)hipFree(d->d_blocks);
  if (d->h_ptrs)
    (void)hipHostFree(d->h_ptrs);
  if (d->d_ptrs)
    (void)hipFree(d->d_ptrs);
  if (d->h_status)
    (void)hipHostFree(d->h_status);
  if (d->d_status)
    (void)hipFree(d->d_status);
  if (d->d_stage)
    (void)hipFree(d->d_stage);
  if (d->d_seqs)
    (void)hipFree(d->d_seqs);
  if (d->d_n_seq)
    (void)hipFree(d->d_n_seq);
  for (hipEvent_t ev : d->events)
    if (ev)
      (void)hipEventDestroy(ev);
  if (d->stream)
    (void)hipStreamDestroy(d->stream);
  delete d;
}

extern "C" const char *sf_hip_decode_last_error(void) {
  return g_decode_error.c_str();
}

// ---------------------------------------------------------- host encoders

extern "C" size_t sf_lz4_compress_block(const uint8_t *src, size_t n,
                                        uint8_t *dst, size_t cap) {
  if ((!src && n) || !dst)
    return 0;
  return lz4_encode(src, n, dst, cap);
}

extern "C" size_t sf_bslz4_bound(size_t n_elems, size_t elem_size,
                                 size_t block_size) {
  if (elem_size == 0)
    return 0;
  if (block_size == 0)
    block_size = default_block(elem_size);
  const size_t blocks = n_elems / block_size + 1;
  return kHeader + blocks * (4 + lz4_bound(block_size * elem_size)) +
         8 * elem_size;
}

extern "C" int sf_bslz4_compress(const void *elements, size_t n_elems,
                                 size_t elem_size, size_t block_size,
                                 uint8_t *out, size_t cap, size_t *written) {
  if ((!elements && n_elems) || !out || !written || elem_size == 0)
    return fail(SF_BAD_ARGUMENT, "sf_bslz4_compress: bad argument");
  if (block_size == 0)
    block_size = default_block(elem_size);
  if (block_size % 8)
    return fail(SF_BAD_ARGUMENT,
                "sf_bslz4_compress: block size not a multiple of 8");
  if (cap < sf_bslz4_bound(n_elems, elem_size, block_size))
    return fail(SF_OVERFLOW, "sf_bslz4_compress: output below the bound");
  const auto *in = static_cast<const uint8_t *>(elements);
  put_be(out, (uint64_t)n_elems * elem_size, 8);
  put_be(out + 8, (uint64_t)(block_size * elem_size), 4);
  size_t o = kHeader, done = 0;
  std::vector<uint8_t> planes(block_size * elem_size);
  const size_t full = n_elems / block_size;
  const size_t last = n_elems % block_size - n_elems % block_size % 8;
  for (size_t k = 0; k < full + (last ? 1 : 0); k++) {
    const size_t count = k < full ? block_size : last;
    bitshuffle_block(in + done * elem_size, count, elem_size, planes.data());
    const size_t comp = lz4_encode(planes.data(), count * elem_size,
                                   out + o + 4, cap - o - 4);
    if (comp == 0)
      return fail(SF_OVERFLOW, "sf_bslz4_compress: output below the bound");
    put_be(out + o, comp, 4);
    o += 4 + comp;
    done += count;
  }
  const size_t leftover = (n_elems - done) * elem_size;
  memcpy(out + o, in + done * elem_size, leftover);
  *written = o + leftover;
  return SF_OK;
}
