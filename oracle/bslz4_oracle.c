/* CPU restatement of the reference's frame decompression -- TEST
 * INFRASTRUCTURE ONLY: the checker for the device decoder, never linked into
 * the product.
 *
 * decompress::image (src/decompress.cc:76-136) takes one HDF5 chunk as
 * stored -- a 12-byte big-endian filter header (uncompressed bytes, block
 * bytes; src/decompress.cc:14-36) and the payload -- and writes the frame.
 * Its three algorithms:
 *   ""       a stored chunk, copied (src/decompress.cc:90-97);
 *   "bslz4"  bitshuffle's bshuf_decompress_lz4 (src/decompress.cc:116-125).
 *            bitshuffle (kiyo-masui/bitshuffle, a submodule absent from the
 *            reference checkout; its pinned commit is not recorded) is
 *            restated here from its published format: blocks of
 *            block_size elements, each a 4-byte big-endian LZ4 length and an
 *            LZ4 block holding the block's bit planes; the last partial block
 *            rounded down to a multiple of 8 elements; the last size % 8
 *            elements stored raw. The bit-plane layout is parity unpinned
 *            against bitshuffle itself: no implementation or fixture of it
 *            is available here, and only a numpy construction of the
 *            published layout and hand-made known answers check it;
 *   "lz4"    the HDF5 LZ4 filter's framing (src/decompress.cc:42-70).
 * The LZ4 block decoder follows the published block format
 * (lz4_Block_format.md, liblz4 1.9): sequences of a token, literals, a
 * 2-byte little-endian offset and a match; the end-of-block rules that
 * LZ4_decompress_safe enforces (the last 5 bytes of the output are literals,
 * no match starts within the last 12).
 *
 * Where the reference reads past a corrupt chunk (bitshuffle trusts the
 * block lengths it reads), this restatement reports the bitshuffle failure
 * code -91 instead.
 */
#include "bslz4_oracle.h"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

enum { kMinMatch = 4, kLastLiterals = 5, kMfLimit = 12 };

/* One LZ4 block into dst[0, cap): the bytes written, or -1 when the block
 * is malformed (LZ4_decompress_safe's contract, with a fixed error value). */
long oracle_lz4_block(const uint8_t *src, size_t n, uint8_t *dst, size_t cap) {
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
    /* A sequence whose literals reach the last 12 output bytes, or leave
     * fewer than 8 input bytes, is the block's last: it ends the input. */
    if (o + lit + kMfLimit > cap || i + lit + 8 > n) {
      if (i + lit != n || o + lit > cap)
        return -1;
      memcpy(dst + o, src + i, lit);
      return (long)(o + lit);
    }
    memcpy(dst + o, src + i, lit);
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
    if (o + len + kLastLiterals > cap)
      return -1;
    for (size_t k = 0; k < len; k++) /* byte by byte: overlap repeats */
      dst[o + k] = dst[o - off + k];
    o += len;
  }
}

static uint32_t be32(const uint8_t *p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) |
         ((uint32_t)p[2] << 8) | p[3];
}

/* bitshuffle's inverse transpose of one block of n elements (n % 8 == 0):
 * bit k of byte j of element i is bit (i % 8) of byte i / 8 of bit plane
 * 8 j + k, each plane n / 8 bytes long. */
void oracle_bitunshuffle(const uint8_t *planes, size_t n, size_t elem_size,
                         uint8_t *out) {
  memset(out, 0, n * elem_size);
  for (size_t i = 0; i < n; i++)
    for (size_t j = 0; j < elem_size; j++)
      for (unsigned k = 0; k < 8; k++) {
        const size_t plane = 8 * j + k;
        const unsigned bit = (planes[plane * (n / 8) + i / 8] >> (i % 8)) & 1;
        out[i * elem_size + j] |= (uint8_t)(bit << k);
      }
}

/* The forward transpose, for building test chunks. */
void oracle_bitshuffle(const uint8_t *in, size_t n, size_t elem_size,
                       uint8_t *planes) {
  memset(planes, 0, n * elem_size);
  for (size_t i = 0; i < n; i++)
    for (size_t j = 0; j < elem_size; j++)
      for (unsigned k = 0; k < 8; k++) {
        const unsigned bit = (in[i * elem_size + j] >> k) & 1;
        planes[(8 * j + k) * (n / 8) + i / 8] |= (uint8_t)(bit << (i % 8));
      }
}

size_t oracle_bshuf_default_block(size_t elem_size) {
  size_t block = 8192 / elem_size;
  block = block / 8 * 8;
  return block < 128 ? 128 : block;
}

/* bshuf_decompress_lz4(in, out, size, elem_size, block_size): the payload
 * bytes consumed, or a negative bitshuffle code (-81 a block size that is not
 * a multiple of 8, -91 a block that does not decompress to its size). */
long long oracle_bshuf_decompress_lz4(const uint8_t *in, size_t in_len,
                                      uint8_t *out, size_t size,
                                      size_t elem_size, size_t block_size) {
  static uint8_t planes[1 << 20];
  if (block_size == 0)
    block_size = oracle_bshuf_default_block(elem_size);
  if (block_size % 8)
    return -81;
  if (block_size * elem_size > sizeof planes)
    return -80;
  size_t read = 0, done = 0;
  const size_t last = size % block_size - size % block_size % 8;
  const size_t n_blocks = size / block_size + (last ? 1 : 0);
  for (size_t b = 0; b < n_blocks; b++) {
    const size_t n = b < size / block_size ? block_size : last;
    if (in_len - read < 4)
      return -91;
    const uint32_t comp = be32(in + read);
    read += 4;
    if (comp > in_len - read)
      return -91;
    const long got = oracle_lz4_block(in + read, comp, planes, n * elem_size);
    if (got != (long)(n * elem_size))
      return -91;
    oracle_bitunshuffle(planes, n, elem_size, out + done * elem_size);
    read += comp;
    done += n;
  }
  const size_t leftover = (size - done) * elem_size;
  if (leftover > in_len - read)
    return -91;
  memcpy(out + done * elem_size, in + read, leftover);
  return (long long)(read + leftover);
}

static int fail(char *msg, size_t cap, const char *fmt, ...) {
  if (msg && cap) {
    va_list args;
    va_start(args, fmt);
    vsnprintf(msg, cap, fmt, args);
    va_end(args);
  }
  return 1;
}

int oracle_decompress_image(const uint8_t *data, size_t len,
                            const char *algorithm, unsigned bit_depth,
                            size_t height, size_t width, uint8_t *out,
                            size_t out_len, char *msg, size_t msg_cap) {
  const size_t expected = height * width * bit_depth / 8;
  const size_t element = bit_depth / 8;
  if (msg && msg_cap)
    msg[0] = 0;
  if (element == 0 || out_len < expected)
    return fail(msg, msg_cap, "output buffer too small for the frame");
  if (algorithm[0] == 0) {
    if (len != expected)
      return fail(msg, msg_cap, "uncompressed frame is %zu bytes, expected %zu",
                  len, expected);
    memcpy(out, data, expected);
    return 0;
  }
  if (len < 12)
    return fail(msg, msg_cap, "too short for a filter header");
  uint64_t total = 0;
  for (int i = 0; i < 8; i++)
    total = (total << 8) | data[i];
  const uint32_t block = be32(data + 8);
  if (total != expected)
    return fail(msg, msg_cap,
                "filter header says %llu bytes uncompressed, expected %zu",
                (unsigned long long)total, expected);
  const uint8_t *payload = data + 12;
  const size_t payload_len = len - 12;
  if (strcmp(algorithm, "bslz4") == 0) {
    const long long used = oracle_bshuf_decompress_lz4(
        payload, payload_len, out, expected / element, element,
        block / element);
    if (used < 0)
      return fail(msg, msg_cap, "bitshuffle decompression failed with code %lld",
                  used);
    return 0;
  }
  if (strcmp(algorithm, "lz4") == 0) {
    const size_t block_size = block > 0 ? block : expected;
    size_t read = 0, written = 0;
    while (written < expected) {
      if (read + 4 > payload_len)
        return fail(msg, msg_cap, "lz4 data truncated");
      const uint32_t comp = be32(payload + read);
      read += 4;
      if (comp == 0 || read + comp > payload_len)
        return fail(msg, msg_cap, "lz4 block size out of range");
      const size_t remaining = expected - written;
      const size_t want = remaining < block_size ? remaining : block_size;
      const long got =
          oracle_lz4_block(payload + read, comp, out + written, want);
      if (got < 0)
        return fail(msg, msg_cap, "lz4 decompression failed with code %ld", got);
      read += comp;
      written += (size_t)got;
    }
    return 0;
  }
  return fail(msg, msg_cap, "unknown compression algorithm %s", algorithm);
}
