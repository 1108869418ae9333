/* CPU restatement of the reference's frame decompression (src/decompress.cc)
 * -- TEST INFRASTRUCTURE ONLY. */
#ifndef BSLZ4_ORACLE_H
#define BSLZ4_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

long oracle_lz4_block(const uint8_t *src, size_t n, uint8_t *dst, size_t cap);
void oracle_bitunshuffle(const uint8_t *planes, size_t n, size_t elem_size,
                         uint8_t *out);
void oracle_bitshuffle(const uint8_t *in, size_t n, size_t elem_size,
                       uint8_t *planes);
size_t oracle_bshuf_default_block(size_t elem_size);
long long oracle_bshuf_decompress_lz4(const uint8_t *in, size_t in_len,
                                      uint8_t *out, size_t size,
                                      size_t elem_size, size_t block_size);
/* decompress::image: 0, or 1 with the reference's message in msg. */
int oracle_decompress_image(const uint8_t *data, size_t len,
                            const char *algorithm, unsigned bit_depth,
                            size_t height, size_t width, uint8_t *out,
                            size_t out_len, char *msg, size_t msg_cap);

#ifdef __cplusplus
}
#endif

#endif
