/* Decoding HDF5 chunks into device frames: the reference's
 * decompress::image (src/decompress.hh:20-30, src/decompress.cc:76-136) on
 * the device, for batches of chunks.
 *
 * A chunk is what HDF5 stored for one frame: for "bslz4" (bitshuffle+LZ4)
 * and "lz4" a 12-byte big-endian header (uncompressed bytes, block bytes)
 * and the blocks, each a 4-byte big-endian length and an LZ4 block; for ""
 * the frame's bytes as they are. The host walks the block lengths
 * (sf_hip_decoder_index) and validates each chunk as decompress::image does,
 * with its messages; the device then decodes every block of the batch
 * (sf_hip_decoder_decode): a pass that records each block's LZ4 sequences,
 * then a wave per bitshuffle block that replays them in LDS and writes the
 * unshuffled elements to the frame.
 *
 * Where the reference reads past the end of a corrupt chunk (bitshuffle
 * trusts the block lengths it reads), the decoder reports the bitshuffle
 * failure code -91 instead. A plain-LZ4 block that decodes to fewer bytes
 * than its place in the frame is an error here ("lz4 decompression failed
 * with code -1"); the reference would shift the rest of the frame.
 */
#ifndef SF_DECODE_HIP_H
#define SF_DECODE_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "dext_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sf_decoder sf_decoder;

/* A decoder for frames of height x width elements of bit_depth bits (8, 16,
 * 32 or 64), up to max_frames chunks a batch, on the current device with a
 * stream of its own. decompress::frame_bytes (src/decompress.cc:71-74) is
 * height * width * bit_depth / 8. Returns SF_OK, SF_BAD_ARGUMENT or
 * SF_DEVICE_ERROR. */
int sf_hip_decoder_create(unsigned bit_depth, size_t height, size_t width,
                          int max_frames, sf_decoder **out);

/* Walks n chunks held in host memory (chunks[f], sizes[f] bytes, all of one
 * algorithm: "bslz4", "lz4" or "") and uploads their block table. Checks
 * each chunk as decompress::image does (src/decompress.cc:83-135): SF_BAD_DATA
 * with the reference's message (prefixed "chunk f: " in a batch of more than
 * one) in sf_hip_decode_last_error(). A bitshuffle block of more than 16384
 * bytes is SF_BAD_ARGUMENT (the default block is 8192 bytes). Waits for the
 * decoder's earlier work first, then returns with the upload in flight. */
int sf_hip_decoder_index(sf_decoder *dec, const uint8_t *const *chunks,
                         const size_t *sizes, int n, const char *algorithm);

/* Decodes the indexed chunks, device_chunks[f] holding the same bytes as
 * chunks[f] did, into device_frames[f] (frame_bytes each), asynchronously on
 * the decoder's stream. Can be repeated on other chunks with the same
 * block lengths (the same chunks, copied elsewhere). */
int sf_hip_decoder_decode(sf_decoder *dec, const uint8_t *const *device_chunks,
                          void *const *device_frames);

/* Waits for the decode; SF_OK, or SF_BAD_DATA when a block failed to decode
 * ("bitshuffle decompression failed with code -91", "lz4 decompression
 * failed with code -1"). */
int sf_hip_decoder_wait(sf_decoder *dec);

/* decompress::image (src/decompress.hh:20-30) into a device frame: one chunk
 * in host memory is walked, copied to the device, decoded and waited for. */
int sf_hip_decoder_image(sf_decoder *dec, const uint8_t *data, size_t size,
                         const char *algorithm, void *device_frame);

/* HIP events on the decoder's stream around each decode's launches (off by
 * default); sf_hip_decoder_last_ms waits for the last timed decode and
 * returns its device time in ms, or -1 when timing is off. */
int sf_hip_decoder_set_timing(sf_decoder *dec, int on);
double sf_hip_decoder_last_ms(sf_decoder *dec);

/* The decoder's stream (a hipStream_t), to order the threshold after it. */
void *sf_hip_decoder_stream(sf_decoder *dec);
void sf_hip_decoder_destroy(sf_decoder *dec);
const char *sf_hip_decode_last_error(void);

/* Host encoders, for writing test chunks and synthetic compressed series
 * (the reference reads chunks that HDF5 filters wrote; it writes none).
 * sf_bslz4_compress writes the HDF5 bitshuffle filter's chunk: the 12-byte
 * header, bitshuffled blocks of block_size elements (0: 8192 bytes' worth,
 * a multiple of 8) LZ4-compressed, the last n % 8 elements raw. Returns
 * SF_OK, SF_BAD_ARGUMENT or SF_OVERFLOW (cap below sf_bslz4_bound). */
size_t sf_bslz4_bound(size_t n_elems, size_t elem_size, size_t block_size);
int sf_bslz4_compress(const void *elements, size_t n_elems, size_t elem_size,
                      size_t block_size, uint8_t *out, size_t cap,
                      size_t *written);
/* One LZ4 block (the published block format): its length, or 0 when it
 * does not fit in cap (n + n / 255 + 16 always fits). */
size_t sf_lz4_compress_block(const uint8_t *src, size_t n, uint8_t *dst,
                             size_t cap);

#ifdef __cplusplus
}
#endif

#endif
