// sf_hip_find_static: the threshold of one device-resident frame with no
// device allocation at all, for callers that cannot allocate (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// It runs the same kernels as sf_hip_find_u16 / sf_hip_find_u32 (the sparse
// passes and their ordering for 16-bit frames, the register sweep and its
// ordering for 32-bit ones). Their working memory -- the frame descriptor,
// the counters, the unordered staging list and the per-frame tables -- lives
// in one file-scope device array, `scratch`, of kScratchBytes (96 MiB). Each device
// holds its own copy of the array for as long as the process runs; its
// address on a device is looked up once (hipGetSymbolAddress, with that
// device current) and kept, and only that address is handed to kernels.
//
// Every call on a device shares that device's array and leaves it holding
// the call's tables: calls that overlap on one device (two streams, two
// threads) race and must be serialised by the caller. Everything a call
// needs cleared it clears itself, with kernels (no memset node), so a call
// may be captured into a graph and replayed.
#include "sf_device.hh"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

namespace sf {
void fused_layout(int pixel_bytes, uint32_t height, uint32_t width,
                  FrameDesc &d);
size_t cell_entries(uint32_t height, uint32_t width);
size_t sparse_bytes(uint32_t height, uint32_t width, uint32_t capacity);
void sparse_carve(uint8_t *base, FrameDesc &d);
hipError_t launch_batch(int pixel_bytes, bool any_vec, bool any_scalar,
                        const FrameDesc *d, int n, uint32_t items,
                        uint32_t max_bands, uint32_t max_cell_rows,
                        hipStream_t stream);
void ensure_device_tables();
void set_last_error(const std::string &message);
} // namespace sf

namespace {

// SF_FIND_STATIC_SCRATCH_BYTES, 96 MiB: a 4362 x 4148 frame (the Eiger2
// 16M) needs 59.9 MB at 16 bits (the sparse tables and the staging list)
// and 73.8 MB at 32 bits (the staging list and the sweep's cell tables).
constexpr size_t kScratchBytes = SF_FIND_STATIC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

__device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_lookup;
uint8_t *g_scratch[kMaxDevices] = {};

// Messages go to the calling thread's sf_hip_last_error().
int refuse(const std::string &what) {
  sf::set_last_error(what);
  return SF_BAD_ARGUMENT;
}

int device_fail(const char *what, hipError_t e) {
  sf::set_last_error(std::string(what) + ": " + hipGetErrorString(e));
  return SF_DEVICE_ERROR;
}

// The current device's copy of `scratch`, looked up on its first use there.
int scratch_of_device(uint8_t **out) {
  int device = 0;
  hipError_t e = hipGetDevice(&device);
  if (e != hipSuccess)
    return device_fail("hipGetDevice", e);
  if (device < 0 || device >= kMaxDevices)
    return refuse("sf_hip_find_static: device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_lookup);
  if (!g_scratch[device]) {
    void *address = nullptr;
    e = hipGetSymbolAddress(&address, HIP_SYMBOL(scratch));
    if (e != hipSuccess)
      return device_fail("hipGetSymbolAddress(scratch)", e);
    g_scratch[device] = static_cast<uint8_t *>(address);
  }
  *out = g_scratch[device];
  return SF_OK;
}

size_t piece(size_t bytes) { return (bytes + 255) & ~(size_t)255; }

// The descriptor into device memory from a kernel argument: no host
// staging buffer and no copy node, so the call can be captured.
__global__ void put_desc(sf::FrameDesc d, sf::FrameDesc *to) {
  if (threadIdx.x == 0)
    *to = d;
}

} // namespace

extern "C" {

int sf_hip_find_static(int pixel_bytes, const void *frame, size_t height,
                       size_t width, sf_signal_pixel *out, size_t capacity,
                       uint32_t *found, void *stream) {
  if (pixel_bytes != 2 && pixel_bytes != 4)
    return refuse("sf_hip_find_static: 16- or 32-bit frames only");
  if (!frame || !out || !found || height == 0 || width == 0 ||
      capacity == 0)
    return refuse("sf_hip_find_static: null pointer, zero size or no list");
  if (height > 0x7fffffffu / width)
    return refuse("sf_hip_find_static: frame larger than INT32_MAX pixels");
  sf::FrameDesc d;
  std::memset(&d, 0, sizeof(d));
  sf::fused_layout(pixel_bytes, (uint32_t)height, (uint32_t)width, d);
  // the list holds the caller's entries, at most the reference's device
  // list (height * width * sizeof(T) / 16, src/dext_cuda.cu:614-616)
  d.capacity = (uint32_t)std::min<size_t>(capacity, d.capacity);
  const size_t cap = std::max<size_t>(d.capacity, 1);
  // the layout of `scratch` for this frame
  size_t at = 0;
  const size_t desc_at = at;
  at += piece(sizeof(sf::FrameDesc));
  const size_t counter_at = at;
  at += piece(sizeof(uint32_t));
  const size_t staging_at = at;
  at += piece(cap * sizeof(sf_signal_pixel));
  size_t tables_at = at, base_at = 0, band_at = 0, count_at = 0;
  if (pixel_bytes == 2) {
    at += piece(sf::sparse_bytes(d.height, d.width, d.capacity));
  } else {
    const size_t cells = sf::cell_entries(d.height, d.width);
    base_at = at;
    at += piece(cells * sizeof(uint32_t));
    band_at = at;
    at += piece(((size_t)d.n_bands + 1) * sizeof(uint32_t));
    count_at = at;
    at += piece(cells);
  }
  if (at > kScratchBytes) {
    char text[200];
    std::snprintf(text, sizeof(text),
                  "sf_hip_find_static: a %zu x %zu frame of %d-byte pixels "
                  "needs %zu bytes of scratch; the library holds %zu",
                  height, width, pixel_bytes, at, kScratchBytes);
    return refuse(text);
  }
  uint8_t *base = nullptr;
  const int status = scratch_of_device(&base);
  if (status != SF_OK)
    return status;
  sf::ensure_device_tables();
  d.pixels = frame;
  const uintptr_t addr = reinterpret_cast<uintptr_t>(frame);
  d.vec = pixel_bytes == 2 ? (addr % 4 == 0 && d.width % 2 == 0 ? 1u : 0u)
                           : (addr % 8 == 0 && d.width % 2 == 0 ? 1u : 0u);
  d.item_begin = 0;
  d.counter = reinterpret_cast<uint32_t *>(base + counter_at);
  d.out_count = found;
  d.staging = reinterpret_cast<sf_signal_pixel *>(base + staging_at);
  d.out = out;
  if (pixel_bytes == 2) {
    sf::sparse_carve(base + tables_at, d);
  } else {
    d.cell_base = reinterpret_cast<uint32_t *>(base + base_at);
    d.band_total = reinterpret_cast<uint32_t *>(base + band_at);
    d.cell_count = base + count_at;
  }
  sf::FrameDesc *desc = reinterpret_cast<sf::FrameDesc *>(base + desc_at);
  const hipStream_t s = static_cast<hipStream_t>(stream);
  put_desc<<<1, 64, 0, s>>>(d, desc);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return device_fail("sf_hip_find_static: writing the descriptor", e);
  e = sf::launch_batch(pixel_bytes, d.vec != 0, d.vec == 0, desc, 1,
                       d.n_strips * d.n_bands, d.n_bands, d.n_cell_rows, s);
  if (e != hipSuccess)
    return device_fail("sf_hip_find_static: launching the threshold", e);
  return SF_OK;
}

} // extern "C"
