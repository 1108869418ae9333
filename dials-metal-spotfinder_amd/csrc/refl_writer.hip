// Writing a DIALS reflection table (.refl): sf_hip_refl_write, the
// reference's refl::write (src/refl.hh:21-78, src/refl.cc:205-351). Host code
// only: the spot and pixel tables are already on the host when a table is
// written, and the work is a stream of small headers around column dumps.
//
// The layout has to be the reference's byte for byte (it is DIALS' format):
// one msgpack document, columns in name order, each column a type name and a
// raw little-endian dump, every header in msgpack's shortest form. The
// shoebox of a spot is its panel, bbox, a version byte (2), then float data,
// a byte mask (valid | foreground at the spot's pixels) and a zero float
// background over the dense (z, y, x) box.
#include <bit>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/refl_hip.h"

namespace {

thread_local std::string g_refl_error;

static_assert(sizeof(sf_spot) == 104 && sizeof(sf_pixel) == 12,
              "table layouts");
static_assert(sizeof(float) == 4 && sizeof(double) == 8 &&
                  sizeof(std::size_t) == 8,
              "the columns are fixed-width memory images");
static_assert(std::endian::native == std::endian::little,
              "the columns are little-endian memory images");

// DIALS' codes: the Strong flag (dials::af::Flags) and the shoebox mask's
// valid and foreground bits (dials/model/data/mask_code.h).
constexpr uint64_t kStrongFlag = 1u << 5;
constexpr uint8_t kValidForeground = (1u << 0) | (1u << 2);

// Buffered output; `ok` goes false at the first failed write.
struct Sink {
  std::FILE *file = nullptr;
  std::vector<uint8_t> buf;
  bool ok = true;

  void flush() {
    if (ok && !buf.empty() &&
        std::fwrite(buf.data(), 1, buf.size(), file) != buf.size())
      ok = false;
    buf.clear();
  }
  void put(const void *data, size_t n) {
    const uint8_t *p = static_cast<const uint8_t *>(data);
    buf.insert(buf.end(), p, p + n);
    if (buf.size() >= (1u << 20))
      flush();
  }
  void byte(uint8_t b) { buf.push_back(b); }
  template <typename T> void raw(T v) { put(&v, sizeof(T)); }
};

// A msgpack header kind: a one-byte form for values below `fix_below`
// (`fix` | value), else the first tagged form whose big-endian length field
// holds the value.
struct Form {
  uint8_t fix;
  uint64_t fix_below;
  int forms;
  uint8_t tag[4];
  int width[4];
};
constexpr Form kUint = {0x00, 0x80, 4, {0xcc, 0xcd, 0xce, 0xcf}, {1, 2, 4, 8}};
constexpr Form kStr = {0xa0, 32, 3, {0xd9, 0xda, 0xdb, 0}, {1, 2, 4, 0}};
constexpr Form kArray = {0x90, 16, 2, {0xdc, 0xdd, 0, 0}, {2, 4, 0, 0}};
constexpr Form kMap = {0x80, 16, 2, {0xde, 0xdf, 0, 0}, {2, 4, 0, 0}};
constexpr Form kBin = {0x00, 0, 3, {0xc4, 0xc5, 0xc6, 0}, {1, 2, 4, 0}};

// Writes the header; false if no form holds the value.
bool header(Sink &s, const Form &f, uint64_t n) {
  if (n < f.fix_below) {
    s.byte((uint8_t)(f.fix | n));
    return true;
  }
  for (int i = 0; i < f.forms; i++) {
    const int w = f.width[i];
    if (w == 8 || n >> (8 * w) == 0) {
      s.byte(f.tag[i]);
      for (int b = w - 1; b >= 0; b--)
        s.byte((uint8_t)(n >> (8 * b)));
      return true;
    }
  }
  return false;
}

void text(Sink &s, const char *t) {
  const size_t n = std::strlen(t);
  header(s, kStr, n);
  s.put(t, n);
}

// A spot's box extents, and whether each of its pixels lies in its box and
// in the pixel table.
struct Box {
  uint64_t x, y, z;
  uint64_t voxels() const { return x * y * z; }
};
bool spot_box(const sf_spot &sp, const sf_pixel *pixels, size_t n_pixels,
              size_t width, Box &box) {
  if (sp.bbox[1] <= sp.bbox[0] || sp.bbox[3] <= sp.bbox[2] ||
      sp.bbox[5] <= sp.bbox[4])
    return false;
  box.x = (uint64_t)((int64_t)sp.bbox[1] - sp.bbox[0]);
  box.y = (uint64_t)((int64_t)sp.bbox[3] - sp.bbox[2]);
  box.z = (uint64_t)((int64_t)sp.bbox[5] - sp.bbox[4]);
  if (sp.first > n_pixels || sp.n_signal > n_pixels - sp.first)
    return false;
  for (uint32_t k = 0; k < sp.n_signal; k++) {
    const sf_pixel &px = pixels[sp.first + k];
    const int64_t x = (int64_t)(px.index % width) - sp.bbox[0];
    const int64_t y = (int64_t)(px.index / width) - sp.bbox[2];
    const int64_t z = (int64_t)px.frame - sp.bbox[4];
    if (x < 0 || y < 0 || z < 0 || (uint64_t)x >= box.x ||
        (uint64_t)y >= box.y || (uint64_t)z >= box.z)
      return false;
  }
  return true;
}

// One column: [ type, [ rows, bin(bytes) ] ] after its name.
bool column(Sink &s, const char *name, const char *type, size_t rows,
            uint64_t bytes) {
  text(s, name);
  header(s, kArray, 2);
  text(s, type);
  header(s, kArray, 2);
  header(s, kUint, rows);
  return header(s, kBin, bytes);
}

int fail(int status, const std::string &what) {
  g_refl_error = what;
  return status;
}

} // namespace

extern "C" int sf_hip_refl_write(const char *path, const sf_spot *spots,
                                 size_t n_spots, const sf_pixel *pixels,
                                 size_t n_pixels, size_t width,
                                 const sf_refl_options *options) {
  if (!path || !options || (n_spots && !spots) || (n_pixels && !pixels))
    return fail(SF_BAD_ARGUMENT, "sf_hip_refl_write: a null argument");
  if (width == 0)
    return fail(SF_BAD_ARGUMENT, "sf_hip_refl_write: a frame width of zero");
  const bool boxes = options->shoeboxes != 0;
  // every check before the file is opened
  uint64_t box_bytes = 0;
  std::vector<Box> box(boxes ? n_spots : 0);
  for (size_t i = 0; boxes && i < n_spots; i++) {
    if (!spot_box(spots[i], pixels, n_pixels, width, box[i]))
      return fail(SF_BAD_ARGUMENT,
                  "sf_hip_refl_write: spot " + std::to_string(i) +
                      " has pixels outside its box or the pixel table");
    box_bytes += 4 + 6 * 4 + 1 + box[i].voxels() * (4 + 1 + 4);
  }
  const uint64_t bin_max = 0xFFFFFFFFull;
  if (box_bytes > bin_max || (uint64_t)n_spots * 24 > bin_max)
    return fail(SF_IO_ERROR,
                "sf_hip_refl_write: a column would pass msgpack's 4 GB binary "
                "limit; write without shoeboxes");

  Sink s;
  s.file = std::fopen(path, "wb");
  if (!s.file)
    return fail(SF_IO_ERROR, std::string("cannot write ") + path);
  s.buf.reserve((1u << 20) + 4096);
  const size_t rows = n_spots;

  header(s, kArray, 3);
  text(s, "dials::af::reflection_table");
  header(s, kUint, 2); // the format version DIALS writes
  header(s, kMap, 3);
  text(s, "identifiers");
  const bool named = options->identifier && options->identifier[0];
  header(s, kMap, named ? 1 : 0);
  if (named) {
    header(s, kUint, (uint64_t)(int64_t)options->id);
    text(s, options->identifier);
  }
  text(s, "nrows");
  header(s, kUint, rows);
  text(s, "data");
  header(s, kMap, boxes ? 10 : 9);

  column(s, "bbox", "int6", rows, (uint64_t)rows * 24);
  for (size_t i = 0; i < rows; i++)
    s.put(spots[i].bbox, 24);
  column(s, "flags", "std::size_t", rows, (uint64_t)rows * 8);
  for (size_t i = 0; i < rows; i++)
    s.raw<uint64_t>(kStrongFlag);
  column(s, "id", "int", rows, (uint64_t)rows * 4);
  for (size_t i = 0; i < rows; i++)
    s.raw<int32_t>(options->id);
  column(s, "intensity.sum.value", "double", rows, (uint64_t)rows * 8);
  for (size_t i = 0; i < rows; i++)
    s.raw<double>(spots[i].intensity);
  column(s, "intensity.sum.variance", "double", rows, (uint64_t)rows * 8);
  for (size_t i = 0; i < rows; i++)
    s.raw<double>(spots[i].intensity_variance);
  column(s, "n_signal", "int", rows, (uint64_t)rows * 4);
  for (size_t i = 0; i < rows; i++)
    s.raw<int32_t>((int32_t)spots[i].n_signal);
  column(s, "panel", "std::size_t", rows, (uint64_t)rows * 8);
  for (size_t i = 0; i < rows; i++)
    s.raw<uint64_t>(options->panel);
  if (boxes) {
    column(s, "shoebox", "Shoebox<>", rows, box_bytes);
    std::vector<float> data;
    std::vector<uint8_t> mask;
    for (size_t i = 0; i < rows; i++) {
      const sf_spot &sp = spots[i];
      const Box &b = box[i];
      data.assign(b.voxels(), 0.0f);
      mask.assign(b.voxels(), 0);
      for (uint32_t k = 0; k < sp.n_signal; k++) {
        const sf_pixel &px = pixels[sp.first + k];
        const uint64_t x = px.index % width - (uint64_t)sp.bbox[0];
        const uint64_t y = px.index / width - (uint64_t)sp.bbox[2];
        const uint64_t z = (uint64_t)((int64_t)px.frame - sp.bbox[4]);
        const uint64_t at = (z * b.y + y) * b.x + x; // z slowest, x fastest
        data[at] = (float)px.value;
        mask[at] = kValidForeground;
      }
      s.raw<uint32_t>((uint32_t)options->panel);
      s.put(sp.bbox, 24);
      s.byte(2); // shoebox version 2: a byte mask
      s.put(data.data(), data.size() * sizeof(float));
      s.put(mask.data(), mask.size());
      data.assign(b.voxels(), 0.0f); // the background, never filled
      s.put(data.data(), data.size() * sizeof(float));
    }
  }
  column(s, "xyzobs.px.value", "vec3<double>", rows, (uint64_t)rows * 24);
  for (size_t i = 0; i < rows; i++)
    s.put(spots[i].position, 24);
  column(s, "xyzobs.px.variance", "vec3<double>", rows, (uint64_t)rows * 24);
  for (size_t i = 0; i < rows; i++)
    s.put(spots[i].variance, 24);

  s.flush();
  const bool closed = std::fclose(s.file) == 0;
  if (!s.ok || !closed)
    return fail(SF_IO_ERROR, std::string("cannot finish writing ") + path);
  return SF_OK;
}

extern "C" const char *sf_hip_refl_last_error(void) {
  return g_refl_error.c_str();
}
