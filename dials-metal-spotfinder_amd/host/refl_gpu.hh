// The reference's reflection-table writer interface (src/refl.hh, namespace
// refl) over sf_hip_refl_write. A caller written against refl::write --
// src/find_spots.cc's table output -- switches by including this header
// instead (with dials_spots_gpu.hh for the tables) and linking -ldext_hip.
//
// The file is byte for byte the one the reference writes for the same
// tables and options. write() throws std::runtime_error where the
// reference's does (a zero width, a file that cannot be written, a column
// past msgpack's 4 GB binary limit: src/refl.cc:160-164, 205-211), and also
// for a spot whose pixels lie outside the pixel table or its box, which the
// reference would read past.
#ifndef SPOTFINDER_REFL_GPU_HH
#define SPOTFINDER_REFL_GPU_HH

#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

#include "dials_spots_gpu.hh"
#include "refl_hip.h"

namespace refl {

// src/refl.hh:52-63: the same fields and defaults.
struct Options {
  std::size_t panel = 0;
  int id = 0;
  std::string identifier;
  bool shoeboxes = true;
};

// src/refl.hh:65-74.
inline void write(const std::string &path,
                  const std::vector<dials_spots::Spot> &spots,
                  const std::vector<dials_spots::Pixel> &pixels,
                  std::size_t width, const Options &options) {
  static_assert(sizeof(dials_spots::Spot) == sizeof(sf_spot) &&
                    sizeof(dials_spots::Pixel) == sizeof(sf_pixel),
                "the tables are the C ABI's, element for element");
  sf_refl_options o;
  o.panel = options.panel;
  o.id = options.id;
  o.shoeboxes = options.shoeboxes ? 1 : 0;
  o.identifier = options.identifier.c_str();
  const int status = sf_hip_refl_write(
      path.c_str(), reinterpret_cast<const sf_spot *>(spots.data()),
      spots.size(), reinterpret_cast<const sf_pixel *>(pixels.data()),
      pixels.size(), width, &o);
  if (status != SF_OK)
    throw std::runtime_error(std::string("refl: ") +
                             sf_hip_refl_last_error());
}

} // namespace refl

#endif // SPOTFINDER_REFL_GPU_HH
