// The reference's grouping interface (src/dials_spots.hh, namespace
// dials_spots), with the labelling and the centroid/intensity pass on the
// device (libdext_hip.so). A caller written against dials_spots::Labeller
// -- src/find_spots.cc:460-462, 580, 636 -- switches by including this
// header instead and linking -ldext_hip.
//
// After finish(), spots(), pixels() and counts() are identical, row for row
// and bit for bit, to the reference Labeller's. Before finish() they hold
// every component that has closed so far, as the reference's do; queries
// label what has been added (a device synchronisation).
//
// add() throws std::runtime_error where the reference's add() does (a frame
// not above the last one, indices not strictly ascending, an index off the
// frame: src/dials_spots.cc:58-91); the constructor throws where the
// reference's does (src/dials_spots.cc:22-30).
#ifndef SPOTFINDER_DIALS_SPOTS_GPU_HH
#define SPOTFINDER_DIALS_SPOTS_GPU_HH

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "dext_gpu.hh"
#include "dext_hip.h"

namespace dials_spots {

// src/dials_spots.hh:48-106: the same fields, types and order.
struct Pixel {
  std::int32_t frame;
  std::uint32_t index;
  std::uint32_t value;
};

struct Spot {
  std::int32_t bbox[6] = {0, 0, 0, 0, 0, 0};
  double position[3] = {0.0, 0.0, 0.0};
  double variance[3] = {0.0, 0.0, 0.0};
  double intensity = 0.0;
  double intensity_variance = 0.0;
  std::uint32_t n_signal = 0;
  std::uint64_t first = 0;
};

struct Options {
  std::size_t min_spot_size = 3;
  std::size_t max_spot_size = 1000;
  double max_separation = 2.0;
  bool two_d = false;
};

struct Counts {
  std::uint64_t signal_pixels = 0;
  std::uint64_t groups = 0;
  std::uint64_t too_small = 0;
  std::uint64_t too_large = 0;
  std::uint64_t separated = 0;
  std::uint64_t accepted = 0;
};

static_assert(sizeof(Pixel) == sizeof(sf_pixel), "Pixel layout");
static_assert(sizeof(Spot) == sizeof(sf_spot), "Spot layout");
static_assert(sizeof(Counts) == sizeof(sf_counts), "Counts layout");

class Labeller {
public:
  Labeller(std::size_t height, std::size_t width, Options options) {
    sf_options o;
    o.min_spot_size = options.min_spot_size;
    o.max_spot_size = options.max_spot_size;
    o.max_separation = options.max_separation;
    o.two_d = options.two_d ? 1 : 0;
    handle_ = sf_hip_labeller_create(height, width, &o);
    if (!handle_)
      throw std::runtime_error(std::string("dials_spots: ") +
                               sf_hip_labeller_last_error());
  }
  ~Labeller() { sf_hip_labeller_destroy(handle_); }
  Labeller(const Labeller &) = delete;
  Labeller &operator=(const Labeller &) = delete;

  void add(std::int64_t frame, const std::vector<SignalPixel> &signal) {
    check(sf_hip_labeller_add(
        handle_, frame,
        reinterpret_cast<const sf_signal_pixel *>(signal.data()),
        signal.size()));
  }

  void finish() { check(sf_hip_labeller_finish(handle_)); }

  const std::vector<Spot> &spots() const {
    refresh();
    return spots_;
  }
  const std::vector<Pixel> &pixels() const {
    refresh();
    return pixels_;
  }
  const Counts &counts() const {
    refresh();
    return counts_;
  }
  std::size_t live() const { return sf_hip_labeller_live(handle_); }

private:
  static void check(int status) {
    if (status != SF_OK)
      throw std::runtime_error(std::string("dials_spots: ") +
                               sf_hip_labeller_last_error());
  }

  void refresh() const {
    std::size_t n_spots = 0, n_pixels = 0;
    check(sf_hip_labeller_sizes(handle_, &n_spots, &n_pixels));
    spots_.resize(n_spots);
    pixels_.resize(n_pixels);
    check(sf_hip_labeller_results(
        handle_, reinterpret_cast<sf_spot *>(spots_.data()),
        reinterpret_cast<sf_pixel *>(pixels_.data()),
        reinterpret_cast<sf_counts *>(&counts_)));
  }

  sf_labeller *handle_ = nullptr;
  mutable std::vector<Spot> spots_;
  mutable std::vector<Pixel> pixels_;
  mutable Counts counts_;
};

} // namespace dials_spots

#endif // SPOTFINDER_DIALS_SPOTS_GPU_HH
