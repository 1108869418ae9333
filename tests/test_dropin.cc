// The C++ drop-in, as a caller of the reference's interfaces would use it:
// gpu::find<T> over the reference's device-test frames under all four window
// combinations (the loop of the reference's tests/test_dext_gpu.cc:179-199),
// and dials_spots::Labeller over a short series, each against the CPU oracle
// (oracle/dext_oracle.c) with no tolerance. Exits 77 (a skip) when there is
// no device, 0 when every comparison agrees, 1 otherwise.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "dext_gpu.hh"
#include "dials_spots_gpu.hh"
#include "refl_gpu.hh"
#include "dext_oracle.h"

namespace {

struct Frame {
  std::string name;
  std::size_t height, width;
  std::vector<std::uint16_t> pixels;
};

Frame make(const char *name, std::size_t h, std::size_t w, std::uint64_t seed,
           double background, bool mask) {
  Frame f{name, h, w, std::vector<std::uint16_t>(h * w)};
  oracle_synthetic_frame(f.pixels.data(), h, w, seed, background, mask ? 1 : 0);
  return f;
}

std::vector<oracle_signal_pixel> oracle(const Frame &f) {
  std::vector<oracle_signal_pixel> out(f.height * f.width);
  std::size_t found = 0;
  oracle_dext_u16(f.pixels.data(), f.height, f.width, out.data(), out.size(),
                  &found);
  out.resize(found);
  return out;
}

bool same(const std::vector<SignalPixel> &a,
          const std::vector<oracle_signal_pixel> &b) {
  return a.size() == b.size() &&
         (a.empty() ||
          std::memcmp(a.data(), b.data(), a.size() * sizeof(SignalPixel)) == 0);
}

} // namespace

int main() {
  if (!gpu::available()) {
    std::printf("no %s device available; skipping\n", gpu::backend());
    return 77;
  }
  gpu::report_windows(false);
  std::vector<Frame> frames;
  frames.push_back(make("exact tile", 32, 32, 1, 12.0, false));
  frames.push_back(make("under one tile", 19, 27, 2, 8.0, false));
  frames.push_back(make("ragged", 200, 301, 3, 15.0, false));
  frames.push_back(make("ragged, masked", 200, 301, 4, 15.0, true));
  frames.push_back(make("dark", 128, 128, 5, 1.0, false));
  frames.push_back(make("module", 512, 1030, 6, 20.0, true));
  Frame all{"entirely masked", 64, 64, std::vector<std::uint16_t>(64 * 64)};
  for (auto &p : all.pixels)
    p = 0xffff;
  frames.push_back(all);

  int failures = 0, checks = 0;
  const gpu::Window windows[] = {gpu::Window::Direct, gpu::Window::Tile};
  std::vector<SignalPixel> device;
  for (const gpu::Window w0 : windows)
    for (const gpu::Window w2 : windows) {
      gpu::stage0_window(w0);
      gpu::stage2_window(w2);
      for (const Frame &f : frames) {
        checks++;
        const int status = gpu::find<std::uint16_t>(f.pixels.data(), device,
                                                    f.height, f.width);
        if (status != 0 || !same(device, oracle(f))) {
          std::printf("FAIL %s (stage0 %s, stage2 %s): status %d, %zu px\n",
                      f.name.c_str(), gpu::name(w0), gpu::name(w2), status,
                      device.size());
          failures++;
        }
      }
    }
  gpu::stage0_window(gpu::Window::Tile);
  gpu::stage2_window(gpu::Window::Tile);

  // -1 for the arguments src/dext_cuda.cu:588-593 refuses
  checks++;
  if (gpu::find<std::uint16_t>(nullptr, device, 4, 4) != -1 ||
      gpu::find<std::uint16_t>(frames[0].pixels.data(), device, 0, 4) != -1) {
    std::printf("FAIL bad arguments are not -1\n");
    failures++;
  }

  // the grouping boundary: five frames of a module-sized series
  {
    const std::size_t h = 512, w = 1030;
    dials_spots::Options options;
    dials_spots::Labeller device_labeller(h, w, options);
    oracle_labeller *cpu = oracle_labeller_new(h, w, 3, 1000, 2.0, 0);
    for (int z = 0; z < 5; z++) {
      const Frame f = make("series", h, w, 60 + (std::uint64_t)z, 20.0, true);
      gpu::find<std::uint16_t>(f.pixels.data(), device, h, w);
      device_labeller.add(z, device);
      const auto expected = oracle(f);
      oracle_labeller_add(cpu, z, expected.data(), expected.size());
    }
    device_labeller.finish();
    oracle_labeller_finish(cpu);
    std::vector<oracle_spot> spots(oracle_labeller_spot_count(cpu));
    std::vector<oracle_pixel> pixels(oracle_labeller_pixel_count(cpu));
    oracle_counts counts;
    oracle_labeller_results(cpu, spots.data(), pixels.data(), &counts);
    checks++;
    const auto &ds = device_labeller.spots();
    const auto &dp = device_labeller.pixels();
    const bool ok =
        ds.size() == spots.size() && dp.size() == pixels.size() &&
        std::memcmp(ds.data(), spots.data(), ds.size() * sizeof(oracle_spot)) ==
            0 &&
        std::memcmp(dp.data(), pixels.data(),
                    dp.size() * sizeof(oracle_pixel)) == 0 &&
        std::memcmp(&device_labeller.counts(), &counts, sizeof(counts)) == 0;
    if (!ok || spots.empty()) {
      std::printf("FAIL labeller: %zu spots vs %zu\n", ds.size(), spots.size());
      failures++;
    }
    // the table writer: the device tables through the refl:: interface,
    // the oracle's through the C entry point, to the same bytes
    {
      const std::string a = "/tmp/sf_dropin_device.refl";
      const std::string b = "/tmp/sf_dropin_oracle.refl";
      refl::Options ro;
      ro.identifier = "dropin";
      refl::write(a, ds, dp, w, ro);
      sf_refl_options co{0, 0, 1, "dropin"};
      const int status = sf_hip_refl_write(
          b.c_str(), reinterpret_cast<const sf_spot *>(spots.data()),
          spots.size(), reinterpret_cast<const sf_pixel *>(pixels.data()),
          pixels.size(), w, &co);
      auto slurp = [](const std::string &path) {
        std::vector<char> bytes;
        if (std::FILE *f = std::fopen(path.c_str(), "rb")) {
          char chunk[4096];
          std::size_t n;
          while ((n = std::fread(chunk, 1, sizeof chunk, f)) > 0)
            bytes.insert(bytes.end(), chunk, chunk + n);
          std::fclose(f);
        }
        return bytes;
      };
      const std::vector<char> fa = slurp(a), fb = slurp(b);
      bool zero_width_threw = false;
      try {
        refl::write(a, ds, dp, 0, ro);
      } catch (const std::exception &) {
        zero_width_threw = true;
      }
      std::remove(a.c_str());
      std::remove(b.c_str());
      checks++;
      if (status != SF_OK || fa.empty() || fa != fb || !zero_width_threw) {
        std::printf("FAIL refl: status %d, %zu vs %zu bytes\n", status,
                    fa.size(), fb.size());
        failures++;
      }
    }
    bool threw = false;
    try {
      device_labeller.add(2, device);
    } catch (const std::exception &) {
      threw = true;
    }
    checks++;
    if (!threw) {
      std::printf("FAIL a frame out of order did not throw\n");
      failures++;
    }
    oracle_labeller_free(cpu);
  }
  std::printf("%s: %d of %d checks agree\n", failures ? "FAILED" : "PASS",
              checks - failures, checks);
  return failures ? 1 : 0;
}
