// Host driver for the sanitizer build of the CPU oracle (oracle/hdsr_oracle.cpp).  Arguments:
// <dir> <width> <height> <sh_degree> <flags...>.  <dir> holds positions.bin, scales.bin,
// rotations.bin, opacities.bin, sorted.bin, materials.bin, shdata.bin and uniforms.bin (raw
// little-endian arrays); every array is read into a heap buffer of exactly its size.  One frame
// is run per flags value and every output is copied out through orc_size / orc_get.  Prints
// "<flags> <refs> <rgba checksum>" per frame; exits 1 if the oracle rejects an input.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <string>
#include <vector>

struct orc_splats {
  const float* pos;
  const float* logscale;
  const float* rot;
  const float* opacity;
  const uint32_t* sorted;
  const uint32_t* material;
  const float* normals;
  uint32_t normals_len;
  const float* shdata;
  uint32_t n;
  uint32_t sh_degree;
};

extern "C" {
void* orc_run(const orc_splats* s, const void* uniforms256, uint32_t width, uint32_t height, uint32_t flags);
uint64_t orc_size(void* h, int which);
int orc_get(void* h, int which, void* dst, uint64_t bytes);
void orc_free(void* h);
}

template <class T>
static std::vector<T> load(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  std::vector<char> raw((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  std::vector<T> out(raw.size() / sizeof(T));
  if (!out.empty()) std::copy(raw.begin(), raw.begin() + out.size() * sizeof(T), reinterpret_cast<char*>(out.data()));
  return out;
}

int main(int argc, char** argv) {
  if (argc < 6) return 2;
  const std::string d = argv[1];
  const uint32_t w = std::strtoul(argv[2], nullptr, 10), h = std::strtoul(argv[3], nullptr, 10);
  const uint32_t deg = std::strtoul(argv[4], nullptr, 10);
  auto pos = load<float>(d + "/positions.bin"), sc = load<float>(d + "/scales.bin");
  auto rot = load<float>(d + "/rotations.bin"), op = load<float>(d + "/opacities.bin");
  auto sh = load<float>(d + "/shdata.bin"), u = load<float>(d + "/uniforms.bin");
  auto sorted = load<uint32_t>(d + "/sorted.bin"), mat = load<uint32_t>(d + "/materials.bin");
  orc_splats s{pos.data(), sc.data(), rot.data(), op.data(), sorted.data(), mat.data(), nullptr, 0, sh.data(),
               static_cast<uint32_t>(pos.size() / 3), deg};
  for (int i = 5; i < argc; ++i) {
    const uint32_t flags = std::strtoul(argv[i], nullptr, 10);
    void* fr = orc_run(&s, u.data(), w, h, flags);
    if (!fr) return 1;
    uint32_t refs = 0;
    double sum = 0.0;
    for (int which = 0; which <= 20; ++which) {
      const uint64_t n = orc_size(fr, which);
      std::vector<uint8_t> buf(n);
      if (orc_get(fr, which, buf.data(), n) != 0) return 1;
      if (which == 16 && n >= 4) refs = *reinterpret_cast<const uint32_t*>(buf.data());
      if (which == 12)
        for (uint64_t k = 0; k + 4 <= n; k += 4) sum += *reinterpret_cast<const float*>(buf.data() + k);
    }
    orc_free(fr);
    std::printf("%u %u %.6f\n", flags, refs, sum);
  }
  return 0;
}
