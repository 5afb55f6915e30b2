// Host driver for the sanitizer build of the PLY decoder (csrc/assets.cpp).  For each file
// argument: copy the bytes into a heap buffer of exactly the file's length (so a read past the
// end is an AddressSanitizer error), inspect the header, then decode into output arrays sized
// exactly from the header's counts.  Prints "<rc> <count> <message>" per file; exits 0 unless a
// sanitizer stops the process.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <vector>

#include "hdsr.h"
#include "hdsr_assets.h"

int main(int argc, char** argv) {
  for (int i = 1; i < argc; ++i) {
    std::ifstream f(argv[i], std::ios::binary);
    std::vector<char> raw((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    uint8_t* bytes = new uint8_t[raw.size() ? raw.size() : 1];
    if (!raw.empty()) std::memcpy(bytes, raw.data(), raw.size());
    char err[256] = {0};
    hdsr_ply_info info;
    int rc = hdsr_ply_inspect(argv[i], bytes, raw.size(), &info, err, sizeof(err));
    uint32_t count = 0;
    if (rc == 0) {
      count = info.count;
      const size_t n = info.count, k = info.sh_coefficient_count;
      std::vector<float> pos(3 * n), col(3 * n), op(n), rad(n), sc(3 * n), rot(4 * n), sh(3 * k * n);
      std::vector<float> nrm, rough, metal, emis, dn;
      std::vector<uint32_t> ids(n);
      hdsr_splat_attributes a;
      std::memset(&a, 0, sizeof(a));
      a.positions = pos.data();
      a.colors = col.data();
      a.opacities = op.data();
      a.radii = rad.data();
      a.scales = sc.data();
      a.rotations = rot.data();
      a.sh_coefficients = k ? sh.data() : nullptr;
      a.original_ids = ids.data();
      if (info.fields & HDSR_PLY_NORMALS) { nrm.resize(3 * n); a.normals = nrm.data(); }
      if (info.fields & HDSR_PLY_ROUGHNESS) { rough.resize(n); a.roughness = rough.data(); }
      if (info.fields & HDSR_PLY_METALNESS) { metal.resize(n); a.metalness = metal.data(); }
      if (info.fields & HDSR_PLY_EMISSIVE) { emis.resize(3 * n); a.emissive = emis.data(); }
      if (info.fields & HDSR_PLY_DETAIL_NORMALS) { dn.resize(3 * n); a.detail_normals = dn.data(); }
      rc = hdsr_ply_decode(argv[i], bytes, raw.size(), &a, err, sizeof(err));
    }
    for (char* c = err; *c; ++c)  // messages may echo header bytes: one line per file
      if (static_cast<unsigned char>(*c) < 0x20 || static_cast<unsigned char>(*c) >= 0x7f) *c = '?';
    std::printf("%d %u %s\n", rc, count, err);
    delete[] bytes;
  }
  return 0;
}
