// Mutation fuzz of the bitshuffle chunk decoder (csrc/bitshuffle.cpp), host
// only, built by tests/test_bitshuffle_fuzz.py with g++ and AddressSanitizer +
// UndefinedBehaviorSanitizer (or ThreadSanitizer).
//
//   bshuf_fuzz <cases file> <iterations> <seed>
//
// The cases file holds valid LZ4-mode containers with their decoded bytes
// (int64 count; per case int64 element size, decoded length, container length,
// then the container and the decoded bytes).  Each valid case must decode to
// its bytes.  Then every iteration takes a case and damages it (bit flips,
// byte overwrites, truncation, appended bytes), sometimes with a wrong output
// length, element size or the uncompressed mode, and decodes it from an
// exact-size heap copy into an exact-size output followed by a guard byte.
// The decoder may accept or refuse; the sanitizers report any read or write
// outside the buffers, and the guard byte any write past the output.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

namespace zdt {
int fail(int code, const char*, ...) { return code; }
}  // namespace zdt

extern "C" int zdt_bshuf_decompress(const void* in, size_t in_len, void* out, size_t out_len, size_t elem_size,
                                    size_t block_size, int compression);

namespace {

struct Case {
  int64_t E = 0;
  std::vector<uint8_t> c, raw;
};

bool read_cases(const char* path, std::vector<Case>& cs) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  int64_t nc = 0;
  bool good = std::fread(&nc, 8, 1, f) == 1 && nc > 0;
  for (int64_t i = 0; good && i < nc; ++i) {
    int64_t h[3];
    good = std::fread(h, 8, 3, f) == 3 && h[0] > 0 && h[1] >= 0 && h[2] >= 0;
    if (!good) break;
    Case k;
    k.E = h[0];
    k.c.resize(static_cast<size_t>(h[2]));
    k.raw.resize(static_cast<size_t>(h[1]));
    good = std::fread(k.c.data(), 1, k.c.size(), f) == k.c.size() &&
           std::fread(k.raw.data(), 1, k.raw.size(), f) == k.raw.size();
    cs.push_back(std::move(k));
  }
  std::fclose(f);
  return good;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s <cases> <iterations> <seed>\n", argv[0]);
    return 2;
  }
  std::vector<Case> cs;
  if (!read_cases(argv[1], cs)) {
    std::fprintf(stderr, "cannot read %s\n", argv[1]);
    return 2;
  }
  const long iters = std::atol(argv[2]);
  std::mt19937_64 rng(std::strtoull(argv[3], nullptr, 10));
  for (const Case& k : cs) {
    std::vector<uint8_t> out(k.raw.size());
    if (zdt_bshuf_decompress(k.c.data(), k.c.size(), out.data(), out.size(), k.E, 0, 2) != 0 || out != k.raw) {
      std::printf("a valid container did not decode to its bytes\n");
      return 1;
    }
  }
  long accepted = 0, refused = 0;
  for (long it = 0; it < iters; ++it) {
    const Case& k = cs[rng() % cs.size()];
    std::vector<uint8_t> in = k.c;
    const int kind = static_cast<int>(rng() % 6);
    const int nm = 1 + static_cast<int>(rng() % 4);
    for (int m = 0; m < nm; ++m) {
      if (kind <= 2 && !in.empty()) in[rng() % in.size()] ^= static_cast<uint8_t>(1u << (rng() % 8));
      else if (kind == 3 && !in.empty()) in.resize(rng() % in.size());
      else if (kind == 4) in.push_back(static_cast<uint8_t>(rng()));
      else if (!in.empty()) in[rng() % in.size()] = static_cast<uint8_t>(rng());
    }
    size_t out_len = k.raw.size();
    size_t E = static_cast<size_t>(k.E);
    if (rng() % 8 == 0) out_len = out_len + E * (rng() % 3) - E * (rng() % 2);
    if (rng() % 16 == 0) E = 1 + rng() % 8;
    const int comp = rng() % 10 == 0 ? 0 : 2;
    const size_t bs = comp == 0 ? (rng() % 3) * 64 : 0;
    std::vector<uint8_t> out(out_len + 1, 0xAB);  // the last byte is a guard
    const int rc = zdt_bshuf_decompress(in.empty() ? nullptr : in.data(), in.size(), out.data(), out_len, E, bs, comp);
    if (out[out_len] != 0xAB) {
      std::printf("iteration %ld wrote past the output\n", it);
      return 1;
    }
    (rc == 0 ? accepted : refused) += 1;
  }
  std::printf("accepted %ld refused %ld\n", accepted, refused);
  return 0;
}
