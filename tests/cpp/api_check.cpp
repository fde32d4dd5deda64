// Host-only checks of include/flsgpu.hpp (no device calls): prints, per
// binding, encoding and type, how many (unpack_n_vectors, unpack_n_values,
// unpacker, patcher) combinations are provided, and round-trips the enum
// spellings. tests/test_cpp_api.py compares the counts with those of the
// reference's generated bindings.
#include <cstdio>
#include <string>

#include "flsgpu.hpp"

namespace B = bindings::detail;

template <typename T>
int count(B::Fn fn, B::Enc enc) {
  int n = 0;
  for (unsigned nvec : {1u, 2u, 4u, 8u})
    for (unsigned nval : {1u, 32u})
      for (int u = 0; u <= int(enums::Unpacker::StatefulBranchless); ++u)
        for (int p = 0; p <= int(enums::Patcher::PrefetchAllBranchless); ++p)
          n += B::provided<T>(fn, enc, nvec, nval, enums::Unpacker(u), enums::Patcher(p));
  return n;
}

int main() {
  const char *fns[] = {"decompress", "query", "compute", "multi"};
  const char *encs[] = {"BP", "FFOR", "ALP", "ALPExtended"};
  for (int f = 0; f < 4; ++f)
    for (int e = 0; e < 4; ++e) {
      const auto fn = B::Fn(f);
      const auto enc = B::Enc(e);
      if (e < 2) {
        std::printf("%s %s uint32_t %d\n", fns[f], encs[e], count<uint32_t>(fn, enc));
        std::printf("%s %s uint64_t %d\n", fns[f], encs[e], count<uint64_t>(fn, enc));
      } else {
        std::printf("%s %s float %d\n", fns[f], encs[e], count<float>(fn, enc));
        std::printf("%s %s double %d\n", fns[f], encs[e], count<double>(fn, enc));
      }
    }
  int bad = 0;
  for (const char *s : {"dummy", "old-fls", "switch-case", "stateless", "stateful-branchless",
                        "stateful-register-branchless-2"})
    bad += enums::string_to_unpacker(s) == enums::Unpacker::Dummy && std::string(s) != "dummy";
  try {
    enums::string_to_patcher("no-such");
    bad += 1;
  } catch (const std::invalid_argument &) {
  }
  std::printf("enum_errors %d\n", bad);
  return 0;
}
