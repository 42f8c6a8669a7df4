// Runs the device-wide exclusive scan (scan_exclusive's launch sequence and grids) on the CPU
// through the HIP-model emulation, in place and out of place, with poisoned partials. Test
// infrastructure only.
// usage: scan_emu <in> <out> [direct_max]; <in> holds int64 n then n int32, <out> gets the in-place
// then the out-of-place result (2n int32); direct_max (default kScanTile) is the largest partial
// count scanned by the two-launch form, 0 forces the three-launch form.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "launch_emu.h"

using namespace clue;
using namespace emu_host;

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int64_t n = 0;
  if (std::fread(&n, sizeof n, 1, f) != 1 || n < 0) return 2;
  std::vector<int> v((size_t)n);
  if (n && std::fread(v.data(), 4, (size_t)n, f) != (size_t)n) return 2;
  std::fclose(f);
  const int P = 0x5A5A5A5A;
  std::vector<int> part((size_t)scan_partials(n > 0 ? n : 1), P);
  const int directMax = argc > 3 ? std::atoi(argv[3]) : kScanTile;
  std::vector<int> inplace = v, out((size_t)n, P);
  scan_exclusive(inplace.data(), inplace.data(), n, part.data(), directMax);
  std::fill(part.begin(), part.end(), P);
  scan_exclusive(v.data(), out.data(), n, part.data(), directMax);
  FILE* g = std::fopen(argv[2], "wb");
  if (!g) return 2;
  std::fwrite(inplace.data(), 4, (size_t)n, g);
  std::fwrite(out.data(), 4, (size_t)n, g);
  std::fclose(g);
  return 0;
}
