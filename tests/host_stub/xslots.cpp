// The X sweep's slot walker (fvs_xslots.hpp) on the host: every lane of
// every persistent wave walks its chunks exactly as SweepXKernel does (one
// Place for the first chunk, Next per stride, lane 0's row as the wave base).
// Checked for each (nx, rows, waves): a valid lane sits on its true slot,
// every load index lies inside the nx * rows cells, the owned real slots
// write each cell exactly once, and no row offset wraps.  Prints the number
// of violations (0 expected).
#include "fvs_xslots.hpp"

#include <cstdio>
#include <vector>

using fvs::XSlot;
using fvs::XSlotWalker;

static long Check(int64_t nx, int64_t rows, int64_t nwaves) {
  const uint32_t S = static_cast<uint32_t>(nx + 2);
  const uint32_t Q = static_cast<uint32_t>(rows) * S;
  const int64_t nchunks = (static_cast<int64_t>(Q) + fvs::kXOwned - 1) / fvs::kXOwned;
  const int64_t ncell = nx * rows;
  std::vector<int> written(ncell, 0);
  long bad = 0;
  for (int64_t w = 0; w < nwaves; ++w) {
    std::vector<XSlotWalker> lanes;
    for (int l = 0; l < 64; ++l) lanes.emplace_back(nx, Q, nwaves, l);
    int64_t c = w;
    for (auto& L : lanes) L.Place(c < nchunks ? c : 0);
    for (; c < nchunks; c += nwaves) {
      bool valid[64];
      uint32_t row[64], sl[64];
      for (int l = 0; l < 64; ++l) lanes[l].RowSlot(c, &valid[l], &row[l], &sl[l]);
      const uint32_t row0 = row[0];
      for (int l = 0; l < 64; ++l) {
        const XSlot s = lanes[l].Make(valid[l], row[l], sl[l], row0);
        if (row[l] < row0) ++bad;  // the unsigned offset would wrap
        if (valid[l] && static_cast<int64_t>(row[l]) * S + sl[l] != lanes[l].Slot(c)) ++bad;
        const int64_t cl = s.cell < 0 ? 0 : (s.cell >= nx ? nx - 1 : s.cell);
        const int64_t ld = s.wbase + static_cast<int64_t>(s.rbase) + cl;
        if (ld < 0 || ld >= ncell) ++bad;
        const bool owned = l >= 2 && l < 2 + fvs::kXOwned;
        const bool real = s.valid && s.cell >= 0 && s.cell < nx;
        if (owned && real) {
          const int64_t st = s.wbase + static_cast<int64_t>(s.rbase) + s.cell;
          if (st < 0 || st >= ncell) ++bad;
          else ++written[st];
        }
      }
      const int64_t cn = c + nwaves;
      for (auto& L : lanes) L.Next(c, cn, nchunks);
    }
  }
  for (int64_t i = 0; i < ncell; ++i)
    if (written[i] != 1) ++bad;
  return bad;
}

int main() {
  const int64_t nxs[] = {1, 2, 3, 5, 57, 58, 59, 60, 61, 62, 63, 64, 65, 100, 127, 256};
  const int64_t rowss[] = {1, 2, 3, 7, 31, 333};
  long bad = 0, cases = 0;
  for (int64_t nx : nxs)
    for (int64_t rows : rowss) {
      const int64_t slots = (nx + 2) * rows;
      const int64_t waves = (slots + fvs::kXOwned - 1) / fvs::kXOwned;
      const int64_t full = ((waves + 3) / 4) * 4;  // blocks of four waves
      const int64_t counts[] = {full, 4, 8, 12, 20, 4096};
      for (int64_t nwaves : counts) {
        const long b = Check(nx, rows, nwaves);
        if (b) std::printf("nx %lld rows %lld waves %lld: %ld violations\n",
                           static_cast<long long>(nx), static_cast<long long>(rows),
                           static_cast<long long>(nwaves), b);
        bad += b;
        ++cases;
      }
    }
  std::printf("%ld %ld\n", bad, cases);
  return bad != 0;
}
