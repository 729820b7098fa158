// Slot walker of the persistent X sweep (SweepXKernel in fvs_sweep.hip).
//
// Every row of nx cells contributes nx + 2 slots: the ghost cell -1, the
// cells 0 .. nx-1 and the ghost cell nx.  The slots of all rows form one
// sequence of Q slots cut into chunks; lane l of chunk c takes slot
// q = c * kXOwned + l - 2, so lanes 2 .. 61 own their slots and lanes 0, 1,
// 62, 63 overlap the neighbouring chunks.  A persistent wave visits the
// chunks c, c + nwaves, c + 2 nwaves, ...; the (row, slot) of its lane
// advances by the constant stride without a division per chunk.
//
// Addresses are a wave-uniform base (the row of lane 0) plus a 32-bit lane
// offset relative to it.  The offset is unsigned, so no lane's row may lie
// below lane 0's: lanes whose slot falls outside the sequence take its
// nearest end (slot 0 before it, slot Q - 1 after it), which keeps rows
// non-decreasing in the lane index.  The same holds after a stride only if
// the lane's slot in the previous chunk was its true slot: lanes 0 and 1 of
// a wave's first chunk sit on the clamped slot 0, and advancing them from
// there can put lane 0 a row beyond lanes 2 .. 63, whose offsets then wrap
// to about 2^32 and send their loads outside the arrays.  Such lanes are
// placed anew (Next).
//
// The header compiles for the device and, with tests/host_stub, for the
// host, where tests/test_xslots.py walks every lane of every wave over
// ragged shapes and checks that each load and store stays inside the
// arrays and that each cell is written exactly once.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fvs {

constexpr int kXOwned = 60;      // slots owned per chunk (lanes 2 .. 61)
constexpr int kXWavesPerCu = 16;  // persistent waves per CU

struct XSlot {
  bool valid;      // the lane's slot lies in the sequence
  uint32_t sl;     // slot within its row (0 .. nx + 1)
  int64_t cell;    // sl - 1: -1 and nx are the ghost slots
  uint32_t rbase;  // (row - row of lane 0) * nx
  int64_t wbase;   // row of lane 0 * nx (wave-uniform)
};

struct XSlotWalker {
  int64_t nx;
  uint32_t S;  // slots per row, nx + 2
  uint32_t Q;  // slots in the sequence
  int lane;
  uint32_t step_rows, step_sl;  // the stride nwaves * kXOwned as (rows, slots)
  uint32_t q_row = 0, q_sl = 0;  // (row, slot) of the lane in the current chunk

  __host__ __device__ XSlotWalker(int64_t nx_, uint32_t Q_, int64_t nwaves, int lane_)
      : nx(nx_), S(static_cast<uint32_t>(nx_ + 2)), Q(Q_), lane(lane_) {
    const uint32_t step = static_cast<uint32_t>(nwaves * kXOwned);
    step_rows = step / S;
    step_sl = step - step_rows * S;
  }

  __host__ __device__ int64_t Slot(int64_t c) const { return c * kXOwned + lane - 2; }

  // (row, slot) of chunk c from one division, clamped into the sequence.
  __host__ __device__ void Place(int64_t c) {
    const int64_t q64 = Slot(c);
    const uint32_t q =
        q64 < 0 ? 0u : static_cast<uint32_t>(q64 < static_cast<int64_t>(Q) ? q64 : Q - 1);
    q_row = q / S;
    q_sl = q - q_row * S;
  }

  // The lane's row and slot in chunk c (Place or Advance must have set the
  // position of c for lanes whose slot is in the sequence).
  __host__ __device__ void RowSlot(int64_t c, bool* valid, uint32_t* row, uint32_t* sl) const {
    const int64_t q64 = Slot(c);
    *valid = q64 >= 0 && q64 < static_cast<int64_t>(Q);
    if (*valid) {
      *row = q_row;
      *sl = q_sl;
    } else {
      const uint32_t q = q64 < 0 ? 0u : Q - 1u;
      *row = q / S;
      *sl = q - *row * S;
    }
  }

  // The slot record given lane 0's row (row0, read across the wave).
  __host__ __device__ XSlot Make(bool valid, uint32_t row, uint32_t sl, uint32_t row0) const {
    XSlot s;
    s.valid = valid;
    s.sl = sl;
    s.cell = static_cast<int64_t>(sl) - 1;
    s.rbase = (row - row0) * static_cast<uint32_t>(nx);
    s.wbase = static_cast<int64_t>(row0) * nx;
    return s;
  }

  // Moves the position by the stride, with one compare.
  __host__ __device__ void Advance() {
    q_sl += step_sl;
    q_row += step_rows;
    if (q_sl >= S) {
      q_sl -= S;
      q_row += 1;
    }
  }

  // Position of the next chunk cn = c + nwaves: advanced from chunk c unless
  // the lane's slot in c lay before the sequence (a clamped position, placed
  // anew).  A lane past the end in c is past it in cn as well, where
  // RowSlot ignores the position.  Past the last chunk the position stays at
  // c.
  __host__ __device__ void Next(int64_t c, int64_t cn, int64_t nchunks) {
    if (cn < nchunks && Slot(c) >= 0) {
      Advance();
    } else {
      Place(cn < nchunks ? cn : c);
    }
  }
};

} // namespace fvs
