// Two-pass H x v (column-panel up pass, row-panel down pass).
#pragma once
#include "chains.hpp"

namespace edl {

struct TwoPassView {
  int dimup, dimdw, kup, kdw;
  int ld;                  // column stride of the vectors (>= dimup)
  const int32_t* map_up;
  const int32_t* map_dw;
  const uint32_t* pk_up;
  const uint32_t* pk_dw;
  const int32_t* cnt_up;
  const int32_t* cnt_dw;
  const double* dg_up;
  const double* dg_dw;
};

// Lanczos phase-A fusion of the down pass (single chain).
struct LzA {
  LzScalars* sc;
  double* alpha;
  double* part;       // [pass_dw_grid]
  unsigned* counter;
};

struct TwoPassCfg {
  int C = 2;            // columns per up-pass panel
  int threads_up = 1024;
  int grid_up = 256;
  int R = 32;           // rows per down-pass panel
  int threads_dw = 256;
  int nchunk = 32;      // column chunks per down-pass panel (one block each)
};

// Blocks of a down-pass launch (8 * nchunk * ceil(panels / 8)); the size of
// LzA::part.
int pass_dw_grid(const TwoPassView& V, const TwoPassCfg& cfg);

bool two_pass_ok(const Sector& s, size_t elem);
// Column stride of the engine's own vectors: DimUp rounded up to a whole
// number of 128-byte cache lines of fp32 (32 elements), so that a 32-row
// fp32 panel segment (16-row fp64) of any column is exactly one line.
int padded_ld(int dimup);
// ld = DimUp for vectors in the natural layout i = iup + idw * DimUp.
TwoPassView two_pass_view(const Sector& s, int ld);

// Operator mode: sc = null, input qa.  Chain mode: input q_k = (k odd ? qb : qa).
template <typename T>
void launch_pass_up(const TwoPassView& V, const InterParams& I, const T* lut, const T* qa, const T* qb, T* tout,
                    const LzScalars* sc, const TwoPassCfg& cfg, hipStream_t st);
// Operator mode (A = null): out = tin + H_dw qa.  Chain mode: q_k and the
// q_{k-1} buffer (overwritten with the phase-A vector) alternate by parity.
template <typename T>
void launch_pass_dw(const TwoPassView& V, const T* lut, T* qa, T* qb, const T* tin, T* out, const LzA* A,
                    const TwoPassCfg& cfg, hipStream_t st);

}  // namespace edl
