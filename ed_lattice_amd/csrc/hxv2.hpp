// Two-pass H x v (column-panel up pass, row-panel down pass).
#pragma once
#include "chains.hpp"

namespace edl {

struct TwoPassView {
  int dimup, dimdw, kup, kdw;
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
  double* part;       // [grid_dw]
  unsigned* counter;
};

struct TwoPassCfg {
  int C = 2;            // columns per up-pass panel
  int threads_up = 1024;
  int grid_up = 256;
  int R = 32;           // rows per down-pass panel
  int threads_dw = 256;
  int grid_dw = 2048;   // multiple of 8
};

bool two_pass_ok(const Sector& s, size_t elem);
TwoPassView two_pass_view(const Sector& s);

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
