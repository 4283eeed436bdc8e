// Batched two-pass Lanczos steps for sectors small enough that a panel of
// whole columns and a slab of whole rows both fit in LDS (the GF / chi
// seed sectors).  One Lanczos step of every chain is four launches:
//   up    (tiles: chain x column panel) phase B of the previous step where
//         one is pending -- r = t - alpha_{k-1} v_{k-1} stored over t and
//         the panel's partial of |r|^2 -- then T = [D + H_up] q_k of the
//         panel;
//   beta  (one block per chain) beta_k = |r| summed over the panels in
//         order, scalars advanced;
//   down  (tiles: chain x row slab) t = (T + H_dw q_k) / s_k
//         - beta_k q_{k-1} / s_{k-1} over q_{k-1}, slab partials of
//         alpha_k = <t, q_k> / s_k;
//   alpha (one block per chain) alpha_k summed over the slabs in order.
// The reductions are separate launches because a fence inside the hot
// kernels (cross-XCD visibility of the partials) writes back and
// invalidates the XCD's L2 on this GPU.
#pragma once
#include "chains.hpp"
#include "hxv2.hpp"

namespace edl {

struct BtChain {
  TwoPassView V;         // sector tables, ld = padded column stride
  const void* lut_up;    // coefficient tables of the chain's scalar type
  const void* lut_dw;
  void* q0;              // stored vectors (shared with the chain's ChainDev)
  void* q1;
  void* tb;              // up-pass intermediate T (ld x dimdw)
  LzScalars* sc;         // pad = 1 while the phase B of step k-1 is pending
  double* alpha;         // [nmax]
  double* beta;          // [nmax]; beta[0] = 0
  double* partD;         // [nslab] partial sums of alpha_k
  double* partU;         // [npanel] partial sums of beta_k^2
  int nslab, npanel, nmax;
  double brk_tol;        // relative breakdown threshold on beta
};

struct BtShape {
  int C = 1;             // columns per up tile
  int RS = 1;            // rows per down slab
  int threads_up = 256;
  int up_rpt = 2;        // rows per thread / hop chunk of the up row loop
  int up_kch = 8;
  int dw_nc = 2;         // columns per thread in the down tile
  int threads_dw = 256;
  int up_h16 = 0;        // 1: the up tiles read the hops two to a word (bt_h16_ok, spin 0)
  int dw_h16 = 0;        // 1: the down tiles likewise (spin 1)
};

// Largest up panel width and down slab height that fit the sectors, or
// RS = 0 when some sector does not fit (the caller then takes the
// single-pass kernels).
BtShape bt_shape(int max_dimup, int max_dimdw, size_t elem);
// LDS bytes of a down tile: coefficient table, RS-row slab (plus a zero
// column with the two-hop table), hop counts.
size_t bt_dw_lds_bytes(int dimdw, int RS, size_t elem, bool h16 = false);

struct BtSet {
  std::vector<BtChain> h;
  BtChain* d = nullptr;
  int64_t* d_tile_up = nullptr;  // [n+1] up-tile offsets
  int64_t* d_tile_dw = nullptr;  // [n+1] down-tile offsets
  int32_t* d_chain_up = nullptr; // [tiles_up] chain of each up tile
  int32_t* d_chain_dw = nullptr; // [tiles_dw] chain of each down tile
  int64_t tiles_up = 0, tiles_dw = 0;
  BtShape shape;
  ~BtSet();
};

// Whether every chain's table of spin `spin` (0 up, 1 down) also exists two
// hops to a word (one hopping magnitude, dim + 32 <= 2^15).
bool bt_h16_ok(const std::vector<BtChain>& chains, int spin);

template <typename T>
void bt_finalize(BtSet& bs);
// One pass over every tile, then (scalars = true) the per-chain scalar
// update that completes the step's half (beta after the up pass, alpha after
// the down pass).  scalars = false launches the tile kernel alone and leaves
// the chain state as it was (schedule timing).
template <typename T>
void bt_launch_up(const BtSet& bs, const InterParams& I, hipStream_t st, bool scalars = true);
template <typename T>
void bt_launch_dw(const BtSet& bs, hipStream_t st, bool scalars = true);

// Persistent form: one block of 1024 threads per chain runs every remaining
// step of its chain (up pass over all panels, beta, down pass over all
// slabs, alpha) in a single launch, the reductions inside the block.  Shapes:
// C 2 or 4 columns per panel, RS 8, 16 (or 32 in fp32) rows per slab.
bool bt_persist_shape_ok(int C, int RS, size_t elem);
size_t bt_persist_lds_bytes(int max_dimup, int max_dimdw, int C, int RS, size_t elem);
template <typename T>
void bt_launch_persist(const BtSet& bs, const InterParams& I, hipStream_t st);

}  // namespace edl
