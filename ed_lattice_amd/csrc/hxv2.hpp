// Two-pass H x v (column-panel up pass, row-panel down pass).
#pragma once
#include "chains.hpp"

namespace edl {

struct TwoPassView {
  int dimup, dimdw, kup, kdw;
  int nv_up;               // distinct up-spin hopping values (1: coefficients +-h, cids 0 and 1)
  int ld;                  // column stride of the vectors (>= dimup)
  const int32_t* map_up;
  const int32_t* map_dw;
  const uint32_t* pk_up;
  const uint32_t* pk_dw;
  const int32_t* cnt_up;
  const int32_t* cnt_dw;
  const double* dg_up;
  const double* dg_dw;
  const uint32_t* csr_dw;  // down-spin packed words grouped by column (no padding)
  const int32_t* off_dw;   // [dimdw+1] their offsets
  // the down-pass column order (SpinTable::ord, csro, offo): position k of a
  // chunk is column ord_dw[k], its words csro_dw[offo_dw[k] ..); null: k itself
  const int32_t* ord_dw;
  const uint32_t* csro_dw;
  const int32_t* offo_dw;
  // bank-scheduled up-spin tables by staged row width W (SpinTable::pkc), or null
  const uint32_t* pkc_up[4];
  const int32_t* pkc_off[4];
  const int32_t* pkc_ns[4];
  // the same two slots to a word (ColorTable::words16), or null
  const uint32_t* pkh_up[4];
  const int32_t* pkh_off[4];
  const int32_t* pkh_ns[4];
  // two hops to a word bank-scheduled by sign (ColorTable::words16s), or null
  const uint32_t* pks_up[4];
  const int32_t* pks_off[4];
  const int32_t* pks_ns[4];
  const int32_t* pks_np[4];
  // the packed up-spin table two hops to a word in table order (SpinTable::pk16), or null
  const uint32_t* pk16_up;
  const uint32_t* pk16_dw;  // the down spin's, likewise (read by the batched down tiles)
  // down-spin column ranges by top-site count b (SpinTable::rng), or null
  const int32_t* rng_lo[kMaxRangeBits + 1];
  const uint32_t* rng_words[kMaxRangeBits + 1];
  const int32_t* rng_off[kMaxRangeBits + 1];
  const int32_t* rng_nloc[kMaxRangeBits + 1];
  int rng_n[kMaxRangeBits + 1];
  int rng_maxcols[kMaxRangeBits + 1];
  int rng_maxwords[kMaxRangeBits + 1];
};

// One Lanczos chain driven by the two passes.  A step is four launches:
//   up     phase B of the previous step where pending (r = t - alpha v over
//          t, block partials of |r|^2, replay term x += c v), then
//          T = [D + H_up] q, unscaled;
//   beta   one block: beta = |r|, scalars advance (the chain stops after
//          nmax steps or on breakdown);
//   down   t = (T + H_dw q) / s_k - beta_k q_{k-1} / s_{k-1} over q_{k-1},
//          block partials of alpha_k;
//   alpha  one block: alpha_k, phase B of step k pending.
// No fence inside the passes: the kernel boundaries publish the partials.
struct LzChain {
  LzScalars* sc;
  double* alpha;          // [nmax]
  double* beta;           // [nmax]; beta[0] = 0
  double* part_up;        // [pass_up_grid]
  double* part_dw;        // [pass_dw_grid]
  double* x;              // fp64 replay accumulator (natural layout), or null
  const double* ccoef;    // replay coefficients c_k, or null
  void* basis;            // [nmax][ld * DimDw] stored q_k (two-pass up pass), or null
  double* bscale;         // [nmax] their norms s_k
  int nmax;
  double brk_tol;         // relative breakdown threshold on beta
};

struct TwoPassCfg {
  int C = 2;            // columns per up-pass panel
  int threads_up = 1024;
  int grid_up = 256;
  int up_rpt = 2;       // rows per thread in the up pass (1 or 2)
  int up_kch = 8;       // hop chunk of the up pass ((1, 8|16) or (2, 4|8))
  int up_col = 0;       // 1: the up pass reads the bank-scheduled hop table (fp32, C 1 or 3);
                        // 2: its form with two hops to a word (one hopping magnitude);
                        // 3: the table order two hops to a word (any precision and width);
                        // 4: two hops to a word bank-scheduled by sign (as 2, no per-hop sign);
                        // 5, 6: forms 2, 4 with the block's next panel preloaded into registers
                        //    (operator mode; DimUp <= 13 * threads_up)
  int R = 64;           // rows per down-pass panel (32 or 64)
  int threads_dw = 256;
  int nchunk = 32;      // column chunks per down-pass panel (one block each)
  int ncol = 1;         // columns per lane in the down pass (1 or 2)
  int kch = 16;         // hop chunk of the down pass (8, 16, 32 with 1 column; 8, 16 with 2)
  int dw_vec = 0;       // 1: vector-gather down pass (16-byte lanes; R 16|32|64|128, 256 in fp32;
                        //    (ncol, kch) in (1, 2|4), (2, 1|2|4))
  int dw_rng = 0;       // b > 0: range-blocked down pass over the column ranges of b top sites
                        //    (R 32|64, threads_dw; ncol, kch, nchunk, dw_vec unused)
  int seg_rank = 0;     // segmented sectors, 1: the up pass finds hop targets by rank arithmetic
                        //    (RankHops) instead of reading the hop table
  int dw_ord = 0;       // 1: the vector-gather down pass takes its columns in the locality order
                        //    (SpinTable::ord; the layout of the vectors is unchanged)
  // Down-pass shape of the Lanczos chains (the pass with the Lanczos terms
  // fused, MODE 1), tuned separately; ch_R = 0: the operator's shape.
  int ch_R = 0, ch_threads = 0, ch_nchunk = 0, ch_ncol = 0, ch_kch = 0, ch_vec = 0, ch_rng = 0, ch_ord = 0;
  // Up-pass shape of the chains (phase B fused), likewise; ch_C = 0: the operator's.
  int ch_C = 0, ch_up_threads = 0, ch_grid_up = 0, ch_up_rpt = 0, ch_up_kch = 0, ch_up_col = 0;
};

// The schedule a Lanczos chain runs: the operator's, with the down-pass
// shape tuned for the chain step where there is one.
inline TwoPassCfg chain_cfg(const TwoPassCfg& c) {
  TwoPassCfg r = c;
  if (c.ch_R > 0) {
    r.R = c.ch_R;
    r.threads_dw = c.ch_threads;
    r.nchunk = c.ch_nchunk;
    r.ncol = c.ch_ncol;
    r.kch = c.ch_kch;
    r.dw_vec = c.ch_vec;
    r.dw_rng = c.ch_rng;
    r.dw_ord = c.ch_ord;
  }
  if (c.ch_C > 0) {
    r.C = c.ch_C;
    r.threads_up = c.ch_up_threads;
    r.grid_up = c.ch_grid_up;
    r.up_rpt = c.ch_up_rpt;
    r.up_kch = c.ch_up_kch;
    r.up_col = c.ch_up_col;
  }
  return r;
}

// Whether the bank-scheduled up-pass table exists for panels of C columns of
// T (staged rows of an odd number of dwords: fp32 with C = 1 or 3), in the
// form `col` (1: a hop to a word, 2: two hops to a word); col 3: the
// two-hop table in table order exists and C columns plus the zero rows fit.
bool up_col_ok(const TwoPassView& V, int C, size_t elem, int col = 1);
// Builds the bank-scheduled up-spin tables of a two-pass sector and the
// two-hop tables in table order of both spins (once).
void build_up_color_tables(Sector& s, hipStream_t st);
// Builds the down-spin column-range tables of a two-pass sector (once).
void build_dw_range_tables(Sector& s, hipStream_t st);
// Builds the down spin's locality order of columns and its reordered hop
// table (SpinTable::ord, csro, offo; once).
void build_dw_order(Sector& s, hipStream_t st);
// Whether the range tables for b top sites exist and a block of R rows of
// the largest range fits in LDS (slab, hop words, per-column offsets).
bool pass_dwr_ok(const TwoPassView& V, int R, int b, size_t elem);

// Blocks of a down-pass launch (8 * nchunk * ceil(panels / 8)), the size of
// LzChain::part_dw; blocks of an up-pass launch, the size of part_up.
int pass_dw_grid(const TwoPassView& V, const TwoPassCfg& cfg);
// Whether a launch with this schedule runs the vector-gather down pass
// (padded layout and 16-byte aligned buffers; otherwise the 4-byte form).
bool pass_dw_vec_ok(const TwoPassView& V, const TwoPassCfg& cfg, size_t elem, const void* const* bufs, int nbufs);
// Dynamic LDS of a down-pass block (column table and offsets).
size_t pass_dw_lds_bytes(const TwoPassView& V, const TwoPassCfg& cfg);
int pass_up_grid(const TwoPassView& V, const TwoPassCfg& cfg);

bool two_pass_ok(const Sector& s, size_t elem);
// x (natural layout, fp64) = sum_{k<m} w_k basis_k for m vectors stored with
// the view's column stride (basis_k at k * ld * DimDw).
template <typename T>
void launch_basis_combine(const T* basis, const TwoPassView& V, const double* w, int m, double* x, hipStream_t st);

// Column stride of the engine's own vectors: DimUp rounded up to a whole
// number of 128-byte cache lines of fp32 (32 elements), so that a 32-row
// fp32 panel segment (16-row fp64) of any column is exactly one line.
int padded_ld(int dimup);
// ld = DimUp for vectors in the natural layout i = iup + idw * DimUp.
TwoPassView two_pass_view(const Sector& s, int ld);

// Operator mode (L = null): tout = [D + H_up] qa.  Chain mode: see LzChain;
// the stored vectors q_k alternate between qa and qb by parity.
template <typename T>
void launch_pass_up(const TwoPassView& V, const InterParams& I, const T* lut, T* qa, T* qb, T* tout, const LzChain* L,
                    const TwoPassCfg& cfg, hipStream_t st);
// Operator mode (A = null): out = tin + H_dw qa.  Chain mode: see LzChain.
template <typename T>
void launch_pass_dw(const TwoPassView& V, const T* lut, T* qa, T* qb, const T* tin, T* out, const LzChain* A,
                    const TwoPassCfg& cfg, hipStream_t st);
// T += N q, the non-local Kanamori block (Jx, Jp; nothing when both are 0):
// operator mode (L = null) with q = qa, chain mode with q = q_k of the chain
// (after its beta launch).
template <typename T>
void launch_nonlocal(const TwoPassView& V, const InterParams& I, T* qa, T* qb, T* tout, const LzChain* L,
                     hipStream_t st);
void launch_lz_beta(const LzChain& L, const TwoPassView& V, const TwoPassCfg& cfg, hipStream_t st);
void launch_lz_alpha(const LzChain& L, const TwoPassView& V, const TwoPassCfg& cfg, hipStream_t st);

// ---- segmented up pass: sectors whose up columns exceed LDS.
// A column is cut into row segments that share the occupation of the top
// up sites (states are sorted, so each is one contiguous index range); hops
// that leave the top sites untouched stay inside a segment.  One block takes
// one (column, segment) tile: the segment is staged in LDS, neighbours
// inside it are gathered there and the others from global memory (L2: the
// tiles of a column run next to each other on one XCD).
//
// Rows per segment at most: the LDS capacity for C columns of elements of
// `elem` bytes, lowered by EDL_UP_SEGMENT_ROWS when set (tests force several
// segments on small sectors with it).
int seg_row_cap(size_t elem, int C);
// Whether the operator and the Lanczos chains of `s` use the segmented up
// pass with the down pass: packed hop tables, and either the columns exceed
// LDS or EDL_UP_SEGMENT_ROWS is set below DimUp.
bool seg_path(const Sector& s, size_t elem);
// Segment bounds of `s` for panels of C columns of `elem` bytes (rebuilt
// when the row cap changes).
const std::vector<int>& seg_bounds(Sector& s, size_t elem, int C);
// Fixed down-pass schedule of segmented sectors (no tuning: one launch of
// such a sector moves gigabytes).
TwoPassCfg seg_dw_cfg(int dimdw);
// Operator mode (L = null): tout = [D + H_up] qa.  Chain mode: T = [D +
// H_up] q_k unscaled for the current k (phase B already done).  Tiles are
// (panel of C columns, segment); seg_bounds(s, sizeof(T), C) must be the
// sector's current plan.
// rank = 1: hop targets by rank arithmetic (needs s.sp[0].rk.ok).
template <typename T>
void launch_up_seg(const TwoPassView& V, const InterParams& I, const T* lut, T* qa, T* qb, T* tout,
                   const LzChain* L, const Sector& s, int C, int rank, hipStream_t st);
// Whether the segmented up pass of `s` can find hop targets by rank arithmetic.
bool seg_rank_ok(const Sector& s);
// Phase B of a chain as a streaming pass (r = t - alpha_k q_k / s_k over t,
// block partials of |r|^2 into L.part_up, replay term x += c_k q_k / s_k),
// for the segmented path; `grid` blocks = the partial count for
// launch_lz_beta_n.
int phase_b_grid(const TwoPassView& V);
template <typename T>
void launch_phase_b(const TwoPassView& V, T* qa, T* qb, const LzChain& L, hipStream_t st);
void launch_lz_beta_n(const LzChain& L, int nparts, hipStream_t st);
// After launch_phase_b: r -= sum_m (x_m . r) x_m against nx fp64 vectors X
// (natural layout, orthonormal), with the |r|^2 partials of L.part_up
// retaken after the projection (part: nx * phase_b_grid doubles, coef: nx).
template <typename T>
void launch_deflate(const TwoPassView& V, T* qa, T* qb, const LzChain& L, const double* X, int nx, double* part,
                    double* coef, hipStream_t st);

}  // namespace edl
