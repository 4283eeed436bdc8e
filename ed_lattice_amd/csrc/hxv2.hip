// Two-pass on-the-fly H x v for sectors whose columns fit in LDS.
//
//   pass U (column panels):  T(:, c) = [D + H_up] v(:, c)
//     a block stages C whole columns of v in LDS, interleaved by row, and
//     every row gathers its up-hop neighbours from LDS (HxV_up.f90,
//     HxV_local.f90); the packed hop word of a row is read once for all C
//     columns.
//   pass D (row panels):     out(r, :) = T(r, :) + [H_dw v](r, :)
//     XCD x takes the row panels p = x, x+8, ... in order, one block per
//     (panel, column chunk); the blocks resident on an XCD work on the same
//     panel, so the panel (R rows x DimDw columns) is read from HBM once
//     and its down-hop neighbour segments (R contiguous rows of another
//     column, HxV_dw.f90) are served from that XCD's L2.  With the padded
//     column stride a 32-row fp32 (16-row fp64) segment is one cache line.
//
// Hop tables are packed words nbr | cid << 24 with the signed coefficient
// lut[cid] kept in LDS.  Pass D carries the Lanczos phase-A fusion
// (- beta v_{k-1}, alpha partial sums) when given a chain.
#include <algorithm>
#include <numeric>

#include "chains.hpp"
#include "hxv2.hpp"

namespace edl {

template <typename T, int C>
struct Vec;
template <typename T>
struct Vec<T, 1> {
  typedef T type;
  static __device__ __forceinline__ T get(const type& v, int) { return v; }
  static __device__ __forceinline__ void set(type& v, int, T x) { v = x; }
};
template <typename T>
struct Vec<T, 2> {
  typedef T type __attribute__((ext_vector_type(2)));
  static __device__ __forceinline__ T get(const type& v, int j) { return j == 0 ? v.x : v.y; }
  static __device__ __forceinline__ void set(type& v, int j, T x) {
    if (j == 0) v.x = x; else v.y = x;
  }
};
template <typename T>
struct Tri {
  T v[3];
};
template <typename T>
struct Vec<T, 3> {  // packed (3 * sizeof(T) bytes): an fp32 panel of 3 columns of 12870 rows fits in LDS
  typedef Tri<T> type;
  static __device__ __forceinline__ T get(const type& v, int j) { return v.v[j]; }
  static __device__ __forceinline__ void set(type& v, int j, T x) { v.v[j] = x; }
};
template <typename T>
struct Vec<T, 4> {
  typedef T type __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ T get(const type& v, int j) { return v[j]; }
  static __device__ __forceinline__ void set(type& v, int j, T x) { v[j] = x; }
};
template <typename T>
struct Vec<T, 8> {  // fp32 only (32-byte rows): the widest panel of the rank-arithmetic segmented up pass
  typedef T type __attribute__((ext_vector_type(8)));
  static __device__ __forceinline__ T get(const type& v, int j) { return v[j]; }
  static __device__ __forceinline__ void set(type& v, int j, T x) { v[j] = x; }
};

// Rows per thread of a panel-staging batch: as many as about `regs` 32-bit
// registers of loads hold (`per` loads per element), 2 to `cap`.  All loads
// of a batch are issued before its LDS writes, so staging a panel costs one
// memory round trip per batch.
template <typename T, int C>
constexpr int stage_rows(int regs, int per, int cap) {
  const int n = regs / (C * per * (int)(sizeof(T) / 4));
  return n > cap ? cap : (n < 2 ? 2 : n);
}

static __device__ __forceinline__ double inter2(const InterParams& I, uint32_t su, uint32_t sd) {
  const uint32_t both = su & sd;
  double e = 0.0;
  for (int k = 0; k < I.norb; ++k) e += I.U[k] * (double)__popc(both & I.orb_mask[k]);
  for (int p = 0; p < I.npairs; ++p) {
    const uint32_t a = I.pi[p], b = I.pj[p];
    e += I.Ust * (double)((((su >> a) & (sd >> b)) & 1u) + (((su >> b) & (sd >> a)) & 1u));
  }
  return e;
}

// h with its sign bit XOR-ed with bit 31 of `sbit` (the other bits are 0).
static __device__ __forceinline__ float flip_sign(float h, uint32_t sbit) { return __uint_as_float(__float_as_uint(h) ^ sbit); }
static __device__ __forceinline__ double flip_sign(double h, uint32_t sbit) {
  return __longlong_as_double(__double_as_longlong(h) ^ ((long long)sbit << 32));
}

// ------------------------------------------------------------------ pass U
// One panel of C columns starting at c0: stage them in LDS (interleaved by
// row), then T(:, c0+j) = [D + H_up] q(:, c0+j) * ts for every row.  Both
// the single-sector kernel and the batched chain kernel run it.
// Registers that hold a block's next panel between its loads and its LDS
// store (PFR rows per thread and column).
template <typename T, int C, int PFR>
struct PanelPrefetch {
  T v[C][PFR > 0 ? PFR : 1];
};
constexpr int kUpPrefetchRows = 13;  // 12 870 strings at 1024 threads (the 4x4 sector)

template <typename T, int C, int RPT = 2, int UKCH = 8, bool LREG = false, bool COL = false, bool H16 = false,
          bool SGN = false>
static __device__ __forceinline__ void up_rows(const TwoPassView& V, const InterParams& I, const T* lut,
                                               const typename Vec<T, C>::type* P, T* __restrict__ tout, int c0,
                                               T ts);

// Columns c0 .. c0+C-1 of q staged in P, rows interleaved (no barrier).
template <typename T, int C>
static __device__ __forceinline__ void stage_plain(const TwoPassView& V, typename Vec<T, C>::type* P,
                                                   const T* __restrict__ q, int c0) {
  typedef typename Vec<T, C>::type VT;
  const int dimup = V.dimup, dimdw = V.dimdw;
  const size_t ld = (size_t)V.ld;
  constexpr int SB = stage_rows<T, C>(32, 1, 16);
  for (int r0 = threadIdx.x; r0 < dimup; r0 += SB * blockDim.x) {
    T a[SB][C];
#pragma unroll
    for (int h = 0; h < SB; ++h) {
      const int r = r0 + h * (int)blockDim.x;
#pragma unroll
      for (int j = 0; j < C; ++j)
        a[h][j] = (r < dimup && c0 + j < dimdw) ? __builtin_nontemporal_load(&q[(size_t)(c0 + j) * ld + r]) : (T)0;
    }
#pragma unroll
    for (int h = 0; h < SB; ++h) {
      const int r = r0 + h * (int)blockDim.x;
      if (r >= dimup) break;
      VT x;
#pragma unroll
      for (int j = 0; j < C; ++j) Vec<T, C>::set(x, j, a[h][j]);
      P[r] = x;
    }
  }
}

template <typename T, int C, int RPT = 2, int UKCH = 8, bool LREG = false, bool COL = false, bool H16 = false,
          bool SGN = false>
static __device__ __forceinline__ void up_panel(const TwoPassView& V, const InterParams& I, const T* lut,
                                                typename Vec<T, C>::type* P, const T* __restrict__ q,
                                                T* __restrict__ tout, int c0, T ts) {
  typedef typename Vec<T, C>::type VT;
  const int dimup = V.dimup, dimdw = V.dimdw;
  const size_t ld = (size_t)V.ld;
  __syncthreads();
  constexpr int SB = stage_rows<T, C>(32, 1, 16);
  for (int r0 = threadIdx.x; r0 < dimup; r0 += SB * blockDim.x) {
    T a[SB][C];
#pragma unroll
    for (int h = 0; h < SB; ++h) {
      const int r = r0 + h * (int)blockDim.x;
#pragma unroll
      for (int j = 0; j < C; ++j)
        a[h][j] = (r < dimup && c0 + j < dimdw) ? __builtin_nontemporal_load(&q[(size_t)(c0 + j) * ld + r]) : (T)0;
    }
#pragma unroll
    for (int h = 0; h < SB; ++h) {
      const int r = r0 + h * (int)blockDim.x;
      if (r >= dimup) break;
      VT x;
#pragma unroll
      for (int j = 0; j < C; ++j) Vec<T, C>::set(x, j, a[h][j]);
      P[r] = x;
    }
  }
  __syncthreads();
  up_rows<T, C, RPT, UKCH, LREG, COL, H16, SGN>(V, I, lut, P, tout, c0, ts);
}

// The row loop of a panel already staged in P: RPT rows per thread, hop
// words fetched UKCH at a time so that RPT*UKCH table loads are in flight
// before the LDS gathers that depend on them.  LREG: the spin has one
// distinct hopping value, the coefficient of cid 0 / 1 is lut[0] / lut[1]
// (held in registers) and padding (cid 255) is 0, so a hop costs no LDS
// read besides its gather.  COL: the hops come in the slots of the bank-
// scheduled table (ColorTable, width W = C * sizeof(T) / 4): in every slot
// the 32 lanes of a half-wave gather from distinct LDS banks; every row of a
// 64-row block runs the block's slot count (unused slots: coefficient 0).
// H16 (with COL and LREG, fp32): two slots to a table word, a half being the
// LDS row and the sign of the coefficient; unused slots gather a zero row
// (rows dimup .. dimup+31 of P hold zeros).
template <typename T, int C, int RPT, int UKCH, bool LREG, bool COL, bool H16, bool SGN>
static __device__ __forceinline__ void up_rows(const TwoPassView& V, const InterParams& I, const T* lut,
                                               const typename Vec<T, C>::type* P, T* __restrict__ tout, int c0,
                                               T ts) {
  typedef typename Vec<T, C>::type VT;
  constexpr int W = C * (int)sizeof(T) / 4;
  const int dimup = V.dimup, dimdw = V.dimdw;
  const size_t ld = (size_t)V.ld;
  const size_t tstride = COL ? 64 : (size_t)dimup;
  uint32_t sd[C], sdm[C];
  double dgd[C];
  // one orbital and no inter-orbital pairs (the Hubbard lattices): the
  // up x down interaction is U * popcount(su & sd & mask), with the column's
  // part of the AND taken once per panel
  const bool one_orb = I.norb == 1 && I.npairs == 0;
  const double u0 = I.U[0];
#pragma unroll
  for (int j = 0; j < C; ++j) {
    const int c = (c0 + j < dimdw) ? c0 + j : c0;
    sd[j] = (uint32_t)V.map_dw[c];
    sdm[j] = sd[j] & I.orb_mask[0];
    dgd[j] = V.dg_dw[c];
  }
  const T l0 = lut[0], l1 = lut[1];
  const int stride = RPT * blockDim.x;
  for (int i0 = threadIdx.x; i0 < dimup; i0 += stride) {
    int row[RPT], nh[RPT], npw[RPT];
    const uint32_t* tb[RPT];
    bool okr[RPT];
    int nmax = 0;
    T acc[RPT][C];
#pragma unroll
    for (int h = 0; h < RPT; ++h) {
      row[h] = i0 + h * (int)blockDim.x;
      okr[h] = row[h] < dimup;
      if (!okr[h]) row[h] = i0;
      npw[h] = 0;
      if (COL && H16) {
        // the 64 rows of a wave are one block of the table (blockDim is a
        // multiple of 64), so its slot count is wave-uniform: the hop loop
        // bounds stay scalar.  Rows past DimUp inside the last block read
        // its zero-row padding (their sums are never stored).
        const int b = __builtin_amdgcn_readfirstlane(i0 + h * (int)blockDim.x) >> 6;
        const bool inb = b < ((dimup + 63) >> 6);
        if (SGN) {
          // split by sign: words [0, npw) take +h, the rest -h
          nh[h] = inb ? V.pks_ns[W][b] : 0;
          npw[h] = inb ? V.pks_np[W][b] : 0;
          tb[h] = V.pks_up[W] + (inb ? V.pks_off[W][b] : 0) + ((i0 + h * (int)blockDim.x) & 63);
        } else {
          nh[h] = inb ? V.pkh_ns[W][b] : 0;
          tb[h] = V.pkh_up[W] + (inb ? V.pkh_off[W][b] : 0) + ((i0 + h * (int)blockDim.x) & 63);
        }
      } else if (COL) {
        const int b = row[h] >> 6;
        nh[h] = okr[h] ? V.pkc_ns[W][b] : 0;
        tb[h] = V.pkc_up[W] + V.pkc_off[W][b] + (row[h] & 63);
      } else if (H16) {
        // the row's hops two to a word, table order ([pair][row])
        nh[h] = okr[h] ? (V.cnt_up[row[h]] + 1) >> 1 : 0;
        tb[h] = V.pk16_up + row[h];
      } else {
        nh[h] = okr[h] ? V.cnt_up[row[h]] : 0;
        tb[h] = V.pk_up + row[h];
      }
      nmax = nh[h] > nmax ? nh[h] : nmax;
      const uint32_t su = (uint32_t)V.map_up[row[h]];
      const double dgu = V.dg_up[row[h]];
      const VT self = P[row[h]];
      if (one_orb) {
#pragma unroll
        for (int j = 0; j < C; ++j)
          acc[h][j] = (T)(dgu + dgd[j] + u0 * (double)__popc(su & sdm[j])) * Vec<T, C>::get(self, j);
      } else {
#pragma unroll
        for (int j = 0; j < C; ++j)
          acc[h][j] = (T)(dgu + dgd[j] + inter2(I, su, sd[j])) * Vec<T, C>::get(self, j);
      }
    }
    const uint32_t zpad = (uint32_t)dimup | ((uint32_t)dimup << 16);
    for (int k0 = 0; k0 < nmax; k0 += UKCH) {
      uint32_t w[RPT][UKCH];
#pragma unroll
      for (int h = 0; h < RPT; ++h)
#pragma unroll
        for (int j = 0; j < UKCH; ++j)
          w[h][j] = (k0 + j < nh[h]) ? tb[h][(size_t)(k0 + j) * tstride] : (H16 ? zpad : (255u << 24));
#pragma unroll
      for (int h = 0; h < RPT; ++h)
#pragma unroll
        for (int j = 0; j < UKCH; ++j) {
          if constexpr (H16) {
            static_assert(LREG && (!COL || sizeof(T) == 4 || C == 1), "two hops to a word: register coefficients");
            // bank-scheduled words with their own signs: the word count is
            // wave-uniform, so a chunk's words past it are skipped by a
            // scalar branch (measured slower on the sign-split words)
            if (COL && !SGN && k0 + j >= nmax) continue;
            const uint32_t wl = w[h][j];
            // sign-split words carry one sign for the whole wave (set by the
            // word index); the other forms' halves carry their own (bit 15)
            const T cw = (k0 + j < npw[h]) ? l0 : l1;
            const T cl = SGN ? cw : flip_sign(l0, (wl << 16) & 0x80000000u);
            const T chi = SGN ? cw : flip_sign(l0, wl & 0x80000000u);
            const VT xl = P[wl & 0x7FFFu];
            const VT xh = P[(wl >> 16) & 0x7FFFu];
#pragma unroll
            for (int c = 0; c < C; ++c) acc[h][c] += cl * Vec<T, C>::get(xl, c);
#pragma unroll
            for (int c = 0; c < C; ++c) acc[h][c] += chi * Vec<T, C>::get(xh, c);
            continue;
          }
          const uint32_t cid = w[h][j] >> 24;
          const T cf = LREG ? (cid == 0u ? l0 : (cid == 1u ? l1 : (T)0)) : lut[cid];
          const VT x = P[w[h][j] & 0xFFFFFFu];
#pragma unroll
          for (int c = 0; c < C; ++c) acc[h][c] += cf * Vec<T, C>::get(x, c);
        }
    }
#pragma unroll
    for (int h = 0; h < RPT; ++h) {
      if (!okr[h]) continue;
#pragma unroll
      for (int j = 0; j < C; ++j)
        if (c0 + j < dimdw) __builtin_nontemporal_store(acc[h][j] * ts, &tout[(size_t)(c0 + j) * ld + row[h]]);
    }
  }
}

static __host__ __device__ constexpr size_t lds_head_bytes(size_t sz) { return 256 * sz + (16 - (256 * sz) % 16) % 16; }

// Stage columns c0 .. c0+C-1 of the up-pass input in P.  Without phase B
// the input is q.  With phase B of step k pending it is r = t - am * q_k
// (t read from qn, r stored back over it unless `keep_t`), and the block's
// share of |r|^2 is returned; x (natural layout) accumulates cx * q_k for
// the eigenvector replay when given, and q_k is copied to `bstore` (same
// layout) when given.  SB rows per thread at a time, every load of those
// rows in flight first.
template <typename T, int C, int SB = 2>
static __device__ __forceinline__ double stage_panel(const TwoPassView& V, typename Vec<T, C>::type* P, T* qn,
                                                     const T* __restrict__ qk, int c0, bool pend, bool keep_t,
                                                     T am, double* __restrict__ x, double cx,
                                                     T* __restrict__ bstore = nullptr) {
  typedef typename Vec<T, C>::type VT;
  const int dimup = V.dimup, dimdw = V.dimdw;
  const size_t ld = (size_t)V.ld;
  double nrm = 0.0;
  for (int r0 = threadIdx.x; r0 < dimup; r0 += SB * blockDim.x) {
    int rr[SB];
#pragma unroll
    for (int h = 0; h < SB; ++h) rr[h] = r0 + h * (int)blockDim.x;
    T a[SB][C], b[SB][C];
#pragma unroll
    for (int h = 0; h < SB; ++h)
#pragma unroll
      for (int j = 0; j < C; ++j) {
        const bool ok = rr[h] < dimup && c0 + j < dimdw;
        const size_t i = (size_t)(c0 + j) * ld + rr[h];
        a[h][j] = ok ? qn[i] : (T)0;
        b[h][j] = (ok && pend) ? qk[i] : (T)0;
      }
#pragma unroll
    for (int h = 0; h < SB; ++h) {
      if (rr[h] >= dimup) break;
      VT v;
#pragma unroll
      for (int j = 0; j < C; ++j) {
        T e = a[h][j];
        if (pend && c0 + j < dimdw) {
          e = a[h][j] - am * b[h][j];
          if (!keep_t) qn[(size_t)(c0 + j) * ld + rr[h]] = e;
          if (bstore) bstore[(size_t)(c0 + j) * ld + rr[h]] = b[h][j];
          nrm += (double)e * (double)e;
          if (x) x[(size_t)(c0 + j) * dimup + rr[h]] += cx * (double)b[h][j];
        }
        Vec<T, C>::set(v, j, e);
      }
      P[rr[h]] = v;
    }
  }
  return nrm;
}

// Operator mode (chain = 0): T = [D + H_up] qa.  Chain mode: phase B of the
// previous step where one is pending (r = t - alpha_k v_k over t, block
// partial of |r|^2 into L.part_up, replay term x += c_k v_k), then
// T = [D + H_up] q (unscaled; the down pass divides by s).  When the chain
// has run its nmax steps only the replay term remains.
template <typename T, int C, int RPT, int UKCH, bool LREG, bool COL, bool H16, bool SGN, int PFR = 0>
static __device__ __forceinline__ void pass_up_body(const TwoPassView& V, const InterParams& I,
                                                    const T* __restrict__ lut_up, T* qa, T* qb,
                                                    T* __restrict__ tout, const LzChain& L, int chain) {
  typedef typename Vec<T, C>::type VT;
  extern __shared__ __align__(16) unsigned char smem[];
  // H16 takes its two coefficients from registers: no LDS table, and the
  // panel starts at the base of the dynamic LDS (small gather offsets fold
  // into the LDS instructions)
  T* lut = H16 ? const_cast<T*>(lut_up) : reinterpret_cast<T*>(smem);
  VT* P = reinterpret_cast<VT*>(smem + (H16 ? 0 : lds_head_bytes(sizeof(T))));
  __shared__ double sh[16];
  const int npan = (V.dimdw + C - 1) / C;
  if (H16 && threadIdx.x < 32) {
    // the zero rows that unused slots gather (the staging below never writes them)
    VT z;
#pragma unroll
    for (int j = 0; j < C; ++j) Vec<T, C>::set(z, j, (T)0);
    P[V.dimup + threadIdx.x] = z;
  }
  if (!chain) {
    if (!H16)
      for (int k = threadIdx.x; k < 256; k += blockDim.x) lut[k] = lut_up[k];
    if constexpr (PFR > 0) {
      // the block's first panel is staged as usual; each later one is
      // loaded into registers (PanelPrefetch) before the barrier that ends
      // the current panel and written to P after it
      const int dimup = V.dimup;
      int pan = blockIdx.x;
      if (pan >= npan) return;
      stage_plain<T, C>(V, P, qa, pan * C);
      __syncthreads();
      for (;;) {
        const int nxt = pan + (int)gridDim.x;
        up_rows<T, C, RPT, UKCH, LREG, COL, H16, SGN>(V, I, lut, P, tout, pan * C, (T)1);
        if (nxt >= npan) break;
        // the next panel's loads go out as soon as this thread's rows are
        // done, so their round trip overlaps the other waves' last rows and
        // the barrier (the waves of a block end an iteration apart)
        PanelPrefetch<T, C, PFR> pf;
        const int c1 = nxt * C;
#pragma unroll
        for (int j = 0; j < C; ++j)
#pragma unroll
          for (int kk = 0; kk < PFR; ++kk) {
            const int r = (int)threadIdx.x + kk * (int)blockDim.x;
            pf.v[j][kk] = (r < dimup && c1 + j < V.dimdw)
                              ? __builtin_nontemporal_load(&qa[(size_t)(c1 + j) * V.ld + r])
                              : (T)0;
          }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < PFR; ++kk) {
          const int r = (int)threadIdx.x + kk * (int)blockDim.x;
          if (r < dimup) {
            VT x;
#pragma unroll
            for (int j = 0; j < C; ++j) Vec<T, C>::set(x, j, pf.v[j][kk]);
            P[r] = x;
          }
        }
        __syncthreads();
        pan = nxt;
      }
      return;
    }
    for (int pan = blockIdx.x; pan < npan; pan += gridDim.x)
      up_panel<T, C, RPT, UKCH, LREG, COL, H16, SGN>(V, I, lut, P, qa, tout, pan * C, (T)1);
    return;
  }
  const LzScalars s = *L.sc;
  if (s.stopped) return;
  const bool pend = s.pad != 0;
  const bool last = pend && s.k + 1 >= L.nmax;
  if (last && !L.x) return;
  if (!H16)
    for (int k = threadIdx.x; k < 256; k += blockDim.x) lut[k] = lut_up[k];
  T* qn = ((pend ? s.k + 1 : s.k) & 1) ? qb : qa;
  const T* qk = (s.k & 1) ? qb : qa;
  const T am = (T)(s.alpha_k / s.s_cur);
  double* x = pend ? L.x : nullptr;
  const double cx = x ? L.ccoef[s.k] / s.s_cur : 0.0;
  // stored basis: q_k (read for phase B) goes to basis[k] with its norm
  T* bst = (pend && L.basis) ? static_cast<T*>(L.basis) + (size_t)s.k * V.ld * V.dimdw : nullptr;
  if (bst && blockIdx.x == 0 && threadIdx.x == 0) L.bscale[s.k] = s.s_cur;
  double nrm = 0.0;
  for (int pan = blockIdx.x; pan < npan; pan += gridDim.x) {
    __syncthreads();
    nrm += stage_panel<T, C, stage_rows<T, C>(32, 2, 8)>(V, P, qn, qk, pan * C, pend, last, am, x, cx, bst);
    __syncthreads();
    if (!last) up_rows<T, C, RPT, UKCH, LREG, COL, H16, SGN>(V, I, lut, P, tout, pan * C, (T)1);
  }
  if (pend && !last) {
    const double v = block_sum(nrm, sh);
    if (threadIdx.x == 0) L.part_up[blockIdx.x] = v;
  }
}

template <typename T, int C, int RPT, int UKCH, bool LREG = false, bool COL = false, bool H16 = false>
__global__ __launch_bounds__(1024) void k_pass_up(TwoPassView V, InterParams I, const T* __restrict__ lut_up, T* qa,
                                                  T* qb, T* __restrict__ tout, LzChain L, int chain) {
  pass_up_body<T, C, RPT, UKCH, LREG, COL, H16, false>(V, I, lut_up, qa, qb, tout, L, chain);
}

// The bank-scheduled two-hop passes (SGN: the words split by sign) with the
// block's next panel loaded into registers as each thread finishes the
// current one (operator mode; DimUp <= kUpPrefetchRows * blockDim; a chain
// step runs the form without it).
template <typename T, int C, int RPT, int UKCH, bool SGN>
__global__ __launch_bounds__(1024) void k_pass_upp(TwoPassView V, InterParams I, const T* __restrict__ lut_up, T* qa,
                                                   T* qb, T* __restrict__ tout, LzChain L, int chain) {
  pass_up_body<T, C, RPT, UKCH, true, true, true, SGN, kUpPrefetchRows>(V, I, lut_up, qa, qb, tout, L, chain);
}

// The same pass reading the bank-scheduled two-hop words split by sign
// (ColorTable::words16s): one coefficient per word for the whole wave, so a
// hop costs no sign decoding.
template <typename T, int C, int RPT, int UKCH>
__global__ __launch_bounds__(1024) void k_pass_ups(TwoPassView V, InterParams I, const T* __restrict__ lut_up, T* qa,
                                                   T* qb, T* __restrict__ tout, LzChain L, int chain) {
  pass_up_body<T, C, RPT, UKCH, true, true, true, true>(V, I, lut_up, qa, qb, tout, L, chain);
}

// ------------------------------------------------------------------ pass D
// MODE 0: out = tin + H_dw q.   MODE 1 (Lanczos phase A): out (the buffer of
// q_{k-1}) = tin + H_dw q / s_k - beta_k q_{k-1} / s_{k-1}, alpha partials.
template <typename T, int R, int MODE, int NCOL, int KCH>
__global__ __launch_bounds__(1024) void k_pass_dw(TwoPassView V, const T* __restrict__ lut_dw, T* qa, T* qb,
                                                  const T* __restrict__ tin, T* outp, LzChain A, int nchunk) {
  __shared__ T lut[256];
  __shared__ double sh[16];
  LzScalars s{};
  const T* __restrict__ q = qa;
  T* __restrict__ out = outp;
  if (MODE == 1) {
    s = *A.sc;
    if (s.stopped) return;
    q = (s.k & 1) ? qb : qa;
    out = (s.k & 1) ? qa : qb;
  }
  for (int k = threadIdx.x; k < 256; k += blockDim.x) lut[k] = lut_dw[k];
  const int dimup = V.dimup, dimdw = V.dimdw;
  const size_t ld = (size_t)V.ld;
  // Block b = 8*i + x runs on XCD x (the hardware deals blocks to the 8 XCDs
  // round robin) and takes column chunk i % nchunk of panel p = 8*(i /
  // nchunk) + x.  Each XCD receives its blocks in increasing i, so the
  // dispatcher itself keeps the blocks resident on an XCD within one or two
  // consecutive panels of that XCD: the panel's neighbour segments stay in
  // the XCD's L2 without any inter-block synchronisation.  Blocks past the
  // last panel only contribute a zero partial.
  const int b = blockIdx.x;
  const int xcd = b % 8, ib = b / 8;
  const int jb = ib % nchunk;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int SLOTS = 64 / R;
  const int r = lane % R, slot = lane / R;
  const int nwaves = blockDim.x >> 6;
  const int npan = (dimup + R - 1) / R;
  const int p0 = 8 * (ib / nchunk) + xcd, pend = p0 < npan ? p0 + 1 : p0;
  const int cb = (int)((int64_t)jb * dimdw / nchunk), ce = (int)((int64_t)(jb + 1) * dimdw / nchunk);
  // Stage the chunk's hop words (one contiguous run of the column-grouped
  // table: no padding, so the table's share of L2 is the exact hop count)
  // and their per-column offsets in LDS.
  extern __shared__ __align__(16) unsigned char dsm[];
  uint32_t* tbl = reinterpret_cast<uint32_t*>(dsm);
  const int ncol = ce - cb;
  const int w0 = V.off_dw[cb], nw = V.off_dw[ce] - w0;
  int* toff = reinterpret_cast<int*>(tbl + nw);  // [ncol+1]
  // (non-temporal: the table should not displace the panel's lines in L2)
  for (int i = threadIdx.x; i < nw; i += blockDim.x) tbl[i] = __builtin_nontemporal_load(&V.csr_dw[w0 + i]);
  for (int c = threadIdx.x; c <= ncol; c += blockDim.x) toff[c] = __builtin_nontemporal_load(&V.off_dw[cb + c]) - w0;
  __syncthreads();
  const double inv_s = (MODE == 1) ? 1.0 / s.s_cur : 1.0;
  const T ts = (T)inv_s;
  const T bo = (MODE == 1) ? (T)(s.beta_k / s.s_old) : (T)0;
  double part = 0.0;
  for (int p = p0; p < pend; ++p) {
    const int iup = p * R + r;
    const bool ok = iup < dimup;
    // NCOL columns per lane.  Their T (and q_{k-1}, q) elements are requested
    // first, then the neighbour segments of a chunk of KCH hops of every
    // column together (NCOL*KCH loads in flight); the chunk count follows the
    // exact hop counts.
    const int cstep = nwaves * SLOTS;
    const int iuc = ok ? iup : 0;
    for (int c0 = cb + wave * SLOTS + slot; c0 < ce; c0 += NCOL * cstep) {
      int col[NCOL], nh[NCOL];
      bool okc[NCOL];
      T t0[NCOL], o0[NCOL], q0v[NCOL], acc[NCOL];
      int nmax = 0;
#pragma unroll
      for (int h = 0; h < NCOL; ++h) {
        col[h] = c0 + h * cstep;
        okc[h] = col[h] < ce;
        if (!okc[h]) col[h] = c0;
        nh[h] = okc[h] ? toff[col[h] - cb + 1] - toff[col[h] - cb] : 0;
        nmax = nh[h] > nmax ? nh[h] : nmax;
        const size_t i = (size_t)col[h] * ld + iup;
        const bool live = ok && okc[h];
        t0[h] = live ? __builtin_nontemporal_load(&tin[i]) : (T)0;
        o0[h] = (MODE == 1 && live) ? __builtin_nontemporal_load(&out[i]) : (T)0;
        q0v[h] = (MODE == 1 && live) ? q[i] : (T)0;
        acc[h] = (T)0;
      }
      for (int k0 = 0; k0 < nmax; k0 += KCH) {
        uint32_t w[NCOL][KCH];
#pragma unroll
        for (int h = 0; h < NCOL; ++h) {
          const uint32_t* tw = tbl + toff[col[h] - cb];
#pragma unroll
          for (int j = 0; j < KCH; ++j)
            w[h][j] = (k0 + j < nh[h]) ? tw[k0 + j] : ((uint32_t)col[h] | (255u << 24));
        }
        T x[NCOL][KCH];
#pragma unroll
        for (int h = 0; h < NCOL; ++h)
#pragma unroll
          for (int j = 0; j < KCH; ++j)
            x[h][j] = (k0 + j < nh[h]) ? q[(size_t)(w[h][j] & 0xFFFFFFu) * ld + iuc] : (T)0;
#pragma unroll
        for (int h = 0; h < NCOL; ++h)
#pragma unroll
          for (int j = 0; j < KCH; ++j) acc[h] += lut[w[h][j] >> 24] * x[h][j];
      }
      if (ok) {
#pragma unroll
        for (int h = 0; h < NCOL; ++h) {
          if (!okc[h]) continue;
          const size_t i = (size_t)col[h] * ld + iup;
          if (MODE == 0) {
            __builtin_nontemporal_store(t0[h] + acc[h], &out[i]);
          } else {
            const T tv = (t0[h] + acc[h]) * ts - bo * o0[h];
            __builtin_nontemporal_store(tv, &out[i]);
            part += (double)tv * (double)q0v[h];
          }
        }
      }
    }
  }
  if (MODE == 1) {
    // block partial of alpha_k; summed in block order by k_lz_alpha
    const double v = block_sum(part * inv_s, sh);
    if (threadIdx.x == 0) A.part_dw[b] = v;
  }
}

// Vector-gather down pass (R-row panels, R = 16 to 128, 256 in fp32; padded
// layout: ld % 32 == 0).  A lane loads VW = 16 / sizeof(T) consecutive rows
// with one 16-byte load, so an R-row neighbour segment takes LPS = R / VW
// lanes and one wave instruction fetches GPW = 64 / LPS segments: lane group
// g takes the hops k = g, g + GPW, g + 2 GPW, ... of the wave's column, and
// the groups' sums meet by lane shuffles.  Against one 4-byte element per
// lane this issues GPW times fewer gather instructions for the same bytes,
// and it admits panels taller than a wave (longer neighbour segments, the
// shape that ran fastest).  Same block mapping, modes and LDS column table
// as k_pass_dw.  A quad that straddles DimUp loads rows of the column's
// padding (inside the column) and stores only rows < DimUp.  With V.ord_dw
// set, chunk positions are taken in the locality order: position k is
// column ord_dw[k] (staged in LDS with the chunk's words, which come from the
// reordered table), so the columns a panel's concurrent blocks work on
// gather from a narrower window of the panel.
template <typename T, int R, int MODE, int NCOL, int KCH>
__global__ __launch_bounds__(1024) void k_pass_dwv(TwoPassView V, const T* __restrict__ lut_dw, T* qa, T* qb,
                                                   const T* __restrict__ tin, T* outp, LzChain A, int nchunk) {
  constexpr int VW = 16 / (int)sizeof(T);
  constexpr int LPS = R / VW;
  constexpr int GPW = 64 / LPS;
  typedef T VT __attribute__((ext_vector_type(VW)));
  __shared__ T lut[256];
  __shared__ double sh[16];
  LzScalars s{};
  const T* __restrict__ q = qa;
  T* __restrict__ out = outp;
  if (MODE == 1) {
    s = *A.sc;
    if (s.stopped) return;
    q = (s.k & 1) ? qb : qa;
    out = (s.k & 1) ? qa : qb;
  }
  for (int k = threadIdx.x; k < 256; k += blockDim.x) lut[k] = lut_dw[k];
  const int dimup = V.dimup, dimdw = V.dimdw;
  const size_t ld = (size_t)V.ld;
  const int b = blockIdx.x;
  // (panel, chunk) units in panel-major order, each XCD a contiguous run of
  // them (block b goes to XCD b % 8): a panel's chunks run on one XCD at a
  // time, and the XCDs' shares differ by at most one unit
  const int64_t n8 = gridDim.x / 8;
  const int64_t u = (int64_t)(b % 8) * n8 + b / 8;
  const int jb = (int)(u % nchunk);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane / LPS, l = lane % LPS;
  const int nwaves = blockDim.x >> 6;
  const int npan = (dimup + R - 1) / R;
  const int p = (int)(u / nchunk);
  const int cb = (int)((int64_t)jb * dimdw / nchunk), ce = (int)((int64_t)(jb + 1) * dimdw / nchunk);
  extern __shared__ __align__(16) unsigned char dsm[];
  uint32_t* tbl = reinterpret_cast<uint32_t*>(dsm);
  const int ncol = ce - cb;
  const int32_t* __restrict__ ord = V.ord_dw;
  const uint32_t* __restrict__ csr = ord ? V.csro_dw : V.csr_dw;
  const int32_t* __restrict__ offs = ord ? V.offo_dw : V.off_dw;
  const int w0 = offs[cb], nw = offs[ce] - w0;
  int* toff = reinterpret_cast<int*>(tbl + nw);
  int* tcol = toff + ncol + 1;
  for (int i = threadIdx.x; i < nw; i += blockDim.x) tbl[i] = __builtin_nontemporal_load(&csr[w0 + i]);
  for (int c = threadIdx.x; c <= ncol; c += blockDim.x) toff[c] = __builtin_nontemporal_load(&offs[cb + c]) - w0;
  for (int c = threadIdx.x; c < ncol; c += blockDim.x) tcol[c] = ord ? __builtin_nontemporal_load(&ord[cb + c]) : cb + c;
  __syncthreads();
  const double inv_s = (MODE == 1) ? 1.0 / s.s_cur : 1.0;
  const T ts = (T)inv_s;
  const T bo = (MODE == 1) ? (T)(s.beta_k / s.s_old) : (T)0;
  double part = 0.0;
  if (p < npan) {
    const int iup0 = p * R + l * VW;
    const bool ok = iup0 < dimup;
    const int iuc = ok ? iup0 : 0;
    const int nval = ok ? min(VW, dimup - iup0) : 0;  // rows of the quad below DimUp
    const bool lead = ok && g == 0;
    for (int c0 = cb + wave; c0 < ce; c0 += NCOL * nwaves) {
      int col[NCOL], nh[NCOL], cx[NCOL];
      bool okc[NCOL];
      VT t0[NCOL], o0[NCOL], q0v[NCOL], acc[NCOL];
      int nmax = 0;
#pragma unroll
      for (int h = 0; h < NCOL; ++h) {
        col[h] = c0 + h * nwaves;
        okc[h] = col[h] < ce;
        if (!okc[h]) col[h] = c0;
        cx[h] = tcol[col[h] - cb];  // the column at this position
        nh[h] = okc[h] ? toff[col[h] - cb + 1] - toff[col[h] - cb] : 0;
        nmax = nh[h] > nmax ? nh[h] : nmax;
        const size_t i = (size_t)cx[h] * ld + iuc;
        const bool live = lead && okc[h];
        t0[h] = live ? __builtin_nontemporal_load(reinterpret_cast<const VT*>(&tin[i])) : (VT)0;
        o0[h] = (MODE == 1 && live) ? __builtin_nontemporal_load(reinterpret_cast<const VT*>(&out[i])) : (VT)0;
        q0v[h] = (MODE == 1 && live) ? *reinterpret_cast<const VT*>(&q[i]) : (VT)0;
        acc[h] = (VT)0;
      }
      for (int k0 = 0; k0 < nmax; k0 += GPW * KCH) {
        uint32_t w[NCOL][KCH];
#pragma unroll
        for (int h = 0; h < NCOL; ++h) {
          const uint32_t* tw = tbl + toff[col[h] - cb];
#pragma unroll
          for (int j = 0; j < KCH; ++j) {
            const int k = k0 + g + GPW * j;
            w[h][j] = (k < nh[h]) ? tw[k] : ((uint32_t)col[h] | (255u << 24));
          }
        }
        VT x[NCOL][KCH];
#pragma unroll
        for (int h = 0; h < NCOL; ++h)
#pragma unroll
          for (int j = 0; j < KCH; ++j) {
            const uint32_t nbr = w[h][j] & 0xFFFFFFu;
            x[h][j] = (k0 + g + GPW * j < nh[h]) ? *reinterpret_cast<const VT*>(&q[(size_t)nbr * ld + iuc]) : (VT)0;
          }
#pragma unroll
        for (int h = 0; h < NCOL; ++h)
#pragma unroll
          for (int j = 0; j < KCH; ++j) acc[h] += lut[w[h][j] >> 24] * x[h][j];
      }
#pragma unroll
      for (int h = 0; h < NCOL; ++h)
#pragma unroll
        for (int e = 0; e < VW; ++e) {
          T v = acc[h][e];
#pragma unroll
          for (int m = LPS; m < 64; m *= 2) v += __shfl_xor(v, m);
          acc[h][e] = v;
        }
      if (lead) {
#pragma unroll
        for (int h = 0; h < NCOL; ++h) {
          if (!okc[h]) continue;
          const size_t i = (size_t)cx[h] * ld + iup0;
          VT tv;
          if (MODE == 0) {
            tv = t0[h] + acc[h];
          } else {
            tv = (t0[h] + acc[h]) * ts - bo * o0[h];
          }
          if (nval == VW) {
            __builtin_nontemporal_store(tv, reinterpret_cast<VT*>(&out[i]));
            if (MODE == 1)
#pragma unroll
              for (int e = 0; e < VW; ++e) part += (double)tv[e] * (double)q0v[h][e];
          } else {
#pragma unroll
            for (int e = 0; e < VW; ++e)
              if (e < nval) {
                out[i + e] = tv[e];
                if (MODE == 1) part += (double)tv[e] * (double)q0v[h][e];
              }
          }
        }
      }
    }
  }
  if (MODE == 1) {
    const double v = block_sum(part * inv_s, sh);
    if (threadIdx.x == 0) A.part_dw[b] = v;
  }
}

// ------------------------------------------ non-local Kanamori block (Jx, Jp)
// T += N q between the up and the down pass: N (HxV_non_local.f90:1-88) moves
// one up and one down electron between two orbitals of a site, so it is
// neither a row nor a column operation; each element gathers its moves
// (nonlocal_gather, the coefficient of the move that starts at it: the block
// is real symmetric).  In a chain q is the current vector q_k (after the
// step's beta, as the down pass reads it).
template <typename T>
__global__ __launch_bounds__(256) void k_nonlocal_add(TwoPassView V, InterParams I, T* qa, T* qb, T* tout, LzChain L,
                                                      int chain) {
  const T* q = qa;
  if (chain) {
    const LzScalars s = *L.sc;
    if (s.stopped) return;
    q = (s.k & 1) ? qb : qa;
  }
  const int64_t n = (int64_t)V.dimup * V.dimdw;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int idw = (int)(e / V.dimup), iup = (int)(e % V.dimup);
    const size_t i = (size_t)idw * V.ld + iup;
    tout[i] += (T)nonlocal_gather<T>(I, (uint32_t)V.map_up[iup], (uint32_t)V.map_dw[idw], q, V.ld, I.Jx, I.Jp);
  }
}

// ---------------------------------------------- pass D, range-blocked form
// Tiles (row panel of R rows, column range g): the columns of a range share
// the occupation of the top b sites of the down string, so the hops that
// leave those sites alone (most of them: 53 % of the 4x4 sector's at b = 5)
// stay inside the range.  A block stages the R-row slab of every column of
// its range in LDS, with the range's hop words regrouped per column (inside
// first, local indices); a lane is a row of one column (R = 64: a wave is
// one column, R = 32: two), so the inside hops are conflict-free LDS reads
// of consecutive rows, and the others are R-row segments of the same panel
// from L2, where the other ranges' blocks of the panel (on the same XCD:
// block b = 8 i + x takes range i % n of panel 8 (i / n) + x) staged them.
// MODE as k_pass_dw.  Traffic per launch: q, T and out once.
template <typename T, int R, int MODE>
__global__ __launch_bounds__(1024) void k_pass_dwr(TwoPassView V, const T* __restrict__ lut_dw, T* qa, T* qb,
                                                   const T* __restrict__ tin, T* outp, LzChain A, int b) {
  constexpr int KCH = 8;
  constexpr int CPW = 64 / R;  // columns per wave
  __shared__ T lut[256];
  __shared__ double sh[16];
  extern __shared__ __align__(16) unsigned char dsm[];
  LzScalars s{};
  const T* __restrict__ q = qa;
  T* __restrict__ out = outp;
  if (MODE == 1) {
    s = *A.sc;
    if (s.stopped) return;
    q = (s.k & 1) ? qb : qa;
    out = (s.k & 1) ? qa : qb;
  }
  for (int k = threadIdx.x; k < 256; k += blockDim.x) lut[k] = lut_dw[k];
  const int dimup = V.dimup;
  const size_t ld = (size_t)V.ld;
  const int ng = V.rng_n[b];
  const int blk = blockIdx.x, xcd = blk % 8, ib = blk / 8;
  const int p = 8 * (ib / ng) + xcd, g = ib % ng;
  const int npan = (dimup + R - 1) / R;
  const double inv_s = (MODE == 1) ? 1.0 / s.s_cur : 1.0;
  const T ts = (T)inv_s;
  const T bo = (MODE == 1) ? (T)(s.beta_k / s.s_old) : (T)0;
  double part = 0.0;
  if (p < npan) {
    const int lo = V.rng_lo[b][g], nc = V.rng_lo[b][g + 1] - lo;
    const int w0 = V.rng_off[b][lo], nw = V.rng_off[b][lo + nc] - w0;
    T* S = reinterpret_cast<T*>(dsm);                                  // [nc][R]
    uint32_t* tw = reinterpret_cast<uint32_t*>(S + (size_t)nc * R);  // [nw]
    int* toff = reinterpret_cast<int*>(tw + nw);                      // [nc + 1]
    int* tnl = toff + nc + 1;                                         // [nc]
    const int r0 = p * R;
    for (int e = threadIdx.x; e < nc * R; e += blockDim.x) {
      const int jl = e / R, r = e % R;
      S[e] = (r0 + r < dimup) ? q[(size_t)(lo + jl) * ld + r0 + r] : (T)0;
    }
    for (int i = threadIdx.x; i < nw; i += blockDim.x) tw[i] = __builtin_nontemporal_load(&V.rng_words[b][w0 + i]);
    for (int c = threadIdx.x; c <= nc; c += blockDim.x) {
      toff[c] = V.rng_off[b][lo + c] - w0;
      if (c < nc) tnl[c] = V.rng_nloc[b][lo + c];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwv = blockDim.x >> 6;
    const int r = lane % R, cs = lane / R;
    const bool rok = r0 + r < dimup;
    for (int jb = wave * CPW; jb < nc; jb += nwv * CPW) {
      const int jl = jb + cs;
      const bool cok = jl < nc && rok;
      const int jc = jl < nc ? jl : jb;
      const size_t iel = (size_t)(lo + jc) * ld + r0 + r;
      const T t0 = cok ? __builtin_nontemporal_load(&tin[iel]) : (T)0;
      const T o0 = (MODE == 1 && cok) ? __builtin_nontemporal_load(&out[iel]) : (T)0;
      const uint32_t* wp = tw + toff[jc];
      const int nl = jl < nc ? tnl[jc] : 0;
      const int nt = jl < nc ? toff[jc + 1] - toff[jc] : 0;
      T acc = (T)0;
      for (int k = 0; k < nl; ++k) {
        const uint32_t w = wp[k];
        acc += lut[w >> 24] * S[(int)(w & 0xFFFFFFu) * R + r];
      }
      for (int k0 = nl; k0 < nt; k0 += KCH) {
        uint32_t w[KCH];
        T x[KCH];
#pragma unroll
        for (int j = 0; j < KCH; ++j) w[j] = (k0 + j < nt) ? wp[k0 + j] : ((uint32_t)(lo + jc) | (255u << 24));
#pragma unroll
        for (int j = 0; j < KCH; ++j)
          x[j] = (rok && k0 + j < nt) ? q[(size_t)(w[j] & 0xFFFFFFu) * ld + r0 + r] : (T)0;
#pragma unroll
        for (int j = 0; j < KCH; ++j) acc += lut[w[j] >> 24] * x[j];
      }
      if (cok) {
        if (MODE == 0) {
          __builtin_nontemporal_store(t0 + acc, &out[iel]);
        } else {
          const T tv = (t0 + acc) * ts - bo * o0;
          __builtin_nontemporal_store(tv, &out[iel]);
          part += (double)tv * (double)S[jc * R + r];
        }
      }
    }
  }
  if (MODE == 1) {
    const double v = block_sum(part * inv_s, sh);
    if (threadIdx.x == 0) A.part_dw[blk] = v;
  }
}

// ------------------------------------------------------ chain reductions
// One block each, between the passes: the kernel boundary makes the block
// partials visible to it (a fence inside the passes would write back and
// invalidate the XCD's L2).  Sums run in block order (deterministic).
static __device__ double ordered_sum_blocks(const double* __restrict__ part, int n, double* sh) {
  const int per = (n + blockDim.x - 1) / blockDim.x;
  const int a = threadIdx.x * per, e = min(a + per, n);
  double x = 0.0;
  for (int k = a; k < e; ++k) x += part[k];
  return block_sum(x, sh);
}

// After the up pass: beta_{k+1} = |r|; the step's scalars advance, or the
// chain stops after nmax steps or on breakdown.
__global__ __launch_bounds__(256) void k_lz_beta(LzChain L, int nparts) {
  __shared__ double sh[16];
  const LzScalars s = *L.sc;
  if (s.stopped || !s.pad) return;
  const bool more = (s.k + 1 < L.nmax);
  const double s2 = more ? ordered_sum_blocks(L.part_up, nparts, sh) : 0.0;
  if (threadIdx.x == 0) {
    const double b = sqrt(s2);
    LzScalars n = s;
    n.pad = 0;
    if (more) L.beta[s.k + 1] = b;
    const bool brk = !(b > L.brk_tol * (fabs(s.alpha_k) + s.beta_k));
    if (!more || brk) {
      n.stopped = 1;
    } else {
      n.s_old = s.s_cur;
      n.s_cur = b;
      n.beta_k = b;
      n.k = s.k + 1;
    }
    *L.sc = n;
  }
}

// After the down pass: alpha_k; phase B of step k becomes pending.
__global__ __launch_bounds__(256) void k_lz_alpha(LzChain L, int nparts) {
  __shared__ double sh[16];
  const LzScalars s = *L.sc;
  if (s.stopped) return;
  const double a = ordered_sum_blocks(L.part_dw, nparts, sh);
  if (threadIdx.x == 0) {
    L.alpha[s.k] = a;
    LzScalars n = s;
    n.alpha_k = a;
    n.nsteps = s.k + 1;
    n.pad = 1;
    *L.sc = n;
  }
}

// ------------------------------------------------ eigenvector from a basis
// x (natural layout, fp64) = sum_k w_k q_k over m stored vectors with column
// stride ld, k in step order (the order in which the replay adds them).
template <typename T>
__global__ __launch_bounds__(256) void k_basis_combine(const T* __restrict__ basis, size_t stride,
                                                       const double* __restrict__ w, int m, int dimup, int64_t n,
                                                       int ld, double* __restrict__ x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t idw = i / dimup;
    const size_t p = (size_t)idw * ld + (size_t)(i - idw * dimup);
    double acc = 0.0;
    int k = 0;
    for (; k + 8 <= m; k += 8) {
      T v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = basis[(size_t)(k + j) * stride + p];
#pragma unroll
      for (int j = 0; j < 8; ++j) acc += w[k + j] * (double)v[j];
    }
    for (; k < m; ++k) acc += w[k] * (double)basis[(size_t)k * stride + p];
    x[i] = acc;
  }
}

template <typename T>
void launch_basis_combine(const T* basis, const TwoPassView& V, const double* w, int m, double* x, hipStream_t st) {
  const int64_t n = (int64_t)V.dimup * V.dimdw;
  const size_t stride = (size_t)V.ld * V.dimdw;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 256 * 64);
  hipLaunchKernelGGL(k_basis_combine<T>, dim3((unsigned)blocks), dim3(256), 0, st, basis, stride, w, m, V.dimup, n,
                     V.ld, x);
  EDL_HIP(hipGetLastError());
}
template void launch_basis_combine<double>(const double*, const TwoPassView&, const double*, int, double*,
                                           hipStream_t);
template void launch_basis_combine<float>(const float*, const TwoPassView&, const double*, int, double*, hipStream_t);

// ---------------------------------------------------------------- host side
static int lds_limit_bytes() { return 160 * 1024; }

// Edge colouring of a bipartite multigraph (lanes x banks) with as many
// colours as its largest degree (Konig's theorem): each edge takes a colour
// free at both ends, after swapping two colours along an alternating path
// when the free colours differ.  color[e] receives the slot of edge e.
static int color_bipartite(const std::vector<int>& el, const std::vector<int>& ek, int nl, int nk,
                           std::vector<int>& color) {
  const int ne = (int)el.size();
  std::vector<int> dl(nl, 0), dk(nk, 0);
  for (int e = 0; e < ne; ++e) {
    ++dl[el[e]];
    ++dk[ek[e]];
  }
  int D = 0;
  for (int v : dl) D = std::max(D, v);
  for (int v : dk) D = std::max(D, v);
  color.assign(ne, -1);
  if (D == 0) return 0;
  std::vector<int> al((size_t)nl * D, -1), ak((size_t)nk * D, -1);
  std::vector<int> path;
  for (int e = 0; e < ne; ++e) {
    const int l = el[e], k = ek[e];
    int a = 0, b = 0;
    while (al[(size_t)l * D + a] >= 0) ++a;
    while (ak[(size_t)k * D + b] >= 0) ++b;
    if (ak[(size_t)k * D + a] >= 0) {
      // the a/b alternating path from bank k: k -a- l1 -b- k2 -a- ...
      path.clear();
      int at = k, col = a;
      bool on_bank = true;
      for (;;) {
        const int f = on_bank ? ak[(size_t)at * D + col] : al[(size_t)at * D + col];
        if (f < 0) break;
        path.push_back(f);
        at = on_bank ? el[f] : ek[f];
        on_bank = !on_bank;
        col = (col == a) ? b : a;
      }
      for (int f : path) {
        al[(size_t)el[f] * D + color[f]] = -1;
        ak[(size_t)ek[f] * D + color[f]] = -1;
      }
      for (int f : path) {
        color[f] = (color[f] == a) ? b : a;
        al[(size_t)el[f] * D + color[f]] = f;
        ak[(size_t)ek[f] * D + color[f]] = f;
      }
    }
    EDL_CHECK(al[(size_t)l * D + a] < 0 && ak[(size_t)k * D + a] < 0, "color_bipartite: no common free colour");
    color[e] = a;
    al[(size_t)l * D + a] = e;
    ak[(size_t)k * D + a] = e;
  }
  return D;
}

// The bank-scheduled table of W-dword rows (see ColorTable): each half-wave
// group of 32 strings is coloured on its own, the block of 64 takes the
// larger slot count of its two halves.
static void build_color_table(SpinTable& t, int W, hipStream_t st) {
  ColorTable& ct = t.pkc[W];
  if (ct.words || !t.packed_ok || t.dim < 1) return;
  const int dim = t.dim, kmax = t.kmax;
  // Lanes served together and the bank classes they must spread over: a
  // row of W dwords, W odd: 32 lanes over the 32 single-dword banks (class
  // W * row mod 32); W = 2: 16 lanes over 16 bank pairs (class row mod 16),
  // a 64-bit read serving 16 lanes of 8 bytes at a time.
  const int L = (W % 2) ? 32 : 16, G = 64 / L;
  auto cls = [&](int64_t r) { return (int)((W % 2) ? (W * r) % 32 : r % 16); };
  std::vector<uint32_t> pk((size_t)std::max(kmax, 1) * dim);
  std::vector<int32_t> cnt(dim);
  EDL_HIP(hipMemcpyAsync(pk.data(), t.packed, sizeof(uint32_t) * pk.size(), hipMemcpyDeviceToHost, st));
  EDL_HIP(hipMemcpyAsync(cnt.data(), t.cnt, sizeof(int32_t) * dim, hipMemcpyDeviceToHost, st));
  EDL_HIP(hipStreamSynchronize(st));
  const int nb = (dim + 63) / 64;
  std::vector<int32_t> off(nb), ns(nb);
  std::vector<uint32_t> words;
  std::vector<int> el, ek, color;
  std::vector<uint32_t> ew;
  for (int b = 0; b < nb; ++b) {
    std::vector<int> nsh(G, 0);
    std::vector<std::vector<int>> col(G), lanes(G);
    std::vector<std::vector<uint32_t>> wds(G);
    for (int h = 0; h < G; ++h) {
      el.clear();
      ek.clear();
      ew.clear();
      for (int l = 0; l < L; ++l) {
        const int r = b * 64 + h * L + l;
        if (r >= dim) break;
        for (int k = 0; k < cnt[r]; ++k) {
          const uint32_t w = pk[(size_t)k * dim + r];
          el.push_back(l);
          ek.push_back(cls(w & 0xFFFFFFu));
          ew.push_back(w);
        }
      }
      nsh[h] = color_bipartite(el, ek, L, L, color);
      col[h] = color;
      wds[h] = ew;
      lanes[h] = el;
    }
    const int nsb = *std::max_element(nsh.begin(), nsh.end());
    off[b] = (int32_t)words.size();
    ns[b] = nsb;
    EDL_CHECK(words.size() + (size_t)nsb * 64 < ((size_t)1 << 31), "build_color_table: table exceeds 2^31 words");
    words.resize(words.size() + (size_t)nsb * 64, 0u);
    uint32_t* blk = words.data() + off[b];
    for (int h = 0; h < G; ++h) {
      std::vector<char> used((size_t)nsb * L, 0);
      std::vector<char> filled((size_t)nsb * L, 0);
      for (size_t e = 0; e < wds[h].size(); ++e) {
        const int s = col[h][e], l = lanes[h][e];
        blk[(size_t)s * 64 + h * L + l] = wds[h][e];
        filled[(size_t)s * L + l] = 1;
        used[(size_t)s * L + cls(wds[h][e] & 0xFFFFFFu)] = 1;
      }
      for (int s = 0; s < nsb; ++s) {
        // padding: cid 255 at one address of a class this slot leaves free
        int freecls = 0;
        while (freecls < L && used[(size_t)s * L + freecls]) ++freecls;
        uint32_t padrow = 0;
        if (freecls < L)
          for (int r = 0; r < std::min(dim, 32 * W + 32); ++r)
            if (cls(r) == freecls) {
              padrow = (uint32_t)r;
              break;
            }
        for (int l = 0; l < L; ++l)
          if (!filled[(size_t)s * L + l]) blk[(size_t)s * 64 + h * L + l] = padrow | (255u << 24);
      }
    }
  }
  ct.slots = (int64_t)words.size() / 64;
  // Two slots to a word where the spin has one hopping magnitude (cids 0
  // and 1 only: lut[1] = -lut[0]) and the rows plus 32 zero rows fit 15
  // bits.  A slot's padding (cid 255) becomes the zero row in the bank of
  // its old padding address, so the bank schedule is kept.
  std::vector<uint32_t> w16;
  std::vector<int32_t> off16(nb), ns16(nb);
  const bool h16 = t.nvals == 1 && dim + 32 <= (1 << 15);
  if (h16) {
    auto half = [&](uint32_t w) -> uint32_t {
      const uint32_t cid = w >> 24, r = w & 0xFFFFFFu;
      if (cid == 255u) {
        const int c = cls(r);
        for (int j = 0; j < 32; ++j)
          if (cls(dim + j) == c) return (uint32_t)(dim + j);
        return (uint32_t)dim;
      }
      return r | ((cid & 1u) << 15);
    };
    for (int b = 0; b < nb; ++b) {
      ns16[b] = (ns[b] + 1) / 2;
      off16[b] = (int32_t)w16.size();
      for (int p = 0; p < ns16[b]; ++p)
        for (int l = 0; l < 64; ++l) {
          const uint32_t* blk = words.data() + off[b];
          const uint32_t lo = half(blk[(size_t)(2 * p) * 64 + l]);
          const uint32_t hi = 2 * p + 1 < ns[b] ? half(blk[(size_t)(2 * p + 1) * 64 + l]) : (uint32_t)dim;
          w16.push_back(lo | (hi << 16));
        }
    }
  }
  // The sign-split form: the +h hops (cid 0) and the -h hops (cid 1) of
  // each half-wave group coloured on their own, so that all lanes of a word
  // share its sign; a slot's unused lanes read the zero row in a bank class
  // the slot leaves free, and a set's odd last slot pairs with zero row dim.
  std::vector<uint32_t> w16s;
  std::vector<int32_t> off16s(nb), ns16s(nb), np16s(nb);
  if (h16) {
    auto zero_row = [&](int c) -> uint32_t {
      for (int j = 0; j < 32; ++j)
        if (cls(dim + j) == c) return (uint32_t)(dim + j);
      return (uint32_t)dim;
    };
    for (int b = 0; b < nb; ++b) {
      int nsl[2] = {0, 0};
      std::vector<std::vector<int>> colS[2], laneS[2];
      std::vector<std::vector<uint32_t>> rowS[2];
      for (int sg = 0; sg < 2; ++sg) {
        colS[sg].resize(G);
        laneS[sg].resize(G);
        rowS[sg].resize(G);
      }
      for (int h = 0; h < G; ++h) {
        std::vector<int> el2[2], ek2[2];
        std::vector<uint32_t> er2[2];
        for (int l = 0; l < L; ++l) {
          const int r = b * 64 + h * L + l;
          if (r >= dim) break;
          for (int k = 0; k < cnt[r]; ++k) {
            const uint32_t w = pk[(size_t)k * dim + r];
            const int sg = (int)((w >> 24) & 1u);
            el2[sg].push_back(l);
            ek2[sg].push_back(cls(w & 0xFFFFFFu));
            er2[sg].push_back(w & 0xFFFFFFu);
          }
        }
        for (int sg = 0; sg < 2; ++sg) {
          nsl[sg] = std::max(nsl[sg], color_bipartite(el2[sg], ek2[sg], L, L, colS[sg][h]));
          laneS[sg][h] = el2[sg];
          rowS[sg][h] = er2[sg];
        }
      }
      off16s[b] = (int32_t)w16s.size();
      np16s[b] = (nsl[0] + 1) / 2;
      ns16s[b] = np16s[b] + (nsl[1] + 1) / 2;
      for (int sg = 0; sg < 2; ++sg) {
        const int nsg = nsl[sg];
        std::vector<uint32_t> slot((size_t)nsg * 64, 0xFFFFFFFFu);
        for (int h = 0; h < G; ++h) {
          std::vector<char> used((size_t)nsg * L, 0);
          for (size_t e = 0; e < rowS[sg][h].size(); ++e) {
            const int sl = colS[sg][h][e], l = laneS[sg][h][e];
            slot[(size_t)sl * 64 + h * L + l] = rowS[sg][h][e];
            used[(size_t)sl * L + cls(rowS[sg][h][e])] = 1;
          }
          for (int sl = 0; sl < nsg; ++sl) {
            int freecls = 0;
            while (freecls < L && used[(size_t)sl * L + freecls]) ++freecls;
            const uint32_t z = zero_row(freecls < L ? freecls : 0);
            for (int l = 0; l < L; ++l)
              if (slot[(size_t)sl * 64 + h * L + l] == 0xFFFFFFFFu) slot[(size_t)sl * 64 + h * L + l] = z;
          }
        }
        for (int p = 0; p < (nsg + 1) / 2; ++p)
          for (int l = 0; l < 64; ++l) {
            const uint32_t lo = slot[(size_t)(2 * p) * 64 + l];
            const uint32_t hi = 2 * p + 1 < nsg ? slot[(size_t)(2 * p + 1) * 64 + l] : (uint32_t)dim;
            w16s.push_back(lo | (hi << 16));
          }
      }
    }
  }
  EDL_HIP(hipMalloc(&ct.words, sizeof(uint32_t) * std::max<size_t>(words.size(), 1)));
  EDL_HIP(hipMalloc(&ct.off, sizeof(int32_t) * nb));
  EDL_HIP(hipMalloc(&ct.ns, sizeof(int32_t) * nb));
  EDL_HIP(hipMemcpyAsync(ct.words, words.data(), sizeof(uint32_t) * words.size(), hipMemcpyHostToDevice, st));
  EDL_HIP(hipMemcpyAsync(ct.off, off.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, st));
  EDL_HIP(hipMemcpyAsync(ct.ns, ns.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, st));
  if (h16) {
    EDL_HIP(hipMalloc(&ct.words16, sizeof(uint32_t) * std::max<size_t>(w16.size(), 1)));
    EDL_HIP(hipMalloc(&ct.off16, sizeof(int32_t) * nb));
    EDL_HIP(hipMalloc(&ct.ns16, sizeof(int32_t) * nb));
    EDL_HIP(hipMemcpyAsync(ct.words16, w16.data(), sizeof(uint32_t) * w16.size(), hipMemcpyHostToDevice, st));
    EDL_HIP(hipMemcpyAsync(ct.off16, off16.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, st));
    EDL_HIP(hipMemcpyAsync(ct.ns16, ns16.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, st));
    EDL_HIP(hipMalloc(&ct.words16s, sizeof(uint32_t) * std::max<size_t>(w16s.size(), 1)));
    EDL_HIP(hipMalloc(&ct.off16s, sizeof(int32_t) * nb));
    EDL_HIP(hipMalloc(&ct.ns16s, sizeof(int32_t) * nb));
    EDL_HIP(hipMalloc(&ct.np16s, sizeof(int32_t) * nb));
    EDL_HIP(hipMemcpyAsync(ct.words16s, w16s.data(), sizeof(uint32_t) * w16s.size(), hipMemcpyHostToDevice, st));
    EDL_HIP(hipMemcpyAsync(ct.off16s, off16s.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, st));
    EDL_HIP(hipMemcpyAsync(ct.ns16s, ns16s.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, st));
    EDL_HIP(hipMemcpyAsync(ct.np16s, np16s.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, st));
  }
  EDL_HIP(hipStreamSynchronize(st));  // the host vectors must outlive the copies
}

// The packed table two hops to a word in table order (SpinTable::pk16).
static void build_pair_table(SpinTable& t, hipStream_t st) {
  if (t.pk16 || !t.packed_ok || t.nvals != 1 || t.dim < 1 || t.dim + 32 > (1 << 15)) return;
  const int dim = t.dim, kmax = std::max(t.kmax, 1), kp = (kmax + 1) / 2;
  std::vector<uint32_t> pk((size_t)kmax * dim);
  std::vector<int32_t> cnt(dim);
  EDL_HIP(hipMemcpyAsync(pk.data(), t.packed, sizeof(uint32_t) * pk.size(), hipMemcpyDeviceToHost, st));
  EDL_HIP(hipMemcpyAsync(cnt.data(), t.cnt, sizeof(int32_t) * dim, hipMemcpyDeviceToHost, st));
  EDL_HIP(hipStreamSynchronize(st));
  const uint32_t zero = (uint32_t)dim;  // the first zero row after the staged strings
  std::vector<uint32_t> w((size_t)kp * dim, zero | (zero << 16));
  for (int a = 0; a < dim; ++a)
    for (int k = 0; k < cnt[a]; ++k) {
      const uint32_t x = pk[(size_t)k * dim + a];
      const uint32_t half = (x & 0xFFFFFFu) | (((x >> 24) & 1u) << 15);
      uint32_t& o = w[(size_t)(k / 2) * dim + a];
      o = (k & 1) ? ((o & 0xFFFFu) | (half << 16)) : ((o & 0xFFFF0000u) | half);
    }
  EDL_HIP(hipMalloc(&t.pk16, sizeof(uint32_t) * w.size()));
  EDL_HIP(hipMemcpyAsync(t.pk16, w.data(), sizeof(uint32_t) * w.size(), hipMemcpyHostToDevice, st));
  EDL_HIP(hipStreamSynchronize(st));  // w must outlive the copy
}

void build_up_color_tables(Sector& s, hipStream_t st) {
  if (s.sp[1].packed_ok) build_pair_table(s.sp[1], st);  // the batched down tiles read it
  if (!s.sp[0].packed_ok) return;
  build_pair_table(s.sp[0], st);
  if (s.sp[0].dim < 64) return;
  build_color_table(s.sp[0], 1, st);
  build_color_table(s.sp[0], 2, st);
  build_color_table(s.sp[0], 3, st);
}

// Column ranges of b top sites for every b whose ranges average at least 16
// columns (RangeTable).
void build_dw_range_tables(Sector& s, hipStream_t st) {
  SpinTable& t = s.sp[1];
  if (!t.packed_ok || t.dim < 256) return;
  const int dim = t.dim;
  std::vector<int32_t> map(dim), off(dim + 1);
  EDL_HIP(hipMemcpyAsync(map.data(), t.map, sizeof(int32_t) * dim, hipMemcpyDeviceToHost, st));
  EDL_HIP(hipMemcpyAsync(off.data(), t.off, sizeof(int32_t) * (dim + 1), hipMemcpyDeviceToHost, st));
  EDL_HIP(hipStreamSynchronize(st));
  std::vector<uint32_t> csr((size_t)off[dim]);
  if (!csr.empty())
    EDL_HIP(hipMemcpyAsync(csr.data(), t.csr, sizeof(uint32_t) * csr.size(), hipMemcpyDeviceToHost, st));
  EDL_HIP(hipStreamSynchronize(st));
  int Ns = 0;
  for (int j = 0; j < dim; ++j)
    while (Ns < 31 && (map[j] >> Ns) != 0) ++Ns;
  for (int b = 1; b <= kMaxRangeBits && b < Ns; ++b) {
    RangeTable& rt = t.rng[b];
    if (rt.lo) continue;
    std::vector<int32_t> lo{0}, rid(dim);
    for (int j = 1; j < dim; ++j)
      if ((map[j] >> (Ns - b)) != (map[j - 1] >> (Ns - b))) lo.push_back(j);
    lo.push_back(dim);
    const int n = (int)lo.size() - 1;
    if (dim / n < 16) break;  // more top sites only make narrower ranges
    int maxc = 0, maxw = 0;
    for (int g = 0; g < n; ++g) {
      for (int j = lo[g]; j < lo[g + 1]; ++j) rid[j] = g;
      maxc = std::max(maxc, lo[g + 1] - lo[g]);
      maxw = std::max(maxw, off[lo[g + 1]] - off[lo[g]]);
    }
    std::vector<uint32_t> words(csr.size());
    std::vector<int32_t> nloc(dim);
    for (int j = 0; j < dim; ++j) {
      int a = off[j], z = off[j + 1];
      for (int k = off[j]; k < off[j + 1]; ++k) {
        const uint32_t w = csr[k];
        const int nb = (int)(w & 0xFFFFFFu);
        if (rid[nb] == rid[j])
          words[a++] = (uint32_t)(nb - lo[rid[j]]) | (w & 0xFF000000u);
        else
          words[--z] = w;
      }
      nloc[j] = a - off[j];
    }
    EDL_HIP(hipMalloc(&rt.lo, sizeof(int32_t) * (n + 1)));
    EDL_HIP(hipMalloc(&rt.words, sizeof(uint32_t) * std::max<size_t>(words.size(), 1)));
    EDL_HIP(hipMalloc(&rt.off, sizeof(int32_t) * (dim + 1)));
    EDL_HIP(hipMalloc(&rt.nloc, sizeof(int32_t) * dim));
    EDL_HIP(hipMemcpyAsync(rt.lo, lo.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, st));
    EDL_HIP(hipMemcpyAsync(rt.words, words.data(), sizeof(uint32_t) * words.size(), hipMemcpyHostToDevice, st));
    EDL_HIP(hipMemcpyAsync(rt.off, off.data(), sizeof(int32_t) * (dim + 1), hipMemcpyHostToDevice, st));
    EDL_HIP(hipMemcpyAsync(rt.nloc, nloc.data(), sizeof(int32_t) * dim, hipMemcpyHostToDevice, st));
    EDL_HIP(hipStreamSynchronize(st));  // the host vectors must outlive the copies
    rt.n = n;
    rt.maxcols = maxc;
    rt.maxwords = maxw;
  }
}

void build_dw_order(Sector& s, hipStream_t st) {
  SpinTable& t = s.sp[1];
  if (!t.packed_ok || t.dim < 256 || t.ord) return;
  const int dim = t.dim;
  std::vector<int32_t> off(dim + 1);
  EDL_HIP(hipMemcpyAsync(off.data(), t.off, sizeof(int32_t) * (dim + 1), hipMemcpyDeviceToHost, st));
  EDL_HIP(hipStreamSynchronize(st));
  std::vector<uint32_t> csr((size_t)off[dim]);
  if (!csr.empty())
    EDL_HIP(hipMemcpyAsync(csr.data(), t.csr, sizeof(uint32_t) * csr.size(), hipMemcpyDeviceToHost, st));
  EDL_HIP(hipStreamSynchronize(st));
  // Cuthill-McKee: breadth first over the hop graph, a string's unvisited
  // neighbours by increasing degree (then index), each component from a
  // pseudo-peripheral string (the last string reached from the component's
  // string of least degree, then the last reached from that one), which
  // keeps the breadth-first levels narrow.
  auto deg = [&](int j) { return off[j + 1] - off[j]; };
  std::vector<int32_t> ord;
  ord.reserve(dim);
  std::vector<int32_t> mark(dim, -1);  // pass that last reached a string
  int pass = 0;
  std::vector<int32_t> level, nb;
  // breadth-first pass `pass` from s over the strings not yet in `ord`,
  // visiting the order into `out`
  auto bfs = [&](int s, std::vector<int32_t>& out) {
    out.clear();
    mark[s] = pass;
    out.push_back(s);
    for (size_t head = 0; head < out.size(); ++head) {
      const int u = out[head];
      nb.clear();
      for (int k = off[u]; k < off[u + 1]; ++k) {
        const int v = (int)(csr[k] & 0xFFFFFFu);
        if (mark[v] != pass && mark[v] != -2) {
          mark[v] = pass;
          nb.push_back(v);
        }
      }
      std::sort(nb.begin(), nb.end(), [&](int x, int y) { return deg(x) != deg(y) ? deg(x) < deg(y) : x < y; });
      out.insert(out.end(), nb.begin(), nb.end());
    }
    ++pass;
  };
  std::vector<int32_t> starts(dim);
  std::iota(starts.begin(), starts.end(), 0);
  std::stable_sort(starts.begin(), starts.end(), [&](int x, int y) { return deg(x) < deg(y); });
  for (int s0 : starts) {
    // (one round per component where the hops go both ways; a one-way hop
    // pattern can leave s0 outside the last pass's strings: then again)
    while (mark[s0] != -2) {
      bfs(s0, level);
      bfs(level.back(), level);
      bfs(level.back(), level);
      for (int v : level) mark[v] = -2;
      ord.insert(ord.end(), level.begin(), level.end());
    }
  }
  EDL_CHECK((int)ord.size() == dim, "build_dw_order: the order does not cover the strings");
  std::vector<int32_t> offo(dim + 1, 0);
  std::vector<uint32_t> csro(csr.size());
  for (int k = 0; k < dim; ++k) {
    const int j = ord[k];
    offo[k + 1] = offo[k] + deg(j);
    std::copy(csr.begin() + off[j], csr.begin() + off[j + 1], csro.begin() + offo[k]);
  }
  int32_t* d_ord = nullptr;
  uint32_t* d_csro = nullptr;
  int32_t* d_offo = nullptr;
  try {
    EDL_HIP(hipMalloc(&d_ord, sizeof(int32_t) * dim));
    EDL_HIP(hipMalloc(&d_csro, sizeof(uint32_t) * std::max<size_t>(csro.size(), 1)));
    EDL_HIP(hipMalloc(&d_offo, sizeof(int32_t) * (dim + 1)));
    EDL_HIP(hipMemcpyAsync(d_ord, ord.data(), sizeof(int32_t) * dim, hipMemcpyHostToDevice, st));
    EDL_HIP(hipMemcpyAsync(d_csro, csro.data(), sizeof(uint32_t) * csro.size(), hipMemcpyHostToDevice, st));
    EDL_HIP(hipMemcpyAsync(d_offo, offo.data(), sizeof(int32_t) * (dim + 1), hipMemcpyHostToDevice, st));
    EDL_HIP(hipStreamSynchronize(st));  // the host vectors must outlive the copies
  } catch (...) {
    if (d_ord) (void)hipFree(d_ord);
    if (d_csro) (void)hipFree(d_csro);
    if (d_offo) (void)hipFree(d_offo);
    throw;
  }
  t.ord = d_ord;
  t.csro = d_csro;
  t.offo = d_offo;
}

static size_t up_lds_bytes(int dimup, int C, size_t sz) {
  const size_t head = 256 * sz;
  return head + (16 - head % 16) % 16 + (size_t)dimup * C * sz;
}

bool two_pass_ok(const Sector& s, size_t elem) {
  return s.sp[0].packed_ok && s.sp[1].packed_ok &&
         up_lds_bytes(s.sp[0].dim, 1, elem) <= (size_t)lds_limit_bytes() - 256;
}

int padded_ld(int dimup) { return (dimup + 31) / 32 * 32; }

TwoPassView two_pass_view(const Sector& s, int ld) {
  EDL_CHECK(ld >= s.sp[0].dim, "two_pass_view: column stride below DimUp");
  TwoPassView v{};
  v.dimup = s.sp[0].dim;
  v.dimdw = s.sp[1].dim;
  v.ld = ld;
  v.kup = s.sp[0].kmax;
  v.kdw = s.sp[1].kmax;
  v.nv_up = s.sp[0].nvals;
  v.map_up = s.sp[0].map;
  v.map_dw = s.sp[1].map;
  v.pk_up = s.sp[0].packed;
  v.pk_dw = s.sp[1].packed;
  v.cnt_up = s.sp[0].cnt;
  v.cnt_dw = s.sp[1].cnt;
  v.dg_up = s.sp[0].diag;
  v.dg_dw = s.sp[1].diag;
  v.csr_dw = s.sp[1].csr;
  v.off_dw = s.sp[1].off;
  v.ord_dw = s.sp[1].ord;
  v.csro_dw = s.sp[1].csro;
  v.offo_dw = s.sp[1].offo;
  for (int W = 0; W < 4; ++W) {
    v.pkc_up[W] = s.sp[0].pkc[W].words;
    v.pkc_off[W] = s.sp[0].pkc[W].off;
    v.pkc_ns[W] = s.sp[0].pkc[W].ns;
    v.pkh_up[W] = s.sp[0].pkc[W].words16;
    v.pkh_off[W] = s.sp[0].pkc[W].off16;
    v.pkh_ns[W] = s.sp[0].pkc[W].ns16;
    v.pks_up[W] = s.sp[0].pkc[W].words16s;
    v.pks_off[W] = s.sp[0].pkc[W].off16s;
    v.pks_ns[W] = s.sp[0].pkc[W].ns16s;
    v.pks_np[W] = s.sp[0].pkc[W].np16s;
  }
  v.pk16_up = s.sp[0].pk16;
  v.pk16_dw = s.sp[1].pk16;
  for (int b = 0; b <= kMaxRangeBits; ++b) {
    const RangeTable& rt = s.sp[1].rng[b];
    v.rng_lo[b] = rt.lo;
    v.rng_words[b] = rt.words;
    v.rng_off[b] = rt.off;
    v.rng_nloc[b] = rt.nloc;
    v.rng_n[b] = rt.n;
    v.rng_maxcols[b] = rt.maxcols;
    v.rng_maxwords[b] = rt.maxwords;
  }
  return v;
}

template <typename T, int C, int RPT, int UKCH, bool LREG, bool COL, bool H16 = false>
static void launch_up_l(const TwoPassView& V, const InterParams& I, const T* lut, T* qa, T* qb, T* tout,
                        const LzChain* L, int grid, int threads, hipStream_t st) {
  const size_t lds = H16 ? (size_t)(V.dimup + 32) * C * sizeof(T) : up_lds_bytes(V.dimup, C, sizeof(T));
  EDL_CHECK(lds <= (size_t)lds_limit_bytes() - 256, "launch_pass_up: panel exceeds LDS");
  set_lds_limit((const void*)k_pass_up<T, C, RPT, UKCH, LREG, COL, H16>, lds_limit_bytes() - 256);
  hipLaunchKernelGGL((k_pass_up<T, C, RPT, UKCH, LREG, COL, H16>), dim3(grid), dim3(threads), lds, st, V, I, lut, qa,
                     qb, tout, L ? *L : LzChain{}, L ? 1 : 0);
  EDL_HIP(hipGetLastError());
}

bool up_col_ok(const TwoPassView& V, int C, size_t elem, int col) {
  const int W = C * (int)elem / 4;
  // 5, 6: the forms 2, 4 with the next panel prefetched (launch checks DimUp
  // against kUpPrefetchRows * threads)
  if (col == 5 || col == 6)
    return (col == 5 ? (elem == 4 ? (C == 1 || C == 3) : C == 1) : true) && up_col_ok(V, C, elem, col == 5 ? 2 : 4);
  if (col == 4) {  // the sign-split words of the two-hop bank-scheduled tables
    if (!((elem == 8 && C == 1) || (elem == 4 && (C == 1 || C == 3)))) return false;
    return V.pks_up[W] != nullptr && V.nv_up == 1 &&
           up_lds_bytes(V.dimup + 32, C, elem) <= (size_t)lds_limit_bytes() - 256;
  }
  if (col == 3)
    return V.pk16_up != nullptr && (size_t)(V.dimup + 32) * C * elem <= (size_t)lds_limit_bytes() - 256;
  if (elem == 8 && C == 1 && col == 2)  // 8-byte rows: 16 lanes over 16 bank pairs
    return V.pkh_up[W] != nullptr && V.nv_up == 1 &&
           up_lds_bytes(V.dimup + 32, C, elem) <= (size_t)lds_limit_bytes() - 256;
  if (elem != 4 || (C != 1 && C != 3)) return false;
  if (col == 2)
    return V.pkh_up[W] != nullptr && V.nv_up == 1 &&
           up_lds_bytes(V.dimup + 32, C, elem) <= (size_t)lds_limit_bytes() - 256;
  return col == 1 && V.pkc_up[W] != nullptr;
}

template <typename T, int C, int RPT, int UKCH>
static void launch_up_s(const TwoPassView& V, const InterParams& I, const T* lut, T* qa, T* qb, T* tout,
                        const LzChain* L, int grid, int threads, hipStream_t st) {
  const size_t lds = (size_t)(V.dimup + 32) * C * sizeof(T);
  EDL_CHECK(lds <= (size_t)lds_limit_bytes() - 256, "launch_pass_up: panel exceeds LDS");
  set_lds_limit((const void*)k_pass_ups<T, C, RPT, UKCH>, lds_limit_bytes() - 256);
  hipLaunchKernelGGL((k_pass_ups<T, C, RPT, UKCH>), dim3(grid), dim3(threads), lds, st, V, I, lut, qa, qb, tout,
                     L ? *L : LzChain{}, L ? 1 : 0);
  EDL_HIP(hipGetLastError());
}

template <typename T, int C, int RPT, int UKCH, bool SGN>
static void launch_up_p(const TwoPassView& V, const InterParams& I, const T* lut, T* qa, T* qb, T* tout,
                        const LzChain* L, int grid, int threads, hipStream_t st) {
  const size_t lds = (size_t)(V.dimup + 32) * C * sizeof(T);
  EDL_CHECK(lds <= (size_t)lds_limit_bytes() - 256, "launch_pass_up: panel exceeds LDS");
  EDL_CHECK(V.dimup <= kUpPrefetchRows * threads, "launch_pass_up: the prefetching form takes DimUp <= kUpPrefetchRows * threads");
  set_lds_limit((const void*)k_pass_upp<T, C, RPT, UKCH, SGN>, lds_limit_bytes() - 256);
  hipLaunchKernelGGL((k_pass_upp<T, C, RPT, UKCH, SGN>), dim3(grid), dim3(threads), lds, st, V, I, lut, qa, qb, tout,
                     L ? *L : LzChain{}, L ? 1 : 0);
  EDL_HIP(hipGetLastError());
}

template <typename T, int C, int RPT, int UKCH>
static void launch_up_k(const TwoPassView& V, const InterParams& I, const T* lut, T* qa, T* qb, T* tout,
                        const LzChain* L, int grid, int threads, int col, hipStream_t st) {
  if (col == 5 || col == 6) {
    EDL_CHECK(up_col_ok(V, C, sizeof(T), col), "launch_pass_up: no two-hop bank-scheduled table for this form");
    if constexpr ((sizeof(T) == 8 && C == 1) || (sizeof(T) == 4 && (C == 1 || C == 3))) {
      if (col == 5) return launch_up_p<T, C, RPT, UKCH, false>(V, I, lut, qa, qb, tout, L, grid, threads, st);
      return launch_up_p<T, C, RPT, UKCH, true>(V, I, lut, qa, qb, tout, L, grid, threads, st);
    }
  }
  if (col == 4) {
    EDL_CHECK(up_col_ok(V, C, sizeof(T), 4), "launch_pass_up: no sign-split two-hop table for this sector and width");
    if constexpr ((sizeof(T) == 8 && C == 1) || (sizeof(T) == 4 && (C == 1 || C == 3)))
      return launch_up_s<T, C, RPT, UKCH>(V, I, lut, qa, qb, tout, L, grid, threads, st);
  }
  if (col == 3) {
    EDL_CHECK(up_col_ok(V, C, sizeof(T), 3), "launch_pass_up: no two-hop table for this sector");
    return launch_up_l<T, C, RPT, UKCH, true, false, true>(V, I, lut, qa, qb, tout, L, grid, threads, st);
  }
  if constexpr (sizeof(T) == 8 && C == 1) {
    if (col == 2) {
      EDL_CHECK(up_col_ok(V, C, sizeof(T), 2), "launch_pass_up: no two-hop bank-scheduled table for this sector");
      return launch_up_l<T, C, RPT, UKCH, true, true, true>(V, I, lut, qa, qb, tout, L, grid, threads, st);
    }
  }
  // fp32 only: in fp64 the two coefficient registers cost more than the
  // LDS reads they save (measured slower)
  if constexpr (sizeof(T) == 4 && (C == 1 || C == 3)) {
    if (col == 2) {
      EDL_CHECK(up_col_ok(V, C, sizeof(T), 2), "launch_pass_up: no two-hop bank-scheduled table for this sector");
      return launch_up_l<T, C, RPT, UKCH, true, true, true>(V, I, lut, qa, qb, tout, L, grid, threads, st);
    }
    if (col) {
      EDL_CHECK(up_col_ok(V, C, sizeof(T)), "launch_pass_up: no bank-scheduled table for this panel width");
      if (V.nv_up == 1) return launch_up_l<T, C, RPT, UKCH, true, true>(V, I, lut, qa, qb, tout, L, grid, threads, st);
      return launch_up_l<T, C, RPT, UKCH, false, true>(V, I, lut, qa, qb, tout, L, grid, threads, st);
    }
  }
  EDL_CHECK(!col, "launch_pass_up: the bank-scheduled table takes fp32 panels of 1 or 3 columns");
  if constexpr (sizeof(T) == 4) {
    if (V.nv_up == 1) return launch_up_l<T, C, RPT, UKCH, true, false>(V, I, lut, qa, qb, tout, L, grid, threads, st);
  }
  launch_up_l<T, C, RPT, UKCH, false, false>(V, I, lut, qa, qb, tout, L, grid, threads, st);
}

template <typename T, int C>
static void launch_up_c(const TwoPassView& V, const InterParams& I, const T* lut, T* qa, T* qb, T* tout,
                        const LzChain* L, int grid, const TwoPassCfg& cfg, hipStream_t st) {
  const int th = cfg.threads_up;
  const int col = cfg.up_col;
  switch (cfg.up_rpt * 100 + cfg.up_kch) {
    case 108: return launch_up_k<T, C, 1, 8>(V, I, lut, qa, qb, tout, L, grid, th, col, st);
    case 116: return launch_up_k<T, C, 1, 16>(V, I, lut, qa, qb, tout, L, grid, th, col, st);
    case 208: return launch_up_k<T, C, 2, 8>(V, I, lut, qa, qb, tout, L, grid, th, col, st);
    case 204: return launch_up_k<T, C, 2, 4>(V, I, lut, qa, qb, tout, L, grid, th, col, st);
    default: throw Error("launch_pass_up: unsupported rows per thread / hop chunk");
  }
}

int pass_up_grid(const TwoPassView& V, const TwoPassCfg& cfg) {
  const int npan = (V.dimdw + cfg.C - 1) / cfg.C;
  return cfg.grid_up < npan ? cfg.grid_up : npan;
}

template <typename T>
void launch_pass_up(const TwoPassView& V, const InterParams& I, const T* lut, T* qa, T* qb, T* tout, const LzChain* L,
                    const TwoPassCfg& cfg, hipStream_t st) {
  const int grid = pass_up_grid(V, cfg);
  switch (cfg.C) {
    case 1: launch_up_c<T, 1>(V, I, lut, qa, qb, tout, L, grid, cfg, st); break;
    case 2: launch_up_c<T, 2>(V, I, lut, qa, qb, tout, L, grid, cfg, st); break;
    case 3: launch_up_c<T, 3>(V, I, lut, qa, qb, tout, L, grid, cfg, st); break;
    case 4: launch_up_c<T, 4>(V, I, lut, qa, qb, tout, L, grid, cfg, st); break;
    default: throw Error("launch_pass_up: unsupported panel width");
  }
}

template <typename T>
void launch_nonlocal(const TwoPassView& V, const InterParams& I, T* qa, T* qb, T* tout, const LzChain* L,
                     hipStream_t st) {
  if (I.Jx == 0.0 && I.Jp == 0.0) return;
  const int64_t n = (int64_t)V.dimup * V.dimdw;
  const int64_t g = std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_nonlocal_add<T>, dim3((unsigned)std::max<int64_t>(g, 1)), dim3(256), 0, st, V, I, qa, qb, tout,
                     L ? *L : LzChain{}, L ? 1 : 0);
  EDL_HIP(hipGetLastError());
}
template void launch_nonlocal<double>(const TwoPassView&, const InterParams&, double*, double*, double*,
                                      const LzChain*, hipStream_t);
template void launch_nonlocal<float>(const TwoPassView&, const InterParams&, float*, float*, float*, const LzChain*,
                                     hipStream_t);

void launch_lz_beta(const LzChain& L, const TwoPassView& V, const TwoPassCfg& cfg, hipStream_t st) {
  hipLaunchKernelGGL(k_lz_beta, dim3(1), dim3(256), 0, st, L, pass_up_grid(V, cfg));
  EDL_HIP(hipGetLastError());
}

void launch_lz_alpha(const LzChain& L, const TwoPassView& V, const TwoPassCfg& cfg, hipStream_t st) {
  hipLaunchKernelGGL(k_lz_alpha, dim3(1), dim3(256), 0, st, L, pass_dw_grid(V, cfg));
  EDL_HIP(hipGetLastError());
}

int pass_dw_grid(const TwoPassView& V, const TwoPassCfg& cfg) {
  const int npan = (V.dimup + cfg.R - 1) / cfg.R;
  if (cfg.dw_vec == 1 && cfg.dw_rng == 0) return (int)(8 * (((int64_t)npan * cfg.nchunk + 7) / 8));
  if (cfg.dw_rng > 0) return 8 * V.rng_n[cfg.dw_rng] * ((npan + 7) / 8);
  return 8 * cfg.nchunk * ((npan + 7) / 8);
}

static size_t pass_dwr_lds(const TwoPassView& V, int R, int b, size_t elem) {
  return elem * (size_t)V.rng_maxcols[b] * R + sizeof(uint32_t) * (size_t)V.rng_maxwords[b] +
         sizeof(int) * (2 * (size_t)V.rng_maxcols[b] + 1);
}

bool pass_dwr_ok(const TwoPassView& V, int R, int b, size_t elem) {
  return b >= 1 && b <= kMaxRangeBits && V.rng_lo[b] != nullptr && (R == 32 || R == 64) &&
         pass_dwr_lds(V, R, b, elem) <= (size_t)lds_limit_bytes() - 4096;
}

template <typename T, int MODE, int NCOL, int KCH>
static void launch_dw_n(const TwoPassView& V, const T* lut, T* qa, T* qb, const T* tin, T* out, const LzChain& A,
                        const TwoPassCfg& cfg, size_t lds, int grid, hipStream_t st) {
  for (auto f : {(const void*)k_pass_dw<T, 32, MODE, NCOL, KCH>, (const void*)k_pass_dw<T, 64, MODE, NCOL, KCH>})
    set_lds_limit(f, lds_limit_bytes() - 4096);
  if (cfg.R == 32)
    hipLaunchKernelGGL((k_pass_dw<T, 32, MODE, NCOL, KCH>), dim3(grid), dim3(cfg.threads_dw), lds, st, V, lut, qa, qb,
                       tin, out, A, cfg.nchunk);
  else if (cfg.R == 64)
    hipLaunchKernelGGL((k_pass_dw<T, 64, MODE, NCOL, KCH>), dim3(grid), dim3(cfg.threads_dw), lds, st, V, lut, qa, qb,
                       tin, out, A, cfg.nchunk);
  else
    throw Error("launch_pass_dw: unsupported panel height");
  EDL_HIP(hipGetLastError());
}

// Dynamic LDS of a down-pass block: the chunk's hop words, their offsets
// and (vector-gather form) its columns, bounded by ncmax * (kdw + 2) words,
// ncmax = ceil(dimdw / nchunk) + 1 (a chunk has at most ncmax - 1 columns).
size_t pass_dw_lds_bytes(const TwoPassView& V, const TwoPassCfg& cfg) {
  if (cfg.dw_rng > 0) return (size_t)-1;  // depends on the element size: see pass_dwr_ok
  const size_t ncmax = (size_t)((V.dimdw + cfg.nchunk - 1) / cfg.nchunk + 1);
  return sizeof(uint32_t) * ncmax * (size_t)(V.kdw + 2);
}

template <typename T, int R, int MODE, int NCOL, int KCH>
static void launch_dwv_k(const TwoPassView& V, const T* lut, T* qa, T* qb, const T* tin, T* out, const LzChain& A,
                         const TwoPassCfg& cfg, size_t lds, int grid, hipStream_t st) {
  set_lds_limit((const void*)k_pass_dwv<T, R, MODE, NCOL, KCH>, lds_limit_bytes() - 4096);
  hipLaunchKernelGGL((k_pass_dwv<T, R, MODE, NCOL, KCH>), dim3(grid), dim3(cfg.threads_dw), lds, st, V, lut, qa, qb,
                     tin, out, A, cfg.nchunk);
  EDL_HIP(hipGetLastError());
}

template <typename T, int MODE, int NCOL, int KCH>
static void launch_dwv_n(const TwoPassView& V, const T* lut, T* qa, T* qb, const T* tin, T* out, const LzChain& A,
                         const TwoPassCfg& cfg, size_t lds, int grid, hipStream_t st) {
  switch (cfg.R) {
    case 16: return launch_dwv_k<T, 16, MODE, NCOL, KCH>(V, lut, qa, qb, tin, out, A, cfg, lds, grid, st);
    case 32: return launch_dwv_k<T, 32, MODE, NCOL, KCH>(V, lut, qa, qb, tin, out, A, cfg, lds, grid, st);
    case 64: return launch_dwv_k<T, 64, MODE, NCOL, KCH>(V, lut, qa, qb, tin, out, A, cfg, lds, grid, st);
    case 128: return launch_dwv_k<T, 128, MODE, NCOL, KCH>(V, lut, qa, qb, tin, out, A, cfg, lds, grid, st);
    case 256:
      if constexpr (sizeof(T) == 4)
        return launch_dwv_k<T, 256, MODE, NCOL, KCH>(V, lut, qa, qb, tin, out, A, cfg, lds, grid, st);
      [[fallthrough]];
    default: throw Error("launch_pass_dw: unsupported panel height of the vector-gather form");
  }
}

bool pass_dw_vec_ok(const TwoPassView& V, const TwoPassCfg& cfg, size_t elem, const void* const* bufs, int nbufs) {
  const bool rok = cfg.R == 16 || cfg.R == 32 || cfg.R == 64 || cfg.R == 128 || (cfg.R == 256 && elem == 4);
  if (!cfg.dw_vec || !rok || V.ld % 32 != 0) return false;
  for (int i = 0; i < nbufs; ++i)
    if (bufs[i] && reinterpret_cast<uintptr_t>(bufs[i]) % 16 != 0) return false;
  return true;
}

template <typename T, int R, int MODE>
static void launch_dwr_k(const TwoPassView& V, const T* lut, T* qa, T* qb, const T* tin, T* out, const LzChain& A,
                         const TwoPassCfg& cfg, hipStream_t st) {
  const size_t lds = pass_dwr_lds(V, R, cfg.dw_rng, sizeof(T));
  set_lds_limit((const void*)k_pass_dwr<T, R, MODE>, lds_limit_bytes() - 4096);
  hipLaunchKernelGGL((k_pass_dwr<T, R, MODE>), dim3(pass_dw_grid(V, cfg)), dim3(cfg.threads_dw), lds, st, V, lut, qa,
                     qb, tin, out, A, cfg.dw_rng);
  EDL_HIP(hipGetLastError());
}

template <typename T, int MODE>
static void launch_dw_r(const TwoPassView& V, const T* lut, T* qa, T* qb, const T* tin, T* out, const LzChain& A,
                        const TwoPassCfg& cfg_in, hipStream_t st) {
  TwoPassCfg cfg = cfg_in;
  if (cfg.dw_rng > 0) {
    EDL_CHECK(pass_dwr_ok(V, cfg.R, cfg.dw_rng, sizeof(T)) && cfg.threads_dw % 64 == 0,
              "launch_pass_dw: no range tables for this top-site count, or the range exceeds LDS");
    if (cfg.R == 32) return launch_dwr_k<T, 32, MODE>(V, lut, qa, qb, tin, out, A, cfg, st);
    return launch_dwr_k<T, 64, MODE>(V, lut, qa, qb, tin, out, A, cfg, st);
  }
  const void* bufs[4] = {qa, qb, tin, out};
  if (cfg.dw_vec && !pass_dw_vec_ok(V, cfg, sizeof(T), bufs, 4)) {
    // natural layout (ld = DimUp) or unaligned caller buffers: the 4-byte form
    // (operator mode only: a chain's partial count follows its own form)
    EDL_CHECK(MODE == 0, "launch_pass_dw: the chain step's vector-gather down pass needs the padded, aligned layout");
    cfg.dw_vec = 0;
    cfg.ncol = 2;
    cfg.kch = 8;
    if (cfg.R != 32 && cfg.R != 64) cfg.R = 64;
  }
  EDL_CHECK(cfg.nchunk >= 1 && cfg.nchunk <= V.dimdw, "launch_pass_dw: bad column chunk count");
  const size_t lds = pass_dw_lds_bytes(V, cfg);
  EDL_CHECK(lds <= (size_t)lds_limit_bytes() - 4096, "launch_pass_dw: column table exceeds LDS");
  const int grid = pass_dw_grid(V, cfg);
  if (cfg.dw_vec) {
    // the locality order of columns where asked for, else the natural one
    TwoPassView Vo = V;
    if (cfg.dw_ord) {
      EDL_CHECK(V.ord_dw && V.csro_dw && V.offo_dw, "launch_pass_dw: no column order built for this sector");
    } else {
      Vo.ord_dw = nullptr;
      Vo.csro_dw = nullptr;
      Vo.offo_dw = nullptr;
    }
    switch (cfg.ncol * 100 + cfg.kch) {
      case 102: return launch_dwv_n<T, MODE, 1, 2>(Vo, lut, qa, qb, tin, out, A, cfg, lds, grid, st);
      case 104: return launch_dwv_n<T, MODE, 1, 4>(Vo, lut, qa, qb, tin, out, A, cfg, lds, grid, st);
      case 201: return launch_dwv_n<T, MODE, 2, 1>(Vo, lut, qa, qb, tin, out, A, cfg, lds, grid, st);
      case 202: return launch_dwv_n<T, MODE, 2, 2>(Vo, lut, qa, qb, tin, out, A, cfg, lds, grid, st);
      case 204: return launch_dwv_n<T, MODE, 2, 4>(Vo, lut, qa, qb, tin, out, A, cfg, lds, grid, st);
      default: throw Error("launch_pass_dw: unsupported columns per wave / hop chunk of the vector-gather form");
    }
  }
  const int code = cfg.ncol * 100 + cfg.kch;
  switch (code) {
    case 108: return launch_dw_n<T, MODE, 1, 8>(V, lut, qa, qb, tin, out, A, cfg, lds, grid, st);
    case 116: return launch_dw_n<T, MODE, 1, 16>(V, lut, qa, qb, tin, out, A, cfg, lds, grid, st);
    case 132: return launch_dw_n<T, MODE, 1, 32>(V, lut, qa, qb, tin, out, A, cfg, lds, grid, st);
    case 208: return launch_dw_n<T, MODE, 2, 8>(V, lut, qa, qb, tin, out, A, cfg, lds, grid, st);
    case 216: return launch_dw_n<T, MODE, 2, 16>(V, lut, qa, qb, tin, out, A, cfg, lds, grid, st);
    default: throw Error("launch_pass_dw: unsupported columns per lane / hop chunk");
  }
}

template <typename T>
void launch_pass_dw(const TwoPassView& V, const T* lut, T* qa, T* qb, const T* tin, T* out, const LzChain* A,
                    const TwoPassCfg& cfg, hipStream_t st) {
  if (A) launch_dw_r<T, 1>(V, lut, qa, qb, tin, out, *A, cfg, st);
  else launch_dw_r<T, 0>(V, lut, qa, qb, tin, out, LzChain{}, cfg, st);
}

// ------------------------------------------------------- segmented up pass
static __device__ __forceinline__ int64_t seg_xcd_tile(int64_t n8) {
  return (int64_t)(blockIdx.x % 8) * n8 + blockIdx.x / 8;
}

template <typename T, int C, int UKCH>
__global__ __launch_bounds__(1024) void k_up_seg(TwoPassView V, InterParams I, const T* __restrict__ lut_up, T* qa,
                                                 T* qb, T* __restrict__ tout, LzChain L, int chain,
                                                 const int* __restrict__ seg, int nseg, int64_t ntiles, int64_t n8) {
  typedef typename Vec<T, C>::type VT;
  extern __shared__ __align__(16) unsigned char smem[];
  T* lut = reinterpret_cast<T*>(smem);
  VT* P = reinterpret_cast<VT*>(smem + lds_head_bytes(sizeof(T)));
  // tiles (panel of C columns, segment), panel-major (the segments of a
  // panel are consecutive), and each XCD takes a contiguous range of them
  const int64_t t = seg_xcd_tile(n8);
  if (t >= ntiles) return;
  const T* __restrict__ q = qa;
  if (chain) {
    const LzScalars s = *L.sc;
    if (s.stopped) return;
    q = (s.k & 1) ? qb : qa;
  }
  const int pan = (int)(t / nseg), sg = (int)(t % nseg);
  const int lo = seg[sg], hi = seg[sg + 1];
  const int dimup = V.dimup, dimdw = V.dimdw;
  const int c0 = pan * C;
  const size_t ld = (size_t)V.ld;
  for (int k = threadIdx.x; k < 256; k += blockDim.x) lut[k] = lut_up[k];
  constexpr int SB = stage_rows<T, C>(32, 1, 16);
  for (int r0 = lo + threadIdx.x; r0 < hi; r0 += SB * blockDim.x) {
    T a[SB][C];
#pragma unroll
    for (int h = 0; h < SB; ++h) {
      const int r = r0 + h * (int)blockDim.x;
#pragma unroll
      for (int j = 0; j < C; ++j) a[h][j] = (r < hi && c0 + j < dimdw) ? q[(size_t)(c0 + j) * ld + r] : (T)0;
    }
#pragma unroll
    for (int h = 0; h < SB; ++h) {
      const int r = r0 + h * (int)blockDim.x;
      if (r >= hi) break;
      VT x;
#pragma unroll
      for (int j = 0; j < C; ++j) Vec<T, C>::set(x, j, a[h][j]);
      P[r - lo] = x;
    }
  }
  __syncthreads();
  uint32_t sd[C], sdm[C];
  double dgd[C];
  const bool one_orb = I.norb == 1 && I.npairs == 0;  // U * popcount(su & sd & mask) only
  const double u0 = I.U[0];
#pragma unroll
  for (int j = 0; j < C; ++j) {
    const int cc = (c0 + j < dimdw) ? c0 + j : c0;
    sd[j] = (uint32_t)V.map_dw[cc];
    sdm[j] = sd[j] & I.orb_mask[0];
    dgd[j] = V.dg_dw[cc];
  }
  for (int row = lo + threadIdx.x; row < hi; row += blockDim.x) {
    const int nh = V.cnt_up[row];
    const uint32_t su = (uint32_t)V.map_up[row];
    const double dgu = V.dg_up[row];
    const VT self = P[row - lo];
    T acc[C];
    if (one_orb) {
#pragma unroll
      for (int j = 0; j < C; ++j)
        acc[j] = (T)(dgu + dgd[j] + u0 * (double)__popc(su & sdm[j])) * Vec<T, C>::get(self, j);
    } else {
#pragma unroll
      for (int j = 0; j < C; ++j) acc[j] = (T)(dgu + dgd[j] + inter2(I, su, sd[j])) * Vec<T, C>::get(self, j);
    }
    for (int k0 = 0; k0 < nh; k0 += UKCH) {
      uint32_t w[UKCH];
#pragma unroll
      for (int j = 0; j < UKCH; ++j) w[j] = (k0 + j < nh) ? V.pk_up[(size_t)(k0 + j) * dimup + row] : (255u << 24);
#pragma unroll
      for (int j = 0; j < UKCH; ++j) {
        if (k0 + j >= nh) continue;
        const int nb = (int)(w[j] & 0xFFFFFFu);
        const T cf = lut[w[j] >> 24];
        VT x;
        if (nb >= lo && nb < hi) {
          x = P[nb - lo];
        } else {
#pragma unroll
          for (int c = 0; c < C; ++c) Vec<T, C>::set(x, c, (c0 + c < dimdw) ? q[(size_t)(c0 + c) * ld + nb] : (T)0);
        }
#pragma unroll
        for (int c = 0; c < C; ++c) acc[c] += cf * Vec<T, C>::get(x, c);
      }
    }
#pragma unroll
    for (int j = 0; j < C; ++j)
      if (c0 + j < dimdw) tout[(size_t)(c0 + j) * ld + row] = acc[j];
  }
}

// Segmented up pass with hop targets by rank arithmetic: the same tiles as
// k_up_seg, but a string's neighbours come from its own bits instead of the
// hop table, which a panel of C columns would otherwise read in full (on
// the 3x6 open cluster the table is 5.3 MB, larger than an XCD's L2, and
// its re-reads were most of the launch's traffic).  For bond {p, q} a row
// string su hops iff exactly one of p, q is occupied; the other string is
// su ^ (1 << p | 1 << q) (the hop is its own inverse, and h is the same in
// both directions), its rank is rlo[..] + rhi[..] (RankHops) read from LDS,
// and the sign is (-1)^popcount(su & sites strictly between p and q) as in
// k_hop_fill.  Bonds come in twelve runs: [local | may leave the segment] x
// [both ends below B | both at or above B | across] x [no site between p
// and q | some], local ones being those below the segments' top-site split,
// whose targets stay in the segment.  Local bonds go four at a time (all
// table reads, then all row reads, each group one LDS round trip); a lane
// whose string does not hop reads the zero row after the segment.
struct RankArgs {
  int n[12];
  int B;
  uint32_t lomask;
  uint32_t m[kMaxBonds];     // 1 << p | 1 << q
  uint32_t mlo4[kMaxBonds];  // (m & lomask) << 2: byte offset change in rlo
  uint32_t mhi4[kMaxBonds];  // (m >> B) << 2: likewise in rhi
  uint32_t btw[kMaxBonds];   // sites strictly between p and q
  double hd[kMaxBonds];
  float hf[kMaxBonds];
  const uint32_t* rlo;
  const uint32_t* rhi;
  int nlo, nhi;
};
constexpr size_t kRankHead = 8 * kRankHalfWords;  // both tables, static LDS of k_up_segr

static __device__ __forceinline__ float rank_h(const RankArgs& R, int b, float) { return R.hf[b]; }
static __device__ __forceinline__ double rank_h(const RankArgs& R, int b, double) { return R.hd[b]; }

// Rank of the string su ^ m (CLS 0: m below B, 1: m at or above B, 2: across).
template <int CLS>
static __device__ __forceinline__ uint32_t rank_target(const uint32_t* rt, const RankArgs& R, int b, uint32_t L4,
                                                       uint32_t H4, uint32_t rl, uint32_t rh) {
  const unsigned char* sm = reinterpret_cast<const unsigned char*>(rt);
  const unsigned char* sh = sm + 4 * kRankHalfWords;
  if (CLS == 0) return *reinterpret_cast<const uint32_t*>(sm + (L4 ^ R.mlo4[b])) + rh;
  if (CLS == 1) return rl + *reinterpret_cast<const uint32_t*>(sh + (H4 ^ R.mhi4[b]));
  return *reinterpret_cast<const uint32_t*>(sm + (L4 ^ R.mlo4[b])) +
         *reinterpret_cast<const uint32_t*>(sh + (H4 ^ R.mhi4[b]));
}

template <typename T, bool SGN>
static __device__ __forceinline__ T rank_coef(const RankArgs& R, int b, uint32_t su) {
  const T h = rank_h(R, b, (T)0);
  return SGN ? flip_sign(h, (uint32_t)__popc(su & R.btw[b]) << 31) : h;
}

template <typename T, int C, int CLS, bool SGN, bool LEAVE>
static __device__ __forceinline__ void rank_bonds(const RankArgs& R, int b0, int b1, uint32_t su, uint32_t L4,
                                                  uint32_t H4, uint32_t rl, uint32_t rh, const uint32_t* rt,
                                                  const typename Vec<T, C>::type* P, uint32_t lo, uint32_t nrows,
                                                  const T* __restrict__ q, size_t ld, int c0, int dimdw, T (&acc)[C]) {
  typedef typename Vec<T, C>::type VT;
  int b = b0;
  if (!LEAVE) {
    constexpr int G = 4;
    for (; b + G <= b1; b += G) {
      uint32_t sel[G];
#pragma unroll
      for (int u = 0; u < G; ++u) {
        const uint32_t m = R.m[b + u], tm = su & m;
        const uint32_t idx = rank_target<CLS>(rt, R, b + u, L4, H4, rl, rh);
        sel[u] = (tm != 0u && tm != m) ? idx - lo : nrows;
      }
      VT x[G];
#pragma unroll
      for (int u = 0; u < G; ++u) x[u] = P[sel[u]];
#pragma unroll
      for (int u = 0; u < G; ++u) {
        const T cf = rank_coef<T, SGN>(R, b + u, su);
#pragma unroll
        for (int c = 0; c < C; ++c) acc[c] += cf * Vec<T, C>::get(x[u], c);
      }
    }
  }
  for (; b < b1; ++b) {
    const uint32_t m = R.m[b], tm = su & m;
    const bool hop = tm != 0u && tm != m;
    const uint32_t idx = rank_target<CLS>(rt, R, b, L4, H4, rl, rh);
    const uint32_t li = idx - lo;
    VT x = P[(hop && li < nrows) ? li : nrows];
    if (LEAVE && hop && li >= nrows) {
#pragma unroll
      for (int c = 0; c < C; ++c) Vec<T, C>::set(x, c, (c0 + c < dimdw) ? q[(size_t)(c0 + c) * ld + idx] : (T)0);
    }
    const T cf = rank_coef<T, SGN>(R, b, su);
#pragma unroll
    for (int c = 0; c < C; ++c) acc[c] += cf * Vec<T, C>::get(x, c);
  }
}

template <typename T, int C>
__global__ __launch_bounds__(1024) void k_up_segr(TwoPassView V, InterParams I, RankArgs R, T* qa, T* qb,
                                                  T* __restrict__ tout, LzChain L, int chain,
                                                  const int* __restrict__ seg, int nseg, int64_t ntiles, int64_t n8) {
  typedef typename Vec<T, C>::type VT;
  __shared__ uint32_t rt[2 * kRankHalfWords];
  extern __shared__ __align__(16) unsigned char smem[];
  VT* P = reinterpret_cast<VT*>(smem);  // segment rows, then one zero row
  const int64_t t = seg_xcd_tile(n8);
  if (t >= ntiles) return;
  const T* __restrict__ q = qa;
  if (chain) {
    const LzScalars s = *L.sc;
    if (s.stopped) return;
    q = (s.k & 1) ? qb : qa;
  }
  const int pan = (int)(t / nseg), sg = (int)(t % nseg);
  const int lo = seg[sg], hi = seg[sg + 1];
  const int dimdw = V.dimdw;
  const int c0 = pan * C;
  const size_t ld = (size_t)V.ld;
  const uint32_t nrows = (uint32_t)(hi - lo);
  for (int k = threadIdx.x; k < R.nlo; k += blockDim.x) rt[k] = R.rlo[k];
  for (int k = threadIdx.x; k < R.nhi; k += blockDim.x) rt[kRankHalfWords + k] = R.rhi[k];
  if (threadIdx.x == 0) {
    VT z;
#pragma unroll
    for (int j = 0; j < C; ++j) Vec<T, C>::set(z, j, (T)0);
    P[nrows] = z;
  }
  constexpr int SB = stage_rows<T, C>(32, 1, 16);
  for (int r0 = lo + threadIdx.x; r0 < hi; r0 += SB * blockDim.x) {
    T a[SB][C];
#pragma unroll
    for (int h = 0; h < SB; ++h) {
      const int r = r0 + h * (int)blockDim.x;
#pragma unroll
      for (int j = 0; j < C; ++j) a[h][j] = (r < hi && c0 + j < dimdw) ? q[(size_t)(c0 + j) * ld + r] : (T)0;
    }
#pragma unroll
    for (int h = 0; h < SB; ++h) {
      const int r = r0 + h * (int)blockDim.x;
      if (r >= hi) break;
      VT x;
#pragma unroll
      for (int j = 0; j < C; ++j) Vec<T, C>::set(x, j, a[h][j]);
      P[r - lo] = x;
    }
  }
  __syncthreads();
  uint32_t sd[C], sdm[C];
  double dgd[C];
  const bool one_orb = I.norb == 1 && I.npairs == 0;
  const double u0 = I.U[0];
#pragma unroll
  for (int j = 0; j < C; ++j) {
    const int cc = (c0 + j < dimdw) ? c0 + j : c0;
    sd[j] = (uint32_t)V.map_dw[cc];
    sdm[j] = sd[j] & I.orb_mask[0];
    dgd[j] = V.dg_dw[cc];
  }
  int e[13];
  e[0] = 0;
#pragma unroll
  for (int k = 0; k < 12; ++k) e[k + 1] = e[k] + R.n[k];
  for (int row = lo + threadIdx.x; row < hi; row += blockDim.x) {
    const uint32_t su = (uint32_t)V.map_up[row];
    const double dgu = V.dg_up[row];
    const VT self = P[row - lo];
    T acc[C];
    if (one_orb) {
#pragma unroll
      for (int j = 0; j < C; ++j)
        acc[j] = (T)(dgu + dgd[j] + u0 * (double)__popc(su & sdm[j])) * Vec<T, C>::get(self, j);
    } else {
#pragma unroll
      for (int j = 0; j < C; ++j) acc[j] = (T)(dgu + dgd[j] + inter2(I, su, sd[j])) * Vec<T, C>::get(self, j);
    }
    const uint32_t L4 = (su & R.lomask) << 2, H4 = (su >> R.B) << 2;
    const uint32_t rl = rt[L4 >> 2], rh = rt[kRankHalfWords + (H4 >> 2)];
#define EDL_RANK_RUN(k, CLS, SGN, LEAVE)                                                                         \
  rank_bonds<T, C, CLS, SGN, LEAVE>(R, e[k], e[k + 1], su, L4, H4, rl, rh, rt, P, (uint32_t)lo, nrows, q, ld, c0, \
                                    dimdw, acc)
    EDL_RANK_RUN(0, 0, false, false);
    EDL_RANK_RUN(1, 0, true, false);
    EDL_RANK_RUN(2, 1, false, false);
    EDL_RANK_RUN(3, 1, true, false);
    EDL_RANK_RUN(4, 2, false, false);
    EDL_RANK_RUN(5, 2, true, false);
    EDL_RANK_RUN(6, 0, false, true);
    EDL_RANK_RUN(7, 0, true, true);
    EDL_RANK_RUN(8, 1, false, true);
    EDL_RANK_RUN(9, 1, true, true);
    EDL_RANK_RUN(10, 2, false, true);
    EDL_RANK_RUN(11, 2, true, true);
#undef EDL_RANK_RUN
#pragma unroll
    for (int j = 0; j < C; ++j)
      if (c0 + j < dimdw) tout[(size_t)(c0 + j) * ld + row] = acc[j];
  }
}

// Phase B of a chain over whole columns (block = columns b, b + grid, ...).
template <typename T>
__global__ __launch_bounds__(256) void k_phase_b(TwoPassView V, T* qa, T* qb, LzChain L) {
  __shared__ double sh[16];
  const LzScalars s = *L.sc;
  if (s.stopped || !s.pad) return;
  const bool last = s.k + 1 >= L.nmax;
  if (last && !L.x) return;
  T* __restrict__ qn = ((s.k + 1) & 1) ? qb : qa;
  const T* __restrict__ qk = (s.k & 1) ? qb : qa;
  const T am = (T)(s.alpha_k / s.s_cur);
  double* __restrict__ x = L.x;
  const double cx = x ? L.ccoef[s.k] / s.s_cur : 0.0;
  const int dimup = V.dimup, dimdw = V.dimdw;
  const size_t ld = (size_t)V.ld;
  double nrm = 0.0;
  for (int col = blockIdx.x; col < dimdw; col += gridDim.x)
    for (int row = threadIdx.x; row < dimup; row += blockDim.x) {
      const size_t i = (size_t)col * ld + row;
      const T b = qk[i];
      if (!last) {
        const T e = qn[i] - am * b;
        qn[i] = e;
        nrm += (double)e * (double)e;
      }
      if (x) x[(size_t)col * dimup + row] += cx * (double)b;
    }
  if (last) return;
  const double v = block_sum(nrm, sh);
  if (threadIdx.x == 0) L.part_up[blockIdx.x] = v;
}

// Deflation of a chain against nx fixed vectors X (fp64, natural layout,
// normalised, mutually orthogonal): after phase B, r -= sum_m (x_m . r) x_m,
// and the block partials of |r|^2 are taken after the projection.  Three
// launches on the phase-B grid: dots, one-block coefficients, update.
// They run only where phase B ran (same scalar tests as k_phase_b).
static __device__ __forceinline__ bool deflate_active(const LzScalars& s, const LzChain& L) {
  return !s.stopped && s.pad && s.k + 1 < L.nmax;
}

template <typename T>
__global__ __launch_bounds__(256) void k_defl_dot(TwoPassView V, T* qa, T* qb, LzChain L,
                                                  const double* __restrict__ X, int nx, double* __restrict__ part) {
  __shared__ double sh[16];
  const LzScalars s = *L.sc;
  if (!deflate_active(s, L)) return;
  const T* __restrict__ r = ((s.k + 1) & 1) ? qb : qa;
  const int dimup = V.dimup, dimdw = V.dimdw;
  const size_t ld = (size_t)V.ld, D = (size_t)dimup * dimdw;
  for (int m = 0; m < nx; ++m) {
    const double* __restrict__ x = X + (size_t)m * D;
    double d = 0.0;
    for (int col = blockIdx.x; col < dimdw; col += gridDim.x)
      for (int row = threadIdx.x; row < dimup; row += blockDim.x)
        d += x[(size_t)col * dimup + row] * (double)r[(size_t)col * ld + row];
    d = block_sum(d, sh);
    if (threadIdx.x == 0) part[(size_t)m * gridDim.x + blockIdx.x] = d;
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_defl_coef(LzChain L, const double* __restrict__ part, int nparts, int nx,
                                                   double* __restrict__ coef) {
  __shared__ double sh[16];
  const LzScalars s = *L.sc;
  if (!deflate_active(s, L)) return;
  for (int m = 0; m < nx; ++m) {
    const double c = ordered_sum_blocks(part + (size_t)m * nparts, nparts, sh);
    if (threadIdx.x == 0) coef[m] = c;
    __syncthreads();
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_defl_update(TwoPassView V, T* qa, T* qb, LzChain L,
                                                     const double* __restrict__ X, int nx,
                                                     const double* __restrict__ coef) {
  __shared__ double sh[16];
  const LzScalars s = *L.sc;
  if (!deflate_active(s, L)) return;
  T* __restrict__ r = ((s.k + 1) & 1) ? qb : qa;
  const int dimup = V.dimup, dimdw = V.dimdw;
  const size_t ld = (size_t)V.ld, D = (size_t)dimup * dimdw;
  double nrm = 0.0;
  for (int col = blockIdx.x; col < dimdw; col += gridDim.x)
    for (int row = threadIdx.x; row < dimup; row += blockDim.x) {
      const size_t n = (size_t)col * dimup + row, i = (size_t)col * ld + row;
      double e = (double)r[i];
      for (int m = 0; m < nx; ++m) e -= coef[m] * X[(size_t)m * D + n];
      const T et = (T)e;
      r[i] = et;
      nrm += (double)et * (double)et;
    }
  const double v = block_sum(nrm, sh);
  if (threadIdx.x == 0) L.part_up[blockIdx.x] = v;
}

__global__ __launch_bounds__(256) void k_lz_beta_n(LzChain L, int nparts) {
  // same as k_lz_beta with an explicit partial count
  __shared__ double sh[16];
  const LzScalars s = *L.sc;
  if (s.stopped || !s.pad) return;
  const bool more = (s.k + 1 < L.nmax);
  const double s2 = more ? ordered_sum_blocks(L.part_up, nparts, sh) : 0.0;
  if (threadIdx.x == 0) {
    const double b = sqrt(s2);
    LzScalars n = s;
    n.pad = 0;
    if (more) L.beta[s.k + 1] = b;
    const bool brk = !(b > L.brk_tol * (fabs(s.alpha_k) + s.beta_k));
    if (!more || brk) {
      n.stopped = 1;
    } else {
      n.s_old = s.s_cur;
      n.s_cur = b;
      n.beta_k = b;
      n.k = s.k + 1;
    }
    *L.sc = n;
  }
}

int seg_row_cap(size_t elem, int C) {
  // room for the larger head of the two kernels (coefficient lut, rank
  // tables) and k_up_segr's zero row
  const size_t head = std::max(lds_head_bytes(elem), kRankHead);
  int cap = (int)(((size_t)lds_limit_bytes() - 256 - head) / (elem * (size_t)C)) - 1;
  if (const char* e = std::getenv("EDL_UP_SEGMENT_ROWS")) {
    const int v = std::atoi(e);
    EDL_CHECK(v >= 1, "EDL_UP_SEGMENT_ROWS must be a positive row count");
    cap = std::min(cap, v);
  }
  return cap;
}

bool seg_path(const Sector& s, size_t elem) {
  if (!(s.sp[0].packed_ok && s.sp[1].packed_ok) || s.sp[0].dim < 1 || s.sp[1].dim < 1)
    return false;
  return !two_pass_ok(s, elem) || seg_row_cap(elem, 1) < s.sp[0].dim;
}

const std::vector<int>& seg_bounds(Sector& s, size_t elem, int C) {
  const int cap = seg_row_cap(elem, C);
  if (s.d_seg && s.seg_cap == cap) return s.seg_h;
  const SpinTable& t = s.sp[0];
  std::vector<int32_t> map(t.dim);
  EDL_HIP(hipMemcpy(map.data(), t.map, sizeof(int32_t) * t.dim, hipMemcpyDeviceToHost));
  int bits = 1;
  while (bits < 31 && (map.back() >> bits) != 0) ++bits;
  // the coarsest split by top-site occupation whose runs all fit
  std::vector<int> runs;
  int sh = bits;
  for (; sh >= 0; --sh) {
    runs.assign(1, 0);
    for (int i = 1; i < t.dim; ++i)
      if ((map[i] >> sh) != (map[i - 1] >> sh)) runs.push_back(i);
    runs.push_back(t.dim);
    int longest = 0;
    for (size_t k = 0; k + 1 < runs.size(); ++k) longest = std::max(longest, runs[k + 1] - runs[k]);
    if (longest <= cap) break;
  }
  EDL_CHECK(sh >= 0, "seg_bounds: no split fits");
  // adjacent runs merged while they fit (fewer segments, more local hops)
  std::vector<int> b(1, 0);
  for (size_t k = 1; k < runs.size(); ++k) {
    if (runs[k] - b.back() > cap) b.push_back(runs[k - 1]);
  }
  b.push_back(t.dim);
  if (s.d_seg) (void)hipFree(s.d_seg);
  s.d_seg = nullptr;
  EDL_HIP(hipMalloc(&s.d_seg, sizeof(int) * b.size()));
  EDL_HIP(hipMemcpy(s.d_seg, b.data(), sizeof(int) * b.size(), hipMemcpyHostToDevice));
  s.seg_h = b;
  s.seg_cap = cap;
  s.seg_sh = sh;
  return s.seg_h;
}

TwoPassCfg seg_dw_cfg(int dimdw) {
  TwoPassCfg p;
  p.R = 64;
  p.threads_dw = 256;
  p.nchunk = std::max(1, std::min(256, dimdw));
  p.ncol = 2;
  p.kch = 8;
  return p;
}

template <typename T, int C>
static void launch_up_seg_c(const TwoPassView& V, const InterParams& I, const T* lut, T* qa, T* qb, T* tout,
                            const LzChain* L, const Sector& s, hipStream_t st) {
  const int nseg = (int)s.seg_h.size() - 1;
  int rows = 0;
  for (int k = 0; k < nseg; ++k) rows = std::max(rows, s.seg_h[k + 1] - s.seg_h[k]);
  const size_t lds = lds_head_bytes(sizeof(T)) + (size_t)rows * C * sizeof(T);
  EDL_CHECK(lds <= (size_t)lds_limit_bytes() - 256, "launch_up_seg: segment exceeds LDS");
  set_lds_limit((const void*)k_up_seg<T, C, 8>, lds_limit_bytes() - 256);
  const int64_t ntiles = (int64_t)((V.dimdw + C - 1) / C) * nseg;
  const int64_t n8 = (ntiles + 7) / 8;
  EDL_CHECK(8 * n8 < ((int64_t)1 << 31), "launch_up_seg: too many tiles");
  hipLaunchKernelGGL((k_up_seg<T, C, 8>), dim3((unsigned)(8 * n8)), dim3(1024), lds, st, V, I, lut, qa, qb, tout,
                     L ? *L : LzChain{}, L ? 1 : 0, s.d_seg, nseg, ntiles, n8);
  EDL_HIP(hipGetLastError());
}

bool seg_rank_ok(const Sector& s) { return s.sp[0].rk.ok; }

// Kernel arguments of k_up_segr: the bonds of the up spin sorted into the
// six runs (local / leaving the segment x below B / at or above B / across).
static RankArgs rank_args(const Sector& s) {
  const RankHops& rk = s.sp[0].rk;
  RankArgs R{};
  R.B = rk.B;
  R.lomask = (uint32_t)rk.nlo - 1u;
  R.rlo = rk.rlo;
  R.rhi = rk.rhi;
  R.nlo = rk.nlo;
  R.nhi = rk.nhi;
  int k = 0;
  for (int run = 0; run < 12; ++run) {
    const bool leave = run >= 6, sgn = run & 1;
    const int cls = (run % 6) / 2;
    for (int b = 0; b < rk.nb; ++b) {
      const int p = rk.p[b], q = rk.q[b];  // p < q
      if ((q >= s.seg_sh) != leave) continue;
      const int c = (q < rk.B) ? 0 : (p >= rk.B ? 1 : 2);
      const uint32_t btw = ((1u << q) - 1u) & ~((2u << p) - 1u);
      if (c != cls || (btw != 0u) != sgn) continue;
      const uint32_t m = (1u << p) | (1u << q);
      R.m[k] = m;
      R.mlo4[k] = (m & R.lomask) << 2;
      R.mhi4[k] = (m >> rk.B) << 2;
      R.btw[k] = btw;
      R.hd[k] = rk.h[b];
      R.hf[k] = (float)rk.h[b];
      ++R.n[run];
      ++k;
    }
  }
  EDL_CHECK(k == rk.nb, "rank_args: bond runs do not cover the bonds");
  return R;
}

template <typename T, int C>
static void launch_up_segr_c(const TwoPassView& V, const InterParams& I, T* qa, T* qb, T* tout, const LzChain* L,
                             const Sector& s, hipStream_t st) {
  const int nseg = (int)s.seg_h.size() - 1;
  int rows = 0;
  for (int k = 0; k < nseg; ++k) rows = std::max(rows, s.seg_h[k + 1] - s.seg_h[k]);
  const size_t lds = (size_t)(rows + 1) * C * sizeof(T);  // dynamic: segment and zero row
  EDL_CHECK(kRankHead + lds <= (size_t)lds_limit_bytes() - 256, "launch_up_seg: segment exceeds LDS");
  const RankArgs R = rank_args(s);
  EDL_CHECK(R.nlo <= kRankHalfWords && R.nhi <= kRankHalfWords, "launch_up_seg: rank tables exceed LDS");
  set_lds_limit((const void*)k_up_segr<T, C>, lds_limit_bytes() - 256 - (int)kRankHead);
  const int64_t ntiles = (int64_t)((V.dimdw + C - 1) / C) * nseg;
  const int64_t n8 = (ntiles + 7) / 8;
  EDL_CHECK(8 * n8 < ((int64_t)1 << 31), "launch_up_seg: too many tiles");
  hipLaunchKernelGGL((k_up_segr<T, C>), dim3((unsigned)(8 * n8)), dim3(1024), lds, st, V, I, R, qa, qb, tout,
                     L ? *L : LzChain{}, L ? 1 : 0, s.d_seg, nseg, ntiles, n8);
  EDL_HIP(hipGetLastError());
}

template <typename T>
void launch_up_seg(const TwoPassView& V, const InterParams& I, const T* lut, T* qa, T* qb, T* tout,
                   const LzChain* L, const Sector& s, int C, int rank, hipStream_t st) {
  EDL_CHECK(s.d_seg && s.seg_h.size() >= 2 && s.seg_cap == seg_row_cap(sizeof(T), C),
            "launch_up_seg: segment bounds not built for this panel width");
  if (rank) {
    EDL_CHECK(seg_rank_ok(s), "launch_up_seg: no rank tables for this sector");
    switch (C) {
      case 1: return launch_up_segr_c<T, 1>(V, I, qa, qb, tout, L, s, st);
      case 2: return launch_up_segr_c<T, 2>(V, I, qa, qb, tout, L, s, st);
      case 3: return launch_up_segr_c<T, 3>(V, I, qa, qb, tout, L, s, st);
      case 4: return launch_up_segr_c<T, 4>(V, I, qa, qb, tout, L, s, st);
      case 8:
        if constexpr (sizeof(T) == 4) return launch_up_segr_c<T, 8>(V, I, qa, qb, tout, L, s, st);
        throw Error("launch_up_seg: 8-column panels are fp32 only");
      default: throw Error("launch_up_seg: unsupported panel width");
    }
  }
  switch (C) {
    case 1: return launch_up_seg_c<T, 1>(V, I, lut, qa, qb, tout, L, s, st);
    case 2: return launch_up_seg_c<T, 2>(V, I, lut, qa, qb, tout, L, s, st);
    case 3: return launch_up_seg_c<T, 3>(V, I, lut, qa, qb, tout, L, s, st);
    case 4: return launch_up_seg_c<T, 4>(V, I, lut, qa, qb, tout, L, s, st);
    default: throw Error("launch_up_seg: unsupported panel width");
  }
}

int phase_b_grid(const TwoPassView& V) { return std::min(V.dimdw, 4096); }

template <typename T>
void launch_phase_b(const TwoPassView& V, T* qa, T* qb, const LzChain& L, hipStream_t st) {
  hipLaunchKernelGGL(k_phase_b<T>, dim3(phase_b_grid(V)), dim3(256), 0, st, V, qa, qb, L);
  EDL_HIP(hipGetLastError());
}

void launch_lz_beta_n(const LzChain& L, int nparts, hipStream_t st) {
  hipLaunchKernelGGL(k_lz_beta_n, dim3(1), dim3(256), 0, st, L, nparts);
  EDL_HIP(hipGetLastError());
}

template void launch_up_seg<double>(const TwoPassView&, const InterParams&, const double*, double*, double*, double*,
                                    const LzChain*, const Sector&, int, int, hipStream_t);
template void launch_up_seg<float>(const TwoPassView&, const InterParams&, const float*, float*, float*, float*,
                                   const LzChain*, const Sector&, int, int, hipStream_t);
template <typename T>
void launch_deflate(const TwoPassView& V, T* qa, T* qb, const LzChain& L, const double* X, int nx, double* part,
                    double* coef, hipStream_t st) {
  const int g = phase_b_grid(V);
  hipLaunchKernelGGL(k_defl_dot<T>, dim3(g), dim3(256), 0, st, V, qa, qb, L, X, nx, part);
  EDL_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_defl_coef, dim3(1), dim3(256), 0, st, L, part, g, nx, coef);
  EDL_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_defl_update<T>, dim3(g), dim3(256), 0, st, V, qa, qb, L, X, nx, coef);
  EDL_HIP(hipGetLastError());
}

template void launch_deflate<double>(const TwoPassView&, double*, double*, const LzChain&, const double*, int,
                                     double*, double*, hipStream_t);
template void launch_deflate<float>(const TwoPassView&, float*, float*, const LzChain&, const double*, int,
                                    double*, double*, hipStream_t);
template void launch_phase_b<double>(const TwoPassView&, double*, double*, const LzChain&, hipStream_t);
template void launch_phase_b<float>(const TwoPassView&, float*, float*, const LzChain&, hipStream_t);

template void launch_pass_up<double>(const TwoPassView&, const InterParams&, const double*, double*, double*,
                                     double*, const LzChain*, const TwoPassCfg&, hipStream_t);
template void launch_pass_up<float>(const TwoPassView&, const InterParams&, const float*, float*, float*, float*,
                                    const LzChain*, const TwoPassCfg&, hipStream_t);
template void launch_pass_dw<double>(const TwoPassView&, const double*, double*, double*, const double*, double*,
                                     const LzChain*, const TwoPassCfg&, hipStream_t);
template void launch_pass_dw<float>(const TwoPassView&, const float*, float*, float*, const float*, float*,
                                    const LzChain*, const TwoPassCfg&, hipStream_t);

}  // namespace edl
