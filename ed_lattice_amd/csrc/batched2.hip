// Batched two-pass Lanczos steps for small sectors (the GF / chi seeds).
//
// Up tiles (chain, panel of C columns) run the same LDS column pass as the
// single-sector operator.  Down tiles (chain, slab of RS rows) stage the
// slab -- RS rows of every column, i.e. RS contiguous elements per column,
// one cache line when RS * sizeof(T) = 128 -- in LDS and gather every
// down-hop neighbour there, so both passes stream q, T and q_{k-1} through
// HBM once and keep all gathers on chip.  The down tile fuses Lanczos
// phase A, the up tile the previous step's phase B; their per-tile partial
// sums are added in tile order by one block per chain in a following
// launch (deterministic).
#include <algorithm>

#include "batched2.hpp"

namespace edl {

// Block b of a grid of 8*n8 blocks takes tile (b % 8) * n8 + b / 8: the
// blocks of one XCD take a contiguous range of the tile list, so the slabs
// of one chain that share cache lines run on the same XCD.
static __device__ __forceinline__ int64_t xcd_tile(int64_t n8) {
  return (int64_t)(blockIdx.x % 8) * n8 + blockIdx.x / 8;
}

// H16: the chains' up-spin hops two to a word (SpinTable::pk16; one hopping
// magnitude), ±h from a register, zero rows after the staged strings.
template <typename T, int C, int RPT, int UKCH, bool H16 = false>
__global__ __launch_bounds__(1024) void k_bt_up(const BtChain* __restrict__ chains, const int64_t* __restrict__ tile0,
                                                const int32_t* __restrict__ tchain, InterParams I, int64_t Tall,
                                                int64_t n8) {
  typedef typename Vec<T, C>::type VT;
  extern __shared__ __align__(16) unsigned char smem[];
  T* lut = reinterpret_cast<T*>(smem);
  VT* P = reinterpret_cast<VT*>(smem + lds_head_bytes(sizeof(T)));
  __shared__ double sh[16];
  const int64_t t = xcd_tile(n8);
  if (t >= Tall) return;
  const int c = tchain[t];
  const BtChain& B = chains[c];
  const LzScalars s = *B.sc;
  if (s.stopped) return;
  const TwoPassView V = B.V;
  const int pan = (int)(t - tile0[c]);
  const int c0 = pan * C;
  const T* lsrc = reinterpret_cast<const T*>(B.lut_up);
  // the coefficient table is requested before the panel and written to LDS
  // after it (its load completes under the panel's)
  const T lv = threadIdx.x < 256 ? lsrc[threadIdx.x] : (T)0;
  // Phase B of step k pending: r = t - alpha_k q_k / s_k, stored over t (the
  // buffer that becomes q_{k+1}); this panel of r is then the up-pass input.
  // Otherwise (first launch) the input is q_0.
  const bool pend = s.pad != 0;
  T* __restrict__ qn = reinterpret_cast<T*>(((pend ? s.k + 1 : s.k) & 1) ? B.q1 : B.q0);
  const T* __restrict__ qk = reinterpret_cast<const T*>((s.k & 1) ? B.q1 : B.q0);
  const T am = (T)(s.alpha_k / s.s_cur);
  const double nrm = stage_panel<T, C, stage_rows<T, C>(32, 2, 8)>(V, P, qn, qk, c0, pend, false, am, nullptr, 0.0);
  if (H16) {
    if (threadIdx.x < 32) {
      VT z;
#pragma unroll
      for (int j = 0; j < C; ++j) Vec<T, C>::set(z, j, (T)0);
      P[V.dimup + threadIdx.x] = z;
    }
  } else {
    if (threadIdx.x < 256) lut[threadIdx.x] = lv;
    for (int k = threadIdx.x + blockDim.x; k < 256; k += blockDim.x) lut[k] = lsrc[k];
  }
  __syncthreads();
  up_rows<T, C, RPT, UKCH, H16, false, H16>(V, I, H16 ? lsrc : lut, P, reinterpret_cast<T*>(B.tb), c0, (T)1);
  if (!pend) return;
  const double v = block_sum(nrm, sh);
  if (threadIdx.x == 0) B.partU[pan] = v;
}

// Deterministic sum of part[0..n) by one block (contiguous slice per thread,
// fixed tree); thread 0 holds the result.
static __device__ double ordered_sum(const double* __restrict__ part, int n, double* sh) {
  const int per = (n + blockDim.x - 1) / blockDim.x;
  const int a = threadIdx.x * per, e = min(a + per, n);
  double x = 0.0;
  for (int k = a; k < e; ++k) x += part[k];
  return block_sum(x, sh);
}

// One block per chain, after the up launch: beta_{k+1} = |r| from the
// panel partials, and the step's scalars advance.  The kernel boundary
// makes every panel's partial visible (no fence inside the hot kernels).
__global__ __launch_bounds__(256) void k_bt_beta(const BtChain* __restrict__ chains) {
  __shared__ double sh[16];
  const BtChain& B = chains[blockIdx.x];
  const LzScalars s = *B.sc;
  if (s.stopped || !s.pad) return;
  const double s2 = ordered_sum(B.partU, B.npanel, sh);
  if (threadIdx.x == 0) {
    const double b = sqrt(s2);
    LzScalars n = s;
    n.pad = 0;
    const bool more = (s.k + 1 < B.nmax);
    if (more) B.beta[s.k + 1] = b;
    const bool brk = !(b > B.brk_tol * (fabs(s.alpha_k) + s.beta_k));
    if (!more || brk) {
      n.stopped = 1;
    } else {
      n.s_old = s.s_cur;
      n.s_cur = b;
      n.beta_k = b;
      n.k = s.k + 1;
    }
    *B.sc = n;
  }
}

// After the beta launch, for models with the non-local Kanamori block: T +=
// N q_k of every chain (blockIdx.y = chain).
template <typename T>
__global__ __launch_bounds__(256) void k_bt_nonlocal(const BtChain* __restrict__ chains, InterParams I) {
  const BtChain& B = chains[blockIdx.y];
  const LzScalars s = *B.sc;
  if (s.stopped) return;
  const TwoPassView V = B.V;
  const T* q = reinterpret_cast<const T*>((s.k & 1) ? B.q1 : B.q0);
  T* tb = reinterpret_cast<T*>(B.tb);
  const int64_t n = (int64_t)V.dimup * V.dimdw;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int idw = (int)(e / V.dimup), iup = (int)(e % V.dimup);
    const size_t i = (size_t)idw * V.ld + iup;
    tb[i] += (T)nonlocal_gather<T>(I, (uint32_t)V.map_up[iup], (uint32_t)V.map_dw[idw], q, V.ld, I.Jx, I.Jp);
  }
}

// One block per chain, after the down launch: alpha_k from the slab
// partials; phase B of step k becomes pending.
__global__ __launch_bounds__(256) void k_bt_alpha(const BtChain* __restrict__ chains) {
  __shared__ double sh[16];
  const BtChain& B = chains[blockIdx.x];
  const LzScalars s = *B.sc;
  if (s.stopped) return;
  const double a = ordered_sum(B.partD, B.nslab, sh);
  if (threadIdx.x == 0) {
    B.alpha[s.k] = a;
    LzScalars n = s;
    n.alpha_k = a;
    n.nsteps = s.k + 1;
    n.pad = 1;
    if (s.k + 1 >= B.nmax) n.stopped = 1;
    *B.sc = n;
  }
}

// Down tile of one slab (rows r0 .. r0+RS-1 of every column): the slab of q
// is staged in S (the hop counts are already in cntS), then t = (T + H_dw q)
// / s_k - beta_k q_{k-1} / s_{k-1} over q_{k-1}; returns this thread's share
// of <t, q_k> (times s_k).  Both the tile kernel and the persistent kernel
// run it.
// Stage the slab of rows r0 .. r0+RS-1 of every column of q in S[dimdw][RS],
// SB elements per thread at a time with every load of a batch issued before
// its LDS writes (one memory round trip per batch).
template <typename T, int RS, int SB = 8>
static __device__ __forceinline__ void bt_stage_slab(const TwoPassView& V, T* S, const T* __restrict__ q, int r0) {
  const int n = V.dimdw * RS;
  const size_t ld = (size_t)V.ld;
  for (int base = 0; base < n; base += SB * (int)blockDim.x) {
    T a[SB];
#pragma unroll
    for (int h = 0; h < SB; ++h) {
      const int i = base + h * (int)blockDim.x + (int)threadIdx.x;
      const int j = i / RS, r = i % RS;
      a[h] = (i < n && r0 + r < V.dimup) ? q[(size_t)j * ld + r0 + r] : (T)0;
    }
#pragma unroll
    for (int h = 0; h < SB; ++h) {
      const int i = base + h * (int)blockDim.x + (int)threadIdx.x;
      if (i < n) S[i] = a[h];
    }
  }
}

// STAGE = false: the caller staged S (bt_stage_slab) and synchronised.
// H16: the down spin's hops two to a word (SpinTable::pk16), +-h from a
// register (lut = the global coefficient table), a missing second hop
// gathering the zero column dimdw of S.
template <typename T, int RS, int NC, bool STAGE = true, bool H16 = false>
static __device__ __forceinline__ double bt_dw_slab(const TwoPassView& V, const T* lut, T* S, const int* cntS,
                                                    const T* __restrict__ q, T* __restrict__ qo,
                                                    const T* __restrict__ tb, int r0, T ts, T bo) {
  const int dimup = V.dimup, dimdw = V.dimdw;
  const size_t ld = (size_t)V.ld;
  if (STAGE) {
    bt_stage_slab<T, RS>(V, S, q, r0);
    __syncthreads();
  }
  // lane = row of the slab (fastest) x column: the global accesses of a wave
  // are 64 / RS whole slab segments
  const int r = threadIdx.x % RS, jl = threadIdx.x / RS, jstep = blockDim.x / RS;
  const bool rok = r0 + r < dimup;
  double part = 0.0;
  constexpr int KCH = 8;
  // NC columns per thread; their T and q_{k-1} elements and NC*KCH hop
  // words are requested before the LDS gathers that need them (the hop
  // counts come from LDS).
  for (int j0 = jl; j0 < dimdw; j0 += NC * jstep) {
    int col[NC], nh[NC];
    bool okc[NC];
    size_t ii[NC];
    T tin[NC], qold[NC], acc[NC];
    int nmax = 0;
#pragma unroll
    for (int h = 0; h < NC; ++h) {
      col[h] = j0 + h * jstep;
      okc[h] = col[h] < dimdw;
      if (!okc[h]) col[h] = j0;
      nh[h] = okc[h] ? (H16 ? (cntS[col[h]] + 1) >> 1 : cntS[col[h]]) : 0;
      nmax = nh[h] > nmax ? nh[h] : nmax;
      ii[h] = (size_t)col[h] * ld + r0 + r;
      tin[h] = (rok && okc[h]) ? tb[ii[h]] : (T)0;
      qold[h] = (rok && okc[h]) ? qo[ii[h]] : (T)0;
      acc[h] = (T)0;
    }
    if constexpr (H16) {
      constexpr int KW = KCH / 2;  // words (two hops each) per chunk
      const T l0 = lut[0];
      const uint32_t zpad = (uint32_t)dimdw | ((uint32_t)dimdw << 16);
      for (int k0 = 0; k0 < nmax; k0 += KW) {
        uint32_t w[NC][KW];
#pragma unroll
        for (int h = 0; h < NC; ++h)
#pragma unroll
          for (int jj = 0; jj < KW; ++jj)
            w[h][jj] = (k0 + jj < nh[h]) ? V.pk16_dw[(size_t)(k0 + jj) * dimdw + col[h]] : zpad;
#pragma unroll
        for (int h = 0; h < NC; ++h)
#pragma unroll
          for (int jj = 0; jj < KW; ++jj) {
            const uint32_t wl = w[h][jj];
            acc[h] += flip_sign(l0, (wl << 16) & 0x80000000u) * S[(int)(wl & 0x7FFFu) * RS + r];
            acc[h] += flip_sign(l0, wl & 0x80000000u) * S[(int)((wl >> 16) & 0x7FFFu) * RS + r];
          }
      }
    } else {
      for (int k0 = 0; k0 < nmax; k0 += KCH) {
        uint32_t w[NC][KCH];
#pragma unroll
        for (int h = 0; h < NC; ++h)
#pragma unroll
          for (int jj = 0; jj < KCH; ++jj)
            w[h][jj] = (k0 + jj < nh[h]) ? V.pk_dw[(size_t)(k0 + jj) * dimdw + col[h]] : (255u << 24);
#pragma unroll
        for (int h = 0; h < NC; ++h)
#pragma unroll
          for (int jj = 0; jj < KCH; ++jj) acc[h] += lut[w[h][jj] >> 24] * S[(int)(w[h][jj] & 0xFFFFFFu) * RS + r];
      }
    }
    if (rok) {
#pragma unroll
      for (int h = 0; h < NC; ++h) {
        if (!okc[h]) continue;
        const T tv = (tin[h] + acc[h]) * ts - bo * qold[h];
        qo[ii[h]] = tv;
        part += (double)tv * (double)S[col[h] * RS + r];
      }
    }
  }
  return part;
}

template <typename T, int RS, int NC, bool H16 = false>
__global__ __launch_bounds__(1024) void k_bt_dw(const BtChain* __restrict__ chains, const int64_t* __restrict__ tile0,
                                                const int32_t* __restrict__ tchain, int64_t Tall, int64_t n8) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* lut = reinterpret_cast<T*>(smem);
  T* S = reinterpret_cast<T*>(smem + lds_head_bytes(sizeof(T)));  // [dimdw][RS]
  __shared__ double sh[16];
  const int64_t t = xcd_tile(n8);
  if (t >= Tall) return;
  const int c = tchain[t];
  const BtChain& B = chains[c];
  const LzScalars s = *B.sc;
  if (s.stopped) return;
  const TwoPassView V = B.V;
  const T* __restrict__ q = reinterpret_cast<const T*>((s.k & 1) ? B.q1 : B.q0);
  T* __restrict__ qo = reinterpret_cast<T*>((s.k & 1) ? B.q0 : B.q1);
  const T* __restrict__ tb = reinterpret_cast<const T*>(B.tb);
  const T* lsrc = reinterpret_cast<const T*>(B.lut_dw);
  const int sl = (int)(t - tile0[c]);
  // [dimdw] hop counts, after the slab (and, with H16, after its zero column)
  int* cntS = reinterpret_cast<int*>(S + (size_t)(V.dimdw + (H16 ? 1 : 0)) * RS);
  // the coefficient table, the hop counts and the slab are requested
  // together and written to LDS after (one memory round trip, not three)
  const T lv = (!H16 && threadIdx.x < 256) ? lsrc[threadIdx.x] : (T)0;
  const int cv = (int)threadIdx.x < V.dimdw ? V.cnt_dw[threadIdx.x] : 0;
  bt_stage_slab<T, RS>(V, S, q, sl * RS);
  if (H16) {
    if (threadIdx.x < RS) S[(size_t)V.dimdw * RS + threadIdx.x] = (T)0;
  } else {
    if (threadIdx.x < 256) lut[threadIdx.x] = lv;
    for (int k = threadIdx.x + blockDim.x; k < 256; k += blockDim.x) lut[k] = lsrc[k];
  }
  if ((int)threadIdx.x < V.dimdw) cntS[threadIdx.x] = cv;
  for (int j = threadIdx.x + blockDim.x; j < V.dimdw; j += blockDim.x) cntS[j] = V.cnt_dw[j];
  __syncthreads();
  const double inv_s = 1.0 / s.s_cur;
  const T ts = (T)inv_s, bo = (T)(s.beta_k / s.s_old);
  const double part = bt_dw_slab<T, RS, NC, false, H16>(V, H16 ? lsrc : lut, S, cntS, q, qo, tb, sl * RS, ts, bo);
  const double v = block_sum(part * inv_s, sh);
  if (threadIdx.x == 0) B.partD[sl] = v;
}

// Every step of one chain in one block (a persistent launch: one block per
// chain, all Lanczos steps before it exits).  The step is the four launches
// of the tile path done by one block in turn -- the up pass over every
// panel (phase B fused), beta, the down pass over every slab (phase A
// fused), alpha -- with the reductions inside the block, so no launch or
// grid-wide synchronisation separates the phases: __syncthreads() orders the
// block's own global writes (r, T, t) before the phase that reads them.
template <typename T, int C, int RPT, int UKCH, int RS, int NC>
__global__ __launch_bounds__(1024) void k_bt_persist(const BtChain* __restrict__ chains, InterParams I) {
  typedef typename Vec<T, C>::type VT;
  extern __shared__ __align__(16) unsigned char smem[];
  T* lut_u = reinterpret_cast<T*>(smem);
  T* lut_d = lut_u + 256;
  unsigned char* body = smem + 2 * lds_head_bytes(sizeof(T));
  VT* P = reinterpret_cast<VT*>(body);  // up phase: [dimup] panel rows
  T* S = reinterpret_cast<T*>(body);    // down phase: [dimdw][RS] slab
  __shared__ double sh[16];
  __shared__ double bc;
  const BtChain& B = chains[blockIdx.x];
  LzScalars s = *B.sc;
  if (s.stopped) return;
  const TwoPassView V = B.V;
  const int dimdw = V.dimdw;
  int* cntS = reinterpret_cast<int*>(body + sizeof(T) * (size_t)dimdw * RS);  // [dimdw], after the slab
  const T* lu = reinterpret_cast<const T*>(B.lut_up);
  const T* ldw = reinterpret_cast<const T*>(B.lut_dw);
  for (int k = threadIdx.x; k < 256; k += blockDim.x) {
    lut_u[k] = lu[k];
    lut_d[k] = ldw[k];
  }
  // (no __restrict__ at this scope: the buffers change roles between phases)
  T* tb = reinterpret_cast<T*>(B.tb);
  const int npan = (dimdw + C - 1) / C;
  for (;;) {
    // ---- up: phase B of the previous step where pending, T = [D + H_up] q
    const bool pend = s.pad != 0;
    T* qn = reinterpret_cast<T*>(((pend ? s.k + 1 : s.k) & 1) ? B.q1 : B.q0);
    const T* qk = reinterpret_cast<const T*>((s.k & 1) ? B.q1 : B.q0);
    const T am = (T)(s.alpha_k / s.s_cur);
    double nrm = 0.0;
    for (int pan = 0; pan < npan; ++pan) {
      __syncthreads();
      nrm += stage_panel<T, C, stage_rows<T, C>(32, 2, 8)>(V, P, qn, qk, pan * C, pend, false, am, nullptr, 0.0);
      __syncthreads();
      up_rows<T, C, RPT, UKCH>(V, I, lut_u, P, tb, pan * C, (T)1);
    }
    if (pend) {
      // ---- beta_{k+1} = |r|, the scalars advance (k_bt_beta)
      const double s2 = block_sum(nrm, sh);
      if (threadIdx.x == 0) bc = s2;
      __syncthreads();
      const double b = sqrt(bc);
      const bool more = (s.k + 1 < B.nmax);
      if (more && threadIdx.x == 0) B.beta[s.k + 1] = b;
      s.pad = 0;
      const bool brk = !(b > B.brk_tol * (fabs(s.alpha_k) + s.beta_k));
      if (!more || brk) {
        s.stopped = 1;
        break;
      }
      s.s_old = s.s_cur;
      s.s_cur = b;
      s.beta_k = b;
      s.k = s.k + 1;
    }
    // ---- down: t = (T + H_dw q)/s_k - beta_k q_{k-1}/s_{k-1}, alpha (k_bt_dw, k_bt_alpha)
    __syncthreads();  // r and T of every panel written before any slab reads them
    if (I.Jx != 0.0 || I.Jp != 0.0) {
      // T += N q_k, the non-local Kanamori block (k_bt_nonlocal)
      const T* qc = reinterpret_cast<const T*>((s.k & 1) ? B.q1 : B.q0);
      const int64_t n = (int64_t)V.dimup * dimdw;
      for (int64_t e = threadIdx.x; e < n; e += blockDim.x) {
        const int idw = (int)(e / V.dimup), iup = (int)(e % V.dimup);
        const size_t i = (size_t)idw * V.ld + iup;
        tb[i] += (T)nonlocal_gather<T>(I, (uint32_t)V.map_up[iup], (uint32_t)V.map_dw[idw], qc, V.ld, I.Jx, I.Jp);
      }
      __syncthreads();
    }
    for (int j = threadIdx.x; j < dimdw; j += blockDim.x) cntS[j] = V.cnt_dw[j];
    const T* q = reinterpret_cast<const T*>((s.k & 1) ? B.q1 : B.q0);
    T* qo = reinterpret_cast<T*>((s.k & 1) ? B.q0 : B.q1);
    const double inv_s = 1.0 / s.s_cur;
    const T ts = (T)inv_s, bo = (T)(s.beta_k / s.s_old);
    double part = 0.0;
    for (int sl = 0; sl < B.nslab; ++sl) {
      __syncthreads();  // the previous slab's gathers are done with S
      part += bt_dw_slab<T, RS, NC>(V, lut_d, S, cntS, q, qo, tb, sl * RS, ts, bo);
    }
    const double a0 = block_sum(part * inv_s, sh);
    if (threadIdx.x == 0) bc = a0;
    __syncthreads();
    const double a = bc;
    if (threadIdx.x == 0) B.alpha[s.k] = a;
    s.alpha_k = a;
    s.nsteps = s.k + 1;
    s.pad = 1;
    if (s.k + 1 >= B.nmax) {
      s.stopped = 1;
      break;
    }
  }
  if (threadIdx.x == 0) *B.sc = s;
}

// ---------------------------------------------------------------- host side
static constexpr size_t kBtUpBudget = 64 * 1024;   // two or more up tiles per CU
static constexpr size_t kBtDwBudget = 150 * 1024;  // one slab per CU at most

size_t bt_dw_lds_bytes(int dimdw, int RS, size_t elem, bool h16) {
  return lds_head_bytes(elem) + (size_t)RS * (dimdw + (h16 ? 1 : 0)) * elem + sizeof(int) * (size_t)dimdw;
}

BtShape bt_shape(int max_dimup, int max_dimdw, size_t elem) {
  BtShape b;
  b.RS = 0;
  const size_t head = lds_head_bytes(elem);
  int C = 0;
  for (int c : {4, 2, 1})
    if (head + (size_t)c * max_dimup * elem <= kBtUpBudget) {
      C = c;
      break;
    }
  if (C == 0 && head + (size_t)max_dimup * elem <= kBtDwBudget) C = 1;
  if (C == 0) return b;
  int RS = 0;
  for (int rs : {32, 16, 8, 4, 2, 1})
    if ((size_t)rs * elem <= 128 && bt_dw_lds_bytes(max_dimdw, rs, elem) <= kBtDwBudget) {
      RS = rs;
      break;
    }
  if (RS == 0) return b;
  b.C = C;
  b.RS = RS;
  const size_t up_lds = head + (size_t)C * max_dimup * elem;
  b.threads_up = up_lds > 40 * 1024 ? 1024 : 256;
  const size_t dw_lds = bt_dw_lds_bytes(max_dimdw, RS, elem);
  b.threads_dw = dw_lds > 40 * 1024 ? 1024 : 256;
  return b;
}

BtSet::~BtSet() {
  if (d) (void)hipFree(d);
  if (d_tile_up) (void)hipFree(d_tile_up);
  if (d_tile_dw) (void)hipFree(d_tile_dw);
  if (d_chain_up) (void)hipFree(d_chain_up);
  if (d_chain_dw) (void)hipFree(d_chain_dw);
}

// LDS limits of every instantiation, set once per device and outside any
// graph capture.
template <typename T>
static void bt_attributes() {
  for (auto f : {(const void*)k_bt_up<T, 3, 2, 8>, (const void*)k_bt_up<T, 3, 2, 4>, (const void*)k_bt_up<T, 3, 1, 8>,
                 (const void*)k_bt_up<T, 1, 2, 8>, (const void*)k_bt_up<T, 2, 2, 8>, (const void*)k_bt_up<T, 4, 2, 8>,
                 (const void*)k_bt_up<T, 1, 2, 4>, (const void*)k_bt_up<T, 2, 2, 4>, (const void*)k_bt_up<T, 4, 2, 4>,
                 (const void*)k_bt_up<T, 1, 1, 8>, (const void*)k_bt_up<T, 2, 1, 8>, (const void*)k_bt_up<T, 4, 1, 8>,
                 (const void*)k_bt_up<T, 3, 2, 8, true>, (const void*)k_bt_up<T, 3, 2, 4, true>,
                 (const void*)k_bt_up<T, 3, 1, 8, true>, (const void*)k_bt_up<T, 1, 2, 8, true>,
                 (const void*)k_bt_up<T, 2, 2, 8, true>, (const void*)k_bt_up<T, 4, 2, 8, true>,
                 (const void*)k_bt_up<T, 1, 2, 4, true>, (const void*)k_bt_up<T, 2, 2, 4, true>,
                 (const void*)k_bt_up<T, 4, 2, 4, true>, (const void*)k_bt_up<T, 1, 1, 8, true>,
                 (const void*)k_bt_up<T, 2, 1, 8, true>, (const void*)k_bt_up<T, 4, 1, 8, true>,
                 (const void*)k_bt_dw<T, 1, 1>, (const void*)k_bt_dw<T, 2, 1>, (const void*)k_bt_dw<T, 4, 1>,
                 (const void*)k_bt_dw<T, 8, 1>, (const void*)k_bt_dw<T, 16, 1>, (const void*)k_bt_dw<T, 32, 1>,
                 (const void*)k_bt_dw<T, 1, 2>, (const void*)k_bt_dw<T, 2, 2>, (const void*)k_bt_dw<T, 4, 2>,
                 (const void*)k_bt_dw<T, 8, 2>, (const void*)k_bt_dw<T, 16, 2>, (const void*)k_bt_dw<T, 32, 2>,
                 (const void*)k_bt_dw<T, 1, 1, true>, (const void*)k_bt_dw<T, 2, 1, true>,
                 (const void*)k_bt_dw<T, 4, 1, true>, (const void*)k_bt_dw<T, 8, 1, true>,
                 (const void*)k_bt_dw<T, 16, 1, true>, (const void*)k_bt_dw<T, 32, 1, true>,
                 (const void*)k_bt_dw<T, 1, 2, true>, (const void*)k_bt_dw<T, 2, 2, true>,
                 (const void*)k_bt_dw<T, 4, 2, true>, (const void*)k_bt_dw<T, 8, 2, true>,
                 (const void*)k_bt_dw<T, 16, 2, true>, (const void*)k_bt_dw<T, 32, 2, true>})
    set_lds_limit(f, (int)kBtDwBudget);
}

template <typename T>
void bt_finalize(BtSet& bs) {
  bt_attributes<T>();
  const size_t n = bs.h.size();
  std::vector<int64_t> tu(n + 1, 0), td(n + 1, 0);
  for (size_t k = 0; k < n; ++k) {
    const BtChain& b = bs.h[k];
    EDL_CHECK(b.nslab == (b.V.dimup + bs.shape.RS - 1) / bs.shape.RS, "bt_finalize: slab count mismatch");
    tu[k + 1] = tu[k] + (b.V.dimdw + bs.shape.C - 1) / bs.shape.C;
    td[k + 1] = td[k] + b.nslab;
  }
  bs.tiles_up = tu[n];
  bs.tiles_dw = td[n];
  EDL_HIP(hipMalloc(&bs.d, sizeof(BtChain) * n));
  EDL_HIP(hipMalloc(&bs.d_tile_up, sizeof(int64_t) * (n + 1)));
  EDL_HIP(hipMalloc(&bs.d_tile_dw, sizeof(int64_t) * (n + 1)));
  EDL_HIP(hipMemcpy(bs.d, bs.h.data(), sizeof(BtChain) * n, hipMemcpyHostToDevice));
  EDL_HIP(hipMemcpy(bs.d_tile_up, tu.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice));
  EDL_HIP(hipMemcpy(bs.d_tile_dw, td.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice));
  std::vector<int32_t> cu((size_t)bs.tiles_up), cdw((size_t)bs.tiles_dw);
  for (size_t k = 0; k < n; ++k) {
    for (int64_t t = tu[k]; t < tu[k + 1]; ++t) cu[(size_t)t] = (int32_t)k;
    for (int64_t t = td[k]; t < td[k + 1]; ++t) cdw[(size_t)t] = (int32_t)k;
  }
  EDL_HIP(hipMalloc(&bs.d_chain_up, sizeof(int32_t) * std::max<size_t>(cu.size(), 1)));
  EDL_HIP(hipMalloc(&bs.d_chain_dw, sizeof(int32_t) * std::max<size_t>(cdw.size(), 1)));
  EDL_HIP(hipMemcpy(bs.d_chain_up, cu.data(), sizeof(int32_t) * cu.size(), hipMemcpyHostToDevice));
  EDL_HIP(hipMemcpy(bs.d_chain_dw, cdw.data(), sizeof(int32_t) * cdw.size(), hipMemcpyHostToDevice));
}

bool bt_h16_ok(const std::vector<BtChain>& chains, int spin) {
  for (const BtChain& b : chains)
    if (!(spin == 0 ? b.V.pk16_up : b.V.pk16_dw)) return false;
  return !chains.empty();
}

static int max_dim(const BtSet& bs, int which) {
  int m = 1;
  for (const BtChain& b : bs.h) m = std::max(m, which == 0 ? b.V.dimup : b.V.dimdw);
  return m;
}

template <typename T, int C, int RPT, int UKCH>
static void bt_up_k(const BtSet& bs, const InterParams& I, hipStream_t st, bool scalars) {
  const bool h16 = bs.shape.up_h16 != 0;
  const size_t lds = lds_head_bytes(sizeof(T)) + (size_t)C * (max_dim(bs, 0) + (h16 ? 32 : 0)) * sizeof(T);
  EDL_CHECK(lds <= kBtDwBudget, "bt_launch_up: panel exceeds LDS");
  EDL_CHECK(!h16 || bt_h16_ok(bs.h, 0), "bt_launch_up: a chain's sector has no two-hop table");
  const int64_t n8 = (bs.tiles_up + 7) / 8;
  if (h16)
    hipLaunchKernelGGL((k_bt_up<T, C, RPT, UKCH, true>), dim3((unsigned)(8 * n8)), dim3(bs.shape.threads_up), lds, st,
                       bs.d, bs.d_tile_up, bs.d_chain_up, I, bs.tiles_up, n8);
  else
    hipLaunchKernelGGL((k_bt_up<T, C, RPT, UKCH>), dim3((unsigned)(8 * n8)), dim3(bs.shape.threads_up), lds, st,
                       bs.d, bs.d_tile_up, bs.d_chain_up, I, bs.tiles_up, n8);
  EDL_HIP(hipGetLastError());
  if (!scalars) return;
  hipLaunchKernelGGL(k_bt_beta, dim3((unsigned)bs.h.size()), dim3(256), 0, st, bs.d);
  EDL_HIP(hipGetLastError());
  if (I.Jx != 0.0 || I.Jp != 0.0) {
    int64_t n = 1;
    for (const BtChain& b : bs.h) n = std::max<int64_t>(n, (int64_t)b.V.dimup * b.V.dimdw);
    const unsigned gx = (unsigned)std::min<int64_t>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(k_bt_nonlocal<T>, dim3(gx, (unsigned)bs.h.size()), dim3(256), 0, st, bs.d, I);
    EDL_HIP(hipGetLastError());
  }
}

template <typename T, int RS, int NC>
static void bt_dw_r(const BtSet& bs, hipStream_t st, bool scalars) {
  const bool h16 = bs.shape.dw_h16 != 0;
  const size_t lds = bt_dw_lds_bytes(max_dim(bs, 1), RS, sizeof(T), h16);
  EDL_CHECK(lds <= kBtDwBudget, "bt_launch_dw: slab exceeds LDS");
  EDL_CHECK(bs.shape.threads_dw % RS == 0, "bt_launch_dw: block size not a multiple of the slab height");
  EDL_CHECK(!h16 || bt_h16_ok(bs.h, 1), "bt_launch_dw: a chain's sector has no two-hop table");
  const int64_t n8 = (bs.tiles_dw + 7) / 8;
  if (h16)
    hipLaunchKernelGGL((k_bt_dw<T, RS, NC, true>), dim3((unsigned)(8 * n8)), dim3(bs.shape.threads_dw), lds, st,
                       bs.d, bs.d_tile_dw, bs.d_chain_dw, bs.tiles_dw, n8);
  else
    hipLaunchKernelGGL((k_bt_dw<T, RS, NC>), dim3((unsigned)(8 * n8)), dim3(bs.shape.threads_dw), lds, st, bs.d,
                       bs.d_tile_dw, bs.d_chain_dw, bs.tiles_dw, n8);
  EDL_HIP(hipGetLastError());
  if (!scalars) return;
  hipLaunchKernelGGL(k_bt_alpha, dim3((unsigned)bs.h.size()), dim3(256), 0, st, bs.d);
  EDL_HIP(hipGetLastError());
}

template <typename T, int C>
static void bt_up_c(const BtSet& bs, const InterParams& I, hipStream_t st, bool scalars) {
  switch (bs.shape.up_rpt * 100 + bs.shape.up_kch) {
    case 208: return bt_up_k<T, C, 2, 8>(bs, I, st, scalars);
    case 204: return bt_up_k<T, C, 2, 4>(bs, I, st, scalars);
    case 108: return bt_up_k<T, C, 1, 8>(bs, I, st, scalars);
    default: throw Error("bt_launch_up: unsupported rows per thread / hop chunk");
  }
}

template <typename T>
void bt_launch_up(const BtSet& bs, const InterParams& I, hipStream_t st, bool scalars) {
  switch (bs.shape.C) {
    case 1: return bt_up_c<T, 1>(bs, I, st, scalars);
    case 2: return bt_up_c<T, 2>(bs, I, st, scalars);
    case 3: return bt_up_c<T, 3>(bs, I, st, scalars);
    case 4: return bt_up_c<T, 4>(bs, I, st, scalars);
    default: throw Error("bt_launch_up: unsupported panel width");
  }
}

template <typename T, int NC>
static void bt_dw_n(const BtSet& bs, hipStream_t st, bool scalars) {
  switch (bs.shape.RS) {
    case 1: return bt_dw_r<T, 1, NC>(bs, st, scalars);
    case 2: return bt_dw_r<T, 2, NC>(bs, st, scalars);
    case 4: return bt_dw_r<T, 4, NC>(bs, st, scalars);
    case 8: return bt_dw_r<T, 8, NC>(bs, st, scalars);
    case 16: return bt_dw_r<T, 16, NC>(bs, st, scalars);
    case 32: return bt_dw_r<T, 32, NC>(bs, st, scalars);
    default: throw Error("bt_launch_dw: unsupported slab height");
  }
}

template <typename T>
void bt_launch_dw(const BtSet& bs, hipStream_t st, bool scalars) {
  if (bs.shape.dw_nc == 1)
    bt_dw_n<T, 1>(bs, st, scalars);
  else if (bs.shape.dw_nc == 2)
    bt_dw_n<T, 2>(bs, st, scalars);
  else
    throw Error("bt_launch_dw: unsupported columns per thread");
}

// ---- persistent launches
size_t bt_persist_lds_bytes(int max_dimup, int max_dimdw, int C, int RS, size_t elem) {
  const size_t up = (size_t)C * max_dimup * elem;
  const size_t dw = (size_t)RS * max_dimdw * elem + sizeof(int) * (size_t)max_dimdw;
  return 2 * lds_head_bytes(elem) + std::max(up, dw);
}

bool bt_persist_shape_ok(int C, int RS, size_t elem) {
  return (C == 2 || C == 4) && (RS == 8 || RS == 16 || (RS == 32 && elem == 4));
}

template <typename T, int C, int RS>
static void bt_persist_k(const BtSet& bs, const InterParams& I, hipStream_t st) {
  const size_t lds = bt_persist_lds_bytes(max_dim(bs, 0), max_dim(bs, 1), C, RS, sizeof(T));
  EDL_CHECK(lds <= kBtDwBudget, "bt_launch_persist: panel or slab exceeds LDS");
  set_lds_limit((const void*)k_bt_persist<T, C, 2, 8, RS, 2>, (int)kBtDwBudget);
  hipLaunchKernelGGL((k_bt_persist<T, C, 2, 8, RS, 2>), dim3((unsigned)bs.h.size()), dim3(1024), lds, st, bs.d, I);
  EDL_HIP(hipGetLastError());
}

template <typename T, int C>
static void bt_persist_c(const BtSet& bs, const InterParams& I, hipStream_t st) {
  switch (bs.shape.RS) {
    case 8: return bt_persist_k<T, C, 8>(bs, I, st);
    case 16: return bt_persist_k<T, C, 16>(bs, I, st);
    case 32:
      if constexpr (sizeof(T) == 4) return bt_persist_k<T, C, 32>(bs, I, st);
      [[fallthrough]];
    default: throw Error("bt_launch_persist: unsupported slab height");
  }
}

template <typename T>
void bt_launch_persist(const BtSet& bs, const InterParams& I, hipStream_t st) {
  EDL_CHECK(bt_persist_shape_ok(bs.shape.C, bs.shape.RS, sizeof(T)), "bt_launch_persist: unsupported shape");
  for (const BtChain& b : bs.h)
    EDL_CHECK(b.nslab == (b.V.dimup + bs.shape.RS - 1) / bs.shape.RS, "bt_launch_persist: slab count mismatch");
  if (bs.shape.C == 2)
    bt_persist_c<T, 2>(bs, I, st);
  else
    bt_persist_c<T, 4>(bs, I, st);
}

template void bt_launch_persist<double>(const BtSet&, const InterParams&, hipStream_t);
template void bt_launch_persist<float>(const BtSet&, const InterParams&, hipStream_t);
template void bt_finalize<double>(BtSet&);
template void bt_finalize<float>(BtSet&);
template void bt_launch_up<double>(const BtSet&, const InterParams&, hipStream_t, bool);
template void bt_launch_up<float>(const BtSet&, const InterParams&, hipStream_t, bool);
template void bt_launch_dw<double>(const BtSet&, hipStream_t, bool);
template void bt_launch_dw<float>(const BtSet&, hipStream_t, bool);

}  // namespace edl
