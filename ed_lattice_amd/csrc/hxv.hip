// Single-pass on-the-fly H x v over a list of Lanczos chains (the batched
// GF/chi seeds, and ground states whose columns do not fit in LDS), and the
// two phases of a Lanczos step.
//
// A vector of sector (Nup, Ndw) is the column-major matrix V(iup, idw),
// element i = iup + idw*DimUp (SECTOR.f90:443-480, 0-based here).  In
// gather form every output element is
//
//   (H v)(iup,idw) = [dg_up(iup) + dg_dw(idw) + Uint(s_up & s_dw)] v(iup,idw)
//                  + sum_k cu_k(iup) v(nu_k(iup), idw)      (HxV_up.f90)
//                  + sum_k cd_k(idw) v(iup, nd_k(idw))      (HxV_dw.f90)
//
// with one writer per element, so no atomics on data and no write races.
//
// Work unit: a tile = one column of one chain.  The tiles of all chains
// form one list.  With a grid of G = 8*nb blocks, XCD x (the hardware sends
// block b to XCD b % 8) owns the list slice [x*T/8, (x+1)*T/8) and its nb
// blocks take tiles x0 + j, x0 + j + nb, ...: the blocks of one XCD sweep
// the same chain together, so the chain's vector (a few MB for the seed
// sectors) stays in that XCD's L2 and the down-hop neighbour columns are
// L2 hits.
#include "chains.hpp"

namespace edl {

static __device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Deterministic block sum (fixed shuffle tree, fixed wave order), thread 0.
static __device__ double block_sum(double v, double* sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < nw; ++w) r += sh[w];
  return r;
}

static __device__ __forceinline__ double inter_term(const InterParams& I, uint32_t su, uint32_t sd) {
  const uint32_t both = su & sd;
  double e = 0.0;
  for (int k = 0; k < I.norb; ++k) e += I.U[k] * (double)__popc(both & I.orb_mask[k]);
  for (int p = 0; p < I.npairs; ++p) {
    const uint32_t a = I.pi[p], b = I.pj[p];
    e += I.Ust * (double)((((su >> a) & (sd >> b)) & 1u) + (((su >> b) & (sd >> a)) & 1u));
  }
  return e;
}

// Fermionic sign of c_pos / cdg_pos on string s (AUX_FUNX.f90:76-102): the
// parity of the occupied sites below pos.
static __device__ __forceinline__ double fsign(uint32_t s, uint32_t pos) {
  return (__popc(s & ((1u << pos) - 1u)) & 1) ? -1.0 : 1.0;
}

// Row (su, sd) of the non-local Kanamori block (HxV_non_local.f90:14-88) in
// gather form.  The reference scatters, from a source string pair, to the
// pair reached by c+_io c_jo (up) c+_jo c_io (down) with Jx (spin exchange)
// or c+_io c_jo (up) c+_io c_jo (down) with Jp (pair hopping), for every
// ordered same-site orbital pair; the block is symmetric, so row i gathers
// with the coefficient of the move that starts at i.
// cx, cp: the coefficients of the exchange and pair-hopping moves (Jx, Jp
// for the operator; 1, 0 and 0, 1 for the expectation values Dse, Dph).
template <typename T>
static __device__ __forceinline__ double nonlocal_gather(const InterParams& I, uint32_t su, uint32_t sd,
                                                         const T* __restrict__ q, size_t ld, double cx, double cp) {
  double acc = 0.0;
  for (int p = 0; p < I.npairs; ++p)
    for (int o = 0; o < 2; ++o) {
      const uint32_t io = o ? I.pj[p] : I.pi[p], jo = o ? I.pi[p] : I.pj[p];
      const uint32_t bio = 1u << io, bjo = 1u << jo;
      const bool up_j = su & bjo, up_i = su & bio, dw_j = sd & bjo, dw_i = sd & bio;
      if (cx != 0.0 && up_j && !up_i && dw_i && !dw_j) {
        const uint32_t k1 = sd ^ bio;                    // c_io (down)
        const double s12 = fsign(sd, io) * fsign(k1, jo);  // then cdg_jo
        const uint32_t k2 = k1 | bjo;
        const uint32_t k3 = su ^ bjo;                    // c_jo (up)
        const double s34 = fsign(su, jo) * fsign(k3, io);  // then cdg_io
        const uint32_t k4 = k3 | bio;
        acc += cx * s12 * s34 * (double)q[(size_t)rank_of(k2) * ld + rank_of(k4)];
      }
      if (cp != 0.0 && up_j && dw_j && !up_i && !dw_i) {
        const uint32_t k1 = sd ^ bjo;                    // c_jo (down)
        const double s12 = fsign(sd, jo) * fsign(k1, io);  // then cdg_io
        const uint32_t k2 = k1 | bio;
        const uint32_t k3 = su ^ bjo;                    // c_jo (up)
        const double s34 = fsign(su, jo) * fsign(k3, io);  // then cdg_io
        const uint32_t k4 = k3 | bio;
        acc += cp * s12 * s34 * (double)q[(size_t)rank_of(k2) * ld + rank_of(k4)];
      }
    }
  return acc;
}

template <typename T>
static __device__ __forceinline__ T nonlocal_row(const InterParams& I, uint32_t su, uint32_t sd,
                                                 const T* __restrict__ q, size_t ld) {
  return (T)nonlocal_gather<T>(I, su, sd, q, ld, I.Jx, I.Jp);
}

// Tile walk of one block: XCD-synchronous strided order (see header).
struct TileWalk {
  int64_t t, end, step;
  __device__ TileWalk(int64_t T) {
    const int G = gridDim.x, b = blockIdx.x;
    if (G >= 8 && G % 8 == 0) {
      const int x = b % 8, j = b / 8;
      const int64_t x0 = (int64_t)x * T / 8, x1 = (int64_t)(x + 1) * T / 8;
      t = x0 + j;
      end = x1;
      step = G / 8;
    } else {
      t = b;
      end = T;
      step = G;
    }
  }
};

static __device__ __forceinline__ int chain_of(const int64_t* __restrict__ tile0, int nchains, int64_t t) {
  int lo = 0, hi = nchains - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tile0[mid] <= t) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// (H q)(:, idw) / s for the rows this thread owns in the column, written
// through `emit(i, h)`.  Up hops gather inside the column (L1/L2 resident
// while the block works on it), down hops read whole neighbour columns.
template <typename T, typename Emit>
static __device__ __forceinline__ void column_hq(const ChainDev& C, const InterParams& I, const T* __restrict__ q,
                                                 int idw, Emit emit) {
  const T* __restrict__ cu = reinterpret_cast<const T*>(C.cu);
  const T* __restrict__ cd = reinterpret_cast<const T*>(C.cd);
  const size_t col = (size_t)idw * C.ld;
  const uint32_t sd = (uint32_t)C.map_dw[idw];
  const double dgd = C.dg_dw[idw];
  const int ndw = C.cnt_dw[idw];
  constexpr int KCH = 8;
  for (int iup = threadIdx.x; iup < C.dimup; iup += blockDim.x) {
    const uint32_t su = (uint32_t)C.map_up[iup];
    T acc = (T)(C.dg_up[iup] + dgd + inter_term(I, su, sd)) * q[col + iup];
    const int nup = C.cnt_up[iup];
    for (int k0 = 0; k0 < nup; k0 += KCH) {
      T x[KCH], c[KCH];
#pragma unroll
      for (int j = 0; j < KCH; ++j) {
        const bool on = k0 + j < nup;
        const size_t o = (size_t)(k0 + j) * C.dimup + iup;
        c[j] = on ? cu[o] : (T)0;
        x[j] = on ? q[col + (size_t)C.nbr_up[o]] : (T)0;
      }
#pragma unroll
      for (int j = 0; j < KCH; ++j) acc += c[j] * x[j];
    }
    for (int k0 = 0; k0 < ndw; k0 += KCH) {
      T x[KCH], c[KCH];
#pragma unroll
      for (int j = 0; j < KCH; ++j) {
        const bool on = k0 + j < ndw;
        const size_t o = (size_t)(k0 + j) * C.dimdw + idw;
        c[j] = on ? cd[o] : (T)0;
        x[j] = on ? q[(size_t)C.nbr_dw[o] * C.ld + iup] : (T)0;
      }
#pragma unroll
      for (int j = 0; j < KCH; ++j) acc += c[j] * x[j];
    }
    if (I.Jx != 0.0 || I.Jp != 0.0) acc += nonlocal_row<T>(I, su, sd, q, C.ld);
    emit(col + iup, acc);
  }
}

// Record a tile's partial sum; returns true in every thread of the block
// that completed the chain's last tile of this launch.
static __device__ bool tile_done(const ChainDev& C, int idw, double acc, double* sh, int* flag) {
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) {
    C.part[idw] = s;
    __threadfence();
    const unsigned old = atomicAdd(C.counter, 1u);
    *flag = (old + 1u == (unsigned)C.ntiles) ? 1 : 0;
  }
  __syncthreads();
  const bool last = *flag != 0;
  __syncthreads();
  return last;
}

// Sum of the chain's tile partials in column order (deterministic), thread 0.
static __device__ double chain_total(const ChainDev& C, double* sh) {
  __threadfence();
  double v = 0.0;
  const int per = (C.ntiles + blockDim.x - 1) / blockDim.x;  // contiguous slice per thread
  const int a = threadIdx.x * per, b = min(a + per, C.ntiles);
  for (int s = a; s < b; ++s) v += __hip_atomic_load(&C.part[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return block_sum(v, sh);
}

// ---------------------------------------------------------------- phase A
// t = H q_k / s_k - beta_k q_{k-1} / s_{k-1}, written over q_{k-1};
// alpha_k = <t, v_k>.
template <typename T>
__global__ __launch_bounds__(1024) void k_lz_a(const ChainDev* __restrict__ chains, const int64_t* __restrict__ tile0,
                                               int nchains, InterParams I, int64_t Tall) {
  __shared__ double sh[16];
  __shared__ int flag;
  for (TileWalk w(Tall); w.t < w.end; w.t += w.step) {
    const int c = chain_of(tile0, nchains, w.t);
    const ChainDev& C = chains[c];
    const int idw = (int)(w.t - C.tile0);
    const LzScalars sc = *C.sc;
    if (sc.stopped) continue;
    const T* __restrict__ qk = reinterpret_cast<const T*>((sc.k & 1) ? C.q1 : C.q0);
    T* __restrict__ qo = reinterpret_cast<T*>((sc.k & 1) ? C.q0 : C.q1);
    const double inv_s = 1.0 / sc.s_cur;
    const T ts = (T)inv_s, tb = (T)(sc.beta_k / sc.s_old);
    double acc = 0.0;
    column_hq<T>(C, I, qk, idw, [&](size_t i, T h) {
      const T tv = h * ts - tb * qo[i];
      qo[i] = tv;
      acc += (double)tv * (double)qk[i];
    });
    if (tile_done(C, idw, acc * inv_s, sh, &flag)) {
      const double a = chain_total(C, sh);
      if (threadIdx.x == 0) {
        C.alpha[sc.k] = a;
        C.sc->alpha_k = a;
        *C.counter = 0u;
        __threadfence();
      }
    }
  }
}

// ---------------------------------------------------------------- phase B
// r = t - alpha_k v_k (kept as q_{k+1}, norm beta_{k+1}); optional replay
// accumulation x += c_k v_k.
template <typename T>
__global__ __launch_bounds__(1024) void k_lz_b(const ChainDev* __restrict__ chains, const int64_t* __restrict__ tile0,
                                               int nchains, int64_t Tall) {
  __shared__ double sh[16];
  __shared__ int flag;
  for (TileWalk w(Tall); w.t < w.end; w.t += w.step) {
    const int c = chain_of(tile0, nchains, w.t);
    const ChainDev& C = chains[c];
    const int idw = (int)(w.t - C.tile0);
    const LzScalars sc = *C.sc;
    if (sc.stopped) continue;
    const T* __restrict__ qk = reinterpret_cast<const T*>((sc.k & 1) ? C.q1 : C.q0);
    T* __restrict__ qo = reinterpret_cast<T*>((sc.k & 1) ? C.q0 : C.q1);
    const double inv_s = 1.0 / sc.s_cur;
    const T am = (T)(sc.alpha_k * inv_s);
    double* __restrict__ x = C.x;
    const double cx = x ? C.ccoef[sc.k] * inv_s : 0.0;
    const size_t col = (size_t)idw * C.ld, colx = (size_t)idw * C.dimup;
    double acc = 0.0;
    for (int iup = threadIdx.x; iup < C.dimup; iup += blockDim.x) {
      const size_t i = col + iup;
      const T qv = qk[i];
      const T r = qo[i] - am * qv;
      qo[i] = r;
      acc += (double)r * (double)r;
      if (x) x[colx + iup] += cx * (double)qv;
    }
    if (tile_done(C, idw, acc, sh, &flag)) {
      const double s2 = chain_total(C, sh);
      if (threadIdx.x == 0) {
        const double b = sqrt(s2);
        LzScalars n = sc;
        n.nsteps = sc.k + 1;
        const bool more = (sc.k + 1 < C.nmax);
        if (more) C.beta[sc.k + 1] = b;
        const bool brk = !(b > C.brk_tol * (fabs(sc.alpha_k) + sc.beta_k));
        if (!more || brk) {
          n.stopped = 1;
        } else {
          n.s_old = sc.s_cur;
          n.s_cur = b;
          n.beta_k = b;
          n.k = sc.k + 1;
        }
        *C.sc = n;
        *C.counter = 0u;
        __threadfence();
      }
    }
  }
}

// -------------------------------------------------------------- operator
template <typename T>
__global__ __launch_bounds__(1024) void k_hxv(ChainDev C, InterParams I, const T* __restrict__ v, T* __restrict__ out,
                                              int64_t Tall) {
  for (TileWalk w(Tall); w.t < w.end; w.t += w.step)
    column_hq<T>(C, I, v, (int)w.t, [&](size_t i, T h) { out[i] = h; });
}

// --------------------------------------------------------------- helpers
__global__ void k_start_vector(double* x, int64_t n, uint64_t seed) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    // SplitMix64 output for state seed + (i+1)*golden gamma, mapped to U[-1,1)
    uint64_t z = seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    x[i] = 2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
  }
}

template <typename T>
__global__ void k_convert(const double* __restrict__ s, T* __restrict__ d, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    d[i] = (T)s[i];
}
// Natural layout (rows x cols, column-major) to column stride ld >= rows.
template <typename T>
__global__ void k_convert_ld(const double* __restrict__ s, T* __restrict__ d, int rows, int64_t n, int64_t ld) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    d[(i / rows) * ld + i % rows] = (T)s[i];
}
// Column stride ld to the natural layout, widened to fp64.
template <typename T>
__global__ void k_to64_ld(const T* __restrict__ s, double* __restrict__ d, int rows, int64_t n, int64_t ld) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    d[i] = (double)s[(i / rows) * ld + i % rows];
}
template <typename T>
__global__ void k_to64(const T* __restrict__ s, double* __restrict__ d, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    d[i] = (double)s[i];
}
__global__ void k_axpy(double* __restrict__ y, const double* __restrict__ x, double a, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] += a * x[i];
}

__global__ void k_scale(double* x, int64_t n, double s) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] *= s;
}
__global__ void k_zero(uint32_t* p, size_t nwords) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}
__global__ void k_dot_partial(const double* __restrict__ a, const double* __restrict__ b, int64_t n, double* part) {
  __shared__ double sh[16];
  double v = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v += a[i] * b[i];
  const double s = block_sum(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
template <typename T>
__global__ void k_dot_partial_t(const T* __restrict__ a, const T* __restrict__ b, int64_t n, double* part) {
  __shared__ double sh[16];
  double v = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v += (double)a[i] * (double)b[i];
  const double s = block_sum(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
template <typename T>
__global__ void k_sumsq_partial(const T* __restrict__ a, int64_t n, double* part) {
  __shared__ double sh[16];
  double v = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v += (double)a[i] * (double)a[i];
  const double s = block_sum(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

static int grid_for(int64_t n) {
  int64_t g = (n + kBlock - 1) / kBlock;
  return (int)(g > 4096 ? 4096 : (g < 1 ? 1 : g));
}

void launch_axpy(double* y, const double* x, double a, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_axpy, dim3(grid_for(n)), dim3(kBlock), 0, st, y, x, a, n);
  EDL_HIP(hipGetLastError());
}

void launch_start_vector(double* d_x, int64_t n, uint64_t seed, hipStream_t st) {
  hipLaunchKernelGGL(k_start_vector, dim3(grid_for(n)), dim3(kBlock), 0, st, d_x, n, seed);
  EDL_HIP(hipGetLastError());
}
template <typename T>
void launch_convert(const double* src, T* dst, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_convert<T>, dim3(grid_for(n)), dim3(kBlock), 0, st, src, dst, n);
  EDL_HIP(hipGetLastError());
}
template <typename T>
void launch_convert_ld(const double* src, T* dst, int rows, int cols, int ld, hipStream_t st) {
  EDL_CHECK(ld >= rows && rows > 0, "launch_convert_ld: bad column stride");
  const int64_t n = (int64_t)rows * cols;
  if (ld == rows) return launch_convert<T>(src, dst, n, st);
  hipLaunchKernelGGL(k_convert_ld<T>, dim3(grid_for(n)), dim3(kBlock), 0, st, src, dst, rows, n, (int64_t)ld);
  EDL_HIP(hipGetLastError());
}
template <typename T>
void launch_to_fp64_ld(const T* src, double* dst, int rows, int cols, int ld, hipStream_t st) {
  EDL_CHECK(ld >= rows && rows > 0, "launch_to_fp64_ld: bad column stride");
  const int64_t n = (int64_t)rows * cols;
  if (ld == rows) return launch_to_fp64<T>(src, dst, n, st);
  hipLaunchKernelGGL(k_to64_ld<T>, dim3(grid_for(n)), dim3(kBlock), 0, st, src, dst, rows, n, (int64_t)ld);
  EDL_HIP(hipGetLastError());
}
template <typename T>
void launch_to_fp64(const T* src, double* dst, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_to64<T>, dim3(grid_for(n)), dim3(kBlock), 0, st, src, dst, n);
  EDL_HIP(hipGetLastError());
}
void launch_scale(double* x, int64_t n, double s, hipStream_t st) {
  hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(kBlock), 0, st, x, n, s);
  EDL_HIP(hipGetLastError());
}
void launch_zero_bytes(void* p, size_t bytes, hipStream_t st) {
  EDL_CHECK(bytes % 4 == 0, "launch_zero_bytes: size not a multiple of 4");
  const size_t nw = bytes / 4;
  if (!nw) return;
  hipLaunchKernelGGL(k_zero, dim3(grid_for((int64_t)nw)), dim3(kBlock), 0, st, (uint32_t*)p, nw);
  EDL_HIP(hipGetLastError());
}

// One block: deterministic total of g block partials (contiguous slice per
// thread, fixed tree) into *out.
__global__ void k_sum_partials(const double* __restrict__ part, int g, double* __restrict__ out) {
  __shared__ double sh[16];
  const int per = (g + blockDim.x - 1) / blockDim.x;
  const int a = threadIdx.x * per, e = min(a + per, g);
  double x = 0.0;
  for (int k = a; k < e; ++k) x += part[k];
  const double t = block_sum(x, sh);
  if (threadIdx.x == 0) *out = t;
}

int partials_size(int64_t n) { return grid_for(n); }

void device_dot_to(const double* a, const double* b, int64_t n, double* part, double* out, hipStream_t st) {
  const int g = grid_for(n);
  hipLaunchKernelGGL(k_dot_partial, dim3(g), dim3(kBlock), 0, st, a, b, n, part);
  EDL_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, st, part, g, out);
  EDL_HIP(hipGetLastError());
}

template <typename T>
void device_sumsq_to(const T* a, int64_t n, double* part, double* out, hipStream_t st) {
  const int g = grid_for(n);
  hipLaunchKernelGGL(k_sumsq_partial<T>, dim3(g), dim3(kBlock), 0, st, a, n, part);
  EDL_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, st, part, g, out);
  EDL_HIP(hipGetLastError());
}
template void device_sumsq_to<double>(const double*, int64_t, double*, double*, hipStream_t);
template void device_sumsq_to<float>(const float*, int64_t, double*, double*, hipStream_t);

// Initial scalars of chain k from its seed norm n2[k] and the sum of squares
// ss[k] of its stored start vector (nothing runs when either is zero).
__global__ void k_init_scalars(LzScalars* const* __restrict__ sc, const double* __restrict__ n2,
                               const double* __restrict__ ss, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  LzScalars s{};
  s.stopped = (n2[k] == 0.0 || ss[k] == 0.0) ? 1 : 0;
  s.s_cur = ss[k] > 0.0 ? sqrt(ss[k]) : 1.0;
  s.s_old = 1.0;
  s.norm2 = n2[k];
  *sc[k] = s;
}

void launch_init_scalars(LzScalars* const* d_sc, const double* n2, const double* ss, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_init_scalars, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, d_sc, n2, ss, n);
  EDL_HIP(hipGetLastError());
}

static double host_sum_partials(const double* part, int g, hipStream_t st) {
  std::vector<double> h(g);
  EDL_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * g, hipMemcpyDeviceToHost, st));
  EDL_HIP(hipStreamSynchronize(st));
  double s = 0.0;
  for (double v : h) s += v;
  return s;
}

double device_dot(Ctx& c, const double* a, const double* b, int64_t n, hipStream_t st) {
  const int g = grid_for(n);
  double* part = nullptr;
  EDL_HIP(hipMalloc(&part, sizeof(double) * g));
  hipLaunchKernelGGL(k_dot_partial, dim3(g), dim3(kBlock), 0, st, a, b, n, part);
  EDL_HIP(hipGetLastError());
  const double s = host_sum_partials(part, g, st);
  EDL_HIP(hipFree(part));
  return s;
}

template <typename T>
double device_sumsq(const T* a, int64_t n, hipStream_t st) {
  const int g = grid_for(n);
  double* part = nullptr;
  EDL_HIP(hipMalloc(&part, sizeof(double) * g));
  hipLaunchKernelGGL(k_sumsq_partial<T>, dim3(g), dim3(kBlock), 0, st, a, n, part);
  EDL_HIP(hipGetLastError());
  const double s = host_sum_partials(part, g, st);
  EDL_HIP(hipFree(part));
  return s;
}
template <typename T>
double device_dot_t(const T* a, const T* b, int64_t n, hipStream_t st) {
  const int g = grid_for(n);
  double* part = nullptr;
  EDL_HIP(hipMalloc(&part, sizeof(double) * g));
  hipLaunchKernelGGL(k_dot_partial_t<T>, dim3(g), dim3(kBlock), 0, st, a, b, n, part);
  EDL_HIP(hipGetLastError());
  const double s = host_sum_partials(part, g, st);
  EDL_HIP(hipFree(part));
  return s;
}
template double device_dot_t<double>(const double*, const double*, int64_t, hipStream_t);
template double device_dot_t<float>(const float*, const float*, int64_t, hipStream_t);
template double device_sumsq<double>(const double*, int64_t, hipStream_t);
template double device_sumsq<float>(const float*, int64_t, hipStream_t);

// ------------------------------------------------------------- schedule
void fill_sector_view(const Sector& s, bool fp32, ChainDev& cd) {
  cd.dimup = s.sp[0].dim;
  cd.dimdw = s.sp[1].dim;
  cd.kup = s.sp[0].kmax;
  cd.kdw = s.sp[1].kmax;
  cd.map_up = s.sp[0].map;
  cd.map_dw = s.sp[1].map;
  cd.nbr_up = s.sp[0].nbr;
  cd.nbr_dw = s.sp[1].nbr;
  cd.cnt_up = s.sp[0].cnt;
  cd.cnt_dw = s.sp[1].cnt;
  cd.cu = fp32 ? (const void*)s.sp[0].coef32 : (const void*)s.sp[0].coef64;
  cd.cd = fp32 ? (const void*)s.sp[1].coef32 : (const void*)s.sp[1].coef64;
  cd.dg_up = s.sp[0].diag;
  cd.dg_dw = s.sp[1].diag;
  cd.ntiles = cd.dimdw;
  cd.ld = cd.dimup;
}

static int pick_grid(Ctx& c, int64_t total_tiles, const Schedule& w) {
  int64_t g = w.grid > 0 ? w.grid : (int64_t)c.num_cu * 8;
  if (g > total_tiles) g = total_tiles;
  if (g >= 8) g -= g % 8;
  return (int)(g < 1 ? 1 : g);
}

template <typename T>
ChainSet<T>::~ChainSet() {
  for (void* p : owned) (void)hipFree(p);
}

template <typename T>
void chainset_finalize(Ctx& c, ChainSet<T>& cs, const Schedule& want) {
  cs.sch = want;
  int64_t tot = 0;
  std::vector<int64_t> t0(cs.h.size() + 1);
  size_t part_total = 0;
  for (size_t k = 0; k < cs.h.size(); ++k) {
    cs.h[k].ntiles = cs.h[k].dimdw;
    cs.h[k].tile0 = tot;
    t0[k] = tot;
    tot += cs.h[k].ntiles;
    part_total += (size_t)cs.h[k].ntiles;
  }
  t0[cs.h.size()] = tot;
  cs.sch.total_tiles = tot;
  cs.sch.grid = pick_grid(c, tot, want);
  double* d_part = nullptr;
  EDL_HIP(hipMalloc(&d_part, sizeof(double) * (part_total ? part_total : 1)));
  cs.owned.push_back(d_part);
  size_t off = 0;
  for (auto& cd : cs.h) {
    cd.part = d_part + off;
    off += (size_t)cd.ntiles;
  }
  EDL_HIP(hipMalloc(&cs.d, sizeof(ChainDev) * cs.h.size()));
  cs.owned.push_back(cs.d);
  EDL_HIP(hipMalloc(&cs.d_tile0, sizeof(int64_t) * t0.size()));
  cs.owned.push_back(cs.d_tile0);
  EDL_HIP(hipMemcpyAsync(cs.d, cs.h.data(), sizeof(ChainDev) * cs.h.size(), hipMemcpyHostToDevice, c.stream));
  EDL_HIP(hipMemcpyAsync(cs.d_tile0, t0.data(), sizeof(int64_t) * t0.size(), hipMemcpyHostToDevice, c.stream));
  // t0 is local: the copies must have read their host sources before return
  EDL_HIP(hipStreamSynchronize(c.stream));
}

template <typename T>
void lz_phase_a(Ctx& c, ChainSet<T>& cs, hipStream_t st) {
  hipLaunchKernelGGL(k_lz_a<T>, dim3(cs.sch.grid), dim3(cs.sch.block_threads), 0, st, cs.d, cs.d_tile0,
                     (int)cs.h.size(), inter_of(c), cs.sch.total_tiles);
  EDL_HIP(hipGetLastError());
}

template <typename T>
void lz_phase_b(Ctx& c, ChainSet<T>& cs, hipStream_t st) {
  hipLaunchKernelGGL(k_lz_b<T>, dim3(cs.sch.grid), dim3(cs.sch.block_threads), 0, st, cs.d, cs.d_tile0,
                     (int)cs.h.size(), cs.sch.total_tiles);
  EDL_HIP(hipGetLastError());
}

template <typename T>
void hxv_operator(Ctx& c, const Sector& s, const T* v, T* out, hipStream_t st, const Schedule& want) {
  ChainDev cd{};
  fill_sector_view(s, sizeof(T) == 4, cd);
  const int G = pick_grid(c, cd.ntiles, want);
  hipLaunchKernelGGL(k_hxv<T>, dim3(G), dim3(want.block_threads), 0, st, cd, inter_of(c), v, out,
                     (int64_t)cd.ntiles);
  EDL_HIP(hipGetLastError());
}

template void chainset_finalize<double>(Ctx&, ChainSet<double>&, const Schedule&);
template void chainset_finalize<float>(Ctx&, ChainSet<float>&, const Schedule&);
template ChainSet<double>::~ChainSet();
template ChainSet<float>::~ChainSet();
template void lz_phase_a<double>(Ctx&, ChainSet<double>&, hipStream_t);
template void lz_phase_a<float>(Ctx&, ChainSet<float>&, hipStream_t);
template void lz_phase_b<double>(Ctx&, ChainSet<double>&, hipStream_t);
template void lz_phase_b<float>(Ctx&, ChainSet<float>&, hipStream_t);
template void hxv_operator<double>(Ctx&, const Sector&, const double*, double*, hipStream_t, const Schedule&);
template void hxv_operator<float>(Ctx&, const Sector&, const float*, float*, hipStream_t, const Schedule&);
template void launch_convert<double>(const double*, double*, int64_t, hipStream_t);
template void launch_convert<float>(const double*, float*, int64_t, hipStream_t);
template void launch_convert_ld<double>(const double*, double*, int, int, int, hipStream_t);
template void launch_convert_ld<float>(const double*, float*, int, int, int, hipStream_t);
template void launch_to_fp64<double>(const double*, double*, int64_t, hipStream_t);
template void launch_to_fp64<float>(const float*, double*, int64_t, hipStream_t);
template void launch_to_fp64_ld<double>(const double*, double*, int, int, int, hipStream_t);
template void launch_to_fp64_ld<float>(const float*, double*, int, int, int, hipStream_t);

}  // namespace edl
