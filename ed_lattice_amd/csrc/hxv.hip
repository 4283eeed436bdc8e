// On-the-fly H x v on the up (x) down factorised basis, and the two fused
// phases of a Lanczos step.
//
// A vector of sector (Nup, Ndw) is the column-major matrix V(iup, idw),
// element i = iup + idw*DimUp (SECTOR.f90:443-480, 0-based here).  In
// gather form every output element is
//
//   (H v)(iup,idw) = [dg_up(iup) + dg_dw(idw) + Uint(s_up & s_dw)] v(iup,idw)
//                  + sum_k cu_k(iup) v(nu_k(iup), idw)      (HxV_up.f90)
//                  + sum_k cd_k(idw) v(iup, nd_k(idw))      (HxV_dw.f90)
//
// with one writer per element, so no atomics and no write races.  The up
// hops gather inside one column; the down hops read whole neighbour
// columns, coalesced across iup.  The neighbour lists of the column are
// uniform over a tile and are read once per tile.
//
// Work is cut into tiles (one column, a range of rows) and handed to a
// persistent grid of logical blocks, each taking a contiguous tile range.
// Physical block b runs logical block (b % 8) * (G / 8) + b / 8 so that the
// contiguous tile ranges of one XCD cover neighbouring columns and their
// up-hop gathers stay in that XCD's L2.
#include "chains.hpp"

namespace edl {

static __device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Deterministic block sum (fixed shuffle tree, fixed wave order).
static __device__ double block_sum(double v, double* sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < nw; ++w) r += sh[w];
  return r;  // valid in thread 0
}

static __device__ __forceinline__ int logical_block(int G) {
  const int b = blockIdx.x;
  return (G >= 8 && (G % 8) == 0) ? (b % 8) * (G / 8) + b / 8 : b;
}

static __device__ __forceinline__ double inter_term(const InterParams& I, uint32_t su, uint32_t sd) {
  const uint32_t both = su & sd;
  double e = 0.0;
  if (I.norb == 1) {
    e = I.U[0] * (double)__popc(both & I.orb_mask[0]);
  } else {
    for (int k = 0; k < I.norb; ++k) e += I.U[k] * (double)__popc(both & I.orb_mask[k]);
  }
  for (int p = 0; p < I.npairs; ++p) {
    const uint32_t a = I.pi[p], b = I.pj[p];
    e += I.Ust * (double)((((su >> a) & (sd >> b)) & 1u) + (((su >> b) & (sd >> a)) & 1u));
  }
  return e;
}

// (H q)(iup, idw) in the scalar type T.
template <typename T>
static __device__ __forceinline__ T hq_elem(const ChainDev& C, const InterParams& I, const T* __restrict__ q,
                                            int iup, int idw, size_t col, uint32_t sd, double dgd) {
  const T* __restrict__ cu = reinterpret_cast<const T*>(C.cu);
  const T* __restrict__ cd = reinterpret_cast<const T*>(C.cd);
  const uint32_t su = (uint32_t)C.map_up[iup];
  const double dg = C.dg_up[iup] + dgd + inter_term(I, su, sd);
  T acc = (T)dg * q[col + iup];
  for (int k = 0; k < C.kup; ++k) {
    const size_t o = (size_t)k * C.dimup + iup;
    acc += cu[o] * q[col + (size_t)C.nbr_up[o]];
  }
  for (int k = 0; k < C.kdw; ++k) {
    const size_t o = (size_t)k * C.dimdw + idw;
    acc += cd[o] * q[(size_t)C.nbr_dw[o] * C.dimup + iup];
  }
  return acc;
}

static __device__ __forceinline__ int chain_of(const int64_t* __restrict__ tile0, int nchains, int64_t t) {
  int lo = 0, hi = nchains - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tile0[mid] <= t) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Flush the block's partial sum of chain C; returns true in every thread of
// the block that completed the chain (the last of its `span` blocks).
static __device__ bool flush_partial(const ChainDev& C, int Lb, double acc, double* sh, int* flag) {
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) {
    C.part[Lb - C.lb_first] = s;
    __threadfence();
    const unsigned old = atomicAdd(C.counter, 1u);
    *flag = (old + 1u == (unsigned)C.span) ? 1 : 0;
  }
  __syncthreads();
  const bool last = *flag != 0;
  __syncthreads();
  return last;
}

// Sum of the chain's partials in slot order (deterministic), thread 0.
static __device__ double gather_partials(const ChainDev& C, double* sh) {
  __threadfence();
  double v = 0.0;
  for (int s = threadIdx.x; s < C.span; s += blockDim.x)
    v += __hip_atomic_load(&C.part[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return block_sum(v, sh);
}

// ---------------------------------------------------------------- phase A
// t = H q_k / s_k - beta_k q_{k-1} / s_{k-1}, written over q_{k-1};
// alpha_k = <t, v_k>.
template <typename T>
__global__ __launch_bounds__(1024) void k_lz_a(const ChainDev* __restrict__ chains, const int64_t* __restrict__ tile0,
                                               int nchains, InterParams I, int G, int64_t Tall) {
  __shared__ double sh[16];
  __shared__ int flag;
  const int Lb = logical_block(G);
  const int64_t tb = (int64_t)Lb * Tall / G, te = (int64_t)(Lb + 1) * Tall / G;
  int c = chain_of(tile0, nchains, tb);
  for (int64_t t = tb; t < te; ++c) {
    const ChainDev& C = chains[c];
    const int64_t tstop = te < tile0[c + 1] ? te : tile0[c + 1];
    const LzScalars sc = *C.sc;
    if (sc.stopped) { t = tstop; continue; }
    const T* __restrict__ qk = reinterpret_cast<const T*>((sc.k & 1) ? C.q1 : C.q0);
    T* __restrict__ qo = reinterpret_cast<T*>((sc.k & 1) ? C.q0 : C.q1);
    const double inv_s = 1.0 / sc.s_cur;
    const T ts = (T)inv_s, tb_o = (T)(sc.beta_k / sc.s_old);
    double acc = 0.0;
    for (; t < tstop; ++t) {
      const int64_t lt = t - C.tile0;
      const int idw = (int)(lt / C.ntr), rt = (int)(lt % C.ntr);
      const int r0 = rt * C.rows_per_tile;
      const int r1 = min(r0 + C.rows_per_tile, C.dimup);
      const size_t col = (size_t)idw * C.dimup;
      const uint32_t sd = (uint32_t)C.map_dw[idw];
      const double dgd = C.dg_dw[idw];
      for (int iup = r0 + threadIdx.x; iup < r1; iup += blockDim.x) {
        const T h = hq_elem<T>(C, I, qk, iup, idw, col, sd, dgd);
        const size_t i = col + iup;
        const T tv = h * ts - tb_o * qo[i];
        qo[i] = tv;
        acc += (double)tv * (double)qk[i];
      }
    }
    if (flush_partial(C, Lb, acc * inv_s, sh, &flag)) {
      const double a = gather_partials(C, sh);
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
                                               int nchains, int G, int64_t Tall) {
  __shared__ double sh[16];
  __shared__ int flag;
  const int Lb = logical_block(G);
  const int64_t tb = (int64_t)Lb * Tall / G, te = (int64_t)(Lb + 1) * Tall / G;
  int c = chain_of(tile0, nchains, tb);
  for (int64_t t = tb; t < te; ++c) {
    const ChainDev& C = chains[c];
    const int64_t tstop = te < tile0[c + 1] ? te : tile0[c + 1];
    const LzScalars sc = *C.sc;
    if (sc.stopped) { t = tstop; continue; }
    const T* __restrict__ qk = reinterpret_cast<const T*>((sc.k & 1) ? C.q1 : C.q0);
    T* __restrict__ qo = reinterpret_cast<T*>((sc.k & 1) ? C.q0 : C.q1);
    const double inv_s = 1.0 / sc.s_cur;
    const T am = (T)(sc.alpha_k * inv_s);
    double* __restrict__ x = C.x;
    const double cx = x ? C.ccoef[sc.k] * inv_s : 0.0;
    double acc = 0.0;
    for (; t < tstop; ++t) {
      const int64_t lt = t - C.tile0;
      const int idw = (int)(lt / C.ntr), rt = (int)(lt % C.ntr);
      const int r0 = rt * C.rows_per_tile;
      const int r1 = min(r0 + C.rows_per_tile, C.dimup);
      const size_t col = (size_t)idw * C.dimup;
      for (int iup = r0 + threadIdx.x; iup < r1; iup += blockDim.x) {
        const size_t i = col + iup;
        const T q = qk[i];
        const T r = qo[i] - am * q;
        qo[i] = r;
        acc += (double)r * (double)r;
        if (x) x[i] += cx * (double)q;
      }
    }
    if (flush_partial(C, Lb, acc, sh, &flag)) {
      const double s2 = gather_partials(C, sh);
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
                                              int G, int64_t Tall) {
  const int Lb = logical_block(G);
  const int64_t tb = (int64_t)Lb * Tall / G, te = (int64_t)(Lb + 1) * Tall / G;
  for (int64_t t = tb; t < te; ++t) {
    const int idw = (int)(t / C.ntr), rt = (int)(t % C.ntr);
    const int r0 = rt * C.rows_per_tile;
    const int r1 = min(r0 + C.rows_per_tile, C.dimup);
    const size_t col = (size_t)idw * C.dimup;
    const uint32_t sd = (uint32_t)C.map_dw[idw];
    const double dgd = C.dg_dw[idw];
    for (int iup = r0 + threadIdx.x; iup < r1; iup += blockDim.x)
      out[col + iup] = hq_elem<T>(C, I, v, iup, idw, col, sd, dgd);
  }
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
template <typename T>
__global__ void k_to64(const T* __restrict__ s, double* __restrict__ d, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    d[i] = (double)s[i];
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
double device_dot(Ctx& c, const double* a, const double* b, int64_t n, hipStream_t st) {
  const int g = grid_for(n);
  double* part = nullptr;
  EDL_HIP(hipMalloc(&part, sizeof(double) * g));
  hipLaunchKernelGGL(k_dot_partial, dim3(g), dim3(kBlock), 0, st, a, b, n, part);
  EDL_HIP(hipGetLastError());
  std::vector<double> h(g);
  EDL_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * g, hipMemcpyDeviceToHost, st));
  EDL_HIP(hipStreamSynchronize(st));
  EDL_HIP(hipFree(part));
  double s = 0.0;
  for (double v : h) s += v;
  return s;
}

template <typename T>
double device_sumsq(const T* a, int64_t n, hipStream_t st) {
  const int g = grid_for(n);
  double* part = nullptr;
  EDL_HIP(hipMalloc(&part, sizeof(double) * g));
  hipLaunchKernelGGL(k_sumsq_partial<T>, dim3(g), dim3(kBlock), 0, st, a, n, part);
  EDL_HIP(hipGetLastError());
  std::vector<double> h(g);
  EDL_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * g, hipMemcpyDeviceToHost, st));
  EDL_HIP(hipStreamSynchronize(st));
  EDL_HIP(hipFree(part));
  double s = 0.0;
  for (double v : h) s += v;
  return s;
}
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
  cd.cu = fp32 ? (const void*)s.sp[0].coef32 : (const void*)s.sp[0].coef64;
  cd.cd = fp32 ? (const void*)s.sp[1].coef32 : (const void*)s.sp[1].coef64;
  cd.dg_up = s.sp[0].diag;
  cd.dg_dw = s.sp[1].diag;
}

static void tile_chain(ChainDev& cd, const Schedule& w) {
  const int per = w.block_threads * w.rows_per_thread;
  cd.ntr = (cd.dimup + per - 1) / per;
  cd.rows_per_tile = (cd.dimup + cd.ntr - 1) / cd.ntr;
  cd.ntiles = (int64_t)cd.ntr * cd.dimdw;
}

static int pick_grid(Ctx& c, int64_t total_tiles, const Schedule& w) {
  int64_t g = w.grid > 0 ? w.grid : (int64_t)c.num_cu * 8;
  if (g > total_tiles) g = total_tiles;
  if (g >= 8) g -= g % 8;
  return (int)(g < 1 ? 1 : g);
}

// first logical block whose tile range contains tile t
static int lb_of(int64_t t, int64_t Tall, int G) {
  int lo = 0, hi = G - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if ((int64_t)mid * Tall / G <= t) lo = mid; else hi = mid - 1;
  }
  return lo;
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
  for (size_t k = 0; k < cs.h.size(); ++k) {
    tile_chain(cs.h[k], want);
    cs.h[k].tile0 = tot;
    t0[k] = tot;
    tot += cs.h[k].ntiles;
  }
  t0[cs.h.size()] = tot;
  cs.sch.total_tiles = tot;
  cs.sch.grid = pick_grid(c, tot, want);
  const int G = cs.sch.grid;
  size_t part_total = 0;
  for (auto& cd : cs.h) {
    cd.lb_first = lb_of(cd.tile0, tot, G);
    cd.span = lb_of(cd.tile0 + cd.ntiles - 1, tot, G) - cd.lb_first + 1;
    part_total += (size_t)cd.span;
  }
  double* d_part = nullptr;
  EDL_HIP(hipMalloc(&d_part, sizeof(double) * (part_total ? part_total : 1)));
  cs.owned.push_back(d_part);
  size_t off = 0;
  for (auto& cd : cs.h) {
    cd.part = d_part + off;
    off += (size_t)cd.span;
  }
  EDL_HIP(hipMalloc(&cs.d, sizeof(ChainDev) * cs.h.size()));
  cs.owned.push_back(cs.d);
  EDL_HIP(hipMalloc(&cs.d_tile0, sizeof(int64_t) * t0.size()));
  cs.owned.push_back(cs.d_tile0);
  EDL_HIP(hipMemcpyAsync(cs.d, cs.h.data(), sizeof(ChainDev) * cs.h.size(), hipMemcpyHostToDevice, c.stream));
  EDL_HIP(hipMemcpyAsync(cs.d_tile0, t0.data(), sizeof(int64_t) * t0.size(), hipMemcpyHostToDevice, c.stream));
}

template <typename T>
void lz_phase_a(Ctx& c, ChainSet<T>& cs, hipStream_t st) {
  hipLaunchKernelGGL(k_lz_a<T>, dim3(cs.sch.grid), dim3(cs.sch.block_threads), 0, st, cs.d, cs.d_tile0,
                     (int)cs.h.size(), inter_of(c), cs.sch.grid, cs.sch.total_tiles);
  EDL_HIP(hipGetLastError());
}

template <typename T>
void lz_phase_b(Ctx& c, ChainSet<T>& cs, hipStream_t st) {
  hipLaunchKernelGGL(k_lz_b<T>, dim3(cs.sch.grid), dim3(cs.sch.block_threads), 0, st, cs.d, cs.d_tile0,
                     (int)cs.h.size(), cs.sch.grid, cs.sch.total_tiles);
  EDL_HIP(hipGetLastError());
}

template <typename T>
void hxv_operator(Ctx& c, const Sector& s, const T* v, T* out, hipStream_t st, const Schedule& want) {
  ChainDev cd{};
  fill_sector_view(s, sizeof(T) == 4, cd);
  tile_chain(cd, want);
  const int G = pick_grid(c, cd.ntiles, want);
  hipLaunchKernelGGL(k_hxv<T>, dim3(G), dim3(want.block_threads), 0, st, cd, inter_of(c), v, out, G, cd.ntiles);
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
template void launch_to_fp64<double>(const double*, double*, int64_t, hipStream_t);
template void launch_to_fp64<float>(const float*, double*, int64_t, hipStream_t);

}  // namespace edl
