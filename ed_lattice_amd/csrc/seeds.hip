// Seed vectors of the GF and chi continued-fraction runs, built on the
// device from the ground state, and ground-state observables.
//
// Reference seed builders, all in scatter form on the host:
//   c^dg_io |gs>, c_io |gs>          GF_ELECTRONS.f90:241-245, 267-271
//   (c^dg_io + c^dg_jo)|gs>, (c_io + c_jo)|gs>   GF_ELECTRONS.f90:343-353, 376-386
//   S^z_io |gs>, (S^z_io + S^z_jo)|gs>           CHI_SPIN_ELECTRONS.f90:216-220, 275-280
// with apply_op_C / apply_op_CDG / apply_op_Sz of SECTOR.f90:217-318.
// Here each element j of the target sector gathers its unique source
// element of the ground-state sector: for c^dg_p the source string is the
// target string with bit p cleared, for c_p the target with bit p set; the
// sign (-1)^popcount(bits below p) is the c/cdg sign of AUX_FUNX.f90:76-102
// (bits below p agree between source and target).  Up and down operators
// commute: no sign from the other spin (SECTOR.f90:233-240).
#include "chains.hpp"
#include "seeds.hpp"

namespace edl {

struct SeedTerm {
  int op;      // 1 = C, 2 = CDG, 3 = SZ
  int ispin;   // 1 up, 2 down (C, CDG)
  int pos;     // 0-based site
};

// One term of a seed at target element (jup, jdw); returns coefficient * gs.
static __device__ __forceinline__ double seed_term(const SeedTerm& T, const double* __restrict__ gs, int gs_dimup,
                                                   uint32_t tu, uint32_t td, int jup, int jdw) {
  if (T.op == 3) {
    const double sz = 0.5 * ((double)((tu >> T.pos) & 1u) - (double)((td >> T.pos) & 1u));
    return sz * gs[(size_t)jdw * gs_dimup + jup];
  }
  const uint32_t t = (T.ispin == 1) ? tu : td;
  const uint32_t bit = 1u << T.pos;
  uint32_t src;
  if (T.op == 2) {  // c^dg_p: target must hold the electron
    if (!(t & bit)) return 0.0;
    src = t & ~bit;
  } else {          // c_p: target must miss it
    if (t & bit) return 0.0;
    src = t | bit;
  }
  const double sg = (__popc(t & (bit - 1u)) & 1) ? -1.0 : 1.0;
  const int32_t r = rank_of(src);
  const size_t o = (T.ispin == 1) ? (size_t)jdw * gs_dimup + (size_t)r : (size_t)r * gs_dimup + (size_t)jup;
  return sg * gs[o];
}

__global__ void k_seed(const double* __restrict__ gs, int gs_dimup, const int32_t* __restrict__ map_up,
                       const int32_t* __restrict__ map_dw, int dimup, int64_t dim, SeedTerm a, SeedTerm b, int nterms,
                       double* __restrict__ out) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < dim; j += (int64_t)gridDim.x * blockDim.x) {
    const int jup = (int)(j % dimup), jdw = (int)(j / dimup);
    const uint32_t tu = (uint32_t)map_up[jup], td = (uint32_t)map_dw[jdw];
    double v = seed_term(a, gs, gs_dimup, tu, td, jup, jdw);
    if (nterms > 1) v += seed_term(b, gs, gs_dimup, tu, td, jup, jdw);
    out[j] = v;
  }
}

// Current seed in gather form: term coef * c+_a c_b reaches target j from
// the source with a (of that spin) moved back to b; the reference's signs
// are those of c_b on the source and then c+_a (AUX_FUNX.f90:76-102).
__global__ void k_seed_current(const double* __restrict__ gs, const int32_t* __restrict__ map_up,
                               const int32_t* __restrict__ map_dw, int dimup, int64_t dim, CurrentSeed J,
                               double* __restrict__ out) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < dim; j += (int64_t)gridDim.x * blockDim.x) {
    const int jup = (int)(j % dimup), jdw = (int)(j / dimup);
    const uint32_t tu = (uint32_t)map_up[jup], td = (uint32_t)map_dw[jdw];
    double v = 0.0;
    for (int k = 0; k < J.n; ++k) {
      const CurrentTerm T = J.t[k];
      const uint32_t t = (T.spin == 1) ? tu : td;
      const uint32_t ba = 1u << T.a, bb = 1u << T.b;
      if (T.a == T.b || !(t & ba) || (t & bb)) continue;
      const uint32_t src = (t & ~ba) | bb;
      const uint32_t k1 = src & ~bb;
      const double sg = ((__popc(src & (bb - 1u)) + __popc(k1 & (ba - 1u))) & 1) ? -1.0 : 1.0;
      const int32_t r = rank_of(src);
      const size_t o = (T.spin == 1) ? (size_t)jdw * dimup + (size_t)r : (size_t)r * dimup + (size_t)jup;
      v += T.coef * sg * gs[o];
    }
    out[j] = v;
  }
}

void launch_current_seed(const CurrentSeed& J, const double* gs, const Sector& s, double* out, hipStream_t st) {
  int64_t g = (s.dim + kBlock - 1) / kBlock;
  if (g > 8192) g = 8192;
  hipLaunchKernelGGL(k_seed_current, dim3((int)g), dim3(kBlock), 0, st, gs, s.sp[0].map, s.sp[1].map, s.sp[0].dim,
                     s.dim, J, out);
  EDL_HIP(hipGetLastError());
}

// Per-site double occupancy sum_i |gs(i)|^2 n_up(p) n_dw(p), and the row /
// column sums of |gs|^2 from which the densities follow.
template <int NS>
__global__ void k_docc(const double* __restrict__ gs, const int32_t* __restrict__ map_up,
                       const int32_t* __restrict__ map_dw, int dimup, int64_t dim, double* __restrict__ part) {
  __shared__ double sh[NS][16];
  double acc[NS];
#pragma unroll
  for (int p = 0; p < NS; ++p) acc[p] = 0.0;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < dim; j += (int64_t)gridDim.x * blockDim.x) {
    const int jup = (int)(j % dimup), jdw = (int)(j / dimup);
    const uint32_t both = (uint32_t)map_up[jup] & (uint32_t)map_dw[jdw];
    const double w = gs[j] * gs[j];
#pragma unroll
    for (int p = 0; p < NS; ++p) acc[p] += (double)((both >> p) & 1u) * w;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int p = 0; p < NS; ++p) {
    double v = acc[p];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) sh[p][wid] = v;
  }
  __syncthreads();
  if (threadIdx.x < NS) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += sh[threadIdx.x][w];
    part[(size_t)blockIdx.x * NS + threadIdx.x] = s;
  }
}

// sum over columns (rowsum[iup]) and over rows (colsum[idw]) of |gs|^2
__global__ void k_colsum(const double* __restrict__ gs, int dimup, int dimdw, double* __restrict__ colsum) {
  __shared__ double sh[16];
  const int idw = blockIdx.x;
  double v = 0.0;
  for (int i = threadIdx.x; i < dimup; i += blockDim.x) {
    const double g = gs[(size_t)idw * dimup + i];
    v += g * g;
  }
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sh[w];
    colsum[idw] = s;
  }
}
__global__ void k_rowsum(const double* __restrict__ gs, int dimup, int dimdw, double* __restrict__ rowsum) {
  const int iup = blockIdx.x * blockDim.x + threadIdx.x;
  if (iup >= dimup) return;
  double v = 0.0;
  for (int idw = 0; idw < dimdw; ++idw) {
    const double g = gs[(size_t)idw * dimup + iup];
    v += g * g;
  }
  rowsum[iup] = v;
}

void launch_seed(const SeedSpec& s, const double* gs, int gs_dimup, const Sector& target, double* out,
                 hipStream_t st) {
  SeedTerm a{s.op, s.ispin, s.io - 1}, b{s.op, s.ispin, (s.jo > 0 ? s.jo : s.io) - 1};
  const int nterms = s.jo > 0 ? 2 : 1;
  int64_t g = (target.dim + kBlock - 1) / kBlock;
  if (g > 8192) g = 8192;
  hipLaunchKernelGGL(k_seed, dim3((int)g), dim3(kBlock), 0, st, gs, gs_dimup, target.sp[0].map, target.sp[1].map,
                     target.sp[0].dim, target.dim, a, b, nterms, out);
  EDL_HIP(hipGetLastError());
}

void device_observables(Ctx& c, const Sector& s, const double* gs, double* dens_up, double* dens_dw, double* docc) {
  const int Ns = c.model.Ns;
  hipStream_t st = c.stream;
  const int dimup = s.sp[0].dim, dimdw = s.sp[1].dim;
  const int nb = 1024;
  double* part = nullptr;
  double* rs = nullptr;
  double* cs = nullptr;
  EDL_HIP(hipMalloc(&part, sizeof(double) * nb * 32));
  EDL_HIP(hipMalloc(&rs, sizeof(double) * dimup));
  EDL_HIP(hipMalloc(&cs, sizeof(double) * dimdw));
  hipLaunchKernelGGL(k_docc<32>, dim3(nb), dim3(kBlock), 0, st, gs, s.sp[0].map, s.sp[1].map, dimup, s.dim, part);
  hipLaunchKernelGGL(k_rowsum, dim3((dimup + kBlock - 1) / kBlock), dim3(kBlock), 0, st, gs, dimup, dimdw, rs);
  hipLaunchKernelGGL(k_colsum, dim3(dimdw), dim3(kBlock), 0, st, gs, dimup, dimdw, cs);
  EDL_HIP(hipGetLastError());
  std::vector<double> hp((size_t)nb * 32), hr(dimup), hc(dimdw);
  std::vector<int32_t> mu(dimup), md(dimdw);
  EDL_HIP(hipMemcpyAsync(hp.data(), part, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, st));
  EDL_HIP(hipMemcpyAsync(hr.data(), rs, sizeof(double) * dimup, hipMemcpyDeviceToHost, st));
  EDL_HIP(hipMemcpyAsync(hc.data(), cs, sizeof(double) * dimdw, hipMemcpyDeviceToHost, st));
  EDL_HIP(hipMemcpyAsync(mu.data(), s.sp[0].map, sizeof(int32_t) * dimup, hipMemcpyDeviceToHost, st));
  EDL_HIP(hipMemcpyAsync(md.data(), s.sp[1].map, sizeof(int32_t) * dimdw, hipMemcpyDeviceToHost, st));
  EDL_HIP(hipStreamSynchronize(st));
  EDL_HIP(hipFree(part));
  EDL_HIP(hipFree(rs));
  EDL_HIP(hipFree(cs));
  for (int p = 0; p < Ns; ++p) {
    double d = 0.0, u = 0.0, w = 0.0;
    for (int b = 0; b < nb; ++b) d += hp[(size_t)b * 32 + p];
    for (int i = 0; i < dimup; ++i) u += ((mu[i] >> p) & 1) ? hr[i] : 0.0;
    for (int i = 0; i < dimdw; ++i) w += ((md[i] >> p) & 1) ? hc[i] : 0.0;
    docc[p] = d;
    dens_up[p] = u;
    dens_dw[p] = w;
  }
}

// A[idw * 32 + p] = sum over the rows of column idw of |gs|^2 * (bit p of
// the up string): the up/down cross term of <n_i n_j> and <Sz_i Sz_j>.
template <int NS>
__global__ void k_colbits(const double* __restrict__ gs, const int32_t* __restrict__ map_up, int dimup,
                          double* __restrict__ A) {
  __shared__ double sh[NS][16];
  const int idw = blockIdx.x;
  double acc[NS];
#pragma unroll
  for (int p = 0; p < NS; ++p) acc[p] = 0.0;
  for (int i = threadIdx.x; i < dimup; i += blockDim.x) {
    const double g = gs[(size_t)idw * dimup + i];
    const uint32_t su = (uint32_t)map_up[i];
#pragma unroll
    for (int p = 0; p < NS; ++p) acc[p] += (double)((su >> p) & 1u) * g * g;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int p = 0; p < NS; ++p) {
    double v = acc[p];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) sh[p][wid] = v;
  }
  __syncthreads();
  if (threadIdx.x < NS) {
    double t = 0.0;
    for (int w = 0; w < nw; ++w) t += sh[threadIdx.x][w];
    A[(size_t)idw * NS + threadIdx.x] = t;
  }
}

// <Sz_i Sz_j> and <n_i n_j> of the installed state (the sz2 and n2 of
// lanc_observables, OBSERVABLES.f90:142-150) from three contractions of
// W = |gs|^2: UU_ij = sum_up rowsum * u_i u_j, DD_ij = sum_dw colsum * d_i d_j
// and UD_ij = sum W u_i d_j; n2 = UU + DD + UD + UD^T, sz2 = (UU + DD - UD -
// UD^T) / 4.  Outputs Ns x Ns, column-major.
void device_correlations(Ctx& c, const Sector& s, const double* gs, double* sz2, double* n2) {
  const int Ns = c.model.Ns;
  hipStream_t st = c.stream;
  const int dimup = s.sp[0].dim, dimdw = s.sp[1].dim;
  struct Buf {  // freed on every exit, errors included
    double* p = nullptr;
    explicit Buf(size_t n) { EDL_HIP(hipMalloc(&p, sizeof(double) * n)); }
    ~Buf() { if (p) (void)hipFree(p); }
  } rs(dimup), cs(dimdw), A((size_t)32 * dimdw);
  hipLaunchKernelGGL(k_rowsum, dim3((dimup + kBlock - 1) / kBlock), dim3(kBlock), 0, st, gs, dimup, dimdw, rs.p);
  hipLaunchKernelGGL(k_colsum, dim3(dimdw), dim3(kBlock), 0, st, gs, dimup, dimdw, cs.p);
  hipLaunchKernelGGL(k_colbits<32>, dim3(dimdw), dim3(kBlock), 0, st, gs, s.sp[0].map, dimup, A.p);
  EDL_HIP(hipGetLastError());
  std::vector<double> hr(dimup), hc(dimdw), ha((size_t)32 * dimdw);
  std::vector<int32_t> mu(dimup), md(dimdw);
  EDL_HIP(hipMemcpyAsync(hr.data(), rs.p, sizeof(double) * dimup, hipMemcpyDeviceToHost, st));
  EDL_HIP(hipMemcpyAsync(hc.data(), cs.p, sizeof(double) * dimdw, hipMemcpyDeviceToHost, st));
  EDL_HIP(hipMemcpyAsync(ha.data(), A.p, sizeof(double) * ha.size(), hipMemcpyDeviceToHost, st));
  EDL_HIP(hipMemcpyAsync(mu.data(), s.sp[0].map, sizeof(int32_t) * dimup, hipMemcpyDeviceToHost, st));
  EDL_HIP(hipMemcpyAsync(md.data(), s.sp[1].map, sizeof(int32_t) * dimdw, hipMemcpyDeviceToHost, st));
  EDL_HIP(hipStreamSynchronize(st));
  std::vector<double> uu((size_t)Ns * Ns, 0.0), dd((size_t)Ns * Ns, 0.0), ud((size_t)Ns * Ns, 0.0);
  for (int a = 0; a < dimup; ++a)
    for (int i = 0; i < Ns; ++i)
      if ((mu[a] >> i) & 1)
        for (int j = 0; j < Ns; ++j)
          if ((mu[a] >> j) & 1) uu[(size_t)i * Ns + j] += hr[a];
  for (int b = 0; b < dimdw; ++b)
    for (int i = 0; i < Ns; ++i) {
      const double ai = ha[(size_t)b * 32 + i];
      for (int j = 0; j < Ns; ++j)
        if ((md[b] >> j) & 1) {
          ud[(size_t)i * Ns + j] += ai;
          if ((md[b] >> i) & 1) dd[(size_t)i * Ns + j] += hc[b];
        }
    }
  for (int j = 0; j < Ns; ++j)
    for (int i = 0; i < Ns; ++i) {
      const size_t ij = (size_t)i * Ns + j, ji = (size_t)j * Ns + i;
      n2[(size_t)j * Ns + i] = uu[ij] + dd[ij] + ud[ij] + ud[ji];
      sz2[(size_t)j * Ns + i] = 0.25 * (uu[ij] + dd[ij] - ud[ij] - ud[ji]);
    }
}

// ------------------------------------------------------------- energies
struct EnergyParams {
  InterParams I;                 // U masks, same-site pairs, Ust, Jx, Jp
  double hd_up[kMaxNs], hd_dw[kMaxNs];  // Re diag Hij(1,:,:), Hij(Nspin,:,:)
  double u_site[kMaxNs];         // Uloc of the orbital each site belongs to
  double Jh;
  int Ns, hfmode;
};

// Per-element terms of lanc_energy (ENERGY.f90:124-305), block partials
// part[8 * block + k] of: <K>, U/Ust/Jh/Jx/Jp energy, <Eloc>, <Ehf>, Dust,
// Dund, Dse, Dph.  The kinetic term gathers each row's hops (the hopping
// block is symmetric), the exchange and pair-hopping terms the moves that
// start at the element, as the reference's sums over i of
// state(i) * conjg(state(j)).
__global__ __launch_bounds__(256) void k_energy(const double* __restrict__ gs, const int32_t* __restrict__ map_up,
                                                const int32_t* __restrict__ map_dw, int dimup, int dimdw,
                                                int64_t dim, const int32_t* __restrict__ nbr_up,
                                                const double* __restrict__ cu, const int32_t* __restrict__ cnt_up,
                                                const int32_t* __restrict__ nbr_dw, const double* __restrict__ cd,
                                                const int32_t* __restrict__ cnt_dw, EnergyParams P,
                                                double* __restrict__ part) {
  __shared__ double sh[16];
  double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < dim; j += (int64_t)gridDim.x * blockDim.x) {
    const int iup = (int)(j % dimup), idw = (int)(j / dimup);
    const uint32_t su = (uint32_t)map_up[iup], sd = (uint32_t)map_dw[idw];
    const double v = gs[j], w = v * v;
    double hk = 0.0;
    for (int k = 0; k < cnt_up[iup]; ++k)
      hk += cu[(size_t)k * dimup + iup] * gs[(size_t)idw * dimup + nbr_up[(size_t)k * dimup + iup]];
    for (int k = 0; k < cnt_dw[idw]; ++k)
      hk += cd[(size_t)k * dimdw + idw] * gs[(size_t)nbr_dw[(size_t)k * dimdw + idw] * dimup + iup];
    a[0] += v * hk;
    double ek = 0.0, eh = 0.0;
    for (int io = 0; io < P.Ns; ++io) {
      const double nu = (double)((su >> io) & 1u), nd = (double)((sd >> io) & 1u);
      ek += P.hd_up[io] * nu + P.hd_dw[io] * nd;
      eh -= 0.5 * P.u_site[io] * (nu + nd);
    }
    double ep = inter_term(P.I, su, sd), dust = 0.0, dund = 0.0;
    for (int p = 0; p < P.I.npairs; ++p) {
      const uint32_t x = P.I.pi[p], y = P.I.pj[p];
      const double ux = (su >> x) & 1u, uy = (su >> y) & 1u, dx = (sd >> x) & 1u, dy = (sd >> y) & 1u;
      dust += ux * dy + uy * dx;
      const double same = ux * uy + dx * dy;
      dund += same;
      ep += (P.I.Ust - P.Jh) * same;
      eh -= (0.5 * P.I.Ust + 0.5 * (P.I.Ust - P.Jh)) * (ux + dx + uy + dy);
    }
    a[1] += w * ep;
    a[2] += w * ek;
    if (P.hfmode) a[3] += w * eh;
    a[4] += w * dust;
    a[5] += w * dund;
    if (P.I.npairs > 0) {
      a[6] += v * nonlocal_gather<double>(P.I, su, sd, gs, (size_t)dimup, 1.0, 0.0);
      a[7] += v * nonlocal_gather<double>(P.I, su, sd, gs, (size_t)dimup, 0.0, 1.0);
    }
  }
  for (int k = 0; k < 8; ++k) {
    const double s = block_sum(a[k], sh);
    if (threadIdx.x == 0) part[(size_t)blockIdx.x * 8 + k] = s;
    __syncthreads();
  }
}

void device_energy(Ctx& c, const Sector& s, const double* gs, double* e11) {
  const Model& M = c.model;
  EnergyParams P{};
  P.I = M.inter;
  P.Ns = M.Ns;
  P.Jh = M.Jh;
  P.hfmode = M.hfmode;
  for (int io = 0; io < M.Ns; ++io) {
    P.hd_up[io] = M.Hij[(size_t)M.Nspin * (io + (size_t)M.Ns * io)];
    P.hd_dw[io] = M.Hij[(size_t)(M.Nspin - 1) + (size_t)M.Nspin * (io + (size_t)M.Ns * io)];
  }
  for (int k = 0; k < M.Norb; ++k)
    for (int isite = 1; isite <= M.Nsites[k]; ++isite) {
      const int io = (k == 0) ? isite : isite + k * M.Nsites[k - 1];  // pack_indices
      if (io >= 1 && io <= M.Ns) P.u_site[io - 1] = M.Uloc[k];
    }
  const int nb = (int)std::min<int64_t>(1024, (s.dim + 255) / 256);
  double* part = nullptr;
  EDL_HIP(hipMalloc(&part, sizeof(double) * 8 * (size_t)nb));
  hipLaunchKernelGGL(k_energy, dim3(nb), dim3(256), 0, c.stream, gs, s.sp[0].map, s.sp[1].map, s.sp[0].dim,
                     s.sp[1].dim, s.dim, s.sp[0].nbr, s.sp[0].coef64, s.sp[0].cnt, s.sp[1].nbr, s.sp[1].coef64,
                     s.sp[1].cnt, P, part);
  EDL_HIP(hipGetLastError());
  std::vector<double> h((size_t)nb * 8);
  EDL_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c.stream));
  EDL_HIP(hipStreamSynchronize(c.stream));
  EDL_HIP(hipFree(part));
  double t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int b = 0; b < nb; ++b)
    for (int k = 0; k < 8; ++k) t[k] += h[(size_t)b * 8 + k];
  const double ekin = t[0], eknot = t[2], ehf = t[3], dse = t[6], dph = t[7];
  const double epot = t[1] + M.inter.Jx * dse + M.inter.Jp * dph + ehf;  // ed_Epot + ed_Ehartree
  const double out[11] = {ekin, epot, epot - ehf, eknot, ehf, t[4], t[5], dse, dph, 0.0, 0.0};
  for (int k = 0; k < 11; ++k) e11[k] = out[k];
}

}  // namespace edl
