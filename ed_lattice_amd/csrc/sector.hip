// Fock-sector construction on the device.
//
// A spin species with N electrons on Ns levels is the ascending list of
// Ns-bit strings with popcount N (SECTOR.f90:94-109 builds it by scanning
// 0..2^Ns-1).  Its position in that list is the combinatorial number
// system rank  r(s) = sum_k C(p_k, k),  p_1 < p_2 < ... the set bits
// (0-based), k = 1..N; unranking inverts it greedily.  Both are exact
// integer maps, so the device lists are bit-identical to the reference's
// maps and r(s)+1 equals binary_search(map, s) (AUX_FUNX.f90:163-180).
//
// For each string the gather-form hopping table lists the strings that
// reach it: output string s receives Hij(p,q) * sign from source
// s ^ (1<<p) ^ (1<<q) whenever bit p of s is set and bit q clear
// (the scatter c_q then c^dg_p of HxV_up.f90:13-27 read from the target's
// side).  sign = (-1)^popcount(s & bits strictly between p and q), which is
// the product sg1*sg2 of c/cdg (AUX_FUNX.f90:76-102).
#include "edl_internal.hpp"

namespace edl {

__constant__ uint32_t c_binom[33][33];

void init_binomials() {
  static bool done = false;
  if (done) return;
  uint32_t h[33][33] = {};
  for (int n = 0; n <= 32; n++) {
    h[n][0] = 1;
    for (int k = 1; k <= n; k++) h[n][k] = h[n - 1][k - 1] + (k <= n - 1 ? h[n - 1][k] : 0);
  }
  EDL_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_binom), h, sizeof(h)));
  done = true;
}

__device__ __forceinline__ uint32_t binom(int n, int k) { return (k < 0 || k > n) ? 0u : c_binom[n][k]; }

__device__ __forceinline__ int32_t rank_of(uint32_t s) {
  uint32_t r = 0;
  int k = 1;
  while (s) {
    int p = __ffs(s) - 1;
    r += binom(p, k);
    ++k;
    s &= s - 1;
  }
  return (int32_t)r;
}

__global__ void k_unrank(int32_t* map, int dim, int N, int Ns) {
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= dim) return;
  uint32_t r = (uint32_t)a, s = 0;
  int hi = Ns - 1;
  for (int k = N; k >= 1; --k) {
    int p = k - 1;
    while (p + 1 <= hi && binom(p + 1, k) <= r) ++p;
    s |= 1u << p;
    r -= binom(p, k);
    hi = p - 1;
  }
  map[a] = (int32_t)s;
}

struct HopPair {
  uint8_t p, q;
  uint8_t v;  // index of h among the spin's distinct hopping values
  double h;
};

__global__ void k_hop_count(const int32_t* map, int dim, const HopPair* pairs, int npairs, int32_t* cnt) {
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= dim) return;
  uint32_t s = (uint32_t)map[a];
  int c = 0;
  for (int k = 0; k < npairs; ++k) {
    uint32_t p = pairs[k].p, q = pairs[k].q;
    c += ((s >> p) & 1u) & (~(s >> q) & 1u);
  }
  cnt[a] = c;
}

// Packed words of string a, k < cnt[a], to csr[off[a] + k].
__global__ void k_ell_to_csr(const uint32_t* packed, const int32_t* cnt, const int32_t* off, int dim,
                             uint32_t* csr) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= dim) return;
  const int n = cnt[a], o = off[a];
  for (int k = 0; k < n; ++k) csr[o + k] = packed[(size_t)k * dim + a];
}

__global__ void k_hop_fill(const int32_t* map, int dim, const HopPair* pairs, int npairs, int kmax,
                           int32_t* nbr, double* coef64, float* coef32, uint32_t* packed) {
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= dim) return;
  uint32_t s = (uint32_t)map[a];
  int c = 0;
  for (int k = 0; k < npairs; ++k) {
    uint32_t p = pairs[k].p, q = pairs[k].q;
    if (((s >> p) & 1u) && !((s >> q) & 1u)) {
      uint32_t src = s ^ (1u << p) ^ (1u << q);
      uint32_t lo = p < q ? p : q, hi = p < q ? q : p;
      uint32_t between = ((1u << hi) - 1u) & ~((2u << lo) - 1u);
      double sg = (__popc(s & between) & 1) ? -1.0 : 1.0;
      double v = pairs[k].h * sg;
      size_t o = (size_t)c * dim + a;
      const int32_t r = rank_of(src);
      nbr[o] = r;
      coef64[o] = v;
      coef32[o] = (float)v;
      if (packed) packed[o] = ((uint32_t)r & 0xFFFFFFu) | ((uint32_t)(2 * pairs[k].v + (sg < 0.0 ? 1 : 0)) << 24);
      ++c;
    }
  }
  for (; c < kmax; ++c) {  // padding: self index, zero coefficient
    size_t o = (size_t)c * dim + a;
    nbr[o] = a;
    coef64[o] = 0.0;
    coef32[o] = 0.0f;
    if (packed) packed[o] = ((uint32_t)a & 0xFFFFFFu) | (255u << 24);
  }
}

// One-spin part of the diagonal of HxV_local.f90 for each string.
struct DiagParams {
  int Ns;
  double hdiag[kMaxNs];     // Re Hij(spin,io,io)
  double shift[kMaxNs];     // -xmu - hfmode*U/2 - hfmode*(2Ust-Jh)/2*(#pairs of io), if ed_filling==0
  int npairs;
  uint8_t pi[kMaxPairs], pj[kMaxPairs];
  double ust_jh;            // (Ust - Jh) same-spin pair term
};

__global__ void k_diag(const int32_t* map, int dim, DiagParams P, double* dg) {
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= dim) return;
  uint32_t s = (uint32_t)map[a];
  double h = 0.0;
  for (int io = 0; io < P.Ns; ++io)
    if ((s >> io) & 1u) h += P.hdiag[io] + P.shift[io];
  for (int k = 0; k < P.npairs; ++k) h += P.ust_jh * (double)(((s >> P.pi[k]) & (s >> P.pj[k])) & 1u);
  dg[a] = h;
}

__global__ void k_rank(const int32_t* states, int64_t n, int32_t* out, int Ns, int N, int dim) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  uint32_t s = (uint32_t)states[t];
  // 1-based position, 0 if not a string of this species (binary_search miss)
  bool ok = (Ns >= 32 || (s >> Ns) == 0u) && __popc(s) == N;
  int32_t r = ok ? rank_of(s) + 1 : 0;
  out[t] = (ok && r <= dim) ? r : 0;
}

static int blocks_for(int64_t n) { return (int)((n + kBlock - 1) / kBlock); }

// Split rank tables and bond list of one spin (RankHops), where its hopping
// matrix is symmetric with at most kMaxBonds bonds on at most 20 sites.  The
// tables are checked against the ascending enumeration of all C(Ns, N)
// strings (Gosper's successor), which is the order of the sector's map.
static void build_rank_hops(SpinTable& t, int Ns, const std::vector<HopPair>& pairs, hipStream_t st) {
  RankHops& r = t.rk;
  r = RankHops{};
  if (Ns > 20 || t.N < 1 || t.N >= Ns) return;
  for (const HopPair& a : pairs) {
    bool twin = false;
    for (const HopPair& b : pairs) twin = twin || (b.p == a.q && b.q == a.p && b.h == a.h);
    if (!twin) return;
    if (a.p < a.q) {
      if (r.nb == kMaxBonds) return;
      r.p[r.nb] = a.p;
      r.q[r.nb] = a.q;
      r.h[r.nb] = a.h;
      ++r.nb;
    }
  }
  r.B = Ns / 2;
  r.nlo = 1 << r.B;
  r.nhi = 1 << (Ns - r.B);
  uint64_t C[33][33] = {};
  for (int n = 0; n <= 32; ++n) {
    C[n][0] = 1;
    for (int k = 1; k <= n; ++k) C[n][k] = C[n - 1][k - 1] + (k <= n - 1 ? C[n - 1][k] : 0);
  }
  std::vector<uint32_t> lo(r.nlo), hi(r.nhi);
  for (int L = 0; L < r.nlo; ++L) {
    uint64_t v = 0;
    int j = 0;
    for (int p = 0; p < r.B; ++p)
      if ((L >> p) & 1) v += C[p][++j];
    lo[L] = (uint32_t)v;
  }
  for (int H = 0; H < r.nhi; ++H) {
    const int nl = t.N - __builtin_popcount((unsigned)H);
    uint64_t v = 0;
    if (nl >= 0) {
      int j = 0;
      for (int p = 0; p < Ns - r.B; ++p)
        if ((H >> p) & 1) {
          ++j;
          v += C[r.B + p][nl + j];
        }
    }
    hi[H] = (uint32_t)v;
  }
  const uint32_t lomask = (uint32_t)r.nlo - 1u;
  uint32_t s = (1u << t.N) - 1u;
  for (int a = 0; a < t.dim; ++a) {
    EDL_CHECK(lo[s & lomask] + hi[s >> r.B] == (uint32_t)a, "build_rank_hops: split rank tables disagree");
    const uint32_t c = s & (0u - s), rr = s + c;  // next string of the same popcount
    s = (((rr ^ s) >> 2) / c) | rr;
  }
  EDL_HIP(hipMalloc(&r.rlo, sizeof(uint32_t) * r.nlo));
  EDL_HIP(hipMalloc(&r.rhi, sizeof(uint32_t) * r.nhi));
  EDL_HIP(hipMemcpyAsync(r.rlo, lo.data(), sizeof(uint32_t) * r.nlo, hipMemcpyHostToDevice, st));
  EDL_HIP(hipMemcpyAsync(r.rhi, hi.data(), sizeof(uint32_t) * r.nhi, hipMemcpyHostToDevice, st));
  EDL_HIP(hipStreamSynchronize(st));  // lo, hi (host) must outlive the copies
  r.ok = true;
}

static void build_spin(Ctx& c, const Model& M, SpinTable& t, int spin /*0 up, 1 dw*/) {
  const int Ns = M.Ns;
  hipStream_t st = c.stream;
  // binomial as the reference computes it (SECTOR.f90:657-675)
  double xh = 1.0;
  for (int i = 1; i <= t.N; i++) xh = xh * (double)(Ns + 1 - i) / (double)i;
  t.dim = (t.N == 0) ? 1 : (int)(xh + 0.5);
  EDL_HIP(hipMalloc(&t.map, sizeof(int32_t) * t.dim));
  hipLaunchKernelGGL(k_unrank, dim3(blocks_for(t.dim)), dim3(kBlock), 0, st, t.map, t.dim, t.N, Ns);
  EDL_HIP(hipGetLastError());

  // hopping pairs of this spin: Hij(Nspin,...) is used for the down spin
  // (HxV_dw.f90:15), Hij(1,...) for the up spin (HxV_up.f90:16)
  const int s_idx = (spin == 0) ? 0 : (M.Nspin - 1);
  std::vector<HopPair> pairs;
  std::vector<double> vals;  // distinct hopping values, in order of appearance
  for (int p = 0; p < Ns; ++p)
    for (int q = 0; q < Ns; ++q) {
      if (p == q) continue;
      double h = M.Hij[(size_t)s_idx + (size_t)M.Nspin * (p + (size_t)Ns * q)];
      if (h == 0.0) continue;
      size_t v = 0;
      while (v < vals.size() && vals[v] != h) ++v;
      if (v == vals.size()) vals.push_back(h);
      pairs.push_back(HopPair{(uint8_t)p, (uint8_t)q, (uint8_t)(v < 255 ? v : 255), h});
    }
  t.nvals = (int)vals.size();
  if (spin == 0) build_rank_hops(t, Ns, pairs, st);  // read by the segmented up pass only
  HopPair* d_pairs = nullptr;
  int32_t* d_cnt = nullptr;
  const int np = (int)pairs.size();
  EDL_HIP(hipMalloc(&d_pairs, sizeof(HopPair) * (np > 0 ? np : 1)));
  if (np) EDL_HIP(hipMemcpyAsync(d_pairs, pairs.data(), sizeof(HopPair) * np, hipMemcpyHostToDevice, st));
  EDL_HIP(hipMalloc(&d_cnt, sizeof(int32_t) * t.dim));
  hipLaunchKernelGGL(k_hop_count, dim3(blocks_for(t.dim)), dim3(kBlock), 0, st, t.map, t.dim, d_pairs, np, d_cnt);
  std::vector<int32_t> cnt(t.dim);
  EDL_HIP(hipMemcpyAsync(cnt.data(), d_cnt, sizeof(int32_t) * t.dim, hipMemcpyDeviceToHost, st));
  EDL_HIP(hipStreamSynchronize(st));
  int kmax = 0;
  for (int v : cnt) kmax = v > kmax ? v : kmax;
  t.kmax = kmax;
  size_t n = (size_t)(kmax > 0 ? kmax : 1) * t.dim;
  EDL_HIP(hipMalloc(&t.nbr, sizeof(int32_t) * n));
  EDL_HIP(hipMalloc(&t.coef64, sizeof(double) * n));
  EDL_HIP(hipMalloc(&t.coef32, sizeof(float) * n));
  t.packed_ok = (t.nvals <= 127) && (t.dim < (1 << 24));
  if (t.packed_ok) {
    EDL_HIP(hipMalloc(&t.packed, sizeof(uint32_t) * n));
    for (int k = 0; k < 256; ++k) t.lut64[k] = 0.0;
    for (int v = 0; v < t.nvals; ++v) {
      t.lut64[2 * v] = vals[v];
      t.lut64[2 * v + 1] = -vals[v];
    }
    for (int k = 0; k < 256; ++k) t.lut32[k] = (float)t.lut64[k];
    EDL_HIP(hipMalloc(&t.d_lut64, sizeof(t.lut64)));
    EDL_HIP(hipMalloc(&t.d_lut32, sizeof(t.lut32)));
    EDL_HIP(hipMemcpyAsync(t.d_lut64, t.lut64, sizeof(t.lut64), hipMemcpyHostToDevice, st));
    EDL_HIP(hipMemcpyAsync(t.d_lut32, t.lut32, sizeof(t.lut32), hipMemcpyHostToDevice, st));
  }
  hipLaunchKernelGGL(k_hop_fill, dim3(blocks_for(t.dim)), dim3(kBlock), 0, st, t.map, t.dim, d_pairs, np, kmax,
                     t.nbr, t.coef64, t.coef32, t.packed);
  EDL_HIP(hipGetLastError());
  if (t.packed_ok) {
    std::vector<int32_t> off((size_t)t.dim + 1, 0);
    int64_t tot = 0;
    for (int a = 0; a < t.dim; ++a) {
      off[a] = (int32_t)tot;
      tot += cnt[a];
    }
    EDL_CHECK(tot < ((int64_t)1 << 31), "build_sector: hop table exceeds 2^31 words");
    off[t.dim] = (int32_t)tot;
    EDL_HIP(hipMalloc(&t.off, sizeof(int32_t) * ((size_t)t.dim + 1)));
    EDL_HIP(hipMalloc(&t.csr, sizeof(uint32_t) * (size_t)(tot > 0 ? tot : 1)));
    EDL_HIP(hipMemcpyAsync(t.off, off.data(), sizeof(int32_t) * off.size(), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_ell_to_csr, dim3(blocks_for(t.dim)), dim3(kBlock), 0, st, t.packed, d_cnt, t.off, t.dim,
                       t.csr);
    EDL_HIP(hipGetLastError());
    EDL_HIP(hipStreamSynchronize(st));  // off (host) must outlive the copy
  }

  // one-spin diagonal (HxV_local.f90:16-18, 47-60, 63-91)
  DiagParams P{};
  P.Ns = Ns;
  bool anyU = false;
  for (double u : M.Uloc) anyU = anyU || (u != 0.0);
  for (int io = 0; io < Ns; ++io) {
    P.hdiag[io] = M.Hij[(size_t)s_idx + (size_t)M.Nspin * (io + (size_t)Ns * io)];
    P.shift[io] = 0.0;
  }
  if (M.ed_filling == 0) {
    for (int io = 0; io < Ns; ++io) P.shift[io] -= M.xmu;
    if (M.hfmode) {
      if (anyU) {
        // -0.5*Uloc(iorb)*n(io) for io = pack_indices(isite,iorb)
        for (int iorb = 0; iorb < M.Norb; ++iorb)
          for (int isite = 1; isite <= M.Nsites[iorb]; ++isite) {
            int io = (iorb == 0) ? isite : isite + iorb * M.Nsites[iorb - 1];
            if (io >= 1 && io <= Ns) P.shift[io - 1] -= 0.5 * M.Uloc[iorb];
          }
      }
      if (M.Norb > 1)
        for (int k = 0; k < M.inter.npairs; ++k) {
          double d = 0.5 * M.Ust + 0.5 * (M.Ust - M.Jh);
          P.shift[M.inter.pi[k]] -= d;
          P.shift[M.inter.pj[k]] -= d;
        }
    }
  }
  P.npairs = M.inter.npairs;
  for (int k = 0; k < P.npairs; ++k) {
    P.pi[k] = M.inter.pi[k];
    P.pj[k] = M.inter.pj[k];
  }
  P.ust_jh = M.Ust - M.Jh;
  EDL_HIP(hipMalloc(&t.diag, sizeof(double) * t.dim));
  hipLaunchKernelGGL(k_diag, dim3(blocks_for(t.dim)), dim3(kBlock), 0, st, t.map, t.dim, P, t.diag);
  EDL_HIP(hipGetLastError());
  EDL_HIP(hipStreamSynchronize(st));
  EDL_HIP(hipFree(d_pairs));
  t.cnt = d_cnt;  // kept: exact hop count per string
}

void build_sector_tables(Ctx& c, Sector& s) {
  init_binomials();
  s.sp[0].N = s.Nup;
  s.sp[1].N = s.Ndw;
  build_spin(c, c.model, s.sp[0], 0);
  build_spin(c, c.model, s.sp[1], 1);
  s.dim = (int64_t)s.sp[0].dim * (int64_t)s.sp[1].dim;
  s.nonlocal = c.model.inter.Jx != 0.0 || c.model.inter.Jp != 0.0;
}

void free_sector(Sector& s) {
  if (s.d_seg) (void)hipFree(s.d_seg);
  s.d_seg = nullptr;
  s.seg_h.clear();
  s.seg_cap = 0;
  for (auto& t : s.sp) {
    if (t.map) (void)hipFree(t.map);
    if (t.nbr) (void)hipFree(t.nbr);
    if (t.coef64) (void)hipFree(t.coef64);
    if (t.coef32) (void)hipFree(t.coef32);
    if (t.diag) (void)hipFree(t.diag);
    if (t.packed) (void)hipFree(t.packed);
    if (t.cnt) (void)hipFree(t.cnt);
    if (t.csr) (void)hipFree(t.csr);
    if (t.off) (void)hipFree(t.off);
    if (t.ord) (void)hipFree(t.ord);
    if (t.csro) (void)hipFree(t.csro);
    if (t.offo) (void)hipFree(t.offo);
    if (t.d_lut64) (void)hipFree(t.d_lut64);
    if (t.d_lut32) (void)hipFree(t.d_lut32);
    for (auto& ct : t.pkc) {
      if (ct.words) (void)hipFree(ct.words);
      if (ct.off) (void)hipFree(ct.off);
      if (ct.ns) (void)hipFree(ct.ns);
      if (ct.words16) (void)hipFree(ct.words16);
      if (ct.off16) (void)hipFree(ct.off16);
      if (ct.ns16) (void)hipFree(ct.ns16);
      if (ct.words16s) (void)hipFree(ct.words16s);
      if (ct.off16s) (void)hipFree(ct.off16s);
      if (ct.ns16s) (void)hipFree(ct.ns16s);
      if (ct.np16s) (void)hipFree(ct.np16s);
    }
    if (t.pk16) (void)hipFree(t.pk16);
    if (t.rk.rlo) (void)hipFree(t.rk.rlo);
    if (t.rk.rhi) (void)hipFree(t.rk.rhi);
    for (auto& rt : t.rng) {
      if (rt.lo) (void)hipFree(rt.lo);
      if (rt.words) (void)hipFree(rt.words);
      if (rt.off) (void)hipFree(rt.off);
      if (rt.nloc) (void)hipFree(rt.nloc);
    }
    t = SpinTable{};
  }
}

void rank_strings(Ctx& c, const SpinTable& t, int Ns, int64_t n, const int32_t* d_states, int32_t* d_out) {
  init_binomials();
  if (n <= 0) return;
  hipLaunchKernelGGL(k_rank, dim3(blocks_for(n)), dim3(kBlock), 0, c.stream, d_states, n, d_out, Ns, t.N, t.dim);
  EDL_HIP(hipGetLastError());
}

}  // namespace edl
