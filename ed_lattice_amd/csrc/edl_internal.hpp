// Internal declarations shared by the HIP translation units of libedlat_hip.
// The public C-ABI is include/edlat_gpu.h.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <vector>

namespace edl {

constexpr int kMaxNs = 31;      // spin strings are held in 32-bit words
constexpr int kMaxOrb = 8;
constexpr int kMaxPairs = 64;   // same-site inter-orbital pairs (Ust/Jh terms)
constexpr int kBlock = 256;     // threads per block of the streaming kernels

struct Error : std::exception {
  std::string msg;
  explicit Error(std::string m) : msg(std::move(m)) {}
  const char* what() const noexcept override { return msg.c_str(); }
};

#define EDL_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      throw ::edl::Error(std::string(#call) + ": " + hipGetErrorString(e_));          \
  } while (0)

#define EDL_CHECK(cond, text)                 \
  do {                                        \
    if (!(cond)) throw ::edl::Error(text);    \
  } while (0)

// Raises a kernel's dynamic-LDS limit on the current device, once per
// (device, kernel): the attribute belongs to the device, so a process with
// contexts on several GPUs sets it on each.
inline void set_lds_limit(const void* kernel, int bytes) {
  static std::mutex m;
  static std::set<std::pair<int, const void*>> done;
  int dev = 0;
  EDL_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(m);
  if (done.count({dev, kernel})) return;
  EDL_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
  done.insert({dev, kernel});
}

// Parameters of the diagonal that are evaluated per element: the density-
// density interaction between the up and the down string of one basis state.
struct InterParams {
  int norb;                  // number of orbital masks
  uint32_t orb_mask[kMaxOrb];  // sites of orbital k
  double U[kMaxOrb];         // Uloc(k)
  int npairs;                // same-site, different-orbital pairs (io, jo)
  uint8_t pi[kMaxPairs], pj[kMaxPairs];  // 0-based positions
  double Ust;
  // Non-local Kanamori block (HxV_non_local.f90), on when Norb > 1 and Jx or
  // Jp is non-zero (Jhflag, SETUP.f90:35-36): spin exchange and pair
  // hopping between the orbitals of a site.
  double Jx, Jp;
};

// Bank-scheduled hop table of the up pass: per block of 64 strings, slots
// [ns[b]][64] of packed words from word off[b] on, ordered so that in every
// slot the 32 lanes of each half-wave read LDS words in distinct banks (bank
// = W * nbr mod 32 for a staged row of W dwords; a bipartite edge colouring
// of lanes x banks, which needs no more slots than the larger of the
// busiest lane's and the busiest bank's hop count).  Slots a lane does not
// use hold cid 255 and the address of a bank that slot leaves free.
struct ColorTable {
  uint32_t* words = nullptr;
  int32_t* off = nullptr;      // [nblocks]
  int32_t* ns = nullptr;       // [nblocks]
  int64_t slots = 0;           // sum of ns (table size / 64 words)
  // The same slots two to a word (strings with one hopping magnitude h and
  // dim + 32 <= 2^15): slot 2p in the low half of pair p, 2p+1 in the high
  // half, each half the LDS row (15 bits) and the sign of h (bit 15).
  // Unused slots read one of 32 zero rows staged after the last string (the
  // one in the bank the slot leaves free).  Null where not built.
  uint32_t* words16 = nullptr;
  int32_t* off16 = nullptr;    // [nblocks] first word of block b
  int32_t* ns16 = nullptr;     // [nblocks] pairs of block b (ceil(ns / 2))
  // Two hops to a word, split by sign (built with words16): the hops of a
  // block that take +h and those that take -h are bank-scheduled
  // separately, each set's slots paired into words (halves = LDS rows, no
  // sign bits), the +h words first, so every lane of a word has its sign.
  uint32_t* words16s = nullptr;
  int32_t* off16s = nullptr;   // [nblocks] first word of block b
  int32_t* ns16s = nullptr;    // [nblocks] words of block b
  int32_t* np16s = nullptr;    // [nblocks] its leading words that take +h
};

// Column ranges of the down spin for the range-blocked down pass: the
// strings sorted ascending share the occupation of their top b sites over
// contiguous index ranges lo[g] .. lo[g+1]-1; the hops of column j are
// regrouped as its hops inside j's range first (word = local index in the
// range | cid << 24, nloc[j] of them) then the others (global index),
// words[off[j] .. off[j+1]).
struct RangeTable {
  int n = 0;                   // ranges
  int maxcols = 0;             // columns of the largest range
  int maxwords = 0;            // hop words of the largest range
  int32_t* lo = nullptr;       // [n+1]
  uint32_t* words = nullptr;   // [total hops]
  int32_t* off = nullptr;      // [dimdw+1]
  int32_t* nloc = nullptr;     // [dimdw]
};
constexpr int kMaxRangeBits = 8;

// Hops by rank arithmetic.  For strings of popcount N the rank splits at
// bit B:  rank(s) = rlo[s & (2^B - 1)] + rhi[s >> B],  with
//   rlo[L] = sum over the j-th set bit p of L   of C(p, j),
//   rhi[H] = sum over the j-th set bit p of H   of C(B + p, N - popcount(H) + j),
// so the target of a hop across bond {p, q} is two table reads from the
// string itself, with no hop table.  Built for spins of at most 20 sites
// (tables of at most 1024 words each), at most kMaxBonds bonds and a
// symmetric hopping matrix (the same h gathered in both directions).
constexpr int kMaxBonds = 48;
constexpr int kRankHalfWords = 1024;
struct RankHops {
  bool ok = false;
  int B = 0, nlo = 0, nhi = 0;   // split bit, table lengths 2^B and 2^(Ns-B)
  int nb = 0;                    // bonds p < q with h = Hij(p,q) = Hij(q,p) != 0
  uint8_t p[kMaxBonds] = {}, q[kMaxBonds] = {};
  double h[kMaxBonds] = {};
  uint32_t* rlo = nullptr;       // device tables
  uint32_t* rhi = nullptr;
};

// One spin species of one sector: sorted strings, hopping table, diagonal.
struct SpinTable {
  int N = 0;           // electrons of this spin
  int dim = 0;         // C(Ns, N)
  int kmax = 0;        // ELL width of the hopping table
  int32_t* map = nullptr;      // [dim]   ascending strings
  int32_t* nbr = nullptr;      // [kmax][dim] neighbour index, 0-based
  double* coef64 = nullptr;    // [kmax][dim] Hij(io,jo) * fermionic sign
  float* coef32 = nullptr;     // same, rounded to fp32
  double* diag = nullptr;      // [dim] one-spin part of the diagonal
  // Packed table: [kmax][dim] words nbr | cid << 24, with the coefficient
  // lut[cid] = h * sign, h the v-th distinct nonzero Hij value of the spin
  // (cid = 2*v + (sign < 0), 255 = padding).  Available when dim < 2^24 and
  // there are at most 127 distinct hopping values.
  uint32_t* packed = nullptr;
  bool packed_ok = false;
  int nvals = 0;               // distinct hopping values (cids 0 .. 2*nvals-1)
  double lut64[256] = {};
  float lut32[256] = {};
  int32_t* cnt = nullptr;      // [dim] number of hops of each string
  // The packed words grouped by string, without padding: the hops of string
  // a are csr[off[a] .. off[a+1]) in table order (with the packed table).
  uint32_t* csr = nullptr;
  int32_t* off = nullptr;      // [dim+1]
  double* d_lut64 = nullptr;   // device copies of the coefficient lut
  float* d_lut32 = nullptr;
  // Bank-scheduled copies of the packed table for the up pass, one per LDS
  // element width W (dwords per staged row: 1 or 3), built with the sector.
  ColorTable pkc[4];
  // The packed table two hops to a word in table order, [pair][dim]: hops
  // 2p and 2p+1 of a string in the low / high half (LDS row in 15 bits, the
  // sign of h in bit 15; a missing second hop reads zero row dim).  Built
  // for spins with one hopping magnitude and dim + 32 <= 2^15, else null.
  uint32_t* pk16 = nullptr;
  // Column ranges by top-site count b (the down spin's, for the
  // range-blocked down pass), built with the sector for b = 1 .. 8 when
  // the ranges are at least a few columns wide.
  RangeTable rng[kMaxRangeBits + 1];
  RankHops rk;
  // The columns of the down spin in a locality order (Cuthill-McKee over the
  // hop graph: breadth first from a string of least degree, neighbours by
  // increasing degree), so that columns near in the order gather from
  // columns near in the order: ord[k] is the k-th column, and csro[offo[k] ..
  // offo[k+1]) its packed hop words (as csr).  Built with the sector for the
  // down spin (dim >= 256, with the packed table), else null.
  int32_t* ord = nullptr;
  uint32_t* csro = nullptr;
  int32_t* offo = nullptr;
};

struct Sector {
  int isector = 0, Nup = 0, Ndw = 0;
  int64_t dim = 0;
  SpinTable sp[2];
  // Row segments of the segmented up pass (columns larger than LDS): bounds
  // seg_h[0..n] of up index ranges that share their top sites' occupation,
  // each at most seg_cap rows; device copy d_seg.  Built on first use.
  std::vector<int> seg_h;
  int* d_seg = nullptr;
  int seg_cap = 0;
  int seg_sh = 0;  // the segments are unions of runs of equal map >> seg_sh
  // The model has the non-local Kanamori block: every operator and chain of
  // this sector runs on the single-pass kernels, which carry that term.
  bool nonlocal = false;
};

struct Model {
  bool set = false;
  int Ns = 0, Nspin = 1, Norb = 1;
  std::vector<int> Nsites;
  std::vector<double> Hij;   // real Hij(Nspin,Ns,Ns), column-major
  std::vector<double> Uloc;
  double Ust = 0, Jh = 0, Jx = 0, Jp = 0, xmu = 0;
  int hfmode = 1, ed_filling = 0;
  InterParams inter{};
};

struct Ctx;

// sector.hip
void build_sector_tables(Ctx& c, Sector& s);
void free_sector(Sector& s);
void rank_strings(Ctx& c, const SpinTable& t, int Ns, int64_t n, const int32_t* d_states, int32_t* d_out);
void init_binomials();

// The launchers of hxv.hip / hxv2.hip / seeds.hip are declared in chains.hpp,
// hxv2.hpp and seeds.hpp.

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string last_error;
  Model model;
  std::map<int, std::unique_ptr<Sector>> sectors;
  int num_cu = 256;
  // ground state kept on the device after edl_lanczos_gs / edl_set_state
  int gs_sector = 0;
  double gs_energy = 0.0;
  double* d_gs = nullptr;   // fp64 [dim]
  int64_t gs_dim = 0;
  // scratch owned by the context, grown on demand
  void* d_scratch = nullptr;
  size_t scratch_bytes = 0;
  // timing of the last operator / Lanczos call (HIP events on `stream`)
  double last_ms = 0.0;
  int64_t last_hxv = 0;
  // buffers handed out by edl_device_alloc
  std::vector<void*> user_bufs;
};

Sector& get_sector(Ctx& c, int isector);

// noalloc.hip: the file-scope device scratch of edl_hxv_noalloc
void noalloc_scratch_setup(int device);  // with `device` current, once per device
void* noalloc_scratch(int device);       // that device's address of the array
size_t noalloc_scratch_bytes();

}  // namespace edl
