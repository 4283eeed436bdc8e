// C-ABI of libedlat_hip (include/edlat_gpu.h): context, Hamiltonian and
// sector set-up, the operator, the Lanczos drivers and the batched seeds.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <numeric>
#include <tuple>

#include "../../include/edlat_gpu.h"
#include "batched2.hpp"
#include "chains.hpp"
#include "hxv2.hpp"
#include "seeds.hpp"
#include "tridiag.hpp"

struct edl_ctx {
  edl::Ctx c;
};

namespace edl {

InterParams& inter_of(Ctx& c) { return c.model.inter; }

Sector& get_sector(Ctx& c, int isector) {
  auto it = c.sectors.find(isector);
  if (it != c.sectors.end()) return *it->second;
  EDL_CHECK(c.model.set, "Hamiltonian not set (call edl_set_hamiltonian first)");
  const int Ns = c.model.Ns;
  EDL_CHECK(isector >= 1 && isector <= (Ns + 1) * (Ns + 1), "sector index out of range");
  auto s = std::make_unique<Sector>();
  s->isector = isector;
  // get_Nup / get_Ndw (SECTOR.f90:401-424)
  s->Ndw = (isector - 1) % (Ns + 1);
  s->Nup = ((isector - 1) / (Ns + 1)) % (Ns + 1);
  build_sector_tables(c, *s);
  try {
    build_up_color_tables(*s, c.stream);
    build_dw_range_tables(*s, c.stream);
    build_dw_order(*s, c.stream);
  } catch (...) {
    free_sector(*s);
    throw;
  }
  Sector& ref = *s;
  c.sectors[isector] = std::move(s);
  return ref;
}

static void release_sectors(Ctx& c) {
  for (auto& kv : c.sectors) free_sector(*kv.second);
  c.sectors.clear();
}

// ------------------------------------------------------------ utilities
struct DevBuf {
  void* p = nullptr;
  DevBuf() = default;
  explicit DevBuf(size_t bytes) { EDL_HIP(hipMalloc(&p, bytes ? bytes : 4)); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p) { o.p = nullptr; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      if (p) (void)hipFree(p);
      p = o.p;
      o.p = nullptr;
    }
    return *this;
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
  // An optional allocation: empty (and no error left pending) when the
  // device cannot provide it.
  static DevBuf try_alloc(size_t bytes) {
    DevBuf b;
    if (hipMalloc(&b.p, bytes ? bytes : 4) != hipSuccess) {
      b.p = nullptr;
      (void)hipGetLastError();
    }
    return b;
  }
};

// Bump allocator over one device allocation: the buffers of a chain group
// cost one hipMalloc.
struct Arena {
  DevBuf buf;
  size_t off = 0, cap = 0;
  static size_t round(size_t b) { return (b + 255) & ~(size_t)255; }
  void reserve(size_t bytes) {
    buf = DevBuf(bytes);
    cap = bytes;
    off = 0;
  }
  void* take(size_t bytes) {
    EDL_CHECK(off + round(bytes) <= cap, "arena: out of reserved memory");
    void* p = static_cast<char*>(buf.p) + off;
    off += round(bytes);
    return p;
  }
};

struct EventTimer {
  hipEvent_t a{}, b{};
  hipStream_t st;
  explicit EventTimer(hipStream_t s) : st(s) {
    EDL_HIP(hipEventCreate(&a));
    EDL_HIP(hipEventCreate(&b));
    EDL_HIP(hipEventRecord(a, st));
  }
  double stop() {
    EDL_HIP(hipEventRecord(b, st));
    EDL_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    EDL_HIP(hipEventElapsedTime(&ms, a, b));
    return ms;
  }
  ~EventTimer() {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  }
};

template <typename T>
double device_sumsq(const T* x, int64_t n, hipStream_t st);

// Context-owned scratch (the intermediate of the two-pass operator).  It is
// only grown between launches, never while a graph that uses it is alive.
static void* scratch(Ctx& c, size_t bytes) {
  if (c.scratch_bytes < bytes) {
    EDL_HIP(hipStreamSynchronize(c.stream));
    if (c.d_scratch) EDL_HIP(hipFree(c.d_scratch));
    c.d_scratch = nullptr;
    c.scratch_bytes = 0;
    EDL_HIP(hipMalloc(&c.d_scratch, bytes));
    c.scratch_bytes = bytes;
  }
  return c.d_scratch;
}

template <typename T>
static const T* lut_dev(const SpinTable& t);
template <>
const double* lut_dev<double>(const SpinTable& t) { return t.d_lut64; }
template <>
const float* lut_dev<float>(const SpinTable& t) { return t.d_lut32; }

// EDL_VERBOSE=1: phase timings of the drivers on stderr.
static bool verbose() {
  const char* e = std::getenv("EDL_VERBOSE");
  return e && *e && *e != '0';
}

// Tuned two-pass schedule per (sector shape, precision); the chain shapes
// (the ch_* fields) also on their own, for pinned operator schedules.
static std::map<std::tuple<int, int, int, int, int>, TwoPassCfg>& tp_cache() {
  static std::map<std::tuple<int, int, int, int, int>, TwoPassCfg> m;
  return m;
}
static std::map<std::tuple<int, int, int, int, int>, TwoPassCfg>& tp_chain_cache() {
  static std::map<std::tuple<int, int, int, int, int>, TwoPassCfg> m;
  return m;
}
// dst takes the chain shapes of src.
static void with_chain_shapes(TwoPassCfg& dst, const TwoPassCfg& src) {
  dst.ch_R = src.ch_R;
  dst.ch_threads = src.ch_threads;
  dst.ch_nchunk = src.ch_nchunk;
  dst.ch_ncol = src.ch_ncol;
  dst.ch_kch = src.ch_kch;
  dst.ch_vec = src.ch_vec;
  dst.ch_C = src.ch_C;
  dst.ch_up_threads = src.ch_up_threads;
  dst.ch_grid_up = src.ch_grid_up;
  dst.ch_up_rpt = src.ch_up_rpt;
  dst.ch_up_kch = src.ch_up_kch;
  dst.ch_up_col = src.ch_up_col;
  dst.ch_rng = src.ch_rng;
  dst.ch_ord = src.ch_ord;
}

// Column-chunk counts the down-pass tuner tries: fixed ones, and for the
// vector-gather form (slots = columns a block takes per round, NCOL x waves)
// the counts that give every chunk one or two full rounds, so that no round
// runs with idle waves (at most kMaxFitChunks).
constexpr int kMaxFitChunks = 1024;
static std::vector<int> dw_chunk_counts(int dimdw, int slots, bool with32 = true) {
  std::vector<int> v = with32 ? std::vector<int>{32, 64, 128, 256} : std::vector<int>{64, 128, 256};
  if (slots > 0)
    for (int k : {1, 2}) {
      const int n = (dimdw + k * slots - 1) / (k * slots);
      if (n <= kMaxFitChunks && std::find(v.begin(), v.end(), n) == v.end()) v.push_back(n);
    }
  return v;
}

// Pick panel widths, block sizes and grids of the two passes by timing the
// operator on the sector itself (3 launches per candidate after a warm-up).
template <typename T>
static TwoPassCfg two_pass_cfg(Ctx& c, const Sector& s) {
  const auto key = std::make_tuple(s.sp[0].dim, s.sp[1].dim, s.sp[0].kmax, s.sp[1].kmax, (int)sizeof(T));
  // timed on the engine's own (padded) layout
  const int ld = padded_ld(s.sp[0].dim);
  const TwoPassView V = two_pass_view(s, ld);
  TwoPassCfg best;
  bool pinned = false;
  // EDL_TWO_PASS_CFG_F32 / _F64 =
  // "C,threads_up,grid_up,R,threads_dw,nchunk[,ncol,kch[,up_rpt,up_kch[,dw_vec[,up_col[,dw_rng[,dw_ord]]]]]]"
  // pins the operator schedule of that precision (reproducible profiles of
  // one configuration), ahead of any tuned schedule; the operator is not
  // tuned then, its chain shapes still are (once per sector shape).
  if (const char* e = std::getenv(sizeof(T) == 4 ? "EDL_TWO_PASS_CFG_F32" : "EDL_TWO_PASS_CFG_F64")) {
    TwoPassCfg p;
    const int n = std::sscanf(e, "%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d", &p.C, &p.threads_up, &p.grid_up,
                              &p.R, &p.threads_dw, &p.nchunk, &p.ncol, &p.kch, &p.up_rpt, &p.up_kch, &p.dw_vec,
                              &p.up_col, &p.dw_rng, &p.dw_ord);
    EDL_CHECK(p.dw_ord == 0 || (p.dw_ord == 1 && p.dw_vec == 1 && p.dw_rng == 0 && V.ord_dw),
              "EDL_TWO_PASS_CFG_F32/_F64: dw_ord 1 needs the vector-gather down pass and a sector with the column "
              "order (down spin of at least 256 strings)");
    EDL_CHECK(p.dw_rng == 0 || pass_dwr_ok(V, p.R, p.dw_rng, sizeof(T)),
              "EDL_TWO_PASS_CFG_F32/_F64: dw_rng needs range tables for that many top sites and R 32|64 within LDS");
    EDL_CHECK(p.up_col == 0 || up_col_ok(V, p.C, sizeof(T), p.up_col),
              "EDL_TWO_PASS_CFG_F32/_F64: up_col 1 needs fp32 panels of 1 or 3 columns on a sector with the "
              "bank-scheduled table, up_col 2 and 4 also one hopping magnitude and DimUp + 32 < 2^15 (fp64: "
              "one column), up_col 3 one hopping magnitude and DimUp + 32 < 2^15");
    EDL_CHECK(n >= 6, "EDL_TWO_PASS_CFG_F32/_F64: expected "
                      "C,threads_up,grid_up,R,threads_dw,nchunk[,ncol,kch[,up_rpt,up_kch[,dw_vec[,up_col[,dw_rng[,"
                      "dw_ord]]]]]]");
    EDL_CHECK((p.up_rpt == 1 && (p.up_kch == 8 || p.up_kch == 16)) || (p.up_rpt == 2 && (p.up_kch == 4 || p.up_kch == 8)),
              "EDL_TWO_PASS_CFG_F32/_F64: (up_rpt, up_kch) must be (1, 8|16) or (2, 4|8)");
    EDL_CHECK(p.dw_rng > 0 || pass_dw_lds_bytes(V, p) <= (size_t)(160 * 1024 - 4096),
              "EDL_TWO_PASS_CFG_F32/_F64: down-pass column table exceeds LDS");
    EDL_CHECK(p.dw_vec == 0 || p.dw_vec == 1, "EDL_TWO_PASS_CFG_F32/_F64: dw_vec must be 0 or 1");
    if (p.dw_vec)
      EDL_CHECK(((p.ncol == 1 && (p.kch == 2 || p.kch == 4)) || (p.ncol == 2 && (p.kch == 1 || p.kch == 2 || p.kch == 4))) &&
                    (p.R == 16 || p.R == 32 || p.R == 64 || p.R == 128 || (p.R == 256 && sizeof(T) == 4)),
                "EDL_TWO_PASS_CFG_F32/_F64: the vector-gather down pass takes R 16|32|64|128 (256 in fp32), "
                "(ncol, kch) in (1, 2|4) or (2, 1|2|4)");
    else
      EDL_CHECK((p.ncol == 1 && (p.kch == 8 || p.kch == 16 || p.kch == 32)) ||
                    (p.ncol == 2 && (p.kch == 8 || p.kch == 16)),
                "EDL_TWO_PASS_CFG_F32/_F64: (ncol, kch) must be (1, 8|16|32) or (2, 8|16)");
    EDL_CHECK(p.C >= 1 && p.C <= 4 && (p.R == 32 || p.R == 64 || p.dw_vec),
              "EDL_TWO_PASS_CFG_F32/_F64: C must be 1 to 4 and R 32 or 64 (or a vector-gather height)");
    EDL_CHECK(p.threads_up % 64 == 0 && p.threads_up >= 64 && p.threads_up <= 1024 && p.threads_dw % 64 == 0 &&
                  p.threads_dw >= 64 && p.threads_dw <= 1024 && p.grid_up >= 1 && p.nchunk >= 1 &&
                  p.nchunk <= V.dimdw,
              "EDL_TWO_PASS_CFG_F32/_F64: bad block size, grid or chunk count");
    EDL_CHECK(256 * sizeof(T) + 16 + (size_t)V.dimup * p.C * sizeof(T) <= (size_t)(160 * 1024),
              "EDL_TWO_PASS_CFG_F32/_F64: up-pass columns exceed LDS");
    pinned = true;
    best = p;
  } else {
    auto it = tp_cache().find(key);
    if (it != tp_cache().end()) return it->second;
  }
  // The chain shapes do not depend on the operator's: a pinned operator
  // schedule still runs its chains on their own tuned shapes.
  auto ct = tp_chain_cache().find(key);
  // A pinned range-blocked down pass (dw_rng > 0) or one in the locality
  // order of columns (dw_ord = 1) pins the chain's down pass to the same
  // form, so that its Lanczos terms are tested pinned too.
  auto pin_chain_range = [&](TwoPassCfg& b) {
    if (!pinned || (b.dw_rng == 0 && b.dw_ord == 0)) return;
    b.ch_R = b.R;
    b.ch_threads = b.threads_dw;
    b.ch_nchunk = b.nchunk;
    b.ch_ncol = b.ncol;
    b.ch_kch = b.kch;
    b.ch_vec = b.dw_rng > 0 ? 0 : b.dw_vec;
    b.ch_rng = b.dw_rng;
    b.ch_ord = b.dw_ord;
  };
  if (pinned && ct != tp_chain_cache().end()) {
    with_chain_shapes(best, ct->second);
    pin_chain_range(best);
    return best;
  }
  hipStream_t st = c.stream;
  const int64_t D = s.dim, Dp = (int64_t)ld * s.sp[1].dim;
  DevBuf d64(sizeof(double) * D), q(sizeof(T) * Dp), t(sizeof(T) * Dp), o(sizeof(T) * Dp);
  launch_start_vector(d64.as<double>(), D, 0x5EEDull, st);
  launch_zero_bytes(q.p, sizeof(T) * Dp, st);
  launch_convert_ld<T>(d64.as<double>(), q.as<T>(), s.sp[0].dim, s.sp[1].dim, ld, st);
  const int lds = 160 * 1024;
  if (!pinned) {
    best.grid_up = c.num_cu;
    best.nchunk = std::min(32, V.dimdw);
    double best_up = 1e300, best_dw = 1e300;
    for (int col : {0, 1, 2, 3, 4, 5, 6})
    for (int C : {1, 2, 3, 4})
      for (int rk : {208, 204, 108, 116})
      for (int th : {512, 1024})
        for (int gm : {1, 2}) {
          const size_t need = 256 * sizeof(T) + (size_t)(V.dimup + (col >= 2 ? 32 : 0)) * C * sizeof(T);
          if (need > (size_t)lds) continue;
          if (gm == 2 && 2 * need > (size_t)lds) continue;
          if (col && !up_col_ok(V, C, sizeof(T), col)) continue;
          if (col >= 5 && V.dimup > 13 * th) continue;  // prefetching forms: 13 rows per thread at most
          TwoPassCfg w = best;
          w.C = C;
          w.up_col = col;
          w.up_rpt = rk / 100;
          w.up_kch = rk % 100;
          w.threads_up = th;
          w.grid_up = c.num_cu * gm;
          launch_pass_up<T>(V, c.model.inter, lut_dev<T>(s.sp[0]), q.as<T>(), q.as<T>(), t.as<T>(), nullptr, w, st);
          EventTimer tm(st);
          for (int r = 0; r < 3; ++r)
            launch_pass_up<T>(V, c.model.inter, lut_dev<T>(s.sp[0]), q.as<T>(), q.as<T>(), t.as<T>(), nullptr, w, st);
          const double ms = tm.stop();
          if (verbose())
            std::fprintf(stderr, "[edl] tune up  fp%d C %d col %d rpt %d kch %2d threads %4d grid %5d: %.4f ms\n",
                         8 * (int)sizeof(T), C, col, w.up_rpt, w.up_kch, th, w.grid_up, ms / 3);
          if (ms < best_up) {
            best_up = ms;
            best.C = C;
            best.up_col = col;
            best.up_rpt = w.up_rpt;
            best.up_kch = w.up_kch;
            best.threads_up = th;
            best.grid_up = w.grid_up;
          }
        }
    // 4-byte (8-byte) gathers, then 16-byte vector gathers (columns in the
    // natural order, then in the locality order)
    for (int vo : {0, 1, 2})
    for (int R : {16, 32, 64, 128, 256})
      for (int nk : {108, 116, 132, 208, 216, 102, 104, 201, 202, 204})
        for (int th : {256, 512, 1024})
          for (int nch : dw_chunk_counts(V.dimdw, vo ? (nk / 100) * (th / 64) : 0)) {
          if (nch > V.dimdw) continue;
          const int nc = nk / 100, kc = nk % 100;
          const bool vk = kc <= 4;
          if (vk != (vo >= 1) || (!vo && (R == 16 || R > 64)) || (R == 256 && sizeof(T) != 4)) continue;
          if (vo == 2 && !V.ord_dw) continue;
          TwoPassCfg w = best;
          w.R = R;
          w.ncol = nc;
          w.kch = kc;
          w.threads_dw = th;
          w.nchunk = nch;
          w.dw_vec = vo >= 1;
          w.dw_ord = vo == 2;
          if (pass_dw_lds_bytes(V, w) > (size_t)lds - 4096) continue;
          launch_pass_dw<T>(V, lut_dev<T>(s.sp[1]), q.as<T>(), q.as<T>(), t.as<T>(), o.as<T>(), nullptr, w, st);
          EventTimer tm(st);
          for (int r = 0; r < 3; ++r)
            launch_pass_dw<T>(V, lut_dev<T>(s.sp[1]), q.as<T>(), q.as<T>(), t.as<T>(), o.as<T>(), nullptr, w, st);
          const double ms = tm.stop();
          if (verbose())
            std::fprintf(stderr, "[edl] tune dw  fp%d vec %d ord %d R %3d ncol %d kch %2d threads %4d chunks %3d: %.4f ms\n",
                         8 * (int)sizeof(T), w.dw_vec, w.dw_ord, R, nc, kc, th, nch, ms / 3);
          if (ms < best_dw) {
            best_dw = ms;
            best.R = R;
            best.ncol = nc;
            best.kch = kc;
            best.threads_dw = th;
            best.nchunk = nch;
            best.dw_vec = w.dw_vec;
            best.dw_ord = w.dw_ord;
            best.dw_rng = 0;
          }
        }
    // range-blocked form: R-row slabs of column ranges sharing b top sites
    for (int b = 1; b <= kMaxRangeBits; ++b)
      for (int R : {32, 64})
        for (int th : {256, 512, 1024}) {
          if (!pass_dwr_ok(V, R, b, sizeof(T))) continue;
          TwoPassCfg w = best;
          w.R = R;
          w.threads_dw = th;
          w.dw_rng = b;
          launch_pass_dw<T>(V, lut_dev<T>(s.sp[1]), q.as<T>(), q.as<T>(), t.as<T>(), o.as<T>(), nullptr, w, st);
          EventTimer tm(st);
          for (int r = 0; r < 3; ++r)
            launch_pass_dw<T>(V, lut_dev<T>(s.sp[1]), q.as<T>(), q.as<T>(), t.as<T>(), o.as<T>(), nullptr, w, st);
          const double ms = tm.stop();
          if (verbose())
            std::fprintf(stderr, "[edl] tune dw  fp%d ranges b %d (%d) R %3d threads %4d: %.4f ms\n",
                         8 * (int)sizeof(T), b, V.rng_n[b], R, th, ms / 3);
          if (ms < best_dw) {
            best_dw = ms;
            best.R = R;
            best.threads_dw = th;
            best.dw_rng = b;
            best.dw_ord = 0;
          }
        }
  }
  // The chain step's down pass (MODE 1) streams two more vectors and holds
  // more registers; its best vector-gather shape is timed on its own, with
  // the chain scalars in a middle-step state that repeated launches leave
  // unchanged (beta = 0, norms 1).
  {
    LzScalars s0{};
    s0.s_cur = 1.0;
    s0.s_old = 1.0;
    DevBuf scd(sizeof(LzScalars)), ab(sizeof(double) * 8);
    EDL_HIP(hipMemcpyAsync(scd.p, &s0, sizeof(s0), hipMemcpyHostToDevice, st));
    launch_zero_bytes(ab.p, sizeof(double) * 8, st);
    size_t maxpart = 1;
    for (int R : {64, 128})
      for (int nch : {64, 128, 256, kMaxFitChunks}) {
        TwoPassCfg w = best;
        w.R = R;
        w.nchunk = std::min(nch, V.dimdw);
        w.dw_rng = 0;
        maxpart = std::max(maxpart, (size_t)pass_dw_grid(V, w));
      }
    for (int b = 1; b <= kMaxRangeBits; ++b)
      for (int R : {32, 64}) {
        if (!pass_dwr_ok(V, R, b, sizeof(T))) continue;
        TwoPassCfg w = best;
        w.R = R;
        w.dw_rng = b;
        maxpart = std::max(maxpart, (size_t)pass_dw_grid(V, w));
      }
    DevBuf part(sizeof(double) * maxpart);
    LzChain A{};
    A.sc = scd.as<LzScalars>();
    A.alpha = ab.as<double>();
    A.beta = ab.as<double>() + 4;
    A.part_dw = part.as<double>();
    A.nmax = 4;
    double bc = 1e300;
    for (int ord : {0, 1})
    for (int R : {64, 128})
      for (int nk : {102, 104, 201, 202, 204})
        for (int th : {256, 512, 1024})
          for (int nch : dw_chunk_counts(V.dimdw, (nk / 100) * (th / 64), false)) {
            if (nch > V.dimdw || (ord && !V.ord_dw)) continue;
            TwoPassCfg w = best;
            w.R = R;
            w.ncol = nk / 100;
            w.kch = nk % 100;
            w.threads_dw = th;
            w.nchunk = nch;
            w.dw_vec = 1;
            w.dw_rng = 0;
            w.dw_ord = ord;
            if (pass_dw_lds_bytes(V, w) > (size_t)lds - 4096) continue;
            launch_pass_dw<T>(V, lut_dev<T>(s.sp[1]), q.as<T>(), o.as<T>(), t.as<T>(), nullptr, &A, w, st);
            EventTimer tm(st);
            for (int r = 0; r < 3; ++r)
              launch_pass_dw<T>(V, lut_dev<T>(s.sp[1]), q.as<T>(), o.as<T>(), t.as<T>(), nullptr, &A, w, st);
            const double ms = tm.stop();
            if (verbose())
              std::fprintf(stderr, "[edl] tune dw chain fp%d ord %d R %3d ncol %d kch %d threads %4d chunks %3d: %.4f ms\n",
                           8 * (int)sizeof(T), ord, R, w.ncol, w.kch, th, nch, ms / 3);
            if (ms < bc) {
              bc = ms;
              best.ch_R = R;
              best.ch_threads = th;
              best.ch_nchunk = nch;
              best.ch_ncol = w.ncol;
              best.ch_kch = w.kch;
              best.ch_vec = 1;
              best.ch_rng = 0;
              best.ch_ord = ord;
            }
          }
    for (int b = 1; b <= kMaxRangeBits; ++b)
      for (int R : {32, 64})
        for (int th : {256, 512, 1024}) {
          if (!pass_dwr_ok(V, R, b, sizeof(T))) continue;
          TwoPassCfg w = best;
          w.R = R;
          w.threads_dw = th;
          w.dw_rng = b;
          w.dw_vec = 0;
          launch_pass_dw<T>(V, lut_dev<T>(s.sp[1]), q.as<T>(), o.as<T>(), t.as<T>(), nullptr, &A, w, st);
          EventTimer tm(st);
          for (int r = 0; r < 3; ++r)
            launch_pass_dw<T>(V, lut_dev<T>(s.sp[1]), q.as<T>(), o.as<T>(), t.as<T>(), nullptr, &A, w, st);
          const double ms = tm.stop();
          if (verbose())
            std::fprintf(stderr, "[edl] tune dw chain fp%d ranges b %d R %3d threads %4d: %.4f ms\n",
                         8 * (int)sizeof(T), b, R, th, ms / 3);
          if (ms < bc) {
            bc = ms;
            best.ch_R = R;
            best.ch_threads = th;
            best.ch_rng = b;
            best.ch_vec = 0;
            best.ch_ord = 0;
          }
        }
    // The chain step's up pass with phase B pending and alpha = 0 (r = t,
    // written back over t: repeated launches leave their input unchanged).
    LzScalars s1 = s0;
    s1.pad = 1;
    EDL_HIP(hipMemcpyAsync(scd.p, &s1, sizeof(s1), hipMemcpyHostToDevice, st));
    DevBuf pu(sizeof(double) * (size_t)c.num_cu * 2);
    A.part_up = pu.as<double>();
    double bu2 = 1e300;
    for (int col : {0, 1, 2, 3, 4})
    for (int C : {1, 2, 3, 4})
      for (int rk : {208, 204, 108, 116})
        for (int th : {512, 1024})
          for (int gm : {1, 2}) {
            const size_t need = 256 * sizeof(T) + (size_t)(V.dimup + (col >= 2 ? 32 : 0)) * C * sizeof(T);
            if (need > (size_t)lds || (gm == 2 && 2 * need > (size_t)lds)) continue;
            if (col && !up_col_ok(V, C, sizeof(T), col)) continue;
            TwoPassCfg w = best;
            w.C = C;
            w.up_col = col;
            w.up_rpt = rk / 100;
            w.up_kch = rk % 100;
            w.threads_up = th;
            w.grid_up = c.num_cu * gm;
            // qa = q_k (q), qb = t (o): phase B reads both and rewrites t
            launch_pass_up<T>(V, c.model.inter, lut_dev<T>(s.sp[0]), q.as<T>(), o.as<T>(), t.as<T>(), &A, w, st);
            EventTimer tm(st);
            for (int r = 0; r < 3; ++r)
              launch_pass_up<T>(V, c.model.inter, lut_dev<T>(s.sp[0]), q.as<T>(), o.as<T>(), t.as<T>(), &A, w, st);
            const double ms = tm.stop();
            if (verbose())
              std::fprintf(stderr, "[edl] tune up chain fp%d C %d col %d rpt %d kch %2d threads %4d grid %5d: %.4f ms\n",
                           8 * (int)sizeof(T), C, col, w.up_rpt, w.up_kch, th, w.grid_up, ms / 3);
            if (ms < bu2) {
              bu2 = ms;
              best.ch_C = C;
              best.ch_up_col = col;
              best.ch_up_threads = th;
              best.ch_grid_up = w.grid_up;
              best.ch_up_rpt = w.up_rpt;
              best.ch_up_kch = w.up_kch;
            }
          }
  }
  tp_chain_cache()[key] = best;
  if (!pinned) tp_cache()[key] = best;
  pin_chain_range(best);
  return best;
}

// H v on one sector: two-pass kernels where the columns fit in LDS,
// otherwise the single-pass gather kernel.
// v and out have column stride ld (DimUp: natural layout; padded_ld: the
// engine's layout, two-pass sectors only).
// Schedule of a segmented sector: the up-pass panel width C (segments are
// re-cut for each C: wider panels amortise the hop words over more columns,
// narrower ones keep more hops inside a segment) and the down-pass shape,
// each timed on two launches of the sector itself -- a few dozen launches in
// all, against the hundreds of the two-pass search, since one launch here
// moves gigabytes.  The up pass is timed with the hop table and, where the
// sector has rank tables, with hop targets by rank arithmetic
// (EDL_UP_SEG_RANK = 0 or 1 pins one of the two; 1 where the sector has rank
// tables).  Cached per sector shape, precision, row cap and pin.
static int seg_rank_pin() {
  const char* e = std::getenv("EDL_UP_SEG_RANK");
  if (!e || !*e) return -1;
  const int v = std::atoi(e);
  EDL_CHECK(v == 0 || v == 1, "EDL_UP_SEG_RANK must be 0 or 1");
  return v;
}
// EDL_UP_SEG_COLS pins the up-pass panel width C of segmented sectors (1 to
// 4, or 8, which runs with fp32 rank arithmetic and as 4 otherwise); 0 when
// unset.
static int seg_cols_pin() {
  const char* e = std::getenv("EDL_UP_SEG_COLS");
  if (!e || !*e) return 0;
  const int v = std::atoi(e);
  EDL_CHECK((v >= 1 && v <= 4) || v == 8, "EDL_UP_SEG_COLS must be 1, 2, 3, 4 or 8");
  return v;
}

template <typename T>
static TwoPassCfg seg_cfg(Ctx& c, Sector& s) {
  static std::map<std::tuple<int, int, int, int, int, int, int, int>, TwoPassCfg> cache;
  const int pin = (seg_rank_pin() == 1 && !seg_rank_ok(s)) ? 0 : seg_rank_pin();
  const int cpin = seg_cols_pin();
  const auto key = std::make_tuple(s.sp[0].dim, s.sp[1].dim, s.sp[0].kmax, s.sp[1].kmax, (int)sizeof(T),
                                   seg_row_cap(sizeof(T), 1), pin, cpin);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  hipStream_t st = c.stream;
  const int ld = padded_ld(s.sp[0].dim);
  const TwoPassView V = two_pass_view(s, ld);
  const size_t bytes = sizeof(T) * (size_t)ld * s.sp[1].dim;
  DevBuf v(bytes), t(bytes), o(bytes);
  launch_zero_bytes(v.p, bytes, st);
  TwoPassCfg best = seg_dw_cfg(s.sp[1].dim);
  best.C = 1;
  best.seg_rank = pin == 1 ? 1 : 0;
  double bu = 1e300;
  for (int rank = 0; rank <= 1; ++rank) {
    if ((pin >= 0 && rank != pin) || (rank == 1 && !seg_rank_ok(s))) continue;
    for (int C : {1, 2, 3, 4, 8}) {
      if (C == 8 && (rank == 0 || sizeof(T) != 4)) continue;  // fp32 rank arithmetic only
      if (cpin && C != ((cpin == 8 && (rank == 0 || sizeof(T) != 4)) ? 4 : cpin)) continue;  // 8: else 4
      seg_bounds(s, sizeof(T), C);
      double ms = 0.0;
      for (int r = 0; r < 2; ++r) {
        EventTimer tm(st);
        launch_up_seg<T>(V, c.model.inter, lut_dev<T>(s.sp[0]), v.as<T>(), v.as<T>(), t.as<T>(), nullptr, s, C, rank,
                         st);
        ms = tm.stop();
      }
      if (verbose())
        std::fprintf(stderr, "[edl] segmented up pass C %d rank %d: %zu segments, %.3f ms\n", C, rank,
                     s.seg_h.size() - 1, ms);
      if (ms < bu) {
        bu = ms;
        best.C = C;
        best.seg_rank = rank;
      }
    }
  }
  // {R, threads, chunks, vector gathers}: 4-byte gathers with 2 columns per
  // lane and hop chunk 8, vector gathers with 2 columns per wave and hop
  // chunk 2 per lane group
  // plus vector-gather shapes whose column chunks fill whole rounds of the
  // block's waves (32 columns for 1024 threads x 2 columns per wave), and the
  // 256-row fp32 panel
  const int fit32 = (s.sp[1].dim + 31) / 32, fit16 = (s.sp[1].dim + 15) / 16;
  std::vector<std::array<int, 4>> shapes = {{64, 256, 256, 0},  {64, 1024, 64, 0},   {64, 512, 128, 0},
                                            {32, 256, 256, 0},  {64, 1024, 128, 1},  {128, 1024, 128, 1},
                                            {128, 1024, 256, 1}, {32, 1024, 128, 1}, {128, 1024, fit32, 1},
                                            {64, 1024, fit32, 1}};
  if (sizeof(T) == 4) shapes.push_back({256, 512, fit16, 1});
  double bd = 1e300;
  TwoPassCfg dw = best;
  for (int ord : {0, 1})  // vector gathers: columns in the natural, then the locality order
  for (const auto& sh : shapes) {
    if (ord && (!sh[3] || !V.ord_dw)) continue;
    TwoPassCfg w = best;
    w.R = sh[0];
    w.threads_dw = sh[1];
    w.nchunk = std::max(1, std::min(sh[2], s.sp[1].dim));
    w.dw_vec = sh[3];
    w.dw_ord = ord;
    w.ncol = 2;
    w.kch = sh[3] ? 2 : 8;
    if (pass_dw_lds_bytes(V, w) > (size_t)(160 * 1024 - 4096)) continue;
    double ms = 0.0;
    for (int r = 0; r < 2; ++r) {
      EventTimer tm(st);
      launch_pass_dw<T>(V, lut_dev<T>(s.sp[1]), v.as<T>(), v.as<T>(), t.as<T>(), o.as<T>(), nullptr, w, st);
      ms = tm.stop();
    }
    if (verbose())
      std::fprintf(stderr, "[edl] segmented sector down pass vec %d ord %d R %d x %d threads, %d chunks: %.3f ms\n",
                   w.dw_vec, w.dw_ord, w.R, w.threads_dw, w.nchunk, ms);
    if (ms < bd) {
      bd = ms;
      dw = w;
    }
  }
  // The chain step's down pass (MODE 1) over a few vector-gather shapes,
  // with the chain scalars in a middle-step state that repeated launches
  // leave unchanged (beta = 0, norms 1), as for two-pass sectors.
  {
    LzScalars s0{};
    s0.s_cur = 1.0;
    s0.s_old = 1.0;
    DevBuf scd(sizeof(LzScalars)), ab(sizeof(double) * 8);
    EDL_HIP(hipMemcpyAsync(scd.p, &s0, sizeof(s0), hipMemcpyHostToDevice, st));
    launch_zero_bytes(ab.p, sizeof(double) * 8, st);
    // {R, columns per wave, hop chunk, threads, chunks}
    const std::vector<std::array<int, 5>> cshapes = {
        {128, 1, 4, 1024, 256}, {128, 2, 2, 1024, 256}, {64, 1, 4, 1024, 128},   {64, 2, 2, 1024, 128},
        {32, 1, 4, 1024, 128},  {32, 2, 2, 1024, 128},  {128, 2, 2, 1024, fit32}, {128, 1, 4, 1024, fit16}};
    size_t maxpart = 1;
    for (const auto& sh : cshapes) {
      TwoPassCfg w = dw;
      w.R = sh[0];
      w.nchunk = std::max(1, std::min(sh[4], s.sp[1].dim));
      maxpart = std::max(maxpart, (size_t)pass_dw_grid(V, w));
    }
    DevBuf part(sizeof(double) * maxpart);
    LzChain A{};
    A.sc = scd.as<LzScalars>();
    A.alpha = ab.as<double>();
    A.beta = ab.as<double>() + 4;
    A.part_dw = part.as<double>();
    A.nmax = 4;
    double bc = 1e300;
    for (int ord : {0, 1})
    for (const auto& sh : cshapes) {
      if (ord && !V.ord_dw) continue;
      TwoPassCfg w = dw;
      w.R = sh[0];
      w.ncol = sh[1];
      w.kch = sh[2];
      w.threads_dw = sh[3];
      w.nchunk = std::max(1, std::min(sh[4], s.sp[1].dim));
      w.dw_vec = 1;
      w.dw_ord = ord;
      if (pass_dw_lds_bytes(V, w) > (size_t)(160 * 1024 - 4096)) continue;
      double ms = 0.0;
      for (int r = 0; r < 2; ++r) {
        EventTimer tm(st);
        launch_pass_dw<T>(V, lut_dev<T>(s.sp[1]), v.as<T>(), o.as<T>(), t.as<T>(), nullptr, &A, w, st);
        ms = tm.stop();
      }
      if (verbose())
        std::fprintf(stderr, "[edl] segmented sector chain down pass ord %d R %d ncol %d kch %d x %d threads, %d chunks: %.3f ms\n",
                     ord, w.R, w.ncol, w.kch, w.threads_dw, w.nchunk, ms);
      if (ms < bc) {
        bc = ms;
        dw.ch_R = w.R;
        dw.ch_threads = w.threads_dw;
        dw.ch_nchunk = w.nchunk;
        dw.ch_ncol = w.ncol;
        dw.ch_kch = w.kch;
        dw.ch_vec = 1;
        dw.ch_ord = ord;
      }
    }
  }
  seg_bounds(s, sizeof(T), dw.C);
  cache[key] = dw;
  return dw;
}

// Column stride of the engine's own vectors on sector s: padded to whole
// cache lines wherever the down pass runs (two-pass and segmented sectors).
template <typename T>
static int engine_ld(const Sector& s) {
  return (two_pass_ok(s, sizeof(T)) || seg_path(s, sizeof(T))) ? padded_ld(s.sp[0].dim) : s.sp[0].dim;
}

template <typename T>
static void apply_h(Ctx& c, Sector& s, const T* v, T* out, int ld, hipStream_t st) {
  if (seg_path(s, sizeof(T))) {
    // columns larger than LDS: segmented up pass, then the down pass
    const TwoPassCfg cfg = seg_cfg<T>(c, s);
    seg_bounds(s, sizeof(T), cfg.C);
    T* tmp = static_cast<T*>(scratch(c, sizeof(T) * (size_t)ld * s.sp[1].dim));
    const TwoPassView V = two_pass_view(s, ld);
    T* vin = const_cast<T*>(v);  // operator mode only reads it
    launch_up_seg<T>(V, c.model.inter, lut_dev<T>(s.sp[0]), vin, vin, tmp, nullptr, s, cfg.C, cfg.seg_rank, st);
    launch_nonlocal<T>(V, c.model.inter, vin, vin, tmp, nullptr, st);
    launch_pass_dw<T>(V, lut_dev<T>(s.sp[1]), vin, vin, tmp, out, nullptr, cfg, st);
  } else if (two_pass_ok(s, sizeof(T))) {
    const TwoPassCfg cfg = two_pass_cfg<T>(c, s);
    T* tmp = static_cast<T*>(scratch(c, sizeof(T) * (size_t)ld * s.sp[1].dim));
    const TwoPassView V = two_pass_view(s, ld);
    T* vin = const_cast<T*>(v);  // operator mode only reads it
    launch_pass_up<T>(V, c.model.inter, lut_dev<T>(s.sp[0]), vin, vin, tmp, nullptr, cfg, st);
    launch_nonlocal<T>(V, c.model.inter, vin, vin, tmp, nullptr, st);
    launch_pass_dw<T>(V, lut_dev<T>(s.sp[1]), const_cast<T*>(v), const_cast<T*>(v), tmp, out, nullptr, cfg, st);
  } else {
    EDL_CHECK(ld == s.sp[0].dim, "apply_h: the single-pass operator takes the natural layout");
    Schedule w;
    w.block_threads = 256;

    w.grid = c.num_cu * 8;
    hxv_operator<T>(c, s, v, out, st, w);
  }
}

// ----------------------------------------------------------- chain runs
struct ChainOut {
  int nsteps = 0;
  std::vector<double> alpha, beta;
  double norm2 = 0.0;
};

// Per-chain device state allocated for a run.
template <typename T>
struct ChainAlloc {
  DevBuf q0, q1, sc, alpha, beta, counter;
  Arena* ar = nullptr;  // carve from this instead of allocating, when set
  void* get(DevBuf& slot, size_t bytes) {
    if (ar) return ar->take(bytes);
    slot = DevBuf(bytes);
    return slot.p;
  }
  // arena bytes of one chain of D stored elements and nmax steps
  static size_t arena_bytes(int64_t D, int nmax) {
    return 2 * Arena::round(sizeof(T) * D) + Arena::round(sizeof(LzScalars)) +
           2 * Arena::round(sizeof(double) * nmax) + Arena::round(sizeof(unsigned));
  }
};

// d_start is in the natural layout; the stored vectors get column stride ld
// (the padding rows are zero and no kernel writes them).
// With d_ss set, the start vector's sum of squares goes to *d_ss on the
// device (partials in `part`, which holds part_n doubles) and the scalars
// are left to launch_init_scalars: no host round trip.
template <typename T>
static void init_chain(Ctx& c, ChainDev& cd, ChainAlloc<T>& A, const Sector& S, const double* d_start, double norm2,
                       int nmax, hipStream_t st, int ld = 0, double* d_ss = nullptr, double* part = nullptr,
                       int64_t part_n = 0) {
  if (ld == 0) ld = S.sp[0].dim;
  const int64_t D = (int64_t)ld * S.sp[1].dim;
  fill_sector_view(S, sizeof(T) == 4, cd);
  cd.ld = ld;
  T* q0 = static_cast<T*>(A.get(A.q0, sizeof(T) * D));
  T* q1 = static_cast<T*>(A.get(A.q1, sizeof(T) * D));
  LzScalars* sc = static_cast<LzScalars*>(A.get(A.sc, sizeof(LzScalars)));
  double* alpha = static_cast<double*>(A.get(A.alpha, sizeof(double) * nmax));
  double* beta = static_cast<double*>(A.get(A.beta, sizeof(double) * nmax));
  unsigned* counter = static_cast<unsigned*>(A.get(A.counter, sizeof(unsigned)));
  if (ld != S.sp[0].dim) launch_zero_bytes(q0, sizeof(T) * D, st);
  launch_convert_ld<T>(d_start, q0, S.sp[0].dim, S.sp[1].dim, ld, st);
  launch_zero_bytes(q1, sizeof(T) * D, st);
  launch_zero_bytes(alpha, sizeof(double) * nmax, st);
  launch_zero_bytes(beta, sizeof(double) * nmax, st);
  launch_zero_bytes(counter, 4, st);
  if (d_ss) {
    EDL_CHECK(part_n >= partials_size(D), "init_chain: partial-sum buffer smaller than the launch grid");
    device_sumsq_to<T>(q0, D, part, d_ss, st);
  } else {
    const double ss = device_sumsq<T>(q0, D, st);
    LzScalars s{};
    s.k = 0;
    s.stopped = (norm2 == 0.0 || ss == 0.0) ? 1 : 0;
    s.nsteps = 0;
    s.s_cur = ss > 0.0 ? std::sqrt(ss) : 1.0;
    s.s_old = 1.0;
    s.beta_k = 0.0;
    s.alpha_k = 0.0;
    s.norm2 = norm2;
    EDL_HIP(hipMemcpyAsync(sc, &s, sizeof(s), hipMemcpyHostToDevice, st));
    // `s` lives on this stack frame: the copy must have read it before return
    EDL_HIP(hipStreamSynchronize(st));
  }
  cd.q0 = q0;
  cd.q1 = q1;
  cd.x = nullptr;
  cd.ccoef = nullptr;
  cd.counter = counter;
  cd.sc = sc;
  cd.alpha = alpha;
  cd.beta = beta;
  cd.nmax = nmax;
  cd.brk_tol = 1e-14;
}

// Autotuned schedule cache: key = (precision, sum of dims, #chains).
struct TuneKey {
  int prec;
  int64_t dims;
  int n;
  bool operator<(const TuneKey& o) const {
    return std::tie(prec, dims, n) < std::tie(o.prec, o.dims, o.n);
  }
};
static std::map<TuneKey, Schedule>& tune_cache() {
  static std::map<TuneKey, Schedule> m;
  return m;
}

// Pick the tile/wave schedule of a chain set by timing phase A at step 0,
// which is idempotent there (beta_0 = 0, q_{-1} = 0).
template <typename T>
static Schedule autotune(Ctx& c, const std::vector<ChainDev>& chains, hipStream_t st) {
  TuneKey key{(int)sizeof(T), 0, (int)chains.size()};
  for (auto& cd : chains) key.dims += (int64_t)cd.dimup * cd.dimdw;
  auto it = tune_cache().find(key);
  if (it != tune_cache().end()) return it->second;
  const int threads[] = {64, 128, 256, 512, 1024};
  const int gmul[] = {2, 4, 8, 16};
  Schedule best;
  best.block_threads = 256;
  best.grid = c.num_cu * 8;
  double best_ms = 1e300;
  {
    // The default schedule first: where one launch already takes longer
    // than kTuneCapMs (sectors of ~1e9 states and more, memory bound), the
    // 80 timing launches of the search would cost minutes for a few percent.
    constexpr double kTuneCapMs = 25.0;
    ChainSet<T> cs;
    cs.h = chains;
    chainset_finalize(c, cs, best);
    EventTimer tm(st);
    lz_phase_a(c, cs, st);
    const double ms = tm.stop();
    if (ms > kTuneCapMs) {
      best = cs.sch;
      if (verbose())
        std::fprintf(stderr, "[edl] single-pass schedule: default (%d threads, grid %d); one launch %.1f ms\n",
                     best.block_threads, best.grid, ms);
      tune_cache()[key] = best;
      return best;
    }
  }
  for (int th : threads)
    for (int gm : gmul) {
        Schedule w;
        w.block_threads = th;
        w.grid = c.num_cu * gm;
        ChainSet<T> cs;
        cs.h = chains;
        chainset_finalize(c, cs, w);
        lz_phase_a(c, cs, st);  // warm-up
        EventTimer tm(st);
        for (int rep = 0; rep < 3; ++rep) lz_phase_a(c, cs, st);
        const double ms = tm.stop();
        if (ms < best_ms) {
          best_ms = ms;
          best = cs.sch;
          best.grid = w.grid;
        }
      }
  tune_cache()[key] = best;
  return best;
}

// Run the chain set until every chain has stopped or `steps` steps, with the
// two phases of a step captured once in a hipGraph and replayed.  With
// check > 0 and a single chain, the lowest Ritz value is checked every
// `check` steps and the run ends when it changes by <= tol*|E|.
// Phase A of a single large chain done by the two-pass operator.
template <typename T>
struct TwoPassStep {
  TwoPassView V;
  TwoPassCfg cfg;
  const T* lut_up;
  const T* lut_dw;
  T* tmp;
  LzChain L;
  const Sector* seg = nullptr;  // set: segmented up pass, phase B as its own launch
  // Deflation against nx fixed vectors X (phase B as its own launch too)
  const double* X = nullptr;
  int nx = 0;
  double* defl_part = nullptr;
  double* defl_coef = nullptr;
};

// EDL_GRAPHS=0: the steps are launched directly, one launch per kernel and
// step, instead of replaying a captured graph (same kernels and arguments;
// for tools that collect per-dispatch counters).
static bool graphs_enabled() {
  const char* e = std::getenv("EDL_GRAPHS");
  return !(e && std::strcmp(e, "0") == 0);
}

// Replays `enqueue` (the launches of one step, all on `st`) through a graph
// captured once, or launches it directly when graphs are off.
struct StepRunner {
  hipStream_t st;
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
  bool graph;
  template <typename F>
  StepRunner(hipStream_t s, F&& enqueue) : st(s), graph(graphs_enabled()) {
    if (!graph) {
      direct = [enqueue]() { enqueue(); };
      return;
    }
    EDL_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
      enqueue();
    } catch (...) {
      // leave the stream out of capture mode before the error propagates
      hipGraph_t part = nullptr;
      (void)hipStreamEndCapture(st, &part);
      if (part) (void)hipGraphDestroy(part);
      throw;
    }
    EDL_HIP(hipStreamEndCapture(st, &g));
    if (hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) {
      (void)hipGraphDestroy(g);
      throw Error("hipGraphInstantiate failed for the Lanczos step");
    }
  }
  void launch() {
    if (graph)
      EDL_HIP(hipGraphLaunch(ge, st));
    else
      direct();
  }
  ~StepRunner() {
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
  }
  std::function<void()> direct;
};

template <typename T>
static int run_steps(Ctx& c, ChainSet<T>& cs, int steps, int check, double tol, hipStream_t st,
                     const TwoPassStep<T>* tp = nullptr) {
  auto enqueue = [&]() {
    if (tp && (tp->seg || tp->nx > 0)) {
      // phase B of the previous step (+ deflation), beta, up (segmented, or the
      // column-panel pass with no phase B pending), down (+ phase A), alpha
      T* q0 = static_cast<T*>(cs.h[0].q0);
      T* q1 = static_cast<T*>(cs.h[0].q1);
      launch_phase_b<T>(tp->V, q0, q1, tp->L, st);
      if (tp->nx > 0) launch_deflate<T>(tp->V, q0, q1, tp->L, tp->X, tp->nx, tp->defl_part, tp->defl_coef, st);
      launch_lz_beta_n(tp->L, phase_b_grid(tp->V), st);
      if (tp->seg)
        launch_up_seg<T>(tp->V, c.model.inter, tp->lut_up, q0, q1, tp->tmp, &tp->L, *tp->seg, tp->cfg.C,
                         tp->cfg.seg_rank, st);
      else
        launch_pass_up<T>(tp->V, c.model.inter, tp->lut_up, q0, q1, tp->tmp, &tp->L, tp->cfg, st);
      launch_nonlocal<T>(tp->V, c.model.inter, q0, q1, tp->tmp, &tp->L, st);
      launch_pass_dw<T>(tp->V, tp->lut_dw, q0, q1, tp->tmp, nullptr, &tp->L, tp->cfg, st);
      launch_lz_alpha(tp->L, tp->V, tp->cfg, st);
    } else if (tp) {
      // up (+ phase B of the previous step), beta, down (+ phase A), alpha
      T* q0 = static_cast<T*>(cs.h[0].q0);
      T* q1 = static_cast<T*>(cs.h[0].q1);
      launch_pass_up<T>(tp->V, c.model.inter, tp->lut_up, q0, q1, tp->tmp, &tp->L, tp->cfg, st);
      launch_lz_beta(tp->L, tp->V, tp->cfg, st);
      launch_nonlocal<T>(tp->V, c.model.inter, q0, q1, tp->tmp, &tp->L, st);
      launch_pass_dw<T>(tp->V, tp->lut_dw, q0, q1, tp->tmp, nullptr, &tp->L, tp->cfg, st);
      launch_lz_alpha(tp->L, tp->V, tp->cfg, st);
    } else {
      lz_phase_a(c, cs, st);
      lz_phase_b(c, cs, st);
    }
  };
  StepRunner run(st, enqueue);
  double e_prev = 0.0;
  bool have_prev = false;
  int done = 0;
  for (int k = 0; k < steps; ++k) {
    run.launch();
    ++done;
    if (check > 0 && tol > 0.0 && (k + 1) % check == 0 && cs.h.size() == 1) {
      LzScalars s{};
      EDL_HIP(hipMemcpyAsync(&s, cs.h[0].sc, sizeof(s), hipMemcpyDeviceToHost, st));
      EDL_HIP(hipStreamSynchronize(st));
      const int m = s.nsteps;
      if (s.stopped) break;
      std::vector<double> a(m), b(m), w(m);
      EDL_HIP(hipMemcpy(a.data(), cs.h[0].alpha, sizeof(double) * m, hipMemcpyDeviceToHost));
      EDL_HIP(hipMemcpy(b.data(), cs.h[0].beta, sizeof(double) * m, hipMemcpyDeviceToHost));
      std::vector<double> e(m, 0.0);
      for (int i = 0; i + 1 < m; ++i) e[i] = b[i + 1];
      tridiag_eig(m, a.data(), e.data(), w.data(), nullptr, nullptr);
      if (have_prev && std::fabs(w[0] - e_prev) <= tol * std::fabs(w[0])) break;
      e_prev = w[0];
      have_prev = true;
    }
  }
  EDL_HIP(hipStreamSynchronize(st));
  return done;
}

// EDL_BATCHED_KERNELS=single keeps batched chains on the single-pass kernels,
// =chain runs them one at a time through the two-pass (or segmented) step
// wherever the sectors allow (tests run every path on the same seeds).
static bool batched_single_pass_forced() {
  const char* e = std::getenv("EDL_BATCHED_KERNELS");
  return e && std::strcmp(e, "single") == 0;
}
static bool batched_chain_forced() {
  const char* e = std::getenv("EDL_BATCHED_KERNELS");
  return e && std::strcmp(e, "chain") == 0;
}
// Sectors from this many states on fill the GPU with one chain: a group of
// them runs its chains one after another through the two-pass step.
static constexpr int64_t kChainAloneDim = 50000000;

// `steps` Lanczos steps of every chain: up pass (with the previous step's
// phase B), down pass + phase A; captured once as a graph and replayed.
template <typename T>
static int run_steps_bt(Ctx& c, const BtSet& bs, int steps, hipStream_t st) {
  StepRunner run(st, [&]() {
    bt_launch_up<T>(bs, c.model.inter, st);
    bt_launch_dw<T>(bs, st);
  });
  for (int k = 0; k < steps; ++k) run.launch();
  EDL_HIP(hipStreamSynchronize(st));
  return steps;
}

// Chain groups of at least this many chains run as parts on several
// streams (the parts' steps overlap: one part's down tiles share the GPU
// with another part's up tiles); EDL_BATCHED_STREAMS=1..4 forces a count.
static constexpr size_t kBtTwoStreamChains = 64;
static constexpr int kBtMaxStreams = 4;
static int bt_streams(size_t nchains) {
  if (const char* e = std::getenv("EDL_BATCHED_STREAMS")) {
    const int v = std::atoi(e);
    EDL_CHECK(v >= 1 && v <= kBtMaxStreams, "EDL_BATCHED_STREAMS must be 1 .. 4");
    return (int)std::min<size_t>((size_t)v, std::max<size_t>(nchains, 1));
  }
  return nchains >= kBtTwoStreamChains ? 2 : 1;
}

// A stream of the default priority (the library's other stream included)
// and an event, released on scope exit.
struct AuxStream {
  hipStream_t s = nullptr;
  hipEvent_t e = nullptr;
  AuxStream() {
    EDL_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      (void)hipStreamDestroy(s);
      throw Error("hipEventCreateWithFlags failed");
    }
  }
  ~AuxStream() {
    (void)hipStreamSynchronize(s);
    (void)hipEventDestroy(e);
    (void)hipStreamDestroy(s);
  }
};

// `steps` steps of independent chain sets, part 0 on `st` and each other
// part on a stream of its own, each replayed from its own graph, launched
// in turn; `st` waits for all of them.
template <typename T>
static int run_steps_btn(Ctx& c, const std::vector<BtSet>& parts, int steps, hipStream_t st) {
  const size_t n = parts.size();
  EDL_CHECK(n >= 1 && n <= (size_t)kBtMaxStreams, "run_steps_btn: 1 .. 4 chain sets");
  std::vector<std::unique_ptr<AuxStream>> aux;
  for (size_t i = 1; i < n; ++i) aux.emplace_back(new AuxStream());
  hipEvent_t e0 = aux.empty() ? nullptr : aux[0]->e;
  if (e0) EDL_HIP(hipEventRecord(e0, st));  // the chains' set-up on st precedes every part
  for (auto& a : aux) EDL_HIP(hipStreamWaitEvent(a->s, e0, 0));
  std::vector<std::unique_ptr<StepRunner>> run;
  for (size_t i = 0; i < n; ++i) {
    const BtSet* bs = &parts[i];
    hipStream_t s = i ? aux[i - 1]->s : st;
    run.emplace_back(new StepRunner(s, [&c, bs, s]() {
      bt_launch_up<T>(*bs, c.model.inter, s);
      bt_launch_dw<T>(*bs, s);
    }));
  }
  for (int k = 0; k < steps; ++k)
    for (auto& r : run) r->launch();
  for (auto& a : aux) {
    EDL_HIP(hipEventRecord(a->e, a->s));
    EDL_HIP(hipStreamWaitEvent(st, a->e, 0));
  }
  EDL_HIP(hipStreamSynchronize(st));
  return steps;
}

template <typename T>
static ChainOut read_chain(const ChainDev& cd, int nmax, hipStream_t st) {
  ChainOut o;
  LzScalars s{};
  EDL_HIP(hipMemcpyAsync(&s, cd.sc, sizeof(s), hipMemcpyDeviceToHost, st));
  o.alpha.assign(nmax, 0.0);
  o.beta.assign(nmax, 0.0);
  EDL_HIP(hipMemcpyAsync(o.alpha.data(), cd.alpha, sizeof(double) * nmax, hipMemcpyDeviceToHost, st));
  EDL_HIP(hipMemcpyAsync(o.beta.data(), cd.beta, sizeof(double) * nmax, hipMemcpyDeviceToHost, st));
  EDL_HIP(hipStreamSynchronize(st));
  o.nsteps = s.nsteps;
  o.norm2 = s.norm2;
  return o;
}

static double lowest_ritz(const ChainOut& o, std::vector<double>* vec) {
  const int m = o.nsteps;
  EDL_CHECK(m > 0, "Lanczos produced no step");
  std::vector<double> e(m, 0.0), w(m), Z;
  for (int i = 0; i + 1 < m; ++i) e[i] = o.beta[i + 1];
  if (vec) Z.resize((size_t)m * m);
  tridiag_eig(m, o.alpha.data(), e.data(), w.data(), vec ? Z.data() : nullptr, nullptr);
  if (vec) vec->assign(Z.begin(), Z.begin() + m);
  return w[0];
}

struct GsRun {
  double e0 = 0.0;
  ChainOut out;
  int64_t hxv = 0;
};

// One ground-state Lanczos run from d_start; the normalised Ritz vector of
// the lowest eigenvalue is rebuilt into d_x (fp64) by replaying the same
// deterministic recurrence with the tridiagonal eigenvector as weights.
template <typename T>
static GsRun gs_run(Ctx& c, Sector& S, const double* d_start, int niter, double tol, int check, double* d_x,
                    const double* d_X = nullptr, int nX = 0) {
  hipStream_t st = c.stream;
  const int nmax = (int)std::min<int64_t>(niter, S.dim);
  GsRun r;
  Schedule sch;
  // Sectors whose columns fit in LDS: the two-pass step (LzChain); larger
  // columns: the same step with the segmented up pass.
  const bool seg = seg_path(S, sizeof(T));
  const bool two = seg || two_pass_ok(S, sizeof(T));
  TwoPassStep<T> tps{};
  DevBuf tp_up, tp_dw;
  // Two-pass chains keep their vectors with the padded column stride.
  const int ld = two ? padded_ld(S.sp[0].dim) : S.sp[0].dim;
  if (two) {
    tps.cfg = chain_cfg(seg ? seg_cfg<T>(c, S) : two_pass_cfg<T>(c, S));
    if (seg) {
      seg_bounds(S, sizeof(T), tps.cfg.C);
      tps.seg = &S;
    }
    tps.V = two_pass_view(S, ld);
    tps.lut_up = lut_dev<T>(S.sp[0]);
    tps.lut_dw = lut_dev<T>(S.sp[1]);
    tps.tmp = static_cast<T*>(scratch(c, sizeof(T) * (size_t)ld * S.sp[1].dim));
    tp_up = DevBuf(sizeof(double) * std::max(phase_b_grid(tps.V), pass_up_grid(tps.V, tps.cfg)));
    tp_dw = DevBuf(sizeof(double) * pass_dw_grid(tps.V, tps.cfg));
    tps.L.part_up = tp_up.as<double>();
    tps.L.part_dw = tp_dw.as<double>();
    sch.block_threads = 256;
    sch.grid = c.num_cu * 8;
  }
  DevBuf defl_part, defl_coef;
  if (nX > 0) {
    EDL_CHECK(two, "deflated Lanczos needs packed hop tables (two-pass or segmented sectors)");
    defl_part = DevBuf(sizeof(double) * (size_t)nX * phase_b_grid(tps.V));
    defl_coef = DevBuf(sizeof(double) * (size_t)nX);
    tps.X = d_X;
    tps.nx = nX;
    tps.defl_part = defl_part.as<double>();
    tps.defl_coef = defl_coef.as<double>();
  }
  auto bind = [&](const ChainDev& cd, int n, double* x, const double* ccoef) {
    tps.L.sc = cd.sc;
    tps.L.alpha = cd.alpha;
    tps.L.beta = cd.beta;
    tps.L.nmax = n;
    tps.L.brk_tol = cd.brk_tol;
    tps.L.x = x;
    tps.L.ccoef = ccoef;
  };
  // Stored basis (two-pass sectors, no deflation, when nmax vectors fit in
  // 70 % of the free device memory; EDL_LANCZOS_BASIS=0 turns it off): the
  // up pass copies each q_k it reads for phase B, and the eigenvector is one
  // pass over the stored vectors instead of a second run of the chain.
  const size_t vbytes = sizeof(T) * (size_t)ld * S.sp[1].dim;
  bool use_basis = false;
  if (d_x && two && !seg && nX == 0) {
    const char* e = std::getenv("EDL_LANCZOS_BASIS");
    size_t free_b = 0, total_b = 0;
    EDL_HIP(hipMemGetInfo(&free_b, &total_b));
    use_basis = !(e && e[0] == '0') && (double)vbytes * nmax <= 0.7 * (double)free_b;
  }
  DevBuf basis, bscale;
  if (use_basis) {
    // the second run of the chain stays the fallback if the memory is gone
    basis = DevBuf::try_alloc(vbytes * (size_t)nmax);
    use_basis = basis.p != nullptr;
    if (use_basis) bscale = DevBuf(sizeof(double) * (size_t)nmax);
  }
  T* last_q = nullptr;
  {
    ChainSet<T> cs;
    ChainAlloc<T> A;
    ChainDev cd{};
    init_chain<T>(c, cd, A, S, d_start, 1.0, nmax, st, ld);
    if (!two) sch = autotune<T>(c, {cd}, st);
    cs.h.push_back(cd);
    chainset_finalize(c, cs, sch);
    bind(cd, nmax, nullptr, nullptr);
    tps.L.basis = use_basis ? basis.p : nullptr;
    tps.L.bscale = use_basis ? bscale.as<double>() : nullptr;
    r.hxv += run_steps(c, cs, nmax, check, tol, st, two ? &tps : nullptr);
    tps.L.basis = nullptr;
    tps.L.bscale = nullptr;
    r.out = read_chain<T>(cs.h[0], nmax, st);
    if (use_basis && r.out.nsteps > 0) {
      // q_{m-1} and its norm (the scalars keep the last completed step) are
      // still in the chain's buffers
      const int m = r.out.nsteps;
      last_q = static_cast<T*>(((m - 1) & 1) ? cd.q1 : cd.q0);
      EDL_HIP(hipMemcpyAsync(static_cast<char*>(basis.p) + vbytes * (size_t)(m - 1), last_q, vbytes,
                             hipMemcpyDeviceToDevice, st));
      EDL_HIP(hipMemcpyAsync(bscale.as<double>() + (m - 1), &cd.sc->s_cur, sizeof(double), hipMemcpyDeviceToDevice,
                             st));
      EDL_HIP(hipStreamSynchronize(st));
    }
  }
  std::vector<double> cvec;
  r.e0 = lowest_ritz(r.out, d_x ? &cvec : nullptr);
  if (d_x && use_basis && r.out.nsteps > 0) {
    const int m = r.out.nsteps;
    std::vector<double> sc(m), w(m);
    EDL_HIP(hipMemcpy(sc.data(), bscale.p, sizeof(double) * m, hipMemcpyDeviceToHost));
    for (int k = 0; k < m; ++k) w[k] = cvec[k] / sc[k];
    DevBuf dw(sizeof(double) * m);
    EDL_HIP(hipMemcpyAsync(dw.p, w.data(), sizeof(double) * m, hipMemcpyHostToDevice, st));
    launch_basis_combine<T>(static_cast<const T*>(basis.p), tps.V, dw.as<double>(), m, d_x, st);
    const double n2 = device_dot(c, d_x, d_x, S.dim, st);
    launch_scale(d_x, S.dim, 1.0 / std::sqrt(n2), st);
    EDL_HIP(hipStreamSynchronize(st));
  } else if (d_x) {
    const int m = r.out.nsteps;
    DevBuf dc(sizeof(double) * m);
    EDL_HIP(hipMemcpyAsync(dc.p, cvec.data(), sizeof(double) * m, hipMemcpyHostToDevice, st));
    launch_zero_bytes(d_x, sizeof(double) * S.dim, st);
    ChainSet<T> cs;
    ChainAlloc<T> A;
    ChainDev cd{};
    init_chain<T>(c, cd, A, S, d_start, 1.0, m, st, ld);
    cd.x = d_x;
    cd.ccoef = dc.as<double>();
    cs.h.push_back(cd);
    chainset_finalize(c, cs, sch);
    bind(cd, m, d_x, dc.as<double>());
    // two-pass: one more launch, whose phase B (in the up pass, or its own
    // launch on segmented sectors) adds the replay term of the last step
    r.hxv += m;
    run_steps(c, cs, two ? m + 1 : m, 0, 0.0, st, two ? &tps : nullptr);
    const double n2 = device_dot(c, d_x, d_x, S.dim, st);
    launch_scale(d_x, S.dim, 1.0 / std::sqrt(n2), st);
    EDL_HIP(hipStreamSynchronize(st));
  }
  return r;
}

struct GsStats {
  int steps_lo = 0, steps_hi = 0;
  double e0_lo = 0.0;
};
static std::map<const Ctx*, GsStats>& gs_stats() {
  static std::map<const Ctx*, GsStats> m;
  return m;
}

// Drop the installed state (a new model invalidates it).
static void clear_gs(Ctx& c) {
  if (c.d_gs) (void)hipFree(c.d_gs);
  c.d_gs = nullptr;
  c.gs_dim = 0;
  c.gs_sector = 0;
  c.gs_energy = 0.0;
}

// Install a computed state: the context takes the buffer over (only after
// the computation succeeded, so a failure leaves the previous state whole).
static void adopt_gs(Ctx& c, int isector, int64_t dim, double energy, DevBuf&& buf) {
  clear_gs(c);
  c.d_gs = buf.as<double>();
  buf.p = nullptr;
  c.gs_dim = dim;
  c.gs_sector = isector;
  c.gs_energy = energy;
}

// The sector of the installed state, checked against the state's length.
static Sector& gs_sector_of(Ctx& c, const char* who) {
  EDL_CHECK(c.d_gs && c.gs_sector > 0, std::string(who) + ": no ground state installed");
  Sector& S = get_sector(c, c.gs_sector);
  EDL_CHECK(S.dim == c.gs_dim, std::string(who) + ": installed state does not match its sector");
  return S;
}

}  // namespace edl

using namespace edl;

// ------------------------------------------------------------------ C-ABI
#define API_BEGIN try {
#define API_END                                          \
  }                                                      \
  catch (const std::exception& e) {                      \
    if (ctx) ctx->c.last_error = e.what();               \
    return 1;                                            \
  }                                                      \
  return 0;

static thread_local std::string g_create_error;

extern "C" int edl_ctx_create(edl_ctx** out, int device) {
  try {
    EDL_CHECK(out != nullptr, "edl_ctx_create: null output pointer");
    int n = 0;
    EDL_HIP(hipGetDeviceCount(&n));
    EDL_CHECK(n > 0, "edl_ctx_create: no HIP device visible");
    EDL_CHECK(device >= 0 && device < n, "edl_ctx_create: device index out of range");
    EDL_HIP(hipSetDevice(device));
    hipDeviceProp_t prop{};
    EDL_HIP(hipGetDeviceProperties(&prop, device));
    init_binomials();
    noalloc_scratch_setup(device);
    // the stream is the last resource taken: nothing after it can throw
    auto* ctx = new edl_ctx();
    ctx->c.device = device;
    ctx->c.num_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    if (hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking) != hipSuccess) {
      delete ctx;
      throw Error("edl_ctx_create: hipStreamCreateWithFlags failed");
    }
    *out = ctx;
    return 0;
  } catch (const std::exception& e) {
    g_create_error = e.what();
    if (out) *out = nullptr;
    return 1;
  }
}

extern "C" int edl_ctx_destroy(edl_ctx* ctx) {
  if (!ctx) return 0;
  API_BEGIN
  EDL_HIP(hipSetDevice(ctx->c.device));
  EDL_HIP(hipStreamSynchronize(ctx->c.stream));
  release_sectors(ctx->c);
  if (ctx->c.d_gs) (void)hipFree(ctx->c.d_gs);
  if (ctx->c.d_scratch) (void)hipFree(ctx->c.d_scratch);
  for (void* p : ctx->c.user_bufs) (void)hipFree(p);
  (void)hipStreamDestroy(ctx->c.stream);
  gs_stats().erase(&ctx->c);
  delete ctx;
  return 0;
  API_END
}

extern "C" const char* edl_last_error(const edl_ctx* ctx) {
  return ctx ? ctx->c.last_error.c_str() : g_create_error.c_str();
}

extern "C" int edl_set_hamiltonian(edl_ctx* ctx, int Ns, int Nspin, int Norb, const int* Nsites, const double* Hre,
                                   const double* Him, const double* Uloc, double Ust, double Jh, double Jx, double Jp,
                                   double xmu, int hfmode, int ed_filling) {
  API_BEGIN
  EDL_CHECK(ctx && Nsites && Hre && Uloc, "edl_set_hamiltonian: null argument");
  EDL_CHECK(Ns >= 1 && Ns <= kMaxNs, "edl_set_hamiltonian: Ns must be in [1, 31]");
  EDL_CHECK(Nspin == 1 || Nspin == 2, "edl_set_hamiltonian: Nspin must be 1 or 2");
  EDL_CHECK(Norb >= 1 && Norb <= kMaxOrb, "edl_set_hamiltonian: Norb out of range");
  int sum = 0;
  for (int k = 0; k < Norb; ++k) sum += Nsites[k];
  EDL_CHECK(sum == Ns, "edl_set_hamiltonian: Ns != sum(Nsites)");
  const size_t n = (size_t)Nspin * Ns * Ns;
  for (size_t k = 0; k < n; ++k)
    EDL_CHECK(!Him || Him[k] == 0.0, "edl_set_hamiltonian: complex Hij is not supported (real kernels)");
  // Hermiticity (GRAPH_MATRIX.f90:324-326, tolerance 1e-6)
  for (int s = 0; s < Nspin; ++s)
    for (int i = 0; i < Ns; ++i)
      for (int j = 0; j < Ns; ++j) {
        const double a = Hre[s + (size_t)Nspin * (i + (size_t)Ns * j)];
        const double b = Hre[s + (size_t)Nspin * (j + (size_t)Ns * i)];
        EDL_CHECK(std::fabs(a - b) <= 1e-6, "Hij_build ERROR: Hij is not Hermitian");
      }
  EDL_HIP(hipSetDevice(ctx->c.device));
  EDL_HIP(hipStreamSynchronize(ctx->c.stream));
  release_sectors(ctx->c);
  // the installed state belongs to the previous model
  clear_gs(ctx->c);
  Model& M = ctx->c.model;
  M.Ns = Ns;
  M.Nspin = Nspin;
  M.Norb = Norb;
  M.Nsites.assign(Nsites, Nsites + Norb);
  M.Hij.assign(Hre, Hre + n);
  M.Uloc.assign(Uloc, Uloc + Norb);
  M.Ust = Ust;
  M.Jh = Jh;
  M.Jx = Jx;
  M.Jp = Jp;
  M.xmu = xmu;
  M.hfmode = hfmode ? 1 : 0;
  M.ed_filling = ed_filling;
  InterParams I{};
  // U n_up n_dw per orbital (HxV_local.f90:23-30), only when any Uloc /= 0
  bool anyU = false;
  for (int k = 0; k < Norb; ++k) anyU = anyU || Uloc[k] != 0.0;
  I.norb = anyU ? Norb : 1;
  for (int k = 0; k < Norb && anyU; ++k) {
    uint32_t mask = 0;
    for (int isite = 1; isite <= Nsites[k]; ++isite) {
      const int io = (k == 0) ? isite : isite + k * Nsites[k - 1];  // pack_indices, AUX_FUNX.f90:42-53
      EDL_CHECK(io >= 1 && io <= Ns, "pack_indices out of range");
      mask |= 1u << (io - 1);
    }
    I.orb_mask[k] = mask;
    I.U[k] = Uloc[k];
  }
  // same-site inter-orbital pairs (HxV_local.f90:34-45)
  int np = 0;
  for (int a = 0; a < Norb; ++a)
    for (int b = a + 1; b < Norb; ++b)
      for (int isite = 1; isite <= std::min(Nsites[a], Nsites[b]); ++isite) {
        const int io = (a == 0) ? isite : isite + a * Nsites[a - 1];
        const int jo = (b == 0) ? isite : isite + b * Nsites[b - 1];
        EDL_CHECK(np < kMaxPairs, "too many inter-orbital pairs");
        EDL_CHECK(io >= 1 && io <= Ns && jo >= 1 && jo <= Ns, "pack_indices out of range");
        I.pi[np] = (uint8_t)(io - 1);
        I.pj[np] = (uint8_t)(jo - 1);
        ++np;
      }
  I.npairs = np;
  I.Ust = Ust;
  // Jhflag = Norb > 1 and (Jx or Jp non-zero) (SETUP.f90:35-36)
  I.Jx = Norb > 1 ? Jx : 0.0;
  I.Jp = Norb > 1 ? Jp : 0.0;
  M.inter = I;
  M.set = true;
  return 0;
  API_END
}

extern "C" int edl_sector_index(int Ns, int nup, int ndw) {
  if (nup < 0 || ndw < 0 || nup > Ns || ndw > Ns) return 0;
  return 1 + nup * (Ns + 1) + ndw;  // get_Sector, SECTOR.f90:370-381
}

extern "C" int edl_sector_qn(int Ns, int isector, int* nup, int* ndw) {
  if (isector < 1 || isector > (Ns + 1) * (Ns + 1) || !nup || !ndw) return 1;
  *ndw = (isector - 1) % (Ns + 1);
  *nup = ((isector - 1) / (Ns + 1)) % (Ns + 1);
  return 0;
}

extern "C" int edl_sector_build(edl_ctx* ctx, int isector, int64_t* dim, int32_t* dimup, int32_t* dimdw) {
  API_BEGIN
  EDL_HIP(hipSetDevice(ctx->c.device));
  Sector& s = get_sector(ctx->c, isector);
  if (dim) *dim = s.dim;
  if (dimup) *dimup = s.sp[0].dim;
  if (dimdw) *dimdw = s.sp[1].dim;
  API_END
}

extern "C" int edl_sector_maps(edl_ctx* ctx, int isector, int32_t* map_up, int32_t* map_dw) {
  API_BEGIN
  EDL_HIP(hipSetDevice(ctx->c.device));
  Sector& s = get_sector(ctx->c, isector);
  if (map_up)
    EDL_HIP(hipMemcpyAsync(map_up, s.sp[0].map, sizeof(int32_t) * s.sp[0].dim, hipMemcpyDeviceToHost, ctx->c.stream));
  if (map_dw)
    EDL_HIP(hipMemcpyAsync(map_dw, s.sp[1].map, sizeof(int32_t) * s.sp[1].dim, hipMemcpyDeviceToHost, ctx->c.stream));
  EDL_HIP(hipStreamSynchronize(ctx->c.stream));
  API_END
}

extern "C" int edl_sector_hops(edl_ctx* ctx, int isector, int ispin, int32_t* kmax, int32_t* nbr, double* coef) {
  API_BEGIN
  EDL_CHECK(ispin == 1 || ispin == 2, "edl_sector_hops: ispin must be 1 or 2");
  EDL_HIP(hipSetDevice(ctx->c.device));
  Sector& s = get_sector(ctx->c, isector);
  const SpinTable& t = s.sp[ispin - 1];
  if (kmax) *kmax = t.kmax;
  const size_t n = (size_t)t.kmax * t.dim;
  if (nbr && n) EDL_HIP(hipMemcpyAsync(nbr, t.nbr, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx->c.stream));
  if (coef && n) EDL_HIP(hipMemcpyAsync(coef, t.coef64, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->c.stream));
  EDL_HIP(hipStreamSynchronize(ctx->c.stream));
  API_END
}

extern "C" int edl_sector_rank(edl_ctx* ctx, int isector, int ispin, int64_t n, const int32_t* states, int32_t* pos) {
  API_BEGIN
  EDL_CHECK(ispin == 1 || ispin == 2, "edl_sector_rank: ispin must be 1 or 2");
  EDL_HIP(hipSetDevice(ctx->c.device));
  Sector& s = get_sector(ctx->c, isector);
  if (n <= 0) return 0;
  DevBuf ds(sizeof(int32_t) * n), dp(sizeof(int32_t) * n);
  EDL_HIP(hipMemcpyAsync(ds.p, states, sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->c.stream));
  rank_strings(ctx->c, s.sp[ispin - 1], ctx->c.model.Ns, n, ds.as<int32_t>(), dp.as<int32_t>());
  EDL_HIP(hipMemcpyAsync(pos, dp.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx->c.stream));
  EDL_HIP(hipStreamSynchronize(ctx->c.stream));
  API_END
}

static Schedule operator_schedule(Ctx& c) {
  Schedule w;
  w.block_threads = 256;

  w.grid = c.num_cu * 8;
  return w;
}

extern "C" int edl_hxv(edl_ctx* ctx, int isector, int prec, const void* v, void* hv) {
  API_BEGIN
  EDL_CHECK(v && hv, "edl_hxv: null vector");
  EDL_CHECK(prec == EDL_FP64 || prec == EDL_FP32, "edl_hxv: prec must be EDL_FP64 or EDL_FP32");
  EDL_HIP(hipSetDevice(ctx->c.device));
  Sector& s = get_sector(ctx->c, isector);
  EventTimer tm(ctx->c.stream);
  if (prec == EDL_FP64)
    apply_h<double>(ctx->c, s, (const double*)v, (double*)hv, s.sp[0].dim, ctx->c.stream);
  else
    apply_h<float>(ctx->c, s, (const float*)v, (float*)hv, s.sp[0].dim, ctx->c.stream);
  ctx->c.last_ms = tm.stop();
  ctx->c.last_hxv = 1;
  API_END
}

// edl_hxv_noalloc: the two passes with the intermediate T in the file-scope
// device scratch, the schedule the tuner cached for this sector shape (a
// fixed one when none is cached: tuning allocates), no allocation, no
// events.
template <typename T>
static void hxv_noalloc_impl(Ctx& c, Sector& s, const T* v, T* hv) {
  EDL_CHECK(!seg_path(s, sizeof(T)) && two_pass_ok(s, sizeof(T)),
            "edl_hxv_noalloc: the sector needs the two-pass kernels (columns within LDS, packed hop tables)");
  const int ld = s.sp[0].dim;
  const size_t bytes = sizeof(T) * (size_t)ld * s.sp[1].dim;
  EDL_CHECK(bytes <= noalloc_scratch_bytes(),
            "edl_hxv_noalloc: the sector's intermediate (" + std::to_string(bytes) +
                " bytes) exceeds the 256 MiB scratch; use edl_hxv");
  const auto key = std::make_tuple(s.sp[0].dim, s.sp[1].dim, s.sp[0].kmax, s.sp[1].kmax, (int)sizeof(T));
  TwoPassCfg cfg;
  auto it = tp_cache().find(key);
  if (it != tp_cache().end()) {
    cfg = it->second;
  } else {
    cfg.C = 1;
    cfg.up_rpt = 2;
    cfg.up_kch = 8;
    cfg.threads_up = 1024;
    cfg.grid_up = c.num_cu;
    cfg.R = 64;
    cfg.threads_dw = 256;
    cfg.nchunk = std::min(32, s.sp[1].dim);
    cfg.ncol = 1;
    cfg.kch = 16;
    cfg.dw_vec = 0;
  }
  const TwoPassView V = two_pass_view(s, ld);
  T* tmp = static_cast<T*>(noalloc_scratch(c.device));
  T* vin = const_cast<T*>(v);  // operator mode only reads it
  launch_pass_up<T>(V, c.model.inter, lut_dev<T>(s.sp[0]), vin, vin, tmp, nullptr, cfg, c.stream);
  launch_nonlocal<T>(V, c.model.inter, vin, vin, tmp, nullptr, c.stream);
  launch_pass_dw<T>(V, lut_dev<T>(s.sp[1]), vin, vin, tmp, hv, nullptr, cfg, c.stream);
  EDL_HIP(hipStreamSynchronize(c.stream));
}

template <typename T>
static void hxv_host_impl(Ctx& c, Sector& s, const double* v, double* hv, int64_t stride) {
  const int64_t D = s.dim;
  hipStream_t st = c.stream;
  std::vector<double> tmp(D);
  for (int64_t i = 0; i < D; ++i) tmp[i] = v[i * stride];
  // through the engine's own layout (padded columns on two-pass sectors)
  const int ld = engine_ld<T>(s);
  const int64_t Dp = (int64_t)ld * s.sp[1].dim;
  DevBuf d64(sizeof(double) * D), dv(sizeof(T) * Dp), dh(sizeof(T) * Dp);
  EDL_HIP(hipMemcpyAsync(d64.p, tmp.data(), sizeof(double) * D, hipMemcpyHostToDevice, st));
  if (ld != s.sp[0].dim) launch_zero_bytes(dv.p, sizeof(T) * Dp, st);
  launch_convert_ld<T>(d64.as<double>(), dv.as<T>(), s.sp[0].dim, s.sp[1].dim, ld, st);
  apply_h<T>(c, s, dv.as<T>(), dh.as<T>(), ld, st);
  launch_to_fp64_ld<T>(dh.as<T>(), d64.as<double>(), s.sp[0].dim, s.sp[1].dim, ld, st);
  EDL_HIP(hipMemcpyAsync(tmp.data(), d64.p, sizeof(double) * D, hipMemcpyDeviceToHost, st));
  EDL_HIP(hipStreamSynchronize(st));
  for (int64_t i = 0; i < D; ++i) hv[i * stride] = tmp[i];
}

extern "C" int edl_hxv_noalloc(edl_ctx* ctx, int isector, int prec, const void* v, void* hv) {
  API_BEGIN
  EDL_CHECK(v && hv, "edl_hxv_noalloc: null vector");
  EDL_CHECK(prec == EDL_FP64 || prec == EDL_FP32, "edl_hxv_noalloc: prec must be EDL_FP64 or EDL_FP32");
  Ctx& c = ctx->c;
  EDL_HIP(hipSetDevice(c.device));
  auto it = c.sectors.find(isector);
  EDL_CHECK(it != c.sectors.end(), "edl_hxv_noalloc: sector not built (call edl_sector_build first)");
  if (prec == EDL_FP64)
    hxv_noalloc_impl<double>(c, *it->second, static_cast<const double*>(v), static_cast<double*>(hv));
  else
    hxv_noalloc_impl<float>(c, *it->second, static_cast<const float*>(v), static_cast<float*>(hv));
  API_END
}

extern "C" int edl_hxv_host(edl_ctx* ctx, int isector, int prec, const double* v, double* hv) {
  API_BEGIN
  EDL_CHECK(v && hv, "edl_hxv_host: null vector");
  EDL_CHECK(prec == EDL_FP64 || prec == EDL_FP32, "edl_hxv_host: prec must be EDL_FP64 or EDL_FP32");
  EDL_HIP(hipSetDevice(ctx->c.device));
  Sector& s = get_sector(ctx->c, isector);
  if (prec == EDL_FP64) hxv_host_impl<double>(ctx->c, s, v, hv, 1);
  else hxv_host_impl<float>(ctx->c, s, v, hv, 1);
  API_END
}

extern "C" int edl_hxv_host_complex(edl_ctx* ctx, int isector, int prec, const double* v, double* hv) {
  API_BEGIN
  EDL_CHECK(v && hv, "edl_hxv_host_complex: null vector");
  EDL_CHECK(prec == EDL_FP64 || prec == EDL_FP32, "edl_hxv_host_complex: prec must be EDL_FP64 or EDL_FP32");
  EDL_HIP(hipSetDevice(ctx->c.device));
  Sector& s = get_sector(ctx->c, isector);
  for (int part = 0; part < 2; ++part) {
    if (prec == EDL_FP64) hxv_host_impl<double>(ctx->c, s, v + part, hv + part, 2);
    else hxv_host_impl<float>(ctx->c, s, v + part, hv + part, 2);
  }
  API_END
}

extern "C" int edl_vec_random(edl_ctx* ctx, int prec, void* dev, int64_t n, uint64_t seed) {
  API_BEGIN
  EDL_CHECK(dev && n >= 0, "edl_vec_random: bad arguments");
  EDL_CHECK(prec == EDL_FP64 || prec == EDL_FP32, "edl_vec_random: prec must be EDL_FP64 or EDL_FP32");
  Ctx& c = ctx->c;
  EDL_HIP(hipSetDevice(c.device));
  if (prec == EDL_FP64) {
    launch_start_vector(static_cast<double*>(dev), n, seed, c.stream);
  } else {
    // in chunks through an fp64 staging buffer
    const int64_t chunk = std::min<int64_t>(n, (int64_t)1 << 27);
    DevBuf tmp(sizeof(double) * (chunk > 0 ? chunk : 1));
    for (int64_t o = 0; o < n; o += chunk) {
      const int64_t m = std::min(chunk, n - o);
      launch_start_vector(tmp.as<double>(), m, seed + (uint64_t)o * 0x9E3779B97F4A7C15ull, c.stream);
      launch_convert<float>(tmp.as<double>(), static_cast<float*>(dev) + o, m, c.stream);
    }
  }
  EDL_HIP(hipStreamSynchronize(c.stream));
  API_END
}

extern "C" int edl_vec_dot(edl_ctx* ctx, int prec, const void* a, const void* b, int64_t n, double* out) {
  API_BEGIN
  EDL_CHECK(a && b && out && n >= 0, "edl_vec_dot: bad arguments");
  EDL_CHECK(prec == EDL_FP64 || prec == EDL_FP32, "edl_vec_dot: prec must be EDL_FP64 or EDL_FP32");
  Ctx& c = ctx->c;
  EDL_HIP(hipSetDevice(c.device));
  *out = (prec == EDL_FP64) ? device_dot_t<double>(static_cast<const double*>(a), static_cast<const double*>(b), n, c.stream)
                            : device_dot_t<float>(static_cast<const float*>(a), static_cast<const float*>(b), n, c.stream);
  API_END
}

extern "C" int edl_device_alloc(edl_ctx* ctx, int64_t bytes, void** dev_ptr) {
  API_BEGIN
  EDL_CHECK(dev_ptr && bytes > 0, "edl_device_alloc: bad arguments");
  EDL_HIP(hipSetDevice(ctx->c.device));
  void* p = nullptr;
  EDL_HIP(hipMalloc(&p, (size_t)bytes));
  ctx->c.user_bufs.push_back(p);
  *dev_ptr = p;
  API_END
}

extern "C" int edl_copy_h2d(edl_ctx* ctx, void* dst, const void* src, int64_t bytes) {
  API_BEGIN
  EDL_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, ctx->c.stream));
  EDL_HIP(hipStreamSynchronize(ctx->c.stream));
  API_END
}

extern "C" int edl_copy_d2h(edl_ctx* ctx, void* dst, const void* src, int64_t bytes) {
  API_BEGIN
  EDL_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, ctx->c.stream));
  EDL_HIP(hipStreamSynchronize(ctx->c.stream));
  API_END
}

extern "C" int edl_lanczos_gs(edl_ctx* ctx, int isector, int prec, int niter, double tol, uint64_t seed,
                              double refine_tol, double* e0, int* nsteps, double* alpha, double* beta,
                              double* gs_host) {
  API_BEGIN
  EDL_CHECK(prec == EDL_FP64 || prec == EDL_FP32 || prec == EDL_MIXED, "edl_lanczos_gs: bad prec");
  EDL_CHECK(niter >= 1, "edl_lanczos_gs: niter must be >= 1");
  Ctx& c = ctx->c;
  EDL_HIP(hipSetDevice(c.device));
  Sector& S = get_sector(c, isector);
  const int64_t D = S.dim;
  hipStream_t st = c.stream;
  DevBuf start(sizeof(double) * D), x(sizeof(double) * D);
  launch_start_vector(start.as<double>(), D, seed, st);
  GsStats stats;
  EventTimer tm(st);
  GsRun r;
  int64_t hxv = 0;
  if (prec == EDL_FP64) {
    r = gs_run<double>(c, S, start.as<double>(), niter, tol, 5, x.as<double>());
    stats.steps_hi = r.out.nsteps;
  } else if (prec == EDL_FP32) {
    r = gs_run<float>(c, S, start.as<double>(), niter, tol, 5, x.as<double>());
    stats.steps_lo = r.out.nsteps;
  } else {
    GsRun lo = gs_run<float>(c, S, start.as<double>(), niter, tol, 5, x.as<double>());
    stats.steps_lo = lo.out.nsteps;
    stats.e0_lo = lo.e0;
    hxv += lo.hxv;
    // FP64 restart from the FP32 Ritz vector (now in x)
    EDL_HIP(hipMemcpyAsync(start.p, x.p, sizeof(double) * D, hipMemcpyDeviceToDevice, st));
    const double rt = refine_tol > 0.0 ? refine_tol : 1e-13;
    r = gs_run<double>(c, S, start.as<double>(), niter, rt, 5, x.as<double>());
    stats.steps_hi = r.out.nsteps;
  }
  hxv += r.hxv;
  c.last_ms = tm.stop();
  c.last_hxv = hxv;
  adopt_gs(c, isector, D, r.e0, std::move(x));
  gs_stats()[&c] = stats;
  if (e0) *e0 = r.e0;
  if (nsteps) *nsteps = r.out.nsteps;
  const int n = std::min<int>(niter, (int)r.out.alpha.size());
  if (alpha) {
    for (int k = 0; k < niter; ++k) alpha[k] = k < n ? r.out.alpha[k] : 0.0;
  }
  if (beta) {
    for (int k = 0; k < niter; ++k) beta[k] = k < n ? r.out.beta[k] : 0.0;
  }
  if (gs_host) {
    EDL_HIP(hipMemcpyAsync(gs_host, c.d_gs, sizeof(double) * D, hipMemcpyDeviceToHost, st));
    EDL_HIP(hipStreamSynchronize(st));
  }
  API_END
}

// Lowest nstates eigenpairs by deflated Lanczos (the role of sp_eigh at
// DIAG.f90:182-203 when lanc_nstates_sector > 1).  State m: the start vector
// and every Lanczos step projected against states 0..m-1, lowest Ritz pair
// rebuilt by the replay, then orthonormalised against the earlier states.
// Each state starts from its own SplitMix64 vector (seed + m): Lanczos sees
// a degenerate level only along the start vector's projection on it, which
// for a shared start vector is the partner already found and deflated.
template <typename T>
static void lanczos_states_impl(Ctx& c, Sector& S, int nstates, int niter, double tol, uint64_t seed, double* X,
                                double* energies, int64_t* hxv) {
  hipStream_t st = c.stream;
  const int64_t D = S.dim;
  DevBuf start(sizeof(double) * D);
  for (int m = 0; m < nstates; ++m) {
    launch_start_vector(start.as<double>(), D, seed + (uint64_t)m, st);
    double* xm = X + (size_t)m * D;
    for (int j = 0; j < m; ++j) {
      const double* xj = X + (size_t)j * D;
      launch_axpy(start.as<double>(), xj, -device_dot(c, xj, start.as<double>(), D, st), D, st);
    }
    GsRun r = gs_run<T>(c, S, start.as<double>(), niter, tol, 5, xm, X, m);
    for (int pass = 0; pass < 2; ++pass)
      for (int j = 0; j < m; ++j) {
        const double* xj = X + (size_t)j * D;
        launch_axpy(xm, xj, -device_dot(c, xj, xm, D, st), D, st);
      }
    const double n2 = device_dot(c, xm, xm, D, st);
    launch_scale(xm, D, 1.0 / std::sqrt(n2), st);
    energies[m] = r.e0;
    *hxv += r.hxv;
  }
  EDL_HIP(hipStreamSynchronize(st));
}

extern "C" int edl_lanczos_states(edl_ctx* ctx, int isector, int prec, int nstates, int niter, double tol,
                                  uint64_t seed, double* energies, double* vecs) {
  API_BEGIN
  EDL_CHECK(prec == EDL_FP64 || prec == EDL_FP32, "edl_lanczos_states: prec must be EDL_FP64 or EDL_FP32");
  EDL_CHECK(niter >= 1 && nstates >= 1 && energies, "edl_lanczos_states: bad arguments");
  Ctx& c = ctx->c;
  EDL_HIP(hipSetDevice(c.device));
  Sector& S = get_sector(c, isector);
  const int64_t D = S.dim;
  EDL_CHECK(nstates <= D, "edl_lanczos_states: more states than the sector dimension");
  DevBuf X(sizeof(double) * (size_t)D * nstates);
  EventTimer tm(c.stream);
  int64_t hxv = 0;
  if (prec == EDL_FP64)
    lanczos_states_impl<double>(c, S, nstates, niter, tol, seed, X.as<double>(), energies, &hxv);
  else
    lanczos_states_impl<float>(c, S, nstates, niter, tol, seed, X.as<double>(), energies, &hxv);
  c.last_ms = tm.stop();
  c.last_hxv = hxv;
  DevBuf x0(sizeof(double) * D);
  EDL_HIP(hipMemcpyAsync(x0.p, X.p, sizeof(double) * D, hipMemcpyDeviceToDevice, c.stream));
  if (vecs) EDL_HIP(hipMemcpyAsync(vecs, X.p, sizeof(double) * (size_t)D * nstates, hipMemcpyDeviceToHost, c.stream));
  EDL_HIP(hipStreamSynchronize(c.stream));
  adopt_gs(c, isector, D, energies[0], std::move(x0));
  API_END
}

extern "C" int edl_lanczos_stats(edl_ctx* ctx, int* steps_lo, int* steps_hi, double* e0_lo) {
  API_BEGIN
  const GsStats s = gs_stats()[&ctx->c];
  if (steps_lo) *steps_lo = s.steps_lo;
  if (steps_hi) *steps_hi = s.steps_hi;
  if (e0_lo) *e0_lo = s.e0_lo;
  API_END
}

extern "C" int edl_set_state(edl_ctx* ctx, int isector, double energy, const double* vec) {
  API_BEGIN
  EDL_CHECK(vec, "edl_set_state: null vector");
  Ctx& c = ctx->c;
  EDL_HIP(hipSetDevice(c.device));
  Sector& S = get_sector(c, isector);
  DevBuf x(sizeof(double) * S.dim);
  EDL_HIP(hipMemcpyAsync(x.p, vec, sizeof(double) * S.dim, hipMemcpyHostToDevice, c.stream));
  EDL_HIP(hipStreamSynchronize(c.stream));
  adopt_gs(c, isector, S.dim, energy, std::move(x));
  API_END
}

extern "C" int edl_get_state(edl_ctx* ctx, double* vec) {
  API_BEGIN
  Ctx& c = ctx->c;
  EDL_CHECK(c.d_gs && vec, "edl_get_state: no state installed or null output");
  EDL_HIP(hipMemcpyAsync(vec, c.d_gs, sizeof(double) * c.gs_dim, hipMemcpyDeviceToHost, c.stream));
  EDL_HIP(hipStreamSynchronize(c.stream));
  API_END
}

// target sector of a seed (getCsector / getCDGsector, SETUP.f90:265-310)
// Terms of the current of orbital iorb (1-based), OC_ELECTRONS.f90:139-197:
// for each site i of the orbital and its ring successor i+1 (the last site
// wraps to the first), both spins, H(s,i,i+1) c+_i c_i+1 where that element
// is non-zero and -conjg(H(s,i+1,i)) c+_i+1 c_i where that one is; the spin-2
// terms read Hij(Nspin,:,:).
static CurrentSeed current_terms(const Model& M, int iorb) {
  auto pack = [&](int isite, int k) { return (k == 0) ? isite : isite + k * M.Nsites[k - 1]; };  // AUX_FUNX.f90:42-53
  auto h = [&](int spin, int io, int jo) {  // Hij(spin, io, jo), 1-based
    return M.Hij[(size_t)(spin - 1) + (size_t)M.Nspin * ((size_t)(io - 1) + (size_t)M.Ns * (size_t)(jo - 1))];
  };
  CurrentSeed J{};
  const int k = iorb - 1, ns = M.Nsites[k];
  for (int isite = 1; isite <= ns; ++isite) {
    const int jsite = (isite == ns) ? 1 : isite + 1;
    const int a = pack(isite, k), b = pack(jsite, k);
    for (int spin = 1; spin <= 2; ++spin) {
      const int hs = (spin == 1) ? 1 : M.Nspin;
      const double fwd = h(hs, a, b), bwd = h(hs, b, a);
      if (fwd != 0.0) {
        EDL_CHECK(J.n < kMaxCurrentTerms, "current seed: too many terms");
        J.t[J.n++] = CurrentTerm{(uint8_t)(a - 1), (uint8_t)(b - 1), (uint8_t)spin, fwd};
      }
      if (bwd != 0.0) {
        EDL_CHECK(J.n < kMaxCurrentTerms, "current seed: too many terms");
        J.t[J.n++] = CurrentTerm{(uint8_t)(b - 1), (uint8_t)(a - 1), (uint8_t)spin, -bwd};
      }
    }
  }
  return J;
}

static int target_sector(const Ctx& c, int isector, const SeedSpec& s) {
  const int Ns = c.model.Ns;
  const int ndw = (isector - 1) % (Ns + 1), nup = ((isector - 1) / (Ns + 1)) % (Ns + 1);
  if (s.op == EDL_OP_SZ || s.op == EDL_OP_J) return isector;
  const int d = (s.op == EDL_OP_CDG) ? +1 : -1;
  int u = nup, w = ndw;
  if (s.ispin == 1) u += d; else w += d;
  if (u < 0 || w < 0 || u > Ns || w > Ns) return 0;
  return 1 + u * (Ns + 1) + w;
}

static std::map<std::tuple<int, int, int, int, int>, BtShape>& bt_cache() {
  static std::map<std::tuple<int, int, int, int, int>, BtShape> m;
  return m;
}

static std::vector<BtChain> bt_resized(std::vector<BtChain> v, const BtShape& s) {
  for (BtChain& b : v) {
    b.nslab = (b.V.dimup + s.RS - 1) / s.RS;
    b.npanel = (b.V.dimdw + s.C - 1) / s.C;
  }
  return v;
}

// Panel width, slab height and block sizes of the batched two-pass kernels,
// by timing the up and down launches on the chain set itself at step 0,
// where both are repeatable (the up pass has no pending phase B; the down
// pass multiplies q_{-1} by beta_0 = 0); the scalars are put back after.
template <typename T>
static BtShape bt_tune(Ctx& c, const std::vector<BtChain>& chains, int max_up, int max_dw, hipStream_t st) {
  const char* fk = std::getenv("EDL_BATCHED_UP_H16");
  const auto key = std::make_tuple((int)sizeof(T), max_up, max_dw, (int)chains.size(), fk ? std::atoi(fk) : -1);
  auto it = bt_cache().find(key);
  if (it != bt_cache().end()) return it->second;
  BtShape best = bt_shape(max_up, max_dw, sizeof(T));
  const size_t head = 256 * sizeof(T) + 16, budget = 150 * 1024;
  // The tile kernels are timed alone, in the state of a middle step: phase B
  // pending (pad = 1) with alpha_k = 0, so that each up launch reads t and
  // q_k and writes r = t and T as in a real step yet leaves its input as it
  // found it.  The chain scalars are restored afterwards; the vectors the
  // timing launches wrote are rewritten by the real first step before they
  // are read (q_{-1} enters step 0 with weight beta_0 = 0 and stays finite).
  std::vector<LzScalars> snap(chains.size());
  for (size_t k = 0; k < chains.size(); ++k)
    EDL_HIP(hipMemcpyAsync(&snap[k], chains[k].sc, sizeof(LzScalars), hipMemcpyDeviceToHost, st));
  EDL_HIP(hipStreamSynchronize(st));
  std::vector<LzScalars> mid(snap);
  for (LzScalars& m : mid) {
    m.pad = 1;
    m.alpha_k = 0.0;
  }
  for (size_t k = 0; k < chains.size(); ++k)
    EDL_HIP(hipMemcpyAsync(chains[k].sc, &mid[k], sizeof(LzScalars), hipMemcpyHostToDevice, st));
  auto time = [&](const BtShape& shp, bool up) {
    BtSet bs;
    bs.shape = shp;
    bs.h = bt_resized(chains, shp);
    bt_finalize<T>(bs);
    for (int r = 0; r < 3; ++r) {
      if (r == 1) EDL_HIP(hipStreamSynchronize(st));
      EventTimer tm(st);
      if (up)
        bt_launch_up<T>(bs, c.model.inter, st, false);
      else
        bt_launch_dw<T>(bs, st, false);
      const double ms = tm.stop();
      if (r == 2) return ms;
    }
    return 0.0;
  };
  double bu = 1e300, bd = 1e300;
  // EDL_BATCHED_UP_H16=0|1 keeps the tuner to one table form in both tiles
  const bool h16_ok = bt_h16_ok(chains, 0), h16_dw_ok = bt_h16_ok(chains, 1);
  const char* fh = std::getenv("EDL_BATCHED_UP_H16");
  const int h16_only = fh ? std::atoi(fh) : -1;
  for (int h16 : {0, 1}) {
    if ((h16 && !h16_ok) || (h16_only >= 0 && h16 != h16_only)) continue;
    for (int C : {1, 2, 3, 4}) {
      if (head + (size_t)C * (max_up + (h16 ? 32 : 0)) * sizeof(T) > budget) continue;
      for (int rk : {208, 204, 108})
        for (int th : {256, 512, 1024}) {
          BtShape w = best;
          w.C = C;
          w.up_rpt = rk / 100;
          w.up_kch = rk % 100;
          w.threads_up = th;
          w.up_h16 = h16;
          const double ms = time(w, true);
          if (ms < bu) {
            bu = ms;
            best.C = C;
            best.up_rpt = w.up_rpt;
            best.up_kch = w.up_kch;
            best.threads_up = th;
            best.up_h16 = h16;
          }
        }
    }
  }
  for (int h16 : {0, 1}) {
    if ((h16 && !h16_dw_ok) || (h16_only >= 0 && h16 != h16_only)) continue;
    for (int RS : {32, 16, 8, 4, 2, 1}) {
      if ((size_t)RS * sizeof(T) > 128 || bt_dw_lds_bytes(max_dw, RS, sizeof(T), h16) > budget) continue;
      for (int nc : {1, 2})
        for (int th : {256, 512, 1024}) {
          BtShape w = best;
          w.RS = RS;
          w.dw_nc = nc;
          w.threads_dw = th;
          w.dw_h16 = h16;
          const double ms = time(w, false);
          if (ms < bd) {
            bd = ms;
            best.RS = RS;
            best.dw_nc = nc;
            best.threads_dw = th;
            best.dw_h16 = h16;
          }
        }
    }
  }
  for (size_t k = 0; k < chains.size(); ++k)
    EDL_HIP(hipMemcpyAsync(chains[k].sc, &snap[k], sizeof(LzScalars), hipMemcpyHostToDevice, st));
  EDL_HIP(hipStreamSynchronize(st));
  if (verbose())
    std::fprintf(stderr,
                 "[edl] batched two-pass shape for %zu chains: C %d rpt %d kch %d%s x %d threads (%.3f ms), "
                 "RS %d nc %d%s x %d (%.3f ms)\n",
                 chains.size(), best.C, best.up_rpt, best.up_kch, best.up_h16 ? " two-hop" : "", best.threads_up, bu,
                 best.RS, best.dw_nc, best.dw_h16 ? " two-hop" : "", best.threads_dw, bd);
  bt_cache()[key] = best;
  return best;
}

// Persistent or tile launches for a batched group, and the persistent shape:
// both forms run the group's first step (no phase B pending yet) on copies of
// the chains with nmax = 1, which writes only q_1, T, alpha_0 and the
// scalars; the scalars are put back after each timing, and the real first
// step rewrites the rest before reading it.  EDL_BATCHED_KERNELS=persist or
// =tiles forces the form.  Cached per group shape.
static std::map<std::tuple<int, int, int, int>, std::pair<bool, BtShape>>& bt_persist_cache() {
  static std::map<std::tuple<int, int, int, int>, std::pair<bool, BtShape>> m;
  return m;
}

template <typename T>
static bool bt_pick_persist(Ctx& c, const std::vector<BtChain>& chains, const BtShape& tiles, int max_up, int max_dw,
                            hipStream_t st, BtShape* pshape) {
  const char* e = std::getenv("EDL_BATCHED_KERNELS");
  const bool force_p = e && std::strcmp(e, "persist") == 0, force_t = e && std::strcmp(e, "tiles") == 0;
  const auto key = std::make_tuple((int)sizeof(T), max_up, max_dw, (int)chains.size());
  auto it = bt_persist_cache().find(key);
  if (it == bt_persist_cache().end()) {
    std::vector<LzScalars> snap(chains.size());
    for (size_t k = 0; k < chains.size(); ++k)
      EDL_HIP(hipMemcpyAsync(&snap[k], chains[k].sc, sizeof(LzScalars), hipMemcpyDeviceToHost, st));
    EDL_HIP(hipStreamSynchronize(st));
    auto restore = [&]() {
      for (size_t k = 0; k < chains.size(); ++k)
        EDL_HIP(hipMemcpyAsync(chains[k].sc, &snap[k], sizeof(LzScalars), hipMemcpyHostToDevice, st));
      EDL_HIP(hipStreamSynchronize(st));  // snap must outlive the copies
    };
    auto one_step = [&](const BtShape& shp, bool persist) {
      BtSet bs;
      bs.shape = shp;
      bs.h = bt_resized(chains, shp);
      for (BtChain& b : bs.h) b.nmax = 1;
      bt_finalize<T>(bs);
      double best = 1e300;
      for (int r = 0; r < 2; ++r) {
        restore();
        EventTimer tm(st);
        if (persist) {
          bt_launch_persist<T>(bs, c.model.inter, st);
        } else {
          bt_launch_up<T>(bs, c.model.inter, st);
          bt_launch_dw<T>(bs, st);
        }
        best = std::min(best, tm.stop());
      }
      restore();
      return best;
    };
    const double tt = one_step(tiles, false);
    double tp = 1e300;
    BtShape bp = tiles;
    for (int C : {2, 4})
      for (int RS : {8, 16, 32}) {
        if (!bt_persist_shape_ok(C, RS, sizeof(T)) ||
            bt_persist_lds_bytes(max_up, max_dw, C, RS, sizeof(T)) > (size_t)150 * 1024)
          continue;
        BtShape w = tiles;
        w.C = C;
        w.RS = RS;
        w.up_rpt = 2;
        w.up_kch = 8;
        w.dw_nc = 2;
        const double ms = one_step(w, true);
        if (verbose())
          std::fprintf(stderr, "[edl] batched persistent C %d RS %d: %.3f ms per step (tiles %.3f ms)\n", C, RS, ms, tt);
        if (ms < tp) {
          tp = ms;
          bp = w;
        }
      }
    it = bt_persist_cache().emplace(key, std::make_pair(tp < tt, bp)).first;
  }
  *pshape = it->second.second;
  if (force_t) return false;
  if (force_p) return bt_persist_shape_ok(pshape->C, pshape->RS, sizeof(T));
  return it->second.first;
}

template <typename T>
static int64_t tridiag_group(Ctx& c, const std::vector<SeedSpec>& seeds, const std::vector<int>& idx,
                             const std::vector<int>& targets, int nlanc, double* alpha, double* beta, double* norm2,
                             int* nsteps, const double* host_seed) {
  hipStream_t st = c.stream;
  std::vector<ChainAlloc<T>> allocs(idx.size());
  std::vector<ChainDev> chains;
  std::vector<BtChain> btchains;
  std::vector<int> nmaxs;
  std::vector<int> which;
  int64_t maxdim = 0;
  int max_up = 1, max_dw = 1;
  bool packed = true;
  for (size_t k = 0; k < idx.size(); ++k) {
    const Sector& J = get_sector(c, targets[idx[k]]);
    maxdim = std::max<int64_t>(maxdim, J.dim);
    max_up = std::max(max_up, J.sp[0].dim);
    max_dw = std::max(max_dw, J.sp[1].dim);
    packed = packed && J.sp[0].packed_ok && J.sp[1].packed_ok;
  }
  // batched two-pass kernels when every target sector fits them; chains on
  // sectors large enough to fill the GPU alone: one at a time through the
  // two-pass or segmented step; otherwise the single-pass batch
  bool bt = packed && bt_shape(max_up, max_dw, sizeof(T)).RS > 0 && !batched_single_pass_forced();
  bool alone = false;
  if (packed && !batched_single_pass_forced() && (batched_chain_forced() || !bt)) {
    alone = true;
    for (size_t k = 0; k < idx.size(); ++k) {
      Sector& J = get_sector(c, targets[idx[k]]);
      const bool fits = two_pass_ok(J, sizeof(T)) || seg_path(J, sizeof(T));
      alone = alone && fits && (batched_chain_forced() || J.dim >= kChainAloneDim);
    }
    if (alone) bt = false;
  }
  // every buffer of the group in one allocation
  Arena ar;
  {
    size_t bytes = 0;
    for (size_t k = 0; k < idx.size(); ++k) {
      const Sector& J = get_sector(c, targets[idx[k]]);
      const int ld = (bt || alone) ? padded_ld(J.sp[0].dim) : J.sp[0].dim;
      const int64_t D = (int64_t)ld * J.sp[1].dim;
      bytes += ChainAlloc<T>::arena_bytes(D, (int)std::min<int64_t>(nlanc, J.dim));
      if (bt)
        bytes += Arena::round(sizeof(T) * D) + Arena::round(sizeof(double) * J.sp[0].dim) +
                 Arena::round(sizeof(double) * J.sp[1].dim);
    }
    ar.reserve(bytes);
  }
  // Seeds are generated one after another into tmp (stream order keeps the
  // reuse safe); their norms stay on the device until every chain is set up.
  // The partial sums serve the seed norm (J.dim elements) and the stored
  // start vector's sum of squares (ld * DimDw elements, ld >= DimUp rounded
  // up to 32 rows: up to 32x J.dim on sectors with one up string).
  DevBuf tmp(sizeof(double) * maxdim);
  const size_t nc = idx.size();
  int64_t maxred = std::max<int64_t>(maxdim, 1);
  for (size_t k = 0; k < idx.size(); ++k) {
    const Sector& J = get_sector(c, targets[idx[k]]);
    const int ld = (bt || alone) ? padded_ld(J.sp[0].dim) : J.sp[0].dim;
    maxred = std::max<int64_t>(maxred, (int64_t)ld * J.sp[1].dim);
  }
  const int64_t part_n = partials_size(maxred);
  DevBuf sums(sizeof(double) * 2 * nc), part(sizeof(double) * (size_t)part_n);
  double* d_n2 = sums.as<double>();
  double* d_ss = d_n2 + nc;
  for (size_t k = 0; k < nc; ++k) {
    const int s = idx[k];
    Sector& J = get_sector(c, targets[s]);
    if (host_seed) {
      EDL_HIP(hipMemcpyAsync(tmp.p, host_seed, sizeof(double) * J.dim, hipMemcpyHostToDevice, st));
    } else {
      const Sector& I = gs_sector_of(c, "edl_tridiag_batched");
      if (seeds[s].op == EDL_OP_J)
        launch_current_seed(current_terms(c.model, seeds[s].io), c.d_gs, J, tmp.as<double>(), st);
      else
        launch_seed(seeds[s], c.d_gs, I.sp[0].dim, J, tmp.as<double>(), st);
    }
    EDL_CHECK(partials_size(J.dim) <= part_n, "tridiag_group: partial-sum buffer smaller than the launch grid");
    device_dot_to(tmp.as<double>(), tmp.as<double>(), J.dim, part.as<double>(), d_n2 + k, st);
    const int nmax = (int)std::min<int64_t>(nlanc, J.dim);
    ChainDev cd{};
    const int ld = (bt || alone) ? padded_ld(J.sp[0].dim) : J.sp[0].dim;
    allocs[k].ar = &ar;
    init_chain<T>(c, cd, allocs[k], J, tmp.as<double>(), 0.0, nmax, st, ld, d_ss + k, part.as<double>(), part_n);
    if (bt) {
      BtChain b{};
      b.V = two_pass_view(J, ld);
      b.lut_up = lut_dev<T>(J.sp[0]);
      b.lut_dw = lut_dev<T>(J.sp[1]);
      b.q0 = cd.q0;
      b.q1 = cd.q1;
      b.nmax = nmax;
      b.brk_tol = cd.brk_tol;
      b.tb = ar.take(sizeof(T) * (size_t)ld * J.sp[1].dim);
      b.partD = static_cast<double*>(ar.take(sizeof(double) * J.sp[0].dim));  // room for any slab height
      b.partU = static_cast<double*>(ar.take(sizeof(double) * J.sp[1].dim));  // and any panel width
      b.sc = cd.sc;
      b.alpha = cd.alpha;
      b.beta = cd.beta;
      btchains.push_back(b);
    }
    chains.push_back(cd);
    nmaxs.push_back(nmax);
    which.push_back(s);
  }
  if (chains.empty()) return 0;
  {
    std::vector<LzScalars*> scp(nc);
    for (size_t k = 0; k < nc; ++k) scp[k] = chains[k].sc;
    DevBuf d_scp(sizeof(LzScalars*) * nc);
    EDL_HIP(hipMemcpyAsync(d_scp.p, scp.data(), sizeof(LzScalars*) * nc, hipMemcpyHostToDevice, st));
    launch_init_scalars(d_scp.as<LzScalars*>(), d_n2, d_ss, (int)nc, st);
    std::vector<double> n2h(nc);
    EDL_HIP(hipMemcpyAsync(n2h.data(), d_n2, sizeof(double) * nc, hipMemcpyDeviceToHost, st));
    EDL_HIP(hipStreamSynchronize(st));
    for (size_t k = 0; k < nc; ++k) norm2[which[k]] = n2h[k];
  }
  const int steps = *std::max_element(nmaxs.begin(), nmaxs.end());
  ChainSet<T> cs;
  cs.h = chains;
  BtShape shp{};
  BtSet bs;
  std::vector<BtSet> parts;
  bool persist = false;
  int nstreams = 1;
  if (bt) {
    shp = bt_tune<T>(c, btchains, max_up, max_dw, st);
    BtShape pshp{};
    persist = bt_pick_persist<T>(c, btchains, shp, max_up, max_dw, st, &pshp);
    if (persist) shp = pshp;
    nstreams = persist ? 1 : bt_streams(btchains.size());
    if (nstreams == 1) {
      bs.shape = shp;
      bs.h = bt_resized(btchains, shp);
      bt_finalize<T>(bs);
    } else {
      const std::vector<BtChain> all = bt_resized(btchains, shp);
      parts.resize(nstreams);
      for (int h = 0; h < nstreams; ++h) {
        parts[h].shape = shp;
        parts[h].h.assign(all.begin() + all.size() * h / nstreams, all.begin() + all.size() * (h + 1) / nstreams);
        bt_finalize<T>(parts[h]);
      }
    }
  }
  EventTimer run_tm(st);
  if (bt && persist) {
    bt_launch_persist<T>(bs, c.model.inter, st);
    EDL_HIP(hipStreamSynchronize(st));
  } else if (bt && nstreams > 1) {
    run_steps_btn<T>(c, parts, steps, st);
  } else if (bt) {
    run_steps_bt<T>(c, bs, steps, st);
  } else if (alone) {
    for (size_t k = 0; k < chains.size(); ++k) {
      Sector& J = get_sector(c, targets[idx[k]]);
      const bool seg = seg_path(J, sizeof(T));
      TwoPassStep<T> tps{};
      tps.cfg = chain_cfg(seg ? seg_cfg<T>(c, J) : two_pass_cfg<T>(c, J));
      if (seg) {
        seg_bounds(J, sizeof(T), tps.cfg.C);
        tps.seg = &J;
      }
      tps.V = two_pass_view(J, chains[k].ld);
      tps.lut_up = lut_dev<T>(J.sp[0]);
      tps.lut_dw = lut_dev<T>(J.sp[1]);
      tps.tmp = static_cast<T*>(scratch(c, sizeof(T) * (size_t)chains[k].ld * J.sp[1].dim));
      DevBuf pu(sizeof(double) * (seg ? phase_b_grid(tps.V) : pass_up_grid(tps.V, tps.cfg)));
      DevBuf pd(sizeof(double) * pass_dw_grid(tps.V, tps.cfg));
      tps.L.part_up = pu.as<double>();
      tps.L.part_dw = pd.as<double>();
      tps.L.sc = chains[k].sc;
      tps.L.alpha = chains[k].alpha;
      tps.L.beta = chains[k].beta;
      tps.L.nmax = nmaxs[k];
      tps.L.brk_tol = chains[k].brk_tol;
      ChainSet<T> one;
      one.h.push_back(chains[k]);
      Schedule sch;
      sch.block_threads = 256;
      sch.grid = c.num_cu * 8;
      chainset_finalize(c, one, sch);
      run_steps(c, one, nmaxs[k], 0, 0.0, st, &tps);
    }
  } else {
    Schedule sch = autotune<T>(c, chains, st);
    chainset_finalize(c, cs, sch);
    run_steps(c, cs, steps, 0, 0.0, st);
  }
  if (verbose()) {
    const double ms = run_tm.stop();
    std::fprintf(stderr, "[edl] batched group: %zu chains, %s kernels (C %d, RS %d, %d stream%s), %d steps in %.3f ms (%.1f us/step)\n",
                 chains.size(), bt ? (persist ? "persistent two-pass" : "two-pass tile") : (alone ? "one-chain two-pass" : "single-pass"), bt ? shp.C : 0,
                 bt ? shp.RS : 0, nstreams, nstreams > 1 ? "s" : "", steps, ms, 1e3 * ms / steps);
  }
  // every chain's scalars and coefficients requested, then one wait
  std::vector<ChainOut> outs(chains.size());
  std::vector<LzScalars> scs(chains.size());
  for (size_t k = 0; k < chains.size(); ++k) {
    outs[k].alpha.assign(nmaxs[k], 0.0);
    outs[k].beta.assign(nmaxs[k], 0.0);
    EDL_HIP(hipMemcpyAsync(&scs[k], cs.h[k].sc, sizeof(LzScalars), hipMemcpyDeviceToHost, st));
    EDL_HIP(hipMemcpyAsync(outs[k].alpha.data(), cs.h[k].alpha, sizeof(double) * nmaxs[k], hipMemcpyDeviceToHost, st));
    EDL_HIP(hipMemcpyAsync(outs[k].beta.data(), cs.h[k].beta, sizeof(double) * nmaxs[k], hipMemcpyDeviceToHost, st));
  }
  EDL_HIP(hipStreamSynchronize(st));
  int64_t hxv = 0;
  for (size_t k = 0; k < chains.size(); ++k) {
    ChainOut& o = outs[k];
    o.nsteps = scs[k].nsteps;
    o.norm2 = scs[k].norm2;
    const int s = which[k];
    for (int j = 0; j < nlanc; ++j) {
      alpha[(size_t)s * nlanc + j] = j < nmaxs[k] ? o.alpha[j] : 0.0;
      beta[(size_t)s * nlanc + j] = j < nmaxs[k] ? o.beta[j] : 0.0;
    }
    if (nsteps) nsteps[s] = o.nsteps;
    hxv += o.nsteps;
  }
  return hxv;
}

extern "C" int edl_tridiag(edl_ctx* ctx, int isector, int prec, const double* seed, int nlanc, double* alpha,
                           double* beta, double* norm2, int* nsteps) {
  API_BEGIN
  EDL_CHECK(seed && alpha && beta && norm2, "edl_tridiag: null argument");
  EDL_CHECK(prec == EDL_FP64 || prec == EDL_FP32, "edl_tridiag: prec must be EDL_FP64 or EDL_FP32");
  EDL_CHECK(nlanc >= 1, "edl_tridiag: nlanc must be >= 1");
  Ctx& c = ctx->c;
  EDL_HIP(hipSetDevice(c.device));
  get_sector(c, isector);
  std::vector<SeedSpec> seeds(1, SeedSpec{0, 0, 0, 0});
  std::vector<int> idx(1, 0), tg(1, isector);
  int ns = 0;
  EventTimer tm(c.stream);
  int64_t h;
  if (prec == EDL_FP64) h = tridiag_group<double>(c, seeds, idx, tg, nlanc, alpha, beta, norm2, &ns, seed);
  else h = tridiag_group<float>(c, seeds, idx, tg, nlanc, alpha, beta, norm2, &ns, seed);
  c.last_ms = tm.stop();
  c.last_hxv = h;
  if (nsteps) *nsteps = ns;
  API_END
}

extern "C" int edl_tridiag_batched(edl_ctx* ctx, int nseeds, const edl_seed_desc* seeds, int nlanc, int prec,
                                   int group, double* alpha, double* beta, double* norm2, int* nsteps,
                                   int* isector_out) {
  API_BEGIN
  EDL_CHECK(nseeds >= 0 && (nseeds == 0 || seeds), "edl_tridiag_batched: bad seed list");
  EDL_CHECK(alpha && beta && norm2, "edl_tridiag_batched: null output");
  EDL_CHECK(prec == EDL_FP64 || prec == EDL_FP32, "edl_tridiag_batched: prec must be EDL_FP64 or EDL_FP32");
  EDL_CHECK(nlanc >= 1, "edl_tridiag_batched: nlanc must be >= 1");
  Ctx& c = ctx->c;
  EDL_HIP(hipSetDevice(c.device));
  gs_sector_of(c, "edl_tridiag_batched");
  const int Ns = c.model.Ns;
  std::vector<SeedSpec> sp(nseeds);
  std::vector<int> targets(nseeds, 0), active;
  for (int s = 0; s < nseeds; ++s) {
    sp[s] = SeedSpec{seeds[s].op, seeds[s].ispin, seeds[s].io, seeds[s].jo};
    EDL_CHECK(sp[s].op >= EDL_OP_C && sp[s].op <= EDL_OP_J, "edl_tridiag_batched: bad op");
    if (sp[s].op == EDL_OP_J) {
      EDL_CHECK(sp[s].io >= 1 && sp[s].io <= c.model.Norb && sp[s].jo == 0,
                "edl_tridiag_batched: current seed needs an orbital index 1..Norb and jo = 0");
    } else {
      EDL_CHECK(sp[s].op == EDL_OP_SZ || sp[s].ispin == 1 || sp[s].ispin == 2, "edl_tridiag_batched: bad ispin");
      EDL_CHECK(sp[s].io >= 1 && sp[s].io <= Ns && sp[s].jo >= 0 && sp[s].jo <= Ns,
                "edl_tridiag_batched: site index out of range");
    }
    targets[s] = target_sector(c, c.gs_sector, sp[s]);
    if (isector_out) isector_out[s] = targets[s];
    for (int j = 0; j < nlanc; ++j) alpha[(size_t)s * nlanc + j] = beta[(size_t)s * nlanc + j] = 0.0;
    norm2[s] = 0.0;
    if (nsteps) nsteps[s] = 0;
    if (targets[s] != 0) active.push_back(s);
  }
  for (int s : active) get_sector(c, targets[s]);
  // group = 0: as many chains per launch as fit in 70% of the free HBM
  // (each chain holds two stored vectors; one fp64 seed buffer is shared).
  int g = (group <= 0) ? (int)active.size() : group;
  if (group <= 0 && !active.empty()) {
    size_t free_b = 0, total_b = 0;
    EDL_HIP(hipMemGetInfo(&free_b, &total_b));
    int64_t maxdim = 0;
    for (int s : active) maxdim = std::max<int64_t>(maxdim, get_sector(c, targets[s]).dim);
    // q0, q1 and the two-pass intermediate, with the padded column stride
    const double per_chain = 3.0 * 1.03 * (double)maxdim * (prec == EDL_FP64 ? 8.0 : 4.0);
    const double budget = 0.7 * (double)free_b - 8.0 * (double)maxdim;
    const int fit = (int)std::max(1.0, budget / per_chain);
    g = std::min<int>(g, fit);
  }
  EventTimer tm(c.stream);
  int64_t h = 0;
  for (size_t b = 0; b < active.size(); b += g) {
    std::vector<int> idx(active.begin() + b, active.begin() + std::min(active.size(), b + g));
    if (prec == EDL_FP64)
      h += tridiag_group<double>(c, sp, idx, targets, nlanc, alpha, beta, norm2, nsteps, nullptr);
    else
      h += tridiag_group<float>(c, sp, idx, targets, nlanc, alpha, beta, norm2, nsteps, nullptr);
  }
  c.last_ms = tm.stop();
  c.last_hxv = h;
  API_END
}

extern "C" int edl_lanczos_poles(int nlanc, const double* alpha, const double* beta, double norm2, double Ei,
                                 int isign, double* poles, double* weights) {
  try {
    EDL_CHECK(nlanc >= 1 && alpha && beta && poles && weights, "edl_lanczos_poles: bad arguments");
    std::vector<double> e(nlanc, 0.0), w(nlanc), z1(nlanc);
    for (int i = 0; i + 1 < nlanc; ++i) e[i] = beta[i + 1];  // subdiag(2:Nlanc) = blanc(2:Nlanc)
    tridiag_eig(nlanc, alpha, e.data(), w.data(), nullptr, z1.data());
    for (int j = 0; j < nlanc; ++j) {
      weights[j] = z1[j] * z1[j] * norm2;
      poles[j] = isign * (w[j] - Ei);
    }
    return 0;
  } catch (const std::exception& e) {
    g_create_error = e.what();
    return 1;
  }
}

extern "C" int edl_observables(edl_ctx* ctx, double* dens_up, double* dens_dw, double* docc) {
  API_BEGIN
  Ctx& c = ctx->c;
  EDL_CHECK(dens_up && dens_dw && docc, "edl_observables: null output");
  EDL_HIP(hipSetDevice(c.device));
  Sector& S = gs_sector_of(c, "edl_observables");
  device_observables(c, S, c.d_gs, dens_up, dens_dw, docc);
  API_END
}

extern "C" int edl_correlations(edl_ctx* ctx, double* sz2, double* n2) {
  API_BEGIN
  Ctx& c = ctx->c;
  EDL_CHECK(sz2 && n2, "edl_correlations: null output");
  EDL_HIP(hipSetDevice(c.device));
  Sector& S = gs_sector_of(c, "edl_correlations");
  device_correlations(c, S, c.d_gs, sz2, n2);
  API_END
}

extern "C" int edl_local_energy(edl_ctx* ctx, double* e) {
  API_BEGIN
  Ctx& c = ctx->c;
  EDL_CHECK(e, "edl_local_energy: null output");
  EDL_HIP(hipSetDevice(c.device));
  Sector& S = gs_sector_of(c, "edl_local_energy");
  device_energy(c, S, c.d_gs, e);
  API_END
}

extern "C" int edl_last_timing(edl_ctx* ctx, double* ms, int64_t* hxv_count) {
  API_BEGIN
  if (ms) *ms = ctx->c.last_ms;
  if (hxv_count) *hxv_count = ctx->c.last_hxv;
  API_END
}

template <typename T>
static double bench_impl(Ctx& c, Sector& s, int reps) {
  hipStream_t st = c.stream;
  const int64_t D = s.dim;
  // the engine's own layout (padded columns on two-pass sectors)
  const int ld = engine_ld<T>(s);
  const int64_t Dp = (int64_t)ld * s.sp[1].dim;
  DevBuf d64(sizeof(double) * D), v(sizeof(T) * Dp), h(sizeof(T) * Dp);
  launch_start_vector(d64.as<double>(), D, 0x5EEDull, st);
  if (ld != s.sp[0].dim) launch_zero_bytes(v.p, sizeof(T) * Dp, st);
  launch_convert_ld<T>(d64.as<double>(), v.as<T>(), s.sp[0].dim, s.sp[1].dim, ld, st);
  apply_h<T>(c, s, v.as<T>(), h.as<T>(), ld, st);
  EventTimer tm(st);
  for (int r = 0; r < reps; ++r) apply_h<T>(c, s, v.as<T>(), h.as<T>(), ld, st);
  return tm.stop() / reps;
}

extern "C" int edl_bench_hxv(edl_ctx* ctx, int isector, int prec, int reps, double* ms_per_hxv) {
  API_BEGIN
  EDL_CHECK(prec == EDL_FP64 || prec == EDL_FP32, "edl_bench_hxv: prec must be EDL_FP64 or EDL_FP32");
  EDL_CHECK(reps >= 1 && ms_per_hxv, "edl_bench_hxv: bad arguments");
  Ctx& c = ctx->c;
  EDL_HIP(hipSetDevice(c.device));
  Sector& s = get_sector(c, isector);
  *ms_per_hxv = (prec == EDL_FP64) ? bench_impl<double>(c, s, reps) : bench_impl<float>(c, s, reps);
  c.last_ms = *ms_per_hxv * reps;
  c.last_hxv = reps;
  API_END
}

// Mean duration of each kernel launch of the operator, bracketed by events.
template <typename T>
static void bench_passes_impl(Ctx& c, Sector& s, int reps, double* ms_pass, int32_t* cfg) {
  hipStream_t st = c.stream;
  const bool seg = seg_path(s, sizeof(T));
  const bool two = !seg && two_pass_ok(s, sizeof(T));
  const int ld = engine_ld<T>(s);
  const int64_t D = s.dim, Dp = (int64_t)ld * s.sp[1].dim;
  DevBuf d64(sizeof(double) * D), v(sizeof(T) * Dp), h(sizeof(T) * Dp);
  launch_start_vector(d64.as<double>(), D, 0x5EEDull, st);
  if (ld != s.sp[0].dim) launch_zero_bytes(v.p, sizeof(T) * Dp, st);
  launch_convert_ld<T>(d64.as<double>(), v.as<T>(), s.sp[0].dim, s.sp[1].dim, ld, st);
  apply_h<T>(c, s, v.as<T>(), h.as<T>(), ld, st);
  std::vector<hipEvent_t> ev(3 * (size_t)reps);
  for (auto& e : ev) EDL_HIP(hipEventCreate(&e));
  TwoPassCfg cf{};
  if (seg) {
    cf = seg_cfg<T>(c, s);
    seg_bounds(s, sizeof(T), cf.C);
    const TwoPassView V = two_pass_view(s, ld);
    T* tmp = static_cast<T*>(scratch(c, sizeof(T) * (size_t)Dp));
    for (int r = 0; r < reps; ++r) {
      EDL_HIP(hipEventRecord(ev[3 * r], st));
      launch_up_seg<T>(V, c.model.inter, lut_dev<T>(s.sp[0]), v.as<T>(), v.as<T>(), tmp, nullptr, s, cf.C, cf.seg_rank,
                       st);
      launch_nonlocal<T>(V, c.model.inter, v.as<T>(), v.as<T>(), tmp, nullptr, st);
      EDL_HIP(hipEventRecord(ev[3 * r + 1], st));
      launch_pass_dw<T>(V, lut_dev<T>(s.sp[1]), v.as<T>(), v.as<T>(), tmp, h.as<T>(), nullptr, cf, st);
      EDL_HIP(hipEventRecord(ev[3 * r + 2], st));
    }
  } else if (two) {
    cf = two_pass_cfg<T>(c, s);
    const TwoPassView V = two_pass_view(s, ld);
    T* tmp = static_cast<T*>(scratch(c, sizeof(T) * (size_t)Dp));
    for (int r = 0; r < reps; ++r) {
      EDL_HIP(hipEventRecord(ev[3 * r], st));
      launch_pass_up<T>(V, c.model.inter, lut_dev<T>(s.sp[0]), v.as<T>(), v.as<T>(), tmp, nullptr, cf, st);
      launch_nonlocal<T>(V, c.model.inter, v.as<T>(), v.as<T>(), tmp, nullptr, st);
      EDL_HIP(hipEventRecord(ev[3 * r + 1], st));
      launch_pass_dw<T>(V, lut_dev<T>(s.sp[1]), v.as<T>(), v.as<T>(), tmp, h.as<T>(), nullptr, cf, st);
      EDL_HIP(hipEventRecord(ev[3 * r + 2], st));
    }
    const void* bufs[3] = {v.p, tmp, h.p};
    if (!pass_dw_vec_ok(V, cf, sizeof(T), bufs, 3)) cf.dw_vec = 0;  // the form that ran
  } else {
    for (int r = 0; r < reps; ++r) {
      EDL_HIP(hipEventRecord(ev[3 * r], st));
      apply_h<T>(c, s, v.as<T>(), h.as<T>(), ld, st);
      EDL_HIP(hipEventRecord(ev[3 * r + 1], st));
      EDL_HIP(hipEventRecord(ev[3 * r + 2], st));
    }
  }
  EDL_HIP(hipStreamSynchronize(st));
  double up = 0.0, dw = 0.0;
  for (int r = 0; r < reps; ++r) {
    float a = 0.f, b = 0.f;
    EDL_HIP(hipEventElapsedTime(&a, ev[3 * r], ev[3 * r + 1]));
    EDL_HIP(hipEventElapsedTime(&b, ev[3 * r + 1], ev[3 * r + 2]));
    up += a;
    dw += b;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  ms_pass[0] = up / reps;
  ms_pass[1] = (two || seg) ? dw / reps : 0.0;
  // the up pass takes its coefficients from registers in fp32 on one-value
  // lattices, and with a two-hop table (up_col 2 or 3) in either precision
  const int32_t lreg = (two && s.sp[0].nvals == 1 && (sizeof(T) == 4 || cf.up_col >= 2)) ? 1 : 0;
  const int32_t out[16] = {two ? 1 : (seg ? 2 : 0), ld,      cf.C,    cf.threads_up, cf.grid_up, cf.R,
                           cf.threads_dw,            cf.nchunk, cf.ncol, cf.kch,        cf.up_rpt,  cf.up_kch,
                           cf.dw_vec ? 1 + cf.dw_ord : 0, lreg, two ? cf.up_col : cf.seg_rank, cf.dw_rng};
  for (int k = 0; k < 16; ++k) cfg[k] = (two || seg) ? out[k] : (k < 2 ? out[k] : 0);
}

extern "C" int edl_bench_passes(edl_ctx* ctx, int isector, int prec, int reps, double* ms_pass, int32_t* cfg) {
  API_BEGIN
  EDL_CHECK(prec == EDL_FP64 || prec == EDL_FP32, "edl_bench_passes: prec must be EDL_FP64 or EDL_FP32");
  EDL_CHECK(reps >= 1 && ms_pass && cfg, "edl_bench_passes: bad arguments");
  Ctx& c = ctx->c;
  EDL_HIP(hipSetDevice(c.device));
  Sector& s = get_sector(c, isector);
  if (prec == EDL_FP64)
    bench_passes_impl<double>(c, s, reps, ms_pass, cfg);
  else
    bench_passes_impl<float>(c, s, reps, ms_pass, cfg);
  API_END
}
