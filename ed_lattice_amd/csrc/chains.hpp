// Device-resident Lanczos chains.  A chain is one three-term Lanczos run
// (a ground-state solve or one GF/chi seed); a ChainSet advances any number
// of chains, possibly on sectors of different dimension, with one launch
// per phase.
#pragma once
#include "edl_internal.hpp"

namespace edl {

// Scalars of one chain, kept on the device so that a step needs no host
// round trip and can be replayed from a captured graph.
struct LzScalars {
  int k;         // current step, 0-based
  int stopped;   // 1 once the chain has ended (all steps done or breakdown)
  int nsteps;    // number of valid alpha entries
  int pad;
  double s_cur;  // norm of the stored vector q_k (v_k = q_k / s_cur)
  double s_old;  // norm of the stored vector q_{k-1}
  double beta_k; // coupling between v_{k-1} and v_k (0 at k = 0)
  double alpha_k;
  double norm2;  // squared norm of the seed before normalisation
};

// One chain as the kernels see it.  Its work is cut into tiles of one
// column (idw) each; tile partial sums are reduced in column order by the
// block that completes the last tile, so results do not depend on which
// block ran which tile.
struct ChainDev {
  // sector view
  int dimup, dimdw, kup, kdw;
  const int32_t* map_up;
  const int32_t* map_dw;
  const int32_t* nbr_up;   // [kup][dimup]
  const int32_t* nbr_dw;   // [kdw][dimdw]
  const int32_t* cnt_up;   // exact hop counts
  const int32_t* cnt_dw;
  const void* cu;          // hop coefficients, scalar type of the chain
  const void* cd;
  const double* dg_up;
  const double* dg_dw;
  // schedule
  int64_t tile0;           // first global tile of the chain
  int ntiles;              // = dimdw
  // state
  int ld;                  // column stride of q0/q1 (>= dimup)
  void* q0;                // stored vectors; q_k lives in q[k & 1]
  void* q1;
  double* x;               // fp64 eigenvector accumulator (replay, natural layout), or null
  const double* ccoef;     // replay coefficients c_k, or null
  double* part;            // [ntiles] tile partial sums
  unsigned* counter;       // tiles completed in the current launch
  LzScalars* sc;
  double* alpha;           // [nmax]
  double* beta;            // [nmax]; beta[0] = 0
  int nmax;
  double brk_tol;          // relative breakdown threshold on beta
};

struct Schedule {
  int block_threads = kBlock;
  int grid = 0;            // blocks; a multiple of 8 uses the XCD-synchronous order
  int64_t total_tiles = 0;
};

template <typename T>
struct ChainSet {
  std::vector<ChainDev> h;
  ChainDev* d = nullptr;
  int64_t* d_tile0 = nullptr;  // [nchains+1] tile offsets
  Schedule sch;
  // owned device allocations
  std::vector<void*> owned;
  ~ChainSet();
};

InterParams& inter_of(Ctx& c);

// Fill the per-chain schedule fields and upload the descriptors.
template <typename T>
void chainset_finalize(Ctx& c, ChainSet<T>& cs, const Schedule& want);

template <typename T>
void lz_phase_a(Ctx& c, ChainSet<T>& cs, hipStream_t st);
template <typename T>
void lz_phase_b(Ctx& c, ChainSet<T>& cs, hipStream_t st);

// Plain operator out = H v on one sector (directMatVec), single pass.
template <typename T>
void hxv_operator(Ctx& c, const Sector& s, const T* v, T* out, hipStream_t st, const Schedule& want);

// helpers
void fill_sector_view(const Sector& s, bool fp32, ChainDev& cd);
void launch_start_vector(double* d_x, int64_t n, uint64_t seed, hipStream_t st);
template <typename T>
void launch_convert(const double* src, T* dst, int64_t n, hipStream_t st);
// src in the natural layout (rows x cols); dst with column stride ld.
template <typename T>
void launch_convert_ld(const double* src, T* dst, int rows, int cols, int ld, hipStream_t st);
template <typename T>
void launch_to_fp64(const T* src, double* dst, int64_t n, hipStream_t st);
// src with column stride ld; dst in the natural layout (rows x cols).
template <typename T>
void launch_to_fp64_ld(const T* src, double* dst, int rows, int cols, int ld, hipStream_t st);
double device_dot(Ctx& c, const double* a, const double* b, int64_t n, hipStream_t st);
template <typename T>
double device_dot_t(const T* a, const T* b, int64_t n, hipStream_t st);
void launch_scale(double* x, int64_t n, double s, hipStream_t st);
// y += a * x
void launch_axpy(double* y, const double* x, double a, int64_t n, hipStream_t st);
// Device-side reductions into device memory (no host round trip); `part`
// holds partials_size(n) doubles.
int partials_size(int64_t n);
void device_dot_to(const double* a, const double* b, int64_t n, double* part, double* out, hipStream_t st);
template <typename T>
void device_sumsq_to(const T* a, int64_t n, double* part, double* out, hipStream_t st);
// *d_sc[k] = initial Lanczos scalars from n2[k], ss[k] (k < n).
void launch_init_scalars(LzScalars* const* d_sc, const double* n2, const double* ss, int n, hipStream_t st);
void launch_zero_bytes(void* p, size_t bytes, hipStream_t st);

}  // namespace edl
