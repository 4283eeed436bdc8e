/*
 * edlat_gpu.h -- C-ABI of libedlat_hip.so, the MI355X (gfx950) engine for
 * the on-the-fly H x v + Lanczos path of aamaricci/ED_Lattice (EDlattice
 * 1.6.0).  Plain pointers and sizes only; callable from Fortran through
 * iso_c_binding (see ed_lattice_amd/fortran/ed_gpu.f90).
 *
 * Conventions
 *   - Every function returns 0 on success and a non-zero status otherwise;
 *     edl_last_error() then describes the failure.  The reference stops
 *     (`stop "..."`) where these return an error, so a Fortran caller turns
 *     a non-zero status into `stop edl_last_error()`.
 *   - Sector labels are the reference's 1-based isector = 1 + Nup*(Ns+1) + Ndw
 *     (SECTOR.f90:370-381); site indices io, jo are 1-based.
 *   - Vectors are in the reference basis order i = iup + (idw-1)*DimUp
 *     (SECTOR.f90:443-480), up index fastest, 0-based in C.
 *   - Hij is the Fortran array Hij(Nspin,Ns,Ns) in column-major order.
 *     The kernels compute in real arithmetic: a non-zero imaginary part is
 *     refused.  Complex vectors of a real Hamiltonian are handled by
 *     edl_hxv_host_complex (real and imaginary parts separately).
 *   - Dimensions are 64-bit: sectors beyond 2^31-1 states are supported.
 *   - One context owns one HIP stream and all device memory it allocates.
 *     A context is not thread-safe (the reference is single-threaded per
 *     process).
 */
#ifndef EDLAT_GPU_H
#define EDLAT_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct edl_ctx edl_ctx;

/* arithmetic of vectors and H x v */
enum { EDL_FP64 = 0, EDL_FP32 = 1, EDL_MIXED = 2 };
/* seed operators */
enum { EDL_OP_C = 1, EDL_OP_CDG = 2, EDL_OP_SZ = 3, EDL_OP_J = 4 };

/* Context on HIP device `device`.  Replaces the module-global state of
 * ED_HAMILTONIAN_COMMON (HAMILTONIAN_COMMON.f90:13-40). */
int edl_ctx_create(edl_ctx** ctx, int device);
int edl_ctx_destroy(edl_ctx* ctx);
const char* edl_last_error(const edl_ctx* ctx);

/* Hamiltonian parameters: Hij as returned by Hij_get (GRAPH_MATRIX.f90:334-345)
 * plus the input variables read by HxV_local / HxV_up / HxV_dw /
 * HxV_non_local (INPUT_VARS.f90:106-166).  With Norb > 1 and Jx or Jp
 * non-zero (Jhflag, SETUP.f90:35-36) the spin-exchange and pair-hopping
 * block is applied too; such models run on the single-pass kernels.
 * Hij must be Hermitian to 1e-6 (GRAPH_MATRIX.f90:324-326). */
int edl_set_hamiltonian(edl_ctx* ctx, int Ns, int Nspin, int Norb, const int* Nsites, const double* Hij_re,
                        const double* Hij_im, const double* Uloc, double Ust, double Jh, double Jx, double Jp,
                        double xmu, int hfmode, int ed_filling);

/* Sector label maths (get_Sector / get_Nup / get_Ndw, SECTOR.f90:370-424). */
int edl_sector_index(int Ns, int nup, int ndw);
int edl_sector_qn(int Ns, int isector, int* nup, int* ndw);

/* build_sector (SECTOR.f90:69-115): enumerates both spin lists on the device
 * and builds the hopping tables.  Returns Dim, DimUp, DimDw. */
int edl_sector_build(edl_ctx* ctx, int isector, int64_t* dim, int32_t* dimup, int32_t* dimdw);
/* Copies of the device maps H(1)%map, H(2)%map (SECTOR.f90:94-109). */
int edl_sector_maps(edl_ctx* ctx, int isector, int32_t* map_up, int32_t* map_dw);
/* Gather-form hopping table of one spin (ispin 1 or 2): kmax, then
 * nbr[kmax*dimσ] (0-based source index) and coef[kmax*dimσ] (Hij*sign),
 * k-major.  Either array may be NULL to query kmax only. */
int edl_sector_hops(edl_ctx* ctx, int isector, int ispin, int32_t* kmax, int32_t* nbr, double* coef);
/* binary_search (AUX_FUNX.f90:163-180) of n strings in the spin-ispin map:
 * 1-based positions, 0 where a string is not in the map. */
int edl_sector_rank(edl_ctx* ctx, int isector, int ispin, int64_t n, const int32_t* states, int32_t* pos);

/* directMatVec (HAMILTONIAN_DIRECT_HxV.f90:19-59), the dd_sparse_HxV
 * operator of VARS_GLOBAL.f90:102-108: hv = H v.  Device pointers to
 * double (EDL_FP64) or float (EDL_FP32) arrays of Dim elements. */
int edl_hxv(edl_ctx* ctx, int isector, int prec, const void* v_dev, void* hv_dev);
/* edl_hxv without any allocation -- no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The two-pass
 * intermediate lives in a file-scope device array of 256 MiB that every GPU
 * the library runs on holds until the process ends; sectors whose
 * intermediate (DimUp * DimDw elements of `prec`) exceeds it, sectors not
 * built yet (edl_sector_build first) and sectors outside the two-pass
 * kernels are refused with an error.  The schedule is the one the tuner
 * cached for the sector shape (after an edl_hxv of that shape), else a fixed
 * one.  Returns when hv is written.  Calls that overlap on one GPU share
 * the array and race. */
int edl_hxv_noalloc(edl_ctx* ctx, int isector, int prec, const void* v_dev, void* hv_dev);
/* Same with host FP64 arrays (converted to `prec` on the device). */
int edl_hxv_host(edl_ctx* ctx, int isector, int prec, const double* v, double* hv);
/* Same with host complex(8) arrays, interleaved (re, im). */
int edl_hxv_host_complex(edl_ctx* ctx, int isector, int prec, const double* v, double* hv);
/* Device vector utilities for edl_hxv callers (no reference counterpart: the
 * reference keeps its vectors on the host): fill with the SplitMix64 start
 * vector U[-1,1) of `seed`, and the FP64-accumulated dot product. */
int edl_vec_random(edl_ctx* ctx, int prec, void* dev, int64_t n, uint64_t seed);
int edl_vec_dot(edl_ctx* ctx, int prec, const void* a_dev, const void* b_dev, int64_t n, double* out);
/* Device buffer owned by the context (for edl_hxv callers); freed with it. */
int edl_device_alloc(edl_ctx* ctx, int64_t bytes, void** dev_ptr);
int edl_copy_h2d(edl_ctx* ctx, void* dst_dev, const void* src_host, int64_t bytes);
int edl_copy_d2h(edl_ctx* ctx, void* dst_host, const void* src_dev, int64_t bytes);

/* Ground state by plain Lanczos: replaces sp_lanc_eigh at DIAG.f90:206-218.
 * Start vector: SplitMix64(seed) mapped to U[-1,1) per element.  Runs niter
 * steps (fewer on breakdown, or when tol > 0 and |dE0| < tol*|E0| between
 * checks).  prec EDL_MIXED: FP32 vectors and H x v with FP64 alpha/beta,
 * then an FP64 Lanczos restarted from the FP32 Ritz vector until
 * |dE0| <= refine_tol*|E0|.  The normalised ground state stays on the
 * device (the eigenspace entry GF/chi seeds start from); gs_host (Dim
 * doubles) receives a copy when non-NULL.  alpha/beta (niter each, beta[0]
 * = 0) receive the tridiagonal of the last Lanczos run. */
int edl_lanczos_gs(edl_ctx* ctx, int isector, int prec, int niter, double tol, uint64_t seed, double refine_tol,
                   double* e0, int* nsteps, double* alpha, double* beta, double* gs_host);
/* Lowest nstates eigenpairs of a sector (the role of sp_eigh / ARPACK at
 * DIAG.f90:182-203 when lanc_nstates_sector > 1): state m is the lowest
 * Ritz pair of a Lanczos run whose start vector and every step are
 * projected against states 0..m-1 (deflation; degenerate states come out
 * one by one: state m starts from the SplitMix64 vector of seed + m).  Each run takes up to niter steps, fewer when |dE| < tol*|E|
 * between checks.  energies[nstates] ascending as found; vecs (nstates*Dim
 * doubles, row m = state m, normalised) when non-NULL.  State 0 is
 * installed as by edl_lanczos_gs.  Two-pass or segmented sectors only (the
 * packed hop tables). */
int edl_lanczos_states(edl_ctx* ctx, int isector, int prec, int nstates, int niter, double tol, uint64_t seed,
                       double* energies, double* vecs);
/* Statistics of the last edl_lanczos_gs: FP32 steps, FP64 steps, and the
 * FP32 energy before refinement. */
int edl_lanczos_stats(edl_ctx* ctx, int* steps_lo, int* steps_hi, double* e0_lo);

/* Install a ground state from the host (an es_add_state equivalent,
 * EIGENSPACE.f90:167-190) or copy the installed one back. */
int edl_set_state(edl_ctx* ctx, int isector, double energy, const double* vec);
int edl_get_state(edl_ctx* ctx, double* vec);

/* tridiag_Hv_sector (HAMILTONIAN.f90:242-275): norm2 = <seed|seed>, seed
 * normalised, Nlanc steps of sp_lanc_tridiag; alpha[Nlanc], beta[Nlanc]
 * zero-initialised, beta[0] unused; nothing runs when norm2 == 0. */
int edl_tridiag(edl_ctx* ctx, int isector, int prec, const double* seed, int nlanc, double* alpha, double* beta,
                double* norm2, int* nsteps);

/* Batched seeds from the installed ground state.  Seed s is
 * {op, ispin, io, jo}: op EDL_OP_C / EDL_OP_CDG acting on spin ispin in the
 * N-1 / N+1 sector (getCsector / getCDGsector, SETUP.f90:265-310), or
 * EDL_OP_SZ in the same sector; jo > 0 adds the jo term (the mixed seeds of
 * GF_ELECTRONS.f90:303-405, CHI_SPIN_ELECTRONS.f90:244-293).  EDL_OP_J is
 * the optical-conductivity current of orbital io (1-based; ispin and jo
 * unused), lanc_ed_build_oc (OC_ELECTRONS.f90:103-213), in the same sector;
 * its poles are E_j - Ei (edl_lanczos_poles with isign = +1).  Each chain
 * runs Nlanc = min(Dim_target, nlanc) steps (SECTOR.f90:112).  Replaces the
 * sequential per-seed loops of GF_ELECTRONS.f90:202-405 and
 * CHI_SPIN_ELECTRONS.f90:189-293.  Outputs: alpha, beta [nseeds*nlanc]
 * (row s), norm2[nseeds], nsteps[nseeds], isector_out[nseeds] (0 where the
 * target sector does not exist).  group = seeds advanced per launch (0 = as
 * many as fit in 70% of the free HBM, i.e. all of them unless the vectors are
 * very large; 1 = one seed at a time, for comparison). */
typedef struct {
  int op, ispin, io, jo;
} edl_seed_desc;
int edl_tridiag_batched(edl_ctx* ctx, int nseeds, const edl_seed_desc* seeds, int nlanc, int prec, int group,
                        double* alpha, double* beta, double* norm2, int* nsteps, int* isector_out);

/* add_to_lanczos_gf_normal (GF_ELECTRONS.f90:417-457) and
 * add_to_lanczos_spinChi (CHI_SPIN_ELECTRONS.f90:305-366) without the
 * storage: eigen-decompose the Nlanc x Nlanc tridiagonal (alpha, beta[1:]),
 * then weight_j = norm2 * Z(1,j)^2 and pole_j = isign * (E_j - Ei).  Host
 * only (no device needed). */
int edl_lanczos_poles(int nlanc, const double* alpha, const double* beta, double norm2, double Ei, int isign,
                      double* poles, double* weights);

/* lanc_observables (OBSERVABLES.f90:78-192) of the installed ground state:
 * dens_up, dens_dw, docc per site (weight 1). */
int edl_observables(edl_ctx* ctx, double* dens_up, double* dens_dw, double* docc);

/* The two-site correlations of lanc_observables (OBSERVABLES.f90:142-150,
 * written to Sz_corr.ed and N_corr.ed at :337-347) of the installed state
 * (weight 1): sz2[i + j*Ns] = <Sz_i Sz_j>, n2[i + j*Ns] = <n_i n_j>. */
int edl_correlations(edl_ctx* ctx, double* sz2, double* n2);

/* lanc_energy (ENERGY.f90:70-353) of the installed ground state (weight 1):
 * e[11] in the column order of energy_last.ed (ENERGY.f90:654-657): <K>,
 * <Hi>, <V> = <Hi> - <Ehf>, <Eloc>, <Ehf>, Dust, Dund, Dse, Dph, Dkxy, Dkz
 * (Dkxy, Dkz: 0, as the reference's Lanczos path leaves them). */
int edl_local_energy(edl_ctx* ctx, double* e);

/* Device time (HIP events on the context stream) of the last edl_hxv,
 * edl_lanczos_gs, edl_tridiag or edl_tridiag_batched call, and the number
 * of H x v applications it performed. */
int edl_last_timing(edl_ctx* ctx, double* ms, int64_t* hxv_count);

/* Time `reps` back-to-back operator applications on device-resident
 * vectors of `isector` (first call untimed); returns the mean ms per H x v
 * measured with HIP events on the launch stream. */
int edl_bench_hxv(edl_ctx* ctx, int isector, int prec, int reps, double* ms_per_hxv);

/* Per-kernel split of the same operator: ms_pass[0] and ms_pass[1] are the
 * mean durations of the up-pass and down-pass launches (HIP events around
 * each launch on the launch stream; a single-pass sector reports its one
 * kernel in ms_pass[0] and 0 in ms_pass[1]).  cfg[16] receives the tuned
 * schedule: {two_pass (1 = LDS column panels, 2 = segmented up pass for
 * columns larger than LDS, 0 = single pass), column stride, up columns per
 * block, up threads, up grid, down panel rows, down threads, down column
 * chunks, down columns per lane, down hop chunk, up rows per thread, up hop
 * chunk, down-pass gathers of 16 bytes per lane (1; 2: the same with the
 * column chunks taken in the down spin's locality order) or one element per
 * lane (0), up-pass hop coefficients from registers (1: fp32, one distinct
 * hopping value) or from the LDS table (0), up-pass hops in the bank-scheduled
 * order (1: a hop to a table word, 2: two hops to a word, 4: two hops to a
 * word scheduled separately for each sign of h; 5, 6: forms 2, 4 with each
 * block's next panel loaded into registers before the barrier that ends the
 * current one) or in table order (0: a hop to a word, 3: two hops to a
 * word) -- on segmented sectors (cfg[0]
 * = 2) this slot holds the form of the segmented up pass instead: 1 = hop
 * targets by rank arithmetic, 0 = the hop table --, down pass over column
 * ranges of that many top sites (b > 0, range-blocked form) or over column
 * chunks (0)}. */
int edl_bench_passes(edl_ctx* ctx, int isector, int prec, int reps, double* ms_pass, int32_t* cfg);

#ifdef __cplusplus
}
#endif
#endif
