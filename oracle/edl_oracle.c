/*
 * edl_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's on-the-fly H x v path of
 * aamaricci/ED_Lattice ("EDlattice" 1.6.0).  It is the checker for the HIP
 * library (tests, __graft_entry__.smoke) and the CPU baseline timed by
 * bench.py.  No product code links, loads or calls it.
 *
 * The reference is Fortran that needs SciFortran (absent here), so it cannot
 * be compiled from its own sources; this file restates its algorithm in C,
 * keeping the loop order, the per-element work and the complex(8) arithmetic
 * of each routine it follows:
 *
 *   binomial            src/SECTOR.f90:657-675   (real(8) product, rounded)
 *   build_sector        src/SECTOR.f90:69-115    (2^Ns scan, ascending maps)
 *   state2indices etc.  src/SECTOR.f90:443-480
 *   pack_indices        src/AUX_FUNX.f90:42-53
 *   c, cdg              src/AUX_FUNX.f90:76-102  (bit-loop fermionic sign)
 *   bdecomp             src/AUX_FUNX.f90:132-142
 *   binary_search       src/AUX_FUNX.f90:163-180 (recursive, 1-based, 0 = miss)
 *   directMatVec        src/DIAG/HAMILTONIAN_DIRECT_HxV.f90:19-59
 *     HxV_local         src/DIAG/direct/HxV_local.f90:1-97
 *     HxV_non_local     src/DIAG/direct/HxV_non_local.f90:1-88
 *     HxV_up            src/DIAG/direct/HxV_up.f90:1-34
 *     HxV_dw            src/DIAG/direct/HxV_dw.f90:1-32
 *   apply_op_C/CDG/Sz   src/SECTOR.f90:217-318
 *   lanc_observables    src/OBSERVABLES.f90:78-192 (dens_up, dens_dw, docc)
 *
 * Conventions at this C boundary: Fortran positions (sites, map indices) stay
 * 1-based inside the loops exactly as in the reference; arrays passed in are
 * 0-based C arrays.  Hij is the Fortran array Hij(Nspin,Ns,Ns) in column-major
 * order, split into real and imaginary parts.  Complex vectors are interleaved
 * (re, im) doubles.  Default-integer Dim is 32-bit in the reference; sectors
 * with Dim > 2^31-1 are refused here as the reference cannot represent them.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex zc;

#define OR_MAXORB 8
#define OR_MAXNS 30

/* ------------------------------------------------------------------ model */
typedef struct {
  int Ns, Nspin, Norb;
  int Nsites[OR_MAXORB];
  const double *Hre, *Him; /* Hij(Nspin,Ns,Ns), Fortran order */
  double Uloc[OR_MAXORB];
  double Ust, Jh, Jx, Jp, xmu;
  int hfmode, ed_filling;
} or_model;

static inline zc Hij_at(const or_model *m, int spin, int io, int jo) {
  /* Hij(spin,io,jo), all 1-based, column-major [Nspin,Ns,Ns] */
  long k = (long)(spin - 1) + (long)m->Nspin * ((io - 1) + (long)m->Ns * (jo - 1));
  return m->Hre[k] + I * (m->Him ? m->Him[k] : 0.0);
}

/* SECTOR.f90:657-675 */
int or_binomial(int n1, int n2) {
  double xh = 1.0;
  if (n2 < 0) return 0;
  if (n2 == 0) return 1;
  for (int i = 1; i <= n2; i++) xh = xh * (double)(n1 + 1 - i) / (double)i;
  return (int)(xh + 0.5);
}

/* AUX_FUNX.f90:42-53 (1-based site and orbital, 1-based result) */
static int pack_indices(const or_model *m, int siteI, int orbI) {
  if (orbI == 1) return siteI;
  return siteI + (orbI - 1) * m->Nsites[orbI - 2];
}

/* AUX_FUNX.f90:132-142: ivec(l+1) = bit l of i */
static inline void bdecomp(int i, int Ntot, int *ivec) {
  for (int l = 0; l < Ntot; l++) ivec[l + 1] = (i >> l) & 1; /* ivec[1..Ntot] */
}

/* AUX_FUNX.f90:76-88.  Returns 0, or -1 where the reference stops. */
int or_c(int pos, int in, int *out, double *fsgn) {
  if (!((in >> (pos - 1)) & 1)) return -1; /* "C error: C_i|...0_i...>" */
  *fsgn = 1.0;
  for (int l = 1; l <= pos - 1; l++)
    if ((in >> (l - 1)) & 1) *fsgn = -*fsgn;
  *out = in & ~(1 << (pos - 1));
  return 0;
}

/* AUX_FUNX.f90:90-102 */
int or_cdg(int pos, int in, int *out, double *fsgn) {
  if ((in >> (pos - 1)) & 1) return -1; /* "C^+ error: C^+_i|...1_i...>" */
  *fsgn = 1.0;
  for (int l = 1; l <= pos - 1; l++)
    if ((in >> (l - 1)) & 1) *fsgn = -*fsgn;
  *out = in | (1 << (pos - 1));
  return 0;
}

/* AUX_FUNX.f90:163-180: recursion on array sections a(:mid-1), a(mid+1:) */
int or_binary_search(const int *a, int n, int value) {
  if (n == 0) return 0;
  int mid = n / 2 + 1; /* 1-based */
  if (a[mid - 1] > value) return or_binary_search(a, mid - 1, value);
  if (a[mid - 1] < value) {
    int r = or_binary_search(a + mid, n - mid, value);
    if (r != 0) r = mid + r;
    return r;
  }
  return mid;
}

/* SECTOR.f90:94-109 (one spin species): ascending strings with popcnt == N */
int or_build_map(int Ns, int N, int *map) {
  int dim = 0;
  for (int s = 0; s <= (1 << Ns) - 1; s++) {
    if (__builtin_popcount((unsigned)s) != N) continue;
    map[dim++] = s;
  }
  return dim;
}

/* ---------------------------------------------------------- sector state */
typedef struct {
  int DimUp, DimDw, Dim;
  int *mup, *mdw; /* H(1)%map, H(2)%map (0-based C arrays) */
} or_sector;

static int sector_init(or_sector *S, int Ns, int Nup, int Ndw) {
  long dimup = or_binomial(Ns, Nup), dimdw = or_binomial(Ns, Ndw);
  if (dimup <= 0 || dimdw <= 0) return -2;
  if (dimup * dimdw > 2147483647L) return -3; /* default integer overflow */
  S->DimUp = (int)dimup;
  S->DimDw = (int)dimdw;
  S->Dim = (int)(dimup * dimdw);
  S->mup = (int *)malloc(sizeof(int) * dimup);
  S->mdw = (int *)malloc(sizeof(int) * dimdw);
  if (or_build_map(Ns, Nup, S->mup) != S->DimUp) return -4;
  if (or_build_map(Ns, Ndw, S->mdw) != S->DimDw) return -4;
  return 0;
}
static void sector_free(or_sector *S) {
  free(S->mup);
  free(S->mdw);
}

static int model_init(or_model *m, int Ns, int Nspin, int Norb, const int *Nsites,
                      const double *Hre, const double *Him, const double *Uloc,
                      double Ust, double Jh, double Jx, double Jp, double xmu,
                      int hfmode, int ed_filling) {
  if (Norb < 1 || Norb > OR_MAXORB || Ns < 1 || Ns > OR_MAXNS) return -1;
  if (Nspin != 1 && Nspin != 2) return -1;
  int s = 0;
  m->Ns = Ns;
  m->Nspin = Nspin;
  m->Norb = Norb;
  for (int k = 0; k < Norb; k++) {
    m->Nsites[k] = Nsites[k];
    m->Uloc[k] = Uloc[k];
    s += Nsites[k];
  }
  if (s != Ns) return -1; /* Ns = eNs = sum(Nsites), SETUP.f90:40-42 */
  m->Hre = Hre;
  m->Him = Him;
  m->Ust = Ust;
  m->Jh = Jh;
  m->Jx = Jx;
  m->Jp = Jp;
  m->xmu = xmu;
  m->hfmode = hfmode;
  m->ed_filling = ed_filling;
  return 0;
}

/* ----------------------------------------------------- the four blocks */
/* The per-element diagonal value htmp of HxV_local.f90:1-92 for basis
 * element i (1-based). */
static double local_htmp(const or_model *m, const or_sector *S, int i, const double *Hdiag) {
  int Ns = m->Ns, Nspin = m->Nspin, Norb = m->Norb;
  int Nup[OR_MAXNS + 1], Ndw[OR_MAXNS + 1];
  int iup = i % S->DimUp;
  if (iup == 0) iup = S->DimUp;
  int idw = (i - 1) / S->DimUp + 1;
  int mup = S->mup[iup - 1], mdw = S->mdw[idw - 1];
  bdecomp(mup, Ns, Nup);
  bdecomp(mdw, Ns, Ndw);
  double htmp = 0.0;
  for (int io = 1; io <= Ns; io++)
    htmp = htmp + Hdiag[0 * Ns + io - 1] * Nup[io] + Hdiag[(Nspin - 1) * Ns + io - 1] * Ndw[io];
  int anyU = 0;
  for (int k = 0; k < Norb; k++)
    if (m->Uloc[k] != 0.0) anyU = 1;
  if (anyU)
    for (int iorb = 1; iorb <= Norb; iorb++)
      for (int isite = 1; isite <= m->Nsites[iorb - 1]; isite++) {
        int io = pack_indices(m, isite, iorb);
        htmp = htmp + m->Uloc[iorb - 1] * Nup[io] * Ndw[io];
      }
  for (int iorb = 1; iorb <= Norb; iorb++)
    for (int jorb = iorb + 1; jorb <= Norb; jorb++)
      for (int isite = 1; isite <= m->Nsites[iorb - 1]; isite++)
        for (int jsite = 1; jsite <= m->Nsites[jorb - 1]; jsite++) {
          if (isite != jsite) continue;
          int io = pack_indices(m, isite, iorb), jo = pack_indices(m, isite, jorb);
          htmp = htmp + m->Ust * (Nup[io] * Ndw[jo] + Nup[jo] * Ndw[io]);
        }
  for (int iorb = 1; iorb <= Norb; iorb++)
    for (int jorb = iorb + 1; jorb <= Norb; jorb++)
      for (int isite = 1; isite <= m->Nsites[iorb - 1]; isite++)
        for (int jsite = 1; jsite <= m->Nsites[jorb - 1]; jsite++) {
          if (isite != jsite) continue;
          int io = pack_indices(m, isite, iorb), jo = pack_indices(m, isite, jorb);
          htmp = htmp + (m->Ust - m->Jh) * (Nup[io] * Nup[jo] + Ndw[io] * Ndw[jo]);
        }
  if (m->ed_filling == 0) {
    for (int io = 1; io <= Ns; io++) htmp = htmp - m->xmu * (Nup[io] + Ndw[io]);
    if (m->hfmode) {
      if (anyU)
        for (int iorb = 1; iorb <= Norb; iorb++)
          for (int isite = 1; isite <= m->Nsites[iorb - 1]; isite++) {
            int io = pack_indices(m, isite, iorb);
            htmp = htmp - 0.5 * m->Uloc[iorb - 1] * (Nup[io] + Ndw[io]);
          }
      if (Norb > 1)
        for (int iorb = 1; iorb <= Norb; iorb++)
          for (int jorb = iorb + 1; jorb <= Norb; jorb++)
            for (int isite = 1; isite <= m->Nsites[iorb - 1]; isite++)
              for (int jsite = 1; jsite <= m->Nsites[jorb - 1]; jsite++) {
                if (isite != jsite) continue;
                int io = pack_indices(m, isite, iorb), jo = pack_indices(m, isite, jorb);
                htmp = htmp - 0.5 * m->Ust * (Nup[io] + Ndw[io] + Nup[jo] + Ndw[jo]);
                htmp = htmp - 0.5 * (m->Ust - m->Jh) * (Nup[io] + Ndw[io] + Nup[jo] + Ndw[jo]);
              }
    }
  }
  return htmp;
}

/* A sink receives each scatter contribution Hv(i) += h * vin(j). */
typedef struct {
  int mode; /* 0: accumulate into hv; 1: record triplets; 2: rows filter */
  const zc *vin;
  zc *hv;
  /* triplets */
  long ntrip, captrip;
  int32_t *ti, *tj;
  double *tre, *tim;
  /* row filter: accumulate only contributions landing on row `target` */
  int target;
  zc rowacc;
} or_sink;

static inline void sink_add(or_sink *k, int i, int j, zc h) {
  if (k->mode == 0) {
    k->hv[i - 1] = k->hv[i - 1] + h * k->vin[j - 1];
  } else if (k->mode == 1) {
    long t = k->ntrip++;
    if (t < k->captrip) {
      k->ti[t] = i - 1;
      k->tj[t] = j - 1;
      k->tre[t] = creal(h);
      k->tim[t] = cimag(h);
    }
  } else {
    if (i == k->target) k->rowacc += h * k->vin[j - 1];
  }
}

static int Jhflag_of(const or_model *m) { return m->Norb > 1 && (m->Jx != 0.0 || m->Jp != 0.0); }

/* HxV_local.f90:95 for element i */
static void block_local_elem(const or_model *m, const or_sector *S, const double *Hdiag, int i, or_sink *k) {
  double htmp = local_htmp(m, S, i, Hdiag);
  sink_add(k, i, i, htmp);
}

/* HxV_non_local.f90:1-88 for source element j */
static int block_nonlocal_elem(const or_model *m, const or_sector *S, int j, or_sink *k) {
  int Ns = m->Ns;
  int nup[OR_MAXNS + 1], ndw[OR_MAXNS + 1];
  int jup = j % S->DimUp;
  if (jup == 0) jup = S->DimUp;
  int jdw = (j - 1) / S->DimUp + 1;
  int mup = S->mup[jup - 1], mdw = S->mdw[jdw - 1];
  bdecomp(mup, Ns, nup);
  bdecomp(mdw, Ns, ndw);
  int k1, k2, k3, k4;
  double sg1, sg2, sg3, sg4;
  if (m->Jx != 0.0) {
    for (int iorb = 1; iorb <= m->Norb; iorb++)
      for (int jorb = 1; jorb <= m->Norb; jorb++)
        for (int isite = 1; isite <= m->Nsites[iorb - 1]; isite++)
          for (int jsite = 1; jsite <= m->Nsites[jorb - 1]; jsite++) {
            if (isite != jsite) continue;
            int io = pack_indices(m, isite, iorb), jo = pack_indices(m, isite, jorb);
            int Jc = (io != jo) && (nup[jo] == 1) && (ndw[io] == 1) && (ndw[jo] == 0) && (nup[io] == 0);
            if (Jc) {
              if (or_c(io, mdw, &k1, &sg1) || or_cdg(jo, k1, &k2, &sg2)) return -1;
              int idw = or_binary_search(S->mdw, S->DimDw, k2);
              if (or_c(jo, mup, &k3, &sg3) || or_cdg(io, k3, &k4, &sg4)) return -1;
              int iup = or_binary_search(S->mup, S->DimUp, k4);
              zc htmp = m->Jx * sg1 * sg2 * sg3 * sg4;
              int i = iup + (idw - 1) * S->DimUp;
              sink_add(k, i, j, htmp);
            }
          }
  }
  if (m->Jp != 0.0) {
    for (int iorb = 1; iorb <= m->Norb; iorb++)
      for (int jorb = 1; jorb <= m->Norb; jorb++)
        for (int isite = 1; isite <= m->Nsites[iorb - 1]; isite++)
          for (int jsite = 1; jsite <= m->Nsites[jorb - 1]; jsite++) {
            if (isite != jsite) continue;
            int io = pack_indices(m, isite, iorb), jo = pack_indices(m, isite, jorb);
            int Jc = (nup[jo] == 1) && (ndw[jo] == 1) && (ndw[io] == 0) && (nup[io] == 0);
            if (Jc) {
              if (or_c(jo, mdw, &k1, &sg1) || or_cdg(io, k1, &k2, &sg2)) return -1;
              int idw = or_binary_search(S->mdw, S->DimDw, k2);
              if (or_c(jo, mup, &k3, &sg3) || or_cdg(io, k3, &k4, &sg4)) return -1;
              int iup = or_binary_search(S->mup, S->DimUp, k4);
              zc htmp = m->Jp * sg1 * sg2 * sg3 * sg4;
              int i = iup + (idw - 1) * S->DimUp;
              sink_add(k, i, j, htmp);
            }
          }
  }
  return 0;
}

/* HxV_up.f90:3-29 for source element (jup, jdw) */
static int block_up_elem(const or_model *m, const or_sector *S, int jup, int jdw, or_sink *k) {
  int Ns = m->Ns;
  int nup[OR_MAXNS + 1];
  int mup = S->mup[jup - 1];
  bdecomp(mup, Ns, nup);
  int j = jup + (jdw - 1) * S->DimUp;
  int k1, k2;
  double sg1, sg2;
  for (int io = 1; io <= Ns; io++)
    for (int jo = 1; jo <= Ns; jo++) {
      zc h = Hij_at(m, 1, io, jo);
      int Jc = (h != 0.0) && (nup[jo] == 1) && (nup[io] == 0);
      if (Jc) {
        if (or_c(jo, mup, &k1, &sg1) || or_cdg(io, k1, &k2, &sg2)) return -1;
        int iup = or_binary_search(S->mup, S->DimUp, k2);
        int idw = jdw;
        int i = iup + (idw - 1) * S->DimUp;
        zc htmp = h * sg1 * sg2;
        sink_add(k, i, j, htmp);
      }
    }
  return 0;
}

/* HxV_dw.f90:3-28 for source element (jup, jdw) */
static int block_dw_elem(const or_model *m, const or_sector *S, int jup, int jdw, or_sink *k) {
  int Ns = m->Ns, Nspin = m->Nspin;
  int ndw[OR_MAXNS + 1];
  int mdw = S->mdw[jdw - 1];
  bdecomp(mdw, Ns, ndw);
  int j = jup + (jdw - 1) * S->DimUp;
  int k1, k2;
  double sg1, sg2;
  for (int io = 1; io <= Ns; io++)
    for (int jo = 1; jo <= Ns; jo++) {
      zc h = Hij_at(m, Nspin, io, jo);
      int Jc = (h != 0.0) && (ndw[jo] == 1) && (ndw[io] == 0);
      if (Jc) {
        if (or_c(jo, mdw, &k1, &sg1) || or_cdg(io, k1, &k2, &sg2)) return -1;
        int idw = or_binary_search(S->mdw, S->DimDw, k2);
        int iup = jup;
        int i = iup + (idw - 1) * S->DimUp;
        zc htmp = h * sg1 * sg2;
        sink_add(k, i, j, htmp);
      }
    }
  return 0;
}

static void make_Hdiag(const or_model *m, double *Hdiag) {
  /* HAMILTONIAN_DIRECT_HxV.f90:35-39: Hdiag(ispin,:) = dreal(diagonal(Hloc)) with
   * Hloc filled by Hij_get, i.e. the whole Hij */
  for (int s = 1; s <= m->Nspin; s++)
    for (int io = 1; io <= m->Ns; io++) Hdiag[(s - 1) * m->Ns + io - 1] = creal(Hij_at(m, s, io, io));
}

/* ------------------------------------------------------------ public API */
#define OR_MODEL_ARGS                                                                   \
  int Ns, int Nspin, int Norb, const int *Nsites, const double *Hre, const double *Him, \
      const double *Uloc, double Ust, double Jh, double Jx, double Jp, double xmu, int hfmode, int ed_filling
#define OR_MODEL_PASS Ns, Nspin, Norb, Nsites, Hre, Him, Uloc, Ust, Jh, Jx, Jp, xmu, hfmode, ed_filling

/* Sector maps (SECTOR.f90:69-115); returns 0 or a negative error. */
int or_sector_maps(int Ns, int Nup, int Ndw, int *map_up, int *map_dw, int *dimup, int *dimdw) {
  *dimup = or_build_map(Ns, Nup, map_up);
  *dimdw = or_build_map(Ns, Ndw, map_dw);
  return 0;
}

/* directMatVec (HAMILTONIAN_DIRECT_HxV.f90:19-59).  Serial, scatter form.
 * frac in (0,1] restricts each block's outer loop to its first frac part
 * (used only to time a bounded sample); frac = 1 is the full product. */
int or_direct_matvec(OR_MODEL_ARGS, int Nup, int Ndw, const double *vin, double *hv, double frac) {
  or_model m;
  or_sector S;
  int rc = model_init(&m, OR_MODEL_PASS);
  if (rc) return rc;
  rc = sector_init(&S, Ns, Nup, Ndw);
  if (rc) return rc;
  double Hdiag[2 * OR_MAXNS];
  make_Hdiag(&m, Hdiag);
  or_sink k = {0};
  k.mode = 0;
  k.vin = (const zc *)vin;
  k.hv = (zc *)hv;
  for (int i = 0; i < S.Dim; i++) k.hv[i] = 0.0; /* Hv=zero */
  int nloc = (int)(frac * S.Dim), ndw_o = (int)(frac * S.DimDw), nup_o = (int)(frac * S.DimUp);
  if (frac >= 1.0) {
    nloc = S.Dim;
    ndw_o = S.DimDw;
    nup_o = S.DimUp;
  }
  for (int i = 1; i <= nloc; i++) block_local_elem(&m, &S, Hdiag, i, &k);
  if (Jhflag_of(&m))
    for (int j = 1; j <= nloc; j++)
      if (block_nonlocal_elem(&m, &S, j, &k)) rc = -5;
  for (int jdw = 1; jdw <= ndw_o; jdw++)
    for (int jup = 1; jup <= S.DimUp; jup++)
      if (block_up_elem(&m, &S, jup, jdw, &k)) rc = -5;
  for (int jup = 1; jup <= nup_o; jup++)
    for (int jdw = 1; jdw <= S.DimDw; jdw++)
      if (block_dw_elem(&m, &S, jup, jdw, &k)) rc = -5;
  sector_free(&S);
  return rc;
}

/* Same product, the outer loop of each block split over OpenMP threads.
 * Up hops of source column jdw only land in column jdw and down hops of
 * source row jup only in row jup, so each split is free of write races. */
int or_direct_matvec_omp(OR_MODEL_ARGS, int Nup, int Ndw, const double *vin, double *hv, double frac,
                         int nthreads) {
  or_model m;
  or_sector S;
  int rc = model_init(&m, OR_MODEL_PASS);
  if (rc) return rc;
  rc = sector_init(&S, Ns, Nup, Ndw);
  if (rc) return rc;
  if (Jhflag_of(&m)) {
    sector_free(&S);
    return -6;
  }
  double Hdiag[2 * OR_MAXNS];
  make_Hdiag(&m, Hdiag);
  zc *H = (zc *)hv;
  const zc *V = (const zc *)vin;
  int nloc = (int)(frac * S.Dim), ndw_o = (int)(frac * S.DimDw), nup_o = (int)(frac * S.DimUp);
  if (frac >= 1.0) {
    nloc = S.Dim;
    ndw_o = S.DimDw;
    nup_o = S.DimUp;
  }
  int err = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    or_sink k = {0};
    k.mode = 0;
    k.vin = V;
    k.hv = H;
#pragma omp for schedule(static)
    for (int i = 0; i < S.Dim; i++) H[i] = 0.0;
#pragma omp for schedule(static)
    for (int i = 1; i <= nloc; i++) block_local_elem(&m, &S, Hdiag, i, &k);
#pragma omp for schedule(dynamic, 4)
    for (int jdw = 1; jdw <= ndw_o; jdw++)
      for (int jup = 1; jup <= S.DimUp; jup++)
        if (block_up_elem(&m, &S, jup, jdw, &k)) err = 1;
#pragma omp for schedule(dynamic, 4)
    for (int jup = 1; jup <= nup_o; jup++)
      for (int jdw = 1; jdw <= S.DimDw; jdw++)
        if (block_dw_elem(&m, &S, jup, jdw, &k)) err = 1;
  }
  sector_free(&S);
  return err ? -5 : 0;
}

/* The matrix of directMatVec as (i, j, H_ij) triplets, 0-based, in the
 * order the scatter visits them (duplicates are summed by the caller).
 * Returns the number of triplets, or a negative error; if cap is too
 * small nothing beyond cap is stored and the full count is returned. */
long or_matvec_triplets(OR_MODEL_ARGS, int Nup, int Ndw, long cap, int32_t *ti, int32_t *tj, double *tre,
                        double *tim) {
  or_model m;
  or_sector S;
  int rc = model_init(&m, OR_MODEL_PASS);
  if (rc) return rc;
  rc = sector_init(&S, Ns, Nup, Ndw);
  if (rc) return rc;
  double Hdiag[2 * OR_MAXNS];
  make_Hdiag(&m, Hdiag);
  or_sink k = {0};
  k.mode = 1;
  k.captrip = cap;
  k.ti = ti;
  k.tj = tj;
  k.tre = tre;
  k.tim = tim;
  long err = 0;
  for (int i = 1; i <= S.Dim; i++) block_local_elem(&m, &S, Hdiag, i, &k);
  if (Jhflag_of(&m))
    for (int j = 1; j <= S.Dim; j++)
      if (block_nonlocal_elem(&m, &S, j, &k)) err = -5;
  for (int jdw = 1; jdw <= S.DimDw; jdw++)
    for (int jup = 1; jup <= S.DimUp; jup++)
      if (block_up_elem(&m, &S, jup, jdw, &k)) err = -5;
  for (int jup = 1; jup <= S.DimUp; jup++)
    for (int jdw = 1; jdw <= S.DimDw; jdw++)
      if (block_dw_elem(&m, &S, jup, jdw, &k)) err = -5;
  sector_free(&S);
  return err ? err : k.ntrip;
}

/* (H v)_i for a list of rows i (0-based), for sectors too large for a whole
 * product.  The sources j that can reach row i are i itself and its hop
 * partners, found with the same c/cdg/binary_search as the reference; the
 * scatter bodies are then run for those sources and only contributions
 * landing on the listed rows are kept.  Non-local (Jhflag) terms refused. */
int or_hxv_rows(OR_MODEL_ARGS, int Nup, int Ndw, int nrows, const int64_t *rows, const double *vin,
                double *out) {
  or_model m;
  or_sector S;
  int rc = model_init(&m, OR_MODEL_PASS);
  if (rc) return rc;
  if (Jhflag_of(&m)) return -6;
  rc = sector_init(&S, Ns, Nup, Ndw);
  if (rc) return rc;
  double Hdiag[2 * OR_MAXNS];
  make_Hdiag(&m, Hdiag);
  or_sink k = {0};
  k.mode = 2;
  k.vin = (const zc *)vin;
  for (int r = 0; r < nrows; r++) {
    int i = (int)rows[r] + 1;
    k.target = i;
    k.rowacc = 0.0;
    int iup = i % S.DimUp;
    if (iup == 0) iup = S.DimUp;
    int idw = (i - 1) / S.DimUp + 1;
    block_local_elem(&m, &S, Hdiag, i, &k);
    /* candidate sources: up partners jup of iup in column idw */
    int bits[OR_MAXNS + 1];
    int s_up = S.mup[iup - 1];
    bdecomp(s_up, Ns, bits);
    for (int a = 1; a <= Ns; a++)
      for (int b = 1; b <= Ns; b++) {
        if (!(bits[a] == 1 && bits[b] == 0)) continue;
        int k1, k2;
        double g1, g2;
        or_c(a, s_up, &k1, &g1);
        or_cdg(b, k1, &k2, &g2);
        int jup = or_binary_search(S.mup, S.DimUp, k2);
        if (jup > 0) block_up_elem(&m, &S, jup, idw, &k);
      }
    int s_dw = S.mdw[idw - 1];
    bdecomp(s_dw, Ns, bits);
    for (int a = 1; a <= Ns; a++)
      for (int b = 1; b <= Ns; b++) {
        if (!(bits[a] == 1 && bits[b] == 0)) continue;
        int k1, k2;
        double g1, g2;
        or_c(a, s_dw, &k1, &g1);
        or_cdg(b, k1, &k2, &g2);
        int jdw = or_binary_search(S.mdw, S.DimDw, k2);
        if (jdw > 0) block_dw_elem(&m, &S, iup, jdw, &k);
      }
    out[2 * r] = creal(k.rowacc);
    out[2 * r + 1] = cimag(k.rowacc);
  }
  sector_free(&S);
  return 0;
}

/* apply_op_C / apply_op_CDG (SECTOR.f90:217-268) on every element of sector I,
 * writing the seed vvinit(j) = sgn*state(i) into sector J as
 * GF_ELECTRONS.f90:241-245 does (assignment, not accumulation).
 * op: 1 = C, 2 = CDG.  ipos, ispin 1-based.  Complex interleaved vectors. */
int or_apply_op_c_cdg(int Ns, int op, int ipos, int ispin, int NupI, int NdwI, int NupJ, int NdwJ,
                      const double *state, double *vvinit) {
  or_sector SI, SJ;
  int rc = sector_init(&SI, Ns, NupI, NdwI);
  if (rc) return rc;
  rc = sector_init(&SJ, Ns, NupJ, NdwJ);
  if (rc) return rc;
  const zc *st = (const zc *)state;
  zc *vv = (zc *)vvinit;
  for (int j = 0; j < SJ.Dim; j++) vv[j] = 0.0;
  int nud[OR_MAXNS + 1];
  for (int i = 1; i <= SI.Dim; i++) {
    int Ind[2], Jnd[2];
    /* state2indices(i,[DimUps,DimDws],Indices), SECTOR.f90:453-464 */
    int count = i - 1;
    Ind[0] = count % SI.DimUp + 1;
    count /= SI.DimUp;
    Ind[1] = count % SI.DimDw + 1;
    int iud = (ispin == 1) ? SI.mup[Ind[0] - 1] : SI.mdw[Ind[1] - 1];
    bdecomp(iud, Ns, nud);
    int r;
    double sgn;
    if (op == 1) {
      if (nud[ipos] != 1) continue;
      or_c(ipos, iud, &r, &sgn);
    } else {
      if (nud[ipos] != 0) continue;
      or_cdg(ipos, iud, &r, &sgn);
    }
    Jnd[0] = Ind[0];
    Jnd[1] = Ind[1];
    Jnd[ispin - 1] = (ispin == 1) ? or_binary_search(SJ.mup, SJ.DimUp, r) : or_binary_search(SJ.mdw, SJ.DimDw, r);
    int j = Jnd[0] + (Jnd[1] - 1) * SJ.DimUp; /* indices2state */
    if (sgn == 0.0 || j == 0) continue;
    vv[j - 1] = sgn * st[i - 1];
  }
  sector_free(&SI);
  sector_free(&SJ);
  return 0;
}

/* apply_op_Sz (SECTOR.f90:271-291): vvinit(i) = sgn*state(i), same sector,
 * with sgn = (Nup(ipos)-Ndw(ipos))/2 (CHI_SPIN_ELECTRONS.f90:216-220). */
int or_apply_op_sz(int Ns, int ipos, int Nup, int Ndw, const double *state, double *vvinit) {
  or_sector S;
  int rc = sector_init(&S, Ns, Nup, Ndw);
  if (rc) return rc;
  const zc *st = (const zc *)state;
  zc *vv = (zc *)vvinit;
  int nu[OR_MAXNS + 1], nd[OR_MAXNS + 1];
  for (int i = 1; i <= S.Dim; i++) {
    int count = i - 1;
    int a = count % S.DimUp + 1;
    count /= S.DimUp;
    int b = count % S.DimDw + 1;
    bdecomp(S.mup[a - 1], Ns, nu);
    bdecomp(S.mdw[b - 1], Ns, nd);
    double sgn = nu[ipos] - nd[ipos];
    sgn = sgn / 2.0;
    vv[i - 1] = sgn * st[i - 1];
  }
  sector_free(&S);
  return 0;
}

/* lanc_observables (OBSERVABLES.f90:129-152), one state of weight 1:
 * dens_up, dens_dw, docc per site. */
int or_observables(int Ns, int Nup, int Ndw, const double *state, double *dens_up, double *dens_dw,
                   double *docc) {
  or_sector S;
  int rc = sector_init(&S, Ns, Nup, Ndw);
  if (rc) return rc;
  const zc *st = (const zc *)state;
  for (int is = 0; is < Ns; is++) dens_up[is] = dens_dw[is] = docc[is] = 0.0;
  int nu[OR_MAXNS + 1], nd[OR_MAXNS + 1];
  for (int i = 1; i <= S.Dim; i++) {
    double w = cabs(st[i - 1]);
    w = w * w;
    int count = i - 1;
    int a = count % S.DimUp + 1;
    count /= S.DimUp;
    int b = count % S.DimDw + 1;
    bdecomp(S.mup[a - 1], Ns, nu);
    bdecomp(S.mdw[b - 1], Ns, nd);
    for (int is = 1; is <= Ns; is++) {
      dens_up[is - 1] += nu[is] * w;
      dens_dw[is - 1] += nd[is] * w;
      docc[is - 1] += (double)(nu[is] * nd[is]) * w;
    }
  }
  sector_free(&S);
  return 0;
}

/* (H v)_i for listed rows of a sector larger than the reference's default
 * integer can index (Dim > 2^31-1): the same restated per-element formulas
 * (HxV_local.f90:1-92, HxV_up.f90:13-27, HxV_dw.f90:12-28), with 64-bit
 * element indices, a real fp32 input vector and fp64 accumulation.  For
 * each row the candidate sources are its hop partners; a source's scatter
 * contributions are kept when their target (found by binary_search as in the
 * reference) is the row.  Non-local (Jhflag) terms refused. */
/* Element k of the input vector: fp32 (vf) or fp64 (vd) storage. */
static inline double vec_at(const float *vf, const double *vd, int64_t k) { return vf ? (double)vf[k] : vd[k]; }

static int hxv_rows_real64_impl(OR_MODEL_ARGS, int Nup, int Ndw, int nrows, const int64_t *rows, const float *vf,
                                const double *vd, double *out) {
  or_model m;
  int rc = model_init(&m, OR_MODEL_PASS);
  if (rc) return rc;
  if (Jhflag_of(&m)) return -6;
  long dimup = or_binomial(Ns, Nup), dimdw = or_binomial(Ns, Ndw);
  int *mup = (int *)malloc(sizeof(int) * dimup), *mdw = (int *)malloc(sizeof(int) * dimdw);
  or_build_map(Ns, Nup, mup);
  or_build_map(Ns, Ndw, mdw);
  double Hdiag[2 * OR_MAXNS];
  make_Hdiag(&m, Hdiag);
  /* a one-column sector view reused for the diagonal: local_htmp indexes
   * element 1 = (iup, idw) of a 1x1 sector built on the fly */
  for (int r = 0; r < nrows; r++) {
    const int64_t i = rows[r];
    const int iup = (int)(i % dimup) + 1, idw = (int)(i / dimup) + 1; /* 1-based */
    double acc = 0.0;
    {
      or_sector S1;
      int su = mup[iup - 1], sd = mdw[idw - 1];
      S1.DimUp = 1;
      S1.DimDw = 1;
      S1.Dim = 1;
      S1.mup = &su;
      S1.mdw = &sd;
      acc += local_htmp(&m, &S1, 1, Hdiag) * vec_at(vf, vd, i);
    }
    int bits[OR_MAXNS + 1];
    /* up hops: sources jup in column idw */
    bdecomp(mup[iup - 1], Ns, bits);
    for (int a = 1; a <= Ns; a++)
      for (int b = 1; b <= Ns; b++) {
        if (!(bits[a] == 1 && bits[b] == 0)) continue;
        int k1, k2;
        double g1, g2;
        or_c(a, mup[iup - 1], &k1, &g1);
        or_cdg(b, k1, &k2, &g2);
        int jup = or_binary_search(mup, (int)dimup, k2);
        if (jup <= 0) continue;
        int nj[OR_MAXNS + 1];
        bdecomp(mup[jup - 1], Ns, nj);
        for (int io = 1; io <= Ns; io++)
          for (int jo = 1; jo <= Ns; jo++) {
            double h = creal(Hij_at(&m, 1, io, jo));
            if (!(h != 0.0 && nj[jo] == 1 && nj[io] == 0)) continue;
            int t1, t2;
            double s1, s2;
            or_c(jo, mup[jup - 1], &t1, &s1);
            or_cdg(io, t1, &t2, &s2);
            if (or_binary_search(mup, (int)dimup, t2) == iup)
              acc += h * s1 * s2 * vec_at(vf, vd, (int64_t)(jup - 1) + (int64_t)(idw - 1) * dimup);
          }
      }
    /* down hops: sources jdw in row iup */
    bdecomp(mdw[idw - 1], Ns, bits);
    for (int a = 1; a <= Ns; a++)
      for (int b = 1; b <= Ns; b++) {
        if (!(bits[a] == 1 && bits[b] == 0)) continue;
        int k1, k2;
        double g1, g2;
        or_c(a, mdw[idw - 1], &k1, &g1);
        or_cdg(b, k1, &k2, &g2);
        int jdw = or_binary_search(mdw, (int)dimdw, k2);
        if (jdw <= 0) continue;
        int nj[OR_MAXNS + 1];
        bdecomp(mdw[jdw - 1], Ns, nj);
        for (int io = 1; io <= Ns; io++)
          for (int jo = 1; jo <= Ns; jo++) {
            double h = creal(Hij_at(&m, m.Nspin, io, jo));
            if (!(h != 0.0 && nj[jo] == 1 && nj[io] == 0)) continue;
            int t1, t2;
            double s1, s2;
            or_c(jo, mdw[jdw - 1], &t1, &s1);
            or_cdg(io, t1, &t2, &s2);
            if (or_binary_search(mdw, (int)dimdw, t2) == idw)
              acc += h * s1 * s2 * vec_at(vf, vd, (int64_t)(iup - 1) + (int64_t)(jdw - 1) * dimup);
          }
      }
    out[r] = acc;
  }
  free(mup);
  free(mdw);
  return 0;
}

int or_hxv_rows_real64(OR_MODEL_ARGS, int Nup, int Ndw, int nrows, const int64_t *rows, const float *v,
                       double *out) {
  return hxv_rows_real64_impl(OR_MODEL_PASS, Nup, Ndw, nrows, rows, v, NULL, out);
}

/* The same rows for a real fp64 input vector (checks of full-size fp64
 * eigenvectors). */
int or_hxv_rows_real64d(OR_MODEL_ARGS, int Nup, int Ndw, int nrows, const int64_t *rows, const double *v,
                        double *out) {
  return hxv_rows_real64_impl(OR_MODEL_PASS, Nup, Ndw, nrows, rows, NULL, v, out);
}
