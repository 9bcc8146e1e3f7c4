/*
 * ORACLE - test infrastructure only.  Nothing in the product library links,
 * loads or calls this file; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, as the checker.
 *
 * CPU restatements of the reference's nonbonded and PME real-space algorithms:
 *
 *  oracle_nb_fp64          FP64 analytic pair forces, energies and virial for
 *                          every non-excluded minimum-image pair with r < rc:
 *                          energy-switched LJ (ComputeNonbondedUtil.C:960-971,
 *                          CudaComputeNonbondedInteractions.h:37-61) and the
 *                          C1-split Ewald real space (CudaComputeNonbondedInteractions.h:107-159;
 *                          fast + slow = erfc(beta r)/r).
 *  oracle_nb_fp64_alch     the same with alchemical partitions (FEP, and the
 *                          TI lambda derivatives):
 *                          soft-core LJ of fep_vdw_forceandenergies
 *                          (ComputeNonbondedFEP.C:156-165, potential switching,
 *                          vdw_switch / vdw_forceandenergy of
 *                          ComputeNonbondedAlch.h:16-48) and lambda-scaled
 *                          electrostatics (ComputeNonbondedBase2.h:367-372,
 *                          :497, :645-686), with the pair classes of
 *                          calcForceEnergyFEP (CudaComputeNonbondedKernel.cu:203-213).
 *  oracle_nb_fp64_vdw_force_switch, oracle_nb_fp64_alch_vdw_force_switch
 *                          the same two with vdW force switching
 *                          (vdwForceSwitching: ComputeNonbondedUtil.C:523-533,
 *                          :900-919; alchemical: ComputeNonbondedFEP.C:136-159,
 *                          ComputeNonbondedTI.C:113-128).
 *  oracle_nb_fp64_cell     oracle_nb_fp64 in a general (triclinic) cell, by
 *                          brute force over pairs and images (energy- or
 *                          force-switched LJ).
 *  oracle_mod_excl_fp64    the modified (scaled 1-4) exclusion correction
 *                          of modifiedExclusionForce
 *                          (ComputeBondedCUDAKernel.cu:316-632) in FP64.
 *  oracle_nb_table         table-mode pair evaluation: the pair loop of
 *                          oracle_nb_fp64 with forces and energies from given
 *                          1/r float4 tables, linearly interpolated in FP64
 *                          as calcForceEnergy does (CudaComputeNonbondedKernel.cu:99-161,
 *                          CudaComputeNonbondedKernelHip.cu:170-220).
 *  oracle_namd_gpu_tables  NAMD's 1/r-indexed float4 force/energy tables
 *                          (CudaNonbondedTables.C:158-203, :263-300) sampled
 *                          from the spline tables above at r = N/(i + 0.5).
 *  oracle_nb_adjoint_fp64  FP64 gradient of a force-matching loss
 *                          L = sum_i |F_i - T_i|^2 (F = fast + slow force) with
 *                          respect to the LJ table entries A_ab, B_ab (ordered
 *                          type pairs) and the scaled charges, given the
 *                          adjoint vectors g_i = dL/dF_i; the pair terms are
 *                          the derivatives of the oracle_nb_fp64 pair force.
 *  oracle_pairs_fp32       the canonical pair-membership rule the kernels use:
 *                          r2 = fma(dz,dz,fma(dy,dy,dx*dx)) in FP32 on exact
 *                          displacements, r2 < (float)rc^2; counts and
 *                          order-independent checksums of the pair sets, or
 *                          the sets themselves.
 *  oracle_pme_spread_gpu / oracle_pme_gather_gpu
 *                          B-spline spread / gather in FP64 with NAMD's GPU
 *                          grid convention (CudaPmeSolverUtilKernel.cu:331-915).
 *  oracle_pme_scale_coordinates / oracle_pme_spread_cpu / oracle_pme_gather_cpu
 *                          NAMD's CPU convention (PmeBase.inl:17-148,
 *                          PmeRealSpace.C:237-425), order 4 with the explicit
 *                          cubics, other orders with compute_b_spline.
 *  oracle_pairs_exact      NAMD's CPU membership rule (r^2 <= rc^2 in FP64,
 *                          exact on 2^-17 A effective positions).
 *  oracle_namd_tables / oracle_namd_cpu
 *                          NAMD's CPU nonbonded path: FP64 r^2 cubic-spline
 *                          tables of ComputeNonbondedUtil::select and the pair
 *                          loop of ComputeNonbondedBase2.h.
 *
 * Parity pins are listed in the test files that use each function.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TWO_OVER_SQRT_PI 1.1283791670955126

static int is_excluded(const int64_t* ptr, const int32_t* idx, int i, int j) {
  for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k)
    if (idx[k] == j) return 1;
  return 0;
}

static inline double min_image(double d, double L) { return d - L * nearbyint(d / L); }

/* ---- FP64 pair physics --------------------------------------------------- */
typedef struct {
  double rc2, rs2, c1, rc, rcinv, rc3inv, beta;
  /* vdW force switching (vdwForceSwitching, ComputeNonbondedUtil.C:523-533,
   * :900-919; Steinbach & Brooks 1994): 0 = energy switching */
  int forceSwitch;
  double ka, kb, c6inv, c3inv, va, vb;
} PairConst;

static void set_force_switch(PairConst* c, double cutoff, double switchdist) {
  const double s3 = switchdist * switchdist * switchdist, c3 = cutoff * cutoff * cutoff;
  const double s6 = s3 * s3, c6 = c3 * c3;
  c->forceSwitch = 1;
  c->va = -1.0 / (s6 * c6);
  c->vb = -1.0 / (s3 * c3);
  c->ka = c6 / (c6 - s6);
  c->kb = c3 / (c3 - s3);
  c->c6inv = 1.0 / c6;
  c->c3inv = 1.0 / c3;
}

static void pair_fp64(const PairConst* c, double r2, double qq, double A, double B, double* f, double* fs,
                      double* ev, double* ee, double* es) {
  const double r = sqrt(r2);
  const double rinv = 1.0 / r;
  const double rinv2 = rinv * rinv;
  const double rinv6 = rinv2 * rinv2 * rinv2;
  /* LJ: E = A r^-12 - B r^-6, times S(r^2) above the switching distance */
  const double E = (A * rinv6 - B) * rinv6;
  const double dE_r = (-12.0 * A * rinv6 + 6.0 * B) * rinv6 * rinv2; /* (1/r) dE/dr */
  double fv;
  if (c->forceSwitch) {
    /* force switching: below the switching distance the plain terms plus a
     * constant, above it k_a (r^-6 - rc^-6)^2 and k_b (r^-3 - rc^-3)^2 */
    if (r2 > c->rs2) {
      const double rinv3 = rinv * rinv2;
      const double ta = rinv6 - c->c6inv, tb = rinv3 - c->c3inv;
      *ev += A * c->ka * ta * ta - B * c->kb * tb * tb;
      /* (1/r) dE/dr = 2 dE/d(r^2) */
      fv = 2.0 * (-6.0 * A * c->ka * ta * rinv6 * rinv2 + 3.0 * B * c->kb * tb * rinv3 * rinv2);
    } else {
      *ev += E + A * c->va - B * c->vb;
      fv = dE_r;
    }
  } else {
    double S = 1.0, dS_r = 0.0;
    if (r2 > c->rs2) {
      const double d = c->rc2 - r2;
      S = d * d * (c->rc2 + 2.0 * r2 - 3.0 * c->rs2) * c->c1;
      /* dS/dr / r = 2 dS/d(r^2) */
      dS_r = 2.0 * c->c1 * (-2.0 * d * (c->rc2 + 2.0 * r2 - 3.0 * c->rs2) + 2.0 * d * d);
    }
    fv = dE_r * S + E * dS_r;
    *ev += E * S;
  }
  /* fast: E = qq (1/r - 3/(2rc) + r^2/(2 rc^3)) */
  *f = fv + qq * (-rinv * rinv2 + c->rc3inv);
  *ee += qq * (rinv - 1.5 * c->rcinv + 0.5 * r2 * c->rc3inv);
  /* slow: E = qq (erfc(beta r)/r - 1/r + 3/(2rc) - r^2/(2 rc^3)) */
  const double a = c->beta * r;
  const double ec = erfc(a);
  const double g = ec + TWO_OVER_SQRT_PI * a * exp(-a * a); /* -r^2 d/dr(erfc(beta r)/r) */
  *fs = qq * ((1.0 - g) * rinv * rinv2 - c->rc3inv);
  *es += qq * ((ec - 1.0) * rinv + 1.5 * c->rcinv - 0.5 * r2 * c->rc3inv);
}

/* Alchemical factors of one lambda window (AlchData's lambda part). */
typedef struct {
  double elecUp, elecDown, vdwUp, vdwDown;
  double elec2Up, elec2Down, vdw2Up, vdw2Down;
  double shUp, shDown, sh2Up, sh2Down;
  int decouple;
  double coeff;  /* alchVdwShiftCoeff: shift = coeff (1 - vdw lambda) */
} AlchF;

/* vdw_switch (ComputeNonbondedAlch.h:35-48): S and switchmul2 = -2 dS/d(r^2) */
static void vdw_switch64(const PairConst* c, double r2, double* S, double* S2) {
  if (r2 <= c->rs2) {
    *S = 1.0;
    *S2 = 0.0;
  } else {
    const double cd = c->rc2 - r2, sd = r2 - c->rs2;
    *S = c->c1 * cd * cd * (cd + 3.0 * sd);
    *S2 = 12.0 * c->c1 * cd * sd;
  }
}

/* vdw_forceandenergy (ComputeNonbondedAlch.h:16-22) at x = r^2 + shift:
 * U = A/x^6 - B/x^3, F = -2 dU/dx */
static void vdw_fe64(double A, double B, double x, double* U, double* F) {
  const double i2 = 1.0 / x, i6 = i2 * i2 * i2;
  *U = i6 * (A * i6 - B);
  *F = 6.0 * i2 * (2.0 * (*U) + B * i6);
}

/* Force-switched soft-core LJ at x = r^2 + shift with the switching distance
 * moved by the same shift, sd = switchdist^2 + shift, and the cutoff not
 * (vdw_fswitch_forceandenergy / vdw_fswitch_shift, ComputeNonbondedAlch.h:56-104,
 * as fep_vdw_forceandenergies and ti_vdw_force_energy_dUdl call them,
 * ComputeNonbondedFEP.C:136-159, ComputeNonbondedTI.C:113-128).  inner: the
 * pair is within the (unshifted) switching distance.  U is the energy (with
 * the constant shift inside), F = -2 dU/dx. */
static void vdw_fswitch64(const PairConst* c, int inner, double A, double B, double x, double sd, double* U,
                          double* F) {
  const double c6 = c->rc2 * c->rc2 * c->rc2, sd6 = sd * sd * sd;
  if (inner) {
    vdw_fe64(A, B, x, U, F);
    *U += -A / (c6 * sd6) + B / sqrt(c6 * sd6);
  } else {
    const double i2 = 1.0 / x, i6 = i2 * i2 * i2, i3 = sqrt(i6);
    const double ic6 = 1.0 / c6, ic3 = sqrt(ic6);
    const double ka = A / (1.0 - sd6 * ic6), kb = B / (1.0 - sqrt(sd6 * ic6));
    const double ta = i6 - ic6, tb = i3 - ic3;
    *U = ka * ta * ta - kb * tb * tb;
    *F = 6.0 * i2 * (2.0 * ka * ta * i6 - kb * tb * i3);
  }
}

/* e[0..5] = vdw, elec, slow, vdw at lambda2, elec at lambda2, slow at lambda2 */
static void pair_fp64_alch(const PairConst* c, double r2, double qq, double A, double B, int pi, int pj,
                           const AlchF* al, double* f, double* fs, double* e) {
  const int ref = (pi == 0 && pj == 0);
  const int cross = (pi != 0 && pj != 0 && pi != pj);
  const int alch = !ref && !cross;
  const int dec = alch && pi == pj && al->decouple;
  const int up = alch && !dec && (pi == 1 || pj == 1);
  const int down = alch && !dec && (pi == 2 || pj == 2);
  if (cross) {
    *f = *fs = 0.0;
    return;
  }
  const double lv = up ? al->vdwUp : down ? al->vdwDown : 1.0;
  const double le = up ? al->elecUp : down ? al->elecDown : 1.0;
  const double lv2 = up ? al->vdw2Up : down ? al->vdw2Down : 1.0;
  const double le2 = up ? al->elec2Up : down ? al->elec2Down : 1.0;
  const double sh = up ? al->shUp : down ? al->shDown : 0.0;
  const double sh2 = up ? al->sh2Up : down ? al->sh2Down : 0.0;
  double S, S2, U, F, U2, F2;
  if (c->forceSwitch) {
    /* force switching: no multiplicative switch, the switched form itself */
    S = 1.0;
    S2 = 0.0;
    const int inner = r2 <= c->rs2;
    vdw_fswitch64(c, inner, A, B, r2 + sh, c->rs2 + sh, &U, &F);
    vdw_fswitch64(c, inner, A, B, r2 + sh2, c->rs2 + sh2, &U2, &F2);
  } else {
    vdw_switch64(c, r2, &S, &S2);
    vdw_fe64(A, B, r2 + sh, &U, &F);
    vdw_fe64(A, B, r2 + sh2, &U2, &F2);
  }
  /* alch_vdw_force = lv (S F + S2 U) is -2 dE/d(r^2); here f = 2 dE/d(r^2) */
  const double fv = -lv * (S * F + S2 * U);
  e[0] += lv * S * U;
  e[3] += lv2 * S * U2;
  const double r = sqrt(r2), rinv = 1.0 / r, rinv2 = rinv * rinv;
  const double fastE = qq * (rinv - 1.5 * c->rcinv + 0.5 * r2 * c->rc3inv);
  *f = fv + le * qq * (-rinv * rinv2 + c->rc3inv);
  e[1] += le * fastE;
  e[4] += le2 * fastE;
  const double a = c->beta * r;
  const double ec = erfc(a);
  const double g = ec + TWO_OVER_SQRT_PI * a * exp(-a * a);
  *fs = le * qq * ((1.0 - g) * rinv * rinv2 - c->rc3inv);
  const double slowE = qq * ((ec - 1.0) * rinv + 1.5 * c->rcinv - 0.5 * r2 * c->rc3inv);
  e[2] += le * slowE;
  e[5] += le2 * slowE;
  /* TI (calcForceEnergyTI, src/CudaComputeNonbondedKernel.cu:358-526): per
   * partition dE/d(lambda): vdw S (U + lv coeff F/2) (the shift's lambda
   * dependence included), elec and slow unscaled; e[6..11] = vdw up, vdw
   * down, elec up, elec down, slow up, slow down */
  if (up || down) {
    const int o = up ? 0 : 1;
    e[6 + o] += S * (U + lv * al->coeff * 0.5 * F);
    e[8 + o] += fastE;
    e[10 + o] += slowE;
  }
}

/* table-mode pair: columns (fast, -vdwB, vdwA, scor); e[0..2] = vdw, elec, slow */
typedef struct {
  const float* ft;
  const float* et;
  int N;
} Tables;

static void pair_table(const Tables* T, double r2, double qq, double A, double B, double* f, double* fs,
                       double* e) {
  const double x = (double)T->N / sqrt(r2) - 0.5;
  const double fl = floor(x);
  const double a = x - fl;
  int i0 = (int)fl;
  if (i0 < 0) i0 = 0;
  if (i0 > T->N - 1) i0 = T->N - 1;
  const int i1 = i0 + 1 < T->N ? i0 + 1 : T->N - 1;
  double t[4], u[4];
  for (int k = 0; k < 4; ++k) {
    t[k] = T->ft[4 * i0 + k] + a * ((double)T->ft[4 * i1 + k] - T->ft[4 * i0 + k]);
    u[k] = T->et[4 * i0 + k] + a * ((double)T->et[4 * i1 + k] - T->et[4 * i0 + k]);
  }
  *f = A * t[2] + B * t[1] + qq * t[0];
  *fs = qq * t[3];
  e[0] += A * u[2] + B * u[1];
  e[1] += qq * u[0];
  e[2] += qq * u[3];
}

/* cell list over an orthorhombic box; calls back for every i<j pair with r2 < cut2 */
typedef struct {
  int n, nc[3];
  double L[3], cs[3];
  int* head;
  int* next;
} Cells;

static void cells_build(Cells* C, int n, const double* pos, const double L[3], double cut) {
  C->n = n;
  for (int d = 0; d < 3; ++d) {
    C->L[d] = L[d];
    C->nc[d] = (int)floor(L[d] / cut);
    if (C->nc[d] < 1) C->nc[d] = 1;
    C->cs[d] = L[d] / C->nc[d];
  }
  const int ncell = C->nc[0] * C->nc[1] * C->nc[2];
  C->head = (int*)malloc(sizeof(int) * ncell);
  C->next = (int*)malloc(sizeof(int) * (n > 0 ? n : 1));
  for (int k = 0; k < ncell; ++k) C->head[k] = -1;
  for (int i = n - 1; i >= 0; --i) {
    int g[3];
    for (int d = 0; d < 3; ++d) {
      double w = pos[3 * i + d] - L[d] * floor(pos[3 * i + d] / L[d]);
      g[d] = (int)floor(w / C->cs[d]);
      if (g[d] >= C->nc[d]) g[d] = C->nc[d] - 1;
      if (g[d] < 0) g[d] = 0;
    }
    const int cidx = (g[0] * C->nc[1] + g[1]) * C->nc[2] + g[2];
    C->next[i] = C->head[cidx];
    C->head[cidx] = i;
  }
}

static void cells_free(Cells* C) {
  free(C->head);
  free(C->next);
}

/* Neighbour cells of cell (gx,gy,gz) as a de-duplicated list (handles nc < 3). */
static int cell_neighbours(const Cells* C, int gx, int gy, int gz, int* out) {
  int m = 0;
  int rng[3][3], cnt[3];
  const int g[3] = {gx, gy, gz};
  for (int d = 0; d < 3; ++d) {
    cnt[d] = 0;
    for (int o = -1; o <= 1; ++o) {
      int v = ((g[d] + o) % C->nc[d] + C->nc[d]) % C->nc[d];
      int dup = 0;
      for (int k = 0; k < cnt[d]; ++k)
        if (rng[d][k] == v) dup = 1;
      if (!dup) rng[d][cnt[d]++] = v;
    }
  }
  for (int a = 0; a < cnt[0]; ++a)
    for (int b = 0; b < cnt[1]; ++b)
      for (int c = 0; c < cnt[2]; ++c) out[m++] = (rng[0][a] * C->nc[1] + rng[1][b]) * C->nc[2] + rng[2][c];
  return m;
}

/*
 * FP64 oracle.  pos: 3n, q: scaled charges (q*sqrt(COULOMB)), type: n,
 * A,B: ntypes^2 row-major.  Outputs (any may be NULL): f, fs (3n), energies[3]
 * = vdw, elec, slow; virial[9], virialSlow[9] (virial[a*3+b] = sum over
 * pairs of F_ij,a * (x_i - x_j)_b, i.e. NAMD's sum_atoms F_a x_b);
 * npairs = number of pairs with r < rc.  nthreads <= 0: all available.
 */
static int nb_fp64_core(int n, const double* pos, const double* q, const int32_t* type, int ntypes,
                        const double* A, const double* B, const int64_t* eptr, const int32_t* eidx,
                        const double L[3], double cutoff, double switchdist, double beta, const int8_t* part,
                        const AlchF* al, const Tables* tabs, double* f, double* fs, double* energies,
                        int nenergies, double* virial, double* virialSlow, int64_t* npairs, int nthreads,
                        int forceSwitch) {
  PairConst c;
  memset(&c, 0, sizeof(c));
  if (forceSwitch) set_force_switch(&c, cutoff, switchdist);
  c.rc = cutoff;
  c.rc2 = cutoff * cutoff;
  c.rs2 = switchdist * switchdist;
  c.c1 = (switchdist < cutoff) ? 1.0 / ((c.rc2 - c.rs2) * (c.rc2 - c.rs2) * (c.rc2 - c.rs2)) : 0.0;
  c.rcinv = 1.0 / cutoff;
  c.rc3inv = c.rcinv * c.rcinv * c.rcinv;
  c.beta = beta;
  Cells C;
  cells_build(&C, n, pos, L, cutoff);
  const int ncell = C.nc[0] * C.nc[1] * C.nc[2];
  int nt = 1;
#ifdef _OPENMP
  nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
  double* fb = (double*)calloc((size_t)nt * 6 * n, sizeof(double));
  double* acc = (double*)calloc((size_t)nt * 30, sizeof(double));  /* 9 + 9 virial, up to 12 energies */
  int64_t* np = (int64_t*)calloc((size_t)nt, sizeof(int64_t));
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
#endif
  for (int cell = 0; cell < ncell; ++cell) {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double* F = fb + (size_t)tid * 6 * n;
    double* S = acc + (size_t)tid * 30;
    int nb[27];
    const int gz = cell % C.nc[2], gy = (cell / C.nc[2]) % C.nc[1], gx = cell / (C.nc[1] * C.nc[2]);
    const int m = cell_neighbours(&C, gx, gy, gz, nb);
    for (int i = C.head[cell]; i >= 0; i = C.next[i]) {
      for (int k = 0; k < m; ++k) {
        for (int j = C.head[nb[k]]; j >= 0; j = C.next[j]) {
          if (j <= i) continue;
          double d[3];
          for (int u = 0; u < 3; ++u) d[u] = min_image(pos[3 * j + u] - pos[3 * i + u], L[u]);
          const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
          if (!(r2 <= c.rc2)) continue;
          if (is_excluded(eptr, eidx, i, j)) continue;
          double ff, ffs;
          const int ab = type[i] * ntypes + type[j];
          if (part)
            pair_fp64_alch(&c, r2, q[i] * q[j], A[ab], B[ab], part[i], part[j], al, &ff, &ffs, &S[18]);
          else if (tabs)
            pair_table(tabs, r2, q[i] * q[j], A[ab], B[ab], &ff, &ffs, &S[18]);
          else
            pair_fp64(&c, r2, q[i] * q[j], A[ab], B[ab], &ff, &ffs, &S[18], &S[19], &S[20]);
          for (int u = 0; u < 3; ++u) {
            F[3 * i + u] += ff * d[u];
            F[3 * j + u] -= ff * d[u];
            F[3 * n + 3 * i + u] += ffs * d[u];
            F[3 * n + 3 * j + u] -= ffs * d[u];
          }
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
              S[a * 3 + b] -= ff * d[a] * d[b];
              S[9 + a * 3 + b] -= ffs * d[a] * d[b];
            }
          np[tid] += 1;
        }
      }
    }
  }
  if (f) memset(f, 0, sizeof(double) * 3 * n);
  if (fs) memset(fs, 0, sizeof(double) * 3 * n);
  double tot[30] = {0};
  int64_t ntot = 0;
  for (int t = 0; t < nt; ++t) {
    const double* F = fb + (size_t)t * 6 * n;
    if (f)
      for (int k = 0; k < 3 * n; ++k) f[k] += F[k];
    if (fs)
      for (int k = 0; k < 3 * n; ++k) fs[k] += F[3 * n + k];
    for (int k = 0; k < 30; ++k) tot[k] += acc[(size_t)t * 30 + k];
    ntot += np[t];
  }
  if (virial) memcpy(virial, tot, 9 * sizeof(double));
  if (virialSlow) memcpy(virialSlow, tot + 9, 9 * sizeof(double));
  if (energies) memcpy(energies, tot + 18, nenergies * sizeof(double));
  if (npairs) *npairs = ntot;
  free(fb);
  free(acc);
  free(np);
  cells_free(&C);
  return 0;
}

int oracle_nb_fp64(int n, const double* pos, const double* q, const int32_t* type, int ntypes, const double* A,
                   const double* B, const int64_t* eptr, const int32_t* eidx, const double L[3], double cutoff,
                   double switchdist, double beta, double* f, double* fs, double* energies, double* virial,
                   double* virialSlow, int64_t* npairs, int nthreads) {
  return nb_fp64_core(n, pos, q, type, ntypes, A, B, eptr, eidx, L, cutoff, switchdist, beta, NULL, NULL, NULL, f,
                      fs, energies, 3, virial, virialSlow, npairs, nthreads, 0);
}

/* The same with vdW force switching instead of energy switching. */
int oracle_nb_fp64_vdw_force_switch(int n, const double* pos, const double* q, const int32_t* type, int ntypes,
                                    const double* A, const double* B, const int64_t* eptr, const int32_t* eidx,
                                    const double L[3], double cutoff, double switchdist, double beta, double* f,
                                    double* fs, double* energies, double* virial, double* virialSlow,
                                    int64_t* npairs, int nthreads) {
  return nb_fp64_core(n, pos, q, type, ntypes, A, B, eptr, eidx, L, cutoff, switchdist, beta, NULL, NULL, NULL, f,
                      fs, energies, 3, virial, virialSlow, npairs, nthreads, 1);
}

/* force4/energy4: N float4 table entries (columns fast, -vdwB, vdwA, scor) */
int oracle_nb_table(int n, const double* pos, const double* q, const int32_t* type, int ntypes, const double* A,
                    const double* B, const int64_t* eptr, const int32_t* eidx, const double L[3], double cutoff,
                    double switchdist, int N, const float* force4, const float* energy4, double* f, double* fs,
                    double* energies, double* virial, double* virialSlow, int64_t* npairs, int nthreads) {
  Tables T;
  T.ft = force4;
  T.et = energy4;
  T.N = N;
  return nb_fp64_core(n, pos, q, type, ntypes, A, B, eptr, eidx, L, cutoff, switchdist, 0.0, NULL, NULL, &T, f, fs,
                      energies, 3, virial, virialSlow, npairs, nthreads, 0);
}

/* lam[12] = elecUp, elecDown, vdwUp, vdwDown, elec2Up, elec2Down, vdw2Up,
 * vdw2Down, shiftUp, shiftDown, shift2Up, shift2Down; energies[6] = vdw, elec,
 * slow and the same at lambda2. */
static int nb_fp64_alch_core(int n, const double* pos, const double* q, const int32_t* type, int ntypes,
                             const double* A, const double* B, const int64_t* eptr, const int32_t* eidx,
                             const double L[3], double cutoff, double switchdist, double beta, const int8_t* part,
                             const double* lam, int decouple, double* f, double* fs, double* energies,
                             double* virial, double* virialSlow, int64_t* npairs, int nthreads, int forceSwitch) {
  AlchF al;
  al.elecUp = lam[0]; al.elecDown = lam[1]; al.vdwUp = lam[2]; al.vdwDown = lam[3];
  al.elec2Up = lam[4]; al.elec2Down = lam[5]; al.vdw2Up = lam[6]; al.vdw2Down = lam[7];
  al.shUp = lam[8]; al.shDown = lam[9]; al.sh2Up = lam[10]; al.sh2Down = lam[11];
  al.decouple = decouple;
  al.coeff = lam[12];
  return nb_fp64_core(n, pos, q, type, ntypes, A, B, eptr, eidx, L, cutoff, switchdist, beta, part, &al, NULL, f,
                      fs, energies, 12, virial, virialSlow, npairs, nthreads, forceSwitch);
}

int oracle_nb_fp64_alch(int n, const double* pos, const double* q, const int32_t* type, int ntypes,
                        const double* A, const double* B, const int64_t* eptr, const int32_t* eidx,
                        const double L[3], double cutoff, double switchdist, double beta, const int8_t* part,
                        const double* lam, int decouple, double* f, double* fs, double* energies, double* virial,
                        double* virialSlow, int64_t* npairs, int nthreads) {
  return nb_fp64_alch_core(n, pos, q, type, ntypes, A, B, eptr, eidx, L, cutoff, switchdist, beta, part, lam,
                           decouple, f, fs, energies, virial, virialSlow, npairs, nthreads, 0);
}

/* The same with vdW force switching (alchVdwForceSwitching of the GPU kernel,
 * vdwForceSwitching of the CPU FEP/TI pair code). */
int oracle_nb_fp64_alch_vdw_force_switch(int n, const double* pos, const double* q, const int32_t* type,
                                         int ntypes, const double* A, const double* B, const int64_t* eptr,
                                         const int32_t* eidx, const double L[3], double cutoff, double switchdist,
                                         double beta, const int8_t* part, const double* lam, int decouple,
                                         double* f, double* fs, double* energies, double* virial,
                                         double* virialSlow, int64_t* npairs, int nthreads) {
  return nb_fp64_alch_core(n, pos, q, type, ntypes, A, B, eptr, eidx, L, cutoff, switchdist, beta, part, lam,
                           decouple, f, fs, energies, virial, virialSlow, npairs, nthreads, 1);
}

/* ---- force-matching adjoint ------------------------------------------------
 * Pair force f (F_i += f d, F_j -= f d, d = x_j - x_i) as in pair_fp64:
 *   f = A cA(r) + B cB(r) + qq fe(r),
 *   cA = r^-12 dS_r - 12 S r^-14,  cB = 6 S r^-8 - r^-6 dS_r,
 *   fe = -G(beta r)/r^3,  G = erfc(a) + 2a/sqrt(pi) exp(-a^2)   (fast + slow)
 * With s = (g_i - g_j) . d = dL/df for this pair:
 *   dL/dA_{t_i t_j} += s cA, dL/dB_{t_i t_j} += s cB (ordered bins),
 *   dL/dq_i += s q_j fe, dL/dq_j += s q_i fe (scaled charges).
 */
int oracle_nb_adjoint_fp64(int n, const double* pos, const double* q, const int32_t* type, int ntypes,
                           const int64_t* eptr, const int32_t* eidx, const double L[3], double cutoff,
                           double switchdist, double beta, const double* g, double* gA, double* gB, double* gq,
                           int nthreads) {
  PairConst c;
  c.rc = cutoff;
  c.rc2 = cutoff * cutoff;
  c.rs2 = switchdist * switchdist;
  c.c1 = (switchdist < cutoff) ? 1.0 / ((c.rc2 - c.rs2) * (c.rc2 - c.rs2) * (c.rc2 - c.rs2)) : 0.0;
  c.rcinv = 1.0 / cutoff;
  c.rc3inv = c.rcinv * c.rcinv * c.rcinv;
  c.beta = beta;
  Cells C;
  cells_build(&C, n, pos, L, cutoff);
  const int ncell = C.nc[0] * C.nc[1] * C.nc[2];
  int nt = 1;
#ifdef _OPENMP
  nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
  const size_t nb2 = (size_t)ntypes * ntypes;
  double* bins = (double*)calloc((size_t)nt * 2 * nb2, sizeof(double));
  double* gqt = (double*)calloc((size_t)nt * n, sizeof(double));
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
#endif
  for (int cell = 0; cell < ncell; ++cell) {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double* bA = bins + (size_t)tid * 2 * nb2;
    double* bB = bA + nb2;
    double* Q = gqt + (size_t)tid * n;
    int nb[27];
    const int gz = cell % C.nc[2], gy = (cell / C.nc[2]) % C.nc[1], gx = cell / (C.nc[1] * C.nc[2]);
    const int m = cell_neighbours(&C, gx, gy, gz, nb);
    for (int i = C.head[cell]; i >= 0; i = C.next[i]) {
      for (int k = 0; k < m; ++k) {
        for (int j = C.head[nb[k]]; j >= 0; j = C.next[j]) {
          if (j <= i) continue;
          double d[3];
          for (int u = 0; u < 3; ++u) d[u] = min_image(pos[3 * j + u] - pos[3 * i + u], L[u]);
          const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
          if (!(r2 <= c.rc2)) continue;
          if (is_excluded(eptr, eidx, i, j)) continue;
          const double sp = (g[3 * i] - g[3 * j]) * d[0] + (g[3 * i + 1] - g[3 * j + 1]) * d[1] +
                            (g[3 * i + 2] - g[3 * j + 2]) * d[2];
          const double r = sqrt(r2), ri2 = 1.0 / r2, ri6 = ri2 * ri2 * ri2;
          double S = 1.0, dS_r = 0.0;
          if (r2 > c.rs2) {
            const double dd = c.rc2 - r2;
            S = dd * dd * (c.rc2 + 2.0 * r2 - 3.0 * c.rs2) * c.c1;
            dS_r = 2.0 * c.c1 * (-2.0 * dd * (c.rc2 + 2.0 * r2 - 3.0 * c.rs2) + 2.0 * dd * dd);
          }
          const double cA = ri6 * ri6 * (dS_r - 12.0 * S * ri2);
          const double cB = ri6 * (6.0 * S * ri2 - dS_r);
          const double a = c.beta * r;
          const double G = erfc(a) + TWO_OVER_SQRT_PI * a * exp(-a * a);
          const double fe = -G * ri2 / r;
          const size_t bin = (size_t)type[i] * ntypes + type[j];
          bA[bin] += sp * cA;
          bB[bin] += sp * cB;
          Q[i] += sp * q[j] * fe;
          Q[j] += sp * q[i] * fe;
        }
      }
    }
  }
  memset(gA, 0, nb2 * sizeof(double));
  memset(gB, 0, nb2 * sizeof(double));
  memset(gq, 0, (size_t)n * sizeof(double));
  for (int t = 0; t < nt; ++t) {
    for (size_t k = 0; k < nb2; ++k) {
      gA[k] += bins[(size_t)t * 2 * nb2 + k];
      gB[k] += bins[(size_t)t * 2 * nb2 + nb2 + k];
    }
    for (int k = 0; k < n; ++k) gq[k] += gqt[(size_t)t * n + k];
  }
  free(bins);
  free(gqt);
  cells_free(&C);
  return 0;
}

/* ---- canonical FP32 pair sets ------------------------------------------- */
static inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

uint64_t oracle_pair_hash(int64_t i, int64_t j, int64_t n) {
  const int64_t a = i < j ? i : j, b = i < j ? j : i;
  return splitmix64((uint64_t)(a * n + b));
}

/*
 * pos must be the effective positions (multiples of 2^-17 A); the minimum-
 * image displacement is then exact in FP64 and in FP32.  For each unordered
 * pair: r2f = fmaf(dz,dz,fmaf(dy,dy,dx*dx)).  Counts:
 *   out[0] non-excluded pairs with r2f < cut2f
 *   out[1] non-excluded pairs with r2f < plcut2f
 *   out[2] excluded pairs with r2f < plcut2f
 * and checksums (sum and xor of oracle_pair_hash) of the sets in out[0], out[1]:
 *   sums[0], xors[0] for the cutoff set, sums[1], xors[1] for the list set.
 * If pairs != NULL (capacity maxpairs pairs), the list set is written as
 * (i, j) with i < j; out[3] = pairs written (stops at capacity).
 */
int oracle_pairs_fp32(int n, const double* pos, const int64_t* eptr, const int32_t* eidx, const double L[3],
                      float cut2f, float plcut2f, int64_t* out, uint64_t* sums, uint64_t* xors, int32_t* pairs,
                      int64_t maxpairs) {
  const double rl = sqrt((double)plcut2f) * 1.0001 + 1e-6;
  Cells C;
  cells_build(&C, n, pos, L, rl);
  const int ncell = C.nc[0] * C.nc[1] * C.nc[2];
  int64_t cnt[3] = {0, 0, 0}, written = 0;
  uint64_t s0 = 0, s1 = 0, x0 = 0, x1 = 0;
  for (int cell = 0; cell < ncell; ++cell) {
    int nb[27];
    const int gz = cell % C.nc[2], gy = (cell / C.nc[2]) % C.nc[1], gx = cell / (C.nc[1] * C.nc[2]);
    const int m = cell_neighbours(&C, gx, gy, gz, nb);
    for (int i = C.head[cell]; i >= 0; i = C.next[i])
      for (int k = 0; k < m; ++k)
        for (int j = C.head[nb[k]]; j >= 0; j = C.next[j]) {
          if (j <= i) continue;
          float df[3];
          for (int u = 0; u < 3; ++u) df[u] = (float)min_image(pos[3 * j + u] - pos[3 * i + u], L[u]);
          const float r2 = fmaf(df[2], df[2], fmaf(df[1], df[1], df[0] * df[0]));
          if (!(r2 < plcut2f)) continue;
          if (is_excluded(eptr, eidx, i, j)) {
            cnt[2]++;
            continue;
          }
          const uint64_t h = oracle_pair_hash(i, j, n);
          cnt[1]++;
          s1 += h;
          x1 ^= h;
          if (pairs && written < maxpairs) {
            pairs[2 * written] = i;
            pairs[2 * written + 1] = j;
            written++;
          }
          if (r2 < cut2f) {
            cnt[0]++;
            s0 += h;
            x0 ^= h;
          }
        }
  }
  out[0] = cnt[0];
  out[1] = cnt[1];
  out[2] = cnt[2];
  out[3] = written;
  if (sums) { sums[0] = s0; sums[1] = s1; }
  if (xors) { xors[0] = x0; xors[1] = x1; }
  cells_free(&C);
  return 0;
}

/* ---- PME real space ------------------------------------------------------ */

/* order-n weights by the B-spline recursion, GPU indexing (theta[k] at frxi + k) */
static void bspline_gpu(double w, int order, double* th, double* dth) {
  for (int k = 0; k < order; ++k) th[k] = 0.0;
  th[0] = 1.0 - w;
  th[1] = w;
  for (int n = 3; n <= order; ++n) {
    if (n == order && dth) {
      dth[0] = -th[0];
      for (int k = 1; k < order; ++k) dth[k] = th[k - 1] - th[k];
    }
    const double div = 1.0 / (n - 1);
    th[n - 1] = div * w * th[n - 2];
    for (int j = 1; j <= n - 2; ++j) th[n - j - 1] = div * ((w + j) * th[n - j - 2] + (n - j - w) * th[n - j - 1]);
    th[0] = div * (1.0 - w) * th[0];
  }
}

/* atoms: float4 (x,y,z fractional in [0,1), q); grid: double[K1*K2*K3], x fastest; accumulates */
int oracle_pme_spread_gpu(int n, const float* atoms, int K1, int K2, int K3, int order, double* grid) {
  double tx[8], ty[8], tz[8];
  for (int a = 0; a < n; ++a) {
    const double q = atoms[4 * a + 3];
    const float frx = (float)K1 * atoms[4 * a], fry = (float)K2 * atoms[4 * a + 1], frz = (float)K3 * atoms[4 * a + 2];
    const int ix = (int)frx, iy = (int)fry, iz = (int)frz;
    bspline_gpu((double)frx - ix, order, tx, NULL);
    bspline_gpu((double)fry - iy, order, ty, NULL);
    bspline_gpu((double)frz - iz, order, tz, NULL);
    for (int c = 0; c < order; ++c)
      for (int b = 0; b < order; ++b)
        for (int d = 0; d < order; ++d) {
          const int x = (ix + d) % K1, y = (iy + b) % K2, z = (iz + c) % K3;
          grid[x + (size_t)K1 * (y + (size_t)K2 * z)] += q * tx[d] * ty[b] * tz[c];
        }
  }
  return 0;
}

/* fractional-gradient forces (-q K1 sum dtx ty tz Q, ...) or Cartesian with recip (row-major a*,b*,c*) */
int oracle_pme_gather_gpu(int n, const float* atoms, int K1, int K2, int K3, int order, const double* grid,
                          const double* recip, double* force) {
  double tx[8], ty[8], tz[8], dx[8], dy[8], dz[8];
  for (int a = 0; a < n; ++a) {
    const double q = atoms[4 * a + 3];
    const float frx = (float)K1 * atoms[4 * a], fry = (float)K2 * atoms[4 * a + 1], frz = (float)K3 * atoms[4 * a + 2];
    const int ix = (int)frx, iy = (int)fry, iz = (int)frz;
    bspline_gpu((double)frx - ix, order, tx, dx);
    bspline_gpu((double)fry - iy, order, ty, dy);
    bspline_gpu((double)frz - iz, order, tz, dz);
    double f1 = 0, f2 = 0, f3 = 0;
    for (int c = 0; c < order; ++c)
      for (int b = 0; b < order; ++b)
        for (int d = 0; d < order; ++d) {
          const int x = (ix + d) % K1, y = (iy + b) % K2, z = (iz + c) % K3;
          const double Q = grid[x + (size_t)K1 * (y + (size_t)K2 * z)];
          f1 += dx[d] * ty[b] * tz[c] * Q;
          f2 += tx[d] * dy[b] * tz[c] * Q;
          f3 += tx[d] * ty[b] * dz[c] * Q;
        }
    const double g1 = -q * K1 * f1, g2 = -q * K2 * f2, g3 = -q * K3 * f3;
    if (recip) {
      force[3 * a] = g1 * recip[0] + g2 * recip[3] + g3 * recip[6];
      force[3 * a + 1] = g1 * recip[1] + g2 * recip[4] + g3 * recip[7];
      force[3 * a + 2] = g1 * recip[2] + g2 * recip[5] + g3 * recip[8];
    } else {
      force[3 * a] = g1;
      force[3 * a + 1] = g2;
      force[3 * a + 2] = g3;
    }
  }
  return 0;
}

/* scale_coordinates (PmeBase.inl:17-57): p = K (s - floor s), s = shift + r.a*, shift = ((K+order-1)/2)/K */
int oracle_pme_scale_coordinates(int n, const double* pos, const double origin[3], const double* recip, int K1,
                                 int K2, int K3, int order, double* p) {
  const int K[3] = {K1, K2, K3};
  for (int d = 0; d < 3; ++d) {
    const double shift = (double)((K[d] + order - 1) / 2) / (double)K[d];
    for (int i = 0; i < n; ++i) {
      const double px = pos[3 * i] - origin[0], py = pos[3 * i + 1] - origin[1], pz = pos[3 * i + 2] - origin[2];
      const double s = shift + px * recip[3 * d] + py * recip[3 * d + 1] + pz * recip[3 * d + 2];
      double v = K[d] * (s - floor(s));
      if (v == K[d]) v = 0;
      p[3 * i + d] = v;
    }
  }
  return 0;
}

/* compute_b_spline (PmeBase.inl:85-148) for one atom; M, dM: 3*order */
static void compute_b_spline_cpu(const double fr[3], double* M, double* dM, int order) {
  for (int d = 0; d < 3; ++d) {
    double* Mx = M + d * order - 1;  /* 1-based like the reference */
    double* dMx = dM + d * order - 1;
    const double x = fr[d];
    const double x1 = 1.0 - x;
    Mx[1] = 0.5 * x1 * x1;
    Mx[2] = x1 * x + 0.5;
    Mx[3] = 0.5 * x * x;
    Mx[order] = 0.0;
    for (int n = 4; n <= order - 1; ++n) {
      const double div = 1.0 / (n - 1);
      Mx[n] = x * div * Mx[n - 1];
      for (int j = 1; j <= n - 2; ++j) Mx[n - j] = ((x + j) * Mx[n - j - 1] + (n - x - j) * Mx[n - j]) * div;
      Mx[1] *= (1.0 - x) * div;
    }
    dMx[1] = -Mx[1];
    for (int j = 2; j <= order; ++j) dMx[j] = Mx[j - 1] - Mx[j];
    const double div = 1.0 / (order - 1);
    Mx[order] = x * div * Mx[order - 1];
    for (int j = 1; j <= order - 2; ++j)
      Mx[order - j] = ((x + j) * Mx[order - j - 1] + (order - x - j) * Mx[order - j]) * div;
    Mx[1] *= (1.0 - x) * div;
  }
}

/*
 * fill_charges (PmeRealSpace.C:237-345) on a plain wrapped grid: grid point
 * ((int)p - order + 1 + k) mod K gets M[k].  order 4 uses the explicit cubics
 * in float as fill_charges_order4 does (fr = (float)(p - (double)u)).
 */
int oracle_pme_spread_cpu(int n, const double* p, const double* q, int K1, int K2, int K3, int order, double* grid) {
  double M[24], dM[24];
  const int K[3] = {K1, K2, K3};
  for (int a = 0; a < n; ++a) {
    int u[3];
    double fr[3];
    for (int d = 0; d < 3; ++d) {
      u[d] = (int)p[3 * a + d];
      fr[d] = (double)(float)(p[3 * a + d] - (double)u[d]);
    }
    if (order == 4) {
      for (int d = 0; d < 3; ++d) {
        const float f = (float)fr[d];
        M[4 * d + 0] = (float)((((-1. / 6.) * f + 0.5) * f - 0.5) * f + (1. / 6.));
        M[4 * d + 1] = (float)(((0.5 * f - 1.0) * f) * f + (2. / 3.));
        M[4 * d + 2] = (float)(((-0.5 * f + 0.5) * f + 0.5) * f + (1. / 6.));
        M[4 * d + 3] = (float)((1. / 6.) * f * f * f);
      }
    } else {
      compute_b_spline_cpu(fr, M, dM, order);
    }
    for (int i = 0; i < order; ++i)
      for (int j = 0; j < order; ++j)
        for (int k = 0; k < order; ++k) {
          int x = u[0] - order + 1 + i, y = u[1] - order + 1 + j, z = u[2] - order + 1 + k;
          x = ((x % K[0]) + K[0]) % K[0];
          y = ((y % K[1]) + K[1]) % K[1];
          z = ((z % K[2]) + K[2]) % K[2];
          grid[x + (size_t)K1 * (y + (size_t)K2 * z)] += q[a] * M[i] * M[order + j] * M[2 * order + k];
        }
  }
  return 0;
}

/* compute_forces (PmeRealSpace.C:347-425) on a plain wrapped grid; fractional-gradient forces */
int oracle_pme_gather_cpu(int n, const double* p, const double* q, int K1, int K2, int K3, int order,
                          const double* grid, double* force) {
  double M[24], dM[24];
  const int K[3] = {K1, K2, K3};
  for (int a = 0; a < n; ++a) {
    int u[3];
    double fr[3];
    for (int d = 0; d < 3; ++d) {
      u[d] = (int)p[3 * a + d];
      fr[d] = p[3 * a + d] - (double)u[d];
    }
    compute_b_spline_cpu(fr, M, dM, order);
    double f1 = 0, f2 = 0, f3 = 0;
    for (int i = 0; i < order; ++i)
      for (int j = 0; j < order; ++j)
        for (int k = 0; k < order; ++k) {
          int x = u[0] - order + 1 + i, y = u[1] - order + 1 + j, z = u[2] - order + 1 + k;
          x = ((x % K[0]) + K[0]) % K[0];
          y = ((y % K[1]) + K[1]) % K[1];
          z = ((z % K[2]) + K[2]) % K[2];
          const double t = grid[x + (size_t)K1 * (y + (size_t)K2 * z)];
          f1 -= K1 * dM[i] * M[order + j] * M[2 * order + k] * t;
          f2 -= K2 * M[i] * dM[order + j] * M[2 * order + k] * t;
          f3 -= K3 * M[i] * M[order + j] * dM[2 * order + k] * t;
        }
    force[3 * a] = q[a] * f1;
    force[3 * a + 1] = q[a] * f2;
    force[3 * a + 2] = q[a] * f3;
  }
  return 0;
}

/* B-spline partition of unity check helper: sum of weights and derivative weights (GPU recursion) */
int oracle_bspline_gpu(double w, int order, double* th, double* dth) {
  bspline_gpu(w, order, th, dth);
  return 0;
}

/* ---- exact pair sets (NAMD CPU rule) -------------------------------------
 * With effective positions on the 2^-17 A grid the FP64 displacement and r^2
 * are exact, so this is NAMD's CPU membership (pairlist build r2 <= plcutoff2,
 * ComputeNonbondedBase.h:1282; per step r2 + r2_delta <= cutoff2 + r2_delta,
 * ComputeNonbondedInl.h:80-117) without rounding.  Same outputs as
 * oracle_pairs_fp32. */
int oracle_pairs_exact(int n, const double* pos, const int64_t* eptr, const int32_t* eidx, const double L[3],
                       double cut2, double plcut2, int64_t* out, uint64_t* sums, uint64_t* xors, int32_t* pairs,
                       int64_t maxpairs) {
  Cells C;
  cells_build(&C, n, pos, L, sqrt(plcut2) * 1.0001 + 1e-6);
  const int ncell = C.nc[0] * C.nc[1] * C.nc[2];
  int64_t cnt[3] = {0, 0, 0}, written = 0;
  uint64_t s0 = 0, s1 = 0, x0 = 0, x1 = 0;
  for (int cell = 0; cell < ncell; ++cell) {
    int nb[27];
    const int gz = cell % C.nc[2], gy = (cell / C.nc[2]) % C.nc[1], gx = cell / (C.nc[1] * C.nc[2]);
    const int m = cell_neighbours(&C, gx, gy, gz, nb);
    for (int i = C.head[cell]; i >= 0; i = C.next[i])
      for (int k = 0; k < m; ++k)
        for (int j = C.head[nb[k]]; j >= 0; j = C.next[j]) {
          if (j <= i) continue;
          double d[3];
          for (int u = 0; u < 3; ++u) d[u] = min_image(pos[3 * j + u] - pos[3 * i + u], L[u]);
          const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
          if (!(r2 <= plcut2)) continue;
          if (is_excluded(eptr, eidx, i, j)) {
            cnt[2]++;
            continue;
          }
          const uint64_t h = oracle_pair_hash(i, j, n);
          cnt[1]++;
          s1 += h;
          x1 ^= h;
          if (pairs && written < maxpairs) {
            pairs[2 * written] = i;
            pairs[2 * written + 1] = j;
            written++;
          }
          if (r2 <= cut2) {
            cnt[0]++;
            s0 += h;
            x0 ^= h;
          }
        }
  }
  out[0] = cnt[0];
  out[1] = cnt[1];
  out[2] = cnt[2];
  out[3] = written;
  if (sums) { sums[0] = s0; sums[1] = s1; }
  if (xors) { xors[0] = x0; xors[1] = x1; }
  cells_free(&C);
  return 0;
}

/* ---- NAMD CPU nonbonded path: FP64 r^2 cubic-spline tables ----------------
 * Restates ComputeNonbondedUtil::select (ComputeNonbondedUtil.C:604-1124) for
 * PME + C1 splitting + LJ energy switching: bins of 64 per octave of
 * r^2 + r2_delta starting at r2_delta = 1/16 (:607-626), node values of
 * fast/scor/slow/vdwa/vdwb energy and gradient (:763-1037; gradients are
 * d/d(r^2)), linear continuation below r2_limit for fast/vdwa/vdwb (:1080-1100),
 * Hermite cubic per bin with two refinement steps (:1101-1124).  The pair loop
 * restates ComputeNonbondedBase2.h (table index from the double's exponent and
 * top 6 mantissa bits, :86-92; the -6d,-4c,-2b,-a table_four layout of
 * ComputeNonbondedUtil.C:1263-1297 and its Horner order, :207-330). */
typedef struct {
  int n;
  double r2_delta;
  int r2_delta_exp;
  double* r2_table;  /* n */
  double* t[5];      /* fast, scor, slow, vdwa, vdwb: 4*n each (a, b, c, d) */
} NamdTables;

static void namd_tables_build(NamdTables* T, double cutoff, double switchOn, double ewaldcof, double limitDist) {
  const double cutoff2 = cutoff * cutoff;
  const double switchOn2 = switchOn * switchOn;
  const double c0 = (switchOn < cutoff) ? 1.0 / (cutoff2 - switchOn2) : 0.0;
  const double c1 = c0 * c0 * c0;
  const double c3 = 3.0 * (cutoff2 - switchOn2);
  const double pi_ewaldcof = 1.12837916709551 * ewaldcof;
  double r2_delta = 1.0;
  int r2_delta_exp = 0;
  while (r2_delta > 0.1) { r2_delta /= 2.0; r2_delta_exp += 1; }
  double r2_tmp = 1.0;
  int cutoff2_exp = 0;
  while ((cutoff2 + r2_delta) > r2_tmp) { r2_tmp *= 2.0; cutoff2_exp += 1; }
  const int n = (r2_delta_exp + cutoff2_exp) * 64 + 1;
  T->n = n;
  T->r2_delta = r2_delta;
  T->r2_delta_exp = r2_delta_exp;
  T->r2_table = (double*)calloc(n + 1, sizeof(double));
  for (int k = 0; k < 5; ++k) T->t[k] = (double*)calloc(4 * (size_t)n + 8, sizeof(double));
  double r2_limit = limitDist * limitDist;
  if (r2_limit < r2_delta) r2_limit = r2_delta;
  int r2_delta_i = 0;
  T->r2_table[0] = r2_delta;
  for (int i = 1; i < n; ++i) {
    const double r2_base = r2_delta * (1 << (i / 64));
    const double r2_del = r2_base / 64.0;
    const double r2 = r2_base - r2_delta + r2_del * (i % 64);
    const double r = sqrt(r2);
    if (r2 <= r2_limit) r2_delta_i = i;
    const double r_1 = 1.0 / r;
    const double r_2 = 1.0 / r2;
    const double tmp_a = r * ewaldcof;
    const double tmp_b = erfc(tmp_a);
    const double corr_energy = tmp_b;
    const double corr_gradient = pi_ewaldcof * exp(-(tmp_a * tmp_a)) * r + tmp_b;
    const double slow_energy = 0.5 / cutoff * (3.0 - (r2 / cutoff2));
    double slow_gradient = -1.0 / cutoff2 * (r / cutoff);
    const double scor_energy = slow_energy + (corr_energy - 1.0) / r;
    double scor_gradient = slow_gradient - (corr_gradient - 1.0) / r2;
    const double fast_energy = 1.0 / r - slow_energy;
    double fast_gradient = -1.0 / r2 - slow_gradient;
    fast_gradient *= 0.5 * r_1;
    scor_gradient *= 0.5 * r_1;
    slow_gradient *= 0.5 * r_1;
    const double r_6 = r_2 * r_2 * r_2;
    const double r_12 = r_6 * r_6;
    const double c2 = cutoff2 - r2;
    const double c4 = c2 * (c3 - 2.0 * c2);
    const double switchVal = (r2 > switchOn2 ? c2 * c4 * c1 : 1.0);
    const double dSwitchVal = (r2 > switchOn2 ? 2 * c1 * (c2 * c2 - c4) : 0.0);
    const double vdwa_energy = switchVal * r_12;
    const double vdwb_energy = switchVal * r_6;
    const double vdwa_gradient = (dSwitchVal - 6.0 * switchVal * r_2) * r_12;
    const double vdwb_gradient = (dSwitchVal - 3.0 * switchVal * r_2) * r_6;
    const double ev[5] = {fast_energy, scor_energy, slow_energy, vdwa_energy, vdwb_energy};
    const double gv[5] = {fast_gradient, scor_gradient, slow_gradient, vdwa_gradient, vdwb_gradient};
    for (int k = 0; k < 5; ++k) {
      T->t[k][4 * i + 0] = ev[k];
      T->t[k][4 * i + 1] = gv[k];
      T->t[k][4 * i + 2] = 0.0;
      T->t[k][4 * i + 3] = 0.0;
    }
    T->r2_table[i] = r2 + r2_delta;
  }
  for (int k = 0; k < 5; ++k) {
    double* t0 = T->t[k];
    const int smooth_short = (k == 0 || k == 3 || k == 4);
    t0[0] = t0[4] - t0[5] * (r2_delta / 64.0);
    t0[1] = t0[5];
    t0[2] = 0;
    t0[3] = 0;
    double energy0 = 0, gradient0 = 0, r20 = 0;
    if (smooth_short) {
      energy0 = t0[4 * r2_delta_i];
      gradient0 = t0[4 * r2_delta_i + 1];
      r20 = T->r2_table[r2_delta_i];
      t0[0] = energy0 - gradient0 * (r20 - T->r2_table[0]);
      t0[1] = gradient0;
    }
    double* t = t0;
    for (int i = 0; i < n - 1; ++i, t += 4) {
      const double x = (r2_delta * (1 << (i / 64))) / 64.0;
      if (smooth_short && i + 1 < r2_delta_i) {
        t[4] = energy0 - gradient0 * (r20 - T->r2_table[i + 1]);
        t[5] = gradient0;
      }
      const double v1 = t[0], g1 = t[1], v2 = t[4], g2 = t[5];
      double c = (3.0 * (v2 - v1) - x * (2.0 * g1 + g2)) / (x * x);
      double d = (-2.0 * (v2 - v1) + x * (g1 + g2)) / (x * x * x);
      for (int it = 0; it < 2; ++it) {
        const double dv = (v1 - v2) + ((d * x + c) * x + g1) * x;
        const double dg = (g1 - g2) + (3.0 * d * x + 2.0 * c) * x;
        c -= (3.0 * dv - x * dg) / (x * x);
        d -= (-2.0 * dv + x * dg) / (x * x * x);
      }
      t[2] = c;
      t[3] = d;
    }
  }
}

static void namd_tables_free(NamdTables* T) {
  free(T->r2_table);
  for (int k = 0; k < 5; ++k) free(T->t[k]);
}

/* Table entries for tests: writes n, r2_table (n) and the 5 tables (5*4*n, order
 * fast, scor, slow, vdwa, vdwb; each entry a,b,c,d) if the pointers are set. */
int oracle_namd_tables(double cutoff, double switchOn, double ewaldcof, double limitDist, int* n, double* r2_table,
                       double* tables) {
  NamdTables T;
  namd_tables_build(&T, cutoff, switchOn, ewaldcof, limitDist);
  *n = T.n;
  if (r2_table) memcpy(r2_table, T.r2_table, sizeof(double) * T.n);
  if (tables)
    for (int k = 0; k < 5; ++k) memcpy(tables + (size_t)k * 4 * T.n, T.t[k], sizeof(double) * 4 * T.n);
  namd_tables_free(&T);
  return 0;
}

/* buildForceAndEnergyTable for the nonbonded set (CudaNonbondedTables.C:158-203
 * and :273-300): entry i (1 <= i < N) samples the splines at
 * r2 = r^2 + r2_delta with r = N / (i + 0.5); zero beyond the cutoff; force
 * column = prefac * 2 * d(spline)/d(r^2), energy column = prefac * spline;
 * columns (x, y, z, w) = (fast, -vdwb, vdwa, scor); entry 0: zero force,
 * energy of entry 1.  force4, energy4: 4*N floats. */
int oracle_namd_gpu_tables(int N, double cutoff, double switchOn, double ewaldcof, float* force4, float* energy4) {
  NamdTables T;
  namd_tables_build(&T, cutoff, switchOn, ewaldcof, 0.0);
  const int r2_delta_expc = 64 * (T.r2_delta_exp - 1023);
  const double* src[4] = {T.t[0], T.t[4], T.t[3], T.t[1]};
  const double prefac[4] = {1.0, -1.0, 1.0, 1.0};
  memset(force4, 0, sizeof(float) * 4 * (size_t)N);
  memset(energy4, 0, sizeof(float) * 4 * (size_t)N);
  for (int i = 1; i < N; ++i) {
    const double r = (double)N / ((double)i + 0.5);
    if (r > cutoff) continue;
    const double r2 = r * r + T.r2_delta;
    uint64_t bits;
    memcpy(&bits, &r2, sizeof bits);
    const int ti = (int)(bits >> 46) + r2_delta_expc;
    const double diffa = r2 - T.r2_table[ti];
    for (int k = 0; k < 4; ++k) {
      const double* t = src[k] + 4 * (size_t)ti;
      const double grad = (3.0 * t[3] * diffa + 2.0 * t[2]) * diffa + t[1];
      const double ener = t[0] + diffa * ((t[3] * diffa + t[2]) * diffa + t[1]);
      force4[4 * (size_t)i + k] = (float)(prefac[k] * 2.0 * grad);
      energy4[4 * (size_t)i + k] = (float)(prefac[k] * ener);
    }
  }
  for (int k = 0; k < 4; ++k) energy4[k] = energy4[4 + k];
  namd_tables_free(&T);
  return 0;
}

/* The CPU pair path over normal (non-excluded) pairs with r2 + r2_delta <=
 * cutoff2 + r2_delta.  Outputs like oracle_nb_fp64: f (vdw + fast), fs (scor),
 * energies (vdw, fast elec, slow), virial, virialSlow, npairs. */
int oracle_namd_cpu(int n, const double* pos, const double* q, const int32_t* type, int ntypes, const double* A,
                    const double* B, const int64_t* eptr, const int32_t* eidx, const double L[3], double cutoff,
                    double switchdist, double beta, double* f, double* fs, double* energies, double* virial,
                    double* virialSlow, int64_t* npairs, int nthreads) {
  NamdTables T;
  namd_tables_build(&T, cutoff, switchdist, beta, 0.0);
  const double r2_delta = T.r2_delta;
  const int r2_delta_expc = 64 * (T.r2_delta_exp - 1023);
  const double cutoff2_delta = cutoff * cutoff + r2_delta;
  /* table_four: 16 per bin, -6d -4c -2b -a for vdwa, vdwb, fast, scor */
  double* four = (double*)malloc(sizeof(double) * 16 * (size_t)T.n);
  const int src[4] = {3, 4, 0, 1};
  for (int i = 0; i < T.n; ++i)
    for (int s = 0; s < 4; ++s) {
      const double* t = T.t[src[s]] + 4 * i;
      four[16 * i + 4 * s + 0] = -6. * t[3];
      four[16 * i + 4 * s + 1] = -4. * t[2];
      four[16 * i + 4 * s + 2] = -2. * t[1];
      four[16 * i + 4 * s + 3] = -1. * t[0];
    }
  Cells C;
  cells_build(&C, n, pos, L, cutoff);
  const int ncell = C.nc[0] * C.nc[1] * C.nc[2];
  int nt = 1;
#ifdef _OPENMP
  nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
  double* fb = (double*)calloc((size_t)nt * 6 * n, sizeof(double));
  double* acc = (double*)calloc((size_t)nt * 30, sizeof(double));  /* 9 + 9 virial, up to 12 energies */
  int64_t* np = (int64_t*)calloc((size_t)nt, sizeof(int64_t));
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
#endif
  for (int cell = 0; cell < ncell; ++cell) {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double* F = fb + (size_t)tid * 6 * n;
    double* S = acc + (size_t)tid * 30;
    int nb[27];
    const int gz = cell % C.nc[2], gy = (cell / C.nc[2]) % C.nc[1], gx = cell / (C.nc[1] * C.nc[2]);
    const int m = cell_neighbours(&C, gx, gy, gz, nb);
    for (int i = C.head[cell]; i >= 0; i = C.next[i])
      for (int k = 0; k < m; ++k)
        for (int j = C.head[nb[k]]; j >= 0; j = C.next[j]) {
          if (j <= i) continue;
          /* p_ij = p_i - p_j, minimum image */
          const double p_ij_x = -min_image(pos[3 * j] - pos[3 * i], L[0]);
          const double p_ij_y = -min_image(pos[3 * j + 1] - pos[3 * i + 1], L[1]);
          const double p_ij_z = -min_image(pos[3 * j + 2] - pos[3 * i + 2], L[2]);
          double r2 = p_ij_x * p_ij_x + r2_delta;
          r2 += p_ij_y * p_ij_y;
          r2 += p_ij_z * p_ij_z;
          if (!(r2 <= cutoff2_delta)) continue;
          if (is_excluded(eptr, eidx, i, j)) continue;
          uint64_t bits;
          memcpy(&bits, &r2, sizeof bits);
          const int table_i = (int)(bits >> 46) + r2_delta_expc;
          const double diffa = r2 - T.r2_table[table_i];
          const double* tf = four + 16 * table_i;
          const int ab = type[i] * ntypes + type[j];
          const double Aij = A[ab], Bij = B[ab];
          const double vdw_d = Aij * tf[0] - Bij * tf[4];
          const double vdw_c = Aij * tf[1] - Bij * tf[5];
          const double vdw_b = Aij * tf[2] - Bij * tf[6];
          const double vdw_a = Aij * tf[3] - Bij * tf[7];
          const double kqq = q[i] * q[j];
          const double fast_d = kqq * tf[8], fast_c = kqq * tf[9], fast_b = kqq * tf[10], fast_a = kqq * tf[11];
          const double slow_d = kqq * tf[12], slow_c = kqq * tf[13], slow_b = kqq * tf[14], slow_a = kqq * tf[15];
          const double vdw_val = ((diffa * vdw_d * (1 / 6.) + vdw_c * (1 / 4.)) * diffa + vdw_b * (1 / 2.)) * diffa + vdw_a;
          const double fast_val = ((diffa * fast_d * (1 / 6.) + fast_c * (1 / 4.)) * diffa + fast_b * (1 / 2.)) * diffa + fast_a;
          const double slow_val = ((diffa * slow_d * (1 / 6.) + slow_c * (1 / 4.)) * diffa + slow_b * (1 / 2.)) * diffa + slow_a;
          S[18] -= vdw_val;
          S[19] -= fast_val;
          S[20] -= slow_val;
          const double force_r = ((diffa * (vdw_d + fast_d) + (vdw_c + fast_c)) * diffa) + (vdw_b + fast_b);
          const double fullforce_r = (diffa * slow_d + slow_c) * diffa + slow_b;
          const double pij[3] = {p_ij_x, p_ij_y, p_ij_z};
          for (int u = 0; u < 3; ++u) {
            F[3 * i + u] += force_r * pij[u];
            F[3 * j + u] -= force_r * pij[u];
            F[3 * n + 3 * i + u] += fullforce_r * pij[u];
            F[3 * n + 3 * j + u] -= fullforce_r * pij[u];
          }
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
              S[a * 3 + b] += force_r * pij[a] * pij[b];
              S[9 + a * 3 + b] += fullforce_r * pij[a] * pij[b];
            }
          np[tid] += 1;
        }
  }
  if (f) memset(f, 0, sizeof(double) * 3 * n);
  if (fs) memset(fs, 0, sizeof(double) * 3 * n);
  double tot[30] = {0};
  int64_t ntot = 0;
  for (int t = 0; t < nt; ++t) {
    const double* F = fb + (size_t)t * 6 * n;
    if (f)
      for (int k = 0; k < 3 * n; ++k) f[k] += F[k];
    if (fs)
      for (int k = 0; k < 3 * n; ++k) fs[k] += F[3 * n + k];
    for (int k = 0; k < 30; ++k) tot[k] += acc[(size_t)t * 30 + k];
    ntot += np[t];
  }
  if (virial) memcpy(virial, tot, 9 * sizeof(double));
  if (virialSlow) memcpy(virialSlow, tot + 9, 9 * sizeof(double));
  if (energies) memcpy(energies, tot + 18, 3 * sizeof(double));
  if (npairs) *npairs = ntot;
  free(fb);
  free(acc);
  free(np);
  free(four);
  cells_free(&C);
  namd_tables_free(&T);
  return 0;
}

/* ---- modified (scaled 1-4) exclusion correction ----------------------------
 * modifiedExclusionForce (src/ComputeBondedCUDAKernel.cu:316-632) with the
 * C1 pair forms of cudaModExclForceMagCalc_VdwEnergySwitch_PMEC1
 * (src/CudaComputeNonbondedInteractions.h:170-254).  The tile kernel treats a
 * modified pair as an ordinary pair; this adds, for each listed pair with
 * r^2 < rc^2 (minimum image):
 *   vdW   the switched LJ of the pair with (A14 - A, B14 - B)
 *         (exclusionVdwCoefTable, src/CudaNonbondedTables.C:119-131);
 *   fast  qq' (1/r - 3/(2rc) + r^2/(2rc^3)),   qq' = -(1 - scale14) qq;
 *   slow  qq' (3/(2rc) - r^2/(2rc^3)).
 * pairs: atom CSR (mptr[n+1], midx), each pair listed once or from both ends
 * (deduplicated here).  forceSwitch selects the force-switched vdW form.
 * energies[3] = vdw, elec, slow; virial / virialSlow as oracle_nb_fp64. */
static int mod_excl_core(int n, const double* pos, const double* q, const int32_t* type, int ntypes,
                         const double* A, const double* B, const double* A14, const double* B14,
                         const int64_t* mptr, const int32_t* midx, const double L[3], double cutoff,
                         double switchdist, double scale14, int forceSwitch, const int8_t* part, const AlchF* al,
                         double* f, double* fs, double* energies, int nenergies, double* virial,
                         double* virialSlow, int64_t* npairs, const double* cell) {
  /* cell (rows a, b, c; NULL: the orthorhombic L): minimum image by
   * rounding fractional coordinates */
  double R[9] = {0};
  if (cell) {
    const double* a = cell;
    const double* b = cell + 3;
    const double* cc = cell + 6;
    const double bxc[3] = {b[1] * cc[2] - b[2] * cc[1], b[2] * cc[0] - b[0] * cc[2], b[0] * cc[1] - b[1] * cc[0]};
    const double cxa[3] = {cc[1] * a[2] - cc[2] * a[1], cc[2] * a[0] - cc[0] * a[2], cc[0] * a[1] - cc[1] * a[0]};
    const double axb[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    const double V = a[0] * bxc[0] + a[1] * bxc[1] + a[2] * bxc[2];
    for (int d = 0; d < 3; ++d) { R[d] = bxc[d] / V; R[3 + d] = cxa[d] / V; R[6 + d] = axb[d] / V; }
  }
  PairConst c;
  memset(&c, 0, sizeof(c));
  if (forceSwitch) set_force_switch(&c, cutoff, switchdist);
  c.rc = cutoff;
  c.rc2 = cutoff * cutoff;
  c.rs2 = switchdist * switchdist;
  c.c1 = (switchdist < cutoff) ? 1.0 / ((c.rc2 - c.rs2) * (c.rc2 - c.rs2) * (c.rc2 - c.rs2)) : 0.0;
  c.rcinv = 1.0 / cutoff;
  c.rc3inv = c.rcinv * c.rcinv * c.rcinv;
  if (f) memset(f, 0, sizeof(double) * 3 * n);
  if (fs) memset(fs, 0, sizeof(double) * 3 * n);
  double e[12] = {0}, vir[9] = {0}, virs[9] = {0};
  int64_t np = 0;
  for (int i = 0; i < n; ++i) {
    for (int64_t k = mptr[i]; k < mptr[i + 1]; ++k) {
      const int j = midx[k];
      if (j == i) continue;
      /* each unordered pair once: skip (i, j) with j < i if (j, i) is listed too */
      if (j < i) {
        int dup = 0;
        for (int64_t m = mptr[j]; m < mptr[j + 1]; ++m)
          if (midx[m] == i) dup = 1;
        if (dup) continue;
      }
      /* drop repeats of the same partner in i's own list */
      int rep = 0;
      for (int64_t m = mptr[i]; m < k; ++m)
        if (midx[m] == j) rep = 1;
      if (rep) continue;
      double d[3];
      if (cell) {
        for (int u = 0; u < 3; ++u) d[u] = pos[3 * j + u] - pos[3 * i + u];
        double sv[3];
        for (int k = 0; k < 3; ++k) sv[k] = nearbyint(R[3 * k] * d[0] + R[3 * k + 1] * d[1] + R[3 * k + 2] * d[2]);
        for (int u = 0; u < 3; ++u) d[u] -= sv[0] * cell[u] + sv[1] * cell[3 + u] + sv[2] * cell[6 + u];
      } else {
        for (int u = 0; u < 3; ++u) d[u] = min_image(pos[3 * j + u] - pos[3 * i + u], L[u]);
      }
      const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (!(r2 < c.rc2)) continue;
      const int ab = type[i] * ntypes + type[j];
      const double qq = -(1.0 - scale14) * q[i] * q[j];
      const double dA = A14[ab] - A[ab], dB = B14[ab] - B[ab];
      /* pair class as the bonded kernel's pswitch (ComputeBondedCUDA.C:97-99,
       * :3126-3129): 0 plain, 1 up, 2 down, 99 (cross) nothing */
      int ps = 0;
      if (part) {
        const int pi = part[i], pj = part[j];
        const int ref = pi == 0 && pj == 0, cross = pi != 0 && pj != 0 && pi != pj;
        const int dec = !ref && !cross && pi == pj && al->decouple;
        ps = (ref || dec) ? 0 : cross ? 99 : (pi == 1 || pj == 1) ? 1 : 2;
      }
      if (ps == 99) continue;
      double ff, ffs, ev = 0.0, ee = 0.0, es = 0.0;
      const double r = sqrt(r2), rinv = 1.0 / r;
      /* electrostatic parts: energies and (1/r) dE/dr */
      const double eeP = qq * (rinv - 1.5 * c.rcinv + 0.5 * r2 * c.rc3inv);
      const double feP = qq * (-rinv * rinv * rinv + c.rc3inv);
      const double esP = qq * (1.5 * c.rcinv - 0.5 * r2 * c.rc3inv);
      const double fsP = -qq * c.rc3inv;
      if (ps == 0) {
        /* vdW with the difference coefficients; pair_fp64 with qq = 0 and beta = 0
         * gives only the switched LJ (its slow term vanishes with qq) */
        pair_fp64(&c, r2, 0.0, dA, dB, &ff, &ffs, &ev, &ee, &es);
        ff += feP;
        ffs = fsP;
        ee = eeP;
        es = esP;
        e[3] += ev;
        e[4] += ee;
        e[5] += es;
      } else {
        /* up / down (modifiedExclusionForce's alchemical route,
         * ComputeBondedCUDAKernel.cu:379-475, :575-586): soft-core vdW with
         * potential switching and the difference coefficients, electrostatic
         * correction scaled by the pair's elec lambda */
        const int up = ps == 1;
        const double lv = up ? al->vdwUp : al->vdwDown, le = up ? al->elecUp : al->elecDown;
        const double lv2 = up ? al->vdw2Up : al->vdw2Down, le2 = up ? al->elec2Up : al->elec2Down;
        const double sh = up ? al->shUp : al->shDown, sh2 = up ? al->sh2Up : al->sh2Down;
        double S, S2, U, F, U2, F2;
        vdw_switch64(&c, r2, &S, &S2);
        vdw_fe64(dA, dB, r2 + sh, &U, &F);
        vdw_fe64(dA, dB, r2 + sh2, &U2, &F2);
        ff = -lv * (S * F + S2 * U) + le * feP;
        ffs = le * fsP;
        ev = lv * S * U;
        ee = le * eeP;
        es = le * esP;
        e[3] += lv2 * S * U2;
        e[4] += le2 * eeP;
        e[5] += le2 * esP;
        const int o = up ? 0 : 1;
        e[6 + o] += S * (U + lv * al->coeff * 0.5 * F);
        e[8 + o] += eeP;
        e[10 + o] += esP;
      }
      e[0] += ev;
      e[1] += ee;
      e[2] += es;
      for (int u = 0; u < 3; ++u) {
        if (f) { f[3 * i + u] += ff * d[u]; f[3 * j + u] -= ff * d[u]; }
        if (fs) { fs[3 * i + u] += ffs * d[u]; fs[3 * j + u] -= ffs * d[u]; }
      }
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          vir[a * 3 + b] -= ff * d[a] * d[b];
          virs[a * 3 + b] -= ffs * d[a] * d[b];
        }
      ++np;
    }
  }
  if (energies) memcpy(energies, e, nenergies * sizeof(double));
  if (virial) memcpy(virial, vir, sizeof(vir));
  if (virialSlow) memcpy(virialSlow, virs, sizeof(virs));
  if (npairs) *npairs = np;
  return 0;
}

int oracle_mod_excl_fp64(int n, const double* pos, const double* q, const int32_t* type, int ntypes,
                         const double* A, const double* B, const double* A14, const double* B14,
                         const int64_t* mptr, const int32_t* midx, const double L[3], double cutoff,
                         double switchdist, double scale14, int forceSwitch, double* f, double* fs,
                         double* energies, double* virial, double* virialSlow, int64_t* npairs) {
  return mod_excl_core(n, pos, q, type, ntypes, A, B, A14, B14, mptr, midx, L, cutoff, switchdist, scale14,
                       forceSwitch, NULL, NULL, f, fs, energies, 3, virial, virialSlow, npairs, NULL);
}

/* The same in a general cell (rows a, b, c). */
int oracle_mod_excl_cell_fp64(int n, const double* pos, const double* q, const int32_t* type, int ntypes,
                              const double* A, const double* B, const double* A14, const double* B14,
                              const int64_t* mptr, const int32_t* midx, const double cell[9], double cutoff,
                              double switchdist, double scale14, int forceSwitch, double* f, double* fs,
                              double* energies, double* virial, double* virialSlow, int64_t* npairs) {
  const double L[3] = {0, 0, 0};
  return mod_excl_core(n, pos, q, type, ntypes, A, B, A14, B14, mptr, midx, L, cutoff, switchdist, scale14,
                       forceSwitch, NULL, NULL, f, fs, energies, 3, virial, virialSlow, npairs, cell);
}

/* With alchemical partitions (lam[13] and energies[12] as oracle_nb_fp64_alch;
 * cross-partition pairs are skipped and not counted). */
int oracle_mod_excl_alch_fp64(int n, const double* pos, const double* q, const int32_t* type, int ntypes,
                              const double* A, const double* B, const double* A14, const double* B14,
                              const int64_t* mptr, const int32_t* midx, const double L[3], double cutoff,
                              double switchdist, double scale14, const int8_t* part, const double* lam,
                              int decouple, double* f, double* fs, double* energies, double* virial,
                              double* virialSlow, int64_t* npairs) {
  AlchF al;
  al.elecUp = lam[0]; al.elecDown = lam[1]; al.vdwUp = lam[2]; al.vdwDown = lam[3];
  al.elec2Up = lam[4]; al.elec2Down = lam[5]; al.vdw2Up = lam[6]; al.vdw2Down = lam[7];
  al.shUp = lam[8]; al.shDown = lam[9]; al.sh2Up = lam[10]; al.sh2Down = lam[11];
  al.decouple = decouple;
  al.coeff = lam[12];
  return mod_excl_core(n, pos, q, type, ntypes, A, B, A14, B14, mptr, midx, L, cutoff, switchdist, scale14, 0,
                       part, &al, f, fs, energies, 12, virial, virialSlow, npairs, NULL);
}

/* ---- general (triclinic) cell ----------------------------------------------
 * oracle_nb_fp64 in a general periodic cell (lattice rows a, b, c, NAMD's
 * Lattice), by brute force: every pair i < j, each of the 27 images around
 * the fractional minimum image that lies within the cutoff (inclusive), the
 * pair physics of pair_fp64.  Shares no cell-list or wrapping code with the
 * layout it checks. */
int oracle_nb_fp64_cell(int n, const double* pos, const double* q, const int32_t* type, int ntypes,
                        const double* A, const double* B, const int64_t* eptr, const int32_t* eidx,
                        const double cell[9], double cutoff, double switchdist, double beta, int forceSwitch,
                        double* f, double* fs, double* energies, double* virial, double* virialSlow,
                        int64_t* npairs, int nthreads) {
  PairConst c;
  memset(&c, 0, sizeof(c));
  if (forceSwitch) set_force_switch(&c, cutoff, switchdist);
  c.rc = cutoff;
  c.rc2 = cutoff * cutoff;
  c.rs2 = switchdist * switchdist;
  c.c1 = (switchdist < cutoff) ? 1.0 / ((c.rc2 - c.rs2) * (c.rc2 - c.rs2) * (c.rc2 - c.rs2)) : 0.0;
  c.rcinv = 1.0 / cutoff;
  c.rc3inv = c.rcinv * c.rcinv * c.rcinv;
  c.beta = beta;
  const double* a = cell;
  const double* b = cell + 3;
  const double* cc = cell + 6;
  const double bxc[3] = {b[1] * cc[2] - b[2] * cc[1], b[2] * cc[0] - b[0] * cc[2], b[0] * cc[1] - b[1] * cc[0]};
  const double cxa[3] = {cc[1] * a[2] - cc[2] * a[1], cc[2] * a[0] - cc[0] * a[2], cc[0] * a[1] - cc[1] * a[0]};
  const double axb[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
  const double V = a[0] * bxc[0] + a[1] * bxc[1] + a[2] * bxc[2];
  double R[9];
  for (int d = 0; d < 3; ++d) { R[d] = bxc[d] / V; R[3 + d] = cxa[d] / V; R[6 + d] = axb[d] / V; }
  int nt = 1;
#ifdef _OPENMP
  nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
  double* fb = (double*)calloc((size_t)nt * 6 * n, sizeof(double));
  double* acc = (double*)calloc((size_t)nt * 21, sizeof(double));
  int64_t* np = (int64_t*)calloc((size_t)nt, sizeof(int64_t));
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 8) num_threads(nt)
#endif
  for (int i = 0; i < n; ++i) {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double* F = fb + (size_t)tid * 6 * n;
    double* S = acc + (size_t)tid * 21;
    for (int j = i + 1; j < n; ++j) {
      double d0[3];
      for (int u = 0; u < 3; ++u) d0[u] = pos[3 * j + u] - pos[3 * i + u];
      double s[3];
      for (int k = 0; k < 3; ++k) s[k] = nearbyint(R[3 * k] * d0[0] + R[3 * k + 1] * d0[1] + R[3 * k + 2] * d0[2]);
      for (int u = 0; u < 3; ++u) d0[u] -= s[0] * a[u] + s[1] * b[u] + s[2] * cc[u];
      int excl = -1;
      for (int ia = -1; ia <= 1; ++ia)
        for (int ib = -1; ib <= 1; ++ib)
          for (int ic = -1; ic <= 1; ++ic) {
            double d[3];
            for (int u = 0; u < 3; ++u) d[u] = d0[u] + ia * a[u] + ib * b[u] + ic * cc[u];
            const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (!(r2 <= c.rc2)) continue;
            if (excl < 0) excl = is_excluded(eptr, eidx, i, j);
            if (excl) continue;
            double ff, ffs;
            const int ab = type[i] * ntypes + type[j];
            pair_fp64(&c, r2, q[i] * q[j], A[ab], B[ab], &ff, &ffs, &S[18], &S[19], &S[20]);
            for (int u = 0; u < 3; ++u) {
              F[3 * i + u] += ff * d[u];
              F[3 * j + u] -= ff * d[u];
              F[3 * n + 3 * i + u] += ffs * d[u];
              F[3 * n + 3 * j + u] -= ffs * d[u];
            }
            for (int x = 0; x < 3; ++x)
              for (int y = 0; y < 3; ++y) {
                S[x * 3 + y] -= ff * d[x] * d[y];
                S[9 + x * 3 + y] -= ffs * d[x] * d[y];
              }
            np[tid] += 1;
          }
    }
  }
  if (f) memset(f, 0, sizeof(double) * 3 * n);
  if (fs) memset(fs, 0, sizeof(double) * 3 * n);
  double tot[21] = {0};
  int64_t ntot = 0;
  for (int t = 0; t < nt; ++t) {
    const double* F = fb + (size_t)t * 6 * n;
    if (f)
      for (int k = 0; k < 3 * n; ++k) f[k] += F[k];
    if (fs)
      for (int k = 0; k < 3 * n; ++k) fs[k] += F[3 * n + k];
    for (int k = 0; k < 21; ++k) tot[k] += acc[(size_t)t * 21 + k];
    ntot += np[t];
  }
  if (virial) memcpy(virial, tot, 9 * sizeof(double));
  if (virialSlow) memcpy(virialSlow, tot + 9, 9 * sizeof(double));
  if (energies) memcpy(energies, tot + 18, 3 * sizeof(double));
  if (npairs) *npairs = ntot;
  free(fb);
  free(acc);
  free(np);
  return 0;
}
