/*
 * qs3_oracle.c -- CPU restatement of QS3's symmetry-adapted Lanczos hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (ed_amd/, include/) links,
 * loads or calls this file.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker / timed CPU
 * baseline.
 *
 * Every routine restates one routine of the reference algorithm literally (same
 * loop orders, same comparisons, same early exits), with 64-bit integers
 * throughout so that L=1000 sectors do not overflow.  Citations are to the
 * QS-Cube/ED sources (paths relative to the reference root).
 *
 * Conventions: sites are 1..NOS; a configuration is the ascending array of its
 * NOD down-spin sites; ranks are 1-based colex ranks; list indices returned by
 * findstate are 1-based (or -1), exactly as in the reference.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

typedef int64_t i64;

/* ------------------------------------------------------------------------ */
/* Model: lattice, bonds, shift tables, binomials                           */
/* ------------------------------------------------------------------------ */
typedef struct {
  i64 NOS, LX, LY, LZ, NOxxz, ncol;
  i64 *p1, *p2;        /* bond endpoints, swapped so p1 <= p2 (input_param.f90:92-103) */
  double *sJ2, *sJ3;   /* Jxy/2 and Jz/4 (input_param.f90:96-97) */
  double *Jz;          /* raw Jz, for the ferromagnetic energy (eigen_solver.f90:341-347) */
  i64 *sx, *sy, *sz;   /* 1-based one-step shift tables, index 1..NOS (state_lists.f90:36-64) */
  i64 *comb;           /* comb[(j-1)*NOS + (i-1)] = C(i,j), i=1..NOS, j=1..ncol (state_lists.f90:6-15) */
} orc_model;

/* f_combination (state_lists.f90:17-33): multiply then divide, C(n<k) = 0. */
i64 orc_f_combination(i64 n, i64 k) {
  if (n < k) return 0;
  i64 c = 1;
  for (i64 i = 1; i <= k; ++i) { c = c * (n - k + i); c = c / i; }
  return c;
}

static inline i64 C_(const orc_model *m, i64 i, i64 j) {
  /* combination(i,j) with 1 <= i <= NOS, 1 <= j <= ncol */
  return m->comb[(j - 1) * m->NOS + (i - 1)];
}

orc_model *orc_model_new(i64 NOS, i64 LX, i64 LY, i64 LZ, i64 NOxxz,
                         const i64 *pi, const i64 *pj, const double *Jxy,
                         const double *Jz, i64 ncol) {
  orc_model *m = (orc_model *)calloc(1, sizeof(orc_model));
  m->NOS = NOS; m->LX = LX; m->LY = LY; m->LZ = LZ; m->NOxxz = NOxxz;
  m->ncol = ncol < 1 ? 1 : ncol;
  m->p1 = malloc(sizeof(i64) * (NOxxz + 1)); m->p2 = malloc(sizeof(i64) * (NOxxz + 1));
  m->sJ2 = malloc(sizeof(double) * (NOxxz + 1)); m->sJ3 = malloc(sizeof(double) * (NOxxz + 1));
  m->Jz = malloc(sizeof(double) * (NOxxz + 1));
  for (i64 b = 0; b < NOxxz; ++b) {
    i64 a = pi[b], c = pj[b];
    if (a > c) { i64 t = a; a = c; c = t; }
    m->p1[b] = a; m->p2[b] = c;
    m->sJ2[b] = Jxy[b] * 0.5;
    m->sJ3[b] = Jz[b] * 0.25;
    m->Jz[b] = Jz[b];
  }
  m->sx = malloc(sizeof(i64) * (NOS + 1)); m->sy = malloc(sizeof(i64) * (NOS + 1));
  m->sz = malloc(sizeof(i64) * (NOS + 1));
  m->sx[0] = m->sy[0] = m->sz[0] = 0;
  for (i64 z = 1; z <= LZ; ++z)
    for (i64 y = 1; y <= LY; ++y)
      for (i64 x = 1; x <= LX; ++x) {
        i64 s = x + LX * ((y - 1) + LY * (z - 1));
        m->sx[s] = (x % LX) + 1 + LX * (y - 1) + LX * LY * (z - 1);
        m->sy[s] = x + LX * (y % LY) + LX * LY * (z - 1);
        m->sz[s] = x + LX * (y - 1) + LX * LY * (z % LZ);
      }
  m->comb = malloc(sizeof(i64) * NOS * m->ncol);
  for (i64 j = 1; j <= m->ncol; ++j)
    for (i64 i = 1; i <= NOS; ++i) m->comb[(j - 1) * NOS + (i - 1)] = orc_f_combination(i, j);
  return m;
}

void orc_model_free(orc_model *m) {
  if (!m) return;
  free(m->p1); free(m->p2); free(m->sJ2); free(m->sJ3); free(m->Jz);
  free(m->sx); free(m->sy); free(m->sz); free(m->comb); free(m);
}

/* Shift tables out (1-based values, arrays of length NOS). */
void orc_shift_tables(const orc_model *m, i64 *sx, i64 *sy, i64 *sz) {
  for (i64 s = 1; s <= m->NOS; ++s) { sx[s - 1] = m->sx[s]; sy[s - 1] = m->sy[s]; sz[s - 1] = m->sz[s]; }
}

/* ------------------------------------------------------------------------ */
/* Rank / unrank / sort / flip / search                                     */
/* ------------------------------------------------------------------------ */

/* insertion_sort (state_lists.f90:115-129); a is 0-based here. */
static void insertion_sort(i64 *a, i64 nod) {
  for (i64 i = 1; i < nod; ++i) {
    i64 t = a[i], j = i - 1;
    while (j >= 0 && a[j] > t) { a[j + 1] = a[j]; --j; }
    a[j + 1] = t;
  }
}

/* inv_list (state_lists.f90:131-139): rank = n1 + sum_{i>=2} C(n_i - 1, i).
 * C(0, i) (n_i = 1) cannot occur for i >= 2 since n_i >= i. */
i64 orc_inv_list(const orc_model *m, const i64 *n, i64 nod) {
  i64 r = n[0];
  for (i64 i = 2; i <= nod; ++i) r += (n[i - 1] - 1 >= 1) ? C_(m, n[i - 1] - 1, i) : 0;
  return r;
}

/* binary_search over a binomial column (state_lists.f90:310-331): returns the
 * matching index or the insertion point bmin, searching indices [ls, le]. */
static i64 binom_search(const orc_model *m, i64 s, i64 col, i64 ls, i64 le) {
  i64 bmin = ls, bmax = le;
  for (;;) {
    i64 b = bmin + (bmax - bmin) / 2;
    i64 v = C_(m, b, col);
    if (s < v) bmax = b - 1;
    else if (v < s) bmin = b + 1;
    else return b;
    if (bmin > bmax) return bmin;
  }
}

/* list_fly (state_lists.f90:294-308): unrank t into n[0..nod-1]. */
void orc_list_fly(const orc_model *m, i64 t, i64 nod, i64 *n) {
  i64 s = t, j = m->NOS - 1;
  for (i64 i = nod; i >= 2; --i) {
    i64 j0 = binom_search(m, s, i, i, j);
    j = j0 - 1;
    n[i - 1] = j0;
    s -= C_(m, j, i);
  }
  n[0] = s;
}

/* j_flip_ni (state_lists.f90:333-362), 1-based semantics kept via explicit
 * Fortran DO-loop exit values. n, nd are 0-based arrays of length nod. */
void orc_j_flip(i64 i, i64 j, const i64 *n, i64 nod, i64 *nd) {
  i64 kr, kl, jin = 0;
  for (kr = nod; kr >= 1; --kr) {
    if (j < n[kr - 1]) continue;
    else if (j > n[kr - 1]) break;
    else { jin = 1; break; }
  }
  /* loop completed without exit leaves kr = 0 */
  if (!jin) { /* S+_i S-_j : the down spin moves from i to j */
    for (kl = 1; kl <= kr; ++kl) if (i == n[kl - 1]) break;
    for (i64 q = kl; q <= kr - 1; ++q) nd[q - 1] = n[q];
    nd[kr - 1] = j;
  } else {    /* S-_i S+_j : the down spin moves from j to i */
    for (kl = 1; kl <= kr; ++kl) if (i < n[kl - 1]) break;
    nd[kl - 1] = i;
    for (i64 q = kl + 1; q <= kr; ++q) nd[q - 1] = n[q - 2];
  }
  for (i64 q = 1; q <= kl - 1; ++q) nd[q - 1] = n[q - 1];
  for (i64 q = kr + 1; q <= nod; ++q) nd[q - 1] = n[q - 1];
}

/* findstate (state_lists.f90:364-382): 1-based index of s in ascending
 * list_s[0..m-1], or -1. */
i64 orc_findstate(i64 s, const i64 *list_s, i64 mlen) {
  i64 bmin = 1, bmax = mlen;
  if (mlen < 1) return -1;
  for (;;) {
    i64 b = bmin + (bmax - bmin) / 2;
    if (s < list_s[b - 1]) bmax = b - 1;
    else if (s > list_s[b - 1]) bmin = b + 1;
    else return b;
    if (bmin > bmax) return -1;
  }
}

/* representative_SQ (state_lists.f90:384-416): n is replaced by the colex-
 * minimal translation (running minimum, strict <, in loop order k,j,i);
 * ell = (i,j,k) of the last strict improvement, or (0,0,0). Returns the rank. */
i64 orc_representative(const orc_model *m, i64 nod, i64 *n, i64 *ell) {
  i64 nd[64];
  ell[0] = ell[1] = ell[2] = 0;
  for (i64 q = 0; q < nod; ++q) nd[q] = n[q];
  for (i64 k = 1; k <= m->LZ; ++k) {
    for (i64 q = 0; q < nod; ++q) nd[q] = m->sz[nd[q]];
    for (i64 j = 1; j <= m->LY; ++j) {
      for (i64 q = 0; q < nod; ++q) nd[q] = m->sy[nd[q]];
      for (i64 i = 1; i <= m->LX; ++i) {
        for (i64 q = 0; q < nod; ++q) nd[q] = m->sx[nd[q]];
        insertion_sort(nd, nod);
        for (i64 l = nod; l >= 1; --l) {
          if (nd[l - 1] > n[l - 1]) break;
          else if (nd[l - 1] < n[l - 1]) {
            ell[0] = i; ell[1] = j; ell[2] = k;
            for (i64 q = 0; q < l; ++q) n[q] = nd[q];
            break;
          }
        }
      }
    }
  }
  return orc_inv_list(m, n, nod);
}

/* checkstate_SQ (state_lists.f90:66-112): ro > 0 iff rank s is the colex-
 * minimum of its orbit and compatible with (pkx,pky,pkz). Returns ro and, via
 * count_out, the number of translations fixing the configuration. */
double orc_checkstate(const orc_model *m, i64 s, i64 nod, double pkx, double pky,
                      double pkz, i64 *count_out) {
  i64 n[64], np[64];
  double complex coef = 0.0;
  i64 count = 0;
  orc_list_fly(m, s, nod, n);
  for (i64 q = 0; q < nod; ++q) np[q] = n[q];
  if (count_out) *count_out = 0;
  for (i64 k = 1; k <= m->LZ; ++k) {
    for (i64 q = 0; q < nod; ++q) np[q] = m->sz[np[q]];
    double p1 = pkz * (double)k;
    for (i64 j = 1; j <= m->LY; ++j) {
      for (i64 q = 0; q < nod; ++q) np[q] = m->sy[np[q]];
      double p2 = pky * (double)j + p1;
      for (i64 i = 1; i <= m->LX; ++i) {
        for (i64 q = 0; q < nod; ++q) np[q] = m->sx[np[q]];
        insertion_sort(np, nod);
        i64 l;
        for (l = nod; l >= 1; --l) {
          if (np[l - 1] < n[l - 1]) return 0.0;
          else if (np[l - 1] > n[l - 1]) break;
        }
        if (l == 0) {
          double x = pkx * (double)i + p2;
          coef += cexp(CMPLX(0.0, -x));
          count += 1;
        }
      }
    }
  }
  if (count_out) *count_out = count;
  double a = cabs(coef);
  return (a * a) * (double)m->NOS / (double)count;
}

/* ------------------------------------------------------------------------ */
/* Sector build                                                             */
/* ------------------------------------------------------------------------ */

/* allocate_lists_omp_SQ (state_lists.f90:225-292): keep every rank 1..THS
 * whose ro > 1e-15, ascending by rank; list_r = sqrt(ro).  The reference's
 * count / critical-append / quicksort sequence produces exactly the ascending
 * scan below.  Returns the dimension; fills the arrays if non-NULL (call once
 * with NULL to size).  flags (THS bytes) is scratch supplied by the caller. */
i64 orc_build_sector(const orc_model *m, i64 nod, double pkx, double pky, double pkz,
                     i64 *list_s, double *list_r, unsigned char *flags, double *ro_scratch) {
  i64 THS = (nod > 0) ? C_(m, m->NOS, nod) : 1;
#pragma omp parallel for schedule(dynamic, 4096)
  for (i64 s = 1; s <= THS; ++s) {
    double r = orc_checkstate(m, s, nod, pkx, pky, pkz, NULL);
    flags[s - 1] = (r > 1.0e-15);
    ro_scratch[s - 1] = r;
  }
  i64 a = 0;
  for (i64 s = 1; s <= THS; ++s) {
    if (flags[s - 1]) {
      if (list_s) list_s[a] = s;
      if (list_r) list_r[a] = sqrt(ro_scratch[s - 1]);
      ++a;
    }
  }
  return a;
}

/* Chain-only fast sector build (LY = LZ = 1), used where the literal scan over
 * all C(L,N) ranks is too slow (L = 1000).  It enumerates only configurations
 * containing site 1 -- every colex-minimal chain configuration does -- in
 * ascending rank order and applies the same checkstate test.  Tests pin it
 * against orc_build_sector on small chains. Same calling convention. */
i64 orc_build_sector_chain(const orc_model *m, i64 nod, double pkx, double pky, double pkz,
                           i64 *list_s, double *list_r) {
  if (m->LY != 1 || m->LZ != 1 || nod < 1) return -1;
  const i64 L = m->NOS;
  /* anchored candidates: subsets of size nod-1 of {2..L}; colex order over the
   * subset equals ascending full rank.  Enumerate by the top element c.
   * Count per top element in parallel, then fill. */
  i64 nb = (nod == 1) ? 1 : L;        /* buckets indexed by largest site */
  i64 *cnt = calloc(nb + 1, sizeof(i64));
  if (nod == 1) {
    i64 cc = 0;
    double r = orc_checkstate(m, 1, 1, pkx, pky, pkz, &cc);
    if (r > 1e-15) { if (list_s) list_s[0] = 1; if (list_r) list_r[0] = sqrt(r); free(cnt); return 1; }
    free(cnt); return 0;
  }
  /* ranks with n1 = 1 and top element top: rank = 1 + sum_{i=2}^{nod} C(n_i-1, i) */
#pragma omp parallel for schedule(dynamic, 1)
  for (i64 top = nod; top <= L; ++top) {
    /* iterate all subsets {n_2 < ... < n_{nod-1}} of {2..top-1} in colex order */
    i64 n[64]; i64 k = nod - 2; /* free elements */
    for (i64 q = 0; q < k; ++q) n[q + 1] = 2 + q;
    n[0] = 1; n[nod - 1] = top;
    i64 c = 0;
    for (;;) {
      i64 r_rank = orc_inv_list(m, n, nod);
      double r = orc_checkstate(m, r_rank, nod, pkx, pky, pkz, NULL);
      if (r > 1e-15) ++c;
      /* next colex subset of the free elements n[1..k] within [2, top-1] */
      i64 q = 1;
      while (q <= k && n[q] + 1 == ((q + 1 <= k) ? n[q + 1] : top)) ++q;
      if (q > k) break;
      n[q] += 1;
      for (i64 p = 1; p < q; ++p) n[p] = 2 + (p - 1);
    }
    cnt[top] = c;
  }
  i64 *off = calloc(nb + 2, sizeof(i64));
  for (i64 top = 1; top <= L; ++top) off[top + 1] = off[top] + cnt[top];
  i64 D = off[L + 1];
  if (list_s || list_r) {
#pragma omp parallel for schedule(dynamic, 1)
    for (i64 top = nod; top <= L; ++top) {
      i64 n[64]; i64 k = nod - 2;
      for (i64 q = 0; q < k; ++q) n[q + 1] = 2 + q;
      n[0] = 1; n[nod - 1] = top;
      i64 a = off[top];
      for (;;) {
        i64 r_rank = orc_inv_list(m, n, nod);
        double r = orc_checkstate(m, r_rank, nod, pkx, pky, pkz, NULL);
        if (r > 1e-15) {
          if (list_s) list_s[a] = r_rank;
          if (list_r) list_r[a] = sqrt(r);
          ++a;
        }
        i64 q = 1;
        while (q <= k && n[q] + 1 == ((q + 1 <= k) ? n[q + 1] : top)) ++q;
        if (q > k) break;
        n[q] += 1;
        for (i64 p = 1; p < q; ++p) n[p] = 2 + (p - 1);
      }
    }
  }
  free(cnt); free(off);
  return D;
}

/* ------------------------------------------------------------------------ */
/* H.v and observables                                                      */
/* ------------------------------------------------------------------------ */

static inline int member(const i64 *n, i64 nod, i64 s) {
  for (i64 q = 0; q < nod; ++q) if (n[q] == s) return 1;
  return 0;
}

/* explist(0:LX,0:LY,0:LZ) column-major */
static inline double complex ex_at(const orc_model *m, const double complex *ex, const i64 *ell) {
  return ex[ell[0] + (m->LX + 1) * (ell[1] + (m->LY + 1) * ell[2])];
}

/* ham_to_vec_wave_vector (ham2vec_v3.f90:37-72) for rows [r0, r1) (0-based).
 * Pull form; every bond of the bond file is scanned for every row. */
void orc_ham_to_vec(const orc_model *m, i64 nod, i64 lv, const i64 *list_s,
                    const double *list_r, const double complex *explist,
                    double complex *v0, const double complex *v1, i64 r0, i64 r1) {
#pragma omp parallel for schedule(dynamic, 256)
  for (i64 i = r0; i < r1; ++i) {
    i64 st[64], ni[64], ell[3];
    double complex acc = 0.0;
    orc_list_fly(m, list_s[i], nod, st);
    for (i64 b = 0; b < m->NOxxz; ++b) {
      int f1 = member(st, nod, m->p1[b]);
      int f2 = member(st, nod, m->p2[b]);
      if (f1 != f2) {
        acc = acc - m->sJ3[b] * v1[i];
        orc_j_flip(m->p1[b], m->p2[b], st, nod, ni);
        i64 s = orc_representative(m, nod, ni, ell);
        i64 id = orc_findstate(s, list_s, lv);
        if (id > 0)
          acc = acc + (m->sJ2[b] * list_r[id - 1] / list_r[i]) * ex_at(m, explist, ell) * v1[id - 1];
      } else {
        acc = acc + m->sJ3[b] * v1[i];
      }
    }
    v0[i] = acc;
  }
}

/* U(1)-only H.v (source_only_u1/ham2vec_v3.f90:35-65): the rows are all
 * C(NOS, nod) configurations, row i (1-based) being colex rank i
 * (list_fly(i), source_only_u1/state_lists.f90); a hop's target row is
 * inv_list of the flipped configuration (no representative, no norms, no
 * phases).  Needs ncol >= nod.  Rows [r0, r1). */
void orc_u1_ham_to_vec(const orc_model *m, i64 nod, double complex *v0, const double complex *v1, i64 r0,
                       i64 r1) {
#pragma omp parallel for schedule(dynamic, 256)
  for (i64 i = r0; i < r1; ++i) {
    i64 st[64], ni[64];
    double complex acc = 0.0;
    orc_list_fly(m, i + 1, nod, st);
    for (i64 b = 0; b < m->NOxxz; ++b) {
      int f1 = member(st, nod, m->p1[b]);
      int f2 = member(st, nod, m->p2[b]);
      if (f1 != f2) {
        acc = acc - m->sJ3[b] * v1[i];
        orc_j_flip(m->p1[b], m->p2[b], st, nod, ni);
        i64 id = orc_inv_list(m, ni, nod);
        acc = acc + m->sJ2[b] * v1[id - 1];
      } else {
        acc = acc + m->sJ3[b] * v1[i];
      }
    }
    v0[i] = acc;
  }
}

/* Number of hop terms with a target inside the sector (nnz of the off-
 * diagonal part), for the algorithmic byte count. Rows [r0, r1). */
i64 orc_count_hops(const orc_model *m, i64 nod, i64 lv, const i64 *list_s, i64 r0, i64 r1) {
  i64 tot = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : tot)
  for (i64 i = r0; i < r1; ++i) {
    i64 st[64], ni[64], ell[3];
    orc_list_fly(m, list_s[i], nod, st);
    for (i64 b = 0; b < m->NOxxz; ++b) {
      int f1 = member(st, nod, m->p1[b]);
      int f2 = member(st, nod, m->p2[b]);
      if (f1 != f2) {
        orc_j_flip(m->p1[b], m->p2[b], st, nod, ni);
        i64 s = orc_representative(m, nod, ni, ell);
        if (orc_findstate(s, list_s, lv) > 0) ++tot;
      }
    }
  }
  return tot;
}

/* mk_list_ij (get_expectation_values.f90:97-131): translation orbit of the
 * ordered pair r; noij = 0 unless r is the orbit's first element. list_ij is
 * 2 x NOS (column-major: list_ij[2*q + 0/1]). */
void orc_mk_list_ij(const orc_model *m, const i64 *r, i64 *list_ij, i64 *noij_out) {
  i64 noij = 1, rd[2];
  list_ij[0] = r[0]; list_ij[1] = r[1];
  rd[0] = r[0]; rd[1] = r[1];
  for (i64 k = 1; k <= m->LZ; ++k) {
    rd[0] = m->sz[rd[0]]; rd[1] = m->sz[rd[1]];
    for (i64 j = 1; j <= m->LY; ++j) {
      rd[0] = m->sy[rd[0]]; rd[1] = m->sy[rd[1]];
      for (i64 i = 1; i <= m->LX; ++i) {
        rd[0] = m->sx[rd[0]]; rd[1] = m->sx[rd[1]];
        i64 l;
        for (l = 2; l >= 1; --l) {
          if (rd[l - 1] < r[l - 1]) { *noij_out = 0; return; }
          else if (rd[l - 1] > r[l - 1]) {
            list_ij[2 * noij + 0] = rd[0]; list_ij[2 * noij + 1] = rd[1];
            noij += 1;
            break;
          }
        }
        if (l == 0) { *noij_out = noij; return; }
      }
    }
  }
  *noij_out = noij;
}

/* calcu_cf_trans_2 (ham2vec_v3.f90:111-174) for one pair orbit: returns the
 * orbit's <SzSz> and <S+S-> (before replication into the NOS^2 arrays). */
void orc_calcu_cf_trans_2(const orc_model *m, i64 nod, i64 lv, const i64 *list_s,
                          const double *list_r, const double complex *explist,
                          const double complex *v1, const i64 *list_ij, i64 noij,
                          double *szsz_out, double complex *spsm_out) {
  double szsz = 0.0;
  double sp_re = 0.0, sp_im = 0.0;
  if (list_ij[0] == list_ij[1]) {
    szsz = 0.25;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : sp_re)
    for (i64 i = 0; i < lv; ++i) {
      i64 st[64];
      orc_list_fly(m, list_s[i], nod, st);
      for (i64 j = 0; j < noij; ++j) {
        if (member(st, nod, list_ij[2 * j + 0])) continue;
        sp_re += creal(v1[i] * conj(v1[i]));
      }
    }
  } else {
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : szsz, sp_re, sp_im)
    for (i64 i = 0; i < lv; ++i) {
      i64 st[64], ni[64], ell[3];
      orc_list_fly(m, list_s[i], nod, st);
      for (i64 j = 0; j < noij; ++j) {
        i64 a = list_ij[2 * j + 0], b = list_ij[2 * j + 1];
        int f1 = member(st, nod, a), f2 = member(st, nod, b);
        double w = creal(conj(v1[i]) * v1[i]);
        if (!f1 && f2) {
          szsz -= w;
        } else if (f1 && !f2) {
          szsz -= w;
          i64 lo = a < b ? a : b, hi = a < b ? b : a;
          orc_j_flip(lo, hi, st, nod, ni);
          i64 s = orc_representative(m, nod, ni, ell);
          i64 id = orc_findstate(s, list_s, lv);
          if (id > 0) {
            double complex t = (list_r[id - 1] / list_r[i]) * ex_at(m, explist, ell) * conj(v1[i]) * v1[id - 1];
            sp_re += creal(t); sp_im += cimag(t);
          }
        } else {
          szsz += w;
        }
      }
    }
    szsz = szsz / (double)noij * 0.25;
  }
  *szsz_out = szsz;
  *spsm_out = CMPLX(sp_re, sp_im) / (double)noij;
}

/* calcu_lm_trans_2 (ham2vec_v3.f90:74-109) for one site orbit list_n[0..non-1]. */
double orc_calcu_lm_trans_2(const orc_model *m, i64 nod, i64 lv, const i64 *list_s,
                            const double complex *v1, const i64 *list_n, i64 non) {
  double ev = 0.0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : ev)
  for (i64 i = 0; i < lv; ++i) {
    i64 st[64];
    orc_list_fly(m, list_s[i], nod, st);
    double w = creal(v1[i] * conj(v1[i]));
    for (i64 j = 0; j < non; ++j) ev += member(st, nod, list_n[j]) ? -w : w;
  }
  return ev / (double)non * 0.5;
}

/* mk_list_n (get_expectation_values.f90:65-95). */
void orc_mk_list_n(const orc_model *m, i64 n, i64 *list_n, i64 *non_out) {
  i64 non = 1, nd = n;
  list_n[0] = n;
  for (i64 k = 1; k <= m->LZ; ++k) {
    nd = m->sz[nd];
    for (i64 j = 1; j <= m->LY; ++j) {
      nd = m->sy[nd];
      for (i64 i = 1; i <= m->LX; ++i) {
        nd = m->sx[nd];
        if (nd < n) { *non_out = 0; return; }
        else if (nd > n) { list_n[non] = nd; non += 1; }
        else { *non_out = non; return; }
      }
    }
  }
  *non_out = non;
}

/* calcu_Sq_a (get_expectation_values.f90:281-377): push-form S^a_q |psi>.
 * spsmsz 1: S+ (remove a down spin), 2: S- (add one), 3: Sz.  expks[i-1] =
 * exp(-i q.R_i).  The old sector is (nod, list_s, list_r); the new one is
 * (nod_new, list_s_new, list_r_new, explist_new).  Serial (small sizes). */
void orc_calcu_Sq_a(const orc_model *m, i64 spsmsz, i64 nod, i64 dim, const i64 *list_s,
                    const double *list_r, i64 nod_new, i64 dim_new, const i64 *list_s_new,
                    const double *list_r_new, const double complex *explist_new,
                    const double complex *expks, const double complex *psi, double complex *fn) {
  for (i64 q = 0; q < dim_new; ++q) fn[q] = 0.0;
  for (i64 i = 1; i <= m->NOS; ++i) {
    for (i64 l = 0; l < dim; ++l) {
      i64 st[64], bp[64], ell[3];
      orc_list_fly(m, list_s[l], nod, st);
      int fi = member(st, nod, i);
      double siz = 1.0;
      if (spsmsz == 1) {
        if (!fi) continue;
        i64 a = 0;
        for (i64 q = 0; q < nod; ++q) if (st[q] != i) bp[a++] = st[q];
      } else if (spsmsz == 2) {
        if (fi) continue;
        i64 a = 0, put = 0;
        for (i64 q = 0; q < nod; ++q) {
          if (!put && st[q] > i) { bp[a++] = i; put = 1; }
          bp[a++] = st[q];
        }
        if (!put) bp[a++] = i;
      } else {
        siz = fi ? -0.5 : 0.5;
        for (i64 q = 0; q < nod; ++q) bp[q] = st[q];
      }
      i64 s = orc_representative(m, nod_new, bp, ell);
      i64 lp = orc_findstate(s, list_s_new, dim_new);
      if (lp > 0) {
        double complex t = psi[l] * list_r_new[lp - 1] / list_r[l] * expks[i - 1] * conj(ex_at(m, explist_new, ell));
        fn[lp - 1] += (spsmsz == 3) ? siz * t : t;
      }
    }
  }
  double sc = 1.0 / sqrt((double)m->NOS);
  for (i64 q = 0; q < dim_new; ++q) fn[q] = sc * fn[q];
}

/* Dense H (make_full_hamiltonian, ham2vec_v3.f90:5-35), row-major lv x lv,
 * H[i*lv + id].  Small sectors only (oracle for the solvers). */
void orc_full_hamiltonian(const orc_model *m, i64 nod, i64 lv, const i64 *list_s,
                          const double *list_r, const double complex *explist, double complex *H) {
  memset(H, 0, sizeof(double complex) * lv * lv);
#pragma omp parallel for schedule(dynamic, 16)
  for (i64 i = 0; i < lv; ++i) {
    i64 st[64], ni[64], ell[3];
    orc_list_fly(m, list_s[i], nod, st);
    for (i64 b = 0; b < m->NOxxz; ++b) {
      int f1 = member(st, nod, m->p1[b]);
      int f2 = member(st, nod, m->p2[b]);
      if (f1 != f2) {
        H[i * lv + i] -= m->sJ3[b];
        orc_j_flip(m->p1[b], m->p2[b], st, nod, ni);
        i64 s = orc_representative(m, nod, ni, ell);
        i64 id = orc_findstate(s, list_s, lv);
        if (id > 0) H[i * lv + (id - 1)] += (m->sJ2[b] * list_r[id - 1] / list_r[i]) * ex_at(m, explist, ell);
      } else {
        H[i * lv + i] += m->sJ3[b];
      }
    }
  }
}

int orc_max_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }
