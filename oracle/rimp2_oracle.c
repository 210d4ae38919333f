/*
 * rimp2_oracle.c -- CPU FP64 restatement of the reference RI-MP2 algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle the GPU results are
 * compared with: tests/, the smoke check and the CPU-baseline leg of bench.py
 * may call it; the product library (rimp2_gpu_amd) never links, loads or
 * calls it.
 *
 * Reference files restated here (januseriksen/rimp2_gpu):
 *   RIMP2.F90:381-417  RIMP2_calc_toccA   -> oracle_toccA
 *   RIMP2.F90:419-453  RIMP2_calc_tvirtA  -> oracle_tvirtA
 *   RIMP2.F90:456-481  RIMP2_calc_toccB   -> oracle_toccB
 *   RIMP2.F90:483-508  RIMP2_calc_tvirtB  -> oracle_tvirtB
 *   RIMP2.F90:510-533  RIMP2_TransAlpha1  -> oracle_transalpha1
 *   RIMP2.F90:536-560  RIMP2_TransAlpha2  -> oracle_transalpha2
 *   RIMP2.F90:562-587  RIMP2_calc_gocc    -> oracle_gocc
 *   RIMP2.F90:589-614  RIMP2_calc_gvirt   -> oracle_gvirt
 *   RIMP2.F90:616-641  RIMP2_calc_gen4DimFO -> oracle_gen4dimfo
 *   RIMP2.F90:67-170   Major Steps 5-8    -> oracle_steps5to8
 *   RIMP2.F90:172-251  first-order djik / blad -> oracle_first_order
 * plus the canonical RI-MP2 pair energy built from toccA (t = g/Delta,
 * Delta = e_I+e_J-e_B-e_A, RIMP2.F90:399-404) and gocc (g = sum_alpha C C,
 * RIMP2.F90:577-581) with U = identity:
 *     eps_ij = sum_ab t^{ij}_{ab} (2 g^{ij}_{ab} - g^{ij}_{ba}),  E = sum_ij eps_ij
 * and the metric step the reference receives pre-fitted (Calpha,
 * RIMP2.F90:48-49): V = L L^T (lower Cholesky), B = L^{-1} X.
 *
 * Pinning (tests/test_oracle.py): closed-form values of the reference's own
 * constant fixtures (generate.F90:77-128 fed through Steps 5-8), brute-force
 * 4-nested loops written independently in numpy on tiny inputs, the
 * identities B^T B = X^T V^{-1} X, L L^T = V, and LAPACK (scipy) Cholesky /
 * triangular solve, E(sB) = s^4 E(B), permutation invariance, nocc=nvirt=1
 * closed form E = V^2/Delta, and eps_ij <= 0 when every Delta < 0.
 *
 * All arrays are column-major (Fortran order) exactly as in the reference.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IX2(a, b, na) ((int64_t)(a) + (int64_t)(na) * (int64_t)(b))
#define IX3(a, b, c, na, nb) ((int64_t)(a) + (int64_t)(na) * ((int64_t)(b) + (int64_t)(nb) * (int64_t)(c)))
#define IX4(a, b, c, d, na, nb, nc) \
  ((int64_t)(a) + (int64_t)(na) * ((int64_t)(b) + (int64_t)(nb) * ((int64_t)(c) + (int64_t)(nc) * (int64_t)(d))))

/* Neumaier-compensated running sum. */
typedef struct { double s, c; } nsum_t;
static inline void nsum_add(nsum_t* a, double x) {
  double t = a->s + x;
  if (fabs(a->s) >= fabs(x)) a->c += (a->s - t) + x;
  else a->c += (x - t) + a->s;
  a->s = t;
}
static inline double nsum_val(const nsum_t* a) { return a->s + a->c; }

/* RIMP2.F90:381-417.  tocc(nocc,noccEOS,nvirt,nvirt); Calpha(NBA,nvirt,nocc). */
void oracle_toccA(int nvirt, int nocc, int noccEOS, int nba, const double* Calpha, const double* EVocc,
                  const double* EVvirt, double* tocc, const double* UoccEOST) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int B = 0; B < nvirt; B++)
    for (int A = 0; A < nvirt; A++) {
      double* tmp = (double*)malloc(sizeof(double) * (size_t)nocc);
      for (int I = 0; I < nocc; I++) {
        for (int J = 0; J < nocc; J++) {
          double g = 0.0;
          for (int al = 0; al < nba; al++)
            g += Calpha[IX3(al, A, I, nba, nvirt)] * Calpha[IX3(al, B, J, nba, nvirt)];
          double d = EVocc[I] + EVocc[J] - EVvirt[B] - EVvirt[A];
          tmp[J] = g / d;
        }
        for (int Jl = 0; Jl < noccEOS; Jl++) {
          double t = 0.0;
          for (int J = 0; J < nocc; J++) t += tmp[J] * UoccEOST[IX2(J, Jl, nocc)];
          tocc[IX4(I, Jl, A, B, nocc, noccEOS, nvirt)] = t;
        }
      }
      free(tmp);
    }
}

/* RIMP2.F90:419-453.  tvirt(nocc,nocc,nvirtEOS,nvirt). */
void oracle_tvirtA(int nvirt, int nocc, int nvirtEOS, int nba, const double* Calpha, const double* EVocc,
                   const double* EVvirt, double* tvirt, const double* UvirtEOST) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int J = 0; J < nocc; J++)
    for (int I = 0; I < nocc; I++) {
      double* tmp = (double*)malloc(sizeof(double) * (size_t)nvirt);
      for (int B = 0; B < nvirt; B++) {
        for (int A = 0; A < nvirt; A++) {
          double g = 0.0;
          for (int al = 0; al < nba; al++)
            g += Calpha[IX3(al, A, I, nba, nvirt)] * Calpha[IX3(al, B, J, nba, nvirt)];
          double d = EVocc[I] + EVocc[J] - EVvirt[B] - EVvirt[A];
          tmp[A] = g / d;
        }
        for (int Al = 0; Al < nvirtEOS; Al++) {
          double t = 0.0;
          for (int A = 0; A < nvirt; A++) t += tmp[A] * UvirtEOST[IX2(A, Al, nvirt)];
          tvirt[IX4(I, J, Al, B, nocc, nocc, nvirtEOS)] = t;
        }
      }
      free(tmp);
    }
}

/* RIMP2.F90:456-481.  tocc(noccEOS,noccEOS,nvirt,nvirt) -> toccEOS(nvirt,noccEOS,nvirt,noccEOS). */
void oracle_toccB(int nvirt, int noccEOS, const double* tocc, const double* UvirtT, double* toccEOS) {
#pragma omp parallel for collapse(3) schedule(static)
  for (int Bl = 0; Bl < nvirt; Bl++)
    for (int Jl = 0; Jl < noccEOS; Jl++)
      for (int Al = 0; Al < nvirt; Al++)
        for (int Il = 0; Il < noccEOS; Il++) {
          double t = 0.0;
          for (int A = 0; A < nvirt; A++)
            t += tocc[IX4(Il, Jl, A, Bl, noccEOS, noccEOS, nvirt)] * UvirtT[IX2(A, Al, nvirt)];
          toccEOS[IX4(Al, Il, Bl, Jl, nvirt, noccEOS, nvirt)] = t;
        }
}

/* RIMP2.F90:483-508.  tvirt(nocc,nocc,nvirtEOS,nvirtEOS) -> tvirtEOS(nvirtEOS,nocc,nvirtEOS,nocc). */
void oracle_tvirtB(int nvirtEOS, int nocc, const double* tvirt, const double* UoccT, double* tvirtEOS) {
#pragma omp parallel for collapse(3) schedule(static)
  for (int Jl = 0; Jl < nocc; Jl++)
    for (int Bl = 0; Bl < nvirtEOS; Bl++)
      for (int Il = 0; Il < nocc; Il++)
        for (int Al = 0; Al < nvirtEOS; Al++) {
          double t = 0.0;
          for (int J = 0; J < nocc; J++)
            t += tvirt[IX4(Il, J, Al, Bl, nocc, nocc, nvirtEOS)] * UoccT[IX2(J, Jl, nocc)];
          tvirtEOS[IX4(Al, Il, Bl, Jl, nvirtEOS, nocc, nvirtEOS)] = t;
        }
}

/* RIMP2.F90:510-533.  AlphaCD5(nba,nvirt,noccEOS) = sum_B UvirtT(B,Bl) AlphaCD4(nba,B,Jl). */
void oracle_transalpha1(int nvirt, int noccEOS, int nba, const double* UvirtT, const double* in, double* out) {
#pragma omp parallel for collapse(3) schedule(static)
  for (int Jl = 0; Jl < noccEOS; Jl++)
    for (int Bl = 0; Bl < nvirt; Bl++)
      for (int al = 0; al < nba; al++) {
        double t = 0.0;
        for (int B = 0; B < nvirt; B++) t += UvirtT[IX2(B, Bl, nvirt)] * in[IX3(al, B, Jl, nba, nvirt)];
        out[IX3(al, Bl, Jl, nba, nvirt)] = t;
      }
}

/* RIMP2.F90:536-560.  AlphaCD5(NBA,n3,n2) = sum_{B<n1} U(B,Bl) AlphaCD4(NBA,B,J). */
void oracle_transalpha2(int n2, int n1, int n3, int nba, const double* U, const double* in, double* out) {
#pragma omp parallel for collapse(3) schedule(static)
  for (int J = 0; J < n2; J++)
    for (int Bl = 0; Bl < n3; Bl++)
      for (int al = 0; al < nba; al++) {
        double t = 0.0;
        for (int B = 0; B < n1; B++) t += U[IX2(B, Bl, n1)] * in[IX3(al, B, J, nba, n1)];
        out[IX3(al, Bl, J, nba, n3)] = t;
      }
}

/* RIMP2.F90:562-587 (gocc) and 589-614 (gvirt) share one body:
 * g(a,i,b,j) = sum_alpha C3(alpha,a,i) C3(alpha,b,j), g(nv,no,nv,no). */
static void g4(int nv, int no, int nba, const double* C3, double* g) {
#pragma omp parallel for collapse(3) schedule(static)
  for (int j = 0; j < no; j++)
    for (int b = 0; b < nv; b++)
      for (int i = 0; i < no; i++)
        for (int a = 0; a < nv; a++) {
          double t = 0.0;
          for (int al = 0; al < nba; al++) t += C3[IX3(al, a, i, nba, nv)] * C3[IX3(al, b, j, nba, nv)];
          g[IX4(a, i, b, j, nv, no, nv)] = t;
        }
}
void oracle_gocc(int nvirt, int noccEOS, int nba, const double* C3, double* goccEOS) {
  g4(nvirt, noccEOS, nba, C3, goccEOS);
}
void oracle_gvirt(int nvirtEOS, int nocc, int nba, const double* C3, double* gvirtEOS) {
  g4(nvirtEOS, nocc, nba, C3, gvirtEOS);
}

/* RIMP2.F90:616-641.  djik(n1,n2,n3,n4) = sum_alpha C3(alpha,A,B) C4(alpha,C,D). */
void oracle_gen4dimfo(int nba, const double* C3, int n1, int n2, const double* C4, int n3, int n4, double* out) {
#pragma omp parallel for collapse(3) schedule(static)
  for (int d = 0; d < n4; d++)
    for (int c = 0; c < n3; c++)
      for (int b = 0; b < n2; b++)
        for (int a = 0; a < n1; a++) {
          double t = 0.0;
          for (int al = 0; al < nba; al++) t += C3[IX3(al, a, b, nba, n1)] * C4[IX3(al, c, d, nba, n3)];
          out[IX4(a, b, c, d, n1, n2, n3)] = t;
        }
}

/* BLAS dgemm semantics (column-major), as called at RIMP2.F90:79,87,110,118,137,160. */
void oracle_dgemm(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
                  const double* B, int64_t ldb, double beta, double* C, int64_t ldc) {
  int at = (ta == 'T' || ta == 't'), bt = (tb == 'T' || tb == 't');
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < n; j++)
    for (int64_t i = 0; i < m; i++) {
      double t = 0.0;
      for (int64_t p = 0; p < k; p++) {
        double a = at ? A[p + lda * i] : A[i + lda * p];
        double b = bt ? B[j + ldb * p] : B[p + ldb * j];
        t += a * b;
      }
      C[i + ldc * j] = alpha * t + (beta == 0.0 ? 0.0 : beta * C[i + ldc * j]);
    }
}

/* RIMP2.F90:67-170: Major Steps 5-8 -> toccEOS(nvirt,noccEOS,nvirt,noccEOS),
 * tvirtEOS(nvirtEOS,nocc,nvirtEOS,nocc), goccEOS(nvirt,noccEOS,nvirt,noccEOS),
 * gvirtEOS(nvirtEOS,nocc,nvirtEOS,nocc). */
void oracle_steps5to8(int nvirt, int nocc, int noccEOS, int nvirtEOS, int nba, const double* Calpha,
                      const double* EVocc, const double* EVvirt, const double* UoccEOST, const double* UvirtT,
                      const double* UoccT, const double* UvirtEOST, double* toccEOS, double* tvirtEOS,
                      double* goccEOS, double* gvirtEOS) {
  size_t nv = nvirt, no = nocc, noE = noccEOS, nvE = nvirtEOS;
  /* Step 5 (RIMP2.F90:71-93) */
  double* tocc = malloc(sizeof(double) * no * noE * nv * nv);
  oracle_toccA(nvirt, nocc, noccEOS, nba, Calpha, EVocc, EVvirt, tocc, UoccEOST);
  double* tocc2 = malloc(sizeof(double) * noE * noE * nv * nv);
  oracle_dgemm('T', 'N', noccEOS, (int64_t)noccEOS * nvirt * nvirt, nocc, 1.0, UoccEOST, nocc, tocc, nocc, 0.0,
               tocc2, noccEOS);
  free(tocc);
  double* tocc3 = malloc(sizeof(double) * noE * noE * nv * nv);
  oracle_dgemm('N', 'N', (int64_t)noccEOS * noccEOS * nvirt, nvirt, nvirt, 1.0, tocc2,
               (int64_t)noccEOS * noccEOS * nvirt, UvirtT, nvirt, 0.0, tocc3, (int64_t)noccEOS * noccEOS * nvirt);
  free(tocc2);
  oracle_toccB(nvirt, noccEOS, tocc3, UvirtT, toccEOS);
  free(tocc3);
  /* Step 6 (RIMP2.F90:99-124) */
  double* tvirt = malloc(sizeof(double) * no * no * nvE * nv);
  oracle_tvirtA(nvirt, nocc, nvirtEOS, nba, Calpha, EVocc, EVvirt, tvirt, UvirtEOST);
  double* tvirt2 = malloc(sizeof(double) * no * no * nvE * nvE);
  oracle_dgemm('N', 'N', (int64_t)nocc * nocc * nvirtEOS, nvirtEOS, nvirt, 1.0, tvirt,
               (int64_t)nocc * nocc * nvirtEOS, UvirtEOST, nvirt, 0.0, tvirt2, (int64_t)nocc * nocc * nvirtEOS);
  free(tvirt);
  double* tvirt3 = malloc(sizeof(double) * no * no * nvE * nvE);
  oracle_dgemm('T', 'N', nocc, (int64_t)nocc * nvirtEOS * nvirtEOS, nocc, 1.0, UoccT, nocc, tvirt2, nocc, 0.0, tvirt3,
               nocc);
  free(tvirt2);
  oracle_tvirtB(nvirtEOS, nocc, tvirt3, UoccT, tvirtEOS);
  free(tvirt3);
  /* Step 7 (RIMP2.F90:136-147) */
  double* c2 = malloc(sizeof(double) * (size_t)nba * nv * no);
  oracle_dgemm('N', 'N', (int64_t)nba * nvirt, noccEOS, nocc, 1.0, Calpha, (int64_t)nba * nvirt, UoccEOST, nocc, 0.0,
               c2, (int64_t)nba * nvirt);
  double* c3 = malloc(sizeof(double) * (size_t)nba * nv * no);
  oracle_transalpha1(nvirt, noccEOS, nba, UvirtT, c2, c3);
  oracle_gocc(nvirt, noccEOS, nba, c3, goccEOS);
  /* Step 8 (RIMP2.F90:158-170) */
  oracle_dgemm('N', 'N', (int64_t)nba * nvirt, nocc, nocc, 1.0, Calpha, (int64_t)nba * nvirt, UoccT, nocc, 0.0, c2,
               (int64_t)nba * nvirt);
  oracle_transalpha2(nocc, nvirt, nvirtEOS, nba, UvirtEOST, c2, c3);
  oracle_gvirt(nvirtEOS, nocc, nba, c3, gvirtEOS);
  free(c2);
  free(c3);
}

/* RIMP2.F90:172-251: the first-order property intermediates.
 *   djik(nvirt,noccEOS,noccEOS,nocctot) and blad(nvirtEOS,nocc,nvirtEOS,nvirt)
 * from Calpha(NBA,nvirt,nocc), CalphaOcc(NBA,nocc,nocctot) and
 * CalphaVV(NBA,nvirt,nvirt), every call with the reference's own arguments:
 * RIMP2.F90:198 multiplies CalphaOcc by UoccT as an nocctot x nocctot matrix
 * (UoccT is nocc x nocc, so only nocctot = nocc is in bounds; the caller
 * guarantees it), and RIMP2.F90:231-235 multiplies CalphaVV by UvirtT with
 * K = ldb = nocc, i.e. UvirtT's storage read as an nocc x nvirt matrix and
 * only the first nocc virtual columns of CalphaVV summed. */
void oracle_first_order(int nvirt, int nocc, int noccEOS, int nvirtEOS, int nba, int nocctot, const double* Calpha,
                        const double* CalphaOcc, const double* CalphaVV, const double* UoccEOST, const double* UvirtT,
                        const double* UoccT, const double* UvirtEOST, double* djik, double* blad) {
  size_t nv = nvirt, no = nocc, noE = noccEOS, nvE = nvirtEOS, nt = nocctot, q = nba;
  /* djik: RIMP2.F90:178-210 */
  double* c2 = malloc(sizeof(double) * q * nv * noE);
  oracle_dgemm('N', 'N', (int64_t)nba * nvirt, noccEOS, nocc, 1.0, Calpha, (int64_t)nba * nvirt, UoccEOST, nocc, 0.0,
               c2, (int64_t)nba * nvirt);
  double* c3 = malloc(sizeof(double) * q * nv * noE);
  oracle_transalpha2(noccEOS, nvirt, nvirt, nba, UvirtT, c2, c3);
  free(c2);
  double* o2 = malloc(sizeof(double) * q * no * nt);
  oracle_dgemm('N', 'N', (int64_t)nba * nocc, nocctot, nocctot, 1.0, CalphaOcc, (int64_t)nba * nocc, UoccT, nocctot,
               0.0, o2, (int64_t)nba * nocc);
  double* c4 = malloc(sizeof(double) * q * noE * nt);
  oracle_transalpha2(nocctot, nocc, noccEOS, nba, UoccEOST, o2, c4);
  free(o2);
  oracle_gen4dimfo(nba, c3, nvirt, noccEOS, c4, noccEOS, nocctot, djik);
  free(c3);
  free(c4);
  /* blad: RIMP2.F90:216-250 */
  double* b2 = malloc(sizeof(double) * q * nv * no);
  oracle_dgemm('N', 'N', (int64_t)nba * nvirt, nocc, nocc, 1.0, Calpha, (int64_t)nba * nvirt, UoccT, nocc, 0.0, b2,
               (int64_t)nba * nvirt);
  double* b3 = malloc(sizeof(double) * q * nvE * no);
  oracle_transalpha2(nocc, nvirt, nvirtEOS, nba, UvirtEOST, b2, b3);
  free(b2);
  double* v2 = malloc(sizeof(double) * q * nv * nv);
  oracle_dgemm('N', 'N', (int64_t)nba * nvirt, nvirt, nocc, 1.0, CalphaVV, (int64_t)nba * nvirt, UvirtT, nocc, 0.0,
               v2, (int64_t)nba * nvirt);
  double* b4 = malloc(sizeof(double) * q * nvE * nv);
  oracle_transalpha2(nvirt, nvirt, nvirtEOS, nba, UvirtEOST, v2, b4);
  free(v2);
  oracle_gen4dimfo(nba, b3, nvirtEOS, nocc, b4, nvirtEOS, nvirt, blad);
  free(b3);
  free(b4);
}

/* ------------------------------------------------------------------------- */
/* Canonical pair energies from the toccA + gocc semantics with U = I.        */
/* ------------------------------------------------------------------------- */

typedef double v4d __attribute__((vector_size(32)));

/* V(a,b) = sum_alpha Bi(alpha,a) Bj(alpha,b) for one pair: the ALPHAAUX loop of
 * RIMP2.F90:400-402 (and 578-580), evaluated in 4x4 (a,b) register blocks with
 * four alpha lanes per accumulator (fixed summation order). */
static void pair_V(int64_t naux, int64_t nvirt, const double* Bi, const double* Bj, double* V) {
  int64_t a4 = nvirt & ~3LL, q4 = naux & ~3LL;
  for (int64_t a0 = 0; a0 < nvirt; a0 += 4)
    for (int64_t b0 = 0; b0 < nvirt; b0 += 4) {
      int full = (a0 < a4) && (b0 < a4);
      if (full) {
        v4d acc[4][4];
        memset(acc, 0, sizeof(acc));
        for (int64_t q = 0; q < q4; q += 4) {
          v4d x[4], y[4];
          for (int u = 0; u < 4; u++) {
            memcpy(&x[u], Bi + (a0 + u) * naux + q, 32);
            memcpy(&y[u], Bj + (b0 + u) * naux + q, 32);
          }
          for (int u = 0; u < 4; u++)
            for (int w = 0; w < 4; w++) acc[u][w] += x[u] * y[w];
        }
        for (int u = 0; u < 4; u++)
          for (int w = 0; w < 4; w++) {
            double s = (acc[u][w][0] + acc[u][w][1]) + (acc[u][w][2] + acc[u][w][3]);
            for (int64_t q = q4; q < naux; q++) s += Bi[(a0 + u) * naux + q] * Bj[(b0 + w) * naux + q];
            V[(a0 + u) + nvirt * (b0 + w)] = s;
          }
      } else {
        for (int64_t a = a0; a < a0 + 4 && a < nvirt; a++)
          for (int64_t b = b0; b < b0 + 4 && b < nvirt; b++) {
            double s = 0.0;
            for (int64_t q = 0; q < naux; q++) s += Bi[a * naux + q] * Bj[b * naux + q];
            V[a + nvirt * b] = s;
          }
      }
    }
}

/* eps_ij = sum_ab t_ab (2 g_ab - g_ba) with t_ab = g_ab / (e_i+e_j-e_b-e_a)
 * (RIMP2.F90:403-404 sign convention), Neumaier-compensated. */
static double pair_energy(int64_t nvirt, const double* V, double ei, double ej, const double* ev) {
  nsum_t acc = {0.0, 0.0};
  for (int64_t b = 0; b < nvirt; b++)
    for (int64_t a = 0; a < nvirt; a++) {
      double g = V[a + nvirt * b], gx = V[b + nvirt * a];
      double d = ei + ej - ev[b] - ev[a];
      double t = g / d;
      nsum_add(&acc, t * (2.0 * g - gx));
    }
  return nsum_val(&acc);
}

/* B(naux,nvirt,nocc) -> eps_pair(nocc,nocc) (symmetric) and E = sum_ij eps_ij. */
void oracle_pair_energies(int64_t nocc, int64_t nvirt, int64_t naux, const double* B, const double* eps_o,
                          const double* eps_v, double* eps_pair, double* e_tot) {
  int64_t npair = nocc * (nocc + 1) / 2;
  double* pe = malloc(sizeof(double) * (size_t)npair);
#pragma omp parallel
  {
    double* V = malloc(sizeof(double) * (size_t)(nvirt * nvirt));
#pragma omp for schedule(dynamic, 1)
    for (int64_t p = 0; p < npair; p++) {
      int64_t j = (int64_t)((sqrt(8.0 * (double)p + 1.0) - 1.0) / 2.0);
      while (j * (j + 1) / 2 > p) j--;
      while ((j + 1) * (j + 2) / 2 <= p) j++;
      int64_t i = p - j * (j + 1) / 2; /* i <= j */
      pair_V(naux, nvirt, B + naux * nvirt * i, B + naux * nvirt * j, V);
      pe[p] = pair_energy(nvirt, V, eps_o[i], eps_o[j], eps_v);
    }
    free(V);
  }
  nsum_t tot = {0.0, 0.0};
  for (int64_t j = 0; j < nocc; j++)
    for (int64_t i = 0; i <= j; i++) {
      double e = pe[j * (j + 1) / 2 + i];
      if (eps_pair) {
        eps_pair[i + nocc * j] = e;
        eps_pair[j + nocc * i] = e;
      }
      nsum_add(&tot, e);
      if (i != j) nsum_add(&tot, e);
    }
  if (e_tot) *e_tot = nsum_val(&tot);
  free(pe);
}

/* Pair energies for an explicit list of pairs (bounded CPU-baseline sample). */
void oracle_pair_energies_subset(int64_t nvirt, int64_t naux, const double* B, const double* eps_o,
                                 const double* eps_v, int64_t npairs, const int64_t* pi, const int64_t* pj,
                                 double* out) {
#pragma omp parallel
  {
    double* V = malloc(sizeof(double) * (size_t)(nvirt * nvirt));
#pragma omp for schedule(dynamic, 1)
    for (int64_t p = 0; p < npairs; p++) {
      pair_V(naux, nvirt, B + naux * nvirt * pi[p], B + naux * nvirt * pj[p], V);
      out[p] = pair_energy(nvirt, V, eps_o[pi[p]], eps_o[pj[p]], eps_v);
    }
    free(V);
  }
}

/* Full V^{ij}_{ab} matrix for one pair (used by the feature / surrogate tests). */
void oracle_pair_V(int64_t nvirt, int64_t naux, const double* B, int64_t i, int64_t j, double* V) {
  pair_V(naux, nvirt, B + naux * nvirt * i, B + naux * nvirt * j, V);
}

/* ------------------------------------------------------------------------- */
/* Metric: V = L L^T (lower Cholesky, column-major) and B = L^{-1} X.        */
/* ------------------------------------------------------------------------- */

/* dot(x[0..n), y[0..n)) with four interleaved partial sums (fixed order). */
static double dot4(const double* x, const double* y, int64_t n) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int64_t k = 0;
  for (; k + 4 <= n; k += 4) {
    s0 += x[k] * y[k];
    s1 += x[k + 1] * y[k + 1];
    s2 += x[k + 2] * y[k + 2];
    s3 += x[k + 3] * y[k + 3];
  }
  for (; k < n; k++) s0 += x[k] * y[k];
  return (s0 + s1) + (s2 + s3);
}

/* Cholesky V = L L^T by the Cholesky-Crout (left-looking) recurrence
 *   L_jj = sqrt(V_jj - sum_{k<j} L_jk^2),
 *   L_ij = (V_ij - sum_{k<j} L_ik L_jk) / L_jj   (i > j),
 * evaluated on a row-major copy (each sum is a contiguous dot product of two
 * rows) in column blocks of 64: the block's diagonal rows sequentially, then
 * every row below it in parallel, each row kept in cache across the block's
 * 64 columns.  L is returned column-major with zeros above the diagonal;
 * returns 0 or (k+1) for a non-positive pivot at k. */
int oracle_cholesky(int64_t n, const double* V, double* L) {
  const int64_t NB = 64;
  double* R = malloc(sizeof(double) * (size_t)(n * n));
  int bad = 0;
  for (int64_t i = 0; i < n; i++)
    for (int64_t j = 0; j <= i; j++) R[i * n + j] = V[i + n * j];
  for (int64_t j0 = 0; j0 < n && !bad; j0 += NB) {
    const int64_t j1 = (j0 + NB < n) ? j0 + NB : n;
    for (int64_t j = j0; j < j1 && !bad; j++) {
      const double d = R[j * n + j] - dot4(R + j * n, R + j * n, j);
      if (!(d > 0.0)) {
        bad = (int)(j + 1);
        break;
      }
      R[j * n + j] = sqrt(d);
      for (int64_t i = j + 1; i < j1; i++) R[i * n + j] = (R[i * n + j] - dot4(R + i * n, R + j * n, j)) / R[j * n + j];
    }
    if (bad) break;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = j1; i < n; i++)
      for (int64_t j = j0; j < j1; j++) R[i * n + j] = (R[i * n + j] - dot4(R + i * n, R + j * n, j)) / R[j * n + j];
  }
  for (int64_t j = 0; j < n; j++)
    for (int64_t i = 0; i < n; i++) L[i + n * j] = (i >= j && !bad) ? R[i * n + j] : 0.0;
  free(R);
  return bad;
}

/* B(:,c) = L^{-1} X(:,c) for each of ncols columns (forward substitution,
 * column-oriented: L is read down its columns, 16 right-hand sides at once). */
void oracle_trsm_lower(int64_t n, int64_t ncols, const double* L, const double* X, double* B) {
  const int64_t NC = 16;
#pragma omp parallel
  {
    double* w = (double*)malloc(sizeof(double) * (size_t)(n * NC));
#pragma omp for schedule(dynamic, 1)
    for (int64_t c0 = 0; c0 < ncols; c0 += NC) {
      const int64_t nc = (ncols - c0 < NC) ? ncols - c0 : NC;
      for (int64_t i = 0; i < n; i++)
        for (int64_t c = 0; c < NC; c++) w[i * NC + c] = (c < nc) ? X[i + n * (c0 + c)] : 0.0;
      for (int64_t k = 0; k < n; k++) {
        const double* lk = L + n * k;
        const double d = lk[k];
        double* wk = w + k * NC;
        for (int64_t c = 0; c < NC; c++) wk[c] /= d;
        for (int64_t i = k + 1; i < n; i++) {
          const double l = lk[i];
          double* wi = w + i * NC;
          for (int64_t c = 0; c < NC; c++) wi[c] -= l * wk[c];
        }
      }
      for (int64_t c = 0; c < nc; c++)
        for (int64_t i = 0; i < n; i++) B[i + n * (c0 + c)] = w[i * NC + c];
    }
    free(w);
  }
}

/* V(naux,naux), X(naux,nvirt*nocc) -> B = L^{-1} X; returns the Cholesky status. */
int oracle_metric_transform(int64_t naux, int64_t ncols, const double* V, const double* X, double* B) {
  double* L = malloc(sizeof(double) * (size_t)(naux * naux));
  int st = oracle_cholesky(naux, V, L);
  if (st == 0) oracle_trsm_lower(naux, ncols, L, X, B);
  free(L);
  return st;
}

int oracle_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
