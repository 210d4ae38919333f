/* Runs every entry point of the CPU oracle (oracle/rimp2_oracle.c,
 * oracle/mlp_oracle.c) on small, ragged shapes, for a build under
 * -fsanitize=address,undefined: each buffer is allocated to the exact size
 * the function documents, so a read or write past it, a use after free, a
 * leak or undefined arithmetic ends the run.  Light identities keep the
 * calls honest (L L^T = V, L B = X, the pair subset equals the full pair
 * matrix, E = sum of pair energies, an indefinite metric is refused).
 * Prints "OK <checks>" and exits 0, or names the first failed check and
 * exits 1. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void oracle_steps5to8(int nvirt, int nocc, int noccEOS, int nvirtEOS, int nba, const double* Calpha,
                      const double* EVocc, const double* EVvirt, const double* UoccEOST, const double* UvirtT,
                      const double* UoccT, const double* UvirtEOST, double* toccEOS, double* tvirtEOS,
                      double* goccEOS, double* gvirtEOS);
void oracle_first_order(int nvirt, int nocc, int noccEOS, int nvirtEOS, int nba, int nocctot, const double* Calpha,
                        const double* CalphaOcc, const double* CalphaVV, const double* UoccEOST, const double* UvirtT,
                        const double* UoccT, const double* UvirtEOST, double* djik, double* blad);
void oracle_pair_energies(int64_t nocc, int64_t nvirt, int64_t naux, const double* B, const double* eps_o,
                          const double* eps_v, double* eps_pair, double* e_tot);
void oracle_pair_energies_subset(int64_t nvirt, int64_t naux, const double* B, const double* eps_o,
                                 const double* eps_v, int64_t npairs, const int64_t* pi, const int64_t* pj,
                                 double* out);
void oracle_pair_V(int64_t nvirt, int64_t naux, const double* B, int64_t i, int64_t j, double* V);
int oracle_cholesky(int64_t n, const double* V, double* L);
void oracle_trsm_lower(int64_t n, int64_t ncols, const double* L, const double* X, double* B);
int oracle_metric_transform(int64_t naux, int64_t ncols, const double* V, const double* X, double* B);
int oracle_num_threads(void);
int64_t oracle_mlp_num_params(void);
double oracle_mlp_loss_grad(int64_t n, const float* feats, const float* target, const float* params,
                            const float* norm, int emulate_bf16, double* grad, double* pred);

static int g_checks = 0;
static uint64_t g_state = 0x9E3779B97F4A7C15ull;

static double urand(void) {
  g_state ^= g_state << 13;
  g_state ^= g_state >> 7;
  g_state ^= g_state << 17;
  return (double)(g_state >> 11) * (1.0 / 9007199254740992.0);
}

static double* vec(size_t n, double lo, double hi) {
  double* p = malloc(sizeof(double) * (n ? n : 1));
  for (size_t k = 0; k < n; k++) p[k] = lo + (hi - lo) * urand();
  return p;
}

static void check(int ok, const char* what, long long a, long long b) {
  g_checks++;
  if (!ok) {
    printf("FAIL %s (%lld, %lld)\n", what, a, b);
    exit(1);
  }
}

/* V = A A^T / n + I, column-major */
static double* spd(int64_t n) {
  double* A = vec((size_t)(n * n), -0.5, 0.5);
  double* V = malloc(sizeof(double) * (size_t)(n * n));
  for (int64_t i = 0; i < n; i++)
    for (int64_t j = 0; j < n; j++) {
      double s = 0.0;
      for (int64_t k = 0; k < n; k++) s += A[i + n * k] * A[j + n * k];
      V[i + n * j] = s / (double)n + (i == j ? 1.0 : 0.0);
    }
  free(A);
  return V;
}

static void metric_case(int64_t n, int64_t ncols) {
  double* V = spd(n);
  double* L = malloc(sizeof(double) * (size_t)(n * n));
  check(oracle_cholesky(n, V, L) == 0, "cholesky of an SPD metric", n, 0);
  double err = 0.0;
  for (int64_t i = 0; i < n; i++)
    for (int64_t j = 0; j <= i; j++) {
      double s = 0.0;
      for (int64_t k = 0; k <= j; k++) s += L[i + n * k] * L[j + n * k];
      err = fmax(err, fabs(s - V[i + n * j]));
    }
  check(err < 1e-12, "L L^T = V", n, 0);
  double* X = vec((size_t)(n * ncols), -1.0, 1.0);
  double* B = malloc(sizeof(double) * (size_t)(n * ncols));
  double* B2 = malloc(sizeof(double) * (size_t)(n * ncols));
  oracle_trsm_lower(n, ncols, L, X, B);
  err = 0.0;
  for (int64_t c = 0; c < ncols; c++)
    for (int64_t i = 0; i < n; i++) {
      double s = 0.0;
      for (int64_t k = 0; k <= i; k++) s += L[i + n * k] * B[k + n * c];
      err = fmax(err, fabs(s - X[i + n * c]));
    }
  check(err < 1e-11, "L B = X", n, ncols);
  check(oracle_metric_transform(n, ncols, V, X, B2) == 0, "metric transform", n, ncols);
  check(memcmp(B, B2, sizeof(double) * (size_t)(n * ncols)) == 0, "transform = cholesky + trsm", n, ncols);
  V[n + 1 < n * n ? n + 1 : 0] = -4.0 * n; /* a negative diagonal element: indefinite */
  check(oracle_cholesky(n, V, L) != 0, "indefinite metric refused", n, 0);
  free(V);
  free(L);
  free(X);
  free(B);
  free(B2);
}

static void pair_case(int64_t nocc, int64_t nvirt, int64_t naux) {
  double* B = vec((size_t)(naux * nvirt * nocc), -0.3, 0.3);
  double* eo = vec((size_t)nocc, -2.0, -0.3);
  double* ev = vec((size_t)nvirt, 0.1, 3.0);
  double* ep = malloc(sizeof(double) * (size_t)(nocc * nocc));
  double et = 0.0;
  oracle_pair_energies(nocc, nvirt, naux, B, eo, ev, ep, &et);
  double s = 0.0;
  for (int64_t k = 0; k < nocc * nocc; k++) s += ep[k];
  check(fabs(s - et) <= 1e-12 * fabs(et) + 1e-300, "E = sum of pair energies", nocc, nvirt);
  const int64_t np = nocc * nocc;
  int64_t* pi = malloc(sizeof(int64_t) * (size_t)np);
  int64_t* pj = malloc(sizeof(int64_t) * (size_t)np);
  double* sub = malloc(sizeof(double) * (size_t)np);
  for (int64_t k = 0; k < np; k++) {
    pi[k] = k % nocc;
    pj[k] = k / nocc;
  }
  oracle_pair_energies_subset(nvirt, naux, B, eo, ev, np, pi, pj, sub);
  double err = 0.0;
  for (int64_t k = 0; k < np; k++) err = fmax(err, fabs(sub[k] - ep[pi[k] + nocc * pj[k]]));
  check(err <= 1e-12 * fabs(et), "pair subset = pair matrix", nocc, nvirt);
  double* V1 = malloc(sizeof(double) * (size_t)(nvirt * nvirt));
  double* V2 = malloc(sizeof(double) * (size_t)(nvirt * nvirt));
  oracle_pair_V(nvirt, naux, B, 0, nocc - 1, V1);
  oracle_pair_V(nvirt, naux, B, nocc - 1, 0, V2);
  err = 0.0;
  for (int64_t a = 0; a < nvirt; a++)
    for (int64_t b = 0; b < nvirt; b++) err = fmax(err, fabs(V1[a + nvirt * b] - V2[b + nvirt * a]));
  check(err < 1e-13, "V^{ij} = (V^{ji})^T", nocc, nvirt);
  free(B);
  free(eo);
  free(ev);
  free(ep);
  free(pi);
  free(pj);
  free(sub);
  free(V1);
  free(V2);
}

static void eos_case(int nvirt, int nocc, int noccEOS, int nvirtEOS, int nba) {
  const size_t nv = nvirt, no = nocc, noE = noccEOS, nvE = nvirtEOS, q = nba;
  double* C = vec(q * nv * no, -1.0, 1.0);
  double* eo = vec(no, -2.0, -0.3);
  double* ev = vec(nv, 0.1, 3.0);
  double* UoE = vec(no * noE, -1.0, 1.0);
  double* Uv = vec(nv * nv, -1.0, 1.0);
  double* Uo = vec(no * no, -1.0, 1.0);
  double* UvE = vec(nv * nvE, -1.0, 1.0);
  double* t1 = malloc(sizeof(double) * nv * noE * nv * noE);
  double* t2 = malloc(sizeof(double) * nvE * no * nvE * no);
  double* g1 = malloc(sizeof(double) * nv * noE * nv * noE);
  double* g2 = malloc(sizeof(double) * nvE * no * nvE * no);
  oracle_steps5to8(nvirt, nocc, noccEOS, nvirtEOS, nba, C, eo, ev, UoE, Uv, Uo, UvE, t1, t2, g1, g2);
  int fin = 1;
  for (size_t k = 0; k < nv * noE * nv * noE; k++) fin &= isfinite(t1[k]) && isfinite(g1[k]);
  for (size_t k = 0; k < nvE * no * nvE * no; k++) fin &= isfinite(t2[k]) && isfinite(g2[k]);
  check(fin, "Steps 5-8 outputs finite", nvirt, nocc);
  if (nocc <= nvirt) { /* the reference's first-order block is in bounds only there (and nocctot = nocc) */
    double* Co = vec(q * no * no, -1.0, 1.0);
    double* Cvv = vec(q * nv * nv, -1.0, 1.0);
    double* djik = malloc(sizeof(double) * nv * noE * noE * no);
    double* blad = malloc(sizeof(double) * nvE * no * nvE * nv);
    oracle_first_order(nvirt, nocc, noccEOS, nvirtEOS, nba, nocc, C, Co, Cvv, UoE, Uv, Uo, UvE, djik, blad);
    fin = 1;
    for (size_t k = 0; k < nv * noE * noE * no; k++) fin &= isfinite(djik[k]);
    for (size_t k = 0; k < nvE * no * nvE * nv; k++) fin &= isfinite(blad[k]);
    check(fin, "first-order outputs finite", nvirt, nocc);
    free(Co);
    free(Cvv);
    free(djik);
    free(blad);
  }
  free(C);
  free(eo);
  free(ev);
  free(UoE);
  free(Uv);
  free(Uo);
  free(UvE);
  free(t1);
  free(t2);
  free(g1);
  free(g2);
}

static void mlp_case(int64_t n) {
  const int64_t np = oracle_mlp_num_params();
  check(np == 16 * 64 + 64 + 64 * 64 + 64 + 64 + 1, "MLP parameter count", np, 0);
  float* P = malloc(sizeof(float) * (size_t)np);
  float* F = malloc(sizeof(float) * (size_t)(n * 16 + 1));
  float* T = malloc(sizeof(float) * (size_t)(n + 1));
  float norm[34];
  for (int64_t k = 0; k < np; k++) P[k] = (float)(0.4 * urand() - 0.2);
  for (int64_t k = 0; k < n * 16; k++) F[k] = (float)(6.0 * urand() - 2.0);
  for (int64_t k = 0; k < n; k++) T[k] = (float)(3.0 + urand());
  for (int k = 0; k < 16; k++) {
    norm[k] = 1.0f;
    norm[16 + k] = 2.0f;
  }
  norm[32] = 3.0f;
  norm[33] = 1.5f;
  double* g = malloc(sizeof(double) * (size_t)np);
  double* pred = malloc(sizeof(double) * (size_t)(n + 1));
  for (int em = 0; em < 2; em++) {
    const double l1 = oracle_mlp_loss_grad(n, F, T, P, norm, em, g, pred);
    const double l2 = oracle_mlp_loss_grad(n, F, T, P, norm, em, NULL, NULL);
    check(n == 0 || (l1 == l2 && l1 > 0.0), "MLP loss with and without gradient", n, em);
    check(oracle_mlp_loss_grad(n, F, NULL, P, norm, em, NULL, pred) == 0.0, "MLP prediction only", n, em);
  }
  free(P);
  free(F);
  free(T);
  free(g);
  free(pred);
}

int main(void) {
  check(oracle_num_threads() >= 1, "thread count", 0, 0);
  const int64_t md[][2] = {{1, 1}, {3, 2}, {5, 7}, {31, 3}, {67, 5}, {130, 9}};
  for (size_t k = 0; k < sizeof(md) / sizeof(md[0]); k++) metric_case(md[k][0], md[k][1]);
  const int64_t pd[][3] = {{1, 1, 1}, {1, 3, 5}, {2, 5, 7}, {3, 4, 16}, {5, 9, 13}, {7, 17, 33}};
  for (size_t k = 0; k < sizeof(pd) / sizeof(pd[0]); k++) pair_case(pd[k][0], pd[k][1], pd[k][2]);
  const int ed[][5] = {{1, 1, 1, 1, 1}, {5, 3, 2, 2, 7}, {9, 4, 3, 5, 7}, {4, 6, 2, 3, 5}, {11, 5, 2, 3, 13}};
  for (size_t k = 0; k < sizeof(ed) / sizeof(ed[0]); k++) eos_case(ed[k][0], ed[k][1], ed[k][2], ed[k][3], ed[k][4]);
  mlp_case(0);
  mlp_case(1);
  mlp_case(37);
  printf("OK %d\n", g_checks);
  return 0;
}
