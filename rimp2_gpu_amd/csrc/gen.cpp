// Seeded synthetic fragment generator (host code, part of the library so that
// benchmarks and callers can synthesise inputs without any other dependency).
//
// Counter-based: every element is a pure function of (seed, stream, index), so
// the output is independent of thread count and identical on every machine.
//   key(seed, stream)  = mix64(seed * phi + (stream + 1) * 0xD1B54A32D192ED03)
//   u64(key, n)        = mix64(key + (n + 1) * phi)            (SplitMix64 step)
//   uniform53(key, n)  = (u64 >> 11) * 2^-53                    in [0, 1)
//   normal(key, n)     = Box-Muller on draws 2*(n/2), 2*(n/2)+1; cos for even n,
//                        sin for odd n
// Streams: 0 = (P|ia), 1 = metric factor A, 2 = occupied energies,
//          3 = virtual energies, 4 = batch fragment shapes.
// Locality-decay variant (lam > 0, rimp2_gen_fragment_local): orbitals sit on
// a 1-D chain, occupied i at x_i = i, auxiliary Q at x_Q = (Q+1/2) nocc/naux,
// virtual a at x_a = (a+1/2) nocc/nvirt;  (P|ia) is multiplied by
// exp(-(|x_Q - x_i| + |x_a - x_i|)/lam) and A[p,q] by exp(-|x_p - x_q|/lam),
// so pair couplings decay with |i - j| as in a spatially extended fragment.
// Distributions (virtual index a is 0-based):
//   (P|ia)[Q,a,i] = normal * 0.1 / sqrt(1 + a/16),  layout [naux, nvirt, nocc]
//   V = A A^T / naux + I,  A[p,q] = normal,          layout [naux, naux]
//   eps_o ~ U[-2.0, -0.3],  eps_v ~ U[0.1, 3.0]
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <vector>

#include "rimp2_common.h"

namespace {

constexpr uint64_t kPhi = 0x9E3779B97F4A7C15ULL;

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
inline uint64_t stream_key(uint64_t seed, uint64_t stream) {
  return mix64(seed * kPhi + (stream + 1) * 0xD1B54A32D192ED03ULL);
}
inline double uniform53(uint64_t key, uint64_t n) {
  return (double)(mix64(key + (n + 1) * kPhi) >> 11) * 0x1.0p-53;
}
inline double normal(uint64_t key, uint64_t n) {
  uint64_t k = n >> 1;
  double u1 = 1.0 - uniform53(key, 2 * k);  // (0, 1]
  double u2 = uniform53(key, 2 * k + 1);
  double r = std::sqrt(-2.0 * std::log(u1));
  double th = 6.283185307179586476925286766559 * u2;
  return (n & 1) ? r * std::sin(th) : r * std::cos(th);
}

}  // namespace

extern "C" {

int rimp2_gen_fragment_local(uint64_t seed, int64_t naux, int64_t nvirt, int64_t nocc, double lam, double* Pia,
                             double* V, double* eps_o, double* eps_v) {
  if (naux <= 0 || nvirt <= 0 || nocc <= 0) return RIMP2_ERR_ARG;
  if (!std::isfinite(lam)) return RIMP2_ERR_ARG;
  const bool local = lam > 0.0;
  if (Pia) {
    const uint64_t kx = stream_key(seed, 0);
    const int64_t ncol = nvirt * nocc;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < ncol; c++) {
      const int64_t a = c % nvirt, i = c / nvirt;
      double s = 0.1 / std::sqrt(1.0 + (double)a / 16.0);
      const double xa = ((double)a + 0.5) * (double)nocc / (double)nvirt;
      if (local) s *= std::exp(-std::fabs(xa - (double)i) / lam);
      double* col = Pia + naux * c;
      for (int64_t q = 0; q < naux; q++) {
        double v = normal(kx, (uint64_t)(q + naux * c)) * s;
        if (local) v *= std::exp(-std::fabs(((double)q + 0.5) * (double)nocc / (double)naux - (double)i) / lam);
        col[q] = v;
      }
    }
  }
  if (V) {
    const uint64_t ka = stream_key(seed, 1);
    // A stored transposed (row p contiguous) so that V_pq = sum_k A_pk A_qk streams.
    std::vector<double> At((size_t)(naux * naux));
    const double h = (double)nocc / (double)naux;
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < naux; p++)
      for (int64_t k = 0; k < naux; k++) {
        double v = normal(ka, (uint64_t)(p + naux * k));
        if (local) v *= std::exp(-std::fabs((double)(p - k)) * h / lam);
        At[(size_t)(k + naux * p)] = v;
      }
    const double inv = 1.0 / (double)naux;
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t q = 0; q < naux; q++) {
      const double* aq = &At[(size_t)(naux * q)];
      for (int64_t p = q; p < naux; p++) {
        const double* ap = &At[(size_t)(naux * p)];
        double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
        int64_t k = 0;
        for (; k + 4 <= naux; k += 4) {
          s0 += ap[k] * aq[k];
          s1 += ap[k + 1] * aq[k + 1];
          s2 += ap[k + 2] * aq[k + 2];
          s3 += ap[k + 3] * aq[k + 3];
        }
        for (; k < naux; k++) s0 += ap[k] * aq[k];
        double v = ((s0 + s1) + (s2 + s3)) * inv + (p == q ? 1.0 : 0.0);
        V[p + naux * q] = v;
        V[q + naux * p] = v;
      }
    }
  }
  if (eps_o) {
    const uint64_t ko = stream_key(seed, 2);
    for (int64_t i = 0; i < nocc; i++) eps_o[i] = -2.0 + 1.7 * uniform53(ko, (uint64_t)i);
  }
  if (eps_v) {
    const uint64_t kv = stream_key(seed, 3);
    for (int64_t a = 0; a < nvirt; a++) eps_v[a] = 0.1 + 2.9 * uniform53(kv, (uint64_t)a);
  }
  return RIMP2_OK;
}

int rimp2_gen_fragment(uint64_t seed, int64_t naux, int64_t nvirt, int64_t nocc, double* Pia, double* V,
                       double* eps_o, double* eps_v) {
  return rimp2_gen_fragment_local(seed, naux, nvirt, nocc, 0.0, Pia, V, eps_o, eps_v);
}

// Shapes of a batch of DEC fragments: nocc ~ U{4..32}, nvirt ~ U{32..256},
// naux = 4 * nvirt.  Fragment f uses seed + f for its contents.
int rimp2_gen_batch_shapes(uint64_t seed, int64_t nfrag, int64_t* nocc, int64_t* nvirt, int64_t* naux) {
  if (nfrag < 0 || !nocc || !nvirt || !naux) return RIMP2_ERR_ARG;
  const uint64_t kb = stream_key(seed, 4);
  for (int64_t f = 0; f < nfrag; f++) {
    int64_t o = 4 + (int64_t)(29.0 * uniform53(kb, (uint64_t)(2 * f)));
    int64_t v = 32 + (int64_t)(225.0 * uniform53(kb, (uint64_t)(2 * f + 1)));
    nocc[f] = o;
    nvirt[f] = v;
    naux[f] = 4 * v;
  }
  return RIMP2_OK;
}

}  // extern "C"
