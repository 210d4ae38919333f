// Plan of the grouped metric pipeline: the job tables that let one launch per
// step serve the Cholesky, triangular inverse, packing and transform of many
// fragments at once.  Plain C++ (no HIP): the library queues the plan's steps
// on a stream, and a host emulation checks the plan's indexing on the CPU.
//
// Per fragment (column-major, n = naux):
//   L slot    n x n     V copied in, factored in place (lower part used)
//   Linv slot n x n     L^{-1}: 64-wide diagonal blocks from the Cholesky step,
//                       off-diagonal blocks by recursive doubling
//   T slot    n x n     products of the doubling levels
//   Lp        L^{-1} in MFMA A-fragment order, Xp  X in B-fragment order
// Steps: copies with finiteness checks (every V into its L slot first),
// denominator checks, Linv cleared, then for each 64-column step k0 of the
// longest fragment: diagonal blocks (one workgroup per fragment still that
// long), panels L21 = A21 L11^{-T}, trailing updates A22 -= L21 L21^T; then
// per doubling level w (blocks of 64 w): T = Q P^{-1}, Linv[Q] = -R^{-1} T;
// then the L^{-1} and X packings and the transform B = L^{-1} X.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

namespace rimp2 {
namespace mg {

constexpr int64_t kPanel = 64;                             // Cholesky / inverse block width
constexpr int64_t kTrTM = 128, kTrTN = 128, kTrKQ = 64;    // transform tile rows, columns, K padding

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline int64_t rup(int64_t a, int64_t b) { return cdiv(a, b) * b; }

// Bytes of L^{-1} in A-fragment order, and of one X in B-fragment order.
inline size_t linv_packed_bytes(int64_t n) {
  return (size_t)(rup(n, kTrKQ) / 4) * (size_t)(rup(n, kTrTM) / 16) * 64 * sizeof(double);
}
inline size_t x_packed_bytes(int64_t n, int64_t nvirt, int64_t nocc) {
  return (size_t)(rup(n, kTrKQ) / 4) * (size_t)(rup(nvirt, kTrTN) / 16) * 64 * sizeof(double) * (size_t)nocc;
}

struct TransformArgs {
  const unsigned char* Lp;  // packed L^{-1}
  const unsigned char* Xp;  // packed X
  double* B;                // [naux, nvirt, nocc]
  int64_t naux, nvirt, nocc;
  int64_t MB, KS;           // L^{-1} row blocks (16) and k-steps (4)
  int64_t NRB;              // X row blocks (16 virtuals) per occupied
  int64_t MT, NT;           // row tiles, column tiles per occupied
  int64_t nwu, chunk;       // work units (row tiles t and MT-1-t of one column tile), per XCD
};

// Column grouping of the transform.  X and B are [n, nvirt, nocc] column-major,
// i.e. [n, nvirt * nocc]: the transform may take its 128-wide column tiles
// per occupied orbital (groups = nocc, cols = nvirt; each occupied orbital's
// packed X is one slab addressed with 32-bit offsets) or over the flat range
// (groups = 1, cols = nvirt * nocc), which pads each fragment by less than one
// tile instead of up to 127 columns per occupied orbital.  Flat wherever the
// one slab stays below 4 GiB.
inline void transform_columns(int64_t n, int64_t nvirt, int64_t nocc, int64_t& cols, int64_t& groups) {
  const size_t flat_slab = (size_t)(rup(n, kTrKQ) / 4) * (size_t)(rup(nvirt * nocc, kTrTN) / 16) * 512;
  if (flat_slab <= 0xFFFFFFFFull) {
    cols = nvirt * nocc;
    groups = 1;
  } else {
    cols = nvirt;
    groups = nocc;
  }
}

inline TransformArgs transform_args(int64_t n, int64_t nvirt, int64_t nocc) {
  TransformArgs a{};
  a.naux = n;
  a.nvirt = nvirt;
  a.nocc = nocc;
  a.MB = rup(n, kTrTM) / 16;
  a.KS = rup(n, kTrKQ) / 4;
  a.NRB = rup(nvirt, kTrTN) / 16;
  a.MT = rup(n, kTrTM) / kTrTM;
  a.NT = a.NRB / 8;
  a.nwu = ((a.MT + 1) / 2) * a.NT * nocc;
  a.chunk = cdiv(a.nwu, 8);
  return a;
}

// C = alpha A op(B) + beta C over 64 x 64 tiles; job j owns the global tiles
// [first, first + tiles); LOWER_C jobs (M = N) list only the lower tiles.
struct GemmJob {
  const double* A;
  const double* B;
  double* C;
  int64_t M, N, K, lda, ldb, ldc;
  double alpha, beta;
  int64_t tm;     // row tiles
  int64_t first;  // first global tile
};
struct PotrfJob {  // one 64-wide diagonal block of one fragment
  double* A;       // L slot (in place), leading dimension n
  double* Linv;    // Linv slot, leading dimension n
  int64_t n, k0;
  int kb;
  int* status;
};
struct CopyJob {  // dst = src (dst == src: check only); bit 4 on a non-finite value
  const double* src;
  double* dst;
  int* status;
  int64_t first;  // first global element
};
struct DenomJob {  // bit 8 when max(eo) >= min(ev)
  const double* eo;
  const double* ev;
  int64_t nocc, nvirt;
  int* status;
};
// L^{-1} of one fragment into A-fragment order Lp[ks][mb][lane] (lane l <->
// row 16 mb + l % 16, column 4 ks + l / 16), in tiles of 64 rows (4 row
// blocks) x 64 columns (16 k-steps); tiles = (MB / 4) * (KS / 16).  The
// transform's K loop for a 128-row tile stops at that tile's last row, so
// tile (mbt, kst) is read only when kst < 2 (mbt / 2 + 1); the others are
// never written (linv_tile_read).
struct LinvPackJob {
  const double* Li;
  double* Lp;
  int64_t n, MB, KS;
  int64_t first;  // first global tile
};
inline bool linv_tile_read(int64_t mbt, int64_t kst) { return kst < 2 * (mbt / 2 + 1); }
// X of one fragment into B-fragment order Xp[g][ks][rb][lane] (lane l <->
// q = 4 ks + l / 16, column 16 rb + l % 16 of group g), in tiles of 64 q x
// 64 columns (16 k-steps x 4 row blocks); bit 4 of status on a non-finite value.
struct XPackJob {
  const double* X;
  double* Xp;
  int64_t naux, cols, nks, nrb;  // cols per group; nks % 16 == 0, nrb % 4 == 0
  int* status;
  int64_t first;  // first global tile; tiles = groups * (nks / 16) * (nrb / 4)
};
struct TransformJob {
  TransformArgs a;
  int64_t first;  // first global work unit
};

// K stages (16 deep) of transform row tile mt, and of work unit t of a column
// tile (row tiles t and MT-1-t): the unit's cost, near enough.
inline int64_t transform_tile_stages(const TransformArgs& a, int64_t mt) {
  return std::min((mt + 1) * kTrTM, a.KS * 4) / 16;
}
inline int64_t transform_unit_stages(const TransformArgs& a, int64_t t) {
  return transform_tile_stages(a, t) + (a.MT - 1 - t != t ? transform_tile_stages(a, a.MT - 1 - t) : 0);
}

// Units [b[x], b[x + 1]) of a grouped transform go to XCD x.  The jobs list
// fragments by decreasing naux, so equal unit counts per XCD would give XCD 0
// the longest units; instead b[x] is the first unit at which the units before
// it reach x/8 of the total cost.
struct XcdRanges {
  int64_t b[9];
};
inline XcdRanges transform_xcd_ranges(const TransformJob* jobs, int njobs, int64_t total) {
  XcdRanges r{};
  std::vector<int64_t> col_cost((size_t)njobs);
  double all = 0.0;
  for (int j = 0; j < njobs; j++) {
    const TransformArgs& a = jobs[j].a;
    int64_t c = 0;
    for (int64_t t = 0; t < (a.MT + 1) / 2; t++) c += transform_unit_stages(a, t);
    col_cost[(size_t)j] = c;
    all += (double)c * (double)(a.nwu / ((a.MT + 1) / 2));
  }
  double acc = 0.0;  // cost of the units before job j
  int x = 1;
  for (int j = 0; j < njobs && x < 8; j++) {
    const TransformArgs& a = jobs[j].a;
    const int64_t mth = (a.MT + 1) / 2, cc = col_cost[(size_t)j];
    const double job = (double)cc * (double)(a.nwu / mth);
    while (x < 8 && all * x / 8.0 <= acc + job) {
      // smallest k with (cost of the job's first k units) >= rem
      const double rem = all * x / 8.0 - acc;
      int64_t full = cc > 0 ? (int64_t)(rem / (double)cc) : 0;
      if (full * mth > a.nwu) full = a.nwu / mth;
      double got = (double)full * (double)cc;
      int64_t k = full * mth, t = 0;
      while (got < rem && k < a.nwu) {
        got += (double)transform_unit_stages(a, t);
        t++;
        k++;
      }
      r.b[x++] = jobs[j].first + k;
    }
    acc += job;
  }
  for (; x < 8; x++) r.b[x] = total;
  r.b[0] = 0;
  r.b[8] = total;
  return r;
}

enum GroupLaunch { GL_COPY, GL_DENOM, GL_ZERO_LINV, GL_POTRF, GL_PANEL, GL_SYRK, GL_TRI_T, GL_TRI_L, GL_LPACK,
                   GL_XPACK, GL_TRANSFORM };

// GL_PANEL: TRANSB, GL_SYRK: TRANSB + LOWER_C, GL_TRI_T: plain, GL_TRI_L: TRI_A.
struct GroupStep {
  GroupLaunch kind;
  size_t off;     // job table offset in the blob (GL_ZERO_LINV: none)
  int njobs;
  int64_t total;  // elements / tiles / units / workgroups of the launch (GL_ZERO_LINV: bytes)
};

struct Item {      // one fragment with a metric (device arrays)
  int64_t n, nvirt, nocc;
  const double* V;  // [n, n]
  const double* X;  // [n, nvirt, nocc]
  double* B;        // [n, nvirt, nocc] out
  int* status;      // 2 ints: bit 2 + [1] first failing pivot; bit 4 non-finite V or X
};
struct Copy {      // dst[0..n) = src[0..n) (dst == src: check only)
  const double* src;
  double* dst;
  int64_t n;
  int* status;
};
struct Denom {
  const double* eo;
  const double* ev;
  int64_t nocc, nvirt;
  int* status;
};

// Arena layout: fragments by decreasing n, every slot 256-byte aligned.
struct Layout {
  std::vector<size_t> order, lo, lpo, xpo;  // lo: L / Linv / T slot offsets (bytes)
  size_t ln = 0, lpn = 0, xpn = 0;          // arena bytes
  int64_t nmax = 0;
};

inline Layout layout(const std::vector<Item>& items) {
  Layout g;
  const size_t ni = items.size();
  g.order.resize(ni);
  std::iota(g.order.begin(), g.order.end(), 0);
  std::stable_sort(g.order.begin(), g.order.end(), [&](size_t x, size_t y) { return items[x].n > items[y].n; });
  g.lo.resize(ni);
  g.lpo.resize(ni);
  g.xpo.resize(ni);
  for (size_t k : g.order) {
    const Item& it = items[k];
    g.lo[k] = g.ln;
    g.ln += (size_t)rup((int64_t)(it.n * it.n * sizeof(double)), 256);
    g.lpo[k] = g.lpn;
    g.lpn += (size_t)rup((int64_t)linv_packed_bytes(it.n), 256);
    int64_t cols, groups;
    transform_columns(it.n, it.nvirt, it.nocc, cols, groups);
    g.xpo[k] = g.xpn;
    g.xpn += (size_t)rup((int64_t)x_packed_bytes(it.n, cols, groups), 256);
    g.nmax = std::max(g.nmax, it.n);
  }
  return g;
}

struct Arenas {
  unsigned char *L, *Linv, *T, *Lp, *Xp;  // at least ln, ln, ln, lpn, xpn bytes
};

// Fills steps and blob (the job tables, each 256-byte aligned in the blob).
// transform = false: the plan ends with the inverse (L^{-1} in the Linv arena);
// the caller transforms by other means and checks X itself.
inline void plan(const std::vector<Item>& items, const std::vector<Copy>& copies, const std::vector<Denom>& denoms,
                 const Layout& g, const Arenas& ar, std::vector<GroupStep>& steps, std::vector<unsigned char>& blob,
                 bool transform = true) {
  steps.clear();
  blob.clear();
  auto Lk = [&](size_t k) { return reinterpret_cast<double*>(ar.L + g.lo[k]); };
  auto Lik = [&](size_t k) { return reinterpret_cast<double*>(ar.Linv + g.lo[k]); };
  auto Tk = [&](size_t k) { return reinterpret_cast<double*>(ar.T + g.lo[k]); };
  auto put = [&](GroupLaunch kind, const void* p, size_t bytes, int njobs, int64_t total) {
    if (njobs == 0 || total == 0) return;
    const size_t off = (size_t)rup((int64_t)blob.size(), 256);
    blob.resize(off + bytes);
    std::memcpy(blob.data() + off, p, bytes);
    steps.push_back(GroupStep{kind, off, njobs, total});
  };
  auto gemm = [](const double* A, int64_t lda, const double* B, int64_t ldb, double* C, int64_t ldc, int64_t M,
                 int64_t N, int64_t K, double alpha, double beta, bool lower, int64_t& first) {
    GemmJob j{A, B, C, M, N, K, lda, ldb, ldc, alpha, beta, cdiv(M, kPanel), first};
    first += lower ? j.tm * (j.tm + 1) / 2 : j.tm * cdiv(N, kPanel);
    return j;
  };
  {  // the caller's copies, then every V into its L slot
    std::vector<CopyJob> cj;
    int64_t e = 0;
    for (const Copy& c : copies) {
      if (c.n <= 0) continue;
      cj.push_back(CopyJob{c.src, c.dst, c.status, e});
      e += c.n;
    }
    for (size_t k : g.order) {
      cj.push_back(CopyJob{items[k].V, Lk(k), items[k].status, e});
      e += items[k].n * items[k].n;
    }
    put(GL_COPY, cj.data(), cj.size() * sizeof(CopyJob), (int)cj.size(), e);
  }
  {
    std::vector<DenomJob> dj;
    for (const Denom& d : denoms) dj.push_back(DenomJob{d.eo, d.ev, d.nocc, d.nvirt, d.status});
    put(GL_DENOM, dj.data(), dj.size() * sizeof(DenomJob), (int)dj.size(), (int64_t)dj.size());
  }
  if (!items.empty()) steps.push_back(GroupStep{GL_ZERO_LINV, 0, 1, (int64_t)g.ln});
  for (int64_t k0 = 0; k0 < g.nmax; k0 += kPanel) {
    std::vector<PotrfJob> pj;
    std::vector<GemmJob> panel, syrk;
    int64_t tp = 0, ts = 0;
    for (size_t k : g.order) {
      const Item& it = items[k];
      const int64_t n = it.n;
      if (n <= k0) break;  // by decreasing n: no later fragment is this long
      const int kb = (int)std::min<int64_t>(kPanel, n - k0);
      pj.push_back(PotrfJob{Lk(k), Lik(k), n, k0, kb, it.status});
      const int64_t r0 = k0 + kb, m = n - r0;
      if (m <= 0) continue;
      double* L21 = Lk(k) + r0 + n * k0;
      panel.push_back(gemm(L21, n, Lik(k) + k0 + n * k0, n, L21, n, m, kb, kb, 1.0, 0.0, false, tp));
      syrk.push_back(gemm(L21, n, L21, n, Lk(k) + r0 + n * r0, n, m, m, kb, -1.0, 1.0, true, ts));
    }
    put(GL_POTRF, pj.data(), pj.size() * sizeof(PotrfJob), (int)pj.size(), (int64_t)pj.size());
    put(GL_PANEL, panel.data(), panel.size() * sizeof(GemmJob), (int)panel.size(), tp);
    put(GL_SYRK, syrk.data(), syrk.size() * sizeof(GemmJob), (int)syrk.size(), ts);
  }
  // for [[P, 0], [Q, R]]: inverse [[P^-1, 0], [-R^-1 Q P^-1, R^-1]]; the
  // groups of a level are independent; a ragged last group's T follows the
  // full groups' T in the slot (nfull half^2 + sr half <= n^2)
  for (int64_t w = 1; w * kPanel < g.nmax; w *= 2) {
    std::vector<GemmJob> tj, lj;
    int64_t tt = 0, tl = 0;
    for (size_t k : g.order) {
      const int64_t n = items[k].n;
      const int64_t half = w * kPanel, span = 2 * half, nfull = n / span;
      if (half >= n) break;
      double *Lf = Lk(k), *Lif = Lik(k), *Tf = Tk(k);
      for (int64_t q = 0; q < nfull; q++) {
        const int64_t l0 = q * span, mid = l0 + half;
        double* Tq = Tf + q * half * half;
        tj.push_back(gemm(Lf + mid + n * l0, n, Lif + l0 + n * l0, n, Tq, half, half, half, half, 1.0, 0.0, false, tt));
        lj.push_back(gemm(Lif + mid * (1 + n), n, Tq, half, Lif + mid + n * l0, n, half, half, half, -1.0, 0.0, false,
                          tl));
      }
      const int64_t l0 = nfull * span, mid = l0 + half;
      if (mid < n) {
        const int64_t sr = n - mid;
        double* Tq = Tf + nfull * half * half;
        tj.push_back(gemm(Lf + mid + n * l0, n, Lif + l0 + n * l0, n, Tq, sr, sr, half, half, 1.0, 0.0, false, tt));
        lj.push_back(gemm(Lif + mid * (1 + n), n, Tq, sr, Lif + mid + n * l0, n, sr, half, sr, -1.0, 0.0, false, tl));
      }
    }
    put(GL_TRI_T, tj.data(), tj.size() * sizeof(GemmJob), (int)tj.size(), tt);
    put(GL_TRI_L, lj.data(), lj.size() * sizeof(GemmJob), (int)lj.size(), tl);
  }
  if (transform) {
    std::vector<LinvPackJob> lp;
    std::vector<XPackJob> xp;
    std::vector<TransformJob> tr;
    int64_t el = 0, ex = 0, eu = 0;
    for (size_t k : g.order) {
      const Item& it = items[k];
      int64_t cols, groups;
      transform_columns(it.n, it.nvirt, it.nocc, cols, groups);
      TransformArgs a = transform_args(it.n, cols, groups);
      a.Lp = ar.Lp + g.lpo[k];
      a.Xp = ar.Xp + g.xpo[k];
      a.B = it.B;
      lp.push_back(LinvPackJob{Lik(k), reinterpret_cast<double*>(ar.Lp + g.lpo[k]), it.n, a.MB, a.KS, el});
      el += (a.MB / 4) * (a.KS / 16);
      xp.push_back(XPackJob{it.X, reinterpret_cast<double*>(ar.Xp + g.xpo[k]), it.n, cols, a.KS, a.NRB, it.status, ex});
      ex += groups * (a.KS / 16) * (a.NRB / 4);
      tr.push_back(TransformJob{a, eu});
      eu += a.nwu;
    }
    put(GL_LPACK, lp.data(), lp.size() * sizeof(LinvPackJob), (int)lp.size(), el);
    put(GL_XPACK, xp.data(), xp.size() * sizeof(XPackJob), (int)xp.size(), ex);
    put(GL_TRANSFORM, tr.data(), tr.size() * sizeof(TransformJob), (int)tr.size(), eu);
  }
}

}  // namespace mg
}  // namespace rimp2
