// Host emulation of the grouped metric pipeline's plan (metric_group.h).
//
// The plan's job tables are built over host arenas, and every step is run
// here with plain loops that read the job tables the way the device kernels
// do (job lookup by `first`, tile decode, TRANSB / TRI_A / LOWER_C, packed
// fragment orders).  Checked:
//   - every access lies inside a registered buffer;
//   - within one launch no tile reads an element another tile writes (the
//     tiles of a launch run concurrently on the device);
//   - L (lower part), L^{-1} (whole slot) and B = L^{-1} X of every fragment
//     against a direct Cholesky, forward substitution and product;
//   - the status words: a non-positive pivot, a non-finite input, a
//     denominator >= 0 each flag exactly their fragment.
// Prints "OK <launches>" and exits 0, or prints the first failure and exits 1.
// With the argument --small only the fragments up to 129 auxiliary functions
// take part (every panel and doubling case, a few seconds under a sanitizer).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <unordered_map>
#include <vector>

#include "../../rimp2_gpu_amd/csrc/metric_group.h"

using namespace rimp2::mg;

namespace {

struct Region {
  const unsigned char* lo;
  size_t bytes;
  const char* name;
};
std::vector<Region> g_regions;

[[noreturn]] void die(const char* msg, long long a = 0, long long b = 0) {
  std::printf("FAIL %s (%lld, %lld)\n", msg, a, b);
  std::exit(1);
}

void reg(const void* p, size_t bytes, const char* name) {
  g_regions.push_back(Region{static_cast<const unsigned char*>(p), bytes, name});
}

void chk(const void* p, size_t bytes, const char* what) {
  const unsigned char* q = static_cast<const unsigned char*>(p);
  for (const Region& r : g_regions)
    if (q >= r.lo && q + bytes <= r.lo + r.bytes) return;
  std::printf("FAIL out-of-bounds access: %s\n", what);
  std::exit(1);
}

template <class T>
T rd(const T* p, const char* what) {
  chk(p, sizeof(T), what);
  return *p;
}
template <class T>
void wr(T* p, T v, const char* what) {
  chk(p, sizeof(T), what);
  *p = v;
}

uint64_t g_rng = 88172645463325252ull;
double urand() {  // xorshift64, uniform in [0, 1)
  g_rng ^= g_rng << 13;
  g_rng ^= g_rng >> 7;
  g_rng ^= g_rng << 17;
  return (double)(g_rng >> 11) * (1.0 / 9007199254740992.0);
}

// Concurrency check of one launch: element -> tile that wrote it.
struct Hazards {
  std::unordered_map<const double*, int64_t> writer;
  std::vector<std::pair<const double*, int64_t>> reads;
  void w(const double* p, int64_t tile) {
    auto it = writer.find(p);
    if (it != writer.end() && it->second != tile) die("two tiles write one element", it->second, tile);
    writer[p] = tile;
  }
  void r(const double* p, int64_t tile) { reads.emplace_back(p, tile); }
  void verify() {
    for (auto& x : reads) {
      auto it = writer.find(x.first);
      if (it != writer.end() && it->second != x.second) die("a tile reads what another tile writes", it->second, x.second);
    }
    writer.clear();
    reads.clear();
  }
};

template <class Job>
int64_t job_end(const Job* J, int nj, int j, int64_t total) {
  return j + 1 < nj ? J[j + 1].first : total;
}

template <class Job>
void check_firsts(const Job* J, int nj, int64_t total) {
  if (J[0].first != 0) die("first job does not start at 0");
  for (int j = 1; j < nj; j++)
    if (J[j].first < J[j - 1].first) die("job firsts not ascending", j);
  if (J[nj - 1].first > total) die("job first beyond total");
}

void tri_decode(int64_t p, int64_t& i, int64_t& j) {  // j(j+1)/2 + i -> (i, j), i <= j
  int64_t jj = (int64_t)((std::sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
  while (jj * (jj + 1) / 2 > p) jj--;
  while ((jj + 1) * (jj + 2) / 2 <= p) jj++;
  j = jj;
  i = p - jj * (jj + 1) / 2;
}

void run_gemm(const GemmJob* J, int nj, int64_t total, bool transb, bool tri_a, bool lower, Hazards& hz) {
  check_firsts(J, nj, total);
  for (int j = 0; j < nj; j++) {
    const GemmJob& g = J[j];
    const int64_t tiles = job_end(J, nj, j, total) - g.first;
    const int64_t tn = cdiv(g.N, kPanel);
    if (tiles != (lower ? g.tm * (g.tm + 1) / 2 : g.tm * tn)) die("gemm job tile count", j, tiles);
    if (lower && g.M != g.N) die("lower gemm job not square", j);
    for (int64_t u = 0; u < tiles; u++) {
      const int64_t tile = g.first + u;
      int64_t mt, nt;
      if (lower) tri_decode(u, nt, mt);
      else {
        mt = u % g.tm;
        nt = u / g.tm;
      }
      const int64_t m0 = mt * kPanel, n0 = nt * kPanel;
      int64_t kend = g.K;
      if (tri_a) kend = std::min(g.K, m0 + kPanel);
      std::vector<std::pair<double*, double>> out;
      for (int64_t m = m0; m < std::min(g.M, m0 + kPanel); m++)
        for (int64_t n = n0; n < std::min(g.N, n0 + kPanel); n++) {
          if (lower && n > m) continue;
          double acc = 0.0;
          for (int64_t k = 0; k < kend; k++) {
            const double* pa = g.A + m + g.lda * k;
            const double* pb = transb ? g.B + n + g.ldb * k : g.B + k + g.ldb * n;
            acc += rd(pa, "gemm A") * rd(pb, "gemm B");
            hz.r(pa, tile);
            hz.r(pb, tile);
          }
          double* pc = g.C + m + g.ldc * n;
          double v = g.alpha * acc;
          if (g.beta != 0.0) {
            v += g.beta * rd(pc, "gemm C read");
            hz.r(pc, tile);
          }
          out.emplace_back(pc, v);
        }
      for (auto& o : out) {  // the epilogue writes after the tile's K loop
        wr(o.first, o.second, "gemm C");
        hz.w(o.first, tile);
      }
    }
  }
  hz.verify();
}

// Diagonal block: L11 (zeros above) back into A, L11^{-1} (zeros above) into Linv.
void run_potrf(const PotrfJob* J, int nj) {
  for (int j = 0; j < nj; j++) {
    const PotrfJob& p = J[j];
    const int kb = p.kb;
    if (kb < 1 || kb > kPanel || p.k0 + kb > p.n) die("potrf block", j, kb);
    std::vector<double> a(kb * kb, 0.0), l(kb * kb, 0.0), x(kb * kb, 0.0);
    for (int r = 0; r < kb; r++)
      for (int q = 0; q <= r; q++) a[r + kb * q] = rd(p.A + (p.k0 + r) + p.n * (p.k0 + q), "potrf A");
    int bad = 0;
    for (int c = 0; c < kb; c++) {
      double d = a[c + kb * c];
      for (int k = 0; k < c; k++) d -= l[c + kb * k] * l[c + kb * k];
      if (!(d > 0.0)) {
        if (!bad) bad = c + 1;
        d = 1.0;
      }
      const double s = std::sqrt(d);
      l[c + kb * c] = s;
      for (int r = c + 1; r < kb; r++) {
        double v = a[r + kb * c];
        for (int k = 0; k < c; k++) v -= l[r + kb * k] * l[c + kb * k];
        l[r + kb * c] = v / s;
      }
    }
    for (int col = 0; col < kb; col++)
      for (int c = col; c < kb; c++) {
        double v = (c == col) ? 1.0 : 0.0;
        for (int k = col; k < c; k++) v -= l[c + kb * k] * x[k + kb * col];
        x[c + kb * col] = v / l[c + kb * c];
      }
    for (int r = 0; r < kb; r++)
      for (int q = 0; q < kb; q++) {
        wr(p.A + (p.k0 + r) + p.n * (p.k0 + q), q <= r ? l[r + kb * q] : 0.0, "potrf A write");
        wr(p.Linv + (p.k0 + r) + p.n * (p.k0 + q), q <= r ? x[r + kb * q] : 0.0, "potrf Linv write");
      }
    if (bad) {
      if (rd(p.status + 1, "status") == 0) wr(p.status + 1, (int)(p.k0 + bad), "status");
      wr(p.status, rd(p.status, "status") | 2, "status");
    }
  }
}

struct Frag {
  int64_t n, nvirt, nocc;
  bool fitted;
  std::vector<double> V, X, B, eo, ev;
  std::unique_ptr<int[]> status;
};

}  // namespace

int main(int argc, char** argv) {
  // (n, nvirt, nocc): below, at and above the 64-wide panel; doubling levels
  // with several full groups and ragged last groups
  const int64_t dims[][3] = {{1, 3, 1},   {5, 17, 2},   {64, 40, 1},  {65, 17, 3},  {100, 130, 2},
                             {128, 40, 1}, {129, 3, 2},  {200, 17, 1}, {257, 40, 2}, {320, 130, 1}};
  const bool small = argc > 1 && std::strcmp(argv[1], "--small") == 0;
  const int nd = small ? 7 : (int)(sizeof(dims) / sizeof(dims[0]));
  std::vector<Frag> F(nd + 1);
  for (int f = 0; f <= nd; f++) {
    Frag& x = F[f];
    x.fitted = (f == nd);
    x.n = x.fitted ? 30 : dims[f][0];
    x.nvirt = x.fitted ? 7 : dims[f][1];
    x.nocc = x.fitted ? 2 : dims[f][2];
    const int64_t n = x.n;
    x.X.resize(n * x.nvirt * x.nocc);
    for (double& v : x.X) v = urand() - 0.5;
    x.B.assign(n * x.nvirt * x.nocc, 0.0);
    x.eo.resize(x.nocc);
    x.ev.resize(x.nvirt);
    for (double& v : x.eo) v = -2.0 + 1.7 * urand();
    for (double& v : x.ev) v = 0.1 + 2.9 * urand();
    x.status.reset(new int[2]{0, 0});
    if (!x.fitted) {  // V = A A^T / n + I
      std::vector<double> A(n * n);
      for (double& v : A) v = urand() - 0.5;
      x.V.assign(n * n, 0.0);
      for (int64_t i = 0; i < n; i++)
        for (int64_t j = 0; j < n; j++) {
          double s = 0.0;
          for (int64_t k = 0; k < n; k++) s += A[i + n * k] * A[j + n * k];
          x.V[i + n * j] = s / (double)n + (i == j ? 1.0 : 0.0);
        }
      reg(x.V.data(), x.V.size() * 8, "V");
    }
    reg(x.X.data(), x.X.size() * 8, "X");
    reg(x.B.data(), x.B.size() * 8, "B");
    reg(x.eo.data(), x.eo.size() * 8, "eo");
    reg(x.ev.data(), x.ev.size() * 8, "ev");
    reg(x.status.get(), 8, "status");
  }
  auto run_case = [&](int bad_frag, int bad_kind, bool transform = true) -> int {
    // bad_kind: 0 none, 1 non-positive pivot, 2 NaN in X, 3 denominator >= 0
    // transform = false: the plan of the int8 transform's callers, which ends
    // with L^{-1}; X is neither read nor checked by it
    const size_t nreg = g_regions.size();  // this case's buffers are unregistered at its end
    std::vector<Item> items;
    std::vector<Copy> copies;
    std::vector<Denom> denoms;
    std::vector<std::vector<double>> Vs(F.size()), Xs(F.size()), evs(F.size());
    for (size_t f = 0; f < F.size(); f++) {
      Frag& x = F[f];
      x.status[0] = x.status[1] = 0;
      std::fill(x.B.begin(), x.B.end(), 0.0);
      Vs[f] = x.V;
      Xs[f] = x.X;
      evs[f] = x.ev;
      if ((int)f == bad_frag && bad_kind == 1) Vs[f][3 + x.n * 3] = -1.0;
      if ((int)f == bad_frag && bad_kind == 2) Xs[f][Xs[f].size() / 2] = NAN;
      if ((int)f == bad_frag && bad_kind == 3)
        for (double& v : evs[f]) v -= 10.0;
      if (!Vs[f].empty()) reg(Vs[f].data(), Vs[f].size() * 8, "V case");
      reg(Xs[f].data(), Xs[f].size() * 8, "X case");
      reg(evs[f].data(), evs[f].size() * 8, "ev case");
      if (x.fitted) copies.push_back(Copy{Xs[f].data(), x.B.data(), (int64_t)x.B.size(), x.status.get()});
      else items.push_back(Item{x.n, x.nvirt, x.nocc, Vs[f].data(), Xs[f].data(), x.B.data(), x.status.get()});
      denoms.push_back(Denom{x.eo.data(), evs[f].data(), x.nocc, x.nvirt, x.status.get()});
    }
    const Layout g = layout(items);
    std::vector<unsigned char> L(g.ln), Li(g.ln), T(g.ln), Lp(g.lpn), Xp(g.xpn);
    for (auto* v : {&L, &Li, &T, &Lp, &Xp}) {
      std::fill(v->begin(), v->end(), (unsigned char)0xFF);  // NaN garbage: unwritten reads show
      reg(v->data(), v->size(), "arena");
    }
    Arenas ar{L.data(), Li.data(), T.data(), Lp.data(), Xp.data()};
    std::vector<GroupStep> steps;
    std::vector<unsigned char> blob;
    plan(items, copies, denoms, g, ar, steps, blob, transform);
    for (const GroupStep& st : steps)
      if (!transform && (st.kind == GL_LPACK || st.kind == GL_XPACK || st.kind == GL_TRANSFORM))
        die("transform step planned with transform = false", (int)st.kind);
    Hazards hz;
    for (const GroupStep& st : steps) {
      const unsigned char* jp = blob.data() + st.off;
      if (st.kind != GL_ZERO_LINV && (st.off % 256 || st.off >= blob.size())) die("job table offset", (long long)st.off);
      switch (st.kind) {
        case GL_COPY: {
          const CopyJob* J = reinterpret_cast<const CopyJob*>(jp);
          check_firsts(J, st.njobs, st.total);
          for (int j = 0; j < st.njobs; j++)
            for (int64_t e = 0; e < job_end(J, st.njobs, j, st.total) - J[j].first; e++) {
              const double v = rd(J[j].src + e, "copy src");
              if (J[j].dst != J[j].src) wr(J[j].dst + e, v, "copy dst");
              if (!std::isfinite(v)) wr(J[j].status, rd(J[j].status, "status") | 4, "status");
            }
          break;
        }
        case GL_DENOM: {
          const DenomJob* J = reinterpret_cast<const DenomJob*>(jp);
          for (int j = 0; j < st.njobs; j++) {
            double mo = -INFINITY, mv = INFINITY;
            for (int64_t i = 0; i < J[j].nocc; i++) mo = std::fmax(mo, rd(J[j].eo + i, "eo"));
            for (int64_t a = 0; a < J[j].nvirt; a++) mv = std::fmin(mv, rd(J[j].ev + a, "ev"));
            if (!(2.0 * mo - 2.0 * mv < 0.0)) wr(J[j].status, rd(J[j].status, "status") | 8, "status");
          }
          break;
        }
        case GL_ZERO_LINV:
          chk(ar.Linv, (size_t)st.total, "Linv clear");
          std::fill(Li.begin(), Li.begin() + st.total, (unsigned char)0);
          break;
        case GL_POTRF: run_potrf(reinterpret_cast<const PotrfJob*>(jp), st.njobs); break;
        case GL_PANEL: run_gemm(reinterpret_cast<const GemmJob*>(jp), st.njobs, st.total, true, false, false, hz); break;
        case GL_SYRK: run_gemm(reinterpret_cast<const GemmJob*>(jp), st.njobs, st.total, true, false, true, hz); break;
        case GL_TRI_T: run_gemm(reinterpret_cast<const GemmJob*>(jp), st.njobs, st.total, false, false, false, hz); break;
        case GL_TRI_L: run_gemm(reinterpret_cast<const GemmJob*>(jp), st.njobs, st.total, false, true, false, hz); break;
        case GL_LPACK: {
          const LinvPackJob* J = reinterpret_cast<const LinvPackJob*>(jp);
          check_firsts(J, st.njobs, st.total);
          // tiles of 64 rows x 64 columns; tiles the transform does not read
          // stay unwritten (NaN in this arena, so a read of one would show)
          for (int j = 0; j < st.njobs; j++) {
            const LinvPackJob& x = J[j];
            const int64_t kt = x.KS / 16, tiles = job_end(J, st.njobs, j, st.total) - x.first;
            if (x.MB % 4 || x.KS % 16 || tiles != (x.MB / 4) * kt) die("L^-1 packing tile count", j, tiles);
            for (int64_t u = 0; u < tiles; u++) {
              const int64_t mbt = u / kt, kst = u % kt;
              if (!linv_tile_read(mbt, kst)) continue;
              for (int64_t ks = kst * 16; ks < kst * 16 + 16; ks++)
                for (int64_t mb = mbt * 4; mb < mbt * 4 + 4; mb++)
                  for (int lane = 0; lane < 64; lane++) {
                    const int64_t m = mb * 16 + (lane & 15), k = ks * 4 + (lane >> 4), n = x.n;
                    wr(x.Lp + (ks * x.MB + mb) * 64 + lane, (m < n && k < n && k <= m) ? rd(x.Li + m + n * k, "Linv") : 0.0,
                       "Lp");
                  }
            }
          }
          break;
        }
        case GL_XPACK: {
          // tiles of 64 q x 64 columns; each (q, column) of group g goes to
          // fragment (ks = q / 4, rb = column / 16), lane (q % 4) 16 + column % 16;
          // every element of the packed X written exactly once
          const XPackJob* J = reinterpret_cast<const XPackJob*>(jp);
          check_firsts(J, st.njobs, st.total);
          for (int j = 0; j < st.njobs; j++) {
            const XPackJob& x = J[j];
            const Item* it = nullptr;
            for (const Item& c : items)
              if (c.X == x.X) it = &c;
            if (!it) die("X packing job of no fragment", j);
            if (x.nks % 16 || x.nrb % 4) die("X packing tile shape", j);
            const int64_t per = (x.nks / 16) * (x.nrb / 4), tiles = job_end(J, st.njobs, j, st.total) - x.first;
            if (tiles % per) die("X packing tile count", j);
            const int64_t groups = tiles / per;
            if (x.cols * groups != it->nvirt * it->nocc || (groups != 1 && x.cols != it->nvirt))
              die("X packing column grouping", j, groups);
            std::vector<unsigned char> seen((size_t)(groups * x.nks * x.nrb * 64), 0);
            for (int64_t g0 = 0; g0 < groups; g0++)
              for (int64_t q = 0; q < x.nks * 4; q++)
                for (int64_t c = 0; c < x.nrb * 16; c++) {
                  const double v =
                      (q < x.naux && c < x.cols) ? rd(x.X + x.naux * x.cols * g0 + q + x.naux * c, "X") : 0.0;
                  const int64_t idx = ((g0 * x.nks + q / 4) * x.nrb + c / 16) * 64 + (q % 4) * 16 + c % 16;
                  if (seen[idx]++) die("X packing writes an element twice", j);
                  wr(x.Xp + idx, v, "Xp");
                  if (!std::isfinite(v)) wr(x.status, rd(x.status, "status") | 4, "status");
                }
          }
          break;
        }
        case GL_TRANSFORM: {
          // B = L^{-1} X from the packed operands; each unit is a 128 x 128
          // tile of one occupied orbital, K up to its last row
          const TransformJob* J = reinterpret_cast<const TransformJob*>(jp);
          check_firsts(J, st.njobs, st.total);
          {
            // the XCD ranges: ascending, covering every unit once, each
            // within one unit's cost of an eighth of the total
            const XcdRanges r = transform_xcd_ranges(J, st.njobs, st.total);
            if (r.b[0] != 0 || r.b[8] != st.total) die("XCD ranges do not cover the units", r.b[0], r.b[8]);
            std::vector<int64_t> ucost((size_t)st.total);
            int64_t all = 0, umax = 0;
            for (int j = 0; j < st.njobs; j++) {
              const TransformArgs& a = J[j].a;
              const int64_t mth = (a.MT + 1) / 2;
              for (int64_t w2 = 0; w2 < a.nwu; w2++) {
                const int64_t c = transform_unit_stages(a, w2 % mth);
                ucost[(size_t)(J[j].first + w2)] = c;
                all += c;
                umax = std::max(umax, c);
              }
            }
            for (int x = 0; x < 8; x++) {
              if (r.b[x + 1] < r.b[x]) die("XCD ranges not ascending", x);
              int64_t c = 0;
              for (int64_t u = r.b[x]; u < r.b[x + 1]; u++) c += ucost[(size_t)u];
              if (std::fabs((double)c - all / 8.0) > (double)umax) die("XCD range cost off balance", x, c);
            }
          }
          for (int j = 0; j < st.njobs; j++) {
            const TransformArgs& a = J[j].a;
            if (job_end(J, st.njobs, j, st.total) - J[j].first != a.nwu) die("transform unit count", j);
            const double* Lpd = reinterpret_cast<const double*>(a.Lp);
            const double* Xpd = reinterpret_cast<const double*>(a.Xp);
            chk(a.Lp, (size_t)a.KS * a.MB * 512, "transform Lp slab");
            chk(a.Xp, (size_t)a.nocc * a.KS * a.NRB * 512, "transform Xp slabs");
            // work unit: row tiles t and MT-1-t of one column tile; every
            // (row tile, column tile) exactly once
            std::vector<int64_t> units;
            const int64_t mth = (a.MT + 1) / 2;
            for (int64_t w2 = 0; w2 < a.nwu; w2++) {
              const int64_t col = w2 / mth, t = w2 % mth;
              units.push_back(col * a.MT + t);
              if (a.MT - 1 - t != t) units.push_back(col * a.MT + a.MT - 1 - t);
            }
            std::vector<int> seen_u((size_t)(a.MT * a.NT * a.nocc), 0);
            for (int64_t wu : units)
              if (wu >= (int64_t)seen_u.size() || seen_u[wu]++) die("transform unit out of range or repeated", j, wu);
            if ((int64_t)units.size() != a.MT * a.NT * a.nocc) die("transform units missing", j);
            for (int64_t wu : units) {
              const int64_t mt = wu % a.MT, col = wu / a.MT, i = col / a.NT, nt = col % a.NT;
              const int64_t kend = std::min((mt + 1) * kTrTM, a.KS * 4);
              for (int64_t q = mt * kTrTM; q < std::min(a.naux, (mt + 1) * kTrTM); q++)
                for (int64_t v = nt * kTrTN; v < std::min(a.nvirt, (nt + 1) * kTrTN); v++) {
                  double acc = 0.0;
                  for (int64_t k = 0; k < kend; k++) {
                    const double l = rd(Lpd + ((k / 4) * a.MB + q / 16) * 64 + (k % 4) * 16 + q % 16, "Lp read");
                    const double x =
                        rd(Xpd + ((i * a.KS + k / 4) * a.NRB + v / 16) * 64 + (k % 4) * 16 + v % 16, "Xp read");
                    acc += l * x;
                  }
                  wr(a.B + q + a.naux * (v + a.nvirt * i), acc, "B");
                }
            }
          }
          break;
        }
      }
    }
    // compare with the direct computation
    for (size_t f = 0; f < F.size(); f++) {
      Frag& x = F[f];
      int want = ((int)f == bad_frag) ? (bad_kind == 1 ? 2 : bad_kind == 2 ? 4 : bad_kind == 3 ? 8 : 0) : 0;
      if (!transform && bad_kind == 2 && !x.fitted) want = 0;  // X is the int8 slicing's to check
      if ((x.status[0] & want) != want || (x.status[0] & ~want)) die("status word", (long long)f, x.status[0]);
      if (want) continue;
      const int64_t n = x.n, nc = x.nvirt * x.nocc;
      double bmax = 0.0, berr = 0.0;
      if (x.fitted) {
        for (size_t e = 0; e < x.B.size(); e++) berr = std::fmax(berr, std::fabs(x.B[e] - x.X[e]));
        if (berr != 0.0) die("fitted copy", (long long)f);
        continue;
      }
      std::vector<double> l(n * n, 0.0), li(n * n, 0.0);
      for (int64_t c = 0; c < n; c++) {
        double d = x.V[c + n * c];
        for (int64_t k = 0; k < c; k++) d -= l[c + n * k] * l[c + n * k];
        const double s = std::sqrt(d);
        l[c + n * c] = s;
        for (int64_t r = c + 1; r < n; r++) {
          double v = x.V[r + n * c];
          for (int64_t k = 0; k < c; k++) v -= l[r + n * k] * l[c + n * k];
          l[r + n * c] = v / s;
        }
      }
      for (int64_t col = 0; col < n; col++)
        for (int64_t c = col; c < n; c++) {
          double v = (c == col) ? 1.0 : 0.0;
          for (int64_t k = col; k < c; k++) v -= l[c + n * k] * li[k + n * col];
          li[c + n * col] = v / l[c + n * c];
        }
      size_t k = 0;
      for (; k < items.size(); k++)
        if (items[k].B == x.B.data()) break;
      const double* Ls = reinterpret_cast<const double*>(L.data() + g.lo[k]);
      const double* Lis = reinterpret_cast<const double*>(Li.data() + g.lo[k]);
      double lerr = 0.0, lierr = 0.0, limax = 0.0;
      for (int64_t c = 0; c < n; c++)
        for (int64_t r = 0; r < n; r++) {
          if (r >= c) lerr = std::fmax(lerr, std::fabs(Ls[r + n * c] - l[r + n * c]));
          lierr = std::fmax(lierr, std::fabs(Lis[r + n * c] - li[r + n * c]));
          limax = std::fmax(limax, std::fabs(li[r + n * c]));
        }
      for (int64_t c = 0; c < (transform ? nc : 0); c++)
        for (int64_t q = 0; q < n; q++) {
          double s = 0.0;
          for (int64_t p = 0; p <= q; p++) s += li[q + n * p] * x.X[p + n * c];
          bmax = std::fmax(bmax, std::fabs(s));
          berr = std::fmax(berr, std::fabs(x.B[q + n * c] - s));
        }
      if (!(lerr <= 1e-12 * 4.0) || !(lierr <= 1e-12 * limax) || !(berr <= 1e-12 * bmax)) {
        std::printf("FAIL fragment %zu n=%lld: L err %.3e, Linv err %.3e (max %.3e), B err %.3e (max %.3e)\n", f,
                    (long long)n, lerr, lierr, limax, berr, bmax);
        std::exit(1);
      }
    }
    g_regions.resize(nreg);
    return (int)steps.size();
  };
  const int launches = run_case(-1, 0);
  run_case(6, 1);   // n = 129: a pivot <= 0 in the first block
  run_case(3, 2);   // NaN in X
  run_case(8, 3);   // denominator >= 0
  run_case(nd, 2);  // NaN in the fitted copy
  // the plan without the transform (exact-slice modes): L and L^{-1} as above,
  // no packing or transform step, no access to the packed-operand arenas
  run_case(-1, 0, false);
  run_case(6, 1, false);
  run_case(3, 2, false);
  std::printf("OK %d\n", launches);
  return 0;
}
