// CPU oracle for the LS-MPS per-step particle pipeline.
//
// TEST INFRASTRUCTURE ONLY (the checker, never the product): only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
//
// A plain single-threaded FP64 restatement of the reference's src/lsmps.cpp.
// Built with -O3 -std=c++17 -ffp-contract=off so that every multiply and add
// rounds separately, as the reference's x86-64 build does. Each function
// names the reference line it follows. The data structures are deliberately
// the reference's kind (hash-map cells, per-particle neighbor vectors) so the
// timing of this oracle is a fair single-core CPU baseline.
#include "lsmps_oracle.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

struct P3 {
  double x, y, z;
};
inline P3 sub(const P3& a, const P3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
// Squared length, summed x then y then z (src/lsmps.cpp:21).
inline double len2(const P3& d) { return d.x * d.x + d.y * d.y + d.z * d.z; }

enum : uint8_t { T_INTERIOR = 0, T_SURFACE = 1, T_SPLASH = 2, T_NEARWALL = 3, T_WALL = 4 };
constexpr int NB = 9;  // second-order basis size in 3-D

// Kernel weight (re/r - 1)^2 inside (1e-12, re) (src/lsmps.cpp:22).
inline double kernel_weight(double r, double re) {
  if (r <= 1e-12 || r >= re) return 0.0;
  const double t = re / r - 1.0;
  return t * t;
}

// Scaled second-order monomials of the offset (src/lsmps.cpp:23).
inline void monomials(const P3& d, double r, double l0, double q[NB]) {
  const double inv = 1.0 / r;
  q[0] = d.x * inv;
  q[1] = d.y * inv;
  q[2] = d.z * inv;
  const double h = 2 * l0 * r, g = l0 * r;
  q[3] = d.x * d.x / h;
  q[4] = d.y * d.y / h;
  q[5] = d.z * d.z / h;
  q[6] = d.x * d.y / g;
  q[7] = d.x * d.z / g;
  q[8] = d.y * d.z / g;
}

// Lower Cholesky with the diagonal-shift ladder 0, reg, 10reg, 100reg, 1000reg
// (src/lsmps.cpp:24). On success a[] holds L (upper part zeroed).
bool chol_ladder(std::array<double, NB * NB>& a, double reg, bool& shifted) {
  shifted = false;
  for (int attempt = 0; attempt < 5; ++attempt) {
    std::array<double, NB * NB> m = a;
    const double add = attempt ? reg * std::pow(10.0, attempt - 1) : 0.0;
    if (add != 0.0) {
      shifted = true;
      for (int i = 0; i < NB; ++i) m[i * NB + i] += add;
    }
    bool good = true;
    for (int i = 0; i < NB && good; ++i) {
      for (int j = 0; j <= i; ++j) {
        double acc = m[i * NB + j];
        for (int k = 0; k < j; ++k) acc -= m[i * NB + k] * m[j * NB + k];
        if (i == j) {
          if (!(acc > 1e-14) || !std::isfinite(acc)) {
            good = false;
            break;
          }
          m[i * NB + i] = std::sqrt(acc);
        } else {
          m[i * NB + j] = acc / m[j * NB + j];
        }
      }
      for (int j = i + 1; j < NB; ++j) m[i * NB + j] = 0.0;
    }
    if (good) {
      a = m;
      return true;
    }
  }
  return false;
}

// L L^T x = b by forward then backward substitution (src/lsmps.cpp:25).
std::array<double, NB> chol_solve(const std::array<double, NB * NB>& L, const double* b) {
  std::array<double, NB> y{}, x{};
  for (int i = 0; i < NB; ++i) {
    double acc = b[i];
    for (int k = 0; k < i; ++k) acc -= L[i * NB + k] * y[k];
    y[i] = acc / L[i * NB + i];
  }
  for (int i = NB - 1; i >= 0; --i) {
    double acc = y[i];
    for (int k = i + 1; k < NB; ++k) acc -= L[k * NB + i] * x[k];
    x[i] = acc / L[i * NB + i];
  }
  return x;
}

// The manufactured quadratic of the reference's WLS test and its gradient
// (src/lsmps.cpp:28). Laplacian = 2*(0.7 - 0.2 + 0.4) = 1.8.
inline double test_field(const P3& r) {
  return 1 + 2 * r.x - 3 * r.y + .5 * r.z + .7 * r.x * r.x - .2 * r.y * r.y + .4 * r.z * r.z +
         .3 * r.x * r.y - .1 * r.x * r.z + .25 * r.y * r.z;
}
inline P3 test_grad(const P3& r) {
  return {2 + 1.4 * r.x + .3 * r.y - .1 * r.z, -3 - .4 * r.y + .3 * r.x + .25 * r.z,
          .5 + .8 * r.z - .1 * r.x + .25 * r.y};
}

struct CellKey {
  int x, y, z;
  bool operator==(const CellKey& o) const { return x == o.x && y == o.y && z == o.z; }
};
struct CellKeyHash {
  size_t operator()(const CellKey& c) const {
    uint64_t h = 0x9E3779B97F4A7C15ull;
    h ^= (uint32_t)c.x + 0x9E3779B9u + (h << 6) + (h >> 2);
    h ^= (uint32_t)c.y + 0x9E3779B9u + (h << 6) + (h >> 2);
    h ^= (uint32_t)c.z + 0x9E3779B9u + (h << 6) + (h >> 2);
    return (size_t)h;
  }
};

struct Csr {
  std::vector<int> row, col;
  std::vector<double> val, b, xe;
};

}  // namespace

struct orc_state {
  std::vector<P3> r;
  std::vector<uint8_t> type;
  std::vector<std::vector<int>> nbr;
  uint64_t pairs = 0, truncated = 0;
  Csr A;
  std::vector<int> rev;
};

namespace {

// One lattice block, k (z) outermost, then j (y), then i (x); three jitter
// draws x, y, z per particle; bottom layer typed Wall (src/lsmps.cpp:68).
void fill_lattice(std::vector<P3>& r, std::vector<uint8_t>& t, size_t target, int nx, int ny, int nz,
                  double l0, double jit, std::mt19937& rng, bool bottom) {
  std::uniform_real_distribution<double> u(-jit, jit);
  for (int k = 0; k < nz && r.size() < target; ++k)
    for (int j = 0; j < ny && r.size() < target; ++j)
      for (int i = 0; i < nx && r.size() < target; ++i) {
        const double ox = u(rng);
        const double oy = u(rng);
        const double oz = u(rng);
        r.push_back({(i + .5 + ox) * l0, (j + .5 + oy) * l0, (k + .5 + oz) * l0});
        t.push_back(bottom && k == 0 ? T_WALL : T_INTERIOR);
      }
}

// Case generators of the reference (src/lsmps.cpp:72).
void generate(const std::string& name, size_t n, uint32_t seed, double l0, int film_layers,
              std::vector<P3>& r, std::vector<uint8_t>& t) {
  std::mt19937 rng(seed);
  r.clear();
  t.clear();
  r.reserve(n);
  t.reserve(n);
  const int side = std::max(4, (int)std::ceil(std::cbrt((double)n)));
  auto plane_side = [](size_t count, int layers) {
    return std::max(8, (int)std::ceil(std::sqrt((double)count / layers)));
  };
  if (name == "wall_film") {
    const int nz = std::max(3, film_layers + 1);
    const int nx = plane_side(n, nz);
    fill_lattice(r, t, n, nx, nx, nz, l0, .04, rng, true);
  } else if (name == "narrow_gap") {
    const int nz = std::max(4, film_layers + 2);
    const int nx = plane_side(n, nz);
    fill_lattice(r, t, n, nx, nx, nz, l0, .02, rng, true);
    const double top = (nz - 1.5) * l0;
    for (size_t i = 0; i < r.size(); ++i)
      if (r[i].z > top) t[i] = T_WALL;
  } else if (name == "rain_injection") {
    const size_t film = (size_t)(.75 * n);
    const int nz = std::max(3, film_layers + 1);
    const int nx = plane_side(film, nz);
    fill_lattice(r, t, film, nx, nx, nz, l0, .05, rng, true);
    std::uniform_real_distribution<double> ux(0, nx * l0), uy(0, nx * l0), uz(4 * l0, 20 * l0);
    while (r.size() < n) {
      const double a = ux(rng);
      const double b = uy(rng);
      const double c = uz(rng);
      r.push_back({a, b, c});
      t.push_back(T_SPLASH);
    }
  } else {
    fill_lattice(r, t, n, side, side, side, l0, .03, rng, false);
  }
}

int cell_of(double x, double re) { return (int)std::floor(x / re); }

// Hash-grid cell list; for each particle visit the 27 cells z-major, then y,
// then x, each cell's members in ascending index, keep the first maxn hits
// of the strict test |rj - ri|^2 < re^2 and count the rest as truncated
// (src/lsmps.cpp:26).
void cell_list(orc_state& s, double re, int maxn) {
  const size_t n = s.r.size();
  std::unordered_map<CellKey, std::vector<int>, CellKeyHash> grid;
  grid.reserve(n * 2);
  std::vector<CellKey> key(n);
  for (size_t i = 0; i < n; ++i) {
    key[i] = {cell_of(s.r[i].x, re), cell_of(s.r[i].y, re), cell_of(s.r[i].z, re)};
    grid[key[i]].push_back((int)i);
  }
  const double re2 = re * re;
  s.nbr.assign(n, {});
  s.pairs = 0;
  s.truncated = 0;
  for (size_t i = 0; i < n; ++i) {
    std::vector<int>& out = s.nbr[i];
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          auto it = grid.find({key[i].x + dx, key[i].y + dy, key[i].z + dz});
          if (it == grid.end()) continue;
          for (int j : it->second) {
            if ((size_t)j == i) continue;
            if (!(len2(sub(s.r[j], s.r[i])) < re2)) continue;
            if ((int)out.size() < maxn)
              out.push_back(j);
            else
              s.truncated++;
          }
        }
    s.pairs += out.size();
  }
}

// Type of one particle from its own type, its (truncated) neighbor list and
// the 14 direction sectors it sees (src/lsmps.cpp:27).
uint8_t classify_one(const orc_state& s, size_t i, double l0) {
  if (s.type[i] == T_WALL) return T_WALL;
  const std::vector<int>& nb = s.nbr[i];
  if (nb.size() < 10) return T_SPLASH;
  unsigned mask = 0;
  for (int j : nb) {
    const P3 d = sub(s.r[j], s.r[i]);
    const double ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    int axis;
    if (ax >= ay && ax >= az)
      axis = d.x >= 0 ? 0 : 1;
    else if (ay >= ax && ay >= az)
      axis = d.y >= 0 ? 2 : 3;
    else
      axis = d.z >= 0 ? 4 : 5;
    const int oct = 6 + (d.x >= 0 ? 1 : 0) + (d.y >= 0 ? 2 : 0) + (d.z >= 0 ? 4 : 0);
    mask |= (1u << axis) | (1u << oct);
  }
  const int sectors = __builtin_popcount(mask);
  if (s.r[i].z < 1.5 * l0) return T_NEARWALL;
  if (sectors < 11) return T_SURFACE;
  return T_INTERIOR;
}

// Moment matrix (all 81 entries, as the reference accumulates them) and RHS
// of particle i for field values f, divided by the weight sum
// (src/lsmps.cpp:28).
double moments(const orc_state& s, const double* f, double re, double l0, size_t i,
               std::array<double, NB * NB>& M, double rhs[NB]) {
  M.fill(0.0);
  for (int a = 0; a < NB; ++a) rhs[a] = 0.0;
  double wsum = 0.0;
  const double fi = f[i];
  for (int j : s.nbr[i]) {
    const P3 d = sub(s.r[j], s.r[i]);
    const double r = std::sqrt(len2(d));
    const double w = kernel_weight(r, re);
    if (w == 0.0) continue;
    double q[NB];
    monomials(d, r, l0, q);
    const double df = f[j] - fi;
    wsum += w;
    for (int a = 0; a < NB; ++a) {
      const double wq = w * q[a];
      rhs[a] += wq * df / r;
      for (int c = 0; c < NB; ++c) M[a * NB + c] += wq * q[c];
    }
  }
  if (wsum > 0) {
    for (double& v : M) v /= wsum;
    for (int a = 0; a < NB; ++a) rhs[a] /= wsum;
  }
  return wsum;
}

// Solves one particle: status + grad + Laplacian (src/lsmps.cpp:24-25,28).
uint8_t wls_one(const orc_state& s, const double* f, double re, double l0, double reg, size_t i,
                double g[3], double* lap) {
  std::array<double, NB * NB> M;
  double rhs[NB];
  const double wsum = moments(s, f, re, l0, i, M, rhs);
  g[0] = g[1] = g[2] = 0.0;
  *lap = 0.0;
  if (wsum <= 0) return ORC_WLS_NO_WEIGHT;
  bool shifted = false;
  if (!chol_ladder(M, reg, shifted)) return ORC_WLS_FAILED;
  const std::array<double, NB> x = chol_solve(M, rhs);
  g[0] = x[0];
  g[1] = x[1];
  g[2] = x[2];
  *lap = (x[3] + x[4] + x[5]) / l0;
  return shifted ? ORC_WLS_REGULARIZED : ORC_WLS_OK;
}

// Sampled quadratic-reproduction test (src/lsmps.cpp:28).
orc_wls_stats wls_sampled(const orc_state& s, double re, double l0, double reg, int samples) {
  orc_wls_stats st{};
  const size_t n = s.r.size();
  const size_t stride = std::max<size_t>(1, n / (size_t)std::max(1, samples));
  std::vector<double> f(n);
  for (size_t i = 0; i < n; ++i) f[i] = test_field(s.r[i]);
  double sse_g = 0, sse_l = 0;
  for (size_t i = 0; i < n && st.sampled < (uint64_t)samples; i += stride) {
    if (s.type[i] == T_WALL || s.type[i] == T_SPLASH) continue;
    std::array<double, NB * NB> M;
    double rhs[NB];
    const double wsum = moments(s, f.data(), re, l0, i, M, rhs);
    if (wsum <= 0) {
      st.failed++;
      continue;
    }
    if (s.type[i] == T_NEARWALL || s.type[i] == T_SURFACE) st.virtual_points++;
    bool shifted = false;
    if (!chol_ladder(M, reg, shifted)) {
      st.failed++;
      continue;
    }
    const std::array<double, NB> x = chol_solve(M, rhs);
    const P3 e = test_grad(s.r[i]);
    const double ge = std::sqrt(len2({x[0] - e.x, x[1] - e.y, x[2] - e.z}));
    const double le = std::fabs((x[3] + x[4] + x[5]) / l0 - 1.8);
    st.sampled++;
    if (shifted)
      st.regularized++;
    else
      st.ok++;
    st.max_gradient_error = std::max(st.max_gradient_error, ge);
    st.max_laplacian_error = std::max(st.max_laplacian_error, le);
    sse_g += ge * ge;
    sse_l += le * le;
  }
  if (st.sampled) {
    st.rms_gradient_error = std::sqrt(sse_g / st.sampled);
    st.rms_laplacian_error = std::sqrt(sse_l / st.sampled);
  }
  return st;
}

// Seeded manufactured solution U(-1, 1) for row k (no reference code): the
// splitmix64 finaliser (Steele, Lea, Flood 2014) of seed * 2^64/phi + k, top
// 53 bits scaled to [0, 1), mapped to [-1, 1).
double hashed_uniform(uint32_t seed, uint64_t k) {
  uint64_t z = (uint64_t)seed * 0x9E3779B97F4A7C15ull + k;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return -1.0 + 2.0 * ((double)(z >> 11) * 0x1.0p-53);
}

// Graph-Laplacian PPE over the first `cap` Interior/NearWall particles with
// row 0 pinned to the identity and b = A x* (src/lsmps.cpp:29-64).
void assemble(orc_state& s, double re, int max_unknowns, int xmode, uint32_t xseed, double shift,
              orc_ppe_stats& st) {
  const auto t0 = Clock::now();
  const size_t n = s.r.size();
  const size_t cap = std::min(n, (size_t)std::max(0, max_unknowns));
  std::vector<int> map(n, -1);
  s.rev.clear();
  for (size_t i = 0; i < n && s.rev.size() < cap; ++i)
    if (s.type[i] == T_INTERIOR || s.type[i] == T_NEARWALL) {
      map[i] = (int)s.rev.size();
      s.rev.push_back((int)i);
    }
  const size_t m = s.rev.size();
  Csr& A = s.A;
  A = Csr{};
  A.row.push_back(0);
  A.b.assign(m, 0.0);
  A.xe.resize(m);
  for (size_t k = 0; k < m; ++k) {
    const P3& p = s.r[s.rev[k]];
    A.xe[k] = xmode == 0 ? p.x * p.x + p.y * p.y : hashed_uniform(xseed, k);
  }
  for (size_t k = 0; k < m; ++k) {
    const int i = s.rev[k];
    if (k == 0) {
      A.col.push_back(0);
      A.val.push_back(1.0);
      A.row.push_back((int)A.col.size());
      continue;
    }
    double diag = 0.0;
    for (int j : s.nbr[i]) {
      const int cj = map[j];
      if (cj < 0) continue;
      const P3 d = sub(s.r[j], s.r[i]);
      const double r = std::sqrt(len2(d));
      const double w = kernel_weight(r, re);
      if (w <= 0) continue;
      const double a = w / (r * r);
      diag += a;
      A.col.push_back(cj);
      A.val.push_back(-a);
    }
    if (diag <= 0 || !std::isfinite(diag)) {
      st.failed_rows++;
      diag = 1.0;
    }
    if (shift != 0.0) diag += shift * diag;
    A.col.push_back((int)k);
    A.val.push_back(diag);
    A.row.push_back((int)A.col.size());
    if (s.type[i] == T_NEARWALL) st.neumann++;
  }
  for (size_t k = 0; k < m; ++k) {
    double acc = 0.0;
    for (int e = A.row[k]; e < A.row[k + 1]; ++e) acc += A.val[e] * A.xe[A.col[e]];
    A.b[k] = acc;
  }
  st.unknowns = m;
  st.nnz = A.val.size();
  st.dirichlet = n - m;
  st.assembly_ms = ms_since(t0);
}

// WLS-Laplacian PPE rows. No C++ reference: the reference plans them
// (README.md:34) and its 2-D model builds them (python/rain_plate_lsmps_cpu.py:
// 204-253: row = w * (lap_row . basis(off)) per neighbor, diagonal minus their
// sum, free-surface/splash neighbors at p = 0, invalid rows the identity).
// Here in 3-D with the reference's weight, basis and ladder (src/lsmps.cpp:
// 22-25) and wls_test's normalisation (:28): the WLS Laplacian at i is
// (x3 + x4 + x5) / l0 with x = M^-1 sum_j w q_j (f_j - f_i) / (r_j ws), so
// c_ij = (z . q_j) w_j / r_j / ws / l0 with z = M^-1 (0,0,0,1,1,1,0,0,0).
// Row i is minus the operator (the graph Laplacian's sign): -c_ij per unknown
// neighbor, d = sum_j c_ij over all weighted neighbors on the diagonal. A row
// whose WLS fails, or with d <= 0 or d < sum |c_ij| over the unknown
// neighbors (a one-sided stencil), takes the graph-Laplacian row of
// assemble() instead (src/lsmps.cpp:41-51). Unknowns, x* and b = A x* as in
// assemble(); row 0 is not pinned.
void assemble_wls(orc_state& s, double re, double l0, double reg, int max_unknowns, int xmode, uint32_t xseed,
                  double shift, orc_ppe_stats& st) {
  const auto t0 = Clock::now();
  const size_t n = s.r.size();
  const size_t cap = std::min(n, (size_t)std::max(0, max_unknowns));
  std::vector<int> map(n, -1);
  s.rev.clear();
  for (size_t i = 0; i < n && s.rev.size() < cap; ++i)
    if (s.type[i] == T_INTERIOR || s.type[i] == T_NEARWALL) {
      map[i] = (int)s.rev.size();
      s.rev.push_back((int)i);
    }
  const size_t m = s.rev.size();
  Csr& A = s.A;
  A = Csr{};
  A.row.push_back(0);
  A.b.assign(m, 0.0);
  A.xe.resize(m);
  for (size_t k = 0; k < m; ++k) {
    const P3& p = s.r[s.rev[k]];
    A.xe[k] = xmode == 0 ? p.x * p.x + p.y * p.y : hashed_uniform(xseed, k);
  }
  const double e3[NB] = {0, 0, 0, 1, 1, 1, 0, 0, 0};
  for (size_t k = 0; k < m; ++k) {
    const int i = s.rev[k];
    const size_t start = A.col.size();
    // Moments with the reference's normalisation (the right-hand side is not needed).
    std::array<double, NB * NB> M{};
    double wsum = 0.0;
    for (int j : s.nbr[i]) {
      const P3 d = sub(s.r[j], s.r[i]);
      const double r = std::sqrt(len2(d));
      const double w = kernel_weight(r, re);
      if (w == 0.0) continue;
      double q[NB];
      monomials(d, r, l0, q);
      wsum += w;
      for (int a = 0; a < NB; ++a) {
        const double wq = w * q[a];
        for (int c = 0; c < NB; ++c) M[a * NB + c] += wq * q[c];
      }
    }
    bool ok = wsum > 0;
    double diag = 0.0, offabs = 0.0;
    if (ok) {
      for (double& v : M) v /= wsum;
      bool shifted = false;
      ok = chol_ladder(M, reg, shifted);
    }
    if (ok) {
      const std::array<double, NB> z = chol_solve(M, e3);
      for (int j : s.nbr[i]) {
        const P3 d = sub(s.r[j], s.r[i]);
        const double r = std::sqrt(len2(d));
        const double w = kernel_weight(r, re);
        if (w == 0.0) continue;
        double q[NB];
        monomials(d, r, l0, q);
        double dot = 0.0;
        for (int a = 0; a < NB; ++a) dot += z[a] * q[a];
        const double c = dot * w / r / wsum / l0;
        diag += c;
        if (map[j] >= 0) {
          A.col.push_back(map[j]);
          A.val.push_back(-c);
          offabs += std::fabs(c);
        }
      }
      ok = std::isfinite(diag) && diag > 0.0 && diag >= offabs;
    }
    if (!ok) {
      // The graph-Laplacian row, as assemble() builds it.
      A.col.resize(start);
      A.val.resize(start);
      st.graph_rows++;
      diag = 0.0;
      for (int j : s.nbr[i]) {
        const int cj = map[j];
        if (cj < 0) continue;
        const P3 d = sub(s.r[j], s.r[i]);
        const double r = std::sqrt(len2(d));
        const double w = kernel_weight(r, re);
        if (w <= 0) continue;
        const double a = w / (r * r);
        diag += a;
        A.col.push_back(cj);
        A.val.push_back(-a);
      }
      if (diag <= 0 || !std::isfinite(diag)) {
        st.failed_rows++;
        diag = 1.0;
      }
    }
    if (shift != 0.0) diag += shift * diag;
    A.col.push_back((int)k);
    A.val.push_back(diag);
    A.row.push_back((int)A.col.size());
    if (s.type[i] == T_NEARWALL) st.neumann++;
  }
  for (size_t k = 0; k < m; ++k) {
    double acc = 0.0;
    for (int e = A.row[k]; e < A.row[k + 1]; ++e) acc += A.val[e] * A.xe[A.col[e]];
    A.b[k] = acc;
  }
  st.unknowns = m;
  st.nnz = A.val.size();
  st.dirichlet = n - m;
  st.assembly_ms = ms_since(t0);
}

// y = A x, row by row in stored order (src/lsmps.cpp:65).
void csr_mul(const Csr& A, const std::vector<double>& x, std::vector<double>& y) {
  const size_t m = A.b.size();
  y.assign(m, 0.0);
  for (size_t i = 0; i < m; ++i)
    for (int e = A.row[i]; e < A.row[i + 1]; ++e) y[i] += A.val[e] * x[A.col[e]];
}
// Left-to-right dot product (src/lsmps.cpp:66).
double dotp(const std::vector<double>& a, const std::vector<double>& b) {
  double acc = 0.0;
  for (size_t i = 0; i < a.size(); ++i) acc += a[i] * b[i];
  return acc;
}

// BiCGStab of van der Vorst with the reference's breakdown guards and
// stopping rule ||.|| <= tol*max(1,||b||) (src/lsmps.cpp:67). With
// jacobi=true it is right-preconditioned by M = diag(A): p^ = M^-1 p,
// s^ = M^-1 s, so r and s stay true residuals of A x = b.
void bicgstab(const orc_state& s, int max_iter, double tol, bool jacobi, const double* x0,
              orc_ppe_stats& st, std::vector<double>& x) {
  const auto t0 = Clock::now();
  const Csr& A = s.A;
  const size_t m = A.b.size();
  x.assign(m, 0.0);
  if (!m) return;
  std::vector<double> dinv(m, 1.0);
  if (jacobi)
    for (size_t i = 0; i < m; ++i)
      for (int e = A.row[i]; e < A.row[i + 1]; ++e)
        if ((size_t)A.col[e] == i) dinv[i] = 1.0 / A.val[e];
  std::vector<double> r = A.b, Ax;
  if (x0) {
    for (size_t i = 0; i < m; ++i) x[i] = x0[i];
    csr_mul(A, x, Ax);
    for (size_t i = 0; i < m; ++i) r[i] = A.b[i] - Ax[i];
  }
  std::vector<double> r0 = r, p(m, 0.0), v(m, 0.0), sv(m), t(m), ph(m), sh(m);
  const double bnorm = std::sqrt(dotp(A.b, A.b));
  st.initial_residual = std::sqrt(dotp(r, r));
  const double stop = tol * std::max(1.0, bnorm);
  double rho_prev = 1, alpha = 1, omega = 1;
  st.iterations = 0;
  st.converged = 0;
  for (int it = 0; it < max_iter; ++it) {
    const double rho = dotp(r0, r);
    if (std::fabs(rho) < 1e-30) break;
    const double beta = (rho / rho_prev) * (alpha / omega);
    for (size_t i = 0; i < m; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
    for (size_t i = 0; i < m; ++i) ph[i] = jacobi ? dinv[i] * p[i] : p[i];
    csr_mul(A, ph, v);
    const double den = dotp(r0, v);
    if (std::fabs(den) < 1e-30) break;
    alpha = rho / den;
    for (size_t i = 0; i < m; ++i) sv[i] = r[i] - alpha * v[i];
    if (std::sqrt(dotp(sv, sv)) <= stop) {
      for (size_t i = 0; i < m; ++i) x[i] += alpha * ph[i];
      st.iterations = it + 1;
      st.converged = 1;
      break;
    }
    for (size_t i = 0; i < m; ++i) sh[i] = jacobi ? dinv[i] * sv[i] : sv[i];
    csr_mul(A, sh, t);
    const double tt = dotp(t, t);
    if (tt < 1e-30) break;
    omega = dotp(t, sv) / tt;
    for (size_t i = 0; i < m; ++i) x[i] += alpha * ph[i] + omega * sh[i];
    for (size_t i = 0; i < m; ++i) r[i] = sv[i] - omega * t[i];
    st.iterations = it + 1;
    if (std::sqrt(dotp(r, r)) <= stop) {
      st.converged = 1;
      break;
    }
    if (std::fabs(omega) < 1e-30) break;
    rho_prev = rho;
  }
  csr_mul(A, x, Ax);
  double res = 0, se = 0, mx = 0, xn = 0;
  for (size_t i = 0; i < m; ++i) {
    const double d = Ax[i] - A.b[i];
    res += d * d;
    const double e = x[i] - A.xe[i];
    se += e * e;
    xn += A.xe[i] * A.xe[i];
    mx = std::max(mx, std::fabs(e));
  }
  st.final_residual = std::sqrt(res);
  st.relative_l2_error = std::sqrt(se / std::max(xn, 1e-30));
  st.max_error = mx;
  st.solve_ms = ms_since(t0);
}

// ---- MLP 16 -> 64 -> 64 -> 2 (tanh), fp32 --------------------------------
constexpr int F0 = 16, H1 = 64, H2 = 64, F3 = 2;
constexpr size_t O_W1 = 0, O_B1 = O_W1 + H1 * F0, O_W2 = O_B1 + H1, O_B2 = O_W2 + H2 * H1,
                 O_W3 = O_B2 + H2, O_B3 = O_W3 + F3 * H2, N_PARAMS = O_B3 + F3;

void mlp_row(const float* P, const float* x, float h1[H1], float h2[H2], float y[F3]) {
  for (int o = 0; o < H1; ++o) {
    float a = P[O_B1 + o];
    for (int k = 0; k < F0; ++k) a += P[O_W1 + o * F0 + k] * x[k];
    h1[o] = std::tanh(a);
  }
  for (int o = 0; o < H2; ++o) {
    float a = P[O_B2 + o];
    for (int k = 0; k < H1; ++k) a += P[O_W2 + o * H1 + k] * h1[k];
    h2[o] = std::tanh(a);
  }
  for (int o = 0; o < F3; ++o) {
    float a = P[O_B3 + o];
    for (int k = 0; k < H2; ++k) a += P[O_W3 + o * H2 + k] * h2[k];
    y[o] = a;
  }
}

}  // namespace

extern "C" {

uint64_t orc_generate_case(const char* case_name, uint64_t particles, uint32_t seed, double l0,
                           int film_layers, double* xyz, uint8_t* type) {
  std::vector<P3> r;
  std::vector<uint8_t> t;
  generate(case_name ? case_name : "", (size_t)particles, seed, l0, film_layers, r, t);
  for (size_t i = 0; i < r.size(); ++i) {
    if (xyz) {
      xyz[3 * i] = r[i].x;
      xyz[3 * i + 1] = r[i].y;
      xyz[3 * i + 2] = r[i].z;
    }
    if (type) type[i] = t[i];
  }
  return r.size();
}

orc_state* orc_new(const double* xyz, const uint8_t* type, uint64_t n) {
  orc_state* s = new orc_state;
  s->r.resize(n);
  s->type.resize(n);
  for (uint64_t i = 0; i < n; ++i) {
    s->r[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    s->type[i] = type ? type[i] : (uint8_t)T_INTERIOR;
  }
  return s;
}
void orc_free(orc_state* s) { delete s; }

void orc_neighbors(orc_state* s, double re, int maxn, uint64_t* pairs, uint64_t* truncated) {
  cell_list(*s, re, maxn);
  if (pairs) *pairs = s->pairs;
  if (truncated) *truncated = s->truncated;
}

void orc_neighbors_brute(orc_state* s, double re, int maxn, uint64_t* pairs, uint64_t* truncated) {
  // Candidates visited in the same order as the cell walk: cell (z, y, x)
  // offset first, then ascending index, so truncation keeps the same subset.
  const size_t n = s->r.size();
  const double re2 = re * re;
  s->nbr.assign(n, {});
  s->pairs = s->truncated = 0;
  std::vector<std::array<int, 3>> c(n);
  for (size_t i = 0; i < n; ++i) c[i] = {cell_of(s->r[i].x, re), cell_of(s->r[i].y, re), cell_of(s->r[i].z, re)};
  for (size_t i = 0; i < n; ++i) {
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx)
          for (size_t j = 0; j < n; ++j) {
            if (j == i) continue;
            if (c[j][0] != c[i][0] + dx || c[j][1] != c[i][1] + dy || c[j][2] != c[i][2] + dz) continue;
            if (!(len2(sub(s->r[j], s->r[i])) < re2)) continue;
            if ((int)s->nbr[i].size() < maxn)
              s->nbr[i].push_back((int)j);
            else
              s->truncated++;
          }
    s->pairs += s->nbr[i].size();
  }
  if (pairs) *pairs = s->pairs;
  if (truncated) *truncated = s->truncated;
}

void orc_get_neighbors(const orc_state* s, int32_t* row_ptr, int32_t* col) {
  int64_t at = 0;
  row_ptr[0] = 0;
  for (size_t i = 0; i < s->nbr.size(); ++i) {
    for (int j : s->nbr[i]) col[at++] = j;
    row_ptr[i + 1] = (int32_t)at;
  }
}

void orc_set_neighbors(orc_state* s, const int32_t* row_ptr, const int32_t* col) {
  const size_t n = s->r.size();
  s->nbr.assign(n, {});
  s->pairs = 0;
  for (size_t i = 0; i < n; ++i) {
    s->nbr[i].assign(col + row_ptr[i], col + row_ptr[i + 1]);
    s->pairs += s->nbr[i].size();
  }
}

void orc_classify(orc_state* s, double l0) {
  std::vector<uint8_t> out(s->r.size());
  for (size_t i = 0; i < s->r.size(); ++i) out[i] = classify_one(*s, i, l0);
  s->type.swap(out);
}
void orc_get_types(const orc_state* s, uint8_t* type) { std::memcpy(type, s->type.data(), s->type.size()); }
double orc_hashed_uniform(uint32_t seed, uint64_t k) { return hashed_uniform(seed, k); }
void orc_set_types(orc_state* s, const uint8_t* type) { std::memcpy(s->type.data(), type, s->type.size()); }

void orc_wls_all(const orc_state* s, const double* f, double re, double l0, double reg, double* grad3,
                 double* lap, uint8_t* status) {
  for (size_t i = 0; i < s->r.size(); ++i) {
    double g[3] = {0, 0, 0}, L = 0;
    uint8_t st = ORC_WLS_SKIPPED;
    if (s->type[i] != T_WALL && s->type[i] != T_SPLASH) st = wls_one(*s, f, re, l0, reg, i, g, &L);
    grad3[3 * i] = g[0];
    grad3[3 * i + 1] = g[1];
    grad3[3 * i + 2] = g[2];
    lap[i] = L;
    status[i] = st;
  }
}

double orc_wls_moments(const orc_state* s, const double* f, double re, double l0, uint64_t i, double* m81,
                       double* b9) {
  std::array<double, NB * NB> M;
  const double w = moments(*s, f, re, l0, (size_t)i, M, b9);
  std::memcpy(m81, M.data(), sizeof(double) * NB * NB);
  return w;
}

void orc_wls_test(const orc_state* s, double re, double l0, double reg, int samples, orc_wls_stats* out) {
  *out = wls_sampled(*s, re, l0, reg, samples);
}

void orc_assemble_ppe(orc_state* s, double re, int max_unknowns, int xmode, uint32_t xseed, double shift,
                      orc_ppe_stats* out) {
  orc_ppe_stats st{};
  assemble(*s, re, max_unknowns, xmode, xseed, shift, st);
  *out = st;
}

void orc_assemble_ppe_wls(orc_state* s, double re, double l0, double reg, int max_unknowns, int xmode,
                          uint32_t xseed, double shift, orc_ppe_stats* out) {
  orc_ppe_stats st{};
  assemble_wls(*s, re, l0, reg, max_unknowns, xmode, xseed, shift, st);
  *out = st;
}

void orc_get_ppe(const orc_state* s, int32_t* row, int32_t* col, double* val, double* b, double* x_exact,
                 int32_t* rev) {
  const Csr& A = s->A;
  if (row) std::copy(A.row.begin(), A.row.end(), row);
  if (col) std::copy(A.col.begin(), A.col.end(), col);
  if (val) std::copy(A.val.begin(), A.val.end(), val);
  if (b) std::copy(A.b.begin(), A.b.end(), b);
  if (x_exact) std::copy(A.xe.begin(), A.xe.end(), x_exact);
  if (rev) std::copy(s->rev.begin(), s->rev.end(), rev);
}

// Features of the learned preconditioner, row by row in reference order
// (definitions in lsmps_oracle.h); plain sequential FP64 sums.
void orc_features(const orc_state* s, double re, float* feat, double* rhs_scale) {
  const Csr& A = s->A;
  const size_t m = s->rev.size();
  std::vector<int> map(s->r.size(), -1);
  for (size_t k = 0; k < m; ++k) map[s->rev[k]] = (int)k;
  std::vector<double> d(m);
  for (size_t k = 0; k < m; ++k) d[k] = A.val[A.row[k + 1] - 1];  // diagonal is stored last
  double ss = 0.0;
  for (size_t k = 0; k < m; ++k) ss += (A.b[k] / d[k]) * (A.b[k] / d[k]);
  double sb = m ? std::sqrt(ss / (double)m) : 0.0;
  if (!(sb > 0.0) || !std::isfinite(sb)) sb = 1.0;
  if (rhs_scale) *rhs_scale = sb;
  for (size_t k = 0; k < m; ++k) {
    const int i = s->rev[k];
    const std::vector<int>& nb = s->nbr[i];
    const size_t kn = nb.size();
    double sr = 0, sr2 = 0, rmin = 1e300, wu = 0, wall = 0, nbs = 0;
    int nu = 0, nw = 0;
    for (int j : nb) {
      const double dx = s->r[j].x - s->r[i].x, dy = s->r[j].y - s->r[i].y, dz = s->r[j].z - s->r[i].z;
      const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
      const double w = kernel_weight(r, re), rn = r / re;
      sr += rn;
      sr2 += rn * rn;
      rmin = std::min(rmin, rn);
      const double a = w > 0 ? w / (r * r) : 0.0;
      wall += a;
      nw += s->type[j] == T_WALL ? 1 : 0;
      if (map[j] >= 0 && w > 0) {
        wu += a;
        ++nu;
        nbs += a * (A.b[map[j]] / d[map[j]]);
      }
    }
    const double ik = kn ? 1.0 / (double)kn : 0.0;
    const bool pinned = k == 0 || (A.row[k + 1] - A.row[k] == 1 && d[k] == 1.0);
    float* f = feat + k * 16;
    for (int t = 0; t < 5; ++t) f[t] = s->type[i] == t ? 1.0f : 0.0f;
    f[5] = (float)(kn / 32.0);
    f[6] = (float)(sr * ik);
    f[7] = (float)(sr2 * ik);
    f[8] = (float)(kn ? rmin : 0.0);
    f[9] = (float)(d[k] * re * re / 32.0);
    f[10] = (float)(wall > 0 ? wu / wall : 0.0);
    f[11] = (float)(nu * ik);
    f[12] = (float)(A.b[k] / d[k] / sb);
    f[13] = pinned ? 0.0f : (float)(nbs / d[k] / sb);
    f[14] = (float)(nw * ik);
    f[15] = pinned ? 1.0f : 0.0f;
  }
}

void orc_solve_ppe(orc_state* s, int max_iter, double tol, int precond, const double* x0, orc_ppe_stats* out,
                   double* x) {
  std::vector<double> sol;
  bicgstab(*s, max_iter, tol, precond == 1, x0, *out, sol);
  if (x) std::copy(sol.begin(), sol.end(), x);
}

void orc_run(orc_state* s, double l0, double re_ratio, int maxn, int wls_samples, int max_unknowns, int max_iter,
             double tol, int do_solve, orc_run_stats* rs, orc_wls_stats* ws, orc_ppe_stats* ps) {
  const double re = re_ratio * l0;
  const size_t n = s->r.size();
  *rs = orc_run_stats{};
  rs->particles = n;
  auto t0 = Clock::now();
  cell_list(*s, re, maxn);
  rs->neighbor_ms = ms_since(t0);
  t0 = Clock::now();
  orc_classify(s, l0);
  rs->classify_ms = ms_since(t0);
  t0 = Clock::now();
  *ws = wls_sampled(*s, re, l0, 1e-8, wls_samples);
  rs->wls_ms = ms_since(t0);
  {
    std::vector<double> f(n), g(3 * n), L(n);
    std::vector<uint8_t> st(n);
    for (size_t i = 0; i < n; ++i) f[i] = test_field(s->r[i]);
    t0 = Clock::now();
    orc_wls_all(s, f.data(), re, l0, 1e-8, g.data(), L.data(), st.data());
    rs->wls_all_ms = ms_since(t0);
  }
  *ps = orc_ppe_stats{};
  assemble(*s, re, max_unknowns, 0, 0, 0.0, *ps);
  if (do_solve) {
    std::vector<double> x;
    bicgstab(*s, max_iter, tol, false, nullptr, *ps, x);
  }
  rs->pairs = s->pairs;
  rs->truncated = s->truncated;
  std::vector<int> cnt(n);
  for (size_t i = 0; i < n; ++i) cnt[i] = (int)s->nbr[i].size();
  if (n) {
    rs->neighbor_mean = std::accumulate(cnt.begin(), cnt.end(), 0.0) / n;
    std::sort(cnt.begin(), cnt.end());
    auto q = [&](double x) { return cnt[(size_t)(x * (n - 1))]; };
    rs->p50 = q(.5);
    rs->p90 = q(.9);
    rs->p99 = q(.99);
  }
  for (size_t i = 0; i < n; ++i) rs->type_count[s->type[i]]++;
}

void orc_mlp_forward(const float* P, const float* feat, uint64_t n, float* out) {
  float h1[H1], h2[H2];
  for (uint64_t i = 0; i < n; ++i) mlp_row(P, feat + i * F0, h1, h2, out + i * F3);
}

double orc_mlp_loss_grad(const float* P, const float* feat, const float* target, const float* wk, uint64_t n,
                         float* grad) {
  std::vector<double> g(N_PARAMS, 0.0);
  double loss = 0.0;
  float h1[H1], h2[H2], y[F3];
  for (uint64_t i = 0; i < n; ++i) {
    const float* x = feat + i * F0;
    mlp_row(P, x, h1, h2, y);
    double dy[F3];
    for (int o = 0; o < F3; ++o) {
      const double e = (double)y[o] - target[i * F3 + o];
      loss += 0.5 * wk[o] * e * e;
      dy[o] = wk[o] * e;
    }
    double d2[H2];
    for (int k = 0; k < H2; ++k) d2[k] = 0.0;
    for (int o = 0; o < F3; ++o) {
      g[O_B3 + o] += dy[o];
      for (int k = 0; k < H2; ++k) {
        g[O_W3 + o * H2 + k] += dy[o] * h2[k];
        d2[k] += dy[o] * P[O_W3 + o * H2 + k];
      }
    }
    double d1[H1];
    for (int k = 0; k < H1; ++k) d1[k] = 0.0;
    for (int o = 0; o < H2; ++o) {
      const double da = d2[o] * (1.0 - (double)h2[o] * h2[o]);
      g[O_B2 + o] += da;
      for (int k = 0; k < H1; ++k) {
        g[O_W2 + o * H1 + k] += da * h1[k];
        d1[k] += da * P[O_W2 + o * H1 + k];
      }
    }
    for (int o = 0; o < H1; ++o) {
      const double da = d1[o] * (1.0 - (double)h1[o] * h1[o]);
      g[O_B1 + o] += da;
      for (int k = 0; k < F0; ++k) g[O_W1 + o * F0 + k] += da * x[k];
    }
  }
  const double inv = n ? 1.0 / (double)n : 0.0;
  for (size_t k = 0; k < N_PARAMS; ++k) grad[k] = (float)(g[k] * inv);
  return loss * inv;
}

}  // extern "C"
