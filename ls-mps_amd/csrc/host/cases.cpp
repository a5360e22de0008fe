// Synthetic particle sets (host side of generate_case).
//
// Reference cases (src/lsmps.cpp:68,72), bit-identical positions: a jittered
// lattice filled z-outermost, then y, then x, with three
// uniform_real_distribution<double> draws (x, y, z) per particle from one
// std::mt19937(seed); wall_film / narrow_gap / rain_injection / default
// dimensions and jitters as the reference chooses them.
//
// Measurement scenes (no reference code):
//   bulk_lattice    jittered block nx x ny x 20, jitter 0.1 l0, no walls
//                   (50,000 particles -> 50 x 50 x 20).
//   rain_plate      plate inclined 30 deg: 15% wall in 3 layers, 55% film in
//                   a 3 l0 layer, 30% droplets (spheres of radius U(2,4) l0,
//                   centres up to 40 l0 above the plate).
//   industrial_mix  70% jittered bulk block, 20% wall film 2-4 l0 thick on
//                   a plane tilted 12 deg (bottom layer Wall), 10% splash by
//                   Poisson-disk dart throwing at density 0.2 / l0^3.
#include <algorithm>
#include <cmath>
#include <random>
#include <string>
#include <unordered_map>
#include <vector>

#include "lsmps.hpp"

namespace lsmps {
namespace detail {

namespace {

constexpr double kPi = 3.14159265358979323846;

void lattice_block(std::vector<Particle>& out, size_t target, int nx, int ny, int nz, double l0, double jit,
                   std::mt19937& rng, bool bottom_wall) {
  std::uniform_real_distribution<double> u(-jit, jit);
  for (int k = 0; k < nz && out.size() < target; ++k)
    for (int j = 0; j < ny && out.size() < target; ++j)
      for (int i = 0; i < nx && out.size() < target; ++i) {
        Particle p;
        const double a = u(rng);
        const double b = u(rng);
        const double c = u(rng);
        p.r = {(i + .5 + a) * l0, (j + .5 + b) * l0, (k + .5 + c) * l0};
        if (bottom_wall && k == 0) p.type = ParticleType::Wall;
        out.push_back(p);
      }
}

int plate_side(size_t count, int layers) {
  return std::max(8, (int)std::ceil(std::sqrt((double)count / layers)));
}

// Plate frame (u along the slope, v across, w normal) to world coordinates
// for a plane through the origin tilted by `theta` about the y axis.
Vec3 plate_to_world(double u, double v, double w, double theta) {
  const double c = std::cos(theta), s = std::sin(theta);
  return {u * c - w * s, v, u * s + w * c};
}

void rain_plate(const Config& cfg, std::vector<Particle>& out) {
  const size_t n = cfg.particles;
  const double l0 = cfg.l0, theta = 30.0 * kPi / 180.0;
  const size_t n_wall = (size_t)(0.15 * n), n_film = (size_t)(0.55 * n);
  const size_t n_drop = n - n_wall - n_film;
  std::mt19937 rng(cfg.seed);
  std::uniform_real_distribution<double> jit(-0.05, 0.05);
  // Film: 3 layers at spacing l0 on an Lu x Lv plate (Lu/Lv = 1.4).
  const int film_layers = 3;
  const double per_layer = std::ceil((double)n_film / film_layers);
  const int nv = std::max(4, (int)std::ceil(std::sqrt(per_layer / 1.4)));
  const int nu = std::max(4, (int)std::ceil(per_layer / nv));
  // Wall: 3 layers under the same plate area; in-plane spacing chosen so
  // 3 layers hold n_wall particles (coarser than l0 when the counts ask so).
  const double wall_per_layer = std::ceil((double)n_wall / 3.0);
  const double sw = std::sqrt((double)nu * nv / wall_per_layer) * l0;
  const int wu = std::max(1, (int)std::floor(nu * l0 / sw)), wv = std::max(1, (int)std::ceil(wall_per_layer / wu));
  size_t placed = 0;
  for (int k = 0; k < 3 && placed < n_wall; ++k)
    for (int j = 0; j < wv && placed < n_wall; ++j)
      for (int i = 0; i < wu && placed < n_wall; ++i) {
        Particle p;
        p.r = plate_to_world((i + .5) * sw + jit(rng) * sw, (j + .5) * sw + jit(rng) * sw, -(k + .5) * l0, theta);
        p.type = ParticleType::Wall;
        out.push_back(p);
        ++placed;
      }
  placed = 0;
  for (int k = 0; k < film_layers && placed < n_film; ++k)
    for (int j = 0; j < nv && placed < n_film; ++j)
      for (int i = 0; i < nu && placed < n_film; ++i) {
        Particle p;
        const double a = jit(rng), b = jit(rng), c = jit(rng);
        p.r = plate_to_world((i + .5 + a) * l0, (j + .5 + b) * l0, (k + .5 + c) * l0, theta);
        out.push_back(p);
        ++placed;
      }
  // Droplets: spheres of lattice points (jittered) until n_drop are placed.
  std::uniform_real_distribution<double> uu(0.0, nu * l0), uv(0.0, nv * l0), ur(2.0 * l0, 4.0 * l0), u01(0.0, 1.0);
  placed = 0;
  while (placed < n_drop) {
    const double R = ur(rng), cu = uu(rng), cv = uv(rng);
    const double wmin = film_layers * l0 + R + 0.5 * l0, wmax = 40.0 * l0;
    const double cw = wmin + u01(rng) * std::max(0.0, wmax - wmin);
    const int m = (int)std::ceil(R / l0);
    for (int c = -m; c <= m && placed < n_drop; ++c)
      for (int b = -m; b <= m && placed < n_drop; ++b)
        for (int a = -m; a <= m && placed < n_drop; ++a) {
          if ((double)(a * a + b * b + c * c) * l0 * l0 > R * R) continue;
          Particle p;
          const double ja = jit(rng), jb = jit(rng), jc = jit(rng);
          p.r = plate_to_world(cu + (a + ja) * l0, cv + (b + jb) * l0, cw + (c + jc) * l0, theta);
          p.u = {0, 0, -5};
          p.type = ParticleType::Splash;
          out.push_back(p);
          ++placed;
        }
  }
}

void industrial_mix(const Config& cfg, std::vector<Particle>& out) {
  const size_t n = cfg.particles;
  const double l0 = cfg.l0;
  const size_t n_bulk = (size_t)(0.7 * n), n_film = (size_t)(0.2 * n), n_splash = n - n_bulk - n_film;
  std::mt19937 rng(cfg.seed);
  // Bulk block: nx = ny, nz ~ nx / 2.
  const int nz = std::max(4, (int)std::ceil(std::cbrt((double)n_bulk / 3.8)));
  const int nx = std::max(4, (int)std::ceil(std::sqrt((double)n_bulk / nz)));
  lattice_block(out, n_bulk, nx, nx, nz, l0, 0.1, rng, false);
  while (out.size() < n_bulk) out.push_back(out.back());  // unreachable: nx*nx*nz >= n_bulk
  // Wall film on a plane tilted 12 deg, starting 4 l0 beyond the block in x;
  // thickness h(u, v) = (3 + sin(2 pi u / 40 l0) cos(2 pi v / 30 l0)) l0.
  const double theta = 12.0 * kPi / 180.0;
  const int fv = 2 * nx;
  const int fu = std::max(8, (int)std::ceil((double)n_film / (3.0 * fv)) + 8);
  const double x0 = (nx + 4) * l0;
  std::uniform_real_distribution<double> jf(-0.05, 0.05);
  size_t placed = 0;
  for (int k = 0; k < 4 && placed < n_film; ++k)
    for (int j = 0; j < fv && placed < n_film; ++j)
      for (int i = 0; i < fu && placed < n_film; ++i) {
        const double u = (i + .5) * l0, v = (j + .5) * l0, w = (k + .5) * l0;
        const double h = (3.0 + std::sin(2 * kPi * u / (40 * l0)) * std::cos(2 * kPi * v / (30 * l0))) * l0;
        if (w >= h) continue;
        Particle p;
        const double a = jf(rng), b = jf(rng), c = jf(rng);
        Vec3 q = plate_to_world(u + a * l0, v + b * l0, w + c * l0, theta);
        q.x += x0;
        p.r = q;
        if (k == 0) p.type = ParticleType::Wall;
        out.push_back(p);
        ++placed;
      }
  // Splash: dart throwing with minimum distance l0 above the block.
  const double h_box = (double)n_splash / (0.2 * nx * nx) * l0;
  const double zb = (nz + 2) * l0, L = nx * l0;
  std::uniform_real_distribution<double> ux(0.0, L), uz(zb, zb + h_box);
  const double cs = l0;  // background cell = min distance
  const int gx = (int)std::ceil(L / cs) + 1, gz = (int)std::ceil(h_box / cs) + 1;
  std::vector<int> grid((size_t)gx * gx * gz, -1);
  std::vector<Vec3> acc;
  acc.reserve(n_splash);
  size_t tries = 0;
  const size_t max_tries = 200 * n_splash + 1000;
  while (acc.size() < n_splash && tries < max_tries) {
    ++tries;
    const double px = ux(rng);
    const double py = ux(rng);
    const double pz = uz(rng);
    const int ix = std::min(gx - 1, (int)(px / cs)), iy = std::min(gx - 1, (int)(py / cs)),
              iz = std::min(gz - 1, (int)((pz - zb) / cs));
    bool ok = true;
    for (int dz = -1; dz <= 1 && ok; ++dz)
      for (int dy = -1; dy <= 1 && ok; ++dy)
        for (int dx = -1; dx <= 1 && ok; ++dx) {
          const int a = ix + dx, b = iy + dy, c = iz + dz;
          if (a < 0 || b < 0 || c < 0 || a >= gx || b >= gx || c >= gz) continue;
          const int q = grid[((size_t)c * gx + b) * gx + a];
          if (q < 0) continue;
          const Vec3 d = acc[q] - Vec3{px, py, pz};
          if (d.x * d.x + d.y * d.y + d.z * d.z < l0 * l0) ok = false;
        }
    if (!ok) continue;
    int& slot = grid[((size_t)iz * gx + iy) * gx + ix];
    if (slot >= 0) continue;  // one point per cell of size l0 (min distance l0)
    slot = (int)acc.size();
    acc.push_back({px, py, pz});
  }
  for (const Vec3& q : acc) {
    Particle p;
    p.r = q;
    p.type = ParticleType::Splash;
    out.push_back(p);
  }
}

}  // namespace

std::vector<Particle> make_case(const Config& c) {
  std::vector<Particle> p;
  p.reserve(c.particles);
  const std::string& name = c.case_name;
  if (name == "wall_film") {
    std::mt19937 rng(c.seed);
    const int nz = std::max(3, c.film_layers + 1), nx = plate_side(c.particles, nz);
    lattice_block(p, c.particles, nx, nx, nz, c.l0, .04, rng, true);
  } else if (name == "narrow_gap") {
    std::mt19937 rng(c.seed);
    const int nz = std::max(4, c.film_layers + 2), nx = plate_side(c.particles, nz);
    lattice_block(p, c.particles, nx, nx, nz, c.l0, .02, rng, true);
    for (Particle& q : p)
      if (q.r.z > (nz - 1.5) * c.l0) q.type = ParticleType::Wall;
  } else if (name == "rain_injection") {
    std::mt19937 rng(c.seed);
    const size_t film = (size_t)(.75 * c.particles);
    const int nz = std::max(3, c.film_layers + 1), nx = plate_side(film, nz);
    lattice_block(p, film, nx, nx, nz, c.l0, .05, rng, true);
    std::uniform_real_distribution<double> ux(0, nx * c.l0), uy(0, nx * c.l0), uz(4 * c.l0, 20 * c.l0);
    while (p.size() < c.particles) {
      Particle q;
      const double a = ux(rng);
      const double b = uy(rng);
      const double d = uz(rng);
      q.r = {a, b, d};
      q.u = {0, 0, -5};
      q.type = ParticleType::Splash;
      p.push_back(q);
    }
  } else if (name == "bulk_lattice") {
    std::mt19937 rng(c.seed);
    const int nz = 20, nx = std::max(4, (int)std::ceil(std::sqrt((double)c.particles / nz)));
    lattice_block(p, c.particles, nx, nx, nz, c.l0, .1, rng, false);
  } else if (name == "rain_plate") {
    rain_plate(c, p);
  } else if (name == "industrial_mix") {
    industrial_mix(c, p);
  } else {
    std::mt19937 rng(c.seed);
    const int side = std::max(4, (int)std::ceil(std::cbrt((double)c.particles)));
    lattice_block(p, c.particles, side, side, side, c.l0, .03, rng, false);
  }
  return p;
}

}  // namespace detail
}  // namespace lsmps
