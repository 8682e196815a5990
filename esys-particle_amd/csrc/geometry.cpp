// geometry.cpp - seeded dense packing of polydisperse spheres in a box
// (setup-time input generation for specimens; not on the per-step path).
//
// Grow-and-relax: centres start uniformly random with radii scaled by s < 1;
// s grows geometrically while overlaps are relaxed by Jacobi sweeps that move
// each overlapping pair apart along their centre line (half the overlap each,
// damped) and keep spheres inside the box.  The result is a jammed-like
// packing whose near-contacts give the bond network of a bonded specimen.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct SplitMix {
  uint64_t s;
  explicit SplitMix(uint64_t seed) : s(seed) {}
  uint64_t next()
  {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct Grid {
  double cs;
  int nx, ny, nz;
  std::vector<int> start, items;
  void build(const std::vector<double>& x, int n, const double* box, double cell)
  {
    cs = cell;
    nx = std::max(1, (int)(box[0] / cs));
    ny = std::max(1, (int)(box[1] / cs));
    nz = std::max(1, (int)(box[2] / cs));
    start.assign((size_t)nx * ny * nz + 1, 0);
    items.resize(n);
    std::vector<int> c(n);
    for (int i = 0; i < n; i++) {
      int ix = std::min(nx - 1, std::max(0, (int)(x[3 * i] / cs)));
      int iy = std::min(ny - 1, std::max(0, (int)(x[3 * i + 1] / cs)));
      int iz = std::min(nz - 1, std::max(0, (int)(x[3 * i + 2] / cs)));
      c[i] = (ix * ny + iy) * nz + iz;
      start[c[i] + 1]++;
    }
    for (size_t k = 1; k < start.size(); k++) start[k] += start[k - 1];
    std::vector<int> fill(start.begin(), start.end() - 1);
    for (int i = 0; i < n; i++) items[fill[c[i]]++] = i;
  }
};

}  // namespace

extern "C" {

// pos: out, 3n doubles.  rad: n radii.  box: extents (origin at 0).
// Returns the final maximum overlap relative to the smallest radius.
// Each sweep computes every particle's displacement from its own overlaps (a
// gather, so the sweep is deterministic under any thread count).
// periodic_x: x is circular (overlaps across the x ends count through the
// nearest image; positions wrap instead of being clamped).  state[0] is the
// radius scale of the growth: <= 0 starts from seeded random positions (scale
// 0.2), otherwise the sweeps continue from pos at that scale; on return it
// holds the scale reached, state[1] the last maximum overlap, and state[2] 1
// once converged.  So a long packing can run as several calls.
double esp_geom_pack_steps(int n, const double* rad, const double* box, int periodic_x,
                           unsigned long long seed, int max_sweeps, double* pos, double* state)
{
  std::vector<double> x(3 * (size_t)n), d(3 * (size_t)n);
  double rmax = 0, rmin = 1e300;
  for (int i = 0; i < n; i++) {
    rmax = std::max(rmax, rad[i]);
    rmin = std::min(rmin, rad[i]);
  }
  double s = state[0];
  if (s <= 0) {
    SplitMix rng(seed);
    for (int i = 0; i < n; i++)
      for (int c = 0; c < 3; c++) x[3 * i + c] = rng.uniform() * box[c];
    s = 0.2;
  } else {
    memcpy(x.data(), pos, sizeof(double) * 3 * (size_t)n);
  }
  state[2] = 0;
  Grid g;
  std::vector<int> cellof(n);
  double maxov = 0;
  for (int sweep = 0; sweep < max_sweeps; sweep++) {
    g.build(x, n, box, 2 * rmax * s * 1.0001);
    maxov = 0;
#pragma omp parallel for schedule(static) reduction(max : maxov)
    for (int i = 0; i < n; i++) {
      const int cx = std::min(g.nx - 1, std::max(0, (int)(x[3 * i] / g.cs)));
      const int cy = std::min(g.ny - 1, std::max(0, (int)(x[3 * i + 1] / g.cs)));
      const int cz = std::min(g.nz - 1, std::max(0, (int)(x[3 * i + 2] / g.cs)));
      double ddx = 0, ddy = 0, ddz = 0;
      // x cells to visit: the stencil, wrapped when x is circular (every cell
      // once when there are fewer than three)
      int xs[3], nxs = 0;
      if (periodic_x && g.nx < 3) {
        for (int c = 0; c < g.nx; c++) xs[nxs++] = c;
      } else {
        for (int ox = -1; ox <= 1; ox++) {
          int x2 = cx + ox;
          if (periodic_x) x2 = (x2 + g.nx) % g.nx;
          else if (x2 < 0 || x2 >= g.nx) continue;
          xs[nxs++] = x2;
        }
      }
      for (int ix = 0; ix < nxs; ix++)
        for (int oy = -1; oy <= 1; oy++)
          for (int oz = -1; oz <= 1; oz++) {
            const int x2 = xs[ix], y2 = cy + oy, z2 = cz + oz;
            if (y2 < 0 || z2 < 0 || y2 >= g.ny || z2 >= g.nz) continue;
            const int c2 = (x2 * g.ny + y2) * g.nz + z2;
            for (int b = g.start[c2]; b < g.start[c2 + 1]; b++) {
              const int j = g.items[b];
              if (j == i) continue;
              double dx = x[3 * j] - x[3 * i];
              if (periodic_x) {
                if (dx > 0.5 * box[0]) dx -= box[0];
                else if (dx < -0.5 * box[0]) dx += box[0];
              }
              const double dy = x[3 * j + 1] - x[3 * i + 1], dz = x[3 * j + 2] - x[3 * i + 2];
              const double dist2 = dx * dx + dy * dy + dz * dz;
              const double rs = (rad[i] + rad[j]) * s;
              if (dist2 >= rs * rs) continue;
              const double dist = std::sqrt(dist2);
              const double ov = rs - dist;
              maxov = std::max(maxov, ov / (rmin * s));
              double ux, uy, uz;
              if (dist > 1e-12) {
                ux = dx / dist; uy = dy / dist; uz = dz / dist;
              } else {  // coincident centres: separate along x by index order
                ux = j > i ? 1.0 : -1.0; uy = 0; uz = 0;
              }
              const double h = 0.5 * ov;
              ddx -= h * ux; ddy -= h * uy; ddz -= h * uz;
            }
          }
      d[3 * i] = ddx; d[3 * i + 1] = ddy; d[3 * i + 2] = ddz;
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
      const double r = rad[i] * s;
      for (int c = 0; c < 3; c++) {
        double v = x[3 * i + c] + 0.7 * d[3 * i + c];
        if (c == 0 && periodic_x) {
          if (v < 0) v += box[0];
          else if (v >= box[0]) v -= box[0];
        } else {
          v = std::min(std::max(v, r), box[c] - r);
        }
        x[3 * i + c] = v;
      }
    }
    if (s >= 1.0 && maxov < 2e-3) {
      state[2] = 1;
      break;
    }
    if (maxov < 0.05) s = std::min(1.0, s * 1.01);
  }
  memcpy(pos, x.data(), sizeof(double) * 3 * (size_t)n);
  state[0] = s;
  state[1] = maxov;
  return maxov;
}

double esp_geom_pack_px(int n, const double* rad, const double* box, int periodic_x,
                        unsigned long long seed, int max_sweeps, double* pos)
{
  double state[3] = {0, 0, 0};
  return esp_geom_pack_steps(n, rad, box, periodic_x, seed, max_sweeps, pos, state);
}

double esp_geom_pack(int n, const double* rad, const double* box, unsigned long long seed,
                     int max_sweeps, double* pos)
{
  return esp_geom_pack_px(n, rad, box, 0, seed, max_sweeps, pos);
}

// Moves centres until every listed (bonded) pair is at exact contact,
// |x_i - x_j| = r_i + r_j, and no other pair overlaps, by Jacobi sweeps (each
// particle gathers half of every residual it takes part in).  The bond network
// of a packing has fewer constraints than degrees of freedom, so the sweeps
// converge.  Returns the final maximum residual relative to r_i + r_j.
// periodic_x: pairs across the x ends through the nearest image; x wraps
double esp_geom_snap_px(int n, double* pos, const double* rad, int nb, const int* bonds,
                        const double* box, int periodic_x, int max_sweeps, double tol)
{
  std::vector<int> off(n + 1, 0), adj(2 * (size_t)nb);
  for (int b = 0; b < nb; b++) {
    off[bonds[2 * b] + 1]++;
    off[bonds[2 * b + 1] + 1]++;
  }
  for (int i = 0; i < n; i++) off[i + 1] += off[i];
  {
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (int b = 0; b < nb; b++) {
      adj[fill[bonds[2 * b]]++] = bonds[2 * b + 1];
      adj[fill[bonds[2 * b + 1]]++] = bonds[2 * b];
    }
    for (int i = 0; i < n; i++) std::sort(adj.begin() + off[i], adj.begin() + off[i + 1]);
  }
  std::vector<double> x(pos, pos + 3 * (size_t)n), d(3 * (size_t)n);
  double rmax = 0;
  for (int i = 0; i < n; i++) rmax = std::max(rmax, rad[i]);
  Grid g;
  double maxres = 0;
  for (int sweep = 0; sweep < max_sweeps; sweep++) {
    g.build(x, n, box, 2 * rmax * 1.0001);
    maxres = 0;
#pragma omp parallel for schedule(static) reduction(max : maxres)
    for (int i = 0; i < n; i++) {
      double ddx = 0, ddy = 0, ddz = 0;
      int active = 0;
      auto push = [&](int j, double target, bool bonded) {
        double dx = x[3 * j] - x[3 * i];
        if (periodic_x) {
          if (dx > 0.5 * box[0]) dx -= box[0];
          else if (dx < -0.5 * box[0]) dx += box[0];
        }
        const double dy = x[3 * j + 1] - x[3 * i + 1], dz = x[3 * j + 2] - x[3 * i + 2];
        const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
        double res = dist - target;  // > 0: too far
        if (!bonded && res >= 0) return;
        maxres = std::max(maxres, std::fabs(res) / target);
        if (dist <= 1e-12) return;
        const double h = 0.5 * res / dist;
        ddx += h * dx; ddy += h * dy; ddz += h * dz;
        active++;
      };
      for (int k = off[i]; k < off[i + 1]; k++) push(adj[k], rad[i] + rad[adj[k]], true);
      const int cx = std::min(g.nx - 1, std::max(0, (int)(x[3 * i] / g.cs)));
      const int cy = std::min(g.ny - 1, std::max(0, (int)(x[3 * i + 1] / g.cs)));
      const int cz = std::min(g.nz - 1, std::max(0, (int)(x[3 * i + 2] / g.cs)));
      int xs[3], nxs = 0;
      if (periodic_x && g.nx < 3) {
        for (int c = 0; c < g.nx; c++) xs[nxs++] = c;
      } else {
        for (int ox = -1; ox <= 1; ox++) {
          int x2 = cx + ox;
          if (periodic_x) x2 = (x2 + g.nx) % g.nx;
          else if (x2 < 0 || x2 >= g.nx) continue;
          xs[nxs++] = x2;
        }
      }
      for (int ix = 0; ix < nxs; ix++)
        for (int oy = -1; oy <= 1; oy++)
          for (int oz = -1; oz <= 1; oz++) {
            const int x2 = xs[ix], y2 = cy + oy, z2 = cz + oz;
            if (y2 < 0 || z2 < 0 || y2 >= g.ny || z2 >= g.nz) continue;
            const int c2 = (x2 * g.ny + y2) * g.nz + z2;
            for (int b = g.start[c2]; b < g.start[c2 + 1]; b++) {
              const int j = g.items[b];
              if (j == i) continue;
              if (std::binary_search(adj.begin() + off[i], adj.begin() + off[i + 1], j)) continue;
              push(j, rad[i] + rad[j], false);
            }
          }
      // averaged (not summed) corrections keep the sweep stable for any degree
      const double w = active > 0 ? 1.0 / active : 0.0;
      d[3 * i] = ddx * w; d[3 * i + 1] = ddy * w; d[3 * i + 2] = ddz * w;
    }
    if (maxres < tol) break;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++)
      for (int c = 0; c < 3; c++) {
        double v = x[3 * i + c] + 1.5 * d[3 * i + c];
        if (c == 0 && periodic_x) {
          if (v < 0) v += box[0];
          else if (v >= box[0]) v -= box[0];
          x[3 * i] = v;
        } else {
          x[3 * i + c] = std::min(std::max(v, rad[i]), box[c] - rad[i]);
        }
      }
  }
  memcpy(pos, x.data(), sizeof(double) * 3 * (size_t)n);
  return maxres;
}

double esp_geom_snap(int n, double* pos, const double* rad, int nb, const int* bonds,
                     const double* box, int max_sweeps, double tol)
{
  return esp_geom_snap_px(n, pos, rad, nb, bonds, box, 0, max_sweeps, tol);
}

// Random sequential insertion: sphere k (radii in the given order) is placed
// at a uniformly random centre in [r, box - r] (x in [0, box_x) when
// periodic_x, distances then taken to the nearest x image) where it overlaps no
// earlier sphere; up to max_tries attempts each.  A hash grid of cell 2 r_max
// holds the placed spheres.  Returns the number placed (pos holds them).
int esp_geom_rsa(int n, const double* rad, const double* box, int periodic_x, unsigned long long seed,
                 long long max_tries, double* pos)
{
  SplitMix rng(seed);
  double rmax = 0.0;
  for (int i = 0; i < n; i++) rmax = std::max(rmax, rad[i]);
  const double cs = 2.0 * rmax;
  const int nx = std::max(1, (int)(box[0] / cs)), ny = std::max(1, (int)(box[1] / cs)),
            nz = std::max(1, (int)(box[2] / cs));
  const double cx = box[0] / nx, cy = box[1] / ny, cz = box[2] / nz;
  std::vector<std::vector<int>> cell((size_t)nx * ny * nz);
  auto cidx = [&](double x, double y, double z, int& ix, int& iy, int& iz) {
    ix = std::min(nx - 1, std::max(0, (int)(x / cx)));
    iy = std::min(ny - 1, std::max(0, (int)(y / cy)));
    iz = std::min(nz - 1, std::max(0, (int)(z / cz)));
  };
  int placed = 0;
  for (int k = 0; k < n; k++) {
    const double r = rad[k];
    bool ok = false;
    for (long long t = 0; t < max_tries && !ok; t++) {
      const double x = periodic_x ? rng.uniform() * box[0] : r + rng.uniform() * (box[0] - 2 * r);
      const double y = r + rng.uniform() * (box[1] - 2 * r);
      const double z = r + rng.uniform() * (box[2] - 2 * r);
      int ix, iy, iz;
      cidx(x, y, z, ix, iy, iz);
      ok = true;
      for (int dx = -1; dx <= 1 && ok; dx++) {
        int jx = ix + dx;
        if (periodic_x) jx = (jx + nx) % nx;
        else if (jx < 0 || jx >= nx) continue;
        if (periodic_x && nx < 3 && dx != 0 && (ix + dx + nx) % nx == ix) continue;
        for (int dy = -1; dy <= 1 && ok; dy++) {
          const int jy = iy + dy;
          if (jy < 0 || jy >= ny) continue;
          for (int dz = -1; dz <= 1 && ok; dz++) {
            const int jz = iz + dz;
            if (jz < 0 || jz >= nz) continue;
            for (int j : cell[((size_t)jx * ny + jy) * nz + jz]) {
              double ddx = x - pos[3 * j];
              if (periodic_x) {
                if (ddx > 0.5 * box[0]) ddx -= box[0];
                else if (ddx < -0.5 * box[0]) ddx += box[0];
              }
              const double ddy = y - pos[3 * j + 1], ddz = z - pos[3 * j + 2];
              const double rr = r + rad[j];
              if (ddx * ddx + ddy * ddy + ddz * ddz < rr * rr) {
                ok = false;
                break;
              }
            }
          }
        }
      }
      if (ok) {
        pos[3 * k] = x;
        pos[3 * k + 1] = y;
        pos[3 * k + 2] = z;
        cell[((size_t)ix * ny + iy) * nz + iz].push_back(k);
        placed = k + 1;
      }
    }
    if (!ok) break;
  }
  return placed;
}

}  // extern "C"
