// Synthetic scene generators (declared in include/rt_scenes.h).
#include "../../include/rt_scenes.h"

#include <math.h>
#include <stdint.h>

#include <unordered_map>
#include <vector>

namespace {

constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;

inline uint64_t mix64(uint64_t z) {  // splitmix64 finaliser (Steele, Lea, Flood 2014)
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
// Draw j of element i: uniform in [0, 1) with 53 bits.
inline double uni(uint64_t seed, uint64_t i, uint64_t j) {
  uint64_t z = mix64(seed + (i * 16 + j + 1) * kGamma);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}
inline double gauss(uint64_t seed, uint64_t i, uint64_t j) {  // Box-Muller, one of the pair
  double u1 = uni(seed, i, j), u2 = uni(seed, i, j + 1);
  if (u1 < 1e-300) u1 = 1e-300;
  return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

struct V {
  double x, y, z;
};
inline V operator+(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V operator-(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V operator*(V a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V cross(V a, V b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double len(V a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
inline float canon(double v) { return (float)v + 0.0f; }  // no negative zeros

// Value noise: hashed lattice values in [-1, 1], smoothstep trilinear.
double lattice(uint64_t seed, int64_t x, int64_t y, int64_t z) {
  uint64_t h = mix64(seed ^ mix64((uint64_t)x * 0x8CB92BA72F3D8DD7ull ^ mix64((uint64_t)y * 0xD6E8FEB86659FD93ull ^
                                                                          mix64((uint64_t)z + 0x1234567ull))));
  return (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}
double value_noise(uint64_t seed, V p) {
  double fx = floor(p.x), fy = floor(p.y), fz = floor(p.z);
  int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
  double tx = p.x - fx, ty = p.y - fy, tz = p.z - fz;
  tx = tx * tx * (3 - 2 * tx);
  ty = ty * ty * (3 - 2 * ty);
  tz = tz * tz * (3 - 2 * tz);
  double c[2][2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int d = 0; d < 2; ++d) c[a][b][d] = lattice(seed, ix + a, iy + b, iz + d);
  double x00 = c[0][0][0] + (c[1][0][0] - c[0][0][0]) * tx, x10 = c[0][1][0] + (c[1][1][0] - c[0][1][0]) * tx;
  double x01 = c[0][0][1] + (c[1][0][1] - c[0][0][1]) * tx, x11 = c[0][1][1] + (c[1][1][1] - c[0][1][1]) * tx;
  double y0 = x00 + (x10 - x00) * ty, y1 = x01 + (x11 - x01) * ty;
  return y0 + (y1 - y0) * tz;
}
double fbm3(uint64_t seed, V p) {
  double s = value_noise(seed, p * 4.0) + 0.5 * value_noise(seed + 1, p * 8.0) + 0.25 * value_noise(seed + 2, p * 16.0);
  return s / 1.75;
}

const double kPhi = 1.6180339887498949;
const V kIco[12] = {{-1, kPhi, 0}, {1, kPhi, 0}, {-1, -kPhi, 0}, {1, -kPhi, 0}, {0, -1, kPhi}, {0, 1, kPhi},
                    {0, -1, -kPhi}, {0, 1, -kPhi}, {kPhi, 0, -1}, {kPhi, 0, 1}, {-kPhi, 0, -1}, {-kPhi, 0, 1}};
const int kIcoF[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11}, {1, 5, 9},  {5, 11, 4},
                          {11, 10, 2}, {10, 7, 6}, {7, 1, 8},  {3, 9, 4},  {3, 4, 2},   {3, 2, 6},  {3, 6, 8},
                          {3, 8, 9},  {4, 9, 5},  {2, 4, 11}, {6, 2, 10}, {8, 6, 7},   {9, 8, 1}};

}  // namespace

extern "C" {

void rt_gen_soup(int32_t n, uint64_t seed, double edge_median, double edge_sigma, const double offset[3],
                 double scale, float* vx, float* vy, float* vz, float* nx, float* ny, float* nz, int32_t* idx) {
  const double lnmed = log(edge_median);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    V c = {uni(seed, i, 0) * scale + offset[0], uni(seed, i, 1) * scale + offset[1], uni(seed, i, 2) * scale + offset[2]};
    double s = exp(lnmed + edge_sigma * gauss(seed, i, 3));
    double u1 = uni(seed, i, 5), u2 = uni(seed, i, 6), u3 = uni(seed, i, 7);
    // Shoemake (1992): uniformly distributed unit quaternion (w, x, y, z).
    double qx = sqrt(1 - u1) * sin(2 * M_PI * u2), qy = sqrt(1 - u1) * cos(2 * M_PI * u2);
    double qz = sqrt(u1) * sin(2 * M_PI * u3), qw = sqrt(u1) * cos(2 * M_PI * u3);
    double R[3][3] = {{1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)},
                      {2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)},
                      {2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)}};
    double circ = s / sqrt(3.0);
    for (int k = 0; k < 3; ++k) {
      double a = 2 * M_PI * k / 3.0;
      V t = {circ * cos(a), circ * sin(a), 0};
      V w = {R[0][0] * t.x + R[0][1] * t.y, R[1][0] * t.x + R[1][1] * t.y, R[2][0] * t.x + R[2][1] * t.y};
      vx[3 * i + k] = canon(c.x + w.x);
      vy[3 * i + k] = canon(c.y + w.y);
      vz[3 * i + k] = canon(c.z + w.z);
      nx[3 * i + k] = canon(R[0][2]);
      ny[3 * i + k] = canon(R[1][2]);
      nz[3 * i + k] = canon(R[2][2]);
      idx[3 * i + k] = (int32_t)(3 * i + k);
    }
  }
}

void rt_gen_icosphere_counts(int32_t freq, int32_t* nv, int32_t* n) {
  *n = 20 * freq * freq;
  *nv = 10 * freq * freq + 2;
}

void rt_gen_icosphere(int32_t freq, double radius, const double center[3], double noise_amp, uint64_t noise_seed,
                      float* vx, float* vy, float* vz, float* nx, float* ny, float* nz, int32_t* idx) {
  const int f = freq;
  std::vector<V> pos;
  pos.reserve(10 * (size_t)f * f + 2);
  for (int k = 0; k < 12; ++k) pos.push_back(kIco[k] * (1.0 / len(kIco[k])));
  std::unordered_map<uint64_t, int32_t> edge_vert;
  auto edge_key = [](int u, int v, int wu) { return ((uint64_t)u << 40) | ((uint64_t)v << 20) | (uint64_t)wu; };
  size_t tri = 0;
  std::vector<int32_t> grid;
  for (int fc = 0; fc < 20; ++fc) {
    const int A = kIcoF[fc][0], B = kIcoF[fc][1], C = kIcoF[fc][2];
    grid.assign((size_t)(f + 1) * (f + 1), -1);
    for (int i = 0; i <= f; ++i)
      for (int j = 0; j <= f - i; ++j) {
        int wa = f - i - j, wb = i, wc = j;
        int id;
        int nz = (wa > 0) + (wb > 0) + (wc > 0);
        if (nz == 1) {
          id = wa ? A : (wb ? B : C);
        } else if (nz == 2) {
          int u, v, wu, wv;
          if (wa == 0) { u = B; v = C; wu = wb; wv = wc; }
          else if (wb == 0) { u = A; v = C; wu = wa; wv = wc; }
          else { u = A; v = B; wu = wa; wv = wb; }
          if (u > v) { int t = u; u = v; v = t; t = wu; wu = wv; wv = t; }
          uint64_t key = edge_key(u, v, wu);
          auto it = edge_vert.find(key);
          if (it != edge_vert.end()) {
            id = it->second;
          } else {
            V p = kIco[u] * ((double)wu / f) + kIco[v] * ((double)wv / f);
            id = (int32_t)pos.size();
            pos.push_back(p * (1.0 / len(p)));
            edge_vert.emplace(key, id);
          }
        } else {
          V p = (kIco[A] * ((double)wa / f) + kIco[B] * ((double)wb / f)) + kIco[C] * ((double)wc / f);
          id = (int32_t)pos.size();
          pos.push_back(p * (1.0 / len(p)));
        }
        grid[(size_t)i * (f + 1) + j] = id;
      }
    auto g = [&](int i, int j) { return grid[(size_t)i * (f + 1) + j]; };
    for (int i = 0; i < f; ++i)
      for (int j = 0; j < f - i; ++j) {
        idx[3 * tri + 0] = g(i, j);
        idx[3 * tri + 1] = g(i + 1, j);
        idx[3 * tri + 2] = g(i, j + 1);
        ++tri;
        if (i + j < f - 1) {
          idx[3 * tri + 0] = g(i + 1, j);
          idx[3 * tri + 1] = g(i + 1, j + 1);
          idx[3 * tri + 2] = g(i, j + 1);
          ++tri;
        }
      }
  }
  const size_t nvert = pos.size();
  std::vector<V> out(nvert);
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < (int64_t)nvert; ++k) {
    V u = pos[k];
    double r = radius;
    if (noise_amp != 0.0) r *= 1.0 + noise_amp * fbm3(noise_seed, u);
    V p = u * r + V{center[0], center[1], center[2]};
    out[k] = {(double)canon(p.x), (double)canon(p.y), (double)canon(p.z)};
    vx[k] = canon(p.x);
    vy[k] = canon(p.y);
    vz[k] = canon(p.z);
  }
  std::vector<V> acc(nvert, V{0, 0, 0});
  for (size_t t = 0; t < tri; ++t) {
    int a = idx[3 * t], b = idx[3 * t + 1], c = idx[3 * t + 2];
    V fnrm = cross(out[b] - out[a], out[c] - out[a]);
    acc[a] = acc[a] + fnrm;
    acc[b] = acc[b] + fnrm;
    acc[c] = acc[c] + fnrm;
  }
  for (size_t k = 0; k < nvert; ++k) {
    double l = len(acc[k]);
    V nn = l > 0 ? acc[k] * (1.0 / l) : V{0, 1, 0};
    nx[k] = canon(nn.x);
    ny[k] = canon(nn.y);
    nz[k] = canon(nn.z);
  }
}

void rt_gen_jitter(int32_t nv, const float* vx, const float* vy, const float* vz, double sigma, uint64_t seed,
                   float* ox, float* oy, float* oz) {
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < nv; ++k) {
    double gx = gauss(seed, k, 0), gy = gauss(seed, k, 2), gz = gauss(seed, k, 4);
    ox[k] = canon((double)vx[k] + sigma * gx);
    oy[k] = canon((double)vy[k] + sigma * gy);
    oz[k] = canon((double)vz[k] + sigma * gz);
  }
}

}  // extern "C"
