// Initial-condition generators (host side of the library).
//
// Two families:
//  * STELLAR_IC_DISK / PLUMMER / COLLIDING: the seeded workloads the engine is
//    measured on.  Every random draw is a pure function of (seed, body index,
//    draw number) through a splitmix64 hash, so a body's state does not
//    depend on n, on thread count or on the order bodies are generated in.
//  * STELLAR_IC_REF_*: the reference's own models
//    (BarnesHutParticleSystem.cpp:390-560), written with the same standard
//    <random> engine and distributions, so they reproduce the reference's
//    initial state exactly when built against the same C++ library.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

#include "stellar.h"

namespace {

inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Uniform in the open interval (0,1) from 53 random bits.
inline double uniform01(uint64_t seed, uint64_t stream, uint64_t body,
                        uint64_t draw) {
  uint64_t h = mix64(seed ^ mix64(stream * 0xD1B54A32D192ED03ull));
  h = mix64(h ^ mix64(body + 0x632BE59BD9B4E019ull));
  h = mix64(h ^ (draw * 0x8CB92BA72F3D8DD7ull));
  return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

inline void normal2(double u1, double u2, double* z0, double* z1) {
  const double r = std::sqrt(-2.0 * std::log(u1));
  const double a = 2.0 * M_PI * u2;
  *z0 = r * std::cos(a);
  *z1 = r * std::sin(a);
}

// One disk of `cnt` bodies starting at index `first`: body `first` is the
// central mass; the rest have mass 1, radius ~ U[rmin,rmax], angle ~
// U[0,2pi), circular speed sqrt(G*(M + n_inside(r))/r) counter-clockwise,
// plus N(0,jitter) per component, shifted by (cx,cy) and (bvx,bvy).
void make_disk(const stellar_ic_params* p, uint64_t stream, int32_t first,
               int32_t cnt, double cx, double cy, double bvx, double bvy,
               float* mass, float* x, float* y, float* vx, float* vy) {
  if (cnt <= 0) return;
  std::vector<double> r(cnt), phi(cnt);
  for (int32_t k = 1; k < cnt; ++k) {
    const uint64_t b = (uint64_t)(first + k);
    r[k] = p->rmin + (p->rmax - p->rmin) * uniform01(p->seed, stream, b, 0);
    phi[k] = 2.0 * M_PI * uniform01(p->seed, stream, b, 1);
  }
  // n_inside(r_k) = number of disk bodies (mass 1) strictly inside r_k.
  std::vector<int32_t> order(cnt > 1 ? cnt - 1 : 0);
  for (int32_t k = 1; k < cnt; ++k) order[k - 1] = k;
  std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
    return r[a] < r[b] || (r[a] == r[b] && a < b);
  });
  std::vector<int32_t> inside(cnt, 0);
  for (size_t q = 0; q < order.size(); ++q) {
    size_t lo = q;
    while (lo > 0 && r[order[lo - 1]] == r[order[q]]) --lo;
    inside[order[q]] = (int32_t)lo;
  }
  mass[first] = (float)p->central_mass;
  x[first] = (float)cx;
  y[first] = (float)cy;
  vx[first] = (float)bvx;
  vy[first] = (float)bvy;
  for (int32_t k = 1; k < cnt; ++k) {
    const uint64_t b = (uint64_t)(first + k);
    const double v =
        std::sqrt(p->gravity * (p->central_mass + (double)inside[k]) / r[k]);
    double z0, z1;
    normal2(uniform01(p->seed, stream, b, 2), uniform01(p->seed, stream, b, 3),
            &z0, &z1);
    const double s = std::sin(phi[k]), c = std::cos(phi[k]);
    mass[first + k] = 1.0f;
    x[first + k] = (float)(cx + r[k] * c);
    y[first + k] = (float)(cy + r[k] * s);
    vx[first + k] = (float)(bvx - v * s + p->jitter * z0);
    vy[first + k] = (float)(bvy + v * c + p->jitter * z1);
  }
}

// Plummer sphere of total mass 1 (masses 1/n), projected to the plane with a
// uniform angle: r = a/sqrt(u^(-2/3)-1) resampled until r < rcut*a; speed
// q*v_esc(r), v_esc = sqrt(2G)(r^2+a^2)^(-1/4), with q drawn from
// g(q) = q^2 (1-q^2)^(7/2) by rejection (Aarseth, Henon & Wielen 1974);
// velocity direction uniform in the plane.
void make_plummer(const stellar_ic_params* p, int32_t n, float* mass, float* x,
                  float* y, float* vx, float* vy) {
  const double a = p->plummer_a, rcut = p->plummer_rcut * p->plummer_a;
  const float m = (float)(1.0 / (double)n);
  for (int32_t i = 0; i < n; ++i) {
    uint64_t d = 0;
    double r;
    do {
      const double u = uniform01(p->seed, 7, (uint64_t)i, d++);
      r = a / std::sqrt(std::pow(u, -2.0 / 3.0) - 1.0);
    } while (!(r < rcut) && d < 1000);
    if (!(r < rcut)) r = 0.5 * rcut;
    const double phi = 2.0 * M_PI * uniform01(p->seed, 7, (uint64_t)i, d++);
    double q, g;
    do {
      q = uniform01(p->seed, 7, (uint64_t)i, d++);
      g = 0.1 * uniform01(p->seed, 7, (uint64_t)i, d++);
    } while (g > q * q * std::pow(1.0 - q * q, 3.5) && d < 100000);
    const double v = q * std::sqrt(2.0 * p->gravity) *
                     std::pow(r * r + a * a, -0.25);
    const double psi = 2.0 * M_PI * uniform01(p->seed, 7, (uint64_t)i, d++);
    mass[i] = m;
    x[i] = (float)(r * std::cos(phi));
    y[i] = (float)(r * std::sin(phi));
    vx[i] = (float)(v * std::cos(psi));
    vy[i] = (float)(v * std::sin(psi));
  }
}

// ---- the reference's models ---------------------------------------------
// BarnesHutParticleSystem.cpp:390-428.  The default-constructed
// std::default_random_engine, float distributions and the mixed float/double
// expressions are those of the reference, so the draws and roundings agree.
void ref_plummer(float gravity, int32_t n, float* mass, float* x, float* y,
                 float* vx, float* vy) {
  (void)gravity;
  const float a = 1.0f;
  const float pi = 3.14159265f;
  std::default_random_engine gen;
  std::uniform_real_distribution<float> dist(0, 1.0);
  std::uniform_real_distribution<float> dist2(0, 0.1);
  std::uniform_real_distribution<float> dist_phi(0.0, 2 * pi);
  std::uniform_real_distribution<float> dist_theta(-1.0, 1.0);
  for (int32_t i = 0; i < n; ++i) {
    float phi = dist_phi(gen);
    float theta = std::acos(dist_theta(gen));
    (void)theta;
    const double u = (double)dist(gen);
    const float r = (float)((double)a / std::sqrt(std::pow(u, -0.666666) - 1));
    mass[i] = 1.0f;
    x[i] = r * std::cos(phi);
    y[i] = r * std::sin(phi);
    float s = 0.0f, t = 0.1f;
    while ((double)t > (double)(s * s) * std::pow(1.0 - (double)(s * s), 3.5)) {
      s = dist(gen);
      t = dist2(gen);
    }
    const float v = (float)((double)(100 * s) * std::sqrt(2.0) *
                            std::pow(1.0 + (double)(r * r), -0.25));
    phi = dist_phi(gen);
    theta = std::acos(dist_theta(gen));
    vx[i] = v * std::cos(phi);
    vy[i] = v * std::sin(phi);
  }
}

// BarnesHutParticleSystem.cpp:439-483.
void ref_disk(float gravity, int32_t n, float* mass, float* x, float* y,
              float* vx, float* vy) {
  const float pi = 3.14159265f;
  std::default_random_engine gen;
  std::uniform_real_distribution<float> dist(1.5, 12.0);
  std::uniform_real_distribution<float> dist_theta(0.0, 2 * pi);
  for (int32_t i = 0; i < n; ++i) {
    const float theta = dist_theta(gen);
    const float r = dist(gen);
    if (i == 0) {
      mass[i] = 200000;
      x[i] = 0;
      y[i] = 0;
    } else {
      mass[i] = 1.0f;
      x[i] = r * std::cos(theta);
      y[i] = r * std::sin(theta);
    }
    const float rotation = -1;
    const float v = (float)(1.0 * std::sqrt((double)gravity * 200000.0 / (double)r));
    if (i == 0) {
      vx[0] = 0;
      vy[0] = 0;
    } else {
      vx[i] = rotation * v * std::sin(theta);
      vy[i] = -rotation * v * std::cos(theta);
    }
  }
}

// BarnesHutParticleSystem.cpp:494-560.
void ref_colliding(float gravity, int32_t n, float* mass, float* x, float* y,
                   float* vx, float* vy) {
  const float pi = 3.14159265f;
  std::default_random_engine gen;
  std::uniform_real_distribution<float> dist1(1.5, 12.0);
  std::uniform_real_distribution<float> dist2(1, 5.0);
  std::uniform_real_distribution<float> dist_theta(0.0, 2 * pi);
  for (int32_t i = 0; i < n; ++i) {
    const float theta = dist_theta(gen);
    const float r1 = dist1(gen);
    const float r2 = dist2(gen);
    if (i == 0) {
      mass[i] = 100000;
      x[i] = 0;
      y[i] = 0;
    } else if (i == 1) {
      mass[i] = 25000;
      x[i] = 20 * std::cos(theta);
      y[i] = 20 * std::sin(theta);
    } else if (i <= 3 * n / 4) {
      mass[i] = 1.0f;
      x[i] = r1 * std::cos(theta);
      y[i] = r1 * std::sin(theta);
    } else {
      mass[i] = 1.0f;
      x[i] = r2 * std::cos(theta) + x[1];
      y[i] = r2 * std::sin(theta) + y[1];
    }
    const float rotation = 1;
    const float v1 = (float)(1.0 * std::sqrt((double)gravity * 100000.0 / (double)r1));
    const float v2 = (float)(1.0 * std::sqrt((double)gravity * 25000.0 / (double)r2));
    if (i == 0 || i == 1) {
      vx[i] = 0.0f;
      vy[i] = 0.0f;
    } else if (i <= 3 * n / 4) {
      vx[i] = rotation * v1 * std::sin(theta);
      vy[i] = -rotation * v1 * std::cos(theta);
    } else {
      vx[i] = rotation * v2 * std::sin(theta);
      vy[i] = -rotation * v2 * std::cos(theta);
    }
  }
}

}  // namespace

extern "C" int stellar_ic_defaults(int32_t workload, stellar_ic_params* p) {
  if (!p) return STELLAR_ERR_ARG;
  std::memset(p, 0, sizeof(*p));
  p->gravity = 1.0;
  p->jitter = 0.01;
  p->plummer_a = 1.0;
  p->plummer_rcut = 20.0;
  switch (workload) {
    case 0:
      p->model = STELLAR_IC_DISK;
      p->seed = 0;
      p->central_mass = 1000.0;
      p->rmin = 2.0;
      p->rmax = 100.0;
      return STELLAR_OK;
    case 1:
      p->model = STELLAR_IC_PLUMMER;
      p->seed = 1;
      return STELLAR_OK;
    case 2:
      p->model = STELLAR_IC_COLLIDING;
      p->seed = 2;
      p->central_mass = 1.0e4;
      p->rmin = 2.0;
      p->rmax = 200.0;
      p->offset_x = 300.0;
      p->bulk_vx = -5.0;
      p->bulk_vy = 1.0;
      return STELLAR_OK;
    case 4:
      p->model = STELLAR_IC_PLUMMER;
      p->seed = 4;
      p->plummer_a = 0.01;
      return STELLAR_OK;
    default:
      return STELLAR_ERR_ARG;
  }
}

extern "C" int stellar_generate_ic(const stellar_ic_params* p, int32_t n,
                                   float* mass, float* x, float* y, float* vx,
                                   float* vy) {
  if (!p || n < 0 || (n > 0 && (!mass || !x || !y || !vx || !vy)))
    return STELLAR_ERR_ARG;
  if (n == 0) return STELLAR_OK;
  const float g = (float)p->gravity;
  switch (p->model) {
    case STELLAR_IC_DISK:
      make_disk(p, 1, 0, n, 0.0, 0.0, 0.0, 0.0, mass, x, y, vx, vy);
      return STELLAR_OK;
    case STELLAR_IC_PLUMMER:
      if (!(p->plummer_a > 0) || !(p->plummer_rcut > 0)) return STELLAR_ERR_ARG;
      make_plummer(p, n, mass, x, y, vx, vy);
      return STELLAR_OK;
    case STELLAR_IC_COLLIDING: {
      const int32_t h = n / 2;
      make_disk(p, 2, 0, h, p->offset_x, 0.0, p->bulk_vx, p->bulk_vy, mass, x,
                y, vx, vy);
      make_disk(p, 3, h, n - h, -p->offset_x, 0.0, -p->bulk_vx, -p->bulk_vy,
                mass, x, y, vx, vy);
      return STELLAR_OK;
    }
    case STELLAR_IC_REF_DISK:
      ref_disk(g, n, mass, x, y, vx, vy);
      return STELLAR_OK;
    case STELLAR_IC_REF_PLUMMER:
      ref_plummer(g, n, mass, x, y, vx, vy);
      return STELLAR_OK;
    case STELLAR_IC_REF_COLLIDING:
      ref_colliding(g, n, mass, x, y, vx, vy);
      return STELLAR_OK;
    default:
      return STELLAR_ERR_ARG;
  }
}
