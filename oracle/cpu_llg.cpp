// CPU restatement of magex's numba fast path for the stochastic Heun step (test infrastructure and the
// CPU baseline only: nothing under magex_amd/ links or loads this file's library).
//
// Restates, function by function:
//   onsite_grad      magex/fast.py:39-64   (A tensor only; cubic / higher uniaxial terms are not restated)
//   hfield           magex/fast.py:86-120  (coupling table of width 1, 4 or 9)
//   xorshift64* RNG  magex/fast.py:123-134 and the 128-layer ziggurat of magex/fast.py:136-187
//   rhs              magex/fast.py:225-236
//   fused1 / fused2  magex/fast.py:336-390 (Heun branch; pass 1 over 256 fixed chunks, one stream each)
//
// Build: make -C oracle   (g++ -O3 -march=x86-64-v3 -fopenmp -shared: AVX2/FMA, runs on any x86-64 host
// of the last decade, including hosts other than the one that built it)
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <omp.h>

namespace {

constexpr double MU_B = 5.7883818060e-2;
constexpr double GAMMA = 0.1760859630;

struct Zig {
  double kn[128], wn[128], fn[128];
  Zig() {
    // Marsaglia-Tsang tables, fast.py:137-158
    double dn = 3.442619855899, tn = dn;
    const double vn = 9.91256303526217e-3, m1 = 2147483648.0;
    const double q = vn / std::exp(-0.5 * dn * dn);
    kn[0] = (dn / q) * m1;
    kn[1] = 0.0;
    wn[0] = q / m1;
    wn[127] = dn / m1;
    fn[0] = 1.0;
    fn[127] = std::exp(-0.5 * dn * dn);
    for (int i = 126; i >= 1; --i) {
      dn = std::sqrt(-2.0 * std::log(vn / dn + std::exp(-0.5 * dn * dn)));
      kn[i + 1] = (dn / tn) * m1;
      tn = dn;
      fn[i] = std::exp(-0.5 * dn * dn);
      wn[i] = dn / m1;
    }
  }
};
const Zig ZIG;

inline uint64_t xs_next(uint64_t& s) {
  uint64_t x = s;
  x ^= x >> 12;
  x ^= x << 25;
  x ^= x >> 27;
  s = x;
  return x * 2685821657736338717ULL;
}
inline double xs_unif(uint64_t& s) { return double(xs_next(s) >> 11) * (1.0 / 9007199254740992.0); }
inline int64_t xs_hz(uint64_t& s) { return int64_t(int32_t(uint32_t(xs_next(s) >> 32))); }

double znorm(uint64_t& s) {
  int64_t hz = xs_hz(s);
  int iz = int(hz & 127);
  if (double(std::llabs(hz)) < ZIG.kn[iz]) return double(hz) * ZIG.wn[iz];
  const double r = 3.442619855899;
  for (;;) {
    double x = double(hz) * ZIG.wn[iz];
    if (iz == 0) {
      double y;
      do {
        x = -std::log(std::fmax(xs_unif(s), 1e-300)) * 0.2904764;
        y = -std::log(std::fmax(xs_unif(s), 1e-300));
      } while (y + y < x * x);
      return hz > 0 ? r + x : -r - x;
    }
    if (ZIG.fn[iz] + xs_unif(s) * (ZIG.fn[iz - 1] - ZIG.fn[iz]) < std::exp(-0.5 * x * x)) return x;
    hz = xs_hz(s);
    iz = int(hz & 127);
    if (double(std::llabs(hz)) < ZIG.kn[iz]) return double(hz) * ZIG.wn[iz];
  }
}

struct Model {
  int64_t N;
  const int64_t* ptr;
  const int32_t* nbr;
  const int32_t* typ;
  const double* table;
  int w;
  const double* A;   // N x 9 or null
  const double* mu;  // N
  double B[3];
  const uint8_t* pinned;  // N or null
};

inline void hfield(const Model& m, int64_t i, const double* e, double& hx, double& hy, double& hz) {
  hx = hy = hz = 0.0;
  const double* T = m.table;
  if (m.w == 1) {
    for (int64_t k = m.ptr[i]; k < m.ptr[i + 1]; ++k) {
      const double* ej = e + 3 * int64_t(m.nbr[k]);
      const double J = T[m.typ[k]];
      hx += J * ej[0];
      hy += J * ej[1];
      hz += J * ej[2];
    }
  } else if (m.w == 4) {
    for (int64_t k = m.ptr[i]; k < m.ptr[i + 1]; ++k) {
      const double* ej = e + 3 * int64_t(m.nbr[k]);
      const double* t = T + 4 * int64_t(m.typ[k]);
      hx += t[0] * ej[0] + (ej[1] * t[3] - ej[2] * t[2]);
      hy += t[0] * ej[1] + (ej[2] * t[1] - ej[0] * t[3]);
      hz += t[0] * ej[2] + (ej[0] * t[2] - ej[1] * t[1]);
    }
  } else {
    for (int64_t k = m.ptr[i]; k < m.ptr[i + 1]; ++k) {
      const double* ej = e + 3 * int64_t(m.nbr[k]);
      const double* t = T + 9 * int64_t(m.typ[k]);
      hx += t[0] * ej[0] + t[1] * ej[1] + t[2] * ej[2];
      hy += t[3] * ej[0] + t[4] * ej[1] + t[5] * ej[2];
      hz += t[6] * ej[0] + t[7] * ej[1] + t[8] * ej[2];
    }
  }
}

inline void field(const Model& m, int64_t i, const double* e, double* out) {
  double hx, hy, hz, gx = 0, gy = 0, gz = 0;
  hfield(m, i, e, hx, hy, hz);
  const double* x = e + 3 * i;
  if (m.A) {
    const double* a = m.A + 9 * i;
    gx = 2.0 * (a[0] * x[0] + a[1] * x[1] + a[2] * x[2]);
    gy = 2.0 * (a[3] * x[0] + a[4] * x[1] + a[5] * x[2]);
    gz = 2.0 * (a[6] * x[0] + a[7] * x[1] + a[8] * x[2]);
  }
  const double inv = 1.0 / (m.mu[i] * MU_B);
  out[0] = (2 * hx - gx) * inv + m.B[0];
  out[1] = (2 * hy - gy) * inv + m.B[1];
  out[2] = (2 * hz - gz) * inv + m.B[2];
}

inline void rhs(const double* e, const double* b, double a, int prec, double* f) {
  const double cx = e[1] * b[2] - e[2] * b[1], cy = e[2] * b[0] - e[0] * b[2], cz = e[0] * b[1] - e[1] * b[0];
  const double dx = e[1] * cz - e[2] * cy, dy = e[2] * cx - e[0] * cz, dz = e[0] * cy - e[1] * cx;
  const double pre = -GAMMA / (1.0 + a * a);
  if (prec) {
    f[0] = pre * (cx + a * dx);
    f[1] = pre * (cy + a * dy);
    f[2] = pre * (cz + a * dz);
  } else {
    f[0] = pre * a * dx;
    f[1] = pre * a * dy;
    f[2] = pre * a * dz;
  }
}

}  // namespace

extern "C" {

// Standard normals from the chunked ziggurat streams (fast.py:189-200): out (n x 3), st[nstreams].
void cpu_gauss3(int64_t n, double* out, uint64_t* st, int nstreams) {
  const int64_t chunk = (n + nstreams - 1) / nstreams;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < nstreams; ++c)
    for (int64_t i = c * chunk; i < n && i < (c + 1) * chunk; ++i)
      for (int a = 0; a < 3; ++a) out[3 * i + a] = znorm(st[c]);
}

// nsteps stochastic Heun steps on e (N x 3), two fused passes per step (fast.py:336-390).
// sig: per-site noise amplitude (T); noise: injected standard normals (nsteps x N x 3) or null (then the
// ziggurat streams st[256] are used when noisy != 0).  ep, k1, b: caller scratch (N x 3).
int cpu_heun_run(int64_t N, const int64_t* ptr, const int32_t* nbr, const int32_t* typ, const double* table,
                 int w, const double* A, const double* mu, const double* B, const uint8_t* pinned,
                 const double* alpha, const double* sig, double dt, int prec, int noisy, int64_t nsteps,
                 const double* noise, uint64_t* st, int nstreams, double* e, double* ep, double* k1, double* b) {
  Model m{N, ptr, nbr, typ, table, w, A, mu, {B[0], B[1], B[2]}, pinned};
  const int64_t chunk = (N + nstreams - 1) / nstreams;
  for (int64_t n = 0; n < nsteps; ++n) {
    const double* inj = noise ? noise + n * N * 3 : nullptr;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nstreams; ++c) {
      for (int64_t i = c * chunk; i < N && i < (c + 1) * chunk; ++i) {
        double* bi = b + 3 * i;
        double* ki = k1 + 3 * i;
        double* pi = ep + 3 * i;
        const double* ei = e + 3 * i;
        if (pinned && pinned[i]) {
          bi[0] = bi[1] = bi[2] = 0.0;
          ki[0] = ki[1] = ki[2] = 0.0;
          pi[0] = ei[0];
          pi[1] = ei[1];
          pi[2] = ei[2];
          continue;
        }
        if (noisy) {
          for (int a = 0; a < 3; ++a) bi[a] = sig[i] * (inj ? inj[3 * i + a] : znorm(st[c]));
        } else {
          bi[0] = bi[1] = bi[2] = 0.0;
        }
        double Bf[3];
        field(m, i, e, Bf);
        Bf[0] += bi[0];
        Bf[1] += bi[1];
        Bf[2] += bi[2];
        rhs(ei, Bf, alpha[i], prec, ki);
        const double x = ei[0] + dt * ki[0], y = ei[1] + dt * ki[1], z = ei[2] + dt * ki[2];
        const double q = 1.0 / std::sqrt(x * x + y * y + z * z);
        pi[0] = x * q;
        pi[1] = y * q;
        pi[2] = z * q;
      }
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
      if (pinned && pinned[i]) continue;
      double Bf[3], f[3];
      field(m, i, ep, Bf);
      Bf[0] += b[3 * i];
      Bf[1] += b[3 * i + 1];
      Bf[2] += b[3 * i + 2];
      rhs(ep + 3 * i, Bf, alpha[i], prec, f);
      const double x = e[3 * i] + 0.5 * dt * (k1[3 * i] + f[0]);
      const double y = e[3 * i + 1] + 0.5 * dt * (k1[3 * i + 1] + f[1]);
      const double z = e[3 * i + 2] + 0.5 * dt * (k1[3 * i + 2] + f[2]);
      const double q = 1.0 / std::sqrt(x * x + y * y + z * z);
      e[3 * i] = x * q;
      e[3 * i + 1] = y * q;
      e[3 * i + 2] = z * q;
    }
  }
  return 0;
}

int cpu_max_threads(void) { return omp_get_max_threads(); }
void cpu_set_threads(int n) { omp_set_num_threads(n); }

}  // extern "C"
