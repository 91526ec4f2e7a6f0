// gfx950 kernels of the atomistic LLG hot path: stencil exchange gather, fused Heun predictor /
// corrector, field, observables, counter-based thermal noise, topological charge.
//
// Data layout (HBM): spins are packed x,y,z per site (12 B FP32 / 24 B FP64: one dwordx3 load, every
// fetched byte used), sites in sublattice-major order s * ncell + cell with cells row-major (c2
// fastest).  Neighbour indices are never stored: a site's neighbours are cell offsets from the stencil
// in the kernel arguments, wrapped on periodic axes.  Couplings are either a small type table (kernel
// arguments, scalar loads) or per-bond tensors stored once per unordered bond, component-major
// (bond[(f * 9 + q) * ncell + cell]) so that a wavefront reads 64 consecutive floats per component.
//
// Workgroups: 256 threads, one site each; a workgroup never mixes sublattices, so the stencil loop is
// wavefront-uniform.  The block -> tile map is XCD-aware: the dispatcher deals workgroups round-robin
// to the 8 XCDs, so tile = xcd * (G / 8) + (block / 8) gives every XCD one contiguous slab of the
// lattice and the neighbour re-reads of that slab hit its own L2.
#include "mxllg_internal.h"

namespace mx {

// ------------------------------------------------------------------------------------------ helpers
__device__ __forceinline__ int xcd_tile(int b, int G) {
  const int x = b & 7, q = b >> 3, per = G >> 3, rem = G & 7;
  return x * per + min(x, rem) + q;
}

struct Site {
  int s, c, c0, c1, c2, i;
};

template <typename R>
__device__ __forceinline__ bool site_of(const KP<R>& p, Site& st) {
  const int b = xcd_tile(blockIdx.x, gridDim.x);
  st.s = b / p.bps;
  st.c = (b - st.s * p.bps) * BLOCK + threadIdx.x;
  if (st.c >= p.ncell) return false;
  st.c2 = st.c % p.L2;
  const int t = st.c / p.L2;
  st.c1 = t % p.L1;
  st.c0 = t / p.L1;
  st.i = st.s * p.ncell + st.c;
  return true;
}

// Philox4x32-10 (Salmon et al., SC'11): counter (c0..c3), key (k0, k1).
__device__ __forceinline__ void philox10(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3, uint32_t k0,
                                         uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
    const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
    const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// Three standard normals for (seed, step, site): 23-bit uniforms u = (2 (w >> 9) + 1) 2^-24 in (0, 1),
// exact in float; Box-Muller with the hardware log2 / sin / cos (v_sin_f32 and v_cos_f32 take their
// argument in revolutions, so cos(2 pi u) is one instruction).  Computed in float in both precisions,
// so FP32 and FP64 runs see identical noise.
__device__ __forceinline__ void gauss3(uint32_t k0, uint32_t k1, uint64_t step, uint32_t site, float& n0, float& n1,
                                       float& n2) {
  uint32_t c0 = site, c1 = uint32_t(step), c2 = uint32_t(step >> 32), c3 = 0u;
  philox10(c0, c1, c2, c3, k0, k1);
  const float s24 = 5.9604644775390625e-8f;  // 2^-24
  const float u0 = float(2u * (c0 >> 9) + 1u) * s24;
  const float u1 = float(2u * (c1 >> 9) + 1u) * s24;
  const float u2 = float(2u * (c2 >> 9) + 1u) * s24;
  const float u3 = float(2u * (c3 >> 9) + 1u) * s24;
  const float m2ln2 = -1.3862943611198906f;  // -2 ln 2: -2 ln u = -2 ln2 log2 u
  const float r0 = __builtin_sqrtf(fmaxf(m2ln2 * __builtin_amdgcn_logf(u0), 0.0f));
  const float r2 = __builtin_sqrtf(fmaxf(m2ln2 * __builtin_amdgcn_logf(u2), 0.0f));
  n0 = r0 * __builtin_amdgcn_cosf(u1);
  n1 = r0 * __builtin_amdgcn_sinf(u1);
  n2 = r2 * __builtin_amdgcn_cosf(u3);
}

template <typename R>
__device__ __forceinline__ R rnorm3(R x, R y, R z) {
  return R(1) / sqrt(x * x + y * y + z * z);
}

// h_i = sum_j J_ij e_j over the stencil of site st (magex/fast.py:86-120 semantics).
template <typename R, int CM>
__device__ __forceinline__ void exchange(const KP<R>& p, const R* __restrict__ e, const Site& st, R& hx, R& hy,
                                         R& hz) {
  hx = hy = hz = R(0);
  const int kb = p.sptr[st.s], ke = p.sptr[st.s + 1];
  for (int k = kb; k < ke; ++k) {
    const Ent en = p.ent[k];
    int n0 = st.c0 + int(int16_t(en.d01 & 0xffff));
    int n1 = st.c1 + (en.d01 >> 16);
    int n2 = st.c2 + en.d2;
    bool ok = true;
    if (p.pbc0) n0 -= (n0 >= p.L0) ? p.L0 : 0; else ok = ok && unsigned(n0) < unsigned(p.L0);
    if (p.pbc1) n1 -= (n1 >= p.L1) ? p.L1 : 0; else ok = ok && unsigned(n1) < unsigned(p.L1);
    if (p.pbc2) n2 -= (n2 >= p.L2) ? p.L2 : 0; else ok = ok && unsigned(n2) < unsigned(p.L2);
    if (!ok) continue;
    const int cn = (n0 * p.L1 + n1) * p.L2 + n2;
    const int j = (en.tb & 0xffff) * p.ncell + cn;
    const R ex = e[3 * j], ey = e[3 * j + 1], ez = e[3 * j + 2];
    if constexpr (CM == COUP_TABLE1) {
      const R J = p.tab[en.slot];
      hx += J * ex;
      hy += J * ey;
      hz += J * ez;
    } else if constexpr (CM == COUP_TABLE4) {
      // J I + antisymmetric part: (J^A e_j) = e_j x D
      const R* T = p.tab + 4 * en.slot;
      hx += T[0] * ex + (ey * T[3] - ez * T[2]);
      hy += T[0] * ey + (ez * T[1] - ex * T[3]);
      hz += T[0] * ez + (ex * T[2] - ey * T[1]);
    } else if constexpr (CM == COUP_TABLE9) {
      const R* T = p.tab + 9 * en.slot;
      hx += T[0] * ex + T[1] * ey + T[2] * ez;
      hy += T[3] * ex + T[4] * ey + T[5] * ez;
      hz += T[6] * ex + T[7] * ey + T[8] * ez;
    } else {
      // per-bond tensor stored once per unordered bond: forward entries read their own cell's copy,
      // backward entries the partner's copy at the neighbour cell, transposed (J_ji = J_ij^T)
      const bool back = (en.tb >> 16) != 0;
      const R* Jb = p.bond + size_t(en.slot) * 9 * p.ncell + (back ? cn : st.c);
      const int n = p.ncell;
      const R J0 = Jb[0], J1 = Jb[n], J2 = Jb[2 * n], J3 = Jb[3 * n], J4 = Jb[4 * n], J5 = Jb[5 * n],
              J6 = Jb[6 * n], J7 = Jb[7 * n], J8 = Jb[8 * n];
      if (!back) {
        hx += J0 * ex + J1 * ey + J2 * ez;
        hy += J3 * ex + J4 * ey + J5 * ez;
        hz += J6 * ex + J7 * ey + J8 * ez;
      } else {
        hx += J0 * ex + J3 * ey + J6 * ez;
        hy += J1 * ex + J4 * ey + J7 * ez;
        hz += J2 * ex + J5 * ey + J8 * ez;
      }
    }
  }
}

// On-site gradient dE/de (magex/fast.py:39-64): 2 A e + cubic + higher uniaxial.
template <typename R>
__device__ __forceinline__ void onsite_grad(const KP<R>& p, const Site& st, R x, R y, R z, R& gx, R& gy, R& gz) {
  gx = gy = gz = R(0);
  if (p.amode == ONSITE_SUB) {
    if (p.has_A) {
      const R* a = p.Asub[st.s];
      gx = R(2) * (a[0] * x + a[1] * y + a[2] * z);
      gy = R(2) * (a[3] * x + a[4] * y + a[5] * z);
      gz = R(2) * (a[6] * x + a[7] * y + a[8] * z);
    }
  } else if (p.amode == ONSITE_SITE6) {
    const int N = p.N, i = st.i;
    const R axx = p.A[i], ayy = p.A[N + i], azz = p.A[2 * N + i];
    const R axy = p.A[3 * N + i], axz = p.A[4 * N + i], ayz = p.A[5 * N + i];
    gx = R(2) * (axx * x + axy * y + axz * z);
    gy = R(2) * (axy * x + ayy * y + ayz * z);
    gz = R(2) * (axz * x + ayz * y + azz * z);
  } else {
    const int N = p.N, i = st.i;
    const R* a = p.A + i;
    gx = R(2) * (a[0] * x + a[N] * y + a[2 * N] * z);
    gy = R(2) * (a[3 * N] * x + a[4 * N] * y + a[5 * N] * z);
    gz = R(2) * (a[6 * N] * x + a[7 * N] * y + a[8 * N] * z);
  }
  if (p.has_c) {
    const R* C = p.C;
    const R a1 = C[0] * x + C[1] * y + C[2] * z;
    const R a2 = C[3] * x + C[4] * y + C[5] * z;
    const R a3 = C[6] * x + C[7] * y + C[8] * z;
    const R q1 = a1 * a1, q2 = a2 * a2, q3 = a3 * a3, kc = p.Kc[st.i];
    const R d1 = R(2) * a1 * (q2 + q3) * kc, d2 = R(2) * a2 * (q1 + q3) * kc, d3 = R(2) * a3 * (q1 + q2) * kc;
    gx += d1 * C[0] + d2 * C[3] + d3 * C[6];
    gy += d1 * C[1] + d2 * C[4] + d3 * C[7];
    gz += d1 * C[2] + d2 * C[5] + d3 * C[8];
  }
  if (p.has_u) {
    const R c = x * p.u[0] + y * p.u[1] + z * p.u[2];
    const R c2 = c * c;
    const R f = R(-4) * p.K2[st.i] * c2 * c - R(6) * p.K3[st.i] * c2 * c2 * c;
    gx += f * p.u[0];
    gy += f * p.u[1];
    gz += f * p.u[2];
  }
}

// On-site energy (magex/fast.py:66-84).
template <typename R>
__device__ __forceinline__ R onsite_energy(const KP<R>& p, const Site& st, R x, R y, R z) {
  R E = R(0);
  if (p.amode == ONSITE_SUB) {
    if (p.has_A) {
      const R* a = p.Asub[st.s];
      E = x * (a[0] * x + a[1] * y + a[2] * z) + y * (a[3] * x + a[4] * y + a[5] * z) +
          z * (a[6] * x + a[7] * y + a[8] * z);
    }
  } else if (p.amode == ONSITE_SITE6) {
    const int N = p.N, i = st.i;
    const R axx = p.A[i], ayy = p.A[N + i], azz = p.A[2 * N + i];
    const R axy = p.A[3 * N + i], axz = p.A[4 * N + i], ayz = p.A[5 * N + i];
    E = axx * x * x + ayy * y * y + azz * z * z + R(2) * (axy * x * y + axz * x * z + ayz * y * z);
  } else {
    const int N = p.N;
    const R* a = p.A + st.i;
    E = x * (a[0] * x + a[N] * y + a[2 * N] * z) + y * (a[3 * N] * x + a[4 * N] * y + a[5 * N] * z) +
        z * (a[6 * N] * x + a[7 * N] * y + a[8 * N] * z);
  }
  if (p.has_c) {
    const R* C = p.C;
    const R a1 = C[0] * x + C[1] * y + C[2] * z;
    const R a2 = C[3] * x + C[4] * y + C[5] * z;
    const R a3 = C[6] * x + C[7] * y + C[8] * z;
    const R q1 = a1 * a1, q2 = a2 * a2, q3 = a3 * a3;
    E += p.Kc[st.i] * (q1 * q2 + q2 * q3 + q3 * q1);
  }
  if (p.has_u) {
    const R c = x * p.u[0] + y * p.u[1] + z * p.u[2];
    const R c4 = c * c * c * c;
    E += -p.K2[st.i] * c4 - p.K3[st.i] * c4 * c * c;
  }
  return E;
}

// de/dt = -gamma/(1+a^2) [e x B + a e x (e x B)] (magex/llg.py:21-28, fast.py:225-236).
template <typename R>
__device__ __forceinline__ void llg_rhs(R ex, R ey, R ez, R bx, R by, R bz, R a, int prec, R& fx, R& fy, R& fz) {
  const R cx = ey * bz - ez * by, cy = ez * bx - ex * bz, cz = ex * by - ey * bx;
  const R dx = ey * cz - ez * cy, dy = ez * cx - ex * cz, dz = ex * cy - ey * cx;
  const R pre = R(-GAMMA) / (R(1) + a * a);
  if (prec) {
    fx = pre * (cx + a * dx);
    fy = pre * (cy + a * dy);
    fz = pre * (cz + a * dz);
  } else {
    fx = pre * a * dx;
    fy = pre * a * dy;
    fz = pre * a * dz;
  }
}

template <typename R>
__device__ __forceinline__ void thermal(const KP<R>& p, const Site& st, uint64_t step, R& bx, R& by, R& bz) {
  const R sg = p.sigma ? p.sigma[st.i] : p.sigma_sub[st.s];
  const int iref = st.c * p.ns + st.s;
  if (p.noise) {
    bx = sg * p.noise[3 * iref];
    by = sg * p.noise[3 * iref + 1];
    bz = sg * p.noise[3 * iref + 2];
  } else {
    float n0, n1, n2;
    gauss3(p.seed_lo, p.seed_hi, step, uint32_t(iref), n0, n1, n2);
    bx = sg * R(n0);
    by = sg * R(n1);
    bz = sg * R(n2);
  }
}

template <typename R>
__device__ __forceinline__ R inv_mu_of(const KP<R>& p, const Site& st) {
  return p.inv_mu ? p.inv_mu[st.i] : p.inv_mu_sub[st.s];
}

// ------------------------------------------------------------------------------------------ Heun
// PASS 1 (magex/fast.py:336-368): noise + field(e) + predictor ep = n(e + dt k1), keeps k1.
// PASS 2 (magex/fast.py:370-390): field(ep) + corrector e = n(e + dt/2 (k1 + f(ep))), in place.
template <typename R, int CM, int PASS>
__global__ __launch_bounds__(BLOCK) void k_heun(const KP<R> p, R* __restrict__ e, R* __restrict__ ep,
                                                R* __restrict__ k1, uint64_t step) {
  Site st;
  if (!site_of(p, st)) return;
  const int i = st.i;
  const bool pin = p.pinned && p.pinned[i];
  if (PASS == 1) {
    const R x = e[3 * i], y = e[3 * i + 1], z = e[3 * i + 2];
    if (pin) {
      ep[3 * i] = x;
      ep[3 * i + 1] = y;
      ep[3 * i + 2] = z;
      k1[3 * i] = k1[3 * i + 1] = k1[3 * i + 2] = R(0);
      return;
    }
    R hx, hy, hz, gx, gy, gz;
    exchange<R, CM>(p, e, st, hx, hy, hz);
    onsite_grad(p, st, x, y, z, gx, gy, gz);
    const R im = inv_mu_of(p, st);
    R bx = (R(2) * hx - gx) * im + p.B[0], by = (R(2) * hy - gy) * im + p.B[1], bz = (R(2) * hz - gz) * im + p.B[2];
    if (p.noisy) {
      R nx, ny, nz;
      thermal(p, st, step, nx, ny, nz);
      bx += nx;
      by += ny;
      bz += nz;
    }
    const R a = p.alpha ? p.alpha[i] : p.alpha_sub[st.s];
    R fx, fy, fz;
    llg_rhs(x, y, z, bx, by, bz, a, p.prec, fx, fy, fz);
    k1[3 * i] = fx;
    k1[3 * i + 1] = fy;
    k1[3 * i + 2] = fz;
    const R px = x + p.dt * fx, py = y + p.dt * fy, pz = z + p.dt * fz;
    const R q = rnorm3(px, py, pz);
    ep[3 * i] = px * q;
    ep[3 * i + 1] = py * q;
    ep[3 * i + 2] = pz * q;
  } else {
    if (pin) return;
    const R x = ep[3 * i], y = ep[3 * i + 1], z = ep[3 * i + 2];
    R hx, hy, hz, gx, gy, gz;
    exchange<R, CM>(p, ep, st, hx, hy, hz);
    onsite_grad(p, st, x, y, z, gx, gy, gz);
    const R im = inv_mu_of(p, st);
    R bx = (R(2) * hx - gx) * im + p.B[0], by = (R(2) * hy - gy) * im + p.B[1], bz = (R(2) * hz - gz) * im + p.B[2];
    if (p.noisy) {
      R nx, ny, nz;
      thermal(p, st, step, nx, ny, nz);
      bx += nx;
      by += ny;
      bz += nz;
    }
    const R a = p.alpha ? p.alpha[i] : p.alpha_sub[st.s];
    R fx, fy, fz;
    llg_rhs(x, y, z, bx, by, bz, a, p.prec, fx, fy, fz);
    const R hdt = R(0.5) * p.dt;
    const R cx = e[3 * i] + hdt * (k1[3 * i] + fx);
    const R cy = e[3 * i + 1] + hdt * (k1[3 * i + 1] + fy);
    const R cz = e[3 * i + 2] + hdt * (k1[3 * i + 2] + fz);
    const R q = rnorm3(cx, cy, cz);
    e[3 * i] = cx * q;
    e[3 * i + 1] = cy * q;
    e[3 * i + 2] = cz * q;
  }
}

// ------------------------------------------------------------------------------------------ field
// kind: MX_FIELD_BEFF (T), MX_FIELD_GRADIENT (meV), MX_FIELD_EXCHANGE (meV); output in magex order.
template <typename R, int CM>
__global__ __launch_bounds__(BLOCK) void k_field(const KP<R> p, const R* __restrict__ e, R* __restrict__ out,
                                                 int kind) {
  Site st;
  if (!site_of(p, st)) return;
  const int i = st.i, o = 3 * (st.c * p.ns + st.s);
  R hx, hy, hz;
  exchange<R, CM>(p, e, st, hx, hy, hz);
  if (kind == MX_FIELD_EXCHANGE) {
    out[o] = hx;
    out[o + 1] = hy;
    out[o + 2] = hz;
    return;
  }
  if (p.pinned && p.pinned[i]) {
    out[o] = out[o + 1] = out[o + 2] = R(0);
    return;
  }
  R gx, gy, gz;
  onsite_grad(p, st, e[3 * i], e[3 * i + 1], e[3 * i + 2], gx, gy, gz);
  const R im = inv_mu_of(p, st);
  if (kind == MX_FIELD_BEFF) {
    out[o] = (R(2) * hx - gx) * im + p.B[0];
    out[o + 1] = (R(2) * hy - gy) * im + p.B[1];
    out[o + 2] = (R(2) * hz - gz) * im + p.B[2];
  } else {
    const R mm = R(1) / im;  // mu mu_B
    out[o] = gx - R(2) * hx - mm * p.B[0];
    out[o + 1] = gy - R(2) * hy - mm * p.B[1];
    out[o + 2] = gz - R(2) * hz - mm * p.B[2];
  }
}

// ------------------------------------------------------------------------------------------ reductions
// Wavefront (64-lane) sum / max by xor shuffles, then across the 4 wavefronts of the block in LDS,
// in a fixed order (bitwise reproducible).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

template <int NV>
__device__ __forceinline__ void block_reduce(double (&v)[NV], int nmax_from, double* __restrict__ dst) {
  __shared__ double red[BLOCK / 64][NV];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = (k >= nmax_from) ? wave_max(v[k]) : wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) red[w][k] = v[k];
  __syncthreads();
  if (threadIdx.x < NV) {
    const int k = threadIdx.x;
    double a = red[0][k];
    for (int q = 1; q < BLOCK / 64; ++q) a = (k >= nmax_from) ? fmax(a, red[q][k]) : a + red[q][k];
    dst[k] = a;
  }
}

// Per-block partials of [E, sum e (3), sum s e (3), max |e x B_eff|] (spinmodel.py:475-484, 517-524;
// llg.py:250-252).
template <typename R, int CM>
__global__ __launch_bounds__(BLOCK) void k_obs(const KP<R> p, const R* __restrict__ e, double* __restrict__ part) {
  double v[MX_OBS_LEN] = {0, 0, 0, 0, 0, 0, 0, 0};
  Site st;
  if (site_of(p, st)) {
    const int i = st.i;
    const R x = e[3 * i], y = e[3 * i + 1], z = e[3 * i + 2];
    R hx, hy, hz, gx, gy, gz;
    exchange<R, CM>(p, e, st, hx, hy, hz);
    onsite_grad(p, st, x, y, z, gx, gy, gz);
    const R im = inv_mu_of(p, st);
    const R Eon = onsite_energy(p, st, x, y, z);
    v[0] = double(Eon) - double(x * hx + y * hy + z * hz) - double(x * p.B[0] + y * p.B[1] + z * p.B[2]) / double(im);
    v[1] = x;
    v[2] = y;
    v[3] = z;
    const R sg = p.sign_site ? p.sign_site[i] : p.sign_sub[st.s];
    v[4] = sg * x;
    v[5] = sg * y;
    v[6] = sg * z;
    if (!(p.pinned && p.pinned[i])) {
      const R bx = (R(2) * hx - gx) * im + p.B[0], by = (R(2) * hy - gy) * im + p.B[1],
              bz = (R(2) * hz - gz) * im + p.B[2];
      const double tx = y * bz - z * by, ty = z * bx - x * bz, tz = x * by - y * bx;
      v[7] = sqrt(tx * tx + ty * ty + tz * tz);
    }
  }
  block_reduce<MX_OBS_LEN>(v, 7, part + size_t(blockIdx.x) * MX_OBS_LEN);
}

// Second stage: one workgroup folds the per-block partials in a fixed order, then scales them.
template <int NV>
__global__ __launch_bounds__(BLOCK) void k_final(const double* __restrict__ part, int nblk, int nmax_from,
                                                 double scale, double* __restrict__ out) {
  __shared__ double res[NV];
  double v[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = 0.0;
  for (int b = threadIdx.x; b < nblk; b += BLOCK)
#pragma unroll
    for (int k = 0; k < NV; ++k)
      v[k] = (k >= nmax_from) ? fmax(v[k], part[size_t(b) * NV + k]) : v[k] + part[size_t(b) * NV + k];
  block_reduce<NV>(v, nmax_from, res);
  __syncthreads();
  if (threadIdx.x < NV) out[threadIdx.x] = res[threadIdx.x] * scale;
}

// ------------------------------------------------------------------------------------------ misc
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_noise(const KP<R> p, R* __restrict__ out, uint64_t step) {
  const int r = blockIdx.x * BLOCK + threadIdx.x;
  if (r >= p.N) return;
  float n0, n1, n2;
  gauss3(p.seed_lo, p.seed_hi, step, uint32_t(r), n0, n1, n2);
  out[3 * r] = R(n0);
  out[3 * r + 1] = R(n1);
  out[3 * r + 2] = R(n2);
}

// magex order (c * ns + s) <-> internal order (s * ncell + c)
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_permute(const R* __restrict__ src, R* __restrict__ dst, int ncell, int ns,
                                                   int to_internal) {
  const int r = blockIdx.x * BLOCK + threadIdx.x;
  if (r >= ncell * ns) return;
  const int c = r / ns, s = r - c * ns, q = s * ncell + c;
  const int a = to_internal ? r : q, b = to_internal ? q : r;
  dst[3 * b] = src[3 * a];
  dst[3 * b + 1] = src[3 * a + 1];
  dst[3 * b + 2] = src[3 * a + 2];
}

template <typename R>
__global__ __launch_bounds__(BLOCK) void k_rhs(const R* __restrict__ e, const R* __restrict__ B, R a, int prec, int n,
                                               R* __restrict__ out) {
  const int r = blockIdx.x * BLOCK + threadIdx.x;
  if (r >= n) return;
  R fx, fy, fz;
  llg_rhs(e[3 * r], e[3 * r + 1], e[3 * r + 2], B[3 * r], B[3 * r + 1], B[3 * r + 2], a, prec, fx, fy, fz);
  out[3 * r] = fx;
  out[3 * r + 1] = fy;
  out[3 * r + 2] = fz;
}

// Berg-Luscher solid angles of counter-clockwise triangles (magex/texture.py:58-63), per-block sums.
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_topo(const R* __restrict__ e, const int32_t* __restrict__ tri, int64_t ntri,
                                                double* __restrict__ part) {
  double v[1] = {0.0};
  const int64_t t = int64_t(blockIdx.x) * BLOCK + threadIdx.x;
  if (t < ntri) {
    const int a = tri[3 * t], b = tri[3 * t + 1], c = tri[3 * t + 2];
    const double s1x = e[3 * a], s1y = e[3 * a + 1], s1z = e[3 * a + 2];
    const double s2x = e[3 * b], s2y = e[3 * b + 1], s2z = e[3 * b + 2];
    const double s3x = e[3 * c], s3y = e[3 * c + 1], s3z = e[3 * c + 2];
    const double num = s1x * (s2y * s3z - s2z * s3y) + s1y * (s2z * s3x - s2x * s3z) + s1z * (s2x * s3y - s2y * s3x);
    const double den = 1.0 + (s1x * s2x + s1y * s2y + s1z * s2z) + (s2x * s3x + s2y * s3y + s2z * s3z) +
                       (s3x * s1x + s3y * s1y + s3z * s1z);
    v[0] = 2.0 * atan2(num, den);
  }
  block_reduce<1>(v, 1, part + blockIdx.x);
}

// ------------------------------------------------------------------------------------------ launchers
static int check_launch(const char* what) {
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    set_error(std::string(what) + ": " + hipGetErrorString(err));
    return MX_EHIP;
  }
  return MX_OK;
}

template <typename R>
void fill_kp(const Model& m, KP<R>& p) {
  p.L0 = m.L[0];
  p.L1 = m.L[1];
  p.L2 = m.L[2];
  p.ncell = m.ncell;
  p.ns = m.ns;
  p.N = m.N;
  p.bps = (m.ncell + BLOCK - 1) / BLOCK;
  p.nblk = p.bps * m.ns;
  p.pbc0 = m.pbc[0];
  p.pbc1 = m.pbc[1];
  p.pbc2 = m.pbc[2];
  for (int s = 0; s <= MAXSUB; ++s) p.sptr[s] = s <= m.ns ? m.sptr[s] : m.sptr[m.ns];
  for (int k = 0; k < MAXENT; ++k) p.ent[k] = k < m.nent ? m.ent[k] : Ent{0, 0, 0, 0};
  for (int k = 0; k < MAXTAB; ++k) p.tab[k] = k < int(m.table.size()) ? R(m.table[k]) : R(0);
  p.bond = static_cast<const R*>(m.bond_d);
  p.amode = m.amode;
  p.has_A = m.has_A;
  p.A = static_cast<const R*>(m.A_d);
  for (int s = 0; s < MAXSUB; ++s)
    for (int q = 0; q < 9; ++q) p.Asub[s][q] = s < m.ns ? R(m.Asub[9 * s + q]) : R(0);
  p.has_c = m.has_c;
  p.has_u = m.has_u;
  p.Kc = static_cast<const R*>(m.Kc_d);
  p.K2 = static_cast<const R*>(m.K2_d);
  p.K3 = static_cast<const R*>(m.K3_d);
  for (int q = 0; q < 9; ++q) p.C[q] = R(m.C[q]);
  for (int q = 0; q < 3; ++q) p.u[q] = R(m.u[q]);
  p.inv_mu = static_cast<const R*>(m.inv_mu_d);
  p.alpha = static_cast<const R*>(m.alpha_d);
  p.sigma = static_cast<const R*>(m.sigma_d);
  p.sign_site = static_cast<const R*>(m.sign_d);
  for (int s = 0; s < MAXSUB; ++s) {
    const bool ok = s < m.ns;
    p.inv_mu_sub[s] = ok ? R(1.0 / (m.mu_sub[s] * MU_B)) : R(0);
    p.sign_sub[s] = ok ? R(m.sign_sub[s]) : R(0);
    p.alpha_sub[s] = ok && !m.alpha_sub_v.empty() ? R(m.alpha_sub_v[s]) : R(0);
    p.sigma_sub[s] = ok && !m.sigma_sub_v.empty() ? R(m.sigma_sub_v[s]) : R(0);
  }
  p.pinned = static_cast<const uint8_t*>(m.pinned_d);
  for (int q = 0; q < 3; ++q) p.B[q] = R(m.B[q]);
  p.dt = R(m.dt);
  p.prec = m.precession;
  p.noisy = m.noisy;
  p.seed_lo = p.seed_hi = 0;
  p.noise = nullptr;
}

static_assert(sizeof(KP<double>) <= 4000, "kernel arguments must stay below 4 KB");

template <typename R, int CM>
static int heun_cm(const KP<R>& p, R* e, R* ep, R* k1, uint64_t step, hipStream_t s) {
  hipLaunchKernelGGL((k_heun<R, CM, 1>), dim3(p.nblk), dim3(BLOCK), 0, s, p, e, ep, k1, step);
  hipLaunchKernelGGL((k_heun<R, CM, 2>), dim3(p.nblk), dim3(BLOCK), 0, s, p, e, ep, k1, step);
  return check_launch("heun step");
}

template <typename R>
static int heun_t(const Model& m, void* e, uint64_t seed, uint64_t step, const void* noise, hipStream_t s) {
  KP<R> p;
  fill_kp(m, p);
  p.seed_lo = uint32_t(seed);
  p.seed_hi = uint32_t(seed >> 32);
  p.noise = static_cast<const R*>(noise);
  R* E = static_cast<R*>(e);
  R* ep = static_cast<R*>(m.ep_d);
  R* k1 = static_cast<R*>(m.k1_d);
  switch (m.coup) {
    case COUP_TABLE1: return heun_cm<R, COUP_TABLE1>(p, E, ep, k1, step, s);
    case COUP_TABLE4: return heun_cm<R, COUP_TABLE4>(p, E, ep, k1, step, s);
    case COUP_TABLE9: return heun_cm<R, COUP_TABLE9>(p, E, ep, k1, step, s);
    default: return heun_cm<R, COUP_BOND>(p, E, ep, k1, step, s);
  }
}

int launch_heun_step(const Model& m, void* e, uint64_t seed, uint64_t step, const void* noise, hipStream_t s) {
  return m.prec == MX_FP64 ? heun_t<double>(m, e, seed, step, noise, s) : heun_t<float>(m, e, seed, step, noise, s);
}

template <typename R>
static int field_t(const Model& m, const void* e, void* out, int kind, hipStream_t s) {
  KP<R> p;
  fill_kp(m, p);
  const R* E = static_cast<const R*>(e);
  R* O = static_cast<R*>(out);
  switch (m.coup) {
    case COUP_TABLE1: hipLaunchKernelGGL((k_field<R, COUP_TABLE1>), dim3(p.nblk), dim3(BLOCK), 0, s, p, E, O, kind); break;
    case COUP_TABLE4: hipLaunchKernelGGL((k_field<R, COUP_TABLE4>), dim3(p.nblk), dim3(BLOCK), 0, s, p, E, O, kind); break;
    case COUP_TABLE9: hipLaunchKernelGGL((k_field<R, COUP_TABLE9>), dim3(p.nblk), dim3(BLOCK), 0, s, p, E, O, kind); break;
    default: hipLaunchKernelGGL((k_field<R, COUP_BOND>), dim3(p.nblk), dim3(BLOCK), 0, s, p, E, O, kind); break;
  }
  return check_launch("field");
}

int launch_field(const Model& m, const void* e, void* out, int kind, hipStream_t s) {
  return m.prec == MX_FP64 ? field_t<double>(m, e, out, kind, s) : field_t<float>(m, e, out, kind, s);
}

template <typename R>
static int obs_t(const Model& m, const void* e, double* out, hipStream_t s) {
  KP<R> p;
  fill_kp(m, p);
  const R* E = static_cast<const R*>(e);
  switch (m.coup) {
    case COUP_TABLE1: hipLaunchKernelGGL((k_obs<R, COUP_TABLE1>), dim3(p.nblk), dim3(BLOCK), 0, s, p, E, m.part_d); break;
    case COUP_TABLE4: hipLaunchKernelGGL((k_obs<R, COUP_TABLE4>), dim3(p.nblk), dim3(BLOCK), 0, s, p, E, m.part_d); break;
    case COUP_TABLE9: hipLaunchKernelGGL((k_obs<R, COUP_TABLE9>), dim3(p.nblk), dim3(BLOCK), 0, s, p, E, m.part_d); break;
    default: hipLaunchKernelGGL((k_obs<R, COUP_BOND>), dim3(p.nblk), dim3(BLOCK), 0, s, p, E, m.part_d); break;
  }
  hipLaunchKernelGGL((k_final<MX_OBS_LEN>), dim3(1), dim3(BLOCK), 0, s, m.part_d, p.nblk, 7, 1.0, out);
  return check_launch("observables");
}

int launch_observables(const Model& m, const void* e, double* out, hipStream_t s) {
  return m.prec == MX_FP64 ? obs_t<double>(m, e, out, s) : obs_t<float>(m, e, out, s);
}

template <typename R>
static int noise_t(const Model& m, void* out, uint64_t seed, uint64_t step, hipStream_t s) {
  KP<R> p;
  fill_kp(m, p);
  p.seed_lo = uint32_t(seed);
  p.seed_hi = uint32_t(seed >> 32);
  hipLaunchKernelGGL((k_noise<R>), dim3((m.N + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, p, static_cast<R*>(out), step);
  return check_launch("noise");
}

int launch_noise(const Model& m, void* out, uint64_t seed, uint64_t step, hipStream_t s) {
  return m.prec == MX_FP64 ? noise_t<double>(m, out, seed, step, s) : noise_t<float>(m, out, seed, step, s);
}

int launch_permute(const Model& m, const void* src, void* dst, int to_internal, hipStream_t s) {
  const dim3 g((m.N + BLOCK - 1) / BLOCK);
  if (m.prec == MX_FP64)
    hipLaunchKernelGGL((k_permute<double>), g, dim3(BLOCK), 0, s, static_cast<const double*>(src),
                       static_cast<double*>(dst), m.ncell, m.ns, to_internal);
  else
    hipLaunchKernelGGL((k_permute<float>), g, dim3(BLOCK), 0, s, static_cast<const float*>(src),
                       static_cast<float*>(dst), m.ncell, m.ns, to_internal);
  return check_launch("permute");
}

int launch_rhs(const Model& m, const void* e, const void* B, double alpha, int prec, void* out, hipStream_t s) {
  const dim3 g((m.N + BLOCK - 1) / BLOCK);
  if (m.prec == MX_FP64)
    hipLaunchKernelGGL((k_rhs<double>), g, dim3(BLOCK), 0, s, static_cast<const double*>(e),
                       static_cast<const double*>(B), alpha, prec, m.N, static_cast<double*>(out));
  else
    hipLaunchKernelGGL((k_rhs<float>), g, dim3(BLOCK), 0, s, static_cast<const float*>(e),
                       static_cast<const float*>(B), float(alpha), prec, m.N, static_cast<float*>(out));
  return check_launch("rhs");
}

int launch_topo(const Model& m, const void* e, const int32_t* tri, int64_t ntri, double hand, double* out,
                hipStream_t s) {
  const int nblk = int((ntri + BLOCK - 1) / BLOCK);
  if (nblk > m.npart) {
    set_error("topological charge: more triangles than the model's reduction buffer holds");
    return MX_EINVAL;
  }
  if (m.prec == MX_FP64)
    hipLaunchKernelGGL((k_topo<double>), dim3(nblk), dim3(BLOCK), 0, s, static_cast<const double*>(e), tri, ntri,
                       m.part_d);
  else
    hipLaunchKernelGGL((k_topo<float>), dim3(nblk), dim3(BLOCK), 0, s, static_cast<const float*>(e), tri, ntri,
                       m.part_d);
  hipLaunchKernelGGL((k_final<1>), dim3(1), dim3(BLOCK), 0, s, m.part_d, nblk, 1, hand / (4.0 * M_PI), out);
  return check_launch("topological charge");
}

template void fill_kp<double>(const Model&, KP<double>&);
template void fill_kp<float>(const Model&, KP<float>&);

}  // namespace mx
