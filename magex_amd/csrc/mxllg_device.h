// Device helpers shared by the kernel sources: XCD-aware block map, Philox noise, LLG right-hand side,
// wavefront / workgroup reductions.
#pragma once
#include "mxllg_internal.h"

namespace mx {

// The dispatcher deals workgroups round-robin to the 8 XCDs (XCD = block mod 8, checked on MI355X);
// this map gives XCD x the contiguous logical tiles [x G/8, (x+1) G/8), so each XCD's L2 sees one slab.
// With rev, each XCD walks its slab from the far end: a kernel that follows one which walked forward
// starts on the tiles whose data the previous kernel left in that XCD's L2 (L2 contents survive a
// kernel boundary), so alternating directions saves about one L2 of refetch per launch.
__device__ __forceinline__ int xcd_tile(int b, int G, bool rev = false) {
  const int x = b & 7, q = b >> 3, per = G >> 3, rem = G & 7;
  const int cnt = per + (x < rem ? 1 : 0);
  return x * per + min(x, rem) + (rev ? cnt - 1 - q : q);
}

// Philox4x32-10 (Salmon et al., SC'11): counter (c0..c3), key (k0, k1).
__device__ __forceinline__ void philox10(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3, uint32_t k0,
                                         uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    // one 32 x 32 -> 64-bit product per multiplier (v_mad_u64_u32) gives both halves
    const uint64_t p0 = uint64_t(0xD2511F53u) * c0, p1 = uint64_t(0xCD9E8D57u) * c2;
    const uint32_t lo0 = uint32_t(p0), hi0 = uint32_t(p0 >> 32);
    const uint32_t lo1 = uint32_t(p1), hi1 = uint32_t(p1 >> 32);
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

// Three standard normals for (seed, step, site) at double precision: two Philox4x32-10 outputs (fourth
// counter word 0 and 1) give four 64-bit words; the top 52 bits k of each make an odd-multiple uniform
// u = (2 k + 1) 2^-53 in (0, 1), exact in double (tails to sqrt(106 ln 2) = 8.6 sigma instead of float's
// 5.8); Box-Muller (1958) with double log and sincospi (sin / cos of pi x, so cos(2 pi u) = cospi(2 u)).
__device__ __forceinline__ void gauss3d(uint32_t k0, uint32_t k1, uint64_t step, uint32_t site, double& n0,
                                        double& n1, double& n2) {
  uint32_t a0 = site, a1 = uint32_t(step), a2 = uint32_t(step >> 32), a3 = 0u;
  uint32_t b0 = site, b1 = a1, b2 = a2, b3 = 1u;
  philox10(a0, a1, a2, a3, k0, k1);
  philox10(b0, b1, b2, b3, k0, k1);
  const double s53 = 1.1102230246251565404e-16;  // 2^-53
  const double u0 = double(2 * ((uint64_t(a0) << 20) | (a1 >> 12)) + 1) * s53;
  const double u1 = double(2 * ((uint64_t(a2) << 20) | (a3 >> 12)) + 1) * s53;
  const double u2 = double(2 * ((uint64_t(b0) << 20) | (b1 >> 12)) + 1) * s53;
  const double u3 = double(2 * ((uint64_t(b2) << 20) | (b3 >> 12)) + 1) * s53;
  const double r0 = sqrt(-2.0 * log(u0)), r2 = sqrt(-2.0 * log(u2));
  double s1, c1, s3, c3;
  sincospi(2.0 * u1, &s1, &c1);
  sincospi(2.0 * u3, &s3, &c3);
  n0 = r0 * c1;
  n1 = r0 * s1;
  n2 = r2 * c3;
}

// floor(n / d) for 0 <= n < 2^24 and d >= 1, with rcp = 1/d rounded to float: the float product is
// within one unit of the quotient, and the two integer corrections make it exact.
__device__ __forceinline__ int fdiv(int n, int d, float rcp) {
  int q = int(float(n) * rcp);
  q -= q * d > n ? 1 : 0;
  q += (q + 1) * d <= n ? 1 : 0;
  return q;
}

template <typename R>
__device__ __forceinline__ R rnorm3(R x, R y, R z) {
  return R(1) / sqrt(x * x + y * y + z * z);
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

// Rodrigues rotation of (x, y, z) by the rotation vector (wx, wy, wz) (angle |w|); the identity below
// |w| = 1e-14 (magex/fast.py:271-282, llg.py:31-36).  1 - cos is formed as 2 sin^2(th / 2), which
// keeps its relative accuracy at the small angles of a time step (it cancels in FP32 otherwise).
template <typename R>
__device__ __forceinline__ void rotate(R& x, R& y, R& z, R wx, R wy, R wz) {
  const R th = sqrt(wx * wx + wy * wy + wz * wz);
  if (!(th >= R(1e-14))) return;
  const R kx = wx / th, ky = wy / th, kz = wz / th;
  const R s = sin(th), h = sin(R(0.5) * th);
  const R omc = R(2) * h * h, c = R(1) - omc;
  const R kd = kx * x + ky * y + kz * z;
  const R cx = ky * z - kz * y, cy = kz * x - kx * z, cz = kx * y - ky * x;
  const R x2 = x * c + cx * s + kx * kd * omc;
  const R y2 = y * c + cy * s + ky * kd * omc;
  const R z2 = z * c + cz * s + kz * kd * omc;
  x = x2;
  y = y2;
  z = z2;
}

// Depondt's rotation vector: de/dt = w x e with w = e x f for the tangent right-hand side f
// (magex/fast.py:284-288).
template <typename R>
__device__ __forceinline__ void omega(R ex, R ey, R ez, R fx, R fy, R fz, R& wx, R& wy, R& wz) {
  wx = ey * fz - ez * fy;
  wy = ez * fx - ex * fz;
  wz = ex * fy - ey * fx;
}

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

template <int NV, int NT = BLOCK>
__device__ __forceinline__ void block_reduce(double (&v)[NV], int nmax_from, double* __restrict__ dst) {
  __shared__ double red[NT / 64][NV];
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
    for (int q = 1; q < NT / 64; ++q) a = (k >= nmax_from) ? fmax(a, red[q][k]) : a + red[q][k];
    dst[k] = a;
  }
}

// Second stage: one workgroup folds the per-block partials in a fixed order, then scales them; with a
// grid of several workgroups, workgroup r folds row r (partials r * nblk .. into out + r * NV).
template <int NV>
__global__ __launch_bounds__(BLOCK) void k_final(const double* __restrict__ part, int nblk, int nmax_from,
                                                 double scale, double* __restrict__ out) {
  part += size_t(blockIdx.x) * nblk * NV;
  out += size_t(blockIdx.x) * NV;
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

}  // namespace mx
