// On-device sampling of m_q = (1/sqrt N) sum_i e_i exp(-i q.r_i) for the dynamical structure factor
// (magex/spectra.py:49-84 `structure_factor`, which forms ph.T @ e with ph = exp(-i r q^T)/sqrt N held
// as an N x nq matrix).  Here the phases are computed where they are used: r_i = c0 a0 + c1 a1 + c2 a2 +
// tau_s from the site's cell and sublattice, one sincos per (site, q) in double precision, so nothing of
// size N x nq is ever stored.  Sums are fixed-order two-stage reductions (bitwise reproducible).
#include <cmath>

#include "mxllg_device.h"

namespace mx {

struct Geom {
  double a[9];           // cell rows a0, a1, a2 (Angstrom)
  double tau[MAXSUB][3];  // sublattice positions
};

constexpr int SQ_NV = 6;  // Re, Im of the three components

// grid (blocks over sites, nq): spins e in magex order (site = cell * ns + s, cell row-major)
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_mq(const R* __restrict__ e, int N, int ns, int L1, int L2, const Geom g,
                                              const double* __restrict__ q, double* __restrict__ part) {
  const int iq = blockIdx.y;
  const double qx = q[3 * iq], qy = q[3 * iq + 1], qz = q[3 * iq + 2];
  double v[SQ_NV] = {0, 0, 0, 0, 0, 0};
  for (int r = blockIdx.x * BLOCK + threadIdx.x; r < N; r += gridDim.x * BLOCK) {
    const int c = r / ns, s = r - c * ns;
    const int c2 = c % L2, t = c / L2, c1 = t % L1, c0 = t / L1;
    const double x = c0 * g.a[0] + c1 * g.a[3] + c2 * g.a[6] + g.tau[s][0];
    const double y = c0 * g.a[1] + c1 * g.a[4] + c2 * g.a[7] + g.tau[s][1];
    const double z = c0 * g.a[2] + c1 * g.a[5] + c2 * g.a[8] + g.tau[s][2];
    const double ph = x * qx + y * qy + z * qz;
    double sn, cs;
    sincos(ph, &sn, &cs);
    const double ex = e[3 * r], ey = e[3 * r + 1], ez = e[3 * r + 2];
    // e exp(-i ph) = e (cos ph - i sin ph)
    v[0] += ex * cs;
    v[1] -= ex * sn;
    v[2] += ey * cs;
    v[3] -= ey * sn;
    v[4] += ez * cs;
    v[5] -= ez * sn;
  }
  __shared__ double res[SQ_NV];
  block_reduce<SQ_NV>(v, SQ_NV, res);
  __syncthreads();
  if (threadIdx.x < SQ_NV) part[(size_t(iq) * gridDim.x + blockIdx.x) * SQ_NV + threadIdx.x] = res[threadIdx.x];
}

// one workgroup per q: fold the partials of q in order and scale by 1/sqrt N
__global__ __launch_bounds__(BLOCK) void k_mq_final(const double* __restrict__ part, int nblk, double scale,
                                                    double* __restrict__ out) {
  const int iq = blockIdx.x;
  double v[SQ_NV] = {0, 0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < nblk; b += BLOCK)
#pragma unroll
    for (int k = 0; k < SQ_NV; ++k) v[k] += part[(size_t(iq) * nblk + b) * SQ_NV + k];
  __shared__ double res[SQ_NV];
  block_reduce<SQ_NV>(v, SQ_NV, res);
  __syncthreads();
  if (threadIdx.x < SQ_NV) out[size_t(iq) * SQ_NV + threadIdx.x] = res[threadIdx.x] * scale;
}

}  // namespace mx

using namespace mx;

struct mx_sampler {
  int device = 0, prec = MX_FP64, N = 0, ns = 1, L[3] = {1, 1, 1}, nq = 0, nblk = 0;
  Geom g{};
  double* q_d = nullptr;
  double* part_d = nullptr;
};

extern "C" {

int mx_sampler_create(const mx_model* m, int32_t nq, const double* q, const double* cell, const double* basis,
                      mx_sampler** out) {
  if (!m || !q || !cell || !basis || !out || nq < 1) {
    set_error("null argument or no q points");
    return MX_EINVAL;
  }
  *out = nullptr;
  auto* s = new mx_sampler();
  s->device = m->device;
  s->prec = m->prec;
  s->N = m->N;
  s->ns = m->ns;
  for (int a = 0; a < 3; ++a) s->L[a] = m->L[a];
  s->nq = nq;
  for (int k = 0; k < 9; ++k) s->g.a[k] = cell[k];
  for (int t = 0; t < m->ns; ++t)
    for (int a = 0; a < 3; ++a) s->g.tau[t][a] = basis[3 * t + a];
  // enough workgroups to fill the chip over all q, each covering a stretch of sites
  s->nblk = std::max(1, std::min((m->N + BLOCK - 1) / BLOCK, std::max(1, 2048 / nq)));
  hipError_t err = hipSetDevice(m->device);
  if (err == hipSuccess) err = hipMalloc(reinterpret_cast<void**>(&s->q_d), sizeof(double) * 3 * nq);
  if (err == hipSuccess)
    err = hipMalloc(reinterpret_cast<void**>(&s->part_d), sizeof(double) * SQ_NV * size_t(nq) * s->nblk);
  if (err == hipSuccess) err = hipMemcpy(s->q_d, q, sizeof(double) * 3 * nq, hipMemcpyHostToDevice);
  if (err != hipSuccess) {
    set_error(std::string("sampler set-up: ") + hipGetErrorString(err));
    mx_sampler_destroy(s);
    return err == hipErrorOutOfMemory ? MX_ENOMEM : MX_EHIP;
  }
  *out = s;
  return MX_OK;
}

void mx_sampler_destroy(mx_sampler* s) {
  if (!s) return;
  (void)hipSetDevice(s->device);
  if (s->q_d) (void)hipFree(s->q_d);
  if (s->part_d) (void)hipFree(s->part_d);
  delete s;
}

int mx_sample_mq(mx_sampler* s, const void* e, double* out, void* stream) {
  if (!s || !e || !out) {
    set_error("null argument");
    return MX_EINVAL;
  }
  hipError_t err = hipSetDevice(s->device);
  if (err != hipSuccess) {
    set_error(std::string("hipSetDevice: ") + hipGetErrorString(err));
    return MX_EHIP;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 g(s->nblk, s->nq), b(BLOCK);
  if (s->prec == MX_FP64)
    hipLaunchKernelGGL((k_mq<double>), g, b, 0, st, static_cast<const double*>(e), s->N, s->ns, s->L[1], s->L[2],
                       s->g, s->q_d, s->part_d);
  else
    hipLaunchKernelGGL((k_mq<float>), g, b, 0, st, static_cast<const float*>(e), s->N, s->ns, s->L[1], s->L[2], s->g,
                       s->q_d, s->part_d);
  hipLaunchKernelGGL(k_mq_final, dim3(s->nq), dim3(BLOCK), 0, st, s->part_d, s->nblk, 1.0 / std::sqrt(double(s->N)),
                     out);
  err = hipGetLastError();
  if (err != hipSuccess) {
    set_error(std::string("launch sampler: ") + hipGetErrorString(err));
    return MX_EHIP;
  }
  return MX_OK;
}

}  // extern "C"
