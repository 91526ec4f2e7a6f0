// Metropolis Monte Carlo colour sweep on the stencil gather (magex/fast.py:393-503 `_mc_group`,
// driven as magex/montecarlo.py:152-179 `MetropolisMC._fast_sweep`).
//
// A colour is a list of sites none of which couple to each other, so all of them move at once: one
// launch per colour, one thread per site.  Per site: H = 2 h + mu mu_B B (the linear local field, meV),
// a proposal, dE = E_onsite(new) - E_onsite(old) - H.(new - old) for the linear moves (the heat-bath and
// over-relaxation proposals already account for H exactly, so only the on-site change remains), and
// acceptance if dE <= 0 or u < exp(-min(beta dE, 700)).
//   move 0 (gaussian): new = n(old + sigma g), g three standard normals
//   move 1 (uniform):  new = n(g)
//   move 2 (heat bath): cos(theta) = 1 + log1p(-r (-expm1(-2x))) / x around H, x = beta |H| (uniform
//                      cos(theta) below x = 1e-8), azimuth 2 pi u in a frame built on H
//   over-relaxation passes: reflection of old about H (skipped where |H|^2 < 1e-28)
// Random numbers: Philox4x32-10 keyed by the seed, counter (site in magex order, sweep lo, sweep hi,
// 'MC' tag | pass << 8 | draw), converted to 53-bit uniforms (two words each), so a sweep does not
// depend on the launch shape or on the cell layout.  The normals, the heat-bath angle and the
// acceptance test are evaluated in the model's precision (double for FP64, which the CPU restatement
// reproduces; float for FP32).
#include <cmath>

#include "mxllg_field.h"

namespace mx {

constexpr uint32_t MC_TAG = 0x4D430000u;

struct Uni4 {
  double u0, u1;  // two 53-bit uniforms in (0, 1) from one Philox call
};

__device__ __forceinline__ Uni4 mc_draw(uint32_t k0, uint32_t k1, uint32_t site, uint64_t sweep, uint32_t tag) {
  uint32_t c0 = site, c1 = uint32_t(sweep), c2 = uint32_t(sweep >> 32), c3 = tag;
  philox10(c0, c1, c2, c3, k0, k1);
  Uni4 u;
  u.u0 = (double(c0 >> 5) * 67108864.0 + double(c1 >> 6) + 0.5) * 0x1.0p-53;
  u.u1 = (double(c2 >> 5) * 67108864.0 + double(c3 >> 6) + 0.5) * 0x1.0p-53;
  return u;
}

// Per-replica parameters of a replica sweep (parallel tempering): replica r owns spins
// e + r * N * 3 and its own inverse temperature, cone width, generator key and acceptance counter.
struct McRep {
  double beta, sigma;
  uint32_t seed_lo, seed_hi;
};

template <typename R, int CM, int GEO>
__global__ __launch_bounds__(BLOCK) void k_mc(const Static<R>* __restrict__ S, Dyn<R> d, R* __restrict__ e,
                                              const int32_t* __restrict__ sites, int n,
                                              const R* __restrict__ beta_site, R beta_u, int move, int pass,
                                              R sigma, uint64_t sweep, unsigned long long* __restrict__ acc,
                                              const McRep* __restrict__ reps) {
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  if (reps) {
    // replica blockIdx.y: wavefront-uniform scalar loads of its parameters
    const McRep* rp = reps + blockIdx.y;
    e += size_t(blockIdx.y) * size_t(S->N) * 3;
    acc += blockIdx.y;
    beta_u = R(rp->beta);
    sigma = R(rp->sigma);
    d.seed_lo = rp->seed_lo;
    d.seed_hi = rp->seed_hi;
  }
  int ok_count = 0;
  if (t < n) {
    Site st;
    site_at<R, GEO>(S, sites[t], st);
    const int i = st.i;
    R hx, hy, hz;
    // a colour is a fraction of the lattice (a few workgroups per CU): one wide batch keeps all of a
    // site's loads in flight at once (measured on bcc64: FP32 per-bond 4 % faster; FP64 per-bond keeps
    // its narrow batches, 22 % slower wide)
    exchange<R, CM, GEO, (CM == COUP_BOND ? (sizeof(R) == 8 ? 2 : 14) : 16)>(S, d, e, st, hx, hy, hz);
    const R zf = R(1) / inv_mu_of(S, d, st);  // mu mu_B
    const R Hx = R(2) * hx + zf * d.B0, Hy = R(2) * hy + zf * d.B1, Hz = R(2) * hz + zf * d.B2;
    const R ox = e[3 * i], oy = e[3 * i + 1], oz = e[3 * i + 2];
    using F = R;  // precision of the random transforms
    const R bt = beta_site ? beta_site[i] : beta_u;
    const uint32_t iref = uint32_t(st.crm * S->ns + st.s);
    const uint32_t tag = MC_TAG | (uint32_t(pass) << 8);
    R nx = ox, ny = oy, nz = oz;
    bool lin = true, live = true;
    F ua;  // acceptance uniform
    if (pass > 0) {
      // over-relaxation: reflection about H (H.e conserved)
      const R HH = Hx * Hx + Hy * Hy + Hz * Hz;
      if (HH < R(1e-28)) {
        live = false;
      } else {
        const R p = R(2) * (ox * Hx + oy * Hy + oz * Hz) / HH;
        nx = p * Hx - ox;
        ny = p * Hy - oy;
        nz = p * Hz - oz;
        const R q = rnorm3(nx, ny, nz);
        nx *= q;
        ny *= q;
        nz *= q;
      }
      lin = false;
      ua = F(mc_draw(d.seed_lo, d.seed_hi, iref, sweep, tag).u0);
    } else if (move == 2) {
      const Uni4 a = mc_draw(d.seed_lo, d.seed_hi, iref, sweep, tag);
      ua = F(mc_draw(d.seed_lo, d.seed_hi, iref, sweep, tag | 1u).u0);
      const F u0 = F(a.u0), u1 = F(a.u1);
      const F H0 = Hx, H1 = Hy, H2 = Hz;
      const F Hn = sqrt(H0 * H0 + H1 * H1 + H2 * H2);
      const F xx = F(bt) * Hn;
      F cth;
      if (xx < F(1e-8) || !isfinite(xx)) cth = xx < F(1e-8) ? F(2) * u0 - F(1) : F(1);
      else cth = F(1) + log1p(-u0 * (-expm1(F(-2) * xx))) / xx;
      cth = fmin(F(1), fmax(F(-1), cth));
      const F ph = F(2 * M_PI) * u1;
      F zx = F(0), zy = F(0), zz = F(1);
      if (Hn > F(0)) {
        zx = H0 / Hn;
        zy = H1 / Hn;
        zz = H2 / Hn;
      }
      F ax = F(0), ay = F(1), az = F(0);
      if (fabs(zx) < F(0.9)) {
        ax = F(1);
        ay = F(0);
      }
      F t1x = zy * az - zz * ay, t1y = zz * ax - zx * az, t1z = zx * ay - zy * ax;
      const F n1 = sqrt(t1x * t1x + t1y * t1y + t1z * t1z);
      t1x /= n1;
      t1y /= n1;
      t1z /= n1;
      const F t2x = zy * t1z - zz * t1y, t2y = zz * t1x - zx * t1z, t2z = zx * t1y - zy * t1x;
      const F sn = sqrt(fmax(F(0), F(1) - cth * cth));
      const F cp = cos(ph), sp = sin(ph);
      nx = cth * zx + sn * (cp * t1x + sp * t2x);
      ny = cth * zy + sn * (cp * t1y + sp * t2y);
      nz = cth * zz + sn * (cp * t1z + sp * t2z);
      lin = false;
    } else {
      const Uni4 a = mc_draw(d.seed_lo, d.seed_hi, iref, sweep, tag);
      const Uni4 b = mc_draw(d.seed_lo, d.seed_hi, iref, sweep, tag | 1u);
      ua = F(mc_draw(d.seed_lo, d.seed_hi, iref, sweep, tag | 2u).u0);
      const F r0 = sqrt(F(-2) * log(F(a.u0))), r2 = sqrt(F(-2) * log(F(b.u0)));
      const F g0 = r0 * cos(F(2 * M_PI) * F(a.u1)), g1 = r0 * sin(F(2 * M_PI) * F(a.u1));
      const F g2 = r2 * cos(F(2 * M_PI) * F(b.u1));
      if (move == 0) {
        nx = ox + sigma * R(g0);
        ny = oy + sigma * R(g1);
        nz = oz + sigma * R(g2);
      } else {
        nx = R(g0);
        ny = R(g1);
        nz = R(g2);
      }
      const R q = rnorm3(nx, ny, nz);
      nx *= q;
      ny *= q;
      nz *= q;
    }
    if (live) {
      R dE = onsite_energy(S, d, st, nx, ny, nz) - onsite_energy(S, d, st, ox, oy, oz);
      if (lin) dE -= Hx * (nx - ox) + Hy * (ny - oy) + Hz * (nz - oz);
      bool ok = dE <= R(0);
      if (!ok) ok = ua < exp(-fmin(F(dE) * F(bt), F(700)));
      if (ok) {
        e[3 * i] = nx;
        e[3 * i + 1] = ny;
        e[3 * i + 2] = nz;
        ok_count = pass == 0 ? 1 : 0;
      }
    }
  }
  // accepted moves: integer counts, so the order of the additions does not matter
  __shared__ int cnt[BLOCK / 64];
  int v = ok_count;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) cnt[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int w = 0; w < BLOCK / 64; ++w) s += cnt[w];
    if (s) atomicAdd(acc, (unsigned long long)s);
  }
}

template <typename R>
static int mc_t(const Model& m, void* e, const int32_t* sites, int n, const void* beta_site, double beta, int move,
                int pass, double sigma, uint64_t seed, uint64_t sweep, unsigned long long* acc, hipStream_t s,
                const McRep* reps, int nrep) {
  if (n <= 0 || nrep <= 0) return MX_OK;
  const Static<R>* S = static_cast<const Static<R>*>(m.static_d);
  Dyn<R> d{};
  d.B0 = R(m.B[0]);
  d.B1 = R(m.B[1]);
  d.B2 = R(m.B[2]);
  d.seed_lo = uint32_t(seed);
  d.seed_hi = uint32_t(seed >> 32);
  d.bond = static_cast<const R*>(m.bond_d);
  d.A = static_cast<const R*>(m.A_d);
  d.Kc = static_cast<const R*>(m.Kc_d);
  d.K2 = static_cast<const R*>(m.K2_d);
  d.K3 = static_cast<const R*>(m.K3_d);
  d.inv_mu = static_cast<const R*>(m.inv_mu_d);
  R* E = static_cast<R*>(e);
  const R* bs = static_cast<const R*>(beta_site);
  const R bu = std::isinf(beta) ? R(INFINITY) : R(beta);
  const dim3 g((n + BLOCK - 1) / BLOCK, reps ? nrep : 1), b(BLOCK);
#define MX_MC_CALL(CM, GEO)                                                                                    \
  hipLaunchKernelGGL((k_mc<R, CM, GEO>), g, b, 0, s, S, d, E, sites, n, bs, bu, move, pass, R(sigma), sweep, acc, \
                     reps)
  switch (m.coup) {
    case COUP_TABLE1:
      if (m.geo == GEO_TILED) MX_MC_CALL(COUP_TABLE1, GEO_TILED);
      else if (m.geo == GEO_CHECKER) MX_MC_CALL(COUP_TABLE1, GEO_CHECKER);
      else if (m.geo == GEO_PERIODIC) MX_MC_CALL(COUP_TABLE1, GEO_PERIODIC);
      else MX_MC_CALL(COUP_TABLE1, GEO_OPEN);
      break;
    case COUP_TABLE4:
      if (m.geo == GEO_TILED) MX_MC_CALL(COUP_TABLE4, GEO_TILED);
      else if (m.geo == GEO_CHECKER) MX_MC_CALL(COUP_TABLE4, GEO_CHECKER);
      else if (m.geo == GEO_PERIODIC) MX_MC_CALL(COUP_TABLE4, GEO_PERIODIC);
      else MX_MC_CALL(COUP_TABLE4, GEO_OPEN);
      break;
    case COUP_TABLE9:
      if (m.geo == GEO_TILED) MX_MC_CALL(COUP_TABLE9, GEO_TILED);
      else if (m.geo == GEO_CHECKER) MX_MC_CALL(COUP_TABLE9, GEO_CHECKER);
      else if (m.geo == GEO_PERIODIC) MX_MC_CALL(COUP_TABLE9, GEO_PERIODIC);
      else MX_MC_CALL(COUP_TABLE9, GEO_OPEN);
      break;
    default:
      if (m.geo == GEO_TILED) MX_MC_CALL(COUP_BOND, GEO_TILED);
      else if (m.geo == GEO_CHECKER) MX_MC_CALL(COUP_BOND, GEO_CHECKER);
      else if (m.geo == GEO_PERIODIC) MX_MC_CALL(COUP_BOND, GEO_PERIODIC);
      else MX_MC_CALL(COUP_BOND, GEO_OPEN);
      break;
  }
#undef MX_MC_CALL
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    set_error(std::string("launch mc: ") + hipGetErrorString(err));
    return MX_EHIP;
  }
  return MX_OK;
}

int launch_mc(const Model& m, void* e, const int32_t* sites, int n, const void* beta_site, double beta, int move,
              int pass, double sigma, uint64_t seed, uint64_t sweep, unsigned long long* acc, hipStream_t s) {
  return m.prec == MX_FP64
             ? mc_t<double>(m, e, sites, n, beta_site, beta, move, pass, sigma, seed, sweep, acc, s, nullptr, 1)
             : mc_t<float>(m, e, sites, n, beta_site, beta, move, pass, sigma, seed, sweep, acc, s, nullptr, 1);
}

}  // namespace mx

// ------------------------------------------------------------------------------------- replica sweeps
using namespace mx;

extern "C" int mx_mc_sweep_replicas(mx_model* m, void* e, int32_t nrep, int64_t nsweeps, int move, int overrelax,
                                    const double* beta, const double* sigma, const uint64_t* seeds, uint64_t sweep0,
                                    void* stream) {
  if (!m || !e || !beta || !sigma || !seeds) {
    set_error("null argument");
    return MX_EINVAL;
  }
  if (nrep < 1 || nsweeps < 0 || overrelax < 0 || move < 0 || move > 2) {
    set_error("bad replica count, sweep count, over-relaxation count or move");
    return MX_EINVAL;
  }
  if (m->mc_off.empty()) {
    set_error("set the colours first (mx_mc_set_colors)");
    return MX_EINVAL;
  }
  if (m->mc_beta_d) {
    set_error("replica sweeps take one temperature per replica: the model has per-site temperatures");
    return MX_EUNSUPPORTED;
  }
  hipError_t err = hipSetDevice(m->device);
  int rc = err == hipSuccess ? sync_model_static(*m) : MX_EHIP;
  if (rc) {
    if (err != hipSuccess) set_error(std::string("hipSetDevice: ") + hipGetErrorString(err));
    return rc;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  std::vector<McRep> h(nrep);
  bool finite = true;
  for (int r = 0; r < nrep; ++r) {
    h[r].beta = beta[r];
    h[r].sigma = sigma[r];
    h[r].seed_lo = uint32_t(seeds[r]);
    h[r].seed_hi = uint32_t(seeds[r] >> 32);
    finite = finite && std::isfinite(beta[r]);
  }
  // parameter block and per-replica counters: reallocated when the replica count grows; the earlier
  // sweeps on this stream finish before the block is rewritten
  err = hipStreamSynchronize(s);
  if (err == hipSuccess && nrep > m->mc_rep_cap) {
    if (m->mc_rep_d) (void)hipFree(m->mc_rep_d);
    if (m->mc_repacc_d) (void)hipFree(m->mc_repacc_d);
    m->mc_rep_d = nullptr;
    m->mc_repacc_d = nullptr;
    m->mc_rep_cap = 0;
    err = hipMalloc(&m->mc_rep_d, sizeof(McRep) * nrep);
    if (err == hipSuccess)
      err = hipMalloc(reinterpret_cast<void**>(&m->mc_repacc_d), sizeof(unsigned long long) * nrep);
    if (err == hipSuccess) err = hipMemset(m->mc_repacc_d, 0, sizeof(unsigned long long) * nrep);
    if (err == hipSuccess) m->mc_rep_cap = nrep;
  }
  if (err == hipSuccess) err = hipMemcpy(m->mc_rep_d, h.data(), sizeof(McRep) * nrep, hipMemcpyHostToDevice);
  // multi-site cells and tiled / checker layouts: the replicas are swept in the internal site order
  const bool perm = needs_permute(*m);
  if (perm && !m->perm_d) {
    set_error("replica sweeps: the model's cell layout is not set");
    return MX_EINVAL;
  }
  const size_t rbytes = size_t(m->N) * 3 * (m->prec == MX_FP64 ? sizeof(double) : sizeof(float));
  if (err == hipSuccess && perm && nrep > m->mc_repe_cap) {
    if (m->mc_repe_d) (void)hipFree(m->mc_repe_d);
    m->mc_repe_d = nullptr;
    m->mc_repe_cap = 0;
    err = hipMalloc(&m->mc_repe_d, rbytes * nrep);
    if (err == hipSuccess) m->mc_repe_cap = nrep;
  }
  if (err != hipSuccess) {
    set_error(std::string("replica sweep set-up: ") + hipGetErrorString(err));
    return err == hipErrorOutOfMemory ? MX_ENOMEM : MX_EHIP;
  }
  const McRep* reps = static_cast<const McRep*>(m->mc_rep_d);
  const int mv = (move == 2 && !finite) ? 0 : move;
  const int nc = int(m->mc_off.size()) - 1;
  char* const user = static_cast<char*>(e);
  if (perm) {
    for (int r = 0; rc == MX_OK && r < nrep; ++r)
      rc = launch_permute(*m, user + r * rbytes, static_cast<char*>(m->mc_repe_d) + r * rbytes, 1, s);
    e = m->mc_repe_d;
  }
  for (int64_t k = 0; rc == MX_OK && k < nsweeps; ++k)
    for (int pass = 0; rc == MX_OK && pass <= overrelax; ++pass)
      for (int c = 0; rc == MX_OK && c < nc; ++c) {
        const int64_t a = m->mc_off[c], b = m->mc_off[c + 1];
        const uint64_t sw = sweep0 + uint64_t(k);
        rc = m->prec == MX_FP64 ? mc_t<double>(*m, e, m->mc_sites_d + a, int(b - a), nullptr, 0.0, mv, pass, 0.0, 0,
                                               sw, m->mc_repacc_d, s, reps, nrep)
                                : mc_t<float>(*m, e, m->mc_sites_d + a, int(b - a), nullptr, 0.0, mv, pass, 0.0, 0,
                                              sw, m->mc_repacc_d, s, reps, nrep);
      }
  if (perm)
    for (int r = 0; rc == MX_OK && r < nrep; ++r)
      rc = launch_permute(*m, static_cast<char*>(m->mc_repe_d) + r * rbytes, user + r * rbytes, 0, s);
  return rc;
}

extern "C" int mx_mc_replica_accepted(mx_model* m, int32_t nrep, uint64_t* out, int reset) {
  if (!m || !out || nrep < 1) {
    set_error("null argument or bad replica count");
    return MX_EINVAL;
  }
  if (nrep > m->mc_rep_cap) {
    for (int r = 0; r < nrep; ++r) out[r] = 0;
    return MX_OK;
  }
  std::vector<unsigned long long> v(nrep);
  hipError_t err = hipSetDevice(m->device);
  if (err == hipSuccess) err = hipMemcpy(v.data(), m->mc_repacc_d, sizeof(unsigned long long) * nrep,
                                         hipMemcpyDeviceToHost);
  if (err == hipSuccess && reset) err = hipMemset(m->mc_repacc_d, 0, sizeof(unsigned long long) * nrep);
  if (err != hipSuccess) {
    set_error(std::string("replica counters: ") + hipGetErrorString(err));
    return MX_EHIP;
  }
  for (int r = 0; r < nrep; ++r) out[r] = uint64_t(v[r]);
  return MX_OK;
}
