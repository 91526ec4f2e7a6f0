// Batched form: many replicas of one unit-cell model on uneven periodic supercells, each with its own
// damping, temperature and noise seed, advanced by one launch per Heun pass for the whole batch.
//
// Layout: replica r's sites sit at [off_r, off_r + N_r) of one packed array, in magex's order
// (cell * ns + sublattice).  A replica table (Rep, device memory) holds each replica's supercell,
// offsets, damping, noise amplitude and seed; tile_rep maps a workgroup to its replica.  Workgroup
// indices are laid out replica after replica and pass through the XCD-aware map, so a replica's
// workgroups run on one or two XCDs and its neighbour gathers stay in that XCD's L2.
//
// Launch shape (threads per workgroup BS, sites per thread spt, captured step loop or not) is chosen by
// the autotuner in magex_amd.batch; results do not depend on it.  The noise of replica r is keyed by
// (seed_r, step, local site), exactly the stream mx_llg_run draws for that replica alone.
#include <cmath>
#include <cstring>
#include <new>
#include <stdexcept>
#include <vector>

#include "mxllg_device.h"

namespace mx {

struct Rep {
  int L0, L1, L2, ncell;
  int off;    // first site of the replica in the packed arrays
  int tile0;  // first logical workgroup of the replica
  int ntile;  // workgroups of the replica
  float rL1, rL2;  // 1 / L1, 1 / L2 for the exact float division of cell indices (< 2^24)
  int pad;
  double alpha;
  double sig[MAXSUB];  // sigma of sublattice s: sqrt(2 alpha k_B T / (gamma mu_s mu_B dt))
  uint32_t seed_lo, seed_hi;
  uint32_t pad2[2];
  uint32_t Q0, Q1, Q2;  // byte period of axes 0, 1, 2 in the replica's spin array
  uint32_t pad3;
};

// Per (replica, stencil entry): byte offset from a site's spin to its neighbour's before wrapping, and
// the cell coordinate from which each axis wraps (L_k - d_k with the offset d reduced into [0, L_k)).
struct REnt {
  uint32_t D;
  int32_t th0, th1, th2;
};

struct BEnt {
  int32_t slot, t, r01, r2;  // coupling slot, target sublattice, raw cell offset (R0 & 0xffff | R1 << 16), R2
};

template <typename R>
struct BStatic {
  int ns, has_A;
  int sptr[MAXSUB + 1];
  BEnt ent[MAXENT];
  R tab[MAXTAB];
  R Asub[MAXSUB][9];
  R inv_mu_sub[MAXSUB], sign_sub[MAXSUB];
};

template <typename R>
struct BDyn {
  R B0, B1, B2, dt;
  int prec, noisy;
  int noise_dbl;  // the prototype model's generator grade
};

struct BSite {
  int s, c, c0, c1, c2, i, iref;
  uint32_t ob;  // byte offset of the site's spin from its replica's first spin
};

// Spin of a site in array a: the replica's base (wavefront-uniform, 64-bit) plus a 32-bit byte offset.
template <typename R>
__device__ __forceinline__ R* rep_at(R* a, const Rep* __restrict__ rp, uint32_t ob) {
  return reinterpret_cast<R*>(reinterpret_cast<char*>(a + size_t(rp->off) * 3) + ob);
}
template <typename R>
__device__ __forceinline__ const R* rep_at(const R* a, const Rep* __restrict__ rp, uint32_t ob) {
  return reinterpret_cast<const R*>(reinterpret_cast<const char*>(a + size_t(rp->off) * 3) + ob);
}

// h_i over the stencil with raw offsets (|R| < L on every axis), accumulated in stencil order.  Each
// batch of CH gathers is branch-free: every load is issued unconditionally (entries past the end re-read
// the last one, a valid address) and a select zeroes what does not count, so the CH loads of a wavefront
// are in flight together; mid() runs between the first batch's loads and their use.  A neighbour's
// address needs no multiply per lane: with the offset d reduced into [0, L) per axis (scalar), its byte
// offset is the site's own plus a wavefront-uniform delta, less one period per axis that wraps
// (c_k >= L_k - d_k).  Offsets are 32-bit from the replica's base (a replica has fewer than 2^24 cells
// of at most 8 sites of 24 B).  CH (6 or 8) is the batch width the host picks for the stencil; a full
// batch skips the selects (only a last, partial batch pays for them).
template <typename R, int CM, int CH, typename Mid>
__device__ __forceinline__ void bexchange(const BStatic<R>* __restrict__ S, const Rep* __restrict__ rp,
                                          const REnt* __restrict__ re, const R* __restrict__ e, const BSite& st,
                                          R& hx, R& hy, R& hz, Mid&& mid) {
  hx = hy = hz = R(0);
  const int kb = S->sptr[st.s], ke = S->sptr[st.s + 1];
  const uint32_t Q0 = rp->Q0, Q1 = rp->Q1, Q2 = rp->Q2;
  if (kb >= ke) {
    mid();
    return;
  }
  const char* eb = reinterpret_cast<const char*>(e + size_t(rp->off) * 3);
  for (int k0 = kb; k0 < ke; k0 += CH) {
    R ex[CH], ey[CH], ez[CH];
    int slot[CH];
    bool live[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      live[u] = k0 + u < ke;
      const int k = min(k0 + u, ke - 1);
      slot[u] = S->ent[k].slot;
      const REnt q = re[k];
      uint32_t o = st.ob + q.D;
      o -= st.c0 >= q.th0 ? Q0 : 0u;
      o -= st.c1 >= q.th1 ? Q1 : 0u;
      o -= st.c2 >= q.th2 ? Q2 : 0u;
      const R* p = reinterpret_cast<const R*>(eb + o);
      ex[u] = p[0];
      ey[u] = p[1];
      ez[u] = p[2];
    }
    if (k0 == kb) mid();
    const bool full = k0 + CH <= ke;  // wavefront-uniform
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      R x = ex[u], y = ey[u], z = ez[u];
      if (!full) {
        x = live[u] ? x : R(0);
        y = live[u] ? y : R(0);
        z = live[u] ? z : R(0);
      }
      if constexpr (CM == COUP_TABLE1) {
        const R Jv = S->tab[slot[u]];
        hx += Jv * x;
        hy += Jv * y;
        hz += Jv * z;
      } else if constexpr (CM == COUP_TABLE4) {
        const R* T = S->tab + 4 * slot[u];
        hx += T[0] * x + (y * T[3] - z * T[2]);
        hy += T[0] * y + (z * T[1] - x * T[3]);
        hz += T[0] * z + (x * T[2] - y * T[1]);
      } else {
        const R* T = S->tab + 9 * slot[u];
        hx += T[0] * x + T[1] * y + T[2] * z;
        hy += T[3] * x + T[4] * y + T[5] * z;
        hz += T[6] * x + T[7] * y + T[8] * z;
      }
    }
  }
}

template <typename R, int CM, int CH>
__device__ __forceinline__ void bexchange(const BStatic<R>* __restrict__ S, const Rep* __restrict__ rp,
                                          const REnt* __restrict__ re, const R* __restrict__ e, const BSite& st,
                                          R& hx, R& hy, R& hz) {
  bexchange<R, CM, CH>(S, rp, re, e, st, hx, hy, hz, [] {});
}

template <typename R>
__device__ __forceinline__ void bonsite(const BStatic<R>* __restrict__ S, int s, R x, R y, R z, R& gx, R& gy, R& gz) {
  gx = gy = gz = R(0);
  if (S->has_A) {
    const R* a = S->Asub[s];
    gx = R(2) * (a[0] * x + a[1] * y + a[2] * z);
    gy = R(2) * (a[3] * x + a[4] * y + a[5] * z);
    gz = R(2) * (a[6] * x + a[7] * y + a[8] * z);
  }
}

// The workgroup's replica and its place in it; st gets the site of (thread, u).
template <typename R, int BS>
__device__ __forceinline__ bool bsite(const BStatic<R>* __restrict__ S, const Rep* __restrict__ rp, int lt, int spt,
                                      int u, BSite& st) {
  const int ns = S->ns;
  const int tq = lt / ns;
  st.s = lt - tq * ns;
  st.c = (tq * spt + u) * BS + int(threadIdx.x);
  if (st.c >= rp->ncell) return false;
  const int L1 = rp->L1, L2 = rp->L2;
  const int t = fdiv(st.c, L2, rp->rL2);
  st.c2 = st.c - t * L2;
  st.c0 = fdiv(t, L1, rp->rL1);
  st.c1 = t - st.c0 * L1;
  st.iref = st.c * ns + st.s;
  st.i = rp->off + st.iref;
  st.ob = uint32_t(st.iref) * uint32_t(3 * sizeof(R));
  return true;
}

template <typename R, int CM, int BS, int PASS, int CH>
__global__ __launch_bounds__(BS) void k_bheun(const BStatic<R>* __restrict__ S, const Rep* __restrict__ reps,
                                              const int32_t* __restrict__ tile_rep, const REnt* __restrict__ rent,
                                              int nent, const BDyn<R> d, R* __restrict__ e, R* __restrict__ ep,
                                              R* __restrict__ k1, R* __restrict__ bn, uint64_t step, int spt,
                                              const uint64_t* __restrict__ step_ptr) {
  const int b = xcd_tile(blockIdx.x, gridDim.x);
  const int r = tile_rep[b];
  const Rep* __restrict__ rp = reps + r;  // descriptor read through the pointer: scalar loads
  const REnt* __restrict__ re = rent + size_t(r) * nent;
  if (step_ptr) step += *step_ptr;
  const int lt = b - rp->tile0;
  const R a = R(rp->alpha);
  for (int u = 0; u < spt; ++u) {
    BSite st;
    if (!bsite<R, BS>(S, rp, lt, spt, u, st)) return;
    const R* src = PASS == 1 ? e : ep;
    const R* sp = rep_at(src, rp, st.ob);
    R* bnp = rep_at(bn, rp, st.ob);
    R* k1p = rep_at(k1, rp, st.ob);
    const R x = sp[0], y = sp[1], z = sp[2];
    R hx, hy, hz, gx, gy, gz, nx = R(0), ny = R(0), nz = R(0);
    // the thermal field is drawn in pass 1 (while the first neighbour batch is in flight) and kept for
    // pass 2: the batched kernels are bound by arithmetic more than by bytes, and the generator is the
    // largest part of it
    if (PASS == 2 && d.noisy) {
      nx = bnp[0];
      ny = bnp[1];
      nz = bnp[2];
    }
    bexchange<R, CM, CH>(S, rp, re, src, st, hx, hy, hz, [&] {
      if (PASS == 1 && d.noisy) {
        const R sg = R(rp->sig[st.s]);
        if (d.noise_dbl) {
          double n0, n1, n2;
          gauss3d(rp->seed_lo, rp->seed_hi, step, uint32_t(st.iref), n0, n1, n2);
          nx = sg * R(n0);
          ny = sg * R(n1);
          nz = sg * R(n2);
        } else {
          float n0, n1, n2;
          gauss3(rp->seed_lo, rp->seed_hi, step, uint32_t(st.iref), n0, n1, n2);
          nx = sg * R(n0);
          ny = sg * R(n1);
          nz = sg * R(n2);
        }
        bnp[0] = nx;
        bnp[1] = ny;
        bnp[2] = nz;
      }
    });
    bonsite(S, st.s, x, y, z, gx, gy, gz);
    const R im = S->inv_mu_sub[st.s];
    const R bx = (R(2) * hx - gx) * im + d.B0 + nx, by = (R(2) * hy - gy) * im + d.B1 + ny,
            bz = (R(2) * hz - gz) * im + d.B2 + nz;
    R fx, fy, fz;
    llg_rhs(x, y, z, bx, by, bz, a, d.prec, fx, fy, fz);
    const R hdt = R(0.5) * d.dt;
    if (PASS == 1) {
      // k1 keeps w = e + dt/2 k1, so pass 2 does not read e again
      k1p[0] = x + hdt * fx;
      k1p[1] = y + hdt * fy;
      k1p[2] = z + hdt * fz;
      const R px = x + d.dt * fx, py = y + d.dt * fy, pz = z + d.dt * fz;
      const R q = rnorm3(px, py, pz);
      R* o = rep_at(ep, rp, st.ob);
      o[0] = px * q;
      o[1] = py * q;
      o[2] = pz * q;
    } else {
      const R cx = k1p[0] + hdt * fx;
      const R cy = k1p[1] + hdt * fy;
      const R cz = k1p[2] + hdt * fz;
      const R q = rnorm3(cx, cy, cz);
      R* o = rep_at(e, rp, st.ob);
      o[0] = cx * q;
      o[1] = cy * q;
      o[2] = cz * q;
    }
  }
}

// Per-workgroup partials of [E, sum e, sum s e, max torque], stored at the workgroup's logical index so
// that a replica's partials are contiguous.
template <typename R, int CM, int BS>
__global__ __launch_bounds__(BS) void k_bobs(const BStatic<R>* __restrict__ S, const Rep* __restrict__ reps,
                                             const int32_t* __restrict__ tile_rep, const REnt* __restrict__ rent,
                                             int nent, const BDyn<R> d, const R* __restrict__ e, int spt,
                                             double* __restrict__ part) {
  const int b = xcd_tile(blockIdx.x, gridDim.x);
  const int r = tile_rep[b];
  const Rep* __restrict__ rp = reps + r;
  const REnt* __restrict__ re = rent + size_t(r) * nent;
  const int lt = b - rp->tile0;
  double v[MX_OBS_LEN] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int u = 0; u < spt; ++u) {
    BSite st;
    if (!bsite<R, BS>(S, rp, lt, spt, u, st)) break;
    const int i = st.i;
    const R x = e[3 * i], y = e[3 * i + 1], z = e[3 * i + 2];
    R hx, hy, hz, gx, gy, gz;
    bexchange<R, CM, 8>(S, rp, re, e, st, hx, hy, hz);
    bonsite(S, st.s, x, y, z, gx, gy, gz);
    const R im = S->inv_mu_sub[st.s];
    const R Eon = (gx * x + gy * y + gz * z) * R(0.5);
    v[0] += double(Eon) - double(x * hx + y * hy + z * hz) - double(x * d.B0 + y * d.B1 + z * d.B2) / double(im);
    v[1] += x;
    v[2] += y;
    v[3] += z;
    const R sg = S->sign_sub[st.s];
    v[4] += sg * x;
    v[5] += sg * y;
    v[6] += sg * z;
    const R bx = (R(2) * hx - gx) * im + d.B0, by = (R(2) * hy - gy) * im + d.B1, bz = (R(2) * hz - gz) * im + d.B2;
    const double tx = y * bz - z * by, ty = z * bx - x * bz, tz = x * by - y * bx;
    v[7] = fmax(v[7], sqrt(tx * tx + ty * ty + tz * tz));
  }
  block_reduce<MX_OBS_LEN, BS>(v, 7, part + size_t(b) * MX_OBS_LEN);
}

// One workgroup per replica folds its partials in a fixed order.
__global__ __launch_bounds__(BLOCK) void k_bfinal(const Rep* __restrict__ reps, const double* __restrict__ part,
                                                  double* __restrict__ out) {
  __shared__ double res[MX_OBS_LEN];
  const Rep* __restrict__ rp = reps + blockIdx.x;
  const int ntile = rp->ntile, tile0 = rp->tile0;
  double v[MX_OBS_LEN] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int t = threadIdx.x; t < ntile; t += BLOCK)
#pragma unroll
    for (int k = 0; k < MX_OBS_LEN; ++k) {
      const double x = part[size_t(tile0 + t) * MX_OBS_LEN + k];
      v[k] = k == 7 ? fmax(v[k], x) : v[k] + x;
    }
  block_reduce<MX_OBS_LEN>(v, 7, res);
  __syncthreads();
  if (threadIdx.x < MX_OBS_LEN) out[size_t(blockIdx.x) * MX_OBS_LEN + threadIdx.x] = res[threadIdx.x];
}

__global__ void k_bcounter(uint64_t* __restrict__ c, uint64_t value, int add) {
  if (threadIdx.x == 0 && blockIdx.x == 0) c[0] = add ? c[0] + value : value;
}

}  // namespace mx

// ============================================================================================ host
using namespace mx;

struct mx_batch {
  int prec = MX_FP64, device = 0, nrep = 0, ns = 1;
  int coup = COUP_TABLE1;
  std::vector<Rep> reps;
  std::vector<double> mu_sub;
  std::vector<char> static_h;  // BStatic<R> bytes
  double B[3] = {0, 0, 0};
  double dt = 1e-3;
  int precession = 1, noisy = 0, dyn_set = 0;
  int noise_dbl = 0;  // generator grade, the prototype model's
  std::vector<double> alpha, T;
  int block = 256, spt = 1, graph_steps = 0;
  int64_t total = 0;
  int ntiles = 0;
  void *static_d = nullptr, *reps_d = nullptr, *ep_d = nullptr, *k1_d = nullptr;
  void* bn_d = nullptr;  // thermal field of the current step, drawn in pass 1 and reused in pass 2
  void* rent_d = nullptr;  // REnt per (replica, stencil entry)
  int nent = 0;
  // gather batch width of the step kernels: 6 where that takes no more batches than 8 for the longest
  // sublattice stencil (the 6-entry triangular and simple-cubic neighbours: one full batch, no dead
  // entries), else 8.  Same summation order either way, so the same result bits.
  int chw = 8;
  int32_t* tile_rep_d = nullptr;
  double* part_d = nullptr;
  size_t part_cap = 0;
  uint64_t* counter_d = nullptr;
  hipStream_t cap = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<char> gkey;
};

namespace {

struct BErr {
  int code;
  std::string msg;
};
[[noreturn]] void bfail(int code, const std::string& m) { throw BErr{code, m}; }
void bok(hipError_t e, const char* what) {
  if (e != hipSuccess) bfail(e == hipErrorOutOfMemory ? MX_ENOMEM : MX_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}
template <typename F>
int bguard(F&& f) {
  try {
    return f();
  } catch (const BErr& e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return MX_ENOMEM;
  }
}
void bfree(void*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}
size_t rsz(const mx_batch& b) { return b.prec == MX_FP64 ? sizeof(double) : sizeof(float); }

void drop_bgraph(mx_batch& b) {
  if (b.gexec) (void)hipGraphExecDestroy(b.gexec);
  b.gexec = nullptr;
  b.gkey.clear();
}

template <typename R>
void build_static(mx_batch& b, const Model& m) {
  BStatic<R> S;
  std::memset(&S, 0, sizeof(S));
  S.ns = m.ns;
  S.has_A = m.has_A;
  for (int s = 0; s <= MAXSUB; ++s) S.sptr[s] = s <= m.ns ? m.sptr[s] : m.sptr[m.ns];
  for (int k = 0; k < m.nent; ++k) {
    const int* r = &m.ent_raw[6 * size_t(m.ent_order[k])];
    S.ent[k] = BEnt{m.ent[k].slot, r[1], (r[2] & 0xffff) | (r[3] << 16), r[4]};
  }
  for (size_t k = 0; k < m.table.size() && k < size_t(MAXTAB); ++k) S.tab[k] = R(m.table[k]);
  for (int s = 0; s < m.ns; ++s) {
    for (int q = 0; q < 9; ++q) S.Asub[s][q] = R(m.Asub[9 * s + q]);
    S.inv_mu_sub[s] = R(1.0 / (m.mu_sub[s] * MU_B));
    S.sign_sub[s] = R(m.sign_sub[s]);
  }
  b.static_h.assign(reinterpret_cast<char*>(&S), reinterpret_cast<char*>(&S) + sizeof(S));
}

// Byte periods of every replica and the REnt table (internal stencil order: entries grouped by source
// sublattice), uploaded once: the neighbour address arithmetic of the kernels is then one table read
// and three conditional subtractions per entry.
void build_rent(mx_batch& b, const Model& m) {
  const uint32_t W = uint32_t(3 * rsz(b));
  b.nent = m.nent;
  int mx = 0;
  for (int s = 0; s < m.ns; ++s) mx = std::max(mx, m.sptr[s + 1] - m.sptr[s]);
  b.chw = (mx + 5) / 6 <= (mx + 7) / 8 ? 6 : 8;
  std::vector<REnt> t(size_t(b.nrep) * std::max(m.nent, 1));
  for (int r = 0; r < b.nrep; ++r) {
    Rep& p = b.reps[r];
    const int L[3] = {p.L0, p.L1, p.L2};
    p.Q2 = uint32_t(p.L2) * uint32_t(m.ns) * W;
    p.Q1 = p.Q2 * uint32_t(p.L1);
    p.Q0 = p.Q1 * uint32_t(p.L0);
    for (int s = 0; s < m.ns; ++s)
      for (int k = m.sptr[s]; k < m.sptr[s + 1]; ++k) {
        const int* raw = &m.ent_raw[6 * size_t(m.ent_order[k])];
        int dd[3] = {raw[2], raw[3], raw[4]};
        for (int a = 0; a < 3; ++a) dd[a] = ((dd[a] % L[a]) + L[a]) % L[a];
        const int64_t cell = (int64_t(dd[0]) * p.L1 + dd[1]) * p.L2 + dd[2];
        REnt& q = t[size_t(r) * m.nent + k];
        q.D = uint32_t(uint64_t(cell * m.ns + raw[1] - s) * W);  // modulo 2^32, as the kernels add it
        q.th0 = p.L0 - dd[0];
        q.th1 = p.L1 - dd[1];
        q.th2 = p.L2 - dd[2];
      }
  }
  bok(hipMalloc(&b.rent_d, sizeof(REnt) * t.size()), "hipMalloc stencil table");
  bok(hipMemcpy(b.rent_d, t.data(), sizeof(REnt) * t.size(), hipMemcpyHostToDevice), "hipMemcpy stencil table");
}

// Workgroup layout for the current launch shape: tile_rep[logical workgroup] = replica.
void layout(mx_batch& b) {
  std::vector<int32_t> map;
  int t = 0;
  for (int r = 0; r < b.nrep; ++r) {
    Rep& p = b.reps[r];
    const int per = b.block * b.spt;
    p.ntile = b.ns * ((p.ncell + per - 1) / per);
    p.tile0 = t;
    t += p.ntile;
    map.insert(map.end(), p.ntile, r);
  }
  b.ntiles = t;
  bfree(reinterpret_cast<void*&>(b.tile_rep_d));
  bok(hipMalloc(reinterpret_cast<void**>(&b.tile_rep_d), sizeof(int32_t) * map.size()), "hipMalloc tile map");
  bok(hipMemcpy(b.tile_rep_d, map.data(), sizeof(int32_t) * map.size(), hipMemcpyHostToDevice), "hipMemcpy tile map");
  if (size_t(t) > b.part_cap) {
    bfree(reinterpret_cast<void*&>(b.part_d));
    bok(hipMalloc(reinterpret_cast<void**>(&b.part_d), sizeof(double) * MX_OBS_LEN * size_t(t)), "hipMalloc partials");
    b.part_cap = t;
  }
  drop_bgraph(b);
}

// Replaced while batched kernels queued on the caller's (possibly non-blocking) stream may still read the
// table: the device finishes its queued work first (set-up calls only).
void upload_reps(mx_batch& b) {
  bok(hipDeviceSynchronize(), "hipDeviceSynchronize");
  if (!b.reps_d) bok(hipMalloc(&b.reps_d, sizeof(Rep) * b.reps.size()), "hipMalloc replicas");
  bok(hipMemcpy(b.reps_d, b.reps.data(), sizeof(Rep) * b.reps.size(), hipMemcpyHostToDevice), "hipMemcpy replicas");
}

template <typename R, int CM, int BS>
void launch_pair(mx_batch& b, R* e, uint64_t step, const uint64_t* step_ptr, hipStream_t s) {
  BDyn<R> d{R(b.B[0]), R(b.B[1]), R(b.B[2]), R(b.dt), b.precession, b.noisy, b.noise_dbl};
  const BStatic<R>* S = static_cast<const BStatic<R>*>(b.static_d);
  const Rep* reps = static_cast<const Rep*>(b.reps_d);
  R* ep = static_cast<R*>(b.ep_d);
  R* k1 = static_cast<R*>(b.k1_d);
  R* bn = static_cast<R*>(b.bn_d);
  const REnt* rent = static_cast<const REnt*>(b.rent_d);
  if (b.chw == 6) {
    hipLaunchKernelGGL((k_bheun<R, CM, BS, 1, 6>), dim3(b.ntiles), dim3(BS), 0, s, S, reps, b.tile_rep_d, rent, b.nent,
                       d, e, ep, k1, bn, step, b.spt, step_ptr);
    hipLaunchKernelGGL((k_bheun<R, CM, BS, 2, 6>), dim3(b.ntiles), dim3(BS), 0, s, S, reps, b.tile_rep_d, rent, b.nent,
                       d, e, ep, k1, bn, step, b.spt, step_ptr);
  } else {
    hipLaunchKernelGGL((k_bheun<R, CM, BS, 1, 8>), dim3(b.ntiles), dim3(BS), 0, s, S, reps, b.tile_rep_d, rent, b.nent,
                       d, e, ep, k1, bn, step, b.spt, step_ptr);
    hipLaunchKernelGGL((k_bheun<R, CM, BS, 2, 8>), dim3(b.ntiles), dim3(BS), 0, s, S, reps, b.tile_rep_d, rent, b.nent,
                       d, e, ep, k1, bn, step, b.spt, step_ptr);
  }
}

template <typename R, int CM>
void step_cm(mx_batch& b, R* e, uint64_t step, const uint64_t* step_ptr, hipStream_t s) {
  switch (b.block) {
    case 64: launch_pair<R, CM, 64>(b, e, step, step_ptr, s); break;
    case 128: launch_pair<R, CM, 128>(b, e, step, step_ptr, s); break;
    case 512: launch_pair<R, CM, 512>(b, e, step, step_ptr, s); break;
    default: launch_pair<R, CM, 256>(b, e, step, step_ptr, s); break;
  }
}

void bstep(mx_batch& b, void* e, uint64_t step, const uint64_t* step_ptr, hipStream_t s) {
  if (b.prec == MX_FP64) {
    double* E = static_cast<double*>(e);
    if (b.coup == COUP_TABLE1) step_cm<double, COUP_TABLE1>(b, E, step, step_ptr, s);
    else if (b.coup == COUP_TABLE4) step_cm<double, COUP_TABLE4>(b, E, step, step_ptr, s);
    else step_cm<double, COUP_TABLE9>(b, E, step, step_ptr, s);
  } else {
    float* E = static_cast<float*>(e);
    if (b.coup == COUP_TABLE1) step_cm<float, COUP_TABLE1>(b, E, step, step_ptr, s);
    else if (b.coup == COUP_TABLE4) step_cm<float, COUP_TABLE4>(b, E, step, step_ptr, s);
    else step_cm<float, COUP_TABLE9>(b, E, step, step_ptr, s);
  }
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) bfail(MX_EHIP, std::string("batch step: ") + hipGetErrorString(err));
}

template <typename R, int CM, int BS>
void obs_launch(mx_batch& b, const R* e, hipStream_t s) {
  BDyn<R> d{R(b.B[0]), R(b.B[1]), R(b.B[2]), R(b.dt), b.precession, b.noisy, b.noise_dbl};
  hipLaunchKernelGGL((k_bobs<R, CM, BS>), dim3(b.ntiles), dim3(BS), 0, s, static_cast<const BStatic<R>*>(b.static_d),
                     static_cast<const Rep*>(b.reps_d), b.tile_rep_d, static_cast<const REnt*>(b.rent_d), b.nent, d,
                     e, b.spt, b.part_d);
}

template <typename R, int CM>
void obs_cm(mx_batch& b, const R* e, hipStream_t s) {
  switch (b.block) {
    case 64: obs_launch<R, CM, 64>(b, e, s); break;
    case 128: obs_launch<R, CM, 128>(b, e, s); break;
    case 512: obs_launch<R, CM, 512>(b, e, s); break;
    default: obs_launch<R, CM, 256>(b, e, s); break;
  }
}

void ensure_bgraph(mx_batch& b, void* e) {
  std::vector<char> key;
  auto put = [&](const void* p, size_t n) { key.insert(key.end(), (const char*)p, (const char*)p + n); };
  put(&e, sizeof(e));
  put(b.B, sizeof(b.B));
  put(&b.dt, sizeof(b.dt));
  const int ints[7] = {b.precession, b.noisy, b.block, b.spt, b.graph_steps, b.ntiles, b.noise_dbl};
  put(ints, sizeof(ints));
  if (b.gexec && key == b.gkey) return;
  drop_bgraph(b);
  if (!b.counter_d) bok(hipMalloc(reinterpret_cast<void**>(&b.counter_d), sizeof(uint64_t)), "hipMalloc counter");
  if (!b.cap) bok(hipStreamCreateWithFlags(&b.cap, hipStreamNonBlocking), "hipStreamCreate");
  bok(hipStreamBeginCapture(b.cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  for (int n = 0; n < b.graph_steps; ++n) bstep(b, e, uint64_t(n), b.counter_d, b.cap);
  hipLaunchKernelGGL(k_bcounter, dim3(1), dim3(64), 0, b.cap, b.counter_d, uint64_t(b.graph_steps), 1);
  hipGraph_t g = nullptr;
  bok(hipStreamEndCapture(b.cap, &g), "hipStreamEndCapture");
  const hipError_t inst = hipGraphInstantiate(&b.gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  bok(inst, "hipGraphInstantiate");
  b.gkey = key;
}

}  // namespace

extern "C" {

int mx_batch_create(const mx_model* proto, int32_t nrep, const int32_t* L, mx_batch** out) {
  return bguard([&]() -> int {
    if (!proto || !L || !out) bfail(MX_EINVAL, "null argument");
    *out = nullptr;
    if (nrep < 1) bfail(MX_EINVAL, "a batch needs at least one replica");
    if (proto->coup == COUP_BOND || proto->amode != ONSITE_SUB || proto->has_c || proto->has_u ||
        !proto->mu_site.empty() || proto->pinned_d)
      bfail(MX_EUNSUPPORTED, "batched replicas carry the unit-cell model only (no per-bond or per-site data)");
    int reach[3] = {0, 0, 0};
    for (int k = 0; k < proto->nent; ++k)
      for (int a = 0; a < 3; ++a) reach[a] = std::max(reach[a], std::abs(proto->ent_raw[6 * size_t(k) + 2 + a]));
    auto* b = new mx_batch();
    try {
      b->prec = proto->prec;
      b->noise_dbl = proto->noise_dbl;
      b->device = proto->device;
      b->nrep = nrep;
      b->ns = proto->ns;
      b->coup = proto->coup;
      b->mu_sub = proto->mu_sub;
      for (int a = 0; a < 3; ++a) b->B[a] = proto->B[a];
      int64_t off = 0;
      b->reps.resize(nrep);
      for (int r = 0; r < nrep; ++r) {
        Rep& p = b->reps[r];
        std::memset(&p, 0, sizeof(p));
        p.L0 = L[3 * r];
        p.L1 = L[3 * r + 1];
        p.L2 = L[3 * r + 2];
        for (int a = 0; a < 3; ++a)
          if (L[3 * r + a] < 1 || L[3 * r + a] <= reach[a]) bfail(MX_EINVAL, "replica supercell shorter than the stencil reach");
        const int64_t nc = int64_t(p.L0) * p.L1 * p.L2;
        if (off + nc * b->ns > MAX_SITES)
          bfail(MX_EUNSUPPORTED, "more than " + std::to_string(MAX_SITES) +
                                     " sites in a batch (per-site offsets up to 9 N are 32-bit in the kernels)");
        if (nc >= (int64_t(1) << 24)) bfail(MX_EUNSUPPORTED, "a replica of 2^24 cells or more");
        p.rL1 = 1.0f / float(p.L1);
        p.rL2 = 1.0f / float(p.L2);
        p.ncell = int(nc);
        p.off = int(off);
        off += nc * b->ns;
      }
      b->total = off;
      bok(hipSetDevice(b->device), "hipSetDevice");
      if (b->prec == MX_FP64) build_static<double>(*b, *proto);
      else build_static<float>(*b, *proto);
      build_rent(*b, *proto);
      bok(hipMalloc(&b->static_d, b->static_h.size()), "hipMalloc batch constants");
      bok(hipMemcpy(b->static_d, b->static_h.data(), b->static_h.size(), hipMemcpyHostToDevice), "hipMemcpy constants");
      bok(hipMalloc(&b->ep_d, size_t(b->total) * 3 * rsz(*b)), "hipMalloc ep");
      bok(hipMalloc(&b->k1_d, size_t(b->total) * 3 * rsz(*b)), "hipMalloc k1");
      bok(hipMalloc(&b->bn_d, size_t(b->total) * 3 * rsz(*b)), "hipMalloc thermal field");
      layout(*b);
      upload_reps(*b);
    } catch (...) {
      mx_batch_destroy(b);
      throw;
    }
    *out = b;
    return MX_OK;
  });
}

void mx_batch_destroy(mx_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  drop_bgraph(*b);
  for (void** p : {&b->static_d, &b->reps_d, &b->rent_d, &b->ep_d, &b->k1_d, &b->bn_d, reinterpret_cast<void**>(&b->tile_rep_d),
                   reinterpret_cast<void**>(&b->part_d), reinterpret_cast<void**>(&b->counter_d)})
    bfree(*p);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  delete b;
}

int mx_batch_offsets(const mx_batch* b, int64_t* off) {
  return bguard([&]() -> int {
    if (!b || !off) bfail(MX_EINVAL, "null argument");
    for (int r = 0; r < b->nrep; ++r) off[r] = b->reps[r].off;
    off[b->nrep] = b->total;
    return MX_OK;
  });
}

int mx_batch_set_dynamics(mx_batch* b, double dt, const double* alpha, const double* T, int precession) {
  return bguard([&]() -> int {
    if (!b || !alpha || !T) bfail(MX_EINVAL, "null argument");
    if (!(dt > 0)) bfail(MX_EINVAL, "dt must be positive");
    b->dt = dt;
    b->precession = precession ? 1 : 0;
    b->noisy = 0;
    for (int r = 0; r < b->nrep; ++r) {
      if (!(T[r] >= 0)) bfail(MX_EINVAL, "temperatures must be >= 0");
      b->reps[r].alpha = alpha[r];
      // sigma_i = sqrt(2 alpha k_B T / (gamma mu_i mu_B dt)) (magex/llg.py:111), one per sublattice
      for (int q = 0; q < b->ns; ++q)
        b->reps[r].sig[q] = std::sqrt(2.0 * alpha[r] * K_B * T[r] / (GAMMA * b->mu_sub[q] * MU_B * dt));
      b->noisy = b->noisy || T[r] > 0;
    }
    bok(hipSetDevice(b->device), "hipSetDevice");
    upload_reps(*b);
    b->dyn_set = 1;
    return MX_OK;
  });
}

int mx_batch_set_launch(mx_batch* b, int32_t block, int32_t sites_per_thread, int32_t graph_steps) {
  return bguard([&]() -> int {
    if (!b) bfail(MX_EINVAL, "null batch");
    if (block != 64 && block != 128 && block != 256 && block != 512) bfail(MX_EINVAL, "block must be 64, 128, 256 or 512");
    if (sites_per_thread < 1 || sites_per_thread > 8) bfail(MX_EINVAL, "sites per thread must lie in [1, 8]");
    if (graph_steps < 0 || graph_steps > 4096) bfail(MX_EINVAL, "graph_steps must lie in [0, 4096]");
    bok(hipSetDevice(b->device), "hipSetDevice");
    const bool relayout = block != b->block || sites_per_thread != b->spt;
    b->block = block;
    b->spt = sites_per_thread;
    b->graph_steps = graph_steps;
    if (relayout) {
      layout(*b);
      upload_reps(*b);
    }
    drop_bgraph(*b);
    return MX_OK;
  });
}

int mx_batch_run(mx_batch* b, void* e, int64_t nsteps, const uint64_t* seeds, uint64_t step0, void* stream) {
  return bguard([&]() -> int {
    if (!b || !e || !seeds) bfail(MX_EINVAL, "null argument");
    if (nsteps < 0) bfail(MX_EINVAL, "negative step count");
    if (!b->dyn_set) bfail(MX_EINVAL, "set the dynamics first");
    bok(hipSetDevice(b->device), "hipSetDevice");
    bool changed = false;
    for (int r = 0; r < b->nrep; ++r) {
      const uint32_t lo = uint32_t(seeds[r]), hi = uint32_t(seeds[r] >> 32);
      changed = changed || lo != b->reps[r].seed_lo || hi != b->reps[r].seed_hi;
      b->reps[r].seed_lo = lo;
      b->reps[r].seed_hi = hi;
    }
    if (changed) upload_reps(*b);
    hipStream_t s = static_cast<hipStream_t>(stream);
    int64_t n = 0;
    const int G = b->graph_steps;
    if (G > 0 && nsteps >= 2 * int64_t(G)) {
      ensure_bgraph(*b, e);
      hipLaunchKernelGGL(k_bcounter, dim3(1), dim3(64), 0, s, b->counter_d, step0, 0);
      for (; n + G <= nsteps; n += G) bok(hipGraphLaunch(b->gexec, s), "hipGraphLaunch");
    }
    for (; n < nsteps; ++n) bstep(*b, e, step0 + uint64_t(n), nullptr, s);
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess) bfail(MX_EHIP, std::string("batch run: ") + hipGetErrorString(err));
    return MX_OK;
  });
}

int mx_batch_observables(mx_batch* b, const void* e, double* out, void* stream) {
  return bguard([&]() -> int {
    if (!b || !e || !out) bfail(MX_EINVAL, "null argument");
    bok(hipSetDevice(b->device), "hipSetDevice");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (b->prec == MX_FP64) {
      const double* E = static_cast<const double*>(e);
      if (b->coup == COUP_TABLE1) obs_cm<double, COUP_TABLE1>(*b, E, s);
      else if (b->coup == COUP_TABLE4) obs_cm<double, COUP_TABLE4>(*b, E, s);
      else obs_cm<double, COUP_TABLE9>(*b, E, s);
    } else {
      const float* E = static_cast<const float*>(e);
      if (b->coup == COUP_TABLE1) obs_cm<float, COUP_TABLE1>(*b, E, s);
      else if (b->coup == COUP_TABLE4) obs_cm<float, COUP_TABLE4>(*b, E, s);
      else obs_cm<float, COUP_TABLE9>(*b, E, s);
    }
    hipLaunchKernelGGL(k_bfinal, dim3(b->nrep), dim3(BLOCK), 0, s, static_cast<const Rep*>(b->reps_d), b->part_d, out);
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess) bfail(MX_EHIP, std::string("batch observables: ") + hipGetErrorString(err));
    return MX_OK;
  });
}

}  // extern "C"
