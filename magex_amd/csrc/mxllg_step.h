// Device helpers of the one-launch Heun steps (k_step and k_walk in mxllg_kernels.hip, the persistent
// k_steps in mxllg_steps.hip) and the host-side helpers their launchers share: plane-counter waits, the
// state a workgroup holds from its predictor to its corrector, spin sources (spin arrays or tagged
// records), the end-of-launch fallback; on the host, launch checks, the per-launch values (Dyn) of a
// model, and the run-time dispatch over coupling modes and geometries.
#pragma once
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "mxllg_field.h"

namespace mx {

// Stores of a spin that other workgroups of the same launch read (agent scope: global_store ... sc1).
template <typename R>
__device__ __forceinline__ void store3_coherent(R* p, R x, R y, R z) {
  __hip_atomic_store(p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(p + 1, y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(p + 2, z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------------------------------------------------ one-launch step
// Wavefront 0 polls the plane counters `cnt` of planes [lo, lo + n) (wrapped onto [0, L0)) until each
// reaches `times` x the number of workgroups whose tiles cover the plane; false after `spin` unsuccessful
// polls or once the launch's abort word (ctl[3]) is raised.
template <typename R>
__device__ __forceinline__ bool wait_planes(const Static<R>* __restrict__ S, const StepSync& sy,
                                            const uint32_t* cnt, uint32_t times, int lo, int n) {
  const int L0 = S->L0, PS = sy.PS, ns = S->ns, lane = threadIdx.x & 63;
  for (uint32_t it = 0;; ++it) {
    int ok = 1;
    for (int k = lane; k < n; k += 64) {
      int P = lo + k;
      P = P >= L0 ? P - L0 : P;
      const uint32_t want = times * uint32_t(ns * (((P + 1) * PS - 1) / BLOCK - (P * PS) / BLOCK + 1));
      const uint32_t c = __hip_atomic_load(const_cast<uint32_t*>(cnt) + P * CNT_STRIDE, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
      ok = ok && c >= want;
    }
    if (__all(ok)) return true;
    if (it >= sy.spin) return false;
    if ((it & 15) == 15 && __hip_atomic_load(sy.ctl + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
    __builtin_amdgcn_s_sleep(8);
  }
}

// State a workgroup of the one-launch kernels keeps per site from the predictor to the corrector: the
// coupling batches (per-bond tensors when NB > 0), the per-site on-site tensor, damping, the spin e, the
// predictor spin ep, w = e + dt/2 k1 (Heun) or the rotation vector (Depondt), the thermal field, and the
// effective field of the last pass (kept for recorded runs).
template <typename R, int CM, int GEO, int NB>
struct Held {
  Batch<R, CM, GEO> bt[NB > 0 ? NB : 1];
  SiteA<R> A;
  R a, x, y, z, px, py, pz, wx, wy, wz, nx, ny, nz, bx, by, bz;
  // tagged gathers of FP32 table couplings over at most two batches: the entries' record indices and
  // live-and-inside masks, computed once per launch (hidx) instead of per gather.  Measured in persistent
  // runs: triangular 256^2 FP32 10.3 -> 8.9 us per step, sc 16^3 FP32 4.5 -> 4.1; in FP64 the registers
  // they take cost more than the arithmetic they save (sc 16^3: 6.1 -> 6.4), so FP64 does not hold them.
  static constexpr bool HOLD = sizeof(R) == 4 && CM != COUP_BOND;
  static constexpr int CHB = Batch<R, CM, GEO>::CH;
  int j0[CHB], j1[CHB];
  uint32_t ok0, ok1;
  bool hidx;
};

// Where the one-launch kernels take the neighbours' spins from: a spin array read with load policy LD
// (SpinsFrom), or tagged records polled until they carry the tag `want` (TaggedFrom).
template <typename R, int LD>
struct SpinsFrom {
  static constexpr bool tagged = false;
  static constexpr int LD_ = LD;
  const R* e;
};
struct TaggedFrom {
  static constexpr bool tagged = true;
  const uint64_t* rec;
  uint32_t want, spin;
  const uint32_t* abort;
};

// Fills h's held record indices when the sublattice's stencil [kb, ke) is one or two table batches.
template <typename R, int CM, int GEO, int NB>
__device__ __forceinline__ void hold_indices(const Static<R>* __restrict__ S, const Site& st, int kb, int ke,
                                             Held<R, CM, GEO, NB>& h) {
  constexpr int CH = Batch<R, CM, GEO>::CH;
  h.hidx = Held<R, CM, GEO, NB>::HOLD && ke > kb && ke - kb <= 2 * CH;  // wavefront-uniform
  if constexpr (!Held<R, CM, GEO, NB>::HOLD) return;
  if (!h.hidx) return;
  Batch<R, CM, GEO>::tagged_at(S, st, kb, ke, h.j0, h.ok0);
  h.ok1 = 0;
  if (kb + CH < ke) Batch<R, CM, GEO>::tagged_at(S, st, kb + CH, ke, h.j1, h.ok1);
}

// The neighbour field of the tagged gathers with held indices (hold_indices): one or two batches, with
// the same loads, order and sums as the batch loop of predictor_held / corrector_held.
template <typename R, int CM, int GEO, int NB, typename Mid>
__device__ __forceinline__ bool tagged_field_held(const Static<R>* __restrict__ S, const TaggedFrom& src, int kb,
                                                  int ke, const Held<R, CM, GEO, NB>& h, R& hx, R& hy, R& hz,
                                                  Mid&& mid) {
  constexpr int CH = Batch<R, CM, GEO>::CH;
  bool ok = true;
  if constexpr (Held<R, CM, GEO, NB>::HOLD) {
    {
      Batch<R, CM, GEO> bt;
      const bool g = bt.gather_tagged_at(S, src.rec, kb, ke, h.j0, h.ok0, src.want, src.spin, src.abort, mid);
      ok = ok && g;
      bt.accumulate(S, hx, hy, hz);
    }
    if (kb + CH < ke) {  // wavefront-uniform
      Batch<R, CM, GEO> bt;
      const bool g = bt.gather_tagged_at(S, src.rec, kb + CH, ke, h.j1, h.ok1, src.want, src.spin, src.abort, [] {});
      ok = ok && g;
      bt.accumulate(S, hx, hy, hz);
    }
  }
  return ok;
}

template <typename R, int CM, int GEO, bool WITH_J, typename Src>
__device__ __forceinline__ bool fill_batch(Batch<R, CM, GEO>& bt, const Static<R>* __restrict__ S, const Dyn<R>& d,
                                           const Src& src, const Site& st, int k0, int ke) {
  if constexpr (Src::tagged) {
    return bt.template gather_tagged<WITH_J>(S, d, src.rec, st, k0, ke, src.want, src.spin, src.abort, [] {});
  } else {
    bt.template gather<Src::LD_, WITH_J>(S, d, src.e, st, k0, ke);
    return true;
  }
}

// Predictor of site st (magex/fast.py:336-368) with the arithmetic of k_heun's pass 1: gathers the
// neighbours' spins from src (with the couplings unless LOADJ is false, i.e. already held), draws the
// thermal field of `step`, and leaves ep and w (or the rotation vector) in h.  h.x, h.y, h.z hold e.
// False if a tagged gather gave up.
template <typename R, int CM, int GEO, int NB, bool LEAN, bool LOADJ, typename Src>
__device__ __forceinline__ bool predictor_held(const Static<R>* __restrict__ S, const Dyn<R>& d, const Src& src,
                                               const Site& st, int kb, int ke, uint64_t step,
                                               const R* __restrict__ noise, Held<R, CM, GEO, NB>& h) {
  constexpr int CH = Batch<R, CM, GEO>::CH;
  R hx = R(0), hy = R(0), hz = R(0), gx, gy, gz, bx, by, bz, fx, fy, fz;
  bool ok = true;
  h.nx = h.ny = h.nz = R(0);
  if constexpr (NB > 0) {
    // the on-site tensor, then every held batch, in flight before the first is used
    if (LOADJ) h.A.load(S, d, st);
#pragma unroll
    for (int n = 0; n < NB; ++n) {
      const bool g = fill_batch<R, CM, GEO, LOADJ>(h.bt[n], S, d, src, st, kb + n * CH, ke);
      ok = ok && g;
    }
    if (d.noisy) thermal<R, LEAN>(S, d, st, step, noise, h.nx, h.ny, h.nz);
#pragma unroll
    for (int n = 0; n < NB; ++n)
      if (kb + n * CH < ke) h.bt[n].accumulate(S, hx, hy, hz);
  } else if constexpr (Src::tagged) {
    // the thermal field and the on-site gradient (the site's own) while the first poll is in flight
    if (LOADJ) h.A.load(S, d, st);
    auto mid = [&] {
      if (d.noisy) thermal<R, LEAN>(S, d, st, step, noise, h.nx, h.ny, h.nz);
      onsite_grad<R, LEAN>(S, d, st, h.A, h.x, h.y, h.z, gx, gy, gz);
    };
    if (kb >= ke) mid();
    if (Held<R, CM, GEO, NB>::HOLD && h.hidx) {
      ok = tagged_field_held<R, CM, GEO, NB>(S, src, kb, ke, h, hx, hy, hz, mid);
    } else {
      for (int k0 = kb; k0 < ke; k0 += CH) {
        Batch<R, CM, GEO> bt;
        const bool g = k0 == kb ? bt.template gather_tagged<true>(S, d, src.rec, st, k0, ke, src.want, src.spin,
                                                                  src.abort, mid)
                                : bt.template gather_tagged<true>(S, d, src.rec, st, k0, ke, src.want, src.spin,
                                                                  src.abort, [] {});
        ok = ok && g;
        bt.accumulate(S, hx, hy, hz);
      }
    }
  } else {
    exchange<R, CM, GEO, 0, Src::LD_>(S, d, src.e, st, hx, hy, hz, [&] {
      if (d.noisy) thermal<R, LEAN>(S, d, st, step, noise, h.nx, h.ny, h.nz);
    });
  }
  const R x = h.x, y = h.y, z = h.z;
  if constexpr (NB > 0) {
    onsite_grad<R, LEAN>(S, d, st, h.A, x, y, z, gx, gy, gz);
  } else if constexpr (!Src::tagged) {
    if (LOADJ) h.A.load(S, d, st);
    onsite_grad<R, LEAN>(S, d, st, h.A, x, y, z, gx, gy, gz);
  }
  field_from<R, LEAN>(S, d, st, x, y, z, hx, hy, hz, gx, gy, gz, h.nx, h.ny, h.nz, d.B0, d.B1, d.B2, bx, by, bz);
  h.bx = bx;
  h.by = by;
  h.bz = bz;
  llg_rhs(x, y, z, bx, by, bz, h.a, d.prec, fx, fy, fz);
  if (!LEAN && d.integ == MX_DEPONDT) {
    omega(x, y, z, fx, fy, fz, h.wx, h.wy, h.wz);
    h.px = x;
    h.py = y;
    h.pz = z;
    rotate(h.px, h.py, h.pz, d.dt * h.wx, d.dt * h.wy, d.dt * h.wz);
  } else {
    const R hdt = R(0.5) * d.dt;
    h.wx = x + hdt * fx;
    h.wy = y + hdt * fy;
    h.wz = z + hdt * fz;
    const R qx = x + d.dt * fx, qy = y + d.dt * fy, qz = z + d.dt * fz;
    const R q = rnorm3(qx, qy, qz);
    h.px = qx * q;
    h.py = qy * q;
    h.pz = qz * q;
  }
  return ok;
}

// Corrector of site st (magex/fast.py:370-390) with the arithmetic of k_heun's pass 2, from the held
// state and the neighbours' predictor spins from src; the new spin goes to h.x, h.y, h.z.  False if a
// tagged gather gave up.
template <typename R, int CM, int GEO, int NB, bool LEAN, typename Src>
__device__ __forceinline__ bool corrector_held(const Static<R>* __restrict__ S, const Dyn<R>& d, const Src& src,
                                               const Site& st, int kb, int ke, Held<R, CM, GEO, NB>& h) {
  constexpr int CH = Batch<R, CM, GEO>::CH;
  R hx = R(0), hy = R(0), hz = R(0), gx, gy, gz, bx, by, bz, fx, fy, fz;
  bool ok = true;
  if constexpr (NB > 0) {
#pragma unroll
    for (int n = 0; n < NB; ++n) {
      const bool g = fill_batch<R, CM, GEO, false>(h.bt[n], S, d, src, st, kb + n * CH, ke);
      ok = ok && g;
    }
#pragma unroll
    for (int n = 0; n < NB; ++n)
      if (kb + n * CH < ke) h.bt[n].accumulate(S, hx, hy, hz);
  } else if constexpr (Src::tagged) {
    // the on-site gradient at the site's own predictor spin while the first poll is in flight
    auto mid = [&] { onsite_grad<R, LEAN>(S, d, st, h.A, h.px, h.py, h.pz, gx, gy, gz); };
    if (kb >= ke) mid();
    if (Held<R, CM, GEO, NB>::HOLD && h.hidx) {
      ok = tagged_field_held<R, CM, GEO, NB>(S, src, kb, ke, h, hx, hy, hz, mid);
    } else {
      for (int k0 = kb; k0 < ke; k0 += CH) {
        Batch<R, CM, GEO> bt;
        const bool g = k0 == kb ? bt.template gather_tagged<true>(S, d, src.rec, st, k0, ke, src.want, src.spin,
                                                                  src.abort, mid)
                                : bt.template gather_tagged<true>(S, d, src.rec, st, k0, ke, src.want, src.spin,
                                                                  src.abort, [] {});
        ok = ok && g;
        bt.accumulate(S, hx, hy, hz);
      }
    }
  } else {
    exchange<R, CM, GEO, 0, Src::LD_>(S, d, src.e, st, hx, hy, hz, [] {});
  }
  const R px = h.px, py = h.py, pz = h.pz;
  if constexpr (!(Src::tagged && NB == 0)) onsite_grad<R, LEAN>(S, d, st, h.A, px, py, pz, gx, gy, gz);
  field_from<R, LEAN>(S, d, st, px, py, pz, hx, hy, hz, gx, gy, gz, h.nx, h.ny, h.nz, d.Bc0, d.Bc1, d.Bc2, bx, by,
                      bz);
  h.bx = bx;
  h.by = by;
  h.bz = bz;
  llg_rhs(px, py, pz, bx, by, bz, h.a, d.prec, fx, fy, fz);
  const R hdt = R(0.5) * d.dt;
  R cx, cy, cz;
  if (!LEAN && d.integ == MX_DEPONDT) {
    R vx, vy, vz;
    omega(px, py, pz, fx, fy, fz, vx, vy, vz);
    cx = h.x;
    cy = h.y;
    cz = h.z;
    rotate(cx, cy, cz, hdt * (h.wx + vx), hdt * (h.wy + vy), hdt * (h.wz + vz));
  } else {
    cx = h.wx + hdt * fx;
    cy = h.wy + hdt * fy;
    cz = h.wz + hdt * fz;
  }
  const R q = rnorm3(cx, cy, cz);
  h.x = cx * q;
  h.y = cy * q;
  h.z = cz * q;
  return ok;
}

// Planes a workgroup's stencil reaches: [lo, lo + n) wrapped onto [0, L0), from its tile's planes
// [pa, pb] and the reach r0 along axis 0.
template <typename R>
__device__ __forceinline__ void wait_window(const Static<R>* __restrict__ S, int pa, int pb, int r0, int& lo, int& n) {
  const int L0 = S->L0;
  lo = pa - r0;
  n = pb - pa + 1 + 2 * r0;
  if (S->pbc0) {
    if (n >= L0) {
      lo = 0;
      n = L0;
    } else if (lo < 0) {
      lo += L0;
    }
  } else {
    const int hi = min(L0 - 1, pb + r0);
    lo = max(lo, 0);
    n = hi - lo + 1;
  }
}

// End of a one-launch step (k_step, k_walk), called by every thread of every workgroup once its own
// work is done (gave: this workgroup gave up a wait and listed tiles in sy.gu): the last workgroup to
// get here runs the correctors of the listed tiles (every predictor is done by then; w is in k1) and
// clears the control words for the next launch (with k_walk's ticket counters, tick).
template <typename R, int CM, int GEO, bool LEAN>
__device__ __forceinline__ void finish_step(const Static<R>* __restrict__ S, const Dyn<R>& d, R* __restrict__ e,
                                            const R* __restrict__ ep, const R* __restrict__ k1, uint64_t step,
                                            const R* __restrict__ noise, const StepSync& sy, bool gave,
                                            int& sh_flag, uint32_t* tick = nullptr) {
  if (threadIdx.x == 0) {
    const uint32_t done = gave ? __hip_atomic_fetch_add(sy.ctl, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT)
                               : __hip_atomic_fetch_add(sy.ctl, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    sh_flag = done == gridDim.x - 1 ? 1 : 0;
  }
  __syncthreads();
  if (!sh_flag) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  const uint32_t ngu = __hip_atomic_load(sy.ctl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (uint32_t g = 0; g < ngu; ++g) {
    const int b = __hip_atomic_load(sy.gu + g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    Site s2;
    int t2;
    if (step_site<R>(S, b, s2, t2)) {
      const int j = s2.i;
      if (LEAN || !(d.pinned && d.pinned[j])) {
        const R a2 = (!LEAN && d.alpha) ? d.alpha[j] : S->alpha_sub[s2.s];
        corrector<R, CM, GEO, LEAN>(S, d, e, ep, k1, s2, a2, step, noise, d.Bc0, d.Bc1, d.Bc2, nullptr);
      }
    }
  }
  for (int P = threadIdx.x; P < S->L0; P += BLOCK)
    __hip_atomic_store(sy.cnt + P * CNT_STRIDE, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (tick && threadIdx.x < 8)
    __hip_atomic_store(tick + threadIdx.x * CNT_STRIDE, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (ngu) __hip_atomic_fetch_add(sy.ctl + 2, ngu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(sy.ctl + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(sy.ctl, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Tile list of a workgroup that gave up a wait: w of the tile is in k1 (a release makes it, and the
// tile's predictor spins, visible to the last workgroup), and the tile joins sy.gu.
__device__ __forceinline__ void give_up_tile(const StepSync& sy, int tile_block) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t k = __hip_atomic_fetch_add(sy.ctl + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(sy.gu + k, int32_t(tile_block), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}


// ------------------------------------------------------------------------------------------ host helpers

static int check_launch(const char* what) {
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    set_error(std::string(what) + ": " + hipGetErrorString(err));
    return MX_EHIP;
  }
  return MX_OK;
}

// Column map of the step kernels (row-major tiles per XCD per axis-0 plane) when an axis-0 plane is a
// whole number of tiles, 8 or more per XCD; 0: slabs.
static int xcol_of(const Model& m) {
  const long plane = long(m.L[1]) * m.L[2];
  return ((m.geo == GEO_PERIODIC || m.geo == GEO_OPEN) && plane % BLOCK == 0 && (plane / BLOCK) % 8 == 0)
             ? int(plane / BLOCK / 8) : 0;
}

template <typename R>
static Dyn<R> make_dyn(const Model& m, uint64_t seed) {
  Dyn<R> d;
  d.B0 = R(m.B[0]);
  d.B1 = R(m.B[1]);
  d.B2 = R(m.B[2]);
  d.dt = R(m.dt);
  d.prec = m.precession;
  d.noisy = m.noisy;
  d.integ = m.integrator;
  d.sot = (m.sot_dl != 0.0 || m.sot_fl != 0.0) ? 1 : 0;
  // measured on bcc 64^3 per-bond: the reversed corrector walk is 7 % faster in FP64, and 2 % in FP32
  // once the per-bond gathers load their next records ahead (56.1 against 57.3 us per step)
  d.rev2 = 1;
  d.sp0 = R(m.sot_p[0]);
  d.sp1 = R(m.sot_p[1]);
  d.sp2 = R(m.sot_p[2]);
  d.bdl = R(m.sot_dl);
  d.bfl = R(m.sot_fl);
  d.seed_lo = uint32_t(seed);
  d.seed_hi = uint32_t(seed >> 32);
  d.bond = static_cast<const R*>(m.bond_d);
  d.A = static_cast<const R*>(m.A_d);
  d.Kc = static_cast<const R*>(m.Kc_d);
  d.K2 = static_cast<const R*>(m.K2_d);
  d.K3 = static_cast<const R*>(m.K3_d);
  d.inv_mu = static_cast<const R*>(m.inv_mu_d);
  d.alpha = static_cast<const R*>(m.alpha_d);
  d.sigma = static_cast<const R*>(m.sigma_d);
  d.sign_site = static_cast<const R*>(m.sign_d);
  d.pinned = static_cast<const uint8_t*>(m.pinned_d);
  d.step_ptr = nullptr;
  d.save_e = nullptr;
  d.save_B = nullptr;
  d.extra = nullptr;
  d.Bc0 = d.B0;
  d.Bc1 = d.B1;
  d.Bc2 = d.B2;
  return d;
}

static inline int nblk_of(const Model& m) { return m.ns * ((m.ncell + BLOCK - 1) / BLOCK); }

// The model the LEAN step kernels are compiled for (injected noise is checked by the caller).
static inline bool lean_model(const Model& m) {
  return m.integrator == MX_HEUN && !m.pinned_d && !m.alpha_d && !m.sigma_d && !m.inv_mu_d && !m.has_c &&
         !m.has_u && m.sot_dl == 0.0 && m.sot_fl == 0.0;
}

// Dispatch a kernel template over (coupling mode, geometry) at run time.
#define MX_GEO(m, CALL, CM)                                       \
  do {                                                            \
    if ((m).geo == GEO_TILED) { CALL(CM, GEO_TILED); }            \
    else if ((m).geo == GEO_CHECKER) { CALL(CM, GEO_CHECKER); }   \
    else if ((m).geo == GEO_PERIODIC) { CALL(CM, GEO_PERIODIC); } \
    else { CALL(CM, GEO_OPEN); }                                  \
  } while (0)
#define MX_DISPATCH(m, CALL)                                      \
  do {                                                            \
    switch ((m).coup) {                                           \
      case COUP_TABLE1: MX_GEO(m, CALL, COUP_TABLE1); break;      \
      case COUP_TABLE4: MX_GEO(m, CALL, COUP_TABLE4); break;      \
      case COUP_TABLE9: MX_GEO(m, CALL, COUP_TABLE9); break;      \
      default: MX_GEO(m, CALL, COUP_BOND); break;                 \
    }                                                             \
  } while (0)

// ------------------------------------------------------------------------------------------ one-launch step (host)
// Stencil reach along axis 0, in planes (offsets on a periodic axis are stored reduced into [0, L0)).
static int reach0(const Model& m) {
  int r = 0;
  for (const Ent& en : m.ent) {
    int d0 = int(int16_t(en.d01 & 0xffff));
    if (m.pbc[0]) d0 = std::min(d0, m.L[0] - d0);
    r = std::max(r, std::abs(d0));
  }
  return std::min(r, m.L[0]);
}

// The walked step's 8 ticket counters (one per 128-byte line) follow the give-up slots.
static size_t stepsync_tick(const Model& m) {
  const size_t w = stepsync_ctl(m) + CTL_WORDS + size_t(nblk_of(m));
  return (w + CNT_STRIDE - 1) / CNT_STRIDE * CNT_STRIDE;
}

}  // namespace mx
