// gfx950 persistent multi-step launch of the atomistic LLG hot path (k_steps): runs of Heun / Depondt
// steps in one cooperative launch for grids the GPU holds at once, with spins exchanged through tagged
// records, and its host launcher.
#include "mxllg_step.h"

namespace mx {

// nsteps stochastic Heun (or Depondt) steps in one launch, for grids that are resident all at once
// (cooperative launch: every workgroup runs from the start).  Each workgroup owns its tile for the whole
// launch and keeps its sites' state (spin, damping, on-site tensor) in registers across every step; the
// couplings are fetched with each gather (at these sizes they stay in cache, and their loads overlap the
// polls).  Holding per-bond tensors across steps as well took every register of the kernel and gave
// wrong results in the non-lean FP32 instantiation, while the lean one and every other form matched.
// Sites exchange spins through tagged records instead of plane
// counters: every site publishes its spin (tag k = tag0 + 2n), its predictor spin (k + 1) and its
// corrected spin (k + 2) with agent-scope 64-bit stores, each word carrying the tag beside its payload,
// and each gather polls its neighbours' records until every word carries the tag it needs -- one
// store-to-load latency per phase, with no acknowledgement wait, barrier or counter.
//
// Why a record is never overwritten before every reader has it: site b publishes its predictor spin of
// step n + 1 only after reading the corrected spins of step n of all its neighbours, and a neighbour c
// publishes that only after reading b's predictor spin of step n; likewise for corrected spins (the
// stencil is symmetric: c reads b whenever b reads c).  So no tag is ever passed over.  Pinned sites take
// part in every gather (their own spin stays as it is) so that this holds for them too.  Lanes of one
// wavefront publish together after each phase, so a lane never waits for a lane of its own wavefront.
// tag0 differs from every tag of the previous launch on these records (a fresh epoch per launch), so no
// record left by an earlier launch matches.
//
// Every poll is bounded: a lane that gives up raises the launch's abort word, which ends every other
// lane's polls; the launch's last workgroup counts the aborted launch and clears the control words.
// Same arithmetic as k_heun: same result bits.
//
// out (null bases: none) receives every step's record as the single-pass kernels store it in a recorded
// replay (k_heun with save_e / save_B): the step's input spin, its predictor spin, w (the rotation vector
// under Depondt), and the effective field of each pass, with a pinned site's fields zero and its w its spin.
template <typename R>
struct RecPtrs {
  R *e, *ep, *w, *b1, *b2;
};

template <typename R, int CM, int GEO, bool LEAN>
__global__ __launch_bounds__(BLOCK) void k_steps(const Static<R>* __restrict__ S, const Dyn<R> d, R* __restrict__ e,
                                                 uint64_t* __restrict__ rec, uint64_t step0, int nsteps,
                                                 const R* __restrict__ noise, uint32_t tag0, const StepSync sy,
                                                 const RecPtrs<R> out) {
  __shared__ int sh_flag;
  constexpr int NW = TAG_WORDS<R>;
  Site st;
  int tile;
  const bool live = step_site<R>(S, blockIdx.x, st, tile);
  const int i = st.i;
  const bool pin = live && !LEAN && d.pinned && d.pinned[i];
  const int kb = S->sptr[st.s], ke = S->sptr[st.s + 1];
  uint64_t* recE = rec;
  uint64_t* recP = rec + size_t(S->N) * NW;
  const uint32_t* abort = sy.ctl + 3;
  const size_t nstride = size_t(S->N) * 3;
  bool ok = true;
  if (live) {
    Held<R, CM, GEO, 0> h;
    hold_indices<R, CM, GEO, 0>(S, st, kb, ke, h);  // table stencils of up to two batches: indices held
    h.a = (!LEAN && d.alpha) ? d.alpha[i] : S->alpha_sub[st.s];
    h.x = e[3 * i];
    h.y = e[3 * i + 1];
    h.z = e[3 * i + 2];
    publish_tagged<R>(recE + size_t(i) * NW, h.x, h.y, h.z, tag0);
    for (int n = 0; n < nsteps; ++n) {
      const uint64_t step = step0 + uint64_t(n);
      const R* nz = noise ? noise + size_t(n) * nstride : nullptr;
      const uint32_t k = tag0 + 2u * uint32_t(n);
      const R x0 = h.x, y0 = h.y, z0 = h.z;
      const TaggedFrom fromE{recE, k, sy.spin, abort};
      bool g = n == 0 ? predictor_held<R, CM, GEO, 0, LEAN, true>(S, d, fromE, st, kb, ke, step, nz, h)
                      : predictor_held<R, CM, GEO, 0, LEAN, false>(S, d, fromE, st, kb, ke, step, nz, h);
      if (pin) {
        h.px = x0;
        h.py = y0;
        h.pz = z0;
      }
      const size_t ro = size_t(n) * nstride + 3 * size_t(i);
      if (out.e) {  // uniform
        out.e[ro] = x0;
        out.e[ro + 1] = y0;
        out.e[ro + 2] = z0;
        out.ep[ro] = h.px;
        out.ep[ro + 1] = h.py;
        out.ep[ro + 2] = h.pz;
        out.w[ro] = pin ? x0 : h.wx;
        out.w[ro + 1] = pin ? y0 : h.wy;
        out.w[ro + 2] = pin ? z0 : h.wz;
        out.b1[ro] = pin ? R(0) : h.bx;
        out.b1[ro + 1] = pin ? R(0) : h.by;
        out.b1[ro + 2] = pin ? R(0) : h.bz;
      }
      publish_tagged<R>(recP + size_t(i) * NW, h.px, h.py, h.pz, k + 1);
      const TaggedFrom fromP{recP, k + 1, sy.spin, abort};
      const bool g2 = corrector_held<R, CM, GEO, 0, LEAN>(S, d, fromP, st, kb, ke, h);
      g = g && g2;
      if (pin) {
        h.x = x0;
        h.y = y0;
        h.z = z0;
      }
      if (out.e) {
        out.b2[ro] = pin ? R(0) : h.bx;
        out.b2[ro + 1] = pin ? R(0) : h.by;
        out.b2[ro + 2] = pin ? R(0) : h.bz;
      }
      publish_tagged<R>(recE + size_t(i) * NW, h.x, h.y, h.z, k + 2);
      if (!g) {
        ok = false;
        break;
      }
    }
    if (!pin) {
      e[3 * i] = h.x;
      e[3 * i + 1] = h.y;
      e[3 * i + 2] = h.z;
    }
  }
  if (!ok) __hip_atomic_store(const_cast<uint32_t*>(abort), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  // ---- the last workgroup to finish counts an aborted launch and clears the control words
  const int bad = __syncthreads_or(ok ? 0 : 1);
  if (threadIdx.x == 0) {
    const uint32_t done = bad ? __hip_atomic_fetch_add(sy.ctl, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT)
                              : __hip_atomic_fetch_add(sy.ctl, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    sh_flag = done == gridDim.x - 1 ? 1 : 0;
  }
  __syncthreads();
  if (!sh_flag || threadIdx.x != 0) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  if (__hip_atomic_load(sy.ctl + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    __hip_atomic_fetch_add(sy.ctl + 2, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(sy.ctl + 3, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(sy.ctl, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------------------------------------------------ launchers
// Bytes one round of a persistent launch's polls reads: every site reads the tagged records of its
// whole stencil.
static size_t poll_bytes(const Model& m) {
  int mx = 0;
  for (int s = 0; s < m.ns; ++s) mx = std::max(mx, m.sptr[s + 1] - m.sptr[s]);
  return size_t(m.N) * size_t(mx) * (m.prec == MX_FP64 ? 4 : 2) * sizeof(uint64_t);
}

// Persistent launches need every workgroup resident at once (at most one per compute unit; each kernel
// holds at least one) and row-major cells.  They pay where a round of polls is small: measured on
// MI355X, sc 16^3 FP64 (0.8 MB per round) 9.5 -> 6.3 us per step against one launch per step,
// triangular 256^2 FP32 (19 MB) 15.1 -> 10.3, but triangular 256^2 FP64 (38 MB) 15.2 -> 18.8 against two
// launches per step: repeated polls of that much state through the agent-coherent path cost more than
// the launches they save.
constexpr size_t PERSISTENT_POLL_BYTES = size_t(24) << 20;
bool persistent_steps(const Model& m) {
  if (m.step_mode < 2 || !(m.geo == GEO_PERIODIC || m.geo == GEO_OPEN)) return false;
  if (poll_bytes(m) > PERSISTENT_POLL_BYTES) return false;
  int dev_cu = 0;
  if (hipDeviceGetAttribute(&dev_cu, hipDeviceAttributeMultiprocessorCount, m.device) != hipSuccess || dev_cu <= 0)
    return false;
  return nblk_of(m) <= dev_cu;
}

template <typename R, int CM, int GEO, bool LEAN>
static hipError_t steps_coop(dim3 g, hipStream_t s, const Static<R>* S, Dyn<R> d, R* E, uint64_t* rec,
                             uint64_t step0, int n, const R* nz, uint32_t tag0, StepSync sy, RecPtrs<R> ro,
                             bool plain) {
  void* args[] = {&S, &d, &E, &rec, &step0, &n, &nz, &tag0, &sy, &ro};
  if (plain)  // with more LDS than half a compute unit's, no two workgroups of the side-by-side launches share one
    return hipLaunchKernel(reinterpret_cast<const void*>(k_steps<R, CM, GEO, LEAN>), g, dim3(BLOCK), args,
                           SIDE_BY_SIDE_LDS, s);
  return hipLaunchCooperativeKernel(k_steps<R, CM, GEO, LEAN>, g, dim3(BLOCK), args, 0, s);
}

template <typename R, int GEO, bool LEAN>
static hipError_t steps_dispatch(const Model& m, dim3 g, hipStream_t s, const Static<R>* S, const Dyn<R>& d, R* E,
                                 uint64_t* rec, uint64_t step0, int n, const R* nz, uint32_t tag0,
                                 const StepSync& sy, const RecPtrs<R>& ro, bool plain) {
  switch (m.coup) {
    case COUP_TABLE1:
      return steps_coop<R, COUP_TABLE1, GEO, LEAN>(g, s, S, d, E, rec, step0, n, nz, tag0, sy, ro, plain);
    case COUP_TABLE4:
      return steps_coop<R, COUP_TABLE4, GEO, LEAN>(g, s, S, d, E, rec, step0, n, nz, tag0, sy, ro, plain);
    case COUP_TABLE9:
      return steps_coop<R, COUP_TABLE9, GEO, LEAN>(g, s, S, d, E, rec, step0, n, nz, tag0, sy, ro, plain);
    default: return steps_coop<R, COUP_BOND, GEO, LEAN>(g, s, S, d, E, rec, step0, n, nz, tag0, sy, ro, plain);
  }
}

template <typename R>
static int steps_t(const Model& m, void* e, uint64_t seed, uint64_t step0, int64_t nsteps, const void* noise,
                   hipStream_t s, const StepBufs* bufs, const StepRecOut* rec, bool plain) {
  uint32_t* sync = bufs ? bufs->sync : m.stepsync_d;
  uint64_t* recs = bufs ? bufs->rec : m.rec_d;
  if (!sync || !recs || !persistent_steps(m)) {
    set_error("persistent steps: the model does not take persistent launches");
    return MX_EINVAL;
  }
  const Static<R>* S = static_cast<const Static<R>*>(m.static_d);
  const Dyn<R> d = make_dyn<R>(m, seed);
  R* E = static_cast<R*>(e);
  StepSync sy;
  sy.cnt = sync;
  sy.ctl = sync + stepsync_ctl(m);
  sy.gu = reinterpret_cast<int32_t*>(sy.ctl + CTL_WORDS);
  sy.r0 = reach0(m);
  sy.PS = m.L[1] * m.L[2];
  sy.spin = m.wait_polls;
  static_assert(2 * STEPS_PER_LAUNCH + 2 < (1 << 12), "step tags overflow their 12 bits");
  const bool lean = lean_model(m) && !noise;
  const dim3 g(nblk_of(m));
  const size_t stride = size_t(m.N) * 3;
  for (int64_t n = 0; n < nsteps; n += STEPS_PER_LAUNCH) {
    const int k = int(std::min<int64_t>(STEPS_PER_LAUNCH, nsteps - n));
    const R* nz = noise ? static_cast<const R*>(noise) + size_t(n) * stride : nullptr;
    const uint64_t st = step0 + uint64_t(n);
    // this launch's records start at its first step
    RecPtrs<R> ro{nullptr, nullptr, nullptr, nullptr, nullptr};
    if (rec) {
      const size_t off = size_t(n) * stride;
      ro = RecPtrs<R>{static_cast<R*>(rec->e) + off, static_cast<R*>(rec->ep) + off, static_cast<R*>(rec->w) + off,
                      static_cast<R*>(rec->b1) + off, static_cast<R*>(rec->b2) + off};
    }
    // 16-bit tags: the launch's epoch (1 .. 15; 0 would match cleared records) in the high 4 bits, the
    // phase within the launch (up to 2 x STEPS_PER_LAUNCH + 2) in the low 12.  A model's records are
    // written only by its own launches, and consecutive launches take different epochs.
    m.rec_epoch = m.rec_epoch % 15u + 1u;
    const uint32_t tag0 = m.rec_epoch << 12;
    hipError_t err;
    if (m.geo == GEO_PERIODIC)
      err = lean ? steps_dispatch<R, GEO_PERIODIC, true>(m, g, s, S, d, E, recs, st, k, nz, tag0, sy, ro, plain)
                 : steps_dispatch<R, GEO_PERIODIC, false>(m, g, s, S, d, E, recs, st, k, nz, tag0, sy, ro, plain);
    else
      err = lean ? steps_dispatch<R, GEO_OPEN, true>(m, g, s, S, d, E, recs, st, k, nz, tag0, sy, ro, plain)
                 : steps_dispatch<R, GEO_OPEN, false>(m, g, s, S, d, E, recs, st, k, nz, tag0, sy, ro, plain);
    if (err != hipSuccess) {
      set_error(std::string("persistent heun steps: ") + hipGetErrorString(err));
      return MX_EHIP;
    }
  }
  return check_launch("persistent heun steps");
}

int launch_heun_steps(const Model& m, void* e, uint64_t seed, uint64_t step0, int64_t nsteps, const void* noise,
                      hipStream_t s, const StepBufs* bufs, const StepRecOut* rec, bool plain) {
  return m.prec == MX_FP64 ? steps_t<double>(m, e, seed, step0, nsteps, noise, s, bufs, rec, plain)
                           : steps_t<float>(m, e, seed, step0, nsteps, noise, s, bufs, rec, plain);
}

}  // namespace mx
