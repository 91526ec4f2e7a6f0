// gfx950 kernels of the atomistic LLG hot path: stencil exchange gather, fused Heun predictor /
// corrector, field, observables, counter-based thermal noise, topological charge.
//
// Data layout (HBM): spins are packed x,y,z per site (12 B FP32 / 24 B FP64: one dwordx3 load, every
// fetched byte used), sites in sublattice-major order s * ncell + cell.  Cells are row-major (c2
// fastest) or, on fully periodic supercells whose sides allow it, grouped into 3D tiles of 256 cells
// with power-of-two sides (4 x 8 x 8 on a 3D lattice, 16 x 16 in 2D), tile-major: one workgroup owns
// one tile of one sublattice, so both ends of almost every bond are read by the same or adjacent
// workgroups within a few microseconds.  Neighbour indices are never stored: a site's neighbours are cell offsets from the stencil
// in the kernel arguments, wrapped on periodic axes.  Couplings are either a small type table (kernel
// constant block, scalar loads) or per-bond tensors stored once per unordered bond as contiguous
// row-major 3x3 blocks (bond[(f * ncell + cell) * 9 + q]: a wavefront reads 64 consecutive 36-B
// blocks with two dwordx4 loads and one dword load per lane).  Model constants sit in a device
// block (Static<R>) read with wavefront-uniform indices, so the kernel arguments stay a few words.
//
// Workgroups: 256 threads, one site each; a workgroup never mixes sublattices, so the stencil loop is
// wavefront-uniform.  The block -> tile map is XCD-aware: the dispatcher deals workgroups round-robin
// to the 8 XCDs, so tile = xcd * (G / 8) + (block / 8) gives every XCD one contiguous slab of the
// lattice and the neighbour re-reads of that slab hit its own L2.  The step kernels use per-XCD columns
// instead where axis-0 planes split into whole tiles (site_of), and the corrector pass walks
// them backwards, starting on the tiles the predictor pass left in L2.
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "mxllg_step.h"

namespace mx {

// ------------------------------------------------------------------------------------------ Heun
// PASS 1 (magex/fast.py:336-368): noise + field(e) + predictor ep = n(e + dt k1); keeps
//         w = e + dt/2 k1 (the corrector's part known after pass 1) in the buffer k1.
// PASS 2 (magex/fast.py:370-390): field(ep) + corrector e = n(w + dt/2 f(ep)), in place; e itself is
//         not read again.  The noise of a step is regenerated in pass 2 from the same counter.
// Depondt (d.integ == MX_DEPONDT, the same passes' depondt branches): pass 1 keeps the rotation vector
//         w1 = e x f(e) in k1 and rotates ep = R(dt w1) e; pass 2 rotates e = n(R(dt/2 (w1 + w2)) e) with
//         w2 = ep x f(ep).
//
// LEAN: the same passes compiled for the plain model (Heun, nothing pinned, damping, moments and
// temperatures uniform per sublattice, no cubic or higher-uniaxial terms, no spin-orbit torque, generated
// noise, no adjoint stores): the host picks it when the model is that, and its shorter code and scalar
// load chain matter on lattices too small to hide them.  Same arithmetic, so the same result bits.

// Pass 1 of site st: ep and w (or the rotation vector) into ep and k1.  PUB: ep stored with agent-scope
// stores (written through to the agent-coherent level), for kernels that publish it to other workgroups
// of the same launch.
template <typename R, int CM, int GEO, bool LEAN, bool PUB = false>
__device__ __forceinline__ void heun_predict(const Static<R>* __restrict__ S, const Dyn<R>& d, const R* __restrict__ e,
                                             R* __restrict__ ep, R* __restrict__ k1, const Site& st, uint64_t step,
                                             const R* __restrict__ noise) {
  const int i = st.i;
  const bool pin = !LEAN && d.pinned && d.pinned[i];
  const R a = (!LEAN && d.alpha) ? d.alpha[i] : S->alpha_sub[st.s];
  auto put_ep = [&](R x, R y, R z) {
    if constexpr (PUB) {
      store3_coherent(ep + 3 * i, x, y, z);
    } else {
      ep[3 * i] = x;
      ep[3 * i + 1] = y;
      ep[3 * i + 2] = z;
    }
  };
  const R x = e[3 * i], y = e[3 * i + 1], z = e[3 * i + 2];
  if (pin) {
    put_ep(x, y, z);
    k1[3 * i] = x;
    k1[3 * i + 1] = y;
    k1[3 * i + 2] = z;
    if (!LEAN && d.save_e) {  // a recorded step (adjoint): the spin, and no field acting on it
      d.save_e[3 * i] = x;
      d.save_e[3 * i + 1] = y;
      d.save_e[3 * i + 2] = z;
      d.save_B[3 * i] = d.save_B[3 * i + 1] = d.save_B[3 * i + 2] = R(0);
    }
    return;
  }
  R bx, by, bz, fx, fy, fz;
  local_field<R, CM, GEO, LEAN>(S, d, e, st, x, y, z, step, noise, d.B0, d.B1, d.B2, bx, by, bz);
  if (!LEAN && d.save_e) {
    d.save_e[3 * i] = x;
    d.save_e[3 * i + 1] = y;
    d.save_e[3 * i + 2] = z;
    d.save_B[3 * i] = bx;
    d.save_B[3 * i + 1] = by;
    d.save_B[3 * i + 2] = bz;
  }
  llg_rhs(x, y, z, bx, by, bz, a, d.prec, fx, fy, fz);
  if (!LEAN && d.integ == MX_DEPONDT) {
    R wx, wy, wz, px = x, py = y, pz = z;
    omega(x, y, z, fx, fy, fz, wx, wy, wz);
    k1[3 * i] = wx;
    k1[3 * i + 1] = wy;
    k1[3 * i + 2] = wz;
    rotate(px, py, pz, d.dt * wx, d.dt * wy, d.dt * wz);
    put_ep(px, py, pz);
    return;
  }
  const R hdt = R(0.5) * d.dt;
  k1[3 * i] = x + hdt * fx;
  k1[3 * i + 1] = y + hdt * fy;
  k1[3 * i + 2] = z + hdt * fz;
  const R px = x + d.dt * fx, py = y + d.dt * fy, pz = z + d.dt * fz;
  const R q = rnorm3(px, py, pz);
  put_ep(px * q, py * q, pz * q);
}

// Pass 2 of site st: the corrector from ep (neighbours and own) and w in k1, into e.
template <typename R, int CM, int GEO, bool LEAN>
__device__ __forceinline__ void heun_correct(const Static<R>* __restrict__ S, const Dyn<R>& d, R* __restrict__ e,
                                             const R* __restrict__ ep, const R* __restrict__ k1, const Site& st,
                                             uint64_t step, const R* __restrict__ noise) {
  const int i = st.i;
  if (!LEAN && d.pinned && d.pinned[i]) {
    if (!LEAN && d.save_B) d.save_B[3 * i] = d.save_B[3 * i + 1] = d.save_B[3 * i + 2] = R(0);
    return;
  }
  const R a = (!LEAN && d.alpha) ? d.alpha[i] : S->alpha_sub[st.s];
  corrector<R, CM, GEO, LEAN>(S, d, e, ep, k1, st, a, step, noise, d.B0, d.B1, d.B2, d.save_B);
}

template <typename R, int CM, int GEO, int PASS, bool LEAN = false>
__global__ __launch_bounds__(BLOCK) void k_heun(const Static<R>* __restrict__ S, const Dyn<R> d, R* __restrict__ e,
                                                R* __restrict__ ep, R* __restrict__ k1, uint64_t step,
                                                const R* __restrict__ noise) {
  Site st;
  if (!site_of<R, GEO>(S, st, PASS == 2 && d.rev2, true)) return;
  if (d.step_ptr) step += *d.step_ptr;  // graph replays: step = device counter + offset of this node
  if (PASS == 1) heun_predict<R, CM, GEO, LEAN>(S, d, e, ep, k1, st, step, noise);
  else heun_correct<R, CM, GEO, LEAN>(S, d, e, ep, k1, st, step, noise);
}

// One stochastic Heun (or Depondt) step in one launch (magex/fast.py:336-390 as one kernel): each
// workgroup runs the predictor of its tile, publishes the predictor spins, waits until the predictor
// spins of every axis-0 plane its stencil reaches are published, and runs the corrector of the same tile.
// Everything the corrector needs of its own sites stays in registers (Held): the coupling batches (for
// per-bond couplings the bond tensors, so each tensor comes from HBM once per step instead of once per
// pass), the on-site tensor, the thermal field and w.  The arithmetic is that of k_heun's two passes,
// so the results are the same bits.
//
// Ordering: a workgroup's predictor reads e of planes within r0 of its own and never waits; its
// corrector writes e only after every predictor that reads those planes has published (the wait covers
// the same planes).  Publication without cache maintenance: the predictor spins are stored with
// agent-scope stores (global_store ... sc1, written through to the agent-coherent level), the wavefront
// waits for their acknowledgement, and only then do the plane counters advance.  The corrector then
// gathers the neighbours' predictor spins with ordinary cached loads: no cache of this launch can hold a
// stale copy of them, because each predictor line is written once per launch (tiles are whole lines)
// and no workgroup loads a line of a plane before the plane is published, while copies from earlier
// launches are made coherent at the kernel boundary (the two-launch step relies on the same).
// Measured on bcc 64^3 FP32: an agent-scope release and acquire per wavefront (buffer_wbl2 /
// buffer_inv of the XCD's L2) made the step 7x slower than two launches, agent-scope (sc1) loads of the
// gathered spins, which bypass L2, 3.8x.  Two single-launch forms for grids larger than the GPU holds
// were built and measured on bcc 64^3 too, both bit-equal to two launches and both slower: G resident
// workgroups walking the tiles in rounds, each running one tile's predictor and the corrector of the tile
// it predicted a round before (81-109 us per step FP32 against 59; 169-176 FP64 against 142), and a plane
// pipeline per XCD in which workgroups take tickets from their XCD's counter and run the corrector a
// fixed lag of tickets behind the predictor (1, 2, 3 workgroups per CU: 102, 78, 80 us FP32; 173, 157,
// 172 FP64).  The pipeline read as many bytes as two launches (FETCH_SIZE 163 101 KiB per step against
// 162 232) with more L2 misses: a lag long enough that correctors do not wait on the predictors running
// beside them puts about 5 MB of tensors between a predictor and its corrector, more than an XCD's 4 MB
// L2, so the re-reads come from HBM anyway and the waits are pure cost.  Every wait is bounded: a
// workgroup that gives up stores w
// (behind a release fence) and leaves its corrector to the launch's last workgroup, which runs it once
// every predictor is done, then clears the counters for the next launch.
template <typename R, int CM, int GEO, int NB, bool LEAN>
__global__ __launch_bounds__(BLOCK) void k_step(const Static<R>* __restrict__ S, const Dyn<R> d, R* __restrict__ e,
                                                R* __restrict__ ep, R* __restrict__ k1, uint64_t step,
                                                const R* __restrict__ noise, const StepSync sy) {
  __shared__ int sh_flag;
  if (d.step_ptr) step += *d.step_ptr;  // graph replays: step = device counter + offset of this node
  Site st;
  int tile;
  const bool live = step_site<R>(S, blockIdx.x, st, tile);
  const int i = st.i;
  const bool pin = live && !LEAN && d.pinned && d.pinned[i];
  const int kb = S->sptr[st.s], ke = S->sptr[st.s + 1];
  Held<R, CM, GEO, NB> h;
  // ---- predictor (pass 1)
  if (live) {
    h.a = (!LEAN && d.alpha) ? d.alpha[i] : S->alpha_sub[st.s];
    h.x = e[3 * i];
    h.y = e[3 * i + 1];
    h.z = e[3 * i + 2];
    if (pin) {
      h.px = h.wx = h.x;
      h.py = h.wy = h.y;
      h.pz = h.wz = h.z;
    } else {
      predictor_held<R, CM, GEO, NB, LEAN, true>(S, d, SpinsFrom<R, 0>{e}, st, kb, ke, step, noise, h);
    }
    store3_coherent(ep + 3 * i, h.px, h.py, h.pz);
  }
  // ---- publish the predictor spins of this tile's planes: stores acknowledged, then the counters
  const int PS = sy.PS;
  const int c_first = tile * BLOCK, c_last = min(S->ncell, c_first + BLOCK) - 1;
  const int pa = c_first / PS, pb = c_last / PS;
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (int(threadIdx.x) <= pb - pa)
    __hip_atomic_fetch_add(sy.cnt + (pa + threadIdx.x) * CNT_STRIDE, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  // ---- wait for the planes within r0
  if (threadIdx.x < 64) {
    int lo, n;
    wait_window<R>(S, pa, pb, sy.r0, lo, n);
    const bool ok = wait_planes<R>(S, sy, sy.cnt, 1u, lo, n);
    if (threadIdx.x == 0) sh_flag = ok ? 1 : 0;
  }
  __syncthreads();
  const bool ok = sh_flag != 0;
  if (ok) {
    // ---- corrector (pass 2) from registers
    if (live && !pin) {
      corrector_held<R, CM, GEO, NB, LEAN>(S, d, SpinsFrom<R, 0>{ep}, st, kb, ke, h);
      e[3 * i] = h.x;
      e[3 * i + 1] = h.y;
      e[3 * i + 2] = h.z;
    }
  } else {
    // gave up: keep w (or the rotation vector) for the last workgroup's corrector
    if (live && !pin) {
      k1[3 * i] = h.wx;
      k1[3 * i + 1] = h.wy;
      k1[3 * i + 2] = h.wz;
    }
    give_up_tile(sy, blockIdx.x);
  }
  __syncthreads();
  finish_step<R, CM, GEO, LEAN>(S, d, e, ep, k1, step, noise, sy, !ok, sh_flag);
}

// One Heun step in one launch for grids larger than the GPU holds at once (magex/fast.py:336-390 as one
// kernel, like k_step).  The grid is the G workgroups the GPU holds at once.  They walk the tiles of the
// per-XCD column map in plane order: the tiles of each XCD are the tickets of its own counter (workgroup
// b counts on counter b mod 8, the XCD the dispatcher deals it to).  A workgroup runs the predictor of its
// ticket's tile, publishes the predictor spins as k_step does, waits for the planes within r0 of the
// tile, takes its next ticket, and runs the corrector from registers.  Every per-bond tensor then comes
// from HBM once per step, and the tensors in flight between a predictor and its corrector sit in the
// registers of the workgroups that hold them rather than in L2.  Same arithmetic as k_heun's two
// passes: same result bits.
//
// Progress: predictors never wait, and a workgroup takes its next ticket only after its wait has ended,
// so a wait needs only the tickets of the planes within r0 of its tile to be taken.  Let P be the tickets
// of one plane on one counter.  With every workgroup resident and more than (2 r0 + 1) P workgroups per
// counter, the counter whose frontier is furthest behind always has a workgroup whose tile's planes are
// all predicted: at most (r0 + 1) P of its held tiles lie within r0 of the frontier, and at most r0 P are
// first-plane tiles whose periodic stencil wraps onto the last planes (they hold their workgroup until
// those are predicted).  Waits are bounded all the same: a workgroup that gives up leaves its corrector
// to the launch's last workgroup, as in k_step.
template <typename R, int CM, int GEO, int NB, bool LEAN>
__global__ __launch_bounds__(BLOCK) void k_walk(const Static<R>* __restrict__ S, const Dyn<R> d, R* __restrict__ e,
                                                R* __restrict__ ep, R* __restrict__ k1, uint64_t step,
                                                const R* __restrict__ noise, const StepSync sy,
                                                uint32_t* __restrict__ tick) {
  __shared__ int sh_flag, sh_next;
  if (d.step_ptr) step += *d.step_ptr;  // graph replays: step = device counter + offset of this node
  const int x = blockIdx.x & 7, per_ctr = int(gridDim.x >> 3), ntick = S->nblk >> 3, PS = sy.PS;
  bool gave = false;
  for (int t = int(blockIdx.x >> 3); t < ntick;) {
    const int b = (t << 3) | x;  // the block of k_step's column walk that this ticket stands for
    Site st;
    int tile;
    const bool live = step_site<R>(S, b, st, tile);
    const int i = st.i;
    const bool pin = live && !LEAN && d.pinned && d.pinned[i];
    const int kb = S->sptr[st.s], ke = S->sptr[st.s + 1];
    Held<R, CM, GEO, NB> h;
    // ---- predictor (pass 1)
    if (live) {
      h.a = (!LEAN && d.alpha) ? d.alpha[i] : S->alpha_sub[st.s];
      h.x = e[3 * i];
      h.y = e[3 * i + 1];
      h.z = e[3 * i + 2];
      if (pin) {
        h.px = h.wx = h.x;
        h.py = h.wy = h.y;
        h.pz = h.wz = h.z;
      } else {
        predictor_held<R, CM, GEO, NB, LEAN, true>(S, d, SpinsFrom<R, 0>{e}, st, kb, ke, step, noise, h);
      }
      store3_coherent(ep + 3 * i, h.px, h.py, h.pz);
    }
    // ---- publish: stores acknowledged, then the plane counters
    const int c_first = tile * BLOCK, c_last = min(S->ncell, c_first + BLOCK) - 1;
    const int pa = c_first / PS, pb = c_last / PS;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (int(threadIdx.x) <= pb - pa)
      __hip_atomic_fetch_add(sy.cnt + (pa + threadIdx.x) * CNT_STRIDE, 1u, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
    // ---- wait for the planes within r0
    if (threadIdx.x < 64) {
      int lo, n;
      wait_window<R>(S, pa, pb, sy.r0, lo, n);
      const bool ok = wait_planes<R>(S, sy, sy.cnt, 1u, lo, n);
      if (threadIdx.x == 0) sh_flag = ok ? 1 : 0;
    }
    __syncthreads();
    const bool ok = sh_flag != 0;
    // ---- the next ticket, taken now that the wait is over; its latency overlaps the corrector
    uint32_t nt = 0;
    if (threadIdx.x == 0) nt = __hip_atomic_fetch_add(tick + x * CNT_STRIDE, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (ok) {
      // ---- corrector (pass 2) from registers
      if (live && !pin) {
        corrector_held<R, CM, GEO, NB, LEAN>(S, d, SpinsFrom<R, 0>{ep}, st, kb, ke, h);
        e[3 * i] = h.x;
        e[3 * i + 1] = h.y;
        e[3 * i + 2] = h.z;
      }
    } else {
      // gave up: keep w (or the rotation vector) for the last workgroup's corrector
      if (live && !pin) {
        k1[3 * i] = h.wx;
        k1[3 * i + 1] = h.wy;
        k1[3 * i + 2] = h.wz;
      }
      give_up_tile(sy, b);
      gave = true;
    }
    // the counter hands out tickets from G / 8 on (the first G / 8 are the workgroups' own)
    if (threadIdx.x == 0) sh_next = per_ctr + int(nt);
    __syncthreads();
    t = sh_next;
  }
  finish_step<R, CM, GEO, LEAN>(S, d, e, ep, k1, step, noise, sy, gave, sh_flag, tick);
}

// ------------------------------------------------------------------------------------------ field
// kind: MX_FIELD_BEFF (T), MX_FIELD_GRADIENT (meV), MX_FIELD_EXCHANGE (meV); output in magex order.
template <typename R, int CM, int GEO>
__global__ __launch_bounds__(BLOCK) void k_field(const Static<R>* __restrict__ S, const Dyn<R> d,
                                                 const R* __restrict__ e, R* __restrict__ out, int kind,
                                                 size_t istride) {
  // image blockIdx.y of a stack of spin configurations istride values apart (one image: y = 0)
  e += blockIdx.y * istride;
  out += blockIdx.y * istride;
  Site st;
  if (!site_of<R, GEO>(S, st)) return;
  const int i = st.i, o = 3 * (st.crm * S->ns + st.s);
  R hx, hy, hz;
  exchange<R, CM, GEO>(S, d, e, st, hx, hy, hz);
  if (kind == MX_FIELD_EXCHANGE) {
    out[o] = hx;
    out[o + 1] = hy;
    out[o + 2] = hz;
    return;
  }
  if (d.pinned && d.pinned[i]) {
    out[o] = out[o + 1] = out[o + 2] = R(0);
    return;
  }
  R gx, gy, gz;
  onsite_grad(S, d, st, e[3 * i], e[3 * i + 1], e[3 * i + 2], gx, gy, gz);
  const R im = inv_mu_of(S, d, st);
  if (kind == MX_FIELD_BEFF) {
    out[o] = (R(2) * hx - gx) * im + d.B0;
    out[o + 1] = (R(2) * hy - gy) * im + d.B1;
    out[o + 2] = (R(2) * hz - gz) * im + d.B2;
  } else {
    const R mm = R(1) / im;  // mu mu_B
    out[o] = gx - R(2) * hx - mm * d.B0;
    out[o + 1] = gy - R(2) * hy - mm * d.B1;
    out[o + 2] = gz - R(2) * hz - mm * d.B2;
  }
}

// ------------------------------------------------------------------------------------------ reductions
// Per-block partials of [E, sum e (3), sum s e (3), max |e x B_eff|] (spinmodel.py:475-484, 517-524;
// llg.py:250-252).
template <typename R, int CM, int GEO>
__global__ __launch_bounds__(BLOCK) void k_obs(const Static<R>* __restrict__ S, const Dyn<R> d,
                                               const R* __restrict__ e, double* __restrict__ part, size_t istride) {
  e += blockIdx.y * istride;
  part += size_t(blockIdx.y) * gridDim.x * MX_OBS_LEN;
  double v[MX_OBS_LEN] = {0, 0, 0, 0, 0, 0, 0, 0};
  Site st;
  if (site_of<R, GEO>(S, st)) {
    const int i = st.i;
    const R x = e[3 * i], y = e[3 * i + 1], z = e[3 * i + 2];
    R hx, hy, hz, gx, gy, gz;
    exchange<R, CM, GEO>(S, d, e, st, hx, hy, hz);
    onsite_grad(S, d, st, x, y, z, gx, gy, gz);
    const R im = inv_mu_of(S, d, st);
    const R Eon = onsite_energy(S, d, st, x, y, z);
    v[0] = double(Eon) - double(x * hx + y * hy + z * hz) - double(x * d.B0 + y * d.B1 + z * d.B2) / double(im);
    v[1] = x;
    v[2] = y;
    v[3] = z;
    const R sg = d.sign_site ? d.sign_site[i] : S->sign_sub[st.s];
    v[4] = sg * x;
    v[5] = sg * y;
    v[6] = sg * z;
    if (!(d.pinned && d.pinned[i])) {
      const R bx = (R(2) * hx - gx) * im + d.B0, by = (R(2) * hy - gy) * im + d.B1, bz = (R(2) * hz - gz) * im + d.B2;
      const double tx = y * bz - z * by, ty = z * bx - x * bz, tz = x * by - y * bx;
      v[7] = sqrt(tx * tx + ty * ty + tz * tz);
    }
  }
  block_reduce<MX_OBS_LEN>(v, 7, part + size_t(blockIdx.x) * MX_OBS_LEN);
}

// ------------------------------------------------------------------------------------------ misc
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_noise(int N, uint32_t seed_lo, uint32_t seed_hi, uint64_t step, int dbl,
                                                 R* __restrict__ out) {
  const int r = blockIdx.x * BLOCK + threadIdx.x;
  if (r >= N) return;
  if (dbl) {
    double n0, n1, n2;
    gauss3d(seed_lo, seed_hi, step, uint32_t(r), n0, n1, n2);
    out[3 * r] = R(n0);
    out[3 * r + 1] = R(n1);
    out[3 * r + 2] = R(n2);
  } else {
    float n0, n1, n2;
    gauss3(seed_lo, seed_hi, step, uint32_t(r), n0, n1, n2);
    out[3 * r] = R(n0);
    out[3 * r + 1] = R(n1);
    out[3 * r + 2] = R(n2);
  }
}

// Step counter of captured step loops: set before the first replay, advanced by each replay.
__global__ void k_counter(uint64_t* __restrict__ c, uint64_t value, int add) {
  if (threadIdx.x == 0 && blockIdx.x == 0) c[0] = add ? c[0] + value : value;
}

// magex order <-> internal order through perm[q] = magex site of internal site q
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_permute(const R* __restrict__ src, R* __restrict__ dst,
                                                   const int32_t* __restrict__ perm, int n, int to_internal) {
  const int q = blockIdx.x * BLOCK + threadIdx.x;
  if (q >= n) return;
  const int r = perm[q];
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


template <typename R>
static void fill_static(const Model& m, Static<R>& S) {
  std::memset(&S, 0, sizeof(S));
  S.L0 = m.L[0];
  S.L1 = m.L[1];
  S.L2 = m.L[2];
  S.ncell = m.ncell;
  S.ns = m.ns;
  S.N = m.N;
  S.bps = (m.ncell + BLOCK - 1) / BLOCK;
  S.nblk = S.bps * m.ns;
  S.pbc0 = m.pbc[0];
  S.pbc1 = m.pbc[1];
  S.pbc2 = m.pbc[2];
  S.allpbc = m.pbc[0] && m.pbc[1] && m.pbc[2];
  S.tb0 = m.tb[0];
  S.tb1 = m.tb[1];
  S.tb2 = m.tb[2];
  S.nt1 = m.L[1] >> m.tb[1];
  S.nt2 = m.L[2] >> m.tb[2];
  S.fdiv = m.ncell < (1 << 24);
  S.rL1 = 1.0f / float(m.L[1]);
  S.rL2 = 1.0f / float(m.L[2]);
  S.xcol = xcol_of(m);
  S.half = m.ncell / 2;
  S.L2h = m.L[2] / 2;
  S.rL2h = S.L2h > 0 ? 1.0f / float(S.L2h) : 0.0f;
  S.amode = m.amode;
  S.has_A = m.has_A;
  S.has_c = m.has_c;
  S.has_u = m.has_u;
  S.noise_dbl = m.noise_dbl;
  for (int s = 0; s <= MAXSUB; ++s) S.sptr[s] = s <= m.ns ? m.sptr[s] : m.sptr[m.ns];
  for (int k = 0; k < m.nent; ++k) S.ent[k] = m.ent[k];
  for (size_t k = 0; k < m.table.size() && k < size_t(MAXTAB); ++k) S.tab[k] = R(m.table[k]);
  for (int s = 0; s < m.ns; ++s) {
    for (int q = 0; q < 9; ++q) S.Asub[s][q] = R(m.Asub[9 * s + q]);
    S.inv_mu_sub[s] = R(1.0 / (m.mu_sub[s] * MU_B));
    S.sign_sub[s] = R(m.sign_sub[s]);
    S.alpha_sub[s] = m.alpha_sub_v.empty() ? R(0) : R(m.alpha_sub_v[s]);
    S.sigma_sub[s] = m.sigma_sub_v.empty() ? R(0) : R(m.sigma_sub_v[s]);
  }
  for (int q = 0; q < 9; ++q) S.C[q] = R(m.C[q]);
  for (int q = 0; q < 3; ++q) S.u[q] = R(m.u[q]);
}

template <typename R>
static int upload_static_t(Model& m) {
  Static<R> S;
  fill_static(m, S);
  if (!m.static_d) {
    const hipError_t err = hipMalloc(&m.static_d, sizeof(Static<R>));
    if (err != hipSuccess) {
      set_error(std::string("hipMalloc model constants: ") + hipGetErrorString(err));
      return MX_ENOMEM;
    }
  }
  const hipError_t err = hipMemcpy(m.static_d, &S, sizeof(S), hipMemcpyHostToDevice);
  if (err != hipSuccess) {
    set_error(std::string("upload model constants: ") + hipGetErrorString(err));
    return MX_EHIP;
  }
  m.static_dirty = 0;
  return MX_OK;
}

int upload_static(Model& m) { return m.prec == MX_FP64 ? upload_static_t<double>(m) : upload_static_t<float>(m); }

size_t stepsync_words(const Model& m) { return stepsync_tick(m) + 8 * CNT_STRIDE; }

// Per-bond couplings keep all their tensors in registers across the two passes: FP32 up to two batches
// of 7 (the 14-entry bcc stencil); FP64 tensors (twice the registers) are not held, and FP64 per-bond
// models keep two launches per step.
static int held_batches(const Model& m) {
  if (m.coup != COUP_BOND) return 0;
  if (m.prec != MX_FP32) return -1;
  int mx = 0;
  for (int s = 0; s < m.ns; ++s) mx = std::max(mx, m.sptr[s + 1] - m.sptr[s]);
  const int ch = Chunk<float, COUP_BOND>::value;
  const int nb = (mx + ch - 1) / ch;
  return nb <= 2 ? std::max(nb, 1) : -1;
}

// Workgroups a block of the one-launch step may wait on beyond its own plane, in dispatch order.
static long step_window(const Model& m) {
  const long PS = long(m.L[1]) * m.L[2];
  return long(m.ns) * ((2L * reach0(m) + 2) * PS / BLOCK + 3);
}

bool fused_step(const Model& m, int* nb_out) {
  if (m.step_mode == 0 || !(m.geo == GEO_PERIODIC || m.geo == GEO_OPEN)) return false;
  const int nb = held_batches(m);
  if (nb < 0) return false;
  int dev_cu = 0;
  if (hipDeviceGetAttribute(&dev_cu, hipDeviceAttributeMultiprocessorCount, m.device) != hipSuccess || dev_cu <= 0)
    return false;
  // every block this many ahead in dispatch order is resident with at least one resident block per CU,
  // so a wait is always for blocks that are already running (the kernels hold 1 to 4 blocks per CU)
  if (step_window(m) > dev_cu) return false;
  // one launch pays where the whole grid is resident at once (one block per CU or fewer): measured on
  // MI355X, sc 16^3 FP64 11.9 -> 8.8 us per step, triangular 256^2 FP32 15.1 -> 15.0.  On larger grids the
  // blocks run in lock-step generations whose waits leave memory idle (bcc 64^3 FP32, 2048 blocks at two
  // per CU: 57 -> 90 us), so they keep two launches per step.
  if (nblk_of(m) > dev_cu) return false;
  // FP64 beyond 16 k sites: the one-launch step lost on triangular 256^2 FP64 (15.2 -> 18.0 us against
  // two launches) while it won on sc 16^3 FP64 (12.5 -> 9.5) and triangular 256^2 FP32 (15.9 -> 15.1)
  if (m.prec == MX_FP64 && m.N > 16384) return false;
  if (nb_out) *nb_out = nb;
  return true;
}

// Workgroups per compute unit of the walked step (the lean and general kernels alike).
template <int GEO, int NB>
static int walk_occupancy() {
  static int occ = -1;
  if (occ < 0) {
    int a = 0, b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_walk<float, COUP_BOND, GEO, NB, true>, BLOCK, 0) !=
        hipSuccess)
      a = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_walk<float, COUP_BOND, GEO, NB, false>, BLOCK, 0) !=
        hipSuccess)
      b = 0;
    occ = std::min(a, b);
  }
  return occ;
}

bool walk_step(const Model& m, int* nb_out, int* grid_out) {
  if (m.step_mode != 3 || !(m.geo == GEO_PERIODIC || m.geo == GEO_OPEN)) return false;
  const int nb = held_batches(m);
  if (nb < 1 || xcol_of(m) == 0) return false;
  int dev_cu = 0;
  if (hipDeviceGetAttribute(&dev_cu, hipDeviceAttributeMultiprocessorCount, m.device) != hipSuccess || dev_cu <= 0)
    return false;
  const int nblk = nblk_of(m);
  if (nblk <= dev_cu) return false;  // the whole grid is resident: k_step
  int occ;
  if (m.geo == GEO_PERIODIC) occ = nb == 1 ? walk_occupancy<GEO_PERIODIC, 1>() : walk_occupancy<GEO_PERIODIC, 2>();
  else occ = nb == 1 ? walk_occupancy<GEO_OPEN, 1>() : walk_occupancy<GEO_OPEN, 2>();
  const int G = std::min(nblk, occ * dev_cu) & ~7;
  // the progress bound of k_walk: more than (2 r0 + 1) x (tickets per plane per counter) per counter
  if (long(G / 8) <= (2L * reach0(m) + 1) * long(xcol_of(m)) * m.ns) return false;
  if (nb_out) *nb_out = nb;
  if (grid_out) *grid_out = G;
  return true;
}

template <typename R, int GEO, int NB>
static void walk_launch(bool lean, dim3 g, hipStream_t s, const Static<R>* S, const Dyn<R>& d, R* E, R* ep, R* k1,
                        uint64_t step, const R* nz, const StepSync& sy, uint32_t* tick) {
  const dim3 b(BLOCK);
  if (lean) hipLaunchKernelGGL((k_walk<R, COUP_BOND, GEO, NB, true>), g, b, 0, s, S, d, E, ep, k1, step, nz, sy, tick);
  else hipLaunchKernelGGL((k_walk<R, COUP_BOND, GEO, NB, false>), g, b, 0, s, S, d, E, ep, k1, step, nz, sy, tick);
}

template <typename R, int CM, int GEO, int NB>
static void step_launch(const Model& m, bool lean, dim3 g, hipStream_t s, const Static<R>* S, const Dyn<R>& d, R* E,
                        R* ep, R* k1, uint64_t step, const R* nz, const StepSync& sy) {
  const dim3 b(BLOCK);
  if (lean) hipLaunchKernelGGL((k_step<R, CM, GEO, NB, true>), g, b, 0, s, S, d, E, ep, k1, step, nz, sy);
  else hipLaunchKernelGGL((k_step<R, CM, GEO, NB, false>), g, b, 0, s, S, d, E, ep, k1, step, nz, sy);
}

template <typename R, int GEO>
static void step_dispatch(const Model& m, int nb, bool lean, dim3 g, hipStream_t s, const Static<R>* S,
                          const Dyn<R>& d, R* E, R* ep, R* k1, uint64_t step, const R* nz, const StepSync& sy) {
  switch (m.coup) {
    case COUP_TABLE1: step_launch<R, COUP_TABLE1, GEO, 0>(m, lean, g, s, S, d, E, ep, k1, step, nz, sy); break;
    case COUP_TABLE4: step_launch<R, COUP_TABLE4, GEO, 0>(m, lean, g, s, S, d, E, ep, k1, step, nz, sy); break;
    case COUP_TABLE9: step_launch<R, COUP_TABLE9, GEO, 0>(m, lean, g, s, S, d, E, ep, k1, step, nz, sy); break;
    default:
      if constexpr (sizeof(R) == 4) {
        if (nb == 1) step_launch<R, COUP_BOND, GEO, 1>(m, lean, g, s, S, d, E, ep, k1, step, nz, sy);
        else step_launch<R, COUP_BOND, GEO, 2>(m, lean, g, s, S, d, E, ep, k1, step, nz, sy);
      }
      break;
  }
}


template <typename R>
static int heun_t(const Model& m, void* e, uint64_t seed, uint64_t step, const void* noise, hipStream_t s,
                  const uint64_t* step_ptr, const StepBufs* bufs) {
  const Static<R>* S = static_cast<const Static<R>*>(m.static_d);
  Dyn<R> d = make_dyn<R>(m, seed);
  d.step_ptr = step_ptr;
  R* E = static_cast<R*>(e);
  R* ep = static_cast<R*>(bufs ? bufs->ep : m.ep_d);
  R* k1 = static_cast<R*>(bufs ? bufs->k1 : m.k1_d);
  uint32_t* sync = bufs ? bufs->sync : m.stepsync_d;
  const R* nz = static_cast<const R*>(noise);
  const dim3 g(nblk_of(m)), b(BLOCK);
  int nb = 0;
  if (sync && fused_step(m, &nb)) {
    StepSync sy;
    sy.cnt = sync;
    sy.ctl = sync + stepsync_ctl(m);
    sy.gu = reinterpret_cast<int32_t*>(sy.ctl + CTL_WORDS);
    sy.r0 = reach0(m);
    sy.PS = m.L[1] * m.L[2];
    sy.spin = m.wait_polls;
    const bool lean = lean_model(m) && !noise;
    if (m.geo == GEO_PERIODIC) step_dispatch<R, GEO_PERIODIC>(m, nb, lean, g, s, S, d, E, ep, k1, step, nz, sy);
    else step_dispatch<R, GEO_OPEN>(m, nb, lean, g, s, S, d, E, ep, k1, step, nz, sy);
    return check_launch("heun step (one launch)");
  }
  int G = 0;
  if constexpr (sizeof(R) == 4) {
    if (sync && walk_step(m, &nb, &G)) {
      StepSync sy;
      sy.cnt = sync;
      sy.ctl = sync + stepsync_ctl(m);
      sy.gu = reinterpret_cast<int32_t*>(sy.ctl + CTL_WORDS);
      sy.r0 = reach0(m);
      sy.PS = m.L[1] * m.L[2];
      sy.spin = m.wait_polls;
      uint32_t* tick = sync + stepsync_tick(m);
      const bool lean = lean_model(m) && !noise;
      const dim3 gw(G);
      if (m.geo == GEO_PERIODIC) {
        if (nb == 1) walk_launch<R, GEO_PERIODIC, 1>(lean, gw, s, S, d, E, ep, k1, step, nz, sy, tick);
        else walk_launch<R, GEO_PERIODIC, 2>(lean, gw, s, S, d, E, ep, k1, step, nz, sy, tick);
      } else {
        if (nb == 1) walk_launch<R, GEO_OPEN, 1>(lean, gw, s, S, d, E, ep, k1, step, nz, sy, tick);
        else walk_launch<R, GEO_OPEN, 2>(lean, gw, s, S, d, E, ep, k1, step, nz, sy, tick);
      }
      return check_launch("heun step (walked one launch)");
    }
  }
  if (lean_model(m) && !noise) {
#define MX_HEUN_CALL(CM, GEO)                                                                   \
  hipLaunchKernelGGL((k_heun<R, CM, GEO, 1, true>), g, b, 0, s, S, d, E, ep, k1, step, nz); \
  hipLaunchKernelGGL((k_heun<R, CM, GEO, 2, true>), g, b, 0, s, S, d, E, ep, k1, step, nz)
    MX_DISPATCH(m, MX_HEUN_CALL);
#undef MX_HEUN_CALL
  } else {
#define MX_HEUN_CALL(CM, GEO)                                                             \
  hipLaunchKernelGGL((k_heun<R, CM, GEO, 1>), g, b, 0, s, S, d, E, ep, k1, step, nz); \
  hipLaunchKernelGGL((k_heun<R, CM, GEO, 2>), g, b, 0, s, S, d, E, ep, k1, step, nz)
    MX_DISPATCH(m, MX_HEUN_CALL);
#undef MX_HEUN_CALL
  }
  return check_launch("heun step");
}

int launch_heun_step(const Model& m, void* e, uint64_t seed, uint64_t step, const void* noise, hipStream_t s,
                     const uint64_t* step_ptr, const StepBufs* bufs) {
  return m.prec == MX_FP64 ? heun_t<double>(m, e, seed, step, noise, s, step_ptr, bufs)
                           : heun_t<float>(m, e, seed, step, noise, s, step_ptr, bufs);
}

__global__ __launch_bounds__(BLOCK) void k_fill_u32(uint32_t* __restrict__ p, size_t n) {
  for (size_t k = size_t(blockIdx.x) * BLOCK + threadIdx.x; k < n; k += size_t(gridDim.x) * BLOCK) p[k] = 0u;
}

int launch_fill_u32(uint32_t* p, size_t n, hipStream_t s) {
  if (n == 0) return MX_OK;
  const int g = int(std::min<size_t>((n + BLOCK - 1) / BLOCK, 1024));
  hipLaunchKernelGGL(k_fill_u32, dim3(g), dim3(BLOCK), 0, s, p, n);
  return check_launch("fill");
}

template <typename R>
static int pass_t(const Model& m, int pass, void* e, void* ep, void* k1, uint64_t seed, uint64_t step, void* save_e,
                  void* save_B, hipStream_t s, const double* B_ext, const void* extra) {
  const Static<R>* S = static_cast<const Static<R>*>(m.static_d);
  Dyn<R> d = make_dyn<R>(m, seed);
  if (B_ext) {
    d.B0 = R(B_ext[0]);
    d.B1 = R(B_ext[1]);
    d.B2 = R(B_ext[2]);
  }
  d.extra = static_cast<const R*>(extra);
  d.save_e = static_cast<R*>(save_e);
  d.save_B = static_cast<R*>(save_B);
  R* E = static_cast<R*>(e);
  R* EP = static_cast<R*>(ep);
  R* K1 = static_cast<R*>(k1);
  const R* nz = nullptr;
  const dim3 g(nblk_of(m)), b(BLOCK);
#define MX_PASS_CALL(CM, GEO)                                                                 \
  if (pass == 1) hipLaunchKernelGGL((k_heun<R, CM, GEO, 1>), g, b, 0, s, S, d, E, EP, K1, step, nz); \
  else hipLaunchKernelGGL((k_heun<R, CM, GEO, 2>), g, b, 0, s, S, d, E, EP, K1, step, nz)
  MX_DISPATCH(m, MX_PASS_CALL);
#undef MX_PASS_CALL
  return check_launch("heun pass");
}

int launch_heun_pass(const Model& m, int pass, void* e, void* ep, void* k1, uint64_t seed, uint64_t step,
                     void* save_e, void* save_B, hipStream_t s, const double* B_ext, const void* extra) {
  return m.prec == MX_FP64 ? pass_t<double>(m, pass, e, ep, k1, seed, step, save_e, save_B, s, B_ext, extra)
                           : pass_t<float>(m, pass, e, ep, k1, seed, step, save_e, save_B, s, B_ext, extra);
}

int launch_counter(uint64_t* c, uint64_t value, int add, hipStream_t s) {
  hipLaunchKernelGGL(k_counter, dim3(1), dim3(64), 0, s, c, value, add);
  return check_launch("step counter");
}

template <typename R>
static int field_t(const Model& m, const void* e, void* out, int kind, int nimg, hipStream_t s) {
  const Static<R>* S = static_cast<const Static<R>*>(m.static_d);
  const Dyn<R> d = make_dyn<R>(m, 0);
  const R* E = static_cast<const R*>(e);
  R* O = static_cast<R*>(out);
  const size_t istride = size_t(m.N) * 3;
  const dim3 g(nblk_of(m), nimg), b(BLOCK);
#define MX_FIELD_CALL(CM, GEO) hipLaunchKernelGGL((k_field<R, CM, GEO>), g, b, 0, s, S, d, E, O, kind, istride)
  MX_DISPATCH(m, MX_FIELD_CALL);
#undef MX_FIELD_CALL
  return check_launch("field");
}

int launch_field(const Model& m, const void* e, void* out, int kind, hipStream_t s, int nimg) {
  return m.prec == MX_FP64 ? field_t<double>(m, e, out, kind, nimg, s) : field_t<float>(m, e, out, kind, nimg, s);
}

template <typename R>
static int obs_t(const Model& m, const void* e, double* out, int nimg, double* part, hipStream_t s) {
  const Static<R>* S = static_cast<const Static<R>*>(m.static_d);
  const Dyn<R> d = make_dyn<R>(m, 0);
  const R* E = static_cast<const R*>(e);
  const int nb = nblk_of(m);
  const size_t istride = size_t(m.N) * 3;
  const dim3 g(nb, nimg), b(BLOCK);
#define MX_OBS_CALL(CM, GEO) hipLaunchKernelGGL((k_obs<R, CM, GEO>), g, b, 0, s, S, d, E, part, istride)
  MX_DISPATCH(m, MX_OBS_CALL);
#undef MX_OBS_CALL
  hipLaunchKernelGGL((k_final<MX_OBS_LEN>), dim3(nimg), dim3(BLOCK), 0, s, part, nb, 7, 1.0, out);
  return check_launch("observables");
}

int launch_observables(const Model& m, const void* e, double* out, hipStream_t s, int nimg, double* part) {
  if (!part) part = m.part_d;
  return m.prec == MX_FP64 ? obs_t<double>(m, e, out, nimg, part, s) : obs_t<float>(m, e, out, nimg, part, s);
}

int launch_noise(const Model& m, void* out, uint64_t seed, uint64_t step, hipStream_t s) {
  const dim3 g((m.N + BLOCK - 1) / BLOCK);
  if (m.prec == MX_FP64)
    hipLaunchKernelGGL((k_noise<double>), g, dim3(BLOCK), 0, s, m.N, uint32_t(seed), uint32_t(seed >> 32), step,
                       m.noise_dbl, static_cast<double*>(out));
  else
    hipLaunchKernelGGL((k_noise<float>), g, dim3(BLOCK), 0, s, m.N, uint32_t(seed), uint32_t(seed >> 32), step,
                       m.noise_dbl, static_cast<float*>(out));
  return check_launch("noise");
}

bool needs_permute(const Model& m) { return m.ns > 1 || m.geo == GEO_TILED || m.geo == GEO_CHECKER; }

int launch_permute(const Model& m, const void* src, void* dst, int to_internal, hipStream_t s) {
  const dim3 g((m.N + BLOCK - 1) / BLOCK);
  if (m.prec == MX_FP64)
    hipLaunchKernelGGL((k_permute<double>), g, dim3(BLOCK), 0, s, static_cast<const double*>(src),
                       static_cast<double*>(dst), m.perm_d, m.N, to_internal);
  else
    hipLaunchKernelGGL((k_permute<float>), g, dim3(BLOCK), 0, s, static_cast<const float*>(src),
                       static_cast<float*>(dst), m.perm_d, m.N, to_internal);
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

}  // namespace mx
