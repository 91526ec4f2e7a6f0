// Site decoding, stencil neighbour gather and local field of the single-model kernels, shared by the
// step / field / observables kernels (mxllg_kernels.hip) and the adjoint kernels (mxllg_adjoint.hip).
#pragma once
#include "mxllg_device.h"

namespace mx {

// ------------------------------------------------------------------------------------------ helpers
struct Site {
  int s, c, c0, c1, c2, i, crm;  // sublattice, internal cell, coordinates, internal site, magex cell
};

// CHECKER layout: coordinates of internal cell st.c (parity half, then row-major with c2 / 2 fastest).
template <typename R>
__device__ __forceinline__ void checker_cell(const Static<R>* __restrict__ S, Site& st) {
  const int half = S->half, L1 = S->L1, L2h = S->L2h;
  const int p = st.c >= half ? 1 : 0, idx = st.c - p * half;
  int t, k;
  if (S->fdiv) {
    t = fdiv(idx, L2h, S->rL2h);
    st.c0 = fdiv(t, L1, S->rL1);
  } else {
    t = idx / L2h;
    st.c0 = t / L1;
  }
  k = idx - t * L2h;
  st.c1 = t - st.c0 * L1;
  st.c2 = 2 * k + ((p + st.c0 + st.c1) & 1);
}

// Row-major coordinates of cell st.c (c2 fastest).
__device__ __forceinline__ void rowmajor_cell(Site& st, int L1, int L2, int use_fdiv, float rL1, float rL2) {
  if (use_fdiv) {
    const int t = fdiv(st.c, L2, rL2);
    st.c2 = st.c - t * L2;
    st.c0 = fdiv(t, L1, rL1);
    st.c1 = t - st.c0 * L1;
  } else {
    st.c2 = st.c % L2;
    const int t = st.c / L2;
    st.c1 = t % L1;
    st.c0 = t / L1;
  }
}

// Two block -> tile maps.  Slabs (xcd_tile): XCD x walks a contiguous run of row-major tiles, i.e. whole
// planes of axis 0.  Columns (col and S->xcol > 0, row-major layouts whose axis-0 planes hold 8 k whole
// tiles): XCD x owns tiles [x k, (x + 1) k) of every plane and walks them plane by plane, so the two
// reads of a bond along axis 0 are k tiles apart in its walk instead of a whole plane (8 k tiles): on
// bcc 64^3 that is 0.6 MB instead of 4.5 MB of FP64 tensors, inside the XCD's 4 MB L2.  rev walks
// either map backwards.  The map only decides which workgroup handles which tile: results do not depend
// on it.
template <typename R, int GEO>
__device__ __forceinline__ bool site_of(const Static<R>* __restrict__ S, Site& st, bool rev = false,
                                        bool col = false) {
  // every header field is read before the first branch, so they arrive in one round of scalar loads
  const int ncell = S->ncell, L1 = S->L1, L2 = S->L2, ns = S->ns, xcol = S->xcol, use_fdiv = S->fdiv;
  const float rL1 = S->rL1, rL2 = S->rL2;
  // sublattices interleave tile by tile, so one XCD works on every sublattice of its slab of cells and
  // the partner reads of per-bond tensors stored by another sublattice hit the same L2
  int tile;
  if (GEO != GEO_TILED && col && xcol > 0) {
    // every XCD gets exactly G / 8 workgroups: planes x 8 k tiles x ns
    const int x = blockIdx.x & 7, cnt = gridDim.x >> 3, q0 = blockIdx.x >> 3;
    const int q = rev ? cnt - 1 - q0 : q0;
    const int per = xcol * ns;  // workgroups of one XCD in one plane
    const int p = q / per, rem = q - p * per, rr = rem / ns;
    st.s = rem - rr * ns;
    tile = (p * 8 + x) * xcol + rr;
  } else {
    const int b = xcd_tile(blockIdx.x, gridDim.x, rev);
    tile = b / ns;
    st.s = b - tile * ns;
  }
  if constexpr (GEO == GEO_TILED) {
    // tile coordinates are wavefront-uniform (scalar); the lane's place in the tile is shifts and masks
    const int b0 = S->tb0, b1 = S->tb1, b2 = S->tb2;
    const int t2 = tile % S->nt2, tq = tile / S->nt2, t1 = tq % S->nt1, t0 = tq / S->nt1;
    const int l = threadIdx.x;
    st.c0 = (t0 << b0) + (l >> (b1 + b2));
    st.c1 = (t1 << b1) + ((l >> b2) & ((1 << b1) - 1));
    st.c2 = (t2 << b2) + (l & ((1 << b2) - 1));
    st.c = tile * BLOCK + l;
  } else if constexpr (GEO == GEO_CHECKER) {
    st.c = tile * BLOCK + threadIdx.x;
    if (st.c >= ncell) return false;
    checker_cell(S, st);
  } else {
    st.c = tile * BLOCK + threadIdx.x;
    if (st.c >= ncell) return false;
    rowmajor_cell(st, L1, L2, use_fdiv, rL1, rL2);
  }
  st.crm = (st.c0 * L1 + st.c1) * L2 + st.c2;
  st.i = st.s * ncell + st.c;
  return true;
}

// The one-launch step's block -> (sublattice, tile) map on row-major layouts: the per-XCD column walk
// where an axis-0 plane splits into 8 k whole tiles, else tiles in order (sublattices interleaved).
// Either way block b works on plane b / (blocks per plane), so blocks are dispatched plane after plane
// and the planes a block waits for are held by blocks dispatched at most about one plane after it.
// False for threads past the last cell (st.s and tile are still set).
template <typename R>
__device__ __forceinline__ bool step_site(const Static<R>* __restrict__ S, int b, Site& st, int& tile) {
  const int ncell = S->ncell, L1 = S->L1, L2 = S->L2, ns = S->ns, xcol = S->xcol, use_fdiv = S->fdiv;
  const float rL1 = S->rL1, rL2 = S->rL2;
  if (xcol > 0) {
    const int x = b & 7, q = b >> 3, per = xcol * ns;
    const int p = q / per, rem = q - p * per, rr = rem / ns;
    st.s = rem - rr * ns;
    tile = (p * 8 + x) * xcol + rr;
  } else {
    tile = b / ns;
    st.s = b - tile * ns;
  }
  st.c = tile * BLOCK + threadIdx.x;
  if (st.c >= ncell) return false;
  rowmajor_cell(st, L1, L2, use_fdiv, rL1, rL2);
  st.crm = (st.c0 * L1 + st.c1) * L2 + st.c2;
  st.i = st.s * ncell + st.c;
  return true;
}

// Site record of internal site i (sublattice-major order), for kernels that take a list of sites.
template <typename R, int GEO>
__device__ __forceinline__ void site_at(const Static<R>* __restrict__ S, int i, Site& st) {
  const int ncell = S->ncell, L1 = S->L1, L2 = S->L2;
  st.s = i / ncell;
  st.c = i - st.s * ncell;
  if constexpr (GEO == GEO_TILED) {
    const int b0 = S->tb0, b1 = S->tb1, b2 = S->tb2;
    const int tile = st.c / BLOCK, l = st.c - tile * BLOCK;
    const int t2 = tile % S->nt2, tq = tile / S->nt2, t1 = tq % S->nt1, t0 = tq / S->nt1;
    st.c0 = (t0 << b0) + (l >> (b1 + b2));
    st.c1 = (t1 << b1) + ((l >> b2) & ((1 << b1) - 1));
    st.c2 = (t2 << b2) + (l & ((1 << b2) - 1));
  } else if constexpr (GEO == GEO_CHECKER) {
    checker_cell(S, st);
  } else {
    st.c2 = st.c % L2;
    const int t = st.c / L2;
    st.c1 = t % L1;
    st.c0 = t / L1;
  }
  st.crm = (st.c0 * L1 + st.c1) * L2 + st.c2;
  st.i = i;
}

// Neighbour cell (internal index) of st along stencil entry en; offsets are reduced, so one conditional
// wrap per periodic axis.  GEO_OPEN: ok = false past an open face, and the returned cell is the site's
// own (a valid address; the caller zeroes its contribution).
template <typename R, int GEO>
__device__ __forceinline__ int neighbour_cell(const Static<R>* __restrict__ S, const Site& st, const Ent& en,
                                              bool& ok) {
  const int L0 = S->L0, L1 = S->L1, L2 = S->L2;
  int n0 = st.c0 + int(int16_t(en.d01 & 0xffff));
  int n1 = st.c1 + (en.d01 >> 16);
  int n2 = st.c2 + en.d2;
  if constexpr (GEO == GEO_OPEN) {
    n0 = (S->pbc0 && n0 >= L0) ? n0 - L0 : n0;
    n1 = (S->pbc1 && n1 >= L1) ? n1 - L1 : n1;
    n2 = (S->pbc2 && n2 >= L2) ? n2 - L2 : n2;
    ok = unsigned(n0) < unsigned(L0) && unsigned(n1) < unsigned(L1) && unsigned(n2) < unsigned(L2);
    return ok ? (n0 * L1 + n1) * L2 + n2 : st.c;
  } else {
    n0 = n0 >= L0 ? n0 - L0 : n0;
    n1 = n1 >= L1 ? n1 - L1 : n1;
    n2 = n2 >= L2 ? n2 - L2 : n2;
    ok = true;
    if constexpr (GEO == GEO_TILED) {
      const int b0 = S->tb0, b1 = S->tb1, b2 = S->tb2;
      const int tile = ((n0 >> b0) * S->nt1 + (n1 >> b1)) * S->nt2 + (n2 >> b2);
      const int l = ((n0 & ((1 << b0) - 1)) << (b1 + b2)) | ((n1 & ((1 << b1) - 1)) << b2) | (n2 & ((1 << b2) - 1));
      return tile * BLOCK + l;
    } else if constexpr (GEO == GEO_CHECKER) {
      return ((n0 + n1 + n2) & 1) * S->half + (n0 * L1 + n1) * S->L2h + (n2 >> 1);
    } else {
      return (n0 * L1 + n1) * L2 + n2;
    }
  }
}

// Spin loads of the neighbour gathers.  LD 0: ordinary cached loads.  LD 1: agent-scope coherent loads
// (global_load ... sc1), for spins that other workgroups of the same launch wrote and published with an
// agent-scope release: they never return a stale line of this XCD's L2, and need no L2 invalidation.
template <int LD, typename R>
__device__ __forceinline__ R spin_load(const R* p) {
  if constexpr (LD == 1) return __hip_atomic_load(const_cast<R*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else return *p;
}

// Tagged spin records of the persistent step launch: a spin's bits (96 in FP32, 192 in FP64) cut into
// 48-bit chunks, each stored in a 64-bit word under a 16-bit tag (the high bits), so that one
// single-copy-atomic 64-bit access carries a chunk and its tag.  FP32: 2 words; FP64: 4.
template <typename R>
constexpr int TAG_WORDS = sizeof(R) == 8 ? 4 : 2;
constexpr uint64_t CHUNK48 = (1ull << 48) - 1;

template <typename R>
__device__ __forceinline__ void tag3(R x, R y, R z, uint32_t tag, uint64_t (&w)[TAG_WORDS<R>]) {
  const uint64_t t = uint64_t(tag) << 48;
  if constexpr (sizeof(R) == 8) {
    const uint64_t b0 = __builtin_bit_cast(uint64_t, x), b1 = __builtin_bit_cast(uint64_t, y),
                   b2 = __builtin_bit_cast(uint64_t, z);
    w[0] = t | (b0 & CHUNK48);
    w[1] = t | (b0 >> 48) | ((b1 & 0xffffffffull) << 16);
    w[2] = t | (b1 >> 32) | ((b2 & 0xffffull) << 32);
    w[3] = t | (b2 >> 16);
  } else {
    const uint64_t u0 = __builtin_bit_cast(uint32_t, x), u1 = __builtin_bit_cast(uint32_t, y),
                   u2 = __builtin_bit_cast(uint32_t, z);
    w[0] = t | u0 | ((u1 & 0xffffull) << 32);
    w[1] = t | (u1 >> 16) | (u2 << 16);
  }
}

template <typename R>
__device__ __forceinline__ void untag3(const uint64_t (&w)[TAG_WORDS<R>], R& x, R& y, R& z) {
  if constexpr (sizeof(R) == 8) {
    x = __builtin_bit_cast(double, (w[0] & CHUNK48) | ((w[1] & 0xffffull) << 48));
    y = __builtin_bit_cast(double, ((w[1] >> 16) & 0xffffffffull) | ((w[2] & 0xffffffffull) << 32));
    z = __builtin_bit_cast(double, ((w[2] >> 32) & 0xffffull) | ((w[3] & CHUNK48) << 16));
  } else {
    x = __builtin_bit_cast(float, uint32_t(w[0]));
    y = __builtin_bit_cast(float, uint32_t(((w[0] >> 32) & 0xffffull) | ((w[1] & 0xffffull) << 16)));
    z = __builtin_bit_cast(float, uint32_t((w[1] >> 16) & 0xffffffffull));
  }
}

// Publish a spin as tagged record p (agent-scope 64-bit stores).
template <typename R>
__device__ __forceinline__ void publish_tagged(uint64_t* p, R x, R y, R z, uint32_t tag) {
  uint64_t w[TAG_WORDS<R>];
  tag3<R>(x, y, z, tag, w);
#pragma unroll
  for (int q = 0; q < TAG_WORDS<R>; ++q) __hip_atomic_store(p + q, w[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Entries per batch of the stencil loop: every load of a batch is issued before the first is used, so
// a wavefront keeps CH gathers (and, per bond, CH tensors) in flight instead of one.  Per-bond batches
// were timed on MI355X with the 14-entry bcc stencil: FP32 7 (two full batches; 5, 6, 8 and 9 were
// 1-15 % slower), FP64 2 (3 and 5 slower: more tensors in flight evict their partners from L2).  Table
// batches of 9 take the 18-entry triangular stencil (three shells) in two round trips instead of three:
// on 256^2 spins, 12.9 -> 11.2 us per Heun step FP32 and 14.4 -> 12.8 FP64, against 2 % more on the
// 6-entry simple cubic (three dead entries); one batch of 18 was slower on both (register pressure).
template <typename R, int CM>
struct Chunk {
  static constexpr int value = CM == COUP_BOND ? (sizeof(R) == 8 ? 2 : 7) : 9;
};

// One batch of CH stencil entries: gather() issues every load of the batch without a branch, accumulate()
// adds the batch to h in stencil order.  Splitting the two lets ALU work (the thermal noise) run while the
// first batch is in flight.  The CH records are read as one contiguous block (merged scalar loads, one
// wait); an entry past the end becomes a self-gather that counts zero.  Table couplings are fetched while
// the gathers are in flight, so accumulate() waits on nothing but the gathers.
// CHW > 0 overrides the batch width (at most ENT_PAD, the padding of the record array).
template <typename R, int CM, int GEO, int CHW = 0>
struct Batch {
  static constexpr int CH = CHW > 0 ? CHW : Chunk<R, CM>::value;
  static_assert(CH <= ENT_PAD, "a batch may read at most ENT_PAD records past its first");
  static constexpr int NJ = CM == COUP_BOND ? 9 : (CM == COUP_TABLE9 ? 9 : (CM == COUP_TABLE4 ? 4 : 1));
  R ex[CH], ey[CH], ez[CH], J[CH][NJ];

  // the CH records from k0 on (wavefront-uniform: scalar loads), for gather()
  static __device__ __forceinline__ void records(const Static<R>* __restrict__ S, int k0, Ent (&raw)[CH]) {
#pragma unroll
    for (int u = 0; u < CH; ++u) raw[u] = S->ent[k0 + u];
  }

  template <int LD = 0, bool WITH_J = true>
  __device__ __forceinline__ void gather(const Static<R>* __restrict__ S, const Dyn<R>& d, const R* __restrict__ e,
                                         const Site& st, int k0, int ke) {
    Ent raw[CH];
    records(S, k0, raw);
    gather<LD, WITH_J>(S, d, e, st, k0, ke, raw);
  }

  // WITH_J = false reloads only the spins (the couplings of these entries are already in J).
  template <int LD = 0, bool WITH_J = true>
  __device__ __forceinline__ void gather(const Static<R>* __restrict__ S, const Dyn<R>& d, const R* __restrict__ e,
                                         const Site& st, int k0, int ke, const Ent (&raw)[CH]) {
    const int ncell = S->ncell;
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const bool live = k0 + u < ke;
      Ent en;
      en.slot = live ? raw[u].slot : 0;
      en.tb = live ? raw[u].tb : st.s;
      en.d01 = live ? raw[u].d01 : 0;
      en.d2 = live ? raw[u].d2 : 0;
      bool ok;
      const int cn = neighbour_cell<R, GEO>(S, st, en, ok);
      ok = ok && live;
      const int j = (en.tb & 0xffff) * ncell + cn;
      ex[u] = spin_load<LD>(e + 3 * j);
      ey[u] = spin_load<LD>(e + 3 * j + 1);
      ez[u] = spin_load<LD>(e + 3 * j + 2);
      ex[u] = ok ? ex[u] : R(0);
      ey[u] = ok ? ey[u] : R(0);
      ez[u] = ok ? ez[u] : R(0);
      if constexpr (!WITH_J) {
      } else if constexpr (CM == COUP_BOND) {
        // per-bond tensor stored once per unordered bond: a forward entry reads its own cell's copy,
        // a backward entry its partner's copy at the neighbour cell, transposed (J_ji = J_ij^T)
        const bool back = (en.tb >> 16) != 0;
        const R* Jb = d.bond + (size_t(en.slot) * ncell + (back ? cn : st.c)) * 9;
        R t[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) t[q] = Jb[q];
#pragma unroll
        for (int q = 0; q < 9; ++q) J[u][q] = back ? t[(q % 3) * 3 + q / 3] : t[q];
      } else {
        const R* T = S->tab + NJ * en.slot;
#pragma unroll
        for (int q = 0; q < NJ; ++q) J[u][q] = T[q];
      }
    }
  }

  // As gather(), with the neighbours' spins taken from tagged records instead of a spin array: record j
  // holds TAG_WORDS<R> words (see tag3), read with agent-scope 64-bit loads (single-copy atomic).  Each
  // live entry is re-read until all its words carry the tag `want`; entries past the stencil's end and
  // neighbours outside an open boundary are not read (their spins count zero).  False when a lane gives
  // up: after `spin` unsuccessful polls, or once *abort is raised.
  // mid() runs once, after the first round of loads is issued and before it is used.
  template <bool WITH_J, typename Mid>
  __device__ __forceinline__ bool gather_tagged(const Static<R>* __restrict__ S, const Dyn<R>& d,
                                                const uint64_t* __restrict__ rec, const Site& st, int k0, int ke,
                                                uint32_t want, uint32_t spin, const uint32_t* abort, Mid&& mid) {
    static_assert(CH <= 32, "pending entries are a 32-bit mask");
    Ent raw[CH];
    records(S, k0, raw);
    const int ncell = S->ncell;
    int jj[CH];
    uint32_t pend = 0;
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const bool live = k0 + u < ke;
      Ent en;
      en.slot = live ? raw[u].slot : 0;
      en.tb = live ? raw[u].tb : st.s;
      en.d01 = live ? raw[u].d01 : 0;
      en.d2 = live ? raw[u].d2 : 0;
      bool ok;
      const int cn = neighbour_cell<R, GEO>(S, st, en, ok);
      if (ok && live) pend |= 1u << u;
      jj[u] = (en.tb & 0xffff) * ncell + cn;
      ex[u] = ey[u] = ez[u] = R(0);
      if constexpr (!WITH_J) {
      } else if constexpr (CM == COUP_BOND) {
        const bool back = (en.tb >> 16) != 0;
        const R* Jb = d.bond + (size_t(en.slot) * ncell + (back ? cn : st.c)) * 9;
        R t[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) t[q] = Jb[q];
#pragma unroll
        for (int q = 0; q < 9; ++q) J[u][q] = back ? t[(q % 3) * 3 + q / 3] : t[q];
      } else {
        const R* T = S->tab + NJ * en.slot;
#pragma unroll
        for (int q = 0; q < NJ; ++q) J[u][q] = T[q];
      }
    }
    return poll_tagged(rec, jj, pend, k0, ke, want, spin, abort, mid);
  }

  // The record indices jj of the batch's entries from k0 and the mask of those live and inside an open
  // face, as gather_tagged computes them (held across a persistent launch's steps).
  static __device__ __forceinline__ void tagged_at(const Static<R>* __restrict__ S, const Site& st, int k0, int ke,
                                                   int (&jj)[CH], uint32_t& okm) {
    Ent raw[CH];
    records(S, k0, raw);
    const int ncell = S->ncell;
    okm = 0;
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const bool live = k0 + u < ke;
      Ent en;
      en.slot = live ? raw[u].slot : 0;
      en.tb = live ? raw[u].tb : st.s;
      en.d01 = live ? raw[u].d01 : 0;
      en.d2 = live ? raw[u].d2 : 0;
      bool ok;
      const int cn = neighbour_cell<R, GEO>(S, st, en, ok);
      if (ok && live) okm |= 1u << u;
      jj[u] = (en.tb & 0xffff) * ncell + cn;
    }
  }

  // As gather_tagged (table couplings), with the entries' record indices jj and live-and-inside mask okm
  // already known (tagged_at): no neighbour arithmetic.
  template <typename Mid>
  __device__ __forceinline__ bool gather_tagged_at(const Static<R>* __restrict__ S, const uint64_t* __restrict__ rec,
                                                   int k0, int ke, const int (&jj)[CH], uint32_t okm, uint32_t want,
                                                   uint32_t spin, const uint32_t* abort, Mid&& mid) {
    static_assert(CM != COUP_BOND, "per-bond tensors are read per step");
    Ent raw[CH];
    records(S, k0, raw);
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      ex[u] = ey[u] = ez[u] = R(0);
      const R* T = S->tab + NJ * (k0 + u < ke ? raw[u].slot : 0);
#pragma unroll
      for (int q = 0; q < NJ; ++q) J[u][q] = T[q];
    }
    return poll_tagged(rec, jj, okm, k0, ke, want, spin, abort, mid);
  }

  // The polls of the tagged gathers: every live entry's words re-read until all carry the tag `want`,
  // pend the entries still to take.
  template <typename Mid>
  __device__ __forceinline__ bool poll_tagged(const uint64_t* __restrict__ rec, const int (&jj)[CH], uint32_t pend,
                                              int k0, int ke, uint32_t want, uint32_t spin, const uint32_t* abort,
                                              Mid&& mid) {
    constexpr int NW = TAG_WORDS<R>;
    for (uint32_t it = 0;; ++it) {
      // every live entry's words in flight at once (a branch per entry would wait on each in turn)
      uint64_t w[CH][NW];
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        if (k0 + u < ke) {  // wavefront-uniform
          const uint64_t* r = rec + size_t(jj[u]) * NW;
#pragma unroll
          for (int q = 0; q < NW; ++q)
            w[u][q] = __hip_atomic_load(const_cast<uint64_t*>(r + q), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      if (it == 0) mid();
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        if (k0 + u < ke) {
          bool g = true;
#pragma unroll
          for (int q = 0; q < NW; ++q) g = g && uint32_t(w[u][q] >> 48) == want;
          R x, y, z;
          untag3<R>(w[u], x, y, z);
          const bool take = g && (pend >> u & 1u);
          ex[u] = take ? x : ex[u];
          ey[u] = take ? y : ey[u];
          ez[u] = take ? z : ez[u];
          pend &= g ? ~(1u << u) : ~0u;
        }
      }
      if (!pend) return true;
      if (it >= spin) return false;
      if ((it & 15) == 15 && __hip_atomic_load(const_cast<uint32_t*>(abort), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        return false;
      __builtin_amdgcn_s_sleep(2);
    }
  }

  __device__ __forceinline__ void accumulate(const Static<R>* __restrict__ S, R& hx, R& hy, R& hz) const {
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const R x = ex[u], y = ey[u], z = ez[u];
      const R* T = J[u];
      if constexpr (CM == COUP_TABLE1) {
        const R Jv = T[0];
        hx += Jv * x;
        hy += Jv * y;
        hz += Jv * z;
      } else if constexpr (CM == COUP_TABLE4) {
        // J I + antisymmetric part: (J^A e_j) = e_j x D
        hx += T[0] * x + (y * T[3] - z * T[2]);
        hy += T[0] * y + (z * T[1] - x * T[3]);
        hz += T[0] * z + (x * T[2] - y * T[1]);
      } else {
        hx += T[0] * x + T[1] * y + T[2] * z;
        hy += T[3] * x + T[4] * y + T[5] * z;
        hz += T[6] * x + T[7] * y + T[8] * z;
      }
    }
  }
};

// h_i = sum_j J_ij e_j over the stencil of site st (magex/fast.py:86-120 semantics), accumulated in
// stencil order; mid() runs after the first batch's loads are issued and before they are used.
// The batch width does not change the result: entries are summed in stencil order either way.
// Per-bond couplings at the default batch width: the records of the next batch are loaded before the
// current batch is waited on, so a later batch issues its gathers without a scalar round trip of its own
// (bcc 64^3, FP32: 58.2 -> 56.7 us per Heun step).  Table couplings keep the plain order: their coupling
// values are scalar loads too, and waiting for them then also waits for the next records (triangular
// 256^2: 11.2 -> 12.0 us).  An explicit width (the Monte Carlo sweep's single batch) keeps it too: that
// kernel compiled 7 % slower with the prefetching loop, which it never runs.
template <typename R, int CM, int GEO, int CHW = 0, int LD = 0, typename Mid>
__device__ __forceinline__ void exchange(const Static<R>* __restrict__ S, const Dyn<R>& d, const R* __restrict__ e,
                                         const Site& st, R& hx, R& hy, R& hz, Mid&& mid) {
  hx = hy = hz = R(0);
  const int kb = S->sptr[st.s], ke = S->sptr[st.s + 1];
  if (kb >= ke) {
    mid();
    return;
  }
  using B = Batch<R, CM, GEO, CHW>;
  B bt;
  bt.template gather<LD>(S, d, e, st, kb, ke);
  if constexpr (CM == COUP_BOND && CHW == 0) {
    Ent next[B::CH];
    if (kb + B::CH < ke) B::records(S, kb + B::CH, next);
    mid();
    bt.accumulate(S, hx, hy, hz);
    for (int k0 = kb + B::CH; k0 < ke; k0 += B::CH) {
      bt.template gather<LD>(S, d, e, st, k0, ke, next);
      if (k0 + B::CH < ke) B::records(S, k0 + B::CH, next);
      bt.accumulate(S, hx, hy, hz);
    }
  } else {
    mid();
    bt.accumulate(S, hx, hy, hz);
    for (int k0 = kb + B::CH; k0 < ke; k0 += B::CH) {
      bt.template gather<LD>(S, d, e, st, k0, ke);
      bt.accumulate(S, hx, hy, hz);
    }
  }
}

template <typename R, int CM, int GEO, int CHW = 0>
__device__ __forceinline__ void exchange(const Static<R>* __restrict__ S, const Dyn<R>& d, const R* __restrict__ e,
                                         const Site& st, R& hx, R& hy, R& hz) {
  exchange<R, CM, GEO, CHW>(S, d, e, st, hx, hy, hz, [] {});
}

// Per-site on-site tensor of one site (ONSITE_SITE6: xx, yy, zz, xy, xz, yz in v[0..5]; ONSITE_SITE9:
// row-major), read once; the one-launch step keeps it for both Heun passes.
template <typename R>
struct SiteA {
  R v[9];
  // Every element is written on every path (zero where the mode has none): stores that differ between
  // the branches would be merged behind a computed address, which keeps a holder of SiteA (and all it
  // is part of) in scratch memory instead of registers.
  __device__ __forceinline__ void load(const Static<R>* __restrict__ S, const Dyn<R>& d, const Site& st) {
    const int amode = S->amode;
    const int nk = amode == ONSITE_SITE6 ? 6 : (amode == ONSITE_SITE9 ? 9 : 0);
    const size_t N = size_t(S->N);
    const R* a = d.A + st.i;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      R t = R(0);
      if (k < nk) {
        if constexpr (sizeof(R) == 8) {
          // FP64: the per-site six-component tensor is read once per pass; streaming it past L2 leaves
          // room for the partner bond tensors (measured 4-6 % faster on bcc64; slower in FP32, where it
          // stays cached)
          if (amode == ONSITE_SITE6) t = __builtin_nontemporal_load(a + size_t(k) * N);
          else t = a[size_t(k) * N];
        } else {
          t = a[size_t(k) * N];
        }
      }
      v[k] = t;
    }
  }
};

// On-site gradient dE/de (magex/fast.py:39-64): 2 A e + cubic + higher uniaxial, with the site's tensor
// already in A (ignored for ONSITE_SUB, whose tensor is in the constant block).
// LEAN (here and below): compiled for the plain model -- no cubic or higher-uniaxial terms, uniform
// moments and temperatures per sublattice, generated noise, no spin-orbit torque fields.
template <typename R, bool LEAN = false>
__device__ __forceinline__ void onsite_grad(const Static<R>* __restrict__ S, const Dyn<R>& d, const Site& st,
                                            const SiteA<R>& A, R x, R y, R z, R& gx, R& gy, R& gz) {
  gx = gy = gz = R(0);
  const int amode = S->amode;
  if (amode == ONSITE_SUB) {
    if (S->has_A) {
      const R* a = S->Asub[st.s];
      gx = R(2) * (a[0] * x + a[1] * y + a[2] * z);
      gy = R(2) * (a[3] * x + a[4] * y + a[5] * z);
      gz = R(2) * (a[6] * x + a[7] * y + a[8] * z);
    }
  } else if (amode == ONSITE_SITE6) {
    const R axx = A.v[0], ayy = A.v[1], azz = A.v[2], axy = A.v[3], axz = A.v[4], ayz = A.v[5];
    gx = R(2) * (axx * x + axy * y + axz * z);
    gy = R(2) * (axy * x + ayy * y + ayz * z);
    gz = R(2) * (axz * x + ayz * y + azz * z);
  } else {
    const R* a = A.v;
    gx = R(2) * (a[0] * x + a[1] * y + a[2] * z);
    gy = R(2) * (a[3] * x + a[4] * y + a[5] * z);
    gz = R(2) * (a[6] * x + a[7] * y + a[8] * z);
  }
  if (LEAN) return;
  if (S->has_c) {
    const R* C = S->C;
    const R a1 = C[0] * x + C[1] * y + C[2] * z;
    const R a2 = C[3] * x + C[4] * y + C[5] * z;
    const R a3 = C[6] * x + C[7] * y + C[8] * z;
    const R q1 = a1 * a1, q2 = a2 * a2, q3 = a3 * a3, kc = d.Kc[st.i];
    const R d1 = R(2) * a1 * (q2 + q3) * kc, d2 = R(2) * a2 * (q1 + q3) * kc, d3 = R(2) * a3 * (q1 + q2) * kc;
    gx += d1 * C[0] + d2 * C[3] + d3 * C[6];
    gy += d1 * C[1] + d2 * C[4] + d3 * C[7];
    gz += d1 * C[2] + d2 * C[5] + d3 * C[8];
  }
  if (S->has_u) {
    const R* u = S->u;
    const R c = x * u[0] + y * u[1] + z * u[2];
    const R c2 = c * c;
    const R f = R(-4) * d.K2[st.i] * c2 * c - R(6) * d.K3[st.i] * c2 * c2 * c;
    gx += f * u[0];
    gy += f * u[1];
    gz += f * u[2];
  }
}

template <typename R, bool LEAN = false>
__device__ __forceinline__ void onsite_grad(const Static<R>* __restrict__ S, const Dyn<R>& d, const Site& st, R x, R y,
                                            R z, R& gx, R& gy, R& gz) {
  SiteA<R> A;
  A.load(S, d, st);
  onsite_grad<R, LEAN>(S, d, st, A, x, y, z, gx, gy, gz);
}

// On-site energy (magex/fast.py:66-84).
template <typename R>
__device__ __forceinline__ R onsite_energy(const Static<R>* __restrict__ S, const Dyn<R>& d, const Site& st, R x, R y,
                                          R z) {
  R E = R(0);
  const int amode = S->amode;
  if (amode == ONSITE_SUB) {
    if (S->has_A) {
      const R* a = S->Asub[st.s];
      E = x * (a[0] * x + a[1] * y + a[2] * z) + y * (a[3] * x + a[4] * y + a[5] * z) +
          z * (a[6] * x + a[7] * y + a[8] * z);
    }
  } else if (amode == ONSITE_SITE6) {
    const int N = S->N, i = st.i;
    const R* A = d.A;
    const R axx = A[i], ayy = A[N + i], azz = A[2 * N + i];
    const R axy = A[3 * N + i], axz = A[4 * N + i], ayz = A[5 * N + i];
    E = axx * x * x + ayy * y * y + azz * z * z + R(2) * (axy * x * y + axz * x * z + ayz * y * z);
  } else {
    const int N = S->N;
    const R* a = d.A + st.i;
    E = x * (a[0] * x + a[N] * y + a[2 * N] * z) + y * (a[3 * N] * x + a[4 * N] * y + a[5 * N] * z) +
        z * (a[6 * N] * x + a[7 * N] * y + a[8 * N] * z);
  }
  if (S->has_c) {
    const R* C = S->C;
    const R a1 = C[0] * x + C[1] * y + C[2] * z;
    const R a2 = C[3] * x + C[4] * y + C[5] * z;
    const R a3 = C[6] * x + C[7] * y + C[8] * z;
    const R q1 = a1 * a1, q2 = a2 * a2, q3 = a3 * a3;
    E += d.Kc[st.i] * (q1 * q2 + q2 * q3 + q3 * q1);
  }
  if (S->has_u) {
    const R c = x * S->u[0] + y * S->u[1] + z * S->u[2];
    const R c4 = c * c * c * c;
    E += -d.K2[st.i] * c4 - d.K3[st.i] * c4 * c * c;
  }
  return E;
}

template <typename R, bool LEAN = false>
__device__ __forceinline__ void thermal(const Static<R>* __restrict__ S, const Dyn<R>& d, const Site& st,
                                        uint64_t step, const R* __restrict__ noise, R& bx, R& by, R& bz) {
  const R sg = (!LEAN && d.sigma) ? d.sigma[st.i] : S->sigma_sub[st.s];
  const int iref = st.crm * S->ns + st.s;
  if (!LEAN && noise) {
    bx = sg * noise[3 * iref];
    by = sg * noise[3 * iref + 1];
    bz = sg * noise[3 * iref + 2];
  } else if (S->noise_dbl) {
    double n0, n1, n2;
    gauss3d(d.seed_lo, d.seed_hi, step, uint32_t(iref), n0, n1, n2);
    bx = sg * R(n0);
    by = sg * R(n1);
    bz = sg * R(n2);
  } else {
    float n0, n1, n2;
    gauss3(d.seed_lo, d.seed_hi, step, uint32_t(iref), n0, n1, n2);
    bx = sg * R(n0);
    by = sg * R(n1);
    bz = sg * R(n2);
  }
}

template <typename R, bool LEAN = false>
__device__ __forceinline__ R inv_mu_of(const Static<R>* __restrict__ S, const Dyn<R>& d, const Site& st) {
  return (!LEAN && d.inv_mu) ? d.inv_mu[st.i] : S->inv_mu_sub[st.s];
}

// B_i = (2 h_i - g_i) / (mu_i mu_B) + B (+ thermal) at spin (x, y, z) (magex/fast.py:309-317), from the
// exchange sum h and the on-site gradient g; Bu is the uniform field of this pass.  A caller-supplied
// field d.extra (magex order, N x 3) is added after the uniform field and before the thermal field, the
// order of magex's unfused path (fast.py:216-222, 250-269).
template <typename R, bool LEAN = false>
__device__ __forceinline__ void field_from(const Static<R>* __restrict__ S, const Dyn<R>& d, const Site& st, R x, R y,
                                           R z, R hx, R hy, R hz, R gx, R gy, R gz, R nx, R ny, R nz, R Bu0, R Bu1,
                                           R Bu2, R& bx, R& by, R& bz) {
  const R im = inv_mu_of<R, LEAN>(S, d, st);
  if (!LEAN && d.extra) {
    const R* x3 = d.extra + 3 * size_t(st.crm * S->ns + st.s);
    bx = (R(2) * hx - gx) * im + Bu0;
    by = (R(2) * hy - gy) * im + Bu1;
    bz = (R(2) * hz - gz) * im + Bu2;
    bx += x3[0];
    by += x3[1];
    bz += x3[2];
    bx += nx;
    by += ny;
    bz += nz;
  } else {
    bx = (R(2) * hx - gx) * im + Bu0 + nx;
    by = (R(2) * hy - gy) * im + Bu1 + ny;
    bz = (R(2) * hz - gz) * im + Bu2 + nz;
  }
  if (!LEAN && d.sot) {
    // spin-orbit (Slonczewski-like) torque as fields: B_DL (e x p) + B_FL p (magex/llg.py:126-128)
    bx += d.bdl * (y * d.sp2 - z * d.sp1) + d.bfl * d.sp0;
    by += d.bdl * (z * d.sp0 - x * d.sp2) + d.bfl * d.sp1;
    bz += d.bdl * (x * d.sp1 - y * d.sp0) + d.bfl * d.sp2;
  }
}

// The local field at spin (x, y, z) of site st, gathering the neighbours' spins from e.  The generator's
// normals are computed while the first neighbour batch is in flight.  The per-site on-site tensor is
// loaded before the first gather, so that its latency hides behind the gathers' instead of adding a
// round trip after them (the loader's branches on the on-site mode kept the compiler from hoisting it).
template <typename R, int CM, int GEO, bool LEAN = false>
__device__ __forceinline__ void local_field(const Static<R>* __restrict__ S, const Dyn<R>& d, const R* __restrict__ e,
                                            const Site& st, R x, R y, R z, uint64_t step,
                                            const R* __restrict__ noise, R Bu0, R Bu1, R Bu2, R& bx, R& by, R& bz) {
  R hx, hy, hz, gx, gy, gz, nx = R(0), ny = R(0), nz = R(0);
  SiteA<R> A;
  A.load(S, d, st);
  exchange<R, CM, GEO>(S, d, e, st, hx, hy, hz, [&] {
    if (d.noisy) thermal<R, LEAN>(S, d, st, step, noise, nx, ny, nz);
  });
  onsite_grad<R, LEAN>(S, d, st, A, x, y, z, gx, gy, gz);
  field_from<R, LEAN>(S, d, st, x, y, z, hx, hy, hz, gx, gy, gz, nx, ny, nz, Bu0, Bu1, Bu2, bx, by, bz);
}

// Heun / Depondt corrector of site st (magex/fast.py:370-390): the field at the predictor ep (uniform
// part Bu), then e = n(w + dt/2 f(ep)) with w = e + dt/2 k1 kept in k1 (Heun), or e = n(R(dt/2 (w1 + w2)) e)
// with the predictor's rotation vector w1 in k1 (Depondt); in place.  save_B (adjoint recordings, may
// be null) receives the field.
template <typename R, int CM, int GEO, bool LEAN = false>
__device__ __forceinline__ void corrector(const Static<R>* __restrict__ S, const Dyn<R>& d, R* __restrict__ e,
                                          const R* __restrict__ ep, const R* __restrict__ k1, const Site& st, R a,
                                          uint64_t step, const R* __restrict__ noise, R Bu0, R Bu1, R Bu2,
                                          R* __restrict__ save_B) {
  const int i = st.i;
  const R x = ep[3 * i], y = ep[3 * i + 1], z = ep[3 * i + 2];
  // w (or the rotation vector), loaded ahead of the field like the on-site tensor
  const R k1x = k1[3 * i], k1y = k1[3 * i + 1], k1z = k1[3 * i + 2];
  R bx, by, bz, fx, fy, fz;
  local_field<R, CM, GEO, LEAN>(S, d, ep, st, x, y, z, step, noise, Bu0, Bu1, Bu2, bx, by, bz);
  if (!LEAN && save_B) {
    save_B[3 * i] = bx;
    save_B[3 * i + 1] = by;
    save_B[3 * i + 2] = bz;
  }
  llg_rhs(x, y, z, bx, by, bz, a, d.prec, fx, fy, fz);
  const R hdt = R(0.5) * d.dt;
  R cx, cy, cz;
  if (!LEAN && d.integ == MX_DEPONDT) {
    R wx, wy, wz;
    omega(x, y, z, fx, fy, fz, wx, wy, wz);
    cx = e[3 * i];
    cy = e[3 * i + 1];
    cz = e[3 * i + 2];
    rotate(cx, cy, cz, hdt * (k1x + wx), hdt * (k1y + wy), hdt * (k1z + wz));
  } else {
    cx = k1x + hdt * fx;
    cy = k1y + hdt * fy;
    cz = k1z + hdt * fz;
  }
  const R q = rnorm3(cx, cy, cz);
  e[3 * i] = cx * q;
  e[3 * i + 1] = cy * q;
  e[3 * i + 2] = cz * q;
}

}  // namespace mx
