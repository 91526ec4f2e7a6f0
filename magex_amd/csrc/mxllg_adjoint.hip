// Differentiable form: reverse-mode adjoint of the stochastic Heun trajectory (and of the Depondt one,
// whose rotations are differentiated in rotate_vjp_v) with respect to the spins and to the parameters
// of a linearly parameterized Hamiltonian.  No counterpart in magex.
//
// Model: couplings J_k = sum_p theta_p U[p][k] on the stencil entries k, on-site A_s = sum_p theta_p
// Aun[p][s]; B_i = 2 (h_i - A_i e_i) / (mu_i mu_B) + B_ext + b_i with h_i = sum_k J_k e_j(k).
// One step:  k1 = f(e, B(e)),  ep = N(e + dt k1),  k2 = f(ep, B(ep)),  e' = N(e + dt/2 (k1 + k2)),
// f(e, B) = pre [prec e x B + alpha e x (e x B)], pre = -gamma / (1 + alpha^2), N(u) = u / |u|.
// Given lambda = dL/de', one adjoint step returns dL/de and adds dL/dtheta:
//   normalisation:  dL/du = (lambda - y (y . lambda)) / |u|,  y = N(u)
//   f:              with a = pre l and w = prec a + alpha (a x e):  dL/de = alpha (c x a) + B x w,
//                   dL/dB = w x e,  c = e x B
//   field:          nu_i = 2 dL/dB_i / (mu_i mu_B);  dL/de_j += sum_i J_ij^T nu_i - A_j^T nu_j, and since
//                   J_ji = J_ij^T the first term is the forward gather applied to nu
//   parameters:     dL/dtheta_p += sum_i nu_i . (sum_k U[p][k] x_j(k) - Aun[p] x_i)  (x = e or ep).
//   dense:          per unordered bond b = (i, j) stored once (J_ij = T_b, J_ji = T_b^T):
//                   dL/dT_b += nu_i x_j^T + x_i nu_j^T;  per site dL/dA_i += -nu_i x_i^T.
// The forward run keeps the spins every C steps; the backward replays each segment from its
// checkpoint (saving e, ep, w = e + dt/2 k1, B(e), B(ep) per step, noise regenerated from the same counters) and
// sweeps it: three kernels per step (one local, two gathers), or on grids the GPU holds at once one
// persistent launch per segment, with the next segment's replay running beside it on a second stream.
#include <cmath>
#include <cstring>
#include <algorithm>
#include <new>
#include <tuple>
#include <vector>

#include "mxllg_step.h"

namespace mx {

constexpr int MAXP = 16;  // parameters (a full 3x3 tensor and a few more)

template <typename R>
struct AStatic {
  int P;
  int fslot[MAXENT];       // forward-bond slot of each stencil entry (library order), -1 for reverse entries
  R U[MAXENT][MAXP][9];    // unit bond tensors per stencil entry (library order) and parameter
  R Aun[MAXSUB][MAXP][9];  // unit on-site tensors per sublattice and parameter
};

// Vector-Jacobian product of f(e, B) with cotangent l (see the header comment).
template <typename R>
__device__ __forceinline__ void rhs_vjp(R ex, R ey, R ez, R bx, R by, R bz, R alpha, int prec, R lx, R ly, R lz,
                                        R& gex, R& gey, R& gez, R& gbx, R& gby, R& gbz) {
  const R pre = R(-GAMMA) / (R(1) + alpha * alpha);
  const R ax = pre * lx, ay = pre * ly, az = pre * lz;
  const R cx = ey * bz - ez * by, cy = ez * bx - ex * bz, cz = ex * by - ey * bx;
  const R p = prec ? R(1) : R(0);
  const R wx = p * ax + alpha * (ay * ez - az * ey);
  const R wy = p * ay + alpha * (az * ex - ax * ez);
  const R wz = p * az + alpha * (ax * ey - ay * ex);
  gex = alpha * (cy * az - cz * ay) + (by * wz - bz * wy);
  gey = alpha * (cz * ax - cx * az) + (bz * wx - bx * wz);
  gez = alpha * (cx * ay - cy * ax) + (bx * wy - by * wx);
  gbx = wy * ez - wz * ey;
  gby = wz * ex - wx * ez;
  gbz = wx * ey - wy * ex;
}

// (v - y (y . v)) / n
template <typename R>
__device__ __forceinline__ void unproject(R yx, R yy, R yz, R n, R& vx, R& vy, R& vz) {
  const R d = yx * vx + yy * vy + yz * vz;
  vx = (vx - yx * d) / n;
  vy = (vy - yy * d) / n;
  vz = (vz - yz * d) / n;
}

// (dg/de)^T v / 2 for the site's on-site gradient g at the state (sx, sy, sz): A_i^T v for the tensor
// (per sublattice, or per site: symmetric 6 x N or full 9 x N), plus the cubic term's Hessian
// C^T M C v / 2 (M_kk = 2 Kc (q - q_k), M_kl = 4 Kc a_k a_l, a = C s, q_k = a_k^2) and the higher
// uniaxial term's f'(c) (u . v) u / 2 (c = u . s, f' = -12 K2 c^2 - 30 K3 c^4); both Hessians are symmetric.
// The spin-orbit field B_DL (e x p) adds -B_DL (p x v) mu mu_B / 2 (it enters B directly, not through g).
// Depondt steps rotate: y = Rot(v) x = x cos th + A (v x x) + Bc v (v . x), th = |v|, A = sin th / th,
// Bc = (1 - cos th) / th^2 (the identity below th = 1e-14, as the forward).  Rot(v)^T = Rot(-v); the
// cotangent of v for the cotangent g of y is
//   v [-(g.x) A + (g.(v x x)) A'/th + (g.v)(v.x) Bc'/th] + A (x x g) + Bc [(v.x) g + (g.v) x],
// with the ratios from their series below th = 1e-2 (FP64) / 0.1 (FP32), where the closed forms cancel.
template <typename R>
__device__ __forceinline__ void rotate_vjp_v(R vx, R vy, R vz, R x, R y, R z, R gx, R gy, R gz, R& ox, R& oy,
                                             R& oz) {
  const R th2 = vx * vx + vy * vy + vz * vz, th = sqrt(th2);
  if (!(th >= R(1e-14))) {
    ox = oy = oz = R(0);
    return;
  }
  R A, Bc, Ap, Bp;  // A, Bc and A'/th, Bc'/th
  if (th < R(sizeof(R) == 8 ? 1e-2 : 1e-1)) {
    const R t4 = th2 * th2;
    A = R(1) - th2 / R(6) + t4 / R(120);
    Bc = R(0.5) - th2 / R(24) + t4 / R(720);
    Ap = R(-1) / R(3) + th2 / R(30) - t4 / R(840);
    Bp = R(-1) / R(12) + th2 / R(180) - t4 / R(6720);
  } else {
    const R sn = sin(th), h = sin(R(0.5) * th), omc = R(2) * h * h, c = R(1) - omc;
    A = sn / th;
    Bc = omc / th2;
    Ap = (th * c - sn) / (th2 * th);
    Bp = (th * sn - R(2) * omc) / (th2 * th2);
  }
  const R gdx = gx * x + gy * y + gz * z, vdx = vx * x + vy * y + vz * z, gdv = gx * vx + gy * vy + gz * vz;
  const R cx = vy * z - vz * y, cy = vz * x - vx * z, cz = vx * y - vy * x;  // v x x
  const R gdc = gx * cx + gy * cy + gz * cz;
  const R k = -gdx * A + gdc * Ap + gdv * vdx * Bp;
  ox = k * vx + A * (y * gz - z * gy) + Bc * (vdx * gx + gdv * x);
  oy = k * vy + A * (z * gx - x * gz) + Bc * (vdx * gy + gdv * y);
  oz = k * vz + A * (x * gy - y * gx) + Bc * (vdx * gz + gdv * z);
}

// LEAN (here and in the kernels below): compiled for the plain model -- Heun, nothing pinned, no cubic,
// higher-uniaxial or spin-orbit terms (on-site tensors per sublattice or per site).
template <typename R, bool LEAN = false>
__device__ __forceinline__ void onsite_T(const Static<R>* __restrict__ S, const Dyn<R>& d, const Site& st, R sx,
                                         R sy, R sz, R x, R y, R z, R& ox, R& oy, R& oz) {
  const int amode = S->amode;
  if (amode == ONSITE_SUB) {
    const R* a = S->Asub[st.s];
    ox = a[0] * x + a[3] * y + a[6] * z;
    oy = a[1] * x + a[4] * y + a[7] * z;
    oz = a[2] * x + a[5] * y + a[8] * z;
  } else {
    const size_t N = size_t(S->N), i = size_t(st.i);
    const R* A = d.A;
    if (amode == ONSITE_SITE6) {
      const R axx = A[i], ayy = A[N + i], azz = A[2 * N + i], axy = A[3 * N + i], axz = A[4 * N + i],
              ayz = A[5 * N + i];
      ox = axx * x + axy * y + axz * z;
      oy = axy * x + ayy * y + ayz * z;
      oz = axz * x + ayz * y + azz * z;
    } else {
      R a[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) a[q] = A[q * N + i];
      ox = a[0] * x + a[3] * y + a[6] * z;
      oy = a[1] * x + a[4] * y + a[7] * z;
      oz = a[2] * x + a[5] * y + a[8] * z;
    }
  }
  if (LEAN) return;
  if (S->has_c) {
    const R* C = S->C;
    const R a1 = C[0] * sx + C[1] * sy + C[2] * sz, a2 = C[3] * sx + C[4] * sy + C[5] * sz,
            a3 = C[6] * sx + C[7] * sy + C[8] * sz;
    const R w1 = C[0] * x + C[1] * y + C[2] * z, w2 = C[3] * x + C[4] * y + C[5] * z, w3 = C[6] * x + C[7] * y + C[8] * z;
    const R kc = d.Kc[st.i], q1 = a1 * a1, q2 = a2 * a2, q3 = a3 * a3;
    // z = M w / 2
    const R z1 = kc * ((q2 + q3) * w1 + R(2) * a1 * (a2 * w2 + a3 * w3));
    const R z2 = kc * ((q1 + q3) * w2 + R(2) * a2 * (a1 * w1 + a3 * w3));
    const R z3 = kc * ((q1 + q2) * w3 + R(2) * a3 * (a1 * w1 + a2 * w2));
    ox += C[0] * z1 + C[3] * z2 + C[6] * z3;
    oy += C[1] * z1 + C[4] * z2 + C[7] * z3;
    oz += C[2] * z1 + C[5] * z2 + C[8] * z3;
  }
  if (S->has_u) {
    const R* u = S->u;
    const R c = u[0] * sx + u[1] * sy + u[2] * sz, c2 = c * c;
    const R h = (R(-6) * d.K2[st.i] * c2 - R(15) * d.K3[st.i] * c2 * c2) * (u[0] * x + u[1] * y + u[2] * z);
    ox += h * u[0];
    oy += h * u[1];
    oz += h * u[2];
  }
  if (d.sot && d.bdl != R(0)) {
    const R k = -d.bdl * R(0.5) / inv_mu_of(S, d, st);
    ox += k * (d.sp1 * z - d.sp2 * y);
    oy += k * (d.sp2 * x - d.sp0 * z);
    oz += k * (d.sp0 * y - d.sp1 * x);
  }
}

// Where the adjoint gathers take the neighbours' nu from: an array (NuArray), or tagged records polled
// until they carry the tag `want` (TaggedFrom, as the persistent forward's spins).
template <typename R>
struct NuArray {
  static constexpr bool tagged = false;
  const R* nu;
};

// Adjoint gather at site st: H = sum_k J_k nu_j (J from the model's width-9 table, or with CM ==
// COUP_BOND its per-bond tensors: a forward entry's own, a reverse entry's partner's transposed), and for every
// parameter t_p = sum_k U[k][p] x_j; then G_p = nu_i . (t_p - Aun[s][p] x_i) is accumulated into gp.
// NP (4, 8 or 16) is the parameter count rounded up; padded parameters have zero units.
// Entries go in batches of CH whose loads are all issued before the first is used (entries past the
// end re-read the last one with weight zero), so a wavefront keeps 2 CH gathers in flight.  From tagged
// records, each batch's live entries are re-read until every one carries its tag (x and the couplings
// are in flight meanwhile); false when a lane gives up (after src.spin polls, or once *src.abort is raised).
// PAR = false leaves the parameter terms out (x is then not read): the persistent sweep adds them later
// (k_par_seg), off its chain of exchanges.
template <typename R, int GEO, int NP, int CM, typename Src, bool PAR = true>
__device__ __forceinline__ bool adj_gather(const Static<R>* __restrict__ S, const AStatic<R>* __restrict__ A,
                                           const Dyn<R>& d, const Src& src, const R* __restrict__ x,
                                           const Site& st,
                                           R nix, R niy, R niz, R xix, R xiy, R xiz, R& hx, R& hy, R& hz,
                                           double (&gp)[NP]) {
  // FP32 per-bond batches of 7 take bcc's 14-entry stencil in two full batches: on bcc 64^3 the backward
  // went from 11.56 to 11.06 ms (one batch of 14: 12.39); the sums run in entry order either way
  constexpr int CH = (CM == COUP_BOND && sizeof(R) == 4) ? 7 : 8;
  const int ncell = S->ncell;
  bool good = true;
  R t[NP][3];
#pragma unroll
  for (int p = 0; p < NP; ++p) t[p][0] = t[p][1] = t[p][2] = R(0);
  hx = hy = hz = R(0);
  const int kb = S->sptr[st.s], ke = S->sptr[st.s + 1];
  for (int k0 = kb; k0 < ke; k0 += CH) {
    R vx[CH], vy[CH], vz[CH], qx[CH], qy[CH], qz[CH], wt[CH];
    R Jb[CM == COUP_BOND ? CH : 1][9];
    int kk[CH], jj[CH];
    uint32_t pend = 0;
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const bool live = k0 + u < ke;
      kk[u] = live ? k0 + u : ke - 1;
      const Ent en = S->ent[kk[u]];
      bool ok;
      const int cn = neighbour_cell<R, GEO>(S, st, en, ok);
      ok = ok && live;
      const int j = (en.tb & 0xffff) * ncell + cn;
      jj[u] = j;
      if constexpr (Src::tagged) {
        vx[u] = vy[u] = vz[u] = R(0);
        if (ok) pend |= 1u << u;
      } else {
        vx[u] = src.nu[3 * j];
        vy[u] = src.nu[3 * j + 1];
        vz[u] = src.nu[3 * j + 2];
      }
      if constexpr (PAR) {
        qx[u] = x[3 * j];
        qy[u] = x[3 * j + 1];
        qz[u] = x[3 * j + 2];
      }
      if constexpr (CM == COUP_BOND) {
        const bool back = (en.tb >> 16) != 0;
        const R* T = d.bond + (size_t(en.slot) * ncell + (back ? cn : st.c)) * 9;
        R tt[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) tt[q] = T[q];
#pragma unroll
        for (int q = 0; q < 9; ++q) Jb[u][q] = back ? tt[(q % 3) * 3 + q / 3] : tt[q];
      }
      wt[u] = ok ? R(1) : R(0);
    }
    if constexpr (Src::tagged) {
      constexpr int NW = TAG_WORDS<R>;
      for (uint32_t it = 0;; ++it) {
        uint64_t w[CH][NW];
#pragma unroll
        for (int u = 0; u < CH; ++u) {
          if (k0 + u < ke) {  // wavefront-uniform
            const uint64_t* r = src.rec + size_t(jj[u]) * NW;
#pragma unroll
            for (int q = 0; q < NW; ++q)
              w[u][q] = __hip_atomic_load(const_cast<uint64_t*>(r + q), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
        }
#pragma unroll
        for (int u = 0; u < CH; ++u) {
          if (k0 + u < ke) {
            bool g = true;
#pragma unroll
            for (int q = 0; q < NW; ++q) g = g && uint32_t(w[u][q] >> 48) == src.want;
            R a, b, c;
            untag3<R>(w[u], a, b, c);
            const bool take = g && (pend >> u & 1u);
            vx[u] = take ? a : vx[u];
            vy[u] = take ? b : vy[u];
            vz[u] = take ? c : vz[u];
            pend &= g ? ~(1u << u) : ~0u;
          }
        }
        if (!pend) break;
        if (it >= src.spin ||
            ((it & 15) == 15 &&
             __hip_atomic_load(const_cast<uint32_t*>(src.abort), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
          good = false;
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      vx[u] *= wt[u];
      vy[u] *= wt[u];
      vz[u] *= wt[u];
      if constexpr (PAR) {
        qx[u] *= wt[u];
        qy[u] *= wt[u];
        qz[u] *= wt[u];
      }
    }
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const R* T;
      if constexpr (CM == COUP_BOND) T = Jb[u];
      else T = S->tab + 9 * S->ent[kk[u]].slot;
      hx += T[0] * vx[u] + T[1] * vy[u] + T[2] * vz[u];
      hy += T[3] * vx[u] + T[4] * vy[u] + T[5] * vz[u];
      hz += T[6] * vx[u] + T[7] * vy[u] + T[8] * vz[u];
      if constexpr (PAR) {
#pragma unroll
        for (int p = 0; p < NP; ++p) {
          const R* U = A->U[kk[u]][p];
          t[p][0] += U[0] * qx[u] + U[1] * qy[u] + U[2] * qz[u];
          t[p][1] += U[3] * qx[u] + U[4] * qy[u] + U[5] * qz[u];
          t[p][2] += U[6] * qx[u] + U[7] * qy[u] + U[8] * qz[u];
        }
      }
    }
  }
  if constexpr (PAR) {
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const R* a = A->Aun[st.s][p];
      const R ux = t[p][0] - (a[0] * xix + a[1] * xiy + a[2] * xiz);
      const R uy = t[p][1] - (a[3] * xix + a[4] * xiy + a[5] * xiz);
      const R uz = t[p][2] - (a[6] * xix + a[7] * xiy + a[8] * xiz);
      gp[p] += double(nix * ux + niy * uy + niz * uz);
    }
  }
  return good;
}

// Neighbour indices of the persistent sweep's tagged gathers, held for a whole launch: FP32 table
// couplings whose sublattice stencil is one batch (as in the persistent forward, FP64 would pay more in
// registers than it saves).  j and the live-and-inside mask ok are computed as adj_gather computes them.
template <typename R, int CM>
struct AdjHeld {
  static constexpr bool HOLD = sizeof(R) == 4 && CM != COUP_BOND;
  int j[8];
  uint32_t ok;
  bool on;
};

template <typename R, int GEO, int CM>
__device__ __forceinline__ void adj_hold(const Static<R>* __restrict__ S, const Site& st, AdjHeld<R, CM>& h) {
  const int kb = S->sptr[st.s], ke = S->sptr[st.s + 1], ncell = S->ncell;
  h.on = AdjHeld<R, CM>::HOLD && ke > kb && ke - kb <= 8;  // wavefront-uniform
  h.ok = 0;
  if constexpr (!AdjHeld<R, CM>::HOLD) return;
  if (!h.on) return;
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const bool live = kb + u < ke;
    const Ent en = S->ent[live ? kb + u : ke - 1];
    bool ok;
    const int cn = neighbour_cell<R, GEO>(S, st, en, ok);
    if (ok && live) h.ok |= 1u << u;
    h.j[u] = (en.tb & 0xffff) * ncell + cn;
  }
}

// adj_gather's tagged path without parameter terms (PAR = false), with held indices: the same polls,
// weights, order and sums.
template <typename R, int CM>
__device__ __forceinline__ bool adj_gather_held(const Static<R>* __restrict__ S, const TaggedFrom& src,
                                                const Site& st, const AdjHeld<R, CM>& h, R& hx, R& hy, R& hz) {
  constexpr int CH = 8;
  constexpr int NW = TAG_WORDS<R>;
  const int kb = S->sptr[st.s], ke = S->sptr[st.s + 1];
  hx = hy = hz = R(0);
  R vx[CH], vy[CH], vz[CH];
  uint32_t pend = h.ok;
#pragma unroll
  for (int u = 0; u < CH; ++u) vx[u] = vy[u] = vz[u] = R(0);
  bool good = true;
  for (uint32_t it = 0;; ++it) {
    uint64_t w[CH][NW];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      if (kb + u < ke) {  // wavefront-uniform
        const uint64_t* r = src.rec + size_t(h.j[u]) * NW;
#pragma unroll
        for (int q = 0; q < NW; ++q)
          w[u][q] = __hip_atomic_load(const_cast<uint64_t*>(r + q), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      if (kb + u < ke) {
        bool g = true;
#pragma unroll
        for (int q = 0; q < NW; ++q) g = g && uint32_t(w[u][q] >> 48) == src.want;
        R a, b, c;
        untag3<R>(w[u], a, b, c);
        const bool take = g && (pend >> u & 1u);
        vx[u] = take ? a : vx[u];
        vy[u] = take ? b : vy[u];
        vz[u] = take ? c : vz[u];
        pend &= g ? ~(1u << u) : ~0u;
      }
    }
    if (!pend) break;
    if (it >= src.spin ||
        ((it & 15) == 15 &&
         __hip_atomic_load(const_cast<uint32_t*>(src.abort), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
      good = false;
      break;
    }
    __builtin_amdgcn_s_sleep(2);
  }
#pragma unroll
  for (int u = 0; u < CH; ++u) {
    const R wt = (h.ok >> u & 1u) ? R(1) : R(0);
    vx[u] *= wt;
    vy[u] *= wt;
    vz[u] *= wt;
    const R* T = S->tab + 9 * S->ent[kb + u < ke ? kb + u : ke - 1].slot;
    hx += T[0] * vx[u] + T[1] * vy[u] + T[2] * vz[u];
    hy += T[3] * vx[u] + T[4] * vy[u] + T[5] * vz[u];
    hz += T[6] * vx[u] + T[7] * vy[u] + T[8] * vz[u];
  }
  return good;
}

// The parameter terms alone, as adj_gather sums them (the same batches, dead-entry weights, order and
// expressions, hence the same bits): G_p = nu_i . (sum_k U[k][p] x_j - Aun[s][p] x_i) added to gp.
template <typename R, int GEO, int NP>
__device__ __forceinline__ void param_terms(const Static<R>* __restrict__ S, const AStatic<R>* __restrict__ A,
                                            const R* __restrict__ x, const Site& st, R nix, R niy, R niz, R xix,
                                            R xiy, R xiz, double (&gp)[NP]) {
  constexpr int CH = 8;
  const int ncell = S->ncell;
  R t[NP][3];
#pragma unroll
  for (int p = 0; p < NP; ++p) t[p][0] = t[p][1] = t[p][2] = R(0);
  const int kb = S->sptr[st.s], ke = S->sptr[st.s + 1];
  for (int k0 = kb; k0 < ke; k0 += CH) {
    R qx[CH], qy[CH], qz[CH], wt[CH];
    int kk[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const bool live = k0 + u < ke;
      kk[u] = live ? k0 + u : ke - 1;
      const Ent en = S->ent[kk[u]];
      bool ok;
      const int cn = neighbour_cell<R, GEO>(S, st, en, ok);
      ok = ok && live;
      const int j = (en.tb & 0xffff) * ncell + cn;
      qx[u] = x[3 * j];
      qy[u] = x[3 * j + 1];
      qz[u] = x[3 * j + 2];
      wt[u] = ok ? R(1) : R(0);
    }
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      qx[u] *= wt[u];
      qy[u] *= wt[u];
      qz[u] *= wt[u];
    }
#pragma unroll
    for (int u = 0; u < CH; ++u) {
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        const R* U = A->U[kk[u]][p];
        t[p][0] += U[0] * qx[u] + U[1] * qy[u] + U[2] * qz[u];
        t[p][1] += U[3] * qx[u] + U[4] * qy[u] + U[5] * qz[u];
        t[p][2] += U[6] * qx[u] + U[7] * qy[u] + U[8] * qz[u];
      }
    }
  }
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const R* a = A->Aun[st.s][p];
    const R ux = t[p][0] - (a[0] * xix + a[1] * xiy + a[2] * xiz);
    const R uy = t[p][1] - (a[3] * xix + a[4] * xiy + a[5] * xiz);
    const R uz = t[p][2] - (a[6] * xix + a[7] * xiy + a[8] * xiz);
    gp[p] += double(nix * ux + niy * uy + niz * uz);
  }
}

// Saved state of one step: e, ep, w = e + dt/2 k1, B(e), B(ep), each N x 3 in internal site order.
template <typename R>
struct StepRec {
  const R *e, *ep, *w, *B1, *B2;
};

// One step's record at one site (e, ep, w, B(e), B(ep)) and the site's constants (damping, the
// 2 / (mu mu_B) scale of nu, pinned), as the adjoint's local cores use them: loaded one step ahead of
// use in the persistent sweep.
template <typename R>
struct SiteRec {
  R e[3], ep[3], w[3], b1[3], b2[3];
};
template <typename R>
__device__ __forceinline__ void load_rec(const StepRec<R>& r, int i, SiteRec<R>& v) {
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    v.e[q] = r.e[3 * i + q];
    v.ep[q] = r.ep[3 * i + q];
    v.w[q] = r.w[3 * i + q];
    v.b1[q] = r.B1[3 * i + q];
    v.b2[q] = r.B2[3 * i + q];
  }
}
template <typename R>
struct SiteK {
  R a, im2;
  bool pin;
};
template <typename R, bool LEAN>
__device__ __forceinline__ SiteK<R> site_k(const Static<R>* __restrict__ S, const Dyn<R>& d, const Site& st) {
  SiteK<R> k;
  k.a = d.alpha ? d.alpha[st.i] : S->alpha_sub[st.s];
  k.im2 = R(2) * (d.inv_mu ? d.inv_mu[st.i] : S->inv_mu_sub[st.s]);
  k.pin = !LEAN && d.pinned && d.pinned[st.i];
  return k;
}

// Step-local part (corrector) at site i: from lambda = dL/de' and the step's record: lambda_u (lu), the
// cotangent of ep through k2 (ge), nu2 (n2) and, under Depondt, the corrector's share of dL/dw1 (gw).
// A pinned spin is carried unchanged (e' = ep = e), so lambda passes through and no field acts on it
// (ge = n2 = gw = 0).
template <typename R, bool LEAN = false>
__device__ __forceinline__ void adj_local_v(const Dyn<R>& d, const SiteK<R>& k, const SiteRec<R>& v, R lx, R ly,
                                            R lz, R (&lu)[3], R (&ge)[3], R (&n2)[3], R (&gw)[3]) {
  if (k.pin) {
    lu[0] = lx;
    lu[1] = ly;
    lu[2] = lz;
    ge[0] = ge[1] = ge[2] = R(0);
    n2[0] = n2[1] = n2[2] = R(0);
    gw[0] = gw[1] = gw[2] = R(0);
    return;
  }
  if (!LEAN && d.integ == MX_DEPONDT) {
    // e' = N(Rot(v) e), v = dt/2 (w1 + w2), w2 = ep x f(ep, B2); the record's k1 slot holds w1.
    // lu = Rot(v)^T dL/du (the direct part of dL/de), gw = dL/dw1 from this stage (= dL/dw2 =
    // dt/2 dL/dv), ge = dL/dep through w2, and n2.
    const R a = k.a;
    const R ex = v.e[0], ey = v.e[1], ez = v.e[2];
    const R px = v.ep[0], py = v.ep[1], pz = v.ep[2];
    const R bx = v.b2[0], by = v.b2[1], bz = v.b2[2];
    R fx, fy, fz, w2x, w2y, w2z;
    llg_rhs(px, py, pz, bx, by, bz, a, d.prec, fx, fy, fz);
    omega(px, py, pz, fx, fy, fz, w2x, w2y, w2z);
    const R hdt = R(0.5) * d.dt;
    const R vx = hdt * (v.w[0] + w2x), vy = hdt * (v.w[1] + w2y), vz = hdt * (v.w[2] + w2z);
    R ux = ex, uy = ey, uz = ez;
    rotate(ux, uy, uz, vx, vy, vz);
    const R n = sqrt(ux * ux + uy * uy + uz * uz);
    unproject(ux / n, uy / n, uz / n, n, lx, ly, lz);  // dL/du
    R tx = lx, ty = ly, tz = lz;
    rotate(tx, ty, tz, -vx, -vy, -vz);
    lu[0] = tx;
    lu[1] = ty;
    lu[2] = tz;
    R gvx, gvy, gvz;
    rotate_vjp_v(vx, vy, vz, ex, ey, ez, lx, ly, lz, gvx, gvy, gvz);
    const R gwx = hdt * gvx, gwy = hdt * gvy, gwz = hdt * gvz;
    gw[0] = gwx;
    gw[1] = gwy;
    gw[2] = gwz;
    // w2 = ep x f2: dL/dep += f2 x gw, dL/df2 = gw x ep
    const R lfx = gwy * pz - gwz * py, lfy = gwz * px - gwx * pz, lfz = gwx * py - gwy * px;
    R gx, gy, gz, hx, hy, hz;
    rhs_vjp(px, py, pz, bx, by, bz, a, d.prec, lfx, lfy, lfz, gx, gy, gz, hx, hy, hz);
    ge[0] = (fy * gwz - fz * gwy) + gx;
    ge[1] = (fz * gwx - fx * gwz) + gy;
    ge[2] = (fx * gwy - fy * gwx) + gz;
    const R im2 = k.im2;
    n2[0] = im2 * hx;
    n2[1] = im2 * hy;
    n2[2] = im2 * hz;
    return;
  }
  const R a = k.a;
  const R px = v.ep[0], py = v.ep[1], pz = v.ep[2];
  const R bx = v.b2[0], by = v.b2[1], bz = v.b2[2];
  R fx, fy, fz;
  llg_rhs(px, py, pz, bx, by, bz, a, d.prec, fx, fy, fz);
  const R hdt = R(0.5) * d.dt;
  // the record's k1 slot holds w = e + dt/2 k1 (the forward's pass-1 output), so u = w + dt/2 k2
  const R ux = v.w[0] + hdt * fx, uy = v.w[1] + hdt * fy, uz = v.w[2] + hdt * fz;
  const R n = sqrt(ux * ux + uy * uy + uz * uz);
  unproject(ux / n, uy / n, uz / n, n, lx, ly, lz);
  lu[0] = lx;
  lu[1] = ly;
  lu[2] = lz;
  gw[0] = gw[1] = gw[2] = R(0);
  R gx, gy, gz, hx, hy, hz;
  rhs_vjp(px, py, pz, bx, by, bz, a, d.prec, hdt * lx, hdt * ly, hdt * lz, gx, gy, gz, hx, hy, hz);
  ge[0] = gx;
  ge[1] = gy;
  ge[2] = gz;
  const R im2 = k.im2;
  n2[0] = im2 * hx;
  n2[1] = im2 * hy;
  n2[2] = im2 * hz;
}

// adj_local_v with its outputs stored at site i of lamu, gep, nu2 and (Depondt) gwb.
template <typename R, bool LEAN = false>
__device__ __forceinline__ void adj_local(const Static<R>* __restrict__ S, const Dyn<R>& d, const Site& st,
                                          const StepRec<R>& r, R lx, R ly, R lz, R* __restrict__ lamu,
                                          R* __restrict__ gep, R* __restrict__ nu2, R* __restrict__ gwb) {
  const int i = st.i;
  SiteRec<R> v;
  load_rec(r, i, v);
  R lu[3], ge[3], n2[3], gw[3];
  adj_local_v<R, LEAN>(d, site_k<R, LEAN>(S, d, st), v, lx, ly, lz, lu, ge, n2, gw);
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    lamu[3 * i + q] = lu[q];
    gep[3 * i + q] = ge[q];
    nu2[3 * i + q] = n2[q];
  }
  if (!LEAN && d.integ == MX_DEPONDT)
#pragma unroll
    for (int q = 0; q < 3; ++q) gwb[3 * i + q] = gw[q];
}

// Predictor part at site i, after the nu2 gather: from l = dL/dep, lambda_u (lu) and the corrector's
// share of dL/dw1 (gw, Depondt): nu1 (n1) and the partial dL/de (pt).
template <typename R, bool LEAN = false>
__device__ __forceinline__ void adj_pred_v(const Dyn<R>& d, const SiteK<R>& k, const SiteRec<R>& v, R lx, R ly,
                                           R lz, const R (&lu)[3], const R (&gw)[3], R (&n1)[3], R (&pt)[3]) {
  const R a = k.a;
  if (k.pin) {
    // ep = e' = e: dL/de gathers lambda (in lu) and dL/dep; the spin feels no field (nu1 = 0)
    n1[0] = n1[1] = n1[2] = R(0);
    pt[0] = lu[0] + lx;
    pt[1] = lu[1] + ly;
    pt[2] = lu[2] + lz;
  } else if (!LEAN && d.integ == MX_DEPONDT) {
    // ep = Rot(dt w1) e, w1 = e x f(e, B1): dL/de = lu + Rot^T dL/dep + f1 x gw1 + rhs part,
    // gw1 = dt dL/dv1 + the corrector's share (gw)
    const R ex = v.e[0], ey = v.e[1], ez = v.e[2];
    const R bx = v.b1[0], by = v.b1[1], bz = v.b1[2];
    R fx, fy, fz;
    llg_rhs(ex, ey, ez, bx, by, bz, a, d.prec, fx, fy, fz);
    const R vx = d.dt * v.w[0], vy = d.dt * v.w[1], vz = d.dt * v.w[2];
    R tx = lx, ty = ly, tz = lz;
    rotate(tx, ty, tz, -vx, -vy, -vz);
    R gvx, gvy, gvz;
    rotate_vjp_v(vx, vy, vz, ex, ey, ez, lx, ly, lz, gvx, gvy, gvz);
    const R gwx = d.dt * gvx + gw[0], gwy = d.dt * gvy + gw[1], gwz = d.dt * gvz + gw[2];
    const R lfx = gwy * ez - gwz * ey, lfy = gwz * ex - gwx * ez, lfz = gwx * ey - gwy * ex;  // gw1 x e
    R gx, gy, gz, qx, qy, qz;
    rhs_vjp(ex, ey, ez, bx, by, bz, a, d.prec, lfx, lfy, lfz, gx, gy, gz, qx, qy, qz);
    const R im2 = k.im2;
    n1[0] = im2 * qx;
    n1[1] = im2 * qy;
    n1[2] = im2 * qz;
    pt[0] = lu[0] + tx + (fy * gwz - fz * gwy) + gx;
    pt[1] = lu[1] + ty + (fz * gwx - fx * gwz) + gy;
    pt[2] = lu[2] + tz + (fx * gwy - fy * gwx) + gz;
  } else {
    const R ex = v.e[0], ey = v.e[1], ez = v.e[2];
    const R bx = v.b1[0], by = v.b1[1], bz = v.b1[2];
    R fx, fy, fz;
    llg_rhs(ex, ey, ez, bx, by, bz, a, d.prec, fx, fy, fz);
    const R vx = ex + d.dt * fx, vy = ey + d.dt * fy, vz = ez + d.dt * fz;
    const R n = sqrt(vx * vx + vy * vy + vz * vz);
    unproject(vx / n, vy / n, vz / n, n, lx, ly, lz);  // dL/dv, v = e + dt k1
    const R ux = lu[0], uy = lu[1], uz = lu[2];
    const R hdt = R(0.5) * d.dt;
    const R kx = d.dt * lx + hdt * ux, ky = d.dt * ly + hdt * uy, kz = d.dt * lz + hdt * uz;  // dL/dk1
    R gx, gy, gz, qx, qy, qz;
    rhs_vjp(ex, ey, ez, bx, by, bz, a, d.prec, kx, ky, kz, gx, gy, gz, qx, qy, qz);
    const R im2 = k.im2;
    n1[0] = im2 * qx;
    n1[1] = im2 * qy;
    n1[2] = im2 * qz;
    pt[0] = ux + lx + gx;
    pt[1] = uy + ly + gy;
    pt[2] = uz + lz + gz;
  }
}

template <typename R, int GEO, bool LEAN>
__global__ __launch_bounds__(BLOCK) void k_adj_a(const Static<R>* __restrict__ S, const Dyn<R> d,
                                                 const R* __restrict__ lam, const StepRec<R> r, R* __restrict__ lamu,
                                                 R* __restrict__ gep, R* __restrict__ nu2, R* __restrict__ gwb) {
  Site st;
  if (!site_of<R, GEO>(S, st)) return;
  const int i = st.i;
  adj_local<R, LEAN>(S, d, st, r, lam[3 * i], lam[3 * i + 1], lam[3 * i + 2], lamu, gep, nu2, gwb);
}

// Partial sums are kept per wavefront, not per workgroup: a wavefront's sum is a shuffle reduction that
// needs no barrier, so no wavefront waits for the others of its workgroup (which would hold every
// exchange of the persistent sweep to the workgroup's slowest wavefront).  Row of this wavefront:
constexpr int WAVES = BLOCK / 64;
__device__ __forceinline__ size_t wave_row() { return size_t(blockIdx.x) * WAVES + (threadIdx.x >> 6); }

// v[lane] for lanes below NV (a select chain: no dynamic register indexing), 0 elsewhere
template <int NV>
__device__ __forceinline__ double lane_pick(const double (&v)[NV]) {
  const int lane = threadIdx.x & 63;
  double r = 0.0;
#pragma unroll
  for (int k = 0; k < NV; ++k) r = lane == k ? v[k] : r;
  return r;
}

// Lane p's entry of the wavefront's row of gpart (stride MAXP), loaded at kernel start so the read is
// off the critical path.
template <int NP>
__device__ __forceinline__ double partial_row(const double* __restrict__ gpart) {
  const int lane = threadIdx.x & 63;
  return lane < NP ? gpart[wave_row() * MAXP + lane] : 0.0;
}

// Adds the wavefront's sums of gp to its running row entry run (lane p holds parameter p).
template <int NP>
__device__ __forceinline__ void add_partials(double (&gp)[NP], double& run) {
#pragma unroll
  for (int p = 0; p < NP; ++p) gp[p] = wave_sum(gp[p]);
  if ((threadIdx.x & 63) < NP) run = run + lane_pick<NP>(gp);
}

// The same, stored to the wavefront's row: old + the wavefront's sums.
template <int NP>
__device__ __forceinline__ void store_partials(double (&gp)[NP], double old, double* __restrict__ gpart) {
  add_partials<NP>(gp, old);
  const int lane = threadIdx.x & 63;
  if (lane < NP) gpart[wave_row() * MAXP + lane] = old;
}

// dL/dB of the uniform field in this pass: the wavefront's sum of nu mu mu_B / 2 (= dL/dB_i) into its row
// of gf (nblk x WAVES x 3), when the caller asks for it (gf is uniform over the grid).
__device__ __forceinline__ void store_field_partials(double (&fv)[3], double* __restrict__ gf) {
  if (!gf) return;
#pragma unroll
  for (int k = 0; k < 3; ++k) fv[k] = wave_sum(fv[k]);
  const int lane = threadIdx.x & 63;
  if (lane < 3) gf[wave_row() * 3 + lane] = lane_pick<3>(fv);
}

// Gather nu2 (dL/dep) and the predictor's adjoint: writes nu1 and the partial dL/de (part).
template <typename R, int GEO, int NP, int CM, bool LEAN>
__global__ __launch_bounds__(BLOCK) void k_adj_b(const Static<R>* __restrict__ S, const AStatic<R>* __restrict__ A,
                                                 const Dyn<R> d, const StepRec<R> r, const R* __restrict__ lamu,
                                                 const R* __restrict__ gep, const R* __restrict__ nu2,
                                                 R* __restrict__ nu1, R* __restrict__ part, double* __restrict__ gpart,
                                                 double* __restrict__ gf, const R* __restrict__ gwb) {
  const double old = partial_row<NP>(gpart);
  double gp[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) gp[p] = 0.0;
  double fv[3] = {0.0, 0.0, 0.0};
  Site st;
  if (site_of<R, GEO>(S, st)) {
    const int i = st.i;
    const R px = r.ep[3 * i], py = r.ep[3 * i + 1], pz = r.ep[3 * i + 2];
    const R vx2 = nu2[3 * i], vy2 = nu2[3 * i + 1], vz2 = nu2[3 * i + 2];
    if (gf) {
      const double h = 0.5 / double(d.inv_mu ? d.inv_mu[i] : S->inv_mu_sub[st.s]);
      fv[0] = h * double(vx2);
      fv[1] = h * double(vy2);
      fv[2] = h * double(vz2);
    }
    R hx, hy, hz, ox, oy, oz;
    adj_gather<R, GEO, NP, CM>(S, A, d, NuArray<R>{nu2}, r.ep, st, vx2, vy2, vz2, px, py, pz, hx, hy, hz, gp);
    onsite_T<R, LEAN>(S, d, st, px, py, pz, vx2, vy2, vz2, ox, oy, oz);
    // dL/dep; no early return below: every lane reaches the wavefront reduction
    const R lx = gep[3 * i] + hx - ox, ly = gep[3 * i + 1] + hy - oy, lz = gep[3 * i + 2] + hz - oz;
    const R lu[3] = {lamu[3 * i], lamu[3 * i + 1], lamu[3 * i + 2]};
    const SiteK<R> sk = site_k<R, LEAN>(S, d, st);
    R gw[3] = {R(0), R(0), R(0)};
    if (!LEAN && d.integ == MX_DEPONDT && !sk.pin)
      for (int q = 0; q < 3; ++q) gw[q] = gwb[3 * i + q];
    SiteRec<R> v;
    load_rec(r, i, v);
    R n1[3], pt[3];
    adj_pred_v<R, LEAN>(d, sk, v, lx, ly, lz, lu, gw, n1, pt);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      nu1[3 * i + q] = n1[q];
      part[3 * i + q] = pt[q];
    }
  }
  store_partials<NP>(gp, old, gpart);
  store_field_partials(fv, gf);
}

// Gather nu1: lambda = dL/de of the step's input, and the parameter terms at e.  With FUSE the
// previous step's local part (adj_local on its record prev) follows at once from lambda, which then
// never goes through memory; otherwise lambda is stored for a later k_adj_a.
template <typename R, int GEO, int NP, bool FUSE, int CM, bool LEAN>
__global__ __launch_bounds__(BLOCK) void k_adj_c(const Static<R>* __restrict__ S, const AStatic<R>* __restrict__ A,
                                                 const Dyn<R> d, const R* __restrict__ e, const R* __restrict__ nu1,
                                                 const R* __restrict__ part, R* __restrict__ lam,
                                                 const StepRec<R> prev, R* __restrict__ lamu, R* __restrict__ gep,
                                                 R* __restrict__ nu2, double* __restrict__ gpart,
                                                 double* __restrict__ gf, R* __restrict__ gwb) {
  const double old = partial_row<NP>(gpart);
  double gp[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) gp[p] = 0.0;
  double fv[3] = {0.0, 0.0, 0.0};
  Site st;
  if (site_of<R, GEO>(S, st)) {
    const int i = st.i;
    const R vx = nu1[3 * i], vy = nu1[3 * i + 1], vz = nu1[3 * i + 2];
    if (gf) {
      const double h = 0.5 / double(d.inv_mu ? d.inv_mu[i] : S->inv_mu_sub[st.s]);
      fv[0] = h * double(vx);
      fv[1] = h * double(vy);
      fv[2] = h * double(vz);
    }
    R hx, hy, hz, ox, oy, oz;
    adj_gather<R, GEO, NP, CM>(S, A, d, NuArray<R>{nu1}, e, st, vx, vy, vz, e[3 * i], e[3 * i + 1], e[3 * i + 2], hx, hy, hz,
                               gp);
    onsite_T<R, LEAN>(S, d, st, e[3 * i], e[3 * i + 1], e[3 * i + 2], vx, vy, vz, ox, oy, oz);
    const R lx = part[3 * i] + hx - ox, ly = part[3 * i + 1] + hy - oy, lz = part[3 * i + 2] + hz - oz;
    if constexpr (FUSE) {
      adj_local<R, LEAN>(S, d, st, prev, lx, ly, lz, lamu, gep, nu2, gwb);
    } else {
      lam[3 * i] = lx;
      lam[3 * i + 1] = ly;
      lam[3 * i + 2] = lz;
    }
  }
  store_partials<NP>(gp, old, gpart);
  store_field_partials(fv, gf);
}

// Steps [n_lo, n_hi) of a replayed segment's adjoint sweep (newest first) in one launch whose workgroups
// are all resident from the start (cooperative, or an ordinary launch of a grid that fits beside the
// replay running next to it), for models that take the persistent forward launch.  Each workgroup keeps
// its sites' adjoint state (lambda, lambda_u, gep, the corrector's share of dL/dw1, the partial dL/de)
// and, for one-batch FP32 table stencils, its neighbour indices (AdjHeld) in registers for the whole
// launch; sites exchange
// nu2 and nu1 through tagged records (recA, recB: 2 x N x TAG_WORDS words) as the persistent forward
// exchanges spins.  Per step, with k = tag0 + 2 (n_hi - 1 - n): the local part publishes nu2 (tag k),
// the nu2 gather publishes nu1 (k + 1), the nu1 gather gives lambda of the step's input.  A record is
// never overwritten before every reader has it: a site publishes nu2 of step n - 1 only after reading
// nu1 of step n of all its neighbours, each of which publishes that only after reading the site's nu2
// of step n (the stencil is symmetric), and likewise for nu1.  Every wait is for an earlier phase of the
// sweep and every workgroup is resident, so the launch always completes; polls are bounded as the
// forward's, and an aborted launch is counted in the model's count of given-up waits (gave_up).
// Same site map and arithmetic as k_adj_a, k_adj_b and k_adj_c: the same result bits.  The parameter
// terms are left out of the exchanges' chain: nu1 and nu2 of step n go to nuh[n] and nuh[C + n], and
// k_par_seg / k_par_fold add the terms afterwards as those kernels would have.  seg: the records, C steps
// apart (e, ep, w, B1, B2); gf (or null): field-gradient rows of the segment's first step.  lambda comes
// from lam and goes back to it.
template <typename R, int GEO, int NP, int CM, bool LEAN>
__global__ __launch_bounds__(BLOCK) void k_adj_seg(const Static<R>* __restrict__ S, const AStatic<R>* __restrict__ A,
                                                   const Dyn<R> d, const R* __restrict__ seg, int C, int n_hi,
                                                   int n_lo, R* __restrict__ lam, R* __restrict__ nuh,
                                                   double* __restrict__ gf,
                                                   uint64_t* __restrict__ rec, uint32_t tag0, const StepSync sy,
                                                   uint32_t* __restrict__ gave_up) {
  __shared__ int sh_flag;
  constexpr int NW = TAG_WORDS<R>;
  Site st;
  const bool live = site_of<R, GEO>(S, st);
  const int i = st.i;
  const size_t N3 = size_t(S->N) * 3;
  uint64_t* recA = rec;
  uint64_t* recB = rec + size_t(S->N) * NW;
  const uint32_t* abort = sy.ctl + 3;
  const size_t frows = size_t(gridDim.x) * WAVES * 3;
  auto rec_of = [&](int n) {
    return StepRec<R>{seg + size_t(n) * N3, seg + (size_t(C) + n) * N3, seg + (2 * size_t(C) + n) * N3,
                      seg + (3 * size_t(C) + n) * N3, seg + (4 * size_t(C) + n) * N3};
  };
  // the site's constants, lambda and the first step's record; each later record is loaded a step ahead
  R lx = R(0), ly = R(0), lz = R(0);
  SiteK<R> sk{R(0), R(0), false};
  SiteRec<R> v = {};
  double hmu = 0.0;
  AdjHeld<R, CM> hj;
  hj.on = false;
  if (live) {
    adj_hold<R, GEO, CM>(S, st, hj);
    sk = site_k<R, LEAN>(S, d, st);
    if (gf) hmu = 0.5 / double(d.inv_mu ? d.inv_mu[i] : S->inv_mu_sub[st.s]);
    lx = lam[3 * i];
    ly = lam[3 * i + 1];
    lz = lam[3 * i + 2];
    load_rec(rec_of(n_hi - 1), i, v);
  }
  bool ok = true;
  uint32_t k = tag0;
  for (int n = n_hi - 1; n >= n_lo; --n, k += 2) {
    const StepRec<R> r = rec_of(n);
    double gp[NP];  // unused: the parameter terms are added by k_par_seg
    double fv[3] = {0.0, 0.0, 0.0};
    R pt[3] = {R(0), R(0), R(0)}, n1[3] = {R(0), R(0), R(0)};
    // ---- local part, then the nu2 gather and the predictor's adjoint (k_adj_a / k_adj_c, k_adj_b)
    if (live) {
      R lu[3], ge[3], n2[3], gw[3];
      adj_local_v<R, LEAN>(d, sk, v, lx, ly, lz, lu, ge, n2, gw);
      publish_tagged<R>(recA + size_t(i) * NW, n2[0], n2[1], n2[2], k);
      if (nuh)
#pragma unroll
        for (int q = 0; q < 3; ++q) nuh[(size_t(C) + n) * N3 + 3 * i + q] = n2[q];
      if (gf) {
        fv[0] = hmu * double(n2[0]);
        fv[1] = hmu * double(n2[1]);
        fv[2] = hmu * double(n2[2]);
      }
      R hx, hy, hz, ox, oy, oz;
      const TaggedFrom fromA{recA, k, sy.spin, abort};
      const bool g = (AdjHeld<R, CM>::HOLD && hj.on)
                         ? adj_gather_held<R, CM>(S, fromA, st, hj, hx, hy, hz)
                         : adj_gather<R, GEO, NP, CM, TaggedFrom, false>(S, A, d, fromA, r.ep, st, n2[0], n2[1],
                                                                         n2[2], v.ep[0], v.ep[1], v.ep[2], hx, hy,
                                                                         hz, gp);
      ok = ok && g;
      onsite_T<R, LEAN>(S, d, st, v.ep[0], v.ep[1], v.ep[2], n2[0], n2[1], n2[2], ox, oy, oz);
      const R mx = ge[0] + hx - ox, my = ge[1] + hy - oy, mz = ge[2] + hz - oz;
      adj_pred_v<R, LEAN>(d, sk, v, mx, my, mz, lu, gw, n1, pt);
      publish_tagged<R>(recB + size_t(i) * NW, n1[0], n1[1], n1[2], k + 1);
      if (nuh)
#pragma unroll
        for (int q = 0; q < 3; ++q) nuh[size_t(n) * N3 + 3 * i + q] = n1[q];
    }
    store_field_partials(fv, gf ? gf + (size_t(n) * 2 + 1) * frows : nullptr);
    // ---- the nu1 gather: lambda of the step's input (k_adj_c); the next step's record is in flight
    fv[0] = fv[1] = fv[2] = 0.0;
    SiteRec<R> vn = v;
    if (live) {
      if (n > n_lo) load_rec(rec_of(n - 1), i, vn);
      if (gf) {
        fv[0] = hmu * double(n1[0]);
        fv[1] = hmu * double(n1[1]);
        fv[2] = hmu * double(n1[2]);
      }
      R hx, hy, hz, ox, oy, oz;
      const TaggedFrom fromB{recB, k + 1, sy.spin, abort};
      const bool g = (AdjHeld<R, CM>::HOLD && hj.on)
                         ? adj_gather_held<R, CM>(S, fromB, st, hj, hx, hy, hz)
                         : adj_gather<R, GEO, NP, CM, TaggedFrom, false>(S, A, d, fromB, r.e, st, n1[0], n1[1],
                                                                         n1[2], v.e[0], v.e[1], v.e[2], hx, hy, hz,
                                                                         gp);
      ok = ok && g;
      onsite_T<R, LEAN>(S, d, st, v.e[0], v.e[1], v.e[2], n1[0], n1[1], n1[2], ox, oy, oz);
      lx = pt[0] + hx - ox;
      ly = pt[1] + hy - oy;
      lz = pt[2] + hz - oz;
    }
    store_field_partials(fv, gf ? gf + size_t(n) * 2 * frows : nullptr);
    v = vn;
  }
  if (live) {
    lam[3 * i] = lx;
    lam[3 * i + 1] = ly;
    lam[3 * i + 2] = lz;
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
    __hip_atomic_fetch_add(gave_up, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(sy.ctl + 3, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(sy.ctl, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Parameter terms of every pass of a segment swept by k_adj_seg, one grid row per pass q in sweep order
// (q = 2 (n_seg - 1 - n): the nu2 pass of step n, at ep; q + 1: its nu1 pass, at e), from the nu1 / nu2
// the sweep saved (nuh[n], nuh[C + n]) and the records (seg, C steps apart).  Each wavefront's sums go to
// sums[q][row][MAXP]: what k_adj_b / k_adj_c add to their rows, computed alike (same site map, terms and
// wavefront sums).
template <typename R, int GEO, int NP>
__global__ __launch_bounds__(BLOCK) void k_par_seg(const Static<R>* __restrict__ S, const AStatic<R>* __restrict__ A,
                                                   const R* __restrict__ seg, const R* __restrict__ nuh, int C,
                                                   int n_seg, int q0, double* __restrict__ sums) {
  const int q = q0 + int(blockIdx.y), n = n_seg - 1 - (q >> 1);
  const size_t N3 = size_t(S->N) * 3, slot = ((q & 1) ? 0 : size_t(C)) + size_t(n);
  const R* x = seg + slot * N3;   // ep (record 1) for the nu2 pass, e (record 0) for the nu1 pass
  const R* nu = nuh + slot * N3;  // nu2 at C + n, nu1 at n
  double gp[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) gp[p] = 0.0;
  Site st;
  if (site_of<R, GEO>(S, st)) {
    const int i = st.i;
    param_terms<R, GEO, NP>(S, A, x, st, nu[3 * i], nu[3 * i + 1], nu[3 * i + 2], x[3 * i], x[3 * i + 1],
                            x[3 * i + 2], gp);
  }
#pragma unroll
  for (int p = 0; p < NP; ++p) gp[p] = wave_sum(gp[p]);
  const int lane = threadIdx.x & 63;
  if (lane < NP) sums[(size_t(q) * gridDim.x * WAVES + wave_row()) * MAXP + lane] = lane_pick<NP>(gp);
}

// Adds npass passes' sums to the rows of gpart in sweep order (thread = row x NP + parameter).
template <int NP>
__global__ void k_par_fold(const double* __restrict__ sums, int npass, int nrow, double* __restrict__ gpart) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nrow * NP) return;
  const int row = t / NP, p = t - row * NP;
  double v = gpart[size_t(row) * MAXP + p];
  // 16 passes' sums in flight at a time, added in pass order
  constexpr int U = 16;
  for (int q0 = 0; q0 < npass; q0 += U) {
    double x[U];
#pragma unroll
    for (int u = 0; u < U; ++u) x[u] = q0 + u < npass ? sums[(size_t(q0 + u) * nrow + row) * MAXP + p] : 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (q0 + u < npass) v = v + x[u];
  }
  gpart[size_t(row) * MAXP + p] = v;
}

template <typename T>
__global__ void k_zero(T* __restrict__ p, size_t n) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = T(0);
}

template <typename T>
__global__ void k_copy(const T* __restrict__ a, T* __restrict__ b, size_t n) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// Dense gradients of one replayed segment of n_seg steps, from its records and the adjoint vectors its
// sweep saved: nu1 (at e) and nu2 (at ep) of step n at nuh[(0 or 1) * C + n], the records' e and ep at
// seg[(0 or 1) * C + n] (N x 3 each, internal order).  Per unordered bond b = (i, j) stored once
// (J_ij = T_b, J_ji = T_b^T), owned by the forward entry of site i:
//   dL/dT_b += nu_i x_j^T + x_i nu_j^T,   per site  dL/dA_i += -nu_i x_i^T,
// pass by pass in the order the sweep runs them (newest step first, its corrector pass -- nu2 at ep --
// before its predictor pass -- nu1 at e), so the sums are those of adding every pass's terms to the
// accumulators as it runs, and every accumulator is read and written once per segment instead of once
// per pass.  One thread per site and group of FG forward entries (grid row blockIdx.y: group y of the
// site's sublattice), the group's accumulators in FP64 registers; group 0 also holds the site's on-site
// term.  Neighbours past an open face count zero.  Accumulators: gb[(f * 9 + q) * ncell + cell],
// ga[q * N + site].
template <typename R, int GEO, int FG>
__global__ __launch_bounds__(BLOCK) void k_dense_seg(const Static<R>* __restrict__ S,
                                                     const AStatic<R>* __restrict__ A, const R* __restrict__ seg,
                                                     const R* __restrict__ nuh, int C, int n_seg,
                                                     double* __restrict__ gb, double* __restrict__ ga) {
  Site st;
  if (!site_of<R, GEO>(S, st)) return;
  const int i = st.i, ncell = S->ncell;
  const size_t N3 = size_t(S->N) * 3, N = size_t(S->N);
  const int kb = S->sptr[st.s], ke = S->sptr[st.s + 1];
  const bool first = blockIdx.y == 0;
  int k = kb;
  // skip the forward entries of the earlier groups (wavefront-uniform)
  for (int skip = int(blockIdx.y) * FG; k < ke && (skip > 0 || A->fslot[k] < 0); ++k)
    if (A->fslot[k] >= 0) --skip;
  {
    // this group's FG forward entries of the sublattice (wavefront-uniform)
    int kk[FG], nf = 0;
#pragma unroll
    for (int u = 0; u < FG; ++u) {
      while (k < ke && A->fslot[k] < 0) ++k;
      kk[u] = k < ke ? k++ : -1;
      nf += kk[u] >= 0 ? 1 : 0;
    }
    if (nf == 0 && !first) return;
    int jj[FG];
    R w[FG];
    double acc[FG][9];
#pragma unroll
    for (int u = 0; u < FG; ++u) {
      jj[u] = i;
      w[u] = R(0);
      if (u < nf) {
        const Ent en = S->ent[kk[u]];
        bool ok;
        const int cn = neighbour_cell<R, GEO>(S, st, en, ok);
        jj[u] = (en.tb & 0xffff) * ncell + cn;
        w[u] = ok ? R(1) : R(0);
        const double* g = gb + size_t(A->fslot[kk[u]]) * 9 * ncell + st.c;
#pragma unroll
        for (int q = 0; q < 9; ++q) acc[u][q] = g[size_t(q) * ncell];
      }
    }
    double aa[9];
    if (first) {
#pragma unroll
      for (int q = 0; q < 9; ++q) aa[q] = ga[q * N + i];
    }
    // the passes in sweep order, q = 0 .. 2 n_seg - 1: step n_seg - 1 - q / 2, its nu2 pass (at ep) before
    // its nu1 pass (at e); each pass's values are loaded while the previous pass is added
    struct Ld {
      R v[3], x[3], vj[FG][3], xj[FG][3];
    };
    auto load = [&](int q, Ld& L) {
      const int n = n_seg - 1 - (q >> 1), pass = 1 - (q & 1);
      const R* x = seg + (size_t(pass) * C + n) * N3;
      const R* v = nuh + (size_t(pass) * C + n) * N3;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        L.v[c] = v[3 * i + c];
        L.x[c] = x[3 * i + c];
      }
#pragma unroll
      for (int u = 0; u < FG; ++u)
        if (u < nf)
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            L.vj[u][c] = v[3 * jj[u] + c];
            L.xj[u][c] = x[3 * jj[u] + c];
          }
    };
    const int npass = 2 * n_seg;
    Ld cur, nxt;
    load(0, cur);
    for (int q = 0; q < npass; ++q) {
      if (q + 1 < npass) load(q + 1, nxt);
      const double ni[3] = {double(cur.v[0]), double(cur.v[1]), double(cur.v[2])};
      const double xi[3] = {double(cur.x[0]), double(cur.x[1]), double(cur.x[2])};
#pragma unroll
      for (int u = 0; u < FG; ++u) {
        if (u < nf) {
          const double nj[3] = {double(cur.vj[u][0] * w[u]), double(cur.vj[u][1] * w[u]), double(cur.vj[u][2] * w[u])};
          const double xj[3] = {double(cur.xj[u][0] * w[u]), double(cur.xj[u][1] * w[u]), double(cur.xj[u][2] * w[u])};
#pragma unroll
          for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[u][3 * r + c] = acc[u][3 * r + c] + (ni[r] * xj[c] + xi[r] * nj[c]);
        }
      }
      if (first) {
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c) aa[3 * r + c] -= ni[r] * xi[c];
      }
      cur = nxt;
    }
#pragma unroll
    for (int u = 0; u < FG; ++u) {
      if (u < nf) {
        double* g = gb + size_t(A->fslot[kk[u]]) * 9 * ncell + st.c;
#pragma unroll
        for (int q = 0; q < 9; ++q) g[size_t(q) * ncell] = acc[u][q];
      }
    }
    if (first) {
#pragma unroll
      for (int q = 0; q < 9; ++q) ga[q * N + i] = aa[q];
    }
  }
}

// Dense gradients from the accumulators (component-major, internal cells) to the caller's layouts:
// bonds [f][magex cell][9], on-site [magex site][9]; cmap[c] is the internal cell of magex cell c
// (null: the identity).
__global__ void k_dense_bonds(const double* __restrict__ gb, const int* __restrict__ cmap, int nfwd, int ncell,
                              double* __restrict__ out) {
  const size_t n = size_t(nfwd) * ncell * 9;
  const size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const size_t fc = idx / 9;
  const int q = int(idx - fc * 9);
  const int f = int(fc / size_t(ncell)), c = int(fc - size_t(f) * ncell);
  out[idx] = gb[(size_t(f) * 9 + q) * ncell + (cmap ? cmap[c] : c)];
}

__global__ void k_dense_onsite(const double* __restrict__ ga, const int* __restrict__ cmap, int ns, int ncell,
                               double* __restrict__ out) {
  const size_t N = size_t(ns) * ncell;
  const size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= N * 9) return;
  const size_t r = idx / 9;
  const int q = int(idx - r * 9);
  const int c = int(r / size_t(ns)), s = int(r - size_t(c) * ns);
  out[idx] = ga[size_t(q) * N + size_t(s) * ncell + (cmap ? cmap[c] : c)];
}

}  // namespace mx

// ============================================================================================ host
using namespace mx;

struct mx_adjoint {
  mx_model* m = nullptr;
  int P = 0, C = 50;
  std::vector<double> U, Aun;  // P x nent x 9 (library order), P x ns x 9
  std::vector<double> theta;
  std::vector<int> fslot;  // forward-bond slot per library entry (-1: reverse entry)
  int nfwd = 0;
  std::vector<double> boff, aoff;  // fixed parts: per bond nfwd x ncell x 9, per site N x 9 (empty: none)
  void* astatic_d = nullptr;
  // forward record
  bool recorded = false;
  int64_t nsteps = 0;
  uint64_t seed = 0, step0 = 0;
  void* ckpt_d = nullptr;  // (nckpt) x N x 3
  int64_t ckpt_cap = 0;
  // segment and adjoint scratch
  void* seg_d = nullptr;  // records of one segment, see ensure_segment
  int seg_cap = 0;        // steps the records hold
  int64_t seg_k = -1;     // segment whose records seg_d holds (-1: none)
  void *lam_d = nullptr, *lamu_d = nullptr, *gep_d = nullptr, *nu1_d = nullptr, *nu2_d = nullptr, *part_d = nullptr;
  void* gw_d = nullptr;  // Depondt: the corrector's share of dL/dw1, N x 3
  void* ework_d = nullptr;
  double* gpart_d = nullptr;  // per-wavefront parameter-gradient partials, nblk x WAVES x MAXP
  double* gout_d = nullptr;   // their fold, MAXP
  int nblk = 0;
  // dense gradients: per forward bond 9 x ncell, per site 9 x N (FP64, internal cells), and the magex
  // cell -> internal cell map of the layout cmap_geo (null for row-major)
  double *gb_d = nullptr, *ga_d = nullptr;
  void* nuh_d = nullptr;  // the sweep's nu1 and nu2 of every step of a segment (2 x nuh_cap x N x 3)
  int nuh_cap = 0;
  int* cmap_d = nullptr;
  int cmap_geo = -1;
  // driven forward: the uniform field of each pass of each step (nsteps x 2 x 3; empty: the model's)
  std::vector<double> drive;
  // dL/dB of the uniform field per step and pass: per-wavefront partials (nsteps x 2) x nblk x WAVES x 3
  double* gf_d = nullptr;
  int64_t gf_rows = 0;
  bool want_field = false;
  int64_t sweep_step0 = 0;  // global step of the first record of the segment being swept
  // overlapped backward (small grids): a second segment buffer, replayed on a stream of its own while the
  // other is swept, that stream and its events
  void* seg2_d = nullptr;
  int seg2_cap = 0;
  int64_t seg2_k = -1;  // segment whose records seg2_d holds (-1: none)
  hipStream_t s2 = nullptr;
  hipEvent_t ev_start = nullptr, ev_rep[2] = {nullptr, nullptr}, ev_sweep[2] = {nullptr, nullptr};
  // the persistent sweep's own tagged records (2 x N x TAG_WORDS words) and control words, apart from
  // the model's, which the replays use at the same time
  uint64_t* srec_d = nullptr;
  uint32_t* sctl_d = nullptr;
  uint32_t srec_epoch = 0;
  // per-pass wavefront sums of the parameter terms of one persistently swept segment (2 x seg_cap passes x
  // nblk x WAVES rows x MAXP)
  double* psum_d = nullptr;
  int psum_cap = 0;
  bool side_by_side = false;  // replays and sweeps now run beside each other: ordinary launches
};

namespace {

struct AErr {
  int code;
  std::string msg;
};
[[noreturn]] void afail(int code, const std::string& msg) { throw AErr{code, msg}; }
void aok(hipError_t e, const char* what) {
  if (e != hipSuccess) afail(e == hipErrorOutOfMemory ? MX_ENOMEM : MX_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}
void aok_rc(int rc, const char* what) {
  if (rc != MX_OK) afail(rc, std::string(what) + ": " + mx_last_error());
}
template <typename F>
int aguard(F&& f) {
  try {
    return f();
  } catch (const AErr& e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return MX_ENOMEM;
  }
}
void afree(void*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

template <typename R>
void upload_astatic(mx_adjoint& a) {
  const Model& m = *a.m;
  AStatic<R> A;
  std::memset(&A, 0, sizeof(A));
  A.P = a.P;
  for (int k = 0; k < MAXENT; ++k) A.fslot[k] = k < m.nent ? a.fslot[k] : -1;
  for (int p = 0; p < a.P; ++p) {
    for (int k = 0; k < m.nent; ++k)
      for (int q = 0; q < 9; ++q) A.U[k][p][q] = R(a.U[(size_t(p) * m.nent + k) * 9 + q]);
    for (int s = 0; s < m.ns; ++s)
      for (int q = 0; q < 9; ++q) A.Aun[s][p][q] = R(a.Aun[(size_t(p) * m.ns + s) * 9 + q]);
  }
  if (!a.astatic_d) aok(hipMalloc(&a.astatic_d, sizeof(A)), "hipMalloc adjoint constants");
  aok(hipMemcpy(a.astatic_d, &A, sizeof(A), hipMemcpyHostToDevice), "hipMemcpy adjoint constants");
}

// The model's couplings become sum_p theta_p U[p] (one width-9 table row per entry) and its on-site
// tensors sum_p theta_p Aun[p].  With fixed parts (boff, aoff) they are per bond, boff_b + sum_p theta_p
// U[p][f(b)], and per site, aoff_i + sum_p theta_p Aun[p][s(i)], uploaded as the model's per-bond and
// per-site tensors.  A stencil too large for the constant table (nent x 9 > MAXTAB values) also takes
// the per-bond form, with a zero fixed part.
void apply_theta(mx_adjoint& a) {
  Model& m = *a.m;
  if (a.boff.empty() && m.nent * 9 <= MAXTAB) {
    m.table.assign(size_t(m.nent) * 9, 0.0);
    for (int k = 0; k < m.nent; ++k) {
      m.ent[k].slot = k;
      for (int p = 0; p < a.P; ++p)
        for (int q = 0; q < 9; ++q) m.table[9 * size_t(k) + q] += a.theta[p] * a.U[(size_t(p) * m.nent + k) * 9 + q];
    }
    m.ntypes = m.nent;
    m.width = 9;
    m.coup = COUP_TABLE9;
  } else {
    std::vector<int> pos_of_f(a.nfwd, -1);
    for (int k = 0; k < m.nent; ++k)
      if (a.fslot[k] >= 0) pos_of_f[a.fslot[k]] = k;
    std::vector<double> J(a.boff);
    if (J.empty()) J.assign(size_t(a.nfwd) * m.ncell * 9, 0.0);
    for (int f = 0; f < a.nfwd; ++f) {
      double u[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (int p = 0; p < a.P; ++p)
        for (int q = 0; q < 9; ++q) u[q] += a.theta[p] * a.U[(size_t(p) * m.nent + pos_of_f[f]) * 9 + q];
      for (int c = 0; c < m.ncell; ++c)
        for (int q = 0; q < 9; ++q) J[(size_t(f) * m.ncell + c) * 9 + q] += u[q];
    }
    const std::vector<int> before = m.ent_order;
    aok_rc(mx_model_set_bond_tensors(a.m, J.data()), "per-bond tensors");
    if (m.ent_order != before) {
      // the per-bond form pairs each same-sublattice bond's two entries (on its first upload): the units
      // and forward slots, held by entry position, move with their entries
      std::vector<int> old_pos(m.nent);
      for (int p = 0; p < m.nent; ++p) old_pos[before[p]] = p;
      std::vector<double> U(a.U.size());
      std::vector<int> fs(m.nent);
      for (int p = 0; p < m.nent; ++p) {
        const int o = old_pos[m.ent_order[p]];
        fs[p] = a.fslot[o];
        for (int r = 0; r < a.P; ++r)
          for (int q = 0; q < 9; ++q) U[(size_t(r) * m.nent + p) * 9 + q] = a.U[(size_t(r) * m.nent + o) * 9 + q];
      }
      a.U.swap(U);
      a.fslot.swap(fs);
      if (m.prec == MX_FP64) upload_astatic<double>(a);
      else upload_astatic<float>(a);
    }
  }
  if (a.aoff.empty()) {
    m.amode = ONSITE_SUB;
    m.has_A = 0;
    std::fill(m.Asub.begin(), m.Asub.end(), 0.0);
    for (int s = 0; s < m.ns; ++s)
      for (int p = 0; p < a.P; ++p)
        for (int q = 0; q < 9; ++q) {
          m.Asub[9 * s + q] += a.theta[p] * a.Aun[(size_t(p) * m.ns + s) * 9 + q];
          if (m.Asub[9 * s + q] != 0.0) m.has_A = 1;
        }
  } else {
    std::vector<double> A(a.aoff);
    for (int r = 0; r < m.N; ++r) {
      const int s = r % m.ns;
      for (int p = 0; p < a.P; ++p)
        for (int q = 0; q < 9; ++q) A[size_t(r) * 9 + q] += a.theta[p] * a.Aun[(size_t(p) * m.ns + s) * 9 + q];
    }
    aok_rc(mx_model_set_onsite(a.m, 1, A.data()), "per-site on-site tensors");
  }
  m.static_dirty = 1;
}

size_t nbytes3(const Model& m) { return size_t(m.N) * 3 * prec_size(m); }

template <typename R>
void copy_dev(const void* src, void* dst, size_t nelem, hipStream_t s) {
  hipLaunchKernelGGL((k_copy<R>), dim3((nelem + 255) / 256), dim3(256), 0, s, static_cast<const R*>(src),
                     static_cast<R*>(dst), nelem);
}

// Segment records for n steps: 5 arrays (e, ep, w, B(e), B(ep)) of n x N x 3, record (which, step)
// at (which * seg_cap + step) * N * 3 values.
void ensure_segment(mx_adjoint& a, int n) {
  if (n <= a.seg_cap) return;
  afree(a.seg_d);
  afree(a.seg2_d);
  a.seg_cap = a.seg2_cap = 0;
  a.seg_k = a.seg2_k = -1;
  aok(hipMalloc(&a.seg_d, 5 * size_t(n) * nbytes3(*a.m)), "hipMalloc segment records");
  a.seg_cap = n;
}

// Record (which, step n) of the segment buffer buf (seg_d or seg2_d, both seg_cap steps long).
char* seg_slot(mx_adjoint& a, void* buf, int which, int n) {
  return static_cast<char*>(buf) + (size_t(which) * a.seg_cap + n) * nbytes3(*a.m);
}

void to_internal(mx_adjoint& a, const void* src, void* dst, hipStream_t s) {
  const Model& m = *a.m;
  const size_t nel = size_t(m.N) * 3;
  if (needs_permute(m)) aok_rc(launch_permute(m, src, dst, 1, s), "permute");
  else if (m.prec == MX_FP64) copy_dev<double>(src, dst, nel, s);
  else copy_dev<float>(src, dst, nel, s);
}

void to_magex(mx_adjoint& a, const void* src, void* dst, hipStream_t s) {
  const Model& m = *a.m;
  const size_t nel = size_t(m.N) * 3;
  if (needs_permute(m)) aok_rc(launch_permute(m, src, dst, 0, s), "permute");
  else if (m.prec == MX_FP64) copy_dev<double>(src, dst, nel, s);
  else copy_dev<float>(src, dst, nel, s);
}

// Runs segment k (n_seg steps) from the magex-order spins src, saving every step's record into the
// segment buffer buf (seg_d or seg2_d) on stream s; the final spins stay in ework_d (internal order).
void record_segment(mx_adjoint& a, const void* src, int64_t k, int n_seg, hipStream_t s, void* buf) {
  Model& m = *a.m;
  void* ew = a.ework_d;
  (buf == a.seg_d ? a.seg_k : a.seg2_k) = -1;
  to_internal(a, src, ew, s);
  // undriven segments of two or more steps on a model that takes persistent launches: one launch writes
  // the same records, bit for bit, as the two single-pass kernels per step below
  if (a.drive.empty() && n_seg >= 2 && persistent_steps(m) && ensure_step_scratch(m) == MX_OK) {
    const StepRecOut ro{seg_slot(a, buf, 0, 0), seg_slot(a, buf, 1, 0), seg_slot(a, buf, 2, 0),
                        seg_slot(a, buf, 3, 0), seg_slot(a, buf, 4, 0)};
    aok_rc(launch_heun_steps(m, ew, a.seed, a.step0 + uint64_t(k * a.C), n_seg, nullptr, s, nullptr, &ro,
                             a.side_by_side),
           "recorded segment");
  } else {
    for (int n = 0; n < n_seg; ++n) {
      const uint64_t step = a.step0 + uint64_t(k * a.C + n);
      const double* B = a.drive.empty() ? nullptr : a.drive.data() + 6 * size_t(k * a.C + n);
      aok_rc(launch_heun_pass(m, 1, ew, seg_slot(a, buf, 1, n), seg_slot(a, buf, 2, n), a.seed, step,
                              seg_slot(a, buf, 0, n), seg_slot(a, buf, 3, n), s, B),
             "segment pass 1");
      aok_rc(launch_heun_pass(m, 2, ew, seg_slot(a, buf, 1, n), seg_slot(a, buf, 2, n), a.seed, step, nullptr,
                              seg_slot(a, buf, 4, n), s, B ? B + 3 : nullptr),
             "segment pass 2");
    }
  }
  (buf == a.seg_d ? a.seg_k : a.seg2_k) = k;
}

void ensure_buffers(mx_adjoint& a) {
  const Model& m = *a.m;
  const size_t b = nbytes3(m);
  for (void** p : {&a.lam_d, &a.lamu_d, &a.gep_d, &a.nu1_d, &a.nu2_d, &a.part_d, &a.ework_d, &a.gw_d})
    if (!*p) aok(hipMalloc(p, b), "hipMalloc adjoint state");
  a.nblk = m.ns * ((m.ncell + BLOCK - 1) / BLOCK);
  if (!a.gpart_d)
    aok(hipMalloc(reinterpret_cast<void**>(&a.gpart_d), sizeof(double) * MAXP * WAVES * a.nblk), "hipMalloc gpart");
  if (!a.gout_d) aok(hipMalloc(reinterpret_cast<void**>(&a.gout_d), sizeof(double) * MAXP), "hipMalloc gout");
}


// The persistent sweep's own tagged records and control words (zeroed when allocated).
void ensure_sweep_sync(mx_adjoint& a) {
  const Model& m = *a.m;
  if (!a.srec_d) {
    const size_t n = 2 * size_t(m.N) * (m.prec == MX_FP64 ? TAG_WORDS<double> : TAG_WORDS<float>) * sizeof(uint64_t);
    aok(hipMalloc(reinterpret_cast<void**>(&a.srec_d), n), "hipMalloc sweep records");
    aok(hipMemset(a.srec_d, 0, n), "hipMemset sweep records");
  }
  if (!a.sctl_d) {
    aok(hipMalloc(reinterpret_cast<void**>(&a.sctl_d), CTL_WORDS * sizeof(uint32_t)), "hipMalloc sweep control");
    aok(hipMemset(a.sctl_d, 0, CTL_WORDS * sizeof(uint32_t)), "hipMemset sweep control");
  }
}

// The per-step nu1 / nu2 slots of a segment (2 x seg_cap x N x 3), for the dense terms and for the
// persistent sweep's parameter terms.
void ensure_nuh(mx_adjoint& a) {
  if (a.nuh_cap >= a.seg_cap) return;
  afree(a.nuh_d);
  a.nuh_cap = 0;
  aok(hipMalloc(&a.nuh_d, 2 * size_t(a.seg_cap) * nbytes3(*a.m)), "hipMalloc adjoint vectors of a segment");
  a.nuh_cap = a.seg_cap;
}

// The whole sweep of a segment of two or more steps in persistent launches of up to STEPS_PER_LAUNCH
// steps (k_adj_seg), on models that take the persistent forward launch; false when the model does not.
// Its tagged records and control words are its own (replays of the next segment may run beside it on
// the model's), consecutive launches take different epochs, and an aborted launch is counted with the
// model's given-up waits.  Every step's nu1 and nu2 go to the segment's slots (nuh), and k_par_seg /
// k_par_fold then add the parameter terms of all the segment's passes to the partial rows.
template <typename R, int GEO, int NP, int CM, bool LEAN>
bool sweep_persistent(mx_adjoint& a, const Dyn<R>& d, const R* seg, int C, int n_seg, hipStream_t s) {
  if constexpr (GEO != GEO_PERIODIC && GEO != GEO_OPEN) {
    return false;
  } else {
    Model& m = *a.m;
    if (n_seg < 2 || !persistent_steps(m) || ensure_step_scratch(m) != MX_OK || !m.stepsync_d) return false;
    ensure_sweep_sync(a);
    ensure_nuh(a);
    const int nrow = a.nblk * WAVES;
    if (a.psum_cap < 2 * a.seg_cap) {
      if (a.psum_d) (void)hipFree(a.psum_d);
      a.psum_d = nullptr;
      a.psum_cap = 0;
      aok(hipMalloc(reinterpret_cast<void**>(&a.psum_d), sizeof(double) * 2 * size_t(a.seg_cap) * nrow * MAXP),
          "hipMalloc parameter sums");
      a.psum_cap = 2 * a.seg_cap;
    }
    R* nuh = static_cast<R*>(a.nuh_d);
    const Static<R>* S = static_cast<const Static<R>*>(m.static_d);
    const AStatic<R>* A = static_cast<const AStatic<R>*>(a.astatic_d);
    StepSync sy{};
    sy.ctl = a.sctl_d;
    sy.spin = m.wait_polls;
    uint32_t* gave_up = m.stepsync_d + stepsync_ctl(m) + 2;
    R* lam = static_cast<R*>(a.lam_d);
    double* gf = a.want_field ? a.gf_d + size_t(a.sweep_step0) * 2 * a.nblk * WAVES * 3 : nullptr;
    uint64_t* rec = a.srec_d;
    static_assert(2 * STEPS_PER_LAUNCH + 2 < (1 << 12), "sweep tags overflow their 12 bits");
    for (int hi = n_seg; hi > 0; hi -= STEPS_PER_LAUNCH) {
      int lo = std::max(0, hi - STEPS_PER_LAUNCH);
      a.srec_epoch = a.srec_epoch % 15u + 1u;
      uint32_t tag0 = a.srec_epoch << 12;
      void* args[] = {&S, &A, const_cast<Dyn<R>*>(&d), &seg, &C, &hi, &lo, &lam, &nuh, &gf, &rec, &tag0, &sy,
                      &gave_up};
      // beside a replay, an ordinary launch: the runtime runs cooperative launches one at a time, and
      // both grids fit on the GPU together (overlap_backward)
      const void* fn = reinterpret_cast<const void*>(k_adj_seg<R, GEO, NP, CM, LEAN>);
      aok(a.side_by_side ? hipLaunchKernel(fn, dim3(a.nblk), dim3(BLOCK), args, SIDE_BY_SIDE_LDS, s)
                         : hipLaunchCooperativeKernel(fn, dim3(a.nblk), dim3(BLOCK), args, 0, s),
          "persistent adjoint sweep");
    }
    for (int q0 = 0; q0 < 2 * n_seg; q0 += 32768)  // grid rows: at most 65535
      hipLaunchKernelGGL((k_par_seg<R, GEO, NP>), dim3(a.nblk, std::min(32768, 2 * n_seg - q0)), dim3(BLOCK), 0, s, S,
                         A, seg, nuh, C, n_seg, q0, a.psum_d);
    hipLaunchKernelGGL((k_par_fold<NP>), dim3((nrow * NP + 255) / 256), dim3(256), 0, s, a.psum_d, 2 * n_seg, nrow,
                       a.gpart_d);
    return true;
  }
}

// Forward entries per k_dense_seg thread, and its grid rows: groups of that many forward entries of the
// largest sublattice stencil (at least one row, which also holds the on-site terms).
constexpr int DENSE_FG = 4;
int dense_rows(const mx_adjoint& a) {
  const Model& m = *a.m;
  int mx = 0;
  for (int sl = 0; sl < m.ns; ++sl) {
    int c = 0;
    for (int k = m.sptr[sl]; k < m.sptr[sl + 1]; ++k) c += a.fslot[k] >= 0 ? 1 : 0;
    mx = std::max(mx, c);
  }
  return std::max(1, (mx + DENSE_FG - 1) / DENSE_FG);
}

// Adjoint sweep over one replayed segment of n_seg steps (records rec(n), newest last): the local
// part of the segment's last step from lam, then per step (newest first) k_adj_b and k_adj_c, the
// latter fused with the next-older step's local part.  Leaves dL/de of the segment's first input in lam.
// With dense gradients (nuh: the segment's nu1 / nu2 slots, seg: its records, C steps apart) every step
// keeps its own nu1 and nu2, and k_dense_seg adds the segment's bond and site terms afterwards.
template <typename R, int GEO, int NP, int CM, bool LEAN>
void sweep_t(mx_adjoint& a, const StepRec<R>* rec, int n_seg, hipStream_t s, R* nuh, const R* seg, int C) {
  const Model& m = *a.m;
  const Static<R>* S = static_cast<const Static<R>*>(m.static_d);
  const AStatic<R>* A = static_cast<const AStatic<R>*>(a.astatic_d);
  Dyn<R> d{};
  d.dt = R(m.dt);
  d.prec = m.precession;
  d.alpha = static_cast<const R*>(m.alpha_d);
  d.inv_mu = static_cast<const R*>(m.inv_mu_d);
  d.bond = static_cast<const R*>(m.bond_d);
  d.A = static_cast<const R*>(m.A_d);
  d.pinned = static_cast<const uint8_t*>(m.pinned_d);
  d.integ = m.integrator;
  d.Kc = static_cast<const R*>(m.Kc_d);
  d.K2 = static_cast<const R*>(m.K2_d);
  d.K3 = static_cast<const R*>(m.K3_d);
  d.sot = (m.sot_dl != 0.0 || m.sot_fl != 0.0) ? 1 : 0;
  d.sp0 = R(m.sot_p[0]);
  d.sp1 = R(m.sot_p[1]);
  d.sp2 = R(m.sot_p[2]);
  d.bdl = R(m.sot_dl);
  d.bfl = R(m.sot_fl);
  R* lam = static_cast<R*>(a.lam_d);
  R* lamu = static_cast<R*>(a.lamu_d);
  R* gep = static_cast<R*>(a.gep_d);
  const size_t N3 = size_t(m.N) * 3;
  // nu1 / nu2 of step n: the shared buffers, or the step's own slots when the dense terms follow
  auto nu1_of = [&](int n) { return nuh ? nuh + size_t(n) * N3 : static_cast<R*>(a.nu1_d); };
  auto nu2_of = [&](int n) { return nuh ? nuh + (size_t(C) + n) * N3 : static_cast<R*>(a.nu2_d); };
  R* part = static_cast<R*>(a.part_d);
  R* gwb = static_cast<R*>(a.gw_d);
  const dim3 g(a.nblk), b(BLOCK);
  // field-gradient rows: (global step * 2 + pass index) x nblk x WAVES x 3, pass index 0 = predictor, 1 = corrector
  auto frow = [&](int n, int pass) -> double* {
    return a.want_field ? a.gf_d + (size_t(a.sweep_step0 + n) * 2 + pass) * a.nblk * WAVES * 3 : nullptr;
  };
  if (sweep_persistent<R, GEO, NP, CM, LEAN>(a, d, seg, C, n_seg, s)) {
    if (nuh)
      hipLaunchKernelGGL((k_dense_seg<R, GEO, DENSE_FG>), dim3(a.nblk, dense_rows(a)), b, 0, s, S, A, seg, nuh, C,
                         n_seg, a.gb_d, a.ga_d);
    return;
  }
  hipLaunchKernelGGL((k_adj_a<R, GEO, LEAN>), g, b, 0, s, S, d, lam, rec[n_seg - 1], lamu, gep, nu2_of(n_seg - 1),
                     gwb);
  for (int n = n_seg - 1; n >= 0; --n) {
    hipLaunchKernelGGL((k_adj_b<R, GEO, NP, CM, LEAN>), g, b, 0, s, S, A, d, rec[n], lamu, gep, nu2_of(n), nu1_of(n),
                       part, a.gpart_d, frow(n, 1), gwb);
    if (n > 0)
      hipLaunchKernelGGL((k_adj_c<R, GEO, NP, true, CM, LEAN>), g, b, 0, s, S, A, d, rec[n].e, nu1_of(n), part, lam,
                         rec[n - 1], lamu, gep, nu2_of(n - 1), a.gpart_d, frow(n, 0), gwb);
    else
      hipLaunchKernelGGL((k_adj_c<R, GEO, NP, false, CM, LEAN>), g, b, 0, s, S, A, d, rec[n].e, nu1_of(n), part, lam,
                         rec[n], lamu, gep, nu2_of(n), a.gpart_d, frow(n, 0), gwb);
  }
  if (nuh)
    hipLaunchKernelGGL((k_dense_seg<R, GEO, DENSE_FG>), dim3(a.nblk, dense_rows(a)), b, 0, s, S, A, seg, nuh, C,
                       n_seg, a.gb_d, a.ga_d);
}

template <typename R, int NP, int CM, bool LEAN>
void sweep_geo(mx_adjoint& a, const StepRec<R>* rec, int n_seg, hipStream_t s, R* nuh, const R* seg, int C) {
  const int geo = a.m->geo;
  if (geo == GEO_OPEN) sweep_t<R, GEO_OPEN, NP, CM, LEAN>(a, rec, n_seg, s, nuh, seg, C);
  else if (geo == GEO_PERIODIC) sweep_t<R, GEO_PERIODIC, NP, CM, LEAN>(a, rec, n_seg, s, nuh, seg, C);
  else sweep_t<R, GEO_TILED, NP, CM, LEAN>(a, rec, n_seg, s, nuh, seg, C);
}

// The plain model (see LEAN above).
bool lean_adjoint(const Model& m) {
  return m.integrator == MX_HEUN && !m.pinned_d && !m.has_c && !m.has_u && m.sot_dl == 0.0 && m.sot_fl == 0.0;
}

template <typename R, int NP>
void sweep_cm(mx_adjoint& a, const StepRec<R>* rec, int n_seg, hipStream_t s, R* nuh, const R* seg, int C) {
  const bool bond = a.m->coup == COUP_BOND;
  if (lean_adjoint(*a.m)) {
    if (bond) sweep_geo<R, NP, COUP_BOND, true>(a, rec, n_seg, s, nuh, seg, C);
    else sweep_geo<R, NP, COUP_TABLE9, true>(a, rec, n_seg, s, nuh, seg, C);
    return;
  }
  if (bond) sweep_geo<R, NP, COUP_BOND, false>(a, rec, n_seg, s, nuh, seg, C);
  else sweep_geo<R, NP, COUP_TABLE9, false>(a, rec, n_seg, s, nuh, seg, C);
}

template <typename R>
void sweep(mx_adjoint& a, char* seg, size_t b, int C, int n_seg, bool dense, hipStream_t s) {
  std::vector<StepRec<R>> rec(n_seg);
  for (int n = 0; n < n_seg; ++n) {
    auto at = [&](int which) { return reinterpret_cast<const R*>(seg + (size_t(which) * C + n) * b); };
    rec[n] = StepRec<R>{at(0), at(1), at(2), at(3), at(4)};
  }
  R* nuh = dense ? static_cast<R*>(a.nuh_d) : nullptr;
  const R* sg = reinterpret_cast<const R*>(seg);
  if (a.P <= 4) sweep_cm<R, 4>(a, rec.data(), n_seg, s, nuh, sg, C);
  else if (a.P <= 8) sweep_cm<R, 8>(a, rec.data(), n_seg, s, nuh, sg, C);
  else sweep_cm<R, 16>(a, rec.data(), n_seg, s, nuh, sg, C);
}

// Accumulators of the dense gradients (cleared), the per-step nu1 / nu2 slots of a segment (2 x seg_cap
// x N x 3) and, for a permuted layout, the cell map.
void ensure_dense(mx_adjoint& a, hipStream_t s) {
  const Model& m = *a.m;
  const size_t nb = size_t(a.nfwd) * 9 * m.ncell, na = size_t(9) * m.N;
  ensure_nuh(a);
  if (!a.gb_d) aok(hipMalloc(reinterpret_cast<void**>(&a.gb_d), sizeof(double) * std::max<size_t>(nb, 1)), "hipMalloc dense bond gradient");
  if (!a.ga_d) aok(hipMalloc(reinterpret_cast<void**>(&a.ga_d), sizeof(double) * na), "hipMalloc dense on-site gradient");
  if (nb) hipLaunchKernelGGL((k_zero<double>), dim3((nb + 255) / 256), dim3(256), 0, s, a.gb_d, nb);
  hipLaunchKernelGGL((k_zero<double>), dim3((na + 255) / 256), dim3(256), 0, s, a.ga_d, na);
  if (m.geo == GEO_TILED && a.cmap_geo != m.geo) {
    std::vector<int> map(m.ncell);
    for (int c = 0; c < m.ncell; ++c) map[c] = cell_to_internal(m, c);
    if (!a.cmap_d) aok(hipMalloc(reinterpret_cast<void**>(&a.cmap_d), sizeof(int) * m.ncell), "hipMalloc cell map");
    aok(hipMemcpy(a.cmap_d, map.data(), sizeof(int) * m.ncell, hipMemcpyHostToDevice), "hipMemcpy cell map");
    a.cmap_geo = m.geo;
  }
}

// Small grids (persistent launches, two of which fit on the GPU at once) replay the next segment on a
// stream of their own while the current one is swept: both are latency-bound launches of a few dozen
// workgroups, so running them side by side hides the replays.
bool overlap_backward(const Model& m) {
  if (!persistent_steps(m)) return false;
  int dev_cu = 0;
  if (hipDeviceGetAttribute(&dev_cu, hipDeviceAttributeMultiprocessorCount, m.device) != hipSuccess) return false;
  return 2 * (m.ns * ((m.ncell + BLOCK - 1) / BLOCK)) <= dev_cu;
}

// The second segment buffer, the replay stream (non-blocking, default priority) and the events.
void ensure_overlap(mx_adjoint& a) {
  if (!a.seg2_d || a.seg2_cap != a.seg_cap) {
    afree(a.seg2_d);
    a.seg2_cap = 0;
    a.seg2_k = -1;
    aok(hipMalloc(&a.seg2_d, 5 * size_t(a.seg_cap) * nbytes3(*a.m)), "hipMalloc second segment records");
    a.seg2_cap = a.seg_cap;
  }
  if (!a.s2) aok(hipStreamCreateWithFlags(&a.s2, hipStreamNonBlocking), "hipStreamCreateWithFlags");
  for (hipEvent_t* e : {&a.ev_start, &a.ev_rep[0], &a.ev_rep[1], &a.ev_sweep[0], &a.ev_sweep[1]})
    if (!*e) aok(hipEventCreateWithFlags(e, hipEventDisableTiming), "hipEventCreateWithFlags");
}

// The adjoint sweep over every segment, newest first; with grad_bonds / grad_onsite also the dense
// gradients (either may be null).
int backward_impl(mx_adjoint* a, const void* grad_e, void* grad_e0, double* grad_theta, double* grad_bonds,
                  double* grad_onsite, double* grad_field, void* stream) {
  if (!a || !grad_e || !grad_theta) afail(MX_EINVAL, "null argument");
  Model& m = *a->m;
  if (!a->recorded) afail(MX_EINVAL, "no forward recorded for the current parameters: run mx_adjoint_forward");
  aok(hipSetDevice(m.device), "hipSetDevice");
  aok_rc(sync_model_static(m), "model constants");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t b = nbytes3(m);
  const bool fp64 = m.prec == MX_FP64;
  const bool dense = grad_bonds || grad_onsite;
  if (dense) ensure_dense(*a, s);
  a->want_field = grad_field != nullptr && a->nsteps > 0;
  if (a->want_field && a->gf_rows < 2 * a->nsteps) {
    if (a->gf_d) (void)hipFree(a->gf_d);
    a->gf_d = nullptr;
    a->gf_rows = 0;
    aok(hipMalloc(reinterpret_cast<void**>(&a->gf_d), sizeof(double) * 2 * size_t(a->nsteps) * a->nblk * WAVES * 3),
        "hipMalloc field-gradient partials");
    a->gf_rows = 2 * a->nsteps;
  }
  // the adjoint works in the library's internal site order; permute at entry and exit
  to_internal(*a, grad_e, a->lam_d, s);
  const int nrow = a->nblk * WAVES;
  hipLaunchKernelGGL((k_zero<double>), dim3((MAXP * nrow + 255) / 256), dim3(256), 0, s, a->gpart_d,
                     size_t(MAXP) * nrow);
  char* ck = static_cast<char*>(a->ckpt_d);
  const int C = a->C;
  const int64_t K = a->nsteps > 0 ? (a->nsteps - 1) / C + 1 : 0;  // segments
  auto n_of = [&](int64_t k) { return int(std::min<int64_t>(C, a->nsteps - k * C)); };
  auto sweep_from = [&](int64_t k, void* buf) {
    a->sweep_step0 = k * C;
    if (fp64) sweep<double>(*a, static_cast<char*>(buf), b, a->seg_cap, n_of(k), dense, s);
    else sweep<float>(*a, static_cast<char*>(buf), b, a->seg_cap, n_of(k), dense, s);
  };
  if (K >= 2 && overlap_backward(m)) {
    // segment k is swept from buffer (K - 1 - k) % 2 on s while segment k - 1 is replayed into the other
    // buffer on s2; a buffer is replayed into only once the sweep that read it (two segments later) has
    // ended, and a sweep starts only once its replay has.  The replays share ework_d, in order on s2.
    ensure_overlap(*a);
    void* buf[2] = {a->seg_d, a->seg2_d};
    if (a->seg_k != K - 1) record_segment(*a, ck + size_t(K - 1) * b, K - 1, n_of(K - 1), s, buf[0]);
    aok(hipEventRecord(a->ev_start, s), "hipEventRecord");
    aok(hipStreamWaitEvent(a->s2, a->ev_start, 0), "hipStreamWaitEvent");
    struct SideBySide {  // set for the loop, cleared on every exit
      mx_adjoint& a;
      explicit SideBySide(mx_adjoint& x) : a(x) { a.side_by_side = true; }
      ~SideBySide() { a.side_by_side = false; }
    } side(*a);
    for (int64_t k = K - 1; k >= 0; --k) {
      const int cur = int((K - 1 - k) & 1), nxt = cur ^ 1;
      if (k >= 1) {
        if (k + 1 <= K - 1) aok(hipStreamWaitEvent(a->s2, a->ev_sweep[nxt], 0), "hipStreamWaitEvent");
        record_segment(*a, ck + size_t(k - 1) * b, k - 1, n_of(k - 1), a->s2, buf[nxt]);
        aok(hipEventRecord(a->ev_rep[nxt], a->s2), "hipEventRecord");
      }
      if (k < K - 1) aok(hipStreamWaitEvent(s, a->ev_rep[cur], 0), "hipStreamWaitEvent");
      sweep_from(k, buf[cur]);
      aok(hipEventRecord(a->ev_sweep[cur], s), "hipEventRecord");
    }
  } else {
    for (int64_t k = K - 1; k >= 0; --k) {
      // replay the segment from its checkpoint unless its records are still held
      if (a->seg_k != k) record_segment(*a, ck + size_t(k) * b, k, n_of(k), s, a->seg_d);
      sweep_from(k, a->seg_d);
    }
  }
  hipLaunchKernelGGL((k_final<MAXP>), dim3(1), dim3(BLOCK), 0, s, a->gpart_d, nrow, MAXP, 1.0, a->gout_d);
  hipLaunchKernelGGL((k_copy<double>), dim3(1), dim3(64), 0, s, a->gout_d, grad_theta, size_t(a->P));
  if (grad_e0) to_magex(*a, a->lam_d, grad_e0, s);
  if (a->want_field)  // every (step, pass) row was written once by the sweep
    hipLaunchKernelGGL((k_final<3>), dim3(unsigned(2 * a->nsteps)), dim3(BLOCK), 0, s, a->gf_d, nrow, 3, 1.0,
                       grad_field);
  a->want_field = false;
  const int* cmap = m.geo == GEO_TILED ? a->cmap_d : nullptr;
  if (grad_bonds && a->nfwd) {
    const size_t n = size_t(a->nfwd) * m.ncell * 9;
    hipLaunchKernelGGL(k_dense_bonds, dim3((n + 255) / 256), dim3(256), 0, s, a->gb_d, cmap, a->nfwd, m.ncell,
                       grad_bonds);
  }
  if (grad_onsite) {
    const size_t n = size_t(m.N) * 9;
    hipLaunchKernelGGL(k_dense_onsite, dim3((n + 255) / 256), dim3(256), 0, s, a->ga_d, cmap, m.ns, m.ncell,
                       grad_onsite);
  }
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) afail(MX_EHIP, std::string("adjoint: ") + hipGetErrorString(err));
  return MX_OK;
}

}  // namespace

extern "C" {

int mx_adjoint_create(mx_model* m, int32_t nparam, const double* unit_bonds, const double* unit_onsite,
                      int32_t checkpoint_every, mx_adjoint** out) {
  return aguard([&]() -> int {
    if (!m || !out || !unit_bonds || !unit_onsite) afail(MX_EINVAL, "null argument");
    *out = nullptr;
    if (nparam < 1 || nparam > MAXP) afail(MX_EUNSUPPORTED, "1 to 16 parameters");
    if (checkpoint_every < 1) afail(MX_EINVAL, "checkpoint interval must be positive");
    if (m->nent < 1) afail(MX_EINVAL, "set the stencil first");
    // the couplings and on-site tensors are replaced by the parameterized ones (and any fixed parts
    // given by mx_adjoint_set_offsets); cubic and higher uniaxial terms stay as the model has them
    if (m->geo == GEO_CHECKER) afail(MX_EUNSUPPORTED, "the adjoint runs on the row-major or tiled cell layouts");
    // unit tensors arrive in the caller's stencil order; reorder into the library's, and check that
    // every entry's partner carries the transposed tensors (the adjoint gather relies on J_ji = J_ij^T)
    const int n = m->nent;
    std::vector<int> pos_of(n);
    for (int pos = 0; pos < n; ++pos) pos_of[m->ent_order[pos]] = pos;
    auto raw = [&](int k, int q) { return m->ent_raw[6 * size_t(k) + q]; };
    for (int k = 0; k < n; ++k) {
      int partner = -1;
      for (int q = 0; q < n; ++q)
        if (raw(q, 0) == raw(k, 1) && raw(q, 1) == raw(k, 0) && raw(q, 2) == -raw(k, 2) && raw(q, 3) == -raw(k, 3) &&
            raw(q, 4) == -raw(k, 4))
          partner = q;
      if (partner < 0) afail(MX_EINVAL, "stencil entry without a reverse partner");
      for (int p = 0; p < nparam; ++p)
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) {
            const double x = unit_bonds[(size_t(p) * n + k) * 9 + 3 * r + c];
            const double y = unit_bonds[(size_t(p) * n + partner) * 9 + 3 * c + r];
            if (std::fabs(x - y) > 1e-12 * (std::fabs(x) + 1.0))
              afail(MX_EINVAL, "unit couplings of a bond and its reverse must be transposes");
          }
    }
    auto* a = new mx_adjoint();
    try {
      a->m = m;
      a->P = nparam;
      a->C = checkpoint_every;
      a->U.assign(size_t(nparam) * n * 9, 0.0);
      for (int p = 0; p < nparam; ++p)
        for (int k = 0; k < n; ++k)
          for (int q = 0; q < 9; ++q)
            a->U[(size_t(p) * n + pos_of[k]) * 9 + q] = unit_bonds[(size_t(p) * n + k) * 9 + q];
      a->Aun.assign(unit_onsite, unit_onsite + size_t(nparam) * m->ns * 9);
      a->theta.assign(nparam, 0.0);
      // forward entries (s < t, or s == t and R > 0) get bond slots in the caller's order
      std::vector<int> fw(n, -1);
      for (int k = 0; k < n; ++k) {
        const std::tuple<int, int, int> R{raw(k, 2), raw(k, 3), raw(k, 4)};
        if (raw(k, 0) < raw(k, 1) || (raw(k, 0) == raw(k, 1) && R > std::tuple<int, int, int>{0, 0, 0}))
          fw[k] = a->nfwd++;
      }
      a->fslot.assign(n, -1);
      for (int k = 0; k < n; ++k) a->fslot[pos_of[k]] = fw[k];
      aok(hipSetDevice(m->device), "hipSetDevice");
      if (m->prec == MX_FP64) upload_astatic<double>(*a);
      else upload_astatic<float>(*a);
      apply_theta(*a);
    } catch (...) {
      mx_adjoint_destroy(a);
      throw;
    }
    *out = a;
    return MX_OK;
  });
}

void mx_adjoint_destroy(mx_adjoint* a) {
  if (!a) return;
  if (a->m) (void)hipSetDevice(a->m->device);
  for (void** p : {&a->astatic_d, &a->ckpt_d, &a->seg_d, &a->lam_d, &a->lamu_d, &a->gep_d, &a->nu1_d, &a->nu2_d,
                   &a->part_d, &a->ework_d, reinterpret_cast<void**>(&a->gpart_d),
                   reinterpret_cast<void**>(&a->gout_d), reinterpret_cast<void**>(&a->gb_d),
                   reinterpret_cast<void**>(&a->ga_d), &a->nuh_d, reinterpret_cast<void**>(&a->cmap_d),
                   reinterpret_cast<void**>(&a->gf_d), &a->gw_d, &a->seg2_d, reinterpret_cast<void**>(&a->srec_d),
                   reinterpret_cast<void**>(&a->sctl_d), reinterpret_cast<void**>(&a->psum_d)})
    afree(*p);
  for (hipEvent_t e : {a->ev_start, a->ev_rep[0], a->ev_rep[1], a->ev_sweep[0], a->ev_sweep[1]})
    if (e) (void)hipEventDestroy(e);
  if (a->s2) (void)hipStreamDestroy(a->s2);
  delete a;
}

int mx_adjoint_set_params(mx_adjoint* a, const double* theta) {
  return aguard([&]() -> int {
    if (!a || !theta) afail(MX_EINVAL, "null argument");
    a->theta.assign(theta, theta + a->P);
    apply_theta(*a);
    a->recorded = false;  // a recorded forward belongs to the previous parameters
    return MX_OK;
  });
}

namespace {

// The recorded forward: nsteps steps of e from step0, checkpoints every C steps; B (host, nsteps x 2 x 3)
// drives the uniform field of each pass, or is null for the model's own field.
int forward_impl(mx_adjoint* a, void* e, int64_t nsteps, const double* B, uint64_t seed, uint64_t step0,
                 void* stream) {
  if (!a || !e) afail(MX_EINVAL, "null argument");
  if (nsteps < 0) afail(MX_EINVAL, "negative step count");
  if (B) a->drive.assign(B, B + 6 * size_t(nsteps));
  else a->drive.clear();
  Model& m = *a->m;
  if (!m.dyn_set) afail(MX_EINVAL, "set the dynamics first");
  aok(hipSetDevice(m.device), "hipSetDevice");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t nck = nsteps / a->C + 1;
  const size_t b = nbytes3(m);
  if (nck > a->ckpt_cap) {
    afree(a->ckpt_d);
    aok(hipMalloc(&a->ckpt_d, size_t(nck) * b), "hipMalloc checkpoints");
    a->ckpt_cap = nck;
  }
  ensure_buffers(*a);
  a->recorded = false;
  a->nsteps = nsteps;
  a->seed = seed;
  a->step0 = step0;
  // checkpoint k holds the (magex-order) spins after k * C steps
  char* ck = static_cast<char*>(a->ckpt_d);
  const size_t nel = size_t(m.N) * 3;
  // the last segment runs with the single-pass kernels and keeps its records, so the backward starts
  // without a replay
  const int64_t k_last = nsteps > 0 ? (nsteps - 1) / a->C : -1;
  ensure_segment(*a, int(std::min<int64_t>(a->C, std::max<int64_t>(nsteps, 1))));
  a->seg_k = -1;
  for (int64_t k = 0; k * a->C <= nsteps; ++k) {
    if (m.prec == MX_FP64) copy_dev<double>(e, ck + size_t(k) * b, nel, s);
    else copy_dev<float>(e, ck + size_t(k) * b, nel, s);
    const int64_t n = std::min<int64_t>(a->C, nsteps - k * a->C);
    if (n <= 0) continue;
    if (k == k_last) {
      aok_rc(sync_model_static(m), "model constants");
      record_segment(*a, e, k, int(n), s, a->seg_d);
      to_magex(*a, a->ework_d, e, s);
    } else if (B) {
      aok_rc(mx_llg_run_driven(a->m, e, n, B + 6 * size_t(k * a->C), seed, step0 + uint64_t(k * a->C), stream),
             "forward");
    } else {
      aok_rc(mx_llg_run(a->m, e, n, seed, step0 + uint64_t(k * a->C), nullptr, stream), "forward");
    }
  }
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) afail(MX_EHIP, std::string("adjoint forward: ") + hipGetErrorString(err));
  a->recorded = true;
  return MX_OK;
}

}  // namespace

int mx_adjoint_forward(mx_adjoint* a, void* e, int64_t nsteps, uint64_t seed, uint64_t step0, void* stream) {
  return aguard([&]() -> int { return forward_impl(a, e, nsteps, nullptr, seed, step0, stream); });
}

int mx_adjoint_forward_driven(mx_adjoint* a, void* e, int64_t nsteps, const double* B, uint64_t seed,
                              uint64_t step0, void* stream) {
  return aguard([&]() -> int {
    if (!B) afail(MX_EINVAL, "null field sequence");
    return forward_impl(a, e, nsteps, B, seed, step0, stream);
  });
}

int mx_adjoint_backward(mx_adjoint* a, const void* grad_e, void* grad_e0, double* grad_theta, void* stream) {
  return aguard([&]() -> int { return backward_impl(a, grad_e, grad_e0, grad_theta, nullptr, nullptr, nullptr, stream); });
}

int mx_adjoint_backward_dense(mx_adjoint* a, const void* grad_e, void* grad_e0, double* grad_theta,
                              double* grad_bonds, double* grad_onsite, void* stream) {
  return aguard(
      [&]() -> int { return backward_impl(a, grad_e, grad_e0, grad_theta, grad_bonds, grad_onsite, nullptr, stream); });
}

int mx_adjoint_backward_field(mx_adjoint* a, const void* grad_e, void* grad_e0, double* grad_theta, double* grad_field,
                              void* stream) {
  return aguard([&]() -> int {
    if (!grad_field) afail(MX_EINVAL, "null field-gradient output");
    return backward_impl(a, grad_e, grad_e0, grad_theta, nullptr, nullptr, grad_field, stream);
  });
}

int mx_adjoint_num_bonds(const mx_adjoint* a) { return a ? a->nfwd : -1; }

int mx_adjoint_set_offsets(mx_adjoint* a, const double* bond_offset, const double* onsite_offset) {
  return aguard([&]() -> int {
    if (!a) afail(MX_EINVAL, "null argument");
    Model& m = *a->m;
    if (bond_offset && a->nfwd == 0) afail(MX_EINVAL, "per-bond offsets need forward stencil entries");
    aok(hipSetDevice(m.device), "hipSetDevice");
    a->boff.clear();
    a->aoff.clear();
    if (bond_offset) a->boff.assign(bond_offset, bond_offset + size_t(a->nfwd) * m.ncell * 9);
    if (onsite_offset) a->aoff.assign(onsite_offset, onsite_offset + size_t(m.N) * 9);
    apply_theta(*a);
    a->recorded = false;
    return MX_OK;
  });
}

}  // extern "C"
