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
#include <cstring>

#include "mxllg_field.h"

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
template <typename R, int CM, int GEO, int PASS, bool LEAN = false>
__global__ __launch_bounds__(BLOCK) void k_heun(const Static<R>* __restrict__ S, const Dyn<R> d, R* __restrict__ e,
                                                R* __restrict__ ep, R* __restrict__ k1, uint64_t step,
                                                const R* __restrict__ noise) {
  Site st;
  if (!site_of<R, GEO>(S, st, PASS == 2 && d.rev2, true)) return;
  if (d.step_ptr) step += *d.step_ptr;  // graph replays: step = device counter + offset of this node
  const int i = st.i;
  const bool pin = !LEAN && d.pinned && d.pinned[i];
  const R a = (!LEAN && d.alpha) ? d.alpha[i] : S->alpha_sub[st.s];
  if (PASS == 1) {
    const R x = e[3 * i], y = e[3 * i + 1], z = e[3 * i + 2];
    if (pin) {
      ep[3 * i] = x;
      ep[3 * i + 1] = y;
      ep[3 * i + 2] = z;
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
    local_field<R, CM, GEO, LEAN>(S, d, e, st, x, y, z, step, noise, bx, by, bz);
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
      ep[3 * i] = px;
      ep[3 * i + 1] = py;
      ep[3 * i + 2] = pz;
      return;
    }
    const R hdt = R(0.5) * d.dt;
    k1[3 * i] = x + hdt * fx;
    k1[3 * i + 1] = y + hdt * fy;
    k1[3 * i + 2] = z + hdt * fz;
    const R px = x + d.dt * fx, py = y + d.dt * fy, pz = z + d.dt * fz;
    const R q = rnorm3(px, py, pz);
    ep[3 * i] = px * q;
    ep[3 * i + 1] = py * q;
    ep[3 * i + 2] = pz * q;
  } else {
    if (pin) {
      if (!LEAN && d.save_B) d.save_B[3 * i] = d.save_B[3 * i + 1] = d.save_B[3 * i + 2] = R(0);
      return;
    }
    const R x = ep[3 * i], y = ep[3 * i + 1], z = ep[3 * i + 2];
    R bx, by, bz, fx, fy, fz;
    local_field<R, CM, GEO, LEAN>(S, d, ep, st, x, y, z, step, noise, bx, by, bz);
    if (!LEAN && d.save_B) {
      d.save_B[3 * i] = bx;
      d.save_B[3 * i + 1] = by;
      d.save_B[3 * i + 2] = bz;
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
      rotate(cx, cy, cz, hdt * (k1[3 * i] + wx), hdt * (k1[3 * i + 1] + wy), hdt * (k1[3 * i + 2] + wz));
    } else {
      cx = k1[3 * i] + hdt * fx;
      cy = k1[3 * i + 1] + hdt * fy;
      cz = k1[3 * i + 2] + hdt * fz;
    }
    const R q = rnorm3(cx, cy, cz);
    e[3 * i] = cx * q;
    e[3 * i + 1] = cy * q;
    e[3 * i + 2] = cz * q;
  }
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
__global__ __launch_bounds__(BLOCK) void k_noise(int N, uint32_t seed_lo, uint32_t seed_hi, uint64_t step,
                                                 R* __restrict__ out) {
  const int r = blockIdx.x * BLOCK + threadIdx.x;
  if (r >= N) return;
  float n0, n1, n2;
  gauss3(seed_lo, seed_hi, step, uint32_t(r), n0, n1, n2);
  out[3 * r] = R(n0);
  out[3 * r + 1] = R(n1);
  out[3 * r + 2] = R(n2);
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
static int check_launch(const char* what) {
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    set_error(std::string(what) + ": " + hipGetErrorString(err));
    return MX_EHIP;
  }
  return MX_OK;
}

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
  {
    // column map of the step kernels when an axis-0 plane is a whole number of tiles, 8 or more per XCD
    const long plane = long(m.L[1]) * m.L[2];
    S.xcol = ((m.geo == GEO_PERIODIC || m.geo == GEO_OPEN) && plane % BLOCK == 0 && (plane / BLOCK) % 8 == 0)
                 ? int(plane / BLOCK / 8) : 0;
  }
  S.half = m.ncell / 2;
  S.L2h = m.L[2] / 2;
  S.rL2h = S.L2h > 0 ? 1.0f / float(S.L2h) : 0.0f;
  S.amode = m.amode;
  S.has_A = m.has_A;
  S.has_c = m.has_c;
  S.has_u = m.has_u;
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

template <typename R>
static int heun_t(const Model& m, void* e, uint64_t seed, uint64_t step, const void* noise, hipStream_t s,
                  const uint64_t* step_ptr) {
  const Static<R>* S = static_cast<const Static<R>*>(m.static_d);
  Dyn<R> d = make_dyn<R>(m, seed);
  d.step_ptr = step_ptr;
  R* E = static_cast<R*>(e);
  R* ep = static_cast<R*>(m.ep_d);
  R* k1 = static_cast<R*>(m.k1_d);
  const R* nz = static_cast<const R*>(noise);
  const dim3 g(nblk_of(m)), b(BLOCK);
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
                     const uint64_t* step_ptr) {
  return m.prec == MX_FP64 ? heun_t<double>(m, e, seed, step, noise, s, step_ptr)
                           : heun_t<float>(m, e, seed, step, noise, s, step_ptr);
}

template <typename R>
static int pass_t(const Model& m, int pass, void* e, void* ep, void* k1, uint64_t seed, uint64_t step, void* save_e,
                  void* save_B, hipStream_t s, const double* B_ext) {
  const Static<R>* S = static_cast<const Static<R>*>(m.static_d);
  Dyn<R> d = make_dyn<R>(m, seed);
  if (B_ext) {
    d.B0 = R(B_ext[0]);
    d.B1 = R(B_ext[1]);
    d.B2 = R(B_ext[2]);
  }
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
                     void* save_e, void* save_B, hipStream_t s, const double* B_ext) {
  return m.prec == MX_FP64 ? pass_t<double>(m, pass, e, ep, k1, seed, step, save_e, save_B, s, B_ext)
                           : pass_t<float>(m, pass, e, ep, k1, seed, step, save_e, save_B, s, B_ext);
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
                       static_cast<double*>(out));
  else
    hipLaunchKernelGGL((k_noise<float>), g, dim3(BLOCK), 0, s, m.N, uint32_t(seed), uint32_t(seed >> 32), step,
                       static_cast<float*>(out));
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
