// Host side of the C ABI declared in include/mxllg.h: model set-up, device buffers, argument checks and
// the step loop.  No allocation happens inside mx_llg_run / mx_field / mx_observables: every buffer
// is made when the model is configured.
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <stdexcept>
#include <tuple>

#include "mxllg_internal.h"

namespace mx {

static thread_local std::string g_err;

static const char* mx_last_error_text() { return g_err.c_str(); }

void set_error(const std::string& msg) { g_err = msg; }

struct Error {
  int code;
  std::string msg;
};

static void fail(int code, const std::string& msg) { throw Error{code, msg}; }

static void hip_ok(hipError_t e, const char* what) {
  if (e != hipSuccess) fail(e == hipErrorOutOfMemory ? MX_ENOMEM : MX_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename F>
static int guarded(F&& f) {
  try {
    return f();
  } catch (const Error& e) {
    g_err = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return MX_ENOMEM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return MX_EHIP;
  }
}

static size_t rsize(const Model& m) { return m.prec == MX_FP64 ? sizeof(double) : sizeof(float); }

// Set-up calls replace device buffers and constant blocks that kernels already queued on the caller's
// stream may still read (the stream may be non-blocking, so a copy on the null stream does not wait for
// it): the device finishes its queued work first.  Only configuration changes pay for this; the step
// loop does not.
static void quiesce() { hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize"); }

static void dfree(void*& p) {
  if (p) {
    quiesce();
    (void)hipFree(p);
  }
  p = nullptr;
}

// Device copy of n host doubles, converted to the model precision.  perm: reorder a magex-order
// per-site array of n = N * comps values (site-major) into internal component-major order.
static void upload(const Model& m, void*& dst, const double* src, size_t n) {
  quiesce();
  dfree(dst);
  hip_ok(hipMalloc(&dst, n * rsize(m)), "hipMalloc");
  if (m.prec == MX_FP64) {
    hip_ok(hipMemcpy(dst, src, n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
  } else {
    std::vector<float> tmp(src, src + n);
    hip_ok(hipMemcpy(dst, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
  }
}

// Internal cell of row-major cell c: itself, its tile-major index in the TILED layout, or its place in
// its parity half in the CHECKER layout.
static inline int internal_cell(const Model& m, int c) {
  if (m.geo != GEO_TILED && m.geo != GEO_CHECKER) return c;
  const int c2 = c % m.L[2], t = c / m.L[2], c1 = t % m.L[1], c0 = t / m.L[1];
  if (m.geo == GEO_CHECKER) return ((c0 + c1 + c2) & 1) * (m.ncell / 2) + t * (m.L[2] / 2) + (c2 >> 1);
  const int b0 = m.tb[0], b1 = m.tb[1], b2 = m.tb[2];
  const int tile = ((c0 >> b0) * (m.L[1] >> b1) + (c1 >> b1)) * (m.L[2] >> b2) + (c2 >> b2);
  const int l = ((c0 & ((1 << b0) - 1)) << (b1 + b2)) | ((c1 & ((1 << b1) - 1)) << b2) | (c2 & ((1 << b2) - 1));
  return tile * BLOCK + l;
}

int cell_to_internal(const Model& m, int c) { return internal_cell(m, c); }

// magex order site r = c * ns + s -> internal order q = s * ncell + internal_cell(c)
static inline int internal_index(const Model& m, int r) {
  const int c = r / m.ns, s = r - c * m.ns;
  return s * m.ncell + internal_cell(m, c);
}

// Tile sides 2^tb[a] with tb0 + tb1 + tb2 = log2(BLOCK), each side dividing L; bits are dealt one at a
// time to the axis with the fewest so far (ties to the faster axis), so tiles come out near-cubic.
static bool choose_tiles(const Model& m, int tb[3]) {
  if (!(m.pbc[0] && m.pbc[1] && m.pbc[2]) || m.ncell < 4 * BLOCK) return false;
  int tz[3];
  for (int a = 0; a < 3; ++a) {
    tz[a] = 0;
    while (tz[a] < 8 && (m.L[a] % (2 << tz[a])) == 0) ++tz[a];
    tb[a] = 0;
  }
  int bits = 0;
  while ((1 << bits) < BLOCK) ++bits;
  for (int k = 0; k < bits; ++k) {
    int best = -1;
    for (int a = 2; a >= 0; --a)
      if (tb[a] < tz[a] && (best < 0 || tb[a] < tb[best])) best = a;
    if (best < 0) return false;
    ++tb[best];
  }
  return true;
}

// Fix the cell layout (want: 0 row-major, 1 tiled, 2 checker; each where the supercell allows it, else
// row-major) and upload perm[q] = magex site of internal site q.
static void set_layout(Model& m, int want_tiled) {
  int tb[3];
  const bool allp = m.pbc[0] && m.pbc[1] && m.pbc[2];
  if (want_tiled == 2 && allp && m.L[2] % 2 == 0) {
    m.geo = GEO_CHECKER;
    m.tb[0] = m.tb[1] = m.tb[2] = 0;
  } else if (want_tiled == 1 && choose_tiles(m, tb)) {
    m.geo = GEO_TILED;
    for (int a = 0; a < 3; ++a) m.tb[a] = tb[a];
  } else {
    m.geo = allp ? GEO_PERIODIC : GEO_OPEN;
    m.tb[0] = m.tb[1] = m.tb[2] = 0;
  }
  std::vector<int32_t> perm(m.N);
  for (int r = 0; r < m.N; ++r) perm[internal_index(m, r)] = r;
  if (!m.perm_d) hip_ok(hipMalloc(reinterpret_cast<void**>(&m.perm_d), sizeof(int32_t) * size_t(m.N)), "hipMalloc perm");
  hip_ok(hipMemcpy(m.perm_d, perm.data(), sizeof(int32_t) * size_t(m.N), hipMemcpyHostToDevice), "hipMemcpy perm");
  m.want_tiled = want_tiled;
  m.static_dirty = 1;
}

// per-site magex-order array (N x comps, site-major) -> internal order, component-major (comps x N)
static std::vector<double> to_internal_soa(const Model& m, const double* src, int comps) {
  std::vector<double> out(size_t(m.N) * comps);
  for (int r = 0; r < m.N; ++r) {
    const int q = internal_index(m, r);
    for (int k = 0; k < comps; ++k) out[size_t(k) * m.N + q] = src[size_t(r) * comps + k];
  }
  return out;
}

static void ensure_scratch(Model& m) {
  if (!m.ep_d) hip_ok(hipMalloc(&m.ep_d, size_t(m.N) * 3 * rsize(m)), "hipMalloc ep");
  if (!m.k1_d) hip_ok(hipMalloc(&m.k1_d, size_t(m.N) * 3 * rsize(m)), "hipMalloc k1");
  if (needs_permute(m) && !m.eint_d) hip_ok(hipMalloc(&m.eint_d, size_t(m.N) * 3 * rsize(m)), "hipMalloc e");
  if (!m.stepsync_d) {
    const size_t n = stepsync_words(m) * sizeof(uint32_t);
    hip_ok(hipMalloc(reinterpret_cast<void**>(&m.stepsync_d), n), "hipMalloc step control words");
    hip_ok(hipMemset(m.stepsync_d, 0, n), "hipMemset step control words");
  }
  if (!m.rec_d && persistent_steps(m)) {
    const size_t n = rec_words(m) * sizeof(uint64_t);
    hip_ok(hipMalloc(reinterpret_cast<void**>(&m.rec_d), n), "hipMalloc tagged spin records");
    hip_ok(hipMemset(m.rec_d, 0, n), "hipMemset tagged spin records");
  }
  if (!m.part_d) {
    const int nblk = m.ns * ((m.ncell + BLOCK - 1) / BLOCK);
    m.npart = std::max(nblk, (2 * m.N + BLOCK - 1) / BLOCK + 1);
    hip_ok(hipMalloc(reinterpret_cast<void**>(&m.part_d), size_t(m.npart) * MX_OBS_LEN * sizeof(double)),
           "hipMalloc partials");
  }
}

int ensure_step_scratch(Model& m) {
  return guarded([&]() -> int {
    ensure_scratch(m);
    return MX_OK;
  });
}

static void use_device(const Model& m) { hip_ok(hipSetDevice(m.device), "hipSetDevice"); }

int sync_model_static(Model& m) { return m.static_dirty ? upload_static(m) : MX_OK; }
size_t prec_size(const Model& m) { return rsize(m); }

// Bring the device constant block up to date (synchronous copy after the device's queued work, which may
// still read the block; only after configuration changes).
static void sync_static(Model& m) {
  if (!m.static_dirty) return;
  quiesce();
  const int rc = upload_static(m);
  if (rc) fail(rc, "model constants: " + std::string(mx_last_error_text()));
}

static void need_model(const mx_model* m) {
  if (!m) fail(MX_EINVAL, "null model");
}

static void recompute_sigma(Model& m);

static void drop_graph(Model& m) {
  if (m.gexec) (void)hipGraphExecDestroy(m.gexec);
  m.gexec = nullptr;
  m.gkey.clear();
}

template <typename T>
static void put(std::vector<char>& k, const T& v) {
  const char* p = reinterpret_cast<const char*>(&v);
  k.insert(k.end(), p, p + sizeof(T));
}

// Capture (once per configuration) graph_steps Heun steps plus the counter advance on the model's
// own stream; everything a kernel argument bakes in is part of the key.
static void ensure_graph(Model& m, void* E, uint64_t seed) {
  std::vector<char> key;
  for (const void* p : {E, m.ep_d, m.k1_d, m.static_d, m.bond_d, m.A_d, m.Kc_d, m.K2_d, m.K3_d, m.inv_mu_d,
                        m.alpha_d, m.sigma_d, m.sign_d, m.pinned_d})
    put(key, p);
  put(key, seed);
  for (double b : m.B) put(key, b);
  for (double v : {m.sot_p[0], m.sot_p[1], m.sot_p[2], m.sot_dl, m.sot_fl}) put(key, v);
  put(key, m.dt);
  for (int v : {m.precession, m.noisy, m.integrator, m.coup, m.pbc[0], m.pbc[1], m.pbc[2], m.graph_steps, m.prec,
                m.step_mode})
    put(key, v);
  if (m.gexec && key == m.gkey) return;
  drop_graph(m);
  if (!m.counter_d) hip_ok(hipMalloc(reinterpret_cast<void**>(&m.counter_d), sizeof(uint64_t)), "hipMalloc counter");
  if (!m.cap_stream) hip_ok(hipStreamCreateWithFlags(&m.cap_stream, hipStreamNonBlocking), "hipStreamCreate");
  hip_ok(hipStreamBeginCapture(m.cap_stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  int rc = MX_OK;
  for (int n = 0; n < m.graph_steps && rc == MX_OK; ++n)
    rc = launch_heun_step(m, E, seed, uint64_t(n), nullptr, m.cap_stream, m.counter_d);
  if (rc == MX_OK) rc = launch_counter(m.counter_d, uint64_t(m.graph_steps), 1, m.cap_stream);
  hipGraph_t g = nullptr;
  const hipError_t end = hipStreamEndCapture(m.cap_stream, &g);
  if (rc != MX_OK) {
    if (g) (void)hipGraphDestroy(g);
    fail(rc, "capturing the step loop: " + g_err);
  }
  hip_ok(end, "hipStreamEndCapture");
  const hipError_t inst = hipGraphInstantiate(&m.gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  hip_ok(inst, "hipGraphInstantiate");
  m.gkey = key;
}

// ---------------------------------------------------------------------- allocation-free step scratch
// The scratch of mx_llg_run_noalloc: one call's predictor buffers, internal-order spin copy, one-launch
// control words and tagged spin records, in a file-scope device array that every GPU the library runs on
// holds until the process ends.  In host code `scratch` is only a stand-in symbol: its device address on each GPU
// comes from hipGetSymbolAddress (once per device, that device current) and only that address is used.
constexpr size_t SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];
constexpr int SCRATCH_DEVICES = 64;
static void* g_scratch_ptr[SCRATCH_DEVICES] = {};
static std::mutex g_scratch_mu;

static unsigned char* scratch_of(int device) {
  if (device < 0 || device >= SCRATCH_DEVICES) fail(MX_EUNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (err != hipSuccess || !p) fail(MX_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(err));
    g_scratch_ptr[device] = p;
  }
  return static_cast<unsigned char*>(g_scratch_ptr[device]);
}

static size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

}  // namespace mx

using namespace mx;

extern "C" {

const char* mx_last_error(void) { return g_err.c_str(); }

int mx_version(void) { return 1; }

int mx_model_create(int precision, int nsub, const int32_t L[3], const int32_t pbc[3], mx_model** out) {
  return guarded([&]() -> int {
    if (!out || !L || !pbc) fail(MX_EINVAL, "null argument");
    *out = nullptr;
    if (precision != MX_FP64 && precision != MX_FP32) fail(MX_EINVAL, "precision must be MX_FP64 or MX_FP32");
    if (nsub < 1 || nsub > MAXSUB) fail(MX_EUNSUPPORTED, "1 to " + std::to_string(MAXSUB) + " sites per cell");
    int64_t n = nsub;
    for (int a = 0; a < 3; ++a) {
      if (L[a] < 1 || L[a] > 32767) fail(MX_EINVAL, "supercell extents must lie in [1, 32767]");
      n *= L[a];
    }
    if (n > MAX_SITES)
      fail(MX_EUNSUPPORTED, "more than " + std::to_string(MAX_SITES) +
                                " sites (per-site offsets up to 9 N are 32-bit in the kernels)");
    auto* m = new mx_model();
    m->prec = precision;
    m->noise_dbl = precision == MX_FP64 ? 1 : 0;
    hip_ok(hipGetDevice(&m->device), "hipGetDevice");
    m->ns = nsub;
    for (int a = 0; a < 3; ++a) {
      m->L[a] = L[a];
      m->pbc[a] = pbc[a] ? 1 : 0;
    }
    m->ncell = L[0] * L[1] * L[2];
    m->N = int(n);
    m->sptr.assign(nsub + 1, 0);
    m->Asub.assign(9 * nsub, 0.0);
    m->mu_sub.assign(nsub, 2.0);
    m->sign_sub.assign(nsub, 1.0);
    m->alpha_h.assign(1, 0.1);
    m->T_h.assign(1, 0.0);
    m->table.assign(1, 0.0);
    m->width = 1;
    m->ntypes = 1;
    try {
      set_layout(*m, 0);
    } catch (...) {
      mx_model_destroy(m);
      throw;
    }
    *out = m;
    return MX_OK;
  });
}

void mx_model_destroy(mx_model* m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  for (void** p : {&m->bond_d, &m->A_d, &m->Kc_d, &m->K2_d, &m->K3_d, &m->inv_mu_d, &m->sign_d, &m->pinned_d,
                   &m->alpha_d, &m->sigma_d, &m->ep_d, &m->k1_d, &m->eint_d, &m->static_d, &m->mc_beta_d,
                   reinterpret_cast<void**>(&m->mc_sites_d), reinterpret_cast<void**>(&m->mc_acc_d), &m->mc_rep_d,
                   reinterpret_cast<void**>(&m->mc_repacc_d), reinterpret_cast<void**>(&m->img_part_d), &m->mc_repe_d})
    dfree(*p);
  if (m->perm_d) (void)hipFree(m->perm_d);
  if (m->stepsync_d) (void)hipFree(m->stepsync_d);
  if (m->rec_d) (void)hipFree(m->rec_d);
  if (m->part_d) (void)hipFree(m->part_d);
  drop_graph(*m);
  if (m->counter_d) (void)hipFree(m->counter_d);
  if (m->cap_stream) (void)hipStreamDestroy(m->cap_stream);
  delete m;
}

int mx_model_info(const mx_model* m, int64_t* n_sites, int32_t* precision, int32_t* nsub) {
  return guarded([&]() -> int {
    need_model(m);
    if (n_sites) *n_sites = m->N;
    if (precision) *precision = m->prec;
    if (nsub) *nsub = m->ns;
    return MX_OK;
  });
}

int mx_model_set_stencil(mx_model* m, int32_t nent, const int32_t* ent, int32_t ntypes, int32_t width,
                         const double* table) {
  return guarded([&]() -> int {
    need_model(m);
    if (nent < 0 || (nent > 0 && !ent)) fail(MX_EINVAL, "bad stencil");
    if (nent > MAXENT) fail(MX_EUNSUPPORTED, "more than " + std::to_string(MAXENT) + " stencil entries");
    if (width != 1 && width != 4 && width != 9) fail(MX_EINVAL, "coupling width must be 1, 4 or 9");
    if (ntypes < 1 || !table) fail(MX_EINVAL, "empty coupling table");
    if (int64_t(ntypes) * width > MAXTAB)
      fail(MX_EUNSUPPORTED, "coupling table larger than " + std::to_string(MAXTAB) + " values");
    // group entries by source sublattice; inside a group, bonds to other sublattices come first, ordered
    // by their pair (the entry and its reverse partner share a key), so the two workgroups that read a
    // per-bond tensor from either end reach it at the same point of their loops and the second read
    // finds it in L2; same-sublattice bonds follow in the caller's order (mx_model_set_bond_tensors
    // pairs them as well)
    std::vector<int> pair_key(nent);
    for (int k = 0; k < nent; ++k) {
      pair_key[k] = k;
      for (int q = 0; q < nent; ++q)
        if (ent[6 * q] == ent[6 * k + 1] && ent[6 * q + 1] == ent[6 * k] && ent[6 * q + 2] == -ent[6 * k + 2] &&
            ent[6 * q + 3] == -ent[6 * k + 3] && ent[6 * q + 4] == -ent[6 * k + 4])
          pair_key[k] = std::min(pair_key[k], q);
    }
    std::vector<int> order;
    for (int s = 0; s < m->ns; ++s) {
      std::vector<int> grp;
      for (int k = 0; k < nent; ++k)
        if (ent[6 * k] == s) grp.push_back(k);
      std::stable_sort(grp.begin(), grp.end(), [&](int a, int b) {
        const bool sa = ent[6 * a + 1] == s, sb = ent[6 * b + 1] == s;
        if (sa != sb) return !sa;
        return sa ? a < b : pair_key[a] < pair_key[b];
      });
      order.insert(order.end(), grp.begin(), grp.end());
    }
    if (int(order.size()) != nent) fail(MX_EINVAL, "stencil source sublattice out of range");
    m->ent.clear();
    m->sptr.assign(m->ns + 1, 0);
    for (int k : order) {
      const int32_t* e = ent + 6 * k;
      const int s = e[0], t = e[1], typ = e[5];
      if (t < 0 || t >= m->ns) fail(MX_EINVAL, "stencil target sublattice out of range");
      if (typ < 0 || typ >= ntypes) fail(MX_EINVAL, "stencil coupling type out of range");
      int d[3];
      bool self = (s == t);
      for (int a = 0; a < 3; ++a) {
        const int R = e[2 + a];
        if (m->pbc[a]) {
          d[a] = ((R % m->L[a]) + m->L[a]) % m->L[a];
          self = self && d[a] == 0;
        } else {
          if (R < -32767 || R > 32767) fail(MX_EINVAL, "stencil offset out of range");
          d[a] = R;  // |R| >= L on an open axis never finds a target: the kernel's bounds check drops it
          self = self && R == 0;
        }
      }
      if (self) fail(MX_EINVAL, "stencil entry couples a site to itself: fold it into the on-site tensor");
      m->ent.push_back(Ent{typ, t, (d[0] & 0xffff) | (d[1] << 16), d[2]});
      m->sptr[s + 1]++;
    }
    for (int s = 0; s < m->ns; ++s) m->sptr[s + 1] += m->sptr[s];
    m->ent_raw.assign(ent, ent + 6 * size_t(nent));
    m->ent_order = order;
    m->nent = nent;
    m->ntypes = ntypes;
    m->width = width;
    m->table.assign(table, table + size_t(ntypes) * width);
    m->coup = width == 1 ? COUP_TABLE1 : width == 4 ? COUP_TABLE4 : COUP_TABLE9;
    dfree(m->bond_d);
    m->nfwd = 0;
    m->static_dirty = 1;
    return MX_OK;
  });
}

int mx_model_set_bond_tensors(mx_model* m, const double* J) {
  return guarded([&]() -> int {
    need_model(m);
    m->layout_frozen = 1;
    if (!J) fail(MX_EINVAL, "null tensors");
    if (m->nent == 0) fail(MX_EINVAL, "set the stencil first");
    use_device(*m);
    // forward flag and partner of every entry (in the caller's order)
    const int n = m->nent;
    auto raw = [&](int k, int q) { return m->ent_raw[6 * size_t(k) + q]; };
    std::vector<int> fwd_slot(n, -1), partner(n, -1);
    int nfwd = 0;
    for (int k = 0; k < n; ++k) {
      const int s = raw(k, 0), t = raw(k, 1);
      const std::tuple<int, int, int> R{raw(k, 2), raw(k, 3), raw(k, 4)};
      const bool f = s < t || (s == t && R > std::tuple<int, int, int>{0, 0, 0});
      if (f) fwd_slot[k] = nfwd++;
      for (int q = 0; q < n; ++q)
        if (raw(q, 0) == t && raw(q, 1) == s && raw(q, 2) == -raw(k, 2) && raw(q, 3) == -raw(k, 3) &&
            raw(q, 4) == -raw(k, 4))
          partner[k] = q;
      if (partner[k] < 0) fail(MX_EINVAL, "stencil entry without a reverse partner: per-bond storage needs J_ji");
    }
    // per-bond tensors: a same-sublattice bond's two entries side by side (in the order of the pair's
    // first entry), so that its owner and its partner read the tensor in the same batch -- FP64 bcc 64^3
    // 815.9 -> 768.2 MB per Heun step.  Table couplings keep the caller's order there: paired, the
    // persistent FP32 table steps measured slower (triangular 256^2 8.80 -> 9.11 us per step).
    for (int s = 0; s < m->ns; ++s) {
      std::vector<int> pos;
      for (int p = m->sptr[s]; p < m->sptr[s + 1]; ++p)
        if (raw(m->ent_order[p], 1) == s) pos.push_back(p);
      std::vector<int> ks;
      for (int p : pos) ks.push_back(m->ent_order[p]);
      std::stable_sort(ks.begin(), ks.end(),
                       [&](int a, int b) { return std::min(a, partner[a]) < std::min(b, partner[b]); });
      std::vector<Ent> es;
      for (int k : ks)
        for (int p : pos)
          if (m->ent_order[p] == k) es.push_back(m->ent[p]);
      for (size_t q = 0; q < pos.size(); ++q) {
        m->ent_order[pos[q]] = ks[q];
        m->ent[pos[q]] = es[q];
      }
    }
    // entries (internal order) point at a forward slot: their own, or the partner's (back)
    for (size_t pos = 0; pos < m->ent.size(); ++pos) {
      const int k = m->ent_order[pos];
      if (fwd_slot[k] >= 0) {
        m->ent[pos].slot = fwd_slot[k];
        m->ent[pos].tb &= 0xffff;
      } else {
        m->ent[pos].slot = fwd_slot[partner[k]];
        m->ent[pos].tb = (m->ent[pos].tb & 0xffff) | (1 << 16);
      }
    }
    // J arrives as [f][magex cell][9]; the device holds [f][internal cell][9]
    std::vector<double> buf(size_t(nfwd) * m->ncell * 9);
    for (int f = 0; f < nfwd; ++f)
      for (int c = 0; c < m->ncell; ++c) {
        const double* src = J + (size_t(f) * m->ncell + c) * 9;
        double* dst = buf.data() + (size_t(f) * m->ncell + internal_cell(*m, c)) * 9;
        for (int q = 0; q < 9; ++q) dst[q] = src[q];
      }
    upload(*m, m->bond_d, buf.data(), buf.size());
    m->nfwd = nfwd;
    m->coup = COUP_BOND;
    m->static_dirty = 1;
    return MX_OK;
  });
}

int mx_model_set_onsite(mx_model* m, int per_site, const double* A) {
  return guarded([&]() -> int {
    need_model(m);
    m->layout_frozen = 1;
    use_device(*m);
    dfree(m->A_d);
    m->amode = ONSITE_SUB;
    m->has_A = 0;
    std::fill(m->Asub.begin(), m->Asub.end(), 0.0);
    if (!A) {
      m->static_dirty = 1;
      return MX_OK;
    }
    if (!per_site) {
      for (int k = 0; k < 9 * m->ns; ++k) {
        m->Asub[k] = A[k];
        if (A[k] != 0.0) m->has_A = 1;
      }
      m->static_dirty = 1;
    return MX_OK;
    }
    bool sym = true, any = false;
    for (int r = 0; r < m->N && sym; ++r) {
      const double* a = A + 9 * size_t(r);
      sym = a[1] == a[3] && a[2] == a[6] && a[5] == a[7];
    }
    for (size_t k = 0; k < size_t(m->N) * 9 && !any; ++k) any = A[k] != 0.0;
    if (!any) {
      m->static_dirty = 1;
      return MX_OK;
    }
    m->has_A = 1;
    if (sym) {
      std::vector<double> six(size_t(m->N) * 6);
      for (int r = 0; r < m->N; ++r) {
        const double* a = A + 9 * size_t(r);
        const double v[6] = {a[0], a[4], a[8], a[1], a[2], a[5]};
        for (int k = 0; k < 6; ++k) six[size_t(r) * 6 + k] = v[k];
      }
      const std::vector<double> soa = to_internal_soa(*m, six.data(), 6);
      upload(*m, m->A_d, soa.data(), soa.size());
      m->amode = ONSITE_SITE6;
    } else {
      const std::vector<double> soa = to_internal_soa(*m, A, 9);
      upload(*m, m->A_d, soa.data(), soa.size());
      m->amode = ONSITE_SITE9;
    }
    m->static_dirty = 1;
    return MX_OK;
  });
}

int mx_model_set_cubic(mx_model* m, const double* Kc, const double C[9]) {
  return guarded([&]() -> int {
    need_model(m);
    m->layout_frozen = 1;
    use_device(*m);
    dfree(m->Kc_d);
    m->has_c = 0;
    if (!Kc) {
      m->static_dirty = 1;
      return MX_OK;
    }
    if (!C) fail(MX_EINVAL, "cubic axes missing");
    const std::vector<double> v = to_internal_soa(*m, Kc, 1);
    upload(*m, m->Kc_d, v.data(), v.size());
    for (int q = 0; q < 9; ++q) m->C[q] = C[q];
    m->has_c = 1;
    m->static_dirty = 1;
    return MX_OK;
  });
}

int mx_model_set_uniaxial_higher(mx_model* m, const double* K2, const double* K3, const double u[3]) {
  return guarded([&]() -> int {
    need_model(m);
    m->layout_frozen = 1;
    use_device(*m);
    dfree(m->K2_d);
    dfree(m->K3_d);
    m->has_u = 0;
    if (!K2) {
      m->static_dirty = 1;
      return MX_OK;
    }
    if (!u) fail(MX_EINVAL, "uniaxial axis missing");
    const std::vector<double> zeros(m->N, 0.0);
    const std::vector<double> a = to_internal_soa(*m, K2, 1);
    const std::vector<double> b = to_internal_soa(*m, K3 ? K3 : zeros.data(), 1);
    upload(*m, m->K2_d, a.data(), a.size());
    upload(*m, m->K3_d, b.data(), b.size());
    for (int q = 0; q < 3; ++q) m->u[q] = u[q];
    m->has_u = 1;
    m->static_dirty = 1;
    return MX_OK;
  });
}

int mx_model_set_moments(mx_model* m, int per_site, const double* mu, const double* signs) {
  return guarded([&]() -> int {
    need_model(m);
    m->layout_frozen = 1;
    if (!mu) fail(MX_EINVAL, "null moments");
    use_device(*m);
    dfree(m->inv_mu_d);
    dfree(m->sign_d);
    m->mu_site.clear();
    m->sign_site_h.clear();
    if (!per_site) {
      for (int s = 0; s < m->ns; ++s) {
        if (!(mu[s] > 0)) fail(MX_EINVAL, "moments must be positive");
        m->mu_sub[s] = mu[s];
        m->sign_sub[s] = signs ? signs[s] : 1.0;
      }
    } else {
      m->mu_site.assign(mu, mu + m->N);
      std::vector<double> inv(m->N);
      for (int r = 0; r < m->N; ++r) {
        if (!(mu[r] > 0)) fail(MX_EINVAL, "moments must be positive");
        inv[r] = 1.0 / (mu[r] * MU_B);
      }
      const std::vector<double> a = to_internal_soa(*m, inv.data(), 1);
      upload(*m, m->inv_mu_d, a.data(), a.size());
      if (signs) {
        m->sign_site_h.assign(signs, signs + m->N);
        const std::vector<double> b = to_internal_soa(*m, signs, 1);
        upload(*m, m->sign_d, b.data(), b.size());
      }
    }
    if (m->dyn_set) recompute_sigma(*m);
    m->static_dirty = 1;
    return MX_OK;
  });
}

int mx_model_set_pinned(mx_model* m, const uint8_t* pinned) {
  return guarded([&]() -> int {
    need_model(m);
    m->layout_frozen = 1;
    use_device(*m);
    dfree(m->pinned_d);
    if (!pinned) {
      m->static_dirty = 1;
      return MX_OK;
    }
    std::vector<uint8_t> v(m->N);
    bool any = false;
    for (int r = 0; r < m->N; ++r) {
      v[internal_index(*m, r)] = pinned[r] ? 1 : 0;
      any = any || pinned[r];
    }
    if (!any) {
      m->static_dirty = 1;
      return MX_OK;
    }
    quiesce();
    hip_ok(hipMalloc(&m->pinned_d, m->N), "hipMalloc pinned");
    hip_ok(hipMemcpy(m->pinned_d, v.data(), m->N, hipMemcpyHostToDevice), "hipMemcpy pinned");
    m->static_dirty = 1;
    return MX_OK;
  });
}

int mx_model_set_field(mx_model* m, const double B[3]) {
  return guarded([&]() -> int {
    need_model(m);
    if (!B) fail(MX_EINVAL, "null field");
    for (int a = 0; a < 3; ++a) m->B[a] = B[a];
    return MX_OK;
  });
}

namespace mx {
// sigma_i = sqrt(2 alpha_i k_B T_i / (gamma mu_i mu_B dt)) (magex/llg.py:111, fast path llg.py:181-182)
static void recompute_sigma(Model& m) {
  const bool site_alpha = m.alpha_h.size() > 1, site_T = m.T_h.size() > 1, site_mu = !m.mu_site.empty();
  dfree(m.alpha_d);
  dfree(m.sigma_d);
  m.noisy = 0;
  for (double t : m.T_h) m.noisy = m.noisy || t > 0;
  if (site_T) m.noisy = 1;  // a site-temperature map always takes the noisy path (llg.py:175)
  m.alpha_sub_v.assign(m.ns, 0.0);
  m.sigma_sub_v.assign(m.ns, 0.0);
  if (!site_alpha && !site_T && !site_mu) {
    for (int s = 0; s < m.ns; ++s) {
      m.alpha_sub_v[s] = m.alpha_h[0];
      m.sigma_sub_v[s] = std::sqrt(2.0 * m.alpha_h[0] * K_B * m.T_h[0] / (GAMMA * m.mu_sub[s] * MU_B * m.dt));
    }
    return;
  }
  std::vector<double> al(m.N), sg(m.N);
  for (int r = 0; r < m.N; ++r) {
    const double a = site_alpha ? m.alpha_h[r] : m.alpha_h[0];
    const double t = site_T ? m.T_h[r] : m.T_h[0];
    const double mu = site_mu ? m.mu_site[r] : m.mu_sub[r % m.ns];
    al[r] = a;
    sg[r] = std::sqrt(2.0 * a * K_B * t / (GAMMA * mu * MU_B * m.dt));
  }
  const std::vector<double> a2 = to_internal_soa(m, al.data(), 1);
  const std::vector<double> s2 = to_internal_soa(m, sg.data(), 1);
  upload(m, m.alpha_d, a2.data(), a2.size());
  upload(m, m.sigma_d, s2.data(), s2.size());
}
}  // namespace mx

int mx_model_set_dynamics(mx_model* m, double dt, int alpha_per_site, const double* alpha, int T_per_site,
                          const double* T, int precession, int integrator) {
  return guarded([&]() -> int {
    need_model(m);
    m->layout_frozen = 1;
    if (!alpha || !T) fail(MX_EINVAL, "null damping or temperature");
    if (!(dt > 0)) fail(MX_EINVAL, "dt must be positive");
    if (integrator != MX_HEUN && integrator != MX_DEPONDT) fail(MX_EINVAL, "integrator must be MX_HEUN or MX_DEPONDT");
    use_device(*m);
    m->dt = dt;
    m->alpha_h.assign(alpha, alpha + (alpha_per_site ? m->N : 1));
    m->T_h.assign(T, T + (T_per_site ? m->N : 1));
    for (double t : m->T_h)
      if (t < 0 || std::isnan(t)) fail(MX_EINVAL, "temperatures must be >= 0");
    m->precession = precession ? 1 : 0;
    m->integrator = integrator;
    recompute_sigma(*m);
    m->dyn_set = 1;
    ensure_scratch(*m);
    m->static_dirty = 1;
    return MX_OK;
  });
}

int mx_model_set_layout(mx_model* m, int tiled) {
  return guarded([&]() -> int {
    need_model(m);
    if (m->layout_frozen) fail(MX_EINVAL, "the cell layout is fixed once site data has been set");
    if (tiled < 0 || tiled > 2) fail(MX_EINVAL, "layout must be 0 (row-major), 1 (tiled) or 2 (checker)");
    use_device(*m);
    set_layout(*m, tiled);
    return MX_OK;
  });
}

int mx_model_layout(const mx_model* m, int32_t tile[3]) {
  return guarded([&]() -> int {
    need_model(m);
    for (int a = 0; a < 3; ++a) tile[a] = m->geo == GEO_TILED ? (1 << m->tb[a]) : (m->geo == GEO_CHECKER ? -1 : 0);
    return MX_OK;
  });
}

int mx_model_set_graph_steps(mx_model* m, int32_t steps) {
  return guarded([&]() -> int {
    need_model(m);
    if (steps < 0 || steps > 4096) fail(MX_EINVAL, "graph_steps must lie in [0, 4096]");
    if (steps != m->graph_steps) {
      use_device(*m);
      drop_graph(*m);
    }
    m->graph_steps = steps;
    return MX_OK;
  });
}

int mx_model_set_noise(mx_model* m, int32_t grade) {
  return guarded([&]() -> int {
    need_model(m);
    if (grade != MX_NOISE_FLOAT && grade != MX_NOISE_DOUBLE)
      fail(MX_EINVAL, "noise grade must be MX_NOISE_FLOAT or MX_NOISE_DOUBLE");
    if (grade != m->noise_dbl) {
      m->noise_dbl = grade;
      m->static_dirty = 1;
      use_device(*m);
      drop_graph(*m);
    }
    return MX_OK;
  });
}

int mx_model_set_wait_limit(mx_model* m, int32_t polls) {
  return guarded([&]() -> int {
    need_model(m);
    if (polls < 0) fail(MX_EINVAL, "the wait limit is a number of polls, 0 or more");
    if (uint32_t(polls) != m->wait_polls) {
      use_device(*m);
      drop_graph(*m);  // a captured step loop bakes the limit into its launches
    }
    m->wait_polls = uint32_t(polls);
    return MX_OK;
  });
}

int mx_model_set_step_mode(mx_model* m, int32_t mode) {
  return guarded([&]() -> int {
    need_model(m);
    if (mode < 0 || mode > 3)
      fail(MX_EINVAL,
           "step mode must be 0 (two launches), 1 (one launch per step), 2 (persistent runs) or 3 (as 2, and the "
           "walked one-launch step on large FP32 per-bond grids)");
    if (mode != m->step_mode) {
      use_device(*m);
      drop_graph(*m);
    }
    m->step_mode = mode;
    return MX_OK;
  });
}

int mx_model_step_info(mx_model* m, int32_t* one_launch, int64_t* given_up) {
  return guarded([&]() -> int {
    need_model(m);
    use_device(*m);
    if (one_launch) *one_launch = persistent_steps(*m) ? 2 : (fused_step(*m) ? 1 : (walk_step(*m) ? 3 : 0));
    if (given_up) {
      uint32_t v = 0;
      if (m->stepsync_d) {
        hip_ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
        hip_ok(hipMemcpy(&v, m->stepsync_d + stepsync_ctl(*m) + 2, sizeof(v), hipMemcpyDeviceToHost),
               "hipMemcpy");
      }
      *given_up = int64_t(v);
    }
    return MX_OK;
  });
}

int mx_llg_run(mx_model* m, void* e, int64_t nsteps, uint64_t seed, uint64_t step0, const void* noise,
               void* stream) {
  return guarded([&]() -> int {
    need_model(m);
    if (!e) fail(MX_EINVAL, "null spins");
    if (nsteps < 0) fail(MX_EINVAL, "negative step count");
    if (!m->dyn_set) fail(MX_EINVAL, "set the dynamics first");
    use_device(*m);
    ensure_scratch(*m);
    sync_static(*m);
    hipStream_t s = static_cast<hipStream_t>(stream);
    void* E = e;
    if (needs_permute(*m)) {
      int rc = launch_permute(*m, e, m->eint_d, 1, s);
      if (rc) return rc;
      E = m->eint_d;
    }
    const size_t stride = size_t(m->N) * 3 * rsize(*m);
    int64_t n = 0;
    const int G = m->graph_steps;
    if (nsteps >= 2 && persistent_steps(*m)) {
      // the whole run in persistent launches (every workgroup resident, couplings held across steps)
      int rc = launch_heun_steps(*m, E, seed, step0, nsteps, noise, s);
      if (rc) return rc;
      n = nsteps;
    } else if (!noise && G > 0 && nsteps >= 2 * int64_t(G)) {
      // replay a captured loop of G steps; the kernels take the step number from a device counter
      ensure_graph(*m, E, seed);
      int rc = launch_counter(m->counter_d, step0, 0, s);
      if (rc) return rc;
      for (; n + G <= nsteps; n += G) hip_ok(hipGraphLaunch(m->gexec, s), "hipGraphLaunch");
    }
    for (; n < nsteps; ++n) {
      const void* nz = noise ? static_cast<const char*>(noise) + size_t(n) * stride : nullptr;
      int rc = launch_heun_step(*m, E, seed, step0 + uint64_t(n), nz, s);
      if (rc) return rc;
    }
    if (needs_permute(*m)) return launch_permute(*m, m->eint_d, e, 0, s);
    return MX_OK;
  });
}

int mx_llg_run_noalloc(mx_model* m, void* e, int64_t nsteps, uint64_t seed, uint64_t step0, const void* noise,
                       void* stream) {
  return guarded([&]() -> int {
    need_model(m);
    if (!e) fail(MX_EINVAL, "null spins");
    if (nsteps < 0) fail(MX_EINVAL, "negative step count");
    if (!m->dyn_set) fail(MX_EINVAL, "set the dynamics first");
    if (!m->static_d || m->static_dirty)
      fail(MX_EINVAL, "model constants not on the device: run any other entry point once after configuring");
    const size_t spins = align256(size_t(m->N) * 3 * rsize(*m));
    const size_t words = stepsync_words(*m);
    const bool persistent = nsteps >= 2 && persistent_steps(*m);
    const size_t recs = persistent ? rec_words(*m) : 0;
    const size_t need = 3 * spins + align256(words * sizeof(uint32_t)) + recs * sizeof(uint64_t);
    if (need > SCRATCH_BYTES)
      fail(MX_EUNSUPPORTED, "model too large for the allocation-free scratch (" + std::to_string(need) + " bytes of " +
                                std::to_string(SCRATCH_BYTES) + "): use mx_llg_run");
    use_device(*m);
    unsigned char* base = scratch_of(m->device);
    StepBufs bufs{base, base + spins, reinterpret_cast<uint32_t*>(base + 3 * spins),
                  reinterpret_cast<uint64_t*>(base + 3 * spins + align256(words * sizeof(uint32_t)))};
    void* eint = base + 2 * spins;
    hipStream_t s = static_cast<hipStream_t>(stream);
    // the control words and the tagged records (which other models' calls may have left) start every
    // call at zero (a kernel, so the call can sit in captured work)
    int rc = launch_fill_u32(bufs.sync, words, s);
    if (rc) return rc;
    if (recs) {
      rc = launch_fill_u32(reinterpret_cast<uint32_t*>(bufs.rec), 2 * recs, s);
      if (rc) return rc;
    }
    void* E = e;
    if (needs_permute(*m)) {
      rc = launch_permute(*m, e, eint, 1, s);
      if (rc) return rc;
      E = eint;
    }
    const size_t stride = size_t(m->N) * 3 * rsize(*m);
    if (persistent) {
      rc = launch_heun_steps(*m, E, seed, step0, nsteps, noise, s, &bufs);
      if (rc) return rc;
      nsteps = 0;
    }
    for (int64_t n = 0; n < nsteps; ++n) {
      const void* nz = noise ? static_cast<const char*>(noise) + size_t(n) * stride : nullptr;
      rc = launch_heun_step(*m, E, seed, step0 + uint64_t(n), nz, s, nullptr, &bufs);
      if (rc) return rc;
    }
    if (needs_permute(*m)) return launch_permute(*m, eint, e, 0, s);
    return MX_OK;
  });
}

int mx_llg_run_driven(mx_model* m, void* e, int64_t nsteps, const double* B, uint64_t seed, uint64_t step0,
                      void* stream) {
  return guarded([&]() -> int {
    need_model(m);
    if (!e || !B) fail(MX_EINVAL, "null spins or fields");
    if (nsteps < 0) fail(MX_EINVAL, "negative step count");
    if (!m->dyn_set) fail(MX_EINVAL, "set the dynamics first");
    use_device(*m);
    ensure_scratch(*m);
    sync_static(*m);
    hipStream_t s = static_cast<hipStream_t>(stream);
    void* E = e;
    if (needs_permute(*m)) {
      int rc = launch_permute(*m, e, m->eint_d, 1, s);
      if (rc) return rc;
      E = m->eint_d;
    }
    for (int64_t n = 0; n < nsteps; ++n) {
      const uint64_t step = step0 + uint64_t(n);
      int rc = launch_heun_pass(*m, 1, E, m->ep_d, m->k1_d, seed, step, nullptr, nullptr, s, B + 6 * n);
      if (rc == MX_OK) rc = launch_heun_pass(*m, 2, E, m->ep_d, m->k1_d, seed, step, nullptr, nullptr, s, B + 6 * n + 3);
      if (rc) return rc;
    }
    if (needs_permute(*m)) return launch_permute(*m, m->eint_d, e, 0, s);
    return MX_OK;
  });
}

int mx_llg_stage(mx_model* m, void* e, int32_t stage, const double* B, const void* extra, uint64_t seed,
                 uint64_t step, void* ep_out, void* stream) {
  return guarded([&]() -> int {
    need_model(m);
    if (!e) fail(MX_EINVAL, "null spins");
    if (stage != 1 && stage != 2) fail(MX_EINVAL, "stage must be 1 (predictor) or 2 (corrector)");
    if (!m->dyn_set) fail(MX_EINVAL, "set the dynamics first");
    use_device(*m);
    ensure_scratch(*m);
    sync_static(*m);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const bool perm = needs_permute(*m);
    void* E = perm ? m->eint_d : e;
    int rc = MX_OK;
    if (perm && stage == 1) rc = launch_permute(*m, e, m->eint_d, 1, s);
    if (rc == MX_OK) rc = launch_heun_pass(*m, stage, E, m->ep_d, m->k1_d, seed, step, nullptr, nullptr, s, B, extra);
    if (rc == MX_OK && stage == 1 && ep_out) {
      if (perm) rc = launch_permute(*m, m->ep_d, ep_out, 0, s);
      else hip_ok(hipMemcpyAsync(ep_out, m->ep_d, size_t(m->N) * 3 * rsize(*m), hipMemcpyDeviceToDevice, s),
                  "hipMemcpyAsync predictor spins");
    }
    if (rc == MX_OK && stage == 2 && perm) rc = launch_permute(*m, m->eint_d, e, 0, s);
    return rc;
  });
}

int mx_model_set_sot(mx_model* m, const double p[3], double B_DL, double B_FL) {
  return guarded([&]() -> int {
    need_model(m);
    if (!p) fail(MX_EINVAL, "null polarisation");
    for (int a = 0; a < 3; ++a) m->sot_p[a] = p[a];
    m->sot_dl = B_DL;
    m->sot_fl = B_FL;
    return MX_OK;
  });
}

int mx_mc_set_colors(mx_model* m, int32_t ncolor, const int64_t* off, const int32_t* sites) {
  return guarded([&]() -> int {
    need_model(m);
    if (ncolor < 0 || (ncolor > 0 && (!off || !sites))) fail(MX_EINVAL, "null colour lists");
    if (ncolor > 0 && off[0] != 0) fail(MX_EINVAL, "colour offsets must start at 0");
    for (int c = 0; c < ncolor; ++c)
      if (off[c + 1] < off[c]) fail(MX_EINVAL, "colour offsets must not decrease");
    const int64_t n = ncolor > 0 ? off[ncolor] : 0;
    if (n > m->N) fail(MX_EINVAL, "more colour entries than sites");
    std::vector<int32_t> q(static_cast<size_t>(n));
    std::vector<char> seen(static_cast<size_t>(m->N), 0);
    for (int64_t k = 0; k < n; ++k) {
      const int32_t r = sites[k];
      if (r < 0 || r >= m->N) fail(MX_EINVAL, "colour site out of range");
      if (seen[r]) fail(MX_EINVAL, "a site appears in two colours");
      seen[r] = 1;
      q[k] = internal_index(*m, r);
    }
    use_device(*m);
    m->layout_frozen = 1;
    if (m->mc_sites_d) hip_ok(hipFree(m->mc_sites_d), "hipFree");
    m->mc_sites_d = nullptr;
    if (n > 0) {
      hip_ok(hipMalloc(reinterpret_cast<void**>(&m->mc_sites_d), sizeof(int32_t) * n), "hipMalloc colours");
      hip_ok(hipMemcpy(m->mc_sites_d, q.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice), "hipMemcpy colours");
    }
    if (ncolor > 0) m->mc_off.assign(off, off + ncolor + 1);
    else m->mc_off.assign(1, 0);
    return MX_OK;
  });
}

int mx_mc_set_temperature(mx_model* m, int per_site, const double* T) {
  return guarded([&]() -> int {
    need_model(m);
    if (!T) fail(MX_EINVAL, "null temperature");
    auto beta = [](double t) { return t > 0 ? 1.0 / (K_B * t) : INFINITY; };
    use_device(*m);
    m->layout_frozen = 1;
    dfree(m->mc_beta_d);
    m->mc_beta_finite = 1;
    if (per_site) {
      std::vector<double> b(static_cast<size_t>(m->N));
      for (int r = 0; r < m->N; ++r) {
        if (!(T[r] >= 0)) fail(MX_EINVAL, "temperatures must be >= 0");
        b[internal_index(*m, r)] = beta(T[r]);
        if (!std::isfinite(b[internal_index(*m, r)])) m->mc_beta_finite = 0;
      }
      hip_ok(hipMalloc(&m->mc_beta_d, b.size() * rsize(*m)), "hipMalloc beta");
      upload(*m, m->mc_beta_d, b.data(), b.size());
    } else {
      if (!(T[0] >= 0)) fail(MX_EINVAL, "temperature must be >= 0");
      m->mc_beta = beta(T[0]);
      m->mc_beta_finite = std::isfinite(m->mc_beta) ? 1 : 0;
    }
    return MX_OK;
  });
}

int mx_mc_sweep(mx_model* m, void* e, int64_t nsweeps, int move, int overrelax, double sigma, uint64_t seed,
                uint64_t sweep0, void* stream) {
  return guarded([&]() -> int {
    need_model(m);
    if (!e) fail(MX_EINVAL, "null spins");
    if (nsweeps < 0 || overrelax < 0) fail(MX_EINVAL, "negative sweep or over-relaxation count");
    if (move < 0 || move > 2) fail(MX_EINVAL, "move must be 0 (gaussian), 1 (uniform) or 2 (heat bath)");
    if (m->mc_off.empty()) fail(MX_EINVAL, "set the colours first (mx_mc_set_colors)");
    use_device(*m);
    ensure_scratch(*m);
    sync_static(*m);
    if (!m->mc_acc_d) {
      hip_ok(hipMalloc(reinterpret_cast<void**>(&m->mc_acc_d), sizeof(unsigned long long)), "hipMalloc counter");
      hip_ok(hipMemset(m->mc_acc_d, 0, sizeof(unsigned long long)), "hipMemset counter");
    }
    // heat bath needs a finite beta everywhere; otherwise gaussian moves (magex/montecarlo.py:166-167)
    const int mv = (move == 2 && !m->mc_beta_finite) ? 0 : move;
    hipStream_t s = static_cast<hipStream_t>(stream);
    void* E = e;
    if (needs_permute(*m)) {
      int rc = launch_permute(*m, e, m->eint_d, 1, s);
      if (rc) return rc;
      E = m->eint_d;
    }
    const int nc = int(m->mc_off.size()) - 1;
    for (int64_t k = 0; k < nsweeps; ++k) {
      const uint64_t sweep = sweep0 + uint64_t(k);
      for (int pass = 0; pass <= overrelax; ++pass)
        for (int c = 0; c < nc; ++c) {
          const int64_t a = m->mc_off[c], b = m->mc_off[c + 1];
          int rc = launch_mc(*m, E, m->mc_sites_d + a, int(b - a), m->mc_beta_d, m->mc_beta, mv, pass, sigma, seed,
                             sweep, m->mc_acc_d, s);
          if (rc) return rc;
        }
    }
    if (needs_permute(*m)) return launch_permute(*m, m->eint_d, e, 0, s);
    return MX_OK;
  });
}

int mx_mc_accepted(mx_model* m, int64_t* out, int reset) {
  return guarded([&]() -> int {
    need_model(m);
    if (!out) fail(MX_EINVAL, "null output");
    use_device(*m);
    unsigned long long v = 0;
    if (m->mc_acc_d) {
      hip_ok(hipMemcpy(&v, m->mc_acc_d, sizeof(v), hipMemcpyDeviceToHost), "hipMemcpy counter");
      if (reset) hip_ok(hipMemset(m->mc_acc_d, 0, sizeof(v)), "hipMemset counter");
    }
    *out = int64_t(v);
    return MX_OK;
  });
}

int mx_noise(mx_model* m, void* out, uint64_t seed, uint64_t step, void* stream) {
  return guarded([&]() -> int {
    need_model(m);
    if (!out) fail(MX_EINVAL, "null output");
    use_device(*m);
    return launch_noise(*m, out, seed, step, static_cast<hipStream_t>(stream));
  });
}

static const void* internal_spins(mx_model* m, const void* e, hipStream_t s, int& rc) {
  rc = MX_OK;
  if (!needs_permute(*m)) return e;
  rc = launch_permute(*m, e, m->eint_d, 1, s);
  return m->eint_d;
}

int mx_field(mx_model* m, const void* e, void* out, int kind, void* stream) {
  return guarded([&]() -> int {
    need_model(m);
    if (!e || !out) fail(MX_EINVAL, "null buffer");
    if (kind < MX_FIELD_BEFF || kind > MX_FIELD_EXCHANGE) fail(MX_EINVAL, "bad field kind");
    use_device(*m);
    ensure_scratch(*m);
    sync_static(*m);
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc;
    const void* E = internal_spins(m, e, s, rc);
    if (rc) return rc;
    return launch_field(*m, E, out, kind, s);
  });
}

int mx_observables(mx_model* m, const void* e, double* out, void* stream) {
  return guarded([&]() -> int {
    need_model(m);
    if (!e || !out) fail(MX_EINVAL, "null buffer");
    use_device(*m);
    ensure_scratch(*m);
    sync_static(*m);
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc;
    const void* E = internal_spins(m, e, s, rc);
    if (rc) return rc;
    return launch_observables(*m, E, out, s);
  });
}

int mx_field_images(mx_model* m, const void* e, int32_t nimg, void* out, int kind, void* stream) {
  return guarded([&]() -> int {
    need_model(m);
    if (!e || !out) fail(MX_EINVAL, "null buffer");
    if (nimg < 1) fail(MX_EINVAL, "no images");
    if (kind < MX_FIELD_BEFF || kind > MX_FIELD_EXCHANGE) fail(MX_EINVAL, "bad field kind");
    use_device(*m);
    ensure_scratch(*m);
    sync_static(*m);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (!needs_permute(*m)) return launch_field(*m, e, out, kind, s, nimg);
    // images in magex order that the kernels read in another order: one image at a time
    const size_t b = size_t(m->N) * 3 * rsize(*m);
    for (int k = 0; k < nimg; ++k) {
      int rc;
      const void* E = internal_spins(m, static_cast<const char*>(e) + k * b, s, rc);
      if (rc == MX_OK) rc = launch_field(*m, E, static_cast<char*>(out) + k * b, kind, s);
      if (rc) return rc;
    }
    return MX_OK;
  });
}

int mx_observables_images(mx_model* m, const void* e, int32_t nimg, double* out, void* stream) {
  return guarded([&]() -> int {
    need_model(m);
    if (!e || !out) fail(MX_EINVAL, "null buffer");
    if (nimg < 1) fail(MX_EINVAL, "no images");
    use_device(*m);
    ensure_scratch(*m);
    sync_static(*m);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (!needs_permute(*m)) {
      if (nimg > m->img_part_cap) {
        if (m->img_part_d) hip_ok(hipFree(m->img_part_d), "hipFree");
        m->img_part_d = nullptr;
        m->img_part_cap = 0;
        const int nblk = m->ns * ((m->ncell + BLOCK - 1) / BLOCK);
        hip_ok(hipMalloc(reinterpret_cast<void**>(&m->img_part_d), sizeof(double) * MX_OBS_LEN * nblk * size_t(nimg)),
               "hipMalloc image partials");
        m->img_part_cap = nimg;
      }
      return launch_observables(*m, e, out, s, nimg, m->img_part_d);
    }
    const size_t b = size_t(m->N) * 3 * rsize(*m);
    for (int k = 0; k < nimg; ++k) {
      int rc;
      const void* E = internal_spins(m, static_cast<const char*>(e) + k * b, s, rc);
      if (rc == MX_OK) rc = launch_observables(*m, E, out + size_t(k) * MX_OBS_LEN, s);
      if (rc) return rc;
    }
    return MX_OK;
  });
}

int mx_rhs(mx_model* m, const void* e, const void* B, double alpha, int precession, void* out, void* stream) {
  return guarded([&]() -> int {
    need_model(m);
    if (!e || !B || !out) fail(MX_EINVAL, "null buffer");
    use_device(*m);
    return launch_rhs(*m, e, B, alpha, precession ? 1 : 0, out, static_cast<hipStream_t>(stream));
  });
}

int mx_topo_charge(mx_model* m, const void* e, const int32_t* tri, int64_t ntri, double hand, double* out,
                   void* stream) {
  return guarded([&]() -> int {
    need_model(m);
    if (!e || !out || (ntri > 0 && !tri)) fail(MX_EINVAL, "null buffer");
    if (ntri < 0) fail(MX_EINVAL, "negative triangle count");
    use_device(*m);
    ensure_scratch(*m);
    return launch_topo(*m, e, tri, ntri, hand, out, static_cast<hipStream_t>(stream));
  });
}

}  // extern "C"
