// Internal declarations shared by the host API (mxllg_api.hip) and the kernels (mxllg_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "mxllg.h"

namespace mx {

constexpr double MU_B = 5.7883818060e-2;  // meV / T
constexpr double K_B = 8.617333262e-2;    // meV / K
constexpr double GAMMA = 0.1760859630;    // rad / (ps T)

constexpr int MAXSUB = 8;    // sublattices per cell
constexpr int MAXENT = 64;   // stencil entries over all sublattices
constexpr int ENT_PAD = 16;  // records readable past the last entry: the widest gather batch
constexpr int MAXTAB = 256;  // coupling-table values (ntypes x width)
constexpr int BLOCK = 256;   // threads per workgroup (4 wavefronts of 64)
// Most sites of a model (or batch): the kernels form element offsets of per-site arrays in 32-bit ints,
// up to 9 N for the row-major per-site on-site tensor (3 N for spins), so 9 N must stay below 2^31.
constexpr int64_t MAX_SITES = (int64_t(1) << 31) / 9 - 1;

enum CoupMode { COUP_TABLE1 = 0, COUP_TABLE4 = 1, COUP_TABLE9 = 2, COUP_BOND = 3 };
// Cell layouts: OPEN (row-major, some axis open), PERIODIC (row-major, all axes periodic), TILED (all
// axes periodic, cells grouped in 3D tiles of BLOCK cells, power-of-two sides, one tile per workgroup),
// CHECKER (all axes periodic, L2 even: the cells of even c0 + c1 + c2, then the odd ones, each half
// row-major with c2 / 2 fastest -- a two-colour Monte Carlo sweep then reads each colour contiguously).
enum Geometry { GEO_OPEN = 0, GEO_PERIODIC = 1, GEO_TILED = 2, GEO_CHECKER = 3 };
enum OnsiteMode { ONSITE_SUB = 0, ONSITE_SITE6 = 1, ONSITE_SITE9 = 2 };

// One stencil entry (16 bytes): slot = coupling slot (type-table row, or forward-bond slot),
// tb = target sublattice | back << 16, d01 = (d0 & 0xffff) | d1 << 16, d2; offsets d are reduced into
// [0, L) on periodic axes.
struct Ent {
  int32_t slot, tb, d01, d2;
};

// Model constants, resident in device memory (one block per model, uploaded when the model is
// configured).  Kernels read it through a const __restrict__ pointer with wavefront-uniform indices,
// so it comes in through the scalar cache; the kernel arguments themselves stay small.
template <typename R>
struct Static {
  int L0, L1, L2, ncell, ns, N, bps, nblk;  // bps: workgroups per sublattice
  int pbc0, pbc1, pbc2, allpbc;
  int tb0, tb1, tb2, nt1, nt2;  // TILED: log2 of the tile sides; tiles along axes 1 and 2
  int fdiv;                     // ncell < 2^24: divide cell indices by L1, L2 through rL1, rL2
  int xcol;                     // step kernels' column map: row-major tiles per XCD per axis-0 plane (0: slabs)
  int half, L2h;                // CHECKER: cells per parity, L2 / 2
  float rL2h;
  float rL1, rL2;
  int amode, has_A, has_c, has_u;
  int noise_dbl;  // thermal normals at double grade (gauss3d) instead of float grade (gauss3)
  int sptr[MAXSUB + 1];
  Ent ent[MAXENT + ENT_PAD];  // padded: a batch of up to ENT_PAD entries may be read past the last one (never used)
  R tab[MAXTAB];  // COUP_TABLE*: ntypes x width
  R Asub[MAXSUB][9];
  R inv_mu_sub[MAXSUB], alpha_sub[MAXSUB], sigma_sub[MAXSUB], sign_sub[MAXSUB];
  R C[9], u[3];
};

// Per-launch values and the model's per-site arrays, passed by value: pointers that arrive as kernel
// arguments are known to be global memory, so their loads compile to global_load (not flat_load).
template <typename R>
struct Dyn {
  R B0, B1, B2, dt;
  int prec, noisy;
  int integ;  // MX_HEUN or MX_DEPONDT
  int sot;    // spin-orbit torque fields on: B += bdl (e x p) + bfl p
  int rev2;   // the corrector pass walks each XCD's slab backwards (starts on the predictor's last tiles)
  R sp0, sp1, sp2, bdl, bfl;
  uint32_t seed_lo, seed_hi;
  const R* bond;  // COUP_BOND: [nfwd][ncell][9], row-major tensor per bond (36 B FP32 / 72 B FP64)
  const R* A;     // ONSITE_SITE6: 6 x N (xx, yy, zz, xy, xz, yz); ONSITE_SITE9: 9 x N (row-major)
  const R* Kc;
  const R* K2;
  const R* K3;
  const R* inv_mu;  // per-site 1 / (mu mu_B), or null -> inv_mu_sub
  const R* alpha;   // per-site damping, or null -> alpha_sub
  const R* sigma;   // per-site thermal amplitude (T), or null -> sigma_sub
  const R* sign_site;
  const uint8_t* pinned;
  const uint64_t* step_ptr;  // captured step loops: device step counter added to the step argument
  R* save_e;                 // adjoint replays: pass 1 stores e here, both passes store B_eff in save_B
  R* save_B;
  const R* extra;            // caller-supplied field of this pass (magex order, N x 3), or null
  R Bc0, Bc1, Bc2;           // uniform field of the corrector in the one-launch step (B0..B2: predictor)
};

// Control words of the one-launch Heun steps (k_step, k_steps) in model-owned device memory: L0 plane
// counters, CTL_WORDS control words, one give-up slot per workgroup.  Every word is zero between
// launches: the last workgroup of a launch clears what the launch used.  The plane counters sit one per
// 128-byte line: hundreds of wavefronts poll the few counters around their plane while others increment
// them.
constexpr int CNT_STRIDE = 32;
constexpr int CTL_WORDS = 4;
struct StepSync {
  uint32_t* cnt;  // per axis-0 plane (stride CNT_STRIDE): workgroups whose predictor spins covering the
                  // plane are published (k_step)
  uint32_t* ctl;  // [0] workgroups finished, [1] waits given up in this launch, [2] given up (k_step) or
                  // launches aborted (k_steps) since creation, [3] abort raised in this launch (k_steps)
  int32_t* gu;    // blocks whose wait gave up (their corrector runs in the launch's last workgroup)
  int r0;         // stencil reach along axis 0, in planes
  int PS;         // cells per axis-0 plane
  uint32_t spin;  // polls of a wait before it gives up
};

// Host-side model.
struct Model {
  int prec;  // MX_FP64 / MX_FP32
  int device;
  int ns, L[3], pbc[3], ncell, N;
  // cell layout: geo (Geometry), tile sides 2^tb; perm_d[q] = magex index of internal site q
  int geo = GEO_PERIODIC, tb[3] = {0, 0, 0}, want_tiled = 1, layout_frozen = 0;
  int32_t* perm_d = nullptr;
  // stencil
  int nent = 0;
  std::vector<int32_t> ent_raw;  // nent x 6 as given
  std::vector<int> sptr;
  std::vector<Ent> ent;        // internal order (sorted by source sublattice)
  std::vector<int> ent_order;  // ent[pos] came from ent_raw entry ent_order[pos]
  int coup = COUP_TABLE1;
  int ntypes = 0, width = 0;
  std::vector<double> table;
  void* bond_d = nullptr;
  int nfwd = 0;
  // on-site
  int amode = ONSITE_SUB, has_A = 0;
  std::vector<double> Asub;  // ns x 9
  void* A_d = nullptr;
  int has_c = 0, has_u = 0;
  void *Kc_d = nullptr, *K2_d = nullptr, *K3_d = nullptr;
  double C[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, u[3] = {0, 0, 1};
  // moments and dynamics
  std::vector<double> mu_sub, sign_sub;
  std::vector<double> mu_site, sign_site_h;  // empty unless per-site
  void* inv_mu_d = nullptr;
  void* sign_d = nullptr;
  void* pinned_d = nullptr;
  double B[3] = {0, 0, 0};
  double sot_p[3] = {0, 1, 0}, sot_dl = 0, sot_fl = 0;  // spin-orbit torque fields (off when both are 0)
  double dt = 1e-3;
  int precession = 1, integrator = MX_HEUN;
  std::vector<double> alpha_h, T_h;  // size 1 (uniform) or N (per site), magex order
  void* alpha_d = nullptr;
  void* sigma_d = nullptr;
  std::vector<double> alpha_sub_v, sigma_sub_v;
  int noisy = 0;
  int dyn_set = 0;
  int noise_dbl = 0;  // generator grade (MX_NOISE_DOUBLE = 1); set to the precision's own at creation
  // device constant block (Static<R>) and its dirty flag
  void* static_d = nullptr;
  int static_dirty = 1;
  // scratch
  void *ep_d = nullptr, *k1_d = nullptr, *eint_d = nullptr;  // predictor ep; k1_d holds w = e + dt/2 k1
  double* part_d = nullptr;
  int npart = 0;
  // Monte Carlo: colour site lists (internal order, colours concatenated), inverse temperatures
  int32_t* mc_sites_d = nullptr;
  std::vector<int64_t> mc_off;  // ncolor + 1 offsets into mc_sites_d
  void* mc_beta_d = nullptr;    // per-site beta (model precision, internal order) or null
  double mc_beta = 0.0;         // uniform beta (1/meV; +inf at T = 0)
  int mc_beta_finite = 1;       // every beta finite (heat bath possible)
  unsigned long long* mc_acc_d = nullptr;
  void* mc_rep_d = nullptr;  // replica sweeps: per-replica parameters and acceptance counters
  unsigned long long* mc_repacc_d = nullptr;
  int mc_rep_cap = 0;
  void* mc_repe_d = nullptr;  // replica spins in internal site order (layouts that permute), mc_repe_cap of them
  int mc_repe_cap = 0;
  double* img_part_d = nullptr;  // partials of image-stack observables
  int img_part_cap = 0;          // images they hold
  // one-launch steps: mode (0: two launches per step; 1: one launch per step where the model allows it;
  // 2, the default: also runs of steps in persistent launches where they pay), control words
  int step_mode = 2;
  // polls after which a wait of the one-launch, persistent and adjoint-sweep kernels gives up (each
  // poll is a round of loads and a short sleep)
  uint32_t wait_polls = 1u << 16;
  uint32_t* stepsync_d = nullptr;  // L0 plane counters, 4 control words, nblk give-up slots
  // persistent launches (k_steps): tagged spin records (2 x N x TAG_WORDS 64-bit words: spins, predictor
  // spins; zeroed when allocated) and the epoch of the last launch on the model (each launch takes the
  // next, 1 .. 15, so that no record a launch finds matches a tag it waits for)
  uint64_t* rec_d = nullptr;
  mutable uint32_t rec_epoch = 0;
  // captured step loop (hipGraph of graph_steps Heun steps + counter advance), keyed by what it bakes in
  int graph_steps = 0;
  uint64_t* counter_d = nullptr;
  hipStream_t cap_stream = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<char> gkey;
};

// Scratch of one step: the predictor spins ep, w = e + dt/2 k1 (or the rotation vector) in k1, and the
// one-launch step's control words (stepsync_words(m) of them, zero between launches).
struct StepBufs {
  void* ep;
  void* k1;
  uint32_t* sync;
  uint64_t* rec;  // tagged spin records of persistent launches (rec_words(m) words, zero or tagged)
};

// Kernel launchers (mxllg_kernels.hip).  All take internal-order device buffers unless noted.
int upload_static(Model& m);
// One step on e; bufs = null: the model's own scratch.
int launch_heun_step(const Model& m, void* e, uint64_t seed, uint64_t step, const void* noise_step, hipStream_t s,
                     const uint64_t* step_ptr = nullptr, const StepBufs* bufs = nullptr);
// Zero n 32-bit words at p (a kernel, so that it can sit in captured work).
int launch_fill_u32(uint32_t* p, size_t n, hipStream_t s);
int launch_counter(uint64_t* c, uint64_t value, int add, hipStream_t s);
// Whether a step of m runs as one launch (k_step), and the number of held coupling batches it uses.
bool fused_step(const Model& m, int* nb = nullptr);
// Whether a step of m runs as one walked launch (k_walk: FP32 per-bond tensors held in registers, the
// per-XCD column map, a grid larger than the GPU holds at once), its held batches and its grid.
bool walk_step(const Model& m, int* nb = nullptr, int* grid = nullptr);
// Whether runs of steps of m run in persistent launches (k_steps: step mode 2, row-major cells, a grid of
// at most one workgroup per compute unit, a small round of polls).
bool persistent_steps(const Model& m);
// nsteps steps on e in persistent launches of up to STEPS_PER_LAUNCH steps (persistent_steps(m) must hold);
// noise: nsteps x N x 3 normals or null; bufs = null: the model's own scratch and records.
// rec (optional): records of every step n, internal order, N x 3 each at n * N * 3 values from each base --
// the step's input spins e, its predictor spins ep, w = e + dt/2 k1 (Depondt: the rotation vector), and the
// effective fields of both passes (zero on pinned sites, whose w is their spin): what the adjoint's replays
// store with the single-pass kernels.
// plain: an ordinary launch instead of a cooperative one, for a caller that runs it beside one other
// small persistent launch (the runtime runs cooperative launches one at a time); both grids must then fit
// on the GPU together.
struct StepRecOut {
  void *e, *ep, *w, *b1, *b2;
};
constexpr int STEPS_PER_LAUNCH = 1024;
// Dynamic LDS (unused) of a plain launch that runs beside another persistent one: more than half of a
// compute unit's 160 KB, so the two grids never share a compute unit.
constexpr size_t SIDE_BY_SIDE_LDS = 96 * 1024;
int launch_heun_steps(const Model& m, void* e, uint64_t seed, uint64_t step0, int64_t nsteps, const void* noise,
                      hipStream_t s, const StepBufs* bufs = nullptr, const StepRecOut* rec = nullptr,
                      bool plain = false);
// The model's own step scratch (predictor buffers, control words, tagged records), allocated if missing
// (MX_OK or an error code with the message set).
int ensure_step_scratch(Model& m);
// Device words of the one-launch steps (allocated and zeroed with the model's scratch: plane counters,
// control words, give-up slots, the walked step's ticket counters), and the offset of the control words
// among them.
size_t stepsync_words(const Model& m);
inline size_t stepsync_ctl(const Model& m) { return size_t(m.L[0]) * CNT_STRIDE; }
// 64-bit words of the tagged spin records of persistent launches.
inline size_t rec_words(const Model& m) { return 2 * size_t(m.N) * (m.prec == MX_FP64 ? 4 : 2); }
// One Heun pass (1 or 2) with optional stores of e and B_eff (magex order is not used: internal order).
// B_ext (3 doubles) replaces the model's uniform field for this pass when non-null; extra (device, N x 3
// in the model precision, magex site order) is added to the field of every site when non-null.
int launch_heun_pass(const Model& m, int pass, void* e, void* ep, void* k1, uint64_t seed, uint64_t step,
                     void* save_e, void* save_B, hipStream_t s, const double* B_ext = nullptr,
                     const void* extra = nullptr);
int sync_model_static(Model& m);
size_t prec_size(const Model& m);
// nimg images of N x 3 values each, one after another (field: in and out; observables: nimg x MX_OBS_LEN
// out, partials in `part`, nimg x blocks x MX_OBS_LEN, default the model's own for one image).
int launch_field(const Model& m, const void* e_int, void* out_ref, int kind, hipStream_t s, int nimg = 1);
int launch_observables(const Model& m, const void* e_int, double* out, hipStream_t s, int nimg = 1,
                       double* part = nullptr);
int launch_noise(const Model& m, void* out_ref, uint64_t seed, uint64_t step, hipStream_t s);
int launch_permute(const Model& m, const void* src, void* dst, int to_internal, hipStream_t s);
bool needs_permute(const Model& m);
int cell_to_internal(const Model& m, int c);  // internal cell of row-major (magex) cell c
int launch_rhs(const Model& m, const void* e, const void* B, double alpha, int prec, void* out, hipStream_t s);
// One Monte Carlo pass over n sites of one colour (mxllg_mc.hip); pass 0 = the move, > 0 = over-relaxation.
int launch_mc(const Model& m, void* e, const int32_t* sites, int n, const void* beta_site, double beta, int move,
              int pass, double sigma, uint64_t seed, uint64_t sweep, unsigned long long* acc, hipStream_t s);
int launch_topo(const Model& m, const void* e_ref, const int32_t* tri, int64_t ntri, double hand, double* out,
                hipStream_t s);

void set_error(const std::string& msg);

}  // namespace mx

struct mx_model : mx::Model {};
