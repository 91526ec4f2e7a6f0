/* mxllg.h: C ABI of the MI355X (gfx950) atomistic LLG hot path.
 *
 * One library, libmxllg.so, replaces the kernels magex calls on this path:
 *   - the numba CPU kernels _fused1 / _fused2 / _field / _gauss3 (magex/fast.py:189-223, 336-390) that
 *     LLGSolver._fast_step / _fast_field drive (magex/llg.py:153-214);
 *   - the numba.cuda kernels _k_pred / _k_corr and their host class CudaLLG
 *     (magex/cuda_backend.py:144-218, 338-377);
 *   - the NumPy field / energy / observables of SpinSystem (magex/spinmodel.py:454-484, 517-524)
 *     and the Berg-Luscher charge of magex/texture.py:30-63.
 *
 * Conventions (magex/spinmodel.py:3-11, magex/llg.py:1-9): energies in meV, fields in T, time in ps,
 * E = sum_i e_i.A_i.e_i - sum_{i!=j} e_i.J_ij.e_j - mu_B sum_i mu_i e_i.B, every ordered pair counted.
 *
 * Memory: every spin array argument is a DEVICE pointer to N x 3 values (x, y, z per site) in the
 * model's precision (double for MX_FP64, float for MX_FP32), sites in magex's order
 * (site = cell_index * nsub + sublattice, cell_index = (c0 * L1 + c1) * L2 + c2; spinmodel.py:197-206).
 * Model set-up calls take HOST arrays of double and copy them.  The model owns its device buffers;
 * spin buffers belong to the caller.  All work is enqueued on the caller's stream (a hipStream_t,
 * passed as void*; NULL = default stream) and nothing synchronises unless stated.
 *
 * Errors: every int-returning call returns MX_OK (0) or an MX_E* code; mx_last_error() gives the
 * message of the last failing call on the calling thread.  A model is not thread-safe: one host
 * thread drives one model at a time (as magex's solver objects, magex/llg.py:96).
 */
#ifndef MXLLG_H
#define MXLLG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { MX_OK = 0, MX_EINVAL = 1, MX_ENOMEM = 2, MX_EHIP = 3, MX_EUNSUPPORTED = 4 };
enum { MX_FP64 = 0, MX_FP32 = 1 };
enum { MX_FIELD_BEFF = 0, MX_FIELD_GRADIENT = 1, MX_FIELD_EXCHANGE = 2 };
enum { MX_HEUN = 0, MX_DEPONDT = 1 };
enum { MX_OBS_E = 0, MX_OBS_M = 1, MX_OBS_ORDER = 4, MX_OBS_TORQUE = 7, MX_OBS_LEN = 8 };
enum { MX_NOISE_FLOAT = 0, MX_NOISE_DOUBLE = 1 };

typedef struct mx_model mx_model;

/* Library identity. */
const char* mx_last_error(void);
int mx_version(void);

/* Model of a periodic (or open) supercell of a unit cell with nsub magnetic sites.
 * Replaces the array set-up of fast.FastModel (magex/fast.py:630-672) and cuda_backend._Device
 * (magex/cuda_backend.py:307-335); the 3N x 3N CSR of SpinSystem (spinmodel.py:211-241) is never built:
 * couplings are a stencil over cells.  L: supercell (cells per axis, each at most 32767); pbc: periodic
 * flags.  At most 238 609 293 sites ((2^31 - 1) / 9: the kernels index per-site arrays of up to 9 values
 * per site with 32-bit offsets); larger models return MX_EUNSUPPORTED. */
int mx_model_create(int precision, int nsub, const int32_t L[3], const int32_t pbc[3], mx_model** out);
void mx_model_destroy(mx_model* m);
int mx_model_info(const mx_model* m, int64_t* n_sites, int32_t* precision, int32_t* nsub);

/* Cell layout inside the library (the ABI's arrays are always in magex's site order): tiled = 1
 * groups the cells of a fully periodic supercell into 3D tiles of 256 cells with power-of-two sides
 * when the supercell allows it (each side divisible, at least 1024 cells); tiled = 2 (checker) stores
 * the cells of even c0 + c1 + c2 first and the odd ones after, on fully periodic supercells with an
 * even axis-2 side, so a two-colour Monte Carlo sweep reads each colour contiguously (the adjoint
 * entry points refuse it); tiled = 0 (the default) keeps them row-major.  A layout the supercell does
 * not allow falls back to row-major.  Results are identical; on MI355X the row-major order measured
 * fastest for the LLG step.  Only callable before any per-site data is set.  mx_model_layout reports
 * the tile sides (0 when row-major, -1 for checker).  No counterpart in magex (its arrays are always
 * row-major, spinmodel.py:197-206). */
int mx_model_set_layout(mx_model* m, int tiled);
int mx_model_layout(const mx_model* m, int32_t tile[3]);

/* Exchange stencil: nent entries of 6 int32 (s, t, R0, R1, R2, type): the coupling J = table[type] acts
 * from site (c, s) on site (c + R, t) (spinmodel.py:214-241).  table: ntypes x width doubles, width 1
 * (J I), 4 (J I + DMI vector D, J^A_ab = eps_abc D_c; fast.py:104-111) or 9 (row-major 3x3).
 * Entries whose target is the source itself on every cell are rejected: fold them into the on-site
 * tensor first (spinmodel.py:224-229). */
int mx_model_set_stencil(mx_model* m, int32_t nent, const int32_t* ent, int32_t ntypes, int32_t width,
                         const double* table);

/* Per-bond 3x3 tensors on the stencil (no translation symmetry; the per-bond storage of
 * fast.py:612-614).  Every stencil entry must have a partner (t, s, -R) so that each unordered bond is
 * stored once and J_ji = J_ij^T.  The stored ("forward") entries are those with s < t, or s == t and
 * (R0, R1, R2) lexicographically positive, in stencil order; J: nfwd x ncell x 9 doubles, J[f, c] =
 * row-major tensor of forward entry f from cell c.  Replaces the stencil's type table. */
int mx_model_set_bond_tensors(mx_model* m, const double* J);

/* On-site tensor A (gradient 2 A e, energy e.A.e; spinmodel.py:462,479).  per_site=0: nsub x 9,
 * per_site=1: N x 9 (site order as above). */
int mx_model_set_onsite(mx_model* m, int per_site, const double* A);

/* Cubic E = Kc (a1^2 a2^2 + a2^2 a3^2 + a3^2 a1^2), a = C e, and higher uniaxial
 * E = -K2 (e.u)^4 - K3 (e.u)^6 (magex/terms.py:35-75, in-kernel as fast.py:47-63).  Per-site arrays
 * of N doubles; pass NULL to switch a term off. */
int mx_model_set_cubic(mx_model* m, const double* Kc, const double C[9]);
int mx_model_set_uniaxial_higher(mx_model* m, const double* K2, const double* K3, const double u[3]);

/* Moments mu (mu_B) and signs of the reference state: per_site=0: nsub values, 1: N values. */
int mx_model_set_moments(mx_model* m, int per_site, const double* mu, const double* signs);

/* Pinned sites (N bytes, nonzero = fixed; NULL = none). */
int mx_model_set_pinned(mx_model* m, const uint8_t* pinned);

/* External field B (T), held by value: no upload per step (cuda_backend.py:331-335 uploads it). */
int mx_model_set_field(mx_model* m, const double B[3]);

/* Dynamics of LLGSolver (llg.py:49-122): dt (ps); damping alpha and temperature T (K) either uniform
 * (per_site=0: one value) or per site (N values); sigma_i = sqrt(2 alpha_i k_B T_i / (gamma mu_i mu_B dt)).
 * precession 0/1; integrator MX_HEUN (stochastic Heun) or MX_DEPONDT (Heun on rotation vectors,
 * magex/fast.py:271-306 and the depondt branches of _fused1/_fused2, 336-390). */
int mx_model_set_dynamics(mx_model* m, double dt, int alpha_per_site, const double* alpha, int T_per_site,
                          const double* T, int precession, int integrator);

/* nsteps steps of the model's integrator (stochastic Heun or Depondt) on e (device, N x 3, in place);
 * replaces one _fused1 + _fused2 pair per step (fast.py:336-390, llg.py:216-240).  Thermal noise: standard normals from a counter-based
 * generator keyed by (seed, step0 + n, site) -- or, if noise != NULL, read from noise (device,
 * nsteps x N x 3, same precision) -- times sigma_i, held over both Heun stages. */
int mx_llg_run(mx_model* m, void* e, int64_t nsteps, uint64_t seed, uint64_t step0, const void* noise,
               void* stream);

/* mx_llg_run for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in magex.  It calls no hipMalloc, hipFree or
 * other allocating routine: its predictor buffers, internal-order spin copy, one-launch control words and
 * (step mode 2) tagged spin records live in a 64 MiB device array of the library that every GPU it runs
 * on holds until the process ends, which bounds the model at 3 x align256(N x 3 x w) +
 * align256(4 x (32 L0 + 4 + blocks)) (+ 8 N x w for persistent runs) bytes (about 1.8 M FP32 or 0.9 M
 * FP64 sites); larger models return MX_EUNSUPPORTED.  Steps launch directly (no captured loop).
 * The model must be configured and its constants on the device (any other entry point run once after
 * the last set-up call), else MX_EINVAL.  Calls that overlap on one GPU (other streams, other threads,
 * other models) share that array and race; results equal mx_llg_run's bit for bit. */
int mx_llg_run_noalloc(mx_model* m, void* e, int64_t nsteps, uint64_t seed, uint64_t step0, const void* noise,
                       void* stream);

/* nsteps steps as mx_llg_run with internal noise, under a time-dependent uniform field: B (host,
 * nsteps x 2 x 3 doubles) gives, for step n, the field of the predictor pass (B[6n..6n+2]) and of the
 * corrector pass (B[6n+3..6n+5]) in place of the model's field -- magex evaluates field_func at t and
 * t + dt (magex/llg.py:129-130, 204-211, 216-227). */
int mx_llg_run_driven(mx_model* m, void* e, int64_t nsteps, const double* B, uint64_t seed, uint64_t step0,
                      void* stream);

/* One stage of a Heun (or Depondt) step with a caller-supplied field, for fields the host evaluates
 * between the stages -- magex's unfused fast path with `extra` (magex/fast.py:202-223, 250-269, as
 * LLGSolver._fast_step drives it with extra_field(e) and an (N, 3) field_func(t), magex/llg.py:152-166,
 * 197-212).  stage 1 (predictor): thermal noise of (seed, step) + field(e) + predictor, kept in the
 * model; if ep_out != NULL the predictor spins are written there (device, N x 3, magex order) for the
 * caller to evaluate the corrector's field.  stage 2 (corrector): field(predictor) + corrector, e updated
 * in place.  B (host, 3 doubles, NULL = the model's field) is the uniform field of the stage; extra
 * (device, N x 3 in the model precision, magex order, NULL = none) is added to every site's field after
 * the uniform field and before the thermal field, as magex adds `extra`.  Pinned sites do not move.
 * Stage 2 must follow stage 1 on the same model and spins with no other stepping call in between. */
int mx_llg_stage(mx_model* m, void* e, int32_t stage, const double* B, const void* extra, uint64_t seed,
                 uint64_t step, void* ep_out, void* stream);

/* Spin-orbit torque as effective fields on the dynamics (not on mx_field / mx_observables, as in
 * magex): B += B_DL (e x p) + B_FL p, p a unit vector (magex/llg.py:83-86, 126-128).  B_DL = B_FL = 0
 * switches it off. */
int mx_model_set_sot(mx_model* m, const double p[3], double B_DL, double B_FL);

/* ---------------------------------------------------------------------------------------------
 * Metropolis Monte Carlo on the same stencil gather: the colour sweep of magex's numba path
 * (magex/fast.py:393-503 `_mc_group`, driven by magex/montecarlo.py:152-179).  A colour is a set of
 * mutually uncoupled sites; colours move one after another, the sites of a colour all at once.
 * off: ncolor + 1 offsets into sites (magex site order; pinned sites left out). */
int mx_mc_set_colors(mx_model* m, int32_t ncolor, const int64_t* off, const int32_t* sites);
/* Temperature of the sampler (K): per_site = 0: one value, 1: N values; beta = 1 / (k_B T), infinite
 * at T = 0 (magex/montecarlo.py:156-158). */
int mx_mc_set_temperature(mx_model* m, int per_site, const double* T);
/* nsweeps sweeps of e (device, in place): per sweep every colour with move 0 (gaussian cone of width
 * sigma), 1 (uniform) or 2 (heat bath; gaussian where some beta is infinite), then overrelax passes of
 * over-relaxation.  Random numbers: counter-based, keyed by seed, sweep0 + k, site. */
int mx_mc_sweep(mx_model* m, void* e, int64_t nsweeps, int move, int overrelax, double sigma, uint64_t seed,
                uint64_t sweep0, void* stream);
/* Accepted moves (not counting over-relaxation) since the last reset; synchronises. */
int mx_mc_accepted(mx_model* m, int64_t* out, int reset);
/* Replica sweeps for parallel tempering (magex/advanced_mc.py:38-94): nrep copies of the model's
 * lattice, replica r's spins at e + r * N * 3 (device), each with its own beta (1/meV, inf at T = 0),
 * cone width sigma and generator key seeds[r] (host arrays of nrep), all in one launch per colour and
 * pass.  Uniform temperatures only.  Spins are in magex site order; where the model's layout differs
 * (multi-site cells, tiled or checker cells) the call sweeps an internal-order copy of the replicas and
 * writes them back.  The acceptance counts of the replicas accumulate on the device;
 * mx_mc_replica_accepted reads them (synchronising). */
int mx_mc_sweep_replicas(mx_model* m, void* e, int32_t nrep, int64_t nsweeps, int move, int overrelax,
                         const double* beta, const double* sigma, const uint64_t* seeds, uint64_t sweep0,
                         void* stream);
int mx_mc_replica_accepted(mx_model* m, int32_t nrep, uint64_t* out, int reset);

/* ---------------------------------------------------------------------------------------------
 * Spin-wave sampler: m_q = (1/sqrt N) sum_i e_i exp(-i q.r_i) for nq wave vectors, the sampling step of
 * the dynamical structure factor (magex/spectra.py:49-84, `mq[s] = ph.T @ e`).  Positions
 * r_i = c0 a0 + c1 a1 + c2 a2 + tau_s come from the cell rows `cell` (9 doubles, Angstrom) and the
 * sublattice positions `basis` (nsub x 3); q: nq x 3 Cartesian (1/Angstrom); host arrays, copied. */
typedef struct mx_sampler mx_sampler;
int mx_sampler_create(const mx_model* m, int32_t nq, const double* q, const double* cell, const double* basis,
                      mx_sampler** out);
void mx_sampler_destroy(mx_sampler* s);
/* m_q of spins e (device, N x 3, magex order, model precision) into out (device, nq x 6 doubles:
 * Re x, Im x, Re y, Im y, Re z, Im z); nothing synchronises. */
int mx_sample_mq(mx_sampler* s, const void* e, double* out, void* stream);

/* Steps per captured step loop (hipGraph replayed by mx_llg_run when no noise is injected and at least
 * two loops fit; default 0 = always launch step by step).  Results are identical either way: the
 * captured kernels read the step number from a device counter, so the noise sequence is the same.
 * No counterpart in magex (its step loop is a Python for loop, llg.py:263-280). */
int mx_model_set_graph_steps(mx_model* m, int32_t steps);

/* How mx_llg_run, mx_llg_run_noalloc and mx_llg_run_driven's uniform-field steps launch: mode 0 = two
 * kernels per step (the predictor pass, then the corrector pass: magex's _fused1 / _fused2 pair,
 * fast.py:336-390); mode 1 = one kernel per step where the model allows it (row-major cells; table
 * couplings, or FP32 per-bond tensors with at most 14 stencil entries per sublattice; a grid of at most
 * one workgroup per compute unit; FP32, or FP64 up to 16384 sites), else two; mode 2 (default) = as
 * mode 1, and mx_llg_run / mx_llg_run_noalloc run two or more steps of a model with row-major cells, at
 * most one workgroup per compute unit and at most 24 MiB of stencil records per poll (N x stencil
 * entries x 16 bytes FP32, 32 bytes FP64; any coupling form) in persistent cooperative launches of up to
 * 1024 steps, each workgroup keeping its sites' state in registers across them and exchanging spins
 * with its neighbours through tagged records.  In one launch per step each workgroup keeps its sites'
 * couplings and state in registers from the predictor to the corrector and waits only for the predictor
 * spins of the axis-0 planes its stencil reaches, so per-bond tensors come from memory once per step
 * instead of twice.  Mode 3 = as mode 2, and steps of FP32 per-bond models whose grid is larger than the
 * GPU holds at once (row-major cells whose axis-0 planes split into 8 k whole tiles) run as one walked
 * launch: the resident workgroups take tiles from per-XCD ticket counters in plane order and hold each
 * tile's tensors in registers from its predictor to its corrector (measured slower than two launches on
 * bcc 64^3, 84 against 57 us per step, hence not the default).  Results are identical in every mode.
 * No counterpart in magex. */
int mx_model_set_step_mode(mx_model* m, int32_t mode);
/* one_launch: how mx_llg_run steps this model (mode and model permitting): 0 two kernels per step, 1 one
 * kernel per step, 2 runs of two or more steps in persistent launches, 3 one walked launch per step;
 * given_up: waits of one-launch
 * steps that timed out since the model was created, each finished by the launch's last workgroup
 * instead, plus persistent launches that stopped early on a timed-out wait (0 in normal operation;
 * synchronises the device).  Either may be NULL. */
int mx_model_step_info(mx_model* m, int32_t* one_launch, int64_t* given_up);
/* Polls after which a wait of the one-launch steps, persistent runs and persistent adjoint sweeps gives
 * up (default 65536, far beyond any wait of normal operation; 0: after the first unsuccessful poll).  A
 * one-launch step's given-up waits are finished by the launch's last workgroup (same results); a
 * persistent launch that gives up stops early and leaves its spins part-way, which given_up reports.
 * For tests of those paths, which then run on every launch.  No counterpart in magex. */
int mx_model_set_wait_limit(mx_model* m, int32_t polls);

/* Grade of the thermal-noise generator (the counter-based stream that replaces magex's
 * PCG64 / ziggurat draws, magex/llg.py:119, fast.py:123-200): MX_NOISE_DOUBLE (default for MX_FP64) draws
 * 53-bit uniforms from two Philox4x32-10 outputs and applies Box-Muller in double precision;
 * MX_NOISE_FLOAT (default for MX_FP32) draws 24-bit uniforms from one output with float Box-Muller.  An FP64
 * model set to MX_NOISE_FLOAT sees exactly the noise of an FP32 model with the same seed (for FP32
 * accuracy measurements against FP64 on identical noise).  Also governs mx_noise, the batched form
 * (from its prototype) and the adjoint's replays. */
int mx_model_set_noise(mx_model* m, int32_t grade);

/* The standard normals mx_llg_run draws for (seed, step): out (device, N x 3). */
int mx_noise(mx_model* m, void* out, uint64_t seed, uint64_t step, void* stream);

/* Field of e: kind MX_FIELD_BEFF (T, SpinSystem.effective_field), MX_FIELD_GRADIENT (meV, dE/de,
 * SpinSystem.gradient) or MX_FIELD_EXCHANGE (meV, h = sum_j J_ij e_j, SpinSystem.exchange_field);
 * spinmodel.py:454-473.  out: device N x 3. */
int mx_field(mx_model* m, const void* e, void* out, int kind, void* stream);

/* The field (as mx_field) and the observables (as mx_observables) of nimg spin configurations stored one
 * after another (e: nimg x N x 3; out: nimg x N x 3, resp. nimg x MX_OBS_LEN doubles), in one launch
 * for all images: the batched energy and gradient of magex's GNEB image chain (magex/gneb.py:171-187,
 * which evaluates the images one by one). */
int mx_field_images(mx_model* m, const void* e, int32_t nimg, void* out, int kind, void* stream);
int mx_observables_images(mx_model* m, const void* e, int32_t nimg, double* out, void* stream);

/* Observables of e into out (device, MX_OBS_LEN doubles): [E (meV, SpinSystem.energy), sum e (3),
 * sum s_i e_i (3, order parameter numerator), max_i |e_i x B_eff,i| (T, LLGSolver.max_torque)];
 * spinmodel.py:475-484,517-524, llg.py:250-252. */
int mx_observables(mx_model* m, const void* e, double* out, void* stream);

/* de/dt = -gamma/(1+alpha^2) [e x B + alpha e x (e x B)] for given fields (llg.py:21-28): e, B, out
 * device N x 3 in the model precision; alpha uniform. */
int mx_rhs(mx_model* m, const void* e, const void* B, double alpha, int precession, void* out, void* stream);

/* Berg-Luscher topological charge (texture.py:30-63) over ntri triangles tri (device, ntri x 3 int32
 * site indices, each counter-clockwise in the layer plane); Q = hand / (4 pi) sum 2 atan2(...).
 * out: device double. */
int mx_topo_charge(mx_model* m, const void* e, const int32_t* tri, int64_t ntri, double hand, double* out,
                   void* stream);

/* ---------------------------------------------------------------------------------------------
 * Batched form: nrep independent replicas of one unit-cell model, each on its own periodic supercell
 * and with its own damping, temperature and noise seed, advanced together (one launch per Heun pass
 * for the whole batch).  No counterpart in magex: its closest analogue is a Python loop over replicas
 * (advanced_mc.parallel_tempering, advanced_mc.py:38-94; GNEB.forces, gneb.py:171-187).
 * Replica r's spins are N_r x 3 values at site offset off[r] of one packed device array, each
 * replica in magex's site order; replica r's noise is the stream mx_llg_run draws with seed seeds[r]
 * on its own supercell, so a batched run equals nrep separate runs bit for bit. */
typedef struct mx_batch mx_batch;

/* Replicas of proto's unit cell (stencil and coupling table, per-sublattice on-site tensor, moments,
 * field, precision) on supercells L (nrep x 3 int32, all axes periodic, each side longer than the
 * stencil's reach).  Per-bond tensors and per-site arrays of proto are not carried over. */
int mx_batch_create(const mx_model* proto, int32_t nrep, const int32_t* L, mx_batch** out);
void mx_batch_destroy(mx_batch* b);
/* off: nrep + 1 site offsets (off[nrep] = total sites). */
int mx_batch_offsets(const mx_batch* b, int64_t* off);
/* dt (ps); alpha and T (K): nrep values each; precession 0/1. */
int mx_batch_set_dynamics(mx_batch* b, double dt, const double* alpha, const double* T, int precession);
/* Launch shape: threads per workgroup (64, 128, 256 or 512), sites per thread (1 to 8), and steps per
 * captured step loop (0 = direct launches).  Results do not depend on the shape. */
int mx_batch_set_launch(mx_batch* b, int32_t block, int32_t sites_per_thread, int32_t graph_steps);
/* nsteps Heun steps of every replica on the packed spins e (device, in place); seeds: nrep keys. */
int mx_batch_run(mx_batch* b, void* e, int64_t nsteps, const uint64_t* seeds, uint64_t step0, void* stream);
/* Per-replica observables (device, nrep x MX_OBS_LEN doubles, layout as mx_observables). */
int mx_batch_observables(mx_batch* b, const void* e, double* out, void* stream);

/* ---------------------------------------------------------------------------------------------
 * Differentiable form: reverse-mode adjoint of a deterministic-noise Heun (or Depondt) trajectory with
 * respect to the initial spins and to P <= 16 parameters theta of a linearly parameterized Hamiltonian,
 * J_k = sum_p theta_p U[p][k] on every stencil entry k and A_s = sum_p theta_p Aun[p][s] on every
 * sublattice (shells, DMI and anisotropy constants).  No counterpart in magex, whose solvers are not
 * differentiable (magex/llg.py:87-120 is the forward it differentiates).
 * The adjoint takes over the model's couplings and on-site tensors (the model keeps its stencil,
 * moments, field, pinned sites and dynamics); fixed per-bond and per-site parts come through
 * mx_adjoint_set_offsets.  Pinned spins are carried unchanged, so their gradient is lambda plus the
 * field terms of their neighbours.  Cubic / higher uniaxial terms are held fixed and enter through
 * their Hessians.  Spins and gradients are in magex site order. */
typedef struct mx_adjoint mx_adjoint;

/* unit_bonds: nparam x nent x 9 (row-major 3x3 per stencil entry, the order given to
 * mx_model_set_stencil; an entry and its reverse partner must carry transposed tensors);
 * unit_onsite: nparam x nsub x 9.  checkpoint_every: steps between stored spin states. */
int mx_adjoint_create(mx_model* m, int32_t nparam, const double* unit_bonds, const double* unit_onsite,
                      int32_t checkpoint_every, mx_adjoint** out);
void mx_adjoint_destroy(mx_adjoint* a);
/* theta: nparam values (host); sets the model's couplings and on-site tensors. */
int mx_adjoint_set_params(mx_adjoint* a, const double* theta);
/* Fixed parts of a disordered model (host arrays, copied; either may be null for none): bond_offset
 * (nbonds x ncell x 9, the layout of mx_model_set_bond_tensors) makes every unordered bond's tensor
 * bond_offset_b + sum_p theta_p unit_bonds[p][f(b)], and onsite_offset (N x 9, magex site order) every
 * site's on-site tensor onsite_offset_i + sum_p theta_p unit_onsite[p][s(i)].  Applies at once, as
 * mx_adjoint_set_params does.  No counterpart in magex (the per-bond storage of fast.py:612-614 with
 * parameters on top). */
int mx_adjoint_set_offsets(mx_adjoint* a, const double* bond_offset, const double* onsite_offset);
/* nsteps Heun steps of e (device, N x 3, in place) as mx_llg_run with the same seed and step0,
 * recording checkpoints for mx_adjoint_backward. */
int mx_adjoint_forward(mx_adjoint* a, void* e, int64_t nsteps, uint64_t seed, uint64_t step0, void* stream);
/* mx_adjoint_forward under a time-dependent uniform field, as mx_llg_run_driven: B (host, nsteps x 2 x 3)
 * is the field of the predictor and corrector pass of each step (magex's field_func at t and t + dt,
 * magex/llg.py:129-130, 204-211); kept for the backward's replays. */
int mx_adjoint_forward_driven(mx_adjoint* a, void* e, int64_t nsteps, const double* B, uint64_t seed,
                              uint64_t step0, void* stream);
/* Given grad_e = dL/de_final (device, N x 3), writes dL/de_0 into grad_e0 (device, N x 3, may be null)
 * and dL/dtheta into grad_theta (device, nparam doubles), for the last mx_adjoint_forward.  On grids the
 * GPU holds twice over at once (persistent launches), segment replays run on a second stream that the
 * handle owns (non-blocking, default priority), ordered against `stream` with events: it waits for the
 * work queued on `stream` before the call, and everything it runs is joined back into `stream` before
 * the call's last work there.  Polls that give up count in mx_model_step_info's given-up waits. */
int mx_adjoint_backward(mx_adjoint* a, const void* grad_e, void* grad_e0, double* grad_theta, void* stream);
/* mx_adjoint_backward plus the dense gradients at the current couplings (either may be null):
 * grad_bonds (device, nbonds x ncell x 9 doubles, the layout of mx_model_set_bond_tensors) receives
 * dL/dT for every unordered bond's tensor T (J_ij = T, J_ji = T^T), and grad_onsite (device, N x 9
 * doubles, magex site order) dL/dA_i for every site's on-site tensor.  The antisymmetric part of
 * dL/dT gives dL/dD of the bond's DMI vector (T_ab = eps_abc D_c: dL/dD_c = eps_abc dL/dT_ab).  No
 * counterpart in magex (its solvers do not differentiate; row-major DMI convention of
 * magex/models.py:37). */
int mx_adjoint_backward_dense(mx_adjoint* a, const void* grad_e, void* grad_e0, double* grad_theta,
                              double* grad_bonds, double* grad_onsite, void* stream);
/* mx_adjoint_backward plus dL/dB of the uniform field in every pass of every step of the last forward
 * (driven or not): grad_field (device, nsteps x 2 x 3 doubles; predictor, then corrector pass), the sum
 * over sites of dL/dB_i.  No counterpart in magex (it does not differentiate). */
int mx_adjoint_backward_field(mx_adjoint* a, const void* grad_e, void* grad_e0, double* grad_theta, double* grad_field,
                              void* stream);
/* nbonds: forward stencil entries (s < t, or s == t and R > 0), as mx_model_set_bond_tensors counts them. */
int mx_adjoint_num_bonds(const mx_adjoint* a);

#ifdef __cplusplus
}
#endif
#endif
