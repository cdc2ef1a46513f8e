/* C ABI of the MI355X nested-dissection transmission engine.
 *
 * It replaces, for a caller in another language (the reference's Rust crates would bind it with
 * `extern "C"`), the nested-dissection path of crates/caroli-green/src/nested.rs:
 *   NestedTwoTerminal::from_model          nested.rs:530   -> cg_system_create / cg_preset_create
 *   NestedTwoTerminal::block_count         nested.rs:613   -> cg_block_count
 *   NestedTwoTerminal::transmission        nested.rs:618   -> cg_transmission
 *   NestedTwoTerminal::transmissions       nested.rs:623   -> cg_transmissions
 *   NestedTwoTerminal::self_energies       nested.rs:652   -> cg_self_energies
 *   NestedTwoTerminal::transmission_with   nested.rs:682   -> cg_transmission_with
 *   ::transmission_with_precision          nested.rs:688   -> cg_transmission_with (precision)
 *   ::transmissions_checked                nested.rs:630   -> cg_batch_transmission_checked (per-point dT)
 *   TwoTerminal::transmission              two_terminal.rs:186 -> cg_transmission (same engine)
 * and adds the batched launch over systems x energies (the callers' rayon fan-out,
 * crates/caroli-cli/src/run.rs:113-127) and the adjoint dT/dV (no counterpart in the reference).
 *
 * Conventions: complex numbers are interleaved (re, im) doubles; matrices are column-major;
 * sizes are 64-bit; every function returns a status code (CG_OK or a CG_ERR_* value), whose
 * meaning follows GreenError (crates/caroli-green/src/lib.rs:35-47). Handles are thread-safe for
 * concurrent calls (device work is serialised internally). */
#ifndef CAROLI_AMD_H
#define CAROLI_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  CG_OK = 0,
  CG_ERR_SINGULAR = 1,       /* GreenError::Singular(E): singular root or lead surface state */
  CG_ERR_EIGEN = 2,          /* GreenError::Eigen */
  CG_ERR_MODE_COUNT = 3,     /* GreenError::ModeCount */
  CG_ERR_NO_CONVERGENCE = 4, /* refinement did not converge (mixed precision) */
  CG_ERR_INVALID = 6,        /* malformed arguments (the reference panics, nested.rs:531-534, :694) */
  CG_ERR_DEVICE = 7,         /* HIP error, or no device */
  CG_ERR_CAPACITY = 8        /* an input larger than a fixed work area (cg_front_schur_noalloc) */
};

/* Front precision (crates/caroli-green/src/precision.rs:24-32, plus the refined mixed mode). */
enum {
  CG_DOUBLE = 0, /* complex f64 fronts (Precision::Double) */
  CG_SINGLE = 1, /* complex f32 fronts, unrefined (Precision::Single) */
  CG_MIXED = 2   /* complex f32 fronts with FP64-residual iterative refinement */
};

typedef struct cg_system cg_system; /* NestedTwoTerminal */
typedef struct cg_sigma cg_sigma;   /* LeadSelfEnergies (nested.rs:439-459) */

/* NestedTwoTerminal::from_model (nested.rs:530-610) from the arrays the models produce:
 * the CSR Hamiltonian of the region (crates/caroli-model/src/sparse.rs:13-19; vals interleaved),
 * the sites (cells: 3 int64 per site; orbital count and first orbital per site), the lattice
 * dimension, the hopping range per axis, the transport axis, and the lead cells of the left
 * (first slice) and right (last slice) leads, nl x nl and nr x nr: h00 = <cell|H|cell> and
 * h01 = <cell|H|next cell outward>, i.e. cell - axis for the left lead and cell + axis for the
 * right one (the Lead pairs of nested.rs:543-546). */
int32_t cg_system_create(uint64_t n, const int64_t* row_ptr, const uint32_t* cols, const double* vals,
                         uint64_t nsites, const int64_t* cells, const int32_t* site_orbitals,
                         const uint64_t* site_offsets, int32_t dim, const int64_t* range, int32_t axis,
                         const double* left_h00, const double* left_h01, uint64_t nl,
                         const double* right_h00, const double* right_h01, uint64_t nr, cg_system** out);

/* Host builders of named workloads (models of crates/caroli-model/src/presets.rs and
 * crates/caroli-bench/src/main.rs:77-92, disorder seeded as crates/caroli-cli/src/build.rs:251-306).
 * model: 0 BHZ, 1 square, 2 zigzag graphene, 3 cubic.  disorder: 0 none, 1 Anderson (strength =
 * width w), 2 polarized magnetic impurities, opposite g (strength = J, concentration).
 * site_potential: optional extra on-site potential per site (null for none). */
int32_t cg_preset_create(int32_t model, const uint64_t* extent, int32_t axis, int32_t disorder, double strength,
                         double concentration, uint64_t seed, const double* site_potential, cg_system** out);

/* BHZ ribbon (BhzParams::default(), axis 0, clean BHZ leads) of extent[0] x extent[1] sites with
 * classical magnetic impurities, as MagneticImpurities::generate (crates/caroli-disorder/src/lib.rs:
 * 18-100) and the thesis setups (crates/caroli-validation/src/thesis.rs:11-16, :66-78) build them:
 * int(concentration * nx_imp * extent[1]) distinct sites among x0 <= x < x0 + nx_imp, all y,
 * listed with x fastest, from Xoshiro256::seed_from_u64(seed) (seed_mode 0) or
 * Xoshiro256::gfortran(seed) (seed_mode 1); moments 0 isotropic, 1 in a cone of half-angle
 * moment_param (radians) around z, 2 polarized with mean m_z = moment_param; coupling 0 opposite g,
 * 1 same g (bhz_spin, crates/caroli-model/src/presets.rs:60-74); exchange J (n . S) per site. */
int32_t cg_bhz_magnetic_create(const uint64_t* extent, int32_t seed_mode, int64_t seed, int64_t x0, int64_t nx_imp,
                               double concentration, int32_t moments, double moment_param, int32_t coupling,
                               double exchange, cg_system** out);

/* One [[disorder]] table of the CLI input (crates/caroli-cli/src/input.rs:131-170). */
typedef struct cg_disorder {
  int32_t kind;          /* 0 anderson, 1 magnetic */
  double strength;       /* anderson: on-site energies uniform in [-W/2, W/2] */
  double concentration;  /* magnetic: fraction of the candidate sites carrying a moment */
  double exchange;       /* magnetic: J of J n . S */
  int32_t moments;       /* 0 isotropic, 1 cone (moment_param = theta0, radians), 2 polarized (moment_param = mz) */
  double moment_param;
  int32_t spin;          /* S: 0 opposite-g, 1 same-g (four-orbital BHZ sites), 2 Pauli (two-orbital sites) */
  int32_t has_cells[3];  /* direction d restricted to cells[d][0] <= cell < cells[d][1] */
  int64_t cells[3][2];
  uint64_t sublattices;  /* bit mask of the sublattices carrying it (0: all) */
  uint64_t seed;         /* configuration c draws from Xoshiro256::seed_from_u64(seed + c) */
} cg_disorder;

/* The CLI's system of configuration `config` (crates/caroli-cli/src/build.rs:251-310, run.rs:63-68):
 * preset `model` (as cg_preset_create) on Region::cuboid(extent), every disorder added in order,
 * leads the clean preset along `axis`. */
int32_t cg_disordered_create(int32_t model, const uint64_t* extent, int32_t axis, int64_t count,
                             const cg_disorder* disorders, uint64_t config, cg_system** out);

void cg_system_destroy(cg_system* s);

/* The random stream of the builders (crates/caroli-disorder/src/rng.rs): `count` uniforms in
 * [0, 1) of xoshiro256** seeded by mode 0 `Xoshiro256::seed_from_u64(seed)`, mode 1
 * `Xoshiro256::gfortran(seed)`, or mode 2 `gfortran_put(put)` with put = 8 int32 values. */
int32_t cg_uniforms(int32_t mode, uint64_t seed, const int32_t* put, int64_t count, double* out);

/* NestedTwoTerminal::block_count (nested.rs:613). */
int32_t cg_block_count(const cg_system* s);

/* Sizes: n orbitals, nnz, number of sites; per block b: nl, nr, tree nodes, tree levels. */
int32_t cg_system_sizes(const cg_system* s, uint64_t* n, uint64_t* nnz, uint64_t* nsites);
int32_t cg_block_sizes(const cg_system* s, int32_t b, uint64_t* nl, uint64_t* nr, uint64_t* nodes, uint64_t* levels);

/* Copies of the system arrays (any pointer may be null): CSR, geometry, lead cells. */
int32_t cg_system_arrays(const cg_system* s, int64_t* row_ptr, uint32_t* cols, double* vals, int64_t* cells,
                         int32_t* site_orbitals, uint64_t* site_offsets, int32_t* dim, int64_t* range,
                         int32_t* axis, double* left_h00, double* left_h01, double* right_h00, double* right_h01);

/* Node k of block b's elimination tree: counts, and the sorted orbitals when arrays are given. */
int32_t cg_tree_node(const cg_system* s, int32_t b, int32_t k, uint64_t* ne, uint64_t* nb, uint64_t* eliminate,
                     uint64_t* boundary);

/* Algorithmic work of one transmission point (complex MAC = 8 flops, 16-byte complex):
 * flops of the fronts and root, and bytes of one assembly + one Schur write/read per front. */
int32_t cg_work(const cg_system* s, double* flops, double* bytes);

/* NestedTwoTerminal::self_energies (nested.rs:652-676): FP64 mode matching on the host. */
int32_t cg_self_energies(const cg_system* s, double energy, cg_sigma** out);
/* SelfEnergyMethod::ModesBroadened { eta } (crates/caroli-green/src/two_terminal.rs:51-53, :157-172):
 * the leads' modes at E + i eta, the device staying at the real energy E (the thesis codes). */
int32_t cg_self_energies_broadened(const cg_system* s, double energy, double eta, cg_sigma** out);
/* Self-energies given by the caller (per block: Sigma_L nl x nl then Sigma_R nr x nr). */
int32_t cg_sigma_from_arrays(const cg_system* s, double energy, const double* sigmas, cg_sigma** out);
/* Copies Sigma_L and Sigma_R of block b (either may be null); channels = propagating modes. */
int32_t cg_sigma_block(const cg_sigma* g, int32_t b, double* sigma_l, double* sigma_r, int32_t* channels_l,
                       int32_t* channels_r);
double cg_sigma_energy(const cg_sigma* g);
void cg_sigma_destroy(cg_sigma* g);

/* NestedTwoTerminal::transmission (nested.rs:618) on the GPU (self-energies on the host). */
int32_t cg_transmission(const cg_system* s, double energy, int32_t precision, double* out_t);
/* transmission_with / transmission_with_precision (nested.rs:682-713); panics of the reference
 * on foreign self-energies (nested.rs:694, :701-707) become CG_ERR_INVALID. */
int32_t cg_transmission_with(const cg_system* s, const cg_sigma* g, int32_t precision, double* out_t);
/* NestedTwoTerminal::transmissions (nested.rs:623-626): one batched launch. */
int32_t cg_transmissions(const cg_system* s, const double* energies, int64_t count, int32_t precision,
                         double* out_t, int32_t* out_status);

/* Statistics of the last batched call on this thread. */
typedef struct cg_batch_stats {
  double seconds;           /* wall time of the device work (inputs resident) */
  double front_flops;       /* algorithmic flops of the batch */
  double front_bytes;       /* algorithmic bytes of the batch */
  int64_t delayed_points;   /* points redone with delayed pivots */
  int64_t refinement_iters; /* mixed precision: total refinement iterations */
  double max_refine_resid;  /* mixed precision: largest final relative correction */
  int64_t recomputed_points; /* mixed precision: points recomputed in complex f64 (their f32
                                fronts delayed or their refinement did not converge) */
  int64_t goal_points;      /* mixed precision: points accepted by the dual-weighted error bound
                               on T before their corrections fell below the relative tolerance */
} cg_batch_stats;

/* The batched launch: point p is (systems[p], sigmas[p]); all energies x configurations (x
 * systems of several shapes) run in one call, level by level over the elimination trees.
 * out_t[p] = T or NaN; out_status[p] = status code. */
int32_t cg_batch_transmission(const cg_system* const* systems, const cg_sigma* const* sigmas, int64_t count,
                              int32_t precision, double* out_t, int32_t* out_status, cg_batch_stats* stats);

/* Checked reduced precision (PrecisionOptions / PrecisionReport / checked,
 * crates/caroli-green/src/precision.rs:36-134; NestedTwoTerminal::transmissions_checked,
 * nested.rs:630-641): every point runs in options->precision; the points that failed there and
 * one point in `every` (at least 1; 1 checks every point) are recomputed in complex f64, their
 * |T - T_double| written to out_dt (NaN for points not compared) and their double values returned;
 * if the largest difference exceeds the tolerance (or a point failed in double only), the whole
 * batch is recomputed in complex f64 and report->fell_back is set. stats: the first run's. */
typedef struct cg_precision_options {
  int32_t precision; /* CG_DOUBLE, CG_SINGLE or CG_MIXED */
  double tolerance;  /* largest accepted |T - T_double| */
  int64_t every;     /* one point in `every` is compared */
} cg_precision_options;
typedef struct cg_precision_report {
  int32_t precision; /* precision of the values returned: CG_DOUBLE after a fallback */
  int64_t checked;   /* points recomputed in complex f64 and compared */
  double median;     /* median and largest difference among them */
  double max;
  int64_t failed;    /* points that failed in the reduced precision, computed in complex f64 */
  int32_t fell_back; /* 1: the whole batch was recomputed in complex f64 */
  char fallback[160]; /* why (empty when it was not) */
} cg_precision_report;
int32_t cg_batch_transmission_checked(const cg_system* const* systems, const cg_sigma* const* sigmas, int64_t count,
                                      const cg_precision_options* options, double* out_t, int32_t* out_status,
                                      double* out_dt, cg_precision_report* report, cg_batch_stats* stats);

/* The front update on the GPU, as caroli_dense::Field::schur (crates/caroli-dense/src/field.rs:44):
 * S = F22 - F21 F11^-1 F12 of the leading ne x ne block of the n x n front f (column-major), with
 * the pivot test of field.rs:61 / :111 on the pivots of columns [0, ne1) (ne1 = ne for a front).
 * precision CG_DOUBLE or CG_SINGLE. out_s receives (n - ne)^2 values; *out_status is 0, or 5 when
 * the block is refused (numerically singular: the reference returns None). */
int32_t cg_front_schur(const double* f, int64_t n, int64_t ne, int64_t ne1, int32_t precision, double* out_s,
                       int32_t* out_status);

/* The front update of cg_front_schur without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Same arguments, results and kernels as
 * cg_front_schur; its work area (the front, staging, pivots, descriptors and tiles) is one
 * static device array of CG_FRONT_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds
 * until the process ends. It calls no hipMalloc / hipFree. Fronts of more than
 * CG_FRONT_NOALLOC_MAX_N variables are refused with CG_ERR_CAPACITY. All calls on one GPU share
 * that GPU's array (each call clears what it needs and keeps the last call's contents): calls
 * that overlap on one GPU race (calls through this library are serialised by its lock). */
#define CG_FRONT_NOALLOC_MAX_N 1024
#define CG_FRONT_NOALLOC_SCRATCH_BYTES (2u * 1024u * 1024u * 16u + (1u << 20))
int32_t cg_front_schur_noalloc(const double* f, int64_t n, int64_t ne, int64_t ne1, int32_t precision, double* out_s,
                               int32_t* out_status);

/* A resident batch: the inputs of the points uploaded once to HBM, for repeated launches
 * (serving loops, benchmarks). cg_batch_run computes T for every point from the resident
 * inputs; the systems and self-energies must outlive the batch. */
typedef struct cg_batch cg_batch;
int32_t cg_batch_create(const cg_system* const* systems, const cg_sigma* const* sigmas, int64_t count,
                        int32_t precision, cg_batch** out);
int32_t cg_batch_run(cg_batch* batch, double* out_t, int32_t* out_status, cg_batch_stats* stats);
void cg_batch_destroy(cg_batch* batch);

/* Kernel choices per front class (no counterpart in the reference, which has one CPU path per
 * front; the choices mirror its split between small_schur, nested.rs:1167, and Field::schur,
 * field.rs:79, at GPU granularity). Defaults: 0, 1, 128, 0; the environment variables
 * CAROLI_SMALL_FRONTS, CAROLI_FRONT_KERNEL and CAROLI_BLOCKED_MIN_NE change them. */
typedef struct cg_tuning {
  int32_t small_fronts;   /* fronts up to 64 (c128) / 90 (c64) variables: 0 the border in LDS,
                             1 the whole front in LDS, 2 in registers (n <= 64) */
  int32_t split;          /* 1: separators up to 64 by an LDS LU and register triangular solves
                             spread over many workgroups; 0: one workgroup per front */
  int32_t blocked_min_ne; /* separators above this many variables by the blocked kernels */
  int32_t unblocked;      /* 1: every large front by rank-1 updates in HBM */
  int32_t schur_tile;     /* Schur-complement MFMA tiles: 64 (64 x 64, four waves), 32 (32 x 32,
                             one wave) or 0 (by level: 32 where its Schur blocks have at most
                             256 rows) */
} cg_tuning;
/* The tuning in force, or set one (NULL: back to the defaults and the variables). Plans built
 * afterwards follow it; a system planned under another is re-planned at its next use. */
int32_t cg_get_tuning(cg_tuning* out);
int32_t cg_set_tuning(const cg_tuning* t);
/* Autotuner (no counterpart in the reference): times whole eager sweeps of the batch's points
 * (the faster of two per candidate, after one that settles) under candidate tunings, by coordinate descent over the four choices from the one in force (a
 * candidate must be 2 % faster to win), and leaves the fastest in force with the batch re-planned
 * under it. out: the choice; out_seconds: its sweep time (either may be NULL). */
int32_t cg_batch_autotune(cg_batch* batch, cg_tuning* out, double* out_seconds);

/* Per-kernel timing with HIP events on the engine's streams (enable = 1 resets the counters).
 * Each record: kernel family name, launches, total seconds, algorithmic flops and bytes. While
 * timing is on, the launch units of a batch run one after another instead of two at a time, so
 * that every timed launch has the GPU to itself. */
typedef struct cg_kernel_stat {
  char name[48];
  int64_t launches;
  double seconds;
  double flops;
  double bytes;
} cg_kernel_stat;
int32_t cg_kernel_timing(int32_t enable);
int32_t cg_kernel_stats(cg_kernel_stat* out, int32_t max_records);

/* Adjoint: T and dT/d(eps_a) for every orbital a of the device (n values, on-site energies),
 * from one backward sweep over the stored fronts (no counterpart in the reference): psi = G E_R W_R
 * and phi = G^T E_L conj(W_L) by the root and backward solves, then t = W_L^H psi_L and
 * dT/d(eps_a) = 2 Re sum_cd psi_ac conj(t_dc) phi_ad on the GPU (one thread per orbital). */
int32_t cg_backward(const cg_system* s, const cg_sigma* g, double* out_t, double* out_grad);
/* The adjoint of `count` points (systems[p], sigmas[p]) of one structure (the same lattice, sizes
 * and leads; energies and disorder free) in one sweep: the factors of every point kept, psi and
 * phi solved for all of them, the contraction on the GPU. out_t[p]; out_grad: count x n, point
 * after point; out_status[p] as cg_backward's status for that point (its T and gradient NaN when
 * not CG_OK). Points are processed in chunks that keep the stored factors within 64 GB. */
int32_t cg_backward_batch(const cg_system* const* systems, const cg_sigma* const* sigmas, int64_t count, double* out_t,
                          double* out_grad, int32_t* out_status);

/* Scattering states (NestedTwoTerminal::scattering_states, crates/caroli-green/src/nested.rs:851):
 * psi_p = G W_p on every device orbital for each lead p (W_p W_p^H = Gamma_p, one column per
 * channel), the blocks' columns one after another. cols[0], cols[1]: channels of the left and
 * right lead (always written). psi_left (n x cols[0]) and psi_right (n x cols[1]): interleaved
 * complex, column-major; with both NULL only the counts are returned. CG_ERR_INVALID for a point
 * whose separators delayed (the solves follow the static plan). */
int32_t cg_scattering_states(const cg_system* s, const cg_sigma* g, int64_t* cols, double* psi_left,
                             double* psi_right);
/* Site densities (site_densities, nested.rs:994-1025): sum_p f[p] psi_p^H O psi_p / 2 pi on the
 * orbitals of every site (sites in the order the system was created with); op: NULL for O = 1, else
 * an op_dim x op_dim complex matrix (interleaved, column-major) for sites of op_dim orbitals. */
int32_t cg_site_densities(const cg_system* s, const double* psi_left, const double* psi_right, const int64_t* cols,
                          const double* f, const double* op, int32_t op_dim, double* out);
/* Bond currents (bond_currents, nested.rs:930-990): 2 Re sum_{a in i, b in j} H_ab G^<_ba with
 * G^< = i sum_p f[p] psi_p psi_p^H, every coupled pair of sites i < j once, ordered by (i, j). With
 * from / to / current NULL, *count receives the number of bonds; else *count is the capacity on entry
 * and the number written on return. */
int32_t cg_bond_currents(const cg_system* s, const double* psi_left, const double* psi_right, const int64_t* cols,
                         const double* f, int64_t* count, int64_t* from, int64_t* to, double* current);
/* Currents of an on-site operator O (operator_currents, nested.rs:925-928, :958-985; a spin current
 * for O = s_z): Re Tr[(O H_ij + H_ij O) G^<_ji] for every coupled pair of sites i < j, with every
 * site of op_dim orbitals (else CG_ERR_INVALID); op: op_dim x op_dim complex, interleaved,
 * column-major. count / from / to / current as for cg_bond_currents. */
int32_t cg_operator_currents(const cg_system* s, const double* psi_left, const double* psi_right, const int64_t* cols,
                             const double* f, const double* op, int32_t op_dim, int64_t* count, int64_t* from,
                             int64_t* to, double* current);

/* Device present and usable (1) or not (0). */
int32_t cg_device_available(void);
/* The message behind the last CG_ERR_DEVICE / CG_ERR_INVALID / CG_ERR_EIGEN that a call of this
 * thread turned an exception into (the HIP call and its error string for a device error), or "".
 * Valid until this thread's next call. */
const char* cg_last_error(void);

/* ---- NestedMultiTerminal (crates/caroli-green/src/nested_multi.rs) ----
 * A device with any number of leads. The sites a lead couples to stay at the root of the
 * elimination tree, whose bordered root front leaves -G on all of their orbitals; T[p][q] =
 * Tr[Gamma_p G_pq Gamma_q G_pq^H] from lead q into lead p. A multi-terminal system is a cg_system
 * handle (cg_block_count, cg_system_sizes, cg_system_arrays, cg_work, cg_tree_node and
 * cg_system_destroy apply); the two-terminal calls refuse it and its self-energies
 * (CG_ERR_INVALID). Precision: CG_DOUBLE or CG_SINGLE (the fronts in complex f32, the root in f64).
 *
 *   NestedMultiTerminal::from_model         nested_multi.rs:60   -> cg_multi_create / cg_multi_model_create
 *   NestedMultiTerminal::lead_count         nested_multi.rs:164  -> cg_multi_lead_count
 *   NestedMultiTerminal::transmissions      nested_multi.rs:174  -> cg_multi_transmissions
 *   ::transmissions_with_precision          nested_multi.rs:180  -> cg_multi_batch_transmissions
 *   ::sector_transmissions                  nested_multi.rs:246  -> cg_multi_self_energies (sectors)
 *                                                                   + cg_multi_batch_transmissions
 *   ::scattering_states                     nested_multi.rs:354  -> cg_multi_scattering_states
 *   ::smatrix                               nested_multi.rs:440  -> cg_multi_smatrix
 *   ::site_densities / bond_currents /      nested_multi.rs:497  -> cg_multi_site_densities /
 *     operator_currents                                             cg_multi_bond_currents */

/* One lead as attached (AttachedLead, crates/caroli-green/src/multi.rs:15-25): a cell of `cell`
 * orbitals with h00 and h01 (the coupling from a cell to the next one further from the device),
 * the `coupled` device orbitals its first cell reaches (increasing) and that coupling (coupled x
 * cell). kinds: optional, the index within its site of each cell orbital (sector labels; NULL:
 * orbital k is kind k). */
typedef struct {
  uint64_t cell;
  const double* h00;
  const double* h01;
  uint64_t coupled;
  const uint64_t* orbitals;
  const double* coupling;
  const int32_t* kinds;
} cg_lead;

/* NestedMultiTerminal::from_model with the device and its leads built by the caller: the device
 * arrays as for cg_system_create (no axis), then the leads. */
int32_t cg_multi_create(uint64_t n, const int64_t* row_ptr, const uint32_t* cols, const double* vals,
                        uint64_t nsites, const int64_t* cells, const int32_t* site_orbitals,
                        const uint64_t* site_offsets, int32_t dim, const int64_t* range, int32_t leads,
                        const cg_lead* lead, cg_system** out);

/* One interface (Interface, multi.rs:66-72): the device sites of the slice (cells, 3 per site, and
 * sublattices) and the lattice step pointing away from the device. */
typedef struct {
  uint64_t sites;
  const int64_t* cells;
  const int32_t* subs;
  int64_t outward[3];
} cg_interface;

/* NestedMultiTerminal::from_model(device, lead, region, interfaces) with a preset model (0 BHZ,
 * 1 square, 2 zigzag graphene, 3 cubic) as the clean lead: region = the listed sites in that order;
 * the device adds site_potential[i] (optional) on every orbital of site i; the leads continue each
 * interface outward (attach_leads, multi.rs:112-140). */
int32_t cg_multi_model_create(int32_t model, uint64_t sites, const int64_t* cells, const int32_t* subs,
                              const double* site_potential, int32_t interfaces, const cg_interface* itf,
                              cg_system** out);
/* The CLI's multi-terminal system of configuration `config` (crates/caroli-cli/src/build.rs:251-310,
 * :371-413, run.rs:130-138): the preset model on Region::cuboid(extent) with every disorder added
 * (as cg_disordered_create), and a clean lead on each face (face_sign[k] = -1 or +1 along face_axis[k];
 * the interface is the region's sites on that face). */
int32_t cg_multi_disordered_create(int32_t model, const uint64_t* extent, int32_t faces, const int32_t* face_sign,
                                   const int32_t* face_axis, int64_t count, const cg_disorder* disorders,
                                   uint64_t config, cg_system** out);

/* Number of leads of a multi-terminal system (-1 for another handle). */
int32_t cg_multi_lead_count(const cg_system* s);
/* Lead p as attached: its cell size and coupled orbitals, and (each optional) h00, h01 (cell x cell),
 * the orbitals and the coupling (coupled x cell), complex interleaved, column-major. */
int32_t cg_multi_lead(const cg_system* s, int32_t p, uint64_t* cell, uint64_t* coupled, double* h00, double* h01,
                      uint64_t* orbitals, double* coupling);

/* Self-energies and broadenings of every lead at E + i eta (eta = 0: mode matching on the real
 * axis; nested_multi.rs:196-232, SelfEnergyMethod::ModesBroadened otherwise). The rows and columns
 * of T are labels (lead, sector): without sectors (sector_len NULL) one label per lead; else lead p
 * with sector_len[p] > 0 is split by the tag sector_tags[...][kind] of each cell orbital (tags of
 * all leads concatenated; sector_transmissions, nested_multi.rs:246-313), and a lead Hamiltonian that
 * couples two sectors is refused. */
int32_t cg_multi_self_energies(const cg_system* s, double energy, double eta, const int32_t* sector_len,
                               const int32_t* sector_tags, cg_sigma** out);
/* Sigma_p = V g V^H of lead p on its coupled orbitals (coupled x coupled, complex interleaved,
 * column-major; zero between orbitals of different blocks). */
int32_t cg_multi_sigma(const cg_system* s, const cg_sigma* g, int32_t p, double* out);
/* The labels of T: *count, then (each optional) the lead and sector of each, in increasing order. */
int32_t cg_multi_labels(const cg_sigma* g, int32_t* count, int32_t* lead, int32_t* sector);
/* transmissions_with_precision for `count` points (systems[p], sigmas[p], every sigma with the same
 * number of labels L) in one batched launch: out receives count x L x L values, point by point,
 * row-major: out[p][a][b] = T from label b into label a (the diagonal is not a reflection,
 * nested_multi.rs:246-252). */
int32_t cg_multi_batch_transmissions(const cg_system* const* systems, const cg_sigma* const* sigmas, int64_t count,
                                     int32_t precision, double* out, int32_t* out_status, cg_batch_stats* stats);
/* NestedMultiTerminal::transmissions at one energy: out is lead_count x lead_count, row-major. */
int32_t cg_multi_transmissions(const cg_system* s, double energy, int32_t precision, double* out);

/* NestedMultiTerminal::smatrix (nested_multi.rs:440-470, crates/caroli-green/src/smatrix.rs): the
 * scattering matrix at `energy` in the basis of the propagating lead modes normalised to unit
 * current, from one complex f64 sweep per block on the GPU (G on the kept orbitals) and the leads'
 * modes on the host. modes (2 x lead_count) receives the outgoing then the incoming mode count of
 * each lead; full (optional, sum(out) x sum(in), complex, column-major): rows the outgoing modes of
 * lead 0, 1, ..., columns the incoming ones, so |block p,q|^2 sums to T[p][q] and full is unitary;
 * lambda_out / lambda_in (optional): their Bloch factors. With all three NULL only the counts (no
 * device work). CG_ERR_INVALID where separators delayed. */
int32_t cg_multi_smatrix(const cg_system* s, double energy, int32_t* modes, double* full, double* lambda_out,
                         double* lambda_in);
/* NestedMultiTerminal::scattering_states (nested_multi.rs:354-413): psi = G W on every orbital for
 * the channels of each label of g (W W^H = Gamma of the label), the blocks' columns one after
 * another. cols (one per label) receives the channel counts; with psi non-NULL, psi[l] (n x cols[l],
 * complex, column-major) the states. CG_ERR_INVALID where separators delayed. */
int32_t cg_multi_scattering_states(const cg_system* s, const cg_sigma* g, int64_t* cols, double* const* psi);
/* Site densities and bond currents of multi-terminal states (nested_multi.rs:497-513, as
 * cg_site_densities / cg_bond_currents / cg_operator_currents with `leads` state arrays psi[p] of
 * cols[p] columns and occupations f[p]); op NULL for the charge current. */
int32_t cg_multi_site_densities(const cg_system* s, int32_t leads, const double* const* psi, const int64_t* cols,
                                const double* f, const double* op, int32_t op_dim, double* out);
int32_t cg_multi_bond_currents(const cg_system* s, int32_t leads, const double* const* psi, const int64_t* cols,
                               const double* f, const double* op, int32_t op_dim, int64_t* count, int64_t* from,
                               int64_t* to, double* current);

#ifdef __cplusplus
}
#endif
#endif
