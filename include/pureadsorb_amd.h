/*
 * pureadsorb_amd.h -- C ABI of the MI355X (gfx950) Widom insertion backend.
 *
 * A drop-in backend for PureAdsorb.jl's Widom path: the Julia (or Python, or C++) host builds
 * the same inputs it gives `FrameworkBatch` and `widom`, and calls these entry points instead
 * of launching the KernelAbstractions kernels. Plain pointers and sizes only; every entry
 * point returns a status (PA_OK = 0) and leaves a thread-local message for pa_last_error().
 *
 * Conventions: indices are 0-based (the Julia shim subtracts 1); 3x3 matrices are
 * column-major with lattice vectors as columns (StaticArrays' SMatrix layout), so Julia can
 * pass `SMatrix{3,3,Float64}` and `Vector{SVector{3,Float64}}` memory unchanged.
 */
#ifndef PUREADSORB_AMD_H
#define PUREADSORB_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. PA_ERR_ARG corresponds to the reference's ArgumentError, PA_ERR_DIM to its
 * DimensionMismatch; messages keep the reference's wording where its tests match substrings. */
#define PA_OK 0
#define PA_ERR_ARG 1
#define PA_ERR_DIM 2
#define PA_ERR_DEVICE 3
#define PA_ERR_INTERNAL 4

/* Precision of pa_widom / pa_widom_energies. */
#define PA_PREC_F64 0   /* FP64 plain path: one work-item per insertion, every pair in FP64 */
#define PA_PREC_MIXED 1 /* FP32 pair + k-space arithmetic, FP64 accumulation and tables */
/* OR-ed into the precision of pa_widom / pa_widom_grad: skip the mixed path's phase-0 hard-core
 * rejection (the reference's widom_singlephase, src/widom.jl:298). Without it the mixed path
 * first rejects every insertion whose sites come closer than rho_at to an atom of type t, with
 * radii from the rigorous bound of src/reject.jl, so that its Boltzmann weight is exactly 0.0 in
 * FP64; results are identical either way. The FP64 plain path always evaluates every insertion. */
#define PA_SINGLE_PHASE 0x100u

/* pa_batch_create flags */
#define PA_BATCH_FULLK 1u     /* keep every k-vector (FrameworkBatch(...; fullk = true)) */
#define PA_BATCH_HOST_ONLY 2u /* build and validate the tables on the host only: no device is
                                 touched, and pa_widom* on the batch fail */

/* One host framework: src/structure.jl:14-24 (`Framework{T}`). `types` are force-field type
 * indices (the shim maps `symbol * "_"` through `typeindex`, src/batch.jl:309). */
typedef struct pa_framework {
    double cell[9];          /* Å, column-major, lattice vectors as columns */
    int32_t natoms;
    const double *frac;      /* natoms x 3 fractional coordinates */
    const int32_t *types;    /* natoms, 0-based force-field type index */
    const double *charges;   /* natoms, e */
    int32_t replication[3];  /* Framework.replication */
} pa_framework;

/* src/forcefield.jl:8-14 (`ForceField{T}`): Lorentz-Berthelot tables already mixed. */
typedef struct pa_forcefield {
    int32_t ntypes;
    const double *sigma;     /* ntypes x ntypes, column-major, Å */
    const double *epsilon;   /* ntypes x ntypes, column-major, eV */
    double cutoff;           /* LJ cutoff, Å */
    int32_t tail;            /* analytic tail correction on/off */
} pa_forcefield;

/* src/forcefield.jl:76-83 (`Guest{T,N}`), geometry and charges only. */
typedef struct pa_guest {
    int32_t nsites;          /* 1..16 */
    const double *sites;     /* nsites x 3, Å, relative to the reference point */
    const int32_t *types;    /* 0-based force-field type index */
    const double *charges;   /* e */
} pa_guest;

/* src/ewald.jl:8-18 (`EwaldParams{T}`). */
typedef struct pa_ewald {
    double cutoff;           /* real-space cutoff, Å */
    double precision;        /* relative precision (default 1e-6) */
} pa_ewald;

/* src/widom.jl:19-28 (`WidomResult`), plus ln_K_H (ln of K_H in Å³/eV), which stays finite
 * when K_H itself overflows Float64. */
typedef struct pa_widom_result {
    double mu_ex, mu_ex_err;
    double K_H, K_H_err;
    double q_st, q_st_err;
    int64_t nsamples;
    int64_t nblocks;
    double ln_K_H;
} pa_widom_result;

/* Per-framework scalars of a built batch (the deduplicated storage of `FrameworkBatch`). */
typedef struct pa_framework_info {
    int32_t natoms;
    int32_t nk;                   /* k-vectors kept (coupled subset, or all with FULLK) */
    int32_t nk_full;              /* k-vectors `kvectors` enumerates for the stored cell */
    double volume;                /* Å³ */
    double alpha;                 /* Ewald splitting parameter, 1/Å */
    double constant_offset;       /* eV, src/batch.jl:396-407 */
    double self_term_halfrange;   /* eV, src/batch.jl:376 */
    double hardcore_bound;        /* eV, B_s of src/reject.jl:101-131 (per-atom terms rounded up on a
                                     1.02-ratio grid of |K|, so never below the reference's value);
                                     +inf disables phase-0 rejection for the framework */
} pa_framework_info;

/* Device timing of the last pa_widom call on a batch, from HIP events on the library's
 * stream. The energy kernel figures cover only the launches of the energy kernel. */
typedef struct pa_timing {
    double total_ms;              /* first launch to last event of the call */
    double energy_ms;             /* summed duration of the energy-kernel launches */
    int64_t energy_launches;
    double reduce_ms;             /* Boltzmann-sum reduction kernels */
    int64_t insertions;
    double binning_ms;            /* mixed path: pose, bin count, scan and scatter launches */
    double grad_ms;               /* pa_widom_grad: weighted-insertion selection and the gradient kernels */
    int64_t grad_insertions;      /* pa_widom_grad: insertions with a weight above e^-40 of the best */
} pa_timing;

typedef struct pa_batch pa_batch;

/* Replaces `FrameworkBatch(fws, ff, guest, ewald; fullk)` (src/batch.jl:237-432): validates,
 * deduplicates frameworks by value, builds every table in FP64 (k-vectors, prefactors, host
 * structure factors, replication check, orientation-averaged guest self term,
 * constant_offset; the structure factors and the self term on the device, or on the host with
 * PA_BATCH_HOST_ONLY) and uploads them once. Input arrays are copied. */
int pa_batch_create(const pa_framework *fws, int32_t nsys, const pa_forcefield *ff, const pa_guest *guest,
                    const pa_ewald *ewald, uint32_t flags, pa_batch **out);
void pa_batch_destroy(pa_batch *batch);

/* A FrameworkBatch the host has already built, as its plain arrays (src/batch.jl:72-104), indices
 * 0-based. Per-framework arrays hold nframeworks entries (one per distinct framework); framework
 * f's atoms are atom_offsets[f] .. atom_offsets[f+1]-1 of positions / types / charges (any order,
 * e.g. the reference's cell-sorted one) and its k-vectors k_offsets[f] .. k_offsets[f+1]-1 of ks /
 * kprefactor / Shost. Volumes and inverse cells are derived from the cells. */
typedef struct pa_batch_tables {
    int32_t nframeworks;
    int32_t nsys;
    const int32_t *framework_of;       /* nsys: framework of each system */
    const int32_t *atom_offsets;       /* nframeworks + 1, atom_offsets[0] = 0 */
    const double *positions;           /* natoms x 3, Cartesian Å */
    const int32_t *types;              /* natoms: compact LJ type */
    const double *charges;             /* natoms, e */
    const double *cells;               /* nframeworks x 9: Å, column-major, lattice vectors as columns */
    const double *alphas;              /* nframeworks: Ewald splitting parameter, Å^-1 */
    const int32_t *k_offsets;          /* nframeworks + 1, k_offsets[0] = 0 */
    const double *ks;                  /* nk x 3: k-vectors, Å^-1 (reciprocal-lattice vectors of the cell) */
    const double *kprefactor;          /* nk */
    const double *Shost;               /* nk x 2: real and imaginary part of the host structure factor */
    const double *constant_offset;     /* nframeworks, eV */
    const double *self_term_halfrange; /* nframeworks, eV */
    int32_t ntypes;                    /* compact LJ types, 1..64 */
    const double *sigma;               /* ntypes x ntypes, Å, column-major */
    const double *epsilon;             /* ntypes x ntypes, eV, column-major */
    const int32_t *compact_to_orig;    /* ntypes: force-field type of each compact type */
    const double *bs;                  /* nframeworks: hard-core bound B_s, eV (+inf: no phase-0 rejection) */
    const double *kmin;                /* nframeworks x nsites x ntypes: K_min(site, type), eV Å */
    double cutoff;                     /* LJ cutoff, Å */
    double ewald_cutoff;               /* Å */
    int32_t fullk;                     /* 1: the k tables hold every k-vector, 0: the coupled subset */
} pa_batch_tables;

/* Takes over a FrameworkBatch built by the reference (src/batch.jl:237-432), so that
 * `widom(batch::FrameworkBatch, guest; backend)` (src/widom.jl:255-301) runs on the device
 * without rebuilding the batch: the tables are copied as they are (constant_offset, B_s, K_min
 * included) and uploaded once. `guest` carries force-field type indices (guest_types_orig), each
 * of which must appear in compact_to_orig. The batch serves pa_widom, pa_widom_noalloc,
 * pa_widom_energies, pa_hardcore_flags and pa_widom_poses; pa_widom_grad and pa_state_create need
 * the force field itself (per-type parameters, tail flag) and refuse it. */
int pa_batch_from_tables(const pa_batch_tables *tables, const pa_guest *guest, pa_batch **out);

/* nframeworks(batch) (src/batch.jl:123) and framework_of (src/batch.jl:103), 0-based. */
int pa_batch_nframeworks(const pa_batch *batch, int32_t *nframeworks, int32_t *nsys);
int pa_batch_framework_of(const pa_batch *batch, int32_t *framework_of /* nsys */);
int pa_batch_framework_info(const pa_batch *batch, int32_t fw, pa_framework_info *out);

/* Replaces `widom(batch, guest; T, ninsert, seed, chunk, nblocks, run, backend)`
 * (src/widom.jl:255-301, :298-442). `guest` must equal the guest the batch was built with.
 * run = 0 selects default_run; chunk = 0 selects the library's chunk. Insertion j (0-based)
 * of system s draws its pose from ChainRNG(splitmix64(seed, s+1), j+1, 0) (src/rng.jl,
 * src/moves.jl:154-161), so results do not depend on chunk or run apart from run's effect on
 * the per-system sample counts (system_counts, src/widom.jl:44-53). out: nsys entries. */
int pa_widom(pa_batch *batch, const pa_guest *guest, double T, int64_t ninsert, uint64_t seed, int32_t nblocks,
             int32_t run, int64_t chunk, uint32_t precision, pa_widom_result *out);

/* Size of the device array behind pa_widom_noalloc, bytes. */
#define PA_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

/* pa_widom that allocates no device memory. No counterpart in the reference: it is for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, kernels and results as pa_widom with the batch's bin scale as it
 * stands (it never runs the bin-scale tuner; fix the scale with pa_batch_set_bin_scale to
 * compare results bit for bit). Every per-call device buffer (energies, bins, pose coordinates,
 * tile sums, scan storage) is carved from one static device array of PA_NOALLOC_SCRATCH_BYTES per
 * device (64 MiB, held until the process ends): 1e6 insertions of a 3-site guest fit with the
 * default chunk. A call that needs more fails with PA_ERR_ARG, stating both sizes, before any
 * launch. All calls on one device share the array: calls that overlap in time on one device
 * race and must be serialised by the caller. The array keeps the last call's contents; every
 * call clears what it needs zeroed. Tables created by pa_batch_create stay the batch's own. */
int pa_widom_noalloc(pa_batch *batch, const pa_guest *guest, double T, int64_t ninsert, uint64_t seed, int32_t nblocks,
                     int32_t run, int64_t chunk, uint32_t precision, pa_widom_result *out);

/* Gradients of ln K_H and q_st for every system of the batch, on the same insertion streams as
 * pa_widom. No counterpart in the reference (it has no derivative path); the quantities are the
 * reweighting estimators
 *   d ln K_H / dθ = -β <w ∂U/∂θ> / <w>,
 *   d q_st / dθ   = -(<w ∂U/∂θ> - β <w U ∂U/∂θ>) / <w> - β (<w U>/<w>) <w ∂U/∂θ> / <w>,
 * over the force field's LJ parameters (per type epsilon_t and sigma_t, Lorentz-Berthelot
 * mixing; entries of types absent from the frameworks are 0) and the framework charges.
 * Charge entries follow the framework atoms as passed; the derivative of atom j moves every
 * replication image of j together (its `replication` claim stays true), which needs atoms in
 * `replicate` order (image-major, atoms inner) when the replication is not (1, 1, 1).
 * Outputs (nullable):
 *   out[nsys]                                      the pa_widom result of the same run,
 *   dlnkh_deps, dlnkh_dsig, dqst_deps, dqst_dsig   nsys x ntypes (force-field types), row per system,
 *   dlnkh_dq, dqst_dq                              sum over systems of their framework's natoms,
 *                                                  system-major. */
int pa_widom_grad(pa_batch *batch, const pa_guest *guest, double T, int64_t ninsert, uint64_t seed, int32_t nblocks,
                  int32_t run, uint32_t precision, pa_widom_result *out, double *dlnkh_deps, double *dlnkh_dsig,
                  double *dqst_deps, double *dqst_dsig, double *dlnkh_dq, double *dqst_dq);

/* Delta U = insertion_energy + constant_offset (eV) of insertions j0 .. j0+n-1 of system `sys`,
 * computed on the device (the widom_kernel! contract, src/widom.jl:157-175). */
int pa_widom_energies(pa_batch *batch, int32_t sys, uint64_t seed, int64_t j0, int64_t n, uint32_t precision,
                      double *dU);

/* Phase-0 hard-core flags of insertions j0 .. j0+n-1 of system `sys` at temperature T (K):
 * flags[i] = 1 when the mixed path rejects the insertion without computing its energy
 * (hardcore_kernel!, src/widom.jl:99-151, radii of build_rejection_tables :202-253). */
int pa_hardcore_flags(pa_batch *batch, int32_t sys, uint64_t seed, double T, int64_t j0, int64_t n, uint8_t *flags);

/* Device-drawn poses of the same insertions: pos (n x 3, Å) and quaternions (n x 4, x y z w). */
int pa_widom_poses(pa_batch *batch, int32_t sys, uint64_t seed, int64_t j0, int64_t n, double *pos, double *quat);

/* ---------------------------------------------------------------------------------------------
 * Move-energy path: NVT and μVT Monte Carlo chains, one per system of a batch built with
 * PA_BATCH_FULLK. Every chain runs on the device; moves are FP64.
 * ------------------------------------------------------------------------------------------- */
#define PA_MOVE_TRANSLATION 1 /* src/moves.jl:69-71 */
#define PA_MOVE_ROTATION 2
#define PA_MOVE_REINSERTION 3
#define PA_MOVE_INSERTION 4
#define PA_MOVE_DELETION 5

typedef struct pa_state pa_state;

/* Replaces SystemState(batch, guest, ncounts, ff; T, seed, capacities) (src/state.jl:328-446):
 * ncounts[n] initial guests (uniform poses redrawn until they clear phase 0 at T; pose of slot i,
 * attempt t from ChainRNG(splitmix64(seed ^ 0x1A17A1D05E5A3E11, n+1), i+1, t)) in capacities[n]
 * slots (NULL: capacities = ncounts), chain seeds splitmix64(seed, n+1), counters 0, running
 * structure factor, cached host energies and total energy set from scratch in FP64. */
int pa_state_create(pa_batch *batch, const pa_guest *guest, const int32_t *ncounts, const int32_t *capacities,
                    double T, uint64_t seed, pa_state **out);
void pa_state_destroy(pa_state *state);

/* mc_step! (src/moves.jl:577-645): one attempt of `movetype` (translation, rotation or
 * reinsertion) on every chain; step_trans (Å, Gaussian s.d.) and step_rot (in [0, 1]) per chain. */
int pa_mc_step(pa_state *state, int32_t movetype, const double *step_trans, const double *step_rot, double T);
/* mc_insert! / mc_delete! (src/moves.jl:1195-1290): one exchange attempt on every chain at
 * fugacity[n] (Pa). An insertion the Metropolis test accepts on a chain at capacity is an error. */
int pa_mc_insert(pa_state *state, const double *fugacity_pa, double T);
int pa_mc_delete(pa_state *state, const double *fugacity_pa, double T);
/* `nmoves` attempts of one move type back to back in one launch (no counterpart: the reference
 * launches three kernels per attempt). trace (nullable): nsys x nmoves x 4 doubles per attempt:
 * move type, slot within the system, dU (eV), accepted (0/1). */
int pa_mc_moves(pa_state *state, int32_t movetype, int64_t nmoves, const double *step_trans, const double *step_rot,
                const double *fugacity_pa, double T, double *trace);

typedef struct pa_chain_info {
    int32_t occupancy, capacity;
    double energy;            /* running total energy, eV */
    double energy_abs_accum;  /* sum of |dU| of accepted moves since the last audit */
    int64_t rng_counter;
    int32_t accepted[5], attempted[5]; /* per move type, PA_MOVE_* order */
    double const_p, const_q;  /* exchange constant term p + q N (src/moves.jl:728-773) */
} pa_chain_info;

int pa_state_info(const pa_state *state, int32_t sys, pa_chain_info *out);
/* Live guests of system `sys`: refpoints occupancy x 3 (Å), orientations occupancy x 4 (x y z w),
 * cached host-guest energies (eV); each nullable. */
int pa_state_guests(const pa_state *state, int32_t sys, double *refpoints, double *orientations, double *host_energy);
/* Running structure factor of system `sys`: nk x 2 (re, im). */
int pa_state_sk(const pa_state *state, int32_t sys, double *sk);
/* total_energy (src/guest.jl:293-343) of every system from its poses, FP64, with the structure
 * factor rebuilt from the poses. */
int pa_total_energy(pa_state *state, double *energy);
/* audit_energy! (src/audit.jl:97-165) on every system, the from-scratch rebuild on the device;
 * errors read as the reference's ("energy audit failed for system n (Float64): ..."). On
 * success resets the running structure factor and energy to the rebuilt values.
 * discrepancy (nullable): recomputed - running, eV. */
int pa_audit_energy(pa_state *state, double *discrepancy);
/* Test-only (no counterpart in the reference; for the audit's fault-injection tests, after
 * test/audit_tests.jl:60-415): overwrites one element of system `sys`'s running state.
 * PA_FIELD_ENERGY: running energy = re (eV). PA_FIELD_SK: running structure factor at k-vector
 * `index` = re + i im. PA_FIELD_HOST_ENERGY: cached host-guest energy of live guest `index` = re
 * (eV). PA_FIELD_OCCUPANCY: occupancy = re (an integer; raising it within the capacity is refused,
 * it would expose unset slots). An occupancy above its capacity makes every entry point that
 * reads the guest slots refuse the state (PA_ERR_ARG) until it is set back; pa_audit_energy
 * reports it. */
#define PA_FIELD_ENERGY 1
#define PA_FIELD_SK 2
#define PA_FIELD_HOST_ENERGY 3
#define PA_FIELD_OCCUPANCY 4
int pa_state_set_value(pa_state *state, int32_t field, int32_t sys, int64_t index, double re, double im);

/* GCMC driver (src/gcmc.jl:161-337). A cycle of chain n is max(its occupancy, min_cycle_length)
 * attempts; attempt c of chain n draws its move type from ChainRNG(chain seed, c, 1) word 0 with
 * probabilities p_translation, p_rotation, p_reinsertion, p_exchange/2 (insertion),
 * p_exchange/2 (deletion); its proposal and acceptance from ChainRNG(chain seed, c, 0) as in
 * pa_mc_step. Audit every n_audit cycles; counters reset at the start of production. */
typedef struct pa_gcmc_params {
    double T;
    int64_t n_warmup, n_production, n_audit, min_cycle_length;
    const double *fugacity_pa, *step_trans, *step_rot; /* per system */
    double p_translation, p_rotation, p_reinsertion, p_exchange;
    int32_t nblocks;
} pa_gcmc_params;

/* GCMCResult (src/gcmc.jl:32-43) plus the production move counters. */
typedef struct pa_gcmc_result {
    double loading, loading_err, energy, energy_err, q_st, q_st_err, corr_UN;
    int32_t max_occupancy, capacity;
    int64_t ncycles;
    int64_t accepted[5], attempted[5];
} pa_gcmc_result;

int pa_gcmc_run(pa_state *state, const pa_gcmc_params *params, pa_gcmc_result *out);
/* Per-cycle production samples of the last pa_gcmc_run: total energy (eV) and loading. */
int pa_gcmc_samples(const pa_state *state, int32_t sys, int64_t n, double *energy, double *loading);

/* NVT driver with Widom insertion along the chain (src/nvt.jl).
 *
 * widom_chain_kernel! (src/nvt.jl:85-119): Delta U (eV) of test insertions at the given poses
 * (pos n x 3 in Å, quat n x 4 as x y z w) into system `sys` as it currently stands: host-guest
 * real space, the reciprocal cross term against the running structure factor, guest-guest real
 * space against every live guest, plus insertion_constant_term(sys, occupancy) (src/nvt.jl:44-66).
 * The state is not changed. FP64. */
int pa_widom_chain_energies(pa_state *state, int32_t sys, int64_t n, const double *pos, const double *quat,
                            double *dU);

/* run_nvt! (src/nvt.jl:192-358) on every chain of `state`. n_warmup cycles are discarded, then
 * n_production cycles are recorded. A cycle is nsteps = max(max occupancy, min_cycle_length)
 * attempts on every chain (the cycle length is one value for the batch); attempt c of chain n
 * draws its move type uniformly among translation, rotation and reinsertion from
 * ChainRNG(chain seed, c, 1) word 0 (the reference draws one type per step for all chains from
 * a host Xoshiro(seed)), its proposal as pa_mc_step. With no guest slot in the batch no move is
 * attempted. After every production cycle, n_widom_per_cycle test insertions per system into the
 * configuration as it then stands; insertion t of production cycle c of system s takes the pose
 * of pa_widom's insertion j = c * n_widom_per_cycle + t of system s under widom_seed, so a batch
 * with no guests reproduces pa_widom(seed = widom_seed, ninsert = nsys * n_widom_per_cycle *
 * n_production, run = 1). The energy audit (pa_audit_energy) runs every n_audit cycles, warmup
 * and production sharing one count; a production cycle that ends on an audit records the audited
 * energy. Acceptance counters restart at the start of production. */
typedef struct pa_nvt_params {
    double T;
    int64_t n_warmup, n_production, n_widom_per_cycle, n_audit, min_cycle_length;
    const double *step_trans, *step_rot; /* per system */
    uint64_t widom_seed;
    int32_t nblocks;
} pa_nvt_params;

/* NVTResult (src/nvt.jl:22-33): energy block statistics over nblocks blocks of cycles, the Widom
 * estimators with every production cycle as one block (_reduce, src/widom.jl:451-492), and the
 * production acceptance rate of translation, rotation and reinsertion. */
typedef struct pa_nvt_result {
    double energy, energy_err;
    double mu_ex, mu_ex_err, K_H, K_H_err, q_st, q_st_err;
    double acceptance[3];
    int64_t ncycles;
} pa_nvt_result;

int pa_nvt_run(pa_state *state, const pa_guest *guest, const pa_nvt_params *params, pa_nvt_result *out);
/* Per-production-cycle records of the last pa_nvt_run on system `sys` (each nullable, n entries):
 * total energy (eV), and the cycle's Widom sums of exp(-dU/kT) and dU exp(-dU/kT). */
int pa_nvt_samples(const pa_state *state, int32_t sys, int64_t n, double *energy, double *sum_w, double *sum_uw);

/* Timing of the last pa_nvt_run or pa_widom_chain_energies on a state: device time of the move
 * launches and of the test-insertion launches (HIP events on the library's stream), and the
 * host wall time of the whole pa_nvt_run (audits, copies and the host reductions included). */
typedef struct pa_nvt_timing {
    double move_ms;
    int64_t move_launches;
    double insertion_ms;
    int64_t insertion_launches;
    int64_t insertions;
    double total_ms;
} pa_nvt_timing;

int pa_nvt_last_timing(const pa_state *state, pa_nvt_timing *out);

/* Mixed-path bin edge multiplier of `batch` (no counterpart in the reference): every system's
 * spatial bin edge is kappa * max(2 Å, (64 V / insertions)^(1/3)). 0 (the initial value) lets the
 * first large enough pa_widom call time a few values of kappa and keep the fastest; a fixed
 * kappa makes results reproducible across processes (the bin only changes FP32 rounding of the
 * coordinates relative to its centre, ~1e-7 relative in dU). */
int pa_batch_set_bin_scale(pa_batch *batch, double kappa);
int pa_batch_bin_scale(const pa_batch *batch, double *kappa);

/* Mixed-path neighbour-list capacity of `batch`, in entries per pass (no counterpart in the
 * reference). 0 (the initial value) sizes it per call from the atom density (1.3x the mean
 * count within the list radius, within the kernel's LDS budget); a fixed capacity must be a
 * multiple of 64 between one list-build slice (768) and the LDS limit. A bin whose list holds
 * more entries is built and swept in several passes; results agree with a single pass to FP64
 * summation order. */
int pa_batch_set_list_capacity(pa_batch *batch, int32_t capacity);
/* Capacity the last mixed-path pa_widom / pa_widom_energies call on `batch` used, and the
 * largest number of passes any of its bins needed (both 0 before the first such call; passes 0
 * after a pa_widom_noalloc call, which keeps no counter). Reads one counter from the device. */
int pa_batch_list_stats(pa_batch *batch, int32_t *capacity, int32_t *max_passes);

/* ---- Input files, host only (no device is touched) ----
 * Each reader parses one file into an opaque handle that owns the parsed values; the accessors
 * copy them into caller buffers sized from the *_info calls, in the layouts of pa_framework,
 * pa_forcefield and pa_guest. Errors are PA_ERR_ARG with the reference's ArgumentError wording
 * (PA_ERR_DIM where it throws DimensionMismatch or indexes past an array). */
typedef struct pa_cif_file pa_cif_file;
typedef struct pa_ff_file pa_ff_file;
typedef struct pa_guest_file pa_guest_file;

/* read_cif (src/structure.jl:37-69): one P1 data block with one _atom_site loop holding label,
 * type_symbol, fract_x/y/z and charge columns. */
int pa_read_cif(const char *path, pa_cif_file **out);
/* natoms and the cell (src/cell.jl:2-12: column-major, lattice vectors as columns); either may
 * be NULL. */
int pa_cif_info(const pa_cif_file *cif, int32_t *natoms, double cell[9]);
/* frac: natoms x 3 fractional coordinates; charges: natoms, e (either may be NULL). */
int pa_cif_atoms(const pa_cif_file *cif, double *frac, double *charges);
#define PA_CIF_LABELS 0
#define PA_CIF_SYMBOLS 1
/* Framework.labels or .symbols as consecutive NUL-terminated strings; *needed = bytes required.
 * buf == NULL only reports the size. */
int pa_cif_strings(const pa_cif_file *cif, int32_t which, char *buf, size_t cap, size_t *needed);
/* pa_framework.types: typeindex(ff, symbol * "_") of every atom (src/batch.jl:309). */
int pa_cif_types(const pa_cif_file *cif, const pa_ff_file *ff, int32_t *types);
void pa_cif_destroy(pa_cif_file *cif);

/* read_forcefield (src/forcefield.jl:54-61): `parameters: name: [sigma, epsilon]` (a null
 * sigma reads as 1, a null epsilon as 0), `cutoff` (required), `tail_correction` (default
 * true). Types keep the file's order. */
int pa_read_forcefield(const char *path, pa_ff_file **out);
int pa_ff_info(const pa_ff_file *ff, int32_t *ntypes, double *cutoff, int32_t *tail);
/* Per-type sigma / epsilon (ntypes) and the Lorentz-Berthelot tables of ForceField
 * (src/forcefield.jl:24-39; ntypes x ntypes, column-major); any pointer may be NULL. */
int pa_ff_params(const pa_ff_file *ff, double *sigma_type, double *epsilon_type, double *sigma, double *epsilon);
/* Type names as consecutive NUL-terminated strings (as pa_cif_strings). */
int pa_ff_names(const pa_ff_file *ff, char *buf, size_t cap, size_t *needed);
/* typeindex (src/forcefield.jl:41-45), 0-based. */
int pa_ff_typeindex(const pa_ff_file *ff, const char *name, int32_t *index);
void pa_ff_destroy(pa_ff_file *ff);

/* read_guest (src/forcefield.jl:91-101): `positions`, `symbols` (looked up in ff), `charges`,
 * `critical_temperature`, `critical_pressure` (digit-group underscores allowed, as
 * parse_number), `acentric_factor`. */
int pa_read_guest(const char *path, const pa_ff_file *ff, pa_guest_file **out);
int pa_guest_info(const pa_guest_file *g, int32_t *nsites, double *tc, double *pc, double *omega);
/* sites: nsites x 3 (Å); types: 0-based force-field indices; charges: e (any may be NULL). */
int pa_guest_sites(const pa_guest_file *g, double *sites, int32_t *types, double *charges);
void pa_guest_destroy(pa_guest_file *g);

/* Timing of the last pa_widom or pa_widom_grad call on `batch`. */
int pa_last_timing(const pa_batch *batch, pa_timing *out);

/* Copies the thread-local message of the last failing call (empty after success). Returns the
 * message length. */
int pa_last_error(char *buf, size_t len);

/* Name of the HIP device the library runs on (0 devices is an error). */
int pa_device_name(char *buf, size_t len);

/* Waits for all work on the device (the reference's KernelAbstractions.synchronize(backend)
 * after each launch, src/widom.jl:408,427). */
int pa_synchronize(void);

#ifdef __cplusplus
}
#endif
#endif
