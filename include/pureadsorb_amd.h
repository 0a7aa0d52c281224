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
    double grad_ms;               /* pa_widom_grad: accumulator clears and the gradient kernel */
    int64_t grad_insertions;      /* pa_widom_grad: insertions with a weight above e^-40 of the best */
} pa_timing;

typedef struct pa_batch pa_batch;

/* Replaces `FrameworkBatch(fws, ff, guest, ewald; fullk)` (src/batch.jl:237-432): validates,
 * deduplicates frameworks by value, builds every table in FP64 on the host (k-vectors,
 * prefactors, host structure factors, replication check, orientation-averaged guest self
 * term, constant_offset) and uploads them once. Input arrays are copied. */
int pa_batch_create(const pa_framework *fws, int32_t nsys, const pa_forcefield *ff, const pa_guest *guest,
                    const pa_ewald *ewald, uint32_t flags, pa_batch **out);
void pa_batch_destroy(pa_batch *batch);

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
