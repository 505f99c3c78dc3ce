/*
 * libsunnyhip — MI355X (gfx950) spin-dynamics hot path of Sunny.jl behind a
 * plain C ABI (no torch or HIP types in any signature).
 *
 * The reference (Sunny.jl 0.9.2) has no FFI layer: its extension point is
 * Julia multiple dispatch on `step!` (src/Integrators.jl:309).  Each entry
 * point below replaces one reference function for a device-resident copy of
 * a `System` (or a batch of replicas of it); the Julia-side binding is a
 * `ccall` per entry point (see the `SunnyHIP` module shipped beside the
 * library).
 *
 * Conventions
 *  - Return value: SNH_OK (0) or an SNH_ERR_* code; never throws.  The text
 *    of the last failure is available from snh_last_error().
 *  - Host arrays use Julia's column-major memory order for
 *    Array{Vec3,4}(L1,L2,L3,na): site = c1 + L1*(c2 + L2*(c3 + L3*i)),
 *    0-based, xyz contiguous per site (src/System/Types.jl:131-136).
 *    Batches concatenate replicas in order.  Coherent states are N
 *    consecutive (re, im) pairs per site.
 *  - Host pointers are borrowed for the duration of the call only.
 *  - One handle = one HIP stream; a handle is not internally synchronised
 *    (the same contract as a Sunny `System`, src/System/System.jl:138-174).
 *    Tiled Langevin batches also use a second stream of the handle, which
 *    the handle's stream waits for before any later work.
 */
#ifndef SUNNY_HIP_H
#define SUNNY_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SNH_OK 0
#define SNH_ERR_INVALID 1      /* invalid argument or state */
#define SNH_ERR_HIP 2          /* HIP runtime failure (no device, out of memory, ...) */
#define SNH_ERR_NOCONVERGE 3   /* implicit midpoint did not converge within max_iters */
#define SNH_ERR_UNSUPPORTED 4  /* configuration outside what this build supports */

typedef struct snh_system snh_system;

/* Description of a homogeneous System (and of R replicas of it).
 * Mirrors the data Sunny keeps in System{N} (src/System/Types.jl:106-140):
 * per-sublattice kappa / g / field, per-sublattice onsite couplings, and the
 * pair couplings of every sublattice (src/System/Types.jl:46-74). */
typedef struct {
    int32_t na;                 /* atoms per chemical cell (Bravais sublattices) */
    int32_t N;                  /* 0 = :dipole; 2..7 = :SUN with N = 2s+1 */
    int32_t precision;          /* 64 (plain FP64) or 32 (FP32 kernels, FP64 norms/residual/energy) */
    const double *kappa;        /* [na] */
    const double *g;            /* [na][3][3] row-major */
    const double *field;        /* [na][3] external field (energy units) */
    const double *stevens;      /* [na][28] c0|c2[5]|c4[9]|c6[13], q descending; :dipole only */
    const int32_t *kmax;        /* [na] 0, 2, 4 or 6; :dipole only */
    const double *Lambda;       /* [na][N][N][2] onsite matrix, row-major (re, im); :SUN only */
    int32_t nbonds;             /* directed bonds, both directions of every pair */
    const int32_t *bond;        /* [nbonds][6] = i, j, n1, n2, n3, isculled (0-based atoms) */
    const double *bilin;        /* [nbonds][9] row-major J of bond i -> j (reverse bond holds J') */
    const double *biquad;       /* [nbonds][25] row-major coupling of Stevens quadrupoles */
    const int32_t *has_biquad;  /* [nbonds] */
    const double *scalar;       /* [nbonds] constant shift */
    int32_t ewald;              /* 1: long-range dipole-dipole (enable_dipole_dipole!) */
    double mu0_muB2;            /* strength, e.g. 0.6745817653324668 meV*A^3 (src/Units.jl:91) */
    double demag[9];            /* demagnetization tensor, row-major */
    double latvecs[9];          /* lattice vectors as columns, row-major */
    const double *positions;    /* [na][3] fractional positions (Ewald only) */
    int32_t nreplicas;          /* R >= 1 */
    const int32_t *dims;        /* [R][3] supercell dims per replica (may differ; with Ewald each run
                                   of consecutive equal dims shares one transformed tensor) */
    int32_t layout;             /* SNH_LAYOUT_AUTO or SNH_LAYOUT_GHOST (below) */
    /* :SUN only: the general part of set_pair_coupling! (the TensorDecomposition
     * of src/System/Types.jl:23-27, PairExchange.jl:85-111): directed bond b
     * carries terms gen_start[b] .. gen_start[b+1]-1, each a pair (A, B) of
     * N x N Hermitian matrices, row-major (re, im), with bond energy
     * sum_t Re<Zi|A_t|Zi> Re<Zj|B_t|Zj> (Interactions.jl:496-503); the
     * reverse bond holds the same terms as (B, A).  gen_start NULL: none. */
    const int32_t *gen_start;   /* [nbonds + 1] */
    const double *gen_A;        /* [nterms][N][N][2] */
    const double *gen_B;        /* [nterms][N][N][2] */
} snh_spec;

/* Device layouts of the :dipole gather.  AUTO picks the LDS-tiled sweep
 * (unpadded planes) whenever its tile fits the LDS, else the ghost-padded
 * global gather; GHOST forces the latter.  Results agree to rounding.  AUTO
 * also takes the ghost layout for one replica of 513 to 8192 sites without
 * Ewald, whose midpoint steps then run in one cooperative launch per call
 * (snh_query "cooperative"). */
#define SNH_LAYOUT_AUTO 0
#define SNH_LAYOUT_GHOST 1

/* Builds the device copy: replaces System construction plus the couplings set
 * by set_exchange!/set_onsite_coupling!/enable_dipole_dipole!
 * (src/System/System.jl:33-93, src/System/Interactions.jl:182-189).  With
 * Ewald, the real-space tensor A and its transform are computed on the GPU
 * (src/System/Ewald.jl:2-22, 61-152). */
int snh_create(const snh_spec *spec, snh_system **out);
int snh_destroy(snh_system *sys);

/* Text of the most recent failure on this handle (or of the last failed
 * snh_create when sys is NULL).  Valid until the next call on the handle. */
const char *snh_last_error(const snh_system *sys);

/* No counterpart in the reference (build provenance): the first 16 hex
 * digits of SHA-256 over the library's sources and this header, in sorted
 * path order, as the build computed them ("unknown" for a build without
 * the id).  The Python loader recomputes it from the sources beside the
 * library and refuses a library built from other sources. */
const char *snh_build_id(void);

/* Total sites over all replicas, and sites of replica r. */
int snh_num_sites(const snh_system *sys, int64_t *total);
int snh_replica_sites(const snh_system *sys, int32_t r, int64_t *nsites);

/* sys.dipoles / sys.coherents of all replicas (src/System/Types.jl:133-134).
 * Setting coherents also sets dipoles = expected_spin(Z). */
int snh_set_dipoles(snh_system *sys, const double *dipoles);
int snh_get_dipoles(snh_system *sys, double *dipoles);
int snh_set_coherents(snh_system *sys, const double *coherents);
int snh_get_coherents(snh_system *sys, double *coherents);

/* step!(sys, Langevin(dt; damping, kT)) repeated nsteps times
 * (src/Integrators.jl:313-357).  kT: [R] per-replica temperatures.
 * noise: NULL for device Philox4x32-10 noise keyed by (seed, replica) and
 * counted by (site, step0 + k); otherwise the raw draws of randn!(rng, xi)
 * for ONE step (nsteps must be 1), laid out like the state:
 * 3 standard normals per site (:dipole) or N complex normals with variance
 * 1/2 per component (:SUN).  Scaling by sqrt(2 dt damping kT) happens here,
 * as in fill_noise! (src/Integrators.jl:288-298).
 * With SNH_GRAPH_STEPS=n > 0 in the environment at snh_create, device noise
 * (or none) and kernel timing off, runs of whole chunks of n steps replay
 * one captured HIP graph per handle (the noise step counter lives in device
 * memory); results are identical to step-by-step calls.  The tiled kernels
 * take lean forms (no Ewald potential, device noise with the step as an
 * argument, onsite terms of rank <= 2) when a call needs nothing else;
 * SNH_LEAN=0 at snh_create keeps the general forms (identical results). */
int snh_step_langevin(snh_system *sys, double dt, double damping, const double *kT, const double *noise,
                      uint64_t seed, uint64_t step0, int32_t nsteps);

/* step!(sys, ImplicitMidpoint(dt; damping, tol); max_iters) repeated nsteps
 * times (src/Integrators.jl:379-452).  iters_out: NULL or [R], the
 * fixed-point iterations of the last step.  A replica that does not converge
 * leaves its state at the last completed step and the call returns
 * SNH_ERR_NOCONVERGE with the reference's message.  Steps are queued in
 * groups of up to 16 with an iteration budget (the previous group's largest
 * count + 1) and one host check per group; a group in which a replica needs
 * more iterations is undone, tried once with two more, then rerun with a
 * check per step.  SNH_MID_SYNC=1 in the environment at snh_create checks
 * every step.  The results do not depend on any of this. */
int snh_step_midpoint(snh_system *sys, double dt, double damping, double tol, int32_t max_iters, int32_t nsteps,
                      int32_t *iters_out);

/* energy(sys) for every replica, FP64 sum (src/System/Interactions.jl:401-508,
 * src/System/Ewald.jl:155-181).  E: [R]. */
int snh_energy(snh_system *sys, double *E);

/* No counterpart in the reference: the variant of snh_energy for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation).  E_dev: DEVICE memory for [R] FP64
 * energies, written asynchronously on the handle's stream (synchronise, or
 * wait on an event recorded on that stream, before reading).  The call
 * allocates nothing, device or host, and does not synchronise: the
 * per-workgroup partial sums live in a static device array of
 * SNH_NOALLOC_PARTIALS doubles (2 MiB, one copy per GPU, held until the
 * process ends), whose address snh_create resolves for its device.  A handle
 * whose energy launch needs more partial sums than that gets
 * SNH_ERR_UNSUPPORTED (snh_energy still serves it).  Every call on one GPU
 * uses that GPU's array: calls that overlap on one GPU (handles on different
 * streams) race.  A handle whose Ewald grid is transformed by hipFFT (see
 * snh_query "ewald_fft_lib") also gets SNH_ERR_UNSUPPORTED. */
#define SNH_NOALLOC_PARTIALS 262144
int snh_energy_noalloc(snh_system *sys, double *E_dev);

/* set_energy_grad_dipoles! (:dipole; out [sites][3]) or
 * set_energy_grad_coherents! (:SUN; out [sites][N][2]) at the current state
 * (src/System/Interactions.jl:516-669). */
int snh_energy_grad(snh_system *sys, double *out);

/* Blocks until all work queued on the handle's stream has finished. */
int snh_synchronize(snh_system *sys);

/* Gradients through a deterministic implicit-midpoint trajectory.
 *
 * snh_midpoint_record runs nsteps of step!(sys, ImplicitMidpoint(dt; tol))
 * with damping 0 (src/Integrators.jl:379-413) from the current state and
 * keeps the state on the device every checkpoint_every steps; the handle
 * ends at the final state, from which the caller evaluates its loss.
 *
 * snh_midpoint_adjoint then sweeps the recorded trajectory backwards
 * (recomputing each segment from its checkpoint) and returns, for the loss
 * gradient dLdS_final at the final state ([sites][3], host layout):
 *   grad_out[p]   dL/dtheta_p for nparams coupling parameters, each given
 *                 as the spec of its unit couplings (same na, kappa and bond
 *                 list as the system's spec; the system's couplings are
 *                 linear in the parameters, as ModelParam,
 *                 src/System/Types.jl:76-86, src/System/Interactions.jl:1-42);
 *   dLdS_initial  dL/dS at the recorded initial state ([sites][3]; may be NULL).
 * The handle's state is left at the final state.
 *
 * Supported: one replica in :dipole mode at precision 64, with every
 * :dipole coupling: bilinear and biquadratic pair couplings, Zeeman, Stevens
 * anisotropy up to order 6, and the Ewald term (its strength is not a
 * parameter).
 * Sunny has no reverse rule for step!; these entry points are what a
 * ChainRulesCore rrule would call (the pattern of src/Fitting.jl:123-152). */
int snh_midpoint_record(snh_system *sys, double dt, double tol, int32_t max_iters, int32_t nsteps,
                        int32_t checkpoint_every);
int snh_midpoint_adjoint(snh_system *sys, int32_t nparams, const snh_spec *params, const double *dLdS_final,
                         double *grad_out, double *dLdS_initial);

/* Diagnostics with no counterpart in the reference: per-kernel timing by HIP
 * events recorded on the handle's stream around every launch of a kernel
 * family ("dipole_predict", "dipole_correct", "sun_predict", "sun_correct",
 * "midpoint_iter", "ewald_field").  enable = 0 stops recording and clears
 * the totals.  snh_kernel_time synchronises and reports the summed
 * milliseconds and the number of launches recorded since enabling. */
int snh_enable_timing(snh_system *sys, int32_t enable);
int snh_kernel_time(snh_system *sys, const char *family, double *ms, int64_t *launches);

/* Diagnostics with no counterpart in the reference: copies the state planes
 * into the stage buffer `repeats` times with a plain streaming kernel
 * (k_copy_planes), whose bytes per launch are returned in *bytes (read +
 * write).  Used to calibrate HBM traffic counters against a known size. */
int snh_copy_calibration(snh_system *sys, int32_t repeats, int64_t *bytes);

/* Sampled dynamical correlations of one replica, on the device (:dipole, or
 * :SUN with the observables acting on the expected spins, as
 * observable_values! reads sys.dipoles).
 *
 * snh_corr_init replaces SampledCorrelations(sys; measure, energies, dt)
 * (src/SampledCorrelations/SampledCorrelations.jl:194-268) for the buffers:
 *   nobs observables, obs [nobs][na][3]: the value of observable a at a site
 *   of atom i is obs[a][i] . S (Sunny's dipole-mode observables,
 *   src/Measurements/MeasureSpec.jl:53-64: rows of g, or of the identity);
 *   ncorr correlation pairs, pairs [ncorr][2] (0-based observable indices);
 *   n_all_omega = 2 (n_omega - 1) time slots (even, <= 4096), measperiod
 *   integration steps per snapshot, window 1 (cosine) or 0 (rectangular).
 * The caller computes dt and measperiod as adjusted_dt_and_downsampling_factor
 * (SampledCorrelations.jl:152-173).  Lattice axes of up to 32 cells are
 * transformed by line transforms in registers, longer axes (or any under the
 * environment variable SNH_CORR_FFT=lib) by a batched hipFFT plan made here
 * (snh_query "corr_fft_lib").
 *
 * snh_corr_add_sample replaces add_sample!(sc, sys)
 * (src/SampledCorrelations/CorrelationSampling.jl:1-175): a trajectory of
 * step!(sys, ImplicitMidpoint(dt)) (damping 0, tolerance tol) records
 * n_all_omega/2 + 1 snapshots of the observables, measperiod steps apart, and
 * accumulates the sample's correlations into the running mean (the handle
 * ends at the trajectory's last state).
 *
 * snh_corr_data returns sc.data, complex [ncorr][na][na][L1][L2][L3][n_all_omega]
 * in Julia's column-major order (ncorr fastest, (re, im) pairs), and the
 * number of samples (either pointer may be NULL). */
int snh_corr_init(snh_system *sys, int32_t nobs, const double *obs, int32_t ncorr, const int32_t *pairs,
                  int32_t n_all_omega, int32_t measperiod, int32_t window);
int snh_corr_add_sample(snh_system *sys, double dt, double tol, int32_t max_iters);
int snh_corr_data(snh_system *sys, double *data, int64_t *nsamples);

/* step!(sys, LocalSampler(; kT, nsweeps, propose)) for every replica
 * (src/MonteCarlo/Samplers.jl:133-153), :dipole or :SUN; with Ewald the
 * trials run one at a time, as described at the end.
 * Proposal p of the mixture (@mix_proposals, Samplers.jl:67-100) has
 * kinds[p] = 0 propose_uniform (randspin, src/System/System.jl:444-447),
 * 1 propose_flip (flip, System.jl:440-442; flip_ket for :SUN, Spin.jl:119-125)
 * or 2 propose_delta(magnitudes[p]) (perturbed_spin, System.jl:454-468; an
 * infinite magnitude is uniform),
 * drawn with probability weights[p] / sum(weights); 1 <= nprop <= 4.
 * kT: [R].  A trial's energy change is local_energy_change
 * (src/System/Interactions.jl:311-380) in FP64, and it is accepted when
 * kT = 0 and dE <= 0, or rand <= exp(-dE/kT).
 * Trial order: the reference draws ceil(nsweeps * nsites) random sites; here
 * ceil(2 nsweeps) passes go colour class by colour class (snh_local_coloring)
 * and try each site with probability 1/2 (the last pass: the remainder), so a
 * call makes nsweeps * nsites trials on average; pass k is numbered
 * P = 2 sweep0 + k.  Draws are Philox4x32-10 keyed by (seed, replica) and
 * counted by (site, P) on a stream disjoint from snh_step_langevin's.
 * With the Ewald term every pair of sites interacts: each replica then makes
 * ceil(nsweeps * nsites) sequential trials at random sites, trial
 * T = sweep0 * nsites + t drawing its site, u_acc and u_mix from counter
 * (T lo, 0x4D430002, T hi, 0) and its normals from (T lo, 0x4D430003, T hi,
 * 0); dE adds ewald_energy_delta (src/System/Ewald.jl:234-241; :SUN: of the
 * expected dipole) from an FP64 potential that every accepted move updates
 * at all sites (O(nsites) per move, as in the reference), 64 trials per
 * launch; a group of 4 or more replicas instead refreshes the potential by
 * one FFT evaluation after every 512 trials, each launch in between adding
 * the moves since the refresh at its own 64 sites (SNH_ERR_UNSUPPORTED for a non-symmetric demagnetization tensor:
 * the tensor's 3x3 blocks are kept as their upper triangles).  dE_dS (may be NULL): [R][4], the sums of dE and of
 * dS = S_new - S_old over the accepted trials of this call (LocalSampler's
 * fields dE, dS accumulate them). */
int snh_step_local(snh_system *sys, const double *kT, int32_t nprop, const int32_t *kinds,
                   const double *magnitudes, const double *weights, uint64_t seed, uint64_t sweep0,
                   double nsweeps, double *dE_dS);

/* The colour classes of snh_step_local (no counterpart in the reference):
 * class (i, r) holds the sites (c, i) with c = r mod m componentwise, and
 * color [na][m3][m2][m1] (may be NULL) its colour in 0..ncolors-1; no bond
 * joins two sites of one colour. */
int snh_local_coloring(snh_system *sys, int32_t m[3], int32_t *ncolors, int32_t *color);

/* Tuning with no counterpart in the reference: times up to `candidates` of
 * the tile plans the cost model ranks best on this handle's replicas (one
 * warm-up, then `launches` launches each of the Heun predictor and of the
 * corrector, the corrector writing to a scratch buffer so the state is
 * untouched) and keeps the fastest plan for each kernel for all later calls
 * (the predictor's plan also serves energy and gradients, the corrector's the
 * midpoint iteration); a later-ranked plan must be 2% faster to be kept.
 * Tiled batches of two or more replicas are then also timed as two replica
 * groups on two streams in anti-phase; that schedule is kept for Langevin
 * steps when it is 2% faster ("pair_on").
 * best_ms (may be NULL): predictor + corrector time of the kept plans.  A
 * handle on the ghost-padded layout is left as it is (best_ms = 0).  The
 * kept ranks are visible through snh_query ("tile_tuned",
 * "tile_tuned_correct").  Before any autotuning both kernels use the cost
 * model's first plan, except that FP32 handles whose tiled kernels take their
 * lean forms over isotropic bonds start both kernels on the same tile with
 * one site per thread when the model ranks that plan.  With the environment
 * variable SNH_AUTOTUNE_VERBOSE=1 every timed plan is printed to stderr. */
int snh_autotune(snh_system *sys, int32_t candidates, int32_t launches, double *best_ms);

/* Diagnostics with no counterpart in the reference: integer facts about the
 * device plan.  Keys: "layout" (SNH_LAYOUT_*, as chosen), "tile_T1",
 * "tile_T2", "tile_threads", "tile_work" (workgroups per tiled launch),
 * "tile_lds" (bytes of LDS per workgroup) of the predictor's plan;
 * "correct_T1", "correct_T2", "correct_sp", "correct_threads",
 * "correct_work" of the corrector's plan; "tile_candidates" (plans ranked by
 * the cost model), "tile_tuned" / "tile_tuned_correct" (ranks kept by
 * snh_autotune, -1 before),
 * "gather_chunk", "graph_steps"
 * (Langevin steps per captured HIP graph; 0 = graphs off, the default; set
 * by the environment variable SNH_GRAPH_STEPS at snh_create), "graph_live"
 * (1 when a captured graph is instantiated), "pair_split" (first replica of
 * the second replica group of a tiled Langevin batch; 0 = no groups, as with
 * SNH_PAIR=0 at snh_create), "pair_on" (1 when the groups step on two
 * streams in anti-phase: chosen by snh_autotune when whole steps are
 * measurably faster that way, or forced by SNH_PAIR=2). "mid_reruns"
 * (midpoint step groups undone after overrunning their iteration budget),
 * "mid_fallbacks" (groups then rerun with a host check per step),
 * "ewald_fft_lib" (1 when the Ewald grid is transformed by batched hipFFT
 * plans: grids whose transform does not fit one workgroup's LDS, an axis
 * longer than 32 cells, or the environment variable SNH_EWALD_FFT=lib at
 * snh_create; 0 for one workgroup per field), "ewald_slab" (groups of
 * replicas whose Ewald transforms are split by axis: per-plane 2D transforms
 * and one workgroup per in-plane wavevector for the third axis and the
 * contraction; groups of at most 2 replicas, SNH_EWALD_SLAB=0 never, =2 up
 * to 4), "corr_fft_lib" (1 when the
 * sampled correlations' lattice transform uses a hipFFT plan). */
int snh_query(snh_system *sys, const char *key, int64_t *value);

#ifdef __cplusplus
}
#endif

#endif /* SUNNY_HIP_H */
