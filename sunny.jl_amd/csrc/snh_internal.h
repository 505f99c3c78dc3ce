// Internal declarations shared by the libsunnyhip translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/sunny_hip.h"
#include "snh_device.h"

namespace snh {

// Coupling tables and index tables in device memory, in the kernel's real
// type T.  Passed to kernels by value.
template <typename T>
struct Couplings {
    const BlockInfo* blocks;
    const ReplicaInfo* reps;
    const int* bstart;   // [na+1] directed-bond range per sublattice (gather form)
    const int* bj;       // [nb]   target sublattice
    const int* bflag;    // [nb]   bit0 isotropic, bit1 biquad, bit2 culled
    const int* bdelta;   // [ndcls][nb] padded-index offset of the bond
    const T* bJ0;        // [nb] J[0][0] (the coupling of an isotropic bond)
    const T* bJ;         // [nb][9] row-major, J for bond i -> j
    const T* bQ;         // [nb][25] row-major quadrupole coupling
    const T* bscalar;    // [nb]
    const T* kappa;      // [na]
    const T* zee;        // [na][3] = g^T B
    const T* stev;       // [na][28]
    const T* stev_q2;    // [na][9] gradient matrix 2 Q2 of a kmax = 2 anisotropy (row-major)
    const int* kmax;     // [na]
    const T* gmat;       // [na][9]
    const T* field;      // [na][3]
    const T* lam;        // [na][N][N][2] SU(N) onsite matrix
    const T* quad_ops;   // [5][N][N][2] Stevens quadrupoles O[2,q]
    const int* gstart;   // [nb+1] general (A, B) terms of each padded bond slot (:SUN)
    const T* gA;         // [nterms][N][N][2]
    const T* gB;         // [nterms][N][N][2]
    int na, nb, N;
    int chunk;           // neighbours gathered per batch of loads (6, 9 or 18); every
                         // sublattice's bond list is padded to a multiple of it
                         // with zero couplings to the site itself (flag culled)
    int iso;             // every bond isotropic and without biquadratic coupling
    int lean_ok = 0;     // tiled kernels may take their lean forms: every onsite
                         // term of rank <= 2 (and SNH_LEAN is not 0)
    uint32_t total;      // plane stride (padded sites over all replicas)
    Ghost G;
};

template <typename T>
struct NoiseArgs {
    int mode;              // 0 zero, 1 host-supplied raw normals, 2 device Philox
    const double* host;    // [total][per_site] raw draws (mode 1), by padded site
    const T* scale;        // [R] sqrt(2 dt damping kT_r)
    unsigned long long seed;
    unsigned long long step;
    // non-null: the step counter lives in device memory (advanced by
    // k_step_advance after each step), so that a captured graph of several
    // steps replays with fresh noise
    const unsigned long long* step_ptr;
};

// Step number of the noise stream (kernel argument, or the device counter).
template <typename T>
__device__ __forceinline__ unsigned long long noise_step(const NoiseArgs<T>& nz)
{
    return nz.step_ptr ? ldc(nz.step_ptr, 0) : nz.step;
}

enum { kBondIso = 1, kBondBiquad = 2, kBondCulled = 4, kBondGeneral = 8 };

// Largest tiled workgroup for SP sites per thread: 1024 threads (<= 128
// VGPRs) for FP32 with one site per thread, else 512.
template <typename T, int SP>
constexpr int tile_threads_max()
{
    return (sizeof(T) == 4 && SP == 1) ? 1024 : 512;
}
// Layer positions fetched per thread: 3 with one site per thread, 6 with
// two.  Every slot costs three loads per layer whether or not a position
// falls in it (idle slots reload position 0), so the count is kept to what
// the common plans fill: a 32 x 2 tile of 16 sublattices has 2176
// positions for 1024 threads (4 slots: 0.727 ms per Heun step on the
// 64-replica pyrochlore batch, 3 slots: 0.685 ms).
constexpr int tile_load_per(int sp)
{
    return sp == 1 ? 3 : 6;
}

// LDS-tiled gather (snh_tile.hip).  State planes use the unpadded layout
// (ghost width 0): site (cell c, sublattice i) of replica r sits at
// off_r + c + ncell_r * i.  A workgroup owns a column of T1 x T2 cells of
// one replica and sweeps it along axis 3 over layers [k0, k1): the layers
// k - w3 .. k + w3 of a tile widened by (w1, w2) cells on each side live in
// an LDS ring of 2 w3 + 1 slots, so every neighbour of a layer-k site is in
// LDS at a bond-constant offset.  Halo cells are fetched with the periodic
// wrap of bonded_site (src/System/System.jl:204-212).
struct TileWork {
    int r;       // replica
    int t1, t2;  // first cell of the tile along axes 1 and 2
    int k0, k1;  // layer range along axis 3
    int pad[3];
};

struct TileArgs {
    const TileWork* work;
    const int* ldsoff;  // [nslot][nb] byte offset of bond b from the site, ring phase p = k mod nslot
    int T1, T2, lt1;    // tile cells along axes 1, 2; T1 = 1 << lt1
    int w1, w2, w3;     // halo widths
    int ROW;            // T1 + 2 w1 positions per tile row
    int nslot;          // 2 w3 + 1
    int cap;            // positions per slot (>= ROWS na ROW)
    int npos;           // (T2 + 2 w2) * na * ROW positions loaded per layer
    int spl;            // compute lanes per sublattice per layer: T1 T2 rounded up to 64
    int nsite;          // na * spl
    int nwork;          // workgroups
    int sp;             // sites computed per thread per layer (1 or 2)
};

// Geometry of the Ewald FFT grid (one replica; all replicas share it).
constexpr int kLmax = 32;  // longest axis of the in-register line transforms

struct EwaldGeom {
    int L1, L2, L3, H1, na;  // H1 = L1/2 + 1 (half spectrum along axis 1)
    long long ncell, nk;     // nk = H1 * L2 * L3
    int P1, P2;              // padded dims of the state layout
    uint32_t pcell;
    int w1, w2, w3;
    int plain;     // 1: library FFT path, spectra [replica][field][k]
    int slab = 0;  // 1: transforms split by axis (few replicas; in-LDS path only)
};

// Parameters of the real-space Ewald tensor precompute (FP64).
struct EwaldAParams {
    double lv[9];       // lattice vectors (columns), row-major
    double SL[9];       // superlattice vectors (columns), row-major
    double RL[9];       // reciprocal superlattice vectors (columns), row-major
    double demag[9];
    double V, sigma, rmax, kmax, mu0_muB2;
    int nmax[3], mmax[3];
    int L1, L2, L3, na;
};

// :dipole mode (snh_dipole.hip).  Buffers are plane sets [3][total].
template <typename T>
void launch_dipole_predict(const Couplings<T>& C, int nblocks, const T* S, T* Sp, T* H, const T* phi,
                           const NoiseArgs<T>& nz, T dt, T lambda, hipStream_t st);
template <typename T>
void launch_dipole_correct(const Couplings<T>& C, int nblocks, const T* Sp, const T* H, T* Sout, const T* phi,
                           T dt, T lambda, hipStream_t st);
template <typename T>
void launch_dipole_midpoint_hat(const Couplings<T>& C, int nblocks, const T* S, const T* S1, T* Shat,
                                const int* conv_iter, hipStream_t st);
template <typename T>
// Shat non-null: gathered with its images (hat_out, when non-null, receives
// the next iteration's Shat with images); null: formed from S and S1.
void launch_dipole_midpoint_iter(const Couplings<T>& C, int nblocks, const T* S, const T* S1, T* S2,
                                 const T* Shat, const T* phi, T dt, T lambda, double* partial,
                                 const int* conv_iter, T* hat_out, hipStream_t st);
// S = kappa n(S'') for replicas that converged at `iter` (0: at any
// iteration); S'' of iteration m is in Sa for odd m, Sb for even m.
// hat_next non-null: also the next step's first Shat (with images).
template <typename T>
void launch_dipole_midpoint_finalize(const Couplings<T>& C, int nblocks, T* S, const T* Sa, const T* Sb,
                                     const int* conv_iter, int iter, T* hat_next, hipStream_t st);
// The same after tiled sweeps with implicit iterates: S'' = 2 mu Shat - S
// from the converged iteration's Shat (Ho / He: odd / even iteration) and mu.
template <typename T>
void launch_dipole_midpoint_finalize_mu(const Couplings<T>& C, int nblocks, T* S, const T* Ho, const T* He,
                                        const T* Mo, const T* Me, const int* conv_iter, T* hat_next, hipStream_t st);
template <typename T>
void launch_dipole_grad(const Couplings<T>& C, int nblocks, const T* S, const T* phi, T* dE, hipStream_t st);
template <typename T>
void launch_dipole_energy(const Couplings<T>& C, int nblocks, const T* S, const T* phi, double* partial,
                          hipStream_t st);
void launch_reduce_partials(const double* partial, const int* rblk, int R, double* out, hipStream_t st);

// Raises a kernel's dynamic-LDS limit to `bytes` (HIP keeps the attribute per
// function), once per kernel and size.  The kernels seen so far sit in a
// fixed table, so a launch never allocates host memory (the entry point that
// allocates nothing launches through here too); past the table's capacity
// the attribute is simply set on every launch.
void allow_dynamic_lds(const void* fn, size_t bytes);

// Static device scratch of snh_energy_noalloc (snh_noalloc.hip): capacity in
// doubles; noalloc_scratch_init resolves the current device's copy (once per
// device); noalloc_scratch returns it (nullptr before init).
constexpr size_t kNoAllocPartials = SNH_NOALLOC_PARTIALS;
hipError_t noalloc_scratch_init();
double* noalloc_scratch(int dev);
void launch_midpoint_check(const double* partial, const int* rblk, int R, const double* atol2, int* conv_iter,
                           int iter, hipStream_t st);
void launch_fill_int(int* p, int n, int v, hipStream_t st);
// fail_step[r] = s where replica r has not converged (first such step only);
// max_iter[r]: largest iteration count of the group's steps 0..s; last_it[r]:
// step s's count; conv_iter cleared for the next step.
void launch_mid_budget(int* conv_iter, int* fail_step, int* max_iter, int* last_it, int R, int s, hipStream_t st);
// Persistent small-replica kernels (one workgroup per replica, rblk: its
// range of logical blocks): nsteps Langevin steps, or nsteps midpoint steps
// iterated to convergence on the device (iters_out / fail_step per replica).
template <typename T>
void launch_dipole_langevin_persist(const Couplings<T>& C, int R, const int* rblk, T* X, T* Xp, T* Xo,
                                   const NoiseArgs<T>& nz, T dt, T lambda, int nsteps, hipStream_t st);
// Largest one-replica handle stepped by the cooperative midpoint launch.
constexpr int kCoopMaxSites = 8192;
// Barrier state of a cooperative launch: count and gen zero at launch;
// *err set when a wait exceeded spin polls.
struct GridBar {
    unsigned* count;
    unsigned* gen;
    int* err;
    unsigned spin;
};
// nsteps midpoint steps of a one-replica handle (ghost layout, no Ewald) in
// one cooperative launch of nblocks workgroups (one per block); partial:
// [2 nblocks].  Returns the launch status (a grid that cannot be resident is
// refused).
template <typename T>
hipError_t launch_dipole_midpoint_coop(const Couplings<T>& C, int nblocks, T* X, T* B1, T* B2, T* Bh, T* Bh2, T dt,
                                       T lambda, const double* atol2, int max_iters, int nsteps, int* iters_out,
                                       int* fail_step, double* partial, const GridBar& bar, hipStream_t st);
template <typename T>
void launch_dipole_midpoint_persist(const Couplings<T>& C, int R, const int* rblk, T* X, T* B1, T* B2, T* Bh, T* Bh2,
                                    T dt, T lambda, const double* atol2, int max_iters, int nsteps, int* iters_out,
                                    int* fail_step, hipStream_t st);

// Sampled correlations (snh_corr.hip): buffers are FP64 complex (re, im).
template <typename T>
void launch_corr_observe(const Couplings<T>& C, int nblocks, const T* S, int nobs, const double* obs, void* samp,
                         int t, double scale, hipStream_t st);
void launch_corr_space_fft(void* buf, long long nfield, int L1, int L2, int L3, hipStream_t st);
// The same transform by a batched hipFFT plan (axes longer than kLmax).
struct CorrLibFft;
CorrLibFft* corr_lib_create(long long nfield, int L1, int L2, int L3, std::string& err);
bool corr_lib_exec(CorrLibFft* L, void* buf, hipStream_t st);
void corr_lib_destroy(CorrLibFft* L);
void launch_corr_tdft(const void* in, void* out, long long nl, int T, int sign, const double* pre, const double* post,
                      long long mean_count, hipStream_t st);
void launch_corr_product(const void* samp, void* P, int T, int nobs, int na, long long ncell, int a, int i, int b,
                         hipStream_t st);

// Metropolis sweeps of :dipole replicas (snh_mc.hip).  A colour class is a
// set (sublattice i, c1 mod m1, c2 mod m2, c3 mod m3) of sites no two of
// which share a bond.
struct McArgs {
    const int* bn;       // [nb][4] gather slot b: cell offset n1, n2, n3 and target sublattice (padding: the site itself)
    const double* kT;    // [R]
    unsigned long long seed;
    int m1, m2, m3;
    int wide;            // some bond offset exceeds a replica's dims: wrap by division
    int nprop;           // proposals of the mixture (1..4)
    int kind[4];         // 0 uniform, 1 flip, 2 Gaussian perturbation
    double mag[4];       // perturbation magnitude (kind 2)
    double cum[4];       // cumulative probabilities
};
// kBlock sites q0.. of class cls = rho1 + m1 (rho2 + m2 (rho3 + m3 i)) of replica r
struct McWork {
    int r, cls, q0, pad;
};
template <typename T>
void launch_mc_color(const Couplings<T>& C, const McArgs& A, T* S, const McWork* work, int nwork,
                     unsigned long long P, double prob, double* acc, hipStream_t st);
template <typename T>
void launch_sun_mc_color(const Couplings<T>& C, const McArgs& A, T* Z, const McWork* work, int nwork,
                         unsigned long long P, double prob, double* acc, hipStream_t st);
void launch_mc_reduce(const double* acc, const int* start, int ncolors, int R, double* out, hipStream_t st);
// Accepted sequential trial of a replica: site (sublattice i, cell c1..c3)
// and the change of its moment dmu = -g dS.
struct McTrial {
    int flag, i, c1, c2, c3, pad;
    double dmu[3];
};
// Sequential trials of the Ewald form run in chunks of at most kMcChunk per
// launch, lane l of a replica's first wavefront holding the chunk's l-th
// trial (:SUN: the l-th accepted move).
constexpr int kMcChunk = 64;
constexpr int kMcThreads = 256;  // per workgroup of the :dipole launch (trial workgroups: four wavefronts)
static_assert(kMcChunk <= 64, "one accepted move per lane of a wavefront");
// Buffers of one launch of sequential trials with the Ewald term.  Its
// trial workgroups see the potential phi_in of the chunk's start without the
// previous chunk's moves (prev, nprev; absent for a group's first chunk) and
// add those at their own sites; its other napp workgroups per replica write
// phi_out = phi_in + the previous chunk's moves at every site, for the next
// launch.  The chunk's accepted moves go to moves, nmoves ([R][kMcChunk],
// [R]).
struct McChunkIO {
    const double* phi_in;
    double* phi_out;
    // moves of nprevl earlier chunks, added at the trial sites (null: none):
    // list s of replica r at prev[(s R + r) kMcChunk], its count nprev[s R + r]
    // (R = nrep_all, the handle's replicas)
    const McTrial* prev;
    const int* nprev;
    int nprevl;
    int nrep_all;
    // non-null: the sums of those moves at the trial sites, formed by
    // launch_mc_prev_at_trials ([R][kMcChunk][3], FP64), used instead
    const double* pprev;
    McTrial* moves;
    int* nmoves;
    int napp;  // apply workgroups per replica (phi_out = phi_in + the one prev list), 0: none
};
constexpr int kMcApplyGroups = 128;  // napp at most
// Ewald groups of at least this many replicas refresh the potential of
// their sequential trials by an FFT evaluation after every kMcFftPeriod
// chunks instead (the chunks between add the moves since the refresh at
// their trial sites only).
constexpr int kMcFftReplicas = 4;
constexpr int kMcFftPeriod = 8;
// out[(r kMcChunk + l) 3 + x]: the moves of io's nprevl lists at the site of
// trial T0 + l of replica r (r0 <= r < r0 + nr), one wavefront per trial.
void launch_mc_prev_at_trials(const McArgs& A, const McChunkIO& io, const ReplicaInfo* reps, const double* Aew,
                              int na, unsigned long long T0, int r0, int nr, double* out, hipStream_t st);

// The real-space tensor's nblk 3x3 blocks A9 (nine doubles each) as the six
// upper-triangle entries xx, xy, xz, yy, yz, zz in A6 (the blocks are
// symmetric); the form the sequential trials read.
void launch_ewald_A_sym(const double* A9, double* A6, long long nblk, hipStream_t st);
// Trials T0 .. T0 + ntrial - 1 (ntrial <= kMcChunk) of replicas r0 .. r0 +
// nr - 1, which share their dims, with the buffers io; Aew
// is the real-space tensor and A0 its [na][9] self blocks A(0, i, i);
// acc[4 r + q] += (dE, dS).
template <typename T>
void launch_mc_ewald_chunk(const Couplings<T>& C, const McArgs& A, T* S, const McChunkIO& io, const double* Aew,
                           const double* A0, unsigned long long T0, int ntrial, double* acc, int r0, int nr,
                           hipStream_t st);
// The same for :SUN kets (the Ewald term acts on expected dipoles).
template <typename T>
void launch_sun_mc_ewald_chunk(const Couplings<T>& C, const McArgs& A, T* Z, const McChunkIO& io, const double* Aew,
                               const double* A0, unsigned long long T0, int ntrial, double* acc, int r0, int nr,
                               hipStream_t st);
template <typename T>
void launch_planes_to_double(const T* src, double* dst, long long n, hipStream_t st);
void launch_fill_double(double* p, long long n, double v, hipStream_t st);

// Device step counter of the noise stream: set, and advance by one.
void launch_step_set(unsigned long long* p, unsigned long long v, hipStream_t st);
void launch_step_advance(unsigned long long* p, hipStream_t st);
void launch_copy_planes(const void* src, void* dst, size_t bytes, hipStream_t st);

// LDS-tiled :dipole kernels (snh_tile.hip); nthreads = workgroup size.
size_t tile_lds_bytes(const TileArgs& A, size_t elem);
template <typename T>
void launch_tile_predict(const Couplings<T>& C, const TileArgs& A, int nthreads, const T* S, T* Sp, T* H,
                         const T* phi, const NoiseArgs<T>& nz, T dt, T lambda, hipStream_t st);
template <typename T>
void launch_tile_correct(const Couplings<T>& C, const TileArgs& A, int nthreads, const T* Sp, const T* H, T* Sout,
                         const T* phi, T dt, T lambda, hipStream_t st);
template <typename T>
// Shat non-null: the ring gathers these planes and the sweep writes the next
// Shat (hat_out) and mu = |(S + S'')/2| / kappa (mu_out) instead of S''
// (mu_prev: the previous iteration's mu, null at the first); null: it forms
// kappa n((S + S1)/2) at load and writes S'' to S2.
void launch_tile_midpoint_iter(const Couplings<T>& C, const TileArgs& A, int nthreads, const T* S, const T* S1,
                               T* S2, const T* Shat, const T* phi, T dt, T lambda, double* partial,
                               const int* conv_iter, T* hat_out, const T* mu_prev, T* mu_out, hipStream_t st);
template <typename T>
void launch_tile_grad(const Couplings<T>& C, const TileArgs& A, int nthreads, const T* S, const T* phi, T* dE,
                      hipStream_t st);
template <typename T>
void launch_tile_energy(const Couplings<T>& C, const TileArgs& A, int nthreads, const T* S, const T* phi,
                        double* partial, hipStream_t st);

// Adjoint of the deterministic :dipole midpoint step (snh_adjoint.hip, FP64).
template <typename T>
void launch_adj_prep(const Couplings<T>& C, int nblocks, const T* S, const T* Sdd, const T* lam, T* Shat, T* mu,
                     hipStream_t st);
template <typename T>
void launch_adj_w(const Couplings<T>& C, int nblocks, const T* nu, const T* Shat, T dt, T* w, hipStream_t st);
template <typename T>
void launch_adj_apply(const Couplings<T>& Clin, int nblocks, const T* w, const T* S, const T* Sdd, const T* g,
                      const T* nu, const T* mu, T dt, T* nu_out, double* partial, const T* phi_w,
                      const Couplings<T>& Cnl, const T* Shat, hipStream_t st);
template <typename T>
void launch_adj_param(const Couplings<T>& Cp, int nblocks, const T* Shat, const T* w, double* partial,
                      hipStream_t st);
template <typename T>
void launch_adj_lambda(const Couplings<T>& C, int nblocks, const T* nu, const T* mu, T* lam, hipStream_t st);
void launch_adj_accumulate(const double* partial, int n, double* out, int slot, hipStream_t st);

// Ewald (snh_ewald.hip).
// Dynamic LDS of one field's 3D FFT workgroup (elem = sizeof of the FFT type).
size_t ewald_fft_lds_bytes(const EwaldGeom& G, size_t elem);
// The matrix-core contraction serves FP32 batches of at least this many replicas.
constexpr int kEwaldMfmaMinReplicas = 8;
bool ewald_use_mfma(const EwaldGeom& G, int R, size_t elem);
size_t ewald_mfma_lds_bytes(const EwaldGeom& G);
// The slab path (per-plane transforms, per-column contraction) serves groups
// of at most kEwaldSlabReplicas replicas by default and kSlabMaxReplicas
// when forced (SNH_EWALD_SLAB=2); its LDS per plane and per column workgroup.
constexpr int kEwaldSlabReplicas = 2;
constexpr int kSlabMaxReplicas = 4;
size_t ewald_slab_plane_lds_bytes(const EwaldGeom& G, size_t elem);
size_t ewald_slab_column_lds_bytes(const EwaldGeom& G, int R, size_t elem);
// Bytes of one set of spectra (all fields of R replicas, blocked wavevectors).
size_t ewald_spectrum_bytes(const EwaldGeom& G, int R, size_t elem);
void launch_ewald_A(const EwaldAParams& P, const double* pos, double* A, hipStream_t st);
template <typename T>
void launch_ewald_fwdA(const EwaldGeom& G, const double* A, void* M, hipStream_t st);
// The library FFT path (G.plain): hipFFT plans of R replicas' fields and
// their buffers; create returns null with err set on failure.
struct EwaldLibFft;
EwaldLibFft* ewald_lib_create(const EwaldGeom& G, int R, size_t elem, std::string& err);
void ewald_lib_destroy(EwaldLibFft* L);
// A -> M through the library (returns an error text, empty on success).
template <typename T>
std::string launch_ewald_fwdA_plain(const EwaldGeom& G, const double* A, void* M, hipStream_t st);
// Potential phi of state S; false when a library transform fails.  With
// phi64 non-null the interior sites' potential goes to phi64 in FP64 instead
// (same layout; the library path, G.plain, refuses and returns false).
template <typename T>
bool launch_ewald_field(const EwaldGeom& G, int R, const ReplicaInfo* reps, uint32_t total, const T* gmat,
                        const T* S, const void* M, void* Fmu, void* Fphi, T* phi, EwaldLibFft* lib, hipStream_t st,
                        double* phi64 = nullptr);

// :SUN mode (snh_sun.hip).  Buffers are plane sets [2N][total]; expected
// dipoles are [3][total] planes with ghost images.
template <typename T>
void launch_sun_expected_spin(const Couplings<T>& C, int nblocks, const T* Z, T* S, hipStream_t st);
// D/Dp (and Dp/Dout) non-null: neighbours are gathered from the expected
// dipoles D of Z (bilinear couplings only) and the stage writes the dipoles
// of its output; null: neighbours are gathered from the kets.
template <typename T>
void launch_sun_predict(const Couplings<T>& C, int nblocks, const T* Z, T* Zp, T* H, const T* phi,
                        const NoiseArgs<T>& nz, T dt, T lambda, const T* D, T* Dp, hipStream_t st);
template <typename T>
void launch_sun_correct(const Couplings<T>& C, int nblocks, const T* Zp, const T* H, T* Zout, const T* phi, T dt,
                        T lambda, const T* Dp, T* Dout, hipStream_t st);
template <typename T>
void launch_sun_midpoint_bar(const Couplings<T>& C, int nblocks, const T* Z, const T* Z1, T* Zbar,
                             const int* conv_iter, hipStream_t st);
template <typename T>
void launch_sun_midpoint_iter(const Couplings<T>& C, int nblocks, const T* Z, const T* Z1, T* Z2, const T* phi,
                              T dt, T lambda, double* partial, const int* conv_iter, hipStream_t st);
template <typename T>
void launch_sun_midpoint_finalize(const Couplings<T>& C, int nblocks, T* Z, const T* Za, const T* Zb,
                                  const int* conv_iter, int iter, hipStream_t st);
template <typename T>
void launch_sun_grad(const Couplings<T>& C, int nblocks, const T* Z, const T* phi, T* HZ, hipStream_t st);
template <typename T>
void launch_sun_energy(const Couplings<T>& C, int nblocks, const T* Z, const T* phi, double* partial,
                       hipStream_t st);

}  // namespace snh
