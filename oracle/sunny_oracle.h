/*
 * FP64 CPU oracle for Sunny.jl's spin-dynamics hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker against which the
 * HIP implementation is compared.  It may be loaded only by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; the product
 * library never links or calls it.
 *
 * It restates, loop for loop, the reference algorithm (Sunny.jl 0.9.2):
 *   src/Integrators.jl:266-452          rhs_dipole!/rhs_sun!/fill_noise!, step!
 *   src/System/Interactions.jl:401-669  energy, set_energy_grad_dipoles!,
 *                                       set_energy_grad_coherents!
 *   src/System/OnsiteCoupling.jl:316-425 classical Stevens anisotropy
 *   src/Operators/Spin.jl:1-167         spin matrices, expected_spin, ...
 *   src/System/Ewald.jl:1-241           A precompute, FFT convolution, energy
 *   src/System/System.jl:204-212,416-468 bonded_site, normalize_*, proposals
 *   src/MonteCarlo/Samplers.jl:1-153    LocalSampler (Metropolis trials)
 *
 * Layout (identical to Julia's column-major Array{Vec3,4}(L1,L2,L3,na)):
 *   site = c1 + L1*(c2 + L2*(c3 + L3*i)), all 0-based
 *   dipoles:   double[nsites][3]
 *   coherents: double[nsites][N][2]   (re, im)
 */
#ifndef SUNNY_ORACLE_H
#define SUNNY_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int L[3];                 /* supercell dims L1, L2, L3 */
    int na;                   /* atoms (Bravais sublattices) per cell */
    int N;                    /* 0 = :dipole mode; N >= 2 = :SUN mode */
    const double *kappa;      /* [na]  |S| = kappa (dipole) or |Z|^2 = kappa (SU(N)) */
    const double *g;          /* [na][3][3] row-major g tensor */
    const double *field;      /* [na][3] external field B (energy units) */
    const double *stevens;    /* [na][28] c0 | c2[5] | c4[9] | c6[13], q descending */
    const int *kmax;          /* [na] 0,2,4,6 */
    const double *Lambda;     /* [na][N][N][2] onsite Hermitian matrix (SU(N)) */
    int nb;                   /* number of directed bonds (both directions) */
    const int *bond;          /* [nb][6] = i, j, n1, n2, n3, isculled */
    const double *bilin;      /* [nb][9] row-major J (bond i -> j) */
    const double *biquad;     /* [nb][25] row-major 5x5 coupling of Stevens quadrupoles */
    const int *has_biquad;    /* [nb] */
    const double *scalar;     /* [nb] constant shift */
    const double *ewald_A;    /* NULL or [na(j)][na(i)][L3][L2][L1][3][3] (A[off,i,j]) */
    /* general pair couplings (SU(N)): bond b has terms gen_start[b]..gen_start[b+1]-1,
       each (A, B) N x N complex [N][N][2] row-major; NULL gen_start: none */
    const int *gen_start;     /* [nb+1] */
    const double *gen_A;      /* [nterms][N][N][2] */
    const double *gen_B;      /* [nterms][N][N][2] */
} so_spec;

/* Stevens classical anisotropy, OnsiteCoupling.jl:316-425. c = [c0|c2|c4|c6]. */
void so_stevens_energy_grad(const double S[3], const double c[28], int kmax,
                            double *E, double grad[3]);

/* Spin matrices in the N-dim irrep, Spin.jl:1-14. Each out is [N][N][2]. */
void so_spin_matrices(int N, double *Sx, double *Sy, double *Sz);
/* Stevens quadrupoles O[2,q], q = 2..-2, Stevens.jl:12-86. out [5][N][N][2]. */
void so_stevens_quadrupoles(int N, double *Q);

/* expected_spin (Spin.jl:72-87) for all sites; Z [ns][N][2] -> S [ns][3]. */
void so_expected_spin(int N, long ns, const double *Z, double *S);

/* set_energy_grad_dipoles!, Interactions.jl:516-592, with Ewald (Ewald.jl:185-209). */
void so_grad_dipoles(const so_spec *sp, const double *S, double *dE);
/* set_energy_grad_coherents!, Interactions.jl:596-669. */
void so_grad_coherents(const so_spec *sp, const double *Z, double *HZ);
/* energy, Interactions.jl:401-508 (+ ewald_energy Ewald.jl:155-181).
   For SU(N), dipoles must equal expected_spin(Z). */
double so_energy(const so_spec *sp, const double *S, const double *Z);

/* One step!. noise is the raw draw of randn!(rng, xi) in Julia memory order:
   dipole: [ns][3] standard normals; SU(N): [ns][N][2], complex normals with
   variance 1/2 per component.  NULL means zeros. Integrators.jl:313-357. */
void so_step_langevin_dipole(const so_spec *sp, double *S, const double *noise,
                             double dt, double damping, double kT);
void so_step_langevin_sun(const so_spec *sp, double *Z, double *S, const double *noise,
                          double dt, double damping, double kT);
/* Integrators.jl:379-452. Returns iterations used (>0), or -1 when the
   fixed point fails to converge within max_iters (the reference errors). */
int so_step_midpoint_dipole(const so_spec *sp, double *S, double dt, double damping,
                            double tol, int max_iters);
int so_step_midpoint_sun(const so_spec *sp, double *Z, double *S, double dt,
                         double damping, double tol, int max_iters);

/* Ewald.jl:61-152 at q = 0: A[j][i][c3][c2][c1][3][3] (row-major 3x3), already
   scaled by mu0_muB2.  latvecs row-major: latvecs[r*3+c] = component r of a_c.
   positions [na][3] fractional. demag row-major 3x3. */
void so_ewald_precompute(const double latvecs[9], int na, const double *positions,
                         const int L[3], const double demag[9], double mu0_muB2,
                         double *A);
/* Ewald energy, Ewald.jl:155-181 (half-spectrum rfft with sqrt(2) weights). */
double so_ewald_energy(const so_spec *sp, const double *S);
/* Field at every site by the brute-force pairwise sum, Ewald.jl:211-230. */
void so_ewald_grad_bruteforce(const so_spec *sp, const double *S, double *dE);

/* Philox4x32-10 (Salmon et al., SC'11) and the Box-Muller map used by the
   device noise generator. Writes n standard normals for (site, step) with the
   layout normal[k] = k-th normal of that site. */
void so_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void so_philox_normals(uint64_t seed, uint64_t step, long nsites, int per_site,
                       double *out);

/* nsteps Langevin steps with the device's Philox noise stream (key, step0 + k),
   energy after every `every` steps into E.  S: dipoles (:SUN: expected spins
   of Zr, updated with it); Zr: NULL in :dipole mode. */
void so_langevin_run(const so_spec *sp, double *S, double *Zr, double dt, double damping, double kT, uint64_t key,
                     uint64_t step0, long nsteps, long every, double *E);
/* so_ewald_precompute with the splitting width sigma = cbrt(V) / sigma_div
   (the reference's sigma_div is 3). */
void so_ewald_precompute_sigma(const double latvecs[9], int na, const double *positions, const int L[3],
                               const double demag[9], double mu0_muB2, double sigma_div, double *A);
/* Planted mistakes, for tests that show a check catches them: 0 none,
   1 noise scaled by sqrt(dt lambda kT), 2 corrector without noise, 3 dipole
   damping with S/kappa, 4 Ewald displacement in fractional coordinates,
   5 Ewald displacement with the two atoms swapped. */
void so_set_mutation(int m);

/* local_energy_change (Interactions.jl:311-380, Ewald.jl:234-241), :dipole
   mode: energy change when the dipole of `site` becomes Snew.  NAN for a
   bond that maps the site onto itself. */
double so_local_energy_change(const so_spec *sp, const double *S, long site, const double Snew[3]);
/* LocalSampler trials (Samplers.jl:1-153) in the order of the device:
   ceil(2 nsweeps) passes over the colour classes color[na][m3][m2][m1], each
   trying every site with probability 1/2 (the last pass: the remainder).
   Proposal p has kind[p] (0 uniform, 1 flip, 2 delta with mag[p]) and
   cumulative probability cum[p].  acc += (dE, dS) of accepted trials. */
int so_local_sweeps(const so_spec *sp, double *S, const int m[3], const int *color, int ncolors, int nprop,
                    const int *kind, const double *mag, const double *cum, double kT, uint64_t key,
                    uint64_t sweep0, double nsweeps, double acc[4]);
/* :SUN versions (Interactions.jl:311-380, Samplers.jl:1-153): Z [ns][N][2],
   S [ns][3] the expected spins of Z (kept in step by the sweeps). */
double so_local_energy_change_sun(const so_spec *sp, const double *Z, const double *S, long site,
                                  const double *Znew);
int so_local_sweeps_sun(const so_spec *sp, double *Z, double *S, const int m[3], const int *color, int ncolors,
                        int nprop, const int *kind, const double *mag, const double *cum, double kT, uint64_t key,
                        uint64_t sweep0, double nsweeps, double acc[4]);
/* Sequential trials at random sites with Philox draws (the device's form
   when the Ewald term is on): trial T = sweep0 * ns + t, counters
   (T lo, 0x4D430002, T hi, 0) for site and uniforms, (T lo, 0x4D430003,
   T hi, 0) for the normals. */
int so_local_sequential(const so_spec *sp, double *S, int nprop, const int *kind, const double *mag,
                        const double *cum, double kT, uint64_t key, uint64_t sweep0, double nsweeps, double acc[4]);
/* :SUN form of so_local_sequential: the normals of trial T from the counters
   (T lo, 0x4D430003 + 2b, T hi, 0), b = 0, 1, ... */
int so_local_sequential_sun(const so_spec *sp, double *Z, double *S, int nprop, const int *kind, const double *mag,
                            const double *cum, double kT, uint64_t key, uint64_t sweep0, double nsweeps,
                            double acc[4]);
/* The same trials at uniformly random sites, as Samplers.jl:133-153 orders
   them (ceil(nsweeps * nsites) trials; splitmix64 draws from seed). */
int so_local_random_sites(const so_spec *sp, double *S, int nprop, const int *kind, const double *mag,
                          const double *cum, double kT, uint64_t seed, double nsweeps, double acc[4]);

#ifdef __cplusplus
}
#endif
#endif
