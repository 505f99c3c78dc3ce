/*
 * FP64 CPU oracle for Sunny.jl's spin-dynamics hot path.  TEST INFRASTRUCTURE
 * ONLY: it is the checker, never the product.  Each function cites the
 * reference lines it restates (paths relative to the Sunny.jl 0.9.2 tree).
 */
#include "sunny_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

/* Planted mistakes for the tests that show a check catches them (0 = none):
   1  fill_noise! scales by sqrt(dt lambda kT), not sqrt(2 dt lambda kT)
   2  the Heun corrector stage drops the noise term
   3  the dipole damping term uses S/kappa in place of S
   4  Ewald pair displacement taken in fractional coordinates
   5  Ewald pair displacement c + p_i - p_j (atoms swapped)
   Never set outside those tests. */
static int g_mutation = 0;
void so_set_mutation(int m) { g_mutation = m; }

/* ------------------------------------------------------------------------ */
/* Indexing: System.jl:204-212 (bonded_site with periodic wrap), Types.jl:131 */

static inline long site_index(const int L[3], int c1, int c2, int c3, int i)
{
    return (long)c1 + (long)L[0] * ((long)c2 + (long)L[1] * ((long)c3 + (long)L[2] * i));
}

static inline int wrap(int x, int L)
{
    int r = x % L;
    return r < 0 ? r + L : r;
}

static inline long ncells(const so_spec *sp) { return (long)sp->L[0] * sp->L[1] * sp->L[2]; }
static inline long nsites(const so_spec *sp) { return ncells(sp) * sp->na; }

/* ------------------------------------------------------------------------ */
/* Stevens anisotropy, OnsiteCoupling.jl:316-425.  c = c0 | c2[5] | c4[9] | c6[13],
   each block in descending q = k..-k.  1-based Julia indices ck[m] map to
   ck[m-1] here.                                                              */

void so_stevens_energy_grad(const double S[3], const double c[28], int kmax,
                            double *Eout, double grad[3])
{
    const double *c2 = c + 1, *c4 = c + 6, *c6 = c + 15;
    double E = c[0];
    double dE_dz = 0.0;
    cplx dE_dJp = 0.0;

    if (kmax == 0) goto done;

    double X = S[0] * S[0] + S[1] * S[1] + S[2] * S[2];
    cplx Jp1 = S[0] + I * S[1];
    double Jz1 = S[2];
    cplx Jp2 = Jp1 * Jp1;
    double Jz2 = Jz1 * Jz1;
    {
        double A1 = 3 * Jz2 - X, A2 = Jz1, A3 = 1;
        double dA1 = 6 * Jz1, dA2 = 1;
        E += (c2[0] * creal(Jp2) + c2[4] * cimag(Jp2)) * A3 +
             (c2[1] * creal(Jp1) + c2[3] * cimag(Jp1)) * A2 + c2[2] * A1;
        dE_dz += (c2[1] * creal(Jp1) + c2[3] * cimag(Jp1)) * dA2 + c2[2] * dA1;
        dE_dJp += (2.0 / 2) * (c2[0] * Jp1 - I * c2[4] * Jp1) * A3 +
                  (1.0 / 2) * (c2[1] - I * c2[3]) * A2;
    }
    if (kmax == 2) goto done;

    double X2 = X * X;
    cplx Jp3 = Jp2 * Jp1;
    double Jz3 = Jz2 * Jz1;
    cplx Jp4 = Jp2 * Jp2;
    double Jz4 = Jz2 * Jz2;
    {
        double A1 = 35 * Jz4 - (30 * X) * Jz2 + (3 * X2);
        double A2 = 7 * Jz3 - (3 * X) * Jz1;
        double A3 = 7 * Jz2 - X;
        double A4 = Jz1;
        double A5 = 1;
        double dA1 = 140 * Jz3 - (60 * X) * Jz1;
        double dA2 = 21 * Jz2 - 3 * X;
        double dA3 = 14 * Jz1;
        double dA4 = 1;
        E += (c4[0] * creal(Jp4) + c4[8] * cimag(Jp4)) * A5 +
             (c4[1] * creal(Jp3) + c4[7] * cimag(Jp3)) * A4 +
             (c4[2] * creal(Jp2) + c4[6] * cimag(Jp2)) * A3 +
             (c4[3] * creal(Jp1) + c4[5] * cimag(Jp1)) * A2 + c4[4] * A1;
        dE_dz += (c4[1] * creal(Jp3) + c4[7] * cimag(Jp3)) * dA4 +
                 (c4[2] * creal(Jp2) + c4[6] * cimag(Jp2)) * dA3 +
                 (c4[3] * creal(Jp1) + c4[5] * cimag(Jp1)) * dA2 + c4[4] * dA1;
        dE_dJp += (4.0 / 2) * (c4[0] * Jp3 - I * c4[8] * Jp3) * A5 +
                  (3.0 / 2) * (c4[1] * Jp2 - I * c4[7] * Jp2) * A4 +
                  (2.0 / 2) * (c4[2] * Jp1 - I * c4[6] * Jp1) * A3 +
                  (1.0 / 2) * (c4[3] - I * c4[5]) * A2;
    }
    if (kmax == 4) goto done;

    double X3 = X2 * X;
    cplx Jp5 = Jp4 * Jp1;
    double Jz5 = Jz4 * Jz1;
    cplx Jp6 = Jp3 * Jp3;
    double Jz6 = Jz3 * Jz3;
    {
        double A1 = 231 * Jz6 - (315 * X) * Jz4 + (105 * X2) * Jz2 - (5 * X3);
        double A2 = 33 * Jz5 - (30 * X) * Jz3 + (5 * X2) * Jz1;
        double A3 = 33 * Jz4 - (18 * X) * Jz2 + X2;
        double A4 = 11 * Jz3 - (3 * X) * Jz1;
        double A5 = 11 * Jz2 - X;
        double A6 = Jz1;
        double A7 = 1;
        double dA1 = 1386 * Jz5 - (1260 * X) * Jz3 + (210 * X2) * Jz1;
        double dA2 = 165 * Jz4 - (90 * X) * Jz2 + 5 * X2;
        double dA3 = 132 * Jz3 - (36 * X) * Jz1;
        double dA4 = 33 * Jz2 - 3 * X;
        double dA5 = 22 * Jz1;
        double dA6 = 1;
        E += (c6[0] * creal(Jp6) + c6[12] * cimag(Jp6)) * A7 +
             (c6[1] * creal(Jp5) + c6[11] * cimag(Jp5)) * A6 +
             (c6[2] * creal(Jp4) + c6[10] * cimag(Jp4)) * A5 +
             (c6[3] * creal(Jp3) + c6[9] * cimag(Jp3)) * A4 +
             (c6[4] * creal(Jp2) + c6[8] * cimag(Jp2)) * A3 +
             (c6[5] * creal(Jp1) + c6[7] * cimag(Jp1)) * A2 + c6[6] * A1;
        dE_dz += (c6[1] * creal(Jp5) + c6[11] * cimag(Jp5)) * dA6 +
                 (c6[2] * creal(Jp4) + c6[10] * cimag(Jp4)) * dA5 +
                 (c6[3] * creal(Jp3) + c6[9] * cimag(Jp3)) * dA4 +
                 (c6[4] * creal(Jp2) + c6[8] * cimag(Jp2)) * dA3 +
                 (c6[5] * creal(Jp1) + c6[7] * cimag(Jp1)) * dA2 + c6[6] * dA1;
        dE_dJp += (6.0 / 2) * (c6[0] * Jp5 - I * c6[12] * Jp5) * A7 +
                  (5.0 / 2) * (c6[1] * Jp4 - I * c6[11] * Jp4) * A6 +
                  (4.0 / 2) * (c6[2] * Jp3 - I * c6[10] * Jp3) * A5 +
                  (3.0 / 2) * (c6[3] * Jp2 - I * c6[9] * Jp2) * A4 +
                  (2.0 / 2) * (c6[4] * Jp1 - I * c6[8] * Jp1) * A3 +
                  (1.0 / 2) * (c6[5] - I * c6[7]) * A2;
    }

done:
    *Eout = E;
    grad[0] = +2 * creal(dE_dJp);
    grad[1] = -2 * cimag(dE_dJp);
    grad[2] = dE_dz;
}

/* ------------------------------------------------------------------------ */
/* Spin matrices, Spin.jl:1-14; Stevens quadrupoles from Stevens.jl:12-86.  */

#define NMAX 16

void so_spin_matrices(int N, double *Sx, double *Sy, double *Sz)
{
    memset(Sx, 0, sizeof(double) * 2 * N * N);
    memset(Sy, 0, sizeof(double) * 2 * N * N);
    memset(Sz, 0, sizeof(double) * 2 * N * N);
    double s = (N - 1) / 2.0;
    for (int j = 1; j <= N - 1; j++) {
        double off = sqrt(2 * (s + 1) * j - j * (j + 1)) / 2;
        /* element (j-1, j) and (j, j-1), 0-based */
        int a = j - 1, b = j;
        Sx[2 * (a * N + b)] = off;
        Sx[2 * (b * N + a)] = off;
        Sy[2 * (a * N + b) + 1] = -off; /* -i*off above the diagonal */
        Sy[2 * (b * N + a) + 1] = +off;
    }
    for (int k = 0; k < N; k++) Sz[2 * (k * N + k)] = s - k;
}

static void cmat_mul(int N, const cplx *A, const cplx *B, cplx *C)
{
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            cplx acc = 0;
            for (int k = 0; k < N; k++) acc += A[i * N + k] * B[k * N + j];
            C[i * N + j] = acc;
        }
}

void so_stevens_quadrupoles(int N, double *Qout)
{
    double sx[2 * NMAX * NMAX], sy[2 * NMAX * NMAX], sz[2 * NMAX * NMAX];
    so_spin_matrices(N, sx, sy, sz);
    cplx Jx[NMAX * NMAX], Jy[NMAX * NMAX], Jz[NMAX * NMAX];
    for (int k = 0; k < N * N; k++) {
        Jx[k] = sx[2 * k] + I * sx[2 * k + 1];
        Jy[k] = sy[2 * k] + I * sy[2 * k + 1];
        Jz[k] = sz[2 * k] + I * sz[2 * k + 1];
    }
    cplx XX[NMAX * NMAX], YY[NMAX * NMAX], ZZ[NMAX * NMAX], XZ[NMAX * NMAX], ZX[NMAX * NMAX],
        YZ[NMAX * NMAX], ZY[NMAX * NMAX], XY[NMAX * NMAX], YX[NMAX * NMAX];
    cmat_mul(N, Jx, Jx, XX);
    cmat_mul(N, Jy, Jy, YY);
    cmat_mul(N, Jz, Jz, ZZ);
    cmat_mul(N, Jx, Jz, XZ);
    cmat_mul(N, Jz, Jx, ZX);
    cmat_mul(N, Jy, Jz, YZ);
    cmat_mul(N, Jz, Jy, ZY);
    cmat_mul(N, Jx, Jy, XY);
    cmat_mul(N, Jy, Jx, YX);
    /* k = 2: O22 = (Jp^2+Jm^2)/2 = Jx^2-Jy^2; O21 = {Jx,Jz}/2; O20 = 3Jz^2 - X;
       O2,-1 = {Jy,Jz}/2; O2,-2 = (Jp^2-Jm^2)/(2i) = {Jx,Jy}. */
    for (int k = 0; k < N * N; k++) {
        cplx q[5];
        q[0] = XX[k] - YY[k];
        q[1] = (XZ[k] + ZX[k]) / 2;
        q[2] = 3 * ZZ[k] - (XX[k] + YY[k] + ZZ[k]);
        q[3] = (YZ[k] + ZY[k]) / 2;
        q[4] = XY[k] + YX[k];
        for (int a = 0; a < 5; a++) {
            Qout[2 * (a * N * N + k)] = creal(q[a]);
            Qout[2 * (a * N * N + k) + 1] = cimag(q[a]);
        }
    }
}

/* quadrupole / grad_quadrupole of a classical dipole, Spin.jl:48-69 */
static void quadrupole(const double *S, double Q[5])
{
    double x = S[0], y = S[1], z = S[2];
    Q[0] = x * x - y * y;
    Q[1] = z * x;
    Q[2] = -x * x - y * y + 2 * z * z;
    Q[3] = z * y;
    Q[4] = 2 * y * x;
}

static void grad_quadrupole(const double *S, double G[5][3])
{
    double x = S[0], y = S[1], z = S[2];
    G[0][0] = 2 * x;  G[0][1] = -2 * y; G[0][2] = 0;
    G[1][0] = z;      G[1][1] = 0;      G[1][2] = x;
    G[2][0] = -2 * x; G[2][1] = -2 * y; G[2][2] = 4 * z;
    G[3][0] = 0;      G[3][1] = z;      G[3][2] = y;
    G[4][0] = 2 * y;  G[4][1] = 2 * x;  G[4][2] = 0;
}

/* ------------------------------------------------------------------------ */
/* SU(N) helpers: Spin.jl:72-104, 131-167.                                   */

typedef struct {
    int N;
    double off[NMAX];   /* Sx superdiagonal */
    double m[NMAX];     /* Sz diagonal */
    cplx S[3][NMAX * NMAX];
    cplx Q[5][NMAX * NMAX];
} sun_ops;

static void sun_ops_init(sun_ops *o, int N)
{
    o->N = N;
    double s = (N - 1) / 2.0;
    for (int j = 1; j <= N - 1; j++) o->off[j - 1] = sqrt(2 * (s + 1) * j - j * (j + 1)) / 2;
    for (int k = 0; k < N; k++) o->m[k] = s - k;
    double sx[2 * NMAX * NMAX], sy[2 * NMAX * NMAX], sz[2 * NMAX * NMAX], q[5 * 2 * NMAX * NMAX];
    so_spin_matrices(N, sx, sy, sz);
    so_stevens_quadrupoles(N, q);
    for (int k = 0; k < N * N; k++) {
        o->S[0][k] = sx[2 * k] + I * sx[2 * k + 1];
        o->S[1][k] = sy[2 * k] + I * sy[2 * k + 1];
        o->S[2][k] = sz[2 * k] + I * sz[2 * k + 1];
        for (int a = 0; a < 5; a++) o->Q[a][k] = q[2 * (a * N * N + k)] + I * q[2 * (a * N * N + k) + 1];
    }
}

static void expected_spin1(const sun_ops *o, const cplx *Z, double S[3])
{
    int N = o->N;
    cplx c = 0;
    for (int k = 0; k < N - 1; k++) c += conj(Z[k]) * (o->off[k] * Z[k + 1]);
    double nz = 0;
    for (int k = 0; k < N; k++) nz += creal(conj(Z[k]) * (o->m[k] * Z[k]));
    S[0] = 2 * creal(c);
    S[1] = 2 * cimag(c);
    S[2] = nz;
}

static void expected_quadrupole1(const sun_ops *o, const cplx *Z, double Q[5])
{
    int N = o->N;
    for (int a = 0; a < 5; a++) {
        cplx acc = 0;
        for (int j = 0; j < N; j++)
            for (int i = 0; i < N; i++) acc += conj(Z[i]) * o->Q[a][i * N + j] * Z[j];
        Q[a] = creal(acc);
    }
}

/* Re<Z|M|Z> of an N x N complex matrix M ([N][N][2]) (dot(Z, M, Z)) */
static double expect_op(int N, const double *M, const cplx *Z)
{
    cplx acc = 0;
    for (int a = 0; a < N; a++)
        for (int b = 0; b < N; b++) acc += conj(Z[a]) * (M[2 * (a * N + b)] + I * M[2 * (a * N + b) + 1]) * Z[b];
    return creal(acc);
}

/* out += f (M Z) */
static void add_op_ket(int N, const double *M, const cplx *Z, double f, cplx *out)
{
    for (int a = 0; a < N; a++) {
        cplx acc = 0;
        for (int b = 0; b < N; b++) acc += (M[2 * (a * N + b)] + I * M[2 * (a * N + b) + 1]) * Z[b];
        out[a] += f * acc;
    }
}

static inline int gen_terms(const so_spec *sp, int b, int *t0, int *t1)
{
    if (!sp->gen_start || sp->N == 0) return 0;
    *t0 = sp->gen_start[b];
    *t1 = sp->gen_start[b + 1];
    return *t1 > *t0;
}

void so_expected_spin(int N, long ns, const double *Zr, double *S)
{
    sun_ops o;
    sun_ops_init(&o, N);
    const cplx *Z = (const cplx *)Zr;
    for (long s = 0; s < ns; s++) expected_spin1(&o, Z + s * N, S + 3 * s);
}

/* ------------------------------------------------------------------------ */
/* Ewald helpers (Ewald.jl:185-209): a separable complex DFT over cells.     */

/* In-place DFT of a complex field f[c3][c2][c1] along each axis.
   sign = -1 forward (FFTW convention), +1 backward (no scaling). */
static void dft3(cplx *f, const int L[3], int sign)
{
    int maxL = L[0];
    if (L[1] > maxL) maxL = L[1];
    if (L[2] > maxL) maxL = L[2];
    cplx *line = malloc(sizeof(cplx) * maxL);
    cplx *out = malloc(sizeof(cplx) * maxL);
    long stride[3] = {1, L[0], (long)L[0] * L[1]};
    for (int ax = 0; ax < 3; ax++) {
        int n = L[ax];
        cplx *tw = malloc(sizeof(cplx) * n);
        for (int k = 0; k < n; k++) tw[k] = cexp(sign * 2 * M_PI * I * k / n);
        int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
        for (int a = 0; a < L[o1]; a++)
            for (int b = 0; b < L[o2]; b++) {
                long base = a * stride[o1] + b * stride[o2];
                for (int x = 0; x < n; x++) line[x] = f[base + x * stride[ax]];
                for (int k = 0; k < n; k++) {
                    cplx acc = 0;
                    for (int x = 0; x < n; x++) acc += line[x] * tw[(long)k * x % n];
                    out[k] = acc;
                }
                for (int x = 0; x < n; x++) f[base + x * stride[ax]] = out[x];
            }
        free(tw);
    }
    free(line);
    free(out);
}

/* F[A] for all (alpha, beta, i, j): FA[j][i][a][b][k3][k2][k1] complex, full spectrum. */
static cplx *ewald_FA(const so_spec *sp)
{
    long nc = ncells(sp);
    int na = sp->na;
    cplx *FA = malloc(sizeof(cplx) * nc * 9 * na * na);
    for (int j = 0; j < na; j++)
        for (int i = 0; i < na; i++)
            for (int ab = 0; ab < 9; ab++) {
                cplx *f = FA + (((long)j * na + i) * 9 + ab) * nc;
                const double *A = sp->ewald_A + ((long)j * na + i) * nc * 9;
                for (long c = 0; c < nc; c++) f[c] = A[c * 9 + ab];
                dft3(f, sp->L, -1);
            }
    return FA;
}

/* F[mu] with mu = -g S: Fmu[j][b][k3][k2][k1] */
static cplx *ewald_Fmu(const so_spec *sp, const double *S)
{
    long nc = ncells(sp);
    int na = sp->na;
    cplx *Fmu = malloc(sizeof(cplx) * nc * 3 * na);
    for (int j = 0; j < na; j++) {
        const double *g = sp->g + 9 * j;
        for (int b = 0; b < 3; b++) {
            cplx *f = Fmu + ((long)j * 3 + b) * nc;
            for (long c = 0; c < nc; c++) {
                const double *s = S + 3 * (c + nc * j);
                f[c] = -(g[3 * b + 0] * s[0] + g[3 * b + 1] * s[1] + g[3 * b + 2] * s[2]);
            }
            dft3(f, sp->L, -1);
        }
    }
    return Fmu;
}

/* accum_ewald_grad!, Ewald.jl:185-209 */
static void ewald_accum_grad(const so_spec *sp, const double *S, double *dE)
{
    long nc = ncells(sp);
    int na = sp->na;
    cplx *FA = ewald_FA(sp);
    cplx *Fmu = ewald_Fmu(sp, S);
    cplx *Fphi = calloc(nc * 3 * na, sizeof(cplx));
    /* Fphi[a,m,i] += conj(FA[a,b,m,i,j]) * Fmu[b,m,j] (loop order of :198) */
    for (int j = 0; j < na; j++)
        for (int i = 0; i < na; i++)
            for (long m = 0; m < nc; m++)
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++)
                        Fphi[((long)i * 3 + a) * nc + m] +=
                            conj(FA[(((long)j * na + i) * 9 + 3 * a + b) * nc + m]) *
                            Fmu[((long)j * 3 + b) * nc + m];
    for (int i = 0; i < na; i++)
        for (int a = 0; a < 3; a++) dft3(Fphi + ((long)i * 3 + a) * nc, sp->L, +1);
    /* dE -= g' phi, with phi = real(ifft)/Ncells */
    for (int i = 0; i < na; i++) {
        const double *g = sp->g + 9 * i;
        for (long c = 0; c < nc; c++) {
            double phi[3];
            for (int a = 0; a < 3; a++) phi[a] = creal(Fphi[((long)i * 3 + a) * nc + c]) / nc;
            double *d = dE + 3 * (c + nc * i);
            for (int b = 0; b < 3; b++)
                d[b] -= g[0 * 3 + b] * phi[0] + g[1 * 3 + b] * phi[1] + g[2 * 3 + b] * phi[2];
        }
    }
    free(FA);
    free(Fmu);
    free(Fphi);
}

/* ewald_energy, Ewald.jl:155-181: half spectrum m1 in 0..L1/2 with sqrt(2)
   weights on the planes whose conjugate partners are missing. */
double so_ewald_energy(const so_spec *sp, const double *S)
{
    long nc = ncells(sp);
    int na = sp->na;
    int L1 = sp->L[0], L2 = sp->L[1], L3 = sp->L[2];
    int m1max = L1 / 2; /* rfft keeps m1 = 0..floor(L1/2) */
    int even = (L1 % 2 == 0);
    cplx *FA = ewald_FA(sp);
    cplx *Fmu = ewald_Fmu(sp, S);
    /* scale Fmu on m1 in [1, end-1] (even) or [1, end] (odd) by sqrt(2) */
    for (int j = 0; j < na; j++)
        for (int b = 0; b < 3; b++)
            for (int m3 = 0; m3 < L3; m3++)
                for (int m2 = 0; m2 < L2; m2++)
                    for (int m1 = 1; m1 <= m1max; m1++) {
                        if (even && m1 == m1max) continue;
                        Fmu[((long)j * 3 + b) * nc + m1 + (long)L1 * (m2 + (long)L2 * m3)] *= sqrt(2.0);
                    }
    double E = 0;
    for (int j = 0; j < na; j++)
        for (int i = 0; i < na; i++)
            for (int m3 = 0; m3 < L3; m3++)
                for (int m2 = 0; m2 < L2; m2++)
                    for (int m1 = 0; m1 <= m1max; m1++) {
                        long m = m1 + (long)L1 * (m2 + (long)L2 * m3);
                        for (int a = 0; a < 3; a++)
                            for (int b = 0; b < 3; b++)
                                E += 0.5 * creal(conj(Fmu[((long)i * 3 + a) * nc + m]) *
                                                 conj(FA[(((long)j * na + i) * 9 + 3 * a + b) * nc + m]) *
                                                 Fmu[((long)j * 3 + b) * nc + m]);
                    }
    free(FA);
    free(Fmu);
    return E / nc;
}

/* ewald_grad_at / ewald_pairwise_grad_at, Ewald.jl:211-230 */
void so_ewald_grad_bruteforce(const so_spec *sp, const double *S, double *dE)
{
    long nc = ncells(sp);
    int na = sp->na;
    const int *L = sp->L;
    for (int i1 = 0; i1 < na; i1++)
        for (int a3 = 0; a3 < L[2]; a3++)
            for (int a2 = 0; a2 < L[1]; a2++)
                for (int a1 = 0; a1 < L[0]; a1++) {
                    double acc[3] = {0, 0, 0};
                    const double *g1 = sp->g + 9 * i1;
                    for (int i2 = 0; i2 < na; i2++) {
                        const double *g2 = sp->g + 9 * i2;
                        for (int b3 = 0; b3 < L[2]; b3++)
                            for (int b2 = 0; b2 < L[1]; b2++)
                                for (int b1 = 0; b1 < L[0]; b1++) {
                                    long off = site_index(L, wrap(b1 - a1, L[0]), wrap(b2 - a2, L[1]),
                                                          wrap(b3 - a3, L[2]), 0);
                                    const double *A = sp->ewald_A + (((long)i2 * na + i1) * nc + off) * 9;
                                    const double *s2 = S + 3 * site_index(L, b1, b2, b3, i2);
                                    double gs[3], Ags[3];
                                    for (int r = 0; r < 3; r++)
                                        gs[r] = g2[3 * r] * s2[0] + g2[3 * r + 1] * s2[1] + g2[3 * r + 2] * s2[2];
                                    for (int r = 0; r < 3; r++)
                                        Ags[r] = A[3 * r] * gs[0] + A[3 * r + 1] * gs[1] + A[3 * r + 2] * gs[2];
                                    for (int r = 0; r < 3; r++)
                                        acc[r] += g1[0 * 3 + r] * Ags[0] + g1[1 * 3 + r] * Ags[1] + g1[2 * 3 + r] * Ags[2];
                                }
                    }
                    double *d = dE + 3 * site_index(L, a1, a2, a3, i1);
                    d[0] = acc[0];
                    d[1] = acc[1];
                    d[2] = acc[2];
                }
}

/* precompute_dipole_ewald_aux at q = 0, Ewald.jl:61-152 */
void so_ewald_precompute(const double lv[9], int na, const double *pos, const int L[3],
                         const double demag[9], double mu0_muB2, double *Aout)
{
    so_ewald_precompute_sigma(lv, na, pos, L, demag, mu0_muB2, 3.0, Aout);
}

/* Ewald.jl:61-152 with the splitting width sigma = cbrt(V) / sigma_div (the
   reference fixes sigma_div = 3); the sum does not depend on it. */
void so_ewald_precompute_sigma(const double lv[9], int na, const double *pos, const int L[3],
                               const double demag[9], double mu0_muB2, double sigma_div, double *Aout)
{
    /* superlattice vectors (columns) and reciprocal vectors (columns) */
    double SL[9], RL[9];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) SL[3 * r + c] = lv[3 * r + c] * L[c];
    /* recipvecs = 2pi inv(latvecs)' */
    double a = SL[0], b = SL[1], c = SL[2], d = SL[3], e = SL[4], f = SL[5], g = SL[6], h = SL[7], k = SL[8];
    double det = a * (e * k - f * h) - b * (d * k - f * g) + c * (d * h - e * g);
    double inv[9] = {(e * k - f * h) / det, -(b * k - c * h) / det, (b * f - c * e) / det,
                     -(d * k - f * g) / det, (a * k - c * g) / det, -(a * f - c * d) / det,
                     (d * h - e * g) / det, -(a * h - b * g) / det, (a * e - b * d) / det};
    for (int r = 0; r < 3; r++)
        for (int cc = 0; cc < 3; cc++) RL[3 * r + cc] = 2 * M_PI * inv[3 * cc + r];
    double V = det;
    double Lc = cbrt(V);
    double sigma = Lc / sigma_div, s2 = sigma * sigma, s3 = sigma * sigma * sigma;
    double rmax = 6 * sqrt(2.0) * sigma, kmax = 6 * sqrt(2.0) / sigma;
    int nmax[3], mmax[3];
    for (int ax = 0; ax < 3; ax++) {
        double av[3] = {SL[ax], SL[3 + ax], SL[6 + ax]};
        double bv[3] = {RL[ax], RL[3 + ax], RL[6 + ax]};
        double bn = sqrt(bv[0] * bv[0] + bv[1] * bv[1] + bv[2] * bv[2]);
        double an = sqrt(av[0] * av[0] + av[1] * av[1] + av[2] * av[2]);
        double adb = (av[0] * bv[0] + av[1] * bv[1] + av[2] * bv[2]);
        nmax[ax] = (int)nearbyint(rmax / (adb / bn) + 1e-6) + 1; /* Julia round: ties to even */
        mmax[ax] = (int)nearbyint(kmax / (adb / an) + 1e-6);
    }
    long nc = (long)L[0] * L[1] * L[2];
#pragma omp parallel for collapse(2) schedule(dynamic)
    for (int j = 0; j < na; j++)
        for (long cell = 0; cell < nc; cell++) {
            int c1 = cell % L[0], c2 = (cell / L[0]) % L[1], c3 = cell / ((long)L[0] * L[1]);
            for (int i = 0; i < na; i++) {
                double acc[9] = {0};
                const int sw = g_mutation == 5; /* planted: atoms swapped */
                const int a0 = sw ? j : i, a1 = sw ? i : j;
                double fr[3] = {c1 + pos[3 * a1] - pos[3 * a0], c2 + pos[3 * a1 + 1] - pos[3 * a0 + 1],
                                c3 + pos[3 * a1 + 2] - pos[3 * a0 + 2]};
                double dr[3];
                for (int r = 0; r < 3; r++) dr[r] = lv[3 * r] * fr[0] + lv[3 * r + 1] * fr[1] + lv[3 * r + 2] * fr[2];
                if (g_mutation == 4) /* planted: fractional displacement */
                    for (int r = 0; r < 3; r++) dr[r] = fr[r];
                for (int n1 = -nmax[0]; n1 <= nmax[0]; n1++)
                    for (int n2 = -nmax[1]; n2 <= nmax[1]; n2++)
                        for (int n3 = -nmax[2]; n3 <= nmax[2]; n3++) {
                            double rv[3];
                            for (int r = 0; r < 3; r++)
                                rv[r] = dr[r] + SL[3 * r] * n1 + SL[3 * r + 1] * n2 + SL[3 * r + 2] * n3;
                            double r2 = rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
                            if (0 < r2 && r2 <= rmax * rmax) {
                                double rr = sqrt(r2), r3 = r2 * rr;
                                double rh[3] = {rv[0] / rr, rv[1] / rr, rv[2] / rr};
                                double erfc0 = erfc(rr / (sqrt(2.0) * sigma));
                                double gauss0 = sqrt(2 / M_PI) * (rr / sigma) * exp(-r2 / (2 * s2));
                                for (int p = 0; p < 3; p++)
                                    for (int q = 0; q < 3; q++)
                                        acc[3 * p + q] += (1 / (4 * M_PI)) *
                                            ((p == q ? 1.0 : 0.0) / r3 * (erfc0 + gauss0) -
                                             (3 * rh[p] * rh[q] / r3) * (erfc0 + (1 + r2 / (3 * s2)) * gauss0));
                            }
                        }
                for (int m1 = -mmax[0]; m1 <= mmax[0]; m1++)
                    for (int m2 = -mmax[1]; m2 <= mmax[1]; m2++)
                        for (int m3 = -mmax[2]; m3 <= mmax[2]; m3++) {
                            double kv[3];
                            for (int r = 0; r < 3; r++)
                                kv[r] = RL[3 * r] * m1 + RL[3 * r + 1] * m2 + RL[3 * r + 2] * m3;
                            double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
                            double eps2 = 1e-16;
                            if (k2 <= eps2) {
                                for (int p = 0; p < 9; p++) acc[p] += demag[p] / V;
                            } else if (k2 <= kmax * kmax) {
                                double phase = cos(-(kv[0] * dr[0] + kv[1] * dr[1] + kv[2] * dr[2]));
                                double w = phase * (1 / V) * (exp(-s2 * k2 / 2) / k2);
                                for (int p = 0; p < 3; p++)
                                    for (int q = 0; q < 3; q++) acc[3 * p + q] += w * kv[p] * kv[q];
                            }
                        }
                if (dr[0] == 0 && dr[1] == 0 && dr[2] == 0)
                    for (int p = 0; p < 3; p++) acc[4 * p] += -1 / (3 * pow(2 * M_PI, 1.5) * s3);
                double *A = Aout + (((long)j * na + i) * nc + cell) * 9;
                for (int p = 0; p < 9; p++) A[p] = acc[p] * mu0_muB2;
            }
        }
}

/* ------------------------------------------------------------------------ */
/* set_energy_grad_dipoles!, Interactions.jl:516-592                         */

static void grad_dipoles_impl(const so_spec *sp, const double *S, double *dE, int sun_mode)
{
    long nc = ncells(sp), ns = nsites(sp);
    const int *L = sp->L;
    memset(dE, 0, sizeof(double) * 3 * ns);
    /* Zeeman, :520-522: dE += g' B */
    for (long s = 0; s < ns; s++) {
        int i = s / nc;
        const double *g = sp->g + 9 * i, *B = sp->field + 3 * i;
        for (int b = 0; b < 3; b++) dE[3 * s + b] += g[0 * 3 + b] * B[0] + g[1 * 3 + b] * B[1] + g[2 * 3 + b] * B[2];
    }
    for (int i = 0; i < sp->na; i++) {
        /* anisotropy, :546-552 (dipole modes only) */
        if (!sun_mode && sp->kmax[i] >= 0) {
            for (long c = 0; c < nc; c++) {
                long s = c + nc * i;
                double E, gr[3];
                so_stevens_energy_grad(S + 3 * s, sp->stevens + 28 * i, sp->kmax[i], &E, gr);
                dE[3 * s] += gr[0];
                dE[3 * s + 1] += gr[1];
                dE[3 * s + 2] += gr[2];
            }
        }
        /* bonds, :554-591: non-culled bonds of sublattice i, scatter both ends */
        for (int b = 0; b < sp->nb; b++) {
            const int *bd = sp->bond + 6 * b;
            if (bd[0] != i || bd[5]) continue;
            const double *J = sp->bilin + 9 * b;
            for (int c3 = 0; c3 < L[2]; c3++)
                for (int c2 = 0; c2 < L[1]; c2++)
                    for (int c1 = 0; c1 < L[0]; c1++) {
                        long si = site_index(L, c1, c2, c3, i);
                        long sj = site_index(L, wrap(c1 + bd[2], L[0]), wrap(c2 + bd[3], L[1]),
                                             wrap(c3 + bd[4], L[2]), bd[1]);
                        const double *Si = S + 3 * si, *Sj = S + 3 * sj;
                        double gi[3], gj[3];
                        for (int r = 0; r < 3; r++) {
                            gi[r] = J[3 * r] * Sj[0] + J[3 * r + 1] * Sj[1] + J[3 * r + 2] * Sj[2];
                            gj[r] = J[r] * Si[0] + J[3 + r] * Si[1] + J[6 + r] * Si[2];
                        }
                        for (int r = 0; r < 3; r++) {
                            dE[3 * si + r] += gi[r];
                            dE[3 * sj + r] += gj[r];
                        }
                        if (!sun_mode && sp->has_biquad[b]) {
                            const double *Bq = sp->biquad + 25 * b;
                            double Qi[5], Qj[5], Gi[5][3], Gj[5][3];
                            quadrupole(Si, Qi);
                            quadrupole(Sj, Qj);
                            grad_quadrupole(Si, Gi);
                            grad_quadrupole(Sj, Gj);
                            /* dE_i += (B Qj)' dQi ; dE_j += (B' Qi)' dQj */
                            double BQj[5], BtQi[5];
                            for (int p = 0; p < 5; p++) {
                                BQj[p] = 0;
                                BtQi[p] = 0;
                                for (int q = 0; q < 5; q++) {
                                    BQj[p] += Bq[5 * p + q] * Qj[q];
                                    BtQi[p] += Bq[5 * q + p] * Qi[q];
                                }
                            }
                            for (int r = 0; r < 3; r++) {
                                double ai = 0, aj = 0;
                                for (int p = 0; p < 5; p++) {
                                    ai += BQj[p] * Gi[p][r];
                                    aj += BtQi[p] * Gj[p][r];
                                }
                                dE[3 * si + r] += ai;
                                dE[3 * sj + r] += aj;
                            }
                        }
                    }
        }
    }
    if (sp->ewald_A) ewald_accum_grad(sp, S, dE);
}

void so_grad_dipoles(const so_spec *sp, const double *S, double *dE)
{
    grad_dipoles_impl(sp, S, dE, sp->N > 0);
}

/* set_energy_grad_coherents!, Interactions.jl:596-669 */
void so_grad_coherents(const so_spec *sp, const double *Zr, double *HZr)
{
    int N = sp->N;
    long nc = ncells(sp), ns = nsites(sp);
    const int *L = sp->L;
    sun_ops o;
    sun_ops_init(&o, N);
    const cplx *Z = (const cplx *)Zr;
    cplx *HZ = (cplx *)HZr;
    memset(HZ, 0, sizeof(cplx) * N * ns);
    double *dip = malloc(sizeof(double) * 3 * ns);
    double *dE = malloc(sizeof(double) * 3 * ns);
    for (long s = 0; s < ns; s++) expected_spin1(&o, Z + N * s, dip + 3 * s);
    grad_dipoles_impl(sp, dip, dE, 1);
    for (int i = 0; i < sp->na; i++) {
        const double *Lam = sp->Lambda + (long)2 * N * N * i;
        for (long c = 0; c < nc; c++) {
            long s = c + nc * i;
            /* mul_spin_matrices: (Lambda + dE.S) Z */
            for (int a = 0; a < N; a++) {
                cplx acc = 0;
                for (int b = 0; b < N; b++) {
                    cplx h = Lam[2 * (a * N + b)] + I * Lam[2 * (a * N + b) + 1];
                    for (int al = 0; al < 3; al++) h += dE[3 * s + al] * o.S[al][a * N + b];
                    acc += h * Z[N * s + b];
                }
                HZ[N * s + a] += acc;
            }
        }
        for (int b = 0; b < sp->nb; b++) {
            const int *bd = sp->bond + 6 * b;
            int t0, t1;
            if (bd[0] != i || bd[5]) continue;
            /* general couplings, Interactions.jl:659-666: HZi += (A Zi) <B>j, HZj += <A>i (B Zj) */
            if (gen_terms(sp, b, &t0, &t1))
                for (long c = 0; c < nc; c++) {
                    const int c1 = (int)(c % L[0]), c2 = (int)((c / L[0]) % L[1]), c3 = (int)(c / ((long)L[0] * L[1]));
                    long si = site_index(L, c1, c2, c3, i);
                    long sj = site_index(L, wrap(c1 + bd[2], L[0]), wrap(c2 + bd[3], L[1]), wrap(c3 + bd[4], L[2]),
                                         bd[1]);
                    for (int t = t0; t < t1; t++) {
                        const double *A = sp->gen_A + (long)2 * N * N * t, *Bm = sp->gen_B + (long)2 * N * N * t;
                        const double Abar = expect_op(N, A, Z + N * si), Bbar = expect_op(N, Bm, Z + N * sj);
                        add_op_ket(N, A, Z + N * si, Bbar, HZ + N * si);
                        add_op_ket(N, Bm, Z + N * sj, Abar, HZ + N * sj);
                    }
                }
            if (!sp->has_biquad[b]) continue;
            const double *Bq = sp->biquad + 25 * b;
            for (int c3 = 0; c3 < L[2]; c3++)
                for (int c2 = 0; c2 < L[1]; c2++)
                    for (int c1 = 0; c1 < L[0]; c1++) {
                        long si = site_index(L, c1, c2, c3, i);
                        long sj = site_index(L, wrap(c1 + bd[2], L[0]), wrap(c2 + bd[3], L[1]),
                                             wrap(c3 + bd[4], L[2]), bd[1]);
                        double Qi[5], Qj[5], dQi[5], dQj[5];
                        expected_quadrupole1(&o, Z + N * si, Qi);
                        expected_quadrupole1(&o, Z + N * sj, Qj);
                        for (int p = 0; p < 5; p++) {
                            dQi[p] = 0;
                            dQj[p] = 0;
                            for (int q = 0; q < 5; q++) {
                                dQi[p] += Bq[5 * p + q] * Qj[q];
                                dQj[p] += Bq[5 * q + p] * Qi[q];
                            }
                        }
                        /* mul_quadrupole_matrices: (dE/dQ . Q) Z, Spin.jl:151-167 */
                        for (int a = 0; a < N; a++) {
                            cplx ai = 0, aj = 0;
                            for (int bb = 0; bb < N; bb++) {
                                cplx hi = 0, hj = 0;
                                for (int p = 0; p < 5; p++) {
                                    hi += dQi[p] * o.Q[p][a * N + bb];
                                    hj += dQj[p] * o.Q[p][a * N + bb];
                                }
                                ai += hi * Z[N * si + bb];
                                aj += hj * Z[N * sj + bb];
                            }
                            HZ[N * si + a] += ai;
                            HZ[N * sj + a] += aj;
                        }
                    }
        }
    }
    free(dip);
    free(dE);
}

/* energy, Interactions.jl:401-508 */
double so_energy(const so_spec *sp, const double *S, const double *Zr)
{
    long nc = ncells(sp), ns = nsites(sp);
    const int *L = sp->L;
    int N = sp->N;
    const cplx *Z = (const cplx *)Zr;
    sun_ops o;
    if (N > 0) sun_ops_init(&o, N);
    double E = 0;
    for (long s = 0; s < ns; s++) {
        int i = s / nc;
        const double *g = sp->g + 9 * i, *B = sp->field + 3 * i, *d = S + 3 * s;
        for (int r = 0; r < 3; r++) E += B[r] * (g[3 * r] * d[0] + g[3 * r + 1] * d[1] + g[3 * r + 2] * d[2]);
    }
    for (int i = 0; i < sp->na; i++) {
        double Ei = 0;
        if (N == 0) {
            for (long c = 0; c < nc; c++) {
                double e, gr[3];
                so_stevens_energy_grad(S + 3 * (c + nc * i), sp->stevens + 28 * i, sp->kmax[i], &e, gr);
                Ei += e;
            }
        } else {
            const double *Lam = sp->Lambda + (long)2 * N * N * i;
            for (long c = 0; c < nc; c++) {
                const cplx *z = Z + N * (c + nc * i);
                cplx acc = 0;
                for (int a = 0; a < N; a++)
                    for (int b = 0; b < N; b++)
                        acc += conj(z[a]) * (Lam[2 * (a * N + b)] + I * Lam[2 * (a * N + b) + 1]) * z[b];
                Ei += creal(acc);
            }
        }
        for (int b = 0; b < sp->nb; b++) {
            const int *bd = sp->bond + 6 * b;
            if (bd[0] != i || bd[5]) continue;
            const double *J = sp->bilin + 9 * b;
            for (int c3 = 0; c3 < L[2]; c3++)
                for (int c2 = 0; c2 < L[1]; c2++)
                    for (int c1 = 0; c1 < L[0]; c1++) {
                        long si = site_index(L, c1, c2, c3, i);
                        long sj = site_index(L, wrap(c1 + bd[2], L[0]), wrap(c2 + bd[3], L[1]),
                                             wrap(c3 + bd[4], L[2]), bd[1]);
                        const double *Si = S + 3 * si, *Sj = S + 3 * sj;
                        if (N > 0)
                            Ei += sp->scalar[b] * sp->kappa[i] * sp->kappa[bd[1]];
                        else
                            Ei += sp->scalar[b];
                        for (int r = 0; r < 3; r++)
                            Ei += Si[r] * (J[3 * r] * Sj[0] + J[3 * r + 1] * Sj[1] + J[3 * r + 2] * Sj[2]);
                        if (sp->has_biquad[b]) {
                            double Qi[5], Qj[5];
                            if (N == 0) {
                                quadrupole(Si, Qi);
                                quadrupole(Sj, Qj);
                            } else {
                                expected_quadrupole1(&o, Z + N * si, Qi);
                                expected_quadrupole1(&o, Z + N * sj, Qj);
                            }
                            const double *Bq = sp->biquad + 25 * b;
                            for (int p = 0; p < 5; p++)
                                for (int q = 0; q < 5; q++) Ei += Qi[p] * Bq[5 * p + q] * Qj[q];
                        }
                        /* general couplings, Interactions.jl:496-503 */
                        int t0, t1;
                        if (gen_terms(sp, b, &t0, &t1))
                            for (int t = t0; t < t1; t++)
                                Ei += expect_op(N, sp->gen_A + (long)2 * N * N * t, Z + N * si) *
                                      expect_op(N, sp->gen_B + (long)2 * N * N * t, Z + N * sj);
                    }
        }
        E += Ei;
    }
    if (sp->ewald_A) E += so_ewald_energy(sp, S);
    return E;
}

/* ------------------------------------------------------------------------ */
/* Integrators.jl:266-452                                                    */

static inline void cross(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* rhs_dipole!, :266-275 */
static void rhs_dipole(const so_spec *sp, long ns, double *dS, const double *S, const double *xi,
                       const double *dE, double dt, double lambda)
{
    const long nc = ncells(sp);
    for (long s = 0; s < ns; s++) {
        const double *Sv = S + 3 * s, *g = dE + 3 * s;
        double v[3];
        if (lambda == 0) {
            for (int r = 0; r < 3; r++) v[r] = dt * g[r];
        } else {
            double sxg[3];
            cross(Sv, g, sxg);
            if (g_mutation == 3) /* planted: damping with S / kappa */
                for (int r = 0; r < 3; r++) sxg[r] /= sp->kappa[s / nc];
            for (int r = 0; r < 3; r++) v[r] = xi[3 * s + r] + dt * g[r] - dt * lambda * sxg[r];
        }
        double o[3];
        cross(Sv, v, o);
        for (int r = 0; r < 3; r++) dS[3 * s + r] = -o[r];
    }
}

/* normalize_dipole, System.jl:421-423 */
static inline void normalize_dipole(double *s, double kappa)
{
    if (kappa == 0) {
        s[0] = s[1] = s[2] = 0;
        return;
    }
    double n = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (int r = 0; r < 3; r++) s[r] = kappa * (s[r] / n);
}

/* normalize_ket, System.jl:416-418 */
static inline void normalize_ket(int N, cplx *z, double kappa)
{
    if (kappa == 0) {
        for (int a = 0; a < N; a++) z[a] = 0;
        return;
    }
    double n2 = 0;
    for (int a = 0; a < N; a++) n2 += creal(conj(z[a]) * z[a]);
    double d = sqrt(n2 / kappa);
    for (int a = 0; a < N; a++) z[a] = z[a] / d;
}

/* fill_noise!, :288-298 */
static void fill_noise(long n, double *xi, const double *raw, double dt, double lambda, double kT)
{
    if (lambda == 0 || kT == 0 || !raw) {
        memset(xi, 0, sizeof(double) * n);
        return;
    }
    double sc = sqrt((g_mutation == 1 ? 1 : 2) * dt * lambda * kT);
    for (long k = 0; k < n; k++) xi[k] = raw[k] * sc;
}

/* step!(::System{0}, ::Langevin), :313-332 */
void so_step_langevin_dipole(const so_spec *sp, double *S, const double *noise, double dt,
                             double damping, double kT)
{
    long ns = nsites(sp), nc = ncells(sp);
    double *Sp = malloc(sizeof(double) * 3 * ns), *dS1 = malloc(sizeof(double) * 3 * ns),
           *dS2 = malloc(sizeof(double) * 3 * ns), *xi = malloc(sizeof(double) * 3 * ns),
           *dE = malloc(sizeof(double) * 3 * ns);
    fill_noise(3 * ns, xi, noise, dt, damping, kT);
    so_grad_dipoles(sp, S, dE);
    rhs_dipole(sp, ns, dS1, S, xi, dE, dt, damping);
    for (long s = 0; s < ns; s++) {
        for (int r = 0; r < 3; r++) Sp[3 * s + r] = S[3 * s + r] + dS1[3 * s + r];
        normalize_dipole(Sp + 3 * s, sp->kappa[s / nc]);
    }
    so_grad_dipoles(sp, Sp, dE);
    if (g_mutation == 2) memset(xi, 0, sizeof(double) * 3 * ns); /* planted: corrector without noise */
    rhs_dipole(sp, ns, dS2, Sp, xi, dE, dt, damping);
    for (long s = 0; s < ns; s++) {
        for (int r = 0; r < 3; r++) S[3 * s + r] = S[3 * s + r] + (dS1[3 * s + r] + dS2[3 * s + r]) / 2;
        normalize_dipole(S + 3 * s, sp->kappa[s / nc]);
    }
    free(Sp);
    free(dS1);
    free(dS2);
    free(xi);
    free(dE);
}

/* rhs_sun!, :277-286, with proj from MathBasics.jl:62 */
static void rhs_sun(int N, long ns, cplx *dZ, const cplx *Z, const cplx *zeta, const cplx *HZ,
                    double dt, double lambda)
{
    for (long s = 0; s < ns; s++) {
        const cplx *z = Z + N * s, *h = HZ + N * s;
        if (lambda == 0) {
            for (int a = 0; a < N; a++) dZ[N * s + a] = -I * dt * h[a];
        } else {
            cplx v[NMAX];
            for (int a = 0; a < N; a++) v[a] = zeta[N * s + a] + dt * (I + lambda) * h[a];
            cplx nv = 0;
            double nn = 0;
            for (int a = 0; a < N; a++) {
                nv += conj(z[a]) * v[a];
                nn += creal(conj(z[a]) * z[a]);
            }
            for (int a = 0; a < N; a++) dZ[N * s + a] = -(v[a] - z[a] * (nv / nn));
        }
    }
}

/* step!(::System{N}, ::Langevin), :335-357 */
void so_step_langevin_sun(const so_spec *sp, double *Zr, double *S, const double *noise, double dt,
                          double damping, double kT)
{
    int N = sp->N;
    long ns = nsites(sp), nc = ncells(sp);
    cplx *Z = (cplx *)Zr;
    cplx *Zp = malloc(sizeof(cplx) * N * ns), *dZ1 = malloc(sizeof(cplx) * N * ns),
         *dZ2 = malloc(sizeof(cplx) * N * ns), *zeta = malloc(sizeof(cplx) * N * ns),
         *HZ = malloc(sizeof(cplx) * N * ns);
    fill_noise(2 * N * ns, (double *)zeta, noise, dt, damping, kT);
    so_grad_coherents(sp, (double *)Z, (double *)HZ);
    rhs_sun(N, ns, dZ1, Z, zeta, HZ, dt, damping);
    for (long s = 0; s < ns; s++) {
        for (int a = 0; a < N; a++) Zp[N * s + a] = Z[N * s + a] + dZ1[N * s + a];
        normalize_ket(N, Zp + N * s, sp->kappa[s / nc]);
    }
    so_grad_coherents(sp, (double *)Zp, (double *)HZ);
    if (g_mutation == 2) memset(zeta, 0, sizeof(cplx) * N * ns); /* planted: corrector without noise */
    rhs_sun(N, ns, dZ2, Zp, zeta, HZ, dt, damping);
    for (long s = 0; s < ns; s++) {
        for (int a = 0; a < N; a++) Z[N * s + a] = Z[N * s + a] + (dZ1[N * s + a] + dZ2[N * s + a]) / 2;
        normalize_ket(N, Z + N * s, sp->kappa[s / nc]);
    }
    so_expected_spin(N, ns, Zr, S);
    free(Zp);
    free(dZ1);
    free(dZ2);
    free(zeta);
    free(HZ);
}

/* fast_isapprox, :361-371 */
static int fast_isapprox(long n, const double *x, const double *y, long comps, double atol)
{
    double acc = 0;
    for (long s = 0; s < n; s++) {
        double d2 = 0;
        for (long r = 0; r < comps; r++) {
            double d = x[comps * s + r] - y[comps * s + r];
            d2 += d * d;
        }
        acc += d2;
        if (acc > atol * atol) return 0;
    }
    return !isnan(acc);
}

/* step!(::System{0}, ::ImplicitMidpoint), :379-413 */
int so_step_midpoint_dipole(const so_spec *sp, double *S, double dt, double damping, double tol,
                            int max_iters)
{
    long ns = nsites(sp), nc = ncells(sp);
    double atol = tol * sqrt((double)ns);
    double *dS = malloc(sizeof(double) * 3 * ns), *Sh = malloc(sizeof(double) * 3 * ns),
           *S1 = malloc(sizeof(double) * 3 * ns), *S2 = malloc(sizeof(double) * 3 * ns),
           *dE = malloc(sizeof(double) * 3 * ns);
    double *xi = calloc(3 * ns, sizeof(double)); /* kT == 0 is enforced by the constructor */
    memcpy(S1, S, sizeof(double) * 3 * ns);
    memcpy(S2, S, sizeof(double) * 3 * ns);
    int ret = -1;
    for (int it = 1; it <= max_iters; it++) {
        for (long s = 0; s < ns; s++) {
            for (int r = 0; r < 3; r++) Sh[3 * s + r] = (S[3 * s + r] + S1[3 * s + r]) / 2;
            normalize_dipole(Sh + 3 * s, sp->kappa[s / nc]);
        }
        so_grad_dipoles(sp, Sh, dE);
        rhs_dipole(sp, ns, dS, Sh, xi, dE, dt, damping);
        for (long k = 0; k < 3 * ns; k++) S2[k] = S[k] + dS[k];
        if (fast_isapprox(ns, S1, S2, 3, atol)) {
            for (long s = 0; s < ns; s++) {
                for (int r = 0; r < 3; r++) S[3 * s + r] = S2[3 * s + r];
                normalize_dipole(S + 3 * s, sp->kappa[s / nc]);
            }
            ret = it;
            break;
        }
        double *t = S1;
        S1 = S2;
        S2 = t;
    }
    free(dS);
    free(Sh);
    free(S1);
    free(S2);
    free(dE);
    free(xi);
    return ret;
}

/* step!(::System{N}, ::ImplicitMidpoint), :422-452 */
int so_step_midpoint_sun(const so_spec *sp, double *Zr, double *S, double dt, double damping,
                         double tol, int max_iters)
{
    int N = sp->N;
    long ns = nsites(sp), nc = ncells(sp);
    double atol = tol * sqrt((double)ns);
    cplx *Z = (cplx *)Zr;
    cplx *dZ = malloc(sizeof(cplx) * N * ns), *Zb = malloc(sizeof(cplx) * N * ns),
         *Z1 = malloc(sizeof(cplx) * N * ns), *Z2 = malloc(sizeof(cplx) * N * ns),
         *HZ = malloc(sizeof(cplx) * N * ns);
    cplx *zeta = calloc(N * ns, sizeof(cplx));
    memcpy(Z1, Z, sizeof(cplx) * N * ns);
    memcpy(Z2, Z, sizeof(cplx) * N * ns);
    int ret = -1;
    for (int it = 1; it <= max_iters; it++) {
        for (long k = 0; k < N * ns; k++) Zb[k] = (Z[k] + Z1[k]) / 2;
        so_grad_coherents(sp, (double *)Zb, (double *)HZ);
        rhs_sun(N, ns, dZ, Zb, zeta, HZ, dt, damping);
        for (long k = 0; k < N * ns; k++) Z2[k] = Z[k] + dZ[k];
        if (fast_isapprox(ns, (double *)Z1, (double *)Z2, 2 * N, atol)) {
            for (long s = 0; s < ns; s++) {
                for (int a = 0; a < N; a++) Z[N * s + a] = Z2[N * s + a];
                normalize_ket(N, Z + N * s, sp->kappa[s / nc]);
            }
            so_expected_spin(N, ns, Zr, S);
            ret = it;
            break;
        }
        cplx *t = Z1;
        Z1 = Z2;
        Z2 = t;
    }
    free(dZ);
    free(Zb);
    free(Z1);
    free(Z2);
    free(HZ);
    free(zeta);
    return ret;
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw, SC'11)
   and the Box-Muller map used for device-side noise.                       */

void so_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void so_philox_normals(uint64_t seed, uint64_t step, long ns, int per_site, double *out)
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    int nblk = (per_site + 3) / 4;
    for (long s = 0; s < ns; s++)
        for (int b = 0; b < nblk; b++) {
            uint32_t ctr[4] = {(uint32_t)s, (uint32_t)b, (uint32_t)step, (uint32_t)(step >> 32)}, u[4];
            so_philox4x32(ctr, key, u);
            for (int p = 0; p < 2; p++) {
                double ua = ((double)u[2 * p] + 1.0) * (1.0 / 4294967296.0);
                double ub = (double)u[2 * p + 1] * (1.0 / 4294967296.0);
                double r = sqrt(-2.0 * log(ua));
                double th = 2.0 * M_PI * ub;
                int k0 = 4 * b + 2 * p;
                if (k0 < per_site) out[(long)per_site * s + k0] = r * cos(th);
                if (k0 + 1 < per_site) out[(long)per_site * s + k0 + 1] = r * sin(th);
            }
        }
}

/* nsteps of step!(sys, Langevin(dt; damping, kT)) (Integrators.jl:313-357)
   drawing the noise of step step0 + k from the device's Philox stream
   (so_philox_normals under `key`: 3 standard normals per site, or 2N normals
   / sqrt(2) as the (re, im) parts of N complex normals), recording
   energy(sys) after every `every` steps into E[nsteps / every].  S holds the
   dipoles (the expected spins in :SUN mode, kept in step with Z). */
void so_langevin_run(const so_spec *sp, double *S, double *Zr, double dt, double damping, double kT, uint64_t key,
                     uint64_t step0, long nsteps, long every, double *E)
{
    const long ns = nsites(sp);
    const int per = sp->N == 0 ? 3 : 2 * sp->N;
    double *raw = malloc(sizeof(double) * per * ns);
    for (long k = 0; k < nsteps; k++) {
        so_philox_normals(key, step0 + (uint64_t)k, ns, per, raw);
        if (sp->N == 0) {
            so_step_langevin_dipole(sp, S, raw, dt, damping, kT);
        } else {
            for (long q = 0; q < per * ns; q++) raw[q] *= M_SQRT1_2;
            so_step_langevin_sun(sp, Zr, S, raw, dt, damping, kT);
        }
        if (every > 0 && (k + 1) % every == 0) E[(k + 1) / every - 1] = so_energy(sp, S, sp->N ? Zr : NULL);
    }
    free(raw);
}

/* ------------------------------------------------------------------------ */
/* Metropolis updates of single dipoles.                                     */

/* ewald_energy_delta, Ewald.jl:234-241: dS . (field at the site by the
   pairwise sum of Ewald.jl:211-230) + dmu . A[0, i, i] dmu / 2, dmu = -g dS. */
static double ewald_delta(const so_spec *sp, const double *S, long site, const double dS[3])
{
    const long nc = ncells(sp);
    const int *L = sp->L;
    const int na = sp->na, i = (int)(site / nc);
    const long c = site % nc;
    const int c1 = (int)(c % L[0]), c2 = (int)((c / L[0]) % L[1]), c3 = (int)(c / ((long)L[0] * L[1]));
    const double *g = sp->g + 9 * i;
    double grad[3] = {0, 0, 0};
    for (int i2 = 0; i2 < na; i2++) {
        const double *g2 = sp->g + 9 * i2;
        for (int b3 = 0; b3 < L[2]; b3++)
            for (int b2 = 0; b2 < L[1]; b2++)
                for (int b1 = 0; b1 < L[0]; b1++) {
                    long off = site_index(L, wrap(b1 - c1, L[0]), wrap(b2 - c2, L[1]), wrap(b3 - c3, L[2]), 0);
                    const double *A = sp->ewald_A + (((long)i2 * na + i) * nc + off) * 9;
                    const double *s2 = S + 3 * site_index(L, b1, b2, b3, i2);
                    double gs[3], Ags[3];
                    for (int r = 0; r < 3; r++) gs[r] = g2[3 * r] * s2[0] + g2[3 * r + 1] * s2[1] + g2[3 * r + 2] * s2[2];
                    for (int r = 0; r < 3; r++) Ags[r] = A[3 * r] * gs[0] + A[3 * r + 1] * gs[1] + A[3 * r + 2] * gs[2];
                    for (int r = 0; r < 3; r++) grad[r] += g[r] * Ags[0] + g[3 + r] * Ags[1] + g[6 + r] * Ags[2];
                }
    }
    double dmu[3], Adm[3], dE = 0;
    for (int r = 0; r < 3; r++) dmu[r] = -(g[3 * r] * dS[0] + g[3 * r + 1] * dS[1] + g[3 * r + 2] * dS[2]);
    const double *A0 = sp->ewald_A + (((long)i * na + i) * nc) * 9;
    for (int r = 0; r < 3; r++) Adm[r] = A0[3 * r] * dmu[0] + A0[3 * r + 1] * dmu[1] + A0[3 * r + 2] * dmu[2];
    for (int r = 0; r < 3; r++) dE += dS[r] * grad[r] + dmu[r] * Adm[r] / 2;
    return dE;
}

/* local_energy_change, Interactions.jl:311-380 (:dipole mode), with
   ewald_energy_delta, Ewald.jl:234-241 (field at the site by the pairwise
   sum of Ewald.jl:211-230).  NAN when a bond maps the site onto itself
   (the reference raises "Energy delta for self-interaction not supported"). */
double so_local_energy_change(const so_spec *sp, const double *S, long site, const double Snew[3])
{
    const long nc = ncells(sp);
    const int *L = sp->L;
    const int i = (int)(site / nc);
    const long c = site % nc;
    const int c1 = (int)(c % L[0]), c2 = (int)((c / L[0]) % L[1]), c3 = (int)(c / ((long)L[0] * L[1]));
    const double *S0 = S + 3 * site;
    double dS[3] = {Snew[0] - S0[0], Snew[1] - S0[1], Snew[2] - S0[2]};
    double dE = 0;
    /* Zeeman: dot(B, g, dS) */
    const double *g = sp->g + 9 * i, *B = sp->field + 3 * i;
    for (int r = 0; r < 3; r++) dE += B[r] * (g[3 * r] * dS[0] + g[3 * r + 1] * dS[1] + g[3 * r + 2] * dS[2]);
    /* single-ion anisotropy */
    double En, Eo, gr[3];
    so_stevens_energy_grad(Snew, sp->stevens + 28 * i, sp->kmax[i], &En, gr);
    so_stevens_energy_grad(S0, sp->stevens + 28 * i, sp->kmax[i], &Eo, gr);
    dE += En - Eo;
    /* every pair coupling of the site's atom, culled or not */
    for (int b = 0; b < sp->nb; b++) {
        const int *bd = sp->bond + 6 * b;
        if (bd[0] != i) continue;
        long sj = site_index(L, wrap(c1 + bd[2], L[0]), wrap(c2 + bd[3], L[1]), wrap(c3 + bd[4], L[2]), bd[1]);
        if (sj == site) return NAN;
        const double *Sj = S + 3 * sj, *J = sp->bilin + 9 * b;
        for (int r = 0; r < 3; r++) dE += dS[r] * (J[3 * r] * Sj[0] + J[3 * r + 1] * Sj[1] + J[3 * r + 2] * Sj[2]);
        if (sp->has_biquad[b]) {
            double Qn[5], Qo[5], Qj[5];
            quadrupole(Snew, Qn);
            quadrupole(S0, Qo);
            quadrupole(Sj, Qj);
            const double *Bq = sp->biquad + 25 * b;
            for (int p = 0; p < 5; p++)
                for (int q = 0; q < 5; q++) dE += (Qn[p] - Qo[p]) * Bq[5 * p + q] * Qj[q];
        }
    }
    if (sp->ewald_A) dE += ewald_delta(sp, S, site, dS);
    return dE;
}

/* The proposal of a LocalSampler (Samplers.jl:1-100): kind 0 propose_uniform
   (randspin, System.jl:444-447), 1 propose_flip (flip, System.jl:440-442),
   2 propose_delta(mag) (perturbed_spin, System.jl:454-460; infinite mag is
   randspin).  n: three standard normals. */
static void local_proposal(int kind, double mag, double kappa, const double *S0, const double n[3], double out[3])
{
    if (kind == 1) {
        for (int r = 0; r < 3; r++) out[r] = -S0[r];
        return;
    }
    if (kind == 0 || isinf(mag)) {
        for (int r = 0; r < 3; r++) out[r] = n[r];
    } else {
        for (int r = 0; r < 3; r++) out[r] = S0[r] + mag * kappa * n[r];
    }
    normalize_dipole(out, kappa);
}

static void box_muller_pair(uint32_t a, uint32_t b, double *x, double *y)
{
    double ua = ((double)a + 1.0) * (1.0 / 4294967296.0);
    double ub = (double)b * (1.0 / 4294967296.0);
    double r = sqrt(-2.0 * log(ua)), th = 2.0 * M_PI * ub;
    *x = r * cos(th);
    *y = r * sin(th);
}

/* One Metropolis trial at `site` with the given draws; accumulates dE, dS of
   accepted moves into acc[4] (LocalSampler's step!, Samplers.jl:133-153).
   Returns -1 when the energy change is undefined (self-interaction). */
static int local_trial(const so_spec *sp, double *S, long site, int nprop, const int *kind, const double *mag,
                        const double *cum, double kT, double u_mix, double u_acc, const double n[3], double acc[4])
{
    int k = nprop - 1;
    for (int p = 0; p < nprop - 1; p++)
        if (u_mix < cum[p]) {
            k = p;
            break;
        }
    const int i = (int)(site / ncells(sp));
    double Sn[3];
    local_proposal(kind[k], mag[k], sp->kappa[i], S + 3 * site, n, Sn);
    const double dE = so_local_energy_change(sp, S, site, Sn);
    if (isnan(dE)) return -1;
    const int accept = kT == 0 ? dE <= 0 : u_acc <= exp(-dE / kT);
    if (accept) {
        acc[0] += dE;
        for (int r = 0; r < 3; r++) {
            acc[1 + r] += Sn[r] - S[3 * site + r];
            S[3 * site + r] = Sn[r];
        }
    }
    return 0;
}

/* Colour-class passes: the device's ordering of LocalSampler trials.  A
   sweep is two passes; a pass visits the colour classes 0..ncolors-1 in turn
   and tries every site of a class with probability 1/2 (no two sites of one
   class share a bond, so their order is immaterial).  Class of site (c, i):
   color[i][c3 % m3][c2 % m2][c1 % m1].  nsweeps may be fractional: there are
   ceil(2 nsweeps) passes, the last with probability (2 nsweeps - passes + 1)/2,
   so a call makes nsweeps * nsites trials on average, as the reference's
   ceil(nsweeps * nsites) does.  (With every site tried in a fixed order,
   Metropolis flips of Ising-like spins keep the domain walls of a chain
   moving deterministically and the sweep is not ergodic.)
   Draws of site s (Julia index) in pass P = 2 sweep0 + k: Philox4x32-10
   under `key` of the counters (s, 0x4D430000, P) -> four normals by
   Box-Muller (the first three are used) and (s, 0x4D430001, P) -> uniforms
   u_acc, u_mix, u_try (each word / 2^32).  Returns -1 on a self-interaction. */
int so_local_sweeps(const so_spec *sp, double *S, const int m[3], const int *color, int ncolors, int nprop,
                    const int *kind, const double *mag, const double *cum, double kT, uint64_t key,
                    uint64_t sweep0, double nsweeps, double acc[4])
{
    const int *L = sp->L;
    const uint32_t kk[2] = {(uint32_t)key, (uint32_t)(key >> 32)};
    const long npass = (long)ceil(2 * nsweeps);
    for (long k = 0; k < npass; k++) {
        const uint64_t P = 2 * sweep0 + (uint64_t)k;
        const double prob = k == npass - 1 ? (2 * nsweeps - (double)(npass - 1)) / 2 : 0.5;
        for (int col = 0; col < ncolors; col++)
            for (int i = 0; i < sp->na; i++)
                for (int c3 = 0; c3 < L[2]; c3++)
                    for (int c2 = 0; c2 < L[1]; c2++)
                        for (int c1 = 0; c1 < L[0]; c1++) {
                            const int cls = c1 % m[0] + m[0] * (c2 % m[1] + m[1] * (c3 % m[2] + m[2] * i));
                            if (color[cls] != col) continue;
                            const long site = site_index(L, c1, c2, c3, i);
                            uint32_t ctr[4] = {(uint32_t)site, 0x4D430000u, (uint32_t)P, (uint32_t)(P >> 32)}, u[4];
                            so_philox4x32(ctr, kk, u);
                            double n[4];
                            box_muller_pair(u[0], u[1], &n[0], &n[1]);
                            box_muller_pair(u[2], u[3], &n[2], &n[3]);
                            ctr[1] = 0x4D430001u;
                            so_philox4x32(ctr, kk, u);
                            const double inv = 1.0 / 4294967296.0;
                            if (!((double)u[2] * inv < prob)) continue;
                            if (local_trial(sp, S, site, nprop, kind, mag, cum, kT, (double)u[1] * inv,
                                            (double)u[0] * inv, n, acc))
                                return -1;
                        }
    }
    return 0;
}

/* LocalSampler's step! as sequential trials at random sites (Samplers.jl:
   133-153), the device's form when the Ewald term couples every pair of
   sites: ceil(nsweeps * ns) trials, trial T = sweep0 * ns + t drawing from
   Philox4x32-10 under `key`: counter (T lo, 0x4D430002, T hi, 0) -> site =
   (w0 * ns) >> 32, u_acc = w1 / 2^32, u_mix = w2 / 2^32; counter
   (T lo, 0x4D430003, T hi, 0) -> four normals (Box-Muller). */
int so_local_sequential(const so_spec *sp, double *S, int nprop, const int *kind, const double *mag,
                        const double *cum, double kT, uint64_t key, uint64_t sweep0, double nsweeps, double acc[4])
{
    const long ns = nsites(sp);
    const long niters = (long)ceil(nsweeps * (double)ns);
    const uint32_t kk[2] = {(uint32_t)key, (uint32_t)(key >> 32)};
    const double inv = 1.0 / 4294967296.0;
    for (long t = 0; t < niters; t++) {
        const uint64_t T = sweep0 * (uint64_t)ns + (uint64_t)t;
        uint32_t ctr[4] = {(uint32_t)T, 0x4D430002u, (uint32_t)(T >> 32), 0u}, u[4];
        so_philox4x32(ctr, kk, u);
        const long site = (long)(((uint64_t)u[0] * (uint64_t)ns) >> 32);
        const double u_acc = (double)u[1] * inv, u_mix = (double)u[2] * inv;
        ctr[1] = 0x4D430003u;
        uint32_t v[4];
        so_philox4x32(ctr, kk, v);
        double n[4];
        box_muller_pair(v[0], v[1], &n[0], &n[1]);
        box_muller_pair(v[2], v[3], &n[2], &n[3]);
        if (local_trial(sp, S, site, nprop, kind, mag, cum, kT, u_mix, u_acc, n, acc)) return -1;
    }
    return 0;
}

/* LocalSampler's step! as the reference orders it (Samplers.jl:133-153):
   ceil(nsweeps * nsites) trials at uniformly random sites.  The draws come
   from splitmix64 seeded by `seed` (the reference draws from sys.rng). */
static uint64_t splitmix64(uint64_t *x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double splitmix_u01(uint64_t *x) { return (double)(splitmix64(x) >> 11) * (1.0 / 9007199254740992.0); }

int so_local_random_sites(const so_spec *sp, double *S, int nprop, const int *kind, const double *mag,
                          const double *cum, double kT, uint64_t seed, double nsweeps, double acc[4])
{
    const long ns = nsites(sp);
    const long niters = (long)ceil(nsweeps * (double)ns);
    uint64_t st = seed;
    for (long t = 0; t < niters; t++) {
        const long site = (long)(splitmix64(&st) % (uint64_t)ns);
        const double u_mix = splitmix_u01(&st);
        double n[4];
        box_muller_pair((uint32_t)splitmix64(&st), (uint32_t)splitmix64(&st), &n[0], &n[1]);
        box_muller_pair((uint32_t)splitmix64(&st), (uint32_t)splitmix64(&st), &n[2], &n[3]);
        const double u_acc = splitmix_u01(&st);
        if (local_trial(sp, S, site, nprop, kind, mag, cum, kT, u_mix, u_acc, n, acc)) return -1;
    }
    return 0;
}

/* local_energy_change for :SUN (Interactions.jl:311-380): the coherent state
   of `site` becomes Zn (|Zn|^2 = kappa); S holds the expected spins of Z.
   Zeeman and bilinear terms act on expected spins, the onsite term is
   Re(Z' Lambda Z), biquadratic terms act on expected quadrupoles.  Writes the
   new expected spin to Sn.  NAN on a self-interaction. */
static double local_delta_sun(const so_spec *sp, const sun_ops *o, const cplx *Z, const double *S, long site,
                              const cplx *Zn, double Sn[3])
{
    const int N = sp->N;
    const long nc = ncells(sp);
    const int *L = sp->L;
    const int i = (int)(site / nc);
    const long c = site % nc;
    const int c1 = (int)(c % L[0]), c2 = (int)((c / L[0]) % L[1]), c3 = (int)(c / ((long)L[0] * L[1]));
    const cplx *Z0 = Z + (long)N * site;
    expected_spin1(o, Zn, Sn);
    const double *S0 = S + 3 * site;
    double dS[3] = {Sn[0] - S0[0], Sn[1] - S0[1], Sn[2] - S0[2]};
    double dE = 0;
    const double *g = sp->g + 9 * i, *B = sp->field + 3 * i;
    for (int r = 0; r < 3; r++) dE += B[r] * (g[3 * r] * dS[0] + g[3 * r + 1] * dS[1] + g[3 * r + 2] * dS[2]);
    const double *Lam = sp->Lambda + (long)2 * N * N * i;
    cplx en = 0, eo = 0;
    for (int a = 0; a < N; a++)
        for (int b = 0; b < N; b++) {
            const cplx l = Lam[2 * (a * N + b)] + I * Lam[2 * (a * N + b) + 1];
            en += conj(Zn[a]) * l * Zn[b];
            eo += conj(Z0[a]) * l * Z0[b];
        }
    dE += creal(en) - creal(eo);
    double Qn[5], Qo[5];
    expected_quadrupole1(o, Zn, Qn);
    expected_quadrupole1(o, Z0, Qo);
    for (int b = 0; b < sp->nb; b++) {
        const int *bd = sp->bond + 6 * b;
        if (bd[0] != i) continue;
        long sj = site_index(L, wrap(c1 + bd[2], L[0]), wrap(c2 + bd[3], L[1]), wrap(c3 + bd[4], L[2]), bd[1]);
        if (sj == site) return NAN;
        const double *Sj = S + 3 * sj, *J = sp->bilin + 9 * b;
        for (int r = 0; r < 3; r++) dE += dS[r] * (J[3 * r] * Sj[0] + J[3 * r + 1] * Sj[1] + J[3 * r + 2] * Sj[2]);
        if (sp->has_biquad[b]) {
            double Qj[5];
            expected_quadrupole1(o, Z + (long)N * sj, Qj);
            const double *Bq = sp->biquad + 25 * b;
            for (int p = 0; p < 5; p++)
                for (int q = 0; q < 5; q++) dE += (Qn[p] - Qo[p]) * Bq[5 * p + q] * Qj[q];
        }
        /* general couplings, Interactions.jl:371-377 */
        int t0, t1;
        if (gen_terms(sp, b, &t0, &t1))
            for (int t = t0; t < t1; t++) {
                const double *A = sp->gen_A + (long)2 * N * N * t;
                dE += (expect_op(N, A, Zn) - expect_op(N, A, Z0)) *
                      expect_op(N, sp->gen_B + (long)2 * N * N * t, Z + (long)N * sj);
            }
    }
    if (sp->ewald_A) dE += ewald_delta(sp, S, site, dS);
    return dE;
}

double so_local_energy_change_sun(const so_spec *sp, const double *Zr, const double *S, long site,
                                  const double *Znr)
{
    sun_ops o;
    sun_ops_init(&o, sp->N);
    double Sn[3];
    return local_delta_sun(sp, &o, (const cplx *)Zr, S, site, (const cplx *)Znr, Sn);
}

/* flip_ket (Spin.jl:119-125): reverse(parity .* conj(Z)), parity (1, -1, 1, ...). */
static void flip_ket(int N, const cplx *Z, cplx *out)
{
    for (int k = 0; k < N; k++) {
        const int j = N - 1 - k;
        out[k] = (j % 2 ? -1.0 : 1.0) * conj(Z[j]);
    }
}

/* One :SUN Metropolis trial at `site`: proposal kind picked by u_mix, its 2N
   normals from Philox4x32-10 of counter ctr (second word advanced by 2 per
   block of four); component a of the complex draw randn(CVec{N}) (variance
   1/2 per part) is (n[2a] + i n[2a+1]) / sqrt(2).  Proposals: randspin
   (System.jl:448-452), flip (flip_ket), perturbed_spin Z + mag sqrt(kappa)
   randn (System.jl:461-468).  Accepted moves update Z, S (the expected spins
   of Z) and acc.  Returns -1 on a self-interaction. */
static int local_trial_sun(const so_spec *sp, const sun_ops *o, cplx *Z, double *S, long site, int nprop,
                           const int *kind, const double *mag, const double *cum, double kT, double u_mix,
                           double u_acc, const uint32_t ctr[4], const uint32_t kk[2], double acc[4])
{
    const int N = sp->N;
    int kp = nprop - 1;
    for (int p = 0; p < nprop - 1; p++)
        if (u_mix < cum[p]) {
            kp = p;
            break;
        }
    const int i = (int)(site / ncells(sp));
    const double kap = sp->kappa[i];
    const cplx *Z0 = Z + (long)N * site;
    cplx Zn[NMAX];
    if (kind[kp] == 1) {
        flip_ket(N, Z0, Zn);
    } else {
        double n[4 * ((2 * NMAX + 3) / 4)];
        for (int b = 0; b < (2 * N + 3) / 4; b++) {
            uint32_t cn[4] = {ctr[0], ctr[1] + 2u * (uint32_t)b, ctr[2], ctr[3]}, w[4];
            so_philox4x32(cn, kk, w);
            box_muller_pair(w[0], w[1], &n[4 * b], &n[4 * b + 1]);
            box_muller_pair(w[2], w[3], &n[4 * b + 2], &n[4 * b + 3]);
        }
        const int unif = kind[kp] == 0 || isinf(mag[kp]);
        for (int a = 0; a < N; a++) {
            const cplx r = (n[2 * a] + I * n[2 * a + 1]) / sqrt(2.0);
            Zn[a] = unif ? r : Z0[a] + mag[kp] * sqrt(kap) * r;
        }
        normalize_ket(N, Zn, kap);
    }
    double Sn[3];
    const double dE = local_delta_sun(sp, o, Z, S, site, Zn, Sn);
    if (isnan(dE)) return -1;
    const int accept = kT == 0 ? dE <= 0 : u_acc <= exp(-dE / kT);
    if (accept) {
        acc[0] += dE;
        for (int r = 0; r < 3; r++) {
            acc[1 + r] += Sn[r] - S[3 * site + r];
            S[3 * site + r] = Sn[r];
        }
        for (int a = 0; a < N; a++) Z[(long)N * site + a] = Zn[a];
    }
    return 0;
}

/* LocalSampler passes for :SUN, the order and draws of so_local_sweeps: the
   normals of site s in pass P come from the counters (s, 0x4D430000 + 2b, P),
   b = 0, 1, ..., the uniforms from (s, 0x4D430001, P). */
int so_local_sweeps_sun(const so_spec *sp, double *Zr, double *S, const int m[3], const int *color, int ncolors,
                        int nprop, const int *kind, const double *mag, const double *cum, double kT, uint64_t key,
                        uint64_t sweep0, double nsweeps, double acc[4])
{
    const int *L = sp->L;
    cplx *Z = (cplx *)Zr;
    sun_ops o;
    sun_ops_init(&o, sp->N);
    const uint32_t kk[2] = {(uint32_t)key, (uint32_t)(key >> 32)};
    const long npass = (long)ceil(2 * nsweeps);
    const double inv = 1.0 / 4294967296.0;
    for (long k = 0; k < npass; k++) {
        const uint64_t P = 2 * sweep0 + (uint64_t)k;
        const double prob = k == npass - 1 ? (2 * nsweeps - (double)(npass - 1)) / 2 : 0.5;
        for (int col = 0; col < ncolors; col++)
            for (int i = 0; i < sp->na; i++)
                for (int c3 = 0; c3 < L[2]; c3++)
                    for (int c2 = 0; c2 < L[1]; c2++)
                        for (int c1 = 0; c1 < L[0]; c1++) {
                            const int cls = c1 % m[0] + m[0] * (c2 % m[1] + m[1] * (c3 % m[2] + m[2] * i));
                            if (color[cls] != col) continue;
                            const long site = site_index(L, c1, c2, c3, i);
                            uint32_t ctr[4] = {(uint32_t)site, 0x4D430001u, (uint32_t)P, (uint32_t)(P >> 32)}, u[4];
                            so_philox4x32(ctr, kk, u);
                            if (!((double)u[2] * inv < prob)) continue;
                            ctr[1] = 0x4D430000u;
                            if (local_trial_sun(sp, &o, Z, S, site, nprop, kind, mag, cum, kT, (double)u[1] * inv,
                                                (double)u[0] * inv, ctr, kk, acc))
                                return -1;
                        }
    }
    return 0;
}

/* so_local_sequential for :SUN (the device's form with the Ewald term): the
   site, u_acc and u_mix of trial T from counter (T lo, 0x4D430002, T hi, 0),
   the normals from (T lo, 0x4D430003 + 2b, T hi, 0). */
int so_local_sequential_sun(const so_spec *sp, double *Zr, double *S, int nprop, const int *kind, const double *mag,
                            const double *cum, double kT, uint64_t key, uint64_t sweep0, double nsweeps,
                            double acc[4])
{
    cplx *Z = (cplx *)Zr;
    sun_ops o;
    sun_ops_init(&o, sp->N);
    const long ns = nsites(sp);
    const long niters = (long)ceil(nsweeps * (double)ns);
    const uint32_t kk[2] = {(uint32_t)key, (uint32_t)(key >> 32)};
    const double inv = 1.0 / 4294967296.0;
    for (long t = 0; t < niters; t++) {
        const uint64_t T = sweep0 * (uint64_t)ns + (uint64_t)t;
        uint32_t ctr[4] = {(uint32_t)T, 0x4D430002u, (uint32_t)(T >> 32), 0u}, u[4];
        so_philox4x32(ctr, kk, u);
        const long site = (long)(((uint64_t)u[0] * (uint64_t)ns) >> 32);
        ctr[1] = 0x4D430003u;
        if (local_trial_sun(sp, &o, Z, S, site, nprop, kind, mag, cum, kT, (double)u[2] * inv, (double)u[1] * inv, ctr,
                            kk, acc))
            return -1;
    }
    return 0;
}
