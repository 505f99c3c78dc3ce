// Long-range dipole-dipole interaction by Ewald summation, evaluated as an FFT
// convolution over Bravais sublattices.
//
// Reference semantics (Sunny.jl 0.9.2):
//   precompute_dipole_ewald_aux (q = 0)   src/System/Ewald.jl:61-152
//   Ewald constructor (FA = rfft(A))      src/System/Ewald.jl:2-22
//   accum_ewald_grad!                     src/System/Ewald.jl:185-209
//   ewald_energy                          src/System/Ewald.jl:155-181
//
// Pipeline per evaluation (one real field per (sublattice, component)):
//   k_ewald_fwd    mu = -g S, 3D r2c DFT of each field inside one workgroup's LDS
//   k_ewald_contract  F[phi]_k = M_k F[mu]_k with M_k = conj(F[A]_k), a
//                  (3na x 3na) complex block per half-spectrum wavevector k,
//                  stored column-major so each column read is coalesced
//   k_ewald_inv    3D c2r DFT of F[phi] (1/Ncell scaling) -> phi planes
// The local-field kernels then add dE/dS -= g' phi and the energy 1/2 mu.phi.
// Wavevectors: k1 in [0, L1/2] (half spectrum along the fastest cell axis,
// as rfft), k2 in [0, L2), k3 in [0, L3); flat index k1 + H1*(k2 + L2*k3).
#include "snh_internal.h"

namespace snh {

constexpr int kLmax = 32;  // longest supported axis of the FFT grid

template <typename T>
struct C2 {
    T re, im;
};



// 1D DFT of one line held in registers, written back to LDS with stride.
// sign = -1 forward, +1 backward.  tw[m] = exp(sign * 2 pi i m / L).
template <typename T>
__device__ inline void dft_line_c2c(C2<T>* buf, long long base, long long stride, int L, const C2<T>* tw)
{
    C2<T> in[kLmax];
#pragma unroll
    for (int x = 0; x < kLmax; x++)
        if (x < L) in[x] = buf[base + x * stride];
    for (int k = 0; k < L; k++) {
        T ar = 0, ai = 0;
        int m = 0;
#pragma unroll
        for (int x = 0; x < kLmax; x++) {
            if (x < L) {
                C2<T> w = tw[m];
                ar += in[x].re * w.re - in[x].im * w.im;
                ai += in[x].re * w.im + in[x].im * w.re;
                m += k;
                if (m >= L) m -= L;
            }
        }
        buf[base + k * stride] = {ar, ai};
    }
}

template <typename T>
__device__ inline void make_twiddles(C2<T>* tw, int L, int sign)
{
    for (int m = threadIdx.x; m < L; m += blockDim.x) {
        double s, c;
        sincospi(2.0 * m / L, &s, &c);
        tw[m] = {(T)c, (T)(sign * s)};
    }
}

// Forward r2c 3D DFT of the field values produced by `load(c1, c2, c3)`,
// leaving the half spectrum in LDS `spec` (layout k1 + H1*(k2 + L2*k3)).
template <typename T, typename Loader>
__device__ inline void fwd3d(const EwaldGeom& G, C2<T>* spec, C2<T>* tw1, C2<T>* tw2, C2<T>* tw3, Loader load)
{
    make_twiddles(tw1, G.L1, -1);
    make_twiddles(tw2, G.L2, -1);
    make_twiddles(tw3, G.L3, -1);
    __syncthreads();
    // axis 1, real -> half complex
    for (int line = threadIdx.x; line < G.L2 * G.L3; line += blockDim.x) {
        int c2 = line % G.L2, c3 = line / G.L2;
        T in[kLmax];
#pragma unroll
        for (int x = 0; x < kLmax; x++)
            if (x < G.L1) in[x] = load(x, c2, c3);
        for (int k = 0; k < G.H1; k++) {
            T ar = 0, ai = 0;
            int m = 0;
#pragma unroll
            for (int x = 0; x < kLmax; x++) {
                if (x < G.L1) {
                    ar += in[x] * tw1[m].re;
                    ai += in[x] * tw1[m].im;
                    m += k;
                    if (m >= G.L1) m -= G.L1;
                }
            }
            spec[k + (long long)G.H1 * (c2 + (long long)G.L2 * c3)] = {ar, ai};
        }
    }
    __syncthreads();
    for (int line = threadIdx.x; line < G.H1 * G.L3; line += blockDim.x) {
        int k1 = line % G.H1, c3 = line / G.H1;
        dft_line_c2c(spec, k1 + (long long)G.H1 * G.L2 * c3, G.H1, G.L2, tw2);
    }
    __syncthreads();
    for (int line = threadIdx.x; line < G.H1 * G.L2; line += blockDim.x)
        dft_line_c2c(spec, line, (long long)G.H1 * G.L2, G.L3, tw3);
    __syncthreads();
}

// Dynamic LDS: spectrum (nk complex) followed by three twiddle tables.
template <typename T>
__device__ inline C2<T>* lds_spec()
{
    extern __shared__ __align__(16) unsigned char snh_lds[];
    return reinterpret_cast<C2<T>*>(snh_lds);
}

// Padded state index of interior cell (c1, c2, c3) of sublattice j.
__device__ __forceinline__ uint32_t ewald_site(const EwaldGeom& G, uint32_t off, int j, int c1, int c2, int c3)
{
    return off + (uint32_t)((c1 + G.w1) + G.P1 * ((c2 + G.w2) + G.P2 * (c3 + G.w3))) + G.pcell * (uint32_t)j;
}

// Forward transform of mu = -g S.  grid = (3 na, R); field f = 3 j + beta.
template <typename T>
__global__ __launch_bounds__(256) void k_ewald_fwd(EwaldGeom G, const ReplicaInfo* __restrict__ reps,
                                                   uint32_t total, const T* __restrict__ gmat,
                                                   const T* __restrict__ S, C2<T>* __restrict__ Fmu)
{
    const int f = blockIdx.x, r = blockIdx.y;
    const int j = f / 3, beta = f % 3;
    C2<T>* spec = lds_spec<T>();
    C2<T>* tw1 = spec + G.nk;
    C2<T>* tw2 = tw1 + kLmax;
    C2<T>* tw3 = tw2 + kLmax;
    const uint32_t off = reps[r].off;
    const T g0 = gmat[9 * j + 3 * beta], g1 = gmat[9 * j + 3 * beta + 1], g2 = gmat[9 * j + 3 * beta + 2];
    fwd3d<T>(G, spec, tw1, tw2, tw3, [&](int c1, int c2, int c3) -> T {
        const uint32_t g = ewald_site(G, off, j, c1, c2, c3);
        return -(g0 * ldg(S, g) + g1 * ldg(S + total, g) + g2 * ldg(S + 2 * (size_t)total, g));
    });
    C2<T>* out = Fmu + ((long long)r * 3 * G.na + f) * G.nk;
    for (long long k = threadIdx.x; k < G.nk; k += blockDim.x) out[k] = spec[k];
}

// Forward transform of the real-space tensor A (FP64), written as the
// conjugated, column-major contraction blocks M_k[col = 3j+beta][row = 3i+alpha].
// grid = 9 na^2: field = ((j*na + i)*3 + alpha)*3 + beta.
template <typename T>
__global__ __launch_bounds__(256) void k_ewald_fwdA(EwaldGeom G, const double* __restrict__ A,
                                                    C2<T>* __restrict__ M)
{
    const int f = blockIdx.x;
    const int beta = f % 3, alpha = (f / 3) % 3, i = (f / 9) % G.na, j = f / (9 * G.na);
    C2<double>* spec = lds_spec<double>();
    C2<double>* tw1 = spec + G.nk;
    C2<double>* tw2 = tw1 + kLmax;
    C2<double>* tw3 = tw2 + kLmax;
    const double* Aij = A + ((long long)j * G.na + i) * G.ncell * 9;
    fwd3d<double>(G, spec, tw1, tw2, tw3, [&](int c1, int c2, int c3) -> double {
        long long c = c1 + (long long)G.L1 * (c2 + (long long)G.L2 * c3);
        return Aij[c * 9 + 3 * alpha + beta];
    });
    const int n3 = 3 * G.na;
    const long long col = 3 * j + beta, row = 3 * i + alpha;
    for (long long k = threadIdx.x; k < G.nk; k += blockDim.x)
        M[(k * n3 + col) * n3 + row] = {(T)spec[k].re, (T)(-spec[k].im)};
}

// F[phi]_k = M_k F[mu]_k for every replica.  One workgroup per k.
template <typename T>
__global__ __launch_bounds__(256) void k_ewald_contract(EwaldGeom G, int R, const C2<T>* __restrict__ M,
                                                        const C2<T>* __restrict__ Fmu, C2<T>* __restrict__ Fphi)
{
    const long long k = blockIdx.x;
    const int n3 = 3 * G.na;
    extern __shared__ __align__(16) unsigned char snh_lds[];
    C2<T>* v = reinterpret_cast<C2<T>*>(snh_lds);  // [R][n3]
    for (int t = threadIdx.x; t < R * n3; t += blockDim.x) {
        int r = t / n3, col = t % n3;
        v[t] = Fmu[((long long)r * n3 + col) * G.nk + k];
    }
    __syncthreads();
    const C2<T>* Mk = M + k * n3 * n3;
    for (int t = threadIdx.x; t < R * n3; t += blockDim.x) {
        int r = t / n3, row = t % n3;
        const C2<T>* vr = v + r * n3;
        T ar = 0, ai = 0;
        for (int col = 0; col < n3; col++) {
            C2<T> m = Mk[col * n3 + row];
            C2<T> x = vr[col];
            ar += m.re * x.re - m.im * x.im;
            ai += m.re * x.im + m.im * x.re;
        }
        Fphi[((long long)r * n3 + row) * G.nk + k] = {ar, ai};
    }
}

// Inverse: c2r 3D DFT of F[phi] scaled by 1/Ncell -> phi plane alpha of
// sublattice i.  grid = (3 na, R); field = 3 i + alpha.
template <typename T>
__global__ __launch_bounds__(256) void k_ewald_inv(EwaldGeom G, const ReplicaInfo* __restrict__ reps, uint32_t total,
                                                   const C2<T>* __restrict__ Fphi, T* __restrict__ phi)
{
    const int f = blockIdx.x, r = blockIdx.y;
    const int i = f / 3, alpha = f % 3;
    C2<T>* spec = lds_spec<T>();
    C2<T>* tw1 = spec + G.nk;
    C2<T>* tw2 = tw1 + kLmax;
    C2<T>* tw3 = tw2 + kLmax;
    make_twiddles(tw1, G.L1, +1);
    make_twiddles(tw2, G.L2, +1);
    make_twiddles(tw3, G.L3, +1);
    const C2<T>* in = Fphi + ((long long)r * 3 * G.na + f) * G.nk;
    for (long long k = threadIdx.x; k < G.nk; k += blockDim.x) spec[k] = in[k];
    __syncthreads();
    for (int line = threadIdx.x; line < G.H1 * G.L2; line += blockDim.x)
        dft_line_c2c(spec, line, (long long)G.H1 * G.L2, G.L3, tw3);
    __syncthreads();
    for (int line = threadIdx.x; line < G.H1 * G.L3; line += blockDim.x) {
        int k1 = line % G.H1, c3 = line / G.H1;
        dft_line_c2c(spec, k1 + (long long)G.H1 * G.L2 * c3, G.H1, G.L2, tw2);
    }
    __syncthreads();
    // axis 1, half complex -> real, using Hermitian symmetry X[L1-k] = conj(X[k]).
    const T scale = (T)(1.0 / (double)G.ncell);
    T* const out = phi + (size_t)alpha * total;
    const uint32_t off = reps[r].off;
    const bool even = (G.L1 % 2) == 0;
    for (int line = threadIdx.x; line < G.L2 * G.L3; line += blockDim.x) {
        int c2 = line % G.L2, c3 = line / G.L2;
        const C2<T>* X = spec + (long long)G.H1 * (c2 + (long long)G.L2 * c3);
        for (int x = 0; x < G.L1; x++) {
            T acc = X[0].re;
            int m = x;  // twiddle index k*x mod L1
            int last = even ? G.H1 - 1 : G.H1;
            for (int k = 1; k < last; k++) {
                C2<T> w = tw1[m];
                acc += 2 * (X[k].re * w.re - X[k].im * w.im);
                m += x;
                if (m >= G.L1) m -= G.L1;
            }
            if (even && G.H1 > 1) acc += X[G.H1 - 1].re * ((x & 1) ? (T)-1 : (T)1);
            stg(out, ewald_site(G, off, i, x, c2, c3), acc * scale);
        }
    }
}

// Real-space Ewald tensor A[off, i, j] (FP64), scaled by mu0_muB2.
// Thread per (cell offset, i); grid.y = j.  Layout A[j][i][cell][3][3].


__global__ __launch_bounds__(256) void k_ewald_A(EwaldAParams P, const double* __restrict__ pos, double* __restrict__ A)
{
    const long long ncell = (long long)P.L1 * P.L2 * P.L3;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (t >= ncell * P.na) return;
    const int i = (int)(t / ncell);
    const long long cell = t % ncell;
    const int c1 = (int)(cell % P.L1), c2 = (int)((cell / P.L1) % P.L2), c3 = (int)(cell / ((long long)P.L1 * P.L2));
    double fr[3] = {c1 + pos[3 * j] - pos[3 * i], c2 + pos[3 * j + 1] - pos[3 * i + 1], c3 + pos[3 * j + 2] - pos[3 * i + 2]};
    double dr[3];
    for (int r = 0; r < 3; r++) dr[r] = P.lv[3 * r] * fr[0] + P.lv[3 * r + 1] * fr[1] + P.lv[3 * r + 2] * fr[2];
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const double s2 = P.sigma * P.sigma, s3 = s2 * P.sigma;
    const double inv4pi = 1.0 / (4.0 * M_PI);
    const double rmax2 = P.rmax * P.rmax, kmax2 = P.kmax * P.kmax;
    // real-space images
    for (int n1 = -P.nmax[0]; n1 <= P.nmax[0]; n1++)
        for (int n2 = -P.nmax[1]; n2 <= P.nmax[1]; n2++)
            for (int n3 = -P.nmax[2]; n3 <= P.nmax[2]; n3++) {
                double rv[3];
                for (int r = 0; r < 3; r++) rv[r] = dr[r] + P.SL[3 * r] * n1 + P.SL[3 * r + 1] * n2 + P.SL[3 * r + 2] * n3;
                double r2 = rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
                if (0 < r2 && r2 <= rmax2) {
                    double rr = sqrt(r2), r3 = r2 * rr;
                    double erfc0 = erfc(rr / (M_SQRT2 * P.sigma));
                    double gauss0 = sqrt(2 / M_PI) * (rr / P.sigma) * exp(-r2 / (2 * s2));
                    double a = (erfc0 + gauss0) / r3;
                    double b = 3 * (erfc0 + (1 + r2 / (3 * s2)) * gauss0) / (r3 * r2);
                    for (int p = 0; p < 3; p++)
                        for (int q = 0; q < 3; q++) acc[3 * p + q] += inv4pi * ((p == q ? a : 0.0) - b * rv[p] * rv[q]);
                }
            }
    // reciprocal-space images
    for (int m1 = -P.mmax[0]; m1 <= P.mmax[0]; m1++)
        for (int m2 = -P.mmax[1]; m2 <= P.mmax[1]; m2++)
            for (int m3 = -P.mmax[2]; m3 <= P.mmax[2]; m3++) {
                double kv[3];
                for (int r = 0; r < 3; r++) kv[r] = P.RL[3 * r] * m1 + P.RL[3 * r + 1] * m2 + P.RL[3 * r + 2] * m3;
                double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
                if (k2 <= 1e-16) {
                    for (int p = 0; p < 9; p++) acc[p] += P.demag[p] / P.V;
                } else if (k2 <= kmax2) {
                    double w = cos(kv[0] * dr[0] + kv[1] * dr[1] + kv[2] * dr[2]) / P.V * exp(-s2 * k2 / 2) / k2;
                    for (int p = 0; p < 3; p++)
                        for (int q = 0; q < 3; q++) acc[3 * p + q] += w * kv[p] * kv[q];
                }
            }
    if (dr[0] == 0 && dr[1] == 0 && dr[2] == 0) {
        double self = -1 / (3 * pow(2 * M_PI, 1.5) * s3);
        acc[0] += self;
        acc[4] += self;
        acc[8] += self;
    }
    double* out = A + (((long long)j * P.na + i) * ncell + cell) * 9;
    for (int p = 0; p < 9; p++) out[p] = acc[p] * P.mu0_muB2;
}

// ---------------------------------------------------------------------------
// Host-side launchers

size_t ewald_fft_lds_bytes(long long nk, size_t elem)
{
    return (size_t)nk * elem + 3 * kLmax * elem;
}

void launch_ewald_A(const EwaldAParams& P, const double* pos, double* A, hipStream_t st)
{
    long long n = (long long)P.L1 * P.L2 * P.L3 * P.na;
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)P.na);
    hipLaunchKernelGGL(k_ewald_A, grid, dim3(256), 0, st, P, pos, A);
}

template <typename T>
void launch_ewald_fwdA(const EwaldGeom& G, const double* A, void* M, hipStream_t st)
{
    size_t lds = ewald_fft_lds_bytes(G.nk, sizeof(C2<double>));
    hipLaunchKernelGGL(k_ewald_fwdA<T>, dim3(9 * G.na * G.na), dim3(256), lds, st, G, A,
                       reinterpret_cast<C2<T>*>(M));
}

template <typename T>
void launch_ewald_field(const EwaldGeom& G, int R, const ReplicaInfo* reps, uint32_t total, const T* gmat,
                        const T* S, const void* M, void* Fmu, void* Fphi, T* phi, hipStream_t st)
{
    size_t lds = ewald_fft_lds_bytes(G.nk, sizeof(C2<T>));
    hipLaunchKernelGGL(k_ewald_fwd<T>, dim3(3 * G.na, R), dim3(256), lds, st, G, reps, total, gmat, S,
                       reinterpret_cast<C2<T>*>(Fmu));
    size_t lds2 = (size_t)R * 3 * G.na * sizeof(C2<T>);
    hipLaunchKernelGGL(k_ewald_contract<T>, dim3((unsigned)G.nk), dim3(256), lds2, st, G, R,
                       reinterpret_cast<const C2<T>*>(M), reinterpret_cast<const C2<T>*>(Fmu),
                       reinterpret_cast<C2<T>*>(Fphi));
    hipLaunchKernelGGL(k_ewald_inv<T>, dim3(3 * G.na, R), dim3(256), lds, st, G, reps, total,
                       reinterpret_cast<const C2<T>*>(Fphi), phi);
}

template void launch_ewald_fwdA<float>(const EwaldGeom&, const double*, void*, hipStream_t);
template void launch_ewald_fwdA<double>(const EwaldGeom&, const double*, void*, hipStream_t);
template void launch_ewald_field<float>(const EwaldGeom&, int, const ReplicaInfo*, uint32_t, const float*,
                                        const float*, const void*, void*, void*, float*, hipStream_t);
template void launch_ewald_field<double>(const EwaldGeom&, int, const ReplicaInfo*, uint32_t, const double*,
                                         const double*, const void*, void*, void*, double*, hipStream_t);

}  // namespace snh
