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
//   k_ewald_fwd    mu = -g S staged in LDS, 3D r2c FFT of each field inside
//                  one workgroup (radix-2 in registers per line for
//                  power-of-two axes, direct DFT otherwise)
//   k_ewald_contract / k_ewald_contract_mfma
//                  F[phi]_k = M_k F[mu]_k with M_k = conj(F[A]_k), a
//                  (3na x 3na) complex block per half-spectrum wavevector k,
//                  stored column-major so each column read is coalesced.
//                  Few replicas: one matrix-vector product per k (streams M,
//                  HBM-bound).  Many replicas: a complex GEMM per k on the
//                  FP32 matrix cores (v_mfma_f32_16x16x4_f32), every M_k read
//                  once for up to 32 replicas.
//   k_ewald_inv    3D c2r FFT of F[phi] (1/Ncell scaling) -> phi planes
// The local-field kernels then add dE/dS -= g' phi and the energy 1/2 mu.phi.
// Wavevectors: k1 in [0, L1/2] (half spectrum along the fastest cell axis,
// as rfft), k2 in [0, L2), k3 in [0, L3); flat index k1 + H1*(k2 + L2*k3).
// Spectra of all fields live in HBM in blocks of kSpecKB wavevectors,
// [k / kSpecKB][replica][field][k % kSpecKB]: each transform workgroup reads
// and writes whole 128-byte runs, and the contraction of a block for a group
// of replicas reads one compact range.
//
// Grids whose one-field transform does not fit one workgroup's LDS (or with
// an axis longer than kLmax) take the library path (G.plain): the fields are
// gathered into [replica][field][cell] real arrays, transformed by batched
// 3D hipFFT plans into plain spectra [replica][field][k], contracted by the
// same kernels and transformed back; the reference transforms with FFTW
// (Ewald.jl:2-22, 185-209) and has no size bound either.
#include "snh_fft.h"

#include <hipfft/hipfft.h>

#include <cstdlib>
#include <map>
#include <mutex>

namespace snh {

constexpr int kFftThreads = 256;
constexpr int kSpecKB = 16;  // wavevectors per spectrum block (128-byte runs in FP32)

// ---------------------------------------------------------------------------
// One workgroup transforms one field.  Dynamic LDS: spectrum [nk] complex,
// three twiddle tables [kLmax], the real field [L2 L3][L1 + 1] (row padded
// by one word against bank conflicts).

template <typename T>
struct FftLds {
    C2<T>* spec;
    C2<T>* tw1;
    C2<T>* tw2;
    C2<T>* tw3;
    T* fld;
    int row;  // L1 + 1
};

template <typename T>
__device__ __forceinline__ FftLds<T> fft_lds(const EwaldGeom& G)
{
    extern __shared__ __align__(16) unsigned char snh_lds[];
    FftLds<T> F;
    F.spec = reinterpret_cast<C2<T>*>(snh_lds);
    F.tw1 = F.spec + G.nk;
    F.tw2 = F.tw1 + kLmax;
    F.tw3 = F.tw2 + kLmax;
    F.fld = reinterpret_cast<T*>(F.tw3 + kLmax);
    F.row = G.L1 + 1;
    return F;
}

// Forward r2c 3D FFT of F.fld (filled by the caller), leaving the half
// spectrum in F.spec.  Twiddles must be ready (sign -1).
template <typename T>
__device__ __forceinline__ void fwd3d(const EwaldGeom& G, const FftLds<T>& F)
{
    for (int line = threadIdx.x; line < G.L2 * G.L3; line += blockDim.x)
        line_r2c<T>(F.fld + line * F.row, F.spec + G.H1 * line, G.L1, F.tw1);
    __syncthreads();
    for (int line = threadIdx.x; line < G.H1 * G.L3; line += blockDim.x) {
        const int k1 = line % G.H1, c3 = line / G.H1;
        line_c2c<T>(F.spec, k1 + G.H1 * G.L2 * c3, G.H1, G.L2, F.tw2);
    }
    __syncthreads();
    for (int line = threadIdx.x; line < G.H1 * G.L2; line += blockDim.x)
        line_c2c<T>(F.spec, line, G.H1 * G.L2, G.L3, F.tw3);
    __syncthreads();
}

// Backward c2r 3D FFT of F.spec into F.fld (unscaled).  Twiddles sign +1.
template <typename T>
__device__ __forceinline__ void inv3d(const EwaldGeom& G, const FftLds<T>& F)
{
    for (int line = threadIdx.x; line < G.H1 * G.L2; line += blockDim.x)
        line_c2c<T>(F.spec, line, G.H1 * G.L2, G.L3, F.tw3);
    __syncthreads();
    for (int line = threadIdx.x; line < G.H1 * G.L3; line += blockDim.x) {
        const int k1 = line % G.H1, c3 = line / G.H1;
        line_c2c<T>(F.spec, k1 + G.H1 * G.L2 * c3, G.H1, G.L2, F.tw2);
    }
    __syncthreads();
    for (int line = threadIdx.x; line < G.L2 * G.L3; line += blockDim.x)
        line_c2r<T>(F.spec + G.H1 * line, F.fld + line * F.row, G.L1, F.tw1);
    __syncthreads();
}

// Position of (replica r, field f, wavevector k) in the blocked spectra.
__device__ __forceinline__ long long spec_index(const EwaldGeom& G, int R, int r, int f, long long k)
{
    if (G.plain) return ((long long)r * (3 * G.na) + f) * G.nk + k;
    return (((k / kSpecKB) * R + r) * (3 * G.na) + f) * kSpecKB + (k % kSpecKB);
}

// Padded state index of interior cell (c1, c2, c3) of sublattice j.
__device__ __forceinline__ uint32_t ewald_site(const EwaldGeom& G, uint32_t off, int j, int c1, int c2, int c3)
{
    return off + (uint32_t)((c1 + G.w1) + G.P1 * ((c2 + G.w2) + G.P2 * (c3 + G.w3))) + G.pcell * (uint32_t)j;
}

// Forward transform of mu = -g S.  grid = (3 na, R); field f = 3 j + beta.
template <typename T>
__global__ __launch_bounds__(kFftThreads) void k_ewald_fwd(EwaldGeom G, int R, const ReplicaInfo* __restrict__ reps,
                                                           uint32_t total, const T* __restrict__ gmat,
                                                           const T* __restrict__ S, C2<T>* __restrict__ Fmu)
{
    const int f = blockIdx.x, r = blockIdx.y;
    const int j = f / 3, beta = f % 3;
    const FftLds<T> F = fft_lds<T>(G);
    const uint32_t off = reps[r].off;
    const T g0 = gmat[9 * j + 3 * beta], g1 = gmat[9 * j + 3 * beta + 1], g2 = gmat[9 * j + 3 * beta + 2];
    // loads in flight per thread: a 16^3 grid in one round
    constexpr int kBatch = sizeof(T) == 4 ? 16 : 8;
    for (int c0 = 0; c0 < G.ncell; c0 += kBatch * kFftThreads) {
        T v[kBatch];
        // unconditional loads (a cell past the end reads cell 0 and is
        // dropped), so that all of them are in flight before the first wait
#pragma unroll
        for (int q = 0; q < kBatch; q++) {
            const int c = c0 + q * kFftThreads + threadIdx.x;
            const int cc = c < G.ncell ? c : 0;
            const int c1 = cc % G.L1, l = cc / G.L1;
            const uint32_t g = ewald_site(G, off, j, c1, l % G.L2, l / G.L2);
            const T m = -(g0 * ldg(S, g) + g1 * ldg(S + total, g) + g2 * ldg(S + 2 * (size_t)total, g));
            v[q] = c < G.ncell ? m : (T)0;
        }
        if (c0 == 0) {  // while the first loads are in flight
            make_twiddles(F.tw1, G.L1, -1);
            make_twiddles(F.tw2, G.L2, -1);
            make_twiddles(F.tw3, G.L3, -1);
        }
#pragma unroll
        for (int q = 0; q < kBatch; q++) {
            const int c = c0 + q * kFftThreads + threadIdx.x;
            if (c < G.ncell) F.fld[(c / G.L1) * F.row + c % G.L1] = v[q];
        }
    }
    __syncthreads();
    fwd3d<T>(G, F);
    for (int k = threadIdx.x; k < G.nk; k += blockDim.x) Fmu[spec_index(G, R, r, f, k)] = F.spec[k];
}

// Forward transform of the real-space tensor A (FP64), written as the
// conjugated, column-major contraction blocks M_k[col = 3j+beta][row = 3i+alpha].
// grid = 9 na^2: field = ((j*na + i)*3 + alpha)*3 + beta.
template <typename T>
__global__ __launch_bounds__(kFftThreads) void k_ewald_fwdA(EwaldGeom G, const double* __restrict__ A,
                                                            C2<T>* __restrict__ M)
{
    const int f = blockIdx.x;
    const int beta = f % 3, alpha = (f / 3) % 3, i = (f / 9) % G.na, j = f / (9 * G.na);
    const FftLds<double> F = fft_lds<double>(G);
    make_twiddles(F.tw1, G.L1, -1);
    make_twiddles(F.tw2, G.L2, -1);
    make_twiddles(F.tw3, G.L3, -1);
    const double* Aij = A + ((long long)j * G.na + i) * G.ncell * 9;
    for (int c = threadIdx.x; c < G.ncell; c += blockDim.x)
        F.fld[(c / G.L1) * F.row + c % G.L1] = Aij[(long long)c * 9 + 3 * alpha + beta];
    __syncthreads();
    fwd3d<double>(G, F);
    const int n3 = 3 * G.na;
    const long long col = 3 * j + beta, row = 3 * i + alpha;
    for (int k = threadIdx.x; k < G.nk; k += blockDim.x)
        M[(k * n3 + col) * n3 + row] = {(T)F.spec[k].re, (T)(-F.spec[k].im)};
}

// F[phi]_k = M_k F[mu]_k for every replica; one workgroup per k (few replicas).
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
        v[t] = Fmu[spec_index(G, R, r, col, k)];
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
        Fphi[spec_index(G, R, r, row, k)] = {ar, ai};
    }
}

// The same contraction on the matrix cores for many replicas: a workgroup
// takes kMfmaKB wavevectors and up to 16 NC replicas, with one wave per
// (wavevector, tile of 16 output fields): wave w computes rows
// 16 (w / kMfmaKB) .. +15 of Y = M_k X for k = k0 + w % kMfmaKB as real
// 16x16x4 MFMA tiles,
//   Yr = Mr Xr - Mi Xi,  Yi = Mi Xr + Mr Xi,
// output rows and the reduction (input fields) padded to 16 with zeros.
// X and Y pass through LDS.  Operand layouts of v_mfma_f32_16x16x4_f32
// (wave of 64, lane l): A[m][kk]: m = l % 16, kk = l / 16; B[kk][n]:
// kk = l / 16, n = l % 16; D[m][n] (4 registers v): m = 4 (l / 16) + v,
// n = l % 16.
constexpr int kMfmaKB = 4;  // wavevectors per workgroup
constexpr int kMfmaMaxRowTiles = 4;  // 3 na <= 64
static_assert(kSpecKB % kMfmaKB == 0, "a contraction block lies inside one spectrum block");

// Workgroup -> block of wavevectors such that the kSpecKB / kMfmaKB blocks
// sharing 128-byte spectrum runs land on the same XCD (workgroups are
// dispatched to the 8 XCDs round robin) one after another, and meet in its L2.
__device__ __forceinline__ long long xcd_block(long long b, long long nb)
{
    const long long x = b % 8, i = b / 8, lo = nb / 8, rem = nb % 8;
    return x * lo + (x < rem ? x : rem) + i;
}

template <int NRT, int NC>
__global__ __launch_bounds__(256 * NRT) void k_ewald_contract_mfma(EwaldGeom G, int R,
                                                                   const C2<float>* __restrict__ M,
                                                                   const C2<float>* __restrict__ Fmu,
                                                                   C2<float>* __restrict__ Fphi)
{
    constexpr int NT = 256 * NRT;  // threads: kMfmaKB x NRT waves
    constexpr int RC = 16 * NC;    // replicas per workgroup
    constexpr int N3R = 16 * NRT;  // rows padded to the tile
    constexpr int KS = 4 * NRT;    // reduction steps of 4 (covers the padded fields)
    constexpr int XR = RC + 1;     // LDS row of one (field, k), padded against bank conflicts
    const int n3 = 3 * G.na;
    const long long k0 = xcd_block(blockIdx.x, gridDim.x) * kMfmaKB;
    const int r0 = blockIdx.y * RC;
    extern __shared__ __align__(16) unsigned char snh_lds[];
    C2<float>* X = reinterpret_cast<C2<float>*>(snh_lds);  // [N3R][kMfmaKB][XR]

    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int kk_w = wave % kMfmaKB, tile = wave / kMfmaKB;
    const long long k = k0 + kk_w;
    const int lm = lane % 16, lk = lane / 16;
    const int row = 16 * tile + lm;
    // this wave's operand A: rows of its tile of M_k, all loads in flight first
    C2<float> a[KS];
    {
        const C2<float>* Mk = M + (k < G.nk ? k : 0) * n3 * n3;
#pragma unroll
        for (int s = 0; s < KS; s++) {
            const int col = 4 * s + lk;
            a[s] = (k < G.nk && row < n3 && col < n3) ? Mk[col * n3 + row] : C2<float>{0.f, 0.f};
        }
    }
    // X block: 4 NC loads per thread, all issued before the first LDS write
    constexpr int PER = kMfmaKB * N3R * RC / NT;
    static_assert(PER * NT == kMfmaKB * N3R * RC, "whole loads per thread");
    C2<float> v[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int t = q * NT + threadIdx.x;
        const int kk = t % kMfmaKB, f = (t / kMfmaKB) % N3R, rr = t / (kMfmaKB * N3R);
        const int r = r0 + rr;
        v[q] = (f < n3 && k0 + kk < G.nk && r < R) ? Fmu[spec_index(G, R, r, f, k0 + kk)] : C2<float>{0.f, 0.f};
    }
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int t = q * NT + threadIdx.x;
        const int kk = t % kMfmaKB, f = (t / kMfmaKB) % N3R, rr = t / (kMfmaKB * N3R);
        X[(f * kMfmaKB + kk) * XR + rr] = v[q];
    }
    __syncthreads();

    typedef float f4 __attribute__((ext_vector_type(4)));
    f4 yr[NC], yi[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        yr[c] = f4{0.f, 0.f, 0.f, 0.f};
        yi[c] = f4{0.f, 0.f, 0.f, 0.f};
    }
    const C2<float>* Xk = X + kk_w * XR;
#pragma unroll
    for (int s = 0; s < KS; s++) {
        const int col = 4 * s + lk;
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const C2<float> b = Xk[col * kMfmaKB * XR + 16 * c + lm];
            yr[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s].re, b.re, yr[c], 0, 0, 0);
            yi[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s].im, b.re, yi[c], 0, 0, 0);
            yr[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s].im, -b.im, yr[c], 0, 0, 0);
            yi[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s].re, b.im, yi[c], 0, 0, 0);
        }
    }
    __syncthreads();  // every wave is done reading X: reuse it for Y
    C2<float>* Yk = X + kk_w * XR;
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
        for (int v4 = 0; v4 < 4; v4++)
            Yk[(16 * tile + 4 * lk + v4) * kMfmaKB * XR + 16 * c + lm] = {yr[c][v4], yi[c][v4]};
    __syncthreads();
    const int nst = kMfmaKB * n3 * RC;
    for (int t = threadIdx.x; t < nst; t += NT) {
        const int kk = t % kMfmaKB, f = (t / kMfmaKB) % n3, rr = t / (kMfmaKB * n3);
        const int r = r0 + rr;
        if (k0 + kk < G.nk && r < R) Fphi[spec_index(G, R, r, f, k0 + kk)] = X[(f * kMfmaKB + kk) * XR + rr];
    }
}

// Inverse: c2r 3D FFT of F[phi] scaled by 1/Ncell -> phi plane alpha of
// sublattice i (stored as O: T, or FP64 for the Metropolis potential).
// grid = (3 na, R); field = 3 i + alpha.
template <typename T, typename O>
__global__ __launch_bounds__(kFftThreads) void k_ewald_inv(EwaldGeom G, int R, const ReplicaInfo* __restrict__ reps,
                                                           uint32_t total, const C2<T>* __restrict__ Fphi,
                                                           O* __restrict__ phi)
{
    const int f = blockIdx.x, r = blockIdx.y;
    const int i = f / 3, alpha = f % 3;
    const FftLds<T> F = fft_lds<T>(G);
    // loads in flight per thread: the half spectrum of a 16^3 grid in one round
    constexpr int kBatch = sizeof(T) == 4 ? 12 : 6;
    for (int k0 = 0; k0 < G.nk; k0 += kBatch * kFftThreads) {
        C2<T> v[kBatch];
        // unconditional loads (past the end: wavevector 0, dropped below)
#pragma unroll
        for (int q = 0; q < kBatch; q++) {
            const int k = k0 + q * kFftThreads + threadIdx.x;
            v[q] = Fphi[spec_index(G, R, r, f, k < G.nk ? k : 0)];
        }
        if (k0 == 0) {  // while the first loads are in flight
            make_twiddles(F.tw1, G.L1, +1);
            make_twiddles(F.tw2, G.L2, +1);
            make_twiddles(F.tw3, G.L3, +1);
        }
#pragma unroll
        for (int q = 0; q < kBatch; q++) {
            const int k = k0 + q * kFftThreads + threadIdx.x;
            if (k < G.nk) F.spec[k] = v[q];
        }
    }
    __syncthreads();
    inv3d<T>(G, F);
    const T scale = (T)(1.0 / (double)G.ncell);
    O* const out = phi + (size_t)alpha * total;
    const uint32_t off = reps[r].off;
    for (int c = threadIdx.x; c < G.ncell; c += blockDim.x) {
        const int c1 = c % G.L1, l = c / G.L1;
        const T v = F.fld[l * F.row + c1] * scale;
        stg(out, ewald_site(G, off, i, c1, l % G.L2, l / G.L2), (O)v);
    }
}

// ---------------------------------------------------------------------------
// Slab path for few replicas.  The per-field transforms above run one
// workgroup per field (3 na of them per replica), which leaves most of the
// chip idle for one replica.  Here the transform is split by axis:
//   k_ewald_fwd_slab    one workgroup per (sublattice, plane c3): mu = -g S of
//                       the plane's cells for the three components, 2D r2c
//                       along axes 1 and 2 -> X[k12][r][f][c3], k12 = k1 + H1 k2
//   k_ewald_contract_z  one workgroup per wavevector k = k12 + H1 L2 k3: the
//                       DFT along axis 3 at k3 of every field's X column,
//                       then F[phi]_k = M_k F[mu]_k -> Y[r][f][k3][k12]
//   k_ewald_inv_slab    one workgroup per (sublattice, plane c3): the inverse
//                       DFT along axis 3 at c3 of Y, 2D c2r, 1/Ncell -> phi
// The axis-3 transforms are direct sums of L3 terms evaluated where they
// are needed; X and Y take the place of the blocked spectra (same size).

// X: position of (in-plane wavevector k12, replica r, field f, plane c3),
// so that the contraction's workgroup at k12 reads one contiguous run.
__device__ __forceinline__ long long slab_x_index(const EwaldGeom& G, int R, int k12, int r, int f, int c3)
{
    return (((long long)k12 * R + r) * (3 * G.na) + f) * G.L3 + c3;
}

// Y: position of (replica r, field f, wavevector k3, k12), so that the
// inverse's workgroups read runs of H1 L2 wavevectors.
__device__ __forceinline__ long long slab_y_index(const EwaldGeom& G, int r, int f, int k3, int k12)
{
    return (((long long)r * (3 * G.na) + f) * G.L3 + k3) * ((long long)G.H1 * G.L2) + k12;
}

// LDS of the plane kernels: three real planes [L2][L1 + 1], three half
// spectra [H1 L2], three twiddle tables.
template <typename T>
struct SlabLds {
    T* fld;
    C2<T>* spec;
    C2<T>* tw1;
    C2<T>* tw2;
    C2<T>* tw3;
    int row, plane, nkp;
};

template <typename T>
__device__ __forceinline__ SlabLds<T> slab_lds(const EwaldGeom& G)
{
    extern __shared__ __align__(16) unsigned char snh_lds[];
    SlabLds<T> F;
    F.row = G.L1 + 1;
    F.plane = G.L2 * F.row;
    F.nkp = G.H1 * G.L2;
    F.spec = reinterpret_cast<C2<T>*>(snh_lds);
    F.tw1 = F.spec + 3 * F.nkp;
    F.tw2 = F.tw1 + kLmax;
    F.tw3 = F.tw2 + kLmax;
    F.fld = reinterpret_cast<T*>(F.tw3 + kLmax);
    return F;
}

size_t ewald_slab_plane_lds_bytes(const EwaldGeom& G, size_t elem)
{
    return ((size_t)3 * G.H1 * G.L2 * 2 + 3 * 2 * kLmax + (size_t)3 * G.L2 * (G.L1 + 1)) * elem;
}

size_t ewald_slab_column_lds_bytes(const EwaldGeom& G, int R, size_t elem)
{
    return ((size_t)R * 3 * G.na * (G.L3 + 1) * 2 + 2 * kLmax) * elem;
}

// grid = (na L3, R): sublattice j = x / L3, plane c3 = x % L3.
template <typename T>
__global__ __launch_bounds__(kFftThreads) void k_ewald_fwd_slab(EwaldGeom G, int R, const ReplicaInfo* __restrict__ reps,
                                                                uint32_t total, const T* __restrict__ gmat,
                                                                const T* __restrict__ S, C2<T>* __restrict__ X)
{
    const int j = blockIdx.x / G.L3, c3 = blockIdx.x % G.L3, r = blockIdx.y;
    const SlabLds<T> F = slab_lds<T>(G);
    const uint32_t off = reps[r].off;
    const int npl = G.L1 * G.L2;  // <= kLmax^2 = 4 kFftThreads
    T v[4][3];
    // unconditional loads (a cell past the end reads cell 0 and is dropped)
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int c = q * kFftThreads + (int)threadIdx.x;
        const int cc = c < npl ? c : 0;
        const uint32_t g = ewald_site(G, off, j, cc % G.L1, cc / G.L1, c3);
        v[q][0] = ldg(S, g);
        v[q][1] = ldg(S + total, g);
        v[q][2] = ldg(S + 2 * (size_t)total, g);
    }
    make_twiddles(F.tw1, G.L1, -1);
    make_twiddles(F.tw2, G.L2, -1);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int c = q * kFftThreads + (int)threadIdx.x;
        if (c >= npl) continue;
        const int c1 = c % G.L1, c2 = c / G.L1;
#pragma unroll
        for (int beta = 0; beta < 3; beta++) {
            const T* gr = gmat + 9 * j + 3 * beta;  // mu_beta = -(g S)_beta, as k_ewald_fwd forms it
            F.fld[beta * F.plane + c2 * F.row + c1] = -(gr[0] * v[q][0] + gr[1] * v[q][1] + gr[2] * v[q][2]);
        }
    }
    __syncthreads();
    for (int line = threadIdx.x; line < 3 * G.L2; line += blockDim.x) {
        const int beta = line / G.L2, c2 = line % G.L2;
        line_r2c<T>(F.fld + beta * F.plane + c2 * F.row, F.spec + beta * F.nkp + G.H1 * c2, G.L1, F.tw1);
    }
    __syncthreads();
    for (int line = threadIdx.x; line < 3 * G.H1; line += blockDim.x) {
        const int beta = line / G.H1, k1 = line % G.H1;
        line_c2c<T>(F.spec + beta * F.nkp, k1, G.H1, G.L2, F.tw2);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 3 * F.nkp; t += blockDim.x) {
        const int beta = t / F.nkp, k12 = t % F.nkp;
        X[slab_x_index(G, R, k12, r, 3 * j + beta, c3)] = F.spec[t];
    }
}

// grid = nk (one workgroup per wavevector), all R replicas.  LDS: the X
// columns xs[R 3na][L3] at k12, the transformed vector v[R 3na], twiddles.
template <typename T>
__global__ __launch_bounds__(256) void k_ewald_contract_z(EwaldGeom G, int R, const C2<T>* __restrict__ M,
                                                          const C2<T>* __restrict__ X, C2<T>* __restrict__ Y)
{
    extern __shared__ __align__(16) unsigned char snh_lds[];
    const long long k = blockIdx.x;
    const int nkp = G.H1 * G.L2, L3 = G.L3;
    const int k12 = (int)(k % nkp), k3 = (int)(k / nkp);
    const int n3 = 3 * G.na, nv = R * n3;
    C2<T>* xs = reinterpret_cast<C2<T>*>(snh_lds);
    C2<T>* v = xs + nv * L3;
    C2<T>* tw = v + nv;
    const C2<T>* Xk = X + slab_x_index(G, R, k12, 0, 0, 0);  // [r n3 + f][c3], contiguous
    for (int t = threadIdx.x; t < nv * L3; t += blockDim.x) xs[t] = Xk[t];
    make_twiddles(tw, L3, -1);
    __syncthreads();
    for (int t = threadIdx.x; t < nv; t += blockDim.x) {
        T ar = 0, ai = 0;
        int m = 0;  // (k3 c3) mod L3
        for (int c3 = 0; c3 < L3; c3++) {
            const C2<T> x = xs[t * L3 + c3], w = tw[m];
            ar += x.re * w.re - x.im * w.im;
            ai += x.re * w.im + x.im * w.re;
            m += k3;
            if (m >= L3) m -= L3;
        }
        v[t] = {ar, ai};
    }
    __syncthreads();
    const C2<T>* Mk = M + k * n3 * n3;
    for (int t = threadIdx.x; t < nv; t += blockDim.x) {
        const int r = t / n3, row = t % n3;
        const C2<T>* vr = v + r * n3;
        T ar = 0, ai = 0;
        for (int col = 0; col < n3; col++) {
            const C2<T> m = Mk[col * n3 + row];
            const C2<T> x = vr[col];
            ar += m.re * x.re - m.im * x.im;
            ai += m.re * x.im + m.im * x.re;
        }
        Y[slab_y_index(G, r, row, k3, k12)] = {ar, ai};
    }
}

// grid = (na L3, R): sublattice i = x / L3, plane c3 = x % L3.
template <typename T, typename O>
__global__ __launch_bounds__(kFftThreads) void k_ewald_inv_slab(EwaldGeom G, const ReplicaInfo* __restrict__ reps,
                                                                uint32_t total, const C2<T>* __restrict__ Y,
                                                                O* __restrict__ phi)
{
    const int i = blockIdx.x / G.L3, c3 = blockIdx.x % G.L3, r = blockIdx.y;
    const SlabLds<T> F = slab_lds<T>(G);
    make_twiddles(F.tw1, G.L1, +1);
    make_twiddles(F.tw2, G.L2, +1);
    make_twiddles(F.tw3, G.L3, +1);
    __syncthreads();
    // the inverse transform along axis 3 at plane c3
    for (int t = threadIdx.x; t < 3 * F.nkp; t += blockDim.x) {
        const int alpha = t / F.nkp, k12 = t % F.nkp;
        T ar = 0, ai = 0;
        int m = 0;  // (k3 c3) mod L3
        const C2<T>* Yc = Y + slab_y_index(G, r, 3 * i + alpha, 0, k12);
        const long long nkp = F.nkp;
#pragma unroll 8
        for (int k3 = 0; k3 < G.L3; k3++) {
            const C2<T> y = Yc[k3 * nkp], w = F.tw3[m];
            ar += y.re * w.re - y.im * w.im;
            ai += y.re * w.im + y.im * w.re;
            m += c3;
            if (m >= G.L3) m -= G.L3;
        }
        F.spec[t] = {ar, ai};
    }
    __syncthreads();
    for (int line = threadIdx.x; line < 3 * G.H1; line += blockDim.x) {
        const int alpha = line / G.H1, k1 = line % G.H1;
        line_c2c<T>(F.spec + alpha * F.nkp, k1, G.H1, G.L2, F.tw2);
    }
    __syncthreads();
    for (int line = threadIdx.x; line < 3 * G.L2; line += blockDim.x) {
        const int alpha = line / G.L2, c2 = line % G.L2;
        line_c2r<T>(F.spec + alpha * F.nkp + G.H1 * c2, F.fld + alpha * F.plane + c2 * F.row, G.L1, F.tw1);
    }
    __syncthreads();
    const T scale = (T)(1.0 / (double)G.ncell);
    const uint32_t off = reps[r].off;
    const int npl = G.L1 * G.L2;
    for (int t = threadIdx.x; t < 3 * npl; t += blockDim.x) {
        const int alpha = t / npl, c = t % npl;
        const int c1 = c % G.L1, c2 = c / G.L1;
        const T v = F.fld[alpha * F.plane + c2 * F.row + c1] * scale;
        stg(phi + (size_t)alpha * total, ewald_site(G, off, i, c1, c2, c3), (O)v);
    }
}

// ---------------------------------------------------------------------------
// Library path: elementwise gather and scatter around the hipFFT transforms.

// fld[r][f][cell] = mu = -g S of field f = 3 j + beta.  grid = (cells / 256, 3 na, R).
template <typename T>
__global__ __launch_bounds__(256) void k_ewald_gather(EwaldGeom G, const ReplicaInfo* __restrict__ reps,
                                                      uint32_t total, const T* __restrict__ gmat,
                                                      const T* __restrict__ S, T* __restrict__ fld)
{
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    const int f = blockIdx.y, r = blockIdx.z;
    if (c >= G.ncell) return;
    const int j = f / 3, beta = f % 3;
    const int c1 = (int)(c % G.L1);
    const long long l = c / G.L1;
    const uint32_t g = ewald_site(G, reps[r].off, j, c1, (int)(l % G.L2), (int)(l / G.L2));
    const T g0 = gmat[9 * j + 3 * beta], g1 = gmat[9 * j + 3 * beta + 1], g2 = gmat[9 * j + 3 * beta + 2];
    fld[((long long)r * 3 * G.na + f) * G.ncell + c] =
        -(g0 * ldg(S, g) + g1 * ldg(S + total, g) + g2 * ldg(S + 2 * (size_t)total, g));
}

// phi plane alpha of sublattice i (field 3 i + alpha) = fld / Ncell.
template <typename T>
__global__ __launch_bounds__(256) void k_ewald_scatter(EwaldGeom G, const ReplicaInfo* __restrict__ reps,
                                                       uint32_t total, const T* __restrict__ fld, T* __restrict__ phi)
{
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    const int f = blockIdx.y, r = blockIdx.z;
    if (c >= G.ncell) return;
    const int i = f / 3, alpha = f % 3;
    const int c1 = (int)(c % G.L1);
    const long long l = c / G.L1;
    const T scale = (T)(1.0 / (double)G.ncell);
    stg(phi + (size_t)alpha * total, ewald_site(G, reps[r].off, i, c1, (int)(l % G.L2), (int)(l / G.L2)),
        fld[((long long)r * 3 * G.na + f) * G.ncell + c] * scale);
}

// The fields of source sublattice j of A [j][i][cell][3][3]:
// fld[(i 3 + alpha) 3 + beta][cell].  grid = (cells / 256, 9 na).
__global__ __launch_bounds__(256) void k_ewald_A_fields(EwaldGeom G, const double* __restrict__ A, int j,
                                                        double* __restrict__ fld)
{
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    const int q = blockIdx.y;  // (i 3 + alpha) 3 + beta
    if (c >= G.ncell) return;
    const int i = q / 9, ab = q % 9;
    fld[(long long)q * G.ncell + c] = A[(((long long)j * G.na + i) * G.ncell + c) * 9 + ab];
}

// M[(k n3 + col) n3 + row] = conj(spectrum of field (j, i, alpha, beta)) with
// col = 3 j + beta, row = 3 i + alpha, from spec[(i 3 + alpha) 3 + beta][k].
template <typename T>
__global__ __launch_bounds__(256) void k_ewald_M_from_spectra(EwaldGeom G, const C2<double>* __restrict__ spec, int j,
                                                              C2<T>* __restrict__ M)
{
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    const int q = blockIdx.y;
    if (k >= G.nk) return;
    const int i = q / 9, alpha = (q / 3) % 3, beta = q % 3;
    const int n3 = 3 * G.na;
    const C2<double> v = spec[(long long)q * G.nk + k];
    M[(k * n3 + 3 * j + beta) * n3 + 3 * i + alpha] = {(T)v.re, (T)(-v.im)};
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

size_t ewald_fft_lds_bytes(const EwaldGeom& G, size_t elem)
{
    return ((size_t)G.nk * 2 + 3 * 2 * kLmax + (size_t)G.L2 * G.L3 * (G.L1 + 1)) * elem;
}

size_t ewald_mfma_lds_bytes(const EwaldGeom& G)
{
    return (size_t)kMfmaKB * ((3 * G.na + 15) & ~15) * 33 * sizeof(C2<float>);
}

size_t ewald_spectrum_bytes(const EwaldGeom& G, int R, size_t elem)
{
    const long long nkp = (G.nk + kSpecKB - 1) / kSpecKB * kSpecKB;
    return (size_t)nkp * R * 3 * G.na * 2 * elem;
}

bool ewald_use_mfma(const EwaldGeom& G, int R, size_t elem)
{
    return elem == sizeof(float) && R >= kEwaldMfmaMinReplicas && 3 * G.na <= 16 * kMfmaMaxRowTiles;
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
    const size_t lds = ewald_fft_lds_bytes(G, sizeof(double));
    allow_dynamic_lds(reinterpret_cast<const void*>(&k_ewald_fwdA<T>), lds);
    hipLaunchKernelGGL(k_ewald_fwdA<T>, dim3(9 * G.na * G.na), dim3(kFftThreads), lds, st, G, A,
                       reinterpret_cast<C2<T>*>(M));
}

// F[phi] = M F[mu]: matrix cores for FP32 batches, else one matrix-vector
// product per wavevector.
template <typename T>
static void ewald_contract(const EwaldGeom& G, int R, const void* M, void* Fmu, void* Fphi, hipStream_t st)
{
    if constexpr (sizeof(T) == sizeof(float)) {
        if (ewald_use_mfma(G, R, sizeof(T))) {
            const size_t lm = ewald_mfma_lds_bytes(G);
            // column tiles of 16 replicas per workgroup: 2 reads each M_k once
            // for up to 32 replicas; SNH_EWALD_NC=1 forces one (tuning)
            static const int force_nc = [] {
                const char* e = std::getenv("SNH_EWALD_NC");
                return e ? std::atoi(e) : 0;
            }();
            const int nc = force_nc == 1 || (force_nc != 2 && R <= 16) ? 1 : 2;
            const dim3 grid((unsigned)((G.nk + kMfmaKB - 1) / kMfmaKB), (unsigned)((R + 16 * nc - 1) / (16 * nc)));
            const auto* m = reinterpret_cast<const C2<float>*>(M);
            const auto* x = reinterpret_cast<const C2<float>*>(Fmu);
            auto* y = reinterpret_cast<C2<float>*>(Fphi);
            const int nrt = (3 * G.na + 15) / 16;
#define SNH_EWM(NRT, NC)                                                                               \
    do {                                                                                               \
        allow_dynamic_lds(reinterpret_cast<const void*>(&k_ewald_contract_mfma<NRT, NC>), lm);           \
        hipLaunchKernelGGL((k_ewald_contract_mfma<NRT, NC>), grid, dim3(256 * NRT), lm, st, G, R, m, x, y); \
    } while (0)
#define SNH_EWM_NC(NRT) \
    if (nc == 1)        \
        SNH_EWM(NRT, 1); \
    else                \
        SNH_EWM(NRT, 2)
            switch (nrt) {
            case 1: SNH_EWM_NC(1); break;
            case 2: SNH_EWM_NC(2); break;
            case 3: SNH_EWM_NC(3); break;
            default: SNH_EWM_NC(4); break;
            }
#undef SNH_EWM_NC
#undef SNH_EWM
            return;
        }
    }
    const size_t lds2 = (size_t)R * 3 * G.na * sizeof(C2<T>);
    allow_dynamic_lds(reinterpret_cast<const void*>(&k_ewald_contract<T>), lds2);
    hipLaunchKernelGGL(k_ewald_contract<T>, dim3((unsigned)G.nk), dim3(256), lds2, st, G, R,
                       reinterpret_cast<const C2<T>*>(M), reinterpret_cast<const C2<T>*>(Fmu),
                       reinterpret_cast<C2<T>*>(Fphi));
}

// Batched 3D plans of the library path, their work area and the real field
// buffer (mu, then phi).
struct EwaldLibFft {
    hipfftHandle fwd = 0, inv = 0;
    void* work = nullptr;
    void* fld = nullptr;
};

// Creates a plan of `batch` 3D transforms over contiguous grids (real
// [L3][L2][L1], half spectrum [L3][L2][H1]) with auto-allocation off; its
// work size goes to *work.
static hipfftResult ewald_plan(hipfftHandle* h, const EwaldGeom& G, int batch, hipfftType type, size_t* work)
{
    int n[3] = {G.L3, G.L2, G.L1};
    int re[3] = {G.L3, G.L2, G.L1};
    int cx[3] = {G.L3, G.L2, G.H1};
    const bool r2c = type == HIPFFT_R2C || type == HIPFFT_D2Z;
    hipfftResult e = hipfftCreate(h);
    if (e != HIPFFT_SUCCESS) return e;
    e = hipfftSetAutoAllocation(*h, 0);
    if (e != HIPFFT_SUCCESS) return e;
    return hipfftMakePlanMany(*h, 3, n, r2c ? re : cx, 1, r2c ? (int)G.ncell : (int)G.nk, r2c ? cx : re, 1,
                              r2c ? (int)G.nk : (int)G.ncell, type, batch, work);
}

void ewald_lib_destroy(EwaldLibFft* L)
{
    if (!L) return;
    if (L->fwd) (void)hipfftDestroy(L->fwd);
    if (L->inv) (void)hipfftDestroy(L->inv);
    if (L->work) (void)hipFree(L->work);
    if (L->fld) (void)hipFree(L->fld);
    delete L;
}

EwaldLibFft* ewald_lib_create(const EwaldGeom& G, int R, size_t elem, std::string& err)
{
    EwaldLibFft* L = new EwaldLibFft;
    const bool f32 = elem == sizeof(float);
    const int batch = R * 3 * G.na;
    size_t w0 = 0, w1 = 0;
    hipfftResult e = ewald_plan(&L->fwd, G, batch, f32 ? HIPFFT_R2C : HIPFFT_D2Z, &w0);
    if (e == HIPFFT_SUCCESS) e = ewald_plan(&L->inv, G, batch, f32 ? HIPFFT_C2R : HIPFFT_Z2D, &w1);
    if (e != HIPFFT_SUCCESS) {
        err = "hipFFT plan creation failed (code " + std::to_string((int)e) + ")";
        ewald_lib_destroy(L);
        return nullptr;
    }
    const size_t w = std::max<size_t>(std::max(w0, w1), 256);
    if (hipMalloc(&L->work, w) != hipSuccess || hipMalloc(&L->fld, (size_t)batch * G.ncell * elem) != hipSuccess) {
        err = "out of device memory for the Ewald FFT buffers";
        ewald_lib_destroy(L);
        return nullptr;
    }
    if (hipfftSetWorkArea(L->fwd, L->work) != HIPFFT_SUCCESS || hipfftSetWorkArea(L->inv, L->work) != HIPFFT_SUCCESS) {
        err = "hipFFT work area not accepted";
        ewald_lib_destroy(L);
        return nullptr;
    }
    return L;
}

// A -> M by the library: per source sublattice j, 9 na FP64 fields
// transformed in one batch.  Temporary buffers are freed before returning.
template <typename T>
static std::string ewald_fwdA_lib(const EwaldGeom& G, const double* A, void* M, hipStream_t st)
{
    const int nf = 9 * G.na;
    hipfftHandle h = 0;
    size_t w = 0;
    double* fld = nullptr;
    C2<double>* spec = nullptr;
    void* work = nullptr;
    std::string err;
    hipfftResult e = ewald_plan(&h, G, nf, HIPFFT_D2Z, &w);
    if (e != HIPFFT_SUCCESS) err = "hipFFT plan creation failed (code " + std::to_string((int)e) + ")";
    if (err.empty() && (hipMalloc(&fld, sizeof(double) * nf * G.ncell) != hipSuccess ||
                        hipMalloc(&spec, sizeof(C2<double>) * nf * G.nk) != hipSuccess ||
                        hipMalloc(&work, std::max<size_t>(w, 256)) != hipSuccess))
        err = "out of device memory for the Ewald tensor transform";
    if (err.empty() && (hipfftSetWorkArea(h, work) != HIPFFT_SUCCESS || hipfftSetStream(h, st) != HIPFFT_SUCCESS))
        err = "hipFFT work area or stream not accepted";
    const dim3 gc((unsigned)((G.ncell + 255) / 256), nf), gk((unsigned)((G.nk + 255) / 256), nf);
    for (int j = 0; err.empty() && j < G.na; j++) {
        hipLaunchKernelGGL(k_ewald_A_fields, gc, dim3(256), 0, st, G, A, j, fld);
        if (hipfftExecD2Z(h, reinterpret_cast<hipfftDoubleReal*>(fld), reinterpret_cast<hipfftDoubleComplex*>(spec)) !=
            HIPFFT_SUCCESS)
            err = "hipFFT execution failed";
        else
            hipLaunchKernelGGL((k_ewald_M_from_spectra<T>), gk, dim3(256), 0, st, G, spec, j,
                               reinterpret_cast<C2<T>*>(M));
    }
    if (hipStreamSynchronize(st) != hipSuccess && err.empty()) err = "Ewald tensor transform failed";
    if (h) (void)hipfftDestroy(h);
    if (fld) (void)hipFree(fld);
    if (spec) (void)hipFree(spec);
    if (work) (void)hipFree(work);
    return err;
}

template <typename T>
std::string launch_ewald_fwdA_plain(const EwaldGeom& G, const double* A, void* M, hipStream_t st)
{
    return ewald_fwdA_lib<T>(G, A, M, st);
}

// Forward transform of every field, contraction, inverse transform.
template <typename T>
static bool ewald_transforms_lib(const EwaldGeom& G, int R, const ReplicaInfo* reps, uint32_t total, const T* gmat,
                                 const T* S, EwaldLibFft* L, void* Fmu, void* Fphi, hipStream_t st, bool fwd)
{
    T* fld = reinterpret_cast<T*>(L->fld);
    const dim3 grid((unsigned)((G.ncell + 255) / 256), 3 * G.na, R);
    if (fwd) {
        hipLaunchKernelGGL(k_ewald_gather<T>, grid, dim3(256), 0, st, G, reps, total, gmat, S, fld);
        if (hipfftSetStream(L->fwd, st) != HIPFFT_SUCCESS) return false;
        if constexpr (sizeof(T) == sizeof(float))
            return hipfftExecR2C(L->fwd, fld, reinterpret_cast<hipfftComplex*>(Fmu)) == HIPFFT_SUCCESS;
        else
            return hipfftExecD2Z(L->fwd, fld, reinterpret_cast<hipfftDoubleComplex*>(Fmu)) == HIPFFT_SUCCESS;
    }
    if (hipfftSetStream(L->inv, st) != HIPFFT_SUCCESS) return false;
    bool ok;
    if constexpr (sizeof(T) == sizeof(float))
        ok = hipfftExecC2R(L->inv, reinterpret_cast<hipfftComplex*>(Fphi), fld) == HIPFFT_SUCCESS;
    else
        ok = hipfftExecZ2D(L->inv, reinterpret_cast<hipfftDoubleComplex*>(Fphi), fld) == HIPFFT_SUCCESS;
    return ok;
}

template <typename T>
bool launch_ewald_field(const EwaldGeom& G, int R, const ReplicaInfo* reps, uint32_t total, const T* gmat,
                        const T* S, const void* M, void* Fmu, void* Fphi, T* phi, EwaldLibFft* lib, hipStream_t st,
                        double* phi64)
{
    if (phi64 && G.plain) return false;  // the library path writes T only
    if (G.plain) {
        if (!ewald_transforms_lib<T>(G, R, reps, total, gmat, S, lib, Fmu, Fphi, st, true)) return false;
        ewald_contract<T>(G, R, M, Fmu, Fphi, st);
        if (!ewald_transforms_lib<T>(G, R, reps, total, gmat, S, lib, Fmu, Fphi, st, false)) return false;
        const dim3 grid((unsigned)((G.ncell + 255) / 256), 3 * G.na, R);
        hipLaunchKernelGGL(k_ewald_scatter<T>, grid, dim3(256), 0, st, G, reps, total,
                           reinterpret_cast<const T*>(lib->fld), phi);
        return true;
    }
    if (G.slab && R <= kSlabMaxReplicas) {
        const size_t lp = ewald_slab_plane_lds_bytes(G, sizeof(T)), lc = ewald_slab_column_lds_bytes(G, R, sizeof(T));
        allow_dynamic_lds(reinterpret_cast<const void*>(&k_ewald_fwd_slab<T>), lp);
        hipLaunchKernelGGL(k_ewald_fwd_slab<T>, dim3(G.na * G.L3, R), dim3(kFftThreads), lp, st, G, R, reps, total,
                           gmat, S, reinterpret_cast<C2<T>*>(Fmu));
        allow_dynamic_lds(reinterpret_cast<const void*>(&k_ewald_contract_z<T>), lc);
        hipLaunchKernelGGL(k_ewald_contract_z<T>, dim3((unsigned)G.nk), dim3(256), lc, st, G, R,
                           reinterpret_cast<const C2<T>*>(M), reinterpret_cast<const C2<T>*>(Fmu),
                           reinterpret_cast<C2<T>*>(Fphi));
        const auto* y = reinterpret_cast<const C2<T>*>(Fphi);
        if (phi64) {
            allow_dynamic_lds(reinterpret_cast<const void*>(&k_ewald_inv_slab<T, double>), lp);
            hipLaunchKernelGGL((k_ewald_inv_slab<T, double>), dim3(G.na * G.L3, R), dim3(kFftThreads), lp, st, G,
                               reps, total, y, phi64);
        } else {
            allow_dynamic_lds(reinterpret_cast<const void*>(&k_ewald_inv_slab<T, T>), lp);
            hipLaunchKernelGGL((k_ewald_inv_slab<T, T>), dim3(G.na * G.L3, R), dim3(kFftThreads), lp, st, G, reps,
                               total, y, phi);
        }
        return true;
    }
    const size_t lds = ewald_fft_lds_bytes(G, sizeof(T));
    allow_dynamic_lds(reinterpret_cast<const void*>(&k_ewald_fwd<T>), lds);
    hipLaunchKernelGGL(k_ewald_fwd<T>, dim3(3 * G.na, R), dim3(kFftThreads), lds, st, G, R, reps, total, gmat, S,
                       reinterpret_cast<C2<T>*>(Fmu));
    ewald_contract<T>(G, R, M, Fmu, Fphi, st);
    const auto* y = reinterpret_cast<const C2<T>*>(Fphi);
    if (phi64) {
        allow_dynamic_lds(reinterpret_cast<const void*>(&k_ewald_inv<T, double>), lds);
        hipLaunchKernelGGL((k_ewald_inv<T, double>), dim3(3 * G.na, R), dim3(kFftThreads), lds, st, G, R, reps,
                           total, y, phi64);
    } else {
        allow_dynamic_lds(reinterpret_cast<const void*>(&k_ewald_inv<T, T>), lds);
        hipLaunchKernelGGL((k_ewald_inv<T, T>), dim3(3 * G.na, R), dim3(kFftThreads), lds, st, G, R, reps, total, y,
                           phi);
    }
    return true;
}

template void launch_ewald_fwdA<float>(const EwaldGeom&, const double*, void*, hipStream_t);
template void launch_ewald_fwdA<double>(const EwaldGeom&, const double*, void*, hipStream_t);
template std::string launch_ewald_fwdA_plain<float>(const EwaldGeom&, const double*, void*, hipStream_t);
template std::string launch_ewald_fwdA_plain<double>(const EwaldGeom&, const double*, void*, hipStream_t);
template bool launch_ewald_field<float>(const EwaldGeom&, int, const ReplicaInfo*, uint32_t, const float*,
                                        const float*, const void*, void*, void*, float*, EwaldLibFft*, hipStream_t,
                                        double*);
template bool launch_ewald_field<double>(const EwaldGeom&, int, const ReplicaInfo*, uint32_t, const double*,
                                         const double*, const void*, void*, void*, double*, EwaldLibFft*,
                                         hipStream_t, double*);

}  // namespace snh
