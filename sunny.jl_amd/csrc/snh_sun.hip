// :SUN-mode kernels: local Schroedinger field H*Z by bond-stencil gather of
// expected dipoles, fused Heun stages, midpoint sweeps, FP64 energy.
//
// Reference semantics (Sunny.jl 0.9.2):
//   set_energy_grad_coherents!   src/System/Interactions.jl:596-669
//   expected_spin                src/Operators/Spin.jl:72-87
//   expected_quadrupole          src/Operators/Spin.jl:90-104
//   mul_spin_matrices            src/Operators/Spin.jl:131-148
//   mul_quadrupole_matrices      src/Operators/Spin.jl:150-167
//   rhs_sun!, proj               src/Integrators.jl:277-286, src/MathBasics.jl:62
//   step!(::System{N}, ...)      src/Integrators.jl:335-357, 422-452
//   normalize_ket                src/System/System.jl:416-418
//
// The spin matrices are tridiagonal (Sx, Sy) and diagonal (Sz) in the
// descending-m basis, so (dE.S) Z is applied in O(N) without a table; the
// onsite matrix Lambda and the quadrupole operators are dense N x N tables.
// Heun folding as in :dipole mode: rhs_sun is linear in the noise,
// -proj(zeta + dt(i+lambda)HZ, Z) = -proj(zeta, Z) - proj(dt(i+lambda)HZ, Z),
// so the predictor stores H = Z + dZ1/2 - proj(zeta, Z')/2.
#include "snh_internal.h"
#include "snh_mc.h"

namespace snh {

template <typename T>
struct Cx {
    T re, im;
};

template <typename T>
__device__ __forceinline__ Cx<T> cmul(Cx<T> a, Cx<T> b)
{
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

template <typename T, int N>
struct Ket {
    Cx<T> z[N];
};

template <typename T, int N>
__device__ __forceinline__ Ket<T, N> load_ket(const T* __restrict__ P, uint32_t total, uint32_t g)
{
    Ket<T, N> k;
#pragma unroll
    for (int a = 0; a < N; a++) {
        k.z[a].re = ldg(P + (size_t)(2 * a) * total, g);
        k.z[a].im = ldg(P + (size_t)(2 * a + 1) * total, g);
    }
    return k;
}

template <typename T, int N>
__device__ __forceinline__ void store_ket(T* __restrict__ P, uint32_t total, uint32_t g, const Ket<T, N>& k)
{
#pragma unroll
    for (int a = 0; a < N; a++) {
        stg(P + (size_t)(2 * a) * total, g, k.z[a].re);
        stg(P + (size_t)(2 * a + 1) * total, g, k.z[a].im);
    }
}

template <typename T, int N>
__device__ __forceinline__ void store_ket_images(T* __restrict__ P, uint32_t total, const SiteCoord& sc,
                                                 const Ghost& G, const Ket<T, N>& k)
{
    store_ket<T, N>(P, total, sc.g, k);
    for_each_ghost(sc, G, [&](uint32_t gi) { store_ket<T, N>(P, total, gi, k); });
}

template <int N>
__device__ __forceinline__ double spin_off(int k)  // k = 1..N-1
{
    const double s = (N - 1) / 2.0;
    return sqrt(2 * (s + 1) * k - (double)k * (k + 1)) / 2;
}

// expected_spin: (2 Re c, 2 Im c, sum m_k |Z_k|^2), c = sum conj(Z_k) off_k Z_{k+1}
template <typename T, int N>
__device__ __forceinline__ void expected_spin(const Ket<T, N>& Z, T& sx, T& sy, T& sz)
{
    T cr = 0, ci = 0, nz = 0;
#pragma unroll
    for (int k = 0; k < N - 1; k++) {
        const T off = (T)spin_off<N>(k + 1);
        cr += off * (Z.z[k].re * Z.z[k + 1].re + Z.z[k].im * Z.z[k + 1].im);
        ci += off * (Z.z[k].re * Z.z[k + 1].im - Z.z[k].im * Z.z[k + 1].re);
    }
#pragma unroll
    for (int k = 0; k < N; k++) {
        const T m = (T)((N - 1) / 2.0 - k);
        nz += m * (Z.z[k].re * Z.z[k].re + Z.z[k].im * Z.z[k].im);
    }
    sx = 2 * cr;
    sy = 2 * ci;
    sz = nz;
}

// expected_quadrupole from the dense O[2,q] table ([5][N][N][2]).
template <typename T, int N>
__device__ __forceinline__ void expected_quadrupole(const Ket<T, N>& Z, const T* __restrict__ Q, T out[5])
{
#pragma unroll
    for (int p = 0; p < 5; p++) {
        T acc = 0;
        for (int a = 0; a < N; a++)
            for (int b = 0; b < N; b++) {
                const T* q = Q + 2 * ((p * N + a) * N + b);
                const Cx<T> qz = cmul(Cx<T>{q[0], q[1]}, Z.z[b]);
                acc += Z.z[a].re * qz.re + Z.z[a].im * qz.im;  // Re(conj(Za) qz)
            }
        out[p] = acc;
    }
}

// Re<Z|M|Z> for an N x N complex matrix M ([N][N][2], row-major, a read-only
// table indexed uniformly over the wavefront).
template <typename T, int N>
__device__ __forceinline__ T expect_op(const T* M, const Ket<T, N>& Z)
{
    T acc = 0;
#pragma unroll
    for (int a = 0; a < N; a++) {
        T pr = 0, pi = 0;
#pragma unroll
        for (int b = 0; b < N; b++) {
            const T mr = ldc(M, 2 * (a * N + b)), mi = ldc(M, 2 * (a * N + b) + 1);
            pr += mr * Z.z[b].re - mi * Z.z[b].im;
            pi += mr * Z.z[b].im + mi * Z.z[b].re;
        }
        acc += Z.z[a].re * pr + Z.z[a].im * pi;
    }
    return acc;
}

// out += f (M Z)
template <typename T, int N>
__device__ __forceinline__ void add_op_ket(const T* M, const Ket<T, N>& Z, T f, Ket<T, N>& out)
{
#pragma unroll
    for (int a = 0; a < N; a++) {
        T pr = 0, pi = 0;
#pragma unroll
        for (int b = 0; b < N; b++) {
            const T mr = ldc(M, 2 * (a * N + b)), mi = ldc(M, 2 * (a * N + b) + 1);
            pr += mr * Z.z[b].re - mi * Z.z[b].im;
            pi += mr * Z.z[b].im + mi * Z.z[b].re;
        }
        out.z[a].re += f * pr;
        out.z[a].im += f * pi;
    }
}

// Source of kets for the gather: stored planes, or Zbar = (Z + Z1)/2.
template <typename T, int N, bool kMid>
struct KetSrc {
    static constexpr bool kKets = true;
    using Raw = Ket<T, N>;
    const T* Z;
    const T* Z1;
    uint32_t total;
    __device__ __forceinline__ Ket<T, N> get(uint32_t g) const
    {
        Ket<T, N> a = load_ket<T, N>(Z, total, g);
        if (!kMid) return a;
        const Ket<T, N> b = load_ket<T, N>(Z1, total, g);
#pragma unroll
        for (int k = 0; k < N; k++) {
            a.z[k].re = (a.z[k].re + b.z[k].re) / 2;
            a.z[k].im = (a.z[k].im + b.z[k].im) / 2;
        }
        return a;
    }
    __device__ __forceinline__ Raw load(uint32_t g) const { return get(g); }
    __device__ __forceinline__ void spin(const Raw& r, T& sx, T& sy, T& sz) const { expected_spin<T, N>(r, sx, sy, sz); }
};

// Neighbour source of stored expected dipoles (3 planes with ghost images),
// written by the kernel that produced the state: the gather reads 3 values
// per neighbour instead of a ket, and recomputes nothing.  Bilinear
// couplings only (a biquadratic bond needs the neighbour's ket).
template <typename T, int N>
struct DipSrc {
    static constexpr bool kKets = false;
    struct Raw {
        T x, y, z;
    };
    const T* D;
    uint32_t total;
    __device__ __forceinline__ Raw load(uint32_t g) const { return {ldg(D, g), ldg(D, g + total), ldg(D, g + 2 * total)}; }
    __device__ __forceinline__ void spin(const Raw& r, T& sx, T& sy, T& sz) const
    {
        sx = r.x;
        sy = r.y;
        sz = r.z;
    }
};

// Stores the expected dipole of ket z at the site and its ghost images.
template <typename T, int N>
__device__ __forceinline__ void store_dipole_images(T* __restrict__ D, uint32_t total, const SiteCoord& sc,
                                                    const Ghost& G, const Ket<T, N>& z)
{
    T sx, sy, sz;
    expected_spin<T, N>(z, sx, sy, sz);
    auto st = [&](uint32_t g) {
        stg(D, g, sx);
        stg(D, g + total, sy);
        stg(D, g + 2 * total, sz);
    };
    st(sc.g);
    for_each_ghost(sc, G, st);
}

// Neighbours gathered per batch of loads (bond lists are padded to a multiple
// of it with zero self-couplings).
constexpr int kSunChunk = 6;

// H Z at one site (set_energy_grad_coherents!): dE/dS from Zeeman, bilinear
// exchange on expected dipoles and Ewald; then (Lambda + dE.S) Z plus
// biquadratic (dE/dQ . Q) Z.
template <typename T, int N, class Src>
__device__ __forceinline__ Ket<T, N> local_hz(const Couplings<T>& C, const SiteCoord& sc, const Src& src,
                                              const Ket<T, N>& Z, const T* __restrict__ phi)
{
    const int i = sc.i;
    T gx = ldc(C.zee, 3 * i), gy = ldc(C.zee, 3 * i + 1), gz = ldc(C.zee, 3 * i + 2);
    T wq[5] = {0, 0, 0, 0, 0};
    bool anyq = false;
    // general couplings: sum over terms of <Zj|B|Zj> A Z (Interactions.jl:659-666)
    Ket<T, N> hg;
#pragma unroll
    for (int a = 0; a < N; a++) hg.z[a] = {0, 0};
    const int* __restrict__ dl = C.bdelta + sc.R.dcls * C.nb;
    const int b0 = ldc(C.bstart, i), b1 = ldc(C.bstart, i + 1);
    for (int c0 = b0; c0 < b1; c0 += kSunChunk) {
        typename Src::Raw nbr[kSunChunk];
#pragma unroll
        for (int k = 0; k < kSunChunk; k++) nbr[k] = src.load(sc.g + (uint32_t)ldc(dl, c0 + k));
#pragma unroll
        for (int k = 0; k < kSunChunk; k++) {
            const int b = c0 + k;
            const int fl = ldc(C.bflag, b);
            T sx, sy, sz;
            src.spin(nbr[k], sx, sy, sz);
            const T* J = C.bJ + 9 * b;
            if (fl & kBondIso) {
                const T J0 = ldc(C.bJ0, b);
                gx += J0 * sx;
                gy += J0 * sy;
                gz += J0 * sz;
            } else {
                gx += ldc(J, 0) * sx + ldc(J, 1) * sy + ldc(J, 2) * sz;
                gy += ldc(J, 3) * sx + ldc(J, 4) * sy + ldc(J, 5) * sz;
                gz += ldc(J, 6) * sx + ldc(J, 7) * sy + ldc(J, 8) * sz;
            }
            if constexpr (Src::kKets) if (fl & kBondBiquad) {
                T Qj[5];
                expected_quadrupole<T, N>(nbr[k], C.quad_ops, Qj);
                const T* Bq = C.bQ + 25 * b;
                for (int p = 0; p < 5; p++)
                    for (int q = 0; q < 5; q++) wq[p] += Bq[5 * p + q] * Qj[q];
                anyq = true;
            }
            if constexpr (Src::kKets) if (fl & kBondGeneral) {
                for (int t = ldc(C.gstart, b); t < ldc(C.gstart, b + 1); t++) {
                    const T Bbar = expect_op<T, N>(C.gB + 2 * N * N * t, nbr[k]);
                    add_op_ket<T, N>(C.gA + 2 * N * N * t, Z, Bbar, hg);
                }
            }
        }
    }
    if (phi) {  // Ewald potential phi of the expected dipoles: dE/dS -= g' phi
        const T* gm = C.gmat + 9 * i;
        const T px = ldg(phi, sc.g), py = ldg(phi + C.total, sc.g), pz = ldg(phi + 2 * (size_t)C.total, sc.g);
        gx -= gm[0] * px + gm[3] * py + gm[6] * pz;
        gy -= gm[1] * px + gm[4] * py + gm[7] * pz;
        gz -= gm[2] * px + gm[5] * py + gm[8] * pz;
    }
    Ket<T, N> H;
    const T* L = C.lam + 2 * N * N * i;
#pragma unroll
    for (int a = 0; a < N; a++) {
        T hr = 0, hi = 0;
#pragma unroll
        for (int b = 0; b < N; b++) {
            const Cx<T> p = cmul(Cx<T>{L[2 * (a * N + b)], L[2 * (a * N + b) + 1]}, Z.z[b]);
            hr += p.re;
            hi += p.im;
        }
        // (dE.S Z)_a = gz m_a Z_a + off_{a+1} (gx - i gy) Z_{a+1} + off_a (gx + i gy) Z_{a-1}
        const T m = (T)((N - 1) / 2.0 - a);
        hr += gz * m * Z.z[a].re;
        hi += gz * m * Z.z[a].im;
        if (a + 1 < N) {
            const T off = (T)spin_off<N>(a + 1);
            hr += off * (gx * Z.z[a + 1].re + gy * Z.z[a + 1].im);
            hi += off * (gx * Z.z[a + 1].im - gy * Z.z[a + 1].re);
        }
        if (a > 0) {
            const T off = (T)spin_off<N>(a);
            hr += off * (gx * Z.z[a - 1].re - gy * Z.z[a - 1].im);
            hi += off * (gx * Z.z[a - 1].im + gy * Z.z[a - 1].re);
        }
        H.z[a].re = hr;
        H.z[a].im = hi;
    }
    if (anyq) {
        for (int a = 0; a < N; a++)
            for (int b = 0; b < N; b++) {
                T qr = 0, qi = 0;
                for (int p = 0; p < 5; p++) {
                    const T* q = C.quad_ops + 2 * ((p * N + a) * N + b);
                    qr += wq[p] * q[0];
                    qi += wq[p] * q[1];
                }
                const Cx<T> v = cmul(Cx<T>{qr, qi}, Z.z[b]);
                H.z[a].re += v.re;
                H.z[a].im += v.im;
            }
    }
#pragma unroll
    for (int a = 0; a < N; a++) {
        H.z[a].re += hg.z[a].re;
        H.z[a].im += hg.z[a].im;
    }
    return H;
}

// -proj(v, Z) = -(v - Z (Z'v)/(Z'Z))
template <typename T, int N>
__device__ __forceinline__ Ket<T, N> neg_proj(const Ket<T, N>& v, const Ket<T, N>& Z)
{
    T nr = 0, ni = 0, nn = 0;
#pragma unroll
    for (int a = 0; a < N; a++) {
        nr += Z.z[a].re * v.z[a].re + Z.z[a].im * v.z[a].im;
        ni += Z.z[a].re * v.z[a].im - Z.z[a].im * v.z[a].re;
        nn += Z.z[a].re * Z.z[a].re + Z.z[a].im * Z.z[a].im;
    }
    const T fr = nr / nn, fi = ni / nn;
    Ket<T, N> d;
#pragma unroll
    for (int a = 0; a < N; a++) {
        d.z[a].re = -(v.z[a].re - (Z.z[a].re * fr - Z.z[a].im * fi));
        d.z[a].im = -(v.z[a].im - (Z.z[a].re * fi + Z.z[a].im * fr));
    }
    return d;
}

// Deterministic part of rhs_sun!: lambda == 0: -i dt HZ; else -proj(dt (i + lambda) HZ, Z).
template <typename T, int N>
__device__ __forceinline__ Ket<T, N> rhs_det(const Ket<T, N>& Z, const Ket<T, N>& H, T dt, T lambda)
{
    Ket<T, N> d;
    if (lambda == (T)0) {
#pragma unroll
        for (int a = 0; a < N; a++) {
            d.z[a].re = dt * H.z[a].im;
            d.z[a].im = -dt * H.z[a].re;
        }
        return d;
    }
    Ket<T, N> v;
#pragma unroll
    for (int a = 0; a < N; a++) {  // (i + lambda) H = (lambda Hr - Hi) + i (Hr + lambda Hi)
        v.z[a].re = dt * (lambda * H.z[a].re - H.z[a].im);
        v.z[a].im = dt * (H.z[a].re + lambda * H.z[a].im);
    }
    return neg_proj<T, N>(v, Z);
}

// normalize_ket: Z / sqrt(Z'Z / kappa), norm in FP64.
template <typename T, int N>
__device__ __forceinline__ Ket<T, N> normalize_ket(Ket<T, N> Z, T kappa)
{
    if (kappa == (T)0) {
#pragma unroll
        for (int a = 0; a < N; a++) Z.z[a] = {0, 0};
        return Z;
    }
    double n2 = 0;
#pragma unroll
    for (int a = 0; a < N; a++) n2 += (double)Z.z[a].re * Z.z[a].re + (double)Z.z[a].im * Z.z[a].im;
    const double f = 1.0 / sqrt(n2 / (double)kappa);
#pragma unroll
    for (int a = 0; a < N; a++) {
        Z.z[a].re = (T)(Z.z[a].re * f);
        Z.z[a].im = (T)(Z.z[a].im * f);
    }
    return Z;
}

// fill_noise! for one site: zeta = sqrt(2 dt lambda kT) * randn(ComplexF64),
// each component with variance 1/2.
template <typename T, int N>
__device__ __forceinline__ Ket<T, N> site_noise_sun(const NoiseArgs<T>& nz, const SiteCoord& sc)
{
    Ket<T, N> k;
#pragma unroll
    for (int a = 0; a < N; a++) k.z[a] = {0, 0};
    if (nz.mode == 0) return k;
    const T s = nz.scale[sc.r];
    if (nz.mode == 1) {
        const double* h = nz.host + 2 * N * (size_t)sc.g;
#pragma unroll
        for (int a = 0; a < N; a++) k.z[a] = {(T)h[2 * a] * s, (T)h[2 * a + 1] * s};
        return k;
    }
    const T s2 = s * (T)0.70710678118654752440;
    const unsigned long long key = nz.seed + (unsigned long long)sc.r * 0x9E3779B97F4A7C15ull;
    const uint32_t site = (uint32_t)(sc.c + sc.R.ncell * sc.i);
#pragma unroll
    for (int blk = 0; blk < (2 * N + 3) / 4; blk++) {
        T n[4];
        philox_normals4<T>(key, noise_step(nz), site, (uint32_t)blk, n);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int idx = 4 * blk + q;
            if (idx < 2 * N) {
                if (idx & 1)
                    k.z[idx >> 1].im = n[q] * s2;
                else
                    k.z[idx >> 1].re = n[q] * s2;
            }
        }
    }
    return k;
}

// ---------------------------------------------------------------------------
// Kernels

template <typename T, int N>
__global__ __launch_bounds__(kBlock) void k_sun_expected_spin(Couplings<T> C, const T* __restrict__ Z,
                                                              T* __restrict__ S)
{
    const SiteCoord sc = site_coord(C.blocks, C.reps, C.G);
    if (!sc.valid) return;
    store_dipole_images<T, N>(S, C.total, sc, C.G, load_ket<T, N>(Z, C.total, sc.g));
}

// H Z at one site with neighbours gathered from their kets (kDip false) or
// from the expected-dipole planes D (kDip true).
template <typename T, int N, bool kDip>
__device__ __forceinline__ Ket<T, N> site_hz(const Couplings<T>& C, const SiteCoord& sc, const T* __restrict__ Z,
                                             const T* __restrict__ D, const Ket<T, N>& z, const T* __restrict__ phi)
{
    if constexpr (kDip)
        return local_hz<T, N>(C, sc, DipSrc<T, N>{D, C.total}, z, phi);
    else
        return local_hz<T, N>(C, sc, KetSrc<T, N, false>{Z, nullptr, C.total}, z, phi);
}

// Stores an updated ket.  With kDip the neighbours of the next stage read
// expected dipoles: the ket goes to its own site only and the dipoles (with
// ghost images) to Dout; otherwise the ket carries its ghost images.
template <typename T, int N, bool kDip>
__device__ __forceinline__ void store_state(T* __restrict__ Zo, T* __restrict__ Dout, const Couplings<T>& C,
                                            const SiteCoord& sc, const Ket<T, N>& z)
{
    if constexpr (kDip) {
        store_ket<T, N>(Zo, C.total, sc.g, z);
        store_dipole_images<T, N>(Dout, C.total, sc, C.G, z);
    } else {
        store_ket_images<T, N>(Zo, C.total, sc, C.G, z);
    }
}

// Heun predictor: dZ1 = rhs(Z, zeta, HZ); Z' = normalize(Z + dZ1);
// H = Z + dZ1/2 - proj(zeta, Z')/2.
template <typename T, int N, bool kDip>
__global__ __launch_bounds__(kBlock) void k_sun_predict(Couplings<T> C, const T* __restrict__ Z, T* __restrict__ Zp,
                                                        T* __restrict__ Hb, const T* __restrict__ phi,
                                                        NoiseArgs<T> nz, T dt, T lambda, const T* __restrict__ D,
                                                        T* __restrict__ Dp)
{
    const SiteCoord sc = site_coord(C.blocks, C.reps, C.G);
    if (!sc.valid) return;
    const Ket<T, N> z = load_ket<T, N>(Z, C.total, sc.g);
    Ket<T, N> d = rhs_det<T, N>(z, site_hz<T, N, kDip>(C, sc, Z, D, z, phi), dt, lambda);
    Ket<T, N> zeta;
    const bool noisy = lambda != (T)0 && nz.mode != 0;
    if (noisy) {
        zeta = site_noise_sun<T, N>(nz, sc);
        const Ket<T, N> pz = neg_proj<T, N>(zeta, z);
#pragma unroll
        for (int a = 0; a < N; a++) {
            d.z[a].re += pz.z[a].re;
            d.z[a].im += pz.z[a].im;
        }
    }
    Ket<T, N> zp;
#pragma unroll
    for (int a = 0; a < N; a++) zp.z[a] = {z.z[a].re + d.z[a].re, z.z[a].im + d.z[a].im};
    zp = normalize_ket<T, N>(zp, ldc(C.kappa, sc.i));
    store_state<T, N, kDip>(Zp, Dp, C, sc, zp);
    Ket<T, N> h;
#pragma unroll
    for (int a = 0; a < N; a++) h.z[a] = {z.z[a].re + d.z[a].re / 2, z.z[a].im + d.z[a].im / 2};
    if (noisy) {
        const Ket<T, N> pz = neg_proj<T, N>(zeta, zp);
#pragma unroll
        for (int a = 0; a < N; a++) {
            h.z[a].re += pz.z[a].re / 2;
            h.z[a].im += pz.z[a].im / 2;
        }
    }
    store_ket<T, N>(Hb, C.total, sc.g, h);
}

// Heun corrector: Z <- normalize(H + rhs_det(Z', HZ')/2).
template <typename T, int N, bool kDip>
__global__ __launch_bounds__(kBlock) void k_sun_correct(Couplings<T> C, const T* __restrict__ Zp,
                                                        const T* __restrict__ Hb, T* __restrict__ Zout,
                                                        const T* __restrict__ phi, T dt, T lambda,
                                                        const T* __restrict__ Dp, T* __restrict__ Dout)
{
    const SiteCoord sc = site_coord(C.blocks, C.reps, C.G);
    if (!sc.valid) return;
    const Ket<T, N> zp = load_ket<T, N>(Zp, C.total, sc.g);
    const Ket<T, N> d = rhs_det<T, N>(zp, site_hz<T, N, kDip>(C, sc, Zp, Dp, zp, phi), dt, lambda);
    Ket<T, N> h = load_ket<T, N>(Hb, C.total, sc.g);
#pragma unroll
    for (int a = 0; a < N; a++) {
        h.z[a].re += d.z[a].re / 2;
        h.z[a].im += d.z[a].im / 2;
    }
    // the state keeps its ket images (energy, gradients and midpoint gather kets)
    const Ket<T, N> zo = normalize_ket<T, N>(h, ldc(C.kappa, sc.i));
    store_ket_images<T, N>(Zout, C.total, sc, C.G, zo);
    if constexpr (kDip) store_dipole_images<T, N>(Dout, C.total, sc, C.G, zo);
}

template <typename T, int N>
__global__ __launch_bounds__(kBlock) void k_sun_mid_bar(Couplings<T> C, const T* __restrict__ Z,
                                                        const T* __restrict__ Z1, T* __restrict__ Zbar,
                                                        const int* __restrict__ conv_iter)
{
    const SiteCoord sc = site_coord(C.blocks, C.reps, C.G);
    if (!sc.valid || conv_iter[sc.r] != 0) return;
    const KetSrc<T, N, true> src{Z, Z1, C.total};
    store_ket_images<T, N>(Zbar, C.total, sc, C.G, src.get(sc.g));
}

template <typename T, int N>
__global__ __launch_bounds__(kBlock) void k_sun_mid_iter(Couplings<T> C, const T* __restrict__ Z,
                                                         const T* __restrict__ Z1, T* __restrict__ Z2,
                                                         const T* __restrict__ phi, T dt, T lambda,
                                                         double* __restrict__ partial,
                                                         const int* __restrict__ conv_iter)
{
    const SiteCoord sc = site_coord(C.blocks, C.reps, C.G);
    if (conv_iter[sc.r] != 0) return;  // block-uniform
    double d2 = 0;
    if (sc.valid) {
        const KetSrc<T, N, true> src{Z, Z1, C.total};
        const Ket<T, N> zb = src.get(sc.g);
        const Ket<T, N> d = rhs_det<T, N>(zb, local_hz<T, N>(C, sc, src, zb, phi), dt, lambda);
        Ket<T, N> z = load_ket<T, N>(Z, C.total, sc.g);
        const Ket<T, N> z1 = load_ket<T, N>(Z1, C.total, sc.g);
#pragma unroll
        for (int a = 0; a < N; a++) {
            z.z[a].re += d.z[a].re;
            z.z[a].im += d.z[a].im;
            const double er = (double)z1.z[a].re - (double)z.z[a].re, ei = (double)z1.z[a].im - (double)z.z[a].im;
            d2 += er * er + ei * ei;
        }
        store_ket_images<T, N>(Z2, C.total, sc, C.G, z);
    }
    const double t = block_sum(d2);
    if (threadIdx.x == 0) partial[sc.blk] = t;
}

template <typename T, int N>
__global__ __launch_bounds__(kBlock) void k_sun_mid_finalize(Couplings<T> C, T* __restrict__ Z,
                                                             const T* __restrict__ Za, const T* __restrict__ Zb,
                                                             const int* __restrict__ conv_iter, int iter)
{
    const SiteCoord sc = site_coord(C.blocks, C.reps, C.G);
    const int c = sc.valid ? conv_iter[sc.r] : 0;  // iter 0: every converged replica
    if (c == 0 || (iter != 0 && c != iter)) return;
    const T* src = (c & 1) ? Za : Zb;
    store_ket_images<T, N>(Z, C.total, sc, C.G,
                           normalize_ket<T, N>(load_ket<T, N>(src, C.total, sc.g), C.kappa[sc.i]));
}

template <typename T, int N>
__global__ __launch_bounds__(kBlock) void k_sun_grad(Couplings<T> C, const T* __restrict__ Z,
                                                     const T* __restrict__ phi, T* __restrict__ HZ)
{
    const SiteCoord sc = site_coord(C.blocks, C.reps, C.G);
    if (!sc.valid) return;
    const KetSrc<T, N, false> src{Z, nullptr, C.total};
    const Ket<T, N> z = load_ket<T, N>(Z, C.total, sc.g);
    store_ket<T, N>(HZ, C.total, sc.g, local_hz<T, N>(C, sc, src, z, phi));
}

template <typename T, int N>
__global__ __launch_bounds__(kBlock) void k_sun_energy(Couplings<T> C, const T* __restrict__ Z,
                                                       const T* __restrict__ phi, double* __restrict__ partial)
{
    const SiteCoord sc = site_coord(C.blocks, C.reps, C.G);
    double e = 0;
    if (sc.valid) {
        const int i = sc.i;
        const Ket<T, N> z = load_ket<T, N>(Z, C.total, sc.g);
        T sx, sy, sz;
        expected_spin<T, N>(z, sx, sy, sz);
        const T* gm = C.gmat + 9 * i;
        const T* B = C.field + 3 * i;
        for (int a = 0; a < 3; a++)
            e += (double)B[a] * ((double)gm[3 * a] * sx + (double)gm[3 * a + 1] * sy + (double)gm[3 * a + 2] * sz);
        const T* L = C.lam + 2 * N * N * i;  // onsite: Re(Z' Lambda Z)
        for (int a = 0; a < N; a++)
            for (int b = 0; b < N; b++) {
                const Cx<T> p = cmul(Cx<T>{L[2 * (a * N + b)], L[2 * (a * N + b) + 1]}, z.z[b]);
                e += (double)z.z[a].re * p.re + (double)z.z[a].im * p.im;
            }
        const int* __restrict__ dl = C.bdelta + sc.R.dcls * C.nb;
        for (int b = C.bstart[i]; b < C.bstart[i + 1]; b++) {
            const int fl = C.bflag[b];
            if (fl & kBondCulled) continue;
            const int j = C.bj[b];
            const Ket<T, N> zj = load_ket<T, N>(Z, C.total, sc.g + (uint32_t)dl[b]);
            T tx, ty, tz;
            expected_spin<T, N>(zj, tx, ty, tz);
            const T* J = C.bJ + 9 * b;
            double v = (double)C.bscalar[b] * (double)C.kappa[i] * (double)C.kappa[j];
            v += (double)sx * ((double)J[0] * tx + (double)J[1] * ty + (double)J[2] * tz);
            v += (double)sy * ((double)J[3] * tx + (double)J[4] * ty + (double)J[5] * tz);
            v += (double)sz * ((double)J[6] * tx + (double)J[7] * ty + (double)J[8] * tz);
            if (fl & kBondBiquad) {
                T Qi[5], Qj[5];
                expected_quadrupole<T, N>(z, C.quad_ops, Qi);
                expected_quadrupole<T, N>(zj, C.quad_ops, Qj);
                const T* Bq = C.bQ + 25 * b;
                for (int p = 0; p < 5; p++)
                    for (int q = 0; q < 5; q++) v += (double)Qi[p] * (double)Bq[5 * p + q] * (double)Qj[q];
            }
            if (fl & kBondGeneral)  // Interactions.jl:496-503
                for (int t = C.gstart[b]; t < C.gstart[b + 1]; t++)
                    v += (double)expect_op<T, N>(C.gA + 2 * N * N * t, z) *
                         (double)expect_op<T, N>(C.gB + 2 * N * N * t, zj);
            e += v;
        }
        if (phi) {  // 1/2 mu.phi with mu = -g <S>
            const double mx = -((double)gm[0] * sx + (double)gm[1] * sy + (double)gm[2] * sz);
            const double my = -((double)gm[3] * sx + (double)gm[4] * sy + (double)gm[5] * sz);
            const double mz = -((double)gm[6] * sx + (double)gm[7] * sy + (double)gm[8] * sz);
            e += 0.5 * (mx * ldg(phi, sc.g) + my * ldg(phi + C.total, sc.g) + mz * ldg(phi + 2 * (size_t)C.total, sc.g));
        }
    }
    const double t = block_sum(e);
    if (threadIdx.x == 0) partial[sc.blk] = t;
}

// ---------------------------------------------------------------------------
// Metropolis trials of one colour class for :SUN (the pass scheme, work
// lists and draws of snh_mc.hip): LocalSampler's step! (Samplers.jl:133-153)
// with randspin (System.jl:448-452), flip (flip_ket, Spin.jl:119-125) or
// perturbed_spin (System.jl:461-468) proposals and local_energy_change
// (Interactions.jl:311-380) in FP64: Zeeman and bilinear terms on expected
// spins, Re(Z' Lambda Z) onsite, biquadratic terms on expected quadrupoles.
// Accepted kets are stored with their ghost images.  Neighbour kets are
// gathered 3, 2 or 1 bonds at a time as N grows (the SU(N) gather lists are
// padded to 6), which keeps the kernel in registers.

// Re(Zn' Lambda Zn) - Re(Z0' Lambda Z0) in one pass over Lambda.
template <typename T, int N>
__device__ __forceinline__ double onsite_delta(const Couplings<T>& C, int i, const Ket<T, N>& zn, const Ket<T, N>& z0)
{
    const T* L = C.lam + 2 * N * N * i;
    double e = 0;
#pragma unroll
    for (int a = 0; a < N; a++) {
        T nr = 0, ni = 0, orr = 0, oi = 0;  // (Lambda Zn)_a, (Lambda Z0)_a
#pragma unroll
        for (int b = 0; b < N; b++) {
            const T lr = ldc(L, 2 * (a * N + b)), li = ldc(L, 2 * (a * N + b) + 1);
            nr += lr * zn.z[b].re - li * zn.z[b].im;
            ni += lr * zn.z[b].im + li * zn.z[b].re;
            orr += lr * z0.z[b].re - li * z0.z[b].im;
            oi += lr * z0.z[b].im + li * z0.z[b].re;
        }
        e += (double)(zn.z[a].re * nr + zn.z[a].im * ni) - (double)(z0.z[a].re * orr + z0.z[a].im * oi);
    }
    return e;
}

// The 2N normals of a ket proposal: Philox4x32-10 of counter c, its second
// word advanced by 2 per block of four, each block two Box-Muller pairs.
template <typename T, int N>
__device__ __forceinline__ void sun_mc_normals(const uint32_t c[4], unsigned long long key, T n[2 * N])
{
    constexpr int NB = (2 * N + 3) / 4;
#pragma unroll
    for (int b = 0; b < NB; b++) {
        uint32_t cn[4] = {c[0], c[1] + 2u * (uint32_t)b, c[2], c[3]};
        philox4x32(cn, (uint32_t)key, (uint32_t)(key >> 32));
        T m[4];
        box_muller(cn[0], cn[1], m[0], m[1]);
        box_muller(cn[2], cn[3], m[2], m[3]);
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (4 * b + q < 2 * N) n[4 * b + q] = m[q];
    }
}

// Proposal k of LocalSampler for ket z0: flip_ket, randspin or
// perturbed_spin Z + mag sqrt(kappa) randn(CVec{N}) from the normals n.
template <typename T, int N>
__device__ __forceinline__ Ket<T, N> sun_mc_propose_from(const McArgs& A, int k, const Ket<T, N>& z0, T kap,
                                                         const T n[2 * N])
{
    Ket<T, N> zn;
    if (A.kind[k] == 1) {  // flip_ket: reverse(parity .* conj(Z))
#pragma unroll
        for (int a = 0; a < N; a++) {
            const int j = N - 1 - a;
            const T sg = (j & 1) ? (T)-1 : (T)1;
            zn.z[a] = {sg * z0.z[j].re, -sg * z0.z[j].im};
        }
        return zn;
    }
    const double mag = A.mag[k];
    const bool unif = A.kind[k] == 0 || __builtin_isinf(mag);
    const T f = unif ? (T)0 : (T)(mag * sqrt((double)kap));
    const T h = (T)0.70710678118654752440;  // randn(ComplexF64): variance 1/2 per part
#pragma unroll
    for (int a = 0; a < N; a++) {
        const T nr = n[2 * a] * h, ni = n[2 * a + 1] * h;
        zn.z[a] = unif ? Cx<T>{nr, ni} : Cx<T>{z0.z[a].re + f * nr, z0.z[a].im + f * ni};
    }
    return normalize_ket<T, N>(zn, kap);
}

// Proposal k of LocalSampler for ket z0, normals from counter c (not drawn
// for flip_ket).
template <typename T, int N>
__device__ __forceinline__ Ket<T, N> sun_mc_propose(const McArgs& A, int k, const Ket<T, N>& z0, T kap,
                                                    const uint32_t c[4], unsigned long long key)
{
    T n[2 * N];
#pragma unroll
    for (int q = 0; q < 2 * N; q++) n[q] = 0;
    if (A.kind[k] != 1) sun_mc_normals<T, N>(c, key, n);
    return sun_mc_propose_from<T, N>(A, k, z0, kap, n);
}

// local_energy_change (Interactions.jl:311-380) of site sc going from z0 to
// zn, without the Ewald term, in FP64: Zeeman and bilinear terms on expected
// spins, Re(Z' Lambda Z) onsite, biquadratic terms on expected quadrupoles,
// general couplings.  d: change of the expected spin.  Neighbour kets are
// gathered G bonds at a time.
template <typename T, int N, int G>
__device__ __forceinline__ double sun_mc_local_delta(const Couplings<T>& C, const McArgs& A, const T* __restrict__ Z,
                                                     const SiteCoord& sc, const Ket<T, N>& z0, const Ket<T, N>& zn,
                                                     bool wide, double d[3])
{
    const int i = sc.i;
    const uint32_t total = C.total;
    T ox, oy, oz, nx, ny, nz;
    expected_spin<T, N>(z0, ox, oy, oz);
    expected_spin<T, N>(zn, nx, ny, nz);
    const double dx = (double)nx - ox, dy = (double)ny - oy, dz = (double)nz - oz;
    d[0] = dx;
    d[1] = dy;
    d[2] = dz;
    double dE = (double)ldc(C.zee, 3 * i) * dx + (double)ldc(C.zee, 3 * i + 1) * dy + (double)ldc(C.zee, 3 * i + 2) * dz;
    dE += onsite_delta<T, N>(C, i, zn, z0);
    T Qn[5], Qo[5];
    bool have_q = false;
    const int b0 = ldc(C.bstart, i), b1 = ldc(C.bstart, i + 1);
    for (int c0 = b0; c0 < b1; c0 += G) {
        Ket<T, N> zj[G];
#pragma unroll
        for (int kk = 0; kk < G; kk++) zj[kk] = load_ket<T, N>(Z, total, mc_neighbour(A, C.G, sc, c0 + kk, wide));
#pragma unroll
        for (int kk = 0; kk < G; kk++) {
            const int b = c0 + kk;
            T sx, sy, sz;
            expected_spin<T, N>(zj[kk], sx, sy, sz);
            const double x = sx, y = sy, z = sz;
            const T* J = C.bJ + 9 * b;
            dE += dx * ((double)ldc(J, 0) * x + (double)ldc(J, 1) * y + (double)ldc(J, 2) * z) +
                  dy * ((double)ldc(J, 3) * x + (double)ldc(J, 4) * y + (double)ldc(J, 5) * z) +
                  dz * ((double)ldc(J, 6) * x + (double)ldc(J, 7) * y + (double)ldc(J, 8) * z);
            if (ldc(C.bflag, b) & kBondBiquad) {
                if (!have_q) {
                    expected_quadrupole<T, N>(zn, C.quad_ops, Qn);
                    expected_quadrupole<T, N>(z0, C.quad_ops, Qo);
                    have_q = true;
                }
                T Qj[5];
                expected_quadrupole<T, N>(zj[kk], C.quad_ops, Qj);
                const T* Bq = C.bQ + 25 * b;
                for (int p = 0; p < 5; p++) {
                    double w = 0;
                    for (int q = 0; q < 5; q++) w += (double)ldc(Bq, 5 * p + q) * (double)Qj[q];
                    dE += ((double)Qn[p] - (double)Qo[p]) * w;
                }
            }
            if (ldc(C.bflag, b) & kBondGeneral)  // Interactions.jl:371-377
                for (int t = ldc(C.gstart, b); t < ldc(C.gstart, b + 1); t++) {
                    const T* A_ = C.gA + 2 * N * N * t;
                    const double dA = (double)expect_op<T, N>(A_, zn) - (double)expect_op<T, N>(A_, z0);
                    dE += dA * (double)expect_op<T, N>(C.gB + 2 * N * N * t, zj[kk]);
                }
        }
    }
    return dE;
}

// Neighbour kets gathered per step of the Metropolis kernels: 3, 2 or 1 as N
// grows (the SU(N) gather lists are padded to 6), which keeps them in
// registers.
template <typename T, int N>
constexpr int sun_mc_gather()
{
    return (sizeof(T) == 4 ? N <= 3 : N <= 2) ? 3 : (N <= 4 ? 2 : 1);
}

template <typename T, int N>
__global__ __launch_bounds__(kBlock) void k_sun_mc_color(Couplings<T> C, McArgs A, T* __restrict__ Z,
                                                         const McWork* __restrict__ work, unsigned long long P,
                                                         double prob, double* __restrict__ acc)
{
    const McWork W = work[blockIdx.x];
    const SiteCoord sc = mc_work_site(C, A, W, (int)threadIdx.x);
    const int i = sc.i;
    const uint32_t total = C.total;
    double v[4] = {0, 0, 0, 0};
    const unsigned long long key = A.seed + (unsigned long long)W.r * 0x9E3779B97F4A7C15ull;
    const uint32_t site = (uint32_t)sc.c + (uint32_t)sc.R.ncell * (uint32_t)i;
    uint32_t un[4] = {site, kMcTagUniforms, (uint32_t)P, (uint32_t)(P >> 32)};
    philox4x32(un, (uint32_t)key, (uint32_t)(key >> 32));
    const double inv32 = 1.0 / 4294967296.0;
    const double u_acc = (double)un[0] * inv32, u_mix = (double)un[1] * inv32, u_try = (double)un[2] * inv32;
    if (sc.valid && u_try < prob) {
        const Ket<T, N> z0 = load_ket<T, N>(Z, total, sc.g);
        const uint32_t cn[4] = {site, kMcTagNormals, (uint32_t)P, (uint32_t)(P >> 32)};
        const Ket<T, N> zn = sun_mc_propose<T, N>(A, mc_pick(A, u_mix), z0, ldc(C.kappa, i), cn, key);
        double d[3];
        const double dE = sun_mc_local_delta<T, N, sun_mc_gather<T, N>()>(C, A, Z, sc, z0, zn, A.wide != 0, d);
        const double kT = ldc(A.kT, W.r);
        const bool accept = kT == 0.0 ? dE <= 0.0 : u_acc <= exp(-dE / kT);
        if (accept) {
            store_ket_images<T, N>(Z, total, sc, C.G, zn);
            v[0] = dE;
            v[1] = d[0];
            v[2] = d[1];
            v[3] = d[2];
        }
    }
    block_sum4(v);
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; q++) acc[4 * (size_t)blockIdx.x + q] += v[q];
}

// Sequential trials T0 .. T0 + ntrial - 1 of the Ewald form for :SUN,
// workgroup b < nrep for replica r0 + b, one wavefront; workgroups b >= nrep
// are the apply part (mc_apply_part).  Lane l first forms the draws of
// trial T0 + l (those of the :dipole form: mc_seq_draw; normals from the
// counters (T lo, kMcTagSequential + 1, T hi, 0)) and the potential at its
// site; then all lanes make the trials in order, each taking its draws from
// the lane that formed them: the energy change of the ket plus
// ewald_energy_delta of its expected dipole (Ewald.jl:234-241) at the
// potential of the chunk's start plus that of the moves accepted since
// (McChunkPot).  Lane 0 stores an accepted ket with its ghost images.  The
// chunk's accepted moves go to io.moves[r][.] and io.nmoves[r];
// acc[4 r + q] += (dE, dS).
template <typename T, int N>
__global__ __launch_bounds__(64) void k_sun_mc_ewald_chunk(Couplings<T> C, McArgs A, T* __restrict__ Z, McChunkIO io,
                                                           const double* __restrict__ Aew,
                                                           const double* __restrict__ A0, unsigned long long T0,
                                                           int ntrial, double* __restrict__ acc, int r0, int nrep)
{
    if ((int)blockIdx.x >= nrep) {  // uniform over the workgroup
        mc_apply_part(io, Aew, C.reps, C.G, C.na, C.total, r0, (int)blockIdx.x - nrep);
        return;
    }
    const int r = r0 + (int)blockIdx.x;
    const ReplicaInfo R = C.reps[r];
    const uint32_t total = C.total;
    const unsigned long long key = A.seed + (unsigned long long)r * 0x9E3779B97F4A7C15ull;
    const double kT = A.kT[r];
    McSeqSite my;
    T my_n[2 * N];
    McChunkPot P;
    {
        const unsigned long long Tl = T0 + threadIdx.x;
        my = mc_seq_site(R, mc_seq_draw(key, Tl, (uint32_t)R.ncell * (uint32_t)C.na), A);
        const uint32_t cn[4] = {(uint32_t)Tl, kMcTagSequential + 1u, (uint32_t)(Tl >> 32), 0u};
        sun_mc_normals<T, N>(cn, key, my_n);
        mc_pot_start(P, io, Aew, C.na, R, r, mc_site(R, C.G, my.i, my.c1, my.c2, my.c3, true).g, total, my.i,
                     my.c1, my.c2, my.c3, (int)threadIdx.x);
    }
    McTrial mv{};
    int nm = 0;
    double sum[4] = {0, 0, 0, 0};
    for (int t = 0; t < ntrial; t++) {
        const McSeqSite d = mc_seq_site_of_lane(my, t);
        T n[2 * N];
#pragma unroll
        for (int q = 0; q < 2 * N; q++) n[q] = lane_value(my_n[q], t);
        const int i = d.i;
        SiteCoord sc = mc_site(R, C.G, i, d.c1, d.c2, d.c3, true);
        sc.r = r;
        const Ket<T, N> z0 = load_ket<T, N>(Z, total, sc.g);
        double p[3];
        mc_pot_take(P, t, p);
        const Ket<T, N> zn = sun_mc_propose_from<T, N>(A, d.k, z0, ldc(C.kappa, i), n);
        double dS[3];
        double dE = sun_mc_local_delta<T, N, sun_mc_gather<T, N>()>(C, A, Z, sc, z0, zn, A.wide != 0, dS);
        double dmu[3];
        dE += mc_ewald_delta(C, i, p, A0, dS, dmu);
        const int accept = __builtin_amdgcn_readfirstlane(kT == 0.0 ? dE <= 0.0 : d.u_acc <= exp(-dE / kT));
        if (accept) {
            if (threadIdx.x == 0) store_ket_images<T, N>(Z, total, sc, C.G, zn);
            if ((int)threadIdx.x == nm) mv = {1, i, d.c1, d.c2, d.c3, 0, {dmu[0], dmu[1], dmu[2]}};
            nm++;
            mc_pot_accept(P, Aew, C.na, R, t, ntrial, i, d.c1, d.c2, d.c3, dmu, my.i, my.c1, my.c2, my.c3);
            sum[0] += dE;
            for (int a = 0; a < 3; a++) sum[1 + a] += dS[a];
        }
        // lane 0's store precedes every lane's loads of the next trial: the
        // lanes of one wavefront see each other's writes in program order
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    if ((int)threadIdx.x < nm) io.moves[(size_t)r * kMcChunk + threadIdx.x] = mv;
    if (threadIdx.x == 0) {
        io.nmoves[r] = nm;
        for (int q = 0; q < 4; q++) acc[4 * r + q] += sum[q];
    }
}

// ---------------------------------------------------------------------------
// Launchers: dispatch the run-time N to a compiled specialisation.

#define SNH_SUN_DISPATCH(N, KERNEL, ...)                             \
    switch (N) {                                                     \
    case 2: hipLaunchKernelGGL((KERNEL<T, 2>), __VA_ARGS__); break;  \
    case 3: hipLaunchKernelGGL((KERNEL<T, 3>), __VA_ARGS__); break;  \
    case 4: hipLaunchKernelGGL((KERNEL<T, 4>), __VA_ARGS__); break;  \
    case 5: hipLaunchKernelGGL((KERNEL<T, 5>), __VA_ARGS__); break;  \
    case 6: hipLaunchKernelGGL((KERNEL<T, 6>), __VA_ARGS__); break;  \
    case 7: hipLaunchKernelGGL((KERNEL<T, 7>), __VA_ARGS__); break;  \
    default: break;                                                  \
    }

// As SNH_SUN_DISPATCH, with one more template argument X after N.
#define SNH_SUN_DISPATCH_X(N, KERNEL, X, ...)                           \
    switch (N) {                                                        \
    case 2: hipLaunchKernelGGL((KERNEL<T, 2, X>), __VA_ARGS__); break;  \
    case 3: hipLaunchKernelGGL((KERNEL<T, 3, X>), __VA_ARGS__); break;  \
    case 4: hipLaunchKernelGGL((KERNEL<T, 4, X>), __VA_ARGS__); break;  \
    case 5: hipLaunchKernelGGL((KERNEL<T, 5, X>), __VA_ARGS__); break;  \
    case 6: hipLaunchKernelGGL((KERNEL<T, 6, X>), __VA_ARGS__); break;  \
    case 7: hipLaunchKernelGGL((KERNEL<T, 7, X>), __VA_ARGS__); break;  \
    default: break;                                                     \
    }

template <typename T>
void launch_sun_expected_spin(const Couplings<T>& C, int nblocks, const T* Z, T* S, hipStream_t st)
{
    SNH_SUN_DISPATCH(C.N, k_sun_expected_spin, dim3(nblocks), dim3(kBlock), 0, st, C, Z, S)
}

template <typename T>
void launch_sun_predict(const Couplings<T>& C, int nblocks, const T* Z, T* Zp, T* H, const T* phi,
                        const NoiseArgs<T>& nz, T dt, T lambda, const T* D, T* Dp, hipStream_t st)
{
#define SNH_SP(KD)                                                                                    \
    SNH_SUN_DISPATCH_X(C.N, k_sun_predict, KD, dim3(nblocks), dim3(kBlock), 0, st, C, Z, Zp, H, phi, nz, dt, \
                       lambda, D, Dp)
    if (D) {
        SNH_SP(true)
    } else {
        SNH_SP(false)
    }
#undef SNH_SP
}

template <typename T>
void launch_sun_correct(const Couplings<T>& C, int nblocks, const T* Zp, const T* H, T* Zout, const T* phi, T dt,
                        T lambda, const T* Dp, T* Dout, hipStream_t st)
{
#define SNH_SC(KD)                                                                                   \
    SNH_SUN_DISPATCH_X(C.N, k_sun_correct, KD, dim3(nblocks), dim3(kBlock), 0, st, C, Zp, H, Zout, phi, dt, \
                       lambda, Dp, Dout)
    if (Dp) {
        SNH_SC(true)
    } else {
        SNH_SC(false)
    }
#undef SNH_SC
}

template <typename T>
void launch_sun_midpoint_bar(const Couplings<T>& C, int nblocks, const T* Z, const T* Z1, T* Zbar,
                             const int* conv_iter, hipStream_t st)
{
    SNH_SUN_DISPATCH(C.N, k_sun_mid_bar, dim3(nblocks), dim3(kBlock), 0, st, C, Z, Z1, Zbar, conv_iter)
}

template <typename T>
void launch_sun_midpoint_iter(const Couplings<T>& C, int nblocks, const T* Z, const T* Z1, T* Z2, const T* phi, T dt,
                              T lambda, double* partial, const int* conv_iter, hipStream_t st)
{
    SNH_SUN_DISPATCH(C.N, k_sun_mid_iter, dim3(nblocks), dim3(kBlock), 0, st, C, Z, Z1, Z2, phi, dt, lambda, partial,
                     conv_iter)
}

template <typename T>
void launch_sun_midpoint_finalize(const Couplings<T>& C, int nblocks, T* Z, const T* Za, const T* Zb,
                                  const int* conv_iter, int iter, hipStream_t st)
{
    SNH_SUN_DISPATCH(C.N, k_sun_mid_finalize, dim3(nblocks), dim3(kBlock), 0, st, C, Z, Za, Zb, conv_iter, iter)
}

template <typename T>
void launch_sun_grad(const Couplings<T>& C, int nblocks, const T* Z, const T* phi, T* HZ, hipStream_t st)
{
    SNH_SUN_DISPATCH(C.N, k_sun_grad, dim3(nblocks), dim3(kBlock), 0, st, C, Z, phi, HZ)
}

template <typename T>
void launch_sun_energy(const Couplings<T>& C, int nblocks, const T* Z, const T* phi, double* partial, hipStream_t st)
{
    SNH_SUN_DISPATCH(C.N, k_sun_energy, dim3(nblocks), dim3(kBlock), 0, st, C, Z, phi, partial)
}

template <typename T>
void launch_sun_mc_color(const Couplings<T>& C, const McArgs& A, T* Z, const McWork* work, int nwork,
                         unsigned long long P, double prob, double* acc, hipStream_t st)
{
    if (nwork <= 0) return;
    SNH_SUN_DISPATCH(C.N, k_sun_mc_color, dim3(nwork), dim3(kBlock), 0, st, C, A, Z, work, P, prob, acc)
}

template <typename T>
void launch_sun_mc_ewald_chunk(const Couplings<T>& C, const McArgs& A, T* Z, const McChunkIO& io, const double* Aew,
                               const double* A0, unsigned long long T0, int ntrial, double* acc, int r0, int nr,
                               hipStream_t st)
{
    SNH_SUN_DISPATCH(C.N, k_sun_mc_ewald_chunk, dim3(nr * (1 + io.napp)), dim3(64), 0, st, C, A, Z, io, Aew, A0, T0,
                     ntrial, acc, r0, nr)
}

#define SNH_SUN_INST(T)                                                                                          \
    template void launch_sun_mc_ewald_chunk<T>(const Couplings<T>&, const McArgs&, T*, const McChunkIO&,        \
                                               const double*, const double*, unsigned long long, int, double*,   \
                                               int, int, hipStream_t);                                           \
    template void launch_sun_mc_color<T>(const Couplings<T>&, const McArgs&, T*, const McWork*, int,            \
                                         unsigned long long, double, double*, hipStream_t);                      \
    template void launch_sun_expected_spin<T>(const Couplings<T>&, int, const T*, T*, hipStream_t);              \
    template void launch_sun_predict<T>(const Couplings<T>&, int, const T*, T*, T*, const T*,                    \
                                        const NoiseArgs<T>&, T, T, const T*, T*, hipStream_t);                   \
    template void launch_sun_correct<T>(const Couplings<T>&, int, const T*, const T*, T*, const T*, T, T,        \
                                        const T*, T*, hipStream_t);                                              \
    template void launch_sun_midpoint_bar<T>(const Couplings<T>&, int, const T*, const T*, T*, const int*,       \
                                             hipStream_t);                                                       \
    template void launch_sun_midpoint_iter<T>(const Couplings<T>&, int, const T*, const T*, T*, const T*, T, T,  \
                                              double*, const int*, hipStream_t);                                 \
    template void launch_sun_midpoint_finalize<T>(const Couplings<T>&, int, T*, const T*, const T*, const int*,  \
                                                  int, hipStream_t);                                             \
    template void launch_sun_grad<T>(const Couplings<T>&, int, const T*, const T*, T*, hipStream_t);             \
    template void launch_sun_energy<T>(const Couplings<T>&, int, const T*, const T*, double*, hipStream_t);

SNH_SUN_INST(float)
SNH_SUN_INST(double)

}  // namespace snh
