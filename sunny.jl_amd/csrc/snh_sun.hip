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
#include "snh_internal.h"

namespace snh {

template <typename T>
struct Cx {
    T re, im;
};

template <typename T>
__device__ inline Cx<T> cmul(Cx<T> a, Cx<T> b)
{
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

template <typename T, int N>
struct Ket {
    Cx<T> z[N];
};

template <typename T, int N>
__device__ inline Ket<T, N> load_ket(const T* __restrict__ P, long long total, long long g)
{
    Ket<T, N> k;
#pragma unroll
    for (int a = 0; a < N; a++) {
        k.z[a].re = P[(2 * a) * total + g];
        k.z[a].im = P[(2 * a + 1) * total + g];
    }
    return k;
}

template <typename T, int N>
__device__ inline void store_ket(T* __restrict__ P, long long total, long long g, const Ket<T, N>& k)
{
#pragma unroll
    for (int a = 0; a < N; a++) {
        P[(2 * a) * total + g] = k.z[a].re;
        P[(2 * a + 1) * total + g] = k.z[a].im;
    }
}

template <int N>
__device__ inline double spin_off(int k)  // k = 1..N-1
{
    const double s = (N - 1) / 2.0;
    return sqrt(2 * (s + 1) * k - (double)k * (k + 1)) / 2;
}

// expected_spin: (2 Re c, 2 Im c, sum m_k |Z_k|^2), c = sum conj(Z_k) off_k Z_{k+1}
template <typename T, int N>
__device__ inline void expected_spin(const Ket<T, N>& Z, T& sx, T& sy, T& sz)
{
    T cr = 0, ci = 0, nz = 0;
#pragma unroll
    for (int k = 0; k < N - 1; k++) {
        const T off = (T)spin_off<N>(k + 1);
        // conj(Zk) * off * Zk+1
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
__device__ inline void expected_quadrupole(const Ket<T, N>& Z, const T* __restrict__ Q, T out[5])
{
#pragma unroll
    for (int p = 0; p < 5; p++) {
        T acc = 0;
        for (int a = 0; a < N; a++)
            for (int b = 0; b < N; b++) {
                const T* q = Q + 2 * ((p * N + a) * N + b);
                Cx<T> qz = cmul(Cx<T>{q[0], q[1]}, Z.z[b]);
                acc += Z.z[a].re * qz.re + Z.z[a].im * qz.im;  // Re(conj(Za) qz)
            }
        out[p] = acc;
    }
}

// Source of kets for the gather: stored planes, or Zbar = (Z + Z1)/2.
template <typename T, int N, bool kMid>
struct KetSrc {
    const T* Z;
    const T* Z1;
    long long total;
    __device__ inline Ket<T, N> get(long long g) const
    {
        Ket<T, N> a = load_ket<T, N>(Z, total, g);
        if (!kMid) return a;
        Ket<T, N> b = load_ket<T, N>(Z1, total, g);
#pragma unroll
        for (int k = 0; k < N; k++) {
            a.z[k].re = (a.z[k].re + b.z[k].re) / 2;
            a.z[k].im = (a.z[k].im + b.z[k].im) / 2;
        }
        return a;
    }
};

struct SiteCoordS {
    int r, i, c, c1, c2, c3, dcls;
    long long g;
    ReplicaInfo R;
    bool valid;
};

__device__ inline SiteCoordS site_coord_s(const BlockInfo* blocks, const ReplicaInfo* reps)
{
    SiteCoordS sc;
    BlockInfo bi = blocks[blockIdx.x];
    sc.r = bi.r;
    sc.i = bi.i;
    sc.dcls = bi.dcls;
    sc.R = reps[bi.r];
    sc.c = bi.c0 + (int)threadIdx.x;
    sc.valid = sc.c < sc.R.ncell;
    int c = sc.valid ? sc.c : 0;
    sc.c1 = c % sc.R.L1;
    int t = c / sc.R.L1;
    sc.c2 = t % sc.R.L2;
    sc.c3 = t / sc.R.L2;
    sc.g = sc.R.off + c + (long long)sc.R.ncell * sc.i;
    return sc;
}

__device__ inline long long neighbor_site_s(const SiteCoordS& sc, const int* __restrict__ n, int j)
{
    int a = wrap_add(sc.c1, n[0], sc.R.L1);
    int b = wrap_add(sc.c2, n[1], sc.R.L2);
    int d = wrap_add(sc.c3, n[2], sc.R.L3);
    return sc.R.off + a + (long long)sc.R.L1 * (b + (long long)sc.R.L2 * d) + (long long)sc.R.ncell * j;
}

// H Z at one site (set_energy_grad_coherents!): dE/dS from Zeeman, bilinear
// exchange on expected dipoles and Ewald; then (Lambda + dE.S) Z plus
// biquadratic (dE/dQ . Q) Z.
template <typename T, int N, bool kMid>
__device__ inline Ket<T, N> local_hz(const Couplings<T>& C, const SiteCoordS& sc, const KetSrc<T, N, kMid>& src,
                                     const Ket<T, N>& Z, const T* __restrict__ ew)
{
    const int i = sc.i;
    T gx = C.zee[3 * i], gy = C.zee[3 * i + 1], gz = C.zee[3 * i + 2];
    T wq[5] = {0, 0, 0, 0, 0};
    bool anyq = false;
    const int* nred = C.bnred + (long long)sc.dcls * C.nb * 3;
    for (int b = C.bstart[i]; b < C.bstart[i + 1]; b++) {
        const int j = C.bj[b];
        const int fl = C.bflag[b];
        const Ket<T, N> Zj = src.get(neighbor_site_s(sc, nred + 3 * b, j));
        T sx, sy, sz;
        expected_spin<T, N>(Zj, sx, sy, sz);
        const T* J = C.bJ + 9 * b;
        if (fl & kBondIso) {
            gx += J[0] * sx;
            gy += J[0] * sy;
            gz += J[0] * sz;
        } else {
            gx += J[0] * sx + J[1] * sy + J[2] * sz;
            gy += J[3] * sx + J[4] * sy + J[5] * sz;
            gz += J[6] * sx + J[7] * sy + J[8] * sz;
        }
        if (fl & kBondBiquad) {
            T Qj[5];
            expected_quadrupole<T, N>(Zj, C.quad_ops, Qj);
            const T* Bq = C.bQ + 25 * b;
            for (int p = 0; p < 5; p++)
                for (int q = 0; q < 5; q++) wq[p] += Bq[5 * p + q] * Qj[q];
            anyq = true;
        }
    }
    if (ew) {  // Ewald potential phi of the expected dipoles: dE/dS -= g' phi
        const T* gm = C.gmat + 9 * i;
        const T px = ew[sc.g], py = ew[C.total + sc.g], pz = ew[2 * C.total + sc.g];
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
            Cx<T> p = cmul(Cx<T>{L[2 * (a * N + b)], L[2 * (a * N + b) + 1]}, Z.z[b]);
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
                Cx<T> v = cmul(Cx<T>{qr, qi}, Z.z[b]);
                H.z[a].re += v.re;
                H.z[a].im += v.im;
            }
    }
    return H;
}

// rhs_sun!: lambda == 0: dZ = -i dt HZ; else dZ = -proj(zeta + dt (i + lambda) HZ, Z).
template <typename T, int N>
__device__ inline Ket<T, N> rhs_sun(const Ket<T, N>& Z, const Ket<T, N>& zeta, const Ket<T, N>& H, T dt, T lambda)
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
    T nr = 0, ni = 0, nn = 0;
#pragma unroll
    for (int a = 0; a < N; a++) {
        // (i + lambda) H = (lambda Hr - Hi) + i (Hr + lambda Hi)
        v.z[a].re = zeta.z[a].re + dt * (lambda * H.z[a].re - H.z[a].im);
        v.z[a].im = zeta.z[a].im + dt * (H.z[a].re + lambda * H.z[a].im);
        nr += Z.z[a].re * v.z[a].re + Z.z[a].im * v.z[a].im;
        ni += Z.z[a].re * v.z[a].im - Z.z[a].im * v.z[a].re;
        nn += Z.z[a].re * Z.z[a].re + Z.z[a].im * Z.z[a].im;
    }
    const T fr = nr / nn, fi = ni / nn;
#pragma unroll
    for (int a = 0; a < N; a++) {
        d.z[a].re = -(v.z[a].re - (Z.z[a].re * fr - Z.z[a].im * fi));
        d.z[a].im = -(v.z[a].im - (Z.z[a].re * fi + Z.z[a].im * fr));
    }
    return d;
}

// normalize_ket: Z / sqrt(Z'Z / kappa), norm in FP64.
template <typename T, int N>
__device__ inline Ket<T, N> normalize_ket(Ket<T, N> Z, T kappa)
{
    if (kappa == (T)0) {
#pragma unroll
        for (int a = 0; a < N; a++) Z.z[a] = {0, 0};
        return Z;
    }
    double n2 = 0;
#pragma unroll
    for (int a = 0; a < N; a++) n2 += (double)Z.z[a].re * Z.z[a].re + (double)Z.z[a].im * Z.z[a].im;
    double f = 1.0 / sqrt(n2 / (double)kappa);
#pragma unroll
    for (int a = 0; a < N; a++) {
        Z.z[a].re = (T)(Z.z[a].re * f);
        Z.z[a].im = (T)(Z.z[a].im * f);
    }
    return Z;
}

template <typename T, int N>
__device__ inline Ket<T, N> site_noise_sun(const NoiseArgs<T>& nz, const SiteCoordS& sc)
{
    Ket<T, N> k;
    if (nz.mode == 0) {
#pragma unroll
        for (int a = 0; a < N; a++) k.z[a] = {0, 0};
        return k;
    }
    const T s = nz.scale[sc.r];
    if (nz.mode == 1) {
        const double* h = nz.host + 2 * N * sc.g;
#pragma unroll
        for (int a = 0; a < N; a++) k.z[a] = {(T)h[2 * a] * s, (T)h[2 * a + 1] * s};
        return k;
    }
    // randn(ComplexF64): each component has variance 1/2
    const T s2 = s * (T)0.70710678118654752440;
    const unsigned long long key = nz.seed + (unsigned long long)sc.r * 0x9E3779B97F4A7C15ull;
    const uint32_t site = (uint32_t)(sc.c + sc.R.ncell * sc.i);
#pragma unroll
    for (int blk = 0; blk < (2 * N + 3) / 4; blk++) {
        T n[4];
        philox_normals4<T>(key, nz.step, site, (uint32_t)blk, n);
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

__device__ inline double block_sum_s(double v)
{
    __shared__ double red[kBlock / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < kBlock / 64; k++) t += red[k];
    return t;
}

// ---------------------------------------------------------------------------
// Kernels

template <typename T, int N>
__global__ __launch_bounds__(kBlock) void k_sun_expected_spin(Couplings<T> C, const T* __restrict__ Z,
                                                              T* __restrict__ S)
{
    SiteCoordS sc = site_coord_s(C.blocks, C.reps);
    if (!sc.valid) return;
    T sx, sy, sz;
    expected_spin<T, N>(load_ket<T, N>(Z, C.total, sc.g), sx, sy, sz);
    S[sc.g] = sx;
    S[C.total + sc.g] = sy;
    S[2 * C.total + sc.g] = sz;
}

template <typename T, int N>
__global__ __launch_bounds__(kBlock) void k_sun_predict(Couplings<T> C, const T* __restrict__ Z, T* __restrict__ Zp,
                                                        const T* __restrict__ ew, NoiseArgs<T> nz, T dt, T lambda)
{
    SiteCoordS sc = site_coord_s(C.blocks, C.reps);
    if (!sc.valid) return;
    KetSrc<T, N, false> src{Z, nullptr, C.total};
    Ket<T, N> z = load_ket<T, N>(Z, C.total, sc.g);
    Ket<T, N> H = local_hz<T, N, false>(C, sc, src, z, ew);
    Ket<T, N> zeta = site_noise_sun<T, N>(nz, sc);
    Ket<T, N> d = rhs_sun<T, N>(z, zeta, H, dt, lambda);
#pragma unroll
    for (int a = 0; a < N; a++) {
        d.z[a].re += z.z[a].re;
        d.z[a].im += z.z[a].im;
    }
    store_ket<T, N>(Zp, C.total, sc.g, normalize_ket<T, N>(d, C.kappa[sc.i]));
}

template <typename T, int N>
__global__ __launch_bounds__(kBlock) void k_sun_correct(Couplings<T> C, const T* __restrict__ Z,
                                                        const T* __restrict__ Zp, T* __restrict__ Zout,
                                                        const T* __restrict__ ewZ, const T* __restrict__ ewZp,
                                                        NoiseArgs<T> nz, T dt, T lambda)
{
    SiteCoordS sc = site_coord_s(C.blocks, C.reps);
    if (!sc.valid) return;
    Ket<T, N> zeta = site_noise_sun<T, N>(nz, sc);
    KetSrc<T, N, false> s0{Z, nullptr, C.total};
    Ket<T, N> z = load_ket<T, N>(Z, C.total, sc.g);
    Ket<T, N> d1 = rhs_sun<T, N>(z, zeta, local_hz<T, N, false>(C, sc, s0, z, ewZ), dt, lambda);
    KetSrc<T, N, false> s1{Zp, nullptr, C.total};
    Ket<T, N> zp = load_ket<T, N>(Zp, C.total, sc.g);
    Ket<T, N> d2 = rhs_sun<T, N>(zp, zeta, local_hz<T, N, false>(C, sc, s1, zp, ewZp), dt, lambda);
#pragma unroll
    for (int a = 0; a < N; a++) {
        z.z[a].re = z.z[a].re + (d1.z[a].re + d2.z[a].re) / 2;
        z.z[a].im = z.z[a].im + (d1.z[a].im + d2.z[a].im) / 2;
    }
    store_ket<T, N>(Zout, C.total, sc.g, normalize_ket<T, N>(z, C.kappa[sc.i]));
}

template <typename T, int N>
__global__ __launch_bounds__(kBlock) void k_sun_mid_bar(Couplings<T> C, const T* __restrict__ Z,
                                                        const T* __restrict__ Z1, T* __restrict__ Zbar,
                                                        const int* __restrict__ conv_iter)
{
    SiteCoordS sc = site_coord_s(C.blocks, C.reps);
    if (!sc.valid || conv_iter[sc.r] != 0) return;
    KetSrc<T, N, true> src{Z, Z1, C.total};
    store_ket<T, N>(Zbar, C.total, sc.g, src.get(sc.g));
}

template <typename T, int N>
__global__ __launch_bounds__(kBlock) void k_sun_mid_iter(Couplings<T> C, const T* __restrict__ Z,
                                                         const T* __restrict__ Z1, T* __restrict__ Z2,
                                                         const T* __restrict__ ew, T dt, T lambda,
                                                         double* __restrict__ partial,
                                                         const int* __restrict__ conv_iter)
{
    SiteCoordS sc = site_coord_s(C.blocks, C.reps);
    if (conv_iter[sc.r] != 0) return;  // block-uniform
    double d2 = 0;
    if (sc.valid) {
        KetSrc<T, N, true> src{Z, Z1, C.total};
        Ket<T, N> zb = src.get(sc.g);
        Ket<T, N> zero;
#pragma unroll
        for (int a = 0; a < N; a++) zero.z[a] = {0, 0};
        Ket<T, N> d = rhs_sun<T, N>(zb, zero, local_hz<T, N, true>(C, sc, src, zb, ew), dt, lambda);
        Ket<T, N> z = load_ket<T, N>(Z, C.total, sc.g);
        Ket<T, N> z1 = load_ket<T, N>(Z1, C.total, sc.g);
#pragma unroll
        for (int a = 0; a < N; a++) {
            z.z[a].re += d.z[a].re;
            z.z[a].im += d.z[a].im;
            double er = (double)z1.z[a].re - (double)z.z[a].re, ei = (double)z1.z[a].im - (double)z.z[a].im;
            d2 += er * er + ei * ei;
        }
        store_ket<T, N>(Z2, C.total, sc.g, z);
    }
    double t = block_sum_s(d2);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

template <typename T, int N>
__global__ __launch_bounds__(kBlock) void k_sun_mid_finalize(Couplings<T> C, T* __restrict__ Z,
                                                             const T* __restrict__ Za, const T* __restrict__ Zb,
                                                             const int* __restrict__ conv_iter, int iter)
{
    SiteCoordS sc = site_coord_s(C.blocks, C.reps);
    int m = conv_iter[sc.r];
    if (!sc.valid || m != iter) return;
    const T* src = (m & 1) ? Za : Zb;
    store_ket<T, N>(Z, C.total, sc.g, normalize_ket<T, N>(load_ket<T, N>(src, C.total, sc.g), C.kappa[sc.i]));
}

template <typename T, int N>
__global__ __launch_bounds__(kBlock) void k_sun_grad(Couplings<T> C, const T* __restrict__ Z,
                                                     const T* __restrict__ ew, T* __restrict__ HZ)
{
    SiteCoordS sc = site_coord_s(C.blocks, C.reps);
    if (!sc.valid) return;
    KetSrc<T, N, false> src{Z, nullptr, C.total};
    Ket<T, N> z = load_ket<T, N>(Z, C.total, sc.g);
    store_ket<T, N>(HZ, C.total, sc.g, local_hz<T, N, false>(C, sc, src, z, ew));
}

template <typename T, int N>
__global__ __launch_bounds__(kBlock) void k_sun_energy(Couplings<T> C, const T* __restrict__ Z,
                                                       const T* __restrict__ ew, double* __restrict__ partial)
{
    SiteCoordS sc = site_coord_s(C.blocks, C.reps);
    double e = 0;
    if (sc.valid) {
        const int i = sc.i;
        Ket<T, N> z = load_ket<T, N>(Z, C.total, sc.g);
        T sx, sy, sz;
        expected_spin<T, N>(z, sx, sy, sz);
        const T* gm = C.gmat + 9 * i;
        const T* B = C.field + 3 * i;
        for (int a = 0; a < 3; a++)
            e += (double)B[a] * ((double)gm[3 * a] * sx + (double)gm[3 * a + 1] * sy + (double)gm[3 * a + 2] * sz);
        // onsite: Re(Z' Lambda Z)
        const T* L = C.lam + 2 * N * N * i;
        for (int a = 0; a < N; a++)
            for (int b = 0; b < N; b++) {
                Cx<T> p = cmul(Cx<T>{L[2 * (a * N + b)], L[2 * (a * N + b) + 1]}, z.z[b]);
                e += (double)z.z[a].re * p.re + (double)z.z[a].im * p.im;
            }
        const int* nred = C.bnred + (long long)sc.dcls * C.nb * 3;
        for (int b = C.bstart[i]; b < C.bstart[i + 1]; b++) {
            const int fl = C.bflag[b];
            if (fl & kBondCulled) continue;
            const int j = C.bj[b];
            Ket<T, N> zj = load_ket<T, N>(Z, C.total, neighbor_site_s(sc, nred + 3 * b, j));
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
            e += v;
        }
        if (ew) {  // 1/2 mu.phi with mu = -g <S>
            double mx = -((double)gm[0] * sx + (double)gm[1] * sy + (double)gm[2] * sz);
            double my = -((double)gm[3] * sx + (double)gm[4] * sy + (double)gm[5] * sz);
            double mz = -((double)gm[6] * sx + (double)gm[7] * sy + (double)gm[8] * sz);
            e += 0.5 * (mx * ew[sc.g] + my * ew[C.total + sc.g] + mz * ew[2 * C.total + sc.g]);
        }
    }
    double t = block_sum_s(e);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// ---------------------------------------------------------------------------
// Launchers: dispatch the run-time N to a compiled specialisation.

#define SNH_SUN_DISPATCH(N, KERNEL, ...)                                                            \
    switch (N) {                                                                                    \
    case 2: hipLaunchKernelGGL((KERNEL<T, 2>), __VA_ARGS__); break;                                 \
    case 3: hipLaunchKernelGGL((KERNEL<T, 3>), __VA_ARGS__); break;                                 \
    case 4: hipLaunchKernelGGL((KERNEL<T, 4>), __VA_ARGS__); break;                                 \
    case 5: hipLaunchKernelGGL((KERNEL<T, 5>), __VA_ARGS__); break;                                 \
    case 6: hipLaunchKernelGGL((KERNEL<T, 6>), __VA_ARGS__); break;                                 \
    case 7: hipLaunchKernelGGL((KERNEL<T, 7>), __VA_ARGS__); break;                                 \
    default: break;                                                                                 \
    }

template <typename T>
void launch_sun_expected_spin(const Couplings<T>& C, int nblocks, const T* Z, T* S, hipStream_t st)
{
    SNH_SUN_DISPATCH(C.N, k_sun_expected_spin, dim3(nblocks), dim3(kBlock), 0, st, C, Z, S)
}

template <typename T>
void launch_sun_predict(const Couplings<T>& C, int nblocks, const T* Z, T* Zp, const T* ew, const NoiseArgs<T>& nz,
                        T dt, T lambda, hipStream_t st)
{
    SNH_SUN_DISPATCH(C.N, k_sun_predict, dim3(nblocks), dim3(kBlock), 0, st, C, Z, Zp, ew, nz, dt, lambda)
}

template <typename T>
void launch_sun_correct(const Couplings<T>& C, int nblocks, const T* Z, const T* Zp, T* Zout, const T* ewZ,
                        const T* ewZp, const NoiseArgs<T>& nz, T dt, T lambda, hipStream_t st)
{
    SNH_SUN_DISPATCH(C.N, k_sun_correct, dim3(nblocks), dim3(kBlock), 0, st, C, Z, Zp, Zout, ewZ, ewZp, nz, dt,
                     lambda)
}

template <typename T>
void launch_sun_midpoint_bar(const Couplings<T>& C, int nblocks, const T* Z, const T* Z1, T* Zbar,
                             const int* conv_iter, hipStream_t st)
{
    SNH_SUN_DISPATCH(C.N, k_sun_mid_bar, dim3(nblocks), dim3(kBlock), 0, st, C, Z, Z1, Zbar, conv_iter)
}

template <typename T>
void launch_sun_midpoint_iter(const Couplings<T>& C, int nblocks, const T* Z, const T* Z1, T* Z2, const T* ew, T dt,
                              T lambda, double* partial, const int* conv_iter, int iter, hipStream_t st)
{
    (void)iter;
    SNH_SUN_DISPATCH(C.N, k_sun_mid_iter, dim3(nblocks), dim3(kBlock), 0, st, C, Z, Z1, Z2, ew, dt, lambda, partial,
                     conv_iter)
}

template <typename T>
void launch_sun_midpoint_finalize(const Couplings<T>& C, int nblocks, T* Z, const T* Za, const T* Zb,
                                  const int* conv_iter, int iter, hipStream_t st)
{
    SNH_SUN_DISPATCH(C.N, k_sun_mid_finalize, dim3(nblocks), dim3(kBlock), 0, st, C, Z, Za, Zb, conv_iter, iter)
}

template <typename T>
void launch_sun_grad(const Couplings<T>& C, int nblocks, const T* Z, const T* ew, T* HZ, hipStream_t st)
{
    SNH_SUN_DISPATCH(C.N, k_sun_grad, dim3(nblocks), dim3(kBlock), 0, st, C, Z, ew, HZ)
}

template <typename T>
void launch_sun_energy(const Couplings<T>& C, int nblocks, const T* Z, const T* ew, double* partial, hipStream_t st)
{
    SNH_SUN_DISPATCH(C.N, k_sun_energy, dim3(nblocks), dim3(kBlock), 0, st, C, Z, ew, partial)
}

#define SNH_SUN_INST(T)                                                                                           \
    template void launch_sun_expected_spin<T>(const Couplings<T>&, int, const T*, T*, hipStream_t);               \
    template void launch_sun_predict<T>(const Couplings<T>&, int, const T*, T*, const T*, const NoiseArgs<T>&, T, \
                                        T, hipStream_t);                                                          \
    template void launch_sun_correct<T>(const Couplings<T>&, int, const T*, const T*, T*, const T*, const T*,     \
                                        const NoiseArgs<T>&, T, T, hipStream_t);                                  \
    template void launch_sun_midpoint_bar<T>(const Couplings<T>&, int, const T*, const T*, T*, const int*,        \
                                             hipStream_t);                                                        \
    template void launch_sun_midpoint_iter<T>(const Couplings<T>&, int, const T*, const T*, T*, const T*, T, T,   \
                                              double*, const int*, int, hipStream_t);                             \
    template void launch_sun_midpoint_finalize<T>(const Couplings<T>&, int, T*, const T*, const T*, const int*,   \
                                                  int, hipStream_t);                                                   \
    template void launch_sun_grad<T>(const Couplings<T>&, int, const T*, const T*, T*, hipStream_t);              \
    template void launch_sun_energy<T>(const Couplings<T>&, int, const T*, const T*, double*, hipStream_t);

SNH_SUN_INST(float)
SNH_SUN_INST(double)

}  // namespace snh
