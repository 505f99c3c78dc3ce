// Per-site vector algebra of the :dipole kernels, shared by the ghost-padded
// gather kernels (snh_dipole.hip) and the LDS-tiled kernels (snh_tile.hip).
//
// Reference semantics (Sunny.jl 0.9.2):
//   rhs_dipole!          src/Integrators.jl:266-275
//   fill_noise!          src/Integrators.jl:288-298
//   normalize_dipole     src/System/System.jl:421-423
//   quadrupole, grad     src/Operators/Spin.jl:48-69
#pragma once
#include "snh_internal.h"

namespace snh {

template <typename T>
struct V3 {
    T x, y, z;
};

template <typename T>
__device__ __forceinline__ V3<T> cross3(V3<T> a, V3<T> b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// kappa * s/|s| (normalize_dipole, System.jl:421-423).  FP64 state:
// correctly rounded sqrt and divide.  FP32 state: FP32 throughout, with the
// hardware reciprocal square root (v_rsq_f32, 1 ulp): |result| is within
// about 2.5 ulps of kappa.
__device__ __forceinline__ V3<double> normalize_dipole(V3<double> s, double kappa)
{
    if (kappa == 0.0) return {0, 0, 0};
    const double f = kappa / sqrt(s.x * s.x + s.y * s.y + s.z * s.z);
    return {s.x * f, s.y * f, s.z * f};
}

__device__ __forceinline__ V3<float> normalize_dipole(V3<float> s, float kappa)
{
    if (kappa == 0.0f) return {0, 0, 0};
    const float f = kappa * __builtin_amdgcn_rsqf(s.x * s.x + s.y * s.y + s.z * s.z);
    return {s.x * f, s.y * f, s.z * f};
}

// normalize_dipole together with mu = |s| / kappa, the factor with s = mu *
// result (0 for kappa = 0): the midpoint sweep keeps mu beside the next
// Shat instead of the iterate S'' = 2 mu Shat - S.
__device__ __forceinline__ V3<double> normalize_dipole_mu(V3<double> s, double kappa, double& mu)
{
    if (kappa == 0.0) {
        mu = 0;
        return {0, 0, 0};
    }
    const double n = sqrt(s.x * s.x + s.y * s.y + s.z * s.z);
    const double f = kappa / n;
    mu = n / kappa;
    return {s.x * f, s.y * f, s.z * f};
}

__device__ __forceinline__ V3<float> normalize_dipole_mu(V3<float> s, float kappa, float& mu)
{
    if (kappa == 0.0f) {
        mu = 0;
        return {0, 0, 0};
    }
    const float f = kappa * __builtin_amdgcn_rsqf(s.x * s.x + s.y * s.y + s.z * s.z);
    mu = __builtin_amdgcn_rcpf(f);
    return {s.x * f, s.y * f, s.z * f};
}

// Stevens quadrupoles of a classical dipole and their gradient, Spin.jl:48-69.
template <typename T>
__device__ __forceinline__ void quadrupole(V3<T> s, T Q[5])
{
    Q[0] = s.x * s.x - s.y * s.y;
    Q[1] = s.z * s.x;
    Q[2] = -s.x * s.x - s.y * s.y + 2 * s.z * s.z;
    Q[3] = s.z * s.y;
    Q[4] = 2 * s.y * s.x;
}

// g += sum_p w[p] * grad Q_p(s)
template <typename T>
__device__ __forceinline__ void add_grad_quadrupole(V3<T> s, const T w[5], V3<T>& g)
{
    g.x += w[0] * 2 * s.x + w[1] * s.z - w[2] * 2 * s.x + w[4] * 2 * s.y;
    g.y += -w[0] * 2 * s.y - w[2] * 2 * s.y + w[3] * s.z + w[4] * 2 * s.x;
    g.z += w[1] * s.x + w[2] * 4 * s.z + w[3] * s.y;
}

// g += J sj for one bond (isotropic: J = J0 * I), plus the biquadratic term.
template <typename T>
__device__ __forceinline__ void add_bond_field(const Couplings<T>& C, int b, V3<T> s, V3<T> sj, V3<T>& g)
{
    const int fl = ldc(C.bflag, b);
    if (fl & kBondIso) {
        const T J0 = ldc(C.bJ0, b);
        g.x += J0 * sj.x;
        g.y += J0 * sj.y;
        g.z += J0 * sj.z;
    } else {
        const T* J = C.bJ + 9 * b;
        g.x += ldc(J, 0) * sj.x + ldc(J, 1) * sj.y + ldc(J, 2) * sj.z;
        g.y += ldc(J, 3) * sj.x + ldc(J, 4) * sj.y + ldc(J, 5) * sj.z;
        g.z += ldc(J, 6) * sj.x + ldc(J, 7) * sj.y + ldc(J, 8) * sj.z;
    }
    if (fl & kBondBiquad) {
        T Qj[5], w[5];
        quadrupole(sj, Qj);
        const T* B = C.bQ + 25 * b;
#pragma unroll
        for (int p = 0; p < 5; p++) {
            T acc = 0;
#pragma unroll
            for (int q = 0; q < 5; q++) acc += ldc(B, 5 * p + q) * Qj[q];
            w[p] = acc;
        }
        add_grad_quadrupole(s, w, g);
    }
}

// Onsite gradient: Stevens anisotropy.  At kmax = 2 the classical
// polynomial's gradient (with |S|^2 held constant, OnsiteCoupling.jl:
// 316-425) is the linear map 2 Q2 S, Q2 prepared on the host.
template <typename T>
__device__ __forceinline__ void add_onsite_field(const Couplings<T>& C, int i, V3<T> s, V3<T>& g)
{
    const int km = ldc(C.kmax, i);
    if (km == 2) {
        const T* q = C.stev_q2 + 9 * i;
        g.x += ldc(q, 0) * s.x + ldc(q, 1) * s.y + ldc(q, 2) * s.z;
        g.y += ldc(q, 3) * s.x + ldc(q, 4) * s.y + ldc(q, 5) * s.z;
        g.z += ldc(q, 6) * s.x + ldc(q, 7) * s.y + ldc(q, 8) * s.z;
    } else if (km > 2) {
        stevens_grad<T>(s.x, s.y, s.z, C.stev + 28 * i, km, g.x, g.y, g.z, nullptr);
    }
}

// Ewald: dE/dS -= g^T phi (accum_ewald_grad!, Ewald.jl:185-209).
template <typename T>
__device__ __forceinline__ void add_ewald_field(const Couplings<T>& C, int i, V3<T> p, V3<T>& g)
{
    const T* gm = C.gmat + 9 * i;
    g.x -= ldc(gm, 0) * p.x + ldc(gm, 3) * p.y + ldc(gm, 6) * p.z;
    g.y -= ldc(gm, 1) * p.x + ldc(gm, 4) * p.y + ldc(gm, 7) * p.z;
    g.z -= ldc(gm, 2) * p.x + ldc(gm, 5) * p.y + ldc(gm, 8) * p.z;
}

// Deterministic part of rhs_dipole!: -S x (dt g - dt lambda S x g).
template <typename T>
__device__ __forceinline__ V3<T> rhs_det(V3<T> s, V3<T> g, T dt, T lambda)
{
    V3<T> v;
    if (lambda == (T)0) {
        v = {dt * g.x, dt * g.y, dt * g.z};
    } else {
        const V3<T> sg = cross3(s, g);
        v = {dt * g.x - dt * lambda * sg.x, dt * g.y - dt * lambda * sg.y, dt * g.z - dt * lambda * sg.z};
    }
    const V3<T> o = cross3(s, v);
    return {-o.x, -o.y, -o.z};
}

// fill_noise! for one site: xi = sqrt(2 dt lambda kT) * randn.  `site` is
// the site's index within its replica in Julia order (cell + ncell * i),
// `g` its index in the device planes (host-supplied draws are stored there).
template <typename T>
__device__ __forceinline__ V3<T> dipole_noise(const NoiseArgs<T>& nz, int r, uint32_t site, uint32_t g)
{
    if (nz.mode == 0) return {0, 0, 0};
    const T s = ldc(nz.scale, r);
    if (nz.mode == 1) {
        const double* h = nz.host + 3 * (size_t)g;
        return {(T)h[0] * s, (T)h[1] * s, (T)h[2] * s};
    }
    T n[4];
    const unsigned long long key = nz.seed + (unsigned long long)r * 0x9E3779B97F4A7C15ull;
    philox_normals4<T>(key, noise_step(nz), site, 0u, n);
    return {n[0] * s, n[1] * s, n[2] * s};
}

// dipole_noise without host-supplied draws and with the step number passed
// as a kernel argument (modes 0 and 2 only).
template <typename T>
__device__ __forceinline__ V3<T> dipole_noise_philox(const NoiseArgs<T>& nz, int r, uint32_t site)
{
    if (nz.mode == 0) return {0, 0, 0};
    const T s = ldc(nz.scale, r);
    T n[4];
    const unsigned long long key = nz.seed + (unsigned long long)r * 0x9E3779B97F4A7C15ull;
    philox_normals4<T>(key, nz.step, site, 0u, n);
    return {n[0] * s, n[1] * s, n[2] * s};
}

// Heun predictor algebra for one site with field g and noise xi:
//   dS1 = -S x xi + rhs_det(S, g);  S' = kappa*normalize(S + dS1);
//   H = S + dS1/2 - (S' x xi)/2  (the corrector's noise term, folded in).
template <typename T>
__device__ __forceinline__ void heun_predict(V3<T> s, V3<T> g, V3<T> xi, bool noisy, T kappa, T dt, T lambda,
                                             V3<T>& sp, V3<T>& h)
{
    V3<T> d = rhs_det(s, g, dt, lambda);
    if (noisy) {
        const V3<T> sx = cross3(s, xi);
        d = {d.x - sx.x, d.y - sx.y, d.z - sx.z};
    }
    sp = normalize_dipole(V3<T>{s.x + d.x, s.y + d.y, s.z + d.z}, kappa);
    if (noisy) {
        const V3<T> spx = cross3(sp, xi);
        h = {s.x + d.x / 2 - spx.x / 2, s.y + d.y / 2 - spx.y / 2, s.z + d.z / 2 - spx.z / 2};
    } else {
        h = {s.x + d.x / 2, s.y + d.y / 2, s.z + d.z / 2};
    }
}

}  // namespace snh
