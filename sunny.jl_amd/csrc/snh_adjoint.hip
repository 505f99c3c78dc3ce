// Reverse (adjoint) sweep of the deterministic spherical-midpoint trajectory
// in :dipole mode, for gradients of a loss of the final state with respect
// to coupling parameters and to the initial state.
//
// Forward step (step!(::System{0}, ::ImplicitMidpoint) with damping 0,
// src/Integrators.jl:379-413), at its converged fixed point S'':
//     m = (S + S'')/2,  Shat = kappa n(m),  g = dE/dS(Shat),
//     S'' = S + rhs(Shat),  rhs(Shat) = -dt Shat x g,
//     S_next = kappa n(S''),                       n(x) = x/|x|.
// Implicit differentiation of S'' - S - rhs(Shat((S + S'')/2)) = 0 with
// A = d rhs/d S'' = d rhs/d S = (1/2) R N(m), R = d rhs/d Shat and
// N(x) = (kappa/|x|)(I - x x'/|x|^2), gives for an adjoint lambda' of S_next:
//     mu = N(S'') lambda',   nu = mu + A' nu  (solved by fixed-point sweeps),
//     lambda = (I + A') nu = 2 nu - mu,
//     dL/dtheta_p += sum_i w_i . G_p,i(Shat),   w = -dt nu x Shat,
// where G_p is the field of parameter p's unit couplings (the couplings are
// linear in the parameters, ModelParam, src/System/Types.jl:76-86).  With
// field g = M Shat + b (bilinear exchange, kmax <= 2 anisotropy, Zeeman),
//     R' v = -dt g x v + M w(v),   w(v) = -dt v x Shat,
// and M w is the same bond gather as the field (M is symmetric: the reverse
// bond carries J').  Sunny has no reverse rule for step!; this is the
// discrete adjoint of the reference's own update.
#include "snh_gather.h"

namespace snh {

// FP64 forward-mode dual number: value and one directional derivative.
struct Dual {
    double v, d;
    __device__ Dual(double a = 0, double b = 0) : v(a), d(b) {}
};
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator-(Dual a) { return {-a.v, -a.d}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
__device__ __forceinline__ Dual operator*(Dual a, double b) { return {a.v * b, a.d * b}; }
__device__ __forceinline__ Dual& operator+=(Dual& a, Dual b)
{
    a = a + b;
    return a;
}

// The directional derivative, along w, of the field terms that are not
// linear in the state, at Shat: onsite Stevens anisotropy of order 4 and 6
// (by dual numbers through stevens_grad) and biquadratic bonds,
//   g_i += sum_p v_p grad Q_p(S_i),  v = B Q(S_j),
//   d g_i = sum_p dv_p grad Q_p(S_i) + v_p grad Q_p(w_i),  dv = B dQ(S_j)[w_j]
// (grad Q_p is linear).  w is tangent to every Shat_i (w = -dt nu x Shat), so
// holding X = |S|^2 constant, as the Stevens gradient does, changes nothing.
__device__ __forceinline__ V3<double> adj_nonlinear_jvp(const Couplings<double>& C, const SiteCoord& sc,
                                                        const double* __restrict__ Shat,
                                                        const double* __restrict__ w)
{
    V3<double> r = {0, 0, 0};
    const int i = sc.i;
    const V3<double> s = planes(Shat, C.total)(sc.g), wi = planes(w, C.total)(sc.g);
    const int km = C.kmax[i];
    if (km > 2) {
        Dual gx = 0, gy = 0, gz = 0;
        stevens_grad<Dual, double>(Dual(s.x, wi.x), Dual(s.y, wi.y), Dual(s.z, wi.z), C.stev + 28 * i, km, gx, gy,
                                   gz, nullptr);
        r = {gx.d, gy.d, gz.d};
    }
    const int b0 = C.bstart[i], b1 = C.bstart[i + 1];
    const int* __restrict__ dl = C.bdelta + sc.R.dcls * C.nb;
    for (int b = b0; b < b1; b++) {
        if (!(C.bflag[b] & kBondBiquad)) continue;
        const uint32_t gj = sc.g + (uint32_t)dl[b];
        const V3<double> sj = planes(Shat, C.total)(gj), wj = planes(w, C.total)(gj);
        double Qj[5];
        quadrupole(sj, Qj);
        const double dQj[5] = {2 * (sj.x * wj.x - sj.y * wj.y), sj.z * wj.x + sj.x * wj.z,
                               -2 * sj.x * wj.x - 2 * sj.y * wj.y + 4 * sj.z * wj.z, sj.z * wj.y + sj.y * wj.z,
                               2 * (sj.y * wj.x + sj.x * wj.y)};
        const double* B = C.bQ + 25 * b;
        double v[5], dv[5];
        for (int p = 0; p < 5; p++) {
            double a = 0, da = 0;
            for (int q = 0; q < 5; q++) {
                a += B[5 * p + q] * Qj[q];
                da += B[5 * p + q] * dQj[q];
            }
            v[p] = a;
            dv[p] = da;
        }
        add_grad_quadrupole(s, dv, r);
        add_grad_quadrupole(wi, v, r);
    }
    return r;
}

// (kappa/|x|) (v - xh (xh . v)), xh = x/|x|: the Jacobian of kappa n(x)
// (symmetric) applied to v, in FP64.
template <typename T>
__device__ __forceinline__ V3<T> norm_jacobian(V3<T> x, V3<T> v, T kappa)
{
    if (kappa == (T)0) return {0, 0, 0};
    const double xx = x.x, xy = x.y, xz = x.z;
    const double n = sqrt(xx * xx + xy * xy + xz * xz);
    const double hx = xx / n, hy = xy / n, hz = xz / n;
    const double d = hx * v.x + hy * v.y + hz * v.z;
    const double f = (double)kappa / n;
    return {(T)(f * (v.x - hx * d)), (T)(f * (v.y - hy * d)), (T)(f * (v.z - hz * d))};
}

// Shat (with ghost images) and mu = N(S'') lambda'.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_adj_prep(Couplings<T> C, const T* __restrict__ S,
                                                     const T* __restrict__ Sdd, const T* __restrict__ lam,
                                                     T* __restrict__ Shat, T* __restrict__ mu)
{
    const SiteCoord sc = site_coord(C.blocks, C.reps, C.G);
    if (!sc.valid) return;
    const V3<T> s = planes(S, C.total)(sc.g), x = planes(Sdd, C.total)(sc.g), l = planes(lam, C.total)(sc.g);
    const T kap = C.kappa[sc.i];
    store3_images(Shat, C.total, sc, C.G,
                  normalize_dipole(V3<T>{(s.x + x.x) / 2, (s.y + x.y) / 2, (s.z + x.z) / 2}, kap));
    store3(mu, C.total, sc.g, norm_jacobian(x, l, kap));
}

// w = -dt nu x Shat (with ghost images: gathered by the next sweep).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_adj_w(Couplings<T> C, const T* __restrict__ nu,
                                                  const T* __restrict__ Shat, T dt, T* __restrict__ w)
{
    const SiteCoord sc = site_coord(C.blocks, C.reps, C.G);
    if (!sc.valid) return;
    const V3<T> c = cross3(planes(nu, C.total)(sc.g), planes(Shat, C.total)(sc.g));
    store3_images(w, C.total, sc, C.G, V3<T>{-dt * c.x, -dt * c.y, -dt * c.z});
}

// One sweep nu' = mu + (1/2) N(m) (-dt g x nu + M w), with the FP64 sum of
// |nu' - nu|^2 per block, where M w is the derivative of the field at Shat
// along w.  Clin is the coupling set without the Zeeman term and without the
// nonlinear terms (biquadratic bonds, Stevens order 4 and 6), so its field of
// w is the linear part of M w; phi_w (may be null) is the Ewald potential of
// w, whose field -g' phi_w is the Ewald part (the Ewald field is linear and
// symmetric in S).  Cnl non-null (the full couplings): the nonlinear part at
// Shat is added (adj_nonlinear_jvp).
template <typename T, int CH>
__global__ __launch_bounds__(kBlock) void k_adj_apply(Couplings<T> Clin, const T* __restrict__ w,
                                                      const T* __restrict__ S, const T* __restrict__ Sdd,
                                                      const T* __restrict__ g, const T* __restrict__ nu,
                                                      const T* __restrict__ mu, T dt, T* __restrict__ nu_out,
                                                      double* __restrict__ partial, const T* __restrict__ phi_w,
                                                      Couplings<T> Cnl, const T* __restrict__ Shat)
{
    const SiteCoord sc = site_coord(Clin.blocks, Clin.reps, Clin.G);
    double r2 = 0;
    if (sc.valid) {
        const uint32_t tot = Clin.total;
        const Planes<T> Wv = planes(w, tot);
        const SpinSrc<T, false> src{Wv, Wv};
        V3<T> hw = local_field<T, CH, false>(Clin, sc, src, Wv(sc.g), phi_w);
        if (Shat) {
            const V3<double> n = adj_nonlinear_jvp(Cnl, sc, Shat, w);
            hw = {hw.x + n.x, hw.y + n.y, hw.z + n.z};
        }
        const V3<T> gv = planes(g, tot)(sc.g), nv = planes(nu, tot)(sc.g);
        const V3<T> gx = cross3(gv, nv);
        const V3<T> t = {hw.x - dt * gx.x, hw.y - dt * gx.y, hw.z - dt * gx.z};
        const V3<T> s = planes(S, tot)(sc.g), x = planes(Sdd, tot)(sc.g);
        const V3<T> a = norm_jacobian(V3<T>{(s.x + x.x) / 2, (s.y + x.y) / 2, (s.z + x.z) / 2}, t,
                                      Clin.kappa[sc.i]);
        const V3<T> m = planes(mu, tot)(sc.g);
        const V3<T> o = {m.x + a.x / 2, m.y + a.y / 2, m.z + a.z / 2};
        store3(nu_out, tot, sc.g, o);
        const double ex = (double)o.x - nv.x, ey = (double)o.y - nv.y, ez = (double)o.z - nv.z;
        r2 = ex * ex + ey * ey + ez * ez;
    }
    const double t = block_sum(r2);
    if (threadIdx.x == 0) partial[sc.blk] = t;
}

// Per block: sum_i w_i . G_i(Shat) for the unit couplings Cp of one parameter.
template <typename T, int CH>
__global__ __launch_bounds__(kBlock) void k_adj_param(Couplings<T> Cp, const T* __restrict__ Shat,
                                                      const T* __restrict__ w, double* __restrict__ partial)
{
    const SiteCoord sc = site_coord(Cp.blocks, Cp.reps, Cp.G);
    double d = 0;
    if (sc.valid) {
        const Planes<T> Hv = planes(Shat, Cp.total);
        const SpinSrc<T, false> src{Hv, Hv};
        const V3<T> G = local_field<T, CH, false>(Cp, sc, src, Hv(sc.g), nullptr);
        const V3<T> wv = planes(w, Cp.total)(sc.g);
        d = (double)G.x * wv.x + (double)G.y * wv.y + (double)G.z * wv.z;
    }
    const double t = block_sum(d);
    if (threadIdx.x == 0) partial[sc.blk] = t;
}

// lambda = 2 nu - mu.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_adj_lambda(Couplings<T> C, const T* __restrict__ nu,
                                                       const T* __restrict__ mu, T* __restrict__ lam)
{
    const SiteCoord sc = site_coord(C.blocks, C.reps, C.G);
    if (!sc.valid) return;
    const V3<T> n = planes(nu, C.total)(sc.g), m = planes(mu, C.total)(sc.g);
    store3(lam, C.total, sc.g, V3<T>{2 * n.x - m.x, 2 * n.y - m.y, 2 * n.z - m.z});
}

// out[slot] += sum of partial[0..n), in one block.
__global__ __launch_bounds__(kBlock) void k_adj_accumulate(const double* __restrict__ partial, int n,
                                                           double* __restrict__ out, int slot)
{
    double v = 0;
    for (int k = threadIdx.x; k < n; k += kBlock) v += partial[k];
    const double t = block_sum(v);
    if (threadIdx.x == 0) out[slot] += t;
}

// ---------------------------------------------------------------------------
// Launchers

template <typename T>
void launch_adj_prep(const Couplings<T>& C, int nblocks, const T* S, const T* Sdd, const T* lam, T* Shat, T* mu,
                     hipStream_t st)
{
    hipLaunchKernelGGL(k_adj_prep<T>, dim3(nblocks), dim3(kBlock), 0, st, C, S, Sdd, lam, Shat, mu);
}

template <typename T>
void launch_adj_w(const Couplings<T>& C, int nblocks, const T* nu, const T* Shat, T dt, T* w, hipStream_t st)
{
    hipLaunchKernelGGL(k_adj_w<T>, dim3(nblocks), dim3(kBlock), 0, st, C, nu, Shat, dt, w);
}

template <typename T>
void launch_adj_apply(const Couplings<T>& Clin, int nblocks, const T* w, const T* S, const T* Sdd, const T* g,
                      const T* nu, const T* mu, T dt, T* nu_out, double* partial, const T* phi_w,
                      const Couplings<T>& Cnl, const T* Shat, hipStream_t st)
{
#define SNH_AA(CH)                                                                                             \
    hipLaunchKernelGGL((k_adj_apply<T, CH>), dim3(nblocks), dim3(kBlock), 0, st, Clin, w, S, Sdd, g, nu, mu, dt, \
                       nu_out, partial, phi_w, Cnl, Shat)
    if (Clin.chunk == 18) SNH_AA(18);
    else if (Clin.chunk == 9) SNH_AA(9);
    else SNH_AA(6);
#undef SNH_AA
}

template <typename T>
void launch_adj_param(const Couplings<T>& Cp, int nblocks, const T* Shat, const T* w, double* partial,
                      hipStream_t st)
{
#define SNH_AP(CH) \
    hipLaunchKernelGGL((k_adj_param<T, CH>), dim3(nblocks), dim3(kBlock), 0, st, Cp, Shat, w, partial)
    if (Cp.chunk == 18) SNH_AP(18);
    else if (Cp.chunk == 9) SNH_AP(9);
    else SNH_AP(6);
#undef SNH_AP
}

template <typename T>
void launch_adj_lambda(const Couplings<T>& C, int nblocks, const T* nu, const T* mu, T* lam, hipStream_t st)
{
    hipLaunchKernelGGL(k_adj_lambda<T>, dim3(nblocks), dim3(kBlock), 0, st, C, nu, mu, lam);
}

void launch_adj_accumulate(const double* partial, int n, double* out, int slot, hipStream_t st)
{
    hipLaunchKernelGGL(k_adj_accumulate, dim3(1), dim3(kBlock), 0, st, partial, n, out, slot);
}

#define SNH_ADJ_INST(T)                                                                                          \
    template void launch_adj_prep<T>(const Couplings<T>&, int, const T*, const T*, const T*, T*, T*, hipStream_t); \
    template void launch_adj_w<T>(const Couplings<T>&, int, const T*, const T*, T, T*, hipStream_t);              \
    template void launch_adj_apply<T>(const Couplings<T>&, int, const T*, const T*, const T*, const T*, const T*,  \
                                      const T*, T, T*, double*, const T*, const Couplings<T>&, const T*,         \
                                      hipStream_t);                                                              \
    template void launch_adj_param<T>(const Couplings<T>&, int, const T*, const T*, double*, hipStream_t);        \
    template void launch_adj_lambda<T>(const Couplings<T>&, int, const T*, const T*, T*, hipStream_t);

SNH_ADJ_INST(double)

}  // namespace snh
