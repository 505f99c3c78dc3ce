// :dipole-mode kernels: local field by bond-stencil gather, fused stochastic
// Heun stages, spherical-midpoint sweeps, FP64 energy.
//
// Reference semantics (Sunny.jl 0.9.2):
//   set_energy_grad_dipoles!  src/System/Interactions.jl:516-592
//   rhs_dipole!               src/Integrators.jl:266-275
//   step!(::System{0}, ::Langevin)          src/Integrators.jl:313-332
//   step!(::System{0}, ::ImplicitMidpoint)  src/Integrators.jl:379-413
//   energy / energy_aux       src/System/Interactions.jl:401-508
//
// The reference scatters each non-culled bond into both ends (J*Sj to i,
// J'*Si to j).  Here every directed bond (i -> j, J) and its reverse
// (j -> i, J') is in the table, and each site gathers over its own bonds: no
// write conflicts, one coalesced read per neighbour plane.
#include "snh_internal.h"

namespace snh {

template <typename T>
struct Vec3T {
    T x, y, z;
};

template <typename T>
__device__ inline Vec3T<T> load3(const T* __restrict__ P, long long total, long long g)
{
    return {P[g], P[total + g], P[2 * total + g]};
}

template <typename T>
__device__ inline void store3(T* __restrict__ P, long long total, long long g, Vec3T<T> v)
{
    P[g] = v.x;
    P[total + g] = v.y;
    P[2 * total + g] = v.z;
}

template <typename T>
__device__ inline Vec3T<T> cross3(Vec3T<T> a, Vec3T<T> b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// kappa * s/|s| with the norm in FP64 (normalize_dipole, System.jl:421-423).
template <typename T>
__device__ inline Vec3T<T> normalize_dipole(Vec3T<T> s, T kappa)
{
    if (kappa == (T)0) return {0, 0, 0};
    double x = s.x, y = s.y, z = s.z;
    double f = (double)kappa / sqrt(x * x + y * y + z * z);
    return {(T)(x * f), (T)(y * f), (T)(z * f)};
}

// Stevens quadrupoles of a classical dipole and their gradient, Spin.jl:48-69.
template <typename T>
__device__ inline void quadrupole(Vec3T<T> s, T Q[5])
{
    Q[0] = s.x * s.x - s.y * s.y;
    Q[1] = s.z * s.x;
    Q[2] = -s.x * s.x - s.y * s.y + 2 * s.z * s.z;
    Q[3] = s.z * s.y;
    Q[4] = 2 * s.y * s.x;
}

// g += sum_p w[p] * grad Q_p(s)
template <typename T>
__device__ inline void add_grad_quadrupole(Vec3T<T> s, const T w[5], Vec3T<T>& g)
{
    g.x += w[0] * 2 * s.x + w[1] * s.z - w[2] * 2 * s.x + w[4] * 2 * s.y;
    g.y += -w[0] * 2 * s.y - w[2] * 2 * s.y + w[3] * s.z + w[4] * 2 * s.x;
    g.z += w[1] * s.x + w[2] * 4 * s.z + w[3] * s.y;
}

struct SiteCoord {
    int r, i, c, c1, c2, c3, dcls;
    long long g;
    ReplicaInfo R;
    bool valid;
};

__device__ inline SiteCoord site_coord(const BlockInfo* blocks, const ReplicaInfo* reps)
{
    SiteCoord sc;
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

__device__ inline long long neighbor_site(const SiteCoord& sc, const int* __restrict__ n, int j)
{
    int a = wrap_add(sc.c1, n[0], sc.R.L1);
    int b = wrap_add(sc.c2, n[1], sc.R.L2);
    int d = wrap_add(sc.c3, n[2], sc.R.L3);
    return sc.R.off + a + (long long)sc.R.L1 * (b + (long long)sc.R.L2 * d) + (long long)sc.R.ncell * j;
}

// Spin source for the gather: either a stored plane set, or the midpoint
// Shat = kappa * normalize((S + S1)/2) formed on the fly.
template <typename T, bool kMid>
struct SpinSrc {
    const T* S;
    const T* S1;
    long long total;
    __device__ inline Vec3T<T> get(long long g, T kappa) const
    {
        Vec3T<T> a = load3(S, total, g);
        if (!kMid) return a;
        Vec3T<T> b = load3(S1, total, g);
        return normalize_dipole(Vec3T<T>{(a.x + b.x) / 2, (a.y + b.y) / 2, (a.z + b.z) / 2}, kappa);
    }
};

// dE/dS at one site: Zeeman + Stevens anisotropy + bonds (+ Ewald from phi planes).
template <typename T, bool kMid>
__device__ inline Vec3T<T> local_field(const Couplings<T>& C, const SiteCoord& sc, const SpinSrc<T, kMid>& src,
                                       Vec3T<T> s, const T* __restrict__ ew)
{
    const int i = sc.i;
    Vec3T<T> g = {C.zee[3 * i], C.zee[3 * i + 1], C.zee[3 * i + 2]};
    stevens_grad<T>(s.x, s.y, s.z, C.stev + 28 * i, C.kmax[i], g.x, g.y, g.z, nullptr);
    const int b0 = C.bstart[i], b1 = C.bstart[i + 1];
    const int* nred = C.bnred + (long long)sc.dcls * C.nb * 3;
    for (int b = b0; b < b1; b++) {
        const int j = C.bj[b];
        const int fl = C.bflag[b];
        const long long gj = neighbor_site(sc, nred + 3 * b, j);
        const Vec3T<T> sj = src.get(gj, C.kappa[j]);
        const T* J = C.bJ + 9 * b;
        if (fl & kBondIso) {
            T J0 = J[0];
            g.x += J0 * sj.x;
            g.y += J0 * sj.y;
            g.z += J0 * sj.z;
        } else {
            g.x += J[0] * sj.x + J[1] * sj.y + J[2] * sj.z;
            g.y += J[3] * sj.x + J[4] * sj.y + J[5] * sj.z;
            g.z += J[6] * sj.x + J[7] * sj.y + J[8] * sj.z;
        }
        if (fl & kBondBiquad) {
            T Qj[5], w[5];
            quadrupole(sj, Qj);
            const T* B = C.bQ + 25 * b;
#pragma unroll
            for (int p = 0; p < 5; p++) {
                T acc = 0;
#pragma unroll
                for (int q = 0; q < 5; q++) acc += B[5 * p + q] * Qj[q];
                w[p] = acc;
            }
            add_grad_quadrupole(s, w, g);
        }
    }
    if (ew) {  // Ewald potential phi of the same configuration: dE/dS -= g' phi
        Vec3T<T> p = load3(ew, C.total, sc.g);
        const T* gm = C.gmat + 9 * i;
        g.x -= gm[0] * p.x + gm[3] * p.y + gm[6] * p.z;
        g.y -= gm[1] * p.x + gm[4] * p.y + gm[7] * p.z;
        g.z -= gm[2] * p.x + gm[5] * p.y + gm[8] * p.z;
    }
    return g;
}

// rhs_dipole!: dS = -S x (xi + dt*g - dt*lambda*(S x g)); lambda == 0: -S x (dt*g).
template <typename T>
__device__ inline Vec3T<T> rhs_dipole(Vec3T<T> s, Vec3T<T> xi, Vec3T<T> g, T dt, T lambda)
{
    Vec3T<T> v;
    if (lambda == (T)0) {
        v = {dt * g.x, dt * g.y, dt * g.z};
    } else {
        Vec3T<T> sg = cross3(s, g);
        v = {xi.x + dt * g.x - dt * lambda * sg.x, xi.y + dt * g.y - dt * lambda * sg.y,
             xi.z + dt * g.z - dt * lambda * sg.z};
    }
    Vec3T<T> o = cross3(s, v);
    return {-o.x, -o.y, -o.z};
}

// fill_noise! (Integrators.jl:288-298) for one site.
template <typename T>
__device__ inline Vec3T<T> site_noise(const NoiseArgs<T>& nz, const SiteCoord& sc)
{
    if (nz.mode == 0) return {0, 0, 0};
    T sc_r = nz.scale[sc.r];
    if (nz.mode == 1) {
        const double* h = nz.host + 3 * sc.g;
        return {(T)h[0] * sc_r, (T)h[1] * sc_r, (T)h[2] * sc_r};
    }
    T n[4];
    unsigned long long key = nz.seed + (unsigned long long)sc.r * 0x9E3779B97F4A7C15ull;
    philox_normals4<T>(key, nz.step, (uint32_t)(sc.c + sc.R.ncell * sc.i), 0u, n);
    return {n[0] * sc_r, n[1] * sc_r, n[2] * sc_r};
}

// ---------------------------------------------------------------------------
// Kernels

// Heun predictor: S' = kappa*normalize(S + dS1), dS1 = rhs(S, xi, dE(S)).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dipole_predict(Couplings<T> C, const T* __restrict__ S,
                                                           T* __restrict__ Sp, const T* __restrict__ ew,
                                                           NoiseArgs<T> nz, T dt, T lambda)
{
    SiteCoord sc = site_coord(C.blocks, C.reps);
    if (!sc.valid) return;
    SpinSrc<T, false> src{S, nullptr, C.total};
    Vec3T<T> s = load3(S, C.total, sc.g);
    Vec3T<T> g = local_field(C, sc, src, s, ew);
    Vec3T<T> xi = site_noise(nz, sc);
    Vec3T<T> d = rhs_dipole(s, xi, g, dt, lambda);
    store3(Sp, C.total, sc.g, normalize_dipole(Vec3T<T>{s.x + d.x, s.y + d.y, s.z + d.z}, C.kappa[sc.i]));
}

// Heun corrector: S <- kappa*normalize(S + (dS1 + dS2)/2).  dS1 is recomputed
// from S (same noise), so no stage buffer besides S' is needed.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dipole_correct(Couplings<T> C, const T* __restrict__ S,
                                                           const T* __restrict__ Sp, T* __restrict__ Sout,
                                                           const T* __restrict__ ewS,
                                                           const T* __restrict__ ewSp, NoiseArgs<T> nz, T dt,
                                                           T lambda)
{
    SiteCoord sc = site_coord(C.blocks, C.reps);
    if (!sc.valid) return;
    Vec3T<T> xi = site_noise(nz, sc);
    SpinSrc<T, false> src0{S, nullptr, C.total};
    Vec3T<T> s = load3(S, C.total, sc.g);
    Vec3T<T> d1 = rhs_dipole(s, xi, local_field(C, sc, src0, s, ewS), dt, lambda);
    SpinSrc<T, false> src1{Sp, nullptr, C.total};
    Vec3T<T> sp = load3(Sp, C.total, sc.g);
    Vec3T<T> d2 = rhs_dipole(sp, xi, local_field(C, sc, src1, sp, ewSp), dt, lambda);
    Vec3T<T> o = {s.x + (d1.x + d2.x) / 2, s.y + (d1.y + d2.y) / 2, s.z + (d1.z + d2.z) / 2};
    store3(Sout, C.total, sc.g, normalize_dipole(o, C.kappa[sc.i]));
}

// Block sum of one double per thread (wavefront shuffles, then LDS).
__device__ inline double block_sum(double v)
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

// Materialised midpoint Shat (used when a global term, Ewald, needs it).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dipole_mid_hat(Couplings<T> C, const T* __restrict__ S,
                                                           const T* __restrict__ S1, T* __restrict__ Shat,
                                                           const int* __restrict__ conv_iter)
{
    SiteCoord sc = site_coord(C.blocks, C.reps);
    if (!sc.valid || conv_iter[sc.r] != 0) return;
    SpinSrc<T, true> src{S, S1, C.total};
    store3(Shat, C.total, sc.g, src.get(sc.g, C.kappa[sc.i]));
}

// One fixed-point sweep of the spherical midpoint method:
//   Shat = kappa*normalize((S + S1)/2);  S2 = S + rhs(Shat, 0, dE(Shat)),
// plus the FP64 residual sum |S1 - S2|^2 per block (fast_isapprox).
template <typename T, bool kHatGiven>
__global__ __launch_bounds__(kBlock) void k_dipole_mid_iter(Couplings<T> C, const T* __restrict__ S,
                                                            const T* __restrict__ S1, T* __restrict__ S2,
                                                            const T* __restrict__ Shat,
                                                            const T* __restrict__ ew, T dt, T lambda,
                                                            double* __restrict__ partial,
                                                            const int* __restrict__ conv_iter)
{
    SiteCoord sc = site_coord(C.blocks, C.reps);
    if (conv_iter[sc.r] != 0) return;  // block-uniform: whole block belongs to replica r
    double d2 = 0;
    if (sc.valid) {
        Vec3T<T> h, g;
        if (kHatGiven) {
            SpinSrc<T, false> src{Shat, nullptr, C.total};
            h = load3(Shat, C.total, sc.g);
            g = local_field(C, sc, src, h, ew);
        } else {
            SpinSrc<T, true> src{S, S1, C.total};
            h = src.get(sc.g, C.kappa[sc.i]);
            g = local_field(C, sc, src, h, ew);
        }
        Vec3T<T> d = rhs_dipole(h, Vec3T<T>{0, 0, 0}, g, dt, lambda);
        Vec3T<T> s = load3(S, C.total, sc.g);
        Vec3T<T> o = {s.x + d.x, s.y + d.y, s.z + d.z};
        store3(S2, C.total, sc.g, o);
        Vec3T<T> p = load3(S1, C.total, sc.g);
        double ex = (double)p.x - (double)o.x, ey = (double)p.y - (double)o.y, ez = (double)p.z - (double)o.z;
        d2 = ex * ex + ey * ey + ez * ez;
    }
    double t = block_sum(d2);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// S = kappa*normalize(S'') for replicas that converged at iteration `iter`;
// S'' of iteration m lives in Sa when m is odd, Sb when m is even.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dipole_mid_finalize(Couplings<T> C, T* __restrict__ S,
                                                                const T* __restrict__ Sa,
                                                                const T* __restrict__ Sb,
                                                                const int* __restrict__ conv_iter, int iter)
{
    SiteCoord sc = site_coord(C.blocks, C.reps);
    int m = conv_iter[sc.r];
    if (!sc.valid || m != iter) return;
    const T* src = (m & 1) ? Sa : Sb;
    store3(S, C.total, sc.g, normalize_dipole(load3(src, C.total, sc.g), C.kappa[sc.i]));
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_dipole_grad(Couplings<T> C, const T* __restrict__ S,
                                                        const T* __restrict__ ew, T* __restrict__ dE)
{
    SiteCoord sc = site_coord(C.blocks, C.reps);
    if (!sc.valid) return;
    SpinSrc<T, false> src{S, nullptr, C.total};
    Vec3T<T> s = load3(S, C.total, sc.g);
    store3(dE, C.total, sc.g, local_field(C, sc, src, s, ew));
}

// energy_aux over non-culled bonds, Zeeman, anisotropy, Ewald (1/2 mu.phi).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dipole_energy(Couplings<T> C, const T* __restrict__ S,
                                                          const T* __restrict__ ew, double* __restrict__ partial)
{
    SiteCoord sc = site_coord(C.blocks, C.reps);
    double e = 0;
    if (sc.valid) {
        const int i = sc.i;
        Vec3T<T> s = load3(S, C.total, sc.g);
        const T* gm = C.gmat + 9 * i;
        const T* B = C.field + 3 * i;
        for (int a = 0; a < 3; a++)
            e += (double)B[a] * ((double)gm[3 * a] * s.x + (double)gm[3 * a + 1] * s.y + (double)gm[3 * a + 2] * s.z);
        T ea, gx = 0, gy = 0, gz = 0;
        stevens_grad<T>(s.x, s.y, s.z, C.stev + 28 * i, C.kmax[i], gx, gy, gz, &ea);
        e += (double)ea;
        const int* nred = C.bnred + (long long)sc.dcls * C.nb * 3;
        for (int b = C.bstart[i]; b < C.bstart[i + 1]; b++) {
            const int fl = C.bflag[b];
            if (fl & kBondCulled) continue;
            const int j = C.bj[b];
            Vec3T<T> sj = load3(S, C.total, neighbor_site(sc, nred + 3 * b, j));
            const T* J = C.bJ + 9 * b;
            double v = (double)C.bscalar[b];
            v += (double)s.x * ((double)J[0] * sj.x + (double)J[1] * sj.y + (double)J[2] * sj.z);
            v += (double)s.y * ((double)J[3] * sj.x + (double)J[4] * sj.y + (double)J[5] * sj.z);
            v += (double)s.z * ((double)J[6] * sj.x + (double)J[7] * sj.y + (double)J[8] * sj.z);
            if (fl & kBondBiquad) {
                T Qi[5], Qj[5];
                quadrupole(s, Qi);
                quadrupole(sj, Qj);
                const T* Bq = C.bQ + 25 * b;
                for (int p = 0; p < 5; p++)
                    for (int q = 0; q < 5; q++) v += (double)Qi[p] * (double)Bq[5 * p + q] * (double)Qj[q];
            }
            e += v;
        }
        if (ew) {  // 1/2 mu.phi with mu = -g S
            Vec3T<T> p = load3(ew, C.total, sc.g);
            double mx = -((double)gm[0] * s.x + (double)gm[1] * s.y + (double)gm[2] * s.z);
            double my = -((double)gm[3] * s.x + (double)gm[4] * s.y + (double)gm[5] * s.z);
            double mz = -((double)gm[6] * s.x + (double)gm[7] * s.y + (double)gm[8] * s.z);
            e += 0.5 * (mx * p.x + my * p.y + mz * p.z);
        }
    }
    double t = block_sum(e);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// One block per replica: out[r] = sum of partial[rblk[r] .. rblk[r+1]).
__global__ __launch_bounds__(kBlock) void k_reduce_partials(const double* __restrict__ partial,
                                                            const int* __restrict__ rblk, double* __restrict__ out)
{
    const int r = blockIdx.x;
    double v = 0;
    for (int k = rblk[r] + threadIdx.x; k < rblk[r + 1]; k += kBlock) v += partial[k];
    double t = block_sum(v);
    if (threadIdx.x == 0) out[r] = t;
}

// fast_isapprox (Integrators.jl:361-371) per replica: converged when the sum is
// <= atol^2 and not NaN.  Records the iteration of convergence.
__global__ __launch_bounds__(kBlock) void k_midpoint_check(const double* __restrict__ partial,
                                                           const int* __restrict__ rblk,
                                                           const double* __restrict__ atol2,
                                                           int* __restrict__ conv_iter, int iter)
{
    const int r = blockIdx.x;
    if (conv_iter[r] != 0) return;  // uniform over the block
    double v = 0;
    for (int k = rblk[r] + threadIdx.x; k < rblk[r + 1]; k += kBlock) v += partial[k];
    double t = block_sum(v);
    if (threadIdx.x == 0 && t <= atol2[r] && !isnan(t)) conv_iter[r] = iter;
}

// ---------------------------------------------------------------------------
// Launchers

template <typename T>
void launch_dipole_predict(const Couplings<T>& C, int nblocks, const T* S, T* Sp, const T* ew,
                           const NoiseArgs<T>& nz, T dt, T lambda, hipStream_t st)
{
    hipLaunchKernelGGL(k_dipole_predict<T>, dim3(nblocks), dim3(kBlock), 0, st, C, S, Sp, ew, nz, dt, lambda);
}

template <typename T>
void launch_dipole_correct(const Couplings<T>& C, int nblocks, const T* S, const T* Sp, T* Sout, const T* ewS,
                           const T* ewSp, const NoiseArgs<T>& nz, T dt, T lambda, hipStream_t st)
{
    hipLaunchKernelGGL(k_dipole_correct<T>, dim3(nblocks), dim3(kBlock), 0, st, C, S, Sp, Sout, ewS, ewSp, nz,
                       dt, lambda);
}

template <typename T>
void launch_dipole_midpoint_hat(const Couplings<T>& C, int nblocks, const T* S, const T* S1, T* Shat,
                                const int* conv_iter, hipStream_t st)
{
    hipLaunchKernelGGL(k_dipole_mid_hat<T>, dim3(nblocks), dim3(kBlock), 0, st, C, S, S1, Shat, conv_iter);
}

template <typename T>
void launch_dipole_midpoint_iter(const Couplings<T>& C, int nblocks, const T* S, const T* S1, T* S2,
                                 const T* Shat, const T* ew, T dt, T lambda, double* partial,
                                 const int* conv_iter, int iter, hipStream_t st)
{
    (void)iter;
    if (Shat)
        hipLaunchKernelGGL((k_dipole_mid_iter<T, true>), dim3(nblocks), dim3(kBlock), 0, st, C, S, S1, S2, Shat,
                           ew, dt, lambda, partial, conv_iter);
    else
        hipLaunchKernelGGL((k_dipole_mid_iter<T, false>), dim3(nblocks), dim3(kBlock), 0, st, C, S, S1, S2,
                           Shat, ew, dt, lambda, partial, conv_iter);
}

template <typename T>
void launch_dipole_midpoint_finalize(const Couplings<T>& C, int nblocks, T* S, const T* Sa, const T* Sb,
                                     const int* conv_iter, int iter, hipStream_t st)
{
    hipLaunchKernelGGL(k_dipole_mid_finalize<T>, dim3(nblocks), dim3(kBlock), 0, st, C, S, Sa, Sb, conv_iter, iter);
}

template <typename T>
void launch_dipole_grad(const Couplings<T>& C, int nblocks, const T* S, const T* ew, T* dE, hipStream_t st)
{
    hipLaunchKernelGGL(k_dipole_grad<T>, dim3(nblocks), dim3(kBlock), 0, st, C, S, ew, dE);
}

template <typename T>
void launch_dipole_energy(const Couplings<T>& C, int nblocks, const T* S, const T* ew, double* partial,
                          hipStream_t st)
{
    hipLaunchKernelGGL(k_dipole_energy<T>, dim3(nblocks), dim3(kBlock), 0, st, C, S, ew, partial);
}

void launch_reduce_partials(const double* partial, const int* rblk, int R, double* out, hipStream_t st)
{
    hipLaunchKernelGGL(k_reduce_partials, dim3(R), dim3(kBlock), 0, st, partial, rblk, out);
}

void launch_midpoint_check(const double* partial, const int* rblk, int R, const double* atol2, int* conv_iter,
                           int iter, hipStream_t st)
{
    hipLaunchKernelGGL(k_midpoint_check, dim3(R), dim3(kBlock), 0, st, partial, rblk, atol2, conv_iter, iter);
}

#define SNH_INST(T)                                                                                              \
    template void launch_dipole_predict<T>(const Couplings<T>&, int, const T*, T*, const T*,                    \
                                           const NoiseArgs<T>&, T, T, hipStream_t);                             \
    template void launch_dipole_correct<T>(const Couplings<T>&, int, const T*, const T*, T*, const T*, const T*, \
                                           const NoiseArgs<T>&, T, T, hipStream_t);                             \
    template void launch_dipole_midpoint_hat<T>(const Couplings<T>&, int, const T*, const T*, T*, const int*,    \
                                                hipStream_t);                                                    \
    template void launch_dipole_midpoint_iter<T>(const Couplings<T>&, int, const T*, const T*, T*, const T*,     \
                                                 const T*, T, T, double*, const int*, int, hipStream_t);         \
    template void launch_dipole_midpoint_finalize<T>(const Couplings<T>&, int, T*, const T*, const T*,           \
                                                     const int*, int, hipStream_t);                                   \
    template void launch_dipole_grad<T>(const Couplings<T>&, int, const T*, const T*, T*, hipStream_t);          \
    template void launch_dipole_energy<T>(const Couplings<T>&, int, const T*, const T*, double*, hipStream_t);

SNH_INST(float)
SNH_INST(double)

}  // namespace snh
