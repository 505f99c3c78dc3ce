// :dipole-mode kernels: local field by bond-stencil gather, fused stochastic
// Heun stages, spherical-midpoint sweeps, FP64 energy.
//
// Reference semantics (Sunny.jl 0.9.2):
//   set_energy_grad_dipoles!  src/System/Interactions.jl:516-592
//   rhs_dipole!               src/Integrators.jl:266-275
//   fill_noise!               src/Integrators.jl:288-298
//   step!(::System{0}, ::Langevin)          src/Integrators.jl:313-332
//   step!(::System{0}, ::ImplicitMidpoint)  src/Integrators.jl:379-413
//   energy / energy_aux       src/System/Interactions.jl:401-508
//
// The reference scatters each non-culled bond into both ends (J*Sj to i,
// J'*Si to j).  Here every directed bond (i -> j, J) and its reverse
// (j -> i, J') is in the table and each site gathers over its own bonds: no
// write conflicts, one coalesced load per neighbour plane.
//
// Heun in two kernels with two gathers per step.  The corrector's update
//   S + (dS1 + dS2)/2,  dS2 = -S' x xi  -  S' x (dt g' - dt lambda S' x g')
// is linear in the noise term, so the predictor, which knows S' and xi,
// stores H = S + dS1/2 - (S' x xi)/2 and the corrector only gathers S'.
#include "snh_gather.h"

namespace snh {

// ---------------------------------------------------------------------------
// Kernels

// Heun predictor at one site: dS1 = rhs(S, xi, dE(S)); S' = kappa*normalize(S + dS1);
// H = S + dS1/2 - (S' x xi)/2.
template <typename T, int CH, bool kIso>
__device__ __forceinline__ void dipole_predict_site(const Couplings<T>& C, const SiteCoord& sc,
                                                    const T* __restrict__ S, T* __restrict__ Sp, T* __restrict__ H,
                                                    const T* __restrict__ phi, const NoiseArgs<T>& nz, T dt,
                                                    T lambda)
{
    const Planes<T> Sv = planes(S, C.total);
    const SpinSrc<T, false> src{Sv, Sv};
    const V3<T> s = Sv(sc.g);
    const V3<T> g = local_field<T, CH, false, kIso>(C, sc, src, s, phi);
    const bool noisy = lambda != (T)0;
    const V3<T> xi = noisy ? dipole_noise(nz, sc.r, (uint32_t)(sc.c + sc.R.ncell * sc.i), sc.g) : V3<T>{0, 0, 0};
    V3<T> sp, h;
    heun_predict(s, g, xi, noisy, C.kappa[sc.i], dt, lambda, sp, h);
    store3_images(Sp, C.total, sc, C.G, sp);
    store3(H, C.total, sc.g, h);
}

template <typename T, int CH, bool kIso>
__global__ __launch_bounds__(kBlock) void k_dipole_predict(Couplings<T> C, const T* __restrict__ S,
                                                           T* __restrict__ Sp, T* __restrict__ H,
                                                           const T* __restrict__ phi, NoiseArgs<T> nz, T dt,
                                                           T lambda)
{
    const SiteCoord sc = site_coord(C.blocks, C.reps, C.G);
    if (sc.valid) dipole_predict_site<T, CH, kIso>(C, sc, S, Sp, H, phi, nz, dt, lambda);
}

// Heun corrector at one site: S <- kappa*normalize(H + rhs_det(S', dE(S'))/2).
template <typename T, int CH, bool kIso>
__device__ __forceinline__ void dipole_correct_site(const Couplings<T>& C, const SiteCoord& sc,
                                                    const T* __restrict__ Sp, const T* __restrict__ H,
                                                    T* __restrict__ Sout, const T* __restrict__ phi, T dt, T lambda)
{
    const Planes<T> Pv = planes(Sp, C.total);
    const SpinSrc<T, false> src{Pv, Pv};
    const V3<T> sp = Pv(sc.g);
    const V3<T> d = rhs_det(sp, local_field<T, CH, false, kIso>(C, sc, src, sp, phi), dt, lambda);
    const V3<T> h = planes(H, C.total)(sc.g);
    store3_images(Sout, C.total, sc, C.G,
                  normalize_dipole(V3<T>{h.x + d.x / 2, h.y + d.y / 2, h.z + d.z / 2}, C.kappa[sc.i]));
}

template <typename T, int CH, bool kIso>
__global__ __launch_bounds__(kBlock) void k_dipole_correct(Couplings<T> C, const T* __restrict__ Sp,
                                                           const T* __restrict__ H, T* __restrict__ Sout,
                                                           const T* __restrict__ phi, T dt, T lambda)
{
    const SiteCoord sc = site_coord(C.blocks, C.reps, C.G);
    if (sc.valid) dipole_correct_site<T, CH, kIso>(C, sc, Sp, H, Sout, phi, dt, lambda);
}

// Materialised midpoint Shat (when a global term, Ewald, needs it).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dipole_mid_hat(Couplings<T> C, const T* __restrict__ S,
                                                           const T* __restrict__ S1, T* __restrict__ Shat,
                                                           const int* __restrict__ conv_iter)
{
    const SiteCoord sc = site_coord(C.blocks, C.reps, C.G);
    if (!sc.valid || conv_iter[sc.r] != 0) return;
    const SpinSrc<T, true> src{planes(S, C.total), planes(S1, C.total)};
    store3_images(Shat, C.total, sc, C.G, src.get(sc.g, C.kappa[sc.i]));
}

// One fixed-point sweep at one site with Shat given (gathered with its
// images): S2 = S + rhs(Shat) at the site, the next iteration's Shat =
// kappa*normalize((S + S2)/2) with images when hat_out is non-null (S2 then
// has no images); returns |S1 - S2|^2 in FP64.
template <typename T, int CH>
__device__ __forceinline__ double dipole_mid_sweep_site(const Couplings<T>& C, const SiteCoord& sc, const T* S,
                                                        const T* S1, T* S2, const T* Shat, const T* phi, T dt,
                                                        T lambda, T* hat_out)
{
    const Planes<T> Hv = planes(Shat, C.total);
    const SpinSrc<T, false> src{Hv, Hv};
    const V3<T> h = Hv(sc.g);
    const V3<T> g = local_field<T, CH>(C, sc, src, h, phi);
    const V3<T> d = rhs_det(h, g, dt, lambda);
    const V3<T> s = planes(S, C.total)(sc.g);
    const V3<T> o = {s.x + d.x, s.y + d.y, s.z + d.z};
    if (hat_out) {
        store3(S2, C.total, sc.g, o);
        store3_images(hat_out, C.total, sc, C.G,
                      normalize_dipole(V3<T>{(s.x + o.x) / 2, (s.y + o.y) / 2, (s.z + o.z) / 2}, ldc(C.kappa, sc.i)));
    } else {
        store3_images(S2, C.total, sc, C.G, o);
    }
    const V3<T> p = planes(S1, C.total)(sc.g);
    const double ex = (double)p.x - (double)o.x, ey = (double)p.y - (double)o.y, ez = (double)p.z - (double)o.z;
    return ex * ex + ey * ey + ez * ez;
}

// One fixed-point sweep of the spherical midpoint method:
//   Shat = kappa*normalize((S + S1)/2);  S2 = S + rhs(Shat, 0, dE(Shat)),
// plus the FP64 residual sum |S1 - S2|^2 per block (fast_isapprox).
// hat_out non-null (with kHatGiven): the next iteration's Shat =
// kappa*normalize((S + S2)/2) is written with its ghost images, and S2 is
// only read at its own site afterwards (no images).
template <typename T, int CH, bool kHatGiven>
__global__ __launch_bounds__(kBlock) void k_dipole_mid_iter(Couplings<T> C, const T* __restrict__ S,
                                                            const T* __restrict__ S1, T* __restrict__ S2,
                                                            const T* __restrict__ Shat, const T* __restrict__ phi,
                                                            T dt, T lambda, double* __restrict__ partial,
                                                            const int* __restrict__ conv_iter,
                                                            T* __restrict__ hat_out)
{
    const SiteCoord sc = site_coord(C.blocks, C.reps, C.G);
    if (conv_iter[sc.r] != 0) return;  // block-uniform: the whole block belongs to replica r
    double d2 = 0;
    if (sc.valid && kHatGiven) {
        d2 = dipole_mid_sweep_site<T, CH>(C, sc, S, S1, S2, Shat, phi, dt, lambda, hat_out);
    } else if (sc.valid) {
        const SpinSrc<T, true> src{planes(S, C.total), planes(S1, C.total)};
        const V3<T> h = src.get(sc.g, C.kappa[sc.i]);
        const V3<T> g = local_field<T, CH>(C, sc, src, h, phi);
        const V3<T> d = rhs_det(h, g, dt, lambda);
        const V3<T> s = planes(S, C.total)(sc.g);
        const V3<T> o = {s.x + d.x, s.y + d.y, s.z + d.z};
        store3_images(S2, C.total, sc, C.G, o);
        const V3<T> p = planes(S1, C.total)(sc.g);
        const double ex = (double)p.x - (double)o.x, ey = (double)p.y - (double)o.y, ez = (double)p.z - (double)o.z;
        d2 = ex * ex + ey * ey + ez * ez;
    }
    const double t = block_sum(d2);
    if (threadIdx.x == 0) partial[sc.blk] = t;
}

// S = kappa*normalize(S'') for replicas that converged at iteration `iter`;
// S'' of iteration m lives in Sa when m is odd, Sb when m is even.  With
// hat_next non-null the first Shat of the next step, kappa*normalize((S +
// S)/2) as k_dipole_mid_hat forms it, is written too (with images).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dipole_mid_finalize(Couplings<T> C, T* __restrict__ S,
                                                                const T* __restrict__ Sa, const T* __restrict__ Sb,
                                                                const int* __restrict__ conv_iter, int iter,
                                                                T* __restrict__ hat_next)
{
    const SiteCoord sc = site_coord(C.blocks, C.reps, C.G);
    const int c = sc.valid ? conv_iter[sc.r] : 0;  // iter 0: every converged replica
    if (c == 0 || (iter != 0 && c != iter)) return;
    const T* src = (c & 1) ? Sa : Sb;
    const T kap = C.kappa[sc.i];
    const V3<T> s = normalize_dipole(planes(src, C.total)(sc.g), kap);
    store3_images(S, C.total, sc, C.G, s);
    if (hat_next)
        store3_images(hat_next, C.total, sc, C.G,
                      normalize_dipole(V3<T>{(s.x + s.x) / 2, (s.y + s.y) / 2, (s.z + s.z) / 2}, kap));
}

// The same for the tiled sweep's implicit iterates: S'' of iteration m is
// 2 mu_m Shat_{m+1} - S, with Shat_{m+1} in Ho (m odd) or He (m even) and
// mu_m in Mo or Me (every converged replica).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dipole_mid_finalize_mu(Couplings<T> C, T* __restrict__ S, const T* Ho,
                                                                   const T* He, const T* __restrict__ Mo,
                                                                   const T* __restrict__ Me,
                                                                   const int* __restrict__ conv_iter, T* hat_next)
{
    const SiteCoord sc = site_coord(C.blocks, C.reps, C.G);
    const int c = sc.valid ? conv_iter[sc.r] : 0;
    if (c == 0) return;
    const V3<T> h = planes((c & 1) ? Ho : He, C.total)(sc.g);
    const T t = 2 * ldg((c & 1) ? Mo : Me, sc.g);
    const V3<T> s0 = planes(S, C.total)(sc.g);
    const T kap = C.kappa[sc.i];
    const V3<T> s = normalize_dipole(V3<T>{t * h.x - s0.x, t * h.y - s0.y, t * h.z - s0.z}, kap);
    store3_images(S, C.total, sc, C.G, s);
    if (hat_next)  // may be the buffer h came from: each site reads its own entry first
        store3_images(hat_next, C.total, sc, C.G,
                      normalize_dipole(V3<T>{(s.x + s.x) / 2, (s.y + s.y) / 2, (s.z + s.z) / 2}, kap));
}

// ---------------------------------------------------------------------------
// Persistent kernels for small replicas: one workgroup of kPersist threads
// per replica runs many steps, its 256-site blocks kPersistLanes at a time,
// with workgroup barriers between stages instead of kernel launches (and,
// for the midpoint, instead of a host convergence check per step).  A
// workgroup runs on one CU, and __syncthreads orders global memory at
// workgroup scope, so one stage's stores are visible to the next.

template <typename T, typename F>
__device__ __forceinline__ void for_replica_sites(const Couplings<T>& C, int b0, int b1, F f)
{
    const int sub = (int)threadIdx.x / kBlock, lane = (int)threadIdx.x % kBlock;
    for (int b = b0 + sub; b < b1; b += kPersistLanes) {
        const SiteCoord sc = site_coord_at(C.blocks, C.reps, C.G, b, lane);
        if (sc.valid) f(sc);
    }
}

// Sum over the workgroup, returned to every thread (two barriers).
__device__ __forceinline__ double persist_sum(double v, double* red)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0;
#pragma unroll
    for (int k = 0; k < kPersist / 64; k++) t += red[k];
    __syncthreads();
    return t;
}

// nsteps Heun steps of replica blockIdx.x (steps nz.step, nz.step + 1, ...).
template <typename T, int CH, bool kIso>
__global__ __launch_bounds__(kPersist) void k_dipole_langevin_persist(Couplings<T> C, const int* __restrict__ rblk,
                                                                      T* X, T* Xp, T* Xo, NoiseArgs<T> nz, T dt,
                                                                      T lambda, int nsteps)
{
    const int b0 = rblk[blockIdx.x], b1 = rblk[blockIdx.x + 1];
    const unsigned long long step0 = nz.step;
    for (int s = 0; s < nsteps; s++) {
        nz.step = step0 + (unsigned long long)s;
        for_replica_sites(C, b0, b1, [&](const SiteCoord& sc) __attribute__((always_inline)) {
            dipole_predict_site<T, CH, kIso>(C, sc, X, Xp, Xo, nullptr, nz, dt, lambda);
        });
        __syncthreads();
        for_replica_sites(C, b0, b1, [&](const SiteCoord& sc) __attribute__((always_inline)) {
            dipole_correct_site<T, CH, kIso>(C, sc, Xp, Xo, X, nullptr, dt, lambda);
        });
        __syncthreads();
    }
}

// nsteps spherical-midpoint steps of replica r = blockIdx.x, each iterated
// to fast_isapprox convergence on the device: Shat of iteration m lives in
// hats[m % 2] (hats = {Bh2, Bh}), S'' in bufs[m % 2] (bufs = {B2, B1}).
// iters_out[r]: iterations of the last step; a step that does not converge
// within max_iters sets fail_step[r] and leaves X at the previous step.
template <typename T, int CH>
__global__ __launch_bounds__(kPersist) void k_dipole_midpoint_persist(Couplings<T> C, const int* __restrict__ rblk,
                                                                      T* X, T* B1, T* B2, T* Bh, T* Bh2, T dt,
                                                                      T lambda, const double* __restrict__ atol2,
                                                                      int max_iters, int nsteps,
                                                                      int* __restrict__ iters_out,
                                                                      int* __restrict__ fail_step)
{
    __shared__ double red[kPersist / 64];
    const int r = blockIdx.x;
    const int b0 = rblk[r], b1 = rblk[r + 1];
    T* const bufs[2] = {B2, B1};
    T* const hats[2] = {Bh2, Bh};
    const double a2 = atol2[r];
    for (int s = 0; s < nsteps; s++) {
        // the first guess S' = S: Shat = kappa n((S + S)/2) = kappa n(S)
        for_replica_sites(C, b0, b1, [&](const SiteCoord& sc) __attribute__((always_inline)) {
            store3_images(Bh, C.total, sc, C.G, normalize_dipole(planes(X, C.total)(sc.g), ldc(C.kappa, sc.i)));
        });
        __syncthreads();
        int it = 0;
        bool conv = false;
        while (!conv && it < max_iters) {
            ++it;
            const T* S1 = it == 1 ? X : bufs[(it + 1) % 2];
            T* S2 = bufs[it % 2];
            const T* hat = hats[it % 2];
            T* hat_next = hats[(it + 1) % 2];
            double d2 = 0;
            for_replica_sites(C, b0, b1, [&](const SiteCoord& sc) __attribute__((always_inline)) {
                d2 += dipole_mid_sweep_site<T, CH>(C, sc, X, S1, S2, hat, nullptr, dt, lambda, hat_next);
            });
            const double t = persist_sum(d2, red);  // uniform: every thread sums the same partials
            conv = t <= a2 && !isnan(t);
        }
        if (!conv) {
            if (threadIdx.x == 0) {
                fail_step[r] = s;
                iters_out[r] = it;
            }
            return;
        }
        const T* S2 = bufs[it % 2];
        for_replica_sites(C, b0, b1, [&](const SiteCoord& sc) __attribute__((always_inline)) {
            store3_images(X, C.total, sc, C.G, normalize_dipole(planes(S2, C.total)(sc.g), ldc(C.kappa, sc.i)));
        });
        __syncthreads();
        if (threadIdx.x == 0) iters_out[r] = it;
    }
}

// ---------------------------------------------------------------------------
// Cooperative midpoint for one small replica: one workgroup per 256-site
// block, all resident (cooperative launch), a grid barrier per fixed-point
// iteration instead of a sweep launch and a check launch.

// Barrier over the gridDim.x workgroups of a cooperative launch: a counter
// and a generation in device memory (zero at launch).  Thread 0 of each
// workgroup arrives after a release fence at agent scope (its workgroup's
// stores become visible device-wide) and waits for the generation to move
// with acquire loads; the last to arrive resets the counter and moves it.
// A wait longer than bar.spin polls sets *bar.err and proceeds, and once
// *bar.err is set no wait lasts, so every wave still exits promptly.
__device__ __forceinline__ void grid_sync(const GridBar& bar, unsigned& epoch)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned arrived = __hip_atomic_fetch_add(bar.count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == gridDim.x - 1) {
            __hip_atomic_store(bar.count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(bar.gen, epoch + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            unsigned n = 0;
            while (__hip_atomic_load(bar.gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == epoch) {
                if (++n > bar.spin || __hip_atomic_load(bar.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    __hip_atomic_store(bar.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    ++epoch;
    __syncthreads();
}

// nsteps spherical-midpoint steps of the handle's one replica, each iterated
// to fast_isapprox convergence as k_dipole_midpoint_persist does, workgroup
// b sweeping block b.  The residual of iteration m: each workgroup's block
// sum goes to partial[(m % 2) nb + b] and, after the barrier, every
// workgroup sums the nb partials as k_midpoint_check does, so all take the
// same decision (the same one the launched sweeps take).
template <typename T, int CH>
__global__ __launch_bounds__(kBlock) void k_dipole_midpoint_coop(Couplings<T> C, T* X, T* B1, T* B2, T* Bh, T* Bh2,
                                                                 T dt, T lambda, const double* __restrict__ atol2,
                                                                 int max_iters, int nsteps,
                                                                 int* __restrict__ iters_out,
                                                                 int* __restrict__ fail_step, double* partial,
                                                                 GridBar bar)
{
    __shared__ double s_tot;
    const int nb = (int)gridDim.x;
    unsigned epoch = 0;
    const SiteCoord sc = site_coord(C.blocks, C.reps, C.G);
    T* const bufs[2] = {B2, B1};
    T* const hats[2] = {Bh2, Bh};
    const double a2 = atol2[0];
    for (int s = 0; s < nsteps; s++) {
        // the first guess S' = S: Shat = kappa n((S + S)/2) = kappa n(S)
        if (sc.valid)
            store3_images(Bh, C.total, sc, C.G, normalize_dipole(planes(X, C.total)(sc.g), ldc(C.kappa, sc.i)));
        grid_sync(bar, epoch);
        int it = 0;
        bool conv = false;
        while (!conv && it < max_iters) {
            ++it;
            const T* S1 = it == 1 ? X : bufs[(it + 1) % 2];
            double d2 = 0;
            if (sc.valid)
                d2 = dipole_mid_sweep_site<T, CH>(C, sc, X, S1, bufs[it % 2], hats[it % 2], nullptr, dt, lambda,
                                                  hats[(it + 1) % 2]);
            // slots alternate: a workgroup rewrites slot it % 2 only after the
            // next barrier, by which every workgroup has read it
            double* part = partial + (it % 2) * nb;
            const double t = block_sum(d2);
            if (threadIdx.x == 0) part[blockIdx.x] = t;
            grid_sync(bar, epoch);
            double v = 0;
            for (int k = (int)threadIdx.x; k < nb; k += kBlock) v += part[k];
            const double tot = block_sum(v);
            if (threadIdx.x == 0) s_tot = tot;
            __syncthreads();
            conv = s_tot <= a2 && !isnan(s_tot);  // uniform over the grid
        }
        if (!conv) {
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                fail_step[0] = s;
                iters_out[0] = it;
            }
            return;
        }
        if (sc.valid)
            store3_images(X, C.total, sc, C.G,
                          normalize_dipole(planes(bufs[it % 2], C.total)(sc.g), ldc(C.kappa, sc.i)));
        if (blockIdx.x == 0 && threadIdx.x == 0) iters_out[0] = it;
    }
}

template <typename T, int CH>
__global__ __launch_bounds__(kBlock) void k_dipole_grad(Couplings<T> C, const T* __restrict__ S,
                                                        const T* __restrict__ phi, T* __restrict__ dE)
{
    const SiteCoord sc = site_coord(C.blocks, C.reps, C.G);
    if (!sc.valid) return;
    const Planes<T> Sv = planes(S, C.total);
    const SpinSrc<T, false> src{Sv, Sv};
    store3(dE, C.total, sc.g, local_field<T, CH>(C, sc, src, Sv(sc.g), phi));
}

// energy_aux over non-culled bonds, Zeeman, anisotropy, Ewald (1/2 mu.phi).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dipole_energy(Couplings<T> C, const T* __restrict__ S,
                                                          const T* __restrict__ phi, double* __restrict__ partial)
{
    const SiteCoord sc = site_coord(C.blocks, C.reps, C.G);
    double e = 0;
    if (sc.valid) {
        const int i = sc.i;
        const Planes<T> Sv = planes(S, C.total);
        const V3<T> s = Sv(sc.g);
        const T* gm = C.gmat + 9 * i;
        const T* B = C.field + 3 * i;
        for (int a = 0; a < 3; a++)
            e += (double)B[a] * ((double)gm[3 * a] * s.x + (double)gm[3 * a + 1] * s.y + (double)gm[3 * a + 2] * s.z);
        T ea, gx = 0, gy = 0, gz = 0;
        stevens_grad<T>(s.x, s.y, s.z, C.stev + 28 * i, C.kmax[i], gx, gy, gz, &ea);
        e += (double)ea;
        const int* __restrict__ dl = C.bdelta + sc.R.dcls * C.nb;
        for (int b = C.bstart[i]; b < C.bstart[i + 1]; b++) {
            const int fl = C.bflag[b];
            if (fl & kBondCulled) continue;
            const V3<T> sj = Sv(sc.g + (uint32_t)dl[b]);
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
        if (phi) {  // 1/2 mu.phi with mu = -g S
            const V3<T> p = planes(phi, C.total)(sc.g);
            const double mx = -((double)gm[0] * s.x + (double)gm[1] * s.y + (double)gm[2] * s.z);
            const double my = -((double)gm[3] * s.x + (double)gm[4] * s.y + (double)gm[5] * s.z);
            const double mz = -((double)gm[6] * s.x + (double)gm[7] * s.y + (double)gm[8] * s.z);
            e += 0.5 * (mx * p.x + my * p.y + mz * p.z);
        }
    }
    const double t = block_sum(e);
    if (threadIdx.x == 0) partial[sc.blk] = t;
}

// One block per replica: out[r] = sum of partial[rblk[r] .. rblk[r+1]).
__global__ __launch_bounds__(kBlock) void k_reduce_partials(const double* __restrict__ partial,
                                                            const int* __restrict__ rblk, double* __restrict__ out)
{
    const int r = blockIdx.x;
    double v = 0;
    for (int k = rblk[r] + threadIdx.x; k < rblk[r + 1]; k += kBlock) v += partial[k];
    const double t = block_sum(v);
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
    const double t = block_sum(v);
    if (threadIdx.x == 0 && t <= atol2[r] && !isnan(t)) conv_iter[r] = iter;
}

// Plain streaming copy (float4 lanes) used to calibrate traffic counters.
__global__ __launch_bounds__(256) void k_copy_planes(const float4* __restrict__ src, float4* __restrict__ dst, size_t n)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        dst[k] = src[k];
}

void launch_copy_planes(const void* src, void* dst, size_t bytes, hipStream_t st)
{
    const size_t n = bytes / sizeof(float4);
    hipLaunchKernelGGL(k_copy_planes, dim3(4096), dim3(256), 0, st, reinterpret_cast<const float4*>(src),
                       reinterpret_cast<float4*>(dst), n);
}

__global__ void k_fill_int(int* __restrict__ p, int n, int v)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}

// After step s of a budgeted group: fail_step[r] = s for replicas that have
// not converged within the budget (the first such step), max_iter[r] = the
// largest iteration count of the group so far, last_it[r] = this step's
// count; conv_iter[r] is cleared for the next step.
__global__ void k_mid_budget(int* __restrict__ conv_iter, int* __restrict__ fail_step, int* __restrict__ max_iter,
                             int* __restrict__ last_it, int R, int s)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const int c = conv_iter[r];
    if (c == 0 && fail_step[r] < 0) fail_step[r] = s;
    if (s == 0 || c > max_iter[r]) max_iter[r] = c;
    last_it[r] = c;
    conv_iter[r] = 0;
}

__global__ void k_step_set(unsigned long long* __restrict__ p, unsigned long long v)
{
    if (threadIdx.x == 0) p[0] = v;
}

__global__ void k_step_advance(unsigned long long* __restrict__ p)
{
    if (threadIdx.x == 0) p[0] += 1;
}

// ---------------------------------------------------------------------------
// Launchers

template <typename T>
void launch_dipole_predict(const Couplings<T>& C, int nblocks, const T* S, T* Sp, T* H, const T* phi,
                           const NoiseArgs<T>& nz, T dt, T lambda, hipStream_t st)
{
#define SNH_PRED(CH, ISO)                                                                                      \
    hipLaunchKernelGGL((k_dipole_predict<T, CH, ISO>), dim3(nblocks), dim3(kBlock), 0, st, C, S, Sp, H, phi, nz, dt, \
                       lambda)
    if (C.chunk == 18) {
        if (C.iso) SNH_PRED(18, true); else SNH_PRED(18, false);
    } else if (C.chunk == 9) {
        if (C.iso) SNH_PRED(9, true); else SNH_PRED(9, false);
    } else {
        if (C.iso) SNH_PRED(6, true); else SNH_PRED(6, false);
    }
#undef SNH_PRED
}

template <typename T>
void launch_dipole_correct(const Couplings<T>& C, int nblocks, const T* Sp, const T* H, T* Sout, const T* phi, T dt,
                           T lambda, hipStream_t st)
{
#define SNH_CORR(CH, ISO)                                                                                      \
    hipLaunchKernelGGL((k_dipole_correct<T, CH, ISO>), dim3(nblocks), dim3(kBlock), 0, st, C, Sp, H, Sout, phi, dt, \
                       lambda)
    if (C.chunk == 18) {
        if (C.iso) SNH_CORR(18, true); else SNH_CORR(18, false);
    } else if (C.chunk == 9) {
        if (C.iso) SNH_CORR(9, true); else SNH_CORR(9, false);
    } else {
        if (C.iso) SNH_CORR(6, true); else SNH_CORR(6, false);
    }
#undef SNH_CORR
}

template <typename T>
void launch_dipole_midpoint_hat(const Couplings<T>& C, int nblocks, const T* S, const T* S1, T* Shat,
                                const int* conv_iter, hipStream_t st)
{
    hipLaunchKernelGGL(k_dipole_mid_hat<T>, dim3(nblocks), dim3(kBlock), 0, st, C, S, S1, Shat, conv_iter);
}

template <typename T>
void launch_dipole_midpoint_iter(const Couplings<T>& C, int nblocks, const T* S, const T* S1, T* S2,
                                 const T* Shat, const T* phi, T dt, T lambda, double* partial, const int* conv_iter,
                                 T* hat_out, hipStream_t st)
{
    if (Shat)
        hipLaunchKernelGGL((k_dipole_mid_iter<T, 6, true>), dim3(nblocks), dim3(kBlock), 0, st, C, S, S1, S2, Shat,
                           phi, dt, lambda, partial, conv_iter, hat_out);
    else
        hipLaunchKernelGGL((k_dipole_mid_iter<T, 6, false>), dim3(nblocks), dim3(kBlock), 0, st, C, S, S1, S2, Shat,
                           phi, dt, lambda, partial, conv_iter, (T*)nullptr);
}

template <typename T>
void launch_dipole_midpoint_finalize(const Couplings<T>& C, int nblocks, T* S, const T* Sa, const T* Sb,
                                     const int* conv_iter, int iter, T* hat_next, hipStream_t st)
{
    hipLaunchKernelGGL(k_dipole_mid_finalize<T>, dim3(nblocks), dim3(kBlock), 0, st, C, S, Sa, Sb, conv_iter, iter,
                       hat_next);
}

template <typename T>
void launch_dipole_midpoint_finalize_mu(const Couplings<T>& C, int nblocks, T* S, const T* Ho, const T* He,
                                        const T* Mo, const T* Me, const int* conv_iter, T* hat_next, hipStream_t st)
{
    hipLaunchKernelGGL(k_dipole_mid_finalize_mu<T>, dim3(nblocks), dim3(kBlock), 0, st, C, S, Ho, He, Mo, Me,
                       conv_iter, hat_next);
}

template <typename T>
void launch_dipole_grad(const Couplings<T>& C, int nblocks, const T* S, const T* phi, T* dE, hipStream_t st)
{
    hipLaunchKernelGGL((k_dipole_grad<T, 6>), dim3(nblocks), dim3(kBlock), 0, st, C, S, phi, dE);
}

template <typename T>
void launch_dipole_energy(const Couplings<T>& C, int nblocks, const T* S, const T* phi, double* partial,
                          hipStream_t st)
{
    hipLaunchKernelGGL(k_dipole_energy<T>, dim3(nblocks), dim3(kBlock), 0, st, C, S, phi, partial);
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

void launch_fill_int(int* p, int n, int v, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_int, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
}

void launch_mid_budget(int* conv_iter, int* fail_step, int* max_iter, int* last_it, int R, int s, hipStream_t st)
{
    hipLaunchKernelGGL(k_mid_budget, dim3((R + 255) / 256), dim3(256), 0, st, conv_iter, fail_step, max_iter,
                       last_it, R, s);
}

void launch_step_set(unsigned long long* p, unsigned long long v, hipStream_t st)
{
    hipLaunchKernelGGL(k_step_set, dim3(1), dim3(64), 0, st, p, v);
}

void launch_step_advance(unsigned long long* p, hipStream_t st)
{
    hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, st, p);
}

template <typename T>
void launch_dipole_langevin_persist(const Couplings<T>& C, int R, const int* rblk, T* X, T* Xp, T* Xo,
                                   const NoiseArgs<T>& nz, T dt, T lambda, int nsteps, hipStream_t st)
{
#define SNH_LP(CH, ISO)                                                                                      \
    hipLaunchKernelGGL((k_dipole_langevin_persist<T, CH, ISO>), dim3(R), dim3(kPersist), 0, st, C, rblk, X, Xp, \
                       Xo, nz, dt, lambda, nsteps)
    if (C.chunk == 18) {
        if (C.iso) SNH_LP(18, true); else SNH_LP(18, false);
    } else if (C.chunk == 9) {
        if (C.iso) SNH_LP(9, true); else SNH_LP(9, false);
    } else {
        if (C.iso) SNH_LP(6, true); else SNH_LP(6, false);
    }
#undef SNH_LP
}

template <typename T>
void launch_dipole_midpoint_persist(const Couplings<T>& C, int R, const int* rblk, T* X, T* B1, T* B2, T* Bh, T* Bh2,
                                    T dt, T lambda, const double* atol2, int max_iters, int nsteps, int* iters_out,
                                    int* fail_step, hipStream_t st)
{
    hipLaunchKernelGGL((k_dipole_midpoint_persist<T, 6>), dim3(R), dim3(kPersist), 0, st, C, rblk, X, B1, B2, Bh,
                       Bh2, dt, lambda, atol2, max_iters, nsteps, iters_out, fail_step);
}

template <typename T>
hipError_t launch_dipole_midpoint_coop(const Couplings<T>& C, int nblocks, T* X, T* B1, T* B2, T* Bh, T* Bh2, T dt,
                                       T lambda, const double* atol2, int max_iters, int nsteps, int* iters_out,
                                       int* fail_step, double* partial, const GridBar& bar, hipStream_t st)
{
    Couplings<T> c = C;
    GridBar b = bar;
    void* args[] = {&c, &X, &B1, &B2, &Bh, &Bh2, &dt, &lambda, &atol2, &max_iters, &nsteps, &iters_out, &fail_step,
                    &partial, &b};
    const void* fn = C.chunk == 18   ? reinterpret_cast<const void*>(&k_dipole_midpoint_coop<T, 18>)
                     : C.chunk == 9 ? reinterpret_cast<const void*>(&k_dipole_midpoint_coop<T, 9>)
                                    : reinterpret_cast<const void*>(&k_dipole_midpoint_coop<T, 6>);
    return hipLaunchCooperativeKernel(fn, dim3(nblocks), dim3(kBlock), args, 0, st);
}

#define SNH_INST(T)                                                                                              \
    template void launch_dipole_predict<T>(const Couplings<T>&, int, const T*, T*, T*, const T*,                \
                                           const NoiseArgs<T>&, T, T, hipStream_t);                             \
    template void launch_dipole_correct<T>(const Couplings<T>&, int, const T*, const T*, T*, const T*, T, T,     \
                                           hipStream_t);                                                         \
    template void launch_dipole_midpoint_hat<T>(const Couplings<T>&, int, const T*, const T*, T*, const int*,    \
                                                hipStream_t);                                                    \
    template hipError_t launch_dipole_midpoint_coop<T>(const Couplings<T>&, int, T*, T*, T*, T*, T*, T, T,       \
                                                       const double*, int, int, int*, int*, double*,             \
                                                       const GridBar&, hipStream_t);                             \
    template void launch_dipole_midpoint_iter<T>(const Couplings<T>&, int, const T*, const T*, T*, const T*,     \
                                                 const T*, T, T, double*, const int*, T*, hipStream_t);                     \
    template void launch_dipole_midpoint_finalize_mu<T>(const Couplings<T>&, int, T*, const T*, const T*,        \
                                                        const T*, const T*, const int*, T*, hipStream_t);       \
    template void launch_dipole_midpoint_finalize<T>(const Couplings<T>&, int, T*, const T*, const T*,           \
                                                     const int*, int, T*, hipStream_t);                              \
    template void launch_dipole_grad<T>(const Couplings<T>&, int, const T*, const T*, T*, hipStream_t);          \
    template void launch_dipole_langevin_persist<T>(const Couplings<T>&, int, const int*, T*, T*, T*,             \
                                                    const NoiseArgs<T>&, T, T, int, hipStream_t);                \
    template void launch_dipole_midpoint_persist<T>(const Couplings<T>&, int, const int*, T*, T*, T*, T*, T*, T, \
                                                    T, const double*, int, int, int*, int*, hipStream_t);        \
    template void launch_dipole_energy<T>(const Couplings<T>&, int, const T*, const T*, double*, hipStream_t);

SNH_INST(float)
SNH_INST(double)

}  // namespace snh
