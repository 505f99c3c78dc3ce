// LDS-tiled :dipole kernels: the bond-stencil gather reads neighbours from
// an LDS ring of cell layers instead of global memory.
//
// A workgroup owns a column of T1 x T2 cells (all sublattices) of one
// replica and sweeps it along axis 3.  For the layer k being computed, the
// layers k - w3 .. k + w3 of the tile widened by the halo (w1, w2) are held
// in LDS, one slot per layer, position-major with the three components of a
// site adjacent.  Each bond (i -> j, n) is then one constant byte offset
// from the site's own LDS address (per ring phase), so a neighbour costs one
// address add and one 12-byte LDS read.  While a layer is computed, the next
// layer's global loads are in flight in registers; two barriers per layer
// retire the oldest slot and publish the new one.
//
// Per spin the global traffic is the compulsory plane reads and writes plus
// the halo re-reads of neighbouring columns (served mostly from L2), instead
// of one global load per bond and component.
//
// Reference semantics (Sunny.jl 0.9.2) are those of snh_dipole.hip:
//   set_energy_grad_dipoles!  src/System/Interactions.jl:516-592
//   step!(::System{0}, ::Langevin)          src/Integrators.jl:313-332
//   step!(::System{0}, ::ImplicitMidpoint)  src/Integrators.jl:379-413
//   energy_aux                src/System/Interactions.jl:435-508
#include "snh_vec.h"

#include <map>
#include <mutex>

namespace snh {

constexpr int kTileThreads = 1024;

// Occupancy request for the tiled kernels (waves per SIMD; 0 = compiler's choice).
#ifndef SNH_TILE_WAVES_PER_EU
#define SNH_TILE_WAVES_PER_EU 0
#endif
#if SNH_TILE_WAVES_PER_EU > 0
#define SNH_TILE_OCC __attribute__((amdgpu_waves_per_eu(SNH_TILE_WAVES_PER_EU)))
#else
#define SNH_TILE_OCC
#endif

__device__ __forceinline__ int wrap(int v, int L)
{
    int m = v % L;
    return m < 0 ? m + L : m;
}

template <typename T>
__device__ __forceinline__ V3<T> lds_v3(const unsigned char* lds, uint32_t byte)
{
    const T* p = reinterpret_cast<const T*>(lds + byte);
    return {p[0], p[1], p[2]};
}

template <typename T>
__device__ __forceinline__ V3<T> plane_v3(const T* __restrict__ P, uint32_t total, uint32_t g)
{
    return {ldg(P, g), ldg(P, g + total), ldg(P, g + 2 * total)};
}

template <typename T>
__device__ __forceinline__ void plane_store(T* __restrict__ P, uint32_t total, uint32_t g, V3<T> v)
{
    stg(P, g, v.x);
    stg(P, g + total, v.y);
    stg(P, g + 2 * total, v.z);
}

// Buffer descriptor over `bytes` bytes from p (raw addressing, no stride),
// and a load of one T at byte offset voff (per lane) + soff (wavefront
// uniform, a scalar register): no per-load address arithmetic in vector
// registers.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void* p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}

template <typename T>
__device__ __forceinline__ T buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    if constexpr (sizeof(T) == 4)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

template <typename T>
__device__ __forceinline__ void buf_st(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, T v)
{
    typedef __attribute__((__vector_size__(2 * sizeof(unsigned int)))) unsigned int u32x2;
    if constexpr (sizeof(T) == 4)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, voff, soff, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 0);
}

// The three planes of a state through a descriptor over all of them: the
// site at byte offset voff of layer 0, layer offset soff (bytes), plane
// stride pb (bytes).
template <typename T>
__device__ __forceinline__ V3<T> buf_v3(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, uint32_t pb)
{
    return {buf_ld<T>(r, voff, soff), buf_ld<T>(r, voff, soff + pb), buf_ld<T>(r, voff, soff + 2 * pb)};
}

template <typename T>
__device__ __forceinline__ void buf_store_v3(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, uint32_t pb,
                                             V3<T> v)
{
    buf_st<T>(r, voff, soff, v.x);
    buf_st<T>(r, voff, soff + pb, v.y);
    buf_st<T>(r, voff, soff + 2 * pb, v.z);
}

// Sum of one double per thread over a workgroup of any multiple of 64 up to 1024.
__device__ __forceinline__ double tile_block_sum(double v)
{
    __shared__ double red[kTileThreads / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); k++) t += red[k];
    return t;
}

// The site a thread computes in every layer of its column.
struct TileSite {
    uint32_t gb;   // plane index in layer 0 (add L1 L2 k for layer k)
    uint32_t own;  // LDS byte offset of the site within a slot
    int i;         // sublattice (uniform over a wavefront)
    bool ok;

    __device__ __forceinline__ TileSite(const TileArgs& A, const TileWork& W, const ReplicaInfo& R, int na, int elem,
                                        int s)
    {
        const int sub = s / A.spl;
        const int q = s - sub * A.spl;
        const int c1l = q & (A.T1 - 1), c2l = q >> A.lt1;
        const int c1 = W.t1 + c1l, c2 = W.t2 + c2l;
        ok = s < A.nsite && q < A.T1 * A.T2 && c1 < R.L1 && c2 < R.L2;
        // every wavefront lies within one sublattice (spl is a multiple of 64)
        i = __builtin_amdgcn_readfirstlane(s < A.nsite ? sub : 0);
        own = ok ? (uint32_t)(3 * elem * ((c1l + A.w1) + A.ROW * (i + na * (c2l + A.w2)))) : 0u;
        // idle lanes load the replica's first site (in bounds) and store nothing
        gb = ok ? R.off + (uint32_t)(c1 + R.L1 * c2) + (uint32_t)R.ncell * (uint32_t)i : R.off;
    }
};

// Sweeps the workgroup's column: loads the ring, and for every layer k calls
// pre(k + 1) (which moves the own-site inputs of layer k to the registers the
// body reads and issues the loads of layer k + 1), issues the next ring
// layer's loads, then calls body(k, phase).  NSRC = 2 loads the midpoint
// state kappa*normalize((S0 + S1)/2) instead of S0.
template <typename T, int NSRC, int SP, typename Pre, typename Body>
__device__ __forceinline__ void tile_sweep(const Couplings<T>& C, const TileArgs& A, const TileWork& W,
                                           const ReplicaInfo& R, const T* __restrict__ s0, const T* __restrict__ s1,
                                           unsigned char* smem, Pre pre, Body body)
{
    constexpr int kLoadPer = tile_load_per(SP);
    T* lds = reinterpret_cast<T*>(smem);
    const int NT = blockDim.x;
    const uint32_t total = C.total;
    const uint32_t lay = (uint32_t)R.L1 * (uint32_t)R.L2;
    constexpr uint32_t esz = sizeof(T);
    // one descriptor per source over its three planes (< 4 GiB, checked on the host)
    const __amdgpu_buffer_rsrc_t r0 = plane_rsrc(s0, 3u * total * esz);
    const __amdgpu_buffer_rsrc_t r1 = plane_rsrc(NSRC == 2 ? s1 : s0, 3u * total * esz);
    uint32_t lg[kLoadPer];  // byte offset of the position in layer 0
    int lp[kLoadPer];
    T lk[kLoadPer];
#pragma unroll
    for (int m = 0; m < kLoadPer; m++) {
        // positions past npos load position 0 (in bounds) and are not stored:
        // every load is unconditional, so waits on own-site inputs can count
        // the prefetch loads issued after them
        const int p = (int)threadIdx.x + m * NT;
        lp[m] = p < A.npos ? p : -1;
        const int pp = p < A.npos ? p : 0;
        const int x = pp % A.ROW, t = pp / A.ROW;
        const int j = t % C.na, row = t / C.na;
        const int c1 = wrap(W.t1 + x - A.w1, R.L1), c2 = wrap(W.t2 + row - A.w2, R.L2);
        lg[m] = (R.off + (uint32_t)(c1 + R.L1 * c2) + (uint32_t)R.ncell * (uint32_t)j) * esz;
        lk[m] = NSRC == 2 ? ldc(C.kappa, j) : (T)0;
    }
    T buf[kLoadPer][3 * NSRC];
    auto fetch = [&](int kk) __attribute__((always_inline)) {
        const uint32_t o = lay * (uint32_t)wrap(kk, R.L3) * esz;
#pragma unroll
        for (int m = 0; m < kLoadPer; m++) {
#pragma unroll
            for (int a = 0; a < 3; a++) buf[m][a] = buf_ld<T>(r0, lg[m], o + a * total * esz);
            if constexpr (NSRC == 2) {
#pragma unroll
                for (int a = 0; a < 3; a++) buf[m][3 + a] = buf_ld<T>(r1, lg[m], o + a * total * esz);
            }
        }
    };
    auto put = [&](int slot) __attribute__((always_inline)) {
#pragma unroll
        for (int m = 0; m < kLoadPer; m++) {
            if (lp[m] < 0) continue;
            V3<T> v = {buf[m][0], buf[m][1], buf[m][2]};
            if constexpr (NSRC == 2)
                v = normalize_dipole(
                    V3<T>{(v.x + buf[m][3]) / 2, (v.y + buf[m][4]) / 2, (v.z + buf[m][5]) / 2}, lk[m]);
            T* d = lds + 3 * (slot * A.cap + lp[m]);
            d[0] = v.x;
            d[1] = v.y;
            d[2] = v.z;
        }
    };
    pre(W.k0);
    for (int d = 0; d < A.nslot; d++) {
        fetch(W.k0 - A.w3 + d);
        put(d);
    }
    __syncthreads();
    for (int k = W.k0; k < W.k1; k++) {
        const int rel = k - W.k0;
        const bool more = k + 1 < W.k1;
        // own-site inputs of layer k arrived with the previous layer's
        // prefetch (vector-memory counters retire in order); issue those of
        // layer k + 1, then the next ring layer
        pre(more ? k + 1 : k);
        fetch(k + A.w3 + 1);  // wraps in bounds; unused after the last layer
        body(k, (rel + A.w3) % A.nslot);
        if (more) {
            __syncthreads();
            put(rel % A.nslot);  // the slot of layer k - w3, now retired
            __syncthreads();
        }
    }
}

// dE/dS of a site whose own LDS byte address is `own`: Zeeman, bonds from
// the ring (CH per batch of LDS reads), onsite anisotropy (kQ2: of rank at
// most 2 on every sublattice).
template <typename T, int CH, bool kIso, bool kQ2 = false>
__device__ __forceinline__ V3<T> tile_field(const Couplings<T>& C, const int* __restrict__ offs,
                                            const unsigned char* lds, uint32_t own, int i, V3<T> s)
{
    V3<T> g = {ldc(C.zee, 3 * i), ldc(C.zee, 3 * i + 1), ldc(C.zee, 3 * i + 2)};
    const int b0 = ldc(C.bstart, i), b1 = ldc(C.bstart, i + 1);
    for (int c0 = b0; c0 < b1; c0 += CH) {
        V3<T> nb[CH];
#pragma unroll
        for (int k = 0; k < CH; k++) nb[k] = lds_v3<T>(lds, own + (uint32_t)ldc(offs, c0 + k));
#pragma unroll
        for (int k = 0; k < CH; k++) {
            const int b = c0 + k;
            if (kIso) {
                const T J0 = ldc(C.bJ0, b);
                g.x += J0 * nb[k].x;
                g.y += J0 * nb[k].y;
                g.z += J0 * nb[k].z;
            } else {
                add_bond_field(C, b, s, nb[k], g);
            }
        }
    }
    if constexpr (kQ2) {
        if (ldc(C.kmax, i) == 2) {
            const T* q = C.stev_q2 + 9 * i;
            g.x += ldc(q, 0) * s.x + ldc(q, 1) * s.y + ldc(q, 2) * s.z;
            g.y += ldc(q, 3) * s.x + ldc(q, 4) * s.y + ldc(q, 5) * s.z;
            g.z += ldc(q, 6) * s.x + ldc(q, 7) * s.y + ldc(q, 8) * s.z;
        }
    } else {
        add_onsite_field(C, i, s, g);
    }
    return g;
}

// ---------------------------------------------------------------------------
// Kernels
//
// Every kernel computes up to SP sites per thread and layer.  Own-site global
// inputs (H, S, S1, the Ewald potential) of layer k + 1 are loaded by the
// sweep's pre hook into the nx registers while layer k is computed, and move
// to the in registers the body reads at the next layer.

template <typename T>
struct OwnIn {
    V3<T> a, b, p;
    T m;
};

#define SNH_TILE_SITES(T_)                                                                          \
    const TileSite ts0(A, W, R, C.na, (int)sizeof(T_), (int)threadIdx.x);                           \
    const TileSite ts1(A, W, R, C.na, (int)sizeof(T_), (int)(threadIdx.x + blockDim.x));            \
    constexpr bool kTwo = SP > 1;                                                                   \
    const uint32_t lay = (uint32_t)R.L1 * (uint32_t)R.L2;                                           \
    const uint32_t slot = 3u * (uint32_t)sizeof(T_) * (uint32_t)A.cap;                              \
    OwnIn<T_> in0{}, in1{}, nx0{}, nx1{}

// Waves per SIMD requested for the predictor: the lean FP32 one-site form
// fits eight (60 registers; some scalar values kept in register lanes), so
// two 1024-thread workgroups share a CU (predictor 0.403 -> 0.372 ms on the
// 64-replica pyrochlore batch); else the compiler's choice.
#ifndef SNH_PRED_WAVES
#define SNH_PRED_WAVES 8
#endif
template <typename T, int SP, bool kLean>
constexpr int predict_waves()
{
    if (SNH_TILE_WAVES_PER_EU > 0) return SNH_TILE_WAVES_PER_EU;
    return (SNH_PRED_WAVES > 0 && kLean && SP == 1 && sizeof(T) == 4) ? SNH_PRED_WAVES : 1;
}

// Heun predictor (see heun_predict): S' and H from S.  kLean: no Ewald
// potential, device noise with the step as an argument (or none), onsite
// terms of rank <= 2; fewer live pointers and branches in the layer loop.
template <typename T, int CH, bool kIso, int SP, bool kLean>
__global__ __attribute__((amdgpu_waves_per_eu(predict_waves<T, SP, kLean>())))
__launch_bounds__((tile_threads_max<T, SP>())) void k_tile_predict(Couplings<T> C, TileArgs A,
                                                                        const T* __restrict__ S, T* __restrict__ Sp,
                                                                        T* __restrict__ H, const T* __restrict__ phi,
                                                                        NoiseArgs<T> nz, T dt, T lambda)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const TileWork W = A.work[logical_block()];
    const ReplicaInfo R = C.reps[W.r];
    SNH_TILE_SITES(T);
    const bool noisy = lambda != (T)0;
    auto load = [&](const TileSite& ts, int k, OwnIn<T>& in) __attribute__((always_inline)) {
        if (!kLean && phi) in.p = plane_v3(phi, C.total, ts.gb + lay * (uint32_t)k);
    };
    auto site = [&](const TileSite& ts, int k, int ph, const int* offs, const OwnIn<T>& in) __attribute__((always_inline)) {
        if (!ts.ok) return;
        const uint32_t own = ts.own + slot * (uint32_t)ph;
        const uint32_t g = ts.gb + lay * (uint32_t)k;
        const int i = ts.i;
        const V3<T> s = lds_v3<T>(smem, own);
        V3<T> f = tile_field<T, CH, kIso, kLean>(C, offs, smem, own, i, s);
        if (!kLean && phi) add_ewald_field(C, i, in.p, f);
        V3<T> xi = {0, 0, 0};
        if (noisy) xi = kLean ? dipole_noise_philox(nz, W.r, g - R.off) : dipole_noise(nz, W.r, g - R.off, g);
        V3<T> sp, h;
        heun_predict(s, f, xi, noisy, ldc(C.kappa, i), dt, lambda, sp, h);
        plane_store(Sp, C.total, g, sp);
        plane_store(H, C.total, g, h);
    };
    tile_sweep<T, 1, SP>(
        C, A, W, R, S, nullptr, smem,
        [&](int k) __attribute__((always_inline)) {
            in0 = nx0;
            in1 = nx1;
            load(ts0, k, nx0);
            if (kTwo) load(ts1, k, nx1);
        },
        [&](int k, int ph) __attribute__((always_inline)) {
            const int* offs = A.ldsoff + ph * C.nb;
            site(ts0, k, ph, offs, in0);
            if (kTwo) site(ts1, k, ph, offs, in1);
        });
}

// Waves per SIMD requested for the corrector and the midpoint iteration: the
// lean FP32 one-site forms fit in 64 registers, so two 1024-thread
// workgroups share a CU.
template <typename T, int SP, bool kLean>
constexpr int correct_waves()
{
    if (SNH_TILE_WAVES_PER_EU > 0) return SNH_TILE_WAVES_PER_EU;  // a build-time override
    return (kLean && SP == 1 && sizeof(T) == 4) ? 8 : 1;
}

// Heun corrector: S <- kappa*normalize(H + rhs_det(S', dE(S'))/2).  kLean:
// no Ewald potential, onsite terms of rank <= 2.
template <typename T, int CH, bool kIso, int SP, bool kLean>
__global__ __attribute__((amdgpu_waves_per_eu(correct_waves<T, SP, kLean>())))
__launch_bounds__((tile_threads_max<T, SP>())) void k_tile_correct(Couplings<T> C, TileArgs A,
                                                                        const T* __restrict__ Sp,
                                                                        const T* __restrict__ H, T* __restrict__ Sout,
                                                                        const T* __restrict__ phi, T dt, T lambda)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const TileWork W = A.work[logical_block()];
    const ReplicaInfo R = C.reps[W.r];
    SNH_TILE_SITES(T);
    const uint32_t pb = C.total * (uint32_t)sizeof(T);
    const __amdgpu_buffer_rsrc_t rH = plane_rsrc(H, 3 * pb), rS = plane_rsrc(Sout, 3 * pb);
    auto load = [&](const TileSite& ts, int k, OwnIn<T>& in) __attribute__((always_inline)) {
        const uint32_t g = ts.gb + lay * (uint32_t)k;
        in.a = buf_v3<T>(rH, ts.gb * (uint32_t)sizeof(T), lay * (uint32_t)k * (uint32_t)sizeof(T), pb);
        if (!kLean && phi) in.p = plane_v3(phi, C.total, g);
    };
    auto site = [&](const TileSite& ts, int k, int ph, const int* offs, const OwnIn<T>& in) __attribute__((always_inline)) {
        if (!ts.ok) return;
        const uint32_t own = ts.own + slot * (uint32_t)ph;
        const int i = ts.i;
        const V3<T> sp = lds_v3<T>(smem, own);
        V3<T> f = tile_field<T, CH, kIso, kLean>(C, offs, smem, own, i, sp);
        if (!kLean && phi) add_ewald_field(C, i, in.p, f);
        const V3<T> d = rhs_det(sp, f, dt, lambda);
        const V3<T> h = in.a;
        buf_store_v3<T>(rS, ts.gb * (uint32_t)sizeof(T), lay * (uint32_t)k * (uint32_t)sizeof(T), pb,
                        normalize_dipole(V3<T>{h.x + d.x / 2, h.y + d.y / 2, h.z + d.z / 2}, ldc(C.kappa, i)));
    };
    tile_sweep<T, 1, SP>(
        C, A, W, R, Sp, nullptr, smem,
        [&](int k) __attribute__((always_inline)) {
            in0 = nx0;
            in1 = nx1;
            load(ts0, k, nx0);
            if (kTwo) load(ts1, k, nx1);
        },
        [&](int k, int ph) __attribute__((always_inline)) {
            const int* offs = A.ldsoff + ph * C.nb;
            site(ts0, k, ph, offs, in0);
            if (kTwo) site(ts1, k, ph, offs, in1);
        });
}

// One fixed-point sweep of the spherical midpoint method on the ring of
// Shat = kappa*normalize((S + S1)/2) (or a materialised Shat when a global
// term needs it), with the FP64 residual sum |S1 - S2|^2 per workgroup.
// kLean: isotropic bilinear bonds, no Ewald potential, onsite rank <= 2.
// With Shat given the iterates stay implicit: the sweep writes the next
// Shat = kappa n(v), v = (S + S'')/2, and mu = |v|/kappa (mu_out), and
// S'' = 2 mu Shat - S is formed where needed (S1 of the residual from the
// previous iteration's mu_prev and this Shat; null mu_prev: S1 = S, the
// first guess): 44 bytes per site instead of 60.
template <typename T, int CH, bool kHatGiven, int SP, bool kLean>
__global__ __attribute__((amdgpu_waves_per_eu(correct_waves<T, SP, kLean && kHatGiven>())))
__launch_bounds__((tile_threads_max<T, SP>())) void k_tile_mid_iter(Couplings<T> C, TileArgs A,
                                                                         const T* __restrict__ S,
                                                                         const T* __restrict__ S1,
                                                                         T* __restrict__ S2,
                                                                         const T* __restrict__ Shat,
                                                                         const T* __restrict__ phi, T dt, T lambda,
                                                                         double* __restrict__ partial,
                                                                         const int* __restrict__ conv_iter,
                                                                         T* __restrict__ hat_out,
                                                                         const T* __restrict__ mu_prev,
                                                                         T* __restrict__ mu_out)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int wg = logical_block();
    const TileWork W = A.work[wg];
    if (conv_iter[W.r] != 0) return;  // uniform over the workgroup
    const ReplicaInfo R = C.reps[W.r];
    SNH_TILE_SITES(T);
    double d2 = 0;
    auto load = [&](const TileSite& ts, int k, OwnIn<T>& in) __attribute__((always_inline)) {
        const uint32_t g = ts.gb + lay * (uint32_t)k;
        in.a = plane_v3(S, C.total, g);
        if (!kHatGiven)
            in.b = plane_v3(S1, C.total, g);
        else if (mu_prev)
            in.m = ldg(mu_prev, g);
        if (!kLean && phi) in.p = plane_v3(phi, C.total, g);
    };
    auto site = [&](const TileSite& ts, int k, int ph, const int* offs, const OwnIn<T>& in) __attribute__((always_inline)) {
        if (!ts.ok) return;
        const uint32_t own = ts.own + slot * (uint32_t)ph;
        const uint32_t g = ts.gb + lay * (uint32_t)k;
        const int i = ts.i;
        const V3<T> hat = lds_v3<T>(smem, own);
        V3<T> f = tile_field<T, CH, kLean, kLean>(C, offs, smem, own, i, hat);
        if (!kLean && phi) add_ewald_field(C, i, in.p, f);
        const V3<T> d = rhs_det(hat, f, dt, lambda);
        const V3<T> o = {in.a.x + d.x, in.a.y + d.y, in.a.z + d.z};
        V3<T> s1 = in.b;
        if (kHatGiven) {
            // the next iteration's Shat = kappa n(v), v = (S + S'')/2, and mu
            T mu;
            plane_store(hat_out, C.total, g,
                        normalize_dipole_mu(V3<T>{(in.a.x + o.x) / 2, (in.a.y + o.y) / 2, (in.a.z + o.z) / 2},
                                            ldc(C.kappa, i), mu));
            stg(mu_out, g, mu);
            // the previous iterate S'' = 2 mu_prev Shat - S (S at the first iteration)
            s1 = in.a;
            if (mu_prev) {
                const T t = 2 * in.m;
                s1 = {t * hat.x - in.a.x, t * hat.y - in.a.y, t * hat.z - in.a.z};
            }
        } else {
            plane_store(S2, C.total, g, o);
        }
        const double ex = (double)s1.x - (double)o.x, ey = (double)s1.y - (double)o.y, ez = (double)s1.z - (double)o.z;
        d2 += ex * ex + ey * ey + ez * ez;
    };
    auto pre = [&](int k) __attribute__((always_inline)) {
        in0 = nx0;
        in1 = nx1;
        load(ts0, k, nx0);
        if (kTwo) load(ts1, k, nx1);
    };
    auto body = [&](int k, int ph) __attribute__((always_inline)) {
        const int* offs = A.ldsoff + ph * C.nb;
        site(ts0, k, ph, offs, in0);
        if (kTwo) site(ts1, k, ph, offs, in1);
    };
    if (kHatGiven)
        tile_sweep<T, 1, SP>(C, A, W, R, Shat, nullptr, smem, pre, body);
    else
        tile_sweep<T, 2, SP>(C, A, W, R, S, S1, smem, pre, body);
    const double t = tile_block_sum(d2);
    if (threadIdx.x == 0) partial[wg] = t;
}

template <typename T, int CH, int SP>
__global__ SNH_TILE_OCC __launch_bounds__((tile_threads_max<T, SP>())) void k_tile_grad(Couplings<T> C, TileArgs A,
                                                                     const T* __restrict__ S,
                                                                     const T* __restrict__ phi, T* __restrict__ dE)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const TileWork W = A.work[logical_block()];
    const ReplicaInfo R = C.reps[W.r];
    SNH_TILE_SITES(T);
    auto load = [&](const TileSite& ts, int k, OwnIn<T>& in) __attribute__((always_inline)) {
        if (phi) in.p = plane_v3(phi, C.total, ts.gb + lay * (uint32_t)k);
    };
    auto site = [&](const TileSite& ts, int k, int ph, const int* offs, const OwnIn<T>& in) __attribute__((always_inline)) {
        if (!ts.ok) return;
        const uint32_t own = ts.own + slot * (uint32_t)ph;
        const uint32_t g = ts.gb + lay * (uint32_t)k;
        const int i = ts.i;
        V3<T> f = tile_field<T, CH, false>(C, offs, smem, own, i, lds_v3<T>(smem, own));
        if (phi) add_ewald_field(C, i, in.p, f);
        plane_store(dE, C.total, g, f);
    };
    tile_sweep<T, 1, SP>(
        C, A, W, R, S, nullptr, smem,
        [&](int k) __attribute__((always_inline)) {
            in0 = nx0;
            in1 = nx1;
            load(ts0, k, nx0);
            if (kTwo) load(ts1, k, nx1);
        },
        [&](int k, int ph) __attribute__((always_inline)) {
            const int* offs = A.ldsoff + ph * C.nb;
            site(ts0, k, ph, offs, in0);
            if (kTwo) site(ts1, k, ph, offs, in1);
        });
}

// energy_aux per workgroup: Zeeman, anisotropy, non-culled bonds, Ewald 1/2 mu.phi.
template <typename T, int SP>
__global__ SNH_TILE_OCC __launch_bounds__((tile_threads_max<T, SP>())) void k_tile_energy(Couplings<T> C, TileArgs A,
                                                                       const T* __restrict__ S,
                                                                       const T* __restrict__ phi,
                                                                       double* __restrict__ partial)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int wg = logical_block();
    const TileWork W = A.work[wg];
    const ReplicaInfo R = C.reps[W.r];
    SNH_TILE_SITES(T);
    double e = 0;
    auto load = [&](const TileSite& ts, int k, OwnIn<T>& in) __attribute__((always_inline)) {
        if (phi) in.p = plane_v3(phi, C.total, ts.gb + lay * (uint32_t)k);
    };
    auto site = [&](const TileSite& ts, int k, int ph, const int* offs, const OwnIn<T>& in) __attribute__((always_inline)) {
        if (!ts.ok) return;
        const uint32_t own = ts.own + slot * (uint32_t)ph;
        const int i = ts.i;
        const V3<T> s = lds_v3<T>(smem, own);
        const T* gm = C.gmat + 9 * i;
        const T* B = C.field + 3 * i;
        for (int a = 0; a < 3; a++)
            e += (double)ldc(B, a) * ((double)ldc(gm, 3 * a) * s.x + (double)ldc(gm, 3 * a + 1) * s.y +
                                      (double)ldc(gm, 3 * a + 2) * s.z);
        T ea, gx = 0, gy = 0, gz = 0;
        stevens_grad<T>(s.x, s.y, s.z, C.stev + 28 * i, ldc(C.kmax, i), gx, gy, gz, &ea);
        e += (double)ea;
        for (int b = ldc(C.bstart, i); b < ldc(C.bstart, i + 1); b++) {
            const int fl = ldc(C.bflag, b);
            if (fl & kBondCulled) continue;
            const V3<T> sj = lds_v3<T>(smem, own + (uint32_t)ldc(offs, b));
            const T* J = C.bJ + 9 * b;
            double v = (double)ldc(C.bscalar, b);
            v += (double)s.x * ((double)ldc(J, 0) * sj.x + (double)ldc(J, 1) * sj.y + (double)ldc(J, 2) * sj.z);
            v += (double)s.y * ((double)ldc(J, 3) * sj.x + (double)ldc(J, 4) * sj.y + (double)ldc(J, 5) * sj.z);
            v += (double)s.z * ((double)ldc(J, 6) * sj.x + (double)ldc(J, 7) * sj.y + (double)ldc(J, 8) * sj.z);
            if (fl & kBondBiquad) {
                T Qi[5], Qj[5];
                quadrupole(s, Qi);
                quadrupole(sj, Qj);
                const T* Bq = C.bQ + 25 * b;
                for (int p = 0; p < 5; p++)
                    for (int q = 0; q < 5; q++) v += (double)Qi[p] * (double)ldc(Bq, 5 * p + q) * (double)Qj[q];
            }
            e += v;
        }
        if (phi) {  // 1/2 mu.phi with mu = -g S
            const V3<T> p = in.p;
            const double mx = -((double)ldc(gm, 0) * s.x + (double)ldc(gm, 1) * s.y + (double)ldc(gm, 2) * s.z);
            const double my = -((double)ldc(gm, 3) * s.x + (double)ldc(gm, 4) * s.y + (double)ldc(gm, 5) * s.z);
            const double mz = -((double)ldc(gm, 6) * s.x + (double)ldc(gm, 7) * s.y + (double)ldc(gm, 8) * s.z);
            e += 0.5 * (mx * p.x + my * p.y + mz * p.z);
        }
    };
    tile_sweep<T, 1, SP>(
        C, A, W, R, S, nullptr, smem,
        [&](int k) __attribute__((always_inline)) {
            in0 = nx0;
            in1 = nx1;
            load(ts0, k, nx0);
            if (kTwo) load(ts1, k, nx1);
        },
        [&](int k, int ph) __attribute__((always_inline)) {
            const int* offs = A.ldsoff + ph * C.nb;
            site(ts0, k, ph, offs, in0);
            if (kTwo) site(ts1, k, ph, offs, in1);
        });
    const double t = tile_block_sum(e);
    if (threadIdx.x == 0) partial[wg] = t;
}

// ---------------------------------------------------------------------------
// Launchers

size_t tile_lds_bytes(const TileArgs& A, size_t elem) { return (size_t)A.nslot * A.cap * 3 * elem; }

template <typename K>
static void launch_tile(K kernel, const TileArgs&, int, size_t lds, hipStream_t)
{
    allow_dynamic_lds(reinterpret_cast<const void*>(kernel), lds);
}

// LDS reads have short latency: batches of 9 (or 6) neighbours, a divisor
// of the bond-list padding.  SP (sites per thread) comes from the plan.
#define SNH_TILE_DISPATCH(MACRO, ISO)                                  \
    do {                                                               \
        if (C.chunk % 9 == 0) {                                        \
            if (A.sp == 2) MACRO(9, ISO, 2); else MACRO(9, ISO, 1);    \
        } else {                                                       \
            if (A.sp == 2) MACRO(6, ISO, 2); else MACRO(6, ISO, 1);    \
        }                                                              \
    } while (0)

template <typename T>
void launch_tile_predict(const Couplings<T>& C, const TileArgs& A, int nthreads, const T* S, T* Sp, T* H,
                         const T* phi, const NoiseArgs<T>& nz, T dt, T lambda, hipStream_t st)
{
    const size_t lds = tile_lds_bytes(A, sizeof(T));
    const bool lean = C.lean_ok && !phi && nz.mode != 1 && !nz.step_ptr;
#define SNH_TP(CH, ISO, SP)                                                                                        \
    do {                                                                                                           \
        if (lean) {                                                                                                \
            launch_tile(k_tile_predict<T, CH, ISO, SP, true>, A, nthreads, lds, st);                               \
            hipLaunchKernelGGL((k_tile_predict<T, CH, ISO, SP, true>), dim3(A.nwork), dim3(nthreads), lds, st, C, A, \
                               S, Sp, H, phi, nz, dt, lambda);                                                     \
        } else {                                                                                                   \
            launch_tile(k_tile_predict<T, CH, ISO, SP, false>, A, nthreads, lds, st);                              \
            hipLaunchKernelGGL((k_tile_predict<T, CH, ISO, SP, false>), dim3(A.nwork), dim3(nthreads), lds, st, C,   \
                               A, S, Sp, H, phi, nz, dt, lambda);                                                  \
        }                                                                                                          \
    } while (0)
    if (C.iso) SNH_TILE_DISPATCH(SNH_TP, true); else SNH_TILE_DISPATCH(SNH_TP, false);
#undef SNH_TP
}

template <typename T>
void launch_tile_correct(const Couplings<T>& C, const TileArgs& A, int nthreads, const T* Sp, const T* H, T* Sout,
                         const T* phi, T dt, T lambda, hipStream_t st)
{
    const size_t lds = tile_lds_bytes(A, sizeof(T));
    const bool lean = C.lean_ok && !phi;
#define SNH_TC(CH, ISO, SP)                                                                                        \
    do {                                                                                                           \
        if (lean) {                                                                                                \
            launch_tile(k_tile_correct<T, CH, ISO, SP, true>, A, nthreads, lds, st);                               \
            hipLaunchKernelGGL((k_tile_correct<T, CH, ISO, SP, true>), dim3(A.nwork), dim3(nthreads), lds, st, C, A, \
                               Sp, H, Sout, phi, dt, lambda);                                                      \
        } else {                                                                                                   \
            launch_tile(k_tile_correct<T, CH, ISO, SP, false>, A, nthreads, lds, st);                              \
            hipLaunchKernelGGL((k_tile_correct<T, CH, ISO, SP, false>), dim3(A.nwork), dim3(nthreads), lds, st, C,   \
                               A, Sp, H, Sout, phi, dt, lambda);                                                   \
        }                                                                                                          \
    } while (0)
    if (C.iso) SNH_TILE_DISPATCH(SNH_TC, true); else SNH_TILE_DISPATCH(SNH_TC, false);
#undef SNH_TC
}

template <typename T>
void launch_tile_midpoint_iter(const Couplings<T>& C, const TileArgs& A, int nthreads, const T* S, const T* S1,
                               T* S2, const T* Shat, const T* phi, T dt, T lambda, double* partial,
                               const int* conv_iter, T* hat_out, const T* mu_prev, T* mu_out, hipStream_t st)
{
    const size_t lds = tile_lds_bytes(A, sizeof(T));
    const bool lean = C.lean_ok && !phi && C.iso;
#define SNH_TM(HAT, SP)                                                                                            \
    do {                                                                                                           \
        if (lean) {                                                                                                \
            launch_tile(k_tile_mid_iter<T, 6, HAT, SP, true>, A, nthreads, lds, st);                               \
            hipLaunchKernelGGL((k_tile_mid_iter<T, 6, HAT, SP, true>), dim3(A.nwork), dim3(nthreads), lds, st, C, A, \
                               S, S1, S2, Shat, phi, dt, lambda, partial, conv_iter, hat_out, mu_prev, mu_out);    \
        } else {                                                                                                   \
            launch_tile(k_tile_mid_iter<T, 6, HAT, SP, false>, A, nthreads, lds, st);                              \
            hipLaunchKernelGGL((k_tile_mid_iter<T, 6, HAT, SP, false>), dim3(A.nwork), dim3(nthreads), lds, st, C,   \
                               A, S, S1, S2, Shat, phi, dt, lambda, partial, conv_iter, hat_out, mu_prev, mu_out); \
        }                                                                                                          \
    } while (0)
    if (Shat) {
        if (A.sp == 2) SNH_TM(true, 2); else SNH_TM(true, 1);
    } else {
        if (A.sp == 2) SNH_TM(false, 2); else SNH_TM(false, 1);
    }
#undef SNH_TM
}

template <typename T>
void launch_tile_grad(const Couplings<T>& C, const TileArgs& A, int nthreads, const T* S, const T* phi, T* dE,
                      hipStream_t st)
{
    const size_t lds = tile_lds_bytes(A, sizeof(T));
    if (A.sp == 2) {
        launch_tile(k_tile_grad<T, 6, 2>, A, nthreads, lds, st);
        hipLaunchKernelGGL((k_tile_grad<T, 6, 2>), dim3(A.nwork), dim3(nthreads), lds, st, C, A, S, phi, dE);
    } else {
        launch_tile(k_tile_grad<T, 6, 1>, A, nthreads, lds, st);
        hipLaunchKernelGGL((k_tile_grad<T, 6, 1>), dim3(A.nwork), dim3(nthreads), lds, st, C, A, S, phi, dE);
    }
}

template <typename T>
void launch_tile_energy(const Couplings<T>& C, const TileArgs& A, int nthreads, const T* S, const T* phi,
                        double* partial, hipStream_t st)
{
    const size_t lds = tile_lds_bytes(A, sizeof(T));
    if (A.sp == 2) {
        launch_tile(k_tile_energy<T, 2>, A, nthreads, lds, st);
        hipLaunchKernelGGL((k_tile_energy<T, 2>), dim3(A.nwork), dim3(nthreads), lds, st, C, A, S, phi, partial);
    } else {
        launch_tile(k_tile_energy<T, 1>, A, nthreads, lds, st);
        hipLaunchKernelGGL((k_tile_energy<T, 1>), dim3(A.nwork), dim3(nthreads), lds, st, C, A, S, phi, partial);
    }
}

#define SNH_TILE_INST(T)                                                                                         \
    template void launch_tile_predict<T>(const Couplings<T>&, const TileArgs&, int, const T*, T*, T*, const T*,   \
                                         const NoiseArgs<T>&, T, T, hipStream_t);                                \
    template void launch_tile_correct<T>(const Couplings<T>&, const TileArgs&, int, const T*, const T*, T*,       \
                                         const T*, T, T, hipStream_t);                                           \
    template void launch_tile_midpoint_iter<T>(const Couplings<T>&, const TileArgs&, int, const T*, const T*, T*, \
                                               const T*, const T*, T, T, double*, const int*, T*, const T*, T*, \
                                               hipStream_t);                                                     \
    template void launch_tile_grad<T>(const Couplings<T>&, const TileArgs&, int, const T*, const T*, T*,          \
                                      hipStream_t);                                                              \
    template void launch_tile_energy<T>(const Couplings<T>&, const TileArgs&, int, const T*, const T*, double*,   \
                                        hipStream_t);

SNH_TILE_INST(float)
SNH_TILE_INST(double)

}  // namespace snh
