// Device-side building blocks shared by the gfx950 kernels of libsunnyhip.
//
// State layout in HBM: structure of arrays, one plane per real component
// (dipoles: x, y, z; coherent states: re/im of each of the N components).
// Each replica occupies a ghost-padded box of (L1+2w1)(L2+2w2)(L3+2w3) cells
// per sublattice, where w is the largest bond offset along each axis.  The
// padded site index of (cell c, sublattice i) is
//     g = off_r + (c1+w1) + P1*((c2+w2) + P2*(c3+w3)) + Pcell*i,
// with the cell order of Sunny's column-major Array{Vec3,4}(L1,L2,L3,na)
// (src/System/Types.jl:131-136).  A bond (i -> j, n) is then the constant
// offset n1 + P1 n2 + P1 P2 n3 + Pcell (j - i) from every interior site: the
// periodic wrap of bonded_site (src/System/System.jl:204-212) is paid once,
// when a kernel writes a site, by also writing its ghost images.
// Consecutive lanes hold consecutive cells of one sublattice, so every plane
// access of the gather is a coalesced 64-lane load with a 32-bit offset from
// an SGPR base.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace snh {

constexpr int kBlock = 256;
// Persistent small-replica kernels: one workgroup of kPersist threads per
// replica, kPersistLanes blocks of kBlock sites in flight.
constexpr int kPersist = 512;
constexpr int kPersistLanes = kPersist / kBlock;

// One work block = kBlock consecutive interior cells of one (replica, sublattice).
struct BlockInfo {
    int r;
    int i;
    int c0;
    int pad;
};

struct ReplicaInfo {
    int L1, L2, L3, ncell;  // interior dims
    int P1, P2, P3;         // padded dims
    int dcls;               // dims class (index into the bond-offset table)
    uint32_t pcell;         // P1 P2 P3
    uint32_t off;           // first padded site of the replica
    float inv1, inv2;       // 1/L1, 1/L2 (rounded up) for division-free cell coordinates
};

// q = n / d for 0 <= n < 2^24 and 1 <= d: float estimate from a reciprocal
// rounded up, then one exact integer correction.
__device__ __forceinline__ int fdiv(int n, int d, float inv)
{
    int q = (int)((float)n * inv);
    const int r = n - q * d;
    q += (r >= d) - (r < 0);
    return q;
}

struct Ghost {
    int w1, w2, w3;
    int general;  // some w exceeds L: ghost images may wrap more than once
};

// 32-bit element offsets from a base pointer (sites < 2^32 / sizeof(T)).
template <typename T>
__device__ __forceinline__ T ldg(const T* __restrict__ base, uint32_t idx)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + idx * (uint32_t)sizeof(T));
}

template <typename T>
__device__ __forceinline__ void stg(T* __restrict__ base, uint32_t idx, T v)
{
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + idx * (uint32_t)sizeof(T)) = v;
}

// Load from a read-only table through the constant address space: the
// compiler may then keep a wavefront-uniform index in scalar registers and
// use scalar loads even where the kernel also stores to global memory.
template <typename X>
__device__ __forceinline__ X ldc(const X* p, int i)
{
    typedef __attribute__((address_space(4))) const X* cptr;
    return ((cptr)p)[i];
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on
// XCD b % 8), each with its own L2.  Remap so that every XCD walks a
// contiguous range of logical blocks: blocks that share neighbour data then
// share an L2.
constexpr int kXcds = 8;

__device__ __forceinline__ int logical_block()
{
    const int b = blockIdx.x, n = gridDim.x;
    const int q = n / kXcds, rem = n % kXcds;
    const int x = b % kXcds, idx = b / kXcds;
    return x < rem ? x * (q + 1) + idx : rem * (q + 1) + (x - rem) * q + idx;
}

struct SiteCoord {
    int blk;  // logical block
    int r, i, c, c1, c2, c3;
    uint32_t g;  // padded global index
    bool valid;
    bool edge;   // has ghost images
    ReplicaInfo R;
};

// Site of lane `lane` (0..kBlock-1) in logical block `blk`.
__device__ __forceinline__ SiteCoord site_coord_at(const BlockInfo* __restrict__ blocks,
                                                   const ReplicaInfo* __restrict__ reps, const Ghost& G, int blk,
                                                   int lane)
{
    SiteCoord sc;
    sc.blk = blk;
    const BlockInfo bi = blocks[sc.blk];
    sc.r = bi.r;
    sc.i = bi.i;
    sc.R = reps[bi.r];
    sc.c = bi.c0 + lane;
    sc.valid = sc.c < sc.R.ncell;
    const int c = sc.valid ? sc.c : 0;
    const int t = fdiv(c, sc.R.L1, sc.R.inv1);
    sc.c1 = c - t * sc.R.L1;
    sc.c3 = fdiv(t, sc.R.L2, sc.R.inv2);
    sc.c2 = t - sc.c3 * sc.R.L2;
    sc.g = sc.R.off + (uint32_t)((sc.c1 + G.w1) + sc.R.P1 * ((sc.c2 + G.w2) + sc.R.P2 * (sc.c3 + G.w3))) +
           sc.R.pcell * (uint32_t)sc.i;
    sc.edge = (sc.c1 < G.w1) | (sc.c1 >= sc.R.L1 - G.w1) | (sc.c2 < G.w2) | (sc.c2 >= sc.R.L2 - G.w2) |
              (sc.c3 < G.w3) | (sc.c3 >= sc.R.L3 - G.w3);
    return sc;
}

__device__ __forceinline__ SiteCoord site_coord(const BlockInfo* __restrict__ blocks,
                                                const ReplicaInfo* __restrict__ reps, const Ghost& G)
{
    return site_coord_at(blocks, reps, G, logical_block(), (int)threadIdx.x);
}

// Range of image shifts m (in units of L) with 0 <= c + w + m L < L + 2w.
__device__ __forceinline__ void image_range(int c, int L, int w, bool general, int& lo, int& hi)
{
    if (!general) {
        lo = (c >= L - w) ? -1 : 0;
        hi = (c < w) ? 1 : 0;
    } else {
        lo = -((c + w) / L);
        hi = (L + w - 1 - c) / L;
    }
}

// Calls f(padded index) for every image of the site other than itself.
template <typename F>
__device__ __forceinline__ void for_each_ghost(const SiteCoord& sc, const Ghost& G, F f)
{
    if (!sc.edge) return;
    int lo1, hi1, lo2, hi2, lo3, hi3;
    const bool gen = G.general != 0;
    image_range(sc.c1, sc.R.L1, G.w1, gen, lo1, hi1);
    image_range(sc.c2, sc.R.L2, G.w2, gen, lo2, hi2);
    image_range(sc.c3, sc.R.L3, G.w3, gen, lo3, hi3);
    const int s2 = sc.R.L2 * sc.R.P1, s3 = sc.R.L3 * sc.R.P1 * sc.R.P2;
    for (int m3 = lo3; m3 <= hi3; m3++)
        for (int m2 = lo2; m2 <= hi2; m2++)
            for (int m1 = lo1; m1 <= hi1; m1++) {
                if (m1 == 0 && m2 == 0 && m3 == 0) continue;
                f(sc.g + (uint32_t)(m1 * sc.R.L1 + m2 * s2 + m3 * s3));
            }
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) + Box-Muller.
//
// Every caller's key must be uniform over its wave (callers keep one replica
// per workgroup, or run one lane): the key is taken from the first active
// lane (readfirstlane) and lives in scalar registers.  The round keys are advanced in place behind an
// empty asm barrier: otherwise the compiler precomputes all twenty of them
// outside the caller's loops, runs out of scalar registers and reloads them
// from vector-register lanes in every round.
__device__ __forceinline__ void philox4x32(uint32_t c[4], uint32_t k0, uint32_t k1)
{
    k0 = __builtin_amdgcn_readfirstlane(k0);
    k1 = __builtin_amdgcn_readfirstlane(k1);
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = (uint32_t)p1;
        c[2] = n2;
        c[3] = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
        asm volatile("" : "+s"(k0), "+s"(k1));
    }
}

// Box-Muller pair: FP64 with libm-grade functions (parity path); FP32 with
// the hardware transcendental units (v_log_f32, v_sin_f32, v_cos_f32, which
// take revolutions).
__device__ __forceinline__ void box_muller(uint32_t a, uint32_t b, double& x, double& y)
{
    const double ua = ((double)a + 1.0) * (1.0 / 4294967296.0);
    const double ub = (double)b * (1.0 / 4294967296.0);
    const double r = sqrt(-2.0 * log(ua));
    double s, c;
    sincospi(2.0 * ub, &s, &c);
    x = r * c;
    y = r * s;
}

__device__ __forceinline__ void box_muller(uint32_t a, uint32_t b, float& x, float& y)
{
    const float ua = ((float)a + 1.0f) * (1.0f / 4294967296.0f);
    const float ub = (float)b * (1.0f / 4294967296.0f);
    // ln(u) = log2(u) * ln 2
    const float r = __builtin_sqrtf(-2.0f * 0.69314718055994531f * __builtin_amdgcn_logf(ua));
    x = r * __builtin_amdgcn_cosf(ub);
    y = r * __builtin_amdgcn_sinf(ub);
}

// Four standard normals for (site, block, step) under key = seed.
template <typename T>
__device__ __forceinline__ void philox_normals4(uint64_t seed, uint64_t step, uint32_t site, uint32_t blk, T out[4])
{
    uint32_t c[4] = {site, blk, (uint32_t)step, (uint32_t)(step >> 32)};
    philox4x32(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    box_muller(c[0], c[1], out[0], out[1]);
    box_muller(c[2], c[3], out[2], out[3]);
}

// ---------------------------------------------------------------------------
// Classical Stevens anisotropy gradient (and energy), the closed form of
// Sunny's OnsiteCoupling.jl:316-425.  c = c0 | c2[5] | c4[9] | c6[13].
// As in the reference, X = |S|^2 is held constant in the gradient.  Tc, the
// coefficients' type, may differ from T (forward-mode dual numbers in T).
template <typename T, typename Tc = T>
__device__ __forceinline__ void stevens_grad(T sx, T sy, T sz, const Tc* __restrict__ c, int kmax, T& gx, T& gy,
                                             T& gz, T* energy)
{
    T E = c[0];
    T dz = 0, jr = 0, ji = 0;  // dE/dz, dE/dJp (re, im)
    if (kmax > 0) {
        const Tc* c2 = c + 1;
        T X = sx * sx + sy * sy + sz * sz;
        T p1r = sx, p1i = sy;
        T z1 = sz, z2 = sz * sz;
        T p2r = p1r * p1r - p1i * p1i, p2i = 2 * p1r * p1i;
        T A1 = 3 * z2 - X;
        E += c2[0] * p2r + c2[4] * p2i + (c2[1] * p1r + c2[3] * p1i) * z1 + c2[2] * A1;
        dz += (c2[1] * p1r + c2[3] * p1i) + c2[2] * 6 * z1;
        // (c2[0] - i c2[4]) Jp + (1/2)(c2[1] - i c2[3]) z
        jr += c2[0] * p1r + c2[4] * p1i + (T)0.5 * c2[1] * z1;
        ji += c2[0] * p1i - c2[4] * p1r - (T)0.5 * c2[3] * z1;
        if (kmax > 2) {
            const Tc* c4 = c + 6;
            T X2 = X * X;
            T z3 = z2 * z1, z4 = z2 * z2;
            T p3r = p2r * p1r - p2i * p1i, p3i = p2r * p1i + p2i * p1r;
            T p4r = p2r * p2r - p2i * p2i, p4i = 2 * p2r * p2i;
            T B1 = 35 * z4 - 30 * X * z2 + 3 * X2, B2 = 7 * z3 - 3 * X * z1, B3 = 7 * z2 - X, B4 = z1;
            T dB1 = 140 * z3 - 60 * X * z1, dB2 = 21 * z2 - 3 * X, dB3 = 14 * z1;
            E += c4[0] * p4r + c4[8] * p4i + (c4[1] * p3r + c4[7] * p3i) * B4 + (c4[2] * p2r + c4[6] * p2i) * B3 +
                 (c4[3] * p1r + c4[5] * p1i) * B2 + c4[4] * B1;
            dz += (c4[1] * p3r + c4[7] * p3i) + (c4[2] * p2r + c4[6] * p2i) * dB3 +
                  (c4[3] * p1r + c4[5] * p1i) * dB2 + c4[4] * dB1;
            // sum_m (m/2) (c_m - i c_-m) Jp^(m-1) A_m
            jr += 2 * (c4[0] * p3r + c4[8] * p3i) + (T)1.5 * (c4[1] * p2r + c4[7] * p2i) * B4 +
                  (c4[2] * p1r + c4[6] * p1i) * B3 + (T)0.5 * c4[3] * B2;
            ji += 2 * (c4[0] * p3i - c4[8] * p3r) + (T)1.5 * (c4[1] * p2i - c4[7] * p2r) * B4 +
                  (c4[2] * p1i - c4[6] * p1r) * B3 - (T)0.5 * c4[5] * B2;
            if (kmax > 4) {
                const Tc* c6 = c + 15;
                T X3 = X2 * X;
                T z5 = z4 * z1, z6 = z3 * z3;
                T p5r = p4r * p1r - p4i * p1i, p5i = p4r * p1i + p4i * p1r;
                T p6r = p3r * p3r - p3i * p3i, p6i = 2 * p3r * p3i;
                T C1 = 231 * z6 - 315 * X * z4 + 105 * X2 * z2 - 5 * X3;
                T C2 = 33 * z5 - 30 * X * z3 + 5 * X2 * z1;
                T C3 = 33 * z4 - 18 * X * z2 + X2;
                T C4 = 11 * z3 - 3 * X * z1;
                T C5 = 11 * z2 - X;
                T C6 = z1;
                T dC1 = 1386 * z5 - 1260 * X * z3 + 210 * X2 * z1;
                T dC2 = 165 * z4 - 90 * X * z2 + 5 * X2;
                T dC3 = 132 * z3 - 36 * X * z1;
                T dC4 = 33 * z2 - 3 * X;
                T dC5 = 22 * z1;
                E += c6[0] * p6r + c6[12] * p6i + (c6[1] * p5r + c6[11] * p5i) * C6 +
                     (c6[2] * p4r + c6[10] * p4i) * C5 + (c6[3] * p3r + c6[9] * p3i) * C4 +
                     (c6[4] * p2r + c6[8] * p2i) * C3 + (c6[5] * p1r + c6[7] * p1i) * C2 + c6[6] * C1;
                dz += (c6[1] * p5r + c6[11] * p5i) + (c6[2] * p4r + c6[10] * p4i) * dC5 +
                      (c6[3] * p3r + c6[9] * p3i) * dC4 + (c6[4] * p2r + c6[8] * p2i) * dC3 +
                      (c6[5] * p1r + c6[7] * p1i) * dC2 + c6[6] * dC1;
                jr += 3 * (c6[0] * p5r + c6[12] * p5i) + (T)2.5 * (c6[1] * p4r + c6[11] * p4i) * C6 +
                      2 * (c6[2] * p3r + c6[10] * p3i) * C5 + (T)1.5 * (c6[3] * p2r + c6[9] * p2i) * C4 +
                      (c6[4] * p1r + c6[8] * p1i) * C3 + (T)0.5 * c6[5] * C2;
                ji += 3 * (c6[0] * p5i - c6[12] * p5r) + (T)2.5 * (c6[1] * p4i - c6[11] * p4r) * C6 +
                      2 * (c6[2] * p3i - c6[10] * p3r) * C5 + (T)1.5 * (c6[3] * p2i - c6[9] * p2r) * C4 +
                      (c6[4] * p1i - c6[8] * p1r) * C3 - (T)0.5 * c6[7] * C2;
            }
        }
    }
    gx += 2 * jr;
    gy += -2 * ji;
    gz += dz;
    if (energy) *energy = E;
}

// Block sum of one double per thread (wavefront shuffles, then LDS).
__device__ __forceinline__ double block_sum(double v)
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

}  // namespace snh
