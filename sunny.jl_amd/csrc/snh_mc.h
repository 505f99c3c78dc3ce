// Device helpers of the Metropolis colour passes (snh_mc.hip for :dipole,
// snh_sun.hip for :SUN).  Draws of site s in pass P: Philox4x32-10 of the
// counters (s, kMcTagNormals + 2 b, P), b = 0, 1, ... (normals) and
// (s, kMcTagUniforms, P) (u_acc, u_mix, u_try).
#pragma once
#include "snh_internal.h"

namespace snh {

constexpr uint32_t kMcTagNormals = 0x4D430000u;
constexpr uint32_t kMcTagUniforms = 0x4D430001u;
constexpr uint32_t kMcTagSequential = 0x4D430002u;  // and + 1: sequential trials (Ewald)

__device__ __forceinline__ int mc_wrap(int v, int L)
{
    const int m = v % L;
    return m < 0 ? m + L : m;
}

// Classical Stevens energy in FP64 from the T-typed coefficients.
template <typename T>
__device__ __forceinline__ double stevens_energy_d(const Couplings<T>& C, int i, double x, double y, double z)
{
    const int km = ldc(C.kmax, i);
    double c[28];
#pragma unroll
    for (int p = 0; p < 28; p++) c[p] = (double)ldc(C.stev + 28 * i, p);
    double gx = 0, gy = 0, gz = 0, E = 0;
    stevens_grad<double>(x, y, z, c, km, gx, gy, gz, &E);
    return E;
}

__device__ __forceinline__ void quadrupole_d(double x, double y, double z, double Q[5])
{
    Q[0] = x * x - y * y;
    Q[1] = z * x;
    Q[2] = -x * x - y * y + 2 * z * z;
    Q[3] = z * y;
    Q[4] = 2 * y * x;
}

// Sum over a kBlock workgroup of four doubles per thread; thread 0 returns it.
__device__ __forceinline__ void block_sum4(double v[4])
{
    __shared__ double red[kBlock / 64][4];
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_down(v[q], o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
        for (int q = 0; q < 4; q++) red[w][q] = v[q];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; q++) {
            double t = 0;
            for (int k = 0; k < kBlock / 64; k++) t += red[k][q];
            v[q] = t;
        }
}

// Periodic wrap of c + n for 0 <= c < L: without a division when |n| <= L
// (every bond of the handle, unless `wide`).
__device__ __forceinline__ int mc_wrap_to(int v, int L, bool wide)
{
    if (wide) return mc_wrap(v, L);
    v = v >= L ? v - L : v;
    return v < 0 ? v + L : v;
}

// Site (cell c1, c2, c3; sublattice i) of replica R as the gather kernels
// see it: padded index and whether it has ghost images.
__device__ __forceinline__ SiteCoord mc_site(const ReplicaInfo& R, const Ghost& G, int i, int c1, int c2, int c3,
                                             bool valid)
{
    SiteCoord sc;
    sc.blk = 0;
    sc.r = 0;
    sc.i = i;
    sc.R = R;
    sc.c1 = c1;
    sc.c2 = c2;
    sc.c3 = c3;
    sc.c = c1 + R.L1 * (c2 + R.L2 * c3);
    sc.valid = valid;
    sc.g = R.off + (uint32_t)((c1 + G.w1) + R.P1 * ((c2 + G.w2) + R.P2 * (c3 + G.w3))) + R.pcell * (uint32_t)i;
    sc.edge = (c1 < G.w1) | (c1 >= R.L1 - G.w1) | (c2 < G.w2) | (c2 >= R.L2 - G.w2) | (c3 < G.w3) |
              (c3 >= R.L3 - G.w3);
    return sc;
}

// Site `lane` of work item W: kBlock consecutive sites q of class
// W.cls = rho1 + m1 (rho2 + m2 (rho3 + m3 i)) of replica W.r, at cell
// rho + m * (q1, q2, q3).
template <typename T>
__device__ __forceinline__ SiteCoord mc_work_site(const Couplings<T>& C, const McArgs& A, const McWork& W, int lane)
{
    const ReplicaInfo R = C.reps[W.r];
    const int M = A.m1 * A.m2 * A.m3;
    const int i = W.cls / M;
    int rho = W.cls - i * M;
    const int r1 = rho % A.m1;
    rho /= A.m1;
    const int r2 = rho % A.m2, r3 = rho / A.m2;
    const int Q1 = R.L1 / A.m1, Q2 = R.L2 / A.m2, nq = R.ncell / M;
    const int q = W.q0 + lane;
    const int qq = q < nq ? q : 0;
    const int t = qq / Q1;
    const int q1 = qq - t * Q1, q3 = t / Q2, q2 = t - q3 * Q2;
    SiteCoord sc = mc_site(R, C.G, i, r1 + A.m1 * q1, r2 + A.m2 * q2, r3 + A.m3 * q3, q < nq);
    sc.r = W.r;
    return sc;
}

// Proposal of the mixture for the uniform u_mix (the first p with
// u_mix < cum[p], else the last), as @mix_proposals picks it.
__device__ __forceinline__ int mc_pick(const McArgs& A, double u_mix)
{
    int k = A.nprop - 1;
    for (int p = 0; p < A.nprop - 1; p++)
        if (u_mix < A.cum[p]) {
            k = p;
            break;
        }
    return k;
}

// Draws of sequential trial T0 (the Ewald form) of one replica of ns sites:
// Philox4x32-10 of counter (T0 lo, kMcTagSequential, T0 hi, 0) gives the
// site (w0 ns) >> 32, u_acc = w1 / 2^32 and u_mix = w2 / 2^32; the normals
// come from the counters (T0 lo, kMcTagSequential + 1 + 2 b, T0 hi, 0).
struct McSeqDraw {
    uint32_t site;
    double u_acc, u_mix;
};

__device__ __forceinline__ McSeqDraw mc_seq_draw(unsigned long long key, unsigned long long T0, uint32_t ns)
{
    uint32_t un[4] = {(uint32_t)T0, kMcTagSequential, (uint32_t)(T0 >> 32), 0u};
    philox4x32(un, (uint32_t)key, (uint32_t)(key >> 32));
    const double inv32 = 1.0 / 4294967296.0;
    return {(uint32_t)(((unsigned long long)un[0] * ns) >> 32), (double)un[1] * inv32, (double)un[2] * inv32};
}

// ewald_energy_delta (Ewald.jl:234-241) of sublattice-i site g whose dipole
// changes by dS: dS . (-g^T phi) + dmu . A(0, i, i) dmu / 2 with
// dmu = -g dS (returned in dmu), phi the site's current FP64 potential and
// A0 the [na][9] self blocks.
template <typename T>
__device__ __forceinline__ double mc_ewald_delta(const Couplings<T>& C, int i, const double* __restrict__ phi,
                                                 uint32_t g, const double* __restrict__ A0, const double dS[3],
                                                 double dmu[3])
{
    const uint32_t total = C.total;
    const double p[3] = {phi[g], phi[g + total], phi[g + 2 * (size_t)total]};
    double gm[9];
#pragma unroll
    for (int q = 0; q < 9; q++) gm[q] = (double)ldc(C.gmat + 9 * i, q);
#pragma unroll
    for (int a = 0; a < 3; a++) dmu[a] = -(gm[3 * a] * dS[0] + gm[3 * a + 1] * dS[1] + gm[3 * a + 2] * dS[2]);
    const double* a0 = A0 + 9 * i;
    double dE = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const double grad = -(gm[a] * p[0] + gm[3 + a] * p[1] + gm[6 + a] * p[2]);
        const double adm = a0[3 * a] * dmu[0] + a0[3 * a + 1] * dmu[1] + a0[3 * a + 2] * dmu[2];
        dE += dS[a] * grad + dmu[a] * adm / 2;
    }
    return dE;
}

// Padded index of the neighbour across gather slot b of the site.
__device__ __forceinline__ uint32_t mc_neighbour(const McArgs& A, const Ghost& G, const SiteCoord& sc, int b,
                                                 bool wide)
{
    const int* tb = A.bn + 4 * b;
    const ReplicaInfo& R = sc.R;
    const int a1 = mc_wrap_to(sc.c1 + ldc(tb, 0), R.L1, wide), a2 = mc_wrap_to(sc.c2 + ldc(tb, 1), R.L2, wide),
              a3 = mc_wrap_to(sc.c3 + ldc(tb, 2), R.L3, wide);
    return R.off + (uint32_t)((a1 + G.w1) + R.P1 * ((a2 + G.w2) + R.P2 * (a3 + G.w3))) +
           R.pcell * (uint32_t)ldc(tb, 3);
}

}  // namespace snh
