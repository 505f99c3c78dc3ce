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

// Change of the classical Stevens energy in FP64, from the T-typed
// coefficients, as the dipole goes from (ox, oy, oz) to (nx, ny, nz); zero
// without anisotropy (the constant term cancels), the coefficients then
// not loaded.
template <typename T>
__device__ __forceinline__ double stevens_delta_d(const Couplings<T>& C, int i, double nx, double ny, double nz,
                                                  double ox, double oy, double oz)
{
    const int km = ldc(C.kmax, i);
    if (km == 0) return 0.0;
    double c[28];
#pragma unroll
    for (int p = 0; p < 28; p++) c[p] = (double)ldc(C.stev + 28 * i, p);
    double gx = 0, gy = 0, gz = 0, En = 0, Eo = 0;
    stevens_grad<double>(nx, ny, nz, c, km, gx, gy, gz, &En);
    stevens_grad<double>(ox, oy, oz, c, km, gx, gy, gz, &Eo);
    return En - Eo;
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

// The value of v in lane l (wave-uniform l), in 32-bit pieces.
template <typename X>
__device__ __forceinline__ X lane_value(X v, int l)
{
    static_assert(sizeof(X) % 4 == 0, "32-bit pieces");
    uint32_t w[sizeof(X) / 4];
    __builtin_memcpy(w, &v, sizeof(X));
#pragma unroll
    for (int q = 0; q < (int)(sizeof(X) / 4); q++) w[q] = (uint32_t)__builtin_amdgcn_readlane((int)w[q], l);
    X out;
    __builtin_memcpy(&out, w, sizeof(X));
    return out;
}

// A sequential trial's site (sublattice i, cell c1..c3 of replica R),
// acceptance uniform and mixture component k (mc_pick of u_mix).
struct McSeqSite {
    int i, c1, c2, c3, k;
    double u_acc;
};

__device__ __forceinline__ McSeqSite mc_seq_site(const ReplicaInfo& R, const McSeqDraw& dr, const McArgs& A)
{
    McSeqSite d;
    d.i = (int)(dr.site / (uint32_t)R.ncell);
    const int c = (int)(dr.site - (uint32_t)d.i * (uint32_t)R.ncell);
    d.c1 = c % R.L1;
    d.c2 = (c / R.L1) % R.L2;
    d.c3 = c / (R.L1 * R.L2);
    d.k = mc_pick(A, dr.u_mix);
    d.u_acc = dr.u_acc;
    return d;
}

__device__ __forceinline__ McSeqSite mc_seq_site_of_lane(const McSeqSite& my, int l)
{
    McSeqSite d;
    d.i = lane_value(my.i, l);
    d.c1 = lane_value(my.c1, l);
    d.c2 = lane_value(my.c2, l);
    d.c3 = lane_value(my.c3, l);
    d.k = lane_value(my.k, l);
    d.u_acc = lane_value(my.u_acc, l);
    return d;
}

// ewald_energy_delta (Ewald.jl:234-241) of a sublattice-i site whose dipole
// changes by dS: dS . (-g^T p) + dmu . A(0, i, i) dmu / 2 with dmu = -g dS
// (returned in dmu), p the site's current FP64 potential and A0 the [na][9]
// self blocks.
template <typename T>
__device__ __forceinline__ double mc_ewald_delta(const Couplings<T>& C, int i, const double p[3],
                                                 const double* __restrict__ A0, const double dS[3], double dmu[3])
{
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

// Sum over the 64 lanes of a wavefront; every lane gets the same total (a
// butterfly: partners add the same two values, so no lane differs).
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block A(c_m - c; i_m, i) of the real-space tensor kept for sequential
// trials: [i_m][i][cell offset] blocks of six doubles xx, xy, xz, yy, yz, zz
// (the 3x3 blocks are symmetric), 48 bytes each, read as three 16-byte
// loads.  Cells c_m and c of replica R both lie in [0, L).
struct ABlock {
    double xx, xy, xz, yy, yz, zz;
};

__device__ __forceinline__ ABlock mc_ablock(const double* __restrict__ Aew, int na, const ReplicaInfo& R, int im,
                                            int cm1, int cm2, int cm3, int i, int c1, int c2, int c3)
{
    const int o1 = mc_wrap_to(cm1 - c1, R.L1, false), o2 = mc_wrap_to(cm2 - c2, R.L2, false),
              o3 = mc_wrap_to(cm3 - c3, R.L3, false);
    const size_t off = (size_t)o1 + (size_t)R.L1 * ((size_t)o2 + (size_t)R.L2 * (size_t)o3);
    const double2* p = reinterpret_cast<const double2*>(Aew + (((size_t)im * na + i) * (size_t)R.ncell + off) * 6);
    const double2 a = p[0], b = p[1], c = p[2];
    return {a.x, a.y, b.x, b.y, c.x, c.y};
}

// out = A v for a symmetric block.
__device__ __forceinline__ void ablock_apply(const ABlock& A, const double v[3], double out[3])
{
    out[0] = A.xx * v[0] + A.xy * v[1] + A.xz * v[2];
    out[1] = A.xy * v[0] + A.yy * v[1] + A.yz * v[2];
    out[2] = A.xz * v[0] + A.yz * v[1] + A.zz * v[2];
}

// s = sum over the moves mv[0..nm) of A(c_m - c; i_m, i) dmu_m at sublattice-i
// site (c1, c2, c3) of replica R (phi = sum_k A mu_k, Ewald.jl:185-209).
__device__ __forceinline__ void mc_moves_sum(const double* __restrict__ Aew, int na, const ReplicaInfo& R,
                                             const McTrial* __restrict__ mv, int nm, int i, int c1, int c2, int c3,
                                             double s[3])
{
    double a0 = 0, a1 = 0, a2 = 0;
#pragma unroll 8
    for (int m = 0; m < nm; m++) {
        const McTrial t = mv[m];
        double o[3];
        ablock_apply(mc_ablock(Aew, na, R, t.i, t.c1, t.c2, t.c3, i, c1, c2, c3), t.dmu, o);
        a0 += o[0];
        a1 += o[1];
        a2 += o[2];
    }
    s[0] = a0;
    s[1] = a1;
    s[2] = a2;
}

// Apply workgroup b of a launch (b < nr napp): phi_out = phi_in + the
// previous chunk's moves at sites j = W (b mod napp) + thread + W napp k
// (W threads per workgroup) of replica r0 + b / napp.  napp is kept small, so the apply streams the
// tensor over the chunk's whole duration rather than competing with the
// serial trials for memory at its start.
__device__ __forceinline__ void mc_apply_part(const McChunkIO& io, const double* __restrict__ Aew,
                                              const ReplicaInfo* __restrict__ reps, const Ghost& G, int na,
                                              uint32_t total, int r0, int b)
{
    const int r = r0 + b / io.napp;
    const ReplicaInfo R = reps[r];
    const uint32_t ncell = (uint32_t)R.ncell;
    const McTrial* mv = io.prev + (size_t)r * kMcChunk;
    const int nm = io.nprev[r];
    for (uint32_t j = (uint32_t)(b % io.napp) * blockDim.x + threadIdx.x; j < ncell * (uint32_t)na;
         j += blockDim.x * (uint32_t)io.napp) {
        const int i = (int)(j / ncell);
        const int c = (int)(j - (uint32_t)i * ncell);
        const int c1 = c % R.L1, c2 = (c / R.L1) % R.L2, c3 = c / (R.L1 * R.L2);
        double s[3];
        mc_moves_sum(Aew, na, R, mv, nm, i, c1, c2, c3, s);
        const uint32_t g = mc_site(R, G, i, c1, c2, c3, true).g;
#pragma unroll
        for (int q = 0; q < 3; q++) io.phi_out[g + q * (size_t)total] = io.phi_in[g + q * (size_t)total] + s[q];
    }
}

// Ewald potential a sequential trial sees, kept per lane for the chunk's
// trials (lane l: trial T0 + l at site (i, c1, c2, c3)): pot starts as the
// potential of the chunk's start at the site (phi_in plus the previous
// chunk's moves, summed as the apply workgroups sum them) and gains
// A(c_m - c; i_m, i) dmu_m for every move m accepted at an earlier trial of
// the chunk.  The block A of a move is loaded when it is accepted and folded
// into pot at the next trial, so its latency overlaps that trial's other
// loads.
struct McChunkPot {
    double pot[3];
    ABlock a;       // block of the pending move at the lane's site (zero: none)
    double dmu[3];  // the pending move's dmu (wave-uniform)
    bool pending;   // wave-uniform
};

__device__ __forceinline__ void mc_pot_start(McChunkPot& P, const McChunkIO& io, const double* __restrict__ Aew,
                                             int na, const ReplicaInfo& R, int r, uint32_t g, uint32_t total, int i,
                                             int c1, int c2, int c3, int trial)
{
    P.pot[0] = P.pot[1] = P.pot[2] = 0;
    if (io.pprev) {  // summed beforehand: trial `trial` of the chunk sits at (i, c1, c2, c3)
        const double* q = io.pprev + ((size_t)r * kMcChunk + trial) * 3;
        P.pot[0] = q[0];
        P.pot[1] = q[1];
        P.pot[2] = q[2];
    }
    for (int l = 0; io.prev && !io.pprev && l < io.nprevl; l++) {
        const size_t lr = (size_t)l * io.nrep_all + r;
        double s[3];
        mc_moves_sum(Aew, na, R, io.prev + lr * kMcChunk, io.nprev[lr], i, c1, c2, c3, s);
#pragma unroll
        for (int q = 0; q < 3; q++) P.pot[q] += s[q];
    }
#pragma unroll
    for (int q = 0; q < 3; q++) P.pot[q] = io.phi_in[g + q * (size_t)total] + P.pot[q];
    P.pending = false;
}

// Move (im; cm1, cm2, cm3; dmu) accepted at trial t: lanes t < l < ntrial
// load their block.
__device__ __forceinline__ void mc_pot_accept(McChunkPot& P, const double* __restrict__ Aew, int na,
                                              const ReplicaInfo& R, int t, int ntrial, int im, int cm1, int cm2,
                                              int cm3, const double dmu[3], int i, int c1, int c2, int c3)
{
    const int l = (int)threadIdx.x;
    if (l > t && l < ntrial)
        P.a = mc_ablock(Aew, na, R, im, cm1, cm2, cm3, i, c1, c2, c3);
    else
        P.a = {0, 0, 0, 0, 0, 0};
    P.dmu[0] = dmu[0];
    P.dmu[1] = dmu[1];
    P.dmu[2] = dmu[2];
    P.pending = true;
}

// Folds the pending move into pot and returns trial t's potential (every
// lane gets it).
__device__ __forceinline__ void mc_pot_take(McChunkPot& P, int t, double p[3])
{
    if (P.pending) {
        double o[3];
        ablock_apply(P.a, P.dmu, o);
#pragma unroll
        for (int q = 0; q < 3; q++) P.pot[q] += o[q];
        P.pending = false;
    }
    p[0] = lane_value(P.pot[0], t);
    p[1] = lane_value(P.pot[1], t);
    p[2] = lane_value(P.pot[2], t);
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
