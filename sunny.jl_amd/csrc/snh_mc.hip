// Metropolis updates of single dipoles (:dipole mode) on the device; the
// :SUN kernel (k_sun_mc_color, snh_sun.hip) shares the scheme.
//
// Reference semantics (Sunny.jl 0.9.2):
//   LocalSampler, step!                       src/MonteCarlo/Samplers.jl:133-153
//   propose_uniform / propose_flip / propose_delta / @mix_proposals
//                                             src/MonteCarlo/Samplers.jl:1-100
//   randspin, flip, perturbed_spin            src/System/System.jl:440-468
//   local_energy_change                       src/System/Interactions.jl:311-380
//
// The reference tries ceil(nsweeps * nsites) sites drawn at random, one at a
// time.  Here a sweep is two passes; a pass visits the colour classes in
// turn and tries each site of a class with probability 1/2.  The host colours
// the classes (sublattice i, c mod m) so that no bond joins two sites of one
// class, hence all sites of a class update in parallel with exactly the
// energy change a sequential update would see.  Each trial is a Metropolis
// step of the reference's proposal and acceptance rule, so the Boltzmann
// distribution stays stationary, and a call makes nsweeps * nsites trials on
// average.  The random subsets matter: with every site tried in a fixed
// order, Metropolis flips of Ising-like spins move the domain walls of a
// chain deterministically (moves with dE = 0 are always accepted) and the
// sweeps are not ergodic.
//
// Draws of site s (index within its replica, Julia order) in pass P come
// from Philox4x32-10 under key = seed + r * 0x9E3779B97F4A7C15:
//   counter (s, 0x4D430000, P) -> four normals (Box-Muller; three are used)
//   counter (s, 0x4D430001, P) -> u_acc, u_mix, u_try (word / 2^32)
// disjoint from the Langevin stream (second counter word 0).
//
// With the Ewald term every pair of sites interacts, so no two trials are
// independent: trials then keep the reference's order, in chunks of up to
// kMcChunk per launch (k_mc_ewald_chunk): each trial sees the FP64 Ewald
// potential of the chunk's start plus the moves accepted since, the chunk's
// trials evaluated in parallel and decided in order with exact corrections.
// The same launch's other workgroups add the previous chunk's moves to the
// potential at every site of the replica (mc_apply_part, O(N) per accepted
// move as in Ewald.jl:212-241) for the next launch, so that work overlaps
// the chunk's trials.
#include "snh_mc.h"
#include "snh_vec.h"

namespace snh {

// The three normals of a proposal: Philox4x32-10 of counter ctr, two
// Box-Muller pairs (four normals, three used).
template <typename T>
__device__ __forceinline__ void mc_normals(uint32_t ctr[4], unsigned long long key, T n[3])
{
    philox4x32(ctr, (uint32_t)key, (uint32_t)(key >> 32));
    T m[4];
    box_muller(ctr[0], ctr[1], m[0], m[1]);
    box_muller(ctr[2], ctr[3], m[2], m[3]);
    n[0] = m[0];
    n[1] = m[1];
    n[2] = m[2];
}

// Proposal k of the mixture for a site holding s0: propose_flip, randspin
// or perturbed_spin (S + mag kappa randn) from the normals n.
template <typename T>
__device__ __forceinline__ V3<T> mc_propose_from(const McArgs& A, int k, V3<T> s0, T kap, const T n[3])
{
    if (A.kind[k] == 1) return {-s0.x, -s0.y, -s0.z};  // propose_flip
    const double mag = A.mag[k];
    if (A.kind[k] == 0 || __builtin_isinf(mag))  // randspin
        return normalize_dipole(V3<T>{n[0], n[1], n[2]}, kap);
    const T f = (T)(mag * (double)kap);  // perturbed_spin
    return normalize_dipole(V3<T>{s0.x + f * n[0], s0.y + f * n[1], s0.z + f * n[2]}, kap);
}

// Proposal of LocalSampler for a site holding s0, chosen by u_mix; normals
// from counter ctr (not drawn for propose_flip).
template <typename T>
__device__ __forceinline__ V3<T> mc_propose(const McArgs& A, V3<T> s0, T kap, double u_mix, uint32_t ctr[4],
                                            unsigned long long key)
{
    const int k = mc_pick(A, u_mix);
    T n[3] = {0, 0, 0};
    if (A.kind[k] != 1) mc_normals<T>(ctr, key, n);
    return mc_propose_from<T>(A, k, s0, kap, n);
}

// local_energy_change in FP64 (Interactions.jl:311-380) of site sc going
// from s0 to sn: Zeeman, Stevens, bilinear and biquadratic bonds; neighbour
// states gathered G bonds at a time.
template <typename T, int G>
__device__ __forceinline__ double mc_local_delta(const Couplings<T>& C, const McArgs& A, const T* __restrict__ S,
                                                 const SiteCoord& sc, V3<T> s0, V3<T> sn, bool wide)
{
    const int i = sc.i;
    const uint32_t total = C.total;
    const double ox = s0.x, oy = s0.y, oz = s0.z, nx = sn.x, ny = sn.y, nz = sn.z;
    const double dx = nx - ox, dy = ny - oy, dz = nz - oz;
    double dE = (double)ldc(C.zee, 3 * i) * dx + (double)ldc(C.zee, 3 * i + 1) * dy +
                (double)ldc(C.zee, 3 * i + 2) * dz;
    dE += stevens_delta_d(C, i, nx, ny, nz, ox, oy, oz);
    double Qn[5], Qo[5];
    quadrupole_d(nx, ny, nz, Qn);
    quadrupole_d(ox, oy, oz, Qo);
    const int b0 = ldc(C.bstart, i), b1 = ldc(C.bstart, i + 1);
    for (int c0 = b0; c0 < b1; c0 += G) {
        T jx[G], jy[G], jz[G];
#pragma unroll
        for (int kk = 0; kk < G; kk++) {
            const uint32_t gj = mc_neighbour(A, C.G, sc, c0 + kk, wide);
            jx[kk] = ldg(S, gj);
            jy[kk] = ldg(S, gj + total);
            jz[kk] = ldg(S, gj + 2 * total);
        }
#pragma unroll
        for (int kk = 0; kk < G; kk++) {
            const int b = c0 + kk;
            const double x = jx[kk], y = jy[kk], z = jz[kk];
            const T* J = C.bJ + 9 * b;
            dE += dx * ((double)ldc(J, 0) * x + (double)ldc(J, 1) * y + (double)ldc(J, 2) * z) +
                  dy * ((double)ldc(J, 3) * x + (double)ldc(J, 4) * y + (double)ldc(J, 5) * z) +
                  dz * ((double)ldc(J, 6) * x + (double)ldc(J, 7) * y + (double)ldc(J, 8) * z);
            if (ldc(C.bflag, b) & kBondBiquad) {
                double Qj[5];
                quadrupole_d(x, y, z, Qj);
                const T* Bq = C.bQ + 25 * b;
                for (int p = 0; p < 5; p++) {
                    double w = 0;
                    for (int qv = 0; qv < 5; qv++) w += (double)ldc(Bq, 5 * p + qv) * Qj[qv];
                    dE += (Qn[p] - Qo[p]) * w;
                }
            }
        }
    }
    return dE;
}

// Writes sn at site sc and its ghost images.
template <typename T>
__device__ __forceinline__ void mc_store(T* __restrict__ S, const Couplings<T>& C, const SiteCoord& sc, V3<T> sn)
{
    const uint32_t total = C.total;
    stg(S, sc.g, sn.x);
    stg(S, sc.g + total, sn.y);
    stg(S, sc.g + 2 * total, sn.z);
    for_each_ghost(sc, C.G, [&](uint32_t gi) {
        stg(S, gi, sn.x);
        stg(S, gi + total, sn.y);
        stg(S, gi + 2 * total, sn.z);
    });
}

// One colour of pass P: workgroup b takes work[b] = kBlock consecutive sites
// q of one class (replica r, sublattice i, cell residue rho mod m), cell
// c = rho + m * (q1, q2, q3); each site is tried with probability prob.
// acc[4 b + q] += (dE, dS) of the workgroup's accepted trials.  Neighbour
// states are gathered G bonds at a time (the bond lists are padded with
// zero couplings of the site to itself to a multiple of G).
template <typename T, int G>
__global__ __launch_bounds__(kBlock) void k_mc_color(Couplings<T> C, McArgs A, T* __restrict__ S,
                                                     const McWork* __restrict__ work, unsigned long long P,
                                                     double prob, double* __restrict__ acc)
{
    const McWork W = work[blockIdx.x];
    const SiteCoord sc = mc_work_site(C, A, W, (int)threadIdx.x);
    const ReplicaInfo& R = sc.R;
    const int i = sc.i;
    const bool wide = A.wide != 0;
    double v[4] = {0, 0, 0, 0};
    const uint32_t total = C.total;
    const unsigned long long key = A.seed + (unsigned long long)W.r * 0x9E3779B97F4A7C15ull;
    const uint32_t site = (uint32_t)sc.c + (uint32_t)R.ncell * (uint32_t)i;
    uint32_t un[4] = {site, kMcTagUniforms, (uint32_t)P, (uint32_t)(P >> 32)};
    philox4x32(un, (uint32_t)key, (uint32_t)(key >> 32));
    const double inv32 = 1.0 / 4294967296.0;
    const double u_acc = (double)un[0] * inv32, u_mix = (double)un[1] * inv32, u_try = (double)un[2] * inv32;
    if (sc.valid && u_try < prob) {
        const V3<T> s0 = {ldg(S, sc.g), ldg(S, sc.g + total), ldg(S, sc.g + 2 * total)};
        uint32_t cn[4] = {site, kMcTagNormals, (uint32_t)P, (uint32_t)(P >> 32)};
        const V3<T> sn = mc_propose<T>(A, s0, ldc(C.kappa, i), u_mix, cn, key);
        const double dE = mc_local_delta<T, G>(C, A, S, sc, s0, sn, wide);
        const double kT = ldc(A.kT, W.r);
        const bool accept = kT == 0.0 ? dE <= 0.0 : u_acc <= exp(-dE / kT);
        if (accept) {
            mc_store(S, C, sc, sn);
            v[0] = dE;
            v[1] = (double)sn.x - (double)s0.x;
            v[2] = (double)sn.y - (double)s0.y;
            v[3] = (double)sn.z - (double)s0.z;
        }
    }
    block_sum4(v);
    if (threadIdx.x == 0)
        for (int qv = 0; qv < 4; qv++) acc[4 * (size_t)blockIdx.x + qv] += v[qv];
}

// Bond b's term of local_energy_change (Interactions.jl:341-368) for a
// site changing by (dx, dy, dz) and quadrupole dQ, neighbour state (x, y, z).
template <typename T>
__device__ __forceinline__ double mc_bond_term(const Couplings<T>& C, int b, double dx, double dy, double dz,
                                               const double dQ[5], double x, double y, double z)
{
    const T* J = C.bJ + 9 * b;
    double e = dx * ((double)J[0] * x + (double)J[1] * y + (double)J[2] * z) +
               dy * ((double)J[3] * x + (double)J[4] * y + (double)J[5] * z) +
               dz * ((double)J[6] * x + (double)J[7] * y + (double)J[8] * z);
    if (C.bflag[b] & kBondBiquad) {
        double Qj[5];
        quadrupole_d(x, y, z, Qj);
        const T* Bq = C.bQ + 25 * b;
        for (int p = 0; p < 5; p++) {
            double w = 0;
            for (int qv = 0; qv < 5; qv++) w += (double)Bq[5 * p + qv] * Qj[qv];
            e += dQ[p] * w;
        }
    }
    return e;
}

// Canonical index i ncell + c of the site across bond b (wrapped
// periodically) from cell (c1, c2, c3) of replica R.
__device__ __forceinline__ int mc_bond_site_id(const McArgs& A, const ReplicaInfo& R, int b, int c1, int c2, int c3)
{
    const int* tb = A.bn + 4 * b;
    const int a1 = mc_wrap(c1 + tb[0], R.L1), a2 = mc_wrap(c2 + tb[1], R.L2), a3 = mc_wrap(c3 + tb[2], R.L3);
    return tb[3] * R.ncell + a1 + R.L1 * (a2 + R.L2 * a3);
}

// dmu_l . A dmu_m for the block A = A(c_m - c_l; i_m, i_l): the Ewald energy
// between two moved dipoles (Ewald.jl:234-241 with phi = sum_k A mu_k,
// Ewald.jl:185-209).
__device__ __forceinline__ double mc_pair_energy(const ABlock& a, const double dl[3], const double dm[3])
{
    double o[3];
    ablock_apply(a, dm, o);
    return dl[0] * o[0] + dl[1] * o[1] + dl[2] * o[2];
}

// Neighbour ids kept per trial in LDS; bonds beyond are found when needed.
constexpr int kMcNbr = 32;

// One sequential trial made by the whole wavefront on the current state:
// site d, normals n, Ewald potential p at the site; the bond gathers split
// over the lanes.  Returns dE (local_energy_change + ewald_energy_delta,
// Interactions.jl:311-380, Ewald.jl:234-241) and sets sn, dS, dmu; every
// lane gets the same values.
template <typename T>
__device__ __forceinline__ double mc_trial_wave(const Couplings<T>& C, const McArgs& A, const T* __restrict__ S,
                                                const SiteCoord& sc, const McSeqSite& d, const T n[3],
                                                const double p[3], const double* __restrict__ A0, bool wide,
                                                V3<T>& sn, double dS[3], double dmu[3])
{
    const uint32_t total = C.total;
    const int i = d.i;
    const V3<T> s0 = {ldg(S, sc.g), ldg(S, sc.g + total), ldg(S, sc.g + 2 * total)};
    const int b0 = ldc(C.bstart, i), b1 = ldc(C.bstart, i + 1);
    const int bl = b0 + (int)threadIdx.x;
    double jx = 0, jy = 0, jz = 0;
    if (bl < b1) {
        const uint32_t gj = mc_neighbour(A, C.G, sc, bl, wide);
        jx = ldg(S, gj);
        jy = ldg(S, gj + total);
        jz = ldg(S, gj + 2 * total);
    }
    sn = mc_propose_from<T>(A, d.k, s0, ldc(C.kappa, i), n);
    const double ox = s0.x, oy = s0.y, oz = s0.z, nx = sn.x, ny = sn.y, nz = sn.z;
    dS[0] = nx - ox;
    dS[1] = ny - oy;
    dS[2] = nz - oz;
    double Qn[5], Qo[5], dQ[5];
    quadrupole_d(nx, ny, nz, Qn);
    quadrupole_d(ox, oy, oz, Qo);
#pragma unroll
    for (int q = 0; q < 5; q++) dQ[q] = Qn[q] - Qo[q];
    double eb = bl < b1 ? mc_bond_term<T>(C, bl, dS[0], dS[1], dS[2], dQ, jx, jy, jz) : 0.0;
    for (int b = bl + 64; b < b1; b += 64) {  // sublattices with more than 64 bonds
        const uint32_t gj = mc_neighbour(A, C.G, sc, b, wide);
        eb += mc_bond_term<T>(C, b, dS[0], dS[1], dS[2], dQ, ldg(S, gj), ldg(S, gj + total),
                              ldg(S, gj + 2 * total));
    }
    double dE = (double)ldc(C.zee, 3 * i) * dS[0] + (double)ldc(C.zee, 3 * i + 1) * dS[1] +
                (double)ldc(C.zee, 3 * i + 2) * dS[2];
    dE += stevens_delta_d(C, i, nx, ny, nz, ox, oy, oz);
    dE += wave_sum(eb);
    dE += mc_ewald_delta(C, i, p, A0, dS, dmu);
    return dE;
}

// Sequential trials T0 .. T0 + ntrial - 1 of the Ewald form, workgroup b <
// nrep for replica r0 + b, one wavefront; workgroups b >= nrep are the apply
// part (mc_apply_part).  The trials keep the reference's order and outcome
// (Samplers.jl:133-153) but are evaluated in parallel:
//  1. lane l of the first wavefront forms the draws of trial T0 + l
//     (mc_seq_draw; normals from counter (T lo, kMcTagSequential + 1, T hi,
//     0)) and evaluates it on the chunk's starting state: E_l =
//     local_energy_change + ewald_energy_delta at the potential p0_l of the
//     chunk's start (Interactions.jl:311-380, Ewald.jl:234-241), while the
//     other wavefronts load the tensor blocks of p0 (the earlier chunks'
//     moves at the trial sites); then all wavefronts share the loads of the
//     pair energies w[m][l] = dmu_l . A(c_m - c_l) dmu_m with every earlier
//     trial m, kept in LDS;
//  2. the trials are then decided in order: trial t's energy change is
//     E_t + sum of w[m][t] over the moves m accepted before it (the exact
//     change of ewald_energy_delta, linear in the potential), unless an
//     accepted move changed its site or a neighbour (a mask of the earlier
//     trials at those sites, formed from site ids in LDS): then the
//     whole wavefront makes trial t again on the current state (mc_trial_wave)
//     and, when it is accepted, w[t][l] is formed again with its new dmu.
// Accepted spins are stored with their ghost images by the lanes that hold
// them, together, before a trial made again and at the chunk's end; the
// decisions are the first wavefront's alone.  The chunk's accepted moves go
// to io.moves[r][.] and io.nmoves[r]; acc[4 r + q] += (dE, dS).
template <typename T, int G>
__global__ __launch_bounds__(kMcThreads) void k_mc_ewald_chunk(Couplings<T> C, McArgs A, T* __restrict__ S,
                                                               McChunkIO io, const double* __restrict__ Aew,
                                                               const double* __restrict__ A0,
                                                               unsigned long long T0, int ntrial,
                                                               double* __restrict__ acc, int r0, int nrep)
{
    if ((int)blockIdx.x >= nrep) {  // uniform over the workgroup
        mc_apply_part(io, Aew, C.reps, C.G, C.na, C.total, r0, (int)blockIdx.x - nrep);
        return;
    }
    const int r = r0 + (int)blockIdx.x;
    const ReplicaInfo R = C.reps[r];
    const uint32_t total = C.total;
    const int na = C.na;
    const unsigned long long key = A.seed + (unsigned long long)r * 0x9E3779B97F4A7C15ull;
    const double kT = A.kT[r];
    const bool wide = A.wide != 0;
    const int lane = (int)threadIdx.x % 64, wv = (int)threadIdx.x / 64;
    constexpr int NW = kMcThreads / 64;
    static_assert(NW >= 2, "waves 1.. sum the earlier moves while wave 0 makes its trials");
    __shared__ int nid[kMcChunk][kMcNbr + 1];  // [l][0]: trial l's site; [l][1 + k]: across bond k of it
    __shared__ double wsh[kMcChunk][kMcChunk];  // w[m][l], m < l
    __shared__ McTrial rec[kMcChunk];           // trial l's site and dmu
    __shared__ double ppart[NW][kMcChunk][3];   // earlier chunks' moves at trial l's site, per wave (0: none)

    // ---- 1. every lane of wave 0 evaluates its own trial on the chunk's
    // starting state while the other waves sum the earlier chunks' moves at
    // the trial sites; all waves then share the loads of the pair blocks
    McSeqSite my{};
    T my_n[3] = {0, 0, 0};
    SiteCoord sc{};
    int b0 = 0, my_nb = 0;
    V3<T> my_sn{};
    double my_dS[3] = {0, 0, 0}, my_dmu[3] = {0, 0, 0}, my_E = 0;
    if (wv == 0) {
        const unsigned long long Tl = T0 + (unsigned long long)lane;
        my = mc_seq_site(R, mc_seq_draw(key, Tl, (uint32_t)R.ncell * (uint32_t)na), A);
        uint32_t cn[4] = {(uint32_t)Tl, kMcTagSequential + 1u, (uint32_t)(Tl >> 32), 0u};
        mc_normals<T>(cn, key, my_n);
        sc = mc_site(R, C.G, my.i, my.c1, my.c2, my.c3, true);
        sc.r = r;
        b0 = C.bstart[my.i];
        my_nb = C.bstart[my.i + 1] - b0;
        nid[lane][0] = my.i * R.ncell + my.c1 + R.L1 * (my.c2 + R.L2 * my.c3);
        for (int k = 0; k < min(my_nb, kMcNbr); k++)
            nid[lane][1 + k] = mc_bond_site_id(A, R, b0 + k, my.c1, my.c2, my.c3);
        const V3<T> s0 = {ldg(S, sc.g), ldg(S, sc.g + total), ldg(S, sc.g + 2 * total)};
        my_sn = mc_propose_from<T>(A, my.k, s0, ldc(C.kappa, my.i), my_n);
        my_dS[0] = (double)my_sn.x - (double)s0.x;
        my_dS[1] = (double)my_sn.y - (double)s0.y;
        my_dS[2] = (double)my_sn.z - (double)s0.z;
        const T* gm = C.gmat + 9 * my.i;  // dmu = -g dS, as mc_ewald_delta forms it
#pragma unroll
        for (int a = 0; a < 3; a++)
            my_dmu[a] = -((double)gm[3 * a] * my_dS[0] + (double)gm[3 * a + 1] * my_dS[1] +
                          (double)gm[3 * a + 2] * my_dS[2]);
        rec[lane] = {1, my.i, my.c1, my.c2, my.c3, 0, {my_dmu[0], my_dmu[1], my_dmu[2]}};
        my_E = mc_local_delta<T, G>(C, A, S, sc, s0, my_sn, wide);
        ppart[0][lane][0] = ppart[0][lane][1] = ppart[0][lane][2] = 0.0;
    } else {
        // meanwhile the other waves sum the earlier chunks' moves m = wv - 1,
        // wv - 1 + (NW - 1), ... at this lane's trial site (formed from the
        // same draw as wave 0 forms it)
        const McSeqSite ms =
            mc_seq_site(R, mc_seq_draw(key, T0 + (unsigned long long)lane, (uint32_t)R.ncell * (uint32_t)na), A);
        double s3[3] = {0, 0, 0};
        if (io.pprev && wv == 1) {
            const double* q = io.pprev + ((size_t)r * kMcChunk + lane) * 3;
            s3[0] = q[0];
            s3[1] = q[1];
            s3[2] = q[2];
        }
        for (int l = 0; io.prev && !io.pprev && l < io.nprevl; l++) {
            const size_t lr = (size_t)l * io.nrep_all + r;
            const McTrial* pv = io.prev + lr * kMcChunk;
            const int np = io.nprev[lr];
#pragma unroll 4
            for (int m = wv - 1; m < np; m += NW - 1) {
                const McTrial t = pv[m];
                double o[3];
                ablock_apply(mc_ablock(Aew, na, R, t.i, t.c1, t.c2, t.c3, ms.i, ms.c1, ms.c2, ms.c3), t.dmu, o);
                s3[0] += o[0];
                s3[1] += o[1];
                s3[2] += o[2];
            }
        }
        ppart[wv][lane][0] = s3[0];
        ppart[wv][lane][1] = s3[1];
        ppart[wv][lane][2] = s3[2];
    }
    __syncthreads();
    {
        const McTrial me = rec[lane];
        // pair energies w[m][l] with the earlier trials m = wv, wv + NW, ...
        for (int m0 = wv; m0 < ntrial; m0 += 4 * NW) {
            ABlock blk[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                // a pair this lane does not use (m >= lane) loads the lane's own
                // offset-0 block instead: na such blocks, shared through the cache
                const int m = m0 + q * NW;
                const McTrial t = m < lane ? rec[min(m, ntrial - 1)] : me;
                blk[q] = mc_ablock(Aew, na, R, t.i, t.c1, t.c2, t.c3, me.i, me.c1, me.c2, me.c3);
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int m = m0 + q * NW;
                if (m < lane && lane < ntrial) wsh[m][lane] = mc_pair_energy(blk[q], me.dmu, rec[m].dmu);
            }
        }
    }
    __syncthreads();
    if (wv != 0) return;  // the rest is wave 0's (no workgroup barrier follows)
    // the potential of the chunk's start at the lane's trial site
    double p0[3];
#pragma unroll
    for (int x = 0; x < 3; x++) {
        double t = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) t += ppart[w][lane][x];
        p0[x] = io.phi_in[sc.g + x * (size_t)total] + t;
    }
    my_E += mc_ewald_delta(C, my.i, p0, A0, my_dS, my_dmu);
    // the Metropolis test u <= exp(-dE / kT) as dE <= -kT log u, decided by
    // the lane itself; within a relative band of 1e-9 around the threshold
    // (rounding of the logarithm) the test itself is made
    const double thr = kT == 0.0 ? 0.0 : -kT * log(my.u_acc);
    auto metropolis = [&](double dE) __attribute__((always_inline)) {
        const double margin = dE - thr;
        if (fabs(margin) > 1e-9 * (fabs(dE) + fabs(thr)) + 1e-300) return margin < 0.0;
        return kT == 0.0 ? dE <= 0.0 : my.u_acc <= exp(-dE / kT);
    };
    // conflicts: the earlier trials m at this lane's site or at one of its
    // neighbours (trial m's site taken from lane m's register)
    unsigned long long conf = 0;
    {
        const int my_id = nid[lane][0];
        int idm[kMcChunk];
#pragma unroll
        for (int m = 0; m < kMcChunk; m++) idm[m] = lane_value(my_id, m);
        for (int k = 0; k <= my_nb; k++) {
            const int id = k <= kMcNbr ? nid[lane][k] : mc_bond_site_id(A, R, b0 + k - 1, my.c1, my.c2, my.c3);
#pragma unroll
            for (int m = 0; m < kMcChunk; m++) conf |= (unsigned long long)(idm[m] == id) << m;
        }
        conf &= (1ull << lane) - 1ull;  // earlier trials only
    }

    // ---- 2. the trials decided in order
    double accw = 0;              // lane l: sum of w[m][l] over the accepted m < l
    double my_dE = 0;             // lane l: its trial's energy change when decided
    unsigned long long accm = 0;  // accepted trials (wave-uniform)
    unsigned long long pend = 0;  // accepted trials whose spin is not stored yet (wave-uniform)
    int nm = 0;
    // the spins of the pending accepted trials, each stored by its own lane
    // (no trial of this chunk reads the state between an acceptance and
    // the next trial made again, or the chunk's end)
    auto flush = [&]() __attribute__((always_inline)) {
        if (pend) {
            if ((pend >> lane) & 1ull) mc_store(S, C, sc, my_sn);
            pend = 0;
        }
    };
    int t = 0;
    while (t < ntrial) {
        // a run of trials whose site and neighbours no accepted move of this
        // chunk changed: trial t's energy change is E_t + accw, decided by
        // lane t; the pair energies of the next trial are read one ahead
        double wn = lane > t ? wsh[t][lane] : 0.0;
        for (; t < ntrial; t++) {
            if (lane_value(conf, t) & accm) break;  // made again below
            const double wt = wn;
            wn = (t + 1 < ntrial && lane > t + 1) ? wsh[t + 1][lane] : 0.0;
            const double e = my_E + accw;
            const int accept = lane_value((int)metropolis(e), t);
            if (lane == t) my_dE = e;
            if (accept) {
                accw += wt;
                accm |= 1ull << t;
                pend |= 1ull << t;
                nm++;
            }
        }
        if (t == ntrial) break;
        // trial t: an accepted move changed its site or a neighbour, so the
        // whole wavefront makes it again on the current state
        flush();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const McSeqSite d = mc_seq_site_of_lane(my, t);
        const T n[3] = {lane_value(my_n[0], t), lane_value(my_n[1], t), lane_value(my_n[2], t)};
        SiteCoord st = mc_site(R, C.G, d.i, d.c1, d.c2, d.c3, true);
        st.r = r;
        // potential at the site: p0 + the accepted moves of this chunk
        double c3v[3] = {0, 0, 0};
        if (lane < t && ((accm >> lane) & 1ull))
            ablock_apply(mc_ablock(Aew, na, R, my.i, my.c1, my.c2, my.c3, d.i, d.c1, d.c2, d.c3), my_dmu, c3v);
        double p[3];
#pragma unroll
        for (int x = 0; x < 3; x++) p[x] = lane_value(p0[x], t) + wave_sum(c3v[x]);
        V3<T> sn;
        double dmu[3], dS[3];
        const double dE = mc_trial_wave<T>(C, A, S, st, d, n, p, A0, wide, sn, dS, dmu);
        const int accept = __builtin_amdgcn_readfirstlane(kT == 0.0 ? dE <= 0.0 : d.u_acc <= exp(-dE / kT));
        if (lane == t) {
            my_dE = dE;
            my_dS[0] = dS[0];
            my_dS[1] = dS[1];
            my_dS[2] = dS[2];
        }
        if (accept) {
            if (lane == 0) mc_store(S, C, st, sn);
            if (lane == t) {
                my_dmu[0] = dmu[0];
                my_dmu[1] = dmu[1];
                my_dmu[2] = dmu[2];
            }
            // the pair energies of this move with the later trials, with its new dmu
            if (lane > t && lane < ntrial) {
                wsh[t][lane] = mc_pair_energy(mc_ablock(Aew, na, R, d.i, d.c1, d.c2, d.c3, my.i, my.c1, my.c2,
                                                        my.c3),
                                              my_dmu, dmu);
                accw += wsh[t][lane];
            }
            accm |= 1ull << t;
            nm++;
        }
        // this trial's stores (LDS and the state) precede the next trial's
        // loads: the lanes of one wavefront see each other's writes in
        // program order, so no wait for the stores is needed
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        t++;
    }
    flush();
    // sums of dE and dS over the accepted trials
    const bool mine = lane < ntrial && ((accm >> lane) & 1ull);
    double sum[4];
    sum[0] = wave_sum(mine ? my_dE : 0.0);
    for (int a = 0; a < 3; a++) sum[1 + a] = wave_sum(mine ? my_dS[a] : 0.0);
    if (lane < ntrial && ((accm >> lane) & 1ull)) {
        const int slot = __popcll(accm & ((1ull << lane) - 1ull));
        io.moves[(size_t)r * kMcChunk + slot] = {1, my.i, my.c1, my.c2, my.c3, 0, {my_dmu[0], my_dmu[1], my_dmu[2]}};
    }
    if (lane == 0) {
        io.nmoves[r] = nm;
        for (int q = 0; q < 4; q++) acc[4 * r + q] += sum[q];
    }
}

// The moves of io's nprevl lists at the site of each trial of the next
// chunk, spread over the whole device: one wavefront per (replica, trial),
// lane j summing moves j, j + 64, ... of each list, then a wave sum.  The
// trial site comes from the same draw as in k_mc_ewald_chunk.
__global__ __launch_bounds__(256) void k_mc_prev_at_trials(McArgs A, McChunkIO io,
                                                           const ReplicaInfo* __restrict__ reps,
                                                           const double* __restrict__ Aew, int na,
                                                           unsigned long long T0, int r0, double* __restrict__ out)
{
    const int lane = (int)threadIdx.x % 64;
    const int w = (int)(blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64);  // r' kMcChunk + l
    const int r = r0 + w / kMcChunk, l = w % kMcChunk;
    const ReplicaInfo R = reps[r];
    const unsigned long long key = A.seed + (unsigned long long)r * 0x9E3779B97F4A7C15ull;
    const McSeqSite ms =
        mc_seq_site(R, mc_seq_draw(key, T0 + (unsigned long long)l, (uint32_t)R.ncell * (uint32_t)na), A);
    double s3[3] = {0, 0, 0};
    for (int q = 0; q < io.nprevl; q++) {
        const size_t lr = (size_t)q * io.nrep_all + r;
        const McTrial* pv = io.prev + lr * kMcChunk;
        const int np = io.nprev[lr];
        for (int m = lane; m < np; m += 64) {
            const McTrial t = pv[m];
            double o[3];
            ablock_apply(mc_ablock(Aew, na, R, t.i, t.c1, t.c2, t.c3, ms.i, ms.c1, ms.c2, ms.c3), t.dmu, o);
            s3[0] += o[0];
            s3[1] += o[1];
            s3[2] += o[2];
        }
    }
#pragma unroll
    for (int x = 0; x < 3; x++) s3[x] = wave_sum(s3[x]);
    if (lane == 0) {
        double* o = out + ((size_t)r * kMcChunk + l) * 3;
        o[0] = s3[0];
        o[1] = s3[1];
        o[2] = s3[2];
    }
}

void launch_mc_prev_at_trials(const McArgs& A, const McChunkIO& io, const ReplicaInfo* reps, const double* Aew,
                              int na, unsigned long long T0, int r0, int nr, double* out, hipStream_t st)
{
    // kMcChunk wavefronts per replica, four per workgroup
    hipLaunchKernelGGL(k_mc_prev_at_trials, dim3((unsigned)(nr * kMcChunk / 4)), dim3(256), 0, st, A, io, reps, Aew,
                       na, T0, r0, out);
}

// The real-space tensor's 3x3 blocks (row-major, nine doubles) as the six
// entries xx, xy, xz, yy, yz, zz of their upper triangle.
__global__ void k_ewald_A_sym(const double* __restrict__ A9, double* __restrict__ A6, long long nblk)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nblk) return;
    const double* a = A9 + 9 * k;
    double2* o = reinterpret_cast<double2*>(A6 + 6 * k);
    o[0] = {a[0], a[1]};
    o[1] = {a[2], a[4]};
    o[2] = {a[5], a[8]};
}

template <typename T>
__global__ void k_planes_to_double(const T* __restrict__ src, double* __restrict__ dst, long long n)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = (double)src[k];
}

// out[4 r + q] = sum of acc[4 k + q] over the work items k of replica r:
// ranges [start[c (R + 1) + r], start[c (R + 1) + r + 1]) of every colour c.
__global__ __launch_bounds__(kBlock) void k_mc_reduce(const double* __restrict__ acc, const int* __restrict__ start,
                                                      int ncolors, int R, double* __restrict__ out)
{
    const int r = blockIdx.x;
    double v[4] = {0, 0, 0, 0};
    for (int c = 0; c < ncolors; c++) {
        const int* st = start + c * (R + 1);
        for (int k = st[r] + threadIdx.x; k < st[r + 1]; k += kBlock)
            for (int q = 0; q < 4; q++) v[q] += acc[4 * (size_t)k + q];
    }
    block_sum4(v);
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; q++) out[4 * r + q] = v[q];
}

__global__ void k_fill_double(double* __restrict__ p, long long n, double v)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}

template <typename T>
void launch_mc_color(const Couplings<T>& C, const McArgs& A, T* S, const McWork* work, int nwork,
                     unsigned long long P, double prob, double* acc, hipStream_t st)
{
    if (nwork <= 0) return;
    if (C.chunk == 6)
        hipLaunchKernelGGL((k_mc_color<T, 6>), dim3(nwork), dim3(kBlock), 0, st, C, A, S, work, P, prob, acc);
    else  // bond lists padded to 9 or 18
        hipLaunchKernelGGL((k_mc_color<T, 9>), dim3(nwork), dim3(kBlock), 0, st, C, A, S, work, P, prob, acc);
}

void launch_ewald_A_sym(const double* A9, double* A6, long long nblk, hipStream_t st)
{
    hipLaunchKernelGGL(k_ewald_A_sym, dim3((unsigned)((nblk + 255) / 256)), dim3(256), 0, st, A9, A6, nblk);
}

template <typename T>
void launch_mc_ewald_chunk(const Couplings<T>& C, const McArgs& A, T* S, const McChunkIO& io, const double* Aew,
                           const double* A0, unsigned long long T0, int ntrial, double* acc, int r0, int nr,
                           hipStream_t st)
{
    if (C.chunk == 6)
        hipLaunchKernelGGL((k_mc_ewald_chunk<T, 6>), dim3(nr * (1 + io.napp)), dim3(kMcThreads), 0, st, C, A, S, io,
                           Aew, A0, T0, ntrial, acc, r0, nr);
    else  // bond lists padded to 9 or 18
        hipLaunchKernelGGL((k_mc_ewald_chunk<T, 9>), dim3(nr * (1 + io.napp)), dim3(kMcThreads), 0, st, C, A, S, io,
                           Aew, A0, T0, ntrial, acc, r0, nr);
}

template <typename T>
void launch_planes_to_double(const T* src, double* dst, long long n, hipStream_t st)
{
    hipLaunchKernelGGL((k_planes_to_double<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, dst, n);
}

void launch_mc_reduce(const double* acc, const int* start, int ncolors, int R, double* out, hipStream_t st)
{
    hipLaunchKernelGGL(k_mc_reduce, dim3(R), dim3(kBlock), 0, st, acc, start, ncolors, R, out);
}

void launch_fill_double(double* p, long long n, double v, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_double, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
}

template void launch_mc_ewald_chunk<float>(const Couplings<float>&, const McArgs&, float*, const McChunkIO&,
                                           const double*, const double*, unsigned long long, int, double*, int, int,
                                           hipStream_t);
template void launch_mc_ewald_chunk<double>(const Couplings<double>&, const McArgs&, double*, const McChunkIO&,
                                            const double*, const double*, unsigned long long, int, double*, int, int,
                                            hipStream_t);
template void launch_planes_to_double<float>(const float*, double*, long long, hipStream_t);
template void launch_planes_to_double<double>(const double*, double*, long long, hipStream_t);
template void launch_mc_color<float>(const Couplings<float>&, const McArgs&, float*, const McWork*, int,
                                     unsigned long long, double, double*, hipStream_t);
template void launch_mc_color<double>(const Couplings<double>&, const McArgs&, double*, const McWork*, int,
                                      unsigned long long, double, double*, hipStream_t);

}  // namespace snh
