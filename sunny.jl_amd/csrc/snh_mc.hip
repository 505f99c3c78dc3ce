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
#include "snh_mc.h"
#include "snh_vec.h"

namespace snh {

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
        const int k = mc_pick(A, u_mix);
        const V3<T> s0 = {ldg(S, sc.g), ldg(S, sc.g + total), ldg(S, sc.g + 2 * total)};
        const T kap = ldc(C.kappa, i);
        V3<T> sn;
        if (A.kind[k] == 1) {  // propose_flip
            sn = {-s0.x, -s0.y, -s0.z};
        } else {
            uint32_t cn[4] = {site, kMcTagNormals, (uint32_t)P, (uint32_t)(P >> 32)};
            philox4x32(cn, (uint32_t)key, (uint32_t)(key >> 32));
            T n[4];
            box_muller(cn[0], cn[1], n[0], n[1]);
            box_muller(cn[2], cn[3], n[2], n[3]);
            const double mag = A.mag[k];
            if (A.kind[k] == 0 || __builtin_isinf(mag))  // randspin
                sn = normalize_dipole(V3<T>{n[0], n[1], n[2]}, kap);
            else {  // perturbed_spin: S + mag kappa randn
                const T f = (T)(mag * (double)kap);
                sn = normalize_dipole(V3<T>{s0.x + f * n[0], s0.y + f * n[1], s0.z + f * n[2]}, kap);
            }
        }
        // local_energy_change in FP64
        const double ox = s0.x, oy = s0.y, oz = s0.z, nx = sn.x, ny = sn.y, nz = sn.z;
        const double dx = nx - ox, dy = ny - oy, dz = nz - oz;
        double dE = (double)ldc(C.zee, 3 * i) * dx + (double)ldc(C.zee, 3 * i + 1) * dy +
                    (double)ldc(C.zee, 3 * i + 2) * dz;
        dE += stevens_energy_d(C, i, nx, ny, nz) - stevens_energy_d(C, i, ox, oy, oz);
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
        const double kT = ldc(A.kT, W.r);
        const bool accept = kT == 0.0 ? dE <= 0.0 : u_acc <= exp(-dE / kT);
        if (accept) {
            stg(S, sc.g, sn.x);
            stg(S, sc.g + total, sn.y);
            stg(S, sc.g + 2 * total, sn.z);
            for_each_ghost(sc, C.G, [&](uint32_t gi) {
                stg(S, gi, sn.x);
                stg(S, gi + total, sn.y);
                stg(S, gi + 2 * total, sn.z);
            });
            v[0] = dE;
            v[1] = dx;
            v[2] = dy;
            v[3] = dz;
        }
    }
    block_sum4(v);
    if (threadIdx.x == 0)
        for (int qv = 0; qv < 4; qv++) acc[4 * (size_t)blockIdx.x + qv] += v[qv];
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

void launch_mc_reduce(const double* acc, const int* start, int ncolors, int R, double* out, hipStream_t st)
{
    hipLaunchKernelGGL(k_mc_reduce, dim3(R), dim3(kBlock), 0, st, acc, start, ncolors, R, out);
}

void launch_fill_double(double* p, long long n, double v, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_double, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
}

template void launch_mc_color<float>(const Couplings<float>&, const McArgs&, float*, const McWork*, int,
                                     unsigned long long, double, double*, hipStream_t);
template void launch_mc_color<double>(const Couplings<double>&, const McArgs&, double*, const McWork*, int,
                                      unsigned long long, double, double*, hipStream_t);

}  // namespace snh
