// Device-side building blocks shared by the gfx950 kernels of libsunnyhip.
//
// State layout in HBM (structure of arrays, one plane per real component):
//   dipoles:   plane a in {x,y,z}:        S[a*total + g]
//   coherents: plane 2k (re), 2k+1 (im):  Z[(2k+re/im)*total + g]
// where g = roff[r] + c + Ncell_r * i is the global site index of site
// (cell c, sublattice i) of replica r; c = c1 + L1*(c2 + L2*c3) as in
// Sunny's column-major Array{Vec3,4}(L1,L2,L3,na) (src/System/Types.jl:131-136).
// Consecutive lanes of a wavefront hold consecutive cells of one sublattice,
// so every plane load of the bond gather is a coalesced 64-lane access.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace snh {

// One work block = 256 consecutive cells of one (replica, sublattice).
// Blocks never straddle a sublattice, so the bond loop and all coupling
// tables are wavefront-uniform (scalar loads).
struct BlockInfo {
    int r;      // replica
    int i;      // sublattice
    int c0;     // first cell
    int dcls;   // index of the replica's dims class (bond offsets reduced mod L)
};

struct ReplicaInfo {
    int L1, L2, L3, ncell;
    long long off;  // first global site index
};

constexpr int kBlock = 256;

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) + Box-Muller.
__device__ inline void philox4x32(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint32_t hi0 = __umulhi(0xD2511F53u, c[0]), lo0 = 0xD2511F53u * c[0];
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]), lo1 = 0xCD9E8D57u * c[2];
        uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// Four standard normals for (site, block, step) under key = seed.
template <typename T>
__device__ inline void philox_normals4(uint64_t seed, uint64_t step, uint32_t site, uint32_t blk, T out[4])
{
    uint32_t c[4] = {site, blk, (uint32_t)step, (uint32_t)(step >> 32)};
    philox4x32(c, (uint32_t)seed, (uint32_t)(seed >> 32));
#pragma unroll
    for (int p = 0; p < 2; p++) {
        T ua = ((T)c[2 * p] + (T)1) * (T)(1.0 / 4294967296.0);
        T ub = (T)c[2 * p + 1] * (T)(1.0 / 4294967296.0);
        T rr = sqrt((T)-2 * log(ua));
        T th = (T)(2.0 * M_PI) * ub;
        out[2 * p] = rr * cos(th);
        out[2 * p + 1] = rr * sin(th);
    }
}

// ---------------------------------------------------------------------------
// Classical Stevens anisotropy gradient (and energy), the closed form of
// Sunny's OnsiteCoupling.jl:316-425.  c = c0 | c2[5] | c4[9] | c6[13].
// As in the reference, X = |S|^2 is held constant in the gradient.
template <typename T>
__device__ inline void stevens_grad(T sx, T sy, T sz, const T* __restrict__ c, int kmax,
                                    T& gx, T& gy, T& gz, T* energy)
{
    T E = c[0];
    T dz = 0, jr = 0, ji = 0;  // dE/dz, dE/dJp (re, im)
    if (kmax > 0) {
        const T* c2 = c + 1;
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
            const T* c4 = c + 6;
            T X2 = X * X;
            T z3 = z2 * z1, z4 = z2 * z2;
            T p3r = p2r * p1r - p2i * p1i, p3i = p2r * p1i + p2i * p1r;
            T p4r = p2r * p2r - p2i * p2i, p4i = 2 * p2r * p2i;
            T B1 = 35 * z4 - 30 * X * z2 + 3 * X2, B2 = 7 * z3 - 3 * X * z1, B3 = 7 * z2 - X, B4 = z1;
            T dB1 = 140 * z3 - 60 * X * z1, dB2 = 21 * z2 - 3 * X, dB3 = 14 * z1;
            E += c4[0] * p4r + c4[8] * p4i + (c4[1] * p3r + c4[7] * p3i) * B4 +
                 (c4[2] * p2r + c4[6] * p2i) * B3 + (c4[3] * p1r + c4[5] * p1i) * B2 + c4[4] * B1;
            dz += (c4[1] * p3r + c4[7] * p3i) + (c4[2] * p2r + c4[6] * p2i) * dB3 +
                  (c4[3] * p1r + c4[5] * p1i) * dB2 + c4[4] * dB1;
            // sum_m (m/2) (c_m - i c_-m) Jp^(m-1) A_m
            jr += 2 * (c4[0] * p3r + c4[8] * p3i) + (T)1.5 * (c4[1] * p2r + c4[7] * p2i) * B4 +
                  (c4[2] * p1r + c4[6] * p1i) * B3 + (T)0.5 * c4[3] * B2;
            ji += 2 * (c4[0] * p3i - c4[8] * p3r) + (T)1.5 * (c4[1] * p2i - c4[7] * p2r) * B4 +
                  (c4[2] * p1i - c4[6] * p1r) * B3 - (T)0.5 * c4[5] * B2;
            if (kmax > 4) {
                const T* c6 = c + 15;
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

// Periodic neighbour: (c1,c2,c3) + reduced offset (each component in [0, L)).
__device__ inline int wrap_add(int c, int n, int L)
{
    int t = c + n;
    return t >= L ? t - L : t;
}

}  // namespace snh
