// Ghost-padded layout helpers of the :dipole kernels: plane access, stores
// with ghost images, and the chunked bond-stencil gather (local_field).
// Shared by snh_dipole.hip (dynamics) and snh_adjoint.hip (adjoint sweeps).
//
// Reference semantics: set_energy_grad_dipoles!, src/System/Interactions.jl:516-592.
#pragma once
#include "snh_vec.h"

namespace snh {

template <typename T>
struct Planes {
    const T* x;
    const T* y;
    const T* z;
    __device__ __forceinline__ V3<T> operator()(uint32_t g) const { return {ldg(x, g), ldg(y, g), ldg(z, g)}; }
};

template <typename T>
__device__ __forceinline__ Planes<T> planes(const T* P, uint32_t total)
{
    return {P, P + total, P + 2 * (size_t)total};
}

template <typename T>
__device__ __forceinline__ void store3(T* P, uint32_t total, uint32_t g, V3<T> v)
{
    stg(P, g, v.x);
    stg(P + total, g, v.y);
    stg(P + 2 * (size_t)total, g, v.z);
}

// Stores v at the site and at all of its ghost images.
template <typename T>
__device__ __forceinline__ void store3_images(T* P, uint32_t total, const SiteCoord& sc, const Ghost& G, V3<T> v)
{
    store3(P, total, sc.g, v);
    for_each_ghost(sc, G, [&](uint32_t gi) { store3(P, total, gi, v); });
}

// Spin source for the gather: stored planes, or the midpoint
// Shat = kappa * normalize((S + S1)/2) formed on the fly.  load() only issues
// the plane loads; finish() does the arithmetic, so a chunk of neighbours is
// fetched before any of them is consumed.
template <typename T, bool kMid>
struct SpinSrc {
    Planes<T> a, b;
    struct Raw {
        V3<T> u, v;
    };
    __device__ __forceinline__ Raw load(uint32_t g) const
    {
        Raw r;
        r.u = a(g);
        if (kMid) r.v = b(g);
        return r;
    }
    __device__ __forceinline__ V3<T> finish(const Raw& r, T kappa) const
    {
        if (!kMid) return r.u;
        return normalize_dipole(V3<T>{(r.u.x + r.v.x) / 2, (r.u.y + r.v.y) / 2, (r.u.z + r.v.z) / 2}, kappa);
    }
    __device__ __forceinline__ V3<T> get(uint32_t g, T kappa) const { return finish(load(g), kappa); }
};

// dE/dS at one site: Zeeman + Stevens anisotropy + bonds (+ Ewald: -g' phi).
// Bonds are gathered CH at a time (bond lists are padded to a multiple of
// CH with zero self-couplings), so each chunk is one branch-free batch of
// loads whose per-bond metadata comes from contiguous scalar loads.
template <typename T, int CH, bool kMid, bool kIso = false>
__device__ __forceinline__ V3<T> local_field(const Couplings<T>& C, const SiteCoord& sc,
                                             const SpinSrc<T, kMid>& src, V3<T> s, const T* __restrict__ phi)
{
    const int i = sc.i;
    V3<T> g = {C.zee[3 * i], C.zee[3 * i + 1], C.zee[3 * i + 2]};
    const int b0 = C.bstart[i], b1 = C.bstart[i + 1];
    const int* __restrict__ dl = C.bdelta + sc.R.dcls * C.nb;
    for (int c0 = b0; c0 < b1; c0 += CH) {
        typename SpinSrc<T, kMid>::Raw raw[CH];
#pragma unroll
        for (int k = 0; k < CH; k++) raw[k] = src.load(sc.g + (uint32_t)dl[c0 + k]);
#pragma unroll
        for (int k = 0; k < CH; k++) {
            const int b = c0 + k;
            const V3<T> sj = src.finish(raw[k], kMid ? C.kappa[C.bj[b]] : (T)0);
            if (kIso) {  // every bond isotropic, none biquadratic
                const T J0 = C.bJ0[b];
                g.x += J0 * sj.x;
                g.y += J0 * sj.y;
                g.z += J0 * sj.z;
                continue;
            }
            add_bond_field(C, b, s, sj, g);
        }
    }
    add_onsite_field(C, i, s, g);
    if (phi) add_ewald_field(C, i, planes(phi, C.total)(sc.g), g);
    return g;
}

}  // namespace snh
