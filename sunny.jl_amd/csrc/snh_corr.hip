// Sampled dynamical correlations of one replica on the device (:SUN states
// are observed through their expected spins).
//
// Reference semantics (Sunny.jl 0.9.2, src/SampledCorrelations/):
//   observable_values!, trajectory!, new_sample!   CorrelationSampling.jl:1-44
//   accum_sample!                                  CorrelationSampling.jl:46-138
//   SampledCorrelations buffers and FFT plans      SampledCorrelations.jl:194-268
//
// Buffers (FP64 complex, cell index q = q1 + L1 (q2 + L2 q3)):
//   samp [T][nobs][na][ncell]   A_a,i(r or q, t or w), T = n_all_omega
//   P, Y [T][na][ncell]         one correlation channel (c, i) over all j
//   data [ncorr][na][na][T][ncell]   running mean of S^{ab}_{ij}(q, w)
// Steps per sample:
//   k_corr_observe   A(r, t) = (o_a,i . S_r) / sqrt(ncell) for each snapshot
//   k_corr_lines     the lattice transform A(q) = sum_r e^{+iqr} A(r), one
//                    axis at a time, every line held in registers (FFTW's
//                    forward transform between two conjugations)
//   k_corr_tdft      A(w) = sum_t e^{-iwt} A(t) (the zero-padded signal; direct
//                    DFT over the T time slots, T = 2(n_omega - 1) of any size)
//   per (c, i):  k_corr_product  P = conj(A_a,i) A_b,j for all j
//                k_corr_tdft     Y = ifft(P) scaled by window / (T n_contrib)
//                k_corr_tdft     Z = fft(Y), accumulated into the running mean
#include "snh_fft.h"

#include <hipfft/hipfft.h>

namespace snh {

using Cd = C2<double>;

constexpr int kCorrThreads = 256;

// A(r, t) for every site and observable of the replica (one snapshot).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_corr_observe(Couplings<T> C, const T* __restrict__ S, int nobs,
                                                         const double* __restrict__ obs, Cd* __restrict__ samp,
                                                         int t, double scale)
{
    const SiteCoord sc = site_coord(C.blocks, C.reps, C.G);
    if (!sc.valid) return;
    const double sx = ldg(S, sc.g), sy = ldg(S + C.total, sc.g), sz = ldg(S + 2 * (size_t)C.total, sc.g);
    const int na = C.na;
    const long long ncell = sc.R.ncell;
    for (int a = 0; a < nobs; a++) {
        const double* o = obs + 3 * (a * na + sc.i);
        const double v = (o[0] * sx + o[1] * sy + o[2] * sz) * scale;
        samp[(((long long)t * nobs + a) * na + sc.i) * ncell + sc.c] = {v, 0.0};
    }
}

// In-place c2c transforms of the lines along `axis` of nfield fields of
// L1 x L2 x L3 cells (sign +1: e^{+2 pi i k x / L}).  One thread per line.
__global__ __launch_bounds__(kCorrThreads) void k_corr_lines(Cd* __restrict__ buf, long long nfield, int L1, int L2,
                                                             int L3, int axis)
{
    __shared__ Cd tw[kLmax];
    const int L = axis == 0 ? L1 : (axis == 1 ? L2 : L3);
    make_twiddles(tw, L, +1);
    __syncthreads();
    const long long ncell = (long long)L1 * L2 * L3;
    const long long lines_per_field = ncell / L;
    const long long line = (long long)blockIdx.x * kCorrThreads + threadIdx.x;
    if (line >= nfield * lines_per_field) return;
    const long long f = line / lines_per_field, l = line % lines_per_field;
    long long base;
    int stride;
    if (axis == 0) {  // l = q2 + L2 q3
        base = l * L1;
        stride = 1;
    } else if (axis == 1) {  // l = q1 + L1 q3
        base = (l % L1) + (long long)L1 * L2 * (l / L1);
        stride = L1;
    } else {  // l = q1 + L1 q2
        base = l;
        stride = L1 * L2;
    }
    line_c2c<double>(buf + f * ncell, (int)base, stride, L, tw);
}

// Direct DFT over the T time slots of nl independent lines ([T][nl] layout):
//   out[k][l] = post_k * sum_t e^{sign 2 pi i k t / T} pre_t in[t][l]
// pre (may be null): per-slot factor on the input; post (may be null): per
// slot factor on the output.  With mean_count > 0, out is a running mean:
// out += (value - out) / mean_count.  One thread per (k, l); e^{...} from a
// table of the T roots of unity in LDS.
__global__ __launch_bounds__(kCorrThreads) void k_corr_tdft(const Cd* __restrict__ in, Cd* __restrict__ out,
                                                            long long nl, int T, int sign,
                                                            const double* __restrict__ pre,
                                                            const double* __restrict__ post, long long mean_count)
{
    extern __shared__ __align__(16) unsigned char corr_lds[];
    Cd* w = reinterpret_cast<Cd*>(corr_lds);  // [T]
    for (int m = threadIdx.x; m < T; m += blockDim.x) {
        double s, c;
        sincospi(2.0 * m / T, &s, &c);
        w[m] = {c, sign * s};
    }
    __syncthreads();
    const long long idx = (long long)blockIdx.x * kCorrThreads + threadIdx.x;
    if (idx >= nl * T) return;
    const int k = (int)(idx / nl);
    const long long l = idx % nl;
    double ar = 0, ai = 0;
    int m = 0;  // k t mod T
    for (int t = 0; t < T; t++) {
        Cd x = in[(long long)t * nl + l];
        if (pre) {
            x.re *= pre[t];
            x.im *= pre[t];
        }
        ar += x.re * w[m].re - x.im * w[m].im;
        ai += x.re * w[m].im + x.im * w[m].re;
        m += k;
        if (m >= T) m -= T;
    }
    if (post) {
        ar *= post[k];
        ai *= post[k];
    }
    Cd* o = out + (long long)k * nl + l;
    if (mean_count > 0) {
        const double inv = 1.0 / (double)mean_count;
        o->re += (ar - o->re) * inv;
        o->im += (ai - o->im) * inv;
    } else {
        *o = {ar, ai};
    }
}

// P[t][j][q] = conj(A_a,i(q, t)) A_b,j(q, t) for every j.
__global__ __launch_bounds__(kCorrThreads) void k_corr_product(const Cd* __restrict__ samp, Cd* __restrict__ P,
                                                               int T, int nobs, int na, long long ncell, int a,
                                                               int i, int b)
{
    const long long idx = (long long)blockIdx.x * kCorrThreads + threadIdx.x;
    const long long per_t = (long long)na * ncell;
    if (idx >= (long long)T * per_t) return;
    const long long t = idx / per_t, r = idx % per_t;
    const long long j = r / ncell, q = r % ncell;
    const Cd x = samp[((t * nobs + a) * na + i) * ncell + q];
    const Cd y = samp[((t * nobs + b) * na + j) * ncell + q];
    P[idx] = {x.re * y.re + x.im * y.im, x.re * y.im - x.im * y.re};
}

// ---------------------------------------------------------------------------
// Launchers

static unsigned corr_grid(long long n) { return (unsigned)((n + kCorrThreads - 1) / kCorrThreads); }

template <typename T>
void launch_corr_observe(const Couplings<T>& C, int nblocks, const T* S, int nobs, const double* obs, void* samp,
                         int t, double scale, hipStream_t st)
{
    hipLaunchKernelGGL(k_corr_observe<T>, dim3(nblocks), dim3(kBlock), 0, st, C, S, nobs, obs,
                       reinterpret_cast<Cd*>(samp), t, scale);
}

void launch_corr_space_fft(void* buf, long long nfield, int L1, int L2, int L3, hipStream_t st)
{
    const long long ncell = (long long)L1 * L2 * L3;
    const int L[3] = {L1, L2, L3};
    for (int axis = 0; axis < 3; axis++) {
        if (L[axis] == 1) continue;
        const long long lines = nfield * (ncell / L[axis]);
        hipLaunchKernelGGL(k_corr_lines, dim3(corr_grid(lines)), dim3(kCorrThreads), 0, st,
                           reinterpret_cast<Cd*>(buf), nfield, L1, L2, L3, axis);
    }
}

void launch_corr_tdft(const void* in, void* out, long long nl, int T, int sign, const double* pre, const double* post,
                      long long mean_count, hipStream_t st)
{
    hipLaunchKernelGGL(k_corr_tdft, dim3(corr_grid(nl * T)), dim3(kCorrThreads), T * sizeof(Cd), st,
                       reinterpret_cast<const Cd*>(in), reinterpret_cast<Cd*>(out), nl, T, sign, pre, post,
                       mean_count);
}

void launch_corr_product(const void* samp, void* P, int T, int nobs, int na, long long ncell, int a, int i, int b,
                         hipStream_t st)
{
    hipLaunchKernelGGL(k_corr_product, dim3(corr_grid((long long)T * na * ncell)), dim3(kCorrThreads), 0, st,
                       reinterpret_cast<const Cd*>(samp), reinterpret_cast<Cd*>(P), T, nobs, na, ncell, a, i, b);
}

// Lattice transforms with an axis longer than kLmax (or under
// SNH_CORR_FFT=lib): one batched in-place 3D Z2Z hipFFT plan per
// corr_init, HIPFFT_BACKWARD (e^{+2 pi i k x / L}, unnormalised, as the
// line transforms), its work area allocated with it.
struct CorrLibFft {
    hipfftHandle plan = 0;
    void* work = nullptr;
};

void corr_lib_destroy(CorrLibFft* L)
{
    if (!L) return;
    if (L->plan) (void)hipfftDestroy(L->plan);
    if (L->work) (void)hipFree(L->work);
    delete L;
}

CorrLibFft* corr_lib_create(long long nfield, int L1, int L2, int L3, std::string& err)
{
    CorrLibFft* L = new CorrLibFft;
    int n[3] = {L3, L2, L1};
    const int ncell = L1 * L2 * L3;
    size_t w = 0;
    hipfftResult e = hipfftCreate(&L->plan);
    if (e == HIPFFT_SUCCESS) e = hipfftSetAutoAllocation(L->plan, 0);
    if (e == HIPFFT_SUCCESS)
        e = hipfftMakePlanMany(L->plan, 3, n, n, 1, ncell, n, 1, ncell, HIPFFT_Z2Z, (int)nfield, &w);
    if (e != HIPFFT_SUCCESS) {
        err = "hipFFT plan creation failed (code " + std::to_string((int)e) + ")";
        corr_lib_destroy(L);
        return nullptr;
    }
    if (hipMalloc(&L->work, std::max<size_t>(w, 256)) != hipSuccess ||
        hipfftSetWorkArea(L->plan, L->work) != HIPFFT_SUCCESS) {
        err = "hipFFT work area not allocated";
        corr_lib_destroy(L);
        return nullptr;
    }
    return L;
}

bool corr_lib_exec(CorrLibFft* L, void* buf, hipStream_t st)
{
    auto* b = reinterpret_cast<hipfftDoubleComplex*>(buf);
    return hipfftSetStream(L->plan, st) == HIPFFT_SUCCESS &&
           hipfftExecZ2Z(L->plan, b, b, HIPFFT_BACKWARD) == HIPFFT_SUCCESS;
}

template void launch_corr_observe<float>(const Couplings<float>&, int, const float*, int, const double*, void*, int,
                                         double, hipStream_t);
template void launch_corr_observe<double>(const Couplings<double>&, int, const double*, int, const double*, void*,
                                          int, double, hipStream_t);

}  // namespace snh
