// Line transforms held in registers, shared by the Ewald convolution and the
// correlation sampling: radix-2 decimation in time for power-of-two lengths
// (compile-time indices), a direct DFT for other lengths up to kLmax.
#pragma once
#include "snh_internal.h"

namespace snh {

template <typename T>
struct C2 {
    T re, im;
};

template <typename T>
__device__ __forceinline__ C2<T> cmul2(C2<T> a, C2<T> b)
{
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// ---------------------------------------------------------------------------
// One line of the 3D transform.  tw[m] = exp(sign 2 pi i m / L), m < L, in LDS.

constexpr int bit_reverse(int i, int L)
{
    int r = 0;
    for (int b = 1; b < L; b <<= 1) {
        r = (r << 1) | (i & 1);
        i >>= 1;
    }
    return r;
}

constexpr int ilog2(int L) { return L <= 1 ? 0 : 1 + ilog2(L >> 1); }

// In-place radix-2 decimation-in-time FFT of L (a power of two) values held
// in registers; all indices are compile-time after unrolling.
template <typename T, int L>
__device__ __forceinline__ void fft_regs(C2<T> (&x)[L], const C2<T>* __restrict__ tw)
{
#pragma unroll
    for (int i = 0; i < L; i++) {
        const int j = bit_reverse(i, L);
        if (j > i) {
            const C2<T> t = x[i];
            x[i] = x[j];
            x[j] = t;
        }
    }
#pragma unroll
    for (int s = 1; s <= ilog2(L); s++) {
        const int len = 1 << s, half = len >> 1, step = L / len;
#pragma unroll
        for (int i = 0; i < L; i += len)
#pragma unroll
            for (int j = 0; j < half; j++) {
                const C2<T> u = x[i + j];
                const C2<T> v = j == 0 ? x[i + j + half] : cmul2(x[i + j + half], tw[j * step]);
                x[i + j] = {u.re + v.re, u.im + v.im};
                x[i + j + half] = {u.re - v.re, u.im - v.im};
            }
    }
}

// c2c transform of the L values buf[base + x stride] in place.
template <typename T, int L>
__device__ __forceinline__ void line_c2c_pow2(C2<T>* buf, int base, int stride, const C2<T>* tw)
{
    C2<T> x[L];
#pragma unroll
    for (int i = 0; i < L; i++) x[i] = buf[base + i * stride];
    fft_regs<T, L>(x, tw);
#pragma unroll
    for (int i = 0; i < L; i++) buf[base + i * stride] = x[i];
}

// Direct DFT of a line of any length L <= kLmax, in place: inputs are read
// from LDS for every output, outputs held in registers until all are formed.
template <typename T>
__device__ __forceinline__ void line_c2c_any(C2<T>* buf, int base, int stride, int L, const C2<T>* tw)
{
    C2<T> out[kLmax];
#pragma unroll
    for (int k = 0; k < kLmax; k++) {
        if (k < L) {
            T ar = 0, ai = 0;
            int m = 0;
            for (int x = 0; x < L; x++) {
                const C2<T> v = buf[base + x * stride], w = tw[m];
                ar += v.re * w.re - v.im * w.im;
                ai += v.re * w.im + v.im * w.re;
                m += k;
                if (m >= L) m -= L;
            }
            out[k] = {ar, ai};
        }
    }
#pragma unroll
    for (int k = 0; k < kLmax; k++)
        if (k < L) buf[base + k * stride] = out[k];
}

template <typename T>
__device__ __forceinline__ void line_c2c(C2<T>* buf, int base, int stride, int L, const C2<T>* tw)
{
    switch (L) {
    case 1: return;
    case 2: line_c2c_pow2<T, 2>(buf, base, stride, tw); return;
    case 4: line_c2c_pow2<T, 4>(buf, base, stride, tw); return;
    case 8: line_c2c_pow2<T, 8>(buf, base, stride, tw); return;
    case 16: line_c2c_pow2<T, 16>(buf, base, stride, tw); return;
    case 32: line_c2c_pow2<T, 32>(buf, base, stride, tw); return;
    default: line_c2c_any<T>(buf, base, stride, L, tw); return;
    }
}

// Real line fin[0..L) -> half spectrum out[0..L/2] (forward).
template <typename T, int L>
__device__ __forceinline__ void line_r2c_pow2(const T* fin, C2<T>* out, const C2<T>* tw)
{
    C2<T> x[L];
#pragma unroll
    for (int i = 0; i < L; i++) x[i] = {fin[i], (T)0};
    fft_regs<T, L>(x, tw);
#pragma unroll
    for (int k = 0; k <= L / 2; k++) out[k] = x[k];
}

template <typename T>
__device__ __forceinline__ void line_r2c_any(const T* fin, C2<T>* out, int L, const C2<T>* tw)
{
    for (int k = 0; k <= L / 2; k++) {
        T ar = 0, ai = 0;
        int m = 0;
        for (int x = 0; x < L; x++) {
            const T v = fin[x];
            ar += v * tw[m].re;
            ai += v * tw[m].im;
            m += k;
            if (m >= L) m -= L;
        }
        out[k] = {ar, ai};
    }
}

template <typename T>
__device__ __forceinline__ void line_r2c(const T* fin, C2<T>* out, int L, const C2<T>* tw)
{
    switch (L) {
    case 1: out[0] = {fin[0], (T)0}; return;
    case 2: line_r2c_pow2<T, 2>(fin, out, tw); return;
    case 4: line_r2c_pow2<T, 4>(fin, out, tw); return;
    case 8: line_r2c_pow2<T, 8>(fin, out, tw); return;
    case 16: line_r2c_pow2<T, 16>(fin, out, tw); return;
    case 32: line_r2c_pow2<T, 32>(fin, out, tw); return;
    default: line_r2c_any<T>(fin, out, L, tw); return;
    }
}

// Half spectrum in[0..L/2] -> real line fout[0..L) (backward, unscaled): the
// real part of the inverse transform of the Hermitian extension
// X[L-k] = conj(X[k]); imaginary parts of X[0] and X[L/2] drop out, as c2r.
template <typename T, int L>
__device__ __forceinline__ void line_c2r_pow2(const C2<T>* in, T* fout, const C2<T>* tw)
{
    C2<T> x[L];
#pragma unroll
    for (int k = 0; k <= L / 2; k++) x[k] = in[k];
#pragma unroll
    for (int k = L / 2 + 1; k < L; k++) x[k] = {x[L - k].re, -x[L - k].im};
    fft_regs<T, L>(x, tw);
#pragma unroll
    for (int i = 0; i < L; i++) fout[i] = x[i].re;
}

template <typename T>
__device__ __forceinline__ void line_c2r_any(const C2<T>* in, T* fout, int L, const C2<T>* tw)
{
    for (int x = 0; x < L; x++) {
        T acc = 0;
        int m = 0;  // k x mod L
        for (int k = 0; k < L; k++) {
            const C2<T> X = k <= L / 2 ? in[k] : C2<T>{in[L - k].re, -in[L - k].im};
            acc += X.re * tw[m].re - X.im * tw[m].im;
            m += x;
            if (m >= L) m -= L;
        }
        fout[x] = acc;
    }
}

template <typename T>
__device__ __forceinline__ void line_c2r(const C2<T>* in, T* fout, int L, const C2<T>* tw)
{
    switch (L) {
    case 1: fout[0] = in[0].re; return;
    case 2: line_c2r_pow2<T, 2>(in, fout, tw); return;
    case 4: line_c2r_pow2<T, 4>(in, fout, tw); return;
    case 8: line_c2r_pow2<T, 8>(in, fout, tw); return;
    case 16: line_c2r_pow2<T, 16>(in, fout, tw); return;
    case 32: line_c2r_pow2<T, 32>(in, fout, tw); return;
    default: line_c2r_any<T>(in, fout, L, tw); return;
    }
}

template <typename T>
__device__ __forceinline__ void make_twiddles(C2<T>* tw, int L, int sign)
{
    for (int m = threadIdx.x; m < L; m += blockDim.x) {
        double s, c;
        sincospi(2.0 * m / L, &s, &c);
        tw[m] = {(T)c, (T)(sign * s)};
    }
}

}  // namespace snh
