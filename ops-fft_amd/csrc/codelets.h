// codelets.h - register-resident DFT codelets for gfx950.
//
// Every codelet computes the forward DFT X[k] = sum_n x[n] e^{-2 pi i nk/R} in place on an array of
// R complex values held in VGPRs (all indices compile-time, so nothing spills to scratch).  The
// backward transform is obtained by the caller through conjugation on load and store, so only one
// direction is compiled.
//
// Radices: 2, 3, 4 explicit; odd R by the symmetric-pair formula (used for 5, 7, 9, 11, 13 ...);
// composite R = R1*R2 by one in-register Cooley-Tukey split with constant twiddles.  All constants
// are evaluated in double precision at compile time (constexpr series with exact quarter-turn
// reduction) and rounded to float once.
#pragma once
#include <hip/hip_runtime.h>

namespace ofx {

// ---------------------------------------------------------------------------------------------
// compile-time trigonometry: cos/sin(2 pi num/den), exact reduction by quarter turns
// ---------------------------------------------------------------------------------------------
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double sin_series(double x) {  // |x| <= pi/4
    double x2 = x * x, term = x, sum = x;
    for (int i = 1; i < 14; ++i) {
        term *= -x2 / double((2 * i) * (2 * i + 1));
        sum += term;
    }
    return sum;
}
constexpr double cos_series(double x) {  // |x| <= pi/4
    double x2 = x * x, term = 1.0, sum = 1.0;
    for (int i = 1; i < 14; ++i) {
        term *= -x2 / double((2 * i - 1) * (2 * i));
        sum += term;
    }
    return sum;
}
struct cs_pair {
    double c, s;
};
// (cos, sin) of 2*pi*num/den for any integer num (den > 0).
constexpr cs_pair cos_sin_2pi(long long num, long long den) {
    long long m = num % den;
    if (m < 0) m += den;
    // angle = (pi/2) * (4m/den) = (pi/2) * (q + r/den)
    long long a = 4 * m;
    long long q = a / den;
    long long r = a % den;
    double c0 = 0, s0 = 0;
    if (2 * r <= den) {  // phi = (pi/2) r/den <= pi/4
        double phi = (kPi / 2) * double(r) / double(den);
        c0 = cos_series(phi);
        s0 = sin_series(phi);
    } else {  // phi > pi/4: use the complement
        double psi = (kPi / 2) * double(den - r) / double(den);
        c0 = sin_series(psi);
        s0 = cos_series(psi);
    }
    switch (q & 3) {
        case 0: return {c0, s0};
        case 1: return {-s0, c0};
        case 2: return {-c0, -s0};
        default: return {s0, -c0};
    }
}

// ---------------------------------------------------------------------------------------------
// complex helpers (float2 = interleaved complex64)
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, a.y * b.y), fmaf(a.y, b.x, -a.x * b.y));
}
__device__ __forceinline__ float2 mul_negi(float2 a) { return make_float2(a.y, -a.x); }  // * (-i)
__device__ __forceinline__ float2 mul_posi(float2 a) { return make_float2(-a.y, a.x); }  // * (+i)
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// multiply by the compile-time constant W_R^e = e^{-2 pi i e/R}
template <int R, int E>
__device__ __forceinline__ float2 twiddle_const(float2 a) {
    constexpr int e = ((E % R) + R) % R;
    if constexpr (e == 0) {
        return a;
    } else if constexpr (4 * e == R) {
        return mul_negi(a);
    } else if constexpr (2 * e == R) {
        return make_float2(-a.x, -a.y);
    } else if constexpr (4 * e == 3 * R) {
        return mul_posi(a);
    } else if constexpr (8 * e == R) {  // (1 - i)/sqrt2
        constexpr float h = 0.70710678118654752440f;
        return make_float2(h * (a.x + a.y), h * (a.y - a.x));
    } else if constexpr (8 * e == 3 * R) {  // (-1 - i)/sqrt2
        constexpr float h = 0.70710678118654752440f;
        return make_float2(h * (a.y - a.x), -h * (a.x + a.y));
    } else {
        constexpr cs_pair w = cos_sin_2pi(e, R);
        constexpr float wr = float(w.c);
        constexpr float wi = float(-w.s);
        return make_float2(fmaf(a.x, wr, -a.y * wi), fmaf(a.x, wi, a.y * wr));
    }
}

// ---------------------------------------------------------------------------------------------
// codelets
// ---------------------------------------------------------------------------------------------
template <int R>
struct Dft;

template <>
struct Dft<1> {
    __device__ __forceinline__ static void run(float2 (&)[1]) {}
};

template <>
struct Dft<2> {
    __device__ __forceinline__ static void run(float2 (&x)[2]) {
        float2 a = x[0], b = x[1];
        x[0] = cadd(a, b);
        x[1] = csub(a, b);
    }
};

template <>
struct Dft<3> {
    __device__ __forceinline__ static void run(float2 (&x)[3]) {
        constexpr float c1 = -0.5f;
        constexpr float s1 = 0.86602540378443864676f;  // sin(2pi/3)
        float2 s = cadd(x[1], x[2]);
        float2 d = csub(x[1], x[2]);
        float2 x0 = x[0];
        x[0] = cadd(x0, s);
        float2 t = make_float2(fmaf(c1, s.x, x0.x), fmaf(c1, s.y, x0.y));
        // -i*s1*d = (s1*d.y, -s1*d.x)
        x[1] = make_float2(fmaf(s1, d.y, t.x), fmaf(-s1, d.x, t.y));
        x[2] = make_float2(fmaf(-s1, d.y, t.x), fmaf(s1, d.x, t.y));
    }
};

template <>
struct Dft<4> {
    __device__ __forceinline__ static void run(float2 (&x)[4]) {
        float2 t0 = cadd(x[0], x[2]);
        float2 t1 = csub(x[0], x[2]);
        float2 t2 = cadd(x[1], x[3]);
        float2 t3 = mul_negi(csub(x[1], x[3]));
        x[0] = cadd(t0, t2);
        x[2] = csub(t0, t2);
        x[1] = cadd(t1, t3);
        x[3] = csub(t1, t3);
    }
};

// Odd R (prime or not) by symmetric pairs:
//   X_k     = x0 + sum_m s_m cos(2 pi mk/R) - i sum_m d_m sin(2 pi mk/R)
//   X_{R-k} = x0 + sum_m s_m cos(2 pi mk/R) + i sum_m d_m sin(2 pi mk/R)
// with s_m = x_m + x_{R-m}, d_m = x_m - x_{R-m}, m, k = 1..(R-1)/2.
template <int R>
struct DftOdd {
    static constexpr int H = (R - 1) / 2;
    __device__ __forceinline__ static void run(float2 (&x)[R]) {
        float2 s[H + 1], d[H + 1];
        float2 x0 = x[0];
        float2 sum = x0;
#pragma unroll
        for (int m = 1; m <= H; ++m) {
            s[m] = cadd(x[m], x[R - m]);
            d[m] = csub(x[m], x[R - m]);
            sum = cadd(sum, s[m]);
        }
        x[0] = sum;
        body<1>(x, x0, s, d);
    }
    template <int K>
    __device__ __forceinline__ static void body(float2 (&x)[R], float2 x0, const float2 (&s)[H + 1],
                                                const float2 (&d)[H + 1]) {
        if constexpr (K <= H) {
            float2 re = x0;
            float2 im = make_float2(0.f, 0.f);
            accum<K, 1>(re, im, s, d);
            // -i*im = (im.y, -im.x)
            x[K] = make_float2(re.x + im.y, re.y - im.x);
            x[R - K] = make_float2(re.x - im.y, re.y + im.x);
            body<K + 1>(x, x0, s, d);
        }
    }
    template <int K, int M>
    __device__ __forceinline__ static void accum(float2& re, float2& im, const float2 (&s)[H + 1],
                                                 const float2 (&d)[H + 1]) {
        if constexpr (M <= H) {
            constexpr cs_pair w = cos_sin_2pi((long long)M * K, R);
            constexpr float c = float(w.c);
            constexpr float sn = float(w.s);
            re.x = fmaf(c, s[M].x, re.x);
            re.y = fmaf(c, s[M].y, re.y);
            im.x = fmaf(sn, d[M].x, im.x);
            im.y = fmaf(sn, d[M].y, im.y);
            accum<K, M + 1>(re, im, s, d);
        }
    }
};

// Composite R = R1*R2: n = R2*n1 + n2, k = k1 + R1*k2.
template <int R1, int R2>
struct DftSplit {
    static constexpr int R = R1 * R2;
    __device__ __forceinline__ static void run(float2 (&x)[R]) {
        float2 y[R2][R1];
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) {
#pragma unroll
            for (int n1 = 0; n1 < R1; ++n1) y[n2][n1] = x[R2 * n1 + n2];
            Dft<R1>::run(y[n2]);
        }
        twid<1, 1>(y);
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) {
            float2 z[R2];
#pragma unroll
            for (int n2 = 0; n2 < R2; ++n2) z[n2] = y[n2][k1];
            Dft<R2>::run(z);
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) x[k1 + R1 * k2] = z[k2];
        }
    }
    // y[n2][k1] *= W_R^{n2*k1}, unrolled at compile time so each factor is a constant
    template <int N2, int K1>
    __device__ __forceinline__ static void twid(float2 (&y)[R2][R1]) {
        if constexpr (N2 < R2) {
            if constexpr (K1 < R1) {
                y[N2][K1] = twiddle_const<R, N2 * K1>(y[N2][K1]);
                twid<N2, K1 + 1>(y);
            } else {
                twid<N2 + 1, 1>(y);
            }
        }
    }
};

template <> struct Dft<5> : DftOdd<5> {};
template <> struct Dft<7> : DftOdd<7> {};
template <> struct Dft<11> : DftOdd<11> {};
template <> struct Dft<13> : DftOdd<13> {};
template <> struct Dft<17> : DftOdd<17> {};
template <> struct Dft<19> : DftOdd<19> {};
template <> struct Dft<23> : DftOdd<23> {};
template <> struct Dft<29> : DftOdd<29> {};
template <> struct Dft<31> : DftOdd<31> {};
template <> struct Dft<37> : DftOdd<37> {};
template <> struct Dft<41> : DftOdd<41> {};
template <> struct Dft<43> : DftOdd<43> {};
template <> struct Dft<47> : DftOdd<47> {};
template <> struct Dft<6> : DftSplit<2, 3> {};
template <> struct Dft<8> : DftSplit<2, 4> {};
template <> struct Dft<9> : DftSplit<3, 3> {};
template <> struct Dft<10> : DftSplit<2, 5> {};
template <> struct Dft<12> : DftSplit<4, 3> {};
template <> struct Dft<15> : DftSplit<3, 5> {};
template <> struct Dft<16> : DftSplit<4, 4> {};
template <> struct Dft<25> : DftSplit<5, 5> {};
template <> struct Dft<27> : DftSplit<3, 9> {};
template <> struct Dft<32> : DftSplit<4, 8> {};
template <> struct Dft<49> : DftSplit<7, 7> {};
template <> struct Dft<64> : DftSplit<8, 8> {};

}  // namespace ofx
