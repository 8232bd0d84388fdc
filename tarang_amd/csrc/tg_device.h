// tg_device.h -- device-side building blocks for the gfx950 pseudo-spectral path.
//
// Complex arithmetic, radix-2/3/4/5/8 butterflies and a register-resident
// Stockham FFT for lines of L = 8*TPL points handled by TPL threads
// (L = 2^a 3^b 5^c with 8 | L).
//
// Register convention used by every kernel in this library: thread j of a
// line holds the 8 points at positions j + TPL*s, s = 0..7.  The first
// radix-8 Stockham pass reads exactly these positions and the last pass
// produces exactly these positions, so a forward transform can feed an
// inverse transform (and pointwise epilogues) without touching LDS.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tg {

template <typename T>
struct cpx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> mk(T a, T b) { cpx<T> r; r.x = a; r.y = b; return r; }
template <typename T>
__device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ cpx<T> operator*(cpx<T> a, cpx<T> b)
{
    return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__device__ __forceinline__ cpx<T> operator*(T s, cpx<T> a) { return mk<T>(s * a.x, s * a.y); }
template <typename T>
__device__ __forceinline__ cpx<T> conj(cpx<T> a) { return mk<T>(a.x, -a.y); }
// multiply by i*s (s real): i*s*(x+iy) = -s*y + i*s*x
template <typename T>
__device__ __forceinline__ cpx<T> mul_i(T s, cpx<T> a) { return mk<T>(-s * a.y, s * a.x); }
// a * conj(w) when dir>0 (inverse), a*w when dir<0 (forward); w from the forward table
template <int DIR, typename T>
__device__ __forceinline__ cpx<T> twmul(cpx<T> a, cpx<T> w)
{
    if (DIR < 0) return a * w;
    return mk<T>(a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y);
}
// multiply by -i (forward) or +i (inverse)
template <int DIR, typename T>
__device__ __forceinline__ cpx<T> rot90(cpx<T> a)
{
    if (DIR < 0) return mk<T>(a.y, -a.x);
    return mk<T>(-a.y, a.x);
}

// r2 = cos(pi/4); s3 = sin(2 pi/3); c5a, s5a = cos, sin(2 pi/5); c5b, s5b =
// cos, sin(4 pi/5)
template <typename T> struct consts;
template <> struct consts<double> {
    static constexpr double r2 = 0.70710678118654752440;
    static constexpr double s3 = 0.86602540378443864676;
    static constexpr double c5a = 0.30901699437494742410, s5a = 0.95105651629515357212;
    static constexpr double c5b = -0.80901699437494742410, s5b = 0.58778525229247312917;
};
template <> struct consts<float> {
    static constexpr float r2 = 0.70710678118654752440f;
    static constexpr float s3 = 0.86602540378443864676f;
    static constexpr float c5a = 0.30901699437494742410f, s5a = 0.95105651629515357212f;
    static constexpr float c5b = -0.80901699437494742410f, s5b = 0.58778525229247312917f;
};

// In-place DFT of size R on v[0..R-1], natural order output.
template <int DIR, typename T>
__device__ __forceinline__ void dft2(cpx<T> &a0, cpx<T> &a1)
{
    cpx<T> t = a0 - a1;
    a0 = a0 + a1;
    a1 = t;
}

template <int DIR, typename T>
__device__ __forceinline__ void dft4(cpx<T> &a0, cpx<T> &a1, cpx<T> &a2, cpx<T> &a3)
{
    cpx<T> b0 = a0 + a2, b2 = a0 - a2, b1 = a1 + a3, b3 = rot90<DIR>(a1 - a3);
    a0 = b0 + b1;
    a2 = b0 - b1;
    a1 = b2 + b3;
    a3 = b2 - b3;
}

template <int DIR, typename T>
__device__ __forceinline__ void dft8(cpx<T> *a)
{
    const T r = consts<T>::r2;
    cpx<T> b0 = a[0] + a[4], b4 = a[0] - a[4];
    cpx<T> b1 = a[1] + a[5], b5 = a[1] - a[5];
    cpx<T> b2 = a[2] + a[6], b6 = a[2] - a[6];
    cpx<T> b3 = a[3] + a[7], b7 = a[3] - a[7];
    // b5 *= W8, b6 *= W8^2, b7 *= W8^3 with W8 = exp(DIR*i*pi/4)
    if (DIR < 0) {
        b5 = mk<T>(r * (b5.x + b5.y), r * (b5.y - b5.x));
        b7 = mk<T>(r * (b7.y - b7.x), -r * (b7.x + b7.y));
    } else {
        b5 = mk<T>(r * (b5.x - b5.y), r * (b5.x + b5.y));
        b7 = mk<T>(-r * (b7.x + b7.y), r * (b7.x - b7.y));
    }
    b6 = rot90<DIR>(b6);
    cpx<T> c0 = b0 + b2, c2 = b0 - b2, c1 = b1 + b3, c3 = rot90<DIR>(b1 - b3);
    cpx<T> c4 = b4 + b6, c6 = b4 - b6, c5 = b5 + b7, c7 = rot90<DIR>(b5 - b7);
    a[0] = c0 + c1;
    a[4] = c0 - c1;
    a[2] = c2 + c3;
    a[6] = c2 - c3;
    a[1] = c4 + c5;
    a[5] = c4 - c5;
    a[3] = c6 + c7;
    a[7] = c6 - c7;
}

// X_k = sum_n a_n w^(nk), w = exp(DIR 2 pi i / 3): X_1,2 = m +- s3 rot90(a1 - a2)
// with m = a0 - (a1 + a2)/2
template <int DIR, typename T>
__device__ __forceinline__ void dft3(cpx<T> *a)
{
    const cpx<T> t = a[1] + a[2];
    const cpx<T> m = a[0] - (T)0.5 * t;
    const cpx<T> d = consts<T>::s3 * rot90<DIR>(a[1] - a[2]);
    a[0] = a[0] + t;
    a[1] = m + d;
    a[2] = m - d;
}

// w = exp(DIR 2 pi i / 5): with t1 = a1 + a4, t2 = a2 + a3, t3 = a1 - a4,
// t4 = a2 - a3, X_1,4 = a0 + c5a t1 + c5b t2 +- rot90(s5a t3 + s5b t4) and
// X_2,3 = a0 + c5b t1 + c5a t2 +- rot90(s5b t3 - s5a t4)
template <int DIR, typename T>
__device__ __forceinline__ void dft5(cpx<T> *a)
{
    using C = consts<T>;
    const cpx<T> t1 = a[1] + a[4], t2 = a[2] + a[3], t3 = a[1] - a[4], t4 = a[2] - a[3];
    const cpx<T> b1 = a[0] + (C::c5a * t1 + C::c5b * t2);
    const cpx<T> b2 = a[0] + (C::c5b * t1 + C::c5a * t2);
    const cpx<T> d1 = rot90<DIR>(C::s5a * t3 + C::s5b * t4);
    const cpx<T> d2 = rot90<DIR>(C::s5b * t3 - C::s5a * t4);
    a[0] = a[0] + (t1 + t2);
    a[1] = b1 + d1;
    a[4] = b1 - d1;
    a[2] = b2 + d2;
    a[3] = b2 - d2;
}

template <int R, int DIR, typename T>
__device__ __forceinline__ void dftR(cpx<T> *a)
{
    if constexpr (R == 8) dft8<DIR>(a);
    else if constexpr (R == 4) dft4<DIR>(a[0], a[1], a[2], a[3]);
    else if constexpr (R == 2) dft2<DIR>(a[0], a[1]);
    else if constexpr (R == 5) dft5<DIR>(a);
    else if constexpr (R == 3) dft3<DIR>(a);
}

// Odd part of a line length (the product of its radix-3 and radix-5 factors).
__host__ __device__ constexpr int odd_part(int L)
{
    while (L % 2 == 0) L /= 2;
    return L;
}
// Lengths the transforms take: 2^a 3^b 5^c with 8 | L.
__host__ __device__ constexpr bool line_length_ok(int L)
{
    if (L < 8 || L % 8 != 0) return false;
    int m = odd_part(L);
    while (m % 5 == 0) m /= 5;
    while (m % 3 == 0) m /= 3;
    return m == 1;
}
// The odd passes come first: radix 5 while it divides, then radix 3.
__host__ __device__ constexpr int odd_radix(int L, int p)
{
    int m = odd_part(L);
    for (int i = 0; i < p; i++) m /= (m % 5 == 0 ? 5 : 3);
    return m % 5 == 0 ? 5 : 3;
}
__host__ __device__ constexpr int num_odd_passes(int L)
{
    int n = 0;
    for (int m = odd_part(L); m > 1; n++) m /= (m % 5 == 0 ? 5 : 3);
    return n;
}
// Radix of power-of-two pass p of a line of L points (over L / odd_part(L)):
// 8 while it divides, then the rest.
__host__ __device__ constexpr int pass_radix(int L, int p)
{
    int rem = L / odd_part(L);
    for (int i = 0; i < p; i++) rem /= 8;
    return rem >= 8 ? 8 : rem;
}
__host__ __device__ constexpr int num_passes(int L)
{
    int n = 0;
    int rem = L / odd_part(L);
    while (rem > 1) { rem = rem >= 8 ? rem / 8 : 1; n++; }
    return n;
}

// Block barrier that orders LDS accesses only.  The pass kernels exchange
// data between threads through LDS alone, so their barriers need not wait
// for global memory: __syncthreads' fence waits for every outstanding global
// load and store (stores count in vmcnt on this architecture, so each
// barrier after a tile store would wait out the write acknowledgements) and
// for direct-to-LDS loads in flight.  Direct-to-LDS tiles are waited for
// explicitly (dma_wait) before the barrier that publishes them.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// lds_barrier for a point where a direct-to-LDS load into a tile that no
// thread touches until its own wait is in flight: the workgroup-scope release
// fence would wait for that load too (it counts as a pending LDS write).  The
// wave's own LDS instructions are complete before it reaches the barrier
// (s_waitcnt lgkmcnt(0)), as with the workgroup fence, and the
// wavefront-scope fences keep the compiler from moving LDS accesses across
// the barrier; the direct-to-LDS tile is published later by its own wait
// (dma_wait) and an lds_barrier.
__device__ __forceinline__ void lds_barrier_async()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt and expcnt at their maxima
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// Synchronisation of the threads that share one line.  A line of L points
// is handled by TPL = L/8 consecutive threads; when TPL divides 64 the line
// lives inside one wavefront, whose LDS operations execute in order, so only
// the compiler has to be kept from moving LDS accesses across the exchange
// (wavefront-scope fence on LDS only: global loads and stores may still be
// moved across it, e.g. the next line's loads ahead of this line's FFT).
// Longer lines span waves, and lines of 3 or 5 times a power of two
// straddle them: those need the block barrier (every thread of the block
// takes part in every line transform, so the barriers match).
template <int TPL>
__device__ __forceinline__ void line_sync()
{
    if constexpr (TPL <= 64 && 64 % TPL == 0) {
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront", "local");
    } else {
        lds_barrier();
    }
}

// LDS position with one pad entry every 16 points (breaks the stride-8
// write pattern of the first radix-8 pass across banks).
__device__ __forceinline__ int lpad(int i) { return i + (i >> 4); }
__host__ __device__ constexpr int lpad_len(int L) { return L + L / 16 + 1; }

// One Stockham pass, radix R, stride NS (product of the earlier radices),
// over a line of L points; thread j of TPL = L/8.  v[] holds positions
// j + TPL*s on entry.  If LAST, the results stay in v[] at the same
// positions; otherwise they are written to the LDS line buffer (the caller
// synchronises and reloads).
template <int L, int R, int NS, int DIR, bool LAST, typename T>
__device__ __forceinline__ void stockham_pass(cpx<T> *v, cpx<T> *lds, const cpx<T> *__restrict__ tw, int j)
{
    constexpr int TPL = L / 8;
    constexpr int Q = 8 / R;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        cpx<T> a[R];
        const int b = j + TPL * q;
        const int k = b % NS;
#pragma unroll
        for (int r = 0; r < R; r++) a[r] = v[q + r * Q];
        if constexpr (NS > 1) {
            const int step = (L / (NS * R)) * k;
#pragma unroll
            for (int r = 1; r < R; r++) a[r] = twmul<DIR>(a[r], tw[r * step]);
        }
        dftR<R, DIR>(a);
        if constexpr (LAST) {
#pragma unroll
            for (int r = 0; r < R; r++) v[q + r * Q] = a[r];
        } else {
            const int base = (b / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) lds[lpad(base + r * NS)] = a[r];
        }
    }
}

// One odd-radix Stockham pass (R = 3 or 5), LDS to LDS: the L/R butterflies
// are spread over the TPL threads (ceil(8/R) each, the last round partial);
// every input is read before the first output is written.
template <int L, int R, int NS, int DIR, typename T>
__device__ __forceinline__ void odd_pass(cpx<T> *lds, const cpx<T> *__restrict__ tw, int j)
{
    constexpr int TPL = L / 8, NB = L / R, QM = (NB + TPL - 1) / TPL;
    cpx<T> a[QM][R];
#pragma unroll
    for (int q = 0; q < QM; q++) {
        const int b = j + TPL * q;
        if (b < NB) {
#pragma unroll
            for (int r = 0; r < R; r++) a[q][r] = lds[lpad(b + r * NB)];
        }
    }
    line_sync<TPL>();
#pragma unroll
    for (int q = 0; q < QM; q++) {
        const int b = j + TPL * q;
        if (b < NB) {
            const int k = b % NS;
            if constexpr (NS > 1) {
                const int step = (L / (NS * R)) * k;
#pragma unroll
                for (int r = 1; r < R; r++) a[q][r] = twmul<DIR>(a[q][r], tw[r * step]);
            }
            dftR<R, DIR>(a[q]);
            const int base = (b / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) lds[lpad(base + r * NS)] = a[q][r];
        }
    }
    line_sync<TPL>();
}

template <int L, int DIR, int P, int NS, typename T>
__device__ __forceinline__ void odd_passes(cpx<T> *lds, const cpx<T> *__restrict__ tw, int j)
{
    if constexpr (P < num_odd_passes(L)) {
        constexpr int R = odd_radix(L, P);
        odd_pass<L, R, NS, DIR>(lds, tw, j);
        odd_passes<L, DIR, P + 1, NS * R>(lds, tw, j);
    }
}

template <int L, int DIR, int P, int NS, typename T>
__device__ __forceinline__ void fft_passes(cpx<T> *v, cpx<T> *lds, const cpx<T> *__restrict__ tw, int j)
{
    constexpr int NP = num_passes(L);
    if constexpr (P < NP) {
        constexpr int R = pass_radix(L, P);
        constexpr bool LAST = (P == NP - 1);
        stockham_pass<L, R, NS, DIR, LAST>(v, lds, tw, j);
        if constexpr (!LAST) {
            constexpr int TPL = L / 8;
            line_sync<TPL>();
#pragma unroll
            for (int s = 0; s < 8; s++) v[s] = lds[lpad(j + TPL * s)];
            line_sync<TPL>();
            fft_passes<L, DIR, P + 1, NS * R>(v, lds, tw, j);
        }
    }
}

// Full line FFT of L = 8*TPL points (line_length_ok(L), up to 4096), values
// in registers at positions j + TPL*s.  Exponent sign DIR (-1 forward, +1
// inverse), unnormalised.  tw[k] = exp(-2 pi i k / L).  Uses the LDS line
// buffer (lpad_len(L) entries) for the exchanges between passes; every
// thread of the line must call it (line_sync); the caller must have
// synchronised if the LDS line was in use before.  The radix-3/5 passes run
// first, through LDS; the power-of-two passes follow, the last one leaving
// the results in registers.
template <int L, int DIR, typename T>
__device__ __forceinline__ void fft_line(cpx<T> *v, cpx<T> *lds, const cpx<T> *__restrict__ tw, int j)
{
    static_assert(line_length_ok(L), "line length must be 2^a 3^b 5^c with 8 | L");
    if constexpr (odd_part(L) > 1) {
        constexpr int TPL = L / 8;
#pragma unroll
        for (int s = 0; s < 8; s++) lds[lpad(j + TPL * s)] = v[s];
        line_sync<TPL>();
        odd_passes<L, DIR, 0, 1>(lds, tw, j);
#pragma unroll
        for (int s = 0; s < 8; s++) v[s] = lds[lpad(j + TPL * s)];
        line_sync<TPL>();
        fft_passes<L, DIR, 0, odd_part(L)>(v, lds, tw, j);
    } else {
        fft_passes<L, DIR, 0, 1>(v, lds, tw, j);
    }
}

}  // namespace tg
