// Batched y transforms of the subdomain solver (DST-I, DCT-II/III and the
// two quarter-wave pairs), as packed complex FFTs staged in LDS.
//
// Reference semantics: PoissonFFT2D_{DD,NN,DN,ND}::transform and
// ::transform_transpose (pe/poisson/poisson_fft2d.cpp:37-57, 105-133,
// 180-205, 242-262).  The reference runs one FFTW r2c per row (two for the
// NN/ND inverse); here two rows a, b of the field form one complex sequence
// z = ext(a) + i ext(b) of length L, one FFT of length L is done in LDS, and
// the two rows' coefficients are split out of the spectrum:
//   DD     : odd extension, L = 2N+2, spectrum purely imaginary per row
//   NN fwd : half-sample even extension, L = 2N, phase exp(-i pi m / 2N)
//   NN inv : Hermitian extension of the phased coefficients, inverse FFT, L = 2N
//   DN, ND : L = 2N+1, split by Z(m) +- conj Z(L-m)
// Each output equals the reference's coefficient (same signs and mode order),
// so transform results are directly comparable with the reference's.
//
// The FFT is a mixed-radix Stockham autosort: radix 4/2/3/5 butterflies in
// registers, any other prime by a two-phase in-LDS DFT.  One workgroup owns
// one row pair of one right-hand side: it loads both rows (coalesced), runs
// all passes between two LDS buffers and writes the coefficients back in place.
#include "fp_dft_tables.h"
#include "fp_internal.h"

#include <algorithm>
#include <cmath>

namespace fp {

template <typename T>
struct alignas(2 * sizeof(T)) C2 {
    T x, y;
};

template <typename T>
__device__ __forceinline__ C2<T> cadd(C2<T> a, C2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ C2<T> csub(C2<T> a, C2<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
// a * conj(b)
template <typename T>
__device__ __forceinline__ C2<T> cmulc(C2<T> a, C2<T> b) { return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y}; }
template <typename T>
__device__ __forceinline__ C2<T> cscale(C2<T> a, T s) { return {a.x * s, a.y * s}; }
// multiply by -i (FWD) or +i (!FWD)
template <bool FWD, typename T>
__device__ __forceinline__ C2<T> mul_mi(C2<T> a) { return FWD ? C2<T>{a.y, -a.x} : C2<T>{-a.y, a.x}; }

// src/dst: the rows read and the coefficients written (the same buffer for an
// in-place transform).  src64/dst64 = 1: that buffer holds doubles (an FP32
// transform then reads or writes FP64 fields directly); 0: it holds T.
template <typename T>
struct YArgs {
    const void*  src;
    void*        dst;
    long long    lds, ldd;
    int          src64, dst64;
    int          nx, N, L, kind, inverse, nrad;
    int          rad[kMaxRadices];
    const C2<T>* W;
    const C2<T>* aux;
    T            c0, c1;
};

// Up to kMaxBatch transforms of different subdomains in one launch, passed
// by value: workgroups start[q] .. start[q+1]-1 belong to problem q, one per
// (row pair, right-hand side).
template <typename T>
struct YBatch {
    int      nprob;
    int      start[kMaxBatch + 1];
    YArgs<T> p[kMaxBatch];
};

// Twiddle factors exp(-2 pi i t / L) from two small LDS tables:
// W^t = Whi[t >> 6] * Wlo[t & 63]  (error of one extra rounded product)
template <typename T>
struct Tw {
    const C2<T>* lo;  // 64 entries: W^m
    const C2<T>* hi;  // L/64 + 1 entries: W^(64 m)
};

template <bool FWD, typename T>
__device__ __forceinline__ C2<T> twiddle(const Tw<T> W, int t, C2<T> v)
{
    const C2<T> w = cmul(W.hi[t >> 6], W.lo[t & 63]);
    return FWD ? cmul(v, w) : cmulc(v, w);
}

// One Stockham pass with a register butterfly of radix R.
template <int R, bool FWD, int NT, typename T>
__device__ void pass_reg(const C2<T>* __restrict__ src, C2<T>* __restrict__ dst, const Tw<T> W, int L, int Ns)
{
    const int nb = L / R, stride = L / (Ns * R);
    for (int jb = threadIdx.x; jb < nb; jb += NT) {
        const int k = jb % Ns;
        C2<T>     v[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            v[r] = src[jb + r * nb];
            if (r > 0 && Ns > 1) v[r] = twiddle<FWD>(W, r * k * stride, v[r]);
        }
        C2<T> o[R];
        if constexpr (R == 2) {
            o[0] = cadd(v[0], v[1]);
            o[1] = csub(v[0], v[1]);
        } else if constexpr (R == 4) {
            const C2<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
            const C2<T> t2 = cadd(v[1], v[3]), t3 = mul_mi<FWD>(csub(v[1], v[3]));
            o[0] = cadd(t0, t2);
            o[2] = csub(t0, t2);
            o[1] = cadd(t1, t3);
            o[3] = csub(t1, t3);
        } else if constexpr (R == 8) {
            const T     h  = T(0.70710678118654752440084436210485);
            const C2<T> a0 = cadd(v[0], v[4]), a1 = csub(v[0], v[4]);
            const C2<T> a2 = cadd(v[2], v[6]), a3 = mul_mi<FWD>(csub(v[2], v[6]));
            const C2<T> a4 = cadd(v[1], v[5]), a5 = csub(v[1], v[5]);
            const C2<T> a6 = cadd(v[3], v[7]), a7 = mul_mi<FWD>(csub(v[3], v[7]));
            const C2<T> b0 = cadd(a0, a2), b2 = csub(a0, a2), b1 = cadd(a1, a3), b3 = csub(a1, a3);
            C2<T>       c0 = cadd(a4, a6), c2 = csub(a4, a6), c1 = cadd(a5, a7), c3 = csub(a5, a7);
            // c1 *= w8, c2 *= w8^2 = -+i, c3 *= w8^3  (w8 = exp(-+ i pi / 4))
            c1 = FWD ? C2<T>{h * (c1.x + c1.y), h * (c1.y - c1.x)} : C2<T>{h * (c1.x - c1.y), h * (c1.y + c1.x)};
            c2 = mul_mi<FWD>(c2);
            c3 = FWD ? C2<T>{h * (c3.y - c3.x), -h * (c3.x + c3.y)} : C2<T>{-h * (c3.x + c3.y), h * (c3.x - c3.y)};
            o[0] = cadd(b0, c0);
            o[4] = csub(b0, c0);
            o[1] = cadd(b1, c1);
            o[5] = csub(b1, c1);
            o[2] = cadd(b2, c2);
            o[6] = csub(b2, c2);
            o[3] = cadd(b3, c3);
            o[7] = csub(b3, c3);
        } else if constexpr (R == 3) {
            const T     h = T(0.86602540378443864676372317075294);
            const C2<T> s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
            const C2<T> m = {v[0].x - T(0.5) * s.x, v[0].y - T(0.5) * s.y};
            const C2<T> e = mul_mi<FWD>(cscale(d, h));
            o[0] = cadd(v[0], s);
            o[1] = cadd(m, e);
            o[2] = csub(m, e);
        } else if constexpr (R == 5) {
            const T     c1 = T(0.30901699437494742410229341718282), c2 = T(-0.80901699437494742410229341718282);
            const T     s1 = T(0.95105651629515357211643933337938), s2 = T(0.58778525229247312916870595463907);
            const C2<T> a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
            const C2<T> a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
            const C2<T> m1 = {v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y};
            const C2<T> m2 = {v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y};
            const C2<T> n1 = mul_mi<FWD>(C2<T>{s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y});
            const C2<T> n2 = mul_mi<FWD>(C2<T>{s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y});
            o[0] = {v[0].x + a1.x + a2.x, v[0].y + a1.y + a2.y};
            o[1] = cadd(m1, n1);
            o[4] = csub(m1, n1);
            o[2] = cadd(m2, n2);
            o[3] = csub(m2, n2);
        }
        const int idxD = (jb / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) dst[idxD + r * Ns] = o[r];
    }
}

// One odd-prime butterfly, the share of thread group g: outputs q = QB..QE-1
// and R-q, plus X_0 for group 0.  With a_r = x_r + x_{R-r}, b_r = x_r - x_{R-r}:
//   X_q = x_0 + A_q -+ i B_q,  X_{R-q} = x_0 + A_q +- i B_q,
//   A_q = sum_r a_r cos(2 pi q r / R),  B_q = sum_r b_r sin(2 pi q r / R).
// The inputs stream through registers one pair at a time; only the group's
// accumulators stay live.  q and r are compile-time, so the cosines are
// immediates.
template <int R, int G, int g, bool FWD, bool TW, typename T>
__device__ __forceinline__ void odd_group(const C2<T>* __restrict__ src, C2<T>* __restrict__ dst, const Tw<T> W,
                                          int jb, int nb, int k, int stride, int base, int Ns)
{
    constexpr int H  = (R - 1) / 2;
    constexpr int QP = (H + G - 1) / G;
    constexpr int QB = 1 + g * QP;
    constexpr int QE = (QB + QP < H + 1) ? QB + QP : H + 1;
    constexpr int NQ = QE - QB;
    const C2<T>   v0 = src[jb];
    T             ax[NQ], ay[NQ], bx[NQ], by[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        ax[q] = v0.x;
        ay[q] = v0.y;
        bx[q] = T(0);
        by[q] = T(0);
    }
    T x0x = v0.x, x0y = v0.y;
#pragma unroll
    for (int r = 1; r <= H; r++) {
        C2<T> x1 = src[jb + r * nb], x2 = src[jb + (R - r) * nb];
        if constexpr (TW) {
            x1 = twiddle<FWD>(W, r * k * stride, x1);
            x2 = twiddle<FWD>(W, (R - r) * k * stride, x2);
        }
        const C2<T> a = cadd(x1, x2), b = csub(x1, x2);
        if constexpr (g == 0) {
            x0x += a.x;
            x0y += a.y;
        }
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const T c  = T(DftTab<R>::c[((QB + q) * r) % R]);
            const T sn = T(DftTab<R>::s[((QB + q) * r) % R]);
            ax[q] += a.x * c;
            ay[q] += a.y * c;
            bx[q] += b.x * sn;
            by[q] += b.y * sn;
        }
        // one input pair in flight at a time: bounds the live registers to the
        // accumulators (other waves hide the LDS latency)
        __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (g == 0) dst[base] = C2<T>{x0x, x0y};
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const C2<T> xq  = FWD ? C2<T>{ax[q] + by[q], ay[q] - bx[q]} : C2<T>{ax[q] - by[q], ay[q] + bx[q]};
        const C2<T> xrq = FWD ? C2<T>{ax[q] - by[q], ay[q] + bx[q]} : C2<T>{ax[q] + by[q], ay[q] - bx[q]};
        dst[base + (QB + q) * Ns]     = xq;
        dst[base + (R - QB - q) * Ns] = xrq;
    }
}

// One Stockham pass of odd prime radix R: G threads share a butterfly, each
// computing about five of its (q, R-q) output pairs.
template <int R, bool FWD, int NT, bool TW, typename T>
__device__ void pass_odd_tw(const C2<T>* __restrict__ src, C2<T>* __restrict__ dst, const Tw<T> W, int L, int Ns)
{
    constexpr int H = (R - 1) / 2;
    constexpr int G = (H + 4) / 5;
    const int     nb = L / R, stride = L / (Ns * R);
    for (int t = threadIdx.x; t < nb * G; t += NT) {
        const int jb = t / G, grp = t - jb * G, k = jb % Ns;
        const int base = (jb / Ns) * Ns * R + k;
        if (grp == 0) odd_group<R, G, 0, FWD, TW>(src, dst, W, jb, nb, k, stride, base, Ns);
        if constexpr (G > 1) if (grp == 1) odd_group<R, G, 1, FWD, TW>(src, dst, W, jb, nb, k, stride, base, Ns);
        if constexpr (G > 2) if (grp == 2) odd_group<R, G, 2, FWD, TW>(src, dst, W, jb, nb, k, stride, base, Ns);
        if constexpr (G > 3) if (grp == 3) odd_group<R, G, 3, FWD, TW>(src, dst, W, jb, nb, k, stride, base, Ns);
        if constexpr (G > 4) if (grp == 4) odd_group<R, G, 4, FWD, TW>(src, dst, W, jb, nb, k, stride, base, Ns);
    }
}

// the first pass (Ns = 1) has unit twiddles: a separate instantiation skips them
template <int R, bool FWD, int NT, typename T>
__device__ void pass_odd(const C2<T>* __restrict__ src, C2<T>* __restrict__ dst, const Tw<T> W, int L, int Ns)
{
    if (Ns == 1) pass_odd_tw<R, FWD, NT, false>(src, dst, W, L, Ns);
    else pass_odd_tw<R, FWD, NT, true>(src, dst, W, L, Ns);
}

// Stockham pass of any radix R: twiddle in place, then each output is an
// R-term sum read from LDS.
template <bool FWD, int NT, typename T>
__device__ void pass_generic(C2<T>* __restrict__ src, C2<T>* __restrict__ dst, const Tw<T> W, int L, int Ns,
                             int R)
{
    const int nb = L / R, stride = L / (Ns * R);
    for (int idx = threadIdx.x; idx < L; idx += NT) {
        const int r = idx / nb, jb = idx - r * nb, k = jb % Ns;
        if (r > 0) src[idx] = twiddle<FWD>(W, r * k * stride, src[idx]);
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < L; idx += NT) {
        const int q = idx / nb, jb = idx - q * nb, k = jb % Ns;
        C2<T>     acc = src[jb];
        int       t   = 0;
        for (int r = 1; r < R; r++) {
            t += q * nb;
            if (t >= L) t -= L;
            acc = cadd(acc, twiddle<FWD>(W, t, src[jb + r * nb]));
        }
        dst[(jb / Ns) * Ns * R + k + q * Ns] = acc;
    }
}

// Radix dispatch: 2, 3, 4, 5, 7, 8 in every kernel, one larger odd prime P1
// per kernel instantiation (keeping one codelet's registers per kernel), any
// other factor through the generic pass.
template <int P1, bool FWD, int NT, typename T>
__device__ C2<T>* run_fft(const YArgs<T>& a, const Tw<T> W, C2<T>* A, C2<T>* B)
{
    C2<T>* src = A;
    C2<T>* dst = B;
    int    Ns  = 1;
    for (int s = 0; s < a.nrad; s++) {
        const int R = a.rad[s];
        if (R == 2) pass_reg<2, FWD, NT>(src, dst, W, a.L, Ns);
        else if (R == 3) pass_reg<3, FWD, NT>(src, dst, W, a.L, Ns);
        else if (R == 4) pass_reg<4, FWD, NT>(src, dst, W, a.L, Ns);
        else if (R == 5) pass_reg<5, FWD, NT>(src, dst, W, a.L, Ns);
        else if (R == 8) pass_reg<8, FWD, NT>(src, dst, W, a.L, Ns);
        else if (R == 7) pass_odd<7, FWD, NT>(src, dst, W, a.L, Ns);
        else if (P1 > 7 && R == P1) pass_odd<(P1 > 7 ? P1 : 11), FWD, NT>(src, dst, W, a.L, Ns);
        else pass_generic<FWD, NT>(src, dst, W, a.L, Ns, R);
        __syncthreads();
        C2<T>* t = src;
        src      = dst;
        dst      = t;
        Ns *= R;
    }
    return src;
}

template <typename T>
__device__ __forceinline__ T sgn_pow(int p) { return (p & 1) ? T(-1) : T(1); }

// values per thread and row held in registers by the load and store phases
constexpr int VPT = 8;

template <int NT, int P1, typename T>
__global__ __launch_bounds__(NT) void k_ytransform(YBatch<T> bt)
{
    int q = 0;
    while (q + 1 < bt.nprob && (int)blockIdx.x >= bt.start[q + 1]) q++;
    const YArgs<T>& a      = bt.p[q];
    const int       npairs = (a.nx + 1) >> 1;
    const int       local  = (int)blockIdx.x - bt.start[q];
    const int       rhs    = local / npairs, pair = local - rhs * npairs;

    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2<T>* A   = reinterpret_cast<C2<T>*>(smem_raw);
    C2<T>* B   = A + a.L;
    C2<T>* tlo = B + a.L;           // 64 twiddles W^m
    C2<T>* thi = tlo + 64;          // L/64 + 1 twiddles W^(64m)
    const int nhi = a.L / 64 + 1;
    T*     y0  = reinterpret_cast<T*>(thi + nhi);  // NN inverse: y_0 of rows a, b
    for (int q = threadIdx.x; q < 64 + nhi; q += NT) tlo[q] = a.W[q];
    const Tw<T> W{tlo, thi};

    const int  N = a.N, L = a.L;
    const int  ra = 2 * pair, rb = ra + 1;
    const bool hb = rb < a.nx;
    const long long ia = (long long)rhs * a.lds + (long long)ra * N;  // row a in src
    const long long oa0 = (long long)rhs * a.ldd + (long long)ra * N;  // row a in dst
    const bool s64 = sizeof(T) == 8 || a.src64, d64 = sizeof(T) == 8 || a.dst64;
    auto ld_src = [&](long long q) -> T {
        return s64 ? (T)static_cast<const double*>(a.src)[q] : static_cast<const T*>(a.src)[q];
    };
    auto st_dst = [&](long long q, T v) {
        if (d64) static_cast<double*>(a.dst)[q] = (double)v;
        else static_cast<T*>(a.dst)[q] = v;
    };

    // ---- load: build the packed extended sequence in A ----
    // The rows are read in chunks of VPT values per thread with every global
    // load of a chunk issued before the first LDS write, so the HBM latency
    // is paid once per chunk instead of once per element.
    const int kind = a.kind, inv = a.inverse;
    if (kind == T_DD) {
        if (threadIdx.x == 0) A[0] = {T(0), T(0)};
        if (threadIdx.x == 1) A[N + 1] = {T(0), T(0)};
    } else if (kind == T_NN && inv) {
        if (threadIdx.x == 0) A[N] = {T(0), T(0)};
    } else if (kind != T_NN) {
        if (threadIdx.x == 0) A[0] = {T(0), T(0)};
        for (int p = N + 1 + threadIdx.x; p < L; p += NT) A[p] = {T(0), T(0)};
    }
    for (int base = 0; base < N; base += VPT * NT) {
        T va[VPT], vb[VPT];
#pragma unroll
        for (int k = 0; k < VPT; k++) {
            const int n = base + k * NT + threadIdx.x;
            va[k]       = n < N ? ld_src(ia + n) : T(0);
            vb[k]       = (n < N && hb) ? ld_src(ia + N + n) : T(0);
        }
#pragma unroll
        for (int k = 0; k < VPT; k++) {
            const int n = base + k * NT + threadIdx.x;
            if (n >= N) break;
            const T x = va[k], y = vb[k];
            if (kind == T_DD) {
                A[n + 1]     = {x, y};
                A[L - 1 - n] = {-x, -y};
            } else if (kind == T_NN && !inv) {
                A[n]             = {x, y};
                A[2 * N - 1 - n] = {x, y};
            } else if (kind == T_NN) {
                const int   m  = N - 1 - n;
                const T     s  = m == 0 ? a.c1 : a.c0;  // sqrt(1/N) or -sqrt(2/N)
                const T     ya = s * x, yb = s * y;
                const C2<T> e  = a.aux[m];              // exp(-i pi m / 2N); Y_m uses conj(e)
                A[m]           = {ya * e.x + yb * e.y, -ya * e.y + yb * e.x};
                if (m > 0) A[2 * N - m] = {ya * e.x - yb * e.y, ya * e.y + yb * e.x};
                else {
                    y0[0] = ya;
                    y0[1] = yb;
                }
            } else if (kind == T_ND && inv) {
                A[n + 1] = {x, y};
            } else {  // DN both ways, ND forward
                const T s = sgn_pow<T>(n + 1) * a.c0;
                A[n + 1]  = {s * x, s * y};
            }
        }
    }
    __syncthreads();

    const bool   fwd = !(a.kind == T_NN && a.inverse);
    const C2<T>* Z   = fwd ? run_fft<P1, true, NT>(a, W, A, B) : run_fft<P1, false, NT>(a, W, A, B);

    // ---- store: split the two rows' coefficients ----
    for (int base = 0; base < N; base += VPT * NT) {
        C2<T> sc[VPT];
        if (kind == T_NN && !inv) {
#pragma unroll
            for (int k = 0; k < VPT; k++) {
                const int j = base + k * NT + threadIdx.x;
                sc[k]       = j < N ? a.aux[N - 1 - j] : C2<T>{T(0), T(0)};
            }
        } else if (kind == T_ND && !inv) {
#pragma unroll
            for (int k = 0; k < VPT; k++) {
                const int j = base + k * NT + threadIdx.x;
                sc[k]       = j < N ? a.aux[j] : C2<T>{T(0), T(0)};
            }
        }
#pragma unroll
        for (int k = 0; k < VPT; k++) {
            const int j = base + k * NT + threadIdx.x;
            if (j >= N) break;
            T oa, ob;
            if (kind == T_DD) {
                const C2<T> z = Z[j + 1];
                oa            = -a.c0 * z.y;
                ob            = a.c0 * z.x;
            } else if (kind == T_NN && !inv) {
                const int   m = N - 1 - j;
                const C2<T> w = cmul(Z[m], sc[k]);  // 2 C_a(m) + 2i C_b(m)
                const T     s = T(0.5) * ((j == N - 1) ? a.c1 : a.c0);
                oa            = s * w.x;
                ob            = s * w.y;
            } else if (kind == T_NN) {
                const C2<T> z = Z[j];
                oa            = T(0.5) * (z.x + y0[0]);
                ob            = T(0.5) * (z.y + y0[1]);
            } else if (kind == T_DN) {
                const C2<T> zm = Z[j + 1], zc = Z[L - 1 - j];
                const T     s  = sgn_pow<T>(j) * T(0.5);
                oa             = s * (zm.y - zc.y);
                ob             = s * (zc.x - zm.x);
            } else if (!inv) {  // ND forward
                const C2<T> zm = Z[j + 1], zc = Z[L - 1 - j];
                const T     rea = T(0.5) * (zm.x + zc.x), ima = T(0.5) * (zm.y - zc.y);
                const T     reb = T(0.5) * (zm.y + zc.y), imb = T(0.5) * (zc.x - zm.x);
                oa              = sc[k].x * rea - sc[k].y * ima;
                ob              = sc[k].x * reb - sc[k].y * imb;
            } else {  // ND inverse
                const int   kk = N - j;
                const C2<T> zk = Z[kk], zc = Z[L - kk];
                const T     s  = sgn_pow<T>(j + 1) * a.c0 * T(0.5);
                oa             = s * (zk.y - zc.y);
                ob             = s * (zc.x - zk.x);
            }
            st_dst(oa0 + j, oa);
            if (hb) st_dst(oa0 + N + j, ob);
        }
    }
}

// ---------------------------------------------------------------------------
// host side

// Radix schedule: the odd primes above 5 first (their passes then run with
// Ns = 1, i.e. without twiddle factors), then 8, 4, 2, 3, 5.
static void factor_radices(int L, int* rad, int& nrad)
{
    nrad  = 0;
    int n = L;
    {
        int m = L, big[kMaxRadices], nb = 0;
        for (int p : {2, 3, 5})
            while (m % p == 0) m /= p;
        for (int p = 7; m > 1; p += 2)
            while (m % p == 0) {
                if (nb >= kMaxRadices) throw Error(FP_E_INVALID, "transform length has too many factors");
                big[nb++] = p;
                m /= p;
            }
        for (int i = nb - 1; i >= 0; i--) {
            rad[nrad++] = big[i];
            n /= big[i];
        }
    }
    while (n % 8 == 0) { rad[nrad++] = 8; n /= 8; }
    while (n % 4 == 0) { rad[nrad++] = 4; n /= 4; }
    while (n % 2 == 0) { rad[nrad++] = 2; n /= 2; }
    while (n % 3 == 0) { rad[nrad++] = 3; n /= 3; }
    while (n % 5 == 0) { rad[nrad++] = 5; n /= 5; }
    if (n != 1) throw Error(FP_E_INVALID, "radix schedule does not cover the transform length");
}

template <typename T>
void TransformPlan<T>::build(int k, int n)
{
    kind = k;
    N    = n;
    if (kind == T_DD) L = 2 * N + 2;
    else if (kind == T_NN) L = 2 * N;
    else L = 2 * N + 1;
    factor_radices(L, rad, nrad);
    p1 = 0;
    for (int i = 0; i < nrad; i++) {
        const int r = rad[i];
        if ((r == 11 || r == 13 || r == 17 || r == 19 || r == 23 || r == 29 || r == 31 || r == 37 || r == 41 ||
             r == 43 || r == 47) && r > p1)
            p1 = r;
    }
    const long double PI  = 3.141592653589793238462643383279502884L;
    const int         nhi = L / 64 + 1;
    std::vector<T>    w(2 * (size_t)(64 + nhi)), x(2 * (size_t)N, T(0));
    for (int q = 0; q < 64 + nhi; q++) {
        const long long   t   = q < 64 ? q : 64LL * (q - 64);
        const long double ang = -2.0L * PI * (long double)(t % L) / (long double)L;
        w[2 * q]              = (T)cosl(ang);
        w[2 * q + 1]          = (T)sinl(ang);
    }
    if (kind == T_NN) {
        for (int m = 0; m < N; m++) {
            const long double ang = -PI * (long double)m / (2.0L * N);
            x[2 * m]              = (T)cosl(ang);
            x[2 * m + 1]          = (T)sinl(ang);
        }
        c0 = (T)(-std::sqrt(2.0 / N));
        c1 = (T)std::sqrt(1.0 / N);
    } else if (kind == T_ND) {
        for (int j = 0; j < N; j++) {
            // theta_j = 2 N pi (j+1) / (2N+1), reduced modulo 2 pi on integers
            const long long   q   = (2LL * N * (j + 1)) % (2LL * (2 * N + 1));
            const long double ang = PI * (long double)q / (long double)(2 * N + 1);
            x[2 * j]              = (T)sinl(ang);
            x[2 * j + 1]          = (T)cosl(ang);
        }
        c0 = (T)std::sqrt(4.0 / (2.0 * N + 1.0));
    } else if (kind == T_DN) {
        c0 = (T)std::sqrt(4.0 / (2.0 * N + 1.0));
    } else {
        c0 = (T)(std::sqrt(2.0 / (N + 1.0)) * 0.5);
    }
    W.alloc(w.size());
    W.upload(w.data(), w.size());
    aux.alloc(2 * (size_t)N);
    aux.upload(x.data(), x.size());
}

template struct TransformPlan<double>;
template struct TransformPlan<float>;

template <typename T>
static YArgs<T> yargs(const TransformJob<T>& j, int inverse)
{
    const TransformPlan<T>& tp = *j.plan;
    YArgs<T>                a{};
    a.src     = j.src ? j.src : j.data;
    a.dst     = j.data;
    a.lds     = j.src ? j.lds : j.ld;
    a.ldd     = j.ld;
    a.src64   = j.src ? j.src64 : (sizeof(T) == 8);
    a.dst64   = j.dst64;
    a.nx      = j.nx;
    a.N       = tp.N;
    a.L       = tp.L;
    a.kind    = tp.kind;
    a.inverse = inverse;
    a.nrad    = tp.nrad;
    for (int i = 0; i < tp.nrad; i++) a.rad[i] = tp.rad[i];
    a.W   = reinterpret_cast<const C2<T>*>(tp.W.p);
    a.aux = reinterpret_cast<const C2<T>*>(tp.aux.p);
    a.c0  = tp.c0;
    a.c1  = tp.c1;
    return a;
}

template <typename T>
static size_t ytransform_lds(int L)
{
    return (2 * (size_t)L + 64 + L / 64 + 1) * sizeof(C2<T>) + 4 * sizeof(T);
}

// one launch of up to kMaxBatch jobs with the register codelet P1
template <typename T>
static void launch_group(const TransformJob<T>* const* jobs, int n, int P1, int inverse, hipStream_t s)
{
    constexpr int NT = 256;
    YBatch<T>     bt{};
    size_t        lds = 0;
    int           blocks = 0;
    bt.nprob           = n;
    for (int q = 0; q < n; q++) {
        const TransformJob<T>& j = *jobs[q];
        bt.p[q]                  = yargs(j, inverse);
        bt.start[q]              = blocks;
        const long long nb       = (long long)((j.nx + 1) / 2) * j.nrhs;
        if (blocks + nb > (1LL << 30)) throw Error(FP_E_INVALID, "too many transform workgroups in one launch");
        blocks += (int)nb;
        lds = std::max(lds, ytransform_lds<T>(j.plan->L));
    }
    bt.start[n] = blocks;
    if (lds > 160 * 1024) throw Error(FP_E_INVALID, "transform length exceeds the LDS of one workgroup");
    if (blocks == 0) return;
    switch (P1) {
#define FP_P1_CASE(P)                                                                                            \
    case P: {                                                                                                    \
        static bool attr_set = false;                                                                            \
        if (lds > 64 * 1024 && !attr_set) {                                                                      \
            FP_HIP(hipFuncSetAttribute((const void*)k_ytransform<NT, P, T>,                                      \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));                 \
            attr_set = true;                                                                                     \
        }                                                                                                        \
        hipLaunchKernelGGL((k_ytransform<NT, P, T>), dim3(blocks), dim3(NT), lds, s, bt);                        \
        break;                                                                                                   \
    }
        FP_P1_CASE(0)
        FP_P1_CASE(11)
        FP_P1_CASE(13)
        FP_P1_CASE(17)
        FP_P1_CASE(19)
        FP_P1_CASE(23)
        FP_P1_CASE(29)
        FP_P1_CASE(31)
        FP_P1_CASE(37)
        FP_P1_CASE(41)
        FP_P1_CASE(43)
        FP_P1_CASE(47)
#undef FP_P1_CASE
        default: throw Error(FP_E_INVALID, "no transform kernel for this radix schedule");
    }
    FP_HIP(hipGetLastError());
}

// Jobs are grouped by their large-prime codelet (jobs without one join the
// first group) and launched kMaxBatch at a time.
template <typename T>
void launch_ytransform_batch(const TransformJob<T>* jobs, int njobs, int inverse, hipStream_t s)
{
    std::vector<int>                           keys;
    std::vector<std::vector<const TransformJob<T>*>> groups;
    std::vector<const TransformJob<T>*>        plain;
    for (int k = 0; k < njobs; k++) {
        if (jobs[k].nrhs <= 0 || jobs[k].nx <= 0) continue;
        const int p1 = jobs[k].plan->p1;
        if (p1 == 0) {
            plain.push_back(&jobs[k]);
            continue;
        }
        size_t g = 0;
        while (g < keys.size() && keys[g] != p1) g++;
        if (g == keys.size()) {
            keys.push_back(p1);
            groups.emplace_back();
        }
        groups[g].push_back(&jobs[k]);
    }
    if (groups.empty()) {
        keys.push_back(0);
        groups.emplace_back();
    }
    groups[0].insert(groups[0].end(), plain.begin(), plain.end());
    for (size_t g = 0; g < groups.size(); g++)
        for (size_t q0 = 0; q0 < groups[g].size(); q0 += kMaxBatch)
            launch_group<T>(groups[g].data() + q0, (int)std::min<size_t>(kMaxBatch, groups[g].size() - q0), keys[g],
                            inverse, s);
}

template <typename T>
void launch_ytransform(const TransformPlan<T>& tp, T* data, long long ld, int nx, int nrhs, int inverse,
                       hipStream_t s)
{
    TransformJob<T> j{};
    j.plan  = &tp;
    j.data  = data;
    j.ld    = ld;
    j.nx    = nx;
    j.nrhs  = nrhs;
    j.dst64 = sizeof(T) == 8;
    launch_ytransform_batch<T>(&j, 1, inverse, s);
}

template void launch_ytransform_batch<double>(const TransformJob<double>*, int, int, hipStream_t);
template void launch_ytransform_batch<float>(const TransformJob<float>*, int, int, hipStream_t);
template void launch_ytransform<double>(const TransformPlan<double>&, double*, long long, int, int, int, hipStream_t);
template void launch_ytransform<float>(const TransformPlan<float>&, float*, long long, int, int, int, hipStream_t);

}  // namespace fp
