// Batched y transforms of the subdomain solver (DST-I, DCT-II/III and the
// two quarter-wave pairs), as packed complex FFTs staged in LDS.
//
// Reference semantics: PoissonFFT2D_{DD,NN,DN,ND}::transform and
// ::transform_transpose (pe/poisson/poisson_fft2d.cpp:37-57, 105-133,
// 180-205, 242-262).  The reference runs one FFTW r2c per row (two for the
// NN/ND inverse); here two rows a, b of the field form one complex sequence
// z = ext(a) + i ext(b) of length L, one FFT of length L is done in LDS, and
// the two rows' coefficients are split out of the spectrum:
//   DD     : L = N+1 (the half-length DST-I): with x_0 = 0, x_j = f(j-1),
//            y_j = sin(pi j/L)(x_j + x_{L-j}) + (x_j - x_{L-j})/2 is FFT'd;
//            even outputs X_2k = -Im F_k, odd outputs X_2k+1 = F_0/2 +
//            Re(F_1 + .. + F_k) by a block prefix sum in LDS
//   NN fwd : half-sample even extension, L = 2N, phase exp(-i pi m / 2N)
//   NN inv : Hermitian extension of the phased coefficients, inverse FFT, L = 2N
//   DN, ND : L = 2N+1, split by Z(m) +- conj Z(L-m)
// Each output equals the reference's coefficient (same signs and mode order),
// so transform results are directly comparable with the reference's.
//
// One workgroup owns one row pair of one right-hand side: it loads both rows
// (coalesced), runs the FFT in LDS and writes the coefficients back in place.
// Three FFT engines share those load/store phases:
//
//   k_ytransform_sb  (every length whose prime factors are at most 47): a
//     mixed-radix Stockham FFT on ONE LDS buffer of L values.  Each pass reads
//     all of its inputs into registers (radix 2/3/4/5/8 butterflies, or the
//     accumulators of an odd-prime butterfly), the workgroup synchronises, and
//     the outputs overwrite the buffer.  Half the LDS of a two-buffer FFT, so
//     twice the workgroups per CU.  Twiddles: one two-level table lookup per
//     butterfly, the other powers by multiplication.  Odd primes 7..47 run a
//     symmetric-pair DFT split over groups of threads (a group is a whole
//     wavefront, so no divergence), its cosines broadcast from an LDS table.
//   k_ytransform_gen (fallback for lengths whose passes do not fit the
//     per-thread register budget): two LDS buffers, any prime through an
//     R-term in-LDS sum per output.
//   k_ytransform_glob (lengths whose buffers exceed one workgroup's LDS,
//     e.g. the FP64 quarter-wave kinds at N = 4096, L = 8193 = 3 x 2731):
//     the two-buffer engine's passes on buffers in device memory, Bluestein
//     for prime factors above 47, a bounded grid looping over row pairs.
#include "fp_internal.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <complex>
#include <cstring>
#include <map>
#include <mutex>
#include <set>

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
    int          F;  // FFT length: L, or the smooth Bluestein length >= 2L-1
    int          rad[kMaxRadices];  // radices of F
    const C2<T>* chirp;  // Bluestein: exp(-i pi n^2 / L), n < L
    const C2<T>* bh;     // Bluestein: DFT_F of the conjugate chirp, over F
    C2<T>*       gs;     // global-memory engine: two buffers of F values per workgroup slot
    const C2<T>* W;    // 64 + L/64 + 1 twiddles, then cos/sin tables of the odd-prime passes
    int          ncs;  // entries of those tables (staged in LDS; FP32 odd-prime passes)
    const double* frag;  // FP64 odd-prime passes: matrix-core operand fragments of the DFT blocks
    int           nfrag; // their count (doubles), staged in LDS
    const C2<T>* aux;
    T            c0, c1;
    // fused boundary assembly and scaling of a raw right-hand side on load
    int           asmb;
    T             hx2;
    const double* g[4];
    T             ga[4];
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

template <typename T>
__device__ __forceinline__ C2<T> troot(const Tw<T> W, int t)
{
    return cmul(W.hi[t >> 6], W.lo[t & 63]);
}

template <bool FWD, typename T>
__device__ __forceinline__ C2<T> twiddle(const Tw<T> W, int t, C2<T> v)
{
    const C2<T> w = troot(W, t);
    return FWD ? cmul(v, w) : cmulc(v, w);
}

// Radix-R butterfly in registers, R in {2, 3, 4, 5, 8}.
template <int R, bool FWD, typename T>
__device__ __forceinline__ void bfly(const C2<T> (&v)[R], C2<T> (&o)[R])
{
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
}

// ---------------------------------------------------------------------------
// single-buffer engine

// LDS position of sequence element i: one padding slot after every 2^kZpad
// elements spreads the strided butterfly writes of the early passes over the
// LDS banks.
#ifndef FP_ZPAD
#define FP_ZPAD 4
#endif
constexpr int kZpad = FP_ZPAD;
// Form of the odd-prime (7..47) passes, a kernel template parameter chosen
// per plan: none in the radix schedule, matrix cores (FP64), VALU pair sums.
// ODD_VALU / ODD_VALU3: VALU pair sums, kNQ / 3 output pairs per work item
enum OddForm : int { ODD_NONE = 0, ODD_MFMA = 1, ODD_VALU = 2, ODD_VALU3 = 3 };
template <int PAD>
__device__ __forceinline__ int zp(int i)
{
    if constexpr (PAD == 0) return i;
    else return i + (i >> PAD);
}

// butterflies of radix R one thread holds in registers in one pass
// (16 complex values at most)
constexpr int kEPT = 16;
// long FFTs (Bluestein lengths above 16 x 512) hold 24 values per thread
constexpr int kEPTLong = 24;
template <int R, int EPT = kEPT>
constexpr int sb_bpt()
{
    return (EPT + R - 1) / R;
}
// odd-prime butterflies: output pairs per work item (VALU form; the VALU3
// form takes 3, more items of shorter latency), items per thread
constexpr int kNQ  = 5;
constexpr int kIPT = 1;
constexpr int kMaxOddPrime = 47;

// The pass functions below are inlined into one loop over the passes; their
// index arithmetic depends only on (L, Ns, R), which do not change inside
// the loop, so the compiler would hoist every pass's addresses out of it and
// keep them all live (about 130 extra VGPRs).  An empty asm that "redefines"
// the lengths at the top of each pass keeps that arithmetic inside the pass.
__device__ __forceinline__ void pass_lengths(int& L, int& Ns)
{
    L  = __builtin_amdgcn_readfirstlane(L);
    Ns = __builtin_amdgcn_readfirstlane(Ns);
    asm volatile("" : "+s"(L), "+s"(Ns));
}
__device__ __forceinline__ void pass_lengths(int& L, int& Ns, int& R)
{
    R = __builtin_amdgcn_readfirstlane(R);
    asm volatile("" : "+s"(R));
    pass_lengths(L, Ns);
}

// One Stockham pass of radix R on the single buffer Z: read every input,
// synchronise, twiddle + butterfly, write, synchronise.  Twiddles of a
// butterfly are w^r, r = 1..R-1, with w = W^(k * stride) from the table and
// its powers by repeated products (depth <= 3 multiplications).
template <int R, bool FWD, int EPT, typename T>
__device__ __forceinline__ void pass_reg_sb(C2<T>* __restrict__ Z, const Tw<T> W, int L, int Ns)
{
    pass_lengths(L, Ns);
    constexpr int BPT = sb_bpt<R, EPT>();
    const int     NT  = blockDim.x;
    const int     nb = L / R, stride = L / (Ns * R);
    C2<T>         v[BPT][R];
#pragma unroll
    for (int b = 0; b < BPT; b++) {
        const int jb = threadIdx.x + b * NT;
        if (jb < nb) {
#pragma unroll
            for (int r = 0; r < R; r++) v[b][r] = Z[zp<kZpad>(jb + r * nb)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < BPT; b++) {
        const int jb = threadIdx.x + b * NT;
        if (jb < nb) {
            const int k = jb % Ns;
            if (Ns > 1) {
                C2<T> pw[R];
                pw[1] = troot(W, k * stride);
#pragma unroll
                for (int r = 2; r < R; r++) pw[r] = (r & 1) ? cmul(pw[r - 1], pw[1]) : cmul(pw[r / 2], pw[r / 2]);
#pragma unroll
                for (int r = 1; r < R; r++) v[b][r] = FWD ? cmul(v[b][r], pw[r]) : cmulc(v[b][r], pw[r]);
            }
            C2<T> o[R];
            bfly<R, FWD>(v[b], o);
            const int idxD = (jb / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) Z[zp<kZpad>(idxD + r * Ns)] = o[r];
        }
    }
    __syncthreads();
}

// Lanes given to the butterflies of one output group in the odd-prime pass:
// whole wavefronts for 33 or more butterflies, else 32 or 16 lanes, so that
// two or four groups share a wavefront (each cosine read then serves two or
// four groups instead of one).
__host__ __device__ __forceinline__ int odd_slots(int nb)
{
    return nb <= 16 ? 16 : nb <= 32 ? 32 : (nb + 63) & ~63;
}

// One Stockham pass of an odd prime radix R (7..47) on the single buffer.
// With a_r = x_r + x_{R-r}, b_r = x_r - x_{R-r} (twiddled inputs):
//   X_q = x_0 + A_q -+ i B_q,  X_{R-q} = x_0 + A_q +- i B_q,
//   A_q = sum_r a_r cos(2 pi q r / R),  B_q = sum_r b_r sin(2 pi q r / R).
// The H = (R-1)/2 output pairs of a butterfly are split over G groups of kNQ
// pairs; work item t = (group, butterfly) with the butterflies padded to a
// multiple of 64 (or to 16 / 32 lanes for few butterflies, odd_slots), so a
// wavefront holds one to four groups.  The cosines come from the R values
// (cos, sin)(2 pi m / R) staged in LDS with the twiddles, at m = q r mod R
// kept by an addition and an unsigned minimum per read (one to four
// addresses per wavefront).  The
// item a thread owns keeps its accumulators in registers across the
// synchronisation.
template <bool FWD, int NQ, typename T>
__device__ __forceinline__ void pass_odd_sb(C2<T>* __restrict__ Z, const Tw<T> W, int L, int Ns, int R,
                                            const C2<T>* __restrict__ csg)
{
    pass_lengths(L, Ns, R);
    const int NT = blockDim.x;
    const int H = (R - 1) >> 1, G = (H + NQ - 1) / NQ;
    const int nb = L / R, stride = L / (Ns * R), nbp = odd_slots(nb);
    T         ax[kIPT][NQ], ay[kIPT][NQ], bx[kIPT][NQ], by[kIPT][NQ];
    C2<T>     x0[kIPT];
#pragma unroll
    for (int it = 0; it < kIPT; it++) {
        const int t   = threadIdx.x + it * NT;
        const int grp = nbp >= 64 ? __builtin_amdgcn_readfirstlane(t / nbp) : t / nbp;
        const int jb  = t - grp * nbp;
        if (grp < G && jb < nb) {
            const int   QB = 1 + grp * NQ;
            const int   k  = jb % Ns;
            const C2<T> v0 = Z[zp<kZpad>(jb)];
            x0[it]         = v0;
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                ax[it][q] = v0.x;
                ay[it][q] = v0.y;
                bx[it][q] = T(0);
                by[it][q] = T(0);
            }
            unsigned m[NQ];  // (QB + q) r mod R
#pragma unroll
            for (int q = 0; q < NQ; q++) m[q] = 0;
            for (int r = 1; r <= H; r++) {
                C2<T> x1 = Z[zp<kZpad>(jb + r * nb)], x2 = Z[zp<kZpad>(jb + (R - r) * nb)];
                if (Ns > 1) {
                    x1 = twiddle<FWD>(W, r * k * stride, x1);
                    x2 = twiddle<FWD>(W, (R - r) * k * stride, x2);
                }
                const C2<T> a = cadd(x1, x2), b = csub(x1, x2);
                x0[it] = cadd(x0[it], a);
#pragma unroll
                for (int q = 0; q < NQ; q++) {
                    // m + QB + q < 2R: the unsigned difference wraps to a
                    // large value unless m + QB + q >= R
                    const unsigned t = m[q] + (unsigned)(QB + q);
                    m[q]             = min(t, t - (unsigned)R);
                    const C2<T> w    = csg[m[q]];
                    ax[it][q] += a.x * w.x;
                    ay[it][q] += a.y * w.x;
                    bx[it][q] += b.x * w.y;
                    by[it][q] += b.y * w.y;
                }
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kIPT; it++) {
        const int t = threadIdx.x + it * NT;
        const int grp = t / nbp, jb = t - grp * nbp;
        if (grp < G && jb < nb) {
            const int QB   = 1 + grp * NQ;
            const int k    = jb % Ns;
            const int base = (jb / Ns) * Ns * R + k;
            if (grp == 0) Z[zp<kZpad>(base)] = x0[it];
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                if (QB + q > H) break;
                const C2<T> xq  = FWD ? C2<T>{ax[it][q] + by[it][q], ay[it][q] - bx[it][q]}
                                      : C2<T>{ax[it][q] - by[it][q], ay[it][q] + bx[it][q]};
                const C2<T> xrq = FWD ? C2<T>{ax[it][q] - by[it][q], ay[it][q] + bx[it][q]}
                                      : C2<T>{ax[it][q] + by[it][q], ay[it][q] - bx[it][q]};
                Z[zp<kZpad>(base + (QB + q) * Ns)]     = xq;
                Z[zp<kZpad>(base + (R - QB - q) * Ns)] = xrq;
            }
        }
    }
    __syncthreads();
}

// FP64 odd-prime pass on the matrix cores.  With the same a_r, b_r as
// pass_odd_sb, the H + 1 sums per butterfly and component are two real
// matrix products,
//   [S0; A_1 .. A_H] = Cm [a_1 .. a_H]^T,  [B_1 .. B_H] = Sm [b_1 .. b_H]^T,
// Cm(q, r) = cos(2 pi q r / R) (row 0 all ones: S0 = sum_r a_r), Sm = sin,
// whose columns are the butterflies' real and imaginary parts (column
// n = 2 jb + component).  v_mfma_f64_16x16x4f64 tiles them: M = q in tiles of
// 16 (two tiles cover R <= 61), K = r in steps of 4, N = columns in tiles of
// 16.  A wavefront owns whole column tiles: per K step each lane loads the
// x_r, x_{R-r} of its (r, column) (the B operand) and the wavefront issues
// 2 x (M tiles) products against Cm/Sm fragments staged in LDS at kernel
// start (the A operand, one double per lane).  The outputs pair the real and imaginary
// columns of a butterfly (neighbouring lanes) to form X_q and X_{R-q}.
// Operand layouts of the instruction (gfx950): A(i, k) in lane i + 16 k,
// B(k, j) in lane j + 16 k, D(i, j) in lane j + 16 (i % 4), element i / 4.
typedef double d4_t __attribute__((ext_vector_type(4)));
constexpr int kMaxNTw = 2;  // column tiles per wavefront

__host__ __device__ __forceinline__ int mfma_ksteps(int R) { return ((R - 1) / 2 + 3) / 4; }

template <bool FWD>
__device__ __forceinline__ void pass_odd_mfma(C2<double>* __restrict__ Z, const Tw<double> W, int L, int Ns, int R,
                                              const double* __restrict__ frag)
{
    pass_lengths(L, Ns, R);
    const int H = (R - 1) >> 1, MT = (H + 16) >> 4, KT = mfma_ksteps(R);
    const int nb = L / R, stride = L / (Ns * R);
    const int NTL  = (2 * nb + 15) >> 4;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int kq = lane >> 4, jn = lane & 15;
    d4_t       Dc[kMaxNTw][2], Ds[kMaxNTw][2];
    C2<double> x0v[kMaxNTw];
#pragma unroll
    for (int tt = 0; tt < kMaxNTw; tt++) {
        const int nt = wave + tt * nw;
        if (nt >= NTL) break;  // uniform over the wavefront
        const int  n = nt * 16 + jn, jb = n >> 1, comp = n & 1;
        const bool colok = jb < nb;
        const int  k     = colok ? jb % Ns : 0;
        x0v[tt]          = colok ? Z[zp<kZpad>(jb)] : C2<double>{0.0, 0.0};
#pragma unroll
        for (int mt = 0; mt < 2; mt++) {
            Dc[tt][mt] = d4_t{0.0, 0.0, 0.0, 0.0};
            Ds[tt][mt] = d4_t{0.0, 0.0, 0.0, 0.0};
        }
        for (int ks = 0; ks < KT; ks++) {
            const int r  = 4 * ks + kq + 1;
            double    av = 0.0, bv = 0.0;
            if (colok && r <= H) {
                C2<double> x1 = Z[zp<kZpad>(jb + r * nb)], x2 = Z[zp<kZpad>(jb + (R - r) * nb)];
                if (Ns > 1) {
                    x1 = twiddle<FWD>(W, r * k * stride, x1);
                    x2 = twiddle<FWD>(W, (R - r) * k * stride, x2);
                }
                av = comp ? x1.y + x2.y : x1.x + x2.x;
                bv = comp ? x1.y - x2.y : x1.x - x2.x;
            }
#pragma unroll
            for (int mt = 0; mt < 2; mt++) {
                if (mt >= MT) break;
                const double ca = frag[((0 * 2 + mt) * KT + ks) * 64 + lane];
                const double sa = frag[((1 * 2 + mt) * KT + ks) * 64 + lane];
                Dc[tt][mt]      = __builtin_amdgcn_mfma_f64_16x16x4f64(ca, av, Dc[tt][mt], 0, 0, 0);
                Ds[tt][mt]      = __builtin_amdgcn_mfma_f64_16x16x4f64(sa, bv, Ds[tt][mt], 0, 0, 0);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int tt = 0; tt < kMaxNTw; tt++) {
        const int nt = wave + tt * nw;
        if (nt >= NTL) break;
        const int  n = nt * 16 + jn, jb = n >> 1, comp = n & 1;
        const bool colok = jb < nb;
        const int  k     = colok ? jb % Ns : 0;
        const int  base  = colok ? (jb / Ns) * Ns * R + k : 0;
        const C2<double> x0 = x0v[tt];
#pragma unroll
        for (int mt = 0; mt < 2; mt++) {
            if (mt >= MT) break;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const double c = Dc[tt][mt][v], sn = Ds[tt][mt][v];
                const double co = __shfl_xor(c, 1, 64), so = __shfl_xor(sn, 1, 64);
                const int    q  = 16 * mt + 4 * v + kq;
                if (!colok || q > H) continue;
                // (re, im) of A_q and B_q from this lane and its neighbour
                const double Ar = comp ? co : c, Ai = comp ? c : co;
                const double Br = comp ? so : sn, Bi = comp ? sn : so;
                if (q == 0) {
                    if (comp == 0) Z[zp<kZpad>(base)] = C2<double>{x0.x + Ar, x0.y + Ai};
                } else if (comp == 0) {
                    Z[zp<kZpad>(base + q * Ns)] = FWD ? C2<double>{x0.x + Ar + Bi, x0.y + Ai - Br}
                                                      : C2<double>{x0.x + Ar - Bi, x0.y + Ai + Br};
                } else {
                    Z[zp<kZpad>(base + (R - q) * Ns)] = FWD ? C2<double>{x0.x + Ar - Bi, x0.y + Ai + Br}
                                                            : C2<double>{x0.x + Ar + Bi, x0.y + Ai - Br};
                }
            }
        }
    }
    __syncthreads();
}

template <bool FWD, int OM, int EPT, typename T>
__device__ __forceinline__ void run_fft_sb(const YArgs<T>& a, const Tw<T> W, C2<T>* Z, const C2<T>* cs,
                                           const double* fragL)
{
    int Ns = 1, co = 0, fo = 0;
    for (int s = 0; s < a.nrad; s++) {
        const int R = a.rad[s];
        if (R == 2) pass_reg_sb<2, FWD, EPT>(Z, W, a.F, Ns);
        else if (R == 3) pass_reg_sb<3, FWD, EPT>(Z, W, a.F, Ns);
        else if (R == 4) pass_reg_sb<4, FWD, EPT>(Z, W, a.F, Ns);
        else if (R == 5) pass_reg_sb<5, FWD, EPT>(Z, W, a.F, Ns);
        else if (R == 8) pass_reg_sb<8, FWD, EPT>(Z, W, a.F, Ns);
        else if constexpr (OM == ODD_MFMA && sizeof(T) == 8) {
            pass_odd_mfma<FWD>(Z, W, a.F, Ns, R, fragL + fo);
            fo += 2 * 2 * mfma_ksteps(R) * 64;
        } else if constexpr (OM == ODD_VALU || OM == ODD_VALU3) {
            pass_odd_sb<FWD, OM == ODD_VALU ? kNQ : 3>(Z, W, a.F, Ns, R, cs + co);
            co += R;
        }
        Ns *= R;
    }
}

// DFT of length L (sign by FWD) of Z[0..L) through two FFTs of the smooth
// length F >= 2L - 1 (Bluestein): with c_n = exp(-i pi n^2 / L),
//   X_k = c_k sum_n (z_n c_n) conj(c_{k-n}),
// a circular convolution of length F with the conjugate chirp, whose DFT bh
// (scaled by 1/F) is precomputed.  The + sign conjugates input and output.
// Every thread reads and writes only its own elements between barriers.
template <bool FWD, int OM, int EPT, typename T>
__device__ __forceinline__ void run_bluestein(const YArgs<T>& a, const Tw<T> W, C2<T>* Z, const C2<T>* cs,
                                              const double* fragL)
{
    const int L = a.L, F = a.F, NT = blockDim.x;
    for (int n = threadIdx.x; n < F; n += NT) {
        C2<T> v{T(0), T(0)};
        if (n < L) {
            v = Z[zp<kZpad>(n)];
            if (!FWD) v.y = -v.y;
            v = cmul(v, a.chirp[n]);
        }
        Z[zp<kZpad>(n)] = v;
    }
    __syncthreads();
    run_fft_sb<true, OM, EPT>(a, W, Z, cs, fragL);
    for (int k = threadIdx.x; k < F; k += NT) Z[zp<kZpad>(k)] = cmul(Z[zp<kZpad>(k)], a.bh[k]);
    __syncthreads();
    run_fft_sb<false, OM, EPT>(a, W, Z, cs, fragL);
    for (int k = threadIdx.x; k < L; k += NT) {
        C2<T> v = cmul(Z[zp<kZpad>(k)], a.chirp[k]);
        if (!FWD) v.y = -v.y;
        Z[zp<kZpad>(k)] = v;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// two-buffer fallback engine

// One Stockham pass with a register butterfly of radix R.
template <int R, bool FWD, typename T>
__device__ void pass_reg(const C2<T>* __restrict__ src, C2<T>* __restrict__ dst, const Tw<T> W, int L, int Ns)
{
    const int NT = blockDim.x;
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
        bfly<R, FWD>(v, o);
        const int idxD = (jb / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) dst[idxD + r * Ns] = o[r];
    }
}

// Stockham pass of any radix R: twiddle in place, then each output is an
// R-term sum read from LDS.
template <bool FWD, typename T>
__device__ void pass_generic(C2<T>* __restrict__ src, C2<T>* __restrict__ dst, const Tw<T> W, int L, int Ns,
                             int R)
{
    const int NT = blockDim.x;
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

template <bool FWD, typename T>
__device__ C2<T>* run_fft_gen(const YArgs<T>& a, const Tw<T> W, C2<T>* A, C2<T>* B)
{
    // FFT of length F (= L for the two-buffer engine; the Bluestein length
    // in the global-memory engine)
    C2<T>* src = A;
    C2<T>* dst = B;
    int    Ns  = 1;
    for (int s = 0; s < a.nrad; s++) {
        const int R = a.rad[s];
        if (R == 2) pass_reg<2, FWD>(src, dst, W, a.F, Ns);
        else if (R == 3) pass_reg<3, FWD>(src, dst, W, a.F, Ns);
        else if (R == 4) pass_reg<4, FWD>(src, dst, W, a.F, Ns);
        else if (R == 5) pass_reg<5, FWD>(src, dst, W, a.F, Ns);
        else if (R == 8) pass_reg<8, FWD>(src, dst, W, a.F, Ns);
        else pass_generic<FWD>(src, dst, W, a.F, Ns, R);
        __syncthreads();
        C2<T>* t = src;
        src      = dst;
        dst      = t;
        Ns *= R;
    }
    return src;
}

// ---------------------------------------------------------------------------
// load and store phases (both engines)

template <typename T>
__device__ __forceinline__ T sgn_pow(int p) { return (p & 1) ? T(-1) : T(1); }

// values per thread and row held in registers by the load and store phases
constexpr int VPT = 8;

template <typename T>
struct RowPair {
    long long ia, oa;  // row a in src / dst
    bool      hb;      // row b exists
    int       ra, rhs; // row a, right-hand side
};

template <typename T>
__device__ __forceinline__ RowPair<T> row_pair(const YArgs<T>& a, int local)
{
    const int  npairs = (a.nx + 1) >> 1;
    const int  rhs = local / npairs, pair = local - rhs * npairs;
    const int  ra = 2 * pair;
    RowPair<T> rp;
    rp.hb  = ra + 1 < a.nx;
    rp.ra  = ra;
    rp.rhs = rhs;
    rp.ia  = (long long)rhs * a.lds + (long long)ra * a.N;
    rp.oa = (long long)rhs * a.ldd + (long long)ra * a.N;
    return rp;
}

// ConcatPoissonSolver2D::boundary_assembly and the hx^2 scaling
// (pe/concat/concat_solver2d.cpp:110-116,159-223) of value v at row i,
// column n: Dirichlet f += -g / h^2, Neumann f += +-g / h on the face line,
// then f *= hx^2 (the coefficients a.ga carry sign and spacing).
template <typename T>
__device__ __forceinline__ T assemble_at(const YArgs<T>& a, int rhs, int i, int n, T v)
{
    if (i == 0 && a.g[XNEG]) v += a.ga[XNEG] * (T)a.g[XNEG][(long long)rhs * a.N + n];
    if (i == a.nx - 1 && a.g[XPOS]) v += a.ga[XPOS] * (T)a.g[XPOS][(long long)rhs * a.N + n];
    if (n == 0 && a.g[YNEG]) v += a.ga[YNEG] * (T)a.g[YNEG][(long long)rhs * a.nx + i];
    if (n == a.N - 1 && a.g[YPOS]) v += a.ga[YPOS] * (T)a.g[YPOS][(long long)rhs * a.nx + i];
    return a.hx2 * v;
}

// Builds the packed extended sequence of the row pair in A (length L).
template <int PAD, typename T>
__device__ __forceinline__ void load_rows(const YArgs<T>& a, const RowPair<T>& rp, C2<T>* A)
{
    const int  NT = blockDim.x;
    const int  N = a.N, L = a.L;
    const bool s64 = sizeof(T) == 8 || a.src64;
    auto       ld_src = [&](long long q) -> T {
        return s64 ? (T) static_cast<const double*>(a.src)[q] : static_cast<const T*>(a.src)[q];
    };
    // The rows are read in chunks of VPT values per thread with every global
    // load of a chunk issued before the first LDS write, so the HBM latency
    // is paid once per chunk instead of once per element.
    const int kind = a.kind, inv = a.inverse;
    if (kind == T_PP && inv) {
        // periodic inverse: for each frequency q <= N/2 the coefficients of its
        // modes (m = 0 for q = N/2; 2a-1, 2a for q = N/2 - a; N-1 for q = 0)
        // give G_q = sum p_m exp(i pi (2m+1)/4); the Hermitian spectrum
        // X_q = G_q / 2i, X_{N-q} = conj X_q (X_0 = Im G_0, X_{N/2} = Im G_{N/2})
        // of both rows packed as X_a + i X_b
        const T r2 = T(0.70710678118654752440);
        for (int q = threadIdx.x; 2 * q <= N; q += NT) {
            const int a1 = N / 2 - q;  // modes 2 a1 - 1 and 2 a1 (when in range)
            T         gar = 0, gai = 0, gbr = 0, gbi = 0;
            for (int m = 2 * a1 - 1; m <= 2 * a1; m++) {
                if (m < 0 || m >= N) continue;
                const T  pa = ld_src(rp.ia + m), pb = rp.hb ? ld_src(rp.ia + N + m) : T(0);
                const int r = m & 3;
                const T   c = (r == 0 || r == 3) ? r2 : -r2, sn = (r < 2) ? r2 : -r2;
                gar += c * pa;
                gai += sn * pa;
                gbr += c * pb;
                gbi += sn * pb;
            }
            if (q == 0 || 2 * q == N) {
                A[zp<PAD>(q)] = {gai, gbi};
            } else {
                A[zp<PAD>(q)]     = {T(0.5) * (gai + gbr), T(0.5) * (gbi - gar)};
                A[zp<PAD>(N - q)] = {T(0.5) * (gai - gbr), T(0.5) * (gar + gbi)};
            }
        }
        return;
    }
    if (kind == T_NN && inv) {
        // DCT pair inverse by Makhoul's N-point algorithm: x = P^T Re FFT(w),
        // w_m = a_m exp(-i pi m / 2N), a_m = s_m p_{N-1-m}; the real part is
        // the FFT of the Hermitian part h_m = (w_m + conj w_{(N-m) mod N}) / 2,
        // so both rows pack as h_a + i h_b
        for (int m = threadIdx.x; m < N; m += NT) {
            const int mr = m == 0 ? 0 : N - m;
            const T   sm = m == 0 ? a.c1 : a.c0, sr = mr == 0 ? a.c1 : a.c0;  // sqrt(1/N) or -sqrt(2/N)
            const C2<T> em = a.aux[m], er = a.aux[mr];                        // exp(-i pi m / 2N)
            const T pam = sm * ld_src(rp.ia + (N - 1 - m)), pbm = rp.hb ? sm * ld_src(rp.ia + N + (N - 1 - m)) : T(0);
            const T par = sr * ld_src(rp.ia + (N - 1 - mr)), pbr = rp.hb ? sr * ld_src(rp.ia + N + (N - 1 - mr)) : T(0);
            const T har = T(0.5) * (pam * em.x + par * er.x), hai = T(0.5) * (pam * em.y - par * er.y);
            const T hbr = T(0.5) * (pbm * em.x + pbr * er.x), hbi = T(0.5) * (pbm * em.y - pbr * er.y);
            A[zp<PAD>(m)] = {har - hbi, hai + hbr};
        }
        return;
    }
    if (kind == T_DD) {
        // x_0 = 0, x_j = row[j-1] (j = 1..N): each thread reads the pair x_j,
        // x_{L-j} of both rows straight from global memory and writes the
        // half-length algorithm's inputs y_j = s_j (x_j + x_{L-j}) + (x_j -
        // x_{L-j}) / 2, y_{L-j} = s_j (x_j + x_{L-j}) - (x_j - x_{L-j}) / 2,
        // s_j = sin(pi j / L) (y_{L/2} = 2 s x_{L/2} for even L)
        if (threadIdx.x == 0) A[zp<PAD>(0)] = {T(0), T(0)};
        constexpr int PPT = VPT / 2;
        for (int base = 1; 2 * base <= L; base += PPT * NT) {
            T xa[PPT], xb[PPT], ua[PPT], ub[PPT];
#pragma unroll
            for (int k = 0; k < PPT; k++) {
                const int  j  = base + k * NT + threadIdx.x;
                const bool ok = 2 * j <= L;
                xa[k]         = ok ? ld_src(rp.ia + (j - 1)) : T(0);
                ua[k]         = ok ? ld_src(rp.ia + (N - j)) : T(0);
                xb[k]         = (ok && rp.hb) ? ld_src(rp.ia + N + (j - 1)) : T(0);
                ub[k]         = (ok && rp.hb) ? ld_src(rp.ia + N + (N - j)) : T(0);
            }
#pragma unroll
            for (int k = 0; k < PPT; k++) {
                const int j = base + k * NT + threadIdx.x;
                if (2 * j > L) break;
                T x = xa[k], y = xb[k], u = ua[k], w = ub[k];
                if (a.asmb) {
                    x = assemble_at(a, rp.rhs, rp.ra, j - 1, x);
                    u = assemble_at(a, rp.rhs, rp.ra, N - j, u);
                    if (rp.hb) {
                        y = assemble_at(a, rp.rhs, rp.ra + 1, j - 1, y);
                        w = assemble_at(a, rp.rhs, rp.ra + 1, N - j, w);
                    }
                }
                const T   sj = a.aux[j].x;
                const int jm = L - j;
                if (j == jm) {
                    A[zp<PAD>(j)] = {T(2) * sj * x, T(2) * sj * y};
                } else {
                    const T sa = x + u, sb = y + w, da = x - u, db = y - w;
                    A[zp<PAD>(j)]  = {sj * sa + T(0.5) * da, sj * sb + T(0.5) * db};
                    A[zp<PAD>(jm)] = {sj * sa - T(0.5) * da, sj * sb - T(0.5) * db};
                }
            }
        }
        return;
    }
    if (kind != T_PP && kind != T_NN) {
        if (threadIdx.x == 0) A[zp<PAD>(0)] = {T(0), T(0)};
        for (int p = N + 1 + threadIdx.x; p < L; p += NT) A[zp<PAD>(p)] = {T(0), T(0)};
    }
    for (int base = 0; base < N; base += VPT * NT) {
        T va[VPT], vb[VPT];
#pragma unroll
        for (int k = 0; k < VPT; k++) {
            const int n = base + k * NT + threadIdx.x;
            va[k]       = n < N ? ld_src(rp.ia + n) : T(0);
            vb[k]       = (n < N && rp.hb) ? ld_src(rp.ia + N + n) : T(0);
        }
#pragma unroll
        for (int k = 0; k < VPT; k++) {
            const int n = base + k * NT + threadIdx.x;
            if (n >= N) break;
            T x = va[k], y = vb[k];
            if (a.asmb) {
                x = assemble_at(a, rp.rhs, rp.ra, n, x);
                if (rp.hb) y = assemble_at(a, rp.rhs, rp.ra + 1, n, y);
            }
            if (kind == T_PP) {
                A[zp<PAD>(n)] = {x, y};
            } else if (kind == T_NN) {
                // Makhoul's reordering: even entries ascending, odd ones descending
                A[zp<PAD>((n & 1) ? N - 1 - (n >> 1) : (n >> 1))] = {x, y};
            } else if (kind == T_ND && inv) {
                A[zp<PAD>(n + 1)] = {x, y};
            } else {  // DN both ways, ND forward
                const T s = sgn_pow<T>(n + 1) * a.c0;
                A[zp<PAD>(n + 1)]  = {s * x, s * y};
            }
        }
    }
}

// Inclusive prefix sums of S[0..n) over the workgroup: a serial run per
// thread chunk, a wavefront scan of the chunk totals, the wavefronts'
// totals (wt, one per wavefront) added in order.  Summation depth is
// logarithmic in n apart from the chunk (ceil(n / blockDim) terms).
template <int PAD, typename T>
__device__ void block_scan(C2<T>* S, int n, C2<T>* wt)
{
    // S[i] lives at S[zp<PAD>(i)] (the transform's own buffer)
    const int NT = blockDim.x, t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int cpt = (n + NT - 1) / NT;
    const int b0 = min(n, t * cpt), b1 = min(n, b0 + cpt);
    C2<T>     run = {T(0), T(0)};
    for (int q = b0; q < b1; q++) {
        run              = cadd(run, S[zp<PAD>(q)]);
        S[zp<PAD>(q)] = run;
    }
    C2<T> v = run;
    for (int d = 1; d < 64; d <<= 1) {
        const C2<T> u = {__shfl_up(v.x, d, 64), __shfl_up(v.y, d, 64)};
        if (lane >= d) v = cadd(v, u);
    }
    C2<T> ex = {__shfl_up(v.x, 1, 64), __shfl_up(v.y, 1, 64)};
    if (lane == 0) ex = {T(0), T(0)};
    if (lane == 63) wt[wave] = v;
    __syncthreads();
    for (int w = 0; w < wave; w++) ex = cadd(wt[w], ex);
    for (int q = b0; q < b1; q++) S[zp<PAD>(q)] = cadd(S[zp<PAD>(q)], ex);
    __syncthreads();
}

// Splits the two rows' coefficients out of the spectrum Z and stores them.
template <int PAD, typename T>
__device__ __forceinline__ void store_rows(const YArgs<T>& a, const RowPair<T>& rp, C2<T>* Z, C2<T>* wt)
{
    const int  NT = blockDim.x;
    const int  N = a.N, L = a.L;
    const int  kind = a.kind, inv = a.inverse;
    const bool d64 = sizeof(T) == 8 || a.dst64;
    auto       st_dst = [&](long long q, T v) {
        if (d64) static_cast<double*>(a.dst)[q] = (double)v;
        else static_cast<T*>(a.dst)[q] = v;
    };
    if (kind == T_DD) {
        // rows a, b: F_a(q) = (Z_q + conj Z_{L-q}) / 2, F_b(q) = (Z_q - conj Z_{L-q}) / 2i
        // odd outputs X_1, X_3, .. X_{2Ko-1}: prefix sums of Re F(q), q < Ko;
        // even outputs X_2 .. X_{2Ke}: -Im F(k), k = 1..Ke.  Each thread
        // reads the pair Z_q, Z_{L-q} and writes its results back into the
        // same two slots (Re F into slot q, X_{2q} into slot L-q), so the
        // scan runs in place in the transform's buffer
        const int Ko = (N + 1) / 2, Ke = N / 2;
        for (int q = threadIdx.x; q <= Ke || q < Ko; q += NT) {
            const C2<T> zq = Z[zp<PAD>(q)];
            const C2<T> zm = q > 0 ? Z[zp<PAD>(L - q)] : C2<T>{T(0), T(0)};  // X_1 starts at Re F(0) / 2
            if (q >= 1) Z[zp<PAD>(L - q)] = {T(0.5) * (zm.y - zq.y), T(0.5) * (zq.x - zm.x)};
            if (q < Ko) Z[zp<PAD>(q)] = {T(0.5) * (zq.x + zm.x), T(0.5) * (zq.y + zm.y)};  // (Re F_a, Re F_b)
        }
        __syncthreads();
        block_scan<PAD>(Z, Ko, wt);
        for (int base = 0; base < N; base += VPT * NT) {
#pragma unroll
            for (int k = 0; k < VPT; k++) {
                const int j = base + k * NT + threadIdx.x;
                if (j >= N) break;
                // X_{j+1}: even ones in slot L - (j+1)/2, odd ones in slot j/2
                const C2<T> X = Z[zp<PAD>((j & 1) ? L - ((j + 1) >> 1) : (j >> 1))];
                st_dst(rp.oa + j, a.c0 * X.x);
                if (rp.hb) st_dst(rp.oa + N + j, a.c0 * X.y);
            }
        }
        return;
    }
    for (int base = 0; base < N; base += VPT * NT) {
        C2<T> sc[VPT];
        if (kind == T_NN && !inv) {
#pragma unroll
            for (int k = 0; k < VPT; k++) {
                const int j = base + k * NT + threadIdx.x;
                sc[k]       = j < N ? a.aux[N - 1 - j] : C2<T>{T(0), T(0)};
            }
        } else if ((kind == T_ND || kind == T_PP) && !inv) {
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
            if (kind == T_NN && !inv) {
                // C(m) = Re(exp(-i pi m / 2N) V_m), V_a / V_b of the rows split
                // from Z_m and Z_{(N-m) mod N}; mode j holds m = N-1-j
                const int   m  = N - 1 - j;
                const C2<T> zm = Z[zp<PAD>(m)], zr = Z[zp<PAD>(m == 0 ? 0 : N - m)];
                const T     var = T(0.5) * (zm.x + zr.x), vai = T(0.5) * (zm.y - zr.y);
                const T     vbr = T(0.5) * (zm.y + zr.y), vbi = T(0.5) * (zr.x - zm.x);
                const T     s   = (j == N - 1) ? a.c1 : a.c0;
                oa              = s * (sc[k].x * var - sc[k].y * vai);
                ob              = s * (sc[k].x * vbr - sc[k].y * vbi);
            } else if (kind == T_NN) {
                const C2<T> z = Z[zp<PAD>((j & 1) ? N - 1 - (j >> 1) : (j >> 1))];
                oa            = z.x;
                ob            = z.y;
            } else if (kind == T_PP && !inv) {
                // mode j of frequency q: sqrt(2/N) Im(w_j conj F_q), F_a / F_b
                // of the rows split from Z_q and Z_{N-q}
                const int   q  = N / 2 - (j + 1) / 2;
                const C2<T> zq = Z[zp<PAD>(q)], zm = Z[zp<PAD>(q == 0 ? 0 : N - q)];
                const T     far = T(0.5) * (zq.x + zm.x), fai = T(0.5) * (zq.y - zm.y);
                const T     fbr = T(0.5) * (zq.y + zm.y), fbi = T(0.5) * (zm.x - zq.x);
                oa              = a.c0 * (sc[k].y * far - sc[k].x * fai);
                ob              = a.c0 * (sc[k].y * fbr - sc[k].x * fbi);
            } else if (kind == T_PP) {
                // the inverse ran the forward FFT: r_n = Z_{(N - n) mod N}, n = j + 1
                const C2<T> z = Z[zp<PAD>(N - 1 - j)];
                oa            = a.c0 * z.x;
                ob            = a.c0 * z.y;
            } else if (kind == T_DN) {
                const C2<T> zm = Z[zp<PAD>(j + 1)], zc = Z[zp<PAD>(L - 1 - j)];
                const T     s  = sgn_pow<T>(j) * T(0.5);
                oa             = s * (zm.y - zc.y);
                ob             = s * (zc.x - zm.x);
            } else if (!inv) {  // ND forward
                const C2<T> zm = Z[zp<PAD>(j + 1)], zc = Z[zp<PAD>(L - 1 - j)];
                const T     rea = T(0.5) * (zm.x + zc.x), ima = T(0.5) * (zm.y - zc.y);
                const T     reb = T(0.5) * (zm.y + zc.y), imb = T(0.5) * (zc.x - zm.x);
                oa              = sc[k].x * rea - sc[k].y * ima;
                ob              = sc[k].x * reb - sc[k].y * imb;
            } else {  // ND inverse
                const int   kk = N - j;
                const C2<T> zk = Z[zp<PAD>(kk)], zc = Z[zp<PAD>(L - kk)];
                const T     s  = sgn_pow<T>(j + 1) * a.c0 * T(0.5);
                oa             = s * (zk.y - zc.y);
                ob             = s * (zc.x - zk.x);
            }
            st_dst(rp.oa + j, oa);
            if (rp.hb) st_dst(rp.oa + N + j, ob);
        }
    }
}

template <typename T>
__device__ __forceinline__ const YArgs<T>& find_problem(const YBatch<T>& bt, int blk, int& local)
{
    int q = 0;
    while (q + 1 < bt.nprob && blk >= bt.start[q + 1]) q++;
    local = blk - bt.start[q];
    return bt.p[q];
}
template <typename T>
__device__ __forceinline__ const YArgs<T>& find_problem(const JobTable<YArgs<T>>& bt, int blk, int& local)
{
    return table_find(bt, blk, local);
}
template <typename T, typename BT>
__device__ __forceinline__ const YArgs<T>& find_problem(const BT& bt, int& local)
{
    return find_problem<T>(bt, (int)blockIdx.x, local);
}

// threads per workgroup: chosen per plan (128, 256 or 512)
constexpr int kMaxNT = 512;

template <typename T, int OM, typename BT, int EPT>
__global__ __launch_bounds__(kMaxNT) void k_ytransform_sb(BT bt)
{
    int             local;
    const YArgs<T>& a  = find_problem<T>(bt, local);
    const RowPair<T> rp = row_pair(a, local);
    const int       NT = blockDim.x;

    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2<T>*    Z   = reinterpret_cast<C2<T>*>(smem_raw);
    C2<T>*    tlo = Z + zp<kZpad>(a.F) + 1;  // 64 twiddles W^m
    const int nhi = a.F / 64 + 1;
    C2<T>*    thi = tlo + 64;           // L/64 + 1 twiddles W^(64m)
    C2<T>*    cs  = thi + nhi;          // cos/sin tables of the odd-prime passes (FP32)
    double*   fr  = reinterpret_cast<double*>(cs + a.ncs);  // matrix-core fragments (FP64)
    C2<T>*    wt  = reinterpret_cast<C2<T>*>(fr + a.nfrag); // DD: wavefront totals of the prefix scan
    // the first table entry of each thread is loaded before the rows (the
    // two latencies overlap; at 128 threads and up that is the whole
    // twiddle and cosine table of a 1025-point plan) and stored once the
    // rows are in
    const int   ntab = 64 + nhi + a.ncs;
    const C2<T> tv   = (int)threadIdx.x < ntab ? a.W[threadIdx.x] : C2<T>{T(0), T(0)};
    load_rows<kZpad>(a, rp, Z);
    if ((int)threadIdx.x < ntab) tlo[threadIdx.x] = tv;
    for (int q = threadIdx.x + NT; q < ntab; q += NT) tlo[q] = a.W[q];
    for (int q = threadIdx.x; q < a.nfrag; q += NT) fr[q] = a.frag[q];
    const Tw<T> W{tlo, thi};
    __syncthreads();
    // every kind runs the forward FFT (the inverses by conjugate symmetry or
    // index reversal)
    const bool fwd = true;
    if (a.F != a.L) {
        if (fwd) run_bluestein<true, OM, EPT>(a, W, Z, cs, fr);
        else run_bluestein<false, OM, EPT>(a, W, Z, cs, fr);
    } else if (fwd) run_fft_sb<true, OM, EPT>(a, W, Z, cs, fr);
    else run_fft_sb<false, OM, EPT>(a, W, Z, cs, fr);
    store_rows<kZpad>(a, rp, Z, wt);
}

template <typename T, typename BT>
__global__ __launch_bounds__(kMaxNT) void k_ytransform_gen(BT bt)
{
    int             local;
    const YArgs<T>& a  = find_problem<T>(bt, local);
    const RowPair<T> rp = row_pair(a, local);
    const int       NT = blockDim.x;

    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2<T>*    A   = reinterpret_cast<C2<T>*>(smem_raw);
    C2<T>*    B   = A + a.L;
    C2<T>*    tlo = B + a.L;
    C2<T>*    thi = tlo + 64;
    const int nhi = a.L / 64 + 1;
    C2<T>*    wt  = thi + nhi;  // DD: wavefront totals of the prefix scan
    for (int q = threadIdx.x; q < 64 + nhi; q += NT) tlo[q] = a.W[q];
    const Tw<T> W{tlo, thi};

    load_rows<0>(a, rp, A);
    __syncthreads();
    C2<T>* Z = const_cast<C2<T>*>(run_fft_gen<true>(a, W, A, B));
    store_rows<0>(a, rp, Z, wt);
}

// ---------------------------------------------------------------------------
// global-memory engine: transforms whose buffers exceed one workgroup's LDS

// The engine keeps the two FFT buffers of a row pair in device memory (a
// plan-owned scratch of kGlobSlots slots of 2F values) and runs the
// two-buffer engine's passes on them; only the twiddle tables and the scan
// totals are in LDS.  A workgroup's global writes are visible to its other
// threads after a barrier (one CU, one L1).  Lengths with a prime factor
// above 47 take Bluestein over the 2,3,5-smooth F >= 2L - 1.  The grid is at
// most kGlobSlots workgroups; workgroup g owns slot g of every job's scratch
// and loops over the row pairs g, g + grid, ...
constexpr int kGlobSlots = 256;
constexpr int kGlobNT    = 512;

// Bluestein in the two device buffers: A holds z (length L) on entry; the
// returned buffer holds its DFT of sign FWD in [0, L).
template <bool FWD, typename T>
__device__ C2<T>* run_bluestein_glob(const YArgs<T>& a, const Tw<T> W, C2<T>* A, C2<T>* B)
{
    const int L = a.L, F = a.F, NT = blockDim.x;
    for (int n = threadIdx.x; n < F; n += NT) {
        C2<T> v{T(0), T(0)};
        if (n < L) {
            v = A[n];
            if (!FWD) v.y = -v.y;
            v = cmul(v, a.chirp[n]);
        }
        A[n] = v;
    }
    __syncthreads();
    C2<T>* P = run_fft_gen<true>(a, W, A, B);
    for (int k = threadIdx.x; k < F; k += NT) P[k] = cmul(P[k], a.bh[k]);
    __syncthreads();
    C2<T>* Q = run_fft_gen<false>(a, W, P, P == A ? B : A);
    for (int k = threadIdx.x; k < L; k += NT) {
        C2<T> v = cmul(Q[k], a.chirp[k]);
        if (!FWD) v.y = -v.y;
        Q[k] = v;
    }
    __syncthreads();
    return Q;
}

template <typename T, typename BT>
__global__ __launch_bounds__(kGlobNT) void k_ytransform_glob(BT bt, int total)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2<T>* tlo = reinterpret_cast<C2<T>*>(smem_raw);
    C2<T>* thi = tlo + 64;
    const int NT = blockDim.x;
    for (int blk = blockIdx.x; blk < total; blk += gridDim.x) {
        int              local;
        const YArgs<T>&  a   = find_problem<T>(bt, blk, local);
        const RowPair<T> rp  = row_pair(a, local);
        const int        nhi = a.F / 64 + 1;
        C2<T>*           wt  = thi + nhi;  // DD: wavefront totals of the prefix scan
        C2<T>*           A   = a.gs + (size_t)blockIdx.x * 2 * (size_t)a.F;
        C2<T>*           B   = A + a.F;
        __syncthreads();  // the previous row pair's readers of the tables and buffers are done
        for (int q = threadIdx.x; q < 64 + nhi; q += NT) tlo[q] = a.W[q];
        const Tw<T> W{tlo, thi};
        load_rows<0>(a, rp, A);
        __syncthreads();
        C2<T>* Z = a.F != a.L ? run_bluestein_glob<true>(a, W, A, B) : run_fft_gen<true>(a, W, A, B);
        store_rows<0>(a, rp, Z, wt);
    }
}

// ---------------------------------------------------------------------------
// host side

// Radix schedule: the odd primes above 5 first, largest first (their passes
// then run with Ns = 1, i.e. without twiddle factors), then 8, 4, 2, 3, 5.
// smallest 2^a 3^b 5^c >= n
static int smooth235(int n)
{
    for (int m = n;; m++) {
        int r = m;
        for (int p : {2, 3, 5})
            while (r % p == 0) r /= p;
        if (r == 1) return m;
    }
}

// Forward DFT (sign -) of a 2,3,5-smooth length in long double, recursive
// decimation in time: X_k = sum_r W^(rk) S_r[k mod m], S_r the DFT of the
// r-th decimated subsequence.  Host-side table construction only.
static void fft_ld(std::vector<std::complex<long double>>& x)
{
    const size_t n = x.size();
    if (n <= 1) return;
    const size_t p = n % 2 == 0 ? 2 : n % 3 == 0 ? 3 : 5;
    const size_t m = n / p;
    std::vector<std::vector<std::complex<long double>>> sub(p, std::vector<std::complex<long double>>(m));
    for (size_t r = 0; r < p; r++)
        for (size_t j = 0; j < m; j++) sub[r][j] = x[j * p + r];
    for (auto& v : sub) fft_ld(v);
    const long double PI = 3.141592653589793238462643383279502884L;
    for (size_t k = 0; k < n; k++) {
        std::complex<long double> acc = sub[0][k % m];
        for (size_t r = 1; r < p; r++) {
            const long double ang = -2.0L * PI * (long double)((r * k) % n) / (long double)n;
            acc += sub[r][k % m] * std::complex<long double>(cosl(ang), sinl(ang));
        }
        x[k] = acc;
    }
}

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

// Whether the single-buffer engine runs this schedule with NT threads: every
// pass's butterflies (or odd-prime work items) fit the per-thread registers.
static int sb_bpt_rt(int R, int ept)
{
    return (R == 2 || R == 3 || R == 4 || R == 5 || R == 8) ? (ept + R - 1) / R : 0;
}
static bool sb_fits(const int* rad, int nrad, int L, int NT, int odd, int ept = kEPT)
{
    const bool mfma = odd == ODD_MFMA;
    for (int s = 0; s < nrad; s++) {
        const int R = rad[s], nb = L / R;
        if (R <= 8 && R != 7) {
            if ((nb + NT - 1) / NT > sb_bpt_rt(R, ept)) return false;
        } else if (R > kMaxOddPrime) {
            return false;
        } else if (mfma) {
            if ((2 * nb + 15) / 16 > kMaxNTw * (NT / 64)) return false;
        } else {
            const int nq = odd == ODD_VALU3 ? 3 : kNQ;
            const int H = (R - 1) / 2, G = (H + nq - 1) / nq, nbp = odd_slots(nb);
            if ((nbp * G + NT - 1) / NT > kIPT) return false;
        }
    }
    return true;
}

// odd-prime forms a plan may use, preferred first
template <typename T>
static std::vector<int> odd_forms(bool has_odd)
{
    if (!has_odd) return {ODD_NONE};
    if (sizeof(T) == 8) return {ODD_MFMA, ODD_VALU, ODD_VALU3};
    return {ODD_VALU, ODD_VALU3};
}

template <typename T>
static size_t lds_bytes(const TransformPlan<T>& tp, int odd)
{
    const size_t ncs   = (odd == ODD_VALU || odd == ODD_VALU3) ? (size_t)tp.ncs : 0;
    const size_t nfrag = odd == ODD_MFMA ? tp.frag.n : 0;
    const size_t nhi = tp.F / 64 + 1;
    const size_t zl  = (size_t)tp.F + (kZpad ? ((size_t)tp.F >> kZpad) : 0) + 1;
    // wavefront totals of the DD prefix scan (which runs inside Z), one per wavefront
    const size_t sc = tp.kind == T_DD ? (size_t)kMaxNT / 64 : 0;
    if (tp.engine == 1) return (zl + 64 + nhi + ncs + sc) * sizeof(C2<T>) + nfrag * sizeof(double);
    if (tp.engine == 2) return (64 + nhi + (size_t)kGlobNT / 64) * sizeof(C2<T>);
    return (2 * (size_t)tp.L + 64 + nhi + sc) * sizeof(C2<T>);
}

template <typename T>
void TransformPlan<T>::build(int k, int n)
{
    kind = k;
    N    = n;
    if (kind == T_DD) L = N + 1;
    else if (kind == T_NN) L = N;  // Makhoul: one N-point FFT per row pair
    else if (kind == T_PP) L = N;
    else L = 2 * N + 1;
    if (kind == T_PP && N % 2 != 0) throw Error(FP_E_INVALID, "the periodic transform needs an even length");
    // prime factors above 47: Bluestein over the smallest 2,3,5-smooth F >=
    // 2L - 1, in LDS when its single buffer fits there, else in the
    // global-memory engine
    F        = L;
    ept      = kEPT;
    engine   = 1;
    bool big = false, glob = false;
    try {
        factor_radices(L, rad, nrad);
        for (int s = 0; s < nrad; s++) big |= rad[s] > kMaxOddPrime;
    } catch (const Error&) {
        big = true;
    }
    if (big) {
        const int    Fb   = smooth235(2 * L - 1);
        const size_t need = ((size_t)Fb + ((size_t)Fb >> kZpad) + 1 + 64 + Fb / 64 + 1) * sizeof(C2<T>) + 4 * sizeof(T);
        int          rb[kMaxRadices], nb = 0;
        factor_radices(Fb, rb, nb);
        int fits = 0;  // values per thread that fit
        for (int e : {kEPT, kEPTLong})
            for (int t : {128, 256, 512})
                if (!fits && sb_fits(rb, nb, Fb, t, ODD_NONE, e)) fits = e;
        F    = Fb;
        nrad = nb;
        for (int s = 0; s < nb; s++) rad[s] = rb[s];
        if (fits && need <= 160 * 1024) ept = fits;
        else glob = true;
    }
    const long double PI  = 3.141592653589793238462643383279502884L;
    const int         nhi = F / 64 + 1;
    // odd-prime passes: (cos, sin)(2 pi m / R), m = 0..R-1, one table per
    // pass in schedule order
    std::vector<long double> cs;
    ncs = 0;
    for (int s = 0; s < nrad; s++) {
        const int R = rad[s];
        if (R <= 8 && R != 7) continue;
        for (int m = 0; m < R; m++) {
            cs.push_back(cosl(2.0L * PI * m / R));
            cs.push_back(sinl(2.0L * PI * m / R));
        }
        ncs += R;
    }
    // FP64: matrix-core fragments [cos/sin][M tile][K step][lane] of every odd-prime pass
    std::vector<double> fr;
    if (sizeof(T) == 8) {
        for (int s = 0; s < nrad; s++) {
            const int R = rad[s];
            if (R <= 8 && R != 7) continue;
            const int H = (R - 1) / 2, KT = mfma_ksteps(R);
            for (int f = 0; f < 2; f++)
                for (int mt = 0; mt < 2; mt++)
                    for (int ks = 0; ks < KT; ks++)
                        for (int l = 0; l < 64; l++) {
                            const int         q = 16 * mt + (l & 15), r = 4 * ks + (l >> 4) + 1;
                            const long double ang = 2.0L * PI * (long double)(((long long)q * r) % R) / R;
                            long double       v = 0.0L;
                            if (q <= H && r <= H) v = f == 0 ? (q == 0 ? 1.0L : cosl(ang)) : (q == 0 ? 0.0L : sinl(ang));
                            fr.push_back((double)v);
                        }
        }
    }
    std::vector<T> w(2 * (size_t)(64 + nhi + ncs)), x(2 * (size_t)(N + 1), T(0));
    for (int q = 0; q < 64 + nhi; q++) {
        const long long   t   = q < 64 ? q : 64LL * (q - 64);
        const long double ang = -2.0L * PI * (long double)(t % F) / (long double)F;
        w[2 * q]              = (T)cosl(ang);
        w[2 * q + 1]          = (T)sinl(ang);
    }
    for (int q = 0; q < 2 * ncs; q++) w[2 * (size_t)(64 + nhi) + q] = (T)cs[q];
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
    } else if (kind == T_PP) {
        // w_m = exp(i (pi (2m+1)/4 + 2 pi q_m / N)), q_m = N/2 - ceil(m/2): the
        // angle in eighths of a turn plus q_m / N turns, reduced exactly
        for (int m = 0; m < N; m++) {
            const long long   q   = N / 2 - (m + 1) / 2;
            const long double ang = PI * (long double)((2LL * m + 1) % 8) / 4.0L + 2.0L * PI * (long double)q / N;
            x[2 * m]              = (T)cosl(ang);
            x[2 * m + 1]          = (T)sinl(ang);
        }
        c0 = (T)std::sqrt(2.0 / N);
    } else {
        // DD: sin(pi j / L), j < L, and the orthonormal scale sqrt(2 / (N+1))
        for (int j = 0; j < L; j++) x[2 * j] = (T)sinl(PI * (long double)j / (long double)L);
        c0 = (T)std::sqrt(2.0 / (N + 1.0));
    }
    if (F != L) {
        // chirp c_n = exp(-i pi n^2 / L) (n^2 reduced mod 2L), then the DFT of
        // length F of b (b_m = conj c_m for |m| < L, wrapped), scaled by 1/F
        std::vector<std::complex<long double>> c(L), b(F, 0.0L);
        for (int n = 0; n < L; n++) {
            const long long   q   = ((long long)n * n) % (2LL * L);
            const long double ang = -PI * (long double)q / (long double)L;
            c[n]                  = {cosl(ang), sinl(ang)};
        }
        for (int m = 0; m < L; m++) {
            b[m] = std::conj(c[m]);
            if (m) b[F - m] = std::conj(c[m]);
        }
        fft_ld(b);
        std::vector<T> t(2 * (size_t)(L + F));
        for (int n = 0; n < L; n++) {
            t[2 * n]     = (T)c[n].real();
            t[2 * n + 1] = (T)c[n].imag();
        }
        for (int k = 0; k < F; k++) {
            t[2 * (size_t)(L + k)]     = (T)(b[k].real() / F);
            t[2 * (size_t)(L + k) + 1] = (T)(b[k].imag() / F);
        }
        blu.alloc(t.size());
        blu.upload(t.data(), t.size());
    }
    W.alloc(w.size());
    W.upload(w.data(), w.size());
    frag.alloc(fr.size());
    if (!fr.empty()) frag.upload(fr.data(), fr.size());
    aux.alloc(x.size());
    aux.upload(x.data(), x.size());
    // engine: single buffer at 256 threads when it fits (the tuner may change
    // the thread count and the odd-prime form among those that fit), else
    // the two-buffer fallback
    bool has_odd = false;
    for (int s = 0; s < nrad; s++) has_odd |= rad[s] == 7 || rad[s] > 8;
    nt          = 256;
    odd         = ODD_NONE;
    bool chosen = false;
    engine      = 1;  // lds_bytes below sizes the single buffer
    for (int form : odd_forms<T>(glob ? false : has_odd)) {
        if (glob) break;
        for (int t : {256, 512, 128})
            if (sb_fits(rad, nrad, F, t, form, ept) && lds_bytes(*this, form) <= 160 * 1024) {
                chosen = true;
                nt     = t;
                odd    = form;
                break;
            }
        if (chosen) break;
    }
    if (!chosen) {
        engine = 0;
        if (glob || lds_bytes(*this, ODD_NONE) > 160 * 1024) {
            // the two buffers in device memory, one pair per workgroup slot
            engine = 2;
            nt     = kGlobNT;
            gscr.alloc((size_t)kGlobSlots * 2 * 2 * (size_t)F);
        }
    }
}

template struct TransformPlan<double>;
template struct TransformPlan<float>;

// odd-prime form a job runs with: the job's override, else its plan's
template <typename T>
static int job_form(const TransformJob<T>& j)
{
    const TransformPlan<T>& tp = *j.plan;
    if (tp.engine != 1) return 0;
    return j.odd >= 0 ? j.odd : tp.odd;
}

template <typename T>
static YArgs<T> yargs(const TransformJob<T>& j, int inverse)
{
    const int form = job_form(j);
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
    a.F       = tp.F;
    a.chirp   = tp.F != tp.L ? reinterpret_cast<const C2<T>*>(tp.blu.p) : nullptr;
    a.bh      = tp.F != tp.L ? reinterpret_cast<const C2<T>*>(tp.blu.p) + tp.L : nullptr;
    a.gs      = tp.engine == 2 ? reinterpret_cast<C2<T>*>(tp.gscr.p) : nullptr;
    a.kind    = tp.kind;
    a.inverse = inverse;
    a.nrad    = tp.nrad;
    for (int i = 0; i < tp.nrad; i++) a.rad[i] = tp.rad[i];
    a.W   = reinterpret_cast<const C2<T>*>(tp.W.p);
    a.ncs   = (tp.engine == 1 && (form == ODD_VALU || form == ODD_VALU3)) ? tp.ncs : 0;
    a.frag  = tp.frag.p;
    a.nfrag = (tp.engine == 1 && form == ODD_MFMA) ? (int)tp.frag.n : 0;
    a.aux = reinterpret_cast<const C2<T>*>(tp.aux.p);
    a.c0  = tp.c0;
    a.c1  = tp.c1;
    a.asmb = j.asmb;
    a.hx2  = (T)j.hx2;
    for (int f = 0; f < 4; f++) {
        a.g[f]  = j.g[f];
        a.ga[f] = (T)j.ga[f];
    }
    return a;
}

// The kernel of an engine and odd-prime form for the batch type B (YBatch
// by value, or a JobTable in device memory).
template <typename T, typename B>
static void launch_kernel(int engine, int odd, int ept, int blocks, int NT, size_t lds, hipStream_t s, const B& bt)
{
    if (engine == 1 && ept != kEPT && odd != ODD_NONE) throw Error(FP_E_INVALID, "long transforms are 2,3,5-smooth");
    if (engine == 2) {
        // at most kGlobSlots workgroups, each looping over row pairs in its own scratch slot
        hipLaunchKernelGGL((k_ytransform_glob<T, B>), dim3(std::min(blocks, kGlobSlots)), dim3(kGlobNT), lds, s, bt,
                           blocks);
        FP_HIP(hipGetLastError());
        return;
    }
    const void* fn = engine == 0         ? (const void*)k_ytransform_gen<T, B>
                     : ept != kEPT       ? (const void*)k_ytransform_sb<T, ODD_NONE, B, kEPTLong>
                     : odd == ODD_NONE   ? (const void*)k_ytransform_sb<T, ODD_NONE, B, kEPT>
                     : odd == ODD_MFMA   ? (const void*)k_ytransform_sb<T, ODD_MFMA, B, kEPT>
                     : odd == ODD_VALU3  ? (const void*)k_ytransform_sb<T, ODD_VALU3, B, kEPT>
                                         : (const void*)k_ytransform_sb<T, ODD_VALU, B, kEPT>;
    // one attribute flag per kernel instantiation
    static std::mutex            mu;
    static std::set<const void*> attr_set;
    if (lds > 64 * 1024) {
        std::lock_guard<std::mutex> lk(mu);
        if (!attr_set.count(fn)) {
            FP_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
            attr_set.insert(fn);
        }
    }
    if (engine == 0) hipLaunchKernelGGL((k_ytransform_gen<T, B>), dim3(blocks), dim3(NT), lds, s, bt);
    else if (ept != kEPT)
        hipLaunchKernelGGL((k_ytransform_sb<T, ODD_NONE, B, kEPTLong>), dim3(blocks), dim3(NT), lds, s, bt);
    else if (odd == ODD_NONE)
        hipLaunchKernelGGL((k_ytransform_sb<T, ODD_NONE, B, kEPT>), dim3(blocks), dim3(NT), lds, s, bt);
    else if (odd == ODD_MFMA)
        hipLaunchKernelGGL((k_ytransform_sb<T, ODD_MFMA, B, kEPT>), dim3(blocks), dim3(NT), lds, s, bt);
    else if (odd == ODD_VALU3)
        hipLaunchKernelGGL((k_ytransform_sb<T, ODD_VALU3, B, kEPT>), dim3(blocks), dim3(NT), lds, s, bt);
    else hipLaunchKernelGGL((k_ytransform_sb<T, ODD_VALU, B, kEPT>), dim3(blocks), dim3(NT), lds, s, bt);
    FP_HIP(hipGetLastError());
}

// One launch for the jobs of one engine, thread count and odd-prime form:
// their descriptors as kernel arguments up to kMaxBatch jobs, else in a
// device-memory table.
template <typename T>
static void launch_group(const TransformJob<T>* const* jobs, int n, int engine, int NT, int odd, int ept,
                         int inverse, hipStream_t s)
{
    std::vector<YArgs<T>> args(n);
    std::vector<int>      start(n + 1);
    size_t                lds    = 0;
    int                   blocks = 0;
    for (int q = 0; q < n; q++) {
        const TransformJob<T>& j = *jobs[q];
        args[q]                  = yargs(j, inverse);
        start[q]                 = blocks;
        const long long nb       = (long long)((j.nx + 1) / 2) * j.nrhs;
        if (blocks + nb > (1LL << 30)) throw Error(FP_E_INVALID, "too many transform workgroups in one launch");
        blocks += (int)nb;
        lds = std::max(lds, lds_bytes(*j.plan, odd));
        // a shape that leaves butterflies without a thread would give wrong
        // coefficients without any error: refuse it
        if (engine == 1 && !sb_fits(j.plan->rad, j.plan->nrad, j.plan->F, NT, odd, j.plan->ept))
            throw Error(FP_E_INVALID, "transform launch shape does not fit the plan");
    }
    start[n] = blocks;
    if (lds > 160 * 1024) throw Error(FP_E_INVALID, "transform length exceeds the LDS of one workgroup");
    if (blocks == 0) return;
    if (engine == 1 && odd == ODD_MFMA && sizeof(T) != 8) throw Error(FP_E_INVALID, "matrix-core passes are FP64");
    if (n <= kMaxBatch) {
        YBatch<T> bt{};
        bt.nprob = n;
        for (int q = 0; q < n; q++) {
            bt.p[q]     = args[q];
            bt.start[q] = start[q];
        }
        bt.start[n] = blocks;
        launch_kernel<T>(engine, odd, ept, blocks, NT, lds, s, bt);
        return;
    }
    void*                  mem = nullptr;
    const JobTable<YArgs<T>> t = table_upload(args, start, s, mem);
    try {
        launch_kernel<T>(engine, odd, ept, blocks, NT, lds, s, t);
    } catch (...) {
        table_release(mem, s);
        throw;
    }
    table_release(mem, s);
}

// Jobs are grouped by (engine, threads per workgroup, odd-prime form), one
// launch per group.
template <typename T>
void launch_ytransform_batch(const TransformJob<T>* jobs, int njobs, int inverse, hipStream_t s)
{
    std::map<std::array<int, 4>, std::vector<const TransformJob<T>*>> groups;
    for (int k = 0; k < njobs; k++) {
        if (jobs[k].nrhs <= 0 || jobs[k].nx <= 0) continue;
        const TransformPlan<T>& tp = *jobs[k].plan;
        groups[{tp.engine, jobs[k].nt ? jobs[k].nt : tp.nt, job_form(jobs[k]), tp.ept}].push_back(&jobs[k]);
    }
    for (auto& kv : groups)
        launch_group<T>(kv.second.data(), (int)kv.second.size(), kv.first[0], kv.first[1], kv.first[2], kv.first[3],
                        inverse, s);
}

// The launches of launch_ytransform_batch, prepared: one per group of
// (engine, threads, odd-prime form, values per thread), each group's jobs in
// decreasing FFT length (the work of one workgroup grows with it).
template <typename T>
std::vector<GroupPlan> plan_ytransform_groups(const TransformJob<T>* jobs, int njobs, int inverse)
{
    std::map<std::array<int, 4>, std::vector<const TransformJob<T>*>> groups;
    for (int k = 0; k < njobs; k++) {
        if (jobs[k].nrhs <= 0 || jobs[k].nx <= 0) continue;
        const TransformPlan<T>& tp = *jobs[k].plan;
        groups[{tp.engine, jobs[k].nt ? jobs[k].nt : tp.nt, job_form(jobs[k]), tp.ept}].push_back(&jobs[k]);
    }
    std::vector<GroupPlan> out;
    for (auto& kv : groups) {
        std::vector<const TransformJob<T>*>& js = kv.second;
        std::stable_sort(js.begin(), js.end(), [](const TransformJob<T>* a, const TransformJob<T>* b) {
            return a->plan->F > b->plan->F;
        });
        GroupPlan g;
        g.engine = kv.first[0];
        g.nt     = kv.first[1];
        g.odd    = kv.first[2];
        g.ept    = kv.first[3];
        g.nprob  = (int)js.size();
        std::vector<YArgs<T>> args(js.size());
        std::vector<int>      start(js.size() + 1);
        long long             blocks = 0;
        for (size_t q = 0; q < js.size(); q++) {
            const TransformJob<T>& j = *js[q];
            args[q]                  = yargs(j, inverse);
            start[q]                 = (int)blocks;
            blocks += (long long)((j.nx + 1) / 2) * j.nrhs;
            if (blocks > (1LL << 30)) throw Error(FP_E_INVALID, "too many transform workgroups in one launch");
            g.lds = std::max(g.lds, lds_bytes(*j.plan, g.odd));
            if (g.engine == 1 && !sb_fits(j.plan->rad, j.plan->nrad, j.plan->F, g.nt, g.odd, j.plan->ept))
                throw Error(FP_E_INVALID, "transform launch shape does not fit the plan");
        }
        start[js.size()] = (int)blocks;
        if (g.lds > 160 * 1024) throw Error(FP_E_INVALID, "transform length exceeds the LDS of one workgroup");
        if (g.engine == 1 && g.odd == ODD_MFMA && sizeof(T) != 8) throw Error(FP_E_INVALID, "matrix-core passes are FP64");
        g.blocks    = (int)blocks;
        g.threads   = g.nt;
        g.start_off = (args.size() * sizeof(YArgs<T>) + 255) & ~(size_t)255;
        g.table.resize(g.start_off + start.size() * sizeof(int));
        std::memcpy(g.table.data(), args.data(), args.size() * sizeof(YArgs<T>));
        std::memcpy(g.table.data() + g.start_off, start.data(), start.size() * sizeof(int));
        if (g.blocks > 0) out.push_back(std::move(g));
    }
    return out;
}

template <typename T>
void launch_ytransform_group(const GroupPlan& g, const unsigned char* dtable, hipStream_t s)
{
    JobTable<YArgs<T>> t;
    t.nprob = g.nprob;
    t.p     = reinterpret_cast<const YArgs<T>*>(dtable);
    t.start = reinterpret_cast<const int*>(dtable + g.start_off);
    launch_kernel<T>(g.engine, g.odd, g.ept, g.blocks, g.nt, g.lds, s, t);
}
template std::vector<GroupPlan> plan_ytransform_groups<double>(const TransformJob<double>*, int, int);
template std::vector<GroupPlan> plan_ytransform_groups<float>(const TransformJob<float>*, int, int);
template void launch_ytransform_group<double>(const GroupPlan&, const unsigned char*, hipStream_t);
template void launch_ytransform_group<float>(const GroupPlan&, const unsigned char*, hipStream_t);

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
    if (nrhs == 1 && tp.nt1) {
        j.nt  = tp.nt1;
        j.odd = tp.odd1;
    }
    launch_ytransform_batch<T>(&j, 1, inverse, s);
}

// Threads per workgroup and odd-prime pass form of a plan's transform,
// chosen by timing the forward transform of an nx-row scratch field with
// every combination that fits, twice: with enough right-hand sides for about
// 32 MB (the throughput regime; nt, odd) and with one (a launch of a few
// workgroups per CU, bound by each workgroup's latency; nt1, odd1, used for
// launches of one field).  Each candidate keeps its best of three rounds.
// Recorded per (kind, N, nx, precision) (tune_store).

template <typename T>
void tune_transform(TransformPlan<T>& tp, int nx, bool measure)
{
    if (tp.engine != 1) return;
    bool has_odd = false;
    for (int s = 0; s < tp.nrad; s++) has_odd |= tp.rad[s] == 7 || tp.rad[s] > 8;
    std::vector<std::array<int, 2>> cand;  // (threads, odd-prime form)
    for (int form : odd_forms<T>(has_odd))
        for (int t : {128, 256, 512})
            if (sb_fits(tp.rad, tp.nrad, tp.F, t, form, tp.ept) && lds_bytes(tp, form) <= 160 * 1024) cand.push_back({t, form});
    // FP_ODD_FORM=<form> (tests): that odd-prime form, at the smallest
    // thread count it fits, for every launch, without timing
    if (const char* env = std::getenv("FP_ODD_FORM")) {
        const int want = std::atoi(env);
        for (const auto& c : cand)
            if (c[1] == want) {
                tp.nt = tp.nt1 = c[0];
                tp.odd = tp.odd1 = want;
                return;
            }
    }
    if (!measure && !cand.empty()) {
        // estimated shape: the commonest winner of the timed choices, 128
        // threads with three pairs per odd-prime work item, where it fits
        std::array<int, 2> e = cand[0];
        for (int form : {(int)ODD_VALU3, (int)ODD_VALU, (int)ODD_MFMA, (int)ODD_NONE}) {
            bool found = false;
            for (const auto& c : cand)
                if (c[1] == form) {
                    e     = c;
                    found = true;
                    break;
                }
            if (found) break;
        }
        tp.nt = tp.nt1 = e[0];
        tp.odd = tp.odd1 = e[1];
        return;
    }
    if (cand.size() < 2 || (long long)nx * tp.N < (1LL << 16)) return;
    const std::string key = "yt " + std::to_string(tp.kind) + " " + std::to_string(tp.N) + " " +
                            std::to_string(nx) + " " + std::to_string(sizeof(T));
    {
        std::vector<int> v;
        if (tune_lookup(key, v) && v.size() == 4) {
            for (const auto& c : cand)
                if (c[0] == v[0] && c[1] == v[1])
                    for (const auto& c1 : cand)
                        if (c1[0] == v[2] && c1[1] == v[3]) {
                            tp.nt   = v[0];
                            tp.odd  = v[1];
                            tp.nt1  = v[2];
                            tp.odd1 = v[3];
                            return;
                        }
        }
    }
    const long long n1   = (long long)nx * tp.N;
    const int       nmax = (int)std::max<long long>(1, std::min<long long>(64, (32LL << 20) / (n1 * sizeof(T))));
    DevBuf<T>       buf((size_t)(n1 * nmax));
    FP_HIP(hipMemset(buf.p, 0, (size_t)(n1 * nmax) * sizeof(T)));
    hipEvent_t e0, e1;
    FP_HIP(hipEventCreate(&e0));
    FP_HIP(hipEventCreate(&e1));
    tp.nt1 = 0;
    auto pick = [&](int nrhs) {
        std::vector<float> best_of(cand.size(), 1e30f);
        for (int round = 0; round < 3; round++)
            for (size_t ci = 0; ci < cand.size(); ci++) {
                tp.nt  = cand[ci][0];
                tp.odd = cand[ci][1];
                launch_ytransform<T>(tp, buf.p, n1, nx, nrhs, 0, nullptr);
                FP_HIP(hipEventRecord(e0, nullptr));
                for (int r = 0; r < 3; r++) launch_ytransform<T>(tp, buf.p, n1, nx, nrhs, 0, nullptr);
                FP_HIP(hipEventRecord(e1, nullptr));
                FP_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                FP_HIP(hipEventElapsedTime(&ms, e0, e1));
                best_of[ci] = std::min(best_of[ci], ms);
            }
        size_t ib = 0;
        for (size_t ci = 1; ci < cand.size(); ci++)
            if (best_of[ci] < best_of[ib]) ib = ci;
        return cand[ib];
    };
    const std::array<int, 2> one = pick(1);
    const std::array<int, 2> many = nmax > 1 ? pick(nmax) : one;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    tp.nt   = many[0];
    tp.odd  = many[1];
    tp.nt1  = one[0];
    tp.odd1 = one[1];
    tune_store(key, {tp.nt, tp.odd, tp.nt1, tp.odd1});
}
// Launch shape for a batch of transforms launched together (the domains of
// a composite solve): every job of the single-buffer engine takes the same
// thread count and odd-prime pass form where they fit (a plan without odd
// primes fits any form), so that jobs whose own tuned shapes differ share
// one launch and run side by side.  The forward batch is timed with each
// candidate, best of three interleaved rounds; {0, -1} (every job its own
// shape) is one of them.  The jobs must point at valid fields.
template <typename T>
bool batch_shape_fits(const TransformPlan<T>& tp, int nt, int form)
{
    if (tp.engine != 1) return false;
    const int f  = form < 0 ? tp.odd : form;
    const int t  = nt ? nt : tp.nt;
    const bool ok = sb_fits(tp.rad, tp.nrad, tp.F, t, f, tp.ept) && lds_bytes(tp, f) <= 160 * 1024;
    return ok && (f != ODD_NONE || tp.odd == ODD_NONE) && (tp.ept == kEPT || f == ODD_NONE);
}
template <typename T>
void apply_batch_shape(TransformJob<T>& j, const std::array<int, 2>& shape)
{
    j.nt  = 0;
    j.odd = -1;
    if (batch_shape_fits(*j.plan, shape[0], shape[1])) {
        j.nt  = shape[0];
        j.odd = shape[1];
    }
}
template bool batch_shape_fits<double>(const TransformPlan<double>&, int, int);
template bool batch_shape_fits<float>(const TransformPlan<float>&, int, int);
template void apply_batch_shape<double>(TransformJob<double>&, const std::array<int, 2>&);
template void apply_batch_shape<float>(TransformJob<float>&, const std::array<int, 2>&);

template <typename T>
std::array<int, 2> tune_batch_shape(std::vector<TransformJob<T>> jobs, hipStream_t s)
{
    if (jobs.size() < 2) return {0, -1};
    std::string key = "ytb " + std::to_string(sizeof(T));
    for (const auto& j : jobs)
        key += " " + std::to_string(j.plan->kind) + ":" + std::to_string(j.plan->N) + ":" + std::to_string(j.nx);
    {
        std::vector<int> v;
        if (tune_lookup(key, v) && v.size() == 2) return {v[0], v[1]};
    }
    std::vector<std::array<int, 2>> cands;
    for (int nt : {0, 128, 256, 512})
        for (int form : {-1, (int)ODD_MFMA, (int)ODD_VALU, (int)ODD_VALU3})
            if (!(form == ODD_MFMA && sizeof(T) != 8)) cands.push_back({nt, form});
    hipEvent_t e0, e1;
    FP_HIP(hipEventCreate(&e0));
    FP_HIP(hipEventCreate(&e1));
    std::vector<float> best_of(cands.size(), 1e30f);
    for (int round = 0; round < 3; round++)
        for (size_t ci = 0; ci < cands.size(); ci++) {
            const int nt = cands[ci][0], form = cands[ci][1];
            bool      any = nt == 0 && form < 0;
            for (auto& j : jobs) {
                apply_batch_shape(j, cands[ci]);
                if (j.nt != 0 || j.odd != -1) any = true;
            }
            if (!any) continue;
            launch_ytransform_batch<T>(jobs.data(), (int)jobs.size(), 0, s);
            FP_HIP(hipEventRecord(e0, s));
            for (int r = 0; r < 5; r++) launch_ytransform_batch<T>(jobs.data(), (int)jobs.size(), 0, s);
            FP_HIP(hipEventRecord(e1, s));
            FP_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            FP_HIP(hipEventElapsedTime(&ms, e0, e1));
            best_of[ci] = std::min(best_of[ci], ms);
        }
    size_t ib = 0;
    for (size_t ci = 1; ci < cands.size(); ci++)
        if (best_of[ci] < best_of[ib]) ib = ci;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    tune_store(key, {cands[ib][0], cands[ib][1]});
    return cands[ib];
}
template std::array<int, 2> tune_batch_shape<double>(std::vector<TransformJob<double>>, hipStream_t);
template std::array<int, 2> tune_batch_shape<float>(std::vector<TransformJob<float>>, hipStream_t);
template void tune_transform<double>(TransformPlan<double>&, int, bool);
template void tune_transform<float>(TransformPlan<float>&, int, bool);

template void launch_ytransform_batch<double>(const TransformJob<double>*, int, int, hipStream_t);
template void launch_ytransform_batch<float>(const TransformJob<float>*, int, int, hipStream_t);
template void launch_ytransform<double>(const TransformPlan<double>&, double*, long long, int, int, int, hipStream_t);
template void launch_ytransform<float>(const TransformPlan<float>&, float*, long long, int, int, int, hipStream_t);

}  // namespace fp
