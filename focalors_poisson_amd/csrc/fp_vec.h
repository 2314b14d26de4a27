// Declarations of the composite-solve vector kernels (fp_vec.hip).
#pragma once

#include "fp_internal.h"

#include <algorithm>

namespace fp {

template <typename T>
struct AsmFaces {
    const T* g[4];  // ghost data per face (nrhs arrays back to back) or null
    T        a[4];  // coefficient per face
};

template <typename T>
void launch_assemble(T* f, int nx, int ny, long long ld, int nrhs, T hx2, const AsmFaces<T>& fc, hipStream_t s);
template <typename T>
void launch_bond_add(T* f, int nx, int ny, long long ldf, const T* nb, int nbx, int nby, long long ldn, int loc, T a,
                     int nrhs, hipStream_t s);
template <typename T>
void launch_fill(T* p, long long n, T v, hipStream_t s);
void launch_face_extract(const double* f, int nx, int ny, long long ld, int face, double a, double* out, int nrhs,
                         hipStream_t s);
void launch_axpby(const double* x, double* y, long long n, double a, double b, hipStream_t s);
void launch_gather(const double* G, long long ldG, const int* cell, int K, const double* v, double* out, long long ldo,
                   double a, double b, int nrhs, hipStream_t s);
void launch_scatter(double* G, long long ldG, const int* ucell, const int* k1, const int* k2, int nu, const double* w,
                    long long ldw, int nrhs, hipStream_t s);
struct SchurChild {
    const double* S;
    int           cn, voff, lboff, row0;
};
constexpr int kMaxSchurChildren = 8;  // children per launch of the Schur products
struct SchurLinesArgs {
    SchurChild    ch[kMaxSchurChildren];
    int           nch, total_rows;
    const double* v;
    long long     ldv;
    double*       lb;
    long long     ldl;
};
void launch_schur_lines(const SchurLinesArgs& a, int nrhs, hipStream_t s);
// y[off_c + i] = sum_j S_c(i, j) x[off_c + j] for every child c (S, cn and
// voff = lboff = off_c of each entry), kMaxSchurChildren children per launch
void launch_schur_gemv_all(const SchurChild* ch, int nch, const double* x, long long ldx, double* y, long long ldy,
                           int nrhs, hipStream_t s);
// partial column sums of a line-sourced sweep (nblk blocks per rhs) -> y-lines
// (half = 0 / 1: only the column-0 / column-(ny-1) sums, nx values)
void launch_colsum(const double* cp, int nblk, int nx, double* yl, long long ldy, int nrhs, hipStream_t s,
                   int half = -1);
template <typename T>
struct ColsumJobT {
    const T*  cp;
    int       nblk, nx, half;
    T*        yl;
    long long ldy;
};
template <typename T>
struct ColsumBatchT {
    int           n = 0;
    ColsumJobT<T> j[8];
};
using ColsumJob   = ColsumJobT<double>;
using ColsumBatch = ColsumBatchT<double>;
template <typename T>
void launch_colsum_batch(const ColsumBatchT<T>& b, int nrhs, hipStream_t s);
template <typename T>
struct GemvJobT {
    const T*  A;  // n x n row-major
    int       n;
    const T*  x;
    long long ldx;
    T*        y;
    long long ldy;
};
template <typename T>
struct GemvBatchT {
    int         n = 0;
    GemvJobT<T> j[8];
};
using GemvJob   = GemvJobT<double>;
using GemvBatch = GemvBatchT<double>;
template <typename T>
void launch_gemv_batch(const GemvBatchT<T>& b, int nrhs, hipStream_t s);
// mode 0: out = v - t (v == nullptr: the unit vector e_{k0 + r} of right-hand
// side r); mode 1: out = t; mode 2: out = v + w - t; t from the x rows xu and
// the column sums of the partials cp (nblk blocks of 2 nx per rhs; unused
// without y-lines)
template <typename TL>
void launch_lines_assemble(const double* v, double* out, int K, const int* kface, const int* kpos, const TL* xu,
                           int ny, const TL* cp, int nblk, int nx, int nrhs, hipStream_t s, int k0 = 0, int mode = 0,
                           const double* w = nullptr);
template <typename T>
struct LineCopyT {
    const T*  src;
    long long lds;
    int       stride, len;
    T*        dst;
    long long ldd;
    T         coef;
};
template <typename T>
struct LineCopyBatchT {
    int          n = 0;
    LineCopyT<T> c[8];
};
using LineCopy      = LineCopyT<double>;
using LineCopyBatch = LineCopyBatchT<double>;
template <typename T>
void launch_line_copy(const LineCopyBatchT<T>& b, int nrhs, hipStream_t s);
// The line exchanges of one step of the composite sweep in one launch: the
// column sums, basis-change products and line copies of the three batches
// (products go to the BLAS when nrhs reaches the GEMM threshold).
template <typename T>
void launch_exchange(const ColsumBatchT<T>& cs, const GemvBatchT<T>& gv, const LineCopyBatchT<T>& lc, int nrhs,
                     hipStream_t s);
// dst[q] = (float)src[q], q < n
void launch_to_f32(const double* src, float* dst, long long n, hipStream_t s);
void launch_lines_from_units(const SchurLinesArgs& a, int k0, int cnt, hipStream_t s);
void launch_dense_gemv(const double* A, int n, const double* x, double* y, double alpha, double beta, int nrhs,
                       hipStream_t s);
void gauss_jordan_invert(double* A, int n, hipStream_t s);
void launch_transpose(const double* in, double* out, int n, hipStream_t s);
void launch_schur_collect(const double* buf, long long ldb, const int* cell, int cn, int i0, int cnt, double* S,
                          hipStream_t s);
void launch_set_units(double* buf, long long ldb, const int* cell, int i0, int cnt, hipStream_t s);

// Device state of a batched interface GMRES (all pointers device memory).
// Segments per Arnoldi dot product (one workgroup each, summed in order by
// the update): at least 2048 entries per segment, at most kDotSeg.
constexpr int kDotSeg = 16;
inline int gm_nseg(int K) { return K <= 2048 ? 1 : (K + 2047) / 2048 < kDotSeg ? (K + 2047) / 2048 : kDotSeg; }
struct GmresDev {
    int     K, nrhs, m, nblk;  // nblk = ceil(K / 64)
    double* V;      // (m+1) x nrhs x K
    double* Wv;     // nrhs x K
    double* X;      // nrhs x K
    double* C;      // nrhs x K (right-hand side)
    double* H;      // nrhs x (m+1) x m
    double* Hacc;   // nrhs x (m+1)
    double* cs;     // nrhs x m
    double* sn;     // nrhs x m
    double* g;      // nrhs x (m+1)
    double* Y;      // nrhs x m
    double* rnorm;  // nrhs: norm of the right-hand side C at GMRES start
    double* part;   // nrhs x nblk partial sums
    double* part2;  // nrhs x (m+1) x kDotSeg: segment sums of the dot products of the current pass
    double* tolabs; // nrhs
    double* resid;  // nrhs
    int*    state;  // nrhs x 4: active, valid columns, status (1 converged, 2 numeric), steps
    int*    hstate = nullptr;  // host-mapped copies written at cycle start (or null)
    double* hres   = nullptr;  // nrhs residuals, then nrhs initial norms
};

void gm_init(const GmresDev& g, double tol, int rel, hipStream_t s);
void gm_publish(const GmresDev& g, int* hstate, double* hres, hipStream_t s);
bool gm_resid(const GmresDev& g, hipStream_t s);
// gm_init and the first gm_resid of a solve in one one-workgroup launch (the
// state published like gm_resid's); false, with nothing launched, when K is
// too large for one workgroup
bool gm_first(const GmresDev& g, double tol, int rel, hipStream_t s);
void gm_arnoldi(const GmresDev& g, int j, hipStream_t s);
void gm_finish_cycle(const GmresDev& g, hipStream_t s);

}  // namespace fp
