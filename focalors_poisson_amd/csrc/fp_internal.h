// Internal declarations of the MI355X solve path: device plans, kernel
// launchers and the error type that the C ABI turns into status codes.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <functional>
#include <vector>

#include "focalors_poisson.h"

namespace fp {

// Boundary codes (PDEBoundaryType, base/location_boundary.h:62-69) and faces
// (LocationType, base/location_boundary.h:8-16).
enum Bc : int { DIRICHLET = 0, NEUMANN = 1, PERIODIC = 2, ADJACENTED = 3, BC_NULL = 4 };
enum Face : int { XNEG = 0, XPOS = 1, YNEG = 2, YPOS = 3 };

// y-transform kinds, chosen by the (yneg, ypos) pair exactly as
// PoissonSolver2DBase::create_fft (pe/poisson/poisson_solver2d_base.cpp:64-91).
enum TKind : int { T_DD = 0, T_NN = 1, T_DN = 2, T_ND = 3, T_PP = 4 };

inline bool dir_like(int b) { return b == DIRICHLET || b == ADJACENTED; }

struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

#define FP_HIP(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            throw ::fp::Error(e_ == hipErrorOutOfMemory ? FP_E_NOMEM : FP_E_HIP,                       \
                              std::string(#call) + ": " + hipGetErrorString(e_));                      \
    } while (0)

#define FP_REQUIRE(cond, msg)                                                                          \
    do {                                                                                               \
        if (!(cond)) throw ::fp::Error(FP_E_INVALID, msg);                                             \
    } while (0)

// Owning device buffer.
template <typename T>
struct DevBuf {
    T*     p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count) FP_HIP(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void upload(const T* h, size_t count) { FP_HIP(hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice)); }
};

// The message of the last failed C ABI call on this thread (fp_last_error).
std::string& last_error_slot();

// Runs fn, turning its exceptions into an fp_status and the message into
// last_error_slot(): no exception crosses the C ABI.
template <typename F>
int guarded(F&& fn)
{
    try {
        fn();
        return FP_OK;
    } catch (const Error& e) {
        last_error_slot() = e.what();
        return e.status;
    } catch (const std::bad_alloc&) {
        last_error_slot() = "host allocation failed";
        return FP_E_NOMEM;
    } catch (const std::exception& e) {
        last_error_slot() = e.what();
        return FP_E_INVALID;
    }
}

constexpr int kMaxRadices = 24;
// subdomain problems whose descriptors travel as kernel arguments; larger
// batches put them in device memory (JobTable) and still take one launch
constexpr int kMaxBatch = 8;

// Descriptors of a batch too large for kernel arguments, in device memory:
// problem q owns workgroups start[q] .. start[q+1]-1.
template <typename P>
struct JobTable {
    int        nprob;
    const int* start;
    const P*   p;
};
// the problem of workgroup blk in a table (binary search over start)
template <typename P>
__device__ __forceinline__ const P& table_find(const JobTable<P>& t, int blk, int& local)
{
    int lo = 0, hi = t.nprob - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (blk >= t.start[mid]) lo = mid;
        else hi = mid - 1;
    }
    local = blk - t.start[lo];
    return t.p[lo];
}
// Pinned host staging (fp_staging.cpp): a block of at least n bytes, free
// for the caller to fill; staging_put queues its reuse after the work queued
// on s so far (call it right after queuing the copy that reads the block).
void* staging_get(size_t n);
void  staging_put(void* p, hipStream_t s);

// A caller-provided region of device memory that the descriptor tables of
// one call are carved from instead of being allocated (the entry point that
// allocates nothing sets it for the duration of its call, on its thread).
struct TableArena {
    unsigned char* base = nullptr;  // device address
    size_t         cap  = 0;
    size_t         used = 0;
};
TableArena*& table_arena();  // this thread's arena, null when none (fp_staging.cpp)

// Copies n descriptors and their workgroup offsets into stream-ordered device
// memory (released with table_release after the launch on the same stream),
// through pinned staging so that the host does not wait for the device.
// With an arena set: into the arena, by a copy from pageable host memory
// (no device or pinned allocation; the runtime reads the source before the
// call returns, so the host waits for the work queued before it).
// Not for a stream under graph capture: the copy reads host memory.
template <typename P>
JobTable<P> table_upload(const std::vector<P>& p, const std::vector<int>& start, hipStream_t s, void*& mem)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    FP_HIP(hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone)
        throw Error(FP_E_INVALID, "batches of more than 8 subdomains cannot be captured");
    const size_t pb = (p.size() * sizeof(P) + 255) & ~(size_t)255;
    const size_t nb = pb + start.size() * sizeof(int);
    if (TableArena* ar = table_arena()) {
        const size_t off = (ar->used + 255) & ~(size_t)255;
        if (off + nb > ar->cap)
            throw Error(FP_E_INVALID, "descriptor tables exceed the scratch of the no-allocation entry point");
        std::vector<unsigned char> h(nb);
        std::memcpy(h.data(), p.data(), p.size() * sizeof(P));
        std::memcpy(h.data() + pb, start.data(), start.size() * sizeof(int));
        FP_HIP(hipMemcpyAsync(ar->base + off, h.data(), nb, hipMemcpyHostToDevice, s));
        ar->used = off + nb;
        mem      = nullptr;  // nothing to release
        JobTable<P> t;
        t.nprob = (int)p.size();
        t.p     = reinterpret_cast<const P*>(ar->base + off);
        t.start = reinterpret_cast<const int*>(ar->base + off + pb);
        return t;
    }
    FP_HIP(hipMallocAsync(&mem, nb, s));
    unsigned char* h = static_cast<unsigned char*>(staging_get(nb));
    std::memcpy(h, p.data(), p.size() * sizeof(P));
    std::memcpy(h + pb, start.data(), start.size() * sizeof(int));
    const hipError_t e = hipMemcpyAsync(mem, h, nb, hipMemcpyHostToDevice, s);
    staging_put(h, s);
    FP_HIP(e);
    JobTable<P> t;
    t.nprob = (int)p.size();
    t.p     = static_cast<const P*>(mem);
    t.start = reinterpret_cast<const int*>(static_cast<unsigned char*>(mem) + pb);
    return t;
}
inline void table_release(void* mem, hipStream_t s)
{
    if (mem) FP_HIP(hipFreeAsync(mem, s));
}

// A batched y transform of one subdomain size and boundary pair: the packed
// complex FFT of length L (two field rows per FFT) with its radix schedule
// and tables.  T is the arithmetic type.
template <typename T>
struct TransformPlan {
    int        kind = T_DD;
    int        N    = 0;  // transform length = ny
    int        L    = 0;  // complex DFT length
    int        F    = 0;  // FFT length: L, or the Bluestein length (prime factors of L above 47)
    int        nrad = 0;
    int        rad[kMaxRadices] = {};
    int        engine = 1;  // 1: single-buffer FFT in LDS, 0: two LDS buffers, 2: two buffers in device memory
                            // (lengths beyond the LDS of one workgroup)
    int        ncs  = 0;  // cos/sin entries of the odd-prime passes (VALU form)
    DevBuf<double> frag;  // FP64 odd-prime passes: matrix-core operand fragments (matrix-core form)
    DevBuf<T>  W;         // twiddles: 64 of W^m, then L/64+1 of W^(64m), W = exp(-2 pi i / L), then
                          // (cos, sin)(2 pi m / R) of every odd-prime pass R
    DevBuf<T>  aux;       // 2N: NN phase exp(-i pi m / 2N) or ND (sin, cos) of 2N pi (j+1)/(2N+1)
    DevBuf<T>  blu;       // Bluestein: chirp (L), then the DFT of its conjugate / F (F)
    DevBuf<T>  gscr;      // engine 2: 256 workgroup slots of two F-value complex buffers; the plan's
                          // launches must be ordered (one stream at a time, as its handle)
    T          c0 = 0, c1 = 0;
    int        nt = 256;  // threads per workgroup (tuned)
    int        odd = 0;   // odd-prime pass form: 0 none, 1 matrix cores (FP64), 2 / 3 VALU pair sums with 5 / 3 pairs per item (tuned)
    int        nt1 = 0, odd1 = -1;  // the same for launches of one field (tuned; 0 / -1: nt, odd)
    int        ept = 16;  // values per thread of the register butterflies (24 for long Bluestein lengths)
    void build(int kind, int N);
};
// Times the plan's transform on an nx-row field and keeps the fastest
// workgroup size (measure = false: the estimated shape, no timing).
template <typename T>
void tune_transform(TransformPlan<T>& tp, int nx, bool measure);

// Per-subdomain plan: PoissonSolver2D's precomputed state on the device.
struct SubPlan {
    int    nx = 0, ny = 0;
    double hx = 0, hy = 0;
    int    bc[4] = {0, 0, 0, 0};
    int    kind = T_DD;
    bool   no_dirichlet = false;
    std::vector<double>   x_diag;   // host copy, reference order: -2 - delta
    std::vector<double>   delta;    // per mode: -(x_diag + 2) = (hx/hy)^2 4 sin^2(.) >= 0, to full precision
    TransformPlan<double> t64;
    TransformPlan<float>  t32;
    DevBuf<double>        xd;       // delta on the device
    DevBuf<double>        pst;      // ny*tri_p: pivot excess e = 1 + 1/pivot before each row chunk (mode-major)
    int                   tri_ch = 16, tri_p = 0;        // rows per chunk of the x sweep, chunks per mode
    int                   mb_field = 4, mb_lines = 4;    // modes per workgroup of the two sweeps
    // the field sweep's shape in batches of many problems (the throughput
    // regime: longer chunks, more modes per workgroup, whole cache lines per
    // row segment), and its chunk-start table when the chunk length differs
    int                   tri_ch_b = 0, tri_p_b = 0, mb_b = 0;
    DevBuf<double>        pst_b;
    // periodic x: per mode A_D^-1 e_0 then A_D^-1 e_{nx-1} (2 nx ny, field layout)
    bool                  cyc = false;
    DevBuf<double>        zc;
    int                   dev = 0;
    bool                  measure = true;  // time the launch shapes (else the estimated ones)
    void build(int nx, int ny, double hx, double hy, const int bc[4]);
    void init_sweep(const std::vector<double>& delta);
    void ensure_f32();
    void set_chunks(int ch);
    void set_batch_chunks();
    void tune_tridiag(bool measure);
};

// 2 - lambda_i of the y (or z) transform kind, i = 1..n, to full precision
std::vector<long double> eigen_gaps(int kind, int n);

// Choices of the autotuners by key (fp_tune.cpp): process-wide, and shared
// across processes through the file FP_TUNE_FILE names, if any.
bool tune_lookup(const std::string& key, std::vector<int>& vals);
void tune_store(const std::string& key, const std::vector<int>& vals);

// ---- kernel launchers (fp_transform.hip / fp_tridiag.hip) ----
// y transform of `nrhs` fields of one subdomain, in place.  inverse = 0: the
// reference's transform; 1: transform_transpose.
template <typename T>
void launch_ytransform(const TransformPlan<T>& tp, T* data, long long ld, int nx, int nrhs, int inverse,
                       hipStream_t s);
// Tridiagonal solves along x for every mode, in place.
template <typename T>
void launch_tridiag(const SubPlan& sp, T* data, long long ld, int nrhs, hipStream_t s);

// Batched forms: the jobs of several subdomains (any sizes and boundary
// pairs) in as few launches as their kernels allow.
// data/ld: the output (and, with src == null, the input: in place).  A
// separate src/lds is read instead when given; src64/dst64 mark buffers of
// doubles (FP32 transforms reading or writing FP64 fields).
template <typename T>
struct TransformJob {
    const TransformPlan<T>* plan;
    void*                   data;
    long long               ld;
    int                     nx, nrhs;
    const void*             src   = nullptr;
    long long               lds   = 0;
    int                     src64 = 0, dst64 = 0;
    // forward transforms of raw right-hand sides: boundary assembly and the
    // hx^2 scaling fused into the load (ghost data g[face], nrhs arrays back
    // to back, times ga[face]; null faces take none)
    int                     asmb  = 0;
    double                  hx2   = 1.0;
    const double*           g[4]  = {nullptr, nullptr, nullptr, nullptr};
    double                  ga[4] = {0.0, 0.0, 0.0, 0.0};
    int                     nt    = 0;   // threads per workgroup (0: the plan's)
    int                     odd   = -1;  // odd-prime pass form (-1: the plan's)
};
template <typename T>
void launch_ytransform_batch(const TransformJob<T>* jobs, int njobs, int inverse, hipStream_t s);
// One launch shape {threads per workgroup, odd-prime form} for a batch
// launched together ({0, -1}: every job its own), chosen by timing the
// forward batch; the jobs' fields are overwritten.
template <typename T>
std::array<int, 2> tune_batch_shape(std::vector<TransformJob<T>> jobs, hipStream_t s);
// Whether a batch launch shape {threads per workgroup (0: the plan's own),
// odd-prime form (-1: the plan's own)} runs plan tp correctly; and the job's
// shape set to it where it does, else to the plan's own.
template <typename T>
bool batch_shape_fits(const TransformPlan<T>& tp, int nt, int form);
template <typename T>
void apply_batch_shape(TransformJob<T>& j, const std::array<int, 2>& shape);
template <typename T>
struct TridiagJob {
    const SubPlan* sp;
    T*             data;
    long long      ld;
    int            nrhs;
    int            thru = 0;  // 1: part of a batch of many problems (the plan's batch shape)
};
template <typename T>
void launch_tridiag_batch(const TridiagJob<T>* jobs, int njobs, hipStream_t s);

// A launch prepared once and replayed: its shape and its descriptor table
// (nprob descriptors from offset 0, then nprob + 1 workgroup offsets from
// start_off), which the owner copies to device memory; launching then reads
// only that copy, so it queues a kernel and nothing else (graph capture
// allowed).  Within a launch the problems are ordered by decreasing work per
// workgroup, so that the longest workgroups are dispatched first.
struct GroupPlan {
    int                        engine = 1, nt = 0, odd = 0, ept = 0;  // transforms
    int                        ch = 0;                               // sweeps: rows per chunk
    int                        blocks = 0, threads = 0, nprob = 0;
    size_t                     lds = 0, start_off = 0;
    std::vector<unsigned char> table;
};
template <typename T>
std::vector<GroupPlan> plan_ytransform_groups(const TransformJob<T>* jobs, int njobs, int inverse);
template <typename T>
void launch_ytransform_group(const GroupPlan& g, const unsigned char* dtable, hipStream_t s);
template <typename T>
std::vector<GroupPlan> plan_tridiag_groups(const TridiagJob<T>* jobs, int njobs);
template <typename T>
void launch_tridiag_group(const GroupPlan& g, const unsigned char* dtable, hipStream_t s);

// Arguments of the line-sourced tridiagonal sweep (interface operator).
template <typename T>
struct LinesArgs {
    const T*  lines;    // per rhs: Xh0 (ny) | Xh1 (ny) | a (nx) | b (nx)
    long long ldl;      // 2 ny + 2 nx
    const T*  phi;      // phi0 (ny) | phi1 (ny)
    const double* xd;   // filled by the launcher from the plan: delta,
    const double* pst;  // chunk-start pivot excesses,
    int       ns, ne;   // Neumann x ends
    T*        xu;       // per rhs: coefficient rows 0 and nx-1 (2 ny)
    T*        colpart;  // per rhs, per workgroup: 2 x nx column partial sums
    int       nx, ny;
    int       xrows, ycols;
};
template <typename T>
int launch_tridiag_lines(const SubPlan& sp, const LinesArgs<T>& a, int nrhs, hipStream_t s);

// General sweep of the composite solve (fp_tridiag.hip, k_sweep): right-hand
// side = (fin ? f : 0) + lines(lbp) - lines(lbm), line buffers packed per
// right-hand side as [x0 (ny) | x1 (ny) | y0 (nx) | y1 (nx)] (x-lines as
// transform coefficients, y-lines physical); outputs: the field (fout, in
// place or to fo), solved rows 0 / nx-1 (xo0 / xo1, ldxo apart per right-hand side)
// and per-(rhs, mode block) partial column sums cp (2 nx each; the blocks per
// right-hand side are sweep_blocks(sp)).
template <typename T>
struct SweepJob {
    const SubPlan* sp;
    int            nrhs;
    T*             f     = nullptr;
    long long      ld    = 0;
    const T*       lbp   = nullptr;
    const T*       lbm   = nullptr;
    long long      ldl   = 0;
    const T*       phi   = nullptr;
    T*             xo0   = nullptr;
    T*             xo1   = nullptr;
    long long      ldxo  = 0;
    T*             cp    = nullptr;
    int            fin   = 1, fout = 1;
    T*             fo    = nullptr;  // field output (fout) elsewhere than f, same ld
};
template <typename T>
void launch_sweep_batch(const SweepJob<T>* jobs, int njobs, hipStream_t s);
int sweep_blocks(const SubPlan& sp);
template <typename T>
int lines_workgroups(const SubPlan& sp);

// The geometry of a composite handle (faces named in a connection are
// ADJACENTED in bc; adj[4 d + face] is the neighbour or -1).
struct CatView {
    int           nd;
    const int*    nx;
    const int*    ny;
    const double* hx;
    const double* hy;
    const int*    bc;
    const int*    adj;
    int           device;
};
CatView cat_view(const fp_cat_s* c);
// The solve recorded into a HIP graph with work before and after it: into
// the handle's captured-solve slot (proj = 0) or its projection slot for the
// projection with id proj (> 0); cat_replay replays a slot (the projection
// slot only for the projection that recorded it).
void    cat_capture(fp_cat_s* cat, const double* const* d_rhs, double* const* d_fields, const double* const* d_ghost,
                    int nrhs, hipStream_t s, const std::function<void(hipStream_t)>& pre,
                    const std::function<void(hipStream_t)>& post, unsigned long long proj);
int     cat_replay(fp_cat_s* cat, unsigned long long proj, void* stream, fp_solve_info* info);

// Full subdomain solve of nrhs fields (transform, sweep, inverse transform).
template <typename T>
void sub_solve(SubPlan& sp, T* data, long long ld, int nrhs, hipStream_t s);
// Several subdomain solves as one batch per phase: problem k solves nrhs[k]
// fields of plans[k] at data[k] (ld[k] apart).
template <typename T>
void sub_solve_batch(SubPlan* const* plans, T* const* data, const long long* ld, const int* nrhs, int n,
                     hipStream_t s, bool thru = false);

}  // namespace fp

// A context: one HIP device (fp_ctx_create).
struct fp_ctx_s {
    int            device  = 0;
    unsigned char* scratch = nullptr;  // this device's copy of the no-allocation scratch array
    bool           measure = true;     // plans built from this context time their launch shapes
};

// A subdomain solver (fp_sub_create).
struct fp_sub_s {
    fp_ctx_s*   ctx = nullptr;
    fp::SubPlan plan;
};
