/*
 * fp_oracle.c -- CPU FP64 restatement of the Focalors_Poisson 2-D solve path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the timed CPU
 * baseline.  Never part of the product library.
 *
 * The reference computes its y transforms with FFTW r2c plans; FFTW is not in
 * this image, so the r2c here is this file's own mixed-radix Stockham FFT
 * (Bluestein for lengths with a prime factor above 64).  Everything above the
 * r2c restates the reference step by step (file:line at each function).
 * OpenMP loops follow the reference's OPENMP_PARALLEL_FOR placement
 * (base/parallel/omp/enable_openmp.h:3-10): rows of a transform, modes of a
 * Thomas sweep, rows of an elementwise field update.
 */
#include "fp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_PI 3.14159265358979323846

/* ------------------------------------------------------------------------ */
/* FFT                                                                       */
/* ------------------------------------------------------------------------ */

typedef struct or_plan {
    int     n;
    int     nf;
    int     fac[40];
    double* w;          /* w[k] = exp(-2 pi i k / n), k < n, interleaved */
    int     bluestein;  /* 1: whole length done through a power-of-two convolution */
    int     M;
    double* chirp;      /* exp(-i pi k^2 / n), k < n */
    double* bhat;       /* FFT_M of the conjugate chirp sequence */
    struct or_plan* sub;/* power-of-two plan of length M */
} or_plan;

static void cexp_tab(double* w, int n)
{
    for (int k = 0; k < n; k++) {
        long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
        w[2 * k]     = (double)cosl(a);
        w[2 * k + 1] = (double)sinl(a);
    }
}

static int largest_prime_factor(int n)
{
    int lp = 1;
    for (int p = 2; (long)p * p <= n; p++)
        while (n % p == 0) { lp = p; n /= p; }
    if (n > 1) lp = n > lp ? n : lp;
    return lp;
}

static or_plan* plan_create(int n);

static void plan_factor(or_plan* p)
{
    int n = p->n;
    p->nf = 0;
    while (n % 4 == 0) { p->fac[p->nf++] = 4; n /= 4; }
    while (n % 2 == 0) { p->fac[p->nf++] = 2; n /= 2; }
    for (int r = 3; n > 1; r += 2)
        while (n % r == 0) { p->fac[p->nf++] = r; n /= r; }
}

static or_plan* plan_create(int n)
{
    or_plan* p = (or_plan*)calloc(1, sizeof(or_plan));
    p->n = n;
    p->w = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    cexp_tab(p->w, n);
    if (n > 64 && largest_prime_factor(n) > 64) {
        int M = 1;
        while (M < 2 * n - 1) M <<= 1;
        p->bluestein = 1;
        p->M         = M;
        p->sub       = plan_create(M);
        p->chirp     = (double*)malloc(sizeof(double) * 2 * (size_t)n);
        for (int k = 0; k < n; k++) {
            /* k^2 mod 2n keeps the angle argument small and exact */
            long long q = ((long long)k * k) % (2LL * n);
            long double a = -3.14159265358979323846264338327950288L * (long double)q / (long double)n;
            p->chirp[2 * k]     = (double)cosl(a);
            p->chirp[2 * k + 1] = (double)sinl(a);
        }
        double* b = (double*)calloc(2 * (size_t)M, sizeof(double));
        for (int k = 0; k < n; k++) {
            b[2 * k]     = p->chirp[2 * k];
            b[2 * k + 1] = -p->chirp[2 * k + 1];
            if (k > 0) {
                b[2 * (M - k)]     = p->chirp[2 * k];
                b[2 * (M - k) + 1] = -p->chirp[2 * k + 1];
            }
        }
        p->bhat = (double*)malloc(sizeof(double) * 2 * (size_t)M);
        or_dft(M, -1, b, p->bhat);
        free(b);
    } else {
        plan_factor(p);
    }
    return p;
}

/* Plan cache: plans are created under a lock and never freed (like FFTW wisdom).
 * The table grows without bound, so every length keeps exactly one plan for the
 * life of the process, however many lengths a test session visits. */
static or_plan** g_plans  = NULL;
static int       g_nplans = 0, g_cap = 0;

static void plan_free(or_plan* p)
{
    if (!p) return;
    plan_free(p->sub);
    free(p->w); free(p->chirp); free(p->bhat); free(p);
}

static or_plan* plan_get(int n)
{
    or_plan* found = NULL;
#pragma omp critical(or_plan_cache)
    {
        for (int i = 0; i < g_nplans; i++)
            if (g_plans[i]->n == n) { found = g_plans[i]; break; }
    }
    if (found) return found;
    /* built outside the lock: a Bluestein plan builds its own sub-plan */
    or_plan* fresh = plan_create(n);
#pragma omp critical(or_plan_cache)
    {
        for (int i = 0; i < g_nplans; i++)
            if (g_plans[i]->n == n) { found = g_plans[i]; break; }
        if (!found) {
            found = fresh;
            fresh = NULL;
            if (g_nplans == g_cap) {
                int       cap = g_cap ? 2 * g_cap : 256;
                or_plan** t   = (or_plan**)realloc(g_plans, sizeof(or_plan*) * (size_t)cap);
                if (!t) abort();
                g_plans = t;
                g_cap   = cap;
            }
            g_plans[g_nplans++] = found;
        }
    }
    if (fresh) plan_free(fresh); /* another thread won the race; keep its plan */
    return found;
}

int or_plan_count(void)
{
    int n;
#pragma omp critical(or_plan_cache)
    n = g_nplans;
    return n;
}

/* One Stockham pass of radix R over n points: combines R sub-transforms of
 * length Ns into transforms of length Ns*R. */
static void stockham_pass(const or_plan* p, int R, int Ns, int sign, const double* a, double* b)
{
    const int n = p->n, nR = n / R, tstep = n / (Ns * R), rstep = n / R;
    double    v[2 * 64], o[2 * 64];
    for (int j = 0; j < nR; j++) {
        const int jm = j % Ns;
        if (R <= 64) {
            for (int r = 0; r < R; r++) {
                double xr = a[2 * (j + r * nR)], xi = a[2 * (j + r * nR) + 1];
                int    t  = (int)(((long long)r * jm * tstep) % n);
                double wr = p->w[2 * t], wi = -sign * p->w[2 * t + 1];
                v[2 * r]     = xr * wr - xi * wi;
                v[2 * r + 1] = xr * wi + xi * wr;
            }
            if (R == 2) {
                o[0] = v[0] + v[2]; o[1] = v[1] + v[3];
                o[2] = v[0] - v[2]; o[3] = v[1] - v[3];
            } else if (R == 4) {
                double ar = v[0] + v[4], ai = v[1] + v[5], br = v[0] - v[4], bi = v[1] - v[5];
                double cr = v[2] + v[6], ci = v[3] + v[7], dr = v[2] - v[6], di = v[3] - v[7];
                /* multiply d by -i*sign... forward (sign=-1): d*(-i) */
                double er = sign < 0 ? di : -di, ei = sign < 0 ? -dr : dr;
                o[0] = ar + cr; o[1] = ai + ci;
                o[2] = br + er; o[3] = bi + ei;
                o[4] = ar - cr; o[5] = ai - ci;
                o[6] = br - er; o[7] = bi - ei;
            } else {
                for (int q = 0; q < R; q++) {
                    double sr = 0, si = 0;
                    for (int r = 0; r < R; r++) {
                        int    t  = (int)(((long long)q * r % R) * rstep);
                        double wr = p->w[2 * t], wi = -sign * p->w[2 * t + 1];
                        sr += v[2 * r] * wr - v[2 * r + 1] * wi;
                        si += v[2 * r] * wi + v[2 * r + 1] * wr;
                    }
                    o[2 * q] = sr; o[2 * q + 1] = si;
                }
            }
            const int idxD = (j / Ns) * Ns * R + jm;
            for (int r = 0; r < R; r++) {
                b[2 * (idxD + r * Ns)]     = o[2 * r];
                b[2 * (idxD + r * Ns) + 1] = o[2 * r + 1];
            }
        }
    }
}

static void dft_exec(const or_plan* p, int sign, const double* in, double* out, double* work);

/* Bluestein: X_k = c_k * sum_n (x_n c_n) conj(c)_{k-n}, c_k = exp(-i pi k^2/n). */
static void bluestein_exec(const or_plan* p, int sign, const double* in, double* out)
{
    const int n = p->n, M = p->M;
    double*   a = (double*)calloc(2 * (size_t)M, sizeof(double));
    double*   A = (double*)malloc(sizeof(double) * 2 * (size_t)M);
    double*   wk = (double*)malloc(sizeof(double) * 2 * (size_t)M);
    for (int k = 0; k < n; k++) {
        double cr = p->chirp[2 * k], ci = -sign * p->chirp[2 * k + 1];
        double xr = in[2 * k], xi = in[2 * k + 1];
        a[2 * k]     = xr * cr - xi * ci;
        a[2 * k + 1] = xr * ci + xi * cr;
    }
    dft_exec(p->sub, -1, a, A, wk);
    /* conv kernel spectrum: for sign=+1 the chirp is conjugated, whose kernel
     * spectrum is the conjugate-reversal of bhat: B'(k) = conj(bhat(-k)). */
    for (int k = 0; k < M; k++) {
        double br, bi;
        if (sign < 0) { br = p->bhat[2 * k]; bi = p->bhat[2 * k + 1]; }
        else { int kk = (M - k) % M; br = p->bhat[2 * kk]; bi = -p->bhat[2 * kk + 1]; }
        double ar = A[2 * k], ai = A[2 * k + 1];
        A[2 * k]     = ar * br - ai * bi;
        A[2 * k + 1] = ar * bi + ai * br;
    }
    dft_exec(p->sub, +1, A, a, wk);
    const double s = 1.0 / M;
    for (int k = 0; k < n; k++) {
        double cr = p->chirp[2 * k], ci = -sign * p->chirp[2 * k + 1];
        double yr = a[2 * k] * s, yi = a[2 * k + 1] * s;
        out[2 * k]     = yr * cr - yi * ci;
        out[2 * k + 1] = yr * ci + yi * cr;
    }
    free(a); free(A); free(wk);
}

static void dft_exec(const or_plan* p, int sign, const double* in, double* out, double* work)
{
    const int n = p->n;
    if (p->bluestein) { bluestein_exec(p, sign, in, out); return; }
    if (p->nf == 0) { out[0] = in[0]; out[1] = in[1]; return; }
    /* ping-pong between out and work so that the last pass lands in out */
    const double* src = in;
    int           Ns  = 1;
    for (int f = 0; f < p->nf; f++) {
        double* dst = (p->nf - 1 - f) % 2 == 0 ? out : work;
        stockham_pass(p, p->fac[f], Ns, sign, src, dst);
        Ns *= p->fac[f];
        src = dst;
    }
    (void)n;
}

int or_dft(int n, int sign, const double* in_ri, double* out_ri)
{
    if (n <= 0) return 1;
    or_plan* p    = plan_get(n);
    double*  work = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    double*  tmp  = NULL;
    if (in_ri == out_ri) {
        tmp = (double*)malloc(sizeof(double) * 2 * (size_t)n);
        memcpy(tmp, in_ri, sizeof(double) * 2 * (size_t)n);
        in_ri = tmp;
    }
    dft_exec(p, sign, in_ri, out_ri, work);
    free(work); free(tmp);
    return 0;
}

/* r2c of length n into n/2+1 bins.  Even n: half-length complex FFT + split;
 * odd n: full complex FFT.  Thread-safe; scratch is per call. */
static void r2c_exec(int n, const double* x, double* X, double* s1, double* s2, double* s3)
{
    if (n % 2 == 0 && n >= 4) {
        const int h = n / 2;
        or_plan*  p = plan_get(h);
        for (int k = 0; k < h; k++) { s1[2 * k] = x[2 * k]; s1[2 * k + 1] = x[2 * k + 1]; }
        dft_exec(p, -1, s1, s2, s3);
        or_plan* pn = plan_get(n);
        for (int k = 0; k <= h; k++) {
            int    k1 = k % h, k2 = (h - k) % h;
            double zr = s2[2 * k1], zi = s2[2 * k1 + 1];
            double cr = s2[2 * k2], ci = -s2[2 * k2 + 1];
            double er = 0.5 * (zr + cr), ei = 0.5 * (zi + ci);        /* even part */
            double orr = 0.5 * (zi - ci), oi = -0.5 * (zr - cr);     /* odd part = (z - conj)/(2i) */
            int    t   = k % n;
            double wr = pn->w[2 * t], wi = pn->w[2 * t + 1];
            X[2 * k]     = er + orr * wr - oi * wi;
            X[2 * k + 1] = ei + orr * wi + oi * wr;
        }
    } else {
        or_plan* p = plan_get(n);
        for (int k = 0; k < n; k++) { s1[2 * k] = x[k]; s1[2 * k + 1] = 0.0; }
        dft_exec(p, -1, s1, s2, s3);
        for (int k = 0; k <= n / 2; k++) { X[2 * k] = s2[2 * k]; X[2 * k + 1] = s2[2 * k + 1]; }
    }
}

int or_r2c(int n, const double* in, double* out_ri)
{
    if (n <= 0) return 1;
    double* s = (double*)malloc(sizeof(double) * 6 * (size_t)(n + 2));
    r2c_exec(n, in, out_ri, s, s + 2 * (n + 2), s + 4 * (n + 2));
    free(s);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* y transforms (pe/poisson/poisson_fft2d.cpp)                                */
/* ------------------------------------------------------------------------ */

static int is_dir_like(int t) { return t == OR_DIRICHLET || t == OR_ADJACENTED; }

enum { T_DD = 0, T_NN = 1, T_DN = 2, T_ND = 3, T_PP = 4, T_BAD = -1 };

/* PoissonSolver2DBase::create_fft, poisson_solver2d_base.cpp:64-91 */
static int transform_kind(int neg, int pos)
{
    if (neg == OR_PERIODIC && pos == OR_PERIODIC) return T_PP;
    if (neg == OR_NEUMANN && pos == OR_NEUMANN) return T_NN;
    if (is_dir_like(neg) && is_dir_like(pos)) return T_DD;
    if (is_dir_like(neg) && pos == OR_NEUMANN) return T_DN;
    if (neg == OR_NEUMANN && is_dir_like(pos)) return T_ND;
    return T_BAD;
}

/* One row of each transform.  ex is zero-initialised scratch of the r2c length,
 * X holds the bins, s is r2c scratch. */
static void row_transform(int kind, int inverse, int ny, const double* in, double* out, double* ex,
                          double* X, double* s)
{
    double* s1 = s;
    double* s2 = s + 2 * (2 * ny + 4);
    double* s3 = s + 4 * (2 * ny + 4);
    if (kind == T_DD) {
        /* poisson_fft2d.cpp:37-57: ex = [0, f, 0 ...] of length 2ny+2;
         * f_hat(j) = -sqrt(2/(ny+1)) Im X[j+1]; the transpose is the same map. */
        const int    L = 2 * ny + 2;
        const double c = sqrt(2.0 / (1.0 + ny));
        memset(ex, 0, sizeof(double) * L);
        for (int j = 0; j < ny; j++) ex[j + 1] = in[j];
        r2c_exec(L, ex, X, s1, s2, s3);
        for (int j = 0; j < ny; j++) out[j] = -c * X[2 * (j + 1) + 1];
    } else if (kind == T_NN) {
        const int    L    = 2 * ny;
        const double sq2n = sqrt(2.0 / ny), sqn = sqrt(1.0 / ny);
        memset(ex, 0, sizeof(double) * L);
        if (!inverse) {
            /* poisson_fft2d.cpp:105-118 */
            for (int j = 0; j < ny; j++) ex[j + 1] = (j % 2 == 0 ? -1.0 : 1.0) * in[j];
            r2c_exec(L, ex, X, s1, s2, s3);
            for (int j = 0; j < ny - 1; j++) {
                double cj = cos(OR_PI * (j + 1.0) / ny / 2), sj = sin(OR_PI * (j + 1.0) / ny / 2);
                out[j] = sq2n * (-1.0 * cj * X[2 * (j + 1) + 1] - sj * X[2 * (j + 1)]);
            }
            double acc = 0.0;
            for (int j = 0; j < ny; j++) acc += sqn * in[j];
            out[ny - 1] = acc;
        } else {
            /* poisson_fft2d.cpp:120-133 */
            for (int j = 0; j < ny; j++) ex[j + 1] = cos(OR_PI * (j + 1.0) / ny / 2) * in[j];
            r2c_exec(L, ex, X, s1, s2, s3);
            for (int j = 0; j < ny; j++) out[j] = (j % 2 == 0 ? 1.0 : -1.0) * sq2n * X[2 * (j + 1) + 1];
            for (int j = 0; j < ny; j++) ex[j + 1] = sin(OR_PI * (j + 1.0) / ny / 2) * in[j];
            r2c_exec(L, ex, X, s1, s2, s3);
            for (int j = 0; j < ny; j++) out[j] += (j % 2 == 0 ? 1.0 : -1.0) * sq2n * X[2 * (j + 1)];
            for (int j = 0; j < ny; j++) out[j] += (sq2n + sqn) * in[ny - 1];
        }
    } else if (kind == T_ND) {
        const int    L   = 2 * ny + 1;
        const double sq4 = sqrt(4.0 / (2.0 * ny + 1.0));
        memset(ex, 0, sizeof(double) * L);
        if (!inverse) {
            /* poisson_fft2d.cpp:180-189 */
            for (int j = 0; j < ny; j++) ex[j + 1] = (j % 2 == 0 ? -1.0 : 1.0) * sq4 * in[j];
            r2c_exec(L, ex, X, s1, s2, s3);
            for (int j = 0; j < ny; j++) {
                double a  = 2.0 * ny * OR_PI * (j + 1.0) / (2.0 * ny + 1.0);
                out[j]    = sin(a) * X[2 * (j + 1)] - cos(a) * X[2 * (j + 1) + 1];
            }
        } else {
            /* poisson_fft2d.cpp:191-205 */
            for (int j = 0; j < ny; j++) {
                double a  = 2.0 * ny * OR_PI * (j + 1.0) / (2.0 * ny + 1.0);
                ex[j + 1] = sq4 * sin(a) * in[j];
            }
            r2c_exec(L, ex, X, s1, s2, s3);
            for (int j = 0; j < ny; j++) out[j] = X[2 * (j + 1)];
            for (int j = 0; j < ny; j++) {
                double a  = 2.0 * ny * OR_PI * (j + 1.0) / (2.0 * ny + 1.0);
                ex[j + 1] = sq4 * cos(a) * in[j];
            }
            r2c_exec(L, ex, X, s1, s2, s3);
            for (int j = 0; j < ny; j++) out[j] -= X[2 * (j + 1) + 1];
            for (int j = 0; j < ny; j++) out[j] = (j % 2 == 0 ? -1.0 : 1.0) * out[j];
        }
    } else if (kind == T_PP) {
        /* even ny; r2c of length ny.  ex is 4 ny zeroed values, X 4 bin sets */
        const int    L  = ny, H = ny / 2;
        const double sq = sqrt(2.0 / ny);
        double*      e0 = ex;
        double*      e1 = ex + ny;
        double*      e2 = ex + 2 * ny;
        double*      e3 = ex + 3 * ny;
        double*      X0 = X;
        double*      X1 = X + 2 * (H + 1);
        double*      X2 = X + 4 * (H + 1);
        double*      X3 = X + 6 * (H + 1);
        memset(ex, 0, sizeof(double) * 4 * (size_t)ny);
        if (!inverse) {
            /* poisson_fft2d.cpp:367-403: r2c of [0, f_0 .. f_{N-2}] and of the
             * same with alternating signs; ts / tc from their bins */
            for (int j = 0; j < ny - 1; j++) e0[j + 1] = in[j];
            for (int j = 0; j < ny - 1; j++) e1[j + 1] = (j % 2 == 0 ? 1.0 : -1.0) * in[j];
            r2c_exec(L, e0, X0, s1, s2, s3);
            r2c_exec(L, e1, X1, s1, s2, s3);
            for (int m = 0; m < ny; m++) {
                double ts = 0.0, tc;
                if (m > 0) ts = -X0[2 * (H - (m + 1) / 2) + 1];  /* m = 2j+1 or 2j+2: bin N/2-j-1 */
                tc = in[ny - 1] - X1[2 * ((m + 1) / 2)];         /* m = 2j: bin j; m = 2j+1: bin j+1 */
                out[m] = sq * (cos(OR_PI * (2.0 * m + 1.0) / 4.0) * ts + sin(OR_PI * (2.0 * m + 1.0) / 4.0) * tc);
            }
        } else {
            /* poisson_fft2d.cpp:406-465: four r2c of even/odd, cosine/sine
             * weighted coefficients */
            for (int j = 0; j < H; j++) {
                e0[j]         = cos(OR_PI * (4.0 * j + 1.0) / 4.0) * in[2 * j];
                e1[j + 1]     = cos(OR_PI * (4.0 * j + 3.0) / 4.0) * in[2 * j + 1];
                e2[H - j]     = sin(OR_PI * (4.0 * j + 1.0) / 4.0) * in[2 * j];
                e3[H - j - 1] = sin(OR_PI * (4.0 * j + 3.0) / 4.0) * in[2 * j + 1];
            }
            r2c_exec(L, e0, X0, s1, s2, s3);
            r2c_exec(L, e1, X1, s1, s2, s3);
            r2c_exec(L, e2, X2, s1, s2, s3);
            r2c_exec(L, e3, X3, s1, s2, s3);
            double* fc = (double*)calloc((size_t)ny, sizeof(double));
            double* fs = (double*)calloc((size_t)ny, sizeof(double));
            double  sg = -1.0;
            for (int j = 0; j < H; j++) {
                fc[j] = sg * (X0[2 * (j + 1) + 1] + X1[2 * (j + 1) + 1]);
                fs[j] = X2[2 * (j + 1)] + X3[2 * (j + 1)];
                sg    = -sg;
            }
            for (int j = 0; j < H - 1; j++) {
                fc[H + j] = -fc[H - 2 - j];
                fs[H + j] = fs[H - 2 - j];
            }
            fs[ny - 1] = X2[0] + X3[0];
            for (int j = 0; j < ny; j++) out[j] = sq * (fc[j] + fs[j]);
            free(fc);
            free(fs);
        }
    } else if (kind == T_DN) {
        /* poisson_fft2d.cpp:242-262: the same map both ways */
        const int    L   = 2 * ny + 1;
        const double sq4 = sqrt(4.0 / (2.0 * ny + 1.0));
        memset(ex, 0, sizeof(double) * L);
        for (int j = 0; j < ny; j++) ex[j + 1] = (j % 2 == 0 ? -1.0 : 1.0) * sq4 * in[j];
        r2c_exec(L, ex, X, s1, s2, s3);
        for (int j = 0; j < ny; j++) out[j] = (j % 2 == 0 ? 1.0 : -1.0) * X[2 * (j + 1) + 1];
    }
}

int or_transform(int bc_neg, int bc_pos, int inverse, int nx, int ny, const double* in, double* out)
{
    const int kind = transform_kind(bc_neg, bc_pos);
    if (kind == T_BAD || nx <= 0 || ny <= 0) return 1;
    if (kind == T_PP && (ny % 2 != 0 || ny < 2)) return 1;
    int ok = 0;
#pragma omp parallel
    {
        double* ex = (double*)malloc(sizeof(double) * (4 * (size_t)ny + 4));
        double* X  = (double*)malloc(sizeof(double) * 2 * (4 * (size_t)ny + 8));
        double* s  = (double*)malloc(sizeof(double) * 6 * (2 * (size_t)ny + 4));
        double* o  = (double*)malloc(sizeof(double) * (size_t)ny);
#pragma omp for schedule(dynamic)
        for (int i = 0; i < nx; i++) {
            row_transform(kind, inverse, ny, in + (size_t)i * ny, o, ex, X, s);
            memcpy(out + (size_t)i * ny, o, sizeof(double) * ny);
        }
        free(ex); free(X); free(s); free(o);
    }
    return ok;
}

/* PoissonSolver2DBase::cal_lambda (poisson_solver2d_base.cpp:28-62) and the
 * pivot of PoissonSolver2D::init (poisson_solver2d.cpp:43-51). */
int or_x_diag(int ny, double hx, double hy, int neg, int pos, double* out)
{
    const int kind = transform_kind(neg, pos);
    if (kind == T_BAD) return 1;
    for (int j = 0; j < ny; j++) {
        const int i = j + 1;
        double    lam;
        if (kind == T_PP) lam = -2.0 * cos(2.0 * OR_PI / ny * floor(i / 2.0));
        else if (kind == T_NN) lam = -2.0 * cos(OR_PI / ny * i);
        else if (kind == T_DD) lam = 2.0 * cos(OR_PI / (ny + 1) * i);
        else lam = -2.0 * cos(2.0 * OR_PI * i / (2 * ny + 1));
        out[j] = -2.0 + hx * hx / hy / hy * (lam - 2.0);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Thomas (pe/poisson/chasing_method_base.cpp)                                */
/* ------------------------------------------------------------------------ */

/* chasing_standard_without_precompute, chasing_method_base.cpp:24-53 */
void or_chasing(double d, int n, int bs, int be, const double* f, double* x)
{
    double* c = (double*)malloc(sizeof(double) * (size_t)(n > 1 ? n : 2));
    double* y = (double*)malloc(sizeof(double) * (size_t)(n > 1 ? n : 2));
    c[0]      = (bs == OR_NEUMANN) ? 1.0 / (d + 1.0) : 1.0 / d;
    for (int i = 1; i < n - 1; i++) c[i] = 1.0 / (d - c[i - 1]);
    y[0] = (bs == OR_NEUMANN) ? f[0] / (d + 1.0) : f[0] / d;
    for (int i = 1; i < n - 1; i++) y[i] = (f[i] - y[i - 1]) / (d - c[i - 1]);
    y[n - 1] = (be == OR_NEUMANN) ? (f[n - 1] - y[n - 2]) / (d + 1.0 - c[n - 2])
                                  : (f[n - 1] - y[n - 2]) / (d - c[n - 2]);
    x[n - 1] = y[n - 1];
    for (int i = n - 2; i >= 0; i--) x[i] = y[i] - c[i] * x[i + 1];
    free(c); free(y);
}

/* chasing_standard_singular with p_const = 0 (chasing_method_base.cpp:146-154,
 * chasing_method_base.h:49) */
void or_chasing_singular(int n, const double* f, double* x)
{
    x[0] = 0.0;
    x[1] = 0.0 + f[0];
    for (int i = 2; i < n; i++) x[i] = f[i - 1] - x[i - 2] + 2.0 * x[i - 1];
}

/* chasing_periodic_precompute + chasing_periodic_with_precompute
 * (chasing_method_base.cpp:178-258): Dirichlet-end Thomas for f and for the
 * unit vectors e_0, e_{n-1} (z_1, z_2), then the Sherman-Morrison
 * correction with the reference's gamma formulas. */
void or_chasing_periodic(double d, int n, const double* f, double* x)
{
    double* y  = (double*)malloc(sizeof(double) * (size_t)n);
    double* z1 = (double*)malloc(sizeof(double) * (size_t)n);
    double* z2 = (double*)malloc(sizeof(double) * (size_t)n);
    double* u  = (double*)calloc((size_t)n, sizeof(double));
    or_chasing(d, n, OR_DIRICHLET, OR_DIRICHLET, f, y);
    u[0] = 1.0;
    or_chasing(d, n, OR_DIRICHLET, OR_DIRICHLET, u, z1);
    u[0]     = 0.0;
    u[n - 1] = 1.0;
    or_chasing(d, n, OR_DIRICHLET, OR_DIRICHLET, u, z2);
    const double g2 = (y[n - 1] / (1.0 + z1[n - 1]) - y[0] / z1[0]) /
                      (z2[n - 1] / (1.0 + z1[n - 1]) - (1.0 + z2[0]) / z1[0]);
    const double g1 = y[0] / z1[0] - g2 * ((1.0 + z2[0]) / z1[0]);
    for (int i = 0; i < n; i++) x[i] = y[i] - g1 * z1[i] - g2 * z2[i];
    free(y); free(z1); free(z2); free(u);
}

/* chasing_periodic_singular (chasing_method_base.cpp:274-285), p_const = 0 */
void or_chasing_periodic_singular(int n, const double* f, double* x)
{
    x[0] = 0.0;
    x[1] = f[0] / n;
    for (int i = 1; i < n - 1; i++) x[1] = x[1] - (n - 1 - i) * f[i] / n;
    for (int i = 2; i < n; i++) x[i] = f[i - 1] - x[i - 2] + 2.0 * x[i - 1];
}

/* ------------------------------------------------------------------------ */
/* Subdomain solver (pe/poisson/poisson_solver2d.cpp:37-90)                   */
/* ------------------------------------------------------------------------ */

typedef struct sub_solver {
    int     nx, ny;
    double  hx, hy;
    int     bc[4];
    double* x_diag;
    int     no_dirichlet;
    double* buf;    /* nx*ny */
    double* buf2;   /* nx*ny */
} sub_solver;

static sub_solver* sub_create(int nx, int ny, double hx, double hy, const int* bc)
{
    if (transform_kind(bc[2], bc[3]) == T_BAD) return NULL;
    if (transform_kind(bc[2], bc[3]) == T_PP && ny % 2 != 0) return NULL;
    /* a periodic x face needs its partner (ChasingMethod2D::init, chasing_method2d.cpp:21) */
    if ((bc[0] == OR_PERIODIC) != (bc[1] == OR_PERIODIC)) return NULL;
    if (nx < 2 || ny < 1) return NULL;
    sub_solver* s = (sub_solver*)calloc(1, sizeof(sub_solver));
    s->nx = nx; s->ny = ny; s->hx = hx; s->hy = hy;
    memcpy(s->bc, bc, sizeof(int) * 4);
    s->x_diag = (double*)malloc(sizeof(double) * ny);
    or_x_diag(ny, hx, hy, bc[2], bc[3], s->x_diag);
    s->no_dirichlet = !(is_dir_like(bc[0]) || is_dir_like(bc[1]) || is_dir_like(bc[2]) || is_dir_like(bc[3]));
    s->buf  = (double*)malloc(sizeof(double) * (size_t)nx * ny);
    s->buf2 = (double*)malloc(sizeof(double) * (size_t)nx * ny);
    return s;
}

static void sub_destroy(sub_solver* s)
{
    if (!s) return;
    free(s->x_diag); free(s->buf); free(s->buf2); free(s);
}

/* transform -> transpose -> chasing over x per mode -> transpose -> inverse */
static void sub_solve(sub_solver* s, double* f)
{
    const int nx = s->nx, ny = s->ny;
    double*   t    = s->buf;
    double*   tt   = s->buf2;
    or_transform(s->bc[2], s->bc[3], 0, nx, ny, f, t);
    /* field2::transpose, field2.cpp:365-380 */
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) tt[(size_t)j * nx + i] = t[(size_t)i * ny + j];
    /* ChasingMethod2D::chasing, chasing_method2d.cpp:67-88 */
#pragma omp parallel
    {
        double* o = (double*)malloc(sizeof(double) * nx);
#pragma omp for schedule(dynamic)
        for (int j = 0; j < ny; j++) {
            const int per = s->bc[0] == OR_PERIODIC;
            if (s->no_dirichlet && j == ny - 1) {
                if (per) or_chasing_periodic_singular(nx, tt + (size_t)j * nx, o);
                else or_chasing_singular(nx, tt + (size_t)j * nx, o);
            } else if (per) {
                or_chasing_periodic(s->x_diag[j], nx, tt + (size_t)j * nx, o);
            } else {
                or_chasing(s->x_diag[j], nx, s->bc[0], s->bc[1], tt + (size_t)j * nx, o);
            }
            memcpy(t + (size_t)j * nx, o, sizeof(double) * nx);
        }
        free(o);
    }
#pragma omp parallel for schedule(dynamic)
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) tt[(size_t)i * ny + j] = t[(size_t)j * nx + i];
    or_transform(s->bc[2], s->bc[3], 1, nx, ny, tt, f);
}

int or_sub_solve(int nx, int ny, double hx, double hy, const int* bc4, double* f)
{
    sub_solver* s = sub_create(nx, ny, hx, hy, bc4);
    if (!s) return 1;
    sub_solve(s, f);
    sub_destroy(s);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* 3-D subdomain solver (pe/poisson/poisson_solver3d.cpp)                     */
/* ------------------------------------------------------------------------ */

/* PoissonSolver3DBase::cal_lambda (poisson_solver3d_base.cpp:30-66): lambda - 2
 * of mode i = 1..N. */
static double lambda3(int kind, int i, int N)
{
    if (kind == T_PP) return -2.0 - 2.0 * cos(2.0 * OR_PI / N * floor(i / 2.0));
    if (kind == T_NN) return -2.0 - 2.0 * cos(OR_PI / N * i);
    if (kind == T_DD) return -2.0 + 2.0 * cos(OR_PI / (N + 1) * i);
    return -2.0 - 2.0 * cos(2.0 * OR_PI * i / (2 * N + 1));
}

/* PoissonSolver3D::init + solve (poisson_solver3d.cpp:49-123) with
 * ChasingMethod3D::chasing (chasing_method3d.cpp:57-103) on an already
 * hx^2-scaled right-hand side f, layout [i][j][k] (k contiguous, field3).
 * Deviation: with a periodic x axis and no Dirichlet-like face the reference
 * applies the periodic singular recurrence to every y mode of the last z mode
 * (chasing_method3d.cpp:66); the zero mode alone is singular, and only it
 * takes the recurrence here (as in the non-periodic branch, :84). */
int or_sub_solve3d(int nx, int ny, int nz, double hx, double hy, double hz, const int* bc, double* f)
{
    const int ky = transform_kind(bc[2], bc[3]), kz = transform_kind(bc[4], bc[5]);
    if (ky == T_BAD || kz == T_BAD) return 1;
    if ((ky == T_PP && ny % 2) || (kz == T_PP && nz % 2)) return 1;
    if ((bc[0] == OR_PERIODIC) != (bc[1] == OR_PERIODIC) || nx < 2) return 1;
    const size_t n  = (size_t)nx * ny * nz;
    double*      t  = (double*)malloc(sizeof(double) * n);
    double*      u  = (double*)malloc(sizeof(double) * n);
    double*      xd = (double*)malloc(sizeof(double) * (size_t)ny * nz);
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            xd[(size_t)k * ny + j] = -2.0 + hx * hx / hy / hy * lambda3(ky, j + 1, ny) +
                                     hx * hx / hz / hz * lambda3(kz, k + 1, nz);
    int no_dir = 1;
    for (int q = 0; q < 6; q++)
        if (is_dir_like(bc[q])) no_dir = 0;
    or_transform(bc[4], bc[5], 0, nx * ny, nz, f, t);
    /* field3::transpose {2, 0, 1}: (i, j, k) -> (k, i, j) */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) u[((size_t)k * nx + i) * ny + j] = t[((size_t)i * ny + j) * nz + k];
    or_transform(bc[2], bc[3], 0, nz * nx, ny, u, t);
    /* x lines of every (k, j) mode pair: ChasingMethod3D::chasing */
#pragma omp parallel
    {
        double* in  = (double*)malloc(sizeof(double) * nx);
        double* out = (double*)malloc(sizeof(double) * nx);
#pragma omp for schedule(dynamic)
        for (int kj = 0; kj < ny * nz; kj++) {
            const int k = kj / ny, j = kj % ny;
            for (int i = 0; i < nx; i++) in[i] = t[((size_t)k * nx + i) * ny + j];
            const int sing = no_dir && k == nz - 1 && j == ny - 1;
            if (bc[0] == OR_PERIODIC) {
                if (sing) or_chasing_periodic_singular(nx, in, out);
                else or_chasing_periodic(xd[kj], nx, in, out);
            } else if (sing) {
                or_chasing_singular(nx, in, out);
            } else {
                or_chasing(xd[kj], nx, bc[0], bc[1], in, out);
            }
            for (int i = 0; i < nx; i++) t[((size_t)k * nx + i) * ny + j] = out[i];
        }
        free(in); free(out);
    }
    or_transform(bc[2], bc[3], 1, nz * nx, ny, t, u);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) t[((size_t)i * ny + j) * nz + k] = u[((size_t)k * nx + i) * ny + j];
    or_transform(bc[4], bc[5], 1, nx * ny, nz, t, f);
    free(t); free(u); free(xd);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Composite solver (pe/concat/, base/domain/geometry2d.cpp, geometry_tree.hpp) */
/* ------------------------------------------------------------------------ */

static int opposite(int loc) { return loc ^ 1; }

typedef struct schur {
    int     child;
    int     loc;   /* child's side of the parent */
    int     cn;
    double* v;     /* value(j,i) row-major cn x cn */
} schur;

typedef struct dsolver {
    int         dom;
    int         gmres;       /* 0: plain PoissonSolver2D; 1: GMRESSolver2D */
    sub_solver* pe;
    int         nS;
    schur       S[4];
    /* GMRES buffers */
    double *x, *r, *w, *ft, *H, *cs, *sn, *g, *y;
    double** V;
    int      last_iters;
    double   last_res;
} dsolver;

typedef struct or_cat {
    int      nd;
    int*     nx; int* ny;
    double*  hx; double* hy;
    int*     bc;        /* nd*4 effective types */
    int*     adj;       /* nd*4 neighbour index or -1 */
    int      root;
    int*     parent; int* parent_loc;  /* parent_loc: side of the child where the parent is */
    int*     level;
    int      nlev;
    int*     order;     /* domains sorted by level (BFS order) */
    int*     lev_start;
    int*     child;     /* nd*4: child at side loc or -1 */
    dsolver** solver;
    double** temp;
    int      m; double tol; int tol_rel; int maxiter;
    int      schur_mode;
    int      last_iters; double last_res;
} or_cat;

static void field_bond_add_line(int loc, double a, double* f, int nx, int ny, const double* nb, int nbx,
                                int nby)
{
    /* field2::{xneg,xpos,yneg,up}_bond_add(a, neighbour), field2.cpp:257-302 */
    if (loc == OR_XNEG)
        for (int j = 0; j < ny; j++) f[j] += a * nb[(size_t)(nbx - 1) * nby + j];
    else if (loc == OR_XPOS)
        for (int j = 0; j < ny; j++) f[(size_t)(nx - 1) * ny + j] += a * nb[j];
    else if (loc == OR_YNEG)
        for (int i = 0; i < nx; i++) f[(size_t)i * ny] += a * nb[(size_t)i * nby + nby - 1];
    else
        for (int i = 0; i < nx; i++) f[(size_t)i * ny + ny - 1] += a * nb[(size_t)i * nby];
    (void)nbx;
}

static void dsolve(or_cat* c, dsolver* d, double* f);

/* SchurMat2D_*::construct, schur_mat2d.cpp:5-16,35-46,65-76,95-106 */
static void schur_unit_solves(or_cat* c, schur* S)
{
    const int ch = S->child, bnx = c->nx[ch], bny = c->ny[ch], cn = S->cn;
    double*   t  = (double*)malloc(sizeof(double) * (size_t)bnx * bny);
    for (int i = 0; i < cn; i++) {
        memset(t, 0, sizeof(double) * (size_t)bnx * bny);
        size_t src;
        if (S->loc == OR_XNEG) src = (size_t)(bnx - 1) * bny + i;
        else if (S->loc == OR_XPOS) src = (size_t)i;
        else if (S->loc == OR_YPOS) src = (size_t)i * bny;
        else src = (size_t)i * bny + bny - 1;
        t[src] = 1.0;
        dsolve(c, c->solver[ch], t);
        for (int j = 0; j < cn; j++) {
            size_t dst;
            if (S->loc == OR_XNEG) dst = (size_t)(bnx - 1) * bny + j;
            else if (S->loc == OR_XPOS) dst = (size_t)j;
            else if (S->loc == OR_YPOS) dst = (size_t)j * bny;
            else dst = (size_t)j * bny + bny - 1;
            S->v[(size_t)j * cn + i] = t[dst];
        }
    }
    free(t);
}

/* Same matrix from the leaf child's eigen-decomposition (not the reference's
 * construction; used at large sizes, checked against schur_unit_solves):
 * x-line (row i0): S = Phi^T diag((T_k^-1)[i0][i0]) Phi
 * y-line (col j0): S = sum_k Phi[k][j0]^2 T_k^-1 */
static void schur_spectral(or_cat* c, schur* S)
{
    const int   ch = S->child, bnx = c->nx[ch], bny = c->ny[ch], cn = S->cn;
    sub_solver* pe = c->solver[ch]->pe;
    const int   bs = pe->bc[0], be = pe->bc[1];
    if (S->loc == OR_XNEG || S->loc == OR_XPOS) {
        const int i0 = (S->loc == OR_XNEG) ? bnx - 1 : 0;
        double*   G  = (double*)malloc(sizeof(double) * bny);
#pragma omp parallel
        {
            double* e = (double*)calloc(bnx, sizeof(double));
            double* x = (double*)malloc(sizeof(double) * bnx);
#pragma omp for
            for (int k = 0; k < bny; k++) {
                memset(e, 0, sizeof(double) * bnx);
                e[i0] = 1.0;
                or_chasing(pe->x_diag[k], bnx, bs, be, e, x);
                G[k] = x[i0];
            }
            free(e); free(x);
        }
        /* Phi rows from transforms of unit vectors: P[n][k] = Phi[k][n] */
        double* I = (double*)calloc((size_t)cn * cn, sizeof(double));
        double* P = (double*)malloc(sizeof(double) * (size_t)cn * cn);
        for (int n = 0; n < cn; n++) I[(size_t)n * cn + n] = 1.0;
        or_transform(pe->bc[2], pe->bc[3], 0, cn, cn, I, P);
        for (int n = 0; n < cn; n++)
            for (int k = 0; k < cn; k++) P[(size_t)n * cn + k] *= G[k];
        /* column n of S = Phi^T (G . Phi e_n) -> row-wise inverse transform */
        or_transform(pe->bc[2], pe->bc[3], 1, cn, cn, P, I);
        /* I[n][j] = S(j, n) */
        for (int n = 0; n < cn; n++)
            for (int j = 0; j < cn; j++) S->v[(size_t)j * cn + n] = I[(size_t)n * cn + j];
        free(I); free(P); free(G);
    } else {
        const int j0 = (S->loc == OR_YPOS) ? 0 : bny - 1;
        double*   E  = (double*)calloc(bny, sizeof(double));
        double*   W  = (double*)malloc(sizeof(double) * bny);
        E[j0]        = 1.0;
        or_transform(pe->bc[2], pe->bc[3], 0, 1, bny, E, W);  /* W[k] = Phi[k][j0] */
        memset(S->v, 0, sizeof(double) * (size_t)cn * cn);
#pragma omp parallel
        {
            double* e = (double*)calloc(bnx, sizeof(double));
            double* x = (double*)malloc(sizeof(double) * bnx);
#pragma omp for schedule(dynamic)
            for (int i = 0; i < cn; i++) {
                for (int k = 0; k < bny; k++) {
                    memset(e, 0, sizeof(double) * bnx);
                    e[i] = 1.0;
                    or_chasing(pe->x_diag[k], bnx, bs, be, e, x);
                    const double w2 = W[k] * W[k];
                    for (int j = 0; j < cn; j++) S->v[(size_t)j * cn + i] += w2 * x[j];
                }
            }
            free(e); free(x);
        }
        free(E); free(W);
    }
}

/* SchurMat2D_*::operator*, schur_mat2d.cpp:18-33 etc.: adds S * (root line) into
 * the same line of out. */
static void schur_apply_add(const schur* S, const double* root, double* out, int rnx, int rny)
{
    const int cn = S->cn;
#pragma omp parallel for
    for (int i = 0; i < cn; i++) {
        double acc = 0.0;
        for (int j = 0; j < cn; j++) {
            double xj;
            if (S->loc == OR_XNEG) xj = root[j];
            else if (S->loc == OR_XPOS) xj = root[(size_t)(rnx - 1) * rny + j];
            else if (S->loc == OR_YPOS) xj = root[(size_t)j * rny + rny - 1];
            else xj = root[(size_t)j * rny];
            acc += xj * S->v[(size_t)i * cn + j];
        }
        if (S->loc == OR_XNEG) out[i] += acc;
        else if (S->loc == OR_XPOS) out[(size_t)(rnx - 1) * rny + i] += acc;
        else if (S->loc == OR_YPOS) out[(size_t)i * rny + rny - 1] += acc;
        else out[(size_t)i * rny] += acc;
    }
}

static double vdot(const double* a, const double* b, size_t n)
{
    double s = 0.0;
    for (size_t i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* GMRESSolver2D::Afun, gmres_solver2d.cpp:90-104: out = x - A^-1 sum_k S_k x */
static void gmres_afun(or_cat* c, dsolver* d, const double* x, double* out)
{
    const int    nx = c->nx[d->dom], ny = c->ny[d->dom];
    const size_t n  = (size_t)nx * ny;
    memset(d->ft, 0, sizeof(double) * n);
    for (int k = 0; k < d->nS; k++) schur_apply_add(&d->S[k], x, d->ft, nx, ny);
    sub_solve(d->pe, d->ft);
#pragma omp parallel for
    for (long i = 0; i < (long)n; i++) out[i] = x[i] - d->ft[i];
}

/* GMRESSolver2D::solve, gmres_solver2d.cpp:123-252, written against an
 * abstract operator so the mock-operator tests exercise the same code. */
typedef void (*afun_t)(void* ctx, const double* x, double* out);
typedef void (*prec_t)(void* ctx, double* b);

static int gmres_core(size_t n, int m, double tol, int tol_rel, int maxiter, afun_t afun, prec_t prec,
                      void* ctx, double* b, double* x, double* r, double* w, double** V, double* H,
                      double* cs, double* sn, double* g, double* y, double* resvec, int* nres, int* iters,
                      double* lastres)
{
    prec(ctx, b);                    /* b <- A^-1 b (gmres_solver2d.cpp:133) */
    if (tol_rel) tol *= sqrt(vdot(b, b, n));
    memcpy(x, b, sizeof(double) * n); /* x0 = b (:139-142) */
    int    nr = 0, total = 0;
    double beta = 0.0;
    for (int outer = 0; outer < maxiter; outer++) {
        afun(ctx, x, w);
#pragma omp parallel for
        for (long i = 0; i < (long)n; i++) r[i] = b[i] - w[i];
        beta = sqrt(vdot(r, r, n));
        if (resvec) resvec[nr] = beta;
        nr++;
        if (beta < tol) {
            memcpy(b, x, sizeof(double) * n);
            if (nres) *nres = nr;
            if (iters) *iters = total;
            if (lastres) *lastres = beta;
            return 0;
        }
        memset(H, 0, sizeof(double) * (size_t)(m + 1) * m);
        memset(cs, 0, sizeof(double) * m);
        memset(sn, 0, sizeof(double) * m);
        memset(g, 0, sizeof(double) * (m + 1));
#pragma omp parallel for
        for (long i = 0; i < (long)n; i++) V[0][i] = r[i] * (1.0 / beta);
        g[0] = beta;
        int j = 0;
        for (; j < m; j++) {
            afun(ctx, V[j], w);
            total++;
            for (int i = 0; i <= j; i++) {
                H[i * m + j] = vdot(V[i], w, n);
                const double h = H[i * m + j];
#pragma omp parallel for
                for (long q = 0; q < (long)n; q++) w[q] += -h * V[i][q];
            }
            const double hn = sqrt(vdot(w, w, n));
            if (hn < 1e-12) break;
            H[(j + 1) * m + j] = hn;
#pragma omp parallel for
            for (long q = 0; q < (long)n; q++) V[j + 1][q] = w[q] * (1.0 / hn);
            for (int i = 0; i < j; i++) {
                const double t     = cs[i] * H[i * m + j] + sn[i] * H[(i + 1) * m + j];
                H[(i + 1) * m + j] = -sn[i] * H[i * m + j] + cs[i] * H[(i + 1) * m + j];
                H[i * m + j]       = t;
            }
            const double hjj = H[j * m + j], hj1 = H[(j + 1) * m + j];
            const double den = hypot(hjj, hj1);
            cs[j]            = hjj / den;
            sn[j]            = hj1 / den;
            H[j * m + j]       = cs[j] * H[j * m + j] + sn[j] * H[(j + 1) * m + j];
            H[(j + 1) * m + j] = 0.0;
            const double t     = cs[j] * g[j];
            g[j + 1]           = -sn[j] * g[j];
            g[j]               = t;
            if (fabs(g[j + 1]) < tol) { j++; break; }
        }
        for (int i = 0; i < m; i++) y[i] = 0.0;
        for (int i = j - 1; i >= 0; i--) {
            y[i] = g[i];
            for (int k = i + 1; k < j; k++) y[i] -= H[i * m + k] * y[k];
            y[i] /= H[i * m + i];
        }
        for (int i = 0; i < j; i++) {
            const double yi = y[i];
#pragma omp parallel for
            for (long q = 0; q < (long)n; q++) x[q] += yi * V[i][q];
        }
    }
    memcpy(b, x, sizeof(double) * n);
    if (nres) *nres = nr;
    if (iters) *iters = total;
    if (lastres) *lastres = beta;
    return 2; /* returned after maxiter without meeting tol (silent in the reference) */
}

typedef struct { or_cat* c; dsolver* d; } gm_ctx;
static void gm_afun(void* ctx, const double* x, double* out)
{
    gm_ctx* g = (gm_ctx*)ctx;
    gmres_afun(g->c, g->d, x, out);
}
static void gm_prec(void* ctx, double* b)
{
    gm_ctx* g = (gm_ctx*)ctx;
    sub_solve(g->d->pe, b);
}

static void dsolve(or_cat* c, dsolver* d, double* f)
{
    if (!d->gmres) { sub_solve(d->pe, f); return; }
    const size_t n  = (size_t)c->nx[d->dom] * c->ny[d->dom];
    gm_ctx       ctx = {c, d};
    int          it  = 0;
    double       res = 0.0;
    gmres_core(n, c->m, c->tol, c->tol_rel, c->maxiter, gm_afun, gm_prec, &ctx, f, d->x, d->r, d->w, d->V,
               d->H, d->cs, d->sn, d->g, d->y, NULL, NULL, &it, &res);
    d->last_iters = it;
    d->last_res   = res;
}

static dsolver* dsolver_create(or_cat* c, int dom)
{
    dsolver* d = (dsolver*)calloc(1, sizeof(dsolver));
    d->dom     = dom;
    d->pe      = sub_create(c->nx[dom], c->ny[dom], c->hx[dom], c->hy[dom], c->bc + 4 * dom);
    if (!d->pe) { free(d); return NULL; }
    int nch = 0;
    for (int loc = 0; loc < 4; loc++)
        if (c->child[4 * dom + loc] >= 0) nch++;
    if (nch == 0) return d;
    d->gmres     = 1;
    const size_t n = (size_t)c->nx[dom] * c->ny[dom];
    const int    m = c->m;
    d->x  = (double*)malloc(sizeof(double) * n);
    d->r  = (double*)malloc(sizeof(double) * n);
    d->w  = (double*)malloc(sizeof(double) * n);
    d->ft = (double*)malloc(sizeof(double) * n);
    d->V  = (double**)malloc(sizeof(double*) * (m + 1));
    for (int i = 0; i <= m; i++) d->V[i] = (double*)malloc(sizeof(double) * n);
    d->H  = (double*)calloc((size_t)(m + 1) * m, sizeof(double));
    d->cs = (double*)calloc(m, sizeof(double));
    d->sn = (double*)calloc(m, sizeof(double));
    d->g  = (double*)calloc(m + 1, sizeof(double));
    d->y  = (double*)calloc(m, sizeof(double));
    /* GMRESSolver2D::schur_mat_construct, gmres_solver2d.cpp:39-89 */
    for (int loc = 0; loc < 4; loc++) {
        const int ch = c->child[4 * dom + loc];
        if (ch < 0) continue;
        schur* S = &d->S[d->nS++];
        S->child = ch;
        S->loc   = loc;
        S->cn    = (loc == OR_XNEG || loc == OR_XPOS) ? c->ny[ch] : c->nx[ch];
        S->v     = (double*)malloc(sizeof(double) * (size_t)S->cn * S->cn);
        /* the spectral form solves the child's x systems with Dirichlet/Neumann
         * ends: a periodic x axis takes the unit solves */
        if (c->schur_mode == OR_SCHUR_SPECTRAL && !c->solver[ch]->gmres && c->bc[4 * ch + OR_XNEG] != OR_PERIODIC)
            schur_spectral(c, S);
        else schur_unit_solves(c, S);
    }
    return d;
}

static void dsolver_destroy(or_cat* c, dsolver* d)
{
    if (!d) return;
    sub_destroy(d->pe);
    if (d->gmres) {
        free(d->x); free(d->r); free(d->w); free(d->ft);
        for (int i = 0; i <= c->m; i++) free(d->V[i]);
        free(d->V); free(d->H); free(d->cs); free(d->sn); free(d->g); free(d->y);
        for (int k = 0; k < d->nS; k++) free(d->S[k].v);
    }
    free(d);
}

/* TreeUtils::findOptimalRoot (geometry_tree.hpp:22-95).  The reference breaks a
 * two-centre tie by the smaller Domain2DUniform address; domain index is used
 * here as the deterministic stand-in for that address order. */
static int find_root(or_cat* c)
{
    const int nd = c->nd;
    if (nd == 1) return 0;
    int* deg   = (int*)calloc(nd, sizeof(int));
    int* queue = (int*)malloc(sizeof(int) * (size_t)nd * 8);
    int  qh = 0, qt = 0;
    for (int a = 0; a < nd; a++)
        for (int l = 0; l < 4; l++)
            if (c->adj[4 * a + l] >= 0) deg[a]++;
    for (int a = 0; a < nd; a++)
        if (deg[a] <= 1) queue[qt++] = a;
    int remaining = nd;
    while (remaining > 2) {
        int cnt = qt - qh;
        if (cnt == 0) break;
        remaining -= cnt;
        for (int i = 0; i < cnt; i++) {
            int leaf = queue[qh++];
            for (int l = 0; l < 4; l++) {
                int nb = c->adj[4 * leaf + l];
                if (nb >= 0 && --deg[nb] == 1) queue[qt++] = nb;
            }
        }
    }
    int root;
    if (qt == qh) root = 0;
    else {
        root = queue[qh++];
        if (qh < qt && queue[qh] < root) root = queue[qh];
    }
    free(deg); free(queue);
    return root;
}

/* buildTreeMapRecursive + buildParentMapFromTree + BFS (geometry_tree.hpp:97-200) */
static void build_tree(or_cat* c)
{
    const int nd = c->nd;
    c->root      = find_root(c);
    for (int i = 0; i < 4 * nd; i++) c->child[i] = -1;
    for (int i = 0; i < nd; i++) { c->parent[i] = -1; c->parent_loc[i] = -1; c->level[i] = -1; }
    int* q  = (int*)malloc(sizeof(int) * nd);
    int  qh = 0, qt = 0;
    q[qt++]           = c->root;
    c->level[c->root] = 0;
    while (qh < qt) {
        int a = q[qh++];
        for (int l = 0; l < 4; l++) {
            int b = c->adj[4 * a + l];
            if (b < 0 || b == c->parent[a]) continue;
            if (c->level[b] >= 0) continue;
            c->child[4 * a + l] = b;
            c->parent[b]        = a;
            c->parent_loc[b]    = opposite(l);
            c->level[b]         = c->level[a] + 1;
            q[qt++]             = b;
        }
    }
    c->nlev = 0;
    for (int i = 0; i < nd; i++)
        if (c->level[i] + 1 > c->nlev) c->nlev = c->level[i] + 1;
    memcpy(c->order, q, sizeof(int) * nd);
    c->lev_start = (int*)calloc(c->nlev + 1, sizeof(int));
    for (int i = 0; i < nd; i++) c->lev_start[c->level[i] + 1]++;
    for (int l = 0; l < c->nlev; l++) c->lev_start[l + 1] += c->lev_start[l];
    free(q);
}

void* or_cat_create(int nd, const int* nx, const int* ny, const double* hx, const double* hy, const int* bc,
                    int nconn, const int* conn, int m, double tol, int tol_rel, int maxiter, int schur_mode)
{
    if (nd <= 0 || m <= 0) return NULL;
    or_cat* c = (or_cat*)calloc(1, sizeof(or_cat));
    c->nd     = nd;
    c->nx = (int*)malloc(sizeof(int) * nd); c->ny = (int*)malloc(sizeof(int) * nd);
    c->hx = (double*)malloc(sizeof(double) * nd); c->hy = (double*)malloc(sizeof(double) * nd);
    c->bc  = (int*)malloc(sizeof(int) * 4 * nd);
    c->adj = (int*)malloc(sizeof(int) * 4 * nd);
    c->child = (int*)malloc(sizeof(int) * 4 * nd);
    c->parent = (int*)malloc(sizeof(int) * nd); c->parent_loc = (int*)malloc(sizeof(int) * nd);
    c->level = (int*)malloc(sizeof(int) * nd); c->order = (int*)malloc(sizeof(int) * nd);
    memcpy(c->nx, nx, sizeof(int) * nd); memcpy(c->ny, ny, sizeof(int) * nd);
    memcpy(c->hx, hx, sizeof(double) * nd); memcpy(c->hy, hy, sizeof(double) * nd);
    memcpy(c->bc, bc, sizeof(int) * 4 * nd);
    for (int i = 0; i < 4 * nd; i++) c->adj[i] = -1;
    for (int k = 0; k < nconn; k++) {
        int a = conn[3 * k], dir = conn[3 * k + 1], b = conn[3 * k + 2];
        c->adj[4 * a + dir]           = b;
        c->adj[4 * b + opposite(dir)] = a;
        /* Variable2D::set_geometry: adjacent faces become Adjacented (variable2d.cpp:25-36) */
        c->bc[4 * a + dir]           = OR_ADJACENTED;
        c->bc[4 * b + opposite(dir)] = OR_ADJACENTED;
    }
    c->m = m; c->tol = tol; c->tol_rel = tol_rel; c->maxiter = maxiter; c->schur_mode = schur_mode;
    build_tree(c);
    c->solver = (dsolver**)calloc(nd, sizeof(dsolver*));
    c->temp   = (double**)calloc(nd, sizeof(double*));
    /* construct_solver_map: deepest level first, root last (concat_solver2d.cpp:88-99) */
    for (int l = c->nlev - 1; l >= 0; l--) {
        for (int p = c->lev_start[l]; p < c->lev_start[l + 1]; p++) {
            int dom        = c->order[p];
            c->solver[dom] = dsolver_create(c, dom);
            if (!c->solver[dom]) { or_cat_destroy(c); return NULL; }
            if (dom != c->root) c->temp[dom] = (double*)malloc(sizeof(double) * (size_t)nx[dom] * ny[dom]);
        }
    }
    return c;
}

int or_cat_root(void* h) { return h ? ((or_cat*)h)->root : -1; }

int or_cat_tree(void* h, int* level, int* parent, int* parent_loc)
{
    or_cat* c = (or_cat*)h;
    if (!c) return 1;
    memcpy(level, c->level, sizeof(int) * c->nd);
    memcpy(parent, c->parent, sizeof(int) * c->nd);
    memcpy(parent_loc, c->parent_loc, sizeof(int) * c->nd);
    return 0;
}

int or_cat_schur(void* h, int child, double* out, int* cn)
{
    or_cat* c = (or_cat*)h;
    if (!c || child < 0 || child >= c->nd || c->parent[child] < 0) return 1;
    dsolver* d = c->solver[c->parent[child]];
    for (int k = 0; k < d->nS; k++)
        if (d->S[k].child == child) {
            *cn = d->S[k].cn;
            if (out) memcpy(out, d->S[k].v, sizeof(double) * (size_t)d->S[k].cn * d->S[k].cn);
            return 0;
        }
    return 1;
}

/* ConcatPoissonSolver2D::boundary_assembly (concat_solver2d.cpp:159-223) */
static void boundary_assembly(or_cat* c, int dom, double* f, double* const* buf)
{
    const int    nx = c->nx[dom], ny = c->ny[dom];
    const double hx = c->hx[dom], hy = c->hy[dom];
    const int*   t  = c->bc + 4 * dom;
    for (int loc = 0; loc < 4; loc++) {
        const double* g = buf ? buf[4 * dom + loc] : NULL;
        if (!g) continue;
        double a;
        if (t[loc] == OR_DIRICHLET) a = (loc < 2) ? -1.0 / hx / hx : -1.0 / hy / hy;
        else if (t[loc] == OR_NEUMANN) {
            if (loc == OR_XNEG) a = 1.0 / hx;
            else if (loc == OR_XPOS) a = -1.0 / hx;
            else if (loc == OR_YNEG) a = 1.0 / hy;
            else a = -1.0 / hy;
        } else continue;
        if (loc == OR_XNEG) for (int j = 0; j < ny; j++) f[j] += a * g[j];
        else if (loc == OR_XPOS) for (int j = 0; j < ny; j++) f[(size_t)(nx - 1) * ny + j] += a * g[j];
        else if (loc == OR_YNEG) for (int i = 0; i < nx; i++) f[(size_t)i * ny] += a * g[i];
        else for (int i = 0; i < nx; i++) f[(size_t)i * ny + ny - 1] += a * g[i];
    }
}

/* ConcatPoissonSolver2D::solve (concat_solver2d.cpp:101-157) */
int or_cat_solve(void* h, double* const* fields, double* const* buf, int* iters, double* res)
{
    or_cat* c = (or_cat*)h;
    if (!c) return 1;
    const int nd = c->nd;
    for (int d = 0; d < nd; d++) boundary_assembly(c, d, fields[d], buf);
    for (int d = 0; d < nd; d++) {
        const double s = c->hx[d] * c->hx[d];
        const size_t n = (size_t)c->nx[d] * c->ny[d];
        for (size_t i = 0; i < n; i++) fields[d][i] *= s;
    }
    /* upward sweep, deepest level to level 1 */
    for (int l = c->nlev - 1; l >= 1; l--) {
        for (int p = c->lev_start[l]; p < c->lev_start[l + 1]; p++) {
            int dom = c->order[p];
            memcpy(c->temp[dom], fields[dom], sizeof(double) * (size_t)c->nx[dom] * c->ny[dom]);
            for (int loc = 0; loc < 4; loc++) {
                int ch = c->child[4 * dom + loc];
                if (ch < 0) continue;
                field_bond_add_line(loc, -1.0, c->temp[dom], c->nx[dom], c->ny[dom], c->temp[ch], c->nx[ch],
                                    c->ny[ch]);
                field_bond_add_line(loc, -1.0, fields[dom], c->nx[dom], c->ny[dom], c->temp[ch], c->nx[ch],
                                    c->ny[ch]);
            }
        }
        for (int p = c->lev_start[l]; p < c->lev_start[l + 1]; p++) {
            int dom = c->order[p];
            dsolve(c, c->solver[dom], c->temp[dom]);
        }
    }
    const int r = c->root;
    for (int loc = 0; loc < 4; loc++) {
        int ch = c->child[4 * r + loc];
        if (ch < 0) continue;
        field_bond_add_line(loc, -1.0, fields[r], c->nx[r], c->ny[r], c->temp[ch], c->nx[ch], c->ny[ch]);
    }
    dsolve(c, c->solver[r], fields[r]);
    c->last_iters = c->solver[r]->last_iters;
    c->last_res   = c->solver[r]->last_res;
    /* downward sweep */
    for (int l = 1; l < c->nlev; l++) {
        for (int p = c->lev_start[l]; p < c->lev_start[l + 1]; p++) {
            int dom = c->order[p], par = c->parent[dom];
            field_bond_add_line(c->parent_loc[dom], -1.0, fields[dom], c->nx[dom], c->ny[dom], fields[par],
                                c->nx[par], c->ny[par]);
        }
        for (int p = c->lev_start[l]; p < c->lev_start[l + 1]; p++) {
            int dom = c->order[p];
            dsolve(c, c->solver[dom], fields[dom]);
        }
    }
    if (iters) *iters = c->last_iters;
    if (res) *res = c->last_res;
    return 0;
}

/* Residual of the assembled composite system, accumulated in long double:
 *   r_d = hx_d^2 F_d - A u,  F_d = f_d + ghost terms (boundary_assembly),
 * A the scaled 5-point operator of every domain (chasing_method_base.cpp:33-46
 * Neumann ends, poisson_solver2d.cpp:50 ratio hx^2/hy^2, periodic pairs of
 * chasing_method2d.cpp:44-66 / poisson_fft2d.cpp:264-465) with the
 * neighbour's adjacent line entering each Adjacented face with coefficient 1
 * (concat_solver2d.cpp:127-128,140,150, field2.cpp:257-302). */
static void cat_residual_ld(const or_cat* c, double* const* u, double* const* f, double* const* buf,
                            double* const* r)
{
    for (int d = 0; d < c->nd; d++) {
        const int    nx = c->nx[d], ny = c->ny[d];
        const double hx = c->hx[d], hy = c->hy[d];
        const double ry = hx * hx / (hy * hy);
        const int*   t  = c->bc + 4 * d;
        const double* ud = u[d];
        /* assembled right-hand side, as or_cat_solve forms it */
        memcpy(r[d], f[d], sizeof(double) * (size_t)nx * ny);
        boundary_assembly((or_cat*)c, d, r[d], buf);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < nx; i++) {
            for (int j = 0; j < ny; j++) {
                const size_t k  = (size_t)i * ny + j;
                long double  au = -(2.0L + 2.0L * ry) * ud[k];
                /* x neighbours; an Adjacented face without a neighbour (a lone
                 * domain given that type) has ghost 0, as a Dirichlet face */
                if (i > 0) au += ud[k - ny];
                else if (t[OR_XNEG] == OR_NEUMANN) au += ud[k];
                else if (t[OR_XNEG] == OR_PERIODIC) au += ud[(size_t)(nx - 1) * ny + j];
                else if (t[OR_XNEG] == OR_ADJACENTED && c->adj[4 * d + OR_XNEG] >= 0) {
                    const int nb = c->adj[4 * d + OR_XNEG];
                    au += u[nb][(size_t)(c->nx[nb] - 1) * c->ny[nb] + j];
                }
                if (i < nx - 1) au += ud[k + ny];
                else if (t[OR_XPOS] == OR_NEUMANN) au += ud[k];
                else if (t[OR_XPOS] == OR_PERIODIC) au += ud[j];
                else if (t[OR_XPOS] == OR_ADJACENTED && c->adj[4 * d + OR_XPOS] >= 0) au += u[c->adj[4 * d + OR_XPOS]][j];
                /* y neighbours */
                if (j > 0) au += (long double)ry * ud[k - 1];
                else if (t[OR_YNEG] == OR_NEUMANN) au += (long double)ry * ud[k];
                else if (t[OR_YNEG] == OR_PERIODIC) au += (long double)ry * ud[k + ny - 1];
                else if (t[OR_YNEG] == OR_ADJACENTED && c->adj[4 * d + OR_YNEG] >= 0) {
                    const int nb = c->adj[4 * d + OR_YNEG];
                    au += u[nb][(size_t)i * c->ny[nb] + c->ny[nb] - 1];
                }
                if (j < ny - 1) au += (long double)ry * ud[k + 1];
                else if (t[OR_YPOS] == OR_NEUMANN) au += (long double)ry * ud[k];
                else if (t[OR_YPOS] == OR_PERIODIC) au += (long double)ry * ud[k - (ny - 1)];
                else if (t[OR_YPOS] == OR_ADJACENTED && c->adj[4 * d + OR_YPOS] >= 0) {
                    const int nb = c->adj[4 * d + OR_YPOS];
                    au += u[nb][(size_t)i * c->ny[nb]];
                }
                r[d][k] = (double)((long double)hx * hx * r[d][k] - au);
            }
        }
    }
}

int or_cat_residual(void* h, double* const* u, double* const* fields, double* const* buf, double* const* r)
{
    or_cat* c = (or_cat*)h;
    if (!c) return 1;
    cat_residual_ld(c, u, fields, buf, r);
    return 0;
}

int or_cat_solve_refined(void* h, double* const* fields, double* const* buf, int steps, double tol_refine,
                         int* iters, double* res)
{
    or_cat* c = (or_cat*)h;
    if (!c || steps < 0) return 1;
    const int nd = c->nd;
    double**  f0 = (double**)calloc(nd, sizeof(double*));
    double**  rr = (double**)calloc(nd, sizeof(double*));
    for (int d = 0; d < nd; d++) {
        const size_t n = (size_t)c->nx[d] * c->ny[d];
        f0[d] = (double*)malloc(sizeof(double) * n);
        rr[d] = (double*)malloc(sizeof(double) * n);
        memcpy(f0[d], fields[d], sizeof(double) * n);
    }
    int st = or_cat_solve(h, fields, buf, iters, res);
    const double tol0 = c->tol;
    const int    rel0 = c->tol_rel;
    for (int s = 0; s < steps && st == 0; s++) {
        cat_residual_ld(c, fields, f0, buf, rr);
        /* A delta = r: r is already scaled by hx^2 and assembled, so it enters
         * or_cat_solve as a raw right-hand side r / hx^2 with no ghost data */
        for (int d = 0; d < nd; d++) {
            const size_t n = (size_t)c->nx[d] * c->ny[d];
            const double s2 = 1.0 / (c->hx[d] * c->hx[d]);
            for (size_t i = 0; i < n; i++) rr[d][i] *= s2;
        }
        c->tol     = tol_refine;
        c->tol_rel = 1;
        st         = or_cat_solve(h, rr, NULL, NULL, NULL);
        c->tol     = tol0;
        c->tol_rel = rel0;
        for (int d = 0; d < nd; d++) {
            const size_t n = (size_t)c->nx[d] * c->ny[d];
            for (size_t i = 0; i < n; i++) fields[d][i] += rr[d][i];
        }
    }
    for (int d = 0; d < nd; d++) { free(f0[d]); free(rr[d]); }
    free(f0); free(rr);
    return st;
}

void or_cat_destroy(void* h)
{
    or_cat* c = (or_cat*)h;
    if (!c) return;
    if (c->solver)
        for (int d = 0; d < c->nd; d++) dsolver_destroy(c, c->solver[d]);
    if (c->temp)
        for (int d = 0; d < c->nd; d++) free(c->temp[d]);
    free(c->solver); free(c->temp);
    free(c->nx); free(c->ny); free(c->hx); free(c->hy); free(c->bc); free(c->adj); free(c->child);
    free(c->parent); free(c->parent_loc); free(c->level); free(c->order); free(c->lev_start);
    free(c);
}

/* ------------------------------------------------------------------------ */
/* GMRES on a mock operator (gmres_reference.cpp-style tests)                 */
/* ------------------------------------------------------------------------ */

typedef struct { int n; const double* ainv; const double* S; double* t; } mock_ctx;
static void mock_prec(void* ctx, double* b)
{
    mock_ctx* m = (mock_ctx*)ctx;
    for (int i = 0; i < m->n; i++) b[i] *= m->ainv[i];
}
static void mock_afun(void* ctx, const double* x, double* out)
{
    mock_ctx* m = (mock_ctx*)ctx;
    for (int i = 0; i < m->n; i++) {
        double s = 0.0;
        for (int j = 0; j < m->n; j++) s += m->S[(size_t)i * m->n + j] * x[j];
        m->t[i] = s;
    }
    mock_prec(ctx, m->t);
    for (int i = 0; i < m->n; i++) out[i] = x[i] - m->t[i];
}

int or_gmres_mock(int n, const double* ainv, const double* S, double* b, int m, double tol, int maxiter,
                  double* resvec, int* nres)
{
    mock_ctx ctx = {n, ainv, S, (double*)malloc(sizeof(double) * n)};
    double*  x   = (double*)malloc(sizeof(double) * n);
    double*  r   = (double*)malloc(sizeof(double) * n);
    double*  w   = (double*)malloc(sizeof(double) * n);
    double** V   = (double**)malloc(sizeof(double*) * (m + 1));
    for (int i = 0; i <= m; i++) V[i] = (double*)malloc(sizeof(double) * n);
    double* H  = (double*)calloc((size_t)(m + 1) * m, sizeof(double));
    double* cs = (double*)calloc(m, sizeof(double));
    double* sn = (double*)calloc(m, sizeof(double));
    double* g  = (double*)calloc(m + 1, sizeof(double));
    double* y  = (double*)calloc(m, sizeof(double));
    int     it = 0;
    double  rr = 0;
    int rc = gmres_core(n, m, tol, 0, maxiter, mock_afun, mock_prec, &ctx, b, x, r, w, V, H, cs, sn, g, y,
                        resvec, nres, &it, &rr);
    for (int i = 0; i <= m; i++) free(V[i]);
    free(V); free(H); free(cs); free(sn); free(g); free(y); free(x); free(r); free(w); free(ctx.t);
    return rc;
}

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
