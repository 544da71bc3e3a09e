/*
 * oracle.c -- CPU restatement of block-DDA's solve path (FP64).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Nothing in the
 * product library links or calls this file.  See oracle.h for the list of
 * reference functions restated here (file:line citations are repeated at
 * each function below).
 *
 * Deviations that do not change results beyond rounding:
 *   - FFTW's 1-D transforms are replaced by orc_fft (own mixed radix with
 *     direct prime butterflies, Bluestein for prime factors > 61; four-step
 *     with OpenMP threads for long lengths, see orc_fft_config).
 *   - BT_pad / BT_reconstruct / BT_fft keep the reference recursions but
 *     write into preallocated output at an offset instead of
 *     conservativeResize-and-copy.
 *   - tau^T * Qnew^T * Qnew (bl_cocg_rq_*:66/73) is evaluated as
 *     tau^T * (Qnew^T * Qnew).
 */
#include "oracle.h"
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex zc;

/* ------------------------------------------------------------------ */
/* application_function  (mvp_fft/application_function.cpp:7-78)      */
/* ------------------------------------------------------------------ */
int orc_application_function(const int* n_ind, int f, double lf, double kre,
                             double kim, double* out_)
{
    zc* out = (zc*)out_;
    const zc k = kre + I * kim;
    /* X = (row0 - row1) * lf  (application_function.cpp:21) */
    double X[3];
    for (int a = 0; a < 3; ++a) X[a] = (double)(n_ind[a] - n_ind[3 + a]) * lf;
    const double r = sqrt(X[0] * X[0] + X[1] * X[1] + X[2] * X[2]);
    if (f != 3 && f != 6) return -1;
    for (int i = 0; i < f * f; ++i) out[i] = 0;
    if (r == 0.0) return 0; /* zero self term (:31-34, :46-49) */
    /* G = e^{ikr}/r [k^2 (I - XX/r^2) - (1 - ikr)/r^2 (I - 3 XX/r^2)]  (:36-38) */
    zc G[3][3];
    const zc e = cexp(I * k * r) / r;
    const zc c1 = k * k;
    const zc c2 = (1.0 - I * k * r) / (r * r);
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            const double xx = X[a] * X[b] / (r * r);
            const double id = (a == b) ? 1.0 : 0.0;
            G[a][b] = (c1 * (id - xx) - c2 * (id - 3.0 * xx)) * e;
        }
    if (f == 3) {
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) out[a + 3 * b] = -G[a][b]; /* (:39) */
        return 0;
    }
    /* f = 6: F = -k^2 e^{ikr}/r (1 - 1/(ikr)) [u]_x   (:50-63) */
    double u[3] = {X[0] / r, X[1] / r, X[2] / r};
    double ux[3][3] = {{0, -u[2], u[1]}, {u[2], 0, -u[0]}, {-u[1], u[0], 0}};
    const zc cf = -k * k * e * (1.0 - 1.0 / (I * k * r));
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            const zc Fab = cf * ux[a][b];
            out[a + 6 * b] = -G[a][b];             /* (0,0) block -G  :64 */
            out[a + 6 * (b + 3)] = -Fab;           /* (0,1) block -F  :65 */
            out[(a + 3) + 6 * b] = Fab;            /* (1,0) block  F  :66 */
            out[(a + 3) + 6 * (b + 3)] = -G[a][b]; /* (1,1) block -G  :67 */
        }
    return 0;
}

/* ------------------------------------------------------------------ */
/* lengths (Solver.cpp:18, BT_pad.cpp:25-40)                            */
/* ------------------------------------------------------------------ */
int64_t orc_fft_length(const int* n, int f)
{
    return (int64_t)f * f * (2 * n[0] - 1) * (int64_t)(2 * n[1] - 1) * (2 * n[2] - 1);
}

int64_t orc_bt_pad_length(const int* n, int f)
{
    return (int64_t)f * f * n[0] * (int64_t)(2 * n[1] - 1) * (2 * n[2] - 1);
}

static int64_t prod_2nm1(const int* n, int from, int M)
{
    int64_t p = 1;
    for (int l = from; l < M; ++l) p *= (2 * n[l] - 1);
    return p;
}

/* ------------------------------------------------------------------ */
/* BT_fft  (mvp_fft/BT_fft.cpp:8-53)                                    */
/* level is 1-based as in the reference; M = 3 levels.                  */
/* ------------------------------------------------------------------ */
struct af_ctx { double lf, kre, kim; };
static void af_elem(const int* n_ind, int f, void* ctx, double* out)
{
    const struct af_ctx* c = (const struct af_ctx*)ctx;
    orc_application_function(n_ind, f, c->lf, c->kre, c->kim, out);
}

static void bt_fft_rec(const int* n, int M, int f, const int* n_ind_in,
                       int level, orc_elem_fn fn, void* ctx, zc* Au)
{
    int n_ind[6];
    memcpy(n_ind, n_ind_in, sizeof(n_ind)); /* passed by value (:8) */
    if (level == M + 1) {
        /* terminal block: flipud, transpose, column-major map (:25-31):
         * Au[r + c f] = Mblk[f-1-c][r]                                  */
        zc blk[36];
        fn(n_ind, f, ctx, (double*)blk);
        for (int c = 0; c < f; ++c)
            for (int r = 0; r < f; ++r) Au[r + c * f] = blk[(f - 1 - c) + f * r];
        return;
    }
    const int this_n = n[level - 1];
    const int64_t b_edge = (int64_t)f * f * prod_2nm1(n, level, M); /* (:37) */
    /* lower triangular and diagonal blocks (:39-44) */
    for (int i = this_n; i > 0; --i) {
        n_ind[3 + level - 1] = 1;
        n_ind[level - 1] = i;
        bt_fft_rec(n, M, f, n_ind, level + 1, fn, ctx, Au + b_edge * (this_n - i));
    }
    /* upper triangular blocks (:46-50); row 0 keeps the value 1 left by the loop above */
    for (int i = 2; i <= this_n; ++i) {
        n_ind[3 + level - 1] = i;
        bt_fft_rec(n, M, f, n_ind, level + 1, fn, ctx, Au + b_edge * (i + this_n - 2));
    }
}

int orc_bt_fft_custom(const int* n, int f, orc_elem_fn fn, void* ctx, double* Au)
{
    int n_ind[6] = {1, 1, 1, 1, 1, 1}; /* MBT_fft_init.cpp:22-23 */
    if (f < 1 || f > 6) return -1;
    bt_fft_rec(n, 3, f, n_ind, 1, fn, ctx, (zc*)Au);
    return 0;
}

int orc_bt_fft(const int* n, int f, double lf, double kre, double kim, double* Au)
{
    struct af_ctx c = {lf, kre, kim};
    if (f != 3 && f != 6) return -1;
    return orc_bt_fft_custom(n, f, af_elem, &c, Au);
}

/* ------------------------------------------------------------------ */
/* BT_pad  (mvp_fft/BT_pad.cpp:6-45): returns length written            */
/* ------------------------------------------------------------------ */
static int64_t bt_pad_rec(const int* n, int M, int f, const zc* x, zc* xz)
{
    if (M < 1) { /* (:18-23) */
        for (int i = 0; i < f; ++i) xz[i] = x[i];
        for (int i = f; i < f * f; ++i) xz[i] = 0;
        return (int64_t)f * f;
    }
    int64_t b_edge = f; /* (:25) */
    for (int l = 1; l < M; ++l) b_edge *= n[l];
    const int this_n = n[0];
    int64_t len = 0;
    for (int i = 1; i <= this_n; ++i) { /* (:28-41) */
        const int64_t lenblk = bt_pad_rec(n + 1, M - 1, f, x + (i - 1) * b_edge, xz + len);
        len += lenblk;
        if (M > 1) {
            const int64_t zpad = (int64_t)(n[1] - 1) * f * f * prod_2nm1(n, 2, M);
            for (int64_t j = 0; j < zpad; ++j) xz[len + j] = 0;
            len += zpad;
        }
    }
    return len;
}

int64_t orc_bt_pad(const int* n, int f, const double* x, double* xz)
{
    return bt_pad_rec(n, 3, f, (const zc*)x, (zc*)xz);
}

/* ------------------------------------------------------------------ */
/* BT_reconstruct  (mvp_fft/BT_reconstruct.cpp:5-36)                    */
/* ------------------------------------------------------------------ */
static int64_t bt_rec_rec(const int* n, int M, int f, const zc* bz, zc* b)
{
    if (M < 1) { /* (:17-22) */
        int c = 0;
        for (int i = f; i <= f * f; i += f) b[c++] = bz[i - 1];
        return f;
    }
    const int64_t b_edge = (int64_t)f * f * prod_2nm1(n, 1, M); /* (:24) */
    const int this_n = n[0];
    int64_t len = 0;
    for (int i = this_n; i <= 2 * this_n - 1; ++i) /* (:26-33) */
        len += bt_rec_rec(n + 1, M - 1, f, bz + (int64_t)(i - 1) * b_edge, b + len);
    return len;
}

int64_t orc_bt_reconstruct(const int* n, int f, const double* bz, double* b)
{
    return bt_rec_rec(n, 3, f, (const zc*)bz, (zc*)b);
}

/* ------------------------------------------------------------------ */
/* FFT: mixed-radix decimation in time, any length                      */
/* ------------------------------------------------------------------ */
#define MAXFAC 64
#define BLUE_MIN 62
typedef struct {
    int64_t p, m;   /* prime, power-of-two convolution length */
    zc* chirp;      /* e^{-i pi (j^2 mod 2p)/p}, j < p */
    zc* hhat[2];    /* FFT_m of the conjugate chirp filter, [0]=fwd [1]=inv */
    zc* w2;         /* e^{-2 pi i j/m}, j < m/2 */
} blue_t;

typedef struct {
    int64_t n;
    int nfac;
    int64_t fac[MAXFAC], rem[MAXFAC];
    zc* t0;         /* e^{-2 pi i j/n}, j < 32768 */
    zc* t1;         /* e^{-2 pi i (j*32768)/n} */
    zc* wp[MAXFAC]; /* e^{-2 pi i e/p}, e < p, for the direct prime factor p = fac[i] */
    int nblue;
    blue_t blue[MAXFAC];
} fftplan;

static zc expi_frac(int64_t num, int64_t den) /* e^{-2 pi i num/den} */
{
    num %= den;
    if (num < 0) num += den;
    const double a = -2.0 * M_PI * (double)num / (double)den;
    return cos(a) + I * sin(a);
}

static inline zc plan_w(const fftplan* pl, int64_t i)
{
    return pl->t1[i >> 15] * pl->t0[i & 32767];
}

static void pow2_fft(zc* a, int64_t m, const zc* w2, int sign)
{
    /* iterative radix-2, bit reversal first */
    for (int64_t i = 1, j = 0; i < m; ++i) {
        int64_t bit = m >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { zc t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    for (int64_t len = 2; len <= m; len <<= 1) {
        const int64_t half = len >> 1, step = m / len;
        for (int64_t i = 0; i < m; i += len)
            for (int64_t j = 0; j < half; ++j) {
                zc w = w2[j * step];
                if (sign > 0) w = conj(w);
                const zc u = a[i + j], v = a[i + j + half] * w;
                a[i + j] = u + v;
                a[i + j + half] = u - v;
            }
    }
}

static void blue_init(blue_t* b, int64_t p)
{
    b->p = p;
    int64_t m = 1;
    while (m < 2 * p - 1) m <<= 1;
    b->m = m;
    b->chirp = (zc*)malloc(sizeof(zc) * p);
    b->w2 = (zc*)malloc(sizeof(zc) * (m / 2 + 1));
    for (int64_t j = 0; j < m / 2; ++j) b->w2[j] = expi_frac(j, m);
    for (int64_t j = 0; j < p; ++j) {
        /* e^{-i pi j^2/p} = e^{-2 pi i (j^2 mod 2p) / (2p)} */
        const int64_t q = (j * j) % (2 * p);
        b->chirp[j] = expi_frac(q, 2 * p);
    }
    for (int s = 0; s < 2; ++s) {
        zc* h = (zc*)calloc(m, sizeof(zc));
        for (int64_t j = 0; j < p; ++j) {
            const zc c = (s == 0) ? b->chirp[j] : conj(b->chirp[j]);
            h[j] = conj(c);
            if (j > 0) h[m - j] = conj(c);
        }
        pow2_fft(h, m, b->w2, -1);
        b->hhat[s] = h;
    }
}

static fftplan* plan_create(int64_t n)
{
    fftplan* pl = (fftplan*)calloc(1, sizeof(fftplan));
    pl->n = n;
    int64_t r = n;
    while (r > 1) {
        int64_t p = 0;
        if (r % 4 == 0) p = 4;
        else if (r % 2 == 0) p = 2;
        else {
            for (int64_t d = 3; d * d <= r; d += 2)
                if (r % d == 0) { p = d; break; }
            if (!p) p = r;
        }
        pl->fac[pl->nfac++] = p;
        r /= p;
    }
    if (pl->nfac == 0) { pl->fac[0] = 1; pl->nfac = 1; }
    int64_t acc = 1;
    for (int i = pl->nfac - 1; i >= 0; --i) { pl->rem[i] = acc; acc *= pl->fac[i]; }
    const int64_t nt1 = (n >> 15) + 2;
    pl->t0 = (zc*)malloc(sizeof(zc) * 32768);
    pl->t1 = (zc*)malloc(sizeof(zc) * nt1);
    for (int64_t j = 0; j < 32768; ++j) pl->t0[j] = expi_frac(j, n);
    for (int64_t j = 0; j < nt1; ++j) pl->t1[j] = expi_frac(j * 32768, n);
    for (int i = 0; i < pl->nfac; ++i) {
        const int64_t p = pl->fac[i];
        if (p == 2 || p == 4 || p >= BLUE_MIN) continue;
        pl->wp[i] = (zc*)malloc(sizeof(zc) * p);
        for (int64_t e = 0; e < p; ++e) pl->wp[i][e] = expi_frac(e, p);
    }
    for (int i = 0; i < pl->nfac; ++i) {
        const int64_t p = pl->fac[i];
        if (p < BLUE_MIN) continue;
        int found = 0;
        for (int j = 0; j < pl->nblue; ++j) if (pl->blue[j].p == p) found = 1;
        if (!found) blue_init(&pl->blue[pl->nblue++], p);
    }
    return pl;
}

static void plan_destroy(fftplan* pl)
{
    if (!pl) return;
    free(pl->t0);
    free(pl->t1);
    for (int i = 0; i < pl->nfac; ++i) free(pl->wp[i]);
    for (int j = 0; j < pl->nblue; ++j) {
        free(pl->blue[j].chirp); free(pl->blue[j].w2);
        free(pl->blue[j].hhat[0]); free(pl->blue[j].hhat[1]);
    }
    free(pl);
}

/* length-p DFT of t (in place), w_p^e = plan_w((e mod p) * n/p) */
static void dft_small(const fftplan* pl, const zc* wp, zc* t, zc* tmp, int64_t p, int sign)
{
    if (p == 2) {
        const zc a = t[0], b = t[1];
        t[0] = a + b; t[1] = a - b;
        return;
    }
    if (p == 4) {
        const zc a = t[0] + t[2], b = t[0] - t[2], c = t[1] + t[3];
        zc d = t[1] - t[3];
        d = (sign < 0) ? (-I * d) : (I * d);
        t[0] = a + c; t[2] = a - c; t[1] = b + d; t[3] = b - d;
        return;
    }
    if (p >= BLUE_MIN) {
        const blue_t* b = NULL;
        for (int j = 0; j < pl->nblue; ++j) if (pl->blue[j].p == p) b = &pl->blue[j];
        const int s = (sign < 0) ? 0 : 1;
        const int64_t m = b->m;
        for (int64_t j = 0; j < m; ++j) tmp[j] = 0;
        for (int64_t j = 0; j < p; ++j) tmp[j] = t[j] * (s ? conj(b->chirp[j]) : b->chirp[j]);
        pow2_fft(tmp, m, b->w2, -1);
        for (int64_t j = 0; j < m; ++j) tmp[j] *= b->hhat[s][j];
        pow2_fft(tmp, m, b->w2, +1);
        for (int64_t j = 0; j < p; ++j) {
            const zc c = s ? conj(b->chirp[j]) : b->chirp[j];
            t[j] = tmp[j] * c / (double)m;
        }
        return;
    }
    /* odd prime p: pair x_q with x_{p-q}.  With s_q = x_q + x_{p-q},
     * d_q = x_q - x_{p-q} (q = 1 .. h, h = (p-1)/2) and W = e^{-2 pi i/p},
     *   y_k = x_0 + sum_q s_q cos(2 pi q k/p) -+ i sum_q d_q sin(2 pi q k/p),
     *   y_{p-k} likewise with the sign of the sine sum flipped,
     * real coefficients only (wp holds cos - i sin). */
    const int64_t h = (p - 1) / 2;
    const zc x0 = t[0];
    zc* sq = tmp;
    zc* dq = tmp + h;
    zc y0 = x0;
    for (int64_t q = 1; q <= h; ++q) {
        sq[q - 1] = t[q] + t[p - q];
        dq[q - 1] = t[q] - t[p - q];
        y0 += sq[q - 1];
    }
    for (int64_t k = 1; k <= h; ++k) {
        double ar = creal(x0), ai = cimag(x0), br = 0, bi = 0;
        int64_t e = 0;          /* q k mod p */
        for (int64_t q = 1; q <= h; ++q) {
            e += k;
            if (e >= p) e -= p;
            const double c = creal(wp[e]), sn = -cimag(wp[e]);
            ar += creal(sq[q - 1]) * c;
            ai += cimag(sq[q - 1]) * c;
            br += creal(dq[q - 1]) * sn;
            bi += cimag(dq[q - 1]) * sn;
        }
        /* forward: y_k = a - i b; inverse: y_k = a + i b (b = sum d sin) */
        const double sg = sign > 0 ? 1.0 : -1.0;
        t[k] = (ar - sg * bi) + I * (ai + sg * br);
        t[p - k] = (ar + sg * bi) + I * (ai - sg * br);
    }
    t[0] = y0;
}

static void fft_rec(const fftplan* pl, zc* out, const zc* in, int64_t fstride,
                    int lev, int sign, zc* t, zc* tmp)
{
    const int64_t p = pl->fac[lev], m = pl->rem[lev];
    if (m == 1) {
        for (int64_t j = 0; j < p; ++j) out[j] = in[j * fstride];
    } else {
        for (int64_t j = 0; j < p; ++j)
            fft_rec(pl, out + j * m, in + j * fstride, fstride * p, lev + 1, sign, t, tmp);
    }
    for (int64_t u = 0; u < m; ++u) {
        for (int64_t q = 0; q < p; ++q) {
            zc v = out[q * m + u];
            if (u && q) {
                zc w = plan_w(pl, q * u * fstride);
                v *= (sign > 0) ? conj(w) : w;
            }
            t[q] = v;
        }
        dft_small(pl, pl->wp[lev], t, tmp, p, sign);
        for (int64_t q = 0; q < p; ++q) out[u + q * m] = t[q];
    }
}

static int64_t plan_scratch(const fftplan* pl)
{
    int64_t s = 64;
    for (int i = 0; i < pl->nfac; ++i) if (pl->fac[i] > s) s = pl->fac[i];
    for (int j = 0; j < pl->nblue; ++j) if (pl->blue[j].m > s) s = pl->blue[j].m;
    return s;
}

/* Four-step FFT (D. H. Bailey, "FFTs in external or hierarchical memory",
 * J. Supercomputing 4, 23-35 (1990)) for long transforms: with n = n1 n2,
 * j = j1 n2 + j2 and k = k1 + n1 k2,
 *   X[k1 + n1 k2] = sum_j2 W_n2^(j2 k2) W_n^(j2 k1) sum_j1 W_n1^(j1 k1) x[j1 n2 + j2].
 * Columns j2 are gathered BLK at a time (contiguous runs of x), transformed
 * with the length-n1 plan, twiddled and stored row-major; rows k1 are then
 * transformed with the length-n2 plan and scattered BLK at a time.  The
 * sub-transforms stay in cache, and both loops run on g_fft_threads OpenMP
 * threads when called outside a parallel region (one thread inside one, as
 * in the block matvec's loop over columns). */
static int64_t g_fourstep_min = (int64_t)1 << 24;
static int g_fft_threads = 1;

void orc_fft_config(int64_t fourstep_min, int threads)
{
    g_fourstep_min = fourstep_min > 0 ? fourstep_min : ((int64_t)1 << 62);
    g_fft_threads = threads > 0 ? threads : 1;
}

/* n1, n2 > 1 with n1 n2 = n from the plan's factors, as close to sqrt(n) as
 * a greedy split gets (0 when n is prime) */
static int fourstep_split(const fftplan* pl, int64_t* n1, int64_t* n2)
{
    int64_t f[MAXFAC];
    int nf = pl->nfac;
    for (int i = 0; i < nf; ++i) f[i] = pl->fac[i];
    for (int i = 0; i < nf; ++i)         /* descending */
        for (int j = i + 1; j < nf; ++j)
            if (f[j] > f[i]) { int64_t t = f[i]; f[i] = f[j]; f[j] = t; }
    int64_t a = 1, b = 1;
    for (int i = 0; i < nf; ++i) {
        if (a <= b) a *= f[i];
        else b *= f[i];
    }
    if (a < 2 || b < 2) return 0;
    *n1 = a;
    *n2 = b;
    return 1;
}

static void fft_fourstep(const fftplan* pl, int64_t n1, int64_t n2, const zc* in, zc* out, int sign)
{
    enum { BLK = 8 };
    const int64_t n = pl->n;
    fftplan* p1 = plan_create(n1);
    fftplan* p2 = plan_create(n2);
    zc* T = (zc*)malloc(sizeof(zc) * n);
    const int64_t nmax = n1 > n2 ? n1 : n2;
    int64_t sc = plan_scratch(p1);
    if (plan_scratch(p2) > sc) sc = plan_scratch(p2);
    const int nt = omp_in_parallel() ? 1 : g_fft_threads;
#pragma omp parallel num_threads(nt)
    {
        zc* col = (zc*)malloc(sizeof(zc) * BLK * nmax);
        zc* res = (zc*)malloc(sizeof(zc) * BLK * nmax);
        zc* t = (zc*)malloc(sizeof(zc) * sc);
        zc* tmp = (zc*)malloc(sizeof(zc) * sc);
#pragma omp for schedule(dynamic, 1)
        for (int64_t jb = 0; jb < n2; jb += BLK) {
            const int nb = (int)(n2 - jb < BLK ? n2 - jb : BLK);
            for (int64_t j1 = 0; j1 < n1; ++j1)
                for (int b = 0; b < nb; ++b) col[b * n1 + j1] = in[j1 * n2 + jb + b];
            for (int b = 0; b < nb; ++b) fft_rec(p1, res + b * n1, col + b * n1, 1, 0, sign, t, tmp);
            for (int64_t k1 = 0; k1 < n1; ++k1)
                for (int b = 0; b < nb; ++b) {
                    zc w = plan_w(pl, (jb + b) * k1);     /* (jb + b) k1 < n2 n1 = n */
                    if (sign > 0) w = conj(w);
                    T[k1 * n2 + jb + b] = res[b * n1 + k1] * w;
                }
        }
#pragma omp for schedule(dynamic, 1)
        for (int64_t kb = 0; kb < n1; kb += BLK) {
            const int nb = (int)(n1 - kb < BLK ? n1 - kb : BLK);
            for (int b = 0; b < nb; ++b) fft_rec(p2, res + b * n2, T + (kb + b) * n2, 1, 0, sign, t, tmp);
            for (int64_t k2 = 0; k2 < n2; ++k2)
                for (int b = 0; b < nb; ++b) out[kb + b + n1 * k2] = res[b * n2 + k2];
        }
        free(col);
        free(res);
        free(t);
        free(tmp);
    }
    free(T);
    plan_destroy(p1);
    plan_destroy(p2);
}

/* out-of-place execute: out must not alias in */
static void plan_exec(const fftplan* pl, const zc* in, zc* out, int sign)
{
    int64_t n1, n2;
    if (pl->n >= g_fourstep_min && fourstep_split(pl, &n1, &n2)) {
        fft_fourstep(pl, n1, n2, in, out, sign);
        return;
    }
    const int64_t s = plan_scratch(pl);
    zc* t = (zc*)malloc(sizeof(zc) * s);
    zc* tmp = (zc*)malloc(sizeof(zc) * s);
    fft_rec(pl, out, in, 1, 0, sign, t, tmp);
    free(t);
    free(tmp);
}

int orc_fft(int64_t len, double* data, int sign)
{
    if (len < 1) return -1;
    fftplan* pl = plan_create(len);
    zc* out = (zc*)malloc(sizeof(zc) * len);
    if (!out) { plan_destroy(pl); return -2; }
    plan_exec(pl, (const zc*)data, out, sign);
    memcpy(data, out, sizeof(zc) * len);
    free(out);
    plan_destroy(pl);
    return 0;
}

/* ------------------------------------------------------------------ */
/* MBT_fft_init (mvp_fft/MBT_fft_init.cpp:10-42)                        */
/* ------------------------------------------------------------------ */
int orc_mbt_fft_init(const int* n, int f, double lf, double kre, double kim, double* Au_til)
{
    const int64_t F = orc_fft_length(n, f);
    zc* Au = (zc*)malloc(sizeof(zc) * F);
    if (!Au) return -2;
    int rc = orc_bt_fft(n, f, lf, kre, kim, (double*)Au);
    if (rc) { free(Au); return rc; }
    zc* out = (zc*)Au_til;
    /* Au.reverse() (:34) then forward FFT (:35) */
    for (int64_t i = 0; i < F / 2; ++i) { zc t = Au[i]; Au[i] = Au[F - 1 - i]; Au[F - 1 - i] = t; }
    fftplan* pl = plan_create(F);
    plan_exec(pl, Au, out, -1);
    plan_destroy(pl);
    free(Au);
    return 0;
}

/* ------------------------------------------------------------------ */
/* MBT_fft_mvp (mvp_fft/MBT_fft_mvp.cpp:8-43), plan supplied by caller   */
/* ------------------------------------------------------------------ */
static int mbt_mvp_with_plan(const fftplan* pl, const int* n, int f, const zc* Au_til,
                             const zc* p_hat, zc* q_hat)
{
    const int64_t F = pl->n;
    zc* a = (zc*)malloc(sizeof(zc) * F);
    zc* b = (zc*)malloc(sizeof(zc) * F);
    if (!a || !b) { free(a); free(b); return -2; }
    const int nt = omp_in_parallel() ? 1 : g_fft_threads;
    const int64_t lp = bt_pad_rec(n, 3, f, p_hat, a);  /* (:21) */
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int64_t i = lp; i < F; ++i) a[i] = 0;         /* (:23-25) */
    plan_exec(pl, a, b, -1);                            /* (:27) */
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int64_t i = 0; i < F; ++i) b[i] *= Au_til[i];  /* (:28) */
    plan_exec(pl, b, a, +1);                            /* (:29) */
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int64_t i = 0; i < F; ++i) a[i] /= (double)F;  /* (:30) */
    bt_rec_rec(n, 3, f, a, q_hat);                      /* (:31) */
    free(a);
    free(b);
    return 0;
}

int orc_mbt_fft_mvp(const int* n, int f, const double* Au_til, const double* p_hat, double* q_hat)
{
    fftplan* pl = plan_create(orc_fft_length(n, f));
    int rc = mbt_mvp_with_plan(pl, n, f, (const zc*)Au_til, (const zc*)p_hat, (zc*)q_hat);
    plan_destroy(pl);
    return rc;
}

/* ------------------------------------------------------------------ */
/* dense helpers (column-major)                                         */
/* ------------------------------------------------------------------ */
/* C(a x b) = op(A)(a x m) * B(m x b), op = transpose (conjA=0) or adjoint */
static void gram(const zc* A, const zc* B, int64_t m, int a, int b, int conjA, zc* C)
{
    for (int j = 0; j < b; ++j)
        for (int i = 0; i < a; ++i) {
            zc s = 0;
            const zc* Ai = A + (int64_t)i * m;
            const zc* Bj = B + (int64_t)j * m;
            if (conjA) for (int64_t r = 0; r < m; ++r) s += conj(Ai[r]) * Bj[r];
            else for (int64_t r = 0; r < m; ++r) s += Ai[r] * Bj[r];
            C[i + (int64_t)j * a] = s;
        }
}

/* C(m x n) = A(m x k) * B(k x n) */
static void gemm(const zc* A, const zc* B, int64_t m, int k, int n, zc* C)
{
    for (int j = 0; j < n; ++j) {
        zc* Cj = C + (int64_t)j * m;
        for (int64_t r = 0; r < m; ++r) Cj[r] = 0;
        for (int l = 0; l < k; ++l) {
            const zc bl = B[l + (int64_t)j * k];
            const zc* Al = A + (int64_t)l * m;
            for (int64_t r = 0; r < m; ++r) Cj[r] += Al[r] * bl;
        }
    }
}

static double fro(const zc* A, int64_t len)
{
    double s = 0;
    for (int64_t i = 0; i < len; ++i) s += creal(A[i]) * creal(A[i]) + cimag(A[i]) * cimag(A[i]);
    return sqrt(s);
}

/* X = A^{-1} B with partial (full=0) or full (full=1) pivoting; A is s x s,
 * B is s x nb.  Mirrors Eigen's PartialPivLU / FullPivLU solve. */
static void lu_solve(const zc* A_, const zc* B, int s, int nb, int full, zc* X)
{
    zc* A = (zc*)malloc(sizeof(zc) * s * s);
    int* rp = (int*)malloc(sizeof(int) * s);
    int* cp = (int*)malloc(sizeof(int) * s);
    memcpy(A, A_, sizeof(zc) * s * s);
    for (int i = 0; i < s; ++i) { rp[i] = i; cp[i] = i; }
    for (int k = 0; k < s; ++k) {
        int pr = k, pc = k;
        double best = -1;
        for (int j = k; j < (full ? s : k + 1); ++j)
            for (int i = k; i < s; ++i) {
                const double v = cabs(A[i + j * s]);
                if (v > best) { best = v; pr = i; pc = j; }
            }
        if (pr != k) {
            for (int j = 0; j < s; ++j) { zc t = A[k + j * s]; A[k + j * s] = A[pr + j * s]; A[pr + j * s] = t; }
            int t = rp[k]; rp[k] = rp[pr]; rp[pr] = t;
        }
        if (pc != k) {
            for (int i = 0; i < s; ++i) { zc t = A[i + k * s]; A[i + k * s] = A[i + pc * s]; A[i + pc * s] = t; }
            int t = cp[k]; cp[k] = cp[pc]; cp[pc] = t;
        }
        const zc piv = A[k + k * s];
        for (int i = k + 1; i < s; ++i) {
            const zc l = A[i + k * s] / piv;
            A[i + k * s] = l;
            for (int j = k + 1; j < s; ++j) A[i + j * s] -= l * A[k + j * s];
        }
    }
    zc* y = (zc*)malloc(sizeof(zc) * s);
    for (int c = 0; c < nb; ++c) {
        for (int i = 0; i < s; ++i) y[i] = B[rp[i] + c * s];
        for (int i = 0; i < s; ++i)
            for (int j = 0; j < i; ++j) y[i] -= A[i + j * s] * y[j];
        for (int i = s - 1; i >= 0; --i) {
            for (int j = i + 1; j < s; ++j) y[i] -= A[i + j * s] * y[j];
            y[i] /= A[i + i * s];
        }
        for (int i = 0; i < s; ++i) X[cp[i] + c * s] = y[i];
    }
    free(y); free(A); free(rp); free(cp);
}

/* ------------------------------------------------------------------ */
/* qr_reduced (bl_krylov_mvp_fft/qr_reduced.cpp:9-31)                  */
/* MGS with one re-orthogonalisation pass; R keeps first-pass           */
/* coefficients (the tt == 0 test of :22 is true only in pass 0).       */
/* ------------------------------------------------------------------ */
static void qr_reduced_z(const zc* A, int64_t m, int n, zc* Q, zc* R)
{
    memcpy(Q, A, sizeof(zc) * m * n);
    for (int i = 0; i < n * n; ++i) R[i] = 0;
    for (int k = 0; k < n; ++k) {
        double tt = 0.0;
        zc* qk = Q + (int64_t)k * m;
        for (int j = 0; j < 2; ++j) {
            for (int i = 0; i < k; ++i) {
                const zc* qi = Q + (int64_t)i * m;
                zc s = 0;
                for (int64_t r = 0; r < m; ++r) s += conj(qi[r]) * qk[r];
                if (tt == 0.0) R[i + k * n] = s;
                for (int64_t r = 0; r < m; ++r) qk[r] -= s * qi[r];
            }
            tt = fro(qk, m);
        }
        R[k + k * n] = tt;
        for (int64_t r = 0; r < m; ++r) qk[r] /= tt;
    }
}

void orc_lu_solve(const double* A, const double* B, int s, int nb, int full, double* X)
{
    lu_solve((const zc*)A, (const zc*)B, s, nb, full, (zc*)X);
}

int orc_qr_reduced(const double* A, int64_t m, int ncol, double* Q, double* R)
{
    qr_reduced_z((const zc*)A, m, ncol, (zc*)Q, (zc*)R);
    return 0;
}

/* ------------------------------------------------------------------ */
/* inline block matvec (e.g. bl_cocg_rq_jacobi_mvp_fft.cpp:43-60)       */
/* ------------------------------------------------------------------ */
typedef struct {
    const int* n; int f; const int* address; int64_t nocc;
    const zc* Au_til; const zc* diag; zc* jpre; int jacobi;
    fftplan* pl; int nthreads;
} bmv_t;

static int bmv_apply(const bmv_t* c, const zc* S, zc* AS, int L)
{
    const int f = c->f;
    const int64_t C = (int64_t)c->n[0] * c->n[1] * c->n[2];
    const int64_t rows = (int64_t)f * c->nocc;
    int err = 0;
#pragma omp parallel for num_threads(c->nthreads) schedule(dynamic, 1)
    for (int l = 0; l < L; ++l) {
        zc* phat = (zc*)calloc(f * C, sizeof(zc));
        zc* qhat = (zc*)malloc(sizeof(zc) * f * C);
        if (!phat || !qhat) { err = -2; free(phat); free(qhat); continue; }
        const zc* Sl = S + l * rows;
        zc* Yl = AS + l * rows;
        for (int64_t m = 0; m < c->nocc; ++m)
            for (int q = 0; q < f; ++q) phat[(int64_t)f * c->address[m] + q] = Sl[f * m + q];
        int rc = mbt_mvp_with_plan(c->pl, c->n, f, c->Au_til, phat, qhat);
        if (rc) err = rc;
        for (int64_t m = 0; m < c->nocc; ++m)
            for (int q = 0; q < f; ++q) {
                const int64_t i = f * m + q;
                zc v = c->diag[i] * Sl[i];
                v += qhat[(int64_t)f * c->address[m] + q];
                if (c->jacobi) v *= c->jpre[i];
                Yl[i] = v;
            }
        free(phat);
        free(qhat);
    }
    return err;
}

static int bmv_setup(bmv_t* c, const int* n, int f, const int* address, int64_t nocc,
                     const zc* Au_til, const zc* diag, int jacobi, int nthreads)
{
    c->n = n; c->f = f; c->address = address; c->nocc = nocc;
    c->Au_til = Au_til; c->diag = diag; c->jacobi = jacobi;
    c->nthreads = nthreads > 0 ? nthreads : 1;
    c->jpre = (zc*)malloc(sizeof(zc) * f * nocc);
    for (int64_t i = 0; i < (int64_t)f * nocc; ++i) c->jpre[i] = 1.0 / diag[i]; /* jpre = 1/DIAG_A */
    c->pl = plan_create(orc_fft_length(n, f));
    return 0;
}

static void bmv_free(bmv_t* c)
{
    free(c->jpre);
    plan_destroy(c->pl);
}

int orc_block_matvec(const int* n, int f, const int* address, int64_t nocc,
                     const double* Au_til, const double* diag, int jacobi,
                     const double* X, double* Y, int L, int nthreads)
{
    bmv_t c;
    bmv_setup(&c, n, f, address, nocc, (const zc*)Au_til, (const zc*)diag, jacobi, nthreads);
    int rc = bmv_apply(&c, (const zc*)X, (zc*)Y, L);
    bmv_free(&c);
    return rc;
}

/* ------------------------------------------------------------------ */
/* solvers                                                              */
/* ------------------------------------------------------------------ */
#define ZALLOC(len) ((zc*)calloc((size_t)(len), sizeof(zc)))

/* bl_cocg_rq_mvp_fft.cpp:12-72 and bl_cocg_rq_jacobi_mvp_fft.cpp:12-79 */
static int solve_cocg_rq(bmv_t* A, const zc* B, int64_t rows, int L, double tol,
                         int itermax, zc* X, int* iters, double* hist)
{
    const int s = L;
    zc* Bp = ZALLOC(rows * L);
    for (int64_t i = 0; i < rows * L; ++i)
        Bp[i] = A->jacobi ? B[i] * A->jpre[i % rows] : B[i];   /* B_pre (:22-27) */
    const double Bnorm = fro(Bp, rows * L);
    zc *Q = ZALLOC(rows * L), *S = ZALLOC(rows * L), *AS = ZALLOC(rows * L);
    zc *Qn = ZALLOC(rows * L), *W = ZALLOC(rows * L), *T = ZALLOC(rows * L);
    zc *xi = ZALLOC(s * s), *tau = ZALLOC(s * s), *alpha = ZALLOC(s * s), *beta = ZALLOC(s * s);
    zc *G1 = ZALLOC(s * s), *G2 = ZALLOC(s * s), *G3 = ZALLOC(s * s), *t1 = ZALLOC(s * s);
    for (int64_t i = 0; i < rows * L; ++i) X[i] = 0;
    qr_reduced_z(Bp, rows, s, Q, xi);
    memcpy(S, Q, sizeof(zc) * rows * L);
    const int full = A->jacobi ? 0 : 1; /* FullPivLU (:56,66) vs PartialPivLU (:63,73) */
    int k, rc = 0;
    for (k = 0; k < itermax; ++k) {
        if ((rc = bmv_apply(A, S, AS, L))) break;
        gram(S, AS, rows, s, s, 0, G1);               /* S^T AS */
        gram(Q, Q, rows, s, s, 0, G2);                /* Q^T Q */
        lu_solve(G1, G2, s, s, full, alpha);
        gemm(S, alpha, rows, s, s, T);                /* (S alpha) xi */
        gemm(T, xi, rows, s, s, W);
        for (int64_t i = 0; i < rows * L; ++i) X[i] += W[i];
        gemm(AS, alpha, rows, s, s, T);               /* Q - AS alpha */
        for (int64_t i = 0; i < rows * L; ++i) T[i] = Q[i] - T[i];
        qr_reduced_z(T, rows, s, Qn, tau);
        gemm(tau, xi, s, s, s, t1);
        memcpy(xi, t1, sizeof(zc) * s * s);
        const double err = fro(xi, s * s) / Bnorm;
        if (hist) hist[k] = err;
        if (err < tol) { ++k; break; }
        gram(Qn, Qn, rows, s, s, 0, G3);              /* Qnew^T Qnew */
        for (int j = 0; j < s; ++j)                   /* tau^T (Qnew^T Qnew) */
            for (int i = 0; i < s; ++i) {
                zc acc = 0;
                for (int l = 0; l < s; ++l) acc += tau[l + i * s] * G3[l + j * s];
                t1[i + j * s] = acc;
            }
        lu_solve(G2, t1, s, s, full, beta);
        memcpy(Q, Qn, sizeof(zc) * rows * L);
        gemm(S, beta, rows, s, s, T);                 /* S = Q + S beta */
        for (int64_t i = 0; i < rows * L; ++i) S[i] = Q[i] + T[i];
    }
    if (iters) *iters = k;
    free(Bp); free(Q); free(S); free(AS); free(Qn); free(W); free(T);
    free(xi); free(tau); free(alpha); free(beta); free(G1); free(G2); free(G3); free(t1);
    return rc;
}

/* bl_bicgstab_mvp_fft.cpp:12-101 and bl_bicgstab_jacobi_mvp_fft.cpp:12-97 */
static int solve_bicgstab(bmv_t* A, const zc* B, int64_t rows, int L, double tol,
                          int itermax, zc* X, int* iters, double* hist)
{
    const int s = L;
    const int64_t N = rows * L;
    zc *R = ZALLOC(N), *P = ZALLOC(N), *R0 = ZALLOC(N), *V = ZALLOC(N);
    zc *T = ZALLOC(N), *Z = ZALLOC(N), *tmp = ZALLOC(N);
    zc *M1 = ZALLOC(s * s), *M2 = ZALLOC(s * s), *alpha = ZALLOC(s * s), *beta = ZALLOC(s * s);
    for (int64_t i = 0; i < N; ++i) R[i] = A->jacobi ? B[i] * A->jpre[i % rows] : B[i];
    const double Bnorm = fro(R, N);
    memcpy(P, R, sizeof(zc) * N);
    memcpy(R0, R, sizeof(zc) * N);
    for (int64_t i = 0; i < N; ++i) X[i] = 0;
    int k, rc = 0;
    for (k = 0; k < itermax; ++k) {
        if ((rc = bmv_apply(A, P, V, L))) break;       /* V = A P */
        gram(R0, V, rows, s, s, 1, M1);                /* R0^H V (LU kept for beta) */
        gram(R0, R, rows, s, s, 1, M2);
        lu_solve(M1, M2, s, s, 0, alpha);
        gemm(V, alpha, rows, s, s, tmp);               /* T = R - V alpha */
        for (int64_t i = 0; i < N; ++i) T[i] = R[i] - tmp[i];
        if ((rc = bmv_apply(A, T, Z, L))) break;       /* Z = A T */
        zc zt = 0, zz = 0;
        for (int64_t i = 0; i < N; ++i) { zt += conj(Z[i]) * T[i]; zz += conj(Z[i]) * Z[i]; }
        const zc qsi = zt / zz;                        /* (:87) */
        gemm(P, alpha, rows, s, s, tmp);               /* X += P alpha + qsi T */
        for (int64_t i = 0; i < N; ++i) X[i] = X[i] + tmp[i] + qsi * T[i];
        for (int64_t i = 0; i < N; ++i) R[i] = T[i] - qsi * Z[i];
        const double err = fro(R, N) / Bnorm;
        if (hist) hist[k] = err;
        if (err < tol) { ++k; break; }
        gram(R0, Z, rows, s, s, 1, M2);                /* beta = LU^{-1}(-R0^H Z) */
        for (int i = 0; i < s * s; ++i) M2[i] = -M2[i];
        lu_solve(M1, M2, s, s, 0, beta);
        for (int64_t i = 0; i < N; ++i) T[i] = P[i] - qsi * V[i];  /* P = R + (P - qsi V) beta */
        gemm(T, beta, rows, s, s, tmp);
        for (int64_t i = 0; i < N; ++i) P[i] = R[i] + tmp[i];
    }
    if (iters) *iters = k;
    free(R); free(P); free(R0); free(V); free(T); free(Z); free(tmp);
    free(M1); free(M2); free(alpha); free(beta);
    return rc;
}

/* bl_bicggr_mvp_fft.cpp:10-119 and bl_bicggr_jacobi_mvp_fft.cpp:10-120 */
static int solve_bicggr(bmv_t* A, const zc* B, int64_t rows, int L, double tol,
                        int itermax, zc* X, int* iters, double* hist)
{
    const int s = L;
    const int64_t N = rows * L;
    zc *R = ZALLOC(N), *P = ZALLOC(N), *R0 = ZALLOC(N), *V = ZALLOC(N), *W = ZALLOC(N);
    zc *U = ZALLOC(N), *Y = ZALLOC(N), *Rn = ZALLOC(N), *tmp = ZALLOC(N);
    zc *M1 = ZALLOC(s * s), *M2 = ZALLOC(s * s), *alpha = ZALLOC(s * s), *gamma = ZALLOC(s * s);
    for (int64_t i = 0; i < N; ++i) R[i] = A->jacobi ? B[i] * A->jpre[i % rows] : B[i];
    const double Bnorm = fro(R, N);
    memcpy(P, R, sizeof(zc) * N);
    for (int64_t i = 0; i < N; ++i) X[i] = 0;
    int k = 0, rc = bmv_apply(A, R, V, L);             /* V = A R (:22-43) */
    memcpy(W, V, sizeof(zc) * N);
    memcpy(R0, R, sizeof(zc) * N);
    for (k = 0; k < itermax && !rc; ++k) {
        gram(R0, V, rows, s, s, 1, M1);
        gram(R0, R, rows, s, s, 1, M2);
        lu_solve(M1, M2, s, s, 0, alpha);              /* (:51) */
        zc wr = 0, ww = 0;
        for (int64_t i = 0; i < N; ++i) { wr += conj(W[i]) * R[i]; ww += conj(W[i]) * W[i]; }
        const zc qsi = wr / ww;                        /* (:52) */
        for (int64_t i = 0; i < N; ++i) tmp[i] = P[i] - qsi * V[i];   /* S */
        gemm(tmp, alpha, rows, s, s, U);               /* U = S alpha */
        if ((rc = bmv_apply(A, U, Y, L))) break;       /* Y = A U */
        for (int64_t i = 0; i < N; ++i) X[i] = X[i] + qsi * R[i] + U[i];
        for (int64_t i = 0; i < N; ++i) Rn[i] = R[i] - qsi * W[i] - Y[i];
        const double err = fro(Rn, N) / Bnorm;
        if (hist) hist[k] = err;
        if (err < tol) { ++k; break; }
        if ((rc = bmv_apply(A, Rn, W, L))) break;      /* W = A Rnew */
        gram(R0, R, rows, s, s, 1, M1);                /* gamma = (R0^H R)^{-1} R0^H Rnew / qsi */
        gram(R0, Rn, rows, s, s, 1, M2);               /* (R0^H Rnew) / qsi */
        for (int i = 0; i < s * s; ++i) M2[i] /= qsi;
        lu_solve(M1, M2, s, s, 0, gamma);
        memcpy(R, Rn, sizeof(zc) * N);
        gemm(U, gamma, rows, s, s, tmp);
        for (int64_t i = 0; i < N; ++i) P[i] = R[i] + tmp[i];
        gemm(Y, gamma, rows, s, s, tmp);
        for (int64_t i = 0; i < N; ++i) V[i] = W[i] + tmp[i];
    }
    if (iters) *iters = k;
    free(R); free(P); free(R0); free(V); free(W); free(U); free(Y); free(Rn); free(tmp);
    free(M1); free(M2); free(alpha); free(gamma);
    return rc;
}

int orc_solve(int method, int jacobi, const int* n, int f, const int* address,
              int64_t nocc, const double* Au_til, const double* diag,
              const double* B, int L, double tol, int itermax, double* X,
              int* iters, double* hist, int nthreads)
{
    if (f != 3 && f != 6) return -1;
    bmv_t A;
    bmv_setup(&A, n, f, address, nocc, (const zc*)Au_til, (const zc*)diag, jacobi, nthreads);
    const int64_t rows = (int64_t)f * nocc;
    int rc;
    switch (method) {
    case ORC_COCG_RQ: rc = solve_cocg_rq(&A, (const zc*)B, rows, L, tol, itermax, (zc*)X, iters, hist); break;
    case ORC_BICGSTAB: rc = solve_bicgstab(&A, (const zc*)B, rows, L, tol, itermax, (zc*)X, iters, hist); break;
    case ORC_BICGGR: rc = solve_bicggr(&A, (const zc*)B, rows, L, tol, itermax, (zc*)X, iters, hist); break;
    default: rc = -1;
    }
    bmv_free(&A);
    return rc;
}
