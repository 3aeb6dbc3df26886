/*
 * sht_oracle.c -- CPU FP64 restatement of SHTns's Gauss-grid transforms.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py.  Nothing in the product links or loads it.
 *
 * What each part restates (citations into the SHTns 3.5.2 sources):
 *   gauss_nodes_()      sht_legendre.c:763-811   Newton iteration on P_n, weights 2(1-z^2)/pp^2
 *   legendre_coeffs_()  sht_legendre.c:584-613   a_mm seed (compensated product, CS phase)
 *                       sht_legendre.c:646-657   classic orthonormal a_l^m, b_l^m
 *   leg_column_()       sht_legendre.c:24-31     y_l = a_l x y_{l-1} + b_l y_{l-2}; the
 *                       kernel_SH_to_spat.gen.c:278-348 extended-exponent start for
 *                       large m (SHTns rescales by 2.9e135; here by powers of two, exact)
 *   synthesis           kernel_SH_to_spat.gen.c:94-202, 345-441 (even/odd parity split,
 *                       north = E+O, south = E-O) + inverse real DFT (sht_init.c:353-518)
 *   analysis            kernel_spat_to_SH.gen.c:89-192, 471-616 with the quadrature
 *                       weights w_k*2pi/nphi (sht_init.c:566-572, 626-627)
 *   vector synthesis    sht_private.h:534-584 (S,T -> V,W through the sin(theta)d/dtheta
 *                       tridiagonal, sht_func.c:400-448), V,W synthesized with
 *                       Y_lm/sin(theta); m=0 via dY_l0/dtheta (kernel_SH_to_spat.gen.c:94-202)
 *   vector analysis     sht_private.h:296-337 (SH_2scal_to_vect with mx_van = 2*ct + st_dt,
 *                       sht_init.c:1382-1386), m=0 via dY/dtheta (kernel_spat_to_SH.gen.c:89-192)
 *   orc_set_polar_opt() sht_init.c:522-557 (PolarOptimize), applied as in
 *                       kernel_SH_to_spat.gen.c:235-257 and kernel_spat_to_SH.gen.c:198-201
 *
 * This restatement uses the classic three-term recurrence for every m (as SHTns's
 * GPU kernels do, sht_gpu_kernels.cu:800-833), not the Ishioka variant of the CPU
 * kernels: both generate the same functions, and keeping the oracle on a different
 * recurrence from the HIP kernels makes agreement between the two meaningful.
 */
#include "sht_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

struct orc_cfg {
    int lmax, mmax, mres, nlat, nlat_2, nphi;
    long nlm;
    double *ct, *st, *w;      /* nlat each */
    double *wp;               /* pair weights: w, halved at the middle row of an odd grid */
    int mtab_n;               /* tables for m = 0..mtab_n-1 (only needed ones filled) */
    double **seed;            /* seed[m] = a_mm (1 value) */
    double **ab;              /* ab[m][2*(l-m)] = b_l, [2*(l-m)+1] = a_l ; l = m..lmax+1 */
    int *tm;                  /* [mmax+1] first latitude pair of order im (polar optimization) */
};

static int g_threads = 0;
void orc_set_threads(int n) { g_threads = n; }

static long lm_index(const orc_cfg *c, int l, int im) {
    return ((long)im * (2 * c->lmax + 2 - (im + 1) * c->mres)) / 2 + l;
}

long orc_nlm(const orc_cfg *c) { return c->nlm; }

/* ---- Gauss-Legendre nodes (sht_legendre.c:763-811) ---- */
static void gauss_nodes_(double *x, double *st, double *w, int n) {
    const double eps = 2.3e-16;
    const long m = n / 2;
    for (long i = 0; i < m; ++i) {
        double z = (1.0 - (n - 1.) / (8. * n * n * n)) * cos((M_PI * (4 * i + 3)) / (4 * n + 2));
        double z1, pp, p1, p2;
        int k = 10;
        do {
            p1 = z;
            p2 = 1.0;
            for (long l = 2; l <= n; ++l) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2 * l - 1) * z * p2 - (l - 1) * p3) / l;
            }
            pp = n * (p2 - z * p1);
            z1 = z;
            z -= p1 * (1. - z * z) / pp;
        } while ((fabs(z - z1) > (z1 + z) * 0.5 * eps) && (--k > 0));
        double s2 = 1. - z * z;
        x[i] = z;
        x[n - 1 - i] = -z;
        w[i] = 2.0 * s2 / (pp * pp);
        w[n - 1 - i] = w[i];
        st[i] = sqrt(s2);
        st[n - 1 - i] = st[i];
    }
    if (n & 1) {   /* odd n: the middle root x = 0, weight 2 / (n P_{n-1}(0))^2 */
        double p1 = 0.0, p2 = 1.0;   /* P_1(0), P_0(0) */
        for (long l = 2; l < n; ++l) {
            double p3 = p2;
            p2 = p1;
            p1 = ((2 * l - 1) * 0.0 * p2 - (l - 1) * p3) / l;
        }
        /* p1 = P_{n-1}(0) */
        x[m] = 0.0;
        st[m] = 1.0;
        w[m] = 2.0 / ((n * p1) * (n * p1));
    }
}

/* ---- recurrence coefficients for one order m (sht_legendre.c:584-613, 646-657) ---- */
static double seed_amm_(int m) {
    /* Y_m^m prefactor: sqrt( prod_{k=1..m} (2k+1)/(2k) / (4 pi) ) * (-1)^m, with the
       compensated product so the seed keeps ~1 ulp accuracy at large m. */
    double t1 = 0.25 / M_PI, e = 0.0;
    for (int k = 1; k <= m; ++k) {
        double x = ((double)k + 0.5) / k;
        double tt = t1 * x;
        double t1e = fma(t1, x, -tt);
        t1 = tt;
        e = fma(e, x, t1e);
    }
    double t2 = sqrt(t1 + e);
    return (m & 1) ? -t2 : t2;
}

static void fill_m_table_(orc_cfg *c, int m) {
    if (c->ab[m]) return;
    const int ltop = c->lmax + 1;
    double *ab = (double *)calloc(2 * (size_t)(ltop - m + 2), sizeof(double));
    double *sd = (double *)malloc(sizeof(double));
    sd[0] = seed_amm_(m);
    if (ltop >= m + 1) ab[3] = sqrt(2.0 * m + 3);   /* a_{m+1} */
    for (long l = m + 2; l <= ltop; ++l) {
        double t1 = (double)((l + m) * (l - m));
        double t2 = (double)((l - 1 + m) * (l - 1 - m));
        ab[2 * (l - m) + 1] = sqrt(((2 * l + 1) * (2 * l - 1)) / t1);
        ab[2 * (l - m)] = -sqrt(((2 * l + 1) * t2) / ((2 * l - 3) * t1));
    }
    c->ab[m] = ab;
    c->seed[m] = sd;
}

/* ---- one column y_l(x) = Y_l^m(x), l=m..ltop, with an exponent carried in powers of two
 *      while the values are below the double range (extended-range start). ---- */
static void leg_column_pow_(const orc_cfg *c, int m, double x, double sint, int ltop, double *y, int sub);
static void leg_column_m_(const orc_cfg *c, int m, double x, double sint, int ltop, double *y) {
    leg_column_pow_(c, m, x, sint, ltop, y, 0);
}
/* Y_l^m(x) / sint^sub: the same recurrence from the seed a_mm sint^(m-sub) (sub <= m), finite
   at the poles (the vector transforms use sub = 1 for their functions Y/sin(theta)) */
static void leg_column_pow_(const orc_cfg *c, int m, double x, double sint, int ltop, double *y, int sub) {
    const double *ab = c->ab[m];
    /* seed a_mm * sint^(m-sub) as mantissa * 2^E by binary powering with renormalisation */
    int E = 0, es;
    double mant = frexp(c->seed[m][0], &es);
    E += es;
    {
        int n = m - sub;
        double b = frexp(sint, &es);
        int eb = es;
        while (n) {
            if (n & 1) {
                mant *= b;
                E += eb;
                mant = frexp(mant, &es);
                E += es;
            }
            n >>= 1;
            if (n) {
                b *= b;
                eb *= 2;
                b = frexp(b, &es);
                eb += es;
            }
        }
    }
    /* carry E as a multiple of 256 below zero */
    int Ec = 0;
    double y0 = mant;
    if (E < 0) {
        Ec = -(((-E) + 255) / 256) * 256;
        y0 = ldexp(mant, E - Ec);
    } else {
        y0 = ldexp(mant, E);
    }
    y[0] = (Ec == 0) ? y0 : ldexp(y0, Ec);
    if (ltop == m) return;
    double y1 = ab[3] * x * y0;
    y[1] = (Ec == 0) ? y1 : ldexp(y1, Ec);
    for (int l = m + 2; l <= ltop; ++l) {
        double a = ab[2 * (l - m) + 1], b = ab[2 * (l - m)];
        double y2 = a * x * y1 + b * y0;
        y0 = y1;
        y1 = y2;
        if (Ec < 0 && fabs(y1) > 0x1p300) {
            y0 = ldexp(y0, -256);
            y1 = ldexp(y1, -256);
            Ec += 256;
        }
        y[l - m] = (Ec == 0) ? y1 : ldexp(y1, Ec);
    }
}

void orc_legendre(const orc_cfg *c, int im, double x, double sint, int ltop, double *yl) {
    leg_column_m_(c, im * c->mres, x, sint, ltop, yl);
}

/* ---- regular grids (sht_init.c:580-597): Fejer's first rule without the poles (nodes
 *      (k+1/2) pi/n) and Clenshaw-Curtis with the poles (nodes k pi/(n-1)), weights from their
 *      closed forms (Waldvogel, BIT 46 (2006)):
 *        Fejer:  w_k = (2/n) [1 - 2 sum_{j=1}^{n/2} cos(2 j t_k) / (4 j^2 - 1)]
 *        CC:     w_k = (c_k/N) [1 - sum_{j=1}^{N/2} b_j cos(2 j t_k) / (4 j^2 - 1)],  N = n-1,
 *                c_0 = c_N = 1, c_k = 2 otherwise; b_j = 1 for j = N/2, 2 otherwise ---- */
static void regular_nodes_(double *x, double *st, double *w, int n, int poles) {
    const int N = poles ? n - 1 : n;
    for (int k = 0; k < n; ++k) {
        const double t = poles ? M_PI * k / N : M_PI * (k + 0.5) / n;
        double sum = 0.0;
        for (int j = 1; j <= N / 2; ++j) {
            const double b = poles ? ((2 * j == N) ? 1.0 : 2.0) : 2.0;
            sum += b * cos(2.0 * j * t) / (4.0 * j * j - 1.0);
        }
        const double ck = poles ? ((k == 0 || k == N) ? 1.0 : 2.0) : 2.0;
        w[k] = ck / N * (1.0 - sum);
        x[k] = cos(t);
        st[k] = sin(t);
    }
    if (poles) { st[0] = 0.0; st[n - 1] = 0.0; x[0] = 1.0; x[n - 1] = -1.0; }
    if (n & 1) { x[n / 2] = 0.0; st[n / 2] = 1.0; }
}

/* ---- create / destroy ---- */
orc_cfg *orc_create(int lmax, int mmax, int mres, int nlat, int nphi) {
    return orc_create_grid(lmax, mmax, mres, nlat, nphi, 0);
}

orc_cfg *orc_create_grid(int lmax, int mmax, int mres, int nlat, int nphi, int grid) {
    if (lmax < 1 || mres < 1 || mmax < 0 || mmax * mres > lmax) return NULL;
    if (nlat <= lmax || nphi <= 2 * mmax || grid < 0 || grid > 2 || nlat < 2) return NULL;
    orc_cfg *c = (orc_cfg *)calloc(1, sizeof(orc_cfg));
    c->lmax = lmax; c->mmax = mmax; c->mres = mres;
    c->nlat = nlat; c->nlat_2 = (nlat + 1) / 2; c->nphi = nphi;
    c->nlm = (long)(mmax + 1) * (lmax + 1) - (long)mres * mmax * (mmax + 1) / 2;
    c->ct = (double *)malloc(4 * sizeof(double) * nlat);
    c->st = c->ct + nlat;
    c->w = c->st + nlat;
    c->wp = c->w + nlat;
    if (grid == 0) gauss_nodes_(c->ct, c->st, c->w, nlat);
    else regular_nodes_(c->ct, c->st, c->w, nlat, grid == 2);
    /* weight of latitude pair k in the analysis sums (N + S) w: the middle row of an odd grid
       is its own mirror, so its weight is halved there (sht_init.c:598) */
    for (int k = 0; k < nlat; ++k) c->wp[k] = (2 * k + 1 == nlat) ? 0.5 * c->w[k] : c->w[k];
    c->mtab_n = mmax * mres + 2;
    c->seed = (double **)calloc(c->mtab_n, sizeof(double *));
    c->ab = (double **)calloc(c->mtab_n, sizeof(double *));
    for (int im = 0; im <= mmax; ++im) fill_m_table_(c, im * mres);
    if (lmax >= 1) fill_m_table_(c, 1);   /* m=1 functions give dY_l0/dtheta */
    c->tm = (int *)calloc(mmax + 1, sizeof(int));
    return c;
}

void orc_destroy(orc_cfg *c) {
    if (!c) return;
    for (int m = 0; m < c->mtab_n; ++m) { free(c->ab[m]); free(c->seed[m]); }
    free(c->ab); free(c->seed); free(c->ct); free(c->tm); free(c);
}

/* ---- polar optimization (sht_init.c:522-557): tm[im] = first latitude index it (from the
 *      pole) with max_l |Y_l^m(ct[it])| >= eps, searched from tm[im-1]; tm[0] = 0.  The
 *      transforms start order im at latitude (tm[im]/2)*2, the even start of SHTns's
 *      kernels without SIMD (kernel_SH_to_spat.gen.c:235-240, kernel_spat_to_SH.gen.c:198-201). */
void orc_set_polar_opt(orc_cfg *c, double eps) {
    for (int im = 0; im <= c->mmax; ++im) c->tm[im] = 0;
    if (!(eps > 0.0)) return;
    double *y = (double *)malloc(sizeof(double) * (c->lmax + 2));
    int it = 0;
    for (int im = 1; im <= c->mmax; ++im) {
        const int m = im * c->mres;
        while (it < c->nlat_2) {
            leg_column_m_(c, m, c->ct[it], c->st[it], c->lmax, y);
            double v = 0.0;
            for (int l = m; l <= c->lmax; ++l) if (fabs(y[l - m]) > v) v = fabs(y[l - m]);
            if (v >= eps) break;
            ++it;
        }
        c->tm[im] = (it / 2) * 2;
    }
    free(y);
}

void orc_polar_tm(const orc_cfg *c, int *tm) {
    for (int im = 0; im <= c->mmax; ++im) tm[im] = c->tm[im];
}

void orc_grid(const orc_cfg *c, double *ct, double *st, double *w) {
    memcpy(ct, c->ct, sizeof(double) * c->nlat);
    memcpy(st, c->st, sizeof(double) * c->nlat);
    memcpy(w, c->w, sizeof(double) * c->nlat);
}

/* ---- complex FFT: recursive mixed-radix decimation in time with a root table
 *      (any n; radices are the prime factors of n).  Unnormalized, sign=-1 forward. ---- */
static int small_factor_(int n) {
    if (n % 2 == 0) return 2;
    for (int f = 3; f * f <= n; f += 2) if (n % f == 0) return f;
    return n;
}

/* out[k] = sum_j in[j*istride] W_n^{jk}, with W_n^e = root[e*rstep] (root: N-th roots). */
static void fft_r_(const double *in, size_t istride, double *out, int n, const double *root, int N, int rstep) {
    if (n == 1) { out[0] = in[0]; out[1] = in[1]; return; }
    const int p = small_factor_(n), q = n / p;
    for (int r = 0; r < p; ++r)
        fft_r_(in + 2 * (size_t)r * istride, istride * p, out + 2 * (size_t)r * q, q, root, N, rstep * p);
    double tr[64], ti[64];
    double *br = tr, *bi = ti;
    if (p > 64) { br = (double *)malloc(sizeof(double) * p); bi = (double *)malloc(sizeof(double) * p); }
    for (int k = 0; k < q; ++k) {
        for (int r = 0; r < p; ++r) { br[r] = out[2 * ((size_t)r * q + k)]; bi[r] = out[2 * ((size_t)r * q + k) + 1]; }
        for (int s = 0; s < p; ++s) {
            const long kk = k + (long)q * s;
            double sr = 0, si = 0;
            for (int r = 0; r < p; ++r) {
                const long e = ((long)r * kk) % n;
                const double *w = root + 2 * ((e * rstep) % N);
                sr += br[r] * w[0] - bi[r] * w[1];
                si += br[r] * w[1] + bi[r] * w[0];
            }
            out[2 * kk] = sr;
            out[2 * kk + 1] = si;
        }
    }
    if (p > 64) { free(br); free(bi); }
}

void orc_fft(double *data, int n, int sign) {
    double *root = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    for (int k = 0; k < n; ++k) {
        const double ang = 2.0 * M_PI * (double)k / n;
        root[2 * k] = cos(ang);
        root[2 * k + 1] = sign * sin(ang);
    }
    double *out = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    fft_r_(data, 1, out, n, root, n, 1);
    memcpy(data, out, sizeof(double) * 2 * (size_t)n);
    free(out);
    free(root);
}

/* ---- Fourier <-> spatial along phi, one latitude at a time ---- */
/* F: [nlat][mmax+1] complex Fourier coefficients (bin im).  V: theta-contiguous. */
static void fourier_to_spat_(const orc_cfg *c, const double *F, double *V, int imlim) {
    const int nphi = c->nphi, nlat = c->nlat, mm1 = c->mmax + 1;
    #pragma omp parallel for schedule(static) num_threads(g_threads > 0 ? g_threads : omp_get_max_threads())
    for (int it = 0; it < nlat; ++it) {
        double *X = (double *)calloc(2 * (size_t)nphi, sizeof(double));
        const double *Fr = F + 2 * (size_t)it * mm1;
        X[0] = Fr[0];
        for (int im = 1; im <= imlim; ++im) {
            X[2 * im] = Fr[2 * im];
            X[2 * im + 1] = Fr[2 * im + 1];
            X[2 * (nphi - im)] = Fr[2 * im];
            X[2 * (nphi - im) + 1] = -Fr[2 * im + 1];
        }
        orc_fft(X, nphi, +1);
        for (int ip = 0; ip < nphi; ++ip) V[(size_t)ip * nlat + it] = X[2 * ip];
        free(X);
    }
}

static void spat_to_fourier_(const orc_cfg *c, const double *V, double *F, int imlim) {
    const int nphi = c->nphi, nlat = c->nlat, mm1 = c->mmax + 1;
    #pragma omp parallel for schedule(static) num_threads(g_threads > 0 ? g_threads : omp_get_max_threads())
    for (int it = 0; it < nlat; ++it) {
        double *X = (double *)calloc(2 * (size_t)nphi, sizeof(double));
        for (int ip = 0; ip < nphi; ++ip) X[2 * ip] = V[(size_t)ip * nlat + it];
        orc_fft(X, nphi, -1);
        double *Fr = F + 2 * (size_t)it * mm1;
        for (int im = 0; im <= c->mmax; ++im) {
            Fr[2 * im] = (im <= imlim) ? X[2 * im] : 0.0;
            Fr[2 * im + 1] = (im <= imlim) ? X[2 * im + 1] : 0.0;
        }
        free(X);
    }
}

static int imlim_(const orc_cfg *c, int ltr) {
    int imlim = ltr / c->mres;
    return imlim < c->mmax ? imlim : c->mmax;
}

/* ---- scalar synthesis ---- */
void orc_SH_to_spat(const orc_cfg *c, const double *Q, double *Vr, int ltr) {
    if (ltr < 0 || ltr > c->lmax) ltr = c->lmax;
    const int nlat = c->nlat, nl2 = c->nlat_2, mm1 = c->mmax + 1;
    const int imlim = imlim_(c, ltr);
    double *F = (double *)calloc(2 * (size_t)nlat * mm1, sizeof(double));
    #pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads > 0 ? g_threads : omp_get_max_threads())
    for (int im = 0; im <= imlim; ++im) {
        const int m = im * c->mres;
        double *y = (double *)malloc(sizeof(double) * (ltr - m + 2));
        const double *Ql = Q + 2 * lm_index(c, 0, im);
        for (int k = c->tm[im]; k < nl2; ++k) {
            leg_column_m_(c, m, c->ct[k], c->st[k], ltr, y);
            double er = 0, ei = 0, orr = 0, oi = 0;
            for (int l = m; l <= ltr; ++l) {
                double qr = Ql[2 * l], qi = (m == 0) ? 0.0 : Ql[2 * l + 1];
                if (((l - m) & 1) == 0) { er += y[l - m] * qr; ei += y[l - m] * qi; }
                else { orr += y[l - m] * qr; oi += y[l - m] * qi; }
            }
            double *Fn = F + 2 * ((size_t)k * mm1 + im);
            double *Fs = F + 2 * ((size_t)(nlat - 1 - k) * mm1 + im);
            Fn[0] = er + orr; Fn[1] = ei + oi;
            Fs[0] = er - orr; Fs[1] = ei - oi;
        }
        free(y);
    }
    fourier_to_spat_(c, F, Vr, imlim);
    free(F);
}

/* ---- scalar analysis ---- */
void orc_spat_to_SH(const orc_cfg *c, const double *Vr, double *Q, int ltr) {
    if (ltr < 0 || ltr > c->lmax) ltr = c->lmax;
    const int nlat = c->nlat, nl2 = c->nlat_2, mm1 = c->mmax + 1;
    const int imlim = imlim_(c, ltr);
    double *F = (double *)malloc(sizeof(double) * 2 * (size_t)nlat * mm1);
    spat_to_fourier_(c, Vr, F, imlim);
    memset(Q, 0, sizeof(double) * 2 * c->nlm);
    const double wfac = 2.0 * M_PI / c->nphi;
    #pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads > 0 ? g_threads : omp_get_max_threads())
    for (int im = 0; im <= imlim; ++im) {
        const int m = im * c->mres;
        double *y = (double *)malloc(sizeof(double) * (ltr - m + 2));
        double *acc = (double *)calloc(2 * (size_t)(ltr - m + 1), sizeof(double));
        for (int k = c->tm[im]; k < nl2; ++k) {
            leg_column_m_(c, m, c->ct[k], c->st[k], ltr, y);
            const double wk = c->wp[k] * wfac;
            const double *Fn = F + 2 * ((size_t)k * mm1 + im);
            const double *Fs = F + 2 * ((size_t)(nlat - 1 - k) * mm1 + im);
            double er = (Fn[0] + Fs[0]) * wk, ei = (Fn[1] + Fs[1]) * wk;
            double orr = (Fn[0] - Fs[0]) * wk, oi = (Fn[1] - Fs[1]) * wk;
            for (int l = m; l <= ltr; ++l) {
                if (((l - m) & 1) == 0) { acc[2 * (l - m)] += y[l - m] * er; acc[2 * (l - m) + 1] += y[l - m] * ei; }
                else { acc[2 * (l - m)] += y[l - m] * orr; acc[2 * (l - m) + 1] += y[l - m] * oi; }
            }
        }
        double *Ql = Q + 2 * lm_index(c, 0, im);
        for (int l = m; l <= ltr; ++l) {
            Ql[2 * l] = acc[2 * (l - m)];
            Ql[2 * l + 1] = (m == 0) ? 0.0 : acc[2 * (l - m) + 1];
        }
        free(acc);
        free(y);
    }
    free(F);
}

/* eps_l^m = sqrt((l-m)(l+m)/((2l+1)(2l-1))): cos(theta) Y_l = eps_{l+1} Y_{l+1} + eps_l Y_{l-1}
   (sht_func.c:400-438, orthonormal branch: 1/a_{l+1}^m). */
static double eps_lm_(int l, int m) {
    if (l <= m) return 0.0;
    return sqrt((double)((long)(l - m) * (l + m)) / (double)((long)(2 * l + 1) * (2 * l - 1)));
}

/* ---- vector synthesis: Vt = dS/dtheta + (1/sin) dT/dphi ; Vp = (1/sin) dS/dphi - dT/dtheta ---- */
void orc_SHsphtor_to_spat(const orc_cfg *c, const double *S, const double *T, double *Vt, double *Vp, int ltr) {
    if (ltr < 0 || ltr > c->lmax) ltr = c->lmax;
    const int nlat = c->nlat, nl2 = c->nlat_2, mm1 = c->mmax + 1;
    const int imlim = imlim_(c, ltr);
    double *Ft = (double *)calloc(2 * (size_t)nlat * mm1, sizeof(double));
    double *Fp = (double *)calloc(2 * (size_t)nlat * mm1, sizeof(double));
    #pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads > 0 ? g_threads : omp_get_max_threads())
    for (int im = 0; im <= imlim; ++im) {
        const int m = im * c->mres;
        const double *Sl = S + 2 * lm_index(c, 0, im);
        const double *Tl = T + 2 * lm_index(c, 0, im);
        const int n = ltr - m + 2;   /* degrees m..ltr+1 */
        double *V = (double *)calloc(2 * (size_t)n, sizeof(double));
        double *W = (double *)calloc(2 * (size_t)n, sizeof(double));
        double *y = (double *)malloc(sizeof(double) * (n + 1));
        int mcol = m;
        if (m == 0) {
            /* dY_l0/dtheta = sqrt(l(l+1)) Y_l^1 : synthesize with the m=1 functions, l=1..ltr */
            for (int l = 1; l <= ltr; ++l) {
                double f = sqrt((double)l * (l + 1));
                V[2 * (l - 1)] = f * Sl[2 * l];        /* Vt = sum S_l dY/dtheta */
                W[2 * (l - 1)] = -f * Tl[2 * l];       /* Vp = -sum T_l dY/dtheta */
            }
            mcol = 1;
        } else {
            /* V_l = i m T_l + (l-1) eps_l S_{l-1} - (l+2) eps_{l+1} S_{l+1}
               W_l = i m S_l - (l-1) eps_l T_{l-1} + (l+2) eps_{l+1} T_{l+1}   (l=m..ltr+1) */
            for (int l = m; l <= ltr + 1; ++l) {
                double vr = 0, vi = 0, wr = 0, wi = 0;
                if (l <= ltr) { vr += -m * Tl[2 * l + 1]; vi += m * Tl[2 * l]; wr += -m * Sl[2 * l + 1]; wi += m * Sl[2 * l]; }
                if (l - 1 >= m) {
                    double f = (l - 1) * eps_lm_(l, m);
                    vr += f * Sl[2 * (l - 1)]; vi += f * Sl[2 * (l - 1) + 1];
                    wr -= f * Tl[2 * (l - 1)]; wi -= f * Tl[2 * (l - 1) + 1];
                }
                if (l + 1 <= ltr) {
                    double f = -(l + 2) * eps_lm_(l + 1, m);
                    vr += f * Sl[2 * (l + 1)]; vi += f * Sl[2 * (l + 1) + 1];
                    wr -= f * Tl[2 * (l + 1)]; wi -= f * Tl[2 * (l + 1) + 1];
                }
                V[2 * (l - m)] = vr; V[2 * (l - m) + 1] = vi;
                W[2 * (l - m)] = wr; W[2 * (l - m) + 1] = wi;
            }
        }
        const int ltop = (m == 0) ? ltr : ltr + 1;
        for (int k = c->tm[im]; k < nl2; ++k) {
            /* Y/sin(theta) for m >= 1 straight from the recurrence (finite at the poles) */
            leg_column_pow_(c, mcol, c->ct[k], c->st[k], ltop, y, (m == 0) ? 0 : 1);
            const double s1 = 1.0;
            double ver = 0, vei = 0, vor = 0, voi = 0, wer = 0, wei = 0, wor = 0, woi = 0;
            for (int l = mcol; l <= ltop; ++l) {
                const int j = l - mcol;
                double yy = y[j];
                double a_r = V[2 * j], a_i = V[2 * j + 1], b_r = W[2 * j], b_i = W[2 * j + 1];
                if ((j & 1) == 0) { ver += yy * a_r; vei += yy * a_i; wer += yy * b_r; wei += yy * b_i; }
                else { vor += yy * a_r; voi += yy * a_i; wor += yy * b_r; woi += yy * b_i; }
            }
            double *Fn = Ft + 2 * ((size_t)k * mm1 + im), *Fs = Ft + 2 * ((size_t)(nlat - 1 - k) * mm1 + im);
            double *Gn = Fp + 2 * ((size_t)k * mm1 + im), *Gs = Fp + 2 * ((size_t)(nlat - 1 - k) * mm1 + im);
            Fn[0] = (ver + vor) * s1; Fn[1] = (vei + voi) * s1;
            Fs[0] = (ver - vor) * s1; Fs[1] = (vei - voi) * s1;
            Gn[0] = (wer + wor) * s1; Gn[1] = (wei + woi) * s1;
            Gs[0] = (wer - wor) * s1; Gs[1] = (wei - woi) * s1;
            if (m == 0) { Fn[1] = Fs[1] = Gn[1] = Gs[1] = 0.0; }
        }
        free(V); free(W); free(y);
    }
    fourier_to_spat_(c, Ft, Vt, imlim);
    fourier_to_spat_(c, Fp, Vp, imlim);
    free(Ft); free(Fp);
}

/* ---- vector analysis ---- */
void orc_spat_to_SHsphtor(const orc_cfg *c, const double *Vt, const double *Vp, double *S, double *T, int ltr) {
    if (ltr < 0 || ltr > c->lmax) ltr = c->lmax;
    const int nlat = c->nlat, nl2 = c->nlat_2, mm1 = c->mmax + 1;
    const int imlim = imlim_(c, ltr);
    double *Ft = (double *)malloc(sizeof(double) * 2 * (size_t)nlat * mm1);
    double *Fp = (double *)malloc(sizeof(double) * 2 * (size_t)nlat * mm1);
    spat_to_fourier_(c, Vt, Ft, imlim);
    spat_to_fourier_(c, Vp, Fp, imlim);
    memset(S, 0, sizeof(double) * 2 * c->nlm);
    memset(T, 0, sizeof(double) * 2 * c->nlm);
    const double wfac = 2.0 * M_PI / c->nphi;
    #pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads > 0 ? g_threads : omp_get_max_threads())
    for (int im = 0; im <= imlim; ++im) {
        const int m = im * c->mres;
        const int mcol = (m == 0) ? 1 : m;
        const int ltop = (m == 0) ? ltr : ltr + 1;
        const int n = ltop - mcol + 1;
        double *v = (double *)calloc(2 * (size_t)(n + 1), sizeof(double));
        double *w = (double *)calloc(2 * (size_t)(n + 1), sizeof(double));
        double *y = (double *)malloc(sizeof(double) * (n + 1));
        for (int k = c->tm[im]; k < nl2; ++k) {
            leg_column_pow_(c, mcol, c->ct[k], c->st[k], ltop, y, (m == 0) ? 0 : 1);
            const double wk = c->wp[k] * wfac;
            const double *Fn = Ft + 2 * ((size_t)k * mm1 + im), *Fs = Ft + 2 * ((size_t)(nlat - 1 - k) * mm1 + im);
            const double *Gn = Fp + 2 * ((size_t)k * mm1 + im), *Gs = Fp + 2 * ((size_t)(nlat - 1 - k) * mm1 + im);
            double ter = (Fn[0] + Fs[0]) * wk, tei = (Fn[1] + Fs[1]) * wk;
            double tor = (Fn[0] - Fs[0]) * wk, toi = (Fn[1] - Fs[1]) * wk;
            double per = (Gn[0] + Gs[0]) * wk, pei = (Gn[1] + Gs[1]) * wk;
            double por = (Gn[0] - Gs[0]) * wk, poi = (Gn[1] - Gs[1]) * wk;
            for (int j = 0; j < n; ++j) {
                double yy = y[j];
                if ((j & 1) == 0) { v[2 * j] += yy * ter; v[2 * j + 1] += yy * tei; w[2 * j] += yy * per; w[2 * j + 1] += yy * pei; }
                else { v[2 * j] += yy * tor; v[2 * j + 1] += yy * toi; w[2 * j] += yy * por; w[2 * j + 1] += yy * poi; }
            }
        }
        double *Sl = S + 2 * lm_index(c, 0, im);
        double *Tl = T + 2 * lm_index(c, 0, im);
        if (m == 0) {
            /* S_l = (1/(l(l+1))) int Vt dY_l0/dtheta ; T_l = -(1/(l(l+1))) int Vp dY_l0/dtheta */
            for (int l = 1; l <= ltr; ++l) {
                double f = sqrt((double)l * (l + 1)) / ((double)l * (l + 1));
                Sl[2 * l] = f * v[2 * (l - 1)];
                Tl[2 * l] = -f * w[2 * (l - 1)];
            }
        } else {
            /* S_l = [ l eps_{l+1} v_{l+1} - (l+1) eps_l v_{l-1} - i m w_l ] / (l(l+1))
               T_l = [ -i m v_l - l eps_{l+1} w_{l+1} + (l+1) eps_l w_{l-1} ] / (l(l+1)) */
            for (int l = m; l <= ltr; ++l) {
                const int j = l - m;
                double inv = 1.0 / ((double)l * (l + 1));
                double sr = m * w[2 * j + 1], si = -m * w[2 * j];      /* -i m w */
                double tr = m * v[2 * j + 1], ti = -m * v[2 * j];      /* -i m v */
                double fu = l * eps_lm_(l + 1, m);
                sr += fu * v[2 * (j + 1)]; si += fu * v[2 * (j + 1) + 1];
                tr -= fu * w[2 * (j + 1)]; ti -= fu * w[2 * (j + 1) + 1];
                if (j >= 1) {
                    double fl = (l + 1) * eps_lm_(l, m);
                    sr -= fl * v[2 * (j - 1)]; si -= fl * v[2 * (j - 1) + 1];
                    tr += fl * w[2 * (j - 1)]; ti += fl * w[2 * (j - 1) + 1];
                }
                Sl[2 * l] = sr * inv; Sl[2 * l + 1] = si * inv;
                Tl[2 * l] = tr * inv; Tl[2 * l + 1] = ti * inv;
            }
        }
        free(v); free(w); free(y);
    }
    free(Ft); free(Fp);
}
