// sht_host.cpp -- host side of shtns_amd: the SHTns C API (shtns.h / shtns_hip.h) on top of
// the gfx950 kernels in sht_kernels.hip.
//
// Host math is FP64 (long double where it buys accuracy for the tables):
//   Gauss nodes/weights          restates sht_legendre.c:763-811 and sht_init.c:560-649
//   a_mm seeds                   restates sht_legendre.c:584-613
//   Ishioka recurrence tables    same functions as sht_legendre.c:673-734, derived here from
//                                x P_l = e_{l+1} P_{l+1} + e_l P_{l-1} (see build_order_tables)
// Every transform runs on the GPU; without a usable GPU the transform entry points report
// an error (shtns_last_error) and leave their outputs untouched -- there is no CPU path.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <mutex>
#include <vector>

#include "sht_internal.h"

using namespace shtns_amd;

// ------------------------------------------------------------------------------------------
// errors and global state
// ------------------------------------------------------------------------------------------
static std::mutex g_mutex;
static shtns_info *g_cfg_list = nullptr;   // every live config (for shtns_reset)
static int g_verbose = 0;
static int g_gpu_request = 0;              // device chosen by shtns_use_gpu (-1: none)
static thread_local int g_last_err = 0;
static thread_local char g_last_msg[256] = "";

namespace shtns_amd {
void set_error(int code, const char *fmt, ...) {
    g_last_err = code;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_msg, sizeof(g_last_msg), fmt, ap);
    va_end(ap);
    if (code) fprintf(stderr, "*** [shtns_amd] %s\n", g_last_msg);
}
void clear_error() {
    g_last_err = 0;
    g_last_msg[0] = 0;
}
}  // namespace shtns_amd

// invalid sizes abort the program, as SHTns's shtns_runerr does (sht_init.c:68-72)
[[noreturn]] static void runerr(const char *msg) {
    printf("*** [shtns_amd] Run-time error : %s\n", msg);
    fflush(stdout);
    exit(1);
}

#define HIP_OK(call)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                          \
            set_error(2, "HIP error %s at %s:%d", hipGetErrorString(e_), __FILE__, __LINE__); \
            return;                                                                      \
        }                                                                                \
    } while (0)
#define HIP_OK_RET(call, ret)                                                            \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                          \
            set_error(2, "HIP error %s at %s:%d", hipGetErrorString(e_), __FILE__, __LINE__); \
            return ret;                                                                  \
        }                                                                                \
    } while (0)

static void prof_clear(shtns_cfg c);

extern "C" int shtns_last_error(void) { return g_last_err; }
extern "C" const char *shtns_last_error_string(void) { return g_last_msg; }

// ------------------------------------------------------------------------------------------
// sizes
// ------------------------------------------------------------------------------------------
extern "C" long nlm_calc(long lmax, long mmax, long mres) {
    if (mmax * mres > lmax) mmax = lmax / mres;
    return (mmax + 1) * (lmax + 1) - ((mmax * mres) * (mmax + 1)) / 2;
}

extern "C" long nlm_cplx_calc(long lmax, long mmax, long mres) {
    if (mmax * mres > lmax) mmax = lmax / mres;
    return (2 * mmax + 1) * (lmax + 1) - (mmax * mres) * (mmax + 1);
}

extern "C" void shtns_verbose(int v) { g_verbose = v; }

extern "C" const char *shtns_get_build_info(void) {
    return "[shtns_amd 0.1] SHTns 3.5.2 API, gfx950 HIP kernels (FP64 MFMA Legendre, LDS mixed-radix FFT), built " __DATE__;
}
extern "C" void shtns_print_version(void) { printf("%s\n", shtns_get_build_info()); }

extern "C" void shtns_print_cfg(shtns_cfg c) {
    printf("Lmax=%d, Mmax*Mres=%d, Mres=%d, Nlm=%d  [GPU %d, orthonormalized]\n", c->lmax, c->mmax * c->mres, c->mres,
           c->nlm, c->device);
    if (c->ct == nullptr) return;
    printf("Gauss grid : Nlat=%d, Nphi=%d  (%s)\n", c->nlat, c->nphi,
           (c->layout & SHT_PHI_CONTIGUOUS) ? "phi-contiguous" : "theta-contiguous");
}

// ------------------------------------------------------------------------------------------
// host math
// ------------------------------------------------------------------------------------------
// Regular grid without the poles (SHTns's sht_reg_fast / sht_reg_dct grid, sht_init.c:580-588):
// theta_k = (k + 1/2) pi / n, north first, with Fejer's first quadrature rule
//   w_k = (2/n) [1 - 2 sum_{j=1}^{n/2} cos(2 j theta_k) / (4 j^2 - 1)],
// exact for polynomials in cos(theta) of degree < n (Fejer 1933; Waldvogel, BIT 46 (2006)).
// The cosines cos(2 j theta_k) come from the Chebyshev recurrence in long double.
static void fejer1_grid(double *x, double *st, double *w, int n) {
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int k = 0; k < (n + 1) / 2; ++k) {
        const long double th = pi * (k + 0.5L) / n;
        const long double c = cosl(th), s = sinl(th);
        x[k] = (double)c;
        x[n - 1 - k] = (double)-c;
        st[k] = st[n - 1 - k] = (double)s;
        const long double c2 = cosl(2 * th);
        long double tm = 1.0L, t = c2, sum = 0.0L;   // cos(2 (j-1) theta), cos(2 j theta)
        for (int j = 1; j <= n / 2; ++j) {
            sum += t / (4.0L * j * j - 1.0L);
            const long double tn = 2 * c2 * t - tm;
            tm = t;
            t = tn;
        }
        w[k] = w[n - 1 - k] = (double)(2.0L / n * (1.0L - 2.0L * sum));
    }
}

// Regular grid including both poles: theta_k = k pi / N, k = 0..N (N = n - 1), with the
// Clenshaw-Curtis rule  w_k = (c_k / N) [1 - sum_{j=1}^{N/2} b_j cos(2 j theta_k) / (4 j^2 - 1)],
// c_0 = c_N = 1, c_k = 2 otherwise, b_j = 1 for j = N/2 and 2 otherwise (Waldvogel, BIT 46
// (2006)); exact for polynomials in cos(theta) of degree <= N.  Cosines from the Chebyshev
// recurrence in long double.
static void clenshaw_curtis_grid(double *x, double *st, double *w, int n) {
    const long double pi = 3.141592653589793238462643383279502884L;
    const int N = n - 1;
    for (int k = 0; k < (n + 1) / 2; ++k) {
        const long double th = pi * k / N;
        const long double c2 = cosl(2 * th);
        long double tm = 1.0L, t = c2, sum = 0.0L;
        for (int j = 1; j <= N / 2; ++j) {
            const long double b = (2 * j == N) ? 1.0L : 2.0L;
            sum += b * t / (4.0L * j * j - 1.0L);
            const long double tn = 2 * c2 * t - tm;
            tm = t;
            t = tn;
        }
        const long double ck = (k == 0) ? 1.0L : 2.0L;
        w[k] = w[n - 1 - k] = (double)(ck / N * (1.0L - sum));
        x[k] = (k == 0) ? 1.0 : (double)cosl(th);
        x[n - 1 - k] = -x[k];
        st[k] = st[n - 1 - k] = (k == 0) ? 0.0 : (double)sinl(th);
    }
    if (n & 1) { x[n / 2] = 0.0; st[n / 2] = 1.0; }
}

// Gauss-Legendre nodes x_k = cos(theta_k) (descending, north first) and weights (sum 2).
// Each root of P_n is found by Newton's method on the colatitude, in long double:
//   f(theta) = P_n(cos theta),  f'(theta) = n (cos(theta) P_n - P_{n-1}) / sin(theta),
// from Tricomi's asymptotic first guess theta = pi (4k+3)/(4n+2) (k-th root from the north
// pole), with P_n and P_{n-1} from the three-term recurrence; it stops when the step falls
// below 1e-18 rad.  The weights are w_k = 2 sin^2(theta_k) / (n P_{n-1}(x_k))^2 and sin(theta)
// comes from theta itself, so the polar values keep full relative precision.
static void gauss_nodes(double *x, double *st, double *w, int n) {
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int k = 0; k < n / 2; ++k) {
        long double th = pi * (4 * k + 3) / (4.0L * n + 2), pn = 0, pn1 = 0;
        for (int it = 0; it < 30; ++it) {
            const long double c = cosl(th), s = sinl(th);
            long double a = 1.0L, b = c;            // P_{l-1}, P_l
            for (int l = 1; l < n; ++l) {
                const long double nx = ((2 * l + 1) * c * b - l * a) / (l + 1);
                a = b;
                b = nx;
            }
            pn = b;
            pn1 = a;
            const long double step = pn * s / (n * (c * pn - pn1));
            th -= step;
            if (fabsl(step) < 1e-18L) break;
        }
        const long double c = cosl(th), s = sinl(th);
        long double a = 1.0L, b = c;
        for (int l = 1; l < n; ++l) {
            const long double nx = ((2 * l + 1) * c * b - l * a) / (l + 1);
            a = b;
            b = nx;
        }
        pn1 = a;
        const long double wk = 2.0L * s * s / ((n * pn1) * (n * pn1));
        x[k] = (double)c;
        x[n - 1 - k] = (double)-c;
        st[k] = st[n - 1 - k] = (double)s;
        w[k] = w[n - 1 - k] = (double)wk;
    }
    if (n & 1) {   // odd n: the root x = 0, weight 2 / (n P_{n-1}(0))^2
        long double a = 1.0L, b = 0.0L;   // P_0(0), P_1(0)
        for (int l = 1; l < n - 1; ++l) {
            const long double nx = -(long double)l * a / (l + 1);   // P_{l+1}(0) = -l P_{l-1}(0) / (l+1)
            a = b;
            b = nx;
        }
        x[n / 2] = 0.0;
        st[n / 2] = 1.0;
        w[n / 2] = (double)(2.0L / ((n * b) * (n * b)));
    }
}

// Y_m^m prefactor sqrt(prod_{k<=m} (2k+1)/(2k) / (4 pi)) with Condon-Shortley sign (-1)^m.
static long double seed_amm(int m) {
    long double t = 0.25L / 3.14159265358979323846264338327950288L;
    for (int k = 1; k <= m; ++k) t *= ((long double)k + 0.5L) / k;
    long double s = sqrtl(t);
    return (m & 1) ? -s : s;
}

static long double eps_l(int l, int m) {
    if (l <= m) return 0.0L;
    return sqrtl((long double)((long)(l - m) * (l + m)) / (long double)((long)(2 * l + 1) * (2 * l - 1)));
}

// Recurrence tables of one order m, steps i = 0..nst-1 (degrees m .. m+2*nst-1).
// With e_l as above, u_i = P_{m+2i}, v_i = P_{m+2i+1}/x:
//     u_i = e_{m+2i+1} v_i + e_{m+2i} v_{i-1},      x^2 v_i = e_{m+2i+2} u_{i+1} + e_{m+2i+1} u_i.
// Writing v_i = d_i z_i with d_i d_{i+1} = (-1)^i / (e_{m+2i+2} e_{m+2i+3}) turns this into
//     z_{i+1} = (alpha_i x^2 + beta_i) z_i + z_{i-1},
//     alpha_i = (-1)^i d_i^2,   beta_i = -alpha_i (e_{m+2i+1}^2 + e_{m+2i+2}^2),
// with z_0 = sin^m(theta), z_{-1} = 0, d_0 = a_mm / e_{m+1}; then
//     P_{m+2i} = ue_i z_i + ue1_{i-1} z_{i-1},  ue_i = e_{m+2i+1} d_i,  ue1_i = e_{m+2i+2} d_i,
//     P_{m+2i+1} = x uo_i z_i,                  uo_i = d_i.
static void build_order_tables(int m, int nst, double2 *ab, double4 *pm) {
    long double d = seed_amm(m) / eps_l(m + 1, m);
    for (int i = 0; i < nst; ++i) {
        const long double e1 = eps_l(m + 2 * i + 1, m), e2 = eps_l(m + 2 * i + 2, m), e3 = eps_l(m + 2 * i + 3, m);
        const long double sgn = (i & 1) ? -1.0L : 1.0L;
        const long double alpha = sgn * d * d;
        const long double beta = -alpha * (e1 * e1 + e2 * e2);
        ab[i] = make_double2((double)alpha, (double)beta);
        pm[i] = make_double4((double)(e1 * d), (double)(e2 * d), (double)d, 0.0);
        d = sgn / (e2 * e3 * d);
    }
}

// ------------------------------------------------------------------------------------------
// config lifetime
// ------------------------------------------------------------------------------------------
extern "C" shtns_cfg shtns_create(int lmax, int mmax, int mres, enum shtns_norm norm) {
    clear_error();
    if (lmax < 2) runerr("lmax must be at least 2");
    if (lmax > 65534) runerr("lmax too large");
    if (mres <= 0) runerr("MRES must be > 0");
    if (mmax < 0 || mmax * mres > lmax) runerr("MMAX*MRES should not exceed LMAX");
    const int nbase = (int)norm & 255;   // sht_orthonormal, sht_fourpi, sht_schmidt, sht_for_rotations
    if (nbase > (int)sht_for_rotations || ((int)norm & ~(255 | SHT_NO_CS_PHASE | SHT_REAL_NORM))) {
        set_error(1, "unknown normalization (norm=%d)", (int)norm);
        return nullptr;
    }
    shtns_info *c = new shtns_info();
    c->lmax = (unsigned short)lmax;
    c->mmax = (unsigned short)mmax;
    c->mres = (unsigned short)mres;
    c->nlm = (unsigned int)nlm_calc(lmax, mmax, mres);
    c->nlm_cplx = (unsigned int)nlm_cplx_calc(lmax, mmax, mres);
    c->norm = (int)norm;
    c->li = new unsigned short[2 * (size_t)c->nlm];
    c->mi = c->li + c->nlm;
    long lm = 0;
    for (int im = 0; im <= mmax; ++im)
        for (int l = im * mres; l <= lmax; ++l) {
            c->li[lm] = (unsigned short)l;
            c->mi[lm] = (unsigned short)(im * mres);
            ++lm;
        }
    // normalization factors relative to orthonormal harmonics with the Condon-Shortley phase
    // (legendre_precomp, sht_legendre.c:586-605, 635-660: Y_00 = 1 for the 4 pi and Schmidt
    // norms, Schmidt functions divided by sqrt(2l+1), (-1)^m optional, m > 0 scaled by
    // sqrt(mpos_renorm) = sqrt(1/2) with SHT_REAL_NORM)
    if ((int)norm != (int)sht_orthonormal) {
        c->nrm.resize(c->nlm);
        for (unsigned i = 0; i < c->nlm; ++i) {
            const int l = c->li[i], m = c->mi[i];
            long double f = 1.0L;
            if (nbase == sht_fourpi) f = sqrtl(4.0L * M_PIl);
            else if (nbase == sht_schmidt || nbase == sht_for_rotations) f = sqrtl(4.0L * M_PIl / (2 * l + 1));
            if (((int)norm & SHT_NO_CS_PHASE) && (m & 1)) f = -f;
            if (((int)norm & SHT_REAL_NORM) && m > 0) f *= sqrtl(0.5L);
            c->nrm[i] = (double)f;
        }
    }
    std::lock_guard<std::mutex> g(g_mutex);
    c->next = g_cfg_list;
    g_cfg_list = c;
    return c;
}

static void free_device(shtns_info *c) {
    if (!c->gpu_ready) return;
    (void)hipSetDevice(c->device);
    if (c->clone_of == nullptr) {
        DeviceTables &d = c->d;
        void *ptrs[] = {d.x, d.x2, d.st, d.wq, d.step_ofs, d.ab, d.pm, d.tm, d.i0, d.z0, d.i0v, d.z0v, d.roots, d.fft_pos,
                        d.sb_ofs, d.eA, d.abk, d.dblk, d.delta, d.nrm};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
    }
    if (c->scratch) (void)hipFree(c->scratch);
    if (c->xscratch) (void)hipFree(c->xscratch);
    if (c->nscratch) (void)hipFree(c->nscratch);
    c->nscratch = nullptr;
    c->nscratch_bytes = 0;
    for (auto &sl : c->hstage)
        if (sl.first) (void)hipFree(sl.first);
    c->hstage.clear();
    c->xscratch = nullptr;
    c->xscratch_bytes = 0;
    if (c->own_streams) {
        if (c->comp_stream) (void)hipStreamDestroy(c->comp_stream);
    }
    if (c->aux_stream) (void)hipStreamDestroy(c->aux_stream);
    c->aux_stream = nullptr;
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    c->ev_pool.clear();
    c->d = DeviceTables();
    c->scratch = nullptr;
    c->scratch_bytes = 0;
    c->gpu_ready = false;
}

// A clone (cushtns_clone, and every configuration of an uneven batch) borrows its source's
// device tables, and batches capture them in HIP graphs.  Freeing or rebuilding the source's
// grid while a clone lives would leave those pointers dangling, so it is refused.
static bool has_live_clones(shtns_cfg c) {
    std::lock_guard<std::mutex> g(g_mutex);
    for (shtns_info *p = g_cfg_list; p; p = p->next)
        if (p->clone_of == c) return true;
    return false;
}

extern "C" void cushtns_release_gpu(shtns_cfg c) {
    clear_error();
    if (has_live_clones(c)) {
        set_error(1, "cushtns_release_gpu: configuration still has live clones (or batches); destroy them first");
        return;
    }
    free_device(c);
}

extern "C" void shtns_unset_grid(shtns_cfg c) {
    if (has_live_clones(c)) {
        set_error(1, "shtns_unset_grid: configuration still has live clones (or batches); destroy them first");
        return;
    }
    prof_clear(c);
    free_device(c);
    delete[] c->ct;
    c->ct = nullptr;
    c->st = nullptr;
    c->wg.clear();
    c->nlat = 0;
    c->nlat_2 = 0;
    c->nphi = 0;
    c->nspat = 0;
    c->nlat_padded = 0;
}

extern "C" void shtns_destroy(shtns_cfg c) {
    if (!c) return;
    {
        std::lock_guard<std::mutex> g(g_mutex);
        for (shtns_info *p = g_cfg_list; p; p = p->next)
            if (p->clone_of == c) runerr("shtns_destroy: config still has live clones");
        shtns_info **pp = &g_cfg_list;
        while (*pp && *pp != c) pp = &(*pp)->next;
        if (*pp) *pp = c->next;
    }
    shtns_unset_grid(c);
    if (c->clone_of == nullptr) delete[] c->li;
    delete c;
}

extern "C" void shtns_reset(void) {
    while (true) {
        shtns_info *c;
        {
            std::lock_guard<std::mutex> g(g_mutex);
            c = g_cfg_list;
            // destroy clones first
            for (shtns_info *p = g_cfg_list; p; p = p->next)
                if (p->clone_of) { c = p; break; }
        }
        if (!c) break;
        shtns_destroy(c);
    }
}

extern "C" int shtns_use_threads(int num_threads) {
    (void)num_threads;   // transforms run on the GPU; host work is negligible
    return 1;
}

extern "C" int shtns_use_gpu(int device_id) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0 || device_id < 0) {
        g_gpu_request = -1;
        return -1;
    }
    g_gpu_request = device_id % n;
    if (hipSetDevice(g_gpu_request) != hipSuccess) {
        g_gpu_request = -1;
        return -1;
    }
    return g_gpu_request;
}

extern "C" void *shtns_malloc(size_t bytes) {
    // pinned host memory when a GPU is present (as SHTns's CUDA build, sht_gpu.cu:67-81)
    void *p = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) == hipSuccess && n > 0) {
        if (hipHostMalloc(&p, bytes, hipHostMallocDefault) == hipSuccess) return p;
        p = nullptr;
    }
    if (posix_memalign(&p, 64, bytes) != 0) return nullptr;
    return p;
}

extern "C" void shtns_free(void *p) {
    if (!p) return;
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, p) == hipSuccess && attr.type == hipMemoryTypeHost) {
        (void)hipHostFree(p);
        return;
    }
    (void)hipGetLastError();
    free(p);
}

// ------------------------------------------------------------------------------------------
// grid
// ------------------------------------------------------------------------------------------
static int factor_fft(int M, FftPlan &plan) {
    plan.M = M;
    plan.nstages = 0;
    int n = M;
    while (n % 4 == 0 && plan.nstages < 24) { plan.radix[plan.nstages++] = 4; n /= 4; }
    const int small[] = {2, 3, 5, 7};
    for (int p : small)
        while (n % p == 0 && plan.nstages < 24) { plan.radix[plan.nstages++] = p; n /= p; }
    return (n == 1) ? 0 : -1;
}

// digit-reversal: position of bin k after the in-place DIF stages of `plan`
static void fft_positions(const FftPlan &plan, std::vector<int> &pos) {
    const int M = plan.M;
    pos.assign(M, 0);
    for (int k = 0; k < M; ++k) {
        int Ns = M, kk = k, p0 = 0;
        for (int s = 0; s < plan.nstages; ++s) {
            const int p = plan.radix[s];
            const int q = Ns / p;
            p0 += (kk % p) * q;
            kk /= p;
            Ns = q;
        }
        pos[k] = p0;
    }
}

static int device_init(shtns_info *c);

// SHTns's polar optimization (sht_init.c:522-557): tm[im] is the first latitude (from the
// pole) where max_l |Y_l^m(cos theta)| >= eps, searched upward from tm[im-1] (tm[0] = 0).
// The transforms skip latitude pairs below it, from an even index as SHTns's kernels do
// without SIMD (kernel_SH_to_spat.gen.c:235-240, kernel_spat_to_SH.gen.c:198-201): rows of
// order im there are zero in synthesis and left out of the analysis.  Slot mmax+1 (the m=1
// functions used by vector transforms at m=0) is never skipped, as SHTns's m=0.
static void polar_optimize(shtns_cfg c, double eps) {
    c->tm.assign(c->mmax + 2, 0);
    if (!(eps > 0.0)) return;
    std::vector<double> y(c->lmax + 1);
    int it = 0;
    for (int im = 1; im <= c->mmax; ++im) {
        const int m = im * c->mres;
        while (it < c->nlat_2) {
            legendre_sphPlm_array(c, c->lmax, im, c->ct[it], y.data());
            double v = 0.0;
            for (int l = m; l <= c->lmax; ++l) v = std::max(v, fabs(y[l - m]));
            if (v >= eps) break;
            ++it;
        }
        c->tm[im] = (it >> 1) * 2;
    }
}

extern "C" int shtns_set_grid_auto(shtns_cfg c, enum shtns_type flags, double eps, int nl_order, int *nlat, int *nphi) {
    clear_error();
    const int grid = (int)flags & 255;
    const int layout = (int)flags & ~255;
    const bool regular = (grid == sht_reg_fast || grid == sht_reg_dct);   // Fejer grid, no poles
    const bool poles = (grid == sht_reg_poles);                            // Clenshaw-Curtis, both poles
    if (!regular && !poles && grid != sht_gauss && grid != sht_auto && grid != sht_quick_init &&
        grid != sht_gauss_fly) {
        set_error(1, "grid type %d is not implemented", grid);
        return 0;
    }
    if (c->clone_of != nullptr) {
        set_error(1, "shtns_set_grid: a clone shares the grid of its source configuration");
        return 0;
    }
    if (has_live_clones(c)) {
        set_error(1, "shtns_set_grid: configuration still has live clones (or batches); destroy them first");
        return 0;
    }
    if (nl_order <= 0) nl_order = 1;
    if (*nlat <= 0) {   // choose sizes like SHTns: Gauss nlat > (nl_order+1)*lmax/2, regular
                        // nlat >= (nl_order+1)*(lmax+1); nphi > (nl_order+1)*mmax
        int n = (regular || poles) ? (nl_order + 1) * (c->lmax + 1) : ((nl_order + 1) * c->lmax) / 2 + 2;
        n += (n & 1);
        if (poles) n += 1;   // 2^k + 1 style: an odd count puts a row on the equator
        *nlat = n;
    }
    if (*nphi <= 0) {
        int n = (nl_order + 1) * c->mmax * c->mres + 2;
        while (true) {   // smallest 2^a 3^b 5^c >= n, even
            int t = n;
            for (int p : {2, 3, 5})
                while (t % p == 0) t /= p;
            if (t == 1 && (n & 1) == 0) break;
            ++n;
        }
        *nphi = n;
    }
    if (*nlat <= c->lmax) runerr("Nlat must be larger than Lmax");
    if (*nphi <= 2 * c->mmax) runerr("Nphi must be larger than 2*Mmax");
    if (*nphi & 1) {
        set_error(1, "odd nphi is not implemented (nphi=%d)", *nphi);
        return 0;
    }
    FftPlan plan{};
    if (factor_fft(*nphi / 2, plan) != 0) {
        set_error(1, "nphi/2 must factor into 2,3,5,7 (nphi=%d)", *nphi);
        return 0;
    }
    shtns_unset_grid(c);
    c->nlat = *nlat;
    c->nlat_2 = (unsigned short)((*nlat + 1) / 2);   // an odd grid's equator row is its own mirror
    c->nphi = *nphi;
    c->nspat = c->nlat * c->nphi;
    c->nlat_padded = c->nlat;
    c->layout = (layout & SHT_PHI_CONTIGUOUS) ? SHT_PHI_CONTIGUOUS : SHT_THETA_CONTIGUOUS;
    if (layout & SHT_SOUTH_POLE_FIRST) c->layout |= SHT_SOUTH_POLE_FIRST;
    c->eps = eps;
    c->load_save = (layout & SHT_LOAD_SAVE_CFG) != 0;
    c->grid_flags = (int)flags & ~SHT_LOAD_SAVE_CFG;
    c->ct = new double[2 * (size_t)c->nlat];
    c->st = c->ct + c->nlat;
    std::vector<double> w(c->nlat);
    if (regular) fejer1_grid(c->ct, c->st, w.data(), c->nlat);
    else if (poles) clenshaw_curtis_grid(c->ct, c->st, w.data(), c->nlat);
    else gauss_nodes(c->ct, c->st, w.data(), c->nlat);
    {   // quadrature sanity sums (sht_init.c:607-624)
        double s = 0, x2 = 0, st2 = 0;
        for (unsigned i = 0; i < c->nlat; ++i) {
            s += w[i];
            x2 += w[i] * c->ct[i] * c->ct[i];
            st2 += w[i] * c->st[i] * c->st[i];
        }
        if (fabs(s - 2.0) + fabs(x2 * 1.5 - 1.) + fabs(st2 * 0.75 - 1.) > 1e-14) runerr("Bad quadrature accuracy.");
    }
    // stored per latitude pair; the equator row of an odd grid is both rows of its pair, so
    // its weight is halved (sht_init.c:598)
    if (c->nlat & 1) w[c->nlat / 2] *= 0.5;
    if (layout & SHT_SOUTH_POLE_FIRST)
        for (unsigned i = 0; i < c->nlat; ++i) c->ct[i] = -c->ct[i];
    c->wg.assign(w.begin(), w.begin() + c->nlat_2);
    c->d.plan = plan;
    polar_optimize(c, eps);

    // GPU tables (a config is usable for queries without a GPU; transforms then fail)
    if (g_gpu_request >= 0) {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0) {
            if (device_init(c) < 0) fprintf(stderr, "*** [shtns_amd] GPU initialization failed: %s\n", g_last_msg);
            else if (c->load_save) (void)tune_cache_apply(c, 0);   // shapes saved by an earlier run
        } else {
            (void)hipGetLastError();
        }
    }
    return (int)c->nspat;
}

extern "C" int shtns_set_grid(shtns_cfg c, enum shtns_type flags, double eps, int nlat, int nphi) {
    if (nlat == 0 || nphi == 0) runerr("nlat or nphi is zero !");
    return shtns_set_grid_auto(c, flags, eps, 0, &nlat, &nphi);
}

extern "C" shtns_cfg shtns_init(enum shtns_type flags, int lmax, int mmax, int mres, int nlat, int nphi) {
    shtns_cfg c = shtns_create(lmax, mmax, mres, sht_orthonormal);
    if (c) shtns_set_grid(c, flags, 1.0e-10, nlat, nphi);
    return c;
}

extern "C" int shtns_gauss_wts(shtns_cfg c, double *wts) {
    if (c->wg.empty()) return 0;
    for (int i = 0; i < c->nlat_2; ++i) wts[i] = c->wg[i];
    return c->nlat_2;
}

// normalization factor of (l, m) (1 for orthonormal harmonics with the Condon-Shortley phase)
static double nrm_of(shtns_cfg c, int l, int m) {
    if (c->nrm.empty() || m > c->mmax * c->mres || m % c->mres || l < m || l > c->lmax) return 1.0;
    return c->nrm[LiM(c, l, m / c->mres)];
}
// coefficients of 1, cos(theta), sin(theta) cos(phi) and of unit energy (sht_init.c:76-94,
// 1180-1195): the orthonormal values divided by the normalization factor
extern "C" double sh00_1(shtns_cfg c) { return sqrt(4. * M_PI) / nrm_of(c, 0, 0); }
extern "C" double sh10_ct(shtns_cfg c) { return sqrt(4. * M_PI / 3.) / nrm_of(c, 1, 0); }
extern "C" double sh11_st(shtns_cfg c) { return -sqrt(2. * M_PI / 3.) / nrm_of(c, 1, c->mres == 1 ? 1 : 0); }
extern "C" double shlm_e1(shtns_cfg c, int l, int m) {
    double x = 1.0;
    if (m != 0) x *= sqrt(0.5);
    return x / fabs(nrm_of(c, l, m));
}

// Normalized associated Legendre functions Y_l^m(x), l = m..lmax (GSL-compatible output,
// shtns.h:219-227): classic three-term recurrence on the host.
extern "C" int legendre_sphPlm_array(shtns_cfg c, const int lmax, const int im, const double x, double *yl) {
    const int m = im * c->mres;
    if (lmax < m || lmax > c->lmax) return -1;
    const double s2 = (1. - x) * (1. + x);
    long double y0 = seed_amm(m) * powl(sqrtl((long double)s2), m);
    yl[0] = (double)y0;
    if (lmax == m) {
        yl[0] *= nrm_of(c, m, m);
        return 1;
    }
    long double y1 = sqrtl(2.0L * m + 3) * x * y0;
    yl[1] = (double)y1;
    for (int l = m + 2; l <= lmax; ++l) {
        const long double t1 = (long double)((long)(l + m) * (l - m));
        const long double t2 = (long double)((long)(l - 1 + m) * (l - 1 - m));
        const long double a = sqrtl(((long double)(2 * l + 1) * (2 * l - 1)) / t1);
        const long double b = -sqrtl(((long double)(2 * l + 1) * t2) / ((long double)(2 * l - 3) * t1));
        const long double y2 = a * x * y1 + b * y0;
        y0 = y1;
        y1 = y2;
        yl[l - m] = (double)y1;
    }
    if (!c->nrm.empty())
        for (int l = m; l <= lmax; ++l) yl[l - m] *= nrm_of(c, l, m);
    return lmax - m + 1;
}

// f(theta,phi) = sum_l Q_l0 Y_l0 + 2 Re sum_{m>0} Q_lm Y_lm e^{i m phi} at one point (sht_func.c:585-622).
extern "C" double SH_to_point(shtns_cfg c, cplx *Qlm, double cost, double phi) {
    std::vector<double> y(c->lmax + 1);
    const double *q = reinterpret_cast<const double *>(Qlm);
    double v = 0.0;
    for (int im = 0; im <= c->mmax; ++im) {
        const int m = im * c->mres;
        legendre_sphPlm_array(c, c->lmax, im, cost, y.data());
        const long base = LiM(c, 0, im);
        double re = 0, ii = 0;
        for (int l = m; l <= c->lmax; ++l) {
            re += y[l - m] * q[2 * (base + l)];
            ii += y[l - m] * q[2 * (base + l) + 1];
        }
        if (m == 0) v += re;
        else v += 2.0 * (re * cos(m * phi) - ii * sin(m * phi));
    }
    return v;
}

// Point and latitude evaluation (sht_func.c:516-885), on the host.  For order m >= 1 the
// recurrence runs on P_lm / sin(theta) (seed sin^(m-1) theta), so the 1/sin(theta) terms of
// the vector fields stay finite at the poles, and
//   dP_lm/dtheta = l cos(theta) (P_l/sin) - sqrt((2l+1)/(2l-1) (l^2 - m^2)) (P_{l-1}/sin);
// for m = 0, dP_l0/dtheta = sqrt(l (l+1)) P_l1 (orthonormal functions, Condon-Shortley phase).
static void plm_over_sin(int m, int ltop, long double x, long double s, long double *y) {
    // y[l - m] = P_lm(x) / sin(theta) for m >= 1, P_l0(x) for m = 0
    long double y0 = (long double)seed_amm(m) * (m >= 1 ? powl(s, m - 1) : 1.0L);
    y[0] = y0;
    if (ltop == m) return;
    long double y1 = sqrtl(2.0L * m + 3) * x * y0;
    y[1] = y1;
    for (int l = m + 2; l <= ltop; ++l) {
        const long double t1 = (long double)((long)(l + m) * (l - m));
        const long double t2 = (long double)((long)(l - 1 + m) * (l - 1 - m));
        const long double a = sqrtl(((long double)(2 * l + 1) * (2 * l - 1)) / t1);
        const long double b = -sqrtl(((long double)(2 * l + 1) * t2) / ((long double)(2 * l - 3) * t1));
        const long double y2 = a * x * y1 + b * y0;
        y0 = y1;
        y1 = y2;
        y[l - m] = y1;
    }
}
// ys[l-m] = P_lm / sin (m >= 1) or P_l0 (m = 0); dy[l-m] = dP_lm/dtheta; l = m..ltop
static void plm_deriv(int m, int ltop, double cost, std::vector<long double> &ys, std::vector<long double> &dy) {
    const long double x = cost, s = sqrtl((1.0L - x) * (1.0L + x));
    ys.assign(ltop - m + 1, 0.0L);
    dy.assign(ltop - m + 1, 0.0L);
    if (ltop < m) return;
    plm_over_sin(m, ltop, x, s, ys.data());
    if (m >= 1) {
        for (int l = m; l <= ltop; ++l) {
            const long double c = sqrtl((long double)(2 * l + 1) / (2 * l - 1) * ((long double)l * l - (long double)m * m));
            dy[l - m] = l * x * ys[l - m] - (l > m ? c * ys[l - m - 1] : 0.0L);
        }
    } else if (ltop >= 1) {
        std::vector<long double> y1(ltop);
        plm_over_sin(1, ltop, x, s, y1.data());
        for (int l = 1; l <= ltop; ++l) dy[l] = sqrtl((long double)l * (l + 1)) * s * y1[l - 1];
    }
}

// Legendre functions and their theta derivatives, l = m..lmax (shtns.h:233,
// sht_legendre.c:356-366): m = 0: yl = P_l0, dyl = dP_l0/dtheta; m > 0: yl = P_lm / sin(theta),
// dyl = dP_lm/dtheta (finite at the poles).  Returns the first degree with a non-zero value.
extern "C" int legendre_sphPlm_deriv_array(shtns_cfg c, const int lmax, const int im, const double x,
                                           const double sint, double *yl, double *dyl) {
    (void)sint;   // sin(theta) is recomputed in long double from x
    const int m = im * c->mres;
    if (im < 0 || im > c->mmax || lmax < m || lmax > c->lmax + 1) return -1;
    std::vector<long double> ys, dy;
    plm_deriv(m, lmax, x, ys, dy);
    int lnz = lmax + 1;
    for (int l = m; l <= lmax; ++l) {
        yl[l - m] = (double)ys[l - m] * nrm_of(c, l, m);
        dyl[l - m] = (double)dy[l - m] * nrm_of(c, l, m);
        if (lnz > lmax && (yl[l - m] != 0.0 || dyl[l - m] != 0.0)) lnz = l;
    }
    return lnz;
}

// Operators coupling degrees l-1 and l+1 of one order (sht_func.c:450-500), orthonormal
// functions: cos(theta) Y_l = a_l Y_{l+1} + a_{l-1} Y_{l-1}, a_l = sqrt(((l+1)^2 - m^2) / ((2l+1)(2l+3))),
// and sin(theta) d/dtheta Y_l = l a_l Y_{l+1} - (l+1) a_{l-1} Y_{l-1}.  mx[2 lm] multiplies
// Q_{l-1}, mx[2 lm + 1] multiplies Q_{l+1}; the coupling to degree lmax+1 is dropped.
extern "C" void mul_ct_matrix(shtns_cfg c, double *mx) {
    for (int im = 0; im <= c->mmax; ++im) {
        const long m = (long)im * c->mres;
        for (int l = (int)m; l <= c->lmax; ++l) {
            const long lm = LiM(c, l, im);
            const double lo = (l > m) ? sqrt((double)((long)l * l - m * m) / ((2.0 * l - 1) * (2.0 * l + 1))) : 0.0;
            const double hi = (l < c->lmax)
                                  ? sqrt((double)((long)(l + 1) * (l + 1) - m * m) / ((2.0 * l + 1) * (2.0 * l + 3)))
                                  : 0.0;
            // in the configuration's normalization: nrm_{l-1} / nrm_l and nrm_{l+1} / nrm_l
            mx[2 * lm] = (l > m) ? lo * nrm_of(c, l - 1, (int)m) / nrm_of(c, l, (int)m) : 0.0;
            mx[2 * lm + 1] = (l < c->lmax) ? hi * nrm_of(c, l + 1, (int)m) / nrm_of(c, l, (int)m) : 0.0;
        }
    }
}
extern "C" void st_dt_matrix(shtns_cfg c, double *mx) {
    mul_ct_matrix(c, mx);
    for (unsigned lm = 0; lm < c->nlm; ++lm) {
        const int l = c->li[lm];
        mx[2 * lm] *= (double)(l - 1);
        mx[2 * lm + 1] *= -(double)(l + 2);
    }
}
// R_lm = mx[2 lm] Q_{lm-1} + mx[2 lm + 1] Q_{lm+1} (Rlm must differ from Qlm)
extern "C" void SH_mul_mx(shtns_cfg c, double *mx, cplx *Qlm, cplx *Rlm) {
    const double *q = reinterpret_cast<const double *>(Qlm);
    double *r = reinterpret_cast<double *>(Rlm);
    const long n = c->nlm;
    for (long lm = 0; lm < n; ++lm) {
        double re = 0.0, ii = 0.0;
        if (lm > 0) { re += mx[2 * lm] * q[2 * lm - 2]; ii += mx[2 * lm] * q[2 * lm - 1]; }
        if (lm + 1 < n) { re += mx[2 * lm + 1] * q[2 * lm + 2]; ii += mx[2 * lm + 1] * q[2 * lm + 3]; }
        r[2 * lm] = re;
        r[2 * lm + 1] = ii;
    }
}

// std::complex<double> and C99 double complex share layout and return convention
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wreturn-type-c-linkage"
extern "C" cplx SH_to_point_cplx(shtns_cfg c, cplx *alm, double cost, double phi) {
    // z = sum_{l, -l<=m<=l} a_lm Y_l^m, Y_l^{-m} = (-1)^m conj(Y_l^m) (sht_func.c:516-580)
    if (c->mres != 1) runerr("complex SH requires mres=1.");
    const double *a = reinterpret_cast<const double *>(alm);
    const int L = c->lmax, M = c->mmax;
    auto lmc = [&](int l, int m) -> long { return (l <= M) ? (long)l * (l + 1) + m : (long)M * (2 * l - M) + l + m; };
    std::vector<double> y(L + 1);
    double zr = 0, zi = 0;
    for (int m = 0; m <= M; ++m) {
        legendre_sphPlm_array(c, L, m, cost, y.data());
        double pr = 0, pi = 0, nr = 0, ni = 0;
        for (int l = m; l <= L; ++l) {
            pr += y[l - m] * a[2 * lmc(l, m)];
            pi += y[l - m] * a[2 * lmc(l, m) + 1];
            if (m > 0) {
                nr += y[l - m] * a[2 * lmc(l, -m)];
                ni += y[l - m] * a[2 * lmc(l, -m) + 1];
            }
        }
        const double cm = cos(m * phi), sm = sin(m * phi), sg = (m & 1) ? -1.0 : 1.0;
        zr += pr * cm - pi * sm;
        zi += pr * sm + pi * cm;
        if (m > 0) {   // e^{-i m phi} (-1)^m (nr + i ni)
            zr += sg * (nr * cm + ni * sm);
            zi += sg * (ni * cm - nr * sm);
        }
    }
    cplx z;
    reinterpret_cast<double *>(&z)[0] = zr;
    reinterpret_cast<double *>(&z)[1] = zi;
    return z;
}
#pragma clang diagnostic pop

// order sums of the real fields at one latitude: for each m, Fq = sum Q P, Fs = sum S dP,
// Ft = sum T dP, Gs = sum S (P/sin), Gt = sum T (P/sin) (m >= 1); any of Q, S, T may be null
struct OrderSums {
    double q[2], ds[2], dt[2], ps[2], pt[2];
};
static OrderSums order_sums(shtns_cfg c, int im, int ltr, const std::vector<long double> &ys,
                            const std::vector<long double> &dy, double s, const cplx *Q, const cplx *S, const cplx *T) {
    OrderSums r{};
    const int m = im * c->mres;
    const long base = LiM(c, 0, im);
    const double *q = reinterpret_cast<const double *>(Q), *sl = reinterpret_cast<const double *>(S),
                 *tl = reinterpret_cast<const double *>(T);
    for (int l = m; l <= ltr; ++l) {
        const long i = 2 * (base + l);
        const double p = (double)(m >= 1 ? ys[l - m] * s : ys[l - m]), d = (double)dy[l - m],
                     po = (double)(m >= 1 ? ys[l - m] : 0.0L);
        for (int k = 0; k < 2; ++k) {
            if (q) r.q[k] += p * q[i + k];
            if (sl) { r.ds[k] += d * sl[i + k]; r.ps[k] += po * sl[i + k]; }
            if (tl) { r.dt[k] += d * tl[i + k]; r.pt[k] += po * tl[i + k]; }
        }
    }
    return r;
}

// Re(c_m F e^{i m phi}) with c_0 = 1, c_m = 2
static inline double re_e(const double *F, int m, double cm, double sm) {
    return (m == 0 ? 1.0 : 2.0) * (F[0] * cm - F[1] * sm);
}

// host spectrum times the normalization factors (Q itself for orthonormal harmonics)
static const cplx *host_scaled(shtns_cfg c, const cplx *Q, std::vector<cplx> &buf) {
    if (c->nrm.empty() || Q == nullptr) return Q;
    buf.assign(Q, Q + c->nlm);
    for (unsigned i = 0; i < c->nlm; ++i) buf[i] *= c->nrm[i];
    return buf.data();
}

// vr = sum Q Y;  vt = dS/dtheta + (1/sin) dT/dphi;  vp = (1/sin) dS/dphi - dT/dtheta
// (sht_func.c:672-715); SH_to_grad_point is the T = 0 case with Q the radial derivative
static void qst_point(shtns_cfg c, const cplx *Q, const cplx *S, const cplx *T, double cost, double phi, double *vr,
                      double *vt, double *vp) {
    std::vector<cplx> bq, bs, bt;
    Q = host_scaled(c, Q, bq);
    S = host_scaled(c, S, bs);
    T = host_scaled(c, T, bt);
    const double s = sqrt((1. - cost) * (1. + cost));
    std::vector<long double> ys, dy;
    double r = 0, t = 0, p = 0;
    for (int im = 0; im <= c->mmax; ++im) {
        const int m = im * c->mres;
        plm_deriv(m, c->lmax, cost, ys, dy);
        const OrderSums o = order_sums(c, im, c->lmax, ys, dy, s, Q, S, T);
        const double cm = cos(m * phi), sm = sin(m * phi);
        // (i m) G: real part -m G.im, imaginary part m G.re
        const double ims[2] = {-m * o.ps[1], m * o.ps[0]}, imt[2] = {-m * o.pt[1], m * o.pt[0]};
        r += re_e(o.q, m, cm, sm);
        t += re_e(o.ds, m, cm, sm) + re_e(imt, m, cm, sm);
        p += re_e(ims, m, cm, sm) - re_e(o.dt, m, cm, sm);
    }
    *vr = r;
    *vt = t;
    *vp = p;
}
extern "C" void SHqst_to_point(shtns_cfg c, cplx *Qlm, cplx *Slm, cplx *Tlm, double cost, double phi, double *vr,
                               double *vt, double *vp) {
    qst_point(c, Qlm, Slm, Tlm, cost, phi, vr, vt, vp);
}
extern "C" void SH_to_grad_point(shtns_cfg c, cplx *DrSlm, cplx *Slm, double cost, double phi, double *gr, double *gt,
                                 double *gp) {
    qst_point(c, DrSlm, Slm, nullptr, cost, phi, gr, gt, gp);
}

// Synthesis on nphi equispaced longitudes phi_j = 2 pi j / nphi of one latitude, degrees
// <= ltr and orders <= mtr (clamped as sht_func.c:740-743); the Fourier sums are evaluated
// directly, with each angle reduced exactly (m j mod nphi).
static void qst_lat(shtns_cfg c, const cplx *Q, const cplx *S, const cplx *T, double cost, double *vr, double *vt,
                    double *vp, int nphi, int ltr, int mtr) {
    std::vector<cplx> bq, bs, bt;
    Q = host_scaled(c, Q, bq);
    S = host_scaled(c, S, bs);
    T = host_scaled(c, T, bt);
    if (nphi <= 0) return;
    if (ltr > c->lmax) ltr = c->lmax;
    if (mtr > c->mmax) mtr = c->mmax;
    if (mtr * c->mres > ltr) mtr = ltr / c->mres;
    if (mtr * 2 * c->mres >= nphi) mtr = (nphi - 1) / (2 * c->mres);
    const double s = sqrt((1. - cost) * (1. + cost));
    for (int j = 0; j < nphi; ++j) {
        if (vr) vr[j] = 0;
        if (vt) vt[j] = 0;
        if (vp) vp[j] = 0;
    }
    if (ltr < 0 || mtr < 0) return;
    std::vector<long double> ys, dy;
    for (int im = 0; im <= mtr; ++im) {
        const int m = im * c->mres;
        plm_deriv(m, ltr, cost, ys, dy);
        const OrderSums o = order_sums(c, im, ltr, ys, dy, s, Q, S, T);
        const double ims[2] = {-m * o.ps[1], m * o.ps[0]}, imt[2] = {-m * o.pt[1], m * o.pt[0]};
        for (int j = 0; j < nphi; ++j) {
            const double ang = 2.0 * M_PI * (double)(((long)m * j) % nphi) / nphi;
            const double cm = cos(ang), sm = sin(ang);
            if (vr) vr[j] += re_e(o.q, m, cm, sm);
            if (vt) vt[j] += re_e(o.ds, m, cm, sm) + re_e(imt, m, cm, sm);
            if (vp) vp[j] += re_e(ims, m, cm, sm) - re_e(o.dt, m, cm, sm);
        }
    }
}
extern "C" void SH_to_lat(shtns_cfg c, cplx *Qlm, double cost, double *vr, int nphi, int ltr, int mtr) {
    qst_lat(c, Qlm, nullptr, nullptr, cost, vr, nullptr, nullptr, nphi, ltr, mtr);
}
extern "C" void SHqst_to_lat(shtns_cfg c, cplx *Qlm, cplx *Slm, cplx *Tlm, double cost, double *vr, double *vt,
                             double *vp, int nphi, int ltr, int mtr) {
    qst_lat(c, Qlm, Slm, Tlm, cost, vr, vt, vp, nphi, ltr, mtr);
}

// Rotation about the z axis by alpha (the frame by -alpha): R_lm = Q_lm e^{-i m alpha}, so the
// rotated field at longitude phi is the original at phi - alpha (sht_func.c:37-50).  Host
// arrays; Rlm may be Qlm.
extern "C" void SH_Zrotate(shtns_cfg c, cplx *Qlm, double alpha, cplx *Rlm) {
    const double *q = reinterpret_cast<const double *>(Qlm);
    double *r = reinterpret_cast<double *>(Rlm);
    for (int im = 0; im <= c->mmax; ++im) {
        const int m = im * c->mres;
        const double ca = cos(m * alpha), sa = sin(m * alpha);
        for (int l = m; l <= c->lmax; ++l) {
            const long i = 2 * LiM(c, l, im);
            const double re = q[i], ii = q[i + 1];
            r[i] = re * ca + ii * sa;
            r[i + 1] = ii * ca - re * sa;
        }
    }
}

// ------------------------------------------------------------------------------------------
// device initialization
// ------------------------------------------------------------------------------------------
template <class T>
static int upload(T **dst, const T *src, size_t n) {
    if (hipMalloc(dst, n * sizeof(T) + 256) != hipSuccess) return -1;
    if (hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return -1;
    return 0;
}

static int device_init(shtns_info *c) {
    int dev = (g_gpu_request >= 0) ? g_gpu_request : 0;
    HIP_OK_RET(hipSetDevice(dev), -1);
    c->device = dev;
    const int nl2 = c->nlat_2, L = c->lmax, mmax = c->mmax, mres = c->mres;
    DeviceTables &d = c->d;
    // latitude tables
    std::vector<double> x(nl2), x2(nl2), st(nl2), wq(nl2);
    for (int k = 0; k < nl2; ++k) {
        x[k] = c->ct[k];
        x2[k] = c->ct[k] * c->ct[k];
        st[k] = c->st[k];
        wq[k] = c->wg[k] * 2.0 * M_PI / c->nphi;
    }
    // recurrence tables: slots 0..mmax (m = im*mres) and slot mmax+1 (m = 1), degrees up to lmax+1
    std::vector<int> ofs(mmax + 3);
    int tot = 0;
    for (int slot = 0; slot <= mmax + 1; ++slot) {
        const int m = (slot <= mmax) ? slot * mres : 1;
        ofs[slot] = tot;
        tot += (L + 1 - m) / 2 + 8;   // rows continue past the last degree: block loads need no clamp
    }
    ofs[mmax + 2] = tot;
    std::vector<double2> ab(tot);
    std::vector<double4> pm(tot);
    for (int slot = 0; slot <= mmax + 1; ++slot) {
        const int m = (slot <= mmax) ? slot * mres : 1;
        build_order_tables(m, ofs[slot + 1] - ofs[slot], &ab[ofs[slot]], &pm[ofs[slot]]);
    }
    c->nsteps_total = tot;
    // FFT tables
    const int N = c->nphi;
    std::vector<double2> roots(N);
    for (int k = 0; k < N; ++k) {
        const long double a = 2.0L * 3.14159265358979323846264338327950288L * k / N;
        roots[k] = make_double2((double)cosl(a), (double)sinl(a));
    }
    std::vector<int> pos;
    fft_positions(d.plan, pos);

    if (upload(&d.x, x.data(), nl2) || upload(&d.x2, x2.data(), nl2) || upload(&d.st, st.data(), nl2) ||
        upload(&d.wq, wq.data(), nl2) || upload(&d.step_ofs, ofs.data(), ofs.size()) ||
        upload(&d.ab, ab.data(), ab.size()) || upload(&d.pm, pm.data(), pm.size()) ||
        upload(&d.roots, roots.data(), roots.size()) || upload(&d.fft_pos, pos.data(), pos.size()) ||
        upload(&d.tm, c->tm.data(), c->tm.size()) || (!c->nrm.empty() && upload(&d.nrm, c->nrm.data(), c->nrm.size()))) {
        set_error(2, "device table allocation failed");
        c->gpu_ready = true;
        free_device(c);
        return -1;
    }
    const size_t nseed = (size_t)(mmax + 2) * nl2;
    if (hipMalloc(&d.i0, nseed * sizeof(int)) != hipSuccess || hipMalloc(&d.z0, nseed * sizeof(double2)) != hipSuccess ||
        hipMalloc(&d.i0v, nseed * sizeof(int)) != hipSuccess || hipMalloc(&d.z0v, nseed * sizeof(double2)) != hipSuccess) {
        set_error(2, "device seed table allocation failed");
        c->gpu_ready = true;
        free_device(c);
        return -1;
    }
    // recurrences start where the functions exceed 1e-280 (their products with any
    // coefficient are then below double resolution of the result); the polar threshold
    // eps acts through the per-order latitude skip c->tm only, as in SHTns
    launch_seed_tables(&d, L, mmax, mres, nl2, 1e-280, nullptr);
    launch_seed_tables(&d, L, mmax, mres, nl2, 1e-280, nullptr, 1);
    HIP_OK_RET(hipGetLastError(), -1);
    HIP_OK_RET(hipDeviceSynchronize(), -1);
    // fixed workspace of cu_SH_to_spat_batch_noalloc: its address on this device (once per
    // device; a failure only disables that entry point)
    if (noalloc_scratch_setup(dev) != 0 && g_verbose)
        fprintf(stderr, "[shtns_amd] no fixed workspace on device %d: cu_SH_to_spat_batch_noalloc disabled\n", dev);
    c->gpu_ready = true;
    return dev;
}

extern "C" int cushtns_init_gpu(shtns_cfg c) {
    clear_error();
    if (c->ct == nullptr) {
        set_error(1, "cushtns_init_gpu: grid not set");
        return -1;
    }
    if (c->gpu_ready) return c->device;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        set_error(3, "no GPU device available");
        return -1;
    }
    return device_init(c);
}

extern "C" void cushtns_set_streams(shtns_cfg c, hipStream_t compute_stream, hipStream_t transfer_stream) {
    if (c->own_streams && c->comp_stream) (void)hipStreamDestroy(c->comp_stream);
    c->own_streams = false;
    c->comp_stream = compute_stream;
    c->xfer_stream = transfer_stream;
}

extern "C" shtns_cfg cushtns_clone(shtns_cfg c, hipStream_t compute_stream, hipStream_t transfer_stream) {
    clear_error();
    if (!c->gpu_ready) {
        set_error(3, "cushtns_clone: source config has no GPU tables");
        return nullptr;
    }
    shtns_info *k = new shtns_info(*c);
    k->clone_of = (c->clone_of) ? c->clone_of : c;
    k->ct = new double[2 * (size_t)c->nlat];
    memcpy(k->ct, c->ct, sizeof(double) * 2 * c->nlat);
    k->st = k->ct + c->nlat;
    k->scratch = nullptr;
    k->xscratch = nullptr;
    k->xscratch_bytes = 0;
    k->nscratch = nullptr;
    k->nscratch_bytes = 0;
    k->scratch_bytes = 0;
    k->own_streams = false;
    k->comp_stream = compute_stream;
    k->xfer_stream = transfer_stream;
    k->aux_stream = nullptr;
    k->ev_pool.clear();
    for (auto &v : k->prof_ev) v.clear();
    k->prof = false;
    std::lock_guard<std::mutex> g(g_mutex);
    k->next = g_cfg_list;
    g_cfg_list = k;
    return k;
}

// ------------------------------------------------------------------------------------------
// transforms
// ------------------------------------------------------------------------------------------
static bool ready(shtns_cfg c, const char *who) {
    if (c == nullptr || c->ct == nullptr) {
        set_error(1, "%s: grid not set", who);
        return false;
    }
    if (!c->gpu_ready) {
        if (cushtns_init_gpu(c) < 0) {
            set_error(3, "%s: no usable GPU (this library has no CPU transform path)", who);
            return false;
        }
    }
    (void)hipSetDevice(c->device);
    return true;
}

static void *scratch(shtns_cfg c, size_t bytes) {
    if (bytes <= c->scratch_bytes) return c->scratch;
    if (c->scratch) {
        (void)hipStreamSynchronize(c->comp_stream);
        (void)hipFree(c->scratch);
    }
    c->scratch = nullptr;
    c->scratch_bytes = 0;
    if (hipMalloc(&c->scratch, bytes) != hipSuccess) {
        set_error(2, "scratch allocation of %zu bytes failed", bytes);
        return nullptr;
    }
    c->scratch_bytes = bytes;
    return c->scratch;
}

static void spat_strides(shtns_cfg c, long &sp, long &st) {
    if (c->layout & SHT_PHI_CONTIGUOUS) { sp = 1; st = c->nphi; }
    else { sp = c->nlat; st = 1; }
}

static LegendreArgs leg_args(shtns_cfg c, int ltr, int imlim, int nfields) {
    LegendreArgs a;
    a.t = &c->d;
    a.lmax = c->lmax;
    a.mmax = c->mmax;
    a.mres = c->mres;
    a.nlat = c->nlat;
    a.nlat_2 = c->nlat_2;
    a.ltr = ltr;
    a.imlim = imlim;
    a.nfields = nfields;
    a.vector_mode = 0;
    a.mpos_scale = 1.0;
    a.lat_weight = 0;
    a.synth_ncg = c->tune_synth_ncg;
    a.anal_nw = c->tune_anal_nw;
    a.anal_ncg = c->tune_anal_ncg;
    return a;
}

static int clamp_ltr(shtns_cfg c, int ltr) { return (ltr < 0 || ltr > c->lmax) ? c->lmax : ltr; }
static int imlim_of(shtns_cfg c, int ltr) {
    const int im = ltr / c->mres;
    return im < c->mmax ? im : c->mmax;
}

// ---- optional kernel timing (HIP events recorded around each launch on the compute stream)
enum { PK_LEG_SYNTH = 0, PK_LEG_ANAL = 1, PK_FFT_SYNTH = 2, PK_FFT_ANAL = 3, PK_VEC = 4, PK_N = 5 };
struct ProfScope {
    shtns_cfg c;
    int k;
    hipStream_t s;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(shtns_cfg c_, int k_, hipStream_t s_ = nullptr, bool use_s = false, bool allowed = true)
        : c(c_), k(k_), s(use_s ? s_ : c_->comp_stream) {
        if (!c->prof || !allowed) return;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) { a = b = nullptr; return; }
        (void)hipEventRecord(a, s);
    }
    ~ProfScope() {
        if (!a) return;
        (void)hipEventRecord(b, s);
        c->prof_ev[k].emplace_back(a, b);
    }
};

// second stream + an event pool for the overlapped (chunked) transforms
static hipStream_t aux_stream(shtns_cfg c) {
    if (!c->aux_stream) {
        if (hipStreamCreateWithFlags(&c->aux_stream, hipStreamNonBlocking) != hipSuccess) c->aux_stream = nullptr;
    }
    return c->aux_stream;
}
static hipEvent_t pool_event(shtns_cfg c, size_t i) {
    while (c->ev_pool.size() <= i) {
        hipEvent_t e;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
        c->ev_pool.push_back(e);
    }
    return c->ev_pool[i];
}
// Overlapping measured slower at lmax=1023 (the analysis blocks use 150 KB of LDS and cannot
// share a CU with FFT blocks), so it is opt-in: SHTNS_AMD_OVERLAP=1.
static bool overlap_enabled() {
    static const bool on = getenv("SHTNS_AMD_OVERLAP") != nullptr;
    return on;
}

static void prof_clear(shtns_cfg c) {
    for (auto &v : c->prof_ev) {
        for (auto &e : v) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        v.clear();
    }
}

// Additive (no SHTns counterpart): enable/disable per-kernel timing; clears recorded events.
extern "C" void shtns_profile(shtns_cfg c, int enable) {
    prof_clear(c);
    c->prof = (enable != 0);
}

// Total milliseconds and launch counts per kernel kind since shtns_profile(c, 1):
// [0] Legendre synthesis, [1] Legendre analysis, [2] FFT synthesis, [3] FFT analysis,
// [4] vector conversions.  Waits for the recorded events.  Returns the number of kinds.
extern "C" int shtns_profile_get(shtns_cfg c, double *ms, int *count) {
    for (int k = 0; k < PK_N; ++k) {
        double t = 0.0;
        for (auto &e : c->prof_ev[k]) {
            float x = 0.f;
            (void)hipEventSynchronize(e.second);
            if (hipEventElapsedTime(&x, e.first, e.second) == hipSuccess) t += x;
        }
        ms[k] = t;
        count[k] = (int)c->prof_ev[k].size();
    }
    return PK_N;
}

// contraction precision for double data / float data (cushtns_set_precision)
static int prec_f64(shtns_cfg c) { return c->prec >= 0 ? c->prec : kPrecF64; }
static int prec_f32(shtns_cfg c) { return c->prec >= 0 ? c->prec : kPrecF32; }

static void *xscratch(shtns_cfg c, size_t bytes) {
    if (bytes <= c->xscratch_bytes) return c->xscratch;
    if (c->xscratch) {
        (void)hipStreamSynchronize(c->comp_stream);
        (void)hipFree(c->xscratch);
    }
    c->xscratch = nullptr;
    c->xscratch_bytes = 0;
    if (hipMalloc(&c->xscratch, bytes) != hipSuccess) {
        set_error(2, "staging allocation of %zu bytes failed", bytes);
        return nullptr;
    }
    c->xscratch_bytes = bytes;
    return c->xscratch;
}

static void *nscratch(shtns_cfg c, size_t bytes) {
    if (bytes <= c->nscratch_bytes) return c->nscratch;
    if (c->nscratch) {
        (void)hipStreamSynchronize(c->comp_stream);
        (void)hipFree(c->nscratch);
    }
    c->nscratch = nullptr;
    c->nscratch_bytes = 0;
    if (hipMalloc(&c->nscratch, bytes) != hipSuccess) {
        set_error(2, "normalization staging of %zu bytes failed", bytes);
        return nullptr;
    }
    c->nscratch_bytes = bytes;
    return c->nscratch;
}

// Non-orthonormal normalizations act on the spectra only: the transforms run on orthonormal
// harmonics with Q nrm (synthesis; Q / nrm for the adjoint of the analysis) as input, and the
// analysis output is divided by nrm (multiplied for the adjoint of the synthesis).
// scaled_input: a scaled copy of nf spectra (slot 0 or 1 of the staging), or Q itself
static const cplx *scaled_input(shtns_cfg c, const cplx *Q, long &spec_dist, int nf, int inv, int slot) {
    if (c->nrm.empty() || nf <= 0) return Q;
    const size_t one = (size_t)nf * c->nlm * sizeof(double2);
    char *b = (char *)nscratch(c, 2 * one);
    if (!b) return nullptr;
    double2 *out = reinterpret_cast<double2 *>(b + slot * one);
    launch_spec_scale(reinterpret_cast<const double2 *>(Q), spec_dist, out, c->nlm, c->nlm, nf, c->d.nrm, inv,
                      c->comp_stream);
    spec_dist = c->nlm;
    return reinterpret_cast<const cplx *>(out);
}
static void scale_output(shtns_cfg c, cplx *Q, long spec_dist, int nf, int inv) {
    if (c->nrm.empty() || nf <= 0 || g_last_err) return;
    launch_spec_scale(reinterpret_cast<const double2 *>(Q), spec_dist, reinterpret_cast<double2 *>(Q), spec_dist,
                      c->nlm, nf, c->d.nrm, inv, c->comp_stream);
}

// synthesis core: Q (nfields, spec_dist) -> V (nfields, spat_dist)
// The Legendre kernel runs in chunks of latitude tiles on the compute stream; the FFT of a
// chunk's latitudes (north and south rows) runs on the second stream as soon as the chunk
// is done, overlapping the next chunk's Legendre work (matrix cores vs LDS/HBM).
// V is double (vf == nullptr) or, for the FP32-data entry points, float: vf != nullptr, with
// vstage an FP64 staging area of nfields * nspat doubles when the register FFT cannot take
// the length (then the FFT writes vstage and a conversion kernel writes vf).
// Tables of the sliced (int8) contraction, built on first use by the configuration that owns
// the device tables (a clone builds them on its source and copies the pointers).
static bool ensure_i8_tables(shtns_cfg c) {
    shtns_info *own = c->clone_of ? c->clone_of : c;
    DeviceTables &d = own->d;
    if (!d.eA) {
        const int nslot = own->mmax + 1;
        std::vector<int> steps(nslot + 2), sbo(nslot + 1);
        if (hipMemcpy(steps.data(), d.step_ofs, sizeof(int) * (nslot + 1), hipMemcpyDeviceToHost) != hipSuccess) {
            set_error(2, "sliced tables: copy failed");
            return false;
        }
        int tot = 0, mx = 0;
        for (int sl = 0; sl < nslot; ++sl) {
            sbo[sl] = tot;
            const int nsb = (steps[sl + 1] - steps[sl] + 127) / 128;
            tot += nsb;
            mx = std::max(mx, nsb);
        }
        sbo[nslot] = tot;
        // balanced recurrence: w_i = 2^D(b, i & 1) z_i in the frame of K block b (32 steps),
        // D(b, p) = ilogb(d) of the middle step of parity p in the block (d_i = uo_i); within a
        // block the z_{i-1} coefficient is then exactly 1, and the state changes frame by a
        // power of two at each block start
        const int nall = own->nsteps_total;
        std::vector<double2> abk((size_t)tot * 128, make_double2(0.0, 0.0));
        std::vector<int2> dbk((size_t)tot * 4, make_int2(0, 0));
        std::vector<int> dl(nall, 0);
        for (int sl = 0; sl < nslot; ++sl) {
            const int m = sl * own->mres, n = steps[sl + 1] - steps[sl];
            if (n <= 0) continue;
            std::vector<double2> ab(n);
            std::vector<double4> pm(n);
            build_order_tables(m, n, ab.data(), pm.data());
            auto ilog = [&](int i) { return (pm[i].z != 0.0) ? ilogb(pm[i].z) : 0; };
            const int nb = (n + 31) / 32;
            for (int b = 0; b < nb; ++b) {
                int dp[2];
                for (int p = 0; p < 2; ++p) {
                    int i = 32 * b + 16 + p;
                    while (i >= n && i - 2 >= 32 * b) i -= 2;
                    dp[p] = ilog(std::min(i, n - 1));
                }
                dbk[(size_t)sbo[sl] * 4 + b] = make_int2(dp[0], dp[1]);
                for (int i = 32 * b; i < std::min(32 * b + 32, n); ++i) {
                    const int sh = dp[(i + 1) & 1] - dp[i & 1];
                    abk[(size_t)sbo[sl] * 128 + i] = make_double2(ldexp(ab[i].x, sh), ldexp(ab[i].y, sh));
                    dl[steps[sl] + i] = dp[i & 1];
                }
            }
        }
        int *dsb = nullptr, *ddl = nullptr;
        double2 *dabk = nullptr;
        int2 *ddbk = nullptr;
        short *dea = nullptr;
        if (upload(&dsb, sbo.data(), sbo.size()) || upload(&dabk, abk.data(), abk.size()) ||
            upload(&ddbk, dbk.data(), dbk.size()) || upload(&ddl, dl.data(), dl.size()) ||
            hipMalloc(&dea, sizeof(short) * (size_t)tot * own->nlat_2 + 256) != hipSuccess) {
            for (void *p : {(void *)dsb, (void *)dabk, (void *)ddbk, (void *)ddl})
                if (p) (void)hipFree(p);
            set_error(2, "sliced tables: allocation failed");
            return false;
        }
        d.sb_ofs = dsb;
        d.sb_total = tot;
        d.sb_max = mx;
        d.abk = dabk;
        d.dblk = ddbk;
        d.delta = ddl;
        d.eA = dea;
        launch_i8_rowexp(&d, own->mmax, own->nlat_2, c->comp_stream);
        HIP_OK_RET(hipGetLastError(), false);
    }
    if (c != own) {
        c->d.sb_ofs = d.sb_ofs;
        c->d.sb_total = d.sb_total;
        c->d.sb_max = d.sb_max;
        c->d.abk = d.abk;
        c->d.dblk = d.dblk;
        c->d.delta = d.delta;
        c->d.eA = d.eA;
    }
    return true;
}

// Interleaved complex spatial fields as the target / source of a batch of real fields: real
// field k is part k % 2 (real, imaginary) of complex field k / 2, at z + (k / 2) * zdist
// complex elements.  The FFTs then write / read the complex arrays directly.
struct SpatPairs {
    double *z;
    long zdist;
};

// whether the synthesis FFTs of this configuration can write interleaved pairs directly
static bool synth_pairs_ok(shtns_cfg c) {
    long sp, st;
    spat_strides(c, sp, st);
    return !overlap_enabled() && launch_fft_synth_pairs(&c->d, c->nlat, c->nphi, c->mmax, 1, nullptr, 1, 1, nullptr, 1,
                                                        sp, st, nullptr, c->tune_fft_synth, true);
}
// longitude FFTs of synthesis: F (nf real fields, field distance fdist) -> V, or into
// interleaved pairs (then synth_pairs_ok(c) holds)
static void fft_synth_to(shtns_cfg c, int imlim, int nf, const double2 *F, long fdist, double *V, long spat_dist,
                         const SpatPairs *zp, hipStream_t s, int kind = 0) {
    long sp, st;
    spat_strides(c, sp, st);
    if (!zp) {
        launch_fft_synth(&c->d, c->nlat, c->nphi, c->mmax, imlim, nf, F, fdist, V, spat_dist, sp, st, s, 0, -1, kind);
        return;
    }
    if (!launch_fft_synth_pairs(&c->d, c->nlat, c->nphi, imlim, nf / 2, F, 2 * fdist, fdist, zp->z, 2 * zp->zdist, sp,
                                st, s, kind))
        set_error(2, "synthesis into interleaved complex fields: no kernel for this size");
}

// longitude FFTs of analysis: V (nf real fields starting at real field k0) -> F
static void fft_anal_from(shtns_cfg c, int imlim, int nf, int k0, const double *V, long spat_dist, const SpatPairs *zp,
                          double2 *F, long four_dist, hipStream_t s, int kind = 0) {
    long sp, st;
    spat_strides(c, sp, st);
    if (!zp) {
        launch_fft_anal(&c->d, c->nlat, c->nphi, c->mmax, imlim, nf, V + (size_t)k0 * spat_dist, spat_dist, F,
                        four_dist, sp, st, s, 0, -1, kind);
        return;
    }
    const double *zb = zp->z + (size_t)(k0 / 2) * 2 * zp->zdist;
    if (launch_fft_anal_pairs(&c->d, c->nlat, c->nphi, c->mmax, imlim, nf / 2, zb, 2 * zp->zdist, F, 2 * four_dist,
                              four_dist, sp, st, s, kind))
        return;   // both parts in one launch
    for (int part = 0; part < 2; ++part)
        launch_fft_anal(&c->d, c->nlat, c->nphi, c->mmax, imlim, nf / 2, zb + part, 2 * zp->zdist,
                        F + (size_t)part * four_dist, 2 * four_dist, sp, st, s, 0, -1, kind, 2);
}

// Scratch bytes of one synth_core call (Fourier rows F, then the pre-mixed coefficients).
static size_t synth_scratch_bytes(shtns_cfg c, int nfields) {
    const size_t fbytes = (size_t)nfields * (c->mmax + 1) * c->nlat * sizeof(double2);
    return fbytes + synth_qmix_doubles(nfields, c->nsteps_total, c->tune_synth_ncg) * sizeof(double);
}

// ext != nullptr: a caller-provided workspace of at least synth_scratch_bytes(c, nfields)
// bytes is used instead of the configuration's scratch; no stream, event or memory is then
// created (no profiling, no overlapped chunks).
static void synth_core_on(shtns_cfg c, const cplx *Q, double *V, int ltr, int nfields, long spec_dist, long spat_dist,
                          double mpos_scale, int lat_weight, int prec, float *vf, double *vstage, void *ext,
                          const SpatPairs *zp);
// zp: the spatial output goes to interleaved complex fields instead of V (nfields even,
// synth_pairs_ok(c))
static void synth_core(shtns_cfg c, const cplx *Q, double *V, int ltr, int nfields, long spec_dist, long spat_dist,
                       double mpos_scale, int lat_weight, int prec, float *vf = nullptr, double *vstage = nullptr,
                       void *ext = nullptr, const SpatPairs *zp = nullptr) {
    if (nfields <= 0) return;
    if (!c->nrm.empty()) {
        if (ext) {
            set_error(1, "a non-orthonormal normalization needs a scaled copy of the spectra (allocating entry points only)");
            return;
        }
        Q = scaled_input(c, Q, spec_dist, nfields, lat_weight /* adjoint of the analysis: 1 / nrm */, 0);
        if (!Q) return;
    }
    synth_core_on(c, Q, V, ltr, nfields, spec_dist, spat_dist, mpos_scale, lat_weight, prec, vf, vstage, ext, zp);
}
static void synth_core_on(shtns_cfg c, const cplx *Q, double *V, int ltr, int nfields, long spec_dist, long spat_dist,
                          double mpos_scale, int lat_weight, int prec, float *vf, double *vstage, void *ext,
                          const SpatPairs *zp) {
    ltr = clamp_ltr(c, ltr);
    const int imlim = imlim_of(c, ltr);
    const long four_dist = (long)(c->mmax + 1) * c->nlat;
    const size_t fbytes = (size_t)nfields * four_dist * sizeof(double2);
    LegendreArgs a = leg_args(c, ltr, imlim, nfields);
    a.mpos_scale = mpos_scale;
    a.lat_weight = lat_weight;
    a.prec = prec;
    long sp, st;
    spat_strides(c, sp, st);
    const hipStream_t S = c->comp_stream;
    const bool prof = !ext;
    if (prec == kPrecSliced && !vf && !ext && !ensure_i8_tables(c)) return;
    // int8 digit planes of the coefficients (lmax <= 4095; above, the FP64 contraction runs)
    if (prec == kPrecSliced && !vf && !ext && i8_synth_lds_bytes(c->d.sb_max) <= kI8MaxLds) {
        a.synth_ncg = 2;   // pre-mix 32 columns per 8 fields
        const size_t qb = synth_qmix_doubles(nfields, c->nsteps_total, 2) * sizeof(double);
        const size_t pb = i8_planes_bytes(nfields, c->d.sb_total), eb = i8_eB_bytes(nfields, c->d.sb_total);
        char *b2 = (char *)scratch(c, fbytes + qb + pb + eb + 512);
        if (!b2) return;
        double2 *F = (double2 *)b2;
        double *q8 = (double *)(b2 + fbytes);
        unsigned char *pl = (unsigned char *)(((uintptr_t)(b2 + fbytes + qb) + 255) & ~(uintptr_t)255);
        int *eB = (int *)(((uintptr_t)(pl + pb) + 255) & ~(uintptr_t)255);
        {
            ProfScope ps(c, PK_LEG_SYNTH, nullptr, false, prof);
            launch_premix(a, reinterpret_cast<const double2 *>(Q), spec_dist, c->lmax, q8, c->nsteps_total, S);
            launch_legendre_synth_i8(a, c->lmax, F, four_dist, q8, c->nsteps_total, pl, eB, S);
        }
        ProfScope ps(c, PK_FFT_SYNTH, nullptr, false, prof);
        fft_synth_to(c, imlim, nfields, F, four_dist, V, spat_dist, zp, S, c->tune_fft_synth);
        HIP_OK(hipGetLastError());
        return;
    }
    char *base = ext ? (char *)ext : (char *)scratch(c, synth_scratch_bytes(c, nfields));
    if (!base) return;
    double2 *F = (double2 *)base;
    double *qmix = (double *)(base + fbytes);
    const int ntiles = (c->nlat_2 + 63) / 64;
    const int chunk_tiles = 4;
    const int nchunks = (ntiles + chunk_tiles - 1) / chunk_tiles;
    hipStream_t A = (overlap_enabled() && nchunks > 1 && !vf && !ext && !zp) ? aux_stream(c) : nullptr;
    {
        ProfScope ps(c, PK_LEG_SYNTH, nullptr, false, prof);
        launch_premix(a, reinterpret_cast<const double2 *>(Q), spec_dist, c->lmax, qmix, c->nsteps_total, S);
    }
    if (!A) {
        {
            ProfScope ps(c, PK_LEG_SYNTH, nullptr, false, prof);
            launch_legendre_synth_tiles(a, c->lmax, F, four_dist, qmix, c->nsteps_total, 0, ntiles, S);
        }
        ProfScope ps(c, PK_FFT_SYNTH, nullptr, false, prof);
        if (!vf) {
            fft_synth_to(c, imlim, nfields, F, four_dist, V, spat_dist, zp, S, c->tune_fft_synth);
        } else if (!vstage) {
            launch_fft_synth_wave(&c->d, c->nlat, c->nphi / 2, imlim, nfields, F, four_dist, vf, spat_dist, sp, st, S,
                                  0, c->nlat);
        } else {
            launch_fft_synth(&c->d, c->nlat, c->nphi, c->mmax, imlim, nfields, F, four_dist, vstage, c->nspat, sp, st,
                             S);
            launch_convert(vstage, c->nspat, vf, spat_dist, c->nspat, nfields, S);
        }
        HIP_OK(hipGetLastError());
        return;
    }
    for (int ch = 0; ch < nchunks; ++ch) {
        const int t0 = ch * chunk_tiles, ntl = std::min(chunk_tiles, ntiles - t0);
        {
            ProfScope ps(c, PK_LEG_SYNTH);
            launch_legendre_synth_tiles(a, c->lmax, F, four_dist, qmix, c->nsteps_total, t0, ntl, S);
        }
        hipEvent_t ev = pool_event(c, ch);
        HIP_OK(hipEventRecord(ev, S));
        HIP_OK(hipStreamWaitEvent(A, ev, 0));
        const int k0 = 64 * t0, k1 = std::min(64 * (t0 + ntl), (int)c->nlat_2);
        ProfScope ps(c, PK_FFT_SYNTH, A, true);
        launch_fft_synth(&c->d, c->nlat, c->nphi, c->mmax, imlim, nfields, F, four_dist, V, spat_dist, sp, st, A, k0,
                         k1 - k0);
        launch_fft_synth(&c->d, c->nlat, c->nphi, c->mmax, imlim, nfields, F, four_dist, V, spat_dist, sp, st, A,
                         c->nlat - k1, k1 - k0);
    }
    hipEvent_t done = pool_event(c, nchunks);
    HIP_OK(hipEventRecord(done, A));
    HIP_OK(hipStreamWaitEvent(S, done, 0));
    HIP_OK(hipGetLastError());
}

// analysis core: V -> Q.  With 16+ fields the FFTs of 8-field chunks run on the second
// stream ahead of the Legendre analysis of the previous chunk on the compute stream.
// V is double (vf == nullptr) or float (vf != nullptr, vstage as in synth_core: when set,
// vf is first converted into it).
static void anal_core_on(shtns_cfg c, const double *V, cplx *Q, int ltr, int nfields, long spat_dist, long spec_dist,
                         double mpos_scale, int lat_weight, int prec, const float *vf, double *vstage,
                         const SpatPairs *zp);
// zp: the spatial input is read from interleaved complex fields instead of V (nfields even)
static void anal_core(shtns_cfg c, const double *V, cplx *Q, int ltr, int nfields, long spat_dist, long spec_dist,
                      double mpos_scale, int lat_weight, int prec, const float *vf = nullptr,
                      double *vstage = nullptr, const SpatPairs *zp = nullptr) {
    if (nfields <= 0) return;
    anal_core_on(c, V, Q, ltr, nfields, spat_dist, spec_dist, mpos_scale, lat_weight, prec, vf, vstage, zp);
    scale_output(c, Q, spec_dist, nfields, lat_weight /* analysis: 1 / nrm; adjoint of the synthesis: nrm */);
}
static void anal_core_on(shtns_cfg c, const double *V, cplx *Q, int ltr, int nfields, long spat_dist, long spec_dist,
                         double mpos_scale, int lat_weight, int prec, const float *vf, double *vstage,
                         const SpatPairs *zp) {
    ltr = clamp_ltr(c, ltr);
    const int imlim = imlim_of(c, ltr);
    const long four_dist = (long)(c->mmax + 1) * c->nlat;
    LegendreArgs a0 = leg_args(c, ltr, imlim, nfields);
    a0.prec = prec;
    const int nsplit = anal_splits(a0);   // latitude-split analysis: partial spectra after F
    const size_t fbytes = (size_t)nfields * four_dist * sizeof(double2);
    const size_t pbytes = (nsplit > 1) ? (size_t)nsplit * nfields * c->nlm * sizeof(double2) : 0;
    char *fb = (char *)scratch(c, fbytes + pbytes);
    if (!fb) return;
    double2 *F = (double2 *)fb;
    double2 *part = pbytes ? (double2 *)(fb + fbytes) : nullptr;
    long sp, st;
    spat_strides(c, sp, st);
    const hipStream_t S = c->comp_stream;
    const int chunk = 8;
    const int nchunks = (nfields + chunk - 1) / chunk;
    hipStream_t A = (overlap_enabled() && nchunks > 1 && !vf) ? aux_stream(c) : nullptr;
    if (!A) {
        {
            ProfScope ps(c, PK_FFT_ANAL);
            if (!vf) {
                fft_anal_from(c, imlim, nfields, 0, V, spat_dist, zp, F, four_dist, S, c->tune_fft_anal);
            } else if (!vstage) {
                launch_fft_anal_wave(&c->d, c->nlat, c->nphi / 2, c->mmax, imlim, nfields, vf, spat_dist, F,
                                     four_dist, sp, st, S, 0, c->nlat);
            } else {
                launch_convert(vf, spat_dist, vstage, c->nspat, c->nspat, nfields, S);
                launch_fft_anal(&c->d, c->nlat, c->nphi, c->mmax, imlim, nfields, vstage, c->nspat, F, four_dist, sp,
                                st, S);
            }
        }
        LegendreArgs a = leg_args(c, ltr, imlim, nfields);
        a.mpos_scale = mpos_scale;
        a.lat_weight = lat_weight;
        a.prec = prec;
        a.part = part;
        a.part_dist = c->nlm;
        ProfScope ps(c, PK_LEG_ANAL);
        launch_legendre_anal(a, F, four_dist, reinterpret_cast<double2 *>(Q), spec_dist, c->lmax, S);
        HIP_OK(hipGetLastError());
        return;
    }
    hipEvent_t ready = pool_event(c, 0);
    HIP_OK(hipEventRecord(ready, S));
    HIP_OK(hipStreamWaitEvent(A, ready, 0));
    for (int ch = 0; ch < nchunks; ++ch) {
        const int f0 = ch * chunk, nf = std::min(chunk, nfields - f0);
        {
            ProfScope ps(c, PK_FFT_ANAL, A, true);
            fft_anal_from(c, imlim, nf, f0, V, spat_dist, zp, F + (size_t)f0 * four_dist, four_dist, A);
        }
        hipEvent_t ev = pool_event(c, 1 + ch);
        HIP_OK(hipEventRecord(ev, A));
        HIP_OK(hipStreamWaitEvent(S, ev, 0));
        LegendreArgs a = leg_args(c, ltr, imlim, nf);
        a.mpos_scale = mpos_scale;
        a.lat_weight = lat_weight;
        a.prec = prec;
        ProfScope ps(c, PK_LEG_ANAL);
        launch_legendre_anal(a, F + (size_t)f0 * four_dist, four_dist, reinterpret_cast<double2 *>(Q) + (size_t)f0 * spec_dist,
                             spec_dist, c->lmax, S);
    }
    HIP_OK(hipGetLastError());
}

// zt, zpp: the two components go to interleaved complex fields instead of Vt, Vp (nfields
// even, synth_pairs_ok(c))
static void vsynth_core(shtns_cfg c, const cplx *S, const cplx *T, double *Vt, double *Vp, int ltr, int nfields,
                        long spec_dist, long spat_dist, const SpatPairs *zt = nullptr, const SpatPairs *zpp = nullptr) {
    if (nfields <= 0) return;
    if (!c->nrm.empty()) {
        long d2 = spec_dist;
        S = scaled_input(c, S, spec_dist, nfields, 0, 0);
        T = scaled_input(c, T, d2, nfields, 0, 1);
        if (!S || !T) return;
    }
    ltr = clamp_ltr(c, ltr);
    const int imlim = imlim_of(c, ltr);
    const long four_dist = (long)(c->mmax + 1) * c->nlat;
    const long vw_dist = nlm_calc(c->lmax + 1, c->mmax, c->mres) + 2;
    const size_t fbytes = (size_t)2 * nfields * four_dist * sizeof(double2);
    const size_t vbytes = (size_t)2 * nfields * vw_dist * sizeof(double2);
    const size_t qbytes = synth_qmix_doubles(2 * nfields, c->nsteps_total, c->tune_synth_ncg) * sizeof(double);
    char *base = (char *)scratch(c, fbytes + vbytes + qbytes);
    if (!base) return;
    double2 *F = (double2 *)base;
    double2 *VW = (double2 *)(base + fbytes);
    double *qmix = (double *)(base + fbytes + vbytes);
    launch_sphtor_to_vw(reinterpret_cast<const double2 *>(S), reinterpret_cast<const double2 *>(T), spec_dist, VW,
                        vw_dist, c->lmax, c->mmax, c->mres, ltr, imlim, nfields, c->comp_stream);
    LegendreArgs a = leg_args(c, ltr + 1, imlim, 2 * nfields);
    a.vector_mode = 1;
    a.prec = prec_f64(c);
    launch_legendre_synth(a, VW, vw_dist, c->lmax + 1, F, four_dist, qmix, c->nsteps_total, c->comp_stream);
    fft_synth_to(c, imlim, nfields, F, 2 * four_dist, Vt, spat_dist, zt, c->comp_stream);
    fft_synth_to(c, imlim, nfields, F + four_dist, 2 * four_dist, Vp, spat_dist, zpp, c->comp_stream);
    HIP_OK(hipGetLastError());
}

// zt, zpp: the two components are read from interleaved complex fields instead of Vt, Vp
static void vanal_core(shtns_cfg c, const double *Vt, const double *Vp, cplx *S, cplx *T, int ltr, int nfields,
                       long spat_dist, long spec_dist, const SpatPairs *zt = nullptr, const SpatPairs *zpp = nullptr) {
    if (nfields <= 0) return;
    ltr = clamp_ltr(c, ltr);
    const int imlim = imlim_of(c, ltr);
    const long four_dist = (long)(c->mmax + 1) * c->nlat;
    const long vw_dist = nlm_calc(c->lmax + 1, c->mmax, c->mres) + 2;
    const size_t fbytes = (size_t)2 * nfields * four_dist * sizeof(double2);
    const size_t vbytes = (size_t)2 * nfields * vw_dist * sizeof(double2);
    LegendreArgs a = leg_args(c, ltr + 1, imlim, 2 * nfields);
    a.vector_mode = 1;
    a.lat_weight = 1;
    a.prec = prec_f64(c);
    const int nsplit = anal_splits(a);
    const size_t pbytes = (nsplit > 1) ? (size_t)nsplit * 2 * nfields * vw_dist * sizeof(double2) : 0;
    char *base = (char *)scratch(c, fbytes + vbytes + pbytes);
    if (!base) return;
    double2 *F = (double2 *)base;
    double2 *VW = (double2 *)(base + fbytes);
    a.part = pbytes ? (double2 *)(base + fbytes + vbytes) : nullptr;
    a.part_dist = vw_dist;
    fft_anal_from(c, imlim, nfields, 0, Vt, spat_dist, zt, F, 2 * four_dist, c->comp_stream);
    fft_anal_from(c, imlim, nfields, 0, Vp, spat_dist, zpp, F + four_dist, 2 * four_dist, c->comp_stream);
    launch_legendre_anal(a, F, four_dist, VW, vw_dist, c->lmax + 1, c->comp_stream);
    launch_vw_to_sphtor(VW, vw_dist, reinterpret_cast<double2 *>(S), reinterpret_cast<double2 *>(T), spec_dist,
                        c->lmax, c->mmax, c->mres, ltr, imlim, nfields, c->comp_stream);
    HIP_OK(hipGetLastError());
    scale_output(c, S, spec_dist, nfields, 1);
    scale_output(c, T, spec_dist, nfields, 1);
}

// ---- device-pointer API ----
extern "C" void cu_SH_to_spat_batch(shtns_cfg c, cplx *Q, double *V, int ltr, int nf, long spec_dist, long spat_dist) {
    clear_error();
    if (!ready(c, "cu_SH_to_spat")) return;
    synth_core(c, Q, V, ltr, nf, spec_dist, spat_dist, 1.0, 0, prec_f64(c));
}
extern "C" void cu_spat_to_SH_batch(shtns_cfg c, double *V, cplx *Q, int ltr, int nf, long spat_dist, long spec_dist) {
    clear_error();
    if (!ready(c, "cu_spat_to_SH")) return;
    anal_core(c, V, Q, ltr, nf, spat_dist, spec_dist, 1.0, 1, prec_f64(c));
}
// Additive, no SHTns counterpart: cu_SH_to_spat_batch with the Fourier rows and pre-mixed
// coefficients in the library's fixed per-GPU workspace (kNoallocScratchBytes), for callers
// that cannot allocate device memory (HIP graph capture, tight serving loops).  It makes no
// device allocation, stream or event.  The configuration must already hold its GPU tables.
// Every call on one GPU shares that workspace: calls that overlap on one GPU race.
extern "C" void cu_SH_to_spat_batch_noalloc(shtns_cfg c, cplx *Q, double *V, int ltr, int nf, long spec_dist,
                                            long spat_dist) {
    clear_error();
    if (c == nullptr || c->ct == nullptr || !c->gpu_ready) {
        set_error(3, "cu_SH_to_spat_batch_noalloc: configuration has no GPU tables (call cushtns_init_gpu first)");
        return;
    }
    if (nf < 0 || (nf > 0 && (spec_dist < (long)c->nlm || spat_dist < (long)c->nspat))) {
        set_error(1, "cu_SH_to_spat_batch_noalloc: bad batch (nfields=%d, spec_dist=%ld, spat_dist=%ld)", nf, spec_dist,
                  spat_dist);
        return;
    }
    if (nf == 0) return;
    const size_t need = synth_scratch_bytes(c, nf);
    if (need > kNoallocScratchBytes) {
        set_error(2, "cu_SH_to_spat_batch_noalloc: needs %zu bytes of workspace, above the fixed %zu bytes "
                     "(fewer fields per call, or cu_SH_to_spat_batch)", need, kNoallocScratchBytes);
        return;
    }
    (void)hipSetDevice(c->device);
    void *ws = noalloc_scratch_ptr(c->device);
    if (ws == nullptr) {
        set_error(3, "cu_SH_to_spat_batch_noalloc: no fixed workspace on device %d", c->device);
        return;
    }
    synth_core(c, Q, V, ltr, nf, spec_dist, spat_dist, 1.0, 0, prec_f64(c), nullptr, nullptr, ws);
}
extern "C" void cu_SH_to_spat(shtns_cfg c, cplx *Q, double *V, int ltr) {
    cu_SH_to_spat_batch(c, Q, V, ltr, 1, c->nlm, c->nspat);
}
extern "C" void cu_spat_to_SH(shtns_cfg c, double *V, cplx *Q, int ltr) {
    cu_spat_to_SH_batch(c, V, Q, ltr, 1, c->nspat, c->nlm);
}
extern "C" void cu_SH_to_spat_float_batch(shtns_cfg c, cplx_f *Q, float *V, int ltr, int nf, long spec_dist,
                                          long spat_dist) {
    clear_error();
    if (!ready(c, "cu_SH_to_spat_float")) return;
    if (nf <= 0) return;
    const bool stage = !fft_wave_supported(c->nphi / 2);
    const size_t qb = (size_t)nf * c->nlm * sizeof(double2);
    char *base = (char *)xscratch(c, qb + (stage ? (size_t)nf * c->nspat * sizeof(double) : 0));
    if (!base) return;
    double2 *qd = (double2 *)base;
    launch_convert(reinterpret_cast<const float *>(Q), 2 * spec_dist, reinterpret_cast<double *>(qd), 2L * c->nlm,
                   2L * c->nlm, nf, c->comp_stream);
    synth_core(c, reinterpret_cast<const cplx *>(qd), nullptr, ltr, nf, c->nlm, spat_dist, 1.0, 0, prec_f32(c), V,
               stage ? (double *)(base + qb) : nullptr);
}
extern "C" void cu_spat_to_SH_float_batch(shtns_cfg c, float *V, cplx_f *Q, int ltr, int nf, long spat_dist,
                                          long spec_dist) {
    clear_error();
    if (!ready(c, "cu_spat_to_SH_float")) return;
    if (nf <= 0) return;
    const bool stage = !fft_wave_supported(c->nphi / 2);
    const size_t qb = (size_t)nf * c->nlm * sizeof(double2);
    char *base = (char *)xscratch(c, qb + (stage ? (size_t)nf * c->nspat * sizeof(double) : 0));
    if (!base) return;
    double2 *qd = (double2 *)base;
    anal_core(c, nullptr, reinterpret_cast<cplx *>(qd), ltr, nf, spat_dist, c->nlm, 1.0, 1, prec_f32(c), V,
              stage ? (double *)(base + qb) : nullptr);
    if (g_last_err) return;
    launch_convert(reinterpret_cast<const double *>(qd), 2L * c->nlm, reinterpret_cast<float *>(Q), 2 * spec_dist,
                   2L * c->nlm, nf, c->comp_stream);
    HIP_OK(hipGetLastError());
}
extern "C" void cu_SH_to_spat_float(shtns_cfg c, cplx_f *Q, float *V, int ltr) {
    cu_SH_to_spat_float_batch(c, Q, V, ltr, 1, c->nlm, c->nspat);
}
extern "C" void cu_spat_to_SH_float(shtns_cfg c, float *V, cplx_f *Q, int ltr) {
    cu_spat_to_SH_float_batch(c, V, Q, ltr, 1, c->nspat, c->nlm);
}
extern "C" int cushtns_set_precision(shtns_cfg c, int prec) {
    clear_error();
    if (c == nullptr || prec < SHT_PREC_AUTO || prec > SHT_PREC_SLICED) {
        set_error(1, "cushtns_set_precision: unknown precision %d", prec);
        return -1;
    }
    c->prec = prec;
    return 0;
}
extern "C" void cu_SHsphtor_to_spat_batch(shtns_cfg c, cplx *S, cplx *T, double *Vt, double *Vp, int ltr, int nf,
                                          long spec_dist, long spat_dist) {
    clear_error();
    if (!ready(c, "cu_SHsphtor_to_spat")) return;
    vsynth_core(c, S, T, Vt, Vp, ltr, nf, spec_dist, spat_dist);
}
extern "C" void cu_spat_to_SHsphtor_batch(shtns_cfg c, double *Vt, double *Vp, cplx *S, cplx *T, int ltr, int nf,
                                          long spat_dist, long spec_dist) {
    clear_error();
    if (!ready(c, "cu_spat_to_SHsphtor")) return;
    vanal_core(c, Vt, Vp, S, T, ltr, nf, spat_dist, spec_dist);
}
extern "C" void cu_SHsphtor_to_spat(shtns_cfg c, cplx *S, cplx *T, double *Vt, double *Vp, int ltr) {
    cu_SHsphtor_to_spat_batch(c, S, T, Vt, Vp, ltr, 1, c->nlm, c->nspat);
}
extern "C" void cu_spat_to_SHsphtor(shtns_cfg c, double *Vt, double *Vp, cplx *S, cplx *T, int ltr) {
    cu_spat_to_SHsphtor_batch(c, Vt, Vp, S, T, ltr, 1, c->nspat, c->nlm);
}
extern "C" void cu_SH_to_spat_adjoint_batch(shtns_cfg c, double *V, cplx *Q, int ltr, int nf, long spat_dist,
                                            long spec_dist) {
    clear_error();
    if (!ready(c, "cu_SH_to_spat_adjoint")) return;
    anal_core(c, V, Q, ltr, nf, spat_dist, spec_dist, 2.0, 0, prec_f64(c));
}
extern "C" void cu_spat_to_SH_adjoint_batch(shtns_cfg c, cplx *Q, double *V, int ltr, int nf, long spec_dist,
                                            long spat_dist) {
    clear_error();
    if (!ready(c, "cu_spat_to_SH_adjoint")) return;
    synth_core(c, Q, V, ltr, nf, spec_dist, spat_dist, 0.5, 1, prec_f64(c));
}

// ---- host-pointer API: copy in, run on the GPU, copy out (synchronous) ----
// The device copies live in the configuration's staging slots (c->hstage): the i-th array a
// call stages reuses slot i, so repeated calls allocate nothing.  A slot is regrown (after
// the compute stream has drained) only when a call needs more than it holds.  Host arrays
// from shtns_malloc (pinned) copy at full link speed; pageable ones are staged by HIP.
struct HostStage {
    shtns_cfg c;
    size_t next = 0;
    void *dev(size_t bytes) {
        auto &slots = c->hstage;
        if (next == slots.size()) slots.emplace_back(nullptr, 0);
        auto &sl = slots[next];
        if (sl.second < bytes) {
            if (sl.first) {
                (void)hipStreamSynchronize(c->comp_stream);
                (void)hipFree(sl.first);
            }
            sl = {nullptr, 0};
            void *p = nullptr;
            if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
            sl = {p, bytes};
        }
        ++next;
        return sl.first;
    }
};

static bool h2d(void *d, const void *h, size_t n, hipStream_t s) {
    return hipMemcpyAsync(d, h, n, hipMemcpyHostToDevice, s) == hipSuccess;
}
static bool d2h(void *h, const void *d, size_t n, hipStream_t s) {
    return hipMemcpyAsync(h, d, n, hipMemcpyDeviceToHost, s) == hipSuccess;
}

extern "C" void SH_to_spat_l(shtns_cfg c, cplx *Qlm, double *Vr, int ltr) {
    clear_error();
    if (!ready(c, "SH_to_spat")) return;
    HostStage st{c, {}};
    const size_t qb = sizeof(cplx) * c->nlm, vb = sizeof(double) * c->nspat;
    cplx *dq = (cplx *)st.dev(qb);
    double *dv = (double *)st.dev(vb);
    if (!dq || !dv) { set_error(2, "SH_to_spat: device allocation failed"); return; }
    if (!h2d(dq, Qlm, qb, c->comp_stream)) { set_error(2, "SH_to_spat: copy failed"); return; }
    synth_core(c, dq, dv, ltr, 1, c->nlm, c->nspat, 1.0, 0, prec_f64(c));
    if (g_last_err) return;
    if (!d2h(Vr, dv, vb, c->comp_stream)) { set_error(2, "SH_to_spat: copy failed"); return; }
    HIP_OK(hipStreamSynchronize(c->comp_stream));
}

extern "C" void spat_to_SH_l(shtns_cfg c, double *Vr, cplx *Qlm, int ltr) {
    clear_error();
    if (!ready(c, "spat_to_SH")) return;
    HostStage st{c, {}};
    const size_t qb = sizeof(cplx) * c->nlm, vb = sizeof(double) * c->nspat;
    cplx *dq = (cplx *)st.dev(qb);
    double *dv = (double *)st.dev(vb);
    if (!dq || !dv) { set_error(2, "spat_to_SH: device allocation failed"); return; }
    if (!h2d(dv, Vr, vb, c->comp_stream)) { set_error(2, "spat_to_SH: copy failed"); return; }
    anal_core(c, dv, dq, ltr, 1, c->nspat, c->nlm, 1.0, 1, prec_f64(c));
    if (g_last_err) return;
    if (!d2h(Qlm, dq, qb, c->comp_stream)) { set_error(2, "spat_to_SH: copy failed"); return; }
    HIP_OK(hipStreamSynchronize(c->comp_stream));
}

extern "C" void SHsphtor_to_spat_l(shtns_cfg c, cplx *S, cplx *T, double *Vt, double *Vp, int ltr) {
    clear_error();
    if (!ready(c, "SHsphtor_to_spat")) return;
    HostStage st{c, {}};
    const size_t qb = sizeof(cplx) * c->nlm, vb = sizeof(double) * c->nspat;
    cplx *ds = (cplx *)st.dev(qb), *dt = (cplx *)st.dev(qb);
    double *dvt = (double *)st.dev(vb), *dvp = (double *)st.dev(vb);
    if (!ds || !dt || !dvt || !dvp) { set_error(2, "SHsphtor_to_spat: device allocation failed"); return; }
    const hipStream_t s = c->comp_stream;
    if (S) { if (!h2d(ds, S, qb, s)) return; } else HIP_OK(hipMemsetAsync(ds, 0, qb, s));
    if (T) { if (!h2d(dt, T, qb, s)) return; } else HIP_OK(hipMemsetAsync(dt, 0, qb, s));
    vsynth_core(c, ds, dt, dvt, dvp, ltr, 1, c->nlm, c->nspat);
    if (g_last_err) return;
    if (!d2h(Vt, dvt, vb, s) || !d2h(Vp, dvp, vb, s)) { set_error(2, "SHsphtor_to_spat: copy failed"); return; }
    HIP_OK(hipStreamSynchronize(s));
}

extern "C" void spat_to_SHsphtor_l(shtns_cfg c, double *Vt, double *Vp, cplx *S, cplx *T, int ltr) {
    clear_error();
    if (!ready(c, "spat_to_SHsphtor")) return;
    HostStage st{c, {}};
    const size_t qb = sizeof(cplx) * c->nlm, vb = sizeof(double) * c->nspat;
    cplx *ds = (cplx *)st.dev(qb), *dt = (cplx *)st.dev(qb);
    double *dvt = (double *)st.dev(vb), *dvp = (double *)st.dev(vb);
    if (!ds || !dt || !dvt || !dvp) { set_error(2, "spat_to_SHsphtor: device allocation failed"); return; }
    const hipStream_t s = c->comp_stream;
    if (!h2d(dvt, Vt, vb, s) || !h2d(dvp, Vp, vb, s)) { set_error(2, "spat_to_SHsphtor: copy failed"); return; }
    vanal_core(c, dvt, dvp, ds, dt, ltr, 1, c->nspat, c->nlm);
    if (g_last_err) return;
    if (!d2h(S, ds, qb, s) || !d2h(T, dt, qb, s)) { set_error(2, "spat_to_SHsphtor: copy failed"); return; }
    HIP_OK(hipStreamSynchronize(s));
}

extern "C" void SH_to_spat(shtns_cfg c, cplx *Qlm, double *Vr) { SH_to_spat_l(c, Qlm, Vr, c->lmax); }
extern "C" void spat_to_SH(shtns_cfg c, double *Vr, cplx *Qlm) { spat_to_SH_l(c, Vr, Qlm, c->lmax); }
extern "C" void SHsphtor_to_spat(shtns_cfg c, cplx *S, cplx *T, double *Vt, double *Vp) {
    SHsphtor_to_spat_l(c, S, T, Vt, Vp, c->lmax);
}
extern "C" void spat_to_SHsphtor(shtns_cfg c, double *Vt, double *Vp, cplx *S, cplx *T) {
    spat_to_SHsphtor_l(c, Vt, Vp, S, T, c->lmax);
}
extern "C" void SHsph_to_spat_l(shtns_cfg c, cplx *S, double *Vt, double *Vp, int ltr) {
    SHsphtor_to_spat_l(c, S, nullptr, Vt, Vp, ltr);
}
extern "C" void SHtor_to_spat_l(shtns_cfg c, cplx *T, double *Vt, double *Vp, int ltr) {
    SHsphtor_to_spat_l(c, nullptr, T, Vt, Vp, ltr);
}
extern "C" void SHsph_to_spat(shtns_cfg c, cplx *S, double *Vt, double *Vp) { SHsph_to_spat_l(c, S, Vt, Vp, c->lmax); }

/* 3-component transforms (radial scalar + tangential vector): the scalar transform of the
   radial component, then the vector transform of the tangential ones, on the same stream
   (SHTns composes them the same way on the GPU, sht_gpu.cu:531-577; host-pointer forms
   sht_com.c:42-48, 77-83). */
extern "C" void cu_SHqst_to_spat(shtns_cfg c, cplx *Q, cplx *S, cplx *T, double *Vr, double *Vt, double *Vp,
                                 int ltr) {
    cu_SH_to_spat(c, Q, Vr, ltr);
    if (g_last_err) return;
    cu_SHsphtor_to_spat(c, S, T, Vt, Vp, ltr);
}
extern "C" void cu_spat_to_SHqst(shtns_cfg c, double *Vr, double *Vt, double *Vp, cplx *Q, cplx *S, cplx *T,
                                 int ltr) {
    cu_spat_to_SH(c, Vr, Q, ltr);
    if (g_last_err) return;
    cu_spat_to_SHsphtor(c, Vt, Vp, S, T, ltr);
}
extern "C" void SHqst_to_spat_l(shtns_cfg c, cplx *Q, cplx *S, cplx *T, double *Vr, double *Vt, double *Vp,
                                int ltr) {
    SH_to_spat_l(c, Q, Vr, ltr);
    if (g_last_err) return;
    SHsphtor_to_spat_l(c, S, T, Vt, Vp, ltr);
}
extern "C" void spat_to_SHqst_l(shtns_cfg c, double *Vr, double *Vt, double *Vp, cplx *Q, cplx *S, cplx *T,
                                int ltr) {
    spat_to_SH_l(c, Vr, Q, ltr);
    if (g_last_err) return;
    spat_to_SHsphtor_l(c, Vt, Vp, S, T, ltr);
}
extern "C" void SHqst_to_spat(shtns_cfg c, cplx *Q, cplx *S, cplx *T, double *Vr, double *Vt, double *Vp) {
    SHqst_to_spat_l(c, Q, S, T, Vr, Vt, Vp, c->lmax);
}
extern "C" void spat_to_SHqst(shtns_cfg c, double *Vr, double *Vt, double *Vp, cplx *Q, cplx *S, cplx *T) {
    spat_to_SHqst_l(c, Vr, Vt, Vp, Q, S, T, c->lmax);
}
extern "C" void SHtor_to_spat(shtns_cfg c, cplx *T, double *Vt, double *Vp) { SHtor_to_spat_l(c, T, Vt, Vp, c->lmax); }

/* Complex-valued scalar fields (mres = 1).  The field's real and imaginary parts are
   transformed as one batch of two real fields; the spectra are split and merged between the
   complex layout (LM_cplx, -l <= m <= l) and the two parts' real-field spectra by small
   kernels.  SHTns transforms orders +m and -m on the CPU with one complex FFT
   (sht_func.c:943-1052); the Legendre work is the same. */
static bool cplx_ready(shtns_cfg c, const char *who) {
    if (!ready(c, who)) return false;
    if (c->mres != 1) runerr("complex SH requires mres=1.");
    return true;
}
static bool cplx_batch_ok(shtns_cfg c, const char *who, int nf, long spec_dist, long spat_dist) {
    if (!cplx_ready(c, who)) return false;
    if (nf > 0 && (spec_dist < (long)c->nlm_cplx || spat_dist < (long)c->nspat)) {
        set_error(1, "%s: spec_dist %ld < nlm_cplx %u or spat_dist %ld < nspat %u", who, spec_dist, c->nlm_cplx,
                  spat_dist, c->nspat);
        return false;
    }
    return nf > 0;
}
/* Batches of complex fields: field f at alm + f*spec_dist and z + f*spat_dist (complex
   elements).  The fields' real and imaginary parts go through the real kernels as one batch
   of 2 nf real fields.  The FFTs write (read) the interleaved complex arrays directly: the
   synthesis in one launch per batch when the cooperative FFT kernels serve the size, else
   through a staging batch and an interleave pass; the analysis in one launch per part. */
extern "C" void cu_SH_to_spat_cplx_batch(shtns_cfg c, cplx *alm, cplx *z, int nf, long spec_dist, long spat_dist) {
    clear_error();
    if (!cplx_batch_ok(c, "cu_SH_to_spat_cplx", nf, spec_dist, spat_dist)) return;
    const bool direct = synth_pairs_ok(c);
    const size_t qb = 2 * sizeof(double2) * c->nlm * nf, vb = direct ? 0 : 2 * sizeof(double) * c->nspat * nf;
    char *base = (char *)xscratch(c, qb + vb);
    if (!base) return;
    double2 *XY = (double2 *)base;
    launch_cplx_split(reinterpret_cast<const double2 *>(alm), spec_dist, XY, c->nlm, c->lmax, c->mmax, nf,
                      c->comp_stream);
    if (direct) {
        const SpatPairs zp{reinterpret_cast<double *>(z), spat_dist};
        synth_core(c, reinterpret_cast<const cplx *>(XY), nullptr, c->lmax, 2 * nf, c->nlm, c->nspat, 1.0, 0,
                   prec_f64(c), nullptr, nullptr, nullptr, &zp);
    } else {
        double *xy = (double *)(base + qb);
        synth_core(c, reinterpret_cast<const cplx *>(XY), xy, c->lmax, 2 * nf, c->nlm, c->nspat, 1.0, 0, prec_f64(c));
        if (g_last_err) return;
        launch_interleave(xy, reinterpret_cast<double2 *>(z), spat_dist, c->nspat, nf, c->comp_stream);
    }
    HIP_OK(hipGetLastError());
}
extern "C" void cu_spat_cplx_to_SH_batch(shtns_cfg c, cplx *z, cplx *alm, int nf, long spat_dist, long spec_dist) {
    clear_error();
    if (!cplx_batch_ok(c, "cu_spat_cplx_to_SH", nf, spec_dist, spat_dist)) return;
    const size_t qb = 2 * sizeof(double2) * c->nlm * nf;
    double2 *XY = (double2 *)xscratch(c, qb);
    if (!XY) return;
    const SpatPairs zp{reinterpret_cast<double *>(z), spat_dist};
    anal_core(c, nullptr, reinterpret_cast<cplx *>(XY), c->lmax, 2 * nf, c->nspat, c->nlm, 1.0, 1, prec_f64(c),
              nullptr, nullptr, &zp);
    if (g_last_err) return;
    launch_cplx_merge(XY, c->nlm, reinterpret_cast<double2 *>(alm), spec_dist, c->lmax, c->mmax, nf, c->comp_stream);
    HIP_OK(hipGetLastError());
}
extern "C" void cu_SH_to_spat_cplx(shtns_cfg c, cplx *alm, cplx *z) {
    cu_SH_to_spat_cplx_batch(c, alm, z, 1, c->nlm_cplx, c->nspat);
}
extern "C" void cu_spat_cplx_to_SH(shtns_cfg c, cplx *z, cplx *alm) {
    cu_spat_cplx_to_SH_batch(c, z, alm, 1, c->nspat, c->nlm_cplx);
}
/* Complex-valued 2D vector fields: S and T are each split into the spectra of the real and
   imaginary parts' vector fields; all parts go through the vector kernels as one batch of
   2 nf (SHTns: orders +m and -m per order, sht_func.c:1060-1200). */
extern "C" void cu_SHsphtor_to_spat_cplx_batch(shtns_cfg c, cplx *slm, cplx *tlm, cplx *zt, cplx *zp, int nf,
                                               long spec_dist, long spat_dist) {
    clear_error();
    if (!cplx_batch_ok(c, "cu_SHsphtor_to_spat_cplx", nf, spec_dist, spat_dist)) return;
    const bool direct = synth_pairs_ok(c);
    const size_t qb = 2 * sizeof(double2) * c->nlm * nf, vb = direct ? 0 : 2 * sizeof(double) * c->nspat * nf;
    char *base = (char *)xscratch(c, 2 * qb + 2 * vb);
    if (!base) return;
    double2 *SS = (double2 *)base, *TT = (double2 *)(base + qb);
    const hipStream_t s = c->comp_stream;
    launch_cplx_split(reinterpret_cast<const double2 *>(slm), spec_dist, SS, c->nlm, c->lmax, c->mmax, nf, s);
    launch_cplx_split(reinterpret_cast<const double2 *>(tlm), spec_dist, TT, c->nlm, c->lmax, c->mmax, nf, s);
    if (direct) {
        const SpatPairs pt{reinterpret_cast<double *>(zt), spat_dist}, pp{reinterpret_cast<double *>(zp), spat_dist};
        vsynth_core(c, reinterpret_cast<cplx *>(SS), reinterpret_cast<cplx *>(TT), nullptr, nullptr, c->lmax, 2 * nf,
                    c->nlm, c->nspat, &pt, &pp);
    } else {
        double *VT = (double *)(base + 2 * qb), *VP = (double *)(base + 2 * qb + vb);
        vsynth_core(c, reinterpret_cast<cplx *>(SS), reinterpret_cast<cplx *>(TT), VT, VP, c->lmax, 2 * nf, c->nlm,
                    c->nspat);
        if (g_last_err) return;
        launch_interleave(VT, reinterpret_cast<double2 *>(zt), spat_dist, c->nspat, nf, s);
        launch_interleave(VP, reinterpret_cast<double2 *>(zp), spat_dist, c->nspat, nf, s);
    }
    HIP_OK(hipGetLastError());
}
extern "C" void cu_spat_cplx_to_SHsphtor_batch(shtns_cfg c, cplx *zt, cplx *zp, cplx *slm, cplx *tlm, int nf,
                                               long spat_dist, long spec_dist) {
    clear_error();
    if (!cplx_batch_ok(c, "cu_spat_cplx_to_SHsphtor", nf, spec_dist, spat_dist)) return;
    const size_t qb = 2 * sizeof(double2) * c->nlm * nf;
    char *base = (char *)xscratch(c, 2 * qb);
    if (!base) return;
    double2 *SS = (double2 *)base, *TT = (double2 *)(base + qb);
    const hipStream_t s = c->comp_stream;
    const SpatPairs pt{reinterpret_cast<double *>(zt), spat_dist}, pp{reinterpret_cast<double *>(zp), spat_dist};
    vanal_core(c, nullptr, nullptr, reinterpret_cast<cplx *>(SS), reinterpret_cast<cplx *>(TT), c->lmax, 2 * nf,
               c->nspat, c->nlm, &pt, &pp);
    if (g_last_err) return;
    launch_cplx_merge(SS, c->nlm, reinterpret_cast<double2 *>(slm), spec_dist, c->lmax, c->mmax, nf, s);
    launch_cplx_merge(TT, c->nlm, reinterpret_cast<double2 *>(tlm), spec_dist, c->lmax, c->mmax, nf, s);
    HIP_OK(hipGetLastError());
}
extern "C" void cu_SHsphtor_to_spat_cplx(shtns_cfg c, cplx *slm, cplx *tlm, cplx *zt, cplx *zp) {
    cu_SHsphtor_to_spat_cplx_batch(c, slm, tlm, zt, zp, 1, c->nlm_cplx, c->nspat);
}
extern "C" void cu_spat_cplx_to_SHsphtor(shtns_cfg c, cplx *zt, cplx *zp, cplx *slm, cplx *tlm) {
    cu_spat_cplx_to_SHsphtor_batch(c, zt, zp, slm, tlm, 1, c->nspat, c->nlm_cplx);
}
extern "C" void SHsphtor_to_spat_cplx(shtns_cfg c, cplx *slm, cplx *tlm, cplx *zt, cplx *zp) {
    clear_error();
    if (!cplx_ready(c, "SHsphtor_to_spat_cplx")) return;
    HostStage st{c, {}};
    const size_t ab = sizeof(cplx) * c->nlm_cplx, zb = sizeof(cplx) * c->nspat;
    cplx *ds = (cplx *)st.dev(ab), *dt = (cplx *)st.dev(ab), *dzt = (cplx *)st.dev(zb), *dzp = (cplx *)st.dev(zb);
    if (!ds || !dt || !dzt || !dzp) { set_error(2, "SHsphtor_to_spat_cplx: device allocation failed"); return; }
    if (!h2d(ds, slm, ab, c->comp_stream) || !h2d(dt, tlm, ab, c->comp_stream)) {
        set_error(2, "SHsphtor_to_spat_cplx: copy failed");
        return;
    }
    cu_SHsphtor_to_spat_cplx(c, ds, dt, dzt, dzp);
    if (g_last_err) return;
    if (!d2h(zt, dzt, zb, c->comp_stream) || !d2h(zp, dzp, zb, c->comp_stream)) {
        set_error(2, "SHsphtor_to_spat_cplx: copy failed");
        return;
    }
    HIP_OK(hipStreamSynchronize(c->comp_stream));
}
extern "C" void spat_cplx_to_SHsphtor(shtns_cfg c, cplx *zt, cplx *zp, cplx *slm, cplx *tlm) {
    clear_error();
    if (!cplx_ready(c, "spat_cplx_to_SHsphtor")) return;
    HostStage st{c, {}};
    const size_t ab = sizeof(cplx) * c->nlm_cplx, zb = sizeof(cplx) * c->nspat;
    cplx *ds = (cplx *)st.dev(ab), *dt = (cplx *)st.dev(ab), *dzt = (cplx *)st.dev(zb), *dzp = (cplx *)st.dev(zb);
    if (!ds || !dt || !dzt || !dzp) { set_error(2, "spat_cplx_to_SHsphtor: device allocation failed"); return; }
    if (!h2d(dzt, zt, zb, c->comp_stream) || !h2d(dzp, zp, zb, c->comp_stream)) {
        set_error(2, "spat_cplx_to_SHsphtor: copy failed");
        return;
    }
    cu_spat_cplx_to_SHsphtor(c, dzt, dzp, ds, dt);
    if (g_last_err) return;
    if (!d2h(slm, ds, ab, c->comp_stream) || !d2h(tlm, dt, ab, c->comp_stream)) {
        set_error(2, "spat_cplx_to_SHsphtor: copy failed");
        return;
    }
    HIP_OK(hipStreamSynchronize(c->comp_stream));
}
/* complex-valued 3-component fields: the scalar then the vector transform
   (sht_func.c:1197-1210) */
extern "C" void SHqst_to_spat_cplx(shtns_cfg c, cplx *qlm, cplx *slm, cplx *tlm, cplx *zr, cplx *zt, cplx *zp);
extern "C" void spat_cplx_to_SHqst(shtns_cfg c, cplx *zr, cplx *zt, cplx *zp, cplx *qlm, cplx *slm, cplx *tlm);
/* Legendre transforms at one order m = im*mres, no FFT (shtns.h:329-334; the reference's
   own round trip time_SHT.c:929-930 and the complex transforms sht_func.c:943-1052 fix the
   scales):  SH_to_spat_ml: V[k] = sum_{l=m}^{ltr} Q[l-m] P_lm(x_k) at every grid latitude;
   spat_to_SH_ml: Q[l-m] = 2 pi sum_k w_k P_lm(x_k) V[k] for l <= ltr, 0 above.  At m = 0
   only real parts are used and the imaginary parts written are 0.  Host arrays; the work
   runs on the GPU. */
struct MlStage {
    HostStage st;
    double *T = nullptr, *D = nullptr, *ct = nullptr, *sn = nullptr, *w = nullptr;
    double2 *ab = nullptr, *Q = nullptr, *V = nullptr, *Q2 = nullptr, *V2 = nullptr;
    int m = 0, nl = 0, nq = 0;
};
// vector: tables P/sin and dP/dtheta (the recurrence runs on order max(m, 1)), two fields
static bool ml_prepare(shtns_cfg c, int im, int ltr, const char *who, MlStage &g, bool vector = false) {
    if (!ready(c, who)) return false;
    if (im < 0 || im > c->mmax) {
        set_error(1, "%s: order index im=%d outside [0, %d]", who, im, c->mmax);
        return false;
    }
    g.m = im * c->mres;
    if (ltr > c->lmax) ltr = c->lmax;
    g.nl = (ltr >= g.m) ? ltr - g.m + 1 : 0;
    g.nq = c->lmax - g.m + 1;
    const int nlat = c->nlat;
    std::vector<double> w(nlat);
    for (int k = 0; k < nlat; ++k) w[k] = c->wg[std::min(k, nlat - 1 - k)] * ((2 * k + 1 == nlat) ? 2.0 : 1.0);
    std::vector<double2> ab(std::max(g.nl, 2));
    const int m = (vector && g.m == 0) ? 1 : g.m;     // order of the recurrence
    ab[0] = make_double2(0.0, 0.0);
    ab[1] = make_double2((double)sqrtl(2.0L * m + 3), 0.0);
    for (int j = 2; j < g.nl; ++j) {
        const int l = m + j;
        const long double t1 = (long double)((long)(l + m) * (l - m));
        const long double t2 = (long double)((long)(l - 1 + m) * (l - 1 - m));
        ab[j] = make_double2((double)sqrtl(((long double)(2 * l + 1) * (2 * l - 1)) / t1),
                             (double)-sqrtl(((long double)(2 * l + 1) * t2) / ((long double)(2 * l - 3) * t1)));
    }
    const size_t nt = (size_t)std::max(g.nl, 1) * nlat;
    g.T = (double *)g.st.dev(nt * sizeof(double));
    g.ct = (double *)g.st.dev(nlat * sizeof(double));
    g.sn = (double *)g.st.dev(nlat * sizeof(double));
    g.w = (double *)g.st.dev(nlat * sizeof(double));
    g.ab = (double2 *)g.st.dev(ab.size() * sizeof(double2));
    g.Q = (double2 *)g.st.dev((size_t)g.nq * sizeof(double2));
    g.V = (double2 *)g.st.dev((size_t)nlat * sizeof(double2));
    if (vector) {
        g.D = (double *)g.st.dev(nt * sizeof(double));
        g.Q2 = (double2 *)g.st.dev((size_t)g.nq * sizeof(double2));
        g.V2 = (double2 *)g.st.dev((size_t)nlat * sizeof(double2));
    }
    if (!g.T || !g.ct || !g.sn || !g.w || !g.ab || !g.Q || !g.V || (vector && (!g.D || !g.Q2 || !g.V2))) {
        set_error(2, "%s: device allocation failed", who);
        return false;
    }
    const hipStream_t s = c->comp_stream;
    if (!h2d(g.ct, c->ct, nlat * sizeof(double), s) || !h2d(g.sn, c->st, nlat * sizeof(double), s) ||
        !h2d(g.w, w.data(), nlat * sizeof(double), s) || !h2d(g.ab, ab.data(), ab.size() * sizeof(double2), s)) {
        set_error(2, "%s: copy failed", who);
        return false;
    }
    // h2d copies from pageable host memory may return before the copy is done: wait before
    // the host vectors go out of scope
    if (hipStreamSynchronize(s) != hipSuccess) {
        set_error(2, "%s: copy failed", who);
        return false;
    }
    const long double amm = seed_amm(m);
    const double lseed = (double)log2l(fabsl(amm)), sgn = amm < 0 ? -1.0 : 1.0;
    if (g.nl > 0) {
        const int ktm = c->tm[im];   // SHTns zeroes these polar rows in its fixed-m kernels too
                                     // (fly_SH_to_spat.gen.c:335-345)
        if (vector) launch_ml_vtable(g.ct, g.sn, nlat, g.m, g.nl, g.ab, lseed, sgn, ktm, g.T, g.D, s);
        else launch_ml_table(g.ct, g.sn, nlat, m, g.nl, g.ab, lseed, sgn, ktm, g.T, s);
    }
    return true;
}
// one-order spectra (degrees m + j, j < n) times / divided by the normalization factors
static cplx *ml_scaled(shtns_cfg c, int im, cplx *Ql, int n, std::vector<cplx> &buf) {
    if (c->nrm.empty() || n <= 0) return Ql;
    const long b = LiM(c, 0, im);
    const int m = im * c->mres;
    buf.assign(Ql, Ql + n);
    for (int j = 0; j < n; ++j) buf[j] *= c->nrm[b + m + j];
    return buf.data();
}
static void ml_unscale(shtns_cfg c, int im, cplx *Ql, int n) {
    if (c->nrm.empty() || g_last_err) return;
    const long b = LiM(c, 0, im);
    const int m = im * c->mres;
    for (int j = 0; j < n; ++j) Ql[j] /= c->nrm[b + m + j];
}
extern "C" void SH_to_spat_ml(shtns_cfg c, int im, cplx *Ql, cplx *Vr, int ltr) {
    clear_error();
    MlStage g{HostStage{c, {}}};
    if (!ml_prepare(c, im, ltr, "SH_to_spat_ml", g)) return;
    const hipStream_t s = c->comp_stream;
    std::vector<cplx> qb;
    Ql = ml_scaled(c, im, Ql, g.nl, qb);
    if (g.nl > 0) {
        if (!h2d(g.Q, Ql, (size_t)g.nl * sizeof(cplx), s)) { set_error(2, "SH_to_spat_ml: copy failed"); return; }
        launch_ml_synth(g.T, c->nlat, g.nl, g.Q, g.m == 0, g.V, s);
    } else {
        HIP_OK(hipMemsetAsync(g.V, 0, (size_t)c->nlat * sizeof(double2), s));
    }
    HIP_OK(hipGetLastError());
    if (!d2h(Vr, g.V, (size_t)c->nlat * sizeof(cplx), s)) { set_error(2, "SH_to_spat_ml: copy failed"); return; }
    HIP_OK(hipStreamSynchronize(s));
}
extern "C" void spat_to_SH_ml(shtns_cfg c, int im, cplx *Vr, cplx *Ql, int ltr) {
    clear_error();
    MlStage g{HostStage{c, {}}};
    if (!ml_prepare(c, im, ltr, "spat_to_SH_ml", g)) return;
    const hipStream_t s = c->comp_stream;
    HIP_OK(hipMemsetAsync(g.Q, 0, (size_t)g.nq * sizeof(double2), s));
    if (g.nl > 0) {
        if (!h2d(g.V, Vr, (size_t)c->nlat * sizeof(cplx), s)) { set_error(2, "spat_to_SH_ml: copy failed"); return; }
        launch_ml_anal(g.T, g.w, c->nlat, g.nl, g.V, g.m == 0, 2.0 * M_PI, g.Q, s);
    }
    HIP_OK(hipGetLastError());
    if (!d2h(Ql, g.Q, (size_t)g.nq * sizeof(cplx), s)) { set_error(2, "spat_to_SH_ml: copy failed"); return; }
    HIP_OK(hipStreamSynchronize(s));
    ml_unscale(c, im, Ql, g.nq);
}

/* Vector Legendre transforms at one order m = im*mres (shtns.h:336-337):
   Vt_k = sum_l S_l dP_lm/dtheta + T_l (i m / sin) P_lm,  Vp_k = sum_l S_l (i m / sin) P_lm - T_l dP_lm/dtheta;
   S_l = 2 pi / (l (l+1)) sum_k w_k [Vt_k dP_lm/dtheta + Vp_k (-i m / sin) P_lm],
   T_l = 2 pi / (l (l+1)) sum_k w_k [Vt_k (-i m / sin) P_lm - Vp_k dP_lm/dtheta]
   (the per-order pieces of the vector transforms; real parts only at m = 0). */
extern "C" void SHsphtor_to_spat_ml(shtns_cfg c, int im, cplx *Sl, cplx *Tl, cplx *Vt, cplx *Vp, int ltr) {
    clear_error();
    MlStage g{HostStage{c, {}}};
    if (!ml_prepare(c, im, ltr, "SHsphtor_to_spat_ml", g, true)) return;
    const hipStream_t s = c->comp_stream;
    const size_t vb = (size_t)c->nlat * sizeof(double2);
    std::vector<cplx> sb, tb;
    Sl = ml_scaled(c, im, Sl, g.nl, sb);
    Tl = ml_scaled(c, im, Tl, g.nl, tb);
    if (g.nl > 0) {
        if (!h2d(g.Q, Sl, (size_t)g.nl * sizeof(cplx), s) || !h2d(g.Q2, Tl, (size_t)g.nl * sizeof(cplx), s)) {
            set_error(2, "SHsphtor_to_spat_ml: copy failed");
            return;
        }
        launch_ml_vsynth(g.T, g.D, c->nlat, g.nl, g.m, g.Q, g.Q2, g.V, g.V2, s);
    } else {
        HIP_OK(hipMemsetAsync(g.V, 0, vb, s));
        HIP_OK(hipMemsetAsync(g.V2, 0, vb, s));
    }
    HIP_OK(hipGetLastError());
    if (!d2h(Vt, g.V, vb, s) || !d2h(Vp, g.V2, vb, s)) { set_error(2, "SHsphtor_to_spat_ml: copy failed"); return; }
    HIP_OK(hipStreamSynchronize(s));
}
extern "C" void spat_to_SHsphtor_ml(shtns_cfg c, int im, cplx *Vt, cplx *Vp, cplx *Sl, cplx *Tl, int ltr) {
    clear_error();
    MlStage g{HostStage{c, {}}};
    if (!ml_prepare(c, im, ltr, "spat_to_SHsphtor_ml", g, true)) return;
    const hipStream_t s = c->comp_stream;
    const size_t qb = (size_t)g.nq * sizeof(double2), vb = (size_t)c->nlat * sizeof(double2);
    HIP_OK(hipMemsetAsync(g.Q, 0, qb, s));
    HIP_OK(hipMemsetAsync(g.Q2, 0, qb, s));
    if (g.nl > 0) {
        if (!h2d(g.V, Vt, vb, s) || !h2d(g.V2, Vp, vb, s)) { set_error(2, "spat_to_SHsphtor_ml: copy failed"); return; }
        launch_ml_vanal(g.T, g.D, g.w, c->nlat, g.nl, g.m, g.V, g.V2, g.Q, g.Q2, s);
    }
    HIP_OK(hipGetLastError());
    if (!d2h(Sl, g.Q, qb, s) || !d2h(Tl, g.Q2, qb, s)) { set_error(2, "spat_to_SHsphtor_ml: copy failed"); return; }
    HIP_OK(hipStreamSynchronize(s));
    ml_unscale(c, im, Sl, g.nq);
    ml_unscale(c, im, Tl, g.nq);
}

// shorthands of the one-order transforms (shtns.h:338-342): a zero spectrum for the missing
// component; the 3-component forms are the scalar then the vector transform
extern "C" void SHsph_to_spat_ml(shtns_cfg c, int im, cplx *Sl, cplx *Vt, cplx *Vp, int ltr) {
    std::vector<cplx> z((size_t)c->lmax + 1);
    SHsphtor_to_spat_ml(c, im, Sl, z.data(), Vt, Vp, ltr);
}
extern "C" void SHtor_to_spat_ml(shtns_cfg c, int im, cplx *Tl, cplx *Vt, cplx *Vp, int ltr) {
    std::vector<cplx> z((size_t)c->lmax + 1);
    SHsphtor_to_spat_ml(c, im, z.data(), Tl, Vt, Vp, ltr);
}
extern "C" void SHqst_to_spat_ml(shtns_cfg c, int im, cplx *Ql, cplx *Sl, cplx *Tl, cplx *Vr, cplx *Vt, cplx *Vp,
                                 int ltr) {
    SH_to_spat_ml(c, im, Ql, Vr, ltr);
    if (g_last_err) return;
    SHsphtor_to_spat_ml(c, im, Sl, Tl, Vt, Vp, ltr);
}
extern "C" void spat_to_SHqst_ml(shtns_cfg c, int im, cplx *Vr, cplx *Vt, cplx *Vp, cplx *Ql, cplx *Sl, cplx *Tl,
                                 int ltr) {
    spat_to_SH_ml(c, im, Vr, Ql, ltr);
    if (g_last_err) return;
    spat_to_SHsphtor_ml(c, im, Vt, Vp, Sl, Tl, ltr);
}

extern "C" void SH_to_spat_cplx(shtns_cfg c, cplx *alm, cplx *z) {
    clear_error();
    if (!cplx_ready(c, "SH_to_spat_cplx")) return;
    HostStage st{c, {}};
    const size_t ab = sizeof(cplx) * c->nlm_cplx, zb = sizeof(cplx) * c->nspat;
    cplx *da = (cplx *)st.dev(ab), *dz = (cplx *)st.dev(zb);
    if (!da || !dz) { set_error(2, "SH_to_spat_cplx: device allocation failed"); return; }
    if (!h2d(da, alm, ab, c->comp_stream)) { set_error(2, "SH_to_spat_cplx: copy failed"); return; }
    cu_SH_to_spat_cplx(c, da, dz);
    if (g_last_err) return;
    if (!d2h(z, dz, zb, c->comp_stream)) { set_error(2, "SH_to_spat_cplx: copy failed"); return; }
    HIP_OK(hipStreamSynchronize(c->comp_stream));
}
extern "C" void spat_cplx_to_SH(shtns_cfg c, cplx *z, cplx *alm) {
    clear_error();
    if (!cplx_ready(c, "spat_cplx_to_SH")) return;
    HostStage st{c, {}};
    const size_t ab = sizeof(cplx) * c->nlm_cplx, zb = sizeof(cplx) * c->nspat;
    cplx *da = (cplx *)st.dev(ab), *dz = (cplx *)st.dev(zb);
    if (!da || !dz) { set_error(2, "spat_cplx_to_SH: device allocation failed"); return; }
    if (!h2d(dz, z, zb, c->comp_stream)) { set_error(2, "spat_cplx_to_SH: copy failed"); return; }
    cu_spat_cplx_to_SH(c, dz, da);
    if (g_last_err) return;
    if (!d2h(alm, da, ab, c->comp_stream)) { set_error(2, "spat_cplx_to_SH: copy failed"); return; }
    HIP_OK(hipStreamSynchronize(c->comp_stream));
}
extern "C" void SHqst_to_spat_cplx(shtns_cfg c, cplx *qlm, cplx *slm, cplx *tlm, cplx *zr, cplx *zt, cplx *zp) {
    SH_to_spat_cplx(c, qlm, zr);
    if (g_last_err) return;
    SHsphtor_to_spat_cplx(c, slm, tlm, zt, zp);
}
extern "C" void spat_cplx_to_SHqst(shtns_cfg c, cplx *zr, cplx *zt, cplx *zp, cplx *qlm, cplx *slm, cplx *tlm) {
    spat_cplx_to_SH(c, zr, qlm);
    if (g_last_err) return;
    spat_cplx_to_SHsphtor(c, zt, zp, slm, tlm);
}
