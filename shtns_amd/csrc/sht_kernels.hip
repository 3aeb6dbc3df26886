// sht_kernels.hip -- gfx950 kernels of shtns_amd.
//
// Legendre step (per order m, all latitudes):
//   The functions are generated on the fly with the two-degree recurrence of Ishioka (2018),
//   in the form SHTns uses for its CPU kernels (sht_legendre.c:673-734,
//   kernel_SH_to_spat.gen.c:388-441):
//       z_{i+1} = (alpha_i x^2 + beta_i) z_i + z_{i-1},
//       P_{m+2i} = ue_i z_i + ue1_{i-1} z_{i-1},      P_{m+2i+1} = x uo_i z_i.
//   Equatorial symmetry: one latitude pair (x, -x) per lane; even degrees give E, odd give O,
//   north = E + O, south = E - O.
//   The contraction over degrees (synthesis) or latitudes (analysis) runs on the FP64 matrix
//   cores (v_mfma_f64_16x16x4f64): columns are 4 fields x (E.re, E.im, O.re, O.im); the
//   recurrence runs on the vector ALUs and hands its values to the matrix cores through LDS.
//   Each lane starts its recurrence at a precomputed step i0 with precomputed values
//   (seed tables), so no extended-exponent arithmetic is needed in the hot loop.
//
// FFT step (per latitude): real FFT of length nphi as a complex FFT of length nphi/2 in LDS,
//   in-place mixed-radix (2,3,4,5) decimation in frequency; theta-contiguous global layout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <unordered_map>
#include <mutex>

#include "sht_legendre.h"

namespace shtns_amd {


// ---------------------------------------------------------------------------------------
// Seed tables: for each order slot and latitude pair, the first recurrence step whose
// functions exceed `thr` in magnitude, and (z_{i0-1}, z_{i0}) there.  The recurrence is
// run with a power-of-two exponent carried beside the values (exact scaling), so orders
// whose sin^m underflows double precision start correctly (SHTns does the same with its
// SHT_SCALE_FACTOR, kernel_SH_to_spat.gen.c:278-348).  Latitude pairs k < tm[slot] (the
// polar optimization, computed on the host) never start: their rows of that order are
// zero in synthesis and left out of the analysis sums.
// ---------------------------------------------------------------------------------------
__global__ void seed_kernel(const double *__restrict__ xk, const double *__restrict__ stk,
                            const int *__restrict__ step_ofs, const double2 *__restrict__ ab,
                            const double4 *__restrict__ pm, const int *__restrict__ tm, int lmax, int mmax,
                            int mres, int nlat_2, double thr, int vec, int *__restrict__ i0_out,
                            double2 *__restrict__ z0_out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int slot = blockIdx.y;
    if (k >= nlat_2) return;
    if (k < tm[slot]) {
        i0_out[(size_t)slot * nlat_2 + k] = step_ofs[slot + 1] - step_ofs[slot];
        z0_out[(size_t)slot * nlat_2 + k] = make_double2(0.0, 0.0);
        return;
    }
    const int m = (slot <= mmax) ? slot * mres : 1;
    const int nst = step_ofs[slot + 1] - step_ofs[slot];
    const double2 *abm = ab + step_ofs[slot];
    const double4 *pmm = pm + step_ofs[slot];
    const double x = xk[k];
    const double x2 = x * x;
    const double s = stk[k];
    // sin^m as mant * 2^E (sin^(m-1) for the vector functions P/sin of orders m >= 1)
    int E = 0, es;
    double mant = 1.0;
    {
        double b = frexp(s, &es);
        int eb = es;
        int n = (vec && slot <= mmax && m >= 1) ? m - 1 : m;
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
    int Ec = 0;
    if (E < 0) Ec = -(((-E) + 255) / 256) * 256;
    double za = 0.0, zb = ldexp(mant, E - Ec);
    const int lthr = ilogb(thr);
    int i0 = nst;
    double2 seed = make_double2(0.0, 0.0);
    for (int i = 0; i < nst; ++i) {
        const double d = pmm[i].z;   // uo_i: P_{m+2i+1} = x uo_i z_i
        const double v = fabs(d * zb);
        bool sig;
        if (Ec == 0) sig = (v >= thr);
        else sig = (v != 0.0) && (ilogb(v) + Ec > lthr + 1);
        if (sig) {
            i0 = i;
            seed = make_double2(ldexp(za, Ec), ldexp(zb, Ec));
            break;
        }
        const double2 c = abm[i];
        const double zn = fma(fma(c.x, x2, c.y), zb, za);
        za = zb;
        zb = zn;
        if (Ec < 0 && fabs(zb) > 0x1p300) {
            za = ldexp(za, -256);
            zb = ldexp(zb, -256);
            Ec += 256;
        }
    }
    i0_out[(size_t)slot * nlat_2 + k] = i0;
    z0_out[(size_t)slot * nlat_2 + k] = seed;
    (void)lmax;
}

void launch_seed_tables(DeviceTables *t, int lmax, int mmax, int mres, int nlat_2, double thr, hipStream_t s,
                        int vec) {
    dim3 grid((nlat_2 + 63) / 64, mmax + 2);
    hipLaunchKernelGGL(seed_kernel, grid, dim3(64), 0, s, t->x, t->st, t->step_ofs, t->ab, t->pm, t->tm, lmax,
                       mmax, mres, nlat_2, thr, vec, vec ? t->i0v : t->i0, vec ? t->z0v : t->z0);
}

// ---------------------------------------------------------------------------------------
// Synthesis, step 1 (pre-mix): the coefficients are rearranged once per transform into the
// layout of the MFMA B operand, so the Legendre kernel reads them with plain coalesced
// loads and needs no block barrier:
//   qmix[g][slot steps][NCOL],  col = 16*cg + 4*field + (E.re, E.im, O.re, O.im),
//   E_i = ue_i Q_{m+2i} + ue1_i Q_{m+2i+2},  O_i = uo_i Q_{m+2i+1}    (g = group of 4*NCG fields)
// ---------------------------------------------------------------------------------------
template <int NCOL>
__global__ __launch_bounds__(256) void premix_kernel(const int *__restrict__ step_ofs, const double4 *__restrict__ pm,
                                                     const double2 *__restrict__ Q, long spec_dist, int lmax_spec,
                                                     double *__restrict__ qmix, int nsteps_total, int mmax, int mres,
                                                     int ltr, int imlim, int nfields, int vector_mode,
                                                     double mpos_scale) {
    constexpr int FPG = NCOL / 4;   // fields per group; one thread per (step, field): 4 columns
    const int im = blockIdx.x;
    const int g = blockIdx.y;
    if (im > imlim) return;
    const OrderInfo o = order_info(im, lmax_spec, mmax, mres, ltr, vector_mode);
    const int nsteps = o.nsteps;
    const double4 *pmm = pm + step_ofs[o.slot];
    double *dst = qmix + ((size_t)g * nsteps_total + step_ofs[o.slot]) * NCOL;
    const double qscale = (im == 0) ? 1.0 : mpos_scale;
    // rows nsteps .. end of the slot are written as zeros: the Legendre kernel reads them
    // unconditionally (clamped to the slot) when it runs past the last step
    const int slot_rows = step_ofs[o.slot + 1] - step_ofs[o.slot];
    for (int e = blockIdx.z * blockDim.x + threadIdx.x; e < slot_rows * FPG; e += gridDim.z * blockDim.x) {
        const int ii = e / FPG, fi = e % FPG;
        const int f = g * FPG + fi;
        double4 out = make_double4(0.0, 0.0, 0.0, 0.0);
        if (ii < nsteps && f < nfields) {
            const double2 *Qf = Q + (size_t)f * spec_dist + o.spec_base;
            const int le = o.mfun + 2 * ii;
            const double4 cf = pmm[ii];
            const double2 q0 = Qf[le];
            const double2 q1 = (le + 1 <= ltr) ? Qf[le + 1] : make_double2(0.0, 0.0);
            const double2 q2 = (le + 2 <= ltr) ? Qf[le + 2] : make_double2(0.0, 0.0);
            // columns (E.re, E.im, O.re, O.im): E = ue Q_l + ue1 Q_{l+2}, O = uo Q_{l+1}
            out.x = (cf.x * q0.x + cf.y * q2.x) * qscale;
            out.y = (cf.x * q0.y + cf.y * q2.y) * qscale;
            out.z = (cf.z * q1.x) * qscale;
            out.w = (cf.z * q1.y) * qscale;
            if (o.real_only) out.y = out.w = 0.0;
        }
        *reinterpret_cast<double4 *>(dst + (size_t)ii * NCOL + 4 * fi) = out;
    }
}

__global__ void spec_scale_kernel(const double2 *__restrict__ in, long in_dist, double2 *__restrict__ out,
                                  long out_dist, int nlm, const double *__restrict__ nrm, int inv) {
    const int lm = blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (lm >= nlm) return;
    const double c = inv ? 1.0 / nrm[lm] : nrm[lm];
    const double2 v = in[(size_t)f * in_dist + lm];
    out[(size_t)f * out_dist + lm] = make_double2(v.x * c, v.y * c);
}

void launch_spec_scale(const double2 *in, long in_dist, double2 *out, long out_dist, int nlm, int nf,
                       const double *nrm, int inv, hipStream_t s) {
    if (nf <= 0 || nlm <= 0) return;
    hipLaunchKernelGGL(spec_scale_kernel, dim3((nlm + 255) / 256, nf), dim3(256), 0, s, in, in_dist, out, out_dist, nlm,
                       nrm, inv);
}

int anal_splits(const LegendreArgs &a) {
    static const bool off = getenv("SHTNS_AMD_ANAL_SPLIT") && atoi(getenv("SHTNS_AMD_ANAL_SPLIT")) == 0;
    const int ntiles = (a.nlat_2 + 63) / 64;
    if (off || synth_force16() || a.prec != kPrecF64 || a.nfields > 2 || ntiles <= 12 || ntiles % 12 != 0) return 1;
    return ntiles / 12;
}

__global__ void sum_splits_kernel(const double2 *__restrict__ part, long part_dist, int nsplit, int nfields,
                                  double2 *__restrict__ Q, long spec_dist, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (i >= n) return;
    double2 acc = part[(size_t)f * part_dist + i];
    for (int k = 1; k < nsplit; ++k) {   // fixed order: deterministic
        const double2 v = part[((size_t)k * nfields + f) * part_dist + i];
        acc.x += v.x;
        acc.y += v.y;
    }
    Q[(size_t)f * spec_dist + i] = acc;
}

void launch_sum_splits(const double2 *part, long part_dist, int nsplit, int nfields, double2 *Q, long spec_dist,
                       long n, hipStream_t s) {
    hipLaunchKernelGGL(sum_splits_kernel, dim3((unsigned)((n + 255) / 256), nfields), dim3(256), 0, s, part,
                       part_dist, nsplit, nfields, Q, spec_dist, n);
}

bool anal_shape_ok(int nw, int nlr) {
    // the (waves, tiles per wave) cases compiled in anal_launch (sht_legendre.h)
    static const int maxr[13] = {0, 6, 6, 6, 6, 6, 6, 6, 6, 1, 1, 1, 1};
    return nw >= 1 && nw <= 12 && nlr >= 1 && nlr <= maxr[nw];
}

size_t synth_qmix_doubles(int nfields, int nsteps_total, int want_ncg) {
    const int ncg = synth_ncg(nfields, want_ncg);
    const int ng = (nfields + 4 * ncg - 1) / (4 * ncg);
    return (size_t)ng * nsteps_total * 16 * ncg;
}

void launch_premix(const LegendreArgs &a, const double2 *Qlm, long spec_dist, int spec_ltop, double *qmix,
                   int nsteps_total, hipStream_t s) {
    const DeviceTables *t = a.t;
    const int ncg = synth_ncg(a.nfields, a.synth_ncg);
    const int ng = (a.nfields + 4 * ncg - 1) / (4 * ncg);
    dim3 grid(a.mmax + 1, ng, 2);
#define PREMIX(NC)                                                                                             \
    hipLaunchKernelGGL(premix_kernel<NC>, grid, dim3(256), 0, s, t->step_ofs, t->pm, Qlm, spec_dist, spec_ltop, qmix, \
                       nsteps_total, a.mmax, a.mres, a.ltr, a.imlim, a.nfields, a.vector_mode, a.mpos_scale)
    if (ncg == 4) PREMIX(64);
    else if (ncg == 2) PREMIX(32);
    else PREMIX(16);
#undef PREMIX
}

void launch_legendre_synth_tiles(const LegendreArgs &a, int spec_ltop, double2 *F, long four_dist, const double *qmix,
                                 int nsteps_total, int tile0, int ntl, hipStream_t s) {
    if (a.prec == kPrecF32) launch_legendre_synth_tiles_f32(a, spec_ltop, F, four_dist, qmix, nsteps_total, tile0, ntl, s);
    else if (a.prec == kPrecSplit)
        launch_legendre_synth_tiles_split(a, spec_ltop, F, four_dist, qmix, nsteps_total, tile0, ntl, s);
    else synth_tiles_launch<kPrecF64>(a, spec_ltop, F, four_dist, qmix, nsteps_total, tile0, ntl, s);
}

void launch_legendre_anal(const LegendreArgs &a, const double2 *F, long four_dist, double2 *Q, long spec_dist,
                          int spec_ltop, hipStream_t s) {
    if (a.prec == kPrecF32) launch_legendre_anal_f32(a, F, four_dist, Q, spec_dist, spec_ltop, s);
    else if (a.prec == kPrecSplit) launch_legendre_anal_split(a, F, four_dist, Q, spec_dist, spec_ltop, s);
    else anal_launch<kPrecF64>(a, F, four_dist, Q, spec_dist, spec_ltop, s);
}

template <class TS, class TD>
__global__ void convert_kernel(const TS *__restrict__ src, long src_dist, TD *__restrict__ dst, long dst_dist, long n) {
    const TS *sf = src + (size_t)blockIdx.y * src_dist;
    TD *df = dst + (size_t)blockIdx.y * dst_dist;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        df[i] = (TD)sf[i];
}

void launch_convert(const float *src, long src_dist, double *dst, long dst_dist, long n, int nfields, hipStream_t s) {
    if (n <= 0 || nfields <= 0) return;
    const int nb = (int)std::min<long>((n + 255) / 256, 2048);
    hipLaunchKernelGGL((convert_kernel<float, double>), dim3(nb, nfields), dim3(256), 0, s, src, src_dist, dst,
                       dst_dist, n);
}
void launch_convert(const double *src, long src_dist, float *dst, long dst_dist, long n, int nfields, hipStream_t s) {
    if (n <= 0 || nfields <= 0) return;
    const int nb = (int)std::min<long>((n + 255) / 256, 2048);
    hipLaunchKernelGGL((convert_kernel<double, float>), dim3(nb, nfields), dim3(256), 0, s, src, src_dist, dst,
                       dst_dist, n);
}

// ---- complex-valued fields (mres = 1) ----
// A complex field z = x + i y is transformed as its two real fields x and y, one batch of 2.
// Spectral layouts: complex-field a[LM_cplx(l, m)], -l <= m <= l (shtns.h:114); real-field
// X[LM(l, m)], m >= 0.  With Y_l^{-m} = (-1)^m conj(Y_l^m), for m > 0 and b = (-1)^m a_{l,-m}:
//   X_lm = (a_lm + conj b) / 2,  Y_lm = (a_lm - conj b) / (2i);  m = 0: X = Re a, Y = Im a
// and back: a_lm = X + i Y, a_{l,-m} = (-1)^m (conj X + i conj Y)  (sht_func.c:900-935).
__device__ __forceinline__ long lm_real(int l, int m, int lmax) { return ((long)m * (2 * lmax + 1 - m)) / 2 + l; }
__device__ __forceinline__ long lm_cplx(int l, int m, int mmax) {
    return (l <= mmax) ? (long)l * (l + 1) + m : (long)mmax * (2 * l - mmax) + l + m;
}

// Batches: field f of the complex spectra at a + f * a_dist; its two real spectra X_f, Y_f at
// XY + 2 f nlm and XY + (2 f + 1) nlm (one batch of 2 nf real fields, field distance nlm).
// A block takes 8 consecutive orders m0 .. m0+7 (lane i = thread % 8) and blockDim/8 degrees
// at a time: for one l the 8 lanes touch 8 consecutive elements (128 bytes) of each l-row of
// the l-major complex layout, and for one order the degrees run contiguously in the m-major
// real layout, so both sides are coalesced.  No block-level synchronization.
__global__ void cplx_split_kernel(const double2 *__restrict__ a, long a_dist, double2 *__restrict__ XY, long nlm,
                                  int lmax, int mmax) {
    const int i = threadIdx.x & 7, m0 = 8 * blockIdx.x, m = m0 + i;
    if (m > mmax) return;
    const double2 *af = a + (size_t)blockIdx.y * a_dist;
    double2 *X = XY + (size_t)blockIdx.y * 2 * nlm, *Y = X + nlm;
    const double sgn = (m & 1) ? -1.0 : 1.0;
    for (int l = m0 + (int)(threadIdx.x >> 3); l <= lmax; l += blockDim.x >> 3) {
        if (l < m) continue;
        const double2 ap = af[lm_cplx(l, m, mmax)];
        const long lm = lm_real(l, m, lmax);
        if (m == 0) {
            X[lm] = make_double2(ap.x, 0.0);
            Y[lm] = make_double2(ap.y, 0.0);
        } else {
            const double2 an = af[lm_cplx(l, -m, mmax)];
            const double bx = sgn * an.x, by = sgn * an.y;
            X[lm] = make_double2(0.5 * (ap.x + bx), 0.5 * (ap.y - by));
            Y[lm] = make_double2(0.5 * (ap.y + by), 0.5 * (bx - ap.x));
        }
    }
}

__global__ void cplx_merge_kernel(const double2 *__restrict__ XY, long nlm, double2 *__restrict__ a, long a_dist,
                                  int lmax, int mmax) {
    const int i = threadIdx.x & 7, m0 = 8 * blockIdx.x, m = m0 + i;
    if (m > mmax) return;
    double2 *af = a + (size_t)blockIdx.y * a_dist;
    const double2 *X = XY + (size_t)blockIdx.y * 2 * nlm, *Y = X + nlm;
    const double sgn = (m & 1) ? -1.0 : 1.0;
    for (int l = m0 + (int)(threadIdx.x >> 3); l <= lmax; l += blockDim.x >> 3) {
        if (l < m) continue;
        const long lm = lm_real(l, m, lmax);
        const double2 x = X[lm], y = Y[lm];
        af[lm_cplx(l, m, mmax)] = make_double2(x.x - y.y, x.y + y.x);
        if (m > 0) af[lm_cplx(l, -m, mmax)] = make_double2(sgn * (x.x + y.y), sgn * (y.x - x.y));
    }
}

// z_f[i] = (x_f[i], y_f[i]); x_f = xy + 2 f n, y_f = x_f + n; z_f = z + f z_dist.  (The
// synthesis writes the two parts into a staging batch and interleaves them here: FFT stores of
// one part straight into the interleaved array would write half cache lines twice.)
__global__ void interleave_kernel(const double *__restrict__ xy, double2 *__restrict__ z, long z_dist, long n) {
    const double *x = xy + (size_t)blockIdx.y * 2 * n;
    double2 *zf = z + (size_t)blockIdx.y * z_dist;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        zf[i] = make_double2(x[i], x[n + i]);
}

void launch_cplx_split(const double2 *a, long a_dist, double2 *XY, long nlm, int lmax, int mmax, int nf, hipStream_t s) {
    if (nf <= 0) return;
    hipLaunchKernelGGL(cplx_split_kernel, dim3((mmax + 8) / 8, nf), dim3(256), 0, s, a, a_dist, XY, nlm, lmax, mmax);
}
void launch_cplx_merge(const double2 *XY, long nlm, double2 *a, long a_dist, int lmax, int mmax, int nf, hipStream_t s) {
    if (nf <= 0) return;
    hipLaunchKernelGGL(cplx_merge_kernel, dim3((mmax + 8) / 8, nf), dim3(256), 0, s, XY, nlm, a, a_dist, lmax, mmax);
}
void launch_interleave(const double *xy, double2 *z, long z_dist, long n, int nf, hipStream_t s) {
    if (n <= 0 || nf <= 0) return;
    const int nb = (int)std::min<long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(interleave_kernel, dim3(nb, nf), dim3(256), 0, s, xy, z, z_dist, n);
}

// ---- Legendre transforms at one order m (no FFT) ----
// T[j][k] = P_{m+j,m}(x_k), j < nl, every grid latitude k: one thread per latitude runs the
// three-term recurrence from the seed a_mm sin^m (a_mm = seed_sign 2^log2_seed, negative for
// odd m with the Condon-Shortley phase), carrying a separate binary exponent so the
// polar values underflow to zero only when they are below the double range.
// ab[j] = (a_j, b_j): P_{m+j} = a_j x P_{m+j-1} + b_j P_{m+j-2} (ab[1].x = sqrt(2m+3)).
// Latitudes within ktm of either pole (the polar optimization of this order) get zeros.
__global__ void ml_table_kernel(const double *__restrict__ ct, const double *__restrict__ st, int nlat, int m, int nl,
                                const double2 *__restrict__ ab, double log2_seed, double seed_sign, int ktm,
                                double *__restrict__ T) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nlat) return;
    const double x = ct[k], s = st[k];
    if ((m > 0 && !(s > 0.0)) || k < ktm || k >= nlat - ktm) {
        for (int j = 0; j < nl; ++j) T[(size_t)j * nlat + k] = 0.0;
        return;
    }
    const double lg = log2_seed + (m > 0 ? m * log2(s) : 0.0);
    int e = (int)floor(lg);
    double y0 = seed_sign * exp2(lg - e), y1 = 0.0;
    T[k] = ldexp(y0, e);
    if (nl > 1) {
        y1 = ab[1].x * x * y0;
        T[(size_t)nlat + k] = ldexp(y1, e);
    }
    for (int j = 2; j < nl; ++j) {
        const double2 c = ab[j];
        const double y2 = c.x * x * y1 + c.y * y0;
        y0 = y1;
        y1 = y2;
        if (fabs(y1) > 0x1p256) {
            y0 *= 0x1p-256;
            y1 *= 0x1p-256;
            e += 256;
        }
        T[(size_t)j * nlat + k] = ldexp(y1, e);
    }
}
// V[k] = sum_j Q[j] T[j][k]; real_only (m = 0): real parts only
__global__ void ml_synth_kernel(const double *__restrict__ T, int nlat, int nl, const double2 *__restrict__ Q,
                                int real_only, double2 *__restrict__ V) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nlat) return;
    double re = 0.0, im = 0.0;
    for (int j = 0; j < nl; ++j) {
        const double t = T[(size_t)j * nlat + k];
        const double2 q = Q[j];
        re = fma(t, q.x, re);
        im = fma(t, q.y, im);
    }
    V[k] = make_double2(re, real_only ? 0.0 : im);
}
// Q[j] = scale sum_k w_k T[j][k] V[k]: one block of 256 threads per degree, fixed-order tree
__global__ __launch_bounds__(256) void ml_anal_kernel(const double *__restrict__ T, const double *__restrict__ w,
                                                      int nlat, const double2 *__restrict__ V, int real_only,
                                                      double scale, double2 *__restrict__ Q) {
    __shared__ double sr[256], si[256];
    const int j = blockIdx.x, t = threadIdx.x;
    double re = 0.0, im = 0.0;
    for (int k = t; k < nlat; k += 256) {
        const double a = w[k] * T[(size_t)j * nlat + k];
        const double2 v = V[k];
        re = fma(a, v.x, re);
        im = fma(a, v.y, im);
    }
    sr[t] = re;
    si[t] = im;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) {
            sr[t] += sr[t + h];
            si[t] += si[t + h];
        }
        __syncthreads();
    }
    if (t == 0) Q[j] = make_double2(scale * sr[0], real_only ? 0.0 : scale * si[0]);
}

// Vector tables at one order: Y[j][k] = P_lm(x_k) / sin(theta_k) (m >= 1; unused at m = 0) and
// D[j][k] = dP_lm/dtheta at l = m + j.  For m >= 1 the recurrence runs on P/sin (seed
// a_mm sin^(m-1)) and dP_l/dtheta = l x (P_l/sin) - sqrt((2l+1)/(2l-1) (l^2-m^2)) (P_{l-1}/sin).
// For m = 0, dP_l0/dtheta = sqrt(l (l+1)) sin (P_l1/sin) from the order-1 recurrence (ab1,
// seed a_11).  Every grid latitude has sin(theta) > 0.
__global__ void ml_vtable_kernel(const double *__restrict__ ct, const double *__restrict__ st, int nlat, int m, int nl,
                                 const double2 *__restrict__ ab, double log2_seed, double seed_sign, int ktm,
                                 double *__restrict__ Y, double *__restrict__ D) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nlat) return;
    const double x = ct[k], s = st[k];
    if (m >= 1 && (k < ktm || k >= nlat - ktm)) {   // polar optimization of this order
        for (int o = 0; o < nl; ++o) {
            Y[(size_t)o * nlat + k] = 0.0;
            D[(size_t)o * nlat + k] = 0.0;
        }
        return;
    }
    const int mr = (m >= 1) ? m : 1;     // order of the recurrence run here
    if (mr > 1 && !(s > 0.0)) {           // P/sin(theta) vanishes at the poles for orders >= 2
        for (int o = 0; o < nl; ++o) {
            Y[(size_t)o * nlat + k] = 0.0;
            D[(size_t)o * nlat + k] = 0.0;
        }
        return;
    }
    const double lg = log2_seed + ((mr > 1) ? (mr - 1) * log2(s) : 0.0);
    int e = (int)floor(lg);
    double y0 = seed_sign * exp2(lg - e), y1 = 0.0;
    // j counts degrees l = mr + j of the recurrence; output index o = l - m
    auto emit = [&](int l, double yl, double ylm1) {
        const double c = sqrt((2.0 * l + 1.0) / (2.0 * l - 1.0) * ((double)l * l - (double)mr * mr));
        if (m >= 1) {
            Y[(size_t)(l - m) * nlat + k] = ldexp(yl, e);
            D[(size_t)(l - m) * nlat + k] = ldexp(l * x * yl - c * ylm1, e);
        } else {
            D[(size_t)l * nlat + k] = sqrt((double)l * (l + 1)) * s * ldexp(yl, e);
        }
    };
    if (m == 0) {
        Y[k] = 0.0;
        D[k] = 0.0;
    }
    const int lend = m + nl - 1;            // last degree written
    if (lend < mr) return;
    emit(mr, y0, 0.0);
    if (lend == mr) return;
    y1 = ab[1].x * x * y0;
    emit(mr + 1, y1, y0);
    for (int l = mr + 2; l <= lend; ++l) {
        const double2 cf = ab[l - mr];
        const double y2 = cf.x * x * y1 + cf.y * y0;
        y0 = y1;
        y1 = y2;
        if (fabs(y1) > 0x1p256) {
            y0 *= 0x1p-256;
            y1 *= 0x1p-256;
            e += 256;
        }
        emit(l, y1, y0);
    }
}
// Vt[k] = sum_j S_j D + T_j (i m) Y;  Vp[k] = sum_j S_j (i m) Y - T_j D  (real parts only at m = 0)
__global__ void ml_vsynth_kernel(const double *__restrict__ Y, const double *__restrict__ D, int nlat, int nl, int m,
                                 const double2 *__restrict__ S, const double2 *__restrict__ T, double2 *__restrict__ Vt,
                                 double2 *__restrict__ Vp) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nlat) return;
    double tr = 0, ti = 0, pr = 0, pi = 0;
    for (int j = 0; j < nl; ++j) {
        const double d = D[(size_t)j * nlat + k];
        const double my = (m >= 1) ? m * Y[(size_t)j * nlat + k] : 0.0;
        const double2 sj = S[j], tj = T[j];
        tr += d * sj.x - my * tj.y;
        ti += d * sj.y + my * tj.x;
        pr += -my * sj.y - d * tj.x;
        pi += my * sj.x - d * tj.y;
    }
    if (m == 0) ti = pi = 0.0;
    Vt[k] = make_double2(tr, ti);
    Vp[k] = make_double2(pr, pi);
}
// S_j = 2 pi / (l (l+1)) sum_k w_k [Vt D + Vp (-i m) Y];  T_j = 2 pi / (l (l+1)) sum_k w_k [Vt (-i m) Y - Vp D]
__global__ __launch_bounds__(256) void ml_vanal_kernel(const double *__restrict__ Y, const double *__restrict__ D,
                                                       const double *__restrict__ w, int nlat, int m,
                                                       const double2 *__restrict__ Vt, const double2 *__restrict__ Vp,
                                                       double2 *__restrict__ S, double2 *__restrict__ T) {
    __shared__ double red[4][256];
    const int j = blockIdx.x, t = threadIdx.x, l = m + j;
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    for (int k = t; k < nlat; k += 256) {
        const double a = w[k] * D[(size_t)j * nlat + k];
        const double b = (m >= 1) ? w[k] * m * Y[(size_t)j * nlat + k] : 0.0;
        const double2 vt = Vt[k], vp = Vp[k];
        a0 += a * vt.x + b * vp.y;      // S.re
        a1 += a * vt.y - b * vp.x;      // S.im
        a2 += b * vt.y - a * vp.x;      // T.re
        a3 += -b * vt.x - a * vp.y;     // T.im
    }
    red[0][t] = a0;
    red[1][t] = a1;
    red[2][t] = a2;
    red[3][t] = a3;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h)
            for (int q = 0; q < 4; ++q) red[q][t] += red[q][t + h];
        __syncthreads();
    }
    if (t == 0) {
        const double sc = (l > 0) ? 2.0 * M_PI / ((double)l * (l + 1)) : 0.0;
        const bool ro = (m == 0);
        S[j] = make_double2(sc * red[0][0], ro ? 0.0 : sc * red[1][0]);
        T[j] = make_double2(sc * red[2][0], ro ? 0.0 : sc * red[3][0]);
    }
}

void launch_ml_vtable(const double *ct, const double *st, int nlat, int m, int nl, const double2 *ab, double log2_seed,
                      double seed_sign, int ktm, double *Y, double *D, hipStream_t s) {
    hipLaunchKernelGGL(ml_vtable_kernel, dim3((nlat + 255) / 256), dim3(256), 0, s, ct, st, nlat, m, nl, ab, log2_seed,
                       seed_sign, ktm, Y, D);
}
void launch_ml_vsynth(const double *Y, const double *D, int nlat, int nl, int m, const double2 *S, const double2 *T,
                      double2 *Vt, double2 *Vp, hipStream_t s) {
    hipLaunchKernelGGL(ml_vsynth_kernel, dim3((nlat + 255) / 256), dim3(256), 0, s, Y, D, nlat, nl, m, S, T, Vt, Vp);
}
void launch_ml_vanal(const double *Y, const double *D, const double *w, int nlat, int nl, int m, const double2 *Vt,
                     const double2 *Vp, double2 *S, double2 *T, hipStream_t s) {
    hipLaunchKernelGGL(ml_vanal_kernel, dim3(nl), dim3(256), 0, s, Y, D, w, nlat, m, Vt, Vp, S, T);
}
void launch_ml_table(const double *ct, const double *st, int nlat, int m, int nl, const double2 *ab, double log2_seed,
                     double seed_sign, int ktm, double *T, hipStream_t s) {
    hipLaunchKernelGGL(ml_table_kernel, dim3((nlat + 255) / 256), dim3(256), 0, s, ct, st, nlat, m, nl, ab, log2_seed,
                       seed_sign, ktm, T);
}
void launch_ml_synth(const double *T, int nlat, int nl, const double2 *Q, int real_only, double2 *V, hipStream_t s) {
    hipLaunchKernelGGL(ml_synth_kernel, dim3((nlat + 255) / 256), dim3(256), 0, s, T, nlat, nl, Q, real_only, V);
}
void launch_ml_anal(const double *T, const double *w, int nlat, int nl, const double2 *V, int real_only, double scale,
                    double2 *Q, hipStream_t s) {
    hipLaunchKernelGGL(ml_anal_kernel, dim3(nl), dim3(256), 0, s, T, w, nlat, V, real_only, scale, Q);
}

void launch_legendre_synth(const LegendreArgs &a, const double2 *Qlm, long spec_dist, int spec_ltop,
                           double2 *F, long four_dist, double *qmix, int nsteps_total, hipStream_t s) {
    launch_premix(a, Qlm, spec_dist, spec_ltop, qmix, nsteps_total, s);
    launch_legendre_synth_tiles(a, spec_ltop, F, four_dist, qmix, nsteps_total, 0, (a.nlat_2 + 63) / 64, s);
}

}  // namespace shtns_amd

namespace shtns_amd {

// ---------------------------------------------------------------------------------------
// FFT along longitude.  A real sequence x of length N = nphi is handled as the complex
// sequence z[n] = x[2n] + i x[2n+1] of length M = N/2:
//   analysis:  Z = FFT_M(z);  X[k] = (Z[k] + conj Z[M-k])/2 + e^{-2 pi i k/N} (Z[k] - conj Z[M-k])/(2i)
//   synthesis: Z[k] = (X[k] + conj X[M-k]) + i e^{+2 pi i k/N} (X[k] - conj X[M-k]);  z = IFFT_M(Z)
// The M-point transform is an in-place mixed-radix decimation-in-frequency FFT in LDS:
// natural-order input, output X[k] at position pos[k] (digit reversal).
// T latitudes per block; consecutive threads take consecutive latitudes for global access.
// ---------------------------------------------------------------------------------------
__device__ inline double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// multiply by s*i (s = +1 or -1)
__device__ inline double2 cmul_si(double2 a, double s) { return make_double2(-s * a.y, s * a.x); }

// one DIF stage over all butterflies of T sequences of length M held in LDS
template <int P>
__device__ inline void dif_butterfly(double2 *a, int q, int n1, int twm, const double2 *__restrict__ roots, int N,
                                     double sgn) {
    double2 x[P];
#pragma unroll
    for (int r = 0; r < P; ++r) x[r] = a[r * q];
    double2 y[P];
    if (P == 2) {
        y[0] = cadd(x[0], x[1]);
        y[1] = csub(x[0], x[1]);
    } else if (P == 4) {
        const double2 t0 = cadd(x[0], x[2]), t1 = csub(x[0], x[2]);
        const double2 t2 = cadd(x[1], x[3]), t3 = cmul_si(csub(x[1], x[3]), sgn);
        y[0] = cadd(t0, t2);
        y[2] = csub(t0, t2);
        y[1] = cadd(t1, t3);
        y[3] = csub(t1, t3);
    } else if (P == 3) {
        const double c = -0.5, s = 0.86602540378443864676 * sgn;
        const double2 t1 = cadd(x[1], x[2]);
        const double2 t2 = make_double2(x[0].x + c * t1.x, x[0].y + c * t1.y);
        const double2 t3 = cmul_si(csub(x[1], x[2]), s);   // i*s*(x1-x2)
        y[0] = cadd(x[0], t1);
        y[1] = cadd(t2, t3);
        y[2] = csub(t2, t3);
    } else {
        // generic P: y_k = sum_r x_r w_P^{r k}, w_P^j = roots[j*N/P]
#pragma unroll
        for (int k = 0; k < P; ++k) {
            double2 acc = x[0];
#pragma unroll
            for (int r = 1; r < P; ++r) {
                const int e = ((r * k) % P) * (N / P);
                double2 w = roots[e];
                w.y *= sgn;
                acc = cadd(acc, cmul(x[r], w));
            }
            y[k] = acc;
        }
    }
    a[0] = y[0];
#pragma unroll
    for (int k = 1; k < P; ++k) {
        const int e = n1 * k * twm;   // < N since n1 < Ns/P, k < P, twm = N/Ns
        double2 w = roots[e];
        w.y *= sgn;
        a[k * q] = cmul(y[k], w);
    }
}

// Blocks are dealt round-robin to the 8 XCDs; give each XCD a contiguous range of latitude
// groups so concurrently running blocks on one XCD write neighbouring latitudes (whose
// theta-contiguous rows share cache lines in that XCD's L2).
__device__ inline int xcd_group(int bid, int G) {
    if (G % 8) return bid;
    return (bid % 8) * (G / 8) + bid / 8;
}

__device__ void fft_dif_inplace(double2 *buf, int T, int M, const FftPlan &plan, const double2 *__restrict__ roots,
                                double sgn) {
    const int N = 2 * M;
    int Ns = M;
    for (int s = 0; s < plan.nstages; ++s) {
        const int p = plan.radix[s];
        const int q = Ns / p;
        const int nb = M / p;
        const int twm = N / Ns;
        for (int b = threadIdx.x; b < T * nb; b += blockDim.x) {
            const int t = b % T, bb = b / T;
            const int g = bb / q, n1 = bb - g * q;
            double2 *a = buf + (size_t)t * M + (size_t)g * Ns + n1;
            switch (p) {
                case 4: dif_butterfly<4>(a, q, n1, twm, roots, N, sgn); break;
                case 2: dif_butterfly<2>(a, q, n1, twm, roots, N, sgn); break;
                case 3: dif_butterfly<3>(a, q, n1, twm, roots, N, sgn); break;
                case 5: dif_butterfly<5>(a, q, n1, twm, roots, N, sgn); break;
                case 7: dif_butterfly<7>(a, q, n1, twm, roots, N, sgn); break;
                default: break;
            }
        }
        __syncthreads();
        Ns = q;
    }
}

// Compile-time radix plan (same factor order as the host's factor_fft: 4s, then 2, 3, 5, 7),
// so kernels specialized on M have all butterfly index arithmetic on constants.
struct CPlan {
    int n;
    int r[24];
};
__host__ __device__ constexpr CPlan make_cplan(int M) {
    CPlan p{0, {}};
    int n = M;
    while (n % 4 == 0) { p.r[p.n++] = 4; n /= 4; }
    while (n % 2 == 0) { p.r[p.n++] = 2; n /= 2; }
    while (n % 3 == 0) { p.r[p.n++] = 3; n /= 3; }
    while (n % 5 == 0) { p.r[p.n++] = 5; n /= 5; }
    while (n % 7 == 0) { p.r[p.n++] = 7; n /= 7; }
    return p;
}
__host__ __device__ constexpr int cplan_ns(int M, int S) {
    CPlan p = make_cplan(M);
    int ns = M;
    for (int s = 0; s < S; ++s) ns /= p.r[s];
    return ns;
}

template <int M, int T, int S>
__device__ inline void fft_dif_stages_c(double2 *buf, const double2 *__restrict__ roots, double sgn) {
    constexpr CPlan P = make_cplan(M);
    if constexpr (S < P.n) {
        constexpr int p = P.r[S];
        constexpr int Ns = cplan_ns(M, S);
        constexpr int q = Ns / p;
        constexpr int N = 2 * M;
        constexpr int twm = N / Ns;
        for (int b = threadIdx.x; b < T * (M / p); b += 256) {
            const int t = b % T, bb = b / T;
            const int g = bb / q, n1 = bb - g * q;
            double2 *a = buf + t * M + g * Ns + n1;
            dif_butterfly<p>(a, q, n1, twm, roots, N, sgn);
        }
        __syncthreads();
        fft_dif_stages_c<M, T, S + 1>(buf, roots, sgn);
    }
}

// MC > 0: length M = MC known at compile time (T latitudes per block, TC); MC == 0: runtime plan.
template <int MC, int TC>
__global__ __launch_bounds__(256) void fft_synth_kernel(const double2 *__restrict__ F, long four_dist,
                                                        double *__restrict__ V, long spat_dist,
                                                        const double2 *__restrict__ roots, const int *__restrict__ pos,
                                                        FftPlan plan, int nlat, int imlim, long sp, long stt, int Trt,
                                                        int lat_base, int lat_end) {
    extern __shared__ double2 buf[];
    const int M = MC ? MC : plan.M;
    const int T = MC ? TC : Trt;
    const int lat0 = lat_base + xcd_group(blockIdx.x, gridDim.x) * T;
    const int f = blockIdx.y;
    const double2 *Ff = F + (size_t)f * four_dist;
    double *Vf = V + (size_t)f * spat_dist;
    const int Tv = min(T, lat_end - lat0);
    // build Z (natural order)
    for (int e = threadIdx.x; e < T * M; e += blockDim.x) {
        const int t = e % T, k = e / T;
        double2 Z = make_double2(0.0, 0.0);
        if (t < Tv) {
            const int lat = lat0 + t;
            const double2 Xk = (k <= imlim) ? Ff[(size_t)k * nlat + lat] : make_double2(0.0, 0.0);
            const int km = M - k;   // k=0 -> M (Nyquist, zero since M > mmax)
            const double2 Xm = (k > 0 && km <= imlim) ? Ff[(size_t)km * nlat + lat] : make_double2(0.0, 0.0);
            const double2 A = cadd(Xk, cconj(Xm));
            const double2 B = csub(Xk, cconj(Xm));
            const double2 wB = cmul(roots[k], B);
            Z = make_double2(A.x - wB.y, A.y + wB.x);   // A + i w B
        }
        buf[(size_t)t * M + k] = Z;
    }
    __syncthreads();
    if constexpr (MC > 0) fft_dif_stages_c<MC, TC, 0>(buf, roots, +1.0);
    else fft_dif_inplace(buf, T, M, plan, roots, +1.0);
    for (int e = threadIdx.x; e < T * M; e += blockDim.x) {
        const int t = e % T, n = e / T;
        if (t < Tv) {
            const double2 z = buf[(size_t)t * M + pos[n]];
            const int lat = lat0 + t;
            Vf[(size_t)(2 * n) * sp + (size_t)lat * stt] = z.x;
            Vf[(size_t)(2 * n + 1) * sp + (size_t)lat * stt] = z.y;
        }
    }
}

template <int MC, int TC>
__global__ __launch_bounds__(256) void fft_anal_kernel(const double *__restrict__ V, long spat_dist,
                                                       double2 *__restrict__ F, long four_dist,
                                                       const double2 *__restrict__ roots, const int *__restrict__ pos,
                                                       FftPlan plan, int nlat, int mmax, int imlim, long sp, long stt,
                                                       int Trt, int lat_base, int lat_end) {
    extern __shared__ double2 buf[];
    const int M = MC ? MC : plan.M;
    const int T = MC ? TC : Trt;
    const int lat0 = lat_base + xcd_group(blockIdx.x, gridDim.x) * T;
    const int f = blockIdx.y;
    const double *Vf = V + (size_t)f * spat_dist;
    double2 *Ff = F + (size_t)f * four_dist;
    const int Tv = min(T, lat_end - lat0);
    for (int e = threadIdx.x; e < T * M; e += blockDim.x) {
        const int t = e % T, n = e / T;
        double2 z = make_double2(0.0, 0.0);
        if (t < Tv) {
            const int lat = lat0 + t;
            z = make_double2(Vf[(size_t)(2 * n) * sp + (size_t)lat * stt], Vf[(size_t)(2 * n + 1) * sp + (size_t)lat * stt]);
        }
        buf[(size_t)t * M + n] = z;
    }
    __syncthreads();
    if constexpr (MC > 0) fft_dif_stages_c<MC, TC, 0>(buf, roots, -1.0);
    else fft_dif_inplace(buf, T, M, plan, roots, -1.0);
    for (int e = threadIdx.x; e < T * (mmax + 1); e += blockDim.x) {
        const int t = e % T, k = e / T;
        if (t >= Tv) continue;
        const int lat = lat0 + t;
        double2 X = make_double2(0.0, 0.0);
        if (k <= imlim) {
            const double2 Zk = buf[(size_t)t * M + pos[k]];
            const double2 Zm = buf[(size_t)t * M + pos[(M - k) % M]];
            const double2 A = make_double2(0.5 * (Zk.x + Zm.x), 0.5 * (Zk.y - Zm.y));
            // B = (Zk - conj Zm) / (2i) = -i/2 (Zk - conj Zm)
            const double2 D = make_double2(Zk.x - Zm.x, Zk.y + Zm.y);
            const double2 B = make_double2(0.5 * D.y, -0.5 * D.x);
            const double2 w = cconj(roots[k]);
            X = cadd(A, cmul(w, B));
            if (k == 0) X.y = 0.0;
        }
        Ff[(size_t)k * nlat + lat] = X;
    }
}

static int fft_lat_per_block(int M) {
    static const int t_env = getenv("SHTNS_AMD_FFT_T") ? atoi(getenv("SHTNS_AMD_FFT_T")) : 0;
    int T = (int)(65536 / (16L * M));   // LDS budget of 64 KB per block by default
    if (t_env > 0) T = t_env;
    while (T > 1 && 16L * M * T > 160 * 1024) --T;
    if (T < 1) T = 1;
    if (T > 8) T = 8;
    return T;
}

// latitudes per block of the specialized kernels: 2 while 2 sequences fit in 64 KB of LDS
static constexpr int fft_tc(int M) { return (32L * M <= 65536) ? 2 : 1; }

#define FFT_SIZES(X) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024) X(1536) X(2048) \
    X(3072) X(4096) X(6144)

bool lds_fft_specialized(int M) {
#define LDS_OK(MM) if (M == MM) return true;
    FFT_SIZES(LDS_OK)
#undef LDS_OK
    return false;
}

// FFT kernel for a length, layout and direction: 1 register FFT, one latitude per 2..256
// lanes (sht_fft.hip); 2 LDS-staged FFT (below); 3 / 4 block-cooperative register FFT with 256
// / 128 threads, 5 / 6 the same with 8-byte LDS staging and 256 / 512 threads (sht_fft.hip).
// The register kernel keeps each latitude's accesses separate, contiguous only in the
// phi-contiguous layout (sp == 1).  In the theta-contiguous layout, measured at lmax=1023 with
// 16 fields, kind 5: synthesis 0.29 ms (two blocks per CU; kind 3 0.37, LDS kernel 0.55),
// analysis 0.33 ms (three blocks per CU; kind 6 0.38, LDS kernel 0.57).
// kind (autotuner) or SHTNS_AMD_FFT=wave|lds|coop|coop128|coops|coops512 override.
// threads per block of the cooperative kinds (negative: 8-byte LDS staging)
static int coop_threads(int kind) { return kind == 4 ? 128 : kind == 5 ? -256 : kind == 6 ? -512 : 256; }

static int fft_choice(int M, long sp, int kind) {
    static const char *env = getenv("SHTNS_AMD_FFT");
    int want = kind;
    if (!want && env)
        want = !strcmp(env, "wave")       ? 1
               : !strcmp(env, "lds")      ? 2
               : !strcmp(env, "coop")     ? 3
               : !strcmp(env, "coop128")  ? 4
               : !strcmp(env, "coops")    ? 5
               : !strcmp(env, "coops512") ? 6
                                          : 0;
    if (want == 2 && lds_fft_specialized(M)) return 2;
    if (want >= 3 && want <= 6 && fft_coop_supported(M)) return want;
    if (want == 1 && fft_wave_supported(M)) return 1;
    if (sp == 1 && fft_wave_supported(M)) return 1;
    if (fft_coop_supported(M)) return 5;
    if (lds_fft_specialized(M)) return 2;
    return fft_wave_supported(M) ? 1 : 2;
}

// hipFuncSetAttribute once per kernel and size: kept out of the launch path so launches can
// be captured into HIP graphs
void ensure_lds_limit(const void *kern, size_t lds) {
    if (lds <= 65536) return;
    static std::mutex mu;
    static std::unordered_map<const void *, size_t> done;
    std::lock_guard<std::mutex> g(mu);
    size_t &have = done[kern];
    if (have >= lds) return;
    (void)hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    have = lds;
}

template <class K>
static void set_lds_limit(K kern, size_t lds) {
    ensure_lds_limit((const void *)kern, lds);
}

void launch_fft_synth(const DeviceTables *t, int nlat, int nphi, int mmax, int imlim, int nfields,
                      const double2 *F, long four_dist, double *V, long spat_dist, long sp, long st, hipStream_t s,
                      int lat_base, int nlat_chunk, int kind) {
    const int M = nphi / 2;
    if (nlat_chunk < 0) { lat_base = 0; nlat_chunk = nlat; }
    const int lat_end = lat_base + nlat_chunk;
    static const bool generic = getenv("SHTNS_AMD_FFT_GENERIC") != nullptr;
    const int choice = generic ? 2 : fft_choice(M, sp, kind);
    if (choice >= 3 && choice <= 6 &&
        launch_fft_synth_coop(t, nlat, M, imlim, nfields, F, four_dist, V, spat_dist, sp, st, s, lat_base, lat_end,
                              coop_threads(choice)))
        return;
    if (choice == 1 &&
        launch_fft_synth_wave(t, nlat, M, imlim, nfields, F, four_dist, V, spat_dist, sp, st, s, lat_base, lat_end))
        return;
#define FFT_SYN_CASE(MM)                                                                                        \
    if (!generic && M == MM) {                                                                                  \
        constexpr int TC = fft_tc(MM);                                                                          \
        const size_t lds = (size_t)TC * MM * sizeof(double2);                                                   \
        set_lds_limit(fft_synth_kernel<MM, TC>, lds);                                                           \
        hipLaunchKernelGGL((fft_synth_kernel<MM, TC>), dim3((nlat_chunk + TC - 1) / TC, nfields), dim3(256), lds, s, \
                           F, four_dist, V, spat_dist, t->roots, t->fft_pos, t->plan, nlat, imlim, sp, st, TC,  \
                           lat_base, lat_end);                                                                  \
        return;                                                                                                 \
    }
    FFT_SIZES(FFT_SYN_CASE)
#undef FFT_SYN_CASE
    const int T = fft_lat_per_block(M);
    const size_t lds = (size_t)T * M * sizeof(double2);
    set_lds_limit(fft_synth_kernel<0, 1>, lds);
    hipLaunchKernelGGL((fft_synth_kernel<0, 1>), dim3((nlat_chunk + T - 1) / T, nfields), dim3(256), lds, s, F,
                       four_dist, V, spat_dist, t->roots, t->fft_pos, t->plan, nlat, imlim, sp, st, T, lat_base, lat_end);
    (void)mmax;
}

bool launch_fft_synth_pairs(const DeviceTables *t, int nlat, int nphi, int imlim, int nf, const double2 *F, long fdist,
                            long fpart, double *V, long spat_dist, long sp, long st, hipStream_t s, int kind, bool dry) {
    const int M = nphi / 2;
    if (getenv("SHTNS_AMD_FFT_GENERIC") || !fft_coop_supported(M)) return false;
    const int choice = fft_choice(M, sp, kind);
    if (choice != 5 && choice != 6) return false;
    return launch_fft_synth_coops_pairs(t, nlat, M, imlim, nf, F, fdist, fpart, V, spat_dist, 2 * sp, 2 * st, s,
                                        coop_threads(choice), dry);
}

bool launch_fft_anal_pairs(const DeviceTables *t, int nlat, int nphi, int mmax, int imlim, int nf, const double *V,
                           long spat_dist, double2 *F, long fdist, long fpart, long sp, long st, hipStream_t s, int kind,
                           bool dry) {
    const int M = nphi / 2;
    if (getenv("SHTNS_AMD_FFT_GENERIC") || !fft_coop_supported(M)) return false;
    const int choice = fft_choice(M, sp, kind);
    if (choice != 5 && choice != 6) return false;
    return launch_fft_anal_coops_pairs(t, nlat, M, mmax, imlim, nf, V, spat_dist, F, fdist, fpart, 2 * sp, 2 * st, s,
                                       coop_threads(choice), dry);
}

void launch_fft_anal(const DeviceTables *t, int nlat, int nphi, int mmax, int imlim, int nfields, const double *V,
                     long spat_dist, double2 *F, long four_dist, long sp, long st, hipStream_t s, int lat_base,
                     int nlat_chunk, int kind, int es) {
    const int M = nphi / 2;
    if (nlat_chunk < 0) { lat_base = 0; nlat_chunk = nlat; }
    const int lat_end = lat_base + nlat_chunk;
    static const bool generic = getenv("SHTNS_AMD_FFT_GENERIC") != nullptr;
    const int choice = generic ? 2 : fft_choice(M, sp, kind);
    sp *= es;
    st *= es;
    if (choice >= 3 && choice <= 6 &&
        launch_fft_anal_coop(t, nlat, M, mmax, imlim, nfields, V, spat_dist, F, four_dist, sp, st, s, lat_base, lat_end,
                             coop_threads(choice)))
        return;
    if (choice == 1 &&
        launch_fft_anal_wave(t, nlat, M, mmax, imlim, nfields, V, spat_dist, F, four_dist, sp, st, s, lat_base, lat_end))
        return;
#define FFT_ANA_CASE(MM)                                                                                          \
    if (!generic && M == MM) {                                                                                    \
        constexpr int TC = fft_tc(MM);                                                                            \
        const size_t lds = (size_t)TC * MM * sizeof(double2);                                                     \
        set_lds_limit(fft_anal_kernel<MM, TC>, lds);                                                              \
        hipLaunchKernelGGL((fft_anal_kernel<MM, TC>), dim3((nlat_chunk + TC - 1) / TC, nfields), dim3(256), lds, s,   \
                           V, spat_dist, F, four_dist, t->roots, t->fft_pos, t->plan, nlat, mmax, imlim, sp, st, TC, \
                           lat_base, lat_end);                                                                     \
        return;                                                                                                   \
    }
    FFT_SIZES(FFT_ANA_CASE)
#undef FFT_ANA_CASE
    const int T = fft_lat_per_block(M);
    const size_t lds = (size_t)T * M * sizeof(double2);
    set_lds_limit(fft_anal_kernel<0, 1>, lds);
    hipLaunchKernelGGL((fft_anal_kernel<0, 1>), dim3((nlat_chunk + T - 1) / T, nfields), dim3(256), lds, s, V,
                       spat_dist, F, four_dist, t->roots, t->fft_pos, t->plan, nlat, mmax, imlim, sp, st, T, lat_base,
                       lat_end);
}

// ---------------------------------------------------------------------------------------
// Vector <-> two spin-1 scalar spectra (SHTns sht_private.h:296-337, 534-584).
//   synthesis  (m>0): V_l = i m T_l + (l-1) e_l S_{l-1} - (l+2) e_{l+1} S_{l+1}
//                     W_l = i m S_l - (l-1) e_l T_{l-1} + (l+2) e_{l+1} T_{l+1},   l = m..ltr+1
//              (m=0): V_l = sqrt(l(l+1)) Re S_l,  W_l = -sqrt(l(l+1)) Re T_l     (m=1 functions)
//   analysis   (m>0): S_l = [ l e_{l+1} v_{l+1} - (l+1) e_l v_{l-1} - i m w_l ] / (l(l+1))
//                     T_l = [ -i m v_l - l e_{l+1} w_{l+1} + (l+1) e_l w_{l-1} ] / (l(l+1))
//              (m=0): S_l = v_l / sqrt(l(l+1)),  T_l = -w_l / sqrt(l(l+1))
//   with e_l = sqrt((l-m)(l+m)/((2l+1)(2l-1))).  V,W live in a spectral layout of degree
//   lmax+1; field f's V is at VW + 2f*vw_dist, W at VW + (2f+1)*vw_dist.
// ---------------------------------------------------------------------------------------
__device__ inline double eps_lm(int l, int m) {
    if (l <= m) return 0.0;
    return sqrt((double)((long)(l - m) * (l + m)) / (double)((long)(2 * l + 1) * (2 * l - 1)));
}

__global__ void sphtor_to_vw_kernel(const double2 *__restrict__ S, const double2 *__restrict__ T, long spec_dist,
                                    double2 *__restrict__ VW, long vw_dist, int lmax, int mres, int ltr, int imlim) {
    const int im = blockIdx.y;
    const int f = blockIdx.z;
    const int m = im * mres;
    const long sb = ((long)im * (2 * lmax + 2 - (im + 1) * mres)) / 2;
    const long vb = ((long)im * (2 * (lmax + 1) + 2 - (im + 1) * mres)) / 2;
    const double2 *Sl = S + (size_t)f * spec_dist + sb;
    const double2 *Tl = T + (size_t)f * spec_dist + sb;
    double2 *Vl = VW + (size_t)(2 * f) * vw_dist + vb;
    double2 *Wl = VW + (size_t)(2 * f + 1) * vw_dist + vb;
    for (int l = m + blockIdx.x * blockDim.x + threadIdx.x; l <= lmax + 1; l += gridDim.x * blockDim.x) {
        double2 v = make_double2(0.0, 0.0), w = make_double2(0.0, 0.0);
        if (im <= imlim) {
            if (m == 0) {
                if (l >= 1 && l <= ltr) {
                    const double c = sqrt((double)l * (l + 1));
                    v.x = c * Sl[l].x;
                    w.x = -c * Tl[l].x;
                }
            } else if (l <= ltr + 1) {
                if (l <= ltr) {
                    const double2 t = Tl[l], s = Sl[l];
                    v = make_double2(-m * t.y, m * t.x);
                    w = make_double2(-m * s.y, m * s.x);
                }
                if (l - 1 >= m) {
                    const double c = (l - 1) * eps_lm(l, m);
                    const double2 s = Sl[l - 1], t = Tl[l - 1];
                    v.x += c * s.x; v.y += c * s.y;
                    w.x -= c * t.x; w.y -= c * t.y;
                }
                if (l + 1 <= ltr) {
                    const double c = -(l + 2) * eps_lm(l + 1, m);
                    const double2 s = Sl[l + 1], t = Tl[l + 1];
                    v.x += c * s.x; v.y += c * s.y;
                    w.x -= c * t.x; w.y -= c * t.y;
                }
            }
        }
        Vl[l] = v;
        Wl[l] = w;
    }
    if (m == 0 && blockIdx.x == 0 && threadIdx.x == 0) {
        Vl[0] = make_double2(0.0, 0.0);
        Wl[0] = make_double2(0.0, 0.0);
    }
}

__global__ void vw_to_sphtor_kernel(const double2 *__restrict__ VW, long vw_dist, double2 *__restrict__ S,
                                    double2 *__restrict__ T, long spec_dist, int lmax, int mres, int ltr, int imlim) {
    const int im = blockIdx.y;
    const int f = blockIdx.z;
    const int m = im * mres;
    const long sb = ((long)im * (2 * lmax + 2 - (im + 1) * mres)) / 2;
    const long vb = ((long)im * (2 * (lmax + 1) + 2 - (im + 1) * mres)) / 2;
    double2 *Sl = S + (size_t)f * spec_dist + sb;
    double2 *Tl = T + (size_t)f * spec_dist + sb;
    const double2 *v = VW + (size_t)(2 * f) * vw_dist + vb;
    const double2 *w = VW + (size_t)(2 * f + 1) * vw_dist + vb;
    for (int l = m + blockIdx.x * blockDim.x + threadIdx.x; l <= lmax; l += gridDim.x * blockDim.x) {
        double2 s = make_double2(0.0, 0.0), t = make_double2(0.0, 0.0);
        if (im <= imlim && l <= ltr && l >= 1) {
            if (m == 0) {
                const double c = 1.0 / sqrt((double)l * (l + 1));
                s.x = c * v[l].x;
                t.x = -c * w[l].x;
            } else {
                const double inv = 1.0 / ((double)l * (l + 1));
                const double2 vl = v[l], wl = w[l], vu = v[l + 1], wu = w[l + 1];
                const double cu = l * eps_lm(l + 1, m);
                s = make_double2(m * wl.y + cu * vu.x, -m * wl.x + cu * vu.y);
                t = make_double2(m * vl.y - cu * wu.x, -m * vl.x - cu * wu.y);
                if (l - 1 >= m) {
                    const double cl = (l + 1) * eps_lm(l, m);
                    const double2 vd = v[l - 1], wd = w[l - 1];
                    s.x -= cl * vd.x; s.y -= cl * vd.y;
                    t.x += cl * wd.x; t.y += cl * wd.y;
                }
                s.x *= inv; s.y *= inv;
                t.x *= inv; t.y *= inv;
            }
        }
        Sl[l] = s;
        Tl[l] = t;
    }
}

void launch_sphtor_to_vw(const double2 *S, const double2 *T, long spec_dist, double2 *VW, long vw_dist, int lmax,
                         int mmax, int mres, int ltr, int imlim, int nfields, hipStream_t s) {
    dim3 grid((lmax + 2 + 255) / 256, mmax + 1, nfields);
    hipLaunchKernelGGL(sphtor_to_vw_kernel, grid, dim3(256), 0, s, S, T, spec_dist, VW, vw_dist, lmax, mres, ltr,
                       imlim);
}

void launch_vw_to_sphtor(const double2 *VW, long vw_dist, double2 *S, double2 *T, long spec_dist, int lmax, int mmax,
                         int mres, int ltr, int imlim, int nfields, hipStream_t s) {
    dim3 grid((lmax + 1 + 255) / 256, mmax + 1, nfields);
    hipLaunchKernelGGL(vw_to_sphtor_kernel, grid, dim3(256), 0, s, VW, vw_dist, S, T, spec_dist, lmax, mres, ltr,
                       imlim);
}

}  // namespace shtns_amd
