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

#include <cstdio>
#include <cstdlib>

#include "sht_internal.h"

namespace shtns_amd {

typedef double double4_t __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------------------
// Seed tables: for each order slot and latitude pair, the first recurrence step whose
// functions exceed `thr` in magnitude, and (z_{i0-1}, z_{i0}) there.  The recurrence is
// run with a power-of-two exponent carried beside the values (exact scaling), so orders
// whose sin^m underflows double precision start correctly (SHTns does the same with its
// SHT_SCALE_FACTOR, kernel_SH_to_spat.gen.c:278-348).
// ---------------------------------------------------------------------------------------
__global__ void seed_kernel(const double *__restrict__ xk, const double *__restrict__ stk,
                            const int *__restrict__ step_ofs, const double2 *__restrict__ ab,
                            const double4 *__restrict__ pm, int lmax, int mmax, int mres, int nlat_2,
                            double thr, int *__restrict__ i0_out, double2 *__restrict__ z0_out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int slot = blockIdx.y;
    if (k >= nlat_2) return;
    const int m = (slot <= mmax) ? slot * mres : 1;
    const int nst = step_ofs[slot + 1] - step_ofs[slot];
    const double2 *abm = ab + step_ofs[slot];
    const double4 *pmm = pm + step_ofs[slot];
    const double x = xk[k];
    const double x2 = x * x;
    const double s = stk[k];
    // sin^m as mant * 2^E
    int E = 0, es;
    double mant = 1.0;
    {
        double b = frexp(s, &es);
        int eb = es;
        int n = m;
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

void launch_seed_tables(DeviceTables *t, int lmax, int mmax, int mres, int nlat_2, double thr, hipStream_t s) {
    dim3 grid((nlat_2 + 63) / 64, mmax + 2);
    hipLaunchKernelGGL(seed_kernel, grid, dim3(64), 0, s, t->x, t->st, t->step_ofs, t->ab, t->pm, lmax, mmax,
                       mres, nlat_2, thr, t->i0, t->z0);
}

// ---------------------------------------------------------------------------------------
// Helpers shared by the Legendre kernels
// ---------------------------------------------------------------------------------------
// Read-only table viewed through the constant address space: loads with wave-uniform
// indices become scalar (SMEM) loads and stay off the vector-memory counters.
struct UniformTable2 {
    const __attribute__((address_space(4))) double *p;
    __device__ explicit UniformTable2(const double2 *q) : p((const __attribute__((address_space(4))) double *)q) {}
    __device__ double2 operator[](int i) const { return make_double2(p[2 * i], p[2 * i + 1]); }
};
struct OrderInfo {
    int slot;       // table slot (im, or mmax+1 for the m=1 functions of vector m=0)
    int mfun;       // order of the generated functions
    int nsteps;     // recurrence steps to evaluate (degrees mfun .. ltr)
    int spec_base;  // index of degree 0 of this order in the spectral layout (may be virtual)
    bool sinv;      // divide functions by sin(theta) (vector V,W spectra, m>0)
    bool real_only; // m = 0: imaginary parts ignored / zero
};

__device__ inline OrderInfo order_info(int im, int lmax_spec, int mmax, int mres, int ltr, int vector_mode) {
    OrderInfo o;
    const int m = im * mres;
    o.real_only = (im == 0);
    if (vector_mode && im == 0) {
        o.slot = mmax + 1;
        o.mfun = 1;
        o.sinv = false;
    } else {
        o.slot = im;
        o.mfun = m;
        o.sinv = (vector_mode != 0);
    }
    o.nsteps = (ltr >= o.mfun) ? (ltr - o.mfun) / 2 + 1 : 0;
    o.spec_base = (int)(((long)im * (2 * lmax_spec + 2 - (im + 1) * mres)) / 2);
    return o;
}

// ---------------------------------------------------------------------------------------
// Synthesis, step 1 (pre-mix): the coefficients are rearranged once per transform into the
// layout of the MFMA B operand, so the Legendre kernel reads them with plain coalesced
// loads and needs no block barrier:
//   qmix[g][slot steps][NCOL],  col = 16*cg + 4*field + (E.re, E.im, O.re, O.im),
//   E_i = ue_i Q_{m+2i} + ue1_i Q_{m+2i+2},  O_i = uo_i Q_{m+2i+1}    (g = group of 4*NCG fields)
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void premix_kernel(const int *__restrict__ step_ofs, const double4 *__restrict__ pm,
                                                     const double2 *__restrict__ Q, long spec_dist, int lmax_spec,
                                                     double *__restrict__ qmix, int nsteps_total, int ncol, int mmax,
                                                     int mres, int ltr, int imlim, int nfields, int vector_mode,
                                                     double mpos_scale) {
    const int im = blockIdx.x;
    const int g = blockIdx.y;
    if (im > imlim) return;
    const OrderInfo o = order_info(im, lmax_spec, mmax, mres, ltr, vector_mode);
    const int nsteps = o.nsteps;
    const double4 *pmm = pm + step_ofs[o.slot];
    double *dst = qmix + ((size_t)g * nsteps_total + step_ofs[o.slot]) * ncol;
    const double qscale = (im == 0) ? 1.0 : mpos_scale;
    const int fpg = ncol / 4;   // fields per group
    // rows nsteps .. end of the slot are written as zeros: the Legendre kernel reads them
    // unconditionally (clamped to the slot) when it runs past the last step
    const int slot_rows = step_ofs[o.slot + 1] - step_ofs[o.slot];
    for (int e = blockIdx.z * blockDim.x + threadIdx.x; e < slot_rows * ncol; e += gridDim.z * blockDim.x) {
        const int ii = e / ncol, col = e - ii * ncol;
        const int f = g * fpg + (col >> 2);
        const int comp = col & 3;
        double v = 0.0;
        if (ii < nsteps && f < nfields && !(o.real_only && (comp & 1))) {
            const double2 *Qf = Q + (size_t)f * spec_dist + o.spec_base;
            const int le = o.mfun + 2 * ii;
            const double4 cf = pmm[ii];
            if (comp < 2) {
                const double2 q0 = Qf[le];
                double a = (comp == 0) ? q0.x : q0.y, b = 0.0;
                if (le + 2 <= ltr) {
                    const double2 q2 = Qf[le + 2];
                    b = (comp == 0) ? q2.x : q2.y;
                }
                v = cf.x * a + cf.y * b;
            } else if (le + 1 <= ltr) {
                const double2 q1 = Qf[le + 1];
                v = cf.z * ((comp == 2) ? q1.x : q1.y);
            }
            v *= qscale;
        }
        dst[e] = v;
    }
}

// ---------------------------------------------------------------------------------------
// Synthesis, step 2: F[f][im][k] = sum_l Q_lm P_lm(x_k)   (both hemispheres)
// One wave = one order im, 4*NCG fields, one latitude tile of 64 pairs (lane = pair); waves
// are independent (no block barrier).  Software pipeline per 4-step block:
//   A operands (z of 16 latitudes x 4 steps) read back from the wave's LDS tile,
//   4 x NCG MFMAs D[16 lat][16 col] += A * B,
//   then, while the matrix cores work, the recurrence of the next 4 steps is advanced on
//   the vector ALUs into the other LDS tile and the next B operands are loaded.
// The accumulators go straight to F: lanes holding E and O of one field are lane^2
// partners (north = E + x O, south = E - x O).
// ---------------------------------------------------------------------------------------
constexpr int kSynthWaves = 4;

template <int NCG>
__global__ __launch_bounds__(256) void legendre_synth_kernel(
    const double *__restrict__ xk, const double *__restrict__ x2k, const double *__restrict__ stk,
    const double *__restrict__ wqk, const int *__restrict__ step_ofs, const double2 *__restrict__ ab,
    const int *__restrict__ i0tab, const double2 *__restrict__ z0tab, const double *__restrict__ qmix,
    int nsteps_total, int lmax_spec, double2 *__restrict__ F, long four_dist, int mmax, int mres, int nlat,
    int nlat_2, int ltr, int imlim, int nfields, int vector_mode, int lat_weight, int tile0, int ntl) {
    constexpr int NCOL = 16 * NCG;
    __shared__ double ytile[kSynthWaves][2][4][64];

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    // heavy orders first: block x -> (order, tile group); all tile groups of an order adjacent
    const int ntg = (ntl + kSynthWaves - 1) / kSynthWaves;   // tile groups of this launch
    const int im = blockIdx.x / ntg;
    const int tile = tile0 + (blockIdx.x - im * ntg) * kSynthWaves + wave;
    const int g = blockIdx.y;
    const int fbase = g * 4 * NCG;
    if (tile >= tile0 + ntl || tile * 64 >= nlat_2) return;   // no block-level synchronization in this kernel
    const int k = tile * 64 + lane;
    const bool kvalid = (k < nlat_2);
    const int kk = kvalid ? k : (nlat_2 - 1);

    double4_t acc[4][NCG];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < NCG; ++c) acc[r][c] = (double4_t){0.0, 0.0, 0.0, 0.0};

    const OrderInfo o = order_info(im, lmax_spec, mmax, mres, (im <= imlim) ? ltr : -1, vector_mode);
    const int nsteps = (im <= imlim) ? o.nsteps : 0;
    if (nsteps > 0) {
        const int sofs = step_ofs[o.slot];
        const int slot_rows = step_ofs[o.slot + 1] - sofs;   // > nsteps: the last rows of qmix are zero
        const UniformTable2 abm(ab + sofs);   // uniform indices -> scalar loads
        const double *qb = qmix + ((size_t)g * nsteps_total + sofs) * NCOL;
        const double x2 = x2k[kk];
        int i0 = kvalid ? i0tab[(size_t)o.slot * nlat_2 + kk] : nsteps;
        double2 sd = z0tab[(size_t)o.slot * nlat_2 + kk];
        if (o.sinv) {
            const double inv = 1.0 / stk[kk];
            sd.x *= inv;
            sd.y *= inv;
        }
        if (i0 > nsteps) i0 = nsteps;
        int wmin = i0, wmax = i0;
        for (int off = 32; off >= 1; off >>= 1) {
            wmin = min(wmin, __shfl_xor(wmin, off));
            wmax = max(wmax, __shfl_xor(wmax, off));
        }
        // wave-uniform: keeps loop control and coefficient loads on the scalar unit
        wmin = __builtin_amdgcn_readfirstlane(wmin);
        wmax = __builtin_amdgcn_readfirstlane(wmax);
        const int ibeg = wmin & ~3;
        double za = 0.0, zb = 0.0;
        if (i0 == ibeg) { za = sd.x; zb = sd.y; }

        // advance 4 recurrence steps starting at step i, writing z_i..z_{i+3} to tile buffer
        auto advance = [&](int i, int buf) {
            double2 cf[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) cf[s] = abm[min(i + s, nsteps - 1)];
            if (i <= wmax) {   // some lane of this wave starts in this block
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    if (i + s == i0) { za = sd.x; zb = sd.y; }
                    ytile[wave][buf][s][lane] = zb;
                    const double zn = fma(fma(cf[s].x, x2, cf[s].y), zb, za);
                    za = zb;
                    zb = zn;
                }
            } else {
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    ytile[wave][buf][s][lane] = zb;
                    const double zn = fma(fma(cf[s].x, x2, cf[s].y), zb, za);
                    za = zb;
                    zb = zn;
                }
            }
        };
        auto load_b = [&](int i, double *b) {
            const int ii = min(i + (lane >> 4), slot_rows - 1);   // rows >= nsteps hold zeros
#pragma unroll
            for (int c = 0; c < NCG; ++c) b[c] = qb[(size_t)ii * NCOL + c * 16 + (lane & 15)];
        };

        double bcur[NCG], bnxt[NCG];
        advance(ibeg, 0);
        load_b(ibeg, bcur);
        int buf = 0;
        for (int i = ibeg; i < nsteps; i += 4) {
            load_b(i + 4, bnxt);
            __builtin_amdgcn_s_waitcnt(0xc07f);   // LDS tile of this block written
            __builtin_amdgcn_wave_barrier();
            double a[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) a[r] = ytile[wave][buf][lane >> 4][r * 16 + (lane & 15)];
#pragma unroll
            for (int c = 0; c < NCG; ++c)
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r], bcur[c], acc[r][c], 0, 0, 0);
            if (i + 4 < nsteps) advance(i + 4, buf ^ 1);
            buf ^= 1;
#pragma unroll
            for (int c = 0; c < NCG; ++c) bcur[c] = bnxt[c];
        }
    }

    // gfx950 D layout of v_mfma_f64_16x16x4f64: register j of lane l holds
    // D[row = 4*j + l/16][col = l%16]; row tile r covers latitudes tile*64 + 16r + row.
    const int col = lane & 15;
    const int comp = col & 3;
    const bool real_only = (im == 0);
#pragma unroll
    for (int c = 0; c < NCG; ++c) {
        const int f = fbase + c * 4 + (col >> 2);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const double mine = acc[r][c][j];
                const double other = __shfl_xor(mine, 2);   // E <-> O partner
                const int kr = tile * 64 + r * 16 + 4 * j + (lane >> 4);
                if (comp < 2 && kr < nlat_2 && f < nfields) {
                    const double x = xk[kr];
                    const double wmul = lat_weight ? wqk[kr] : 1.0;
                    double n = (mine + x * other) * wmul;
                    double sth = (mine - x * other) * wmul;
                    if (real_only && comp == 1) { n = 0.0; sth = 0.0; }
                    double *Ff = reinterpret_cast<double *>(F + (size_t)f * four_dist + (size_t)im * nlat);
                    Ff[2 * kr + comp] = n;
                    Ff[2 * (nlat - 1 - kr) + comp] = sth;
                }
            }
        }
    }
}

int synth_ncg(int nfields) { return nfields > 8 ? 4 : (nfields > 4 ? 2 : 1); }

size_t synth_qmix_doubles(int nfields, int nsteps_total) {
    const int ncg = synth_ncg(nfields);
    const int ng = (nfields + 4 * ncg - 1) / (4 * ncg);
    return (size_t)ng * nsteps_total * 16 * ncg;
}

void launch_premix(const LegendreArgs &a, const double2 *Qlm, long spec_dist, int spec_ltop, double *qmix,
                   int nsteps_total, hipStream_t s) {
    const DeviceTables *t = a.t;
    const int ncg = synth_ncg(a.nfields);
    const int ng = (a.nfields + 4 * ncg - 1) / (4 * ncg);
    dim3 grid(a.mmax + 1, ng, 4);
    hipLaunchKernelGGL(premix_kernel, grid, dim3(256), 0, s, t->step_ofs, t->pm, Qlm, spec_dist, spec_ltop, qmix,
                       nsteps_total, 16 * ncg, a.mmax, a.mres, a.ltr, a.imlim, a.nfields, a.vector_mode, a.mpos_scale);
}

void launch_legendre_synth_tiles(const LegendreArgs &a, int spec_ltop, double2 *F, long four_dist, const double *qmix,
                                 int nsteps_total, int tile0, int ntl, hipStream_t s) {
    const DeviceTables *t = a.t;
    const int ntg = (ntl + kSynthWaves - 1) / kSynthWaves;
    const int ncg = synth_ncg(a.nfields);
    const int ng = (a.nfields + 4 * ncg - 1) / (4 * ncg);
    dim3 grid((a.mmax + 1) * ntg, ng);
#define SYNTH_LAUNCH(NC)                                                                                         \
    hipLaunchKernelGGL((legendre_synth_kernel<NC>), grid, dim3(64 * kSynthWaves), 0, s, t->x, t->x2, t->st, t->wq, \
                       t->step_ofs, t->ab, t->i0, t->z0, qmix, nsteps_total, spec_ltop, F, four_dist, a.mmax,    \
                       a.mres, a.nlat, a.nlat_2, a.ltr, a.imlim, a.nfields, a.vector_mode, a.lat_weight, tile0, ntl)
    if (ncg == 4) SYNTH_LAUNCH(4);
    else if (ncg == 2) SYNTH_LAUNCH(2);
    else SYNTH_LAUNCH(1);
#undef SYNTH_LAUNCH
}

void launch_legendre_synth(const LegendreArgs &a, const double2 *Qlm, long spec_dist, int spec_ltop,
                           double2 *F, long four_dist, double *qmix, int nsteps_total, hipStream_t s) {
    launch_premix(a, Qlm, spec_dist, spec_ltop, qmix, nsteps_total, s);
    launch_legendre_synth_tiles(a, spec_ltop, F, four_dist, qmix, nsteps_total, 0, (a.nlat_2 + 63) / 64, s);
}

// ---------------------------------------------------------------------------------------
// Analysis: Q_lm = sum_k w_k P_lm(x_k) F_m(x_k)
// Block = NW waves for one order im and 4*NCG fields; wave w owns latitude tiles
// w + NW*r (r < NLR) of 64 pairs, lane = latitude pair.  The weighted E/O values of the
// block's fields for the owned latitudes are held in registers as MFMA B operands.
// Per 16-step block: each lane advances its recurrences into a wave-private LDS tile, the
// tile is read back as the A operand (16 steps x 4 latitudes) and
// D[16 step][16 col] += A * B runs over all owned K-groups; the waves' partial D go to a
// double-buffered LDS area, one barrier, then every thread sums one (step, column) over
// the waves and applies the post-mix (even: ue_i acc_i + ue1_{i-1} acc_{i-1}; odd:
// uo_i acc_i) and stores Q.  Fixed summation order: results are deterministic.
// ---------------------------------------------------------------------------------------
template <int NW, int NLR, int NCG>
__global__ __launch_bounds__(64 * NW) void legendre_anal_kernel(
    const double *__restrict__ xk, const double *__restrict__ x2k, const double *__restrict__ stk,
    const double *__restrict__ wqk, const int *__restrict__ step_ofs, const double2 *__restrict__ ab,
    const double4 *__restrict__ pm, const int *__restrict__ i0tab, const double2 *__restrict__ z0tab,
    const double2 *__restrict__ F, long four_dist, double2 *__restrict__ Q, long spec_dist, int lmax_spec,
    int spec_ltop, int mmax, int mres, int nlat, int nlat_2, int ltr, int imlim, int nfields, int vector_mode,
    double mpos_scale, int lat_weight) {
    constexpr int NCOL = 16 * NCG;
    __shared__ double ytile[NW][16][65];
    __shared__ double red[NW][16][NCOL + 1];   // single-buffered: two barriers per block, 2 blocks/CU fit
    __shared__ double carry[2][NCOL];
    __shared__ int s_ibeg[NW];

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int im = blockIdx.x;
    const int fbase = blockIdx.y * 4 * NCG;

    const OrderInfo o = order_info(im, lmax_spec, mmax, mres, (im <= imlim) ? ltr : -1, vector_mode);
    const int nsteps = (im <= imlim) ? o.nsteps : 0;
    const int sofs = step_ofs[o.slot];
    const UniformTable2 abm(ab + sofs);   // uniform indices -> scalar loads
    const double4 *pmm = pm + sofs;

    double za[NLR], zb[NLR], x2v[NLR], sdx[NLR], sdy[NLR];
    int i0v[NLR];
    double Ereg[NLR][NCG][16];
    int my_ibeg = nsteps, my_imax = 0;
#pragma unroll
    for (int r = 0; r < NLR; ++r) {
        const int k = (wave + NW * r) * 64 + lane;
        const bool kv = (k < nlat_2);
        const int kk = kv ? k : 0;
        x2v[r] = x2k[kk];
        int i0 = kv ? i0tab[(size_t)o.slot * nlat_2 + kk] : nsteps;
        if (i0 > nsteps) i0 = nsteps;
        double2 sd = z0tab[(size_t)o.slot * nlat_2 + kk];
        if (o.sinv) {
            const double inv = 1.0 / stk[kk];
            sd.x *= inv;
            sd.y *= inv;
        }
        i0v[r] = i0;
        sdx[r] = sd.x;
        sdy[r] = sd.y;
        my_ibeg = min(my_ibeg, i0);
        my_imax = max(my_imax, kv ? i0 : 0);
        // B operand for K-group kg: latitude tile_base + 4*kg + lane/16, column lane%16 of group c
#pragma unroll
        for (int c = 0; c < NCG; ++c) {
            const int col = lane & 15;
            const int f = fbase + 4 * c + (col >> 2);
            const int comp = col & 3;
#pragma unroll
            for (int kg = 0; kg < 16; ++kg) {
                const int kl = (wave + NW * r) * 64 + 4 * kg + (lane >> 4);
                double v = 0.0;
                if (kl < nlat_2 && f < nfields && nsteps > 0 && !(o.real_only && (comp & 1))) {
                    const double2 *Ff = F + (size_t)f * four_dist + (size_t)im * nlat;
                    const double2 fn = Ff[kl];
                    const double2 fs = Ff[nlat - 1 - kl];
                    const double w = lat_weight ? wqk[kl] : 1.0;
                    if (comp == 0) v = (fn.x + fs.x) * w;
                    else if (comp == 1) v = (fn.y + fs.y) * w;
                    else if (comp == 2) v = (fn.x - fs.x) * w * xk[kl];
                    else v = (fn.y - fs.y) * w * xk[kl];
                }
                Ereg[r][c][kg] = v;
            }
        }
    }
    for (int off = 32; off >= 1; off >>= 1) {
        my_ibeg = min(my_ibeg, __shfl_xor(my_ibeg, off));
        my_imax = max(my_imax, __shfl_xor(my_imax, off));
    }
    my_imax = __builtin_amdgcn_readfirstlane(my_imax);
    if (lane == 0) s_ibeg[wave] = my_ibeg;
    if (threadIdx.x < 2 * NCOL) (&carry[0][0])[threadIdx.x] = 0.0;
    __syncthreads();
    int ibeg = nsteps;
    for (int w = 0; w < NW; ++w) ibeg = min(ibeg, s_ibeg[w]);
    ibeg = __builtin_amdgcn_readfirstlane(ibeg) & ~15;   // block-uniform, scalar from here on
#pragma unroll
    for (int r = 0; r < NLR; ++r) {
        za[r] = 0.0;
        zb[r] = 0.0;
        if (i0v[r] == ibeg) { za[r] = sdx[r]; zb[r] = sdy[r]; }
    }

    const double qscale = (im == 0) ? 1.0 : mpos_scale;
    const int nsteps_out = (spec_ltop >= o.mfun) ? (spec_ltop - o.mfun) / 2 + 1 : 0;
    double2 *Qbase = Q + o.spec_base;

    // post-mix work split: thread handles elements e = threadIdx.x + j*64*NW (j < NE) of the
    // 16 x NCOL block; their coefficient rows are prefetched before the MFMA phase.
    constexpr int NE = (16 * NCOL + 64 * NW - 1) / (64 * NW);
    int b = 0;
    for (int i = ibeg; i < nsteps; i += 16, ++b) {
        double pcx[NE], pcz[NE], pcp[NE];
#pragma unroll
        for (int j = 0; j < NE; ++j) {
            const int e = threadIdx.x + j * 64 * NW;
            const int row = min(e / NCOL, 15);
            const int ii = min(i + row, nsteps - 1);
            const double4 cf4 = pmm[ii];
            pcx[j] = cf4.x;
            pcz[j] = cf4.z;
            pcp[j] = (ii > 0) ? pmm[ii - 1].y : 0.0;
        }
        double2 cf[16];
#pragma unroll
        for (int s = 0; s < 16; ++s) cf[s] = abm[min(i + s, nsteps - 1)];   // uniform: scalar loads
        double4_t acc[NCG], acc2[NCG];   // two chains per column group (even / odd K-groups)
#pragma unroll
        for (int c = 0; c < NCG; ++c) acc[c] = acc2[c] = (double4_t){0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int r = 0; r < NLR; ++r) {
            if (i <= my_imax) {   // some lane of this wave starts its recurrence here
#pragma unroll
                for (int s = 0; s < 16; ++s) {
                    const int ii = i + s;
                    if (ii == i0v[r]) { za[r] = sdx[r]; zb[r] = sdy[r]; }
                    ytile[wave][s][lane] = zb[r];
                    const double zn = fma(fma(cf[s].x, x2v[r], cf[s].y), zb[r], za[r]);
                    za[r] = zb[r];
                    zb[r] = zn;
                }
            } else {
#pragma unroll
                for (int s = 0; s < 16; ++s) {
                    ytile[wave][s][lane] = zb[r];
                    const double zn = fma(fma(cf[s].x, x2v[r], cf[s].y), zb[r], za[r]);
                    za[r] = zb[r];
                    zb[r] = zn;
                }
            }
            __builtin_amdgcn_s_waitcnt(0xc07f);
            __builtin_amdgcn_wave_barrier();
            double a[16];
#pragma unroll
            for (int kg = 0; kg < 16; ++kg) a[kg] = ytile[wave][lane & 15][4 * kg + (lane >> 4)];
            __builtin_amdgcn_s_waitcnt(0xc07f);   // all A operands in registers before the tile is rewritten
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int kg = 0; kg < 16; kg += 2)
#pragma unroll
                for (int c = 0; c < NCG; ++c) {
                    acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kg], Ereg[r][c][kg], acc[c], 0, 0, 0);
                    acc2[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kg + 1], Ereg[r][c][kg + 1], acc2[c], 0, 0, 0);
                }
        }
#pragma unroll
        for (int c = 0; c < NCG; ++c) acc[c] += acc2[c];
        // D layout (gfx950): acc[c][j] = D[step row 4*j + lane/16][col lane%16]
        const int pb = b & 1;
        __syncthreads();   // previous block's post-mix has finished reading red
#pragma unroll
        for (int c = 0; c < NCG; ++c)
#pragma unroll
            for (int j = 0; j < 4; ++j) red[wave][4 * j + (lane >> 4)][16 * c + (lane & 15)] = acc[c][j];
        __syncthreads();
        // sum over waves + post-mix
#pragma unroll
        for (int j = 0; j < NE; ++j) {
            const int e = threadIdx.x + j * 64 * NW;
            if (e >= 16 * NCOL) break;
            const int row = e / NCOL, col = e - row * NCOL;
            const int comp = col & 3;
            const int f = fbase + (col >> 2);
            double sum = 0.0, sprev = 0.0;
#pragma unroll
            for (int w = 0; w < NW; ++w) sum += red[w][row][col];
            if (row == 15 && comp < 2) carry[pb][col] = sum;
            const int ii = i + row;
            if (f >= nfields || ii >= nsteps_out) continue;
            const int le = o.mfun + 2 * ii;
            double v = 0.0;
            if (ii < nsteps) {
                if (comp < 2) {
                    if (row > 0) {
#pragma unroll
                        for (int w = 0; w < NW; ++w) sprev += red[w][row - 1][col];
                    } else {
                        sprev = carry[pb ^ 1][col];
                    }
                    v = pcx[j] * sum + pcp[j] * sprev;
                } else {
                    v = pcz[j] * sum;
                }
            }
            const int ld = (comp < 2) ? le : le + 1;
            if (ld <= spec_ltop) {
                double out = (ld <= ltr) ? v * qscale : 0.0;
                if (o.real_only && (comp & 1)) out = 0.0;
                reinterpret_cast<double *>(&Qbase[(size_t)f * spec_dist + ld])[comp & 1] = out;
            }
        }
    }
    // degrees not reached by the step loop: before ibeg (negligible functions) and after nsteps
    for (int ii = threadIdx.x; ii < nsteps_out; ii += 64 * NW) {
        if (ii >= ibeg && ii < ibeg + ((nsteps - ibeg + 15) & ~15)) continue;
        const int le = o.mfun + 2 * ii;
        for (int f = fbase; f < min(nfields, fbase + 4 * NCG); ++f) {
            double2 *Qf = Qbase + (size_t)f * spec_dist;
            if (le <= spec_ltop) Qf[le] = make_double2(0.0, 0.0);
            if (le + 1 <= spec_ltop) Qf[le + 1] = make_double2(0.0, 0.0);
        }
    }
    if (vector_mode && im == 0 && threadIdx.x == 0) {   // degree 0 of the vector m=0 spectra
        for (int f = fbase; f < min(nfields, fbase + 4 * NCG); ++f) Qbase[(size_t)f * spec_dist] = make_double2(0.0, 0.0);
    }
}

template <int NW, int NLR, int NCG>
static void launch_anal_t(const LegendreArgs &a, const double2 *F, long four_dist, double2 *Q, long spec_dist,
                          int spec_ltop, hipStream_t s) {
    const DeviceTables *t = a.t;
    dim3 grid(a.mmax + 1, (a.nfields + 4 * NCG - 1) / (4 * NCG));
    hipLaunchKernelGGL((legendre_anal_kernel<NW, NLR, NCG>), grid, dim3(64 * NW), 0, s, t->x, t->x2, t->st, t->wq,
                       t->step_ofs, t->ab, t->pm, t->i0, t->z0, F, four_dist, Q, spec_dist,
                       spec_ltop == a.lmax + 1 ? a.lmax + 1 : a.lmax, spec_ltop, a.mmax, a.mres, a.nlat, a.nlat_2,
                       a.ltr, a.imlim, a.nfields, a.vector_mode, a.mpos_scale, a.lat_weight);
}

void launch_legendre_anal(const LegendreArgs &a, const double2 *F, long four_dist, double2 *Q, long spec_dist,
                          int spec_ltop, hipStream_t s) {
    const int ntiles = (a.nlat_2 + 63) / 64;
    // waves per block: one latitude tile per wave up to 12 waves (measured fastest: lowest
    // register count, 3 waves per SIMD); beyond, a divisor of the tile count, at most 8
    int nw = 1;
    if (ntiles <= 12) nw = ntiles;
    else
        for (int d = 8; d >= 1; --d)
            if (ntiles % d == 0) { nw = d; break; }
    static const int nw_env = getenv("SHTNS_AMD_ANAL_NW") ? atoi(getenv("SHTNS_AMD_ANAL_NW")) : 0;
    if (nw_env > 0 && nw_env <= 16 && ntiles % nw_env == 0) nw = nw_env;
    int nlr = ntiles / nw;
    if (nlr > 6) { nw = 8; nlr = (ntiles + 7) / 8; }
    static const int ncg_env = getenv("SHTNS_AMD_ANAL_NCG") ? atoi(getenv("SHTNS_AMD_ANAL_NCG")) : 0;
    bool wide = (a.nfields > 4) && (nlr <= 3);   // 8 fields per block when registers allow
    if (ncg_env == 1) wide = false;
    if (ncg_env == 2 && nlr <= 3) wide = (a.nfields > 4);
#define ANAL_CASE(W, R)                                                                              \
    if (nw == W && nlr <= R) {                                                                       \
        if (wide) launch_anal_t<W, R, 2>(a, F, four_dist, Q, spec_dist, spec_ltop, s);              \
        else launch_anal_t<W, R, 1>(a, F, four_dist, Q, spec_dist, spec_ltop, s);                   \
        return;                                                                                      \
    }
    ANAL_CASE(1, 1) ANAL_CASE(1, 2) ANAL_CASE(1, 3) ANAL_CASE(1, 6)
    ANAL_CASE(2, 1) ANAL_CASE(2, 3) ANAL_CASE(2, 6)
    ANAL_CASE(3, 1) ANAL_CASE(3, 2) ANAL_CASE(3, 4) ANAL_CASE(3, 6)
    ANAL_CASE(4, 1) ANAL_CASE(4, 2) ANAL_CASE(4, 3) ANAL_CASE(4, 6)
    ANAL_CASE(5, 1) ANAL_CASE(5, 3) ANAL_CASE(5, 6)
    ANAL_CASE(6, 1) ANAL_CASE(6, 2) ANAL_CASE(6, 3) ANAL_CASE(6, 6)
    ANAL_CASE(7, 1) ANAL_CASE(7, 3) ANAL_CASE(7, 6)
    ANAL_CASE(8, 1) ANAL_CASE(8, 2) ANAL_CASE(8, 3) ANAL_CASE(8, 4) ANAL_CASE(8, 6)
    ANAL_CASE(9, 1) ANAL_CASE(10, 1) ANAL_CASE(11, 1) ANAL_CASE(12, 1)
#undef ANAL_CASE
    fprintf(stderr, "shtns_amd: no analysis kernel for nlat_2=%d\n", a.nlat_2);
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

template <class K>
static void set_lds_limit(K kern, size_t lds) {
    if (lds > 65536) (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

void launch_fft_synth(const DeviceTables *t, int nlat, int nphi, int mmax, int imlim, int nfields,
                      const double2 *F, long four_dist, double *V, long spat_dist, long sp, long st, hipStream_t s,
                      int lat_base, int nlat_chunk) {
    const int M = nphi / 2;
    if (nlat_chunk < 0) { lat_base = 0; nlat_chunk = nlat; }
    const int lat_end = lat_base + nlat_chunk;
    static const bool generic = getenv("SHTNS_AMD_FFT_GENERIC") != nullptr;
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

void launch_fft_anal(const DeviceTables *t, int nlat, int nphi, int mmax, int imlim, int nfields, const double *V,
                     long spat_dist, double2 *F, long four_dist, long sp, long st, hipStream_t s, int lat_base,
                     int nlat_chunk) {
    const int M = nphi / 2;
    if (nlat_chunk < 0) { lat_base = 0; nlat_chunk = nlat; }
    const int lat_end = lat_base + nlat_chunk;
    static const bool generic = getenv("SHTNS_AMD_FFT_GENERIC") != nullptr;
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
