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
// Synthesis: F[f][im][k] = sum_l Q_lm P_lm(x_k)   (both hemispheres)
// Block = 4 waves; wave w handles latitude tile (blockIdx.z*4 + w) of 64 latitude pairs,
// for order im = blockIdx.x and fields 4*blockIdx.y .. +3.
// Per 4-step block: lanes compute z_i for their latitude (4 steps) -> LDS; MFMA
// D[16 lat][16 col] += Y[16 lat][4 step] * B[4 step][16 col] for the 4 row tiles.
// ---------------------------------------------------------------------------------------
constexpr int kSynthWaves = 4;

__global__ __launch_bounds__(256) void legendre_synth_kernel(
    const double *__restrict__ xk, const double *__restrict__ x2k, const double *__restrict__ stk,
    const double *__restrict__ wqk, const int *__restrict__ step_ofs, const double2 *__restrict__ ab,
    const double4 *__restrict__ pm, const int *__restrict__ i0tab, const double2 *__restrict__ z0tab,
    const double2 *__restrict__ Q, long spec_dist, int lmax_spec, double2 *__restrict__ F, long four_dist,
    int mmax, int mres, int nlat, int nlat_2, int ltr, int imlim, int nfields, int vector_mode,
    double mpos_scale, int lat_weight) {
    __shared__ double ytile[kSynthWaves][4][64];
    __shared__ double dout[kSynthWaves][64][17];

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int im = blockIdx.x;
    const int fg = blockIdx.y;
    const int tile = blockIdx.z * kSynthWaves + wave;
    const int k = tile * 64 + lane;              // latitude pair owned by this lane
    if (tile * 64 >= nlat_2) return;             // whole wave idle (no barriers below)
    const bool kvalid = (k < nlat_2);
    const int kk = kvalid ? k : (nlat_2 - 1);

    // column handled by this lane in the B operand / D columns
    const int col = lane & 15;
    const int fcol = fg * 4 + (col >> 2);
    const int comp = col & 3;
    const bool fvalid = (fcol < nfields);

    const double x = xk[kk];
    double4_t acc[4];
    for (int r = 0; r < 4; ++r) acc[r] = (double4_t){0.0, 0.0, 0.0, 0.0};

    if (im <= imlim) {
        const OrderInfo o = order_info(im, lmax_spec, mmax, mres, ltr, vector_mode);
        const int sofs = step_ofs[o.slot];
        const double2 *abm = ab + sofs;
        const double4 *pmm = pm + sofs;
        const double x2 = x2k[kk];
        // recurrence start for this lane
        int i0 = kvalid ? i0tab[(size_t)o.slot * nlat_2 + kk] : o.nsteps;
        double2 sd = z0tab[(size_t)o.slot * nlat_2 + kk];
        if (o.sinv) {
            const double inv = 1.0 / stk[kk];
            sd.x *= inv;
            sd.y *= inv;
        }
        if (i0 > o.nsteps) i0 = o.nsteps;
        // wave-uniform range of steps
        int ibeg = i0;
        for (int off = 32; off >= 1; off >>= 1) ibeg = min(ibeg, __shfl_xor(ibeg, off));
        int imax0 = i0;
        for (int off = 32; off >= 1; off >>= 1) imax0 = max(imax0, __shfl_xor(imax0, off));
        ibeg &= ~3;
        double za = 0.0, zb = 0.0;
        if (i0 == ibeg) { za = sd.x; zb = sd.y; }

        const double2 *Qf = Q + (size_t)(fvalid ? fcol : 0) * spec_dist + o.spec_base;
        const double qscale = (im == 0) ? 1.0 : mpos_scale;
        const int ltop = ltr;
        // B operand for step i (this lane: step i + lane/16, column col)
        auto load_b = [&](int i) -> double {
            const int ii = i + (lane >> 4);
            if (!fvalid || ii >= o.nsteps) return 0.0;
            const int le = o.mfun + 2 * ii;
            const double4 c = pmm[ii];
            double v;
            if (comp < 2) {
                const double2 q0 = Qf[le];
                double a = (comp == 0) ? q0.x : q0.y;
                double b = 0.0;
                if (le + 2 <= ltop) {
                    const double2 q2 = Qf[le + 2];
                    b = (comp == 0) ? q2.x : q2.y;
                }
                v = c.x * a + c.y * b;
            } else {
                double a = 0.0;
                if (le + 1 <= ltop) {
                    const double2 q1 = Qf[le + 1];
                    a = (comp == 2) ? q1.x : q1.y;
                }
                v = c.z * a;
            }
            if (o.real_only && (comp & 1)) v = 0.0;
            return v * qscale;
        };

        double bnext = load_b(ibeg);
        for (int i = ibeg; i < o.nsteps; i += 4) {
            const double bcur = bnext;
            // 4 recurrence steps for this lane's latitude
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int ii = i + s;
                if (ii < imax0) {
                    if (ii == i0) { za = sd.x; zb = sd.y; }
                }
                ytile[wave][s][lane] = (ii >= i0) ? zb : 0.0;
                const double2 c = abm[min(ii, o.nsteps - 1)];
                const double zn = fma(fma(c.x, x2, c.y), zb, za);
                za = zb;
                zb = zn;
            }
            bnext = load_b(i + 4);
            __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): LDS writes visible within the wave
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double a = ytile[wave][lane >> 4][r * 16 + (lane & 15)];
                acc[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bcur, acc[r], 0, 0, 0);
            }
            __builtin_amdgcn_wave_barrier();
        }
    }

    // gfx950 D layout of v_mfma_f64_16x16x4f64: register j of lane l holds
    // D[row = 4*j + l/16][col = l%16]  (row tile r covers latitudes 16r..16r+15).
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int j = 0; j < 4; ++j) dout[wave][r * 16 + 4 * j + (lane >> 4)][lane & 15] = acc[r][j];
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __builtin_amdgcn_wave_barrier();

    if (!kvalid) return;
    const double wmul = lat_weight ? wqk[k] : 1.0;
    const bool real_only = (im == 0);
    for (int f4 = 0; f4 < 4; ++f4) {
        const int f = fg * 4 + f4;
        if (f >= nfields) break;
        const double er = dout[wave][lane][4 * f4 + 0];
        const double ei = dout[wave][lane][4 * f4 + 1];
        const double orr = dout[wave][lane][4 * f4 + 2] * x;
        const double oi = dout[wave][lane][4 * f4 + 3] * x;
        double2 n = make_double2((er + orr) * wmul, (ei + oi) * wmul);
        double2 s = make_double2((er - orr) * wmul, (ei - oi) * wmul);
        if (real_only) { n.y = 0.0; s.y = 0.0; }
        double2 *Ff = F + (size_t)f * four_dist + (size_t)im * nlat;
        Ff[k] = n;
        Ff[nlat - 1 - k] = s;
    }
}

void launch_legendre_synth(const LegendreArgs &a, const double2 *Qlm, long spec_dist, int spec_ltop,
                           double2 *F, long four_dist, hipStream_t s) {
    const DeviceTables *t = a.t;
    const int ntiles = (a.nlat_2 + 63) / 64;
    dim3 grid(a.mmax + 1, (a.nfields + 3) / 4, (ntiles + kSynthWaves - 1) / kSynthWaves);
    hipLaunchKernelGGL(legendre_synth_kernel, grid, dim3(64 * kSynthWaves), 0, s, t->x, t->x2, t->st, t->wq,
                       t->step_ofs, t->ab, t->pm, t->i0, t->z0, Qlm, spec_dist, spec_ltop, F, four_dist, a.mmax,
                       a.mres, a.nlat, a.nlat_2, a.ltr, a.imlim, a.nfields, a.vector_mode, a.mpos_scale,
                       a.lat_weight);
}

// ---------------------------------------------------------------------------------------
// Analysis: Q_lm = sum_k w_k P_lm(x_k) F_m(x_k)
// Block = NW waves (all latitudes of one order, fields 4*blockIdx.y..+3); each wave owns
// NLR latitude tiles of 64.  Per 16-step block: lanes compute z for their latitudes ->
// LDS, MFMA D[16 step][16 col] += Y[16 step][4 lat] * E[4 lat][16 col], then the waves'
// partial D are summed through LDS and post-mixed into Q (even: ue_i acc_i + ue1_{i-1}
// acc_{i-1}; odd: uo_i acc_i).  Deterministic: fixed summation order, no atomics.
// ---------------------------------------------------------------------------------------
template <int NW, int NLR>
__global__ __launch_bounds__(64 * NW) void legendre_anal_kernel(
    const double *__restrict__ xk, const double *__restrict__ x2k, const double *__restrict__ stk,
    const double *__restrict__ wqk, const int *__restrict__ step_ofs, const double2 *__restrict__ ab,
    const double4 *__restrict__ pm, const int *__restrict__ i0tab, const double2 *__restrict__ z0tab,
    const double2 *__restrict__ F, long four_dist, double2 *__restrict__ Q, long spec_dist, int lmax_spec,
    int spec_ltop, int mmax, int mres, int nlat, int nlat_2, int ltr, int imlim, int nfields, int vector_mode,
    double mpos_scale, int lat_weight) {
    __shared__ double ytile[NW][16][65];
    __shared__ double red[NW][16][17];
    __shared__ double carry[16];
    __shared__ int s_ibeg[NW];

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int im = blockIdx.x;
    const int fg = blockIdx.y;
    const int col = lane & 15;
    const int fcol = fg * 4 + (col >> 2);
    const int comp = col & 3;
    const bool fvalid = (fcol < nfields);
    const int m = im * mres;

    const OrderInfo o = order_info(im, lmax_spec, mmax, mres, (im <= imlim) ? ltr : -1, vector_mode);
    const int nsteps = (im <= imlim) ? o.nsteps : 0;
    const int sofs = step_ofs[o.slot];
    const double2 *abm = ab + sofs;
    const double4 *pmm = pm + sofs;

    // per owned latitude tile: recurrence state
    double za[NLR], zb[NLR], x2v[NLR], sdx[NLR], sdy[NLR];
    int i0v[NLR];
    // E operand registers: Ereg[r][kg] for latitude tile r, K-group kg (4 latitudes)
    double Ereg[NLR][16];
    int my_ibeg = nsteps;
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
        // B operand (E/O values) for K-group kg: latitude tile_base + 4*kg + lane/16, column col
#pragma unroll
        for (int kg = 0; kg < 16; ++kg) {
            const int kl = (wave + NW * r) * 64 + 4 * kg + (lane >> 4);
            double v = 0.0;
            if (kl < nlat_2 && fvalid && im <= imlim) {
                const double2 *Ff = F + (size_t)fcol * four_dist + (size_t)im * nlat;
                const double2 fn = Ff[kl];
                const double2 fs = Ff[nlat - 1 - kl];
                const double w = lat_weight ? wqk[kl] : 1.0;
                if (comp == 0) v = (fn.x + fs.x) * w;
                else if (comp == 1) v = (fn.y + fs.y) * w;
                else if (comp == 2) v = (fn.x - fs.x) * w * xk[kl];
                else v = (fn.y - fs.y) * w * xk[kl];
                if (o.real_only && (comp & 1)) v = 0.0;
            }
            Ereg[r][kg] = v;
        }
    }
    // block-uniform starting step
    for (int off = 32; off >= 1; off >>= 1) my_ibeg = min(my_ibeg, __shfl_xor(my_ibeg, off));
    if (lane == 0) s_ibeg[wave] = my_ibeg;
    __syncthreads();
    int ibeg = nsteps;
    for (int w = 0; w < NW; ++w) ibeg = min(ibeg, s_ibeg[w]);
    ibeg &= ~15;
#pragma unroll
    for (int r = 0; r < NLR; ++r) {
        za[r] = 0.0;
        zb[r] = 0.0;
        if (i0v[r] == ibeg) { za[r] = sdx[r]; zb[r] = sdy[r]; }
    }
    if (threadIdx.x < 16) carry[threadIdx.x] = 0.0;

    const double qscale = (im == 0) ? 1.0 : mpos_scale;
    // output: zero the steps below ibeg (functions negligible there)
    // degrees written: o.mfun .. spec_ltop (zeros above ltr)
    const int nsteps_out = (spec_ltop >= o.mfun) ? (spec_ltop - o.mfun) / 2 + 1 : 0;

    for (int i = 0; i < nsteps; i += 16) {
        double4_t acc = (double4_t){0.0, 0.0, 0.0, 0.0};
        if (i + 16 > ibeg) {
#pragma unroll
            for (int r = 0; r < NLR; ++r) {
#pragma unroll
                for (int s = 0; s < 16; ++s) {
                    const int ii = i + s;
                    if (ii == i0v[r] && ii != ibeg) { za[r] = sdx[r]; zb[r] = sdy[r]; }
                    ytile[wave][s][lane] = (ii >= i0v[r] && ii < nsteps) ? zb[r] : 0.0;
                    const double2 c = abm[min(ii, nsteps - 1)];
                    const double zn = fma(fma(c.x, x2v[r], c.y), zb[r], za[r]);
                    za[r] = zb[r];
                    zb[r] = zn;
                }
                __builtin_amdgcn_s_waitcnt(0xc07f);
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int kg = 0; kg < 16; ++kg) {
                    const double a = ytile[wave][lane & 15][4 * kg + (lane >> 4)];
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Ereg[r][kg], acc, 0, 0, 0);
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
        // D layout (gfx950): acc[j] = D[step row 4*j + lane/16][col lane%16]
#pragma unroll
        for (int j = 0; j < 4; ++j) red[wave][4 * j + (lane >> 4)][lane & 15] = acc[j];
        __syncthreads();
        // threads 0..255 of the block reduce 16x16 values; then post-mix
        for (int e = threadIdx.x; e < 256; e += 64 * NW) {
            const int row = e >> 4, c = e & 15;
            double sum = 0.0;
#pragma unroll
            for (int w = 0; w < NW; ++w) sum += red[w][row][c];
            red[0][row][c] = sum;   // safe: each (row,c) read by exactly this thread above
        }
        __syncthreads();
        // post-mix and store: thread handles (step row, field, re/im)
        for (int e = threadIdx.x; e < 16 * 4 * 2; e += 64 * NW) {
            const int row = e >> 3, f4 = (e >> 1) & 3, ri = e & 1;
            const int f = fg * 4 + f4;
            const int ii = i + row;
            if (f < nfields && ii < nsteps_out) {
                const int le = o.mfun + 2 * ii;
                double qe = 0.0, qo = 0.0;
                if (ii < nsteps) {
                    const double4 c = pmm[ii];
                    const double ae = red[0][row][4 * f4 + ri];
                    const double aprev = (row > 0) ? red[0][row - 1][4 * f4 + ri] : carry[4 * f4 + ri];
                    const double cprev = (ii > 0) ? pmm[ii - 1].y : 0.0;
                    qe = c.x * ae + cprev * aprev;
                    qo = c.z * red[0][row][4 * f4 + 2 + ri];
                }
                // step nsteps (just past ltr) still receives the ue1 tail of step nsteps-1
                double2 *Qf = Q + (size_t)f * spec_dist + o.spec_base;
                if (le <= spec_ltop) {
                    double v = (le <= ltr) ? qe * qscale : 0.0;
                    if (o.real_only && ri) v = 0.0;
                    reinterpret_cast<double *>(&Qf[le])[ri] = v;
                }
                if (le + 1 <= spec_ltop) {
                    double v = (le + 1 <= ltr) ? qo * qscale : 0.0;
                    if (o.real_only && ri) v = 0.0;
                    reinterpret_cast<double *>(&Qf[le + 1])[ri] = v;
                }
            }
        }
        __syncthreads();
        if (threadIdx.x < 16) carry[threadIdx.x] = red[0][15][threadIdx.x];
        __syncthreads();
    }
    // orders/degrees with no step computed: zero-fill the remaining output degrees
    for (int ii = ((nsteps + 15) & ~15) + threadIdx.x; ii < nsteps_out; ii += 64 * NW) {
        const int le = o.mfun + 2 * ii;
        for (int f = fg * 4; f < min(nfields, fg * 4 + 4); ++f) {
            double2 *Qf = Q + (size_t)f * spec_dist + o.spec_base;
            if (le <= spec_ltop) Qf[le] = make_double2(0.0, 0.0);
            if (le + 1 <= spec_ltop) Qf[le + 1] = make_double2(0.0, 0.0);
        }
    }
    // the degree slot below mfun for vector m=0 (l=0) is zero
    if (vector_mode && im == 0 && threadIdx.x == 0) {
        for (int f = fg * 4; f < min(nfields, fg * 4 + 4); ++f) Q[(size_t)f * spec_dist + o.spec_base] = make_double2(0.0, 0.0);
    }
    (void)m;
}

template <int NW, int NLR>
static void launch_anal_t(const LegendreArgs &a, const double2 *F, long four_dist, double2 *Q, long spec_dist,
                          int spec_ltop, hipStream_t s) {
    const DeviceTables *t = a.t;
    dim3 grid(a.mmax + 1, (a.nfields + 3) / 4);
    hipLaunchKernelGGL((legendre_anal_kernel<NW, NLR>), grid, dim3(64 * NW), 0, s, t->x, t->x2, t->st, t->wq,
                       t->step_ofs, t->ab, t->pm, t->i0, t->z0, F, four_dist, Q, spec_dist,
                       spec_ltop == a.lmax + 1 ? a.lmax + 1 : a.lmax, spec_ltop, a.mmax, a.mres, a.nlat, a.nlat_2,
                       a.ltr, a.imlim, a.nfields, a.vector_mode, a.mpos_scale, a.lat_weight);
}

void launch_legendre_anal(const LegendreArgs &a, const double2 *F, long four_dist, double2 *Q, long spec_dist,
                          int spec_ltop, hipStream_t s) {
    const int ntiles = (a.nlat_2 + 63) / 64;
    // waves per block: a divisor of the tile count, at most 8
    int nw = 1;
    for (int d = 8; d >= 1; --d)
        if (ntiles % d == 0) { nw = d; break; }
    int nlr = ntiles / nw;
    if (nlr > 6) { nw = 8; nlr = (ntiles + 7) / 8; }
#define ANAL_CASE(W, R) if (nw == W && nlr <= R) { launch_anal_t<W, R>(a, F, four_dist, Q, spec_dist, spec_ltop, s); return; }
    ANAL_CASE(1, 1) ANAL_CASE(1, 2) ANAL_CASE(1, 3) ANAL_CASE(1, 5) ANAL_CASE(1, 7)
    ANAL_CASE(2, 1) ANAL_CASE(2, 3) ANAL_CASE(2, 5)
    ANAL_CASE(3, 1) ANAL_CASE(3, 3) ANAL_CASE(3, 5)
    ANAL_CASE(4, 1) ANAL_CASE(4, 2) ANAL_CASE(4, 3) ANAL_CASE(4, 5)
    ANAL_CASE(5, 1) ANAL_CASE(5, 3)
    ANAL_CASE(6, 1) ANAL_CASE(6, 2) ANAL_CASE(6, 3) ANAL_CASE(6, 5)
    ANAL_CASE(7, 1) ANAL_CASE(7, 3)
    ANAL_CASE(8, 1) ANAL_CASE(8, 2) ANAL_CASE(8, 3) ANAL_CASE(8, 4) ANAL_CASE(8, 6) ANAL_CASE(8, 8)
    ANAL_CASE(8, 12) ANAL_CASE(8, 16)
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
        const int e = (int)(((long)n1 * k * twm) % N);
        double2 w = roots[e];
        w.y *= sgn;
        a[k * q] = cmul(y[k], w);
    }
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

__global__ __launch_bounds__(256) void fft_synth_kernel(const double2 *__restrict__ F, long four_dist,
                                                        double *__restrict__ V, long spat_dist,
                                                        const double2 *__restrict__ roots, const int *__restrict__ pos,
                                                        FftPlan plan, int nlat, int imlim, long sp, long stt, int T) {
    extern __shared__ double2 buf[];
    const int M = plan.M;
    const int N = 2 * M;
    const int lat0 = blockIdx.x * T;
    const int f = blockIdx.y;
    const double2 *Ff = F + (size_t)f * four_dist;
    double *Vf = V + (size_t)f * spat_dist;
    const int Tv = min(T, nlat - lat0);
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
    fft_dif_inplace(buf, T, M, plan, roots, +1.0);
    for (int e = threadIdx.x; e < T * M; e += blockDim.x) {
        const int t = e % T, n = e / T;
        if (t < Tv) {
            const double2 z = buf[(size_t)t * M + pos[n]];
            const int lat = lat0 + t;
            Vf[(size_t)(2 * n) * sp + (size_t)lat * stt] = z.x;
            Vf[(size_t)(2 * n + 1) * sp + (size_t)lat * stt] = z.y;
        }
    }
    (void)N;
}

__global__ __launch_bounds__(256) void fft_anal_kernel(const double *__restrict__ V, long spat_dist,
                                                       double2 *__restrict__ F, long four_dist,
                                                       const double2 *__restrict__ roots, const int *__restrict__ pos,
                                                       FftPlan plan, int nlat, int mmax, int imlim, long sp, long stt,
                                                       int T) {
    extern __shared__ double2 buf[];
    const int M = plan.M;
    const int lat0 = blockIdx.x * T;
    const int f = blockIdx.y;
    const double *Vf = V + (size_t)f * spat_dist;
    double2 *Ff = F + (size_t)f * four_dist;
    const int Tv = min(T, nlat - lat0);
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
    fft_dif_inplace(buf, T, M, plan, roots, -1.0);
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
    // LDS budget of 64 KB per block
    int T = (int)(65536 / (16L * M));
    if (T < 1) T = 1;
    if (T > 4) T = 4;
    return T;
}

void launch_fft_synth(const DeviceTables *t, int nlat, int nphi, int mmax, int imlim, int nfields,
                      const double2 *F, long four_dist, double *V, long spat_dist, long sp, long st, hipStream_t s) {
    const int M = nphi / 2;
    const int T = fft_lat_per_block(M);
    dim3 grid((nlat + T - 1) / T, nfields);
    const size_t lds = (size_t)T * M * sizeof(double2);
    if (lds > 65536) (void)hipFuncSetAttribute((const void *)fft_synth_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(fft_synth_kernel, grid, dim3(256), lds, s, F, four_dist, V, spat_dist, t->roots, t->fft_pos,
                       t->plan, nlat, imlim, sp, st, T);
    (void)mmax;
}

void launch_fft_anal(const DeviceTables *t, int nlat, int nphi, int mmax, int imlim, int nfields, const double *V,
                     long spat_dist, double2 *F, long four_dist, long sp, long st, hipStream_t s) {
    const int M = nphi / 2;
    const int T = fft_lat_per_block(M);
    dim3 grid((nlat + T - 1) / T, nfields);
    const size_t lds = (size_t)T * M * sizeof(double2);
    if (lds > 65536) (void)hipFuncSetAttribute((const void *)fft_anal_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(fft_anal_kernel, grid, dim3(256), lds, s, V, spat_dist, F, four_dist, t->roots, t->fft_pos,
                       t->plan, nlat, mmax, imlim, sp, st, T);
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
