// sht_legendre_i8.hip -- Legendre synthesis with the degree contraction done exactly on the
// int8 matrix cores (precision SHT_PREC_SLICED).
//
// Both operands of the contraction F[lat][col] = sum_i z_i(lat) b_i(col) are cut into five
// signed 8-bit digits with a power-of-two scale per latitude and per column that changes
// every 128 recurrence steps (a "super-block"):
//     z = 2^(eA-31) (d0 2^24 + d1 2^16 + d2 2^8 + d3 + d4 2^-8),   |z| < 2^(eA-1),
// and likewise b with eB.  The products of digit planes p, q with p + q <= 5 (19 of the 25)
// are summed exactly in int32 by v_mfma_i32_32x32x32_i8 (|sum| < 2^24 per super-block), one
// accumulator per diagonal p + q, and converted to FP64 with weight 2^(eA+eB-14-8(p+q)) at the
// end of each super-block.  Dropped products have weight <= 2^-62 relative to the leading
// one.  The recurrence, its seeds and tables, the FFT and all sums across super-blocks stay
// FP64; only the contraction changes.
//
// Digits: t = z 2^(31-eA) (|t| < 2^30) is rounded with the 1.5*2^52 trick; the low word of
// t + 1.5*2^52 is h = rint(t) in two's complement, and (h + 0x80808080) ^ 0x80808080 holds
// the four balanced base-256 digits of h (d0 in the top byte).  The fifth digit is
// rint(256 (t - h)) clamped to [-128, 127].
//
// Balancing: the recurrence values z_i alternate in size with the parity of i (z_i = v_i / d_i
// with d_i d_{i+1} ~ 4 / (e e'), ratios up to ~10^3 near l ~ m), and the pre-mixed
// coefficients alternate inversely, so one scale per super-block would waste bits on every
// other step.  The products are unchanged under z_i -> 2^(delta_i) z_i, b_i -> 2^(-delta_i) b_i
// with delta_i = ilogb(d_i): this kernel runs the recurrence of w_i = 2^(delta_i) z_i,
//     w_{i+1} = (a'_i x^2 + b'_i) w_i + g'_i w_{i-1},   a' = 2^(delta_{i+1}-delta_i) alpha_i, ...,
// g'_i = 2^(delta_{i+1}-delta_{i-1}); being exact powers of two, these reproduce 2^(delta) z
// bit for bit.  The coefficient planes are cut from 2^(-delta_i) b_i.
//
// Work decomposition: one wave = 32 latitude pairs x 16 fields (two 32-column tiles of
// 8 fields x (E.re, E.im, O.re, O.im)); lane l holds latitude l & 31 and, for each 32-step
// K block, the steps of parity l >> 5 (byte j <-> step 2j + (l >> 5); the coefficient planes
// use the same map).  Both lane halves run the latitude's recurrence; each keeps every other
// value.  Super-block scales of the recurrence values come from a table built once per grid
// (their magnitudes do not depend on the data); the recurrence state is rescaled by the
// exact power of two at each super-block boundary, so t is the recurrence value itself.
#include "sht_legendre.h"

namespace shtns_amd {

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x16 __attribute__((ext_vector_type(16)));

constexpr int kI8KB = 32;                  // steps per MFMA K block
constexpr int kI8SB = 128;                 // steps per super-block
constexpr int kI8KBperSB = kI8SB / kI8KB;
constexpr int kI8Lat = 32;                 // latitude pairs per wave
constexpr int kI8Waves = 4;                // waves per block
constexpr int kI8ND = 6;                   // diagonals p + q = 0..5
constexpr int kI8Planes = 5;

__device__ __forceinline__ int lo_word(double v) { return __double2loint(v); }

// balanced digits of t (|t| < 2^30): packed top four digits (d0 in byte 3) and the fifth
__device__ __forceinline__ void slice5(double t, unsigned &p4, int &d4) {
    const double M = 6755399441055744.0;   // 1.5 * 2^52
    const double tm = t + M;
    const int h = lo_word(tm);
    const double r = t - (tm - M);
    const int e = lo_word(fma(r, 256.0, M));
    d4 = min(max(e, -128), 127);
    p4 = ((unsigned)h + 0x80808080u) ^ 0x80808080u;
}

// bytes b of four packed words -> one word (byte k from word k)
__device__ __forceinline__ void transpose4(unsigned a, unsigned b, unsigned c, unsigned d, unsigned out[4]) {
    const unsigned tlo = __builtin_amdgcn_perm(b, a, 0x05010400u);   // a0 b0 a1 b1
    const unsigned thi = __builtin_amdgcn_perm(b, a, 0x07030602u);   // a2 b2 a3 b3
    const unsigned ulo = __builtin_amdgcn_perm(d, c, 0x05010400u);   // c0 d0 c1 d1
    const unsigned uhi = __builtin_amdgcn_perm(d, c, 0x07030602u);   // c2 d2 c3 d3
    out[0] = __builtin_amdgcn_perm(ulo, tlo, 0x05040100u);           // a0 b0 c0 d0
    out[1] = __builtin_amdgcn_perm(ulo, tlo, 0x07060302u);           // a1 b1 c1 d1
    out[2] = __builtin_amdgcn_perm(uhi, thi, 0x05040100u);
    out[3] = __builtin_amdgcn_perm(uhi, thi, 0x07060302u);
}
__device__ __forceinline__ unsigned pack_low_bytes(int a, int b, int c, int d) {
    const unsigned ab = __builtin_amdgcn_perm((unsigned)b, (unsigned)a, 0x0c0c0400u);   // a0 b0 0 0
    const unsigned cd = __builtin_amdgcn_perm((unsigned)d, (unsigned)c, 0x04000c0cu);   // 0 0 c0 d0
    return ab | cd;
}

// ---------------------------------------------------------------------------------------
// Row-exponent table: for every (order slot, latitude pair) and super-block, eA with
// max |z_i| < 2^(eA-1) over the super-block's steps i >= i0 (0 where none).  Same
// arithmetic as the synthesis recurrence.
// ---------------------------------------------------------------------------------------
// balanced seeds (w_{i0-1}, w_{i0}) of a lane from the FP64 seeds (z_{i0-1}, z_{i0})
__device__ __forceinline__ double2 balanced_seed(double2 sd, int i0, int nst, const int *__restrict__ dl) {
    const int i = min(i0, nst - 1);
    return make_double2(ldexp(sd.x, dl[max(i - 1, 0)]), ldexp(sd.y, dl[i]));
}

__global__ void i8_rowexp_kernel(const double *__restrict__ x2k, const int *__restrict__ step_ofs,
                                 const double4 *__restrict__ abg, const int *__restrict__ delta,
                                 const int *__restrict__ i0tab, const double2 *__restrict__ z0tab,
                                 const int *__restrict__ sb_ofs, int mmax, int nlat_2, short *__restrict__ eA) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int slot = blockIdx.y;
    if (k >= nlat_2 || slot > mmax) return;
    const int nst = step_ofs[slot + 1] - step_ofs[slot];
    const double4 *abm = abg + step_ofs[slot];
    const double x2 = x2k[k];
    const int i0 = i0tab[(size_t)slot * nlat_2 + k];
    const double2 sd = balanced_seed(z0tab[(size_t)slot * nlat_2 + k], i0, nst, delta + step_ofs[slot]);
    short *out = eA + (size_t)sb_ofs[slot] * nlat_2 + k;
    const int nsb = sb_ofs[slot + 1] - sb_ofs[slot];
    double za = sd.x, zb = sd.y;
    for (int sb = 0; sb < nsb; ++sb) {
        double mx = 0.0;
        const int i1 = min((sb + 1) * kI8SB, nst);
        for (int i = max(sb * kI8SB, i0); i < i1; ++i) {
            mx = fmax(mx, fabs(zb));
            const double4 c = abm[i];
            const double zn = fma(fma(c.x, x2, c.y), zb, c.z * za);
            za = zb;
            zb = zn;
        }
        int e = 0;
        if (mx > 0.0) {
            (void)frexp(mx, &e);   // mx < 2^e
            e += 1;
        }
        out[(size_t)sb * nlat_2] = (short)e;
    }
}

// ---------------------------------------------------------------------------------------
// Coefficient planes: qmix (FP64, 32 columns per 8-field group, pre-mixed as for the FP64
// kernel) -> five digit planes per 32-step K block in the MFMA B layout and a column
// exponent per super-block.  Block = (field group, order slot, super-block), 256 threads:
// thread t owns column t & 31 and steps (t >> 5) + 8 s.
//   planes[fg][kblock][p][lane][16 bytes],  lane = 32 (step & 1) + col,  byte = (step & 31) >> 1
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void i8_slice_coeff_kernel(const double *__restrict__ qmix, int nsteps_total,
                                                             const int *__restrict__ step_ofs,
                                                             const int *__restrict__ delta,
                                                             const int *__restrict__ sb_ofs, int sb_total,
                                                             int nfg_valid, unsigned char *__restrict__ planes,
                                                             int *__restrict__ eB) {
    __shared__ unsigned char st[kI8KBperSB][kI8Planes][64 * 16];
    __shared__ double cmax[8][32];
    const int fg = blockIdx.z, slot = blockIdx.y, sb = blockIdx.x;
    const int nsb = sb_ofs[slot + 1] - sb_ofs[slot];
    if (sb >= nsb) return;
    const int t = threadIdx.x, col = t & 31, r0 = t >> 5;
    const int rows = step_ofs[slot + 1] - step_ofs[slot];
    const double *q = qmix + ((size_t)fg * nsteps_total + step_ofs[slot]) * 32;
    double v[16];
    double mx = 0.0;
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        const int i = sb * kI8SB + r0 + 8 * s;
        v[s] = (i < rows && fg < nfg_valid) ? ldexp(q[(size_t)i * 32 + col], -delta[step_ofs[slot] + i]) : 0.0;
        mx = fmax(mx, fabs(v[s]));
    }
    cmax[r0][col] = mx;
    __syncthreads();
    mx = 0.0;
#pragma unroll
    for (int r = 0; r < 8; ++r) mx = fmax(mx, cmax[r][col]);
    int e = 0;
    if (mx > 0.0) {
        (void)frexp(mx, &e);
        e += 1;
    }
    if (r0 == 0) eB[((size_t)fg * sb_total + sb_ofs[slot] + sb) * 32 + col] = e;
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        const int i = r0 + 8 * s;          // step within the super-block
        unsigned p4;
        int d4;
        slice5(ldexp(v[s], 31 - e), p4, d4);
        const int kb = i >> 5, sr = i & 31;
        const int lane = 32 * (sr & 1) + col, byte = sr >> 1;
#pragma unroll
        for (int p = 0; p < 4; ++p) st[kb][p][lane * 16 + byte] = (unsigned char)(p4 >> (8 * (3 - p)));
        st[kb][4][lane * 16 + byte] = (unsigned char)d4;
    }
    __syncthreads();
    // 20 KB out, 16 bytes per store
    const size_t kb0 = (size_t)4 * (sb_ofs[slot] + sb);
    uint4 *dst = reinterpret_cast<uint4 *>(planes + ((size_t)fg * 4 * sb_total + kb0) * kI8Planes * 1024);
    const uint4 *src = reinterpret_cast<const uint4 *>(&st[0][0][0]);
    for (int w = t; w < kI8KBperSB * kI8Planes * 64; w += 256) dst[w] = src[w];
}

// ---------------------------------------------------------------------------------------
// Synthesis kernel.  Grid: x = order x tile group (tile groups of one order on one XCD),
// y = group of 8 NCT fields.  Block = 4 waves, each 32 latitude pairs x NCT column tiles.
// The K blocks of a wave run as one loop; the coefficient planes of the next K block are
// loaded while the recurrence of the current one runs.
// ---------------------------------------------------------------------------------------
template <int NCT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NCT == 1 ? 2 : 1, NCT == 1 ? 2 : 1))) void
legendre_synth_i8_kernel(
    const double *__restrict__ xk, const double *__restrict__ x2k, const double *__restrict__ wqk,
    const int *__restrict__ step_ofs, const double4 *__restrict__ abg, const int *__restrict__ delta,
    const int *__restrict__ i0tab, const double2 *__restrict__ z0tab, const int *__restrict__ sb_ofs, int sb_total,
    const short *__restrict__ eAtab, const unsigned char *__restrict__ planes, const int *__restrict__ eBtab,
    int lmax_spec, double2 *__restrict__ F, long four_dist, int mmax, int mres, int nlat, int nlat_2, int ltr,
    int imlim, int nfields, int lat_weight) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int ntiles = (nlat_2 + kI8Lat - 1) / kI8Lat;
    const int ntg = (ntiles + kI8Waves - 1) / kI8Waves;
    const int bx = xcd_cluster(blockIdx.x, gridDim.x, ntg);
    const int im = bx / ntg;
    const int tile = (bx - im * ntg) * kI8Waves + wave;
    if (tile >= ntiles) return;   // no block-level synchronization in this kernel
    const int fg0 = NCT * blockIdx.y;               // first 8-field group of this wave
    const int kr = tile * kI8Lat + (lane & 31);     // this lane's latitude pair
    const bool kvalid = kr < nlat_2;
    const int kk = kvalid ? kr : nlat_2 - 1;
    const int par = lane >> 5;

    double out[NCT][16];
#pragma unroll
    for (int c = 0; c < NCT; ++c)
#pragma unroll
        for (int j = 0; j < 16; ++j) out[c][j] = 0.0;

    const OrderInfo o = order_info(im, lmax_spec, mmax, mres, (im <= imlim) ? ltr : -1, 0);
    const int nsteps = (im <= imlim) ? o.nsteps : 0;
    if (nsteps > 0) {
        const int sofs = step_ofs[o.slot];
        const __attribute__((address_space(4))) double *abm =
            (const __attribute__((address_space(4))) double *)(abg + sofs);   // uniform: scalar loads
        const double x2 = x2k[kk];
        int i0 = kvalid ? i0tab[(size_t)o.slot * nlat_2 + kk] : nsteps;
        if (i0 > nsteps) i0 = nsteps;
        const double2 sd = balanced_seed(z0tab[(size_t)o.slot * nlat_2 + kk], i0, nsteps, delta + sofs);
        int wmin = i0, wmax = i0;
        for (int off = 32; off >= 1; off >>= 1) {
            wmin = min(wmin, __shfl_xor(wmin, off));
            wmax = max(wmax, __shfl_xor(wmax, off));
        }
        wmin = __builtin_amdgcn_readfirstlane(wmin);
        wmax = __builtin_amdgcn_readfirstlane(wmax);
        const int sbofs = sb_ofs[o.slot];
        const short *eArow = eAtab + (size_t)sbofs * nlat_2 + kk;
        const unsigned char *pl0 = planes + ((size_t)fg0 * 4 * sb_total + 4 * (size_t)sbofs) * kI8Planes * 1024 +
                                   (size_t)lane * 16;
        const size_t fg_stride = (size_t)4 * sb_total * kI8Planes * 1024;
        const int *eB0 = eBtab + ((size_t)fg0 * sb_total + sbofs) * 32 + (lane & 31);
        auto load_b = [&](int gb, i32x4 (&B)[NCT][kI8Planes]) {
#pragma unroll
            for (int c = 0; c < NCT; ++c)
#pragma unroll
                for (int q = 0; q < kI8Planes; ++q)
                    B[c][q] = *reinterpret_cast<const i32x4 *>(pl0 + c * fg_stride +
                                                               ((size_t)gb * kI8Planes + q) * 1024);
        };
        double za = 0.0, zb = 0.0;
        int e = 31, eprev = 31;            // state scale: values are z * 2^(31 - e)
        double seedscale = 1.0;
        const int gb_first = wmin / kI8KB, gb_end = (nsteps + kI8KB - 1) / kI8KB;
        i32x16 acc[NCT][kI8ND];
        i32x4 Bc[NCT][kI8Planes], Bn[NCT][kI8Planes];
        load_b(gb_first, Bc);
        for (int gb = gb_first; gb < gb_end; ++gb) {
            const int sb = gb / kI8KBperSB;
            if (gb == gb_first || gb % kI8KBperSB == 0) {   // a new super-block: scale, zero sums
                e = kvalid ? (int)eArow[(size_t)sb * nlat_2] : 0;
                const int sh = eprev - e;
                za = ldexp(za, sh);
                zb = ldexp(zb, sh);
                eprev = e;
                seedscale = ldexp(1.0, 31 - e);
#pragma unroll
                for (int c = 0; c < NCT; ++c)
#pragma unroll
                    for (int d = 0; d < kI8ND; ++d) acc[c][d] = (i32x16){0};
            }
            if (gb + 1 < gb_end) load_b(gb + 1, Bn);   // in flight during the recurrence
            const int ibase = gb * kI8KB;
            // recurrence over the block's 32 steps; this lane keeps the steps of its parity
            double keep[16];
            auto steps = [&](auto inject) {
#pragma unroll
                for (int s = 0; s < kI8KB; ++s) {
                    const int i = ibase + s;
                    if (decltype(inject)::value && i == i0) {
                        za = sd.x * seedscale;
                        zb = sd.y * seedscale;
                    }
                    if ((s & 1) == 0) keep[s >> 1] = zb;
                    else keep[s >> 1] = par ? zb : keep[s >> 1];
                    const int ic = 4 * min(i, nsteps - 1);
                    const double zn = fma(fma(abm[ic], x2, abm[ic + 1]), zb, abm[ic + 2] * za);
                    za = zb;
                    zb = zn;
                }
            };
            if (ibase <= wmax) steps(std::true_type{});   // some lane of the wave starts here
            else steps(std::false_type{});
            // digit planes of the kept values: A[p] = 16 bytes
            i32x4 A[kI8Planes];
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                unsigned p4[4];
                int d4[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) slice5(keep[4 * g + u], p4[u], d4[u]);
                unsigned tr[4];
                transpose4(p4[0], p4[1], p4[2], p4[3], tr);
#pragma unroll
                for (int p = 0; p < 4; ++p) A[p][g] = (int)tr[3 - p];   // plane 0 = top digits
                A[4][g] = (int)pack_low_bytes(d4[0], d4[1], d4[2], d4[3]);
            }
#pragma unroll
            for (int c = 0; c < NCT; ++c)
#pragma unroll
                for (int p = 0; p < kI8Planes; ++p)
#pragma unroll
                    for (int q = 0; q < kI8Planes; ++q)
                        if (p + q < kI8ND)
                            acc[c][p + q] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[p], Bc[c][q], acc[c][p + q], 0, 0, 0);
            if (gb % kI8KBperSB == kI8KBperSB - 1 || gb + 1 == gb_end) {
                // super-block result into FP64: D row (register j) = (j & 3) + 8 (j >> 2) + 4 (lane >> 5),
                // column = lane & 31
#pragma unroll
                for (int c = 0; c < NCT; ++c) {
                    const int eb = eB0[(size_t)c * sb_total * 32 + (size_t)sb * 32];
#pragma unroll
                    for (int j = 0; j < 16; ++j) {
                        const int er = __shfl(e, (j & 3) + 8 * (j >> 2) + 4 * par);
                        double t = (double)acc[c][kI8ND - 1][j];
#pragma unroll
                        for (int d = kI8ND - 2; d >= 0; --d) t = fma(t, 0x1p-8, (double)acc[c][d][j]);
                        out[c][j] += ldexp(t, er + eb - 14);
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < NCT; ++c)
#pragma unroll
                for (int q = 0; q < kI8Planes; ++q) Bc[c][q] = Bn[c][q];
        }
    }

    // Epilogue: a quad of lanes (columns 4f..4f+3) holds one field's E.re, E.im, O.re, O.im
    const int col = lane & 31;
    const int reim = col & 1;
    const bool o_lane = (col & 2) != 0;
    const bool zero_im = (im == 0) && reim;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int krow = tile * kI8Lat + (j & 3) + 8 * (j >> 2) + 4 * par;
        const int kc = min(krow, nlat_2 - 1);
        const double x = xk[kc];
        const double xs = o_lane ? -x : x;
        const double wz = zero_im ? 0.0 : (lat_weight ? wqk[kc] * ((2 * krow + 1 == nlat) ? 2.0 : 1.0) : 1.0);
        const int row = o_lane ? (nlat - 1 - krow) : krow;
#pragma unroll
        for (int c = 0; c < NCT; ++c) {
            const double ov = quad_perm<0xEE>(out[c][j]);   // lanes (2, 3, 2, 3)
            const double ev = quad_perm<0x44>(out[c][j]);   // lanes (0, 1, 0, 1)
            const double val = fma(xs, ov, ev) * wz;
            const int f = 8 * (fg0 + c) + (col >> 2);
            if (krow < nlat_2 && f < nfields) {
                double *Ff = reinterpret_cast<double *>(F + (size_t)f * four_dist + (size_t)im * nlat) + reim;
                Ff[2 * row] = val;
            }
        }
    }
}

// ---------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------
void launch_i8_rowexp(const DeviceTables *t, int mmax, int nlat_2, hipStream_t s) {
    dim3 grid((nlat_2 + 63) / 64, mmax + 1);
    hipLaunchKernelGGL(i8_rowexp_kernel, grid, dim3(64), 0, s, t->x2, t->step_ofs, t->abg, t->delta, t->i0, t->z0,
                       t->sb_ofs, mmax, nlat_2, t->eA);
}

size_t i8_planes_bytes(int nfields, int sb_total) {
    const int nfg = (nfields + 15) / 16 * 2;
    return (size_t)nfg * 4 * sb_total * kI8Planes * 1024;
}
size_t i8_eB_bytes(int nfields, int sb_total) {
    const int nfg = (nfields + 15) / 16 * 2;
    return (size_t)nfg * sb_total * 32 * sizeof(int);
}

void launch_legendre_synth_i8(const LegendreArgs &a, int spec_ltop, double2 *F, long four_dist, const double *qmix,
                              int nsteps_total, unsigned char *planes, int *eB, hipStream_t s) {
    const DeviceTables *t = a.t;
    const int nfg = (a.nfields + 15) / 16 * 2;   // 8-field groups (qmix holds them at 32 columns)
    const int mslots = a.mmax + 1;
    const int nfg_valid = (a.nfields + 7) / 8;   // groups the pre-mix wrote
    hipLaunchKernelGGL(i8_slice_coeff_kernel, dim3(t->sb_max, mslots, nfg), dim3(256), 0, s, qmix, nsteps_total,
                       t->step_ofs, t->delta, t->sb_ofs, t->sb_total, nfg_valid, planes, eB);
    const int ntiles = (a.nlat_2 + kI8Lat - 1) / kI8Lat;
    const int ntg = (ntiles + kI8Waves - 1) / kI8Waves;
    // one 8-field column tile per wave keeps the kernel within 256 registers (2 waves per SIMD);
    // two tiles per wave spill
    dim3 grid((a.mmax + 1) * ntg, nfg);
    hipLaunchKernelGGL(legendre_synth_i8_kernel<1>, grid, dim3(64 * kI8Waves), 0, s, t->x, t->x2, t->wq, t->step_ofs,
                       t->abg, t->delta, t->i0, t->z0, t->sb_ofs, t->sb_total, t->eA, planes, eB, spec_ltop, F, four_dist,
                       a.mmax, a.mres, a.nlat, a.nlat_2, a.ltr, a.imlim, a.nfields, a.lat_weight);
}

}  // namespace shtns_amd
