// sht_legendre_i8.hip -- Legendre synthesis with the degree contraction done exactly on the
// int8 matrix cores (precision SHT_PREC_SLICED).
//
// Both operands of the contraction F[lat][col] = sum_i z_i(lat) b_i(col) are cut into five
// signed 8-bit digits with a power-of-two scale per latitude and per column that changes
// every 128 recurrence steps (a "super-block"):
//     z = 2^(eA-39) (d0 2^32 + d1 2^24 + d2 2^16 + d3 2^8 + d4),   |z| < 2^(eA-1),
// and likewise b with eB.  The products of digit planes p, q with p + q <= 5 (19 of the 25)
// are summed exactly in int32 by v_mfma_i32_32x32x32_i8 (|sum| < 2^24 per super-block), one
// accumulator per diagonal p + q, and converted to FP64 with weight 2^(eA+eB-14-8(p+q)) at the
// end of each super-block.  Dropped products have weight <= 2^-62 relative to the leading
// one.  The recurrence, its seeds and tables, the FFT and all sums across super-blocks stay
// FP64; only the contraction changes.
//
// Digits: t = z 2^(39-eA) (|t| < 2^38) plus the constant M = 1.5*2^52 + 0x8080808080 is
// rounded to an integer by one FP64 add; the low 40 bits of the sum's bit pattern are then
// rint(t) + 0x8080808080, whose bytes k = 0..4 are the balanced base-256 digits of rint(t)
// plus 128 (each byte XOR 0x80 is the signed digit).
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
// Work decomposition: one wave = 32 latitude pairs x 8 fields (one 32-column tile of
// 8 fields x (E.re, E.im, O.re, O.im)); lane l holds latitude l & 31 and, for each 32-step
// K block, the steps of parity l >> 5 (byte j <-> step 2j + (l >> 5); the coefficient planes
// use the same map).  Both lane halves run the latitude's recurrence; each keeps every other
// value.  Two waves share a SIMD, so one wave's MFMAs run under the other's VALU work.
// Super-block scales of the recurrence values come from a table built once per grid (their
// magnitudes do not depend on the data); the recurrence state is rescaled by the exact power
// of two at each super-block boundary, so t is the recurrence value itself.
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

// t (|t| < 2^38) -> rint(t) + 0x8080808080 in the low 40 bits of (hi:lo)
__device__ __forceinline__ void slice40(double t, unsigned &lo, unsigned &hi) {
    constexpr double M = 6755399441055744.0 + (double)0x8080808080ull;   // 1.5 * 2^52 + bias
    const double tm = t + M;
    lo = (unsigned)__double2loint(tm);
    hi = (unsigned)__double2hiint(tm);
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
__device__ __forceinline__ unsigned pack_low_bytes(unsigned a, unsigned b, unsigned c, unsigned d) {
    const unsigned ab = __builtin_amdgcn_perm(b, a, 0x0c0c0400u);   // a0 b0 0 0
    const unsigned cd = __builtin_amdgcn_perm(d, c, 0x04000c0cu);   // 0 0 c0 d0
    return ab | cd;
}

// five signed digit planes (plane 0 = most significant) of four values, one byte each
__device__ __forceinline__ void planes4(const unsigned lo[4], const unsigned hi[4], unsigned pl[kI8Planes]) {
    unsigned tr[4];
    transpose4(lo[0], lo[1], lo[2], lo[3], tr);
    pl[0] = pack_low_bytes(hi[0], hi[1], hi[2], hi[3]) ^ 0x80808080u;
#pragma unroll
    for (int p = 1; p < kI8Planes; ++p) pl[p] = tr[4 - p] ^ 0x80808080u;
}

// ---------------------------------------------------------------------------------------
// Row-exponent table: for every (order slot, latitude pair) and super-block, eA with
// max |w_i| < 2^(eA-1) over the super-block's steps i >= i0 (0 where none).  The same
// arithmetic as the synthesis kernel: frames change at every 32-step block start.
// ---------------------------------------------------------------------------------------
// balanced seeds (w_{i0-1}, w_{i0}) of a lane, in the frame of the block holding step i0,
// from the FP64 seeds (z_{i0-1}, z_{i0}); db = the order's block exponents
__device__ __forceinline__ double2 balanced_seed(double2 sd, int i0, int nst, const int2 *__restrict__ db) {
    const int i = min(i0, nst - 1);
    const int2 d = db[i / kI8KB];
    return make_double2(ldexp(sd.x, (i & 1) ? d.x : d.y), ldexp(sd.y, (i & 1) ? d.y : d.x));
}

__global__ void i8_rowexp_kernel(const double *__restrict__ x2k, const int *__restrict__ step_ofs,
                                 const double2 *__restrict__ abk, const int2 *__restrict__ dblk,
                                 const int *__restrict__ i0tab, const double2 *__restrict__ z0tab,
                                 const int *__restrict__ sb_ofs, int mmax, int nlat_2, short *__restrict__ eA) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int slot = blockIdx.y;
    if (k >= nlat_2 || slot > mmax) return;
    const int nst = step_ofs[slot + 1] - step_ofs[slot];
    const double2 *abm = abk + (size_t)sb_ofs[slot] * kI8SB;
    const int2 *db = dblk + (size_t)sb_ofs[slot] * kI8KBperSB;
    const double x2 = x2k[k];
    const int i0 = i0tab[(size_t)slot * nlat_2 + k];
    const double2 sd = balanced_seed(z0tab[(size_t)slot * nlat_2 + k], i0, nst, db);
    short *out = eA + (size_t)sb_ofs[slot] * nlat_2 + k;
    const int nsb = sb_ofs[slot + 1] - sb_ofs[slot];
    double za = sd.x, zb = sd.y;
    for (int sb = 0; sb < nsb; ++sb) {
        double mx = 0.0;
        const int i1 = min((sb + 1) * kI8SB, nst);
        for (int i = max(sb * kI8SB, i0); i < i1; ++i) {
            if (i > i0 && i % kI8KB == 0) {   // into the frame of block i / 32
                const int2 dn = db[i / kI8KB], dp = db[i / kI8KB - 1];
                za = ldexp(za, dn.y - dp.y);
                zb = ldexp(zb, dn.x - dp.x);
            }
            mx = fmax(mx, fabs(zb));
            const double2 c = abm[i];
            const double zn = fma(fma(c.x, x2, c.y), zb, za);
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
// exponent per super-block.  Block = (super-block, order slot, field group), 256 threads:
// thread t owns column t & 31 and steps (t >> 5) + 8 s of the super-block.
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
        unsigned lo, hi;
        slice40(ldexp(v[s], 39 - e), lo, hi);
        const int kb = i >> 5, sr = i & 31;
        const int lane = 32 * (sr & 1) + col, byte = sr >> 1;
        st[kb][0][lane * 16 + byte] = (unsigned char)((hi & 0xffu) ^ 0x80u);
#pragma unroll
        for (int p = 1; p < kI8Planes; ++p)
            st[kb][p][lane * 16 + byte] = (unsigned char)(((lo >> (8 * (4 - p))) & 0xffu) ^ 0x80u);
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
// y = 8-field group.  Block = 4 waves of 32 latitude pairs.  K blocks are software-pipelined:
// the 19 MFMAs of block gb are issued between the steps of block gb + 1 (five after each
// group of eight steps) and the coefficient planes of block gb + 1 are in flight meanwhile.
// The block first copies the order's recurrence coefficients (a table padded with zeros to
// whole super-blocks per order, row sb_ofs[slot] * 128 + i) to LDS; every step reads them
// there (one broadcast address per wave).
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void legendre_synth_i8_kernel(
    const double *__restrict__ xk, const double *__restrict__ x2k, const double *__restrict__ wqk,
    const int *__restrict__ step_ofs, const double2 *__restrict__ abk, const int2 *__restrict__ dblk,
    const int *__restrict__ i0tab, const double2 *__restrict__ z0tab, const int *__restrict__ sb_ofs, int sb_total,
    const short *__restrict__ eAtab, const unsigned char *__restrict__ planes, const int *__restrict__ eBtab,
    int lmax_spec, double2 *__restrict__ F, long four_dist, int mmax, int mres, int nlat, int nlat_2, int ltr,
    int imlim, int nfields, int lat_weight) {
    // LDS: FP64 sums over super-blocks (one column per lane), the order's recurrence
    // coefficients (a', b') per step and its block exponents
    extern __shared__ double lds[];
    double(*sums)[16][64] = reinterpret_cast<double(*)[16][64]>(lds);
    double2 *cfl = reinterpret_cast<double2 *>(lds + kI8Waves * 16 * 64);
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int ntiles = (nlat_2 + kI8Lat - 1) / kI8Lat;
    const int ntg = (ntiles + kI8Waves - 1) / kI8Waves;
    const int bx = xcd_cluster(blockIdx.x, gridDim.x, ntg);
    const int im = bx / ntg;
    const int tile = (bx - im * ntg) * kI8Waves + wave;
    const OrderInfo o = order_info(im, lmax_spec, mmax, mres, (im <= imlim) ? ltr : -1, 0);
    const int nsteps = (im <= imlim) ? o.nsteps : 0;
    const int gb_end = (nsteps + kI8KB - 1) / kI8KB;
    int2 *dbl = reinterpret_cast<int2 *>(cfl + gb_end * kI8KB);
    if (nsteps > 0) {
        const int sbofs = sb_ofs[o.slot];
        const double2 *src = abk + (size_t)sbofs * kI8SB;   // zero past the order's end
        for (int i = threadIdx.x; i < gb_end * kI8KB; i += blockDim.x) cfl[i] = src[i];
        for (int i = threadIdx.x; i < gb_end; i += blockDim.x) dbl[i] = dblk[(size_t)sbofs * kI8KBperSB + i];
    }
    __syncthreads();
    if (tile >= ntiles) return;   // the only block-level synchronization is above
    const int fg = blockIdx.y;
    const int kr = tile * kI8Lat + (lane & 31);   // this lane's latitude pair
    const bool kvalid = kr < nlat_2;
    const int kk = kvalid ? kr : nlat_2 - 1;
    const int half = lane >> 5;
    double *outl = &sums[wave][0][lane];
#pragma unroll
    for (int j = 0; j < 16; ++j) outl[64 * j] = 0.0;

    if (nsteps > 0) {
        const double x2 = x2k[kk];
        int i0 = kvalid ? i0tab[(size_t)o.slot * nlat_2 + kk] : nsteps;
        if (i0 > nsteps) i0 = nsteps;
        const int sbofs = sb_ofs[o.slot];
        const double2 sd = balanced_seed(z0tab[(size_t)o.slot * nlat_2 + kk], i0, nsteps,
                                         dblk + (size_t)sbofs * kI8KBperSB);
        int wmin = i0, wmax = i0;
        for (int off = 32; off >= 1; off >>= 1) {
            wmin = min(wmin, __shfl_xor(wmin, off));
            wmax = max(wmax, __shfl_xor(wmax, off));
        }
        wmin = __builtin_amdgcn_readfirstlane(wmin);
        wmax = __builtin_amdgcn_readfirstlane(wmax);
        const short *eArow = eAtab + (size_t)sbofs * nlat_2 + kk;
        const unsigned char *pl0 = planes + ((size_t)fg * 4 * sb_total + 4 * (size_t)sbofs) * kI8Planes * 1024 +
                                   (size_t)lane * 16;
        const int *eB0 = eBtab + ((size_t)fg * sb_total + sbofs) * 32 + (lane & 31);
        const int gb_first = wmin / kI8KB, gb_end = (nsteps + kI8KB - 1) / kI8KB;
        if (gb_first < gb_end) {
            double za = 0.0, zb = 0.0;
            int e = 39;                    // state scale: values are w * 2^(39 - e)
            const int nsb = (gb_end + kI8KBperSB - 1) / kI8KBperSB;
            int e_sb = -1;                 // super-block of e; e_next = scale of the following one
            int e_next = kvalid ? (int)eArow[(size_t)(gb_first / kI8KBperSB) * nlat_2] : 39;
            i32x16 acc[kI8ND];
            i32x4 A[kI8Planes], Bc[kI8Planes];
            auto load_b = [&](int gb, i32x4 (&B)[kI8Planes]) {
#pragma unroll
                for (int q = 0; q < kI8Planes; ++q)
                    B[q] = *reinterpret_cast<const i32x4 *>(pl0 + ((size_t)gb * kI8Planes + q) * 1024);
            };
            // MFMA k of a K block (k = 0..18): coefficient plane q, value plane p, p + q <= 5;
            // consecutive k use different accumulators
            auto mfma_k = [&](auto kc) {
                constexpr int k = decltype(kc)::value;
                constexpr int q = k < 5 ? 0 : k < 10 ? 1 : k < 14 ? 2 : k < 17 ? 3 : 4;
                constexpr int p = k - (q == 0 ? 0 : q == 1 ? 5 : q == 2 ? 10 : q == 3 ? 14 : 17);
                acc[p + q] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[p], Bc[q], acc[p + q], 0, 0, 0);
            };
            constexpr int kMfma = 19;
            auto mfma_range = [&](auto k0c) {
                constexpr int k0 = decltype(k0c)::value;
                if constexpr (k0 + 0 < kMfma) mfma_k(std::integral_constant<int, k0 + 0>{});
                if constexpr (k0 + 1 < kMfma) mfma_k(std::integral_constant<int, k0 + 1>{});
                if constexpr (k0 + 2 < kMfma) mfma_k(std::integral_constant<int, k0 + 2>{});
                if constexpr (k0 + 3 < kMfma) mfma_k(std::integral_constant<int, k0 + 3>{});
                if constexpr (k0 + 4 < kMfma) mfma_k(std::integral_constant<int, k0 + 4>{});
            };
            // recurrence values of block gb -> A (and the super-block rescale).  Lane half h
            // keeps the values of steps 2j + h (byte j).  With WITH_MFMA the MFMAs of the previous
            // block (operands A, Bc) are issued in between, five after each group of eight steps.
            auto gen = [&](int gb, auto inject, auto with_mfma) {
                int enew = e;
                if (gb / kI8KBperSB != e_sb) {   // a new super-block: its scale was loaded ahead
                    e_sb = gb / kI8KBperSB;
                    enew = e_next;
                    e_next = kvalid ? (int)eArow[(size_t)min(e_sb + 1, nsb - 1) * nlat_2] : e_next;
                }
                const int sh = e - enew;
                const int2 dn = dbl[gb], dp = dbl[max(gb - 1, 0)];   // into the frame of block gb
                za = ldexp(za, sh + dn.y - dp.y);
                zb = ldexp(zb, sh + dn.x - dp.x);
                e = enew;
                const double seedscale = ldexp(1.0, 39 - e);
                const int ibase = gb * kI8KB;
                const double2 *cb = cfl + ibase;
                unsigned An[kI8Planes][4];
                auto group = [&](auto gc) {
                    constexpr int g = decltype(gc)::value;
                    unsigned lo[4], hi[4];
                    double v = 0.0;
                    double2 cg[8];   // the group's coefficients, requested together
#pragma unroll
                    for (int u = 0; u < 8; ++u) cg[u] = cb[8 * g + u];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int s = 8 * g + u;
                        if (decltype(inject)::value && ibase + s == i0) {
                            za = sd.x * seedscale;
                            zb = sd.y * seedscale;
                        }
                        if (u & 1) {
                            v = half ? zb : v;
                            slice40(v, lo[u >> 1], hi[u >> 1]);
                        } else {
                            v = zb;
                        }
                        const double2 c = cg[u];
                        const double zn = fma(fma(c.x, x2, c.y), zb, za);
                        za = zb;
                        zb = zn;
                    }
                    unsigned pl[kI8Planes];
                    planes4(lo, hi, pl);
#pragma unroll
                    for (int p = 0; p < kI8Planes; ++p) An[p][g] = pl[p];
                    if constexpr (decltype(with_mfma)::value) mfma_range(std::integral_constant<int, 5 * g>{});
                    __builtin_amdgcn_sched_barrier(0);   // bound the live range of hoisted loads
                };
                group(std::integral_constant<int, 0>{});
                group(std::integral_constant<int, 1>{});
                group(std::integral_constant<int, 2>{});
                group(std::integral_constant<int, 3>{});
#pragma unroll
                for (int p = 0; p < kI8Planes; ++p)
#pragma unroll
                    for (int w = 0; w < 4; ++w) A[p][w] = (int)An[p][w];
            };
            if (gb_first * kI8KB <= wmax) gen(gb_first, std::true_type{}, std::false_type{});
            else gen(gb_first, std::false_type{}, std::false_type{});
            load_b(gb_first, Bc);
            int gb = gb_first;
            for (int sb = gb_first / kI8KBperSB; sb * kI8KBperSB < gb_end; ++sb) {
                const int emf = e;   // scale of this super-block's values
#pragma unroll
                for (int d = 0; d < kI8ND; ++d) acc[d] = (i32x16){0};
                const int gend = min((sb + 1) * kI8KBperSB, gb_end);
                for (; gb < gend; ++gb) {
                    // the other wave on the SIMD runs its VALU work while these MFMAs execute
                    const bool more = gb + 1 < gb_end;
                    mfma_range(std::integral_constant<int, 0>{});
                    mfma_range(std::integral_constant<int, 5>{});
                    mfma_range(std::integral_constant<int, 10>{});
                    mfma_range(std::integral_constant<int, 15>{});
                    if (more) {
                        load_b(gb + 1, Bc);   // in flight during the next block's recurrence
                        if ((gb + 1) * kI8KB <= wmax) gen(gb + 1, std::true_type{}, std::false_type{});
                        else gen(gb + 1, std::false_type{}, std::false_type{});
                    }
                }
                // super-block result into FP64: D row (register j) = (j & 3) + 8 (j >> 2) + 4 half,
                // column = lane & 31
                const int eb = eB0[(size_t)sb * 32];
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    const int er = __shfl(emf, (j & 3) + 8 * (j >> 2) + 4 * half);
                    double t = (double)acc[kI8ND - 1][j];
#pragma unroll
                    for (int d = kI8ND - 2; d >= 0; --d) t = fma(t, 0x1p-8, (double)acc[d][j]);
                    outl[64 * j] += ldexp(t, er + eb - 14);
                }
            }
        }
    }

    // Epilogue: a quad of lanes (columns 4f..4f+3) holds one field's E.re, E.im, O.re, O.im
    const int col = lane & 31;
    const int reim = col & 1;
    const bool o_lane = (col & 2) != 0;
    const bool zero_im = (im == 0) && reim;
    const int f = 8 * fg + (col >> 2);
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int krow = tile * kI8Lat + (j & 3) + 8 * (j >> 2) + 4 * half;
        const int kc = min(krow, nlat_2 - 1);
        const double x = xk[kc];
        const double xs = o_lane ? -x : x;
        const double wz = zero_im ? 0.0 : (lat_weight ? wqk[kc] * ((2 * krow + 1 == nlat) ? 2.0 : 1.0) : 1.0);
        const int row = o_lane ? (nlat - 1 - krow) : krow;
        const double sum = outl[64 * j];
        const double ov = quad_perm<0xEE>(sum);   // lanes (2, 3, 2, 3)
        const double ev = quad_perm<0x44>(sum);   // lanes (0, 1, 0, 1)
        const double val = fma(xs, ov, ev) * wz;
        if (krow < nlat_2 && f < nfields) {
            double *Ff = reinterpret_cast<double *>(F + (size_t)f * four_dist + (size_t)im * nlat) + reim;
            Ff[2 * row] = val;
        }
    }
}

// ---------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------
void launch_i8_rowexp(const DeviceTables *t, int mmax, int nlat_2, hipStream_t s) {
    dim3 grid((nlat_2 + 63) / 64, mmax + 1);
    hipLaunchKernelGGL(i8_rowexp_kernel, grid, dim3(64), 0, s, t->x2, t->step_ofs, t->abk, t->dblk, t->i0, t->z0,
                       t->sb_ofs, mmax, nlat_2, t->eA);
}

size_t i8_synth_lds_bytes(int sb_max) {
    return sizeof(double) * kI8Waves * 16 * 64 + (size_t)sb_max * (kI8SB * sizeof(double2) + kI8KBperSB * sizeof(int2));
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
    dim3 grid((a.mmax + 1) * ntg, nfg_valid);
    const size_t lds = i8_synth_lds_bytes(t->sb_max);
    ensure_lds_limit(reinterpret_cast<const void *>(legendre_synth_i8_kernel), lds);
    hipLaunchKernelGGL(legendre_synth_i8_kernel, grid, dim3(64 * kI8Waves), lds, s, t->x, t->x2, t->wq, t->step_ofs,
                       t->abk, t->dblk, t->i0, t->z0, t->sb_ofs, t->sb_total, t->eA, planes, eB, spec_ltop, F, four_dist,
                       a.mmax, a.mres, a.nlat, a.nlat_2, a.ltr, a.imlim, a.nfields, a.lat_weight);
}

}  // namespace shtns_amd
