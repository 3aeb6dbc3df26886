// sht_legendre.h -- Legendre-step kernels of shtns_amd as templates, instantiated per
// contraction precision in sht_kernels.hip (FP64), sht_legendre_f32.hip (FP32) and
// sht_legendre_split.hip (FP32 hi/lo split).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <type_traits>

#include "sht_internal.h"

namespace shtns_amd {

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

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
    bool sinv;      // divide the seeds by sin(theta) (unused: vector seeds have their own table)
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
        o.sinv = false;   // vector mode reads seeds of P/sin(theta) (from sin^(m-1), finite at the poles)
    }
    o.nsteps = (ltr >= o.mfun) ? (ltr - o.mfun) / 2 + 1 : 0;
    o.spec_base = (int)(((long)im * (2 * lmax_spec + 2 - (im + 1) * mres)) / 2);
    return o;
}

// Workgroups are dealt round-robin to the 8 XCDs (block b runs on XCD b % 8).  Map physical
// block b to a logical index so that each run of `gsize` consecutive logical blocks lands on
// one XCD, in dispatch order: XCD x takes groups x, x + 8, x + 16, ...  Identity unless nb is
// a multiple of 8 * gsize.
__device__ inline int xcd_cluster(int b, int nb, int gsize) {
    if (gsize <= 1 || nb % (8 * gsize)) return b;
    const int x = b & 7, s = b >> 3;
    return (x + 8 * (s / gsize)) * gsize + s % gsize;
}

// v of quad lane SEL (DPP quad_perm: lane 4q + i reads lane 4q + sel_i, sel_i = (SEL >> 2i) & 3;
// no LDS traffic).  All lanes of the wave must be active.
template <int SEL>
__device__ __forceinline__ double quad_perm(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), SEL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), SEL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// v of lane l - 1 within each row of 16 lanes (DPP row_shr:1); lanes 0, 16, 32, 48 get 0.
// All lanes of the wave must be active.
__device__ __forceinline__ double row_shr1(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x111, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x111, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// ---------------------------------------------------------------------------------------
// Contraction precision PREC (the recurrence, seeds, tables and the final sums stay FP64):
//   0  FP64 matrix cores: v_mfma_f64_16x16x4f64 on z and the coefficients;
//   1  FP32 matrix cores: z and the coefficients rounded to FP32, v_mfma_f32_16x16x4f32,
//      FP32 partial sums over 16 degrees-pairs added into FP64;
//   2  FP32 split: z = z_hi + z_lo and b = b_hi + b_lo (FP32 pairs) with the three products
//      z_hi b_hi + z_lo b_hi + z_hi b_lo on the FP32 matrix cores, same FP32 partial sums.
// A recurrence value sits in one 8-byte LDS slot either as a double or as an FP32 (hi, lo) pair.
// ---------------------------------------------------------------------------------------
template <int PREC>
struct Contract {
    using acc_t = std::conditional_t<PREC == 0, double4_t, float4_t>;
    using a_t = std::conditional_t<PREC == 0, double, float2>;    // A operand (recurrence value)
    using b_t = std::conditional_t<PREC == 0, double, float2>;    // B operand (coefficient or data)
    __device__ static acc_t zero() {
        if constexpr (PREC == 0) return (double4_t){0.0, 0.0, 0.0, 0.0};
        else return (float4_t){0.0f, 0.0f, 0.0f, 0.0f};
    }
    __device__ static void store_z(double *slot, double z) {
        if constexpr (PREC == 0) {
            *slot = z;
        } else {
            const float hi = (float)z;
            *reinterpret_cast<float2 *>(slot) = make_float2(hi, (float)(z - (double)hi));
        }
    }
    __device__ static a_t load_z(const double *slot) {
        if constexpr (PREC == 0) return *slot;
        else return *reinterpret_cast<const float2 *>(slot);
    }
    __device__ static b_t split(double b) {
        if constexpr (PREC == 0) {
            return b;
        } else {
            const float hi = (float)b;
            return make_float2(hi, (float)(b - (double)hi));
        }
    }
    // acc += A * B for one 16x16x4 tile
    __device__ static acc_t mma(a_t a, b_t b, acc_t acc) {
        if constexpr (PREC == 0) {
            return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
        } else if constexpr (PREC == 1) {
            return __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.x, acc, 0, 0, 0);
        } else {
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, b.x, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.y, acc, 0, 0, 0);
            return __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.x, acc, 0, 0, 0);
        }
    }
    // D layout on gfx950: register j of lane l holds D[row][l % 16] with
    //   f64 16x16x4: row = 4 j + l / 16;   f32 16x16x4: row = 4 (l / 16) + j.
    __device__ static int d_row(int j, int lane) {
        if constexpr (PREC == 0) return 4 * j + (lane >> 4);
        else return 4 * (lane >> 4) + j;
    }
};

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

// NF4 > 0 (FP64, NCG = 1, at most 3 fields): field c of the group is contracted with
// v_mfma_f64_4x4x4f64 (four 4x4x4 blocks: 16 latitudes x 4 steps x the field's 4 columns),
// so no matrix-core columns are spent on absent fields.  gfx950 lane layouts of that
// instruction: A lane l = (block (l>>2)&3, row l&3, k l>>4), B lane l = (block (l>>2)&3,
// k l>>4, col l&3), D lane l = (block (l>>2)&3, row l>>4, col l&3); with block b and row i
// taken as latitude 4b + i of a 16-latitude row tile, the A operand is the 16x16x4 one.
template <int NCG, int PREC, int NF4 = 0>
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
    // heavy orders first: block x -> (order, tile group); the tile groups of one order run on
    // one XCD (they read the same coefficient rows, which then hit in that XCD's L2)
    const int ntg = (ntl + kSynthWaves - 1) / kSynthWaves;   // tile groups of this launch
    const int bx = xcd_cluster(blockIdx.x, gridDim.x, ntg);
    const int im = bx / ntg;
    const int tile = tile0 + (bx - im * ntg) * kSynthWaves + wave;
    const int g = blockIdx.y;
    const int fbase = g * 4 * NCG;
    if (tile >= tile0 + ntl || tile * 64 >= nlat_2) return;   // no block-level synchronization in this kernel
    const int k = tile * 64 + lane;
    const bool kvalid = (k < nlat_2);
    const int kk = kvalid ? k : (nlat_2 - 1);

    using C = Contract<PREC>;
    static_assert(NF4 == 0 || (PREC == 0 && NCG == 1 && NF4 <= 3), "4x4x4 path: FP64, one group of <= 3 fields");
    constexpr int NA = NF4 ? NF4 : NCG;   // accumulator sets: fields (4x4x4) or column groups
    using acc64_t = std::conditional_t<NF4 == 0, double4_t, double>;
    acc64_t acc[4][NA];   // FP64 sums
    typename C::acc_t part[4][NCG];   // FP32 partial sums over 16 steps (PREC > 0; unused at PREC 0)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
        for (int c = 0; c < NA; ++c) acc[r][c] = acc64_t{};
#pragma unroll
        for (int c = 0; c < NCG; ++c) part[r][c] = C::zero();
    }
    auto flush = [&]() {
        if constexpr (PREC > 0) {
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < NCG; ++c) {
                    acc[r][c] += __builtin_convertvector(part[r][c], double4_t);
                    part[r][c] = C::zero();
                }
        }
    };

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
        // coefficients of steps i .. i+3: the table holds at least nsteps + 7 rows per order
        // (continued past the last degree), so no clamp is needed and the four pairs are one
        // 64-byte scalar load
        auto load_cf = [&](int i, double2 *cf) {
#pragma unroll
            for (int s = 0; s < 4; ++s) cf[s] = abm[i + s];
        };
        auto advance = [&](int i, int buf, const double2 *cf) {
            if (i <= wmax) {   // some lane of this wave starts in this block
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    if (i + s == i0) { za = sd.x; zb = sd.y; }
                    C::store_z(&ytile[wave][buf][s][lane], zb);
                    const double zn = fma(fma(cf[s].x, x2, cf[s].y), zb, za);
                    za = zb;
                    zb = zn;
                }
            } else {
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    C::store_z(&ytile[wave][buf][s][lane], zb);
                    const double zn = fma(fma(cf[s].x, x2, cf[s].y), zb, za);
                    za = zb;
                    zb = zn;
                }
            }
        };
        // B operand of column group (or field) c for the block at step i
        // (a 32-bit byte offset from the wave-uniform base: one address register)
        auto load_b1 = [&](int i, int c) {
            const int ii = min(i + (lane >> 4), slot_rows - 1);   // rows >= nsteps hold zeros
            const unsigned e = NF4 > 0 ? (unsigned)(ii * NCOL + c * 4 + (lane & 3)) : (unsigned)(ii * NCOL + c * 16 + (lane & 15));
            return *(const double *)((const char *)qb + 8u * e);
        };

        double bcur[NA];
        {
            double2 cf0[4];
            load_cf(ibeg, cf0);
            advance(ibeg, 0, cf0);
        }
        // in column-group order, as in the loop (the first MFMAs then wait for the first load only)
#pragma unroll
        for (int c = 0; c < NA; ++c) {
            bcur[c] = load_b1(ibeg, c);
            __builtin_amdgcn_sched_barrier(0);
        }
        // MFMAs of column group (or field) c on this block's A operands
        auto mma_group = [&](int c, const typename C::a_t *a) {
            if constexpr (NF4 > 0) {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    acc[r][c] = __builtin_amdgcn_mfma_f64_4x4x4f64(a[r], bcur[c], acc[r][c], 0, 0, 0);
            } else {
                const typename C::b_t bb = C::split(bcur[c]);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    if constexpr (PREC == 0) acc[r][c] = C::mma(a[r], bb, acc[r][c]);
                    else part[r][c] = C::mma(a[r], bb, part[r][c]);
                }
            }
        };
        int buf = 0, blk = 0;
        for (int i = ibeg; i < nsteps; i += 4, ++blk) {
            __builtin_amdgcn_s_waitcnt(0xc07f);   // LDS tile of this block written
            __builtin_amdgcn_wave_barrier();
            typename C::a_t a[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) a[r] = C::load_z(&ytile[wave][buf][lane >> 4][r * 16 + (lane & 15)]);
            mma_group(0, a);
            // The next block's coefficients: a scalar load issued behind the first MFMAs (scalar
            // and LDS loads share one counter, so issued before the A reads it would delay
            // them), in flight while the other MFMAs issue.
            __builtin_amdgcn_sched_barrier(0);
            double2 cfn[4];
            load_cf(i + 4, cfn);
            bcur[0] = load_b1(i + 4, 0);   // next block's B operands, each once its MFMAs have issued
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int c = 1; c < NA; ++c) {
                mma_group(c, a);
                bcur[c] = load_b1(i + 4, c);
                __builtin_amdgcn_sched_barrier(0);
            }
            advance(i + 4, buf ^ 1, cfn);   // also after the last block (in bounds, unused)
            buf ^= 1;
            if ((blk & 3) == 3) flush();
        }
        flush();
    }

    // Epilogue.  D[row][col]: the lane of column col = lane & 15 holds rows C::d_row(j, lane) of
    // row tile r (latitudes tile*64 + 16r + row).  Columns are 4 fields x (E.re, E.im, O.re,
    // O.im), so a quad of lanes is one field's (E, O) pairs: every lane reads its pair's O and
    // E values (DPP), E lanes write north = E + x O and O lanes south = E - x O.
    const int col = NF4 ? (lane & 3) : (lane & 15);
    const int reim = col & 1;
    const bool o_lane = (col & 2) != 0;
    const bool zero_im = (im == 0) && reim;   // order 0 has no imaginary part
    if constexpr (NF4 > 0) {
        // D lane l: latitude tile*64 + 16 r + 4 ((l >> 2) & 3) + (l >> 4), column l & 3
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int kr = tile * 64 + r * 16 + 4 * ((lane >> 2) & 3) + (lane >> 4);
            const int kc = min(kr, nlat_2 - 1);
            const double x = xk[kc];
            const double xs = o_lane ? -x : x;
            const double wz = zero_im ? 0.0 : (lat_weight ? wqk[kc] * ((2 * kr + 1 == nlat) ? 2.0 : 1.0) : 1.0);
            const int row = o_lane ? (nlat - 1 - kr) : kr;
#pragma unroll
            for (int c = 0; c < NA; ++c) {
                const int f = fbase + c;
                const double o = quad_perm<0xEE>(acc[r][c]);
                const double e = quad_perm<0x44>(acc[r][c]);
                const double val = fma(xs, o, e) * wz;
                if (kr < nlat_2 && f < nfields)
                    reinterpret_cast<double *>(F + (size_t)f * four_dist + (size_t)im * nlat)[2 * row + reim] = val;
            }
        }
        return;
    } else {
    // every latitude of the tile and every field of the group exist (wave-uniform)
    const bool full = (tile * 64 + 64 <= nlat_2) && (fbase + 4 * NCG <= nfields);
    auto store_tile = [&](auto check) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int row[4];
            bool ok[4];
            double xs[4], wz[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int kr = tile * 64 + r * 16 + C::d_row(j, lane);
                const int kc = min(kr, nlat_2 - 1);
                const double x = xk[kc];
                xs[j] = o_lane ? -x : x;
                // the middle row of an odd grid is its own mirror: its stored pair weight is halved
                wz[j] = zero_im ? 0.0 : (lat_weight ? wqk[kc] * ((2 * kr + 1 == nlat) ? 2.0 : 1.0) : 1.0);
                row[j] = o_lane ? (nlat - 1 - kr) : kr;
                ok[j] = kr < nlat_2;
            }
#pragma unroll
            for (int c = 0; c < NCG; ++c) {
                const int f = fbase + c * 4 + (col >> 2);
                double *Ff = reinterpret_cast<double *>(F + (size_t)min(f, nfields - 1) * four_dist +
                                                        (size_t)im * nlat) + reim;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const double o = quad_perm<0xEE>(acc[r][c][j]);   // lanes (2, 3, 2, 3)
                    const double e = quad_perm<0x44>(acc[r][c][j]);   // lanes (0, 1, 0, 1)
                    const double val = fma(xs[j], o, e) * wz[j];
                    if (!decltype(check)::value || (ok[j] && f < nfields)) Ff[2 * row[j]] = val;
                }
            }
        }
    };
    if (full) store_tile(std::false_type{});
    else store_tile(std::true_type{});
    }
}

// SHTNS_AMD_SYNTH_16=1: keep the 16x16x4 kernels for 1 to 3 fields (for comparisons)
inline bool synth_force16() {
    static const bool f = getenv("SHTNS_AMD_SYNTH_16") != nullptr;
    return f;
}

inline int synth_ncg(int nfields, int want = 0) {
    if (want == 1 || want == 2 || want == 4) return want;
    return nfields > 8 ? 4 : (nfields > 4 ? 2 : 1);
}

template <int PREC>
void synth_tiles_launch(const LegendreArgs &a, int spec_ltop, double2 *F, long four_dist, const double *qmix,
                        int nsteps_total, int tile0, int ntl, hipStream_t s) {
    const DeviceTables *t = a.t;
    const int ntg = (ntl + kSynthWaves - 1) / kSynthWaves;
    const int ncg = synth_ncg(a.nfields, a.synth_ncg);
    const int ng = (a.nfields + 4 * ncg - 1) / (4 * ncg);
    dim3 grid((a.mmax + 1) * ntg, ng);
#define SYNTH_LAUNCH(NC, ...)                                                                                    \
    hipLaunchKernelGGL((legendre_synth_kernel<NC, PREC, ##__VA_ARGS__>), grid, dim3(64 * kSynthWaves), 0, s, t->x, t->x2, t->st, t->wq, \
                       t->step_ofs, t->ab, a.vector_mode ? t->i0v : t->i0, a.vector_mode ? t->z0v : t->z0, qmix, \
                       nsteps_total, spec_ltop, F, four_dist, a.mmax,                                              \
                       a.mres, a.nlat, a.nlat_2, a.ltr, a.imlim, a.nfields, a.vector_mode, a.lat_weight, tile0, ntl)
    if constexpr (PREC == 0) {   // 1 to 3 fields: 4x4x4 blocks, no idle matrix-core columns
        if (a.nfields <= 3 && ncg == 1 && !synth_force16()) {
            if (a.nfields == 1) SYNTH_LAUNCH(1, 1);
            else if (a.nfields == 2) SYNTH_LAUNCH(1, 2);
            else SYNTH_LAUNCH(1, 3);
            return;
        }
    }
    if (ncg == 4) SYNTH_LAUNCH(4);
    else if (ncg == 2) SYNTH_LAUNCH(2);
    else SYNTH_LAUNCH(1);
#undef SYNTH_LAUNCH
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
// NF4 > 0 (FP64, NCG = 1, at most 3 fields): field c is contracted with v_mfma_f64_4x4x4f64
// (four 4x4x4 blocks = 16 steps x 4 latitudes x the field's 4 columns; the A operand is the
// 16x16x4 one, B lane l holds column l & 3 of latitude 4 kg + (l >> 4) in every block, and
// D lane l holds step 4 ((l >> 2) & 3) + (l >> 4), column l & 3).
// NG > 1 (one tile per wave): the block holds NG field groups of the same order, waves
// w = tw + NW g (tile tw, group g); each group's NW partial tiles are reduced separately.
// Six-tile grids then run 12-wave blocks (three waves on every SIMD) instead of 6-wave ones.
// GR > 1 (4x4x4 path, one tile per wave): GR 16-step groups share one barrier pair and one
// reduction round (16 GR rows of partial tiles), for batches whose MFMA work per group is
// too small to cover those fixed costs.
template <int NW, int NLR, int NCG, int PREC, int NF4 = 0, int NG = 1, int GR = 1>
__global__ __launch_bounds__(64 * NW * NG) void legendre_anal_kernel(
    const double *__restrict__ xk, const double *__restrict__ x2k, const double *__restrict__ stk,
    const double *__restrict__ wqk, const int *__restrict__ step_ofs, const double2 *__restrict__ ab,
    const double4 *__restrict__ pm, const int *__restrict__ i0tab, const double2 *__restrict__ z0tab,
    const double2 *__restrict__ F, long four_dist, double2 *__restrict__ Q, long spec_dist, int lmax_spec,
    int spec_ltop, int mmax, int mres, int nlat, int nlat_2, int ltr, int imlim, int nfields, int vector_mode,
    double mpos_scale, int lat_weight, int cf_lds, int tile_base) {
    static_assert(NF4 == 0 || (PREC == 0 && NCG == 1 && NF4 <= 3), "4x4x4 path: FP64, one group of <= 3 fields");
    static_assert(NG == 1 || (NF4 == 0 && NLR == 1), "several field groups per block: one tile per wave");
    static_assert(GR == 1 || (GR == 2 && NF4 > 0 && NLR == 1 && NG == 1), "two groups per round: 4x4x4 path, one tile");
    constexpr int RR = 16 * GR;                   // steps per reduction round
    constexpr int NA = NF4 ? NF4 : NCG;           // operand / accumulator sets
    constexpr int NCOL = NF4 ? 4 * NF4 : 16 * NCG;
    constexpr int NT = 64 * NW * NG;              // threads per block
    __shared__ double ytile[NW * NG][16][65];
    // cf_lds: the order's recurrence coefficients (alpha_i, beta_i), i < nsteps, are copied to
    // LDS once per block (dynamic LDS), so each 16-step group reads them at LDS latency instead
    // of waiting on scalar loads that miss to HBM
    extern __shared__ double2 cfl[];
    // single-buffered: two barriers per block.  Rows padded to NCOL + 2 doubles: a post-mix read
    // (32 lanes = steps 0..15 of two adjacent columns) then falls on 32 distinct bank pairs
    __shared__ double red[NG][NW][RR][NCOL + 2];
    __shared__ double carry[NG][2][NCOL];
    __shared__ int s_ibeg[NW * NG];

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int tw = NG == 1 ? wave : wave % NW, gw = NG == 1 ? 0 : wave / NW;   // tile wave, field group
    // block -> (order, field groups), field groups fastest: blocks are dispatched in index
    // order, so all field groups of the heavy low orders start first
    const int ngr = gridDim.x / (mmax + 1);
    const int im = blockIdx.x / ngr;
    const int fblk = (blockIdx.x - im * ngr) * 4 * NCG * NG;   // first field of the block
    const int fbase = fblk + gw * 4 * NCG;                    // first field of this wave's group

    const OrderInfo o = order_info(im, lmax_spec, mmax, mres, (im <= imlim) ? ltr : -1, vector_mode);
    const int nsteps = (im <= imlim) ? o.nsteps : 0;
    const int sofs = step_ofs[o.slot];
    const UniformTable2 abm(ab + sofs);   // uniform indices -> scalar loads
    const double4 *pmm = pm + sofs;

    // (multi-tile waves keep their registers for the operands: scalar loads there)
    if constexpr (NLR >= 3) cf_lds = 0;
    // one-tile waves load the coefficient-table copy before the operands below: global loads
    // complete in issue order, so its LDS store (before the first barrier) then waits only for
    // these (multi-tile waves have no registers to spare and load it afterwards)
    constexpr int NCF = (NLR == 1) ? 4 : 0;   // per thread; the rest of the table, if any, below
    double2 cfr[NCF > 0 ? NCF : 1];
    if (NCF > 0 && cf_lds) {
#pragma unroll
        for (int q = 0; q < NCF; ++q) {
            const int e = threadIdx.x + q * NT;
            cfr[q] = (e < nsteps) ? ab[sofs + e] : make_double2(0.0, 0.0);
        }
    }

    double za[NLR], zb[NLR], x2v[NLR], sdx[NLR], sdy[NLR];
    int i0v[NLR];
    using C = Contract<PREC>;
    typename C::b_t Ereg[NLR][NA][16];
    int my_ibeg = nsteps, my_imax = 0;
#pragma unroll
    for (int r = 0; r < NLR; ++r) {
        const int k = (tile_base + tw + NW * r) * 64 + lane;
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
        // (E.re, E.im, O.re, O.im = weighted N+S, N+S, (N-S) x, (N-S) x).  Branch-free, every
        // load issued before the first use: the block's first MFMA waits for these.
        const int col = NF4 ? (lane & 3) : (lane & 15), comp = col & 3;
        const bool keep = nsteps > 0 && !(o.real_only && (comp & 1));
        const int kl0 = (tile_base + tw + NW * r) * 64 + (lane >> 4);   // latitude of K-group 0
        if constexpr (NLR >= 3) {
#pragma unroll
            for (int c = 0; c < NA; ++c) {
                const int f = NF4 ? fbase + c : fbase + 4 * c + (col >> 2);
                const bool fv = keep && f < nfields;
                const double2 *Ff = F + (size_t)(fv ? f : 0) * four_dist + (size_t)im * nlat;
#pragma unroll
                for (int kg = 0; kg < 16; ++kg) {
                    // several tiles per wave: one K-group at a time (loads hoisted together
                    // would need more registers than the tiles' operands leave)
                    const int kl = kl0 + 4 * kg;
                    double v = 0.0;
                    if (fv && kl < nlat_2) {
                        const double2 fn = Ff[kl], fs = Ff[nlat - 1 - kl];
                        const double w = lat_weight ? wqk[kl] : ((2 * kl + 1 == nlat) ? 0.5 : 1.0);   // odd grid: equator row once
                        v = (comp == 0 ? fn.x + fs.x : comp == 1 ? fn.y + fs.y
                             : comp == 2 ? (fn.x - fs.x) * xk[kl] : (fn.y - fs.y) * xk[kl]) * w;
                    }
                    Ereg[r][c][kg] = C::split(v);
                }
            }
        } else {
            // Batches of KB K-groups: every load of a batch is issued before its first use
            // (clamped indices, no branches; each lane loads only the real or imaginary part
            // its column takes), and the batches are fenced so the scheduler does not hoist
            // later batches' loads above them.  The arithmetic is that of the branchy form
            // above, rounding for rounding: (a + b) w for E, ((a - b) x) w for O, with a - b
            // as one fma by -1.
            constexpr int KB = 4;
            const bool neg = comp & 2;
            const double sgn = neg ? -1.0 : 1.0;
            const double *Fc[NA];
            bool fvc[NA];
#pragma unroll
            for (int c = 0; c < NA; ++c) {
                const int f = NF4 ? fbase + c : fbase + 4 * c + (col >> 2);
                fvc[c] = keep && f < nfields;
                Fc[c] = reinterpret_cast<const double *>(F + (size_t)(fvc[c] ? f : 0) * four_dist + (size_t)im * nlat) +
                        (comp & 1);
            }
#pragma unroll
            for (int kb = 0; kb < 16; kb += KB) {
                double xq[KB], wq[KB], aq[NA][KB], bq[NA][KB];
#pragma unroll
                for (int q = 0; q < KB; ++q) {
                    const int kl = kl0 + 4 * (kb + q);
                    const int klx = (kl < nlat_2) ? kl : 0;
                    xq[q] = xk[klx];
                    const double wl = wqk[klx];
                    wq[q] = lat_weight ? wl : ((2 * klx + 1 == nlat) ? 0.5 : 1.0);   // odd grid: equator row once
#pragma unroll
                    for (int c = 0; c < NA; ++c) {
                        const int klc = (fvc[c] && kl < nlat_2) ? kl : 0;
                        aq[c][q] = Fc[c][2 * klc];
                        bq[c][q] = Fc[c][2 * (nlat - 1 - klc)];
                    }
                }
#pragma unroll
                for (int q = 0; q < KB; ++q) {
                    const int kl = kl0 + 4 * (kb + q);
#pragma unroll
                    for (int c = 0; c < NA; ++c) {
                        const double s = fma(sgn, bq[c][q], aq[c][q]);
                        double v = (fvc[c] && kl < nlat_2) ? (neg ? s * xq[q] : s) * wq[q] : 0.0;
                        // computed here, not sunk towards the MFMAs (which would keep every
                        // batch's loaded values live across the prologue)
                        asm volatile("" : "+v"(v));
                        Ereg[r][c][kb + q] = C::split(v);
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    }
    for (int off = 32; off >= 1; off >>= 1) {
        my_ibeg = min(my_ibeg, __shfl_xor(my_ibeg, off));
        my_imax = max(my_imax, __shfl_xor(my_imax, off));
    }
    my_imax = __builtin_amdgcn_readfirstlane(my_imax);
    if (lane == 0) s_ibeg[wave] = my_ibeg;
    if (threadIdx.x < 2 * NCOL * NG) (&carry[0][0][0])[threadIdx.x] = 0.0;
    if (cf_lds) {
#pragma unroll
        for (int q = 0; q < NCF; ++q) {
            const int e = threadIdx.x + q * NT;
            if (e < nsteps) cfl[e] = cfr[q];
        }
        for (int e = threadIdx.x + NCF * NT; e < nsteps; e += NT) cfl[e] = ab[sofs + e];   // rest, if any
    }
    __syncthreads();
    int ibeg = nsteps;
    for (int w = 0; w < NW * NG; ++w) ibeg = min(ibeg, s_ibeg[w]);
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

    // post-mix work split: thread handles elements e = threadIdx.x + j*NT (j < NE) of the
    // 16 x NCOL block, step row e % 16 of column e / 16, each for all NG field groups (the
    // same coefficient row); the coefficient rows are prefetched before the MFMA phase.
    // Each 16-lane DPP row holds the 16 steps of one column, so the sum of step row - 1 (the
    // even degrees' second term) is the neighbouring lane's, and a wave's stores cover the
    // 32 consecutive degrees of one field (16 NCOL is a multiple of 64: whole waves take part
    // or skip an element together).
    constexpr int NE = (RR * NCOL + NT - 1) / NT;
    auto cf_at = [&](auto from_lds, int idx) {
        idx = min(idx, nsteps - 1);
        if constexpr (decltype(from_lds)::value) return cfl[idx];   // uniform: LDS broadcast
        else return abm[idx];                                        // uniform: scalar load
    };
    // recurrence of the 16-step group from ig of tile r into this wave's tile buffer
    auto recur_plain = [&](int r, int ig, const double2 *cfg) {
        if (ig <= my_imax) {   // some lane of this wave starts its recurrence here
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                if (ig + s == i0v[r]) { za[r] = sdx[r]; zb[r] = sdy[r]; }
                C::store_z(&ytile[wave][s][lane], zb[r]);
                const double zn = fma(fma(cfg[s].x, x2v[r], cfg[s].y), zb[r], za[r]);
                za[r] = zb[r];
                zb[r] = zn;
            }
        } else {
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                C::store_z(&ytile[wave][s][lane], zb[r]);
                const double zn = fma(fma(cfg[s].x, x2v[r], cfg[s].y), zb[r], za[r]);
                za[r] = zb[r];
                zb[r] = zn;
            }
        }
    };
    auto load_cf16 = [&](double2 *cf, int ig) {
        if (cf_lds) {
#pragma unroll
            for (int s = 0; s < 16; ++s) cf[s] = cf_at(std::true_type{}, ig + s);
        } else {
#pragma unroll
            for (int s = 0; s < 16; ++s) cf[s] = cf_at(std::false_type{}, ig + s);
        }
    };
    int b = 0;
    for (int i = ibeg; i < nsteps; i += RR, ++b) {
        double pcx[NE], pcz[NE], pcp[NE];
#pragma unroll
        for (int j = 0; j < NE; ++j) {
            const int e = threadIdx.x + j * NT;
            const int row = e & (RR - 1);
            const int ii = min(i + row, nsteps - 1);
            const double4 cf4 = pmm[ii];
            pcx[j] = cf4.x;
            pcz[j] = cf4.z;
            pcp[j] = (ii > 0) ? pmm[ii - 1].y : 0.0;
        }
        using accA_t = std::conditional_t<NF4 == 0, typename C::acc_t, double>;
        accA_t acc[GR][NA], acc2[GR][NA];   // two chains per column group (even / odd K-groups)
#pragma unroll
        for (int q = 0; q < GR; ++q)
#pragma unroll
            for (int c = 0; c < NA; ++c) acc[q][c] = acc2[q][c] = accA_t{};
#pragma unroll
        for (int q = 0; q < GR; ++q) {
            double2 cf[16];
            load_cf16(cf, i + 16 * q);
#pragma unroll
            for (int r = 0; r < NLR; ++r) {
                recur_plain(r, i + 16 * q, cf);
                typename C::a_t a[16];
                __builtin_amdgcn_s_waitcnt(0xc07f);
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int kg = 0; kg < 16; ++kg) a[kg] = C::load_z(&ytile[wave][lane & 15][4 * kg + (lane >> 4)]);
                __builtin_amdgcn_s_waitcnt(0xc07f);   // all A operands in registers before the tile is rewritten
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int kg = 0; kg < 16; kg += 2)
#pragma unroll
                    for (int c = 0; c < NA; ++c) {
                        if constexpr (NF4 > 0) {
                            acc[q][c] = __builtin_amdgcn_mfma_f64_4x4x4f64(a[kg], Ereg[r][c][kg], acc[q][c], 0, 0, 0);
                            acc2[q][c] = __builtin_amdgcn_mfma_f64_4x4x4f64(a[kg + 1], Ereg[r][c][kg + 1], acc2[q][c], 0, 0, 0);
                        } else {
                            acc[q][c] = C::mma(a[kg], Ereg[r][c][kg], acc[q][c]);
                            acc2[q][c] = C::mma(a[kg + 1], Ereg[r][c][kg + 1], acc2[q][c]);
                        }
                    }
            }
        }
#pragma unroll
        for (int q = 0; q < GR; ++q)
#pragma unroll
            for (int c = 0; c < NA; ++c) acc[q][c] += acc2[q][c];
        const int pb = b & 1;
        __syncthreads();   // previous block's post-mix has finished reading red
        if constexpr (NF4 > 0) {
            // 4x4x4 D: step 4 ((lane >> 2) & 3) + (lane >> 4), column lane & 3 of field c
#pragma unroll
            for (int q = 0; q < GR; ++q)
#pragma unroll
                for (int c = 0; c < NA; ++c)
                    red[gw][tw][16 * q + 4 * ((lane >> 2) & 3) + (lane >> 4)][4 * c + (lane & 3)] = acc[q][c];
        } else {
            // D layout (gfx950): acc[c][j] = D[step row 4*j + lane/16][col lane%16]
#pragma unroll
            for (int c = 0; c < NCG; ++c)
#pragma unroll
                for (int j = 0; j < 4; ++j) red[gw][tw][C::d_row(j, lane)][16 * c + (lane & 15)] = acc[0][c][j];
        }
        __syncthreads();
        // sum over each group's waves + post-mix
#pragma unroll
        for (int j = 0; j < NE; ++j) {
            const int e = threadIdx.x + j * NT;
            if (e >= RR * NCOL) break;   // wave-uniform
            const int row = e & (RR - 1), col = e / RR;
            const int comp = col & 3;
            const int ii = i + row;
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                const int f = fblk + g * 4 * NCG + (col >> 2);
                double sum = 0.0;
#pragma unroll
                for (int w = 0; w < NW; ++w) sum += red[g][w][row][col];
                // step row - 1 of this column: lane - 1 (the whole wave is active here; within
                // a DPP row for 16-row rounds, a wave-wide shift for 32), or the previous
                // round's last row
                double sprev = (GR == 1) ? row_shr1(sum) : __shfl_up(sum, 1);
                if (row == 0) sprev = carry[g][pb ^ 1][col];
                if (row == RR - 1 && comp < 2) carry[g][pb][col] = sum;
                if (f >= nfields || ii >= nsteps_out) continue;
                const int le = o.mfun + 2 * ii;
                double v = 0.0;
                if (ii < nsteps) {
                    if (comp < 2) {
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
    }
    // degrees not reached by the step loop: before ibeg (negligible functions) and after nsteps
    for (int ii = threadIdx.x; ii < nsteps_out; ii += NT) {
        if (ii >= ibeg && ii < ibeg + ((nsteps - ibeg + RR - 1) / RR) * RR) continue;
        const int le = o.mfun + 2 * ii;
        for (int f = fblk; f < min(nfields, fblk + 4 * NCG * NG); ++f) {
            double2 *Qf = Qbase + (size_t)f * spec_dist;
            if (le <= spec_ltop) Qf[le] = make_double2(0.0, 0.0);
            if (le + 1 <= spec_ltop) Qf[le + 1] = make_double2(0.0, 0.0);
        }
    }
    if (vector_mode && im == 0 && threadIdx.x == 0) {   // degree 0 of the vector m=0 spectra
        for (int f = fblk; f < min(nfields, fblk + 4 * NCG * NG); ++f) Qbase[(size_t)f * spec_dist] = make_double2(0.0, 0.0);
    }
}

// static LDS of legendre_anal_kernel, and whether the order's coefficient table (ltr/2 + 2
// pairs) fits beside it in the 160 KB of a CU
constexpr size_t anal_static_lds(int NW, int NG, int NCOL, int GR) {
    return sizeof(double) * NG * (NW * 16 * 65 + NW * 16 * GR * (NCOL + 2) + 2 * NCOL) + sizeof(int) * NW * NG;
}
inline bool anal_cf_fits(size_t static_lds, int ltr) {
    return static_lds + sizeof(double2) * (size_t)(ltr / 2 + 2) + 1024 <= 160 * 1024;
}

template <int NW, int NLR, int NCG, int PREC, int NF4 = 0, int NG = 1, int GR = 1>
void launch_anal_t(const LegendreArgs &a, const double2 *F, long four_dist, double2 *Q, long spec_dist,
                 int spec_ltop, hipStream_t s, int tile_base = 0) {
    const DeviceTables *t = a.t;
    dim3 grid((a.mmax + 1) * ((a.nfields + 4 * NCG * NG - 1) / (4 * NCG * NG)));
    // static LDS of the kernel, and the coefficient table if it fits beside it in 160 KB
    constexpr int NCOL = NF4 ? 4 * NF4 : 16 * NCG;
    constexpr size_t static_lds = anal_static_lds(NW, NG, NCOL, GR);
    const size_t cf_bytes = sizeof(double2) * (size_t)(a.ltr / 2 + 2);
    const int cf_lds = anal_cf_fits(static_lds, a.ltr) ? 1 : 0;
    const size_t dyn = cf_lds ? cf_bytes : 0;
    ensure_lds_limit((const void *)legendre_anal_kernel<NW, NLR, NCG, PREC, NF4, NG, GR>, dyn);
    hipLaunchKernelGGL((legendre_anal_kernel<NW, NLR, NCG, PREC, NF4, NG, GR>), grid, dim3(64 * NW * NG), dyn, s, t->x, t->x2, t->st,
                       t->wq, t->step_ofs, t->ab, t->pm, a.vector_mode ? t->i0v : t->i0,
                       a.vector_mode ? t->z0v : t->z0, F, four_dist, Q, spec_dist,
                       spec_ltop == a.lmax + 1 ? a.lmax + 1 : a.lmax, spec_ltop, a.mmax, a.mres, a.nlat, a.nlat_2,
                       a.ltr, a.imlim, a.nfields, a.vector_mode, a.mpos_scale, a.lat_weight, cf_lds, tile_base);
}

// Q[f][i] = sum_s part[s][f][i], i < n (the latitude splits' partial spectra)
void launch_sum_splits(const double2 *part, long part_dist, int nsplit, int nfields, double2 *Q, long spec_dist,
                       long n, hipStream_t s);

template <int PREC>
void anal_launch(const LegendreArgs &a, const double2 *F, long four_dist, double2 *Q, long spec_dist,
                 int spec_ltop, hipStream_t s) {
    const int ntiles = (a.nlat_2 + 63) / 64;
    if constexpr (PREC == 0) {
        const int ns = anal_splits(a);
        if (ns > 1 && a.part) {   // 12-wave blocks over a latitude split each, then the sum
            for (int k = 0; k < ns; ++k) {
                double2 *P = a.part + (size_t)k * a.nfields * a.part_dist;
                // one field: two 16-step groups per barrier round where the coefficient table
                // still fits in LDS beside the larger partial-tile buffer
                if (a.nfields == 1 && anal_cf_fits(anal_static_lds(12, 1, 4, 2), a.ltr))
                    launch_anal_t<12, 1, 1, 0, 1, 1, 2>(a, F, four_dist, P, a.part_dist, spec_ltop, s, 12 * k);
                else if (a.nfields == 1) launch_anal_t<12, 1, 1, 0, 1>(a, F, four_dist, P, a.part_dist, spec_ltop, s, 12 * k);
                else launch_anal_t<12, 1, 1, 0, 2>(a, F, four_dist, P, a.part_dist, spec_ltop, s, 12 * k);
            }
            launch_sum_splits(a.part, a.part_dist, ns, a.nfields, Q, spec_dist, a.part_dist, s);
            return;
        }
    }
    // waves per block: one latitude tile per wave up to 12 waves (measured fastest: lowest
    // register count, 3 waves per SIMD); beyond, a divisor of the tile count, at most 8
    int nw = 1;
    if (ntiles <= 12) nw = ntiles;
    else
        for (int d = 8; d >= 1; --d)
            if (ntiles % d == 0) { nw = d; break; }
    static const int nw_env = getenv("SHTNS_AMD_ANAL_NW") ? atoi(getenv("SHTNS_AMD_ANAL_NW")) : 0;
    if (nw_env > 0 && nw_env <= 16 && ntiles % nw_env == 0) nw = nw_env;
    if (a.anal_nw > 0 && ntiles % a.anal_nw == 0 && anal_shape_ok(a.anal_nw, ntiles / a.anal_nw)) nw = a.anal_nw;
    int nlr = ntiles / nw;
    if (nlr > 6) { nw = 8; nlr = (ntiles + 7) / 8; }
    static const int ncg_env = getenv("SHTNS_AMD_ANAL_NCG") ? atoi(getenv("SHTNS_AMD_ANAL_NCG")) : 0;
    const int ncg_want = a.anal_ncg ? a.anal_ncg : ncg_env;
    bool wide = (a.nfields > 4) && (nlr <= 3);   // 8 fields per block when registers allow
    if (ncg_want == 1) wide = false;
    if (ncg_want == 2 && nlr <= 3) wide = (a.nfields > 4);
    // one or two fields (FP64): 4x4x4 blocks, no idle matrix-core columns
    const int nf4 = (PREC == 0 && a.nfields <= 2 && !synth_force16()) ? a.nfields : 0;
    // six one-tile waves and more than one 8-field group: two groups per block, 12 waves (three
    // on every SIMD instead of 1.5); SHTNS_AMD_ANAL_NG=1 keeps the 6-wave blocks
    static const int ng_env = getenv("SHTNS_AMD_ANAL_NG") ? atoi(getenv("SHTNS_AMD_ANAL_NG")) : 0;
    if (nf4 == 0 && nw == 6 && nlr == 1 && wide && a.nfields > 8 && ng_env != 1) {
        launch_anal_t<6, 1, 2, PREC, 0, 2>(a, F, four_dist, Q, spec_dist, spec_ltop, s);
        return;
    }
#define ANAL_CASE(W, R)                                                                              \
    if (nw == W && nlr <= R) {                                                                       \
        if constexpr (PREC == 0) {                                                                   \
            if (nf4 == 1) { launch_anal_t<W, R, 1, 0, 1>(a, F, four_dist, Q, spec_dist, spec_ltop, s); return; } \
            if (nf4 == 2) { launch_anal_t<W, R, 1, 0, 2>(a, F, four_dist, Q, spec_dist, spec_ltop, s); return; } \
        }                                                                                            \
        if (wide) launch_anal_t<W, R, 2, PREC>(a, F, four_dist, Q, spec_dist, spec_ltop, s);        \
        else launch_anal_t<W, R, 1, PREC>(a, F, four_dist, Q, spec_dist, spec_ltop, s);             \
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
