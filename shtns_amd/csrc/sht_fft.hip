// sht_fft.hip -- longitude FFTs of the transforms (SHTns does these with FFTW/cuFFT:
// sht_init.c:353-518 plans nlat c2c DFTs of nphi/2 points with two real rows packed as one
// complex row; sht_gpu.cu:279-303 uses cuFFT Z2Z the same way).
//
// Each latitude row of nphi real values is transformed as one complex sequence of
// M = nphi/2 points z_n = V[2n] + i V[2n+1] (analysis) or produced from one (synthesis).
// A sequence lives in the registers of LQ lanes, R = M/LQ points per lane (R = 24 when 3
// divides M, else 16), lane l holding points l + LQ*j, j < R.  The transform is a Stockham
// (self-sorting, decimation in time) sequence of radix-P passes; the butterflies of a pass
// whose inputs are a lane's own points run in registers, and the data moves between lanes
// through LDS only between passes (M = 1536: radices 24, 8, 8 -> two LDS exchanges).  The
// first pass reads HBM directly and the last pass leaves its outputs in the lane that
// stores them, so nothing else is staged.
#include <algorithm>
#include <utility>

#include "sht_internal.h"

namespace shtns_amd {
namespace {

// ---- compile-time helpers ---------------------------------------------------------------
template <class F, int... I>
__device__ __forceinline__ void sfor_(F &&f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
// calls f(integral_constant<int, i>) for i = 0 .. N-1
template <int N, class F>
__device__ __forceinline__ void sfor(F &&f) {
    sfor_(f, std::make_integer_sequence<int, N>{});
}

// cos(2 pi j / 96)
constexpr double cos96(int j) {
    constexpr double c[25] = {1.0,
                              0.997858923238603506738,
                              0.991444861373810411145,
                              0.980785280403230449126,
                              0.96592582628906828675,
                              0.946930129495105664256,
                              0.923879532511286756128,
                              0.896872741532688303894,
                              0.866025403784438646764,
                              0.831469612302545237079,
                              0.79335334029123516458,
                              0.751839807478977396408,
                              0.707106781186547524401,
                              0.659345815100068868425,
                              0.608761429008720639416,
                              0.555570233019602224743,
                              0.5,
                              0.442288690219001281995,
                              0.382683432365089771728,
                              0.321439465303161580701,
                              0.258819045102520762349,
                              0.195090322016128267848,
                              0.130526192220051591548,
                              0.0654031292301430668153,
                              0.0};
    j = ((j % 96) + 96) % 96;
    return j <= 24 ? c[j] : j <= 48 ? -c[48 - j] : j <= 72 ? -c[j - 48] : c[96 - j];
}
constexpr double sin96(int j) { return cos96(j - 24); }

// a * exp(SGN * 2 pi i J / 96), quarter turns without multiplies
template <int J, int SGN>
__device__ __forceinline__ double2 rot96(double2 a) {
    constexpr int j = ((J % 96) + 96) % 96;
    if constexpr (j == 0) {
        return a;
    } else if constexpr (j == 48) {
        return make_double2(-a.x, -a.y);
    } else if constexpr ((j == 24 && SGN > 0) || (j == 72 && SGN < 0)) {
        return make_double2(-a.y, a.x);   // * i
    } else if constexpr (j == 24 || j == 72) {
        return make_double2(a.y, -a.x);   // * -i
    } else {
        constexpr double c = cos96(j), s = SGN * sin96(j);
        return make_double2(a.x * c - a.y * s, a.x * s + a.y * c);
    }
}

__device__ __forceinline__ double2 c_add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 c_sub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 c_mul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 c_conj(double2 a) { return make_double2(a.x, -a.y); }

// In-register DFT of P points: y_k = sum_n x_n exp(SGN 2 pi i n k / P), P | 96.
// Composite P = p1 * p2 splits as n = p2 n1 + n2, k = k1 + p1 k2.
template <int P, int SGN>
__device__ __forceinline__ void dft(double2 *x) {
    if constexpr (P == 1) {
    } else if constexpr (P == 2) {
        const double2 a = x[0], b = x[1];
        x[0] = c_add(a, b);
        x[1] = c_sub(a, b);
    } else if constexpr (P == 3) {
        constexpr double h = 0.86602540378443864676 * SGN;   // sin(2 pi / 3)
        const double2 s = c_add(x[1], x[2]), d = c_sub(x[1], x[2]);
        const double2 t = make_double2(x[0].x - 0.5 * s.x, x[0].y - 0.5 * s.y);
        const double2 u = make_double2(-h * d.y, h * d.x);
        x[0] = c_add(x[0], s);
        x[1] = c_add(t, u);
        x[2] = c_sub(t, u);
    } else if constexpr (P == 4) {
        const double2 a = c_add(x[0], x[2]), b = c_sub(x[0], x[2]);
        const double2 c = c_add(x[1], x[3]), d = rot96<24, SGN>(c_sub(x[1], x[3]));
        x[0] = c_add(a, c);
        x[2] = c_sub(a, c);
        x[1] = c_add(b, d);
        x[3] = c_sub(b, d);
    } else {
        static_assert(96 % P == 0, "radix must divide 96");
        constexpr int p1 = (P % 4 == 0) ? 4 : (P % 3 == 0 ? 3 : 2);
        constexpr int p2 = P / p1;
        double2 t[P];
        sfor<p2>([&](auto n2c) {
            constexpr int n2 = decltype(n2c)::value;
            double2 u[p1];
            sfor<p1>([&](auto n1c) { u[decltype(n1c)::value] = x[p2 * decltype(n1c)::value + n2]; });
            dft<p1, SGN>(u);
            sfor<p1>([&](auto k1c) {
                constexpr int k1 = decltype(k1c)::value;
                t[n2 * p1 + k1] = rot96<(96 / P) * n2 * k1, SGN>(u[k1]);
            });
        });
        sfor<p1>([&](auto k1c) {
            constexpr int k1 = decltype(k1c)::value;
            double2 u[p2];
            sfor<p2>([&](auto n2c) { u[decltype(n2c)::value] = t[decltype(n2c)::value * p1 + k1]; });
            dft<p2, SGN>(u);
            sfor<p2>([&](auto k2c) { x[k1 + p1 * decltype(k2c)::value] = u[decltype(k2c)::value]; });
        });
    }
}

// ---- plan ---------------------------------------------------------------------------------
constexpr int wave_R(int M) { return (M % 3 == 0) ? 24 : 16; }
constexpr bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
constexpr bool wave_fft_ok(int M) { return M % wave_R(M) == 0 && is_pow2(M / wave_R(M)) && M / wave_R(M) <= 256; }

struct WPlan {
    int n;
    int r[8];
};
// radices for M points with R points per lane: each radix divides R (so a lane's butterflies
// of a pass read only its own points) and the largest one is taken first
constexpr WPlan wave_plan(int M, int R) {
    WPlan p{0, {}};
    constexpr int cand[8] = {24, 16, 12, 8, 6, 4, 3, 2};
    int rem = M;
    while (rem > 1 && p.n < 8) {
        int pick = 0;
        for (int c : cand)
            if (R % c == 0 && rem % c == 0) { pick = c; break; }
        if (!pick) return WPlan{-1, {}};
        p.r[p.n++] = pick;
        rem /= pick;
    }
    return p;
}
constexpr int plan_p(int M, int R, int S) {   // product of the radices of passes < S
    WPlan w = wave_plan(M, R);
    int p = 1;
    for (int s = 0; s < S; ++s) p *= w.r[s];
    return p;
}

// LDS slot of point j of a sequence: one spare slot every 16 points against bank conflicts
__device__ __forceinline__ int lds_slot(int j) { return j + (j >> 4); }
constexpr int lds_seq_stride(int M) { return M + M / 16 + 1; }
// lds_slot(j0 + STEP i) for j0 + STEP i >= 0.  When 16 divides STEP the spare slots move in
// step with i, the slot is linear in i and a compile-time i folds into the LDS instruction's
// offset (one address register for all of a lane's points).
template <int STEP>
__device__ __forceinline__ int slot_at(int j0, int i) {
    if constexpr (STEP % 16 == 0) return lds_slot(j0) + i * (STEP + STEP / 16);
    else return lds_slot(j0 + STEP * i);
}

// Element at byte offset `off` from a block-uniform base: the access then uses the
// scalar-base + 32-bit-offset address form (one VGPR per access, no 64-bit arithmetic).
template <class E>
__device__ __forceinline__ E &at(E *base, unsigned off) {
    return *(E *)((char *)base + off);
}
template <class E>
__device__ __forceinline__ const E &at(const E *base, unsigned off) {
    return *(const E *)((const char *)base + off);
}

template <int M>
struct WaveGeom {
    static constexpr int R = wave_R(M);
    static constexpr int LQ = M / R;                   // lanes per sequence
    static constexpr int NT = LQ >= 64 ? LQ : 64;      // threads per block
    static constexpr int SPB = NT / LQ;                // sequences (latitudes) per block
    static constexpr WPlan PL = wave_plan(M, R);
    static constexpr size_t lds_bytes = (size_t)SPB * lds_seq_stride(M) * sizeof(double2);
};

// Twiddles w^r = exp(SGN 2 pi i r e / (2M)), 0 < r < P, of a butterfly with base exponent e
// (roots[e] = exp(2 pi i e / (2M))): the power-of-two r are read from the table, the others
// are products of two earlier ones (r = h + (r - h), h the highest power of two in r), at
// most two roundings from the table values.  Fewer loads in flight keeps the passes'
// register peak down.
template <int P, int SGN>
__device__ __forceinline__ void twiddles(double2 (&w)[P], const double2 *__restrict__ roots, int e) {
    sfor<P>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        if constexpr (r > 0) {
            constexpr int h = 1 << (31 - __builtin_clz(r));
            if constexpr (h == r) {
                w[r] = roots[r * e];
                if constexpr (SGN < 0) w[r].y = -w[r].y;
            } else {
                w[r] = c_mul(w[h], w[r - h]);
            }
        }
    });
}

// Passes S.. of the Stockham FFT on the lane's points v[j] = x[l + LQ j].  Pass S (radix P,
// p = product of the earlier radices) runs butterflies i = l + LQ b, b < R/P, on inputs
// x[i + r M/P] = v[b + r R/P], multiplies input r by exp(SGN 2 pi i r k/(pP)), k = i mod p,
// and sends output r to position (i - k) P + k + r p.  On the last pass that position is the
// lane's own v[b + r R/P], so it stays in registers.
template <int M, int SGN, int S>
__device__ __forceinline__ void wave_passes(double2 (&v)[WaveGeom<M>::R], double2 *seq, int l,
                                            const double2 *__restrict__ roots) {
    using G = WaveGeom<M>;
    constexpr int R = G::R, LQ = G::LQ;
    if constexpr (S < G::PL.n) {
        constexpr int P = G::PL.r[S];
        constexpr int p = plan_p(M, R, S);
        constexpr int NB = R / P;
        constexpr bool last = (S + 1 == G::PL.n);
        constexpr int tw_step = 2 * M / (p * P);   // roots hold exp(2 pi i e / (2M))
        sfor<NB>([&](auto bc) {
            constexpr int b = decltype(bc)::value;
            const int i = l + LQ * b;
            const int k = (p == 1) ? 0 : i % p;
            double2 u[P], w[P];
            if constexpr (p > 1) twiddles<P, SGN>(w, roots, k * tw_step);
            sfor<P>([&](auto rc) {
                constexpr int r = decltype(rc)::value;
                u[r] = v[b + r * NB];
                if constexpr (p > 1 && r > 0) u[r] = c_mul(u[r], w[r]);
            });
            dft<P, SGN>(u);
            if constexpr (last) {
                sfor<P>([&](auto rc) { v[b + decltype(rc)::value * NB] = u[decltype(rc)::value]; });
            } else {
                const int base = (i - k) * P + k;
                sfor<P>([&](auto rc) { seq[slot_at<p>(base, decltype(rc)::value)] = u[decltype(rc)::value]; });
            }
        });
        if constexpr (!last) {
            __syncthreads();
            sfor<R>([&](auto jc) { v[decltype(jc)::value] = seq[slot_at<LQ>(l, decltype(jc)::value)]; });
            __syncthreads();
            wave_passes<M, SGN, S + 1>(v, seq, l, roots);
        }
    }
}

// The same passes with each exchange moved through an 8-byte-per-point LDS buffer: real
// parts, then imaginary parts (half the LDS of wave_passes, two more barriers per exchange).
template <int M, int SGN, int S>
__device__ __forceinline__ void wave_passes_split(double2 (&v)[WaveGeom<M>::R], double *seq, int l,
                                                  const double2 *__restrict__ roots) {
    using G = WaveGeom<M>;
    constexpr int R = G::R, LQ = G::LQ;
    if constexpr (S < G::PL.n) {
        constexpr int P = G::PL.r[S];
        constexpr int p = plan_p(M, R, S);
        constexpr int NB = R / P;
        constexpr bool last = (S + 1 == G::PL.n);
        constexpr int tw_step = 2 * M / (p * P);
        // butterflies in place: outputs of butterfly b replace its inputs v[b + r NB]
        sfor<NB>([&](auto bc) {
            constexpr int b = decltype(bc)::value;
            const int i = l + LQ * b;
            const int k = (p == 1) ? 0 : i % p;
            double2 u[P], w[P];
            if constexpr (p > 1) twiddles<P, SGN>(w, roots, k * tw_step);
            sfor<P>([&](auto rc) {
                constexpr int r = decltype(rc)::value;
                u[r] = v[b + r * NB];
                if constexpr (p > 1 && r > 0) u[r] = c_mul(u[r], w[r]);
            });
            dft<P, SGN>(u);
            sfor<P>([&](auto rc) { v[b + decltype(rc)::value * NB] = u[decltype(rc)::value]; });
        });
        if constexpr (!last) {
            // output r of butterfly b goes to point (i - k) P + k + r p
            auto put = [&](auto part) {
                sfor<NB>([&](auto bc) {
                    constexpr int b = decltype(bc)::value;
                    const int i = l + LQ * b;
                    const int k = (p == 1) ? 0 : i % p;
                    const int base = (i - k) * P + k;
                    sfor<P>([&](auto rc) {
                        constexpr int r = decltype(rc)::value;
                        seq[slot_at<p>(base, r)] = decltype(part)::value ? v[b + r * NB].y : v[b + r * NB].x;
                    });
                });
            };
            double nx[R];
            put(std::integral_constant<int, 0>{});
            __syncthreads();
            sfor<R>([&](auto jc) { nx[decltype(jc)::value] = seq[slot_at<LQ>(l, decltype(jc)::value)]; });
            __syncthreads();
            put(std::integral_constant<int, 1>{});
            __syncthreads();
            sfor<R>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                v[j] = make_double2(nx[j], seq[slot_at<LQ>(l, j)]);
            });
            __syncthreads();
            wave_passes_split<M, SGN, S + 1>(v, seq, l, roots);
        }
    }
}

// Blocks are dealt round-robin to the 8 XCDs; give each XCD a contiguous range of latitude
// groups so that the blocks running together on one XCD touch neighbouring latitudes (which
// share cache lines of that XCD's L2 in the theta-contiguous layout).
__device__ __forceinline__ int xcd_group_w(int bid, int G) {
    if (G % 8) return bid;
    return (bid % 8) * (G / 8) + bid / 8;
}

// Synthesis: Fourier coefficients F[f][k][lat] (k <= imlim, zero above) -> real rows
// V[f] at (2n)*sp + lat*stt, (2n+1)*sp + lat*stt.  The packed sequence is
// Z_k = A_k + i w_k B_k with A_k = X_k + conj X_{M-k}, B_k = X_k - conj X_{M-k}, w_k = e^{i pi k/M},
// and z_n = sum_k Z_k e^{2 pi i n k/M} = V[2n] + i V[2n+1].
template <int M, class TV>
__global__ __launch_bounds__(WaveGeom<M>::NT) void fft_synth_wave(const double2 *__restrict__ F, long four_dist,
                                                                  TV *__restrict__ V, long spat_dist,
                                                                  const double2 *__restrict__ roots, int nlat,
                                                                  int imlim, long sp, long stt, int lat_base,
                                                                  int lat_end) {
    using G = WaveGeom<M>;
    constexpr int R = G::R, LQ = G::LQ;
    extern __shared__ double2 lds[];
    const int sq = threadIdx.x / LQ, l = threadIdx.x % LQ;
    const int lat = lat_base + xcd_group_w(blockIdx.x, gridDim.x) * G::SPB + sq;
    const bool live = lat < lat_end;
    const double2 *Ff = F + (size_t)blockIdx.y * four_dist + lat;
    double2 *seq = lds + sq * lds_seq_stride(M);
    const double2 wl = roots[l];   // w_k = w_l * exp(i pi j / R) for k = l + LQ j
    double2 v[R];
    sfor<R>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        const int k = l + LQ * j, km = M - k;
        double2 xk = make_double2(0.0, 0.0), xm = make_double2(0.0, 0.0);
        if (live && k <= imlim) xk = Ff[(size_t)k * nlat];
        if (live && k > 0 && km <= imlim) xm = Ff[(size_t)km * nlat];
        const double2 A = c_add(xk, c_conj(xm)), B = c_sub(xk, c_conj(xm));
        const double2 w = rot96<(96 / (2 * R)) * j, 1>(wl);
        const double2 wB = c_mul(w, B);
        v[j] = make_double2(A.x - wB.y, A.y + wB.x);
    });
    wave_passes<M, 1, 0>(v, seq, l, roots);
    if (live) {
        TV *Vf = V + (size_t)blockIdx.y * spat_dist + (size_t)lat * stt;
        sfor<R>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            const long n = l + LQ * j;
            Vf[(2 * n) * sp] = (TV)v[j].x;
            Vf[(2 * n + 1) * sp] = (TV)v[j].y;
        });
    }
}

// Analysis: real rows -> F[f][k][lat], k <= mmax (zero above imlim).  With Z_k the forward
// DFT (sign -1) of z_n = V[2n] + i V[2n+1]: X_k = A + conj(w_k) B, A = (Z_k + conj Z_{M-k})/2,
// B = (Z_k - conj Z_{M-k})/(2i); Im X_0 = 0.
template <int M, class TV>
__global__ __launch_bounds__(WaveGeom<M>::NT) void fft_anal_wave(const TV *__restrict__ V, long spat_dist,
                                                                 double2 *__restrict__ F, long four_dist,
                                                                 const double2 *__restrict__ roots, int nlat,
                                                                 int mmax, int imlim, long sp, long stt,
                                                                 int lat_base, int lat_end) {
    using G = WaveGeom<M>;
    constexpr int R = G::R, LQ = G::LQ;
    extern __shared__ double2 lds[];
    const int sq = threadIdx.x / LQ, l = threadIdx.x % LQ;
    const int lat = lat_base + xcd_group_w(blockIdx.x, gridDim.x) * G::SPB + sq;
    const bool live = lat < lat_end;
    double2 *seq = lds + sq * lds_seq_stride(M);
    double2 v[R];
    {
        const TV *Vf = V + (size_t)blockIdx.y * spat_dist + (size_t)(live ? lat : 0) * stt;
        sfor<R>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            const long n = l + LQ * j;
            v[j] = live ? make_double2((double)Vf[(2 * n) * sp], (double)Vf[(2 * n + 1) * sp]) : make_double2(0.0, 0.0);
        });
    }
    const double2 wl = roots[l];
    wave_passes<M, -1, 0>(v, seq, l, roots);
    // Z_{M-k} sits in another lane: exchange through LDS
    sfor<R>([&](auto jc) { seq[slot_at<LQ>(l, decltype(jc)::value)] = v[decltype(jc)::value]; });
    __syncthreads();
    if (live) {
        double2 *Ff = F + (size_t)blockIdx.y * four_dist + lat;
        sfor<R>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            const int k = l + LQ * j;
            if (k <= mmax) {
                double2 X = make_double2(0.0, 0.0);
                if (k <= imlim) {
                    const double2 Zk = v[j];
                    const double2 Zm = seq[(j == 0 && l == 0) ? 0 : slot_at<-LQ>(M - l, j)];   // Z_{(M-k) mod M}
                    const double2 A = make_double2(0.5 * (Zk.x + Zm.x), 0.5 * (Zk.y - Zm.y));
                    const double2 D = make_double2(Zk.x - Zm.x, Zk.y + Zm.y);
                    const double2 B = make_double2(0.5 * D.y, -0.5 * D.x);
                    const double2 w = c_conj(rot96<(96 / (2 * R)) * j, 1>(wl));
                    X = c_add(A, c_mul(w, B));
                    if (k == 0) X.y = 0.0;
                }
                Ff[(size_t)k * nlat] = X;
            }
        });
    }
}

// ---- block-cooperative variant for the theta-contiguous layout ---------------------------
// The register kernels above give each latitude its own lanes, so in the theta-contiguous
// layout (consecutive latitudes adjacent in memory) every lane's load or store touches a
// different cache line.  Here a block of NT threads holds T = NT/LQ consecutive latitudes
// and moves data between HBM and its LDS sequence buffers with the latitude index fastest
// across threads (runs of T*16 B of F and T*8 B of V per cache line), then each sequence's
// lanes run the same Stockham passes as above.
//
// Sequence-buffer stride (elements): at least lds_seq_stride(M), and such that the T buffers
// start in different LDS bank groups where the latitude-fastest accesses touch them.  With
// 16-byte elements 8 lanes share an LDS cycle (4 of the 32 banks each), so the T latitudes
// x 8/T rows of one cycle are conflict-free for a stride = 8/T mod 8; with 8-byte elements
// 16 lanes share a cycle (2 banks each) and the stride is 16/T mod 16.
constexpr int coop_stride(int M, int T, int lanes) {
    const int want = T >= lanes ? 1 : lanes / T;
    int s = lds_seq_stride(M);
    while (s % lanes != want % lanes) ++s;
    return s;
}

template <int M, int NTB>
struct CoopGeom {
    using G = WaveGeom<M>;
    static constexpr int LQ = G::LQ, R = G::R;
    static constexpr int NT = NTB;         // threads per block (128 or 256)
    static constexpr int T = NT / LQ;      // latitudes per block
    static constexpr int ROWS = NT / T;    // rows (points) per latitude-fastest access
    static constexpr int STR = coop_stride(M, T, 8);
    static constexpr size_t lds_bytes = (size_t)T * STR * sizeof(double2);
};

// X_k of latitudes lat0 .. lat0+T-1 of one field (Ff = the field's F) for a cooperative
// block: R 16-byte loads per thread with the latitude fastest, all issued before any is used.
template <int M, int NTB>
__device__ __forceinline__ void coop_load_F(double2 (&xr)[WaveGeom<M>::R], const double2 *__restrict__ Ff,
                                            int lat0, int lat_end, int nlat, int imlim) {
    using CG = CoopGeom<M, NTB>;
    constexpr int R = CG::R, T = CG::T, ROWS = CG::ROWS;
    const int tl = threadIdx.x % T, row = threadIdx.x / T;
    const bool live = lat0 + tl < lat_end;
    const unsigned off = 16u * (unsigned)(lat0 + tl + row * nlat), step = 16u * (unsigned)(ROWS * nlat);
    sfor<R>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        xr[i] = (live && row + ROWS * i <= imlim) ? at(Ff, off + i * step) : make_double2(0.0, 0.0);
    });
}

// Points of a cooperative block's latitude-fastest loads into the T LDS sequence buffers.
template <int M, int NTB>
__device__ __forceinline__ void coop_stage(const double2 (&xr)[WaveGeom<M>::R], double2 *lds) {
    using CG = CoopGeom<M, NTB>;
    const int tl = threadIdx.x % CG::T, row = threadIdx.x / CG::T;
    double2 *dst = lds + tl * CG::STR;
    sfor<CG::R>([&](auto ic) { dst[slot_at<CG::ROWS>(row, decltype(ic)::value)] = xr[decltype(ic)::value]; });
}

// Cooperative synthesis of one latitude group whose X_k are staged in LDS (after a barrier):
// packed sequences, passes, z_n out to the field's V (Vf) with the latitude fastest.
template <int M, int NTB>
__device__ __forceinline__ void coop_synth_staged(double2 *lds, double *__restrict__ Vf,
                                                  const double2 *__restrict__ roots, double2 wl, int imlim, long sp,
                                                  long stt, int lat0, int lat_end) {
    using CG = CoopGeom<M, NTB>;
    constexpr int R = CG::R, LQ = CG::LQ, T = CG::T, ROWS = CG::ROWS, STR = CG::STR;
    const int sq = threadIdx.x / LQ, l = threadIdx.x % LQ;
    double2 *seq = lds + sq * STR;
    double2 v[R];
    sfor<R>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        const int k = l + LQ * j, km = M - k;
        const double2 xk = (k <= imlim) ? seq[slot_at<LQ>(l, j)] : make_double2(0.0, 0.0);
        const double2 xm = (k > 0 && km <= imlim) ? seq[slot_at<-LQ>(M - l, j)] : make_double2(0.0, 0.0);
        const double2 A = c_add(xk, c_conj(xm)), B = c_sub(xk, c_conj(xm));
        const double2 w = rot96<(96 / (2 * R)) * j, 1>(wl);
        const double2 wB = c_mul(w, B);
        v[j] = make_double2(A.x - wB.y, A.y + wB.x);
    });
    __syncthreads();
    wave_passes<M, 1, 0>(v, seq, l, roots);
    sfor<R>([&](auto jc) { seq[slot_at<LQ>(l, decltype(jc)::value)] = v[decltype(jc)::value]; });
    __syncthreads();
    const int tl = threadIdx.x % T, row = threadIdx.x / T;
    if (lat0 + tl < lat_end) {
        const double2 *src = lds + tl * STR;
        const unsigned off = 8u * (unsigned)((lat0 + tl) * stt + 2 * row * sp);
        const unsigned step = 8u * (unsigned)(2 * ROWS * sp), odd = 8u * (unsigned)sp;
        sfor<R>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            const double2 z = src[slot_at<ROWS>(row, i)];
            at(Vf, off + i * step) = z.x;
            at(Vf, off + i * step + odd) = z.y;
        });
    }
}

template <int M, int NTB>
__global__ __launch_bounds__(NTB) void fft_synth_coop(const double2 *__restrict__ F, long four_dist,
                                                      double *__restrict__ V, long spat_dist,
                                                      const double2 *__restrict__ roots, int nlat, int imlim, long sp,
                                                      long stt, int lat_base, int lat_end) {
    using CG = CoopGeom<M, NTB>;
    constexpr int R = CG::R, LQ = CG::LQ, T = CG::T;
    extern __shared__ double2 lds[];
    const int lat0 = lat_base + xcd_group_w(blockIdx.x, gridDim.x) * T;
    // X_k of the block's T latitudes into LDS, latitude fastest
    double2 xr[R];
    coop_load_F<M, NTB>(xr, F + (size_t)blockIdx.y * four_dist, lat0, lat_end, nlat, imlim);
    coop_stage<M, NTB>(xr, lds);
    __syncthreads();
    coop_synth_staged<M, NTB>(lds, V + (size_t)blockIdx.y * spat_dist, roots, roots[threadIdx.x % LQ], imlim, sp,
                              stt, lat0, lat_end);
}

// z_n = V[2n] + i V[2n+1] of latitudes lat0 .. lat0+T-1 of one field (Vf = the field's V)
// for a cooperative block: 2R 8-byte loads per thread with the latitude fastest.
template <int M, int NTB>
__device__ __forceinline__ void coop_load_V(double2 (&zr)[WaveGeom<M>::R], const double *__restrict__ Vf,
                                            int lat0, int lat_end, long sp, long stt) {
    using CG = CoopGeom<M, NTB>;
    constexpr int R = CG::R, T = CG::T, ROWS = CG::ROWS;
    const int tl = threadIdx.x % T, row = threadIdx.x / T;
    const bool live = lat0 + tl < lat_end;
    const unsigned off = 8u * (unsigned)((lat0 + tl) * stt + 2 * row * sp);
    const unsigned step = 8u * (unsigned)(2 * ROWS * sp), odd = 8u * (unsigned)sp;
    sfor<R>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        zr[i] = live ? make_double2(at(Vf, off + i * step), at(Vf, off + i * step + odd)) : make_double2(0.0, 0.0);
    });
}

// Cooperative analysis of one latitude group whose z_n are staged in LDS (after a barrier):
// passes, real-FFT post-twiddle, X_k out to the field's F (Ff) with the latitude fastest.
template <int M, int NTB>
__device__ __forceinline__ void coop_anal_staged(double2 *lds, double2 *__restrict__ Ff,
                                                 const double2 *__restrict__ roots, double2 wl, int nlat, int mmax,
                                                 int imlim, int lat0, int lat_end) {
    using CG = CoopGeom<M, NTB>;
    constexpr int R = CG::R, LQ = CG::LQ, T = CG::T, ROWS = CG::ROWS, STR = CG::STR;
    const int sq = threadIdx.x / LQ, l = threadIdx.x % LQ;
    double2 *seq = lds + sq * STR;
    double2 v[R];
    sfor<R>([&](auto jc) { v[decltype(jc)::value] = seq[slot_at<LQ>(l, decltype(jc)::value)]; });
    __syncthreads();
    wave_passes<M, -1, 0>(v, seq, l, roots);
    // Z_n to LDS (Z_{M-k} sits in another lane), X_k in registers, X_k back to LDS
    sfor<R>([&](auto jc) { seq[slot_at<LQ>(l, decltype(jc)::value)] = v[decltype(jc)::value]; });
    __syncthreads();
    sfor<R>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        const int k = l + LQ * j;
        double2 X = make_double2(0.0, 0.0);
        if (k <= imlim) {
            const double2 Zk = v[j];
            const double2 Zm = seq[(j == 0 && l == 0) ? 0 : slot_at<-LQ>(M - l, j)];   // Z_{(M-k) mod M}
            const double2 A = make_double2(0.5 * (Zk.x + Zm.x), 0.5 * (Zk.y - Zm.y));
            const double2 D = make_double2(Zk.x - Zm.x, Zk.y + Zm.y);
            const double2 B = make_double2(0.5 * D.y, -0.5 * D.x);
            const double2 w = c_conj(rot96<(96 / (2 * R)) * j, 1>(wl));
            X = c_add(A, c_mul(w, B));
            if (k == 0) X.y = 0.0;
        }
        v[j] = X;
    });
    __syncthreads();
    sfor<R>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        if (l + LQ * j <= mmax) seq[slot_at<LQ>(l, j)] = v[j];
    });
    __syncthreads();
    const int tl = threadIdx.x % T, row = threadIdx.x / T;
    if (lat0 + tl < lat_end) {
        const double2 *src = lds + tl * STR;
        const unsigned off = 16u * (unsigned)(lat0 + tl + row * nlat), step = 16u * (unsigned)(ROWS * nlat);
        sfor<R>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            if (row + ROWS * i <= mmax) at(Ff, off + i * step) = src[slot_at<ROWS>(row, i)];
        });
    }
}

template <int M, int NTB>
__global__ __launch_bounds__(NTB) void fft_anal_coop(const double *__restrict__ V, long spat_dist,
                                                     double2 *__restrict__ F, long four_dist,
                                                     const double2 *__restrict__ roots, int nlat, int mmax, int imlim,
                                                     long sp, long stt, int lat_base, int lat_end) {
    using CG = CoopGeom<M, NTB>;
    constexpr int R = CG::R, LQ = CG::LQ, T = CG::T;
    extern __shared__ double2 lds[];
    const int lat0 = lat_base + xcd_group_w(blockIdx.x, gridDim.x) * T;
    // z_n of the block's T latitudes into LDS, latitude fastest
    double2 zr[R];
    coop_load_V<M, NTB>(zr, V + (size_t)blockIdx.y * spat_dist, lat0, lat_end, sp, stt);
    coop_stage<M, NTB>(zr, lds);
    __syncthreads();
    coop_anal_staged<M, NTB>(lds, F + (size_t)blockIdx.y * four_dist, roots, roots[threadIdx.x % LQ], nlat, mmax,
                             imlim, lat0, lat_end);
}

// Cooperative kernels with 8-byte LDS staging (real and imaginary parts in turn): T
// latitudes in half the LDS of fft_*_coop (256 threads, M = 1536: T = 4 in 52 KB; 512
// threads: T = 8 in 105 KB).  HBM accesses keep their full width (16-byte F, 8-byte V).
template <int M, int NTB>
struct CoopSGeom {
    using G = WaveGeom<M>;
    static constexpr int LQ = G::LQ, R = G::R;
    static constexpr int NT = NTB;
    static constexpr int T = NT / LQ;
    static constexpr int ROWS = NT / T;
    static constexpr int STR = coop_stride(M, T, 16);   // doubles per sequence buffer
    static constexpr size_t lds_bytes = (size_t)T * STR * sizeof(double);
};

// fpart > 0: the two parts (real, imaginary) of interleaved complex fields.  Field y's part p
// reads F + y four_dist + p fpart and writes V + y spat_dist + p (strides sp, stt already
// those of the doubles of the complex array).  gridDim.x is twice the latitude groups, and
// the part is bit 3 of the block index: blocks b and b + 8 run on the same XCD, one after
// the other, and write the two halves of the same cache lines, which merge in that L2.
template <int M, int NTB>
__global__ __launch_bounds__(NTB) void fft_synth_coops(const double2 *__restrict__ F, long four_dist,
                                                       double *__restrict__ V, long spat_dist,
                                                       const double2 *__restrict__ roots, int nlat, int imlim,
                                                       long sp, long stt, int lat_base, int lat_end, long fpart) {
    using CG = CoopSGeom<M, NTB>;
    constexpr int R = CG::R, LQ = CG::LQ, T = CG::T, STR = CG::STR, ROWS = CG::ROWS;
    extern __shared__ double ldsd[];
    int bx = blockIdx.x, nbx = gridDim.x, part = 0;
    if (fpart > 0) {
        part = (bx >> 3) & 1;
        bx = ((bx >> 4) << 3) | (bx & 7);
        nbx >>= 1;
    }
    const int lat0 = lat_base + xcd_group_w(bx, nbx) * T;
    const int tl = threadIdx.x % T, row = threadIdx.x / T;
    const bool live = lat0 + tl < lat_end;
    double *mine = ldsd + tl * STR;                // staging rows of latitude tl
    const int sq = threadIdx.x / LQ, l = threadIdx.x % LQ;
    double *seq = ldsd + sq * STR;                 // sequence sq
    // X_k of the block's latitudes, 16-byte loads with the latitude fastest
    double2 xr[R];
    {
        const double2 *Ff = F + (size_t)blockIdx.y * four_dist + (size_t)part * fpart;
        const unsigned off = 16u * (unsigned)(lat0 + tl + row * nlat), step = 16u * (unsigned)(ROWS * nlat);
        sfor<R>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            xr[i] = (live && row + ROWS * i <= imlim) ? at(Ff, off + i * step) : make_double2(0.0, 0.0);
        });
    }
    // real parts: keep s = Re X_k + Re X_{M-k}, d = Re X_k - Re X_{M-k}
    sfor<R>([&](auto ic) { mine[slot_at<ROWS>(row, decltype(ic)::value)] = xr[decltype(ic)::value].x; });
    __syncthreads();
    double sr[R], dr[R];
    sfor<R>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        const int k = l + LQ * j, km = M - k;
        const double a = (k <= imlim) ? seq[slot_at<LQ>(l, j)] : 0.0;
        const double b = (k > 0 && km <= imlim) ? seq[slot_at<-LQ>(M - l, j)] : 0.0;
        sr[j] = a + b;
        dr[j] = a - b;
    });
    __syncthreads();
    sfor<R>([&](auto ic) { mine[slot_at<ROWS>(row, decltype(ic)::value)] = xr[decltype(ic)::value].y; });
    __syncthreads();
    // Z_k = A + i w B, A = X_k + conj X_{M-k}, B = X_k - conj X_{M-k}, w = e^{i pi k/M}
    const double2 wl = roots[l];
    double2 v[R];
    sfor<R>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        const int k = l + LQ * j, km = M - k;
        const double a = (k <= imlim) ? seq[slot_at<LQ>(l, j)] : 0.0;
        const double b = (k > 0 && km <= imlim) ? seq[slot_at<-LQ>(M - l, j)] : 0.0;
        const double2 A = make_double2(sr[j], a - b), B = make_double2(dr[j], a + b);
        const double2 w = rot96<(96 / (2 * R)) * j, 1>(wl);
        const double2 wB = c_mul(w, B);
        v[j] = make_double2(A.x - wB.y, A.y + wB.x);
    });
    __syncthreads();
    wave_passes_split<M, 1, 0>(v, seq, l, roots);
    // z_n out: real parts to rows 2n, then imaginary parts to rows 2n+1
    const unsigned off = 8u * (unsigned)((live ? lat0 + tl : 0) * stt + 2 * row * sp);
    const unsigned step = 8u * (unsigned)(2 * ROWS * sp), odd = 8u * (unsigned)sp;
    double *Vf = V + (size_t)blockIdx.y * spat_dist + part;
    sfor<R>([&](auto jc) { seq[slot_at<LQ>(l, decltype(jc)::value)] = v[decltype(jc)::value].x; });
    __syncthreads();
    if (live)
        sfor<R>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            at(Vf, off + i * step) = mine[slot_at<ROWS>(row, i)];
        });
    __syncthreads();
    sfor<R>([&](auto jc) { seq[slot_at<LQ>(l, decltype(jc)::value)] = v[decltype(jc)::value].y; });
    __syncthreads();
    if (live)
        sfor<R>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            at(Vf, off + i * step + odd) = mine[slot_at<ROWS>(row, i)];
        });
}

// With 256 threads (52 KB of LDS at M = 1536) three blocks fit in a CU's LDS; the register
// budget of three waves per SIMD (168 VGPRs) lets them all be resident.
template <int M, int NTB>
__global__ __launch_bounds__(NTB) __attribute__((amdgpu_waves_per_eu(NTB == 256 ? 3 : 1))) void fft_anal_coops(const double *__restrict__ V, long spat_dist,
                                                      double2 *__restrict__ F, long four_dist,
                                                      const double2 *__restrict__ roots, int nlat, int mmax,
                                                      int imlim, long sp, long stt, int lat_base, int lat_end,
                                                      long fpart) {
    using CG = CoopSGeom<M, NTB>;
    constexpr int R = CG::R, LQ = CG::LQ, T = CG::T, STR = CG::STR, ROWS = CG::ROWS;
    extern __shared__ double ldsd[];
    // fpart > 0: the two parts of interleaved complex fields, as in fft_synth_coops (part =
    // block-index bit 3: the blocks reading the two halves of a line run on one XCD together)
    int bx = blockIdx.x, nbx = gridDim.x, part = 0;
    if (fpart > 0) {
        part = (bx >> 3) & 1;
        bx = ((bx >> 4) << 3) | (bx & 7);
        nbx >>= 1;
    }
    const int lat0 = lat_base + xcd_group_w(bx, nbx) * T;
    const int tl = threadIdx.x % T, row = threadIdx.x / T;
    const bool live = lat0 + tl < lat_end;
    double *mine = ldsd + tl * STR;
    const int sq = threadIdx.x / LQ, l = threadIdx.x % LQ;
    double *seq = ldsd + sq * STR;
    // z_n = V[2n] + i V[2n+1] of the block's latitudes, latitude fastest
    double zx[R], zy[R];
    {
        const double *Vf = V + (size_t)blockIdx.y * spat_dist + part;
        const unsigned off = 8u * (unsigned)((live ? lat0 + tl : 0) * stt + 2 * row * sp);
        const unsigned step = 8u * (unsigned)(2 * ROWS * sp), odd = 8u * (unsigned)sp;
        sfor<R>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            zx[i] = live ? at(Vf, off + i * step) : 0.0;
            zy[i] = live ? at(Vf, off + i * step + odd) : 0.0;
        });
    }
    sfor<R>([&](auto ic) { mine[slot_at<ROWS>(row, decltype(ic)::value)] = zx[decltype(ic)::value]; });
    __syncthreads();
    double vx[R];
    sfor<R>([&](auto jc) { vx[decltype(jc)::value] = seq[slot_at<LQ>(l, decltype(jc)::value)]; });
    __syncthreads();
    sfor<R>([&](auto ic) { mine[slot_at<ROWS>(row, decltype(ic)::value)] = zy[decltype(ic)::value]; });
    __syncthreads();
    double2 v[R];
    sfor<R>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        v[j] = make_double2(vx[j], seq[slot_at<LQ>(l, j)]);
    });
    __syncthreads();
    wave_passes_split<M, -1, 0>(v, seq, l, roots);
    // X_k = A + conj(w_k) B with A = (Z_k + conj Z_{M-k})/2, B = (Z_k - conj Z_{M-k})/(2i):
    // Z_{M-k} from LDS, real then imaginary parts
    const double2 wl = roots[l];   // loaded here, not before the passes: their register peak
    double zmx[R];
    sfor<R>([&](auto jc) { seq[slot_at<LQ>(l, decltype(jc)::value)] = v[decltype(jc)::value].x; });
    __syncthreads();
    sfor<R>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        zmx[j] = seq[(j == 0 && l == 0) ? 0 : slot_at<-LQ>(M - l, j)];
    });
    __syncthreads();
    sfor<R>([&](auto jc) { seq[slot_at<LQ>(l, decltype(jc)::value)] = v[decltype(jc)::value].y; });
    __syncthreads();
    sfor<R>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        const int k = l + LQ * j;
        const double2 Zm = make_double2(zmx[j], seq[(j == 0 && l == 0) ? 0 : slot_at<-LQ>(M - l, j)]);
        double2 X = make_double2(0.0, 0.0);
        if (k <= imlim) {
            const double2 Zk = v[j];
            const double2 A = make_double2(0.5 * (Zk.x + Zm.x), 0.5 * (Zk.y - Zm.y));
            const double2 D = make_double2(Zk.x - Zm.x, Zk.y + Zm.y);
            const double2 B = make_double2(0.5 * D.y, -0.5 * D.x);
            const double2 w = c_conj(rot96<(96 / (2 * R)) * j, 1>(wl));
            X = c_add(A, c_mul(w, B));
            if (k == 0) X.y = 0.0;
        }
        v[j] = X;
    });
    __syncthreads();
    // F out with 16-byte stores, latitude fastest: real parts through LDS, then imaginary parts
    double fx[R];
    sfor<R>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        if (l + LQ * j <= mmax) seq[slot_at<LQ>(l, j)] = v[j].x;
    });
    __syncthreads();
    sfor<R>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        fx[i] = (row + ROWS * i <= mmax) ? mine[slot_at<ROWS>(row, i)] : 0.0;
    });
    __syncthreads();
    sfor<R>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        if (l + LQ * j <= mmax) seq[slot_at<LQ>(l, j)] = v[j].y;
    });
    __syncthreads();
    if (live) {
        double2 *Ff = F + (size_t)blockIdx.y * four_dist + (size_t)part * fpart;
        const unsigned off = 16u * (unsigned)(lat0 + tl + row * nlat), step = 16u * (unsigned)(ROWS * nlat);
        sfor<R>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            if (row + ROWS * i <= mmax) at(Ff, off + i * step) = make_double2(fx[i], mine[slot_at<ROWS>(row, i)]);
        });
    }
}

template <class K>
void lds_limit(K kern, size_t lds) {
    ensure_lds_limit((const void *)kern, lds);
}

}  // namespace

#define WAVE_FFT_SIZES(X) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) \
    X(1024) X(1536) X(2048) X(3072) X(4096) X(6144)

bool fft_wave_supported(int M) {
#define OK_CASE(MM) if (M == MM) return true;
    WAVE_FFT_SIZES(OK_CASE)
#undef OK_CASE
    return false;
}

namespace {
template <class TV>
bool synth_wave(const DeviceTables *t, int nlat, int M, int imlim, int nfields, const double2 *F, long four_dist,
                TV *V, long spat_dist, long sp, long st, hipStream_t s, int lat_base, int lat_end) {
#define SYN_CASE(MM)                                                                                          \
    if (M == MM) {                                                                                            \
        static_assert(wave_fft_ok(MM) && WaveGeom<MM>::PL.n > 0, "unsupported wave FFT size");                \
        using G = WaveGeom<MM>;                                                                               \
        lds_limit(fft_synth_wave<MM, TV>, G::lds_bytes);                                                      \
        const int nb = (lat_end - lat_base + G::SPB - 1) / G::SPB;                                           \
        hipLaunchKernelGGL((fft_synth_wave<MM, TV>), dim3(nb, nfields), dim3(G::NT), G::lds_bytes, s, F,      \
                           four_dist, V, spat_dist, t->roots, nlat, imlim, sp, st, lat_base, lat_end);        \
        return true;                                                                                          \
    }
    WAVE_FFT_SIZES(SYN_CASE)
#undef SYN_CASE
    return false;
}

template <class TV>
bool anal_wave(const DeviceTables *t, int nlat, int M, int mmax, int imlim, int nfields, const TV *V, long spat_dist,
               double2 *F, long four_dist, long sp, long st, hipStream_t s, int lat_base, int lat_end) {
#define ANA_CASE(MM)                                                                                          \
    if (M == MM) {                                                                                            \
        using G = WaveGeom<MM>;                                                                               \
        lds_limit(fft_anal_wave<MM, TV>, G::lds_bytes);                                                       \
        const int nb = (lat_end - lat_base + G::SPB - 1) / G::SPB;                                           \
        hipLaunchKernelGGL((fft_anal_wave<MM, TV>), dim3(nb, nfields), dim3(G::NT), G::lds_bytes, s, V,       \
                           spat_dist, F, four_dist, t->roots, nlat, mmax, imlim, sp, st, lat_base, lat_end);  \
        return true;                                                                                          \
    }
    WAVE_FFT_SIZES(ANA_CASE)
#undef ANA_CASE
    return false;
}
}  // namespace

bool launch_fft_synth_wave(const DeviceTables *t, int nlat, int M, int imlim, int nfields, const double2 *F,
                           long four_dist, double *V, long spat_dist, long sp, long st, hipStream_t s, int lat_base,
                           int lat_end) {
    return synth_wave(t, nlat, M, imlim, nfields, F, four_dist, V, spat_dist, sp, st, s, lat_base, lat_end);
}

#define COOP_FFT_SIZES(X) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024) X(1536)

bool fft_coop_supported(int M) {
#define OK_CASE(MM) if (M == MM) return true;
    COOP_FFT_SIZES(OK_CASE)
#undef OK_CASE
    return false;
}

template <int NTB>
bool synth_coop(const DeviceTables *t, int nlat, int M, int imlim, int nfields, const double2 *F, long four_dist,
                double *V, long spat_dist, long sp, long st, hipStream_t s, int lat_base, int lat_end) {
#define SYN_COOP(MM)                                                                                          \
    if (M == MM) {                                                                                            \
        using CG = CoopGeom<MM, NTB>;                                                                         \
        static_assert(CG::LQ <= 64 && CG::T >= 2 && CG::T * CG::LQ == CG::NT, "cooperative FFT shape");      \
        lds_limit(fft_synth_coop<MM, NTB>, CG::lds_bytes);                                                    \
        const int nb = (lat_end - lat_base + CG::T - 1) / CG::T;                                             \
        hipLaunchKernelGGL((fft_synth_coop<MM, NTB>), dim3(nb, nfields), dim3(CG::NT), CG::lds_bytes, s, F,   \
                           four_dist, V, spat_dist, t->roots, nlat, imlim, sp, st, lat_base, lat_end);        \
        return true;                                                                                          \
    }
    COOP_FFT_SIZES(SYN_COOP)
#undef SYN_COOP
    return false;
}

template <int NTB>
bool anal_coop(const DeviceTables *t, int nlat, int M, int mmax, int imlim, int nfields, const double *V,
               long spat_dist, double2 *F, long four_dist, long sp, long st, hipStream_t s, int lat_base, int lat_end) {
#define ANA_COOP(MM)                                                                                          \
    if (M == MM) {                                                                                            \
        using CG = CoopGeom<MM, NTB>;                                                                         \
        lds_limit(fft_anal_coop<MM, NTB>, CG::lds_bytes);                                                     \
        const int nb = (lat_end - lat_base + CG::T - 1) / CG::T;                                             \
        hipLaunchKernelGGL((fft_anal_coop<MM, NTB>), dim3(nb, nfields), dim3(CG::NT), CG::lds_bytes, s, V,    \
                           spat_dist, F, four_dist, t->roots, nlat, mmax, imlim, sp, st, lat_base, lat_end);  \
        return true;                                                                                          \
    }
    COOP_FFT_SIZES(ANA_COOP)
#undef ANA_COOP
    return false;
}

template <int NTB>
bool synth_coops(const DeviceTables *t, int nlat, int M, int imlim, int nfields, const double2 *F, long four_dist,
                 double *V, long spat_dist, long sp, long st, hipStream_t s, int lat_base, int lat_end,
                 long fpart = 0, bool dry = false) {
#define SYN_COOPS(MM)                                                                                         \
    if (M == MM) {                                                                                            \
        using CG = CoopSGeom<MM, NTB>;                                                                        \
        static_assert(CG::LQ <= 64 && CG::T >= 2 && CG::T * CG::LQ == CG::NT, "cooperative FFT shape");      \
        lds_limit(fft_synth_coops<MM, NTB>, CG::lds_bytes);                                                   \
        const int nb = (lat_end - lat_base + CG::T - 1) / CG::T;                                             \
        if (fpart > 0 && nb % 8) return false;   /* the part / XCD decoding needs whole groups of 8 */        \
        if (dry) return true;                                                                                 \
        hipLaunchKernelGGL((fft_synth_coops<MM, NTB>), dim3(fpart > 0 ? 2 * nb : nb, nfields), dim3(CG::NT),  \
                           CG::lds_bytes, s, F, four_dist, V, spat_dist, t->roots, nlat, imlim, sp, st,       \
                           lat_base, lat_end, fpart);                                                         \
        return true;                                                                                          \
    }
    COOP_FFT_SIZES(SYN_COOPS)
#undef SYN_COOPS
    return false;
}

template <int NTB>
bool anal_coops(const DeviceTables *t, int nlat, int M, int mmax, int imlim, int nfields, const double *V,
                long spat_dist, double2 *F, long four_dist, long sp, long st, hipStream_t s, int lat_base, int lat_end,
                long fpart = 0, bool dry = false) {
#define ANA_COOPS(MM)                                                                                         \
    if (M == MM) {                                                                                            \
        using CG = CoopSGeom<MM, NTB>;                                                                        \
        lds_limit(fft_anal_coops<MM, NTB>, CG::lds_bytes);                                                    \
        const int nb = (lat_end - lat_base + CG::T - 1) / CG::T;                                             \
        if (fpart > 0 && nb % 8) return false;   /* the part / XCD decoding needs whole groups of 8 */        \
        if (dry) return true;                                                                                 \
        hipLaunchKernelGGL((fft_anal_coops<MM, NTB>), dim3(fpart > 0 ? 2 * nb : nb, nfields), dim3(CG::NT),   \
                           CG::lds_bytes, s, V, spat_dist, F, four_dist, t->roots, nlat, mmax, imlim, sp, st, \
                           lat_base, lat_end, fpart);                                                         \
        return true;                                                                                          \
    }
    COOP_FFT_SIZES(ANA_COOPS)
#undef ANA_COOPS
    return false;
}

// The cooperative kernels address V and F with 32-bit byte offsets from each field's base:
// false when a field's V (nlat rows of 2M points, strides sp, st) or F (rows k <= kmax) would
// reach 4 GiB.
static bool coop_offsets_fit(int nlat, int M, int kmax, long sp, long st) {
    const double v_bytes = 8.0 * ((double)(nlat - 1) * (double)st + (double)(2 * M - 1) * (double)sp + 2.0);
    const double f_bytes = 16.0 * ((double)kmax + 1.0) * (double)nlat;
    return v_bytes < 4294967296.0 && f_bytes < 4294967296.0;
}

bool launch_fft_synth_coop(const DeviceTables *t, int nlat, int M, int imlim, int nfields, const double2 *F,
                           long four_dist, double *V, long spat_dist, long sp, long st, hipStream_t s, int lat_base,
                           int lat_end, int nthreads) {
    if (!coop_offsets_fit(nlat, M, imlim, sp, st)) return false;
    if (nthreads == -256)
        return synth_coops<256>(t, nlat, M, imlim, nfields, F, four_dist, V, spat_dist, sp, st, s, lat_base, lat_end);
    if (nthreads == -512)
        return synth_coops<512>(t, nlat, M, imlim, nfields, F, four_dist, V, spat_dist, sp, st, s, lat_base, lat_end);
    if (nthreads == 128)
        return synth_coop<128>(t, nlat, M, imlim, nfields, F, four_dist, V, spat_dist, sp, st, s, lat_base, lat_end);
    return synth_coop<256>(t, nlat, M, imlim, nfields, F, four_dist, V, spat_dist, sp, st, s, lat_base, lat_end);
}

bool launch_fft_synth_coops_pairs(const DeviceTables *t, int nlat, int M, int imlim, int nf, const double2 *F,
                                  long fdist, long fpart, double *V, long spat_dist, long sp, long st, hipStream_t s,
                                  int nthreads, bool dry) {
    if (fpart <= 0 || !coop_offsets_fit(nlat, M, imlim, sp, st)) return false;
    if (nthreads == -512)
        return synth_coops<512>(t, nlat, M, imlim, nf, F, fdist, V, spat_dist, sp, st, s, 0, nlat, fpart, dry);
    if (nthreads == -256)
        return synth_coops<256>(t, nlat, M, imlim, nf, F, fdist, V, spat_dist, sp, st, s, 0, nlat, fpart, dry);
    return false;
}

bool launch_fft_anal_coops_pairs(const DeviceTables *t, int nlat, int M, int mmax, int imlim, int nf, const double *V,
                                 long spat_dist, double2 *F, long fdist, long fpart, long sp, long st, hipStream_t s,
                                 int nthreads, bool dry) {
    if (fpart <= 0 || !coop_offsets_fit(nlat, M, std::max(mmax, imlim), sp, st)) return false;
    if (nthreads == -512)
        return anal_coops<512>(t, nlat, M, mmax, imlim, nf, V, spat_dist, F, fdist, sp, st, s, 0, nlat, fpart, dry);
    if (nthreads == -256)
        return anal_coops<256>(t, nlat, M, mmax, imlim, nf, V, spat_dist, F, fdist, sp, st, s, 0, nlat, fpart, dry);
    return false;
}

bool launch_fft_anal_coop(const DeviceTables *t, int nlat, int M, int mmax, int imlim, int nfields, const double *V,
                          long spat_dist, double2 *F, long four_dist, long sp, long st, hipStream_t s, int lat_base,
                          int lat_end, int nthreads) {
    if (!coop_offsets_fit(nlat, M, std::max(mmax, imlim), sp, st)) return false;
    if (nthreads == -256)
        return anal_coops<256>(t, nlat, M, mmax, imlim, nfields, V, spat_dist, F, four_dist, sp, st, s, lat_base,
                               lat_end);
    if (nthreads == -512)
        return anal_coops<512>(t, nlat, M, mmax, imlim, nfields, V, spat_dist, F, four_dist, sp, st, s, lat_base,
                               lat_end);
    if (nthreads == 128)
        return anal_coop<128>(t, nlat, M, mmax, imlim, nfields, V, spat_dist, F, four_dist, sp, st, s, lat_base,
                              lat_end);
    return anal_coop<256>(t, nlat, M, mmax, imlim, nfields, V, spat_dist, F, four_dist, sp, st, s, lat_base, lat_end);
}
bool launch_fft_synth_wave(const DeviceTables *t, int nlat, int M, int imlim, int nfields, const double2 *F,
                           long four_dist, float *V, long spat_dist, long sp, long st, hipStream_t s, int lat_base,
                           int lat_end) {
    return synth_wave(t, nlat, M, imlim, nfields, F, four_dist, V, spat_dist, sp, st, s, lat_base, lat_end);
}
bool launch_fft_anal_wave(const DeviceTables *t, int nlat, int M, int mmax, int imlim, int nfields, const double *V,
                          long spat_dist, double2 *F, long four_dist, long sp, long st, hipStream_t s, int lat_base,
                          int lat_end) {
    return anal_wave(t, nlat, M, mmax, imlim, nfields, V, spat_dist, F, four_dist, sp, st, s, lat_base, lat_end);
}
bool launch_fft_anal_wave(const DeviceTables *t, int nlat, int M, int mmax, int imlim, int nfields, const float *V,
                          long spat_dist, double2 *F, long four_dist, long sp, long st, hipStream_t s, int lat_base,
                          int lat_end) {
    return anal_wave(t, nlat, M, mmax, imlim, nfields, V, spat_dist, F, four_dist, sp, st, s, lat_base, lat_end);
}

}  // namespace shtns_amd
