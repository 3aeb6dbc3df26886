// sht_internal.h -- private configuration and kernel-launch interface of shtns_amd.
//
// The configuration begins with the exact member layout of SHTns's public struct
// (shtns.h:79-94) so callers that read cfg->nlm, cfg->ct, ... through the public header
// keep working; everything after it is private to this library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#define SHTNS_PRIVATE
#include "../../include/shtns.h"
#include "../../include/shtns_hip.h"

namespace shtns_amd {

// Kernel tile constants shared by host and device code.
constexpr int kColsPerTile = 16;        // MFMA N: 4 fields x (E.re, E.im, O.re, O.im)
constexpr int kFieldsPerTile = 4;
constexpr int kStepBlock = 16;          // recurrence steps per analysis tile (MFMA M)

// Mixed-radix plan of the M = nphi/2 point complex FFT (stage s has radix radix[s]).
struct FftPlan {
    int M;
    int nstages;
    int radix[24];
};

// Device tables built once per grid (see build_device_tables in sht_host.cpp).
struct DeviceTables {
    // per latitude pair k < nlat_2
    double *x = nullptr;       // cos(theta_k)
    double *x2 = nullptr;      // cos^2
    double *st = nullptr;      // sin(theta_k)
    double *wq = nullptr;      // Gauss weight w_k * 2pi/nphi (both hemispheres share it)
    // per order im: offsets into the per-step coefficient arrays
    int *step_ofs = nullptr;   // [mmax+2] start of order im in alpha/beta/pre-mix arrays
    double2 *ab = nullptr;     // [nsteps] (alpha_i, beta_i): z_{i+1} = (alpha x^2 + beta) z_i + z_{i-1}
    double4 *pm = nullptr;     // [nsteps] pre/post-mix (ue_i, ue1_i, uo_i, 0):
                               //   P_{m+2i} = ue_i z_i + ue1_{i-1} z_{i-1},  P_{m+2i+1} = x uo_i z_i
    // per (order, latitude pair): recurrence start (scalar functions, and /sin for vectors)
    int *tm = nullptr;         // [mmax+2] polar optimization: latitude pairs k < tm[slot] skipped
    int *i0 = nullptr;         // [(mmax+2) * nlat_2] first step with significant values
    double2 *z0 = nullptr;     // [(mmax+2) * nlat_2] (z_{i0-1}, z_{i0})
    // the same for the vector transforms' functions P/sin(theta) (seeds from sin^(m-1), so
    // finite at the poles); slot mmax+1 (the m=1 functions of vector m=0) as for scalars
    int *i0v = nullptr;
    double2 *z0v = nullptr;
    // sliced (int8) contraction, built on first use: per order slot, super-blocks of 128
    // recurrence steps (sb_ofs[slot] = first, sb_total in all, sb_max in one slot) and the
    // power-of-two scale of the recurrence values per (super-block, latitude pair)
    int *sb_ofs = nullptr;     // [mmax+2]
    int sb_total = 0, sb_max = 0;
    short *eA = nullptr;       // [sb_total * nlat_2]
    // balanced recurrence w_i = 2^D(b, i & 1) z_i in the frame of 32-step block b:
    // w_{i+1} = (alpha' x^2 + beta') w_i + w_{i-1} inside a block; abk[sb_ofs[slot] * 128 + i] =
    // (alpha', beta') (zero-padded to whole super-blocks), dblk[sb_ofs[slot] * 4 + b] =
    // (D(b, 0), D(b, 1)), delta[step_ofs[slot] + i] = D(block of i, i & 1)
    double2 *abk = nullptr;
    int2 *dblk = nullptr;
    // per-(l, m) factor c_lm of a non-orthonormal normalization (null: orthonormal with the
    // Condon-Shortley phase), index lm as the spectra
    double *nrm = nullptr;
    int *delta = nullptr;
    // FFT: roots e^{2 pi i k/nphi}, k < nphi
    double2 *roots = nullptr;
    int *fft_pos = nullptr;    // [M] position of output bin k after the in-place DIF FFT
    FftPlan plan{};
    // vector transforms
};

}  // namespace shtns_amd

struct shtns_info {
    // ---- public prefix: identical to SHTns's shtns.h:79-94 ----
    unsigned int nlm;
    unsigned short lmax, mmax, mres, nlat_2;
    unsigned int nlat, nphi, nspat;
    unsigned short *li, *mi;
    double *ct, *st;
    unsigned int nlat_padded, nlm_cplx;
    // ---- private ----
    int norm;
    int layout;                // SHT_THETA_CONTIGUOUS (native) or SHT_PHI_CONTIGUOUS
    double eps;                // polar threshold given to shtns_set_grid
    std::vector<int> tm;       // [mmax+2] first latitude pair of each order slot (polar optimization)
    std::vector<double> wg;    // Gauss weights (raw, sum 2), nlat_2 values (north half)
    int nsteps_total = 0;      // coefficient slots over all orders (scalar and vector use)
    int device = -1;
    bool gpu_ready = false;
    hipStream_t comp_stream = nullptr;
    hipStream_t xfer_stream = nullptr;
    bool own_streams = false;
    shtns_amd::DeviceTables d;
    // scratch: Fourier buffers, vector conversion buffers (grown on demand)
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    shtns_info *next = nullptr;
    shtns_info *clone_of = nullptr;   // tables borrowed from this config
    // second stream and events used to overlap FFT kernels with Legendre kernels of another
    // chunk (created on first use; each clone has its own)
    hipStream_t aux_stream = nullptr;
    std::vector<hipEvent_t> ev_pool;
    // optional per-kernel timing with HIP events on the compute stream (shtns_profile)
    bool prof = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> prof_ev[8];
    // contraction precision chosen with cushtns_set_precision (-1: FP64 for double data,
    // FP32 for float data)
    int prec = -1;
    // device copies of the host arrays of the host-pointer API (SH_to_spat, ...): slot i is
    // the i-th array a call stages; kept and reused by later calls, grown when too small
    std::vector<std::pair<void *, size_t>> hstage;
    // staging for the FP32-data entry points (FP64 copies of spectra / spatial fields)
    void *xscratch = nullptr;
    size_t xscratch_bytes = 0;
    // kernel shapes picked by cushtns_tune (0: automatic)
    int tune_synth_ncg = 0, tune_anal_nw = 0, tune_anal_ncg = 0, tune_fft_synth = 0, tune_fft_anal = 0;
    // normalization (shtns_create's norm argument): Y^norm_lm = nrm[lm] Y^orthonormal_lm with
    // nrm = sqrt(4 pi) (sht_fourpi) or sqrt(4 pi / (2l+1)) (sht_schmidt), times (-1)^m without
    // the Condon-Shortley phase and sqrt(1/2) for m > 0 with SHT_REAL_NORM; empty for the
    // orthonormal default.  Synthesis uses Q nrm, analysis returns Q / nrm.
    std::vector<double> nrm;
    void *nscratch = nullptr;   // scaled copies of input spectra (grown on demand)
    size_t nscratch_bytes = 0;
    // SHT_LOAD_SAVE_CFG given to shtns_set_grid: kernel shapes are read from / appended to
    // the tuning file; grid_flags = the other flags of that call (part of the file's key)
    bool load_save = false;
    int grid_flags = 0;
};

namespace shtns_amd {

// out[f][lm] = in[f][lm] * nrm[lm] (inv = 0) or / nrm[lm] (inv = 1), f < nf (sht_kernels.hip)
void launch_spec_scale(const double2 *in, long in_dist, double2 *out, long out_dist, int nlm, int nf,
                       const double *nrm, int inv, hipStream_t s);

// tuning file (sht_batch.cpp): apply the entry saved for this configuration and batch size
// (nfields <= 0: the last entry for this configuration, any batch size); true if one was found
bool tune_cache_apply(shtns_cfg c, int nfields);

// kernel launchers (sht_kernels.hip)
struct LegendreArgs {
    const DeviceTables *t;
    int lmax, mmax, mres, nlat, nlat_2, ltr, imlim;
    int nfields;
    int vector_mode;          // 0: scalar functions, 1: Y/sin(theta) (spin-1 V,W spectra)
    double mpos_scale;        // factor on m>0 orders (adjoints use 2 or 0.5)
    int lat_weight;           // synthesis: multiply output by wq_k; analysis: 0 => unit weights
    int prec = 0;             // contraction precision: 0 FP64, 1 FP32, 2 FP32 hi/lo split (sht_legendre.h)
    // kernel shapes (0: automatic choice; set by the autotuner, cushtns_tune)
    int synth_ncg = 0;        // synthesis field groups of 4 per wave: 1, 2 or 4
    int anal_nw = 0;          // analysis waves per block (divides the number of latitude tiles)
    int anal_ncg = 0;         // analysis field groups of 4 per block: 1 or 2
    // latitude-split analysis (anal_splits > 1): per-split partial spectra, part_dist doubles2
    // per field, summed into the output by a second kernel
    double2 *part = nullptr;
    long part_dist = 0;
};

// number of latitude splits of the analysis for these arguments (1: none).  With few fields
// and many latitude tiles an order's tiles are split over several blocks of 12 waves (one
// tile each); the caller then provides part (splits x nfields x part_dist double2).
int anal_splits(const LegendreArgs &a);

// status of the last C-API call of this thread (shtns_last_error); set_error also prints
void set_error(int code, const char *fmt, ...);
void clear_error();

// Contraction precisions of the Legendre kernels (SHT_PREC_* in shtns_hip.h)
constexpr int kPrecF64 = 0, kPrecF32 = 1, kPrecSplit = 2, kPrecSliced = 3;

// Fixed device workspace of cu_SH_to_spat_batch_noalloc (sht_noalloc.hip): one array per
// GPU, kept until the process ends.  noalloc_scratch_setup resolves the array's device
// address on device `dev` (which must be current) once per device; noalloc_scratch_ptr
// returns it (nullptr if the set-up failed or never ran on that device).
constexpr size_t kNoallocScratchBytes = (size_t)64 << 20;
int noalloc_scratch_setup(int dev);
void *noalloc_scratch_ptr(int dev);

// Fourier buffer layout: F[field][im][nlat] complex (north to south).
// qmix: scratch of synth_qmix_doubles(nfields, nsteps_total) doubles (pre-mixed coefficients)
void launch_legendre_synth(const LegendreArgs &a, const double2 *Qlm, long spec_dist, int spec_ltop,
                           double2 *F, long four_dist, double *qmix, int nsteps_total, hipStream_t s);
size_t synth_qmix_doubles(int nfields, int nsteps_total, int want_ncg = 0);
// analysis block shape (waves, latitude tiles per wave) with a compiled kernel
bool anal_shape_ok(int nw, int nlr);
void launch_legendre_anal(const LegendreArgs &a, const double2 *F, long four_dist,
                          double2 *Qlm, long spec_dist, int spec_ltop, hipStream_t s);
// Spatial element (ip, it) of a field sits at ip*sp + it*st (theta-contiguous: sp=nlat, st=1).
// lat_base / nlat_chunk restrict the transform to latitudes [lat_base, lat_base + nlat_chunk)
// (nlat_chunk < 0: all latitudes).
// kind: 0 automatic, 1 register FFT, 2 LDS-staged FFT (where a kernel for the length exists),
// 3 / 4 block-cooperative register FFT with 256 / 128 threads per block, 5 / 6 the same with
// 8-byte LDS staging and 256 / 512 threads.
// es (analysis): element stride of V in doubles (2: one part of interleaved complex values);
// the kernel choice follows the layout's strides sp, st, and the kernels step sp * es, st * es.
void launch_fft_synth(const DeviceTables *t, int nlat, int nphi, int mmax, int imlim, int nfields,
                      const double2 *F, long four_dist, double *V, long spat_dist, long sp, long st, hipStream_t s,
                      int lat_base = 0, int nlat_chunk = -1, int kind = 0);
void launch_fft_anal(const DeviceTables *t, int nlat, int nphi, int mmax, int imlim, int nfields,
                     const double *V, long spat_dist, double2 *F, long four_dist, long sp, long st, hipStream_t s,
                     int lat_base = 0, int nlat_chunk = -1, int kind = 0, int es = 1);
// Synthesis FFTs straight into interleaved complex fields: complex field y's parts p = 0, 1
// from F + y fdist + p fpart into the doubles V + y spat_dist + p (sp, st: the real layout's
// strides).  Only the cooperative split-exchange kernels do this; false (nothing launched)
// when they are not the choice for this size and layout.  dry: only report that.
bool launch_fft_synth_pairs(const DeviceTables *t, int nlat, int nphi, int imlim, int nf, const double2 *F, long fdist,
                            long fpart, double *V, long spat_dist, long sp, long st, hipStream_t s, int kind = 0,
                            bool dry = false);
bool launch_fft_synth_coops_pairs(const DeviceTables *t, int nlat, int M, int imlim, int nf, const double2 *F,
                                  long fdist, long fpart, double *V, long spat_dist, long sp, long st, hipStream_t s,
                                  int nthreads, bool dry);
// The analysis counterpart: complex field y's parts p read from the doubles V + y spat_dist + p
// into F + y fdist + p fpart, one launch.
bool launch_fft_anal_pairs(const DeviceTables *t, int nlat, int nphi, int mmax, int imlim, int nf, const double *V,
                           long spat_dist, double2 *F, long fdist, long fpart, long sp, long st, hipStream_t s,
                           int kind = 0, bool dry = false);
bool launch_fft_anal_coops_pairs(const DeviceTables *t, int nlat, int M, int mmax, int imlim, int nf, const double *V,
                                 long spat_dist, double2 *F, long fdist, long fpart, long sp, long st, hipStream_t s,
                                 int nthreads, bool dry);
// Register-resident FFT kernels (sht_fft.hip) for M = nphi/2 in {16, 24} x 2^k up to 6144,
// latitudes [lat_base, lat_end); return false (nothing launched) for other M.
bool launch_fft_synth_wave(const DeviceTables *t, int nlat, int M, int imlim, int nfields, const double2 *F,
                           long four_dist, double *V, long spat_dist, long sp, long st, hipStream_t s, int lat_base,
                           int lat_end);
bool launch_fft_anal_wave(const DeviceTables *t, int nlat, int M, int mmax, int imlim, int nfields, const double *V,
                          long spat_dist, double2 *F, long four_dist, long sp, long st, hipStream_t s, int lat_base,
                          int lat_end);
// the same for FP32 spatial arrays (transform in FP64, values converted at load / store)
bool launch_fft_synth_wave(const DeviceTables *t, int nlat, int M, int imlim, int nfields, const double2 *F,
                           long four_dist, float *V, long spat_dist, long sp, long st, hipStream_t s, int lat_base,
                           int lat_end);
bool launch_fft_anal_wave(const DeviceTables *t, int nlat, int M, int mmax, int imlim, int nfields, const float *V,
                          long spat_dist, double2 *F, long four_dist, long sp, long st, hipStream_t s, int lat_base,
                          int lat_end);
bool fft_wave_supported(int M);
// block-cooperative register FFT (sht_fft.hip): T = 256/LQ latitudes per block, HBM accesses
// with the latitude fastest (theta-contiguous layout); M = nphi/2 in {32, 48, ..., 1536}
bool fft_coop_supported(int M);
bool launch_fft_synth_coop(const DeviceTables *t, int nlat, int M, int imlim, int nfields, const double2 *F,
                           long four_dist, double *V, long spat_dist, long sp, long st, hipStream_t s, int lat_base,
                           int lat_end, int nthreads = 256);
bool launch_fft_anal_coop(const DeviceTables *t, int nlat, int M, int mmax, int imlim, int nfields, const double *V,
                          long spat_dist, double2 *F, long four_dist, long sp, long st, hipStream_t s, int lat_base,
                          int lat_end, int nthreads = 256);
bool lds_fft_specialized(int M);   // an LDS-staged FFT kernel compiled for this length
// raise a kernel's dynamic LDS limit once (outside any launch sequence that may be captured)
void ensure_lds_limit(const void *kern, size_t lds);
// element-wise FP32 <-> FP64 copies of nfields arrays of n scalars (sht_kernels.hip)
void launch_convert(const float *src, long src_dist, double *dst, long dst_dist, long n, int nfields, hipStream_t s);
void launch_convert(const double *src, long src_dist, float *dst, long dst_dist, long n, int nfields, hipStream_t s);
// complex-valued fields: spectra split into / merged from the real and imaginary parts'
// real-field spectra (nf fields; field f's real-part / imaginary-part real fields are entries
// 2f and 2f+1 of a batch), and the spatial interleave of the synthesis (the analysis FFTs
// read the interleaved complex arrays directly)
void launch_cplx_split(const double2 *a, long a_dist, double2 *XY, long nlm, int lmax, int mmax, int nf, hipStream_t s);
void launch_cplx_merge(const double2 *XY, long nlm, double2 *a, long a_dist, int lmax, int mmax, int nf, hipStream_t s);
void launch_interleave(const double *xy, double2 *z, long z_dist, long n, int nf, hipStream_t s);
// Legendre transforms at one order (no FFT): table of P_lm at every latitude, sum over degrees,
// weighted sum over latitudes
// sliced int8 synthesis (sht_legendre_i8.hip): row-exponent table, scratch sizes, launch
// (qmix from launch_premix with synth_ncg = 2: 32 columns per 8 fields)
void launch_i8_rowexp(const DeviceTables *t, int mmax, int nlat_2, hipStream_t s);
size_t i8_planes_bytes(int nfields, int sb_total);
size_t i8_eB_bytes(int nfields, int sb_total);
// LDS of the sliced synthesis kernel (sums and one order's recurrence coefficients); the
// sliced path runs when this is within kI8MaxLds
size_t i8_synth_lds_bytes(int sb_max);
constexpr size_t kI8MaxLds = 80 * 1024;
void launch_legendre_synth_i8(const LegendreArgs &a, int spec_ltop, double2 *F, long four_dist, const double *qmix,
                              int nsteps_total, unsigned char *planes, int *eB, hipStream_t s);
// ktm: latitudes k < ktm and k >= nlat - ktm get zero tables (polar optimization of the order)
void launch_ml_table(const double *ct, const double *st, int nlat, int m, int nl, const double2 *ab, double log2_seed,
                     double seed_sign, int ktm, double *T, hipStream_t s);
void launch_ml_synth(const double *T, int nlat, int nl, const double2 *Q, int real_only, double2 *V, hipStream_t s);
void launch_ml_anal(const double *T, const double *w, int nlat, int nl, const double2 *V, int real_only, double scale,
                    double2 *Q, hipStream_t s);
void launch_ml_vtable(const double *ct, const double *st, int nlat, int m, int nl, const double2 *ab, double log2_seed,
                      double seed_sign, int ktm, double *Y, double *D, hipStream_t s);
void launch_ml_vsynth(const double *Y, const double *D, int nlat, int nl, int m, const double2 *S, const double2 *T,
                      double2 *Vt, double2 *Vp, hipStream_t s);
void launch_ml_vanal(const double *Y, const double *D, const double *w, int nlat, int nl, int m, const double2 *Vt,
                     const double2 *Vp, double2 *S, double2 *T, hipStream_t s);
// the two halves of launch_legendre_synth: coefficient pre-mix, then the Legendre kernel for
// latitude tiles [tile0, tile0 + ntl) (64 latitude pairs per tile)
void launch_premix(const LegendreArgs &a, const double2 *Qlm, long spec_dist, int spec_ltop, double *qmix,
                   int nsteps_total, hipStream_t s);
void launch_legendre_synth_tiles(const LegendreArgs &a, int spec_ltop, double2 *F, long four_dist, const double *qmix,
                                 int nsteps_total, int tile0, int ntl, hipStream_t s);
// per-precision instantiations (sht_legendre_f32.hip, sht_legendre_split.hip)
void launch_legendre_synth_tiles_f32(const LegendreArgs &a, int spec_ltop, double2 *F, long four_dist,
                                     const double *qmix, int nsteps_total, int tile0, int ntl, hipStream_t s);
void launch_legendre_synth_tiles_split(const LegendreArgs &a, int spec_ltop, double2 *F, long four_dist,
                                       const double *qmix, int nsteps_total, int tile0, int ntl, hipStream_t s);
void launch_legendre_anal_f32(const LegendreArgs &a, const double2 *F, long four_dist, double2 *Qlm, long spec_dist,
                              int spec_ltop, hipStream_t s);
void launch_legendre_anal_split(const LegendreArgs &a, const double2 *F, long four_dist, double2 *Qlm,
                                long spec_dist, int spec_ltop, hipStream_t s);
// seeds of the scalar functions (vec = 0: t->i0, t->z0) or of the vector ones (vec = 1)
void launch_seed_tables(DeviceTables *t, int lmax, int mmax, int mres, int nlat_2, double thr, hipStream_t s,
                        int vec = 0);
void launch_sphtor_to_vw(const double2 *S, const double2 *T, long spec_dist, double2 *VW, long vw_dist,
                         int lmax, int mmax, int mres, int ltr, int imlim, int nfields, hipStream_t s);
void launch_vw_to_sphtor(const double2 *VW, long vw_dist, double2 *S, double2 *T, long spec_dist,
                         int lmax, int mmax, int mres, int ltr, int imlim, int nfields, hipStream_t s);

}  // namespace shtns_amd
