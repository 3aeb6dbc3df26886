/*
 * shtns_hip.h -- device-pointer transforms of shtns_amd (stands in for SHTns 3.5.2's
 * shtns_cuda.h; hipStream_t replaces cudaStream_t, both opaque stream handles).
 *
 * All transforms take device pointers, enqueue their kernels on the configuration's
 * compute stream and return without synchronizing.  As in SHTns, one configuration must
 * not be used from two host threads at the same time (shtns_cuda.h:28-31); clone it.
 * Layouts are SHTns's: spectral cplx[nlm] (LM order), spatial double[nphi][nlat]
 * (theta-contiguous, north first).  Batched entry points process nfields independent
 * fields whose arrays sit dist elements apart.
 */
#ifndef SHTNS_AMD_SHTNS_HIP_H
#define SHTNS_AMD_SHTNS_HIP_H

#include <hip/hip_runtime_api.h>

#include "shtns.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- transforms on device memory (shtns_cuda.h:38-52) ---- */
void cu_spat_to_SH(shtns_cfg shtns, double *Vr, cplx *Qlm, int ltr);                       /* :38 */
void cu_spat_to_SH_float(shtns_cfg shtns, float *Vr, cplx_f *Qlm, int ltr);                /* :40 */
void cu_SH_to_spat(shtns_cfg shtns, cplx *Qlm, double *Vr, int ltr);                       /* :42 */
void cu_SH_to_spat_float(shtns_cfg shtns, cplx_f *Qlm, float *Vr, int ltr);                /* :44 */
void cu_spat_to_SHsphtor(shtns_cfg, double *Vt, double *Vp, cplx *Slm, cplx *Tlm, int ltr); /* :46 */
void cu_SHsphtor_to_spat(shtns_cfg, cplx *Slm, cplx *Tlm, double *Vt, double *Vp, int ltr); /* :48 */
void cu_spat_to_SHqst(shtns_cfg, double *Vr, double *Vt, double *Vp, cplx *Qlm, cplx *Slm, cplx *Tlm,
                      int ltr);                                                             /* :50 */
void cu_SHqst_to_spat(shtns_cfg, cplx *Qlm, cplx *Slm, cplx *Tlm, double *Vr, double *Vt, double *Vp,
                      int ltr);                                                             /* :52 */

/* Complex-valued scalar and 2D vector fields on device arrays (additive: SHTns's complex
 * transforms, sht_func.c:943-1200, run on the CPU only).  Same layouts as SH_to_spat_cplx /
 * SHsphtor_to_spat_cplx and their inverses; asynchronous on the config's compute stream;
 * mres = 1. */
void cu_SH_to_spat_cplx(shtns_cfg, cplx *alm, cplx *z);
void cu_spat_cplx_to_SH(shtns_cfg, cplx *z, cplx *alm);
void cu_SHsphtor_to_spat_cplx(shtns_cfg, cplx *Slm, cplx *Tlm, cplx *Vt, cplx *Vp);
void cu_spat_cplx_to_SHsphtor(shtns_cfg, cplx *Vt, cplx *Vp, cplx *Slm, cplx *Tlm);
/* batched forms: field f at alm + f*spec_dist (spec_dist >= nlm_cplx) and z + f*spat_dist
 * (spat_dist >= nspat), in complex elements */
void cu_SH_to_spat_cplx_batch(shtns_cfg, cplx *alm, cplx *z, int nfields, long spec_dist, long spat_dist);
void cu_spat_cplx_to_SH_batch(shtns_cfg, cplx *z, cplx *alm, int nfields, long spat_dist, long spec_dist);
void cu_SHsphtor_to_spat_cplx_batch(shtns_cfg, cplx *Slm, cplx *Tlm, cplx *Vt, cplx *Vp, int nfields, long spec_dist,
                                    long spat_dist);
void cu_spat_cplx_to_SHsphtor_batch(shtns_cfg, cplx *Vt, cplx *Vp, cplx *Slm, cplx *Tlm, int nfields, long spat_dist,
                                    long spec_dist);

/* ---- initialization (shtns_cuda.h:63-72, sht_gpu.cu:227-336) ---- */
shtns_cfg cushtns_clone(shtns_cfg shtns, hipStream_t compute_stream, hipStream_t transfer_stream); /* :63 */
int cushtns_init_gpu(shtns_cfg shtns);                                                     /* :69 */
void cushtns_set_streams(shtns_cfg shtns, hipStream_t compute_stream, hipStream_t transfer_stream); /* :72 */
void cushtns_release_gpu(shtns_cfg shtns);                                                 /* sht_gpu.cu:137 */

/* ---- batched transforms (additive: SHTns batches only inside its kernels, NFIELDS<=2,
 *      sht_gpu.cu:477-493).  Field f reads/writes Qlm + f*spec_dist and Vr + f*spat_dist
 *      (in elements of cplx / double).  spec_dist >= nlm, spat_dist >= nspat. ---- */
void cu_SH_to_spat_batch(shtns_cfg, cplx *Qlm, double *Vr, int ltr, int nfields, long spec_dist, long spat_dist);
void cu_spat_to_SH_batch(shtns_cfg, double *Vr, cplx *Qlm, int ltr, int nfields, long spat_dist, long spec_dist);
void cu_SHsphtor_to_spat_batch(shtns_cfg, cplx *Slm, cplx *Tlm, double *Vt, double *Vp, int ltr,
                               int nfields, long spec_dist, long spat_dist);
void cu_spat_to_SHsphtor_batch(shtns_cfg, double *Vt, double *Vp, cplx *Slm, cplx *Tlm, int ltr,
                               int nfields, long spat_dist, long spec_dist);

/* No counterpart in SHTns.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): cu_SH_to_spat_batch
 * with its work arrays (Fourier rows and pre-mixed coefficients) in a fixed 64 MiB device
 * workspace that the library holds on each GPU until the process ends.  It allocates no
 * device memory and creates no stream or event; the configuration must already hold its GPU
 * tables (shtns_set_grid on a GPU, or cushtns_init_gpu).  A call needing more than 64 MiB
 * (e.g. more than one field at lmax = 1023, nlat = 1536) fails with shtns_last_error() = 2
 * and does nothing.  All calls on one GPU share the workspace: calls that overlap on one GPU
 * (other streams, other threads) race and must be serialized by the caller. */
void cu_SH_to_spat_batch_noalloc(shtns_cfg, cplx *Qlm, double *Vr, int ltr, int nfields, long spec_dist,
                                 long spat_dist);

/* Single-precision data, batched (additive; the unbatched _float calls above are SHTns's).
 * Arrays are FP32; the recurrence, seeds and Fourier transform run in FP64 and the
 * degree/latitude contraction at the configuration's precision (cushtns_set_precision;
 * default SHT_PREC_F32 for these calls). */
void cu_SH_to_spat_float_batch(shtns_cfg, cplx_f *Qlm, float *Vr, int ltr, int nfields, long spec_dist, long spat_dist);
void cu_spat_to_SH_float_batch(shtns_cfg, float *Vr, cplx_f *Qlm, int ltr, int nfields, long spat_dist, long spec_dist);

/* Precision of the Legendre contraction (additive, no SHTns counterpart).  The recurrence,
 * its seeds and tables, the Gauss nodes and weights, the FFT and the final sums stay FP64.
 *   SHT_PREC_F64    FP64 matrix cores (default for double data);
 *   SHT_PREC_F32    recurrence values and coefficients rounded to FP32, FP32 matrix cores,
 *                   FP32 partial sums over 16 degree pairs accumulated in FP64 (default for
 *                   float data);
 *   SHT_PREC_SPLIT  both operands as FP32 hi+lo pairs, three FP32 matrix-core products
 *                   (hi*hi + lo*hi + hi*lo), same accumulation.
 *   SHT_PREC_SLICED both operands as five signed 8-bit digits with power-of-two scales per
 *                   latitude / column and per 128 recurrence steps, exact int32 sums of the
 *                   19 leading digit products on the int8 matrix cores (scalar synthesis of
 *                   double data; other transforms run the FP64 contraction in this mode).
 * SHT_PREC_AUTO restores the defaults.  Returns 0, or -1 for an unknown value. */
#define SHT_PREC_AUTO (-1)
#define SHT_PREC_F64 0
#define SHT_PREC_F32 1
#define SHT_PREC_SPLIT 2
#define SHT_PREC_SLICED 3
int cushtns_set_precision(shtns_cfg, int prec);

/* Autotuner (additive, no SHTns counterpart).  Times the candidate kernel shapes of this
 * configuration on a batch of nfields fields -- synthesis field groups per wave, analysis
 * waves per block and field groups per block, register or LDS-staged FFT -- and keeps the
 * fastest for later calls on this configuration (clones made afterwards inherit them).
 * Allocates and frees temporary device buffers; synchronizes the compute stream.
 * cushtns_tuning reads the five choices (0 = automatic):
 * {synthesis groups, analysis waves, analysis groups, synthesis FFT, analysis FFT}, with FFT
 * 1 = register kernel, 2 = LDS-staged kernel, 3 / 4 = block-cooperative register kernel with
 * 256 / 128 threads per block, 5 / 6 = the same with 8-byte LDS staging, 256 / 512 threads.
 * Both return 0, or -1 on error. */
int cushtns_tune(shtns_cfg, int nfields);
int cushtns_tuning(shtns_cfg, int *choices);
/* Set the five choices by hand (same meaning; a choice without a compiled kernel for this
 * size falls back to the automatic one).  Returns 0, or -1 for a value out of range. */
int cushtns_set_tuning(shtns_cfg, const int *choices);

/* Uneven batches (additive, no SHTns counterpart): n scalar transforms, transform i on
 * configuration cfgs[i] (any mix of lmax / grids, grids already set).  The batch owns the
 * device arrays: transforms sharing a configuration are stored contiguously and run as one
 * batched launch, and the groups of different configurations run as parallel branches of
 * one HIP graph, captured when the batch is created and replayed by each call.
 *   prec  contraction precision of every transform (SHT_PREC_*);
 *   tune  nonzero: run cushtns_tune for each group's configuration and batch size first.
 * The batch clones the configurations (their streams and tuning are not changed).
 * cushtns_batch_arrays gives the device arrays of transform i: Qlm[nlm] (spectral) and
 * Vr[nspat] (spatial, that configuration's layout).  The transforms read and write them;
 * the calls enqueue the graph on `stream` and return without synchronizing. */
typedef struct shtns_batch_info *shtns_batch;
shtns_batch cushtns_batch_create(int n, const shtns_cfg *cfgs, int prec, int tune);
int cushtns_batch_arrays(shtns_batch, int i, cplx **Qlm, double **Vr);
int cushtns_batch_SH_to_spat(shtns_batch, hipStream_t stream);
int cushtns_batch_spat_to_SH(shtns_batch, hipStream_t stream);
void cushtns_batch_destroy(shtns_batch);

/* Adjoint (transpose) operators, for reverse-mode differentiation (additive).
 *   cu_SH_to_spat_adjoint:  Qlm = S^T(Vr)   with S the synthesis operator of SH_to_spat
 *   cu_spat_to_SH_adjoint:  Vr  = A^T(Qlm)  with A the analysis operator of spat_to_SH
 * for the real inner products <Q,Q'> = sum Re(conj(Q) Q') over stored (l,m) and
 * <V,V'> = sum V V' over grid points.  Batched like the functions above. */
void cu_SH_to_spat_adjoint_batch(shtns_cfg, double *Vr, cplx *Qlm, int ltr, int nfields, long spat_dist, long spec_dist);
void cu_spat_to_SH_adjoint_batch(shtns_cfg, cplx *Qlm, double *Vr, int ltr, int nfields, long spec_dist, long spat_dist);

/* Per-kernel timing with HIP events on the configuration's compute stream (additive, no
 * SHTns counterpart; used by the benchmark).  shtns_profile(cfg, 1) clears and enables;
 * shtns_profile_get fills ms[k] (total milliseconds) and count[k] (launches) for
 * k = 0 Legendre synthesis, 1 Legendre analysis, 2 FFT synthesis, 3 FFT analysis,
 * 4 vector conversions, waiting for the recorded events; returns the number of kinds (5). */
void shtns_profile(shtns_cfg, int enable);
int shtns_profile_get(shtns_cfg, double *ms, int *count);

#ifdef __cplusplus
}
#endif
#endif
