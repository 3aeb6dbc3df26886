/*
 * shtns.h -- public C API of shtns_amd, a spherical harmonic transform engine for one
 * AMD MI355X (gfx950).  It exports the same names, signatures, flag values and public
 * struct prefix as SHTns 3.5.2's shtns.h so existing callers link against it unchanged.
 * Each declaration names the SHTns interface it stands in for (file:line of shtns.h
 * in SHTns 3.5.2).
 *
 * Scope: Gauss and regular (Fejer) grids; the scalar transforms SH_to_spat/spat_to_SH, the
 * 2D vector transforms SHsphtor_to_spat/spat_to_SHsphtor and the 3-component transforms
 * SHqst_to_spat/spat_to_SHqst (plus their _l truncated forms) and the complex-valued
 * scalar and 2D vector transforms (SH_to_spat_cplx, SHsphtor_to_spat_cplx and their
 * inverses), computed on the GPU by
 * hand-written HIP kernels.  Host-pointer transforms copy to the device,
 * run there and copy back; the device-pointer variants are in shtns_hip.h.
 */
#ifndef SHTNS_AMD_SHTNS_H
#define SHTNS_AMD_SHTNS_H

#include <stddef.h>

#ifdef __cplusplus
#include <complex>
typedef std::complex<double> cplx;   /* shtns.h:22-30: 2 doubles, re then im */
typedef std::complex<float> cplx_f;
extern "C" {
#else
#include <complex.h>
typedef complex double cplx;
typedef complex float cplx_f;
#endif

#define SHTNS_INTERFACE 0x30500                 /* shtns.h:38 */

typedef struct shtns_info *shtns_cfg;           /* shtns.h:41 */

/* normalizations (shtns.h:48-55).  Only sht_orthonormal is implemented here; the
   others are accepted by shtns_create and rejected with an error message. */
enum shtns_norm { sht_orthonormal, sht_fourpi, sht_schmidt, sht_for_rotations };
#define SHT_NO_CS_PHASE (256 * 4)
#define SHT_REAL_NORM (256 * 8)

/* grid / algorithm selectors (shtns.h:58-75).  sht_gauss, sht_auto, sht_quick_init and
   sht_gauss_fly select the Gauss grid; sht_reg_fast and sht_reg_dct the regular grid without
   poles (equispaced theta, Fejer quadrature).  sht_reg_poles is not implemented. */
enum shtns_type { sht_gauss, sht_auto, sht_reg_fast, sht_reg_dct, sht_quick_init, sht_reg_poles, sht_gauss_fly };
#define SHT_NATIVE_LAYOUT 0
#define SHT_THETA_CONTIGUOUS 256
#define SHT_PHI_CONTIGUOUS (256 * 2)
#define SHT_SOUTH_POLE_FIRST (256 * 32)
#define SHT_SCALAR_ONLY (256 * 16)
#define SHT_LOAD_SAVE_CFG (256 * 64)
#define SHT_ALLOW_GPU (256 * 128)
#define SHT_ALLOW_PADDING (256 * 256)

#ifndef SHTNS_PRIVATE
/* Read-only public prefix of the configuration (shtns.h:79-94), same member order and types. */
struct shtns_info {
    const unsigned int nlm;
    const unsigned short lmax;
    const unsigned short mmax;
    const unsigned short mres;
    const unsigned short nlat_2;
    const unsigned int nlat;
    const unsigned int nphi;
    const unsigned int nspat;
    const unsigned short *const li;
    const unsigned short *const mi;
    const double *const ct;
    const double *const st;
    const unsigned int nlat_padded;
    const unsigned int nlm_cplx;
};
#endif

/* index macros (shtns.h:104-106, 120) */
#define LiM(shtns, l, im) ((((im) * (2 * shtns->lmax + 2 - ((im) + 1) * shtns->mres)) >> 1) + (l))
#define LM(shtns, l, m) ((((((unsigned short)(m)) / shtns->mres) * (2 * shtns->lmax + 2 - ((m) + shtns->mres))) >> 1) + (l))
/* index of (l, m), -l <= m <= l, in a complex-valued field's spectrum, mres = 1 (shtns.h:114) */
#define LM_cplx(shtns, l, m) \
    (((l) <= shtns->mmax) ? (l) * ((l) + 1) + (m) : shtns->mmax * (2 * (l) - shtns->mmax) + (l) + (m))
/* loops over every (l, m) of a spectrum: lm (declared by the macro) is the index; LM_L_LOOP
   also declares the degree l (shtns.h:107-112) */
#define LM_LOOP(shtns, action) { for (int lm = 0; lm < (int)shtns->nlm; ++lm) { action } }
#define LM_L_LOOP(shtns, action) { for (int lm = 0; lm < (int)shtns->nlm; ++lm) { int l = shtns->li[lm]; action } }
#define NSPAT_ALLOC(shtns) (shtns->nspat)
#define PHI_DEG(shtns, ip) (360. / (shtns->nphi * shtns->mres)) * (ip)
#define PHI_RAD(shtns, ip) (2. * M_PI / (shtns->nphi * shtns->mres)) * (ip)

/* sizes and information (shtns.h:131-142, 146) */
long nlm_calc(long lmax, long mmax, long mres);                 /* shtns.h:131 */
long nlm_cplx_calc(long lmax, long mmax, long mres);            /* shtns.h:134 */
void shtns_verbose(int);                                        /* shtns.h:136 */
void shtns_print_version(void);                                 /* shtns.h:137 */
const char *shtns_get_build_info(void);                         /* shtns.h:138 */
void shtns_print_cfg(shtns_cfg);                                /* shtns.h:142 */

/* initialization (shtns.h:147-177) */
shtns_cfg shtns_init(enum shtns_type flags, int lmax, int mmax, int mres, int nlat, int nphi);   /* :149 */
shtns_cfg shtns_create(int lmax, int mmax, int mres, enum shtns_norm norm);                     /* :151 */
int shtns_set_grid(shtns_cfg, enum shtns_type flags, double eps, int nlat, int nphi);           /* :153 */
int shtns_set_grid_auto(shtns_cfg, enum shtns_type flags, double eps, int nl_order, int *nlat, int *nphi); /* :155 */
int shtns_use_threads(int num_threads);                         /* :159 */
int shtns_use_gpu(int device_id);                               /* :161 */
void shtns_reset(void);                                         /* :163 */
void shtns_destroy(shtns_cfg);                                  /* :164 */
void shtns_unset_grid(shtns_cfg);                               /* :165 */
void *shtns_malloc(size_t bytes);                               /* :170 */
void shtns_free(void *p);                                       /* :171 */

/* special values (shtns.h:178-186) */
double sh00_1(shtns_cfg);                                       /* :180 */
double sh10_ct(shtns_cfg);                                      /* :181 */
double sh11_st(shtns_cfg);                                      /* :182 */
double shlm_e1(shtns_cfg, int l, int m);                        /* :183 */
int shtns_gauss_wts(shtns_cfg, double *wts);                    /* :185 */

/* Legendre functions (shtns.h:227-228) */
int legendre_sphPlm_array(shtns_cfg shtns, const int lmax, const int im, const double x, double *yl);
/* m = 0: P_l0 and dP_l0/dtheta; m > 0: P_lm/sin(theta) and dP_lm/dtheta (shtns.h:233) */
int legendre_sphPlm_deriv_array(shtns_cfg shtns, const int lmax, const int im, const double x, const double sint,
                                double *yl, double *dyl);

/* operators coupling degrees l-1 and l+1, mx of 2*nlm doubles (shtns.h:240-245) */
void mul_ct_matrix(shtns_cfg, double *mx);
void st_dt_matrix(shtns_cfg, double *mx);
void SH_mul_mx(shtns_cfg, double *mx, cplx *Qlm, cplx *Rlm);

/* scalar transforms (shtns.h:262-267).  spat_to_SH leaves Vr unchanged here
   (SHTns only promises nothing about it). */
void spat_to_SH(shtns_cfg shtns, double *Vr, cplx *Qlm);
void SH_to_spat(shtns_cfg shtns, cplx *Qlm, double *Vr);

/* Legendre transforms at order m = im*mres, no FFT (shtns.h:329-334): Vr has nlat complex
   values (grid latitudes), Ql lmax+1-m (degrees m..lmax; spat_to_SH_ml writes 0 above ltr).
   V_k = sum_l Q_l P_lm(x_k);  Q_l = 2 pi sum_k w_k P_lm(x_k) V_k.  At m = 0 only real parts. */
void spat_to_SH_ml(shtns_cfg, int im, cplx *Vr, cplx *Ql, int ltr);
void SH_to_spat_ml(shtns_cfg, int im, cplx *Ql, cplx *Vr, int ltr);
/* vector forms (shtns.h:336-337): Vt_k = sum_l S_l dP/dtheta + T_l (i m/sin) P,
   Vp_k = sum_l S_l (i m/sin) P - T_l dP/dtheta; the analysis is the matching Gauss projection
   divided by l (l+1), times 2 pi */
void spat_to_SHsphtor_ml(shtns_cfg, int im, cplx *Vt, cplx *Vp, cplx *Sl, cplx *Tl, int ltr);
void SHsphtor_to_spat_ml(shtns_cfg, int im, cplx *Sl, cplx *Tl, cplx *Vt, cplx *Vp, int ltr);
void SHsph_to_spat_ml(shtns_cfg, int im, cplx *Sl, cplx *Vt, cplx *Vp, int ltr);       /* :338 */
void SHtor_to_spat_ml(shtns_cfg, int im, cplx *Tl, cplx *Vt, cplx *Vp, int ltr);       /* :339 */
void spat_to_SHqst_ml(shtns_cfg, int im, cplx *Vr, cplx *Vt, cplx *Vp, cplx *Ql, cplx *Sl, cplx *Tl, int ltr); /* :341 */
void SHqst_to_spat_ml(shtns_cfg, int im, cplx *Ql, cplx *Sl, cplx *Tl, cplx *Vr, cplx *Vt, cplx *Vp, int ltr); /* :342 */

/* complex-valued scalar fields, mres = 1 (shtns.h:272, 277).  alm[LM_cplx(l, m)] holds
   -l <= m <= l, nlm_cplx entries; z has the real transforms' layout with complex elements. */
void SH_to_spat_cplx(shtns_cfg shtns, cplx *alm, cplx *z);
void spat_cplx_to_SH(shtns_cfg shtns, cplx *z, cplx *alm);
/* complex-valued 2D vector fields, mres = 1 (shtns.h:292, 294): Slm, Tlm in the LM_cplx
   layout; Vt, Vp complex with the real transforms' layout */
void spat_cplx_to_SHsphtor(shtns_cfg, cplx *Vt, cplx *Vp, cplx *Slm, cplx *Tlm);
void SHsphtor_to_spat_cplx(shtns_cfg, cplx *Slm, cplx *Tlm, cplx *Vt, cplx *Vp);
/* complex-valued 3-component fields (shtns.h:308-309) */
void spat_cplx_to_SHqst(shtns_cfg, cplx *Vr, cplx *Vt, cplx *Vp, cplx *Qlm, cplx *Slm, cplx *Tlm);
void SHqst_to_spat_cplx(shtns_cfg, cplx *Qlm, cplx *Slm, cplx *Tlm, cplx *Vr, cplx *Vt, cplx *Vp);

/* 2D vector transforms (shtns.h:283-289) */
void spat_to_SHsphtor(shtns_cfg, double *Vt, double *Vp, cplx *Slm, cplx *Tlm);
void SHsphtor_to_spat(shtns_cfg, cplx *Slm, cplx *Tlm, double *Vt, double *Vp);
void SHsph_to_spat(shtns_cfg, cplx *Slm, double *Vt, double *Vp);
void SHtor_to_spat(shtns_cfg, cplx *Tlm, double *Vt, double *Vp);

/* gradient of a scalar field = the spheroidal transform (shtns.h:297, 327, 345) */
#define SH_to_grad_spat(shtns, S, Gt, Gp) SHsph_to_spat(shtns, S, Gt, Gp)
#define SH_to_grad_spat_l(shtns, S, Gt, Gp, ltr) SHsph_to_spat_l(shtns, S, Gt, Gp, ltr)
#define SH_to_grad_spat_ml(shtns, im, S, Gt, Gp, ltr) SHsph_to_spat_ml(shtns, im, S, Gt, Gp, ltr)

/* 3-component transforms: radial scalar Q and tangential spheroidal/toroidal S, T
   (shtns.h:303, 306) */
void spat_to_SHqst(shtns_cfg, double *Vr, double *Vt, double *Vp, cplx *Qlm, cplx *Slm, cplx *Tlm);
void SHqst_to_spat(shtns_cfg, cplx *Qlm, cplx *Slm, cplx *Tlm, double *Vr, double *Vt, double *Vp);

/* truncated transforms at degree ltr <= lmax (shtns.h:315-324) */
void spat_to_SH_l(shtns_cfg, double *Vr, cplx *Qlm, int ltr);
void SH_to_spat_l(shtns_cfg, cplx *Qlm, double *Vr, int ltr);
void SHsphtor_to_spat_l(shtns_cfg, cplx *Slm, cplx *Tlm, double *Vt, double *Vp, int ltr);
void SHsph_to_spat_l(shtns_cfg, cplx *Slm, double *Vt, double *Vp, int ltr);
void SHtor_to_spat_l(shtns_cfg, cplx *Tlm, double *Vt, double *Vp, int ltr);
void spat_to_SHsphtor_l(shtns_cfg, double *Vt, double *Vp, cplx *Slm, cplx *Tlm, int ltr);
void spat_to_SHqst_l(shtns_cfg, double *Vr, double *Vt, double *Vp, cplx *Qlm, cplx *Slm, cplx *Tlm, int ltr);
void SHqst_to_spat_l(shtns_cfg, cplx *Qlm, cplx *Slm, cplx *Tlm, double *Vr, double *Vt, double *Vp, int ltr);

/* rotation about the z axis by alpha, on the host (shtns.h:188-191); Rlm may be Qlm */
void SH_Zrotate(shtns_cfg, cplx *Qlm, double alpha, cplx *Rlm);
/* rotations about y by alpha, about y and x by 90 degrees (shtns.h:193-198); need mres = 1,
   mmax = lmax.  A rotation by +a maps the field f to f(R^-1 x) (active, right-handed). */
void SH_Yrotate(shtns_cfg, cplx *Qlm, double alpha, cplx *Rlm);
void SH_Yrotate90(shtns_cfg, cplx *Qlm, cplx *Rlm);
void SH_Xrotate90(shtns_cfg, cplx *Qlm, cplx *Rlm);

/* rotation objects (shtns.h:200-218), orthonormal harmonics (norm may add SHT_NO_CS_PHASE,
   SHT_REAL_NORM).  set_angles_ZYZ(a, b, g): rotate by a about z, then b about y, then g
   about z, i.e. f'(x) = f(R^-1 x) with R = Rz(g) Ry(b) Rz(a); set_angles_ZXZ likewise with
   x; set_angle_axis: by theta about the axis (Vx, Vy, Vz).  wigner_d_matrix fills
   mx[(m'+l)(2l+1) + (m+l)] = d^l_{m'm}(b) and returns 2l+1 (0 on error).  apply_real takes
   nlm_calc(lmax, mmax, 1) coefficients of a real field, apply_cplx nlm_cplx coefficients of
   a complex field (LM_cplx layout); Rlm may be the input array. */
typedef struct shtns_rot_ *shtns_rot;
shtns_rot shtns_rotation_create(const int lmax, const int mmax, int norm);
void shtns_rotation_destroy(shtns_rot r);
void shtns_rotation_set_angles_ZYZ(shtns_rot r, double alpha, double beta, double gamma);
void shtns_rotation_set_angles_ZXZ(shtns_rot r, double alpha, double beta, double gamma);
void shtns_rotation_set_angle_axis(shtns_rot r, double theta, double Vx, double Vy, double Vz);
int shtns_rotation_wigner_d_matrix(shtns_rot r, const int l, double *mx);
void shtns_rotation_apply_cplx(shtns_rot r, cplx *Zlm, cplx *Rlm);
void shtns_rotation_apply_real(shtns_rot r, cplx *Qlm, cplx *Rlm);

/* point and latitude evaluation, on the host; no grid needed (shtns.h:352-362) */
double SH_to_point(shtns_cfg, cplx *Qlm, double cost, double phi);
#if defined(__cplusplus) && defined(__clang__)
/* std::complex<double> and C99 double complex share layout and return convention */
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wreturn-type-c-linkage"
#endif
cplx SH_to_point_cplx(shtns_cfg, cplx *alm, double cost, double phi);
#if defined(__cplusplus) && defined(__clang__)
#pragma clang diagnostic pop
#endif
void SH_to_grad_point(shtns_cfg, cplx *DrSlm, cplx *Slm, double cost, double phi, double *vr, double *vt, double *vp);
void SHqst_to_point(shtns_cfg, cplx *Qlm, cplx *Slm, cplx *Tlm, double cost, double phi, double *vr, double *vt,
                    double *vp);
void SH_to_lat(shtns_cfg shtns, cplx *Qlm, double cost, double *vr, int nphi, int ltr, int mtr);
void SHqst_to_lat(shtns_cfg, cplx *Qlm, cplx *Slm, cplx *Tlm, double cost, double *vr, double *vt, double *vp, int nphi,
                  int ltr, int mtr);

/* Status of the last call made from this thread: 0 = success.  Additive: SHTns has no
   status getter (its GPU calls print and return void, sht_gpu.cu:566, 610-615). */
int shtns_last_error(void);
const char *shtns_last_error_string(void);

#ifdef __cplusplus
}
#endif
#endif
