/*
 * sht_oracle.h -- CPU FP64 restatement of SHTns's Gauss-grid spherical harmonic
 * transforms (scalar and spheroidal/toroidal vector).
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker used by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py.  The product
 * library (shtns_amd/) never links, loads or calls it.
 *
 * Conventions (same as SHTns, shtns.h:48-55, doc/main.dox:263-303):
 *   - orthonormal spherical harmonics with Condon-Shortley phase;
 *   - real field f = sum_l Q_l0 Y_l0 + 2 Re sum_{m>0} Q_lm Y_lm (Im Q_l0 ignored);
 *   - spectral array: complex (re,im interleaved) Qlm[LM(l,m)], l contiguous per m,
 *     LM(l,m) = (m/mres)*(2*lmax+2-(m+mres))/2 + l   (shtns.h:104-106);
 *   - spatial array: theta-contiguous V[ip*nlat + it], it=0 next to the north pole
 *     (cos(theta) > 0 first), phi_ip = 2*pi*ip/(nphi*mres).
 */
#ifndef SHT_ORACLE_H
#define SHT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_cfg orc_cfg;

/* lmax>=1, mmax*mres<=lmax, nlat > lmax, nphi > 2*mmax. Returns NULL on bad sizes.
 * grid: 0 Gauss (orc_create), 1 regular without poles (Fejer), 2 regular with both poles
 * (Clenshaw-Curtis).  Odd nlat puts a row on the equator. */
orc_cfg *orc_create(int lmax, int mmax, int mres, int nlat, int nphi);
orc_cfg *orc_create_grid(int lmax, int mmax, int mres, int nlat, int nphi, int grid);
void orc_destroy(orc_cfg *c);
long orc_nlm(const orc_cfg *c);
void orc_set_threads(int n);

/* Gauss nodes: cos(theta), sin(theta), raw Gauss-Legendre weights (sum 2), all of size nlat. */
void orc_grid(const orc_cfg *c, double *ct, double *st, double *w);

/* Normalized associated Legendre functions Y_l^m(x), l=m..ltop (ltop<=lmax+1), m=im*mres. */
void orc_legendre(const orc_cfg *c, int im, double x, double sint, int ltop, double *yl);

/* Polar optimization of SHTns (PolarOptimize, sht_init.c:522-557): with eps > 0, order im
 * skips the latitude pairs below tm[im] (rounded down to even); eps <= 0 (the default after
 * orc_create) skips nothing.  orc_polar_tm copies the mmax+1 start indices. */
void orc_set_polar_opt(orc_cfg *c, double eps);
void orc_polar_tm(const orc_cfg *c, int *tm);

/* Transforms; ltr<0 means lmax.  Inputs are never modified. */
void orc_SH_to_spat(const orc_cfg *c, const double *Qlm, double *Vr, int ltr);
void orc_spat_to_SH(const orc_cfg *c, const double *Vr, double *Qlm, int ltr);
void orc_SHsphtor_to_spat(const orc_cfg *c, const double *Slm, const double *Tlm,
                          double *Vt, double *Vp, int ltr);
void orc_spat_to_SHsphtor(const orc_cfg *c, const double *Vt, const double *Vp,
                          double *Slm, double *Tlm, int ltr);

/* Complex FFT used by the oracle (exposed so tests can pin it against numpy.fft).
 * data: n complex (interleaved), sign -1 forward (e^{-i...}), +1 backward; unnormalized. */
void orc_fft(double *data, int n, int sign);

#ifdef __cplusplus
}
#endif
#endif
