/* A C99 caller written the way SHTns users write them (shtns.h only, public struct
 * fields and macros), built with gcc against libshtns_amd.so.  Runs host-side calls only,
 * so it works without a GPU; prints key=value lines checked by tests/test_abi.py. */
#include <math.h>
#include <stddef.h>
#include <stdio.h>

#include "shtns.h"

int main(void) {
    shtns_cfg c = shtns_create(31, 31, 1, sht_orthonormal);
    int nspat = shtns_set_grid(c, sht_gauss | SHT_THETA_CONTIGUOUS, 0.0, 32, 64);
    double w[16];
    int nw = shtns_gauss_wts(c, w);
    double yl[32];
    legendre_sphPlm_array(c, 31, 2, 0.3, yl);
    printf("nlm=%u lmax=%u mmax=%u mres=%u nlat=%u nlat_2=%u nphi=%u nspat=%u ret=%d\n", c->nlm, c->lmax, c->mmax,
           c->mres, c->nlat, c->nlat_2, c->nphi, c->nspat, nspat);
    printf("LM11=%d LiM=%d nlm_calc=%ld nlm_cplx=%u LMc=%d\n", (int)LM(c, 1, 1), (int)LiM(c, 5, 3), nlm_calc(31, 31, 1),
           c->nlm_cplx, (int)LM_cplx(c, 5, -3));
    printf("ct0=%.17g st0=%.17g li5=%u mi40=%u nw=%d w0=%.17g y2=%.17g\n", c->ct[0], c->st[0], c->li[5], c->mi[40], nw,
           w[0], yl[0]);
    printf("sh00=%.17g sh10=%.17g sh11=%.17g\n", sh00_1(c), sh10_ct(c), sh11_st(c));
    printf("off_nlat=%zu off_li=%zu off_ct=%zu off_nlm_cplx=%zu\n", offsetof(struct shtns_info, nlat),
           offsetof(struct shtns_info, li), offsetof(struct shtns_info, ct), offsetof(struct shtns_info, nlm_cplx));
    {   /* loop macros; the gradient aliases expand to declared functions (compiled, not run:
           they need a GPU) */
        long n = 0, lsum = 0;
        LM_LOOP(c, n += 1 + 0 * lm;)
        LM_L_LOOP(c, lsum += l + 0 * lm;)
        printf("loopn=%ld lsum=%ld\n", n, lsum);
        if (n < 0) {
            SH_to_grad_spat(c, NULL, NULL, NULL);
            SH_to_grad_spat_l(c, NULL, NULL, NULL, 3);
            SH_to_grad_spat_ml(c, 1, NULL, NULL, NULL, 3);
        }
    }
    {   /* double complex return across the C ABI: a_{1,-1} = 1 gives Y_1^{-1}(theta, phi) */
        cplx *alm = (cplx *)shtns_malloc(sizeof(cplx) * c->nlm_cplx);
        for (unsigned i = 0; i < c->nlm_cplx; ++i) alm[i] = 0;
        alm[LM_cplx(c, 1, -1)] = 1.0;
        cplx z = SH_to_point_cplx(c, alm, 0.3, 0.5);
        printf("zre=%.17g zim=%.17g\n", creal(z), cimag(z));
        shtns_free(alm);
    }
    shtns_destroy(c);
    shtns_reset();
    return 0;
}
