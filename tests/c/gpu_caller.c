/* A C99 caller of the transforms, written the way SHTns users write them against shtns.h
 * and shtns_cuda.h (here shtns_hip.h; time_SHT.c:304-315, cudamaxrelerror.cu:43-68), built
 * with gcc and linked to libshtns_amd.so and the HIP runtime.
 *   gpu_caller <lmax> <nlat> <nphi> <in.bin> <out.bin>
 * in.bin: nlm complex doubles (Qlm).  out.bin, in this order:
 *   V_host[nspat]  (SH_to_spat, shtns_malloc'd arrays)
 *   Q_host[nlm]    (spat_to_SH of V_host)
 *   V_dev[nspat]   (cu_SH_to_spat on hipMalloc'd arrays)
 *   Q_dev[nlm]     (cu_spat_to_SH of V_dev)
 *   V_f32[nspat]   (cu_SH_to_spat_float; floats widened to double)
 *   Q_f32[nlm]     (cu_spat_to_SH_float of V_f32)
 * Exit status 0 on success; prints the failing call otherwise. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "shtns.h"
#include "shtns_hip.h"

#define CHECK(x)                                                                  \
    do {                                                                          \
        if ((x) != hipSuccess) {                                                  \
            fprintf(stderr, "HIP call failed: %s\n", #x);                         \
            return 2;                                                             \
        }                                                                         \
    } while (0)

int main(int argc, char **argv) {
    if (argc != 6) {
        fprintf(stderr, "usage: gpu_caller lmax nlat nphi in.bin out.bin\n");
        return 1;
    }
    const int lmax = atoi(argv[1]), nlat = atoi(argv[2]), nphi = atoi(argv[3]);
    shtns_cfg c = shtns_create(lmax, lmax, 1, sht_orthonormal);
    if (!c || shtns_set_grid(c, sht_gauss | SHT_THETA_CONTIGUOUS, 0.0, nlat, nphi) <= 0) {
        fprintf(stderr, "config failed\n");
        return 1;
    }
    const size_t nlm = c->nlm, nspat = c->nspat;
    cplx *Q = (cplx *)shtns_malloc(sizeof(cplx) * nlm);
    cplx *Q2 = (cplx *)shtns_malloc(sizeof(cplx) * nlm);
    double *V = (double *)shtns_malloc(sizeof(double) * nspat);
    FILE *f = fopen(argv[4], "rb");
    if (!f || fread(Q, sizeof(cplx), nlm, f) != nlm) {
        fprintf(stderr, "cannot read %s\n", argv[4]);
        return 1;
    }
    fclose(f);
    FILE *out = fopen(argv[5], "wb");
    if (!out) return 1;

    /* host-pointer transforms */
    SH_to_spat(c, Q, V);
    spat_to_SH(c, V, Q2);
    if (shtns_last_error()) {
        fprintf(stderr, "host transforms: %s\n", shtns_last_error_string());
        return 3;
    }
    fwrite(V, sizeof(double), nspat, out);
    fwrite(Q2, sizeof(cplx), nlm, out);

    /* device-pointer transforms */
    cplx *dQ;
    double *dV;
    CHECK(hipMalloc((void **)&dQ, sizeof(cplx) * nlm));
    CHECK(hipMalloc((void **)&dV, sizeof(double) * nspat));
    CHECK(hipMemcpy(dQ, Q, sizeof(cplx) * nlm, hipMemcpyHostToDevice));
    cu_SH_to_spat(c, dQ, dV, lmax);
    CHECK(hipDeviceSynchronize());
    CHECK(hipMemcpy(V, dV, sizeof(double) * nspat, hipMemcpyDeviceToHost));
    CHECK(hipMemset(dQ, 0, sizeof(cplx) * nlm));
    cu_spat_to_SH(c, dV, dQ, lmax);
    CHECK(hipDeviceSynchronize());
    CHECK(hipMemcpy(Q2, dQ, sizeof(cplx) * nlm, hipMemcpyDeviceToHost));
    if (shtns_last_error()) {
        fprintf(stderr, "device transforms: %s\n", shtns_last_error_string());
        return 3;
    }
    fwrite(V, sizeof(double), nspat, out);
    fwrite(Q2, sizeof(cplx), nlm, out);

    /* single-precision device transforms (shtns_cuda.h:40-44) */
    cplx_f *Qf = (cplx_f *)malloc(sizeof(cplx_f) * nlm);
    float *Vf = (float *)malloc(sizeof(float) * nspat);
    for (size_t i = 0; i < nlm; ++i) Qf[i] = (cplx_f)Q[i];
    cplx_f *dQf;
    float *dVf;
    CHECK(hipMalloc((void **)&dQf, sizeof(cplx_f) * nlm));
    CHECK(hipMalloc((void **)&dVf, sizeof(float) * nspat));
    CHECK(hipMemcpy(dQf, Qf, sizeof(cplx_f) * nlm, hipMemcpyHostToDevice));
    cu_SH_to_spat_float(c, dQf, dVf, lmax);
    CHECK(hipDeviceSynchronize());
    CHECK(hipMemcpy(Vf, dVf, sizeof(float) * nspat, hipMemcpyDeviceToHost));
    cu_spat_to_SH_float(c, dVf, dQf, lmax);
    CHECK(hipDeviceSynchronize());
    CHECK(hipMemcpy(Qf, dQf, sizeof(cplx_f) * nlm, hipMemcpyDeviceToHost));
    if (shtns_last_error()) {
        fprintf(stderr, "float transforms: %s\n", shtns_last_error_string());
        return 3;
    }
    for (size_t i = 0; i < nspat; ++i) V[i] = Vf[i];
    for (size_t i = 0; i < nlm; ++i) Q2[i] = Qf[i];
    fwrite(V, sizeof(double), nspat, out);
    fwrite(Q2, sizeof(cplx), nlm, out);
    fclose(out);

    CHECK(hipFree(dQ));
    CHECK(hipFree(dV));
    CHECK(hipFree(dQf));
    CHECK(hipFree(dVf));
    free(Qf);
    free(Vf);
    shtns_free(Q);
    shtns_free(Q2);
    shtns_free(V);
    shtns_destroy(c);
    return 0;
}
