/* A C99 caller of libiblgf_gpu.so through include/iblgf_gpu.h only, the way
 * the reference's C++ would bind it.  Mode "cpu" checks the status codes that
 * need no device; mode "gpu" also solves through iblgf_gpu_apply_lgf_typed and
 * iblgf_gpu_apply_lgf and requires the two results to be bitwise equal.
 * Prints one line per check; exit status 0 when every check holds. */
#include "iblgf_gpu.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int failures = 0;

static void check(int cond, const char* what)
{
    printf("%s: %s\n", cond ? "ok" : "FAIL", what);
    if (!cond) ++failures;
}

static iblgf_tree* uniform_tree(int subtract_non_leaf)
{
    const int   base[3] = {-16, -16, -16}, ext[3] = {32, 32, 32};
    const int   region[6] = {-16, -16, -16, 32, 32, 32};
    iblgf_tree* t = NULL;
    if (iblgf_tree_create(base, ext, 8, 1.0 / 32, 1, region, &t) != IBLGF_OK) return NULL;
    if (iblgf_tree_init_refine(t, 0, 0, NULL, NULL) != IBLGF_OK) return NULL;
    if (iblgf_tree_build_masks(t, subtract_non_leaf) != IBLGF_OK) return NULL;
    return t;
}

int main(int argc, char** argv)
{
    const int   gpu = argc > 1 && strcmp(argv[1], "gpu") == 0;
    iblgf_tree* t = uniform_tree(0);
    iblgf_tree* tsub = uniform_tree(1);
    int         f;
    check(t != NULL && tsub != NULL, "tree create / init_refine / build_masks");
    check(iblgf_gpu_apply_lgf(NULL, 0, 1, IBLGF_KERNEL_LAPLACE, 0.0) == IBLGF_ERR_ARG, "null context is an argument error");
    for (f = 0; f <= 7; ++f)
    {
        const int rc = iblgf_gpu_apply_lgf_typed(NULL, 0, 1, IBLGF_KERNEL_LAPLACE, 0.0, f);
        const int built = f == IBLGF_FMM_AMR2AMR || f == IBLGF_FMM_IB2xIB || f == IBLGF_FMM_xIB2IB ||
                          f == IBLGF_FMM_STREAM || f == IBLGF_FMM_IB2AMR;
        char      what[96];
        snprintf(what, sizeof what, "fmm_type %d without context -> %s", f, built ? "ERR_ARG" : "ERR_UNSUPPORTED");
        check(rc == (built ? IBLGF_ERR_ARG : IBLGF_ERR_UNSUPPORTED), what);
    }
    check(iblgf_gpu_apply_lgf_typed(NULL, 0, 1, 0, 0.0, 8) == IBLGF_ERR_ARG, "unknown fmm_type is an argument error");
    check(iblgf_gpu_tree_upload(NULL, tsub) == IBLGF_ERR_UNSUPPORTED, "subtract_non_leaf tree refused on upload");
    check(strstr(iblgf_last_error(), "subtract_non_leaf") != NULL, "refusal names subtract_non_leaf");
    if (gpu)
    {
        iblgf_gpu_ctx* ctx = NULL;
        int            nb = 0, bl, dp, be, nl, fs, d1, d2, q, n = 10;
        double         dxb;
        double *       src, *r1, *r2;
        size_t         npts;
        check(iblgf_gpu_create(0, &ctx) == IBLGF_OK, "context on device 0");
        check(iblgf_gpu_tree_upload(ctx, tsub) == IBLGF_ERR_UNSUPPORTED, "subtract_non_leaf refused with a context");
        check(iblgf_gpu_tree_upload(ctx, t) == IBLGF_OK, "tree upload");
        iblgf_tree_info(t, &nb, &bl, &dp, &be, &dxb, &nl);
        npts = (size_t)nb * n * n * n;
        src = (double*)calloc(npts, sizeof(double));
        r1 = (double*)calloc(npts, sizeof(double));
        r2 = (double*)calloc(npts, sizeof(double));
        for (q = 0; q < (int)npts; ++q) src[q] = (double)((q * 7919) % 1013) / 1013.0 - 0.5;
        check(iblgf_gpu_field_create(ctx, &fs) == IBLGF_OK && iblgf_gpu_field_create(ctx, &d1) == IBLGF_OK &&
                  iblgf_gpu_field_create(ctx, &d2) == IBLGF_OK,
            "three fields");
        iblgf_gpu_field_upload(ctx, fs, src);
        check(iblgf_gpu_apply_lgf(ctx, fs, d1, IBLGF_KERNEL_LAPLACE, 0.0) == IBLGF_OK, "apply_lgf");
        check(iblgf_gpu_apply_lgf_typed(ctx, fs, d2, IBLGF_KERNEL_LAPLACE, 0.0, IBLGF_FMM_AMR2AMR) == IBLGF_OK,
            "apply_lgf_typed AMR2AMR");
        iblgf_gpu_field_download(ctx, d1, r1);
        iblgf_gpu_field_download(ctx, d2, r2);
        check(memcmp(r1, r2, npts * sizeof(double)) == 0, "typed AMR2AMR equals apply_lgf bit for bit");
        check(iblgf_gpu_apply_lgf_typed(ctx, fs, d2, IBLGF_KERNEL_IF, 0.5, IBLGF_FMM_STREAM) ==
                  IBLGF_ERR_UNSUPPORTED,
            "STREAM with the IF kernel refused with a context");
        check(iblgf_gpu_apply_lgf_typed(ctx, fs, d2, IBLGF_KERNEL_LAPLACE, 0.0, IBLGF_FMM_STREAM) == IBLGF_OK,
            "STREAM Laplace solve");
        check(iblgf_gpu_apply_lgf_typed(ctx, fs, d2, IBLGF_KERNEL_IF, 0.5, IBLGF_FMM_IB2AMR) == IBLGF_ERR_UNSUPPORTED,
            "IB2AMR with the IF kernel refused with a context");
        check(iblgf_gpu_apply_lgf_typed(ctx, fs, d2, IBLGF_KERNEL_LAPLACE, 0.0, IBLGF_FMM_LAPLACIAN_BC) ==
                  IBLGF_ERR_UNSUPPORTED,
            "Laplacian_BC refused with a context");
        free(src);
        free(r1);
        free(r2);
        iblgf_gpu_destroy(ctx);
    }
    iblgf_tree_destroy(t);
    iblgf_tree_destroy(tsub);
    printf("%d failures\n", failures);
    return failures != 0;
}
