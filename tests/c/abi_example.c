/* The calling sequence a C host uses (the one INTEGRATION.md documents),
 * compiled as plain C against include/cvt.h and linked to libcvt_hip.so.
 * Without a GPU, cvt_create reports an error and the program says so. */
#include <stdio.h>
#include <stdlib.h>

#include "cvt.h"

int main(void)
{
    const int n = 64, S = 16;
    float *rho = malloc(sizeof(float) * n * n), xy[2 * 16];
    for (int i = 0; i < n * n; ++i) rho[i] = 1.0f;
    for (int i = 0; i < S; ++i) { xy[2 * i] = 4.0f + 3.5f * i; xy[2 * i + 1] = 60.0f - 3.5f * i; }
    printf("version %d\n", cvt_version());
    cvt_ctx *ctx = NULL;
    int rc = cvt_create(&ctx, n, n, S);
    if (rc != CVT_OK) {
        printf("no engine: %d %s\n", rc, cvt_last_error());
        free(rho);
        return 0;
    }
    double e = 0.0, g = 0.0;
    int32_t *labels = malloc(sizeof(int32_t) * n * n);
    if ((rc = cvt_set_density(ctx, rho)) || (rc = cvt_set_sites(ctx, xy, NULL, S)) ||
        (rc = cvt_lloyd(ctx, 5, 1.0f)) || (rc = cvt_energy(ctx, &e, &g)) ||
        (rc = cvt_get_sites(ctx, xy)) || (rc = cvt_voronoi(ctx)) || (rc = cvt_get_labels(ctx, labels)))
        printf("error %d: %s\n", rc, cvt_last_error());
    else
        printf("energy %.6f gradnorm %.6f label0 %d\n", e, g, labels[0]);
    cvt_destroy(ctx);
    free(labels);
    free(rho);
    return rc != CVT_OK;
}
