/* A caller of the reference's C ABI written the way its wrappers are
 * (src/cpp/src/visqol_compat.cc:83-105 VisqolManager::Init, :128-150 Run on
 * sample spans; src/python/visqol/visqol_lib_py.py:76-77): options from
 * visqol_options_default, the backend from VISQOL_BACKEND (unset or unknown
 * -> VISQOL_BACKEND_CPU), then visqol_create and visqol_measure.  Nothing in
 * it knows of this library's HIP slot.
 *
 * usage: default_flow MODE LATTICE SAMPLE_RATE REF.f64 DEG.f64
 * The signals are raw little-endian doubles.  Prints "create RC", and on
 * success "measure RC" and the scores as hexadecimal floats (exact bits). */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "visqol.h"

static uint32_t backend_from_env(void) {
    const char *b = getenv("VISQOL_BACKEND");
    if (b == NULL) return VISQOL_BACKEND_CPU;
    if (strcmp(b, "metal") == 0) return VISQOL_BACKEND_METAL;
    if (strcmp(b, "cuda") == 0) return VISQOL_BACKEND_CUDA;
    if (strcmp(b, "vulkan") == 0) return VISQOL_BACKEND_VULKAN;
    return VISQOL_BACKEND_CPU;
}

static double *read_f64(const char *path, size_t *n) {
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    fseek(f, 0, SEEK_END);
    const long bytes = ftell(f);
    fseek(f, 0, SEEK_SET);
    double *x = (double *)malloc(bytes > 0 ? (size_t)bytes : 1);
    *n = fread(x, sizeof(double), (size_t)bytes / sizeof(double), f);
    fclose(f);
    return x;
}

int main(int argc, char **argv) {
    if (argc != 6) {
        fprintf(stderr, "usage: %s MODE LATTICE SAMPLE_RATE REF.f64 DEG.f64\n", argv[0]);
        return 2;
    }
    const int speech = atoi(argv[1]) == 1;
    visqol_options opts;
    visqol_options_default(&opts);
    opts.mode = speech ? VISQOL_MODE_SPEECH : VISQOL_MODE_AUDIO;
    opts.search_window_radius = 60;
    opts.use_lattice_model = speech && atoi(argv[2]) != 0;
    opts.model_path = NULL;
    opts.backend = backend_from_env();
    visqol_ctx *ctx = NULL;
    int rc = visqol_create(&opts, &ctx);
    printf("backends %u\n", visqol_available_backends());
    printf("create %d\n", rc);
    if (rc != VISQOL_OK) {
        printf("error %s\n", visqol_last_error());
        return 0;
    }
    size_t nr = 0, nd = 0;
    double *ref = read_f64(argv[4], &nr), *deg = read_f64(argv[5], &nd);
    if (!ref || !deg) return 2;
    visqol_result *r = NULL;
    rc = visqol_measure(ctx, ref, nr, deg, nd, (uint32_t)atoi(argv[3]), &r);
    printf("measure %d\n", rc);
    if (rc == VISQOL_OK) {
        printf("moslqo %a\nvnsim %a\nlag %a\n", r->moslqo, r->vnsim, r->alignment_lag_s);
        printf("fvnsim");
        for (size_t b = 0; b < r->num_bands; b++) printf(" %a", r->fvnsim[b]);
        printf("\npatches %zu\n", r->num_patches);
        visqol_result_free(r);
    } else {
        printf("error %s\n", visqol_last_error());
    }
    visqol_destroy(ctx);
    free(ref);
    free(deg);
    return 0;
}
