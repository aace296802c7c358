/* Oracle internals shared between its translation units (test infra only). */
#ifndef VO_INTERNAL_H
#define VO_INTERNAL_H
#define _GNU_SOURCE
#include "vo.h"

#include <math.h>
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

double vo_spl_factor_for(const double *ref, size_t nr, const double *deg, size_t nd);

/* nsim.zig:12-21 (PatchSimilarity) on caller-owned arrays. */
typedef struct {
    double *means, *stddevs, *energy;
    double mean_of_means;
} vo_psim;

void vo_patch_similarity_into(const double *ref, const double *deg, size_t nrows, size_t ncols, vo_psim *out);

/* dp.zig:16-29 (Match) with fixed-size per-band storage. */
typedef struct {
    double similarity, ref_start, ref_end, deg_start, deg_end;
    double means[VO_MAX_BANDS], stddevs[VO_MAX_BANDS], energy[VO_MAX_BANDS];
    int valid;
} vo_match;

/* realign.zig:39 */
typedef struct { size_t pad, base, len; } vo_layout;
vo_layout vo_layout_of(size_t signal_len, uint32_t sr, double start, double end);

/* Parsed mapping models (svr.zig:16-34; lattice.zig:14-21). */
typedef struct vo_svr vo_svr;
vo_svr *vo_svr_parse(const char *text, size_t len, int *err);
void vo_svr_free(vo_svr *m);
double vo_svr_eval(const vo_svr *m, const double *obs, size_t n);

typedef struct vo_lattice vo_lattice;
vo_lattice *vo_lattice_load(const uint8_t *bytes, size_t len, size_t nb, int *err);
void vo_lattice_free(vo_lattice *m);
int vo_lattice_eval(const vo_lattice *m, const double *f0, const double *f1, const double *f2, const double *f3,
                    double *y);

void vo_set_error(const char *fmt, ...);
#endif
