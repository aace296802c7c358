// Host-side pieces of the gradient ∂moslqo/∂deg (backend.zig:101-110): the
// discrete choices of the forward pass held fixed, f64 throughout.
#pragma once

#include <functional>
#include <vector>

#include "engine.h"

namespace vq {

// ∂moslqo/∂(fvnsim, fvnsim10, fstdnsim, fvdegenergy) at the given features,
// zero where the score is clamped (visqol.zig:103-146).
using MapperGrad = std::function<void(const double* fvnsim, const double* fvnsim10, const double* fstdnsim,
                                      const double* fvdegenergy, double* g_fvnsim, double* g_fvnsim10,
                                      double* g_fstdnsim, double* g_fvdegenergy)>;

// ∂moslqo/∂(band mean, band stddev, band deg energy) of every match, as the
// aggregation maps them (visqol.zig:714-800): out[(match·nb + band)·3 + k].
// Zero when vnsim < 0.15 (the score is pinned).
void result_grad(const std::vector<Match>& matches, int nb, double frame_duration, const MapperGrad& mapper,
                 std::vector<double>& out);

// measurePatchSimilarity (nsim.zig:427-510) differentiated: from the per-band
// mean, stddev and deg energy slopes (any may be NULL) to the comparison-ready
// ref and deg patch cells, row-major [nrows][ncols]; adds into g_ref, g_deg.
void patch_similarity_grad(const double* ref, const double* deg, int nrows, int ncols, const double* g_means,
                           const double* g_stddevs, const double* g_energy, double* g_ref, double* g_deg);

// prepareForComparison (spectrogram.zig:14-80) differentiated: the raw
// spectrograms' slopes from the prepared ones' (overwrites g_raw_*).  Cells at
// the absolute floor or zero pass nothing; cells raised to a frame's floor pass
// theirs to the frame's maximum; the joint minimum's cell takes minus the sum
// over all cells.
void prepare_grad(const double* raw_ref, const double* raw_deg, int nb, int nfr, int nfd, const double* g_ref_prep,
                  const double* g_deg_prep, double* g_raw_ref, double* g_raw_deg);

// The input's slopes from the aligned, SPL-scaled degraded signal's
// (aligned[t] = factor · deg[t - shift]; when `scaled`, factor varies as
// deg_energy^(-1/2), deg_energy the sum of squares of the samples kept).
void aligned_to_input_grad(const double* deg, long long deg_len, long long shift, long long aligned_len,
                           double factor, double deg_energy, bool scaled, const double* g_aligned, double* g_deg);

}  // namespace vq
