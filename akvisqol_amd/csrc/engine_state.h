// Per-pair state of one chunk of Engine::run, shared by the forward pass
// (engine.cpp) and the gradient pass (engine_grad.cpp).
#pragma once

#include <vector>

#include "engine.h"

namespace vq {

// A realignment row the gradient may pass through (realign.zig:262-281,
// pipeline.zig:986-1013): its spectrograms on the device and where its
// degraded frames sit in the aligned degraded signal.
struct RowTrace {
    int match = -1;
    int nf = 0;
    const double *spec_r = nullptr, *spec_d = nullptr;  // comparison-ready [nb][nf]
    const double *raw_r = nullptr, *raw_d = nullptr;    // the same before the dB preparation
    bool deg_reused = false;  // degraded frames copied from the full-signal frames g0 .. g0+nf
    long long g0 = 0;
    long long first = 0;               // aligned-signal index of row frame 0's first sample
    long long valid_lo = 0, valid_hi = 0;  // aligned samples the row's slice holds (the rest is padding)
};

struct PairState {
    bool live = true;
    const PreparedRef* prep = nullptr;
    const double* dref = nullptr;
    const double* ddeg = nullptr;
    long long rl = 0, dl = 0;
    int lagjob = -1, vadjob = -1;
    long long lag = 0, shift = 0, dl2 = 0;
    double lag_s = 0, factor = 1.0;
    double e_ref = 0, e_deg = 0;  // Σ ref², Σ kept deg² (the SPL match's energies)
    int nf_ref = 0, nf_deg = 0, np = 0, width = 0;
    std::vector<int> idx, lo, hi;
    std::vector<int> match_state;  // DP state per patch: 0 not reached, 1 valid, 2 no real match
    VadPlan vp;
    size_t vad_off = 0;
    double* degs = nullptr;
    double *spec_ref = nullptr, *spec_deg = nullptr, *raw_ref = nullptr, *raw_deg = nullptr;
    double *plane = nullptr, *sims = nullptr, *dp = nullptr;
    int *bt = nullptr, *d_idx = nullptr, *d_lo = nullptr, *d_hi = nullptr, *d_off = nullptr, *d_state = nullptr;
    std::vector<RowTrace> rows;  // realignment rows, kept when gradients are asked for
};

}  // namespace vq
