// Helpers shared by the C ABI translation units (capi.cpp, server.cpp).
#pragma once

#include <memory>
#include <string>

#include "../../include/visqol.h"
#include "engine.h"

namespace vq_capi {

// Thread-local message for visqol_last_error (capi.zig:85-122).
void set_error(const char* fmt, ...);
int invalid(const char* msg);
// Heap result owning every array it exposes, freed by visqol_result_free.
visqol_result* box(const vq::PairResult& r);
// A gfx950 HIP device is current.
bool device_ok(std::string& why);
// Engine and models for validated options; a visqol.h code.
// `interleave`: batches of several chunks run every other chunk on a helper
// engine (off for the server, whose workers already overlap).
int build_engine(const visqol_options& o, vq::EngineOptions& eo, std::unique_ptr<vq::Engine>& engine,
                 bool interleave = true);

}  // namespace vq_capi
