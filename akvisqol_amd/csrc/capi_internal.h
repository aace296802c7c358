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
// Inside a catch block: the visqol.h code and last-error message of the
// exception being handled (BackendError keeps its code, std::bad_alloc is
// E_OUT_OF_MEMORY, any other std::exception E_BACKEND, anything else
// E_INTERNAL), so no exception crosses a C ABI function.
int exception_code();
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
