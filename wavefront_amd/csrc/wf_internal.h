// Internal helpers shared by the host and device translation units.
#pragma once

#include <cstdarg>
#include <cstdint>

#include "wavefront_hip.h"

namespace wf {

// Records a thread-local message for wf_last_error() and returns `code`.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

// Env hook for error-path tests: WF_FAULT_INJECT=<name> makes the named
// entry point return WF_ERR_HIP before touching the device ("short_write"
// makes the snapshot writer report a short write instead of writing).
bool fault_injected(const char *name);

// Reference constants (src/layout.rs:29-31, std::f32::consts::PI).
constexpr float kEps0 = 8.854188e-12f;
constexpr float kMu0 = 1.2566371e-6f;
constexpr float kPiF = 3.14159265358979323846f;

}  // namespace wf
