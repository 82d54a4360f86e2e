// Composition compiler (compiler.cpp): constraint programs compiled into HIP kernels at run time.
#pragma once
#include "runtime.h"

namespace sh {

// The number of constraints a program declares (header check only).
uint32_t compiled_component_constraints(const uint32_t* words, size_t n_words);
// evaluate_constraint_quotients_on_domain for a constraint program: compiles it for `device` on
// first use (cached by content), then launches it on `stream`. Device tables: column addresses,
// the parameter table (entry 0 the cumulative-sum shift), coefficient powers and the vanishing
// inverses; acc (4 coordinate columns of 2^eval_log) is accumulated into.
void evaluate_compiled_component(int device, const uint32_t* words, size_t n_words, uint32_t eval_log,
                                 uint32_t trace_log, const uint32_t* const* d_trace, uint32_t n_trace,
                                 const uint32_t* const* d_inter, uint32_t n_inter, const qm31* d_params,
                                 uint32_t n_params, const qm31* d_pows, uint32_t n_constraints,
                                 const uint32_t* d_dinv, Cols4 acc, hipStream_t stream);

}  // namespace sh
