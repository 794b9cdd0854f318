// tsi_jit.h -- per-tape straight-line evaluation kernels (tsi_jit.cpp).
#pragma once
#include <string>
#include <vector>

#include "tsi.h"
#include "tsi_tape.h"

namespace tsi {

// HIP source of the level kernel for one tape: tsi_tape_eval_fwd (forward-only frontier list)
// or tsi_tape_eval_rev (forward with saved operands, then reverse).
std::string jit_source(const tsi_gpu_instruction *tape, uint32_t n, const std::vector<FwdIns> &fwd,
                       const std::vector<RevIns> &rev, const Lowered &L, bool rev_kernel);

// Compiles `src` for gfx950 with hiprtc (the library's kernel headers are embedded), or finds
// the code object in this process's cache or the on-disk cache ($TSI_JIT_CACHE, default
// $XDG_CACHE_HOME/tsi-jit or ~/.cache/tsi-jit).  Returns TSI_OK or TSI_E_JIT (log filled).
int jit_compile(const std::string &src, std::vector<char> &code, std::string &log, bool &cache_hit, double &ms);

}  // namespace tsi
