// tsi_jit.h -- per-tape straight-line evaluation kernels (tsi_jit.cpp).
#pragma once
#include <string>
#include <vector>

#include "tsi.h"
#include "tsi_tape.h"

namespace tsi {

// Layout of the per-cell liveness records the pruned straight-line kernels exchange between
// levels: t-window lo/hi, region_words words of region bits (one bit per `group` records),
// mode_words words of MIN/MAX choices (2 bits each, 16 per word).
struct JitLayout {
    uint32_t group = 16, regions = 0, region_words = 0, minmax = 0, mode_words = 0, record_words = 0;
};
// (prune: the layout of pruned modules, the only ones that exchange records; unpruned modules
// use longer regions)
JitLayout jit_layout(const std::vector<FwdIns> &fwd, const Lowered &L, bool prune, uint32_t group = 0);

// The domain a tape's kernels may be specialised for: every cell box inside `box` (lo xyz, hi
// xyz), every t-window inside [t0, t1].  fix_free marks the records whose value no cell in it can
// take outside [-1e6, 1e6] (fix then leaves the value unchanged and is not emitted).
struct JitDomain { float box[6]; float t0, t1; };
std::vector<char> fix_free(const tsi_gpu_instruction *tape, uint32_t n, const std::vector<FwdIns> &fwd,
                           const JitDomain &d);

// HIP source of the code of one list for one tape: the forward-only frontier list, or
// (rev_kernel) forward with saved operands, then reverse; with a kernel of its own
// (tsi_tape_eval_fwd / tsi_tape_eval_rev), which jit_source_level replaces by one for both.  prune: regions dead for
// all of a wavefront's parents are skipped and decided MIN/MAX become copies (liveness records
// written for enqueued cells).  skip: the reverse pass of an unpruned module skips subtrees whose
// sensitivity is [0, 0] in the whole wavefront.  group: records per region (0: the default of
// pruned or unpruned modules).
std::string jit_source(const tsi_gpu_instruction *tape, uint32_t n, const std::vector<FwdIns> &fwd,
                       const std::vector<RevIns> &rev, const Lowered &L, bool rev_kernel, bool prune,
                       const JitDomain *domain = nullptr, bool skip = true, uint32_t group = 0);
// Both level kernels' code in one module with one kernel, tsi_tape_eval_level(KArgs, level,
// time), that runs a level's forward-only and reverse lists side by side (blockDim 256).
std::string jit_source_level(const tsi_gpu_instruction *tape, uint32_t n, const std::vector<FwdIns> &fwd,
                             const std::vector<RevIns> &rev, const Lowered &L, bool prune,
                             const JitDomain *domain = nullptr);

// One tape of a batch module: the tape, its lowering.
struct BatchTape {
    const tsi_gpu_instruction *tape;
    uint32_t n;
    const std::vector<FwdIns> *fwd;
    const std::vector<RevIns> *rev;
    const Lowered *low;
};
// The persistent batch kernel (tsi_batch.h) over several tapes: each tape's unpruned code of both
// lists in a namespace of its own (tb<i>), and one kernel, tsi_batch_run(BatchArgs), that runs a
// session's work units with the code of the session's tape.
std::string jit_source_batch(const std::vector<BatchTape> &tapes, const JitDomain *domain);

// Checks every saved-operand stack row, liveness-record word and choice column the source of a
// module addresses against what the runtime allocates for it: stack_rows rows per wavefront
// (Lowered::stack + 1), record_words and mode_words of the pruned layout (jit_layout(.., true)).
// Returns an empty string, or the first violation.
std::string jit_check_bounds(const std::string &src, uint32_t stack_rows, uint32_t record_words,
                             uint32_t mode_words);

// Compiles `src` for gfx950 with hiprtc (the library's kernel headers are embedded), or finds
// the code object in this process's cache or the on-disk cache ($TSI_JIT_CACHE, default
// $XDG_CACHE_HOME/tsi-jit or ~/.cache/tsi-jit).  prune: the source is a pruned module (unpruned
// ones read the stack as global memory and are compiled without machine sinking).  Returns
// TSI_OK or TSI_E_JIT (log filled).
int jit_compile(const std::string &src, bool prune, std::vector<char> &code, std::string &log, bool &cache_hit,
                double &ms);

}  // namespace tsi
