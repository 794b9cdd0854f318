// tsi_runtime.hip -- the C ABI of include/tsi.h: context, device buffers, frame protocol.
//
// Replaces the wgpu half of src/main.rs: State::new's buffers (208-303), State::compute's
// first-frame seeding and `main` dispatch (724-814) and the compute half of State::render
// (649-656 rt_main, 683-690 buf_main).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "tsi.h"
#include "tsi_jit.h"
#include "tsi_launch.h"
#include "tsi_tape.h"

namespace tsi {
int lower(const tsi_gpu_instruction *tape, uint32_t n, std::vector<FwdIns> &fwd, std::vector<RevIns> &rev,
          Lowered &L);
uint32_t fwd_cost(uint32_t op, uint32_t side);
}  // namespace tsi

using namespace tsi;

namespace {

// Staging buffer of tsi_frame_rgba8_noalloc in the library's static device memory: each device
// the library runs on holds one copy (32 MiB) until the process ends.  Host code never uses the
// array's name as a device address: tsi_create looks up each device's copy once, with that
// device current, and every use goes through the pointer it stored.
constexpr size_t kStaticRgba8Pixels = size_t(TSI_STATIC_RGBA8_MAX_PIXELS);
static __device__ __align__(256) uchar4 scratch[kStaticRgba8Pixels];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uchar4 *g_scratch_ptr[kMaxDevices] = {};

// The current device's copy of `scratch` (looked up on the first call for that device).
int scratch_setup(int device, uchar4 **out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return TSI_E_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return TSI_E_HIP;
        g_scratch_ptr[device] = static_cast<uchar4 *>(p);
    }
    *out = g_scratch_ptr[device];
    return TSI_OK;
}

}  // namespace

struct tsi_ctx {
    tsi_params p{};
    int device = 0;
    int num_cu = 256;
    uchar4 *scratch_rgba8 = nullptr;    // this device's copy of the static staging buffer
    hipStream_t stream = nullptr;
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    uint64_t nodes = 0;
    uint32_t nblocks = 0, list_cap = 0, dlist_cap = 0;
    float4 *st = nullptr, *blo = nullptr, *bhi = nullptr, *grad = nullptr;
    uint32_t *lists[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    uint32_t *dlists[2] = {nullptr, nullptr};
    uint32_t *pending = nullptr, *free_stack = nullptr;
    Counters *cnt = nullptr;
    float2 *stack = nullptr;
    size_t stack_bytes = 0;
    FwdIns *fprog = nullptr;
    RevIns *rprog = nullptr;
    // pruned programs
    uint32_t *prune_rec[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    uint32_t *gdesc[2] = {nullptr, nullptr};
    uint32_t *choice = nullptr;
    FwdIns *pool = nullptr;
    uint32_t prune_cap = 0, nchunks = 0, choice_len = 0, pool_cap = 0;
    bool prune = false;
    // speculative evaluation of the shallow levels (KArgs::spec_*): the listed parents, their
    // count, the records, the last frame's tag, the levels asked for
    uint32_t *spec_list = nullptr;
    uint32_t *spec_cnt = nullptr;
    SpecCell *spec_res = nullptr;
    uint32_t spec_cap = 0, spec_tag = 0, spec_max = 0;
    // events around each level's launch (tsi_stats.eval_kernel_ms as the levels' summed kernel
    // time); $TSI_LEVEL_EVENTS=0 leaves them out (eval_kernel_ms is then the levels' span)
    bool level_events = true;
    // The trace's form.  A full launch of 16x16-pixel blocks fills every CU's wave slots while
    // blocks are pending, and the next frame's level kernels, launched beside it, wait for slots;
    // the persistent form on one block per CU leaves them free but, with fewer wavefronts, takes
    // ~15 % longer.  It is chosen while its trace time stays under 1.6 set inversions (the two
    // trace slots then keep up), from the last frame's measured times; $TSI_TRACE_BLOCKS fixes
    // it (0: full launches; n: the persistent form on n blocks per CU).
    int trace_blocks_env = -1;
    bool trace_persistent = false;
    float last_si_ms = 0.0f;
    // the pass runs on its own stream with its own saved-operand stack (one block per CU), after
    // the previous frame's levels up to spec_max have run (spec_ready, recorded by that frame)
    hipStream_t spec_stream = nullptr;
    hipEvent_t spec_ready = nullptr;
    bool spec_ready_recorded = false;
    float2 *spec_stack = nullptr;
    size_t spec_stack_bytes = 0;
    // compiled per-tape evaluation kernels
    hipModule_t mod[2] = {nullptr, nullptr};
    hipFunction_t eval_fn[2] = {nullptr, nullptr};   // forward-only list, reverse list
    tsi_tape_info info{};
    std::vector<tsi_gpu_instruction> tape;
    Lowered low{};
    bool have_tape = false;
    tsi_uniforms u{};
    uint32_t grid_fwd = 0, grid_rev = 0, grid_misc = 0;
    uint64_t last_visits = 0;
    double last_trace_ms = 0.0;
    bool last_trace_persistent = false;
    uint64_t used_nodes = 8;            // nodes in use after the last collected set inversion
    double trace_ms_total = 0.0;
    uint64_t trace_visits_total = 0, trace_count = 0;
    // work in flight: frames (set inversions) and traces, each a ring of slots with their events
    // and pinned host copies of what the host reads back
    struct FrameSlot {
        hipEvent_t begin = nullptr, end = nullptr, eval0 = nullptr, eval1 = nullptr;
        hipEvent_t level[2 * kMaxLevels] = {};   // around each level's evaluation launches
        hipEvent_t spec0 = nullptr, spec1 = nullptr;   // around the speculative pass
        bool spec = false;                             // the frame had one
        Counters *h = nullptr;
    };
    // a trace runs on its slot's stream over its slot's device-side snapshot of the octree, so it
    // overlaps the next frame's set inversion on the main stream and the previous trace's last
    // wavefronts on the other slot's stream
    struct TraceSlot {
        hipEvent_t begin = nullptr, end = nullptr, snap = nullptr;
        unsigned long long *visits_h = nullptr;
        hipStream_t stream = nullptr;
        float4 *st = nullptr, *blo = nullptr, *bhi = nullptr;   // snapshot, voxel_capacity nodes
        float4 *grad = nullptr;                                 // gradient normals only
        float4 *rgba = nullptr;                                 // the frame it writes
        float *depth = nullptr;
        size_t px = 0;
        unsigned long long *visits = nullptr;
        bool persistent = false;   // traced by k_trace_quads
    };
    static constexpr uint32_t kFrameSlots = 2, kTraceSlots = 2;
    FrameSlot fslot[kFrameSlots];
    TraceSlot tslot[kTraceSlots];
    uint64_t f_issued = 0, f_done = 0, t_issued = 0, t_done = 0;
    uint32_t last_tslot = 0;   // the slot of the last trace issued: the frame read back
    // the domain the loaded tape kernels are specialised for (they leave out fix where it is the
    // identity there), the root box of the last reset and the t-windows seen since
    bool jit_specialized = false;
    float root_box[6] = {-5.0f, -5.0f, -5.0f, 5.0f, 5.0f, 5.0f};
    float t_seen_lo = 1.0f, t_seen_hi = 0.0f;   // empty
};

namespace {

#define HIP_OK(x) do { if ((x) != hipSuccess) return TSI_E_HIP; } while (0)

int harvest_trace(tsi_ctx *c);
int harvest_all(tsi_ctx *c);

template <class T>
int dmalloc(T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    if (hipMalloc(reinterpret_cast<void **>(p), count * sizeof(T)) != hipSuccess) {
        *p = nullptr;
        return TSI_E_OOM;
    }
    return TSI_OK;
}

template <class T>
void dfree(T *&p) {
    if (p) hipFree(p);
    p = nullptr;
}

// Whether frames evaluate their shallow levels speculatively (KArgs::spec_*): the compiled
// engine, unpruned, and a depth that leaves a level to take the results.
bool spec_active(const tsi_ctx *c) {
    return c->spec_max > 0 && c->eval_fn[0] != nullptr && !c->prune && c->p.localsubdiv >= 2;
}

KArgs make_args(const tsi_ctx *c) {
    KArgs a{};
    a.st = c->st; a.blo = c->blo; a.bhi = c->bhi; a.grad = c->grad;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) a.lists[i][j] = c->lists[i][j];
    a.dlists[0] = c->dlists[0];
    a.dlists[1] = c->dlists[1];
    a.pending_free = c->pending;
    a.free_stack = c->free_stack;
    a.cnt = c->cnt;
    a.stack = c->stack;
    a.fprog = c->fprog;
    a.rprog = c->rprog;
    a.nfwd = c->low.nfwd;
    a.nrev = c->low.nrev;
    a.stack_len = c->low.stack + 1;
    a.out_slot = c->low.out_slot;
    for (int k = 0; k < 4; ++k) a.arg_slot[k] = c->low.arg_slot[k];
    a.nblocks = c->nblocks;
    a.list_cap = c->list_cap;
    a.dlist_cap = c->dlist_cap;
    a.epsilon = 10.0f * float(std::ldexp(1.0, -c->p.localsubdiv));   // compute.wgsl:3521
    a.delta = c->p.delta;
    a.fast_multiplier = c->p.fast_multiplier;
    a.swept_window = c->p.swept_window;
    a.swept = c->p.swept_volume;
    a.point_sample = c->p.point_sample;
    a.lipschitz = c->p.lipschitz_corrected;
    a.fastenabled = c->p.fastenabled;
    a.force_deriv = (c->p.shading & (TSI_SHADE_NORMALS | TSI_SHADE_CINEMATIC)) && (c->p.shading & TSI_SHADE_GRADIENT) ? 1 : 0;
    a.prune = c->prune ? 1 : 0;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) a.prune_rec[i][j] = c->prune_rec[i][j];
    a.prune_cap = c->prune_cap;
    a.nchunks = c->nchunks;
    a.gdesc[0] = c->gdesc[0];
    a.gdesc[1] = c->gdesc[1];
    a.prune_live_pct = c->p.prune_live_max ? (c->p.prune_live_max < 100u ? c->p.prune_live_max : 100u) : 75u;
    a.choice = c->choice;
    a.choice_len = c->choice_len;
    a.pool = c->pool;
    a.pool_cap = c->pool_cap;
    return a;
}

void release_prune(tsi_ctx *c) {
    for (auto &l : c->prune_rec) for (auto &x : l) dfree(x);
    dfree(c->gdesc[0]); dfree(c->gdesc[1]); dfree(c->choice); dfree(c->pool);
    c->prune = false;
    c->prune_cap = c->nchunks = c->choice_len = c->pool_cap = 0;
}

// Whether a tape is evaluated with pruning under these flags (tsi.h TSI_FLAG_*).
bool prune_wanted(uint32_t flags, const Lowered &L) {
    if (L.prune_class == kPruneOff || (flags & TSI_FLAG_NO_PRUNE)) return false;
    return !(L.prune_class == kPruneApprox && (flags & TSI_FLAG_PRUNE_EXACT_ONLY));
}

// Buffers of pruned evaluation for a lowered tape: liveness records for prune_cap entries of
// each frontier list (two level parities) and per-wavefront MIN/MAX choice columns; for the
// interpreter also group descriptors and the program pool (each level reuses it from 0).
int alloc_prune(tsi_ctx *c, const Lowered &L, const std::vector<FwdIns> &fwd) {
    release_prune(c);
    if (!prune_wanted(c->p.flags, L)) return TSI_OK;
    const bool interp = (c->p.flags & TSI_FLAG_INTERPRET) != 0;
    uint64_t cap = c->p.prune_capacity ? c->p.prune_capacity : (uint64_t(1) << 19);
    if (cap > c->list_cap) cap = c->list_cap;
    cap = (cap + 7) / 8 * 8;
    const uint32_t nchunks = (L.nfwd + 31) / 32;
    uint64_t words = uint64_t(kRecHead + 3 * nchunks) * cap;
    uint32_t choice_len = (L.nminmax + 15) / 16 ? (L.nminmax + 15) / 16 : 1;
    if (!interp) {   // the compiled kernels' record layout (tsi_jit.h)
        const JitLayout Y = jit_layout(fwd, L, true);
        words = uint64_t(Y.record_words) * cap;
        choice_len = Y.mode_words ? Y.mode_words : 1;
    }
    const uint32_t waves = (c->grid_fwd + c->grid_rev) * 4u;   // the compiled level kernel's grid
    // worst case: every group of every list keeps the whole program, twice (fwd + rev)
    const uint64_t per_group = 2ull * ((L.nfwd + 3) / 4 * 4 + 8);
    uint64_t pool = 2ull * (cap / 8) * per_group;
    const uint64_t pool_limit = uint64_t(1) << 27;   // 2 GiB of 16-byte records
    if (pool > pool_limit) pool = pool_limit;
    int rc;
    for (auto &l : c->prune_rec)
        for (auto &x : l)
            if ((rc = dmalloc(&x, words))) { release_prune(c); return rc; }
    if ((rc = dmalloc(&c->choice, size_t(waves) * 64u * choice_len))) {
        release_prune(c);
        return rc;
    }
    if (interp && ((rc = dmalloc(&c->gdesc[0], 2 * (c->list_cap / 8 + 1))) ||
                   (rc = dmalloc(&c->gdesc[1], 2 * (c->list_cap / 8 + 1))) || (rc = dmalloc(&c->pool, pool)))) {
        release_prune(c);
        return rc;
    }
    if (!interp) pool = 0;
    c->prune_cap = uint32_t(cap);
    c->nchunks = nchunks;
    c->choice_len = choice_len;
    c->pool_cap = uint32_t(pool);
    c->prune = true;
    return TSI_OK;
}

void release_jit(tsi_ctx *c) {
    for (int k = 0; k < 2; ++k) {
        if (c->mod[k]) hipModuleUnload(c->mod[k]);
        c->mod[k] = nullptr;
        c->eval_fn[k] = nullptr;
    }
}

// Compiles (or finds cached) the tape's two level kernels and loads them on this device.
// The root box of the reference (main.rs) and a t-range around the frame loop's times (about a
// minute of frames at 1/60): the tape kernels are compiled for it, and recompiled without that
// assumption if a frame leaves it.
const JitDomain kDefaultDomain{{-5.0f, -5.0f, -5.0f, 5.0f, 5.0f, 5.0f}, -2.0f, 64.0f};

// The generated module addresses no stack row, liveness-record word or choice column beyond what
// alloc_prune and tsi_upload_tape allocate (a module that would is refused, not run).
int check_module(const std::string &src, const std::vector<FwdIns> &fwd, const Lowered &L, std::string &log) {
    const JitLayout Y = jit_layout(fwd, L, true);
    const std::string bad = jit_check_bounds(src, L.stack + 1u, Y.record_words, Y.mode_words);
    if (bad.empty()) return TSI_OK;
    log = "generated module out of bounds: " + bad;
    return TSI_E_JIT;
}

int load_jit(tsi_ctx *c, const tsi_gpu_instruction *tape, uint32_t n, const std::vector<FwdIns> &fwd,
             const std::vector<RevIns> &rev, const Lowered &L, const JitDomain *domain) {
    release_jit(c);
    c->jit_specialized = false;
    std::string src, log;
    std::vector<char> code;
    bool hit = false;
    double ms = 0.0;
    int rc = TSI_OK;
    try {
        src = jit_source_level(tape, n, fwd, rev, L, c->prune, domain);
        rc = check_module(src, fwd, L, log);
        if (rc == TSI_OK) rc = jit_compile(src, c->prune, code, log, hit, ms);
    } catch (const std::runtime_error &) {   // the generator cannot lay out this tape
        return TSI_E_JIT;
    } catch (...) {
        return TSI_E_OOM;
    }
    c->info.compile_ms = ms;
    c->info.cache_hit = hit;
    if (rc != TSI_OK) {
        std::strncpy(c->info.log, log.c_str(), sizeof c->info.log - 1);
        c->info.log[sizeof c->info.log - 1] = 0;
        return rc;
    }
    // one kernel runs both lists of a level (eval_fn[1] stays null)
    if (hipModuleLoadData(&c->mod[0], code.data()) != hipSuccess ||
        hipModuleGetFunction(&c->eval_fn[0], c->mod[0], "tsi_tape_eval_level") != hipSuccess) {
        release_jit(c);
        return TSI_E_HIP;
    }
    c->jit_specialized = domain != nullptr;
    c->info.specialized = domain ? 1u : 0u;
    c->info.fix_kept = n - 4u;
    if (domain) {
        const std::vector<char> ff = fix_free(tape, n, fwd, *domain);
        for (uint32_t j = 4; j < n; ++j) c->info.fix_kept -= ff[j] ? 1u : 0u;
    }
    return TSI_OK;
}

void release(tsi_ctx *c) {
    if (c->spec_stream) hipStreamSynchronize(c->spec_stream);
    if (c->stream) hipStreamSynchronize(c->stream);
    for (auto &t : c->tslot)
        if (t.stream) hipStreamSynchronize(t.stream);
    release_jit(c);
    release_prune(c);
    for (auto &f : c->fslot) {
        for (hipEvent_t *e : {&f.begin, &f.end, &f.eval0, &f.eval1, &f.spec0, &f.spec1})
            if (*e) { hipEventDestroy(*e); *e = nullptr; }
        for (hipEvent_t &e : f.level)
            if (e) { hipEventDestroy(e); e = nullptr; }
        if (f.h) { hipHostFree(f.h); f.h = nullptr; }
    }
    for (auto &t : c->tslot) {
        for (hipEvent_t *e : {&t.begin, &t.end, &t.snap})
            if (*e) { hipEventDestroy(*e); *e = nullptr; }
        if (t.visits_h) { hipHostFree(t.visits_h); t.visits_h = nullptr; }
        dfree(t.st); dfree(t.blo); dfree(t.bhi); dfree(t.grad); dfree(t.rgba); dfree(t.depth); dfree(t.visits);
        if (t.stream) { hipStreamDestroy(t.stream); t.stream = nullptr; }
    }
    dfree(c->st); dfree(c->blo); dfree(c->bhi); dfree(c->grad);
    for (auto &l : c->lists) for (auto &x : l) dfree(x);
    dfree(c->dlists[0]); dfree(c->dlists[1]);
    dfree(c->pending); dfree(c->free_stack); dfree(c->cnt); dfree(c->stack);
    dfree(c->spec_list); dfree(c->spec_cnt); dfree(c->spec_res); dfree(c->spec_stack);
    if (c->spec_ready) { hipEventDestroy(c->spec_ready); c->spec_ready = nullptr; }
    if (c->spec_stream) { hipStreamDestroy(c->spec_stream); c->spec_stream = nullptr; }
    dfree(c->fprog); dfree(c->rprog);
    for (auto &e : c->ev) if (e) { hipEventDestroy(e); e = nullptr; }
    if (c->stream) { hipStreamDestroy(c->stream); c->stream = nullptr; }
}

}  // namespace

namespace tsi {

int read_octree(const float4 *st_dev, const float4 *grad_dev, const Counters *cnt_dev, uint64_t nodes, tsi_node *out,
                uint64_t cap, uint64_t *count) {
    Counters h;
    HIP_OK(hipMemcpy(&h, cnt_dev, sizeof h, hipMemcpyDeviceToHost));
    uint64_t used = (uint64_t(h.bump) + 1u) * 8u;
    if (used > nodes) used = nodes;
    std::vector<float4> st, gr;
    try {
        st.resize(used);
        gr.resize(used);
    } catch (...) {
        return TSI_E_OOM;
    }
    HIP_OK(hipMemcpy(st.data(), st_dev, used * sizeof(float4), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(gr.data(), grad_dev, used * sizeof(float4), hipMemcpyDeviceToHost));
    struct Fr { uint32_t node, d, i, j, k; };
    std::vector<Fr> stack{{0, 0, 0, 0, 0}};
    uint64_t n = 0;
    while (!stack.empty()) {
        Fr f = stack.back();
        stack.pop_back();
        if (f.node >= used) return TSI_E_ARG;
        const float4 s = st[f.node];
        uint32_t meta;
        std::memcpy(&meta, &s.w, 4);
        const uint32_t child = meta & 0x7FFFFFFFu;
        if (out && n < cap) {
            tsi_node &o = out[n];
            o.depth = f.d; o.i = f.i; o.j = f.j; o.k = f.k;
            o.f = s.x; o.t = s.y; o.time_err = s.z;
            o.fast = (meta & 0x80000000u) ? 1.0f : 0.0f;
            o.has_children = child != 0;
            o.dx = gr[f.node].x; o.dy = gr[f.node].y; o.dz = gr[f.node].z;
        }
        ++n;
        if (child) {
            for (uint32_t l = 0; l < 8; ++l)
                stack.push_back({(child - 1u) * 8u + l, f.d + 1, 2 * f.i + (l & 1u), 2 * f.j + ((l >> 1) & 1u),
                                 2 * f.k + (l >> 2)});
        }
    }
    if (count) *count = n;
    return TSI_OK;
}

}  // namespace tsi

extern "C" {

void tsi_default_params(tsi_params *p, int32_t localsubdiv) {
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    p->localsubdiv = localsubdiv;          // compute.wgsl:3502 uses 8
    p->delta = 0.01f;
    p->point_sample = 1;
    p->swept_volume = 0;
    p->lipschitz_corrected = 0;
    p->fastenabled = 1;
    p->fast_multiplier = 1.0f;
    p->swept_window = 1.0f / 60.0f;
    p->voxel_capacity = 0;
    p->queue_capacity = 0;
}

const char *tsi_strerror(int code) {
    switch (code) {
    case TSI_OK: return "ok";
    case TSI_E_OOM: return "out of memory";
    case TSI_E_QUEUE_FULL: return "frontier list overflow";
    case TSI_E_POOL_EXHAUSTED: return "voxel pool exhausted";
    case TSI_E_TAPE_TOO_LONG: return "tape longer than 4096 instructions";
    case TSI_E_HIP: return "HIP runtime error";
    case TSI_E_ARG: return "invalid argument";
    case TSI_E_PARSE: return "cannot parse .ll text";
    case TSI_E_NO_TAPE: return "no tape uploaded";
    case TSI_E_UNSUPPORTED: return "tape needs more live values than the register file holds";
    case TSI_E_BUDGET: return "request exceeds the admission budget";
    case TSI_E_JIT: return "the tape's evaluation kernels did not compile";
    case TSI_E_PENDING: return "frames in flight: collect tsi_set_inversion_result first";
    case TSI_E_SCRATCH: return "frame larger than the static staging buffer (TSI_STATIC_RGBA8_MAX_PIXELS)";
    default: return "unknown error";
    }
}

float tsi_camera_zdist(uint32_t height) {
    // compute.wgsl:4592: z = f32(res.y) / tan(radians(vfov) / 2.0), radians per 2432, fov 90.
    const float rad = 90.0f * 0.017453292519943295474f;
    return float(height) / std::tan(rad / 2.0f);
}

int tsi_create(const tsi_params *params, tsi_ctx **out) {
    if (!params || !out) return TSI_E_ARG;
    *out = nullptr;
    if (params->localsubdiv < 1 || params->localsubdiv > kMaxLevels - 4) return TSI_E_ARG;
    if (params->shading & ~kShadeBits) return TSI_E_ARG;
    tsi_ctx *c = new (std::nothrow) tsi_ctx();
    if (!c) return TSI_E_OOM;
    c->p = *params;
    int rc = TSI_OK;
    do {
        if (hipGetDevice(&c->device) != hipSuccess) { rc = TSI_E_HIP; break; }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, c->device) != hipSuccess) { rc = TSI_E_HIP; break; }
        c->num_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
        if ((rc = scratch_setup(c->device, &c->scratch_rgba8))) break;
        bool sok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess &&
                   hipStreamCreateWithFlags(&c->spec_stream, hipStreamNonBlocking) == hipSuccess &&
                   hipEventCreateWithFlags(&c->spec_ready, hipEventDisableTiming) == hipSuccess;
        for (auto &t : c->tslot)
            sok = sok && hipStreamCreateWithFlags(&t.stream, hipStreamNonBlocking) == hipSuccess &&
                  hipEventCreateWithFlags(&t.snap, hipEventDisableTiming) == hipSuccess;
        if (!sok) {
            rc = TSI_E_HIP;
            break;
        }
        bool evok = true;
        for (auto &e : c->ev) evok = evok && hipEventCreate(&e) == hipSuccess;
        for (auto &f : c->fslot) {
            for (hipEvent_t *e : {&f.begin, &f.end, &f.spec1})   // spec1: the frame's level 0 waits for it
                evok = evok && hipEventCreate(e) == hipSuccess;
            // events that only time the levels (nothing waits on them): without the system-scope
            // fence, whose cache writeback and invalidation between two levels cost ~2 % of a frame
            for (hipEvent_t *e : {&f.eval0, &f.eval1, &f.spec0})
                evok = evok && hipEventCreateWithFlags(e, hipEventDisableSystemFence) == hipSuccess;
            for (hipEvent_t &e : f.level)
                evok = evok && hipEventCreateWithFlags(&e, hipEventDisableSystemFence) == hipSuccess;
            evok = evok && hipHostMalloc((void **)&f.h, sizeof(Counters), hipHostMallocDefault) == hipSuccess;
        }
        for (auto &t : c->tslot) {
            for (hipEvent_t *e : {&t.begin, &t.end}) evok = evok && hipEventCreate(e) == hipSuccess;
            evok = evok && hipHostMalloc((void **)&t.visits_h, sizeof(unsigned long long), hipHostMallocDefault) == hipSuccess;
        }
        if (!evok) { rc = TSI_E_HIP; break; }
        uint64_t nodes = params->voxel_capacity ? params->voxel_capacity : (uint64_t(1) << 24);
        nodes = (nodes + 7) / 8 * 8;
        if (nodes < 64) nodes = 64;
        if (nodes / 8 > 0x7FFFFFF0ull) { rc = TSI_E_ARG; break; }
        c->nodes = nodes;
        c->nblocks = uint32_t(nodes / 8);
        uint64_t lc = params->queue_capacity ? params->queue_capacity : nodes / 4;
        if (lc > 0xFFFFFFF0ull) lc = 0xFFFFFFF0ull;
        c->list_cap = uint32_t(lc < 16 ? 16 : lc);
        c->dlist_cap = c->nblocks;
        if ((rc = dmalloc(&c->st, nodes)) || (rc = dmalloc(&c->blo, nodes)) || (rc = dmalloc(&c->bhi, nodes)) ||
            (rc = dmalloc(&c->grad, nodes)))
            break;
        for (auto &l : c->lists)
            for (auto &x : l)
                if (!rc) rc = dmalloc(&x, c->list_cap);
        if (rc) break;
        if (const char *e = std::getenv("TSI_LEVEL_EVENTS")) c->level_events = std::strtol(e, nullptr, 10) != 0;
        if (const char *e = std::getenv("TSI_TRACE_BLOCKS")) {
            const long v = std::strtol(e, nullptr, 10);
            c->trace_blocks_env = v >= 0 && v <= 16 ? int(v) : 0;
        }
        // speculative pass: up to 2^16 listed parents (8 records of 64 B each: 32 MiB)
        c->spec_cap = 1u << 16;
        {
            const char *e = std::getenv("TSI_SPEC_LEVELS");
            const long v = e ? std::strtol(e, nullptr, 10) : kSpecLevelsDefault;
            c->spec_max = (params->flags & TSI_FLAG_NO_SPECULATE) || v < 0 ? 0u : uint32_t(v < kMaxLevels ? v : kMaxLevels);
        }
        if (c->spec_max &&
            ((rc = dmalloc(&c->spec_list, c->spec_cap)) || (rc = dmalloc(&c->spec_cnt, 1)) ||
             (rc = dmalloc(&c->spec_res, size_t(c->spec_cap) * 8u))))
            break;
        if ((rc = dmalloc(&c->dlists[0], c->dlist_cap)) || (rc = dmalloc(&c->dlists[1], c->dlist_cap)) ||
            (rc = dmalloc(&c->pending, c->nblocks)) || (rc = dmalloc(&c->free_stack, c->nblocks)) ||
            (rc = dmalloc(&c->cnt, 1)))
            break;
        for (auto &t : c->tslot)
            if (!rc) rc = dmalloc(&t.visits, 2);   // the visit count, then the persistent trace's quad counter
        if (rc) break;
        if (hipMemsetAsync(c->cnt, 0, sizeof(Counters), c->stream) != hipSuccess) { rc = TSI_E_HIP; break; }
        // level kernels stride over their frontier; 4 blocks (16 wavefronts) per CU for both lists
        // (measured: the reverse list of a large frontier needs 4 wavefronts per SIMD to hide its
        // latency, more only grows the saved-operand stack; the forward-only list is insensitive)
        c->grid_fwd = uint32_t(c->num_cu) * 4u;
        c->grid_rev = uint32_t(c->num_cu) * 4u;
        c->grid_misc = uint32_t(c->num_cu) * 4u;
        if (hipStreamSynchronize(c->stream) != hipSuccess) { rc = TSI_E_HIP; break; }
    } while (0);
    if (rc != TSI_OK) {
        release(c);
        delete c;
        return rc;
    }
    *out = c;
    return TSI_OK;
}

void tsi_destroy(tsi_ctx *c) {
    if (!c) return;
    if (c->stream) hipStreamSynchronize(c->stream);
    release(c);
    delete c;
}

int tsi_upload_tape(tsi_ctx *c, const tsi_gpu_instruction *tape, uint32_t n) {
    if (!c || !tape) return TSI_E_ARG;
    if (c->f_issued != c->f_done) return TSI_E_PENDING;
    if (harvest_all(c)) return TSI_E_HIP;
    HIP_OK(hipStreamSynchronize(c->stream));   // buffers below are replaced
    if (n > 4096) return TSI_E_TAPE_TOO_LONG;
    std::vector<FwdIns> fwd;
    std::vector<RevIns> rev;
    Lowered L;
    int rc;
    try {
        rc = lower(tape, n, fwd, rev, L);
    } catch (...) {
        return TSI_E_OOM;
    }
    if (rc != TSI_OK) return rc;
    HIP_OK(hipStreamSynchronize(c->stream));
    dfree(c->fprog);
    dfree(c->rprog);
    // zero padding records: the interpreter fetches program records ahead of execution
    fwd.resize(program_stride(uint32_t(fwd.size())), FwdIns{});
    rev.resize(program_stride(uint32_t(rev.size())), RevIns{});
    if ((rc = dmalloc(&c->fprog, fwd.size())) || (rc = dmalloc(&c->rprog, rev.size()))) return rc;
    HIP_OK(hipMemcpy(c->fprog, fwd.data(), fwd.size() * sizeof(FwdIns), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c->rprog, rev.data(), rev.size() * sizeof(RevIns), hipMemcpyHostToDevice));
    // per wavefront: L.stack rows of 64 lanes, plus one spare row (the reverse pass loads two
    // intervals per record whatever the record saved)
    // (one per wavefront of the compiled level kernel's grid, which any list may fill)
    size_t need = size_t(c->grid_fwd + c->grid_rev) * 4u * 64u * size_t(L.stack + 1) * sizeof(float2);
    if (need > c->stack_bytes) {
        dfree(c->stack);
        c->stack_bytes = 0;
        if ((rc = dmalloc(&c->stack, need / sizeof(float2)))) return rc;
        c->stack_bytes = need;
    }
    // the speculative pass's own stack: one block of 4 wavefronts per CU
    const size_t spec_need = c->spec_max ? size_t(c->num_cu) * 4u * 64u * size_t(L.stack + 1) * sizeof(float2) : 0;
    if (spec_need > c->spec_stack_bytes) {
        HIP_OK(hipStreamSynchronize(c->spec_stream));
        dfree(c->spec_stack);
        c->spec_stack_bytes = 0;
        if ((rc = dmalloc(&c->spec_stack, spec_need / sizeof(float2)))) return rc;
        c->spec_stack_bytes = spec_need;
    }
    if ((rc = alloc_prune(c, L, fwd))) return rc;
    std::memset(&c->info, 0, sizeof c->info);
    c->info.n = n;
    c->info.stack = L.stack;
    c->info.slots = L.slots;
    c->info.prune_class = uint32_t(L.prune_class);
    if (!(c->p.flags & TSI_FLAG_INTERPRET)) {
        fwd.resize(L.nfwd);
        rev.resize(L.nrev);
        if ((rc = load_jit(c, tape, n, fwd, rev, L, &kDefaultDomain))) {
            c->have_tape = false;
            return rc;
        }
        c->info.engine = 1;
    } else {
        release_jit(c);
    }
    c->tape.assign(tape, tape + n);
    c->low = L;
    c->have_tape = true;
    return TSI_OK;
}

int tsi_reset(tsi_ctx *c, const float root_box[6]) {
    if (!c || !root_box) return TSI_E_ARG;
    if (c->f_issued != c->f_done) return TSI_E_PENDING;
    if (harvest_all(c)) return TSI_E_HIP;
    KArgs a = make_args(c);
    if (launch_reset(a, root_box, 8, c->stream) != 0) return TSI_E_HIP;
    HIP_OK(hipStreamSynchronize(c->stream));
    std::memcpy(c->root_box, root_box, sizeof c->root_box);
    c->t_seen_lo = 1.0f;
    c->t_seen_hi = 0.0f;
    return TSI_OK;
}

int tsi_set_uniforms(tsi_ctx *c, const tsi_uniforms *u) {
    if (!c || !u) return TSI_E_ARG;
    c->u = *u;
    return TSI_OK;
}

int tsi_enqueue_root(tsi_ctx *c) {
    if (!c) return TSI_E_ARG;
    if (launch_enqueue_root(make_args(c), c->stream) != 0) return TSI_E_HIP;
    return TSI_OK;
}

namespace {

// Every cell box lies in the root box and every t-window in the span of the frames' times since
// the reset (each swept window reaching swept_window below its frame's time): when these leave
// the domain the kernels were specialised for, the kernels are rebuilt without the assumption.
int keep_domain(tsi_ctx *c) {
    const float t = c->u.time;
    const float lo = c->p.swept_volume ? t - c->p.swept_window : t;
    if (c->t_seen_lo > c->t_seen_hi) {
        c->t_seen_lo = lo;
        c->t_seen_hi = t;
    } else {
        c->t_seen_lo = std::fmin(c->t_seen_lo, lo);
        c->t_seen_hi = std::fmax(c->t_seen_hi, t);
    }
    if (!c->jit_specialized) return TSI_OK;
    const JitDomain &d = kDefaultDomain;
    bool inside = c->t_seen_lo >= d.t0 && c->t_seen_hi <= d.t1;
    for (int k = 0; k < 3; ++k) inside = inside && c->root_box[k] >= d.box[k] && c->root_box[3 + k] <= d.box[3 + k];
    if (inside) return TSI_OK;
    HIP_OK(hipStreamSynchronize(c->stream));   // frames in flight use the loaded kernels
    for (auto &t : c->tslot) HIP_OK(hipStreamSynchronize(t.stream));
    std::vector<FwdIns> fwd;
    std::vector<RevIns> rev;
    Lowered L;
    int rc = lower(c->tape.data(), uint32_t(c->tape.size()), fwd, rev, L);
    if (rc) return rc;
    fwd.resize(L.nfwd);
    rev.resize(L.nrev);
    return load_jit(c, c->tape.data(), uint32_t(c->tape.size()), fwd, rev, L, nullptr);
}

}  // namespace

int tsi_set_inversion_async(tsi_ctx *c) {
    if (!c) return TSI_E_ARG;
    if (!c->have_tape) return TSI_E_NO_TAPE;
    if (c->f_issued - c->f_done >= tsi_ctx::kFrameSlots) return TSI_E_PENDING;
    if (c->eval_fn[0]) {
        const int rc = keep_domain(c);
        if (rc) return rc;
    }
    tsi_ctx::FrameSlot &f = c->fslot[c->f_issued % tsi_ctx::kFrameSlots];
    KArgs a = make_args(c);
    SpecLaunch sl{};
    f.spec = spec_active(c);
    if (f.spec) {
        a.spec_levels = std::min<uint32_t>(c->spec_max, uint32_t(c->p.localsubdiv - 1));
        a.spec_cap = c->spec_cap;
        if (++c->spec_tag == 0u) c->spec_tag = 1u;
        a.spec_tag = c->spec_tag;
        a.spec_list = c->spec_list;
        a.spec_count = c->spec_cnt;
        a.spec_res = c->spec_res;
        // the pass reads the octree down to depth spec_levels: it starts once the previous frame
        // has run its levels 0 .. spec_max (its last levels run beside it)
        sl.stream = c->spec_stream;
        sl.wait = c->spec_ready_recorded ? c->spec_ready : nullptr;
        sl.begin = f.spec0;
        sl.done = f.spec1;
        sl.ready = c->spec_ready;
        sl.args = a;
        sl.args.stack = c->spec_stack;
        sl.blocks = uint32_t(c->num_cu);
        sl.ready_level = int(c->spec_max);
        c->spec_ready_recorded = true;
    }
    FrameLaunch fl{c->p.localsubdiv, c->u.time, c->grid_fwd, c->grid_rev, c->grid_misc,
                   (void *)c->eval_fn[0], f.spec ? &sl : nullptr};
    HIP_OK(hipEventRecord(f.begin, c->stream));
    if (launch_frame(a, fl, c->stream, f.eval0, f.eval1, c->level_events ? f.level : nullptr) != 0) return TSI_E_HIP;
    HIP_OK(hipMemcpyAsync(f.h, c->cnt, sizeof(Counters), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipEventRecord(f.end, c->stream));
    c->f_issued += 1;
    return TSI_OK;
}

int tsi_set_inversion_result(tsi_ctx *c, double *gpu_ms, tsi_stats *stats) {
    if (!c) return TSI_E_ARG;
    if (c->f_done == c->f_issued) return TSI_E_ARG;   // no frame in flight
    tsi_ctx::FrameSlot &f = c->fslot[c->f_done % tsi_ctx::kFrameSlots];
    c->f_done += 1;
    HIP_OK(hipEventSynchronize(f.end));
    const Counters &h = *f.h;
    float ms = 0.0f, ems = 0.0f;
    hipEventElapsedTime(&ms, f.begin, f.end);
    c->last_si_ms = ms;
    hipEventElapsedTime(&ems, f.eval0, f.eval1);
    if (gpu_ms) *gpu_ms = ms;
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->gpu_ms = ms;
        stats->eval_ms = ems;
        double kms = c->level_events ? 0.0 : double(ems);
        for (int d = 0; c->level_events && d < c->p.localsubdiv; ++d) {
            float x = 0.0f;
            hipEventElapsedTime(&x, f.level[2 * d], f.level[2 * d + 1]);
            kms += x;
        }
        if (f.spec) {   // the speculative pass (on its own stream, beside the previous frame)
            float x = 0.0f;
            hipEventElapsedTime(&x, f.spec0, f.spec1);
            kms += x;
        }
        stats->eval_kernel_ms = kms;
        stats->cell_evals = h.stats[kStatCells];
        stats->rev_evals = h.stats[kStatRevCells];
        stats->fwd_ops = h.stats[kStatFwdOps];
        stats->pruned_instr = h.stats[kStatPruned];
        stats->spec_cells = h.stats[kStatSpec];
        stats->rev_ops = h.stats[kStatRevOps];
        stats->parents = h.stats[kStatParents];
        stats->deallocs = h.stats[kStatDeallocs];
        stats->enqueued = h.stats[kStatEnqueued];
        stats->allocated_nodes = h.stats[kStatLive] * 8u;
        // the compiled engine runs both lists of a level in one launch (and the speculative pass
        // in one more), the interpreter in two
        stats->eval_launches = uint64_t((c->eval_fn[0] ? 1 : 2) * c->p.localsubdiv) + (spec_active(c) ? 1u : 0u);
        stats->prune_launches = c->prune && !c->eval_fn[0] ? uint64_t(2 * (c->p.localsubdiv - 1)) : 0u;
    }
    c->used_nodes = (uint64_t(h.bump) + 1u) * 8u;
    if (c->used_nodes > c->nodes) c->used_nodes = c->nodes;
    if (h.frame_error & kErrPool) return TSI_E_POOL_EXHAUSTED;
    if (h.frame_error & kErrQueue) return TSI_E_QUEUE_FULL;
    return TSI_OK;
}

int tsi_set_inversion(tsi_ctx *c, double *gpu_ms, tsi_stats *stats) {
    if (!c) return TSI_E_ARG;
    if (c->f_issued != c->f_done) return TSI_E_PENDING;
    int rc = tsi_set_inversion_async(c);
    if (rc) return rc;
    return tsi_set_inversion_result(c, gpu_ms, stats);
}

namespace {

// Waits for the oldest trace in flight and records its device time and node visits.
int harvest_trace(tsi_ctx *c) {
    if (c->t_done == c->t_issued) return TSI_OK;
    tsi_ctx::TraceSlot &t = c->tslot[c->t_done % tsi_ctx::kTraceSlots];
    c->t_done += 1;
    HIP_OK(hipEventSynchronize(t.end));
    float tms = 0.0f;
    hipEventElapsedTime(&tms, t.begin, t.end);
    c->last_visits = *t.visits_h;
    c->last_trace_ms = tms;
    c->last_trace_persistent = t.persistent;
    c->trace_ms_total += tms;
    c->trace_visits_total += *t.visits_h;
    c->trace_count += 1;
    return TSI_OK;
}

int harvest_all(tsi_ctx *c) {
    int rc;
    while (c->t_done != c->t_issued)
        if ((rc = harvest_trace(c))) return rc;
    return TSI_OK;
}

}  // namespace

int tsi_trace_async(tsi_ctx *c) {
    if (!c) return TSI_E_ARG;
    const uint32_t W = c->u.resolution[0], H = c->u.resolution[1];
    if (W == 0 || H == 0) return TSI_E_ARG;
    int rc;
    // the slot's previous trace is collected first: its snapshot and frame are then free
    if (c->t_issued - c->t_done >= tsi_ctx::kTraceSlots && (rc = harvest_trace(c))) return rc;
    const uint32_t k = uint32_t(c->t_issued % tsi_ctx::kTraceSlots);
    tsi_ctx::TraceSlot &ts = c->tslot[k];
    const size_t px = size_t(W) * H;
    if (px > ts.px) {
        dfree(ts.rgba);
        dfree(ts.depth);
        ts.px = 0;
        if ((rc = dmalloc(&ts.rgba, px)) || (rc = dmalloc(&ts.depth, px))) return rc;
        ts.px = px;
    }
    if (!ts.st) {
        if ((rc = dmalloc(&ts.st, c->nodes)) || (rc = dmalloc(&ts.blo, c->nodes)) || (rc = dmalloc(&ts.bhi, c->nodes))) {
            dfree(ts.st); dfree(ts.blo); dfree(ts.bhi);
            return rc;
        }
    }
    KArgs a = make_args(c);
    // snapshot of the nodes in use, in stream order after the last set inversion
    const bool gradient_normals = (c->p.shading & (TSI_SHADE_NORMALS | TSI_SHADE_CINEMATIC)) && (c->p.shading & TSI_SHADE_GRADIENT);
    if (gradient_normals && !ts.grad && (rc = dmalloc(&ts.grad, c->nodes))) return rc;
    if (launch_snapshot(a, ts.st, ts.blo, ts.bhi, gradient_normals ? ts.grad : nullptr, c->nodes, c->grid_misc,
                        c->stream) != 0)
        return TSI_E_HIP;
    HIP_OK(hipEventRecord(ts.snap, c->stream));
    HIP_OK(hipStreamWaitEvent(ts.stream, ts.snap, 0));
    TraceArgs t{};
    std::memcpy(t.view, c->u.view, sizeof t.view);
    t.width = W;
    t.height = H;
    t.time = c->u.time;
    t.zdist = tsi_camera_zdist(H);
    t.rgba = ts.rgba;
    t.depth = ts.depth;
    t.visits = ts.visits;
    t.shading = c->p.shading;
    t.grad = ts.grad;
    a.st = ts.st;
    a.blo = ts.blo;
    a.bhi = ts.bhi;
    HIP_OK(hipMemsetAsync(ts.visits, 0, 2 * sizeof(unsigned long long), ts.stream));
    HIP_OK(hipEventRecord(ts.begin, ts.stream));
    uint32_t blocks = 0;
    if (c->trace_blocks_env >= 0) {
        blocks = uint32_t(c->trace_blocks_env) * uint32_t(c->num_cu);
    } else if (c->last_si_ms > 0.0f && c->last_trace_ms > 0.0f) {
        // the persistent form's expected time, from the last trace in either form (+/-10 %
        // hysteresis around the switch)
        const float tp = c->last_trace_persistent ? c->last_trace_ms : 1.15f * c->last_trace_ms;
        const float limit = 1.6f * c->last_si_ms * (c->trace_persistent ? 1.1f : 0.9f);
        c->trace_persistent = tp < limit;
        blocks = c->trace_persistent ? uint32_t(c->num_cu) : 0u;
    }
    ts.persistent = blocks != 0;
    if (launch_trace(a, t, ts.stream, blocks, reinterpret_cast<uint32_t *>(ts.visits + 1)) != 0)
        return TSI_E_HIP;
    HIP_OK(hipMemcpyAsync(ts.visits_h, ts.visits, sizeof(unsigned long long), hipMemcpyDeviceToHost, ts.stream));
    HIP_OK(hipEventRecord(ts.end, ts.stream));
    c->last_tslot = k;
    c->t_issued += 1;
    return TSI_OK;
}

int tsi_trace_wait(tsi_ctx *c, float *rgba_out, float *depth_out) {
    if (!c) return TSI_E_ARG;
    int rc;
    if ((rc = harvest_all(c))) return rc;
    const tsi_ctx::TraceSlot &ts = c->tslot[c->last_tslot];
    const size_t px = size_t(c->u.resolution[0]) * c->u.resolution[1];
    if ((rgba_out || depth_out) && (c->t_issued == 0 || px > ts.px)) return TSI_E_ARG;   // nothing traced at this size
    if (rgba_out) HIP_OK(hipMemcpy(rgba_out, ts.rgba, px * sizeof(float4), hipMemcpyDeviceToHost));
    if (depth_out) HIP_OK(hipMemcpy(depth_out, ts.depth, px * sizeof(float), hipMemcpyDeviceToHost));
    return TSI_OK;
}

int tsi_trace(tsi_ctx *c, float *rgba_out, float *depth_out) {
    int rc = tsi_trace_async(c);
    if (rc) return rc;
    return tsi_trace_wait(c, rgba_out, depth_out);
}

int tsi_trace_totals(tsi_ctx *c, double *trace_ms, uint64_t *node_visits, uint64_t *traces) {
    if (!c) return TSI_E_ARG;
    if (trace_ms) *trace_ms = c->trace_ms_total;
    if (node_visits) *node_visits = c->trace_visits_total;
    if (traces) *traces = c->trace_count;
    return TSI_OK;
}

int tsi_frame_rgba8(tsi_ctx *c, uint8_t *rgba8_out) {
    if (!c || !rgba8_out) return TSI_E_ARG;
    int rc;
    if ((rc = harvest_all(c))) return rc;
    const tsi_ctx::TraceSlot &ts = c->tslot[c->last_tslot];
    const size_t px = size_t(c->u.resolution[0]) * c->u.resolution[1];
    if (px == 0 || c->t_issued == 0 || px > ts.px) return TSI_E_ARG;   // nothing traced at this size
    uchar4 *tmp = nullptr;
    if ((rc = dmalloc(&tmp, px))) return rc;
    rc = launch_rgba8(ts.rgba, tmp, uint32_t(px), c->stream) == 0 ? TSI_OK : TSI_E_HIP;
    if (rc == TSI_OK && hipMemcpyAsync(rgba8_out, tmp, px * 4u, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        rc = TSI_E_HIP;
    if (rc == TSI_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = TSI_E_HIP;
    dfree(tmp);
    return rc;
}

int tsi_frame_rgba8_noalloc(tsi_ctx *c, uint8_t *rgba8_out) {
    if (!c || !rgba8_out) return TSI_E_ARG;
    const size_t px = size_t(c->u.resolution[0]) * c->u.resolution[1];
    if (px > kStaticRgba8Pixels) return TSI_E_SCRATCH;
    if (!c->scratch_rgba8) return TSI_E_HIP;
    int rc;
    if ((rc = harvest_all(c))) return rc;
    const tsi_ctx::TraceSlot &ts = c->tslot[c->last_tslot];
    if (px == 0 || c->t_issued == 0 || px > ts.px) return TSI_E_ARG;   // nothing traced at this size
    // every pixel of the staging buffer that is copied out is written first: nothing to clear
    if (launch_rgba8(ts.rgba, c->scratch_rgba8, uint32_t(px), c->stream) != 0) return TSI_E_HIP;
    HIP_OK(hipMemcpyAsync(rgba8_out, c->scratch_rgba8, px * 4u, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return TSI_OK;
}

int tsi_save_frame_png(tsi_ctx *c, const char *path) {
    if (!c || !path) return TSI_E_ARG;
    const uint32_t W = c->u.resolution[0], H = c->u.resolution[1];
    std::vector<uint8_t> px;
    try {
        px.resize(size_t(W) * H * 4u);
    } catch (...) {
        return TSI_E_OOM;
    }
    int rc = tsi_frame_rgba8(c, px.data());
    if (rc) return rc;
    return tsi_write_png(path, px.data(), W, H);
}

int tsi_frame_device(tsi_ctx *c, const float **rgba_dev, const float **depth_dev) {
    if (!c) return TSI_E_ARG;
    const tsi_ctx::TraceSlot &ts = c->tslot[c->last_tslot];
    if (rgba_dev) *rgba_dev = reinterpret_cast<const float *>(ts.rgba);
    if (depth_dev) *depth_dev = ts.depth;
    return TSI_OK;
}

int tsi_read_tape(tsi_ctx *c, tsi_gpu_instruction *out, uint32_t cap, uint32_t *n) {
    if (!c) return TSI_E_ARG;
    if (!c->have_tape) return TSI_E_NO_TAPE;
    if (n) *n = uint32_t(c->tape.size());
    if (out) {
        uint32_t m = uint32_t(c->tape.size()) < cap ? uint32_t(c->tape.size()) : cap;
        std::memcpy(out, c->tape.data(), m * sizeof(tsi_gpu_instruction));
    }
    return TSI_OK;
}

int tsi_precompile_tape(const tsi_gpu_instruction *tape, uint32_t n, uint32_t flags, tsi_tape_info *out) {
    if (!tape) return TSI_E_ARG;
    if (n > 4096) return TSI_E_TAPE_TOO_LONG;
    std::vector<FwdIns> fwd;
    std::vector<RevIns> rev;
    Lowered L;
    tsi_tape_info info{};
    int rc;
    try {
        rc = lower(tape, n, fwd, rev, L);
        if (rc == TSI_OK) {
            info.n = n;
            info.stack = L.stack;
            info.slots = L.slots;
            info.prune_class = uint32_t(L.prune_class);
            info.engine = 1;
            std::string src, log;
            std::vector<char> code;
            bool hit = false;
            double ms = 0.0;
            const bool prune = prune_wanted(flags, L);
            src = jit_source_level(tape, n, fwd, rev, L, prune, &kDefaultDomain);
            info.specialized = 1;
            info.fix_kept = n - 4u;
            const std::vector<char> ff = fix_free(tape, n, fwd, kDefaultDomain);
            for (uint32_t j = 4; j < n; ++j) info.fix_kept -= ff[j] ? 1u : 0u;
            rc = check_module(src, fwd, L, log);
            if (rc == TSI_OK) rc = jit_compile(src, prune, code, log, hit, ms);
            info.compile_ms = ms;
            info.cache_hit = hit;
            if (rc != TSI_OK) std::strncpy(info.log, log.c_str(), sizeof info.log - 1);
        }
    } catch (const std::runtime_error &) {
        return TSI_E_JIT;
    } catch (...) {
        return TSI_E_OOM;
    }
    if (out) *out = info;
    return rc;
}

int tsi_tape_kernel_source(const tsi_gpu_instruction *tape, uint32_t n, uint32_t flags, char *out, size_t cap,
                           size_t *len, tsi_tape_info *info) {
    if (!tape) return TSI_E_ARG;
    if (n > 4096) return TSI_E_TAPE_TOO_LONG;
    std::vector<FwdIns> fwd;
    std::vector<RevIns> rev;
    Lowered L;
    std::string src;
    try {
        const int rc = lower(tape, n, fwd, rev, L);
        if (rc != TSI_OK) return rc;
        src = jit_source_level(tape, n, fwd, rev, L, prune_wanted(flags, L), &kDefaultDomain);
    } catch (const std::runtime_error &) {
        return TSI_E_JIT;
    } catch (...) {
        return TSI_E_OOM;
    }
    if (info) {
        std::memset(info, 0, sizeof *info);
        info->n = n;
        info->stack = L.stack;
        info->slots = L.slots;
        info->prune_class = uint32_t(L.prune_class);
        info->engine = 1;
        info->specialized = 1;
    }
    if (len) *len = src.size();
    if (out && cap) {
        const size_t m = src.size() < cap - 1 ? src.size() : cap - 1;
        std::memcpy(out, src.data(), m);
        out[m] = 0;
    }
    return TSI_OK;
}

int tsi_last_trace_info(tsi_ctx *c, double *trace_ms, uint64_t *node_visits) {
    if (!c) return TSI_E_ARG;
    if (trace_ms) *trace_ms = c->last_trace_ms;
    if (node_visits) *node_visits = c->last_visits;
    return TSI_OK;
}

int tsi_tape_info_get(tsi_ctx *c, tsi_tape_info *out) {
    if (!c || !out) return TSI_E_ARG;
    *out = c->info;
    return TSI_OK;
}

int tsi_read_nodes(tsi_ctx *c, tsi_node *out, uint64_t cap, uint64_t *count) {
    if (!c) return TSI_E_ARG;
    HIP_OK(hipStreamSynchronize(c->stream));
    return read_octree(c->st, c->grad, c->cnt, c->nodes, out, cap, count);
}

int tsi_set_params(tsi_ctx *c, const tsi_params *p) {
    if (!c || !p) return TSI_E_ARG;
    if (c->f_issued != c->f_done) return TSI_E_PENDING;
    const tsi_params &o = c->p;
    if (p->voxel_capacity != o.voxel_capacity || p->queue_capacity != o.queue_capacity || p->flags != o.flags ||
        p->prune_capacity != o.prune_capacity || p->prune_live_max != o.prune_live_max || p->shading != o.shading)
        return TSI_E_ARG;
    if (p->localsubdiv < 1 || p->localsubdiv > kMaxLevels - 4 || !std::isfinite(p->delta) ||
        !std::isfinite(p->swept_window) || !std::isfinite(p->fast_multiplier))
        return TSI_E_ARG;
    c->p = *p;
    return TSI_OK;
}

}  // extern "C"
