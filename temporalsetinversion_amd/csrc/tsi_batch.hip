// tsi_batch.hip -- the batched and served forms of include/tsi.h: sessions in a device arena, one
// persistent set-inversion launch (tsi_batch.h) and one tile-queue trace launch per submit.
//
// No counterpart in the reference: its host (src/main.rs:724-879) dispatches one scene's `main`
// per frame.  A session holds what a tsi_ctx holds for one scene (node pool, frontier lists,
// dealloc lists, free stack, counters, frame), carved out of one arena allocation, so that
// opening and closing sessions never calls hipMalloc and admission is a question of free arena
// space (the HBM budget).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "tsi.h"
#include "tsi_batch.h"
#include "tsi_jit.h"
#include "tsi_launch.h"
#include "tsi_tape.h"

namespace tsi {
int lower(const tsi_gpu_instruction *tape, uint32_t n, std::vector<FwdIns> &fwd, std::vector<RevIns> &rev,
          Lowered &L);
}  // namespace tsi

using namespace tsi;

namespace {

#define HIP_OK(x) do { if ((x) != hipSuccess) return TSI_E_HIP; } while (0)

constexpr uint64_t kAlign = 256;
constexpr uint32_t kDefaultWidth = 1920, kDefaultHeight = 1080;
// The batch kernels leave out fix() where a host bound over this domain proves it the identity
// (tsi_jit.cpp fix_free): cell boxes inside the reference's root box, t-windows inside [-2, 64].
const JitDomain kBatchDomain{{-5.0f, -5.0f, -5.0f, 5.0f, 5.0f, 5.0f}, -2.0f, 64.0f};

struct Session {
    bool open = false;
    uint32_t tape = 0;
    tsi_params p{};
    float box[6] = {};
    uint64_t off = 0, bytes = 0;       // arena region
    uint64_t nodes = 0;
    uint32_t nblocks = 0, list_cap = 0;
    KArgs a{};
    float4 *rgba = nullptr;
    float *depth = nullptr;
    size_t px_cap = 0;
    uint32_t w = 0, h = 0;             // last traced frame
};

// A session's arena layout: offsets of its arrays from the region's start.
struct Layout {
    uint64_t st, blo, bhi, grad, lists[4], dlists[2], pending, free_stack, cnt, rgba, depth;
    uint64_t spec_list, spec_cnt, spec_res;   // the speculative pass's (spec_cap > 0)
    uint64_t total;
};

uint64_t up(uint64_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

bool session_sizes(const tsi_session_desc &d, uint64_t &nodes, uint32_t &list_cap, size_t &px) {
    nodes = d.params.voxel_capacity ? d.params.voxel_capacity : (uint64_t(1) << 22);
    nodes = (nodes + 7) / 8 * 8;
    if (nodes < 64) nodes = 64;
    if (nodes / 8 > 0x7FFFFFF0ull) return false;
    uint64_t lc = d.params.queue_capacity ? d.params.queue_capacity : nodes / 4;
    if (lc > 0xFFFFFFF0ull) lc = 0xFFFFFFF0ull;
    list_cap = uint32_t(lc < 16 ? 16 : lc);
    const uint32_t W = d.max_width ? d.max_width : kDefaultWidth, H = d.max_height ? d.max_height : kDefaultHeight;
    px = size_t(W) * H;
    return true;
}

Layout layout(uint64_t nodes, uint32_t list_cap, size_t px, uint32_t spec_cap) {
    Layout L{};
    uint64_t o = 0;
    auto take = [&](uint64_t bytes) { const uint64_t at = o; o = up(o + bytes); return at; };
    L.st = take(nodes * 16);
    L.blo = take(nodes * 16);
    L.bhi = take(nodes * 16);
    L.grad = take(nodes * 16);
    for (auto &x : L.lists) x = take(uint64_t(list_cap) * 4);
    const uint64_t nblocks = nodes / 8;
    for (auto &x : L.dlists) x = take(nblocks * 4);
    L.pending = take(nblocks * 4);
    L.free_stack = take(nblocks * 4);
    L.cnt = take(sizeof(Counters));
    L.rgba = take(uint64_t(px) * 16);
    L.depth = take(uint64_t(px) * 4);
    if (spec_cap) {
        L.spec_list = take(uint64_t(spec_cap) * 4);
        L.spec_cnt = take(4);
        L.spec_res = take(uint64_t(spec_cap) * 8u * sizeof(SpecCell));
    }
    L.total = o;
    return L;
}

bool params_ok(const tsi_params &p) {
    return p.localsubdiv >= 1 && p.localsubdiv <= kMaxLevels - 4 && std::isfinite(p.delta) &&
           std::isfinite(p.swept_window) && std::isfinite(p.fast_multiplier) && (p.shading & ~kShadeBits) == 0;
}

}  // namespace

struct tsi_batch {
    int device = 0, num_cu = 256;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;
    tsi_batch_config cfg{};
    tsi_batch_info info{};
    std::vector<std::vector<tsi_gpu_instruction>> tapes;
    std::vector<Lowered> low;
    uint32_t grid = 0, stack_rows = 0, max_bpc = 1, wps = 1;
    // speculative pass of the sessions' shallow levels (KArgs::spec_*): levels asked for
    // ($TSI_SPEC_LEVELS), listed nodes per session, the last submit's tag
    uint32_t spec_max = 0, spec_cap = 0, spec_tag = 0;
    float2 *stack = nullptr;
    BatchSession *sess_d = nullptr, *sess_h = nullptr;
    BatchCtl *ctl = nullptr;
    Counters *out_d = nullptr, *out_h = nullptr;
    TraceJob *jobs_d = nullptr, *jobs_h = nullptr;
    unsigned long long *visits_d = nullptr, *visits_h = nullptr;
    char *arena = nullptr;
    uint64_t arena_bytes = 0, arena_used = 0;
    std::map<uint64_t, uint64_t> free_ranges;   // offset -> bytes
    std::vector<Session> sessions;
    // the submit in flight
    bool pending = false;
    std::vector<tsi_frame_req> reqs;
    std::vector<int32_t> job_of;                // trace job of each request, or -1
    std::vector<uint32_t> order;                // request of each session-table slot
    bool traced = false;
};

namespace {

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

template <class T>
int hmalloc(T **p, size_t count) {
    *p = nullptr;
    if (hipHostMalloc(reinterpret_cast<void **>(p), (count ? count : 1) * sizeof(T), hipHostMallocDefault) !=
        hipSuccess) {
        *p = nullptr;
        return TSI_E_OOM;
    }
    return TSI_OK;
}

template <class T>
void hfree(T *&p) {
    if (p) hipHostFree(p);
    p = nullptr;
}

void release(tsi_batch *b) {
    if (b->stream) hipStreamSynchronize(b->stream);
    if (b->mod) hipModuleUnload(b->mod);
    b->mod = nullptr;
    dfree(b->stack); dfree(b->sess_d); dfree(b->ctl); dfree(b->out_d); dfree(b->jobs_d);
    dfree(b->visits_d); dfree(b->arena);
    hfree(b->sess_h); hfree(b->out_h); hfree(b->jobs_h); hfree(b->visits_h);
    for (auto &e : b->ev)
        if (e) { hipEventDestroy(e); e = nullptr; }
    if (b->stream) { hipStreamDestroy(b->stream); b->stream = nullptr; }
}

// First fit in the arena's free ranges; returns false when nothing fits.
bool arena_take(tsi_batch *b, uint64_t bytes, uint64_t &off) {
    for (auto it = b->free_ranges.begin(); it != b->free_ranges.end(); ++it) {
        if (it->second < bytes) continue;
        off = it->first;
        const uint64_t rest = it->second - bytes;
        b->free_ranges.erase(it);
        if (rest) b->free_ranges[off + bytes] = rest;
        b->arena_used += bytes;
        return true;
    }
    return false;
}

void arena_give(tsi_batch *b, uint64_t off, uint64_t bytes) {
    b->arena_used -= bytes;
    auto it = b->free_ranges.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != b->free_ranges.end() && it->first + it->second == next->first) {
        it->second += next->second;
        b->free_ranges.erase(next);
    }
    if (it != b->free_ranges.begin()) {
        auto prev = std::prev(it);
        if (prev->first + prev->second == it->first) {
            prev->second += it->second;
            b->free_ranges.erase(it);
        }
    }
}

// The kernel arguments of a session (make_args of tsi_runtime.hip for a context).
void session_args(const tsi_batch *b, Session &s, const Layout &L) {
    char *base = b->arena + s.off;
    KArgs a{};
    a.st = reinterpret_cast<float4 *>(base + L.st);
    a.blo = reinterpret_cast<float4 *>(base + L.blo);
    a.bhi = reinterpret_cast<float4 *>(base + L.bhi);
    a.grad = reinterpret_cast<float4 *>(base + L.grad);
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) a.lists[i][j] = reinterpret_cast<uint32_t *>(base + L.lists[2 * i + j]);
    a.dlists[0] = reinterpret_cast<uint32_t *>(base + L.dlists[0]);
    a.dlists[1] = reinterpret_cast<uint32_t *>(base + L.dlists[1]);
    a.pending_free = reinterpret_cast<uint32_t *>(base + L.pending);
    a.free_stack = reinterpret_cast<uint32_t *>(base + L.free_stack);
    a.cnt = reinterpret_cast<Counters *>(base + L.cnt);
    a.stack = nullptr;
    const Lowered &low = b->low[s.tape];
    a.nfwd = low.nfwd;
    a.nrev = low.nrev;
    a.stack_len = b->stack_rows;
    a.out_slot = low.out_slot;
    for (int k = 0; k < 4; ++k) a.arg_slot[k] = low.arg_slot[k];
    a.nblocks = s.nblocks;
    a.list_cap = s.list_cap;
    a.dlist_cap = s.nblocks;
    a.epsilon = 10.0f * float(std::ldexp(1.0, -s.p.localsubdiv));   // compute.wgsl:3521
    a.delta = s.p.delta;
    a.fast_multiplier = s.p.fast_multiplier;
    a.swept_window = s.p.swept_window;
    a.swept = s.p.swept_volume;
    a.point_sample = s.p.point_sample;
    a.lipschitz = s.p.lipschitz_corrected;
    a.fastenabled = s.p.fastenabled;
    a.force_deriv = (s.p.shading & (TSI_SHADE_NORMALS | TSI_SHADE_CINEMATIC)) && (s.p.shading & TSI_SHADE_GRADIENT) ? 1 : 0;
    a.prune = 0;
    if (b->spec_max && s.p.localsubdiv >= 2) {
        a.spec_levels = std::min<uint32_t>(b->spec_max, uint32_t(s.p.localsubdiv - 1));
        a.spec_cap = b->spec_cap;
        a.spec_list = reinterpret_cast<uint32_t *>(base + L.spec_list);
        a.spec_count = reinterpret_cast<uint32_t *>(base + L.spec_cnt);
        a.spec_res = reinterpret_cast<SpecCell *>(base + L.spec_res);
    }
    s.a = a;
    s.rgba = reinterpret_cast<float4 *>(base + L.rgba);
    s.depth = reinterpret_cast<float *>(base + L.depth);
}

int reset_session(tsi_batch *b, Session &s) {
    HIP_OK(hipMemsetAsync(s.a.cnt, 0, sizeof(Counters), b->stream));
    if (launch_reset(s.a, s.box, 8, b->stream) != 0) return TSI_E_HIP;
    return TSI_OK;
}

// Lowers the tapes and compiles (or finds in the code-object cache) the batch module.
int compile_batch(const tsi_gpu_instruction *const *tapes, const uint32_t *lens, uint32_t ntapes,
                  std::vector<std::vector<tsi_gpu_instruction>> &copies, std::vector<Lowered> &low,
                  uint32_t &stack_rows, std::vector<char> &code, tsi_batch_info &info) {
    std::vector<std::vector<FwdIns>> fwd(ntapes);
    std::vector<std::vector<RevIns>> rev(ntapes);
    low.assign(ntapes, Lowered{});
    copies.assign(ntapes, {});
    stack_rows = 0;
    int rc = TSI_OK;
    try {
        for (uint32_t i = 0; i < ntapes; ++i) {
            if (!tapes[i]) return TSI_E_ARG;
            if (lens[i] > 4096) return TSI_E_TAPE_TOO_LONG;
            copies[i].assign(tapes[i], tapes[i] + lens[i]);
            if ((rc = lower(tapes[i], lens[i], fwd[i], rev[i], low[i]))) return rc;
            fwd[i].resize(low[i].nfwd);
            rev[i].resize(low[i].nrev);
            stack_rows = std::max(stack_rows, low[i].stack + 1u);
        }
        std::vector<BatchTape> bt;
        for (uint32_t i = 0; i < ntapes; ++i)
            bt.push_back(BatchTape{copies[i].data(), lens[i], &fwd[i], &rev[i], &low[i]});
        const std::string src = jit_source_batch(bt, &kBatchDomain);
        std::string log;
        bool hit = false;
        double ms = 0.0;
        rc = jit_compile(src, false, code, log, hit, ms);
        info.compile_ms = ms;
        info.cache_hit = hit ? 1u : 0u;
        info.tapes = ntapes;
        info.stack_rows = stack_rows;
        if (rc != TSI_OK) std::strncpy(info.log, log.c_str(), sizeof info.log - 1);
    } catch (const std::runtime_error &) {
        return TSI_E_JIT;
    } catch (...) {
        return TSI_E_OOM;
    }
    return rc;
}

}  // namespace

extern "C" {

int tsi_batch_precompile(const tsi_gpu_instruction *const *tapes, const uint32_t *lens, uint32_t ntapes,
                         tsi_batch_info *out) {
    if (!tapes || !lens || ntapes == 0 || ntapes > 256) return TSI_E_ARG;
    std::vector<std::vector<tsi_gpu_instruction>> copies;
    std::vector<Lowered> low;
    uint32_t rows = 0;
    std::vector<char> code;
    tsi_batch_info info{};
    const int rc = compile_batch(tapes, lens, ntapes, copies, low, rows, code, info);
    if (out) *out = info;
    return rc;
}

int tsi_batch_create(const tsi_gpu_instruction *const *tapes, const uint32_t *lens, uint32_t ntapes,
                     const tsi_batch_config *cfg, tsi_batch **out) {
    if (!tapes || !lens || !out || ntapes == 0 || ntapes > 256) return TSI_E_ARG;
    *out = nullptr;
    tsi_batch *b = new (std::nothrow) tsi_batch();
    if (!b) return TSI_E_OOM;
    if (cfg) b->cfg = *cfg;
    tsi_batch_config &C = b->cfg;
    if (C.max_sessions == 0 || C.max_sessions > kBatchMaxSessions) C.max_sessions = kBatchMaxSessions;
    if (C.waves_per_simd == 0) C.waves_per_simd = 2;
    if (C.waves_per_simd > 4) C.waves_per_simd = 4;
    if (C.claim_units == 0) C.claim_units = 1;
    {
        const char *e = std::getenv("TSI_SPEC_LEVELS");
        const long v = e ? std::strtol(e, nullptr, 10) : kSpecLevelsDefault;
        b->spec_max = v <= 0 ? 0u : uint32_t(v < kMaxLevels ? v : kMaxLevels);
        b->spec_cap = b->spec_max ? (1u << 13) : 0u;   // 8 192 nodes: 4 MiB of records per session
    }
    int rc = TSI_OK;
    do {
        // ---- tapes: lowering, then the module of all of them ----
        std::vector<char> code;
        if ((rc = compile_batch(tapes, lens, ntapes, b->tapes, b->low, b->stack_rows, code, b->info))) break;
        if (hipModuleLoadData(&b->mod, code.data()) != hipSuccess ||
            hipModuleGetFunction(&b->fn, b->mod, "tsi_batch_run") != hipSuccess) {
            rc = TSI_E_HIP;
            break;
        }
        // ---- device, grid, buffers ----
        if (hipGetDevice(&b->device) != hipSuccess) { rc = TSI_E_HIP; break; }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, b->device) != hipSuccess) { rc = TSI_E_HIP; break; }
        b->num_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
        // one block per CU of 4 x waves_per_simd wavefronts: the most that fit a CU is how
        // many 256-thread blocks would (each is one wavefront per SIMD), at most 4 (the kernel's
        // launch bound); the stack is sized for that (tsi_batch_tune may change the block)
        int per_cu = 0;
        if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, b->fn, 256, 0) != hipSuccess || per_cu < 1) {
            rc = TSI_E_HIP;
            break;
        }
        b->max_bpc = std::min<uint32_t>(4u, uint32_t(per_cu));
        const uint32_t bpc = std::min<uint32_t>(C.waves_per_simd, b->max_bpc);
        b->wps = bpc;
        b->grid = uint32_t(b->num_cu);
        if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) { rc = TSI_E_HIP; break; }
        for (auto &e : b->ev)
            if (hipEventCreate(&e) != hipSuccess) { rc = TSI_E_HIP; break; }
        if (rc) break;
        const size_t stack_elems = size_t(b->num_cu) * b->max_bpc * 4u * 64u * b->stack_rows;
        if ((rc = dmalloc(&b->stack, stack_elems)) || (rc = dmalloc(&b->sess_d, C.max_sessions)) ||
            (rc = dmalloc(&b->ctl, 1)) ||
            (rc = dmalloc(&b->out_d, C.max_sessions)) || (rc = dmalloc(&b->jobs_d, C.max_sessions)) ||
            (rc = dmalloc(&b->visits_d, C.max_sessions)) || (rc = hmalloc(&b->sess_h, C.max_sessions)) ||
            (rc = hmalloc(&b->out_h, C.max_sessions)) || (rc = hmalloc(&b->jobs_h, C.max_sessions)) ||
            (rc = hmalloc(&b->visits_h, C.max_sessions)))
            break;
        uint64_t arena = C.arena_bytes;
        if (arena == 0) {
            size_t fr = 0, tot = 0;
            if (hipMemGetInfo(&fr, &tot) != hipSuccess) { rc = TSI_E_HIP; break; }
            arena = fr / 2;
        }
        arena = arena / kAlign * kAlign;
        if ((rc = dmalloc(&b->arena, arena))) break;
        b->arena_bytes = arena;
        b->free_ranges[0] = arena;
        b->sessions.resize(C.max_sessions);
        b->info.tapes = ntapes;
        b->info.grid_blocks = b->grid;
        b->info.waves_per_simd = bpc;
        b->info.stack_rows = b->stack_rows;
        b->info.arena_bytes = arena;
        b->info.stack_bytes = stack_elems * sizeof(float2);
    } while (0);
    if (rc != TSI_OK) {
        release(b);
        delete b;
        return rc;
    }
    *out = b;
    return TSI_OK;
}

void tsi_batch_destroy(tsi_batch *b) {
    if (!b) return;
    release(b);
    delete b;
}

int tsi_batch_tune(tsi_batch *b, const tsi_batch_config *cfg) {
    if (!b || !cfg) return TSI_E_ARG;
    if (b->pending) return TSI_E_PENDING;
    if (cfg->waves_per_simd > b->max_bpc || cfg->claim_units > 64u) return TSI_E_ARG;
    tsi_batch_config &C = b->cfg;
    if (cfg->waves_per_simd) {
        C.waves_per_simd = cfg->waves_per_simd;
        b->wps = cfg->waves_per_simd;
        b->info.waves_per_simd = cfg->waves_per_simd;
    }
    if (cfg->claim_units) C.claim_units = cfg->claim_units;
    return TSI_OK;
}

int tsi_batch_info_get(tsi_batch *b, tsi_batch_info *out) {
    if (!b || !out) return TSI_E_ARG;
    b->info.arena_used = b->arena_used;
    uint32_t open = 0;
    for (const Session &s : b->sessions) open += s.open ? 1u : 0u;
    b->info.sessions_open = open;
    b->info.max_waves_per_simd = b->max_bpc;
    b->info.claim_units = b->cfg.claim_units;
    *out = b->info;
    return TSI_OK;
}

int tsi_batch_footprint(tsi_batch *b, const tsi_session_desc *d, uint64_t *bytes) {
    if (!b || !d || !bytes) return TSI_E_ARG;
    uint64_t nodes;
    uint32_t lc;
    size_t px;
    if (!session_sizes(*d, nodes, lc, px)) return TSI_E_ARG;
    *bytes = layout(nodes, lc, px, b->spec_cap).total;
    return TSI_OK;
}

int tsi_batch_open(tsi_batch *b, const tsi_session_desc *d, uint32_t *session) {
    if (!b || !d || !session) return TSI_E_ARG;
    if (d->tape >= b->tapes.size() || !params_ok(d->params)) return TSI_E_ARG;
    for (int k = 0; k < 3; ++k)
        if (!(d->root_box[k] >= kBatchDomain.box[k] && d->root_box[3 + k] <= kBatchDomain.box[3 + k] &&
              d->root_box[k] < d->root_box[3 + k]))
            return TSI_E_ARG;
    uint64_t nodes;
    uint32_t lc;
    size_t px;
    if (!session_sizes(*d, nodes, lc, px)) return TSI_E_ARG;
    uint32_t slot = 0;
    while (slot < b->sessions.size() && b->sessions[slot].open) ++slot;
    if (slot == b->sessions.size()) return TSI_E_BUDGET;
    const Layout L = layout(nodes, lc, px, b->spec_cap);
    uint64_t off = 0;
    if (!arena_take(b, L.total, off)) return TSI_E_BUDGET;
    Session &s = b->sessions[slot];
    s = Session{};
    s.tape = d->tape;
    s.p = d->params;
    std::memcpy(s.box, d->root_box, sizeof s.box);
    s.off = off;
    s.bytes = L.total;
    s.nodes = nodes;
    s.nblocks = uint32_t(nodes / 8);
    s.list_cap = lc;
    s.px_cap = px;
    session_args(b, s, L);
    int rc = reset_session(b, s);
    if (rc) {
        arena_give(b, off, L.total);
        return rc;
    }
    s.open = true;
    *session = slot;
    return TSI_OK;
}

int tsi_batch_close(tsi_batch *b, uint32_t session) {
    if (!b || session >= b->sessions.size() || !b->sessions[session].open) return TSI_E_ARG;
    if (b->pending) return TSI_E_PENDING;
    Session &s = b->sessions[session];
    HIP_OK(hipStreamSynchronize(b->stream));   // work queued for it (its reset) has finished
    arena_give(b, s.off, s.bytes);
    s = Session{};
    return TSI_OK;
}

int tsi_batch_reset(tsi_batch *b, uint32_t session) {
    if (!b || session >= b->sessions.size() || !b->sessions[session].open) return TSI_E_ARG;
    if (b->pending) return TSI_E_PENDING;
    return reset_session(b, b->sessions[session]);
}

int tsi_batch_submit(tsi_batch *b, const tsi_frame_req *reqs, uint32_t n) {
    if (!b || (!reqs && n)) return TSI_E_ARG;
    if (b->pending) return TSI_E_PENDING;
    if (n == 0 || n > b->sessions.size()) return TSI_E_ARG;
    std::vector<char> seen(b->sessions.size(), 0);
    int32_t levels = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const tsi_frame_req &r = reqs[i];
        if (r.session >= b->sessions.size() || !b->sessions[r.session].open || seen[r.session]) return TSI_E_ARG;
        seen[r.session] = 1;
        const Session &s = b->sessions[r.session];
        // the kernels' domain: t-windows inside [-2, 64]
        const float lo = s.p.swept_volume ? r.time - s.p.swept_window : r.time;
        if (!(lo >= kBatchDomain.t0 && r.time <= kBatchDomain.t1)) return TSI_E_ARG;
        if (size_t(r.width) * r.height > s.px_cap) return TSI_E_ARG;
        levels = std::max(levels, s.p.localsubdiv);
    }
    b->reqs.assign(reqs, reqs + n);
    b->job_of.assign(n, -1);
    // the kernel's queue takes the sessions in table order: sessions of one tape next to each
    // other, so that wavefronts running at the same time mostly run the same tape's code
    // (instruction cache); slot j of the table holds request order[j]
    b->order.resize(n);
    for (uint32_t i = 0; i < n; ++i) b->order[i] = i;
    std::stable_sort(b->order.begin(), b->order.end(), [&](uint32_t x, uint32_t y) {
        return b->sessions[reqs[x].session].tape < b->sessions[reqs[y].session].tape;
    });
    uint32_t njobs = 0;
    bool any_spec = false;
    for (uint32_t j = 0; j < n; ++j) {
        const uint32_t i = b->order[j];
        const tsi_frame_req &r = reqs[i];
        Session &s = b->sessions[r.session];
        BatchSession &S = b->sess_h[j];
        std::memset(&S, 0, sizeof S);
        S.a = s.a;
        S.a.spec_tag = b->spec_tag + 1u ? b->spec_tag + 1u : 1u;
        any_spec = any_spec || S.a.spec_levels != 0;
        S.time = r.time;
        S.tape = s.tape;
        S.levels = s.p.localsubdiv;
        S.reset = (r.flags & TSI_FRAME_RESET) ? 1u : 0u;
        std::memcpy(S.box, s.box, sizeof S.box);
        if (r.width && r.height) {
            TraceJob &J = b->jobs_h[njobs];
            std::memset(&J, 0, sizeof J);
            J.st = s.a.st;
            J.blo = s.a.blo;
            J.bhi = s.a.bhi;
            J.grad = s.a.grad;
            std::memcpy(J.t.view, r.view, sizeof J.t.view);
            J.t.width = r.width;
            J.t.height = r.height;
            J.t.time = r.time;
            J.t.zdist = tsi_camera_zdist(r.height);
            J.t.rgba = s.rgba;
            J.t.depth = s.depth;
            J.t.visits = b->visits_d + njobs;
            J.t.shading = s.p.shading;
            J.t.grad = s.a.grad;
            J.blocks_x = (r.width + 15u) / 16u;
            b->job_of[i] = int32_t(njobs);
            s.w = r.width;
            s.h = r.height;
            ++njobs;
        }
    }
    HIP_OK(hipMemcpyAsync(b->sess_d, b->sess_h, n * sizeof(BatchSession), hipMemcpyHostToDevice, b->stream));
    if (launch_fill_u32(reinterpret_cast<uint32_t *>(b->ctl), 0u, sizeof(BatchCtl) / 4, b->stream)) return TSI_E_HIP;
    BatchArgs args{};
    args.sess = b->sess_d;
    args.ctl = b->ctl;
    args.out = b->out_d;
    args.stack = b->stack;
    args.stack_len = b->stack_rows;
    args.nsess = n;
    args.phases = levels + 2;
    args.first_phase = any_spec ? -1 : 0;
    b->spec_tag = b->spec_tag + 1u ? b->spec_tag + 1u : 1u;
    args.claim_units = b->cfg.claim_units;
    args.wait_ticks = uint64_t(100000000) * 10u;   // 10 s at the 100 MHz real-time counter
    void *kp[] = {&args};
    HIP_OK(hipEventRecord(b->ev[0], b->stream));
    HIP_OK(hipModuleLaunchCooperativeKernel(b->fn, b->grid, 1, 1, 256u * b->wps, 1, 1, 0, b->stream, kp));
    HIP_OK(hipEventRecord(b->ev[1], b->stream));
    b->traced = njobs > 0;
    if (njobs) {
        // cinematic frames after the others, each group with its own launch (the path tracer's
        // kernel holds more registers; the plain trace keeps its own)
        TraceJob *jb = b->jobs_h;
        std::stable_partition(jb, jb + njobs, [](const TraceJob &J) { return !(J.t.shading & TSI_SHADE_CINEMATIC); });
        uint32_t nplain = 0, bplain = 0, bcin = 0;
        for (uint32_t j = 0; j < njobs; ++j) {
            const bool cin = (jb[j].t.shading & TSI_SHADE_CINEMATIC) != 0;
            uint32_t &acc = cin ? bcin : bplain;
            jb[j].block0 = acc;
            acc += jb[j].blocks_x * ((jb[j].t.height + 15u) / 16u);
            nplain += cin ? 0u : 1u;
        }
        HIP_OK(hipMemcpyAsync(b->jobs_d, b->jobs_h, njobs * sizeof(TraceJob), hipMemcpyHostToDevice, b->stream));
        if (launch_fill_u32(reinterpret_cast<uint32_t *>(b->visits_d), 0u, 2u * njobs, b->stream)) return TSI_E_HIP;
        HIP_OK(hipEventRecord(b->ev[2], b->stream));
        if (nplain && launch_trace_batch(b->jobs_d, nplain, bplain, false, b->stream)) return TSI_E_HIP;
        if (nplain < njobs && launch_trace_batch(b->jobs_d + nplain, njobs - nplain, bcin, true, b->stream))
            return TSI_E_HIP;
        HIP_OK(hipMemcpyAsync(b->visits_h, b->visits_d, njobs * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                              b->stream));
    }
    HIP_OK(hipMemcpyAsync(b->out_h, b->out_d, n * sizeof(Counters), hipMemcpyDeviceToHost, b->stream));
    HIP_OK(hipEventRecord(b->ev[3], b->stream));
    b->pending = true;
    return TSI_OK;
}

int tsi_batch_query(tsi_batch *b) {
    if (!b) return TSI_E_ARG;
    if (!b->pending) return TSI_OK;
    const hipError_t e = hipEventQuery(b->ev[3]);
    if (e == hipSuccess) return TSI_OK;
    return e == hipErrorNotReady ? TSI_E_PENDING : TSI_E_HIP;
}

int tsi_batch_wait(tsi_batch *b, tsi_frame_result *results, double *set_ms, double *trace_ms) {
    if (!b) return TSI_E_ARG;
    if (!b->pending) return TSI_E_ARG;
    b->pending = false;
    HIP_OK(hipEventSynchronize(b->ev[3]));
    float sms = 0.0f, tms = 0.0f;
    hipEventElapsedTime(&sms, b->ev[0], b->ev[1]);
    if (b->traced) hipEventElapsedTime(&tms, b->ev[2], b->ev[3]);
    if (set_ms) *set_ms = sms;
    if (trace_ms) *trace_ms = tms;
    // a barrier that timed out abandons the whole run
    BatchCtl ctl;
    HIP_OK(hipMemcpy(&ctl, b->ctl, sizeof ctl, hipMemcpyDeviceToHost));
    const uint32_t n = uint32_t(b->reqs.size());
    for (uint32_t j = 0; j < n; ++j) {
        const uint32_t i = b->order[j];
        const Counters &h = b->out_h[j];
        if (!results) continue;
        tsi_frame_result &r = results[i];
        std::memset(&r, 0, sizeof r);
        r.status = (h.frame_error & kErrPool) ? TSI_E_POOL_EXHAUSTED
                 : (h.frame_error & kErrQueue) ? TSI_E_QUEUE_FULL : TSI_OK;
        if (ctl.error[0]) r.status = TSI_E_HIP;
        r.stats.gpu_ms = sms;
        r.stats.eval_ms = sms;
        r.stats.cell_evals = h.stats[kStatCells];
        r.stats.rev_evals = h.stats[kStatRevCells];
        r.stats.fwd_ops = h.stats[kStatFwdOps];
        r.stats.rev_ops = h.stats[kStatRevOps];
        r.stats.pruned_instr = h.stats[kStatPruned];
        r.stats.spec_cells = h.stats[kStatSpec];
        r.stats.parents = h.stats[kStatParents];
        r.stats.deallocs = h.stats[kStatDeallocs];
        r.stats.enqueued = h.stats[kStatEnqueued];
        r.stats.allocated_nodes = h.stats[kStatLive] * 8u;
        r.stats.eval_launches = 1;
        r.stats.eval_kernel_ms = sms;
        if (b->job_of[i] >= 0) r.node_visits = b->visits_h[b->job_of[i]];
    }
    return ctl.error[0] ? TSI_E_HIP : TSI_OK;
}

int tsi_batch_read_nodes(tsi_batch *b, uint32_t session, tsi_node *out, uint64_t cap, uint64_t *count) {
    if (!b || session >= b->sessions.size() || !b->sessions[session].open) return TSI_E_ARG;
    if (b->pending) return TSI_E_PENDING;
    const Session &s = b->sessions[session];
    HIP_OK(hipStreamSynchronize(b->stream));
    return read_octree(s.a.st, s.a.grad, s.a.cnt, s.nodes, out, cap, count);
}

int tsi_batch_frame(tsi_batch *b, uint32_t session, float *rgba_out, float *depth_out) {
    if (!b || session >= b->sessions.size() || !b->sessions[session].open) return TSI_E_ARG;
    if (b->pending) return TSI_E_PENDING;
    const Session &s = b->sessions[session];
    if (s.w == 0 || s.h == 0) return TSI_E_ARG;   // nothing traced yet
    HIP_OK(hipStreamSynchronize(b->stream));
    const size_t px = size_t(s.w) * s.h;
    if (rgba_out) HIP_OK(hipMemcpy(rgba_out, s.rgba, px * sizeof(float4), hipMemcpyDeviceToHost));
    if (depth_out) HIP_OK(hipMemcpy(depth_out, s.depth, px * sizeof(float), hipMemcpyDeviceToHost));
    return TSI_OK;
}

}  // extern "C"
