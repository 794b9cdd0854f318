// tsi_kernels.hip -- CDNA4 (gfx950) kernels of the temporal set-inversion path.
//
// The reference runs one persistent WGSL kernel fed by a lock-free broker queue, one
// parent voxel per 8-thread workgroup, with every lane keeping a private 131 KB tape in
// global memory (compute.wgsl:3513-4056).  Here one frame is a fixed sequence of launches
// over a level-synchronous frontier:
//   k_frame_begin            seed the root (buf_main / first-frame seeding)
//   k_eval<REV> x levels     one 64-lane wavefront = 8 parents x 8 child octants; the tape
//                            is interpreted wave-uniformly (instructions via scalar loads,
//                            live values in a uniformly indexed VGPR file), then the cell is
//                            classified; subdivision / dealloc outputs are compacted with
//                            wavefront ballots + one atomic per wave and list
//   k_dealloc x levels       the dealloc cascade of compute.wgsl:3533-3543 / 4036-4044
//   k_frame_end/commit       free-list merge
// Frontier lists are split by whether the children will need the reverse pass, so the
// forward-only kernel carries no reverse-mode code or stack traffic.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "tsi.h"
#include "tsi_eval.h"
#include "tsi_interp.h"
#include "tsi_launch.h"

namespace tsi {

// ---------------------------------------------------------------- frame kernels

// One level of one frontier list with the interpreter (eval_level, tsi_eval.h).
template <bool REV>
__global__ void __launch_bounds__(256) k_eval(KArgs a, int level, float time) {
    eval_level<REV, InterpEngine>(a, level, time);
}

__global__ void k_frame_begin(KArgs a) { frame_begin(a, threadIdx.x, blockDim.x); }

// The speculative pass's node list (spec_gather, tsi_eval.h): one block.
__global__ void __launch_bounds__(1024) k_spec_gather(KArgs a) { spec_gather(a, threadIdx.x, blockDim.x); }
// every parent decided.  Groups with an entry lacking a record run the full program.
__global__ void __launch_bounds__(256) k_prune(KArgs a, int level, int which) {
    Counters *cnt = a.cnt;
    const uint32_t count = min(cnt->list_count[level][which], a.list_cap);
    if (count == 0) return;
    const uint32_t lane = lane_id();
    const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    const uint32_t *prec = a.prune_rec[level & 1][which];
    uint32_t *desc = a.gdesc[which];
    const uint32_t nc = a.nchunks, cap = a.prune_cap;
    const uint32_t ngroups = (count + 7u) / 8u;
    const uint32_t *ffull = reinterpret_cast<const uint32_t *>(a.fprog);
    const uint32_t *rfull = reinterpret_cast<const uint32_t *>(a.rprog);
    for (uint32_t g = wave; g < ngroups; g += nwaves) {
        const uint32_t e0 = g * 8u;
        const uint32_t ne = count - e0 < 8u ? count - e0 : 8u;
        if (e0 + ne > a.prune_cap) {
            if (lane == 0) { desc[2u * g] = 0u; desc[2u * g + 1u] = 0u; }
            continue;
        }
        // union of the entries' live words; lane t holds chunks t and t + 64
        uint32_t u0 = 0, u1 = 0;
        const uint32_t c0 = lane, c1 = lane + 64u;
        for (uint32_t i = 0; i < ne; ++i) {
            const uint32_t *r = prec + size_t(kRecHead) * cap + e0 + i;
            if (c0 < nc) u0 |= r[size_t(c0) * cap];
            if (c1 < nc) u1 |= r[size_t(c1) * cap];
        }
        // exclusive prefix of per-chunk record counts, in chunk order
        const uint32_t n0 = __popc(u0), n1 = __popc(u1);
        uint32_t s0 = n0, s1 = n1;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y0 = uint32_t(__shfl_up(int(s0), off, 64)), y1 = uint32_t(__shfl_up(int(s1), off, 64));
            if (lane >= uint32_t(off)) { s0 += y0; s1 += y1; }
        }
        const uint32_t tot0 = bcast(s0, 63), tot1 = bcast(s1, 63);
        const uint32_t m = tot0 + tot1;
        const uint32_t off0 = s0 - n0, off1 = tot0 + s1 - n1;
        const uint32_t pad = program_stride(m);
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&cnt->pool_used[level], 2u * pad);
        base = bcast(base, 0);
        if (base + 2u * pad > a.pool_cap) {
            if (lane == 0) { desc[2u * g] = 0u; desc[2u * g + 1u] = 0u; }
            continue;
        }
        uint32_t *fw = reinterpret_cast<uint32_t *>(a.pool + base);
        uint32_t *rw = reinterpret_cast<uint32_t *>(a.pool + base + pad);
        for (int half = 0; half < 2; ++half) {
            const uint32_t c = half ? c1 : c0;
            uint32_t u = half ? u1 : u0, at = half ? off1 : off0;
            if (c >= nc) continue;
            uint32_t Lw[8], Rw[8], Vw[8];
#pragma unroll
            for (uint32_t i = 0; i < 8; ++i) {
                const uint32_t *r = prec + size_t(kRecHead + c) * cap + e0 + (i < ne ? i : 0u);
                Vw[i] = i < ne ? r[0] : 0u;
                Lw[i] = i < ne ? r[size_t(nc) * cap] : 0u;
                Rw[i] = i < ne ? r[size_t(2u * nc) * cap] : 0u;
            }
            while (u) {
                const uint32_t b = __builtin_ctz(u);
                u &= u - 1u;
                const uint32_t k = c * 32u + b;
                uint32_t modes = 0;
#pragma unroll
                for (uint32_t i = 0; i < 8; ++i) {
                    const uint32_t live = (Vw[i] >> b) & 1u;
                    const uint32_t md = live ? (((Lw[i] >> b) & 1u) | (((Rw[i] >> b) & 1u) << 1)) : 3u;
                    modes |= md << (2u * i);
                }
                const uint32_t *f = ffull + size_t(k) * 4u;
                uint32_t *fo = fw + size_t(at) * 4u;
                fo[0] = f[0];
                fo[1] = f[1];
                fo[2] = f[2];
                fo[3] = (f[3] & 0xFFFFu) | (modes << 16);
                const uint32_t *rr = rfull + size_t(a.nfwd - 1u - k) * 4u;
                uint32_t *ro = rw + size_t(m - 1u - at) * 4u;
                ro[0] = rr[0];
                ro[1] = rr[1];
                ro[2] = rr[2];
                ro[3] = (rr[3] & 0xFFFFu) | (modes << 16);
                ++at;
            }
        }
        // zero padding records after both programs (the interpreter prefetches past the end)
        for (uint32_t i = m * 4u + lane; i < pad * 4u; i += 64u) {
            fw[i] = 0u;
            rw[i] = 0u;
        }
        if (lane == 0) { desc[2u * g] = base; desc[2u * g + 1u] = m; }
    }
}

// The dealloc cascade's steps after the last level (dealloc_step, tsi_eval.h).
__global__ void __launch_bounds__(256) k_dealloc(KArgs a, int step) {
    const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    uint64_t freed = wave_sum64(dealloc_step(a, step, wave, nwaves));
    if (lane_id() == 0 && freed) atomicAdd((unsigned long long *)&a.cnt->stats[kStatDeallocs], (unsigned long long)freed);
}

// Free-list merge and counter commit at the end of a frame (frame_end / frame_commit, tsi_eval.h).
__global__ void __launch_bounds__(256) k_frame_end(KArgs a) {
    frame_end(a, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

__global__ void k_frame_commit(KArgs a) { frame_commit(a); }

__global__ void k_clear_nodes(KArgs a, uint64_t n) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
        a.st[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

__global__ void k_reset(KArgs a, float x0, float y0, float z0, float x1, float y1, float z1) {
    Counters *c = a.cnt;
    a.st[0] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);   // main.rs:749-763: f=0, t=0, no children
    a.blo[0] = make_float4(x0, y0, z0, 0.0f);
    a.bhi[0] = make_float4(x1, y1, z1, 0.0f);
    a.grad[0] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    c->root_pending = 1;
    c->alloc_tickets = 0;
    c->free_top = 0;
    c->bump = 0;
    c->pending = 0;
    c->error = 0;
    c->frame_error = 0;
}

__global__ void k_enqueue_root(KArgs a) { a.cnt->root_pending = 1; }

// ---------------------------------------------------------------- trace (rt_main)

struct V3 { float x, y, z; };
TSI_DI V3 v3(float x, float y, float z) { return V3{x, y, z}; }
// (a sum of squares: +0 or more, +inf or NaN, where sqrt_nonneg equals sqrtf bit for bit)
TSI_DI float vlength(V3 a) { return sqrt_nonneg((a.x * a.x + a.y * a.y) + a.z * a.z); }
TSI_DI float wsign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

constexpr float kTrEps = 0.0001f;     // compute.wgsl:4140
constexpr float kDarken = 0.98f;      // compute.wgsl:3507

// compute.wgsl:4059-4070
TSI_DI V3 efficient_slabs(V3 p0, V3 p1, V3 o, V3 inv) {
    const float t0x = (p0.x - o.x) * inv.x, t0y = (p0.y - o.y) * inv.y, t0z = (p0.z - o.z) * inv.z;
    const float t1x = (p1.x - o.x) * inv.x, t1y = (p1.y - o.y) * inv.y, t1z = (p1.z - o.z) * inv.z;
    const float d1 = wmax(wmin(t0x, t1x), wmax(wmin(t0y, t1y), wmin(t0z, t1z)));
    const float d2 = wmin(wmax(t0x, t1x), wmin(wmax(t0y, t1y), wmax(t0z, t1z)));
    return v3((d1 <= d2 && d2 >= 0.0f) ? 1.0f : 0.0f, wmax(d1, 0.0f), wmax(d2, 0.0f));
}

TSI_DI float mylerp(float a, float b, float t) {   // compute.wgsl:4116
    const float tt = wmin(wmax(t, 0.0f), 1.0f);
    return a * (1.0f - tt) + b * tt;
}

// normalize (WGSL), as v / length(v)
TSI_DI V3 vnormalize(V3 v) {
    const float l = vlength(v);
    return v3(v.x / l, v.y / l, v.z / l);
}

// trace (compute.wgsl:4189-4403); the stack lives in LDS.  CIN: cinematic colouring (no octree
// darkening, movement (0.7, 1e-4, 1e-4) fading to base 0.1, 4346-4374) and normal_at_hit in nrm:
// finite_diff_normal's (1, 1, 1) for a tape scene, or the voxel's gradient with
// TSI_SHADE_GRADIENT (fast_normals = false, 4333-4342).
template <bool CIN>
TSI_DI void trace_ray(const KArgs &a, uint32_t *stk, V3 ro, V3 rd, float time, V3 &col, float &dist,
                      uint32_t &visits, uint32_t shading, const float4 *grad, V3 &nrm) {
    const bool normal_viz = (shading & TSI_SHADE_NORMALS) != 0;
    nrm = v3(0.0f, 0.0f, 0.0f);
    constexpr uint32_t FAIL = 0xFFFFFFFFu;
    uint32_t sp = 0;
    V3 c = v3(1.0f, 1.0f, 1.0f);
    V3 o2 = ro;
    bool intersected = false;
    const V3 inv = v3(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
    auto push = [&](uint32_t v) { if (sp != 16u) { stk[sp * 256u] = v; ++sp; } };
    auto peak = [&]() -> uint32_t { return sp == 0u ? FAIL : stk[(sp - 1u) * 256u]; };
    auto box = [&](uint32_t n, V3 &p0, V3 &p1) {
        const float4 lo = a.blo[n], hi = a.bhi[n];
        p0 = v3(lo.x, lo.y, lo.z);
        p1 = v3(hi.x, hi.y, hi.z);
    };
    push(0u);
    uint32_t region = 0u;
    V3 p0, p1;
    box(region, p0, p1);
    // q0/q1: the stored (unwidened) box of the node whose box p0/p1 was last set from: a child's
    // box is computed from it exactly as the set inversion computed the stored one
    // (lo + d * octant, d = (hi - lo) / 2), so the children's boxes are not loaded
    V3 q0 = p0, q1 = p1;
    uint32_t qnode = 0u;   // the node q0/q1 belong to
    // meta of the node entered by the last descent (its st was just read)
    uint32_t entered = FAIL, entered_meta = 0u;
    V3 hit = efficient_slabs(p0, p1, o2, inv);
    if (hit.x == 0.0f) { col = c; dist = -1.0f; return; }
    float sms = 0.0f;
    for (uint32_t iter = 0; iter < 5000u; ++iter) {
        if (intersected) break;
        if (sp == 0u) { col = c; dist = -1.0f; return; }
        --sp;
        region = stk[sp * 256u];
        if (hit.x == 0.0f) {
            const uint32_t top = peak();
            if (top != FAIL) {
                box(top, p0, p1);
                q0 = p0;
                q1 = p1;
                qnode = top;
                hit = efficient_slabs(p0, p1, o2, inv);
            }
            continue;
        }
        const float ex = p1.x - p0.x, ey = p1.y - p0.y, ez = p1.z - p0.z;
        const float rox = ((o2.x + (hit.y + kTrEps) * rd.x) - p0.x) * 2.0f / ex;
        const float roy = ((o2.y + (hit.y + kTrEps) * rd.y) - p0.y) * 2.0f / ey;
        const float roz = ((o2.z + (hit.y + kTrEps) * rd.z) - p0.z) * 2.0f / ez;
        const float rdx = rd.x * (2.0f / ex), rdy = rd.y * (2.0f / ey), rdz = rd.z * (2.0f / ez);
        int32_t mx = f2i_sat(floorf(rox)), my = f2i_sat(floorf(roy)), mz = f2i_sat(floorf(roz));
        const float ddx = 1.0f / fabsf(rdx), ddy = 1.0f / fabsf(rdy), ddz = 1.0f / fabsf(rdz);
        const float sgx = wsign(rdx), sgy = wsign(rdy), sgz = wsign(rdz);
        const int32_t stx = f2i_sat(sgx), sty = f2i_sat(sgy), stz = f2i_sat(sgz);
        float sdx = ((sgx * (float(mx) - rox)) + sgx * 0.5f + 0.5f) * ddx;
        float sdy = ((sgy * (float(my) - roy)) + sgy * 0.5f + 0.5f) * ddy;
        float sdz = ((sgz * (float(mz) - roz)) + sgz * 0.5f + 0.5f) * ddz;
        float mkx = 0.0f, mky = 0.0f, mkz = 0.0f;
        const V3 old = o2;
        const uint32_t rmeta = (region == entered ? entered_meta : meta_of(a.st[region])) & kChildMask;
        entered = FAIL;
        // (after a DDA that left the 2x2x2 grid the popped node keeps the previous box, as in the
        // reference; its children's boxes are then loaded)
        const bool qown = qnode == region;
        const float hx = (q1.x - q0.x) / 2.0f, hy = (q1.y - q0.y) / 2.0f, hz = (q1.z - q0.z) / 2.0f;
        for (int i = 0; i < 32; ++i) {
            if (!CIN && !normal_viz && wmax(p1.x - p0.x, wmax(p1.y - p0.y, p1.z - p0.z)) < 1.0f) {   // 4265
                c.x = c.x * kDarken; c.y = c.y * kDarken; c.z = c.z * kDarken;
            }
            const float t = ((vlength(v3(mkx * (sdx - ddx), mky * (sdy - ddy), mkz * (sdz - ddz))) + hit.y) + kTrEps) + sms;
            o2 = v3(old.x + t * rd.x, old.y + t * rd.y, old.z + t * rd.z);
            sms = 0.0f;
            if (t >= hit.z - kTrEps) {
                o2 = v3(old.x + hit.z * rd.x, old.y + hit.z * rd.y, old.z + hit.z * rd.z);
                const uint32_t top = peak();
                if (top != FAIL) {
                    box(top, p0, p1);
                    q0 = p0;
                    q1 = p1;
                    qnode = top;
                    hit = efficient_slabs(p0, p1, o2, inv);
                }
                break;
            }
            const int32_t ci = int32_t(uint32_t(mx) + 2u * uint32_t(my) + 4u * uint32_t(mz));
            if (ci < 0 || ci >= 8) break;
            if (rmeta != 0u) {
                const uint32_t child = (rmeta - 1u) * 8u + uint32_t(ci);
                const float4 cs = a.st[child];
                ++visits;
                if (cs.x <= 0.0f) {
                    // the child's box as eval_level stored it (tsi_eval.h), from its parent's
                    const float ox = float(uint32_t(ci) & 1u), oy = float((uint32_t(ci) >> 1) & 1u),
                                oz = float(uint32_t(ci) >> 2);
                    V3 lo = v3(q0.x + hx * ox, q0.y + hy * oy, q0.z + hz * oz);
                    V3 hi = v3(q0.x + hx * (1.0f + ox), q0.y + hy * (1.0f + oy), q0.z + hz * (1.0f + oz));
                    if (!qown) box(child, lo, hi);
                    const V3 tp0 = v3(lo.x - kTrEps, lo.y - kTrEps, lo.z - kTrEps);
                    const V3 tp1 = v3(hi.x + kTrEps, hi.y + kTrEps, hi.z + kTrEps);
                    const V3 th = efficient_slabs(tp0, tp1, o2, inv);
                    if (th.x == 1.0f) {
                        if ((meta_of(cs) & kChildMask) != 0u) {
                            push(region);
                            push(child);
                            p0 = tp0;
                            p1 = tp1;
                            q0 = lo;
                            q1 = hi;
                            qnode = child;
                            entered = child;
                            entered_meta = meta_of(cs);
                            hit = th;
                            break;
                        } else {
                            intersected = true;
                            if (normal_viz || CIN) {
                                V3 n = v3(1.0f, 1.0f, 1.0f);   // finite_diff_normal of a tape scene
                                if (shading & TSI_SHADE_GRADIENT) {
                                    const float4 gn = grad[child];
                                    n = v3(gn.x, gn.y, gn.z);
                                }
                                nrm = n;
                                if (normal_viz) {   // 4333-4377
                                    const V3 nv = vnormalize(v3(n.x * 1.0f, n.y * 1.0f, n.z * 1.0f));
                                    c = v3(nv.x * 0.5f + 0.5f, nv.y * 0.5f + 0.5f, nv.z * 0.5f + 0.5f);
                                    break;
                                }
                            }
                            const float tt = (time - cs.y) * 60.0f;
                            if (CIN) {
                                c = v3(c.x - mylerp(1.0f - 0.7f, 1.0f - 0.1f, tt), c.y - mylerp(1.0f - 0.0001f, 1.0f - 0.1f, tt),
                                       c.z - mylerp(1.0f - 0.0001f, 1.0f - 0.1f, tt));
                            } else {
                                c = v3(c.x - mylerp(0.0f, 0.0f, tt), c.y - mylerp(1.0f, 0.0f, tt), c.z - mylerp(1.0f, 0.0f, tt));
                            }
                            break;
                        }
                    }
                }
            }
            const bool bx = sdx <= wmin(sdy, sdz), by = sdy <= wmin(sdz, sdx), bz = sdz <= wmin(sdx, sdy);
            mkx = bx ? 1.0f : 0.0f;
            mky = by ? 1.0f : 0.0f;
            mkz = bz ? 1.0f : 0.0f;
            sdx = sdx + mkx * ddx;
            sdy = sdy + mky * ddy;
            sdz = sdz + mkz * ddz;
            mx = int32_t(uint32_t(mx) + uint32_t(bx ? stx : 0));
            my = int32_t(uint32_t(my) + uint32_t(by ? sty : 0));
            mz = int32_t(uint32_t(mz) + uint32_t(bz ? stz : 0));
        }
    }
    c = v3(wmax(wmin(c.x, 1.0f), 0.0f), wmax(wmin(c.y, 1.0f), 0.0f), wmax(wmin(c.z, 1.0f), 0.0f));
    float d = fabsf(vlength(v3(o2.x - ro.x, o2.y - ro.y, o2.z - ro.z)));
    if (!intersected) d = -1.0f;
    col = c;
    dist = d;
}

// One pixel of rt_main (compute.wgsl:4650-4688): ray_direction (4590-4594), view * vec4(d, 0)
// (4630), trace, get_color_single's normal_viz miss colour (4510-4514) and the 0.45 gamma.
TSI_DI void trace_pixel(const KArgs &a, const TraceArgs &t, uint32_t x, uint32_t y, uint32_t *stk, uint32_t &visits) {
    const float xyx = (float(x) - float(t.width) / 2.0f) + 0.5f;
    const float xyy = (float(y) - float(t.height) / 2.0f) + 0.5f;
    V3 d = v3(xyx, -xyy, -t.zdist);
    const float len = vlength(d);
    d = v3(d.x / len, d.y / len, d.z / len);
    const float *m = t.view;
    const V3 g = v3(((m[0] * d.x + m[4] * d.y) + m[8] * d.z) + m[12] * 0.0f,
                    ((m[1] * d.x + m[5] * d.y) + m[9] * d.z) + m[13] * 0.0f,
                    ((m[2] * d.x + m[6] * d.y) + m[10] * d.z) + m[14] * 0.0f);
    const V3 eye = v3(m[12], m[13], m[14]);
    V3 c, n;
    float dist;
    trace_ray<false>(a, stk, eye, g, t.time, c, dist, visits, t.shading, t.grad, n);
    const size_t pi = size_t(y) * t.width + x;
    float ca = 1.0f;
    if ((t.shading & TSI_SHADE_NORMALS) && !(dist >= 0.0f)) {
        c = v3(10.0f, 10.0f, 10.0f);
        ca = 0.0f;
    }
    const float r = (0.0f + c.x) / 1.0f, gg = (0.0f + c.y) / 1.0f, b = (0.0f + c.z) / 1.0f;
    t.rgba[pi] = make_float4(det_pow045(r), det_pow045(gg), det_pow045(b), (0.0f + ca) / 1.0f);
    t.depth[pi] = dist;
}

// ---------------------------------------------------------------- cinematic shading

TSI_DI float vdot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
TSI_DI V3 vcross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }

constexpr float kPiW = 3.14159274101257324f;   // compute.wgsl:2476 PI as f32

// frand0to1 (compute.wgsl:2466-2471): the LCG step seed * 747796405 + 1, then PCG's RXS-M-XS
// output permutation, mapped to [0, 1] by / 4294967295.0 (2^32 as f32); frand (2454-2459): [-1, 1].
TSI_DI float frand01(uint32_t &seed) {
    seed = seed * 747796405u + 1u;
    uint32_t w = ((seed >> ((seed >> 28u) + 4u)) ^ seed) * 277803737u;
    w = (w >> 22u) ^ w;
    return float(w) / 4294967295.0f;
}
TSI_DI float frand(uint32_t &seed) { return frand01(seed) * 2.0f - 1.0f; }

// getSampleBiased (4405-4434): a cosine-power lobe of exponent `power` around d
TSI_DI V3 sample_biased(V3 d, float power, uint32_t &seed) {
    const V3 dir = vnormalize(d);
    const V3 ot = fabsf(dir.x) > fabsf(dir.z) ? v3(-dir.y, dir.x, 0.0f) : v3(0.0f, -dir.z, dir.y);   // ortho
    const V3 o1 = vnormalize(ot);
    const V3 o2 = vnormalize(vcross(dir, o1));
    float rx = frand01(seed);
    float ry = frand01(seed);
    rx = rx * 2.0f * kPiW;
    ry = det_pow(ry, 1.0f / (power + 1.0f));
    const float om = sqrtf(1.0f - ry * ry);
    const float cx = det_cos(rx) * om, sy = det_sin(rx) * om;
    return v3((o1.x * cx + o2.x * sy) + dir.x * ry, (o1.y * cx + o2.y * sy) + dir.y * ry,
              (o1.z * cx + o2.z * sy) + dir.z * ry);
}

// fromSpherical (494-518)
TSI_DI V3 from_spherical(float px, float py) {
    return v3(det_cos(px) * det_sin(py), det_sin(px) * det_sin(py), det_cos(py));
}

// sunsky (520-640; Fragmentarium's Sunsky.frag) with turbidity = sunSize = 1 and the sky's sun
// intensity 1000 (423-427), RayleighPhase (429), totalMie (437), hgPhase (458), SunIntensity
// (478); pow(x, 2.0) is x * x.  The sun disk it computes is not in its result.
TSI_DI V3 sunsky(V3 view, float spx, float spy) {
    const float sadc = det_cos((1.0f * 3.1415927410125732f) / 180.0f);
    const V3 sd = vnormalize(from_spherical((spx - 0.0f) * 6.28000020980835f, (spy - 0.5f) * 3.140000104904175f));
    const V3 up = v3(0.0f, 0.0f, 1.0f);
    const float cvs = vdot(view, sd);
    const float csu = vdot(sd, up);
    const float cuv = vdot(up, view);
    if (sadc == 1.0f) return v3(1.0f, 0.0f, 0.0f);
    const float sunE = 1000.0f * wmax(0.0f, 1.0f - det_exp(-((1.6110731611233942f - det_acos(csu)) / 1.5f)));
    const V3 ray = v3(5.176821105123963e-6f, 1.2785348189936485e-5f, 2.8530756026157178e-5f);
    // totalMie(lambda = (680, 550, 450) nm, K = (0.686, 0.678, 0.666), T = 1) * 0.005
    const float mc = (0.20000000298023224f * 1.0f) * 9.99999983775159e-18f;
    const float mk = (0.4339999854564667f * mc) * 3.1415927410125732f;
    const float tp = 2.0f * 3.1415927410125732f;
    const float qx = tp / 6.800000278417428e-7f, qy = tp / 5.499999815583578e-7f, qz = tp / 4.4999998749517545e-7f;
    const V3 mie = v3(((mk * (qx * qx)) * 0.6859999895095825f) * 0.004999999888241291f,
                      ((mk * (qy * qy)) * 0.6779999732971191f) * 0.004999999888241291f,
                      ((mk * (qz * qz)) * 0.6660000085830688f) * 0.004999999888241291f);
    const float zen = wmax(0.0f, cuv);
    const float rol = 8400.0f / zen, mol = 1250.0f / zen;
    const V3 fex = v3(det_exp(-((ray.x * rol) + (mie.x * mol))), det_exp(-((ray.y * rol) + (mie.y * mol))),
                      det_exp(-((ray.z * rol) + (mie.z * mol))));
    const float rp = (3.0f / (16.0f * 3.1415927410125732f)) * (1.0f + cvs * cvs);
    const float g = 0.800000011920929f;
    const float hp = (1.0f / (4.0f * 3.1415927410125732f)) * ((1.0f - g * g) / det_pow((1.0f - (2.0f * g) * cvs) + g * g, 1.5f));
    const V3 lx = v3(ray.x * rp + mie.x * hp, ray.y * rp + mie.y * hp, ray.z * rp + mie.z * hp);
    const V3 tl = v3(ray.x + mie.x, ray.y + mie.y, ray.z + mie.z);
    const V3 se = v3(sunE * (lx.x / tl.x), sunE * (lx.y / tl.y), sunE * (lx.z / tl.z));
    const float f = wmin(wmax(det_pow(1.0f - vdot(up, sd), 5.0f), 0.0f), 1.0f);
    const V3 mx = v3(1.0f * (1.0f - f) + det_pow(se.x * fex.x, 0.5f) * f, 1.0f * (1.0f - f) + det_pow(se.y * fex.y, 0.5f) * f,
                     1.0f * (1.0f - f) + det_pow(se.z * fex.z, 0.5f) * f);
    return v3(((se.x * (1.0f - fex.x)) * mx.x) * 0.009999999776482582f, ((se.y * (1.0f - fex.y)) * mx.y) * 0.009999999776482582f,
              ((se.z * (1.0f - fex.z)) * mx.z) * 0.009999999776482582f);
}

// LinearToSRGB (4600-4607) per channel, after clamping to [0, 1]
TSI_DI float linear_to_srgb(float x) {
    const float r = wmin(wmax(x, 0.0f), 1.0f);
    const float a = det_pow(r, 0.416666666666666667f) * 1.055f - 0.055f;
    const float b = r * 12.92f;
    const float t = r < 0.0031308f ? 1.0f : 0.0f;
    return a * (1.0f - t) + b * t;
}

// One pixel of rt_main with cinematic = true (compute.wgsl:4651-4688, 4620-4648, 4440-4523):
// seed3 from the pixel index and the time, 8 samples jittered by frand() * 0.5, each
// get_color_single's path of two bounces, each bounce a sun shadow ray (traced only when its
// result is read: sunlight > 0, 4501), a miss after the first bounce sees the sky; the mean goes
// through LinearToSRGB and the 0.45 gamma, alpha is the mean alpha.  The lighting is the
// scene-independent one (sun at (-0.1, 0.7), intensity 3, specular 1, albedo 1, sun size 0.01).
// Every ray of the pixel goes through the one trace_ray call below: a lane's rays follow each
// other in a small state machine (sample, bounce, shadow), so a wavefront's lanes trace together
// wherever their paths are.  depth: the first sample's first ray.
TSI_DI void trace_pixel_cin(const KArgs &a, const TraceArgs &t, uint32_t x, uint32_t y, uint32_t *stk, uint32_t &visits) {
    constexpr float spx = -0.1f, spy = 0.7f, sun_i = 3.0f, specular_i = 1.0f, albedo = 1.0f, sunsize = 0.01f,
                    bgmul = 1.0f;
    const uint32_t pidx = y * t.width + x;
    uint32_t seed = pidx + uint32_t(12345.0f * (det_sin(t.time * 12345.0f) * 0.5f + 0.5f));
    const float *m = t.view;
    const V3 eye = v3(m[12], m[13], m[14]);
    const V3 sdn = vnormalize(from_spherical((spx - 0.0f) * 6.28f, (spy - 0.5f) * 3.14f));
    const V3 sundir = v3(sdn.x, sdn.z, sdn.y);
    float ax = 0.0f, ay = 0.0f, az = 0.0f, aw = 0.0f;
    float depth0 = -1.0f;
    V3 frm = eye, dir = eye, lum = eye, direct = eye, old_rd = eye, nh = eye, ssd = eye;
    float sunlight = 0.0f;
    uint32_t sample = 0, bounce = 0;
    bool shadow = false;
    auto begin_sample = [&]() {
        const float jx = frand(seed);
        const float jy = frand(seed);
        const float px = float(x) + (jx * 1.0f) * 0.5f, py = float(y) + (jy * 1.0f) * 0.5f;
        const float xyx = (px - float(t.width) / 2.0f) + 0.5f;
        const float xyy = (py - float(t.height) / 2.0f) + 0.5f;
        V3 d = v3(xyx, -xyy, -t.zdist);
        const float len = vlength(d);
        d = v3(d.x / len, d.y / len, d.z / len);
        dir = v3(((m[0] * d.x + m[4] * d.y) + m[8] * d.z) + m[12] * 0.0f,
                 ((m[1] * d.x + m[5] * d.y) + m[9] * d.z) + m[13] * 0.0f,
                 ((m[2] * d.x + m[6] * d.y) + m[10] * d.z) + m[14] * 0.0f);
        frm = eye;
        lum = v3(1.0f, 1.0f, 1.0f);
        direct = v3(0.0f, 0.0f, 0.0f);
        bounce = 0;
        shadow = false;
    };
    begin_sample();
    for (uint32_t it = 0; it < 32u; ++it) {   // 8 samples x 2 bounces x (ray, shadow ray)
        V3 c, n;
        float d;
        trace_ray<true>(a, stk, frm, shadow ? ssd : dir, t.time, c, d, visits, t.shading, t.grad, n);
        bool end_bounce = false, end_sample = false;
        float rx = 0.0f, ry = 0.0f, rz = 0.0f, rw = 0.0f;
        if (shadow) {
            shadow = false;
            if (d < 0.0f) {   // not in shadow
                direct = v3(direct.x + (lum.x * sunlight) * sun_i, direct.y + (lum.y * sunlight) * sun_i,
                            direct.z + (lum.z * sunlight) * sun_i);
                const V3 half = vnormalize(v3(-old_rd.x + sundir.x, -old_rd.y + sundir.y, -old_rd.z + sundir.z));
                const float spec = (100.0f * det_pow(wmax(vdot(nh, half), 0.0f), 30.0f)) * specular_i;
                direct = v3(direct.x + lum.x * spec, direct.y + lum.y * spec, direct.z + lum.z * spec);
            }
            end_bounce = true;
        } else {
            if (sample == 0u && bounce == 0u) depth0 = d;
            if (d >= 0.0f) {
                const V3 hp = v3(frm.x + dir.x * d, frm.y + dir.y * d, frm.z + dir.z * d);
                old_rd = dir;
                dir = sample_biased(n, 3.0f, seed);
                lum = v3((lum.x * c.x) * albedo, (lum.y * c.y) * albedo, (lum.z * c.z) * albedo);
                frm = v3(hp.x + n.x * 0.05f, hp.y + n.y * 0.05f, hp.z + n.z * 0.05f);
                const float j0 = frand(seed);
                const float j1 = frand(seed);
                const float j2 = frand(seed);
                ssd = vnormalize(v3(sundir.x + j0 * sunsize, sundir.y + j1 * sunsize, sundir.z + j2 * sunsize));
                sunlight = vdot(n, ssd);
                nh = n;
                if (sunlight > 0.0f) shadow = true;
                else end_bounce = true;
            } else if (bounce == 0u) {   // see_background = false (3506)
                rx = 10.0f; ry = 10.0f; rz = 10.0f; rw = 0.0f;
                end_sample = true;
            } else {
                const V3 bg = sunsky(v3(dir.x, dir.z, dir.y), spx, spy);
                rx = (direct.x + lum.x * bg.x) * bgmul;
                ry = (direct.y + lum.y * bg.y) * bgmul;
                rz = (direct.z + lum.z * bg.z) * bgmul;
                rw = 1.0f * bgmul;
                end_sample = true;
            }
        }
        if (end_bounce && ++bounce == 2u) {
            rx = direct.x; ry = direct.y; rz = direct.z; rw = 1.0f;
            end_sample = true;
        }
        if (end_sample) {
            ax = ax + rx; ay = ay + ry; az = az + rz; aw = aw + rw;
            if (++sample == 8u) break;
            begin_sample();
        }
    }
    ax = ax / 8.0f; ay = ay / 8.0f; az = az / 8.0f; aw = aw / 8.0f;
    const size_t pi = size_t(pidx);
    t.rgba[pi] = make_float4(det_pow045(linear_to_srgb(ax)), det_pow045(linear_to_srgb(ay)),
                             det_pow045(linear_to_srgb(az)), aw);
    t.depth[pi] = depth0;
}

template <bool CIN>
TSI_DI void shade_pixel(const KArgs &a, const TraceArgs &t, uint32_t x, uint32_t y, uint32_t *stk, uint32_t &visits) {
    if (CIN) trace_pixel_cin(a, t, x, y, stk, visits);
    else trace_pixel(a, t, x, y, stk, visits);
}

template <bool CIN>
__global__ void __launch_bounds__(256) k_trace(KArgs a, TraceArgs t) {
    __shared__ uint32_t stack_lds[16 * 256];
    // a 16x16 block of pixels; each wavefront takes one 8x8 quadrant (square tiles keep the
    // rays of a wavefront close together, so their traversals diverge less)
    const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63u;
    const uint32_t x = blockIdx.x * 16u + (w & 1u) * 8u + (l & 7u);
    const uint32_t y = blockIdx.y * 16u + (w >> 1) * 8u + (l >> 3);
    uint32_t visits = 0;
    if (x < t.width && y < t.height) shade_pixel<CIN>(a, t, x, y, stack_lds + threadIdx.x, visits);
    uint64_t v = wave_sum64(visits);
    if (lane_id() == 0 && v) atomicAdd((unsigned long long *)t.visits, (unsigned long long)v);
}

// The frames of a batch in one launch: block i is a 16x16-pixel block of the frame whose blocks
// [jobs[j].block0, jobs[j + 1].block0) hold it, row-major over ceil(width / 16) blocks per row,
// and its wavefronts take the 8x8 quadrants, as k_trace's do (a persistent grid taking tiles
// from a counter measured slower: a frame's slowest tiles then ran one after another in a few
// wavefronts, where the hardware dispatcher starts each block on the first free slot).
// The same pixels on a grid of a few blocks per CU: each wavefront claims 8x8-pixel quadrants
// (row-major over the frame) from a counter until none is left.  A full launch of 16x16 blocks
// keeps every CU's wave slots filled with trace wavefronts while any block is pending, so the
// next frame's level kernels, launched beside it, wait for slots; this leaves the CUs most of
// their slots.  Each pixel is shaded by the same code as in k_trace: the frame is the same.
template <bool CIN>
__global__ void __launch_bounds__(256) k_trace_quads(KArgs a, TraceArgs t, uint32_t *next_quad, uint32_t quads_x,
                                                     uint32_t nquads) {
    __shared__ uint32_t stack_lds[16 * 256];
    const uint32_t l = threadIdx.x & 63u;
    uint32_t visits = 0;
    for (;;) {
        uint32_t q = 0;
        if (l == 0) q = atomicAdd(next_quad, 1u);
        q = __builtin_amdgcn_readfirstlane(q);
        if (q >= nquads) break;
        const uint32_t x = (q % quads_x) * 8u + (l & 7u), y = (q / quads_x) * 8u + (l >> 3);
        if (x < t.width && y < t.height) shade_pixel<CIN>(a, t, x, y, stack_lds + threadIdx.x, visits);
    }
    uint64_t v = wave_sum64(visits);
    if (l == 0 && v) atomicAdd((unsigned long long *)t.visits, (unsigned long long)v);
}

template <bool CIN>
__global__ void __launch_bounds__(256) k_trace_batch(const TraceJob *jobs, uint32_t njobs) {
    __shared__ uint32_t stack_lds[16 * 256];
    const uint32_t blk = blockIdx.x;
    uint32_t lo = 0, hi = njobs;   // last job with block0 <= blk
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (as_const(jobs[mid]).block0 <= blk) lo = mid; else hi = mid;
    }
    // the job table is written before the launch and only read here: scalar loads (as_const)
    const auto &J = as_const(jobs[__builtin_amdgcn_readfirstlane(lo)]);
    KArgs a{};
    a.st = const_cast<float4 *>(J.st);
    a.blo = const_cast<float4 *>(J.blo);
    a.bhi = const_cast<float4 *>(J.bhi);
    TraceArgs t;
    for (int k = 0; k < 16; ++k) t.view[k] = J.t.view[k];
    t.width = J.t.width;
    t.height = J.t.height;
    t.time = J.t.time;
    t.zdist = J.t.zdist;
    t.rgba = J.t.rgba;
    t.depth = J.t.depth;
    t.visits = J.t.visits;
    t.shading = J.t.shading;
    t.grad = J.t.grad;
    const uint32_t local = blk - J.block0;
    const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63u;
    const uint32_t x = (local % J.blocks_x) * 16u + (w & 1u) * 8u + (l & 7u);
    const uint32_t y = (local / J.blocks_x) * 16u + (w >> 1) * 8u + (l >> 3);
    uint32_t visits = 0;
    if (x < t.width && y < t.height) shade_pixel<CIN>(a, t, x, y, stack_lds + threadIdx.x, visits);
    const uint64_t v = wave_sum64(visits);
    if (lane_id() == 0 && v) atomicAdd((unsigned long long *)t.visits, (unsigned long long)v);
}

// RGBA32F -> RGBA8 as the reference's frame dump converts it (src/main.rs:709-711, image 0.24
// into_rgba8): clamp to [0, 1], scale by 255, round half away from zero.
__global__ void __launch_bounds__(256) k_rgba8(const float4 *in, uchar4 *out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 v = in[i];
    auto q = [](float x) -> unsigned char {
        return (unsigned char)roundf(wmin(wmax(x, 0.0f), 1.0f) * 255.0f);
    };
    out[i] = make_uchar4(q(v.x), q(v.y), q(v.z), q(v.w));
}

// ---------------------------------------------------------------- launchers

int launch_rgba8(const float4 *in, uchar4 *out, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_rgba8, dim3((n + 255u) / 256u), dim3(256), 0, s, in, out, n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Blocks (4 wavefronts, 8 entries each) a list of level d can fill: it holds at most the 8^d nodes
// at depth d (and its dealloc list as many blocks), so the top levels launch a few blocks instead
// of the full grid (whose idle blocks would wait for dispatch behind a trace running beside).
uint32_t level_blocks(int d, uint32_t full) {
    if (d >= 7) return full;
    const uint64_t nodes = uint64_t(1) << (3 * d);
    const uint64_t groups = (nodes + 7u) / 8u;
    const uint64_t blocks = (groups + 3u) / 4u;
    return blocks < full ? uint32_t(blocks) : full;
}

int launch_frame(const KArgs &a, const FrameLaunch &fl, hipStream_t s, hipEvent_t e_eval0, hipEvent_t e_eval1,
                 const hipEvent_t *e_level) {
    const SpecLaunch *sp = fl.tape_eval ? fl.spec : nullptr;
    if (sp) {   // the speculative pass, beside the previous frame's last levels
        if (sp->wait) hipStreamWaitEvent(sp->stream, sp->wait, 0);
        hipLaunchKernelGGL(k_spec_gather, dim3(1), dim3(1024), 0, sp->stream, sp->args);
        hipEventRecord(sp->begin, sp->stream);   // times the tape kernel's launch only
        KArgs ka = sp->args;
        int level = -1;
        float t = fl.time;
        void *args[] = {&ka, &level, &t};
        if (hipModuleLaunchKernel((hipFunction_t)fl.tape_eval, sp->blocks, 1, 1, 256, 1, 1, 0, sp->stream, args,
                                  nullptr) != hipSuccess)
            return -1;
        hipEventRecord(sp->done, sp->stream);
    }
    hipLaunchKernelGGL(k_frame_begin, dim3(1), dim3(64), 0, s, a);
    if (sp) hipStreamWaitEvent(s, sp->done, 0);
    if (e_eval0) hipEventRecord(e_eval0, s);
    for (int d = 0; d < fl.levels; ++d) {
        const bool compiled = fl.tape_eval != nullptr;
        if (e_level) hipEventRecord(e_level[2 * d], s);
        if (a.prune && d > 0 && !compiled) {   // the interpreter's group programs
            hipLaunchKernelGGL(k_prune, dim3(fl.grid_misc), dim3(256), 0, s, a, d, 0);
            hipLaunchKernelGGL(k_prune, dim3(fl.grid_misc), dim3(256), 0, s, a, d, 1);
        }
        if (compiled) {
            KArgs ka = a;
            int level = d;
            float t = fl.time;
            void *args[] = {&ka, &level, &t};
            // both lists of the level in one launch (tsi_tape_eval_level splits the grid: each
            // list gets at least one block)
            const uint32_t nb = std::min(fl.grid_fwd + fl.grid_rev, 2u * level_blocks(d, fl.grid_rev));
            if (hipModuleLaunchKernel((hipFunction_t)fl.tape_eval, nb, 1, 1, 256, 1, 1, 0, s, args, nullptr) !=
                hipSuccess)
                return -1;
        } else {
            hipLaunchKernelGGL(k_eval<false>, dim3(level_blocks(d, fl.grid_fwd)), dim3(256), 0, s, a, d, fl.time);
            hipLaunchKernelGGL(k_eval<true>, dim3(level_blocks(d, fl.grid_rev)), dim3(256), 0, s, a, d, fl.time);
        }
        if (e_level) hipEventRecord(e_level[2 * d + 1], s);
        if (sp && d == (sp->ready_level < fl.levels - 1 ? sp->ready_level : fl.levels - 1))
            hipEventRecord(sp->ready, s);
    }
    if (e_eval1) hipEventRecord(e_eval1, s);
    // levels 1..L-1 ran the cascade's steps 1..L-1 (eval_level); blocks dropped by the last
    // level and the descendants of earlier drops need at most two more
    for (int c = fl.levels; c <= fl.levels + 1; ++c)
        hipLaunchKernelGGL(k_dealloc, dim3(fl.grid_misc), dim3(256), 0, s, a, c);
    hipLaunchKernelGGL(k_frame_end, dim3(fl.grid_misc), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_frame_commit, dim3(1), dim3(1), 0, s, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_reset(const KArgs &a, const float box[6], uint64_t nodes, hipStream_t s) {
    hipLaunchKernelGGL(k_clear_nodes, dim3(1024), dim3(256), 0, s, a, nodes);
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(1), 0, s, a, box[0], box[1], box[2], box[3], box[4], box[5]);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_enqueue_root(const KArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_enqueue_root, dim3(1), dim3(1), 0, s, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Copy of the nodes in use for an asynchronous trace; the count is read on the device (the
// host does not wait for the set inversion that produced it).
// (s_grad: the gradients too, for gradient-normal shading; may be null)
__global__ void __launch_bounds__(256) k_snapshot(const Counters *cnt, const float4 *st, const float4 *blo,
                                                  const float4 *bhi, const float4 *grad, float4 *s_st, float4 *s_blo,
                                                  float4 *s_bhi, float4 *s_grad, uint64_t nodes) {
    uint64_t used = (uint64_t(cnt->bump) + 1u) * 8u;
    if (used > nodes) used = nodes;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < used; i += uint64_t(gridDim.x) * blockDim.x) {
        s_st[i] = st[i];
        s_blo[i] = blo[i];
        s_bhi[i] = bhi[i];
        if (s_grad) s_grad[i] = grad[i];
    }
}

int launch_snapshot(const KArgs &a, float4 *s_st, float4 *s_blo, float4 *s_bhi, float4 *s_grad, uint64_t nodes,
                    uint32_t grid, hipStream_t s) {
    hipLaunchKernelGGL(k_snapshot, dim3(grid), dim3(256), 0, s, a.cnt, a.st, a.blo, a.bhi, a.grad, s_st, s_blo, s_bhi,
                       s_grad, nodes);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_trace_batch(const TraceJob *jobs, uint32_t njobs, uint32_t total_blocks, bool cinematic, hipStream_t s) {
    if (cinematic) hipLaunchKernelGGL(k_trace_batch<true>, dim3(total_blocks), dim3(256), 0, s, jobs, njobs);
    else hipLaunchKernelGGL(k_trace_batch<false>, dim3(total_blocks), dim3(256), 0, s, jobs, njobs);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

int launch_fill_u32(uint32_t *p, uint32_t v, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u32, dim3((n + 255u) / 256u < 1024u ? (n + 255u) / 256u : 1024u), dim3(256), 0, s, p,
                       v, n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_trace(const KArgs &a, const TraceArgs &t, hipStream_t s, uint32_t blocks, uint32_t *next_quad) {
    if (blocks && next_quad) {   // persistent: `blocks` blocks, *next_quad zeroed on the stream before
        const uint32_t qx = (t.width + 7u) / 8u, nq = qx * ((t.height + 7u) / 8u);
        const uint32_t nb = std::min(blocks, (nq + 3u) / 4u);
        if (t.shading & TSI_SHADE_CINEMATIC)
            hipLaunchKernelGGL(k_trace_quads<true>, dim3(nb), dim3(256), 0, s, a, t, next_quad, qx, nq);
        else
            hipLaunchKernelGGL(k_trace_quads<false>, dim3(nb), dim3(256), 0, s, a, t, next_quad, qx, nq);
        return hipGetLastError() == hipSuccess ? 0 : -1;
    }
    dim3 grid((t.width + 15u) / 16u, (t.height + 15u) / 16u);
    if (t.shading & TSI_SHADE_CINEMATIC) hipLaunchKernelGGL(k_trace<true>, grid, dim3(256), 0, s, a, t);
    else hipLaunchKernelGGL(k_trace<false>, grid, dim3(256), 0, s, a, t);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace tsi

// ---- check of the interval square root against the compiler's correctly rounded sqrtf ----

namespace tsi {

__global__ void __launch_bounds__(256) k_check_sqrt(uint32_t first, uint32_t count, unsigned long long *bad,
                                                    uint32_t *first_bad, uint32_t *done) {
    const uint32_t stride = gridDim.x * blockDim.x;   // count <= 2^31: i + stride never wraps
    uint32_t n = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride, ++n) {
        const uint32_t u = first + i;
        const float x = __uint_as_float(u);
        if (__float_as_uint(sqrt_nonneg(x)) != __float_as_uint(sqrtf(x))) {
            atomicAdd(bad, 1ull);
            atomicMin(first_bad, u);
        }
    }
    if (n) atomicAdd(done, n);
}

}  // namespace tsi

// Diagnostic export (no counterpart in the reference; not part of the tsi.h interface): compares
// sqrt_nonneg, the square root of SQRT_I, with sqrtf bit for bit for the binary32 values with bit
// patterns first .. first + count - 1 on the current device.  *mismatches receives the number of
// differences, *first_bad the smallest differing pattern (0xFFFFFFFF: none), *checked the number
// of values compared (count, when every thread ran).
extern "C" int tsi_check_sqrt(uint32_t first, uint32_t count, uint64_t *mismatches, uint32_t *first_bad,
                              uint32_t *checked) {
    if (!mismatches || !first_bad || !checked || count == 0 || count > 0x80000000u ||
        uint64_t(first) + count > (1ull << 32))
        return TSI_E_ARG;
    void *buf = nullptr;
    if (hipMalloc(&buf, 16) != hipSuccess) return TSI_E_OOM;
    unsigned long long *bad = static_cast<unsigned long long *>(buf);
    uint32_t *fb = reinterpret_cast<uint32_t *>(static_cast<char *>(buf) + 8);
    const uint32_t init[4] = {0u, 0u, 0xFFFFFFFFu, 0u};
    int rc = TSI_OK;
    if (hipMemcpy(buf, init, sizeof init, hipMemcpyHostToDevice) != hipSuccess) rc = TSI_E_HIP;
    if (rc == TSI_OK) {
        hipLaunchKernelGGL(tsi::k_check_sqrt, dim3(4096), dim3(256), 0, 0, first, count, bad, fb, fb + 1);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = TSI_E_HIP;
    }
    uint32_t out[4] = {0u, 0u, 0u, 0u};
    if (rc == TSI_OK && hipMemcpy(out, buf, sizeof out, hipMemcpyDeviceToHost) != hipSuccess) rc = TSI_E_HIP;
    hipFree(buf);
    if (rc != TSI_OK) return rc;
    *mismatches = uint64_t(out[0]) | (uint64_t(out[1]) << 32);
    *first_bad = out[2];
    *checked = out[3];
    return TSI_OK;
}
