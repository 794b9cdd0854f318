// tsi_wave.h -- wavefront (64-lane) helpers and the node-state bit layout shared by the
// set-inversion kernels and the per-tape evaluation kernels.
#pragma once
#include "tsi_interval.h"
#include "tsi_kernels.h"

namespace tsi {

TSI_DI uint32_t lane_id() { return threadIdx.x & 63u; }
TSI_DI uint64_t ballot(bool p) { return __ballot(p); }
TSI_DI uint32_t popc64(uint64_t m) { return uint32_t(__popcll(m)); }
TSI_DI uint32_t rank_in(uint64_t m) {
    // number of set bits of m below this lane
    return uint32_t(__popcll(m & ((1ull << lane_id()) - 1ull)));
}
TSI_DI uint32_t bcast(uint32_t v, int src) { return uint32_t(__shfl(int(v), src, 64)); }

// Appends one element per lane where p holds to list[0 .. cap), count at *cnt.  One atomic per
// wavefront; returns the lane's position (0xFFFFFFFF if none, or on overflow, which is flagged).
TSI_DI uint32_t wave_append(bool p, uint32_t value, uint32_t *list, uint32_t *cnt, uint32_t cap, uint32_t *err,
                            uint32_t errbit) {
    uint64_t m = ballot(p);
    if (m == 0) return 0xFFFFFFFFu;
    uint32_t n = popc64(m), base = 0;
    if (lane_id() == 0) base = atomicAdd(cnt, n);
    base = bcast(base, 0);
    uint32_t at = 0xFFFFFFFFu;
    if (p) {
        at = base + rank_in(m);
        if (at < cap) list[at] = value;
        else { atomicOr(err, errbit); at = 0xFFFFFFFFu; }
    }
    return at;
}

TSI_DI uint64_t wave_sum64(uint64_t v) {
    for (int off = 32; off > 0; off >>= 1) {
        uint32_t lo = uint32_t(v), hi = uint32_t(v >> 32);
        lo = uint32_t(__shfl_xor(int(lo), off, 64));
        hi = uint32_t(__shfl_xor(int(hi), off, 64));
        v += (uint64_t(hi) << 32) | lo;
    }
    return v;
}

// A struct the kernel only reads (a table written before the launch), seen through the constant
// address space: reading a field is a scalar load, as for a launch's own arguments.
template <class T>
TSI_DI const __attribute__((address_space(4))) T &as_const(const T &x) {
    return *(const __attribute__((address_space(4))) T *)&x;
}

// node state (st[].w)
TSI_DI uint32_t meta_of(float4 s) { return __float_as_uint(s.w); }
constexpr uint32_t kFastBit = 0x80000000u;
constexpr uint32_t kChildMask = 0x7FFFFFFFu;   // child block + 1, 0 = no children

}  // namespace tsi
