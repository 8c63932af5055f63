// Register sequences of LFSRCirculant (LFSRCirculant.pyx:159-237, 316-407).
//
// The circulant matrix of a +-1 shift-register sequence is applied as
// scatter -> Walsh-Hadamard transform -> gather (the reference's _core,
// LFSRCirculant.pyx:316-396): input row t goes to the register state reached
// after t generator steps, the output reads the states reached by the
// tap-side recurrence from 1.  This file produces those two state sequences
// on the host (the recurrence is sequential); the apply itself runs on the
// library's scatter, Hadamard and gather kernels.
#include "fm_common.hpp"

namespace {

// one step of the generating register: the new top bit is the parity of the
// tapped bits (LFSRCirculant.pyx:31-40)
inline uint32_t gen_step(uint32_t state, uint32_t poly, uint32_t mask) {
    if (__builtin_parity(state & poly)) state |= mask;
    return state >> 1;
}

// one step of the tap-side register (LFSRCirculant.pyx:42-48)
inline uint32_t tap_step(uint32_t state, uint32_t poly, uint32_t mask) {
    state <<= 1;
    if (state & mask) state ^= (poly | mask);
    return state;
}

}  // namespace

extern "C" fm_status fm_lfsr_sequence(uint32_t polynomial, uint32_t start, int order,
                                      int64_t *period, int64_t *states, int64_t *taps,
                                      int64_t capacity) {
    FM_REQUIRE(order >= 1 && order <= 31, FM_ERR_VALUE,
               "Only polynomials of order 1 to 31 are supported.");
    FM_REQUIRE(period != nullptr, FM_ERR_VALUE, "NULL period");
    const uint32_t mask = uint32_t(1) << order;
    const uint32_t st = start & (mask - 1);
    FM_REQUIRE(st != 0, FM_ERR_VALUE, "Initial state must be non-zero.");
    uint32_t s = gen_step(st, polynomial, mask);
    int64_t p = 1;
    while (s != st) {
        s = gen_step(s, polynomial, mask);
        ++p;
        FM_REQUIRE(p < int64_t(mask) && s != 0, FM_ERR_VALUE,
                   "Register configuration produces invalid sequence.");
    }
    *period = p;
    if (!states && !taps) return FM_SUCCESS;
    FM_REQUIRE(capacity >= p, FM_ERR_VALUE, "sequence buffers hold %lld < %lld entries",
               (long long)capacity, (long long)p);
    if (states) {
        uint32_t a = st;
        for (int64_t t = 0; t < p; ++t) {
            states[t] = int64_t(a & (mask - 1));
            a = gen_step(a, polynomial, mask);
        }
    }
    if (taps) {
        uint32_t b = 1;
        for (int64_t t = 0; t < p; ++t) {
            taps[t] = int64_t(b);
            b = tap_step(b, polynomial, mask);
        }
    }
    return FM_SUCCESS;
}
