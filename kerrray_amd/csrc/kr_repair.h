// Verify/repair of a planned trace's mask, run to a fixed point.
//
// After the trace, a pixel must be re-traced in FP64 when it was not traced
// in FP64 and its photon ended neither captured nor escaped, or some pixel
// within band_px (Chebyshev distance) differs from it in captured-ness: its
// class could be one that FP32 round-off moved across the boundary.  That is
// the verify pass.  A re-traced pixel whose class changes moves the boundary,
// so the FP32 pixels within band_px of it are re-traced in the next round,
// and so on until a round changes no class.  Only pixels within band_px of a
// changed pixel can have gained a differing neighbour, so the rounds after
// the first look at those alone.  A pixel is listed at most once over all
// rounds (its mark byte records FP64 and "queued"), which bounds every list
// by the pixel count.
//
// The same functions build the GPU kernels (kr_mixed.hip) and, with
// KR_HOST_BUILD, a host simulation of the rounds used by the test suite.
#pragma once

#include "kr_device.h"

namespace kr {

// Per-pixel mark byte of a planned trace.
enum : uint8_t { MARK_FP64 = 1, MARK_QUEUED = 2 };

// Repair rounds after the verify pass (KR_REPAIR_ROUNDS of the C ABI).
constexpr int REPAIR_ROUNDS = 4;

// Verify one pixel of an n x n image (row-major states and marks).
KR_HD bool verify_pixel(const int64_t* state, const uint8_t* marks, int64_t n, int band_px, int64_t idx) {
    if (marks[idx] & MARK_FP64) return false;
    const int64_t s = state[idx];
    if (s != CAPTURED && s != ESCAPED) return true;
    const bool cap = s == CAPTURED;
    const int64_t i = idx / n, j = idx - (idx / n) * n;
    for (int di = -band_px; di <= band_px; ++di) {
        const int64_t ii = i + di;
        if (ii < 0 || ii >= n) continue;
        for (int dj = -band_px; dj <= band_px; ++dj) {
            const int64_t jj = j + dj;
            if (jj < 0 || jj >= n) continue;
            if ((state[ii * n + jj] == CAPTURED) != cap) return true;
        }
    }
    return false;
}

// Calls f(pos) for every pixel within band_px of idx, idx excluded, inside
// the image.
template <class F>
KR_HD void for_each_neighbour(int64_t n, int band_px, int64_t idx, F&& f) {
    const int64_t i = idx / n, j = idx - (idx / n) * n;
    for (int di = -band_px; di <= band_px; ++di) {
        const int64_t ii = i + di;
        if (ii < 0 || ii >= n) continue;
        for (int dj = -band_px; dj <= band_px; ++dj) {
            const int64_t jj = j + dj;
            if (jj < 0 || jj >= n || (di == 0 && dj == 0)) continue;
            f(ii * n + jj);
        }
    }
}

// Sets `bits` in the mark byte of pixel pos and returns the byte as it was.
// On the device the marks are updated by 32-bit atomics on the aligned word
// holding the byte (marks must be 4-byte aligned), since lanes of one round
// mark neighbouring pixels concurrently.
KR_HD uint8_t mark_or(uint8_t* marks, int64_t pos, uint8_t bits) {
#if defined(KR_HOST_BUILD) || !defined(__HIP_DEVICE_COMPILE__)
    const uint8_t old = marks[pos];
    marks[pos] = static_cast<uint8_t>(old | bits);
    return old;
#else
    unsigned int* word = reinterpret_cast<unsigned int*>(marks + (pos & ~int64_t(3)));
    const int shift = static_cast<int>(pos & 3) * 8;
    const unsigned int old = atomicOr(word, static_cast<unsigned int>(bits) << shift);
    return static_cast<uint8_t>((old >> shift) & 0xffu);
#endif
}

}  // namespace kr
