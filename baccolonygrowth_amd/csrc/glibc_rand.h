// glibc_rand.h - the glibc rand() stream (random_r TYPE_3), with jump-ahead.
//
// The reference draws every random number from the process-wide rand()
// (ABM/CellMD3D.cpp:23; ABM/Forces.cpp:102; ABM/grow.cpp:109-152;
// ABM/Simulation.cpp:166). This is the published glibc algorithm: seed
// r0 = s (0 -> 1), r_i = 16807 r_{i-1} mod (2^31 - 1) for i = 1..30,
// r_i = r_{i-31} for i = 31..33, then r_k = r_{k-3} + r_{k-31} (mod 2^32)
// with the first 310 values discarded; each draw is r_k >> 1.
//
// The recurrence is linear over Z/2^32 with characteristic polynomial
// x^31 - x^28 - 1, so skipping n draws is x^n mod that polynomial
// (square-and-multiply) applied to the last 31 values: O(31^2 log n).
#pragma once

#include <stdint.h>
#include <string.h>

namespace bcg {

class GlibcRand {
   public:
    void seed(uint32_t s) {
        if (s == 0) s = 1;
        r_[0] = s;
        for (int i = 1; i < 31; i++) {
            // 16807 * r mod (2^31 - 1) on the signed 31-bit value
            long long w = (16807LL * (int32_t)r_[i - 1]) % 2147483647LL;
            if (w < 0) w += 2147483647LL;
            r_[i] = (uint32_t)w;
        }
        for (int i = 31; i < 34; i++) r_[i] = r_[i - 31];
        k_ = 0;
        for (int i = 34; i < 344; i++) (void)next_raw();
    }

    // One draw, as rand().
    int32_t next() { return (int32_t)(next_raw() >> 1); }

    // n draws as rand() in a row, into out: the recurrence over a straight
    // buffer (no ring indexing), then the ring rewound to its last 34 values.
    void fill(int32_t* out, size_t n) {
        constexpr size_t B = 4096;
        uint32_t W[34 + B];
        while (n > 0) {
            const size_t m = n < B ? n : B;
            for (int j = 0; j < 34; j++) W[j] = r_[(k_ + j) % 34];  // oldest first
            for (size_t j = 34; j < 34 + m; j++) {
                W[j] = W[j - 31] + W[j - 3];
                out[j - 34] = (int32_t)(W[j] >> 1);
            }
            for (int j = 0; j < 34; j++) r_[j] = W[m + j];
            k_ = 0;
            out += m;
            n -= m;
        }
    }

    // ((float)rand()/RAND_MAX - 0.5) exactly as the reference evaluates it.
    double centered() {
        float f = (float)next() / (float)2147483647;
        return (double)f - 0.5;
    }

    void discard(uint64_t n) {
        if (n < 1024) {
            while (n--) (void)next_raw();
            return;
        }
        uint32_t c[31];
        xpow(n, c);
        uint32_t W[61];
        for (int j = 0; j < 31; j++) W[j] = r_[(k_ + 34 - 31 + j) % 34];
        for (int j = 31; j < 61; j++) W[j] = W[j - 3] + W[j - 31];
        uint32_t nw[31];
        for (int j = 0; j < 31; j++) {
            uint32_t s = 0;
            for (int i = 0; i < 31; i++) s += c[i] * W[i + j];
            nw[j] = s;
        }
        k_ = (int)((k_ + n) % 34);
        for (int j = 0; j < 31; j++) r_[(k_ + 34 - 31 + j) % 34] = nw[j];
    }

   private:
    uint32_t r_[34];
    int k_;  // ring slot of the next value

    uint32_t next_raw() {
        uint32_t v = r_[(k_ + 3) % 34] + r_[(k_ + 31) % 34];
        r_[k_] = v;
        k_ = (k_ + 1) % 34;
        return v;
    }

    // a <- a*b mod (x^31 - x^28 - 1)
    static void mulmod(uint32_t* a, const uint32_t* b) {
        uint32_t t[61];
        memset(t, 0, sizeof t);
        for (int i = 0; i < 31; i++)
            for (int j = 0; j < 31; j++) t[i + j] += a[i] * b[j];
        for (int d = 60; d >= 31; d--) {
            t[d - 3] += t[d];
            t[d - 31] += t[d];
        }
        memcpy(a, t, 31 * sizeof(uint32_t));
    }

    // c <- x^n mod (x^31 - x^28 - 1)
    static void xpow(uint64_t n, uint32_t* c) {
        uint32_t base[31], res[31];
        memset(base, 0, sizeof base);
        memset(res, 0, sizeof res);
        base[1] = 1;
        res[0] = 1;
        while (n) {
            if (n & 1) mulmod(res, base);
            n >>= 1;
            if (n) mulmod(base, base);
        }
        memcpy(c, res, sizeof res);
    }
};

}  // namespace bcg
