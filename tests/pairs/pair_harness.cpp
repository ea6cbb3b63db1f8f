// Host harness of the list build's FP32 pair classification (test
// infrastructure): the classifier the kernels use, compiled for the host, over
// a batch of pairs. Reciprocal mode 0 is exact; 1 and 2 round it one ulp up or
// down, as an approximate reciprocal may; 3 is 30 % low, which moves the
// closest-point parameters far off (the bounds must hold for any parameters).
#include "../../baccolonygrowth_amd/csrc/pair_class.h"

#include <math.h>

namespace {
struct HostRcp {
    int mode;
    float operator()(float x) const {
        const float r = 1.0f / x;
        if (mode == 1) return nextafterf(r, INFINITY);
        if (mode == 2) return nextafterf(r, -INFINITY);
        if (mode == 3) return 0.7f * r;
        return r;
    }
};
}  // namespace

extern "C" {
// rec: n x 16 floats (a_p, a_q, b_p, b_q as float4); thr: n x 2 (re, rd);
// shift: n x 2 (sx, sy); out: n verdicts.
int pc_classify_batch(int n, const float* rec, const float* thr, const float* shift, int mode, int* out) {
    const HostRcp rcp{mode};
    for (int i = 0; i < n; i++) {
        const float* r = rec + 16 * (long)i;
        const float4 ap = make_float4(r[0], r[1], r[2], r[3]), aq = make_float4(r[4], r[5], r[6], r[7]);
        const float4 bp = make_float4(r[8], r[9], r[10], r[11]), bq = make_float4(r[12], r[13], r[14], r[15]);
        out[i] = bcg::classify_pair(ap, aq, thr[2 * i], thr[2 * i + 1], bp, bq, shift[2 * i], shift[2 * i + 1], rcp);
    }
    return 0;
}

// The list build's end-point prefilter over a batch: ends n x 6 (the
// candidate's end points p, q with x, y relative to its column origin, FP64),
// centre n x 3 (A's centre, x and y relative to the same origin, FP64), reach n
// (A's L/2 + Rc R, FP64); out: 1 when the key's sign bit is set (kept).
int pc_prefilter_batch(int n, const double* ends, const double* centre, const double* reach, int* out) {
    for (int i = 0; i < n; i++) {
        const double* e = ends + 6 * (long)i;
        const double* c = centre + 3 * (long)i;
        const float4 p = bcg::prefilter_point(e[0], e[1], e[2]), q = bcg::prefilter_point(e[3], e[4], e[5]);
        const bcg::PreCol col =
            bcg::prefilter_column((float)c[0], (float)c[1], (float)c[2], bcg::prefilter_reach(reach[i]));
        out[i] = signbit(col.key(p, q)) ? 1 : 0;
    }
    return 0;
}

// surely_apart over a batch: rec as above (end points relative to a common
// origin), sigma: n values.
int pc_apart_batch(int n, const float* rec, const float* sigma, int mode, int* out) {
    const HostRcp rcp{mode};
    for (int i = 0; i < n; i++) {
        const float* r = rec + 16 * (long)i;
        const float4 ap = make_float4(r[0], r[1], r[2], r[3]), aq = make_float4(r[4], r[5], r[6], r[7]);
        const float4 bp = make_float4(r[8], r[9], r[10], r[11]), bq = make_float4(r[12], r[13], r[14], r[15]);
        out[i] = bcg::surely_apart(ap, aq, bp, bq, sigma[i], rcp) ? 1 : 0;
    }
    return 0;
}
}
