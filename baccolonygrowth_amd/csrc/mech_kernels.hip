// mech_kernels.hip - gfx950 kernels of the mechanics step.
//
//   bins      counting sort of cells by UniformGrid box (atomic counts,
//             wave-level scan, scatter, per-bin sort to ascending ID)
//   lists     27-box gather + reduceNeighbours filter, one thread per cell,
//             ELL layout [k][slot] so a wave's list reads coalesce
//   move_grow both sum_forces passes, the two half kicks and the drift,
//             then grow and the division flag; one thread per cell, no
//             force atomics, neighbour records read as 128-byte lines
//   compact   ascending list of dividing cells from the division bitmap
//   divide    one thread per division
//   patch     the reference's post-division list patch, replayed in
//             division order by one workgroup
#include "mech_kernels.h"
#include "pair_class.h"

#include <climits>

namespace bcg {

// ------------------------------------------------------------- scans ---
// Inclusive wave scan (64 lanes) of one int per lane.
__device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}

// Exclusive block scan of one int per thread (blockDim multiple of 64,
// at most 1024). Returns the exclusive prefix; *total gets the block sum.
__device__ __forceinline__ int block_excl_scan(int v, int* lds, int* total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    int inc = wave_incl_scan(v);
    if (lane == 63) lds[wid] = inc;
    __syncthreads();
    if (wid == 0) {
        int w = lane < nw ? lds[lane] : 0;
        int wi = wave_incl_scan(w);
        if (lane < nw) lds[32 + lane] = wi - w;
        if (lane == nw - 1) lds[31] = wi;
    }
    __syncthreads();
    int r = inc - v + lds[32 + wid];
    if (total) *total = lds[31];
    __syncthreads();
    return r;
}

constexpr int SCAN_T = 256, SCAN_I = 8, SCAN_B = SCAN_T * SCAN_I;

__global__ void __launch_bounds__(SCAN_T) k_scan_reduce(const int* __restrict__ in, int n, int* __restrict__ tot) {
    __shared__ int lds[96];
    const int base = blockIdx.x * SCAN_B + threadIdx.x * SCAN_I;
    int s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_I; i++) s += (base + i < n) ? in[base + i] : 0;
    int t;
    block_excl_scan(s, lds, &t);
    if (threadIdx.x == 0) tot[blockIdx.x] = t;
}

__global__ void __launch_bounds__(1024) k_scan_totals(int* __restrict__ tot, int nb, int* __restrict__ grand) {
    __shared__ int lds[96];
    const int per = (nb + 1023) / 1024;
    const int base = threadIdx.x * per;
    int s = 0;
    for (int i = 0; i < per; i++) s += (base + i < nb) ? tot[base + i] : 0;
    int t;
    int ex = block_excl_scan(s, lds, &t);
    for (int i = 0; i < per; i++)
        if (base + i < nb) {
            int v = tot[base + i];
            tot[base + i] = ex;
            ex += v;
        }
    if (threadIdx.x == 0 && grand) *grand = t;
}

__global__ void __launch_bounds__(SCAN_T) k_scan_apply(const int* __restrict__ in, int n, const int* __restrict__ tot,
                                                       int* __restrict__ out) {
    __shared__ int lds[96];
    const int base = blockIdx.x * SCAN_B + threadIdx.x * SCAN_I;
    int v[SCAN_I];
    int s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_I; i++) {
        v[i] = (base + i < n) ? in[base + i] : 0;
        s += v[i];
    }
    int ex = block_excl_scan(s, lds, nullptr) + tot[blockIdx.x];
#pragma unroll
    for (int i = 0; i < SCAN_I; i++)
        if (base + i < n) {
            out[base + i] = ex;
            ex += v[i];
        }
}

// Two-launch scan: tile sums, then every tile adds the sums of the tiles
// before it (at most a few thousand, read from L2) and scans its own items.
// A thread owns 16 consecutive items, read and written as four 16-byte
// vectors when the arrays are 16-byte aligned. In place (in == out) is
// allowed: a tile reads all its items before it writes any.
constexpr int TSCAN_I = TSCAN_ITEMS / TSCAN_T;
static_assert(TSCAN_I == 16, "four 16-byte vectors a thread");

__device__ __forceinline__ void tscan_load(const int* __restrict__ in, int n, int base, bool vec, int (&v)[TSCAN_I]) {
    if (vec && base + TSCAN_I <= n) {
        const int4* p = reinterpret_cast<const int4*>(in + base);
#pragma unroll
        for (int q = 0; q < TSCAN_I / 4; q++) {
            const int4 x = p[q];
            v[4 * q] = x.x;
            v[4 * q + 1] = x.y;
            v[4 * q + 2] = x.z;
            v[4 * q + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < TSCAN_I; i++) v[i] = base + i < n ? in[base + i] : 0;
    }
}

// Sum of one int per thread over the block (blockDim = TSCAN_T).
__device__ __forceinline__ int block_sum(int s, int* lds) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
    __syncthreads();
    int t = 0;
#pragma unroll
    for (int w = 0; w < TSCAN_T / 64; w++) t += lds[w];
    __syncthreads();
    return t;
}

__global__ void __launch_bounds__(TSCAN_T) k_tscan_sums(const int* __restrict__ in, int n, int* __restrict__ tot) {
    __shared__ int lds[TSCAN_T / 64];
    const bool vec = (reinterpret_cast<uintptr_t>(in) & 15) == 0;
    int v[TSCAN_I];
    tscan_load(in, n, blockIdx.x * TSCAN_ITEMS + threadIdx.x * TSCAN_I, vec, v);
    int s = 0;
#pragma unroll
    for (int i = 0; i < TSCAN_I; i++) s += v[i];
    s = block_sum(s, lds);
    if (threadIdx.x == 0) tot[blockIdx.x] = s;
}

__global__ void __launch_bounds__(TSCAN_T) k_tscan_apply(const int* in, int n, const int* __restrict__ tot, int* out,
                                                         int* __restrict__ grand) {
    __shared__ int lds[96];
    const bool vec = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
    const int b = blockIdx.x;
    int p = 0;
    for (int i = threadIdx.x; i < b; i += TSCAN_T) p += tot[i];
    const int prefix = block_sum(p, lds);
    const int base = b * TSCAN_ITEMS + threadIdx.x * TSCAN_I;
    int v[TSCAN_I];
    tscan_load(in, n, base, vec, v);
    int s = 0;
#pragma unroll
    for (int i = 0; i < TSCAN_I; i++) s += v[i];
    int total;
    int ex = prefix + block_excl_scan(s, lds, &total);
#pragma unroll
    for (int i = 0; i < TSCAN_I; i++) {
        const int x = v[i];
        v[i] = ex;
        ex += x;
    }
    if (vec && base + TSCAN_I <= n) {
        int4* q = reinterpret_cast<int4*>(out + base);
#pragma unroll
        for (int j = 0; j < TSCAN_I / 4; j++) q[j] = make_int4(v[4 * j], v[4 * j + 1], v[4 * j + 2], v[4 * j + 3]);
    } else {
#pragma unroll
        for (int i = 0; i < TSCAN_I; i++)
            if (base + i < n) out[base + i] = v[i];
    }
    if (grand && b == (int)gridDim.x - 1 && threadIdx.x == 0) *grand = prefix + total;
}

// ------------------------------------------------------------- bins ---
// Origin of bin column (x, y) for the FP32 prefilter records (any per-column
// origin does; this one keeps local coordinates within about a bin).
__device__ __forceinline__ double col_ox(const PhysParams& P, int x) { return (double)(x - P.BoxX / 2) * P.BoxLength; }
__device__ __forceinline__ double col_oy(const PhysParams& P, int y) { return (double)(y - P.BoxY / 2) * P.BoxLength; }

// FP32 prefilter record of an end point (prefilter_point): x, y relative to
// its column origin (ox, oy).
__device__ __forceinline__ float4 local_point(d3 e, double ox, double oy) {
    return prefilter_point(e.x - ox, e.y - oy, e.z);
}

__global__ void k_bin_count(const double* __restrict__ hot, int n, PhysParams P, int* __restrict__ key,
                            int* __restrict__ cnt, int* __restrict__ slot, float4* __restrict__ cmL,
                            int* __restrict__ err) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Seg s;
    load_seg(hot, i, s);
    const d3 cm = cm_of(s);
    int x, y, z;
    int a = bin_of(P, cm, x, y, z);
    if (a < 0) {
        // reported as an error; the cell is filed in the nearest bin so that
        // every later kernel still sees a complete permutation
        atomicOr(err, ERRBIT_DOMAIN);
        x = x < 0 ? 0 : (x >= P.BoxX ? P.BoxX - 1 : x);
        y = y < 0 ? 0 : (y >= P.BoxY ? P.BoxY - 1 : y);
        z = z < 0 ? 0 : (z >= P.BoxZ ? P.BoxZ - 1 : z);
        a = x * (P.BoxY * P.BoxZ) + y * P.BoxZ + z;
    }
    key[i] = a;
    // FP32 end points for the list build (its walk and pair classification)
    const double ox = col_ox(P, x), oy = col_oy(P, y);
    cmL[2 * i] = local_point(s.p, ox, oy);
    cmL[2 * i + 1] = local_point(s.q, ox, oy);
    slot[i] = atomicAdd(&cnt[a], 1);  // arrival order in the bin
}

// Bin order with the cells of a bin in arrival order.
__global__ void k_bin_place(const int* __restrict__ key, const int* __restrict__ slot, int n,
                            const int* __restrict__ start, int* __restrict__ scattered) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) scattered[start[key[i]] + slot[i]] = i;
}

// Ascending IDs inside each bin: the cell at position p of the arrival order
// goes to its bin's start plus the number of cells of the bin with a smaller
// ID (bins hold tens of cells; a thread per cell), with its candidate record
// (the FP32 end points).
__global__ void k_bin_rank(const int* __restrict__ scattered, int n, const int* __restrict__ key,
                           const int* __restrict__ start, const float4* __restrict__ cmL, int* __restrict__ sorted,
                           float4* __restrict__ cand) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int id = scattered[p];
    const int a = key[id];
    const int s = start[a], e = start[a + 1];
    int r = 0;
    for (int q = s; q < e; q++) r += scattered[q] < id;
    sorted[s + r] = id;
    cand[2 * (s + r)] = cmL[2 * id];
    cand[2 * (s + r) + 1] = cmL[2 * id + 1];
}

// ------------------------------------------------------------ lists ---
// Blocks are dispatched round-robin over the 8 XCDs; remap so that each XCD
// (one L2) owns a contiguous range of bin-ordered slots, i.e. an x-slab.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    constexpr int X = 8;
    const int per = nb / X, rem = nb % X;
    const int xcd = b % X, k = b / X;
    return xcd * per + (xcd < rem ? xcd : rem) + k;
}

// UniformGrid::GetNeighbours (ABM/UniformGrid.h:141-170) + reduceNeighbours
// (ABM/Neighbours.cpp:57-93) for the cell in `slot`, against the current bins.
//
// Two phases so that the FP64 predicate runs converged: the FP32 prefilter
// scan queues surviving candidate positions in LDS (GQ per thread, column
// layout queue[e * blockDim + tid]); a full queue, and the end of the scan,
// flush it through the exact test. Without the queue every lane would run
// the masked FP64 test whenever any lane of its wave had a survivor.
constexpr int GQ = 16;

// The exact test (reduceNeighbours' predicate in FP64) of the queued
// candidate cells, in queue order; kept ones are appended to the cell's list.
// The cell itself, queued like any other candidate, is skipped. The next
// candidate's record is requested before this one is tested.
__device__ __forceinline__ int flush_queue(const PhysParams& P, const double* __restrict__ hot, int self, const Seg& A,
                                           d3 cmA, const int* queue, int qstride, int nq, int* __restrict__ nl,
                                           int stride, int slot, int cap, int m) {
    if (nq == 0) return m;
    int j = queue[0];
    Seg B;
    load_seg(hot, j, B);
    for (int e = 0; e < nq; e++) {
        const int jn = queue[min(e + 1, nq - 1) * qstride];
        Seg Bn;
        load_seg(hot, jn, Bn);
        if ((j != self) & keep_neighbour(A, cmA, B, P.Rc)) {
            if (m < cap) nl[(size_t)m * stride + slot] = j;
            m++;
        }
        j = jn;
        B = Bn;
    }
    return m;
}

// The FP32 form of the first test of reduceNeighbours (prefilter_column and
// PreCol in pair_class.h), against the cell A of a lane: A's centre in FP64,
// its z and reach (prefilter_reach) in FP32.
struct Pre {
    double cx, cy;  // centre of A
    float cz, r;    // centre z; reach plus margin
};
__device__ __forceinline__ Pre prefilter_of(const PhysParams& P, const Seg& A, d3 cmA) {
    Pre f;
    f.cx = cmA.x;
    f.cy = cmA.y;
    f.cz = (float)cmA.z;
    f.r = prefilter_reach(A.L / 2.0 + P.Rc * A.R);
    return f;
}
// The test constants of A against bin column (x, y).
__device__ __forceinline__ PreCol pre_col(const PhysParams& P, const Pre& f, int x, int y) {
    return prefilter_column((float)(f.cx - col_ox(P, x)), (float)(f.cy - col_oy(P, y)), f.cz, f.r);
}
// The verdict: the sign bit of the key (set when an end point is within reach).
__device__ __forceinline__ bool pre_near(const PreCol& c, const float4 p, const float4 q) {
    return (__float_as_uint(c.key(p, q)) >> 31) != 0;
}


// v_rcp_f32: the approximate reciprocal of the FP32 pair classification.
struct DevRcp {
    __device__ float operator()(float x) const { return __builtin_amdgcn_rcpf(x); }
};


__device__ __forceinline__ int gather_one(const PhysParams& P, const double* __restrict__ hot, int id, int slot,
                                          const int* __restrict__ start, const int* __restrict__ sorted,
                                          const float4* __restrict__ cand, int* __restrict__ nl, int stride, int cap,
                                          int* __restrict__ err, int* __restrict__ queue, int qstride) {
    Seg A;
    load_seg(hot, id, A);
    d3 cmA = cm_of(A);
    int x0, y0, z0;
    if (bin_of(P, cmA, x0, y0, z0) < 0) {
        atomicOr(err, ERRBIT_DOMAIN);
        return 0;
    }
    const Pre pre = prefilter_of(P, A, cmA);
    int m = 0, nq = 0;
    auto flush = [&]() { m = flush_queue(P, hot, id, A, cmA, queue, qstride, nq, nl, stride, slot, cap, m); nq = 0; };
    auto push = [&](int k) { queue[nq++ * qstride] = sorted[k]; };
    const int zlo = max(z0 - 1, 0), zhi = min(z0 + 2, P.BoxZ);
    const int ylo = max(y0 - 1, 0), yhi = min(y0 + 2, P.BoxY);
    const int xlo = max(x0 - 1, 0), xhi = min(x0 + 2, P.BoxX);
    // The 27-box set (the list is a set; its order is free): bins adjacent in
    // z are adjacent in the sorted order, so each (x, y) column is one range.
    for (int xa = xlo; xa < xhi; xa++)
        for (int ya = ylo; ya < yhi; ya++) {
            const PreCol col = pre_col(P, pre, xa, ya);
            auto near = [&](int k) { return pre_near(col, cand[2 * k], cand[2 * k + 1]); };
            const int a = xa * (P.BoxY * P.BoxZ) + ya * P.BoxZ;
            const int k1 = start[a + zhi];
            int k = start[a + zlo];
            for (; k + 4 <= k1; k += 4) {
                // flush together: a queue full in one lane empties every queue of the wave
                if (__any(nq > GQ - 4)) flush();
                const bool n0 = near(k), n1 = near(k + 1), n2 = near(k + 2), n3 = near(k + 3);
                if (n0) push(k);
                if (n1) push(k + 1);
                if (n2) push(k + 2);
                if (n3) push(k + 3);
            }
            for (; k < k1; k++) {
                if (__any(nq == GQ)) flush();
                if (near(k)) push(k);
            }
        }
    flush();
    if (m > cap) {
        atomicOr(err, ERRBIT_NLIST);
        m = cap;
    }
    return m;
}

// Orders one wave's LDS writes before its later LDS reads (the lanes of a
// wave run in lockstep; the fences keep the compiler from moving accesses).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ int wave_min(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}

// The list build of the cells of one wave (consecutive in bin order): the
// wave walks the union of their 27-bin blocks once, (x, y) column by column,
// WALK_GROUP candidates at a time. A group's
// records are the same for every lane, so they come by scalar loads, and each
// lane tests them all against its FP32 prefilter, shifting the verdicts into
// a bit mask (one funnel shift of the sign bit a candidate). The mask is then cut to the
// lane's own bins of the column (the position range [klo, khi): its bins
// z0-1..z0+1, empty when the column is not next to its own (x0, y0)) and its set bits
// are queued in ascending order. A lane meets its candidates in ascending bin
// order, as in gather_one, so the lists are the same. Lanes with live = false
// take part in the walk and keep nothing.
//
// The exact test runs on the whole wave's queues at once: its sum of queue
// lengths (cell, candidate) pairs are dealt out evenly over the 64 lanes and
// decided from the FP32 candidate records (classify_pair: the owner's end
// points from LDS, the candidate's from the stream just walked). Each verdict
// is written back into the queue slot (the candidate's id, -1, or for the few
// pairs the FP32 values leave open the marker -2 - k), and each lane then
// appends its kept ids and markers in queue order. After the walk, lanes with
// markers settle them with the reference's FP64 predicate and close the gaps,
// so the FP64 registers are not live during the walk.
__device__ __forceinline__ int gather_wave(const PhysParams& P, const double* __restrict__ hot, bool live, int slot,
                                           int x0, int y0, int z0, int xlo, int xhi, int ylo, int yhi, int zlo, int zhi,
                                           const int* __restrict__ start, const int* __restrict__ sorted,
                                           const float4* __restrict__ cand, int* __restrict__ nl, int stride, int cap,
                                           int* __restrict__ queue, int qstride, unsigned char* own, int* offs,
                                           const float4* op, const float4* oq, const float2* orr, const Pre& pre) {
    int m = 0, nq = 0, namb = 0;
    const int BYZ = P.BoxY * P.BoxZ;
    const int lane = __lane_id();
    const int base = slot - lane;  // the wave's first position in bin order
    const float BL = (float)P.BoxLength;
    int* const qw = queue - lane;  // entry e of lane l: qw[e * qstride + l]
    auto flush = [&]() {
        // exclusive scan of the queue lengths (0..GQ) by bit planes
        int off = 0, tot = 0;
#pragma unroll
        for (int b = 0; (1 << b) <= GQ; b++) {
            const unsigned long long v = __ballot((nq >> b) & 1);
            off += (int)__builtin_amdgcn_mbcnt_hi((unsigned)(v >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)v, 0u)) << b;
            tot += __popcll(v) << b;
        }
        if (tot == 0) return;
        offs[lane] = off;
        for (int e = 0; e < nq; e++) own[off + e] = (unsigned char)lane;
        wave_lds_sync();
        // pair p's owner lane, queue slot, candidate position, column code
        // and records
        struct PairOps {
            int o, k, code, j;
            int* q;
            float4 bp, bq;
        };
        auto pair_ops = [&](int p) {
            PairOps r;
            r.o = own[p];
            r.q = qw + (p - offs[r.o]) * qstride + r.o;
            const unsigned e = (unsigned)*r.q;
            r.k = (int)(e & QK_MASK);
            r.code = (int)(e >> QK_BITS);
            r.bp = cand[2 * r.k];
            r.bq = cand[2 * r.k + 1];
            r.j = sorted[r.k];
            return r;
        };
        if (lane < tot) {
            PairOps c = pair_ops(lane);
            for (int p = lane; p < tot; p += 64) {
                // the lane's next pair is requested before this one is decided
                // (past the last pair, the last pair again: read, not used;
                // the slots read are never those written in this round)
                const PairOps nx = pair_ops(min(p + 64, tot - 1));
                const int dx = c.code / 3 - 1, dy = c.code % 3 - 1;
                int cls = 0;
                if (c.k != base + c.o) {  // the cell itself is skipped
                    const float2 r = orr[c.o];
                    cls = classify_pair(op[c.o], oq[c.o], r.x, r.y, c.bp, c.bq, (float)dx * BL, (float)dy * BL,
                                        DevRcp{});
                }
                *c.q = cls > 0 ? c.j : (cls < 0 ? -2 - c.k : -1);
                c = nx;
            }
        }
        wave_lds_sync();
        for (int e = 0; e < nq; e++) {
            const int j = queue[e * qstride];
            if (j != -1) {
                if (m < cap) nl[(size_t)m * stride + slot] = j;
                m++;
                namb += j < -1;
            }
        }
        nq = 0;
        wave_lds_sync();
    };
    const int zl = max(z0 - 1, 0), zh = min(z0 + 1, P.BoxZ - 1) + 1;
    for (int xa = xlo; xa <= xhi; xa++)
        for (int ya = ylo; ya <= yhi; ya++) {
            const bool ady = live && (abs(xa - x0) <= 1) & (abs(ya - y0) <= 1);
            if (!__any(ady)) continue;
            // the column's bins zlo..zhi are one range of the bin order
            const int a = __builtin_amdgcn_readfirstlane(xa * BYZ + ya * P.BoxZ);
            const int k0 = __builtin_amdgcn_readfirstlane(start[a + zlo]);
            const int k1 = __builtin_amdgcn_readfirstlane(start[a + zhi + 1]);
            const int klo = ady ? start[a + zl] : 0;
            const int khi = ady ? start[a + zh] : 0;
            const PreCol col = pre_col(P, pre, xa, ya);
            const int qcode = ady ? (((xa - x0 + 1) * 3 + (ya - y0 + 1)) << QK_BITS) : 0;
            for (int g = k0; g < k1; g += WALK_GROUP) {
                // the group's records are the same for every lane: scalar loads
                const float4* __restrict__ cg = cand + 2 * (size_t)g;
                // the group's candidates in steps of WALK_STEP (a short last
                // group of a column stops at the next multiple of the step)
                const int n = min(WALK_GROUP, (k1 - g + WALK_STEP - 1) / WALK_STEP * WALK_STEP);
                unsigned bits = 0;
                for (int u0 = 0; u0 < n; u0 += WALK_STEP) {
#pragma unroll
                    for (int u = u0; u < u0 + WALK_STEP; u++)  // bits << 1 | the verdict's sign bit
                        bits = __builtin_amdgcn_alignbit(bits, __float_as_uint(col.key(cg[2 * u], cg[2 * u + 1])), 31);
                }
                // bit WALK_GROUP-1-u: candidate g+u passed the prefilter
                if (n < WALK_GROUP) bits <<= WALK_GROUP - n;
                // only the lane's own bins: positions [klo, khi) of the group
                const int lo = min(max(klo - g, 0), WALK_GROUP), hi = min(max(khi - g, 0), WALK_GROUP);
                const unsigned below_hi = hi >= WALK_GROUP ? ~0u : ~(~0u >> hi);  // bits of u < hi
                const unsigned below_lo = lo >= WALK_GROUP ? ~0u : ~(~0u >> lo);  // bits of u < lo
                bits &= below_hi & ~below_lo;
                // queue the survivors in ascending order: room is made once a
                // group (a lane with more than a queue's worth fills and
                // flushes in turns), then each lane pushes its own
                if (!__any(bits != 0)) continue;
                const int qb = qcode + g;
                auto push = [&]() {
                    const int u = __builtin_clz(bits);
                    bits &= ~(0x80000000u >> u);
                    queue[nq * qstride] = qb + u;
                    nq++;
                };
                if (__any(nq + __popc(bits) > GQ)) {
                    flush();
                    while (__any(__popc(bits) > GQ)) {
                        for (int e = 0; e < GQ && bits != 0; e++) push();
                        flush();
                    }
                }
                while (bits != 0) push();
            }
        }
    flush();
    if (namb > 0) {
        // the FP64 predicate for the pairs the FP32 records left open, in
        // list order, closing the gaps of the rejected ones
        Seg A;
        load_seg(hot, sorted[slot], A);
        const d3 cmA = cm_of(A);
        const int stored = min(m, cap);
        int w = 0;
        for (int e = 0; e < stored; e++) {
            int j = nl[(size_t)e * stride + slot];
            if (j < -1) {
                j = sorted[-2 - j];
                Seg B;
                load_seg(hot, j, B);
                if (!keep_neighbour(A, cmA, B, P.Rc)) continue;
            }
            nl[(size_t)w++ * stride + slot] = j;
        }
        m = w + (m - stored);  // a list past its capacity stays past it (an error)
    }
    return m;
}

// The lists of the cells in bin-order slots [0, n) (order, slot_of and
// nl_count by slot); n <= QK_MASK + 1 (the queue entries' position field).
__global__ void __launch_bounds__(256) k_gather_all(const double* __restrict__ hot, int n, PhysParams P,
                                                    const int* __restrict__ start, const int* __restrict__ sorted,
                                                    const float4* __restrict__ cand, int* __restrict__ nl,
                                                    int* __restrict__ nl_count, int stride, int cap,
                                                    int* __restrict__ order, int* __restrict__ slot_of,
                                                    int* __restrict__ block_max, int* __restrict__ err) {
    __shared__ int queue[GQ * 256];
    __shared__ unsigned char owner[GQ * 256];  // per wave: owner lane of each queued pair
    __shared__ int qoff[256];
    __shared__ float4 own_p[256], own_q[256];  // each cell's FP32 end points (its candidate record)
    __shared__ float2 own_r[256];              // each cell's L/2 + Rc R and Rc R
    __shared__ int wave_lmax[4];
    const int t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int lane = __lane_id(), w = threadIdx.x / 64;
    // every lane of a wave stays to the end (the wave walk uses cross-lane
    // operations); lanes past the last cell are not live
    const bool live = t < n;
    const int id = live ? sorted[t] : 0;
    int x0 = 0, y0 = 0, z0 = 0;
    Pre pre;
    bool inside;
    {
        Seg A;
        load_seg(hot, id, A);
        const d3 cmA = cm_of(A);
        inside = live && bin_of(P, cmA, x0, y0, z0) >= 0;
        pre = prefilter_of(P, A, cmA);
        if (live) {
            own_p[threadIdx.x] = cand[2 * t];
            own_q[threadIdx.x] = cand[2 * t + 1];
            own_r[threadIdx.x] = make_float2((float)(A.L / 2.0 + P.Rc * A.R), (float)(P.Rc * A.R));
        }
    }
    if (live && !inside) atomicOr(err, ERRBIT_DOMAIN);
    const bool walk = inside;
    // the union of the lanes' blocks
    const int xmin = wave_min(walk ? x0 : INT_MAX);
    int m = 0;
    if (xmin != INT_MAX) {  // (no lane inside: nothing to walk)
        // wave-uniform bounds (readfirstlane: the walk's loops and its record
        // addresses are then uniform, so the records come by scalar loads)
        const int xlo = __builtin_amdgcn_readfirstlane(max(xmin - 1, 0));
        const int xhi = __builtin_amdgcn_readfirstlane(min(wave_max(walk ? x0 : INT_MIN) + 1, P.BoxX - 1));
        const int ylo = __builtin_amdgcn_readfirstlane(max(wave_min(walk ? y0 : INT_MAX) - 1, 0));
        const int yhi = __builtin_amdgcn_readfirstlane(min(wave_max(walk ? y0 : INT_MIN) + 1, P.BoxY - 1));
        const int zlo = __builtin_amdgcn_readfirstlane(max(wave_min(walk ? z0 : INT_MAX) - 1, 0));
        const int zhi = __builtin_amdgcn_readfirstlane(min(wave_max(walk ? z0 : INT_MIN) + 1, P.BoxZ - 1));
        wave_lds_sync();  // the owners' records are in LDS
        m = gather_wave(P, hot, walk, t, x0, y0, z0, xlo, xhi, ylo, yhi, zlo, zhi, start, sorted, cand, nl, stride,
                        cap, queue + threadIdx.x, 256, owner + GQ * 64 * w, qoff + 64 * w,
                        own_p + 64 * w, own_q + 64 * w, own_r + 64 * w, pre);
    }
    if (m > cap) {
        atomicOr(err, ERRBIT_NLIST);
        m = cap;
    }
    if (live) {
        order[t] = id;
        slot_of[id] = t;
        nl_count[t] = m;
    }
    // the block's longest list (k_reduce_max folds the blocks' values; one
    // device atomic per wave on a single address cost ~0.15 ms at 2^20 cells)
    const int wm = wave_max(live ? m : 0);
    if (lane == 0) wave_lmax[w] = wm;
    __syncthreads();
    if (threadIdx.x == 0)
        block_max[blockIdx.x] = max(max(wave_lmax[0], wave_lmax[1]), max(wave_lmax[2], wave_lmax[3]));
}

// Folds per-block values into one statistic (one workgroup, one atomic). The
// values are read as 16-byte vectors, four a thread at once (the loads of a
// thread are independent: their latencies overlap); `part` is a whole device
// allocation, so 16-byte aligned.
__global__ void __launch_bounds__(1024) k_reduce_max(const int* __restrict__ part, int n, int* __restrict__ out) {
    __shared__ int red[16];
    int v = 0;
    const int n4 = n >> 2;
    const int4* __restrict__ p4 = reinterpret_cast<const int4*>(part);
    const int T0 = (int)blockDim.x;
    for (int i0 = threadIdx.x; i0 < n4; i0 += 4 * T0) {
        int4 x[4];
#pragma unroll
        for (int k = 0; k < 4; k++) x[k] = p4[min(i0 + k * T0, n4 - 1)];  // (in bounds: i0 < n4)
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (i0 + k * T0 < n4) v = max(max(v, max(x[k].x, x[k].y)), max(x[k].z, x[k].w));
    }
    const int i = 4 * n4 + (int)threadIdx.x;
    if (i < n) v = max(v, part[i]);
    v = wave_max(v);
    if (__lane_id() == 0) red[threadIdx.x / 64] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x / 64); w++) v = max(v, red[w]);
        atomicMax(out, v);
    }
}
__global__ void __launch_bounds__(1024) k_reduce_sum(const unsigned* __restrict__ part, int n,
                                                     unsigned long long* __restrict__ out) {
    __shared__ unsigned long long red[16];
    unsigned long long v = 0;
    const int n4 = n >> 2;
    const int4* __restrict__ p4 = reinterpret_cast<const int4*>(part);
    const int T0 = (int)blockDim.x;
    for (int i0 = threadIdx.x; i0 < n4; i0 += 4 * T0) {
        int4 x[4];
#pragma unroll
        for (int k = 0; k < 4; k++) x[k] = p4[min(i0 + k * T0, n4 - 1)];  // (in bounds: i0 < n4)
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (i0 + k * T0 < n4)
                v += (unsigned long long)(unsigned)x[k].x + (unsigned)x[k].y + (unsigned)x[k].z + (unsigned)x[k].w;
    }
    const int i = 4 * n4 + (int)threadIdx.x;
    if (i < n) v += part[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (__lane_id() == 0) red[threadIdx.x / 64] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x / 64); w++) v += red[w];
        if (v) atomicAdd(out, v);
    }
}

__global__ void __launch_bounds__(256) k_gather_ids(const double* __restrict__ hot, const int* __restrict__ ids, int nid,
                                                    PhysParams P, const int* __restrict__ start,
                                                    const int* __restrict__ sorted, const float4* __restrict__ cand,
                                                    int* __restrict__ nl, int* __restrict__ nl_count, int stride,
                                                    int cap, const int* __restrict__ slot_of,
                                                    int* __restrict__ maxlist, int* __restrict__ err) {
    __shared__ int queue[GQ * 256];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nid) return;
    const int id = ids[i];
    const int t = slot_of[id];
    int m = gather_one(P, hot, id, t, start, sorted, cand, nl, stride, cap, err, queue + threadIdx.x, 256);
    nl_count[t] = m;
    atomicMax(maxlist, m);
}

// -------------------------------------------------------- move + grow ---
// The rest of sum_forces after the cell-cell loop (ABM/Forces.cpp:349-380):
// wall contact, viscous drag, surface tension, planar constraint. F and T
// hold the cell-cell sums on entry.
__device__ __forceinline__ void finish_forces(const PhysParams& P, const Seg& S, const Kin& K, d3 cm, double wall_z,
                                              const SurfOps& so, d3& F, d3& T, d3& sf, d3& df, int& draws) {
    sf = mk(0, 0, 0);
    df = mk(0, 0, 0);
    if (dmin(S.q.z, S.p.z) < 1.2 * S.R) {
        // F_cw: ABM/Forces.cpp:88-204 (one rand() draw per call, counted)
        double L = S.L + 4.0 / 3 * S.R;
        double M = L * S.R * S.R * kPI;
        d3 r1, r2;
        d3 F1 = F_cw_locus(P, S, K, cm, S.p, wall_z, M, r1, sf, df);
        d3 F2 = F_cw_locus(P, S, K, cm, S.q, wall_z, M, r2, sf, df);
        draws++;
        F = add(F, F1);
        T = add(T, cross(sub(r1, cm), F1));
        F = add(F, F2);
        T = add(T, cross(sub(r2, cm), F2));
    }
    // F_v: ABM/Forces.cpp:320-329
    F = add(F, scl(K.v, -P.viscosity * 6.0 * kPI * S.R));
    T = add(T, scl(K.w, -P.viscosity * 6.0 * kPI * S.R));
    d3 Fs, Ts;
    F_surf_tension_ops(P, S, cm, so, Fs, Ts);
    F = add(F, Fs);
    T = add(T, Ts);
    if (P.planar) {
        T.z = 0;
        F.y = 0;
    }
}

// sum_forces (ABM/Forces.cpp:332-381) of cell S against its list.
__device__ __forceinline__ void sum_forces(const PhysParams& P, const Seg& S, const Kin& K, const double* __restrict__ hot,
                                           const int* __restrict__ nl, int stride, int slot, int cnt, double wall_z,
                                           const double* __restrict__ H, const double* __restrict__ N, d3& F, d3& T,
                                           d3& sf, d3& df, int& draws) {
    F = mk(0, 0, 0);
    T = mk(0, 0, 0);
    const d3 cm = cm_of(S);
    // the next neighbour's record is requested before this one is used
    int j = cnt > 0 ? nl[slot] : 0;
    Seg B;
    if (cnt > 0) load_seg(hot, j, B);
    for (int k = 0; k < cnt; k++) {
        int jn = j;
        Seg Bn = B;
        if (k + 1 < cnt) {
            jn = nl[(size_t)(k + 1) * stride + slot];
            load_seg(hot, jn, Bn);
        }
        F_cc_acc(P, S, K, cm, B, hot, j, F, T);
        j = jn;
        B = Bn;
    }
    finish_forces(P, S, K, cm, wall_z, surf_ops(P, S, cm, H, N), F, T, sf, df, draws);
}

// Cell state of one block of k_move_grow staged in LDS for its pair phases
// (structure of arrays, one column per cell of the block).
// Pairs tested for contact at once (a super-chunk of PAIR_SUPER chunks).
constexpr int PAIR_SUPER = 8;
constexpr int PAIR_BATCH = 4;  // chunks whose loads are issued together (divides PAIR_SUPER)
static_assert(PAIR_SUPER % PAIR_BATCH == 0, "whole batches");
static_assert(MOVE_BLOCK == 64, "the pair dealing of pair_sums is one wave");

struct MoveLds {
    double seg[8][MOVE_BLOCK];  // p.x p.y p.z q.x q.y q.z L R
    double kin[6][MOVE_BLOCK];  // v.x v.y v.z w.x w.y w.z
    double con[6][MOVE_BLOCK];  // one chunk of pair contributions: F.xyz, torque.xyz
    int off[MOVE_BLOCK + 1];    // exclusive scan of the cells' list lengths
    int scan[48];
    int draws[MOVE_BLOCK / 64];
    int cp_j[PAIR_SUPER * MOVE_BLOCK];            // a super-chunk's pairs that may touch: neighbour id,
    unsigned char cp_c[PAIR_SUPER * MOVE_BLOCK];  // and owning cell, in pair order
};

__device__ __forceinline__ void lds_put_seg(MoveLds& L, int c, const Seg& S) {
    L.seg[0][c] = S.p.x;
    L.seg[1][c] = S.p.y;
    L.seg[2][c] = S.p.z;
    L.seg[3][c] = S.q.x;
    L.seg[4][c] = S.q.y;
    L.seg[5][c] = S.q.z;
    L.seg[6][c] = S.L;
    L.seg[7][c] = S.R;
}
__device__ __forceinline__ void lds_put_kin(MoveLds& L, int c, const Kin& K) {
    L.kin[0][c] = K.v.x;
    L.kin[1][c] = K.v.y;
    L.kin[2][c] = K.v.z;
    L.kin[3][c] = K.w.x;
    L.kin[4][c] = K.w.y;
    L.kin[5][c] = K.w.z;
}
__device__ __forceinline__ Seg lds_seg(const MoveLds& L, int c) {
    Seg S;
    S.p = mk(L.seg[0][c], L.seg[1][c], L.seg[2][c]);
    S.q = mk(L.seg[3][c], L.seg[4][c], L.seg[5][c]);
    S.L = L.seg[6][c];
    S.R = L.seg[7][c];
    return S;
}
__device__ __forceinline__ Kin lds_kin(const MoveLds& L, int c) {
    Kin K;
    K.v = mk(L.kin[0][c], L.kin[1][c], L.kin[2][c]);
    K.w = mk(L.kin[3][c], L.kin[4][c], L.kin[5][c]);
    return K;
}

// The cell-cell sums (the neighbour loop of ABM/Forces.cpp:341-347) of every
// cell of the block against the segments and kinematics staged in L. The
// block's (cell, list entry) pairs are spread over its threads, a super-chunk
// of PAIR_SUPER * MOVE_BLOCK pairs at a time. Every pair is first tested in
// FP32 for "surely not in contact" (surely_apart: F_cc would then add only
// signed zeros, which leave a sum that started at +0 unchanged). The others,
// about one in six on a dense colony, are compacted in pair order and
// evaluated by F_cc, one per thread, MOVE_BLOCK at a time into the chunk
// buffer; each cell's own thread then adds its pairs in list order. These are
// the reference's additions in the reference's order, so F and T are
// bit-identical to the sequential loop, while no lane idles behind a longer
// list elsewhere in its wave.
__device__ __forceinline__ void pair_sums(const PhysParams& P, MoveLds& L, const double* __restrict__ hot,
                                          const int* __restrict__ nl, int stride, int slot0, d3& F, d3& T) {
    __syncthreads();
    const int tid = threadIdx.x;
    const int total = L.off[MOVE_BLOCK];
    const int lo = L.off[tid], hi = L.off[tid + 1];
    F = mk(0, 0, 0);
    T = mk(0, 0, 0);
    auto owner = [&](int p) {  // the largest c with off[c] <= p (off[0] = 0)
        int c = 0;
#pragma unroll
        for (int s = MOVE_BLOCK / 2; s > 0; s >>= 1)
            if (L.off[c + s] <= p) c += s;
        return c;
    };
    constexpr int SC = PAIR_SUPER * MOVE_BLOCK;
    for (int sbase = 0; sbase < total; sbase += SC) {
        unsigned long long touch[PAIR_SUPER];
        int nc = 0;
        // PAIR_BATCH chunks at a time: their owners, list entries and
        // records are all requested before any is tested (one round trip to
        // memory a batch; a pair past the end reads the last pair's and
        // keeps nothing)
#pragma unroll
        for (int r0 = 0; r0 < PAIR_SUPER; r0 += PAIR_BATCH) {
            if (sbase + r0 * MOVE_BLOCK >= total) {  // (uniform)
#pragma unroll
                for (int b = 0; b < PAIR_BATCH; b++) touch[r0 + b] = 0;
                continue;
            }
            int cb[PAIR_BATCH], jb[PAIR_BATCH];
#pragma unroll
            for (int b = 0; b < PAIR_BATCH; b++) {
                const int p = min(sbase + (r0 + b) * MOVE_BLOCK + tid, total - 1);
                cb[b] = owner(p);
                jb[b] = nl[(size_t)(p - L.off[cb[b]]) * stride + slot0 + cb[b]];
            }
            Seg Bb[PAIR_BATCH];
#pragma unroll
            for (int b = 0; b < PAIR_BATCH; b++) load_seg(hot, jb[b], Bb[b]);
#pragma unroll
            for (int b = 0; b < PAIR_BATCH; b++) {
                const int r = r0 + b;
                const int c = cb[b], j = jb[b];
                bool t = false;
                if (sbase + r * MOVE_BLOCK + tid < total) {
                    const Seg& B = Bb[b];
                    const Seg A = lds_seg(L, c);
                    const d3 o = cm_of(A);
                    auto rel = [&](d3 e) {
                        return make_float4((float)(e.x - o.x), (float)(e.y - o.y), (float)(e.z - o.z), 0.f);
                    };
                    t = !surely_apart(rel(A.p), rel(A.q), rel(B.p), rel(B.q), (float)(A.R + B.R), DevRcp{});
                }
                touch[r] = __ballot(t);
                if (t) {
                    const int k = nc + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(touch[r] >> 32),
                                                                     __builtin_amdgcn_mbcnt_lo((unsigned)touch[r], 0u));
                    L.cp_j[k] = j;
                    L.cp_c[k] = (unsigned char)c;
                }
                nc += __popcll(touch[r]);
            }
        }
        // this cell's pairs of the super-chunk that may touch: [cs, ce) of the compact list
        auto before = [&](int p) {  // pairs that may touch in [sbase, p), sbase <= p <= sbase + SC
            const int q = p - sbase;
            int m = 0;
#pragma unroll
            for (int r = 0; r < PAIR_SUPER; r++) {
                const int bq = min(max(q - r * MOVE_BLOCK, 0), MOVE_BLOCK);
                m += __popcll(bq >= 64 ? touch[r] : (touch[r] & ((1ull << bq) - 1ull)));
            }
            return m;
        };
        const int cs = before(min(max(lo, sbase), sbase + SC)), ce = before(min(max(hi, sbase), sbase + SC));
        __syncthreads();
        for (int base = 0; base < nc; base += MOVE_BLOCK) {
            const int k = base + tid;
            if (k < nc) {
                const int c = L.cp_c[k], j = L.cp_j[k];
                Seg B;
                load_seg(hot, j, B);
                Kin KB;  // requested with the segment (F_cc reads it when the cells touch)
                load_kin(hot, j, KB);
                const Seg A = lds_seg(L, c);
                d3 f, tc;
                F_cc(
                    P, A, cm_of(A), B, [&]() { return lds_kin(L, c); }, [&]() { return KB; }, f, tc);
                L.con[0][tid] = f.x;
                L.con[1][tid] = f.y;
                L.con[2][tid] = f.z;
                L.con[3][tid] = tc.x;
                L.con[4][tid] = tc.y;
                L.con[5][tid] = tc.z;
            }
            __syncthreads();
            const int a = cs > base ? cs : base;
            const int b = ce < base + MOVE_BLOCK ? ce : base + MOVE_BLOCK;
            for (int q = a - base; q < b - base; q++) {
                F = add(F, mk(L.con[0][q], L.con[1][q], L.con[2][q]));
                T = add(T, mk(L.con[3][q], L.con[4][q], L.con[5][q]));
            }
            __syncthreads();
        }
    }
}

// MoveCell + GrowCell for every cell (ABM/Simulation.cpp:330-335): integrate
// (ABM/Integrate.cpp:11-31) with both force sums, then grow. Block b owns the
// bin-ordered slots [b' * MOVE_BLOCK, (b' + 1) * MOVE_BLOCK), b' its XCD remap.
__global__ void __launch_bounds__(MOVE_BLOCK) k_move_grow(
    const double* __restrict__ hot_old, double* __restrict__ hot_new, int n, PhysParams P, double dt,
    const int* __restrict__ order, const int* __restrict__ nl, const int* __restrict__ nl_count, int stride,
    const double* __restrict__ H, const double* __restrict__ Nrm, const double* __restrict__ wall,
    const double* __restrict__ gr1, const double* __restrict__ gr2, const int* __restrict__ type,
    double* __restrict__ fric, int* __restrict__ fric_id, unsigned* __restrict__ divbits,
    unsigned* __restrict__ block_draws, int* __restrict__ err) {
    __shared__ MoveLds L;
    const int tid = threadIdx.x;
    const int slot0 = xcd_block(blockIdx.x, gridDim.x) * MOVE_BLOCK;
    const int t = slot0 + tid;
    int draws = 0;
    bool live = false;  // a cell inside the bin grid
    SurfOps so{};       // F_surf_tension's operands at the state of the next pass
    int a2 = -1;        // bin of the half-step position (where the cell grows)
    double GRa = 0;     // the growth rate there (read before pass 2)
    int id = 0, cnt = 0;
    double Ldiv = 0, wall_z = 0;
    if (t < n) {
        id = order[t];
        Seg S;
        Kin K;
        load_seg(hot_old, id, S);
        load_kin(hot_old, id, K);
        Ldiv = hot_old[(size_t)id * HOT + 14];
        int x, y, z;
        if (bin_of(P, cm_of(S), x, y, z) < 0) {  // MoveCell oldAddress: ABM/Cells.cpp:25-26
            // reported; the cell is carried over unchanged and does not divide
            atomicOr(err, ERRBIT_DOMAIN);
            const double2* src = reinterpret_cast<const double2*>(hot_old + (size_t)id * HOT);
            double2* dst = reinterpret_cast<double2*>(hot_new + (size_t)id * HOT);
            for (int c = 0; c < HOT / 2; c++) dst[c] = src[c];
            double2* fo = reinterpret_cast<double2*>(fric + 6 * (size_t)t);
            fo[0] = fo[1] = fo[2] = make_double2(0.0, 0.0);
            fric_id[t] = id;
        } else {
            live = true;
            wall_z = wall[x * P.BoxY + y];
            cnt = nl_count[t];
            lds_put_seg(L, tid, S);
            lds_put_kin(L, tid, K);
            so = surf_ops(P, S, cm_of(S), H, Nrm);  // (its loads overlap the pair sums)
        }
    }
    int total;
    const int o = block_excl_scan(cnt, L.scan, &total);
    L.off[tid] = o;
    if (tid == 0) L.off[MOVE_BLOCK] = total;

    // pass 1 at the old state
    d3 F, T, sf, df;
    pair_sums(P, L, hot_old, nl, stride, slot0, F, T);
    if (live) {
        const Seg S = lds_seg(L, tid);
        const Kin K = lds_kin(L, tid);
        const d3 cm = cm_of(S);
        finish_forces(P, S, K, cm, wall_z, so, F, T, sf, df, draws);
        // UpdatePositions: ABM/Integrate.cpp:34-90
        Kin K1;
        half_kick(dt, F, T, S, K, K1);
        Seg S1;
        S1.p = add(S.p, scl(add(K1.v, cross(K1.w, sub(S.p, cm))), dt));
        S1.q = add(S.q, scl(add(K1.v, cross(K1.w, sub(S.q, cm))), dt));
        S1.L = S.L;
        S1.R = S.R;
        if (P.planar) {
            S1.p.y = 0;
            S1.q.y = 0;
        }
        lds_put_seg(L, tid, S1);
        lds_put_kin(L, tid, K1);
        so = surf_ops(P, S1, cm_of(S1), H, Nrm);
        // pass 2 changes only the velocities: the position the cell grows
        // at is known here
        int x2, y2, z2;
        a2 = bin_of(P, cm_of(S1), x2, y2, z2);
        if (a2 >= 0) GRa = (P.crossFeeding && type[id] != 1) ? gr2[a2] : gr1[a2];
    }
    // pass 2 at the half step, against the neighbours' old records
    pair_sums(P, L, hot_old, nl, stride, slot0, F, T);
    if (live) {
        Seg S1 = lds_seg(L, tid);
        const Kin K1 = lds_kin(L, tid);
        finish_forces(P, S1, K1, cm_of(S1), wall_z, so, F, T, sf, df, draws);
        // UpdateVelocities: ABM/Integrate.cpp:93-127
        Kin K2;
        half_kick(dt, F, T, S1, K1, K2);
        if (P.planar) K2.v.y = 0;
        // grow: ABM/grow.cpp:17-102 at the MoveCell newAddress (ABM/Cells.cpp:32)
        double GR = 0;
        if (a2 < 0) {
            atomicOr(err, ERRBIT_DOMAIN);
        } else {
            GR = GRa * 1.5850;
            d3 v = sub(S1.q, S1.p);
            double dL = S1.L * dt * GR;
            d3 dv = scl(v, dL * 0.5 / S1.L);
            S1.L += dL;
            S1.p = sub(S1.p, dv);
            S1.q = add(S1.q, dv);
        }
        double2* out = reinterpret_cast<double2*>(hot_new + (size_t)id * HOT);
        out[0] = make_double2(S1.p.x, S1.p.y);
        out[1] = make_double2(S1.p.z, S1.q.x);
        out[2] = make_double2(S1.q.y, S1.q.z);
        out[3] = make_double2(K2.v.x, K2.v.y);
        out[4] = make_double2(K2.v.z, K2.w.x);
        out[5] = make_double2(K2.w.y, K2.w.z);
        out[6] = make_double2(S1.L, S1.R);
        out[7] = make_double2(Ldiv, GR);
        // DynFric, StaFric in slot order (coalesced), with the cell's ID
        double2* fo = reinterpret_cast<double2*>(fric + 6 * (size_t)t);
        fo[0] = make_double2(df.x, df.y);
        fo[1] = make_double2(df.z, sf.x);
        fo[2] = make_double2(sf.y, sf.z);
        fric_id[t] = id;
        if (S1.L >= Ldiv) atomicOr(&divbits[id >> 5], 1u << (id & 31));  // GrowCell: ABM/Cells.cpp:48
    }
    // the block's count of F_cw draws (integers: the fold's order is free)
    int w = draws;
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) w += __shfl_xor(w, s, 64);
    if ((tid & 63) == 0) L.draws[tid >> 6] = w;
    __syncthreads();
    if (tid == 0) {
        int s = 0;
        for (int i = 0; i < MOVE_BLOCK / 64; i++) s += L.draws[i];
        block_draws[blockIdx.x] = (unsigned)s;  // folded by k_reduce_sum
    }
}

// Pass-1 forces only, in ID order (bcg_sum_forces).
__global__ void __launch_bounds__(MOVE_BLOCK) k_sum_forces(const double* __restrict__ hot, int n, PhysParams P,
                                                           const int* __restrict__ order, const int* __restrict__ nl,
                                                           const int* __restrict__ nl_count, int stride,
                                                           const double* __restrict__ H, const double* __restrict__ Nrm,
                                                           const double* __restrict__ wall, double* __restrict__ Fo,
                                                           double* __restrict__ To, int* __restrict__ err) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int id = order[t];
    Seg S;
    Kin K;
    load_seg(hot, id, S);
    load_kin(hot, id, K);
    int x, y, z;
    if (bin_of(P, cm_of(S), x, y, z) < 0) {
        atomicOr(err, ERRBIT_DOMAIN);
        return;
    }
    d3 F, T, sf, df;
    int draws = 0;
    sum_forces(P, S, K, hot, nl, stride, t, nl_count[t], wall[x * P.BoxY + y], H, Nrm, F, T, sf, df, draws);
    Fo[3 * (size_t)id] = F.x;
    Fo[3 * (size_t)id + 1] = F.y;
    Fo[3 * (size_t)id + 2] = F.z;
    To[3 * (size_t)id] = T.x;
    To[3 * (size_t)id + 1] = T.y;
    To[3 * (size_t)id + 2] = T.z;
}

// mean_stress (ABM/Compute.cpp:18-117) of every cell at the current state:
// the diagonal of its stress tensor from the contact forces of its
// neighbours (F_cc, in list order), of the agar (F_cw, one rand() draw per
// call, counted) and of the surface tension, each times its point of action
// relative to the centre, over the cell volume; the radial and azimuthal
// components; and the net force. S5 = (xx, yy, zz, rr, tt), by cell ID.
__global__ void __launch_bounds__(MOVE_BLOCK) k_mean_stress(const double* __restrict__ hot, int n, PhysParams P,
                                                            const int* __restrict__ order, const int* __restrict__ nl,
                                                            const int* __restrict__ nl_count, int stride,
                                                            const double* __restrict__ H, const double* __restrict__ Nrm,
                                                            const double* __restrict__ wall, double* __restrict__ S5,
                                                            double* __restrict__ F3,
                                                            unsigned* __restrict__ wave_draws,
                                                            int* __restrict__ err) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int id = order[t];
    Seg S;
    Kin K;
    load_seg(hot, id, S);
    load_kin(hot, id, K);
    const d3 cm = cm_of(S);
    int x, y, z;
    if (bin_of(P, cm, x, y, z) < 0) {
        atomicOr(err, ERRBIT_DOMAIN);
        return;
    }
    const double V = S.L * kPI * S.R * S.R + 4.0 / 3.0 * kPI * S.R * S.R * S.R;
    double xx = 0, yy = 0, zz = 0;
    d3 Fn = mk(0, 0, 0);
    bool drew = false;
    const int cnt = nl_count[t];
    for (int k = 0; k < cnt; k++) {
        const int j = nl[(size_t)k * stride + t];
        Seg B;
        load_seg(hot, j, B);
        d3 F, r;
        F_cc_r(
            P, S, cm, B, [&]() { return K; },
            [&]() {
                Kin KB;
                load_kin(hot, j, KB);
                return KB;
            },
            F, r);
        Fn = add(Fn, F);
        xx -= F.x * r.x;
        yy -= F.y * r.y;
        zz -= F.z * r.z;
    }
    if (dmin(S.q.z, S.p.z) < 1.2 * S.R) {
        const double L = S.L + 4.0 / 3 * S.R;
        const double M = L * S.R * S.R * kPI;
        d3 r1, r2, sf = mk(0, 0, 0), df = mk(0, 0, 0);
        const double wall_z = wall[x * P.BoxY + y];
        const d3 F1 = F_cw_locus(P, S, K, cm, S.p, wall_z, M, r1, sf, df);
        const d3 F2 = F_cw_locus(P, S, K, cm, S.q, wall_z, M, r2, sf, df);
        drew = true;
        Fn = add(Fn, F1);
        r1 = sub(r1, cm);
        xx -= F1.x * r1.x;
        yy -= F1.y * r1.y;
        zz -= F1.z * r1.z;
        Fn = add(Fn, F2);
        r2 = sub(r2, cm);
        xx -= F2.x * r2.x;
        yy -= F2.y * r2.y;
        zz -= F2.z * r2.z;
    }
    d3 Fs, Ts;
    F_surf_tension(P, S, cm, H, Nrm, Fs, Ts);
    Fn = add(Fn, Fs);
    d3 r = sub(S.p, cm);
    r.z += S.R;
    d3 r2 = sub(S.q, cm);
    r2.z += S.R;
    xx -= Fs.x * r.x * 0.5 - Fs.x * r2.x * 0.5;
    yy -= Fs.y * r.y * 0.5 - Fs.y * r2.y * 0.5;
    zz -= Fs.z * r.z * 0.5 - Fs.z * r2.z * 0.5;
    xx /= V;
    yy /= V;
    zz /= V;
    double rr, tt;
    const double Lcm = sqrt(cm.x * cm.x + cm.y * cm.y);
    if (Lcm > DBL_EPSILON) {
        const double s = 1.0 / Lcm;
        const double ux = cm.x * s, uy = cm.y * s;
        rr = xx * ux + yy * uy;
        tt = -xx * uy + yy * ux;
    } else {
        rr = xx + yy;
        tt = 0.0;
    }
    double* o = S5 + 5 * (size_t)id;
    o[0] = xx;
    o[1] = yy;
    o[2] = zz;
    o[3] = rr;
    o[4] = tt;
    F3[3 * (size_t)id] = Fn.x;
    F3[3 * (size_t)id + 1] = Fn.y;
    F3[3 * (size_t)id + 2] = Fn.z;
    // the wave's count of F_cw draws, by its first remaining lane (the
    // caller zeroes the counts; k_reduce_sum folds them)
    const unsigned long long d = __ballot(drew);
    if (__lane_id() == __builtin_ctzll(__ballot(1))) wave_draws[t / 64] = (unsigned)__popcll(d);
}

// ----------------------------------------------------------- compact ---
// Dividing cells from the flag bitmap (a word of 32 IDs a thread): the
// number per block, then the IDs in ascending order (block offset + rank in
// the block); the second pass clears the words it read.
__global__ void __launch_bounds__(256) k_flag_count(const unsigned* __restrict__ bits, int nw, int* __restrict__ tot) {
    __shared__ int lds[4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    int c = i < nw ? __popc(bits[i]) : 0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) tot[blockIdx.x] = lds[0] + lds[1] + lds[2] + lds[3];
}

__global__ void __launch_bounds__(256) k_flag_write(unsigned* __restrict__ bits, int nw, const int* __restrict__ off,
                                                    int* __restrict__ ids) {
    __shared__ int lds[96];
    const int i = blockIdx.x * 256 + threadIdx.x;
    unsigned w = i < nw ? bits[i] : 0u;
    int o = off[blockIdx.x] + block_excl_scan(__popc(w), lds, nullptr);
    if (w) bits[i] = 0u;
    while (w) {
        const int b = __builtin_ctz(w);
        ids[o++] = (i << 5) + b;
        w &= w - 1;
    }
}

// ------------------------------------------------------------ divide ---
// divide (ABM/grow.cpp:107-186) + DivideCell bookkeeping, one thread per
// division; u holds the 9 rand() values of each division's draws, mapped
// here to ((float)rand()/RAND_MAX - 0.5) as the reference evaluates it
// ((float)RAND_MAX is 2^31, so the division is an exact scaling).
__global__ void k_divide(double* __restrict__ hot, int n0, int D, const int* __restrict__ div_ids,
                         const int32_t* __restrict__ u, PhysParams P, double t, double* __restrict__ mother_pre,
                         int* __restrict__ div_index_of, double* __restrict__ time_p, double* __restrict__ time_q,
                         int* __restrict__ age_p, int* __restrict__ age_q, int* __restrict__ type,
                         int* __restrict__ ancestor, double* __restrict__ fric, int* __restrict__ fric_id,
                         int* __restrict__ order, int* __restrict__ slot_of, int* __restrict__ nl_count) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= D) return;
    const int m = div_ids[k];
    const int d = n0 + k;
    double* hm = hot + (size_t)m * HOT;
    double* hd = hot + (size_t)d * HOT;
    for (int i = 0; i < HOT; i++) mother_pre[(size_t)k * HOT + i] = hm[i];
    div_index_of[m] = k;
    double r[9];
#pragma unroll
    for (int q = 0; q < 9; q++) r[q] = (double)((float)u[9 * (size_t)k + q] / (float)2147483647) - 0.5;
    d3 mp = mk(hm[0], hm[1], hm[2]), mq = mk(hm[3], hm[4], hm[5]);
    const d3 mv = mk(hm[6], hm[7], hm[8]);
    double mL = hm[12], mR = hm[13];
    const double GR = hm[15];

    double dl = r[0] * P.varL;
    double dangle = r[1] * P.varAngle;
    d3 pq = sub(mp, mq);
    d3 ddir1 = mk(r[2], r[3], r[4]);
    ddir1 = sub(ddir1, scl(pq, dot(ddir1, pq) / dot(pq, pq)));
    ddir1 = scl(ddir1, 1 / sqrt(dot(ddir1, ddir1)));
    d3 dOrien1 = scl(ddir1, dangle);
    d3 ddir2 = mk(r[5], r[6], r[7]);
    ddir2 = sub(ddir2, scl(pq, dot(ddir2, pq) / dot(pq, pq)));
    ddir2 = scl(ddir2, 1 / sqrt(dot(ddir2, ddir2)));
    d3 dOrien2 = scl(ddir2, dangle);

    d3 uv = sub(mp, mq);
    double L = mL;
    uv = scl(uv, 1 / L);
    d3 dq = mq;
    double dR = P.cellRadius;
    mL = L / 2 - mR + dl;
    double dLen = L / 2 + mR - 2 * dR - dl;
    double dwz = r[8] * 0.001;
    mq = sub(mp, scl(uv, mL));
    time_q[d] = time_q[m];
    time_q[m] = t;
    time_p[d] = t;
    age_p[m] = age_p[m] + 1;
    age_q[d] = age_q[m] + 1;
    age_p[d] = 0;  // the reference leaves it as found in fresh (zeroed) storage
    d3 dp = sub(mq, scl(uv, dR + mR));
    dq = sub(dp, scl(uv, dLen));
    mp = add(mp, scl(dOrien1, mL));
    mq = sub(mq, scl(dOrien1, mL));
    dp = add(dp, scl(dOrien2, dLen));
    dq = sub(dq, scl(dOrien2, dLen));
    type[d] = type[m];
    ancestor[d] = ancestor[m];
    const double mLdiv = P.L_divide / 2 + P.cellRadius + mL;
    const double dLdiv = P.L_divide / 2 + P.cellRadius + dLen;
    for (int c = 0; c < 6; c++) fric[6 * (size_t)d + c] = 0;  // the daughter's slot is its ID
    fric_id[d] = d;
    hm[0] = mp.x; hm[1] = mp.y; hm[2] = mp.z;
    hm[3] = mq.x; hm[4] = mq.y; hm[5] = mq.z;
    hm[9] = 0; hm[10] = 0; hm[11] = 0;
    hm[12] = mL;
    hm[14] = mLdiv;
    hd[0] = dp.x; hd[1] = dp.y; hd[2] = dp.z;
    hd[3] = dq.x; hd[4] = dq.y; hd[5] = dq.z;
    hd[6] = mv.x; hd[7] = mv.y; hd[8] = mv.z;
    hd[9] = 0; hd[10] = 0; hd[11] = dwz;
    hd[12] = dLen;
    hd[13] = dR;
    hd[14] = dLdiv;
    hd[15] = GR;
    order[d] = d;
    slot_of[d] = d;
    nl_count[d] = 0;
}

// ------------------------------------------------------------- patch ---
__device__ __forceinline__ void vload(const Versioned& V, int j, int k, Seg& s) {
    const int kj = V.div_index_of[j];
    if (kj > k) load_seg(V.mother_pre, kj, s);
    else load_seg(V.hot, j, s);
}

__device__ int gather_versioned(const PhysParams& P, const Versioned& V, int c, int k, int D,
                                const int* __restrict__ div_ids, const int* __restrict__ start,
                                const int* __restrict__ sorted, int* __restrict__ nl, int stride, int cap, int slot,
                                int* __restrict__ err) {
    Seg A;
    vload(V, c, k, A);
    const d3 cmA = cm_of(A);
    int x0, y0, z0;
    if (bin_of(P, cmA, x0, y0, z0) < 0) {
        atomicOr(err, ERRBIT_DOMAIN);
        return 0;
    }
    const int zlo = max(z0 - 1, 0), zhi = min(z0 + 2, P.BoxZ);
    const int ylo = max(y0 - 1, 0), yhi = min(y0 + 2, P.BoxY);
    const int xlo = max(x0 - 1, 0), xhi = min(x0 + 2, P.BoxX);
    int m = 0;
    for (int za = zlo; za < zhi; za++)
        for (int ya = ylo; ya < yhi; ya++)
            for (int xa = xlo; xa < xhi; xa++) {
                const int a = xa * (P.BoxY * P.BoxZ) + ya * P.BoxZ + za;
                for (int q = start[a]; q < start[a + 1]; q++) {
                    const int j = sorted[q];
                    if (j == c || j > V.n0 + k || V.div_index_of[j] > k) continue;
                    Seg B;
                    load_seg(V.hot, j, B);
                    if (keep_neighbour(A, cmA, B, P.Rc)) {
                        if (m < cap) nl[(size_t)m * stride + slot] = j;
                        m++;
                    }
                }
            }
    // mothers of later divisions, in the boxes they occupied before dividing
    for (int kk = k + 1; kk < D; kk++) {
        const int j = div_ids[kk];
        if (j == c) continue;
        Seg B;
        load_seg(V.mother_pre, kk, B);
        int bx, by, bz;
        if (bin_of(P, cm_of(B), bx, by, bz) < 0) continue;
        if (bx < xlo || bx >= xhi || by < ylo || by >= yhi || bz < zlo || bz >= zhi) continue;
        if (keep_neighbour(A, cmA, B, P.Rc)) {
            if (m < cap) nl[(size_t)m * stride + slot] = j;
            m++;
        }
    }
    if (m > cap) {
        atomicOr(err, ERRBIT_NLIST);
        m = cap;
    }
    return m;
}

// ABM/Simulation.cpp:395-403 for every division in order: the mother, the
// mother's current neighbours and the daughter are re-gathered against the
// grid as it stands after that division.
__global__ void __launch_bounds__(PATCH_BLOCK) k_patch(PhysParams P, Versioned V, int D, const int* __restrict__ div_ids,
                                                       const int* __restrict__ start, const int* __restrict__ sorted,
                                                       int* __restrict__ nl, int* __restrict__ nl_count, int stride,
                                                       int cap, const int* __restrict__ slot_of,
                                                       int* __restrict__ maxlist, int* __restrict__ err) {
    extern __shared__ int members[];  // cap + 2
    __shared__ int nmem;
    for (int k = 0; k < D; k++) {
        if (threadIdx.x == 0) {
            const int m = div_ids[k];
            const int sm = slot_of[m];
            const int c = nl_count[sm];
            members[0] = m;
            for (int i = 0; i < c; i++) members[1 + i] = nl[(size_t)i * stride + sm];
            members[c + 1] = V.n0 + k;
            nmem = c + 2;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < nmem; i += blockDim.x) {
            const int c = members[i];
            const int s = slot_of[c];
            const int cnt = gather_versioned(P, V, c, k, D, div_ids, start, sorted, nl, stride, cap, s, err);
            nl_count[s] = cnt;
            atomicMax(maxlist, cnt);
        }
        __syncthreads();
    }
}

__global__ void k_reset_div_index(int* __restrict__ div_index_of, const int* __restrict__ div_ids, int D) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < D) div_index_of[div_ids[k]] = -1;
}

}  // namespace bcg
