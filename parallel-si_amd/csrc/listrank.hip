// List ranking on gfx950: Helman-JaJa sublist walks and Wyllie pointer
// jumping, chosen by mean list length.
//
// Reference roles: the splitter pick, sublist walk, reduced-list walk and
// broadcast of gpu_list_ranking.cu:96-110, :153-193, :197-240 (host),
// :244-277 and cpu_list_ranking.cpp:108-239; the tree-Wyllie pointer jumping of
// gpu_parity.cu:79-106.  The design here differs:
//   * splitters are one per aligned block of D node ids, at a hashed offset
//     inside the block, so "is x a splitter" and "which sublist does it start"
//     are pure arithmetic on x (no marker writes into succ, which stays
//     read-only), and structured lists (Euler tours) get evenly mixed splitters;
//   * every list head is a splitter of its own (a "head slot");
//   * a walk longer than Hc hops is cut and continued as a new overflow
//     sublist, which bounds the local rank so (sublist, local rank) packs into
//     the 32-bit output word during the walk;
//   * the reduced list is ranked on the device, recursively by the same walk
//     with weights, and the top level (<= a few thousand nodes) by pointer
//     jumping inside one workgroup's LDS: no host round trip;
//   * Wyllie ranks along predecessors with the (pointer, prefix) pair packed in
//     one 64-bit word per node (ping-pong buffers), with a device-side count of
//     still-active nodes so rounds after convergence exit at once.
// Output: rank[i] = number of nodes before i on its list (0 at the head), or -1
// when no given head reaches i (cycles, or heads that were not passed).
#include "common.h"

#include <algorithm>
#include <vector>

namespace psi {
namespace lr {

constexpr int kTopThreads = 1024;

// Defaults measured on the 16,777,216-node list (MI355X, us per ranking,
// level-0 / reduced-level splitter densities): 2^-4 / 2^-3: 753-779,
// 2^-4 / 2^-2: 768, 2^-4 / 2^-4: 785, 2^-5 / 2^-3: 760, 2^-5 / 2^-2: 743-747,
// 2^-5 / 2^-4: 802, 2^-3 / 2^-3: 794.  The reduced levels' walks are bound by
// their longest sublist, the level-0 walk by its random line traffic.
struct Tuning {
    int log_d0 = 5;      // level-0 splitter density: one per 2^log_d0 ids (2^24+ nodes; level0_log_d)
    int local_bits = 12; // level-0 hop cap 2^local_bits (also the packed local field)
    int log_dk = 2;      // reduced-level splitter density
    int hop_cap_k = 1024;
    int walk_mean_max = 64; // AUTO: mean list length at or below this -> head walks only
    // AUTO: at most this many nodes -> Wyllie (heads as given).  Measured on
    // one list (MI355X, us per ranking, HJ with the densities above / Wyllie):
    // 1024 nodes 28 / 11, 4096: 42 / 19, 8192: 41 / 28, 16384: 45 / 50,
    // 65536: 54 / 74, 2^20: 128 / 404.
    int wyllie_max = 8192;
    uint32_t salt = 0x2545F491u;
};
Tuning g_tune;

__device__ __forceinline__ bool is_split(uint32_t x, int log_d, uint32_t salt)
{
    uint32_t mask = (1u << log_d) - 1u;
    return (x & mask) == (fmix32((x >> log_d) ^ salt) & mask);
}

__device__ __forceinline__ int block_splitter(int t, int log_d, uint32_t salt)
{
    return (t << log_d) + (int)(fmix32((uint32_t)t ^ salt) & ((1u << log_d) - 1u));
}

template <class T> __global__ void fill_kernel(T* p, long long n, T v)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <class T> int fill(T* p, long long n, T v, hipStream_t s)
{
    if (n <= 0) return PSI_OK;
    fill_kernel<T><<<nblocks(n, 256), 256, 0, s>>>(p, n, v);
    PSI_LAUNCH_CHECK();
    return PSI_OK;
}

// Several fills of int or int2 arrays in one launch (each launch costs a few
// microseconds, more than filling a level of the reduced list).
constexpr int kFillSegs = 16;
struct FillSet {
    void* p[kFillSegs];
    long long start[kFillSegs + 1];
    long long v[kFillSegs];       // int2 (x, y) as y << 32 | x for 8-byte segments
    int wide[kFillSegs];          // 1: 8-byte elements
    int k;
};

__global__ void __launch_bounds__(256) fill_set_kernel(FillSet fs)
{
    const long long total = fs.start[fs.k];
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        int j = 0;
        while (i >= fs.start[j + 1]) j++;
        const long long e = i - fs.start[j];
        if (fs.wide[j]) ((long long*)fs.p[j])[e] = fs.v[j];
        else ((int*)fs.p[j])[e] = (int)fs.v[j];
    }
}

struct Fills {
    FillSet fs{};
    hipStream_t s;
    explicit Fills(hipStream_t st) : s(st) { fs.start[0] = 0; }
    int seg(void* p, long long n, long long v, int wide)
    {
        if (n <= 0) return PSI_OK;
        if (fs.k == kFillSegs)
            if (int rc = launch()) return rc;
        fs.p[fs.k] = p;
        fs.v[fs.k] = v;
        fs.wide[fs.k] = wide;
        fs.start[fs.k + 1] = fs.start[fs.k] + n;
        fs.k++;
        return PSI_OK;
    }
    int add(int* p, long long n, int v) { return seg(p, n, v, 0); }
    int add(int2* p, long long n, int x, int y)
    {
        return seg(p, n, (long long)(((unsigned long long)(uint32_t)y << 32) | (uint32_t)x), 1);
    }
    int launch()
    {
        if (fs.k == 0) return PSI_OK;
        const long long total = fs.start[fs.k];
        const unsigned grid = (unsigned)std::min<long long>(nblocks(total, 256), 4096);
        fill_set_kernel<<<grid, 256, 0, s>>>(fs);
        PSI_LAUNCH_CHECK();
        fs.k = 0;
        return PSI_OK;
    }
};

// ctl[0]: overflow sublists allocated; ctl[1]: error bits (1 bad index, 2 overflow exhausted)

// Head walks alone (no splitters): one lane per head walks its whole list and
// writes each node's rank, its hop count from the head, at once; no reduced
// list and no final pass.  A walk still going after N hops went round a
// cycle reachable from a head (error bit 2); a successor or head out of range
// sets bit 1.
__global__ void __launch_bounds__(256) walk_heads(const int* __restrict__ succ, int N,
                                                  const int* __restrict__ heads, int H,
                                                  int* __restrict__ rank, int* __restrict__ ctl)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= H) return;
    int cur = heads[k];
    if ((unsigned)cur >= (unsigned)N) {
        atomicOr(&ctl[1], 1);
        return;
    }
    for (int j = 0;;) {
        rank[cur] = j;
        int nxt = succ[cur];
        if (nxt < 0 || nxt >= N) {
            if (nxt != -1) atomicOr(&ctl[1], 1);
            return;
        }
        if (++j >= N) {
            atomicOr(&ctl[1], 2);
            return;
        }
        cur = nxt;
    }
}

// Level-0 walk with unit weights.  One lane per sublist start; writes
// packed[x] = (sublist << LB) | position-in-sublist for every node it visits,
// and each sublist's (next sublist, length) to rnw.
__global__ void __launch_bounds__(256) walk0_unit(const int* __restrict__ succ, int N,
                                                  const int* __restrict__ heads, int H, int B,
                                                  int log_d, uint32_t salt, int LB,
                                                  int* __restrict__ packed, int2* __restrict__ rnw,
                                                  int* __restrict__ ctl, int over_base, int over_cap)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B + H) return;
    const int hc = 1 << LB;
    int rid = t, cur;
    if (t < B) {
        cur = block_splitter(t, log_d, salt);
        if (cur >= N) return;                      // slot stays unused (-2)
    } else {
        cur = heads[t - B];
        if ((unsigned)cur >= (unsigned)N) {
            atomicOr(&ctl[1], 1);
            return;
        }
        if (B > 0 && is_split((uint32_t)cur, log_d, salt)) {
            rnw[rid] = make_int2(cur >> log_d, 0);
            return;
        }
    }
    int local = 0;
    for (;;) {
        packed[cur] = (rid << LB) | local;
        local++;
        int nxt = succ[cur];
        if (nxt < 0 || nxt >= N) {
            if (nxt != -1) atomicOr(&ctl[1], 1);
            rnw[rid] = make_int2(-1, local);
            return;
        }
        if (B > 0 && is_split((uint32_t)nxt, log_d, salt)) {
            rnw[rid] = make_int2(nxt >> log_d, local);
            return;
        }
        if (local == hc) {
            int o = atomicAdd(&ctl[0], 1);
            if (o >= over_cap) {
                atomicOr(&ctl[1], 2);
                rnw[rid] = make_int2(-1, local);
                return;
            }
            int nr = over_base + o;
            rnw[rid] = make_int2(nr, local);
            rid = nr;
            local = 0;
        }
        cur = nxt;
    }
}

// Weighted walk of a reduced level.  Heads are the slots [hb, hb+H).  A node's
// (next, weight) is one 8-byte record and its (sublist, offset) in the next
// level another, so each hop is one random load and one random store.
__global__ void __launch_bounds__(256) walk_w(const int2* __restrict__ nw_in, int R, int hb, int H,
                                              int B, int log_d, uint32_t salt, int hc,
                                              int2* __restrict__ rl_out, int2* __restrict__ nw_out,
                                              int* __restrict__ ctl, int over_base, int over_cap)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B + H) return;
    int rid = t, cur;
    if (t < B) {
        cur = block_splitter(t, log_d, salt);
        if (cur >= R) return;
    } else {
        cur = hb + (t - B);
        if (B > 0 && is_split((uint32_t)cur, log_d, salt)) {
            nw_out[rid] = make_int2(cur >> log_d, 0);
            return;
        }
    }
    int acc = 0, hops = 0;
    for (;;) {
        const int2 c = nw_in[cur];
        rl_out[cur] = make_int2(rid, acc);
        acc += c.y;
        hops++;
        const int nxt = c.x;
        if (nxt < 0 || nxt >= R) {
            nw_out[rid] = make_int2(-1, acc);
            return;
        }
        if (B > 0 && is_split((uint32_t)nxt, log_d, salt)) {
            nw_out[rid] = make_int2(nxt >> log_d, acc);
            return;
        }
        if (hops == hc) {
            int o = atomicAdd(&ctl[0], 1);
            if (o >= over_cap) {
                atomicOr(&ctl[1], 2);
                nw_out[rid] = make_int2(-1, acc);
                return;
            }
            int nr = over_base + o;
            nw_out[rid] = make_int2(nr, acc);
            rid = nr;
            acc = 0;
            hops = 0;
        }
        cur = nxt;
    }
}

// Top level: exclusive weighted prefix from the head slot of each list, by
// pointer jumping along predecessors inside one workgroup's LDS.
// pred state: >= 0 predecessor still to absorb, -1 resolved, -3 unreachable.
constexpr int kTopCap = 16384;
__global__ void __launch_bounds__(kTopThreads) top_lds(int R, const int2* __restrict__ nw,
                                                       int hb, int H, int* __restrict__ rok,
                                                       int* __restrict__ roff)
{
    __shared__ int P[kTopCap];
    __shared__ int S[kTopCap];
    constexpr int per = kTopCap / kTopThreads;
    const int tid = threadIdx.x;
    int2 v[per];
#pragma unroll
    for (int k = 0; k < per; k++) {                // a lane's loads in flight together
        const int i = tid + k * kTopThreads;
        if (i < R) v[k] = nw[i];
    }
    for (int i = tid; i < R; i += kTopThreads) {
        P[i] = -3;
        S[i] = 0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < per; k++) {                // a node's weight starts its successor's prefix
        const int i = tid + k * kTopThreads;
        if (i < R && v[k].x >= 0 && v[k].x < R) {
            P[v[k].x] = i;
            S[v[k].x] = v[k].y;
        }
    }
    __syncthreads();
    for (int i = tid; i < R; i += kTopThreads)
        if (P[i] < 0) P[i] = (i >= hb && i < hb + H) ? -1 : -3;
    __syncthreads();
    int rounds = ceil_log2_dev(R) + 1;
    for (int r = 0; r < rounds; r++) {
        int np[per], ns[per];
#pragma unroll
        for (int k = 0; k < per; k++) {
            int i = tid + k * kTopThreads;
            np[k] = -1;
            ns[k] = 0;
            if (i < R) {
                int p = P[i], s = S[i];
                if (p >= 0) {
                    int pp = P[p];
                    s += S[p];
                    p = pp;
                }
                np[k] = p;
                ns[k] = s;
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < per; k++) {
            int i = tid + k * kTopThreads;
            if (i < R) {
                P[i] = np[k];
                S[i] = ns[k];
            }
        }
        __syncthreads();
    }
    for (int i = tid; i < R; i += kTopThreads) {
        rok[i] = P[i] == -1 ? 1 : 0;
        roff[i] = P[i] == -1 ? S[i] : -1;          // -1: unreachable
    }
}

// Top level by sublist walks inside one workgroup's LDS (for at most
// kTopHjHeads heads): splitters one per aligned block of 2^kTopLogD slots and
// the head slots start sublists; each lane walks its two sublists side by
// side, leaving in every node its sublist (as -(s + 3)) and its offset in it;
// the at most 2048 sublists are then ranked by pointer jumping and each
// node's offset is its sublist's plus its own.  Work O(R) instead of top_lds'
// O(R log R) LDS accesses.  A walk longer than R hops (a cycle a head reaches)
// is cut; its nodes after the cut read -1.
constexpr int kTopLogD = 4;
constexpr int kTopHjHeads = 1024;
constexpr int kTopSubs = (kTopCap >> kTopLogD) + kTopHjHeads;
static_assert(kTopSubs == 2 * kTopThreads, "two sublists per lane");
__global__ void __launch_bounds__(kTopThreads) top_hj(int R, const int2* __restrict__ nw, int hb,
                                                      int H, uint32_t salt, int* __restrict__ rok,
                                                      int* __restrict__ roff)
{
    __shared__ int NX[kTopCap];     // successor, then -(sublist + 3) once walked
    __shared__ int WL[kTopCap];     // weight, then offset inside the sublist
    __shared__ int P[kTopSubs];     // sublist pointer jumping: predecessor state
    __shared__ int S[kTopSubs];     // and prefix; at the end the sublist's offset
    const int tid = threadIdx.x;
    const int NB = (R + (1 << kTopLogD) - 1) >> kTopLogD;
    const int M = NB + H;
    {
        // all of a lane's loads in flight together, then into LDS
        constexpr int per = kTopCap / kTopThreads;
        int2 v[per];
#pragma unroll
        for (int k = 0; k < per; k++) {
            const int i = tid + k * kTopThreads;
            if (i < R) v[k] = nw[i];
        }
#pragma unroll
        for (int k = 0; k < per; k++) {
            const int i = tid + k * kTopThreads;
            if (i < R) {
                NX[i] = v[k].x;
                WL[i] = v[k].y;
            }
        }
    }
    for (int s = tid; s < kTopSubs; s += kTopThreads) {
        P[s] = -3;
        S[s] = 0;
    }
    __syncthreads();
    int cur[2], sn[2], acc[2], hops[2];
    bool act[2];
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int s = tid + k * kTopThreads;
        act[k] = false;
        sn[k] = -2;                                // no sublist
        acc[k] = 0;
        hops[k] = 0;
        cur[k] = 0;
        if (s < NB) {
            cur[k] = block_splitter(s, kTopLogD, salt);
            act[k] = cur[k] < R;
        } else if (s < M) {
            cur[k] = hb + (s - NB);
            if (is_split((uint32_t)cur[k], kTopLogD, salt))
                sn[k] = cur[k] >> kTopLogD;        // the head is a splitter: its sublist follows
            else
                act[k] = true;
        }
    }
    while (act[0] || act[1]) {
#pragma unroll
        for (int k = 0; k < 2; k++) {
            if (!act[k]) continue;
            const int s = tid + k * kTopThreads;
            const int c = cur[k];
            const int x = NX[c];
            const int w = WL[c];
            NX[c] = -(s + 3);
            WL[c] = acc[k];
            acc[k] += w;
            if (x < 0 || x >= R || ++hops[k] > R) {
                sn[k] = -1;
                act[k] = false;
            } else if (is_split((uint32_t)x, kTopLogD, salt)) {
                sn[k] = x >> kTopLogD;
                act[k] = false;
            } else {
                cur[k] = x;
            }
        }
    }
    __syncthreads();
    // predecessors of the sublists; a sublist's prefix starts with the weight
    // of the one before it
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int s = tid + k * kTopThreads;
        if (s < M && sn[k] >= 0 && sn[k] < NB) {
            P[sn[k]] = s;
            S[sn[k]] = acc[k];
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int s = tid + k * kTopThreads;
        if (s >= NB && s < M && P[s] == -3) P[s] = -1;   // a head: offset 0
    }
    __syncthreads();
    const int rounds = ceil_log2_dev(M) + 1;
    for (int r = 0; r < rounds; r++) {
        int np[2], ns[2];
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const int s = tid + k * kTopThreads;
            np[k] = -3;
            ns[k] = 0;
            if (s < M) {
                int p = P[s], v = S[s];
                if (p >= 0) {
                    v += S[p];
                    p = P[p];
                }
                np[k] = p;
                ns[k] = v;
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const int s = tid + k * kTopThreads;
            if (s < M) {
                P[s] = np[k];
                S[s] = ns[k];
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int s = tid + k * kTopThreads;
        if (s < M && P[s] != -1) S[s] = -1;       // not reached from a head
    }
    __syncthreads();
    for (int i = tid; i < R; i += kTopThreads) {
        const int c = NX[i];
        const int o = c <= -3 ? S[-c - 3] : -1;
        rok[i] = o >= 0 ? 1 : 0;
        roff[i] = o >= 0 ? o + WL[i] : -1;
    }
}

// Level-0 Wyllie for a list of at most kTopCap nodes inside one workgroup's
// LDS: predecessors and prefixes never leave LDS, ceil(log2 N) + 1 rounds,
// one launch (the gpu mode's pointer jumping, gpu_parity.cu:79-106, on a
// list).  use_heads: the heads[] nodes start the lists and any other node
// without a predecessor is unreachable (-1), as for the walks; otherwise
// every node without a predecessor is a head.  ctl[1] |= 1 for an index out
// of range, 2 for a node with two predecessors (not a forest of lists).
__global__ void __launch_bounds__(kTopThreads) wy_lds0(const int* __restrict__ succ, int N,
                                                       const int* __restrict__ heads, int H,
                                                       int use_heads, int* __restrict__ rank,
                                                       int* __restrict__ ctl, int ctl_zero)
{
    __shared__ int P[kTopCap];
    __shared__ int S[kTopCap];
    constexpr int per = kTopCap / kTopThreads;
    const int tid = threadIdx.x;
    int bits = 0;
    if (tid < ctl_zero) ctl[tid] = 0;             // this call's control words (one workgroup)
    int nxv[per];
#pragma unroll
    for (int k = 0; k < per; k++) {                // a lane's loads in flight together
        const int i = tid + k * kTopThreads;
        nxv[k] = i < N ? succ[i] : -1;
    }
    for (int i = tid; i < N; i += kTopThreads) P[i] = -3;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < per; k++) {
        const int i = tid + k * kTopThreads;
        const int nx = nxv[k];
        if (i >= N) continue;
        if (nx >= 0 && nx < N) {
            if (atomicExch(&P[nx], i) != -3) bits |= 2;
        } else if (nx != -1) {
            bits |= 1;
        }
    }
    __syncthreads();
    if (use_heads) {
        for (int k = tid; k < H; k += kTopThreads) {
            int h = heads[k];
            if (h >= 0 && h < N) {
                if (P[h] == -3) P[h] = -2;        // a given head
            } else {
                bits |= 1;
            }
        }
        __syncthreads();
    }
    for (int i = tid; i < N; i += kTopThreads) {
        int p = P[i];
        if (p >= 0) {
            S[i] = 1;
        } else {
            P[i] = (p == -2 || !use_heads) ? -1 : -3;
            S[i] = 0;
        }
    }
    __syncthreads();
    const int rounds = ceil_log2_dev(N) + 1;
    for (int r = 0; r < rounds; r++) {
        int np[per], ns[per];
#pragma unroll
        for (int k = 0; k < per; k++) {
            int i = tid + k * kTopThreads;
            np[k] = -1;
            ns[k] = 0;
            if (i < N) {
                int p = P[i], s = S[i];
                if (p >= 0) {
                    s += S[p];
                    p = P[p];
                }
                np[k] = p;
                ns[k] = s;
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < per; k++) {
            int i = tid + k * kTopThreads;
            if (i < N) {
                P[i] = np[k];
                S[i] = ns[k];
            }
        }
        __syncthreads();
    }
    for (int i = tid; i < N; i += kTopThreads) rank[i] = P[i] == -1 ? S[i] : -1;
    __threadfence_block();
    __syncthreads();                              // the zeroing above precedes the bits
    if (bits) atomicOr(&ctl[1], bits);
}

// Wyllie pointer jumping along predecessors.  PS[i] = (pred-state, prefix).
// hb < 0: every node without a predecessor is a head (level-0 Wyllie).
// ctl (level 0 only, else null): bit 2 if a node has two predecessors.
// nxt[i * stride]: the successor of node i (stride 2 reads the x of int2
// records).
__global__ void wy_pred(const int* __restrict__ nxt, int stride, int R, int* __restrict__ pred,
                        int* __restrict__ ctl)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= R) return;
    int nx = nxt[(long long)i * stride];
    if (nx >= 0 && nx < R) {
        if (ctl) {
            if (atomicExch(&pred[nx], i) != -1) atomicOr(&ctl[1], 2);
        } else {
            pred[nx] = i;
        }
    } else if (ctl && nx != -1) {
        atomicOr(&ctl[1], 1);
    }
}

// Level-0 heads given by a list: hmark[h] = 1.
__global__ void wy_mark_heads(const int* __restrict__ heads, int H, int N, int* __restrict__ hmark,
                              int* __restrict__ ctl)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= H) return;
    int h = heads[k];
    if (h >= 0 && h < N) hmark[h] = 1;
    else atomicOr(&ctl[1], 1);
}

// hmark (level 0 with given heads, else null): a node without a predecessor
// is a head only if marked.
// w[p * 2]: the weight of node p (the y of the level's int2 records).
template <bool UNIT>
__global__ void wy_init(const int* __restrict__ pred, const int* __restrict__ w, int R, int hb,
                        int H, const int* __restrict__ hmark, int2* __restrict__ ps)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= R) return;
    int p = pred[i];
    int2 v;
    if (p >= 0) {
        v = make_int2(p, UNIT ? 1 : w[(long long)p * 2]);
    } else {
        bool head = hmark ? hmark[i] != 0 : (hb < 0 ? true : (i >= hb && i < hb + H));
        v = make_int2(head ? -1 : -3, 0);
    }
    ps[i] = v;
}

// active: kWySlots counters per round (one atomic per block into slot
// blockIdx % kWySlots), then active[64 * kWySlots] = last round that ran.
constexpr int kWySlots = 64;
constexpr int kWyRounds = 64;

// Grid-stride over at most kWyGrid workgroups, so that the rounds launched
// after convergence (which exit at once) cost little to dispatch.
constexpr int kWyGrid = 2048;

__global__ void __launch_bounds__(256) wy_round(const int2* __restrict__ in, int2* __restrict__ out,
                                                int R, int* __restrict__ active, int r)
{
    __shared__ unsigned int cnt;
    // The round decision: a wave reduction over the previous round's slot
    // words (written by the previous launch, the same for every wave).
    if (r > 0 && !__any(active[(r - 1) * kWySlots + (threadIdx.x & (kWySlots - 1))] != 0))
        return;                                   // converged: nothing left to jump
    if (threadIdx.x == 0) {
        cnt = 0;
        if (blockIdx.x == 0) active[kWyRounds * kWySlots] = r;
    }
    __syncthreads();
    bool act = false;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < R;
         i += (long long)gridDim.x * blockDim.x) {
        int2 v = in[i];
        if (v.x >= 0) {
            int2 u = in[v.x];
            v.y += u.y;
            v.x = u.x;
            act |= v.x >= 0;
        }
        out[i] = v;
    }
    unsigned long long m = __ballot(act);
    if ((threadIdx.x & (kWave - 1)) == 0 && m) atomicAdd(&cnt, 1u);
    __syncthreads();
    if (threadIdx.x == 0 && cnt) atomicAdd(&active[r * kWySlots + (blockIdx.x & (kWySlots - 1))], 1);
}

// Reads the result from the buffer the last round that ran wrote.
__global__ void wy_finish(const int2* __restrict__ a, const int2* __restrict__ b, int R,
                          const int* __restrict__ active, int* __restrict__ rok,
                          int* __restrict__ roff)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= R) return;
    int last = active[kWyRounds * kWySlots];      // round r writes buffer (r+1)&1
    const int2* src = ((last + 1) & 1) ? b : a;
    int2 v = src[i];
    rok[i] = v.x == -1 ? 1 : 0;
    roff[i] = v.x == -1 ? v.y : -1;
}

__global__ void prop_kernel(int R, const int2* __restrict__ rl, const int* __restrict__ roff_up,
                            int* __restrict__ rok, int* __restrict__ roff)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const int2 v = rl[r];
    int o = v.x < 0 ? -1 : roff_up[v.x];
    rok[r] = o >= 0;
    roff[r] = o >= 0 ? o + v.y : -1;
}

__global__ void final0(int N, int LB, int R1, int* __restrict__ rank, const int* __restrict__ rok1,
                       const int* __restrict__ roff1)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= N) return;
    int p = rank[x];
    if (p < 0) return;                             // never visited: stays -1
    int id = p >> LB;
    int o = id < R1 ? roff1[id] : -1;             // one gather: -1 marks unreachable
    rank[x] = o >= 0 ? o + (p & ((1 << LB) - 1)) : -1;
}

__device__ __forceinline__ int final_one(int p, int LB, int R1, const int* __restrict__ roff1)
{
    if (p < 0) return p;
    int id = p >> LB;
    int o = id < R1 ? roff1[id] : -1;
    return o >= 0 ? o + (p & ((1 << LB) - 1)) : -1;
}

// final0 four nodes per lane (16-byte loads and stores of rank; the roff1
// gathers stay in L2): rank[4i .. 4i+3] for i < N4.
__global__ void __launch_bounds__(256) final0_v4(int N4, int LB, int R1, int4* __restrict__ rank4,
                                                 const int* __restrict__ roff1)
{
    typedef int v4i __attribute__((ext_vector_type(4)));
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N4) return;
    // rank streams through once: non-temporal, so that it does not push the
    // gathered offset table out of L2
    v4i* q = (v4i*)rank4;
    v4i p = __builtin_nontemporal_load(q + i);
    v4i o = {final_one(p.x, LB, R1, roff1), final_one(p.y, LB, R1, roff1),
             final_one(p.z, LB, R1, roff1), final_one(p.w, LB, R1, roff1)};
    __builtin_nontemporal_store(o, q + i);
}

// ---------------------------------------------------------------------------
// Host orchestration.

static size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

// Device workspace carved front to back.  A take beyond cap (a planning
// mistake) sets `over` and returns the start of the buffer, so no kernel
// touches memory outside it; the caller then reports an error.
struct Arena {
    void* buf = nullptr;
    size_t cap = 0;
    size_t used = 0;
    bool over = false;
    template <class T> T* take(long long n)
    {
        size_t bytes = al256((size_t)std::max(n, 1LL) * sizeof(T));
        if (used + bytes > cap) {
            over = true;
            return (T*)buf;
        }
        T* p = (T*)((char*)buf + used);
        used += bytes;
        return p;
    }
};

// How a reduced level is ranked.
enum LevelKind { kTop, kWyllie, kWalk };

struct Level {
    int R = 0, hb = 0, H = 0;
    int2* nw = nullptr;                            // this level's nodes: (next, weight)
    int2* rl = nullptr;                            // (sublist, offset) in the next level, if walked
    int *rok = nullptr, *roff = nullptr;           // final offsets of this level's nodes
    LevelKind kind = kTop;
    int B = 0, log_d = 0, hc = 0, over = 0;        // kWalk: splitters, hop cap, overflow slots
    int* pred = nullptr;                           // kWyllie: predecessors and ping-pong pairs
    int2 *ps0 = nullptr, *ps1 = nullptr;
};

struct Plan {
    int N, H, LB, B0, over0;
    std::vector<Level> lv;                         // lv[0] is reduced level 1
};

static size_t level_bytes(long long R)
{
    const size_t r = (size_t)std::max(R, 1LL);
    return 2 * al256(r * 8) + 2 * al256(r * 4);
}

Arena g_arena;           // psi_list_rank's workspace, grown with hipMalloc
int* g_ctl = nullptr;    // its control words (2 ints per level + 64 Wyllie round counters)
int* g_last_ctl = nullptr;   // control words of the last call (psi_list_rank_status)
constexpr int kMaxLevels = 8;
constexpr int kCtlInts = kMaxLevels * 2 + 64 * 64 + 1;
constexpr size_t kCtlBytes = (kCtlInts * sizeof(int) + 255) & ~(size_t)255;

static int ensure_ctl()
{
    if (g_ctl) return PSI_OK;
    PSI_HIP(hipMalloc(&g_ctl, kCtlInts * sizeof(int)));
    return PSI_OK;
}

// Workspace the levels below a level of R nodes and H heads take (the level's
// own arrays excluded), following the same choices plan_levels makes.
static size_t rank_level_bytes(long long R, long long H, int k)
{
    const Tuning& tu = g_tune;
    if (R <= kTopCap) return 0;
    if (H * 4 >= R || k + 1 >= kMaxLevels) return al256((size_t)R * 4) + 2 * al256((size_t)R * 8);
    long long B = (R + (1LL << tu.log_dk) - 1) >> tu.log_dk;
    long long R2 = B + H + R / tu.hop_cap_k + 1;
    return level_bytes(R2) + rank_level_bytes(R2, H, k + 1);
}

// Chooses how each reduced level from lv[k] down is ranked, takes its arrays
// from `ar` and adds every fill they need to `f`, so that one launch clears
// all levels before the level-0 walk.
static int plan_levels(Plan& pl, int k, Arena& ar, int* ctl0, Fills& f)
{
    const Tuning& tu = g_tune;
    for (;; k++) {
        Level& L = pl.lv[k];
        if (L.R <= kTopCap) {
            L.kind = kTop;
            return PSI_OK;
        }
        if ((long long)L.H * 4 >= L.R || k + 1 >= kMaxLevels) {
            // Global Wyllie on this level (a list of it has at most R nodes).
            L.kind = kWyllie;
            L.pred = ar.take<int>(L.R);
            L.ps0 = ar.take<int2>(L.R);
            L.ps1 = ar.take<int2>(L.R);
            if (ar.over) return fail(PSI_ERR_INVALID, "list ranking: workspace plan exceeded");
            if (int rc = f.add(L.pred, L.R, -1)) return rc;
            return f.add(ctl0 + kMaxLevels * 2, kWyRounds * kWySlots + 1, 0);
        }
        // Walk this level into a new reduced level.
        L.kind = kWalk;
        L.log_d = tu.log_dk;
        L.B = (int)((L.R + (1LL << L.log_d) - 1) >> L.log_d);
        L.hc = tu.hop_cap_k;
        L.over = L.R / L.hc + 1;
        Level nx;
        nx.R = L.B + L.H + L.over;
        nx.hb = L.B;
        nx.H = L.H;
        nx.nw = ar.take<int2>(nx.R);
        nx.rl = ar.take<int2>(nx.R);
        nx.rok = ar.take<int>(nx.R);
        nx.roff = ar.take<int>(nx.R);
        if (ar.over) return fail(PSI_ERR_INVALID, "list ranking: workspace plan exceeded");
        if (int rc = f.add(nx.nw, nx.R, -2, 0)) return rc;   // unused slots: no successor
        if (int rc = f.add(L.rl, L.R, -1, 0)) return rc;     // not reached by this level's walks
        pl.lv.push_back(nx);
    }
}

// Ranks the reduced level lv[k] and the levels below it as planned (offsets
// into lv[k].rok/roff); control words from ctl0, zeroed before the walks.
static int rank_level(Plan& pl, int k, int* ctl0, hipStream_t s)
{
    Level& L = pl.lv[k];
    if (L.kind == kTop) {
        if (L.H <= kTopHjHeads)
            top_hj<<<1, kTopThreads, 0, s>>>(L.R, L.nw, L.hb, L.H, g_tune.salt, L.rok, L.roff);
        else
            top_lds<<<1, kTopThreads, 0, s>>>(L.R, L.nw, L.hb, L.H, L.rok, L.roff);
        PSI_LAUNCH_CHECK();
        return PSI_OK;
    }
    if (L.kind == kWyllie) {
        int* active = ctl0 + kMaxLevels * 2;
        int rounds = ceil_log2(L.R) + 1;
        wy_pred<<<nblocks(L.R, 256), 256, 0, s>>>((const int*)L.nw, 2, L.R, L.pred, nullptr);
        wy_init<false><<<nblocks(L.R, 256), 256, 0, s>>>(L.pred, (const int*)L.nw + 1, L.R, L.hb,
                                                         L.H, nullptr, L.ps0);
        const unsigned wg = (unsigned)std::min<long long>(nblocks(L.R, 256), kWyGrid);
        for (int r = 0; r < rounds; r++)
            wy_round<<<wg, 256, 0, s>>>((r & 1) ? L.ps1 : L.ps0, (r & 1) ? L.ps0 : L.ps1, L.R,
                                        active, r);
        wy_finish<<<nblocks(L.R, 256), 256, 0, s>>>(L.ps0, L.ps1, L.R, active, L.rok, L.roff);
        PSI_LAUNCH_CHECK();
        return PSI_OK;
    }
    int* ctl = ctl0 + 2 * (k + 1);
    const Level& N2 = pl.lv[k + 1];
    if (L.B + L.H > 0)
        walk_w<<<nblocks(L.B + L.H, 256), 256, 0, s>>>(L.nw, L.R, L.hb, L.H, L.B, L.log_d,
                                                       g_tune.salt, L.hc, L.rl, N2.nw, ctl,
                                                       L.B + L.H, L.over);
    PSI_LAUNCH_CHECK();
    if (int rc = rank_level(pl, k + 1, ctl0, s)) return rc;
    const Level& L2 = pl.lv[k];
    const Level& N3 = pl.lv[k + 1];
    prop_kernel<<<nblocks(L2.R, 256), 256, 0, s>>>(L2.R, L2.rl, N3.roff, L2.rok, L2.roff);
    PSI_LAUNCH_CHECK();
    return PSI_OK;
}

// Workspace of each method for N nodes and H heads (what the runs below take).
static size_t wyllie0_bytes(long long N, bool use_heads)
{
    if (N <= kTopCap) return 0;
    return (use_heads ? 3 : 2) * al256((size_t)N * 4) + 2 * al256((size_t)N * 8);
}

struct Hj0Shape {
    int log_d = 0, B0 = 0, LB = 0, over0 = 0;
    long long R1 = 0;
};

// Level-0 splitter density for N nodes: the tuned one from 2^24 nodes on,
// one step denser below 2^24 and two below 9 * 2^20 (shorter lists are bound by
// their longest sublist rather than by the walk's line traffic).  Measured on
// random lists, us per ranking at level-0 densities 2^-3 / 2^-4 / 2^-5:
// 65,536 nodes 54 / 68 / 128, 2^20: 129 / 133 / 182, 2^22: 243 / 270 / 305,
// 8,388,610: 406 / 451 / 451, 10 * 2^20: 508 / 483 / 507, 12 * 2^20:
// 586 / 577 / 607, 14 * 2^20: 690 / 640 / 689, 2^24: 782-792 / 717-741 / 727-728.
static int level0_log_d(int N)
{
    const int steps = N >= (1 << 24) ? 0 : N >= 9 * (1 << 20) ? 1 : 2;
    return std::max(1, g_tune.log_d0 - steps);
}

static int hj0_shape(int N, int H, Hj0Shape* sh)
{
    const Tuning& tu = g_tune;
    sh->log_d = level0_log_d(N);
    sh->B0 = (int)((N + (1LL << sh->log_d) - 1) >> sh->log_d);
    int LB = tu.local_bits;
    for (;;) {
        sh->over0 = (int)(N >> LB) + 1;
        sh->R1 = (long long)sh->B0 + H + sh->over0;
        if (sh->R1 < (1LL << (31 - LB))) break;
        if (--LB < 4) return fail(PSI_ERR_INVALID, "list ranking: too many heads/sublists to pack");
    }
    sh->LB = LB;
    return PSI_OK;
}

static size_t hj0_bytes(const Hj0Shape& sh, int H)
{
    return level_bytes(sh.R1) + rank_level_bytes(sh.R1, H, 0);
}

// Sizes `ar` for `need` bytes: grows it (hipMalloc) when `grow`, else fails
// if its fixed capacity is smaller.
static int arena_prepare(Arena& ar, size_t need, bool grow)
{
    if (grow) {
        if (int rc = ensure(&ar.buf, &ar.cap, need)) return rc;
    } else if (need > ar.cap) {
        char msg[160];
        snprintf(msg, sizeof msg, "list ranking: needs %zu bytes of workspace, the fixed "
                 "workspace holds %zu", need, ar.cap);
        return fail(PSI_ERR_INVALID, msg);
    }
    ar.used = 0;
    ar.over = false;
    return PSI_OK;
}

// Level-0 Wyllie.  use_heads: heads[] start the lists (AUTO's choice for
// short lists, same results as the walks); else every node without a
// predecessor is a head (PSI_LR_WYLLIE).  Lists of at most kTopCap nodes are
// ranked in one workgroup's LDS (one launch, no workspace).
static int wyllie0(const int* succ, int N, const int* heads, int H, bool use_heads, int* rank,
                   Arena& ar, int* ctl0, bool grow, hipStream_t s)
{
    if (N <= kTopCap) {
        wy_lds0<<<1, kTopThreads, 0, s>>>(succ, N, heads, H, use_heads ? 1 : 0, rank, ctl0,
                                          kMaxLevels * 2);
        PSI_LAUNCH_CHECK();
        return PSI_OK;
    }
    if (int rc = arena_prepare(ar, wyllie0_bytes(N, use_heads), grow)) return rc;
    int* pred = ar.take<int>(N);
    int* rok = ar.take<int>(N);
    int2* ps0 = ar.take<int2>(N);
    int2* ps1 = ar.take<int2>(N);
    int* hmark = use_heads ? ar.take<int>(N) : nullptr;
    if (ar.over) return fail(PSI_ERR_INVALID, "list ranking: workspace plan exceeded");
    int* active = ctl0 + kMaxLevels * 2;
    int rounds = ceil_log2(N) + 1;
    if (rounds > 64) rounds = 64;
    {
        Fills f(s);
        f.add(ctl0, kMaxLevels * 2, 0);
        f.add(pred, N, -1);
        f.add(active, kWyRounds * kWySlots + 1, 0);
        if (use_heads) f.add(hmark, N, 0);
        if (int rc = f.launch()) return rc;
    }
    if (use_heads && H > 0) wy_mark_heads<<<nblocks(H, 256), 256, 0, s>>>(heads, H, N, hmark, ctl0);
    wy_pred<<<nblocks(N, 256), 256, 0, s>>>(succ, 1, N, pred, ctl0);
    wy_init<true><<<nblocks(N, 256), 256, 0, s>>>(pred, nullptr, N, -1, 0, hmark, ps0);
    const unsigned wg = (unsigned)std::min<long long>(nblocks(N, 256), kWyGrid);
    for (int r = 0; r < rounds; r++)
        wy_round<<<wg, 256, 0, s>>>((r & 1) ? ps1 : ps0, (r & 1) ? ps0 : ps1, N, active, r);
    wy_finish<<<nblocks(N, 256), 256, 0, s>>>(ps0, ps1, N, active, rok, rank);
    PSI_LAUNCH_CHECK();
    return PSI_OK;
}

// Head walks alone (PSI_LR_WALK): ranks written during the walk, no workspace.
static int walk_direct(const int* succ, int N, const int* heads, int H, int* rank, bool checked,
                       int* ctl0, hipStream_t s)
{
    Fills f(s);
    if (int rc = f.add(ctl0, kMaxLevels * 2, 0)) return rc;
    if (checked)
        if (int rc = f.add(rank, N, -1)) return rc;
    if (int rc = f.launch()) return rc;
    if (H > 0) {
        walk_heads<<<nblocks(H, 256), 256, 0, s>>>(succ, N, heads, H, rank, ctl0);
        PSI_LAUNCH_CHECK();
    }
    return PSI_OK;
}

static int hj0(const int* succ, int N, const int* heads, int H, int* rank, bool checked,
               Arena& ar, int* ctl0, bool grow, hipStream_t s)
{
    const Tuning& tu = g_tune;
    Hj0Shape sh;
    if (int rc = hj0_shape(N, H, &sh)) return rc;
    if (int rc = arena_prepare(ar, hj0_bytes(sh, H), grow)) return rc;
    Plan pl;
    pl.N = N;
    pl.H = H;
    const int log_d = sh.log_d;
    pl.B0 = sh.B0;
    pl.over0 = sh.over0;
    pl.LB = sh.LB;
    const int LB = sh.LB;
    const long long R1 = sh.R1;
    Level L1;
    L1.R = (int)R1;
    L1.hb = pl.B0;
    L1.H = H;
    L1.nw = ar.take<int2>(R1);
    L1.rl = ar.take<int2>(R1);
    L1.rok = ar.take<int>(R1);
    L1.roff = ar.take<int>(R1);
    if (ar.over) return fail(PSI_ERR_INVALID, "list ranking: workspace plan exceeded");
    pl.lv.push_back(L1);
    {
        // one launch clears the control words, every level's arrays and (checked) rank
        Fills f(s);
        if (int rc = f.add(ctl0, kMaxLevels * 2, 0)) return rc;
        if (int rc = f.add(pl.lv[0].nw, R1, -2, 0)) return rc;
        if (checked)
            if (int rc = f.add(rank, N, -1)) return rc;
        if (int rc = plan_levels(pl, 0, ar, ctl0, f)) return rc;
        if (int rc = f.launch()) return rc;
    }
    if (pl.B0 + H > 0) {
        walk0_unit<<<nblocks((long long)pl.B0 + H, 256), 256, 0, s>>>(
            succ, N, heads, H, pl.B0, log_d, tu.salt, LB, rank, pl.lv[0].nw, ctl0, pl.B0 + H,
            pl.over0);
        PSI_LAUNCH_CHECK();
    }
    if (int rc = rank_level(pl, 0, ctl0, s)) return rc;
    const int* roff1 = pl.lv[0].roff;
    if (((uintptr_t)rank & 15) == 0 && N >= 1024) {
        const int N4 = N / 4;
        final0_v4<<<nblocks(N4, 256), 256, 0, s>>>(N4, LB, (int)R1, (int4*)rank, roff1);
        if (N4 * 4 < N)
            final0<<<1, 256, 0, s>>>(N - N4 * 4, LB, (int)R1, rank + N4 * 4, pl.lv[0].rok, roff1);
    } else {
        final0<<<nblocks(N, 256), 256, 0, s>>>(N, LB, (int)R1, rank, pl.lv[0].rok, roff1);
    }
    PSI_LAUNCH_CHECK();
    return PSI_OK;
}

// One list ranking on stream s with workspace `ar` and control words ctl0
// (kCtlInts ints); the method dispatch of psi_list_rank.
static int run(const int* d_succ, int n, const int* d_heads, int n_heads, int* d_rank, int method,
               Arena& ar, int* ctl0, bool grow, hipStream_t s)
{
    if (n < 0 || n_heads < 0) return fail(PSI_ERR_INVALID, "psi_list_rank: negative size");
    if (n == 0) return PSI_OK;
    int flags = method & ~0xff;
    method &= 0xff;
    bool checked = (flags & PSI_LR_CHECKED) != 0;
    g_last_ctl = ctl0;        // each method zeroes ctl0[0, 2 * kMaxLevels) in its first launch
    if (method == PSI_LR_AUTO) {
        // by list length: Wyllie for short lists (few nodes: its log n rounds
        // beat the walks' launches and their longest sublist), head walks for
        // many short lists, Helman-JaJa otherwise
        if (n <= g_tune.wyllie_max)
            return wyllie0(d_succ, n, d_heads, n_heads, true, d_rank, ar, ctl0, grow, s);
        long long mean = n_heads ? (long long)n / n_heads : (long long)n;
        method = mean <= g_tune.walk_mean_max ? PSI_LR_WALK : PSI_LR_HJ;
    }
    switch (method) {
    case PSI_LR_HJ: return hj0(d_succ, n, d_heads, n_heads, d_rank, checked, ar, ctl0, grow, s);
    case PSI_LR_WALK: return walk_direct(d_succ, n, d_heads, n_heads, d_rank, checked, ctl0, s);
    case PSI_LR_WYLLIE: return wyllie0(d_succ, n, d_heads, n_heads, false, d_rank, ar, ctl0, grow, s);
    default: return fail(PSI_ERR_INVALID, "psi_list_rank: unknown method");
    }
}

} // namespace lr
} // namespace psi

using namespace psi;

// psi_list_rank's calls take turns (one workspace, on the library stream).
static std::mutex g_lr_mu;

extern "C" int psi_list_rank(const int* d_succ, int n, const int* d_heads, int n_heads, int* d_rank,
                             int method)
{
    if (n < 0 || n_heads < 0) return fail(PSI_ERR_INVALID, "psi_list_rank: negative size");
    if (n == 0) return PSI_OK;
    std::lock_guard<std::mutex> lk(g_lr_mu);
    if (int rc = lr::ensure_ctl()) return rc;
    return lr::run(d_succ, n, d_heads, n_heads, d_rank, method, lr::g_arena, lr::g_ctl, true,
                   stream());
}

// ---- the same list ranking with a fixed, statically allocated workspace ----
// The control words and the arena of psi_list_rank live here instead of in
// hipMalloc'd memory: one array per GPU, held until the process ends.
constexpr size_t kScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices];   // device address of `scratch` on each GPU
}

namespace psi {
// Device address of the current GPU's scratch array (resolved once per
// device, with that device current).  The array's name is only a host-side
// stand-in; kernels and fills get this address.
int lr_scratch_address(void** out)
{
    int dev = 0;
    PSI_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(PSI_ERR_INVALID, "list ranking: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
        g_scratch_dev[dev] = p;
    }
    if (out) *out = g_scratch_dev[dev];
    return PSI_OK;
}
} // namespace psi

extern "C" int psi_list_rank_noalloc(const int* d_succ, int n, const int* d_heads, int n_heads,
                                     int* d_rank, int method)
{
    if (n < 0 || n_heads < 0) return fail(PSI_ERR_INVALID, "psi_list_rank_noalloc: negative size");
    if (n == 0) return PSI_OK;
    void* base = nullptr;
    if (int rc = lr_scratch_address(&base)) return rc;
    lr::Arena ar;
    ar.buf = (char*)base + lr::kCtlBytes;
    ar.cap = kScratchBytes - lr::kCtlBytes;
    int rc = lr::run(d_succ, n, d_heads, n_heads, d_rank, method, ar, (int*)base, false, stream());
    if (rc == PSI_ERR_INVALID && strstr(psi_last_error(), "fixed workspace")) {
        std::string msg = std::string("psi_list_rank_noalloc: ") + psi_last_error() +
                          " (its maximum; use psi_list_rank for this input)";
        return fail(PSI_ERR_INVALID, msg);
    }
    return rc;
}

// Reads the device error bits of the last list ranking (either entry point).
extern "C" int psi_list_rank_status(int* error_bits, int* overflow_used)
{
    int* ctl = lr::g_last_ctl;
    if (!ctl) {
        if (int rc = lr::ensure_ctl()) return rc;
        ctl = lr::g_ctl;
        PSI_HIP(hipMemsetAsync(ctl, 0, lr::kMaxLevels * 2 * sizeof(int), stream()));
    }
    int h[2];
    PSI_HIP(hipMemcpyAsync(h, ctl, sizeof h, hipMemcpyDeviceToHost, stream()));
    PSI_HIP(hipStreamSynchronize(stream()));
    if (overflow_used) *overflow_used = h[0];
    int bits = 0;
    std::vector<int> all(lr::kMaxLevels * 2);
    PSI_HIP(hipMemcpyAsync(all.data(), ctl, all.size() * sizeof(int), hipMemcpyDeviceToHost, stream()));
    PSI_HIP(hipStreamSynchronize(stream()));
    for (int k = 0; k < lr::kMaxLevels; k++) bits |= all[2 * k + 1];
    if (error_bits) *error_bits = bits;
    return PSI_OK;
}

extern "C" int psi_list_rank_host(const int* succ, int n, const int* heads, int n_heads, int* rank,
                                  int method)
{
    if (n <= 0) return n == 0 ? PSI_OK : fail(PSI_ERR_INVALID, "psi_list_rank_host: negative size");
    hipStream_t s = stream();
    int *ds = nullptr, *dh = nullptr, *dr = nullptr;
    PSI_HIP(hipMalloc(&ds, (size_t)n * 4));
    PSI_HIP(hipMalloc(&dh, (size_t)std::max(n_heads, 1) * 4));
    PSI_HIP(hipMalloc(&dr, (size_t)n * 4));
    PSI_HIP(hipMemcpyAsync(ds, succ, (size_t)n * 4, hipMemcpyHostToDevice, s));
    if (n_heads) PSI_HIP(hipMemcpyAsync(dh, heads, (size_t)n_heads * 4, hipMemcpyHostToDevice, s));
    int rc = psi_list_rank(ds, n, dh, n_heads, dr, method | PSI_LR_CHECKED);
    if (rc == PSI_OK) {
        PSI_HIP(hipMemcpyAsync(rank, dr, (size_t)n * 4, hipMemcpyDeviceToHost, s));
        PSI_HIP(hipStreamSynchronize(s));
        int bits = 0;
        rc = psi_list_rank_status(&bits, nullptr);
        if (rc == PSI_OK && bits) rc = fail(PSI_ERR_INVALID, bits & 1 ? "list ranking: successor or head out of range"
                                                                      : "list ranking: walk did not end (cycle reachable from a head?)");
    }
    (void)hipFree(ds);
    (void)hipFree(dh);
    (void)hipFree(dr);
    return rc;
}

extern "C" int psi_list_rank_tune(int param, int value)
{
    std::lock_guard<std::mutex> lk(g_lr_mu);     // not while psi_list_rank plans with it
    auto& t = lr::g_tune;
    switch (param) {
    case PSI_LRT_LOG_D0: if (value < 1 || value > 16) break; t.log_d0 = value; return PSI_OK;
    case PSI_LRT_LOCAL_BITS: if (value < 4 || value > 16) break; t.local_bits = value; return PSI_OK;
    case PSI_LRT_LOG_DK: if (value < 1 || value > 16) break; t.log_dk = value; return PSI_OK;
    case PSI_LRT_HOP_CAP_K: if (value < 16) break; t.hop_cap_k = value; return PSI_OK;
    case PSI_LRT_WALK_MEAN_MAX: if (value < 1) break; t.walk_mean_max = value; return PSI_OK;
    case PSI_LRT_WYLLIE_MAX: if (value < 0) break; t.wyllie_max = value; return PSI_OK;
    default: break;
    }
    return fail(PSI_ERR_INVALID, "psi_list_rank_tune: bad parameter or value");
}
