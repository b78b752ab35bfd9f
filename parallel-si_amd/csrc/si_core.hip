// Device-resident strategy improvement with exact compact valuations.
//
// Reference semantics: CPUParity (cpu_parity.cpp:54-307) driven by solve_si
// (gpumain.cpp:18-53).  Reference GPU roles: GPUParity (gpu_parity.cu:16-354):
// d_init_strat :38-56, the tree-Wyllie valuation d_init_val/d_val_step/
// d_mark_infinite :61-187, compare_valuation :191-262, d_switch_strategy
// :264-317, d_mark_solved :321-336, get_winning :340-354.
//
// Valuation.  A finite vertex's valuation is the multiset of priorities on its
// strategy path to the sink.  Instead of the reference's dense count vector
// per vertex (n * (max_pri+1) ints, gpu_parity.cu:25-29) each vertex keeps
//   par  - its strategy successor at valuation time (the sink tree),
//   dep  - the length of its path to the sink,
//   key  - (label + 1) << 30 | the largest priority on that path,
// computed by pointer jumping over the strategy graph: (next, max, count)
// records combine associatively, so ceil(log2 d) + 1 rounds resolve every
// vertex at distance d from the sink or from a solved vertex.  A round that
// resolves nothing proves that every still-jumping vertex leads into a cycle;
// from then on the rounds only have to make the jump window cover the longest
// cycle, after which the window maximum at the cycle vertex gives the cycle's
// largest priority and so the label (cpu_parity.cpp:35-52, :81-89).
// Solved vertices are terminals with their frozen label (cpu_parity.cpp:116-127).
//
// Compare (total preorder of cpu_parity.cpp:154-236).  +inf > finite > -inf.
// Two finite valuations whose path maxima differ are ordered by that maximum
// alone: the larger one appears only on one side, so it is the maxdiff
// priority.  Equal maxima (a few percent of compares on random games) take the
// exact path: walk both paths to their lowest common ancestor with par/dep,
// then compare the two private segments from the top priority down,
// (max, count) level by level.  All integer, no approximation.
//
// Control.  Every game has a GameCtl block.  One SI step = valuation, br++,
// mark_solved(1), switch_strategy(1), and for games whose player-1 switch was
// empty: mark_solved(0), iter++, switch_strategy(0), done if empty.  Kernels
// read the blocks and skip finished games, so steps can be enqueued (and
// captured in a hipGraph) without any host round trip.
#include "si_core.h"

#include <hip/hip_ext.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <thread>

namespace psi {

constexpr int kBS = 256;
constexpr int kRows = kBS / kWave;   // rows of 64 vertices in a block
constexpr int kInitRows = 2;         // rows of a block initialised together
constexpr int kPriMask = 0x3FFFFFFF;
// Jump levels 0..kLiftCap are kept as binary-lifting tables; later levels
// rotate through three more buffers (a launch may read level r while writing
// levels r + 1 and r + 2).
constexpr int kLiftCap = 12;
constexpr int kRotating = 3;
__host__ __device__ __forceinline__ int level_slot(int j)
{
    return j <= kLiftCap ? j : kLiftCap + 1 + (j - kLiftCap - 1) % kRotating;
}

__device__ __forceinline__ int pri_of(int vi) { return vi & kPriMask; }

// Jump record (8 bytes): x = the 2^j-th successor, or, once the path ends
// within 2^j vertices, a terminal code -(1 + kind + 8 * count) with kind 0
// sink, 1 solved +inf, 2 solved -inf, 3 finished game, and count = vertices on
// the path before the terminal (a record that has not ended spans exactly 2^j
// vertices, so it needs no count); y = largest priority on the span (0 when
// empty) | bit 30: some link on the span changed since the last valuation.
constexpr int kChangedBit = 1 << 30;
__host__ __device__ __forceinline__ int term_code(int kind, int count) { return -(1 + kind + 8 * count); }
__device__ __forceinline__ int term_kind(int x) { return (-x - 1) & 7; }
__device__ __forceinline__ int term_count(int x) { return (-x - 1) >> 3; }
__device__ __forceinline__ int owner_of(int vi) { return (vi >> 30) & 1; }

__device__ __forceinline__ unsigned long long hmix_i32(int i, int v)
{
    return splitmix64(((unsigned long long)(uint32_t)i << 32) | (unsigned long long)(uint32_t)v);
}

__device__ __forceinline__ unsigned long long hmix_u64(int i, unsigned long long v)
{
    return splitmix64(v ^ splitmix64((unsigned long long)(uint32_t)i + 0x51ED27ull));
}

// Adds val to dst (per lane); one atomic per wave when the wave's lanes all
// target the same address.  Must be called by every lane of the wave.
__device__ __forceinline__ void wave_add_u64(unsigned long long* dst, unsigned long long val)
{
    unsigned long long d0 = __shfl((unsigned long long)dst, 0);
    if (__all((unsigned long long)dst == d0)) {
        for (int o = 32; o > 0; o >>= 1) val += __shfl_xor(val, o);
        if ((threadIdx.x & (kWave - 1)) == 0 && dst) atomicAdd(dst, val);
    } else if (dst) {
        atomicAdd(dst, val);
    }
}

// Round summary word (written by the launch that decided round r, read by the
// next one): bit 63 halt, bit 62 the launch counted cycle-label mismatches,
// bits 32..60 stop + 1, bits 0..31 cyc.
constexpr int kSumChecked = 62;
__device__ __forceinline__ int sum_stop(unsigned long long sm) { return (int)((sm >> 32) & 0x1FFFFFFFu) - 1; }

// Phase bookkeeping for round r: stop = first round that resolved nothing, or
// after which nothing was still jumping (-1 if none yet), cyc = vertices still
// jumping after it (all lead into cycles).  The state through round r-2 comes
// from the summary written by round r-1; round r-1 itself is read from its
// slots.  No kernel reads a word that its own blocks write.
__device__ __forceinline__ void round_state(int r, const JumpCnt* jc, int* stop, int* cyc, int* halt)
{
    // summary[(r-1)&1] was written by block 0 of round r-1 (a previous kernel)
    unsigned long long sm = r >= 1 ? jc->summary[(r - 1) & 1] : 0ull;
    *halt = (int)(sm >> 63);
    int st = sum_stop(sm), cy = (int)(sm & 0xFFFFFFFFu);
    if (r >= 1 && st < 0 && !*halt) {
        unsigned long long res = 0, act = 0;
        for (int k = 0; k < kSlots; k++) {
            unsigned long long x = jc->slot[r - 1][k];
            res += x >> 32;
            act += x & 0xFFFFFFFFu;
        }
        if (res == 0 || act == 0) {
            st = r - 1;
            cy = (int)act;
        }
    }
    *stop = st;
    *cyc = cy;
}

// Whether round r runs: always while some vertex still reaches a terminal;
// afterwards until the window 2^r covers the cycle-bound vertices.  Once a
// round does not run, no later round does (halt).
__device__ __forceinline__ int decide_round(int r, const JumpCnt* jc, int maxn, int* stop,
                                            int* cyc, int* halt)
{
    round_state(r, jc, stop, cyc, halt);
    if (*halt || r >= kMaxRounds) return 0;
    if (r == 0 || *stop < 0) return 1;
    return (r < 30 && (1 << r) < min(*cyc, maxn)) ? 1 : 0;
}

static_assert(kSlots == kWave, "a wave reads one slot row");

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long x)
{
    for (int o = kWave / 2; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

// decide_round by a whole wave: sm = the summary word of round r-1, x = this
// lane's slot word of round r-1 (lane k: slot k; kSlots == kWave), both loaded
// together; mis = nonzero if launch r-1 found a cycle-label mismatch (the same
// in every lane).  Every lane returns the same.
//
// Cycle labels (two-level launches, jump2_body).  For a vertex v still jumping
// at the launch's end, let lambda(v) be the parity of the largest priority in
// the level-r window that starts at x_r(v), v's 2^r-th successor: the label a
// valuation ending at level r gives v.  The launch's first two gather steps
// read the level-r records at x_r(v) and at x_r(x_r(v)), so it compares
// lambda(v) with lambda(x_r(v)) at no cost and flags any difference.  Once the
// stop has happened by the launch's end, the vertices still jumping are
// exactly the cycle-bound ones, and with no difference every label is already
// right: along a cycle the level-r windows at x_r(v), x_r(x_r(v)), ... follow
// one another and cover the cycle, so one of them holds the cycle's largest
// priority and the equal bits carry its parity to every cycle vertex; a tail
// vertex reaches the cycle through x_r, so it has the same bit.  A window at a
// later level is a union of such windows, so the labels stay right at every
// later level, and the rounds stop.
__device__ __forceinline__ int decide_wave(int r, unsigned long long sm, unsigned long long x, unsigned mis,
                                           int maxn, int* stop, int* cyc)
{
    const int halt = (int)(sm >> 63);
    const int checked = (int)((sm >> kSumChecked) & 1);
    int st = sum_stop(sm), cy = (int)(sm & 0xFFFFFFFFu);
    if (r >= 1 && st < 0 && !halt) {               // uniform: sm is the same in every lane
        const unsigned long long res = wave_sum_u64(x >> 32), act = wave_sum_u64(x & 0xFFFFFFFFu);
        if (res == 0 || act == 0) {
            st = r - 1;
            cy = (int)act;
        }
    }
    *stop = st;
    *cyc = cy;
    if (halt || r >= kMaxRounds) return 0;
    if (r == 0 || st < 0) return 1;
    if (checked && !mis) return 0;                // cycle labels consistent
    return (r < 30 && (1 << r) < min(cy, maxn)) ? 1 : 0;
}

__device__ __forceinline__ unsigned long long summary_word(int run, int stop, int cyc, int checked = 0)
{
    return ((unsigned long long)(run ? 0 : 1) << 63) | ((unsigned long long)checked << kSumChecked) |
           ((unsigned long long)(stop + 1) << 32) | (unsigned)cyc;
}

// ---------------------------------------------------------------------------
// init_strat (gpu_parity.cu:38-56; cpu_parity.cpp:276-286)
__global__ void k_init_strat(long long N, const int* __restrict__ vinfo, const int* __restrict__ off,
                             const int* __restrict__ tgt, int* __restrict__ strat,
                             int8_t* __restrict__ solved, int8_t* __restrict__ lab)
{
    long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    strat[v] = owner_of(vinfo[v]) == 0 ? -1 : tgt[off[v]];
    solved[v] = 0;
    lab[v] = 0;
}

// Restarts the control blocks of games [g0, g1).
__global__ void k_reset_ctl(int g0, int g1, GameCtl* gc)
{
    int g = g0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= g1) return;
    GameCtl c = gc[g];
    c.done = 0;
    c.br = 0;
    c.iter = 0;
    c.sw1 = 0;
    c.sw0 = 0;
    c.outer = 0;
    c.reset_pending = 0;
    c.capped = 0;
    c.clear0 = 0;
    c.bf_init = 0;
    c.err = 0;
    gc[g] = c;
}

// reset_opponent_strategy (cpu_parity.cpp:300-307)
__global__ void k_reset_opp(long long N, const int* __restrict__ vinfo, const int* __restrict__ off,
                            const int* __restrict__ tgt, int* __restrict__ strat)
{
    long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    if (owner_of(vinfo[v]) == 1) strat[v] = tgt[off[v]];
}

// Uses the values here (waiting for them): the loads that produced them stay
// above this point, in flight together, rather than being moved by the
// compiler next to their first use (under a branch, where each is waited
// for on its own).
__device__ __forceinline__ void issued(int a, int b, int c, int d, int e, int f)
{
    asm volatile("" ::"v"(a), "v"(b), "v"(c), "v"(d), "v"(e), "v"(f));
}

// Tag of the d1 marks of a valuation (game 0's br at its INIT): 1..255, so
// marks left by earlier valuations need no clearing (a stale mark whose tag
// recurs only adds work: the vertex is computed rather than kept).
__host__ __device__ __forceinline__ int d1_tag(int br) { return br % 255 + 1; }

// Marks v (whose link changed for the coming valuation) and its
// link-children in the d1 map with the valuation's tag: the in-edge sources
// s of v whose link is v (unsolved, strategy v; the strategies are final
// during an INIT).  Plain byte stores: writers of one byte agree.
__device__ __forceinline__ void d1_mark(int v, int tag, uint8_t* __restrict__ d1, const int* __restrict__ rin_off,
                                        const int* __restrict__ rin_src, const int* __restrict__ own0,
                                        const int* __restrict__ own1, const int* __restrict__ strat,
                                        const int8_t* __restrict__ solved)
{
    d1[v] = (uint8_t)tag;
    const int e1 = rin_off[v + 1];
    for (int e = rin_off[v]; e < e1; e += 8) {
        int src[8];
#pragma unroll
        for (int j = 0; j < 8; j++) src[j] = e + j < e1 ? rin_src[e + j] : -1;
#pragma unroll
        for (int j = 0; j < 8; j++) src[j] = src[j] < 0 ? -1 : ((src[j] & 1) ? own1 : own0)[src[j] >> 1];
        int st[8], so[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            st[j] = src[j] >= 0 ? strat[src[j]] : -1;
            so[j] = src[j] >= 0 ? solved[src[j]] : 1;
        }
#pragma unroll
        for (int j = 0; j < 8; j++)
            if (st[j] == v && !so[j]) d1[src[j]] = (uint8_t)tag;
    }
}

// Level-0 jump records.  S = (next or terminal code, max priority, count, -).
// Terminal codes: -1 sink, -2 solved +inf, -3 solved -inf, -4 finished game.
__device__ __forceinline__ void val_init_body(
    long long N, const int* __restrict__ vinfo, const int* __restrict__ off,
    const int* __restrict__ tgt, const int* __restrict__ gid, const GameCtl* __restrict__ gc,
    int* __restrict__ strat, const int8_t* __restrict__ solved, const int8_t* __restrict__ lab,
    int* __restrict__ par, int2* __restrict__ S0, unsigned long long* __restrict__ Z0,
    JumpCnt* __restrict__ jc, uint8_t* __restrict__ blive, uint8_t* __restrict__ bact0,
    uint8_t* __restrict__ blast, int trace, const int* __restrict__ hi, uint8_t* __restrict__ d1 = nullptr,
    const int* __restrict__ rin_off = nullptr, const int* __restrict__ rin_src = nullptr,
    const int* __restrict__ own0 = nullptr, const int* __restrict__ own1 = nullptr,
    uint8_t* __restrict__ bchg = nullptr)
{
    if (blockIdx.x == 0) {
        unsigned long long* sl = &jc->slot[0][0];
        for (int i = threadIdx.x; i < kMaxRounds * kSlots; i += blockDim.x) sl[i] = 0;
        if (threadIdx.x == 0) {
            jc->summary[0] = jc->summary[1] = 0;
            jc->last = -1;                         // jc->err stays sticky
        }
    }
    // Vertices at or above *hi belong to no unfinished game; no kernel of the
    // step visits their blocks.
    const long long lim = min(N, (long long)*hi);
    const long long nbk = (lim + kBS - 1) / kBS;
    const int lane = (int)(threadIdx.x & (kWave - 1));
    const long long nwv = (long long)gridDim.x * (kBS / kWave);
    // A wave takes a whole block, kInitRows rows of 64 vertices at a time
    // with their loads in flight together: the loads are unconditional, with
    // clamped indices (a load under a branch is waited for where the branch
    // joins), and no barrier separates the waves.
    for (long long b = (long long)blockIdx.x * (kBS / kWave) + threadIdx.x / kWave; b < nbk; b += nwv) {
        const long long base = b * kBS + lane;
        unsigned long long mlive[kRows];
        unsigned long long any = 0, anyj = 0;
        unsigned linkchg = 0;                     // bit per row: this lane's vertex has a new link
        unsigned long long tchg = 0;              // a terminal whose record has the change bit
#pragma unroll
        for (int h = 0; h < kRows; h += kInitRows) {
            int g[kInitRows], vi[kInitRows], s[kInitRows], old[kInitRows], sv[kInitRows], lv[kInitRows];
#pragma unroll
            for (int q = 0; q < kInitRows; q++) {
                const long long vc = min(base + (h + q) * kWave, N - 1);
                g[q] = gid ? gid[vc] : 0;        // no gid: one game
                vi[q] = vinfo[vc];
                s[q] = strat[vc];
                old[q] = par[vc];
                sv[q] = solved[vc];
                lv[q] = lab[vc];
            }
#pragma unroll
            for (int q = 0; q < kInitRows; q++) issued(g[q], vi[q], s[q], old[q], sv[q], lv[q]);
            // then the games' control words
            int done[kInitRows], rp[kInitRows], br[kInitRows], lsem[kInitRows];
#pragma unroll
            for (int q = 0; q < kInitRows; q++) {
                done[q] = gc[g[q]].done;
                rp[q] = gc[g[q]].reset_pending;
                br[q] = gc[g[q]].br;
                lsem[q] = gc[g[q]].bf;
            }
#pragma unroll
            for (int q = 0; q < kInitRows; q++) issued(done[q], rp[q], br[q], lsem[q], 0, 0);
#pragma unroll
            for (int q = 0; q < kInitRows; q++) {
                const long long v = base + (h + q) * kWave;
                const bool live = v < N && !done[q];
                bool jumping = false;
                if (live) {
                    int sq = s[q];
                    if (rp[q] && owner_of(vi[q]) == 1) {   // main.cpp:25-26
                        sq = tgt[off[v]];
                        strat[v] = sq;
                    }
                    // w = 1 if this vertex's link differs from the previous
                    // valuation's (the OR along a path says the path changed).
                    const int first = br[q] == 0;
                    // solved vertices are terminals with their frozen label,
                    // except under CPUList's semantics (bf 3), whose valuation
                    // never reads the solved flags (cpu_list_ranking.cpp:242-255)
                    if (sv[q] && lsem[q] != 3) {
                        const int code = term_code(lv[q] == 1 ? 1 : 2, 0);
                        if (first | (old[q] != code)) tchg |= 1ull;
                        S0[v] = make_int2(code, (first | (old[q] != code)) ? kChangedBit : 0);
                        if (old[q] != code) par[v] = code;   // (unchanged links: no store)
                        if (old[q] != code) linkchg |= 1u << (h + q);
                        if (trace) Z0[v] = 0;
                    } else {
                        if (old[q] != sq) par[v] = sq;
                        if (old[q] != sq) linkchg |= 1u << (h + q);
                        if (sq < 0 && (first | (old[q] != sq))) tchg |= 1ull;
                        S0[v] = make_int2(sq >= 0 ? sq : term_code(0, 1),
                                          pri_of(vi[q]) | ((first | (old[q] != sq)) ? kChangedBit : 0));
                        if (trace) Z0[v] = splitmix64((unsigned long long)(uint32_t)pri_of(vi[q]));
                        jumping = sq >= 0;
                    }
                }
                mlive[h + q] = __ballot(live);
                any |= mlive[h + q];
                anyj |= __ballot(jumping);
            }
        }
        // Blocks that hold no vertex of an unfinished game are skipped by the
        // jump and label kernels of this valuation; round 0 visits only
        // blocks with a vertex that still has to jump.  Other vertices of a
        // visited block hold a terminal record, so that the jump rounds never
        // follow what a finished game, or a free range of a serving pool, left
        // in the level buffers.
        if (any) {
#pragma unroll
            for (int q = 0; q < kRows; q++) {
                const long long v = base + q * kWave;
                if (v < N && !((mlive[q] >> lane) & 1)) S0[v] = make_int2(term_code(3, 0), 0);
            }
        }
        const bool tc = __any(tchg != 0);
        if (lane == 0) {
            blive[b] = (uint8_t)(any != 0);
            bact0[b] = (uint8_t)(anyj != 0);
            blast[b] = 0;
            if (bchg) bchg[b] = (uint8_t)tc;      // the jump rounds add resolved vertices
        }
        if (d1)
#pragma unroll 1
            for (unsigned m = linkchg; m; m &= m - 1)
                d1_mark((int)(base + (__ffs(m) - 1) * kWave), d1_tag(gc[0].br), d1, rin_off, rin_src, own0, own1,
                        strat, solved);
    }
}

__global__ void __launch_bounds__(kBS) k_val_init(
    long long N, const int* __restrict__ vinfo, const int* __restrict__ off,
    const int* __restrict__ tgt, const int* __restrict__ gid, const GameCtl* __restrict__ gc,
    int* __restrict__ strat, const int8_t* __restrict__ solved, const int8_t* __restrict__ lab,
    int* __restrict__ par, int2* __restrict__ S0, unsigned long long* __restrict__ Z0,
    JumpCnt* __restrict__ jc, uint8_t* __restrict__ blive, uint8_t* __restrict__ bact0,
    uint8_t* __restrict__ blast, int trace, const int* __restrict__ hi)
{
    val_init_body(N, vinfo, off, tgt, gid, gc, strat, solved, lab, par, S0, Z0, jc, blive, bact0,
                  blast, trace, hi);
}

// One pointer-jumping round over the blocks that still hold a jumping vertex
// (grid-stride, so launches after convergence cost a few thousand one-word
// reads).  A block that is not visited keeps its last written level; its
// records there are final, which is where gathers and the label kernel read
// them (level min(r, blast[block])).  Lifting compares only read a vertex's
// levels below its resolution level, which its block always wrote.
//
// Round r of a valuation that runs it.  cnt: two zeroed shared counters;
// the block's resolved/still-jumping counts go to slot_row.
__device__ __forceinline__ void jump_body(int r, long long N, int2* __restrict__ Sl,
                                          unsigned long long* __restrict__ Zl,
                                          unsigned long long* __restrict__ slot_row,
                                          const uint8_t* __restrict__ blive,
                                          const uint8_t* __restrict__ bact_in,
                                          uint8_t* __restrict__ bact_out,
                                          uint8_t* __restrict__ blast, int trace,
                                          const int* __restrict__ hi, unsigned int* cnt,
                                          uint8_t* __restrict__ bchg = nullptr)
{
    unsigned int& cnt_res = cnt[0];
    unsigned int& cnt_act = cnt[1];
    const int2* Sin = Sl + (size_t)level_slot(r) * N;
    int2* Sout = Sl + (size_t)level_slot(r + 1) * N;
    const int span = r < 28 ? 1 << r : 0;         // vertices of an unfinished level-r record
    const unsigned long long* Zin = Zl + (size_t)level_slot(r) * N;
    unsigned long long* Zout = Zl + (size_t)level_slot(r + 1) * N;
    unsigned my_res = 0, my_act = 0;
    const long long lim = min(N, (long long)*hi);
    const long long nbk = (lim + kBS - 1) / kBS;  // blocks below the bound
    const long long G = gridDim.x;
    const int lane = (int)(threadIdx.x & (kWave - 1));
    const int wid = (int)(threadIdx.x / kWave);
    // This workgroup's blocks are b0, b0 + G, ...; their flags are read 64 at
    // a time (one per lane, every wave alike), so blocks with nothing to do
    // cost no serial round trip.  The blocks with work are dealt to the
    // workgroup's waves in turn: a wave runs a whole block, the loads of its
    // kRows rows in flight together, and the waves' blocks overlap (no
    // barrier between them).
    for (long long b0 = blockIdx.x; b0 < nbk; b0 += (long long)kWave * G) {
        const long long bq = b0 + (long long)lane * G;
        const long long bc = min(bq, nbk - 1);    // both flags loaded together
        const int f_live = blive[bc], f_act = bact_in[bc];
        const bool on = bq < nbk && f_live && f_act;
        if (threadIdx.x < kWave && bq < nbk && !on) bact_out[bq] = 0;
        unsigned long long m = __ballot(on);
        for (int k = 0; m; m &= m - 1, k++) {
            if ((k & (kRows - 1)) != wid) continue;
            const long long b = b0 + (long long)(__ffsll((long long)m) - 1) * G;
            const long long base = b * kBS + lane;
            if (trace) {
                // trace digests: the same gathers, row by row, before the
                // records below replace level r + 1
#pragma unroll 1
                for (int q = 0; q < kRows; q++) {
                    const long long v = base + q * kWave;
                    if (v >= N) continue;
                    const int2 s0 = Sin[v];
                    unsigned long long z = Zin[v];
                    if (s0.x >= 0) {
                        const int bz = min((int)blast[s0.x / kBS], r);
                        z += Zl[(size_t)level_slot(bz) * N + s0.x];
                    }
                    Zout[v] = z;
                }
            }
            // Every load below is unconditional, with clamped indices (rows past
            // N read vertex N - 1, non-jumping rows read vertex 0), so that a
            // row's loads are all in flight before any result is waited for.
            int2 s[kRows], t[kRows];
            int bl[kRows];
#pragma unroll
            for (int q = 0; q < kRows; q++) {
                const long long v = base + q * kWave;
                s[q] = Sin[min(v, N - 1)];
                if (v >= N) s[q].x = -1;
            }
            // The successor's record at level r, loaded together with its
            // block's last level; a block that stopped earlier (its records
            // final) is read again at that level.
#pragma unroll
            for (int q = 0; q < kRows; q++) {
                const int x = max(s[q].x, 0);
                bl[q] = blast[x / kBS];
                t[q] = Sl[(size_t)level_slot(r) * N + x];
            }
#pragma unroll
            for (int q = 0; q < kRows; q++) {
                const bool again = s[q].x >= 0 && bl[q] < r;
                if (__any(again)) {
                    const int2 u = Sl[(size_t)level_slot(again ? bl[q] : r) * N + (again ? s[q].x : 0)];
                    if (again) t[q] = u;
                }
            }
            unsigned long long any = 0, tchg = 0;  // tchg: a record that ended with the change bit
#pragma unroll
            for (int q = 0; q < kRows; q++) {
                const long long v = base + q * kWave;
                bool res = false, act = false;
                if (s[q].x >= 0) {
                    const int2 u = t[q];
                    s[q].x = u.x >= 0 ? u.x : term_code(term_kind(u.x), span + term_count(u.x));
                    s[q].y = max(s[q].y & kPriMask, u.y & kPriMask) | ((s[q].y | u.y) & kChangedBit);
                    res = s[q].x < 0;
                    act = !res;
                }
                if (v < N) Sout[v] = s[q];
                const unsigned long long mr = __ballot(res), ma = __ballot(act);
                my_res += (unsigned)__popcll(mr);
                my_act += (unsigned)__popcll(ma);
                any |= ma;
                if (bchg) tchg |= __ballot(v < N && s[q].x < 0 && (s[q].y & kChangedBit));
            }
            if (lane == 0) {
                bact_out[b] = (uint8_t)(any != 0);
                blast[b] = (uint8_t)(r + 1);
                if (tchg) bchg[b] = 1;
            }
        }
    }
    if ((threadIdx.x & (kWave - 1)) == 0 && (my_res | my_act)) {
        atomicAdd(&cnt_res, my_res);
        atomicAdd(&cnt_act, my_act);
    }
    __syncthreads();
    if (threadIdx.x == 0 && (cnt_res | cnt_act))
        atomicAdd(&slot_row[blockIdx.x & (kSlots - 1)], ((unsigned long long)cnt_res << 32) | cnt_act);
}

// A record extended by the record of its end vertex: s spans `span` vertices
// and has not ended.
__device__ __forceinline__ int2 join_rec(int2 s, int2 t, int span)
{
    int2 o;
    o.x = t.x >= 0 ? t.x : term_code(term_kind(t.x), span + term_count(t.x));
    o.y = max(s.y & kPriMask, t.y & kPriMask) | ((s.y | t.y) & kChangedBit);
    return o;
}

// Two rounds in one launch: three dependent gather steps instead of two
// launches (one gather step each, after the first round few vertices have
// resolved).  From level r, level r + 1
// = s o rec(x1) and level r + 2 = (level r + 1) o rec(x2) o rec(x3), every
// gathered record at level r (written by an earlier launch; x1, x2, x3 the
// ends of the spans so far), with the same re-read at a stopped block's
// level as one round (levels r, r + 1 and r + 2 are in different buffers).
// The counts are round r + 1's (resolved at level r + 2,
// still jumping after it): a stop seen there is a stop at r or r + 1, and
// only whether there was one is ever used.  A vertex still jumping at level
// r + 2 whose cycle-label bit differs from that of its 2^r-th successor
// (decide_wave) sets mis_row[block % kSlots]; cnt[2] is a zeroed shared flag.
__device__ __forceinline__ void jump2_body(int r, long long N, int2* __restrict__ Sl,
                                           unsigned long long* __restrict__ Zl,
                                           unsigned long long* __restrict__ slot_row,
                                           const uint8_t* __restrict__ blive,
                                           const uint8_t* __restrict__ bact_in,
                                           uint8_t* __restrict__ bact_out,
                                           uint8_t* __restrict__ blast, int trace,
                                           const int* __restrict__ hi, unsigned int* cnt,
                                           unsigned* __restrict__ mis_row, const uint8_t* __restrict__ d1 = nullptr,
                                           uint8_t* __restrict__ l12ok = nullptr, int tag = 0,
                                           uint8_t* __restrict__ bchg = nullptr)
{
    // Round 0 with l12ok (one game, step program): the launch records in
    // l12ok which blocks it wrote levels 1 and 2 of; with d1 as well (the
    // INIT marked the vertices with a new link and their link-children), a
    // vertex of such a block whose level-1 and level-2 windows hold no new
    // link (it is unmarked and so is its 2nd successor, read from its
    // level-1 record) keeps both records from the last valuation, with the
    // change bit cleared.
    const bool rec_ok = r == 0 && l12ok != nullptr;
    unsigned int& cnt_res = cnt[0];
    unsigned int& cnt_act = cnt[1];
    const int2* Sin = Sl + (size_t)level_slot(r) * N;
    int2* Sout1 = Sl + (size_t)level_slot(r + 1) * N;
    int2* Sout2 = Sl + (size_t)level_slot(r + 2) * N;
    const int span = 1 << r;
    const unsigned long long* Zin = Zl + (size_t)level_slot(r) * N;
    unsigned long long* Zout1 = Zl + (size_t)level_slot(r + 1) * N;
    unsigned long long* Zout2 = Zl + (size_t)level_slot(r + 2) * N;
    unsigned my_res = 0, my_act = 0;
    const long long lim = min(N, (long long)*hi);
    const long long nbk = (lim + kBS - 1) / kBS;
    const long long G = gridDim.x;
    const int lane = (int)(threadIdx.x & (kWave - 1));
    const int wid = (int)(threadIdx.x / kWave);
    for (long long b0 = blockIdx.x; b0 < nbk; b0 += (long long)kWave * G) {
        const long long bq = b0 + (long long)lane * G;
        const long long bc = min(bq, nbk - 1);
        const int f_live = blive[bc], f_act = bact_in[bc];
        const int f_ok = rec_ok && d1 ? l12ok[bc] : 0;
        const bool on = bq < nbk && f_live && f_act;
        if (threadIdx.x < kWave && bq < nbk && !on) {
            bact_out[bq] = 0;
            if (rec_ok) l12ok[bq] = 0;
        }
        unsigned long long m = __ballot(on);
        const unsigned long long mok = __ballot(f_ok != 0);
        for (int k = 0; m; m &= m - 1, k++) {
            if ((k & (kRows - 1)) != wid) continue;
            const int bi = __ffsll((long long)m) - 1;
            const long long b = b0 + (long long)bi * G;
            const long long base = b * kBS + lane;
            if (trace) {
                // trace digests, row by row, from level r alone (as below)
#pragma unroll 1
                for (int q = 0; q < kRows; q++) {
                    const long long v = base + q * kWave;
                    if (v >= N) continue;
                    int2 c = Sin[v];
                    unsigned long long z = Zin[v];
#pragma unroll 1
                    for (int step = 0; step < 3 && c.x >= 0; step++) {
                        const int bz = min((int)blast[c.x / kBS], r);
                        const size_t at = (size_t)level_slot(bz) * N + c.x;
                        z += Zl[at];
                        c = join_rec(c, Sl[at], (step == 2 ? 3 : step + 1) * span);
                        if (step == 0) Zout1[v] = z;
                    }
                    if (Sin[v].x < 0) Zout1[v] = z;
                    Zout2[v] = z;
                }
            }
            // The rows' records, then three steps of gathers for all rows at
            // once (unconditional, clamped indices, as in one round).
            int2 s[kRows];
            bool act1[kRows];                     // still jumping at level r + 1
            int lp = 0;                           // bit q: row q's lambda(v) xor lambda(x_r(v))
            unsigned cmp = (1u << kRows) - 1;     // bit q: row q's vertex is computed here
#pragma unroll
            for (int q = 0; q < kRows; q++) {
                const long long v = base + q * kWave;
                s[q] = Sin[min(v, N - 1)];
                if (v >= N) s[q].x = -1;
            }
            if ((mok >> bi) & 1) {
                // The marks and level-1 records of the rows, then the marks
                // of the 2nd successors, then the level-2 records: every load
                // unconditional (the vertex's own, or vertex 0's mark for a
                // path that ends within two vertices), so that the rows'
                // loads are in flight together.  A terminal keeps its records
                // too (copies of its unchanged level-0 record).
                int m0[kRows], m2[kRows];
                int2 o1[kRows], o2[kRows];
#pragma unroll
                for (int q = 0; q < kRows; q++) {
                    const long long vc = min(base + q * kWave, N - 1);
                    m0[q] = d1[vc];
                    o1[q] = Sl[(size_t)level_slot(1) * N + vc];
                }
#pragma unroll
                for (int q = 0; q < kRows; q++) m2[q] = d1[max(o1[q].x, 0)];
                unsigned keep = 0;
#pragma unroll
                for (int q = 0; q < kRows; q++) {
                    const long long v = base + q * kWave;
                    if (v < N && m0[q] != tag && (o1[q].x < 0 || m2[q] != tag)) keep |= 1u << q;
                    o2[q] = Sl[(size_t)level_slot(2) * N + min(v, N - 1)];
                }
#pragma unroll
                for (int q = 0; q < kRows; q++) {
                    if (!((keep >> q) & 1)) continue;
                    const long long v = base + q * kWave;
                    if (o1[q].y & kChangedBit) Sl[(size_t)level_slot(1) * N + v].y = o1[q].y & ~kChangedBit;
                    if (o2[q].y & kChangedBit) Sl[(size_t)level_slot(2) * N + v].y = o2[q].y & ~kChangedBit;
                    act1[q] = o1[q].x >= 0;
                    s[q] = o2[q];
                }
                cmp &= ~keep;
            }
#pragma unroll 1
            for (int step = 0; step < 3; step++) {
                int2 t[kRows];
                int bl[kRows];
#pragma unroll
                for (int q = 0; q < kRows; q++) {
                    const int x = (cmp >> q) & 1 ? max(s[q].x, 0) : 0;
                    bl[q] = blast[x / kBS];
                    t[q] = Sl[(size_t)level_slot(r) * N + x];
                }
#pragma unroll
                for (int q = 0; q < kRows; q++) {
                    const bool again = (cmp >> q) & 1 && s[q].x >= 0 && bl[q] < r;
                    if (__any(again)) {
                        const int2 u = Sl[(size_t)level_slot(again ? bl[q] : r) * N + (again ? s[q].x : 0)];
                        if (again) t[q] = u;
                    }
                }
                const int sp = (step == 2 ? 3 : step + 1) * span;   // vertices s spans so far
                // the level-r windows at x_r(v) (step 0) and x_r(x_r(v)) (step 1)
                if (step < 2)
#pragma unroll
                    for (int q = 0; q < kRows; q++) lp ^= (t[q].y & 1) << q;
#pragma unroll
                for (int q = 0; q < kRows; q++)
                    if ((cmp >> q) & 1 && s[q].x >= 0) s[q] = join_rec(s[q], t[q], sp);
                if (step == 0)
#pragma unroll
                    for (int q = 0; q < kRows; q++) {
                        const long long v = base + q * kWave;
                        if ((cmp >> q) & 1) {
                            if (v < N) Sout1[v] = s[q];       // level r + 1
                            act1[q] = s[q].x >= 0;
                        }
                    }
                bool more = false;
#pragma unroll
                for (int q = 0; q < kRows; q++) more |= (cmp >> q) & 1 && s[q].x >= 0;
                if (!__any(more)) break;
            }
            unsigned long long any = 0, tchg = 0;  // tchg: a computed record that ended with the change bit
            bool mis = false;
#pragma unroll
            for (int q = 0; q < kRows; q++) {
                const long long v = base + q * kWave;
                if (v < N && (cmp >> q) & 1) Sout2[v] = s[q];   // level r + 2
                const bool act = s[q].x >= 0;
                // resolved in round r + 1: jumping at level r + 1, not at r + 2
                const bool res = act1[q] && !act;
                const unsigned long long ma = __ballot(act);
                my_res += (unsigned)__popcll(__ballot(res));
                my_act += (unsigned)__popcll(ma);
                any |= ma;
                // a kept record has no cycle-label check: counted as a mismatch
                mis |= act && (((lp >> q) & 1) || !((cmp >> q) & 1));
                if (bchg) tchg |= __ballot(v < N && (cmp >> q) & 1 && !act && (s[q].y & kChangedBit));
            }
            if (__any(mis) && lane == 0) cnt[2] = 1;
            if (lane == 0) {
                bact_out[b] = (uint8_t)(any != 0);
                blast[b] = (uint8_t)(r + 2);
                if (rec_ok) l12ok[b] = 1;
                if (tchg) bchg[b] = 1;
            }
        }
    }
    if ((threadIdx.x & (kWave - 1)) == 0 && (my_res | my_act)) {
        atomicAdd(&cnt_res, my_res);
        atomicAdd(&cnt_act, my_act);
    }
    __syncthreads();
    if (threadIdx.x == 0 && (cnt_res | cnt_act))
        atomicAdd(&slot_row[blockIdx.x & (kSlots - 1)], ((unsigned long long)cnt_res << 32) | cnt_act);
    if (threadIdx.x == 0 && cnt[2]) mis_row[blockIdx.x & (kSlots - 1)] = 1;
}

__global__ void __launch_bounds__(kBS) k_jump(int r, long long N, int maxn, int2* __restrict__ Sl,
                                              unsigned long long* __restrict__ Zl,
                                              JumpCnt* __restrict__ jc,
                                              const uint8_t* __restrict__ blive,
                                              const uint8_t* __restrict__ bact_in,
                                              uint8_t* __restrict__ bact_out,
                                              uint8_t* __restrict__ blast, int trace,
                                              const int* __restrict__ hi)
{
    __shared__ int run_s;
    __shared__ unsigned int cnt[2];
    if (threadIdx.x < kWave) {                    // the first wave decides
        const unsigned long long sm = r >= 1 ? jc->summary[(r - 1) & 1] : 0ull;
        const unsigned long long x = r >= 1 ? jc->slot[r - 1][threadIdx.x] : 0ull;
        int stop, cyc;
        const int run = decide_wave(r, sm, x, 1u, maxn, &stop, &cyc);
        if (threadIdx.x == 0) {
            run_s = run;
            cnt[0] = cnt[1] = 0;
            if (blockIdx.x == 0) {
                jc->summary[r & 1] = summary_word(run, stop, cyc);
                if (run) jc->last = r;
            }
        }
    }
    __syncthreads();
    if (!run_s) return;
    jump_body(r, N, Sl, Zl, jc->slot[r], blive, bact_in, bact_out, blast, trace, hi, cnt);
}

// Labels (d_mark_infinite role), compare keys, trace hashes; in the loop form
// also mark_solved(1) (cpu_parity.cpp:265-274), which solve_si calls right
// after every valuation.
//
// check_rounds: flag the valuation if a round at lmax would still have run
// (the launch sequence has lmax rounds).
__device__ __forceinline__ void label_body(
    long long N, int maxn, int lmax, const int2* __restrict__ Sl,
    const unsigned long long* __restrict__ Zl, const uint8_t* __restrict__ blast,
    const int* __restrict__ gid, GameCtl* __restrict__ gc, const int* __restrict__ strat,
    int8_t* __restrict__ solved, int8_t* __restrict__ lab, int* __restrict__ dep,
    int* __restrict__ key, unsigned long long* __restrict__ dig, unsigned long long* __restrict__ tr,
    int trace_cap, int loop_form, int trace, JumpCnt* __restrict__ jc,
    const uint8_t* __restrict__ blive, const int* __restrict__ hi,
    unsigned long long* __restrict__ chg, unsigned long long* __restrict__ chg0, int check_rounds,
    unsigned* __restrict__ ccnt, const uint8_t* __restrict__ bchg = nullptr,
    const uint8_t* __restrict__ bfin = nullptr)
{
    // Quiet blocks (one game, plain semantics, not its first valuation, no
    // trace; bchg and bfin given): a block none of whose records ended with
    // the change bit (bchg: set by the INIT for its terminals and by the
    // jump launches for what they resolved) and with no vertex still jumping
    // after the last round (bfin: none cycle-bound) holds no vertex whose
    // valuation, label or solved flag changes, so only its change words are
    // written (chg zero; chg0 zero after an outer step, else as it was).
    const int quiet_clr = bchg ? gc[0].clear0 : 0;
    const int last = jc->last;
    unsigned wchg = 0;                            // this wave's changed vertices (ccnt)
    const int F = last + 1;
    if (check_rounds && blockIdx.x == 0 && threadIdx.x == 0) {
        int stop, cyc, halt;
        if (decide_round(lmax, jc, maxn, &stop, &cyc, &halt)) jc->err = 1;
    }
    const long long lim = min(N, (long long)*hi);
    const long long nbk = (lim + kBS - 1) / kBS;
    const long long G = gridDim.x;
    const int lane = (int)(threadIdx.x & (kWave - 1));
    const int wid = (int)(threadIdx.x / kWave);
    // Blocks b0, b0 + G, ... with their live flags read 64 at a time; the
    // live ones are dealt to the workgroup's waves in turn, a wave taking a
    // whole block, kInitRows rows at a time with their loads in flight
    // together (unconditional, clamped indices; no barrier between waves).
    for (long long b0 = blockIdx.x; b0 < nbk; b0 += (long long)kWave * G) {
        const long long bq = b0 + (long long)lane * G;
        const long long bc = min(bq, nbk - 1);
        const bool bl_live = blive[bc] != 0;
        const bool quiet = bchg && !bchg[bc] && !bfin[bc];
        unsigned long long m = __ballot(bq < nbk && bl_live);
        const unsigned long long mq = __ballot(quiet);
        for (int k = 0; m; m &= m - 1, k++) {
            if ((k & (kRows - 1)) != wid) continue;
            const int bi = __ffsll((long long)m) - 1;
            const long long b = b0 + (long long)bi * G;
            if ((mq >> bi) & 1) {
                if (lane < kRows) {
                    chg[b * kRows + lane] = 0;
                    if (quiet_clr) chg0[b * kRows + lane] = 0;
                }
                continue;
            }
            // this block's final level (it may have stopped jumping earlier)
            const int Fb = min(F, (int)blast[b]);
            const int2* SF = Sl + (size_t)level_slot(Fb) * N;
            const unsigned long long* ZF = Zl + (size_t)level_slot(Fb) * N;
#pragma unroll
            for (int h = 0; h < kRows; h += kInitRows) {
                int g[kInitRows], sol[kInitRows], lold[kInitRows];
                int2 s[kInitRows];
                unsigned long long c0w[kInitRows];
#pragma unroll
                for (int q = 0; q < kInitRows; q++) {
                    const long long v = b * kBS + (h + q) * kWave + lane;
                    const long long vc = min(v, N - 1);
                    g[q] = gid ? gid[vc] : 0;    // no gid: one game
                    sol[q] = solved[vc];
                    lold[q] = lab[vc];
                    s[q] = SF[vc];
                    c0w[q] = chg0[v >> 6];            // the row's word (every lane alike)
                }
#pragma unroll
                for (int q = 0; q < kInitRows; q++) issued(g[q], sol[q], lold[q], s[q].x, s[q].y, (int)c0w[q]);
                // then the games' control words
                int done[kInitRows], vbase[kInitRows], clr0[kInitRows], bf[kInitRows], bfi[kInitRows];
                int first[kInitRows];
#pragma unroll
                for (int q = 0; q < kInitRows; q++) {
                    done[q] = gc[g[q]].done;
                    vbase[q] = gc[g[q]].vbase;
                    clr0[q] = gc[g[q]].clear0;
                    bf[q] = gc[g[q]].bf;
                    bfi[q] = gc[g[q]].bf_init;
                    first[q] = gc[g[q]].br == 0;
                }
#pragma unroll
                for (int q = 0; q < kInitRows; q++) issued(done[q], vbase[q], clr0[q], bf[q], bfi[q], first[q]);
#pragma unroll
                for (int q = 0; q < kInitRows; q++) {
                    const long long v = b * kBS + (h + q) * kWave + lane;
                    const bool live = v < N && !done[q];
                    const int kind = s[q].x < 0 ? term_kind(s[q].x) : -1;
                    // A cycle-bound vertex: its 2^F-th successor is on the
                    // cycle, whose largest priority is that successor's
                    // window maximum at its block's final level (loads for
                    // the whole row, other lanes reading their own vertex).
                    int cyc_pri = 0;
                    const bool cyc = live && !sol[q] && kind < 0 && bf[q] != 3;
                    if (__any(cyc)) {
                        const int x = cyc ? s[q].x : (int)min(v, N - 1);
                        const int Fy = min(F, (int)blast[x / kBS]);
                        cyc_pri = Sl[(size_t)level_slot(Fy) * N + x].y;
                    }
                    int l = 0, loc = 0;
                    unsigned long long d = 0;
                    bool ch = false;
                    if (live) {
                        loc = (int)(v - vbase[q]);
                        // Key and depth are stored only in a game's first
                        // valuation and when the valuation changed: a finite
                        // path with no changed link is the same path (same
                        // length and maximum), an infinite label that stayed
                        // has the same key, a solved vertex keeps its label.
                        int kv, dv;
                        if (sol[q] && bf[q] != 3) {
                            l = lold[q];
                            kv = (l + 1) << 30;
                            dv = 0;
                        } else {
                            if (kind == 0) {                   // reaches the sink: finite
                                l = 0;
                                dv = term_count(s[q].x);
                                kv = (1 << 30) | (s[q].y & kPriMask);
                                ch = (s[q].y & kChangedBit) != 0; // a link on the path changed
                                if (trace) d = ZF[v];
                            } else {
                                if (kind == 1) l = 1;
                                else if (kind == 2) l = -1;
                                else if (kind < 0) {           // largest priority on the cycle
                                    l = (cyc_pri & 1) ? -1 : 1;
                                    // solve_bf (bf 1): the relaxation never closes a cycle
                                    // (such vertices stay infinite, cpu_bf.cpp:156-160);
                                    // the first best response stops on an odd one.
                                    // CPUBV solve (bf 2): every valuation stops on one
                                    // (cpu_bv.cpp:73-84).  CPUList (bf 3): every vertex
                                    // that does not reach the sink is +inf
                                    // (broadcast_val, cpu_list_ranking.cpp:208-239).
                                    if (bf[q] == 3) {
                                        l = 1;
                                    } else if (bf[q]) {
                                        if (bfi[q] && bf[q] == 1) l = 1;
                                        else if (l == -1) gc[g[q]].err = 1;
                                    }
                                } else {
                                    l = 1;                     // kind 3 (finished game): never on a live path
                                }
                                kv = (l + 1) << 30;
                                dv = 0;
                            }
                            ch = ch || l != lold[q];
                            if (l != lold[q]) lab[v] = (int8_t)l;
                            if (loop_form && l == -1) solved[v] = 1;
                        }
                        if (first[q] || ch) {
                            key[v] = kv;
                            dep[v] = dv;
                        }
                        if (trace) dig[v] = d;
                    }
                    // Valuation changed since the previous one (one word per
                    // row), and since the last switch(0) pass of the vertex's
                    // game: that word is cleared for games that took the outer
                    // step in the previous step.
                    const unsigned long long cm = __ballot(ch);
                    const unsigned long long clr = __ballot(live && clr0[q] != 0);
                    wchg += (unsigned)__popcll(cm);
                    if (lane == 0) {
                        chg[v >> 6] = cm;
                        chg0[v >> 6] = (c0w[q] & ~clr) | cm;
                    }
                    if (trace) {
                        const int gq = live ? g[q] : -1;
                        int br = live ? gc[gq].br : 0;
                        bool rec = live && br < trace_cap;
                        unsigned long long* base = rec ? tr + ((size_t)gq * trace_cap + br) * 4 : nullptr;
                        int sv = rec ? strat[v] : -1;
                        if (sv >= 0) sv -= vbase[q];              // local ids, as a game alone
                        const int so = live ? sol[q] : 0;
                        wave_add_u64(rec ? base + 0 : nullptr, rec ? hmix_i32(loc, sv) : 0);
                        wave_add_u64(rec ? base + 1 : nullptr, rec ? hmix_i32(loc, so) : 0);
                        wave_add_u64(rec ? base + 2 : nullptr, rec ? hmix_i32(loc, l) : 0);
                        wave_add_u64(rec ? base + 3 : nullptr, rec ? hmix_u64(loc, d) : 0);
                    }
                }
            }
        }
    }
    if (ccnt) {                                   // one atomic per workgroup, a line per slot
        __shared__ unsigned wg_chg;
        if (threadIdx.x == 0) wg_chg = 0;
        __syncthreads();
        if (lane == 0 && wchg) atomicAdd(&wg_chg, wchg);
        __syncthreads();
        if (threadIdx.x == 0 && wg_chg) atomicAdd(&ccnt[(blockIdx.x & (kSlots - 1)) * kCntStride], wg_chg);
    }
}

__global__ void __launch_bounds__(kBS) k_label(
    long long N, int maxn, int lmax, const int2* __restrict__ Sl,
    const unsigned long long* __restrict__ Zl, const uint8_t* __restrict__ blast,
    const int* __restrict__ gid, GameCtl* __restrict__ gc, const int* __restrict__ strat,
    int8_t* __restrict__ solved, int8_t* __restrict__ lab, int* __restrict__ dep,
    int* __restrict__ key, unsigned long long* __restrict__ dig, unsigned long long* __restrict__ tr,
    int trace_cap, int loop_form, int trace, JumpCnt* __restrict__ jc,
    const uint8_t* __restrict__ blive, const int* __restrict__ hi,
    unsigned long long* __restrict__ chg, unsigned long long* __restrict__ chg0)
{
    label_body(N, maxn, lmax, Sl, Zl, blast, gid, gc, strat, solved, lab, dep, key, dig, tr, trace_cap,
               loop_form, trace, jc, blive, hi, chg, chg0, 1, nullptr);
}

// Exact compare of two finite vertices whose path maxima are equal.  The jump
// records of every level are binary-lifting tables (level j: 2^j-th
// successor and the largest priority over those 2^j vertices), so the lowest
// common ancestor and the largest priority of each private segment take
// O(log depth) gathers.  Returns the sign when the two segment maxima differ;
// otherwise 2, with the lengths of the private segments (u's first la
// vertices and w's first lb vertices, before the common part), whose
// multisets coop_scan compares.
__device__ __forceinline__ int cmp_lca(int u, int w, const int* __restrict__ dep,
                                    const int2* __restrict__ Sl, const int* __restrict__ vinfo,
                                    long long N, int F, int* la_out, int* lb_out)
{
    const int T = min(F, kLiftCap);                   // highest stored table level
    auto rec = [&](int j, int x) { return Sl[(size_t)j * N + x]; };
    int a = u, b = w, da = dep[u], db = dep[w];
    int mA = -1, mB = -1;
    // Bring the deeper one up to the other's depth.
    while (da - db >= (1 << T)) { int2 t = rec(T, a); mA = max(mA, t.y & kPriMask); a = t.x; da -= 1 << T; }
    while (db - da >= (1 << T)) { int2 t = rec(T, b); mB = max(mB, t.y & kPriMask); b = t.x; db -= 1 << T; }
    for (int j = T - 1; j >= 0; j--) {
        if (da - db >= (1 << j)) { int2 t = rec(j, a); mA = max(mA, t.y & kPriMask); a = t.x; da -= 1 << j; }
        if (db - da >= (1 << j)) { int2 t = rec(j, b); mB = max(mB, t.y & kPriMask); b = t.x; db -= 1 << j; }
    }
    int dl;                                           // depth of the common ancestor
    if (a == b) {
        dl = da;
    } else {
        // Jump both while their ancestors differ (only levels shorter than the
        // remaining path are read: longer windows end at the sink for both).
        while ((1 << T) < da) {
            int2 ta = rec(T, a), tb = rec(T, b);
            if (ta.x == tb.x) break;
            mA = max(mA, ta.y & kPriMask); mB = max(mB, tb.y & kPriMask); a = ta.x; b = tb.x; da -= 1 << T;
        }
        for (int j = T - 1; j >= 0; j--) {
            if ((1 << j) >= da) continue;
            int2 ta = rec(j, a), tb = rec(j, b);
            if (ta.x != tb.x) {
                mA = max(mA, ta.y & kPriMask); mB = max(mB, tb.y & kPriMask); a = ta.x; b = tb.x; da -= 1 << j;
            }
        }
        mA = max(mA, pri_of(vinfo[a]));
        mB = max(mB, pri_of(vinfo[b]));
        dl = da - 1;
    }
    if (mA != mB) {
        if (mA > mB) return (mA & 1) ? -1 : 1;
        return (mB & 1) ? 1 : -1;
    }
    *la_out = dep[u] - dl;
    *lb_out = dep[w] - dl;
    return 2;
}

// The vertex d steps along x's path (d below the path's length).
__device__ __forceinline__ int path_step(int x, int d, const int2* __restrict__ Sl, long long N, int T)
{
    while (d >= (1 << T)) { x = Sl[(size_t)T * N + x].x; d -= 1 << T; }
    for (int j = T - 1; j >= 0; j--)
        if (d >= (1 << j)) { x = Sl[(size_t)j * N + x].x; d -= 1 << j; }
    return x;
}

__device__ __forceinline__ int wave_max(int x)
{
    for (int o = kWave / 2; o > 0; o >>= 1) x = max(x, __shfl_xor(x, o));
    return x;
}

__device__ __forceinline__ int wave_sum(int x)
{
    for (int o = kWave / 2; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

// Compares the priority multisets of u's first la and w's first lb path
// vertices from the top down (compare_valuation's finite case,
// cpu_parity.cpp:194-236): the largest priority whose counts differ decides,
// even priorities counting for player 0.  Run by the whole wave: each lane
// walks 1/64 of each segment (its start found through the lifting tables), and
// each step of the top-down compare is a wave reduction.  Every lane returns
// the result.
__device__ __forceinline__ int coop_scan(int u, int la, int w, int lb, const int* __restrict__ par,
                                      const int* __restrict__ vinfo, const int2* __restrict__ Sl,
                                      long long N, int F, unsigned long long* __restrict__ prof)
{
    const int T = min(F, kLiftCap);
    const int lane = threadIdx.x & (kWave - 1);
    const int pa = (int)(((long long)lane * la) >> 6), na = (int)(((long long)(lane + 1) * la) >> 6) - pa;
    const int pb = (int)(((long long)lane * lb) >> 6), nb = (int)(((long long)(lane + 1) * lb) >> 6) - pb;
    const int xa = na > 0 ? path_step(u, pa, Sl, N, T) : -1;
    const int xb = nb > 0 ? path_step(w, pb, Sl, N, T) : -1;
    if (prof && lane == 0) atomicAdd(prof + 1, 1ull);
    int thr = 0x7FFFFFFF;
    for (;;) {
        if (prof && lane == 0) atomicAdd(prof + 2, (unsigned long long)(la + lb));
        int ma = -1, ca = 0, mb = -1, cb = 0;
        int x = xa;
        for (int i = 0; i < na; i++) {
            const int p = pri_of(vinfo[x]);
            if (p < thr) {
                if (p > ma) { ma = p; ca = 1; }
                else if (p == ma) ca++;
            }
            x = par[x];
        }
        x = xb;
        for (int i = 0; i < nb; i++) {
            const int p = pri_of(vinfo[x]);
            if (p < thr) {
                if (p > mb) { mb = p; cb = 1; }
                else if (p == mb) cb++;
            }
            x = par[x];
        }
        const int MA = wave_max(ma), MB = wave_max(mb);
        if (MA != MB) {
            if (MA > MB) return (MA & 1) ? -1 : 1;
            return (MB & 1) ? 1 : -1;
        }
        if (MA < 0) return 0;
        const int CA = wave_sum(ma == MA ? ca : 0), CB = wave_sum(mb == MB ? cb : 0);
        if (CA != CB) {
            if ((MA & 1) == 0) return CA > CB ? 1 : -1;
            return CA < CB ? 1 : -1;
        }
        thr = MA;
    }
}

// Sign of compare_valuation(u, w) (cpu_parity.cpp:194-236) from the two keys;
// -1 is the sink (finite, empty path).  *tie: both finite with equal path
// maxima (the exact compare decides).
__device__ __forceinline__ int cmp_fast(int u, int ku, int w, int kw, bool* tie)
{
    *tie = false;
    if (u == w) return 0;
    int cu = (int)((unsigned)ku >> 30), cw = (int)((unsigned)kw >> 30);   // +inf keys are negative ints
    if (cu != cw) return cu > cw ? 1 : -1;
    if (cu != 1) return 0;
    int pu = u < 0 ? -1 : (ku & kPriMask);
    int pw = w < 0 ? -1 : (kw & kPriMask);
    if (pu != pw) {
        if (pu > pw) return (pu & 1) ? -1 : 1;
        return (pw & 1) ? 1 : -1;
    }
    *tie = true;
    return 0;
}

__device__ __forceinline__ int key_of(const int* __restrict__ key, int u)
{
    return u < 0 ? (1 << 30) : key[u];
}

// switch_strategy(player) (cpu_parity.cpp:239-263, gpu_parity.cu:264-317):
// keep the current successor unless one is strictly better; the first
// strictly-better-than-current edge in CSR order wins each step of the scan.
// Loop form with player 0 also applies mark_solved(0) first.
// Successors compared per batch of loads (registers for targets and keys).
constexpr int kSwBatch = 5;

__device__ __forceinline__ void switch_body(
    int player, long long N, const int* __restrict__ own, long long own_n,
    const int* __restrict__ vinfo, const int* __restrict__ off,
    const int* __restrict__ tgt, const int* __restrict__ gid, GameCtl* __restrict__ gc,
    int* __restrict__ strat, const int* __restrict__ key, const int* __restrict__ dep,
    const int* __restrict__ par, const int2* __restrict__ Sl, const JumpCnt* __restrict__ jc,
    int loop_form, int* __restrict__ flag, unsigned long long* __restrict__ ties,
    const int* __restrict__ hi, const unsigned long long* __restrict__ chg,
    const unsigned long long* __restrict__ chg0, uint8_t* __restrict__ dq, int use_d)
{
    // Loop form: a vertex left at the scan's fixed point stays there unless
    // the valuation of its current successor or of a candidate changed since
    // its last scan; only such vertices gather keys.  (The result of a scan
    // is at least as good as every successor, so scanning it again switches
    // nothing.)  Player 1 scans after every valuation (chg: changed at the
    // last one); player 0 at outer steps (chg0: changed since then).
    const bool incremental = loop_form != 0;
    const unsigned long long* __restrict__ cb = player == 1 ? chg : chg0;
    const int F = jc->last + 1;                   // levels of the last valuation
    // A serving pool's owner lists are indexed like its vertices (holes are
    // -1), so nothing at or above the occupied bound is a vertex.
    // Waves take runs of 64 owner-list entries, grid-stride.
    const int lim = (int)min(own_n, (long long)*hi);
    const int nruns = (lim + kWave - 1) / kWave;
    const int nw = (int)(gridDim.x * (blockDim.x / kWave));
    const int lane = (int)(threadIdx.x & (kWave - 1));
    for (int w = (int)(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave); w < nruns; w += nw) {
        const int i = w * kWave + lane;
        // Change marks (dq, one game): an entry is marked when the valuation
        // of one of its successors changed since its last scan.  The marks
        // are consumed here; with use_d only marked entries are visited, and
        // they are exactly the entries the change bits would select (with
        // the current successor among the successors), plus at most entries
        // whose scan then compares no candidate.
        bool dv = false;
        if (dq) {
            dv = i < lim && dq[i] != 0;
            if (dv) dq[i] = 0;
            if (use_d && !__any(dv)) continue;
        }
        bool sw = false, live = false;
        // Only the successors whose valuation changed can beat a current
        // successor whose valuation did not (the others were no better than it
        // at the last scan, which left it a maximum), so such a vertex compares
        // just those, in CSR order: the result is the full scan's.
        bool only_changed = false;
        int g = -1, c0 = -1, eb = 0, ee = 0;
        const int v = i < lim && (!use_d || dv) ? own[i] : -1;
        if (v >= 0) {
            g = gid ? gid[v] : 0;                 // no gid: one game
            c0 = strat[v];                        // loaded together with the game's control
            eb = off[v];
            ee = off[v + 1];
            live = !gc[g].done && (!loop_form || player == 1 || gc[g].outer);
            if (live && incremental && !(player == 1 && gc[g].reset_pending)) {
                bool dirty = c0 >= 0 && ((cb[c0 >> 6] >> (c0 & 63)) & 1);
                only_changed = !dirty;
                if (use_d) dirty = true;          // marked: some successor changed
                const int e1 = ee;
                for (int e0 = eb; e0 < e1 && !dirty; e0 += 8) {
                    const int cnt = min(8, e1 - e0);
                    int us[8];
#pragma unroll
                    for (int k = 0; k < 8; k++) us[k] = k < cnt ? tgt[e0 + k] : -1;
#pragma unroll
                    for (int k = 0; k < 8; k++)
                        if (us[k] >= 0) dirty |= ((cb[us[k] >> 6] >> (us[k] & 63)) & 1) != 0;
                }
                live = dirty;
            }
        }
        // Successors in batches of kSwBatch: the target and key loads of a batch are
        // independent, so they are in flight together; the scan itself is
        // sequential in CSR order (the reference's tie-break).  The loop runs
        // for the whole wave, so that exact compares whose segment maxima tie
        // are done by all lanes together (coop_scan), one at a time.
        int cur = live ? c0 : -1;
        // The current successor is one of the successors, so its key arrives
        // with the first batch's keys (no load of its own) unless it sits
        // further on in a longer edge list.
        int kcur = live && cur < 0 ? (1 << 30) : 0;
        // First the best candidate by key alone.  The exact order refines the
        // key order, and the scan ends at the first candidate (current
        // successor first) that is a maximum of the exact order, so its result
        // has the best key: when no two candidates share that key (no tie
        // against the best so far), the best by key is the result; otherwise
        // the scan runs over just the candidates with the best key.
        int best = cur, kbest = kcur;
        bool exact = false;
        {
            int e = live ? eb : 0;
            const int e1 = live ? ee : 0;
            bool first = cur >= 0;
            while (e < e1) {
                const int cnt = min(kSwBatch, e1 - e);
                int us[kSwBatch], ks[kSwBatch];
#pragma unroll
                for (int k = 0; k < kSwBatch; k++) us[k] = k < cnt ? tgt[e + k] : -1;
                if (only_changed) {                   // the current successor stays: its key is needed
#pragma unroll
                    for (int k = 0; k < kSwBatch; k++)
                        if (us[k] >= 0 && us[k] != cur && !((cb[us[k] >> 6] >> (us[k] & 63)) & 1))
                            us[k] = -2;
                }
#pragma unroll
                for (int k = 0; k < kSwBatch; k++) ks[k] = us[k] >= 0 ? key[us[k]] : 0;
                if (first) {
                    bool found = false;
#pragma unroll
                    for (int k = 0; k < kSwBatch; k++)
                        if (us[k] == cur) {
                            kcur = ks[k];
                            found = true;
                        }
                    if (!found) kcur = key[cur];
                    kbest = kcur;
                    first = false;
                }
#pragma unroll
                for (int k = 0; k < kSwBatch; k++) {
                    if (us[k] < 0) continue;
                    bool tie;
                    const int c = cmp_fast(us[k], ks[k], best, kbest, &tie);
                    if (player == 0 ? c > 0 : c < 0) {
                        best = us[k];
                        kbest = ks[k];
                        exact = false;
                    } else if (tie) {
                        exact = true;
                    }
                }
                e += kSwBatch;
            }
        }
        int e = live && exact ? eb : 0;
        const int e1 = live && exact ? ee : 0;
        while (__any(e < e1)) {
            const int cnt = e < e1 ? min(kSwBatch, e1 - e) : 0;
            int us[kSwBatch], ks[kSwBatch];
#pragma unroll
            for (int k = 0; k < kSwBatch; k++) us[k] = k < cnt ? tgt[e + k] : -1;
            if (only_changed) {                   // -2: unchanged, skipped
#pragma unroll
                for (int k = 0; k < kSwBatch; k++)
                    if (us[k] >= 0 && !((cb[us[k] >> 6] >> (us[k] & 63)) & 1)) us[k] = -2;
            }
#pragma unroll
            for (int k = 0; k < kSwBatch; k++) ks[k] = us[k] >= 0 ? key[us[k]] : 0;
#pragma unroll
            for (int k = 0; k < kSwBatch; k++) {  // -3: worse key than the best, skipped
                bool tie;
                const int c = us[k] >= 0 ? cmp_fast(us[k], ks[k], best, kbest, &tie) : 0;
                if (player == 0 ? c < 0 : c > 0) us[k] = -3;
            }
            // One compare at a time (a rolled loop, so the exact-compare code
            // exists once; k is uniform, so us[k] and ks[k] are register reads).
#pragma unroll 1
            for (int k = 0; k < kSwBatch && __any(k < cnt); k++) {
                const int u = us[k], ku = ks[k];
                int c = 0, la = 0, lb = 0;
                bool tie = false;
                if (u >= 0) {
                    c = cmp_fast(u, ku, cur, kcur, &tie);
                    if (tie) {
                        if (ties) atomicAdd(ties, 1ull);
                        c = cmp_lca(u, cur, dep, Sl, vinfo, N, F, &la, &lb);
                    }
                }
                unsigned long long need = __ballot(c == 2);
                while (need) {
                    const int L = __ffsll((long long)need) - 1;
                    const int r = coop_scan(__shfl(u, L), __shfl(la, L), __shfl(cur, L),
                                            __shfl(lb, L), par, vinfo, Sl, N, F, ties);
                    if ((int)(threadIdx.x & (kWave - 1)) == L) c = r;
                    need &= need - 1;
                }
                if (u >= 0 && (player == 0 ? c > 0 : c < 0)) {
                    cur = u;
                    kcur = ku;
                    sw = true;
                }
            }
            e += kSwBatch;
        }
        if (live && !exact) cur = best;
        sw = live && cur != c0;
        if (sw) strat[v] = cur;
        // one store per wave and game: the flag only has to become nonzero
        unsigned long long m = __ballot(sw);
        if (m) {
            int g0 = __shfl(g, __ffsll((long long)m) - 1);
            bool uni = __all(!sw || g == g0);
            if (uni) {
                if ((threadIdx.x & (kWave - 1)) == (unsigned)(__ffsll((long long)m) - 1)) {
                    if (loop_form) {
                        if (player == 1) gc[g0].sw1 = 1; else gc[g0].sw0 = 1;
                    } else {
                        *flag = 1;
                    }
                }
            } else if (sw) {
                if (loop_form) {
                    if (player == 1) gc[g].sw1 = 1; else gc[g].sw0 = 1;
                } else {
                    *flag = 1;
                }
            }
        }
    }
}

__global__ void __launch_bounds__(kBS) k_switch(
    int player, long long N, const int* __restrict__ own, long long own_n,
    const int* __restrict__ vinfo, const int* __restrict__ off,
    const int* __restrict__ tgt, const int* __restrict__ gid, GameCtl* __restrict__ gc,
    int* __restrict__ strat, const int* __restrict__ key, const int* __restrict__ dep,
    const int* __restrict__ par, const int2* __restrict__ Sl, const JumpCnt* __restrict__ jc,
    int loop_form, int* __restrict__ flag, unsigned long long* __restrict__ ties,
    const int* __restrict__ hi, const unsigned long long* __restrict__ chg,
    const unsigned long long* __restrict__ chg0)
{
    switch_body(player, N, own, own_n, vinfo, off, tgt, gid, gc, strat, key, dep, par, Sl, jc,
                loop_form, flag, ties, hi, chg, chg0, nullptr, 0);
}

// mark_solved(0) of the games that take the outer step (fused form of
// cpu_parity.cpp:265-274 as solve_si calls it, gpumain.cpp:44).
// solve_bf games after a relaxation also record succ: the chosen successor of
// every unsolved Odd vertex with a finite valuation (cpu_bf.cpp:120-147).
__device__ __forceinline__ void mark0_body(long long N, const int* __restrict__ gid,
                                           const GameCtl* __restrict__ gc,
                                           const int8_t* __restrict__ lab, int8_t* __restrict__ solved,
                                           const int* __restrict__ hi, const int* __restrict__ vinfo,
                                           const int* __restrict__ strat, int* __restrict__ succ)
{
    const long long lim = min(N, (long long)*hi);
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < lim; v += stride) {
        int g = gid[v];
        if (gc[g].done || !gc[g].outer) continue;
        const int l = lab[v];
        if (succ && gc[g].bf == 1 && gc[g].bf_init && l == 0 && !solved[v] && owner_of(vinfo[v]) == 1)
            succ[v] = strat[v];
        if (l == 1) solved[v] = 1;
    }
}

__global__ void k_mark0_outer(long long N, const int* __restrict__ gid, const GameCtl* __restrict__ gc,
                              const int8_t* __restrict__ lab, int8_t* __restrict__ solved,
                              const int* __restrict__ hi, const int* __restrict__ vinfo,
                              const int* __restrict__ strat, int* __restrict__ succ)
{
    mark0_body(N, gid, gc, lab, solved, hi, vinfo, strat, succ);
}

// succ starts at the game's vertex 0 (cpu_bf.cpp:17).
__global__ void k_succ_init(long long N, const int* __restrict__ gid, const GameCtl* __restrict__ gc,
                            int* __restrict__ succ)
{
    long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < N) succ[v] = gc[gid[v]].vbase;
}

// End of solve_bf: Odd's strategy is succ once a relaxation has run
// (cpu_bf.cpp:161-164; solve_bf's first best response is an SI inner loop).
__global__ void k_bf_final(long long N, const int* __restrict__ gid, const GameCtl* __restrict__ gc,
                           const int* __restrict__ vinfo, const int* __restrict__ succ,
                           int* __restrict__ strat)
{
    long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const int g = gid[v];
    if (gc[g].bf && gc[g].iter >= 2 && !gc[g].err && owner_of(vinfo[v]) == 1) strat[v] = succ[v];
}

// mark_solved(player) as a separate step (cpu_parity.cpp:265-274).
__global__ void k_mark_solved(int player, long long N, const int8_t* __restrict__ lab,
                              int8_t* __restrict__ solved)
{
    long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    if ((player == 0 && lab[v] == 1) || (player == 1 && lab[v] == -1)) solved[v] = 1;
}

// After switch_strategy(1): count the round; the inner loop ends when no
// player-1 vertex switched (gpumain.cpp:34-38).  A game that does not take the
// outer step is checked against its round cap here (switch(0) and its control
// leave it unchanged).  pg (the step program, or null): games taking the outer
// step are flagged, finished games counted down.
__device__ __forceinline__ void ctl1_body(int G, GameCtl* __restrict__ gc, StepProg* __restrict__ pg)
{
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < G; g += gridDim.x * blockDim.x) {
        GameCtl c = gc[g];
        if (c.done) continue;
        if (c.err) {               // solve_bf: odd strategy cycle (the reference exits)
            c.done = 1;
            gc[g] = c;
            if (pg) atomicSub(&pg->live, 1);
            continue;
        }
        c.br++;
        c.outer = c.sw1 == 0;
        c.sw1 = 0;
        c.reset_pending = 0;
        c.clear0 = 0;
        if (!c.outer && c.max_rounds > 0 && c.br >= c.max_rounds) {
            c.capped = 1;
            c.done = 1;
            if (pg) atomicSub(&pg->live, 1);
        }
        if (pg && c.outer) atomicOr(&pg->any_outer, 1);
        gc[g] = c;
    }
}

__global__ void k_ctl1(int G, GameCtl* gc) { ctl1_body(G, gc, nullptr); }

// After switch_strategy(0) (gpumain.cpp:44-50).
__device__ __forceinline__ void ctl0_body(int G, GameCtl* __restrict__ gc, StepProg* __restrict__ pg)
{
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < G; g += gridDim.x * blockDim.x) {
        GameCtl c = gc[g];
        if (c.done) continue;
        if (c.outer) {
            c.iter++;
            c.clear0 = 1;          // the next valuation restarts the switch(0) record
            c.bf_init = 1;         // solve_bf: later best responses relax
            if (c.sw0 == 0) c.done = 1;
            else if (c.reset) c.reset_pending = 1;
        }
        c.sw0 = 0;
        c.outer = 0;
        if (!c.done && c.max_rounds > 0 && c.br >= c.max_rounds) {
            c.capped = 1;
            c.done = 1;
        }
        if (pg && c.done) atomicSub(&pg->live, 1);
        gc[g] = c;
    }
    // the switch(0) pass before this launch consumed its marks
    if (pg && blockIdx.x == 0 && threadIdx.x == 0) pg->full0 = 0;
}

__global__ void k_ctl0(int G, GameCtl* gc) { ctl0_body(G, gc, nullptr); }

// ---------------------------------------------------------------------------
// Step program: the device-resident loop as identical launches of k_phase.
// Launch i reads the phase launch i-1 ran and runs the next one:
//   INIT (val_init) -> JUMP 0 -> JUMP 2 -> ... (rounds r and r + 1 per
//   launch, while the decision runs round r) -> LABEL -> SW1 -> CTL1 ->
//   [MARK0+SW0 -> CTL0, only when some
//   game takes the outer step] -> INIT ...
// so no launch is spent on a round after convergence or on an outer-step
// phase that no game takes; a graph of these launches can stop anywhere and
// the next graph continues.  Once no game is live every launch exits at once.
enum Phase {
    PH_START = 0, PH_INIT, PH_JUMP, PH_LABEL, PH_SW1, PH_CTL1, PH_MARK0, PH_SW0, PH_CTL0, PH_DONE, PH_ACT,
    PH_PUSH
};

// One game, after the labels, when at most push_max vertices changed their
// valuation (the step program runs this phase only then): mark the
// owner-list entries of their in-neighbours (plain byte stores of 1: writers
// of one mark agree) for the next switch(1) (player 1) and switch(0) (player
// 0, kept until that pass).  After a valuation with more changes, the next
// switch(1) and switch(0) scan every entry instead.
__device__ __forceinline__ void push_body(long long N, const unsigned long long* __restrict__ chg,
                                          const int* __restrict__ rin_off,
                                          const int* __restrict__ rin_src, uint8_t* __restrict__ d0,
                                          uint8_t* __restrict__ d1)
{
    const long long nw = (N + 63) / 64;
    for (long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x; w < nw;
         w += (long long)gridDim.x * blockDim.x) {
        for (unsigned long long x = chg[w]; x; x &= x - 1) {
            const int u = (int)(w * 64 + __ffsll((long long)x) - 1);
            const int e1 = rin_off[u + 1];
            for (int e = rin_off[u]; e < e1; e++) {
                const int c = rin_src[e];
                (c & 1 ? d1 : d0)[c >> 1] = 1;
            }
        }
    }
}

// Serving pool: games the host admitted (done, activate set) become live,
// at a step boundary; the occupied bound grows to cover them.
__device__ __forceinline__ void act_body(int G, GameCtl* __restrict__ gc, StepProg* __restrict__ pg,
                                         int* __restrict__ hi)
{
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < G; g += gridDim.x * blockDim.x) {
        if (!gc[g].activate) continue;
        gc[g].activate = 0;
        gc[g].done = 0;
        atomicAdd(&pg->live, 1);
        atomicMax(hi, gc[g].vbase + gc[g].n + 1);
    }
}

struct StepArgs {
    long long N, nb, own_n0, own_n1;
    int G, maxn, lmax, trace, trace_cap;
    int two_levels;  // jump launches advance two levels (else one)
    const int *vinfo, *off, *tgt, *gid, *own0, *own1, *hi;
    const int *rin_off, *rin_src;
    uint8_t *dm0, *dm1;   // change marks of the owner lists
    long long push_max;   // 0: no marks (PH_PUSH never runs)
    uint8_t *d1, *l12ok;  // reuse of level-1/2 records (with push_max; else null)
    uint8_t* bchg;        // blocks whose records ended with the change bit (with push_max)
    GameCtl* gc;
    int *strat, *par, *dep, *key, *flag, *succ;
    int8_t *solved, *lab;
    uint8_t *blive, *bact, *blast;
    int2* Sl;
    unsigned long long *Zl, *dig, *tr, *chg, *chg0;
    JumpCnt* jc;
};

// A pointer field of the arguments, read as a pointer to global memory:
// loads and stores through it (after inlining, also through what is derived
// from it) are then global instructions; a pointer read from memory is
// otherwise generic, and flat instructions need 64-bit vector addresses and
// wait on two counters.
template <class T> __device__ __forceinline__ T* glb(T* const& field)
{
    using GlobalPtr = __attribute__((address_space(1))) T*;
    return (T*)__builtin_bit_cast(GlobalPtr, field);
}

// the head of StepProg as 32-bit words: st[2], any_outer, live, host_seq,
// dev_seq[2], (padding), psum[2], pushed, bf0
constexpr int kProgHeadWords = 14;
static_assert(offsetof(StepProg, any_outer) == 8 && offsetof(StepProg, live) == 12 &&
                  offsetof(StepProg, host_seq) == 16 && offsetof(StepProg, dev_seq) == 20 &&
                  offsetof(StepProg, psum) == 32 && offsetof(StepProg, pushed) == 48 &&
                  offsetof(StepProg, bf0) == 52 && offsetof(StepProg, pslot) == 4 * kProgHeadWords,
              "StepProg head layout");

// The arguments live in device memory (one pointer argument): each phase
// loads the fields it uses, which keeps scalar registers free.
__global__ void __launch_bounds__(kBS) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_phase(int i, const StepArgs* __restrict__ ap, StepProg* __restrict__ pg)
{
    const StepArgs& a = *ap;
    __shared__ int ph_s, r_s, dbl_s, dpar_s, marked_s, reuse_s;
    __shared__ unsigned int cnt[3];
    if (threadIdx.x < kWave) {                    // the first wave decides
        const int lane = threadIdx.x;
        // Every word launch i-1 left (phase, round summary, control words)
        // and its slot row, in two vector loads issued together: lane k < 14
        // holds word k of the state's head, every lane its slot word.
        const int hw = reinterpret_cast<const int*>(pg)[min(lane, kProgHeadWords - 1)];
        const unsigned long long x = pg->pslot[(i + 3) & 3][lane];
        const unsigned mi = pg->pmis[(i + 3) & 3][lane];
        const int maxn = a.maxn, lmax = a.lmax, two = a.two_levels;
        const long long push_max = a.push_max;
        // the last valuation's changed-vertex count, one slot per lane
        // (written by its label launch, zeroed by the next val_init)
        const unsigned cc = pg->ccnt[lane * kCntStride];
        // all these loads issued before any is waited for
        asm volatile("" ::"v"(hw), "v"(x), "v"(mi), "v"(cc), "s"(maxn), "s"(lmax), "s"(two));
        const int p = (i + 1) & 1;
        // lane k's word as a scalar (the same in every lane)
        auto head = [&](int k) { return __builtin_amdgcn_readlane(hw, k); };
        const int prev = head(p);                                         // st[p]
        const int any_outer = head(2), live = head(3);
        const int hseq = head(4), dseq = head(5 + p);                     // host_seq, dev_seq[p]
        const unsigned long long sm =
            ((unsigned long long)(unsigned)head(9 + 2 * p) << 32) | (unsigned)head(8 + 2 * p);   // psum[p]
        // st = phase | bact parity offset << 7 | round << 8 (an INIT's
        // "round" says that it marked the changed links in d1)
        int ph = prev & 0x7f, r = prev >> 8, dbl = 0, dpar = (prev >> 7) & 1;
        const int reuse = ph == PH_INIT && (r & 1);   // this launch is that valuation's round 0
        switch (ph) {
        case PH_INIT: ph = PH_JUMP; r = 0; dpar = 0; break;
        case PH_JUMP: r++; break;
        case PH_LABEL: {
            // marks for the switches when few valuations changed (every
            // workgroup sums the same slots, so all agree)
            unsigned t = cc;
            for (int o = kSlots / 2; o; o >>= 1) t += __shfl_xor(t, o);
            ph = push_max && (long long)t <= push_max ? PH_PUSH : PH_SW1;
            break;
        }
        case PH_PUSH: ph = PH_SW1; break;
        case PH_SW1: ph = PH_CTL1; break;
        case PH_CTL1: ph = any_outer ? PH_MARK0 : PH_INIT; break;
        case PH_MARK0: ph = PH_CTL0; break;   // switch(0) ran in the same launch
        case PH_SW0: ph = PH_CTL0; break;
        case PH_DONE: break;
        default: ph = PH_INIT; break;             // PH_START, PH_CTL0, PH_ACT
        }
        // at a step boundary: admitted games first (a serving pool), then
        // the next step if any game is live
        if ((ph == PH_INIT || ph == PH_DONE) && hseq != dseq) ph = PH_ACT;
        if (ph == PH_INIT) {
            r = 0;
            if (live <= 0) ph = PH_DONE;
            // after a switch(1) on change marks (few valuations changed) and
            // no outer step: mark the changed links for the first round's
            // record reuse (one game, plain semantics, no trace)
            if (ph == PH_INIT && a.d1 && !a.trace && (prev & 0x7f) == PH_CTL1 && head(12) && head(13) == 0)
                r = 1;
        }
        if (ph == PH_JUMP) {
            // round r's decision from round r-1 (launch i-1); a round at lmax
            // would mean the valuation needs more rounds than it has levels
            int stop, cyc;
            unsigned mo = mi;                     // launch i-1's mismatch flags, any slot
            for (int o = kSlots / 2; o; o >>= 1) mo |= __shfl_xor(mo, o);
            int run = decide_wave(r, r >= 1 ? sm : 0ull, x, mo, maxn, &stop, &cyc);
            if (run && r >= lmax) {
                if (blockIdx.x == 0 && lane == 0) a.jc->err = 1;
                run = 0;
            }
            // Rounds r and r + 1 in this launch (jump2_body) while level
            // r + 2 exists; the next launch then takes round r + 2.
            dbl = run && two && r + 2 <= lmax;
            if (blockIdx.x == 0 && lane == 0) {
                pg->psum[i & 1] = summary_word(run, stop, cyc, dbl);
                if (run) a.jc->last = r + dbl;
            }
            if (!run) ph = PH_LABEL;
        }
        if (blockIdx.x == 0) {
            pg->pslot[(i + 2) & 3][lane] = 0;     // launch i+2 counts its round there
            pg->pmis[(i + 2) & 3][lane] = 0;
            if (lane == 0) {
                pg->st[i & 1] = ph | ((dpar ^ dbl) << 7) | ((r + dbl) << 8);
                if (ph == PH_SW1) {
                    pg->any_outer = 0;                 // set again by CTL1
                    pg->pushed = (prev & 0x7f) == PH_PUSH;
                }
                pg->dev_seq[i & 1] = ph == PH_ACT ? hseq : dseq;
            }
        }
        if (lane == 0) {
            ph_s = ph;
            r_s = r;
            dbl_s = dbl;
            dpar_s = dpar;
            marked_s = (prev & 0x7f) == PH_PUSH;  // this switch(1) has marks
            cnt[0] = cnt[1] = cnt[2] = 0;
            reuse_s = reuse;
        }
    }
    __syncthreads();
    const int ph = ph_s, r = r_s;
    // one game: every vertex is in game 0 and the bodies skip the gid loads
    const int* gid1 = a.G == 1 ? nullptr : glb(a.gid);
    switch (ph) {
    case PH_INIT:
        // the coming valuation counts its changes afresh (every reader of the
        // last count ran before this launch)
        if (a.push_max && blockIdx.x == 0 && threadIdx.x < kSlots) pg->ccnt[threadIdx.x * kCntStride] = 0;
        val_init_body(a.N, glb(a.vinfo), glb(a.off), glb(a.tgt), gid1, glb(a.gc), glb(a.strat),
                      glb(a.solved), glb(a.lab), glb(a.par), glb(a.Sl), glb(a.Zl), glb(a.jc),
                      glb(a.blive), glb(a.bact), glb(a.blast), a.trace, glb(a.hi),
                      r ? glb(a.d1) : nullptr, glb(a.rin_off), glb(a.rin_src), glb(a.own0), glb(a.own1),
                      a.bchg ? glb(a.bchg) : nullptr);
        break;
    case PH_JUMP: {
        // block flags of round r: after an odd number of two-round launches
        // in this valuation, the parity of the flag buffers is offset by one
        const int pin = (r + dpar_s) & 1;
        if (dbl_s)
            jump2_body(r, a.N, glb(a.Sl), glb(a.Zl), pg->pslot[i & 3], glb(a.blive),
                       glb(a.bact) + pin * a.nb, glb(a.bact) + (pin ^ 1) * a.nb, glb(a.blast),
                       a.trace, glb(a.hi), cnt, pg->pmis[i & 3], reuse_s ? glb(a.d1) : nullptr,
                       a.l12ok && !a.trace ? glb(a.l12ok) : nullptr, reuse_s ? d1_tag(glb(a.gc)[0].br) : 0,
                       a.bchg ? glb(a.bchg) : nullptr);
        else
            jump_body(r, a.N, glb(a.Sl), glb(a.Zl), pg->pslot[i & 3], glb(a.blive),
                      glb(a.bact) + pin * a.nb, glb(a.bact) + (pin ^ 1) * a.nb, glb(a.blast),
                      a.trace, glb(a.hi), cnt, a.bchg ? glb(a.bchg) : nullptr);
        break;
    }
    case PH_LABEL: {
        // quiet blocks: one game with change flags, plain semantics, not its
        // first valuation, no trace
        const GameCtl* g0 = glb(a.gc);
        const bool quiet = a.bchg && !a.trace && a.G == 1 && g0->bf == 0 && g0->br > 0;
        label_body(a.N, a.maxn, a.lmax, glb(a.Sl), glb(a.Zl), glb(a.blast), gid1, glb(a.gc),
                   glb(a.strat), glb(a.solved), glb(a.lab), glb(a.dep), glb(a.key), glb(a.dig),
                   glb(a.tr), a.trace_cap, 1, a.trace, glb(a.jc), glb(a.blive), glb(a.hi), glb(a.chg),
                   glb(a.chg0), 0, a.push_max ? pg->ccnt : nullptr,
                   quiet ? glb(a.bchg) : nullptr, glb(a.bact) + ((r + dpar_s) & 1) * a.nb);
        break;
    }
    case PH_PUSH:
        push_body(a.N, glb(a.chg), glb(a.rin_off), glb(a.rin_src), glb(a.dm0), glb(a.dm1));
        break;
    case PH_MARK0:
        // mark_solved(0), then switch(0) in the same launch: the switch reads
        // nothing the marking writes (solved flags; succ of Odd's vertices,
        // whose strategies switch(0) leaves alone), and the marking reads no
        // strategy switch(0) writes
        mark0_body(a.N, glb(a.gid), glb(a.gc), glb(a.lab), glb(a.solved), glb(a.hi), glb(a.vinfo),
                   glb(a.strat), a.succ ? glb(a.succ) : nullptr);
        [[fallthrough]];
    case PH_SW1:
    case PH_SW0: {                                // one inlined copy of the switch
        const int pl = ph == PH_SW1 ? 1 : 0;
        // change marks (one game): only marked entries unless a valuation
        // since this player's last pass set none, or player 1's strategy
        // was reset (then every entry is scanned)
        uint8_t* dq = nullptr;
        int use_d = 0;
        if (a.push_max) {
            dq = pl ? glb(a.dm1) : glb(a.dm0);
            use_d = (pl ? marked_s : !pg->full0) && !(pl && glb(a.gc)[0].reset_pending);
            // a valuation without marks: switch(0) scans everything at its next pass
            if (pl && !marked_s && blockIdx.x == 0 && threadIdx.x == 0) pg->full0 = 1;
        }
        switch_body(pl, a.N, pl ? glb(a.own1) : glb(a.own0), pl ? a.own_n1 : a.own_n0, glb(a.vinfo),
                    glb(a.off), glb(a.tgt), gid1, glb(a.gc), glb(a.strat), glb(a.key),
                    glb(a.dep), glb(a.par), glb(a.Sl), glb(a.jc), 1, glb(a.flag), nullptr, glb(a.hi),
                    glb(a.chg), glb(a.chg0), dq, use_d);
        break;
    }
    case PH_CTL1: ctl1_body(a.G, glb(a.gc), pg); break;
    case PH_CTL0: ctl0_body(a.G, glb(a.gc), pg); break;
    case PH_ACT: act_body(a.G, glb(a.gc), pg, glb(a.flag) + 2); break;
    default: break;
    }
}

// Step program start: the next launch runs INIT; live = unfinished games.
__global__ void k_prog_reset(int G, const GameCtl* __restrict__ gc, StepProg* __restrict__ pg,
                             uint8_t* __restrict__ l12ok = nullptr, long long nb = 0)
{
    // no block's level-1 and level-2 records are taken as the last valuation's
    for (long long b = threadIdx.x; l12ok && b < nb; b += blockDim.x) l12ok[b] = 0;
    __shared__ int acc;
    if (threadIdx.x == 0) acc = 0;
    __syncthreads();
    int live = 0;
    for (int g = threadIdx.x; g < G; g += blockDim.x) live += gc[g].done ? 0 : 1;
    atomicAdd(&acc, live);
    __syncthreads();
    if (threadIdx.x == 0) {
        pg->st[0] = PH_START;
        pg->st[1] = PH_START;
        pg->host_seq = pg->dev_seq[0] = pg->dev_seq[1] = 0;
        pg->psum[0] = pg->psum[1] = 0;
        pg->any_outer = 0;
        pg->live = acc;
        pg->full0 = 1;             // marks start over: the first passes scan everything
        pg->pushed = 0;
        pg->bf0 = G >= 1 ? gc[0].bf : 0;
    }
    for (int k = threadIdx.x; k < 4 * kSlots; k += blockDim.x) {
        pg->pslot[k / kSlots][k % kSlots] = 0;
        pg->pmis[k / kSlots][k % kSlots] = 0;
    }
    for (int k = threadIdx.x; k < kSlots; k += blockDim.x) pg->ccnt[k * kCntStride] = 0;
}

__global__ void k_count_live(int G, const GameCtl* gc, int* out)
{
    __shared__ int acc;
    if (threadIdx.x == 0) acc = 0;
    __syncthreads();
    int live = 0;
    for (int g = threadIdx.x; g < G; g += blockDim.x) live += gc[g].done ? 0 : 1;
    atomicAdd(&acc, live);
    __syncthreads();
    if (threadIdx.x == 0) *out = acc;
}

// Serving pool: one past the last vertex of an unfinished game, so that the
// step's full-range kernels stop there (flag[2]).
__global__ void k_pool_hi(int G, const GameCtl* __restrict__ gc, int* __restrict__ hi)
{
    __shared__ int acc;
    if (threadIdx.x == 0) acc = 0;
    __syncthreads();
    int h = 0;
    for (int g = threadIdx.x; g < G; g += blockDim.x)
        if (!gc[g].done) h = max(h, gc[g].vbase + gc[g].n + 1);
    atomicMax(&acc, h);
    __syncthreads();
    if (threadIdx.x == 0) *hi = acc;
}

template <class T> __global__ void k_fill(T* p, long long n, T v)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Host side.

template <class T> static int dalloc(T** p, long long n)
{
    std::lock_guard<std::recursive_mutex> lk(setup_mutex());
    PSI_HIP(hipMalloc((void**)p, (size_t)std::max(n, 1LL) * sizeof(T)));
    return PSI_OK;
}

const char* core_validate_game(const GameView& gv, int* max_pri)
{
    const int n = gv.n;
    const long long m = gv.m;
    if (n <= 0) return "game with no vertices";
    if (n >= (1 << 28)) return "game too large (at most 2^28 - 1 vertices)";
    if (gv.off[0] != 0 || (long long)gv.off[n] != m)
        return "offsets must start at 0 and end at the edge count";
    int mp = 0;
    for (int v = 0; v < n; v++) {
        if (gv.off[v + 1] <= gv.off[v]) return "sinks are not allowed!";
        int o = gv.owner[v], p = gv.pri[v];
        if (o != 0 && o != 1) return "invalid player";
        if (p < 0 || p > kPriMask) return "priority out of range [0, 2^30)";
        mp = std::max(mp, p);
    }
    for (long long e = 0; e < m; e++)
        if ((unsigned)gv.tgt[e] >= (unsigned)n) return "successor out of range";
    *max_pri = mp;
    return nullptr;
}

// PSI_TIMING=1 in the environment prints the phases of solver creation.
struct PhaseTimer {
    bool on;
    std::chrono::steady_clock::time_point t0;
    PhaseTimer() : on(getenv("PSI_TIMING") != nullptr), t0(std::chrono::steady_clock::now()) {}
    void mark(const char* what)
    {
        if (!on) return;
        auto t1 = std::chrono::steady_clock::now();
        fprintf(stderr, "[psi] %s %.3f s\n", what, std::chrono::duration<double>(t1 - t0).count());
        t0 = t1;
    }
};

// ---- device-side packing of a batch ----------------------------------------
__device__ __forceinline__ int game_of(const long long* __restrict__ base, int G, long long x)
{
    int lo = 0, hi = G - 1;                       // last g with base[g] <= x
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (base[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// owner/priority were copied into own_tmp/pri_tmp; offsets (local) into off.
__global__ void k_pack_vertices(long long N, int G, const long long* __restrict__ vbase,
                                const long long* __restrict__ ebase, const int* __restrict__ own_tmp,
                                const int* __restrict__ pri_tmp, int* __restrict__ vinfo,
                                int* __restrict__ off, int* __restrict__ gid)
{
    long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    int g = game_of(vbase, G, v);
    gid[v] = g;
    vinfo[v] = pri_tmp[v] | (own_tmp[v] << 30);
    off[v] = (int)(off[v] + ebase[g]);
}

__global__ void k_globalize_targets(long long N, const int* __restrict__ off, const int* __restrict__ gid,
                                    const GameCtl* __restrict__ gc, int* __restrict__ tgt)
{
    long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    int vb = gc[gid[v]].vbase;
    for (int e = off[v]; e < off[v + 1]; e++) tgt[e] += vb;
}

// Owner lists (vertex ids of each player, ascending) by a block-count /
// scan / scatter compaction.
__global__ void k_owner_count(long long N, const int* __restrict__ vinfo, int* __restrict__ cnt1)
{
    long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int o = v < N ? owner_of(vinfo[v]) : 0;
    int c = __syncthreads_count(o);
    if (threadIdx.x == 0) cnt1[blockIdx.x] = c;
}

__global__ void __launch_bounds__(1024) k_owner_scan(int nb, int* __restrict__ cnt1)
{
    // exclusive scan of nb block counts in place, one workgroup
    __shared__ int part[1024];
    const int per = (nb + 1023) / 1024;
    const int t = threadIdx.x, lo = t * per, hi = min(nb, lo + per);
    int sum = 0;
    for (int i = lo; i < hi; i++) sum += cnt1[i];
    part[t] = sum;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        int x = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += x;
        __syncthreads();
    }
    int run = part[t] - sum;
    for (int i = lo; i < hi; i++) {
        int c = cnt1[i];
        cnt1[i] = run;
        run += c;
    }
}

__global__ void k_owner_scatter(long long N, const int* __restrict__ vinfo, const int* __restrict__ pre1,
                                int* __restrict__ own0, int* __restrict__ own1)
{
    __shared__ int wsum[kBS / kWave];
    long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int o = v < N ? owner_of(vinfo[v]) : 0;
    unsigned long long m = __ballot(o);
    int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    if (lane == 0) wsum[w] = __popcll(m);
    __syncthreads();
    int before = 0;
    for (int k = 0; k < w; k++) before += wsum[k];
    int in_wave = __popcll(m & ((1ull << lane) - 1ull));
    int r1 = pre1[blockIdx.x] + before + in_wave;            // ones before v
    if (v < N) {
        if (o) own1[r1] = (int)v;
        else own0[v - r1] = (int)v;
    }
}

// In-edge lists of one game (the switch's change marks): in-degree counts,
// an exclusive scan, then each edge placed through an atomic cursor per
// target (the order within a list does not matter); an in-edge is stored as
// its source's owner-list entry, index << 1 | owner.
__global__ void k_rin_count(long long N, const int* __restrict__ off, const int* __restrict__ tgt,
                            int* __restrict__ cnt)
{
    long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    for (int e = off[v]; e < off[v + 1]; e++) atomicAdd(&cnt[tgt[e]], 1);
}

__global__ void k_owner_index(long long n, const int* __restrict__ own, int* __restrict__ idx)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) idx[own[i]] = (int)i;
}

__global__ void k_rin_fill(long long N, const int* __restrict__ vinfo, const int* __restrict__ off,
                           const int* __restrict__ tgt, const int* __restrict__ idx,
                           int* __restrict__ cur, int* __restrict__ src)
{
    long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const int code = (idx[v] << 1) | owner_of(vinfo[v]);
    for (int e = off[v]; e < off[v + 1]; e++) src[atomicAdd(&cur[tgt[e]], 1)] = code;
}

// In-edge lists and change marks of a one-game core; on failure (memory)
// the core runs without marks.
static void build_in_edges(Core& c, hipStream_t s)
{
    const long long N = c.N, M = c.M;
    if (N <= 0 || N >= (1LL << 30)) return;
    // One allocation for the lists, the marks and the build's scratch (kept
    // until the handle is destroyed: no free, hence no wait, at creation).
    size_t tmp_bytes = 0;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, (int*)nullptr, (int*)nullptr, (int)(N + 1), s) !=
        hipSuccess)
        return;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t b_off = al((N + 1) * 4), b_src = al(std::max(M, 1LL) * 4);
    const size_t b_d0 = al(c.own_n[0] + 64), b_d1 = al(c.own_n[1] + 64);
    const size_t b_cnt = al((N + 1) * 4), b_idx = al(N * 4), b_tmp = al(std::max<size_t>(tmp_bytes, 1));
    const size_t b_m1 = al(N), b_ok = al(nblocks(N, kBS));   // d1 marks, l12ok flags
    const size_t b_bc = al(nblocks(N, kBS));                  // bchg flags
    if (dalloc(&c.rin_mem, (long long)(b_off + b_src + b_d0 + b_d1 + b_m1 + b_ok + b_bc + b_cnt + b_idx + b_tmp)) !=
        PSI_OK) {
        (void)hipGetLastError();
        c.rin_mem = nullptr;
        return;
    }
    unsigned char* q = c.rin_mem;
    c.rin_off = (int*)q;            q += b_off;
    c.rin_src = (int*)q;            q += b_src;
    c.dmark[0] = q;                 q += b_d0;
    c.dmark[1] = q;                 q += b_d1;
    c.d1 = q;                       q += b_m1;
    c.l12ok = q;                    q += b_ok;
    c.bchg = q;                     q += b_bc;
    int* cnt = (int*)q;             q += b_cnt;
    int* idx = (int*)q;             q += b_idx;
    unsigned char* tmp = q;
    const unsigned nb = nblocks(N, kBS);
    bool ok = hipMemsetAsync(cnt, 0, (N + 1) * 4, s) == hipSuccess &&
              hipMemsetAsync(c.dmark[0], 0, b_d0 + b_d1 + b_m1 + b_ok + b_bc, s) == hipSuccess;
    k_rin_count<<<nb, kBS, 0, s>>>(N, c.off, c.tgt, cnt);
    ok = ok && hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, cnt, c.rin_off, (int)(N + 1), s) == hipSuccess;
    ok = ok && hipMemcpyAsync(cnt, c.rin_off, N * 4, hipMemcpyDeviceToDevice, s) == hipSuccess;
    for (int p = 0; p < 2; p++)
        if (c.own_n[p] > 0) k_owner_index<<<nblocks(c.own_n[p], kBS), kBS, 0, s>>>(c.own_n[p], c.own[p], idx);
    k_rin_fill<<<nb, kBS, 0, s>>>(N, c.vinfo, c.off, c.tgt, idx, cnt, c.rin_src);
    ok = ok && hipGetLastError() == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        (void)hipFree(c.rin_mem);
        c.rin_mem = nullptr;
        c.rin_off = c.rin_src = nullptr;
        c.dmark[0] = c.dmark[1] = nullptr;
        c.d1 = c.l12ok = c.bchg = nullptr;
        return;
    }
    c.push_ok = true;
}

int core_create(Core& c, int G, const GameView* games, bool trace, int trace_cap, hipStream_t s)
{
    PhaseTimer pt;
    if (G <= 0) return fail(PSI_ERR_INVALID, "no games");
    std::vector<long long> vbase(G + 1, 0), ebase(G + 1, 0);
    int maxn = 0;
    for (int g = 0; g < G; g++) {
        vbase[g + 1] = vbase[g] + std::max(games[g].n, 0);
        ebase[g + 1] = ebase[g] + std::max(games[g].m, 0LL);
        maxn = std::max(maxn, games[g].n);
    }
    const long long N = vbase[G], M = ebase[G];
    if (N >= (1LL << 31) - 1 || M >= (1LL << 31) - 1)
        return fail(PSI_ERR_INVALID, "total vertices and edges must fit in int32");
    std::vector<const char*> errs(G, nullptr);
    std::vector<int> mp(G, 0);
    parallel_games(G, [&](int g) { errs[g] = core_validate_game(games[g], &mp[g]); });
    for (int g = 0; g < G; g++)
        if (errs[g]) return fail(PSI_ERR_INVALID, errs[g]);
    pt.mark("validate");
    c.G = G;
    c.N = N;
    c.M = M;
    c.s = s;
    c.trace = trace;
    c.trace_cap = trace ? std::max(trace_cap, 1) : 0;
    c.max_pri = *std::max_element(mp.begin(), mp.end());
    c.lmax = ceil_log2(maxn) + 2;
    c.levels = std::min(c.lmax + 1, kLiftCap + 1 + kRotating);
    c.h_n.resize(G);
    c.h_m.resize(G);
    c.h_vbase.resize(G);
    std::vector<GameCtl> gc(G);
    for (int g = 0; g < G; g++) {
        c.h_n[g] = games[g].n;
        c.h_m[g] = games[g].m;
        c.h_vbase[g] = (int)vbase[g];
        std::memset(&gc[g], 0, sizeof(GameCtl));
        gc[g].vbase = (int)vbase[g];
        gc[g].n = games[g].n;
    }
    if (int rc = dalloc(&c.vinfo, N)) return rc;
    if (int rc = dalloc(&c.off, N + 1)) return rc;
    if (int rc = dalloc(&c.tgt, M)) return rc;
    if (int rc = dalloc(&c.gid, N)) return rc;
    if (int rc = dalloc(&c.strat, N)) return rc;
    if (int rc = dalloc(&c.par, N)) return rc;
    if (int rc = dalloc(&c.dep, N)) return rc;
    if (int rc = dalloc(&c.key, N)) return rc;
    if (int rc = dalloc(&c.solved, N)) return rc;
    if (int rc = dalloc(&c.lab, N)) return rc;
    if (int rc = dalloc(&c.blive, nblocks(N, kBS))) return rc;
    if (int rc = dalloc(&c.bact, 2LL * nblocks(N, kBS))) return rc;
    if (int rc = dalloc(&c.blast, nblocks(N, kBS))) return rc;
    if (int rc = dalloc(&c.Sl, N * c.levels)) return rc;
    c.chg_words = 4LL * nblocks(N, kBS);
    if (int rc = dalloc(&c.prog, 1)) return rc;
    if (int rc = dalloc(&c.chg, 2 * c.chg_words)) return rc;
    if (trace) {
        if (int rc = dalloc(&c.Zl, N * c.levels)) return rc;
        if (int rc = dalloc(&c.dig, N)) return rc;
        if (int rc = dalloc(&c.tr, (long long)G * c.trace_cap * 4)) return rc;
    }
    if (int rc = dalloc(&c.gctl, G)) return rc;
    if (int rc = dalloc(&c.jc, 1)) return rc;
    if (int rc = dalloc(&c.flag, 4)) return rc;
    long long *d_vbase = nullptr, *d_ebase = nullptr;
    int* d_cnt = nullptr;
    const unsigned nb = nblocks(N, kBS);
    if (int rc = dalloc(&d_vbase, G + 1)) return rc;
    if (int rc = dalloc(&d_ebase, G + 1)) return rc;
    if (int rc = dalloc(&d_cnt, nb)) return rc;
    pt.mark("alloc");
    // Each game's arrays go straight to their place (large totals through
    // pinned chunks); owner and priority are staged in par/dep, which the
    // first valuation overwrites.
    if ((size_t)(3 * N + M) * 4 >= (64ull << 20)) {
        std::vector<HostPart> own(G), pri(G), offs(G), tg(G);
        for (int g = 0; g < G; g++) {
            const GameView& gv = games[g];
            own[g] = HostPart{gv.owner, (size_t)gv.n * 4};
            pri[g] = HostPart{gv.pri, (size_t)gv.n * 4};
            offs[g] = HostPart{gv.off, (size_t)gv.n * 4};
            tg[g] = HostPart{gv.tgt, (size_t)gv.m * 4};
        }
        if (int rc = staged_upload(s, c.par, own.data(), G)) return rc;
        if (int rc = staged_upload(s, c.dep, pri.data(), G)) return rc;
        if (int rc = staged_upload(s, c.off, offs.data(), G)) return rc;
        if (int rc = staged_upload(s, c.tgt, tg.data(), G)) return rc;
    } else {
        for (int g = 0; g < G; g++) {
            const GameView& gv = games[g];
            PSI_HIP(hipMemcpyAsync(c.par + vbase[g], gv.owner, (size_t)gv.n * 4, hipMemcpyHostToDevice, s));
            PSI_HIP(hipMemcpyAsync(c.dep + vbase[g], gv.pri, (size_t)gv.n * 4, hipMemcpyHostToDevice, s));
            PSI_HIP(hipMemcpyAsync(c.off + vbase[g], gv.off, (size_t)gv.n * 4, hipMemcpyHostToDevice, s));
            PSI_HIP(hipMemcpyAsync(c.tgt + ebase[g], gv.tgt, (size_t)gv.m * 4, hipMemcpyHostToDevice, s));
        }
    }
    int Mi = (int)M;
    PSI_HIP(hipMemcpyAsync(c.off + N, &Mi, 4, hipMemcpyHostToDevice, s));
    PSI_HIP(hipMemcpyAsync(d_vbase, vbase.data(), (G + 1) * 8, hipMemcpyHostToDevice, s));
    PSI_HIP(hipMemcpyAsync(d_ebase, ebase.data(), (G + 1) * 8, hipMemcpyHostToDevice, s));
    PSI_HIP(hipMemcpyAsync(c.gctl, gc.data(), G * sizeof(GameCtl), hipMemcpyHostToDevice, s));
    pt.mark("copy");
    k_pack_vertices<<<nb, kBS, 0, s>>>(N, G, d_vbase, d_ebase, c.par, c.dep, c.vinfo, c.off, c.gid);
    k_globalize_targets<<<nb, kBS, 0, s>>>(N, c.off, c.gid, c.gctl, c.tgt);
    k_owner_count<<<nb, kBS, 0, s>>>(N, c.vinfo, d_cnt);
    k_owner_scan<<<1, 1024, 0, s>>>((int)nb, d_cnt);
    PSI_LAUNCH_CHECK();
    // number of player-1 vertices = last prefix + last block's count
    int last_pre = 0;
    PSI_HIP(hipMemcpyAsync(&last_pre, d_cnt + nb - 1, 4, hipMemcpyDeviceToHost, s));
    PSI_HIP(hipStreamSynchronize(s));
    int last_cnt = 0;
    {
        // count of the last block, recomputed on the host side from vinfo tail
        long long lo = (long long)(nb - 1) * kBS;
        std::vector<int> tail(N - lo);
        PSI_HIP(hipMemcpyAsync(tail.data(), c.vinfo + lo, (N - lo) * 4, hipMemcpyDeviceToHost, s));
        PSI_HIP(hipStreamSynchronize(s));
        for (int x : tail) last_cnt += (x >> 30) & 1;
    }
    c.own_n[1] = (long long)last_pre + last_cnt;
    c.own_n[0] = N - c.own_n[1];
    if (int rc = dalloc(&c.own[0], c.own_n[0])) return rc;
    if (int rc = dalloc(&c.own[1], c.own_n[1])) return rc;
    k_owner_scatter<<<nb, kBS, 0, s>>>(N, c.vinfo, d_cnt, c.own[0], c.own[1]);
    k_fill<int><<<1, 64, 0, s>>>(c.flag + 2, 1, (int)N);   // every vertex is occupied
    PSI_LAUNCH_CHECK();
    PSI_HIP(hipStreamSynchronize(s));
    (void)hipFree(d_vbase);
    (void)hipFree(d_ebase);
    (void)hipFree(d_cnt);
    pt.mark("pack on device");
    if (G == 1 && g_si_tune.push > 0 && N >= g_si_tune.push_min_n) {
        build_in_edges(c, s);
        pt.mark("in-edge lists");
    }
    return PSI_OK;
}

// ---- serving pool: games admitted into and released from fixed-capacity
// device arrays while the SI steps of other games run.  Admission work runs
// on a second stream; a region only joins the steps when its control block is
// written on the step stream, between two step batches.  Until then (and
// after it finishes) its vertices belong to a finished game, which every step
// kernel skips.
__global__ void k_pool_pack(int slot, long long vbase, int n, long long ebase, long long m,
                            int sentinel, const int* __restrict__ own_tmp,
                            const int* __restrict__ pri_tmp, int* __restrict__ vinfo,
                            int* __restrict__ off, int* __restrict__ gid)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    long long v = vbase + i;
    if (i == n) {                                 // the region's closing offset
        off[v] = (int)(ebase + m);
        gid[v] = sentinel;
        return;
    }
    vinfo[v] = pri_tmp[v] | (own_tmp[v] << 30);
    off[v] = (int)(off[v] + ebase);
    gid[v] = slot;
}

__global__ void k_pool_targets(long long ebase, long long m, long long vbase, int* __restrict__ tgt)
{
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < m) tgt[ebase + e] += (int)vbase;
}

// init_strat of one region (gpu_parity.cu:38-56).
__global__ void k_pool_init(long long vbase, int n, const int* __restrict__ vinfo,
                            const int* __restrict__ off, const int* __restrict__ tgt,
                            int* __restrict__ strat, int8_t* __restrict__ solved,
                            int8_t* __restrict__ lab)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    long long v = vbase + i;
    strat[v] = owner_of(vinfo[v]) == 0 ? -1 : tgt[off[v]];
    solved[v] = 0;
    lab[v] = 0;
}

// Owner lists of one region, written at the region's own positions:
// own1[vbase + k] = k-th player-1 vertex, own0 likewise; the rest stay -1.
__global__ void k_pool_owner_count(long long vbase, int n, const int* __restrict__ vinfo,
                                   int* __restrict__ cnt1)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int o = i < n ? owner_of(vinfo[vbase + i]) : 0;
    int c = __syncthreads_count(o);
    if (threadIdx.x == 0) cnt1[blockIdx.x] = c;
}

__global__ void k_pool_owner_scatter(long long vbase, int n, const int* __restrict__ vinfo,
                                     const int* __restrict__ pre1, int* __restrict__ own0,
                                     int* __restrict__ own1)
{
    __shared__ int wsum[kBS / kWave];
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int o = i < n ? owner_of(vinfo[vbase + i]) : 0;
    unsigned long long m = __ballot(o);
    int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    if (lane == 0) wsum[w] = __popcll(m);
    __syncthreads();
    int before = 0;
    for (int k = 0; k < w; k++) before += wsum[k];
    int r1 = pre1[blockIdx.x] + before + __popcll(m & ((1ull << lane) - 1ull));
    if (i < n) {
        if (o) own1[vbase + r1] = (int)(vbase + i);
        else own0[vbase + (i - r1)] = (int)(vbase + i);
    }
}

// Strategy of a finished region in the game's own vertex ids.
__global__ void k_pool_localize(long long vbase, int n, int* __restrict__ strat)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = strat[vbase + i];
    if (s >= 0) strat[vbase + i] = (int)(s - vbase);
}

__global__ void k_pool_release(long long vbase, long long cnt, int sentinel, int* __restrict__ gid,
                               int* __restrict__ own0, int* __restrict__ own1)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    gid[vbase + i] = sentinel;
    own0[vbase + i] = -1;
    own1[vbase + i] = -1;
}

long long core_pool_bytes_per_vertex(int maxn)
{
    const int lmax = ceil_log2(maxn) + 2;
    const int levels = std::min(lmax + 1, kLiftCap + 1 + kRotating);
    // vinfo off gid strat par dep key own0 own1 (ints), solved lab, records
    return 9 * 4 + 2 + 8LL * levels + 1;           // + block flags, change bits, rounded up
}

int core_create_pool(Core& c, long long vcap, long long ecap, int slots, int maxn, hipStream_t s)
{
    if (vcap <= 1 || ecap <= 0 || slots <= 0 || maxn <= 0)
        return fail(PSI_ERR_INVALID, "pool: capacities must be positive");
    if (vcap >= (1LL << 31) - 1 || ecap >= (1LL << 31) - 1)
        return fail(PSI_ERR_INVALID, "pool: capacities must fit in int32");
    const long long N = vcap, M = ecap;
    const int G = slots + 1;                      // the last slot is a finished sentinel
    c.G = G;
    c.N = N;
    c.M = M;
    c.s = s;
    c.pool = true;
    c.maxn_cap = maxn;
    c.trace = false;
    c.trace_cap = 0;
    c.lmax = ceil_log2(maxn) + 2;
    c.levels = std::min(c.lmax + 1, kLiftCap + 1 + kRotating);
    c.h_n.assign(G, 0);
    c.h_m.assign(G, 0);
    c.h_vbase.assign(G, 0);
    if (int rc = dalloc(&c.vinfo, N)) return rc;
    if (int rc = dalloc(&c.off, N + 1)) return rc;
    if (int rc = dalloc(&c.tgt, M)) return rc;
    if (int rc = dalloc(&c.gid, N)) return rc;
    if (int rc = dalloc(&c.strat, N)) return rc;
    if (int rc = dalloc(&c.par, N)) return rc;
    if (int rc = dalloc(&c.dep, N)) return rc;
    if (int rc = dalloc(&c.key, N)) return rc;
    if (int rc = dalloc(&c.solved, N)) return rc;
    if (int rc = dalloc(&c.lab, N)) return rc;
    if (int rc = dalloc(&c.blive, nblocks(N, kBS))) return rc;
    if (int rc = dalloc(&c.bact, 2LL * nblocks(N, kBS))) return rc;
    if (int rc = dalloc(&c.blast, nblocks(N, kBS))) return rc;
    if (int rc = dalloc(&c.Sl, N * c.levels)) return rc;
    c.chg_words = 4LL * nblocks(N, kBS);
    if (int rc = dalloc(&c.chg, 2 * c.chg_words)) return rc;
    if (int rc = dalloc(&c.own[0], N)) return rc;
    if (int rc = dalloc(&c.own[1], N)) return rc;
    if (int rc = dalloc(&c.gctl, G)) return rc;
    if (int rc = dalloc(&c.jc, 1)) return rc;
    if (int rc = dalloc(&c.flag, 4)) return rc;
    if (int rc = dalloc(&c.cnt_scratch, nblocks((long long)maxn + 1, kBS))) return rc;
    if (int rc = dalloc(&c.prog, 1)) return rc;
    c.own_n[0] = c.own_n[1] = N;
    std::vector<GameCtl> gc(G);
    for (auto& x : gc) {
        std::memset(&x, 0, sizeof x);
        x.done = 1;
    }
    PSI_HIP(hipMemcpyAsync(c.gctl, gc.data(), G * sizeof(GameCtl), hipMemcpyHostToDevice, s));
    k_fill<int><<<nblocks(N, kBS), kBS, 0, s>>>(c.gid, N, slots);
    k_fill<int><<<nblocks(N, kBS), kBS, 0, s>>>(c.own[0], N, -1);
    k_fill<int><<<nblocks(N, kBS), kBS, 0, s>>>(c.own[1], N, -1);
    k_fill<int><<<1, 64, 0, s>>>(c.flag, 4, 0);
    const long long jn = (long long)(sizeof(JumpCnt) / 4);
    k_fill<int><<<nblocks(jn, kBS), kBS, 0, s>>>((int*)c.jc, jn, 0);
    k_prog_reset<<<1, 256, 0, s>>>(G, c.gctl, c.prog);
    PSI_LAUNCH_CHECK();
    PSI_HIP(hipStreamSynchronize(s));
    return PSI_OK;
}

int core_pool_admit(Core& c, int slot, long long vbase, long long ebase, const GameView& gv,
                    hipStream_t s2)
{
    const int n = gv.n;
    const long long m = gv.m;
    if (slot < 0 || slot >= c.G - 1 || vbase < 0 || vbase + n + 1 > c.N || ebase < 0 || ebase + m > c.M)
        return fail(PSI_ERR_INVALID, "pool: region out of range");
    PSI_HIP(hipMemcpyAsync(c.par + vbase, gv.owner, (size_t)n * 4, hipMemcpyHostToDevice, s2));
    PSI_HIP(hipMemcpyAsync(c.dep + vbase, gv.pri, (size_t)n * 4, hipMemcpyHostToDevice, s2));
    PSI_HIP(hipMemcpyAsync(c.off + vbase, gv.off, (size_t)n * 4, hipMemcpyHostToDevice, s2));
    PSI_HIP(hipMemcpyAsync(c.tgt + ebase, gv.tgt, (size_t)m * 4, hipMemcpyHostToDevice, s2));
    const unsigned nb = nblocks((long long)n, kBS);
    k_pool_pack<<<nblocks((long long)n + 1, kBS), kBS, 0, s2>>>(slot, vbase, n, ebase, m, c.G - 1,
                                                               c.par, c.dep, c.vinfo, c.off, c.gid);
    k_pool_targets<<<nblocks(m, kBS), kBS, 0, s2>>>(ebase, m, vbase, c.tgt);
    k_pool_init<<<nb, kBS, 0, s2>>>(vbase, n, c.vinfo, c.off, c.tgt, c.strat, c.solved, c.lab);
    k_pool_owner_count<<<nb, kBS, 0, s2>>>(vbase, n, c.vinfo, c.cnt_scratch);
    k_owner_scan<<<1, 1024, 0, s2>>>((int)nb, c.cnt_scratch);
    k_pool_owner_scatter<<<nb, kBS, 0, s2>>>(vbase, n, c.vinfo, c.cnt_scratch, c.own[0], c.own[1]);
    PSI_LAUNCH_CHECK();
    c.h_n[slot] = n;
    c.h_m[slot] = m;
    c.h_vbase[slot] = (int)vbase;
    return PSI_OK;
}

int core_pool_localize(Core& c, int slot, hipStream_t s2)
{
    k_pool_localize<<<nblocks((long long)c.h_n[slot], kBS), kBS, 0, s2>>>(c.h_vbase[slot], c.h_n[slot],
                                                                        c.strat);
    PSI_LAUNCH_CHECK();
    return PSI_OK;
}

int core_pool_release(Core& c, int slot, hipStream_t s2)
{
    long long vb = c.h_vbase[slot], cnt = (long long)c.h_n[slot] + 1;
    k_pool_release<<<nblocks(cnt, kBS), kBS, 0, s2>>>(vb, cnt, c.G - 1, c.gid, c.own[0], c.own[1]);
    PSI_LAUNCH_CHECK();
    c.h_n[slot] = 0;
    c.h_m[slot] = 0;
    return PSI_OK;
}

void core_destroy(Core& c)
{
    if (c.graph) (void)hipGraphExecDestroy(c.graph);
    if (c.pgraph) (void)hipGraphExecDestroy(c.pgraph);
    void* ptrs[] = {c.vinfo, c.off, c.tgt, c.gid, c.strat, c.par, c.dep, c.key, c.solved, c.lab, c.blive,
                    c.bact, c.blast, c.Sl, c.own[0], c.own[1], c.Zl, c.dig, c.tr, c.gctl, c.jc, c.flag,
                    c.cnt_scratch, c.chg, c.succ, c.prog, c.pargs, c.rin_mem};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    c = Core();
}

#define GRID(n) nblocks((n), kBS), kBS, 0, c.s
// Jump rounds: capped grid (grid-stride inside) so idle launches stay cheap.
#define JGRID(n) std::min(nblocks((n), kBS), jump_grid_cap()), kBS, 0, c.s
// Step kernels: capped grid-stride launches, so that a serving pool's large
// capacity costs nothing above its occupied bound.
#define SGRID(n) std::min(nblocks((n), kBS), c.pool ? g_si_tune.pool_grid : g_si_tune.step_grid), kBS, 0, c.s

// Launch-shape knobs (psi_si_tune; the autotuner sets them).  They apply to
// graphs captured afterwards.
SiTuning g_si_tune;

static unsigned jump_grid_cap() { return g_si_tune.jump_grid; }

static int maxn_of(const Core& c)
{
    if (c.pool) return c.maxn_cap;               // any admitted game is at most this long
    int m = 0;
    for (int n : c.h_n) m = std::max(m, n);
    return m;
}

static void launch_val_init(Core& c)
{
    k_val_init<<<SGRID(c.N)>>>(c.N, c.vinfo, c.off, c.tgt, c.gid, c.gctl, c.strat, c.solved, c.lab,
                              c.par, c.Sl, c.Zl, c.jc, c.blive, c.bact, c.blast, c.trace, c.flag + 2);
}

static void launch_jump(Core& c, int r, int maxn)
{
    const long long nb = nblocks(c.N, kBS);
    k_jump<<<JGRID(c.N)>>>(r, c.N, maxn, c.Sl, c.Zl, c.jc, c.blive, c.bact + (r & 1) * nb,
                           c.bact + ((r + 1) & 1) * nb, c.blast, c.trace, c.flag + 2);
}

// The same launch with events that take the kernel's own start and end
// times (the dispatch packet's timestamps, as a profiler reads them), not
// the stream's position around it.
static void launch_jump_timed(Core& c, int r, int maxn, hipEvent_t start, hipEvent_t stop)
{
    const long long nb = nblocks(c.N, kBS);
    hipExtLaunchKernelGGL(k_jump, dim3((unsigned)std::min(nblocks(c.N, kBS), jump_grid_cap())), dim3(kBS),
                          0, c.s, start, stop, 0, r, c.N, maxn, c.Sl, c.Zl, c.jc, (const uint8_t*)c.blive,
                          (const uint8_t*)(c.bact + (r & 1) * nb), c.bact + ((r + 1) & 1) * nb,
                          c.blast, (int)c.trace, (const int*)(c.flag + 2));
}

static void launch_label(Core& c, int maxn, bool loop_form)
{
    k_label<<<SGRID(c.N)>>>(c.N, maxn, c.lmax, c.Sl, c.Zl, c.blast, c.gid, c.gctl, c.strat,
                           c.solved, c.lab, c.dep, c.key, c.dig, c.tr, c.trace_cap, loop_form,
                           c.trace, c.jc, c.blive, c.flag + 2, c.chg, c.chg + c.chg_words);
}

static void launch_switch(Core& c, int player, bool loop_form)
{
    if (loop_form && player == 0)
        k_mark0_outer<<<SGRID(c.N)>>>(c.N, c.gid, c.gctl, c.lab, c.solved, c.flag + 2, c.vinfo,
                                      c.strat, c.succ);
    if (c.own_n[player] == 0) return;
    k_switch<<<SGRID(c.own_n[player])>>>(player, c.N, c.own[player], c.own_n[player], c.vinfo,
                                        c.off, c.tgt, c.gid, c.gctl, c.strat, c.key, c.dep,
                                        c.par, c.Sl, c.jc, loop_form, c.flag, c.prof_ties,
                                        c.flag + 2, c.chg, c.chg + c.chg_words);
}

int core_set_params(Core& c, int reset, int max_rounds, int bf)
{
    std::vector<GameCtl> gc(c.G);
    PSI_HIP(hipMemcpyAsync(gc.data(), c.gctl, c.G * sizeof(GameCtl), hipMemcpyDeviceToHost, c.s));
    PSI_HIP(hipStreamSynchronize(c.s));
    for (auto& x : gc) {
        x.reset = reset;
        x.max_rounds = max_rounds;
        x.bf = bf;
    }
    if (bf == 1 && !c.succ) {
        if (int rc = dalloc(&c.succ, c.N)) return rc;
        // captured launches hold the (null) succ they were captured with
        std::lock_guard<std::recursive_mutex> lk(setup_mutex());
        if (c.graph) (void)hipGraphExecDestroy(c.graph);
        if (c.pgraph) (void)hipGraphExecDestroy(c.pgraph);
        c.graph = nullptr;
        c.pgraph = nullptr;
    }
    if (bf == 1) k_succ_init<<<GRID(c.N)>>>(c.N, c.gid, c.gctl, c.succ);
    PSI_HIP(hipMemcpyAsync(c.gctl, gc.data(), c.G * sizeof(GameCtl), hipMemcpyHostToDevice, c.s));
    PSI_HIP(hipStreamSynchronize(c.s));
    return PSI_OK;
}

int core_init_strat(Core& c)
{
    k_init_strat<<<GRID(c.N)>>>(c.N, c.vinfo, c.off, c.tgt, c.strat, c.solved, c.lab);
    k_reset_ctl<<<GRID(c.G)>>>(0, c.G, c.gctl);
    k_fill<int><<<GRID((long long)(sizeof(JumpCnt) / 4))>>>((int*)c.jc, (long long)(sizeof(JumpCnt) / 4), 0);
    if (c.trace) k_fill<unsigned long long><<<GRID((long long)c.G * c.trace_cap * 4)>>>(
        c.tr, (long long)c.G * c.trace_cap * 4, 0ull);
    PSI_LAUNCH_CHECK();
    return PSI_OK;
}

int core_valuation(Core& c, bool loop_form)
{
    int maxn = maxn_of(c);
    launch_val_init(c);
    for (int r = 0; r < c.lmax; r++) launch_jump(c, r, maxn);
    launch_label(c, maxn, loop_form);
    PSI_LAUNCH_CHECK();
    return PSI_OK;
}

int core_reset_opp(Core& c)
{
    k_reset_opp<<<GRID(c.N)>>>(c.N, c.vinfo, c.off, c.tgt, c.strat);
    PSI_LAUNCH_CHECK();
    return PSI_OK;
}

// A state loaded from the host (strategies, solved flags, labels) starts the
// game's run afresh: its control block restarts, so that a finished game takes
// part in the next valuation again and that valuation, as the first of the
// run, stores every key, depth and link (later ones store only changes).
int core_state_loaded(Core& c, int game)
{
    k_reset_ctl<<<1, 64, 0, c.s>>>(game, game + 1, c.gctl);
    PSI_LAUNCH_CHECK();
    return PSI_OK;
}

int core_mark_solved(Core& c, int player)
{
    k_mark_solved<<<GRID(c.N)>>>(player, c.N, c.lab, c.solved);
    PSI_LAUNCH_CHECK();
    return PSI_OK;
}

int core_switch(Core& c, int player, int* switched_host)
{
    k_fill<int><<<1, 64, 0, c.s>>>(c.flag, 1, 0);
    launch_switch(c, player, false);
    PSI_LAUNCH_CHECK();
    int h = 0;
    PSI_HIP(hipMemcpyAsync(&h, c.flag, sizeof(int), hipMemcpyDeviceToHost, c.s));
    PSI_HIP(hipStreamSynchronize(c.s));
    *switched_host = h;
    return PSI_OK;
}

int core_enqueue_steps(Core& c, int steps)
{
    for (int k = 0; k < steps; k++) {
        if (c.pool) k_pool_hi<<<1, 256, 0, c.s>>>(c.G, c.gctl, c.flag + 2);
        if (int rc = core_valuation(c, true)) return rc;
        launch_switch(c, 1, true);
        k_ctl1<<<GRID(c.G)>>>(c.G, c.gctl);
        launch_switch(c, 0, true);
        k_ctl0<<<GRID(c.G)>>>(c.G, c.gctl);
        PSI_LAUNCH_CHECK();
    }
    return PSI_OK;
}

// The same loop launched kernel by kernel with HIP events between launches, to
// attribute device time per kernel.  times[k], counts[k] for k in
// PSI_K_VAL_INIT..PSI_K_CTL; a jump launch that found nothing to do is counted
// under PSI_K_JUMP_IDLE; counts[PSI_K_GATHERS] = successor records gathered.
int core_solve_profiled(Core& c, double* times, long long* counts)
{
    for (int k = 0; k < PSI_K_COUNT; k++) {
        times[k] = 0;
        counts[k] = 0;
    }
    int maxn = maxn_of(c);
    const int nev = c.lmax + 8;
    std::vector<hipEvent_t> ev(nev), jev(2 * c.lmax);   // jev: each round's kernel start and end
    for (auto& e : ev) PSI_HIP(hipEventCreate(&e));
    for (auto& e : jev) PSI_HIP(hipEventCreate(&e));
    auto span = [&](int a, int b) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, ev[a], ev[b]);
        return (double)ms;
    };
    int rc = PSI_OK;
    unsigned long long* ties = nullptr;
    {
        std::lock_guard<std::recursive_mutex> lk(setup_mutex());
        PSI_HIP(hipMalloc(&ties, 24));
    }
    PSI_HIP(hipMemsetAsync(ties, 0, 24, c.s));
    c.prof_ties = ties;
    c.prof_rounds.clear();
    for (int step = 0; step < (1 << 20); step++) {
        int k = 0;
        PSI_HIP(hipEventRecord(ev[k++], c.s));
        launch_val_init(c);
        PSI_HIP(hipEventRecord(ev[k++], c.s));
        for (int r = 0; r < c.lmax; r++) {
            launch_jump_timed(c, r, maxn, jev[2 * r], jev[2 * r + 1]);
            PSI_HIP(hipEventRecord(ev[k++], c.s));
        }
        launch_label(c, maxn, true);
        PSI_HIP(hipEventRecord(ev[k++], c.s));
        launch_switch(c, 1, true);
        PSI_HIP(hipEventRecord(ev[k++], c.s));
        k_ctl1<<<GRID(c.G)>>>(c.G, c.gctl);
        PSI_HIP(hipEventRecord(ev[k++], c.s));
        launch_switch(c, 0, true);
        PSI_HIP(hipEventRecord(ev[k++], c.s));
        k_ctl0<<<GRID(c.G)>>>(c.G, c.gctl);
        PSI_HIP(hipEventRecord(ev[k++], c.s));
        PSI_LAUNCH_CHECK();
        PSI_HIP(hipStreamSynchronize(c.s));
        JumpCnt jc;
        PSI_HIP(hipMemcpyAsync(&jc, c.jc, sizeof jc, hipMemcpyDeviceToHost, c.s));
        PSI_HIP(hipStreamSynchronize(c.s));
        int e = 0;
        c.prof_rounds.push_back(jc.last + 1);
        times[PSI_K_VAL_INIT] += span(e, e + 1);
        counts[PSI_K_VAL_INIT]++;
        e++;
        unsigned long long prev_act = 0;
        for (int r = 0; r < c.lmax; r++, e++) {
            int cls = r <= jc.last ? PSI_K_JUMP : PSI_K_JUMP_IDLE;
            float kms = 0;
            (void)hipEventElapsedTime(&kms, jev[2 * r], jev[2 * r + 1]);
            times[cls] += kms;
            counts[cls]++;
            unsigned long long res = 0, act = 0;
            if (r < kMaxRounds)
                for (int q = 0; q < kSlots; q++) {
                    res += jc.slot[r][q] >> 32;
                    act += jc.slot[r][q] & 0xFFFFFFFFu;
                }
            if (r <= jc.last) counts[PSI_K_GATHERS] += (long long)(r == 0 ? res + act : prev_act);
            prev_act = act;
        }
        times[PSI_K_LABEL] += span(e, e + 1); counts[PSI_K_LABEL]++; e++;
        times[PSI_K_SWITCH1] += span(e, e + 1); counts[PSI_K_SWITCH1]++; e++;
        times[PSI_K_CTL] += span(e, e + 1); counts[PSI_K_CTL]++; e++;
        times[PSI_K_SWITCH0] += span(e, e + 1); counts[PSI_K_SWITCH0]++; e++;
        times[PSI_K_CTL] += span(e, e + 1); counts[PSI_K_CTL]++; e++;
        times[PSI_K_TOTAL] += span(0, e);
        counts[PSI_K_TOTAL]++;
        int done = 0;
        if ((rc = core_all_done(c, &done))) break;
        if (done) break;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    for (auto& e : jev) (void)hipEventDestroy(e);
    unsigned long long h_ties[3] = {0, 0, 0};
    (void)hipMemcpyAsync(h_ties, ties, 24, hipMemcpyDeviceToHost, c.s);
    (void)hipStreamSynchronize(c.s);
    counts[PSI_K_TIES] = (long long)h_ties[0];
    counts[PSI_K_SCANS] = (long long)h_ties[1];
    counts[PSI_K_SCAN_STEPS] = (long long)h_ties[2];
    c.prof_ties = nullptr;
    {
        std::lock_guard<std::recursive_mutex> lk(setup_mutex());
        (void)hipFree(ties);
    }
    return rc;
}

int core_all_done(Core& c, int* done)
{
    k_count_live<<<1, 256, 0, c.s>>>(c.G, c.gctl, c.flag + 1);
    PSI_LAUNCH_CHECK();
    int live = 0;
    PSI_HIP(hipMemcpyAsync(&live, c.flag + 1, sizeof(int), hipMemcpyDeviceToHost, c.s));
    PSI_HIP(hipStreamSynchronize(c.s));
    *done = live == 0;
    return PSI_OK;
}

// Captures `steps` SI steps once into a graph and replays it until every game
// is done.  No host round trip inside a batch of steps.
int core_launch_steps(Core& c, int steps_per_check)
{
    if (steps_per_check < 1) steps_per_check = 1;
    if (!c.graph || c.graph_steps != steps_per_check) {
        if (c.graph) {
            (void)hipGraphExecDestroy(c.graph);
            c.graph = nullptr;
        }
        hipGraph_t gr;
        std::lock_guard<std::recursive_mutex> lk(setup_mutex());
        PSI_HIP(hipStreamBeginCapture(c.s, hipStreamCaptureModeThreadLocal));
        int rc = core_enqueue_steps(c, steps_per_check);
        hipError_t e = hipStreamEndCapture(c.s, &gr);
        if (e != hipSuccess) return rc ? rc : hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
        if (rc) {                                 // the captured graph is not used
            (void)hipGraphDestroy(gr);
            return rc;
        }
        e = hipGraphInstantiate(&c.graph, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
        c.graph_steps = steps_per_check;
    }
    PSI_HIP(hipGraphLaunch(c.graph, c.s));
    return PSI_OK;
}

// Change marks are set when at most this many vertices changed (0: never).
static long long push_max_of(const Core& c)
{
    if (!c.rin_off || !c.push_ok || g_si_tune.push <= 0) return 0;
    return std::max(1LL, c.N / g_si_tune.push);
}

static StepArgs step_args(Core& c)
{
    StepArgs a;
    a.N = c.N;
    a.nb = nblocks(c.N, kBS);
    a.own_n0 = c.own_n[0];
    a.own_n1 = c.own_n[1];
    a.G = c.G;
    a.maxn = maxn_of(c);
    a.lmax = c.lmax;
    a.two_levels = g_si_tune.two_levels;
    a.trace = c.trace;
    a.trace_cap = c.trace_cap;
    a.vinfo = c.vinfo;
    a.off = c.off;
    a.tgt = c.tgt;
    a.gid = c.gid;
    a.own0 = c.own[0];
    a.own1 = c.own[1];
    a.rin_off = c.rin_off;
    a.rin_src = c.rin_src;
    a.dm0 = c.dmark[0];
    a.dm1 = c.dmark[1];
    a.push_max = push_max_of(c);
    a.d1 = a.push_max ? c.d1 : nullptr;
    a.l12ok = a.push_max ? c.l12ok : nullptr;
    a.bchg = a.push_max ? c.bchg : nullptr;
    a.hi = c.flag + 2;
    a.gc = c.gctl;
    a.strat = c.strat;
    a.par = c.par;
    a.dep = c.dep;
    a.key = c.key;
    a.flag = c.flag;
    a.succ = c.succ;
    a.solved = c.solved;
    a.lab = c.lab;
    a.blive = c.blive;
    a.bact = c.bact;
    a.blast = c.blast;
    a.Sl = c.Sl;
    a.Zl = c.Zl;
    a.dig = c.dig;
    a.tr = c.tr;
    a.chg = c.chg;
    a.chg0 = c.chg + c.chg_words;
    a.jc = c.jc;
    return a;
}

// Copies the step-program arguments to the device (before a capture).
static int upload_step_args(Core& c, StepArgs** out)
{
    if (!c.pargs) {
        StepArgs* p = nullptr;
        if (int rc = dalloc(&p, 1)) return rc;
        c.pargs = p;
    }
    const StepArgs a = step_args(c);
    PSI_HIP(hipMemcpyAsync(c.pargs, &a, sizeof a, hipMemcpyHostToDevice, c.s));
    PSI_HIP(hipStreamSynchronize(c.s));
    c.pargs_two = a.two_levels;
    c.pargs_push = a.push_max;
    *out = (StepArgs*)c.pargs;
    return PSI_OK;
}

// Workgroups of a step-program launch: at most the tuned grid, and at most
// one per two 256-vertex blocks, so that on small games every wave of a
// workgroup has a block to run (a workgroup's blocks are dealt to its four
// waves) and fewer workgroups are dispatched per launch.
static unsigned program_grid(const Core& c)
{
    return std::max(1u, std::min((nblocks(c.N, kBS) + 1) / 2, g_si_tune.phase_grid));
}

// Launches one graph of `launches` (a multiple of 4) step-program launches,
// captured on first use.
static int launch_program(Core& c, int launches)
{
    if (c.pgraph && (c.pargs_two != g_si_tune.two_levels || c.pargs_push != push_max_of(c))) {
        // tuned (or a state loaded) since the capture
        StepArgs* a = nullptr;
        if (int rc = upload_step_args(c, &a)) return rc;
    }
    if (!c.pgraph || c.pgraph_launches != launches || c.pgraph_grid != program_grid(c)) {
        if (c.pgraph) {
            (void)hipGraphExecDestroy(c.pgraph);
            c.pgraph = nullptr;
        }
        StepArgs* a = nullptr;
        if (int rc = upload_step_args(c, &a)) return rc;
        const unsigned grid = program_grid(c);
        hipGraph_t gr;
        std::lock_guard<std::recursive_mutex> lk(setup_mutex());
        PSI_HIP(hipStreamBeginCapture(c.s, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < launches; i++) k_phase<<<grid, kBS, 0, c.s>>>(i, a, c.prog);
        hipError_t e = hipGetLastError();
        hipError_t e2 = hipStreamEndCapture(c.s, &gr);
        if (e2 != hipSuccess) {
            return e != hipSuccess ? hip_fail(e, "k_phase launch", __FILE__, __LINE__)
                                   : hip_fail(e2, "hipStreamEndCapture", __FILE__, __LINE__);
        }
        if (e != hipSuccess) {                    // the captured graph is not used
            (void)hipGraphDestroy(gr);
            return hip_fail(e, "k_phase launch", __FILE__, __LINE__);
        }
        e = hipGraphInstantiate(&c.pgraph, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
        c.pgraph_launches = launches;
        c.pgraph_grid = grid;
    }
    PSI_HIP(hipGraphLaunch(c.pgraph, c.s));
    return PSI_OK;
}

// The step program launched one launch at a time with HIP events around each
// (no graph), reading back which phase each launch ran: device milliseconds
// and launch counts per phase class (the PSI_K_* classes of
// core_solve_profiled; mark_solved(0) counts with switch(0)).
int core_solve_program_profiled(Core& c, double* times, long long* counts)
{
    for (int k = 0; k < PSI_K_COUNT; k++) {
        times[k] = 0;
        counts[k] = 0;
    }
    if (c.pool || !c.prog) return fail(PSI_ERR_INVALID, "step program: not available on this handle");
    StepArgs* a = nullptr;
    if (int rc = upload_step_args(c, &a)) return rc;
    const unsigned grid = program_grid(c);
    hipEvent_t e0, e1;
    PSI_HIP(hipEventCreate(&e0));
    PSI_HIP(hipEventCreate(&e1));
    k_prog_reset<<<1, 256, 0, c.s>>>(c.G, c.gctl, c.prog, c.l12ok, nblocks(c.N, kBS));
    int rc = PSI_OK;
    c.prof_phase.clear();
    c.prof_us.clear();
    for (long long i = 0; i < (1LL << 26); i++) {
        const int par = (int)(i & 1);
        // events attached to the dispatch itself: the kernel's own start and
        // end, as a profiler's kernel trace reads them
        hipExtLaunchKernelGGL(k_phase, dim3(grid), dim3(kBS), 0, c.s, e0, e1, 0, (int)(i & 3), a,
                              c.prog);
        if (hipGetLastError() != hipSuccess) { rc = fail(PSI_ERR_HIP, "step program launch"); break; }
        int st = 0;
        if (hipMemcpyAsync(&st, &c.prog->st[par], sizeof st, hipMemcpyDeviceToHost, c.s) != hipSuccess ||
            hipStreamSynchronize(c.s) != hipSuccess) {
            rc = fail(PSI_ERR_HIP, "step program: launch failed");
            break;
        }
        const int ph = st & 0x7f;
        if (ph == PH_DONE) break;
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        c.prof_phase.push_back(st);
        c.prof_us.push_back(ms * 1000.0f);
        int cls = PSI_K_CTL;
        switch (ph) {
        case PH_INIT: cls = PSI_K_VAL_INIT; counts[PSI_K_TOTAL]++; break;
        case PH_JUMP: cls = PSI_K_JUMP; break;
        case PH_LABEL: cls = PSI_K_LABEL; break;
        case PH_SW1: cls = PSI_K_SWITCH1; break;
        case PH_MARK0: case PH_SW0: cls = PSI_K_SWITCH0; break;
        case PH_PUSH: cls = PSI_K_PUSH; break;
        default: break;
        }
        times[cls] += ms;
        counts[cls]++;
        times[PSI_K_TOTAL] += ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int core_pool_launch_program(Core& c, int steps)
{
    return launch_program(c, 16 * std::max(1, steps));
}

int core_solve(Core& c, int steps_per_check)
{
    if (!c.pool && c.prog && g_si_tune.program) {
        // ~16 launches per SI step on random games; an even count keeps the
        // state slots alternating across graphs
        const int launches = 16 * std::max(1, steps_per_check);   // a multiple of 4
        k_prog_reset<<<1, 256, 0, c.s>>>(c.G, c.gctl, c.prog, c.l12ok, nblocks(c.N, kBS));
        PSI_LAUNCH_CHECK();
        for (;;) {
            if (int rc = launch_program(c, launches)) return rc;
            int done = 0;
            if (int rc = core_all_done(c, &done)) return rc;
            if (done) break;
        }
    } else {
        for (;;) {
            if (int rc = core_launch_steps(c, steps_per_check)) return rc;
            int done = 0;
            if (int rc = core_all_done(c, &done)) return rc;
            if (done) break;
        }
    }
    JumpCnt jc;
    PSI_HIP(hipMemcpyAsync(&jc, c.jc, sizeof jc, hipMemcpyDeviceToHost, c.s));
    PSI_HIP(hipStreamSynchronize(c.s));
    if (jc.err) return fail(PSI_ERR_INVALID, "internal: jump rounds exhausted");
    return PSI_OK;
}

int core_solve_bf(Core& c, int steps_per_check)
{
    if (int rc = core_solve(c, steps_per_check)) return rc;
    k_bf_final<<<GRID(c.N)>>>(c.N, c.gid, c.gctl, c.vinfo, c.succ, c.strat);
    PSI_LAUNCH_CHECK();
    PSI_HIP(hipStreamSynchronize(c.s));
    return PSI_OK;
}

} // namespace psi
