// Seeded workload generators (host only; no HIP calls).
//
// Every generator uses the same explicit PRNG so that a (seed, size) pair
// names one input on every machine: xoshiro256** seeded from a splitmix64
// stream, bounded integers by Lemire's multiply-shift with rejection.
//
// Random game (the generator the benchmark configurations describe):
//   for v in 0..n-1, in order:
//     owner(v)    = top bit of next()                       Bernoulli(1/2)
//     priority(v) = bounded(pri_bound)                       U[0, pri_bound)
//     d           = deg_lo + bounded(deg_hi - deg_lo + 1)    U[deg_lo, deg_hi]
//     successors  = draws bounded(n), a draw equal to an earlier successor of
//                   v is redrawn, until d distinct (self-loops allowed);
//                   CSR order = draw order (it is the switch tie-break order).
// Chain game: vertex i has priority i and owner i%2; its first successor is
// i+1 (the last vertex points to 0); with probability 1/10 it gets one extra
// successor drawn uniformly (skipped if equal to the first).
// Random list: succ[perm[i]] = perm[i+1] for a Fisher-Yates permutation.
// Forest of lists: lengths log-uniform in [1, len_hi], rescaled to the
// requested node total, laid over one random permutation.

#include "../../include/psi.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct Rng {
    uint64_t s[4];
    static uint64_t sm(uint64_t& x)
    {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed)
    {
        uint64_t x = seed;
        for (auto& w : s) w = sm(x);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    // Uniform in [0, range), range >= 1 (Lemire 2019).
    uint64_t bounded(uint64_t range)
    {
        uint64_t x = next();
        __uint128_t m = (__uint128_t)x * range;
        uint64_t l = (uint64_t)m;
        if (l < range) {
            uint64_t t = (0 - range) % range;
            while (l < t) {
                x = next();
                m = (__uint128_t)x * range;
                l = (uint64_t)m;
            }
        }
        return (uint64_t)(m >> 64);
    }
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

} // namespace

extern "C" void psi_set_error_(const char* msg);

extern "C" int psi_gen_random_game(uint64_t seed, int n, int pri_bound, int deg_lo, int deg_hi,
                                   int* owner, int* priority, int* offsets, int* targets,
                                   int64_t targets_cap, int64_t* m_out)
{
    if (n <= 0 || pri_bound <= 0 || deg_lo < 1 || deg_hi < deg_lo) {
        psi_set_error_("psi_gen_random_game: bad arguments");
        return PSI_ERR_INVALID;
    }
    Rng r(seed);
    int64_t m = 0;
    offsets[0] = 0;
    int picks[64];
    for (int v = 0; v < n; v++) {
        owner[v] = (int)(r.next() >> 63);
        priority[v] = (int)r.bounded((uint64_t)pri_bound);
        int d = deg_lo + (int)r.bounded((uint64_t)(deg_hi - deg_lo + 1));
        if (d > n) d = n;
        if (d > 64) d = 64;
        if (m + d > targets_cap || m + d > INT32_MAX) {
            psi_set_error_("psi_gen_random_game: targets capacity exceeded");
            return PSI_ERR_INVALID;
        }
        int k = 0;
        while (k < d) {
            int t = (int)r.bounded((uint64_t)n);
            bool dup = false;
            for (int j = 0; j < k; j++) dup |= picks[j] == t;
            if (!dup) picks[k++] = t;
        }
        for (int j = 0; j < d; j++) targets[m + j] = picks[j];
        m += d;
        offsets[v + 1] = (int)m;
    }
    *m_out = m;
    return PSI_OK;
}

extern "C" int psi_gen_chain_game(uint64_t seed, int n, int* owner, int* priority, int* offsets,
                                  int* targets, int64_t targets_cap, int64_t* m_out)
{
    if (n <= 1) {
        psi_set_error_("psi_gen_chain_game: n must be > 1");
        return PSI_ERR_INVALID;
    }
    Rng r(seed);
    int64_t m = 0;
    offsets[0] = 0;
    for (int v = 0; v < n; v++) {
        owner[v] = v & 1;
        priority[v] = v;
        int first = v + 1 < n ? v + 1 : 0;
        if (m + 2 > targets_cap) {
            psi_set_error_("psi_gen_chain_game: targets capacity exceeded");
            return PSI_ERR_INVALID;
        }
        targets[m++] = first;
        if (r.bounded(10) == 0) {
            int t = (int)r.bounded((uint64_t)n);
            if (t != first) targets[m++] = t;
        }
        offsets[v + 1] = (int)m;
    }
    *m_out = m;
    return PSI_OK;
}

extern "C" int psi_gen_sizes(uint64_t seed, int count, double lo, double hi, int* out)
{
    if (count < 0 || lo < 1 || hi < lo) {
        psi_set_error_("psi_gen_sizes: bad arguments");
        return PSI_ERR_INVALID;
    }
    Rng r(seed);
    double a = std::log(lo), b = std::log(hi);
    for (int i = 0; i < count; i++) {
        double x = std::exp(a + r.unit() * (b - a));
        long long v = std::llround(x);
        if (v < (long long)lo) v = (long long)lo;
        if (v > (long long)hi) v = (long long)hi;
        out[i] = (int)v;
    }
    return PSI_OK;
}

extern "C" int psi_gen_permutation(uint64_t seed, int n, int* perm)
{
    Rng r(seed);
    for (int i = 0; i < n; i++) perm[i] = i;
    for (int i = n - 1; i > 0; i--) {
        int j = (int)r.bounded((uint64_t)i + 1);
        int t = perm[i];
        perm[i] = perm[j];
        perm[j] = t;
    }
    return PSI_OK;
}

extern "C" int psi_gen_list(uint64_t seed, int n, int* succ, int* head)
{
    if (n <= 0) {
        psi_set_error_("psi_gen_list: n must be positive");
        return PSI_ERR_INVALID;
    }
    std::vector<int> perm(n);
    psi_gen_permutation(seed, n, perm.data());
    for (int i = 0; i + 1 < n; i++) succ[perm[i]] = perm[i + 1];
    succ[perm[n - 1]] = -1;
    *head = perm[0];
    return PSI_OK;
}

// Forest of n_lists lists with n_nodes nodes in total.  Raw lengths are
// log-uniform in [1, len_hi]; they are scaled so their mean is n_nodes/n_lists,
// rounded, clamped to >= 1, and the remaining difference to n_nodes is spread
// one node at a time over the lists in index order (skipping lists of length 1
// when removing).  lengths_out may be null.
extern "C" int psi_gen_forest(uint64_t seed, int n_nodes, int n_lists, int len_hi, int* succ,
                              int* heads, int* lengths_out)
{
    if (n_lists <= 0 || n_nodes < n_lists || len_hi < 1) {
        psi_set_error_("psi_gen_forest: bad arguments");
        return PSI_ERR_INVALID;
    }
    Rng r(seed);
    std::vector<double> raw(n_lists);
    double lh = std::log((double)len_hi), sum = 0;
    for (int i = 0; i < n_lists; i++) {
        raw[i] = std::exp(r.unit() * lh);
        sum += raw[i];
    }
    double scale = (double)n_nodes / sum;
    std::vector<int> len(n_lists);
    long long tot = 0;
    for (int i = 0; i < n_lists; i++) {
        long long L = std::llround(raw[i] * scale);
        if (L < 1) L = 1;
        len[i] = (int)L;
        tot += L;
    }
    long long diff = (long long)n_nodes - tot;
    for (int i = 0; diff != 0; i = (i + 1) % n_lists) {
        if (diff > 0) { len[i]++; diff--; }
        else if (len[i] > 1) { len[i]--; diff++; }
    }
    std::vector<int> perm(n_nodes);
    psi_gen_permutation(seed ^ 0xF0E5ull, n_nodes, perm.data());
    long long o = 0;
    for (int i = 0; i < n_lists; i++) {
        heads[i] = perm[o];
        for (int j = 0; j + 1 < len[i]; j++) succ[perm[o + j]] = perm[o + j + 1];
        succ[perm[o + len[i] - 1]] = -1;
        if (lengths_out) lengths_out[i] = len[i];
        o += len[i];
    }
    return PSI_OK;
}
