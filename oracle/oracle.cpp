// CPU restatement of the reference's strategy-improvement path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle: tests, the smoke check and
// the CPU-baseline leg of the benchmark call it to check or time the HIP
// product path.  Nothing in the product links against it.
//
// What it restates (reference file:line):
//   * PGSolver parser ............ game.cpp:51-107 (grammar and error strings)
//   * CPUParity semantics ........ cpu_parity.cpp:9-307
//       - init_strat ............. cpu_parity.cpp:276-286
//       - compute_valuation ...... cpu_parity.cpp:54-133 (recursive DFS, restated
//                                  here as an explicit stack; identical results)
//       - even_cycle ............. cpu_parity.cpp:35-52
//       - compare_valuation ...... cpu_parity.cpp:154-236
//       - switch_strategy ........ cpu_parity.cpp:239-263
//       - mark_solved ............ cpu_parity.cpp:265-274
//       - get_winning ............ cpu_parity.cpp:288-297
//       - reset_opponent_strategy  cpu_parity.cpp:300-307
//   * solve_si driver loop ....... main.cpp:20-55 (== gpumain.cpp:18-53)
//   * CPU_BellmanFord + solve_bf . cpu_bf.cpp:8-166, main.cpp:57-77 (dense rows,
//                                  in-place relaxation in vertex order)
//   * list ranking ............... the walk of cpu_list_ranking.cpp:151-204 with
//                                  unit weights is the sequential distance from
//                                  the head; restated as a brute-force walk.
//
// Two exact valuation representations are provided:
//   dense  : one int count per priority per vertex, as the reference stores it
//            (cpu_parity.cpp:9-32).  Needs n*(max_pri+1) ints.
//   sparse : a finite vertex's valuation is the multiset of priorities on its
//            strategy path to the sink; it is never materialised.  Two finite
//            valuations are compared by walking both paths up to their lowest
//            common ancestor in the sink tree, sorting the two private parts and
//            comparing them from the top.  This is the same total preorder as the
//            dense maxdiff scan (the common part cancels in every count).
// Both are exact integer computations; tests check that they agree.
//
// Parity pinning (how each function is checked, see tests/):
//   parser        - reference error strings + golden fixtures from the compiled
//                   reference harness (oracle/ref_harness.cpp)
//   solve (dense) - per-round strategy/label/valuation-digest traces of the
//                   compiled reference CPUParity on the committed fixtures
//   solve (sparse)- equality with dense on every fixture and on random games
//   list ranking  - closed-form invariants (a permutation list ranks to the
//                   inverse permutation) and hand-made forests
//   zielonka      - textbook algorithm; cross-checked against the winning
//                   regions of the reference fixtures
//   solve_bf      - the compiled reference's cpubf run (final strategy,
//                   labels, iter and br counts) on the committed fixtures

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

namespace {

static thread_local std::string g_err;

inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Weight of one occurrence of priority p in a valuation digest.
inline uint64_t pri_weight(int p) { return splitmix64((uint64_t)(uint32_t)p); }

// Order-independent hash of an int32 array (sum of per-slot mixes, mod 2^64).
uint64_t hash_i32(const int* a, int n)
{
    uint64_t h = 0;
    for (int i = 0; i < n; i++)
        h += splitmix64(((uint64_t)(uint32_t)i << 32) | (uint64_t)(uint32_t)a[i]);
    return h;
}

uint64_t hash_u64(const uint64_t* a, int n)
{
    uint64_t h = 0;
    for (int i = 0; i < n; i++)
        h += splitmix64(a[i] ^ splitmix64((uint64_t)i + 0x51ED27u));
    return h;
}

struct Game {
    int n = 0;
    int max_pri = 0;
    std::vector<int> off, tgt, owner, pri;
};

// ---------------------------------------------------------------------------
// Parser: restates Game::parse_pgsolver(istream&) (game.cpp:51-107).
// Returns 0 or -1 with g_err set to the reference's message.
int parse_text(const std::string& text, Game& g)
{
    std::istringstream in(text);
    std::string line;
    std::vector<std::vector<int>> edges;
    g = Game();
    if (!std::getline(in, line)) return 0;          // header line discarded (game.cpp:56)
    while (std::getline(in, line)) {
        std::stringstream ss(line);
        std::string token;
        if (!(ss >> token)) continue;               // empty line (game.cpp:62)
        int id;
        try {
            id = std::stoi(token);
        } catch (...) {
            g_err = "invalid id";
            return -1;
        }
        if (id < 0) { g_err = "invalid id"; return -1; }
        int priority;
        if (!(ss >> priority)) { g_err = "missing priority"; return -1; }
        int player;
        if (!(ss >> player)) { g_err = "missing player"; return -1; }
        if (player != 0 && player != 1) { g_err = "invalid player"; return -1; }
        int assigned = g.n++;
        edges.emplace_back();
        g.owner.push_back(player);
        g.pri.push_back(priority);
        if (priority > g.max_pri) g.max_pri = priority;
        if (id != assigned) { g_err = "vertices must be listed in order starting at 0!"; return -1; }
        for (;;) {
            int to;
            if (!(ss >> to)) { g_err = "missing successor"; return -1; }
            edges[id].push_back(to);
            char ch;
            if (!(ss >> ch)) { g_err = "missing ; to end line"; return -1; }
            if (ch != ',') break;
        }
        if (edges[id].empty()) { g_err = "sinks are not allowed!"; return -1; }
    }
    g.off.assign(g.n + 1, 0);
    for (int v = 0; v < g.n; v++) g.off[v + 1] = g.off[v] + (int)edges[v].size();
    g.tgt.clear();
    for (int v = 0; v < g.n; v++) g.tgt.insert(g.tgt.end(), edges[v].begin(), edges[v].end());
    return 0;
}

int load_csr(int n, const int* off, const int* tgt, const int* owner, const int* pri, Game& g)
{
    g = Game();
    g.n = n;
    g.off.assign(off, off + n + 1);
    g.tgt.assign(tgt, tgt + off[n]);
    g.owner.assign(owner, owner + n);
    g.pri.assign(pri, pri + n);
    g.max_pri = 0;
    for (int v = 0; v < n; v++) {
        if (pri[v] < 0) { g_err = "negative priority"; return -1; }
        if (owner[v] != 0 && owner[v] != 1) { g_err = "invalid player"; return -1; }
        if (off[v + 1] <= off[v]) { g_err = "sinks are not allowed!"; return -1; }
        if (pri[v] > g.max_pri) g.max_pri = pri[v];
    }
    for (int e = 0; e < off[n]; e++)
        if (tgt[e] < 0 || tgt[e] >= n) { g_err = "successor out of range"; return -1; }
    return 0;
}

// ---------------------------------------------------------------------------
// CPUParity restatement.
struct Parity {
    const Game& g;
    int n, P;
    bool dense;
    std::vector<int> strat, inf, solved, done;
    std::vector<int> vals;          // dense: n*P
    std::vector<int> depth;         // sparse: path length to the sink (finite vertices)
    std::vector<int> vstrat;        // sparse: the strategy the valuation was computed for
                                    // (switch_strategy edits strat while comparing)
    std::vector<uint64_t> digest;   // sum of pri_weight over the path (finite vertices)
    std::vector<int> bufA, bufB;    // sparse compare scratch
    std::vector<int> stack;

    Parity(const Game& g_, bool dense_) : g(g_), n(g_.n), P(g_.max_pri + 1), dense(dense_)
    {
        strat.assign(n, 0);
        inf.assign(n, 0);
        solved.assign(n, 0);
        done.assign(n, 0);
        depth.assign(n, 0);
        digest.assign(n, 0);
        if (dense) vals.assign((size_t)n * P, 0);
    }

    // cpu_parity.cpp:276-286
    void init_strat()
    {
        for (int v = 0; v < n; v++) strat[v] = g.owner[v] == 0 ? -1 : g.tgt[g.off[v]];
    }

    // cpu_parity.cpp:300-307
    void reset_opponent_strategy()
    {
        for (int v = 0; v < n; v++)
            if (g.owner[v] == 1) strat[v] = g.tgt[g.off[v]];
    }

    // cpu_parity.cpp:35-52: parity of the largest priority on the strategy cycle
    // through v (the running maximum starts at 0).
    bool even_cycle(int v) const
    {
        int cur = v, mp = 0;
        for (;;) {
            cur = strat[cur];
            if (g.pri[cur] > mp) mp = g.pri[cur];
            if (cur == v) break;
        }
        return mp % 2 == 0;
    }

    void set_sink_val(int v)
    {
        if (dense) {
            int* row = &vals[(size_t)v * P];
            std::fill(row, row + P, 0);
            row[g.pri[v]] = 1;
        }
        depth[v] = 1;
        digest[v] = pri_weight(g.pri[v]);
    }

    void set_succ_val(int v, int next)
    {
        if (dense) {
            int* row = &vals[(size_t)v * P];
            const int* nrow = &vals[(size_t)next * P];
            std::copy(nrow, nrow + P, row);
            row[g.pri[v]] += 1;
        }
        depth[v] = depth[next] + 1;
        digest[v] = digest[next] + pri_weight(g.pri[v]);
    }

    // Finishing step of the recursive compute_valuation(v) after the recursive
    // call on next = strat[v] has returned (cpu_parity.cpp:98-114).
    void finish(int v)
    {
        int next = strat[v];
        if (inf[next] != 0) {
            inf[v] = inf[next];
        } else {
            set_succ_val(v, next);
        }
        done[v] = 2;
    }

    // compute_valuation(v) of cpu_parity.cpp:54-114 with the recursion replaced
    // by an explicit stack of vertices whose done flag is 1.
    void valuate_from(int v0)
    {
        stack.clear();
        int v = v0;
        for (;;) {
            bool leaf = true;
            if (inf[v] == 1) {                      // cpu_parity.cpp:61-65
                done[v] = 2;
            } else if (strat[v] == -1) {            // cpu_parity.cpp:67-76
                set_sink_val(v);
                done[v] = 2;
            } else {
                int next = strat[v];
                if (done[next] == 1) {              // cpu_parity.cpp:81-89
                    inf[v] = even_cycle(v) ? 1 : -1;
                    done[v] = 2;
                } else if (done[next] == 0) {       // cpu_parity.cpp:92-96
                    done[v] = 1;
                    stack.push_back(v);
                    v = next;
                    leaf = false;
                } else {
                    finish(v);
                }
            }
            if (!leaf) continue;
            while (!stack.empty()) {
                int u = stack.back();
                stack.pop_back();
                finish(u);
            }
            return;
        }
    }

    // cpu_parity.cpp:116-133
    void compute_valuation()
    {
        for (int v = 0; v < n; v++) {
            if (!solved[v]) {
                done[v] = 0;
                inf[v] = 0;
            } else {
                done[v] = 2;
            }
        }
        for (int v = 0; v < n; v++)
            if (!done[v]) valuate_from(v);
        if (!dense) vstrat = strat;
    }

    // cpu_parity.cpp:154-192 (dense rows; row == nullptr is the zero vector)
    int compare_dense(const int* a, const int* b) const
    {
        int maxdiff = -1;
        for (int i = P - 1; i >= 0; i--) {
            int x = a ? a[i] : 0, y = b ? b[i] : 0;
            if (x != y) { maxdiff = i; break; }
        }
        if (maxdiff == -1) return 0;
        int vc = a ? a[maxdiff] : 0, uc = b ? b[maxdiff] : 0;
        if (maxdiff % 2 == 0) return vc > uc ? 1 : (vc == uc ? 0 : -1);
        return vc < uc ? 1 : (vc == uc ? 0 : -1);
    }

    // Exact compare of two finite valuations by their private path parts.
    int compare_sparse(int u, int w)
    {
        if (u == w) return 0;
        bufA.clear();
        bufB.clear();
        int a = u, b = w;
        int da = a < 0 ? 0 : depth[a], db = b < 0 ? 0 : depth[b];
        while (da > db) { bufA.push_back(g.pri[a]); a = vstrat[a]; da--; }
        while (db > da) { bufB.push_back(g.pri[b]); b = vstrat[b]; db--; }
        while (a != b) {
            bufA.push_back(g.pri[a]); a = vstrat[a];
            bufB.push_back(g.pri[b]); b = vstrat[b];
        }
        std::sort(bufA.begin(), bufA.end(), std::greater<int>());
        std::sort(bufB.begin(), bufB.end(), std::greater<int>());
        size_t i = 0;
        for (;; i++) {
            int x = i < bufA.size() ? bufA[i] : -1;
            int y = i < bufB.size() ? bufB[i] : -1;
            if (x == -1 && y == -1) return 0;
            if (x != y) {
                if (x > y) return x % 2 == 0 ? 1 : -1;
                return y % 2 == 0 ? -1 : 1;
            }
        }
    }

    // cpu_parity.cpp:194-236
    int compare(int v, int u)
    {
        if (v != -1 && u != -1 && inf[v] && inf[u]) {
            if (inf[v] == 1 && inf[u] == -1) return 1;
            if (inf[v] == -1 && inf[u] == 1) return -1;
            return 0;
        }
        if ((v != -1 && inf[v]) && (u == -1 || !inf[u])) return inf[v] == 1 ? 1 : -1;
        if ((u != -1 && inf[u]) && (v == -1 || !inf[v])) return inf[u] == 1 ? -1 : 1;
        if (dense)
            return compare_dense(v == -1 ? nullptr : &vals[(size_t)v * P],
                                 u == -1 ? nullptr : &vals[(size_t)u * P]);
        return compare_sparse(v, u);
    }

    // cpu_parity.cpp:239-263
    int switch_strategy(int player)
    {
        int total = 0;
        for (int v = 0; v < n; v++) {
            if (g.owner[v] != player) continue;
            int cur = strat[v];
            for (int e = g.off[v]; e < g.off[v + 1]; e++) {
                int u = g.tgt[e];
                int c = compare(u, cur);
                if ((player == 0 && c == 1) || (player == 1 && c == -1)) {
                    strat[v] = u;
                    cur = u;
                    total++;
                }
            }
        }
        return total;
    }

    // cpu_parity.cpp:265-274
    void mark_solved(int player)
    {
        for (int v = 0; v < n; v++) {
            if (player == 0 && inf[v] == 1) solved[v] = 1;
            else if (player == 1 && inf[v] == -1) solved[v] = 1;
        }
    }

    // Digest of the current valuation of v: 0 unless v is finite.
    uint64_t val_digest(int v) const { return inf[v] == 0 ? digest[v] : 0; }
};

struct Trace {
    uint64_t* hashes;   // 4 per round: strategy valued, solved, labels, digests
    int* strat_rounds;  // n per round (optional)
    int* inf_rounds;
    uint64_t* dig_rounds;
    int cap;
    int count;
};

void record(Parity& s, Trace* t, std::vector<uint64_t>& dig)
{
    if (!t) return;
    int r = t->count++;
    if (r >= t->cap) return;
    int n = s.n;
    for (int v = 0; v < n; v++) dig[v] = s.val_digest(v);
    if (t->hashes) {
        t->hashes[4 * r + 0] = hash_i32(s.strat.data(), n);
        t->hashes[4 * r + 1] = hash_i32(s.solved.data(), n);
        t->hashes[4 * r + 2] = hash_i32(s.inf.data(), n);
        t->hashes[4 * r + 3] = hash_u64(dig.data(), n);
    }
    if (t->strat_rounds) std::copy(s.strat.begin(), s.strat.end(), t->strat_rounds + (size_t)r * n);
    if (t->inf_rounds) std::copy(s.inf.begin(), s.inf.end(), t->inf_rounds + (size_t)r * n);
    if (t->dig_rounds) std::copy(dig.begin(), dig.end(), t->dig_rounds + (size_t)r * n);
}

// ---------------------------------------------------------------------------
// CPUList<ArrayVal> (cpu_list_ranking.cpp:15-276), the reference's threaded
// CPU path: each valuation is an Euler tour of the strategy forest ranked by
// Helman-JaJa, every phase forked onto T std::threads and joined as launch()
// does (:40-71).  Semantics (cpulist): a vertex whose strategy reaches the
// sink is finite, every other vertex +inf (broadcast_val, :208-239); solved
// flags are kept (CPUBV::mark_solved, cpu_bv.cpp:222-231) but the valuation
// never reads them.  Values are carried as (path length, digest) pairs, both
// additive like the reference's count vectors (+ at an up node, - at a down
// node), so the sublist walk, reduced-list prefix and broadcast are the
// reference's; compares use the exact sparse form on the valuation-time
// strategy tree.  The O(P^2) copy in update_val (:138-143) is not repeated.
struct ListSolver {
    Parity s;
    int T, n, l, S;
    std::vector<int> succ, vert_list, scratch, split_succ, split_inf;
    std::vector<int> dep_scr, split_dep, split_dep_sum;
    std::vector<uint64_t> dig_scr, split_dig, split_dig_sum;

    ListSolver(const Game& g, int threads) : s(g, false), T(std::max(1, threads))
    {
        n = g.n;
        l = 2 * n + 2;
        S = (l - 1) / 128 + 1;                      // cpu_list_ranking.cpp:21
        succ.assign(l, 0);
        vert_list.assign(n + 1, -1);
        scratch.assign(S, 0);
        split_succ.assign(S, 0);
        split_inf.assign(S, 1);
        dep_scr.assign(n, 0);
        dig_scr.assign(n, 0);
        split_dep.assign(S, 0);
        split_dep_sum.assign(S, 0);
        split_dig.assign(S, 0);
        split_dig_sum.assign(S, 0);
    }

    // cpu_list_ranking.cpp:40-71: T threads over [0, count) in equal ranges,
    // the last taking the remainder; serial when count < T.
    template <class F> void launch(F f, int count)
    {
        if (count < T) {
            if (count > 0) f(0, count - 1);
            return;
        }
        std::vector<std::thread> th;
        th.reserve(T);
        for (int i = 0; i < T; i++) {
            int start = i * (count / T);
            int end = i == T - 1 ? count - 1 : (i + 1) * (count / T) - 1;
            th.emplace_back(f, start, end);
        }
        for (auto& t : th) t.join();
    }

    // :74-85: up(i) = 2i+1 -> down(i) = 2i; the sink is vertex n, its down ends the list
    void init_list(int a, int b)
    {
        for (int i = a; i <= b; i++) {
            succ[2 * i + 1] = 2 * i;
            if (i == n) succ[2 * n] = -1;
            else vert_list[i] = -1;
        }
    }

    // :87-106: push up(i) onto the child list of its strategy successor
    void build_list(int a, int b)
    {
        for (int i = a; i <= b; i++) {
            int next = s.strat[i];
            int next_up = next == -1 ? 2 * n + 1 : 2 * next + 1;
            int my_up = 2 * i + 1;
            succ[2 * i] = __atomic_exchange_n(&succ[next_up], my_up, __ATOMIC_RELAXED);
        }
    }

    // :108-121: evenly spaced splitters, the last one at the head up(sink);
    // integer positions (the reference's float product picks the same kind of
    // spread; which nodes are splitters does not change any valuation)
    void pick_splitters(int a, int b)
    {
        for (int i = a; i <= b; i++) {
            int pos = i == S - 1 ? l - 1 : (int)((long long)(l - 1) * i / S);
            scratch[i] = succ[pos];
            succ[pos] = -100 - i;
        }
    }

    // :123-170: walk each sublist, carrying the running value
    void traverse_sublists(int a, int b)
    {
        for (int idx = a; idx <= b; idx++) {
            int cur = scratch[idx];
            int cd = 0;
            uint64_t cg = 0;
            while (cur >= 0) {
                int v = cur / 2;
                if (v != n) {                        // the sink has no priority (:127-128)
                    if (cur & 1) {                   // up: add, and record at v
                        cd += 1;
                        cg += pri_weight(s.g.pri[v]);
                        dep_scr[v] = cd;
                        dig_scr[v] = cg;
                        vert_list[v] = idx;
                    } else {                         // down: remove
                        cd -= 1;
                        cg -= pri_weight(s.g.pri[v]);
                    }
                }
                cur = succ[cur];
            }
            split_dep[idx] = cd;
            split_dig[idx] = cg;
            split_succ[idx] = cur;
        }
    }

    // :172-204: serial walk of the reduced list from the head's sublist
    void process_reduced()
    {
        std::fill(split_inf.begin(), split_inf.end(), 1);
        std::fill(split_dep_sum.begin(), split_dep_sum.end(), 0);
        std::fill(split_dig_sum.begin(), split_dig_sum.end(), 0);
        int cur = S - 1, d = 0;
        uint64_t g = 0;
        while (cur != -1) {
            split_dep_sum[cur] = d;
            split_dig_sum[cur] = g;
            d += split_dep[cur];
            g += split_dig[cur];
            split_inf[cur] = 0;
            int next = split_succ[cur];
            if (next == -1) break;
            cur = -(next + 100);
        }
    }

    // :208-239
    void broadcast_val(int a, int b)
    {
        for (int v = a; v <= b; v++) {
            if (s.inf[v] == -1) continue;
            int k = vert_list[v];
            if (k == -1) { s.inf[v] = 1; continue; }
            if (split_inf[k] == 0) {
                s.depth[v] = split_dep_sum[k] + dep_scr[v];
                s.digest[v] = split_dig_sum[k] + dig_scr[v];
                s.inf[v] = 0;
            } else {
                s.inf[v] = 1;
            }
        }
    }

    // :242-255
    void compute_valuation()
    {
        launch([this](int a, int b) { init_list(a, b); }, n + 1);
        launch([this](int a, int b) { build_list(a, b); }, n);
        launch([this](int a, int b) { pick_splitters(a, b); }, S);
        launch([this](int a, int b) { traverse_sublists(a, b); }, S);
        process_reduced();
        launch([this](int a, int b) { broadcast_val(a, b); }, n);
        s.vstrat = s.strat;
    }

    // Exact compare of two finite valuations (private path parts), with
    // per-thread scratch.
    int compare_sparse(int u, int w, std::vector<int>& A, std::vector<int>& B) const
    {
        if (u == w) return 0;
        A.clear();
        B.clear();
        int a = u, b = w;
        int da = a < 0 ? 0 : s.depth[a], db = b < 0 ? 0 : s.depth[b];
        while (da > db) { A.push_back(s.g.pri[a]); a = s.vstrat[a]; da--; }
        while (db > da) { B.push_back(s.g.pri[b]); b = s.vstrat[b]; db--; }
        while (a != b) {
            A.push_back(s.g.pri[a]); a = s.vstrat[a];
            B.push_back(s.g.pri[b]); b = s.vstrat[b];
        }
        std::sort(A.begin(), A.end(), std::greater<int>());
        std::sort(B.begin(), B.end(), std::greater<int>());
        for (size_t i = 0;; i++) {
            int x = i < A.size() ? A[i] : -1, y = i < B.size() ? B[i] : -1;
            if (x == -1 && y == -1) return 0;
            if (x != y) {
                if (x > y) return x % 2 == 0 ? 1 : -1;
                return y % 2 == 0 ? -1 : 1;
            }
        }
    }

    // CPUBV::compare_valuation (cpu_bv.cpp:144-186)
    int compare(int v, int u, std::vector<int>& A, std::vector<int>& B) const
    {
        const std::vector<int>& inf = s.inf;
        if (v != -1 && u != -1 && inf[v] && inf[u]) {
            if (inf[v] == 1 && inf[u] == -1) return 1;
            if (inf[v] == -1 && inf[u] == 1) return -1;
            return 0;
        }
        if ((v != -1 && inf[v]) && (u == -1 || !inf[u])) return inf[v] == 1 ? 1 : -1;
        if ((u != -1 && inf[u]) && (v == -1 || !inf[v])) return inf[u] == 1 ? -1 : 1;
        return compare_sparse(v, u, A, B);
    }

    // CPUList::switch_strategy (:262-276) over CPUBV::switch_vertex
    // (cpu_bv.cpp:188-207); per-thread counts are added after the join.
    int switch_strategy(int player)
    {
        std::vector<long long> part(T, 0);
        std::atomic<int> slot{0};
        launch([&](int a, int b) {
            std::vector<int> A, B;
            long long total = 0;
            for (int v = a; v <= b; v++) {
                if (s.g.owner[v] != player) continue;
                int cur = s.strat[v];
                for (int e = s.g.off[v]; e < s.g.off[v + 1]; e++) {
                    int u = s.g.tgt[e];
                    int c = compare(u, cur, A, B);
                    if ((player == 0 && c == 1) || (player == 1 && c == -1)) {
                        s.strat[v] = u;
                        cur = u;
                        total++;
                    }
                }
            }
            part[slot.fetch_add(1) % T] += total;
        }, n);
        long long t = 0;
        for (long long x : part) t += x;
        return (int)std::min<long long>(t, 1 << 30);
    }
};

// ---------------------------------------------------------------------------
// CPU_BellmanFord<ArrayVal> (cpu_bf.cpp:8-166) driven by solve_bf (main.cpp:57-77),
// on top of CPUBV<ArrayVal> (cpu_bv.cpp), whose init_strat / compute_valuation /
// compare / switch / mark_solved are CPUParity's (the Parity struct, dense
// rows) except that compute_valuation aborts on an odd strategy cycle
// (cpu_bv.cpp:73-84).
struct BellmanFord {
    Parity s;
    std::vector<int> bf_inf, succ;
    bool init = false;
    int status = 0;                  // 1: odd cycle (the reference exits), 3: a relaxation hit n rounds
    std::vector<int> tmp;

    explicit BellmanFord(const Game& g) : s(g, true)
    {
        bf_inf.assign(s.n, 1);       // cpu_bf.cpp:13-17 (p1_vertices is never read)
        succ.assign(s.n, 0);
        tmp.assign(s.P, 0);
    }

    // ArrayVal: a row plus the priority of v (add_vertex, valuation.cpp)
    void row_plus(const int* from, int v, int* out) const
    {
        if (from) std::copy(from, from + s.P, out);
        else std::fill(out, out + s.P, 0);
        out[s.g.pri[v]] += 1;
    }

    // cpu_bf.cpp:25-49
    void best_response(int* br)
    {
        if (!init) {
            for (;;) {
                s.compute_valuation();
                for (int v = 0; v < s.n; v++)
                    if (!s.solved[v] && s.inf[v] == -1) { status = 1; return; }   // cpu_bv.cpp:79-83
                s.mark_solved(1);
                if (s.switch_strategy(1) == 0) break;
            }
            init = true;
            return;
        }
        std::fill(bf_inf.begin(), bf_inf.end(), 1);
        best_response_alg(br);
    }

    // cpu_bf.cpp:51-166: in-place (Gauss-Seidel) relaxation in vertex order
    void best_response_alg(int* br)
    {
        const int n = s.n, P = s.P;
        std::fill(s.vals.begin(), s.vals.end(), 0);
        bool done = false;
        for (int i = 0; i < n; i++) {
            (*br)++;
            done = true;
            for (int v = 0; v < n; v++) {
                if (s.solved[v]) continue;
                int* vrow = &s.vals[(size_t)v * P];
                if (s.g.owner[v] == 0) {
                    const int u = s.strat[v];
                    if (u != -1 && bf_inf[u]) continue;
                    if (bf_inf[v]) { bf_inf[v] = 0; done = false; }
                    row_plus(u == -1 ? nullptr : &s.vals[(size_t)u * P], v, tmp.data());
                    if (!std::equal(tmp.begin(), tmp.end(), vrow)) {
                        std::copy(tmp.begin(), tmp.end(), vrow);
                        done = false;
                    }
                    continue;
                }
                for (int e = s.g.off[v]; e < s.g.off[v + 1]; e++) {
                    const int u = s.g.tgt[e];
                    if (bf_inf[u]) continue;
                    if (bf_inf[v]) {
                        row_plus(&s.vals[(size_t)u * P], v, vrow);
                        bf_inf[v] = 0;
                        succ[v] = u;
                        done = false;
                        continue;
                    }
                    row_plus(&s.vals[(size_t)u * P], v, tmp.data());
                    if (s.compare_dense(vrow, tmp.data()) == 1) {
                        std::copy(tmp.begin(), tmp.end(), vrow);
                        succ[v] = u;
                        done = false;
                    }
                }
            }
            if (done) break;
        }
        if (!done) status = 3;
        for (int v = 0; v < n; v++) {
            if (bf_inf[v] && !s.solved[v]) s.inf[v] = 1;
            if (s.g.owner[v] == 1) s.strat[v] = succ[v];
        }
    }
};

// ---------------------------------------------------------------------------
// Zielonka's recursive algorithm (max-parity: the largest priority seen
// infinitely often decides; even -> player 0).  Independent check on winners.
struct Zielonka {
    const Game& g;
    std::vector<std::vector<int>> pred;
    explicit Zielonka(const Game& g_) : g(g_), pred(g_.n)
    {
        for (int v = 0; v < g.n; v++)
            for (int e = g.off[v]; e < g.off[v + 1]; e++) pred[g.tgt[e]].push_back(v);
    }

    // Attractor of U for player i inside the subgame marked by in[].
    std::vector<char> attr(const std::vector<char>& in, const std::vector<int>& U, int i)
    {
        int n = g.n;
        std::vector<char> A(n, 0);
        std::vector<int> cnt(n, 0), q;
        for (int v = 0; v < n; v++) {
            if (!in[v]) continue;
            int c = 0;
            for (int e = g.off[v]; e < g.off[v + 1]; e++) c += in[g.tgt[e]] ? 1 : 0;
            cnt[v] = c;
        }
        for (int u : U) if (!A[u]) { A[u] = 1; q.push_back(u); }
        for (size_t h = 0; h < q.size(); h++) {
            int u = q[h];
            for (int p : pred[u]) {
                if (!in[p] || A[p]) continue;
                if (g.owner[p] == i) { A[p] = 1; q.push_back(p); }
                else if (--cnt[p] == 0) { A[p] = 1; q.push_back(p); }
            }
        }
        return A;
    }

    // Returns winner per vertex (0/1) for vertices in in[]; others untouched.
    void solve(std::vector<char> in, std::vector<int>& win)
    {
        int n = g.n;
        int maxp = -1;
        for (int v = 0; v < n; v++) if (in[v] && g.pri[v] > maxp) maxp = g.pri[v];
        if (maxp < 0) return;
        int i = maxp % 2;
        std::vector<int> U;
        for (int v = 0; v < n; v++) if (in[v] && g.pri[v] == maxp) U.push_back(v);
        std::vector<char> A = attr(in, U, i);
        std::vector<char> sub(n, 0);
        for (int v = 0; v < n; v++) sub[v] = in[v] && !A[v];
        std::vector<int> w1(n, -1);
        solve(sub, w1);
        std::vector<int> opp;
        for (int v = 0; v < n; v++) if (sub[v] && w1[v] == 1 - i) opp.push_back(v);
        if (opp.empty()) {
            for (int v = 0; v < n; v++) if (in[v]) win[v] = i;
            return;
        }
        std::vector<char> B = attr(in, opp, 1 - i);
        std::vector<char> sub2(n, 0);
        for (int v = 0; v < n; v++) sub2[v] = in[v] && !B[v];
        std::vector<int> w2(n, -1);
        solve(sub2, w2);
        for (int v = 0; v < n; v++) {
            if (!in[v]) continue;
            win[v] = B[v] ? 1 - i : w2[v];
        }
    }
};

} // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

uint64_t orc_splitmix64(uint64_t x) { return splitmix64(x); }
uint64_t orc_hash_i32(const int* a, int n) { return hash_i32(a, n); }
uint64_t orc_hash_u64(const uint64_t* a, int n) { return hash_u64(a, n); }

// Parse PGSolver text.  On success returns 0 and fills *n, *m, *max_pri; the
// arrays are then read with orc_parsed_arrays (same thread).
static thread_local Game t_parsed;
int orc_parse(const char* text, long len, int* n, int* m, int* max_pri)
{
    int rc = parse_text(std::string(text, (size_t)len), t_parsed);
    if (rc) return rc;
    *n = t_parsed.n;
    *m = (int)t_parsed.tgt.size();
    *max_pri = t_parsed.max_pri;
    return 0;
}

void orc_parsed_arrays(int* owner, int* pri, int* off, int* tgt)
{
    std::copy(t_parsed.owner.begin(), t_parsed.owner.end(), owner);
    std::copy(t_parsed.pri.begin(), t_parsed.pri.end(), pri);
    std::copy(t_parsed.off.begin(), t_parsed.off.end(), off);
    std::copy(t_parsed.tgt.begin(), t_parsed.tgt.end(), tgt);
}

// Full solve_si (main.cpp:20-55) under CPUParity semantics.
//   rep: 0 dense, 1 sparse.   reset: the cpu-reset mode.
//   max_rounds: cap on valuation rounds (br); returns 1 if hit.
//   hashes (4*cap u64), strat_rounds/inf_rounds (cap*n i32), dig_rounds (cap*n u64)
//   are optional per-round traces.
// Outputs: final strategy, labels (inf), solved flags, iter and br counts,
// and the solve time in seconds (the region test() times, main.cpp:80-95).
int orc_solve(int n, const int* off, const int* tgt, const int* owner, const int* pri,
              int rep, int reset, int max_rounds,
              int* strat_out, int* inf_out, int* solved_out, int* iter_out, int* br_out,
              uint64_t* hashes, int* strat_rounds, int* inf_rounds, uint64_t* dig_rounds,
              int trace_cap, int* n_rounds, double* seconds)
{
    Game g;
    if (load_csr(n, off, tgt, owner, pri, g)) return -1;
    if (rep == 0 && (double)n * (g.max_pri + 1) > 4.0e9) { g_err = "dense valuation too large"; return -1; }
    Parity s(g, rep == 0);
    Trace tr{hashes, strat_rounds, inf_rounds, dig_rounds, trace_cap, 0};
    Trace* tp = (hashes || strat_rounds || inf_rounds || dig_rounds) ? &tr : nullptr;
    std::vector<uint64_t> dig(n);
    int br = 0, iter = 0, capped = 0;
    auto t0 = std::chrono::steady_clock::now();
    s.init_strat();
    bool first = true;
    for (;;) {
        if (!first && reset) s.reset_opponent_strategy();
        for (;;) {
            if (max_rounds > 0 && br >= max_rounds) { capped = 1; break; }
            s.compute_valuation();          // compute_first_val == compute_valuation here
            record(s, tp, dig);
            br++;
            s.mark_solved(1);
            if (s.switch_strategy(1) == 0) break;
        }
        if (capped) break;
        first = false;
        s.mark_solved(0);
        iter++;
        if (s.switch_strategy(0) == 0) break;
    }
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    if (strat_out) std::copy(s.strat.begin(), s.strat.end(), strat_out);
    if (inf_out) std::copy(s.inf.begin(), s.inf.end(), inf_out);
    if (solved_out) std::copy(s.solved.begin(), s.solved.end(), solved_out);
    if (iter_out) *iter_out = iter;
    if (br_out) *br_out = br;
    if (n_rounds) *n_rounds = tp ? tr.count : br;
    return capped;
}

// solve_si (main.cpp:20-55, the cpulist mode of main.cpp:178-191) with the
// CPUList restatement on `threads` threads.  Returns 1 if max_rounds was
// reached (cpulist need not terminate: on games where Odd closes an odd
// cycle its labels oscillate).  Outputs and traces as orc_solve; seconds =
// the solve region.
int orc_solve_list(int n, const int* off, const int* tgt, const int* owner, const int* pri,
                   int threads, int max_rounds, int* strat_out, int* inf_out, int* solved_out,
                   int* iter_out, int* br_out, uint64_t* hashes, int trace_cap, int* n_rounds,
                   double* seconds)
{
    Game g;
    if (load_csr(n, off, tgt, owner, pri, g)) return -1;
    ListSolver L(g, threads);
    Parity& s = L.s;
    Trace tr{hashes, nullptr, nullptr, nullptr, trace_cap, 0};
    Trace* tp = hashes ? &tr : nullptr;
    std::vector<uint64_t> dig(n);
    int br = 0, iter = 0, capped = 0;
    auto t0 = std::chrono::steady_clock::now();
    s.init_strat();                                  // CPUBV::init_strat (cpu_bv.cpp:233-243)
    for (;;) {
        for (;;) {
            if (max_rounds > 0 && br >= max_rounds) { capped = 1; break; }
            L.compute_valuation();                   // compute_first_val is the same (:257-260)
            record(s, tp, dig);
            br++;
            s.mark_solved(1);
            if (L.switch_strategy(1) == 0) break;
        }
        if (capped) break;
        s.mark_solved(0);
        iter++;
        if (L.switch_strategy(0) == 0) break;
    }
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    if (strat_out) std::copy(s.strat.begin(), s.strat.end(), strat_out);
    if (inf_out) std::copy(s.inf.begin(), s.inf.end(), inf_out);
    if (solved_out) std::copy(s.solved.begin(), s.solved.end(), solved_out);
    if (iter_out) *iter_out = iter;
    if (br_out) *br_out = br;
    if (n_rounds) *n_rounds = tp ? tr.count : br;
    return capped;
}

// solve_si (main.cpp:20-55, the cpubv mode of main.cpp:163-167) with
// CPUBV<ArrayVal>: CPUParity's init_strat, valuation, compare, switch and
// mark_solved (cpu_bv.cpp:30-254 restate them for a templated value type),
// except that compute_valuation stops the program when a strategy cycle's
// largest priority is odd (cpu_bv.cpp:73-84: "Game contains an odd cycle that
// uses only odd nodes", exit(1)).  status: 0 completed, 1 stopped there; then
// iter and br count the outer iterations and valuations completed before the
// stopping one (the reference prints neither).  Completed games are CPUParity
// solves in which no -1 label ever appears.
int orc_solve_bv(int n, const int* off, const int* tgt, const int* owner, const int* pri,
                 int rep, int max_rounds, int* strat_out, int* inf_out, int* solved_out,
                 int* iter_out, int* br_out, int* status_out)
{
    Game g;
    if (load_csr(n, off, tgt, owner, pri, g)) return -1;
    if (rep == 0 && (double)n * (g.max_pri + 1) > 4.0e9) { g_err = "dense valuation too large"; return -1; }
    Parity s(g, rep == 0);
    int br = 0, iter = 0, status = 0, capped = 0;
    s.init_strat();
    for (;;) {
        for (;;) {
            if (max_rounds > 0 && br >= max_rounds) { capped = 1; break; }
            s.compute_valuation();
            for (int v = 0; v < n && !status; v++)
                if (!s.solved[v] && s.inf[v] == -1) status = 1;     // cpu_bv.cpp:79-83
            if (status) break;
            br++;
            s.mark_solved(1);
            if (s.switch_strategy(1) == 0) break;
        }
        if (capped || status) break;
        s.mark_solved(0);
        iter++;
        if (s.switch_strategy(0) == 0) break;
    }
    if (strat_out) std::copy(s.strat.begin(), s.strat.end(), strat_out);
    if (inf_out) std::copy(s.inf.begin(), s.inf.end(), inf_out);
    if (solved_out) std::copy(s.solved.begin(), s.solved.end(), solved_out);
    if (iter_out) *iter_out = iter;
    if (br_out) *br_out = br;
    if (status_out) *status_out = status;
    return capped;
}

// solve_bf (main.cpp:57-77) with CPU_BellmanFord<ArrayVal>.  status: 0 ok,
// 1 odd strategy cycle in the first best response (the reference prints
// "Game contains an odd cycle that uses only odd nodes" and exits), 2 stopped
// at max_iters, 3 some relaxation ran n rounds without converging.
// dig_out: digest of each final valuation (sum count[p] * pri_weight(p), 0 if
// infinite).
int orc_solve_bf(int n, const int* off, const int* tgt, const int* owner, const int* pri,
                 int max_iters, int* strat_out, int* inf_out, int* solved_out, uint64_t* dig_out,
                 int* iter_out, int* br_out, int* status_out)
{
    Game g;
    if (load_csr(n, off, tgt, owner, pri, g)) return -1;
    if ((double)n * (g.max_pri + 1) > 4.0e9) { g_err = "dense valuation too large"; return -1; }
    BellmanFord bf(g);
    Parity& s = bf.s;
    int br = 0, iter = 0;
    s.init_strat();
    for (;;) {
        if (max_iters > 0 && iter >= max_iters) { bf.status = bf.status ? bf.status : 2; break; }
        bf.best_response(&br);
        if (bf.status == 1) break;
        s.mark_solved(0);
        iter++;
        if (s.switch_strategy(0) == 0) break;
    }
    if (strat_out) std::copy(s.strat.begin(), s.strat.end(), strat_out);
    if (inf_out) std::copy(s.inf.begin(), s.inf.end(), inf_out);
    if (solved_out) std::copy(s.solved.begin(), s.solved.end(), solved_out);
    if (dig_out)
        for (int v = 0; v < n; v++) {
            uint64_t d = 0;
            if (s.inf[v] == 0)
                for (int p = 0; p < s.P; p++) d += (uint64_t)s.vals[(size_t)v * s.P + p] * pri_weight(p);
            dig_out[v] = d;
        }
    if (iter_out) *iter_out = iter;
    if (br_out) *br_out = br;
    if (status_out) *status_out = bf.status;
    return 0;
}

// One valuation of a given strategy with a given solved set (labels and
// digests out).  Used to pin single kernels: valuation only.
int orc_valuate(int n, const int* off, const int* tgt, const int* owner, const int* pri,
                const int* strat, const int* solved, const int* inf_in,
                int* inf_out, uint64_t* dig_out, int* depth_out)
{
    Game g;
    if (load_csr(n, off, tgt, owner, pri, g)) return -1;
    Parity s(g, false);
    std::copy(strat, strat + n, s.strat.begin());
    std::copy(solved, solved + n, s.solved.begin());
    std::copy(inf_in, inf_in + n, s.inf.begin());
    s.compute_valuation();
    for (int v = 0; v < n; v++) {
        inf_out[v] = s.inf[v];
        dig_out[v] = s.val_digest(v);
        if (depth_out) depth_out[v] = s.inf[v] == 0 ? s.depth[v] : 0;
    }
    return 0;
}

// One switch_strategy(player) on a valued state; returns the switch count.
int orc_switch(int n, const int* off, const int* tgt, const int* owner, const int* pri,
               int rep, const int* strat, const int* solved, const int* inf_in,
               int player, int* strat_out)
{
    Game g;
    if (load_csr(n, off, tgt, owner, pri, g)) return -1;
    Parity s(g, rep == 0);
    std::copy(strat, strat + n, s.strat.begin());
    std::copy(solved, solved + n, s.solved.begin());
    std::copy(inf_in, inf_in + n, s.inf.begin());
    s.compute_valuation();
    int c = s.switch_strategy(player);
    std::copy(s.strat.begin(), s.strat.end(), strat_out);
    return c;
}

// Pairwise compare of two vertices (-1 = sink) after valuing `strat`.
// Fills out[k] = compare(us[k], ws[k]) for k < q.
int orc_compare_pairs(int n, const int* off, const int* tgt, const int* owner, const int* pri,
                      int rep, const int* strat, const int* us, const int* ws, int q, int* out)
{
    Game g;
    if (load_csr(n, off, tgt, owner, pri, g)) return -1;
    Parity s(g, rep == 0);
    std::copy(strat, strat + n, s.strat.begin());
    s.compute_valuation();
    for (int k = 0; k < q; k++) out[k] = s.compare(us[k], ws[k]);
    return 0;
}

// Brute-force list ranking.  succ[i] in [-1, n); heads lists the first node of
// every list.  rank[i] = distance from its head, or -1 if no head reaches i.
// Returns -1 if a walk revisits a node (cycle reachable from a head).
int orc_list_rank(const int* succ, int n, const int* heads, int n_heads, int* rank)
{
    for (int i = 0; i < n; i++) rank[i] = -1;
    for (int h = 0; h < n_heads; h++) {
        int cur = heads[h], r = 0;
        while (cur != -1) {
            if (cur < 0 || cur >= n) { g_err = "successor out of range"; return -1; }
            if (rank[cur] != -1) { g_err = "node reached twice"; return -1; }
            rank[cur] = r++;
            cur = succ[cur];
        }
    }
    return 0;
}

// Weighted prefix along lists: out[i] = sum of w over the nodes from the head
// up to and including i (int64), or 0 if unreached.
int orc_list_scan(const int* succ, int n, const int* heads, int n_heads, const int64_t* w,
                  int64_t* out)
{
    std::vector<char> seen(n, 0);
    for (int i = 0; i < n; i++) out[i] = 0;
    for (int h = 0; h < n_heads; h++) {
        int cur = heads[h];
        int64_t acc = 0;
        while (cur != -1) {
            if (seen[cur]) { g_err = "node reached twice"; return -1; }
            seen[cur] = 1;
            acc += w[cur];
            out[cur] = acc;
            cur = succ[cur];
        }
    }
    return 0;
}

// Winning regions by Zielonka's algorithm: win[v] = 0 or 1.
int orc_zielonka(int n, const int* off, const int* tgt, const int* owner, const int* pri, int* win)
{
    Game g;
    if (load_csr(n, off, tgt, owner, pri, g)) return -1;
    Zielonka z(g);
    std::vector<int> w(n, -1);
    z.solve(std::vector<char>(n, 1), w);
    std::copy(w.begin(), w.end(), win);
    return 0;
}

} // extern "C"
