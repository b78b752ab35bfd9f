// Driver for the reference's own CPU solvers, used only to write golden
// fixtures in the build container.  It is compiled together with the
// reference's CPU sources where they lie (see oracle/Makefile, target `ref`);
// none of those sources are copied.  The reference's main.cpp needs boost, so
// this file restates its solve_si / solve_bf loops (main.cpp:20-77) with a
// round cap and prints one JSON object per valuation round.
//
// usage: ref_harness <cpu|cpu-reset|cpubv|cpulist|cpubf> <game.gm> [max_rounds] [threads]
//
// Output (JSON lines):
//   {"round":r,"strat":[..],"solved":[..],"inf":[..],"dig":[..]}   per valuation
//   {"final":1,"status":"ok"|"capped","iter":i,"br":b,"strat":[..],"inf":[..],
//    "w0":[..],"w1":[..]}
// dig[v] = sum_p count_v[p] * splitmix64(p) mod 2^64 for finite v, else 0.

#include "game.h"
#include "cpu_parity.h"
#include "cpu_bv.h"
#include "cpu_list_ranking.h"
#include "cpu_bf.h"
#include "valuation.h"

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <sstream>
#include <string>
#include <vector>

static uint64_t mix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void put_ints(const char* key, const std::vector<int>& a, bool comma = true)
{
    printf("\"%s\":[", key);
    for (size_t i = 0; i < a.size(); i++) printf(i ? ",%d" : "%d", a[i]);
    printf(comma ? "]," : "]");
}

static void put_u64(const char* key, const std::vector<uint64_t>& a)
{
    printf("\"%s\":[", key);
    for (size_t i = 0; i < a.size(); i++) printf(i ? ",%llu" : "%llu", (unsigned long long)a[i]);
    printf("]");
}

// Digest of a dense count vector given as ArrayVal::to_string() text "[c0 c1 ... ]".
static uint64_t digest_from_text(const std::string& s)
{
    std::string body = s;
    for (char& c : body) if (c == '[' || c == ']') c = ' ';
    std::istringstream in(body);
    uint64_t d = 0;
    long long c;
    int p = 0;
    while (in >> c) {
        d += (uint64_t)c * mix64((uint64_t)p);
        p++;
    }
    return d;
}

struct ParityProbe : public CPUParity {
    explicit ParityProbe(Game& g) : CPUParity(g) {}
    void dump(int r)
    {
        int n = g.get_n_vertices();
        std::vector<uint64_t> dig(n, 0);
        for (int v = 0; v < n; v++) {
            if (infinite[v] != 0) continue;
            uint64_t d = 0;
            for (size_t p = 0; p < vals[v].size(); p++) d += (uint64_t)vals[v][p] * mix64(p);
            dig[v] = d;
        }
        printf("{\"round\":%d,", r);
        put_ints("strat", strat);
        put_ints("solved", solved);
        put_ints("inf", infinite);
        put_u64("dig", dig);
        printf("}\n");
        fflush(stdout);
    }
    std::vector<int>& S() { return strat; }
    std::vector<int>& I() { return infinite; }
};

template <typename Base> struct BVProbe : public Base {
    template <typename... A> explicit BVProbe(A&&... a) : Base(a...) {}
    void dump(int r)
    {
        int n = this->g.get_n_vertices();
        std::vector<uint64_t> dig(n, 0);
        for (int v = 0; v < n; v++) {
            if (this->infinite[v] != 0) continue;
            dig[v] = digest_from_text(this->vals[v].to_string());
        }
        printf("{\"round\":%d,", r);
        put_ints("strat", this->strat);
        put_ints("solved", this->solved);
        put_ints("inf", this->infinite);
        put_u64("dig", dig);
        printf("}\n");
        fflush(stdout);
    }
    std::vector<int>& S() { return this->strat; }
    std::vector<int>& I() { return this->infinite; }
};

template <typename T> static void finish(T& si, const char* status, int iter, int br)
{
    std::vector<int> w0, w1;
    si.get_winning(w0, 0);
    si.get_winning(w1, 1);
    printf("{\"final\":1,\"status\":\"%s\",\"iter\":%d,\"br\":%d,", status, iter, br);
    put_ints("strat", si.S());
    put_ints("inf", si.I());
    put_ints("w0", w0);
    put_ints("w1", w1, false);
    printf("}\n");
    fflush(stdout);
}

// Restatement of solve_si (main.cpp:20-55) with a cap on valuation rounds.
template <typename T> static void run_si(T& si, bool reset, int max_rounds)
{
    int br = 0, iter = 0;
    si.init_strat();
    bool first = true;
    for (;;) {
        if (!first && reset) si.reset_opponent_strategy();
        for (;;) {
            if (br >= max_rounds) { finish(si, "capped", iter, br); return; }
            if (first) si.compute_first_val();
            else si.compute_valuation();
            si.dump(br);
            br++;
            si.mark_solved(1);
            if (si.switch_strategy(1) == 0) break;
        }
        first = false;
        si.mark_solved(0);
        iter++;
        if (si.switch_strategy(0) == 0) break;
    }
    finish(si, "ok", iter, br);
}

// Restatement of solve_bf (main.cpp:57-77); br is counted inside best_response.
template <typename T> static void run_bf(T& si, int max_iters)
{
    int br = 0, iter = 0;
    si.init_strat();
    for (;;) {
        if (iter >= max_iters) { finish(si, "capped", iter, br); return; }
        si.best_response(&br);
        si.dump(iter);
        si.mark_solved(0);
        iter++;
        if (si.switch_strategy(0) == 0) break;
    }
    finish(si, "ok", iter, br);
}

int main(int argc, char** argv)
{
    if (argc < 3) {
        fprintf(stderr, "usage: %s <cpu|cpu-reset|cpubv|cpulist|cpubf> <game.gm> [max_rounds] [threads]\n", argv[0]);
        return 2;
    }
    std::string mode = argv[1];
    int max_rounds = argc > 3 ? atoi(argv[3]) : 100000;
    int threads = argc > 4 ? atoi(argv[4]) : 8;
    Game g;
    try {
        g.parse_pgsolver(std::string(argv[2]));
    } catch (const char* err) {
        printf("{\"parse_error\":\"%s\"}\n", err);
        return 3;
    }
    if (mode == "cpu" || mode == "cpu-reset") {
        ParityProbe si(g);
        run_si(si, mode == "cpu-reset", max_rounds);
    } else if (mode == "cpubv") {
        BVProbe<CPUBV<ArrayVal>> si(g);
        run_si(si, false, max_rounds);
    } else if (mode == "cpulist") {
        BVProbe<CPUList<ArrayVal>> si(g, threads);
        run_si(si, false, max_rounds);
    } else if (mode == "cpubf") {
        BVProbe<CPU_BellmanFord<ArrayVal>> si(g);
        run_bf(si, max_rounds);
    } else {
        fprintf(stderr, "unknown mode %s\n", mode.c_str());
        return 2;
    }
    return 0;
}
