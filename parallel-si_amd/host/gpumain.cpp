// gpuparsi: the reference's GPU command line (gpumain.cpp:102-146) on MI355X.
//
//   gpuparsi [options] <gpu|gpulist|gpubf|gpubv> <game.gm | game.gm.gz>
//
// Prints "<seconds> <iter_count> <br_count>" for the solve_si region, as the
// reference's test() does (gpumain.cpp:56-71).  By default the whole solve_si
// loop runs on the device; --host-loop drives it from the host through the
// StrategyImprovement calls exactly as gpumain.cpp:18-53 does.  gpubf runs
// solve_bf (main.cpp:57-77, the CPU command line's cpubf) on the device; an
// odd strategy cycle in its first best response prints the reference's
// message and exits with status 1, as cpu_bv.cpp:73-84 does; its br counts
// device valuation rounds, not the reference's relaxation sweeps (said on
// stderr, and as "br_kind" in the JSON line).  gpubv runs solve_si with CPUBV
// semantics (main.cpp:163-167, the CPU command line's cpubv) on the device:
// the same stop on the first odd strategy cycle of any valuation.
// Options: --host-loop, --reset (the cpu-reset schedule), --winning (print W0
// and W1), --strategy (print the final strategy), --json (one JSON line),
// --cpulist[=N] (the CPU command line's cpulist semantics, CPUList: vertices
// not reaching the sink are +inf; such runs need not end, so they stop after
// N valuations, default 100000, with a note on stderr).
// Exit status: 0 ok, 1 usage, 255 parse error ("parsing error: ..." on stderr,
// as gpumain.cpp:124-127), 2 device error.
#include "game.h"
#include "gpu_game.h"
#include "gpu_parity.h"
#include "gpu_list_ranking.h"

#include <zlib.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

static bool ends_with(const std::string& s, const char* suf)
{
    size_t n = strlen(suf);
    return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

// bzip2's stdio-style reading interface (libbz2.so.1, the bzlib 1.0 API's
// zlib-compatibility functions).  The image has the library but not its
// header, so the three entry points used are declared by their documented
// signatures.
extern "C" {
typedef void BZFILE;
BZFILE* BZ2_bzopen(const char* path, const char* mode);
int BZ2_bzread(BZFILE* b, void* buf, int len);
void BZ2_bzclose(BZFILE* b);
}

static bool read_bz2(const std::string& path, std::string& out)
{
    BZFILE* f = BZ2_bzopen(path.c_str(), "rb");
    if (!f) return false;
    char buf[1 << 16];
    int k;
    while ((k = BZ2_bzread(f, buf, sizeof buf)) > 0) out.append(buf, k);
    BZ2_bzclose(f);
    return k == 0;
}

static bool read_gz(const std::string& path, std::string& out)
{
    gzFile f = gzopen(path.c_str(), "rb");
    if (!f) return false;
    char buf[1 << 16];
    int k;
    while ((k = gzread(f, buf, sizeof buf)) > 0) out.append(buf, k);
    gzclose(f);
    return k == 0;
}

// The reference's solve_si (gpumain.cpp:18-53), driven from the host.
static void solve_si(StrategyImprovement& si, int* br_count, int* iter_count, bool reset)
{
    si.init_strat();
    bool first = true;
    for (;;) {
        if (!first && reset) si.reset_opponent_strategy();
        for (;;) {
            if (first) si.compute_first_val();
            else si.compute_valuation();
            (*br_count)++;
            si.mark_solved(1);
            if (si.switch_strategy(1) == 0) break;
        }
        first = false;
        si.mark_solved(0);
        (*iter_count)++;
        if (si.switch_strategy(0) == 0) break;
    }
}

static void print_list(const char* name, const std::vector<int>& v, bool json)
{
    if (json) {
        printf("\"%s\":[", name);
        for (size_t i = 0; i < v.size(); i++) printf(i ? ",%d" : "%d", v[i]);
        printf("]");
    } else {
        printf("%s:", name);
        for (int x : v) printf(" %d", x);
        printf("\n");
    }
}

int main(int argc, char** argv)
{
    bool host_loop = false, reset = false, winning = false, strategy = false, json = false;
    int cpulist_cap = 0;             // > 0: CPUList's semantics with this round cap
    std::vector<std::string> pos;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        if (a == "--host-loop") host_loop = true;
        else if (a == "--reset") reset = true;
        else if (a == "--winning") winning = true;
        else if (a == "--strategy") strategy = true;
        else if (a == "--json") json = true;
        else if (a == "--cpulist") cpulist_cap = 100000;
        else if (a.rfind("--cpulist=", 0) == 0) cpulist_cap = std::max(1, atoi(a.c_str() + 10));
        else pos.push_back(a);
    }
    if (pos.size() != 2) {
        std::cout << "usage: " << argv[0] << " [--host-loop] [--reset] [--winning] [--strategy] [--json]"
                  << " [--cpulist[=N]]"
                  << " [algorithm] [game]" << std::endl
                  << std::endl
                  << "    algorithms: gpu, gpulist, gpubf, gpubv" << std::endl
                  << std::endl;
        return 1;
    }
    const std::string& algorithm = pos[0];
    const std::string& filename = pos[1];
    if (algorithm != "gpu" && algorithm != "gpulist" && algorithm != "gpubf" && algorithm != "gpubv") {
        std::cout << "unknown algorithm: " << algorithm << std::endl;
        return 1;
    }
    Game g;
    try {
        // decompression by extension, as gpumain.cpp:118-120
        if (ends_with(filename, ".gz") || ends_with(filename, ".bz2")) {
            std::string text;
            const bool ok = ends_with(filename, ".gz") ? read_gz(filename, text) : read_bz2(filename, text);
            if (!ok) throw "Unable to open file";
            std::istringstream in(text);
            g.parse_pgsolver(in);
        } else {
            g.parse_pgsolver(filename);
        }
    } catch (const char* err) {
        std::cerr << "parsing error: " << err << std::endl;
        return -1;
    }
    try {
        GPUGame gpu_g(g);
        GPUParity* si = algorithm == "gpulist" ? new GPUList(g, gpu_g) : new GPUParity(g, gpu_g);
        int iter = 0, br = 0;
        auto t0 = std::chrono::steady_clock::now();
        if (algorithm == "gpubf" || algorithm == "gpubv") {
            try {
                if (algorithm == "gpubf") {
                    si->solve_bf(&iter, &br);
                    std::cerr << "gpubf: br counts device valuation rounds, not the reference's "
                                 "relaxation sweeps" << std::endl;
                } else {
                    si->solve_bv(&iter, &br);
                }
            } catch (const std::exception& e) {
                if (std::strstr(e.what(), "odd cycle")) {
                    std::cout << "Game contains an odd cycle that uses only odd nodes" << std::endl;
                    delete si;
                    return 1;
                }
                throw;
            }
        } else if (cpulist_cap > 0) {
            if (!si->solve_list(cpulist_cap, &iter, &br))
                std::cerr << "cpulist semantics: stopped at the round cap (" << cpulist_cap
                          << " valuations)" << std::endl;
        } else if (host_loop) {
            solve_si(*si, &br, &iter, reset);
        } else {
            si->solve(reset, &iter, &br);
        }
        auto t1 = std::chrono::steady_clock::now();
        double sec = std::chrono::duration<double>(t1 - t0).count();
        std::vector<int> w0, w1;
        if (winning) {
            si->get_winning(w0, 0);
            si->get_winning(w1, 1);
        }
        std::vector<int> st;
        if (strategy) st = si->strategy();
        if (json) {
            printf("{\"seconds\":%.6f,\"iter\":%d,\"br\":%d", sec, iter, br);
            if (algorithm == "gpubf") printf(",\"br_kind\":\"device valuation rounds\"");
            if (winning) {
                printf(",");
                print_list("w0", w0, true);
                printf(",");
                print_list("w1", w1, true);
            }
            if (strategy) {
                printf(",");
                print_list("strategy", st, true);
            }
            printf("}\n");
        } else {
            printf("%f %d %d\n", sec, iter, br);
            if (winning) {
                print_list("W0", w0, false);
                print_list("W1", w1, false);
            }
            if (strategy) print_list("strategy", st, false);
        }
        delete si;
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 2;
    }
    return 0;
}
