// si_driver: a program written only against the reference's C++ surface
// (game.h, gpu_game.h, gpu_parity.h, gpu_list_ranking.h, si.h), the way
// gpumain.cpp uses it: the solve_si loop over StrategyImprovement&
// (gpumain.cpp:18-53), the timed test() (gpumain.cpp:56-71) and the solver
// choice of main (gpumain.cpp:102-146).  Built with the library to show that
// code of that shape compiles and runs on these headers unchanged.
//
//   si_driver <gpu|gpulist> <game.gm>
//
// Prints "<seconds> <iter_count> <br_count>" (as test() does), then the
// winning sets from get_winning(out, 0) and get_winning(out, 1), then
// "csr <0|1>": 1 if GPUGame's public device arrays (player, priority,
// vertices, edges; gpu_game.h:11-16) hold the game's CSR.
#include "game.h"
#include "gpu_game.h"
#include "gpu_parity.h"
#include "gpu_list_ranking.h"
#include "si.h"

#include <sys/time.h>

#include <iostream>
#include <string>
#include <vector>

using namespace std;

void solve_si(StrategyImprovement& si, int* br_count, int* iter_count, bool reset)
{
    si.init_strat();
    bool first = true;
    while (true) {
        if (!first && reset) si.reset_opponent_strategy();
        while (true) {
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

template <typename T> double test(T& si, void (*f)(T&, int*, int*, bool), bool reset)
{
    timeval start;
    gettimeofday(&start, nullptr);
    int iter_count = 0;
    int br_count = 0;
    (*f)(si, &br_count, &iter_count, reset);
    timeval end;
    gettimeofday(&end, nullptr);
    double solve_time = (end.tv_sec - start.tv_sec) + (end.tv_usec - start.tv_usec) / 1000000.0;
    cout << solve_time << " " << iter_count << " " << br_count << endl;
    return solve_time;
}

static bool device_csr_matches(Game& g, GPUGame& d)
{
    int n = d.num_vert, m = d.num_edge;
    if (n != g.get_n_vertices()) return false;
    vector<int> player(n), priority(n), vertices(n + 1), edges(m);
    if (psi_memcpy_d2h(player.data(), d.player, 4 * (size_t)n) != PSI_OK ||
        psi_memcpy_d2h(priority.data(), d.priority, 4 * (size_t)n) != PSI_OK ||
        psi_memcpy_d2h(vertices.data(), d.vertices, 4 * (size_t)(n + 1)) != PSI_OK ||
        psi_memcpy_d2h(edges.data(), d.edges, 4 * (size_t)m) != PSI_OK)
        return false;
    int e = 0;
    for (int v = 0; v < n; v++) {
        if (player[v] != g.get_player(v) || priority[v] != g.get_priority(v) || vertices[v] != e)
            return false;
        for (int u : g.get_edges(v))
            if (edges[e++] != u) return false;
    }
    return vertices[n] == m && e == m;
}

int main(int argc, char** argv)
{
    if (argc != 3) {
        cout << "usage: " << argv[0] << " [algorithm] [game]" << endl;
        return 1;
    }
    string algorithm = argv[1];
    Game g;
    try {
        g.parse_pgsolver(string(argv[2]));
    } catch (const char* err) {
        cout << "parsing error: " << err << endl;
        return -1;
    }
    GPUGame gpu_g(g);
    StrategyImprovement* si = nullptr;
    if (algorithm == "gpu")
        si = new GPUParity(g, gpu_g);
    else if (algorithm == "gpulist")
        si = new GPUList(g, gpu_g);
    else {
        cout << "unknown algorithm: " << algorithm << endl;
        return 1;
    }
    test(*si, solve_si, false);
    for (int p = 0; p < 2; p++) {
        vector<int> out;
        si->get_winning(out, p);
        cout << "W" << p << ":";
        for (int v : out) cout << " " << v;
        cout << endl;
    }
    cout << "csr " << (device_csr_matches(g, gpu_g) ? 1 : 0) << endl;
    delete si;
    return 0;
}
