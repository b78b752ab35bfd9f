#include "gpu_parity.h"

#include <stdexcept>
#include <string>

static void check(int rc, const char* what)
{
    if (rc != PSI_OK) throw std::runtime_error(std::string(what) + ": " + psi_last_error());
}

GPUParity::GPUParity(Game& game, GPUGame& dev) : g(game), g_dev(dev)
{
    check(psi_solver_create(g_dev.handle, 0, 0, &s), "GPUParity");
}

GPUParity::~GPUParity() { psi_solver_destroy(s); }

void GPUParity::init_strat() { check(psi_init_strat(s), "init_strat"); }
void GPUParity::compute_valuation() { check(psi_compute_valuation(s), "compute_valuation"); }
void GPUParity::mark_solved(int player) { check(psi_mark_solved(s, player), "mark_solved"); }

int GPUParity::switch_strategy(int player)
{
    int sw = 0;
    check(psi_switch_strategy(s, player, &sw), "switch_strategy");
    return sw;
}

void GPUParity::get_winning(std::vector<int>& out, int player)
{
    std::vector<int> w(g_dev.num_vert);
    int k = 0;
    check(psi_get_winning(s, 0, player, w.data(), &k), "get_winning");
    out.insert(out.end(), w.begin(), w.begin() + k);
}

void GPUParity::reset_opponent_strategy()
{
    check(psi_reset_opponent_strategy(s), "reset_opponent_strategy");
}

void GPUParity::solve(bool reset, int* iter_count, int* br_count)
{
    check(psi_solve(s, reset ? 1 : 0, 0, 4), "solve");
    int capped = 0;
    check(psi_get_stats(s, 0, iter_count, br_count, &capped), "solve");
}

void GPUParity::solve_bf(int* iter_count, int* br_count)
{
    check(psi_solve_bf(s, 0, 4), "solve_bf");
    int capped = 0;
    check(psi_get_stats(s, 0, iter_count, br_count, &capped), "solve_bf");
}

void GPUParity::solve_bv(int* iter_count, int* br_count)
{
    check(psi_solve_bv(s, 0, 4), "solve_bv");
    int capped = 0;
    check(psi_get_stats(s, 0, iter_count, br_count, &capped), "solve_bv");
}

bool GPUParity::solve_list(int max_rounds, int* iter_count, int* br_count)
{
    int rc = psi_solve_list(s, max_rounds, 4);
    if (rc != PSI_ERR_CAP) check(rc, "solve_list");
    int capped = 0;
    check(psi_get_stats(s, 0, iter_count, br_count, &capped), "solve_list");
    return rc == PSI_OK;
}

std::vector<int> GPUParity::strategy()
{
    std::vector<int> out(g_dev.num_vert);
    check(psi_get_strategy(s, 0, out.data()), "strategy");
    return out;
}

std::vector<int> GPUParity::labels()
{
    std::vector<int> out(g_dev.num_vert);
    check(psi_get_labels(s, 0, out.data()), "labels");
    return out;
}
