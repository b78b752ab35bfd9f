// MI355X solver behind the reference's StrategyImprovement interface:
//   GPUParity(Game&, GPUGame&)  - the reference's gpu mode  (gpu_parity.h:8-38)
// (GPUList, the gpulist mode, is declared in gpu_list_ranking.h.)
// It computes CPUParity's semantics (cpu_parity.cpp), the reference's
// canonical CPU path, bit-exact: mark_solved freezes solved vertices (the
// reference's GPUParity::mark_solved is a no-op, gpu_parity.cu:333-336) and
// every vertex not reaching the sink is labelled by its cycle's largest
// priority.
// switch_strategy returns 1 if any vertex switched (the reference's GPU
// returns this flag too; its CPU returns a count; solve_si only tests 0).
// solve() runs the whole solve_si loop on the device instead of from the host.
// Failures throw std::runtime_error with the library's message.
#ifndef PSI_GPU_PARITY_H
#define PSI_GPU_PARITY_H

#include "gpu_game.h"
#include "si.h"

class GPUParity : public StrategyImprovement
{
protected:
    Game& g;
    GPUGame& g_dev;
    psi_solver* s = nullptr;

public:
    GPUParity(Game& g, GPUGame& g_dev);
    ~GPUParity() override;
    GPUParity(const GPUParity&) = delete;
    GPUParity& operator=(const GPUParity&) = delete;

    void init_strat() override;
    void compute_valuation() override;
    void mark_solved(int player) override;
    int switch_strategy(int player) override;
    void get_winning(std::vector<int>& out, int player) override;
    void reset_opponent_strategy() override;

    // Device-resident solve_si; returns (iter_count, br_count).
    void solve(bool reset, int* iter_count, int* br_count);
    // Device-resident solve_bf (main.cpp:57-77): the cpu_bf role, Bellman-Ford
    // best responses for Odd.  Throws on an odd strategy cycle in the first
    // best response (the reference prints the message and exits).
    void solve_bf(int* iter_count, int* br_count);
    // Device-resident solve_si with CPUBV semantics (main.cpp:163-167, the
    // cpubv mode): throws on the first odd strategy cycle, where the
    // reference prints the message and exits (cpu_bv.cpp:73-84).
    void solve_bv(int* iter_count, int* br_count);
    // Device-resident solve_si with CPUList's semantics (the cpulist mode,
    // main.cpp:178-191): false if it stopped at max_rounds valuations.
    bool solve_list(int max_rounds, int* iter_count, int* br_count);
    std::vector<int> strategy();
    std::vector<int> labels();
    psi_solver* handle() { return s; }
};

#endif
