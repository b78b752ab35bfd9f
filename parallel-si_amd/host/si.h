// The strategy-improvement interface of the reference (si.h:6-16), unchanged,
// so drivers written against it (solve_si, gpumain.cpp:18-53) run on the
// MI355X solvers below.
#ifndef PSI_SI_H
#define PSI_SI_H

#include <vector>

class StrategyImprovement
{
public:
    virtual ~StrategyImprovement() = default;
    virtual void init_strat() = 0;
    virtual void compute_valuation() = 0;
    virtual void compute_first_val() { compute_valuation(); }
    virtual void mark_solved(int player) = 0;
    virtual int switch_strategy(int player) = 0;
    virtual void get_winning(std::vector<int>& out, int player) = 0;
    virtual void reset_opponent_strategy() {}
};

#endif
