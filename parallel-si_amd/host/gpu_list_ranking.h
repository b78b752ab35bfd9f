// The reference's gpulist solver class (gpu_list_ranking.h:11-41), the header
// gpumain.cpp:4 includes.  GPUList(Game&, GPUGame&) derives from GPUParity
// and, like it, computes CPUParity's semantics (cpu_parity.cpp) bit-exact.
// The reference's GPUList labels every vertex whose strategy does not reach
// the sink +inf (gpu_list_ranking.cu:244-277, cpu_list_ranking.cpp:208-239),
// which makes it switch forever on games where Odd closes an odd cycle (its
// own test2.gm); here such vertices take their cycle's label, so both modes
// of gpuparsi terminate with the canonical result.
#ifndef PSI_GPU_LIST_RANKING_H
#define PSI_GPU_LIST_RANKING_H

#include "game.h"
#include "gpu_game.h"
#include "gpu_parity.h"
#include "si.h"

class GPUList : public GPUParity
{
public:
    GPUList(Game& g, GPUGame& g_dev) : GPUParity(g, g_dev) {}
};

#endif
