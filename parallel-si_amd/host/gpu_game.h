// Device game (the reference's GPUGame, gpu_game.h:6-21): the int32 CSR of a
// Game -- vertices[n+1] offsets, edges[m] targets in file order, player[n],
// priority[n].  Here the CSR is validated on construction (successor ids in
// range, priorities in [0, 2^30), at least one successor) and held as a
// psi_game; the solver that uses it owns the device copy.  Failures throw
// std::runtime_error with the library's message.
#ifndef PSI_GPU_GAME_H
#define PSI_GPU_GAME_H

#include "game.h"
#include "../../include/psi.h"

class GPUGame
{
public:
    Game& g;
    int num_vert = 0;
    int num_edge = 0;
    psi_game* handle = nullptr;

    explicit GPUGame(Game& g);
    ~GPUGame();
    GPUGame(const GPUGame&) = delete;
    GPUGame& operator=(const GPUGame&) = delete;
};

#endif
