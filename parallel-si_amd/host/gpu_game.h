// Device game (the reference's GPUGame, gpu_game.h:6-21): the int32 CSR of a
// Game -- vertices[n+1] offsets, edges[m] targets in file order, player[n],
// priority[n] -- uploaded to device memory by the constructor, as
// gpu_game.cpp:11-41 does, and exposed through the same public members.
// Here the CSR is validated on construction (successor ids in range,
// priorities in [0, 2^30), at least one successor) and held as a psi_game
// that owns the device arrays (all four are freed by the destructor; the
// reference frees only player and priority, gpu_game.cpp:43-47).  Solvers
// built on it keep their own device state.  Failures throw
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

    int* player = nullptr;    // device pointers, int32
    int* priority = nullptr;
    int* vertices = nullptr;  // offsets, vertices[num_vert] = num_edge
    int* edges = nullptr;     // the target of each edge

    psi_game* handle = nullptr;

    explicit GPUGame(Game& g);
    ~GPUGame();
    GPUGame(const GPUGame&) = delete;
    GPUGame& operator=(const GPUGame&) = delete;
};

#endif
