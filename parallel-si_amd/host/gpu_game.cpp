#include "gpu_game.h"

#include <stdexcept>
#include <string>
#include <vector>

GPUGame::GPUGame(Game& game) : g(game)
{
    std::vector<int> owner, prio, off, tgt;
    g.to_csr(owner, prio, off, tgt);
    num_vert = (int)owner.size();
    num_edge = (int)tgt.size();
    int rc = psi_game_create(num_vert, num_edge, owner.data(), prio.data(), off.data(), tgt.data(),
                             &handle);
    if (rc == PSI_OK) rc = psi_game_device_csr(handle, &player, &priority, &vertices, &edges);
    if (rc != PSI_OK) {
        std::string msg = std::string("GPUGame: ") + psi_last_error();
        psi_game_destroy(handle);
        handle = nullptr;
        throw std::runtime_error(msg);
    }
}

GPUGame::~GPUGame() { psi_game_destroy(handle); }
