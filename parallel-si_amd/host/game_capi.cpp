// C-ABI entries over the C++ game model: parse PGSolver text, read a game's
// CSR back.
#include "game.h"
#include "../../include/psi.h"

#include <sstream>
#include <string>
#include <vector>

extern "C" void psi_set_error_(const char* msg);

extern "C" int psi_game_from_gm(const char* text, size_t len, psi_game** out)
{
    *out = nullptr;
    Game g;
    try {
        std::istringstream in(std::string(text, len));
        g.parse_pgsolver(in);
    } catch (const char* err) {
        psi_set_error_(err);
        return PSI_ERR_INVALID;
    }
    std::vector<int> owner, prio, off, tgt;
    g.to_csr(owner, prio, off, tgt);
    return psi_game_create((int)owner.size(), (int)tgt.size(), owner.data(), prio.data(), off.data(),
                           tgt.data(), out);
}
