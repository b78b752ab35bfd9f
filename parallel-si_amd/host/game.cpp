#include "game.h"

#include <fstream>
#include <sstream>
#include <stdexcept>

int Game::add_vertex(int pl, int pri)
{
    int id = n_vertices++;
    edges.emplace_back();
    player.push_back(pl);
    priority.push_back(pri);
    if (pri > max_pri) max_pri = pri;
    return id;
}

void Game::add_edge(int from, int to) { edges[from].push_back(to); }

void Game::parse_pgsolver(std::string filename)
{
    std::ifstream in(filename.c_str());
    if (!in) throw "Unable to open file";
    parse_pgsolver(in);
}

// PGSolver text: the first line is a header and is ignored; every other
// non-empty line is "id priority owner s1,s2,...[ "label"];" with ids in order
// from 0.  Successors are separated by commas; the first character after a
// successor that is not a comma ends the list (game.cpp:51-107).
void Game::parse_pgsolver(std::istream& in)
{
    std::string line;
    if (!std::getline(in, line)) return;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::string first;
        if (!(ss >> first)) continue;
        int id;
        try {
            id = std::stoi(first);
        } catch (const std::exception&) {
            throw "invalid id";
        }
        if (id < 0) throw "invalid id";
        int pri, pl;
        if (!(ss >> pri)) throw "missing priority";
        if (!(ss >> pl)) throw "missing player";
        if (pl != 0 && pl != 1) throw "invalid player";
        if (add_vertex(pl, pri) != id) throw "vertices must be listed in order starting at 0!";
        for (;;) {
            int to;
            if (!(ss >> to)) throw "missing successor";
            add_edge(id, to);
            char sep;
            if (!(ss >> sep)) throw "missing ; to end line";
            if (sep != ',') break;
        }
    }
}

void Game::to_csr(std::vector<int>& owner, std::vector<int>& prio, std::vector<int>& offsets,
                  std::vector<int>& targets) const
{
    owner = player;
    prio = priority;
    offsets.assign(1, 0);
    targets.clear();
    for (int v = 0; v < n_vertices; v++) {
        targets.insert(targets.end(), edges[v].begin(), edges[v].end());
        offsets.push_back((int)targets.size());
    }
}
