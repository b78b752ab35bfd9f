// Game model and PGSolver loader: the reference's Game (game.h:9-34) with the
// same member functions, so code written against it compiles unchanged.
//
// Differences, all on inputs the reference does not define:
//   * parse errors are thrown as const char* with the reference's messages
//     (game.cpp:51-107); a non-numeric vertex id throws "invalid id" (the
//     reference lets std::stoi's exception escape);
//   * parse_pgsolver(filename) throws "Unable to open file" instead of
//     printing and calling exit(1) (game.cpp:41-49);
//   * unique_priority_transform (game.cpp:110-134) is not provided: no
//     solver mode uses it.
#ifndef PSI_GAME_H
#define PSI_GAME_H

#include <istream>
#include <string>
#include <vector>

class GPUGame;

class Game
{
    friend class GPUGame;

    int n_vertices = 0;
    int max_pri = 0;
    std::vector<std::vector<int>> edges;
    std::vector<int> player;
    std::vector<int> priority;

public:
    Game() = default;

    int add_vertex(int player, int priority);
    void add_edge(int from, int to);

    int get_max_pri() { return max_pri; }
    int get_n_vertices() { return n_vertices; }
    int get_player(int v) { return player[v]; }
    int get_priority(int v) { return priority[v]; }
    std::vector<int>& get_edges(int v) { return edges[v]; }

    void parse_pgsolver(std::string filename);
    void parse_pgsolver(std::istream& in);

    // int32 CSR view (the GPUGame layout, gpu_game.cpp:11-41).
    void to_csr(std::vector<int>& owner, std::vector<int>& prio, std::vector<int>& offsets,
                std::vector<int>& targets) const;
};

#endif
