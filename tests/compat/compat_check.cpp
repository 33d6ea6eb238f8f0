// compat_check -- calls one method of the drop-in evaluator classes
// (include/graphgolf_compat.hpp) on a graph file in the reference's text
// formats and prints what it returns, so tests can compare the shims'
// values and sentinels with the oracle.  Test infrastructure.
//   compat_check conv-calc|conv-diam|conv-wvc|piece|box|graph FILE
// Output: one line of space-separated values, doubles with 17 digits.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include "graphgolf_compat.hpp"
#include "gg_formats.hpp"

int main(int argc, char **argv) {
    if (argc != 3) {
        std::cerr << "usage: compat_check conv-calc|conv-diam|conv-wvc|piece|box|graph FILE\n";
        return 1;
    }
    const std::string mode = argv[1];
    std::ifstream in(argv[2]);
    if (!in.is_open()) { std::cerr << "cannot open " << argv[2] << "\n"; return 1; }
    try {
        if (mode.rfind("conv-", 0) == 0) {
            graphgolf::part p;
            p.load(in);
            graphgolf::cudaASPLconv cu(p.N, p.M, p.degree, 0);
            if (mode == "conv-calc") {
                std::printf("%.17g\n", cu.calc(p));
            } else if (mode == "conv-diam") {
                auto r = cu.diameterASPL(p);
                std::printf("%d %.17g\n", r.first, r.second);
            } else if (mode == "conv-wvc") {
                auto r = cu.WVCdiameterASPL(p);
                std::printf("%d %d %.17g\n", std::get<0>(r), std::get<1>(r), std::get<2>(r));
            } else {
                return 1;
            }
        } else if (mode == "piece") {
            graphgolf::piece p;
            p.load(in);
            graphgolf::cudaASPLpiece cu(p.Nx, p.Ny, p.Mx, p.My, p.degree, 0);
            auto r = cu.diameterASPL(p);
            std::printf("%d %.17g\n", r.first, r.second);
        } else if (mode == "box") {
            graphgolf::box p;
            p.load(in);
            graphgolf::cudaASPLbox cu(p.Nx, p.Ny, p.Nz, p.Mx, p.My, p.Mz, p.degree, 0);
            auto r = cu.diameterASPL(p);
            std::printf("%d %.17g\n", r.first, r.second);
        } else if (mode == "graph") {
            graphgolf::graph g;
            g.load(in);
            graphgolf::cudaASPLgraph cu(g.N, g.degree_max, 0);
            auto r = cu.calc(g);
            std::printf("%d %lld\n", r.first, (long long)r.second);
        } else {
            std::cerr << "unknown mode " << mode << "\n";
            return 1;
        }
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << "\n";
        return 2;
    }
    return 0;
}
