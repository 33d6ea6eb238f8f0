// gg-convert -- a symmetric graph expanded to the reference's edge-list
// format: every vertex's neighbours, each undirected edge once as "a b" with
// a < b, sorted (cppcode/piece2graph.cpp:84-103, box2graph.cpp:85-106; -f part
// does the same for the 1-D form).  Vertex ids: part v; piece x*Ny + y; box
// (x*Ny + y)*Nz + z.
//
//   gg-convert -f part|piece|box [-i IN] [-o OUT]
#include <algorithm>
#include <fstream>
#include <iostream>
#include <string>
#include <utility>
#include <vector>
#include "gg_formats.hpp"

static int usage() {
    std::cerr << "usage: gg-convert -f part|piece|box [-i IN] [-o OUT]\n";
    return 1;
}

static long long wrap(long long x, long long n) {
    x %= n;
    return x < 0 ? x + n : x;
}

int main(int argc, char **argv) {
    std::string fmt = "piece", in_path, out_path;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        if (i + 1 >= argc) return usage();
        if (a == "-f" || a == "--format") fmt = argv[++i];
        else if (a == "-i" || a == "--input") in_path = argv[++i];
        else if (a == "-o" || a == "--output") out_path = argv[++i];
        else return usage();
    }
    std::ifstream fin;
    if (!in_path.empty()) {
        fin.open(in_path);
        if (!fin.is_open()) { std::cerr << "error: " << in_path << " is not open\n"; return 1; }
    }
    std::istream &ist = fin.is_open() ? fin : std::cin;
    std::vector<std::pair<long long, long long>> E;
    auto add = [&](long long a, long long b) {
        if (a > b) std::swap(a, b);
        E.emplace_back(a, b);
    };
    if (fmt == "part") {
        graphgolf::part p;
        p.load(ist);
        for (long long v = 0; v < p.N; v++)
            for (int off : p.edges[v % p.M]) add(v, wrap(v + off, p.N));
    } else if (fmt == "piece") {
        graphgolf::piece p;
        p.load(ist);
        for (long long x = 0; x < p.Nx; x++)
            for (long long y = 0; y < p.Ny; y++)
                for (const auto &e : p.edges[x % p.Mx][y % p.My])
                    add(x * p.Ny + y, wrap(x + e.first, p.Nx) * p.Ny + wrap(y + e.second, p.Ny));
    } else if (fmt == "box") {
        graphgolf::box p;
        p.load(ist);
        for (long long x = 0; x < p.Nx; x++)
            for (long long y = 0; y < p.Ny; y++)
                for (long long z = 0; z < p.Nz; z++)
                    for (const auto &e : p.edges[x % p.Mx][y % p.My][z % p.Mz]) {
                        const long long u = (wrap(x + std::get<0>(e), p.Nx) * p.Ny +
                                             wrap(y + std::get<1>(e), p.Ny)) * p.Nz +
                                            wrap(z + std::get<2>(e), p.Nz);
                        add((x * p.Ny + y) * p.Nz + z, u);
                    }
    } else {
        return usage();
    }
    std::sort(E.begin(), E.end());
    E.erase(std::unique(E.begin(), E.end()), E.end());
    std::ofstream fout;
    if (!out_path.empty()) {
        fout.open(out_path);
        if (!fout.is_open()) { std::cerr << "error: " << out_path << " is not open\n"; return 1; }
    }
    std::ostream &ost = fout.is_open() ? fout : std::cout;
    for (const auto &e : E) ost << e.first << ' ' << e.second << '\n';
    return 0;
}
