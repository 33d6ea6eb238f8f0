// gg-gen -- seeded cyclically symmetric regular graphs in the reference's text
// formats: the options of cppcode/RNDpiecegen.cpp:12-22 and RNDboxgen.cpp
// (--nx --ny --nz --mx --my --mz -d -s -o), plus -f part (-n N -m M) for the
// 1-D form.  The pairing with a random lift per axis is the reference's
// (RNDpiecegen.cpp:89-127); pairs that would give a self-loop or a repeated
// offset are repaired (gg_gen_lattice), where the reference keeps them.
//
//   gg-gen -f part  -n N -m M -d DEG [-s SEED] [-o OUT]
//   gg-gen -f piece --nx NX --ny NY --mx MX --my MY -d DEG [-s SEED] [-o OUT]
//   gg-gen -f box   --nx NX --ny NY --nz NZ --mx MX --my MY --mz MZ -d DEG [-s SEED] [-o OUT]
#include <fstream>
#include <iostream>
#include <random>
#include <string>
#include <vector>
#include "gg_formats.hpp"
#include "graphgolf_mi355x.h"
#include "graphgolf_search.h"

static int usage(int code) {
    std::cout << "gg-gen -f part  -n N -m M -d DEG [-s SEED] [-o OUT]\n"
                 "gg-gen -f piece --nx NX --ny NY --mx MX --my MY -d DEG [-s SEED] [-o OUT]\n"
                 "gg-gen -f box   --nx NX --ny NY --nz NZ --mx MX --my MY --mz MZ -d DEG "
                 "[-s SEED] [-o OUT]\n";
    return code;
}

int main(int argc, char **argv) {
    std::string fmt = "piece", out_path;
    int n[3] = {0, 0, 0}, m[3] = {0, 0, 0}, d = 0;
    bool have_seed = false;
    unsigned long long seed = 0;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) { std::cerr << "missing value for " << a << "\n"; std::exit(1); }
            return argv[++i];
        };
        if (a == "-h" || a == "--help") return usage(0);
        else if (a == "-f" || a == "--format") fmt = next();
        else if (a == "-n" || a == "--number" || a == "--nx") n[0] = std::stoi(next());
        else if (a == "--ny") n[1] = std::stoi(next());
        else if (a == "--nz") n[2] = std::stoi(next());
        else if (a == "-m" || a == "--module" || a == "--mx") m[0] = std::stoi(next());
        else if (a == "--my") m[1] = std::stoi(next());
        else if (a == "--mz") m[2] = std::stoi(next());
        else if (a == "-d" || a == "--degree") d = std::stoi(next());
        else if (a == "-s" || a == "--seed") { seed = std::stoull(next()); have_seed = true; }
        else if (a == "-o" || a == "--output") out_path = next();
        else { std::cout << "unknown option " << a << "\n"; return usage(1); }
    }
    const int dims = fmt == "part" ? 1 : fmt == "piece" ? 2 : fmt == "box" ? 3 : 0;
    if (!dims) return usage(1);
    if (!have_seed) seed = std::random_device{}();
    int M = 1;
    for (int k = 0; k < dims; k++) M *= m[k] > 0 ? m[k] : 0;
    std::vector<int32_t> t((size_t)(M > 0 ? M : 1) * (d > 0 ? d : 1) * dims);
    const int st = gg_gen_lattice(dims, n, m, d, seed, t.data());
    if (st != GG_OK) {
        std::cerr << "error: " << gg_strerror(st) << " (sizes must divide, M*d even)\n";
        return 1;
    }
    std::ofstream fout;
    if (!out_path.empty()) {
        fout.open(out_path);
        if (!fout.is_open()) { std::cerr << "error: " << out_path << " is not open\n"; return 1; }
    }
    std::ostream &ost = fout.is_open() ? fout : std::cout;
    if (dims == 1) {
        graphgolf::part p;
        p.N = n[0];
        p.M = m[0];
        p.degree = d;
        p.edges.assign(M, std::vector<int>(d));
        for (int r = 0; r < M; r++)
            for (int j = 0; j < d; j++) p.edges[r][j] = t[(size_t)r * d + j];
        p.print(ost);
    } else if (dims == 2) {
        graphgolf::piece p;
        p.Nx = n[0]; p.Ny = n[1]; p.Mx = m[0]; p.My = m[1]; p.degree = d;
        p.edges.assign(p.Mx, std::vector<std::vector<std::pair<int, int>>>(p.My));
        for (int x = 0; x < p.Mx; x++)
            for (int y = 0; y < p.My; y++)
                for (int j = 0; j < d; j++) {
                    const size_t q = (((size_t)x * p.My + y) * d + j) * 2;
                    p.edges[x][y].emplace_back(t[q], t[q + 1]);
                }
        p.print(ost);
    } else {
        graphgolf::box p;
        p.Nx = n[0]; p.Ny = n[1]; p.Nz = n[2]; p.Mx = m[0]; p.My = m[1]; p.Mz = m[2];
        p.degree = d;
        p.edges.assign(p.Mx, std::vector<std::vector<std::vector<std::tuple<int, int, int>>>>(
                                 p.My, std::vector<std::vector<std::tuple<int, int, int>>>(p.Mz)));
        for (int x = 0; x < p.Mx; x++)
            for (int y = 0; y < p.My; y++)
                for (int z = 0; z < p.Mz; z++)
                    for (int j = 0; j < d; j++) {
                        const size_t q = ((((size_t)x * p.My + y) * p.Mz + z) * d + j) * 3;
                        p.edges[x][y][z].emplace_back(t[q], t[q + 1], t[q + 2]);
                    }
        p.print(ost);
    }
    return 0;
}
