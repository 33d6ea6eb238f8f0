// gg-inittemp -- initial annealing temperature as cppcode/initTemp.cpp and
// pieceinitTemp.cpp estimate it: over random graphs, the mean increase of the
// distance total over the neighbour moves that make it worse, and
// inittemp = -avg / ln 0.4 (initTemp.cpp:399-430), printed as its last lines
// ("cnt:", "avg:", "inittemp:").  -m M < N: orbit graphs with N/M copies and
// orbit moves (gg_init_temp_part); -m N: random regular graphs and 2-opt swaps
// (gg_init_temp).  The graphs and moves are this library's seeded ones.
//
//   gg-inittemp [-n N] [-d DEG] [-m M] [-s SEED] [-D DEVICE] [--graphs G]
//               [--neighbours K] [-o OUT]
#include <fstream>
#include <iostream>
#include <random>
#include <string>
#include "graphgolf_mi355x.h"
#include "graphgolf_search.h"

static int usage(int code) {
    std::cout << "gg-inittemp [-n N] [-d DEG] [-m M] [-s SEED] [-D DEVICE] [--graphs G] "
                 "[--neighbours K] [-o OUT]\n";
    return code;
}

int main(int argc, char **argv) {
    // defaults of initTemp.cpp:17-20 and its 100 x 100 samples (:399, :417)
    int N = 1000, d = 16, M = 20, device = 0, graphs = 100, neighbours = 100;
    bool have_seed = false;
    unsigned long long seed = 0;
    std::string out_path;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) { std::cerr << "missing value for " << a << "\n"; std::exit(1); }
            return argv[++i];
        };
        if (a == "-h" || a == "--help") return usage(0);
        else if (a == "-n" || a == "--number") N = std::stoi(next());
        else if (a == "-d" || a == "--degree") d = std::stoi(next());
        else if (a == "-m" || a == "--module") M = std::stoi(next());
        else if (a == "-s" || a == "--seed") { seed = std::stoull(next()); have_seed = true; }
        else if (a == "-D" || a == "--device") device = std::stoi(next());
        else if (a == "--graphs") graphs = std::stoi(next());
        else if (a == "--neighbours") neighbours = std::stoi(next());
        else if (a == "-o" || a == "--output") out_path = next();
        else { std::cout << "unknown option " << a << "\n"; return usage(1); }
    }
    if (!have_seed) seed = std::random_device{}();
    gg_init_temp_result r{};
    const int st = M >= N ? gg_init_temp(N, d, graphs, neighbours, seed, device, &r)
                          : gg_init_temp_part(N, M, d, graphs, neighbours, seed, device, &r);
    if (st != GG_OK) { std::cerr << "error: " << gg_strerror(st) << "\n"; return 1; }
    std::ofstream fout;
    if (!out_path.empty()) {
        fout.open(out_path);
        if (!fout.is_open()) { std::cerr << "error: " << out_path << " is not open\n"; return 1; }
    }
    std::ostream &ost = fout.is_open() ? fout : std::cout;
    ost.precision(10);
    ost << "cnt: " << r.count << "\n";
    ost << "avg: " << r.mean_increase_total << "\n";
    ost << "inittemp: " << r.temp_total << "\n";
    return 0;
}
