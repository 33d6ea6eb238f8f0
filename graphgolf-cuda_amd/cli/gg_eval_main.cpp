// gg-eval: one exact evaluation of a graph file, printed as the reference's
// evaluator CLIs print it (cppcode/main.cpp:99-116 for part with WVC,
// piece2aspl.cpp:96-110, box2aspl.cpp, aspl_cuda.cpp:82-105 for edge lists).
//   gg-eval [-f part|piece|box|graph] [-i FILE] [-o FILE] [-D DEVICE]
#include <chrono>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include "graphgolf_compat.hpp"
#include "gg_formats.hpp"

static int usage() {
    std::cerr << "usage: gg-eval [-f part|piece|box|graph] [-i FILE] [-o FILE] [-D DEVICE]\n";
    return 1;
}

int main(int argc, char **argv) {
    std::string fmt = "part", in_path, out_path;
    int device = 0;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto next = [&](std::string &dst) {
            if (i + 1 >= argc) return false;
            dst = argv[++i];
            return true;
        };
        std::string v;
        if ((a == "-f" || a == "--format") && next(fmt)) continue;
        if ((a == "-i" || a == "--input") && next(in_path)) continue;
        if ((a == "-o" || a == "--output") && next(out_path)) continue;
        if ((a == "-D" || a == "--device") && next(v)) { device = std::stoi(v); continue; }
        return usage();
    }
    std::ifstream fin;
    std::ofstream fout;
    if (!in_path.empty()) {
        fin.open(in_path);
        if (!fin.is_open()) { std::cerr << "error: " << in_path << " is not open\n"; return 1; }
    }
    if (!out_path.empty()) {
        fout.open(out_path);
        if (!fout.is_open()) { std::cerr << "error: " << out_path << " is not open\n"; return 1; }
    }
    std::istream &ist = fin.is_open() ? fin : std::cin;
    std::ostream &ost = fout.is_open() ? fout : std::cout;
    ost.precision(11);
    std::cout.precision(11);
    auto emit = [&](const std::string &line) {
        std::cout << line << std::endl;
        if (fout.is_open()) ost << line << std::endl;
    };
    auto num = [](double x) {
        std::ostringstream s;
        s.precision(11);
        s << x;
        return s.str();
    };
    try {
        auto t0 = std::chrono::steady_clock::now();
        if (fmt == "part") {
            graphgolf::part p;
            p.load(ist);
            graphgolf::cudaASPLconv cu(p.N, p.M, p.degree, device);
            t0 = std::chrono::steady_clock::now();
            auto r = cu.WVCdiameterASPL(p);
            auto t1 = std::chrono::steady_clock::now();
            emit("ASPL: " + num(std::get<2>(r)));
            emit("Diameter: " + std::to_string(std::get<1>(r)));
            emit("WVC: " + std::to_string(std::get<0>(r)));
            emit("TIME: " + num((double)std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count()) + " msec.");
        } else if (fmt == "piece") {
            graphgolf::piece p;
            p.load(ist);
            graphgolf::cudaASPLpiece cu(p.Nx, p.Ny, p.Mx, p.My, p.degree, device);
            t0 = std::chrono::steady_clock::now();
            auto r = cu.diameterASPL(p);
            auto t1 = std::chrono::steady_clock::now();
            emit("ASPL: " + num(r.second));
            emit("Diameter: " + std::to_string(r.first));
            emit("TIME: " + num((double)std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count()) + " msec.");
        } else if (fmt == "box") {
            graphgolf::box p;
            p.load(ist);
            graphgolf::cudaASPLbox cu(p.Nx, p.Ny, p.Nz, p.Mx, p.My, p.Mz, p.degree, device);
            t0 = std::chrono::steady_clock::now();
            auto r = cu.diameterASPL(p);
            auto t1 = std::chrono::steady_clock::now();
            emit("ASPL: " + num(r.second));
            emit("Diameter: " + std::to_string(r.first));
            emit("TIME: " + num((double)std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count()) + " msec.");
        } else if (fmt == "graph") {
            // aspl_cuda.cpp:82-105
            graphgolf::graph g;
            g.load(ist);
            std::cout << "Nodes: " << g.N << " Degrees_max: " << g.degree_max << std::endl;
            graphgolf::cudaASPLgraph cu(g.N, g.degree_max, device);
            t0 = std::chrono::steady_clock::now();
            auto r = cu.calc(g);
            auto t1 = std::chrono::steady_clock::now();
            double aspl = double(r.second) / ((int64_t)(g.N - 1) * g.N);
            emit("Diameter: " + std::to_string(r.first));
            std::cout << "ASPL: " << num(aspl) << " (" << r.second / 2 << "/"
                      << (int64_t)g.N * (g.N - 1) / 2 << ")" << std::endl;
            if (fout.is_open()) ost << "ASPL: " << num(aspl) << std::endl;
            emit("TIME: " + num((double)std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count()) + " msec.");
        } else {
            return usage();
        }
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 2;
    }
    return 0;
}
