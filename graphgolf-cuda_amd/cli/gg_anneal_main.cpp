// gg-anneal -- simulated annealing / hill climbing driver on the MI355X
// evaluator with the command line, input/output text format and log lines
// of the reference's drivers (cppcode/annealing.cpp:14-105 options and
// set-up, :293-330 initial evaluation and log header, :396-452 loop, log and
// result; cppcode/HillClimb.cpp for --hill).
//
//   gg-anneal [-c COUNT] [-s SEED] [-i IN] [-o OUT] [-l LOG] [-v VERBOSE] [-D DEVICE]
//             [-t TEMP] [--hill] [--mode plain|batched|learned] [-K K] [-P P]
//             [--schedule step|geometric] [--temp-end T] [-f part|piece|box] [--wvc] [--nr]
//
// Input and output are `part` tables (part.hpp:13-43) by default.  M = N: an
// explicit graph, 2-opt swaps (gg_search); M < N: a cyclically symmetric
// graph, orbit moves (gg_orbit_search).  -f piece / -f box read and write the
// 2-D / 3-D forms (piece.hpp:18-52, box.hpp:37-80) and search them with the
// moves of pieceannealing.cpp / boxHillClimb.cpp (gg_piece_search_create,
// gg_box_search_create; plain mode).  TEMP is the reference's initial temperature in
// total-distance units (default 46.1793, annealing.cpp:314); the default
// schedule is the reference's step schedule (annealing.cpp:442).  The moves
// are this library's (gg_search.cpp / gg_orbit.cpp), so a run does not
// reproduce the reference's random trajectory; the log and console lines do
// have the reference's layout.  The printed solution is the library's best by
// (diameter, ASPL); the reference keeps the accepted solution of least ASPL
// (annealing.cpp:429-436), which differs only when a later solution of
// smaller diameter has a larger ASPL.  A best-solution dump in the verbose
// file is written once per 1000 iterations (for the last improvement among
// them) rather than at every improvement.  --wvc compares candidates on
// (diameter, WVC, ASPL) as annealingWVC.cpp:412-417 / HillClimbWVC.cpp:240
// do (part input, plain mode), with annealingWVC's schedule by default:
// temp = TEMP^(1 - i/COUNT), geometric from TEMP to 1 (annealingWVC.cpp:411).
// --nr accepts a symmetric part table whose rows have different numbers of
// entries, as annealingNR does: slots are drawn per row among its entries
// (annealingNR.cpp:117,146-147), rows keep their lengths, short rows are
// evaluated padded with offset 0 (cudaASPLconv.cu:199-201).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <random>
#include <string>
#include <vector>
#include "gg_formats.hpp"
#include "graphgolf_mi355x.h"
#include "graphgolf_search.h"

namespace {

int usage(int code) {
    std::cout << "gg-anneal [-c COUNT] [-s SEED] [-i IN] [-o OUT] [-l LOG] [-v VERBOSE] [-D DEVICE]\n"
                 "          [-t TEMP] [--hill] [--mode plain|batched|learned] [-K K] [-P P]\n"
                 "          [--schedule step|geometric] [--temp-end T] [-f part|piece|box]\n"
                 "          [--wvc] [--nr]\n";
    return code;
}

const char *kClear = "\x1b[F\x1b[E\x1b[K";  // the reference's line rewrite (annealing.cpp:418)

}  // namespace

int main(int argc, char **argv) {
    auto t_init = std::chrono::steady_clock::now();
    int count = 1000, device = 0, K = 32, P = 1024;
    bool have_seed = false, hill = false, wvc = false, have_schedule = false, nr = false;
    unsigned int seed = 0;
    double inittemp = 46.1793, temp_end = -1;
    std::string in_path, out_path, log_path, verbose_path, mode = "plain", schedule = "step";
    std::string fmt = "part";
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) { std::cerr << "missing value for " << a << "\n"; std::exit(1); }
            return argv[++i];
        };
        if (a == "-h" || a == "--help") return usage(0);
        else if (a == "-c" || a == "--count") count = std::stoi(next());
        else if (a == "-s" || a == "--seed") { seed = (unsigned)std::stoul(next()); have_seed = true; }
        else if (a == "-i" || a == "--input") in_path = next();
        else if (a == "-o" || a == "--output") out_path = next();
        else if (a == "-l" || a == "--log") log_path = next();
        else if (a == "-v" || a == "--verbose") verbose_path = next();
        else if (a == "-D" || a == "--device") device = std::stoi(next());
        else if (a == "-t" || a == "--temp") inittemp = std::stod(next());
        else if (a == "--hill") hill = true;
        else if (a == "--mode") mode = next();
        else if (a == "-K") K = std::stoi(next());
        else if (a == "-P") P = std::stoi(next());
        else if (a == "--schedule") { schedule = next(); have_schedule = true; }
        else if (a == "--wvc") wvc = true;
        else if (a == "--nr") nr = true;
        else if (a == "--temp-end") temp_end = std::stod(next());
        else if (a == "-f" || a == "--format") fmt = next();
        else { std::cout << "unknown option " << a << "\n"; return usage(1); }
    }
    std::ifstream inputfs;
    std::ofstream outputfs, logfs, verbosefs;
    if (!in_path.empty()) {
        inputfs.open(in_path, std::ios::binary | std::ios::in);
        if (!inputfs.is_open()) { std::cerr << "error: " << in_path << " is not open\n"; return 1; }
    }
    if (!out_path.empty()) {
        outputfs.open(out_path, std::ios::binary | std::ios::out);
        if (!outputfs.is_open()) { std::cerr << "error: " << out_path << " is not open\n"; return 1; }
    }
    const bool logging = !log_path.empty(), verbose = !verbose_path.empty();
    if (logging) {
        logfs.open(log_path, std::ios::binary | std::ios::out);
        if (!logfs.is_open()) { std::cerr << "error: " << log_path << " is not open\n"; return 1; }
        logfs.precision(10);
    }
    if (verbose) {
        verbosefs.open(verbose_path, std::ios::binary | std::ios::out);
        if (!verbosefs.is_open()) {
            std::cerr << "error: " << verbose_path << " is not open\n";
            return 1;
        }
        verbosefs.precision(10);
    }
    std::istream &ist = inputfs.is_open() ? inputfs : std::cin;
    std::ostream &ost = outputfs.is_open() ? outputfs : std::cout;
    for (std::ostream *f : {static_cast<std::ostream *>(&verbosefs), static_cast<std::ostream *>(&logfs)}) {
        if (f == &verbosefs && !verbose) continue;
        if (f == &logfs && !logging) continue;
        if (!in_path.empty()) *f << "#input file: " << in_path << std::endl;
        if (!out_path.empty()) *f << "#output file: " << out_path << std::endl;
    }
    if (fmt != "part" && fmt != "piece" && fmt != "box") return usage(1);
    // the graph as a flat table: part [M][d]; piece / box [M][d][dims] in the
    // evaluators' layout (cudaASPLpiece.cu:163-174)
    graphgolf::part init;
    graphgolf::piece init_piece;
    graphgolf::box init_box;
    int dims = 1, N = 0, M = 0, d = 0;
    bool regular = true;
    std::vector<int32_t> tab, lens;
    if (fmt == "part") {
        init.load(ist);
        regular = init.regular;
        N = init.N; M = init.M; d = init.degree;
        tab.assign((size_t)M * d, 0);
        if (regular || nr) {
            // --nr: each row's entries first, the rest offset 0
            for (int r = 0; r < M; r++) {
                lens.push_back((int32_t)init.edges[r].size());
                for (size_t j = 0; j < init.edges[r].size(); j++)
                    tab[(size_t)r * d + j] = init.edges[r][j];
            }
            if (!regular && M == N) {
                std::cout << "--nr needs a symmetric table (M < N)" << std::endl;
                return 1;
            }
            regular = true;
        }
    } else if (fmt == "piece") {
        init_piece.load(ist);
        regular = init_piece.regular;
        dims = 2;
        N = init_piece.Nx * init_piece.Ny; M = init_piece.Mx * init_piece.My;
        d = init_piece.degree;
        tab.resize((size_t)M * d * 2);
        if (regular)
            for (int x = 0; x < init_piece.Mx; x++)
                for (int y = 0; y < init_piece.My; y++)
                    for (int j = 0; j < d; j++) {
                        const size_t o = (((size_t)x * init_piece.My + y) * d + j) * 2;
                        tab[o] = init_piece.edges[x][y][j].first;
                        tab[o + 1] = init_piece.edges[x][y][j].second;
                    }
    } else {
        init_box.load(ist);
        regular = init_box.regular;
        dims = 3;
        N = init_box.N; M = init_box.M; d = init_box.degree;
        tab.resize((size_t)M * d * 3);
        if (regular)
            for (int x = 0; x < init_box.Mx; x++)
                for (int y = 0; y < init_box.My; y++)
                    for (int z = 0; z < init_box.Mz; z++)
                        for (int j = 0; j < d; j++) {
                            const size_t o =
                                ((((size_t)x * init_box.My + y) * init_box.Mz + z) * d + j) * 3;
                            const auto &e = init_box.edges[x][y][z][j];
                            tab[o] = std::get<0>(e);
                            tab[o + 1] = std::get<1>(e);
                            tab[o + 2] = std::get<2>(e);
                        }
    }
    if (!regular) {
        std::cout << "non-regular graph" << std::endl;
        return 1;
    }
    if (verbose) verbosefs << "#N: " << N << " M: " << M << std::endl;
    if (!have_seed) seed = std::random_device{}();
    if (verbose) verbosefs << "#seed: " << seed << std::endl;
    const double NN1 = (double)N * (N - 1);
    // prints a table the search returns in the input's format (explicit
    // graphs come back as neighbour ids: offsets v -> u as u - v)
    auto print_table = [&](std::ostream &o, const std::vector<int32_t> &t) {
        if (dims == 1) {
            graphgolf::part p = init;
            for (int r = 0; r < M; r++)
                for (size_t j = 0; j < p.edges[r].size(); j++)
                    p.edges[r][j] = M == N ? t[(size_t)r * d + j] - r : t[(size_t)r * d + j];
            p.print(o);
        } else if (dims == 2) {
            graphgolf::piece p = init_piece;
            for (int x = 0; x < p.Mx; x++)
                for (int y = 0; y < p.My; y++)
                    for (int j = 0; j < d; j++) {
                        const size_t q = (((size_t)x * p.My + y) * d + j) * 2;
                        p.edges[x][y][j] = std::make_pair(t[q], t[q + 1]);
                    }
            p.print(o);
        } else {
            graphgolf::box p = init_box;
            for (int x = 0; x < p.Mx; x++)
                for (int y = 0; y < p.My; y++)
                    for (int z = 0; z < p.Mz; z++)
                        for (int j = 0; j < d; j++) {
                            const size_t q = ((((size_t)x * p.My + y) * p.Mz + z) * d + j) * 3;
                            p.edges[x][y][z][j] = std::make_tuple(t[q], t[q + 1], t[q + 2]);
                        }
            p.print(o);
        }
    };

    gg_search_cfg cfg{};
    cfg.mode = mode == "learned" ? GG_MODE_LEARNED : mode == "batched" ? GG_MODE_BATCHED : GG_MODE_PLAIN;
    cfg.rule = wvc ? (hill ? GG_RULE_HILL_WVC : GG_RULE_ANNEAL_WVC)
                   : (hill ? GG_RULE_HILL : GG_RULE_ANNEAL);
    if (wvc && !have_schedule) schedule = "geometric";
    if (wvc && temp_end <= 0) temp_end = 1.0;
    cfg.K = K;
    cfg.P = P;
    cfg.total_steps = count > 0 ? count : 1;
    cfg.warmup_pairs = 2048;
    cfg.train_every = 1;
    cfg.train_batch = 256;
    cfg.T0 = inittemp / NN1;
    cfg.T1 = temp_end > 0 ? temp_end / NN1 : cfg.T0 * 1e-3;
    cfg.lr = 1e-3;
    cfg.seed = seed;
    cfg.schedule = schedule == "geometric" ? GG_SCHED_GEOMETRIC : GG_SCHED_STEP;

    // evaluator of the initial solution (annealing.cpp:296-312)
    gg_eval *ev = nullptr;
    int st = dims == 1 ? gg_create_part(N, M, d, device, &ev)
             : dims == 2 ? gg_create_piece(init_piece.Nx, init_piece.Ny, init_piece.Mx,
                                           init_piece.My, d, device, &ev)
                         : gg_create_box(init_box.Nx, init_box.Ny, init_box.Nz, init_box.Mx,
                                         init_box.My, init_box.Mz, d, device, &ev);
    if (st != GG_OK) { std::cerr << "error: " << gg_strerror(st) << "\n"; return 1; }
    gg_result r0;
    st = dims == 1 ? gg_eval_part(ev, tab.data(), &r0)
         : dims == 2 ? gg_eval_piece(ev, tab.data(), &r0)
                     : gg_eval_box(ev, tab.data(), &r0);
    gg_destroy(ev);
    if (st != GG_OK) { std::cerr << "error: " << gg_strerror(st) << "\n"; return 1; }
    if (r0.status != GG_OK) {
        if (logging) logfs << "#initial solution is unconnected" << std::endl;
        std::cout << "initial solution is unconnected" << std::endl;
        return 0;
    }
    std::cout.precision(10);
    const double init_aspl = gg_aspl(&r0, N);
    std::cout << "ASPL(init_x): " << init_aspl << std::endl;
    if (verbose) verbosefs << "#ASPL(init_x): " << init_aspl << std::endl;
    if (logging) {
        logfs << "#iteration ASPL(x_best) ASPL(x) ASPL(y) temp" << std::endl;
        logfs << 0 << ' ' << init_aspl << ' ' << init_aspl << ' ' << init_aspl << ' ' << inittemp
              << std::endl;
    }

    // the search: explicit graph (2-opt) or symmetric table (orbit moves)
    gg_search *s = nullptr;
    gg_orbit_search *os = nullptr;
    std::vector<int32_t> adj;
    if (dims == 2) {
        cfg.mode = GG_MODE_PLAIN;
        st = gg_piece_search_create(init_piece.Nx, init_piece.Ny, init_piece.Mx, init_piece.My, d,
                                    tab.data(), &cfg, device, &os);
    } else if (dims == 3) {
        cfg.mode = GG_MODE_PLAIN;
        st = gg_box_search_create(init_box.Nx, init_box.Ny, init_box.Nz, init_box.Mx, init_box.My,
                                  init_box.Mz, d, tab.data(), &cfg, device, &os);
    } else if (M == N) {
        adj.resize((size_t)N * d);
        for (int v = 0; v < N; v++)
            for (int j = 0; j < d; j++) {
                long long u = ((long long)v + tab[(size_t)v * d + j]) % N;
                adj[(size_t)v * d + j] = (int32_t)(u < 0 ? u + N : u);
            }
        st = gg_search_create(N, d, adj.data(), &cfg, device, &s);
    } else {
        cfg.mode = GG_MODE_PLAIN;
        st = gg_orbit_search_create(N, M, d, tab.data(), &cfg, device, &os);
        if (st == GG_OK && nr) st = gg_orbit_search_set_row_lengths(os, lens.data());
    }
    if (st != GG_OK) { std::cerr << "error: " << gg_strerror(st) << "\n"; return 1; }

    double fx = init_aspl, fx_best = init_aspl;
    int dx = r0.diameter, dx_best = r0.diameter;
    const int chunk = 1000;
    std::vector<int64_t> log((size_t)chunk * 10);
    for (int done = 0; done < count;) {
        const int n = count - done < chunk ? count - done : chunk;
        auto t0 = std::chrono::steady_clock::now();
        gg_search_stats stats;
        int got = 0;
        if (s) {
            st = gg_search_set_log(s, n);
            if (st == GG_OK) st = gg_search_run(s, n, &stats);
            if (st == GG_OK) st = gg_search_get_log(s, log.data(), n, &got);
        } else {
            st = gg_orbit_search_set_log(os, n);
            if (st == GG_OK) st = gg_orbit_search_run(os, n, &stats);
            if (st == GG_OK) st = gg_orbit_search_get_log(os, log.data(), n, &got);
        }
        if (st != GG_OK) { std::cerr << "error: " << gg_strerror(st) << "\n"; return 1; }
        auto t1 = std::chrono::steady_clock::now();
        const double ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / (n ? n : 1);
        bool improved = false;
        for (int e = 0; e < got; e++) {
            const int64_t *rec = &log[(size_t)e * 10];
            const int64_t i = rec[0];
            const bool ok = rec[8] == GG_OK, accept = rec[9] != 0;
            const double fy = ok ? (double)rec[6] / NN1 : 1e8;
            const int dy = ok ? (int)rec[7] : 100000000;
            const double temp = (cfg.schedule == GG_SCHED_STEP
                                     ? cfg.T0 * std::pow(0.995, (double)((i - 1) / 10000))
                                     : cfg.T0 * std::pow(cfg.T1 / cfg.T0, (double)i / cfg.total_steps)) *
                                NN1;
            if (accept) {
                if (fx != fy) {
                    std::cout << kClear << "iteration: " << i << " fx_best: " << fx_best
                              << " fx: " << fx << " time: " << ms << "ms" << std::flush;
                    if (verbose)
                        verbosefs << "iteration: " << i << " fx_best: " << fx_best << " fx: " << fx
                                  << " temp: " << temp << std::endl;
                    if (logging) logfs << i << ' ' << fx_best << ' ' << fx << ' ' << temp << std::endl;
                }
                const bool best = fy < fx_best && dy <= dx_best;
                fx = fy;
                dx = dy;
                if (best) {
                    std::cout << std::endl;
                    fx_best = fy;
                    dx_best = dy;
                    improved = true;
                }
            }
            if (i % 1000 == 0)
                std::cout << kClear << "iteration: " << i << " fx_best: " << fx_best << " fx: " << fx
                          << " time: " << ms << "ms" << std::flush;
        }
        done += n;
        if (improved && verbose) {
            std::vector<int32_t> out((size_t)M * d * dims);
            if (s) gg_search_graph(s, out.data(), 1);
            else gg_orbit_search_table(os, out.data(), 1);
            verbosefs << "#update. new solution:" << std::endl;
            print_table(verbosefs, out);
        }
    }
    const double elapsed =
        std::chrono::duration_cast<std::chrono::seconds>(std::chrono::steady_clock::now() - t_init)
            .count();
    std::cout << std::endl << "total elapsed time: " << elapsed << "s" << std::endl;
    if (verbose) verbosefs << "#total elapsed time: " << elapsed << "s" << std::endl;
    // best solution in the input's format (offsets v -> u as u - v for M = N)
    std::vector<int32_t> out((size_t)M * d * dims);
    if (s) gg_search_graph(s, out.data(), 1);
    else gg_orbit_search_table(os, out.data(), 1);
    print_table(ost, out);
    if (s) gg_search_destroy(s);
    if (os) gg_orbit_search_destroy(os);
    (void)dx;
    return 0;
}
