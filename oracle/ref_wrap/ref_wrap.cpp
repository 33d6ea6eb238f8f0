// ref_wrap.cpp -- TEST INFRASTRUCTURE ONLY (checker side).
//
// Builds the reference's own CPU oracles (cppcode/cpuASPLqueue.cpp,
// cppcode/cpuASPLpiece.cpp, cppcode/cpuASPLbox.cpp with cppcode/part.hpp,
// piece.hpp, box.hpp, arrayQueue.hpp) straight from the reference tree, which
// the Makefile next to this file puts on the include path.  Nothing of the
// reference is copied here: this unit only fills the reference's own structs
// from flat arrays and calls its diameterASPL, so the tests can use it to pin
// the plain-C restatement in gg_oracle.c.  Output goes to oracle/_ref/.
#include <cstdint>
#include <cmath>
#include <vector>
#include <tuple>
#include "cpuASPLqueue.cpp"
#include "cpuASPLpiece.cpp"
#include "cpuASPLbox.cpp"

namespace {
template <class Ev, class G>
int run(Ev &ev, G &g, int64_t nn, int *diam, double *aspl) {
    auto r = ev.diameterASPL(g);
    *diam = r.first;
    *aspl = r.second;
    (void)nn;
    return 0;
}
}

extern "C" {

// part: M rows, row r has lens[r] offsets at offs[r*stride ...].
int ref_part_diameter_aspl(int N, int M, int stride, const int32_t *lens,
                           const int32_t *offs, int *diam, double *aspl) {
    graphgolf::part p;
    p.N = N; p.M = M; p.edges.assign(M, {});
    p.degree = 0; p.regular = true;
    for (int r = 0; r < M; r++)
        for (int i = 0; i < lens[r]; i++) p.edges[r].push_back(offs[(size_t)r * stride + i]);
    if (N <= 1024) { static graphgolf::cpuASPLqueue<1024> ev; return run(ev, p, N, diam, aspl); }
    if (N <= 16384) { static graphgolf::cpuASPLqueue<16384> ev; return run(ev, p, N, diam, aspl); }
    if (N <= 65536) { static graphgolf::cpuASPLqueue<65536> ev; return run(ev, p, N, diam, aspl); }
    return -1;
}

int ref_piece_diameter_aspl(int Nx, int Ny, int Mx, int My, int stride,
                            const int32_t *lens, const int32_t *offs, int *diam,
                            double *aspl) {
    graphgolf::piece p(Nx, Ny, Mx, My);
    for (int x = 0; x < Mx; x++)
        for (int y = 0; y < My; y++) {
            int r = x * My + y;
            for (int i = 0; i < lens[r]; i++)
                p.edges[x][y].emplace_back(offs[((size_t)r * stride + i) * 2],
                                           offs[((size_t)r * stride + i) * 2 + 1]);
        }
    int N = Nx * Ny;
    if (N <= 1024) { static graphgolf::cpuASPLpiece<1024> ev; return run(ev, p, N, diam, aspl); }
    if (N <= 16384) { static graphgolf::cpuASPLpiece<16384> ev; return run(ev, p, N, diam, aspl); }
    if (N <= 65536) { static graphgolf::cpuASPLpiece<65536> ev; return run(ev, p, N, diam, aspl); }
    return -1;
}

int ref_box_diameter_aspl(int Nx, int Ny, int Nz, int Mx, int My, int Mz, int stride,
                          const int32_t *lens, const int32_t *offs, int *diam,
                          double *aspl) {
    graphgolf::box p(Nx, Ny, Nz, Mx, My, Mz);
    for (int x = 0; x < Mx; x++)
        for (int y = 0; y < My; y++)
            for (int z = 0; z < Mz; z++) {
                int r = (x * My + y) * Mz + z;
                for (int i = 0; i < lens[r]; i++) {
                    const int32_t *o = offs + ((size_t)r * stride + i) * 3;
                    p.edges[x][y][z].emplace_back(o[0], o[1], o[2]);
                }
            }
    int N = Nx * Ny * Nz;
    if (N <= 1024) { static graphgolf::cpuASPLbox<1024> ev; return run(ev, p, N, diam, aspl); }
    if (N <= 16384) { static graphgolf::cpuASPLbox<16384> ev; return run(ev, p, N, diam, aspl); }
    if (N <= 65536) { static graphgolf::cpuASPLbox<65536> ev; return run(ev, p, N, diam, aspl); }
    return -1;
}
}
