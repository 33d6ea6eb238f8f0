// graphgolf_compat.hpp -- drop-in replacements for the reference's evaluator
// classes, over the C ABI of libgg_mi355x.
//
// A driver of the reference (cppcode/annealing.cpp, HillClimb.cpp,
// pieceannealing.cpp, boxHillClimb.cpp, main.cpp, piece2aspl.cpp, box2aspl.cpp,
// aspl_cuda.cpp) compiles unchanged against this header instead of
// cudaASPLconv.hpp / cudaASPLpiece.hpp / cudaASPLbox.hpp / cudaASPLgraph.hpp:
// same namespace, class names, constructor arguments, method signatures and
// sentinels.  The graph arguments are duck-typed (any type with the
// reference's public fields: part{N,M,edges}, piece{edges[x][y] of pairs},
// box{edges[x][y][z] of tuples}, graph{edges}), exactly as the reference's
// evaluators take them.
//
// Errors the reference never checks (no device, HIP failures) throw
// std::runtime_error here.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>
#include "graphgolf_mi355x.h"

namespace graphgolf {

namespace detail {
inline void check(int st, const char *what) {
    if (st < 0) throw std::runtime_error(std::string(what) + ": " + gg_strerror(st));
}
// Reference sentinels: diameterASPL (cudacode/cudaASPLconv.cu:229,234),
// WVCdiameterASPL (:293,298).
constexpr int kBadDiameter = 100000000;
constexpr double kBadASPL = 100000000.0;
}  // namespace detail

// cudacode/cudaASPLconv.hpp:6-23
class cudaASPLconv {
    gg_eval *ev_ = nullptr;
    std::vector<int32_t> tab_;

    template <class Part>
    gg_result run(Part &p, int level_limit) {
        // packing of cudaASPLconv.cu:198-202: rows padded with offset 0
        for (int i = 0; i < M; i++)
            for (int j = 0; j < degree; j++)
                tab_[(size_t)i * degree + j] = j < (int)p.edges[i].size() ? p.edges[i][j] : 0;
        detail::check(gg_eval_set_option(ev_, GG_OPT_LEVEL_LIMIT, level_limit), "level limit");
        gg_result r;
        detail::check(gg_eval_part(ev_, tab_.data(), &r), "gg_eval_part");
        return r;
    }

public:
    int N, M, degree;
    cudaASPLconv(int N, int M, int degree, int device = 0) : N(N), M(M), degree(degree) {
        detail::check(gg_create_part(N, M, degree, device, &ev_), "gg_create_part");
        tab_.resize((size_t)M * degree);
    }
    ~cudaASPLconv() { gg_destroy(ev_); }
    cudaASPLconv(const cudaASPLconv &) = delete;
    cudaASPLconv &operator=(const cudaASPLconv &) = delete;

    // cudaASPLconv.cu:151-196: ASPL, or 1e9 when unconnected (:182) / too deep
    template <class Part>
    double calc(Part &p) {
        gg_result r = run(p, GG_LEVEL_LIMIT);
        if (r.status != GG_OK) return 1e9;
        return gg_aspl(&r, N);
    }
    // cudaASPLconv.cu:197-243
    template <class Part>
    std::pair<int, double> diameterASPL(Part &p) {
        gg_result r = run(p, GG_LEVEL_LIMIT);
        if (r.status != GG_OK) return {detail::kBadDiameter, detail::kBadASPL};
        return {r.diameter, gg_aspl(&r, N)};
    }
    // cudaASPLconv.cu:244-308 (levels up to 100, :264)
    template <class Part>
    std::tuple<int, int, double> WVCdiameterASPL(Part &p) {
        gg_result r = run(p, 100);
        if (r.status != GG_OK)
            return std::make_tuple(detail::kBadDiameter, detail::kBadDiameter, detail::kBadASPL);
        return std::make_tuple(r.wvc, r.diameter, gg_aspl(&r, N));
    }
};

// cudacode/cudaASPLpiece.hpp:6-21
class cudaASPLpiece {
    gg_eval *ev_ = nullptr;
    std::vector<int32_t> tab_;

public:
    int Nx, Ny, Mx, My, degree;
    cudaASPLpiece(int Nx, int Ny, int Mx, int My, int degree, int device = 0)
        : Nx(Nx), Ny(Ny), Mx(Mx), My(My), degree(degree) {
        detail::check(gg_create_piece(Nx, Ny, Mx, My, degree, device, &ev_), "gg_create_piece");
        tab_.resize((size_t)Mx * My * degree * 2);
    }
    ~cudaASPLpiece() { gg_destroy(ev_); }
    cudaASPLpiece(const cudaASPLpiece &) = delete;
    cudaASPLpiece &operator=(const cudaASPLpiece &) = delete;

    // cudaASPLpiece.cu:159-216; packing :163-174 (rows truncated / padded to degree)
    template <class Piece>
    std::pair<int, double> diameterASPL(Piece &p) {
        for (int x = 0; x < Mx; x++)
            for (int y = 0; y < My; y++)
                for (int e = 0; e < degree; e++) {
                    int dx = 0, dy = 0;
                    if (e < (int)p.edges[x][y].size()) std::tie(dx, dy) = p.edges[x][y][e];
                    tab_[((size_t)(x * My + y) * degree + e) * 2] = dx;
                    tab_[((size_t)(x * My + y) * degree + e) * 2 + 1] = dy;
                }
        gg_result r;
        detail::check(gg_eval_piece(ev_, tab_.data(), &r), "gg_eval_piece");
        if (r.status != GG_OK) return {detail::kBadDiameter, detail::kBadASPL};
        return {r.diameter, gg_aspl(&r, Nx * Ny)};
    }
};

// cudacode/cudaASPLbox.hpp:6-21
class cudaASPLbox {
    gg_eval *ev_ = nullptr;
    std::vector<int32_t> tab_;

public:
    int Nx, Ny, Nz, Mx, My, Mz, degree;
    cudaASPLbox(int Nx, int Ny, int Nz, int Mx, int My, int Mz, int degree, int device = 0)
        : Nx(Nx), Ny(Ny), Nz(Nz), Mx(Mx), My(My), Mz(Mz), degree(degree) {
        detail::check(gg_create_box(Nx, Ny, Nz, Mx, My, Mz, degree, device, &ev_),
                      "gg_create_box");
        tab_.resize((size_t)Mx * My * Mz * degree * 3);
    }
    ~cudaASPLbox() { gg_destroy(ev_); }
    cudaASPLbox(const cudaASPLbox &) = delete;
    cudaASPLbox &operator=(const cudaASPLbox &) = delete;

    // cudaASPLbox.cu:166-226
    template <class Box>
    std::pair<int, double> diameterASPL(Box &p) {
        for (int x = 0; x < Mx; x++)
            for (int y = 0; y < My; y++)
                for (int z = 0; z < Mz; z++)
                    for (int e = 0; e < degree; e++) {
                        int dx = 0, dy = 0, dz = 0;
                        if (e < (int)p.edges[x][y][z].size())
                            std::tie(dx, dy, dz) = p.edges[x][y][z][e];
                        size_t o = ((size_t)((x * My + y) * Mz + z) * degree + e) * 3;
                        tab_[o] = dx;
                        tab_[o + 1] = dy;
                        tab_[o + 2] = dz;
                    }
        gg_result r;
        detail::check(gg_eval_box(ev_, tab_.data(), &r), "gg_eval_box");
        if (r.status != GG_OK) return {detail::kBadDiameter, detail::kBadASPL};
        return {r.diameter, gg_aspl(&r, Nx * Ny * Nz)};
    }
};

// cudacode/cudaASPLgraph.hpp:5-19: explicit adjacency, all N sources,
// returns (diameter, unscaled total); (-1,-1) unconnected, (-2,-2) too deep
// (cudaASPLgraph.cu:182,190); rows padded with self-loops (:144-153).
class cudaASPLgraph {
    gg_eval *ev_ = nullptr;
    std::vector<int32_t> adj_;
    int N_, dmax_;

public:
    cudaASPLgraph(int N, int degree_max, int device = 0) : N_(N), dmax_(degree_max) {
        detail::check(gg_create_adj(N, degree_max, device, &ev_), "gg_create_adj");
        detail::check(gg_eval_set_option(ev_, GG_OPT_LEVEL_LIMIT, 199), "level limit");
        adj_.resize((size_t)N * degree_max);
    }
    ~cudaASPLgraph() { gg_destroy(ev_); }
    cudaASPLgraph(const cudaASPLgraph &) = delete;
    cudaASPLgraph &operator=(const cudaASPLgraph &) = delete;

    template <class Graph>
    std::pair<int, int64_t> calc(Graph &g) {
        for (int i = 0; i < N_; i++)
            for (int j = 0; j < dmax_; j++)
                adj_[(size_t)i * dmax_ + j] = j < (int)g.edges[i].size() ? g.edges[i][j] : i;
        gg_result r;
        detail::check(gg_eval_adj(ev_, adj_.data(), &r), "gg_eval_adj");
        if (r.status == GG_UNCONNECTED) return {-1, -1};
        if (r.status == GG_DIAMETER_LIMIT) return {-2, -2};
        return {r.diameter, r.total_reps};
    }
};

}  // namespace graphgolf
