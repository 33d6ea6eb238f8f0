// gg_formats.hpp -- the reference's text formats for the four graph
// encodings, with the public fields its drivers and evaluators use.
// Formats (whitespace separated):
//   part  : "N M", then M rows "E off_1 .. off_E"              (cppcode/part.hpp:13-43)
//   piece : "Nx Ny Mx My", then Mx*My rows "E dx dy .. "      (cppcode/piece.hpp:18-52)
//   box   : "Nx Ny Nz Mx My Mz", then rows "E dx dy dz .."    (cppcode/box.hpp:37-80)  
//   graph : one edge "u v" per line, N = max id + 1          (cppcode/graph.hpp:15-51)
// 'degree' is the largest row length for part and the first row's length for
// piece and box, as the reference computes it (part.hpp:21-25, piece.hpp:28-31,
// box.hpp:51-54).
#pragma once
#include <algorithm>
#include <iostream>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

namespace graphgolf {

struct part {
    int N = 0, M = 0, degree = 0;
    bool regular = true;
    std::vector<std::vector<int>> edges;
    void load(std::istream &in) {
        in >> N >> M;
        edges.assign(M, {});
        degree = 0;
        regular = true;
        for (int i = 0; i < M; i++) {
            int E;
            in >> E;
            if (!degree) degree = E;
            else if (degree != E) { regular = false; degree = std::max(degree, E); }
            edges[i].resize(E);
            for (int j = 0; j < E; j++) in >> edges[i][j];
        }
    }
    void print(std::ostream &out) const {
        out << N << ' ' << M << '\n';
        for (const auto &r : edges) {
            out << r.size();
            for (int x : r) out << ' ' << x;
            out << '\n';
        }
    }
};

struct piece {
    int Nx = 0, Ny = 0, Mx = 0, My = 0, degree = 0;
    bool regular = true;
    std::vector<std::vector<std::vector<std::pair<int, int>>>> edges;
    void load(std::istream &in) {
        in >> Nx >> Ny >> Mx >> My;
        edges.assign(Mx, std::vector<std::vector<std::pair<int, int>>>(My));
        degree = 0;
        for (int x = 0; x < Mx; x++)
            for (int y = 0; y < My; y++) {
                int E;
                in >> E;
                if (!degree) degree = E;
                else if (degree != E) regular = false;
                for (int e = 0; e < E; e++) {
                    int dx, dy;
                    in >> dx >> dy;
                    edges[x][y].emplace_back(dx, dy);
                }
            }
    }
    // piece.hpp:41-52
    void print(std::ostream &out) const {
        out << Nx << ' ' << Ny << ' ' << Mx << ' ' << My << '\n';
        for (int x = 0; x < Mx; x++)
            for (int y = 0; y < My; y++) {
                out << edges[x][y].size();
                for (const auto &e : edges[x][y]) out << ' ' << e.first << ' ' << e.second;
                out << '\n';
            }
    }
};

struct box {
    int Nx = 0, Ny = 0, Nz = 0, Mx = 0, My = 0, Mz = 0, N = 0, M = 0, degree = 0;
    bool regular = true;
    std::vector<std::vector<std::vector<std::vector<std::tuple<int, int, int>>>>> edges;
    void load(std::istream &in) {
        in >> Nx >> Ny >> Nz >> Mx >> My >> Mz;
        N = Nx * Ny * Nz;
        M = Mx * My * Mz;
        edges.assign(Mx, std::vector<std::vector<std::vector<std::tuple<int, int, int>>>>(
                             My, std::vector<std::vector<std::tuple<int, int, int>>>(Mz)));
        for (int x = 0; x < Mx; x++)
            for (int y = 0; y < My; y++)
                for (int z = 0; z < Mz; z++) {
                    int E;
                    in >> E;
                    if (!degree) degree = E;
                    else if (degree != E) regular = false;
                    for (int e = 0; e < E; e++) {
                        int dx, dy, dz;
                        in >> dx >> dy >> dz;
                        edges[x][y][z].emplace_back(dx, dy, dz);
                    }
                }
    }
    // box.hpp:65-80
    void print(std::ostream &out) const {
        out << Nx << ' ' << Ny << ' ' << Nz << ' ' << Mx << ' ' << My << ' ' << Mz << '\n';
        for (int x = 0; x < Mx; x++)
            for (int y = 0; y < My; y++)
                for (int z = 0; z < Mz; z++) {
                    out << edges[x][y][z].size();
                    for (const auto &e : edges[x][y][z])
                        out << ' ' << std::get<0>(e) << ' ' << std::get<1>(e) << ' ' << std::get<2>(e);
                    out << '\n';
                }
    }
};

struct graph {
    int N = 0, degree_max = 0;
    std::vector<std::vector<int>> edges;
    void load(std::istream &in) {
        std::string line;
        std::vector<std::pair<int, int>> es;
        N = 0;
        while (std::getline(in, line)) {
            std::stringstream ss(line);
            int u, v;
            if (!(ss >> u >> v)) continue;
            es.emplace_back(u, v);
            N = std::max({N, u, v});
        }
        N++;
        edges.assign(N, {});
        for (auto &e : es) {
            edges[e.first].push_back(e.second);
            edges[e.second].push_back(e.first);
        }
        degree_max = 0;
        for (auto &r : edges) degree_max = std::max(degree_max, (int)r.size());
    }
};

}  // namespace graphgolf
