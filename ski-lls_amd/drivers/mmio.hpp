// mmio.hpp -- minimal MatrixMarket reader for the command-line drivers.
// Supports "matrix array real general" and "matrix coordinate
// {real,integer,pattern} {general,symmetric}".  Matrices with more columns
// than rows are transposed, as the reference drivers do
// (test/ls_blendenpik_hashing_sparse_input.cpp:43-48).
#pragma once

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

struct CscMatrix {
    long m = 0, n = 0;
    std::vector<long> p, i;
    std::vector<double> x;
};

inline CscMatrix read_matrix_market(const std::string &path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open " + path);
    std::string line;
    std::getline(in, line);
    std::string lower = line;
    std::transform(lower.begin(), lower.end(), lower.begin(), ::tolower);
    if (lower.rfind("%%matrixmarket", 0) != 0) throw std::runtime_error("not a MatrixMarket file");
    bool array = lower.find(" array") != std::string::npos;
    bool pattern = lower.find(" pattern") != std::string::npos;
    bool symmetric = lower.find(" symmetric") != std::string::npos;
    if (lower.find(" complex") != std::string::npos)
        throw std::runtime_error("complex matrices are not supported");
    do {
        std::getline(in, line);
    } while (in && (line.empty() || line[0] == '%'));
    std::istringstream hdr(line);
    long m, n, nz = 0;
    hdr >> m >> n;
    if (!array) hdr >> nz;
    std::vector<long> ri, ci;
    std::vector<double> vv;
    if (array) {
        for (long j = 0; j < n; ++j)
            for (long r = 0; r < m; ++r) {
                double v;
                in >> v;
                if (v != 0.0) {
                    ri.push_back(r);
                    ci.push_back(j);
                    vv.push_back(v);
                }
            }
    } else {
        for (long e = 0; e < nz; ++e) {
            long r, c;
            double v = 1.0;
            in >> r >> c;
            if (!pattern) in >> v;
            ri.push_back(r - 1);
            ci.push_back(c - 1);
            vv.push_back(v);
            if (symmetric && r != c) {
                ri.push_back(c - 1);
                ci.push_back(r - 1);
                vv.push_back(v);
            }
        }
    }
    if (!in && !in.eof()) throw std::runtime_error("malformed MatrixMarket data in " + path);
    bool transpose = m < n;
    CscMatrix A;
    A.m = transpose ? n : m;
    A.n = transpose ? m : n;
    if (transpose) std::swap(ri, ci);
    // counting sort by column, rows in increasing order within a column
    std::vector<long> order(vv.size());
    for (size_t e = 0; e < order.size(); ++e) order[e] = (long)e;
    std::stable_sort(order.begin(), order.end(), [&](long a, long b) {
        return ci[a] != ci[b] ? ci[a] < ci[b] : ri[a] < ri[b];
    });
    A.p.assign(A.n + 1, 0);
    for (long c : ci) A.p[c + 1]++;
    for (long j = 0; j < A.n; ++j) A.p[j + 1] += A.p[j];
    for (long e : order) {
        A.i.push_back(ri[e]);
        A.x.push_back(vv[e]);
    }
    return A;
}

inline std::vector<double> to_dense(const CscMatrix &A) {
    std::vector<double> D((size_t)A.m * A.n, 0.0);
    for (long j = 0; j < A.n; ++j)
        for (long q = A.p[j]; q < A.p[j + 1]; ++q) D[(size_t)(A.i[q] + j * A.m)] += A.x[q];
    return D;
}
