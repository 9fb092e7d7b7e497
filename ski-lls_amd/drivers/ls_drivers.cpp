// ls_drivers.cpp -- command-line drivers over the ski-lls.h entry points, with
// the parameters, right-hand side (b = ones) and output lines of the
// reference's drivers:
//   ls_blendenpik_hashing <matrix.mtx> [gamma] [k]
//       test/ls_blendenpik_hashing.cpp (dense robust solver, k=1, gamma=1.7)
//   ls_blendenpik_hashing_sparse_input <matrix.mtx> [gamma] [k]
//       test/ls_blendenpik_hashing_sparse_input.cpp (same on a densified matrix)
//   ls_blendenpik <matrix.mtx> [gamma]
//       test/ls_blendenpik.cpp (ls_dense_blendenpik, gamma=2.2)
//   ls_qr <matrix.mtx> [gamma] [k] [abs_tol] [ordering]
//       test/ls_qr.cpp (ls_sparse_spqr, k=2, gamma=1.4)
// The program name (argv[0]'s basename) selects the driver; the first
// argument "--driver=<name>" overrides it.
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "../../include/ski-lls.h"
#include "mmio.hpp"

static double now() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

static int dense_driver(const char *label, int argc, char **argv) {
    CscMatrix S = read_matrix_market(argv[1]);
    std::cout << "number of rows is: " << S.m << std::endl;
    std::cout << "number of columns is: " << S.n << std::endl;
    std::cout << "number of non-zeros is: " << S.i.size() << std::endl;
    std::vector<double> data = to_dense(S);
    Mat_d A(S.m, S.n, data.data());
    std::vector<double> bd((size_t)S.m, 1.0);
    Vec_d b(S.m, bd.data());
    long k = 1;
    double gamma = 1.7;
    long max_it = MAX_IT;
    double it_tol = IT_TOL, rcond = 1e-12, abs_tol = ABS_TOL;
    int wisdom = WISDOM, debug = 0;
    if (argc >= 3) gamma = std::atof(argv[2]);
    if (argc >= 4) k = std::atol(argv[3]);
    std::cout << "Solver Name: Ls_blendenpik_hashing" << std::endl;
    std::cout << "NNZ_PER_COLUMN is: " << k << std::endl;
    std::cout << "OVER_SAMPLING_RATIO is: " << gamma << std::endl;
    long it, rank;
    int flag;
    Vec_d x(A.n());
    double t0 = now();
    ls_dense_hashing_blendenpik(A, b, x, rank, flag, it, gamma, k, abs_tol, rcond, it_tol, max_it,
                                debug, wisdom);
    double t1 = now();
    std::cout << "Iteration flag is: " << flag << std::endl;
    std::cout << "Time taken by " << label << " is: " << (t1 - t0) << std::endl;
    A.mv('n', 1, x, -1, b);
    double r = 0;
    for (long i = 0; i < b.n(); ++i) r += b(i) * b(i);
    std::cout.precision(6);
    std::cout << "Residual of " << label << " is: " << std::sqrt(r) << std::endl;
    std::cout << "Iteration of " << label << " is: " << it << std::endl;
    std::cout << "Rank of " << label << " is: " << rank << std::endl;
    return 0;
}

static int blendenpik_driver(int argc, char **argv) {
    CscMatrix S = read_matrix_market(argv[1]);
    std::cout << "number of rows is: " << S.m << std::endl;
    std::cout << "number of columns is: " << S.n << std::endl;
    std::cout << "number of non-zeros is: " << S.i.size() << std::endl;
    std::vector<double> data = to_dense(S);
    Mat_d A(S.m, S.n, data.data());
    std::vector<double> bd((size_t)S.m, 1.0);
    Vec_d b(S.m, bd.data());
    double gamma = 2.2;
    long max_it = MAX_IT;
    double it_tol = IT_TOL;
    int wisdom = WISDOM, debug = 0;
    if (argc >= 3) gamma = std::atof(argv[2]);
    std::cout << "Solver Name: Ls_blendenpik" << std::endl;
    std::cout << "OVER_SAMPLING_RATIO is: " << gamma << std::endl;
    long it, rank;
    int flag;
    Vec_d x(A.n());
    double t0 = now();
    ls_dense_blendenpik(A, b, x, rank, flag, it, gamma, it_tol, max_it, debug, wisdom);
    double t1 = now();
    std::cout << "Iteration flag is: " << flag << std::endl;
    std::cout << "Time taken by ls_blendenpik is: " << (t1 - t0) << std::endl;
    A.mv('n', 1, x, -1, b);
    double r = 0;
    for (long i = 0; i < b.n(); ++i) r += b(i) * b(i);
    std::cout.precision(6);
    std::cout << "Residual of ls_blendenpik is: " << std::sqrt(r) << std::endl;
    std::cout << "Iteration of ls_blendenpik is: " << it << std::endl;
    return 0;
}

static int sparse_driver(int argc, char **argv) {
    CscMatrix S = read_matrix_market(argv[1]);
    std::cout << "number of rows is: " << S.m << std::endl;
    std::cout << "number of columns is: " << S.n << std::endl;
    std::cout << "number of non-zeros is: " << S.i.size() << std::endl;
    cs_dl A = {(long)S.i.size(), S.m, S.n, S.p.data(), S.i.data(), S.x.data(), -1};
    std::vector<double> bd((size_t)S.m, 1.0);
    Vec_d b(S.m, bd.data());
    long k = 2;
    double gamma = 1.4;
    long max_it = MAX_IT;
    double it_tol = IT_TOL, abs_tol = ABS_TOL;
    int ordering = SPQR_ORDERING;
    if (argc >= 3) gamma = std::atof(argv[2]);
    if (argc >= 4) k = std::atol(argv[3]);
    if (argc >= 5) abs_tol = std::atof(argv[4]);
    if (argc >= 6) ordering = std::atoi(argv[5]);
    std::cout << "Solver Name: Ls_qr" << std::endl;
    std::cout << "NNZ_PER_COLUMN is: " << k << std::endl;
    std::cout << "OVER_SAMPLING_RATIO is: " << gamma << std::endl;
    long it, rank;
    int flag;
    Vec_d x(S.n);
    double t0 = now();
    ls_sparse_spqr(A, b, x, rank, flag, it, gamma, k, abs_tol, ordering, it_tol, max_it,
                   RCOND_THRESHOLD, PERTURB, 0);
    double t1 = now();
    std::cout << "Iteration flag is: " << flag << std::endl;
    std::cout << "Time taken by ls_sparse_spqr is: " << (t1 - t0) << std::endl;
    // residual b - A x (b = ones)
    std::vector<double> r((size_t)S.m, -1.0);
    for (long j = 0; j < S.n; ++j)
        for (long q = S.p[j]; q < S.p[j + 1]; ++q) r[(size_t)S.i[q]] += S.x[q] * x(j);
    double s = 0;
    for (double v : r) s += v * v;
    std::cout.precision(6);
    std::cout << "Residual of ls_sparse_spqr is: " << std::sqrt(s) << std::endl;
    std::cout << "Iteration of ls_sparse_spqr is: " << it << std::endl;
    std::cout << "Rank of ls_sparse_spqr is: " << rank << std::endl;
    return 0;
}

int main(int argc, char **argv) {
    std::string name = argv[0];
    size_t slash = name.find_last_of('/');
    if (slash != std::string::npos) name = name.substr(slash + 1);
    if (argc >= 2 && std::strncmp(argv[1], "--driver=", 9) == 0) {
        name = argv[1] + 9;
        ++argv;
        --argc;
    }
    if (argc < 2) {
        std::cerr << "usage: " << name << " <matrix.mtx> [gamma] [k] ..." << std::endl;
        return 2;
    }
    try {
        if (name.find("ls_qr") != std::string::npos) return sparse_driver(argc, argv);
        if (name.find("hashing") == std::string::npos &&
            name.find("ls_blendenpik") != std::string::npos)
            return blendenpik_driver(argc, argv);
        if (name.find("sparse_input") != std::string::npos)
            return dense_driver("ls_blendenpik_hashing", argc, argv);
        return dense_driver("ls_blendenpik_hashing", argc, argv);
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
}
