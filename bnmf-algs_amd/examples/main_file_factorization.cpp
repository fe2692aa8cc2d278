// main_file_factorization.cpp — command-line factorization of a matrix file
// with the library's GPU solvers (the bld_mult, bld_add and bld_appr branches
// of the reference's example/main_file_factorization.cpp:19-119).
//
//   main_file_factorization ALG FILE N_COMPONENTS BETA MAX_ITER
//
// ALG is bld_mult, bld_add or bld_appr (the reference's other algorithms,
// nmf, seq_greedy_bld, collapsed_gibbs and collapsed_icm, are not part of
// this library). FILE holds one matrix row per line, entries separated by
// whitespace. The model parameters are the example's: a = 100, b = 1,
// alpha = 0.05 (x entries), beta = 10 (z entries, the last 60); bld_mult runs
// with psi_appr; each solver is followed by bld_fact. BETA (the NMF beta of
// the reference's nmf branch) is accepted and unused.
//
// S.txt, W.txt, H.txt and L.txt are written to the working directory in the
// layout the reference's `file << tensor` / `file << matrix` produces
// (Eigen's default stream format): one line per index of the first
// dimension (S: x lines of y*z values, j outer and k inner; W: x lines of z;
// H: z lines of y; L: one line of y), values in the stream's default
// notation and precision (6 significant digits), each right-aligned to the
// widest value of the object and separated by one space, lines joined by
// newlines with none after the last. BNMF_CLI_PRECISION=<digits> (e.g. 17)
// writes more digits in the same layout.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "bnmf_algs/bnmf_algs.hpp"

using namespace bnmf_algs;

namespace {

bool read_matrix(const std::string& path, std::vector<double>& data, long& rows, long& cols) {
    std::ifstream in(path);
    if (!in) return false;
    std::string line;
    rows = 0;
    cols = -1;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        double v;
        long n = 0;
        while (ls >> v) {
            data.push_back(v);
            ++n;
        }
        if (n == 0) continue;
        if (cols >= 0 && n != cols) return false;
        cols = n;
        ++rows;
    }
    return rows > 0;
}

// Eigen's default stream output of a rows x cols row-major block: every
// value formatted as the stream would, padded on the left to the widest one
void write_eigen(const std::string& path, const double* m, long rows, long cols, int precision) {
    std::ofstream out(path, std::ios_base::trunc);
    if (rows * cols == 0) return;
    std::vector<std::string> txt((size_t)(rows * cols));
    size_t width = 0;
    for (long t = 0; t < rows * cols; ++t) {
        std::ostringstream v;
        v.precision(precision);
        v << m[t];
        txt[(size_t)t] = v.str();
        width = std::max(width, txt[(size_t)t].size());
    }
    for (long r = 0; r < rows; ++r) {
        for (long c = 0; c < cols; ++c) {
            if (c) out << ' ';
            out << std::setw((int)width) << txt[(size_t)(r * cols + c)];
        }
        if (r + 1 < rows) out << '\n';
    }
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 6) {
        std::cout << "usage: " << argv[0] << " alg filename n_components beta max_iter\n\n"
                  << "where alg is one of <bld_mult | bld_add | bld_appr>" << std::endl;
        return -1;
    }
    const std::string alg(argv[1]);
    if (alg != "bld_mult" && alg != "bld_add" && alg != "bld_appr") {
        std::cout << "Invalid algorithm: this library provides bld_mult, bld_add and bld_appr"
                  << std::endl;
        return -1;
    }
    const std::string filename(argv[2]);
    const size_t z = std::stoul(argv[3]);
    (void)std::stod(argv[4]);
    const size_t max_iter = std::stoul(argv[5]);

    std::vector<double> data;
    long x = 0, y = 0;
    if (!read_matrix(filename, data, x, y)) {
        std::cout << "cannot read a rectangular matrix from " << filename << std::endl;
        return -1;
    }
    matrixd X(x, y);
    for (long t = 0; t < x * y; ++t) X.data()[t] = data[(size_t)t];

    std::vector<double> alpha_dirichlet((size_t)x, 0.05);
    std::vector<double> beta_dirichlet(z, 10);
    beta_dirichlet.back() = 60;
    alloc_model::Params<double> params(100, 1, alpha_dirichlet, beta_dirichlet);

    std::cout << "Computing the factorization" << std::endl;
    const auto t0 = std::chrono::high_resolution_clock::now();
    tensord<3> S;
    matrixd W, H;
    vectord L;
    try {
        if (alg == "bld_mult")
            S = bld::bld_mult(X, z, params, max_iter, true);
        else if (alg == "bld_add")
            S = bld::bld_add(X, z, params, max_iter);
        else
            S = std::get<0>(bld::bld_appr(X, z, params, max_iter));
        std::tie(W, H, L) = bld::bld_fact(S, params);
    } catch (const std::exception& e) {
        std::cout << "error: " << e.what() << std::endl;
        return -1;
    }
    const auto t1 = std::chrono::high_resolution_clock::now();
    std::cout << "Total time: "
              << std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count()
              << " milliseconds" << std::endl;

    int precision = 6;
    if (const char* e = std::getenv("BNMF_CLI_PRECISION")) precision = std::max(1, std::atoi(e));
    write_eigen("S.txt", S.data(), x, y * (long)z, precision);
    write_eigen("W.txt", W.data(), x, (long)z, precision);
    write_eigen("H.txt", H.data(), (long)z, y, precision);
    write_eigen("L.txt", L.data(), 1, y, precision);
    return 0;
}
