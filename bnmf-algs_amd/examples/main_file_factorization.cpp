// main_file_factorization.cpp — command-line factorization of a matrix file
// with the GPU bld_mult solver (the bld_mult branch of the reference's
// example/main_file_factorization.cpp:19-119).
//
//   main_file_factorization bld_mult FILE N_COMPONENTS BETA MAX_ITER
//
// FILE holds one matrix row per line, entries separated by whitespace. The
// model parameters are the example's: a = 100, b = 1, alpha = 0.05 (x
// entries), beta = 10 (z entries, the last 60); bld_mult runs with
// psi_appr, then bld_fact. Writes S.txt (x*y lines of z values: fiber S_ij:
// per line, rows i outer), W.txt (x lines of z), H.txt (z lines of y) and
// L.txt (one line of y) to the working directory. BETA (the NMF beta of the
// reference's other algorithms) is accepted and unused by bld_mult.
#include <chrono>
#include <cstdio>
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

template <typename M>
void write_rows(const std::string& path, const M& m, long rows, long cols) {
    std::ofstream out(path, std::ios_base::trunc);
    out << std::setprecision(17);
    for (long r = 0; r < rows; ++r) {
        for (long c = 0; c < cols; ++c) out << (c ? " " : "") << m.data()[r * cols + c];
        out << '\n';
    }
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 6) {
        std::cout << "usage: " << argv[0] << " alg filename n_components beta max_iter\n\n"
                  << "where alg is bld_mult (the GPU solver of this library)" << std::endl;
        return -1;
    }
    const std::string alg(argv[1]);
    if (alg != "bld_mult") {
        std::cout << "Invalid algorithm: this build provides bld_mult only" << std::endl;
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
        S = bld::bld_mult(X, z, params, max_iter, true);
        std::tie(W, H, L) = bld::bld_fact(S, params);
    } catch (const std::exception& e) {
        std::cout << "error: " << e.what() << std::endl;
        return -1;
    }
    const auto t1 = std::chrono::high_resolution_clock::now();
    std::cout << "Total time: "
              << std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count()
              << " milliseconds" << std::endl;

    write_rows("S.txt", S, x * y, (long)z);
    write_rows("W.txt", W, x, (long)z);
    write_rows("H.txt", H, (long)z, y);
    write_rows("L.txt", L, 1, y);
    return 0;
}
