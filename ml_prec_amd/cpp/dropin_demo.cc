// A driver written against the C++ drop-in API in the same sequence as the
// reference's project/main.cc:46-187: load the matrix (CSR BINARY, transformed
// to Skyline), build the preconditioner from a Config, run the outer solver.
//
// usage: dropin_demo <matrix.bin> <rhs.txt> <out_x.bin> <uh_cheb|multi_split|diag|id> <pcg|gmres>
//                    <eps> [prec_conf_file] [fp32]
// Prints "niter <n> true_rel_res <r>" on success.
#include "ml_prec.h"

#include <fstream>
#include <iostream>

static void load_vector(Vector& v, const std::string& filename) {   // ASCII vector (vector.cc:52-70)
    std::ifstream is(filename.c_str());
    if (!is.good()) THROW_EXCEPTION("Problem reading file \"" << filename << "\"");
    std::string line;
    uint n;
    std::getline(is, line);
    is >> n;
    v.resize(n);
    std::getline(is, line);
    std::getline(is, line);
    for (uint i = 0; i < n; i++) is >> v[i];
}

int main(int argc, char** argv) {
    if (argc < 7) {
        std::cerr << "usage: dropin_demo matrix.bin rhs.txt out_x.bin prec solver eps [conf] [fp32]" << std::endl;
        return 2;
    }
    try {
        Config cfg;
        std::string prec = argv[4], solver = argv[5];
        double eps = std::atof(argv[6]);
        if (argc > 7 && std::string(argv[7]) != "-") cfg.prec_conf_file = argv[7];
        if (argc > 8 && std::string(argv[8]) == "fp32") cfg.gpu_precision = MLP_FP32;

        SkylineMatrix A;
        A.load(argv[1], BINARY, true);
        Vector b;
        load_vector(b, argv[2]);

        PrecBase* B = nullptr;
        if (prec == "uh_cheb") B = new Prec(A, cfg);
        else if (prec == "multi_split") B = new MultiSplitPrec(A, cfg);
        else if (prec == "diag") B = new DiagPrec(A);
        else if (prec == "id") B = new IdPrec(A);
        else if (prec == "gs") B = new GSPrec(A);
        else if (prec == "comp") B = new CompositePrec(A, cfg);
        else if (prec == "sym_split") B = new SymPrec(A, cfg);
        else THROW_EXCEPTION("Unknown prec type \"" << prec << "\"");

        Vector x(A.size());
        SolverStats stats;
        if (solver == "pcg") PCGSolver(A, b, *B, x, stats, eps, NORM_L2, true);
        else if (solver == "gmres") GMRESSolver(A, b, *B, x, stats, eps, NORM_L2, true);
        else if (solver == "simple") SimpleSolver(A, b, *B, x, stats, eps, NORM_L2, true);
        else if (solver == "cheb") {
            // ChebSolver needs the spectral bounds of the Chebyshev preconditioner (project/main.cc:142-149)
            Prec& Bcheb = dynamic_cast<Prec&>(*B);
            ChebSolver(A, Bcheb.lmin(), Bcheb.lmax(), b, Bcheb, x, eps, NORM_L2, true);
        } else THROW_EXCEPTION("Unknown solver type \"" << solver << "\"");

        Vector z(A.size());
        B->solve(b, z);                    // one standalone preconditioner apply, written after x
        std::ofstream os(argv[3], std::ofstream::binary);
        os.write(reinterpret_cast<const char*>(x.data()), x.size() * sizeof(double));
        os.write(reinterpret_cast<const char*>(z.data()), z.size() * sizeof(double));
        std::cout << "niter " << stats.niter << " true_rel_res " << stats.true_rel_res << std::endl;
        delete B;
    } catch (const Exception& e) {
        std::cerr << "Exception: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
