// spca_console: command-line front end of the MI355X solver, a drop-in for the reference's
// gpu_console and multicore_console (c++/src/frontends/gpu_console.cu,
// c++/src/frontends/multicore_console.cpp).  Same options, CSV input and output files; the
// solve runs on the GPU through the C ABI (include/spca.h).  Starting points are the
// reference CPU method's (rand_r per starting point, c++/src/gpower/gpower_commons.h:21-47),
// drawn in float, as multicore_console draws them without -d; gpu_console's unseeded rand()
// draws are not reproduced.
//
// Options (reference parser c++/src/utils/option_console_parser.h:43-143, same getopt string
// plus -p and -e):
//   -i FILE  input CSV (required)          -o FILE  output prefix (required)
//   -f 1..8  formulation (required): 1 L0 cons / L1 var, 2 L0 cons / L2 var, 3 L1 cons / L1,
//            4 L1 cons / L2, 5 L0 pen / L1, 6 L0 pen / L2, 7 L1 pen / L1, 8 L1 pen / L2
//   -s N     constraint s          -g X  penalty gamma        -m N  max iterations
//   -t X     tolerance             -l N  starting points      -r N  batch size
//   -u N     OTF (0/1)             -d X  double precision: rejected (exit 2), see below
//   -v X     verbose               -x N  distributed grid (accepted, not used: one GPU)
//   -p P     products: fp32 (default), bf16x3, fp16 (no reference counterpart)
//   -e       echo the parsed settings and the data shape, then exit without a GPU call
//            (no reference counterpart)
// Defaults are the reference OptimizationSettings constructor's
// (c++/src/class/optimization_settings.h:100-122).
//
// -d: the reference runs its whole solve in double then (gpu_console.cu:97-101,
// multicore_console.cpp:52-56).  This solver's products are FP32, bf16x3 or fp16 MFMA and
// its iterates FP32, so a double solve does not exist here; -d exits with status 2 and a
// message instead of returning FP32 results under a "Double precision: 1" label.  The FP32
// form (-p fp32) with the FP64 explained variance and FP64 support re-solve is the closest
// this solver offers.
//
// Outputs (reference c++/src/utils/file_reader.h:100-146):
//   <prefix>_x                       "index,value" for every nonzero of x (16 digits)
//   <prefix>_optimizationStatistics  settings, timing and result, in the reference's lines;
//     times in seconds as the labels say (multicore_console writes seconds; gpu_console
//     writes its cudaEvent loop time, milliseconds, under the same "sec" label,
//     c++/src/gpugpower/gpu_sparse_PCA_solver.h:188-193).
#include <getopt.h>
#include <unistd.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "spca.h"

namespace {

struct ConsoleOptions {
    spca_settings st;
    std::string input, output;
    bool have_input = false, have_output = false, have_formulation = false;
    bool use_double = false, verbose = false, echo = false;
    int grid = 0;
};

void print_usage() {
    std::cout << "Usage: spca_console -i data.csv -o result_prefix -f 1..8 [-s s] [-g gamma] [-m maxIter]\n"
                 "       [-t tol] [-l startingPoints] [-r batchSize] [-u otf] [-d 1] [-v 1] [-p fp32|bf16x3|fp16] [-e]\n";
}

// reference -f numbering -> spca_formulation (option_console_parser.h:103-137)
int formulation_of_flag(int f) {
    switch (f) {
        case 1: return L0_constrained_L1_PCA;
        case 2: return L0_constrained_L2_PCA;
        case 3: return L1_constrained_L1_PCA;
        case 4: return L1_constrained_L2_PCA;
        case 5: return L0_penalized_L1_PCA;
        case 6: return L0_penalized_L2_PCA;
        case 7: return L1_penalized_L1_PCA;
        case 8: return L1_penalized_L2_PCA;
        default: return -1;
    }
}

int parse(ConsoleOptions& o, int argc, char** argv) {
    spca_settings_default(&o.st);
    // reference defaults
    o.st.tolerance = 0.01;
    o.st.constraint = 10;
    o.st.penalty = 0;
    o.st.total_sp = 64;
    o.st.batch = 64;
    o.st.use_otf = 0;
    o.st.max_iter = 20;
    o.st.start = SPCA_START_REFERENCE;
    int c;
    while ((c = getopt(argc, argv, "i:f:o:m:t:l:r:u:v:d:s:g:x:p:e")) != -1) {
        switch (c) {
            case 'x': o.grid = atoi(optarg); break;
            case 'l': o.st.total_sp = atoi(optarg); break;
            case 'r': o.st.batch = atoi(optarg); break;
            case 's': o.st.constraint = (unsigned)atoi(optarg); break;
            case 'g': o.st.penalty = atof(optarg); break;
            case 'm': o.st.max_iter = atoi(optarg); break;
            case 'i': o.input = optarg; o.have_input = true; break;
            case 'o': o.output = optarg; o.have_output = true; break;
            case 'u': o.st.use_otf = atoi(optarg); break;
            case 't': o.st.tolerance = atof(optarg); break;
            case 'd': o.use_double = true; break;
            case 'v': o.verbose = true; break;
            case 'f': {
                const int f = formulation_of_flag(atoi(optarg));
                if (f >= 0) {
                    o.st.formulation = f;
                    o.have_formulation = true;
                }
                break;
            }
            case 'p': {
                const std::string p = optarg;
                if (p == "fp32" || p == "0") o.st.precision = SPCA_PREC_FP32;
                else if (p == "bf16x3" || p == "1") o.st.precision = SPCA_PREC_BF16X3;
                else if (p == "fp16" || p == "2") o.st.precision = SPCA_PREC_FP16;
                else {
                    std::cerr << "unknown products form: " << p << "\n";
                    return 1;
                }
                break;
            }
            case 'e': o.echo = true; break;
            default: break;
        }
    }
    if (!o.have_input || !o.have_output || !o.have_formulation) {
        print_usage();
        return 1;
    }
    if (o.st.batch > o.st.total_sp) o.st.total_sp = o.st.batch;
    return 0;
}

// Reference CSV reader (file_reader.h:31-86): m = lines, n = comma-separated cells of the
// first line; values read with operator>> in the reading precision, row by row, skipping
// one separator character after every value but the last of a row.
template <class F>
bool read_csv(const std::string& path, unsigned& m, unsigned& n, std::vector<float>& B) {
    m = n = 0;
    {
        std::ifstream data(path);
        if (!data.is_open()) return false;
        std::string line;
        while (std::getline(data, line)) {
            if (m == 0) {
                std::stringstream ls(line);
                std::string cell;
                while (std::getline(ls, cell, ',')) n++;
            }
            m++;
        }
    }
    B.assign((size_t)m * n, 0.f);
    std::ifstream in(path);
    if (!in.is_open()) return false;
    for (unsigned row = 0; row < m; row++)
        for (unsigned col = 0; col < n; col++) {
            F v = 0;
            in >> v;
            char sep;
            if (col < n - 1) in >> sep;
            B[row + (size_t)col * m] = (float)v;
        }
    return true;
}

std::string suffixed(const std::string& base, const char* suffix) { return base + "_" + suffix; }

void save_x(const std::string& prefix, const std::vector<float>& x) {
    std::ofstream f(suffixed(prefix, "x"));
    for (size_t i = 0; i < x.size(); i++)
        if (x[i] != 0) f << i << "," << std::setprecision(16) << x[i] << '\n';
}

void save_statistics(const std::string& prefix, const ConsoleOptions& o, const spca_settings& st,
                     const spca_stats& s, double elapsed_s) {
    std::ofstream f(suffixed(prefix, "optimizationStatistics"));
    f << "Solver options " << '\n';
    f << "Formulation: " << st.formulation << '\n';
    if (spca_is_constrained(st.formulation)) f << "Constraint parameter: " << st.constraint << '\n';
    else f << "penaltyParameter parameter: " << st.penalty << '\n';
    f << "Max iterations per starting point: " << st.max_iter << '\n';
    f << "Starting points: " << st.total_sp << '\n';
    f << "Batch size: " << st.batch << '\n';
    f << "Batching strategy (OTF): " << st.use_otf << '\n';
    f << "Double precision: " << 0 << '\n';
    f << "tolerance: " << st.tolerance << '\n';
    f << '\n' << "Timing " << '\n';
    f << "Total computational time: " << std::setprecision(16) << s.loop_ms * 1e-3 << " sec" << '\n';
    f << "Total elapsed time: " << std::setprecision(16) << elapsed_s << " sec" << '\n';
    f << '\n' << "Result " << '\n';
    f << "Objective value: " << std::setprecision(16) << s.fval << '\n';
    f << "Elapsed it (total): " << s.it << '\n';
    f << "Average it (per starting point): " << std::setprecision(16)
      << s.it * st.batch / (0.0 + st.total_sp) << '\n';
}

}  // namespace

int main(int argc, char** argv) {
    const auto t0 = std::chrono::steady_clock::now();
    ConsoleOptions o;
    if (parse(o, argc, argv)) return 1;
    unsigned m = 0, n = 0;
    std::vector<float> B;
    const bool ok = read_csv<float>(o.input, m, n, B);
    if (!ok || m == 0 || n == 0) {
        std::cerr << "cannot read a matrix from " << o.input << "\n";
        return 1;
    }
    if (o.use_double) {
        std::cerr << "double precision (-d) is not provided: products run in FP32, bf16x3 or fp16 (-p) with FP32 "
                     "iterates; run without -d\n";
        return 2;
    }
    spca_settings st = o.st;
    spca_settings_check(&st, n);  // the reference clamp: s <= n, batches, total rounded up
    if (o.echo) {
        std::cout << "m " << m << " n " << n << " formulation " << st.formulation << " s " << st.constraint
                  << " gamma " << st.penalty << " max_iter " << st.max_iter << " tol " << st.tolerance
                  << " total_sp " << st.total_sp << " batch " << st.batch << " otf " << st.use_otf << " double 0"
                  << " precision " << st.precision << "\n";
        return 0;
    }
    spca_ctx ctx = nullptr;
    int rc = spca_ctx_create(0, 0, &ctx);
    if (rc != SPCA_OK) {
        std::cerr << "spca_ctx_create failed (" << rc << ")\n";
        return 2;
    }
    std::vector<float> x(n, 0.f);
    spca_stats s;
    double fval = 0;
    rc = spca_dense_solver_host(ctx, B.data(), (int)m, x.data(), m, n, &st, &s, &fval);
    if (rc != SPCA_OK) {
        std::cerr << "solve failed: " << spca_last_error(ctx) << "\n";
        spca_ctx_destroy(ctx);
        return 2;
    }
    if (o.verbose)
        std::cout << "objective " << std::setprecision(16) << s.fval << " iterations " << s.it << " loop " << s.loop_ms
                  << " ms\n";
    spca_ctx_destroy(ctx);
    const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    save_x(o.output, x);
    save_statistics(o.output, o, st, s, elapsed);
    return 0;
}
