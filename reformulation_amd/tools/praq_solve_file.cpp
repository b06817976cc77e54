// File CLI of the PRAQ solver: the wire format, usage, stdout lines and exit
// statuses of qubo/praq-cuda/tools/praq_solve_file.cu:1-31, so the harness
// backend that drives that program (qubo/harness/backends/praq_cuda.py:26-35)
// drives this one unchanged.
//
// Instance file (little-endian): int64 magic 0x51415250, int64 n, int64 m,
//   double diag[n], int64 rows[m], int64 cols[m], double vals[m].
// Result file: int64 n, int64 iterations, int64 status, double lower,
//   double upper, double x[n]; written only when the solver ran.
// Usage: praq_solve_file <instance> <result> [tol] [max_iter] [f32]
// Exit status: 0 converged, 1 iteration limit, 2 usage or unreadable input,
//   3 solver error.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "praq_hip.h"

namespace {

constexpr int64_t kMagic = 0x51415250;

double since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

template <class T>
bool read_n(FILE* f, std::vector<T>& v, int64_t count) {
    v.resize(size_t(count));
    return count == 0 || std::fread(v.data(), sizeof(T), v.size(), f) == v.size();
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3 || argc > 6) {
        std::fprintf(stderr, "usage: %s <instance> <result> [tol] [max_iter] [f32]\n", argv[0]);
        return 2;
    }
    PraqParameters p = praq_default_parameters();
    if (argc > 3) p.tolerance = std::atof(argv[3]);
    if (argc > 4) p.max_iter = std::atoll(argv[4]);
    if (argc > 5) p.use_float32 = std::strcmp(argv[5], "f32") == 0;

    auto t0 = std::chrono::steady_clock::now();
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) {
        std::fprintf(stderr, "praq_solve_file: cannot open %s\n", argv[1]);
        return 2;
    }
    int64_t head[3] = {0, 0, 0};
    std::vector<double> diag, vals;
    std::vector<int64_t> rows, cols;
    bool ok = std::fread(head, sizeof(int64_t), 3, f) == 3 && head[0] == kMagic && head[1] >= 0 &&
              head[2] >= 0 && read_n(f, diag, head[1]) && read_n(f, rows, head[2]) &&
              read_n(f, cols, head[2]) && read_n(f, vals, head[2]);
    std::fclose(f);
    if (!ok) {
        std::fprintf(stderr, "praq_solve_file: %s is not a PRAQ instance file\n", argv[1]);
        return 2;
    }
    PraqInstance inst{head[1], head[2], diag.data(), rows.data(), cols.data(), vals.data()};
    (void)hipFree(nullptr);  // the runtime's one-off start belongs to setup
    const double setup = since(t0);

    std::vector<double> x(size_t(inst.n));
    PraqResult res;
    t0 = std::chrono::steady_clock::now();
    const int status = praq_solve(&inst, &p, x.data(), &res);
    const double solve = since(t0);
    if (status < 0) {
        std::fprintf(stderr, "praq_solve_file: praq_solve failed with status %d\n", status);
        return 3;
    }

    FILE* out = std::fopen(argv[2], "wb");
    if (!out) {
        std::fprintf(stderr, "praq_solve_file: cannot write %s\n", argv[2]);
        return 2;
    }
    const int64_t words[3] = {inst.n, res.iterations, int64_t(res.status)};
    const double bounds[2] = {res.lower, res.upper};
    ok = std::fwrite(words, sizeof(int64_t), 3, out) == 3 &&
         std::fwrite(bounds, sizeof(double), 2, out) == 2 &&
         std::fwrite(x.data(), sizeof(double), x.size(), out) == x.size();
    ok = std::fclose(out) == 0 && ok;
    if (!ok) {
        std::fprintf(stderr, "praq_solve_file: short write to %s\n", argv[2]);
        return 2;
    }
    std::printf("n: %lld\nedges: %lld\nprecision: %s\nstatus: %s\niterations: %lld\n",
                (long long)inst.n, (long long)inst.m, p.use_float32 ? "float32" : "float64",
                res.status == 0 ? "converged" : "iteration_limit", (long long)res.iterations);
    std::printf("lower: %.17g\nupper: %.17g\nsetup_seconds: %.6f\nsolve_seconds: %.6f\n", res.lower,
                res.upper, setup, solve);
    return res.status;
}
