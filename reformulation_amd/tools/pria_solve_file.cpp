// File CLI of the PRIA solver: the wire format, usage, flags, stdout lines
// and exit statuses of jrp/pria-cuda/tools/pria_solve_file.cu:1-43, so the
// harness backend that drives that program (jrp/harness/backends/pria_cuda.py)
// drives this one unchanged.
//
// Instance file (little-endian): int64 magic 0x41495250, int64 n, int64 t,
//   double k0, double k[n], double h[n t t] (row-major, upper triangle read).
// Result file: int64 n, int64 t, int64 iterations, int64 status, double lower,
//   double upper, double gap [, double x[n t t], y[n t t], z[n t t] with the
//   `solution` flag]; status 0 when the gap reached tol, else 1.
// Usage: pria_solve_file <instance> <result> [tol] [max_iter]
//        [f32|vanilla|verbose|solution]...
// Exit status: 0 converged, 1 iteration limit, 2 usage or unreadable input,
//   3 solver error.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "pria_hip.h"

namespace {

constexpr int64_t kMagic = 0x41495250;

double since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr,
                     "usage: %s <instance> <result> [tol] [max_iter] "
                     "[f32|vanilla|verbose|solution]...\n",
                     argv[0]);
        return 2;
    }
    double tol = 1e-4;
    int64_t max_iter = -1;
    bool f32 = false, vanilla = false, verbose = false, solution = false;
    if (argc > 3) tol = std::atof(argv[3]);
    if (argc > 4) max_iter = std::atoll(argv[4]);
    for (int q = 5; q < argc; ++q) {
        if (!std::strcmp(argv[q], "f32")) f32 = true;
        else if (!std::strcmp(argv[q], "vanilla")) vanilla = true;
        else if (!std::strcmp(argv[q], "verbose")) verbose = true;
        else if (!std::strcmp(argv[q], "solution")) solution = true;
        else {
            std::fprintf(stderr, "pria_solve_file: unknown flag %s\n", argv[q]);
            return 2;
        }
    }

    auto t0 = std::chrono::steady_clock::now();
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) {
        std::fprintf(stderr, "pria_solve_file: cannot open %s\n", argv[1]);
        return 2;
    }
    int64_t head[3] = {0, 0, 0};
    double k0 = 0.0;
    std::vector<double> k, h;
    bool ok = std::fread(head, sizeof(int64_t), 3, f) == 3 && head[0] == kMagic && head[1] > 0 &&
              head[2] > 0 && std::fread(&k0, sizeof(double), 1, f) == 1;
    if (ok) {
        k.resize(size_t(head[1]));
        h.resize(size_t(head[1]) * size_t(head[2]) * size_t(head[2]));
        ok = std::fread(k.data(), sizeof(double), k.size(), f) == k.size() &&
             std::fread(h.data(), sizeof(double), h.size(), f) == h.size();
    }
    std::fclose(f);
    if (!ok) {
        std::fprintf(stderr, "pria_solve_file: %s is not a PRIA instance file\n", argv[1]);
        return 2;
    }
    const int64_t n = head[1], t = head[2];
    const size_t size = size_t(n) * size_t(t) * size_t(t);
    pria_params p = pria_params_auto(max_iter, tol);
    p.use_halpern = vanilla ? 0 : 1;
    p.verbose = verbose ? 1 : 0;
    (void)hipFree(nullptr);  // the runtime's one-off start belongs to setup
    const double setup = since(t0);

    std::vector<double> x, y, z;
    if (solution) {
        x.resize(size);
        y.resize(size);
        z.resize(size);
    }
    // the reference's own entry points, as its CLI calls them: FP64 inputs
    // and iterates, or the FP32 entry with inputs rounded to float
    pria_result_info info;
    t0 = std::chrono::steady_clock::now();
    int status;
    if (f32) {
        std::vector<float> kf(k.begin(), k.end()), hf(h.begin(), h.end());
        std::vector<float> xf, yf, zf;
        if (solution) {
            xf.resize(size);
            yf.resize(size);
            zf.resize(size);
        }
        status = pria_solve_f32(n, t, kf.data(), float(k0), hf.data(), &p, nullptr, nullptr,
                                nullptr, solution ? xf.data() : nullptr,
                                solution ? yf.data() : nullptr, solution ? zf.data() : nullptr,
                                &info);
        for (size_t g = 0; solution && g < size; ++g) {
            x[g] = xf[g];
            y[g] = yf[g];
            z[g] = zf[g];
        }
    } else {
        status = pria_solve_f64(n, t, k.data(), k0, h.data(), &p, nullptr, nullptr, nullptr,
                                solution ? x.data() : nullptr, solution ? y.data() : nullptr,
                                solution ? z.data() : nullptr, &info);
    }
    const double solve = since(t0);
    if (status != 0) {
        std::fprintf(stderr, "pria_solve_file: pria_solve failed with status %d\n", status);
        return 3;
    }
    // primal_objective is the objective at the returned point (the upper
    // bound), dual_objective the lower bound
    const double lower = info.dual_objective, upper = info.primal_objective;
    const int64_t converged = std::isfinite(info.relative_gap) && info.relative_gap <= tol ? 0 : 1;

    FILE* out = std::fopen(argv[2], "wb");
    if (!out) {
        std::fprintf(stderr, "pria_solve_file: cannot write %s\n", argv[2]);
        return 2;
    }
    const int64_t words[4] = {n, t, info.iterations, converged};
    const double bounds[3] = {lower, upper, info.relative_gap};
    ok = std::fwrite(words, sizeof(int64_t), 4, out) == 4 &&
         std::fwrite(bounds, sizeof(double), 3, out) == 3;
    if (ok && solution)
        ok = std::fwrite(x.data(), sizeof(double), size, out) == size &&
             std::fwrite(y.data(), sizeof(double), size, out) == size &&
             std::fwrite(z.data(), sizeof(double), size, out) == size;
    ok = std::fclose(out) == 0 && ok;
    if (!ok) {
        std::fprintf(stderr, "pria_solve_file: short write to %s\n", argv[2]);
        return 2;
    }
    std::printf("items: %lld\nperiods: %lld\nprecision: %s\nstatus: %s\niterations: %lld\n",
                (long long)n, (long long)t, f32 ? "float32" : "float64",
                converged == 0 ? "converged" : "iteration_limit", (long long)info.iterations);
    std::printf("lower: %.17g\nupper: %.17g\nrelative_gap: %.6e\nsetup_seconds: %.6f\n"
                "solve_seconds: %.6f\n",
                lower, upper, info.relative_gap, setup, solve);
    return int(converged);
}
