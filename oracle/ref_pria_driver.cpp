// Driver around the reference's own CPU PRIA solver (ref_impl.hpp, compiled
// from the reference tree by oracle/Makefile).  TEST INFRASTRUCTURE ONLY: it
// produces golden vectors that pin the oracle; nothing shipped runs it.
//
//   pria_ref gen SEED N T OUT      write ref::random_instance(SEED, N, T)
//   pria_ref solve IN TOL MAXIT OUT  solve IN with pria_params_auto(MAXIT, TOL)
//
// Instance file: int64 magic 0x41495250, n, t; double k0; double k[n];
// double h[n t t] -- the wire format of jrp/harness/backends/pria_cuda.py:25-40.
// Result file: int64 n, t, iterations, status; double primal, dual, gap;
// double x[n t t], y[n t t], z[n t t].
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "pria/pria.h"

/* pria_params_auto lives in pria.cu, which needs CUDA; ref_impl only needs
 * the struct, so the defaults are filled here from pria.h's documentation
 * of them (Parameters.with_auto_step). */
#include "ref_impl.hpp"

static pria_params defaults(int64_t max_iter, double tol) {
    pria_params p;
    p.omega = 0.10835;
    p.eta = 0.999 / std::sqrt(2.0);
    p.beta_sufficient = std::exp(-1.0);
    p.beta_artificial = 0.2468;
    p.beta_necessary = 0.5706;
    p.theta_initial = 0.81930;
    p.theta_target = 0.42818;
    p.theta_decay = 0.90474;
    p.tolerance = tol;
    p.max_iter = max_iter;
    p.use_halpern = 1;
    p.warm_steps = 0;
    p.op_steps = 16;
    p.verbose = 0;
    return p;
}

int main(int argc, char** argv) {
    if (argc >= 6 && !std::strcmp(argv[1], "gen")) {
        ref::Instance in = ref::random_instance(std::strtoull(argv[2], nullptr, 10),
                                                std::atoll(argv[3]), std::atoll(argv[4]));
        FILE* f = std::fopen(argv[5], "wb");
        if (!f) return 2;
        int64_t head[3] = {0x41495250, in.n, in.t};
        std::fwrite(head, sizeof(int64_t), 3, f);
        std::fwrite(&in.k0, sizeof(double), 1, f);
        std::fwrite(in.k.data(), sizeof(double), in.k.size(), f);
        std::fwrite(in.h.data(), sizeof(double), in.h.size(), f);
        std::fclose(f);
        return 0;
    }
    if (argc >= 6 && !std::strcmp(argv[1], "solve")) {
        FILE* f = std::fopen(argv[2], "rb");
        if (!f) return 2;
        int64_t head[3];
        if (std::fread(head, sizeof(int64_t), 3, f) != 3 || head[0] != 0x41495250) return 2;
        ref::Instance in;
        in.n = head[1];
        in.t = head[2];
        in.k.resize(in.n);
        in.h.resize(in.n * in.t * in.t);
        if (std::fread(&in.k0, sizeof(double), 1, f) != 1) return 2;
        if (std::fread(in.k.data(), sizeof(double), in.k.size(), f) != in.k.size()) return 2;
        if (std::fread(in.h.data(), sizeof(double), in.h.size(), f) != in.h.size()) return 2;
        std::fclose(f);
        pria_params p = defaults(std::atoll(argv[4]), std::atof(argv[3]));
        ref::Result r = ref::solve(in, p);
        FILE* o = std::fopen(argv[5], "wb");
        if (!o) return 2;
        int64_t rh[4] = {in.n, in.t, r.iterations, r.gap <= p.tolerance ? 0 : 1};
        double rd[3] = {r.primal, r.dual, r.gap};
        std::fwrite(rh, sizeof(int64_t), 4, o);
        std::fwrite(rd, sizeof(double), 3, o);
        std::fwrite(r.x.data(), sizeof(double), r.x.size(), o);
        std::fwrite(r.y.data(), sizeof(double), r.y.size(), o);
        std::fwrite(r.z.data(), sizeof(double), r.z.size(), o);
        std::fclose(o);
        return 0;
    }
    std::fprintf(stderr, "usage: pria_ref gen SEED N T OUT | solve IN TOL MAXIT OUT\n");
    return 2;
}
