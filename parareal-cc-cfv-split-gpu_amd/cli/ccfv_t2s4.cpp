// ccfv_t2s4 — command-line driver with the reference program's interface
// (T2S4_Parareal.cu:989-1366), on the C ABI of libccfv.
//
//   ccfv_t2s4 config.txt [--slices N] [--kmax K] [--tol TOL] [--no-files]
//
// The config file holds "key:value" lines with the reference's keys T_min,
// T_max, N (Nx = Ny = N), Nt, T_span and K (:1026-1049).  The run is the
// reference's: the reference problem on [-1,1]^2 with velocity (-4y, 4x)
// and the rotating Gaussian (:27-56); the initial cell averages go to
// data_T2S4_000000.txt and the serial fine solution every T_span steps to
// data_T2S4_%06d.txt (:93-114, :1234-1243); then Parareal with 40 slices of
// J = Nt/40 fine steps and 8 iterations (:1250-1253), one solve, printing
// per iteration the max-norm distance of U^k(T) to the serial fine solution
// (:1340-1346), and the closing lines of :1359-1366.  The initial data are
// the exact solution at t = 0 whatever T_min is (:1147).  Unlike the
// reference, a key missing
// from the config file is an error (the reference reads an uninitialised
// value), --slices/--kmax/--tol change the Parareal plan, and the time line
// reports wall-clock seconds.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "../../include/ccfv.h"

namespace {

int die(const char *what)
{
    std::fprintf(stderr, "%s failed: %s\n", what, ccfv_last_error());
    return 2;
}

// data_T2S4_%06d.txt: "x y C" per cell, x slow, a blank line after each x
// (T2S4_Parareal.cu:93-114)
bool write_field(const std::vector<double> &u, int step, int nx, int ny, double xlo,
                 double xhi, double ylo, double yhi)
{
    char name[64];
    std::snprintf(name, sizeof(name), "data_T2S4_%06d.txt", step);
    FILE *fp = std::fopen(name, "w");
    if (!fp) return false;
    const double dx = (xhi - xlo) / nx, dy = (yhi - ylo) / ny;
    for (int i = 0; i < nx; i++) {
        const double x = xlo + dx * (i + 0.5);
        for (int j = 0; j < ny; j++) {
            const double y = ylo + dy * (j + 0.5);
            std::fprintf(fp, "%.6lf %.6lf %.4e\n", x, y, u[(size_t)i * ny + j]);
        }
        std::fprintf(fp, "\n");
    }
    return std::fclose(fp) == 0;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::cout << "please input config file name" << std::endl;
        return 1;
    }
    int slices = 40, kmax = 8;
    double tol = 0.0;
    bool files = true;
    for (int a = 2; a < argc; a++) {
        if (!std::strcmp(argv[a], "--slices") && a + 1 < argc) slices = std::atoi(argv[++a]);
        else if (!std::strcmp(argv[a], "--kmax") && a + 1 < argc) kmax = std::atoi(argv[++a]);
        else if (!std::strcmp(argv[a], "--tol") && a + 1 < argc) tol = std::atof(argv[++a]);
        else if (!std::strcmp(argv[a], "--no-files")) files = false;
        else {
            std::fprintf(stderr, "unknown argument: %s\n", argv[a]);
            return 1;
        }
    }
    // configuration (T2S4_Parareal.cu:1018-1055)
    std::ifstream cfg(argv[1]);
    if (!cfg.is_open()) {
        std::cout << "Cannot open config file!" << std::endl;
        return 1;
    }
    std::cout << "open config file ok" << std::endl;
    double T_min = 0, T_max = 0, K = 0;
    int N = 0, Nt = 0, T_span = 0;
    std::map<std::string, bool> seen;
    std::string line;
    while (std::getline(cfg, line)) {
        if (line.empty()) continue;
        const size_t pos = line.find(':');
        if (pos == std::string::npos) continue;
        const std::string key = line.substr(0, pos), val = line.substr(pos + 1);
        if (key == "T_min") seen[key] = std::sscanf(val.c_str(), "%lf", &T_min) == 1;
        if (key == "T_max") seen[key] = std::sscanf(val.c_str(), "%lf", &T_max) == 1;
        if (key == "N") seen[key] = std::sscanf(val.c_str(), "%d", &N) == 1;
        if (key == "Nt") seen[key] = std::sscanf(val.c_str(), "%d", &Nt) == 1;
        if (key == "T_span") seen[key] = std::sscanf(val.c_str(), "%d", &T_span) == 1;
        if (key == "K") seen[key] = std::sscanf(val.c_str(), "%lf", &K) == 1;
    }
    for (const char *k : {"T_min", "T_max", "N", "Nt", "T_span", "K"})
        if (!seen[k]) {
            std::fprintf(stderr, "config: missing or unreadable key %s\n", k);
            return 1;
        }
    if (Nt < 1 || T_span < 1 || !(T_max > T_min) || slices < 1 || Nt / slices < 1 ||
        kmax < 0 || kmax > CCFV_MAX_ITERS) {
        std::fprintf(stderr, "config: need Nt >= slices >= 1, T_span >= 1, T_max > T_min, "
                             "0 <= kmax <= 64\n");
        return 1;
    }

    ccfv_problem P;
    ccfv_ic ic;
    if (ccfv_problem_reference(&P, N, K) || ccfv_ic_reference(&ic, K)) return die("setup");
    ccfv_ctx *c = nullptr;
    if (ccfv_create(&P, &c)) return die("ccfv_create");
    const double dx = (P.x_hi - P.x_lo) / P.nx, dy = (P.y_hi - P.y_lo) / P.ny;
    const double dt = (T_max - T_min) / Nt;  // :1060
    std::vector<double> u((size_t)P.nx * P.ny), u0(u.size()), uref(u.size());
    const auto t_start = std::chrono::steady_clock::now();

    // initial cell averages of the exact solution at t = 0 (:1147-1152)
    if (ccfv_fill_exact(c, &ic, 0.0) || ccfv_get_state(c, u0.data())) return die("initial data");
    if (files) write_field(u0, 0, P.nx, P.ny, P.x_lo, P.x_hi, P.y_lo, P.y_hi);

    // serial fine solution, output every T_span steps (:1234-1243)
    for (int n0 = 0; n0 < Nt;) {
        const int n1 = std::min(Nt, (n0 / T_span + 1) * T_span);
        if (ccfv_advance(c, T_min, dt, n0, n1)) return die("ccfv_advance");
        if (n1 % T_span == 0 && files) {
            if (ccfv_get_state(c, u.data())) return die("ccfv_get_state");
            write_field(u, n1, P.nx, P.ny, P.x_lo, P.x_hi, P.y_lo, P.y_hi);
        }
        n0 = n1;
    }
    if (ccfv_get_state(c, uref.data())) return die("ccfv_get_state");

    // Parareal: slices of J = Nt/slices fine steps of DT/J (:1250-1253), one
    // solve keeping the end state of every iteration (:1296-1346)
    const int J = Nt / slices;
    if (ccfv_set_state(c, u0.data())) return die("ccfv_set_state");
    if (ccfv_set_option(c, CCFV_OPT_KEEP_ITERATES, 1)) return die("ccfv_set_option");
    ccfv_parareal_report rep;
    const int rc = ccfv_parareal_solve(c, T_min, T_max, J * slices, slices, kmax, tol, &rep);
    if (rc && rc != CCFV_ENOTCONV) return die("ccfv_parareal_solve");
    for (int k = 1; k <= rep.iterations; k++) {
        if (ccfv_parareal_iterate(c, k, u.data())) return die("ccfv_parareal_iterate");
        double err = 0.0;
        for (size_t i = 0; i < u.size(); i++) err = std::fmax(err, std::fabs(u[i] - uref[i]));
        std::printf("Parareal iter %2d: err(T) = %.3e\n", k, err);
    }
    std::fflush(stdout);
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();

    // closing lines (:1359-1366)
    std::cout << "K:" << K << std::fixed << std::setprecision(16) << std::endl;
    std::cout << "dx:" << dx << std::endl;
    std::cout << "dy:" << dy << std::endl;
    std::cout << "dt:" << dt << std::endl;
    std::cout << "Nx:" << P.nx << std::endl;
    std::cout << "Ny:" << P.ny << std::endl;
    std::cout << "Nt:" << Nt << std::endl;
    std::cout << "Total time (including Parareal): " << secs << " s" << std::endl;
    ccfv_destroy(c);
    return 0;
}
