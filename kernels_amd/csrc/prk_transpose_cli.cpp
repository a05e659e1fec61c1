// prk-transpose: the driver of Cxx11/transpose-hip.cc with the same command
// line, checks, messages and report, running on the gfx950 library through
// its C ABI.
//
//   prk-transpose <# iterations> <matrix order> [variant (0/1/2)]
//
// Argument handling follows Cxx11/transpose-hip.cc:136-167, the run
// :205-219, the validation and report :235-266.  The reference's three
// kernel variants (naive, coalesced, no bank conflicts) are accepted and
// reported; every one runs the library's LDS-tiled kernel, which also
// handles orders that are not a multiple of the tile (the reference's
// launch grid covers only order/64 tiles per side).
#include <hip/hip_runtime.h>

#include <array>
#include <cstdlib>
#include <iostream>
#include <string>

#include "prk_stencil.h"

namespace {

const int kMaxOrder = 46340;  // floor(sqrt(INT_MAX)), Cxx11/prk_util.h:373-377

int fail(const char *what, int rc)
{
    std::cout << "ERROR: " << what << ": " << prk_last_error() << " (" << rc << ")" << std::endl;
    return 1;
}

}  // namespace

int main(int argc, char *argv[])
{
    std::cout << "Parallel Research Kernels" << std::endl;
    std::cout << "MI355X/HIP Matrix transpose: B = A^T" << std::endl;

    int iterations, order, variant;
    try {
        if (argc < 3) throw "Usage: <# iterations> <matrix order> [variant (0/1/2)]";
        iterations = std::atoi(argv[1]);
        if (iterations < 1) throw "ERROR: iterations must be >= 1";
        order = std::atoi(argv[2]);
        if (order <= 0) throw "ERROR: Matrix Order must be greater than 0";
        else if (order > kMaxOrder) throw "ERROR: matrix dimension too large - overflow risk";
        variant = 2;
        if (argc > 3) variant = std::atoi(argv[3]);
        if (variant < 0 || variant > 2)
            throw "Please select a valid variant (0: naive 1: coalesced, 2: no bank conflicts)";
    } catch (const char *e) {
        std::cout << e << std::endl;
        return 1;
    }

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        std::cout << "ERROR: no HIP device" << std::endl;
        return 1;
    }
    const std::array<std::string, 3> vnames = {"naive", "coalesced", "no bank conflicts"};
    std::cout << "Number of iterations  = " << iterations << std::endl;
    std::cout << "Matrix order          = " << order << std::endl;
    std::cout << "Variant               = " << vnames[variant] << " (LDS-tiled gfx950 kernel)" << std::endl;

    const size_t bytes = (size_t)order * (size_t)order * sizeof(double);
    double *A = nullptr, *B = nullptr;
    if (hipMalloc(&A, bytes) != hipSuccess || hipMalloc(&B, bytes) != hipSuccess) {
        std::cout << "ERROR: device allocation of " << 2 * bytes << " bytes failed" << std::endl;
        return 1;
    }
    int rc = prk_transpose_init(order, A, B, nullptr);
    if (rc) return fail("init", rc);
    double avgtime = 0.0;
    rc = prk_transpose_run(order, iterations, A, B, nullptr, &avgtime);
    if (rc) return fail("run", rc);
    double abserr = 0.0;
    rc = prk_transpose_check(order, iterations, B, &abserr, nullptr);
    if (rc) return fail("check", rc);
    (void)hipFree(A);
    (void)hipFree(B);

    const double epsilon = 1.0e-8;
    if (abserr < epsilon) {
        std::cout << "Solution validates" << std::endl;
        std::cout << "Rate (MB/s): " << 1.0e-6 * (4.0 * (double)bytes) / avgtime << " Avg time (s): " << avgtime
                  << std::endl;
    } else {
        std::cout << "ERROR: Aggregate squared error " << abserr << " exceeds threshold " << epsilon << std::endl;
        return 1;
    }
    return 0;
}
