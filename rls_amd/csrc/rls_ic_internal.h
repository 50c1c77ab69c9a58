// The incomplete-Cholesky factor object behind the opaque rls_ic handle (built in rls_ic.cpp,
// uploaded by rls_set_ic in rls_engine.hip).
#pragma once
#include <complex>
#include <cstdint>
#include <vector>

struct rls_ic {
    int64_t n = 0;
    std::vector<int64_t> colptr, rowind;   // 0-based CSC of L, diagonal first per column
    std::vector<std::complex<double>> values;
    std::vector<int64_t> p;                // column ordering, p[k] = original index at k
    double alpha = 0.0;                    // diagonal shift finally used
    int n_shifts = 0;                      // shift increases after breakdowns
};
