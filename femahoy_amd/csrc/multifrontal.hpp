// The direct solver for large sparse matrices: a multifrontal LDL^T on the device, in the
// nested-dissection order and with the fundamental supernodes of the reference's sparse Cholesky
// (crates/fem/src/cholesky.rs:44-160, symbolic.rs:41-150, order.rs:19-53).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <memory>
#include <vector>

#include "csr_host.hpp"
#include "symbolic.hpp"

namespace fem {

class Multifrontal {
   public:
    // Factors the symmetric positive definite matrix a (the CSR of the free unknowns, both
    // triangles) in the nested-dissection order of the points xyz (one per unknown), on stream s.
    // A pivot at or below 1e-10 of its diagonal fails with the reference's text
    // (cholesky.rs:92-99). The fronts stay in device memory: at most `budget` bytes.
    Multifrontal(const csr::Csr& a, const std::vector<std::array<double, 3>>& xyz, hipStream_t s, uint64_t budget);
    ~Multifrontal();
    Multifrontal(const Multifrontal&) = delete;
    Multifrontal& operator=(const Multifrontal&) = delete;

    // x = A^-1 b for device vectors of n values, on stream s (b and x may not alias).
    void solve(const double* b, double* x, hipStream_t s) const;

    uint64_t n() const;
    const sym::Counts& counts() const;  // the factor's entries and flops (symbolic.rs:162-183)
    uint64_t front_bytes() const;       // device bytes of the fronts
    uint64_t supernodes() const;
    uint64_t levels() const;            // levels of the supernode tree

   private:
    struct Impl;
    std::unique_ptr<Impl> p_;
};

}  // namespace fem
