// The solver of the assembled boundary: Solver::new(a, unknowns, method, crossover, None) of
// crates/fem/src/solve.rs:92-183 for a CSR matrix of the free unknowns, on the device.
#pragma once

#include <cstdint>
#include <memory>
#include <string>

#include "../../include/fem_solver.h"
#include "csr_host.hpp"
#include "solver_common.hpp"
#include "symbolic.hpp"

namespace fem {

// The unknowns of the boundary (solve.rs:64-77) on the host.
struct CsrUnknowns {
    int kind = FEM_UNKNOWNS_GRID;
    csr::GridLevel grid;                // GRID
    csr::Rect p;                        // QUADRATIC: the prolongation from the linear unknowns
    bool coarse_grid = false;           // QUADRATIC: the linear unknowns' grid level is given
    csr::GridLevel coarse;              // ... and here it is
    std::vector<std::array<double, 3>> xyz;  // POINTS, QUADRATIC: the coordinates of each unknown, if given
    bool cells = false;                      // Solver::new's cells were given
    csr::Cells cell_data;                    // ... and here they are
};

// The coordinates of each unknown of a grid level (solve.rs:143-150: the planes of its point).
std::vector<std::array<double, 3>> grid_xyz(const csr::GridLevel& g);

// The direct-or-multigrid choice of Solver::new for FEM_METHOD_AUTO (solve.rs:106-118, 158-170):
// Crossover::Unknowns, or Crossover::Flops on the counts of the nested-dissection factor of `a`
// (which are returned in `counts`, with `analysed` set).
bool auto_picks_direct(const csr::Csr& a, const std::vector<std::array<double, 3>>& xyz, const Options& o,
                       sym::Counts& counts, bool& analysed);

class CsrSolverImpl;

class CsrSolver {
   public:
    CsrSolver(csr::Csr a, CsrUnknowns u, const Options& o);
    ~CsrSolver();
    int method() const;
    uint64_t n() const;
    Options& options();
    void solve_device(const double* d_b, double* d_x, int (*cancelled)(void*), void* ctx, fem_stats* st);
    void solve_host(const double* b, double* x, int (*cancelled)(void*), void* ctx, fem_stats* st);
    void apply_host(const double* x, double* y);
    void precondition_host(const double* r, double* z);
    void level_info(int level, uint64_t* unknowns, double* lmax) const;
    const char* device_name() const;
    // The factor counts a FLOPS crossover compared; false when it did not analyse the matrix.
    bool crossover_counts(uint64_t* entries, double* flops) const;
    // The host hierarchy (levels' matrices and prolongations) the device levels were made from.
    const csr::GridHierarchy& host_levels() const;

   private:
    std::unique_ptr<CsrSolverImpl> impl_;
};

// The boundary's C structs on the host (validated by the solver).
csr::Csr csr_from(const fem_csr* a);
csr::Rect rect_from(const fem_csr* p);
csr::GridLevel grid_from(const fem_grid_level* g, uint64_t n);
csr::Cells cells_from(const fem_cells* c);
CsrUnknowns unknowns_from(const fem_unknowns* u, uint64_t n);

}  // namespace fem
