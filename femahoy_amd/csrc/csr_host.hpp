// The assembled form of the solver boundary, on the host: the reference's CSR matrix of the free
// unknowns and the description of its unknowns (crates/fem/src/solve.rs:64-77 Unknowns), and the
// setup steps of its grid multigrid that produce integer structures (crates/fem/src/multigrid.rs:
// 84-139 the level loop, 329-452 coarsen; crates/fem/src/sparse.rs:233-292 the Galerkin product).
// These are written to give the reference's integer output bit for bit (prolongation sparsity,
// first-seen coarse numbering, coarse patterns) and its FP64 values in the reference's order of
// operations, so this file is compiled without FMA contraction.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace fem {
namespace csr {

// A square CSR matrix with both triangles, columns sorted in each row (sparse.rs:12-17).
struct Csr {
    uint64_t n = 0;
    std::vector<uint64_t> ptr;
    std::vector<uint32_t> idx;
    std::vector<double> val;
    uint64_t nnz() const { return ptr.empty() ? 0 : ptr[n]; }
};

// A rectangular CSR matrix, rows x cols (sparse.rs:202-208): the prolongation.
struct Rect {
    uint64_t rows = 0, cols = 0;
    std::vector<uint64_t> ptr;
    std::vector<uint32_t> idx;
    std::vector<double> val;
};

// The unknowns of a grid level (multigrid.rs:55-59): the plane coordinates of each axis, and the
// grid point (plane indices) and component of each unknown.
struct GridLevel {
    std::array<std::vector<double>, 3> planes;
    std::vector<std::array<uint32_t, 3>> points;
    std::vector<uint8_t> comp;
};

// Checks a matrix at the boundary: ptr of n + 1 non-decreasing offsets, columns below n and
// increasing in each row, a diagonal entry in every row. Throws fem::Error.
void validate(const Csr& a);
void validate(const Rect& p);
void validate(const GridLevel& g, uint64_t n);

// multigrid.rs:329-452: the next coarser level of `fine` for the matrix `a` and the prolongation
// to `fine`; false when no axis has more than 2 planes. Every other plane plus the last is kept on
// an axis of more than 2 planes. The weights of a point between two coarse planes are the
// geometric ones, or for a scalar the operator's (its row's couplings to the plane below and
// above) when both are positive. Coarse unknowns are numbered in the order the fine rows first
// reach them; each row of the prolongation is sorted by column.
bool coarsen(const GridLevel& fine, const Csr& a, Rect& p, GridLevel& coarse,
             std::array<std::vector<size_t>, 3>* kept = nullptr);

// multigrid.rs:67-70 Cells: the lowest grid point and the 24 x 24 element matrix (unknown 3 n + c,
// corners in hex.rs:34 BITS order) of each cell of a displacement level.
struct Cells {
    std::vector<std::array<uint32_t, 3>> origin;
    std::vector<std::array<double, 576>> k;
};
void validate(const Cells& cells, const GridLevel& g);

// multigrid.rs:458-671 by_cells (with modes and condense): the coarse matrix of a displacement
// from the cells of the fine level: each coarse cell sums its fine cells' energies on its 8
// trilinear corner fields and 9 incompatible modes at the free fine unknowns, drops the modes the
// corner fields span there (Gram-Cholesky test, 1e-6), condenses the rest and is assembled on
// the coarse unknowns. `next` gets the coarse cells and matrices for the next level. The fine
// cells of a coarse cell are summed in increasing fine-cell order.
Csr by_cells(const GridLevel& fine, const Cells& cells, const GridLevel& coarse, const std::array<std::vector<size_t>, 3>& kept,
             Cells& next);

// sparse.rs:233-292: P^T A P by Gustavson's method (a dense accumulator per row of P^T), rows
// sorted. Rows are independent; `threads` host threads share them (the result does not depend
// on the count).
Csr galerkin(const Rect& p, const Csr& a, int threads);

// P^T as a CSR of p.cols rows (the restriction), rows in increasing fine index.
Rect transpose(const Rect& p);

// The diagonal (sparse.rs:89); 0 where a row has no diagonal entry.
std::vector<double> diagonal(const Csr& a);

// The grid multigrid hierarchy of multigrid.rs:84-139: Galerkin coarse matrices, or with
// `cells` (a displacement) those of by_cells; coarsen while the level has more than `coarsest`
// unknowns, fewer than 30 prolongations exist, a coarsening exists and keeps at most 90 % of the
// unknowns. max_levels > 0 also stops at that many levels.
struct GridHierarchy {
    std::vector<Csr> a;        // level 0 (the input) to the coarsest
    std::vector<Rect> p;       // p[l]: level l + 1 to level l
    std::vector<GridLevel> level;
};
GridHierarchy grid_hierarchy(const Csr& a0, const GridLevel& level0, uint64_t coarsest, int max_levels, int threads,
                             const Cells* cells = nullptr);

}  // namespace csr
}  // namespace fem
