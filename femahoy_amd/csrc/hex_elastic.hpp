// Linear elasticity on the box grid's hex cells, as the reference's box-grid caller builds it
// (crates/fem/src/elastic.rs:287-547): the 8-node hex with 9 incompatible modes condensed to its
// corners (crates/fem/src/hex.rs:146-279), the assembled matrix of the free unknowns, the right
// side of the body force and the face tractions, the grid level of the unknowns and the cells
// that the multigrid's coarse matrices are made from (multigrid.rs:458-671). Host code, FP64, in
// the reference's order of operations and without FMA contraction, so that the matrix and the
// right side equal a sequential assembly in the reference's order bit for bit.
#pragma once

#include <array>
#include <vector>

#include "csr_host.hpp"
#include "model.hpp"

namespace fem {

// The stiffness of an isotropic material by the inverse of its compliance (hex.rs:58-115):
// normal[3][3] row-major, then G_yz, G_xz, G_xy. A material outside the reference's range throws
// its texts (hex.rs:79, hex.rs:93).
std::array<double, 12> isotropic_stiffness(double e, double nu);
// hex.rs:66-115 Orthotropic::stiffness: m = E_x, E_y, E_z, nu_xy, nu_xz, nu_yz, G_xy, G_xz, G_yz;
// the isotropic stiffness is this one with E, nu and E / (2 (1 + nu)) on every axis.
std::array<double, 12> orthotropic_stiffness(const double m[9]);

// hex.rs:199-279 Hex::new (its matrix k): the 33 x 33 stiffness of the 8 corner functions and the
// 3 mode functions of each component on the 2 x 2 x 2 Gauss rule, the modes condensed by a 9 x 9
// Cholesky factor. Row-major, unknown 3 corner + component, corners in the order of a VTK hexahedron.
std::array<double, 576> hex_element(const std::array<double, 3>& size, const std::array<double, 12>& d);

// hex.rs:199-279 with expansion: the stiffness and the load of a unit temperature change at each
// corner (24 x 8, row-major), the modes' share condensed with them, for the expansion coefficients
// alpha along x, y and z.
void hex_element_thermal(const std::array<double, 3>& size, const std::array<double, 12>& d,
                         const std::array<double, 3>& alpha, std::array<double, 576>& k,
                         std::array<double, 192>& load);

// hex.rs:313-340 strain_stress: the engineering strain (xx, yy, zz, yz, xz, xy) and the stress in
// the same order at the local point xi (each coordinate -1 to 1) from the corner displacements u
// (24) and temperature changes t (8): the modes' amplitudes recovered, the strain of all 11
// functions, D times it less the thermal stress of the interpolated temperature change.
void hex_strain_stress(const std::array<double, 3>& size, const std::array<double, 12>& d,
                       const std::array<double, 3>& alpha, const double xi[3], const double u[24], const double t[8],
                       double strain[6], double stress[6]);

// The reduced system of a box model of kind ElasticHex8.
struct HexBoxSystem {
    csr::Csr a;             // Csr::pattern + add_element in cell order, then submatrix(free)
    std::vector<double> rhs;
    csr::GridLevel level;   // the grid point and component of each free unknown
    csr::Cells cells;       // each cell's lowest grid point and matrix, in cell order
    double aspect = 1.0;    // the largest ratio of a cell's longest side to its shortest (elastic.rs:525-536)
};
HexBoxSystem hex_box_system(const BoxModel& m);

// The planes of axis d of a box model: h i for i = 0 .. n (the cell sizes are their differences).
std::vector<double> box_planes(const BoxModel& m, int d);

}  // namespace fem
