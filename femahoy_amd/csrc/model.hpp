// The structured box model on the host: the reference's numbering of points and unknowns, the
// free unknowns, the materials in grid layout and the loads. Pure integer and FP64 host work,
// done once per solver.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace fem {

enum class Kind { HeatHex8 = 0, HeatTet4 = 1, ElasticTet4 = 2, ElasticTet10 = 3, ElasticHex8 = 4 };

// Kinds on the box cells themselves (hex8) rather than their 6 Kuhn tetrahedra.
inline bool hex_kind(Kind k) { return k == Kind::HeatHex8 || k == Kind::ElasticHex8; }

struct BoxInput {
    Kind kind = Kind::HeatHex8;
    int n[3] = {1, 1, 1};
    double len[3] = {1, 1, 1};
    std::vector<double> coef0, coef1;  // reference order (cells or tetrahedra)
    bool fixed[6] = {};
    double value[6] = {};
    double source = 0;
    double body[3] = {};
    double traction[6][3] = {};
};

struct BoxModel {
    BoxInput in;
    int comps = 1;
    bool quadratic = false;
    int gn[3] = {0, 0, 0};  // level-0 grid nodes per axis (n + 1, or 2 n + 1 for tet10)
    double h[3] = {0, 0, 0};
    uint32_t fix = 0;       // face mask of the fixed faces
    size_t points = 0;      // points of the reference numbering (= grid nodes of level 0)
    size_t free_count = 0;  // free unknowns (known from the shape alone)
    std::vector<uint32_t> point_grid;  // reference point -> level-0 grid node
    std::vector<int64_t> free;         // free unknowns, increasing (reference numbering)
    std::vector<int64_t> free_grid;    // free unknown k -> level-0 grid unknown (comps * node + c)
    // materials in grid layout, FP64
    std::vector<double> cell_k;        // hex8: k (heat) or Young's modulus (elasticity) per cell (x fastest)
    std::vector<double> cell_nu;       // hex8 elasticity: Poisson ratio per cell
    std::vector<double> tet_coef;      // Kuhn tets: [(q * 6 + type) * cells + cell], q = 0 (k or lambda), 1 (mu)
    std::vector<uint32_t> tet_cell;    // tetrahedron (reference order) -> cell (x fastest)
    std::vector<uint8_t> tet_type;     // tetrahedron (reference order) -> Kuhn type (chain axis order)
    // loads over all level-0 grid unknowns, and the fixed values
    std::vector<double> load;
    std::vector<double> fixed_value;   // over level-0 grid unknowns; nonzero only on fixed faces

    size_t cells() const { return (size_t)in.n[0] * in.n[1] * in.n[2]; }
    size_t grid_nodes() const { return (size_t)gn[0] * gn[1] * gn[2]; }
    size_t grid_unknowns() const { return grid_nodes() * comps; }
    size_t unknowns() const { return points * comps; }
};

// Builds the model. Throws fem::Error with a message on bad input.
BoxModel build_model(const BoxInput& in);

// The shape of the model alone (grid, cell sizes, fixed faces, the count of free unknowns), with
// the input checked, without the numbering: what admission by footprint needs, in O(1).
BoxModel box_shape(const BoxInput& in);

// The prism split of the reference (crates/fem/src/prism.rs:131-154): the 3 tetrahedra of a
// prism with the global indices g of its 6 corners, as local corners.
std::array<std::array<int, 4>, 3> prism_split(const std::array<uint64_t, 6>& g);

}  // namespace fem
