// The fill-reducing order and the factor counts that the reference's crossover reads
// (crates/fem/src/solve.rs:106-118, 158-170: Crossover::Flops), on the host: nested dissection
// with coordinate bisection (order.rs:19-179) and the counts of the symbolic analysis
// (symbolic.rs:41-192: elimination tree, postorder, column counts, fundamental supernodes).
// Integer work; the results must equal the reference's exactly, so the steps are the reference's.
#pragma once

#include <array>
#include <cstdint>
#include <vector>

#include "csr_host.hpp"

namespace fem {
namespace sym {

// order.rs:19-53: item k is the old index of the new node k. `xyz` holds one point per node.
std::vector<uint64_t> nested_dissection(const csr::Csr& a, const std::vector<std::array<double, 3>>& xyz);

// What Symbolic::analyse (symbolic.rs:41-150) gives the crossover: entries() (symbolic.rs:162-172),
// flops() (174-183), largest_front() (186-192), and the supernode count.
struct Counts {
    uint64_t entries = 0;
    double flops = 0.0;
    uint64_t largest_front = 0;
    uint64_t supernodes = 0;
};
Counts analyse(const csr::Csr& a, const std::vector<uint64_t>& order);

// The structure of the supernodal factor that the multifrontal factor fills (symbolic.rs:41-150):
// the fill-reducing order with its elimination-tree postorder (perm: new -> old), the fundamental
// supernodes (columns start[s] .. start[s + 1] - 1 in the new order), the rows of each (rows[rptr[s]
// .. rptr[s + 1]), new indices, increasing; the first start[s + 1] - start[s] are its own columns),
// the supernode tree (sparent, -1 at a root; a supernode's number is above its children's), and
// the counts of `analyse`.
struct Supernodal {
    uint64_t n = 0;
    std::vector<uint64_t> perm, iperm;
    std::vector<uint64_t> start;
    std::vector<uint64_t> rptr, rows;
    std::vector<int64_t> sparent;
    Counts counts;
    uint64_t supernodes() const { return start.empty() ? 0 : start.size() - 1; }
};
Supernodal supernodal(const csr::Csr& a, const std::vector<uint64_t>& order);
// The multifrontal factor's joins: a supernode whose columns end where its parent's begin (its
// parent's last child) joins its parent when the two hold at most kmax pivots; its rows become
// both pivot ranges and the parent's rows past its pivots. Same order, same counts; kmax 0 (or 1)
// joins nothing.
Supernodal amalgamate(const Supernodal& s, uint64_t kmax);

// Crossover::Flops (solve.rs:113-117): direct when the factor has at most `entries` entries and
// (`flat`, or its flops are at most `flops` per unknown).
inline bool flops_picks_direct(const Counts& c, uint64_t n, double flops, uint64_t entries, bool flat) {
    return c.entries <= entries && (flat || c.flops <= flops * (double)n);
}

}  // namespace sym
}  // namespace fem
