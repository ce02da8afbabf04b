// Nested dissection and the counts of the symbolic analysis (symbolic.hpp). The steps follow
// crates/fem/src/order.rs and symbolic.rs so that the order, and hence the entry and flop counts
// the crossover compares, are the reference's exactly.
#include "symbolic.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

#include "common.hpp"

namespace fem {
namespace sym {

namespace {

constexpr uint64_t NONE = std::numeric_limits<uint64_t>::max();
constexpr size_t LEAF = 64;                               // order.rs:15
constexpr double CUTS[5] = {0.3, 0.4, 0.5, 0.6, 0.7};    // order.rs:56

// f64::total_cmp (IEEE total order: -0.0 before +0.0), the sort of order.rs:80.
bool total_less(double x, double y) {
    int64_t a, b;
    std::memcpy(&a, &x, 8);
    std::memcpy(&b, &y, 8);
    a ^= (int64_t)((uint64_t)(a >> 63) >> 1);
    b ^= (int64_t)((uint64_t)(b >> 63) >> 1);
    return a < b;
}

bool touches(const csr::Csr& a, uint64_t v, const std::vector<uint8_t>& side, uint8_t other) {
    for (uint64_t k = a.ptr[v]; k < a.ptr[v + 1]; ++k)
        if (side[a.idx[k]] == other) return true;
    return false;
}

// order.rs:143-179: the separator size of a trial cut and the size of the smaller side.
void separator(const csr::Csr& a, const std::vector<uint64_t>& set, const std::vector<std::array<double, 3>>& xyz,
               std::vector<uint8_t>& side, int axis, double cut, size_t& sep, size_t& small) {
    size_t nleft = 0;
    for (uint64_t v : set) {
        if (xyz[v][axis] < cut) {
            side[v] = 1;
            ++nleft;
        } else {
            side[v] = 2;
        }
    }
    size_t sl = 0, sr = 0;
    for (uint64_t v : set) {
        const uint8_t mine = side[v];
        if (touches(a, v, side, (uint8_t)(3 - mine))) {
            if (mine == 1) ++sl;
            else ++sr;
        }
    }
    for (uint64_t v : set) side[v] = 0;
    sep = std::min(sl, sr);
    small = std::min(nleft, set.size() - nleft);
}

// order.rs:66-139. False when all the nodes are at one point.
bool split(const csr::Csr& a, const std::vector<uint64_t>& set, const std::vector<std::array<double, 3>>& xyz,
           std::vector<uint8_t>& side, std::vector<uint64_t>& left, std::vector<uint64_t>& right,
           std::vector<uint64_t>& sep) {
    double lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = std::numeric_limits<double>::infinity();
        hi[d] = -std::numeric_limits<double>::infinity();
    }
    for (uint64_t v : set)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::fmin(lo[d], xyz[v][d]);
            hi[d] = std::fmax(hi[d], xyz[v][d]);
        }
    bool have = false;
    double best_cost = 0.0, best_cut = 0.0;
    int best_axis = 0;
    std::vector<double> values;
    for (int axis = 0; axis < 3; ++axis) {
        if (hi[axis] <= lo[axis]) continue;
        values.clear();
        for (uint64_t v : set) values.push_back(xyz[v][axis]);
        std::sort(values.begin(), values.end(), total_less);
        for (double q : CUTS) {
            size_t at = (size_t)((double)values.size() * q);
            if (at > values.size() - 1) at = values.size() - 1;
            double cut = values[at];
            if (cut <= lo[axis]) {
                // more than the fraction q of the nodes lie on the lowest plane: cut above it
                bool found = false;
                for (double x : values)
                    if (x > lo[axis]) {
                        cut = x;
                        found = true;
                        break;
                    }
                if (!found) continue;
            }
            size_t s = 0, small = 0;
            separator(a, set, xyz, side, axis, cut, s, small);
            const double cost = (double)s / std::sqrt(2.0 * (double)small / (double)set.size());
            if (!have || cost < best_cost) {
                have = true;
                best_cost = cost;
                best_cut = cut;
                best_axis = axis;
            }
        }
    }
    if (!have) return false;
    left.clear();
    right.clear();
    for (uint64_t v : set) (xyz[v][best_axis] < best_cut ? left : right).push_back(v);
    for (uint64_t v : left) side[v] = 1;
    for (uint64_t v : right) side[v] = 2;
    std::vector<uint64_t> sep_left, sep_right;
    for (uint64_t v : left)
        if (touches(a, v, side, 2)) sep_left.push_back(v);
    for (uint64_t v : right)
        if (touches(a, v, side, 1)) sep_right.push_back(v);
    for (uint64_t v : set) side[v] = 0;
    const bool from_left = sep_left.size() <= sep_right.size();
    sep = from_left ? std::move(sep_left) : std::move(sep_right);
    std::vector<uint64_t> in_sep = sep;
    std::sort(in_sep.begin(), in_sep.end());
    auto rest = [&](std::vector<uint64_t>& part) {
        std::vector<uint64_t> out;
        out.reserve(part.size());
        for (uint64_t v : part)
            if (!std::binary_search(in_sep.begin(), in_sep.end(), v)) out.push_back(v);
        part = std::move(out);
    };
    if (from_left) rest(left);
    else rest(right);
    return true;
}

std::vector<uint64_t> inverse(const std::vector<uint64_t>& p) {
    std::vector<uint64_t> ip(p.size());
    for (uint64_t k = 0; k < p.size(); ++k) ip[p[k]] = k;
    return ip;
}

// symbolic.rs:195-215: the elimination tree of P A P^T with path compression.
std::vector<uint64_t> etree(const csr::Csr& a, const std::vector<uint64_t>& order, const std::vector<uint64_t>& iorder) {
    const uint64_t n = a.n;
    std::vector<uint64_t> parent(n, NONE), ancestor(n, NONE);
    for (uint64_t k = 0; k < n; ++k) {
        const uint64_t old = order[k];
        for (uint64_t e = a.ptr[old]; e < a.ptr[old + 1]; ++e) {
            uint64_t i = iorder[a.idx[e]];
            while (i != NONE && i < k) {
                const uint64_t next = ancestor[i];
                ancestor[i] = k;
                if (next == NONE) parent[i] = k;
                i = next;
            }
        }
    }
    return parent;
}

// symbolic.rs:218-245: a postorder, children visited in increasing order, roots in increasing order.
std::vector<uint64_t> postorder(const std::vector<uint64_t>& parent) {
    const uint64_t n = parent.size();
    std::vector<uint64_t> head(n, NONE), next(n, NONE);
    for (uint64_t j = n; j-- > 0;)
        if (parent[j] != NONE) {
            next[j] = head[parent[j]];
            head[parent[j]] = j;
        }
    std::vector<uint64_t> post;
    post.reserve(n);
    std::vector<uint64_t> stack;
    for (uint64_t root = 0; root < n; ++root) {
        if (parent[root] != NONE) continue;
        stack.push_back(root);
        while (!stack.empty()) {
            const uint64_t top = stack.back();
            const uint64_t child = head[top];
            if (child == NONE) {
                stack.pop_back();
                post.push_back(top);
            } else {
                head[top] = next[child];
                stack.push_back(child);
            }
        }
    }
    return post;
}

// symbolic.rs:249-266: the entries of each column of L (diagonal included) from the row subtrees.
std::vector<uint64_t> column_counts(const csr::Csr& a, const std::vector<uint64_t>& perm,
                                    const std::vector<uint64_t>& iperm, const std::vector<uint64_t>& parent) {
    const uint64_t n = a.n;
    std::vector<uint64_t> counts(n, 1), mark(n, NONE);
    for (uint64_t i = 0; i < n; ++i) {
        mark[i] = i;
        const uint64_t old = perm[i];
        for (uint64_t e = a.ptr[old]; e < a.ptr[old + 1]; ++e) {
            uint64_t j = iperm[a.idx[e]];
            while (j < i && mark[j] != i) {
                counts[j] += 1;
                mark[j] = i;
                j = parent[j];
            }
        }
    }
    return counts;
}

}  // namespace

std::vector<uint64_t> nested_dissection(const csr::Csr& a, const std::vector<std::array<double, 3>>& xyz) {
    const uint64_t n = a.n;
    if (xyz.size() != n) throw Error("nested dissection needs 1 point for each unknown");
    std::vector<uint64_t> order;
    order.reserve(n);
    std::vector<uint8_t> side(n, 0);
    struct Work {
        bool emit;
        std::vector<uint64_t> set;
    };
    std::vector<Work> stack;
    {
        std::vector<uint64_t> all(n);
        for (uint64_t k = 0; k < n; ++k) all[k] = k;
        stack.push_back({false, std::move(all)});
    }
    std::vector<uint64_t> left, right, sep;
    while (!stack.empty()) {
        Work w = std::move(stack.back());
        stack.pop_back();
        if (w.emit || w.set.size() <= LEAF) {
            order.insert(order.end(), w.set.begin(), w.set.end());
            continue;
        }
        if (!split(a, w.set, xyz, side, left, right, sep)) {
            order.insert(order.end(), w.set.begin(), w.set.end());
            continue;
        }
        // the stack runs the last push first: left, then right, then the separator
        stack.push_back({true, std::move(sep)});
        stack.push_back({false, std::move(right)});
        stack.push_back({false, std::move(left)});
        left = {};
        right = {};
        sep = {};
    }
    return order;
}

Counts analyse(const csr::Csr& a, const std::vector<uint64_t>& order) {
    const uint64_t n = a.n;
    if (order.size() != n) throw Error("the order must hold each column once");
    const std::vector<uint64_t> iorder = inverse(order);
    const std::vector<uint64_t> parent0 = etree(a, order, iorder);
    const std::vector<uint64_t> post = postorder(parent0);
    std::vector<uint64_t> perm(n);
    for (uint64_t k = 0; k < n; ++k) perm[k] = order[post[k]];
    const std::vector<uint64_t> iperm = inverse(perm), ipost = inverse(post);
    std::vector<uint64_t> parent(n);
    for (uint64_t k = 0; k < n; ++k) parent[k] = parent0[post[k]] == NONE ? NONE : ipost[parent0[post[k]]];
    const std::vector<uint64_t> counts = column_counts(a, perm, iperm, parent);
    std::vector<uint64_t> nchild(n, 0);
    for (uint64_t p : parent)
        if (p != NONE) nchild[p] += 1;
    // fundamental supernodes (symbolic.rs:73-80); within one, column first + j has m - j entries,
    // m the count of its first column (the length of its row list)
    Counts out;
    uint64_t first = 0;
    auto close = [&](uint64_t end) {
        const uint64_t k = end - first, m = counts[first];
        out.entries += k * m - k * (k - 1) / 2;
        double f = 0.0;
        for (uint64_t j = 0; j < k; ++j) {
            const double r = (double)(m - j);
            f += r * r;
        }
        out.flops += f;
        out.largest_front = std::max(out.largest_front, m * m);
        out.supernodes += 1;
        first = end;
    };
    for (uint64_t j = 1; j < n; ++j) {
        const bool joins = parent[j - 1] == j && counts[j - 1] == counts[j] + 1 && nchild[j] == 1;
        if (!joins) close(j);
    }
    if (n > 0) close(n);
    return out;
}

Supernodal supernodal(const csr::Csr& a, const std::vector<uint64_t>& order) {
    const uint64_t n = a.n;
    if (order.size() != n) throw Error("the order must hold each column once");
    Supernodal S;
    S.n = n;
    S.counts = analyse(a, order);
    const std::vector<uint64_t> iorder = inverse(order);
    const std::vector<uint64_t> parent0 = etree(a, order, iorder);
    const std::vector<uint64_t> post = postorder(parent0);
    S.perm.resize(n);
    for (uint64_t k = 0; k < n; ++k) S.perm[k] = order[post[k]];
    S.iperm = inverse(S.perm);
    const std::vector<uint64_t> ipost = inverse(post);
    std::vector<uint64_t> parent(n);
    for (uint64_t k = 0; k < n; ++k) parent[k] = parent0[post[k]] == NONE ? NONE : ipost[parent0[post[k]]];
    const std::vector<uint64_t> counts = column_counts(a, S.perm, S.iperm, parent);
    std::vector<uint64_t> nchild(n, 0);
    for (uint64_t p : parent)
        if (p != NONE) nchild[p] += 1;
    // the fundamental supernodes of analyse (symbolic.rs:73-80)
    S.start.push_back(0);
    for (uint64_t j = 1; j < n; ++j) {
        const bool joins = parent[j - 1] == j && counts[j - 1] == counts[j] + 1 && nchild[j] == 1;
        if (!joins) S.start.push_back(j);
    }
    if (n > 0) S.start.push_back(n);
    const uint64_t ns = S.supernodes();
    std::vector<uint64_t> sup_of(n);
    for (uint64_t s = 0; s < ns; ++s)
        for (uint64_t j = S.start[s]; j < S.start[s + 1]; ++j) sup_of[j] = s;
    S.sparent.assign(ns, -1);
    std::vector<std::vector<uint64_t>> kids(ns);
    for (uint64_t s = 0; s < ns; ++s) {
        const uint64_t p = parent[S.start[s + 1] - 1];
        if (p != NONE) {
            S.sparent[s] = (int64_t)sup_of[p];
            kids[sup_of[p]].push_back(s);
        }
    }
    // the rows of each supernode: its columns' entries of A at or below its first column, and its
    // children's rows past their own columns (all at or below its first column)
    S.rptr.assign(ns + 1, 0);
    std::vector<uint64_t> mark(n, NONE), list;
    for (uint64_t s = 0; s < ns; ++s) {
        const uint64_t first = S.start[s], end = S.start[s + 1];
        list.clear();
        for (uint64_t j = first; j < end; ++j) {
            mark[j] = s;
            list.push_back(j);
        }
        for (uint64_t j = first; j < end; ++j) {
            const uint64_t old = S.perm[j];
            for (uint64_t e = a.ptr[old]; e < a.ptr[old + 1]; ++e) {
                const uint64_t i = S.iperm[a.idx[e]];
                if (i >= first && mark[i] != s) {
                    mark[i] = s;
                    list.push_back(i);
                }
            }
        }
        for (uint64_t c : kids[s]) {
            const uint64_t kc = S.start[c + 1] - S.start[c];
            for (uint64_t t = S.rptr[c] + kc; t < S.rptr[c + 1]; ++t) {
                const uint64_t i = S.rows[t];
                if (mark[i] != s) {
                    mark[i] = s;
                    list.push_back(i);
                }
            }
        }
        std::sort(list.begin(), list.end());
        if (list.size() != counts[first]) throw Error("internal: a supernode's rows differ from its column count");
        S.rows.insert(S.rows.end(), list.begin(), list.end());
        S.rptr[s + 1] = S.rows.size();
    }
    return S;
}

}  // namespace sym
// Amalgamation of small supernodes: a supernode whose columns end where its parent's begin (its
// parent's last child in the post-order) joins its parent when the two hold at most `kmax` pivots
// together. The joined front's rows are the pivots of both, then the parent's rows past its pivots
// (the child's rows past its pivots are among them: its update goes to its parent), so the child's
// columns only gain entries that stay exactly 0 in the factor, and the order is unchanged. Chains
// of 1- and 2-pivot supernodes of the separators then take one level of the tree, not one each.
sym::Supernodal sym::amalgamate(const sym::Supernodal& S, uint64_t kmax) {
    const uint64_t ns = S.supernodes();
    std::vector<uint64_t> first(S.start.begin(), S.start.end() - 1);  // first column, moves down on a join
    std::vector<int64_t> into(ns, -1);                                 // the supernode a dead one joined
    auto live = [&](int64_t q) {
        while (q >= 0 && into[(size_t)q] >= 0) q = into[(size_t)q];
        return q;
    };
    for (uint64_t q = 0; q < ns; ++q) {
        const int64_t p = S.sparent[q];
        if (p < 0) continue;
        const uint64_t kq = S.start[q + 1] - first[q], kp = S.start[(size_t)p + 1] - first[(size_t)p];
        if (S.start[q + 1] == first[(size_t)p] && kq + kp <= kmax) {
            first[(size_t)p] = first[q];
            into[q] = p;
        }
    }
    sym::Supernodal R;
    R.n = S.n;
    R.perm = S.perm;
    R.iperm = S.iperm;
    R.counts = S.counts;
    std::vector<int64_t> index(ns, -1);
    for (uint64_t q = 0; q < ns; ++q)
        if (into[q] < 0) index[q] = (int64_t)R.start.size(), R.start.push_back(first[q]);
    R.start.push_back(S.n);
    R.rptr.push_back(0);
    for (uint64_t q = 0; q < ns; ++q) {
        if (into[q] >= 0) continue;
        const uint64_t kq = S.start[q + 1] - S.start[q];
        for (uint64_t c = first[q]; c < S.start[q + 1]; ++c) R.rows.push_back(c);
        R.rows.insert(R.rows.end(), S.rows.begin() + (ptrdiff_t)(S.rptr[q] + kq), S.rows.begin() + (ptrdiff_t)S.rptr[q + 1]);
        R.rptr.push_back(R.rows.size());
        const int64_t p = live(S.sparent[q]);
        R.sparent.push_back(p < 0 ? -1 : index[(size_t)p]);
    }
    return R;
}

}  // namespace fem
