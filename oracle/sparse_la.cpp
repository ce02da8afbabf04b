// TEST INFRASTRUCTURE ONLY (the oracle). Restates crates/fem/src/sparse.rs, order.rs,
// symbolic.rs and cholesky.rs. Each function cites the lines it follows.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>

#include "oracle.hpp"

namespace orc {

// ------------------------------------------------------------------ sparse.rs:22-65
Csr Csr::pattern(usize n, const std::vector<usize>& eptr, const std::vector<u32>& conn) {
    const usize ne = eptr.size() - 1;
    std::vector<usize> count(n + 1, 0);
    for (usize k = 0; k < ne; ++k)
        for (usize t = eptr[k]; t < eptr[k + 1]; ++t) count[conn[t] + 1] += 1;
    for (usize i = 0; i < n; ++i) count[i + 1] += count[i];
    std::vector<u32> of_node(count[n]);
    std::vector<usize> next = count;
    for (usize k = 0; k < ne; ++k)
        for (usize t = eptr[k]; t < eptr[k + 1]; ++t) of_node[next[conn[t]]++] = (u32)k;
    Csr out;
    out.n = n;
    out.ptr.reserve(n + 1);
    out.ptr.push_back(0);
    std::vector<usize> mark(n, NONE);
    std::vector<u32> row;
    for (usize i = 0; i < n; ++i) {
        row.clear();
        mark[i] = i;
        row.push_back((u32)i);
        for (usize t = count[i]; t < count[i + 1]; ++t) {
            const usize k = of_node[t];
            for (usize s = eptr[k]; s < eptr[k + 1]; ++s) {
                const u32 b = conn[s];
                if (mark[b] != i) {
                    mark[b] = i;
                    row.push_back(b);
                }
            }
        }
        std::sort(row.begin(), row.end());
        out.idx.insert(out.idx.end(), row.begin(), row.end());
        out.ptr.push_back(out.idx.size());
    }
    out.val.assign(out.idx.size(), 0.0);
    return out;
}

// sparse.rs:68-71
usize Csr::find(usize i, usize j) const {
    auto b = idx.begin() + ptr[i], e = idx.begin() + ptr[i + 1];
    auto it = std::lower_bound(b, e, (u32)j);
    if (it != e && *it == (u32)j) return (usize)(it - idx.begin());
    return NONE;
}

// sparse.rs:75-86
void Csr::add_element(const u32* nodes, usize m, const double* ke) {
    for (usize a = 0; a < m; ++a) {
        const usize i = nodes[a];
        for (usize b = 0; b < m; ++b) {
            const usize k = find(i, nodes[b]);
            if (k == NONE) throw Error("the pattern holds each element pair");
            val[k] += ke[a * m + b];
        }
    }
}

// sparse.rs:89-93
std::vector<double> Csr::diagonal() const {
    std::vector<double> d(n);
    for (usize i = 0; i < n; ++i) {
        const usize k = find(i, i);
        d[i] = k == NONE ? 0.0 : val[k];
    }
    return d;
}

// sparse.rs:96-104
void Csr::mul(const double* x, double* y) const {
    for (usize i = 0; i < n; ++i) {
        double s = 0.0;
        for (usize k = ptr[i]; k < ptr[i + 1]; ++k) s += val[k] * x[idx[k]];
        y[i] = s;
    }
}

// sparse.rs:138-148
static void sort_row(u32* idx, double* val, usize len) {
    bool sorted = true;
    for (usize k = 1; k < len; ++k)
        if (!(idx[k - 1] < idx[k])) {
            sorted = false;
            break;
        }
    if (sorted) return;
    std::vector<std::pair<u32, double>> pairs(len);
    for (usize k = 0; k < len; ++k) pairs[k] = {idx[k], val[k]};
    std::sort(pairs.begin(), pairs.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    for (usize k = 0; k < len; ++k) {
        idx[k] = pairs[k].first;
        val[k] = pairs[k].second;
    }
}

// sparse.rs:108-135
Csr Csr::submatrix(const std::vector<usize>& keep) const {
    std::vector<u32> newi(n, UINT32_MAX);
    for (usize k = 0; k < keep.size(); ++k) newi[keep[k]] = (u32)k;
    Csr out;
    out.n = keep.size();
    out.ptr.reserve(keep.size() + 1);
    out.ptr.push_back(0);
    for (usize i : keep) {
        const usize start = out.idx.size();
        for (usize k = ptr[i]; k < ptr[i + 1]; ++k) {
            const u32 j = newi[idx[k]];
            if (j != UINT32_MAX) {
                out.idx.push_back(j);
                out.val.push_back(val[k]);
            }
        }
        sort_row(out.idx.data() + start, out.val.data() + start, out.idx.size() - start);
        out.ptr.push_back(out.idx.size());
    }
    return out;
}

// sparse.rs:212-220
void Rect::mul(const double* x, double* y) const {
    for (usize i = 0; i < rows; ++i) {
        double s = 0.0;
        for (usize k = ptr[i]; k < ptr[i + 1]; ++k) s += val[k] * x[idx[k]];
        y[i] = s;
    }
}

// sparse.rs:223-230
void Rect::mul_transpose(const double* x, double* y) const {
    for (usize c = 0; c < cols; ++c) y[c] = 0.0;
    for (usize i = 0; i < rows; ++i) {
        const double xi = x[i];
        for (usize k = ptr[i]; k < ptr[i + 1]; ++k) y[idx[k]] += val[k] * xi;
    }
}

// sparse.rs:233-292
Csr Rect::galerkin(const Csr& a) const {
    const usize nc = cols;
    std::vector<usize> tptr(nc + 1, 0);
    for (u32 j : idx) tptr[j + 1] += 1;
    for (usize j = 0; j < nc; ++j) tptr[j + 1] += tptr[j];
    std::vector<u32> tidx(idx.size());
    std::vector<double> tval(idx.size());
    std::vector<usize> next = tptr;
    for (usize i = 0; i < rows; ++i)
        for (usize k = ptr[i]; k < ptr[i + 1]; ++k) {
            const usize j = idx[k];
            tidx[next[j]] = (u32)i;
            tval[next[j]] = val[k];
            next[j] += 1;
        }
    std::vector<double> acc(nc, 0.0);
    std::vector<usize> mark(nc, NONE);
    Csr out;
    out.n = nc;
    out.ptr.reserve(nc + 1);
    out.ptr.push_back(0);
    std::vector<u32> row;
    for (usize c = 0; c < nc; ++c) {
        row.clear();
        for (usize t = tptr[c]; t < tptr[c + 1]; ++t) {
            const usize i = tidx[t];
            const double pic = tval[t];
            for (usize ka = a.ptr[i]; ka < a.ptr[i + 1]; ++ka) {
                const usize l = a.idx[ka];
                const double aval = a.val[ka] * pic;
                for (usize kp = ptr[l]; kp < ptr[l + 1]; ++kp) {
                    const usize d = idx[kp];
                    if (mark[d] != c) {
                        mark[d] = c;
                        acc[d] = 0.0;
                        row.push_back((u32)d);
                    }
                    acc[d] += aval * val[kp];
                }
            }
        }
        std::sort(row.begin(), row.end());
        for (u32 d : row) {
            out.idx.push_back(d);
            out.val.push_back(acc[d]);
        }
        out.ptr.push_back(out.idx.size());
    }
    return out;
}

// ------------------------------------------------------------------ order.rs
static const usize LEAF = 64;                                   // order.rs:15
static const double CUTS[5] = {0.3, 0.4, 0.5, 0.6, 0.7};        // order.rs:56

static bool touches_other(const Csr& a, usize v, const std::vector<std::uint8_t>& side, std::uint8_t other) {
    for (usize k = a.ptr[v]; k < a.ptr[v + 1]; ++k)
        if (side[a.idx[k]] == other) return true;
    return false;
}

// order.rs:143-179
static std::pair<usize, usize> separator(const Csr& a, const std::vector<usize>& set, const std::vector<Vec3>& xyz,
                                         std::vector<std::uint8_t>& side, usize axis, double cut) {
    usize nleft = 0;
    for (usize v : set) {
        if (xyz[v][axis] < cut) {
            nleft += 1;
            side[v] = 1;
        } else {
            side[v] = 2;
        }
    }
    usize sl = 0, sr = 0;
    for (usize v : set) {
        const std::uint8_t mine = side[v];
        const std::uint8_t other = 3 - mine;
        if (touches_other(a, v, side, other)) {
            if (mine == 1) sl += 1;
            else sr += 1;
        }
    }
    for (usize v : set) side[v] = 0;
    return {std::min(sl, sr), std::min(nleft, set.size() - nleft)};
}

// Rust's f64::total_cmp for the finite coordinates used here, with -0.0 before +0.0.
static bool total_less(double a, double b) {
    if (a == b) return std::signbit(a) && !std::signbit(b);
    return a < b;
}

struct Split {
    std::vector<usize> left, right, sep;
};

// order.rs:66-139
static std::optional<Split> split(const Csr& a, const std::vector<usize>& set, const std::vector<Vec3>& xyz,
                                  std::vector<std::uint8_t>& side) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (usize v : set)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::fmin(lo[d], xyz[v][d]);
            hi[d] = std::fmax(hi[d], xyz[v][d]);
        }
    bool have = false;
    double best_cost = 0, best_cut = 0;
    usize best_axis = 0;
    std::vector<double> values;
    for (usize axis = 0; axis < 3; ++axis) {
        if (hi[axis] <= lo[axis]) continue;
        values.clear();
        for (usize v : set) values.push_back(xyz[v][axis]);
        std::sort(values.begin(), values.end(), total_less);
        for (double q : CUTS) {
            // `(len as f64 * q) as usize`, truncation toward 0.
            usize at = (usize)((double)values.size() * q);
            double cut = values[std::min(at, values.size() - 1)];
            if (cut <= lo[axis]) {
                bool found = false;
                for (double x : values)
                    if (x > lo[axis]) {
                        cut = x;
                        found = true;
                        break;
                    }
                if (!found) continue;
            }
            auto [sep, small] = separator(a, set, xyz, side, axis, cut);
            const double cost = (double)sep / std::sqrt(2.0 * (double)small / (double)set.size());
            if (!have || cost < best_cost) {
                have = true;
                best_cost = cost;
                best_cut = cut;
                best_axis = axis;
            }
        }
    }
    if (!have) return std::nullopt;
    const double cut = best_cut;
    const usize axis = best_axis;
    std::vector<usize> left, right;
    for (usize v : set) (xyz[v][axis] < cut ? left : right).push_back(v);
    for (usize v : left) side[v] = 1;
    for (usize v : right) side[v] = 2;
    std::vector<usize> sep_left, sep_right;
    for (usize v : left)
        if (touches_other(a, v, side, 2)) sep_left.push_back(v);
    for (usize v : right)
        if (touches_other(a, v, side, 1)) sep_right.push_back(v);
    for (usize v : set) side[v] = 0;
    const bool from_left = sep_left.size() <= sep_right.size();
    std::vector<usize> sep = from_left ? sep_left : sep_right;
    std::vector<usize> in_sep = sep;
    std::sort(in_sep.begin(), in_sep.end());
    auto rest = [&](const std::vector<usize>& part) {
        std::vector<usize> out;
        for (usize v : part)
            if (!std::binary_search(in_sep.begin(), in_sep.end(), v)) out.push_back(v);
        return out;
    };
    Split s;
    if (from_left) {
        s.left = rest(left);
        s.right = std::move(right);
    } else {
        s.left = std::move(left);
        s.right = rest(right);
    }
    s.sep = std::move(sep);
    return s;
}

// order.rs:19-53
std::vector<usize> nested_dissection(const Csr& a, const std::vector<Vec3>& xyz) {
    if (a.n != xyz.size()) throw Error("1 point for each node");
    std::vector<usize> order;
    order.reserve(a.n);
    std::vector<std::uint8_t> side(a.n, 0);
    struct Work {
        bool emit;
        std::vector<usize> set;
    };
    std::vector<Work> stack;
    std::vector<usize> all(a.n);
    for (usize i = 0; i < a.n; ++i) all[i] = i;
    stack.push_back({false, std::move(all)});
    while (!stack.empty()) {
        Work w = std::move(stack.back());
        stack.pop_back();
        if (w.emit) {
            order.insert(order.end(), w.set.begin(), w.set.end());
            continue;
        }
        if (w.set.size() <= LEAF) {
            order.insert(order.end(), w.set.begin(), w.set.end());
            continue;
        }
        auto s = split(a, w.set, xyz, side);
        if (!s) {
            order.insert(order.end(), w.set.begin(), w.set.end());
            continue;
        }
        stack.push_back({true, std::move(s->sep)});
        stack.push_back({false, std::move(s->right)});
        stack.push_back({false, std::move(s->left)});
    }
    return order;
}

// ------------------------------------------------------------------ symbolic.rs
static std::vector<usize> inverse(const std::vector<usize>& p) {
    std::vector<usize> ip(p.size());
    for (usize k = 0; k < p.size(); ++k) ip[p[k]] = k;
    return ip;
}

// symbolic.rs:195-215
static std::vector<usize> etree(const Csr& a, const std::vector<usize>& order, const std::vector<usize>& iorder) {
    const usize n = a.n;
    std::vector<usize> parent(n, NONE), ancestor(n, NONE);
    for (usize k = 0; k < n; ++k) {
        const usize old = order[k];
        for (usize t = a.ptr[old]; t < a.ptr[old + 1]; ++t) {
            usize i = iorder[a.idx[t]];
            while (i != NONE && i < k) {
                const usize next = ancestor[i];
                ancestor[i] = k;
                if (next == NONE) parent[i] = k;
                i = next;
            }
        }
    }
    return parent;
}

// symbolic.rs:218-245
static std::vector<usize> postorder(const std::vector<usize>& parent) {
    const usize n = parent.size();
    std::vector<usize> head(n, NONE), next(n, NONE);
    for (usize j = n; j-- > 0;) {
        if (parent[j] != NONE) {
            next[j] = head[parent[j]];
            head[parent[j]] = j;
        }
    }
    std::vector<usize> post;
    post.reserve(n);
    std::vector<usize> stack;
    for (usize root = 0; root < n; ++root) {
        if (parent[root] != NONE) continue;
        stack.push_back(root);
        while (!stack.empty()) {
            const usize top = stack.back();
            const usize child = head[top];
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

// symbolic.rs:249-266
static std::vector<usize> column_counts(const Csr& a, const std::vector<usize>& perm, const std::vector<usize>& iperm,
                                        const std::vector<usize>& parent) {
    const usize n = a.n;
    std::vector<usize> counts(n, 1), mark(n, NONE);
    for (usize i = 0; i < n; ++i) {
        mark[i] = i;
        const usize old = perm[i];
        for (usize t = a.ptr[old]; t < a.ptr[old + 1]; ++t) {
            usize j = iperm[a.idx[t]];
            while (j < i && mark[j] != i) {
                counts[j] += 1;
                mark[j] = i;
                j = parent[j];
            }
        }
    }
    return counts;
}

// symbolic.rs:41-150
Symbolic Symbolic::analyse(const Csr& a, const std::vector<usize>& order) {
    const usize n = a.n;
    if (order.size() != n) throw Error("the order has each column once");
    const std::vector<usize> iorder = inverse(order);
    const std::vector<usize> parent0 = etree(a, order, iorder);
    const std::vector<usize> post = postorder(parent0);
    Symbolic s;
    s.n = n;
    s.perm.resize(n);
    for (usize k = 0; k < n; ++k) s.perm[k] = order[post[k]];
    s.iperm = inverse(s.perm);
    const std::vector<usize> ipost = inverse(post);
    std::vector<usize> parent(n);
    for (usize k = 0; k < n; ++k) parent[k] = parent0[post[k]] == NONE ? NONE : ipost[parent0[post[k]]];
    const std::vector<usize> counts = column_counts(a, s.perm, s.iperm, parent);
    std::vector<usize> nchild(n, 0);
    for (usize p : parent)
        if (p != NONE) nchild[p] += 1;
    s.start.push_back(0);
    for (usize j = 1; j < n; ++j) {
        const bool joins = parent[j - 1] == j && counts[j - 1] == counts[j] + 1 && nchild[j] == 1;
        if (!joins) s.start.push_back(j);
    }
    s.start.push_back(n);
    const usize ns = s.start.size() - 1;
    std::vector<usize> of_col(n, 0);
    for (usize sn = 0; sn < ns; ++sn)
        for (usize j = s.start[sn]; j < s.start[sn + 1]; ++j) of_col[j] = sn;
    std::vector<usize> sn_parent(ns);
    for (usize sn = 0; sn < ns; ++sn) {
        const usize p = parent[s.start[sn + 1] - 1];
        sn_parent[sn] = p == NONE ? NONE : of_col[p];
    }
    s.children.assign(ns, 0);
    std::vector<std::vector<usize>> kids(ns);
    for (usize sn = 0; sn < ns; ++sn)
        if (sn_parent[sn] != NONE) {
            s.children[sn_parent[sn]] += 1;
            kids[sn_parent[sn]].push_back(sn);
        }
    s.rows_ptr.push_back(0);
    std::vector<usize> mark(n, NONE);
    std::vector<u32> list;
    for (usize sn = 0; sn < ns; ++sn) {
        list.clear();
        const usize first = s.start[sn], end = s.start[sn + 1];
        for (usize j = first; j < end; ++j) {
            mark[j] = sn;
            list.push_back((u32)j);
        }
        for (usize j = first; j < end; ++j) {
            const usize old = s.perm[j];
            for (usize t = a.ptr[old]; t < a.ptr[old + 1]; ++t) {
                const usize i = s.iperm[a.idx[t]];
                if (i >= end && mark[i] != sn) {
                    mark[i] = sn;
                    list.push_back((u32)i);
                }
            }
        }
        for (usize c : kids[sn])
            for (usize t = s.rows_ptr[c]; t < s.rows_ptr[c + 1]; ++t) {
                const usize i = s.rows[t];
                if (i >= end && mark[i] != sn) {
                    mark[i] = sn;
                    list.push_back((u32)i);
                }
            }
        std::sort(list.begin(), list.end());
        s.rows.insert(s.rows.end(), list.begin(), list.end());
        s.rows_ptr.push_back(s.rows.size());
    }
    s.parent = std::move(sn_parent);
    return s;
}

// symbolic.rs:162-170
std::uint64_t Symbolic::entries() const {
    std::uint64_t total = 0;
    for (usize s = 0; s < supernodes(); ++s) {
        const std::uint64_t k = start[s + 1] - start[s];
        const std::uint64_t m = rows_ptr[s + 1] - rows_ptr[s];
        total += k * m - k * (k - 1) / 2;
    }
    return total;
}

// symbolic.rs:174-182
double Symbolic::flops() const {
    double total = 0.0;
    for (usize s = 0; s < supernodes(); ++s) {
        const usize k = start[s + 1] - start[s];
        const usize m = rows_ptr[s + 1] - rows_ptr[s];
        double part = 0.0;
        for (usize j = 0; j < k; ++j) {
            const double v = (double)(m - j);
            part += v * v;
        }
        total += part;
    }
    return total;
}

// symbolic.rs:185-191
std::uint64_t Symbolic::largest_front() const {
    std::uint64_t best = 0;
    for (usize s = 0; s < supernodes(); ++s) {
        const std::uint64_t m = rows_ptr[s + 1] - rows_ptr[s];
        best = std::max(best, m * m);
    }
    return best;
}

// ------------------------------------------------------------------ cholesky.rs
static const double PIVOT = 1e-10;  // cholesky.rs:22

// cholesky.rs:173-225. Returns the local column of a failed pivot, or NONE.
static usize partial_cholesky(double* f, usize m, usize k, const double* diag) {
    for (usize j = 0; j < k; ++j) {
        double* col = f + j * m;
        const double d = col[j];
        if (!(d > PIVOT * std::fabs(diag[j]))) return j;
        const double l = std::sqrt(d);
        col[j] = l;
        const double inv = 1.0 / l;
        for (usize r = j + 1; r < m; ++r) col[r] *= inv;
        for (usize c = j + 1; c < k; ++c) {
            double* target = f + c * m;
            const double factor = col[c];
            for (usize r = c; r < m; ++r) target[r] -= col[r] * factor;
        }
    }
    double* l = f;
    for (usize c = k; c < m; ++c) {
        double* target = f + c * m;
        usize p = 0;
        while (p + 4 <= k) {
            const double* c0 = l + p * m;
            const double* c1 = l + (p + 1) * m;
            const double* c2 = l + (p + 2) * m;
            const double* c3 = l + (p + 3) * m;
            const double f0 = c0[c], f1 = c1[c], f2 = c2[c], f3 = c3[c];
            for (usize r = c; r < m; ++r) target[r] -= c0[r] * f0 + c1[r] * f1 + c2[r] * f2 + c3[r] * f3;
            p += 4;
        }
        while (p < k) {
            const double* c0 = l + p * m;
            const double f0 = c0[c];
            for (usize r = c; r < m; ++r) target[r] -= c0[r] * f0;
            p += 1;
        }
    }
    return NONE;
}

// cholesky.rs:44-122
Cholesky Cholesky::factor(const Symbolic& sym, const Csr& a) {
    const usize n = sym.n;
    const usize ns = sym.supernodes();
    Cholesky out;
    out.offset.reserve(ns + 1);
    out.offset.push_back(0);
    for (usize s = 0; s < ns; ++s) {
        const usize k = sym.start[s + 1] - sym.start[s];
        const usize m = sym.rows_len(s);
        out.offset.push_back(out.offset[s] + k * m);
    }
    out.values.assign(out.offset[ns], 0.0);
    std::vector<usize> pos(n, NONE);
    struct Update {
        std::vector<u32> rows;
        std::vector<double> values;
    };
    std::vector<Update> stack;
    std::vector<double> front, diag;
    for (usize s = 0; s < ns; ++s) {
        const u32* rows = sym.rows_of(s);
        const usize first = sym.start[s], end = sym.start[s + 1];
        const usize m = sym.rows_len(s), k = end - first;
        for (usize l = 0; l < m; ++l) pos[rows[l]] = l;
        front.assign(m * m, 0.0);
        diag.assign(k, 0.0);
        for (usize j = first; j < end; ++j) {
            const usize old = sym.perm[j];
            const usize col = (j - first) * m;
            for (usize e = a.ptr[old]; e < a.ptr[old + 1]; ++e) {
                const usize i = sym.iperm[a.idx[e]];
                if (i >= j) front[col + pos[i]] += a.val[e];
                if (i == j) diag[j - first] = a.val[e];
            }
        }
        for (usize c = 0; c < sym.children[s]; ++c) {
            if (stack.empty()) throw Error("each child pushed its update");
            Update u = std::move(stack.back());
            stack.pop_back();
            const usize mu = u.rows.size();
            std::vector<usize> at(mu);
            for (usize t = 0; t < mu; ++t) at[t] = pos[u.rows[t]];
            for (usize b = 0; b < mu; ++b) {
                const usize col = at[b] * m;
                const double* src = u.values.data() + b * mu;
                for (usize aa = b; aa < mu; ++aa) front[col + at[aa]] += src[aa];
            }
        }
        const usize bad = partial_cholesky(front.data(), m, k, diag.data());
        if (bad != NONE) {
            const usize column = sym.perm[first + bad];
            char buf[256];
            // Rust's `{:e}` of an f64: the shortest round-trip digits in scientific notation.
            char num[64];
            std::snprintf(num, sizeof num, "%.17g", diag[bad]);
            double back = diag[bad];
            for (int prec = 0; prec <= 17; ++prec) {
                std::snprintf(num, sizeof num, "%.*e", prec, diag[bad]);
                if (std::strtod(num, nullptr) == back) break;
            }
            // Rust writes 1e0 / 1.5e-3 (no '+' and no leading zeros in the exponent).
            std::string t(num);
            const usize epos = t.find('e');
            if (epos != std::string::npos) {
                std::string mant = t.substr(0, epos);
                int ex = std::atoi(t.c_str() + epos + 1);
                t = mant + "e" + std::to_string(ex);
            }
            std::snprintf(buf, sizeof buf,
                          "the matrix is not positive definite: the pivot of column %zu is 0 or less to the "
                          "precision of the arithmetic (its diagonal is %s)",
                          column, t.c_str());
            throw Error(buf);
        }
        std::copy(front.begin(), front.begin() + k * m, out.values.begin() + out.offset[s]);
        if (m > k) {
            const usize mu = m - k;
            Update u;
            u.values.assign(mu * mu, 0.0);
            for (usize b = 0; b < mu; ++b) {
                const usize src = (k + b) * m + k;
                for (usize t = b; t < mu; ++t) u.values[b * mu + t] = front[src + t];
            }
            u.rows.assign(rows + k, rows + m);
            stack.push_back(std::move(u));
        }
        for (usize l = 0; l < m; ++l) pos[rows[l]] = NONE;
    }
    out.sym = sym;
    return out;
}

// cholesky.rs:125-160
void Cholesky::solve(double* b) const {
    const usize n = sym.n;
    std::vector<double> y(n);
    for (usize k = 0; k < n; ++k) y[k] = b[sym.perm[k]];
    const usize ns = sym.supernodes();
    for (usize s = 0; s < ns; ++s) {
        const u32* rows = sym.rows_of(s);
        const usize first = sym.start[s], end = sym.start[s + 1];
        const usize m = sym.rows_len(s);
        const double* block = values.data() + offset[s];
        for (usize j = 0; j < end - first; ++j) {
            const double* col = block + j * m;
            const double v = y[first + j] / col[j];
            y[first + j] = v;
            for (usize r = j + 1; r < m; ++r) y[rows[r]] -= col[r] * v;
        }
    }
    for (usize s = ns; s-- > 0;) {
        const u32* rows = sym.rows_of(s);
        const usize first = sym.start[s], end = sym.start[s + 1];
        const usize m = sym.rows_len(s);
        const double* block = values.data() + offset[s];
        for (usize j = end - first; j-- > 0;) {
            const double* col = block + j * m;
            double v = y[first + j];
            for (usize r = j + 1; r < m; ++r) v -= col[r] * y[rows[r]];
            y[first + j] = v / col[j];
        }
    }
    for (usize k = 0; k < n; ++k) b[sym.perm[k]] = y[k];
}

}  // namespace orc
