// Host-side (CPU) constraint matrix builder.
//
// Reference: build_constraint_matrix / reduce_constraints
// (tectosaur/fast_constraints.cpp:120-300).  Linear equality constraints
// sum_k val_k u_{dof_k} = rhs are eliminated one at a time, ordered by
// (number of terms, largest dof); each is first rewritten in terms of the
// constraints already eliminated, then solved for its largest dof.  After
// a final back-substitution pass every eliminated dof is a combination of
// free dofs plus a combination of constraint right-hand sides.  The result
// is the map u = cm x + rhs_mat c of tectosaur/constraints.py:10-17.
//
// The elimination order, term merging, zero filtering (|val| < 1e-13) and
// pivot choice are the reference's, so the same constraint list yields the
// same cm; the data structures are this file's own.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <numeric>
#include <vector>

#include "../../include/tct_hip.h"

namespace tct {
int set_error(int code, const char *fmt, ...);
}

namespace {

struct Term {
    double val;
    int64_t dof;
};
using Terms = std::vector<Term>;

struct Eq {        // sum terms = sum rhs (rhs terms index constraint numbers)
    Terms terms;
    Terms rhs;
};

struct Isolated {  // lhs_dof = sum eq.terms + sum eq.rhs
    int64_t lhs_dof;
    Eq eq;
};

using Reduced = std::map<int64_t, Isolated>;

Terms merge(const Terms &tv) {
    Terms out;
    for (const Term &t : tv) {
        bool found = false;
        for (Term &o : out)
            if (o.dof == t.dof) {
                o.val += t.val;
                found = true;
                break;
            }
        if (!found) out.push_back(t);
    }
    return out;
}

Terms drop_zeros(const Terms &tv) {
    Terms out;
    for (const Term &t : tv)
        if (!(std::fabs(t.val) < 1e-13)) out.push_back(t);
    return out;
}

Eq tidy(const Eq &e) { return Eq{drop_zeros(merge(e.terms)), drop_zeros(merge(e.rhs))}; }

// (largest dof, its index); the first index on ties, 0 for an empty list
std::pair<int64_t, size_t> largest_dof(const Terms &t) {
    int64_t best = 0;
    size_t idx = 0;
    for (size_t i = 0; i < t.size(); i++)
        if (t[i].dof > best) {
            best = t[i].dof;
            idx = i;
        }
    return {best, idx};
}

Isolated isolate(const Eq &e, size_t k) {
    const Term lhs = e.terms[k];
    Isolated out;
    out.lhs_dof = lhs.dof;
    for (size_t i = 0; i < e.terms.size(); i++)
        if (i != k) out.eq.terms.push_back({-e.terms[i].val / lhs.val, e.terms[i].dof});
    for (const Term &r : e.rhs) out.eq.rhs.push_back({r.val / lhs.val, r.dof});
    return out;
}

// Replace term k of e (whose dof is sub.lhs_dof) by sub's expression.
Eq substitute(const Eq &e, size_t k, const Isolated &sub, double rhs_factor) {
    const double f = e.terms[k].val;
    Eq out;
    for (size_t i = 0; i < e.terms.size(); i++)
        if (i != k) out.terms.push_back(e.terms[i]);
    for (const Term &t : sub.eq.terms) out.terms.push_back({t.val * f, t.dof});
    out.rhs = e.rhs;
    for (const Term &r : sub.eq.rhs) out.rhs.push_back({-rhs_factor * f * r.val, r.dof});
    return out;
}

Eq reduce(Eq e, const Reduced &m, double rhs_factor) {
    for (;;) {
        bool changed = false;
        for (size_t i = 0; i < e.terms.size(); i++) {
            auto it = m.find(e.terms[i].dof);
            if (it != m.end()) {
                e = tidy(substitute(e, i, it->second, rhs_factor));
                changed = true;
                break;
            }
        }
        if (!changed) return e;
    }
}

}  // namespace

// Constraints arrive flattened: for constraint i, n_terms[i] terms taken in
// order from term_dofs / term_vals, right-hand side rhs[i].
// Outputs (caller-allocated, capacities given; *_n receive the counts):
//   cm COO (cm_rows, cm_cols, cm_vals), rhs_mat COO (r_rows, r_cols, r_vals),
//   *n_eliminated = number of eliminated dofs (cm has n_total - n_eliminated
//   columns).  When a capacity is too small nothing is written and the
//   counts tell the caller the sizes to allocate.
extern "C" int tct_build_constraint_matrix(int64_t n_cs, const int64_t *n_terms,
                                           const int64_t *term_dofs, const double *term_vals,
                                           int64_t n_total_dofs, int64_t cap_cm,
                                           int64_t *cm_rows, int64_t *cm_cols, double *cm_vals,
                                           int64_t *cm_n, int64_t cap_r, int64_t *r_rows,
                                           int64_t *r_cols, double *r_vals, int64_t *r_n,
                                           int64_t *n_eliminated) {
    if (n_cs < 0 || n_total_dofs < 0 || !cm_n || !r_n || !n_eliminated ||
        (n_cs > 0 && (!n_terms || !term_dofs || !term_vals)))
        return tct::set_error(TCT_EARG, "build_constraint_matrix: bad arguments");
    std::vector<Eq> cs(n_cs);
    int64_t off = 0;
    for (int64_t i = 0; i < n_cs; i++) {
        for (int64_t k = 0; k < n_terms[i]; k++) {
            const int64_t d = term_dofs[off + k];
            if (d < 0 || d >= n_total_dofs)
                return tct::set_error(TCT_EARG, "constraint %lld: dof %lld out of range",
                                      (long long)i, (long long)d);
            cs[i].terms.push_back({term_vals[off + k], d});
        }
        off += n_terms[i];
        cs[i].rhs = {Term{1.0, i}};
    }
    std::vector<size_t> order(n_cs);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        if (cs[a].terms.size() == cs[b].terms.size())
            return largest_dof(cs[a].terms).first < largest_dof(cs[b].terms).first;
        return cs[a].terms.size() < cs[b].terms.size();
    });
    Reduced lower;
    for (size_t oi = 0; oi < order.size(); oi++) {
        Eq e = reduce(tidy(cs[order[oi]]), lower, 1.0);
        if (e.terms.empty()) continue;
        Isolated iso = isolate(e, largest_dof(e.terms).second);
        lower[iso.lhs_dof] = iso;
    }
    for (auto &kv : lower) kv.second.eq = reduce(kv.second.eq, lower, -1.0);

    std::vector<int64_t> new_dof(n_total_dofs, -1);
    int64_t next = 0, ncm = 0, nr = 0;
    for (int64_t i = 0; i < n_total_dofs; i++) {
        auto it = lower.find(i);
        if (it == lower.end()) {
            new_dof[i] = next++;
            ncm++;
        } else {
            ncm += (int64_t)it->second.eq.terms.size();
            nr += (int64_t)it->second.eq.rhs.size();
        }
    }
    *cm_n = ncm;
    *r_n = nr;
    *n_eliminated = (int64_t)lower.size();
    if (cap_cm < ncm || cap_r < nr || !cm_rows || !cm_cols || !cm_vals ||
        (nr > 0 && (!r_rows || !r_cols || !r_vals)))
        return TCT_OK;
    int64_t a = 0, b = 0;
    for (int64_t i = 0; i < n_total_dofs; i++) {
        auto it = lower.find(i);
        if (it == lower.end()) {
            cm_rows[a] = i;
            cm_cols[a] = new_dof[i];
            cm_vals[a] = 1.0;
            a++;
            continue;
        }
        for (const Term &t : it->second.eq.terms) {
            if (new_dof[t.dof] < 0)
                return tct::set_error(TCT_EHIP, "constraint reduction left dof %lld dependent",
                                      (long long)t.dof);
            cm_rows[a] = i;
            cm_cols[a] = new_dof[t.dof];
            cm_vals[a] = t.val;
            a++;
        }
        for (const Term &t : it->second.eq.rhs) {
            r_rows[b] = i;
            r_cols[b] = t.dof;
            r_vals[b] = t.val;
            b++;
        }
    }
    return TCT_OK;
}
