// CPU oracle for the t-J ladder exact-diagonalisation hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, step by step and in plain
// host C++, the reference package's CPU algorithms so that the HIP product
// path can be checked against it.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load the library built from it, and only as
// the reference results to compare with (or as the timed CPU baseline), never as
// the thing that ships.
//
// Every function names the reference file:line it follows (paths relative to
// the reference package root src/exactdiag/).  Arithmetic is kept in the
// reference's order (naive complex products, sequential sums, first-seen
// de-duplication) and the file must be compiled with -ffp-contract=off so
// that no product is fused into an FMA.
//
// Parity pinning (see tests/):
//   * basis sizes: 68 golden files + the two sum rules of the reference tests;
//   * orbit / representative / amplitude: the reference's own
//     symmetry_ladder.h compiled into oracle/_ref (bit-exact comparison), and
//     brute-force permutation action on tiny ladders;
//   * column matrix elements: brute-force Fock-space Hamiltonian projected
//     on the symmetric basis, and the 51 golden eigenvalue files;
//   * sparse dot: dense product of toarray();
//   * continued fraction: exact Lehmann sums on small sectors.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <memory>
#include <numbers>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using cplx = std::complex<double>;
using state_t = short;
constexpr state_t HOLE = 0, DOWN = 1, UP = 2;

thread_local std::string g_err;

int py_mod(int a, int b) {  // symmetry_utils.h:15-22
    int r = a % b;
    r += b * (r < 0);
    return r;
}

// Binomial table, general/matrix_rules_utils.pyx:122-134: comb[j][i] = C(i, j)
// built with the reference's floating running product and round-half-even.
std::vector<int64_t> make_comb(int size) {
    int w = size + 1;
    std::vector<int64_t> comb(size_t(w) * w, 0);
    for (int i = 0; i <= size; ++i) {
        double t = 1.0;
        for (int j = 0; j <= i / 2; ++j) {
            if (i != j && i * j != 0) t *= double(i + 1 - j) / double(j);
            int64_t v = int64_t(std::nearbyint(t));
            comb[size_t(j) * w + i] = v;
            comb[size_t(i - j) * w + i] = v;
        }
    }
    return comb;
}

struct Image {
    std::vector<state_t> st;
    cplx amp;
};

// State ordering, general/group/commutation_counter.h:45-69 with the
// hierarchy {HOLE, DOWN, UP}: 0 equal, 1 first lower, 2 second lower.
int compare_states(const state_t* a, const state_t* b, int n) {
    const state_t order[2] = {HOLE, DOWN};
    for (state_t pref : order) {
        for (int i = n - 1; i >= 0; --i) {
            bool bp = b[i] == pref, ap = a[i] == pref;
            if (bp && !ap) return 1;
            if (!bp && ap) return 2;
        }
    }
    return 0;
}

struct Ladder {
    int L, N, kx, ky;
    cplx rung_exp, mirror_exp;

    Ladder(int L_, int kx_, int ky_) : L(L_), N(2 * L_), kx(kx_), ky(ky_) {
        // tJ_spin_half_ladder/symmetry_ladder.h:134-138
        const cplx i2pi(0.0, 2.0 * std::numbers::pi);
        mirror_exp = std::exp(-i2pi * cplx(double(ky) / double(2)));
        rung_exp = std::exp(-i2pi * cplx(double(kx) / double(L)));
    }

    // tJ_spin_half_ladder/symmetry_ladder.h:113-235.  Returns the orbit size
    // (0 = not the lowest state when check_lowest, or incompatible with the
    // sector's characters).  images[0] is the input state.
    int orbit(const state_t* s, bool check_lowest, std::vector<Image>& images) const {
        images.clear();
        bool ne_even = true;
        int paired = 0;
        for (int i = 0; i < N; ++i) {
            if (i % 2 == 0 && s[i] != HOLE && s[i + 1] != HOLE) ++paired;
            ne_even ^= (s[i] != HOLE);
        }
        const cplx mirror_fac = mirror_exp * cplx(paired % 2 == 0 ? 1 : -1);
        cplx trans_amp = 1;
        bool mirror_sym = false;
        std::vector<state_t> rep(2 * N), mrep(2 * N);
        for (int i = 0; i < N; ++i) rep[i] = rep[i + N] = s[i];
        for (int i = 0; i < N; i += 2) {
            mrep[i] = s[i + 1];
            mrep[i + 1] = s[i];
        }
        for (int i = 0; i < N; ++i) mrep[i + N] = mrep[i];
        images.push_back({std::vector<state_t>(s, s + N), cplx(1)});
        const double eps = 1e-8;
        for (int p = 1; p < N; ++p) {
            const int m = (p / 2) % L;
            const bool mirrored = (p % 2) == 1;
            const state_t* t;
            cplx amp;
            if (mirrored) {
                if (mirror_sym) continue;
                t = &mrep[N - 2 * m];
                amp = trans_amp * mirror_fac;
            } else {
                t = &rep[N - 2 * m];
                trans_amp *= rung_exp;
                if (ne_even && ((t[0] == HOLE) ^ (t[1] == HOLE))) trans_amp *= -1;
                amp = trans_amp;
            }
            if (check_lowest && compare_states(s, t, N) == 2) {
                images.clear();
                return 0;
            }
            if (m < L / 2 + 1) {
                if (std::equal(t, t + N, images[0].st.begin())) {
                    if (std::abs(amp - cplx(1, 0)) > eps) {
                        images.clear();
                        return 0;
                    }
                    if (p == 1) {
                        mirror_sym = true;
                        continue;
                    }
                    break;
                }
                if (images.size() != 1 && std::equal(t, t + N, images[1].st.begin())) {
                    if (std::abs(amp - images[1].amp) > eps) {
                        images.clear();
                        return 0;
                    }
                    break;
                }
            }
            images.push_back({std::vector<state_t>(t, t + N), amp});
        }
        const int n = int(images.size());
        for (auto& im : images) im.amp /= std::sqrt(n);
        return n;
    }
};

struct Sector {
    int L, N, h, ne, nd, kx, ky;
    int64_t n_spin, n_hole;
    std::vector<int64_t> comb;
    Ladder sym;
    std::vector<int64_t> dense_to_sparse;
    std::vector<int32_t> orbit_size;

    int64_t C(int j, int i) const { return comb[size_t(j) * (N + 1) + i]; }

    Sector(int L_, int h_, int nd_, int kx_, int ky_)
        : L(L_), N(2 * L_), h(h_), ne(2 * L_ - h_), nd(nd_), kx(kx_), ky(ky_), sym(L_, kx_, ky_) {
        comb = make_comb(N);
        n_spin = C(nd, ne);  // tJ_spin_half_ladder/configs.py:100-101
        n_hole = C(h, N);    // tJ_spin_half_ladder/matrix_setup.pyx:402
    }

    // tJ_spin_half_ladder/symmetry_ladder.h:314-344
    void unrank(int64_t sparse, state_t* st) const {
        int rh = h, rs = nd;
        int64_t hi = sparse / n_spin, si = sparse - hi * n_spin;
        for (int j = N - 1; j >= 0; --j) {
            if (rh > 0) {
                int64_t c = C(rh, j);
                if (hi >= c) {
                    st[j] = HOLE;
                    --rh;
                    hi -= c;
                    continue;
                }
            }
            if (rs > 0) {
                int64_t c = C(rs, j - rh);
                if (si >= c) {
                    st[j] = DOWN;
                    --rs;
                    si -= c;
                    continue;
                }
            }
            st[j] = UP;
        }
    }

    // tJ_spin_half_ladder/symmetry_ladder.h:346-361
    int64_t rank(const state_t* st) const {
        int64_t hi = 0, si = 0;
        int ah = st[0] == HOLE, ad = st[0] == DOWN;
        for (int i = 1; i < N; ++i) {
            if (st[i] == HOLE) {
                ++ah;
                hi += C(ah, i);
            } else if (st[i] == DOWN) {
                ++ad;
                si += C(ad, i - ah);
            }
        }
        return si + hi * n_spin;
    }

    // tJ_spin_half_ladder/symmetry_ladder.h:363-405
    bool hole_index_eligible(int64_t hole_index) const {
        if (h <= 1 || h >= N - 1) return hole_index == 0;
        std::vector<int> pos(h);
        int rem = h;
        int64_t t = hole_index;
        for (int j = N - 1; j >= 0; --j) {
            int64_t c = C(rem, j);
            if (t >= c) {
                pos[rem - 1] = j;
                --rem;
                t -= c;
                if (rem == 0) break;
            }
        }
        std::vector<state_t> st(N, UP);
        for (int p : pos) st[p] = HOLE;
        for (int i = 0; i < N; ++i)
            if (st[i] == UP) {
                st[i] = DOWN;
                break;
            }
        std::vector<Image> im;
        return sym.orbit(st.data(), true, im) > 0;
    }

    // general/basis_indexing.h:188-242 with general/symmetry.h:199-224
    void build_basis(int threads) {
        std::vector<int64_t> elig;
        for (int64_t i = 0; i < n_hole; ++i)
            if (hole_index_eligible(i)) elig.push_back(i);
        std::vector<std::vector<int64_t>> idx(elig.size());
        std::vector<std::vector<int32_t>> cnt(elig.size());
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
        for (long long e = 0; e < (long long)elig.size(); ++e) {
            std::vector<state_t> st(N);
            std::vector<Image> im;
            const int64_t lo = elig[e] * n_spin, hi = lo + n_spin;
            for (int64_t s = lo; s < hi; ++s) {
                unrank(s, st.data());
                int n = sym.orbit(st.data(), true, im);
                if (n) {
                    idx[e].push_back(s);
                    cnt[e].push_back(n);
                }
            }
        }
        for (size_t e = 0; e < elig.size(); ++e) {
            dense_to_sparse.insert(dense_to_sparse.end(), idx[e].begin(), idx[e].end());
            orbit_size.insert(orbit_size.end(), cnt[e].begin(), cnt[e].end());
        }
    }

    int64_t dim() const { return int64_t(dense_to_sparse.size()); }

    // general/basis_indexing.h:152-168
    int64_t dense(int64_t sparse) const {
        if (sparse < 0) return -1;
        auto it = std::lower_bound(dense_to_sparse.begin(), dense_to_sparse.end(), sparse);
        if (it == dense_to_sparse.end() || *it != sparse) return -1;
        return int64_t(it - dense_to_sparse.begin());
    }

    // general/symmetry.h:240-269
    void lowest(const state_t* st, int64_t& ind, cplx& amp) const {
        std::vector<Image> im;
        int n = sym.orbit(st, false, im);
        if (n == 0) {
            ind = -1;
            amp = 0;
            return;
        }
        const state_t* low = st;
        cplx c = im[0].amp;
        for (int i = 1; i < n; ++i)
            if (compare_states(low, im[i].st.data(), N) == 2) {
                low = im[i].st.data();
                c = im[i].amp;
            }
        ind = rank(low);
        amp = c;
    }
};

// ---------------------------------------------------------------------------
// Operator rules (general/column_functions.h, tJ_spin_half_ladder/
// matrix_element_functions.pyx, tJ_spin_half_ladder/matrix_setup.pyx).

enum Valid { V_HOP, V_SWAP, V_NO_HOLES, V_ALL_UP, V_ALL_DOWN, V_ALL_HOLES, V_UNEQUAL_SPINS_2HOLES };
enum NewState { N_SWAP, N_SAME, N_ADD_DOWN, N_ADD_UP, N_ADD_HOLE, N_REMOVE2_ADD_UP_DOWN };

struct Rule {
    std::vector<std::vector<int>> combos;  // reversed index lists
    std::vector<std::pair<cplx, cplx>> weights;
    Valid valid;
    NewState make;
    bool fermion_sign;  // count_commutations over non-hole sites
    bool spin_projection_weight;  // get_weights_spin_projection
    bool density_weight = false;  // N(q): no reference counterpart
    bool singlet_weight = false;  // get_weights_annihilation_part_singlet_singlet
};

// tJ_spin_half_ladder/symmetry_ladder.h:96-101
int shifted_index(int N, int i, int s0, int s1) {
    i = py_mod(i, N);
    int leg = i / 2 + s0;
    int rung = py_mod(i + (s1 % 2), 2);
    return py_mod(2 * leg + rung, N);
}

// general/matrix_rules_utils.pyx:15-87 (+ Validator_Unique_Pairs :136-148)
// for two-site combinations starting at every site.
void pair_combos(int N, const std::vector<std::pair<int, int>>& shifts, int validator,
                 cplx w0, cplx w1, Rule& r) {
    std::vector<std::pair<int, int>> seen;
    for (int i = 0; i < N; ++i)
        for (auto sh : shifts) {
            std::pair<int, int> pr{i, shifted_index(N, i, sh.first, sh.second)};
            if (validator) {
                bool dup = std::find(seen.begin(), seen.end(), pr) != seen.end();
                if (validator == 2)
                    dup = dup || std::find(seen.begin(), seen.end(),
                                           std::make_pair(pr.second, pr.first)) != seen.end();
                if (dup) continue;
                seen.push_back(pr);
            }
            r.combos.push_back({pr.second, pr.first});
            r.weights.push_back({w0, w1});
        }
}

// tJ_spin_half_ladder/matrix_setup.pyx:80-118
Rule hamiltonian_rule(int N, int term) {
    Rule r;
    r.spin_projection_weight = false;
    r.make = N_SWAP;
    if (term == 0) {  // tl
        pair_combos(N, {{1, 0}, {-1, 0}}, 0, 0.0, -1.0, r);
    } else if (term == 1) {  // tr, ordered validator
        pair_combos(N, {{0, 1}, {0, -1}}, 1, 0.0, -1.0, r);
    } else if (term == 2) {  // jl
        pair_combos(N, {{1, 0}}, 0, -0.5, 0.5, r);
    } else {  // jr, unordered validator
        pair_combos(N, {{0, 1}}, 2, -0.5, 0.5, r);
    }
    if (term < 2) {
        r.valid = V_HOP;
        r.fermion_sign = true;
    } else {
        r.valid = V_SWAP;
        r.fermion_sign = false;
    }
    return r;
}

// general/group/symmetry_generator.h:87-102: the per-axis term is an int*int
// divided by a float in single precision, then summed and scaled in double.
double q_dot_r(int site, int qx, int qy, int L) {
    int sx = site / 2, sy = site % 2;
    double acc = 0.0;
    acc += double(float(qx * sx) / float(L));
    acc += double(float(qy * sy) / float(2));
    return acc * 2.0 * std::numbers::pi;
}

// general/matrix_rules_utils.pyx:90-119: base * exp(-1j*q.r)
std::pair<cplx, cplx> iqr_weight(int site, int qx, int qy, int L, cplx b0, cplx b1) {
    const double qr = q_dot_r(site, qx, qy, L);
    const cplx e = std::exp(cplx(0.0, -qr));
    return {b0 * e, b1 * e};
}

// Operators of tJ_spin_half_ladder/matrix_setup.pyx:235-325 (SZQ, A+, A-),
// :248-279 (rung / leg currents) and the density operator N(q) (new).
enum OpKind { OP_SZQ = 0, OP_NQ = 1, OP_APLUS_DOWN = 2, OP_APLUS_UP = 3, OP_AMINUS_DOWN = 4,
              OP_AMINUS_UP = 5, OP_CURRENT_RUNG = 6, OP_CURRENT_LEG = 7 };

Rule operator_rule(int L, int op, int qx, int qy) {
    const int N = 2 * L;
    Rule r;
    r.spin_projection_weight = false;
    r.fermion_sign = false;
    if (op == OP_SZQ || op == OP_NQ) {
        const double b = (op == OP_SZQ ? 0.5 : 1.0) / std::sqrt(double(N));
        for (int i = 0; i < N; ++i) {
            r.combos.push_back({i});
            r.weights.push_back(iqr_weight(i, qx, qy, L, 0.0, b));
        }
        r.valid = V_NO_HOLES;
        r.make = N_SAME;
        r.spin_projection_weight = (op == OP_SZQ);
        r.density_weight = (op == OP_NQ);
    } else if (op <= OP_AMINUS_UP) {
        const double b = 1.0 / std::sqrt(double(N));
        for (int i = 0; i < N; ++i) {
            r.combos.push_back({i});
            r.weights.push_back(iqr_weight(i, qx, qy, L, 0.0, b));
        }
        r.fermion_sign = true;
        if (op == OP_APLUS_DOWN || op == OP_APLUS_UP) {
            r.valid = V_ALL_HOLES;
            r.make = op == OP_APLUS_DOWN ? N_ADD_DOWN : N_ADD_UP;
        } else {
            r.valid = op == OP_AMINUS_DOWN ? V_ALL_DOWN : V_ALL_UP;
            r.make = N_ADD_HOLE;
        }
    } else if (op == OP_CURRENT_RUNG) {
        // matrix_setup.pyx:248-263 (qy is shifted by one by the caller)
        for (int i = 0; i < N; ++i) {
            auto sw = iqr_weight(i, qx, qy, L, 0.0, cplx(0, 1));
            int j = shifted_index(N, i, 0, 1);
            r.combos.push_back({j, i});
            r.weights.push_back({sw.first * cplx(0), sw.second * cplx(1)});
        }
        r.valid = V_HOP;
        r.make = N_SWAP;
    } else {  // current leg, matrix_setup.pyx:265-279
        const int sh[2] = {1, -1};
        const double ew[2] = {1, -1};
        for (int i = 0; i < N; ++i) {
            auto sw = iqr_weight(i, qx, qy, L, 0.0, cplx(0, 1));
            for (int k = 0; k < 2; ++k) {
                int j = shifted_index(N, i, sh[k], 0);
                r.combos.push_back({j, i});
                r.weights.push_back({sw.first * cplx(0), sw.second * cplx(ew[k])});
            }
        }
        r.valid = V_HOP;
        r.make = N_SWAP;
        r.fermion_sign = true;
    }
    return r;
}

// Position-correlation operators (tJ_spin_half_ladder/matrix_setup.pyx:409-480):
// kind 0 hole projection, 1 S^z product, 2 singlet-singlet.  `sh` holds the
// config's fixed distances followed by the free shift, as (leg, rung) pairs;
// combinations start at every site (combinations_from_start_end_points,
// general/matrix_rules_utils.pyx:15-87, reverse=False).
Rule position_rule(int L, int kind, const std::vector<std::pair<int, int>>& sh) {
    const int N = 2 * L;
    Rule r;
    r.fermion_sign = false;
    r.spin_projection_weight = false;
    if (kind == 2) {
        if (sh.size() != 3) throw std::runtime_error("singlet-singlet needs two fixed distances and a free shift");
        const auto f0 = sh[0], f1 = sh[1], f2 = sh[2];
        const std::pair<int, int> ends[2][3] = {
            {f0, f2, {f1.first + f2.first, f1.second + f2.second}},
            {{-f0.first, -f0.second},
             {f2.first - f0.first, f2.second - f0.second},
             {f1.first + f2.first - f0.first, f1.second + f2.second - f0.second}}};
        for (int i = 0; i < N; ++i)
            for (const auto& e : ends) {
                std::vector<int> c = {i};
                for (const auto& s : e) c.push_back(shifted_index(N, i, s.first, s.second));
                r.combos.push_back(c);
                r.weights.push_back({0.0, 0.5 / N});
            }
        r.valid = V_UNEQUAL_SPINS_2HOLES;
        r.make = N_REMOVE2_ADD_UP_DOWN;
        r.fermion_sign = true;
        r.singlet_weight = true;
    } else {
        for (int i = 0; i < N; ++i) {
            std::vector<int> c = {i};
            for (const auto& s : sh) c.push_back(shifted_index(N, i, s.first, s.second));
            r.combos.push_back(c);
            r.weights.push_back({1.0, 0.0});
        }
        r.valid = kind == 0 ? V_ALL_HOLES : V_NO_HOLES;
        r.make = N_SAME;
        r.spin_projection_weight = kind == 1;
    }
    return r;
}

bool rule_valid(const Rule& r, const state_t* s, const std::vector<int>& c) {
    // tJ_spin_half_ladder/matrix_element_functions.pyx:16-51
    switch (r.valid) {
        case V_HOP: return s[c[0]] == HOLE && s[c[1]] != HOLE;
        case V_SWAP: return s[c[0]] != s[c[1]] && s[c[0]] != HOLE && s[c[1]] != HOLE;
        case V_NO_HOLES:
            for (int i : c) if (s[i] == HOLE) return false;
            return true;
        case V_ALL_UP:
            for (int i : c) if (s[i] != UP) return false;
            return true;
        case V_ALL_DOWN:
            for (int i : c) if (s[i] != DOWN) return false;
            return true;
        case V_ALL_HOLES:
            for (int i : c) if (s[i] != HOLE) return false;
            return true;
        case V_UNEQUAL_SPINS_2HOLES: {  // matrix_element_functions.pyx:53-62
            if (c[0] == c[1]) return false;
            const bool first = c[0] == c[2] || c[0] == c[3] || s[c[0]] == HOLE;
            const bool second = c[1] == c[2] || c[1] == c[3] || s[c[1]] == HOLE;
            return first && second && s[c[2]] != s[c[3]] && s[c[2]] != HOLE && s[c[3]] != HOLE;
        }
    }
    return false;
}

// general/column_functions.h:317-346 with counted value HOLE, compare_equals=false
int count_commutations(const state_t* s, int n, const std::vector<int>& idx) {
    std::vector<int> opn(n, 0);
    for (int i = 0; i < n; ++i)
        if (s[i] != HOLE)
            for (int j = 0; j < i; ++j) opn[j] += 1;
    int count = 0;
    for (int k = int(idx.size()) - 1; k >= 0; --k) {
        count += opn[idx[k]];
        for (int j = 0; j < idx[k]; ++j) opn[j] += 1;
    }
    return count;
}

// general/column_functions.h:48-60
void add_result(cplx v, int64_t j, std::vector<cplx>& vals, std::vector<int64_t>& inds) {
    for (size_t k = 0; k < inds.size(); ++k)
        if (inds[k] == j) {
            vals[k] += v;
            return;
        }
    vals.push_back(v);
    inds.push_back(j);
}

// general/column_functions.h:150-207 (commutes_with_symmetries branch), with
// add_generic_matrix_element :99-148 and order_nonzero_value_index_pairs :79-97.
void column(const Sector& init, const Sector& fin, const Rule& r, bool triangle, int64_t i,
            std::vector<cplx>& vals, std::vector<int64_t>& inds) {
    vals.clear();
    inds.clear();
    const int N = init.N;
    std::vector<state_t> st(N), ns(N);
    init.unrank(init.dense_to_sparse[i], st.data());
    const int n = init.orbit_size[i];
    const cplx dw0 = 1;
    const cplx ow0 = std::sqrt(n);
    for (size_t c = 0; c < r.combos.size(); ++c) {
        const auto& cb = r.combos[c];
        cplx dw = dw0 * r.weights[c].first;
        cplx ow = ow0 * r.weights[c].second;
        if (!rule_valid(r, st.data(), cb)) continue;
        if (r.spin_projection_weight) {  // matrix_element_functions.pyx:71-82
            cplx f = 1;
            for (int p : cb)
                if (st[p] == DOWN) f *= -1;
            dw *= f;
            ow *= f;
        }
        if (r.singlet_weight && st[cb[2]] == DOWN) {  // matrix_element_functions.pyx:84-95
            dw *= -1;
            ow *= -1;
        }
        if (std::abs(dw) > 1e-8) add_result(dw, i, vals, inds);
        if (std::abs(ow) < 1e-8) continue;
        ns = st;
        switch (r.make) {
            case N_SWAP:
                ns[cb[0]] = st[cb[1]];
                ns[cb[1]] = st[cb[0]];
                break;
            case N_SAME: break;
            case N_ADD_DOWN: for (int p : cb) ns[p] = DOWN; break;
            case N_ADD_UP: for (int p : cb) ns[p] = UP; break;
            case N_ADD_HOLE: for (int p : cb) ns[p] = HOLE; break;
            case N_REMOVE2_ADD_UP_DOWN: {  // matrix_element_functions.pyx:125-137, last index first
                const state_t changes[4] = {HOLE, HOLE, UP, DOWN};
                const int len = int(cb.size());
                for (int k = 0; k < len; ++k) ns[cb[len - 1 - k]] = changes[k % 4];
                break;
            }
        }
        int64_t sp;
        cplx amp;
        fin.lowest(ns.data(), sp, amp);
        int64_t j = fin.dense(sp);
        if (j < 0 || (triangle && j < i)) continue;
        cplx v = ow * amp;
        if (r.fermion_sign) v *= (count_commutations(st.data(), N, cb) % 2 == 0 ? 1 : -1);
        add_result(v, j, vals, inds);
    }
    // drop near-zeros, sort by index
    std::vector<std::pair<int64_t, cplx>> z;
    for (size_t k = 0; k < inds.size(); ++k)
        if (!(std::abs(vals[k]) < 1e-8)) z.push_back({inds[k], vals[k]});
    std::sort(z.begin(), z.end(), [](auto& a, auto& b) { return a.first < b.first; });
    vals.resize(z.size());
    inds.resize(z.size());
    for (size_t k = 0; k < z.size(); ++k) {
        inds[k] = z[k].first;
        vals[k] = z[k].second;
    }
}

// The density operator N(q) multiplies each site weight by the occupation
// (1 for an electron); it reuses the Szq path with spin_projection off.

// general/cython_sparse_matrices.pyx:58-155 (row-major Hermitian storage):
// the column-major lower triangle is kept as-is and its unique-value table
// conjugated.
struct Csr {
    std::vector<cplx> values;
    std::vector<int16_t> value_inds;
    std::vector<int64_t> minor, major;
};

void dedup_append(Csr& m, cplx v) {
    for (size_t u = 0; u < m.values.size(); ++u)
        if (std::abs(v - m.values[u]) < 1e-6) {
            m.value_inds.push_back(int16_t(u));
            return;
        }
    if (m.values.size() == 32767) throw std::runtime_error("too many unique values");
    m.value_inds.push_back(int16_t(m.values.size()));
    m.values.push_back(v);
}

struct Columns {
    std::vector<int64_t> ptr, rows;
    std::vector<cplx> vals;
};

Columns all_columns(const Sector& init, const Sector& fin, const Rule& r, bool triangle, int threads) {
    const int64_t D = init.dim();
    std::vector<std::vector<cplx>> cv(D);
    std::vector<std::vector<int64_t>> ci(D);
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads)
    for (long long i = 0; i < D; ++i) column(init, fin, r, triangle, i, cv[i], ci[i]);
    Columns out;
    out.ptr.assign(D + 1, 0);
    for (int64_t i = 0; i < D; ++i) out.ptr[i + 1] = out.ptr[i] + int64_t(ci[i].size());
    out.rows.reserve(out.ptr[D]);
    out.vals.reserve(out.ptr[D]);
    for (int64_t i = 0; i < D; ++i) {
        out.rows.insert(out.rows.end(), ci[i].begin(), ci[i].end());
        out.vals.insert(out.vals.end(), cv[i].begin(), cv[i].end());
    }
    return out;
}

Csr to_csr(const Columns& c, bool hermitian_row_major) {
    Csr m;
    m.major = c.ptr;
    m.minor = c.rows;
    for (const cplx& v : c.vals) dedup_append(m, v);
    if (m.values.empty()) {  // cython_sparse_matrices.pyx:123-129
        m.values.push_back(0);
        m.value_inds.push_back(0);
        m.minor.push_back(0);
        m.major.back() = 1;
        // the reference keeps an explicit 0 at (0, 0)
    }
    if (hermitian_row_major)
        for (auto& v : m.values) v = std::conj(v);
    return m;
}

// general/cython_sparse_dot.pyx:40-80 + general/cpp_sparse_dot.h:59-71 and
// :104-138 for row-major, sorted, Hermitian-triangle storage.  The result
// does not depend on the thread count: every output element receives its
// explicit-pass sum first and then Hermitian-pass terms in ascending row order.
void csr_dot_add(const Csr& m, const cplx* x, cplx* y, int64_t D, cplx w, bool first, int threads) {
    if (m.minor.empty()) return;
    if (first)
        for (int64_t i = 0; i < D; ++i) y[i] = 0;
#pragma omp parallel num_threads(threads)
    {
#pragma omp for schedule(dynamic, 10000)
        for (long long r = 0; r < D; ++r) {
            cplx acc = 0;
            for (int64_t j = m.major[r]; j < m.major[r + 1]; ++j) acc += x[m.minor[j]] * m.values[m.value_inds[j]];
            y[r] += w * acc;
        }
        int nt = 1, tid = 0;
#ifdef _OPENMP
        nt = omp_get_num_threads();
        tid = omp_get_thread_num();
#endif
        int64_t step = D / nt + 1;
        int64_t lo = step * tid, hi = tid == nt - 1 ? D : std::min<int64_t>(D, step * (tid + 1));
        for (int64_t r = 0; r < hi; ++r)
            for (int64_t j = m.major[r]; j < m.major[r + 1]; ++j) {
                int64_t c = m.minor[j];
                if (c >= hi) break;
                if (c < lo || c == r) continue;
                y[c] += x[r] * std::conj(m.values[m.value_inds[j]]) * w;
            }
    }
}

// term[0..3]: the Hamiltonian terms in the default layout; term[EXTRA]: the
// last matrix built by orc_build_layout_csr or orc_build_op_csr.
constexpr int EXTRA = 4;
struct Hamiltonian {
    Csr term[5];
    bool built[5] = {false, false, false, false, false};
};

struct Handle {
    std::unique_ptr<Sector> sec;
    Hamiltonian ham;
};

}  // namespace

// ---------------------------------------------------------------------------
// C ABI used by the Python test helpers (oracle/oracle.py).

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

void* orc_sector_new(int L, int h, int nd, int kx, int ky, int threads) {
    try {
        auto* hd = new Handle;
        hd->sec.reset(new Sector(L, h, nd, kx, ky));
        hd->sec->build_basis(threads);
        return hd;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void orc_sector_free(void* p) { delete static_cast<Handle*>(p); }

int64_t orc_sector_dim(void* p) { return static_cast<Handle*>(p)->sec->dim(); }

void orc_sector_basis(void* p, int64_t* sparse, int32_t* counts) {
    auto& s = *static_cast<Handle*>(p)->sec;
    std::copy(s.dense_to_sparse.begin(), s.dense_to_sparse.end(), sparse);
    std::copy(s.orbit_size.begin(), s.orbit_size.end(), counts);
}

int64_t orc_comb(int size, int j, int i) {
    auto c = make_comb(size);
    return c[size_t(j) * (size + 1) + i];
}

void orc_unrank(void* p, int64_t sparse, short* st) { static_cast<Handle*>(p)->sec->unrank(sparse, st); }

int64_t orc_rank(void* p, const short* st) { return static_cast<Handle*>(p)->sec->rank(st); }

int64_t orc_dense(void* p, int64_t sparse) { return static_cast<Handle*>(p)->sec->dense(sparse); }

int orc_hole_eligible(void* p, int64_t hole_index) {
    return static_cast<Handle*>(p)->sec->hole_index_eligible(hole_index) ? 1 : 0;
}

// Orbit of a state: returns n, fills images (n*N shorts) and amplitudes (2n doubles).
int orc_orbit(void* p, const short* st, int check_lowest, short* out_states, double* out_amp) {
    auto& s = *static_cast<Handle*>(p)->sec;
    std::vector<Image> im;
    int n = s.sym.orbit(st, check_lowest != 0, im);
    for (int k = 0; k < n; ++k) {
        std::copy(im[k].st.begin(), im[k].st.end(), out_states + size_t(k) * s.N);
        out_amp[2 * k] = im[k].amp.real();
        out_amp[2 * k + 1] = im[k].amp.imag();
    }
    return n;
}

void orc_lowest(void* p, const short* st, int64_t* ind, double* amp) {
    cplx a;
    static_cast<Handle*>(p)->sec->lowest(st, *ind, a);
    amp[0] = a.real();
    amp[1] = a.imag();
}

// Columns of a Hamiltonian term (0..3 = tl, tr, jl, jr) before
// de-duplication.  Call with rows == nullptr to get nnz.
int64_t orc_term_columns(void* p, int term, int triangle, int threads, int64_t* ptr, int64_t* rows,
                         double* vals) {
    auto& s = *static_cast<Handle*>(p)->sec;
    Rule r = hamiltonian_rule(s.N, term);
    Columns c = all_columns(s, s, r, triangle != 0, threads);
    if (rows) {
        std::copy(c.ptr.begin(), c.ptr.end(), ptr);
        std::copy(c.rows.begin(), c.rows.end(), rows);
        std::memcpy(vals, c.vals.data(), c.vals.size() * sizeof(cplx));
    }
    return int64_t(c.rows.size());
}

// Operator columns between two sectors (op = OpKind).
int64_t orc_op_columns(void* pi, void* pf, int op, int qx, int qy, int threads, int64_t* ptr, int64_t* rows,
                       double* vals) {
    auto& si = *static_cast<Handle*>(pi)->sec;
    auto& sf = *static_cast<Handle*>(pf)->sec;
    Rule r = operator_rule(si.L, op, qx, qy);
    Columns c = all_columns(si, sf, r, false, threads);
    if (rows) {
        std::copy(c.ptr.begin(), c.ptr.end(), ptr);
        std::copy(c.rows.begin(), c.rows.end(), rows);
        std::memcpy(vals, c.vals.data(), c.vals.size() * sizeof(cplx));
    }
    return int64_t(c.rows.size());
}

// Operator site weights as the reference builds them (complex pairs).
// q.r of general/group/symmetry_generator.h:87-102 for one site (radians).
double orc_q_dot_r(int L, int site, int qx, int qy) { return q_dot_r(site, qx, qy, L); }

void orc_op_weights(int L, int op, int qx, int qy, double* out) {
    Rule r = operator_rule(L, op, qx, qy);
    for (size_t c = 0; c < r.weights.size(); ++c) {
        out[4 * c + 0] = r.weights[c].first.real();
        out[4 * c + 1] = r.weights[c].first.imag();
        out[4 * c + 2] = r.weights[c].second.real();
        out[4 * c + 3] = r.weights[c].second.imag();
    }
}

// Position-correlation operator columns on one sector (same layout as orc_op_columns).
int64_t orc_position_columns(void* p, int kind, const int* shifts, int n_shifts, int threads, int64_t* ptr,
                             int64_t* rows, double* vals) {
    try {
        auto& s = *static_cast<Handle*>(p)->sec;
        std::vector<std::pair<int, int>> sh;
        for (int k = 0; k < n_shifts; ++k) sh.push_back({shifts[2 * k], shifts[2 * k + 1]});
        Rule r = position_rule(s.L, kind, sh);
        Columns c = all_columns(s, s, r, false, threads);
        if (rows) {
            std::copy(c.ptr.begin(), c.ptr.end(), ptr);
            std::copy(c.rows.begin(), c.rows.end(), rows);
            std::memcpy(vals, c.vals.data(), c.vals.size() * sizeof(cplx));
        }
        return int64_t(c.rows.size());
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// calculate_hole_spin_projection_correlations / calculate_position_correlations
// (tJ_spin_half_ladder/position_correlations.pyx:84-127, :186-230): out[d] for
// every node index d (N values).  kind 0: holes (aggregated per hole pattern,
// Hole_Index_Aggregator), 1: S^z products (each dense state).
int orc_position_correlations(void* p, int kind, const int* fixed, int n_fixed, const double* psi, double* out) {
    try {
        auto& s = *static_cast<Handle*>(p)->sec;
        const int N = s.N;
        const int64_t D = s.dim();
        const cplx* v = reinterpret_cast<const cplx*>(psi);
        std::vector<state_t> st(N);
        for (int d = 0; d < N; ++d) {
            // positions[r] = (r, r shifted by the unit shifts of d, r shifted by each fixed distance)
            std::vector<std::vector<int>> pos(N);
            for (int r = 0; r < N; ++r) {
                pos[r] = {r, shifted_index(N, r, d / 2, d % 2)};
                for (int k = 0; k < n_fixed; ++k) pos[r].push_back(shifted_index(N, r, fixed[2 * k], fixed[2 * k + 1]));
            }
            double corr = 0;
            int64_t i = 0;
            while (i < D) {
                const int64_t sparse = s.dense_to_sparse[i];
                cplx amp;
                int64_t next;
                if (kind == 0) {
                    const int64_t hole = sparse / s.n_spin;
                    const int64_t top = std::min<int64_t>(D, i + s.n_spin);
                    double value = 0;
                    next = top;
                    for (int64_t k = i; k < top; ++k) {
                        if (s.dense_to_sparse[k] / s.n_spin != hole) {
                            next = k;
                            break;
                        }
                        value += std::norm(v[k]);
                    }
                    amp = std::sqrt(cplx(value));
                } else {
                    amp = v[i];
                    next = i + 1;
                }
                s.unrank(sparse, st.data());
                const double an = std::norm(amp);
                for (int r = 0; r < N; ++r) {
                    int val = 1;
                    for (int q : pos[r]) {
                        if (kind == 0) {
                            if (st[q] != HOLE) {
                                val = 0;
                                break;
                            }
                        } else {
                            if (st[q] == HOLE) {
                                val = 0;
                                break;
                            }
                            if (st[q] == DOWN) val *= -1;
                        }
                    }
                    corr += an * val;
                }
                i = next;
            }
            out[d] = corr;
        }
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// Build and keep the reference's stored CSR of a Hamiltonian term.
int orc_build_term_csr(void* p, int term, int threads) {
    try {
        auto* hd = static_cast<Handle*>(p);
        Rule r = hamiltonian_rule(hd->sec->N, term);
        Columns c = all_columns(*hd->sec, *hd->sec, r, true, threads);
        hd->ham.term[term] = to_csr(c, true);
        hd->ham.built[term] = true;
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// A Hamiltonian term in any of the reference's layouts (config.triangle_only,
// config.major; cython_sparse_matrices.pyx:110-184): the column arrays are
// kept and the value table is conjugated for row-major storage.  Kept in the
// EXTRA slot.
int orc_build_layout_csr(void* p, int term, int triangle, int row_major, int threads) {
    try {
        auto* hd = static_cast<Handle*>(p);
        Rule r = hamiltonian_rule(hd->sec->N, term);
        Columns c = all_columns(*hd->sec, *hd->sec, r, triangle != 0, threads);
        hd->ham.term[EXTRA] = to_csr(c, row_major != 0);
        hd->ham.built[EXTRA] = true;
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// An excitation operator as setup_excitation_operator stores it
// (column-major, tJ_spin_half_ladder/matrix_setup.pyx:340-371), kept in the
// EXTRA slot of the initial sector's handle.
int orc_build_op_csr(void* pi, void* pf, int op, int qx, int qy, int threads) {
    try {
        auto* hi = static_cast<Handle*>(pi);
        auto& sf = *static_cast<Handle*>(pf)->sec;
        Rule r = operator_rule(hi->sec->L, op, qx, qy);
        Columns c = all_columns(*hi->sec, sf, r, false, threads);
        hi->ham.term[EXTRA] = to_csr(c, false);
        hi->ham.built[EXTRA] = true;
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// The reference's stored CSR of one term restricted to rows [lo, hi) of the
// row-major upper triangle (a bounded sample for the CPU baseline).
int orc_build_term_csr_rows(void* p, int term, int64_t lo, int64_t hi, int threads) {
    try {
        auto* hd = static_cast<Handle*>(p);
        const Sector& s = *hd->sec;
        Rule r = hamiltonian_rule(s.N, term);
        const int64_t n = hi - lo;
        std::vector<std::vector<cplx>> cv(n);
        std::vector<std::vector<int64_t>> ci(n);
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads)
        for (long long k = 0; k < n; ++k) column(s, s, r, true, lo + k, cv[k], ci[k]);
        Columns c;
        c.ptr.assign(n + 1, 0);
        for (int64_t k = 0; k < n; ++k) c.ptr[k + 1] = c.ptr[k] + int64_t(ci[k].size());
        for (int64_t k = 0; k < n; ++k) {
            c.rows.insert(c.rows.end(), ci[k].begin(), ci[k].end());
            c.vals.insert(c.vals.end(), cv[k].begin(), cv[k].end());
        }
        hd->ham.term[term] = to_csr(c, true);
        hd->ham.built[term] = true;
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// Sizes of the stored CSR of one term: returns the stored elements; n_unique
// receives the unique values and n_rows the rows it covers (all rows, or the
// rows of the last orc_build_term_csr_rows call).
int64_t orc_term_csr_sizes(void* p, int term, int64_t* n_unique, int64_t* n_rows) {
    auto& m = static_cast<Handle*>(p)->ham.term[term];
    *n_unique = int64_t(m.values.size());
    *n_rows = m.major.empty() ? 0 : int64_t(m.major.size()) - 1;
    return int64_t(m.minor.size());
}

void orc_term_csr_copy(void* p, int term, double* values, int16_t* value_inds, int64_t* minor, int64_t* major) {
    auto& m = static_cast<Handle*>(p)->ham.term[term];
    std::memcpy(values, m.values.data(), m.values.size() * sizeof(cplx));
    std::copy(m.value_inds.begin(), m.value_inds.end(), value_inds);
    std::copy(m.minor.begin(), m.minor.end(), minor);
    std::copy(m.major.begin(), m.major.end(), major);
}

// general/sparse_matrices.py:233-237: y = sum_T w_T * M_T x over terms with
// nonzero weight, in tl, tr, jl, jr order.  Weights are complex pairs.
void orc_hv(void* p, const double* w, const double* x, double* y, int threads) {
    auto* hd = static_cast<Handle*>(p);
    const int64_t D = hd->sec->dim();
    bool first = true;
    for (int t = 0; t < 4; ++t) {
        cplx wt(w[2 * t], w[2 * t + 1]);
        if (wt == cplx(0) || !hd->ham.built[t]) continue;
        if (std::abs(wt) < 1e-6) continue;
        csr_dot_add(hd->ham.term[t], reinterpret_cast<const cplx*>(x), reinterpret_cast<cplx*>(y), D, wt, first,
                    threads);
        first = false;
    }
    if (first)
        for (int64_t i = 0; i < 2 * D; ++i) y[i] = 0;
}

// general/cython_lanczos_diagonalization.pyx:42-48: Im(1/det) of the
// continued fraction at each z.
void orc_cf(const double* a, const double* b, int n, const double* z, int nz, double* out) {
    for (int k = 0; k < nz; ++k) {
        cplx zz(z[2 * k], z[2 * k + 1]);
        cplx det = zz - a[n - 1];
        for (int i = 1; i < n; ++i) det = zz - a[n - 1 - i] - b[n - i] * b[n - i] / det;
        out[k] = (1.0 / det).imag();
    }
}

// general/cython_lanczos_diagonalization.pyx:50-77 with the stored CSR
// Hamiltonian; sequential sums stand in for BLAS dot and nrm2.  Returns the
// number of recorded coefficients.
int orc_lanczos(void* p, const double* w, const double* v0, int nsteps, double* a, double* b, int threads) {
    auto* hd = static_cast<Handle*>(p);
    const int64_t D = hd->sec->dim();
    std::vector<cplx> vn(D), vm(D, 0.0), hv(D), vp(D);
    const cplx* in = reinterpret_cast<const cplx*>(v0);
    double nrm = 0;
    for (int64_t i = 0; i < D; ++i) nrm += std::norm(in[i]);
    nrm = std::sqrt(nrm);
    for (int64_t i = 0; i < D; ++i) vn[i] = in[i] / nrm;
    double bn = 0;
    b[0] = 0;
    int done = 0;
    for (int it = 0; it < nsteps; ++it) {
        orc_hv(p, w, reinterpret_cast<double*>(vn.data()), reinterpret_cast<double*>(hv.data()), threads);
        cplx dot = 0;
        for (int64_t i = 0; i < D; ++i) dot += std::conj(vn[i]) * hv[i];
        double an = dot.real();
        double s = 0;
        for (int64_t i = 0; i < D; ++i) {
            vp[i] = hv[i] - an * vn[i] - bn * vm[i];
            s += std::norm(vp[i]);
        }
        double bp = std::sqrt(s);
        a[it] = an;
        b[it + 1] = bp;
        ++done;
        if (bp < 1e-8) break;
        for (int64_t i = 0; i < D; ++i) vp[i] /= bp;
        vm.swap(vn);
        vn.swap(vp);
        bn = bp;
    }
    return done;
}

}  // extern "C"
