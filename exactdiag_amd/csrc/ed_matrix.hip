// Matrix elements, matrix-free H·v and excitation operators for gfx950.
//
// Two kernels compute matrix elements:
//  * column_kernel: one lane per column, the reference's per-column
//    algorithm step by step (general/column_functions.h:99-207): rules in
//    combination order, the orbit walk for every final state, duplicate
//    merging in order, near-zero removal, sort by row.  Bit-exact values;
//    used for parity, for operator matrices and for small sectors.
//  * hv_kernel: one lane per row in gather form, y_f = sum_j conj(H_jf) x_j,
//    which by hermiticity equals row f of the reference's matrix.  Bonds are
//    found with whole-mask bit operations, representatives through the
//    hole-pattern table and indices through combinatorial ranking, so no
//    orbit walk and no search runs in the hot loop.  No atomics; the
//    per-workgroup partial of <x, y> feeds the Lanczos coefficient a_n.

#include <algorithm>
#include <cstdlib>
#include <vector>

#include "ed_internal.h"
#include "ed_orbit.h"
#include "ed_matrix.h"

namespace ed {

namespace {

constexpr int MAXC = 72;  // per-column capacity of the exact kernel

struct RuleView {
    int n_combo, arity, valid, make, spin_proj, density, fermion, singlet, max_per_col;
    const int8_t* idx;
    const double* w;
};

__device__ inline int site_value(uint32_t H, uint32_t Dn, int i) {  // 0 hole, 1 down, 2 up
    const uint32_t b = 1u << i;
    return (H & b) ? 0 : ((Dn & b) ? 1 : 2);
}

// Validity of a combination of r.arity sites c[0..arity)
// (tJ_spin_half_ladder/matrix_element_functions.pyx:16-62).
__device__ inline bool rule_valid(const RuleView& r, uint32_t H, uint32_t Dn, const int* c) {
    switch (r.valid) {
        case RV_HOP: return site_value(H, Dn, c[0]) == 0 && site_value(H, Dn, c[1]) != 0;
        case RV_SWAP: {
            const int va = site_value(H, Dn, c[0]), vb = site_value(H, Dn, c[1]);
            return va != vb && va != 0 && vb != 0;
        }
        case RV_NO_HOLES:
        case RV_ALL_UP:
        case RV_ALL_DOWN:
        case RV_ALL_HOLES: {
            for (int k = 0; k < r.arity; ++k) {
                const int v = site_value(H, Dn, c[k]);
                if (r.valid == RV_NO_HOLES ? v == 0 : v != (r.valid == RV_ALL_UP ? 2 : r.valid == RV_ALL_DOWN ? 1 : 0))
                    return false;
            }
            return true;
        }
        case RV_UNEQUAL_SPINS_2HOLES: {
            if (c[0] == c[1]) return false;
            const bool first = c[0] == c[2] || c[0] == c[3] || site_value(H, Dn, c[0]) == 0;
            const bool second = c[1] == c[2] || c[1] == c[3] || site_value(H, Dn, c[1]) == 0;
            const int v2 = site_value(H, Dn, c[2]), v3 = site_value(H, Dn, c[3]);
            return first && second && v2 != v3 && v2 != 0 && v3 != 0;
        }
    }
    return false;
}

__device__ inline void set_site(uint32_t& H, uint32_t& Dn, int i, int v) {
    const uint32_t b = 1u << i;
    H = v == 0 ? (H | b) : (H & ~b);
    Dn = v == 1 ? (Dn | b) : (Dn & ~b);
}

// general/column_functions.h:317-346: operators act from the last index to
// the first; each counts the electrons above it plus the operators already
// placed above it (a site placed twice counts twice).
__device__ inline int commutation_parity(uint32_t E, int arity, const int* idx) {
    int count = 0;
    for (int k = arity - 1; k >= 0; --k) {
        const int p = idx[k];
        const uint32_t above = p >= 31 ? 0u : (~0u << (p + 1));
        count += popc32(E & above);
        for (int m = k + 1; m < arity; ++m) count += idx[m] > p;
    }
    return count & 1;
}

__device__ inline bool cabs_gt(C2 v, double thr2) { return v.re * v.re + v.im * v.im > thr2; }

// The elements of column i of the rule's matrix (rows of sector fin), merged
// by row, near-zeros dropped, sorted by row: cv / ci[0..return value);
// overflow is set when more than MAXC distinct rows were reached.
__device__ int column_elements(const SectorView& ini, const SectorView& fin, const RuleView& r, int triangle,
                               int64_t i, C2* cv, int64_t* ci, bool& overflow) {
    const uint64_t key = ini.key[i];
    const uint32_t H = uint32_t(key >> 32), Dn = uint32_t(key);
    const uint32_t E = ~H & ini.full;
    const int n = ini.orbit[i];
    const C2 dw0{1.0, 0.0};
    const C2 ow0{ini.sqrt_tab[n], 0.0};
    int cnt = 0;
    overflow = false;
    auto add = [&](C2 v, int64_t j) {
        for (int k = 0; k < cnt; ++k)
            if (ci[k] == j) {
                cv[k] = cadd(cv[k], v);
                return;
            }
        if (cnt < MAXC) {
            cv[cnt] = v;
            ci[cnt] = j;
            ++cnt;
        } else {
            overflow = true;
        }
    };
    // state-preserving rules reach the same final state from every combination
    bool same_cached = false;
    int64_t same_j = -1;
    C2 same_c{0, 0};
    for (int c = 0; c < r.n_combo; ++c) {
        int cs[RULE_W];
        for (int k = 0; k < RULE_W; ++k) cs[k] = r.idx[RULE_W * c + k];
        const int a = cs[0], b = cs[1];
        C2 dw = cmul(dw0, C2{r.w[4 * c], r.w[4 * c + 1]});
        C2 ow = cmul(ow0, C2{r.w[4 * c + 2], r.w[4 * c + 3]});
        if (!rule_valid(r, H, Dn, cs)) continue;
        if (r.spin_proj) {  // tJ_spin_half_ladder/matrix_element_functions.pyx:71-82
            C2 f{1.0, 0.0};
            for (int k = 0; k < r.arity; ++k)
                if (site_value(H, Dn, cs[k]) == 1) f = cmul(f, C2{-1.0, 0.0});
            dw = cmul(dw, f);
            ow = cmul(ow, f);
        }
        if (r.singlet && site_value(H, Dn, cs[2]) == 1) {  // matrix_element_functions.pyx:84-95
            dw = cmul(dw, C2{-1.0, 0.0});
            ow = cmul(ow, C2{-1.0, 0.0});
        }
        if (cabs_gt(dw, 1e-16)) add(dw, i);
        if (ow.re * ow.re + ow.im * ow.im < 1e-16) continue;
        uint32_t H2 = H, D2 = Dn;
        int64_t j;
        C2 cc;
        if (r.make == RM_SAME && same_cached) {
            j = same_j;
            cc = same_c;
        } else {
            switch (r.make) {
                case RM_SWAP: {
                    const int va = site_value(H, Dn, a), vb = site_value(H, Dn, b);
                    set_site(H2, D2, a, vb);
                    set_site(H2, D2, b, va);
                    break;
                }
                case RM_SAME: break;
                case RM_ADD_DOWN:
                    set_site(H2, D2, a, 1);
                    if (r.arity > 1) set_site(H2, D2, b, 1);
                    break;
                case RM_ADD_UP:
                    set_site(H2, D2, a, 2);
                    if (r.arity > 1) set_site(H2, D2, b, 2);
                    break;
                case RM_ADD_HOLE:
                    set_site(H2, D2, a, 0);
                    if (r.arity > 1) set_site(H2, D2, b, 0);
                    break;
                case RM_REMOVE2_ADD_UP_DOWN: {  // matrix_element_functions.pyx:125-137, last index first
                    const int changes[4] = {0, 0, 2, 1};
                    for (int k = 0; k < r.arity; ++k) set_site(H2, D2, cs[r.arity - 1 - k], changes[k & 3]);
                    break;
                }
            }
            const OrbitResult o = orbit_walk(fin, H2, D2, false);
            j = -1;
            cc = C2{0, 0};
            if (o.n > 0) {
                const Found fd = lookup_rep(fin, uint32_t(o.best >> 32), uint32_t(o.best));
                j = fd.j;
                cc = lowest_amplitude(fin, H2, o.best_p, o.n);
            }
            if (r.make == RM_SAME) {
                same_cached = true;
                same_j = j;
                same_c = cc;
            }
        }
        if (j < 0 || (triangle && j < i)) continue;
        C2 v = cmul(ow, cc);
        if (r.fermion && commutation_parity(E, r.arity, cs)) v = C2{v.re * -1.0, v.im * -1.0};
        add(v, j);
    }
    // drop near-zero entries, sort by row (general/column_functions.h:79-97)
    int m = 0;
    for (int k = 0; k < cnt; ++k)
        if (!(cv[k].re * cv[k].re + cv[k].im * cv[k].im < 1e-16)) {
            cv[m] = cv[k];
            ci[m] = ci[k];
            ++m;
        }
    for (int k = 1; k < m; ++k) {
        const int64_t ki = ci[k];
        const C2 kv = cv[k];
        int q = k - 1;
        while (q >= 0 && ci[q] > ki) {
            ci[q + 1] = ci[q];
            cv[q + 1] = cv[q];
            --q;
        }
        ci[q + 1] = ki;
        cv[q + 1] = kv;
    }
    return m;
}

// Columns [lo, lo + n) of the initial sector; column lo + t is written at
// slot t of count / rows / vals.
__global__ void __launch_bounds__(128) column_kernel(SectorView ini, SectorView fin, RuleView r, int triangle,
                                                      int64_t lo, int64_t n_cols, int32_t* count, int64_t* rows,
                                                      double* vals) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n_cols) return;
    C2 cv[MAXC];
    int64_t ci[MAXC];
    bool overflow;
    const int m = column_elements(ini, fin, r, triangle, lo + t, cv, ci, overflow);
    count[t] = overflow ? MAXC + 1 : m;
    const int keep = m < r.max_per_col ? m : r.max_per_col;
    for (int k = 0; k < keep; ++k) {
        rows[t * r.max_per_col + k] = ci[k];
        vals[2 * (t * r.max_per_col + k)] = cv[k].re;
        vals[2 * (t * r.max_per_col + k) + 1] = cv[k].im;
    }
}

// ---------------------------------------------------------------------------
// Matrix-free H·v (gather form).

template <class VT>
__device__ inline C2 load_c(const VT* p, int64_t i) {
    const VT v = p[i];
    return {double(v.x), double(v.y)};
}

template <class VT>
__device__ inline void store_c(VT* p, int64_t i, C2 v) {
    VT o;
    o.x = decltype(o.x)(v.re);
    o.y = decltype(o.y)(v.im);
    p[i] = o;
}

__device__ inline int64_t lookup_in_block(const SectorView& s, int blk, uint32_t Hc, uint32_t Dr, int& n) {
    const int64_t sr = rank_spins(Hc, Dr, s.comb);
    const int64_t off = s.block_table[blk];
    if (off < 0) {
        n = s.N;
        return s.block_base[blk] + sr;
    }
    const int32_t j = s.table[off + sr];
    n = j >= 0 ? int(s.orbit[j]) : 0;
    return j;
}

// conj(val) * x_j accumulated with weight; val = scale * sign * A[p][n_j].
template <class VT>
__device__ inline void accumulate(const SectorView& s, const VT* __restrict__ x, uint32_t Hsrc, int p, int blk,
                                  uint32_t Hc, uint32_t Dr, double scale, C2 W, C2& acc) {
    int nj;
    const int64_t j = lookup_in_block(s, blk, Hc, Dr, nj);
    if (j < 0) return;
    const C2 a = amp_at(s, p, nj);
    const double sg = double(image_sign(Hsrc, p, s.L, s.ne_even, s.full)) * scale;
    const C2 v{sg * a.re, -(sg * a.im)};  // conj(val)
    const C2 xj = load_c(x, j);
    const C2 t = cmul(v, xj);
    acc = cadd(acc, cmul(W, t));
}

template <class VT, int BS>
__global__ void __launch_bounds__(BS) hv_kernel(SectorView s, HvWeights w, const VT* __restrict__ x,
                                                VT* __restrict__ y, double* __restrict__ partial, int remap) {
    int bid = blockIdx.x;
    if (remap) {
        // contiguous row ranges per XCD: consecutive workgroups go round
        // robin over the 8 XCDs, so give XCD k the k-th eighth of the rows
        const int nwg = gridDim.x;
        const int per = nwg / 8;
        if (bid < per * 8) bid = (bid & 7) * per + (bid >> 3);
    }
    const int64_t f = int64_t(bid) * BS + threadIdx.x;
    C2 acc{0.0, 0.0}, dot{0.0, 0.0};
    if (f < s.D) {
        const uint64_t key = s.key[f];
        const uint32_t H = uint32_t(key >> 32), Dn = uint32_t(key);
        const uint32_t full = s.full;
        const uint32_t E = ~H & full, Up = E & ~Dn;
        const double sqf = s.sqrt_tab[s.orbit[f]];
        const C2 xf = load_c(x, f);
        const int64_t hr = rank_holes(H, s.comb);
        const uint32_t gm_self = s.canon_gmask[hr];
        const int blk_self = s.canon_block[hr];
        // exchange: the hole pattern is unchanged, only its stabiliser can
        // lower the new down pattern
        auto exchange = [&](uint32_t D2, C2 W) {
            uint32_t gm = gm_self;
            int bp = ctz32(gm);
            uint32_t best = image_mask(D2, bp, s.N, full);
            gm &= gm - 1;
            while (gm) {
                const int p = ctz32(gm);
                const uint32_t d = image_mask(D2, p, s.N, full);
                if (d < best) {
                    best = d;
                    bp = p;
                }
                gm &= gm - 1;
            }
            accumulate(s, x, H, bp, blk_self, H, best, 0.5 * sqf, W, acc);
        };
        if (w.active & 4) {  // jl: bonds (i, i + 2), general/matrix_setup.pyx:95-101
            const uint32_t d2 = rotl_n(Dn, s.N - 2, s.N, full), u2 = rotl_n(Up, s.N - 2, s.N, full);
            uint32_t anti = ((Dn & u2) | (Up & d2)) & full;
            if (s.N == 2) anti = 0;
            const C2 W = w.w[2];
            const double diag = -0.5 * double(popc32(anti));
            acc = cadd(acc, cmul(W, C2{diag * xf.re, diag * xf.im}));
            while (anti) {
                const int i = ctz32(anti);
                anti &= anti - 1;
                const int k = (i + 2) % s.N;
                exchange(Dn ^ (1u << i) ^ (1u << k), W);
            }
        }
        if (w.active & 8) {  // jr: bonds (2x, 2x + 1)
            uint32_t anti = (Dn ^ (Dn >> 1)) & E & (E >> 1) & EVEN_BITS & full;
            const C2 W = w.w[3];
            const double diag = -0.5 * double(popc32(anti));
            acc = cadd(acc, cmul(W, C2{diag * xf.re, diag * xf.im}));
            while (anti) {
                const int i = ctz32(anti);
                anti &= anti - 1;
                exchange(Dn ^ (3u << i), W);
            }
        }
        if (w.active & 3) {  // hopping onto each hole, tl from both leg neighbours, tr from the rung partner
            uint32_t holes = H;
            while (holes) {
                const int t = ctz32(holes);
                holes &= holes - 1;
                for (int dir = 0; dir < 3; ++dir) {
                    int src;
                    C2 W;
                    if (dir < 2) {
                        if (!(w.active & 1)) continue;
                        src = dir == 0 ? (t + 2) % s.N : (t - 2 + s.N) % s.N;
                        W = w.w[0];
                    } else {
                        if (!(w.active & 2)) continue;
                        src = t ^ 1;
                        W = w.w[1];
                    }
                    if (!(E & (1u << src))) continue;
                    const uint32_t bt = 1u << t, bs = 1u << src;
                    const uint32_t H2 = H ^ bt ^ bs;
                    const uint32_t D2 = (Dn & bs) ? (Dn ^ bs ^ bt) : Dn;
                    const uint32_t above_s = src >= 31 ? 0u : (~0u << (src + 1));
                    const uint32_t above_t = t >= 31 ? 0u : (~0u << (t + 1));
                    const int par = (popc32(E & above_s) + popc32(E & above_t) + (t < src ? 1 : 0)) & 1;
                    const int64_t hr2 = rank_holes(H2, s.comb);
                    const int blk = s.canon_block[hr2];
                    if (blk < 0) continue;
                    uint32_t gm = s.canon_gmask[hr2];
                    int bp = ctz32(gm);
                    const uint32_t Hc = image_mask(H2, bp, s.N, full);
                    uint32_t best = image_mask(D2, bp, s.N, full);
                    gm &= gm - 1;
                    while (gm) {
                        const int p = ctz32(gm);
                        const uint32_t d = image_mask(D2, p, s.N, full);
                        if (d < best) {
                            best = d;
                            bp = p;
                        }
                        gm &= gm - 1;
                    }
                    accumulate(s, x, H2, bp, blk, Hc, best, par ? sqf : -sqf, W, acc);
                }
            }
        }
        store_c(y, f, acc);
        dot = cmul(C2{xf.re, -xf.im}, acc);
    }
    if (partial) {
        // deterministic workgroup reduction of <x, y>
        workgroup_sum2<BS>(dot.re, dot.im);
        if (threadIdx.x == 0) {
            partial[2 * bid] = dot.re;
            partial[2 * bid + 1] = dot.im;
        }
    }
}

// Change of the spin rank when the spin pattern Dn becomes D2 by a swap
// inside the site window [lo, hi] (no other site changes): only downs in the
// window change their colex term, downs above keep their index.
__device__ inline int64_t window_rank_delta(uint32_t H, uint32_t Dn, uint32_t D2, int lo, int hi,
                                            const uint32_t* comb) {
    const uint32_t below = (1u << lo) - 1u;
    int kb = popc32(Dn & below), ka = kb, hb = popc32(H & below);
    int64_t before = 0, after = 0;
    for (int q = lo; q <= hi; ++q) {
        const uint32_t bit = 1u << q;
        if (H & bit) {
            ++hb;
            continue;
        }
        const int c = q - hb;
        if (Dn & bit) before += comb[(++kb) * COMB_W + c];
        if (D2 & bit) after += comb[(++ka) * COMB_W + c];
    }
    return after - before;
}

template <class VT>
__device__ inline void accumulate_at(const SectorView& s, const VT* __restrict__ x, int64_t j, int nj, uint32_t Hsrc,
                                     int p, double scale, C2 W, C2& acc) {
    const C2 a = amp_at(s, p, nj);
    const double sg = double(image_sign(Hsrc, p, s.L, s.ne_even, s.full)) * scale;
    const C2 v{sg * a.re, -(sg * a.im)};
    const C2 t = cmul(v, load_c(x, j));
    acc = cadd(acc, cmul(W, t));
}

// H·v with every lane of a wavefront walking the bonds in the same fixed
// order.  Consecutive rows of a hole-pattern block share the colex high part
// of their spin pattern, so an exchange across two high sites sends all lanes
// to one contiguous range of x (coalesced), and exchanges on low sites stay
// inside the same small slice.  On blocks whose hole pattern has no
// stabiliser, the exchange target is row + (window rank change): no search,
// no full rank.
template <class VT, int BS>
__global__ void __launch_bounds__(BS) hv_lockstep_kernel(SectorView s, HvWeights w, const VT* __restrict__ x,
                                                         VT* __restrict__ y, double* __restrict__ partial,
                                                         int remap) {
    int bid = blockIdx.x;
    if (remap) {
        const int nwg = gridDim.x;
        const int per = nwg / 8;
        if (bid < per * 8) bid = (bid & 7) * per + (bid >> 3);
    }
    const int64_t f = int64_t(bid) * BS + threadIdx.x;
    C2 acc{0.0, 0.0}, dot{0.0, 0.0};
    if (f < s.D) {
        const uint64_t key = s.key[f];
        const uint32_t H = uint32_t(key >> 32), Dn = uint32_t(key);
        const uint32_t full = s.full;
        const uint32_t E = ~H & full, Up = E & ~Dn;
        const double sqf = s.sqrt_tab[s.orbit[f]];
        const C2 xf = load_c(x, f);
        const int64_t hr = rank_holes(H, s.comb);
        const uint32_t gm_self = s.canon_gmask[hr];
        const int blk_self = s.canon_block[hr];
        const bool generic = (gm_self & (gm_self - 1u)) == 0u;
        auto exchange = [&](uint32_t D2, int lo, int hi, bool wrap, C2 W) {
            if (generic) {
                const int64_t j = wrap ? s.block_base[blk_self] + rank_spins(H, D2, s.comb)
                                       : f + window_rank_delta(H, Dn, D2, lo, hi, s.comb);
                accumulate_at(s, x, j, s.N, H, 0, 0.5 * sqf, W, acc);
                return;
            }
            uint32_t gm = gm_self;
            int bp = ctz32(gm);
            uint32_t best = image_mask(D2, bp, s.N, full);
            gm &= gm - 1;
            while (gm) {
                const int p = ctz32(gm);
                const uint32_t d = image_mask(D2, p, s.N, full);
                if (d < best) {
                    best = d;
                    bp = p;
                }
                gm &= gm - 1;
            }
            accumulate(s, x, H, bp, blk_self, H, best, 0.5 * sqf, W, acc);
        };
        if (w.active & 4) {  // jl: bonds (i, i + 2)
            const uint32_t d2 = rotl_n(Dn, s.N - 2, s.N, full), u2 = rotl_n(Up, s.N - 2, s.N, full);
            uint32_t anti = ((Dn & u2) | (Up & d2)) & full;
            if (s.N == 2) anti = 0;
            const C2 W = w.w[2];
            const double diag = -0.5 * double(popc32(anti));
            acc = cadd(acc, cmul(W, C2{diag * xf.re, diag * xf.im}));
            for (int i = 0; i < s.N; ++i) {
                if (!((anti >> i) & 1u)) continue;
                const int k = (i + 2) % s.N;
                const bool wrap = k < i;
                exchange(Dn ^ (1u << i) ^ (1u << k), wrap ? k : i, wrap ? i : k, wrap || s.N == 4, W);
            }
        }
        if (w.active & 8) {  // jr: bonds (2x, 2x + 1)
            const uint32_t anti = (Dn ^ (Dn >> 1)) & E & (E >> 1) & EVEN_BITS & full;
            const C2 W = w.w[3];
            const double diag = -0.5 * double(popc32(anti));
            acc = cadd(acc, cmul(W, C2{diag * xf.re, diag * xf.im}));
            for (int i = 0; i < s.N; i += 2) {
                if (!((anti >> i) & 1u)) continue;
                exchange(Dn ^ (3u << i), i, i + 1, false, W);
            }
        }
        if (w.active & 3) {
            uint32_t holes = H;
            while (holes) {
                const int t = ctz32(holes);
                holes &= holes - 1;
                for (int dir = 0; dir < 3; ++dir) {
                    int src;
                    C2 W;
                    if (dir < 2) {
                        if (!(w.active & 1)) continue;
                        src = dir == 0 ? (t + 2) % s.N : (t - 2 + s.N) % s.N;
                        W = w.w[0];
                    } else {
                        if (!(w.active & 2)) continue;
                        src = t ^ 1;
                        W = w.w[1];
                    }
                    if (!(E & (1u << src))) continue;
                    const uint32_t bt = 1u << t, bs = 1u << src;
                    const uint32_t H2 = H ^ bt ^ bs;
                    const uint32_t D2 = (Dn & bs) ? (Dn ^ bs ^ bt) : Dn;
                    const uint32_t above_s = src >= 31 ? 0u : (~0u << (src + 1));
                    const uint32_t above_t = t >= 31 ? 0u : (~0u << (t + 1));
                    const int par = (popc32(E & above_s) + popc32(E & above_t) + (t < src ? 1 : 0)) & 1;
                    const int64_t hr2 = rank_holes(H2, s.comb);
                    const int blk = s.canon_block[hr2];
                    if (blk < 0) continue;
                    uint32_t gm = s.canon_gmask[hr2];
                    int bp = ctz32(gm);
                    const uint32_t Hc = image_mask(H2, bp, s.N, full);
                    uint32_t best = image_mask(D2, bp, s.N, full);
                    gm &= gm - 1;
                    while (gm) {
                        const int p = ctz32(gm);
                        const uint32_t d = image_mask(D2, p, s.N, full);
                        if (d < best) {
                            best = d;
                            bp = p;
                        }
                        gm &= gm - 1;
                    }
                    accumulate(s, x, H2, bp, blk, Hc, best, par ? sqf : -sqf, W, acc);
                }
            }
        }
        store_c(y, f, acc);
        dot = cmul(C2{xf.re, -xf.im}, acc);
    }
    if (partial) {
        workgroup_sum2<BS>(dot.re, dot.im);
        if (threadIdx.x == 0) {
            partial[2 * bid] = dot.re;
            partial[2 * bid + 1] = dot.im;
        }
    }
}

// Szq and N(q) in gather form: final row j has the same representative as
// initial column i; the element is the reference's per-site sum for column i
// (matrix_setup.pyx:235-246), computed with the same operations.
__global__ void diag_op_gather_kernel(SectorView ini, SectorView fin, RuleView r, const double* __restrict__ x,
                                      double* __restrict__ y) {
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= fin.D) return;
    const uint64_t key = fin.key[j];
    const uint32_t H = uint32_t(key >> 32), Dn = uint32_t(key);
    const Found fi = lookup_rep(ini, H, Dn);
    if (fi.j < 0) {
        y[2 * j] = 0.0;
        y[2 * j + 1] = 0.0;
        return;
    }
    const int nf = fin.orbit[j];
    const C2 cc = amp_at(fin, 0, nf);
    const C2 ow0{ini.sqrt_tab[fi.n], 0.0};
    C2 acc{0.0, 0.0};
    bool first = true;
    for (int c = 0; c < r.n_combo; ++c) {
        const int a = r.idx[RULE_W * c];
        C2 ow = cmul(ow0, C2{r.w[4 * c + 2], r.w[4 * c + 3]});
        const int va = site_value(H, Dn, a);
        if (va == 0) continue;
        if (r.spin_proj) {
            C2 f{1.0, 0.0};
            if (va == 1) f = cmul(f, C2{-1.0, 0.0});
            ow = cmul(ow, f);
        }
        if (!(ow.re * ow.re + ow.im * ow.im >= 1e-16)) continue;
        const C2 v = cmul(ow, cc);
        acc = first ? v : cadd(acc, v);
        first = false;
    }
    if (first || acc.re * acc.re + acc.im * acc.im < 1e-16) {
        y[2 * j] = 0.0;
        y[2 * j + 1] = 0.0;
        return;
    }
    const C2 xi{x[2 * fi.j], x[2 * fi.j + 1]};
    const C2 o = cmul(xi, acc);
    y[2 * j] = o.re;
    y[2 * j + 1] = o.im;
}

// <psi|O|psi> = sum_i sum_j conj(psi_j) O_ji psi_i, one column per lane in a
// grid-stride loop; per-workgroup partials (re, im) in a fixed order.
__global__ void __launch_bounds__(128) expect_kernel(SectorView s, RuleView r, const double* __restrict__ psi,
                                                     double* partial) {
    __shared__ double red[2 * 128];
    double re = 0.0, im = 0.0;
    C2 cv[MAXC];
    int64_t ci[MAXC];
    for (int64_t i = int64_t(blockIdx.x) * 128 + threadIdx.x; i < s.D; i += int64_t(gridDim.x) * 128) {
        bool overflow;
        const int m = column_elements(s, s, r, 0, i, cv, ci, overflow);
        const C2 xi{psi[2 * i], psi[2 * i + 1]};
        for (int k = 0; k < m; ++k) {
            const C2 xj{psi[2 * ci[k]], -psi[2 * ci[k] + 1]};
            const C2 v = cmul(cmul(xj, cv[k]), xi);
            re += v.re;
            im += v.im;
        }
    }
    red[2 * threadIdx.x] = re;
    red[2 * threadIdx.x + 1] = im;
    __syncthreads();
    for (int st = 64; st > 0; st >>= 1) {
        if (int(threadIdx.x) < st) {
            red[2 * threadIdx.x] += red[2 * (threadIdx.x + st)];
            red[2 * threadIdx.x + 1] += red[2 * (threadIdx.x + st) + 1];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = red[0];
        partial[2 * blockIdx.x + 1] = red[1];
    }
}

// Shifts of a position list for the correlation kernel: (legs, rungs) with
// legs in [0, L), rungs in {0, 1}.
struct PosShifts {
    int n;
    int leg[8], rung[8];
};

// Bit r of the result = bit (r moved by `leg` legs and `rung` rungs) of X
// (get_index_by_relative_shift on the ladder, symmetry_ladder.h:103-111).
__device__ inline uint32_t moved_mask(uint32_t X, int leg, int rung, int N, uint32_t full) {
    if (rung) X = leg_swap(X);
    const int s = 2 * leg;
    return s ? rotl_n(X, N - s, N, full) : X;
}

// calculate_hole_spin_projection_correlations (position_correlations.pyx:84-127,
// :186-230) in bit-parallel form: for every node index d, the sum over dense
// states of |psi|^2 times the sum over reference sites r of the value at the
// positions (r, r + unit shifts of d, r + each fixed distance).  With X_k the
// state's masks moved by the k-th shift, a whole sum over r is a popcount:
// holes: popc(AND_k H_k); S^z: the product over the positions is +1 or -1 where
// every position holds an electron (A = AND_k E_k) and the parity of the down
// spins is P = XOR_k Dn_k (a position listed twice cancels), so the sum is
// popc(A) - 2 popc(A & P).  NP >= N, a multiple of 8.
template <int NP>
__global__ void __launch_bounds__(256) position_corr_kernel(SectorView s, int kind, PosShifts fixed,
                                                            const double* __restrict__ psi, double* partial) {
    __shared__ double red[256];
    const int N = s.N;
    double acc[NP];
#pragma unroll
    for (int d = 0; d < NP; ++d) acc[d] = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < s.D; i += int64_t(gridDim.x) * 256) {
        const uint64_t key = s.key[i];
        const uint32_t H = uint32_t(key >> 32), Dn = uint32_t(key);
        const uint32_t E = ~H & s.full;
        const double p = psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1];
        uint32_t FA = (kind == ED_POS_HOLES ? H : E), FP = (kind == ED_POS_HOLES ? 0u : Dn);
        for (int k = 0; k < fixed.n; ++k) {
            FA &= moved_mask(kind == ED_POS_HOLES ? H : E, fixed.leg[k], fixed.rung[k], N, s.full);
            if (kind != ED_POS_HOLES) FP ^= moved_mask(Dn, fixed.leg[k], fixed.rung[k], N, s.full);
        }
        // fully unrolled with d a constant, so that acc stays in registers
#pragma unroll
        for (int d = 0; d < NP; ++d) {
            if (d < N) {
                int v;
                if (kind == ED_POS_HOLES) {
                    v = popc32(FA & moved_mask(H, d >> 1, d & 1, N, s.full));
                } else {
                    const uint32_t A = FA & moved_mask(E, d >> 1, d & 1, N, s.full);
                    const uint32_t P = FP ^ moved_mask(Dn, d >> 1, d & 1, N, s.full);
                    v = popc32(A) - 2 * popc32(A & P);
                }
                acc[d] += p * double(v);
            }
        }
    }
#pragma unroll
    for (int d = 0; d < NP; ++d) {
        if (d >= N) continue;  // uniform: N is the same for every lane
        red[threadIdx.x] = acc[d];
        __syncthreads();
        for (int st = 128; st > 0; st >>= 1) {
            if (int(threadIdx.x) < st) red[threadIdx.x] += red[threadIdx.x + st];
            __syncthreads();
        }
        if (threadIdx.x == 0) partial[int64_t(blockIdx.x) * N + d] = red[0];
        __syncthreads();
    }
}

__global__ void scatter_columns_kernel(int64_t D, int max_per_col, const int32_t* count, const int64_t* rows,
                                       const double* vals, const double* __restrict__ x, double* y) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= D) return;
    const C2 xi{x[2 * i], x[2 * i + 1]};
    const int c = count[i] < max_per_col ? count[i] : max_per_col;  // a column beyond capacity is cut
    for (int k = 0; k < c; ++k) {
        const int64_t e = i * max_per_col + k;
        const C2 v = cmul(xi, C2{vals[2 * e], vals[2 * e + 1]});
        atomicAdd(&y[2 * rows[e]], v.re);
        atomicAdd(&y[2 * rows[e] + 1], v.im);
    }
}

__global__ void zero_kernel(double* p, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// Sum of n (re, im) partials in a fixed order: 1024 threads, each with four
// independent 16-byte loads in flight, then a tree.
__global__ void __launch_bounds__(1024) reduce_partials_kernel(const double* partial, int64_t n, double* out) {
    double re, im;
    sum_partials<1024>(partial, n, re, im);
    workgroup_sum2<1024>(re, im);
    if (threadIdx.x == 0) {
        out[0] = re;
        out[1] = im;
    }
}

struct RuleDev {
    RuleView v{};
    int8_t* d_idx = nullptr;
    double* d_w = nullptr;
    RuleDev(const RuleHost& h, hipStream_t st) {
        d_idx = dev_upload(h.idx, st);
        d_w = dev_upload(h.w, st);
        v.n_combo = h.n_combo();
        v.arity = h.arity;
        v.valid = h.valid;
        v.make = h.make;
        v.spin_proj = h.spin_proj;
        v.density = h.density;
        v.fermion = h.fermion;
        v.singlet = h.singlet;
        v.max_per_col = h.max_per_col;
        v.idx = d_idx;
        v.w = d_w;
    }
    ~RuleDev() {
        dev_free(d_idx);
        dev_free(d_w);
    }
};

void run_columns(const ed_sector* ini, const ed_sector* fin, const RuleHost& rule, int triangle, int max_per_col,
                 int32_t* d_count, int64_t* d_rows, double* d_vals, hipStream_t st) {
    ED_REQUIRE(ini && ini->D >= 0 && fin && fin->D >= 0, ED_ERR_STATE, "basis not built");
    launch_columns(ini, fin, rule, triangle, max_per_col, 0, ini->D, d_count, d_rows, d_vals, st);
}

}  // namespace

void launch_columns(const ed_sector* ini, const ed_sector* fin, const RuleHost& rule, int triangle, int max_per_col,
                    int64_t lo, int64_t n, int32_t* d_count, int64_t* d_rows, double* d_vals, hipStream_t st) {
    ED_REQUIRE(ini && ini->D >= 0 && fin && fin->D >= 0, ED_ERR_STATE, "basis not built");
    ED_REQUIRE(ini->L == fin->L, ED_ERR_ARG, "sectors of different ladders");
    ED_REQUIRE(max_per_col >= 1 && max_per_col <= MAXC, ED_ERR_ARG, "max_per_col out of range");
    ED_REQUIRE(lo >= 0 && n >= 0 && lo + n <= ini->D, ED_ERR_ARG, "column range outside the sector");
    if (n == 0) return;
    RuleDev rd(rule, st);
    rd.v.max_per_col = max_per_col;
    const int T = 128;
    hipLaunchKernelGGL(column_kernel, dim3(unsigned((n + T - 1) / T)), dim3(T), 0, st, ini->view(), fin->view(),
                       rd.v, triangle, lo, n, d_count, d_rows, d_vals);
    ED_HIP(hipGetLastError());
    ED_HIP(hipStreamSynchronize(st));
}

int column_capacity() { return MAXC; }

// y = O x from the rule's columns: padded columns, then a scatter of x_i times
// column i (atomic adds into y).
void apply_columns(const ed_sector* ini, const ed_sector* fin, const RuleHost& rule, const double* d_x, double* d_y,
                   hipStream_t st) {
    const int T = 256;
    const int mpc = std::min(rule.max_per_col, MAXC);
    int32_t* cnt = dev_alloc<int32_t>(size_t(std::max<int64_t>(ini->D, 1)));
    int64_t* rows = dev_alloc<int64_t>(size_t(std::max<int64_t>(ini->D, 1) * mpc));
    double* vals = dev_alloc<double>(size_t(std::max<int64_t>(ini->D, 1) * mpc * 2));
    try {
        launch_columns(ini, fin, rule, 0, mpc, 0, ini->D, cnt, rows, vals, st);
        hipLaunchKernelGGL(zero_kernel, dim3(unsigned((2 * fin->D + T - 1) / T)), dim3(T), 0, st, d_y, 2 * fin->D);
        if (ini->D > 0)
            hipLaunchKernelGGL(scatter_columns_kernel, dim3(unsigned((ini->D + T - 1) / T)), dim3(T), 0, st, ini->D,
                               mpc, cnt, rows, vals, d_x, d_y);
        ED_HIP(hipGetLastError());
        ED_HIP(hipStreamSynchronize(st));
    } catch (...) {
        dev_free(cnt);
        dev_free(rows);
        dev_free(vals);
        throw;
    }
    dev_free(cnt);
    dev_free(rows);
    dev_free(vals);
}

// Launch helper shared with the Lanczos driver; returns the number of
// workgroup partials written.  ED_HV_KERNEL selects a kernel for
// measurements: 2 (default) LDS tables + persistent workgroups, 1 lock-step
// bonds, 0 per-lane bond loop; 2 falls back to 1 when the tables exceed LDS.
int64_t launch_hv(const ed_sector* s, const HvWeights& w, const void* x, void* y, double* partial, int precision,
                  int remap, hipStream_t st) {
    constexpr int BS = 256;
    const int64_t nwg = (s->D + BS - 1) / BS;
    if (nwg == 0) return 0;
    static const int variant = [] {
        const char* e = std::getenv("ED_HV_KERNEL");
        return e ? std::atoi(e) : 2;
    }();
    if (variant == 2 || s->flip) {
        const int64_t n = launch_hv_lds(s, w, x, y, partial, precision, st);
        if (n >= 0) return n;
    }
    if (variant == 0) {
        if (precision == 0)
            hipLaunchKernelGGL((hv_kernel<double2, BS>), dim3(unsigned(nwg)), dim3(BS), 0, st, s->view(), w,
                               static_cast<const double2*>(x), static_cast<double2*>(y), partial, remap);
        else
            hipLaunchKernelGGL((hv_kernel<float2, BS>), dim3(unsigned(nwg)), dim3(BS), 0, st, s->view(), w,
                               static_cast<const float2*>(x), static_cast<float2*>(y), partial, remap);
    } else {
        if (precision == 0)
            hipLaunchKernelGGL((hv_lockstep_kernel<double2, BS>), dim3(unsigned(nwg)), dim3(BS), 0, st, s->view(),
                               w, static_cast<const double2*>(x), static_cast<double2*>(y), partial, remap);
        else
            hipLaunchKernelGGL((hv_lockstep_kernel<float2, BS>), dim3(unsigned(nwg)), dim3(BS), 0, st, s->view(),
                               w, static_cast<const float2*>(x), static_cast<float2*>(y), partial, remap);
    }
    ED_HIP(hipGetLastError());
    return nwg;
}

int64_t hv_partials(const ed_sector* s) { return (s->D + 255) / 256 + hv_fix_partials(s); }

void launch_reduce_partials(const double* partial, int64_t n, double* out, hipStream_t st) {
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(1024), 0, st, partial, n, out);
    ED_HIP(hipGetLastError());
}

void require_flip_operator(const ed_sector* ini, const ed_sector* fin, int op) {
    if (!ini->flip && !fin->flip) return;
    const bool keeps = op == ED_OP_NQ || op == ED_OP_CURRENT_RUNG || op == ED_OP_CURRENT_LEG;
    ED_REQUIRE(op == ED_OP_SZQ || keeps, ED_ERR_ARG,
               "between spin-flip sectors only Szq, N(q) and the currents (A+- change S^z)");
    ED_REQUIRE(ini->flip && fin->flip && fin->flip == (keeps ? ini->flip : -ini->flip), ED_ERR_ARG,
               "the final spin-flip parity must be the initial one (N(q), currents) or its opposite (Szq)");
}

void require_no_flip(const ed_sector* s, const char* what) {
    ED_REQUIRE(!s || !s->flip, ED_ERR_ARG, std::string(what) + " is not available for sectors with the spin flip");
}

HvWeights make_weights(const double* w) {
    HvWeights hw{};
    for (int t = 0; t < 4; ++t) {
        hw.w[t] = C2{w[2 * t], w[2 * t + 1]};
        // general/sparse_matrices.py:89: terms with |weight| < 1e-6 are skipped
        if (w[2 * t] * w[2 * t] + w[2 * t + 1] * w[2 * t + 1] >= 1e-12) hw.active |= 1 << t;
    }
    return hw;
}

}  // namespace ed

using namespace ed;

extern "C" {

int ed_term_columns(const ed_sector* s, int term, int triangle, int max_per_col, int32_t* d_count, int64_t* d_rows,
                    double* d_vals, void* stream) {
    return guarded([&] {
        ED_REQUIRE(s, ED_ERR_ARG, "sector is null");
        run_columns(s, s, hamiltonian_rule(s->L, term), triangle, max_per_col, d_count, d_rows, d_vals,
                    static_cast<hipStream_t>(stream));
    });
}

int ed_op_columns(const ed_sector* ini, const ed_sector* fin, int op, int qx, int qy, int max_per_col,
                  int32_t* d_count, int64_t* d_rows, double* d_vals, void* stream) {
    return guarded([&] {
        ED_REQUIRE(ini && fin, ED_ERR_ARG, "sector is null");
        require_flip_operator(ini, fin, op);
        ED_REQUIRE(ini && fin, ED_ERR_ARG, "sector is null");
        run_columns(ini, fin, operator_rule(ini->L, op, qx, qy), 0, max_per_col, d_count, d_rows, d_vals,
                    static_cast<hipStream_t>(stream));
    });
}

int ed_hv(const ed_sector* s, const double* w, const void* d_x, void* d_y, double* d_dot, int precision,
          void* stream) {
    return guarded([&] {
        ED_REQUIRE(s && s->D >= 0, ED_ERR_STATE, "basis not built");
        ED_REQUIRE(precision == 0 || precision == 1, ED_ERR_ARG, "precision must be 0 or 1");
        hipStream_t st = static_cast<hipStream_t>(stream);
        if (s->D == 0) return;
        const HvWeights hw = make_weights(w);
        const int64_t np = launch_hv(s, hw, d_x, d_y, d_dot ? s->d_partial : nullptr, precision, 1, st);
        if (d_dot) launch_reduce_partials(s->d_partial, np, d_dot, st);
    });
}

int ed_op_apply(const ed_sector* ini, const ed_sector* fin, int op, int qx, int qy, const double* d_x, double* d_y,
                void* stream) {
    return guarded([&] {
        ED_REQUIRE(ini && fin, ED_ERR_ARG, "sector is null");
        require_flip_operator(ini, fin, op);
        ED_REQUIRE(ini && fin && ini->D >= 0 && fin->D >= 0, ED_ERR_STATE, "basis not built");
        hipStream_t st = static_cast<hipStream_t>(stream);
        const RuleHost rule = operator_rule(ini->L, op, qx, qy);
        const int T = 256;
        if (fin->D == 0) return;
        if (op == ED_OP_SZQ || op == ED_OP_NQ) {
            ED_REQUIRE(ini->h == fin->h && ini->nd == fin->nd, ED_ERR_ARG, "Szq/N(q) keep holes and spins");
            RuleDev rd(rule, st);
            hipLaunchKernelGGL(diag_op_gather_kernel, dim3(unsigned((fin->D + T - 1) / T)), dim3(T), 0, st,
                               ini->view(), fin->view(), rd.v, d_x, d_y);
            ED_HIP(hipGetLastError());
            ED_HIP(hipStreamSynchronize(st));
            return;
        }
        apply_columns(ini, fin, rule, d_x, d_y, st);
    });
}

int ed_position_op_columns(const ed_sector* s, int kind, const int* shifts, int n_shifts, int max_per_col,
                           int32_t* d_count, int64_t* d_rows, double* d_vals, void* stream) {
    return guarded([&] {
        require_no_flip(s, "position correlations");
        ED_REQUIRE(s && (shifts || n_shifts == 0), ED_ERR_ARG, "null argument");
        run_columns(s, s, position_rule(s->L, kind, shifts, n_shifts), 0, max_per_col, d_count, d_rows, d_vals,
                    static_cast<hipStream_t>(stream));
    });
}

int ed_position_op_apply(const ed_sector* s, int kind, const int* shifts, int n_shifts, const double* d_x,
                         double* d_y, void* stream) {
    return guarded([&] {
        require_no_flip(s, "position correlations");
        ED_REQUIRE(s && s->D >= 0, ED_ERR_STATE, "basis not built");
        ED_REQUIRE(shifts || n_shifts == 0, ED_ERR_ARG, "null shifts");
        if (s->D == 0) return;
        apply_columns(s, s, position_rule(s->L, kind, shifts, n_shifts), d_x, d_y, static_cast<hipStream_t>(stream));
    });
}

int ed_position_expectation(const ed_sector* s, int kind, const int* shifts, int n_shifts, const double* d_psi,
                            double* out, void* stream) {
    return guarded([&] {
        require_no_flip(s, "position correlations");
        ED_REQUIRE(s && s->D >= 0, ED_ERR_STATE, "basis not built");
        ED_REQUIRE(out && (shifts || n_shifts == 0), ED_ERR_ARG, "null argument");
        out[0] = out[1] = 0.0;
        if (s->D == 0) return;
        hipStream_t st = static_cast<hipStream_t>(stream);
        const RuleHost rule = position_rule(s->L, kind, shifts, n_shifts);
        ED_REQUIRE(rule.n_combo() + 1 <= MAXC, ED_ERR_CAPACITY, "too many combinations per column");
        RuleDev rd(rule, st);
        const int64_t nwg = std::min<int64_t>((s->D + 127) / 128, 4096);
        double* partial = dev_alloc<double>(size_t(2 * nwg + 2));
        try {
            hipLaunchKernelGGL(expect_kernel, dim3(unsigned(nwg)), dim3(128), 0, st, s->view(), rd.v, d_psi, partial);
            launch_reduce_partials(partial, nwg, partial + 2 * nwg, st);
            ED_HIP(hipMemcpyAsync(out, partial + 2 * nwg, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
            ED_HIP(hipStreamSynchronize(st));
        } catch (...) {
            dev_free(partial);
            throw;
        }
        dev_free(partial);
    });
}

int ed_position_correlations(const ed_sector* s, int kind, const int* fixed, int n_fixed, const double* d_psi,
                             double* out, void* stream) {
    return guarded([&] {
        require_no_flip(s, "position correlations");
        ED_REQUIRE(s && s->D >= 0, ED_ERR_STATE, "basis not built");
        ED_REQUIRE(out && (fixed || n_fixed == 0), ED_ERR_ARG, "null argument");
        ED_REQUIRE(kind == ED_POS_HOLES || kind == ED_POS_SZ, ED_ERR_ARG, "kind must be holes or S^z");
        ED_REQUIRE(n_fixed >= 0 && n_fixed <= 8, ED_ERR_ARG, "at most 8 fixed distances");
        const int N = s->N;
        PosShifts fx{};
        fx.n = n_fixed;
        for (int k = 0; k < n_fixed; ++k) {
            fx.leg[k] = ((fixed[2 * k] % s->L) + s->L) % s->L;
            fx.rung[k] = ((fixed[2 * k + 1] % 2) + 2) % 2;
        }
        for (int d = 0; d < N; ++d) out[d] = 0.0;
        if (s->D == 0) return;
        hipStream_t st = static_cast<hipStream_t>(stream);
        const int64_t nwg = std::min<int64_t>((s->D + 255) / 256, 2048);
        double* partial = dev_alloc<double>(size_t(nwg * N));
        try {
            if (N <= 8)
                hipLaunchKernelGGL(position_corr_kernel<8>, dim3(unsigned(nwg)), dim3(256), 0, st, s->view(), kind, fx,
                                   d_psi, partial);
            else if (N <= 16)
                hipLaunchKernelGGL(position_corr_kernel<16>, dim3(unsigned(nwg)), dim3(256), 0, st, s->view(), kind,
                                   fx, d_psi, partial);
            else if (N <= 24)
                hipLaunchKernelGGL(position_corr_kernel<24>, dim3(unsigned(nwg)), dim3(256), 0, st, s->view(), kind,
                                   fx, d_psi, partial);
            else
                hipLaunchKernelGGL(position_corr_kernel<32>, dim3(unsigned(nwg)), dim3(256), 0, st, s->view(), kind,
                                   fx, d_psi, partial);
            ED_HIP(hipGetLastError());
            std::vector<double> h(size_t(nwg * N));
            ED_HIP(hipMemcpyAsync(h.data(), partial, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
            ED_HIP(hipStreamSynchronize(st));
            for (int64_t b = 0; b < nwg; ++b)
                for (int d = 0; d < N; ++d) out[d] += h[size_t(b * N + d)];
        } catch (...) {
            dev_free(partial);
            throw;
        }
        dev_free(partial);
    });
}

}  // extern "C"
