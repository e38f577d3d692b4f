// Sector set-up and symmetric basis enumeration for gfx950.
//
// Host part: binomial table, amplitude tables and the hole-pattern tables
// (eligibility prefilter, minimising group elements, blocks).  Device part:
// candidate scan (one lane per spin configuration of an eligible hole
// pattern) -> per-chunk counts -> host prefix -> fill of the dense arrays.

#include <algorithm>
#include <cmath>
#include <complex>
#include <numbers>
#include <numeric>

#include "ed_internal.h"
#include "ed_matrix.h"
#include "ed_orbit.h"

namespace ed {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 4;
constexpr int64_t CHUNK = SCAN_THREADS * SCAN_ITEMS;

uint32_t unrank_holes_host(int64_t r, int h, int N, const std::vector<uint32_t>& comb) {
    uint32_t H = 0;
    int rem = h;
    for (int j = N - 1; j >= 0 && rem > 0; --j) {
        const int64_t c = comb[rem * COMB_W + j];
        if (r >= c) {
            H |= 1u << j;
            r -= c;
            --rem;
        }
    }
    return H;
}

// Candidate scan: flag[sr] = orbit size (0 = not a basis state) for every
// spin configuration of one eligible hole pattern (grid.y = block).
__global__ void __launch_bounds__(SCAN_THREADS) basis_count_kernel(SectorView s, uint8_t* flags, int32_t* counts,
                                                                    int64_t chunks_per_block) {
    const int blk = blockIdx.y;
    const int64_t chunk = blockIdx.x;
    const uint32_t H = s.block_hole[blk];
    int c = 0;
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const int64_t sr = chunk * CHUNK + int64_t(threadIdx.x) * SCAN_ITEMS + k;
        if (sr >= s.n_spin) break;
        const uint32_t Dn = unrank_spins(H, sr, s.nd, s.N, s.comb);
        const OrbitResult r = orbit_walk(s, H, Dn, true);
        flags[int64_t(blk) * s.n_spin + sr] = uint8_t(r.n);
        c += r.n > 0;
    }
    __shared__ int red[SCAN_THREADS];
    red[threadIdx.x] = c;
    __syncthreads();
    for (int w = SCAN_THREADS / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) counts[int64_t(blk) * chunks_per_block + chunk] = red[0];
}

// Fill: dense index = chunk offset + exclusive prefix inside the chunk.
__global__ void __launch_bounds__(SCAN_THREADS) basis_fill_kernel(SectorView s, const uint8_t* flags,
                                                                   const int64_t* offsets, int64_t chunks_per_block,
                                                                   uint64_t* key, uint8_t* orbit, int32_t* table) {
    const int blk = blockIdx.y;
    const int64_t chunk = blockIdx.x;
    const uint32_t H = s.block_hole[blk];
    const int64_t tab = s.block_table[blk];
    uint8_t f[SCAN_ITEMS];
    int c = 0;
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const int64_t sr = chunk * CHUNK + int64_t(threadIdx.x) * SCAN_ITEMS + k;
        f[k] = sr < s.n_spin ? flags[int64_t(blk) * s.n_spin + sr] : 0;
        c += f[k] > 0;
    }
    __shared__ int pre[SCAN_THREADS];
    pre[threadIdx.x] = c;
    __syncthreads();
    // Hillis-Steele inclusive scan over the workgroup
    for (int w = 1; w < SCAN_THREADS; w <<= 1) {
        const int v = int(threadIdx.x) >= w ? pre[threadIdx.x - w] : 0;
        __syncthreads();
        pre[threadIdx.x] += v;
        __syncthreads();
    }
    int64_t d = offsets[int64_t(blk) * chunks_per_block + chunk] + pre[threadIdx.x] - c;
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const int64_t sr = chunk * CHUNK + int64_t(threadIdx.x) * SCAN_ITEMS + k;
        if (sr >= s.n_spin) break;
        if (f[k]) {
            const uint32_t Dn = unrank_spins(H, sr, s.nd, s.N, s.comb);
            key[d] = make_key(H, Dn);
            orbit[d] = f[k];
            if (tab >= 0) table[tab + sr] = int32_t(d);
            ++d;
        } else if (tab >= 0) {
            table[tab + sr] = -1;
        }
    }
}

// Orbit size beside each table entry (0 where the spin rank is no basis
// state), so that a lookup through the table needs no second, dependent load.
__global__ void table_orbit_kernel(const int32_t* table, const uint8_t* orbit, uint8_t* out, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t j = table[i];
    out[i] = j >= 0 ? orbit[j] : uint8_t(0);
}

__global__ void basis_export_kernel(SectorView s, int64_t* sparse, int32_t* orb) {
    const int64_t f = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (f >= s.D) return;
    const uint64_t k = s.key[f];
    const uint32_t H = uint32_t(k >> 32), Dn = uint32_t(k);
    if (sparse) sparse[f] = rank_holes(H, s.comb) * s.n_spin + rank_spins(H, Dn, s.comb);
    if (orb) orb[f] = s.orbit[f];
}

__device__ inline uint32_t unrank_holes_dev(int64_t r, int h, int N, const uint32_t* comb) {
    uint32_t H = 0;
    int rem = h;
    for (int j = N - 1; j >= 0 && rem > 0; --j) {
        const int64_t c = comb[rem * COMB_W + j];
        if (r >= c) {
            H |= 1u << j;
            r -= c;
            --rem;
        }
    }
    return H;
}

__global__ void lookup_kernel(SectorView s, const int64_t* sparse, int64_t* dense, int64_t n) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t sp = sparse[t];
    int64_t out = -1;
    if (sp >= 0) {
        const int64_t hr = sp / s.n_spin, sr = sp - hr * s.n_spin;
        if (hr < s.n_hole_ranks) {
            const int blk = s.block_of_rank[hr];
            if (blk >= 0) {
                const int64_t off = s.block_table[blk];
                out = off < 0 ? s.block_base[blk] + sr : int64_t(s.table[off + sr]);
            }
        }
    }
    dense[t] = out;
}

__global__ void lowest_kernel(SectorView s, const int64_t* sparse, int64_t* rep, double* amp, int64_t n) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t sp = sparse[t];
    const int64_t hr = sp / s.n_spin, sr = sp - hr * s.n_spin;
    const uint32_t H = unrank_holes_dev(hr, s.h, s.N, s.comb);
    const uint32_t Dn = unrank_spins(H, sr, s.nd, s.N, s.comb);
    const OrbitResult r = orbit_walk(s, H, Dn, false);
    if (r.n == 0) {
        rep[t] = -1;
        amp[2 * t] = 0.0;
        amp[2 * t + 1] = 0.0;
        return;
    }
    const uint32_t Hb = uint32_t(r.best >> 32), Db = uint32_t(r.best);
    rep[t] = rank_holes(Hb, s.comb) * s.n_spin + rank_spins(Hb, Db, s.comb);
    const C2 c = lowest_amplitude(s, H, r.best_p, r.n);
    amp[2 * t] = c.re;
    amp[2 * t + 1] = c.im;
}

}  // namespace

// ---------------------------------------------------------------------------
// Rules (host): restatement of tJ_spin_half_ladder/matrix_setup.pyx:80-118 and
// :235-325 with general/matrix_rules_utils.pyx:15-148.

namespace {

int py_mod(int a, int b) {
    int r = a % b;
    return r + b * (r < 0);
}

int shifted_index(int N, int i, int s0, int s1) {  // symmetry_ladder.h:96-101
    i = py_mod(i, N);
    const int leg = i / 2 + s0;
    const int rung = py_mod(i + (s1 % 2), 2);
    return py_mod(2 * leg + rung, N);
}

void add_combo(RuleHost& r, int a, int b, std::complex<double> w0, std::complex<double> w1) {
    r.idx.push_back(int8_t(a));
    r.idx.push_back(int8_t(b));
    for (int k = 2; k < RULE_W; ++k) r.idx.push_back(0);
    r.w.push_back(w0.real());
    r.w.push_back(w0.imag());
    r.w.push_back(w1.real());
    r.w.push_back(w1.imag());
}

void pair_combos(RuleHost& r, int N, const std::vector<std::pair<int, int>>& shifts, int validator, double w0,
                 double w1) {
    std::vector<std::pair<int, int>> seen;
    for (int i = 0; i < N; ++i)
        for (auto sh : shifts) {
            std::pair<int, int> pr{i, shifted_index(N, i, sh.first, sh.second)};
            if (validator) {
                bool dup = std::find(seen.begin(), seen.end(), pr) != seen.end();
                if (validator == 2)
                    dup = dup || std::find(seen.begin(), seen.end(), std::make_pair(pr.second, pr.first)) != seen.end();
                if (dup) continue;
                seen.push_back(pr);
            }
            add_combo(r, pr.second, pr.first, w0, w1);
        }
}

// general/group/symmetry_generator.h:87-102 (single-precision per-axis term).
double q_dot_r(int site, int qx, int qy, int L) {
    double acc = 0.0;
    acc += double(float(qx * (site / 2)) / float(L));
    acc += double(float(qy * (site % 2)) / float(2));
    return acc * 2.0 * std::numbers::pi;
}

}  // namespace

RuleHost hamiltonian_rule(int L, int term) {
    const int N = 2 * L;
    RuleHost r;
    r.arity = 2;
    r.make = RM_SWAP;
    if (term == 0) pair_combos(r, N, {{1, 0}, {-1, 0}}, 0, 0.0, -1.0);
    else if (term == 1) pair_combos(r, N, {{0, 1}, {0, -1}}, 1, 0.0, -1.0);
    else if (term == 2) pair_combos(r, N, {{1, 0}}, 0, -0.5, 0.5);
    else if (term == 3) pair_combos(r, N, {{0, 1}}, 2, -0.5, 0.5);
    else throw Fail{ED_ERR_ARG, "term must be 0..3"};
    r.valid = term < 2 ? RV_HOP : RV_SWAP;
    r.fermion = term < 2;
    r.max_per_col = r.n_combo() + 1;
    return r;
}

RuleHost operator_rule(int L, int op, int qx, int qy) {
    using cd = std::complex<double>;
    const int N = 2 * L;
    RuleHost r;
    auto iqr = [&](int site, cd b0, cd b1) {
        const cd e = std::exp(cd(0.0, -q_dot_r(site, qx, qy, L)));
        return std::make_pair(b0 * e, b1 * e);
    };
    if (op == ED_OP_SZQ || op == ED_OP_NQ) {
        const double b = (op == ED_OP_SZQ ? 0.5 : 1.0) / std::sqrt(double(N));
        for (int i = 0; i < N; ++i) {
            auto w = iqr(i, 0.0, b);
            add_combo(r, i, 0, w.first, w.second);
        }
        r.arity = 1;
        r.valid = RV_NO_HOLES;
        r.make = RM_SAME;
        r.spin_proj = op == ED_OP_SZQ;
        r.density = op == ED_OP_NQ;
        r.max_per_col = 1;
    } else if (op >= ED_OP_APLUS_DOWN && op <= ED_OP_AMINUS_UP) {
        const double b = 1.0 / std::sqrt(double(N));
        for (int i = 0; i < N; ++i) {
            auto w = iqr(i, 0.0, b);
            add_combo(r, i, 0, w.first, w.second);
        }
        r.arity = 1;
        r.fermion = 1;
        if (op == ED_OP_APLUS_DOWN || op == ED_OP_APLUS_UP) {
            r.valid = RV_ALL_HOLES;
            r.make = op == ED_OP_APLUS_DOWN ? RM_ADD_DOWN : RM_ADD_UP;
        } else {
            r.valid = op == ED_OP_AMINUS_DOWN ? RV_ALL_DOWN : RV_ALL_UP;
            r.make = RM_ADD_HOLE;
        }
        r.max_per_col = N;
    } else if (op == ED_OP_CURRENT_RUNG) {
        for (int i = 0; i < N; ++i) {
            auto sw = iqr(i, 0.0, cd(0, 1));
            add_combo(r, shifted_index(N, i, 0, 1), i, sw.first * cd(0), sw.second * cd(1));
        }
        r.arity = 2;
        r.valid = RV_HOP;
        r.make = RM_SWAP;
        r.max_per_col = N;
    } else if (op == ED_OP_CURRENT_LEG) {
        const int sh[2] = {1, -1};
        const double ew[2] = {1, -1};
        for (int i = 0; i < N; ++i) {
            auto sw = iqr(i, 0.0, cd(0, 1));
            for (int k = 0; k < 2; ++k)
                add_combo(r, shifted_index(N, i, sh[k], 0), i, sw.first * cd(0), sw.second * cd(ew[k]));
        }
        r.arity = 2;
        r.valid = RV_HOP;
        r.make = RM_SWAP;
        r.fermion = 1;
        r.max_per_col = 2 * N;
    } else {
        throw Fail{ED_ERR_ARG, "unknown operator id"};
    }
    return r;
}

int shifted_site(int N, int i, int leg, int rung) { return shifted_index(N, i, leg, rung); }

// get_position_correlation_operator (tJ_spin_half_ladder/matrix_setup.pyx:409-480):
// combinations start at every site (combinations_from_start_end_points with
// reverse=False).  Holes / S^z: (first, first + each shift), weights (1, 0),
// diagonal only.  Singlet-singlet, shifts = (y, x, j) for
// Delta_i^dagger(y) Delta_{i+j}(x): end points (y, j, x + j) and
// (-y, j - y, x + j - y), weights (0, 0.5 / N).
RuleHost position_rule(int L, int kind, const int* sh, int n_shifts) {
    const int N = 2 * L;
    RuleHost r;
    ED_REQUIRE(n_shifts >= 0 && (kind != ED_POS_SINGLET || n_shifts == 3), ED_ERR_ARG,
               "singlet-singlet needs two fixed distances and the free shift");
    ED_REQUIRE(kind == ED_POS_SINGLET || n_shifts + 1 <= RULE_W, ED_ERR_ARG,
               "at most three shifts for hole or S^z correlations");
    ED_REQUIRE(kind >= ED_POS_HOLES && kind <= ED_POS_SINGLET, ED_ERR_ARG, "unknown position-correlation kind");
    auto push = [&](const int* c, int n, std::complex<double> w0, std::complex<double> w1) {
        for (int k = 0; k < RULE_W; ++k) r.idx.push_back(int8_t(k < n ? c[k] : 0));
        r.w.insert(r.w.end(), {w0.real(), w0.imag(), w1.real(), w1.imag()});
    };
    if (kind == ED_POS_SINGLET) {
        const int y0 = sh[0], y1 = sh[1], x0 = sh[2], x1 = sh[3], j0 = sh[4], j1 = sh[5];
        const int ends[2][3][2] = {{{y0, y1}, {j0, j1}, {x0 + j0, x1 + j1}},
                                   {{-y0, -y1}, {j0 - y0, j1 - y1}, {x0 + j0 - y0, x1 + j1 - y1}}};
        for (int i = 0; i < N; ++i)
            for (const auto& e : ends) {
                int c[4] = {i, 0, 0, 0};
                for (int k = 0; k < 3; ++k) c[k + 1] = shifted_index(N, i, e[k][0], e[k][1]);
                push(c, 4, 0.0, 0.5 / N);
            }
        r.arity = 4;
        r.valid = RV_UNEQUAL_SPINS_2HOLES;
        r.make = RM_REMOVE2_ADD_UP_DOWN;
        r.fermion = 1;
        r.singlet = 1;
        r.max_per_col = r.n_combo() + 1;
    } else {
        for (int i = 0; i < N; ++i) {
            int c[RULE_W] = {i, 0, 0, 0};
            for (int k = 0; k < n_shifts; ++k) c[k + 1] = shifted_index(N, i, sh[2 * k], sh[2 * k + 1]);
            push(c, n_shifts + 1, 1.0, 0.0);
        }
        r.arity = n_shifts + 1;
        r.valid = kind == ED_POS_HOLES ? RV_ALL_HOLES : RV_NO_HOLES;
        r.make = RM_SAME;
        r.spin_proj = kind == ED_POS_SZ;
        r.max_per_col = 1;
    }
    return r;
}

}  // namespace ed

using namespace ed;

ed::SectorView ed_sector::host_view() const {
    SectorView v{};
    v.L = L;
    v.N = N;
    v.h = h;
    v.ne = ne;
    v.nd = nd;
    v.ne_even = (ne % 2) == 0;
    v.full = N == 32 ? 0xffffffffu : ((1u << N) - 1u);
    v.D = D;
    v.n_spin = n_spin;
    v.n_hole_ranks = int(n_hole);
    v.n_blocks = int(block_rank.size());
    v.any_k = any_k;
    v.flip = flip;
    v.G = flip ? 2 * N : N;
    v.NO = v.G;
    v.comb = comb.data();
    v.amp = amp.data();
    v.sqrt_tab = sqrt_tab.data();
    v.rank_tab = rank_tab.data();
    return v;
}

ed::SectorView ed_sector::view() const {
    SectorView v = host_view();
    v.comb = d_comb;
    v.canon_gmask = d_canon_gmask;
    v.canon_block = d_canon_block;
    v.block_of_rank = d_block_of_rank;
    v.block_base = d_block_base;
    v.block_table = d_block_table;
    v.block_hole = d_block_hole;
    v.table = d_table;
    v.key = d_key;
    v.orbit = d_orbit;
    v.amp = d_amp;
    v.sqrt_tab = d_sqrt;
    v.rank_tab = d_rank_tab;
    v.table_orbit = d_table_orbit;
    return v;
}

void ed_sector::release() {
    for (void* p : {(void*)d_comb, (void*)d_amp, (void*)d_sqrt, (void*)d_canon_gmask, (void*)d_canon_block,
                    (void*)d_block_of_rank, (void*)d_block_base, (void*)d_block_table, (void*)d_block_hole,
                    (void*)d_table, (void*)d_key, (void*)d_orbit, (void*)d_partial, (void*)d_rank_tab, (void*)d_table_orbit,
                    (void*)d_fix})
        dev_free(p);
    d_fix = nullptr;
    n_fix = 0;
    d_comb = nullptr;
    d_amp = d_sqrt = nullptr;
    d_rank_tab = nullptr;
    d_table_orbit = nullptr;
    d_canon_gmask = nullptr;
    d_canon_block = d_block_of_rank = nullptr;
    d_block_base = d_block_table = nullptr;
    d_block_hole = nullptr;
    d_table = nullptr;
    d_key = nullptr;
    d_orbit = nullptr;
    d_partial = nullptr;
}

extern "C" {

const char* ed_last_error(void) { return ed::g_last_error.c_str(); }

int ed_version(void) { return 1; }

int ed_cache_release(void) {
    return guarded([&] {
        int dev = 0;
        ED_HIP(hipGetDevice(&dev));
        dev_cache_release(dev);
    });
}

}  // extern "C"

namespace ed {
namespace {

int sector_create(int L, int h, int nd, int kx, int ky, int any_k, int flip, ed_sector** out) {
    return guarded([&] {
        ED_REQUIRE(out != nullptr, ED_ERR_ARG, "out is null");
        ED_REQUIRE(L >= 1 && 2 * L <= MAX_N, ED_ERR_ARG, "num_rungs must be in 1..16");
        const int N = 2 * L;
        ED_REQUIRE(h >= 0 && h <= N, ED_ERR_ARG, "num_holes out of range");
        ED_REQUIRE(nd >= 0 && nd <= N - h, ED_ERR_ARG, "num_down_spins out of range");
        ED_REQUIRE(ky == 0 || ky == 1, ED_ERR_ARG, "rung quantum number must be 0 or 1");
        ED_REQUIRE(flip == 0 || flip == 1 || flip == -1, ED_ERR_ARG, "spin-flip parity must be +1 or -1 (0: none)");
        ED_REQUIRE(!flip || (2 * nd == N - h && nd > 0), ED_ERR_ARG,
                   "the spin flip needs S^z = 0 (as many down as up spins) and at least one electron");
        ED_REQUIRE(!(flip && any_k), ED_ERR_ARG, "no spin flip in the union of momentum sectors");
        auto* s = new ed_sector;
        s->flip = flip;
        s->L = L;
        s->N = N;
        s->h = h;
        s->ne = N - h;
        s->nd = nd;
        s->kx = kx;
        s->ky = ky;
        s->any_k = any_k;
        // binomials C(n, k), exact integers (the reference's table,
        // general/matrix_rules_utils.pyx:122-134, holds the same values)
        s->comb.assign(COMB_W * COMB_W, 0);
        for (int n = 0; n < COMB_W; ++n) {
            s->comb[0 * COMB_W + n] = 1;
            for (int k = 1; k <= n; ++k)
                s->comb[k * COMB_W + n] = s->comb[(k - 1) * COMB_W + n - 1] + (k <= n - 1 ? s->comb[k * COMB_W + n - 1] : 0);
        }
        s->n_spin = s->comb[nd * COMB_W + s->ne];
        s->n_hole = s->comb[h * COMB_W + N];
        ED_REQUIRE(s->n_hole <= (int64_t(1) << 24), ED_ERR_ARG, "too many hole patterns for the pattern tables");
        // amplitude tables, tJ_spin_half_ladder/symmetry_ladder.h:134-175
        using cd = std::complex<double>;
        const cd i2pi(0.0, 2.0 * std::numbers::pi);
        const cd mirror_exp = std::exp(-i2pi * cd(double(ky) / double(2)));
        const cd rung_exp = std::exp(-i2pi * cd(double(kx) / double(L)));
        std::vector<cd> E(L + 1);
        E[0] = 1;
        for (int m = 1; m <= L; ++m) E[m] = E[m - 1] * rung_exp;
        // with the spin flip: elements p + N (the flip after p) with character z F[p], orbits up to 2N
        const int G = flip ? 2 * N : N, NO = G;
        s->amp.assign(size_t(G) * (NO + 1) * 2, 0.0);
        s->sqrt_tab.assign(NO + 1, 0.0);
        for (int n = 0; n <= NO; ++n) s->sqrt_tab[n] = std::sqrt(double(n));
        {
            const RankGeom g = rank_geom(s->ne, nd);
            s->rank_tab.assign(size_t(g.entries), 0u);
            for (int c = 0; c < g.n; ++c) {
                const int kmax = std::min(8 * c, nd);
                for (int kb = 0; kb <= kmax; ++kb)
                    for (uint32_t bits = 0; bits < (1u << g.width[c]); ++bits) {
                        uint32_t r = 0;
                        int j = 0;
                        for (int i = 0; i < g.width[c]; ++i)
                            if ((bits >> i) & 1u) {
                                const int k = kb + ++j;
                                r += k <= COMB_W - 1 ? s->comb[k * COMB_W + 8 * c + i] : 0u;
                            }
                        s->rank_tab[size_t(g.base[c] + (kb << g.width[c]) + int(bits))] = r;
                    }
            }
        }
        for (int f = 0; f < (flip ? 2 : 1); ++f)
            for (int m = 0; m < L; ++m)
                for (int b = 0; b < 2; ++b) {
                    cd F = b ? E[m] * mirror_exp : E[m];
                    if (f && flip < 0) F = -F;
                    for (int n = 1; n <= NO; ++n) {
                        cd a = F;
                        a /= std::sqrt(n);
                        const size_t idx = (size_t(f) * N + size_t(m) * 2 + b) * (NO + 1) + n;
                        s->amp[2 * idx] = a.real();
                        s->amp[2 * idx + 1] = a.imag();
                    }
                }
        // hole patterns: eligibility prefilter (symmetry_ladder.h:363-405),
        // minimising group elements, blocks
        SectorView hv = s->host_view();
        s->hole_by_rank.resize(s->n_hole);
        s->canon_gmask.resize(s->n_hole);
        s->canon_block.assign(s->n_hole, -1);
        s->block_of_rank.assign(s->n_hole, -1);
        std::vector<int64_t> canon_rank(s->n_hole);
        for (int64_t r = 0; r < s->n_hole; ++r) {
            const uint32_t H = unrank_holes_host(r, h, N, s->comb);
            s->hole_by_rank[r] = H;
            uint32_t best = 0xffffffffu, gm = 0;
            for (int p = 0; p < N; ++p) {
                const uint32_t im = image_mask(H, p, N, hv.full);
                if (im < best) {
                    best = im;
                    gm = 1u << p;
                } else if (im == best) {
                    gm |= 1u << p;
                }
            }
            s->canon_gmask[r] = gm;
            canon_rank[r] = rank_holes(best, s->comb.data());
            bool eligible;
            if (flip) {
                eligible = best == H;  // the pattern is the minimum of its orbit; states are tested one by one
            } else if (h <= 1 || h >= N - 1) {
                eligible = r == 0;
            } else {
                const uint32_t free = ~H & hv.full;
                const uint32_t Dn = free & (~free + 1u);  // first non-hole site holds the one down spin
                eligible = orbit_walk(hv, H, Dn, true).n > 0;
            }
            if (eligible) {
                s->block_of_rank[r] = int32_t(s->block_rank.size());
                s->block_rank.push_back(r);
                s->block_hole.push_back(H);
                bool sym = false;
                for (int p = 1; p < N; ++p) sym = sym || image_mask(H, p, N, hv.full) == H;
                s->block_sym.push_back(sym ? 1 : 0);
            }
        }
        for (int64_t r = 0; r < s->n_hole; ++r) s->canon_block[r] = s->block_of_rank[canon_rank[r]];
        const size_t nb = s->block_rank.size();
        s->block_base.assign(nb, 0);
        s->block_table.assign(nb, -1);
        int64_t tab = 0;
        for (size_t b = 0; b < nb; ++b)
            if (s->block_sym[b]) {
                s->block_table[b] = tab;
                tab += s->n_spin;
            }
        s->table_size = tab;
        *out = s;
    });
}

}  // namespace
}  // namespace ed

extern "C" {

int ed_sector_create(int L, int h, int nd, int kx, int ky, ed_sector** out) {
    return sector_create(L, h, nd, kx, ky, 0, 0, out);
}

int ed_sector_create_flip(int L, int h, int nd, int kx, int ky, int z, ed_sector** out) {
    return sector_create(L, h, nd, kx, ky, 0, z, out);
}

int ed_sector_flip(const ed_sector* s) { return s ? s->flip : 0; }

int ed_sector_create_union(int L, int h, int nd, ed_sector** out) { return sector_create(L, h, nd, 0, 0, 1, 0, out); }

int ed_sector_destroy(ed_sector* s) {
    return guarded([&] {
        if (!s) return;
        s->release();
        delete s;
    });
}

int ed_basis_build(ed_sector* s, void* stream) {
    return guarded([&] {
        ED_REQUIRE(s != nullptr, ED_ERR_ARG, "sector is null");
        hipStream_t st = static_cast<hipStream_t>(stream);
        s->release();
        s->d_comb = dev_upload(s->comb, st);
        s->d_amp = dev_upload(s->amp, st);
        s->d_sqrt = dev_upload(s->sqrt_tab, st);
        s->d_rank_tab = dev_upload(s->rank_tab, st);
        s->d_canon_gmask = dev_upload(s->canon_gmask, st);
        s->d_canon_block = dev_upload(s->canon_block, st);
        s->d_block_of_rank = dev_upload(s->block_of_rank, st);
        s->d_block_table = dev_upload(s->block_table, st);
        s->d_block_hole = dev_upload(s->block_hole, st);
        const int64_t nb = int64_t(s->block_rank.size());
        const int64_t cpb = (s->n_spin + CHUNK - 1) / CHUNK;
        if (nb == 0 || s->n_spin == 0) {
            s->D = 0;
            s->d_block_base = dev_upload(s->block_base, st);
            s->d_key = dev_alloc<uint64_t>(1);
            s->d_orbit = dev_alloc<uint8_t>(1);
            s->d_table = dev_alloc<int32_t>(1);
            s->d_table_orbit = dev_alloc<uint8_t>(1);
            ED_HIP(hipStreamSynchronize(st));
            return;
        }
        ED_REQUIRE(cpb <= 0x7fffffff && nb <= 65535, ED_ERR_ARG, "sector too large for the scan grid");
        uint8_t* flags = dev_alloc<uint8_t>(size_t(nb * s->n_spin));
        int32_t* counts = dev_alloc<int32_t>(size_t(nb * cpb));
        SectorView v = s->view();
        const dim3 grid{unsigned(cpb), unsigned(nb), 1u};
        hipLaunchKernelGGL(basis_count_kernel, grid, dim3(SCAN_THREADS), 0, st, v, flags, counts, cpb);
        ED_HIP(hipGetLastError());
        std::vector<int32_t> hc(nb * cpb);
        ED_HIP(hipMemcpyAsync(hc.data(), counts, hc.size() * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        ED_HIP(hipStreamSynchronize(st));
        std::vector<int64_t> off(nb * cpb);
        int64_t acc = 0;
        for (int64_t b = 0; b < nb; ++b) {
            s->block_base[b] = acc;
            int64_t in_block = 0;
            for (int64_t c = 0; c < cpb; ++c) {
                off[b * cpb + c] = acc;
                acc += hc[b * cpb + c];
                in_block += hc[b * cpb + c];
            }
            // a pattern without stabiliser keeps every spin configuration; with the spin flip those whose top
            // electron is up (the lower of each flipped pair), the first C(ne - 1, nd) spin ranks
            const int64_t generic = s->flip ? int64_t(s->comb[s->nd * COMB_W + s->ne - 1]) : s->n_spin;
            if (!s->block_sym[b] && in_block != generic) {
                dev_free(flags);
                dev_free(counts);
                throw Fail{ED_ERR_STATE, "hole pattern without stabiliser lost spin configurations"};
            }
        }
        s->D = acc;
        ED_REQUIRE(s->D < (int64_t(1) << 31), ED_ERR_ARG, "basis larger than 2^31 states");
        s->d_block_base = dev_upload(s->block_base, st);
        int64_t* d_off = dev_upload(off, st);
        s->d_key = dev_alloc<uint64_t>(size_t(s->D));
        s->d_orbit = dev_alloc<uint8_t>(size_t(s->D));
        s->d_table = dev_alloc<int32_t>(size_t(std::max<int64_t>(s->table_size, 1)));
        v = s->view();
        hipLaunchKernelGGL(basis_fill_kernel, grid, dim3(SCAN_THREADS), 0, st, v, flags, d_off, cpb, s->d_key,
                           s->d_orbit, s->d_table);
        ED_HIP(hipGetLastError());
        s->d_table_orbit = dev_alloc<uint8_t>(size_t(std::max<int64_t>(s->table_size, 1)));
        if (s->table_size > 0) {
            hipLaunchKernelGGL(table_orbit_kernel, dim3(unsigned((s->table_size + 255) / 256)), dim3(256), 0, st,
                               s->d_table, s->d_orbit, s->d_table_orbit, s->table_size);
            ED_HIP(hipGetLastError());
        }
        ED_HIP(hipStreamSynchronize(st));
        dev_free(flags);
        dev_free(counts);
        dev_free(d_off);
        hv_fix_list(s, st);
        s->d_partial = dev_alloc<double>(size_t(2 * hv_partials(s)));
    });
}

int64_t ed_basis_size(const ed_sector* s) { return s ? s->D : -1; }

int ed_basis_export(const ed_sector* s, int64_t* d_sparse, int32_t* d_orbit, void* stream) {
    return guarded([&] {
        ED_REQUIRE(s && s->D >= 0, ED_ERR_STATE, "basis not built");
        if (s->D == 0) return;
        hipStream_t st = static_cast<hipStream_t>(stream);
        const int T = 256;
        hipLaunchKernelGGL(basis_export_kernel, dim3(unsigned((s->D + T - 1) / T)), dim3(T), 0, st, s->view(),
                           d_sparse, d_orbit);
        ED_HIP(hipGetLastError());
    });
}

int ed_lookup(const ed_sector* s, const int64_t* d_sparse, int64_t* d_dense, int64_t n, void* stream) {
    return guarded([&] {
        ED_REQUIRE(s && s->D >= 0, ED_ERR_STATE, "basis not built");
        if (n == 0) return;
        hipStream_t st = static_cast<hipStream_t>(stream);
        const int T = 256;
        hipLaunchKernelGGL(lookup_kernel, dim3(unsigned((n + T - 1) / T)), dim3(T), 0, st, s->view(), d_sparse,
                           d_dense, n);
        ED_HIP(hipGetLastError());
    });
}

int ed_lowest(const ed_sector* s, const int64_t* d_sparse, int64_t* d_rep, double* d_amp, int64_t n, void* stream) {
    return guarded([&] {
        ED_REQUIRE(s && s->D >= 0, ED_ERR_STATE, "basis not built");
        if (n == 0) return;
        hipStream_t st = static_cast<hipStream_t>(stream);
        const int T = 256;
        hipLaunchKernelGGL(lowest_kernel, dim3(unsigned((n + T - 1) / T)), dim3(T), 0, st, s->view(), d_sparse,
                           d_rep, d_amp, n);
        ED_HIP(hipGetLastError());
    });
}

}  // extern "C"
