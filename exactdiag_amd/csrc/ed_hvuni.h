// The wavefront-uniform row walk of the matrix-free gather-form H·v (gfx950),
// shared by the single-sector kernel (ed_hv.hip) and the batched engine
// (ed_batch.hip): for a wavefront whose 64 rows share one hole pattern, the
// hole-only work is done once per wavefront, exchange targets in blocks
// without stabiliser come from closed-form rank changes, hop targets from the
// chunked spin-rank table (LdsLayout::rt), and gathers use 32-bit offsets
// (vectors under 4 GiB).
#pragma once

#include <cstdlib>

#include "ed_hvrow.h"

namespace ed {

// The down pattern D with the sites of the hole mask Hm taken out (electron
// order).  Hm is the same in every lane, so the loop runs on the scalar unit.
__device__ inline uint32_t electron_order(uint32_t D, uint32_t Hm) {
    while (Hm) {
        const int p = 31 - __builtin_clz(Hm);
        const uint32_t lo = (1u << p) - 1u;
        D = (D & lo) | ((D >> 1) & ~lo);
        Hm &= lo;
    }
    return D;
}

// Colex rank of an electron-order down pattern from the chunked table
// (RankGeom): one table read per 8-bit chunk instead of one binomial per
// down spin.
__device__ inline int32_t rank_by_chunks(uint32_t c, const uint32_t* rt, const LdsLayout& lay) {
    uint32_t r = 0;
    int kb = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k < lay.rt_n) {
            const uint32_t bits = (c >> (8 * k)) & 0xffu;
            r += rt[lay.rt_base[k] + (kb << lay.rt_width[k]) + int(bits)];
            kb += popc32(bits);
        }
    }
    return int32_t(r);
}

// Spin-rank change of an exchange on the bond (i, i + span), span 1 (rung) or
// 2 (leg, without wrap), both ends electrons with opposite spins: the down
// spin at electron position c = i - (holes below i), with kd downs below it,
// moves to the other end (Dn has site i) or comes from it.  Between the ends
// sits at most one site (span 2): a hole, or an electron whose spin md shifts
// the second binomial.  Equal to the window rank change of the general walk.
__device__ inline int32_t exchange_delta(uint32_t H, uint32_t Dn, int i, int span, const uint32_t* comb) {
    const uint32_t below = (1u << i) - 1u;
    const int c = i - popc32(H & below);
    const int kd = popc32(Dn & below);
    int32_t d = int32_t(comb[kd * COMB_W + c]);
    if (span == 2 && !((H >> (i + 1)) & 1u)) d += int32_t(comb[(kd + int((Dn >> (i + 1)) & 1u)) * COMB_W + c + 1]);
    return ((Dn >> i) & 1u) ? d : -d;
}

// x_j at a 32-bit byte offset (the uniform walk needs D * sizeof(VT) < 2^32):
// the address is the scalar base plus the lane's offset.
template <class VT>
__device__ inline C2 ldx32(const VT* __restrict__ x, uint32_t j) {
    const VT v = *reinterpret_cast<const VT*>(reinterpret_cast<const char*>(x) + uint64_t(j * uint32_t(sizeof(VT))));
    return {double(v.x), double(v.y)};
}

// Sum of x_j over one group's elements, loads issued for the lanes that
// have an element before any is used (empty slots add zero).
template <int K, class VT>
__device__ inline void gather_sum(const VT* __restrict__ x, const uint32_t (&j)[K], const bool (&act)[K], C2& sum) {
    C2 v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        v[k] = C2{0.0, 0.0};
        if (act[k]) v[k] = ldx32(x, j[k]);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) sum = cadd(sum, v[k]);
}

// Exchange elements of one term on the bonds listed in `bonds` (bit i: bond
// (i, i + span), the same in every lane), G at a time: every element of the
// term has the same coefficient, so the x_j are summed and multiplied once.
template <int G, int SPAN, class VT>
__device__ inline C2 exchange_sum(const VT* __restrict__ x, const Tabs& T, uint32_t H, uint32_t Dn, uint32_t anti,
                                  uint32_t bonds, uint32_t f) {
    C2 sum{0.0, 0.0};
    while (bonds) {
        uint32_t j[G];
        bool act[G];
#pragma unroll
        for (int k = 0; k < G; ++k) {
            j[k] = f;
            act[k] = false;
            if (bonds) {
                const int i = ctz32(bonds);
                bonds &= bonds - 1;
                act[k] = (anti >> i) & 1u;
                j[k] = f + uint32_t(exchange_delta(H, Dn, i, SPAN, T.comb));
            }
        }
        gather_sum<G>(x, j, act, sum);
    }
    return sum;
}

// Exchange elements of one term in a block whose hole pattern H has a
// stabiliser (gm: its group elements, identity included; the block's spin
// ranks map to rows through the table at `off`): the target (H, D2) is
// brought to its lowest image over gm (image_mask(H, p) = H for those p,
// the first minimum wins as in lowest_downs), ranked, looked up with its
// orbit size, and weighted per lane.  Table reads of a group are issued
// together, then its gathers.  SPAN 2 includes the bonds that wrap round.
template <int G, int SPAN, class VT>
__device__ inline void exchange_sym(const SectorView& s, const VT* __restrict__ x, const Tabs& T,
                                    const LdsLayout& lay, uint32_t H, uint32_t Dn, uint32_t anti, uint32_t bonds,
                                    uint32_t gm, int64_t off, double half_sqf, C2 W, C2& acc) {
    const int N = s.N;
    const uint32_t full = s.full;
    while (bonds) {
        uint32_t idx[G];
        int bp[G];
        bool act[G];
#pragma unroll
        for (int k = 0; k < G; ++k) {
            idx[k] = 0;
            bp[k] = 0;
            act[k] = false;
            if (bonds) {
                const int i = ctz32(bonds);
                bonds &= bonds - 1;
                const int kk = i + SPAN < N ? i + SPAN : i + SPAN - N;
                act[k] = (anti >> i) & 1u;
                const uint32_t D2 = Dn ^ (1u << i) ^ (1u << kk);
                uint32_t best = 0;
                uint32_t g = gm;
                bool first = true;
                while (g) {
                    const int p = ctz32(g);
                    g &= g - 1;
                    const uint32_t d = image_mask(D2, p, N, full);
                    if (first || d < best) {
                        best = d;
                        bp[k] = p;
                    }
                    first = false;
                }
                idx[k] = uint32_t(off) + uint32_t(rank_by_chunks(electron_order(best, H), T.rt, lay));
            }
        }
        int32_t j[G];
        int nj[G];
#pragma unroll
        for (int k = 0; k < G; ++k) {
            j[k] = -1;
            nj[k] = 0;
            if (act[k]) {
                j[k] = s.table[idx[k]];
                nj[k] = s.table_orbit[idx[k]];
            }
        }
        C2 v[G];
#pragma unroll
        for (int k = 0; k < G; ++k) {
            act[k] = act[k] && j[k] >= 0;
            v[k] = C2{0.0, 0.0};
            if (act[k]) v[k] = ldx32(x, uint32_t(j[k]));
        }
#pragma unroll
        for (int k = 0; k < G; ++k)
            if (act[k]) acc = cadd(acc, cmul(weighted_conj<false>(s, T, H, bp[k], nj[k], half_sqf, W), v[k]));
    }
}

// Row f of the gather form when every lane of the wavefront has the same hole
// pattern H (a scalar), in a sector without spin flip: the elements and
// coefficients of walk_row_t, with the hole-only work (bond sites, hop
// targets' hole patterns, group elements, signs and coefficients) done once
// per wavefront and hop targets' ranks from the chunked table.  Without
// stabiliser (SYM false: block blk_self holds every spin configuration,
// orbit size N) exchange targets come from exchange_delta and each exchange
// term's x_j are summed before its one coefficient is applied; with one
// (SYM: stabiliser elements gm_self, rows through the block's table)
// exchange_sym finds and weights them per lane.
template <class VT, int G, bool SYM>
__device__ inline C2 hv_row_uniform(const SectorView& s, const Tabs& T, const LdsLayout& lay, const HvWeights& w,
                                    const VT* __restrict__ x, uint32_t f, uint32_t H, uint32_t Dn, int blk_self,
                                    uint32_t gm_self, C2& xf) {
    const int N = s.N, L = s.L;
    const uint32_t full = s.full;
    const uint32_t E = ~H & full, Up = E & ~Dn;
    GatherSink<VT> sink{x, ldx32(x, f), w, C2{0.0, 0.0}};
    const double sqf = SYM ? T.sq[s.orbit[f]] : T.sq[N];
    const uint32_t base_self = uint32_t(T.bb[blk_self]);
    const int64_t off_self = T.bt[blk_self];
    if (SYM && (w.active & 4)) {  // jl, every bond by the stabiliser search
        const uint32_t d2 = rotl_n(Dn, N - 2, N, full), u2 = rotl_n(Up, N - 2, N, full);
        const uint32_t anti = N == 2 ? 0u : (((Dn & u2) | (Up & d2)) & full);
        const uint32_t both = N == 2 ? 0u : (E & rotl_n(E, N - 2, N, full));
        sink.diag(2, -0.5 * double(popc32(anti)));
        exchange_sym<G, 2>(s, x, T, lay, H, Dn, anti, both, gm_self, off_self, 0.5 * sqf, w.w[2], sink.acc);
    }
    if (SYM && (w.active & 8)) {  // jr
        const uint32_t anti = (Dn ^ (Dn >> 1)) & E & (E >> 1) & EVEN_BITS & full;
        const uint32_t both = E & (E >> 1) & EVEN_BITS & full;
        sink.diag(3, -0.5 * double(popc32(anti)));
        exchange_sym<G, 1>(s, x, T, lay, H, Dn, anti, both, gm_self, off_self, 0.5 * sqf, w.w[3], sink.acc);
    }
    if (!SYM && (w.active & 4)) {  // jl: bonds (i, i + 2)
        const uint32_t d2 = rotl_n(Dn, N - 2, N, full), u2 = rotl_n(Up, N - 2, N, full);
        const uint32_t anti = N == 2 ? 0u : (((Dn & u2) | (Up & d2)) & full);
        const uint32_t both = N == 2 ? 0u : (E & rotl_n(E, N - 2, N, full));
        const C2 cg = weighted_conj<false>(s, T, H, 0, N, 0.5 * sqf, w.w[2]);
        sink.diag(2, -0.5 * double(popc32(anti)));
        // bonds i <= N - 3 by the window rank change, the two that wrap round by the full rank
        const uint32_t inner = both & (full >> 2);
        C2 sum = exchange_sum<G, 2>(x, T, H, Dn, anti, inner, f);
        uint32_t j[2];
        bool act[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int i = N - 2 + k, kk = k;
            j[k] = f;
            act[k] = false;
            if ((both >> i) & 1u) {
                act[k] = (anti >> i) & 1u;
                const uint32_t D2 = act[k] ? Dn ^ (1u << i) ^ (1u << kk) : Dn;
                j[k] = base_self + uint32_t(rank_by_chunks(electron_order(D2, H), T.rt, lay));
            }
        }
        gather_sum<2>(x, j, act, sum);
        sink.acc = cadd(sink.acc, cmul(cg, sum));
    }
    if (!SYM && (w.active & 8)) {  // jr: bonds (2x, 2x + 1)
        const uint32_t anti = (Dn ^ (Dn >> 1)) & E & (E >> 1) & EVEN_BITS & full;
        const uint32_t both = E & (E >> 1) & EVEN_BITS & full;
        const C2 cg = weighted_conj<false>(s, T, H, 0, N, 0.5 * sqf, w.w[3]);
        sink.diag(3, -0.5 * double(popc32(anti)));
        const C2 sum = exchange_sum<G, 1>(x, T, H, Dn, anti, both, f);
        sink.acc = cadd(sink.acc, cmul(cg, sum));
    }
    if (w.active & 3) {  // hops onto each hole: leg neighbours (tl) and rung partner (tr)
        uint32_t holes = H;
        while (holes) {
            const int t = ctz32(holes);
            holes &= holes - 1;
            uint32_t j[3];
            C2 c[3];
            bool act[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                j[k] = f;
                c[k] = C2{0.0, 0.0};
                act[k] = false;
                if (k < 2 && !(w.active & 1)) continue;
                if (k == 2 && !(w.active & 2)) continue;
                const int src = k == 0 ? (t + 2 < N ? t + 2 : t + 2 - N) : k == 1 ? (t >= 2 ? t - 2 : t - 2 + N)
                                                                                   : (t ^ 1);
                if (!(E & (1u << src))) continue;
                const C2 W = k < 2 ? w.w[0] : w.w[1];
                const uint32_t bt = 1u << t, bs = 1u << src;
                const uint32_t H2 = H ^ bt ^ bs;
                const uint32_t D2 = (Dn & bs) ? (Dn ^ bs ^ bt) : Dn;
                const uint32_t above_s = src >= 31 ? 0u : (~0u << (src + 1));
                const uint32_t above_t = t >= 31 ? 0u : (~0u << (t + 1));
                const int par = (popc32(E & above_s) + popc32(E & above_t) + (t < src ? 1 : 0)) & 1;
                const int32_t hr2 = rank_holes32(H2, T.comb);
                const int blk = T.cb[hr2];
                if (blk < 0) continue;
                const uint32_t gm = T.gm[hr2];
                if ((gm & (gm - 1u)) == 0u && T.bt[blk] < 0) {
                    // target pattern without stabiliser: one group element for every lane
                    const int p = ctz32(gm);
                    const uint32_t Hc = image_mask(H2, p, N, full);
                    j[k] = uint32_t(T.bb[blk]) +
                           uint32_t(rank_by_chunks(electron_order(image_mask(D2, p, N, full), Hc), T.rt, lay));
                    c[k] = weighted_conj<false>(s, T, H2, p, N, par ? sqf : -sqf, W);
                    act[k] = true;
                } else {
                    int p, nj;
                    int64_t jj;
                    if (locate<false>(s, T, H2, D2, gm, blk, p, jj, nj)) {
                        j[k] = uint32_t(jj);
                        c[k] = weighted_conj<false>(s, T, H2, p, nj, par ? sqf : -sqf, W);
                        act[k] = true;
                    }
                }
            }
            C2 v[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                v[k] = C2{0.0, 0.0};
                if (act[k]) v[k] = ldx32(x, j[k]);
            }
#pragma unroll
            for (int k = 0; k < 3; ++k)
                if (act[k]) sink.acc = cadd(sink.acc, cmul(c[k], v[k]));
        }
    }
    xf = sink.xf;
    return sink.acc;
}

// Whether the wavefront of rows f0 .. f0 + 63 (f0 < D) can take
// hv_row_uniform: every live row (below D) has the same hole pattern H, whose
// block is without stabiliser (generic) or has a table.  Lanes past the last
// row (live false) read row f0.  Outputs describe the wavefront's rows when
// it returns true.  The list of wavefronts that fail it (hv_fix_list) is
// built with this same function.
__device__ inline bool wave_uniform(const SectorView& s, const Tabs& T, int64_t f, int64_t f0, bool live, uint32_t& H,
                                    uint32_t& Dn, int& blk, uint32_t& gm, bool& generic) {
    const uint64_t key = s.key[live ? f : f0];
    const uint32_t Hl = uint32_t(key >> 32);
    H = __builtin_amdgcn_readfirstlane(Hl);
    Dn = uint32_t(key);
    if (!__all(Hl == H)) return false;
    const int32_t hr = rank_holes32(H, T.comb);
    gm = T.gm[hr];
    blk = T.cb[hr];
    generic = (gm & (gm - 1u)) == 0u && T.bt[blk] < 0;
    return generic || T.bt[blk] >= 0;
}

// Row f through hv_row_uniform when its wavefront (rows from f0, f0 < D)
// passes wave_uniform (every wavefront but those where a block ends); lanes
// past the last row (live false) walk row f0 and keep the result to
// themselves.  Returns false, with nothing computed, otherwise.
template <class VT, int G>
__device__ inline bool uniform_row(const SectorView& s, const Tabs& T, const LdsLayout& lay, const HvWeights& w,
                                   const VT* __restrict__ x, int64_t f, int64_t f0, bool live, C2& acc, C2& xf) {
    uint32_t H, Dn, gm;
    int blk;
    bool generic;
    if (!wave_uniform(s, T, f, f0, live, H, Dn, blk, gm, generic)) return false;
    const uint32_t fe = uint32_t(live ? f : f0);
    acc = generic ? hv_row_uniform<VT, G, false>(s, T, lay, w, x, fe, H, Dn, blk, gm, xf)
                  : hv_row_uniform<VT, G, true>(s, T, lay, w, x, fe, H, Dn, blk, gm, xf);
    return true;
}

// The wavefront-uniform walk unless ED_HV_UNIFORM=0.
inline bool hv_uniform_enabled() {
    static const bool on = [] {
        const char* e = std::getenv("ED_HV_UNIFORM");
        return !(e && e[0] == '0');
    }();
    return on;
}

}  // namespace ed
