// One row of the gather-form H·v, shared by the single-sector kernel and the
// batched kernels (gfx950).
//
// Row f of y = H x in the gather form y_f = sum_j conj(H_jf) x_j, H_jf the
// reference's matrix element <j~|H|f~> (general/column_functions.h:99-148 for
// one element, tJ_spin_half_ladder/matrix_setup.pyx:80-118 for the four
// terms).  The element computation reads only small per-sector tables
// (binomials, amplitude table, sqrt(n), per-hole-pattern group elements and
// blocks, block offsets); the kernels stage them in LDS, or point at their
// copies in global memory when they do not fit.  Ranks are 32-bit.  An
// exchange on a hole pattern without stabiliser gets its target as row +
// window rank change, with the same coefficient for every bond of a term.
// Gathers of x_j are issued G at a time before any is consumed.
#pragma once

#include <algorithm>

#include "ed_internal.h"
#include "ed_matrix.h"
#include "ed_orbit.h"

namespace ed {

// Byte offsets of the staged tables; rt >= 0 adds the chunked spin-rank
// table (RankGeom) for the wavefront-uniform row walk.  comb_bytes: the
// leading rows of the binomial table that are staged; amp < 0: the
// amplitude table is read from global memory, not staged.
struct LdsLayout {
    int comb, amp, sq, gm, cb, bb, bt, bytes;
    int comb_bytes = 0;
    int rt = -1;
    int rt_n = 0, rt_entries = 0;
    int rt_base[4] = {0, 0, 0, 0};
    int rt_width[4] = {0, 0, 0, 0};
};

// lean: only the binomial rows 0 .. max(h, nd) + 1 (every rank, rank
// change and hole rank reads rows k <= max(h, nd)), and no amplitude table.
inline LdsLayout lds_layout(const ed_sector* s, bool lean = false) {
    LdsLayout l{};
    int off = 0;
    auto take = [&](int bytes) {
        const int at = off;
        off += (bytes + 15) & ~15;
        return at;
    };
    l.comb_bytes = (lean ? std::min(COMB_W, std::max(s->h, s->nd) + 2) : COMB_W) * COMB_W * 4;
    l.comb = take(l.comb_bytes);
    const int G = s->flip ? 2 * s->N : s->N;
    l.amp = lean ? -1 : take(G * (G + 1) * 16);
    l.sq = take((G + 1) * 8);
    l.gm = take(int(s->n_hole) * 4);
    l.cb = take(int(s->n_hole) * 4);
    l.bb = take(int(s->block_rank.size()) * 8);
    l.bt = take(int(s->block_rank.size()) * 8);
    l.bytes = off;
    return l;
}

constexpr int LDS_CAP = 60 * 1024;

// lds_layout with the chunked spin-rank table appended when the sector can
// take the wavefront-uniform walk (no spin flip, one momentum, vectors under
// 4 GiB) and all of it fits in LDS_CAP; else rt = -1.
inline LdsLayout lds_layout_ranks(const ed_sector* s) {
    const LdsLayout full = lds_layout(s);
    LdsLayout l = lds_layout(s, true);
    const RankGeom g = rank_geom(s->ne, s->nd);
    const int bytes = (g.entries * 4 + 15) & ~15;
    // 32-bit byte offsets of x_j for either vector type
    const bool small = s->D * 16 < (int64_t(1) << 32);
    if (s->flip || s->any_k || !small || s->rank_tab.empty() || l.bytes + bytes > LDS_CAP) return full;
    l.rt = l.bytes;
    l.bytes += bytes;
    l.rt_n = g.n;
    l.rt_entries = g.entries;
    for (int c = 0; c < 4; ++c) {
        l.rt_base[c] = g.base[c];
        l.rt_width[c] = g.width[c];
    }
    return l;
}

struct Tabs {
    const uint32_t* comb;
    const double* amp;
    const double* sq;
    const uint32_t* gm;
    const int32_t* cb;
    const int64_t* bb;
    const int64_t* bt;
    const uint32_t* rt;  // chunked spin ranks (LdsLayout::rt >= 0 only)
};

__device__ inline Tabs global_tabs(const SectorView& s) {
    return Tabs{s.comb, s.amp, s.sqrt_tab, s.canon_gmask, s.canon_block, s.block_base, s.block_table, s.rank_tab};
}

// Copy `bytes` bytes into LDS with 16-byte loads (dst and src 16-byte
// aligned), the tail with 4-byte ones.
template <int BS>
__device__ inline void stage_bytes(unsigned char* dst, const void* src, int bytes) {
    const int nv = bytes >> 4;
    const uint4* s4 = static_cast<const uint4*>(src);
    uint4* d4 = reinterpret_cast<uint4*>(dst);
    for (int i = threadIdx.x; i < nv; i += BS) d4[i] = s4[i];
    const int nt = (bytes - (nv << 4)) >> 2;
    if (int(threadIdx.x) < nt)
        reinterpret_cast<uint32_t*>(dst + (nv << 4))[threadIdx.x] =
            reinterpret_cast<const uint32_t*>(static_cast<const unsigned char*>(src) + (nv << 4))[threadIdx.x];
}

// Copy the tables into LDS (all threads of the workgroup take part; the
// caller synchronises before use).  The device tables are separate
// allocations (256-byte aligned) and the layout's offsets are multiples of 16.
template <int BS>
__device__ inline Tabs stage_tabs(const SectorView& s, const LdsLayout& lay, unsigned char* lds) {
    stage_bytes<BS>(lds + lay.comb, s.comb, lay.comb_bytes);
    const double* amp = s.amp;
    if (lay.amp >= 0) {
        stage_bytes<BS>(lds + lay.amp, s.amp, s.G * (s.NO + 1) * 16);
        amp = reinterpret_cast<const double*>(lds + lay.amp);
    }
    stage_bytes<BS>(lds + lay.sq, s.sqrt_tab, (s.NO + 1) * 8);
    stage_bytes<BS>(lds + lay.gm, s.canon_gmask, s.n_hole_ranks * 4);
    stage_bytes<BS>(lds + lay.cb, s.canon_block, s.n_hole_ranks * 4);
    stage_bytes<BS>(lds + lay.bb, s.block_base, s.n_blocks * 8);
    stage_bytes<BS>(lds + lay.bt, s.block_table, s.n_blocks * 8);
    const uint32_t* rt = nullptr;
    if (lay.rt >= 0) {
        stage_bytes<BS>(lds + lay.rt, s.rank_tab, lay.rt_entries * 4);
        rt = reinterpret_cast<const uint32_t*>(lds + lay.rt);
    }
    return Tabs{reinterpret_cast<const uint32_t*>(lds + lay.comb), amp,
                reinterpret_cast<const double*>(lds + lay.sq), reinterpret_cast<const uint32_t*>(lds + lay.gm),
                reinterpret_cast<const int32_t*>(lds + lay.cb), reinterpret_cast<const int64_t*>(lds + lay.bb),
                reinterpret_cast<const int64_t*>(lds + lay.bt), rt};
}

__device__ inline int32_t rank_holes32(uint32_t H, const uint32_t* comb) {
    int32_t r = 0;
    int k = 1;
    while (H) {
        r += int32_t(comb[k * COMB_W + ctz32(H)]);
        ++k;
        H &= H - 1;
    }
    return r;
}

__device__ inline int32_t rank_spins32(uint32_t H, uint32_t Dn, const uint32_t* comb) {
    int32_t r = 0;
    int k = 1;
    while (Dn) {
        const int p = ctz32(Dn);
        r += int32_t(comb[k * COMB_W + p - popc32(H & ((1u << p) - 1u))]);
        ++k;
        Dn &= Dn - 1;
    }
    return r;
}

// Spin-rank change when the down pattern Dn becomes D2 and the two differ
// only inside sites [lo, hi].
__device__ inline int32_t window_delta32(uint32_t H, uint32_t Dn, uint32_t D2, int lo, int hi, const uint32_t* comb) {
    const uint32_t below = (1u << lo) - 1u;
    int kb = popc32(Dn & below), ka = kb, hb = popc32(H & below);
    int32_t d = 0;
    for (int q = lo; q <= hi; ++q) {
        const uint32_t bit = 1u << q;
        if (H & bit) {
            ++hb;
            continue;
        }
        const int c = q - hb;
        if (Dn & bit) d -= int32_t(comb[(++kb) * COMB_W + c]);
        if (D2 & bit) d += int32_t(comb[(++ka) * COMB_W + c]);
    }
    return d;
}

// One matrix element prepared for the gather: target row and the weighted
// conjugate amplitude c, so that acc += c * x[j].
struct Elem {
    int64_t j;
    C2 c;
    uint32_t code;  // stored-form element code (elem_code), NO_ELEM for an empty slot
};

template <bool FLIP>
__device__ inline C2 weighted_conj(const SectorView& s, const Tabs& T, uint32_t Hsrc, int p, int nj, double scale,
                                   C2 W) {
    const int idx = p * ((FLIP ? s.NO : s.N) + 1) + nj;
    const double ar = T.amp[2 * idx], ai = T.amp[2 * idx + 1];
    const double sg = double(group_sign<FLIP>(s, Hsrc, p)) * scale;
    return cmul(W, C2{sg * ar, -(sg * ai)});
}

// Lowest image among the group elements gm (hole part already minimal),
// then its row: generic blocks by rank, symmetric blocks by table.
template <bool FLIP>
__device__ inline bool locate(const SectorView& s, const Tabs& T, uint32_t Hsrc, uint32_t D2, uint32_t gm, int blk,
                              int& p_out, int64_t& j_out, int& nj_out) {
    const uint32_t Hc = image_mask(Hsrc, ctz32(gm), s.N, s.full);
    uint32_t best;
    int bp;
    lowest_downs<FLIP>(s, Hsrc, D2, gm, best, bp);
    const int32_t sr = rank_spins32(Hc, best, T.comb);
    const int64_t off = T.bt[blk];
    p_out = bp;
    if (off < 0) {
        j_out = T.bb[blk] + sr;
        nj_out = FLIP ? s.NO : s.N;
        return true;
    }
    const int32_t j = s.table[off + sr];
    if (j < 0) return false;
    j_out = j;
    nj_out = int(s.orbit[j]);
    return true;
}

template <class VT>
__device__ inline C2 ldx(const VT* __restrict__ x, int64_t j) {
    const VT v = x[j];
    return {double(v.x), double(v.y)};
}

// Element code for the stored form: combo = (term * G + p) * (NO + 1) + n_j
// (term 0..3 = tl, tr, jl, jr; p the group element of G; n_j the target's
// orbit size, at most NO) in the low 15 bits, bit 15 set when the element's
// sign is negative.  NO_ELEM marks an empty slot of a group.
constexpr uint32_t NO_ELEM = 0xffffffffu;

template <bool FLIP>
__device__ inline uint32_t elem_code(const SectorView& s, int term, int p, int nj, bool neg) {
    const int G = FLIP ? s.G : s.N, NO = FLIP ? s.NO : s.N;
    return uint32_t((term * G + p) * (NO + 1) + nj) | (neg ? 0x8000u : 0u);
}

// Walk the elements of row f in a fixed order (jl bonds, jr bonds, hops onto
// each hole), G exchange elements or three hops per group, and hand each
// group to the sink: sink.diag(term, -count/2) for the exchange diagonal of
// term 2 (jl) and 3 (jr), sink.group(e) for the elements.
template <int G, bool FLIP, class Sink>
__device__ inline void walk_row_t(const SectorView& s, const Tabs& T, const HvWeights& w, int64_t f, Sink& sink) {
    const int N = s.N, L = s.L, ne_even = s.ne_even;
    const uint32_t full = s.full;
    const uint64_t key = s.key[f];
    const uint32_t H = uint32_t(key >> 32), Dn = uint32_t(key);
    const uint32_t E = ~H & full, Up = E & ~Dn;
    const double sqf = T.sq[s.orbit[f]];
    const int32_t hr = rank_holes32(H, T.comb);
    const uint32_t gm_self = T.gm[hr];
    const int blk_self = T.cb[hr];
    const bool generic = (gm_self & (gm_self - 1u)) == 0u;
    // with the spin flip a generic representative has its top electron up: an exchange that turns that spin
    // down reaches the flipped pattern instead
    const uint32_t top_e = FLIP ? (1u << (31 - __clz(int(E)))) : 0u;
    const int NO = FLIP ? s.NO : N;
    // an exchange element: generic blocks by window rank change (coefficient cg, the same for every bond of the
    // term), others by the stabiliser search
    auto exchange = [&](int term, uint32_t D2, int lo, int hi, bool wrap, C2 W, C2 cg, Elem& e) {
        if (generic && !(D2 & top_e)) {
            e.j = wrap ? T.bb[blk_self] + rank_spins32(H, D2, T.comb) : f + window_delta32(H, Dn, D2, lo, hi, T.comb);
            e.c = cg;
            e.code = elem_code<FLIP>(s, term, 0, NO, false);
            return;
        }
        int p, nj;
        int64_t j;
        if (locate<FLIP>(s, T, H, D2, gm_self, blk_self, p, j, nj)) {
            e.j = j;
            e.c = weighted_conj<FLIP>(s, T, H, p, nj, 0.5 * sqf, W);
            e.code = elem_code<FLIP>(s, term, p, nj, group_sign<FLIP>(s, H, p) < 0);
        }
    };
    if (w.active & 4) {  // jl: bonds (i, i + 2)
        const uint32_t d2 = rotl_n(Dn, N - 2, N, full), u2 = rotl_n(Up, N - 2, N, full);
        const uint32_t anti = N == 2 ? 0u : (((Dn & u2) | (Up & d2)) & full);
        const C2 W = w.w[2];
        const C2 cg = weighted_conj<FLIP>(s, T, H, 0, NO, 0.5 * sqf, W);
        sink.diag(2, -0.5 * double(popc32(anti)));
        for (int g = 0; g < N; g += G) {
            Elem e[G];
#pragma unroll
            for (int k = 0; k < G; ++k) {
                e[k] = Elem{f, C2{0.0, 0.0}, NO_ELEM};
                const int i = g + k;
                if (i < N && ((anti >> i) & 1u)) {
                    const int kk = i + 2 < N ? i + 2 : i + 2 - N;
                    const bool wrap = kk < i;
                    exchange(2, Dn ^ (1u << i) ^ (1u << kk), wrap ? kk : i, wrap ? i : kk, wrap, W, cg, e[k]);
                }
            }
            sink.group(e);
        }
    }
    if (w.active & 8) {  // jr: bonds (2x, 2x + 1)
        const uint32_t anti = (Dn ^ (Dn >> 1)) & E & (E >> 1) & EVEN_BITS & full;
        const C2 W = w.w[3];
        const C2 cg = weighted_conj<FLIP>(s, T, H, 0, NO, 0.5 * sqf, W);
        sink.diag(3, -0.5 * double(popc32(anti)));
        for (int g = 0; g < L; g += G) {
            Elem e[G];
#pragma unroll
            for (int k = 0; k < G; ++k) {
                e[k] = Elem{f, C2{0.0, 0.0}, NO_ELEM};
                const int i = 2 * (g + k);
                if (g + k < L && ((anti >> i) & 1u)) exchange(3, Dn ^ (3u << i), i, i + 1, false, W, cg, e[k]);
            }
            sink.group(e);
        }
    }
    if (w.active & 3) {  // hops onto each hole: leg neighbours (tl) and rung partner (tr)
        uint32_t holes = H;
        while (holes) {
            const int t = ctz32(holes);
            holes &= holes - 1;
            Elem e[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                e[k] = Elem{f, C2{0.0, 0.0}, NO_ELEM};
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
                int p, nj;
                int64_t j;
                if (locate<FLIP>(s, T, H2, D2, T.gm[hr2], blk, p, j, nj)) {
                    e[k].j = j;
                    e[k].c = weighted_conj<FLIP>(s, T, H2, p, nj, par ? sqf : -sqf, W);
                    e[k].code = elem_code<FLIP>(s, k < 2 ? 0 : 1, p, nj, (group_sign<FLIP>(s, H2, p) < 0) == bool(par));
                }
            }
            sink.group(e);
        }
    }
}

// walk_row_t for the sector's group (the spin-flip code only where the
// sector has the spin flip; the branch is the same for every lane).
template <int G, class Sink>
__device__ inline void walk_row(const SectorView& s, const Tabs& T, const HvWeights& w, int64_t f, Sink& sink) {
    if (s.flip)
        walk_row_t<G, true>(s, T, w, f, sink);
    else
        walk_row_t<G, false>(s, T, w, f, sink);
}

// Matrix-free sink: gathers of one group are issued together, then
// accumulated in element order.
template <class VT>
struct GatherSink {
    const VT* __restrict__ x;
    C2 xf;
    HvWeights w;
    C2 acc;

    __device__ inline void diag(int term, double d) {
        acc = cadd(acc, cmul(w.w[term], C2{d * xf.re, d * xf.im}));
    }

    template <int G>
    __device__ inline void group(const Elem (&e)[G]) {
        C2 v[G];
#pragma unroll
        for (int k = 0; k < G; ++k) v[k] = ldx(x, e[k].j);
#pragma unroll
        for (int k = 0; k < G; ++k) acc = cadd(acc, cmul(e[k].c, v[k]));
    }
};

// (H x)_f; also returns x_f for the fused <x, Hx>.
template <class VT, int G>
__device__ inline C2 hv_row(const SectorView& s, const Tabs& T, const HvWeights& w, const VT* __restrict__ x,
                            int64_t f, C2& xf) {
    GatherSink<VT> sink{x, ldx(x, f), w, C2{0.0, 0.0}};
    walk_row<G>(s, T, w, f, sink);
    xf = sink.xf;
    return sink.acc;
}

}  // namespace ed
