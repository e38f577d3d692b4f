// Orbit, representative and basis lookup on bit-mask states (host and device).
#pragma once

#include "ed_common.h"

namespace ed {

// Amplitude table access: A[p][n] = F[p] / sqrt(n), F[p][1] == F; with the
// spin flip F[p + N] = z F[p].
__host__ __device__ inline C2 amp_at(const SectorView& s, int p, int n) {
    const int idx = p * (s.NO + 1) + n;
    return {s.amp[2 * idx], s.amp[2 * idx + 1]};
}

// Fermion sign of group element p (the spin flip relabels spins without
// reordering any operator: sign +1).  FLIP: the sector may have the spin flip.
template <bool FLIP = true>
__host__ __device__ inline int group_sign(const SectorView& s, uint32_t H, int p) {
    if constexpr (FLIP) p = p < s.N ? p : p - s.N;
    return image_sign(H, p, s.L, s.ne_even, s.full);
}

struct OrbitResult {
    int n;          // orbit size, 0 = rejected (not lowest / incompatible)
    int best_p;     // first group element mapping the state to its lowest image
    uint64_t best;  // key of the lowest image
};

// The orbit under translations x leg mirror x spin flip (no counterpart in the
// reference): every one of the 2N elements, p + N being the flip after p.
// The stabiliser elements (images equal to the state) must all have
// character x sign 1, else the state has no component in the sector; the
// orbit size is 2N / |stabiliser|; the first element reaching the lowest
// image is kept.
__host__ __device__ inline OrbitResult orbit_walk_flip(const SectorView& s, uint32_t H, uint32_t Dn,
                                                       bool check_lowest) {
    const uint64_t k0 = make_key(H, Dn);
    const uint32_t Df = flip_downs(H, Dn, s.full);
    uint64_t best = k0;
    int best_p = 0, stab = 1;
    for (int q = 1; q < 2 * s.N; ++q) {
        const int p = q < s.N ? q : q - s.N;
        const uint64_t kt = make_key(image_mask(H, p, s.N, s.full), image_mask(q < s.N ? Dn : Df, p, s.N, s.full));
        if (check_lowest && kt < k0) return {0, 0, 0};
        if (kt == k0) {
            const C2 F = amp_at(s, q, 1);
            const double sg = double(image_sign(H, p, s.L, s.ne_even, s.full));
            const double dr = sg * F.re - 1.0, di = sg * F.im;
            if (!s.any_k && dr * dr + di * di > 1e-16) return {0, 0, 0};
            ++stab;
        } else if (kt < best) {
            best = kt;
            best_p = q;
        }
    }
    return {2 * s.N / stab, best_p, best};
}

// The reference's orbit walk, tJ_spin_half_ladder/symmetry_ladder.h:113-235:
// group elements in the order p = 1 .. N-1, optional rejection as soon as an
// image is lower than the input (check_lowest), repeat detection only for
// rung shifts m <= L/2, compatibility tested on the amplitude of the first
// repeat (|a - a_ref| > 1e-8 rejects), and the mirrored half skipped once the
// state is mirror symmetric.  With any_k the compatibility test is skipped
// (every lowest image is kept, with the same orbit size).  Amplitudes are
// sign * F[m][b] exactly as the reference's running product produces them
// (the running product differs from the unsigned one only by exact sign
// flips).
// With the spin flip, orbit_walk_flip instead.
__host__ __device__ inline OrbitResult orbit_walk(const SectorView& s, uint32_t H, uint32_t Dn, bool check_lowest) {
    if (s.flip) return orbit_walk_flip(s, H, Dn, check_lowest);
    const uint64_t k0 = make_key(H, Dn);
    bool mirror_sym = false;
    int count = 1;
    uint64_t k1 = 0;
    double a1r = 0.0, a1i = 0.0;
    uint64_t best = k0;
    int best_p = 0;
    const int half = s.L / 2 + 1;
    for (int p = 1; p < s.N; ++p) {
        const int m = p >> 1, b = p & 1;
        if (b && mirror_sym) continue;
        const uint64_t kt = make_key(image_mask(H, p, s.N, s.full), image_mask(Dn, p, s.N, s.full));
        if (check_lowest && kt < k0) return {0, 0, 0};
        const C2 F = amp_at(s, p, 1);
        const double sg = double(image_sign(H, p, s.L, s.ne_even, s.full));
        const double ar = sg * F.re, ai = sg * F.im;
        if (m < half) {
            if (kt == k0) {
                const double dr = ar - 1.0, di = ai;
                if (!s.any_k && dr * dr + di * di > 1e-16) return {0, 0, 0};
                if (p == 1) {
                    mirror_sym = true;
                    continue;
                }
                break;
            }
            if (count != 1 && kt == k1) {
                const double dr = ar - a1r, di = ai - a1i;
                if (!s.any_k && dr * dr + di * di > 1e-16) return {0, 0, 0};
                break;
            }
        }
        if (count == 1) {
            k1 = kt;
            a1r = ar;
            a1i = ai;
        }
        if (kt < best) {
            best = kt;
            best_p = p;
        }
        ++count;
    }
    return {count, best_p, best};
}

// Normalised amplitude of the lowest image reached from (H, .) by p, as the
// reference returns it from get_lowest_sparse_amplitude_from_state
// (general/symmetry.h:240-269): sign(H, p) * F[m][b] / sqrt(n).
__host__ __device__ inline C2 lowest_amplitude(const SectorView& s, uint32_t H, int p, int n) {
    const C2 a = amp_at(s, p, n);
    const double sg = double(group_sign(s, H, p));
    return {sg * a.re, sg * a.im};
}

struct Found {
    int64_t j;  // dense index, -1 = not a basis state
    int n;      // orbit size of the basis state
};

// Dense index of a representative (its hole pattern is the orbit minimum).
// Generic blocks (hole pattern without a stabiliser) hold every spin
// configuration, so dense = base + spin rank; symmetric blocks keep a table.
__device__ inline Found lookup_rep(const SectorView& s, uint32_t Hr, uint32_t Dr) {
    const int64_t hr = rank_holes(Hr, s.comb);
    const int blk = s.block_of_rank[hr];
    if (blk < 0) return {-1, 0};
    const int64_t sr = rank_spins(Hr, Dr, s.comb);
    const int64_t off = s.block_table[blk];
    if (off < 0) return {s.block_base[blk] + sr, s.NO};
    const int32_t j = s.table[off + sr];
    if (j < 0) return {-1, 0};
    return {int64_t(j), int(s.orbit[j])};
}

struct Rep {
    uint32_t H, Dn;
    int p;  // group element (first in enumeration order) giving the lowest image
    int blk;
};

// Lowest image of the down pattern among the group elements gm (bit p: image
// p gives the minimal hole pattern): compared in increasing p, first minimum
// wins, which is the element the reference's orbit walk stores first.  With
// the spin flip each element's flipped image (element p + N) follows it.
template <bool FLIP = true>
__host__ __device__ inline void lowest_downs(const SectorView& s, uint32_t H, uint32_t Dn, uint32_t gm,
                                             uint32_t& best, int& bp) {
    const uint32_t Ef = FLIP && s.flip ? ~H & s.full : 0u;
    bp = -1;
    while (gm) {
        const int p = ctz32(gm);
        const uint32_t d = image_mask(Dn, p, s.N, s.full);
        if (bp < 0 || d < best) {
            best = d;
            bp = p;
        }
        if (FLIP && s.flip) {
            const uint32_t df = image_mask(Ef, p, s.N, s.full) & ~d;
            if (df < best) {
                best = df;
                bp = p + s.N;
            }
        }
        gm &= gm - 1;
    }
}

// Lowest image of an arbitrary state using the per-hole-rank table of group
// elements that minimise the hole pattern.
__device__ inline Rep find_rep(const SectorView& s, uint32_t H, uint32_t Dn) {
    const int64_t hr = rank_holes(H, s.comb);
    const uint32_t gm = s.canon_gmask[hr];
    const int blk = s.canon_block[hr];
    const uint32_t Hc = image_mask(H, ctz32(gm), s.N, s.full);
    uint32_t best;
    int bp;
    lowest_downs(s, H, Dn, gm, best, bp);
    return {Hc, best, bp, blk};
}

}  // namespace ed
