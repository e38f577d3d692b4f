// Orbit, representative and basis lookup on bit-mask states (host and device).
#pragma once

#include "ed_common.h"

namespace ed {

// Amplitude table access: A[m][b][n] = F[m][b] / sqrt(n), F[m][b][1] == F.
__host__ __device__ inline C2 amp_at(const SectorView& s, int p, int n) {
    const int idx = ((p >> 1) * 2 + (p & 1)) * (s.N + 1) + n;
    return {s.amp[2 * idx], s.amp[2 * idx + 1]};
}

struct OrbitResult {
    int n;          // orbit size, 0 = rejected (not lowest / incompatible)
    int best_p;     // first group element mapping the state to its lowest image
    uint64_t best;  // key of the lowest image
};

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
__host__ __device__ inline OrbitResult orbit_walk(const SectorView& s, uint32_t H, uint32_t Dn, bool check_lowest) {
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
    const double sg = double(image_sign(H, p, s.L, s.ne_even, s.full));
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
    if (off < 0) return {s.block_base[blk] + sr, s.N};
    const int32_t j = s.table[off + sr];
    if (j < 0) return {-1, 0};
    return {int64_t(j), int(s.orbit[j])};
}

struct Rep {
    uint32_t H, Dn;
    int p;  // group element (first in enumeration order) giving the lowest image
    int blk;
};

// Lowest image of an arbitrary state using the per-hole-rank table of group
// elements that minimise the hole pattern.  Among them the down pattern is
// compared in increasing p, first minimum wins, which is the element the
// reference's orbit walk stores first.
__device__ inline Rep find_rep(const SectorView& s, uint32_t H, uint32_t Dn) {
    const int64_t hr = rank_holes(H, s.comb);
    uint32_t gm = s.canon_gmask[hr];
    const int blk = s.canon_block[hr];
    const int p0 = ctz32(gm);
    const uint32_t Hc = image_mask(H, p0, s.N, s.full);
    uint32_t best = image_mask(Dn, p0, s.N, s.full);
    int bp = p0;
    gm &= gm - 1;
    while (gm) {
        const int p = ctz32(gm);
        const uint32_t d = image_mask(Dn, p, s.N, s.full);
        if (d < best) {
            best = d;
            bp = p;
        }
        gm &= gm - 1;
    }
    return {Hc, best, bp, blk};
}

}  // namespace ed
