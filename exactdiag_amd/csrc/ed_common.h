// Shared host/device definitions for the gfx950 t-J ladder kernels.
//
// State encoding.  A ladder of L rungs has N = 2L sites, site i = 2x + y
// (x = rung along the legs, y = leg), as in the reference
// (tJ_spin_half_ladder/symmetry_ladder.h:96-111).  A state is two N-bit
// masks: H (bit i set = hole at site i) and Dn (bit i set = spin down).  The
// reference's state order (general/group/commutation_counter.h:45-69, hole
// pattern first, then down pattern, both compared from the top site down) is
// the numeric order of the 64-bit key (H << 32) | Dn, and equals the order of
// the reference's sparse index (tJ_spin_half_ladder/symmetry_ladder.h:346-361).
//
// Group elements.  p = 2m + b (m = rung translation, b = leg mirror) in the
// enumeration order of symmetry_ladder.h:157-176.  Image of a mask under p:
// rotate left by 2m inside N bits after an optional swap of the two legs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ed {

constexpr int MAX_N = 32;
constexpr int COMB_W = 33;  // binomial table row width: comb[k * COMB_W + n] = C(n, k)
constexpr uint32_t EVEN_BITS = 0x55555555u;

struct SectorView {
    int L, N, h, ne, nd;
    int ne_even;       // electron count even: the translation wrap sign applies
    uint32_t full;     // N low bits set
    int64_t D;         // dense dimension
    int64_t n_spin;    // C(ne, nd)
    int n_hole_ranks;  // C(N, h)
    int n_blocks;
    int any_k;         // 1: every lowest image is a basis state (no momentum compatibility test)
    int flip;          // spin-flip parity z = +1 or -1 (S^z = 0 only), 0 = no spin flip (the reference's group)
    int G;             // group elements: N, or 2 N with the spin flip (p + N = the flip after p)
    int NO;            // largest orbit size (N, or 2 N), the amplitude table's n range
    const uint32_t* comb;         // C(n, k) for n, k <= 32
    const uint32_t* canon_gmask;  // per hole rank: bit p set if image p gives the minimal hole mask
    const int32_t* canon_block;   // per hole rank: block of the minimal image (-1 = no basis states)
    const int32_t* block_of_rank; // per hole rank: block id when the rank itself is an eligible pattern
    const int64_t* block_base;    // per block: first dense index
    const int64_t* block_table;   // per block: offset into table (-1 = generic block, dense = base + spin rank)
    const uint32_t* block_hole;   // per block: hole mask
    const int32_t* table;         // spin rank -> dense index or -1, for blocks with a symmetric hole pattern
    const uint64_t* key;          // per dense index: (H << 32) | Dn of the representative
    const uint8_t* orbit;         // per dense index: orbit size n
    const double* amp;            // A[(p * (NO + 1) + n) * 2 + {re, im}] = F[p] / sqrt(n), p = 2m + b (+ N)
    const double* sqrt_tab;       // sqrt(n) for n = 0..NO, computed on the host
    const uint32_t* rank_tab;     // spin ranks by 8-bit chunks (RankGeom of ne, nd)
    const uint8_t* table_orbit;   // per table entry: orbit size of its state, 0 where table holds -1
};

__host__ __device__ inline uint32_t rotl_n(uint32_t x, int s, int N, uint32_t full) {
    if (s == 0) return x;
    uint64_t y = (uint64_t(x) << s) | (uint64_t(x) >> (N - s));
    return uint32_t(y) & full;
}

__host__ __device__ inline uint32_t leg_swap(uint32_t x) { return ((x & EVEN_BITS) << 1) | ((x >> 1) & EVEN_BITS); }

__host__ __device__ inline uint32_t image_mask(uint32_t x, int p, int N, uint32_t full) {
    const int m = p >> 1;
    if (p & 1) x = leg_swap(x);
    return rotl_n(x, 2 * m, N, full);
}

__host__ __device__ inline uint64_t make_key(uint32_t H, uint32_t Dn) { return (uint64_t(H) << 32) | Dn; }

// Down mask after flipping every spin of the state (holes stay).
__host__ __device__ inline uint32_t flip_downs(uint32_t H, uint32_t Dn, uint32_t full) { return ~H & full & ~Dn; }

// Workgroup w of a grid of nwg runs on XCD w % 8.  xcd_order gives the piece
// of work it takes: XCD x walks chunks x, x + 8, ... of `chunk` consecutive
// pieces, so that the eight XCDs work on neighbouring chunks at the same time
// (chunk <= 0 or >= nwg / 8: one contiguous eighth per XCD).  A bijection on
// [0, nwg).
__host__ __device__ inline int xcd_order(int w, int nwg, int chunk) {
    const int per = nwg / 8;
    if (per <= 0 || w >= per * 8) return w;
    const int x = w & 7, l = w >> 3;
    const int S = chunk > 0 && chunk < per ? chunk : per;
    const int full = per / S;
    if (l < full * S) return ((l / S) * 8 + x) * S + l % S;
    return full * S * 8 + x * (per - full * S) + (l - full * S);
}

__host__ __device__ inline int popc32(uint32_t x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __popc(x);
#else
    return __builtin_popcount(x);
#endif
}

__host__ __device__ inline int ctz32(uint32_t x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __ffs(int(x)) - 1;
#else
    return __builtin_ctz(x);
#endif
}

// Sign of image p relative to the state (H, .) as the reference accumulates
// it (symmetry_ladder.h:125-175): -1 per singly occupied rung wrapped from
// the top to the front when the electron count is even, and for mirrored
// images -1 per doubly occupied rung.
__host__ __device__ inline int image_sign(uint32_t H, int p, int L, int ne_even, uint32_t full) {
    const int m = p >> 1;
    int par = 0;
    if (ne_even && m > 0) {
        const uint32_t single = (H ^ (H >> 1)) & EVEN_BITS & full;
        const uint32_t top = full & ~((1u << (2 * (L - m))) - 1u);
        par += popc32(single & top);
    }
    if (p & 1) {
        const uint32_t paired = (~H) & (~H >> 1) & EVEN_BITS & full;
        par += popc32(paired);
    }
    return (par & 1) ? -1 : 1;
}

// Colex rank of the hole pattern: sum_k C(p_k, k + 1) (symmetry_ladder.h:346-361).
__host__ __device__ inline int64_t rank_holes(uint32_t H, const uint32_t* comb) {
    int64_t r = 0;
    int k = 1;
    while (H) {
        const int p = ctz32(H);
        r += comb[k * COMB_W + p];
        ++k;
        H &= H - 1;
    }
    return r;
}

// Colex rank of the down spins counted over non-hole sites only.
__host__ __device__ inline int64_t rank_spins(uint32_t H, uint32_t Dn, const uint32_t* comb) {
    int64_t r = 0;
    int k = 1;
    while (Dn) {
        const int p = ctz32(Dn);
        const int c = p - popc32(H & ((1u << p) - 1u));
        r += comb[k * COMB_W + c];
        ++k;
        Dn &= Dn - 1;
    }
    return r;
}

// Inverse of rank_spins for a fixed hole pattern (symmetry_ladder.h:314-344).
__host__ __device__ inline uint32_t unrank_spins(uint32_t H, int64_t sr, int nd, int N, const uint32_t* comb) {
    uint32_t Dn = 0;
    int rem = nd;
    int holes_below = popc32(H);
    for (int j = N - 1; j >= 0 && rem > 0; --j) {
        if (H & (1u << j)) {
            --holes_below;
            continue;
        }
        const int c = j - holes_below;
        const int64_t cost = comb[rem * COMB_W + c];
        if (sr >= cost) {
            Dn |= 1u << j;
            sr -= cost;
            --rem;
        }
    }
    return Dn;
}

// Complex product in the naive form used by std::complex / Cython, with
// every product rounded separately (callers compile with -ffp-contract=off).
struct C2 {
    double re, im;
};

__host__ __device__ inline C2 cmul(C2 a, C2 b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

__host__ __device__ inline C2 cadd(C2 a, C2 b) { return {a.re + b.re, a.im + b.im}; }

}  // namespace ed
