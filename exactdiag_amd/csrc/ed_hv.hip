// Matrix-free H·v with the sector tables staged in LDS (gfx950).
//
// Gather form y_f = sum_j conj(H_jf) x_j, one lane per row, as the other
// H·v kernels, but organised for the CDNA4 memory system:
//  * every table the element computation touches (binomials, amplitude
//    table, sqrt(n), per-hole-pattern group elements and blocks, block
//    offsets) is copied once per workgroup into LDS, so the inner loop issues
//    LDS reads instead of scattered vector-L1 loads;
//  * workgroups are persistent: each walks a contiguous range of row tiles,
//    so the table copy is amortised and the dot-product partials shrink to
//    one per workgroup; workgroup w is placed so consecutive workgroups on
//    one XCD (round-robin dispatch over 8 XCDs) take neighbouring ranges;
//  * ranks are 32-bit; an exchange on a hole pattern without stabiliser
//    gets its target as row + window rank change (no full rank);
//  * gathers of x_j are issued four at a time before any is consumed, so
//    four loads per lane are in flight.

#include <algorithm>

#include "ed_internal.h"
#include "ed_matrix.h"
#include "ed_orbit.h"

namespace ed {
namespace {

struct LdsLayout {
    int comb, amp, sq, gm, cb, bb, bt, bytes;
};

LdsLayout lds_layout(const ed_sector* s) {
    LdsLayout l{};
    int off = 0;
    auto take = [&](int bytes) {
        const int at = off;
        off += (bytes + 15) & ~15;
        return at;
    };
    l.comb = take(COMB_W * COMB_W * 4);
    l.amp = take(s->L * 2 * (s->N + 1) * 16);
    l.sq = take((s->N + 1) * 8);
    l.gm = take(int(s->n_hole) * 4);
    l.cb = take(int(s->n_hole) * 4);
    l.bb = take(int(s->block_rank.size()) * 8);
    l.bt = take(int(s->block_rank.size()) * 8);
    l.bytes = off;
    return l;
}

constexpr int LDS_CAP = 60 * 1024;

struct Tabs {
    const uint32_t* comb;
    const double* amp;
    const double* sq;
    const uint32_t* gm;
    const int32_t* cb;
    const int64_t* bb;
    const int64_t* bt;
};

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
};

__device__ inline C2 weighted_conj(const Tabs& T, int N, int L, int ne_even, uint32_t full, uint32_t Hsrc, int p,
                                   int nj, double scale, C2 W) {
    const int idx = ((p >> 1) * 2 + (p & 1)) * (N + 1) + nj;
    const double ar = T.amp[2 * idx], ai = T.amp[2 * idx + 1];
    const double sg = double(image_sign(Hsrc, p, L, ne_even, full)) * scale;
    return cmul(W, C2{sg * ar, -(sg * ai)});
}

// Lowest image among the group elements gm (hole part already minimal),
// then its row: generic blocks by rank, symmetric blocks by table.
__device__ inline bool locate(const SectorView& s, const Tabs& T, uint32_t Hsrc, uint32_t D2, uint32_t gm, int blk,
                              int& p_out, int64_t& j_out, int& nj_out) {
    int bp = ctz32(gm);
    const uint32_t Hc = image_mask(Hsrc, bp, s.N, s.full);
    uint32_t best = image_mask(D2, bp, s.N, s.full);
    gm &= gm - 1;
    while (gm) {
        const int p = ctz32(gm);
        const uint32_t d = image_mask(D2, p, s.N, s.full);
        if (d < best) {
            best = d;
            bp = p;
        }
        gm &= gm - 1;
    }
    const int32_t sr = rank_spins32(Hc, best, T.comb);
    const int64_t off = T.bt[blk];
    p_out = bp;
    if (off < 0) {
        j_out = T.bb[blk] + sr;
        nj_out = s.N;
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

template <class VT>
__device__ inline void flush4(const VT* __restrict__ x, const Elem (&e)[4], C2& acc) {
    C2 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = ldx(x, e[k].j);
#pragma unroll
    for (int k = 0; k < 4; ++k) acc = cadd(acc, cmul(e[k].c, v[k]));
}

template <class VT, int BS>
__global__ void __launch_bounds__(BS) hv_lds_kernel(SectorView s, HvWeights w, const VT* __restrict__ x,
                                                    VT* __restrict__ y, double* __restrict__ partial, LdsLayout lay,
                                                    int64_t ntiles, int64_t tiles_per_wg) {
    extern __shared__ __align__(16) unsigned char lds[];
    uint32_t* s_comb = reinterpret_cast<uint32_t*>(lds + lay.comb);
    double* s_amp = reinterpret_cast<double*>(lds + lay.amp);
    double* s_sq = reinterpret_cast<double*>(lds + lay.sq);
    uint32_t* s_gm = reinterpret_cast<uint32_t*>(lds + lay.gm);
    int32_t* s_cb = reinterpret_cast<int32_t*>(lds + lay.cb);
    int64_t* s_bb = reinterpret_cast<int64_t*>(lds + lay.bb);
    int64_t* s_bt = reinterpret_cast<int64_t*>(lds + lay.bt);
    {
        const int t = threadIdx.x;
        for (int i = t; i < COMB_W * COMB_W; i += BS) s_comb[i] = s.comb[i];
        const int namp = s.L * 2 * (s.N + 1) * 2;
        for (int i = t; i < namp; i += BS) s_amp[i] = s.amp[i];
        for (int i = t; i <= s.N; i += BS) s_sq[i] = s.sqrt_tab[i];
        for (int i = t; i < s.n_hole_ranks; i += BS) {
            s_gm[i] = s.canon_gmask[i];
            s_cb[i] = s.canon_block[i];
        }
        for (int i = t; i < s.n_blocks; i += BS) {
            s_bb[i] = s.block_base[i];
            s_bt[i] = s.block_table[i];
        }
    }
    __syncthreads();
    const Tabs T{s_comb, s_amp, s_sq, s_gm, s_cb, s_bb, s_bt};
    // neighbouring row ranges for consecutive workgroups of one XCD
    const int nwg = gridDim.x;
    int wg = blockIdx.x;
    const int per = nwg / 8;
    if (per > 0 && wg < per * 8) wg = (wg & 7) * per + (wg >> 3);
    const int64_t t0 = int64_t(wg) * tiles_per_wg;
    const int64_t t1 = t0 + tiles_per_wg < ntiles ? t0 + tiles_per_wg : ntiles;
    const int N = s.N, L = s.L, ne_even = s.ne_even;
    const uint32_t full = s.full;
    C2 dot{0.0, 0.0};
    for (int64_t tile = t0; tile < t1; ++tile) {
        const int64_t f = tile * BS + threadIdx.x;
        if (f >= s.D) continue;
        const uint64_t key = s.key[f];
        const uint32_t H = uint32_t(key >> 32), Dn = uint32_t(key);
        const uint32_t E = ~H & full, Up = E & ~Dn;
        const double sqf = T.sq[s.orbit[f]];
        const C2 xf = ldx(x, f);
        const int32_t hr = rank_holes32(H, T.comb);
        const uint32_t gm_self = T.gm[hr];
        const int blk_self = T.cb[hr];
        const bool generic = (gm_self & (gm_self - 1u)) == 0u;
        C2 acc{0.0, 0.0};
        // an exchange element: generic blocks by window rank change, others by the stabiliser search
        auto exchange = [&](uint32_t D2, int lo, int hi, bool wrap, C2 W, Elem& e) {
            if (generic) {
                e.j = wrap ? T.bb[blk_self] + rank_spins32(H, D2, T.comb) : f + window_delta32(H, Dn, D2, lo, hi, T.comb);
                e.c = weighted_conj(T, N, L, ne_even, full, H, 0, N, 0.5 * sqf, W);
                return;
            }
            int p, nj;
            int64_t j;
            if (locate(s, T, H, D2, gm_self, blk_self, p, j, nj)) {
                e.j = j;
                e.c = weighted_conj(T, N, L, ne_even, full, H, p, nj, 0.5 * sqf, W);
            }
        };
        if (w.active & 4) {  // jl: bonds (i, i + 2)
            const uint32_t d2 = rotl_n(Dn, N - 2, N, full), u2 = rotl_n(Up, N - 2, N, full);
            const uint32_t anti = N == 2 ? 0u : (((Dn & u2) | (Up & d2)) & full);
            const C2 W = w.w[2];
            const double diag = -0.5 * double(popc32(anti));
            acc = cadd(acc, cmul(W, C2{diag * xf.re, diag * xf.im}));
            for (int g = 0; g < N; g += 4) {
                Elem e[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    e[k] = Elem{f, C2{0.0, 0.0}};
                    const int i = g + k;
                    if (i < N && ((anti >> i) & 1u)) {
                        const int kk = i + 2 < N ? i + 2 : i + 2 - N;
                        const bool wrap = kk < i;
                        exchange(Dn ^ (1u << i) ^ (1u << kk), wrap ? kk : i, wrap ? i : kk, wrap, W, e[k]);
                    }
                }
                flush4(x, e, acc);
            }
        }
        if (w.active & 8) {  // jr: bonds (2x, 2x + 1)
            const uint32_t anti = (Dn ^ (Dn >> 1)) & E & (E >> 1) & EVEN_BITS & full;
            const C2 W = w.w[3];
            const double diag = -0.5 * double(popc32(anti));
            acc = cadd(acc, cmul(W, C2{diag * xf.re, diag * xf.im}));
            for (int g = 0; g < L; g += 4) {
                Elem e[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    e[k] = Elem{f, C2{0.0, 0.0}};
                    const int i = 2 * (g + k);
                    if (g + k < L && ((anti >> i) & 1u)) exchange(Dn ^ (3u << i), i, i + 1, false, W, e[k]);
                }
                flush4(x, e, acc);
            }
        }
        if (w.active & 3) {  // hops onto each hole: leg neighbours (tl) and rung partner (tr)
            uint32_t holes = H;
            while (holes) {
                const int t = ctz32(holes);
                holes &= holes - 1;
                Elem e[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    e[k] = Elem{f, C2{0.0, 0.0}};
                    if (k == 3) continue;
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
                    if (locate(s, T, H2, D2, T.gm[hr2], blk, p, j, nj)) {
                        e[k].j = j;
                        e[k].c = weighted_conj(T, N, L, ne_even, full, H2, p, nj, par ? sqf : -sqf, W);
                    }
                }
                flush4(x, e, acc);
            }
        }
        VT o;
        o.x = decltype(o.x)(acc.re);
        o.y = decltype(o.y)(acc.im);
        y[f] = o;
        dot = cadd(dot, cmul(C2{xf.re, -xf.im}, acc));
    }
    if (partial) {
        __shared__ double red[2 * BS];
        red[2 * threadIdx.x] = dot.re;
        red[2 * threadIdx.x + 1] = dot.im;
        __syncthreads();
        for (int st = BS / 2; st > 0; st >>= 1) {
            if (int(threadIdx.x) < st) {
                red[2 * threadIdx.x] += red[2 * (threadIdx.x + st)];
                red[2 * threadIdx.x + 1] += red[2 * (threadIdx.x + st) + 1];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            partial[2 * wg] = red[0];
            partial[2 * wg + 1] = red[1];
        }
    }
}

}  // namespace

// Persistent LDS-table H·v; returns the number of partials written, or -1
// when the sector's tables do not fit in LDS (caller falls back).
int64_t launch_hv_lds(const ed_sector* s, const HvWeights& w, const void* x, void* y, double* partial, int precision,
                      hipStream_t st) {
    constexpr int BS = 256;
    const LdsLayout lay = lds_layout(s);
    if (lay.bytes > LDS_CAP) return -1;
    const int64_t ntiles = (s->D + BS - 1) / BS;
    if (ntiles == 0) return 0;
    int dev = 0;
    ED_HIP(hipGetDevice(&dev));
    static int cus = 0;
    if (cus == 0) ED_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    // enough workgroups to fill every CU several times over, each with >= 4 tiles
    const int64_t target = int64_t(cus) * 8;
    const int64_t tpw = std::max<int64_t>(4, (ntiles + target - 1) / target);
    const int64_t nwg = (ntiles + tpw - 1) / tpw;
    if (precision == 0)
        hipLaunchKernelGGL((hv_lds_kernel<double2, BS>), dim3(unsigned(nwg)), dim3(BS), lay.bytes, st, s->view(), w,
                           static_cast<const double2*>(x), static_cast<double2*>(y), partial, lay, ntiles, tpw);
    else
        hipLaunchKernelGGL((hv_lds_kernel<float2, BS>), dim3(unsigned(nwg)), dim3(BS), lay.bytes, st, s->view(), w,
                           static_cast<const float2*>(x), static_cast<float2*>(y), partial, lay, ntiles, tpw);
    ED_HIP(hipGetLastError());
    return nwg;
}

}  // namespace ed
