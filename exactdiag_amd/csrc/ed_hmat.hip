// Stored Hamiltonian for repeated H·v (gfx950).
//
// The reference builds H once per sector as four upper-triangle CSR matrices
// with int16 indices into a table of unique values (dedup at 1e-6,
// general/cython_sparse_matrices.pyx:58-241) and applies them with
// sparse_matrix_dot (general/cpp_sparse_dot.h:59-138) at every Lanczos step.
// Here the matrix is also built once, on the GPU, for one set of term
// weights, and laid out for 64-wide wavefronts:
//  * gather form (row f holds conj(H_jf) for every j, all four terms and the
//    diagonal), so the product is written once per row, no atomics;
//  * sliced ELL with slices of 64 consecutive rows: element k of the rows of
//    slice s sits at sptr[s] + 64 k + lane, so the 64 lanes of a wavefront
//    read consecutive words (coalesced) at every k;
//  * per element a column and a code = index into a table of exact unique
//    values plus a negate bit: packed into one 32-bit word when the column
//    and the code fit (4 bytes per element, e.g. 23 + 9 bits for the 2x12
//    sector), else an int32 column and a uint16 code (6 bytes); per row the
//    element count (uint8) and the diagonal's value index (uint16);
//  * the common factor sqrt(n_f) of row f's off-diagonal elements is applied
//    once per row (from the orbit table), so the value table holds
//    W_t s_t conj(A[p][n_j]) per (term, group element, target orbit size),
//    s_t = 1 for hops and 1/2 for exchange.
// The elements come from the matrix-free walker (ed_hvrow.h), in the same
// order, in two passes (count, fill).

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <map>
#include <type_traits>
#include <utility>
#include <vector>

#include "ed_hmatrow.h"
#include "ed_hvrow.h"

namespace ed {
namespace {

constexpr int BS = 256;

struct CountSink {
    int n;
    int djl, djr;

    __device__ inline void diag(int term, double d) {
        const int c = int(-2.0 * d + 0.5);
        if (term == 2)
            djl = c;
        else
            djr = c;
    }

    template <int G>
    __device__ inline void group(const Elem (&e)[G]) {
#pragma unroll
        for (int k = 0; k < G; ++k) n += e[k].code != NO_ELEM;
    }
};

struct FillSink {
    int32_t* col;         // 6-byte form: column ...
    uint16_t* code;       // ... and value index | negate << 15
    uint32_t* pack;       // 4-byte form: column << shift | value index << 1 | negate
    int shift;            // 0: 6-byte form
    const uint16_t* map;  // combo -> value index
    int64_t base;         // sptr[slice] + lane
    int k;

    __device__ inline void diag(int, double) {}

    template <int G>
    __device__ inline void group(const Elem (&e)[G]) {
#pragma unroll
        for (int q = 0; q < G; ++q) {
            if (e[q].code == NO_ELEM) continue;
            const int64_t at = base + int64_t(k) * 64;
            const uint32_t vi = map[e[q].code & 0x7fffu];
            const uint32_t neg = (e[q].code >> 15) & 1u;
            if (shift) {
                pack[at] = (uint32_t(e[q].j) << shift) | (vi << 1) | neg;
            } else {
                col[at] = int32_t(e[q].j);
                code[at] = uint16_t(vi | (neg << 15));
            }
            ++k;
        }
    }
};

__global__ void __launch_bounds__(BS) hm_count_kernel(SectorView s, HvWeights w, LdsLayout lay, int lds,
                                                      const uint16_t* __restrict__ diag_map, uint8_t* cnt,
                                                      uint16_t* diag) {
    extern __shared__ __align__(16) unsigned char smem[];
    const Tabs T = lds ? stage_tabs<BS>(s, lay, smem) : global_tabs(s);
    __syncthreads();
    const int64_t f = int64_t(blockIdx.x) * BS + threadIdx.x;
    if (f >= s.D) return;
    CountSink c{0, 0, 0};
    walk_row<4>(s, T, w, f, c);
    cnt[f] = uint8_t(c.n);
    diag[f] = diag_map[c.djl * (s.L + 1) + c.djr];
}

// width[slice] = largest element count of its 64 rows (low 16 bits) and
// their sum (high 16 bits: at most 64 x 255)
__global__ void __launch_bounds__(BS) hm_width_kernel(const uint8_t* cnt, int64_t D, int32_t* width, int64_t nslices) {
    const int64_t sl = int64_t(blockIdx.x) * (BS / 64) + threadIdx.x / 64;
    const int64_t f = sl * 64 + (threadIdx.x & 63);
    int c = f < D ? int(cnt[f]) : 0;
    int t = c;
    for (int off = 32; off > 0; off >>= 1) {
        c = max(c, __shfl_xor(c, off, 64));
        t += __shfl_xor(t, off, 64);
    }
    if ((threadIdx.x & 63) == 0 && sl < nslices) width[sl] = c | (t << 16);
}

__global__ void __launch_bounds__(BS) hm_fill_kernel(SectorView s, HvWeights w, LdsLayout lay, int lds,
                                                     const uint16_t* __restrict__ map,
                                                     const int64_t* __restrict__ sptr, int32_t* col, uint16_t* code,
                                                     uint32_t* pack, int shift) {
    extern __shared__ __align__(16) unsigned char smem[];
    const Tabs T = lds ? stage_tabs<BS>(s, lay, smem) : global_tabs(s);
    __syncthreads();
    const int64_t f = int64_t(blockIdx.x) * BS + threadIdx.x;
    if (f >= s.D) return;
    FillSink sink{col, code, pack, shift, map, sptr[f >> 6] + (f & 63), 0};
    walk_row<4>(s, T, w, f, sink);
}

// present[n] = 1 for every orbit size n in the basis.
__global__ void __launch_bounds__(BS) hm_orbit_sizes_kernel(const uint8_t* __restrict__ orbit, int64_t D, int NO,
                                                            int* present) {
    for (int64_t f = int64_t(blockIdx.x) * BS + threadIdx.x; f < D; f += int64_t(gridDim.x) * BS) {
        const int n = orbit[f];
        if (n <= NO && !present[n]) present[n] = 1;
    }
}

// Order each row's stored elements so that the 64 lanes of a wavefront, which
// read slot k of 64 consecutive rows f, gather x[f + d] with one offset d where
// possible: their addresses are then consecutive.  One wavefront per 64-row
// slice, one lane per row.  ALIGNED (full slices whose rows hold at most 32
// elements): every (offset d = column - f, code) pair of row 0 that all 64
// rows hold is ranked by its place in row 0, and each row puts its elements
// with those pairs first, in that order (exchange on bonds above the spins that
// vary within the slice has the same rank change in every row), then the rest
// by column (a pair that row 0 holds twice counts once); hm_uniform_kernel
// then finds the shared leading slots.  Otherwise
// the rows are sorted by column alone.  A row of up to 32 elements is sorted
// by a bitonic network in registers; a longer one (0.07 % of the rows of the
// 2x12, 2-hole sector) is insertion-sorted in place in HBM.  The key is
// rank << 32 | packed word, or rank << 48 | column << 16 | code for the
// two-array form (rank 255 for elements sorted by column).
template <bool PACKED, bool ALIGNED>
__global__ void __launch_bounds__(64) hm_sort_kernel(int64_t D, const int64_t* __restrict__ sptr,
                                                     const uint8_t* __restrict__ cnt, int32_t* col, uint16_t* code,
                                                     uint32_t* pack, int shift) {
    constexpr int HB = PACKED ? 32 : 48;  // where the rank sits above the word
    const uint64_t LOW = (uint64_t(1) << HB) - 1u;
    const int lane = threadIdx.x;
    const int64_t f = int64_t(blockIdx.x) * 64 + lane;
    const bool live = f < D;
    const int64_t base = sptr[blockIdx.x] + lane;
    const int n = live ? int(cnt[f]) : 0;
    auto word_at = [&](int k) -> uint64_t {  // packed word, or column << 16 | code
        const int64_t at = base + int64_t(k) * 64;
        if constexpr (PACKED) return pack[at];
        else return (uint64_t(uint32_t(col[at])) << 16) | code[at];
    };
    auto put = [&](int k, uint64_t key) {
        const int64_t at = base + int64_t(k) * 64;
        if constexpr (PACKED) {
            pack[at] = uint32_t(key & LOW);
        } else {
            col[at] = int32_t(uint32_t((key >> 16) & 0xffffffffu));
            code[at] = uint16_t(key & 0xffffu);
        }
    };
    // (offset, code) of a word for row `row`
    auto offset_of = [&](uint64_t w, int64_t row) -> int32_t {
        if constexpr (PACKED) return int32_t(int64_t(uint32_t(w) >> shift) - row);
        else return int32_t(int64_t(w >> 16) - row);
    };
    auto code_of = [&](uint64_t w) -> uint32_t {
        if constexpr (PACKED) return uint32_t(w) & ((1u << shift) - 1u);
        else return uint32_t(w & 0xffffu);
    };
    int nmax = n;
    for (int o = 32; o > 0; o >>= 1) nmax = max(nmax, __shfl_xor(nmax, o, 64));
    const bool align = ALIGNED && nmax <= 32 && int64_t(blockIdx.x) * 64 + 64 <= D;
    if (n <= 32 || (align && live)) {
        // up to 32 keys: a bitonic network in registers (slots past n padded with the largest key)
        uint64_t r[32];
#pragma unroll
        for (int k = 0; k < 32; ++k) r[k] = k < n ? ((uint64_t(255) << HB) | word_at(k)) : ~uint64_t(0);
        if (align) {
            // row 0's (offset, code) pairs, the candidates, in a 64-slot LDS hash table (key -> place in row 0;
            // a pair row 0 holds twice keeps its first place); every lane then looks its own pairs up, and the
            // candidates that all 64 rows hold (the AND of the lanes' masks) get their place as rank
            __shared__ uint64_t hkey[64];
            __shared__ int hidx[64];
            auto pair_key = [&](uint64_t w, int64_t row) -> uint64_t {
                return (uint64_t(uint32_t(offset_of(w, row))) << 32) | code_of(w);
            };
            auto slot_of = [](uint64_t key) -> int {
                const uint32_t hsh = uint32_t(key >> 32) * 0x9E3779B1u ^ uint32_t(key) * 0x85EBCA6Bu;
                return int(hsh >> 26);
            };
            hidx[lane] = -1;
            hkey[lane] = 0;
            __syncthreads();
            const int64_t f0 = int64_t(blockIdx.x) * 64;
            if (lane == 0) {
#pragma unroll
                for (int k = 0; k < 32; ++k) {
                    if (k < n) {
                        const uint64_t key = pair_key(r[k] & LOW, f0);
                        int h = slot_of(key);
                        while (hidx[h] >= 0 && hkey[h] != key) h = (h + 1) & 63;
                        if (hidx[h] < 0) {
                            hkey[h] = key;
                            hidx[h] = k;
                        }
                    }
                }
            }
            __syncthreads();
            uint32_t have = 0;
            int cand[32];
#pragma unroll
            for (int k = 0; k < 32; ++k) {
                cand[k] = -1;
                if (k < n) {
                    const uint64_t key = pair_key(r[k] & LOW, f);
                    int h = slot_of(key);
                    int c = -1;
                    while (hidx[h] >= 0) {  // at most 32 of the 64 slots are taken, so an empty one ends the probe
                        if (hkey[h] == key) {
                            c = hidx[h];
                            break;
                        }
                        h = (h + 1) & 63;
                    }
                    if (c >= 0 && !((have >> c) & 1u)) {
                        have |= 1u << c;
                        cand[k] = c;
                    }
                }
            }
            uint32_t all = have;
            for (int o = 32; o > 0; o >>= 1) all &= __shfl_xor(all, o, 64);
#pragma unroll
            for (int k = 0; k < 32; ++k)
                if (cand[k] >= 0 && ((all >> cand[k]) & 1u)) r[k] = (uint64_t(cand[k]) << HB) | (r[k] & LOW);
        }
#pragma unroll
        for (int size = 2; size <= 32; size <<= 1) {
#pragma unroll
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
#pragma unroll
                for (int i = 0; i < 32; ++i) {
                    const int j = i ^ stride;
                    if (j > i) {
                        const uint64_t a = r[i], b = r[j];
                        const bool up = (i & size) == 0;
                        r[i] = up ? (a < b ? a : b) : (a < b ? b : a);
                        r[j] = up ? (a < b ? b : a) : (a < b ? a : b);
                    }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 32; ++k)
            if (k < n) put(k, r[k]);
    } else {
        for (int k = 1; k < n; ++k) {
            const uint64_t v = word_at(k);
            int q = k - 1;
            while (q >= 0 && word_at(q) > v) {
                put(q + 1, word_at(q));
                --q;
            }
            put(q + 1, v);
        }
    }
}

// After the ALIGNED sort: u[s] = the number of leading slots k at which every
// row of the full slice s holds the same (column offset, code); of each row's
// remaining elements the local ones have |column - row| <= radius.  lwid[s] =
// the most local elements of a row, width2[s] = lwid[s] | (the most far ones)
// << 8 | (sum of the remaining counts) << 16.  One wavefront per slice.  A
// slice with fewer than 64 rows keeps all of its elements per row.
template <bool PACKED>
__global__ void __launch_bounds__(64) hm_uniform_kernel(int64_t D, const int64_t* __restrict__ sptr,
                                                        const uint8_t* __restrict__ cnt, const int32_t* __restrict__ col,
                                                        const uint16_t* __restrict__ code,
                                                        const uint32_t* __restrict__ pack, int shift, int radius,
                                                        int32_t* ucount, uint8_t* lwid, int32_t* width2) {
    const int lane = threadIdx.x;
    const int64_t s = blockIdx.x;
    const int64_t f = s * 64 + lane;
    const bool full = s * 64 + 64 <= D;
    const int n = f < D ? int(cnt[f]) : 0;
    const int64_t base = sptr[s] + lane;
    auto elem = [&](int k, int64_t& c, uint32_t& cd) {
        const int64_t at = base + int64_t(k) * 64;
        if constexpr (PACKED) {
            const uint32_t w = pack[at];
            c = int64_t(w >> shift);
            cd = w & ((1u << shift) - 1u);
        } else {
            c = col[at];
            cd = code[at];
        }
    };
    int nmin = n;
    for (int o = 32; o > 0; o >>= 1) nmin = min(nmin, __shfl_xor(nmin, o, 64));
    int u = 0;
    if (full) {
        for (; u < nmin; ++u) {
            int64_t c;
            uint32_t cd;
            elem(u, c, cd);
            const int32_t d = int32_t(c - f);
            const bool same = d == __shfl(d, 0, 64) && cd == uint32_t(__shfl(int(cd), 0, 64));
            if (!__all(same)) break;
        }
    }
    int nl = 0;
    for (int k = u; k < n; ++k) {
        int64_t c;
        uint32_t cd;
        elem(k, c, cd);
        nl += (c - f <= radius && f - c <= radius) ? 1 : 0;
    }
    int nfar = n - u - nl, wl = nl, wf = nfar, sum = n - u;
    for (int o = 32; o > 0; o >>= 1) {
        wl = max(wl, __shfl_xor(wl, o, 64));
        wf = max(wf, __shfl_xor(wf, o, 64));
        sum += __shfl_xor(sum, o, 64);
    }
    if (lane == 0) {
        ucount[s] = u;
        lwid[s] = uint8_t(wl);
        width2[s] = wl | (wf << 8) | (sum << 16);
    }
}

// Move the uniform slots of slice s into uni[uptr[s] ..] (from row s * 64),
// each row's remaining local elements into slots [0, lwid[s]) and its far
// ones into the slots after them of the new sliced-ELL arrays, keeping their
// order; cnt[f] becomes the far count and lcnt[f] the local one.
template <bool PACKED>
__global__ void __launch_bounds__(64) hm_compact_kernel(int64_t D, const int64_t* __restrict__ sptr,
                                                        const int64_t* __restrict__ sptr2,
                                                        const int64_t* __restrict__ uptr,
                                                        const uint8_t* __restrict__ lwid, int radius, uint8_t* cnt,
                                                        uint8_t* lcnt, const int32_t* __restrict__ col,
                                                        const uint16_t* __restrict__ code,
                                                        const uint32_t* __restrict__ pack, int shift, int32_t* col2,
                                                        uint16_t* code2, uint32_t* pack2, uint2* uni) {
    const int lane = threadIdx.x;
    const int64_t s = blockIdx.x;
    const int64_t f = s * 64 + lane;
    const int u = int(uptr[s + 1] - uptr[s]);
    for (int k = lane; k < u; k += 64) {
        const int64_t at = sptr[s] + int64_t(k) * 64;  // row s * 64 (lane 0 of the slice)
        uint32_t d, cd;
        if constexpr (PACKED) {
            const uint32_t w = pack[at];
            d = uint32_t(int32_t(int64_t(w >> shift) - s * 64));
            cd = w & ((1u << shift) - 1u);
        } else {
            d = uint32_t(int32_t(int64_t(col[at]) - s * 64));
            cd = code[at];
        }
        uni[uptr[s] + k] = uint2{d, cd};
    }
    if (f >= D) return;
    const int n = cnt[f];
    const int64_t from = sptr[s] + lane, to = sptr2[s] + lane;
    int il = 0, ifar = int(lwid[s]);
    for (int k = u; k < n; ++k) {
        const int64_t src = from + int64_t(k) * 64;
        int64_t c;
        if constexpr (PACKED)
            c = int64_t(pack[src] >> shift);
        else
            c = col[src];
        const bool local = c - f <= radius && f - c <= radius;
        const int64_t dst = to + int64_t(local ? il++ : ifar++) * 64;
        if constexpr (PACKED) {
            pack2[dst] = pack[src];
        } else {
            col2[dst] = col[src];
            code2[dst] = code[src];
        }
    }
    cnt[f] = uint8_t(ifar - int(lwid[s]));
    lcnt[f] = uint8_t(il);
}

// The product over all rows (stored_row, ed_hmatrow.h).  Persistent
// workgroups walk `tiles` consecutive 256-row tiles (consecutive workgroups of
// one XCD take neighbouring ranges, so their gathers share that XCD's L2);
// with NT the output is stored non-temporally as well.  WIN: for each tile,
// x over the tile's rows and m.radius rows either side is staged in LDS
// first, and the local elements read it there.  partial[workgroup] = sum
// conj(x_f) y_f over its rows.
template <class VT, int U, bool NT, bool PACKED, bool WIN>
__global__ void __launch_bounds__(BS) hm_spmv_kernel(StoredView m, int64_t D, const VT* __restrict__ x,
                                                     VT* __restrict__ y, double* __restrict__ partial,
                                                     int64_t tiles, int chunk) {
    extern __shared__ __align__(16) unsigned char smem[];
    double* V = reinterpret_cast<double*>(smem);
    VT* win = reinterpret_cast<VT*>(smem + ((size_t(m.n_vals) * 16 + 15) / 16) * 16);
    for (int i = threadIdx.x; i < 2 * m.n_vals; i += BS) V[i] = m.vals[i];
    __syncthreads();
    const int nwg = gridDim.x;
    int wg = blockIdx.x;
    wg = xcd_order(wg, nwg, chunk);
    const int64_t f0 = int64_t(wg) * tiles * BS;
    const int64_t f1 = f0 + tiles * BS < D ? f0 + tiles * BS : D;
    C2 dot{0.0, 0.0};
    for (int64_t t0 = f0; t0 < f1; t0 += BS) {
        const int64_t w0 = t0 - m.radius;
        if constexpr (WIN) {
            const int64_t wlen = BS + 2 * int64_t(m.radius);
            for (int64_t i = threadIdx.x; i < wlen; i += BS) {
                const int64_t j = w0 + i;
                if (j >= 0 && j < D) win[i] = x[j];
            }
            __syncthreads();
        }
        const int64_t f = t0 + threadIdx.x;
        if (f < f1) {
            C2 xf;
            const C2 yf = stored_row<VT, U, NT, PACKED, WIN>(m, V, x, f, xf, win, w0);
            if constexpr (is_real_vec<VT>::value) {
                if constexpr (NT)
                    __builtin_nontemporal_store(VT(yf.re), y + f);
                else
                    y[f] = VT(yf.re);
                dot.re += xf.re * yf.re;
            } else {
                VT o;
                o.x = decltype(o.x)(yf.re);
                o.y = decltype(o.y)(yf.im);
                if constexpr (NT) {
                    using S = decltype(o.x);
                    typedef S S2 __attribute__((ext_vector_type(2)));
                    S2 v;
                    v.x = o.x;
                    v.y = o.y;
                    __builtin_nontemporal_store(v, reinterpret_cast<S2*>(y + f));
                } else {
                    y[f] = o;
                }
                dot = cadd(dot, cmul(C2{xf.re, -xf.im}, yf));
            }
        }
        if constexpr (WIN) __syncthreads();  // the next tile's window overwrites this one
    }
    if (partial) {
        workgroup_sum2<BS>(dot.re, dot.im);
        if (threadIdx.x == 0) {
            partial[2 * wg] = dot.re;
            partial[2 * wg + 1] = dot.im;
        }
    }
}

// Read-rate probe, the ceiling the stored product is measured against: every
// lane loads 16 bytes, either each from its own 128-byte line at a hashed
// place of a buffer of mask + 1 elements (SCATTER, as the far gathers do) or
// 64 consecutive elements per wavefront (eight whole lines).  Four loads are
// in flight per lane; the sum only keeps the loads from being removed.
__device__ inline uint32_t probe_hash(uint32_t a) {
    a ^= a >> 16;
    a *= 0x7feb352du;
    a ^= a >> 15;
    a *= 0x846ca68bu;
    a ^= a >> 16;
    return a;
}

template <bool SCATTER>
__global__ void __launch_bounds__(BS) line_probe_kernel(const double2* __restrict__ x, uint32_t mask, int iters,
                                                        double* __restrict__ out) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = (blockIdx.x * BS + threadIdx.x) >> 6;
    double acc = 0.0;
    for (int it = 0; it < iters; it += 4) {
        double2 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t seed = wave * 1315423911u + uint32_t(it + u) * 2654435761u;
            const uint32_t j = SCATTER ? (probe_hash(seed ^ (lane * 97u)) & mask)
                                       : ((probe_hash(seed) & mask) & ~63u) + lane;
            v[u] = x[j];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) acc += v[u].x + v[u].y;
    }
    if (acc == 1.0e300) out[0] = acc;  // never true for the zeroed buffer
}

uint64_t bits(double v) {
    uint64_t u;
    std::memcpy(&u, &v, 8);
    return u;
}

}  // namespace
}  // namespace ed

using namespace ed;

struct ed_hmat {
    const ed_sector* s = nullptr;
    HvWeights w{};
    int64_t D = 0, nslices = 0, slots = 0, nnz = 0;
    int n_vals = 0;
    int tiles = 0;   // row tiles per workgroup of the product (0 = default)
    int unroll = 4;  // gathers in flight per lane (4 or 8)
    int nontemporal = -1;  // stream elements and output non-temporally: 0 no, otherwise yes (-1: default, yes)
    int64_t* d_sptr = nullptr;
    int32_t* d_col = nullptr;    // 6-byte form
    uint16_t* d_code = nullptr;  // 6-byte form
    uint32_t* d_pack = nullptr;  // 4-byte form
    int shift = 0;               // bits of value index + sign in the 4-byte form; 0 = 6-byte form
    bool real = false;           // every stored value real: real vectors allowed
    uint8_t* d_cnt = nullptr;
    uint16_t* d_diag = nullptr;
    double* d_vals = nullptr;
    double* d_partial = nullptr;
    int64_t* d_uptr = nullptr;  // slice-uniform entries (nullptr: none)
    uint2* d_uni = nullptr;
    int64_t n_uni = 0;
    uint8_t* d_lcnt = nullptr;  // per row: local elements, read from the staged window (nullptr: none)
    uint8_t* d_lwid = nullptr;  // per slice: slots of the local segment
    int radius = 0;

    int element_bytes() const { return shift ? 4 : 6; }
    int64_t bytes() const {
        return slots * element_bytes() + D * 3 + (nslices + 1) * 8 + int64_t(n_vals) * 16 +
               (d_uptr ? (nslices + 1) * 8 + n_uni * 8 : 0) + (d_lcnt ? D + nslices : 0);
    }

    ~ed_hmat() {
        for (void* p : {(void*)d_sptr, (void*)d_col, (void*)d_code, (void*)d_pack, (void*)d_cnt, (void*)d_diag,
                        (void*)d_vals, (void*)d_partial, (void*)d_uptr, (void*)d_uni, (void*)d_lcnt, (void*)d_lwid})
            dev_free(p);
    }
};

namespace ed {

const ed_sector* hmat_sector(const ed_hmat* m) { return m->s; }

HvWeights hmat_weights(const ed_hmat* m) { return m->w; }

StoredView hmat_view(const ed_hmat* m) {
    return StoredView{m->d_sptr, m->d_col, m->d_code, m->d_pack, m->shift, m->n_vals, m->d_cnt, m->d_diag,
                      m->s->d_orbit, m->s->d_sqrt, m->d_vals, m->d_uptr, m->d_uni, m->d_lcnt, m->d_lwid,
                      m->d_lcnt ? m->radius : 0};
}

// One 256-row tile per workgroup: the fastest setting measured for the 2x12
// sector (ed_hmat_tune, product plus the reduction of its partials).
// Order of each row's elements after the fill (ED_HMAT_ROW_ORDER, read at
// ed_hmat_create): 0 keeps the walk order, 1 (default) sorts by column, 2
// puts offsets shared across the slice first (hm_sort_kernel).  On the 2x12,
// 2-hole sector the column sort took the product from 0.529 ms to 0.471 ms:
// the 64 lanes reading slot k gather from nearby columns more often.
int hmat_row_order() {
    const char* e = std::getenv("ED_HMAT_ROW_ORDER");
    const int v = e ? std::atoi(e) : 1;
    return v < 0 || v > 2 ? 1 : v;
}

int hmat_default_tiles(int64_t) { return 1; }

// Dynamic LDS of the product: the value table, then (radius >= 0) the window
// of 256 + 2 radius vector elements.
size_t hmat_lds_bytes(int n_vals, int radius, size_t vec_bytes) {
    const size_t vals = (size_t(n_vals) * 16 + 15) / 16 * 16;
    return radius < 0 ? vals : vals + (size_t(BS) + 2 * size_t(radius)) * vec_bytes;
}

// Reach of the local elements read from the staged window (ED_HMAT_WINDOW,
// read at ed_hmat_create; 0 stages none).
int hmat_window_radius() {
    const char* e = std::getenv("ED_HMAT_WINDOW");
    const int v = e ? std::atoi(e) : 128;
    return std::max(0, std::min(v, 4096));
}

// Slice-uniform entries kept once per slice (the default; ED_HMAT_UNIFORM=0,
// read at ed_hmat_create, keeps every element per row).
bool hmat_uniform() {
    const char* e = std::getenv("ED_HMAT_UNIFORM");
    return !(e && std::atoi(e) == 0);
}

// Moves the leading slots that every row of a slice shares into the per-slice
// table (hm_uniform_kernel, hm_compact_kernel) and rebuilds the sliced-ELL
// arrays for the remaining elements.
void extract_uniform(ed_hmat* m, hipStream_t st) {
    const int64_t D = m->D, ns = m->nslices;
    std::vector<void*> tmp;      // the new arrays, owned by the matrix once complete
    std::vector<void*> scratch;  // freed in any case
    auto cleanup = [&] {
        for (void* p : tmp) dev_free(p);
        for (void* p : scratch) dev_free(p);
        tmp.clear();
        scratch.clear();
    };
    try {
        int32_t* d_u = dev_alloc<int32_t>(size_t(ns));
        scratch.push_back(d_u);
        int32_t* d_w = dev_alloc<int32_t>(size_t(ns));
        scratch.push_back(d_w);
        uint8_t* d_lwid = dev_alloc<uint8_t>(size_t(ns));
        tmp.push_back(d_lwid);
        uint8_t* d_lcnt = dev_alloc<uint8_t>(size_t(std::max<int64_t>(D, 1)));
        tmp.push_back(d_lcnt);
        const int radius = m->radius;
        hipLaunchKernelGGL(m->shift ? hm_uniform_kernel<true> : hm_uniform_kernel<false>, dim3(unsigned(ns)), dim3(64),
                           0, st, D, m->d_sptr, m->d_cnt, m->d_col, m->d_code, m->d_pack, m->shift, radius, d_u,
                           d_lwid, d_w);
        ED_HIP(hipGetLastError());
        std::vector<int32_t> u(static_cast<size_t>(ns)), w(static_cast<size_t>(ns));
        ED_HIP(hipMemcpyAsync(u.data(), d_u, u.size() * 4, hipMemcpyDeviceToHost, st));
        ED_HIP(hipMemcpyAsync(w.data(), d_w, w.size() * 4, hipMemcpyDeviceToHost, st));
        ED_HIP(hipStreamSynchronize(st));
        std::vector<int64_t> uptr(static_cast<size_t>(ns + 1), 0), sptr2(static_cast<size_t>(ns + 1), 0);
        for (int64_t i = 0; i < ns; ++i) {
            uptr[size_t(i + 1)] = uptr[size_t(i)] + u[size_t(i)];
            sptr2[size_t(i + 1)] = sptr2[size_t(i)] + int64_t((w[size_t(i)] & 0xff) + ((w[size_t(i)] >> 8) & 0xff)) * 64;
        }
        m->n_uni = uptr.back();
        const int64_t slots2 = sptr2.back();
        int64_t* d_uptr = dev_upload(uptr, st);
        tmp.push_back(d_uptr);
        int64_t* d_sptr2 = dev_upload(sptr2, st);
        tmp.push_back(d_sptr2);
        uint2* d_uni = dev_alloc<uint2>(size_t(std::max<int64_t>(m->n_uni, 1)));
        tmp.push_back(d_uni);
        int32_t* col2 = nullptr;
        uint16_t* code2 = nullptr;
        uint32_t* pack2 = nullptr;
        if (m->shift) {
            pack2 = dev_alloc<uint32_t>(size_t(std::max<int64_t>(slots2, 1)));
            tmp.push_back(pack2);
        } else {
            col2 = dev_alloc<int32_t>(size_t(std::max<int64_t>(slots2, 1)));
            tmp.push_back(col2);
            code2 = dev_alloc<uint16_t>(size_t(std::max<int64_t>(slots2, 1)));
            tmp.push_back(code2);
        }
        hipLaunchKernelGGL(m->shift ? hm_compact_kernel<true> : hm_compact_kernel<false>, dim3(unsigned(ns)), dim3(64),
                           0, st, D, m->d_sptr, d_sptr2, d_uptr, d_lwid, radius, m->d_cnt, d_lcnt, m->d_col,
                           m->d_code, m->d_pack, m->shift, col2, code2, pack2, d_uni);
        ED_HIP(hipGetLastError());
        ED_HIP(hipStreamSynchronize(st));
        for (void* p : {(void*)m->d_sptr, (void*)m->d_col, (void*)m->d_code, (void*)m->d_pack}) dev_free(p);
        m->d_sptr = d_sptr2;
        m->d_col = col2;
        m->d_code = code2;
        m->d_pack = pack2;
        m->d_uptr = d_uptr;
        m->d_uni = d_uni;
        if (radius > 0) {
            m->d_lcnt = d_lcnt;
            m->d_lwid = d_lwid;
        } else {  // every remaining element counted far: the local arrays are all zero
            dev_free(d_lcnt);
            dev_free(d_lwid);
        }
        m->slots = slots2;
        tmp.clear();  // now owned by the matrix
        cleanup();
    } catch (...) {
        cleanup();
        throw;
    }
}

template <class VT, int U>
void launch_spmv(const ed_hmat* m, const void* x, void* y, double* partial, int64_t nwg, int64_t tiles,
                 hipStream_t st) {
    // the staged window pays for complex128 vectors only (2x12 sector: 0.393 ms with a radius of 128 against
    // 0.403 ms without; complex64 0.329 against 0.318 ms), so the other types gather their local elements
    const bool win = m->d_lcnt != nullptr && sizeof(VT) == 16;
    const size_t lds = hmat_lds_bytes(m->n_vals, win ? m->radius : -1, sizeof(VT));
    // non-temporal streaming by default: with the chunked XCD order it gains for both vector types on the
    // 2x12 sector (ed_hmat_tune: complex64 0.340 ms with it, 0.352 ms without)
    const bool nt = m->nontemporal != 0;
    const StoredView v = hmat_view(m);
    auto go = [&](auto kernel) {
        // XCD chunk of 512 workgroups (131,072 rows with one tile each): measured on the 2x12 sector, the
        // eight XCDs working on neighbouring chunks beat one contiguous eighth each by 10 % (complex128
        // 0.442 ms instead of 0.486 ms, complex64 0.352 instead of 0.388); ED_HMAT_XCD_CHUNK overrides it
        // (0 = one eighth each)
        static const int chunk = [] {
            const char* e = std::getenv("ED_HMAT_XCD_CHUNK");
            return e ? std::atoi(e) : 512;
        }();
        hipLaunchKernelGGL(kernel, dim3(unsigned(nwg)), dim3(BS), lds, st, v, m->D, static_cast<const VT*>(x),
                           static_cast<VT*>(y), partial, tiles, chunk);
    };
    auto pick = [&](auto w) {
        constexpr bool W = decltype(w)::value;
        if (m->shift) {
            if (nt)
                go(hm_spmv_kernel<VT, U, true, true, W>);
            else
                go(hm_spmv_kernel<VT, U, false, true, W>);
        } else {
            if (nt)
                go(hm_spmv_kernel<VT, U, true, false, W>);
            else
                go(hm_spmv_kernel<VT, U, false, false, W>);
        }
    };
    if (win)
        pick(std::true_type{});
    else
        pick(std::false_type{});
}

// Launches the product; returns the number of partials written (one per workgroup).
int64_t hmat_max_partials(const ed_hmat* m) { return std::max<int64_t>((m->D + BS - 1) / BS, 1); }

int64_t launch_hmat(const ed_hmat* m, const void* x, void* y, double* partial, int precision, hipStream_t st) {
    if (m->D == 0) return 0;
    const int64_t tiles = m->tiles > 0 ? m->tiles : hmat_default_tiles(m->D);
    const int64_t nwg = ((m->D + BS - 1) / BS + tiles - 1) / tiles;
    const bool u8 = m->unroll == 8;
    switch (precision) {
        case ED_PREC_C128:
            u8 ? launch_spmv<double2, 8>(m, x, y, partial, nwg, tiles, st)
               : launch_spmv<double2, 4>(m, x, y, partial, nwg, tiles, st);
            break;
        case ED_PREC_C64:
            u8 ? launch_spmv<float2, 8>(m, x, y, partial, nwg, tiles, st)
               : launch_spmv<float2, 4>(m, x, y, partial, nwg, tiles, st);
            break;
        case ED_PREC_R64:
            u8 ? launch_spmv<double, 8>(m, x, y, partial, nwg, tiles, st)
               : launch_spmv<double, 4>(m, x, y, partial, nwg, tiles, st);
            break;
        default:
            u8 ? launch_spmv<float, 8>(m, x, y, partial, nwg, tiles, st)
               : launch_spmv<float, 4>(m, x, y, partial, nwg, tiles, st);
            break;
    }
    ED_HIP(hipGetLastError());
    return nwg;
}

// True when every value of the stored table is real (to 1e-13 of the largest
// magnitude): H is then a real symmetric matrix in the sector's basis, which
// holds for momenta kx = 0 and kx = L/2.
bool hmat_is_real(const ed_hmat* m) { return m->real; }

void hmat_require_precision(const ed_hmat* m, int precision) {
    ED_REQUIRE(precision >= ED_PREC_C128 && precision <= ED_PREC_R32, ED_ERR_ARG, "precision must be 0, 1, 2 or 3");
    ED_REQUIRE(precision <= ED_PREC_C64 || m->real, ED_ERR_ARG,
               "real vectors need a real stored matrix (momentum kx = 0 or L/2)");
}

// Stored structure of sector s.  coded = 0: each element's code indexes the
// exact unique values (the product's table).  coded = 1: an element's code is
// term * 32 + p (the term and the group element mapping its target to the
// representative), the sign is kept, a row's diagonal code is
// jl_count * (L + 1) + jr_count (its antiparallel leg and rung bonds) and the
// value table holds 128 unused entries; values, momentum phases and
// orbit-size factors are left to the caller (the multi-momentum product).
ed_hmat* hmat_build(const ed_sector* s, const double* w, int coded, hipStream_t st) {
    ED_REQUIRE(s && s->D >= 0 && w, ED_ERR_ARG, "bad stored-matrix arguments");
    auto* m = new ed_hmat();
    std::vector<void*> tmp;
    auto cleanup_tmp = [&] {
        for (void* p : tmp) dev_free(p);
        tmp.clear();
    };
    try {
        m->s = s;
        m->w = make_weights(w);
        m->D = s->D;
        const int N = s->N, L = s->L;
        // exact unique values: off-diagonal combos, then diagonal (jl count, jr count) pairs
        std::vector<double> vals;
        std::map<std::pair<uint64_t, uint64_t>, int> uid;
        auto intern = [&](C2 v) {
            const auto key = std::make_pair(bits(v.re), bits(v.im));
            auto it = uid.find(key);
            if (it != uid.end()) return it->second;
            const int id = int(vals.size() / 2);
            vals.push_back(v.re);
            vals.push_back(v.im);
            uid.emplace(key, id);
            return id;
        };
        // element codes (elem_code): (term * G + p) * (NO + 1) + n_j, G group elements and orbits up to NO
        const int G = s->flip ? 2 * N : N, NO = G;
        const int ncombo = 4 * G * (NO + 1);
        std::vector<uint16_t> map(size_t(ncombo), 0);
        if (coded) {
            ED_REQUIRE(N <= 32 && !s->flip, ED_ERR_ARG,
                       "coded elements need at most 32 group elements (no spin flip)");
            vals.assign(2 * 128, 0.0);
            for (int t = 0; t < 4; ++t)
                for (int p = 0; p < N; ++p)
                    for (int nj = 0; nj <= N; ++nj) map[size_t((t * N + p) * (N + 1) + nj)] = uint16_t(t * 32 + p);
        }
        // orbit sizes a target can have: those of the basis states (NO for generic blocks); values of other
        // sizes are never addressed, and leaving them out keeps the table small enough for 4-byte elements
        std::vector<int> nj_present(size_t(NO) + 1, 0);
        nj_present[size_t(NO)] = 1;
        if (!coded && s->D > 0) {
            int* d_present = dev_alloc<int>(size_t(NO) + 1);
            tmp.push_back(d_present);
            ED_HIP(hipMemsetAsync(d_present, 0, (size_t(NO) + 1) * sizeof(int), st));
            hipLaunchKernelGGL(hm_orbit_sizes_kernel, dim3(unsigned(std::min<int64_t>((s->D + BS - 1) / BS, 4096))),
                               dim3(BS), 0, st, s->d_orbit, s->D, NO, d_present);
            ED_HIP(hipGetLastError());
            std::vector<int> present(size_t(NO) + 1, 0);
            ED_HIP(hipMemcpyAsync(present.data(), d_present, present.size() * sizeof(int), hipMemcpyDeviceToHost, st));
            ED_HIP(hipStreamSynchronize(st));
            for (int n = 0; n <= NO; ++n) nj_present[size_t(n)] |= present[size_t(n)];
        }
        for (int t = 0; t < 4 && !coded; ++t) {
            if (!(m->w.active & (1 << t))) continue;
            const double st_ = t < 2 ? 1.0 : 0.5;
            for (int p = 0; p < G; ++p)
                for (int nj = 0; nj <= NO; ++nj) {
                    if (!nj_present[size_t(nj)]) continue;
                    const int idx = p * (NO + 1) + nj;
                    const double ar = s->amp[2 * idx], ai = s->amp[2 * idx + 1];
                    map[size_t((t * G + p) * (NO + 1) + nj)] =
                        uint16_t(intern(cmul(m->w.w[t], C2{st_ * ar, -(st_ * ai)})));
                }
        }
        const int n_off = int(vals.size() / 2);  // off-diagonal values come first
        std::vector<uint16_t> dmap(size_t((N + 1) * (L + 1)), 0);
        for (int a = 0; a <= N; ++a)
            for (int b = 0; b <= L; ++b) {
                C2 d{0.0, 0.0};
                if (m->w.active & 4) d = cadd(d, cmul(m->w.w[2], C2{-0.5 * a, 0.0}));
                if (m->w.active & 8) d = cadd(d, cmul(m->w.w[3], C2{-0.5 * b, 0.0}));
                dmap[size_t(a * (L + 1) + b)] = coded ? uint16_t(a * (L + 1) + b) : uint16_t(intern(d));
            }
        m->n_vals = int(vals.size() / 2);
        ED_REQUIRE(m->n_vals < 0x8000, ED_ERR_CAPACITY, "more than 32767 unique matrix values");
        if (!coded) {
            // the product stages the value table (and the window) in LDS, next to its reduction buffer
            int dev = 0, max_lds = 0;
            ED_HIP(hipGetDevice(&dev));
            ED_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
            const size_t cap = size_t(max_lds) > 2 * BS * sizeof(double) ? size_t(max_lds) - 2 * BS * sizeof(double) : 0;
            ED_REQUIRE(hmat_lds_bytes(m->n_vals, -1, 16) <= cap, ED_ERR_CAPACITY,
                       "the stored matrix's value table does not fit in LDS");
            int radius = hmat_window_radius();
            while (radius > 0 && hmat_lds_bytes(m->n_vals, radius, 16) > cap) radius /= 2;
            m->radius = radius;
        }
        if (!coded) {
            double big = 0.0, imag = 0.0;
            for (int i = 0; i < m->n_vals; ++i) {
                big = std::max({big, std::fabs(vals[size_t(2 * i)]), std::fabs(vals[size_t(2 * i + 1)])});
                imag = std::max(imag, std::fabs(vals[size_t(2 * i + 1)]));
            }
            m->real = imag <= 1e-13 * big;
        }
        // 4-byte elements when column, value index and sign fit 32 bits (ED_HMAT_ELEMENT_BYTES=6 forces the
        // 6-byte form)
        {
            int col_bits = 1;
            while (col_bits < 31 && (int64_t(1) << col_bits) < s->D) ++col_bits;
            const int code_bits = 32 - col_bits;
            const char* force = std::getenv("ED_HMAT_ELEMENT_BYTES");
            const bool six = force && std::atoi(force) == 6;
            if (!six && code_bits >= 2 && int64_t(n_off) <= (int64_t(1) << (code_bits - 1))) m->shift = code_bits;
        }
        m->d_vals = dev_upload(vals, st);
        const int64_t D = s->D;
        m->nslices = (D + 63) / 64;
        m->d_cnt = dev_alloc<uint8_t>(size_t(std::max<int64_t>(D, 1)));
        m->d_diag = dev_alloc<uint16_t>(size_t(std::max<int64_t>(D, 1)));
        // partials of the product (one per workgroup, at most one per tile) + one dot for the tuner
        m->d_partial = dev_alloc<double>(size_t(2 * std::max<int64_t>((D + BS - 1) / BS, 1) + 2));
        std::vector<int64_t> sptr(static_cast<size_t>(m->nslices + 1), 0);
        if (D > 0) {
            uint16_t* d_map = dev_upload(map, st);
            tmp.push_back(d_map);
            uint16_t* d_dmap = dev_upload(dmap, st);
            tmp.push_back(d_dmap);
            int32_t* d_width = dev_alloc<int32_t>(size_t(m->nslices));
            tmp.push_back(d_width);
            const LdsLayout lay = lds_layout(s);
            const int lds = lay.bytes <= LDS_CAP;
            const unsigned nwg = unsigned((D + BS - 1) / BS);
            const SectorView v = s->view();
            hipLaunchKernelGGL(hm_count_kernel, dim3(nwg), dim3(BS), lds ? lay.bytes : 0, st, v, m->w, lay, lds,
                               d_dmap, m->d_cnt, m->d_diag);
            hipLaunchKernelGGL(hm_width_kernel, dim3(unsigned((m->nslices + 3) / 4)), dim3(BS), 0, st, m->d_cnt,
                               D, d_width, m->nslices);
            ED_HIP(hipGetLastError());
            std::vector<int32_t> width(static_cast<size_t>(m->nslices));
            ED_HIP(hipMemcpyAsync(width.data(), d_width, width.size() * 4, hipMemcpyDeviceToHost, st));
            ED_HIP(hipStreamSynchronize(st));
            for (int64_t i = 0; i < m->nslices; ++i) {
                const int32_t wv = width[size_t(i)];
                sptr[size_t(i + 1)] = sptr[size_t(i)] + int64_t(wv & 0xffff) * 64;
                m->nnz += int64_t(uint32_t(wv) >> 16);
            }
            m->slots = sptr.back();
            m->d_sptr = dev_upload(sptr, st);
            if (m->shift) {
                m->d_pack = dev_alloc<uint32_t>(size_t(std::max<int64_t>(m->slots, 1)));
            } else {
                m->d_col = dev_alloc<int32_t>(size_t(std::max<int64_t>(m->slots, 1)));
                m->d_code = dev_alloc<uint16_t>(size_t(std::max<int64_t>(m->slots, 1)));
            }
            hipLaunchKernelGGL(hm_fill_kernel, dim3(nwg), dim3(BS), lds ? lay.bytes : 0, st, v, m->w, lay, lds,
                               d_map, m->d_sptr, m->d_col, m->d_code, m->d_pack, m->shift);
            ED_HIP(hipGetLastError());
            const bool uniform = !coded && hmat_uniform();
            const int order = uniform ? 2 : hmat_row_order();
            if (order) {
                auto kernel = m->shift ? (order == 2 ? hm_sort_kernel<true, true> : hm_sort_kernel<true, false>)
                                       : (order == 2 ? hm_sort_kernel<false, true> : hm_sort_kernel<false, false>);
                hipLaunchKernelGGL(kernel, dim3(unsigned(m->nslices)), dim3(64), 0, st, D, m->d_sptr, m->d_cnt,
                                   m->d_col, m->d_code, m->d_pack, m->shift);
            }
            ED_HIP(hipGetLastError());
            if (uniform) extract_uniform(m, st);
            ED_HIP(hipGetLastError());
            ED_HIP(hipStreamSynchronize(st));
        } else {
            m->d_sptr = dev_upload(sptr, st);
            ED_HIP(hipStreamSynchronize(st));
        }
        cleanup_tmp();
    } catch (...) {
        cleanup_tmp();
        delete m;
        throw;
    }
    return m;
}

}  // namespace ed

extern "C" {

int ed_hmat_create(const ed_sector* s, const double* w, ed_hmat** out, void* stream) {
    return guarded([&] {
        ED_REQUIRE(out, ED_ERR_ARG, "bad stored-matrix arguments");
        *out = nullptr;
        *out = hmat_build(s, w, 0, static_cast<hipStream_t>(stream));
    });
}

int ed_hmat_destroy(ed_hmat* m) {
    delete m;
    return ED_OK;
}

int ed_hmat_info(const ed_hmat* m, int64_t* bytes, int64_t* nnz, int64_t* slots, int* n_values) {
    return guarded([&] {
        ED_REQUIRE(m, ED_ERR_ARG, "null matrix");
        if (bytes) *bytes = m->bytes();
        if (nnz) *nnz = m->nnz;
        if (slots) *slots = m->slots;
        if (n_values) *n_values = m->n_vals;
    });
}

int ed_hmat_element_bytes(const ed_hmat* m) { return m ? m->element_bytes() : ED_ERR_ARG; }

int ed_hmat_set_config(ed_hmat* m, int tiles, int unroll, int nontemporal) {
    return guarded([&] {
        ED_REQUIRE(m, ED_ERR_ARG, "null matrix");
        ED_REQUIRE(tiles >= 0 && tiles <= 1024, ED_ERR_ARG, "tiles must be in [0, 1024] (0 = default)");
        ED_REQUIRE(unroll == 4 || unroll == 8, ED_ERR_ARG, "unroll must be 4 or 8");
        ED_REQUIRE(nontemporal >= -1 && nontemporal <= 1, ED_ERR_ARG, "nontemporal must be -1, 0 or 1");
        m->tiles = tiles;
        m->unroll = unroll;
        m->nontemporal = nontemporal;
    });
}

int ed_hmat_tune(ed_hmat* m, const double* d_x, double* d_y, int precision, int* tiles, int* unroll,
                 int* nontemporal, double* best_ms, void* stream) {
    return guarded([&] {
        ED_REQUIRE(m && d_x && d_y, ED_ERR_ARG, "null argument");
        hmat_require_precision(m, precision);
        hipStream_t st = static_cast<hipStream_t>(stream);
        double best = 1e300;
        int bt = 0, bu = 4, bn = 0;
        if (m->D > 0) {
            hipEvent_t e0, e1;
            ED_HIP(hipEventCreate(&e0));
            ED_HIP(hipEventCreate(&e1));
            for (int t : {1, 2, 4, 8, 16, 32}) {
                for (int u : {4, 8})
                for (int nt : {0, 1}) {
                    m->tiles = t;
                    m->unroll = u;
                    m->nontemporal = nt;
                    // the product with its fused dot, as ed_hmat_hv runs it
                    auto run = [&] {
                        const int64_t np = launch_hmat(m, d_x, d_y, m->d_partial, precision, st);
                        launch_reduce_partials(m->d_partial, np, m->d_partial + 2 * hmat_max_partials(m), st);
                    };
                    run();
                    ED_HIP(hipEventRecord(e0, st));
                    for (int r = 0; r < 3; ++r) run();
                    ED_HIP(hipEventRecord(e1, st));
                    ED_HIP(hipEventSynchronize(e1));
                    float ms = 0.f;
                    ED_HIP(hipEventElapsedTime(&ms, e0, e1));
                    if (ms / 3 < best) {
                        best = ms / 3;
                        bt = t;
                        bu = u;
                        bn = nt;
                    }
                }
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        }
        m->tiles = bt;
        m->unroll = bu;
        m->nontemporal = bn;
        if (tiles) *tiles = bt;
        if (unroll) *unroll = bu;
        if (nontemporal) *nontemporal = bn;
        if (best_ms) *best_ms = best < 1e300 ? best : 0.0;
    });
}

int ed_hmat_is_real(const ed_hmat* m) { return m && m->real ? 1 : 0; }

int ed_hmat_hv(const ed_hmat* m, const double* d_x, double* d_y, double* d_dot, int precision, void* stream) {
    return guarded([&] {
        ED_REQUIRE(m && d_x && d_y, ED_ERR_ARG, "null argument");
        hmat_require_precision(m, precision);
        if (m->D == 0) return;
        hipStream_t st = static_cast<hipStream_t>(stream);
        const int64_t np = launch_hmat(m, d_x, d_y, d_dot ? m->d_partial : nullptr, precision, st);
        if (d_dot) launch_reduce_partials(m->d_partial, np, d_dot, st);
    });
}

int ed_hmat_line_probe(int scattered, int64_t span_bytes, double* lines_per_s, void* stream) {
    return guarded([&] {
        ED_REQUIRE(lines_per_s, ED_ERR_ARG, "null argument");
        ED_REQUIRE(span_bytes >= (int64_t(1) << 20) && span_bytes <= (int64_t(1) << 32) &&
                       (span_bytes & (span_bytes - 1)) == 0,
                   ED_ERR_ARG, "span_bytes must be a power of two in [1 MiB, 4 GiB]");
        hipStream_t st = static_cast<hipStream_t>(stream);
        const size_t n = size_t(span_bytes) / sizeof(double2);
        double2* x = dev_alloc<double2>(n);
        double* out = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        try {
            out = dev_alloc<double>(1);
            ED_HIP(hipMemsetAsync(x, 0, n * sizeof(double2), st));
            ED_HIP(hipEventCreate(&e0));
            ED_HIP(hipEventCreate(&e1));
            const unsigned nwg = 8192;
            const int iters = 256;
            const uint32_t mask = uint32_t(n - 1);
            auto run = [&] {
                if (scattered)
                    hipLaunchKernelGGL(line_probe_kernel<true>, dim3(nwg), dim3(BS), 0, st, x, mask, iters, out);
                else
                    hipLaunchKernelGGL(line_probe_kernel<false>, dim3(nwg), dim3(BS), 0, st, x, mask, iters, out);
                ED_HIP(hipGetLastError());
            };
            run();
            float best = 1e30f;
            for (int r = 0; r < 3; ++r) {
                ED_HIP(hipEventRecord(e0, st));
                run();
                ED_HIP(hipEventRecord(e1, st));
                ED_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                ED_HIP(hipEventElapsedTime(&ms, e0, e1));
                best = std::min(best, ms);
            }
            // wave-loads per launch x lines per wave-load (64 scattered, 8 for 1 KB in a row)
            const double lines = double(nwg) * (BS / 64) * iters * (scattered ? 64.0 : 8.0);
            *lines_per_s = lines / (double(best) * 1e-3);
        } catch (...) {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            dev_free(out);
            dev_free(x);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        dev_free(out);
        dev_free(x);
    });
}

}  // extern "C"
