// Several momentum sectors of one ladder in one product (gfx950).
//
// The sectors k of a ladder (num_rungs, num_holes, num_down_spins) share
// their representatives: a state is in sector k when its orbit is compatible
// with k (tJ_spin_half_ladder/symmetry_ladder.h:134-175), and every
// compatible state has the same orbit size n in all of them.  An element of
// H between representatives, <j~|H|f~> (general/column_functions.h:99-148
// for one element, tJ_spin_half_ladder/matrix_setup.pyx:80-118 for the
// terms), is
//     sqrt(n_f) * sign * w_t * s_t * conj(F_k[p]) / sqrt(n_j)
// with p the group element that maps the target to its representative and
// F_k[p] = exp(-2 pi i (kx m / L + ky b / 2)) for p = 2 m + b.  Only the phase
// depends on k.  So one stored structure over the union of the sectors (every
// lowest image, ed_sector_create_union) holds all of them: an element keeps
// (column, term, p, sign), and slot k's value is Q[t][p][k] = w_t s_t
// conj(F_k[p]).  In scaled coordinates x'_f = x_f / sqrt(n_f) the orbit
// factors drop out of the product,
//     y'_f = sum_e sign_e Q[t_e][p_e][k] x'_{j_e} + d_f x'_f,
// and inner products carry the weight n_f: <x, y> = sum_f n_f conj(x'_f) y'_f.
//
// Vectors are blocks of 16 slots, element (f, k) at f * 16 + k, so the 16
// lanes of one row gather one contiguous 16-slot line per element.  A slot is
// masked to its sector's states (bit k of mask[f]); rows outside it are zero.
//
// The Lanczos recurrence (general/cython_lanczos_diagonalization.pyx:18-77)
// runs for all slots at once with per-slot coefficients, stopping each slot
// at b < 1e-8 as the reference does.

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <complex>
#include <numbers>
#include <vector>

#include "ed_hmatrow.h"
#include "ed_internal.h"
#include "ed_matrix.h"
#include "ed_orbit.h"

struct ed_hmat;

namespace ed {

ed_hmat* hmat_build(const ed_sector* s, const double* w, int coded, hipStream_t st);

namespace {

// acc += (vr + i vi) x as fused multiply-adds (the stored products' accumulation)
__device__ inline void mk_fma(C2& acc, double vr, double vi, const C2& x) {
    acc.re = __builtin_fma(vr, x.re, __builtin_fma(-vi, x.im, acc.re));
    acc.im = __builtin_fma(vr, x.im, __builtin_fma(vi, x.re, acc.im));
}

constexpr int MK = 16;          // slots per block
constexpr int MBS = 256;        // threads per workgroup: 16 rows x 16 slots
constexpr int MK_CODES = 128;   // term * 32 + p
constexpr int MK_WGS = 2048;    // workgroups of the product and vector kernels (multiple of 8)
constexpr int RBS = 1024;

struct MkTabs {
    const double* q;   // [MK_CODES][MK] complex
    const double* wd;  // [MK][4]: slot's jl and jr weights (re, im), zero when the term is off
    int L1;            // L + 1: diagonal code = jl_count * (L + 1) + jr_count
};

template <class VT>
__device__ inline C2 ldv(const VT* p, int64_t i) {
    const VT v = p[i];
    return {double(v.x), double(v.y)};
}

template <class VT>
__device__ inline void stv(VT* p, int64_t i, C2 v) {
    VT o;
    o.x = decltype(o.x)(v.re);
    o.y = decltype(o.y)(v.im);
    p[i] = o;
}

// Rows of XCD x (workgroups blockIdx % 8 == x run on one XCD): row groups
// [x G / 8, (x + 1) G / 8), walked by its workgroups in turn, so that the
// groups one XCD holds at a time are neighbours and share its L2.
struct GroupRange {
    int64_t lo, hi, step, first;
};

__device__ inline GroupRange xcd_groups(int64_t G) {
    const int xcd = blockIdx.x & 7;
    const int64_t per = gridDim.x >> 3;
    return {G * xcd / 8, G * (xcd + 1) / 8, per, int64_t(blockIdx.x >> 3)};
}

// The product's walk: chunks of C groups dealt round-robin to the XCDs
// (chunk c to XCD c % 8), each chunk walked by its XCD's workgroups in turn,
// so that the eight XCDs work on neighbouring chunks at the same time.
struct ChunkWalk {
    int64_t G, C, per;
    int xcd;
    int64_t first;
};

__device__ inline ChunkWalk chunk_walk(int64_t G, int64_t C) {
    const int64_t per = gridDim.x >> 3;
    return {G, C > per ? C : per, per, int(blockIdx.x & 7), int64_t(blockIdx.x >> 3)};  // every workgroup busy
}

// Sum over the rows of the workgroup (NW wavefronts of 4 rows) for each
// slot; lane k of the first wavefront holds slot k's total.
template <int NW = 4>
__device__ inline void slot_sum(double& re, double& im, double (*red)[MK][2]) {
    re += __shfl_xor(re, 16, 64);
    im += __shfl_xor(im, 16, 64);
    re += __shfl_xor(re, 32, 64);
    im += __shfl_xor(im, 32, 64);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane < MK) {
        red[wave][lane][0] = re;
        red[wave][lane][1] = im;
    }
    __syncthreads();
    if (threadIdx.x < MK) {
        re = 0.0;
        im = 0.0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            re += red[w][threadIdx.x][0];
            im += red[w][threadIdx.x][1];
        }
    }
}

template <bool PACKED>
__device__ inline void mk_elem(const StoredView& m, int64_t at, int32_t& j, int& code, bool& neg) {
    if constexpr (PACKED) {
        const uint32_t e = __builtin_nontemporal_load(m.pack + at);
        j = int32_t(e >> m.shift);
        code = int((e & ((1u << m.shift) - 1u)) >> 1);
        neg = e & 1u;
    } else {
        j = __builtin_nontemporal_load(m.col + at);
        const uint16_t c = __builtin_nontemporal_load(m.code + at);
        code = c & 0x7fff;
        neg = c >> 15;
    }
}

// y' = H' x' for all slots; partial[wg][k] = sum_f n_f conj(x'_f) y'_f.
// NW wavefronts per workgroup, 4 rows each; U gathers in flight per lane.
template <class VT, bool PACKED, int U, int NW>
__global__ void __launch_bounds__(64 * NW) mk_spmv_kernel(StoredView m, int64_t D, MkTabs tb,
                                                          const uint16_t* __restrict__ mask, const VT* __restrict__ x,
                                                          VT* __restrict__ y, double* __restrict__ partial,
                                                          int64_t chunk_rows) {
    constexpr int NT = 64 * NW, RW = 4 * NW;
    __shared__ double2 Q[MK_CODES * MK];
    __shared__ double red[NW][MK][2];
    for (int i = threadIdx.x; i < MK_CODES * MK; i += NT) Q[i] = double2{tb.q[2 * i], tb.q[2 * i + 1]};
    __syncthreads();
    const int k = threadIdx.x & (MK - 1);
    const int r = threadIdx.x >> 4;
    const C2 wjl{tb.wd[4 * k], tb.wd[4 * k + 1]}, wjr{tb.wd[4 * k + 2], tb.wd[4 * k + 3]};
    double dre = 0.0, dim = 0.0;
    const int64_t G = (D + RW - 1) / RW;
    const ChunkWalk cw = chunk_walk(G, chunk_rows > 0 ? (chunk_rows + RW - 1) / RW : (G + 7) / 8);
    for (int64_t c = cw.xcd; c * cw.C < G; c += 8)
    for (int64_t grp = c * cw.C + cw.first, hi = (c + 1) * cw.C < G ? (c + 1) * cw.C : G; grp < hi; grp += cw.per) {
        const int64_t f = grp * RW + r;
        if (f >= D) continue;
        const int64_t base = m.sptr[f >> 6] + (f & 63);
        const int n = __builtin_nontemporal_load(m.cnt + f);
        C2 acc{0.0, 0.0};
        int e = 0;
        for (; e + U <= n; e += U) {
            int32_t j[U];
            int c[U];
            bool ng[U];
#pragma unroll
            for (int u = 0; u < U; ++u) mk_elem<PACKED>(m, base + int64_t(e + u) * 64, j[u], c[u], ng[u]);
            C2 xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) xv[u] = ldv(x, int64_t(j[u]) * MK + k);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const double2 q = Q[c[u] * MK + k];
                const double sg = ng[u] ? -1.0 : 1.0;
                mk_fma(acc, sg * q.x, sg * q.y, xv[u]);
            }
        }
        for (; e < n; ++e) {
            int32_t j;
            int c;
            bool ng;
            mk_elem<PACKED>(m, base + int64_t(e) * 64, j, c, ng);
            const C2 xv = ldv(x, int64_t(j) * MK + k);
            const double2 q = Q[c * MK + k];
            const double sg = ng ? -1.0 : 1.0;
            mk_fma(acc, sg * q.x, sg * q.y, xv);
        }
        const C2 xf = ldv(x, f * MK + k);
        // diagonal: -1/2 per antiparallel bond times the slot's jl / jr weight, as the single-sector table has it
        const int dc = __builtin_nontemporal_load(m.diag + f);
        const int na = dc / tb.L1, nb = dc - na * tb.L1;
        C2 d = cadd(C2{0.0, 0.0}, cmul(wjl, C2{-0.5 * na, 0.0}));
        d = cadd(d, cmul(wjr, C2{-0.5 * nb, 0.0}));
        C2 yf = cadd(acc, cmul(d, xf));
        if (!((mask[f] >> k) & 1u)) yf = C2{0.0, 0.0};
        stv(y, f * MK + k, yf);
        const double nf = double(m.orbit[f]);
        dre += nf * (xf.re * yf.re + xf.im * yf.im);
        dim += nf * (xf.re * yf.im - xf.im * yf.re);
    }
    slot_sum<NW>(dre, dim, red);
    if (threadIdx.x < MK) {
        partial[2 * (int64_t(blockIdx.x) * MK + threadIdx.x)] = dre;
        partial[2 * (int64_t(blockIdx.x) * MK + threadIdx.x) + 1] = dim;
    }
}

// out[k] = sum over workgroups of partial[wg][k]; one workgroup per slot.
__global__ void __launch_bounds__(RBS) mk_reduce_kernel(const double* partial, int64_t nwg, double* out) {
    __shared__ double red[2 * RBS];
    const int k = blockIdx.x;
    double re = 0.0, im = 0.0;
    for (int64_t w = threadIdx.x; w < nwg; w += RBS) {
        re += partial[2 * (w * MK + k)];
        im += partial[2 * (w * MK + k) + 1];
    }
    red[2 * threadIdx.x] = re;
    red[2 * threadIdx.x + 1] = im;
    __syncthreads();
    for (int s = RBS / 2; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) {
            red[2 * threadIdx.x] += red[2 * (threadIdx.x + s)];
            red[2 * threadIdx.x + 1] += red[2 * (threadIdx.x + s) + 1];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * k] = red[0];
        out[2 * k + 1] = red[1];
    }
}

// Place a sector vector (its own dense order) into slot `slot` of a block in
// scaled coordinates and mark the slot's states.  Sector states are
// representatives, so each is found in the union; a state that is not, or
// whose orbit size differs, sets err.
__global__ void __launch_bounds__(MBS) mk_attach_kernel(SectorView u, SectorView sk, int slot, const double* phi,
                                                        double* block, uint16_t* mask, int* err) {
    const int64_t i = int64_t(blockIdx.x) * MBS + threadIdx.x;
    if (i >= sk.D) return;
    const uint64_t key = sk.key[i];
    const Found fd = lookup_rep(u, uint32_t(key >> 32), uint32_t(key));
    if (fd.j < 0 || fd.n != int(sk.orbit[i])) {
        atomicOr(err, 1);
        return;
    }
    const double inv = 1.0 / std::sqrt(double(fd.n));
    block[2 * (fd.j * MK + slot)] = phi[2 * i] * inv;
    block[2 * (fd.j * MK + slot) + 1] = phi[2 * i + 1] * inv;
    mask[fd.j] = uint16_t(mask[fd.j] | (1u << slot));  // one sector state per union state: no race
}

// The sector vector of slot `slot` (its own dense order) from a block.
__global__ void __launch_bounds__(MBS) mk_extract_kernel(SectorView u, SectorView sk, int slot, const double* block,
                                                         double* out, int* err) {
    const int64_t i = int64_t(blockIdx.x) * MBS + threadIdx.x;
    if (i >= sk.D) return;
    const uint64_t key = sk.key[i];
    const Found fd = lookup_rep(u, uint32_t(key >> 32), uint32_t(key));
    if (fd.j < 0) {
        atomicOr(err, 1);
        out[2 * i] = out[2 * i + 1] = 0.0;
        return;
    }
    const double sq = std::sqrt(double(fd.n));
    out[2 * i] = block[2 * (fd.j * MK + slot)] * sq;
    out[2 * i + 1] = block[2 * (fd.j * MK + slot) + 1] * sq;
}

// -- Lanczos over all slots ----------------------------------------------------

struct MkLz {
    double* a;  // [MK][nsteps]
    double* b;  // [MK][nsteps + 1]
    double* s;  // [MK][nsteps + 1]: 1 / ||w_n||, the scale of the stored unnormalised vector
    int* flag;  // [MK][2]: stopped, coefficient count
    int nsteps;
};

template <class VT>
__global__ void __launch_bounds__(MBS) mk_copy_in_kernel(int64_t D, const uint8_t* __restrict__ orbit,
                                                         const double* __restrict__ src, VT* __restrict__ dst,
                                                         double* __restrict__ partial) {
    __shared__ double red[4][MK][2];
    const int k = threadIdx.x & (MK - 1), r = threadIdx.x >> 4;
    double n2 = 0.0, z = 0.0;
    const GroupRange g = xcd_groups((D + 15) / 16);
    for (int64_t grp = g.lo + g.first; grp < g.hi; grp += g.step) {
        const int64_t f = grp * 16 + r;
        if (f >= D) continue;
        const int64_t at = f * MK + k;
        stv(dst, at, C2{src[2 * at], src[2 * at + 1]});
        const C2 v = ldv(dst, at);
        n2 += double(orbit[f]) * (v.re * v.re + v.im * v.im);
    }
    slot_sum(n2, z, red);
    if (threadIdx.x < MK) {
        partial[2 * (int64_t(blockIdx.x) * MK + threadIdx.x)] = n2;
        partial[2 * (int64_t(blockIdx.x) * MK + threadIdx.x) + 1] = 0.0;
    }
}

__device__ inline double reduce_slot(const double* partial, int64_t nwg, int k, double* red) {
    double re = 0.0;
    for (int64_t w = threadIdx.x; w < nwg; w += RBS) re += partial[2 * (w * MK + k)];
    red[threadIdx.x] = re;
    __syncthreads();
    for (int s = RBS / 2; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    return red[0];
}

__global__ void __launch_bounds__(RBS) mk_lz_init_kernel(MkLz lz, const double* partial, int64_t nwg) {
    __shared__ double red[RBS];
    const int k = blockIdx.x;
    const double n2 = reduce_slot(partial, nwg, k, red);
    if (threadIdx.x == 0) {
        lz.b[k * (lz.nsteps + 1)] = 0.0;
        lz.s[k * (lz.nsteps + 1)] = n2 > 0.0 ? 1.0 / std::sqrt(n2) : 0.0;
        lz.flag[2 * k] = n2 > 0.0 ? 0 : 1;
        lz.flag[2 * k + 1] = 0;
    }
}

__global__ void __launch_bounds__(RBS) mk_lz_a_kernel(MkLz lz, const double* partial, int64_t nwg, int step) {
    __shared__ double red[RBS];
    const int k = blockIdx.x;
    if (lz.flag[2 * k]) return;
    const double re = reduce_slot(partial, nwg, k, red);
    if (threadIdx.x == 0) {
        const double sn = lz.s[k * (lz.nsteps + 1) + step];
        lz.a[k * lz.nsteps + step] = sn * sn * re;
    }
}

// w = s_n y - a_n s_n u - b_n s_{n-1} u_prev into u_prev, with its weighted
// squared norm per slot; a stopped slot is set to zero.  With acc (second
// pass of two-pass Lanczos), acc += ritz[k][n] v_n in FP64.
template <class VT>
__global__ void __launch_bounds__(MBS) mk_lz_update_kernel(MkLz lz, int step, int64_t D,
                                                           const uint8_t* __restrict__ orbit,
                                                           const VT* __restrict__ y, const VT* __restrict__ u,
                                                           VT* __restrict__ up, double* __restrict__ partial,
                                                           double2* __restrict__ acc, const double* __restrict__ ritz) {
    __shared__ double red[4][MK][2];
    const int k = threadIdx.x & (MK - 1), r = threadIdx.x >> 4;
    const int S = lz.nsteps + 1;
    const bool stopped = lz.flag[2 * k] != 0;
    const double sn = stopped ? 0.0 : lz.s[k * S + step];
    const double an = stopped ? 0.0 : lz.a[k * lz.nsteps + step];
    const double bn = stopped ? 0.0 : lz.b[k * S + step];
    const double snm = (stopped || step == 0) ? 0.0 : lz.s[k * S + step - 1];
    const double c = (acc && !stopped) ? ritz[k * lz.nsteps + step] * sn : 0.0;
    double n2 = 0.0, z = 0.0;
    const GroupRange g = xcd_groups((D + 15) / 16);
    for (int64_t grp = g.lo + g.first; grp < g.hi; grp += g.step) {
        const int64_t f = grp * 16 + r;
        if (f >= D) continue;
        const int64_t at = f * MK + k;
        const C2 yv = ldv(y, at), uv = ldv(u, at), pv = ldv(up, at);
        if (acc) {
            const double2 av = acc[at];
            acc[at] = double2{av.x + c * uv.re, av.y + c * uv.im};
        }
        const double vr = sn * uv.re, vi = sn * uv.im;
        const double wr = sn * yv.re - an * vr - bn * (snm * pv.re);
        const double wi = sn * yv.im - an * vi - bn * (snm * pv.im);
        stv(up, at, C2{wr, wi});
        const C2 wv = ldv(up, at);
        n2 += double(orbit[f]) * (wv.re * wv.re + wv.im * wv.im);
    }
    slot_sum(n2, z, red);
    if (threadIdx.x < MK) {
        partial[2 * (int64_t(blockIdx.x) * MK + threadIdx.x)] = n2;
        partial[2 * (int64_t(blockIdx.x) * MK + threadIdx.x) + 1] = 0.0;
    }
}

__global__ void __launch_bounds__(RBS) mk_lz_b_kernel(MkLz lz, const double* partial, int64_t nwg, int step) {
    __shared__ double red[RBS];
    const int k = blockIdx.x;
    if (lz.flag[2 * k]) return;
    const double n2 = reduce_slot(partial, nwg, k, red);
    if (threadIdx.x == 0) {
        const int S = lz.nsteps + 1;
        const double bnp = std::sqrt(n2);
        lz.b[k * S + step + 1] = bnp;
        lz.s[k * S + step + 1] = 1.0 / bnp;
        lz.flag[2 * k + 1] = step + 1;
        if (bnp < 1e-8) lz.flag[2 * k] = 1;
    }
}

// partial[wg][k] = sum_f n_f |y_k(f) - e_k x_k(f)|^2 (complex128 blocks; y null: |x|^2).
__global__ void __launch_bounds__(MBS) mk_resid_kernel(int64_t D, const uint8_t* __restrict__ orbit,
                                                       const double2* __restrict__ x, const double2* __restrict__ y,
                                                       const double* __restrict__ e, double* __restrict__ partial) {
    __shared__ double red[4][MK][2];
    const int k = threadIdx.x & (MK - 1), r = threadIdx.x >> 4;
    const double ek = e ? e[k] : 0.0;
    double n2 = 0.0, z = 0.0;
    const GroupRange g = xcd_groups((D + 15) / 16);
    for (int64_t grp = g.lo + g.first; grp < g.hi; grp += g.step) {
        const int64_t f = grp * 16 + r;
        if (f >= D) continue;
        const double2 xv = x[f * MK + k];
        double rr = xv.x, ri = xv.y;
        if (y) {
            const double2 yv = y[f * MK + k];
            rr = yv.x - ek * xv.x;
            ri = yv.y - ek * xv.y;
        }
        n2 += double(orbit[f]) * (rr * rr + ri * ri);
    }
    slot_sum(n2, z, red);
    if (threadIdx.x < MK) {
        partial[2 * (int64_t(blockIdx.x) * MK + threadIdx.x)] = n2;
        partial[2 * (int64_t(blockIdx.x) * MK + threadIdx.x) + 1] = 0.0;
    }
}

__global__ void __launch_bounds__(MBS) mk_scale_kernel(int64_t D, double2* __restrict__ x, const double* __restrict__ s) {
    const int k = threadIdx.x & (MK - 1), r = threadIdx.x >> 4;
    const double sk = s[k];
    const GroupRange g = xcd_groups((D + 15) / 16);
    for (int64_t grp = g.lo + g.first; grp < g.hi; grp += g.step) {
        const int64_t f = grp * 16 + r;
        if (f >= D) continue;
        const double2 v = x[f * MK + k];
        x[f * MK + k] = double2{sk * v.x, sk * v.y};
    }
}

}  // namespace
}  // namespace ed

using namespace ed;

struct ed_mk {
    const ed_sector* u = nullptr;
    ed_hmat* hm = nullptr;
    int n_slots = 0;
    int kx[MK] = {}, ky[MK] = {};
    int waves = 8;   // wavefronts per workgroup of the product (4 or 8)
    int unroll = 4;  // gathers in flight per lane (4 or 8)
    int64_t D = 0;
    double* d_q = nullptr;
    double* d_wd = nullptr;  // [MK][4] per-slot diagonal weights
    uint16_t* d_mask = nullptr;
    double* d_partial = nullptr;  // [MK_WGS][MK] complex
    double* d_aux = nullptr;      // [4 * MK]: per-slot host inputs (e or s) and reduced outputs
    int* d_err = nullptr;

    ~ed_mk() {
        for (void* p : {(void*)d_q, (void*)d_wd, (void*)d_mask, (void*)d_partial, (void*)d_aux, (void*)d_err})
            dev_free(p);
        if (hm) ed_hmat_destroy(hm);
    }
};

namespace ed {
namespace {

MkTabs mk_tabs(const ed_mk* m) { return MkTabs{m->d_q, m->d_wd, m->u->L + 1}; }

template <class VT, bool PACKED, int U, int NW>
void launch_mk_t(const ed_mk* m, const StoredView& v, const MkTabs& tb, const VT* x, VT* y, double* partial,
                 hipStream_t st) {
    // chunks of 16,384 rows: on the 14 momenta of the 2x12 sector 4.40 ms instead of 4.72 ms per complex64
    // product (7.30 -> 6.62 ms complex128) against one contiguous eighth per XCD; ED_MK_XCD_CHUNK overrides it
    // (rows, 0 = one eighth each)
    static const int64_t chunk_rows = [] {
        const char* e = std::getenv("ED_MK_XCD_CHUNK");
        return e ? int64_t(std::atoll(e)) : int64_t(16384);
    }();
    hipLaunchKernelGGL((mk_spmv_kernel<VT, PACKED, U, NW>), dim3(MK_WGS), dim3(64 * NW), 0, st, v, m->D, tb,
                       m->d_mask, x, y, partial, chunk_rows);
}

template <class VT, bool PACKED>
void launch_mk_p(const ed_mk* m, const StoredView& v, const MkTabs& tb, const VT* x, VT* y, double* partial,
                 hipStream_t st) {
    const int cfg = (m->waves == 8 ? 2 : 0) + (m->unroll == 8 ? 1 : 0);
    switch (cfg) {
        case 0: launch_mk_t<VT, PACKED, 4, 4>(m, v, tb, x, y, partial, st); break;
        case 1: launch_mk_t<VT, PACKED, 8, 4>(m, v, tb, x, y, partial, st); break;
        case 2: launch_mk_t<VT, PACKED, 4, 8>(m, v, tb, x, y, partial, st); break;
        default: launch_mk_t<VT, PACKED, 8, 8>(m, v, tb, x, y, partial, st); break;
    }
}

template <class VT>
void launch_mk(const ed_mk* m, const VT* x, VT* y, double* partial, hipStream_t st) {
    const StoredView v = hmat_view(m->hm);
    const MkTabs tb = mk_tabs(m);
    if (v.shift)
        launch_mk_p<VT, true>(m, v, tb, x, y, partial, st);
    else
        launch_mk_p<VT, false>(m, v, tb, x, y, partial, st);
    ED_HIP(hipGetLastError());
}

void check_slot_sector(const ed_mk* m, int slot, const ed_sector* sk) {
    ED_REQUIRE(slot >= 0 && slot < m->n_slots, ED_ERR_ARG, "slot out of range");
    ED_REQUIRE(sk && sk->D >= 0 && sk->d_key, ED_ERR_STATE, "sector basis not built");
    require_no_flip(sk, "the multi-momentum product");
    ED_REQUIRE(sk->L == m->u->L && sk->h == m->u->h && sk->nd == m->u->nd && !sk->any_k, ED_ERR_ARG,
               "sector is not one of this ladder's momentum sectors");
    const int L = m->u->L;
    ED_REQUIRE(((sk->kx - m->kx[slot]) % L + L) % L == 0 && sk->ky == m->ky[slot], ED_ERR_ARG,
               "sector momentum differs from the slot's");
}

template <class VT>
void mk_lanczos_t(const ed_mk* m, const double* d_v0, int nsteps, double* a, double* b, int* n_out, hipStream_t st,
                  const double* ritz = nullptr, double2* acc = nullptr) {
    const int64_t D = m->D;
    const size_t nv = size_t(D) * MK;
    VT* buf[3] = {dev_alloc<VT>(nv), dev_alloc<VT>(nv), dev_alloc<VT>(nv)};
    MkLz lz{dev_alloc<double>(size_t(MK) * nsteps), dev_alloc<double>(size_t(MK) * (nsteps + 1)),
            dev_alloc<double>(size_t(MK) * (nsteps + 1)), dev_alloc<int>(2 * MK), nsteps};
    double* d_ritz = nullptr;
    auto cleanup = [&] {
        for (auto* p : buf) dev_free(p);
        dev_free(lz.a);
        dev_free(lz.b);
        dev_free(lz.s);
        dev_free(lz.flag);
        dev_free(d_ritz);
    };
    try {
        VT* u = buf[0];
        VT* up = buf[1];
        VT* y = buf[2];
        const uint8_t* orbit = m->u->d_orbit;
        ED_HIP(hipMemsetAsync(up, 0, nv * sizeof(VT), st));
        ED_HIP(hipMemsetAsync(lz.a, 0, size_t(MK) * nsteps * sizeof(double), st));
        ED_HIP(hipMemsetAsync(lz.b, 0, size_t(MK) * (nsteps + 1) * sizeof(double), st));
        if (acc) {
            std::vector<double> r(size_t(MK) * nsteps, 0.0);
            std::copy(ritz, ritz + size_t(m->n_slots) * nsteps, r.begin());
            d_ritz = dev_upload(r, st);
            ED_HIP(hipMemsetAsync(acc, 0, nv * sizeof(double2), st));
        }
        hipLaunchKernelGGL((mk_copy_in_kernel<VT>), dim3(MK_WGS), dim3(MBS), 0, st, D, orbit, d_v0, u, m->d_partial);
        hipLaunchKernelGGL(mk_lz_init_kernel, dim3(MK), dim3(RBS), 0, st, lz, m->d_partial, int64_t(MK_WGS));
        ED_HIP(hipGetLastError());
        int flag[2 * MK];
        for (int n = 0; n < nsteps; ++n) {
            launch_mk<VT>(m, u, y, m->d_partial, st);
            hipLaunchKernelGGL(mk_lz_a_kernel, dim3(MK), dim3(RBS), 0, st, lz, m->d_partial, int64_t(MK_WGS), n);
            hipLaunchKernelGGL((mk_lz_update_kernel<VT>), dim3(MK_WGS), dim3(MBS), 0, st, lz, n, D, orbit, y, u, up,
                               m->d_partial, acc, d_ritz);
            hipLaunchKernelGGL(mk_lz_b_kernel, dim3(MK), dim3(RBS), 0, st, lz, m->d_partial, int64_t(MK_WGS), n);
            ED_HIP(hipGetLastError());
            std::swap(u, up);  // w (u_{n+1}) now lives in u
            if ((n & 63) == 63) {
                ED_HIP(hipMemcpyAsync(flag, lz.flag, sizeof(flag), hipMemcpyDeviceToHost, st));
                ED_HIP(hipStreamSynchronize(st));
                bool all = true;
                for (int k = 0; k < MK; ++k) all = all && flag[2 * k];
                if (all) break;
            }
        }
        ED_HIP(hipMemcpyAsync(flag, lz.flag, sizeof(flag), hipMemcpyDeviceToHost, st));
        for (int k = 0; k < m->n_slots && a && b; ++k) {
            ED_HIP(hipMemcpyAsync(a + size_t(k) * nsteps, lz.a + size_t(k) * nsteps, size_t(nsteps) * sizeof(double),
                                  hipMemcpyDeviceToHost, st));
            ED_HIP(hipMemcpyAsync(b + size_t(k) * (nsteps + 1), lz.b + size_t(k) * (nsteps + 1),
                                  size_t(nsteps + 1) * sizeof(double), hipMemcpyDeviceToHost, st));
        }
        ED_HIP(hipStreamSynchronize(st));
        for (int k = 0; k < m->n_slots && n_out; ++k) n_out[k] = flag[2 * k + 1];
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
}

}  // namespace
}  // namespace ed

extern "C" {

int ed_mk_create(const ed_sector* u, const double* w, int n_slots, const int* kx, const int* ky, ed_mk** out,
                 void* stream) {
    return guarded([&] {
        ED_REQUIRE(out, ED_ERR_ARG, "out is null");
        *out = nullptr;
        ED_REQUIRE(u && u->any_k, ED_ERR_ARG, "the union sector (ed_sector_create_union) is required");
        ED_REQUIRE(u->D >= 0 && u->d_key, ED_ERR_STATE, "union basis not built");
        ED_REQUIRE(n_slots >= 1 && n_slots <= MK && kx && ky && w, ED_ERR_ARG, "1..16 momentum slots required");
        hipStream_t st = static_cast<hipStream_t>(stream);
        auto* m = new ed_mk();
        try {
            m->u = u;
            m->D = u->D;
            m->n_slots = n_slots;
            const int L = u->L, N = u->N;
            for (int k = 0; k < n_slots; ++k) {
                ED_REQUIRE(ky[k] == 0 || ky[k] == 1, ED_ERR_ARG, "rung quantum number must be 0 or 1");
                m->kx[k] = kx[k];
                m->ky[k] = ky[k];
            }
            // the structure holds every term that some slot uses
            std::vector<HvWeights> hw(static_cast<size_t>(n_slots));
            double wb[8] = {0.0};
            for (int k = 0; k < n_slots; ++k) {
                hw[size_t(k)] = make_weights(w + 8 * k);
                for (int t = 0; t < 4; ++t)
                    if (hw[size_t(k)].active & (1 << t)) wb[2 * t] = 1.0;
            }
            m->hm = hmat_build(u, wb, 1, st);
            // Q[t * 32 + p][k] = w_t s_t conj(F_k[p]), F_k[p] built as for the sector's amplitude table
            using cd = std::complex<double>;
            const cd i2pi(0.0, 2.0 * std::numbers::pi);
            std::vector<double> q(size_t(MK_CODES) * MK * 2, 0.0);
            std::vector<double> wd(size_t(MK) * 4, 0.0);
            for (int k = 0; k < n_slots; ++k) {
                const HvWeights& h = hw[size_t(k)];
                const cd mirror_exp = std::exp(-i2pi * cd(double(ky[k]) / double(2)));
                const cd rung_exp = std::exp(-i2pi * cd(double(kx[k]) / double(L)));
                std::vector<cd> E(size_t(L) + 1);
                E[0] = 1;
                for (int mm = 1; mm <= L; ++mm) E[size_t(mm)] = E[size_t(mm - 1)] * rung_exp;
                for (int t = 0; t < 4; ++t) {
                    if (!(h.active & (1 << t))) continue;
                    const double st_ = t < 2 ? 1.0 : 0.5;
                    for (int p = 0; p < N; ++p) {
                        const int mm = p >> 1, bb = p & 1;
                        const cd F = bb ? E[size_t(mm)] * mirror_exp : E[size_t(mm)];
                        const C2 val = cmul(h.w[t], C2{st_ * F.real(), -(st_ * F.imag())});
                        const size_t at = size_t((t * 32 + p) * MK + k);
                        q[2 * at] = val.re;
                        q[2 * at + 1] = val.im;
                    }
                }
                if (h.active & 4) {
                    wd[size_t(4 * k)] = h.w[2].re;
                    wd[size_t(4 * k + 1)] = h.w[2].im;
                }
                if (h.active & 8) {
                    wd[size_t(4 * k + 2)] = h.w[3].re;
                    wd[size_t(4 * k + 3)] = h.w[3].im;
                }
            }
            m->d_wd = dev_upload(wd, st);
            m->d_q = dev_upload(q, st);
            m->d_mask = dev_alloc<uint16_t>(size_t(std::max<int64_t>(m->D, 1)));
            ED_HIP(hipMemsetAsync(m->d_mask, 0, size_t(std::max<int64_t>(m->D, 1)) * 2, st));
            m->d_partial = dev_alloc<double>(size_t(2) * MK_WGS * MK);
            m->d_aux = dev_alloc<double>(size_t(4) * MK);
            m->d_err = dev_alloc<int>(1);
            ED_HIP(hipStreamSynchronize(st));
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
    });
}

int ed_mk_set_config(ed_mk* m, int waves, int unroll) {
    return guarded([&] {
        ED_REQUIRE(m, ED_ERR_ARG, "null matrix");
        ED_REQUIRE((waves == 4 || waves == 8) && (unroll == 4 || unroll == 8), ED_ERR_ARG,
                   "waves and unroll must be 4 or 8");
        m->waves = waves;
        m->unroll = unroll;
    });
}

int ed_mk_destroy(ed_mk* m) {
    delete m;
    return ED_OK;
}

int ed_mk_info(const ed_mk* m, int64_t* dim, int64_t* nnz, int64_t* bytes) {
    return guarded([&] {
        ED_REQUIRE(m, ED_ERR_ARG, "null matrix");
        int64_t b = 0, z = 0, slots = 0;
        int nv = 0;
        ED_REQUIRE(ed_hmat_info(m->hm, &b, &z, &slots, &nv) == ED_OK, ED_ERR_STATE, "stored structure");
        if (dim) *dim = m->D;
        if (nnz) *nnz = z;
        if (bytes) *bytes = b + m->D * 2;
    });
}

int ed_mk_attach(ed_mk* m, int slot, const ed_sector* sk, const double* d_phi, double* d_block, void* stream) {
    return guarded([&] {
        ED_REQUIRE(m, ED_ERR_ARG, "null matrix");
        check_slot_sector(m, slot, sk);
        ED_REQUIRE(sk->D == 0 || (d_phi && d_block), ED_ERR_ARG, "null vector");
        if (sk->D == 0) return;
        hipStream_t st = static_cast<hipStream_t>(stream);
        ED_HIP(hipMemsetAsync(m->d_err, 0, sizeof(int), st));
        hipLaunchKernelGGL(mk_attach_kernel, dim3(unsigned((sk->D + MBS - 1) / MBS)), dim3(MBS), 0, st, m->u->view(),
                           sk->view(), slot, d_phi, d_block, m->d_mask, m->d_err);
        ED_HIP(hipGetLastError());
        int err = 0;
        ED_HIP(hipMemcpyAsync(&err, m->d_err, sizeof(int), hipMemcpyDeviceToHost, st));
        ED_HIP(hipStreamSynchronize(st));
        ED_REQUIRE(err == 0, ED_ERR_STATE, "a sector state is missing from the union basis");
    });
}

int ed_mk_extract(const ed_mk* m, int slot, const ed_sector* sk, const double* d_block, double* d_out,
                  void* stream) {
    return guarded([&] {
        ED_REQUIRE(m, ED_ERR_ARG, "null matrix");
        check_slot_sector(m, slot, sk);
        ED_REQUIRE(sk->D == 0 || (d_out && d_block), ED_ERR_ARG, "null vector");
        if (sk->D == 0) return;
        hipStream_t st = static_cast<hipStream_t>(stream);
        ED_HIP(hipMemsetAsync(m->d_err, 0, sizeof(int), st));
        hipLaunchKernelGGL(mk_extract_kernel, dim3(unsigned((sk->D + MBS - 1) / MBS)), dim3(MBS), 0, st, m->u->view(),
                           sk->view(), slot, d_block, d_out, m->d_err);
        ED_HIP(hipGetLastError());
        int err = 0;
        ED_HIP(hipMemcpyAsync(&err, m->d_err, sizeof(int), hipMemcpyDeviceToHost, st));
        ED_HIP(hipStreamSynchronize(st));
        ED_REQUIRE(err == 0, ED_ERR_STATE, "a sector state is missing from the union basis");
    });
}

int ed_mk_hv(const ed_mk* m, const double* d_x, double* d_y, double* d_dot, int precision, void* stream) {
    return guarded([&] {
        ED_REQUIRE(m, ED_ERR_ARG, "null matrix");
        ED_REQUIRE(precision == 0 || precision == 1, ED_ERR_ARG, "precision must be 0 (c128) or 1 (c64)");
        if (m->D == 0) return;
        ED_REQUIRE(d_x && d_y, ED_ERR_ARG, "null vector");
        hipStream_t st = static_cast<hipStream_t>(stream);
        if (precision == 0)
            launch_mk<double2>(m, reinterpret_cast<const double2*>(d_x), reinterpret_cast<double2*>(d_y), m->d_partial,
                               st);
        else
            launch_mk<float2>(m, reinterpret_cast<const float2*>(d_x), reinterpret_cast<float2*>(d_y), m->d_partial,
                              st);
        if (d_dot) {
            hipLaunchKernelGGL(mk_reduce_kernel, dim3(MK), dim3(RBS), 0, st, m->d_partial, int64_t(MK_WGS), d_dot);
            ED_HIP(hipGetLastError());
        }
    });
}

int ed_mk_lanczos_run(const ed_mk* m, const double* d_v0, int nsteps, double* a, double* b, int* n_out,
                      int precision, void* stream) {
    return guarded([&] {
        ED_REQUIRE(m && a && b && n_out, ED_ERR_ARG, "null argument");
        ED_REQUIRE(nsteps >= 1, ED_ERR_ARG, "nsteps must be positive");
        ED_REQUIRE(precision == 0 || precision == 1, ED_ERR_ARG, "precision must be 0 (c128) or 1 (c64)");
        for (int k = 0; k < m->n_slots; ++k) n_out[k] = 0;
        if (m->D == 0) return;
        ED_REQUIRE(d_v0, ED_ERR_ARG, "null start block");
        hipStream_t st = static_cast<hipStream_t>(stream);
        if (precision == 0)
            mk_lanczos_t<double2>(m, d_v0, nsteps, a, b, n_out, st);
        else
            mk_lanczos_t<float2>(m, d_v0, nsteps, a, b, n_out, st);
    });
}

int ed_mk_lanczos_ritz(const ed_mk* m, const double* d_v0, int nsteps, const double* ritz, double* d_x, int precision,
                       void* stream) {
    return guarded([&] {
        ED_REQUIRE(m && ritz, ED_ERR_ARG, "null argument");
        ED_REQUIRE(nsteps >= 1, ED_ERR_ARG, "nsteps must be positive");
        ED_REQUIRE(precision == 0 || precision == 1, ED_ERR_ARG, "precision must be 0 (c128) or 1 (c64)");
        if (m->D == 0) return;
        ED_REQUIRE(d_v0 && d_x, ED_ERR_ARG, "null block");
        hipStream_t st = static_cast<hipStream_t>(stream);
        double2* acc = reinterpret_cast<double2*>(d_x);
        if (precision == 0)
            mk_lanczos_t<double2>(m, d_v0, nsteps, nullptr, nullptr, nullptr, st, ritz, acc);
        else
            mk_lanczos_t<float2>(m, d_v0, nsteps, nullptr, nullptr, nullptr, st, ritz, acc);
    });
}

int ed_mk_residual_norms(const ed_mk* m, const double* d_x, const double* d_hx, const double* e, double* out,
                         void* stream) {
    return guarded([&] {
        ED_REQUIRE(m && out, ED_ERR_ARG, "null argument");
        for (int k = 0; k < m->n_slots; ++k) out[k] = 0.0;
        if (m->D == 0) return;
        ED_REQUIRE(d_x, ED_ERR_ARG, "null block");
        hipStream_t st = static_cast<hipStream_t>(stream);
        double h[2 * MK] = {0.0};
        const bool with_e = d_hx && e;
        if (with_e) {
            std::copy(e, e + m->n_slots, h);
            ED_HIP(hipMemcpyAsync(m->d_aux, h, MK * sizeof(double), hipMemcpyHostToDevice, st));
        }
        hipLaunchKernelGGL(mk_resid_kernel, dim3(MK_WGS), dim3(MBS), 0, st, m->D, m->u->d_orbit,
                           reinterpret_cast<const double2*>(d_x), reinterpret_cast<const double2*>(d_hx),
                           with_e ? m->d_aux : nullptr, m->d_partial);
        hipLaunchKernelGGL(mk_reduce_kernel, dim3(MK), dim3(RBS), 0, st, m->d_partial, int64_t(MK_WGS),
                           m->d_aux + 2 * MK);
        ED_HIP(hipGetLastError());
        ED_HIP(hipMemcpyAsync(h, m->d_aux + 2 * MK, sizeof(h), hipMemcpyDeviceToHost, st));
        ED_HIP(hipStreamSynchronize(st));
        for (int k = 0; k < m->n_slots; ++k) out[k] = h[2 * k];
    });
}

int ed_mk_scale(const ed_mk* m, double* d_x, const double* s, void* stream) {
    return guarded([&] {
        ED_REQUIRE(m && s, ED_ERR_ARG, "null argument");
        if (m->D == 0) return;
        ED_REQUIRE(d_x, ED_ERR_ARG, "null block");
        hipStream_t st = static_cast<hipStream_t>(stream);
        double h[MK] = {0.0};
        std::copy(s, s + m->n_slots, h);
        ED_HIP(hipMemcpyAsync(m->d_aux, h, sizeof(h), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(mk_scale_kernel, dim3(MK_WGS), dim3(MBS), 0, st, m->D, reinterpret_cast<double2*>(d_x),
                           m->d_aux);
        ED_HIP(hipGetLastError());
        ED_HIP(hipStreamSynchronize(st));  // h is read by the copy before it goes out of scope
    });
}

}  // extern "C"
