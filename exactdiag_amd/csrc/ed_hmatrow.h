// One row of the stored-H product, shared by the single-sector kernel
// (ed_hmat.hip) and the batched kernels (ed_batch.hip).
//
// Row f of the sliced-ELL gather form: y_f = sqrt(n_f) * sum_k (+-V[code_k])
// x[col_k] + V[diag_f] x_f, element k at sptr[f / 64] + 64 k + f % 64 (local
// segment first, stored_row), as
// packed 32-bit words (column << shift | value index << 1 | negate) or as an
// int32 column with a uint16 code (bit 15 = negate).  Elements that every row
// of a 64-row slice has with the same column offset d = column - f and the
// same code are kept once per slice, in uni[uptr[s] .. uptr[s + 1]) as
// (d, code): the 64 lanes of the wavefront that owns the slice read each such
// entry at one address and gather x[f + d] from 64 consecutive states.  U
// elements are loaded, then their U gathers issued together.  V is the value
// table (staged in LDS by the caller).  The caller's wavefronts must each
// cover the rows of one slice (row ranges start at multiples of 64).
#pragma once

#include "ed_internal.h"

namespace ed {

struct StoredView {
    const int64_t* sptr;
    const int32_t* col;
    const uint16_t* code;
    const uint32_t* pack;
    int shift;  // 0: int32 column + uint16 code
    int n_vals;
    const uint8_t* cnt;
    const uint16_t* diag;
    const uint8_t* orbit;
    const double* sq;
    const double* vals;
    const int64_t* uptr;  // per slice: first slice-uniform entry (nullptr: none stored)
    const uint2* uni;     // slice-uniform entries (column offset, code in the element's format)
    const uint8_t* lcnt;  // per row: local elements (|column - row| <= radius); nullptr: none
    const uint8_t* lwid;  // per slice: slots of the local segment (the far segment follows)
    int radius;           // reach of the local elements and of the staged window
};

template <bool PACKED>
__device__ inline void decode_code(uint32_t c, int& vi, double& sg) {
    if constexpr (PACKED) {
        vi = int(c >> 1);
        sg = (c & 1u) ? -1.0 : 1.0;
    } else {
        vi = int(c & 0x7fffu);
        sg = (c & 0x8000u) ? -1.0 : 1.0;
    }
}

// Range [u0, u1) of the slice-uniform entries of the wavefront's slice (one value for all its lanes).
__device__ inline void uniform_range(const StoredView& m, int64_t f, int64_t& u0, int64_t& u1) {
    u0 = u1 = 0;
    if (!m.uptr) return;
    const int s = __builtin_amdgcn_readfirstlane(int(f >> 6));
    u0 = m.uptr[s];
    u1 = m.uptr[s + 1];
}

// Read-once streams can be loaded with a non-temporal hint (NT) so that L2
// keeps the gathered x_j.
template <bool NT, class T>
__device__ inline T ld_stream(const T* p) {
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

template <bool PACKED, bool NT>
__device__ inline void ld_elem(const StoredView& m, int64_t at, int32_t& j, int& vi, double& sg) {
    if constexpr (PACKED) {
        const uint32_t e = ld_stream<NT>(m.pack + at);
        j = int32_t(e >> m.shift);
        vi = int((e & ((1u << m.shift) - 1u)) >> 1);
        sg = (e & 1u) ? -1.0 : 1.0;
    } else {
        j = ld_stream<NT>(m.col + at);
        const uint16_t c = ld_stream<NT>(m.code + at);
        vi = c & 0x7fff;
        sg = (c & 0x8000) ? -1.0 : 1.0;
    }
}

template <class VT>
__device__ inline C2 ld_vec(const VT* __restrict__ x, int64_t j) {
    const VT v = x[j];
    if constexpr (is_real_vec<VT>::value)
        return {double(v), 0.0};
    else
        return {double(v.x), double(v.y)};
}

// y_f of the stored product; xf receives x_f for the fused <x, Hx>.
// Real vectors (VT double or float) use only the real parts of the values.
// Row f's per-lane elements sit in two segments of its slice: local ones
// (|column - f| <= m.radius) in slots [0, lwid[s]), far ones from slot
// lwid[s] on.  WIN: x[win0 + i] is staged in LDS as win[i] for every column
// the local elements, the local slice-uniform entries and x_f can reach; the
// other elements gather from x.  Without WIN every element gathers from x.
template <class VT, int U, bool NT, bool PACKED, bool WIN = false>
__device__ inline C2 stored_row(const StoredView& m, const double* V, const VT* __restrict__ x, int64_t f, C2& xf,
                                const VT* win = nullptr, int64_t win0 = 0) {
    constexpr bool REAL = is_real_vec<VT>::value;
    const int64_t s = f >> 6;
    const int64_t base = m.sptr[s] + (f & 63);
    const int nf = ld_stream<NT>(m.cnt + f);
    const int nl = m.lcnt ? int(ld_stream<NT>(m.lcnt + f)) : 0;
    const int64_t far0 = base + (m.lwid ? int64_t(m.lwid[s]) * 64 : 0);
    int64_t u0, u1;
    uniform_range(m, f, u0, u1);
    C2 acc{0.0, 0.0};
    // the values are exact table entries; each element's product enters the sum as two fused multiply-adds
    // per component (fewer FP64 instructions than a rounded product and a sum, and no less accurate)
    auto add = [&](int vi, double sg, const C2& xv) {
        const double vr = sg * V[2 * vi];
        if constexpr (REAL) {
            acc.re = __builtin_fma(vr, xv.re, acc.re);
        } else {
            const double vim = sg * V[2 * vi + 1];
            acc.re = __builtin_fma(vr, xv.re, __builtin_fma(-vim, xv.im, acc.re));
            acc.im = __builtin_fma(vr, xv.im, __builtin_fma(vim, xv.re, acc.im));
        }
    };
    // slice-uniform entries (offset and code the same for the whole wavefront, so is the branch)
    int64_t q = u0;
    for (; q + U <= u1; q += U) {
        int vi[U];
        double sg[U];
        int32_t d[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint2 e = m.uni[q + u];
            d[u] = int32_t(e.x);
            decode_code<PACKED>(e.y, vi[u], sg[u]);
        }
        C2 xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int32_t j = int32_t(f) + d[u];
            if (WIN && d[u] >= -m.radius && d[u] <= m.radius)
                xv[u] = ld_vec(win, int64_t(j) - win0);
            else
                xv[u] = ld_vec(x, j);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) add(vi[u], sg[u], xv[u]);
    }
    for (; q < u1; ++q) {
        const uint2 e = m.uni[q];
        int vi;
        double sg;
        decode_code<PACKED>(e.y, vi, sg);
        const int32_t d = int32_t(e.x), j = int32_t(f) + d;
        add(vi, sg, (WIN && d >= -m.radius && d <= m.radius) ? ld_vec(win, int64_t(j) - win0) : ld_vec(x, j));
    }
    // per-lane elements: the local segment (from the window when staged), then the far one
    auto segment = [&](int64_t first, int n, bool local) {
        int k = 0;
        for (; k + U <= n; k += U) {
            int32_t j[U];
            int vi[U];
            double sg[U];
#pragma unroll
            for (int u = 0; u < U; ++u) ld_elem<PACKED, NT>(m, first + int64_t(k + u) * 64, j[u], vi[u], sg[u]);
            C2 xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) xv[u] = (WIN && local) ? ld_vec(win, int64_t(j[u]) - win0) : ld_vec(x, j[u]);
#pragma unroll
            for (int u = 0; u < U; ++u) add(vi[u], sg[u], xv[u]);
        }
        for (; k < n; ++k) {
            int32_t j;
            int vi;
            double sg;
            ld_elem<PACKED, NT>(m, first + int64_t(k) * 64, j, vi, sg);
            add(vi, sg, (WIN && local) ? ld_vec(win, int64_t(j) - win0) : ld_vec(x, j));
        }
    };
    segment(base, nl, true);
    segment(far0, nf, false);
    const double sqf = m.sq[m.orbit[f]];
    xf = WIN ? ld_vec(win, f - win0) : ld_vec(x, f);
    const int dv = ld_stream<NT>(m.diag + f);
    if constexpr (REAL) return C2{sqf * acc.re + V[2 * dv] * xf.re, 0.0};
    return cadd(C2{sqf * acc.re, sqf * acc.im}, cmul(C2{V[2 * dv], V[2 * dv + 1]}, xf));
}

}  // namespace ed
