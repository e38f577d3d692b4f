// One row of the stored-H product, shared by the single-sector kernel
// (ed_hmat.hip) and the batched kernels (ed_batch.hip).
//
// Row f of the sliced-ELL gather form: y_f = sqrt(n_f) * sum_k (+-V[code_k])
// x[col_k] + V[diag_f] x_f, element k at sptr[f / 64] + 64 k + f % 64, as
// packed 32-bit words (column << shift | value index << 1 | negate) or as an
// int32 column with a uint16 code (bit 15 = negate).  U elements are loaded,
// then their U gathers issued together.  V is the value table (staged in LDS
// by the caller).
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
};

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
template <class VT, int U, bool NT, bool PACKED>
__device__ inline C2 stored_row(const StoredView& m, const double* V, const VT* __restrict__ x, int64_t f, C2& xf) {
    const int64_t base = m.sptr[f >> 6] + (f & 63);
    const int n = ld_stream<NT>(m.cnt + f);
    if constexpr (is_real_vec<VT>::value) {
        double acc = 0.0;
        int k = 0;
        for (; k + U <= n; k += U) {
            int32_t j[U];
            int vi[U];
            double sg[U];
#pragma unroll
            for (int u = 0; u < U; ++u) ld_elem<PACKED, NT>(m, base + int64_t(k + u) * 64, j[u], vi[u], sg[u]);
            double xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) xv[u] = double(x[j[u]]);
#pragma unroll
            for (int u = 0; u < U; ++u) acc += sg[u] * V[2 * vi[u]] * xv[u];
        }
        for (; k < n; ++k) {
            int32_t j;
            int vi;
            double sg;
            ld_elem<PACKED, NT>(m, base + int64_t(k) * 64, j, vi, sg);
            acc += sg * V[2 * vi] * double(x[j]);
        }
        const double sqf = m.sq[m.orbit[f]];
        xf = C2{double(x[f]), 0.0};
        const int dv = ld_stream<NT>(m.diag + f);
        return C2{sqf * acc + V[2 * dv] * xf.re, 0.0};
    }
    C2 acc{0.0, 0.0};
    int k = 0;
    for (; k + U <= n; k += U) {
        int32_t j[U];
        int vi[U];
        double sg[U];
#pragma unroll
        for (int u = 0; u < U; ++u) ld_elem<PACKED, NT>(m, base + int64_t(k + u) * 64, j[u], vi[u], sg[u]);
        C2 xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = ld_vec(x, j[u]);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const C2 v{sg[u] * V[2 * vi[u]], sg[u] * V[2 * vi[u] + 1]};
            acc = cadd(acc, cmul(v, xv[u]));
        }
    }
    for (; k < n; ++k) {
        int32_t j;
        int vi;
        double sg;
        ld_elem<PACKED, NT>(m, base + int64_t(k) * 64, j, vi, sg);
        const C2 xv = ld_vec(x, j);
        acc = cadd(acc, cmul(C2{sg * V[2 * vi], sg * V[2 * vi + 1]}, xv));
    }
    const double sqf = m.sq[m.orbit[f]];
    xf = ld_vec(x, f);
    const int dv = ld_stream<NT>(m.diag + f);
    return cadd(C2{sqf * acc.re, sqf * acc.im}, cmul(C2{V[2 * dv], V[2 * dv + 1]}, xf));
}

}  // namespace ed
