// Matrix-free H·v with the sector tables staged in LDS (gfx950).
//
// Gather form y_f = sum_j conj(H_jf) x_j, one lane per row (ed_hvrow.h),
// organised for the CDNA4 memory system:
//  * every table the element computation touches is copied once per
//    workgroup into LDS, so the inner loop issues LDS reads instead of
//    scattered vector-L1 loads;
//  * each workgroup walks a short contiguous range of row tiles (two by
//    default), so the table copy is shared and the dot-product partials are
//    one per workgroup; workgroup w is placed so consecutive workgroups on
//    one XCD (round-robin dispatch over 8 XCDs) take neighbouring ranges;
//  * a wavefront whose 64 rows share one hole pattern (all but those where a
//    block ends) does the hole-only work once, on the scalar unit, finds
//    hop targets by chunked ranks and, in blocks without stabiliser,
//    exchange targets by closed-form rank changes (hv_row_uniform); the
//    other wavefronts, listed once per sector (hv_fix_list), take the
//    per-lane walk with two gathers in flight in the first workgroups of the
//    same launch (they start first and overlap the rest), so the kernel
//    needs fewer registers than with a per-lane walk of four gathers in
//    every workgroup (65 instead of 95 VGPRs for complex128: seven
//    wavefronts per SIMD instead of five, with the lean LDS layout);
//  * the gather group size G and the tiles per workgroup are tunable per
//    sector (ed_hv_tune times the candidates and keeps the fastest).

#include <algorithm>
#include <cstdlib>
#include <vector>

#include "ed_hvuni.h"

namespace ed {
namespace {

// UNIFORM: the first fix_wgs workgroups (a multiple of 8) walk the listed
// wavefronts fix[0 .. n_fix) per lane, BS / 64 of them each; the others walk
// row tiles, taking the wavefronts that pass wave_uniform and skipping the
// rest.  partial[workgroup] = sum conj(x_f) y_f over its rows.
template <class VT, int BS, int G, bool USE_LDS, bool UNIFORM>
__global__ void __launch_bounds__(BS) hv_lds_kernel(SectorView s, HvWeights w, const VT* __restrict__ x,
                                                    VT* __restrict__ y, double* __restrict__ partial, LdsLayout lay,
                                                    int64_t ntiles, int64_t tiles_per_wg, int chunk,
                                                    const int32_t* __restrict__ fix, int64_t n_fix, int fix_wgs) {
    extern __shared__ __align__(16) unsigned char lds[];
    // a compile-time choice, so that the table reads compile to LDS reads
    const Tabs T = USE_LDS ? stage_tabs<BS>(s, lay, lds) : global_tabs(s);
    __syncthreads();
    if (UNIFORM && int(blockIdx.x) < fix_wgs) {
        const int64_t k = int64_t(blockIdx.x) * (BS / 64) + (threadIdx.x >> 6);
        C2 dot{0.0, 0.0};
        if (k < n_fix) {
            const int64_t f = int64_t(fix[k]) * 64 + (threadIdx.x & 63);
            if (f < s.D) {
                C2 xf;
                const C2 acc = hv_row<VT, 2>(s, T, w, x, f, xf);
                VT o;
                o.x = decltype(o.x)(acc.re);
                o.y = decltype(o.y)(acc.im);
                y[f] = o;
                dot = cmul(C2{xf.re, -xf.im}, acc);
            }
        }
        if (partial) {
            workgroup_sum2<BS>(dot.re, dot.im);
            if (threadIdx.x == 0) {
                partial[2 * blockIdx.x] = dot.re;
                partial[2 * blockIdx.x + 1] = dot.im;
            }
        }
        return;
    }
    // the eight XCDs on neighbouring chunks of row ranges (fix_wgs is a multiple of 8, so workgroup
    // fix_wgs + i runs on the XCD that workgroup i would)
    const int wg = fix_wgs + xcd_order(int(blockIdx.x) - fix_wgs, int(gridDim.x) - fix_wgs, chunk);
    const int64_t tw = wg - fix_wgs;
    const int64_t t0 = tw * tiles_per_wg;
    const int64_t t1 = t0 + tiles_per_wg < ntiles ? t0 + tiles_per_wg : ntiles;
    C2 dot{0.0, 0.0};
    for (int64_t tile = t0; tile < t1; ++tile) {
        const int64_t f = tile * BS + threadIdx.x;
        if constexpr (UNIFORM) {
            const int64_t f0 = tile * BS + int(threadIdx.x & ~63u);
            if (f0 >= s.D) continue;
            const bool live = f < s.D;
            C2 acc, xf;
            if (uniform_row<VT, G>(s, T, lay, w, x, f, f0, live, acc, xf) && live) {
                VT o;
                o.x = decltype(o.x)(acc.re);
                o.y = decltype(o.y)(acc.im);
                y[f] = o;
                dot = cadd(dot, cmul(C2{xf.re, -xf.im}, acc));
            }
            continue;  // the other wavefronts are the first workgroups'
        }
        if (f >= s.D) continue;
        C2 xf;
        const C2 acc = hv_row<VT, G>(s, T, w, x, f, xf);
        VT o;
        o.x = decltype(o.x)(acc.re);
        o.y = decltype(o.y)(acc.im);
        y[f] = o;
        dot = cadd(dot, cmul(C2{xf.re, -xf.im}, acc));
    }
    if (partial) {
        workgroup_sum2<BS>(dot.re, dot.im);
        if (threadIdx.x == 0) {
            partial[2 * wg] = dot.re;
            partial[2 * wg + 1] = dot.im;
        }
    }
}

// flag[v] = 1 when 64-row wavefront v fails wave_uniform (the same test the
// main kernel applies), one wavefront per 64 threads.
__global__ void __launch_bounds__(256) hv_wave_flags_kernel(SectorView s, int64_t nwaves, uint8_t* flag) {
    const Tabs T = global_tabs(s);
    const int64_t v = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (v >= nwaves) return;  // whole wavefronts
    const int64_t f0 = v * 64, f = f0 + (threadIdx.x & 63);
    uint32_t H, Dn, gm;
    int blk;
    bool generic;
    const bool u = wave_uniform(s, T, f, f0, f < s.D, H, Dn, blk, gm, generic);
    if ((threadIdx.x & 63) == 0) flag[v] = u ? 0 : 1;
}

template <class VT>
using HvKernel = void (*)(SectorView, HvWeights, const VT*, VT*, double*, LdsLayout, int64_t, int64_t, int,
                          const int32_t*, int64_t, int);

// The layout of the sector's H·v and whether it takes the wavefront-uniform kernel.
LdsLayout hv_layout(const ed_sector* s) { return hv_uniform_enabled() ? lds_layout_ranks(s) : lds_layout(s); }

bool hv_uniform_kernel(const LdsLayout& lay) { return lay.bytes <= LDS_CAP && lay.rt >= 0; }

// Workgroups for n_fix listed wavefronts: four per workgroup, rounded up to a multiple of the 8 XCDs.
int64_t fix_workgroups(int64_t n_fix) { return n_fix > 0 ? ((n_fix + 3) / 4 + 7) / 8 * 8 : 0; }

template <class VT>
HvKernel<VT> pick_kernel(int G, const LdsLayout& lay) {
    constexpr int BS = 256;
    if (lay.bytes > LDS_CAP) return hv_lds_kernel<VT, BS, 4, false, false>;  // spin flip: tables in global memory
    if (lay.rt >= 0) {
        if (G == 2) return hv_lds_kernel<VT, BS, 2, true, true>;
        if (G == 8) return hv_lds_kernel<VT, BS, 8, true, true>;
        return hv_lds_kernel<VT, BS, 4, true, true>;
    }
    if (G == 2) return hv_lds_kernel<VT, BS, 2, true, false>;
    if (G == 8) return hv_lds_kernel<VT, BS, 8, true, false>;
    return hv_lds_kernel<VT, BS, 4, true, false>;
}

// tiles_per_wg = 0: one to four row tiles per workgroup, so that there are
// at least about 24 workgroups per CU.  Many short workgroups let the
// dispatcher balance rows of unequal cost (on the 2x12 ladder 0.60 ms per
// complex128 product with four tiles, against 1.35 ms with one round of
// resident workgroups walking 25 tiles each).  Returns the number of
// workgroups (partials).
template <class VT>
int64_t launch_g(int G, int64_t tiles_per_wg, const LdsLayout& lay, hipStream_t st, const SectorView& v,
                 const HvWeights& w, const void* x, void* y, double* partial, int64_t ntiles, const int32_t* d_fix,
                 int64_t n_fix) {
    constexpr int BS = 256;
    const HvKernel<VT> k = pick_kernel<VT>(G, lay);
    const size_t bytes = lay.bytes > LDS_CAP ? 0 : size_t(lay.bytes);
    static const int cus = [] {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return 256;
        return n;
    }();
    // the uniform kernel (lean LDS, seven wavefronts per SIMD): at most two tiles (2x12 ladder: 0.54 ms with two
    // tiles, 0.58 ms with four)
    const int64_t max_tpw = hv_uniform_kernel(lay) ? 2 : 4;
    const int64_t tpw =
        tiles_per_wg > 0 ? tiles_per_wg : std::clamp<int64_t>(ntiles / (24 * int64_t(cus)), 1, max_tpw);
    const int64_t nwg = (ntiles + tpw - 1) / tpw;
    // XCD chunks of about 131,072 rows (ED_HV_XCD_CHUNK_ROWS overrides, 0 = one eighth per XCD)
    static const int64_t chunk_rows = [] {
        const char* e = std::getenv("ED_HV_XCD_CHUNK_ROWS");
        return e ? int64_t(std::atoll(e)) : int64_t(131072);
    }();
    const int chunk = chunk_rows > 0 ? int(std::max<int64_t>(1, chunk_rows / (tpw * BS))) : 0;
    const int fix_wgs = hv_uniform_kernel(lay) ? int(fix_workgroups(n_fix)) : 0;
    hipLaunchKernelGGL(k, dim3(unsigned(fix_wgs + nwg)), dim3(BS), bytes, st, v, w, static_cast<const VT*>(x),
                       static_cast<VT*>(y), partial, lay, ntiles, tpw, chunk, d_fix, n_fix, fix_wgs);
    return fix_wgs + nwg;
}

}  // namespace

// Persistent LDS-table H·v; returns the number of partials written, or -1
// when the sector's tables do not fit in LDS (caller falls back).  Sectors
// with the spin flip, whose amplitude table is four times larger, read their
// tables from global memory instead (the fallback kernels have no spin flip).
int64_t launch_hv_lds(const ed_sector* s, const HvWeights& w, const void* x, void* y, double* partial, int precision,
                      hipStream_t st) {
    constexpr int BS = 256;
    const LdsLayout lay = hv_layout(s);
    if (lay.bytes > LDS_CAP && !s->flip) return -1;
    const int64_t ntiles = (s->D + BS - 1) / BS;
    if (ntiles == 0) return 0;
    const int G = s->tune_gather > 0 ? s->tune_gather : 4;
    const int64_t nwg = precision == 0 ? launch_g<double2>(G, s->tune_tiles, lay, st, s->view(), w, x, y, partial,
                                                           ntiles, s->d_fix, s->n_fix)
                                       : launch_g<float2>(G, s->tune_tiles, lay, st, s->view(), w, x, y, partial,
                                                          ntiles, s->d_fix, s->n_fix);
    ED_HIP(hipGetLastError());
    return nwg;
}

void hv_fix_list(ed_sector* s, hipStream_t st) {
    dev_free(s->d_fix);
    s->d_fix = nullptr;
    s->n_fix = 0;
    if (s->D <= 0 || !hv_uniform_kernel(hv_layout(s))) return;
    const int64_t nw = (s->D + 63) / 64;
    uint8_t* d_flag = dev_alloc<uint8_t>(size_t(nw));
    std::vector<uint8_t> flag(static_cast<size_t>(nw));
    hipLaunchKernelGGL(hv_wave_flags_kernel, dim3(unsigned((nw + 3) / 4)), dim3(256), 0, st, s->view(), nw, d_flag);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(flag.data(), d_flag, size_t(nw), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    dev_free(d_flag);
    ED_HIP(e);
    std::vector<int32_t> list;
    for (int64_t v = 0; v < nw; ++v)
        if (flag[size_t(v)]) list.push_back(int32_t(v));
    if (list.empty()) return;
    s->d_fix = dev_upload(list, st);
    s->n_fix = int64_t(list.size());
    ED_HIP(hipStreamSynchronize(st));  // the list outlives its upload
}

int64_t hv_fix_partials(const ed_sector* s) { return fix_workgroups(s->n_fix); }

}  // namespace ed
