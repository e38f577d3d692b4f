// Matrix-free H·v with the sector tables staged in LDS (gfx950).
//
// Gather form y_f = sum_j conj(H_jf) x_j, one lane per row (ed_hvrow.h),
// organised for the CDNA4 memory system:
//  * every table the element computation touches is copied once per
//    workgroup into LDS, so the inner loop issues LDS reads instead of
//    scattered vector-L1 loads;
//  * workgroups are persistent: each walks a contiguous range of row tiles,
//    so the table copy is amortised and the dot-product partials shrink to
//    one per workgroup; workgroup w is placed so consecutive workgroups on
//    one XCD (round-robin dispatch over 8 XCDs) take neighbouring ranges;
//  * the gather group size G and the tiles per workgroup are tunable per
//    sector (ed_hv_tune times the candidates and keeps the fastest).

#include <algorithm>
#include <cstdlib>

#include "ed_hvrow.h"

namespace ed {
namespace {

template <class VT, int BS, int G, bool USE_LDS>
__global__ void __launch_bounds__(BS) hv_lds_kernel(SectorView s, HvWeights w, const VT* __restrict__ x,
                                                    VT* __restrict__ y, double* __restrict__ partial, LdsLayout lay,
                                                    int64_t ntiles, int64_t tiles_per_wg, int chunk) {
    extern __shared__ __align__(16) unsigned char lds[];
    // a compile-time choice, so that the table reads compile to LDS reads
    const Tabs T = USE_LDS ? stage_tabs<BS>(s, lay, lds) : global_tabs(s);
    __syncthreads();
    // the eight XCDs on neighbouring chunks of row ranges
    const int wg = xcd_order(blockIdx.x, gridDim.x, chunk);
    const int64_t t0 = int64_t(wg) * tiles_per_wg;
    const int64_t t1 = t0 + tiles_per_wg < ntiles ? t0 + tiles_per_wg : ntiles;
    C2 dot{0.0, 0.0};
    for (int64_t tile = t0; tile < t1; ++tile) {
        const int64_t f = tile * BS + threadIdx.x;
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

template <class VT>
void launch_g(int G, int64_t nwg, const LdsLayout& lay, hipStream_t st, const SectorView& v, const HvWeights& w,
              const void* x, void* y, double* partial, int64_t ntiles, int64_t tpw) {
    constexpr int BS = 256;
    // XCD chunks of about 131,072 rows (ED_HV_XCD_CHUNK_ROWS overrides, 0 = one eighth per XCD)
    static const int64_t chunk_rows = [] {
        const char* e = std::getenv("ED_HV_XCD_CHUNK_ROWS");
        return e ? int64_t(std::atoll(e)) : int64_t(131072);
    }();
    const int chunk = chunk_rows > 0 ? int(std::max<int64_t>(1, chunk_rows / (tpw * BS))) : 0;
    const VT* xx = static_cast<const VT*>(x);
    VT* yy = static_cast<VT*>(y);
    const dim3 grid{unsigned(nwg), 1u, 1u};
    auto go = [&](auto kernel, size_t bytes) {
        hipLaunchKernelGGL(kernel, grid, dim3(BS), bytes, st, v, w, xx, yy, partial, lay, ntiles, tpw, chunk);
    };
    if (lay.bytes > LDS_CAP)  // only sectors with the spin flip get here: tables from global memory
        go(hv_lds_kernel<VT, BS, 4, false>, 0);
    else if (G == 2)
        go(hv_lds_kernel<VT, BS, 2, true>, size_t(lay.bytes));
    else if (G == 8)
        go(hv_lds_kernel<VT, BS, 8, true>, size_t(lay.bytes));
    else
        go(hv_lds_kernel<VT, BS, 4, true>, size_t(lay.bytes));
}

}  // namespace

int hv_default_tiles(const ed_sector* s) {
    constexpr int BS = 256;
    const int64_t ntiles = (s->D + BS - 1) / BS;
    int dev = 0;
    ED_HIP(hipGetDevice(&dev));
    static int cus = 0;
    if (cus == 0) ED_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    // enough workgroups to fill every CU several times over, each with >= 4 tiles
    const int64_t target = int64_t(cus) * 8;
    return int(std::max<int64_t>(4, (ntiles + target - 1) / target));
}

// Persistent LDS-table H·v; returns the number of partials written, or -1
// when the sector's tables do not fit in LDS (caller falls back).  Sectors
// with the spin flip, whose amplitude table is four times larger, read their
// tables from global memory instead (the fallback kernels have no spin flip).
int64_t launch_hv_lds(const ed_sector* s, const HvWeights& w, const void* x, void* y, double* partial, int precision,
                      hipStream_t st) {
    constexpr int BS = 256;
    const LdsLayout lay = lds_layout(s);
    if (lay.bytes > LDS_CAP && !s->flip) return -1;
    const int64_t ntiles = (s->D + BS - 1) / BS;
    if (ntiles == 0) return 0;
    const int64_t tpw = s->tune_tiles > 0 ? s->tune_tiles : hv_default_tiles(s);
    const int G = s->tune_gather > 0 ? s->tune_gather : 4;
    const int64_t nwg = (ntiles + tpw - 1) / tpw;
    if (precision == 0)
        launch_g<double2>(G, nwg, lay, st, s->view(), w, x, y, partial, ntiles, tpw);
    else
        launch_g<float2>(G, nwg, lay, st, s->view(), w, x, y, partial, ntiles, tpw);
    ED_HIP(hipGetLastError());
    return nwg;
}

}  // namespace ed
