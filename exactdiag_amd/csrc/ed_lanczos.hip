// Lanczos recurrence, continued fraction and dense vector helpers for gfx950.
//
// The recurrence keeps its coefficients on the device: per step one fused
// H·v (with the workgroup partials of <u, Hu>), a one-workgroup reduction
// that first writes b_n and s_n = 1 / b_n from the previous update's
// partials and then a_n, and one fused update w = s_n Hu - a_n s_n u_n - b_n
// s_{n-1} u_{n-1} (with partials of ||w||^2); b of the last step is reduced
// after the loop.  Vectors are never rescaled in place: u_n is the
// unnormalised Lanczos vector and s_n its scale, which saves a full pass per
// step.  No host synchronisation inside the loop except a stop check every
// 64 steps; with a stop rule (ed_hmat_lanczos_run_tol) the host reads the
// coefficients of each chunk while the GPU runs the next one.  All
// reductions are fixed-order, so runs are reproducible.

#include <algorithm>
#include <cmath>
#include <complex>
#include <mutex>
#include <type_traits>
#include <vector>

#include "ed_internal.h"
#include "ed_matrix.h"

namespace ed {
namespace {

constexpr int BS = 256;

// One device block: flag (two ints in the first double), a, b, then s, so
// that flag, a and b reach the host in one copy.
struct LzDev {
    double* a;  // [nsteps]
    double* b;  // [nsteps + 1]
    double* s;  // [nsteps + 1]
    int* flag;  // [0] stopped, [1] count
};

inline size_t lz_doubles(int nsteps) { return size_t(1) + size_t(nsteps) + size_t(nsteps + 1) + size_t(nsteps + 1); }

inline LzDev lz_view(double* block, int nsteps) {
    return LzDev{block + 1, block + 1 + nsteps, block + 2 + 2 * size_t(nsteps), reinterpret_cast<int*>(block)};
}

// Vector elements: complex (double2, float2) or real (double, float; the
// imaginary part reads as zero and is not stored).
template <class VT>
__device__ inline void ldc(const VT* p, int64_t i, double& re, double& im) {
    const VT v = p[i];
    if constexpr (is_real_vec<VT>::value) {
        re = double(v);
        im = 0.0;
    } else {
        re = double(v.x);
        im = double(v.y);
    }
}

// x rounded to the component type of VT, as stc stores it
template <class VT>
__device__ inline double rounded(double x) {
    if constexpr (std::is_same_v<VT, float2> || std::is_same_v<VT, float>)
        return double(float(x));
    else
        return x;
}

template <class VT>
__device__ inline void stc(VT* p, int64_t i, double re, double im) {
    if constexpr (is_real_vec<VT>::value) {
        p[i] = VT(re);
    } else {
        VT o;
        o.x = decltype(o.x)(re);
        o.y = decltype(o.y)(im);
        p[i] = o;
    }
}

constexpr int RBS = 1024;     // threads of the one-workgroup reductions
constexpr int UPD_WGS = 2048;  // workgroups of the grid-stride vector kernels (fewer partials to reduce)

template <int NT = BS>
__device__ inline void block_sum2(double& re, double& im) {
    workgroup_sum2<NT>(re, im);
}

// partial[2 wg] = squared norm of the stored start vector; for real vectors
// partial[2 wg + 1] = squared norm of the imaginary part that was dropped.
template <class VT>
__global__ void __launch_bounds__(BS) copy_in_kernel(const double* src, VT* dst, int64_t D, double* partial) {
    double n2 = 0.0, lost = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * BS + threadIdx.x; i < D; i += int64_t(gridDim.x) * BS) {
        const double re = src[2 * i], im = src[2 * i + 1];
        stc(dst, i, re, im);
        double r2, i2;
        ldc(dst, i, r2, i2);
        n2 += r2 * r2 + i2 * i2;
        if constexpr (is_real_vec<VT>::value) lost += im * im;
    }
    block_sum2(n2, lost);
    if (threadIdx.x == 0) partial[2 * blockIdx.x] = n2, partial[2 * blockIdx.x + 1] = lost;
}

// flag[1] = -1 marks a start vector whose dropped imaginary part exceeds
// 1e-10 of its norm (real vectors only); the run then stops at once.
__global__ void __launch_bounds__(RBS) lz_init_kernel(LzDev lz, const double* partial, int64_t n) {
    double re = 0.0, im = 0.0;
    for (int64_t k = threadIdx.x; k < n; k += RBS) {
        re += partial[2 * k];
        im += partial[2 * k + 1];
    }
    block_sum2<RBS>(re, im);
    if (threadIdx.x == 0) {
        const bool lost = im > 1e-20 * (re + im);
        lz.b[0] = 0.0;
        lz.s[0] = 1.0 / std::sqrt(re);
        lz.flag[0] = (re > 0.0 && !lost) ? 0 : 1;
        lz.flag[1] = lost ? -1 : 0;
    }
}

// b_step and s_step from the update partials of step - 1 (step > 0), as
// lz_reduce_b_kernel; then, unless the space is exhausted, a_step from the
// product's partials.
__global__ void __launch_bounds__(RBS) lz_reduce_ba_kernel(LzDev lz, const double* pupd, int64_t nupd,
                                                           const double* partial, int64_t n, int step) {
    if (lz.flag[0]) return;
    __shared__ double sn_shared;
    __shared__ int stopped;
    if (step > 0) {
        double re = 0.0, im = 0.0;
        for (int64_t k = threadIdx.x; k < nupd; k += RBS) re += pupd[2 * k];
        block_sum2<RBS>(re, im);
        if (threadIdx.x == 0) {
            const double bn = std::sqrt(re);
            lz.b[step] = bn;
            lz.s[step] = 1.0 / bn;
            lz.flag[1] = step;
            stopped = bn < 1e-8;
            if (stopped) lz.flag[0] = 1;
            sn_shared = 1.0 / bn;
        }
    } else if (threadIdx.x == 0) {
        stopped = 0;
        sn_shared = lz.s[0];
    }
    __syncthreads();
    if (stopped) return;
    double re, im;
    sum_partials<RBS>(partial, n, re, im);
    block_sum2<RBS>(re, im);
    if (threadIdx.x == 0) lz.a[step] = sn_shared * sn_shared * re;
}

template <class VT>
__global__ void __launch_bounds__(BS) lz_update_kernel(LzDev lz, int step, int64_t D, const VT* __restrict__ y,
                                                       const VT* __restrict__ u, VT* __restrict__ up,
                                                       VT* __restrict__ basis, double2* __restrict__ acc,
                                                       const double* __restrict__ ritz, double* partial) {
    if (lz.flag[0]) return;
    const double sn = lz.s[step], an = lz.a[step], bn = lz.b[step];
    const double snm = step > 0 ? lz.s[step - 1] : 0.0;
    const double c = acc ? ritz[step] * sn : 0.0;
    double n2 = 0.0;
    int64_t first = 0;  // elements before `first` were done 16 bytes at a time
    if (!basis && !acc) {
        // the plain recurrence: E elements per 16-byte load and store (vectors are 256-byte aligned)
        constexpr int E = int(16 / sizeof(VT));
        const int64_t nv = D / E;
        const uint4* y4 = reinterpret_cast<const uint4*>(y);
        const uint4* u4 = reinterpret_cast<const uint4*>(u);
        uint4* p4 = reinterpret_cast<uint4*>(up);
        for (int64_t v = int64_t(blockIdx.x) * BS + threadIdx.x; v < nv; v += int64_t(gridDim.x) * BS) {
            const uint4 yv = y4[v], uv = u4[v], pv = p4[v];
            const VT* ye = reinterpret_cast<const VT*>(&yv);
            const VT* ue = reinterpret_cast<const VT*>(&uv);
            const VT* pe = reinterpret_cast<const VT*>(&pv);
            uint4 wv;
            VT* we = reinterpret_cast<VT*>(&wv);
#pragma unroll
            for (int e = 0; e < E; ++e) {
                double yr, yi, ur, ui, pr, pi;
                ldc(ye, e, yr, yi);
                ldc(ue, e, ur, ui);
                ldc(pe, e, pr, pi);
                const double wr = sn * yr - an * (sn * ur) - bn * (snm * pr);
                const double wi = sn * yi - an * (sn * ui) - bn * (snm * pi);
                stc(we, e, wr, wi);
                const double r2 = rounded<VT>(wr), i2 = is_real_vec<VT>::value ? 0.0 : rounded<VT>(wi);
                n2 += r2 * r2 + i2 * i2;
            }
            p4[v] = wv;
        }
        first = nv * E;
    }
    for (int64_t i = first + int64_t(blockIdx.x) * BS + threadIdx.x; i < D; i += int64_t(gridDim.x) * BS) {
        double yr, yi, ur, ui, pr, pi;
        ldc(y, i, yr, yi);
        ldc(u, i, ur, ui);
        ldc(up, i, pr, pi);
        if (acc) {  // second pass: x += y_n v_n in FP64
            const double2 av = acc[i];
            acc[i] = double2{av.x + c * ur, av.y + c * ui};
        }
        const double vr = sn * ur, vi = sn * ui;
        const double hr = sn * yr, hi = sn * yi;
        const double qr = snm * pr, qi = snm * pi;
        const double wr = hr - an * vr - bn * qr;
        const double wi = hi - an * vi - bn * qi;
        stc(up, i, wr, wi);
        if (basis) stc(basis, i, vr, vi);
        // the norm of w as stored (rounded to the vector type), from registers
        const double r2 = rounded<VT>(wr), i2 = is_real_vec<VT>::value ? 0.0 : rounded<VT>(wi);
        n2 += r2 * r2 + i2 * i2;
    }
    double z = 0.0;
    block_sum2(n2, z);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = n2;
        partial[2 * blockIdx.x + 1] = 0.0;
    }
}

__global__ void __launch_bounds__(RBS) lz_reduce_b_kernel(LzDev lz, const double* partial, int64_t n, int step) {
    if (lz.flag[0]) return;
    double re = 0.0, im = 0.0;
    for (int64_t k = threadIdx.x; k < n; k += RBS) re += partial[2 * k];
    block_sum2<RBS>(re, im);
    if (threadIdx.x == 0) {
        const double bnp = std::sqrt(re);
        lz.b[step + 1] = bnp;
        lz.s[step + 1] = 1.0 / bnp;
        lz.flag[1] = step + 1;
        if (bnp < 1e-8) lz.flag[0] = 1;
    }
}

// general/cython_lanczos_diagonalization.pyx:42-48 per omega.  With b_inf > 0
// the last level is closed by the tail of a constant continuation (a_inf,
// b_inf): t(z) = (z - a_inf - sqrt((z - a_inf)^2 - 4 b_inf^2)) / (2 b_inf^2),
// the branch with Im t <= 0, so det_{n-1} = z - a_{n-1} - b_inf^2 t(z).
__global__ void cf_kernel(const double* a, const double* b, int n, const double* ws, int nw, double e0, double eta,
                          double factor, double a_inf, double b_inf, double* out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nw) return;
    const double zr = ws[k] + e0, zi = eta;
    double dr = zr - a[n - 1], di = zi;
    if (b_inf > 0.0) {
        const double ur = zr - a_inf, ui = zi;
        // w = u^2 - 4 b^2, s = sqrt(w) (principal), t = (u - s) / (2 b^2); pick the sign with Im t <= 0
        const double wr = ur * ur - ui * ui - 4.0 * b_inf * b_inf, wi = 2.0 * ur * ui;
        const double mod = std::sqrt(wr * wr + wi * wi);
        double sr = std::sqrt(0.5 * (mod + wr));
        double si = std::copysign(std::sqrt(0.5 * (mod - wr)), wi);
        double tr = (ur - sr) / (2.0 * b_inf * b_inf), ti = (ui - si) / (2.0 * b_inf * b_inf);
        if (ti > 0.0) {
            tr = (ur + sr) / (2.0 * b_inf * b_inf);
            ti = (ui + si) / (2.0 * b_inf * b_inf);
        }
        dr -= b_inf * b_inf * tr;
        di -= b_inf * b_inf * ti;
    }
    for (int i = 1; i < n; ++i) {
        const double bb = b[n - i] * b[n - i];
        const double den = dr * dr + di * di;
        const double qr = bb * dr / den, qi = -bb * di / den;  // bb / det
        dr = zr - a[n - 1 - i] - qr;
        di = zi - qi;
    }
    out[k] = factor * (-di / (dr * dr + di * di));
}

__global__ void __launch_bounds__(BS) dots_kernel(const double* V, int64_t ld, const double* w, int64_t D,
                                                  double* partial, int nwg) {
    const int k = blockIdx.y;
    double re = 0.0, im = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * BS + threadIdx.x; i < D; i += int64_t(nwg) * BS) {
        const double vr = V[2 * (k * ld + i)], vi = V[2 * (k * ld + i) + 1];
        const double xr = w[2 * i], xi = w[2 * i + 1];
        re += vr * xr + vi * xi;
        im += vr * xi - vi * xr;
    }
    block_sum2(re, im);
    if (threadIdx.x == 0) {
        partial[2 * (int64_t(k) * nwg + blockIdx.x)] = re;
        partial[2 * (int64_t(k) * nwg + blockIdx.x) + 1] = im;
    }
}

__global__ void dots_reduce_kernel(const double* partial, int nwg, double* out) {
    const int k = blockIdx.x;
    double re = 0.0, im = 0.0;
    for (int j = threadIdx.x; j < nwg; j += BS) {
        re += partial[2 * (int64_t(k) * nwg + j)];
        im += partial[2 * (int64_t(k) * nwg + j) + 1];
    }
    block_sum2(re, im);
    if (threadIdx.x == 0) {
        out[2 * k] = re;
        out[2 * k + 1] = im;
    }
}

__global__ void combine_kernel(const double* V, int nv, int64_t ld, const double* c, double* out, int64_t D,
                               int accumulate_negative) {
    const int64_t i = int64_t(blockIdx.x) * BS + threadIdx.x;
    if (i >= D) return;
    double re = 0.0, im = 0.0;
    for (int k = 0; k < nv; ++k) {
        const double vr = V[2 * (k * ld + i)], vi = V[2 * (k * ld + i) + 1];
        const double cr = c[2 * k], ci = c[2 * k + 1];
        re += cr * vr - ci * vi;
        im += cr * vi + ci * vr;
    }
    if (accumulate_negative) {
        out[2 * i] -= re;
        out[2 * i + 1] -= im;
    } else {
        out[2 * i] = re;
        out[2 * i + 1] = im;
    }
}

__global__ void scale_kernel(double* w, int64_t D, double cr, double ci) {
    const int64_t i = int64_t(blockIdx.x) * BS + threadIdx.x;
    if (i >= D) return;
    const double re = w[2 * i], im = w[2 * i + 1];
    w[2 * i] = cr * re - ci * im;
    w[2 * i + 1] = cr * im + ci * re;
}

__global__ void zero2_kernel(double2* p, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * BS + threadIdx.x;
    if (i < n) p[i] = double2{0.0, 0.0};
}

// Eigenvalues below x of the symmetric tridiagonal with diagonal d[0..n) and
// off-diagonal e[1..n) (e[i] couples i - 1 and i): the negative pivots of
// its LDL^T factorisation (Sturm count); a pivot that vanishes is replaced
// by -pivmin.
int sturm_below(const double* d, const double* e, int n, double x, double pivmin) {
    int c = 0;
    double q = 1.0;
    for (int i = 0; i < n; ++i) {
        q = (d[i] - x) - (i ? e[i] * e[i] / q : 0.0);
        if (std::fabs(q) < pivmin) q = -pivmin;
        if (q < 0.0) ++c;
    }
    return c;
}

// Smallest and largest eigenvalue of that tridiagonal, by bisection inside
// its Gershgorin interval to a relative width of 1e-14.
void extreme_eigenvalues(const double* d, const double* e, int n, double& lo, double& hi) {
    double gl = d[0], gu = d[0], emax = 1.0;
    for (int i = 0; i < n; ++i) {
        const double r = (i > 0 ? std::fabs(e[i]) : 0.0) + (i + 1 < n ? std::fabs(e[i + 1]) : 0.0);
        gl = std::min(gl, d[i] - r);
        gu = std::max(gu, d[i] + r);
        if (i > 0) emax = std::max(emax, e[i] * e[i]);
    }
    const double pivmin = 1e-300 * emax;
    auto kth = [&](int k) {  // the (k + 1)-th smallest eigenvalue
        double l = gl, u = gu;
        for (int it = 0; it < 200; ++it) {
            const double m = 0.5 * (l + u);
            if (u - l <= 1e-14 * std::max(std::fabs(l), std::fabs(u)) || m <= l || m >= u) break;
            if (sturm_below(d, e, n, m, pivmin) > k)
                u = m;
            else
                l = m;
        }
        return 0.5 * (l + u);
    };
    lo = kth(0);
    hi = kth(n - 1);
}

// Im(1 / det(z)) of the first n levels, closed by the square-root terminator
// at the bandwidth of their tridiagonal (as cf_kernel with a_inf = centre,
// b_inf = half-width / 2), at z = w + i eta for the grid points of `w`.
void terminated_spectrum(const double* a, const double* b, int n, const std::vector<double>& w, double eta,
                         std::vector<double>& out) {
    double lo, hi;
    extreme_eigenvalues(a, b, n, lo, hi);
    const double a_inf = 0.5 * (hi + lo), b_inf = 0.5 * std::max(0.5 * (hi - lo), 1e-12);
    const double bb = b_inf * b_inf;
    const size_t m = w.size();
    std::vector<double> dr(m), di(m);
    for (size_t k = 0; k < m; ++k) {
        const std::complex<double> z(w[k], eta), u = z - a_inf;
        const std::complex<double> s = std::sqrt(u * u - 4.0 * bb);
        std::complex<double> t = (u - s) / (2.0 * bb);
        if (t.imag() > 0.0) t = (u + s) / (2.0 * bb);
        const std::complex<double> det0 = z - a[n - 1] - bb * t;
        dr[k] = det0.real();
        di[k] = det0.imag();
    }
    // the levels in real arithmetic, all grid points per level: det <- z - a_i - b_{i+1}^2 / det
    for (int i = 1; i < n; ++i) {
        const double ai = a[n - 1 - i], b2 = b[n - i] * b[n - i];
        for (size_t k = 0; k < m; ++k) {
            const double q = b2 / (dr[k] * dr[k] + di[k] * di[k]);
            dr[k] = w[k] - ai - q * dr[k];
            di[k] = eta + q * di[k];
        }
    }
    out.resize(m);
    for (size_t k = 0; k < m; ++k) out[k] = -di[k] / (dr[k] * dr[k] + di[k] * di[k]);
}

// The larger relative L1 distance between the terminated spectrum of n
// levels and those of n - delta and n - delta / 2 levels, on the rule's
// strided grid (weights: central differences of the grid, one-sided at its
// ends).  The shorter comparison catches a spectrum that passes close to
// its n - delta form while still moving.
double spectrum_change(const double* a, const double* b, int n, int delta, const ed_lanczos_stop& r) {
    std::vector<double> w;
    for (int k = 0; k < r.n_w; k += r.stride) w.push_back(r.ws[k] + r.e0);
    const size_t m = w.size();
    std::vector<double> dw(m, 1.0);
    if (m > 1) {
        dw[0] = w[1] - w[0];
        dw[m - 1] = w[m - 1] - w[m - 2];
        for (size_t k = 1; k + 1 < m; ++k) dw[k] = 0.5 * (w[k + 1] - w[k - 1]);
    }
    std::vector<double> s1, s2;
    terminated_spectrum(a, b, n, w, r.eta, s1);
    double worst = 0.0;
    for (const int back : {delta, delta / 2}) {
        if (back < 1) continue;
        terminated_spectrum(a, b, n - back, w, r.eta, s2);
        double num = 0.0, den = 0.0;
        for (size_t k = 0; k < m; ++k) {
            num += std::fabs(s1[k] - s2[k]) * dw[k];
            den += std::fabs(s2[k]) * dw[k];
        }
        worst = std::max(worst, den == 0.0 ? (num == 0.0 ? 0.0 : HUGE_VAL) : num / den);
    }
    return worst;
}

void check_stop_rule(const ed_lanczos_stop* r) {
    ED_REQUIRE(r && r->ws && r->n_w >= 2 && r->stride >= 1 && r->eta > 0.0 && r->tol > 0.0 && r->check_every >= 1 &&
                   r->delta >= 1 && r->min_steps >= 1,
               ED_ERR_ARG, "bad stop rule: need a grid of >= 2 points, stride >= 1, eta > 0, tol > 0, "
                           "check_every >= 1, delta >= 1, min_steps >= 1");
}

// acc (optional, c128) receives sum_n ritz[n] v_n: the second pass of
// two-pass Lanczos (the recurrence is deterministic, so it sees the vectors
// of the first pass).
template <class VT>
void lanczos_run_t(const ed_sector* s, const double* w, const double* d_v0, int nsteps, double* a, double* b,
                   int* n_out, VT* d_basis, int precision, hipStream_t st, double2* acc = nullptr,
                   const double* ritz = nullptr, const ed_hmat* m = nullptr, const ed_lanczos_stop* stop = nullptr,
                   int* reason = nullptr, int64_t basis_ld = 0) {
    if constexpr (is_real_vec<VT>::value) ED_REQUIRE(m, ED_ERR_ARG, "real vectors need a stored matrix");
    const int64_t D = s->D;
    const int64_t nwg = (D + BS - 1) / BS;
    const int64_t nupd = std::min<int64_t>(nwg, UPD_WGS);  // grid-stride update: at most UPD_WGS partials
    const HvWeights hw = make_weights(w);
    VT* buf[3] = {dev_alloc<VT>(size_t(D)), dev_alloc<VT>(size_t(D)), dev_alloc<VT>(size_t(D))};
    double* partial = dev_alloc<double>(size_t(2 * std::max<int64_t>(nwg, hv_partials(s))));
    double* pupd = dev_alloc<double>(size_t(2 * nupd));  // the update's partials, kept until the next step
    double* lz_block = dev_alloc<double>(lz_doubles(nsteps));
    const LzDev lz = lz_view(lz_block, nsteps);
    double* d_ritz = nullptr;
    // the stop rule's two pinned host copies of the coefficients and the flag, one per chunk in flight
    struct Slot {
        double* block = nullptr;  // host copy of flag, a and b (the layout of lz_view)
        const double* a = nullptr;
        const double* b = nullptr;
        const int* flag = nullptr;
        hipEvent_t ev = nullptr;
    } slot[2];
    auto cleanup = [&] {
        if (stop) (void)hipStreamSynchronize(st);
        for (auto* p : buf) dev_free(p);
        dev_free(partial);
        dev_free(pupd);
        dev_free(lz_block);
        dev_free(d_ritz);
        for (Slot& sl : slot) {
            if (sl.block) (void)hipHostFree(sl.block);
            if (sl.ev) (void)hipEventDestroy(sl.ev);
        }
    };
    try {
        VT* u = buf[0];
        VT* up = buf[1];
        VT* y = buf[2];
        if (acc) {
            d_ritz = dev_alloc<double>(size_t(nsteps));
            ED_HIP(hipMemcpyAsync(d_ritz, ritz, size_t(nsteps) * sizeof(double), hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(zero2_kernel, dim3(unsigned(nwg)), dim3(BS), 0, st, acc, D);
        }
        ED_HIP(hipMemsetAsync(up, 0, size_t(D) * sizeof(VT), st));
        hipLaunchKernelGGL((copy_in_kernel<VT>), dim3(unsigned(nupd)), dim3(BS), 0, st, d_v0, u, D, partial);
        hipLaunchKernelGGL(lz_init_kernel, dim3(1), dim3(RBS), 0, st, lz, partial, nupd);
        ED_HIP(hipGetLastError());
        int flag[2] = {0, 0};
        if constexpr (is_real_vec<VT>::value) {
            ED_HIP(hipMemcpyAsync(flag, lz.flag, sizeof(flag), hipMemcpyDeviceToHost, st));
            ED_HIP(hipStreamSynchronize(st));
            ED_REQUIRE(flag[1] >= 0, ED_ERR_ARG,
                       "start vector not real (imaginary part above 1e-10 of its norm): rotate its phase first");
        }
        auto launch_step = [&](int n) {
            const int64_t np =
                m ? launch_hmat(m, u, y, partial, precision, st) : launch_hv(s, hw, u, y, partial, precision, 1, st);
            hipLaunchKernelGGL(lz_reduce_ba_kernel, dim3(1), dim3(RBS), 0, st, lz, pupd, nupd, partial, np, n);
            hipLaunchKernelGGL((lz_update_kernel<VT>), dim3(unsigned(nupd)), dim3(BS), 0, st, lz, n, D, y, u, up,
                               d_basis ? d_basis + int64_t(n) * (basis_ld > 0 ? basis_ld : D) : nullptr, acc, d_ritz,
                               pupd);
            ED_HIP(hipGetLastError());
            VT* t = up;  // w (u_{n+1}) now lives in up
            up = u;
            u = t;
        };
        if (stop) {
            // chunks of check_every steps; the host checks the coefficients of chunk c (copied to pinned memory)
            // while the GPU runs chunk c + 1
            const size_t head = size_t(2 * nsteps + 2);  // flag, a and b: the first doubles of the block
            for (Slot& sl : slot) {
                ED_HIP(hipHostMalloc(reinterpret_cast<void**>(&sl.block), head * sizeof(double)));
                const LzDev h = lz_view(sl.block, nsteps);
                sl.a = h.a;
                sl.b = h.b;
                sl.flag = h.flag;
                ED_HIP(hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming));
            }
            auto post = [&](Slot& sl) {
                ED_HIP(hipMemcpyAsync(sl.block, lz_block, head * sizeof(double), hipMemcpyDeviceToHost, st));
                ED_HIP(hipEventRecord(sl.ev, st));
            };
            int n = 0;
            const int first = std::min(std::max(stop->min_steps, stop->delta + 1), nsteps);
            while (n < first) launch_step(n++);
            post(slot[0]);
            int cur = 0, result = 0, why = 0;
            for (;;) {
                const int pend = n;  // steps that slot[cur] covers
                if (n < nsteps) {
                    const int next = std::min(n + stop->check_every, nsteps);
                    while (n < next) launch_step(n++);
                    post(slot[1 - cur]);
                }
                ED_HIP(hipEventSynchronize(slot[cur].ev));
                const Slot& sl = slot[cur];
                if (sl.flag[0]) {  // breakdown (or a zero start vector): the fraction is exact, no terminator
                    result = std::max(sl.flag[1], 0);
                    why = 2;
                    break;
                }
                const bool met = pend > stop->delta && spectrum_change(sl.a, sl.b, pend, stop->delta, *stop) < stop->tol;
                if (met || pend >= nsteps) {
                    result = pend;
                    why = met ? 1 : 0;
                    break;
                }
                cur = 1 - cur;
            }
            ED_HIP(hipStreamSynchronize(st));
            if (result > 0) {
                std::copy(slot[cur].a, slot[cur].a + result, a);
                std::copy(slot[cur].b, slot[cur].b + result, b);
            }
            *n_out = result;
            if (reason) *reason = why;
            cleanup();
            return;
        }
        int last = -1;
        for (int n = 0; n < nsteps; ++n) {
            launch_step(n);
            last = n;
            if ((n & 63) == 63) {
                ED_HIP(hipMemcpyAsync(flag, lz.flag, sizeof(flag), hipMemcpyDeviceToHost, st));
                ED_HIP(hipStreamSynchronize(st));
                if (flag[0]) break;
            }
        }
        if (last >= 0) {  // b of the last step
            hipLaunchKernelGGL(lz_reduce_b_kernel, dim3(1), dim3(RBS), 0, st, lz, pupd, nupd, last);
            ED_HIP(hipGetLastError());
        }
        ED_HIP(hipMemcpyAsync(flag, lz.flag, sizeof(flag), hipMemcpyDeviceToHost, st));
        ED_HIP(hipStreamSynchronize(st));
        const int cnt = flag[1];
        if (cnt > 0 && a && b) {
            ED_HIP(hipMemcpyAsync(a, lz.a, size_t(cnt) * sizeof(double), hipMemcpyDeviceToHost, st));
            ED_HIP(hipMemcpyAsync(b, lz.b, size_t(cnt) * sizeof(double), hipMemcpyDeviceToHost, st));
            ED_HIP(hipStreamSynchronize(st));
        }
        *n_out = cnt;
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
}

// Calls f with a value of the vector type of `precision` (ed_precision).
template <class F>
void hmat_lanczos(int precision, F&& f) {
    switch (precision) {
        case ED_PREC_C128: f(double2{}); break;
        case ED_PREC_C64: f(float2{}); break;
        case ED_PREC_R64: f(double{}); break;
        default: f(float{}); break;
    }
}

}  // namespace
}  // namespace ed

using namespace ed;

extern "C" {

int ed_lanczos_run(const ed_sector* s, const double* w, const double* d_v0, int nsteps, double* a, double* b,
                   int* n_out, double* d_basis, int precision, void* stream) {
    return guarded([&] {
        ED_REQUIRE(s && s->D >= 0, ED_ERR_STATE, "basis not built");
        ED_REQUIRE(nsteps >= 1 && a && b && n_out, ED_ERR_ARG, "bad Lanczos arguments");
        ED_REQUIRE(precision == 0 || precision == 1, ED_ERR_ARG, "precision must be 0 or 1");
        *n_out = 0;
        if (s->D == 0) return;
        hipStream_t st = static_cast<hipStream_t>(stream);
        if (precision == 0)
            lanczos_run_t<double2>(s, w, d_v0, nsteps, a, b, n_out, reinterpret_cast<double2*>(d_basis), 0, st);
        else
            lanczos_run_t<float2>(s, w, d_v0, nsteps, a, b, n_out, reinterpret_cast<float2*>(d_basis), 1, st);
    });
}

int ed_lanczos_ritz(const ed_sector* s, const double* w, const double* d_v0, int nsteps, const double* ritz,
                    double* d_x, int precision, void* stream) {
    return guarded([&] {
        ED_REQUIRE(s && s->D >= 0, ED_ERR_STATE, "basis not built");
        ED_REQUIRE(nsteps >= 1 && ritz && d_x, ED_ERR_ARG, "bad Ritz-vector arguments");
        ED_REQUIRE(precision == 0 || precision == 1, ED_ERR_ARG, "precision must be 0 or 1");
        if (s->D == 0) return;
        hipStream_t st = static_cast<hipStream_t>(stream);
        int n = 0;
        auto* acc = reinterpret_cast<double2*>(d_x);
        if (precision == 0)
            lanczos_run_t<double2>(s, w, d_v0, nsteps, nullptr, nullptr, &n, nullptr, 0, st, acc, ritz);
        else
            lanczos_run_t<float2>(s, w, d_v0, nsteps, nullptr, nullptr, &n, nullptr, 1, st, acc, ritz);
    });
}

int ed_hmat_lanczos_run(const ed_hmat* m, const double* d_v0, int nsteps, double* a, double* b, int* n_out,
                        int precision, void* stream) {
    return guarded([&] {
        ED_REQUIRE(m, ED_ERR_ARG, "null matrix");
        const ed_sector* s = hmat_sector(m);
        ED_REQUIRE(nsteps >= 1 && a && b && n_out, ED_ERR_ARG, "bad Lanczos arguments");
        hmat_require_precision(m, precision);
        *n_out = 0;
        if (s->D == 0) return;
        hipStream_t st = static_cast<hipStream_t>(stream);
        const double w0[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        hmat_lanczos(precision, [&](auto vt) {
            using VT = decltype(vt);
            lanczos_run_t<VT>(s, w0, d_v0, nsteps, a, b, n_out, nullptr, precision, st, nullptr, nullptr, m);
        });
    });
}

int ed_hmat_lanczos_run_basis(const ed_hmat* m, const double* d_v0, int nsteps, double* a, double* b, int* n_out,
                              void* d_basis, int64_t ld, int precision, void* stream) {
    return guarded([&] {
        ED_REQUIRE(m, ED_ERR_ARG, "null matrix");
        const ed_sector* s = hmat_sector(m);
        ED_REQUIRE(nsteps >= 1 && a && b && n_out && d_basis && ld >= s->D, ED_ERR_ARG,
                   "bad Lanczos arguments (the basis needs nsteps rows of ld >= D elements)");
        hmat_require_precision(m, precision);
        *n_out = 0;
        if (s->D == 0) return;
        hipStream_t st = static_cast<hipStream_t>(stream);
        const double w0[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        hmat_lanczos(precision, [&](auto vt) {
            using VT = decltype(vt);
            lanczos_run_t<VT>(s, w0, d_v0, nsteps, a, b, n_out, static_cast<VT*>(d_basis), precision, st, nullptr,
                              nullptr, m, nullptr, nullptr, ld);
        });
    });
}

int ed_hmat_lanczos_run_tol(const ed_hmat* m, const double* d_v0, int max_steps, const ed_lanczos_stop* stop,
                            double* a, double* b, int* n_out, int* reason, int precision, void* stream) {
    return guarded([&] {
        ED_REQUIRE(m, ED_ERR_ARG, "null matrix");
        const ed_sector* s = hmat_sector(m);
        ED_REQUIRE(max_steps >= 1 && a && b && n_out && reason, ED_ERR_ARG, "bad Lanczos arguments");
        check_stop_rule(stop);
        hmat_require_precision(m, precision);
        *n_out = 0;
        *reason = 2;
        if (s->D == 0) return;
        hipStream_t st = static_cast<hipStream_t>(stream);
        const double w0[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        hmat_lanczos(precision, [&](auto vt) {
            using VT = decltype(vt);
            lanczos_run_t<VT>(s, w0, d_v0, max_steps, a, b, n_out, nullptr, precision, st, nullptr, nullptr, m, stop,
                              reason);
        });
    });
}

int ed_spectrum_change(const double* a, const double* b, int n, int delta, const ed_lanczos_stop* stop,
                       double* out) {
    return guarded([&] {
        check_stop_rule(stop);
        ED_REQUIRE(a && b && out && delta >= 1 && n > delta, ED_ERR_ARG, "need n > delta >= 1 coefficients");
        *out = spectrum_change(a, b, n, delta, *stop);
    });
}

int ed_hmat_lanczos_ritz(const ed_hmat* m, const double* d_v0, int nsteps, const double* ritz, double* d_x,
                         int precision, void* stream) {
    return guarded([&] {
        ED_REQUIRE(m, ED_ERR_ARG, "null matrix");
        const ed_sector* s = hmat_sector(m);
        ED_REQUIRE(nsteps >= 1 && ritz && d_x, ED_ERR_ARG, "bad Ritz-vector arguments");
        hmat_require_precision(m, precision);
        if (s->D == 0) return;
        hipStream_t st = static_cast<hipStream_t>(stream);
        const double w0[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        int n = 0;
        auto* acc = reinterpret_cast<double2*>(d_x);
        hmat_lanczos(precision, [&](auto vt) {
            using VT = decltype(vt);
            lanczos_run_t<VT>(s, w0, d_v0, nsteps, nullptr, nullptr, &n, nullptr, precision, st, acc, ritz, m);
        });
    });
}

int ed_cf_eval_tail(const double* a, const double* b, int n, const double* ws, int n_w, double e0, double eta,
                    double factor, double a_inf, double b_inf, double* out);

int ed_cf_eval(const double* a, const double* b, int n, const double* ws, int n_w, double e0, double eta,
               double factor, double* out) {
    return ed_cf_eval_tail(a, b, n, ws, n_w, e0, eta, factor, 0.0, 0.0, out);
}

int ed_cf_eval_tail(const double* a, const double* b, int n, const double* ws, int n_w, double e0, double eta,
                    double factor, double a_inf, double b_inf, double* out) {
    return guarded([&] {
        ED_REQUIRE(n >= 1 && n_w >= 0, ED_ERR_ARG, "bad continued-fraction sizes");
        if (n_w == 0) return;
        hipStream_t st = nullptr;
        std::vector<double> ab(a, a + n), bb(b, b + n), wv(ws, ws + n_w);
        double* da = dev_upload(ab, st);
        double* db = dev_upload(bb, st);
        double* dw = dev_upload(wv, st);
        double* dout = dev_alloc<double>(size_t(n_w));
        hipLaunchKernelGGL(cf_kernel, dim3(unsigned((n_w + 127) / 128)), dim3(128), 0, st, da, db, n, dw, n_w, e0, eta,
                           factor, a_inf, b_inf, dout);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpy(out, dout, size_t(n_w) * sizeof(double), hipMemcpyDeviceToHost);
        dev_free(da);
        dev_free(db);
        dev_free(dw);
        dev_free(dout);
        ED_HIP(e);
    });
}

}  // extern "C"

namespace {

constexpr int DOTS_MAX_WGS = 1024;  // workgroups of dots_kernel per vector

int dots_workgroups(int64_t D) { return int(std::min<int64_t>(std::max<int64_t>((D + BS - 1) / BS, 1), DOTS_MAX_WGS)); }

// out[k] = <V_k, w>: partials (2 * nwg * nv doubles) and the reduced dots (2 * nv) in device memory the caller
// provides, then copied to the host out.
hipError_t vec_dots_into(const double* d_V, int nv, int64_t ld, const double* d_w, int64_t D, double* partial,
                         double* dout, double* out, hipStream_t st) {
    const int nwg = dots_workgroups(D);
    hipLaunchKernelGGL(dots_kernel, dim3(unsigned(nwg), unsigned(nv)), dim3(BS), 0, st, d_V, ld, d_w, D, partial, nwg);
    hipLaunchKernelGGL(dots_reduce_kernel, dim3(unsigned(nv)), dim3(BS), 0, st, partial, nwg, dout);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, size_t(2 * nv) * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e;
}

// Fixed scratch of ed_vec_dots_noalloc: the partials of at most NOALLOC_MAX_VECTORS vectors, then their
// reduced dots (2 MiB + 2 KiB).  Every GPU the library runs on holds its own copy until the process ends.
constexpr int NOALLOC_MAX_VECTORS = 128;
constexpr int64_t NOALLOC_PARTIALS = int64_t(2) * DOTS_MAX_WGS * NOALLOC_MAX_VECTORS;
static __device__ __align__(256) double scratch[NOALLOC_PARTIALS + 2 * NOALLOC_MAX_VECTORS];

// Device address of `scratch` on the current device, looked up once per device.  The array's name on the
// host is only a stand-in for the symbol: kernels and copies get this address.
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[MAX_DEVICES] = {};

double* noalloc_scratch() {
    int dev = 0;
    ED_HIP(hipGetDevice(&dev));
    ED_REQUIRE(dev >= 0 && dev < MAX_DEVICES, ED_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        ED_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ED_REQUIRE(p, ED_ERR_HIP, "no device address for the scratch array");
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    return g_scratch_ptr[dev];
}

}  // namespace

extern "C" {

int ed_vec_dots(const double* d_V, int nv, int64_t ld, const double* d_w, int64_t D, double* out, void* stream) {
    return guarded([&] {
        ED_REQUIRE(nv >= 0 && nv <= 65535 && D >= 0, ED_ERR_ARG, "bad sizes");
        if (nv == 0) return;
        hipStream_t st = static_cast<hipStream_t>(stream);
        const int nwg = dots_workgroups(D);
        double* partial = dev_alloc<double>(size_t(2) * nwg * nv);
        double* dout = dev_alloc<double>(size_t(2 * nv));
        const hipError_t e = vec_dots_into(d_V, nv, ld, d_w, D, partial, dout, out, st);
        dev_free(partial);
        dev_free(dout);
        ED_HIP(e);
    });
}

int ed_vec_dots_noalloc(const double* d_V, int nv, int64_t ld, const double* d_w, int64_t D, double* out,
                        void* stream) {
    return guarded([&] {
        ED_REQUIRE(nv >= 0 && D >= 0, ED_ERR_ARG, "bad sizes");
        ED_REQUIRE(nv <= NOALLOC_MAX_VECTORS, ED_ERR_CAPACITY,
                   "ed_vec_dots_noalloc takes at most 128 vectors (its fixed scratch holds their partials)");
        if (nv == 0) return;
        double* scr = noalloc_scratch();
        // dots_kernel writes every partial it reads back, so nothing in the scratch needs clearing
        ED_HIP(vec_dots_into(d_V, nv, ld, d_w, D, scr, scr + NOALLOC_PARTIALS, out, static_cast<hipStream_t>(stream)));
    });
}

static int combine_impl(const double* d_V, int nv, int64_t ld, const double* c, double* d_out, int64_t D,
                        int negative, void* stream) {
    return guarded([&] {
        if (D == 0) return;
        hipStream_t st = static_cast<hipStream_t>(stream);
        std::vector<double> cv(c, c + 2 * nv);
        if (cv.empty()) cv.assign(2, 0.0);
        double* dc = dev_upload(cv, st);
        hipLaunchKernelGGL(combine_kernel, dim3(unsigned((D + BS - 1) / BS)), dim3(BS), 0, st, d_V, nv, ld, dc, d_out,
                           D, negative);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        dev_free(dc);
        ED_HIP(e);
    });
}

int ed_vec_axpys(const double* d_V, int nv, int64_t ld, const double* c, double* d_w, int64_t D, void* stream) {
    return combine_impl(d_V, nv, ld, c, d_w, D, 1, stream);
}

int ed_vec_combine(const double* d_V, int nv, int64_t ld, const double* c, double* d_out, int64_t D, void* stream) {
    return combine_impl(d_V, nv, ld, c, d_out, D, 0, stream);
}

int ed_vec_norm2(const double* d_w, int64_t D, double* out, void* stream) {
    double z[2];
    int rc = ed_vec_dots(d_w, 1, 0, d_w, D, z, stream);
    *out = z[0];
    return rc;
}

int ed_vec_scale(double* d_w, int64_t D, double cr, double ci, void* stream) {
    return guarded([&] {
        if (D == 0) return;
        hipStream_t st = static_cast<hipStream_t>(stream);
        hipLaunchKernelGGL(scale_kernel, dim3(unsigned((D + BS - 1) / BS)), dim3(BS), 0, st, d_w, D, cr, ci);
        ED_HIP(hipGetLastError());
    });
}

}  // extern "C"
