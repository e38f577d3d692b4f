// Batched Lanczos over sectors of uneven size, and the H·v autotuner (gfx950).
//
// The reference computes every momentum sector (tJ_spin_half_ladder/api.py:45-59)
// and every q of a spectrum (api.py:116-161) one after the other, each with its
// own matrices and Lanczos loop.  Here a batch holds many sectors (instances)
// at once:
//  * the vectors of all instances live in one pool, instance i at rows
//    [off_i, off_i + D_i);
//  * the work is cut into jobs (instance, row range) of `tiles` row tiles of
//    256 each; every batched kernel runs one persistent workgroup per job, so
//    one launch covers all instances whatever their sizes, and a workgroup
//    stages its instance's tables in LDS once for all its rows;
//  * one Lanczos step of all instances is four launches (H·v with the
//    <u, Hu> partials, per-instance reduction for a_n, fused update with the
//    ||w||^2 partials, per-instance reduction for b_{n+1}); an instance that
//    has reached its own step count or broken down (b < 1e-8) is skipped by
//    every kernel through its device flag;
//  * the whole run, all steps of all instances, is captured once into a HIP
//    graph and replayed as one launch (no memset nodes: clearing is a kernel);
//  * a second pass with per-instance coefficients y_n accumulates
//    x = sum_n y_n v_n, which gives Ritz vectors without storing the basis
//    (two-pass Lanczos: the recurrence is deterministic, so pass two sees
//    the same vectors as pass one).
//
// The autotuner times the batched H·v (and the single-sector one) over the
// gather group size G (2, 4, 8 gathers in flight per lane) and the tiles per
// workgroup (1..32, i.e. 64..2048 states per wavefront) and keeps the fastest.

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <type_traits>
#include <vector>

#include "ed_hmatrow.h"
#include "ed_hvuni.h"

namespace ed {
namespace {

constexpr int BS = 256;

struct BInst {
    SectorView s;
    LdsLayout lay;
    HvWeights w;
    StoredView m;  // the instance's stored H when `stored`
    int64_t voff;  // first pool row of the instance
    int job0, job1;
    int lds;     // tables staged in LDS (matrix-free instances)
    int stored;  // rows from the stored H instead of the matrix-free walk
};

struct BJob {
    int inst;
    int pad;
    int64_t r0, r1;  // rows [r0, r1) of the instance
};

struct BLz {
    double* a;     // [n_inst * ld]
    double* b;     // [n_inst * (ld + 1)]
    double* s;     // [n_inst * (ld + 1)]
    int* flag;     // [2 * n_inst]: stopped, steps done
    int* nsteps;   // [n_inst]
    double* ritz;  // [n_inst * ld] or null
    int ld;
};

__device__ inline int job_index(int nblocks) {
    int wg = blockIdx.x;
    const int per = nblocks / 8;
    if (per > 0 && wg < per * 8) wg = (wg & 7) * per + (wg >> 3);
    return wg;
}

__device__ inline void wg_sum2(double& re, double& im) { workgroup_sum2<BS>(re, im); }

__device__ inline bool inactive(const BLz& lz, int inst, int step) {
    return lz.flag && (lz.flag[2 * inst] || step >= lz.nsteps[inst]);
}

// y = H u for every job's rows; partial[job] = <u, Hu> over them.
template <int G>
__global__ void __launch_bounds__(BS) bhv_kernel(const BInst* __restrict__ insts, const BJob* __restrict__ jobs,
                                                 int njobs, BLz lz, int step, const double2* __restrict__ x,
                                                 double2* __restrict__ y, double* __restrict__ partial, int chunk) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int jb = xcd_order(blockIdx.x, njobs, chunk);  // the eight XCDs on neighbouring chunks of jobs
    const BJob job = jobs[jb];
    const BInst& I = insts[job.inst];
    if (inactive(lz, job.inst, step)) {
        if (threadIdx.x == 0 && partial) partial[2 * jb] = partial[2 * jb + 1] = 0.0;
        return;
    }
    const double2* xi = x + I.voff;
    double2* yi = y + I.voff;
    C2 dot{0.0, 0.0};
    if (I.stored) {
        const StoredView m = I.m;
        double* V = reinterpret_cast<double*>(lds);
        for (int i = threadIdx.x; i < 2 * m.n_vals; i += BS) V[i] = m.vals[i];
        __syncthreads();
        auto rows = [&](auto packed) {
            for (int64_t f = job.r0 + threadIdx.x; f < job.r1; f += BS) {
                C2 xf;
                const C2 acc = stored_row<double2, 4, false, decltype(packed)::value>(m, V, xi, f, xf);
                yi[f] = double2{acc.re, acc.im};
                dot = cadd(dot, cmul(C2{xf.re, -xf.im}, acc));
            }
        };
        if (m.shift)
            rows(std::true_type{});
        else
            rows(std::false_type{});
    } else {
        const SectorView s = I.s;
        const Tabs T = I.lds ? stage_tabs<BS>(s, I.lay, lds) : global_tabs(s);
        __syncthreads();
        const HvWeights w = I.w;
        // the wavefront-uniform walk where the instance's tables (with the chunked ranks) are staged; a job's
        // rows start at a multiple of 256, so each wavefront covers 64 consecutive rows
        const bool uni = I.lds && I.lay.rt >= 0;
        for (int64_t r = job.r0; r < job.r1; r += BS) {
            const int64_t f = r + threadIdx.x;
            if (uni) {
                const int64_t f0 = r + int(threadIdx.x & ~63u);
                if (f0 >= job.r1) continue;
                const bool live = f < job.r1;
                C2 acc, xf;
                if (uniform_row<double2, G>(s, T, I.lay, w, xi, f, f0, live, acc, xf)) {
                    if (live) {
                        yi[f] = double2{acc.re, acc.im};
                        dot = cadd(dot, cmul(C2{xf.re, -xf.im}, acc));
                    }
                    continue;
                }
            }
            if (f >= job.r1) continue;
            C2 xf;
            const C2 acc = hv_row<double2, G>(s, T, w, xi, f, xf);
            yi[f] = double2{acc.re, acc.im};
            dot = cadd(dot, cmul(C2{xf.re, -xf.im}, acc));
        }
    }
    if (partial) {
        wg_sum2(dot.re, dot.im);
        if (threadIdx.x == 0) {
            partial[2 * jb] = dot.re;
            partial[2 * jb + 1] = dot.im;
        }
    }
}

// Per instance: sum of its jobs' partials (fixed order).
__device__ inline void inst_sum(const BInst& I, const double* partial, double& re, double& im) {
    re = im = 0.0;
    for (int k = I.job0 + int(threadIdx.x); k < I.job1; k += BS) {
        re += partial[2 * k];
        im += partial[2 * k + 1];
    }
    wg_sum2(re, im);
}

__global__ void __launch_bounds__(BS) bsum_kernel(const BInst* insts, const double* partial, double* out) {
    double re, im;
    inst_sum(insts[blockIdx.x], partial, re, im);
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = re;
        out[2 * blockIdx.x + 1] = im;
    }
}

// u = v0 on the job's rows, the previous vector and the accumulator cleared;
// partial = ||u||^2.
__global__ void __launch_bounds__(BS) binit_kernel(const BInst* insts, const BJob* jobs, int njobs, BLz lz,
                                                   const double2* v0, double2* u, double2* up, double2* acc,
                                                   double* partial) {
    const int jb = job_index(njobs);
    const BJob job = jobs[jb];
    const BInst& I = insts[job.inst];
    const int64_t o = I.voff;
    double n2 = 0.0;
    for (int64_t f = job.r0 + threadIdx.x; f < job.r1; f += BS) {
        const double2 v = v0[o + f];
        u[o + f] = v;
        up[o + f] = double2{0.0, 0.0};
        if (acc) acc[o + f] = double2{0.0, 0.0};
        n2 += v.x * v.x + v.y * v.y;
    }
    double z = 0.0;
    wg_sum2(n2, z);
    if (threadIdx.x == 0) {
        partial[2 * jb] = n2;
        partial[2 * jb + 1] = 0.0;
    }
}

__global__ void __launch_bounds__(BS) binit_reduce_kernel(const BInst* insts, BLz lz, const double* partial) {
    const int i = blockIdx.x;
    double re, im;
    inst_sum(insts[i], partial, re, im);
    if (threadIdx.x == 0) {
        const int ld1 = lz.ld + 1;
        lz.b[i * ld1] = 0.0;
        lz.s[i * ld1] = re > 0.0 ? 1.0 / std::sqrt(re) : 0.0;
        lz.flag[2 * i] = (re > 0.0 && lz.nsteps[i] > 0) ? 0 : 1;
        lz.flag[2 * i + 1] = 0;
    }
}

__global__ void __launch_bounds__(BS) breduce_a_kernel(const BInst* insts, BLz lz, const double* partial, int step) {
    const int i = blockIdx.x;
    if (inactive(lz, i, step)) return;
    double re, im;
    inst_sum(insts[i], partial, re, im);
    if (threadIdx.x == 0) {
        const double sn = lz.s[i * (lz.ld + 1) + step];
        lz.a[i * lz.ld + step] = sn * sn * re;
    }
}

// w = s_n Hu - a_n s_n u_n - b_n s_{n-1} u_{n-1} into up (the caller swaps
// the roles of u and up after the step); acc += y_n s_n u_n in pass two;
// partial = ||w||^2.
__global__ void __launch_bounds__(BS) bupdate_kernel(const BInst* insts, const BJob* jobs, int njobs, BLz lz, int step,
                                                     const double2* __restrict__ y, const double2* __restrict__ u,
                                                     double2* __restrict__ up, double2* __restrict__ acc,
                                                     double* __restrict__ partial) {
    const int jb = job_index(njobs);
    const BJob job = jobs[jb];
    const int i = job.inst;
    if (inactive(lz, i, step)) {
        if (threadIdx.x == 0) partial[2 * jb] = partial[2 * jb + 1] = 0.0;
        return;
    }
    const int ld1 = lz.ld + 1;
    const double sn = lz.s[i * ld1 + step], an = lz.a[i * lz.ld + step], bn = lz.b[i * ld1 + step];
    const double snm = step > 0 ? lz.s[i * ld1 + step - 1] : 0.0;
    const double cy = (acc && lz.ritz) ? lz.ritz[i * lz.ld + step] * sn : 0.0;
    const int64_t o = insts[i].voff;
    double n2 = 0.0;
    for (int64_t f = job.r0 + threadIdx.x; f < job.r1; f += BS) {
        const double2 yv = y[o + f], uv = u[o + f], pv = up[o + f];
        const double vr = sn * uv.x, vi = sn * uv.y;
        const double hr = sn * yv.x, hi = sn * yv.y;
        const double qr = snm * pv.x, qi = snm * pv.y;
        const double wr = hr - an * vr - bn * qr;
        const double wi = hi - an * vi - bn * qi;
        up[o + f] = double2{wr, wi};
        if (acc) {
            const double2 av = acc[o + f];
            acc[o + f] = double2{av.x + cy * uv.x, av.y + cy * uv.y};
        }
        n2 += wr * wr + wi * wi;
    }
    double z = 0.0;
    wg_sum2(n2, z);
    if (threadIdx.x == 0) {
        partial[2 * jb] = n2;
        partial[2 * jb + 1] = 0.0;
    }
}

__global__ void __launch_bounds__(BS) breduce_b_kernel(const BInst* insts, BLz lz, const double* partial, int step) {
    const int i = blockIdx.x;
    if (inactive(lz, i, step)) return;
    double re, im;
    inst_sum(insts[i], partial, re, im);
    if (threadIdx.x == 0) {
        const int ld1 = lz.ld + 1;
        const double bnp = std::sqrt(re);
        lz.b[i * ld1 + step + 1] = bnp;
        lz.s[i * ld1 + step + 1] = 1.0 / bnp;
        lz.flag[2 * i + 1] = step + 1;
        if (bnp < 1e-8) lz.flag[2 * i] = 1;
    }
}

template <class F>
double time_ms(hipStream_t st, int reps, F&& launch) {
    hipEvent_t e0, e1;
    ED_HIP(hipEventCreate(&e0));
    ED_HIP(hipEventCreate(&e1));
    launch();  // warm-up
    ED_HIP(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r) launch();
    ED_HIP(hipEventRecord(e1, st));
    ED_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    ED_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return double(ms) / reps;
}

constexpr int TUNE_TILES[] = {1, 2, 4, 8, 16, 32};
constexpr int TUNE_GATHER[] = {2, 4, 8};

}  // namespace
}  // namespace ed

using namespace ed;

struct ed_batch {
    std::vector<const ed_sector*> sectors;
    std::vector<BInst> inst;
    std::vector<BJob> jobs;
    std::vector<int64_t> offsets;  // n + 1
    int tiles = 4, gather = 4;
    int lds_bytes = 0;
    BInst* d_inst = nullptr;
    BJob* d_jobs = nullptr;
    double* d_partial = nullptr;
    size_t partial_cap = 0;
    hipStream_t capture = nullptr;

    int n() const { return int(sectors.size()); }

    void build_jobs() {
        jobs.clear();
        const int64_t rows = int64_t(tiles) * BS;
        for (int i = 0; i < n(); ++i) {
            inst[i].job0 = int(jobs.size());
            const int64_t D = sectors[i]->D;
            for (int64_t r = 0; r < D; r += rows) jobs.push_back(BJob{i, 0, r, std::min(D, r + rows)});
            inst[i].job1 = int(jobs.size());
        }
        ED_HIP(hipDeviceSynchronize());  // no kernel may still read the old lists
        dev_free(d_inst);
        dev_free(d_jobs);
        d_inst = nullptr;
        d_jobs = nullptr;
        d_inst = dev_upload(inst, nullptr);
        d_jobs = dev_upload(jobs, nullptr);
        ED_HIP(hipDeviceSynchronize());
        const size_t need = 2 * std::max<size_t>(jobs.size(), 1);
        if (need > partial_cap) {
            dev_free(d_partial);
            d_partial = nullptr;
            d_partial = dev_alloc<double>(need);
            partial_cap = need;
        }
    }

    void launch_hv(const double2* x, double2* y, const BLz& lz, int step, hipStream_t st) const {
        const int nj = int(jobs.size());
        if (nj == 0) return;
        const dim3 grid{unsigned(nj), 1u, 1u};
        // XCD chunks of about 131,072 rows (ED_BATCH_XCD_CHUNK_ROWS overrides, 0 = one eighth per XCD)
        static const int64_t chunk_rows = [] {
            const char* e = std::getenv("ED_BATCH_XCD_CHUNK_ROWS");
            return e ? int64_t(std::atoll(e)) : int64_t(131072);
        }();
        const int chunk = chunk_rows > 0 ? int(std::max<int64_t>(1, chunk_rows / (int64_t(tiles) * BS))) : 0;
        if (gather == 2)
            hipLaunchKernelGGL(bhv_kernel<2>, grid, dim3(BS), lds_bytes, st, d_inst, d_jobs, nj, lz, step, x, y,
                               d_partial, chunk);
        else if (gather == 8)
            hipLaunchKernelGGL(bhv_kernel<8>, grid, dim3(BS), lds_bytes, st, d_inst, d_jobs, nj, lz, step, x, y,
                               d_partial, chunk);
        else
            hipLaunchKernelGGL(bhv_kernel<4>, grid, dim3(BS), lds_bytes, st, d_inst, d_jobs, nj, lz, step, x, y,
                               d_partial, chunk);
        ED_HIP(hipGetLastError());
    }

    ~ed_batch() {
        dev_free(d_inst);
        dev_free(d_jobs);
        dev_free(d_partial);
        if (capture) (void)hipStreamDestroy(capture);
    }
};

extern "C" {

int ed_batch_create(ed_sector* const* sectors, const double* weights, int n, ed_batch** out) {
    return guarded([&] {
        ED_REQUIRE(sectors && weights && out && n >= 1, ED_ERR_ARG, "bad batch arguments");
        *out = nullptr;
        auto* b = new ed_batch();
        try {
            int64_t off = 0;
            for (int i = 0; i < n; ++i) {
                ed_sector* s = sectors[i];
                ED_REQUIRE(s && s->D >= 0, ED_ERR_STATE, "batch sector without a basis");
                b->sectors.push_back(s);
                BInst I{};
                I.s = s->view();
                I.lay = hv_uniform_enabled() ? lds_layout_ranks(s) : lds_layout(s);
                I.lds = I.lay.bytes <= LDS_CAP;
                if (I.lds) b->lds_bytes = std::max(b->lds_bytes, I.lay.bytes);
                I.w = make_weights(weights + 8 * i);
                I.voff = off;
                b->offsets.push_back(off);
                off += s->D;
                b->inst.push_back(I);
            }
            b->offsets.push_back(off);
            b->build_jobs();
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

int ed_batch_create_stored(const ed_hmat* const* mats, int n, ed_batch** out) {
    return guarded([&] {
        ED_REQUIRE(mats && out && n >= 1, ED_ERR_ARG, "bad batch arguments");
        *out = nullptr;
        auto* b = new ed_batch();
        try {
            int64_t off = 0;
            for (int i = 0; i < n; ++i) {
                ED_REQUIRE(mats[i], ED_ERR_ARG, "null stored matrix");
                const ed_sector* s = hmat_sector(mats[i]);
                b->sectors.push_back(s);
                BInst I{};
                I.s = s->view();
                I.lay = lds_layout(s);
                I.w = hmat_weights(mats[i]);
                I.m = hmat_view(mats[i]);
                I.stored = 1;
                b->lds_bytes = std::max(b->lds_bytes, I.m.n_vals * 16);
                I.voff = off;
                b->offsets.push_back(off);
                off += s->D;
                b->inst.push_back(I);
            }
            b->offsets.push_back(off);
            b->build_jobs();
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

int ed_batch_destroy(ed_batch* b) {
    delete b;
    return ED_OK;
}

int ed_batch_offsets(const ed_batch* b, int64_t* offsets) {
    return guarded([&] {
        ED_REQUIRE(b && offsets, ED_ERR_ARG, "null argument");
        std::copy(b->offsets.begin(), b->offsets.end(), offsets);
    });
}

int ed_batch_set_config(ed_batch* b, int tiles, int gather) {
    return guarded([&] {
        ED_REQUIRE(b, ED_ERR_ARG, "null batch");
        ED_REQUIRE(tiles >= 1 && tiles <= 1024, ED_ERR_ARG, "tiles must be in [1, 1024]");
        ED_REQUIRE(gather == 2 || gather == 4 || gather == 8, ED_ERR_ARG, "gather must be 2, 4 or 8");
        b->gather = gather;
        if (tiles != b->tiles) {
            b->tiles = tiles;
            b->build_jobs();
        }
    });
}

int ed_batch_get_config(const ed_batch* b, int* tiles, int* gather, int* n_jobs) {
    return guarded([&] {
        ED_REQUIRE(b, ED_ERR_ARG, "null batch");
        if (tiles) *tiles = b->tiles;
        if (gather) *gather = b->gather;
        if (n_jobs) *n_jobs = int(b->jobs.size());
    });
}

int ed_batch_hv(ed_batch* b, const double* d_x, double* d_y, double* dots, void* stream) {
    return guarded([&] {
        ED_REQUIRE(b && d_x && d_y, ED_ERR_ARG, "null argument");
        hipStream_t st = static_cast<hipStream_t>(stream);
        BLz none{};
        b->launch_hv(reinterpret_cast<const double2*>(d_x), reinterpret_cast<double2*>(d_y), none, 0, st);
        if (dots) {
            double* d_out = dev_alloc<double>(size_t(2 * b->n()));
            hipLaunchKernelGGL(bsum_kernel, dim3(unsigned(b->n())), dim3(BS), 0, st, b->d_inst, b->d_partial, d_out);
            hipError_t e = hipGetLastError();
            if (e == hipSuccess)
                e = hipMemcpyAsync(dots, d_out, size_t(2 * b->n()) * sizeof(double), hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            dev_free(d_out);
            ED_HIP(e);
        }
    });
}

int ed_batch_tune(ed_batch* b, const double* d_x, double* d_y, int* tiles, int* gather, double* best_ms,
                  void* stream) {
    return guarded([&] {
        ED_REQUIRE(b && d_x && d_y, ED_ERR_ARG, "null argument");
        hipStream_t st = static_cast<hipStream_t>(stream);
        const auto* x = reinterpret_cast<const double2*>(d_x);
        auto* y = reinterpret_cast<double2*>(d_y);
        BLz none{};
        double best = 1e300;
        int bt = b->tiles, bg = b->gather;
        for (int t : TUNE_TILES) {
            b->tiles = t;
            b->build_jobs();
            for (int g : TUNE_GATHER) {
                b->gather = g;
                const double ms = time_ms(st, 3, [&] { b->launch_hv(x, y, none, 0, st); });
                if (ms < best) {
                    best = ms;
                    bt = t;
                    bg = g;
                }
            }
        }
        b->tiles = bt;
        b->gather = bg;
        b->build_jobs();
        if (tiles) *tiles = bt;
        if (gather) *gather = bg;
        if (best_ms) *best_ms = best;
    });
}

// Lanczos on every instance: nsteps[i] steps from the pool vector d_v0 (any
// norm; instance rows at ed_batch_offsets).  a[i * ld + n] and
// b[i * (ld + 1) + n] (b_0 = 0, b_{n_out} = norm of the last residual)
// receive the coefficients and n_out[i] the steps done.
// With ritz (host, [n * ld]) and d_acc, also d_acc = sum_n ritz[i * ld + n] v_n.
int ed_batch_lanczos(ed_batch* b, const double* d_v0, const int* nsteps, int ld, const double* ritz, double* d_acc,
                     double* a, double* bcoef, int* n_out, int use_graph, void* stream) {
    return guarded([&] {
        ED_REQUIRE(b && d_v0 && nsteps && a && bcoef && n_out, ED_ERR_ARG, "null argument");
        ED_REQUIRE(ld >= 1, ED_ERR_ARG, "ld must be >= 1");
        const int n = b->n();
        int maxs = 0;
        for (int i = 0; i < n; ++i) {
            ED_REQUIRE(nsteps[i] >= 0 && nsteps[i] <= ld, ED_ERR_ARG, "nsteps must be in [0, ld]");
            maxs = std::max(maxs, nsteps[i]);
        }
        ED_REQUIRE((ritz == nullptr) == (d_acc == nullptr), ED_ERR_ARG, "ritz and acc go together");
        hipStream_t st = static_cast<hipStream_t>(stream);
        const int64_t total = b->offsets.back();
        std::vector<void*> owned;
        auto alloc = [&](size_t bytes) {
            void* p = dev_alloc<unsigned char>(bytes);
            owned.push_back(p);
            return p;
        };
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        auto cleanup = [&] {
            if (exec) (void)hipGraphExecDestroy(exec);
            if (graph) (void)hipGraphDestroy(graph);
            for (void* p : owned) dev_free(p);
        };
        try {
            auto* pool = static_cast<double2*>(alloc(size_t(3 * std::max<int64_t>(total, 1)) * sizeof(double2)));
            double2* u = pool;
            double2* up = pool + total;
            double2* y = pool + 2 * total;
            BLz lz{};
            lz.ld = ld;
            lz.a = static_cast<double*>(alloc(size_t(n) * ld * sizeof(double)));
            lz.b = static_cast<double*>(alloc(size_t(n) * (ld + 1) * sizeof(double)));
            lz.s = static_cast<double*>(alloc(size_t(n) * (ld + 1) * sizeof(double)));
            lz.flag = static_cast<int*>(alloc(size_t(2 * n) * sizeof(int)));
            lz.nsteps = static_cast<int*>(alloc(size_t(n) * sizeof(int)));
            ED_HIP(hipMemcpyAsync(lz.nsteps, nsteps, size_t(n) * sizeof(int), hipMemcpyHostToDevice, st));
            if (ritz) {
                lz.ritz = static_cast<double*>(alloc(size_t(n) * ld * sizeof(double)));
                ED_HIP(hipMemcpyAsync(lz.ritz, ritz, size_t(n) * ld * sizeof(double), hipMemcpyHostToDevice, st));
            }
            ED_HIP(hipStreamSynchronize(st));
            auto* acc = reinterpret_cast<double2*>(d_acc);
            const int nj = int(b->jobs.size());
            const dim3 gj{unsigned(std::max(nj, 1)), 1u, 1u}, gi{unsigned(n), 1u, 1u};
            hipStream_t cs = st;
            if (use_graph) {
                if (!b->capture) ED_HIP(hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking));
                cs = b->capture;
                ED_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
            }
            try {
                if (nj > 0) {
                    hipLaunchKernelGGL(binit_kernel, gj, dim3(BS), 0, cs, b->d_inst, b->d_jobs, nj, lz,
                                       reinterpret_cast<const double2*>(d_v0), u, up, acc, b->d_partial);
                    hipLaunchKernelGGL(binit_reduce_kernel, gi, dim3(BS), 0, cs, b->d_inst, lz, b->d_partial);
                    for (int step = 0; step < maxs; ++step) {
                        b->launch_hv(u, y, lz, step, cs);
                        hipLaunchKernelGGL(breduce_a_kernel, gi, dim3(BS), 0, cs, b->d_inst, lz, b->d_partial, step);
                        hipLaunchKernelGGL(bupdate_kernel, gj, dim3(BS), 0, cs, b->d_inst, b->d_jobs, nj, lz, step, y,
                                           u, up, acc, b->d_partial);
                        hipLaunchKernelGGL(breduce_b_kernel, gi, dim3(BS), 0, cs, b->d_inst, lz, b->d_partial, step);
                        std::swap(u, up);
                    }
                    ED_HIP(hipGetLastError());
                }
            } catch (...) {
                if (use_graph) {
                    hipGraph_t g = nullptr;
                    (void)hipStreamEndCapture(cs, &g);
                    if (g) (void)hipGraphDestroy(g);
                }
                throw;
            }
            if (use_graph) {
                ED_HIP(hipStreamEndCapture(cs, &graph));
                ED_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
                ED_HIP(hipGraphLaunch(exec, st));
            }
            std::vector<int> flag(size_t(2 * n), 0);
            std::vector<double> ha(size_t(n) * ld), hb(size_t(n) * (ld + 1));
            if (nj > 0) {
                ED_HIP(hipMemcpyAsync(flag.data(), lz.flag, flag.size() * sizeof(int), hipMemcpyDeviceToHost, st));
                ED_HIP(hipMemcpyAsync(ha.data(), lz.a, ha.size() * sizeof(double), hipMemcpyDeviceToHost, st));
                ED_HIP(hipMemcpyAsync(hb.data(), lz.b, hb.size() * sizeof(double), hipMemcpyDeviceToHost, st));
            }
            ED_HIP(hipStreamSynchronize(st));
            for (int i = 0; i < n; ++i) {
                const int cnt = nj > 0 ? flag[2 * i + 1] : 0;
                n_out[i] = cnt;
                for (int k = 0; k < ld; ++k) a[size_t(i) * ld + k] = k < cnt ? ha[size_t(i) * ld + k] : 0.0;
                for (int k = 0; k <= ld; ++k)
                    bcoef[size_t(i) * (ld + 1) + k] = k <= cnt ? hb[size_t(i) * (ld + 1) + k] : 0.0;
            }
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
    });
}

// Single-sector autotuner: times the LDS H·v over the same candidates and
// stores the fastest (tiles, gather) in the sector for later H·v calls.
int ed_hv_tune(ed_sector* s, const double* w, int precision, const double* d_x, double* d_y, int* tiles, int* gather,
               double* best_ms, void* stream) {
    return guarded([&] {
        ED_REQUIRE(s && s->D >= 0 && w && d_x && d_y, ED_ERR_ARG, "bad tune arguments");
        ED_REQUIRE(precision == 0 || precision == 1, ED_ERR_ARG, "precision must be 0 or 1");
        hipStream_t st = static_cast<hipStream_t>(stream);
        const HvWeights hw = make_weights(w);
        double best = 1e300;
        int bt = 0, bg = 0;
        if (s->D > 0 && lds_layout(s).bytes <= LDS_CAP) {
            for (int t : TUNE_TILES) {
                for (int g : TUNE_GATHER) {
                    s->tune_tiles = t;
                    s->tune_gather = g;
                    const double ms =
                        time_ms(st, 3, [&] { launch_hv_lds(s, hw, d_x, d_y, s->d_partial, precision, st); });
                    if (ms < best) {
                        best = ms;
                        bt = t;
                        bg = g;
                    }
                }
            }
        }
        s->tune_tiles = bt;
        s->tune_gather = bg;
        if (tiles) *tiles = bt;
        if (gather) *gather = bg;
        if (best_ms) *best_ms = best < 1e300 ? best : 0.0;
    });
}

}  // extern "C"
