// Host-side sector object and error plumbing shared by the library's
// translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "ed_common.h"
#include "exactdiag_hip.h"

namespace ed {

// Vector element types: complex (double2, float2) or real (double, float).
// Real vectors serve sectors whose H is real (momenta kx = 0 or L/2): their
// values have no imaginary part, so a real start vector stays real.
template <class VT>
struct is_real_vec {
    static constexpr bool value = false;
};
template <>
struct is_real_vec<double> {
    static constexpr bool value = true;
};
template <>
struct is_real_vec<float> {
    static constexpr bool value = true;
};

// This thread's share of the sum of n (re, im) partials (16-byte aligned
// pairs), NT threads striding over them with four independent loads in
// flight; the order is fixed, so a tree over the threads gives reproducible
// sums.
template <int NT>
__device__ inline void sum_partials(const double* partial, int64_t n, double& re, double& im) {
    const double2* p = reinterpret_cast<const double2*>(partial);
    double r[4] = {0.0, 0.0, 0.0, 0.0}, i[4] = {0.0, 0.0, 0.0, 0.0};
    int64_t k = threadIdx.x;
    for (; k + 3 * NT < n; k += 4 * NT) {
        double2 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = p[k + u * NT];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            r[u] += v[u].x;
            i[u] += v[u].y;
        }
    }
    for (; k < n; k += NT) {
        const double2 v = p[k];
        r[0] += v.x;
        i[0] += v.y;
    }
    re = (r[0] + r[1]) + (r[2] + r[3]);
    im = (i[0] + i[1]) + (i[2] + i[3]);
}

// Colex rank of an ne-bit pattern with nd ones (the spin rank of a down
// pattern in electron order, holes removed) by 8-bit chunks: chunk c (bits
// 8c .. 8c + width[c] - 1) with kb ones below it adds
// tab[base[c] + (kb << width[c]) + chunk bits], the sum over its set bits
// (the j-th of the chunk at bit i) of C(8c + i, kb + j + 1); kb runs to
// min(8c, nd).  entries = size of the table.
struct RankGeom {
    int n = 0;
    int base[4] = {0, 0, 0, 0};
    int width[4] = {0, 0, 0, 0};
    int entries = 0;
};

inline RankGeom rank_geom(int ne, int nd) {
    RankGeom g;
    g.n = (ne + 7) / 8;
    for (int c = 0; c < g.n; ++c) {
        g.width[c] = ne - 8 * c < 8 ? ne - 8 * c : 8;
        g.base[c] = g.entries;
        const int kmax = 8 * c < nd ? 8 * c : nd;
        g.entries += (kmax + 1) << g.width[c];
    }
    return g;
}

void set_error(const std::string& msg);

struct Fail {
    int status;
    std::string msg;
};

#define ED_HIP(call)                                                                                    \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            throw ::ed::Fail{ED_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)};            \
    } while (0)

#define ED_REQUIRE(cond, status, msg)              \
    do {                                           \
        if (!(cond)) throw ::ed::Fail{status, msg}; \
    } while (0)

template <class F>
int guarded(F&& f) {
    try {
        f();
        return ED_OK;
    } catch (const Fail& e) {
        set_error(e.msg);
        return e.status;
    } catch (const std::exception& e) {
        set_error(e.what());
        return ED_ERR_ARG;
    }
}

// Device memory of the library's buffers.  Freeing a block of tens of MB
// with hipFree costs about 0.25 ms whatever its size, and one spectrum of a
// large sector frees a dozen such blocks (basis arrays, stored-H arrays and
// their build scratch, Lanczos vectors); so blocks of at least
// DEV_CACHE_MIN bytes are kept after dev_free, per device and up to
// DEV_CACHE_CAP bytes, and a later request takes the smallest kept block of
// its device that is at least as large and at most 1/8 larger.  dev_free
// first waits for the device, as hipFree does, so no queued work still
// uses a block when it is handed out again.  When hipMalloc fails, the kept
// blocks of the device are released and the allocation is retried;
// ed_cache_release returns them on request.
constexpr size_t DEV_CACHE_MIN = size_t(64) << 10;
constexpr size_t DEV_CACHE_CAP = size_t(24) << 30;

struct DevCache {
    std::mutex mu;
    std::unordered_map<void*, std::pair<int, size_t>> live;  // block -> (device, bytes)
    std::multimap<size_t, std::pair<int, void*>> idle;       // bytes -> (device, block)
    size_t idle_bytes = 0;
};

inline DevCache& dev_cache() {
    static DevCache c;
    return c;
}

// Release the kept blocks of device `dev` (all devices when dev < 0).
inline void dev_cache_release(int dev) {
    DevCache& c = dev_cache();
    std::vector<void*> out;
    {
        std::lock_guard<std::mutex> g(c.mu);
        for (auto it = c.idle.begin(); it != c.idle.end();) {
            if (dev < 0 || it->second.first == dev) {
                out.push_back(it->second.second);
                c.idle_bytes -= it->first;
                it = c.idle.erase(it);
            } else {
                ++it;
            }
        }
    }
    for (void* p : out) (void)hipFree(p);
}

inline void* dev_alloc_bytes(size_t n) {
    n = n ? n : 1;
    void* p = nullptr;
    if (n < DEV_CACHE_MIN) {
        ED_HIP(hipMalloc(&p, n));
        return p;
    }
    const size_t grain = size_t(2) << 20;
    const size_t want = n < grain ? (n + 4095) / 4096 * 4096 : (n + grain - 1) / grain * grain;
    int dev = 0;
    ED_HIP(hipGetDevice(&dev));
    DevCache& c = dev_cache();
    {
        std::lock_guard<std::mutex> g(c.mu);
        for (auto it = c.idle.lower_bound(want); it != c.idle.end() && it->first <= want + want / 8; ++it) {
            if (it->second.first != dev) continue;
            p = it->second.second;
            c.live.emplace(p, std::make_pair(dev, it->first));
            c.idle_bytes -= it->first;
            c.idle.erase(it);
            return p;
        }
    }
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        dev_cache_release(dev);
        e = hipMalloc(&p, want);
    }
    ED_HIP(e);
    std::lock_guard<std::mutex> g(c.mu);
    c.live.emplace(p, std::make_pair(dev, want));
    return p;
}

template <class T>
T* dev_alloc(size_t n) {
    return static_cast<T*>(dev_alloc_bytes((n ? n : 1) * sizeof(T)));
}

template <class T>
T* dev_upload(const std::vector<T>& v, hipStream_t st) {
    T* p = dev_alloc<T>(v.size());
    if (!v.empty()) ED_HIP(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
    return p;
}

// Sum of (re, im) over a workgroup of NT threads (a multiple of 64) in a
// fixed order: each wavefront adds its 64 values by lane shuffles (xor
// butterfly; lane 0's association order is fixed), then the NT / 64
// wavefront sums are added in wavefront order from LDS.  Every thread
// returns the total; the trailing barrier lets a kernel call it again.
template <int NT>
__device__ inline void workgroup_sum2(double& re, double& im) {
    static_assert(NT % 64 == 0, "whole wavefronts");
    for (int o = 32; o > 0; o >>= 1) {
        re += __shfl_xor(re, o, 64);
        im += __shfl_xor(im, o, 64);
    }
    __shared__ double wsum[2 * (NT / 64)];
    const int w = int(threadIdx.x) >> 6;
    if ((threadIdx.x & 63) == 0) {
        wsum[2 * w] = re;
        wsum[2 * w + 1] = im;
    }
    __syncthreads();
    re = wsum[0];
    im = wsum[1];
    for (int k = 1; k < NT / 64; ++k) {
        re += wsum[2 * k];
        im += wsum[2 * k + 1];
    }
    __syncthreads();
}

inline void dev_free(void* p) {
    if (!p) return;
    DevCache& c = dev_cache();
    int dev = -1;
    (void)hipGetDevice(&dev);
    std::unique_lock<std::mutex> g(c.mu);
    const auto it = c.live.find(p);
    if (it == c.live.end()) {  // a small block, not kept
        g.unlock();
        (void)hipFree(p);
        return;
    }
    const int owner = it->second.first;
    const size_t bytes = it->second.second;
    c.live.erase(it);
    const bool keep = owner == dev && c.idle_bytes + bytes <= DEV_CACHE_CAP;
    g.unlock();
    if (!keep || hipDeviceSynchronize() != hipSuccess) {
        (void)hipFree(p);
        return;
    }
    g.lock();
    c.idle.emplace(bytes, std::make_pair(owner, p));
    c.idle_bytes += bytes;
}

// Combination list of one rule: reversed site pairs (or single sites) with
// (diagonal, off-diagonal) weights, as combinations_from_start_end_points
// returns them (general/matrix_rules_utils.pyx:15-87).
constexpr int RULE_W = 4;  // sites per combination slot (arity <= 4)
struct RuleHost {
    int arity = 2;
    std::vector<int8_t> idx;     // n_combo * RULE_W (unused slots 0)
    std::vector<double> w;       // n_combo * 4: diag re, im, off re, im
    int valid = 0, make = 0, spin_proj = 0, density = 0, fermion = 0;
    int singlet = 0;             // get_weights_annihilation_part_singlet_singlet
    int max_per_col = 1;
    int n_combo() const { return int(idx.size() / RULE_W); }
};

enum RuleValid {
    RV_HOP = 0, RV_SWAP = 1, RV_NO_HOLES = 2, RV_ALL_UP = 3, RV_ALL_DOWN = 4, RV_ALL_HOLES = 5,
    RV_UNEQUAL_SPINS_2HOLES = 6
};
enum RuleMake { RM_SWAP = 0, RM_SAME = 1, RM_ADD_DOWN = 2, RM_ADD_UP = 3, RM_ADD_HOLE = 4, RM_REMOVE2_ADD_UP_DOWN = 5 };

RuleHost hamiltonian_rule(int L, int term);
RuleHost operator_rule(int L, int op, int qx, int qy);
// Position-correlation operators (kind ED_POS_HOLES / ED_POS_SZ / ED_POS_SINGLET);
// shifts: (leg, rung) pairs, the fixed distances followed by the free shift.
RuleHost position_rule(int L, int kind, const int* shifts, int n_shifts);
int shifted_site(int N, int i, int leg, int rung);

}  // namespace ed

struct ed_sector {
    int L = 0, N = 0, h = 0, ne = 0, nd = 0, kx = 0, ky = 0;
    int any_k = 0;  // union of the momentum sectors: see ed_sector_create_union
    int flip = 0;   // spin-flip parity (+1 / -1), 0 = without the spin flip: see ed_sector_create_flip
    int64_t n_spin = 0, n_hole = 0;
    std::vector<uint32_t> comb;        // COMB_W * COMB_W
    std::vector<double> amp;           // see SectorView::amp
    std::vector<double> sqrt_tab;      // NO + 1 (SectorView::NO)
    std::vector<uint32_t> rank_tab;    // spin ranks by chunks (ed::RankGeom of ne, nd)
    std::vector<uint32_t> hole_by_rank;
    std::vector<uint32_t> canon_gmask;
    std::vector<int32_t> canon_block, block_of_rank;
    std::vector<int64_t> block_rank, block_base, block_table;
    std::vector<uint32_t> block_hole;
    std::vector<uint8_t> block_sym;
    int64_t D = -1;
    int64_t table_size = 0;
    int tune_tiles = 0;   // H·v row tiles per workgroup (0 = default heuristic)
    int tune_gather = 0;  // H·v gathers in flight per lane (0 = 4)

    uint32_t* d_comb = nullptr;
    double* d_amp = nullptr;
    double* d_sqrt = nullptr;
    uint32_t* d_rank_tab = nullptr;
    uint32_t* d_canon_gmask = nullptr;
    int32_t* d_canon_block = nullptr;
    int32_t* d_block_of_rank = nullptr;
    int64_t* d_block_base = nullptr;
    int64_t* d_block_table = nullptr;
    uint32_t* d_block_hole = nullptr;
    int32_t* d_table = nullptr;
    uint8_t* d_table_orbit = nullptr;  // orbit size of each table entry's state (0: none)
    uint64_t* d_key = nullptr;
    uint8_t* d_orbit = nullptr;
    double* d_partial = nullptr;  // per-workgroup partials of the fused H·v dot (calls on one sector serialise on it)
    // 64-row wavefronts (row f0 / 64) that the wavefront-uniform H·v leaves to the per-lane walk (hv_fix_list)
    int32_t* d_fix = nullptr;
    int64_t n_fix = 0;

    ed::SectorView host_view() const;  // host pointers (tables only)
    ed::SectorView view() const;       // device pointers
    void release();
};
