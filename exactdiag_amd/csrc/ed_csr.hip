// The reference's stored matrices in HBM, and their product.
//
// calc_sparse_matrix (general/cython_sparse_matrices.pyx:185-241) computes a
// matrix column by column, keeps a table of unique values (an element joins
// the first table entry within 1e-6 of it, else is appended; pyx:88-99) and
// writes values complex128[U], value_inds int16[nnz], minor_inds int64[nnz]
// and major_pointers int64[major + 1].  Column-major matrices keep the
// column arrays; the Hermitian row-major layout of the Hamiltonian terms
// also keeps them (column i of the lower triangle is row i of the upper one)
// and conjugates the value table (pyx:131-145).  ed_csr_from_sector /
// ed_csr_from_operator build exactly those arrays on the GPU:
//   1. the exact per-column elements (launch_columns), in chunks of columns
//      so that the padded scratch stays bounded;
//   2. an exclusive scan of the column counts -> major pointers;
//   3. compaction into minor indices and element values;
//   4. unique values: a hash set over the elements' exact bit patterns keeps
//      where each pattern first appears; the host walks the distinct patterns
//      (a few hundred) in order of first appearance with the reference's
//      tolerance rule.  This gives the reference's indices: a pattern's index
//      is fixed when it first appears, and every later element with that
//      pattern searches the same table prefix and finds the same entry;
//   5. value indices per element from its pattern (checked bit for bit
//      against the pattern's first element).
// ed_csr_create uploads arrays read from a folder the reference wrote.
//
// ed_csr_sum_* applies sum_T w_T M_T as Added_Sparse_Matrices.dot over
// sparse_matrix_dot does (general/sparse_matrices.py:83-105, :233-237;
// general/cython_sparse_dot.pyx:40-100; general/cpp_sparse_dot.h:59-138):
// explicit elements with w v, Hermitian-pass elements (stored off-diagonal
// elements applied transposed) with w conj(v) -- the weight is not
// conjugated -- and terms with |w| < 1e-6 skipped.  The sum is gathered per
// output row: each output row lists its elements from every term and both
// passes sorted by (column, value code), one thread sums a row in that fixed
// order, so the product is deterministic and needs no atomics.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "ed_internal.h"
#include "ed_matrix.h"

using namespace ed;

struct ed_csr {
    int64_t rows = 0, cols = 0, nnz = 0;
    int row_major = 1, herm_tri = 0;
    int n_values = 0;
    int max_count = 0;                // longest major slice
    std::vector<double> h_values;     // 2 * n_values (re, im)
    int64_t* d_major = nullptr;       // n_major + 1
    int64_t* d_minor = nullptr;       // nnz
    int16_t* d_vinds = nullptr;       // nnz
    double* d_values = nullptr;       // 2 * n_values
    int64_t n_major() const { return row_major ? rows : cols; }
    int64_t n_minor() const { return row_major ? cols : rows; }
    void release() {
        dev_free(d_major);
        dev_free(d_minor);
        dev_free(d_vinds);
        dev_free(d_values);
        d_major = d_minor = nullptr;
        d_vinds = nullptr;
        d_values = nullptr;
    }
};

// Sliced layout of the summed operator: rows in slices of 64 (one
// wavefront), element k of row 64 s + l at sptr[s] + 64 k + l as (column,
// code), code = 2 * value + (Hermitian pass), PAD_CODE in the padding.
struct ed_csr_sum {
    int64_t rows = 0, cols = 0, nnz = 0, slots = 0;
    int n_table = 0;
    int64_t* d_sptr = nullptr;    // slices + 1
    int2* d_elem = nullptr;       // slots
    double* d_table = nullptr;    // 2 * n_table complex: (w v, w conj(v)) per value
    void release() {
        dev_free(d_sptr);
        dev_free(d_elem);
        dev_free(d_table);
        d_sptr = nullptr;
        d_elem = nullptr;
        d_table = nullptr;
    }
};

namespace {

constexpr int MAX_UNIQUE = 32767;  // MAX_NUM_UNIQUIE_VALUES = iinfo(int16).max (types_and_constants.pyx:24)
constexpr double UNIQUE_TOL = 1e-6;

// ---------------------------------------------------------------------------
// Exclusive scan of int32 counts into int64 offsets (n + 1 entries).

constexpr int SCAN_T = 1024;
constexpr int SCAN_PER = 4;
constexpr int64_t SCAN_TILE = int64_t(SCAN_T) * SCAN_PER;

__device__ int64_t wg_exclusive_scan(int64_t v, int64_t* sh, int64_t& total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_T; off <<= 1) {
        const int64_t add = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    total = sh[SCAN_T - 1];
    const int64_t ex = sh[t] - v;
    __syncthreads();
    return ex;
}

__global__ void __launch_bounds__(SCAN_T) scan_tiles_kernel(const int32_t* cnt, int64_t n, int64_t* tile_sum) {
    __shared__ int64_t sh[SCAN_T];
    const int64_t base = int64_t(blockIdx.x) * SCAN_TILE + int64_t(threadIdx.x) * SCAN_PER;
    int64_t s = 0;
    for (int k = 0; k < SCAN_PER; ++k)
        if (base + k < n) s += cnt[base + k];
    int64_t total;
    (void)wg_exclusive_scan(s, sh, total);
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

// One workgroup: tile sums -> tile offsets in place; out_total = grand total.
__global__ void __launch_bounds__(SCAN_T) scan_sums_kernel(int64_t* tile_sum, int64_t nt, int64_t* out_total) {
    __shared__ int64_t sh[SCAN_T];
    int64_t carry = 0;
    for (int64_t c = 0; c < nt; c += SCAN_T) {
        const int64_t i = c + threadIdx.x;
        const int64_t v = i < nt ? tile_sum[i] : 0;
        int64_t total;
        const int64_t ex = wg_exclusive_scan(v, sh, total);
        if (i < nt) tile_sum[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) *out_total = carry;
}

__global__ void __launch_bounds__(SCAN_T) scan_write_kernel(const int32_t* cnt, int64_t n, const int64_t* tile_off,
                                                            int64_t* out) {
    __shared__ int64_t sh[SCAN_T];
    const int64_t base = int64_t(blockIdx.x) * SCAN_TILE + int64_t(threadIdx.x) * SCAN_PER;
    int64_t v[SCAN_PER];
    int64_t s = 0;
    for (int k = 0; k < SCAN_PER; ++k) {
        v[k] = base + k < n ? cnt[base + k] : 0;
        s += v[k];
    }
    int64_t total;
    int64_t ex = wg_exclusive_scan(s, sh, total) + tile_off[blockIdx.x];
    for (int k = 0; k < SCAN_PER; ++k) {
        if (base + k < n) out[base + k] = ex;
        ex += v[k];
    }
}

// out[0..n) = exclusive prefix sums of cnt, out[n] = total (returned).
int64_t scan_counts(const int32_t* d_cnt, int64_t n, int64_t* d_out, hipStream_t st) {
    const int64_t nt = (n + SCAN_TILE - 1) / SCAN_TILE;
    ED_REQUIRE(nt < (int64_t(1) << 31), ED_ERR_ARG, "too many rows for the scan");
    int64_t* tiles = dev_alloc<int64_t>(size_t(std::max<int64_t>(nt, 1)));
    try {
        if (nt > 0) hipLaunchKernelGGL(scan_tiles_kernel, dim3(unsigned(nt)), dim3(SCAN_T), 0, st, d_cnt, n, tiles);
        hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(SCAN_T), 0, st, tiles, nt, d_out + n);
        if (nt > 0)
            hipLaunchKernelGGL(scan_write_kernel, dim3(unsigned(nt)), dim3(SCAN_T), 0, st, d_cnt, n, tiles, d_out);
        ED_HIP(hipGetLastError());
        int64_t total = 0;
        ED_HIP(hipMemcpyAsync(&total, d_out + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        ED_HIP(hipStreamSynchronize(st));
        dev_free(tiles);
        return total;
    } catch (...) {
        dev_free(tiles);
        throw;
    }
}

__global__ void max_count_kernel(const int32_t* cnt, int64_t n, int* out) {
    __shared__ int sh[256];
    int m = 0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        m = max(m, cnt[i]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) sh[threadIdx.x] = max(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, sh[0]);
}

// ---------------------------------------------------------------------------
// Build kernels.

__global__ void compact_kernel(int64_t n, int cap, const int32_t* cnt, const int64_t* major, const int64_t* rows,
                               const double2* vals, int64_t* minor, double2* ev) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int c = cnt[t];
    const int64_t o = major[t];
    for (int k = 0; k < c; ++k) {
        minor[o + k] = rows[t * cap + k];
        ev[o + k] = vals[t * cap + k];
    }
}

__device__ inline uint64_t mix64(uint64_t x) {  // splitmix64 finaliser
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

__device__ inline uint64_t pattern_key(double2 v) {
    const uint64_t a = uint64_t(__double_as_longlong(v.x));
    const uint64_t b = uint64_t(__double_as_longlong(v.y));
    const uint64_t k = mix64(a ^ mix64(b + 0x9e3779b97f4a7c15ull));
    return k ? k : 1;  // 0 marks an empty slot
}

constexpr int HASH_LDS = 512;     // per-workgroup table of patterns (power of two)
constexpr int HASH_ROUNDS = 64;   // elements per lane: a workgroup takes 256 * HASH_ROUNDS consecutive elements

__device__ inline int global_insert(unsigned long long k, uint32_t mask, unsigned long long* key, int* err) {
    uint32_t h = uint32_t(k ^ (k >> 32)) & mask;
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        unsigned long long cur = key[h];
        if (cur == 0) cur = atomicCAS(&key[h], 0ull, k);
        if (cur == 0 || cur == k) return int(h);
        h = (h + 1) & mask;
    }
    atomicOr(err, 1);
    return -1;
}

// Insert every element's pattern into the global set: slot[e] = its slot,
// first[slot] = the smallest element index with that pattern.  Each
// workgroup first collects its patterns in LDS (with their smallest index),
// so the global set sees one insertion per pattern and workgroup; patterns
// that do not fit the LDS table go to the global set directly.
__global__ void __launch_bounds__(256) hash_kernel(int64_t nnz, const double2* ev, uint32_t mask,
                                                   unsigned long long* key, unsigned long long* first,
                                                   int32_t* slot, int* err) {
    __shared__ unsigned long long lkey[HASH_LDS];
    __shared__ unsigned long long lfirst[HASH_LDS];
    __shared__ int lglob[HASH_LDS];
    for (int i = threadIdx.x; i < HASH_LDS; i += 256) {
        lkey[i] = 0;
        lfirst[i] = ~0ull;
        lglob[i] = -1;
    }
    __syncthreads();
    const int64_t chunk0 = int64_t(blockIdx.x) * 256 * HASH_ROUNDS;
    for (int round = 0; round < HASH_ROUNDS; ++round) {
        const int64_t e = chunk0 + int64_t(round) * 256 + threadIdx.x;
        int ls = -1;
        unsigned long long k = 0;
        if (e < nnz) {
            k = pattern_key(ev[e]);
            uint32_t h = uint32_t(k ^ (k >> 32)) & (HASH_LDS - 1);
            for (int probe = 0; probe < HASH_LDS; ++probe) {
                const unsigned long long cur = atomicCAS(&lkey[h], 0ull, k);
                if (cur == 0 || cur == k) {
                    ls = int(h);
                    break;
                }
                h = (h + 1) & (HASH_LDS - 1);
            }
            if (ls >= 0) atomicMin(&lfirst[ls], (unsigned long long)e);
        }
        __syncthreads();
        // one lane per new LDS entry inserts its pattern into the global set
        if (ls >= 0 && lglob[ls] == -1 && atomicCAS(&lglob[ls], -1, -2) == -1)
            lglob[ls] = global_insert(k, mask, key, err);
        __syncthreads();
        if (e < nnz) {
            if (ls >= 0) {
                slot[e] = lglob[ls];
            } else {
                const int g = global_insert(k, mask, key, err);
                slot[e] = g;
                if (g >= 0) atomicMin(&first[g], (unsigned long long)e);
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < HASH_LDS; i += 256)
        if (lkey[i] != 0 && lglob[i] >= 0) atomicMin(&first[lglob[i]], lfirst[i]);
}

__global__ void list_kernel(uint32_t cap, const unsigned long long* key, const unsigned long long* first,
                            const double2* ev, int* counter, int32_t* slot_idx, int64_t* list_first,
                            double2* list_val) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= cap || key[s] == 0) return;
    const int i = atomicAdd(counter, 1);
    slot_idx[s] = i;
    list_first[i] = int64_t(first[s]);
    list_val[i] = ev[first[s]];
}

__global__ void assign_kernel(int64_t nnz, const int32_t* slot, const int32_t* slot_idx, const int16_t* vi_of_list,
                              const unsigned long long* first, const double2* ev, int16_t* vinds, int* err) {
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < nnz; e += int64_t(gridDim.x) * blockDim.x) {
        const int32_t s = slot[e];
        vinds[e] = vi_of_list[slot_idx[s]];
        const double2 a = ev[e], b = ev[first[s]];
        if (__double_as_longlong(a.x) != __double_as_longlong(b.x) ||
            __double_as_longlong(a.y) != __double_as_longlong(b.y))
            atomicOr(err, 4);  // two patterns share a hash key
    }
}

unsigned grid_for(int64_t n, int T, int64_t cap_blocks = 8192) {
    return unsigned(std::max<int64_t>(1, std::min<int64_t>((n + T - 1) / T, cap_blocks)));
}

// The stored CSR of the elements of `rule` from sector ini to sector fin:
// columns of ini, rows of fin (upper part only with `triangle`), kept as
// column arrays; row_major / herm_tri say how they are read back and
// conj_values conjugates the value table (the Hermitian row-major layout).
ed_csr* build_csr(const ed_sector* ini, const ed_sector* fin, const RuleHost& rule, int triangle, int row_major,
                  int herm_tri, int conj_values, hipStream_t st) {
    ED_REQUIRE(ini && fin && ini->D >= 0 && fin->D >= 0, ED_ERR_STATE, "basis not built");
    const int64_t D = ini->D;  // columns computed
    const int cap = std::min(rule.max_per_col, column_capacity());
    auto* m = new ed_csr;
    m->rows = fin->D;
    m->cols = ini->D;
    m->row_major = row_major;
    m->herm_tri = herm_tri;
    int32_t* cnt = nullptr;
    int64_t* rows = nullptr;
    double* vals = nullptr;
    double2* ev = nullptr;
    unsigned long long *key = nullptr, *first = nullptr;
    int32_t *slot = nullptr, *slot_idx = nullptr;
    int64_t* list_first = nullptr;
    double2* list_val = nullptr;
    int16_t* vi_dev = nullptr;
    int* flags = nullptr;  // [0] max count, [1] error bits, [2] list counter
    auto cleanup = [&] {
        dev_free(cnt);
        dev_free(rows);
        dev_free(vals);
        dev_free(ev);
        dev_free(key);
        dev_free(first);
        dev_free(slot);
        dev_free(slot_idx);
        dev_free(list_first);
        dev_free(list_val);
        dev_free(vi_dev);
        dev_free(flags);
    };
    try {
        // 1. counts (and elements, when one chunk holds every column)
        const int64_t chunk_cap = std::max<int64_t>(1, (int64_t(1) << 30) / (int64_t(cap) * 24));
        const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(D, chunk_cap));
        cnt = dev_alloc<int32_t>(size_t(std::max<int64_t>(D, 1)));
        rows = dev_alloc<int64_t>(size_t(chunk * cap));
        vals = dev_alloc<double>(size_t(2 * chunk * cap));
        flags = dev_alloc<int>(3);
        ED_HIP(hipMemsetAsync(flags, 0, 3 * sizeof(int), st));
        for (int64_t lo = 0; lo < D; lo += chunk) {
            const int64_t n = std::min(chunk, D - lo);
            launch_columns(ini, fin, rule, triangle, cap, lo, n, cnt + lo, rows, vals, st);
        }
        if (D > 0) hipLaunchKernelGGL(max_count_kernel, dim3(grid_for(D, 256, 1024)), dim3(256), 0, st, cnt, D, flags);
        int hf[3] = {0, 0, 0};
        ED_HIP(hipMemcpyAsync(hf, flags, 3 * sizeof(int), hipMemcpyDeviceToHost, st));
        ED_HIP(hipStreamSynchronize(st));
        ED_REQUIRE(hf[0] <= cap, ED_ERR_CAPACITY,
                   "a column has " + std::to_string(hf[0]) + " elements, more than the capacity " +
                       std::to_string(cap));
        m->max_count = hf[0];

        // 2. major pointers
        m->d_major = dev_alloc<int64_t>(size_t(D + 1));
        const int64_t nnz = scan_counts(cnt, D, m->d_major, st);

        // 3. compaction (the columns are computed again when they took several chunks)
        const int64_t nalloc = std::max<int64_t>(nnz, 1);
        m->d_minor = dev_alloc<int64_t>(size_t(nalloc));
        ev = dev_alloc<double2>(size_t(nalloc));
        for (int64_t lo = 0; lo < D; lo += chunk) {
            const int64_t n = std::min(chunk, D - lo);
            if (chunk < D) launch_columns(ini, fin, rule, triangle, cap, lo, n, cnt + lo, rows, vals, st);
            hipLaunchKernelGGL(compact_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, n, cap, cnt + lo,
                               m->d_major + lo, rows, reinterpret_cast<const double2*>(vals), m->d_minor, ev);
            ED_HIP(hipGetLastError());
        }
        dev_free(rows);
        rows = nullptr;
        dev_free(vals);
        vals = nullptr;
        dev_free(cnt);
        cnt = nullptr;

        m->d_vinds = dev_alloc<int16_t>(size_t(nalloc));
        std::vector<double> uniq;  // (re, im) pairs, in table order
        if (nnz == 0) {
            // an empty matrix is written as one explicit zero in the last major slice (pyx:123-129)
            const int64_t zero = 0, one = 1;
            const int16_t z16 = 0;
            ED_HIP(hipMemcpyAsync(m->d_minor, &zero, sizeof(int64_t), hipMemcpyHostToDevice, st));
            ED_HIP(hipMemcpyAsync(m->d_vinds, &z16, sizeof(int16_t), hipMemcpyHostToDevice, st));
            ED_HIP(hipMemcpyAsync(m->d_major + D, &one, sizeof(int64_t), hipMemcpyHostToDevice, st));
            ED_HIP(hipStreamSynchronize(st));
            uniq = {0.0, 0.0};
            m->nnz = 1;
        } else {
            // 4. distinct patterns and their first elements
            uint32_t hcap = 1024;
            while (int64_t(hcap) < 2 * std::min<int64_t>(nnz, 1 << 19)) hcap <<= 1;
            key = dev_alloc<unsigned long long>(hcap);
            first = dev_alloc<unsigned long long>(hcap);
            slot = dev_alloc<int32_t>(size_t(nnz));
            slot_idx = dev_alloc<int32_t>(hcap);
            list_first = dev_alloc<int64_t>(hcap);
            list_val = dev_alloc<double2>(hcap);
            ED_HIP(hipMemsetAsync(key, 0, hcap * sizeof(unsigned long long), st));
            ED_HIP(hipMemsetAsync(first, 0xff, hcap * sizeof(unsigned long long), st));
            const int64_t hblocks = (nnz + 256 * HASH_ROUNDS - 1) / (256 * HASH_ROUNDS);
            ED_REQUIRE(hblocks < (int64_t(1) << 31), ED_ERR_ARG, "too many elements");
            hipLaunchKernelGGL(hash_kernel, dim3(unsigned(hblocks)), dim3(256), 0, st, nnz, ev, hcap - 1, key, first,
                               slot, flags + 1);
            hipLaunchKernelGGL(list_kernel, dim3((hcap + 255) / 256), dim3(256), 0, st, hcap, key, first, ev,
                               flags + 2, slot_idx, list_first, list_val);
            ED_HIP(hipGetLastError());
            ED_HIP(hipMemcpyAsync(hf, flags, 3 * sizeof(int), hipMemcpyDeviceToHost, st));
            ED_HIP(hipStreamSynchronize(st));
            ED_REQUIRE(!(hf[1] & 1), ED_ERR_CAPACITY, "too many distinct element values");
            const int nd = hf[2];
            std::vector<int64_t> lf(static_cast<size_t>(nd));
            std::vector<double> lv(size_t(2 * nd));
            ED_HIP(hipMemcpyAsync(lf.data(), list_first, size_t(nd) * sizeof(int64_t), hipMemcpyDeviceToHost, st));
            ED_HIP(hipMemcpyAsync(lv.data(), list_val, size_t(nd) * sizeof(double2), hipMemcpyDeviceToHost, st));
            ED_HIP(hipStreamSynchronize(st));
            // the reference's walk: in order of first appearance, join the first entry within the tolerance
            std::vector<int> order(static_cast<size_t>(nd));
            std::iota(order.begin(), order.end(), 0);
            std::sort(order.begin(), order.end(), [&](int a, int b) { return lf[a] < lf[b]; });
            std::vector<int16_t> vi(static_cast<size_t>(nd));
            for (int i : order) {
                const std::complex<double> v(lv[2 * i], lv[2 * i + 1]);
                int found = -1;
                for (size_t u = 0; u < uniq.size() / 2; ++u)
                    if (std::abs(v - std::complex<double>(uniq[2 * u], uniq[2 * u + 1])) < UNIQUE_TOL) {
                        found = int(u);
                        break;
                    }
                if (found < 0) {
                    ED_REQUIRE(int(uniq.size() / 2) < MAX_UNIQUE, ED_ERR_CAPACITY, "too many unique values");
                    found = int(uniq.size() / 2);
                    uniq.push_back(v.real());
                    uniq.push_back(v.imag());
                }
                vi[size_t(i)] = int16_t(found);
            }
            // 5. value index per element
            vi_dev = dev_upload(vi, st);
            hipLaunchKernelGGL(assign_kernel, dim3(grid_for(nnz, 256)), dim3(256), 0, st, nnz, slot, slot_idx, vi_dev,
                               first, ev, m->d_vinds, flags + 1);
            ED_HIP(hipGetLastError());
            ED_HIP(hipMemcpyAsync(hf, flags, 3 * sizeof(int), hipMemcpyDeviceToHost, st));
            ED_HIP(hipStreamSynchronize(st));
            ED_REQUIRE(!(hf[1] & 4), ED_ERR_STATE, "element value hash collision");
            m->nnz = nnz;
        }
        if (conj_values)
            for (size_t u = 1; u < uniq.size(); u += 2) uniq[u] = -uniq[u];
        m->n_values = int(uniq.size() / 2);
        m->h_values = uniq;
        m->d_values = dev_upload(uniq, st);
        ED_HIP(hipStreamSynchronize(st));
        cleanup();
        return m;
    } catch (...) {
        cleanup();
        m->release();
        delete m;
        throw;
    }
}

// ---------------------------------------------------------------------------
// Weighted sum of stored matrices, gathered per output row.

// Element (major a, minor b) of a stored matrix: the output rows it feeds.
//   row-major:    explicit -> row a (column b); Hermitian pass, b != a -> row b (column a)
//   column-major: explicit -> row b (column a); Hermitian pass, b != a -> row a (column b)
__global__ void sum_count_kernel(int64_t n_major, const int64_t* major, const int64_t* minor, int row_major, int herm,
                                 int32_t* rowcnt) {
    const int64_t a = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (a >= n_major) return;
    int own = 0;
    for (int64_t e = major[a]; e < major[a + 1]; ++e) {
        const int64_t b = minor[e];
        if (row_major) {
            ++own;
            if (herm && b != a) atomicAdd(&rowcnt[b], 1);
        } else {
            atomicAdd(&rowcnt[b], 1);
            if (herm && b != a) ++own;
        }
    }
    if (own) atomicAdd(&rowcnt[a], own);
}

__global__ void sum_fill_kernel(int64_t n_major, const int64_t* major, const int64_t* minor, const int16_t* vinds,
                                int row_major, int herm, uint32_t table_off, unsigned long long* cursor, int32_t* col,
                                uint32_t* code) {
    const int64_t a = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (a >= n_major) return;
    for (int64_t e = major[a]; e < major[a + 1]; ++e) {
        const int64_t b = minor[e];
        const uint32_t u = 2u * (table_off + uint32_t(uint16_t(vinds[e])));
        const int64_t r0 = row_major ? a : b, c0 = row_major ? b : a;
        const unsigned long long p0 = atomicAdd(&cursor[r0], 1ull);
        col[p0] = int32_t(c0);
        code[p0] = u;
        if (herm && b != a) {
            const unsigned long long p1 = atomicAdd(&cursor[c0], 1ull);
            col[p1] = int32_t(r0);
            code[p1] = u + 1u;
        }
    }
}

// Each output row's elements in (column, code) order: one wavefront per row,
// a bitonic sort of the 64-bit keys (column << 32 | code) across the lanes;
// rows longer than a wavefront are sorted by one lane in place.
__global__ void __launch_bounds__(256) sum_sort_kernel(int64_t rows, const int64_t* ptr, int32_t* col,
                                                       uint32_t* code) {
    const int64_t r = (int64_t(blockIdx.x) * 256 + threadIdx.x) / 64;
    const int lane = threadIdx.x & 63;
    if (r >= rows) return;  // uniform across the wavefront
    const int64_t s = ptr[r], n = ptr[r + 1] - s;
    if (n <= 1) return;
    if (n > 64) {
        if (lane != 0) return;
        for (int64_t k = s + 1; k < s + n; ++k) {
            const int32_t ck = col[k];
            const uint32_t dk = code[k];
            int64_t q = k - 1;
            while (q >= s && (col[q] > ck || (col[q] == ck && code[q] > dk))) {
                col[q + 1] = col[q];
                code[q + 1] = code[q];
                --q;
            }
            col[q + 1] = ck;
            code[q + 1] = dk;
        }
        return;
    }
    unsigned long long key = lane < n ? (uint64_t(uint32_t(col[s + lane])) << 32) | code[s + lane] : ~0ull;
    for (int k = 2; k <= 64; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            const unsigned long long other = __shfl_xor(key, j, 64);
            const bool keep_min = ((lane & j) == 0) == ((lane & k) == 0);
            key = keep_min ? (other < key ? other : key) : (other > key ? other : key);
        }
    if (lane < n) {
        col[s + lane] = int32_t(key >> 32);
        code[s + lane] = uint32_t(key);
    }
}

constexpr int SLICE = 64;
constexpr int PAD_CODE = -1;

// Slice widths (longest row of each slice) times SLICE: the slice sizes.
__global__ void slice_size_kernel(int64_t rows, int64_t n_slices, const int64_t* ptr, int32_t* size) {
    const int64_t sl = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (sl >= n_slices) return;
    int64_t w = 0;
    const int64_t r1 = min(rows, (sl + 1) * SLICE);
    for (int64_t r = sl * SLICE; r < r1; ++r) w = max(w, ptr[r + 1] - ptr[r]);
    size[sl] = int32_t(w * SLICE);
}

__global__ void slice_fill_kernel(int64_t rows, int64_t n_slices, const int64_t* ptr, const int32_t* col,
                                  const uint32_t* code, const int64_t* sptr, int2* elem) {
    const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= n_slices * SLICE) return;
    const int64_t sl = r / SLICE, lane = r % SLICE;
    const int64_t base = sptr[sl], width = (sptr[sl + 1] - base) / SLICE;
    const int64_t e0 = r < rows ? ptr[r] : 0;
    const int64_t len = r < rows ? ptr[r + 1] - e0 : 0;
    for (int64_t k = 0; k < width; ++k)
        elem[base + k * SLICE + lane] = k < len ? make_int2(col[e0 + k], int(code[e0 + k])) : make_int2(0, PAD_CODE);
}

template <bool LDS>
__global__ void __launch_bounds__(256) csr_sum_hv_kernel(int64_t rows, const int64_t* sptr, const int2* elem,
                                                         const double2* table, int n_entries, const double2* x,
                                                         double2* y) {
    extern __shared__ double2 tab[];
    if (LDS) {
        for (int i = threadIdx.x; i < n_entries; i += 256) tab[i] = table[i];
        __syncthreads();
    }
    // row tiles dealt to the XCDs in neighbouring chunks of 512 (the stored product's order)
    const int64_t r = int64_t(xcd_order(blockIdx.x, gridDim.x, 512)) * 256 + threadIdx.x;
    if (r >= rows) return;
    const int64_t base = sptr[r / SLICE];
    const int64_t width = (sptr[r / SLICE + 1] - base) / SLICE;
    const int2* p = elem + base + (r % SLICE);
    double re = 0.0, im = 0.0;
#pragma unroll 4
    for (int64_t k = 0; k < width; ++k) {
        const int2 e = p[k * SLICE];
        if (e.y != PAD_CODE) {
            const double2 t = LDS ? tab[e.y] : table[e.y];
            const double2 v = x[e.x];
            re += t.x * v.x - t.y * v.y;
            im += t.x * v.y + t.y * v.x;
        }
    }
    y[r] = make_double2(re, im);
}

}  // namespace

extern "C" {

int ed_csr_from_sector(const ed_sector* s, int term, int triangle, int row_major, ed_csr** out, void* stream) {
    return guarded([&] {
        ED_REQUIRE(s && out, ED_ERR_ARG, "null argument");
        require_no_flip(s, "the reference's matrix folders");
        ED_REQUIRE(term >= 0 && term < 4, ED_ERR_ARG, "term must be 0..3 (tl, tr, jl, jr)");
        *out = build_csr(s, s, hamiltonian_rule(s->L, term), triangle ? 1 : 0, row_major ? 1 : 0, triangle ? 1 : 0,
                         row_major ? 1 : 0, static_cast<hipStream_t>(stream));
    });
}

int ed_csr_from_operator(const ed_sector* init, const ed_sector* fin, int op, int qx, int qy, ed_csr** out,
                         void* stream) {
    return guarded([&] {
        ED_REQUIRE(init && fin && out, ED_ERR_ARG, "null argument");
        require_no_flip(init, "the reference's matrix folders");
        require_no_flip(fin, "the reference's matrix folders");
        ED_REQUIRE(init->L == fin->L, ED_ERR_ARG, "sectors of different ladders");
        *out = build_csr(init, fin, operator_rule(init->L, op, qx, qy), 0, 0, 0, 0, static_cast<hipStream_t>(stream));
    });
}

int ed_csr_from_position_operator(const ed_sector* s, int kind, const int* shifts, int n_shifts, ed_csr** out,
                                  void* stream) {
    return guarded([&] {
        ED_REQUIRE(s && out && (shifts || n_shifts == 0), ED_ERR_ARG, "null argument");
        require_no_flip(s, "the reference's matrix folders");
        *out = build_csr(s, s, position_rule(s->L, kind, shifts, n_shifts), 0, 0, 0, 0,
                         static_cast<hipStream_t>(stream));
    });
}

int ed_csr_create(int64_t rows, int64_t cols, int row_major, int hermitian_triangle, int64_t nnz, int n_values,
                  const double* values, const int16_t* value_inds, const int64_t* minor, const int64_t* major,
                  ed_csr** out, void* stream) {
    return guarded([&] {
        ED_REQUIRE(out, ED_ERR_ARG, "null argument");
        ED_REQUIRE(rows >= 0 && cols >= 0 && nnz >= 0 && n_values >= 0, ED_ERR_ARG, "negative size");
        ED_REQUIRE(n_values <= MAX_UNIQUE, ED_ERR_ARG, "more unique values than int16 indices reach");
        ED_REQUIRE(!hermitian_triangle || rows == cols, ED_ERR_ARG, "a Hermitian triangle must be square");
        ED_REQUIRE((nnz == 0 || (value_inds && minor)) && major && (n_values == 0 || values), ED_ERR_ARG,
                   "null array");
        hipStream_t st = static_cast<hipStream_t>(stream);
        auto* m = new ed_csr;
        m->rows = rows;
        m->cols = cols;
        m->row_major = row_major ? 1 : 0;
        m->herm_tri = hermitian_triangle ? 1 : 0;
        m->nnz = nnz;
        m->n_values = n_values;
        try {
            const int64_t nm = m->n_major(), nmin = m->n_minor();
            ED_REQUIRE(nmin <= INT32_MAX && nm <= INT32_MAX, ED_ERR_ARG, "dimension beyond int32");
            ED_REQUIRE(major[0] == 0 && major[nm] == nnz, ED_ERR_ARG, "major pointers do not span the elements");
            int64_t longest = 0;
            for (int64_t a = 0; a < nm; ++a) {
                ED_REQUIRE(major[a + 1] >= major[a], ED_ERR_ARG, "major pointers decrease");
                longest = std::max(longest, major[a + 1] - major[a]);
            }
            for (int64_t e = 0; e < nnz; ++e) {
                ED_REQUIRE(minor[e] >= 0 && minor[e] < nmin, ED_ERR_ARG, "minor index out of range");
                ED_REQUIRE(value_inds[e] >= 0 && value_inds[e] < n_values, ED_ERR_ARG, "value index out of range");
            }
            m->max_count = int(std::min<int64_t>(longest, INT32_MAX));
            m->h_values.assign(values, values + 2 * size_t(n_values));
            m->d_major = dev_alloc<int64_t>(size_t(nm + 1));
            m->d_minor = dev_alloc<int64_t>(size_t(std::max<int64_t>(nnz, 1)));
            m->d_vinds = dev_alloc<int16_t>(size_t(std::max<int64_t>(nnz, 1)));
            m->d_values = dev_alloc<double>(size_t(std::max(2 * n_values, 2)));
            ED_HIP(hipMemcpyAsync(m->d_major, major, size_t(nm + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
            if (nnz) {
                ED_HIP(hipMemcpyAsync(m->d_minor, minor, size_t(nnz) * sizeof(int64_t), hipMemcpyHostToDevice, st));
                ED_HIP(hipMemcpyAsync(m->d_vinds, value_inds, size_t(nnz) * sizeof(int16_t), hipMemcpyHostToDevice,
                                      st));
            }
            if (n_values)
                ED_HIP(hipMemcpyAsync(m->d_values, values, size_t(2 * n_values) * sizeof(double),
                                      hipMemcpyHostToDevice, st));
            ED_HIP(hipStreamSynchronize(st));
        } catch (...) {
            m->release();
            delete m;
            throw;
        }
        *out = m;
    });
}

int ed_csr_info(const ed_csr* m, int64_t* rows, int64_t* cols, int64_t* nnz, int* n_values, int* max_count) {
    return guarded([&] {
        ED_REQUIRE(m, ED_ERR_ARG, "null matrix");
        if (rows) *rows = m->rows;
        if (cols) *cols = m->cols;
        if (nnz) *nnz = m->nnz;
        if (n_values) *n_values = m->n_values;
        if (max_count) *max_count = m->max_count;
    });
}

int ed_csr_copy_out(const ed_csr* m, double* values, int16_t* value_inds, int64_t* minor, int64_t* major,
                    void* stream) {
    return guarded([&] {
        ED_REQUIRE(m, ED_ERR_ARG, "null matrix");
        hipStream_t st = static_cast<hipStream_t>(stream);
        if (values && m->n_values)
            ED_HIP(hipMemcpyAsync(values, m->d_values, size_t(2 * m->n_values) * sizeof(double),
                                  hipMemcpyDeviceToHost, st));
        if (value_inds && m->nnz)
            ED_HIP(hipMemcpyAsync(value_inds, m->d_vinds, size_t(m->nnz) * sizeof(int16_t), hipMemcpyDeviceToHost,
                                  st));
        if (minor && m->nnz)
            ED_HIP(hipMemcpyAsync(minor, m->d_minor, size_t(m->nnz) * sizeof(int64_t), hipMemcpyDeviceToHost, st));
        if (major)
            ED_HIP(hipMemcpyAsync(major, m->d_major, size_t(m->n_major() + 1) * sizeof(int64_t),
                                  hipMemcpyDeviceToHost, st));
        ED_HIP(hipStreamSynchronize(st));
    });
}

int ed_csr_destroy(ed_csr* m) {
    if (m) {
        m->release();
        delete m;
    }
    return ED_OK;
}

int ed_csr_sum_create(const ed_csr* const* terms, const double* weights, int n, ed_csr_sum** out, void* stream) {
    return guarded([&] {
        ED_REQUIRE(terms && weights && out && n >= 1, ED_ERR_ARG, "null argument");
        for (int t = 0; t < n; ++t) ED_REQUIRE(terms[t], ED_ERR_ARG, "null matrix");
        const int64_t rows = terms[0]->rows, cols = terms[0]->cols;
        for (int t = 1; t < n; ++t)
            ED_REQUIRE(terms[t]->rows == rows && terms[t]->cols == cols, ED_ERR_ARG, "incompatible shapes");
        ED_REQUIRE(rows <= INT32_MAX && cols <= INT32_MAX, ED_ERR_ARG, "dimension beyond int32");
        hipStream_t st = static_cast<hipStream_t>(stream);
        auto* m = new ed_csr_sum;
        m->rows = rows;
        m->cols = cols;
        int32_t* rowcnt = nullptr;
        unsigned long long* cursor = nullptr;
        int64_t* d_ptr = nullptr;
        int32_t* d_col = nullptr;
        uint32_t* d_code = nullptr;
        auto free_tmp = [&] {
            dev_free(rowcnt);
            dev_free(cursor);
            dev_free(d_ptr);
            dev_free(d_col);
            dev_free(d_code);
            rowcnt = nullptr;
            cursor = nullptr;
            d_ptr = nullptr;
            d_col = nullptr;
            d_code = nullptr;
        };
        try {
            // value table: (w v, w conj(v)) per stored value of every term kept
            std::vector<double> table;
            std::vector<int> keep;
            std::vector<uint32_t> off;
            for (int t = 0; t < n; ++t) {
                const std::complex<double> w(weights[2 * t], weights[2 * t + 1]);
                if (std::abs(w) < 1e-6) continue;  // general/sparse_matrices.py:96
                keep.push_back(t);
                off.push_back(uint32_t(table.size() / 4));
                const auto& hv = terms[t]->h_values;
                for (int u = 0; u < terms[t]->n_values; ++u) {
                    const std::complex<double> v(hv[2 * u], hv[2 * u + 1]);
                    const std::complex<double> a = w * v, b = w * std::conj(v);
                    table.insert(table.end(), {a.real(), a.imag(), b.real(), b.imag()});
                }
            }
            m->n_table = int(table.size() / 4);
            m->d_table = dev_upload(table.empty() ? std::vector<double>{0.0, 0.0} : table, st);
            rowcnt = dev_alloc<int32_t>(size_t(std::max<int64_t>(rows, 1)));
            ED_HIP(hipMemsetAsync(rowcnt, 0, size_t(std::max<int64_t>(rows, 1)) * sizeof(int32_t), st));
            for (int t : keep) {
                const ed_csr* c = terms[t];
                const int64_t nm = c->n_major();
                if (nm > 0)
                    hipLaunchKernelGGL(sum_count_kernel, dim3(unsigned((nm + 255) / 256)), dim3(256), 0, st, nm,
                                       c->d_major, c->d_minor, c->row_major, c->herm_tri, rowcnt);
            }
            ED_HIP(hipGetLastError());
            // per-row lists, sorted
            d_ptr = dev_alloc<int64_t>(size_t(rows + 1));
            m->nnz = scan_counts(rowcnt, rows, d_ptr, st);
            d_col = dev_alloc<int32_t>(size_t(std::max<int64_t>(m->nnz, 1)));
            d_code = dev_alloc<uint32_t>(size_t(std::max<int64_t>(m->nnz, 1)));
            cursor = dev_alloc<unsigned long long>(size_t(std::max<int64_t>(rows, 1)));
            if (rows > 0)
                ED_HIP(hipMemcpyAsync(cursor, d_ptr, size_t(rows) * sizeof(int64_t), hipMemcpyDeviceToDevice, st));
            for (size_t k = 0; k < keep.size(); ++k) {
                const ed_csr* c = terms[keep[k]];
                const int64_t nm = c->n_major();
                if (nm > 0)
                    hipLaunchKernelGGL(sum_fill_kernel, dim3(unsigned((nm + 255) / 256)), dim3(256), 0, st, nm,
                                       c->d_major, c->d_minor, c->d_vinds, c->row_major, c->herm_tri, off[k], cursor,
                                       d_col, d_code);
            }
            if (rows > 0)
                hipLaunchKernelGGL(sum_sort_kernel, dim3(unsigned((rows + 3) / 4)), dim3(256), 0, st, rows, d_ptr,
                                   d_col, d_code);
            ED_HIP(hipGetLastError());
            // sliced layout (rowcnt reused for the slice sizes)
            const int64_t n_slices = (rows + SLICE - 1) / SLICE;
            if (n_slices > 0)
                hipLaunchKernelGGL(slice_size_kernel, dim3(unsigned((n_slices + 255) / 256)), dim3(256), 0, st, rows,
                                   n_slices, d_ptr, rowcnt);
            m->d_sptr = dev_alloc<int64_t>(size_t(n_slices + 1));
            m->slots = scan_counts(rowcnt, n_slices, m->d_sptr, st);
            m->d_elem = dev_alloc<int2>(size_t(std::max<int64_t>(m->slots, 1)));
            if (n_slices > 0)
                hipLaunchKernelGGL(slice_fill_kernel, dim3(unsigned((n_slices * SLICE + 255) / 256)), dim3(256), 0,
                                   st, rows, n_slices, d_ptr, d_col, d_code, m->d_sptr, m->d_elem);
            ED_HIP(hipGetLastError());
            ED_HIP(hipStreamSynchronize(st));
            free_tmp();
        } catch (...) {
            free_tmp();
            m->release();
            delete m;
            throw;
        }
        *out = m;
    });
}

int ed_csr_sum_info(const ed_csr_sum* m, int64_t* rows, int64_t* cols, int64_t* nnz, int64_t* slots) {
    return guarded([&] {
        ED_REQUIRE(m, ED_ERR_ARG, "null operator");
        if (rows) *rows = m->rows;
        if (cols) *cols = m->cols;
        if (nnz) *nnz = m->nnz;
        if (slots) *slots = m->slots;
    });
}

int ed_csr_sum_hv(const ed_csr_sum* m, const double* d_x, double* d_y, void* stream) {
    return guarded([&] {
        ED_REQUIRE(m, ED_ERR_ARG, "null operator");
        if (m->rows == 0) return;  // empty vectors may come with null pointers
        ED_REQUIRE(d_y && (d_x || m->cols == 0), ED_ERR_ARG, "null vector");
        hipStream_t st = static_cast<hipStream_t>(stream);
        const unsigned grid = unsigned((m->rows + 255) / 256);
        const int entries = 2 * m->n_table;
        const auto* x = reinterpret_cast<const double2*>(d_x);
        auto* y = reinterpret_cast<double2*>(d_y);
        const auto* tab = reinterpret_cast<const double2*>(m->d_table);
        if (entries > 0 && size_t(entries) * sizeof(double2) <= 32768)
            hipLaunchKernelGGL(csr_sum_hv_kernel<true>, dim3(grid), dim3(256), size_t(entries) * sizeof(double2), st,
                               m->rows, m->d_sptr, m->d_elem, tab, entries, x, y);
        else
            hipLaunchKernelGGL(csr_sum_hv_kernel<false>, dim3(grid), dim3(256), 0, st, m->rows, m->d_sptr, m->d_elem,
                               tab, entries, x, y);
        ED_HIP(hipGetLastError());
    });
}

int ed_csr_sum_destroy(ed_csr_sum* m) {
    if (m) {
        m->release();
        delete m;
    }
    return ED_OK;
}

}  // extern "C"
