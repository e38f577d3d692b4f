// Learned continued-fraction closure: a small MLP trained and evaluated with
// hand-written BF16 MFMA kernels on gfx950.  No counterpart in the reference.
//
// Network: n_in -> [n_hidden, GELU] x n_layers -> n_out (linear).  FP32
// master weights and Adam state; every matrix product runs on
// v_mfma_f32_32x32x16_bf16 with BF16 operands and FP32 accumulation.
// Row-major layouts: activations [batch][features], weights W[k][n].
//
// MFMA 32x32x16 BF16 lane layout (64 lanes): lane l supplies
// A[m = l % 32][k = 8 (l / 32) + i] and B[k = 8 (l / 32) + i][n = l % 32] for
// i = 0..7 and receives D[m = 8 (j / 4) + 4 (l / 32) + j % 4][n = l % 32] in
// accumulator j = 0..15.

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <random>
#include <type_traits>
#include <vector>

#include "ed_internal.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace ed {
namespace {

__host__ __device__ inline uint16_t f32_to_bf16(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7f800000u) == 0x7f800000u) return uint16_t((u >> 16) | ((u & 0xffff) ? 0x40 : 0));  // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return uint16_t(u >> 16);
}

__host__ __device__ inline float bf16_to_f32(uint16_t h) {
    return __builtin_bit_cast(float, uint32_t(h) << 16);
}

// C[M][N] (fp32) = alpha * opA(A)[M][K] . opB(B)[K][N] (+ beta * C), BF16 operands
// addressed with element strides so transposed operands need no copies.
// One wavefront per 32x32 tile of C.
__global__ void __launch_bounds__(64) gemm_bf16_kernel(int M, int N, int K, const uint16_t* __restrict__ A,
                                                       int64_t sam, int64_t sak, const uint16_t* __restrict__ B,
                                                       int64_t sbk, int64_t sbn, float* __restrict__ C,
                                                       int64_t ldc, int accumulate) {
    const int lane = threadIdx.x;
    const int m0 = blockIdx.y * 32, n0 = blockIdx.x * 32;
    const int r = lane & 31, half = lane >> 5;
    f32x16 acc;
    for (int j = 0; j < 16; ++j) acc[j] = 0.0f;
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x8 a, b;
        for (int i = 0; i < 8; ++i) {
            const int k = k0 + 8 * half + i;
            const int m = m0 + r, n = n0 + r;
            const uint16_t av = (m < M && k < K) ? A[int64_t(m) * sam + int64_t(k) * sak] : uint16_t(0);
            const uint16_t bv = (n < N && k < K) ? B[int64_t(k) * sbk + int64_t(n) * sbn] : uint16_t(0);
            a[i] = __builtin_bit_cast(__bf16, av);
            b[i] = __builtin_bit_cast(__bf16, bv);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    for (int j = 0; j < 16; ++j) {
        const int m = m0 + 8 * (j >> 2) + 4 * half + (j & 3);
        const int n = n0 + r;
        if (m < M && n < N) {
            float* c = &C[int64_t(m) * ldc + n];
            *c = accumulate ? *c + acc[j] : acc[j];
        }
    }
}

// The same product on (32 TM WM) x (32 TN WN) tiles of C per workgroup of WM x WN wavefronts, staged through
// LDS.  Each wavefront owns a (32 TM) x (32 TN) piece (TM x TN MFMA tiles); a wider piece reads fewer LDS
// operands per MFMA (TM + TN reads for TM TN products), and a larger workgroup tile writes fewer LDS tile
// elements per MFMA.  Per 32-wide step of K the workgroup loads the A and B tiles
// with 16-byte global loads of 8 consecutive elements along whichever dimension is contiguous in memory
// (A_K: A's k stride is 1, else its m stride is 1; B_K: B's k stride is 1, else its n stride is 1) and stores
// them in LDS as As[m][k] and Bs[n][k] (k contiguous, row pitch GLDK), so that every MFMA operand is one
// 16-byte LDS read.  The next step's global loads are issued before the current step's MFMAs and stored to
// the second LDS buffer after them.  The 64 x 128 pieces are held to two wavefronts per SIMD (no
// accumulation registers apart from the VGPRs), which they need to beat 64 x 64 pieces.  Chunks that
// cross the matrix edge are read element by element with zero fill; the host launches this kernel only when
// the operands are 16-byte aligned and their other stride is a multiple of 8.
constexpr int GBK = 32, GLDK = GBK + 8;
// LDS offset of tile row r: pitch GLDK (80 bytes), and for a tile stashed as k-pairs (PAD) 8 more elements
// (16 bytes) after every 16 rows.  Measured on gfx950 (cycles per wavefront instruction, 16-byte operand
// reads of 32 rows x 2 chunks / 4-byte k-pair writes of rows 8 a + e, a = 0..15): pitch 80 alone 6.0 / 32.6,
// with 16 bytes per 16 rows 8.9 / 4.5, with 16 bytes per 8 rows 14.0 / 4.4.  A tile written with 16-byte
// stores (k-contiguous operand) has no pair writes and keeps the plain pitch.
template <bool PAD>
__device__ inline int lrow(int r) { return r * GLDK + (PAD ? (r >> 4) * 8 : 0); }
template <bool PAD>
__host__ __device__ constexpr int lds_rows(int rows) { return rows * GLDK + (PAD ? (rows / 16) * 8 : 0); }

typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

// 8 consecutive elements p[0], p[step], ... with p + i * step in range for i < n_in (n_in >= 8: one vector load)
__device__ inline u16x8 load8(const uint16_t* __restrict__ p, int n_in, bool ok) {
    u16x8 v;
    if (ok && n_in >= 8) return *reinterpret_cast<const u16x8*>(p);
    for (int i = 0; i < 8; ++i) v[i] = (ok && i < n_in) ? p[i] : uint16_t(0);
    return v;
}

// One operand's tile of R rows (m for A, n for B) x GBK k per step, R / 64 chunks of 8 elements per thread.
// K_CONTIG: chunk c = tid + 256 i holds 8 consecutive k of row c / 4 (k = 8 (c % 4)).  Otherwise chunks 2 g
// and 2 g + 1 of thread t hold rows 128 g + 8 (t % 16) .. + 7 at k = 2 (t / 16) and 2 (t / 16) + 1 (16-byte
// loads along the contiguous dimension); the stash writes each row's pair of k as one 4-byte word, so the
// tile reaches LDS transposed (k contiguous) in eight writes per pair of chunks.
template <int R, bool K_CONTIG, int NT>
struct TileIO {
    static constexpr int NC = 4 * R / NT;
    static constexpr bool PAD = !K_CONTIG;
    static constexpr int LDS = lds_rows<PAD>(R);
    static_assert(NC >= 1 && (K_CONTIG || NC % 2 == 0), "tile too small for the thread count");
    __device__ static void fetch(u16x8 (&v)[NC], const uint16_t* __restrict__ X, int64_t s_row, int64_t s_k,
                                 int row0, int rows, int k0, int K, int tid) {
        for (int i = 0; i < NC; ++i) {
            if constexpr (K_CONTIG) {
                const int c = tid + NT * i;
                const int row = row0 + c / 4, k = k0 + 8 * (c % 4);
                v[i] = load8(X + int64_t(row) * s_row + k, K - k, row < rows && k < K);
            } else {
                const int k = k0 + 2 * ((tid / 16) % 16) + (i & 1);
                const int row = row0 + 128 * (tid / 256 + (NT / 256) * (i >> 1)) + 8 * (tid % 16);
                v[i] = load8(X + int64_t(k) * s_k + row, rows - row, k < K && row < rows);
            }
        }
    }
    __device__ static void stash(uint16_t* S, const u16x8 (&v)[NC], int tid) {
        if constexpr (K_CONTIG) {
            for (int i = 0; i < NC; ++i) {
                const int c = tid + NT * i;
                *reinterpret_cast<u16x8*>(&S[lrow<PAD>(c / 4) + 8 * (c % 4)]) = v[i];
            }
        } else {
            const int k = 2 * ((tid / 16) % 16);
            for (int g = 0; g < NC / 2; ++g) {
                const int r = 128 * (tid / 256 + (NT / 256) * g) + 8 * (tid % 16);
                for (int e = 0; e < 8; ++e)
                    *reinterpret_cast<uint32_t*>(&S[lrow<PAD>(r + e) + k]) =
                        uint32_t(v[2 * g][e]) | (uint32_t(v[2 * g + 1][e]) << 16);
            }
        }
    }
};

template <bool A_K, bool B_K, int TM, int TN, int WM, int WN>
__global__ void __launch_bounds__(64 * WM * WN, (WM * WN == 4 && TM * TN > 4 ? 2 : 1))
    gemm_bf16_tiled_kernel(int M, int N, int K, const uint16_t* __restrict__ A, int64_t sam, int64_t sak,
                           const uint16_t* __restrict__ B, int64_t sbk, int64_t sbn, float* __restrict__ C,
                           int64_t ldc, int accumulate) {
    constexpr int NT = 64 * WM * WN, BM = 32 * TM * WM, BN = 32 * TN * WN;
    using IA = TileIO<BM, A_K, NT>;
    using IB = TileIO<BN, B_K, NT>;
    // two LDS buffers: step s computes from buffer s % 2 while the tiles of step s + 1 are written to the
    // other, so each step needs one barrier
    __shared__ __align__(16) uint16_t As[2][IA::LDS];
    __shared__ __align__(16) uint16_t Bs[2][IB::LDS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int wm = (wave / WN) * 32 * TM, wn = (wave % WN) * 32 * TN;
    const int r = lane & 31, half = lane >> 5;
    f32x16 acc[TM][TN];
    for (int i = 0; i < TM; ++i)
        for (int j = 0; j < TN; ++j)
            for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.0f;
    u16x8 ra[IA::NC], rb[IB::NC];
    auto fetch = [&](int k0) {
        IA::fetch(ra, A, sam, sak, m0, M, k0, K, tid);
        IB::fetch(rb, B, sbn, sbk, n0, N, k0, K, tid);
    };
    fetch(0);
    IA::stash(As[0], ra, tid);
    IB::stash(Bs[0], rb, tid);
    __syncthreads();
    int cur = 0;
    for (int k0 = 0; k0 < K; k0 += GBK) {
        const bool more = k0 + GBK < K;
        if (more) fetch(k0 + GBK);
        const uint16_t* as = As[cur];
        const uint16_t* bs = Bs[cur];
        for (int kk = 0; kk < GBK; kk += 16) {
            bf16x8 a[TM], b[TN];
            for (int t = 0; t < TM; ++t)
                a[t] = __builtin_bit_cast(
                    bf16x8, *reinterpret_cast<const u16x8*>(&as[lrow<IA::PAD>(wm + 32 * t + r) + kk + 8 * half]));
            for (int t = 0; t < TN; ++t)
                b[t] = __builtin_bit_cast(
                    bf16x8, *reinterpret_cast<const u16x8*>(&bs[lrow<IB::PAD>(wn + 32 * t + r) + kk + 8 * half]));
            for (int i = 0; i < TM; ++i)
                for (int j = 0; j < TN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        // the other buffer was last read in the previous step, before the barrier that ended it
        if (more) {
            IA::stash(As[cur ^ 1], ra, tid);
            IB::stash(Bs[cur ^ 1], rb, tid);
        }
        __syncthreads();
        cur ^= 1;
    }
    for (int i = 0; i < TM; ++i)
        for (int j = 0; j < TN; ++j)
            for (int e = 0; e < 16; ++e) {
                const int m = m0 + wm + 32 * i + 8 * (e >> 2) + 4 * half + (e & 3);
                const int n = n0 + wn + 32 * j + r;
                if (m < M && n < N) {
                    float* c = &C[int64_t(m) * ldc + n];
                    *c = accumulate ? *c + acc[i][j][e] : acc[i][j][e];
                }
            }
}

__device__ inline float gelu(float x) {
    const float k = 0.7978845608028654f;  // sqrt(2 / pi)
    return 0.5f * x * (1.0f + tanhf(k * (x + 0.044715f * x * x * x)));
}

__device__ inline float gelu_grad(float x) {
    const float k = 0.7978845608028654f;
    const float u = k * (x + 0.044715f * x * x * x);
    const float t = tanhf(u);
    return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * k * (1.0f + 3.0f * 0.044715f * x * x);
}

__global__ void to_bf16_kernel(const float* in, uint16_t* out, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = f32_to_bf16(in[i]);
}

// z += bias; hidden layers also write a = gelu(z) in fp32 and bf16.
__global__ void bias_act_kernel(float* z, const float* bias, int rows, int cols, int act, float* a, uint16_t* a16) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= int64_t(rows) * cols) return;
    const float v = z[i] + bias[i % cols];
    z[i] = v;
    if (act) {
        const float g = gelu(v);
        a[i] = g;
        a16[i] = f32_to_bf16(g);
    }
}

// dY = 2 (Y - T) / count; per-block partial of sum (Y - T)^2.
__global__ void mse_grad_kernel(const float* y, const float* t, int64_t n, float scale, float* dy, uint16_t* dy16,
                                double* partial) {
    __shared__ double red[256];
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    double s = 0.0;
    if (i < n) {
        const float d = y[i] - t[i];
        const float g = 2.0f * d * scale;
        dy[i] = g;
        dy16[i] = f32_to_bf16(g);
        s = double(d) * double(d);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (int(threadIdx.x) < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// dZ = dA * gelu'(Z) in fp32 and bf16.
__global__ void gelu_back_kernel(const float* da, const float* z, int64_t n, float* dz, uint16_t* dz16) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = da[i] * gelu_grad(z[i]);
    dz[i] = g;
    dz16[i] = f32_to_bf16(g);
}

// db[n] = sum over rows of dZ[row][n] in a fixed order: 32 columns per 256-thread workgroup, the rows split
// into 8 interleaved groups (thread g of a column sums rows g, g + 8, ...), then the 8 group sums added in
// group order.  (One thread per column walking all rows took 55 us for 256 x 256, half the training step.)
__global__ void __launch_bounds__(256) bias_grad_kernel(const float* dz, int rows, int cols, float* db) {
    __shared__ float part[8][32];
    const int cl = threadIdx.x & 31, g = threadIdx.x >> 5;
    const int c = blockIdx.x * 32 + cl;
    float sum = 0.0f;
    if (c < cols)
        for (int r = g; r < rows; r += 8) sum += dz[int64_t(r) * cols + c];
    part[g][cl] = sum;
    __syncthreads();
    if (g == 0 && c < cols) {
        float t = part[0][cl];
        for (int k = 1; k < 8; ++k) t += part[k][cl];
        db[c] = t;
    }
}

__global__ void adam_kernel(float* w, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
                            float eps, float c1, float c2, uint16_t* w16) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    w16[i] = f32_to_bf16(wi);
}

__global__ void reduce_double_kernel(const double* p, int n, double* out) {
    __shared__ double red[256];
    double s = 0.0;
    for (int k = threadIdx.x; k < n; k += 256) s += p[k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (int(threadIdx.x) < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

inline unsigned nblk(int64_t n, int t = 256) { return unsigned((n + t - 1) / t); }

// Which GEMM kernel a product uses: 1 = LDS-tiled with 128 x 128 tiles, 2 = 128 x 256 tiles, 3 = 256 x 256
// tiles, 0 = one wavefront per 32 x 32 tile.  ED_GEMM_KERNEL forces one (for measurements); otherwise the tiled
// kernel whenever the operands allow its 16-byte loads, its tile chosen by the size of C.
int gemm_kernel_choice() {
    static const int v = [] {
        const char* e = std::getenv("ED_GEMM_KERNEL");
        return e ? std::atoi(e) : -1;
    }();
    return v;
}

void gemm(hipStream_t st, int M, int N, int K, const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B,
          int64_t sbk, int64_t sbn, float* C, int64_t ldc, int accumulate) {
    // 16-byte loads of 8 elements: one unit stride per operand, the other a multiple of 8, aligned bases
    const bool a_k = sak == 1 && sam % 8 == 0, a_m = sam == 1 && sak % 8 == 0;
    const bool b_k = sbk == 1 && sbn % 8 == 0, b_n = sbn == 1 && sbk % 8 == 0;
    const bool aligned = (reinterpret_cast<uintptr_t>(A) % 16 == 0) && (reinterpret_cast<uintptr_t>(B) % 16 == 0);
    const int choice = gemm_kernel_choice();
    if (choice != 0 && aligned && (a_k || a_m) && (b_k || b_n)) {
        // 256 x 256 tiles (eight wavefronts of 64 x 128) once C has 256 of them (one per CU), 128 x 256 (four
        // wavefronts of 64 x 128) once it has 128, else 128 x 128; measured at 8192^3 with B n-contiguous: 787,
        // 739 and 629 TFLOP/s
        const int64_t mn = int64_t(M) * N;
        const bool big = choice == 3 || (choice < 0 && mn >= (int64_t(1) << 24));
        const bool wide = choice == 2 || (choice < 0 && mn >= (int64_t(1) << 22));
        auto go = [&](auto kernel, int bm, int bn, int nt) {
            const dim3 grid{unsigned((N + bn - 1) / bn), unsigned((M + bm - 1) / bm), 1u};
            hipLaunchKernelGGL(kernel, grid, dim3(nt), 0, st, M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, accumulate);
        };
        auto pick = [&](auto a_kc, auto b_kc) {
            constexpr bool AK = decltype(a_kc)::value, BK = decltype(b_kc)::value;
            if (big)
                go(gemm_bf16_tiled_kernel<AK, BK, 2, 4, 4, 2>, 256, 256, 512);
            else if (wide)
                go(gemm_bf16_tiled_kernel<AK, BK, 2, 4, 2, 2>, 128, 256, 256);
            else
                go(gemm_bf16_tiled_kernel<AK, BK, 2, 2, 2, 2>, 128, 128, 256);
        };
        if (a_k && b_k)
            pick(std::true_type{}, std::true_type{});
        else if (a_k)
            pick(std::true_type{}, std::false_type{});
        else if (b_k)
            pick(std::false_type{}, std::true_type{});
        else
            pick(std::false_type{}, std::false_type{});
        return;
    }
    const dim3 grid{unsigned((N + 31) / 32), unsigned((M + 31) / 32), 1u};
    hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(64), 0, st, M, N, K, A, sam, sak, B, sbk, sbn, C, ldc,
                       accumulate);
}

}  // namespace
}  // namespace ed

using namespace ed;

struct ed_mlp {
    std::vector<int> dims;         // n_in, hidden..., n_out
    std::vector<int64_t> w_off;    // offset of W_l in the parameter vector
    std::vector<int64_t> b_off;    // offset of b_l
    int64_t n_params = 0;
    float *d_p = nullptr, *d_g = nullptr, *d_m = nullptr, *d_v = nullptr;
    uint16_t* d_p16 = nullptr;
    int64_t step = 0;
    int cap = 0;                   // batch capacity of the activation buffers
    std::vector<float*> z, a, dz;  // per layer (index l = 1..L)
    std::vector<uint16_t*> a16, dz16;
    uint16_t* x16 = nullptr;
    float* da = nullptr;
    double* partial = nullptr;
    double* loss = nullptr;

    int layers() const { return int(dims.size()) - 1; }

    void free_acts() {
        for (auto* p : z) dev_free(p);
        for (auto* p : a) dev_free(p);
        for (auto* p : dz) dev_free(p);
        for (auto* p : a16) dev_free(p);
        for (auto* p : dz16) dev_free(p);
        dev_free(x16);
        dev_free(da);
        dev_free(partial);
        dev_free(loss);
        z.clear();
        a.clear();
        dz.clear();
        a16.clear();
        dz16.clear();
        x16 = nullptr;
        da = nullptr;
        partial = nullptr;
        loss = nullptr;
        cap = 0;
    }

    void ensure(int batch) {
        if (batch <= cap) return;
        free_acts();
        const int L = layers();
        int wmax = 0;
        for (int d : dims) wmax = std::max(wmax, d);
        z.assign(L + 1, nullptr);
        a.assign(L + 1, nullptr);
        dz.assign(L + 1, nullptr);
        a16.assign(L + 1, nullptr);
        dz16.assign(L + 1, nullptr);
        for (int l = 1; l <= L; ++l) {
            z[l] = dev_alloc<float>(size_t(batch) * dims[l]);
            a[l] = dev_alloc<float>(size_t(batch) * dims[l]);
            dz[l] = dev_alloc<float>(size_t(batch) * dims[l]);
            a16[l] = dev_alloc<uint16_t>(size_t(batch) * dims[l]);
            dz16[l] = dev_alloc<uint16_t>(size_t(batch) * dims[l]);
        }
        x16 = dev_alloc<uint16_t>(size_t(batch) * dims[0]);
        da = dev_alloc<float>(size_t(batch) * wmax);
        partial = dev_alloc<double>(size_t(nblk(int64_t(batch) * dims.back())) + 1);
        loss = dev_alloc<double>(1);
        cap = batch;
    }

    void release() {
        free_acts();
        dev_free(d_p);
        dev_free(d_g);
        dev_free(d_m);
        dev_free(d_v);
        dev_free(d_p16);
    }

    // forward pass for `batch` rows of x (fp32, device); output in z[L]
    void forward(const float* x, int batch, hipStream_t st) {
        ensure(batch);
        const int L = layers();
        hipLaunchKernelGGL(to_bf16_kernel, dim3(nblk(int64_t(batch) * dims[0])), dim3(256), 0, st, x, x16,
                           int64_t(batch) * dims[0]);
        const uint16_t* in16 = x16;
        for (int l = 1; l <= L; ++l) {
            const int K = dims[l - 1], N = dims[l];
            gemm(st, batch, N, K, in16, K, 1, d_p16 + w_off[l - 1], N, 1, z[l], N, 0);
            hipLaunchKernelGGL(bias_act_kernel, dim3(nblk(int64_t(batch) * N)), dim3(256), 0, st, z[l],
                               d_p + b_off[l - 1], batch, N, l < L ? 1 : 0, a[l], a16[l]);
            in16 = a16[l];
        }
    }

    // forward + backward of the mean squared error; gradients into d_g
    void backward(const float* x, const float* t, int batch, hipStream_t st) {
        forward(x, batch, st);
        const int L = layers();
        const int64_t n_out = int64_t(batch) * dims[L];
        const unsigned nb = nblk(n_out);
        hipLaunchKernelGGL(mse_grad_kernel, dim3(nb), dim3(256), 0, st, z[L], t, n_out, 1.0f / float(n_out), dz[L],
                           dz16[L], partial);
        hipLaunchKernelGGL(reduce_double_kernel, dim3(1), dim3(256), 0, st, partial, int(nb), loss);
        for (int l = L; l >= 1; --l) {
            const int K = dims[l - 1], N = dims[l];
            const uint16_t* in16 = l > 1 ? a16[l - 1] : x16;
            // dW[k][n] = sum_b in[b][k] dZ[b][n]
            gemm(st, K, N, batch, in16, 1, K, dz16[l], N, 1, d_g + w_off[l - 1], N, 0);
            hipLaunchKernelGGL(bias_grad_kernel, dim3(unsigned((N + 31) / 32)), dim3(256), 0, st, dz[l], batch, N,
                               d_g + b_off[l - 1]);
            if (l > 1) {
                // dA[b][k] = sum_n dZ[b][n] W[k][n]
                gemm(st, batch, K, N, dz16[l], N, 1, d_p16 + w_off[l - 1], 1, N, da, K, 0);
                hipLaunchKernelGGL(gelu_back_kernel, dim3(nblk(int64_t(batch) * K)), dim3(256), 0, st, da, z[l - 1],
                                   int64_t(batch) * K, dz[l - 1], dz16[l - 1]);
            }
        }
    }
};

extern "C" {

int ed_mlp_create(const int* dims, int n_dims, uint64_t seed, ed_mlp** out) {
    return guarded([&] {
        ED_REQUIRE(out && dims && n_dims >= 2, ED_ERR_ARG, "need at least input and output widths");
        auto* m = new ed_mlp;
        m->dims.assign(dims, dims + n_dims);
        for (int d : m->dims) ED_REQUIRE(d >= 1 && d <= 16384, ED_ERR_ARG, "layer width out of range");
        int64_t off = 0;
        for (int l = 1; l < n_dims; ++l) {
            m->w_off.push_back(off);
            off += int64_t(m->dims[l - 1]) * m->dims[l];
            m->b_off.push_back(off);
            off += m->dims[l];
        }
        m->n_params = off;
        // He-style initialisation on the host, seeded
        std::vector<float> p(size_t(off), 0.0f);
        std::mt19937_64 rng(seed);
        for (int l = 1; l < n_dims; ++l) {
            std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / float(m->dims[l - 1])));
            for (int64_t i = 0; i < int64_t(m->dims[l - 1]) * m->dims[l]; ++i) p[m->w_off[l - 1] + i] = nd(rng);
        }
        std::vector<uint16_t> p16(p.size());
        for (size_t i = 0; i < p.size(); ++i) p16[i] = f32_to_bf16(p[i]);
        hipStream_t st = nullptr;
        m->d_p = dev_upload(p, st);
        m->d_p16 = dev_upload(p16, st);
        std::vector<float> zero(p.size(), 0.0f);
        m->d_g = dev_upload(zero, st);
        m->d_m = dev_upload(zero, st);
        m->d_v = dev_upload(zero, st);
        ED_HIP(hipStreamSynchronize(st));
        *out = m;
    });
}

int ed_mlp_destroy(ed_mlp* m) {
    return guarded([&] {
        if (!m) return;
        m->release();
        delete m;
    });
}

int64_t ed_mlp_num_params(const ed_mlp* m) { return m ? m->n_params : -1; }

int ed_mlp_get_params(const ed_mlp* m, float* host, int64_t n) {
    return guarded([&] {
        ED_REQUIRE(m && n == m->n_params, ED_ERR_ARG, "parameter count mismatch");
        ED_HIP(hipMemcpy(host, m->d_p, size_t(n) * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int ed_mlp_set_params(ed_mlp* m, const float* host, int64_t n) {
    return guarded([&] {
        ED_REQUIRE(m && n == m->n_params, ED_ERR_ARG, "parameter count mismatch");
        std::vector<float> p(host, host + n);
        std::vector<uint16_t> p16(p.size());
        for (size_t i = 0; i < p.size(); ++i) p16[i] = f32_to_bf16(p[i]);
        ED_HIP(hipMemcpy(m->d_p, p.data(), size_t(n) * sizeof(float), hipMemcpyHostToDevice));
        ED_HIP(hipMemcpy(m->d_p16, p16.data(), size_t(n) * sizeof(uint16_t), hipMemcpyHostToDevice));
        std::vector<float> zero(p.size(), 0.0f);
        ED_HIP(hipMemcpy(m->d_m, zero.data(), size_t(n) * sizeof(float), hipMemcpyHostToDevice));
        ED_HIP(hipMemcpy(m->d_v, zero.data(), size_t(n) * sizeof(float), hipMemcpyHostToDevice));
        m->step = 0;
    });
}

int ed_mlp_forward(ed_mlp* m, const float* d_x, int batch, float* d_y, void* stream) {
    return guarded([&] {
        ED_REQUIRE(m && batch >= 1, ED_ERR_ARG, "bad batch");
        hipStream_t st = static_cast<hipStream_t>(stream);
        m->forward(d_x, batch, st);
        ED_HIP(hipMemcpyAsync(d_y, m->z[m->layers()], size_t(batch) * m->dims.back() * sizeof(float),
                              hipMemcpyDeviceToDevice, st));
        ED_HIP(hipGetLastError());
    });
}

int ed_mlp_gradients(ed_mlp* m, const float* d_x, const float* d_t, int batch, float* host_grad, double* loss,
                     void* stream) {
    return guarded([&] {
        ED_REQUIRE(m && batch >= 1, ED_ERR_ARG, "bad batch");
        hipStream_t st = static_cast<hipStream_t>(stream);
        m->backward(d_x, d_t, batch, st);
        ED_HIP(hipGetLastError());
        if (host_grad)
            ED_HIP(hipMemcpyAsync(host_grad, m->d_g, size_t(m->n_params) * sizeof(float), hipMemcpyDeviceToHost, st));
        double l = 0.0;
        ED_HIP(hipMemcpyAsync(&l, m->loss, sizeof(double), hipMemcpyDeviceToHost, st));
        ED_HIP(hipStreamSynchronize(st));
        if (loss) *loss = l / double(int64_t(batch) * m->dims.back());
    });
}

int ed_mlp_train_step(ed_mlp* m, const float* d_x, const float* d_t, int batch, float lr, double* loss,
                      void* stream) {
    return guarded([&] {
        ED_REQUIRE(m && batch >= 1, ED_ERR_ARG, "bad batch");
        hipStream_t st = static_cast<hipStream_t>(stream);
        m->backward(d_x, d_t, batch, st);
        m->step += 1;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.0f - std::pow(b1, float(m->step)), c2 = 1.0f - std::pow(b2, float(m->step));
        hipLaunchKernelGGL(adam_kernel, dim3(nblk(m->n_params)), dim3(256), 0, st, m->d_p, m->d_g, m->d_m, m->d_v,
                           m->n_params, lr, b1, b2, 1e-8f, c1, c2, m->d_p16);
        ED_HIP(hipGetLastError());
        if (loss) {
            double l = 0.0;
            ED_HIP(hipMemcpyAsync(&l, m->loss, sizeof(double), hipMemcpyDeviceToHost, st));
            ED_HIP(hipStreamSynchronize(st));
            *loss = l / double(int64_t(batch) * m->dims.back());
        }
    });
}

// C = A . B with BF16 operands given as raw uint16 and element strides.
int ed_gemm_bf16(int M, int N, int K, const uint16_t* d_A, int64_t sam, int64_t sak, const uint16_t* d_B, int64_t sbk,
                 int64_t sbn, float* d_C, int64_t ldc, void* stream) {
    return guarded([&] {
        ED_REQUIRE(M >= 0 && N >= 0 && K >= 0, ED_ERR_ARG, "negative GEMM size");
        if (M == 0 || N == 0) return;
        gemm(static_cast<hipStream_t>(stream), M, N, K, d_A, sam, sak, d_B, sbk, sbn, d_C, ldc, 0);
        ED_HIP(hipGetLastError());
    });
}

}  // extern "C"
