// Harness around the reference package's own C++ headers (not copied: the
// Makefile in this folder compiles them where they lie under the reference
// checkout).  TEST INFRASTRUCTURE ONLY: tests/ use it to pin the oracle's
// orbit, representative, amplitude, rank and basis-scan functions
// bit-for-bit, and bench.py may time its sparse dot as the reference CPU
// baseline.  Only the headers that build with this image's toolchain are
// used: tJ_spin_half_ladder/symmetry_ladder.h (with its includes) and
// general/cpp_sparse_dot.h.  general/basis_indexing.h needs <format>, which
// this image does not ship, so calc_basis_map's driver loop
// (general/basis_indexing.h:188-242) is restated here around the
// reference's per-hole-pattern functions.

#include <cstdint>
#include <memory>
#include <vector>
#include <omp.h>

#include "tJ_spin_half_ladder/symmetry_ladder.h"
#include "general/cpp_sparse_dot.h"

namespace {

struct RefSector {
    int L, N, h, nd;
    std::shared_ptr<I_Symmetry_Generator> sym;
    std::shared_ptr<State_Index_Amplitude_Translator> tr;
    std::vector<int64_t> sparse;
    std::vector<int32_t> counts;
};

std::vector<std::vector<MATRIX_INDEX_TYPE>> binomials(int n) {
    std::vector<std::vector<MATRIX_INDEX_TYPE>> c(n + 1, std::vector<MATRIX_INDEX_TYPE>(n + 1, 0));
    for (int i = 0; i <= n; ++i) {
        c[0][i] = 1;
        for (int j = 1; j <= i; ++j) c[j][i] = c[j - 1][i - 1] + (j <= i - 1 ? c[j][i - 1] : 0);
    }
    return c;
}

}  // namespace

extern "C" {

void* ref_sector_new(int L, int h, int nd, int kx, int ky, int threads) {
    auto* s = new RefSector;
    s->L = L;
    s->N = 2 * L;
    s->h = h;
    s->nd = nd;
    auto comb = binomials(2 * L);
    int n_spin = int(comb[nd][2 * L - h]);
    s->sym = std::make_shared<Symmetries_Single_Spin_Half_2leg_Ladder>(L, std::vector<int>{kx, ky});
    s->tr = std::make_shared<State_Index_Amplitude_Translator>(h, nd, n_spin, comb, s->sym);
    const long long n_hole = comb[h][2 * L];
    std::vector<MATRIX_INDEX_TYPE> elig(n_hole);
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
    for (long long i = 0; i < n_hole; ++i) elig[i] = s->tr->check_major_index_supports_lowest_states(i);
    std::vector<MATRIX_INDEX_TYPE> kept;
    for (auto e : elig)
        if (e != -1) kept.push_back(e);
    std::vector<Indices_Counts> parts(kept.size());
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
    for (long long i = 0; i < (long long)kept.size(); ++i)
        parts[i] = s->tr->get_lowest_sparse_indices_from_major_index(kept[i]);
    for (auto& p : parts) {
        s->sparse.insert(s->sparse.end(), p.indices.begin(), p.indices.end());
        s->counts.insert(s->counts.end(), p.counts.begin(), p.counts.end());
    }
    return s;
}

void ref_sector_free(void* p) { delete static_cast<RefSector*>(p); }

int64_t ref_sector_dim(void* p) { return int64_t(static_cast<RefSector*>(p)->sparse.size()); }

void ref_sector_basis(void* p, int64_t* sparse, int32_t* counts) {
    auto* s = static_cast<RefSector*>(p);
    std::copy(s->sparse.begin(), s->sparse.end(), sparse);
    std::copy(s->counts.begin(), s->counts.end(), counts);
}

int64_t ref_hole_eligible(void* p, int64_t hole_index) {
    return static_cast<RefSector*>(p)->tr->check_major_index_supports_lowest_states(hole_index);
}

void ref_unrank(void* p, int64_t sparse, short* st) {
    auto* s = static_cast<RefSector*>(p);
    auto v = s->tr->sparse_index_to_state(sparse);
    std::copy(v.begin(), v.end(), st);
}

int64_t ref_rank(void* p, const short* st) { return static_cast<RefSector*>(p)->tr->state_to_sparse_index(st); }

int ref_orbit(void* p, const short* st, int check_lowest, short* out_states, double* out_amp) {
    auto* s = static_cast<RefSector*>(p);
    std::vector<State_Amplitude> im(s->N, State_Amplitude());
    int n = s->sym->get_symmetry_states_from_state(st, check_lowest != 0, im.data());
    for (int k = 0; k < n; ++k) {
        std::copy(im[k].state, im[k].state + s->N, out_states + size_t(k) * s->N);
        out_amp[2 * k] = im[k].amplitude.real();
        out_amp[2 * k + 1] = im[k].amplitude.imag();
    }
    return n;
}

void ref_lowest(void* p, const short* st, int64_t* ind, double* amp) {
    Index_Amplitude ia;
    static_cast<RefSector*>(p)->tr->get_lowest_sparse_amplitude_from_state(st, ia);
    *ind = ia.ind;
    amp[0] = ia.amplitude.real();
    amp[1] = ia.amplitude.imag();
}

// The reference's I_Symmetry_Generator::get_q_dot_r
// (general/group/symmetry_generator.h:87-102) on the ladder's generator, for
// one site and one integer q: out = {real, imag}.
void ref_q_dot_r(int L, int site, int qx, int qy, double* out) {
    Symmetries_Single_Spin_Half_2leg_Ladder sym(L, std::vector<int>{0, 0});
    const std::complex<double> v = sym.get_q_dot_r(pos_int(site), std::vector<int>{qx, qy});
    out[0] = v.real();
    out[1] = v.imag();
}

// Row-major Hermitian-triangle sparse dot as driven by
// general/cython_sparse_dot.pyx:40-80 (driver restated), calling the
// reference kernels of general/cpp_sparse_dot.h:59-71 and :104-138.
void ref_csr_dot(int64_t D, const int64_t* major, const short* value_inds, const int64_t* minor,
                 const double* values, const double* x, double* y, double wr, double wi, int first, int threads) {
    using c = std::complex<double>;
    const c* vals = reinterpret_cast<const c*>(values);
    const c* vin = reinterpret_cast<const c*>(x);
    c* vout = reinterpret_cast<c*>(y);
    const c w(wr, wi);
#pragma omp parallel num_threads(threads)
    {
#pragma omp for schedule(dynamic, 10000)
        for (long long r = 0; r < D; ++r) {
            if (first) vout[r] = 0;
        }
#pragma omp for schedule(dynamic, 10000)
        for (long long r = 0; r < D; ++r)
            vout[r] += w * matvec_explicit_elements_pass_row_major(r, (const long long*)major, vals, value_inds,
                                                                   (const long long*)minor, vin);
        const long long nt = omp_get_num_threads(), t = omp_get_thread_num();
        const long long lo = std::min<long long>(D, (D / nt + 1) * t);
        const long long hi = t == nt - 1 ? D : std::min<long long>(D, (D / nt + 1) * (t + 1));
        matvec_hermitian_conjugate_pass_row_major(lo, hi, D, (const long long*)major, vals,
                                                  value_inds, (const long long*)minor, vin, true, vout, w);
    }
}

// Rows [row_lo, row_lo + n_rows) of a row-major Hermitian-triangle CSR
// (major has n_rows + 1 entries): the explicit pass for those rows and their
// Hermitian-conjugate scatter, each thread owning a range of output
// indices as in general/cython_sparse_dot.pyx:58-79.
void ref_csr_dot_rows(int64_t D, int64_t row_lo, int64_t n_rows, const int64_t* major, const short* value_inds,
                      const int64_t* minor, const double* values, const double* x, double* y, double wr, double wi,
                      int threads) {
    using c = std::complex<double>;
    const c* vals = reinterpret_cast<const c*>(values);
    const c* vin = reinterpret_cast<const c*>(x);
    c* vout = reinterpret_cast<c*>(y);
    const c w(wr, wi);
#pragma omp parallel num_threads(threads)
    {
#pragma omp for schedule(dynamic, 10000)
        for (long long k = 0; k < n_rows; ++k)
            vout[row_lo + k] += w * matvec_explicit_elements_pass_row_major(k, (const long long*)major, vals,
                                                                            value_inds, (const long long*)minor, vin);
        const long long nt = omp_get_num_threads(), t = omp_get_thread_num();
        const long long lo = std::min<long long>(D, (D / nt + 1) * t);
        const long long hi = t == nt - 1 ? D : std::min<long long>(D, (D / nt + 1) * (t + 1));
        for (long long k = 0; k < n_rows; ++k) {
            const long long row = row_lo + k;
            for (long long j = major[k]; j < major[k + 1]; ++j) {
                const long long col = minor[j];
                if (col >= hi) break;
                if (col < lo || col == row) continue;
                vout[col] += vin[row] * std::conj(vals[value_inds[j]]) * w;
            }
        }
    }
}

}  // extern "C"
