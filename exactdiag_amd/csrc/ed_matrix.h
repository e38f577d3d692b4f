// Internal interface between the matrix kernels and the Lanczos driver.
#pragma once

#include "ed_internal.h"

struct ed_hmat;

namespace ed {

struct HvWeights {
    C2 w[4];     // tl, tr, jl, jr
    int active;  // bit t set when |w_t| >= 1e-6
};

HvWeights make_weights(const double* w);
int64_t launch_hv(const ed_sector* s, const HvWeights& w, const void* x, void* y, double* partial, int precision,
                  int remap, hipStream_t st);
int64_t launch_hv_lds(const ed_sector* s, const HvWeights& w, const void* x, void* y, double* partial, int precision,
                      hipStream_t st);
int64_t hv_partials(const ed_sector* s);  // the most partials any H·v launch on s writes
// After the basis: the wavefronts that the wavefront-uniform H·v leaves to the per-lane walk (s->d_fix, n_fix);
// none when the sector does not take that walk.
void hv_fix_list(ed_sector* s, hipStream_t st);
int64_t hv_fix_partials(const ed_sector* s);  // the workgroups (partials) that walk those wavefronts
int64_t launch_hmat(const ed_hmat* m, const void* x, void* y, double* partial, int precision, hipStream_t st);
const ed_sector* hmat_sector(const ed_hmat* m);
bool hmat_is_real(const ed_hmat* m);
void hmat_require_precision(const ed_hmat* m, int precision);  // throws ED_ERR_ARG
HvWeights hmat_weights(const ed_hmat* m);
struct StoredView;
StoredView hmat_view(const ed_hmat* m);
void launch_reduce_partials(const double* partial, int64_t n, double* out, hipStream_t st);
// Exact per-column elements of columns [lo, lo + n) (the reference's
// calc_column_matrix_elements); count[t] > column_capacity() flags overflow.
void launch_columns(const ed_sector* ini, const ed_sector* fin, const RuleHost& rule, int triangle, int max_per_col,
                    int64_t lo, int64_t n, int32_t* d_count, int64_t* d_rows, double* d_vals, hipStream_t st);
int column_capacity();
void apply_columns(const ed_sector* ini, const ed_sector* fin, const RuleHost& rule, const double* d_x, double* d_y,
                   hipStream_t st);
// Sectors with the spin flip: an excitation operator between them must carry a definite spin-flip parity
// (Szq reverses z, N(q) and the currents keep it; A+- change S^z); other entry points take none (ED_ERR_ARG).
void require_flip_operator(const ed_sector* ini, const ed_sector* fin, int op);
void require_no_flip(const ed_sector* s, const char* what);

}  // namespace ed
