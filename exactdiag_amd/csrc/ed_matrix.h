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
int64_t hv_partials(const ed_sector* s);
int hv_default_tiles(const ed_sector* s);
int64_t launch_hmat(const ed_hmat* m, const void* x, void* y, double* partial, int precision, hipStream_t st);
const ed_sector* hmat_sector(const ed_hmat* m);
HvWeights hmat_weights(const ed_hmat* m);
struct StoredView;
StoredView hmat_view(const ed_hmat* m);
void launch_reduce_partials(const double* partial, int64_t n, double* out, hipStream_t st);

}  // namespace ed
