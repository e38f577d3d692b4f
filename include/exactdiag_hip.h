/* C ABI of the MI355X (gfx950) t-J ladder exact-diagonalisation library
 * (libexactdiag_hip.so).  Plain pointers and sizes only.  Every function
 * returns 0 on success and a negative status on failure; ed_last_error()
 * gives the thread's last message.  Device pointers are HIP device memory
 * owned by the caller; `stream` is a hipStream_t (0 = null stream).
 *
 * Complex vectors are interleaved (re, im) doubles ("c128") or floats
 * ("c64").  Dense basis indices follow the reference's order (ascending
 * sparse index).
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference package root src/exactdiag/).
 */
#ifndef EXACTDIAG_HIP_H
#define EXACTDIAG_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ed_sector ed_sector;

enum ed_status {
    ED_OK = 0,
    ED_ERR_ARG = -1,      /* invalid argument (the reference raises ValueError) */
    ED_ERR_HIP = -2,      /* a HIP runtime call failed */
    ED_ERR_STATE = -3,    /* call order violated (e.g. basis not built) */
    ED_ERR_CAPACITY = -4  /* output capacity too small (the reference raises IndexError) */
};

/* Vector types of the products and Lanczos runs ("precision" arguments).
 * The reference computes in complex128 throughout.  The real types serve
 * stored matrices that ed_hmat_is_real reports real (momenta kx = 0, L/2):
 * there a real start vector keeps every Lanczos vector real. */
enum ed_precision {
    ED_PREC_C128 = 0,  /* complex128 vectors */
    ED_PREC_C64 = 1,   /* complex64 vectors, FP64 sums and reductions */
    ED_PREC_R64 = 2,   /* float64 vectors (stored matrices only) */
    ED_PREC_R32 = 3    /* float32 vectors, FP64 sums and reductions (stored matrices only) */
};

/* Operator ids for ed_op_columns / ed_op_apply. */
enum ed_operator {
    ED_OP_SZQ = 0,          /* tJ_spin_half_ladder/matrix_setup.pyx:235-246 */
    ED_OP_NQ = 1,           /* density N(q): no counterpart in the reference */
    ED_OP_APLUS_DOWN = 2,   /* matrix_setup.pyx:281-300, adds a down spin */
    ED_OP_APLUS_UP = 3,     /* matrix_setup.pyx:281-300, adds an up spin */
    ED_OP_AMINUS_DOWN = 4,  /* matrix_setup.pyx:302-325, removes a down spin */
    ED_OP_AMINUS_UP = 5,    /* matrix_setup.pyx:302-325, removes an up spin */
    ED_OP_CURRENT_RUNG = 6, /* matrix_setup.pyx:248-263 (caller passes qy already shifted by one) */
    ED_OP_CURRENT_LEG = 7   /* matrix_setup.pyx:265-279 */
};

/* Position-correlation kinds (tJ_spin_half_ladder/configs.py:236-241). */
enum ed_position_kind {
    ED_POS_HOLES = 0,   /* "hole_correlations" */
    ED_POS_SZ = 1,      /* "Sz_correlations" */
    ED_POS_SINGLET = 2  /* "singlet-singlet" */
};

const char* ed_last_error(void);
int ed_version(void);
/* The library keeps the device blocks of 64 KiB or more that it frees (up to
 * 24 GiB per device) for its next allocations, since hipFree of a large
 * block costs about 0.25 ms; this returns the kept blocks of the current
 * device to HIP.  No counterpart in the reference (CPU memory). */
int ed_cache_release(void);

/* Sector = (num_rungs L, num_holes h, num_down_spins nd, k = (kx, ky)).
 * Replaces Symmetries_Single_Spin_Half_2leg_Ladder + State_Index_Amplitude_
 * Translator (tJ_spin_half_ladder/symmetry_ladder.h:55-406) as built by
 * py_get_ladder_translators (tJ_spin_half_ladder/matrix_setup.pyx:377-406).
 * Host tables (binomials, amplitudes, hole patterns) are built here. */
int ed_sector_create(int L, int h, int nd, int kx, int ky, ed_sector** out);
int ed_sector_destroy(ed_sector* s);
/* A sector of the group extended by the global spin flip, with parity z = +1
 * or -1 (S^z = 0 only).  No counterpart in the reference, whose ladder group is
 * leg translations x leg mirror (tJ_spin_half_ladder/symmetry_ladder.h:55-70):
 * optional, for callers that want the halved dimension; the sectors of
 * ed_sector_create keep the reference's bases.  The spin flip relabels spins
 * without a fermion sign; the H·v, stored-H, Lanczos and batched entry points
 * and the operators Szq (z -> -z), N(q) and the currents (z -> z) take these
 * sectors; A+-, position correlations, the reference's matrix folders and
 * the multi-momentum product refuse them (ED_ERR_ARG). */
int ed_sector_create_flip(int L, int h, int nd, int kx, int ky, int z, ed_sector** out);
int ed_sector_flip(const ed_sector* s); /* z, or 0 without the spin flip */

/* Enumerate the symmetric basis on the GPU.  Replaces calc_basis_map
 * (general/basis_indexing.h:188-242). */
int ed_basis_build(ed_sector* s, void* stream);
int64_t ed_basis_size(const ed_sector* s); /* Basis_Index_Map::get_num_states, basis_indexing.h:182-184 */

/* Basis in the reference's format: sparse index (int64) and orbit size
 * (int32) per dense index; both device pointers.  Replaces
 * Basis_Index_Map::get_sparse / get_num_sparse_in_dense (basis_indexing.h:148-150, :170-172). */
int ed_basis_export(const ed_sector* s, int64_t* d_sparse, int32_t* d_orbit, void* stream);

/* Sparse index -> dense index (-1 when absent) for n device-resident
 * indices.  Replaces Basis_Index_Map::get_dense (basis_indexing.h:152-168). */
int ed_lookup(const ed_sector* s, const int64_t* d_sparse, int64_t* d_dense, int64_t n, void* stream);

/* Lowest image of n arbitrary states given by sparse index: out sparse index
 * of the representative (-1 if incompatible) and its normalised amplitude
 * (c128).  Replaces get_lowest_sparse_amplitude_from_state (general/symmetry.h:240-269). */
int ed_lowest(const ed_sector* s, const int64_t* d_sparse, int64_t* d_rep_sparse, double* d_amp, int64_t n,
              void* stream);

/* Matrix elements of one Hamiltonian term (0 tl, 1 tr, 2 jl, 3 jr) column by
 * column, before the reference's unique-value table: for column i,
 * d_count[i] entries at d_rows/d_vals[i * max_per_col ...], sorted by row.
 * triangle = keep rows >= i only.  Replaces calc_column_matrix_elements with
 * the rules of setup_hamiltonian (general/column_functions.h:150-207,
 * tJ_spin_half_ladder/matrix_setup.pyx:80-129). */
int ed_term_columns(const ed_sector* s, int term, int triangle, int max_per_col, int32_t* d_count, int64_t* d_rows,
                    double* d_vals, void* stream);

/* Same for an excitation operator from sector `init` to sector `fin`
 * (setup_excitation_operator, matrix_setup.pyx:195-374). */
int ed_op_columns(const ed_sector* init, const ed_sector* fin, int op, int qx, int qy, int max_per_col,
                  int32_t* d_count, int64_t* d_rows, double* d_vals, void* stream);

/* Position-correlation operators of one sector
 * (get_position_correlation_operator, tJ_spin_half_ladder/matrix_setup.pyx:409-480):
 * kind ED_POS_HOLES / ED_POS_SZ project on holes / multiply S^z signs at
 * (r, r + each shift) summed over r (diagonal); ED_POS_SINGLET is
 * (0.5 / N) sum_i Delta_i^dagger(y) Delta_{i+j}(x).  shifts: n_shifts
 * (leg, rung) int pairs in host memory, the config's fixed distances then
 * the free shift (singlet: y, x, j).  ed_position_op_columns as
 * ed_op_columns; ed_position_op_apply: d_y = O d_x (c128); 
 * ed_position_expectation: out[0..1] = <psi|O|psi> (host), as
 * calculate_singlet_singlet_correlations takes np.vdot(eigvec,
 * operator.dot(eigvec)) (tJ_spin_half_ladder/position_correlations.pyx:129-142). */
int ed_position_op_columns(const ed_sector* s, int kind, const int* shifts, int n_shifts, int max_per_col,
                           int32_t* d_count, int64_t* d_rows, double* d_vals, void* stream);
int ed_position_op_apply(const ed_sector* s, int kind, const int* shifts, int n_shifts, const double* d_x,
                         double* d_y, void* stream);
int ed_position_expectation(const ed_sector* s, int kind, const int* shifts, int n_shifts, const double* d_psi,
                            double* out, void* stream);

/* Hole or S^z position correlations of a state (c128 device vector), one
 * value per node index d in out[0..2L) (host): sum over basis states of
 * |psi|^2 times the number of reference sites r whose positions (r, r + the
 * unit shifts of d, r + each fixed distance) all hold holes (kind
 * ED_POS_HOLES), or the product of the S^z signs there (ED_POS_SZ, 0 with a
 * hole).  Replaces calculate_hole_spin_projection_correlations and
 * calculate_position_correlations
 * (tJ_spin_half_ladder/position_correlations.pyx:84-127, :186-230). */
int ed_position_correlations(const ed_sector* s, int kind, const int* fixed, int n_fixed, const double* d_psi,
                             double* out, void* stream);

/* Matrix-free y = sum_T w_T H_T x over the four terms (w: 4 complex
 * weights as 8 doubles, host memory).  Optional d_dot (>= 1 c128 on device):
 * receives <x, y>.  precision: 0 = c128 vectors, 1 = c64 vectors (FP64 sums
 * inside a row, FP32 storage).  Replaces Added_Sparse_Matrices.dot
 * (general/sparse_matrices.py:233-237) over sparse_matrix_dot
 * (general/cython_sparse_dot.pyx:40-100). */
int ed_hv(const ed_sector* s, const double* w, const void* d_x, void* d_y, double* d_dot, int precision,
          void* stream);

/* y = O x for an excitation operator (x in sector init, y in sector fin).
 * Replaces excitation_operator.dot in lanczos_dynamical_response
 * (general/cython_lanczos_diagonalization.pyx:36). */
int ed_op_apply(const ed_sector* init, const ed_sector* fin, int op, int qx, int qy, const double* d_x, double* d_y,
                void* stream);

/* Plain Lanczos recurrence from start vector d_v0 (c128, need not be
 * normalised): a[0..n), b[0..n) with b[0] = 0, stopping when b < 1e-8, as
 * _lanczos_get_anbn (general/cython_lanczos_diagonalization.pyx:50-77).
 * Returns the number of coefficients in *n_out.  d_basis (optional,
 * nsteps * D c128) receives the normalised Lanczos vectors.  precision as in
 * ed_hv; the coefficient reductions are FP64 in both. */
int ed_lanczos_run(const ed_sector* s, const double* w, const double* d_v0, int nsteps, double* a, double* b,
                   int* n_out, double* d_basis, int precision, void* stream);

/* Second pass of two-pass Lanczos: rerun the recurrence of ed_lanczos_run
 * from the same d_v0 and accumulate d_x = sum_n ritz[n] v_n (c128, FP64
 * accumulation also when precision = 1), e.g. the Ritz vector of a
 * tridiagonal eigenvector.  No counterpart in the reference (which keeps no
 * Lanczos vectors; its eigenvectors come from ARPACK,
 * general/lanczos_diagonalization.py:81-121). */
int ed_lanczos_ritz(const ed_sector* s, const double* w, const double* d_v0, int nsteps, const double* ritz,
                    double* d_x, int precision, void* stream);

/* Stored Hamiltonian of one sector and one set of weights, built once on the
 * GPU for repeated products.  Replaces calc_sparse_matrix
 * (general/cython_sparse_matrices.pyx:58-241: four upper-triangle CSR
 * matrices with int16 indices into unique values) and Sparse_Matrix.dot /
 * Added_Sparse_Matrices.dot (general/sparse_matrices.py:83-105, 233-237;
 * general/cpp_sparse_dot.h:59-138).  Layout: gather form, sliced ELL of
 * 64-row slices, int32 column + uint16 code into a table of exact unique
 * values per element.  The sector must outlive the matrix.  ed_hmat_info:
 * device bytes, stored elements, slots (with padding), unique values.
 * ed_hmat_hv as ed_hv; ed_hmat_lanczos_run / _ritz as ed_lanczos_run /
 * ed_lanczos_ritz with the stored matrix as H. */
typedef struct ed_hmat ed_hmat;
int ed_hmat_create(const ed_sector* s, const double* w, ed_hmat** out, void* stream);
int ed_hmat_destroy(ed_hmat* m);
int ed_hmat_info(const ed_hmat* m, int64_t* bytes, int64_t* nnz, int64_t* slots, int* n_values);
/* Bytes per stored element: 4 when column, value index and sign fit one
 * 32-bit word, else 6 (the environment variable ED_HMAT_ELEMENT_BYTES=6,
 * read at ed_hmat_create, forces 6). */
int ed_hmat_element_bytes(const ed_hmat* m);
/* precision ED_PREC_R64 / ED_PREC_R32: d_x, d_y are float64 / float32
 * arrays of D values and d_dot receives <x, y> with a zero imaginary part;
 * ED_ERR_ARG unless ed_hmat_is_real. */
int ed_hmat_hv(const ed_hmat* m, const double* d_x, double* d_y, double* d_dot, int precision, void* stream);
/* 1 when every stored value is real (to 1e-13 of the largest magnitude), so
 * that H is a real symmetric matrix in the sector's basis: the momenta kx = 0
 * and kx = L/2 (the reference's characters exp(-i 2 pi kx r / L) are then
 * +-1; it still computes these sectors in complex128,
 * tJ_spin_half_ladder/symmetry_ladder.h:113-235).  0 otherwise. */
int ed_hmat_is_real(const ed_hmat* m);
/* Autotuner for the stored product (no counterpart in the reference): row
 * tiles of 256 per workgroup (1..32) x gathers in flight per lane (4, 8) x
 * plain or non-temporal streaming of the stored elements and the output;
 * times the product with its fused dot and keeps the fastest in the matrix.
 * ed_hmat_set_config sets the same three choices (tiles 0 = default;
 * nontemporal -1 = the default: non-temporal). */
int ed_hmat_tune(ed_hmat* m, const double* d_x, double* d_y, int precision, int* tiles, int* unroll,
                 int* nontemporal, double* best_ms, void* stream);
int ed_hmat_set_config(ed_hmat* m, int tiles, int unroll, int nontemporal);
/* Read-rate probe (no counterpart in the reference; measurement only): the
 * rate, in 128-byte lines per second, at which the device serves 16-byte
 * loads from a buffer of span_bytes (a power of two in [1 MiB, 4 GiB]),
 * scattered = 1: every lane of a wavefront from its own line at a hashed
 * place (as the stored product's far gathers), 0: 64 consecutive elements per
 * wavefront.  Best of three launches of 8.4 M wavefront loads each, timed with
 * HIP events on stream; the buffer is taken from the device memory cache. */
int ed_hmat_line_probe(int scattered, int64_t span_bytes, double* lines_per_s, void* stream);
/* With a real precision (ED_PREC_R64, ED_PREC_R32) the start vector d_v0
 * (complex128) must be real: ED_ERR_ARG when its imaginary part exceeds
 * 1e-10 of its norm (rotate a global phase away first) or when the matrix is
 * not real.  The Ritz vector d_x stays complex128 (zero imaginary part). */
int ed_hmat_lanczos_run(const ed_hmat* m, const double* d_v0, int nsteps, double* a, double* b, int* n_out,
                        int precision, void* stream);
int ed_hmat_lanczos_ritz(const ed_hmat* m, const double* d_v0, int nsteps, const double* ritz, double* d_x,
                         int precision, void* stream);

/* ed_hmat_lanczos_run that also keeps the normalised Lanczos vectors: row n
 * of d_basis (device, rows of ld >= D elements of the precision's vector
 * type) receives v_n, so that a Ritz vector is one combination of the rows
 * instead of the second pass of ed_hmat_lanczos_ritz.  No counterpart in the
 * reference (scipy's ARPACK keeps its own basis,
 * general/lanczos_diagonalization.py:81-121); meant for sectors whose nsteps
 * vectors fit in HBM beside the matrix. */
int ed_hmat_lanczos_run_basis(const ed_hmat* m, const double* d_v0, int nsteps, double* a, double* b, int* n_out,
                              void* d_basis, int64_t ld, int precision, void* stream);

/* Stop rule of a time-to-accuracy Lanczos run (no counterpart in the
 * reference, whose lanczos_dynamical_response runs a fixed number of steps,
 * general/cython_lanczos_diagonalization.pyx:18-77).  From min_steps on,
 * after every check_every steps, the spectrum of the first n coefficients
 * closed by the square-root terminator (ed_spectrum_change) is compared
 * with those of the first n - delta and n - delta / 2; the run stops at the
 * first n where the larger relative L1 distance on the grid ws[0],
 * ws[stride], ... (z = w + e0 + i eta) is below tol. */
typedef struct ed_lanczos_stop {
    const double* ws; /* host grid of n_w frequencies */
    int n_w;
    int stride;       /* >= 1: every stride-th frequency enters the estimate */
    double e0, eta, tol;
    int min_steps, check_every, delta;
} ed_lanczos_stop;
/* ed_hmat_lanczos_run that stops by the rule (at most max_steps).  The check
 * of each chunk runs on the host while the GPU computes the next one, so a
 * run that stops at n has also computed up to check_every further steps;
 * only the first n coefficients are returned (*n_out = n).  *reason: 0 =
 * max_steps reached, 1 = the rule was met, 2 = the Krylov space was
 * exhausted (b below 1e-8; then no terminator applies). */
int ed_hmat_lanczos_run_tol(const ed_hmat* m, const double* d_v0, int max_steps, const ed_lanczos_stop* stop,
                            double* a, double* b, int* n_out, int* reason, int precision, void* stream);
/* The rule's estimate on the host (no GPU): the larger relative L1 distance
 * between the terminated spectrum of the first n coefficients and those of
 * the first n - delta and n - delta / 2 (a[0..n), b[0..n), b[0] unused).
 * Each terminator takes the centre and
 * half-width w of the extreme eigenvalues of its tridiagonal (bisection),
 * a_inf = centre, b_inf = w / 2; the L1 weights are the grid's central
 * differences. */
int ed_spectrum_change(const double* a, const double* b, int n, int delta, const ed_lanczos_stop* stop,
                       double* out);

/* The reference's on-disk matrices, held in HBM.  An ed_csr is one stored
 * matrix in the reference's layout: values complex128[n_values] (re, im
 * pairs), value_inds int16[nnz], minor int64[nnz], major int64[n_major + 1],
 * n_major = rows for row-major matrices, cols for column-major ones.
 *
 * ed_csr_from_sector builds Hamiltonian term `term` (0 tl, 1 tr, 2 jl, 3 jr)
 * of a sector as calc_sparse_matrix writes it for setup_hamiltonian
 * (general/cython_sparse_matrices.pyx:58-241 with the rules of
 * tJ_spin_half_ladder/matrix_setup.pyx:41-129): sorted, unique values
 * joined within 1e-6 in order of first appearance; triangle = the
 * config's triangle_only (upper triangle, read back as a Hermitian
 * triangle), row_major = config.major == "row" (value table conjugated,
 * pyx:139-145).  The reference's default is triangle = row_major = 1.  ed_csr_from_operator builds an
 * excitation operator from sector init to sector fin as
 * setup_excitation_operator stores it (matrix_setup.pyx:195-374):
 * column-major, shape (fin rows, init cols).  Both return ED_ERR_CAPACITY
 * where the reference raises IndexError (more than 32767 unique values).
 *
 * ed_csr_create uploads arrays read from a matrix folder
 * (Sparse_Matrix.__init__, general/sparse_matrices.py:27-80); it checks the
 * pointers and indices and rejects arrays that do not form a matrix.
 * ed_csr_info / ed_csr_copy_out read a matrix back (host pointers; any may
 * be null).  max_count = the longest major slice. */
typedef struct ed_csr ed_csr;
int ed_csr_from_sector(const ed_sector* s, int term, int triangle, int row_major, ed_csr** out, void* stream);
int ed_csr_from_operator(const ed_sector* init, const ed_sector* fin, int op, int qx, int qy, ed_csr** out,
                         void* stream);
/* A position-correlation operator (see ed_position_op_columns) stored as
 * get_position_correlation_operator writes it: column-major, sorted. */
int ed_csr_from_position_operator(const ed_sector* s, int kind, const int* shifts, int n_shifts, ed_csr** out,
                                  void* stream);
int ed_csr_create(int64_t rows, int64_t cols, int row_major, int hermitian_triangle, int64_t nnz, int n_values,
                  const double* values, const int16_t* value_inds, const int64_t* minor, const int64_t* major,
                  ed_csr** out, void* stream);
int ed_csr_info(const ed_csr* m, int64_t* rows, int64_t* cols, int64_t* nnz, int* n_values, int* max_count);
int ed_csr_copy_out(const ed_csr* m, double* values, int16_t* value_inds, int64_t* minor, int64_t* major,
                    void* stream);
int ed_csr_destroy(ed_csr* m);

/* Weighted sum of stored matrices of one shape (weights: n complex values
 * as 2n doubles, host memory), applied as Added_Sparse_Matrices.dot over
 * Sparse_Matrix.dot / sparse_matrix_dot (general/sparse_matrices.py:83-105,
 * :233-237; general/cython_sparse_dot.pyx:40-100;
 * general/cpp_sparse_dot.h:59-138): explicit elements times w, elements of
 * a Hermitian triangle applied transposed with conj(value) times w (w not
 * conjugated), terms with |w| < 1e-6 left out.  ed_csr_sum_hv: d_y = sum
 * (complex128 device vectors, d_y of `rows`, d_x of `cols` entries).  The
 * matrices may be destroyed after ed_csr_sum_create.  Layout: rows in
 * slices of 64, 8 bytes (column, value code) per element; ed_csr_sum_info
 * gives the elements (nnz) and the slots with the slice padding. */
typedef struct ed_csr_sum ed_csr_sum;
int ed_csr_sum_create(const ed_csr* const* terms, const double* weights, int n, ed_csr_sum** out, void* stream);
int ed_csr_sum_info(const ed_csr_sum* m, int64_t* rows, int64_t* cols, int64_t* nnz, int64_t* slots);
int ed_csr_sum_hv(const ed_csr_sum* m, const double* d_x, double* d_y, void* stream);
int ed_csr_sum_destroy(ed_csr_sum* m);

/* Batched Lanczos over many sectors of uneven size (instances), each with
 * its own weights (8 doubles per instance as in ed_hv).  Replaces the
 * one-sector-at-a-time loops of get_all_k_eigenpairs
 * (tJ_spin_half_ladder/api.py:45-59) and of the spectra drivers
 * (api.py:116-161), each of which runs calculate_eigenpairs or
 * lanczos_dynamical_response (general/lanczos_diagonalization.py:53-121,
 * general/cython_lanczos_diagonalization.pyx:18-77) per sector.
 * Vectors of all instances form one c128 pool; instance i owns rows
 * [offsets[i], offsets[i + 1]) (ed_batch_offsets writes n + 1 values).
 * The sectors must outlive the batch. */
typedef struct ed_batch ed_batch;
int ed_batch_create(ed_sector* const* sectors, const double* weights, int n, ed_batch** out);
/* The same batch over stored matrices (ed_hmat_create): each instance's H·v
 * reads its stored H.  The matrices and their sectors must outlive the batch. */
int ed_batch_create_stored(const ed_hmat* const* mats, int n, ed_batch** out);
int ed_batch_destroy(ed_batch* b);
int ed_batch_offsets(const ed_batch* b, int64_t* offsets);
/* d_y = H d_x for every instance; dots (host, 2n doubles, optional) gets
 * <x_i, H x_i> per instance. */
int ed_batch_hv(ed_batch* b, const double* d_x, double* d_y, double* dots, void* stream);
/* Lanczos on every instance, nsteps[i] <= ld steps from the pool vector d_v0
 * (any norm), stopping an instance when b < 1e-8, all steps of all
 * instances in one HIP graph launch when use_graph.  a[i*ld + n] and
 * b[i*(ld+1) + n] (b_0 = 0; b at n_out[i] is the norm of the last
 * residual); n_out[i] = steps done.  With ritz (host, n*ld) and d_acc
 * (pool), d_acc = sum_n ritz[i*ld + n] v_n (second pass of two-pass
 * Lanczos; the recurrence is deterministic). */
int ed_batch_lanczos(ed_batch* b, const double* d_v0, const int* nsteps, int ld, const double* ritz, double* d_acc,
                     double* a, double* bcoef, int* n_out, int use_graph, void* stream);
/* Autotuner (no counterpart in the reference): time the batched H·v over
 * the tiles per workgroup (1..32) and gathers in flight per lane (2, 4, 8),
 * keep the fastest.  ed_batch_set_config / get_config set and read it. */
int ed_batch_tune(ed_batch* b, const double* d_x, double* d_y, int* tiles, int* gather, double* best_ms,
                  void* stream);
int ed_batch_set_config(ed_batch* b, int tiles, int gather);
int ed_batch_get_config(const ed_batch* b, int* tiles, int* gather, int* n_jobs);
/* The same autotuner for the single-sector H·v of ed_hv / ed_lanczos_run;
 * the choice is stored in the sector. */
int ed_hv_tune(ed_sector* s, const double* w, int precision, const double* d_x, double* d_y, int* tiles, int* gather,
               double* best_ms, void* stream);

/* Several momentum sectors of one ladder in one product.  Replaces, for up
 * to 16 sectors k of the same (L, h, nd) at once, the per-sector H.dot
 * (general/sparse_matrices.py:83-105, :233-237) and the Lanczos recurrence
 * (general/cython_lanczos_diagonalization.pyx:18-77) that get_spectrum runs
 * once per q (tJ_spin_half_ladder/api.py:116-132).
 *
 * ed_sector_create_union: the union of the momentum sectors, every lowest
 * image under the ladder's group (symmetry_ladder.h:113-235 without the
 * compatibility test); build its basis with ed_basis_build.
 * ed_mk_create: one stored structure over the union, slot k carrying the
 * momentum (kx[k], ky[k]) and the term weights w[8k .. 8k+7] (tl, tr, jl, jr
 * as re, im pairs); the union sector must outlive the matrix.  Vectors are
 * blocks complex[D_union][16] in scaled
 * coordinates x'_f = x_f / sqrt(n_f), n_f the orbit size; slot k is masked
 * to sector k's states.
 * ed_mk_attach: write sector vector d_phi (the sector's own dense order) into
 * slot `slot` of d_block (zero the block first) and mark the slot's states;
 * ed_mk_extract: the reverse.  The sector must have the slot's momentum.
 * ed_mk_hv: d_y = H d_x for all slots (precision 0: complex128 blocks, 1:
 * complex64 blocks); d_dot (device, 32 doubles, optional) gets <x_k, H x_k>.
 * ed_mk_lanczos_run: plain Lanczos from the complex128 block d_v0 for all
 * slots, a[k*nsteps + n], b[k*(nsteps+1) + n] (b_0 = 0), n_out[k] = steps
 * done, each slot stopping at b < 1e-8; vectors complex128 (precision 0) or
 * complex64 with FP64 reductions (precision 1).
 * ed_mk_lanczos_ritz: the same recurrence again (second pass of two-pass
 * Lanczos) accumulating d_x (complex128 block) = sum_n ritz[k*nsteps + n]
 * v_n(k) for every slot in FP64.
 * ed_mk_residual_norms: out[k] = sum_f n_f |hx_k(f) - e[k] x_k(f)|^2 for
 * complex128 blocks (d_hx null: the squared norm of x_k); ed_mk_scale: slot k
 * of a complex128 block times s[k].  Calls on one matrix share its scratch and
 * must not overlap. */
typedef struct ed_mk ed_mk;
int ed_sector_create_union(int L, int h, int nd, ed_sector** out);
int ed_mk_create(const ed_sector* u, const double* w, int n_slots, const int* kx, const int* ky, ed_mk** out,
                 void* stream);
int ed_mk_destroy(ed_mk* m);
/* Launch shape of the product (no counterpart in the reference): waves = 4
 * or 8 wavefronts per workgroup (16 or 32 rows), unroll = 4 or 8 gathers in
 * flight per lane. */
int ed_mk_set_config(ed_mk* m, int waves, int unroll);
int ed_mk_info(const ed_mk* m, int64_t* dim, int64_t* nnz, int64_t* bytes);
int ed_mk_attach(ed_mk* m, int slot, const ed_sector* sk, const double* d_phi, double* d_block, void* stream);
int ed_mk_extract(const ed_mk* m, int slot, const ed_sector* sk, const double* d_block, double* d_out, void* stream);
int ed_mk_hv(const ed_mk* m, const double* d_x, double* d_y, double* d_dot, int precision, void* stream);
int ed_mk_lanczos_run(const ed_mk* m, const double* d_v0, int nsteps, double* a, double* b, int* n_out,
                      int precision, void* stream);
int ed_mk_lanczos_ritz(const ed_mk* m, const double* d_v0, int nsteps, const double* ritz, double* d_x, int precision,
                       void* stream);
int ed_mk_residual_norms(const ed_mk* m, const double* d_x, const double* d_hx, const double* e, double* out,
                         void* stream);
int ed_mk_scale(const ed_mk* m, double* d_x, const double* s, void* stream);

/* Continued fraction of general/cython_lanczos_diagonalization.pyx:42-48
 * evaluated on the GPU: out[k] = factor * Im(1 / det(z_k)) for z_k = ws[k] +
 * e0 + i*eta.  Host arrays in and out. */
int ed_cf_eval(const double* a, const double* b, int n, const double* ws, int n_w, double e0, double eta,
               double factor, double* out);
/* As ed_cf_eval, with the fraction closed after level n-1 by the tail of a
 * constant continuation (a_inf, b_inf) (square-root terminator; b_inf = 0
 * gives ed_cf_eval).  No counterpart in the reference: used by the
 * learned-closure form. */
int ed_cf_eval_tail(const double* a, const double* b, int n, const double* ws, int n_w, double e0, double eta,
                    double factor, double a_inf, double b_inf, double* out);

/* Dense vector helpers for the Krylov eigensolver (all c128, device).
 * ed_vec_dots: out[k] = <V_k, w> for k < nv (host out, FP64 deterministic).
 * ed_vec_axpys: w -= sum_k c[k] V_k (c host).  ed_vec_combine: out = sum_k
 * c[k] V_k (c host).  ed_vec_norm2: host out = ||w||^2.  ed_vec_scale: w *= c. */
int ed_vec_dots(const double* d_V, int nv, int64_t ld, const double* d_w, int64_t D, double* out, void* stream);
/* ed_vec_dots without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * partials live in a fixed 2 MiB array per GPU, held until the process ends;
 * at most 128 vectors (ED_ERR_CAPACITY beyond).  Results equal ed_vec_dots bit
 * for bit.  Calls that overlap on one GPU (other threads or streams) share
 * that array and race. */
int ed_vec_dots_noalloc(const double* d_V, int nv, int64_t ld, const double* d_w, int64_t D, double* out,
                        void* stream);
int ed_vec_axpys(const double* d_V, int nv, int64_t ld, const double* c, double* d_w, int64_t D, void* stream);
int ed_vec_combine(const double* d_V, int nv, int64_t ld, const double* c, double* d_out, int64_t D, void* stream);
int ed_vec_norm2(const double* d_w, int64_t D, double* out, void* stream);
int ed_vec_scale(double* d_w, int64_t D, double cr, double ci, void* stream);

/* Learned continued-fraction closure (no counterpart in the reference): an
 * MLP dims[0] -> dims[1] ... -> dims[n-1] with GELU on hidden layers, BF16
 * weights in MFMA GEMMs with FP32 accumulation, FP32 master weights and
 * Adam.  Parameters are laid out as W_1 (dims[0] x dims[1], row-major), b_1,
 * W_2, b_2, ...  Inputs/targets are FP32 row-major [batch][features] on the
 * device.  The loss is the mean squared error over all outputs. */
typedef struct ed_mlp ed_mlp;
int ed_mlp_create(const int* dims, int n_dims, uint64_t seed, ed_mlp** out);
int ed_mlp_destroy(ed_mlp* m);
int64_t ed_mlp_num_params(const ed_mlp* m);
int ed_mlp_get_params(const ed_mlp* m, float* host, int64_t n);
int ed_mlp_set_params(ed_mlp* m, const float* host, int64_t n);
int ed_mlp_forward(ed_mlp* m, const float* d_x, int batch, float* d_y, void* stream);
int ed_mlp_gradients(ed_mlp* m, const float* d_x, const float* d_t, int batch, float* host_grad, double* loss,
                     void* stream);
int ed_mlp_train_step(ed_mlp* m, const float* d_x, const float* d_t, int batch, float lr, double* loss,
                      void* stream);
/* C[M][N] = A[M][K] . B[K][N] with BF16 operands (raw 16-bit, element
 * strides) and FP32 output, on v_mfma_f32_32x32x16_bf16. */
int ed_gemm_bf16(int M, int N, int K, const uint16_t* d_A, int64_t sam, int64_t sak, const uint16_t* d_B, int64_t sbk,
                 int64_t sbn, float* d_C, int64_t ldc, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* EXACTDIAG_HIP_H */
