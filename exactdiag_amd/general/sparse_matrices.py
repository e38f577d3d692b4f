"""GPU operators with the reference's matrix interface (reference: general/sparse_matrices.py).

The reference stores each Hamiltonian term as a CSR on disk and sums them in
``Added_Sparse_Matrices.dot`` (general/sparse_matrices.py:210-243).  Here the
weighted sum is stored once in HBM (sliced ELL, built on the GPU the first
time the operator is applied) when the sector is large enough for that to
pay and the device has room for it; otherwise it is applied matrix-free.
Both are HIP kernels of the library.  The objects keep the ``.dot``,
``.shape``, ``.weights``, ``.toarray()`` surface.  ``dot`` accepts numpy
arrays (copied to and from the device) or device tensors (kept there).
"""

from __future__ import annotations

import dataclasses
import json
import logging
import os
import pathlib

import numpy as np

from exactdiag_amd.sector import TERMS, Sector, weights_array

_logger = logging.getLogger(__name__)

VALUE_TYPE = np.complex128
STORED_MIN_DIM = 2048  # smaller sectors: the build does not pay for itself


def stored_bytes_bound(sector) -> int:
    """Upper bound of the stored H's size: every row as long as the longest possible one."""
    N, L, h = sector.N, sector.L, sector.h
    longest = N + L + 3 * min(h, N - h)
    return int(sector.dim * (3 + 6 * longest) + 8 * (sector.dim // 64 + 1))


def _to_device(vec):
    import torch

    if isinstance(vec, torch.Tensor):
        if not vec.is_cuda:
            return vec.to("cuda"), "torch_cpu"
        return vec, "torch"
    return torch.as_tensor(np.ascontiguousarray(vec, np.complex128), device="cuda"), "numpy"


def _from_device(t, kind):
    if kind == "torch":
        return t
    if kind == "torch_cpu":
        return t.cpu()
    return t.cpu().numpy()


class Hamiltonian_Operator:
    """sum_T w_T H_T on one sector, applied on the GPU (replaces Added_Sparse_Matrices)."""

    def __init__(self, sector: Sector, weights):
        self.sector = sector
        if isinstance(weights, dict):
            self.names = [t for t in TERMS if weights.get(t, 0) != 0]
            w = [weights.get(t, 0) for t in TERMS]
        else:
            self.names = list(TERMS)
            w = list(weights)
        self.weights = np.asarray(w, np.complex128)
        self._w = weights_array(self.weights)
        self.shape = sector.shape
        self.num_matrices = len(self.names)
        self._stored = None
        self.use_stored = True
        # Lanczos runs on a real stored H (momenta kx = 0, L/2) use real vectors when the start vector is real
        # up to a global phase: the same coefficients to rounding, about 25 % less time per step
        self.use_real_vectors = True

    @property
    def stored(self):
        """The stored form, built on first use, or None (small sector, or not enough free HBM)."""
        if self._stored is None and self.use_stored and self.shape[0] >= STORED_MIN_DIM:
            import torch

            free, _ = torch.cuda.mem_get_info()
            if stored_bytes_bound(self.sector) < 0.5 * free:
                self._stored = self.sector.store(self._w)
            else:
                self.use_stored = False
        return self._stored

    def apply(self, x):
        """H x for a device vector (complex128 or complex64)."""
        m = self.stored
        return m.hv(x) if m is not None else self.sector.hv(self._w, x)

    def lanczos(self, v0, nsteps, precision=0):
        """Plain Lanczos coefficients (a, b) from the device vector v0 (precision 0: complex128 vectors, 1:
        complex64; float64 / float32 instead when the stored H is real and v0 is real up to a phase)."""
        m = self.stored
        if m is not None:
            if self.use_real_vectors and m.is_real and precision in (0, 1):
                from exactdiag_amd._lib import PREC_R32, PREC_R64
                from exactdiag_amd.sector import real_phase

                r = real_phase(v0)
                if r is not None:
                    return m.lanczos(r, nsteps, precision=PREC_R64 if precision == 0 else PREC_R32)
            return m.lanczos(v0, nsteps, precision=precision)
        return self.sector.lanczos(self._w, v0, nsteps, precision=precision)

    def dot(self, vec):
        if self.shape[0] == 0:
            return np.zeros(0, VALUE_TYPE)
        x, kind = _to_device(vec)
        import torch

        if x.dtype not in (torch.complex128, torch.complex64):
            x = x.to(torch.complex128)
        return _from_device(self.apply(x.contiguous()), kind)

    def dot_with_inner(self, x):
        """(H x, <x, H x>) with the inner product fused into the kernel."""
        m = self.stored
        return m.hv(x, with_dot=True) if m is not None else self.sector.hv(self._w, x, with_dot=True)

    def matrix(self):
        """Dense matrix of ``dot`` (small sectors only): columns H e_k from the GPU kernel."""
        import torch

        D = self.shape[0]
        eye = torch.eye(D, dtype=torch.complex128, device="cuda")
        cols = [self.sector.hv(self._w, eye[:, k].contiguous()) for k in range(D)]
        return torch.stack(cols, dim=1).cpu().numpy()

    def toarray(self):
        """Dense array in the reference's convention.

        Sparse_Matrix.toarray (general/sparse_matrices.py:110-125) writes a
        row-major entry at [minor, major], so for the Hermitian row-major
        Hamiltonian it returns the complex conjugate of the matrix that
        ``dot`` applies.  Kept as is for drop-in compatibility.
        """
        return np.conj(self.matrix())


class Excitation_Operator:
    """Operator between an initial and a final sector (replaces the column-major Sparse_Matrix)."""

    def __init__(self, initial: Sector, final: Sector, op: int, qx: int, qy: int, name: str = ""):
        self.initial, self.final = initial, final
        self.op, self.qx, self.qy = int(op), int(qx), int(qy)
        self.name = name
        self.shape = (final.dim, initial.dim)

    def dot(self, vec):
        x, kind = _to_device(vec)
        y = self.initial.op_apply(self.final, self.op, self.qx, self.qy, x)
        return _from_device(y, kind)

    def columns(self):
        return self.initial.op_columns(self.final, self.op, self.qx, self.qy)

    def toarray(self):
        ptr, rows, vals = self.columns()
        mat = np.zeros(self.shape, VALUE_TYPE)
        for i in range(self.shape[1]):
            for k in range(ptr[i], ptr[i + 1]):
                mat[rows[k], i] += vals[k]
        return mat


# ---------------------------------------------------------------------------
# The reference's matrix folders (general/sparse_matrices.py:17-125,
# general/cython_sparse_matrices.pyx:29-57 and :110-184).

MATRIX_INDEX_TYPE = np.int64
VALUE_INDEX_TYPE = np.int16
FILE_NAMES = {
    "info_name": "matrix_info.json",
    "values_name": "values.npy",
    "minor_inds_name": "minor_inds.npy",
    "value_inds_name": "value_inds.npy",
    "major_pointers_name": "major_pointers.npy",
}
_SIGNATURE_FIELDS = ("matrix_name", "shape", "symmetry_strings", "max_values_per_column", "major", "values_name",
                     "minor_inds_name", "value_inds_name", "major_pointers_name", "folder_name", "info_name",
                     "initial_system_info", "final_system_info")


@dataclasses.dataclass
class Signature:
    """JSON metadata of a stored matrix, field for field as the reference's (cython_sparse_matrices.pyx:29-57)."""

    matrix_name: str
    shape: tuple
    symmetry_strings: tuple
    max_values_per_column: int
    major: str
    values_name: str
    minor_inds_name: str
    value_inds_name: str
    major_pointers_name: str
    folder_name: str
    info_name: str
    initial_system_info: dict
    final_system_info: "dict | None" = None

    def __post_init__(self):
        self.shape = tuple(int(v) for v in self.shape)
        self.symmetry_strings = tuple(str(v) for v in self.symmetry_strings)
        self.max_values_per_column = int(self.max_values_per_column)
        if self.major not in ("row", "column"):
            raise ValueError(f"major must be 'row' or 'column', not {self.major!r}")
        if "triangle_only" in self.symmetry_strings and "hermitian" not in self.symmetry_strings:
            raise ValueError("symmetry option triangle_only requires the hermitian option")

    def __getitem__(self, key):
        return getattr(self, key)

    def as_json(self) -> dict:
        d = dataclasses.asdict(self)
        d["shape"] = list(self.shape)
        d["symmetry_strings"] = list(self.symmetry_strings)
        return {k: d[k] for k in _SIGNATURE_FIELDS}

    @property
    def is_row_major(self) -> bool:
        return self.major == "row"

    @property
    def is_hermitian_triangle(self) -> bool:
        return "triangle_only" in self.symmetry_strings and "hermitian" in self.symmetry_strings


def _as_signature(signature) -> Signature:
    if isinstance(signature, Signature):
        return signature
    extra = set(signature) - set(_SIGNATURE_FIELDS)
    if extra:
        raise ValueError(f"unknown signature fields {sorted(extra)}")
    return Signature(**signature)


def read_signature(folder) -> Signature:
    """The signature a folder holds (its matrix_info.json)."""
    folder = pathlib.Path(folder)
    with open(folder / FILE_NAMES["info_name"], encoding="utf-8") as fp:
        return _as_signature(json.load(fp))


def write_matrix_folder(signature, values, value_inds, minor_inds, major_pointers) -> pathlib.Path:
    """Write a matrix folder as _Tmp_To_Perm_Transcriber.rewrite_to_perm does (pyx:110-184).

    matrix_info.json (indent 4) plus four raw little-endian arrays without a header, under the file names
    the signature gives (the reference names them .npy but writes them with np.memmap).
    """
    sig = _as_signature(signature)
    folder = pathlib.Path(sig.folder_name)
    folder.mkdir(parents=True, exist_ok=True)
    arrays = (
        (sig.values_name, np.ascontiguousarray(values, "<c16")),
        (sig.value_inds_name, np.ascontiguousarray(value_inds, "<i2")),
        (sig.minor_inds_name, np.ascontiguousarray(minor_inds, "<i8")),
        (sig.major_pointers_name, np.ascontiguousarray(major_pointers, "<i8")),
    )
    n_major = sig.shape[0] if sig.is_row_major else sig.shape[1]
    if len(arrays[3][1]) != n_major + 1:
        raise ValueError(f"major_pointers has {len(arrays[3][1])} entries, the shape needs {n_major + 1}")
    for name, arr in arrays:
        tmp = folder / (name + ".tmp")
        arr.tofile(tmp)
        os.replace(tmp, folder / name)
    with open(folder / sig.info_name, "w", encoding="utf-8") as fp:
        json.dump(sig.as_json(), fp, indent=4)
    return folder


def read_matrix_folder(signature):
    """(values, value_inds, minor_inds, major_pointers) memory-mapped read-only from a folder."""
    sig = _as_signature(signature)
    folder = pathlib.Path(sig.folder_name)

    def mm(name, dtype):
        path = folder / name
        if path.stat().st_size == 0:
            return np.zeros(0, dtype)
        return np.memmap(path, dtype=dtype, mode="r")

    return (mm(sig.values_name, "<c16"), mm(sig.value_inds_name, "<i2"), mm(sig.minor_inds_name, "<i8"),
            mm(sig.major_pointers_name, "<i8"))


class Sparse_Matrix:
    """A stored matrix folder, applied on the GPU (reference: general/sparse_matrices.py:27-140).

    The constructor follows the reference: the folder's matrix_info.json wins over a conflicting requested
    signature (with a warning).  The arrays are memory-mapped and uploaded to HBM on the first product;
    ``dot`` is the GPU product with the reference's semantics (explicit elements, plus the transposed
    conjugate of the off-diagonal ones for a Hermitian triangle; weights below 1e-6 skip the matrix).
    """

    def __init__(self, signature, dtype=VALUE_INDEX_TYPE, num_threads=12):
        requested = _as_signature(signature)
        folder = pathlib.Path(requested.folder_name)
        loaded = read_signature(folder)
        conflicts = [(k, loaded[k], requested[k]) for k in _SIGNATURE_FIELDS if loaded[k] != requested[k]]
        if conflicts:
            _logger.warning(f"{requested.matrix_name} matrix is already saved with different signature. conflicts "
                            f"(key, saved value, requested value): {conflicts}. continuing with the saved one.")
        self.signature = loaded
        self.name = loaded.matrix_name
        self.symmetry_strings = loaded.symmetry_strings
        self.num_threads = num_threads
        if np.dtype(dtype) != np.dtype(VALUE_INDEX_TYPE):
            raise ValueError("value indices are int16 in the stored layout")
        self.vals, self.value_inds, self.minor_inds, self.major_pointers = read_matrix_folder(
            dataclasses.replace(loaded, folder_name=str(folder)))
        self.shape = list(loaded.shape) if loaded.shape else [len(self.major_pointers) - 1] * 2
        self.is_row_major = loaded.is_row_major
        self.is_sorted = "sorted" in self.symmetry_strings
        self.is_hermitian_triangle = loaded.is_hermitian_triangle
        self._device = None
        self._sums = {}

    @classmethod
    def from_name(cls, name: str, config, num_threads=None):
        """The matrix <calc folder>/<name> of a config (sparse_matrices.py:127-140).

        The reference finds the folder because computing the eigenpairs wrote it.  Here nothing is written
        as a side effect, so a missing Hamiltonian-term folder is built on the GPU and written first.
        """
        folder = config.get_calc_folder() / name
        if not (folder / FILE_NAMES["info_name"]).exists() and name in TERMS and hasattr(config, "weights"):
            from exactdiag_amd.tJ_spin_half_ladder.matrix_setup import write_hamiltonian_term

            write_hamiltonian_term(config, name)
        sig = read_signature(config.get_calc_folder() / name)
        sig = dataclasses.replace(sig, folder_name=str(config.get_calc_folder() / name))
        return cls(sig, num_threads=config.num_threads if num_threads is None else num_threads)

    @classmethod
    def from_device(cls, stored, signature):
        """Write a matrix held in HBM (``StoredCSR``) as a folder and return it opened."""
        write_matrix_folder(signature, *stored.arrays())
        return cls(signature)

    def device(self):
        """The matrix in HBM (uploaded once)."""
        if self._device is None:
            from exactdiag_amd.csr import StoredCSR

            self._device = StoredCSR.from_arrays(self.shape, self.is_row_major, self.is_hermitian_triangle,
                                                 np.asarray(self.vals), np.asarray(self.value_inds),
                                                 np.asarray(self.minor_inds), np.asarray(self.major_pointers))
        return self._device

    def _sum(self, weight):
        from exactdiag_amd.csr import CSRSum

        key = complex(weight)
        if key not in self._sums:
            self._sums[key] = CSRSum([self.device()], [key])
        return self._sums[key]

    def dot(self, vec, weight=1, matvec=None):
        """weight * M vec (+ matvec, which is updated and returned when given)."""
        if len(self.minor_inds) == 0 or abs(weight) < 1e-6:
            return np.zeros(self.shape[0], VALUE_TYPE) if matvec is None else matvec
        x, kind = _to_device(vec)
        y = self._sum(weight).hv(x.to(_torch_dtype()).contiguous())
        if matvec is None:
            return _from_device(y, kind)
        if kind == "numpy":
            matvec += y.cpu().numpy()
        else:
            matvec += y.to(matvec.device)
        return matvec

    def shift_dot(self, vin, shift):
        return self.dot(vin) - shift * vin

    def toarray(self):
        """Dense array with the reference's index convention (sparse_matrices.py:110-125); small matrices only."""
        mat = np.zeros(self.shape, VALUE_TYPE)
        vals = np.asarray(self.vals)
        for i in range(len(self.major_pointers) - 1):
            for j in range(self.major_pointers[i], self.major_pointers[i + 1]):
                u = self.minor_inds[j]
                mat[u, i] += vals[self.value_inds[j]]
                if self.is_hermitian_triangle and u != i:
                    mat[i, u] += np.conj(vals[self.value_inds[j]])
        return mat


def _torch_dtype():
    import torch

    return torch.complex128


class Added_Sparse_Matrices:
    """Weighted sum of stored matrices on the GPU (reference: general/sparse_matrices.py:210-243)."""

    def __init__(self, mats, weights=None):
        self.refs = list(mats)
        if any(list(m.shape) != list(self.refs[0].shape) for m in self.refs[1:]):
            raise ValueError("incompatible shapes")
        self.shape = self.refs[0].shape
        self.num_matrices = len(self.refs)
        if weights is None:
            weights = [1] * self.num_matrices
        elif len(weights) != self.num_matrices:
            raise ValueError("wrong weights supplied")
        self.weights = np.asarray(weights, VALUE_TYPE)
        self.num_threads = max(m.num_threads for m in self.refs)
        self._op = None

    def add_matrix(self, matrix, weight):
        if list(self.shape) != list(matrix.shape):
            raise ValueError(f"Nonmatching shapes {self.shape} and {matrix.shape}.")
        self.weights = np.append(self.weights, VALUE_TYPE(weight))
        self.num_matrices += 1
        self.refs.append(matrix)
        self._op = None

    def operator(self):
        from exactdiag_amd.csr import CSRSum

        if self._op is None:
            self._op = CSRSum([m.device() for m in self.refs], self.weights)
        return self._op

    def apply(self, x):
        return self.operator().hv(x)

    def dot(self, vec):
        x, kind = _to_device(vec)
        return _from_device(self.apply(x.to(_torch_dtype()).contiguous()), kind)

    def lanczos(self, v0, nsteps, precision=0):
        """Plain Lanczos coefficients over the stored matrices (complex128 only)."""
        from exactdiag_amd.general.krylov import plain_lanczos

        if precision:
            raise ValueError("the stored reference layout is applied in complex128 only")
        return plain_lanczos(self.apply, v0.to(_torch_dtype()), nsteps)

    def toarray(self):
        mat = np.zeros(self.shape, VALUE_TYPE)
        for w, m in zip(self.weights, self.refs):
            mat += w * m.toarray()
        return mat
