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

import numpy as np

from exactdiag_amd.sector import TERMS, Sector, weights_array

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
        """Plain Lanczos coefficients (a, b) from the device vector v0."""
        m = self.stored
        if m is not None:
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
