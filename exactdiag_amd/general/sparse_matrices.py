"""GPU operators with the reference's matrix interface (reference: general/sparse_matrices.py).

The reference stores each Hamiltonian term as a CSR on disk and sums them in
``Added_Sparse_Matrices.dot`` (general/sparse_matrices.py:210-243).  Here the
Hamiltonian is applied matrix-free by the HIP kernel; the objects keep the
``.dot``, ``.shape``, ``.weights``, ``.toarray()`` surface.  ``dot`` accepts
numpy arrays (copied to and from the device) or device tensors (kept there).
"""

from __future__ import annotations

import numpy as np

from exactdiag_amd.sector import TERMS, Sector, weights_array

VALUE_TYPE = np.complex128


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

    def dot(self, vec):
        if self.shape[0] == 0:
            return np.zeros(0, VALUE_TYPE)
        x, kind = _to_device(vec)
        if x.dtype not in (np.complex64, np.complex128):
            import torch

            if x.dtype not in (torch.complex128, torch.complex64):
                x = x.to(torch.complex128)
        y = self.sector.hv(self._w, x)
        return _from_device(y, kind)

    def dot_with_inner(self, x):
        """(H x, <x, H x>) with the inner product fused into the kernel."""
        return self.sector.hv(self._w, x, with_dot=True)

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
