"""Sectors with the spin flip (optional, no counterpart in the reference) against the brute force.

The reference's ladder group is leg translations x leg mirror (tJ_spin_half_ladder/symmetry_ladder.h:55-70);
ed_sector_create_flip adds the global spin flip with parity z.  Its bases, Hamiltonian terms and operators are
pinned here by the second-quantised brute force of tests/fock.py (symmetric states from the projector over the
extended group) on tiny ladders, and at larger sizes by the physics the two groups share: the z = +-1 blocks
together hold the momentum sector's spectrum, and a spectrum computed in the flip blocks equals the one computed
in the momentum sector.
"""

import numpy as np
import pytest

from fock import Fock, _fock_operator

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


def _dense(torch, apply, D):
    eye = torch.eye(D, dtype=torch.complex128, device="cuda")
    return np.stack([apply(eye[:, k].contiguous()).cpu().numpy() for k in range(D)], axis=1)


@pytest.mark.parametrize("L,h,nd", [(3, 0, 3), (3, 2, 2), (4, 2, 3), (4, 0, 4), (5, 2, 4)])
def test_flip_basis_and_hamiltonian_against_brute_force(torch, oracle_mod, L, h, nd):
    from exactdiag_amd.sector import Sector, weights_array

    fk = Fock(L, h, nd)
    ranker = oracle_mod.OracleSector(L, h, nd, 0, 0)
    checked = 0
    for kx in range(L):
        for ky in (0, 1):
            for z in (1, -1):
                reps, B, sizes = fk.symmetric_basis(kx, ky, z, orbit_sizes=True)
                s = Sector(L, h, nd, kx, ky, flip=z)
                print("flip sector", (L, h, nd, kx, ky, z), "dim", s.dim, flush=True)
                assert s.dim == len(reps)
                if s.dim == 0:
                    continue
                sp, orb = s.basis()
                assert list(sp) == [ranker.rank(np.array(r, np.int16)) for r in reps]
                assert list(orb) == sizes
                for t in range(4):
                    w = np.zeros(4)
                    w[t] = 1.0
                    Hs = B.conj().T @ fk.hamiltonian(w) @ B
                    wa = weights_array(w)
                    M = _dense(torch, lambda v: s.hv(wa, v), s.dim)
                    assert np.max(np.abs(M - Hs)) < 1e-12, (t, kx, ky, z)
                w = [1.0, 0.7, 0.4, 0.3]
                Hs = B.conj().T @ fk.hamiltonian(w) @ B
                m = s.store(w)
                x = np.random.default_rng(kx).normal(size=s.dim) + 1j * np.random.default_rng(ky).normal(size=s.dim)
                y = m.hv(torch.as_tensor(x, device="cuda")).cpu().numpy()
                assert np.max(np.abs(y - Hs @ x)) < 1e-12 * max(1.0, np.max(np.abs(Hs @ x)))
                m.close()
                checked += 1
    assert checked >= 2 * L


OPS = [("Szq", (1, 1), -1), ("Nq", (2, 0), 1), ("leg", (1, 1), 1), ("rung", (1, 0), 1)]


@pytest.mark.parametrize("L,h,nd", [(4, 2, 3), (3, 2, 2)])
@pytest.mark.parametrize("name,q,parity", OPS)
def test_flip_operators_against_brute_force(torch, L, h, nd, name, q, parity):
    """Szq takes parity z to -z, N(q) and the currents keep it (the rung current is called with qy + 1 and ends
    in ky + qy + 1, as setup_excitation_operator does)."""
    from exactdiag_amd import _lib
    from exactdiag_amd.allq import final_momentum
    from exactdiag_amd.sector import Sector

    op = {"Szq": _lib.OP_SZQ, "Nq": _lib.OP_NQ, "leg": _lib.OP_CURRENT_LEG, "rung": _lib.OP_CURRENT_RUNG}[name]
    qx, qy = q
    qy_call = qy + 1 if name == "rung" else qy
    fk = Fock(L, h, nd)
    n_checked = 0
    for kx in range(L):
        for ky in (0, 1):
            for z in (1, -1):
                kf = final_momentum(L, kx, ky, qx, qy_call)
                _, Bi = fk.symmetric_basis(kx, ky, z)
                _, Bf = fk.symmetric_basis(kf[0], kf[1], parity * z)
                if Bi.shape[1] == 0 or Bf.shape[1] == 0:
                    continue
                Os = Bf.conj().T @ _fock_operator(fk, fk, op, qx, qy) @ Bi
                si, sf = Sector(L, h, nd, kx, ky, flip=z), Sector(L, h, nd, *kf, flip=parity * z)
                ptr, rows, vals = si.op_columns(sf, op, qx, qy_call)
                M = np.zeros((sf.dim, si.dim), np.complex128)
                for i in range(si.dim):
                    for k in range(ptr[i], ptr[i + 1]):
                        M[rows[k], i] += vals[k]
                # the reference's single-precision q.r in the weights (about 1e-7)
                assert np.max(np.abs(M - Os)) < 2e-6, (kx, ky, z)
                x = np.random.default_rng(3).normal(size=si.dim) + 0.5j
                y = si.op_apply(sf, op, qx, qy_call, torch.as_tensor(x, device="cuda")).cpu().numpy()
                assert np.max(np.abs(y - M @ x)) < 1e-12 * max(1.0, np.max(np.abs(M @ x)))
                n_checked += np.max(np.abs(Os)) > 1e-3
    assert n_checked >= 2
    # an operator that changes S^z, or the wrong final parity, is refused
    si, sf = Sector(L, h, nd, 0, 0, flip=1), Sector(L, h, nd, *final_momentum(L, 0, 0, qx, qy_call), flip=-parity)
    with pytest.raises(ValueError, match="spin-flip parity"):
        si.op_columns(sf, op, qx, qy_call)


def test_flip_blocks_hold_the_sector_spectrum_and_its_spectra(torch):
    """2x8 ladder, 2 holes: the z = +-1 blocks of k = (0, 0) split its dimension, the lower block's ground state
    is the sector's, and S(q = (pi, pi), w) computed in the flip blocks equals the momentum sector's."""
    from exactdiag_amd import _lib
    from exactdiag_amd.general import krylov
    from exactdiag_amd.sector import Sector, cf_eval

    L, h, nd, w = 8, 2, 7, [1.0, 1.0, 0.4, 0.4]
    s0 = Sector(L, h, nd, 0, 0)
    blocks = {z: Sector(L, h, nd, 0, 0, flip=z) for z in (1, -1)}
    assert blocks[1].dim + blocks[-1].dim == s0.dim and abs(blocks[1].dim - blocks[-1].dim) < 0.01 * s0.dim
    gs = {}
    for key, s in [("all", s0)] + list(blocks.items()):
        m = s.store(w)
        vals, vecs, _ = krylov.lowest_eigenpairs(m.hv, s.dim, 2, tol=1e-11)
        gs[key] = (vals, vecs[0], s, m)
        print(key, "dim", s.dim, "lowest", vals, flush=True)
    e_blocks = np.sort(np.concatenate([gs[1][0], gs[-1][0]]))
    assert abs(gs["all"][0][0] - e_blocks[0]) < 1e-10
    z0 = 1 if gs[1][0][0] < gs[-1][0][0] else -1
    ws = np.linspace(0, 4, 2001)
    q = (4, 1)
    runs = []
    for sec, zf, e0, psi in ((s0, 0, gs["all"][0][0], gs["all"][1]), (blocks[z0], -z0, gs[z0][0][0], gs[z0][1])):
        fin = Sector(L, h, nd, 4, 1, flip=zf)
        phi = sec.op_apply(fin, _lib.OP_SZQ, *q, psi)
        n2 = krylov.vec_norm(phi) ** 2
        mf = fin.store(w)
        a, b = mf.lanczos(phi, 300)
        runs.append((n2, a, b, cf_eval(a, b, ws, e0, 0.05, -n2 / np.pi)))
        print("final dim", fin.dim, "|phi|^2", n2, "steps", len(a), flush=True)
    (n2a, aa, ba, sa), (n2b, ab, bb, sb) = runs
    assert abs(n2a - n2b) < 1e-9
    # the same recurrence in another basis: the coefficients agree until orthogonality is lost, the spectra
    # to the truncation level of the fraction
    # (the two ground states agree to their tolerance, 1e-11, and the recurrences amplify the difference by
    # orders of magnitude within a few tens of steps: about 1e-13 at step 5, 1e-9 near step 30)
    assert np.max(np.abs(aa[:20] - ab[:20])) < 1e-9 and np.max(np.abs(ba[:20] - bb[:20])) < 1e-9
    dev = np.max(np.abs(sa - sb)) / np.max(sa)
    print("max |S - S_flip| / max S", dev, flush=True)
    assert dev < 1e-3
