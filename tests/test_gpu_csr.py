"""The reference's stored matrices built on the GPU, read from folders, and applied there.

calc_sparse_matrix (general/cython_sparse_matrices.pyx:58-241) is restated in
the oracle (ed_oracle.cpp: column, dedup_append, to_csr); its products go
through the reference's own compiled sparse_matrix_dot in
tests/test_oracle_ref.py.  Here the GPU build must give the oracle's arrays
exactly (pointers, indices and value-table order bit for bit, values equal
as numbers), the folders
must hold them in the reference's file layout, and the GPU product of the
stored matrices must equal the oracle's CSR dot and the stored-H product.
"""

import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SECTORS = [
    (2, 0, 2, 0, 0), (2, 1, 1, 1, 1), (3, 1, 2, 0, 0), (3, 0, 3, 1, 1), (4, 2, 3, 1, 1), (4, 3, 2, 2, 0),
    (4, 0, 4, 2, 1), (5, 2, 4, 2, 0), (6, 2, 5, 0, 0), (6, 1, 5, 3, 1), (6, 4, 4, 0, 1), (7, 2, 6, 3, 0),
    (4, 2, 3, -1, 0), (6, 0, 6, 7, 0),
]
WEIGHTS = [[1.0, 1.0, 0.4, 0.4], [0.45, 0.36, 0.15, 0.12], [1.0 + 0.3j, 0.7 - 0.2j, 0.4j, 0.25], [0, 1, 0, 1]]


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


def _same(a, b):
    """Equal arrays: integers bit for bit, values exactly (a zero's sign may differ: the GPU's complex products
    can give -0.0 where the oracle's give +0.0, which no product or comparison can tell apart)."""
    a, b = np.asarray(a), np.asarray(b)
    return a.dtype == b.dtype and a.shape == b.shape and np.array_equal(a, b)


def _assert_csr(got, ref, what):
    names = ("values", "value_inds", "minor_inds", "major_pointers")
    for name, g, r in zip(names, got, ref):
        assert _same(g, r), (what, name, len(g), len(r))


def test_term_csr_bit_exact(oracle_mod, torch):
    from exactdiag_amd.csr import StoredCSR
    from exactdiag_amd.sector import Sector

    for p in SECTORS:
        print("sector", p, flush=True)
        s, o = Sector(*p), oracle_mod.OracleSector(*p)
        for term in range(4):
            m = StoredCSR.from_sector(s, term)
            assert m.shape == (s.dim, s.dim)
            _assert_csr(m.arrays(), o.term_csr(term), (p, term))
            m.close()


@pytest.mark.parametrize("triangle,row_major", [(False, True), (True, False), (False, False)])
def test_other_layouts_bit_exact(oracle_mod, torch, triangle, row_major):
    from exactdiag_amd.csr import StoredCSR
    from exactdiag_amd.sector import Sector

    for p in [(4, 2, 3, 1, 1), (6, 1, 5, 3, 1), (5, 2, 4, 2, 0)]:
        s, o = Sector(*p), oracle_mod.OracleSector(*p)
        for term in range(4):
            m = StoredCSR.from_sector(s, term, triangle_only=triangle, row_major=row_major)
            _assert_csr(m.arrays(), o.layout_csr(term, triangle, row_major), (p, term, triangle, row_major))


OP_CASES = [
    ((4, 2, 3, 0, 0), 0, (2, 1), (4, 2, 3, 2, 1)),
    ((6, 2, 5, 1, 0), 1, (2, 0), (6, 2, 5, 3, 0)),
    ((5, 1, 2, 0, 0), 2, (1, 0), (5, 0, 3, 1, 0)),
    ((5, 2, 4, 1, 1), 3, (2, 1), (5, 1, 4, -2, 0)),
    ((5, 1, 4, 0, 0), 4, (-1, 1), (5, 2, 3, -1, 1)),
    ((4, 1, 3, 0, 1), 5, (1, 0), (4, 2, 3, 1, 1)),
    ((5, 2, 4, 0, 0), 6, (1, 1), (5, 2, 4, 1, 1)),
    ((5, 2, 4, 0, 0), 7, (2, 0), (5, 2, 4, 2, 0)),
]


@pytest.mark.parametrize("case", OP_CASES)
def test_operator_csr_bit_exact(oracle_mod, torch, case):
    from exactdiag_amd.csr import CSRSum, StoredCSR
    from exactdiag_amd.sector import Sector

    pi, op, (qx, qy), pf = case
    print("operator case", case, flush=True)
    gi, gf = Sector(*pi), Sector(*pf)
    oi, of = oracle_mod.OracleSector(*pi), oracle_mod.OracleSector(*pf)
    m = StoredCSR.from_operator(gi, gf, op, qx, qy)
    assert m.shape == (gf.dim, gi.dim)
    ref = oi.op_csr(of, op, qx, qy)
    _assert_csr(m.arrays(), ref, case)
    # column-major product (cython_sparse_dot.pyx:82-100): y[minor] += value * x[major]
    values, vinds, minor, major = ref
    x = np.random.default_rng(4).normal(size=gi.dim) + 1j * np.random.default_rng(5).normal(size=gi.dim)
    y_ref = np.zeros(gf.dim, np.complex128)
    for i in range(gi.dim):
        for k in range(major[i], major[i + 1]):
            y_ref[minor[k]] += values[vinds[k]] * x[i]
    y = CSRSum([m], [1.0]).hv(torch.as_tensor(x, device="cuda")).cpu().numpy()
    assert np.max(np.abs(y - y_ref)) <= 1e-13 * max(1.0, np.max(np.abs(y_ref)))
    # the stored values are the first of each group within 1e-6 (the reference's unique-value table), so the
    # stored operator differs from the exact one by less than 1e-6 per element
    y_op = gi.op_apply(gf, op, qx, qy, torch.as_tensor(x, device="cuda")).cpu().numpy()
    assert np.max(np.abs(y - y_op)) <= 1e-6 * 2 * gi.N * np.max(np.abs(x))  # at most 2N elements per row


def test_sum_matches_oracle_dot(oracle_mod, torch):
    """Weighted sum over the stored terms = the oracle's restatement of Added_Sparse_Matrices.dot."""
    from exactdiag_amd.csr import CSRSum, StoredCSR
    from exactdiag_amd.sector import Sector

    for p in SECTORS:
        s, o = Sector(*p), oracle_mod.OracleSector(*p)
        if s.dim == 0:
            continue
        mats = [StoredCSR.from_sector(s, t) for t in range(4)]
        for wi, w in enumerate(WEIGHTS):
            x = np.random.default_rng(wi).normal(size=s.dim) + 1j * np.random.default_rng(wi + 7).normal(size=s.dim)
            ref = o.hv(w, x)
            y = CSRSum(mats, w).hv(torch.as_tensor(x, device="cuda")).cpu().numpy()
            assert np.max(np.abs(y - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref))), (p, wi)


def test_sum_skips_small_weights_and_checks_shapes(torch):
    from exactdiag_amd.csr import CSRSum, StoredCSR
    from exactdiag_amd.sector import Sector

    s = Sector(4, 2, 3, 1, 1)
    mats = [StoredCSR.from_sector(s, t) for t in range(4)]
    x = torch.randn(s.dim, dtype=torch.complex128, device="cuda")
    y_all = CSRSum(mats, [1, 1, 0.4, 0.4]).hv(x)
    assert torch.equal(CSRSum(mats, [1, 1, 0.4, 0.4]).hv(x), y_all)  # deterministic
    y_skip = CSRSum(mats, [1, 1, 0.4, 5e-7]).hv(x)  # |w| < 1e-6 leaves the term out
    y_three = CSRSum(mats[:3], [1, 1, 0.4]).hv(x)
    assert float((y_skip - y_three).abs().max()) <= 1e-14 * float(y_three.abs().max())
    other = StoredCSR.from_sector(Sector(5, 2, 4, 0, 0), 0)
    with pytest.raises(ValueError):
        CSRSum([mats[0], other], [1, 1])


def test_uploaded_arrays_are_checked(torch):
    from exactdiag_amd.csr import StoredCSR

    vals = np.array([1.0 + 0j, 2.0])
    ok = StoredCSR.from_arrays((3, 3), True, True, vals, [0, 1, 0], [0, 2, 1], [0, 2, 2, 3])
    assert ok.nnz == 3 and ok.max_count == 2
    bad_cases = [
        dict(vinds=[0, 2, 0], minor=[0, 2, 1], major=[0, 2, 2, 3]),  # value index out of range
        dict(vinds=[0, 1, 0], minor=[0, 3, 1], major=[0, 2, 2, 3]),  # column out of range
        dict(vinds=[0, 1, 0], minor=[0, 2, 1], major=[0, 2, 1, 3]),  # decreasing pointers
        dict(vinds=[0, 1, 0], minor=[0, 2, 1], major=[1, 2, 2, 3]),  # pointers do not start at 0
    ]
    for c in bad_cases:
        with pytest.raises(ValueError):
            StoredCSR.from_arrays((3, 3), True, True, vals, c["vinds"], c["minor"], c["major"])


def test_folders_round_trip_and_match_stored_h(tmp_path, monkeypatch, oracle_mod, torch):
    """get_hamiltonian_matrices writes the reference's folders; reopened, they give the stored-H product."""
    from exactdiag_amd.general.sparse_matrices import FILE_NAMES, Sparse_Matrix
    from exactdiag_amd.tJ_spin_half_ladder import configs, matrix_setup

    monkeypatch.chdir(tmp_path)
    ham = configs.Hamiltonian_Config(num_rungs=6, num_holes=2, weights={"tl": 0.45, "tr": 0.36, "jl": 0.15,
                                                                         "jr": 0.12},
                                     symmetry_qs={"leg": 1, "rung": 1})
    added = matrix_setup.get_hamiltonian_matrices(ham)
    o = oracle_mod.OracleSector(6, 2, 5, 1, 1)
    for t, name in enumerate(("tl", "tr", "jl", "jr")):
        folder = ham.get_calc_folder() / name
        info = json.loads((folder / FILE_NAMES["info_name"]).read_text())
        assert list(info) == ["matrix_name", "shape", "symmetry_strings", "max_values_per_column", "major",
                              "values_name", "minor_inds_name", "value_inds_name", "major_pointers_name",
                              "folder_name", "info_name", "initial_system_info", "final_system_info"]
        assert info["shape"] == [o.dim, o.dim] and info["major"] == "row"
        assert info["symmetry_strings"] == ["hermitian", "sorted", "triangle_only"]
        values, vinds, minor, major = o.term_csr(t)
        # raw arrays, no header (np.memmap files)
        assert (folder / "values.npy").stat().st_size == 16 * len(values)
        assert _same(np.fromfile(folder / "values.npy", "<c16"), values)
        assert _same(np.fromfile(folder / "value_inds.npy", "<i2"), vinds)
        assert _same(np.fromfile(folder / "minor_inds.npy", "<i8"), minor)
        assert _same(np.fromfile(folder / "major_pointers.npy", "<i8"), major)
    x = np.random.default_rng(3).normal(size=o.dim) + 1j * np.random.default_rng(4).normal(size=o.dim)
    ref = o.hv([0.45, 0.36, 0.15, 0.12], x)
    for mats in (added, matrix_setup.get_hamiltonian_matrices(ham)):  # built, then reopened
        y = mats.dot(x)
        assert np.max(np.abs(y - ref)) <= 1e-13 * np.max(np.abs(ref))
    y_h = ham.setup()().dot(x)
    assert np.max(np.abs(added.dot(x) - y_h)) <= 1e-12 * np.max(np.abs(ref))
    # one term, with a weight and an accumulator, as Sparse_Matrix.dot
    jl = Sparse_Matrix.from_name("jl", ham)
    acc = np.ones(o.dim, np.complex128)
    out = jl.dot(x, 0.15, acc)
    assert out is acc
    assert np.max(np.abs(acc - 1 - 0.15 * o.hv([0, 0, 1, 0], x))) <= 1e-13 * np.max(np.abs(ref))
    # dense form in the reference's convention
    small = configs.Hamiltonian_Config(num_rungs=3, num_holes=1, weights={"tl": 1, "tr": 1, "jl": 0.4, "jr": 0.4},
                                       symmetry_qs={"leg": 0, "rung": 0})
    dense = matrix_setup.get_hamiltonian_matrices(small).toarray()
    assert np.max(np.abs(dense - oracle_mod.OracleSector(3, 1, 2, 0, 0).toarray([1, 1, 0.4, 0.4]))) < 1e-14


def test_lanczos_over_folders_matches_stored_h(tmp_path, monkeypatch, torch):
    from exactdiag_amd.tJ_spin_half_ladder import configs, matrix_setup

    monkeypatch.chdir(tmp_path)
    ham = configs.Hamiltonian_Config(num_rungs=7, num_holes=2, weights={"tl": 1, "tr": 1, "jl": 0.4, "jr": 0.4},
                                     symmetry_qs={"leg": 3, "rung": 0})
    added = matrix_setup.get_hamiltonian_matrices(ham)
    op = ham.setup()()
    v0 = torch.as_tensor(np.random.default_rng(9).normal(size=op.shape[0]) + 0j, device="cuda")
    a1, b1 = added.lanczos(v0, 60)
    a2, b2 = op.lanczos(v0, 60)
    assert len(a1) == len(a2) == 60
    np.testing.assert_allclose(a1, a2, rtol=0, atol=1e-10)
    np.testing.assert_allclose(b1, b2, rtol=0, atol=1e-10)


def test_reference_load_matrix_example_flow(tmp_path, monkeypatch, torch):
    """examples/tJ_spin_half_ladder/load_matrix.py of the reference: eigenpairs, then the term matrices by
    name summed with the config's weights; |E0 gs - H gs| must be small."""
    from exactdiag_amd.general.sparse_matrices import Added_Sparse_Matrices, Sparse_Matrix
    from exactdiag_amd.tJ_spin_half_ladder import api, configs

    monkeypatch.chdir(tmp_path)
    config = configs.Eigenpair_Config(
        hamiltonian={"num_rungs": 5, "num_holes": 2, "weights": {"tl": 0.45, "tr": 0.36, "jl": 0.15, "jr": 0.12},
                     "symmetry_qs": {"leg": 2, "rung": 1}, "num_threads": 8},
        eigenpair={"num_eigenpairs": 5})
    eigvals, eigvecs = api.get_eigenpairs(config)
    gs = eigvecs[:, 0]
    weights = {n: getattr(config.hamiltonian.weights, n) for n in ("tl", "tr", "jl", "jr")}
    matrices = Added_Sparse_Matrices(*zip(*[[Sparse_Matrix.from_name(name=n, config=config.hamiltonian), w]
                                            for n, w in weights.items()]))
    trial = matrices.dot(gs)
    assert np.linalg.norm(eigvals[0] * gs - trial) < 1e-6


def test_excitation_operator_folder(tmp_path, monkeypatch, oracle_mod, torch):
    from exactdiag_amd.tJ_spin_half_ladder import configs, matrix_setup

    monkeypatch.chdir(tmp_path)
    cfg = configs.Limited_Spectrum_Config(
        hamiltonian={"num_rungs": 5, "num_holes": 2, "weights": {"tl": 1, "tr": 1, "jl": 0.4, "jr": 0.4},
                     "symmetry_qs": {"leg": 0, "rung": 0}},
        spectrum={"name": "Szq", "operator_symmetry_qs": {"leg": 2, "rung": 1}, "broadening": 0.05,
                  "omega_max": 4.0},
    )
    m = matrix_setup.get_excitation_operator_matrix(cfg)
    sig = m.signature
    assert sig.matrix_name == "Szq_[2 1]" and sig.major == "column" and sig.symmetry_strings == ("sorted",)
    oi, of = oracle_mod.OracleSector(5, 2, 4, 0, 0), oracle_mod.OracleSector(5, 2, 4, 2, 1)
    ref = oi.op_csr(of, 0, 2, 1)
    got = (np.asarray(m.vals), np.asarray(m.value_inds), np.asarray(m.minor_inds), np.asarray(m.major_pointers))
    _assert_csr(got, ref, "Szq folder")
    assert list(m.shape) == [of.dim, oi.dim]
    again = matrix_setup.get_excitation_operator_matrix(cfg)  # reopened from the folder
    x = np.random.default_rng(2).normal(size=oi.dim) + 0j
    assert np.array_equal(again.dot(x), m.dot(x))


def test_large_sector_prefix_and_table_invariants(oracle_mod, torch):
    """2x10, 2 holes, k=(3, pi) (D = 461,890): the first 20,000 rows equal the oracle's build of those rows;
    over the whole matrix the value table keeps the reference's walk: entries pairwise at least 1e-6 apart,
    index u first used after index u - 1, every element within 1e-6 of its entry."""
    from exactdiag_amd.csr import StoredCSR
    from exactdiag_amd.sector import Sector

    p = (10, 2, 9, 3, 1)
    s, o = Sector(*p), oracle_mod.OracleSector(*p)
    rows = 20000
    for t in range(4):
        print("term", t, flush=True)
        vg, ig, mg, pg = StoredCSR.from_sector(s, t).arrays()
        o.build_csr_rows(t, 0, rows)
        vo, io, mo, po = o.term_csr(t)
        if po[-1] == 1 and not po[:-1].any() and len(vo) == 1 and vo[0] == 0:
            assert pg[rows] == 0  # no elements in these rows: the oracle wrote the empty-matrix marker
        else:
            n = int(po[-1])
            assert _same(pg[: rows + 1], po) and _same(mg[:n], mo) and _same(ig[:n], io)
            assert _same(vg[: len(vo)], vo)
        assert np.all(np.abs(vg[:, None] - vg[None, :])[np.triu_indices(len(vg), 1)] >= 1e-6)
        u, first = np.unique(ig, return_index=True)
        assert np.array_equal(u, np.arange(len(vg))) and np.all(np.diff(first) > 0)
        assert np.all(pg[1:] >= pg[:-1]) and pg[-1] == len(mg)
        rows_of = np.repeat(np.arange(s.dim), np.diff(pg))
        assert np.all(mg >= rows_of)  # upper triangle


@pytest.mark.parametrize("p", [(2, 2, 0, 0, 0), (3, 0, 0, 1, 0), (2, 4, 0, 0, 0), (3, 6, 0, 0, 0)])
def test_empty_and_tiny_sectors(oracle_mod, torch, p):
    """Sectors without states or with one state: the stored arrays are the oracle's (the empty-matrix marker
    included), the summed product has the sector's size, and position correlations of any state are finite."""
    from exactdiag_amd.csr import CSRSum, StoredCSR
    from exactdiag_amd.sector import Sector

    s, o = Sector(*p), oracle_mod.OracleSector(*p)
    assert s.dim == o.dim
    mats = []
    for t in range(4):
        m = StoredCSR.from_sector(s, t)
        _assert_csr(m.arrays(), o.term_csr(t), (p, t))
        mats.append(m)
    op = CSRSum(mats, [1, 1, 0.4, 0.4])
    x = torch.ones(s.dim, dtype=torch.complex128, device="cuda")
    y = op.hv(x)
    assert y.shape == (s.dim,)
    if s.dim:
        np.testing.assert_allclose(y.cpu().numpy(), o.hv([1, 1, 0.4, 0.4], x.cpu().numpy()), atol=1e-13)
        corr = s.position_correlations(0, np.zeros((0, 2), np.int32), x / x.norm())
        assert np.all(np.isfinite(corr))
