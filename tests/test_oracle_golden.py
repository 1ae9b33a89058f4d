"""The C oracle against the golden fixtures of the independent numpy restatement
(tests/golden/make_golden.py) and against closed forms.  CPU only."""
import os

import numpy as np
import pytest

import oracle as O

CASES = ["cdl_tiny", "cdl_small", "cdl_mixedfl", "cdl_long"]


def _rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(np.asarray(a, np.float64) - b) / (nb if nb > 0 else 1.0)


def _run(oc, sched):
    outer, ncsc, ndict = sched
    objs, imps = [], []
    for _ in range(outer):
        oc.code_update(int(ncsc))
        o, imp = oc.update_best()
        objs.append(o)
        imps.append(imp)
        oc.dict_update(int(ndict))
    return np.array(objs), np.array(imps)


@pytest.mark.parametrize("case", CASES)
def test_oracle_f64_matches_numpy_restatement(case, golden_dir):
    z = np.load(os.path.join(golden_dir, case + ".npz"))
    rho, sigma, lam = z["params"]
    oc = O.OracleCDL(z["s"], z["D0"], z["X0"], z["fl"], rho, sigma, lam, "f64")
    objs, imps = _run(oc, z["sched"])
    np.testing.assert_allclose(objs, z["obj"], rtol=1e-10)
    assert (imps == z["improved"]).all()
    for f in ("X", "DX", "d", "G1", "G3", "Xbest", "Dbest", "y1", "y2", "y3", "Y1", "g1", "g2", "g3", "Td"):
        assert _rel(oc.get(f), z[f]) < 1e-10, f
    th, cnt = oc.code_thresh(1.0)
    assert cnt == int(z["thresh_count"])
    assert abs(th - float(z["thresh"])) <= 1e-6 * abs(float(z["thresh"]))   # th is stored via REAL(sqrt)


@pytest.mark.parametrize("case", CASES)
def test_oracle_f32_close_to_f64(case, golden_dir):
    z = np.load(os.path.join(golden_dir, case + ".npz"))
    rho, sigma, lam = z["params"]
    oc = O.OracleCDL(z["s"], z["D0"], z["X0"], z["fl"], rho, sigma, lam, "f32")
    objs, _ = _run(oc, z["sched"])
    np.testing.assert_allclose(objs, z["obj"], rtol=1e-4)
    for f in ("X", "DX", "G1", "G3"):
        assert _rel(oc.get(f), z[f]) < 1e-4, f


@pytest.mark.parametrize("n", [1, 2, 3, 7, 24, 104, 432, 440, 632, 840, 848, 1240, 2032, 2040, 2048])
def test_oracle_fft_is_the_dft(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    if n <= 128:   # the DFT definition, directly
        k = np.arange(n)
        ref = np.exp(-2j * np.pi * np.outer(k, k) / n) @ x
    else:
        ref = np.fft.fft(x)
    got = O.fft(x)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max() * max(1, np.log2(n))
    back = O.fft(got, inverse=True)
    assert np.abs(back - x).max() <= 1e-12 * max(1, np.log2(n))
    # Parseval and Hermitian symmetry of a real input's spectrum
    xr = rng.standard_normal(n)
    X = O.fft(xr)
    assert abs(np.sum(np.abs(X) ** 2) / n - np.sum(xr ** 2)) < 1e-9 * np.sum(xr ** 2)
    if n > 1:
        assert np.abs(X[1:] - np.conj(X[1:][::-1])).max() < 1e-10 * max(1.0, np.abs(X).max())


def test_csc_solve_is_the_dense_solve():
    """One CSC step from a hand-made state: X_hat must equal solve(conj(g) g^T + 2I, R)
    per frequency (Sherman-Morrison, csc_kernels.jl:19-49) -- checked through the spectra
    of X that the oracle produces.  Uses the FP64 oracle on a tiny problem."""
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "cdl_tiny.npz"))
    rho, sigma, lam = z["params"]
    oc = O.OracleCDL(z["s"], z["D0"], z["X0"], z["fl"], rho, sigma, lam, "f64")
    # state before the step
    X, DX, Y1 = oc.get("X"), oc.get("DX"), oc.get("Y1")
    y1, y2, y3, Y3 = oc.get("y1"), oc.get("y2"), oc.get("y3"), oc.get("Y3")
    G1 = oc.get("G1")
    oc.code_update(1)
    Xn = oc.get("X")
    t = X + y2
    Y2 = np.sign(t) * np.maximum(np.abs(t) - lam, 0)
    y1n, y2n, y3n = y1 + DX - Y1, y2 + X - Y2, y3 + X - Y3
    F = lambda a: np.fft.fft(a, axis=-1)
    G = F(G1)                                   # (M, Lg)
    N, M, Lg = X.shape
    for n in range(N):
        R = np.conj(G) * F(Y1[n] - y1n[n])[None] + F(Y2[n] - y2n[n]) + F(Y3[n] - y3n[n])
        for l in range(Lg):
            g = G[:, l]
            A = np.outer(np.conj(g), g) + 2.0 * np.eye(M)
            sol = np.linalg.solve(A, R[:, l])
            got = F(Xn[n])[:, l]
            assert np.allclose(got, sol, atol=1e-10 * max(1, np.abs(sol).max()))


def test_dict_combine_is_twice_the_solve():
    """Quirk Q3 (cdl_kernels.jl:24-33): d_hat equals 2x the Sherman-Morrison solution of
    (conj(y) y^T + 2I) d = R'."""
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "cdl_tiny.npz"))
    rho, sigma, lam = z["params"]
    oc = O.OracleCDL(z["s"], z["D0"], z["X0"], z["fl"], rho, sigma, lam, "f64")
    oc.code_update(2)
    Y3, G1, G3 = oc.get("Y3"), oc.get("G1"), oc.get("G3")
    G2, g2, g1, g3 = oc.get("G2"), oc.get("g2"), oc.get("g1"), oc.get("g3")
    oc.dict_update(1)
    d = oc.get("d")
    F = lambda a: np.fft.fft(a, axis=-1)
    N, M, Lg = Y3.shape
    for n in range(N):
        Yh = F(Y3[n])
        R = np.conj(Yh) * F(G2[n] - g2[n])[None] + (F(G1) - F(g1[n])) + (F(G3) - F(g3[n]))
        dh = F(d[n])
        for l in range(Lg):
            y = Yh[:, l]
            A = np.outer(np.conj(y), y) + 2.0 * np.eye(M)
            sol = np.linalg.solve(A, R[:, l])
            assert np.allclose(dh[:, l], 2 * sol, atol=1e-9 * max(1, np.abs(sol).max()))


def test_prox_simplex_quirk_rows():
    """Q4 (cdl_kernels.jl:59-60): with aligned 4-groups summing to 1, rows with 0-based
    l % 4 != 3 are unchanged and each l % 4 == 3 row moves by -0.25*(next group's sum - 1)."""
    rng = np.random.default_rng(5)
    M, Lg, fl = 2, 32, 12
    P = np.zeros((M, Lg)); P[:, :fl] = 1
    v = rng.random((M, Lg))
    v.reshape(M, -1, 4)[:] /= v.reshape(M, -1, 4).sum(axis=2, keepdims=True)   # groups sum to 1
    v[:, 12:16] *= 1.5                                                         # group 3 sums to 1.5
    v[:, 8:12] *= 0.5                                                          # group 2 sums to 0.5
    out = O.prox_simplex(v, P, "f64")
    exp = v.copy()
    exp[:, fl:] = 0
    # row 3 (0-based) uses group 1 (sum 1) -> unchanged; row 7 uses group 2 (sum 0.5) -> +0.125;
    # row 11 uses group 3 (sum 1.5) -> -0.125; rows 8..10 use their own group 2 -> +0.125.
    exp[:, 7] += 0.125
    exp[:, 8:11] += 0.125
    exp[:, 11] -= 0.125
    np.testing.assert_allclose(out, exp, atol=1e-12)
    # FP32 form: group sum in float32, combination in float64, stored as float32
    o32 = O.prox_simplex(v.astype(np.float32), P.astype(np.float32), "f32")
    np.testing.assert_allclose(o32, exp, atol=1e-6)


def test_box_projection():
    P = np.array([[1, 1, 1, 1, 0, 0]], float)
    v = np.array([[-0.5, 0.0, 0.3, 1.7, 0.4, 2.0]])
    out = O.box(v, P, "f64")
    np.testing.assert_array_equal(out, [[0, 0, 0.3, 1.0, 0, 0]])


C2_SLICE = "c2_slice.npz"


def test_c2_slice_fixture_is_consistent(golden_dir):
    """The 1000-sequence, 200-iteration oracle fixture: its inputs rebuild bit for bit from the
    generators and seeds in make_c2_slice.py (the GPU test rebuilds them the same way), the FP32
    and FP64 oracle objective traces agree to 1e-4, and the stored filters satisfy the
    projections (G1, G3 vanish off the support of length 20, G3 in [0, 1])."""
    path = os.path.join(golden_dir, C2_SLICE)
    if not os.path.exists(path):
        pytest.skip("c2_slice.npz not generated")
    import sys
    sys.path.insert(0, golden_dir)
    import make_c2_slice as mk
    z = np.load(path)
    s, D0, X0 = mk.inputs()
    for nm, a in (("s", s), ("D0", D0), ("X0", X0)):
        assert mk.digest(a) == str(z["digest_" + nm]), nm
    assert z["trace_f64"].shape == (mk.ITER_NUM,)
    np.testing.assert_allclose(z["trace_f32"], z["trace_f64"], rtol=1e-4)
    for p in ("f32", "f64"):
        G1, G3 = z["G1_" + p], z["G3_" + p]
        assert G1.shape == (mk.M, 4 * 200 + 2 * mk.FLEN)
        assert np.all(G1[:, mk.FLEN:] == 0) and np.all(G3[:, mk.FLEN:] == 0)
        assert G3.min() >= 0 and G3.max() <= 1
        assert z["DX_" + p].shape == (mk.N_DX, G1.shape[1])


def test_g1_minus_g3_is_the_same_for_every_sequence():
    """The dictionary duals g1 and g3 (train.jl:70,72) start at 0 and receive the same d, so
    g1 - g3 depends on the filters only and is equal across sequences; the library holds it
    once per filter.  Checked on the FP64 oracle, which keeps both per sequence as the reference
    does, after three outer iterations."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    from make_golden import make_cdl_inputs
    fl = [8, 12, 8]
    s, D0, X0, _ = make_cdl_inputs(4, 6, 20, 3, fl)
    oc = O.OracleCDL(s, D0, X0, fl, 1.0, 100.0, 0.05, "f64")
    _run(oc, (3, 4, 2))
    diff = oc.get("g1") - oc.get("g3")          # (N, M, Lg)
    assert np.abs(diff).max() > 1e-3            # not trivially zero
    np.testing.assert_allclose(diff, np.broadcast_to(diff[:1], diff.shape), rtol=0, atol=1e-12 * np.abs(diff).max())
