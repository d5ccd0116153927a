"""LGL basis data and the reference's file formats, pinned by closed forms and numpy.polynomial."""
import numpy as np
import numpy.polynomial.legendre as npleg
import pytest

from helpers import dg3d, problem


@pytest.mark.parametrize("N", range(1, 9))
def test_lgl_nodes_weights(N):
    x, w, M, S, D = dg3d.basis_generate(N)
    c = np.zeros(N + 1)
    c[N] = 1
    interior = np.sort(npleg.legroots(npleg.legder(c)))
    ref = np.concatenate([[-1.0], interior, [1.0]])
    assert np.abs(x - ref).max() < 5e-15
    # LGL quadrature is exact for degree <= 2N-1
    for k in range(2 * N):
        exact = (1 - (-1) ** (k + 1)) / (k + 1)
        assert abs(np.dot(w, x ** k) - exact) < 1e-13


def lagrange_poly(x, j):
    n = len(x)
    others = [x[m] for m in range(n) if m != j]
    p = np.polynomial.Polynomial.fromroots(others)
    return p / p(x[j])


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 7])
def test_mass_and_stiffness_closed_forms(N):
    x, w, M, S, D = dg3d.basis_generate(N)
    n = N + 1
    L = [lagrange_poly(x, j) for j in range(n)]
    Mex = np.zeros((n, n))
    Sex = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            pm = (L[i] * L[j]).integ()
            Mex[i, j] = 0.5 * (pm(1) - pm(-1))  # unit interval = half of [-1,1]
            ps = (L[i].deriv() * L[j]).integ()
            Sex[i, j] = ps(1) - ps(-1)
    assert np.abs(M - Mex).max() < 1e-13  # monomial-basis integration loses ~1e-14 at N=7
    assert np.abs(S - Sex).max() < 3e-13  # (degree-2N monomial coefficients lose ~1e-13 at N=7)
    assert abs(M.sum() - 1.0) < 1e-14
    e = np.zeros(n)
    e[0], e[-1] = -1.0, 1.0
    assert np.abs(S.sum(axis=1) - e).max() < 1e-13       # int l_i' = l_i(1) - l_i(-1)
    assert np.abs(S + S.T - np.diag(e)).max() < 1e-13    # summation by parts
    assert np.all(np.linalg.eigvalsh(M) > 0)


@pytest.mark.parametrize("N", [1, 3, 4, 7])
def test_line_operator_differentiates_polynomials(N):
    """M^{-1}(S f - (f_N e_N - f_0 e_0)) = -f' on the unit interval for deg f <= N."""
    x, w, M, S, D = dg3d.basis_generate(N)
    xu = x / 2 + 0.5
    Mi = np.linalg.inv(M)
    for k in range(N + 1):
        f = xu ** k
        df = k * xu ** (k - 1) if k else np.zeros_like(xu)
        bnd = np.zeros(N + 1)
        bnd[0], bnd[-1] = -f[0], f[-1]
        assert np.abs(Mi @ (S @ f - bnd) + df).max() < 1e-11 * max(1, k) ** 2


def test_files_round_trip(tmp_path):
    dg3d.basis_write_files(tmp_path, 8)
    for N in range(1, 9):
        x, w, M, S = dg3d.basis_read_files(tmp_path / "hardcoded_coeff.txt", tmp_path / "matrices.txt", N)
        x2, w2, M2, S2, _ = dg3d.basis_generate(N)
        assert np.array_equal(x, x2) and np.array_equal(w, w2)
        assert np.array_equal(M, M2) and np.array_equal(S, S2)


def test_file_format_matches_reference_parsers(tmp_path):
    """Basis.cpp:135-165: a 2-character prefix then the count; entry index = order.
    Grid.cpp:969-983: 'Order=<N>' then n rows of M then n rows of S; 16 hex digits per value."""
    dg3d.basis_write_files(tmp_path, 3)
    coeff = (tmp_path / "hardcoded_coeff.txt").read_text().splitlines()
    assert coeff[0] == "n=1" and coeff[3] == "n=2" and coeff[9] == "n=4"
    assert len(coeff[10].split()) == 4
    assert all(len(tok) == 16 for tok in coeff[10].split())
    mats = (tmp_path / "matrices.txt").read_text().splitlines()
    assert mats[0] == "Order=1" and mats[5] == "Order=2"
    v = int(mats[1].split()[0], 16)
    assert np.frombuffer(v.to_bytes(8, "little"), dtype="<f8")[0] == pytest.approx(1 / 3, rel=1e-15)
    dpsi = (tmp_path / "hardcoded_coeff2.txt").read_text().splitlines()
    assert dpsi[0].startswith("p") and dpsi[0] == "p=1"


def test_shipped_data_files_match_generator():
    for N in range(1, 8):
        x, w, M, S = problem.load_basis(N)
        x2, w2, M2, S2, _ = dg3d.basis_generate(N)
        assert np.array_equal(M, M2) and np.array_equal(S, S2) and np.array_equal(x, x2)


def test_missing_file_is_an_io_error(tmp_path):
    with pytest.raises(dg3d.Dg3dError) as ei:
        dg3d.basis_read_files(tmp_path / "nope.txt", tmp_path / "nope2.txt", 3)
    assert ei.value.code == dg3d.E_IO
