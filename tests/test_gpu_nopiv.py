"""GPU: complex64 LU without pivoting + FP64 refinement (FRICO_SOLVE_C64_NP_IR), the refinement form
whose factorisation needs no grid-wide pivot search.  Held to numpy's solve and to the complex128 LU;
its fall-backs to the pivoted LU are exercised."""
import numpy as np
import pytest

from frico_amd import meshes

pytestmark = pytest.mark.gpu

F0 = 300e6


def relmax(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


def exact_lu_system(n, seed, per_row=2, nrhs=4):
    """Z = L U with L unit lower and U upper, both with per_row entries of {+-1, +-j} a row off the
    diagonal, and U's diagonal of units {+-1, +-j}: unpivoted elimination recovers L and U with every
    intermediate a small Gaussian integer, so complex64 is exact and the first iterate equals the
    integer solution bit for bit."""
    rng = np.random.default_rng(seed)
    L = np.eye(n, dtype=np.complex128)
    U = np.diag(rng.choice([1, -1, 1j, -1j], size=n)).astype(np.complex128)
    for i in range(1, n):
        cols = rng.choice(i, size=min(per_row, i), replace=False)
        L[i, cols] = rng.choice([-1, 1], size=cols.size) * rng.choice([1, 1j], size=cols.size)
    for i in range(n - 1):
        cols = i + 1 + rng.choice(n - 1 - i, size=min(per_row, n - 1 - i), replace=False)
        U[i, cols] = rng.choice([-1, 1], size=cols.size) * rng.choice([1, 1j], size=cols.size)
    X = (rng.integers(-3, 4, (n, nrhs)) + 1j * rng.integers(-3, 4, (n, nrhs))).astype(np.complex128)
    A = L @ U
    return A, A @ X, X


# one to four right-hand sides take the chained solves (one launch a triangle), five the blocked ones;
# n = 130 ends in a short chunk, n = 64 is a single chunk (blocked)
@pytest.mark.parametrize("nb", [32, 64, 128])
@pytest.mark.parametrize("carried", [True, False])
@pytest.mark.parametrize("n,nrhs", [(700, 4), (700, 1), (130, 2), (1000, 5), (64, 3)])
def test_unpivoted_lu_is_exact_on_integer_factors(nb, carried, n, nrhs):
    import frico_amd
    A, B, X = exact_lu_system(n, 31, nrhs=nrhs)
    assert np.max(np.abs(A)) < 100
    with frico_amd.Context(0) as c:
        c.set_block(nb)
        c.set_matrix(A)
        if carried:
            c.set_rhs(B)
        c.factor("c64_np_ir")
        assert c.factored_form == frico_amd.SOLVE_C64_NP_IR
        if not carried:
            c.set_rhs(B)
        c.solve(tol=1.0, maxit=1)
        assert np.array_equal(c.solution(), X)


@pytest.mark.parametrize("n,nb", [(1, 128), (63, 64), (200, 64), (517, 128), (1300, 128), (2000, 64),
                                  (45, 32), (700, 32)])
@pytest.mark.parametrize("nrhs", [1, 3, 6])
def test_unpivoted_lu_solves_dominant_systems(n, nb, nrhs):
    import frico_amd
    rng = np.random.default_rng(n + 1)
    A = rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n)) + 4.0 * np.sqrt(n) * np.eye(n)
    B = rng.uniform(-1, 1, (n, nrhs)) + 1j * rng.uniform(-1, 1, (n, nrhs))
    print(f"unpivoted n={n} nb={nb} nrhs={nrhs}", flush=True)
    with frico_amd.Context(0) as c:
        c.set_block(nb)
        c.set_matrix(A)
        c.set_rhs(B)
        c.factor("c64_np_ir")
        info = c.solve(tol=1e-12, maxit=10)
        assert info.converged and c.factored_form == frico_amd.SOLVE_C64_NP_IR and info.iterations <= 4
        assert relmax(c.solution(), np.linalg.solve(A, B)) <= 1e-10


@pytest.mark.parametrize("diag", [4.0, 1.0])
def test_unpivoted_lu_with_long_block_columns(diag):
    """n = 4700 at nb = 128: the first block columns have more than 4096 rows below the diagonal
    block, so their L21 takes the transposed path (col_transpose, trsm_block and the joining
    product, split_f16) instead of the fused block solve; the weaker diagonal gives the diagonal
    blocks larger condition numbers."""
    import frico_amd
    n, nrhs = 4700, 3
    rng = np.random.default_rng(4700)
    A = rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n)) + diag * np.sqrt(n) * np.eye(n)
    B = rng.uniform(-1, 1, (n, nrhs)) + 1j * rng.uniform(-1, 1, (n, nrhs))
    print(f"unpivoted long block columns n={n} diag={diag}", flush=True)
    with frico_amd.Context(0) as c:
        c.set_block(128)
        c.set_matrix(A)
        c.set_rhs(B)
        c.factor("c64_np_ir")
        info = c.solve(tol=1e-12, maxit=10)
        assert info.converged and c.factored_form == frico_amd.SOLVE_C64_NP_IR and info.iterations <= 4
        assert relmax(c.solution(), np.linalg.solve(A, B)) <= 1e-10


def test_unpivoted_factors_do_not_depend_on_the_lookahead(monkeypatch):
    import frico_amd
    rng = np.random.default_rng(3)
    n = 1100
    A = rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n)) + 2.0 * np.sqrt(n) * np.eye(n)
    B = rng.uniform(-1, 1, (n, 5)) + 1j * rng.uniform(-1, 1, (n, 5))
    out = []
    for la in ("1", "0"):
        monkeypatch.setenv("FRICO_LU_LOOKAHEAD", la)
        with frico_amd.Context(0) as c:
            c.set_block(128)
            c.set_matrix(A)
            c.set_rhs(B)
            c.factor("c64_np_ir")
            c.solve(tol=1.0, maxit=1)
            out.append(c.solution())
    assert np.array_equal(out[0], out[1])


def test_zero_pivot_falls_back_to_the_pivoted_lu():
    import frico_amd
    rng = np.random.default_rng(4)
    n = 300
    A = rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n)) + 2.0 * np.sqrt(n) * np.eye(n)
    A[0, 0] = 0.0
    B = rng.uniform(-1, 1, (n, 2)) + 1j * rng.uniform(-1, 1, (n, 2))
    with frico_amd.Context(0) as c:
        c.set_matrix(A)
        c.set_rhs(B)
        c.factor("c64_np_ir")
        assert c.factored_form == frico_amd.SOLVE_C64_IR
        assert c.solve(tol=1e-12).converged
        assert relmax(c.solution(), np.linalg.solve(A, B)) <= 1e-10


@pytest.mark.parametrize("z", [40, 150, 255, 383])
def test_zero_pivot_inside_a_diagonal_block_falls_back(z):
    """Block-diagonal A = diag(A1, [[0, 1], [1, 0]], A2) with the 2 x 2 swap at column z: the
    unpivoted elimination of A1 never reaches the swap block, so its pivot at z is exactly zero.
    z = 40, 150, 255, 383 put the zero in the first and second 128-column diagonal blocks, in a
    middle sub-panel, at a block's last column and at the last column of the matrix's third block
    (the swap then straddles two diagonal blocks)."""
    import frico_amd
    rng = np.random.default_rng(z)
    n = 520
    A = np.zeros((n, n), dtype=np.complex128)
    A[:z, :z] = rng.uniform(-1, 1, (z, z)) + 1j * rng.uniform(-1, 1, (z, z)) + 2.0 * np.sqrt(z) * np.eye(z)
    A[z, z + 1] = A[z + 1, z] = 1.0
    m = n - z - 2
    A[z + 2:, z + 2:] = rng.uniform(-1, 1, (m, m)) + 1j * rng.uniform(-1, 1, (m, m)) + 2.0 * np.sqrt(m) * np.eye(m)
    B = rng.uniform(-1, 1, (n, 3)) + 1j * rng.uniform(-1, 1, (n, 3))
    with frico_amd.Context(0) as c:
        c.set_block(128)
        c.set_matrix(A)
        c.set_rhs(B)
        c.factor("c64_np_ir")
        assert c.factored_form == frico_amd.SOLVE_C64_IR
        assert c.solve(tol=1e-12).converged
        assert relmax(c.solution(), np.linalg.solve(A, B)) <= 1e-10


def test_growth_falls_back_after_the_refinement_stalls():
    import frico_amd
    rng = np.random.default_rng(5)
    n = 400
    A = rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n)) + 2.0 * np.sqrt(n) * np.eye(n)
    A[0, 0] = 1e-20
    B = rng.uniform(-1, 1, (n, 2)) + 1j * rng.uniform(-1, 1, (n, 2))
    with frico_amd.Context(0) as c:
        c.set_matrix(A)
        c.set_rhs(B)
        c.factor("c64_np_ir")
        assert c.factored_form == frico_amd.SOLVE_C64_NP_IR
        info = c.solve(tol=1e-12, maxit=10)
        assert info.converged and c.factored_form == frico_amd.SOLVE_C64_IR
        assert relmax(c.solution(), np.linalg.solve(A, B)) <= 1e-10


def test_quadrature_plate_monostatic_rcs(gpu_ctx):
    """Quadrature-mode plate (20 x 20 squares, N = 1160), 7 incidences: the unpivoted form needs
    no fall-back and its RCS equals the complex128 LU's to 1e-9."""
    import frico_amd
    lam = meshes.wavelength(F0)
    sm = meshes.plate(1.5 * lam, 20)
    th = np.radians(np.arange(0.0, 91.0, 15.0))
    gpu_ctx.load_mesh(sm, 5)
    s128, _ = gpu_ctx.monostatic_sweep([F0], th, 0.0, form="c128")
    s64, info = gpu_ctx.monostatic_sweep([F0], th, 0.0, form="c64_np_ir", tol=1e-12)
    assert gpu_ctx.factored_form == frico_amd.SOLVE_C64_NP_IR
    assert info[0].converged and info[0].iterations <= 3
    assert relmax(s64, s128) <= 1e-9


@pytest.mark.parametrize("form", ["c64_np_ir", "c64_sym_ir", "c64_ir"])
@pytest.mark.parametrize("bad", [np.nan, np.inf])
def test_non_finite_matrix_is_never_reported_converged(form, bad):
    """A NaN or infinite entry of Z spreads NaN through the factors and the residual.  The
    refinement must count such a residual as the worst possible one: the unpivoted forms fall back
    to the pivoted LU (whose pivot search takes a NaN as an ordinary, largest pivot), and the solve
    ends in FRICO_ERR_NOCONV instead of reporting convergence."""
    import frico_amd
    n = 300
    rng = np.random.default_rng(5)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = A + A.T + n * np.eye(n)
    A[17, 42] = A[42, 17] = bad
    B = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    with frico_amd.Context(0) as c:
        c.set_matrix(A)
        c.set_rhs(B)
        c.factor(form)
        with pytest.raises(frico_amd.NotConvergedError):
            c.solve(tol=1e-10, maxit=3)
        assert c.factored_form == frico_amd.SOLVE_C64_IR
