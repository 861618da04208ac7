"""The device Davidson-Liu solver with dense test matrices, after the reference's
tests/pytest/dl_solver/test_dl.py and test_dl_errors.py."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

M4 = np.array([[-1, 1, 1, 1], [1, 0, 1, 1], [1, 1, 0, 1], [1, 1, 1, 0]], dtype=float)


def synthetic(size):
    M = np.zeros((size, size))
    for i in range(size):
        M[i, i] = -1.0 + i * 0.1
        for j in range(i):
            M[i, j] = M[j, i] = 0.05 / (1.0 + abs(i - j))
    return M


def solve_dl(size, nroot):
    import forte_amd
    M = synthetic(size)
    s = forte_amd.DavidsonLiuSolver(size, nroot)
    s.add_h_diag(np.diag(M).copy())
    s.add_guesses([[(i, 1.0)] for i in range(nroot)])
    s.add_test_sigma_builder(M)
    s.solve()
    assert np.allclose(s.eigenvalues(), np.linalg.eigvalsh(M)[:nroot])


def test_dl_1():
    import forte_amd
    s = forte_amd.DavidsonLiuSolver(4, 1)
    s.add_h_diag(np.diag(M4).copy())
    s.add_guesses([[(0, 1.0)], [(1, 1.0)], [(2, 1.0)], [(3, 1.0)]])
    s.add_test_sigma_builder(M4)
    s.solve()
    assert np.isclose(s.eigenvalues()[0], np.linalg.eigvalsh(M4)[0])


def test_dl_2():
    import forte_amd
    s = forte_amd.DavidsonLiuSolver(4, 1)
    s.add_h_diag(np.diag(M4).copy())
    s.add_guesses([[(0, 0.1)]])
    s.add_test_sigma_builder(M4)
    s.solve()
    assert np.isclose(s.eigenvalues()[0], np.linalg.eigvalsh(M4)[0])


def test_dl_3():
    for nroot in (1, 2, 5, 10):
        solve_dl(10, nroot)
        solve_dl(100, nroot)
        solve_dl(1000, nroot)


def test_dl_4():
    for size in range(1, 20):
        for nroot in range(1, size + 1):
            solve_dl(size, nroot)


def test_project_out():
    import forte_amd
    w, V = np.linalg.eigh(M4)
    s = forte_amd.DavidsonLiuSolver(4, 1)
    s.add_test_sigma_builder(M4)
    s.add_h_diag(np.diag(M4).copy())
    s.add_project_out_vectors([[(i, v) for i, v in enumerate(V[:, 0])]])
    s.solve()
    assert np.isclose(s.eigenvalues()[0], w[1])


def test_dl_restart():
    import forte_amd
    s = forte_amd.DavidsonLiuSolver(4, 1)
    s.add_test_sigma_builder(M4)
    s.add_h_diag(np.diag(M4).copy())
    s.add_guesses([[(0, 1.0)]])
    s.solve()
    s.solve()
    assert np.isclose(s.eigenvalues()[0], np.linalg.eigvalsh(M4)[0])
    M2 = M4.copy()
    M2[0, 0] = -2
    s.add_test_sigma_builder(M2)
    s.solve()
    assert np.isclose(s.eigenvalues()[0], np.linalg.eigvalsh(M2)[0])


def test_dl_errors():
    import forte_amd
    s = forte_amd.DavidsonLiuSolver(4, 1)
    s.add_h_diag(np.array([-1.0, 0, 0, 0]))
    s.add_guesses([[(0, 0.0)]])
    s.add_test_sigma_builder(M4)
    with pytest.raises(RuntimeError):
        s.solve()
    s = forte_amd.DavidsonLiuSolver(4, 2)
    s.add_h_diag(np.diag(M4).copy())
    s.add_guesses([[(0, 0.1)]])
    s.add_test_sigma_builder(M4)
    with pytest.raises(RuntimeError):
        s.solve()
    s = forte_amd.DavidsonLiuSolver(4, 1)
    s.add_h_diag(np.diag(M4).copy())
    s.add_guesses([[(0, 0.1)]])
    with pytest.raises(RuntimeError):
        s.solve()
    s = forte_amd.DavidsonLiuSolver(4, 1)
    s.add_test_sigma_builder(M4)
    with pytest.raises(RuntimeError):
        s.solve()
    with pytest.raises(RuntimeError):
        forte_amd.DavidsonLiuSolver(4, 1, 2, 2)
    with pytest.raises(RuntimeError):
        forte_amd.DavidsonLiuSolver(0, 1)
