"""The one-kernel PCG iteration (kPcgStep, drape-2_amd/csrc/kernels.hip) restated in numpy FP64 in the kernel's
order of operations, against the oracle's solveLinear (cpu_solver.cpp:169-213).

The kernel takes two scalars of standard PCG from exact expansions over the previous iteration's vectors
(r_k.z_k = r.zr - 2 alpha r.t + alpha^2 ap.t and r_k.r_k = r.r - 2 alpha r.ap + alpha^2 ap.ap), rebuilds each
gathered column's z as zr - alpha t, and applies the block-Jacobi inverse (stored as its upper triangle) with the
contact filter after it, filter(M x) (cpu_solver.cpp:179-181).
These tests pin that algebra and the stopping rules (r0 = 0, p.Ap <= 0 before an update, tolerance and iteration
cap after one) against the oracle's PCG, with and without held contacts: same number of updates, same step.
"""
import numpy as np
import pytest

from test_oracle_terms import _rig


def _dense_blocks(rs, ci, bl, n):
    A = np.zeros((3 * n, 3 * n))
    for i in range(n):
        for s in range(rs[i], rs[i + 1]):
            j = ci[s]
            A[3 * i:3 * i + 3, 3 * j:3 * j + 3] = bl[s]
    return A


def one_kernel_pcg(oracle, rs, ci, bl, pin, active, normals, rhs, tol, max_cg):
    """Returns (step, updates) of the one-kernel form, kernel by kernel."""
    n = len(rs) - 1
    A = _dense_blocks(rs, ci, bl, n)
    free = np.repeat(pin == 0, 3).astype(float)
    nrm = np.where(active[:, None] != 0, normals, 0.0)

    def S(x):
        x = x.reshape(n, 3)
        return (x - nrm * np.sum(nrm * x, axis=1, keepdims=True)).reshape(-1)

    # Block-Jacobi inverse M (identity when the block is not positive definite, zero if pinned), symmetrised as
    # the kernel stores it (upper triangle); applied as filter(M x).
    M = np.zeros((n, 3, 3))
    for i in range(n):
        if pin[i]:
            continue
        d = np.where(ci[rs[i]:rs[i + 1]] == i)[0][0] + rs[i]
        ok, inv = oracle.invert_spd3(bl[d])
        Mi = inv if ok else np.eye(3)
        M[i] = np.triu(Mi) + np.triu(Mi, 1).T

    def mul(x):
        return S(np.einsum("nij,nj->ni", M, x.reshape(n, 3)).reshape(-1))

    def SA(p):
        return S((A @ (p * free)) * free)

    # Init kernel: outputs of "kernel -1" (grad = -rhs, r_0 = S(-grad)).
    r_prev = S(np.asarray(rhs, float).reshape(-1))
    ap_prev = np.zeros(3 * n)
    p_prev = np.zeros(3 * n)
    zr_prev = mul(r_prev)
    t_prev = np.zeros(3 * n)
    parts = (0.0, r_prev @ zr_prev, 0.0, 0.0, 0.0, 0.0, r_prev @ r_prev)
    step = np.zeros(3 * n)
    k, updates, r0 = 0, 0, 0.0
    while True:
        pAp, rz, rt, apt, rap, apap, rr = parts
        update, alpha, beta = False, 0.0, 0.0
        if k == 0:
            r0 = np.sqrt(rr)
            conv = not (r0 > 0) or max_cg == 0
        elif not (pAp > 0):
            conv = True
        else:
            update = True
            alpha = rz / pAp
            rr_next = rr - 2 * alpha * rap + alpha * alpha * apap
            rz_next = rz - 2 * alpha * rt + alpha * alpha * apt
            conv = np.sqrt(max(rr_next, 0.0)) <= tol * r0 or k >= max_cg
            if not conv:
                beta = rz_next / rz
        if update:
            step = step + alpha * p_prev
            updates += 1
        if conv:
            return step.reshape(n, 3), updates
        r = r_prev - alpha * ap_prev
        p = (zr_prev - alpha * t_prev) + beta * p_prev
        ap = SA(p)
        zr, t = mul(r), mul(ap)
        parts = (p @ ap, r @ zr, r @ t, ap @ t, r @ ap, ap @ ap, r @ r)
        r_prev, ap_prev, p_prev, zr_prev, t_prev = r, ap, p, zr, t
        k += 1


def _anisotropic(rs, ci, bl, seed):
    """The rig with full (non-isotropic) SPD diagonal blocks, so that M r leaves a held row's contact plane
    unless the preconditioner is filtered."""
    rng = np.random.default_rng(seed)
    bl = bl.copy()
    n = len(rs) - 1
    for i in range(n):
        d = np.where(ci[rs[i]:rs[i + 1]] == i)[0][0] + rs[i]
        q = rng.uniform(-1, 1, (3, 3))
        bl[d] = bl[d] + 1.5 * q @ q.T
    return bl


def _contacts(n, rows, seed):
    rng = np.random.default_rng(seed)
    active = np.zeros(n, np.uint8)
    active[list(rows)] = 1
    normals = rng.normal(size=(n, 3))
    normals /= np.linalg.norm(normals, axis=1, keepdims=True)
    return active, normals


@pytest.mark.parametrize("tol,max_cg", [(1e-2, 50), (1e-6, 200), (1e-12, 500), (1e-12, 3), (1e-6, 0)])
@pytest.mark.parametrize("with_contact", [False, True], ids=["free", "contact"])
def test_one_kernel_form_reproduces_oracle_pcg(oracle, tol, max_cg, with_contact):
    rs, ci, bl, pin, M, free = _rig()
    bl = _anisotropic(rs, ci, bl, 13)
    n = len(rs) - 1
    rng = np.random.default_rng(5)
    b = rng.uniform(-1, 1, (n, 3))
    b[pin == 1] = 0
    active, normals = _contacts(n, (0, 5, 11, 22, 31, 38) if with_contact else (), 9)
    x_ref, it_ref, _ = oracle.bsr_pcg_contact(rs, ci, bl, pin, active, normals, b, tol, max_cg)
    x, it = one_kernel_pcg(oracle, rs, ci, bl, pin, active, normals, b, tol, max_cg)
    assert it == it_ref
    assert np.max(np.abs(x - x_ref)) <= 1e-10 * max(1.0, np.max(np.abs(x_ref)))
    if with_contact:
        # held rows move only in their contact planes
        assert np.max(np.abs(np.sum(x[active == 1] * normals[active == 1], axis=1))) < 1e-12


def test_one_kernel_form_stops_where_oracle_finds_negative_curvature(oracle):
    """An indefinite system (one diagonal block negative: identity preconditioner there): both forms stop before
    the update whose p.Ap is not positive."""
    rs, ci, bl, pin, M, free = _rig(seed=91)
    bl = bl.copy()
    n = len(rs) - 1
    d = np.where(ci[rs[7]:rs[8]] == 7)[0][0] + rs[7]
    bl[d] = -8.0 * np.eye(3)
    b = np.random.default_rng(4).uniform(-1, 1, (n, 3))
    b[pin == 1] = 0
    active, normals = _contacts(n, (), 1)
    x_ref, it_ref, _ = oracle.bsr_pcg_contact(rs, ci, bl, pin, active, normals, b, 1e-12, 500)
    x, it = one_kernel_pcg(oracle, rs, ci, bl, pin, active, normals, b, 1e-12, 500)
    assert it == it_ref < 500
    assert np.max(np.abs(x - x_ref)) <= 1e-10 * max(1.0, np.max(np.abs(x_ref)))


def test_one_kernel_form_zero_rhs_does_nothing(oracle):
    rs, ci, bl, pin, M, free = _rig()
    n = len(rs) - 1
    active, normals = _contacts(n, (2,), 3)
    x, it = one_kernel_pcg(oracle, rs, ci, bl, pin, active, normals, np.zeros((n, 3)), 1e-6, 50)
    assert it == 0 and not np.any(x)
