"""Host logic of the CG loop used by the tolerance solves (cagr_amd/solve.py),
on the CPU with dense operators standing in for the treecode normal
operator: convergence to the requested recurrence residual, the iteration
bound, the periodic residual checks and the stall rule."""
import torch

from cagr_amd.solve import _cg as _cg_lib


def torch_step(x, r, p, q, rr, rr_new):
    """The CG recurrence of cagr_cg_step in torch (the library's step runs on
    a GPU; these tests drive the loop's host logic with it on the CPU)."""
    pq = torch.dot(p.view(-1), q.view(-1))
    alpha = rr / pq if float(pq) != 0.0 else torch.zeros_like(pq)
    x.add_(alpha * p)
    r.sub_(alpha * q)
    rr_new.copy_(torch.dot(r.view(-1), r.view(-1)))
    p.mul_(rr_new / rr if float(rr) != 0.0 else 0.0).add_(r)


def _cg(*a, **k):
    return _cg_lib(*a, step=torch_step, **k)


class DenseNormal:
    """op.normal(p, precision) = M p, optionally with a fixed relative
    perturbation that is not symmetric (a stand-in for an operator whose
    forward and adjoint are transposes of each other only to rounding)."""

    def __init__(self, m, perturb=0.0, seed=0):
        self.m = m
        g = torch.Generator().manual_seed(seed)
        e = torch.randn(m.shape, generator=g, dtype=m.dtype)
        self.mp = m + perturb * float(torch.linalg.matrix_norm(m, 2)) * e / m.shape[0] ** 0.5
        self.calls = 0

    def normal(self, p, precision):
        self.calls += 1
        return self.mp @ p


def spd(n, cond, seed=0):
    g = torch.Generator().manual_seed(seed)
    q, _ = torch.linalg.qr(torch.randn(n, n, generator=g, dtype=torch.float64))
    ev = torch.logspace(0, float(torch.log10(torch.tensor(cond))), n, dtype=torch.float64)
    return (q * ev) @ q.T


def test_cg_reaches_tolerance_and_solution():
    m = spd(200, 1e3)
    b = torch.randn(200, 3, dtype=torch.float64, generator=torch.Generator().manual_seed(1))
    op = DenseNormal(m)
    tol = 1e-10 * float(b.norm())
    x, it = _cg(op, b, torch.zeros_like(b), tol, 10_000, "fp64")
    ref = torch.linalg.solve(m, b)
    assert float((m @ x - b).norm()) <= 10 * tol
    assert float((x - ref).norm() / ref.norm()) < 1e-6
    # the host checks the residual every 10 iterations: at most 9 extra
    assert it % 10 == 0 and op.calls == it
    x1, it1 = _cg(DenseNormal(m), b, torch.zeros_like(b), tol, 10_000, "fp64", check_every=1)
    assert it1 <= it < it1 + 10
    assert torch.allclose(x[:, :], x1, rtol=0, atol=1e-8 * float(ref.abs().max()))


def test_cg_iteration_bound_and_warm_start():
    m = spd(150, 1e4)
    b = torch.ones(150, 1, dtype=torch.float64)
    x, it = _cg(DenseNormal(m), b, torch.zeros_like(b), 0.0, 37, "fp64")
    assert it == 37                       # not a multiple of the check interval
    # a warm start from the solution needs no iteration
    ref = torch.linalg.solve(m, b)
    x2, it2 = _cg(DenseNormal(m), b, ref, 1e-6, 100, "fp64")
    assert it2 == 0 and torch.equal(x2, ref)


def test_cg_stall_rule_stops_a_stagnating_run():
    m = spd(200, 1e8)
    b = torch.randn(200, 1, dtype=torch.float64, generator=torch.Generator().manual_seed(2))
    logs = []
    op = DenseNormal(m, perturb=1e-6)
    x, it = _cg(op, b, torch.zeros_like(b), 1e-14 * float(b.norm()), 20_000, "fp32c",
                log=logs.append, stall=300)
    assert it < 20_000
    assert any("stalled" in l for l in logs)
    # without the rule the same run goes to the bound
    _, it0 = _cg(DenseNormal(m, perturb=1e-6), b, torch.zeros_like(b), 1e-14 * float(b.norm()),
                 3000, "fp32c")
    assert it0 == 3000
