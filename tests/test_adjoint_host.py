"""Host side of the gradient harness: the torch continued fraction against the oracle's, and torch.autograd of the contour
objective against central finite differences (no GPU)."""
import numpy as np
import torch

from asr_amd import adjoint as AD
from asr_amd import host
from asr_amd.augcls import Member


def _chains(rng, nch, m):
    a = rng.uniform(-0.3, 0.3, (nch, m))
    b = rng.uniform(0.2, 0.5, (nch, m))
    return a, b


def test_cf_green_matches_oracle_avg2():
    from oracle.oracle import np_cf_species
    rng = np.random.default_rng(0)
    a, b = _chains(rng, 4, 12)
    z = np.array([-0.9 + 0.05j, -0.2 + 0.3j, 0.1 + 0.01j, 0.7 + 0.4j, 1.8 + 0.2j])
    g = AD.cf_green(torch.from_numpy(a), torch.from_numpy(b), torch.from_numpy(z)).numpy()
    for c in range(4):
        for j, zz in enumerate(z):
            ref, _ = np_cf_species(a[c], b[c], np.zeros(12), 12, zz, tail="avg2")
            assert abs(g[c, j] - ref) < 1e-12 * max(1.0, abs(ref))


def test_cf_green_is_a_herglotz_function():
    """Im G < 0 in the upper half plane (a resolvent's diagonal element), for any chain and the avg2 tail."""
    rng = np.random.default_rng(1)
    a, b = _chains(rng, 6, 9)
    z = torch.from_numpy(rng.uniform(-2, 2, 40) + 1j * rng.uniform(1e-3, 1.0, 40))
    g = AD.cf_green(torch.from_numpy(a), torch.from_numpy(b), z)
    assert torch.all(g.imag < 0)


def test_contour_objective_gradient_fd():
    rng = np.random.default_rng(2)
    nch, m = 3, 10
    a0, b0 = _chains(rng, nch, m)
    z, w = host.contour_closed(-1.5, 0.1, 0.5, 16, 8)
    zt, wt = torch.from_numpy(z), torch.from_numpy(w)
    weight = torch.tensor([0.5, 0.3, 0.2], dtype=torch.float64)
    q0 = torch.tensor([0.4, 0.6, 0.5], dtype=torch.float64)

    def L(a, b):
        return AD.contour_objective(a, b, zt, wt, weight, q0, lam=2.0)

    a = torch.tensor(a0, requires_grad=True)
    b = torch.tensor(b0, requires_grad=True)
    L(a, b).backward()
    h = 1e-6
    for arr, grad, is_a in ((a0, a.grad.numpy(), True), (b0, b.grad.numpy(), False)):
        for c in range(nch):
            for k in range(m):
                p = arr.copy(); p[c, k] += h
                q = arr.copy(); q[c, k] -= h
                if is_a:
                    fd = (L(torch.from_numpy(p), torch.from_numpy(b0)) - L(torch.from_numpy(q), torch.from_numpy(b0))) / (2 * h)
                else:
                    fd = (L(torch.from_numpy(a0), torch.from_numpy(p)) - L(torch.from_numpy(a0), torch.from_numpy(q))) / (2 * h)
                assert abs(float(fd) - grad[c, k]) < 1e-6 * max(1.0, abs(grad[c, k])), (is_a, c, k)
    # b_{m-1} does not enter the fraction
    assert np.all(b.grad.numpy()[:, m - 1] == 0.0)


def test_members_at_spreads_theta_over_l_shells():
    theta = np.arange(18, dtype=np.float64).reshape(2, 3, 3)
    off = np.full((1, 3, 3), 0.5)
    m = Member(l0=2, c=np.zeros((3, 9)), d=np.zeros((3, 9)), o=np.zeros((3, 9)), e=np.zeros((3, 9)), frozen=1)
    (mm,) = AD.members_at([m], theta, off)
    assert mm.l0 == 2 and mm.frozen == 1
    assert np.array_equal(mm.c[1], [3.5, 4.5, 4.5, 4.5, 5.5, 5.5, 5.5, 5.5, 5.5])
    assert np.array_equal(mm.d[2], [15, 16, 16, 16, 17, 17, 17, 17, 17])
