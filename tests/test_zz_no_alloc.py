"""asr_gpu_kpm_moments_noalloc (the entry point that allocates no device memory; its work arrays live in the library's
fixed device scratch) against asr_gpu_kpm_moments on the same supercell and Rademacher vectors: the moments bit for bit,
on two calls in a row, the second on a smaller supercell with fewer vectors and moments (the scratch keeps the first
call's contents), with and without moment doubling; sizes beyond the scratch are refused before any GPU work."""
import numpy as np
import pytest

from asr_amd import workloads as W

pytestmark = pytest.mark.gpu


def _engine():
    from asr_amd import engine
    assert engine.device_count() >= 1, "no GPU visible"
    return engine


@pytest.mark.parametrize("doubling", [False, True])
def test_noalloc_moments_equal_the_allocating_entry(doubling):
    import torch
    eng = _engine()
    cases = [(6, 64, 96, 5), (3, 40, 30, 6)]  # (supercell L, vectors, moments, seed): the second call is smaller
    for L, nvec, nmom, seed in cases:
        g, species, par, block, block0 = W.kpm_supercell(L=L, seed=seed)
        R = eng.GpuRealization(g.nbr, species, par, block, block0)
        R.set_doubling(doubling)
        lo, hi, _, _ = R.spectral_bounds(40, 0.05)
        c, s = 0.5 * (lo + hi), 0.5 * (hi - lo) / 0.99
        ref = R.kpm_moments(c, s, nmom, nvec, seed)["mu"]
        mu = torch.full((nmom,), float("nan"), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        print(f"noalloc: L={L} ({g.n} sites), {nvec} vectors, {nmom} moments, doubling {doubling}", flush=True)
        R.kpm_moments_noalloc(c, s, nmom, nvec, seed, mu.data_ptr())
        got = mu.cpu().numpy()
        assert np.array_equal(got, ref), float(np.max(np.abs(got - ref)))
        R.close()


def test_noalloc_refuses_sizes_beyond_the_scratch():
    import torch
    eng = _engine()
    g, species, par, block, block0 = W.kpm_supercell(L=12, seed=3)  # 6,912 sites: 3 x 62,208 rows x 64 vectors > 64 MiB
    R = eng.GpuRealization(g.nbr, species, par, block, block0)
    mu = torch.zeros(16, dtype=torch.float64, device="cuda")
    with pytest.raises(eng.EngineError, match="exceed the fixed device scratch"):
        R.kpm_moments_noalloc(0.0, 2.0, 16, 64, 1, mu.data_ptr())
    R.close()
