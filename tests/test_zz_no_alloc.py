"""qmc_energy_noalloc (staging in the library's fixed device scratch array)
against qmc_energy (staging in allocated device buffers), through the C ABI:
two calls in a row, the second on a smaller input, bit-exact energies; an
input beyond the scratch limit is refused before any device work."""
import numpy as np
import pytest

from oracle import pyoracle as po
from qiotoolkit_amd import _qmc, workloads as wl

pytestmark = pytest.mark.gpu


def _problem(ctx, kind, terms):
    op = po.Problem(kind, terms)
    off, ids, c = op.edges()
    return op, _qmc.Problem(ctx, _qmc.QMC_ISING if kind == "ising" else _qmc.QMC_PUBO, op.n, off, ids, c)


@pytest.mark.parametrize("kind,terms", [("ising", lambda: wl.er_ising(3000, 6.0, seed=61)),
                                        ("pubo", lambda: wl.random_pubo(800, 4000, seed=62)),
                                        ("ising", lambda: wl.lattice3d(8, seed=63))])
def test_energy_noalloc_matches_energy(kind, terms):
    ctx = _qmc.Context(0)
    op, qp = _problem(ctx, kind, terms())
    rng = np.random.default_rng(64)
    for n_states in (700, 37):  # second call smaller: stale bytes of the first stay behind it
        states = rng.integers(0, 2, size=(n_states, op.n), dtype=np.uint8)
        want = qp.energy(states)
        got = qp.energy(states, noalloc=True)
        assert np.array_equal(got, want), n_states
        np.testing.assert_allclose(got[:5], [op.cost(s) for s in states[:5]], rtol=1e-12)
    # beyond the limit: refused on the host with the reason, nothing launched
    limit = _qmc.NOALLOC_SCRATCH_BYTES
    too_many = limit // op.n + 1
    big = np.zeros((too_many, op.n), dtype=np.uint8)
    with pytest.raises(_qmc.QmcError) as e:
        qp.energy(big, noalloc=True)
    assert e.value.code == _qmc.QMC_INVALID_ARGUMENT and "scratch" in str(e.value)
    # and the next call still works
    states = rng.integers(0, 2, size=(3, op.n), dtype=np.uint8)
    assert np.array_equal(qp.energy(states, noalloc=True), qp.energy(states))
    qp.close()
    ctx.close()
