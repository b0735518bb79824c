"""phx_gather_noalloc, the gather entry point that allocates nothing (its staging is a fixed device
array of the library): the same E/B as phx_gather, bit for bit, on two calls in a row (the second
with a smaller array, so that what the first call left in the shared array must not leak into the
second), on linear and on cell-segmented storage (empty slots skipped), and the capacity limit."""
import sys

import numpy as np
import pytest

import parity_util as U
from phare_amd import phx

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _level(dim, order, nc, rng):
    g = phx.Level(dim, order, nc, [0.3, 0.25, 0.2][:dim], capacity=20000)
    for k in range(6):
        a = 0.3 * rng.standard_normal(g.shape(phx.F_EAVG + k))
        g.set(phx.F_EAVG + k, a + (1.0 if k == 3 else 0.0))
    return g


@pytest.mark.parametrize("dim,order", [(1, 1), (2, 2), (3, 3)])
def test_noalloc_gather_matches_gather_twice(dim, order):
    _say(f"[gpu] gather_noalloc dim={dim} order={order}")
    rng = np.random.default_rng(900 + 10 * dim + order)
    nc = [[23], [14, 11], [9, 10, 11]][dim - 1]
    g = _level(dim, order, nc, rng)
    for n in (6000, 1500):  # the second call on a smaller array
        p = U.random_particles(rng, dim, n, nc, [0] * dim, 0, vscale=0.5)
        g.set_particles(0, phx.DOMAIN, p)
        E0, B0 = g.gather(phx.EM_AVG)
        E1, B1 = g.gather_noalloc(phx.EM_AVG)
        assert E1.shape == (n, 3)
        assert np.array_equal(E0, E1) and np.array_equal(B0, B1), n
    g.close()


def test_noalloc_gather_on_sorted_storage_skips_empty_slots():
    _say("[gpu] gather_noalloc on segmented storage")
    rng = np.random.default_rng(77)
    nc = [20, 12]
    g = _level(2, 2, nc, rng)
    p = U.random_particles(rng, 2, 5000, nc, [0, 0], 0, vscale=0.5)
    g.set_particles(0, phx.DOMAIN, p)
    g.sort_particles()
    n_live = g.count(0, phx.DOMAIN)
    E0, B0 = g.gather(phx.EM_AVG)
    E1, B1 = g.gather_noalloc(phx.EM_AVG)
    assert E1.shape == (n_live, 3) == E0.shape
    assert np.array_equal(E0, E1) and np.array_equal(B0, B1)
    g.close()


def test_noalloc_gather_refuses_arrays_beyond_its_scratch():
    _say("[gpu] gather_noalloc capacity limit")
    n = (1 << 20) + 1
    nc = [64]
    g = phx.Level(1, 1, nc, [0.2], capacity=n)
    p = {"icell": np.zeros((n, 1), np.int32), "delta": np.full((n, 1), 0.5),
         "v": np.zeros((n, 3)), "weight": np.ones(n), "charge": np.ones(n)}
    g.set_particles(0, phx.DOMAIN, p)
    with pytest.raises(phx.PhxError) as e:
        g.gather_noalloc(phx.EM_AVG)
    assert e.value.status == 5 and "scratch" in str(e.value)
    g.close()
