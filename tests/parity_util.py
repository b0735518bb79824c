"""Comparison helpers shared by the parity tests."""
import numpy as np

# tolerance of the plain FP64 path against the oracle: max-norm relative error of a whole array
RTOL = 1e-12


def max_rel(a, b):
    """max |a-b| relative to max |b| (max-norm relative error of a whole array)"""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.abs(b).max() if b.size else 0.0
    if scale == 0.0:
        return float(np.abs(a).max()) if a.size else 0.0
    return float(np.abs(a - b).max() / scale)


def assert_close(a, b, rtol=RTOL, what=""):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    assert np.all(np.isfinite(a)) == np.all(np.isfinite(b)), what
    e = max_rel(a, b)
    assert e <= rtol, f"{what}: max relative error {e:.3e} > {rtol:.1e}"


def sort_particles(p):
    """canonical order of a particle set: by cell, then by delta components, then velocity"""
    keys = [p["v"][:, 2], p["v"][:, 1], p["v"][:, 0]]
    for d in reversed(range(p["delta"].shape[1])):
        keys.append(p["delta"][:, d])
    for d in reversed(range(p["icell"].shape[1])):
        keys.append(p["icell"][:, d])
    idx = np.lexsort(keys)
    return {k: v[idx] for k, v in p.items()}


def assert_same_particles(a, b, rtol=RTOL, exact=False, what="particles"):
    assert len(a["weight"]) == len(b["weight"]), (what, len(a["weight"]), len(b["weight"]))
    a = sort_particles(a)
    b = sort_particles(b)
    assert np.array_equal(a["icell"], b["icell"]), f"{what}: cell indices differ"
    if exact:
        for k in ("delta", "v", "weight", "charge"):
            assert np.array_equal(a[k], b[k]), f"{what}: {k} differs"
    else:
        for k in ("delta", "v", "weight", "charge"):
            assert_close(a[k], b[k], rtol, f"{what}.{k}")


def random_particles(rng, dim, n, ncells, lower, spread, vscale=1.0):
    """n particles in the domain grown by `spread` cells, deltas in [0,1)"""
    ic = np.stack([rng.integers(lower[d] - spread, lower[d] + ncells[d] + spread, n)
                   for d in range(dim)], 1).astype(np.int32)
    return dict(icell=ic, delta=rng.random((n, dim)), v=vscale * rng.standard_normal((n, 3)),
                weight=0.5 + rng.random(n), charge=1.0 + 0.5 * rng.random(n))
