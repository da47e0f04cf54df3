"""CPU checks of pfast_source_order, the internal source order of the batched 3D shot
sets (host code of libpfast; no GPU is touched).

Eight consecutive internal sources share one sweep (a chunk).  The order must be a
permutation, deterministic, the identity in mode 0, Morton order of the nodes in mode 1
(checked against a from-scratch bit interleave), and in mode 2 balanced clusters whose
members lie closer to their centroid than Morton chunks do.
"""
import numpy as np
import pytest


def _spread(src, perm, spacing, C=8):
    P = src[perm] * np.array([spacing[1], spacing[0], spacing[2]])
    d = []
    for c in range(0, len(perm), C):
        g = P[c:c + C]
        d.append(np.sqrt(((g - g.mean(0)) ** 2).sum(1)).mean())
    return float(np.mean(d))


def _morton_ref(src):
    def key(s):
        z = 0
        for b in range(21):
            z |= ((int(s[0]) >> b) & 1) << (3 * b + 2)
            z |= ((int(s[1]) >> b) & 1) << (3 * b + 1)
            z |= ((int(s[2]) >> b) & 1) << (3 * b)
        return z
    return np.array(sorted(range(len(src)), key=lambda i: key(src[i])), np.int32)


@pytest.mark.parametrize("n,seed", [(400, 1), (37, 2), (1, 3), (8, 4), (1003, 5)])
def test_source_orders(n, seed):
    import pfast_amd
    rng = np.random.default_rng(seed)
    src = np.stack([rng.integers(0, 201, n), rng.integers(0, 201, n),
                    rng.integers(0, 3, n)], 1).astype(np.int32)
    sp = (25.0, 20.0, 10.0)
    ident = pfast_amd.source_order(src, sp, "caller")
    assert np.array_equal(ident, np.arange(n))
    mort = pfast_amd.source_order(src, sp, "morton")
    assert np.array_equal(mort, _morton_ref(src))
    clus = pfast_amd.source_order(src, sp, "clusters")
    assert sorted(clus.tolist()) == list(range(n))
    assert np.array_equal(clus, pfast_amd.source_order(src, sp, "clusters"))
    if n > 8:
        s_id, s_m, s_c = (_spread(src, p, sp) for p in (ident, mort, clus))
        print(f"n {n}: member-centroid distance caller {s_id:.0f}, morton {s_m:.0f}, "
              f"clusters {s_c:.0f}")
        assert s_c <= s_m <= s_id


def test_source_order_rejects_bad_mode():
    import pfast_amd
    with pytest.raises(pfast_amd.PfastError):
        pfast_amd.source_order(np.zeros((3, 3), np.int32), mode=3)
