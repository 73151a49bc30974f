"""The batching form's grouping (gmc_batch_plan, host-only): the systems
sorted by decreasing order and cut into contiguous groups of at most
256 / nalpha systems, padded to a multiple of 8 above the group's largest
order, minimising the summed cost model c0 + c1 npad + c2 B npad^2 +
c3 B npad^3 over the groups.  Checked against brute force over every cut of
small inputs, and on the 64-system workload's shape."""
import itertools

import numpy as np
import pytest


def _cost(c, nd_sorted, nalpha, cut):
    tot = 0.0
    for i, j in zip(cut[:-1], cut[1:]):
        npad = -(-nd_sorted[i] // 8) * 8
        b = (j - i) * nalpha
        tot += c[0] + c[1] * npad + b * npad ** 2 * (c[2] + c[3] * npad)
    return tot


def _brute(nd, nalpha, c):
    nd_sorted = sorted(nd, reverse=True)
    m, gmax = len(nd), max(1, 256 // nalpha)
    best = None
    for k in range(m):
        for inner in itertools.combinations(range(1, m), k):
            cut = (0,) + inner + (m,)
            if any(j - i > gmax for i, j in zip(cut[:-1], cut[1:])):
                continue
            v = _cost(c, nd_sorted, nalpha, cut)
            if best is None or v < best[0]:
                best = (v, cut)
    return best


def _plan_cost(nd, nalpha, c, grp, npad):
    tot = 0.0
    for g in range(grp.max() + 1):
        sel = grp == g
        b = int(sel.sum()) * nalpha
        p = int(npad[sel][0])
        assert np.all(npad[sel] == p) and p % 8 == 0 and p >= max(np.asarray(nd)[sel])
        tot += c[0] + c[1] * p + b * p ** 2 * (c[2] + c[3] * p)
    return tot


@pytest.mark.parametrize("seed", range(6))
def test_plan_is_optimal_against_brute_force(seed):
    from gammcor_amd import host
    rng = np.random.default_rng(seed)
    m = int(rng.integers(2, 9))
    nd = rng.integers(20, 2000, m).tolist()
    nalpha = int(rng.choice([16, 64, 128]))
    c = np.array([rng.uniform(0.5, 5), rng.uniform(0, 0.1), rng.uniform(0, 2e-6), rng.uniform(1e-10, 1e-9)])
    grp, npad = host.batch_plan(nd, nalpha, c)
    best, cut = _brute(nd, nalpha, c)
    assert abs(_plan_cost(nd, nalpha, c, grp, npad) - best) <= 1e-9 * best
    # groups are contiguous runs of the order sorted by decreasing size
    order = np.argsort(-np.asarray(nd), kind="stable")
    assert np.all(np.diff(grp[order]) >= 0)


def test_plan_of_the_64_system_workload():
    """The workload's orders (NDimX of the 64 systems) under the measured
    cost model: every group within 16 systems and padded above its members."""
    import accas_oracle as O
    from gammcor_amd import host, synth
    nd = []
    for cfg in synth.batch_configs():
        s = synth.make_system(with_integrals=False, **cfg)
        nd.append(len(O.select_pairs(s.occ, s.nele)))
    nd = np.asarray(nd)
    grp, npad = host.batch_plan(nd, 16)
    assert np.all(np.bincount(grp) <= 16)
    assert np.all(npad >= nd) and np.all(npad % 8 == 0)
    for g in range(grp.max() + 1):
        assert npad[grp == g][0] == -(-nd[grp == g].max() // 8) * 8
    print(f"{grp.max() + 1} groups, padded orders {sorted(set(npad.tolist()), reverse=True)}")


def test_cost_model_needs_four_coefficients():
    from gammcor_amd import host
    with pytest.raises(ValueError):
        host.batch_plan([100, 200], 16, [1.0, 0.0, 1e-9])
