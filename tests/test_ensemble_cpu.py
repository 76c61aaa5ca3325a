"""The uneven-ensemble workload on the CPU: its random numbers, and that each
domain is a valid cyclic MPDATA problem (divergence-free advector accepted by
the oracle's check, duplicated edge cells, conservation and positivity over
a step)."""
import numpy as np
import pytest

import libmpdataxx_amd as M
import oracle as O
from libmpdataxx_amd import workloads as W


def test_mt19937_64_matches_the_cpp_standard():
    # [rand.predef]: the 10000th consecutive invocation of a default-constructed
    # std::mt19937_64 produces 9981545732273789042
    g = W.MT19937_64()
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042


@pytest.fixture(scope="module")
def domains():
    return W.ensemble_domains()


def test_ensemble_sizes_and_options(domains):
    assert len(domains) == 64
    for d in domains[:32]:
        assert len(d["grid"]) == 2 and all(64 <= n <= 512 for n in d["grid"])
    for d in domains[32:]:
        assert len(d["grid"]) == 3 and all(32 <= n <= 128 for n in d["grid"])
    assert {d["n_iters"] for d in domains} == {2, 3, 4}
    assert {d["fct"] for d in domains} == {True, False}
    again = W.ensemble_domains()
    for a, b in zip(domains, again):
        assert a["grid"] == b["grid"] and np.array_equal(a["psi"], b["psi"])


@pytest.mark.parametrize("i", [0, 5, 31, 32, 40, 63])
def test_ensemble_domain_is_a_valid_cyclic_problem(domains, i):
    d = domains[i]
    nd = len(d["grid"])
    psi = d["psi"]
    for k in range(nd):
        first = [slice(None)] * nd
        last = [slice(None)] * nd
        first[k], last[k] = 0, -1
        np.testing.assert_array_equal(psi[tuple(first)], psi[tuple(last)])
    assert max(np.abs(g).max() for g in d["gc"]) <= 0.5 + 1e-12
    o = O.Oracle(d["grid"], opts=O.FCT if d["fct"] else 0, n_iters=d["n_iters"],
                 bcs=[(O.CYCLIC, O.CYCLIC)] * nd)
    o.set_advectee(psi)
    for k in range(nd):
        o.set_advector(d["gc"][k], k)
    o.advance(1)      # raises if the advector is divergent
    a = o.advectee()
    u = tuple(slice(0, n - 1) for n in d["grid"])
    assert abs(a[u].sum() - psi[u].sum()) <= 1e-12 * psi[u].sum()
    assert a.min() > 0


def test_ensemble_api_declared():
    for name in ("mpdx_ensemble_create", "mpdx_ensemble_advance", "mpdx_ensemble_tune", "mpdx_set_x_chunk"):
        assert name in M.EXPORTED
