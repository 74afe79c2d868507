"""mtgp*_init and mtgp*_init_by_array computed on the GPU: the same windows
as the host init family (itself pinned to the reference's sequences) and
as the oracle, for every table, DC sets, many generators, and key lengths
below, at and above the state size."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TABLES = [(32, 11213), (32, 23209), (32, 44497), (64, 23209), (64, 44497), (64, 110503)]


def _specs():
    specs = []
    for t, (bits, mexp) in enumerate(TABLES):
        specs += [(bits, mexp, i) for i in (0, 1, 77, 127)]
    return specs


def test_device_init_seed_matches_host_and_oracle(mtgp, oracle):
    specs = _specs() * 20  # 480 generators
    rng = np.random.default_rng(3)
    seeds = [int(v) for v in rng.integers(0, 2**63, len(specs))]
    host = mtgp.GenSet(specs)
    host.init_seed(seeds)
    dev = mtgp.GenSet(specs)
    dev.init_seed(seeds, device=True)
    for g in range(len(specs)):
        assert (dev.state(g) == host.state(g)).all(), (g, specs[g])
    for g in range(0, len(specs), 37):
        bits, mexp, idx = specs[g]
        p = oracle.params(bits, mexp, idx)
        s = seeds[g] & 0xFFFFFFFF if bits == 32 else seeds[g]
        assert (dev.state(g) == oracle.init_state(p, s)).all(), g


def test_device_init_array_matches_host_and_oracle(mtgp, oracle):
    specs = _specs()
    rng = np.random.default_rng(4)
    keys = []
    for g, (bits, mexp, idx) in enumerate(specs):
        n = mexp // bits + 1
        length = [0, 1, 4, n - 1, n, n + 3, 2 * n][g % 7]
        keys.append([int(v) for v in rng.integers(0, 2**63, length)])
    host = mtgp.GenSet(specs)
    host.init_array(keys)
    dev = mtgp.GenSet(specs)
    dev.init_array(keys, device=True)
    for g, (bits, mexp, idx) in enumerate(specs):
        assert (dev.state(g) == host.state(g)).all(), (g, specs[g], len(keys[g]))
        p = oracle.params(bits, mexp, idx)
        k = [v & 0xFFFFFFFF for v in keys[g]] if bits == 32 else keys[g]
        assert (dev.state(g) == oracle.init_by_array(p, k)).all(), (g, len(k))


def test_device_init_dc_sets(mtgp):
    recs = [mtgp.dc_params(32, i) for i in (0, 16, 199)]
    host, dev = mtgp.GenSet(recs), mtgp.GenSet(recs)
    host.init_seed([7, 8, 9])
    dev.init_seed([7, 8, 9], device=True)
    for g in range(3):
        assert (dev.state(g) == host.state(g)).all()
