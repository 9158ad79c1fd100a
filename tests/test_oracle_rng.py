"""Pins the oracle's MRG32k3a streams (RngStream.cpp restated in oracle/ffea_oracle.c).

- the known answer of the reference's own test (tests/consistency/rngstream/
  testRngStream.cpp:129, "39.697547445251") replayed call for call on the oracle;
- draw-for-draw equality with the reference's RngStream.cpp built under oracle/_ref
  (skipped where the reference tree was not available to build it), and with the
  committed golden fixture made from it (tests/golden/rng_streams.json);
- the glibc rand() sequence behind the seed derivation (World.cpp:272-276),
  computed independently from glibc's published TYPE_3 additive generator.
"""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "rng_streams.json")


def test_testrngstream_known_answer():
    """testRngStream.cpp:26-129, same calls on the oracle's streams"""
    O = oracle
    pk = O.Package()
    g1, g2, g3 = O.Rng(package=pk), O.Rng(package=pk), O.Rng(package=pk)
    s = g2.u01() + g3.u01()
    g1.advance(5, 3)
    s += g1.u01()
    g1.reset_start_stream()
    for _ in range(35):
        g1.advance(0, 1)
    s += g1.u01()
    g1.reset_start_stream()
    sumi = 0
    for _ in range(35):
        sumi += g1.rand_int(1, 10)
    s += sumi / 100.0
    sum3 = 0.0
    for _ in range(100):
        sum3 += g3.u01()
    s += sum3 / 10.0
    g3.reset_start_stream()
    for _ in range(5):
        s += g3.u01()
    for _ in range(4):
        g3.reset_next_substream()
    for _ in range(5):
        s += g3.u01()
    g3.reset_start_substream()
    for _ in range(5):
        s += g3.u01()
    g2.reset_next_substream()
    sum3 = 0.0
    for v in g2.fill(100000):
        sum3 += v
    s += sum3 / 10000.0
    g3.set_antithetic(True)
    sum3 = 0.0
    for v in g3.fill(100000):
        sum3 += v
    s += sum3 / 10000.0
    pk.set_seed([1, 1, 1, 1, 1, 1])
    gar = [O.Rng(package=pk) for _ in range(4)]
    for g in gar:
        s += g.u01()
    gar[2].advance(-127, 0)
    s += gar[2].u01()
    gar[2].reset_next_substream()
    gar[2].increased_precis(True)
    sum3 = 0.0
    for v in gar[2].fill(100000):
        sum3 += v
    s += sum3 / 10000.0
    gar[2].set_antithetic(True)
    sum3 = 0.0
    for v in gar[2].fill(100000):
        sum3 += v
    s += sum3 / 10000.0
    gar[2].set_antithetic(False)
    gar[2].increased_precis(False)
    for g in gar:
        s += g.u01()
    assert f"{s:.14g}" == "39.697547445251"


@pytest.mark.skipif(not os.path.exists(oracle.REF_TEST_RNG), reason="oracle/_ref not built (no reference tree)")
def test_reference_own_rng_test_passes():
    assert subprocess.run([oracle.REF_TEST_RNG]).returncode == 0


def _oracle_stream(seed, stream, substreams, n):
    pk = oracle.Package()
    assert pk.set_seed(seed)
    gs = [oracle.Rng(package=pk) for _ in range(stream + 1)]
    g = gs[stream]
    for _ in range(substreams):
        g.reset_next_substream()
    return g.fill(n), g.state()


CASES = [([12345] * 6, 0, 0), ([1, 2, 3, 4, 5, 6], 0, 0), ([4294967086, 7, 9, 4294944442, 11, 1], 3, 2),
         ([2024, 77, 5, 190, 2, 99999], 1, 1)]


@pytest.mark.skipif(oracle.ref_lib() is None, reason="oracle/_ref not built (no reference tree)")
@pytest.mark.parametrize("seed,stream,subs", CASES)
def test_streams_equal_reference_rngstream(seed, stream, subs):
    R = oracle.ref_lib()
    n = 2000
    out = np.zeros(n)
    st = np.zeros(6, np.uint32)
    assert R.ref_rng_stream_draws(np.array(seed, np.uint32), stream, subs, n, out, st) == 0
    mine, mst = _oracle_stream(seed, stream, subs, n)
    assert np.array_equal(out, mine)
    assert np.array_equal(st, mst)


@pytest.mark.skipif(oracle.ref_lib() is None, reason="oracle/_ref not built (no reference tree)")
@pytest.mark.parametrize("e,c", [(0, 7), (10, 0), (20, -3), (-5, 11), (70, 123456)])
def test_advance_state_equal_reference(e, c):
    R = oracle.ref_lib()
    seed = np.array([31337, 1, 2, 3, 4, 5], np.uint32)
    out = np.zeros(50)
    st = np.zeros(6, np.uint32)
    assert R.ref_rng_advance_draws(seed, e, c, 50, out, st) == 0
    g = oracle.Rng(state=seed)
    g.advance(e, c)
    assert np.array_equal(g.fill(50), out)
    assert np.array_equal(g.state(), st)


def test_streams_equal_golden_fixture():
    with open(GOLDEN) as f:
        gold = json.load(f)
    for case in gold["streams"]:
        mine, st = _oracle_stream(case["seed"], case["stream"], case["substreams"], len(case["draws_hex"]))
        assert [v.hex() for v in mine] == case["draws_hex"]
        assert st.tolist() == case["state_after"]


def _glibc_rand(seed, count):
    """glibc random(3) TYPE_3 (x^31 + x^3 + 1) as published: r[i] = 16807 r[i-1] mod
    (2^31 - 1) for i < 31, r[i] = r[i-31] for 31 <= i < 34, then r[i] = r[i-3] + r[i-31]
    mod 2^32; output r[i] >> 1 for i >= 344."""
    r = [0] * (344 + count)
    r[0] = seed if seed != 0 else 1
    for i in range(1, 31):
        hi, lo = divmod(r[i - 1], 127773)
        w = 16807 * lo - 2836 * hi
        r[i] = w + 2147483647 if w < 0 else w
    for i in range(31, 34):
        r[i] = r[i - 31]
    for i in range(34, 344 + count):
        r[i] = (r[i - 31] + r[i - 3]) & 0xFFFFFFFF
    return [v >> 1 for v in r[344:]]


@pytest.mark.parametrize("rng_seed", [1, 44, 2024, 123456789])
def test_sixseed_from_glibc_rand(rng_seed):
    rs = _glibc_rand(rng_seed, 12)
    m = [4294967087, 4294967087, 4294967087, 4294944443, 4294944443, 4294944443]
    expect = [((rs[2 * i] + rs[2 * i + 1]) & 0xFFFFFFFF) % m[i] for i in range(6)]
    assert oracle.sixseed(rng_seed).tolist() == expect
