"""Oracle RNG, multiplier table and stream seeding (CPU).

How each part is pinned (the reference holds no vectors for this path; its multiplier file
safeprimes_base32.txt is not in the repository):
- MWC step, gpumcml_rng/gpumcml_rng.cu:40: Python integer arithmetic of the recurrence.
- rand_MWC_co (:41): round-toward-zero of a u32 to binary32 built in Python from the IEEE
  definition (largest float32 <= u), then an exact 2^-32 scale; rand_MWC_oc (:53): float32
  1 - co with round-to-nearest (numpy float32 arithmetic).
- Multipliers: the defining property (a*2^32-1 and a*2^31-1 prime) re-checked by a Python
  Miller-Rabin, plus a gap check that no valid multiplier was skipped at the top of the range.
- init_RNG (:59-118): every stream is rebuilt in Python from the begin-multiplier MWC stream (the
  recurrence pinned above): the carry from one output, floored after a binary64 product, the value
  from the next, and the pairs that break the state invariants of :83-85 drawn again.  Tiny
  multipliers force the rejections; a seed built to put u*a/2^32 just below an integer separates
  binary64 rounding from an exact or binary32 floor.  Product and oracle must equal that bit for bit.
"""
import json
import math
import os

import numpy as np
import pytest

import ymc3d_amd as y
from oracle import oracle as O

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))
M32 = (1 << 32) - 1


def rz_float32(u: int) -> np.float32:
    """Largest binary32 value <= u (u a non-negative integer)."""
    f = np.float32(u)
    if int(f) > u:
        f = np.nextafter(f, np.float32(0))
    return f


def py_stream(x: int, a: int, n: int):
    co, oc = [], []
    for _ in range(n):
        x = (x & M32) * a + (x >> 32)
        c = np.float32(rz_float32(x & M32) * np.float32(2.0 ** -32))
        co.append(c)
        oc.append(np.float32(np.float32(1.0) - c))
    return np.array(co, np.float32), np.array(oc, np.float32), x


def test_mwc_closed_form():
    rng = np.random.default_rng(1)
    tab = GOLDEN["multipliers_first_65"]
    xs = [int(v) for v in rng.integers(1, 2 ** 62, 40, dtype=np.uint64)]
    a = [tab[1 + i % 64] for i in range(40)]
    co, xf = O.mwc_fill(np.array(xs, np.uint64), np.array(a, np.uint32), 50, 0)
    oc, _ = O.mwc_fill(np.array(xs, np.uint64), np.array(a, np.uint32), 50, 1)
    for s in range(40):
        pco, poc, px = py_stream(xs[s], a[s], 50)
        assert np.array_equal(co[:, s].view(np.uint32), pco.view(np.uint32))
        assert np.array_equal(oc[:, s].view(np.uint32), poc.view(np.uint32))
        assert int(xf[s]) == px


def test_rz_conversion_edges():
    lib = O.lib()
    cases = [0, 1, 2, (1 << 24) - 1, 1 << 24, (1 << 24) + 1, (1 << 24) + 3, (1 << 25) + 3, 0x7FFFFFFF, 0x80000000,
             0x80000001, 0xFFFFFF7F, 0xFFFFFF80, 0xFFFFFFFF]
    rng = np.random.default_rng(2)
    cases += [int(v) for v in rng.integers(0, 2 ** 32, 2000, dtype=np.uint64)]
    for u in cases:
        got = np.float32(lib.ymc_oracle_u32_to_float_rz(u))
        assert got == rz_float32(u), u
        assert int(got) <= u


def test_draw_ranges():
    x, a = y.seed_streams(np.array(GOLDEN["multipliers_first_65"], np.uint32), 7, 64)
    co, _ = O.mwc_fill(x, a, 2000, 0)
    oc, _ = O.mwc_fill(x, a, 2000, 1)
    assert co.min() >= 0.0 and co.max() < 1.0
    assert oc.min() > 0.0 and oc.max() <= 1.0
    # uniformity: mean and variance of 128k draws
    assert abs(co.mean() - 0.5) < 0.005 and abs(co.var() - 1 / 12) < 0.002


def test_mwc_golden():
    g = GOLDEN["mwc"]
    co, xf = O.mwc_fill(np.array(g["x0"], np.uint64), np.array(g["a"], np.uint32), 16, 0)
    oc, _ = O.mwc_fill(np.array(g["x0"], np.uint64), np.array(g["a"], np.uint32), 16, 1)
    assert co.view(np.uint32).T.tolist() == g["co_bits"]
    assert oc.view(np.uint32).T.tolist() == g["oc_bits"]
    assert [int(v) for v in xf] == g["x_final"]


def _is_prime(n: int) -> bool:
    if n < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for b in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        x = pow(b, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def _is_multiplier(a: int) -> bool:
    return _is_prime(a * 2 ** 31 - 1) and _is_prime(a * 2 ** 32 - 1)


def test_multiplier_table_property(multipliers):
    tab = [int(v) for v in multipliers]
    assert len(tab) == (1 << 20) + 1
    assert all(tab[i] > tab[i + 1] for i in range(len(tab) - 1))  # descending, unique
    rng = np.random.default_rng(3)
    sample = list(range(200)) + [int(i) for i in rng.integers(0, len(tab), 300)]
    for i in sample:
        assert _is_multiplier(tab[i]), tab[i]
        assert O.lib().ymc_oracle_is_multiplier(tab[i]) == 1
    assert tab[:65] == GOLDEN["multipliers_first_65"]


def test_multiplier_table_has_no_gaps_at_the_top(multipliers):
    tab = [int(v) for v in multipliers[:40]]
    skipped = [a for a in range(tab[0] + 1, 2 ** 32) if _is_multiplier(a)]
    assert skipped == []
    for hi, lo in zip(tab, tab[1:]):
        assert not any(_is_multiplier(a) for a in range(lo + 1, hi))


def test_oracle_multiplier_check_rejects_composites():
    assert O.lib().ymc_oracle_is_multiplier(4294967118) == 1
    assert O.lib().ymc_oracle_is_multiplier(4294967117) == _is_multiplier(4294967117)
    assert O.lib().ymc_oracle_is_prime(2 ** 61 - 1) == 1 and O.lib().ymc_oracle_is_prime(2 ** 61 + 1) == 0


def _check_invariants(x, a):
    c = x >> np.uint64(32)
    lo = x & np.uint64(M32)
    assert np.all(x != 0)
    assert np.all(c < a.astype(np.uint64) - 1)
    assert np.all(lo != M32)


@pytest.mark.parametrize("seed", [1, 12345, 987654321, (7 << 32) | 99, 2 ** 40 + 17])
def test_seeding_product_equals_oracle(multipliers, seed):
    tab = multipliers[:1025]
    x, a = y.seed_streams(tab, seed, 5000)  # cycles through the table ~5 times
    st, xo, ao = O.init_rng(tab, seed, 5000)
    assert st == 0
    assert np.array_equal(x, xo) and np.array_equal(a, ao)
    _check_invariants(x, a)
    assert np.array_equal(a[:1024], tab[1:1025]) and np.array_equal(a[1024:2048], tab[1:1025])


def test_seeding_rejection_branch():
    """Tiny multipliers make the rejection loop (gpumcml_rng.cu:100) fire often."""
    tab = np.array([13, 3, 5, 7, 4], np.uint32)
    for seed in (1, 2, 3, 5, 11):
        x, a = y.seed_streams(tab, seed, 400)
        st, xo, ao = O.init_rng(tab, seed, 400)
        assert st == 0 and np.array_equal(x, xo) and np.array_equal(a, ao)
        _check_invariants(x, a)


def test_seeding_rejects_invalid_seeds(multipliers):
    tab = multipliers[:9]
    begin = int(tab[0])
    for bad in (0, (begin - 1) << 32, 0xFFFFFFFF, ((begin + 5) << 32) | 3):
        with pytest.raises(y.YmcError) as ei:
            y.seed_streams(tab, bad, 4)
        assert ei.value.status == y._abi.ERR_SEED
        assert O.init_rng(tab, bad, 4)[0] == y._abi.ERR_SEED


def test_seeding_golden():
    g = GOLDEN["seeding"]
    tab = np.array(GOLDEN["multipliers_first_65"], np.uint32)
    x, a = y.seed_streams(tab, g["seed"], 100)
    assert [int(v) for v in x] == g["x"] and [int(v) for v in a] == g["a"]


def test_multiplier_text_reader(tmp_path):
    p = tmp_path / "safeprimes_base32.txt"
    p.write_text("4294967118 0 0\n4294966893 1 2\n4294966830 3 4\n")
    tab = y.read_multipliers_text(str(p))
    assert tab.tolist() == [4294967118, 4294966893, 4294966830]
    (tmp_path / "bad.txt").write_text("12\n")
    with pytest.raises(y.YmcError):
        y.read_multipliers_text(str(tmp_path / "bad.txt"))
    with pytest.raises(y.YmcError) as ei:
        y.read_multipliers_text(str(tmp_path / "missing.txt"))
    assert ei.value.status == y._abi.ERR_IO


# ---------------------------------------------------------------------------------------------
# init_RNG (gpumcml_rng.cu:88-112) tied to the begin-multiplier MWC stream in Python integers.
#
# The begin stream is the MWC recurrence pinned above (multiplier table[0], state = the seed).
# Stream i takes the next two outputs u, v of that stream: carry c = floor((u / 2^32) * a_i) with
# the product rounded to the nearest binary64 (the C expression of :106 is evaluated in double;
# Python floats are IEEE binary64 with round-to-nearest), value x = v (:109-110); the pair is
# drawn again while the state violates 0 < state, c < a_i - 1, x < 2^32 - 1 (:100).


def _begin_outputs(seed: int, begin: int):
    x = seed
    while True:
        x = (x & M32) * begin + (x >> 32)
        yield x & M32


def _py_init_rng(table, seed: int, n: int, carry_rule=None, swap=False):
    begin = int(table[0])
    outs = _begin_outputs(seed, begin)
    carry_rule = carry_rule or (lambda u, a: int(math.floor((float(u) / 4294967296.0) * float(a))))
    xs, As, rejected = [], [], 0
    for i in range(n):
        a = int(table[1 + i % (len(table) - 1)])
        while True:
            u, v = next(outs), next(outs)
            if swap:
                u, v = v, u
            c = carry_rule(u, a) & M32
            state = (c << 32) | v
            if state != 0 and c < a - 1 and v < M32:
                break
            rejected += 1
        xs.append(state)
        As.append(a)
    return np.array(xs, np.uint64), np.array(As, np.uint32), rejected


@pytest.mark.parametrize("seed", [1, 12345, (7 << 32) | 99, 2 ** 40 + 17])
def test_seeding_is_the_begin_stream_split_into_carry_and_value(multipliers, seed):
    tab = [int(v) for v in multipliers[:257]]
    want_x, want_a, _ = _py_init_rng(tab, seed, 600)
    x, a = y.seed_streams(np.array(tab, np.uint32), seed, 600)
    st, xo, ao = O.init_rng(np.array(tab, np.uint32), seed, 600)
    assert st == 0
    assert np.array_equal(xo, want_x) and np.array_equal(ao, want_a)
    assert np.array_equal(x, want_x) and np.array_equal(a, want_a)
    # the plausible slips give other states: c and x drawn in the other order, the carry floored
    # from a binary32 product
    swapped, _, _ = _py_init_rng(tab, seed, 600, swap=True)
    single = lambda u, a: int(math.floor(float(np.float32(np.float32(u) / np.float32(4294967296.0)) * np.float32(a))))
    f32, _, _ = _py_init_rng(tab, seed, 600, carry_rule=single)
    assert np.count_nonzero(swapped != xo) > 590
    assert np.count_nonzero(f32 != xo) > 300


def test_seeding_rejections_follow_the_begin_stream():
    """Tiny multipliers reject most pairs (gpumcml_rng.cu:100): the streams must skip exactly those."""
    tab = [13, 3, 5, 7, 4]
    for seed in (1, 2, 3, 5, 11):
        want_x, want_a, rejected = _py_init_rng(tab, seed, 400)
        assert rejected > 100
        st, xo, ao = O.init_rng(np.array(tab, np.uint32), seed, 400)
        x, a = y.seed_streams(np.array(tab, np.uint32), seed, 400)
        assert st == 0
        assert np.array_equal(xo, want_x) and np.array_equal(ao, want_a)
        assert np.array_equal(x, want_x) and np.array_equal(a, want_a)


def _seed_with_first_output(u: int, begin: int) -> int:
    """A valid init_RNG seed whose begin stream's first output is u."""
    for lo in range(1, 1 << 20):
        hi = (u - lo * begin) % (1 << 32)
        if hi < begin - 1:
            return (hi << 32) | lo
    raise AssertionError("no seed found")


def test_seeding_carry_is_floored_after_binary64_rounding(multipliers):
    """u * a / 2^32 just below an integer k: the binary64 product of :106 rounds to k, so the carry
    is k; the exact quotient would floor to k - 1."""
    begin, a = int(multipliers[0]), int(multipliers[1])
    assert a % 2 == 1
    inv = pow(a, -1, 1 << 32)
    u = ((1 << 32) - 1) * inv % (1 << 32)  # u * a = -1 (mod 2^32)
    exact = (u * a) >> 32
    rounded = int(math.floor((float(u) / 4294967296.0) * float(a)))
    assert rounded == exact + 1 and exact + 1 < a - 1
    seed = _seed_with_first_output(u, begin)
    tab = np.array([begin, a], np.uint32)
    st, xo, _ = O.init_rng(tab, seed, 1)
    x, _ = y.seed_streams(tab, seed, 1)
    assert st == 0 and int(x[0]) == int(xo[0])
    assert int(xo[0]) >> 32 == exact + 1
    want, _, _ = _py_init_rng([begin, a], seed, 1)
    assert int(want[0]) == int(xo[0])
