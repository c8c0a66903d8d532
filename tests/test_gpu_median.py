"""The energy median on the GPU (ofx_median: the radix select the engine runs every inner
iteration) against a sort, bit for bit, on the cases of the reference's own tests
(rust/crates/optflow-core/src/select.rs:173-285): spread values, heavy duplicates, signed zeros,
tiny and nearly equal values, two separated halves, all-equal arrays, NaN at any position, and the
sign of zero in the mean of the two middle values."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    import optflow_soil_amd as ofx
    e = ofx.Engine(0)
    yield e
    e.close()


def lcg(seed=99):
    s = seed

    def rnd():
        nonlocal s
        s = (s * 6364136223846793005 + 1442695040888963407) % (1 << 64)
        return (s >> 11) / float(1 << 53)
    return rnd


def sorted_median(v):
    """The median of a sort by total order (np.sort orders -0.0 and 0.0 as equal, so sort the keys)."""
    b = v.view(np.int64)
    key = np.where(b < 0, ~b, b | np.int64(-2 ** 63)).view(np.uint64)
    s = v[np.argsort(key, kind="stable")]
    n = len(s)
    return s[n // 2] if n % 2 else (s[n // 2 - 1] + s[n // 2]) / 2.0


def bits(x):
    return np.float64(x).view(np.uint64)


def reference_sets():
    rnd = lcg()
    r = lambda n: np.array([rnd() for _ in range(n)])
    sets = {
        "spread": r(200_001) * 3.0 - 1.0,
        "duplicates": np.floor(r(200_000) * 7.0) - 2.0,
        "signed zeros and tiny": np.array([(0.0, -0.0, 1e-300 * rnd(), -rnd())[i % 4] for i in range(70_000)]),
        "nearly equal": 1.0 + r(65_537) * 1e-12,
        "even, middle distinct": np.array([(i % 1000) + 0.5 for i in range(262_144)]),
        "two halves": np.array([rnd() if i % 2 == 0 else 1e3 + rnd() for i in range(80_000)]),
        "zero and two": np.array([-0.0 if i % 2 == 0 else 2.0 for i in range(80_000)]),
        "all equal": np.full(70_000, 7.0),
        "one": np.array([2.5]),
        "odd small": np.array([3.0, 1.0, 2.0]),
        "even small": np.array([4.0, 1.0, 3.0, 2.0]),
        "ties": np.array([2.0, 2.0, 2.0, 9.0]),
    }
    return sets


@pytest.mark.parametrize("name", list(reference_sets()))
def test_median_equals_the_sort(eng, name):
    v = reference_sets()[name]
    got = eng.median(v)
    want = sorted_median(v)
    print(f"median {name} n={len(v)}: {got!r} {want!r}", flush=True)
    assert bits(got) == bits(want)
    assert got == np.median(v) or (np.isnan(got) and np.isnan(np.median(v)))


def test_median_sign_of_zero(eng):
    """select.rs:269-285: the mean of -0.0 and -0.0 is -0.0, any other pair with a 0.0 is 0.0."""
    assert bits(eng.median(np.array([-0.0, -0.0]))) == bits(-0.0)
    assert bits(eng.median(np.array([-0.0, 0.0]))) == bits(0.0)
    assert bits(eng.median(np.array([0.0, -0.0, 0.0, 0.0]))) == bits(0.0)
    assert bits(eng.median(np.array([-1.0, -0.0, -0.0, 5.0]))) == bits(-0.0)


@pytest.mark.parametrize("n", [7, 1000, 65_537, 4 * 2048 * 2048 // 4 + 1])
def test_median_sees_a_nan_anywhere(eng, n):
    """select.rs:254-267 and :300-330: a single NaN at the start, the middle or the end makes the
    median NaN; without it the median is finite; NaNs in several places count together."""
    v = ((np.arange(n) * 7919) % 1009).astype(np.float64)
    assert not np.isnan(eng.median(v))
    for at in (0, n // 2, n - 1):
        w = v.copy()
        w[at] = np.nan
        assert np.isnan(eng.median(w)), at
    w = v.copy()
    w[::max(1, n // 16)] = np.nan
    assert np.isnan(eng.median(w))


def test_median_of_nothing_is_nan(eng):
    assert np.isnan(eng.median(np.array([], np.float64)))


def test_median_at_frame_size(eng):
    """A 2048 x 2048 energy-like array (positive, heavy-tailed) and its negation, odd and even."""
    rng = np.random.default_rng(3)
    v = rng.lognormal(0.0, 2.0, 2048 * 2048)
    for arr in (v, -v, v[:-1]):
        assert bits(eng.median(arr)) == bits(sorted_median(arr))
