"""Pins the shared scalar primitives (hdsr_fmath.h) against float64 libm and numpy.

The kernels and the oracle both use these functions, so their agreement proves nothing about
them; these tests do: exp/ln against float64 within 2 ulp over the ranges the path uses and
beyond, binary16 packing against numpy's IEEE round-to-nearest-even conversion (including
subnormals, ties, overflow and NaN), the WGSL float->u32 saturation, and the Morton code as a
bijection against a bit-by-bit definition.
"""
import numpy as np
import pytest

from oracle import oracle


def ulp_err(approx32: np.ndarray, exact64: np.ndarray) -> np.ndarray:
    a = approx32.astype(np.float64)
    spacing = np.spacing(np.abs(exact64).astype(np.float32)).astype(np.float64)
    return np.abs(a - exact64) / spacing


def run_n(fn, x):
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty_like(x)
    fn(x.ctypes.data, y.ctypes.data, x.size)
    return y


@pytest.fixture(scope="module")
def lib():
    return oracle.load()


def test_expf_within_2ulp(lib):
    rng = np.random.default_rng(0)
    x = np.concatenate([
        np.linspace(-4.0, 0.5, 200_001, dtype=np.float32),      # composite power range
        rng.normal(-4.0, 0.5, 100_000).astype(np.float32),      # log-scales
        np.linspace(-86.9, 87.9, 100_001, dtype=np.float32),
        np.array([0.0, -0.0, 1.0, -1.0, 1e-30, -1e-30, 0.34657359, -0.34657359], np.float32),
    ])
    y = run_n(lib.orc_expf_n, x)
    e = ulp_err(y, np.exp(x.astype(np.float64)))
    assert e.max() <= 2.0, f"max ulp {e.max()} at x={x[np.argmax(e)]}"
    assert lib.orc_expf(0.0) == 1.0


def test_expf_gauss_within_2ulp(lib):
    """The compositor's exponential (rounding constant for k, 2^k added to the exponent bits): within
    2 ulp of float64 exp over the power range it is read on and its whole stated range, exact at 0,
    and equal to hdsr_expf except where x * log2(e) lies within rounding of a half-integer."""
    rng = np.random.default_rng(3)
    x = np.concatenate([
        np.linspace(-4.0, 0.0, 400_001, dtype=np.float32),      # the powers it is read on
        -rng.exponential(1.0, 200_000).astype(np.float32),
        np.linspace(-86.0, 88.0, 200_001, dtype=np.float32),
        np.array([0.0, -0.0, -1e-30, -0.34657359, 0.34657359, -3.4657359], np.float32),
    ])
    y = run_n(lib.orc_expf_gauss_n, x)
    e = ulp_err(y, np.exp(x.astype(np.float64)))
    assert e.max() <= 2.0, f"max ulp {e.max()} at x={x[np.argmax(e)]}"
    assert run_n(lib.orc_expf_gauss_n, np.zeros(1))[0] == 1.0
    z = run_n(lib.orc_expf_n, x)
    diff = x[y != z].astype(np.float64)
    assert diff.size < 1e-4 * x.size
    t = diff * np.log2(np.e)  # the differences sit where k's two roundings can part
    assert np.all(np.abs(np.abs(t - np.floor(t)) - 0.5) < 1e-6 * np.maximum(1.0, np.abs(t)))


def test_expf_special_values(lib):
    x = np.array([np.nan, 100.0, 88.5, -100.0, -87.5, np.inf, -np.inf], np.float32)
    y = run_n(lib.orc_expf_n, x)
    assert np.isnan(y[0])
    assert y[1] == np.inf and y[2] == np.inf and y[5] == np.inf
    assert y[3] == 0.0 and y[4] == 0.0 and y[6] == 0.0


def test_logf_within_2ulp(lib):
    rng = np.random.default_rng(1)
    x = np.concatenate([
        np.geomspace(1e-6, 1.0, 200_001).astype(np.float32),   # NDC depths
        rng.uniform(0.5, 2.0, 100_000).astype(np.float32),
        np.geomspace(1e-37, 3e38, 100_001).astype(np.float32),
        np.array([1.0, 2.0, 0.5, 1.4142135, 1.4142137, 0.70710677, 1e-6], np.float32),
    ])
    y = run_n(lib.orc_logf_n, x)
    ex = np.log(x.astype(np.float64))
    nz = ex != 0
    e = ulp_err(y[nz], ex[nz])
    assert e.max() <= 2.0, f"max ulp {e.max()} at x={x[nz][np.argmax(e)]}"
    assert lib.orc_logf(1.0) == 0.0


def test_logf_special_and_subnormal(lib):
    x = np.array([0.0, -1.0, np.inf, np.nan, 1e-40, 1.4e-45], np.float32)
    y = run_n(lib.orc_logf_n, x)
    assert y[0] == -np.inf and np.isnan(y[1]) and y[2] == np.inf and np.isnan(y[3])
    for i in (4, 5):
        assert abs(float(y[i]) - np.log(float(x[i]))) <= 2 * np.spacing(np.float32(abs(np.log(float(x[i])))))


def test_logf_monotone_on_depths(lib):
    # per-tile keys rely on ln being non-decreasing over the depth range
    x = np.sort(np.random.default_rng(2).uniform(1e-6, 1.0, 500_000).astype(np.float32))
    y = run_n(lib.orc_logf_n, x)
    assert np.all(np.diff(y) >= 0)


def test_f16_rne_matches_numpy(lib):
    rng = np.random.default_rng(3)
    specials = np.array([0.0, -0.0, 1.0, -1.0, 65504.0, 65519.99, 65520.0, 1e6, -1e6, np.inf, -np.inf,
                         6.1035156e-05, 6.0975552e-05, 5.9604645e-08, 2.9802322e-08, 2.9802326e-08,
                         1.0009765625, 1.00048828125, 1.00146484375, 0.99, 0.3, 1 / 255], np.float32)
    x = np.concatenate([specials, rng.normal(0, 1, 200_000).astype(np.float32),
                        rng.uniform(-7e4, 7e4, 100_000).astype(np.float32),
                        (rng.uniform(0, 1, 100_000) * 1e-4).astype(np.float32),
                        rng.integers(0, 2**32, 200_000, dtype=np.uint64).astype(np.uint32).view(np.float32)])
    x = x[~np.isnan(x)]
    y = np.empty(x.size, np.uint32)
    lib.orc_f16_bits_n(np.ascontiguousarray(x).ctypes.data, y.ctypes.data, x.size)
    ref = x.astype(np.float16).view(np.uint16).astype(np.uint32)
    bad = np.nonzero(y != ref)[0]
    assert bad.size == 0, f"f16 mismatch at {x[bad[:5]]}: {y[bad[:5]]} vs {ref[bad[:5]]}"
    nan = lib.orc_f16_bits(np.float32(np.nan))
    assert (nan & 0x7C00) == 0x7C00 and (nan & 0x3FF) != 0


def test_f16_decode_exhaustive(lib):
    h = np.arange(65536, dtype=np.uint32)
    ref = h.astype(np.uint16).view(np.float16).astype(np.float32)
    got = np.array([lib.orc_f16_to_f32(int(v)) for v in h[::7]], np.float32)
    r = ref[::7]
    ok = (got == r) | (np.isnan(got) & np.isnan(r))
    assert ok.all()


def test_u32sat(lib):
    cases = {0.0: 0, -1.0: 0, 0.99: 0, 1.0: 1, 3.7: 3, 4294967040.0: 4294967040, 4294967296.0: 4294967295,
             1e20: 4294967295, float("nan"): 0, float("-inf"): 0, float("inf"): 4294967295}
    for x, want in cases.items():
        assert lib.orc_u32sat(x) == want, x


def test_q20_with_folded_scale(lib):
    # the tile-list write multiplies by scale * 2^13 at once; the key must equal
    # min(u32sat((d * scale) * 2^13), 2^20 - 1), two roundings, as the sort stages define it
    rng = np.random.default_rng(3)
    n = 400_000
    d = (rng.choice([-1.0, 1.0], n) * 10.0 ** rng.uniform(-45, 39, n)).astype(np.float32)
    d[: n // 4] = rng.uniform(-1.0, 129.0, n // 4).astype(np.float32)  # ln-depth offsets of a tile
    sc = (10.0 ** rng.uniform(0.0, np.log10(1.28e12), n)).astype(np.float32)
    sc[::7] = 0.0
    sc[1::11] = np.float32(1.28e12)
    d[2::13] = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1e-45, -1e-45, 3.4e38, 2.0 ** -140,
                         127.99999, 128.0, 1048575.0 / 8192.0, 1048576.0 / 8192.0], np.float32)[
        np.arange(len(d[2::13])) % 13]
    out = np.empty(n, np.uint32)
    lib.orc_q20_folded_n(d.ctypes.data, sc.ctypes.data, out.ctypes.data, n)
    with np.errstate(over="ignore", invalid="ignore", under="ignore"):
        y = (d * sc).astype(np.float32) * np.float32(8192.0)
        want = np.where(~(y > 0), 0.0, np.where(y >= 4294967296.0, 4294967295.0, np.trunc(np.nan_to_num(y))))
    want = np.minimum(want, 0xFFFFF).astype(np.uint32)
    bad = np.flatnonzero(out != want)
    assert bad.size == 0, (d[bad[:5]], sc[bad[:5]], out[bad[:5]], want[bad[:5]])


def test_morton_bijection(lib):
    def ref(x, y):
        out = 0
        for b in range(16):
            out |= ((x >> b) & 1) << (2 * b)
            out |= ((y >> b) & 1) << (2 * b + 1)
        return out

    rng = np.random.default_rng(4)
    for x, y in [(0, 0), (1, 0), (0, 1), (255, 255), (39, 29), (119, 67), (239, 134), (65535, 65535)]:
        assert lib.orc_morton2(x, y) == ref(x, y)
    for x, y in rng.integers(0, 65536, (2000, 2)):
        assert lib.orc_morton2(int(x), int(y)) == ref(int(x), int(y))
    seen = {lib.orc_morton2(x, y) for x in range(64) for y in range(64)}
    assert len(seen) == 64 * 64 and max(seen) == 64 * 64 - 1
