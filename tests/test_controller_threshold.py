"""The rounding claim behind the solve kernel's tiny-error path (CPU, IEEE
double as on the GPU).

The plain RKF78 controller (Boost.Odeint's rules: accept, and grow by
0.9 max(err, 5^-8)^(-1/8) for err < 0.5) only needs to know whether the error
ratio err = max_i |e_i| / sc_i is at most 5^-8.  The kernel decides that
without dividing when sc_i >= 2^-1000 and |e_i| < fl(sc_i * K), with
K = fl(fl(1/390625) * (1 - 2^-50)).  This holds only if every admitted |e_i|
gives fl(|e_i| / sc_i) <= fl(1/390625); checked here at the boundary (the
largest admitted |e_i|) over scales spanning the admitted range, and shown to
fail without the margin."""
import numpy as np

CAP = 1.0 / 390625.0
K = CAP * (1.0 - 2.0**-50)


def _scales(n=1_000_000, seed=17):
    rng = np.random.default_rng(seed)
    sc = np.ldexp(1.0 + rng.random(n), rng.integers(-1000, 1000, n))
    powers = np.ldexp(1.0, np.arange(-1000, 1001))
    return np.concatenate([sc, powers, np.nextafter(powers[1:], 0.0), np.nextafter(powers, np.inf)])


def _largest_admitted(sc, k):
    lim = sc * k
    return np.nextafter(lim, 0.0), lim


def test_admitted_errors_round_to_at_most_the_growth_threshold():
    sc = _scales()
    e, lim = _largest_admitted(sc, K)
    assert np.all(lim >= np.finfo(np.float64).tiny)  # sc * K is a normal number
    assert np.all(e < lim)
    assert np.all(e / sc <= CAP)
    # the threshold admits practically everything below the cap
    assert np.all(e / sc >= CAP * (1.0 - 2.0**-48))


def test_threshold_without_its_margin_is_caught():
    sc = _scales()
    # dropping the margin and admitting the boundary (|e| <= sc * 5^-8), or the
    # margin with the wrong sign, admits quotients above the cap
    lim = sc * CAP
    assert np.any(lim / sc > CAP)
    e, _ = _largest_admitted(sc, CAP * (1.0 + 2.0**-50))
    assert np.any(e / sc > CAP)


def test_zero_nan_and_infinite_scales_are_not_admitted():
    with np.errstate(invalid="ignore", over="ignore"):
        for sc, e in ((0.0, 0.0), (np.inf, np.inf), (np.nan, 0.0), (1.0, np.nan), (2.0**-1001, 0.0)):
            admitted = sc >= 2.0**-1000 and abs(e) < sc * K
            assert not admitted, (sc, e)
        # an infinite scale with a finite error is admitted, and the exact ratio is 0 too
        assert np.inf >= 2.0**-1000 and 1.0 < np.inf * K and 1.0 / np.inf == 0.0
