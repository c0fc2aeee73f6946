"""GPU tests of the batched book (one persistent, graph-captured launch over many contracts).

Bars: every contract's result equals a per-contract mcdx_price_gbm call bit for bit, under
every schedule and after autotuning; graph replays interleaved with other GPU work return
the same bits every time; FP64 books match the oracle's prices to 1e-12 relative.
"""
import ctypes as C

import numpy as np
import pytest

import mcdx
import oracle_lib as O

pytestmark = pytest.mark.gpu
FP32, FP64 = mcdx.PREC_FP32, mcdx.PREC_FP64


def small_book():
    cs = mcdx.c4_book(n_contracts=96, seed=7, log2_min=8, log2_max=14)
    cs[0].antithetic = 1
    cs[1].product = mcdx.LOOKBACK
    cs[1].strike_style = 1
    cs[1].monitoring_points = 17
    cs[2].product = mcdx.BARRIER
    cs[2].barrier = 130.0
    cs[2].brownian_bridge = 1
    cs[2].monitoring_points = 40
    cs[3].product = mcdx.ASIAN
    cs[3].control_variate = 1
    cs[3].monitoring_points = 30
    return cs


@pytest.mark.parametrize("prec", [FP32, FP64])
def test_book_equals_per_contract_calls_bitwise(prec):
    cs = small_book()
    book = mcdx.Book(cs, prec)
    got = book.price()
    for c, g in zip(cs, got):
        assert g == mcdx.price(c, prec)
    book.close()


def test_schedules_and_autotune_do_not_change_bits():
    cs = small_book()
    book = mcdx.Book(cs, FP32)
    ref = book.price()
    for bpc in (1, 2, 3, 8):
        for order in (0, 1):
            book.set_schedule(bpc, order)
            assert book.price() == ref
    ms = book.autotune()
    assert ms > 0
    assert book.price() == ref
    book.close()


def test_graph_replays_with_other_work_between():
    cs = small_book()
    book = mcdx.Book(cs, FP32)
    ref = book.price()
    for i in range(4):
        mcdx.price(cs[i], FP32)          # eager work on another stream in between
        book.enqueue()
        assert book.results() == ref
    book.close()


def test_fp64_book_matches_oracle():
    cs = small_book()[:12]
    book = mcdx.Book(cs, FP64)
    got = book.price()
    for c, g in zip(cs, got):
        p = O.product(c.product, spot=c.spot, strike=c.strike, barrier=c.barrier,
                      rebate=c.rebate, rate=c.rate, carry_yield=c.carry_yield, vol=c.vol,
                      time=c.time, cash_amount=c.cash_amount,
                      monitoring_points=c.monitoring_points, type=c.type,
                      direction=c.direction, knock=c.knock, strike_style=c.strike_style,
                      average_style=c.average_style, digital_style=c.digital_style,
                      antithetic=c.antithetic, control_variate=c.control_variate,
                      brownian_bridge=c.brownian_bridge)
        op, ose, on = O.price(p, c.path_count, c.seed)
        assert g[2] == on
        assert abs(g[0] - op) <= 1e-12 * max(1.0, abs(op)), (c.product, g, op)
    book.close()


def test_book_rejects_bad_contracts():
    cs = small_book()[:3]
    cs[1].monitoring_points = 0
    arr = (mcdx.GbmContract * 3)(*cs)
    h = C.c_void_p()
    assert mcdx.lib().mcdx_book_create(arr, 3, FP32, C.byref(h)) == mcdx.EINVAL
    assert mcdx.lib().mcdx_book_create(arr, 0, FP32, C.byref(h)) == mcdx.EINVAL


def crn_book(n=120, seed=11):
    """Common seed across the book, so contracts of one family and step count share draws."""
    cs = mcdx.c4_book(n_contracts=n, seed=seed, log2_min=8, log2_max=14, mc_seed=5)
    cs[0].product = mcdx.LOOKBACK
    cs[0].strike_style = 1
    for c in cs[1:4]:  # a lookback group with ragged path counts
        c.product = mcdx.LOOKBACK
        c.strike_style = 0
        c.monitoring_points = cs[0].monitoring_points
    cs[4].control_variate = 1 if cs[4].product == mcdx.ASIAN else 0
    return cs


@pytest.mark.parametrize("flags", [0, mcdx.Book.NO_SHARING])
def test_shared_draw_book_equals_per_contract_calls_bitwise(flags):
    cs = crn_book()
    book = mcdx.Book(cs, FP32, flags)
    got = book.price()
    for c, g in zip(cs, got):
        assert g == mcdx.price(c, FP32), (c.product, c.path_count)
    for bpc, order in ((1, 0), (2, 1)):
        book.set_schedule(bpc, order)
        assert book.price() == got
    book.close()


def test_book_greeks_are_per_path_central_differences():
    cs = crn_book(24, seed=3)
    book = mcdx.Book(cs, FP32, mcdx.Book.GREEKS)
    prices = book.price()
    greeks = book.greeks()
    for c, p, (d, g, v) in zip(cs, prices, greeks):
        assert p == mcdx.price(c, FP32)
        hs, hv = max(0.01 * c.spot, 1e-4), max(0.01 * c.vol, 1e-4)

        def bumped(**kw):
            b = mcdx.GbmContract.from_buffer_copy(c)
            for k, x in kw.items():
                setattr(b, k, x)
            return mcdx.price(b, FP32)[0]
        up, dn = bumped(spot=c.spot + hs), bumped(spot=c.spot - hs)
        vu, vd = bumped(vol=c.vol + hv), bumped(vol=c.vol - hv)
        # the mean of per-path differences equals the difference of the bumped prices
        tol = 1e-9 * (abs(up) + abs(dn) + 1.0)
        assert abs(d[0] - (up - dn) / (2 * hs)) <= tol / hs
        assert abs(g[0] - (up - 2 * p[0] + dn) / hs ** 2) <= 4 * tol / hs ** 2
        assert abs(v[0] - (vu - vd) / (2 * hv)) <= tol / hv
        assert d[2] == g[2] == v[2] == c.path_count
        assert d[1] >= 0 and v[1] >= 0
    book.close()


def test_book_greeks_european_match_black_scholes():
    import mcd
    cs = [mcdx.contract(mcdx.EUROPEAN, strike=k, vol=0.25, time=1.0, path_count=1 << 20, seed=9)
          for k in (90.0, 100.0, 110.0)]
    book = mcdx.Book(cs, FP32, mcdx.Book.GREEKS)
    book.price()
    for c, (d, g, v) in zip(cs, book.greeks()):
        h = 1e-3
        f = lambda s=c.spot, vol=c.vol: mcd.black_scholes_merton(s, c.strike, c.rate, 0.0, vol, 1.0,
                                                                 mcd.OptionType.Call)
        bd = (f(s=c.spot + h) - f(s=c.spot - h)) / (2 * h)
        bg = (f(s=c.spot + h) - 2 * f() + f(s=c.spot - h)) / h ** 2
        bv = (f(vol=c.vol + h) - f(vol=c.vol - h)) / (2 * h)
        assert abs(d[0] - bd) <= 4 * d[1] + 1e-3
        assert abs(g[0] - bg) <= 4 * g[1] + 2e-3
        assert abs(v[0] - bv) <= 4 * v[1] + 0.05
    book.close()


def pathwise_book():
    E, A = mcdx.EUROPEAN, mcdx.ASIAN
    kw = dict(rate=0.05, carry_yield=0.01, vol=0.25, time=1.0, path_count=40_000, seed=21)
    cs = [mcdx.contract(E, strike=95.0, type=0, **kw),
          mcdx.contract(E, strike=105.0, type=1, **kw)]
    for ss, avs, typ in [(0, 0, 0), (0, 1, 1), (1, 0, 1), (1, 1, 0)]:
        cs.append(mcdx.contract(A, strike=98.0, type=typ, strike_style=ss, average_style=avs,
                                monitoring_points=19, **kw))
    cs.append(mcdx.contract(mcdx.DIGITAL, **kw))
    cs.append(mcdx.contract(mcdx.BARRIER, barrier=130.0, monitoring_points=12, **kw))
    cs.append(mcdx.contract(A, monitoring_points=12, control_variate=1, **kw))
    return cs


def test_book_pathwise_matches_per_contract_pathwise():
    # The book's pathwise delta and vega (FP32 draws on the book's paths) against the
    # per-contract pathwise entry points (the reference's FP64 arithmetic on the same Philox
    # draws, pinned to the oracle in test_gpu_greeks.py): the same estimator on the same
    # paths, so they agree far inside one standard error.
    cs = pathwise_book()
    book = mcdx.Book(cs, FP32, mcdx.Book.GREEKS)
    book.price()
    pw = book.pathwise()
    for c, (d, v) in zip(cs, pw):
        if c.product == mcdx.EUROPEAN:
            ref = mcdx.greeks("mcdx_pathwise_european", c.spot, c.strike, c.rate, c.carry_yield,
                              c.vol, c.time, c.type, c.path_count, c.seed, n_out=3)
        elif c.product == mcdx.ASIAN and not c.control_variate:
            ref = mcdx.greeks("mcdx_pathwise_asian", c.spot, c.strike, c.rate, c.carry_yield,
                              c.vol, c.time, c.type, c.strike_style, c.average_style,
                              c.monitoring_points, c.path_count, c.seed, n_out=3)
        else:
            assert np.isnan(d[0]) and np.isnan(v[0]) and d[2] == v[2] == 0
            continue
        for got, want in ((d, ref[0]), (v, ref[1])):
            assert got[2] == c.path_count
            assert abs(got[0] - want[0]) <= 0.02 * want[1] + 1e-9, (c.product, got, want)
            assert abs(got[1] - want[1]) <= 0.01 * want[1] + 1e-12, (c.product, got, want)
    book.close()


def test_book_pathwise_european_near_black_scholes_and_fd():
    import mcd
    cs = [mcdx.contract(mcdx.EUROPEAN, strike=k, vol=0.25, time=1.0, path_count=1 << 20, seed=5)
          for k in (90.0, 110.0)]
    book = mcdx.Book(cs, FP32, mcdx.Book.GREEKS)
    book.price()
    for c, (d, g, v), (pd, pv) in zip(cs, book.greeks(), book.pathwise()):
        h = 1e-3
        f = lambda s=c.spot, vol=c.vol: mcd.black_scholes_merton(s, c.strike, c.rate, 0.0, vol,
                                                                 1.0, mcd.OptionType.Call)
        bd = (f(s=c.spot + h) - f(s=c.spot - h)) / (2 * h)
        bv = (f(vol=c.vol + h) - f(vol=c.vol - h)) / (2 * h)
        assert abs(pd[0] - bd) <= 4 * pd[1], (pd, bd)
        assert abs(pv[0] - bv) <= 4 * pv[1], (pv, bv)
        # pathwise and finite differences on the same paths estimate the same Greeks
        assert abs(pd[0] - d[0]) <= 4 * np.hypot(pd[1], d[1])
        assert abs(pv[0] - v[0]) <= 4 * np.hypot(pv[1], v[1])
    book.close()


def test_greeks_book_rejects_unsupported_contracts():
    cs = crn_book(8)[:4]
    cs[0].antithetic = 1
    arr = (mcdx.GbmContract * 4)(*cs)
    h = C.c_void_p()
    assert mcdx.lib().mcdx_book_create_ex(arr, 4, FP32, 2, C.byref(h)) == mcdx.EINVAL
    assert mcdx.lib().mcdx_book_create_ex(arr, 4, FP64, 2, C.byref(h)) == mcdx.EINVAL


def test_c4_full_book_equals_per_contract_calls_bitwise():
    # The C4 workload itself (4096 contracts, about 1e9 paths, one Monte Carlo seed so that
    # families share draws): every contract's result equals its own launch bit for bit.
    cs = mcdx.c4_book(mc_seed=42)
    book = mcdx.Book(cs, FP32)
    got = book.price()
    for c, g in zip(cs, got):
        assert g == mcdx.price(c, FP32)
    book.close()


def greeks_mix_book():
    # Greeks units of every family with ragged path counts, so groups of 2 and 3 units hold
    # members whose tiles end at different points.
    cs = crn_book(30, seed=11)
    for c in cs:
        c.antithetic = 0
        c.brownian_bridge = 0
    return cs


def test_greeks_book_bits_independent_of_fusion_width_and_tiles_per_item():
    cs = greeks_mix_book()
    book = mcdx.Book(cs, FP32, mcdx.Book.GREEKS)
    book.set_schedule2(0, 1, 1, 1)
    ref = (book.price(), book.greeks(), book.pathwise())
    for c, p in zip(cs, ref[0]):
        assert p == mcdx.price(c, FP32)
    for width in (1, 2, 3):
        for tpi in (1, 2, 4):
            for bpc, order in ((0, 1), (1, 0)):
                book.set_schedule2(bpc, order, width, tpi)
                assert book.schedule2()[3:] == (width, tpi)
                got = (book.price(), book.greeks(), book.pathwise())
                assert got[0] == ref[0] and got[1] == ref[1]
                # NaN (no pathwise estimator) compares by its place
                assert [[tuple(np.nan_to_num(x)) for x in r] for r in got[2]] == \
                       [[tuple(np.nan_to_num(x)) for x in r] for r in ref[2]]
    with pytest.raises(mcdx.McdxError):
        book.set_schedule2(0, 1, 4, 1)
    with pytest.raises(mcdx.McdxError):
        book.set_schedule2(0, 1, 1, 3)
    book.close()


def test_price_book_tiles_per_item_do_not_change_bits():
    cs = crn_book()
    book = mcdx.Book(cs, FP32)
    ref = book.price()
    for tpi in (2, 4, 8):
        book.set_schedule2(0, 1, 0, tpi)
        assert book.price() == ref
    book.close()


def test_schedule_cache_file_round_trip_and_damage(tmp_path, monkeypatch):
    path = tmp_path / "sched.txt"
    monkeypatch.setenv("MCDX_BOOK_SCHEDULE_CACHE", str(path))
    cs = greeks_mix_book()
    book = mcdx.Book(cs, FP32, mcdx.Book.GREEKS)
    ref = (book.price(), book.greeks())
    book.autotune()
    tuned = book.schedule2()
    assert (book.price(), book.greeks()) == ref
    book.close()
    lines = path.read_text().splitlines()
    assert lines and lines[-1].startswith("mcdx-book-schedule 1 ")
    key = lines[-1].split()[2]
    # a new book of the same shape starts from the tuned schedule
    book = mcdx.Book(cs, FP32, mcdx.Book.GREEKS)
    assert book.schedule2()[3:] == tuned[3:]
    book.close()
    # a damaged cache (garbage, values out of range, wrong version) falls back to defaults
    path.write_text("garbage\n" + f"mcdx-book-schedule 1 {key} 2 1 9 1\n" +
                    f"mcdx-book-schedule 2 {key} 1 0 1 1\n" + f"mcdx-book-schedule 1 {key} 1 1 1\n")
    monkeypatch.setenv("MCDX_BOOK_SCHEDULE_CACHE", str(tmp_path / "other.txt"))
    mcdx.Book(cs[:3], FP32, mcdx.Book.GREEKS).close()  # the cache is re-read on a path change
    monkeypatch.setenv("MCDX_BOOK_SCHEDULE_CACHE", str(path))
    book = mcdx.Book(cs, FP32, mcdx.Book.GREEKS)
    assert book.schedule2()[3:] == (2, 1)  # the default width and tiles per item
    assert (book.price(), book.greeks()) == ref
    book.close()
    # a valid line for this shape is taken
    path.write_text(f"mcdx-book-schedule 1 {key} 1 0 3 4\n")
    monkeypatch.setenv("MCDX_BOOK_SCHEDULE_CACHE", str(tmp_path / "x.txt"))
    b2 = mcdx.Book(cs[:3], FP32, mcdx.Book.GREEKS)  # another shape: reads x.txt (absent)
    b2.close()
    monkeypatch.setenv("MCDX_BOOK_SCHEDULE_CACHE", str(path))
    book = mcdx.Book(cs, FP32, mcdx.Book.GREEKS)
    assert book.schedule2()[1] == 0 and book.schedule2()[3:] == (3, 4)
    assert (book.price(), book.greeks()) == ref
    book.close()
