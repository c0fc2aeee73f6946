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
