// pybind11 module `mcd`: the reference's Python API (bindings/python/src/module.cpp:64-293)
// with the same names, classes, attributes, argument names and defaults, backed by the MI355X
// engine. The GIL is released around every pricer. Where the reference returns silently,
// this module does too for valid inputs; a device error (no GPU, allocation failure) raises
// RuntimeError instead of returning NaN.
#include <pybind11/pybind11.h>

#include "mcd/engine.hpp"
#include <pybind11/stl.h>

#include <new>
#include <string>

#include "mcd/core/sobol.hpp"
#include "mcd/core/timing.hpp"
#include "mcd/core/types.hpp"
#include "mcd/greeks/finite_difference.hpp"
#include "mcd/greeks/likelihood_ratio.hpp"
#include "mcd/greeks/pathwise.hpp"
#include "mcd/pricers/analytic.hpp"
#include "mcd/pricers/binomial.hpp"
#include "mcd/pricers/heston.hpp"
#include "mcd/pricers/lsm.hpp"
#include "mcd/pricers/monte_carlo.hpp"
#include "mcd/pricers/qmc.hpp"
#include "mcdx.h"

namespace py = pybind11;
using namespace mcd;
namespace pricers = mcd::pricers;
namespace greeks = mcd::greeks;

namespace {

const char* status_text(int s) {
    switch (s) {
    case MCDX_EINVAL: return "invalid argument";
    case MCDX_EDEVICE: return "HIP device error";
    case MCDX_ENOMEM: return "device memory exhausted";
    case MCDX_ENODEV: return "no HIP device visible (the mcd engine needs an MI355X)";
    default: return "engine error";
    }
}

// Run an engine-backed call with the GIL released; raise if the engine reported an error.
template <typename F>
auto engine(F&& f) {
    decltype(f()) out;
    {
        py::gil_scoped_release release;
        out = f();
    }
    const int st = mcd::last_engine_status();
    if (st != MCDX_OK) {
        if (st == MCDX_EINVAL) throw py::value_error(status_text(st));
        if (st == MCDX_ENOMEM) throw std::bad_alloc(); // MemoryError, as a host allocation's
        throw std::runtime_error(status_text(st));
    }
    return out;
}

py::dict benchmark_european(double spot, double strike, double rate, double carry_yield,
                            double vol, double time, OptionType type, std::uint64_t path_count,
                            std::uint64_t seed, pricers::McOptions options) {
    // module.cpp:29-60: time_call around the European pricer, GIL released only there.
    mcd::TimingResult<pricers::McResult> timed{};
    {
        py::gil_scoped_release release;
        timed = mcd::time_call([&] {
            return pricers::monte_carlo_european(spot, strike, rate, carry_yield, vol, time, type,
                                                 path_count, seed, options);
        });
    }
    if (mcd::last_engine_status() != MCDX_OK)
        throw std::runtime_error(status_text(mcd::last_engine_status()));
    py::dict out;
    out["price"] = timed.value.price;
    out["standard_error"] = timed.value.standard_error;
    out["path_count"] = timed.value.path_count;
    out["elapsed_seconds"] = timed.elapsed_seconds;
    out["paths_per_second"] = timed.elapsed_seconds > 0.0
                                  ? static_cast<double>(timed.value.path_count) / timed.elapsed_seconds
                                  : 0.0;
    out["simd_enabled"] = false; // kHasNeon && !antithetic; this build has no NEON path
    return out;
}

} // namespace

PYBIND11_MODULE(mcd, m) {
    m.doc() = "mcd Monte Carlo derivatives pricing engine (MI355X device backend)";

    py::enum_<OptionType>(m, "OptionType").value("Call", OptionType::Call).value("Put", OptionType::Put);
    py::enum_<BarrierDirection>(m, "BarrierDirection")
        .value("Up", BarrierDirection::Up)
        .value("Down", BarrierDirection::Down);
    py::enum_<BarrierKnock>(m, "BarrierKnock").value("In", BarrierKnock::In).value("Out", BarrierKnock::Out);
    py::enum_<StrikeStyle>(m, "StrikeStyle")
        .value("Fixed", StrikeStyle::Fixed)
        .value("Floating", StrikeStyle::Floating);
    py::enum_<DigitalStyle>(m, "DigitalStyle")
        .value("CashOrNothing", DigitalStyle::CashOrNothing)
        .value("AssetOrNothing", DigitalStyle::AssetOrNothing);
    py::enum_<AverageStyle>(m, "AverageStyle")
        .value("Arithmetic", AverageStyle::Arithmetic)
        .value("Geometric", AverageStyle::Geometric);

    py::class_<pricers::McResult>(m, "McResult")
        .def_readonly("price", &pricers::McResult::price)
        .def_readonly("standard_error", &pricers::McResult::standard_error)
        .def_readonly("path_count", &pricers::McResult::path_count);
    py::class_<pricers::McOptions>(m, "McOptions")
        .def(py::init<>())
        .def_readwrite("antithetic", &pricers::McOptions::antithetic)
        .def_readwrite("control_variate", &pricers::McOptions::control_variate)
        .def_readwrite("brownian_bridge", &pricers::McOptions::brownian_bridge)
        .def_readwrite("num_threads", &pricers::McOptions::num_threads);
    py::class_<pricers::BinomialResult>(m, "BinomialResult")
        .def_readonly("price", &pricers::BinomialResult::price)
        .def_readonly("risk_neutral_probability", &pricers::BinomialResult::risk_neutral_probability)
        .def_readonly("up_factor", &pricers::BinomialResult::up_factor)
        .def_readonly("down_factor", &pricers::BinomialResult::down_factor);
    py::class_<pricers::LsmResult>(m, "LsmResult")
        .def_readonly("price", &pricers::LsmResult::price)
        .def_readonly("standard_error", &pricers::LsmResult::standard_error);
    py::class_<greeks::EuropeanGreeks>(m, "EuropeanGreeks")
        .def_readonly("delta", &greeks::EuropeanGreeks::delta)
        .def_readonly("gamma", &greeks::EuropeanGreeks::gamma)
        .def_readonly("vega", &greeks::EuropeanGreeks::vega)
        .def_readonly("theta", &greeks::EuropeanGreeks::theta)
        .def_readonly("rho", &greeks::EuropeanGreeks::rho);
    py::class_<greeks::BumpSizes>(m, "BumpSizes")
        .def(py::init<>())
        .def_readwrite("spot", &greeks::BumpSizes::spot)
        .def_readwrite("vol", &greeks::BumpSizes::vol)
        .def_readwrite("rate", &greeks::BumpSizes::rate)
        .def_readwrite("time", &greeks::BumpSizes::time);
    py::class_<greeks::LrGreeksResult>(m, "LrGreeksResult")
        .def_readonly("value", &greeks::LrGreeksResult::value)
        .def_readonly("standard_error", &greeks::LrGreeksResult::standard_error);
    py::class_<greeks::LrGreeks>(m, "LrGreeks")
        .def_readonly("delta", &greeks::LrGreeks::delta)
        .def_readonly("gamma", &greeks::LrGreeks::gamma)
        .def_readonly("vega", &greeks::LrGreeks::vega)
        .def_readonly("rho", &greeks::LrGreeks::rho)
        .def_readonly("theta", &greeks::LrGreeks::theta);
    py::class_<greeks::PathwiseGreeksResult>(m, "PathwiseGreeksResult")
        .def_readonly("value", &greeks::PathwiseGreeksResult::value)
        .def_readonly("standard_error", &greeks::PathwiseGreeksResult::standard_error);
    py::class_<greeks::PathwiseGreeks>(m, "PathwiseGreeks")
        .def_readonly("delta", &greeks::PathwiseGreeks::delta)
        .def_readonly("vega", &greeks::PathwiseGreeks::vega)
        .def_readonly("rho", &greeks::PathwiseGreeks::rho);
    py::class_<pricers::QmcResult>(m, "QmcResult").def_readonly("price", &pricers::QmcResult::price);
    py::class_<models::HestonParams>(m, "HestonParams")
        .def(py::init<>())
        .def_readwrite("spot", &models::HestonParams::spot)
        .def_readwrite("rate", &models::HestonParams::rate)
        .def_readwrite("carry_yield", &models::HestonParams::carry_yield)
        .def_readwrite("v0", &models::HestonParams::v0)
        .def_readwrite("kappa", &models::HestonParams::kappa)
        .def_readwrite("theta", &models::HestonParams::theta)
        .def_readwrite("xi", &models::HestonParams::xi)
        .def_readwrite("rho", &models::HestonParams::rho)
        .def_readwrite("time", &models::HestonParams::time);
    py::class_<pricers::HestonMcResult>(m, "HestonMcResult")
        .def_readonly("price", &pricers::HestonMcResult::price)
        .def_readonly("standard_error", &pricers::HestonMcResult::standard_error)
        .def_readonly("path_count", &pricers::HestonMcResult::path_count);

    m.attr("SOBOL_MAX_DIMENSIONS") = kSobolMaxDimensions;
    m.attr("SOBOL_EXTENDED_DIMENSIONS") = kSobolExtendedDimensions;
    m.attr("HAS_NEON") = false;

    const auto rel = py::call_guard<py::gil_scoped_release>();
    m.def("black_scholes_merton", &pricers::black_scholes_merton, rel);
    m.def("black76", &pricers::black76, rel);
    m.def("kemna_vorst", &pricers::kemna_vorst, rel);
    m.def("reiner_rubinstein", &pricers::reiner_rubinstein, rel, py::arg("spot"),
          py::arg("strike"), py::arg("barrier"), py::arg("rate"), py::arg("carry_yield"),
          py::arg("vol"), py::arg("time"), py::arg("type"), py::arg("direction"),
          py::arg("knock"), py::arg("rebate") = 0.0);
    m.def("lookback_fixed_strike", &pricers::lookback_fixed_strike, rel);
    m.def("lookback_floating_strike", &pricers::lookback_floating_strike, rel);
    m.def("digital", &pricers::digital, rel, py::arg("spot"), py::arg("strike"), py::arg("rate"),
          py::arg("carry_yield"), py::arg("vol"), py::arg("time"), py::arg("type"),
          py::arg("style"), py::arg("cash_amount") = 1.0);
    m.def("forward_price", &pricers::forward_price, rel);
    m.def("forward_value", &pricers::forward_value, rel);
    m.def("futures_mark_to_market", &pricers::futures_mark_to_market, rel);
    m.def("crr_binomial", &pricers::crr_binomial, rel);
    m.def("crr_binomial_american", &pricers::crr_binomial_american, rel);

    m.def(
        "monte_carlo_european",
        [](double s, double k, double r, double q, double v, double t, OptionType type,
           std::uint64_t n, std::uint64_t seed, pricers::McOptions o) {
            return engine([&] { return pricers::monte_carlo_european(s, k, r, q, v, t, type, n, seed, o); });
        },
        py::arg("spot"), py::arg("strike"), py::arg("rate"), py::arg("carry_yield"),
        py::arg("vol"), py::arg("time"), py::arg("type"), py::arg("path_count"), py::arg("seed"),
        py::arg("options") = pricers::McOptions{});
    m.def(
        "monte_carlo_digital",
        [](double s, double k, double r, double q, double v, double t, OptionType type,
           DigitalStyle style, double cash, std::uint64_t n, std::uint64_t seed,
           pricers::McOptions o) {
            return engine([&] {
                return pricers::monte_carlo_digital(s, k, r, q, v, t, type, style, cash, n, seed, o);
            });
        },
        py::arg("spot"), py::arg("strike"), py::arg("rate"), py::arg("carry_yield"),
        py::arg("vol"), py::arg("time"), py::arg("type"), py::arg("style"),
        py::arg("cash_amount"), py::arg("path_count"), py::arg("seed"),
        py::arg("options") = pricers::McOptions{});
    m.def(
        "monte_carlo_asian",
        [](double s, double k, double r, double q, double v, double t, OptionType type,
           StrikeStyle ss, AverageStyle as, int mp, std::uint64_t n, std::uint64_t seed,
           pricers::McOptions o) {
            return engine([&] {
                return pricers::monte_carlo_asian(s, k, r, q, v, t, type, ss, as, mp, n, seed, o);
            });
        },
        py::arg("spot"), py::arg("strike"), py::arg("rate"), py::arg("carry_yield"),
        py::arg("vol"), py::arg("time"), py::arg("type"), py::arg("strike_style"),
        py::arg("average_style"), py::arg("monitoring_points"), py::arg("path_count"),
        py::arg("seed"), py::arg("options") = pricers::McOptions{});
    m.def(
        "monte_carlo_barrier",
        [](double s, double k, double h, double r, double q, double v, double t, OptionType type,
           BarrierDirection d, BarrierKnock kn, double rebate, int mp, std::uint64_t n,
           std::uint64_t seed, pricers::McOptions o) {
            return engine([&] {
                return pricers::monte_carlo_barrier(s, k, h, r, q, v, t, type, d, kn, rebate, mp, n,
                                                    seed, o);
            });
        },
        py::arg("spot"), py::arg("strike"), py::arg("barrier"), py::arg("rate"),
        py::arg("carry_yield"), py::arg("vol"), py::arg("time"), py::arg("type"),
        py::arg("direction"), py::arg("knock"), py::arg("rebate"), py::arg("monitoring_points"),
        py::arg("path_count"), py::arg("seed"), py::arg("options") = pricers::McOptions{});
    m.def(
        "monte_carlo_lookback",
        [](double s, double k, double r, double q, double v, double t, OptionType type,
           StrikeStyle style, int mp, std::uint64_t n, std::uint64_t seed, pricers::McOptions o) {
            return engine([&] {
                return pricers::monte_carlo_lookback(s, k, r, q, v, t, type, style, mp, n, seed, o);
            });
        },
        py::arg("spot"), py::arg("strike"), py::arg("rate"), py::arg("carry_yield"),
        py::arg("vol"), py::arg("time"), py::arg("type"), py::arg("style"),
        py::arg("monitoring_points"), py::arg("path_count"), py::arg("seed"),
        py::arg("options") = pricers::McOptions{});
    m.def("monte_carlo_lsm_american",
          [](double s, double k, double r, double q, double v, double t, OptionType type, int mp,
             std::uint64_t n, std::uint64_t seed) {
              return engine([&] {
                  return pricers::monte_carlo_lsm_american(s, k, r, q, v, t, type, mp, n, seed);
              });
          });

    m.def("default_bump_sizes", &greeks::default_bump_sizes, rel);
    m.def("finite_difference_european",
          [](double s, double k, double r, double q, double v, double t, OptionType type,
             std::uint64_t n, std::uint64_t seed, greeks::BumpSizes b) {
              return engine([&] {
                  return greeks::finite_difference_european(s, k, r, q, v, t, type, n, seed, b);
              });
          });
    m.def("likelihood_ratio_european",
          [](double s, double k, double r, double q, double v, double t, OptionType type,
             std::uint64_t n, std::uint64_t seed) {
              return engine([&] { return greeks::likelihood_ratio_european(s, k, r, q, v, t, type, n, seed); });
          });
    m.def("likelihood_ratio_digital",
          [](double s, double k, double r, double q, double v, double t, OptionType type,
             DigitalStyle style, double cash, std::uint64_t n, std::uint64_t seed) {
              return engine([&] {
                  return greeks::likelihood_ratio_digital(s, k, r, q, v, t, type, style, cash, n, seed);
              });
          });
    m.def("likelihood_ratio_barrier",
          [](double s, double k, double h, double r, double q, double v, double t, OptionType type,
             BarrierDirection d, BarrierKnock kn, double rebate, int mp, std::uint64_t n,
             std::uint64_t seed) {
              return engine([&] {
                  return greeks::likelihood_ratio_barrier(s, k, h, r, q, v, t, type, d, kn, rebate,
                                                          mp, n, seed);
              });
          });
    m.def("pathwise_european",
          [](double s, double k, double r, double q, double v, double t, OptionType type,
             std::uint64_t n, std::uint64_t seed) {
              return engine([&] { return greeks::pathwise_european(s, k, r, q, v, t, type, n, seed); });
          });
    m.def("pathwise_asian",
          [](double s, double k, double r, double q, double v, double t, OptionType type,
             StrikeStyle ss, AverageStyle as, int mp, std::uint64_t n, std::uint64_t seed) {
              return engine([&] {
                  return greeks::pathwise_asian(s, k, r, q, v, t, type, ss, as, mp, n, seed);
              });
          });

    m.def("qmc_sobol_european",
          [](double s, double k, double r, double q, double v, double t, OptionType type,
             std::uint64_t n) {
              return engine([&] { return pricers::qmc_sobol_european(s, k, r, q, v, t, type, n); });
          });
    m.def(
        "qmc_sobol_asian",
        [](double s, double k, double r, double q, double v, double t, OptionType type,
           StrikeStyle ss, int mp, std::uint64_t n) {
            if (mp < 1 || mp > static_cast<int>(kSobolMaxDimensions)) // module.cpp:268-272
                throw py::value_error("monitoring_points must be between 1 and " +
                                      std::to_string(kSobolMaxDimensions));
            return engine([&] { return pricers::qmc_sobol_asian(s, k, r, q, v, t, type, ss, mp, n); });
        },
        py::arg("spot"), py::arg("strike"), py::arg("rate"), py::arg("carry_yield"),
        py::arg("vol"), py::arg("time"), py::arg("type"), py::arg("strike_style"),
        py::arg("monitoring_points"), py::arg("path_count"));

    m.def("heston_qe_european",
          [](const models::HestonParams& p, double k, OptionType type, int mp, std::uint64_t n,
             std::uint64_t seed) {
              return engine([&] { return pricers::heston_qe_european(p, k, type, mp, n, seed); });
          });
    m.def("heston_semi_analytic", &pricers::heston_semi_analytic, rel);

    m.def("benchmark_european", &benchmark_european, py::arg("spot"), py::arg("strike"),
          py::arg("rate"), py::arg("carry_yield"), py::arg("vol"), py::arg("time"),
          py::arg("type"), py::arg("path_count"), py::arg("seed"),
          py::arg("options") = pricers::McOptions{});
}
