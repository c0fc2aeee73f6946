// mcd_cli on the MI355X engine: one flat JSON request on stdin, one flat JSON response on
// stdout; exit 0, or {"error": "..."} and exit 1.
//
// Request fields, response schema and error contract follow the reference's
// apps/mcd_cli/main.cpp:229-556 (flat JSON as src/core/json.cpp:287-319): products forward,
// binomial_european/american, european, digital, asian, barrier, lookback, american, heston;
// requests price (default), greeks, lr_greeks, pathwise_greeks, qmc_sobol, semi_analytic.
// Every Monte Carlo price carries standard_error, ci_95_low/high, path_count, seed,
// elapsed_seconds and paths_per_second. Pricing goes through the drop-in C++ API
// (include/mcd/**, libmcd.so) and so through the GPU engine; an engine failure (no device,
// out of memory) becomes an error response rather than a NaN price.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <iterator>
#include <stdexcept>
#include <string>
#include <string_view>
#include <utility>
#include <vector>

#include "mcd/core/sobol.hpp"
#include "mcd/core/timing.hpp"
#include "mcd/core/types.hpp"
#include "mcd/engine.hpp"
#include "mcd/greeks/finite_difference.hpp"
#include "mcd/greeks/likelihood_ratio.hpp"
#include "mcd/greeks/pathwise.hpp"
#include "mcd/pricers/analytic.hpp"
#include "mcd/pricers/binomial.hpp"
#include "mcd/pricers/heston.hpp"
#include "mcd/pricers/lsm.hpp"
#include "mcd/pricers/monte_carlo.hpp"
#include "mcd/pricers/qmc.hpp"

namespace {

// ---- flat JSON: one object of string / number / bool / null members ----------------------
struct Value {
    enum Kind { Null, Bool, Number, String } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    static Value number(double d) { Value v; v.kind = Number; v.num = d; return v; }
    static Value boolean(bool x) { Value v; v.kind = Bool; v.b = x; return v; }
    static Value string(std::string s) { Value v; v.kind = String; v.str = std::move(s); return v; }
};
using Object = std::vector<std::pair<std::string, Value>>;

struct Reader {
    std::string_view s;
    size_t i = 0;

    [[noreturn]] void fail(const std::string& what) const {
        throw std::runtime_error("invalid JSON at offset " + std::to_string(i) + ": " + what);
    }
    void ws() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) ++i;
    }
    bool take(char c) {
        ws();
        if (i < s.size() && s[i] == c) { ++i; return true; }
        return false;
    }
    std::string string_lit() {
        if (!take('"')) fail("expected a string");
        std::string out;
        while (true) {
            if (i >= s.size()) fail("unterminated string");
            const char c = s[i++];
            if (c == '"') break;
            if ((unsigned char)c < 0x20) fail("control character in string");
            if (c != '\\') { out.push_back(c); continue; }
            if (i >= s.size()) fail("unterminated escape");
            switch (s[i++]) {
            case '"': out.push_back('"'); break;
            case '\\': out.push_back('\\'); break;
            case '/': out.push_back('/'); break;
            case 'b': out.push_back('\b'); break;
            case 'f': out.push_back('\f'); break;
            case 'n': out.push_back('\n'); break;
            case 'r': out.push_back('\r'); break;
            case 't': out.push_back('\t'); break;
            default: fail("unsupported escape (\\u escapes are not accepted)");
            }
        }
        return out;
    }
    double number_lit() {
        const size_t start = i;
        if (i < s.size() && s[i] == '-') ++i;
        if (i >= s.size() || !(s[i] >= '0' && s[i] <= '9')) fail("expected a value");
        if (s[i] == '0') ++i;
        else while (i < s.size() && s[i] >= '0' && s[i] <= '9') ++i;
        if (i < s.size() && s[i] == '.') {
            ++i;
            if (i >= s.size() || !(s[i] >= '0' && s[i] <= '9')) fail("digit expected after '.'");
            while (i < s.size() && s[i] >= '0' && s[i] <= '9') ++i;
        }
        if (i < s.size() && (s[i] == 'e' || s[i] == 'E')) {
            ++i;
            if (i < s.size() && (s[i] == '+' || s[i] == '-')) ++i;
            if (i >= s.size() || !(s[i] >= '0' && s[i] <= '9')) fail("digit expected in exponent");
            while (i < s.size() && s[i] >= '0' && s[i] <= '9') ++i;
        }
        const std::string text(s.substr(start, i - start));
        return std::strtod(text.c_str(), nullptr);
    }
    bool word(std::string_view w) {
        if (s.substr(i, w.size()) == w) { i += w.size(); return true; }
        return false;
    }
    Value value() {
        ws();
        if (i >= s.size()) fail("expected a value");
        const char c = s[i];
        if (c == '"') return Value::string(string_lit());
        if (c == '{' || c == '[') fail("nested objects and arrays are not supported");
        if (word("true")) return Value::boolean(true);
        if (word("false")) return Value::boolean(false);
        if (word("null")) return Value{};
        return Value::number(number_lit());
    }
    Object object() {
        Object out;
        if (!take('{')) fail("expected '{'");
        if (take('}')) return out;
        do {
            ws();
            std::string key = string_lit();
            if (!take(':')) fail("expected ':'");
            out.emplace_back(std::move(key), value());
        } while (take(','));
        if (!take('}')) fail("expected ',' or '}'");
        ws();
        if (i != s.size()) fail("trailing content after the object");
        return out;
    }
};

void write_string(std::string& out, const std::string& s) {
    out.push_back('"');
    for (const char c : s) {
        switch (c) {
        case '"': out += "\\\""; break;
        case '\\': out += "\\\\"; break;
        case '\n': out += "\\n"; break;
        case '\r': out += "\\r"; break;
        case '\t': out += "\\t"; break;
        default:
            if ((unsigned char)c < 0x20) {
                char buf[8];
                std::snprintf(buf, sizeof buf, "\\u%04x", (unsigned)(unsigned char)c);
                out += buf;
            } else {
                out.push_back(c);
            }
        }
    }
    out.push_back('"');
}

std::string write(const Object& obj) {
    std::string out = "{";
    for (size_t k = 0; k < obj.size(); ++k) {
        if (k) out.push_back(',');
        write_string(out, obj[k].first);
        out.push_back(':');
        const Value& v = obj[k].second;
        switch (v.kind) {
        case Value::Null: out += "null"; break;
        case Value::Bool: out += v.b ? "true" : "false"; break;
        case Value::String: write_string(out, v.str); break;
        case Value::Number:
            if (std::isfinite(v.num)) {
                char buf[40];
                std::snprintf(buf, sizeof buf, "%.10g", v.num); // the reference's display precision
                out += buf;
            } else {
                out += "null";
            }
            break;
        }
    }
    out.push_back('}');
    return out;
}

// ---- request fields ------------------------------------------------------------------------
const Value* find(const Object& o, std::string_view key) {
    for (const auto& kv : o)
        if (kv.first == key) return &kv.second;
    return nullptr;
}

const Value& need(const Object& o, std::string_view key) {
    const Value* v = find(o, key);
    if (!v) throw std::runtime_error("missing required field '" + std::string(key) + "'");
    return *v;
}

double num(const Object& o, std::string_view key) {
    const Value& v = need(o, key);
    if (v.kind != Value::Number) throw std::runtime_error("field '" + std::string(key) + "' must be a number");
    return v.num;
}

double num_or(const Object& o, std::string_view key, double fallback) {
    return find(o, key) ? num(o, key) : fallback;
}

std::string str(const Object& o, std::string_view key) {
    const Value& v = need(o, key);
    if (v.kind != Value::String) throw std::runtime_error("field '" + std::string(key) + "' must be a string");
    return v.str;
}

std::string str_or(const Object& o, std::string_view key, std::string fallback) {
    return find(o, key) ? str(o, key) : fallback;
}

// Seeds and path counts travel as JSON numbers (doubles): only integers up to 2^53 are exact.
std::uint64_t count(const Object& o, std::string_view key) {
    const double d = num(o, key);
    if (!(d >= 0.0) || d != std::floor(d) || d > 9007199254740992.0)
        throw std::runtime_error("field '" + std::string(key) +
                                 "' must be a non-negative integer no larger than 2^53");
    return (std::uint64_t)d;
}

int integer(const Object& o, std::string_view key) {
    const double d = num(o, key);
    if (d != std::floor(d) || d < -2147483648.0 || d > 2147483647.0)
        throw std::runtime_error("field '" + std::string(key) + "' must be an integer");
    return (int)d;
}

template <typename E>
E choose(const Object& o, std::string_view key, std::string_view a, std::string_view b, E ea, E eb) {
    const std::string s = str(o, key);
    if (s == a) return ea;
    if (s == b) return eb;
    throw std::runtime_error("'" + std::string(key) + "' must be '" + std::string(a) + "' or '" +
                             std::string(b) + "'");
}

mcd::OptionType option_type(const Object& o) {
    return choose(o, "type", "call", "put", mcd::OptionType::Call, mcd::OptionType::Put);
}

// A noexcept pricer that failed on the device reports a NaN price; surface the status.
void check_engine() {
    static const char* const what[] = {"ok", "invalid argument", "device error",
                                       "out of device or pinned memory", "no GPU device visible"};
    const int st = mcd::last_engine_status();
    if (st != 0)
        throw std::runtime_error(std::string("engine: ") + (st > 0 && st < 5 ? what[st] : "error") +
                                 " (status " + std::to_string(st) + ")");
}

// ---- responses -------------------------------------------------------------------------------
constexpr double kZ975 = 1.959963985;

Object mc_response(double price, double se, std::uint64_t paths, std::uint64_t seed, double elapsed) {
    Object out;
    out.emplace_back("price", Value::number(price));
    out.emplace_back("standard_error", Value::number(se));
    out.emplace_back("ci_95_low", Value::number(price - kZ975 * se));
    out.emplace_back("ci_95_high", Value::number(price + kZ975 * se));
    out.emplace_back("path_count", Value::number((double)paths));
    out.emplace_back("seed", Value::number((double)seed));
    out.emplace_back("elapsed_seconds", Value::number(elapsed));
    out.emplace_back("paths_per_second", Value::number(elapsed > 0.0 ? (double)paths / elapsed : 0.0));
    return out;
}

Object price_response(double price, double elapsed) {
    return Object{{"price", Value::number(price)}, {"elapsed_seconds", Value::number(elapsed)}};
}

Object qmc_response(double price, std::uint64_t paths, double elapsed) {
    Object out;
    out.emplace_back("price", Value::number(price));
    out.emplace_back("path_count", Value::number((double)paths));
    out.emplace_back("elapsed_seconds", Value::number(elapsed));
    out.emplace_back("paths_per_second", Value::number(elapsed > 0.0 ? (double)paths / elapsed : 0.0));
    out.emplace_back("note", Value::string("Sobol QMC is deterministic: no standard_error or "
                                           "confidence interval is reported."));
    return out;
}

void add_greek(Object& out, const char* name, double value, double se) {
    out.emplace_back(name, Value::number(value));
    out.emplace_back(std::string(name) + "_se", Value::number(se));
}

Object lr_response(const mcd::greeks::LrGreeks& g, double elapsed) {
    Object out;
    add_greek(out, "delta", g.delta.value, g.delta.standard_error);
    add_greek(out, "gamma", g.gamma.value, g.gamma.standard_error);
    add_greek(out, "vega", g.vega.value, g.vega.standard_error);
    add_greek(out, "rho", g.rho.value, g.rho.standard_error);
    if (g.theta) add_greek(out, "theta", g.theta->value, g.theta->standard_error);
    out.emplace_back("elapsed_seconds", Value::number(elapsed));
    return out;
}

Object pathwise_response(const mcd::greeks::PathwiseGreeks& g, double elapsed) {
    Object out;
    add_greek(out, "delta", g.delta.value, g.delta.standard_error);
    add_greek(out, "vega", g.vega.value, g.vega.standard_error);
    add_greek(out, "rho", g.rho.value, g.rho.standard_error);
    out.emplace_back("elapsed_seconds", Value::number(elapsed));
    return out;
}

// ---- dispatch --------------------------------------------------------------------------------
struct Market {
    double spot, strike, rate, carry, vol, time;
    mcd::OptionType type;
};

Market market(const Object& r, bool strike_required = true) {
    return Market{num(r, "spot"), strike_required ? num(r, "strike") : num_or(r, "strike", 0.0),
                  num(r, "rate"), num(r, "carry_yield"), num(r, "vol"), num(r, "time"),
                  option_type(r)};
}

template <typename F>
auto timed(F&& f) {
    auto t = mcd::time_call(std::forward<F>(f));
    check_engine();
    return t;
}

Object handle(const Object& r) {
    namespace P = mcd::pricers;
    namespace G = mcd::greeks;
    const std::string product = str(r, "product");
    const std::string kind = str_or(r, "request", "price");

    if (product == "forward") {
        const double s = num(r, "spot"), rate = num(r, "rate"), q = num(r, "carry_yield"),
                     t = num(r, "time");
        auto res = mcd::time_call([&] { return P::forward_price(s, rate, q, t); });
        return price_response(res.value, res.elapsed_seconds);
    }
    if (product == "binomial_european" || product == "binomial_american") {
        const Market m = market(r);
        const int steps = integer(r, "steps");
        if (product == "binomial_american") {
            auto res = mcd::time_call([&] {
                return P::crr_binomial_american(m.spot, m.strike, m.rate, m.carry, m.vol, m.time, steps, m.type);
            });
            return price_response(res.value, res.elapsed_seconds);
        }
        auto res = mcd::time_call([&] {
            return P::crr_binomial(m.spot, m.strike, m.rate, m.carry, m.vol, m.time, steps, m.type);
        });
        Object out = price_response(res.value.price, res.elapsed_seconds);
        out.emplace_back("risk_neutral_probability", Value::number(res.value.risk_neutral_probability));
        out.emplace_back("up_factor", Value::number(res.value.up_factor));
        out.emplace_back("down_factor", Value::number(res.value.down_factor));
        return out;
    }
    if (product == "european") {
        const Market m = market(r);
        if (kind == "qmc_sobol") {
            const std::uint64_t n = count(r, "path_count");
            auto res = timed([&] {
                return P::qmc_sobol_european(m.spot, m.strike, m.rate, m.carry, m.vol, m.time, m.type, n);
            });
            return qmc_response(res.value.price, n, res.elapsed_seconds);
        }
        const std::uint64_t n = count(r, "path_count"), seed = count(r, "seed");
        if (kind == "greeks") {
            const auto bumps = G::default_bump_sizes(m.spot, m.vol, m.time);
            auto res = timed([&] {
                return G::finite_difference_european(m.spot, m.strike, m.rate, m.carry, m.vol, m.time,
                                                     m.type, n, seed, bumps);
            });
            return Object{{"delta", Value::number(res.value.delta)},
                          {"gamma", Value::number(res.value.gamma)},
                          {"vega", Value::number(res.value.vega)},
                          {"theta", Value::number(res.value.theta)},
                          {"rho", Value::number(res.value.rho)},
                          {"elapsed_seconds", Value::number(res.elapsed_seconds)}};
        }
        if (kind == "lr_greeks") {
            auto res = timed([&] {
                return G::likelihood_ratio_european(m.spot, m.strike, m.rate, m.carry, m.vol, m.time, m.type, n, seed);
            });
            return lr_response(res.value, res.elapsed_seconds);
        }
        if (kind == "pathwise_greeks") {
            auto res = timed([&] {
                return G::pathwise_european(m.spot, m.strike, m.rate, m.carry, m.vol, m.time, m.type, n, seed);
            });
            return pathwise_response(res.value, res.elapsed_seconds);
        }
        auto res = timed([&] {
            return P::monte_carlo_european(m.spot, m.strike, m.rate, m.carry, m.vol, m.time, m.type, n, seed);
        });
        Object out = mc_response(res.value.price, res.value.standard_error, res.value.path_count, seed,
                                 res.elapsed_seconds);
        out.emplace_back("simd_enabled", Value::boolean(false)); // no NEON path in this build
        return out;
    }
    if (product == "digital") {
        const Market m = market(r);
        const auto style = choose(r, "digital_style", "cash_or_nothing", "asset_or_nothing",
                                  mcd::DigitalStyle::CashOrNothing, mcd::DigitalStyle::AssetOrNothing);
        const double cash = num_or(r, "cash_amount", 1.0);
        const std::uint64_t n = count(r, "path_count"), seed = count(r, "seed");
        if (kind == "lr_greeks") {
            auto res = timed([&] {
                return G::likelihood_ratio_digital(m.spot, m.strike, m.rate, m.carry, m.vol, m.time, m.type,
                                                   style, cash, n, seed);
            });
            return lr_response(res.value, res.elapsed_seconds);
        }
        auto res = timed([&] {
            return P::monte_carlo_digital(m.spot, m.strike, m.rate, m.carry, m.vol, m.time, m.type, style,
                                          cash, n, seed);
        });
        Object out = mc_response(res.value.price, res.value.standard_error, res.value.path_count, seed,
                                 res.elapsed_seconds);
        out.emplace_back("simd_enabled", Value::boolean(false));
        return out;
    }
    if (product == "asian") {
        const Market m = market(r);
        const auto ss = choose(r, "strike_style", "fixed", "floating", mcd::StrikeStyle::Fixed,
                               mcd::StrikeStyle::Floating);
        const auto avg = choose(r, "average_style", "arithmetic", "geometric",
                                mcd::AverageStyle::Arithmetic, mcd::AverageStyle::Geometric);
        const int mpts = integer(r, "monitoring_points");
        const std::uint64_t n = count(r, "path_count");
        if (kind == "qmc_sobol") {
            if (avg != mcd::AverageStyle::Arithmetic)
                throw std::runtime_error("'qmc_sobol' asian only supports 'arithmetic' average_style");
            if (mpts < 1 || mpts > (int)mcd::kSobolMaxDimensions)
                throw std::runtime_error("'monitoring_points' for 'qmc_sobol' asian must be between 1 and " +
                                         std::to_string(mcd::kSobolMaxDimensions));
            auto res = timed([&] {
                return P::qmc_sobol_asian(m.spot, m.strike, m.rate, m.carry, m.vol, m.time, m.type, ss, mpts, n);
            });
            return qmc_response(res.value.price, n, res.elapsed_seconds);
        }
        const std::uint64_t seed = count(r, "seed");
        if (kind == "pathwise_greeks") {
            auto res = timed([&] {
                return G::pathwise_asian(m.spot, m.strike, m.rate, m.carry, m.vol, m.time, m.type, ss, avg,
                                         mpts, n, seed);
            });
            return pathwise_response(res.value, res.elapsed_seconds);
        }
        auto res = timed([&] {
            return P::monte_carlo_asian(m.spot, m.strike, m.rate, m.carry, m.vol, m.time, m.type, ss, avg,
                                        mpts, n, seed);
        });
        return mc_response(res.value.price, res.value.standard_error, res.value.path_count, seed,
                           res.elapsed_seconds);
    }
    if (product == "barrier") {
        const Market m = market(r);
        const double h = num(r, "barrier");
        const auto dir = choose(r, "direction", "up", "down", mcd::BarrierDirection::Up,
                                mcd::BarrierDirection::Down);
        const auto knock = choose(r, "knock", "in", "out", mcd::BarrierKnock::In, mcd::BarrierKnock::Out);
        const double rebate = num_or(r, "rebate", 0.0);
        const int mpts = integer(r, "monitoring_points");
        const std::uint64_t n = count(r, "path_count"), seed = count(r, "seed");
        if (kind == "lr_greeks") {
            auto res = timed([&] {
                return G::likelihood_ratio_barrier(m.spot, m.strike, h, m.rate, m.carry, m.vol, m.time, m.type,
                                                   dir, knock, rebate, mpts, n, seed);
            });
            return lr_response(res.value, res.elapsed_seconds);
        }
        auto res = timed([&] {
            return P::monte_carlo_barrier(m.spot, m.strike, h, m.rate, m.carry, m.vol, m.time, m.type, dir,
                                          knock, rebate, mpts, n, seed);
        });
        return mc_response(res.value.price, res.value.standard_error, res.value.path_count, seed,
                           res.elapsed_seconds);
    }
    if (product == "lookback") {
        const Market m = market(r, false);
        const auto style = choose(r, "style", "fixed", "floating", mcd::StrikeStyle::Fixed,
                                  mcd::StrikeStyle::Floating);
        const int mpts = integer(r, "monitoring_points");
        const std::uint64_t n = count(r, "path_count"), seed = count(r, "seed");
        auto res = timed([&] {
            return P::monte_carlo_lookback(m.spot, m.strike, m.rate, m.carry, m.vol, m.time, m.type, style,
                                           mpts, n, seed);
        });
        return mc_response(res.value.price, res.value.standard_error, res.value.path_count, seed,
                           res.elapsed_seconds);
    }
    if (product == "american") {
        const Market m = market(r);
        const int mpts = integer(r, "monitoring_points");
        const std::uint64_t n = count(r, "path_count"), seed = count(r, "seed");
        auto res = timed([&] {
            return P::monte_carlo_lsm_american(m.spot, m.strike, m.rate, m.carry, m.vol, m.time, m.type,
                                               mpts, n, seed);
        });
        return mc_response(res.value.price, res.value.standard_error, n, seed, res.elapsed_seconds);
    }
    if (product == "heston") {
        const double spot = num(r, "spot"), strike = num(r, "strike");
        const mcd::models::HestonParams hp{spot, num(r, "rate"), num(r, "carry_yield"), num(r, "v0"),
                                           num(r, "kappa"), num(r, "theta"), num(r, "xi"), num(r, "rho"),
                                           num(r, "time")};
        const mcd::OptionType type = option_type(r);
        if (kind == "semi_analytic") {
            auto res = mcd::time_call([&] { return P::heston_semi_analytic(hp, strike, type); });
            return price_response(res.value, res.elapsed_seconds);
        }
        const int mpts = integer(r, "monitoring_points");
        const std::uint64_t n = count(r, "path_count"), seed = count(r, "seed");
        auto res = timed([&] { return P::heston_qe_european(hp, strike, type, mpts, n, seed); });
        return mc_response(res.value.price, res.value.standard_error, res.value.path_count, seed,
                           res.elapsed_seconds);
    }
    throw std::runtime_error("unknown 'product': '" + product + "'");
}

} // namespace

int main() {
    const std::string input((std::istreambuf_iterator<char>(std::cin)), std::istreambuf_iterator<char>());
    try {
        Reader rd{input};
        const Object req = rd.object();
        std::cout << write(handle(req)) << "\n";
        return 0;
    } catch (const std::exception& e) {
        std::cout << write(Object{{"error", Value::string(e.what())}}) << "\n";
        return 1;
    }
}
