// Drop-in counterpart of the reference's include/mcd/core/gauss_legendre.hpp:22-33.
#pragma once

#include <cstddef>
#include <vector>

namespace mcd {

struct GaussLegendreRule {
    std::vector<double> nodes;   // on [-1, 1], ascending
    std::vector<double> weights; // positive, one per node
};

// The `order`-point rule (exact for polynomials of degree <= 2 order - 1), nodes found by
// Newton's method on the Legendre polynomial P_order.
[[nodiscard]] GaussLegendreRule gauss_legendre_rule(int order);

// The integral of f over [a, b] by the `order`-point rule mapped affinely onto [a, b].
template <typename F>
[[nodiscard]] double gauss_legendre_integrate(const F& f, double a, double b, int order) {
    const GaussLegendreRule rule = gauss_legendre_rule(order);
    const double half = 0.5 * (b - a), mid = 0.5 * (a + b);
    double acc = 0.0;
    for (std::size_t k = 0; k < rule.nodes.size(); ++k) acc += rule.weights[k] * f(mid + half * rule.nodes[k]);
    return half * acc;
}

} // namespace mcd
