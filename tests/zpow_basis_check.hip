// Host check of the propagation's z-power basis (compiled and run by
// test_zpow_basis.py): for random unit directions, the scaled real harmonics
// Rt_l,+-m = C_m Pt_l^m, S_m Pt_l^m formed by their recurrences must equal
// sum_j M_aj Z_j with M = ZPowBasis<L> and Z_l,+-m = (C_m, S_m) z^(l-m) taken
// directly, and pair_zpow must form those Z.  Prints one line per order L.
#include "../doi_amd/csrc/rte_device.hpp"

#include <cmath>
#include <cstdio>
#include <random>

using namespace rte;

template <int L>
static void check()
{
    constexpr int K = (L + 1) * (L + 1);
    static constexpr ZPowBasis<L> zb{};
    std::mt19937 gen(7 + L);
    std::normal_distribution<double> nd;
    double worst = 0.0, worst_pair = 0.0;
    for (int t = 0; t < 4000; ++t) {
        double x = nd(gen), y = nd(gen), z = nd(gen);
        const double r = std::sqrt(x * x + y * y + z * z);
        x /= r;
        y /= r;
        z /= r;
        // (x + i y)^m from the polar form, Pt_l^m by the monic recurrence
        const double rho = std::sqrt(x * x + y * y), phi = std::atan2(y, x);
        double R[K], Z[K];
        for (int m = 0; m <= L; ++m) {
            const double C = std::pow(rho, m) * std::cos(m * phi), S = std::pow(rho, m) * std::sin(m * phi);
            double P[L + 2] = {};
            P[m] = 1.0;
            if (m + 1 <= L) P[m + 1] = z;
            for (int l = m + 2; l <= L; ++l) P[l] = z * P[l - 1] - sh_gamma<double>(l, m) * P[l - 2];
            for (int l = m; l <= L; ++l) {
                R[l * l + l + m] = C * P[l];
                Z[l * l + l + m] = C * std::pow(z, l - m);
                if (m) {
                    R[l * l + l - m] = S * P[l];
                    Z[l * l + l - m] = S * std::pow(z, l - m);
                }
            }
        }
        for (int a = 0; a < K; ++a) {
            double s = 0.0;
            for (int j = 0; j < K; ++j) s += zb.m[a][j] * Z[j];
            worst = std::fmax(worst, std::fabs(s - R[a]));
        }
        PairZ<double, L> G;
        pair_zpow<double, L>(G, x, y, z);
        for (int j = 1; j < K; ++j) worst_pair = std::fmax(worst_pair, std::fabs(G.R[j] - Z[j]));
    }
    std::printf("L %d basis %.3e pair %.3e\n", L, worst, worst_pair);
}

int main()
{
    check<0>();
    check<1>();
    check<2>();
    check<3>();
    check<4>();
    check<5>();
    check<6>();
    check<7>();
    return 0;
}
