// Seeded synthetic heat-step workload (host). The reference has no such
// generator; this one draws every quantity from the reference's keyed
// counter RNG (sim/core/random/counter_rng.hpp:19-45) on the validation
// stream, with the quantity as the tick and the cell or edge id as the key:
//
//   K_e  = k_base (1 + 0.5 sin^2 lat_e) exp(sigma z)       tick 1, key = edge
//   land = harmonic field of degrees 1..4 (24 N(0,1) coefficients, tick 2)
//          above its area-weighted 70th percentile (30 % of the area)
//   C    = 1e7 J/m^2/K on land, 4e8 on ocean
//   T    = 288 - 40 sin^2 lat + 2 z                          tick 3, key = cell
//   S    = 100 cos lat + 10 z                               tick 16 + step
//
// z are Box-Muller normals of samples 0 and 1; lat_e is the latitude of the
// sum of the edge's two cell centres. Identical bits to the oracle's
// generator, which the tests check.
#include "internal.hpp"
#include "workload.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>

namespace psim {
namespace {

constexpr uint64_t splitmix(uint64_t v) {
    v += 0x9e3779b97f4a7c15ULL;
    v = (v ^ (v >> 30U)) * 0xbf58476d1ce4e5b9ULL;
    v = (v ^ (v >> 27U)) * 0x94d049bb133111ebULL;
    return v ^ (v >> 31U);
}

double unit_draw(uint64_t seed, int64_t tick, uint32_t key, uint32_t sample) {
    uint64_t k = splitmix(seed);
    k = splitmix(k ^ uint64_t{1});   // validation stream
    k = splitmix(k ^ static_cast<uint64_t>(tick));
    k = splitmix(k ^ uint64_t{key});
    k = splitmix(k ^ uint64_t{sample});
    return static_cast<double>(k >> 11U) * (1.0 / 9007199254740992.0);
}

double gaussian(uint64_t seed, int64_t tick, uint32_t key) {
    const double u1 = unit_draw(seed, tick, key, 0), u2 = unit_draw(seed, tick, key, 1);
    return std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(6.283185307179586 * u2);
}

// latitude_rad / longitude_rad of local_tangent_basis.cpp:26-34 (the
// argument is normalised first).
void lat_lon(double x, double y, double z, double& lat, double* lon) {
    const double len = std::sqrt(x * x + y * y + z * z);
    const double ux = x / len, uy = y / len, uz = z / len;
    lat = std::asin(std::clamp(uz, -1.0, 1.0));
    if (lon != nullptr) *lon = std::atan2(uy, ux);
}

// Orthonormal real spherical harmonic (without 1/sqrt(4 pi)) of degree l,
// order m at sin(latitude) s and longitude lon, by the upward recurrence in
// degree of the associated Legendre functions.
double real_harmonic(int l, int m, double s, double lon) {
    const int a = std::abs(m);
    const double c = std::sqrt((1.0 - s) * (1.0 + s));
    double pa = 1.0;
    for (int i = 1; i <= a; ++i) pa *= static_cast<double>(2 * i - 1) * c;
    double value = pa;
    if (l > a) {
        double lower = pa, upper = s * static_cast<double>(2 * a + 1) * pa;
        for (int d = a + 2; d <= l; ++d) {
            const double next = (s * static_cast<double>(2 * d - 1) * upper - static_cast<double>(d + a - 1) * lower) /
                                static_cast<double>(d - a);
            lower = upper;
            upper = next;
        }
        value = upper;
    }
    double ratio = 1.0;
    for (int i = l - a + 1; i <= l + a; ++i) ratio /= static_cast<double>(i);
    const double scale = std::sqrt(static_cast<double>(2 * l + 1) * ratio);
    if (m > 0) return scale * value * std::cos(static_cast<double>(a) * lon);
    if (m < 0) return scale * value * std::sin(static_cast<double>(a) * lon);
    return scale * value;
}

}  // namespace

uint64_t instance_seed(uint64_t base, uint64_t index) { return splitmix(base ^ index); }

void make_heat_instance(const HostMesh& m, uint64_t seed, double k_base, double sigma, double* conductance,
                        double* capacity_area, uint8_t* land, double* temperature) {
    const int32_t n = m.cells;
    for (int32_t e = 0; e < m.edges; ++e) {
        const uint32_t a = m.first[e], b = m.second[e];
        double lat = 0.0;
        lat_lon(m.center[3 * a] + m.center[3 * b], m.center[3 * a + 1] + m.center[3 * b + 1],
                m.center[3 * a + 2] + m.center[3 * b + 2], lat, nullptr);
        const double s = std::sin(lat);
        conductance[e] = k_base * (1.0 + 0.5 * (s * s)) * std::exp(sigma * gaussian(seed, 1, static_cast<uint32_t>(e)));
    }
    double coefficient[24];
    for (uint32_t q = 0; q < 24; ++q) coefficient[q] = gaussian(seed, 2, q);
    std::vector<double> field(static_cast<size_t>(n));
    std::vector<double> sin_lat(static_cast<size_t>(n));
    for (int32_t c = 0; c < n; ++c) {
        double lat = 0.0, lon = 0.0;
        lat_lon(m.center[3 * c], m.center[3 * c + 1], m.center[3 * c + 2], lat, &lon);
        const double s = std::sin(lat);
        sin_lat[c] = s;
        double f = 0.0;
        int q = 0;
        for (int l = 1; l <= 4; ++l)
            for (int mm = -l; mm <= l; ++mm) f += coefficient[q++] * real_harmonic(l, mm, s, lon);
        field[c] = f;
    }
    std::vector<int32_t> rank(static_cast<size_t>(n));
    std::iota(rank.begin(), rank.end(), 0);
    std::sort(rank.begin(), rank.end(), [&](int32_t a, int32_t b) { return field[a] != field[b] ? field[a] > field[b] : a < b; });
    double total = 0.0;
    for (int32_t c = 0; c < n; ++c) total += m.area[c];
    const double quota = 0.3 * total;
    double taken = 0.0;
    for (int32_t c : rank) {
        land[c] = taken < quota ? 1 : 0;
        if (land[c] != 0) taken += m.area[c];
    }
    for (int32_t c = 0; c < n; ++c) {
        capacity_area[c] = (land[c] != 0 ? 1.0e7 : 4.0e8) * m.area[c];
        const double s = sin_lat[c];
        temperature[c] = 288.0 - 40.0 * (s * s) + 2.0 * gaussian(seed, 3, static_cast<uint32_t>(c));
    }
}

void make_heat_forcing(const HostMesh& m, uint64_t seed, uint32_t step, double* forcing) {
    for (int32_t c = 0; c < m.cells; ++c) {
        double lat = 0.0;
        lat_lon(m.center[3 * c], m.center[3 * c + 1], m.center[3 * c + 2], lat, nullptr);
        forcing[c] = 100.0 * std::cos(lat) + 10.0 * gaussian(seed, 16 + static_cast<int64_t>(step), static_cast<uint32_t>(c));
    }
}

}  // namespace psim
