// generate.hip -- host-side synthetic inputs (no reference counterpart).
//
// dem_generate_pour: polydisperse loose packing by random sequential addition.  Radii are
// log-normal (median, sigma) clipped to [rmin, rmax]; spheres are placed largest first at uniform
// random positions inside the box (keeping a radius of clearance from every face), rejecting any
// position that overlaps an already placed sphere; the output order is a random permutation so
// that particle index carries no spatial information.  Deterministic for a given seed.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "../../include/demhip.h"

namespace {

struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    double normal() {  // Box-Muller
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
    }
};

}  // namespace

extern "C" int32_t dem_generate_pour(int64_t n, const double lo[3], const double hi[3],
                                     double median_r, double sigma, double rmin, double rmax,
                                     double density, uint64_t seed, dem_grain_in_t* out) {
    if (n <= 0 || !lo || !hi || !out || !(rmin > 0) || !(rmax >= rmin) || !(median_r > 0))
        return DEM_ERR_INVALID;
    for (int a = 0; a < 3; ++a)
        if (!(hi[a] - lo[a] > 2.0 * rmax)) return DEM_ERR_INVALID;
    Rng rng{seed};
    std::vector<double> r(n);
    for (int64_t i = 0; i < n; ++i)
        r[i] = std::min(std::max(median_r * exp(sigma * rng.normal()), rmin), rmax);
    std::vector<int64_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return r[a] > r[b]; });

    const double cs = 2.0 * rmax;
    int64_t dims[3];
    for (int a = 0; a < 3; ++a) dims[a] = std::max<int64_t>(1, (int64_t)ceil((hi[a] - lo[a]) / cs));
    const int64_t ncell = dims[0] * dims[1] * dims[2];
    std::vector<int64_t> head(ncell, -1), next(n, -1);
    std::vector<double> px(n), py(n), pz(n);
    auto cell = [&](double x, double y, double z, int64_t c[3]) {
        c[0] = std::min<int64_t>(dims[0] - 1, std::max<int64_t>(0, (int64_t)((x - lo[0]) / cs)));
        c[1] = std::min<int64_t>(dims[1] - 1, std::max<int64_t>(0, (int64_t)((y - lo[1]) / cs)));
        c[2] = std::min<int64_t>(dims[2] - 1, std::max<int64_t>(0, (int64_t)((z - lo[2]) / cs)));
    };
    for (int64_t k = 0; k < n; ++k) {
        const int64_t i = order[k];
        const double ri = r[i];
        bool placed = false;
        for (int trial = 0; trial < 100000 && !placed; ++trial) {
            const double x = lo[0] + ri + rng.uniform() * (hi[0] - lo[0] - 2.0 * ri);
            const double y = lo[1] + ri + rng.uniform() * (hi[1] - lo[1] - 2.0 * ri);
            const double z = lo[2] + ri + rng.uniform() * (hi[2] - lo[2] - 2.0 * ri);
            int64_t c[3];
            cell(x, y, z, c);
            bool ok = true;
            for (int64_t dz = -1; dz <= 1 && ok; ++dz)
                for (int64_t dy = -1; dy <= 1 && ok; ++dy)
                    for (int64_t dx = -1; dx <= 1 && ok; ++dx) {
                        const int64_t cx = c[0] + dx, cy = c[1] + dy, cz = c[2] + dz;
                        if (cx < 0 || cy < 0 || cz < 0 || cx >= dims[0] || cy >= dims[1] || cz >= dims[2])
                            continue;
                        for (int64_t j = head[(cz * dims[1] + cy) * dims[0] + cx]; j >= 0; j = next[j]) {
                            const double ex = px[j] - x, ey = py[j] - y, ez = pz[j] - z;
                            const double rs = (ri + r[j]) * (1.0 + 1e-9);
                            if (ex * ex + ey * ey + ez * ez <= rs * rs) { ok = false; break; }
                        }
                    }
            if (!ok) continue;
            px[i] = x; py[i] = y; pz[i] = z;
            const int64_t id = (c[2] * dims[1] + c[1]) * dims[0] + c[0];
            next[i] = head[id];
            head[id] = i;
            placed = true;
        }
        if (!placed) return DEM_ERR_CAPACITY;
    }
    // random output order (Fisher-Yates)
    std::vector<int64_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    for (int64_t k = n - 1; k > 0; --k) {
        const int64_t j = (int64_t)(rng.next() % (uint64_t)(k + 1));
        std::swap(perm[k], perm[j]);
    }
    for (int64_t k = 0; k < n; ++k) {
        const int64_t i = perm[k];
        dem_grain_in_t& g = out[k];
        memset(&g, 0, sizeof(g));
        g.id = (int32_t)(k + 1);
        g.gid = 0;
        g.V = 4.0 / 3.0 * M_PI * r[i] * r[i] * r[i];
        g.m = density * g.V;
        g.k[0] = px[i]; g.k[1] = py[i]; g.k[2] = pz[i];
    }
    return DEM_OK;
}
