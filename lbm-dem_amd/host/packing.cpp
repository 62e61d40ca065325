// Seeded dense sphere packings for the benchmark workloads (no reference counterpart: the
// reference reads input/particles.dat as given). Random sequential addition at a reduced radius
// followed by radius growth with overlap relaxation, inside an axis-aligned box.
//
//   int lbdem_pack_spheres(int n, double r, const double lo[3], const double hi[3],
//                          unsigned long long seed, int relax_steps, double* x /*3n*/,
//                          double* max_overlap)
// Returns 0, or 1 when RSA could not place n spheres at the reduced radius.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace {
struct Rng {  // splitmix64
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uni() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct Grid {
    double lo[3], cell;
    int n[3];
    std::vector<std::vector<int>> items;
    Grid(const double* l, const double* h, double c) : cell(c) {
        for (int a = 0; a < 3; ++a) { lo[a] = l[a]; n[a] = std::max(1, int(std::ceil((h[a] - l[a]) / c))); }
        items.assign(size_t(n[0]) * n[1] * n[2], {});
    }
    int key(const double* p, int* c) const {
        for (int a = 0; a < 3; ++a) c[a] = std::min(n[a] - 1, std::max(0, int((p[a] - lo[a]) / cell)));
        return (c[2] * n[1] + c[1]) * n[0] + c[0];
    }
};
}  // namespace

extern "C" int lbdem_pack_spheres(int n, double r, const double lo[3], const double hi[3], unsigned long long seed,
                                  int relax_steps, double* x, double* max_overlap) {
    Rng rng(seed);
    const double vol = (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
    const double phi = n * 4.0 / 3.0 * M_PI * r * r * r / vol;
    const double r0 = phi > 0.28 ? r * std::cbrt(0.28 / phi) : r;   // RSA radius well below the jamming limit
    // RSA with a cell grid
    Grid g(lo, hi, 2.0 * r);
    int placed = 0;
    long long tries = 0;
    while (placed < n && tries < 4000ll * n) {
        ++tries;
        double p[3];
        for (int a = 0; a < 3; ++a) p[a] = lo[a] + r0 + (hi[a] - lo[a] - 2 * r0) * rng.uni();
        int c[3];
        g.key(p, c);
        bool ok = true;
        for (int dz = -1; dz <= 1 && ok; ++dz)
            for (int dy = -1; dy <= 1 && ok; ++dy)
                for (int dx = -1; dx <= 1 && ok; ++dx) {
                    const int cx = c[0] + dx, cy = c[1] + dy, cz = c[2] + dz;
                    if (cx < 0 || cy < 0 || cz < 0 || cx >= g.n[0] || cy >= g.n[1] || cz >= g.n[2]) continue;
                    for (int q : g.items[(size_t(cz) * g.n[1] + cy) * g.n[0] + cx]) {
                        const double d0 = p[0] - x[3 * q], d1 = p[1] - x[3 * q + 1], d2 = p[2] - x[3 * q + 2];
                        if (d0 * d0 + d1 * d1 + d2 * d2 < 4 * r0 * r0) { ok = false; break; }
                    }
                }
        if (!ok) continue;
        for (int a = 0; a < 3; ++a) x[3 * placed + a] = p[a];
        g.items[g.key(p, c)].push_back(placed);
        ++placed;
    }
    if (placed < n) return 1;
    // growth to r with overlap relaxation (Jacobi sweeps of half-overlap displacements)
    std::vector<double> dx(3 * size_t(n));
    double worst = 0.0;
    for (int it = 0; it <= relax_steps; ++it) {
        const double rc = it >= relax_steps ? r : r0 + (r - r0) * std::min(1.0, double(it) / (0.8 * relax_steps));
        Grid gg(lo, hi, 2.0 * r);
        int c[3];
        for (int i = 0; i < n; ++i) gg.items[gg.key(&x[3 * i], c)].push_back(i);
        std::fill(dx.begin(), dx.end(), 0.0);
        worst = 0.0;
        for (int i = 0; i < n; ++i) {
            gg.key(&x[3 * i], c);
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dxc = -1; dxc <= 1; ++dxc) {
                        const int cx = c[0] + dxc, cy = c[1] + dy, cz = c[2] + dz;
                        if (cx < 0 || cy < 0 || cz < 0 || cx >= gg.n[0] || cy >= gg.n[1] || cz >= gg.n[2]) continue;
                        for (int q : gg.items[(size_t(cz) * gg.n[1] + cy) * gg.n[0] + cx]) {
                            if (q == i) continue;
                            double d[3] = {x[3 * i] - x[3 * q], x[3 * i + 1] - x[3 * q + 1], x[3 * i + 2] - x[3 * q + 2]};
                            const double dist = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                            const double ov = 2 * rc * 1.002 - dist;
                            if (ov > 0 && dist > 1e-12) {
                                worst = std::max(worst, ov);
                                for (int a = 0; a < 3; ++a) dx[3 * i + a] += 0.5 * ov * d[a] / dist;
                            }
                        }
                    }
        }
        for (int i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) {
                double v = x[3 * i + a] + dx[3 * i + a];
                v = std::min(hi[a] - rc * 1.001, std::max(lo[a] + rc * 1.001, v));
                x[3 * i + a] = v;
            }
    }
    if (max_overlap) *max_overlap = worst;
    return 0;
}
