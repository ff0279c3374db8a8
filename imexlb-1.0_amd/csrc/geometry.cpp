// geometry.cpp -- host-side lattice set-up (see geometry.hpp).
#include "geometry.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <unordered_map>

namespace imexlb {

void LatticeTable::build(int d)
{
    dim = d;
    nq = d == 2 ? 9 : 27;
    // D2Q9: q = (ex+1)*3 + (ey+1)            (initialization.f90:115-122)
    // D3Q27: q = (ez+1)*9 + (ex+1)*3 + (ey+1) (initialization_3d.f90:111-121)
    for (int q = 0; q < nq; ++q) {
        e[q][0] = (q % 9) / 3 - 1;
        e[q][1] = q % 3 - 1;
        e[q][2] = d == 3 ? q / 9 - 1 : 0;
    }
}

void RefGeometry::init(int d, int nx, int ny, int nz)
{
    dim = d;
    nq = d == 2 ? 9 : 27;
    n[0] = nx;
    n[1] = ny;
    n[2] = d == 3 ? nz : 0;
    for (int i = 0; i < 3; ++i) L[i] = n[i] + 1;
    // ghost width 3 on every side (cart_mpi.f90:10, 133-136)
    dy = L[0] + 6;
    dz = d == 3 ? dy * int64_t(L[1] + 6) : 0;
    A = dy * int64_t(L[1] + 6) * (d == 3 ? int64_t(L[2] + 6) : 1);
    lat.build(d);
}

// SetGeometry, initialization.f90:167-255 / initialization_3d.f90:170-279.
void RefGeometry::build(double pos, double radius, double aspect, double angle_deg)
{
    geo.assign(size_t(A), 0);
    fluid_id.clear(); bl.clear(); br.clear(); bu.clear(); bd.clear(); bf.clear(); bb.clear(); b_user.clear();
    scan_oob = 0;
    const int nx = n[0], ny = n[1], nz = n[2];
    const double yc = double(ny) / 2.0, zc = double(nz) / 2.0, r2 = radius * radius;
    const bool ellipse = dim == 2 && (aspect != 1.0 || angle_deg != 0.0);
    const double th = angle_deg * (3.14159265358979323846 / 180.0);
    const double ct = std::cos(th), st = std::sin(th);
    const double semi_a2 = radius * radius * aspect, semi_b2 = radius * radius / aspect;
    const int zlo = dim == 3 ? -1 : 0, zhi = dim == 3 ? L[2] + 1 : 0;

    // pass 1: interior mask + fluid list, ghost-ring wall lists (scan order z, y, x)
    for (int z = zlo; z <= zhi; ++z)
        for (int y = -1; y <= L[1] + 1; ++y)
            for (int x = -1; x <= L[0] + 1; ++x) {
                const int64_t c = id(x, y, z);
                if (interior(x, y, z)) {
                    const double a = double(x) - pos, b = double(y) - yc;
                    if (ellipse) {
                        const double u = ct * a + st * b, v = ct * b - st * a;
                        geo[c] = u * u / semi_a2 + v * v / semi_b2 > 1.0 ? 1 : 0;
                    } else {
                        double s = a * a + b * b;
                        if (dim == 3) {
                            const double w = double(z) - zc;
                            s = s + w * w;
                        }
                        geo[c] = s > r2 ? 1 : 0;
                    }
                    if (geo[c]) fluid_id.push_back(int32_t(c));
                    continue;
                }
                if (x == -1) bl.push_back(int32_t(c));
                if (x == nx + 1) br.push_back(int32_t(c));
                if (dim == 2) {
                    if (y == -1) bd.push_back(int32_t(c));
                    if (y == ny + 1) bu.push_back(int32_t(c));
                } else {
                    if (y == -1) bf.push_back(int32_t(c));
                    if (y == ny + 1) bb.push_back(int32_t(c));
                    if (z == -1) bd.push_back(int32_t(c));
                    if (z == nz + 1) bu.push_back(int32_t(c));
                }
            }

    // PassInt(geo) at one rank: every ghost cell takes its periodic image
    auto wrap = [](int v, int len) { return ((v % len) + len) % len; };
    const int zg0 = dim == 3 ? -3 : 0, zg1 = dim == 3 ? L[2] + 2 : 0;
    for (int z = zg0; z <= zg1; ++z)
        for (int y = -3; y <= L[1] + 2; ++y) {
            const bool row_ghost = y < 0 || y >= L[1] || (dim == 3 && (z < 0 || z >= L[2]));
            for (int x = -3; x <= L[0] + 2; x = (row_ghost || x >= L[0] || x < -1) ? x + 1 : L[0]) {
                geo[id(x, y, z)] = geo[id(wrap(x, L[0]), wrap(y, L[1]), dim == 3 ? wrap(z, L[2]) : 0)];
            }
        }

    // pass 2: body list.  Per x-row keep the first solid cell whose cell at
    // flat offset e_x(q) + dy*(q*dim+1) [+ dz*(q*dim+2)] is fluid for some q
    // (the offsets use q itself, initialization.f90:240, _3d:263).
    for (int z = zlo; z <= zhi; ++z)
        for (int y = -1; y <= L[1] + 1; ++y)
            for (int x = -1; x <= L[0] + 1; ++x) {
                const int64_t c = id(x, y, z);
                if (geo[c] != 0) continue;
                bool hit = false;
                for (int q = 0; q < nq && !hit; ++q) {
                    int64_t t = c + lat.e[q][0] + dy * int64_t(q * dim + 1);
                    if (dim == 3) t += dz * int64_t(q * dim + 2);
                    if (t < 0 || t >= A) {
                        ++scan_oob;
                        continue;
                    }
                    hit = geo[t] == 1;
                }
                if (hit) {
                    b_user.push_back(int32_t(c));
                    break;  // leave this x-row
                }
            }
}

namespace {

// Phases of BoundaryCondition after the collision, in execution order:
// 0 PassF periodic fill, 1 Up wall (bu), 2 Down wall (bd), 3 User (b_user),
// 4 Inlet (bl), 5 Outlet (br).  lbm.f90:100-153.
struct Resolver {
    const RefGeometry &g;
    const std::vector<int32_t> *lists[6];
    int64_t order_sensitive = 0;

    // Every list is built in scan order, which is increasing cell id, so a
    // cell's position in its list is found by binary search.
    explicit Resolver(const RefGeometry &geom) : g(geom), lists{nullptr, &g.bu, &g.bd, &g.b_user, &g.bl, &g.br}
    {
    }

    bool ghost(int64_t c) const
    {
        int x, y, z;
        g.coords(c, x, y, z);
        return !g.interior(x, y, z);
    }

    bool member(int k, int64_t c, int64_t &p) const
    {
        if (k == 0) {
            p = 0;
            return ghost(c);
        }
        const std::vector<int32_t> &v = *lists[k];
        auto it = std::lower_bound(v.begin(), v.end(), c, [](int32_t a, int64_t b) { return int64_t(a) < b; });
        if (it == v.end() || int64_t(*it) != c) return false;
        p = int64_t(it - v.begin());
        return true;
    }

    int64_t image(int64_t c) const
    {
        int x, y, z;
        g.coords(c, x, y, z);
        auto wrap = [](int v, int len) { return ((v % len) + len) % len; };
        return g.id(wrap(x, g.L[0]), wrap(y, g.L[1]), g.dim == 3 ? wrap(z, g.L[2]) : 0);
    }

    // source slot phase k reads for slot (c, q); false when past the array
    bool source(int k, int64_t c, int q, int64_t &s, int &qs) const
    {
        switch (k) {
        case 1:
        case 2:
        case 3:  // mirror: fb(c,q) = fb(c + 2 e_q, nq-1-q)
            s = c + g.eoff(q, 2);
            qs = g.nq - 1 - q;
            break;
        case 4:  // inlet: fb(c,q) = fb(c+1, q)
            s = c + 1;
            qs = q;
            break;
        default:  // 5, outlet: fb(c,q) = fb(c + e_q, q)
            s = c + g.eoff(q, 1);
            qs = q;
            break;
        }
        return s >= 0 && s < g.A;
    }

    // Value of slot (c,q) after phases 0..k, in FillEntry::ref_src encoding.
    // Before PassF (k < 0) an interior cell holds its post-collision value if
    // fluid, the previous step's value if a body cell, 0 otherwise.
    int64_t value(int64_t c, int q, int k, bool &undefined)
    {
        for (;;) {
            if (k < 0) {
                if (ghost(c)) {  // a stale ghost value: PassF refreshes every ghost cell first
                    undefined = true;
                    return -1;
                }
                if (g.geo[c] == 1) return c * g.nq + q;
                int64_t pu;
                return member(3, c, pu) ? encode_prev(c * g.nq + q) : -1;
            }
            int64_t pc;
            if (!member(k, c, pc)) {
                --k;
                continue;
            }
            if (k == 0) {  // periodic image, sent from the interior
                c = image(c);
                k = -1;
                continue;
            }
            int64_t s;
            int qs;
            if (!source(k, c, q, s, qs)) {
                undefined = true;
                return -1;
            }
            int64_t ps;
            const bool same_phase = member(k, s, ps);
            if (same_phase && ps != pc) ++order_sensitive;
            const bool already = same_phase && (ps < pc || (ps == pc && qs < q));
            c = s;
            q = qs;
            if (!already) --k;
        }
    }
};

}  // namespace

BcTable resolve_boundary(const RefGeometry &g)
{
    BcTable t;
    Resolver r(g);
    std::vector<char> sensitive, pulled;  // per entry
    auto add = [&](int64_t s, int q, bool is_pulled) {
        bool undefined = false;
        const int64_t before = r.order_sensitive;
        const int64_t v = r.value(s, q, 5, undefined);
        if (undefined) {
            t.undefined = true;
            t.error = "boundary value depends on a read past the end of the reference arrays";
        }
        t.entries.push_back(FillEntry{s * g.nq + q, v});
        sensitive.push_back(r.order_sensitive != before);
        pulled.push_back(is_pulled);
    };
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const int zmax = g.dim == 3 ? nz : 0;
    // (a) fluid nodes on the domain shell pull from ghost cells
    for (int z = 0; z <= zmax; ++z)
        for (int y = 0; y <= ny; ++y) {
            const bool full_row = y == 0 || y == ny || (g.dim == 3 && (z == 0 || z == nz));
            for (int x = 0; x <= nx; x = (full_row || x == nx) ? x + 1 : nx) {
                const int64_t c = g.id(x, y, z);
                if (g.geo[c] != 1) continue;
                for (int q = 0; q < g.nq; ++q) {
                    const int sx = x - g.lat.e[q][0], sy = y - g.lat.e[q][1], sz = z - g.lat.e[q][2];
                    if (!g.interior(sx, sy, sz)) add(c - g.eoff(q, 1), q, true);
                }
            }
        }
    // (b) every direction of every interior body cell of the b_user list:
    // fluid nodes pull some of them now, and all of them carry over as the
    // next step's previous values
    for (int32_t s : g.b_user) {
        int x, y, z;
        g.coords(s, x, y, z);
        if (!g.interior(x, y, z)) continue;  // ghost body cells are covered by (a)
        for (int q = 0; q < g.nq; ++q) {
            const int px = x + g.lat.e[q][0], py = y + g.lat.e[q][1], pz = z + g.lat.e[q][2];
            add(s, q, g.interior(px, py, pz) && g.geo[g.id(px, py, pz)] == 1);
        }
    }
    // Order-sensitive reads matter when they reach a fluid node: directly,
    // or through a previous-step body value that was itself order-sensitive.
    std::unordered_map<int64_t, size_t> at;
    for (size_t i = 0; i < t.entries.size(); ++i) at.emplace(t.entries[i].ref_slot, i);
    for (bool changed = true; changed;) {
        changed = false;
        for (size_t i = 0; i < t.entries.size(); ++i) {
            const int64_t src = t.entries[i].ref_src;
            if (sensitive[i] || src > -2) continue;
            auto it = at.find(decode_prev(src));
            if (it != at.end() && sensitive[it->second]) sensitive[i] = changed = true;
        }
    }
    t.order_sensitive = 0;
    for (size_t i = 0; i < t.entries.size(); ++i) t.order_sensitive += sensitive[i] && pulled[i];
    return t;
}

}  // namespace imexlb
