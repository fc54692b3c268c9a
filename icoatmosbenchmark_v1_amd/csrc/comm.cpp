// Region topology and halo map of the icosahedral grid, for the on-device halo
// fill that replaces the reference's MPI halo exchange.
//
// The rules are those of the reference's communication kernel (paths relative
// to kernels/NICAM/communication/src/):
//   region link table      mod_adm.f90:629-801   (RGNMNG_generate_ico)
//   vertex walk-around     mod_adm.f90:808-894   (RGNMNG_vertex_walkaround)
//   singular-point flags   mod_adm.f90:257-265
//   region -> region halo  mod_comm.f90:213-609  (COMM_list_generate)
//   pole <-> region        mod_comm.f90:924-1200 (COMM_sortdest_pl)
//   singular points        mod_comm.f90:1384-1502 (COMM_sortdest_singular),
//                          applied after the exchange (mod_comm.f90:1868-1881)
// With every region on one GPU there is no message to send: the exchange is
// one gather of halo points from their source points.  A singular-point copy
// reads a halo point that the exchange itself fills, so its source is
// resolved here to that point's own source, and one gather pass suffices.
#include <cstring>
#include <map>
#include <vector>

#include "internal.h"

namespace nicam {
namespace {

enum { SW = 1, NW = 2, NE = 3, SE = 4 };  // edge directions (mod_adm.f90:39-42)
enum { VW = 1, VN = 2, VE = 3, VS = 4 };  // vertex directions (mod_adm.f90:45-48)
constexpr int kVlink = 5;                 // ADM_vlink

struct Topology {
    int rall = 0;
    // [rgn][dir] (1-based) -> neighbour region and the direction seen from it
    std::vector<int> edge_rgn, edge_dir;
    // [rgn][vertex] -> number of regions around the vertex, and the walk
    std::vector<int> vert_num;
    std::vector<int> vert_rgn, vert_dir;  // [rgn][vertex][n]
    int pl_rgn[2][kVlink + 1];            // pole walks (NPL, SPL)
    int pl_dir[2][kVlink + 1];
    int E(int r, int d) const { return (r - 1) * 4 + (d - 1); }
    int V(int r, int v) const { return (r - 1) * 4 + (v - 1); }
    int VT(int r, int v, int n) const { return ((r - 1) * 4 + (v - 1)) * (kVlink + 1) + (n - 1); }
};

// mod_adm.f90:629-801: the 10 diamonds of the icosahedron, each split into
// 2^rlevel x 2^rlevel regions, and their edge neighbours.
Topology build_topology(int rlevel)
{
    static const int dmd[10][4] = {{6, 5, 2, 10}, {10, 1, 3, 9}, {9, 2, 4, 8}, {8, 3, 5, 7}, {7, 4, 1, 6},
                                   {7, 5, 1, 10}, {8, 4, 5, 6}, {9, 3, 4, 7}, {10, 2, 3, 8}, {6, 1, 2, 9}};
    Topology T;
    const int n1 = 1 << rlevel, n2 = n1 * n1;
    T.rall = 10 * n2;
    T.edge_rgn.assign(T.rall * 4, 0);
    T.edge_dir.assign(T.rall * 4, 0);
    auto id = [&](int d, int i, int j) { return (d - 1) * n2 + (j - 1) * n1 + i; };
    for (int d = 1; d <= 10; ++d)
        for (int i = 1; i <= n1; ++i)
            for (int j = 1; j <= n1; ++j) {
                const int r = id(d, i, j);
                const bool north = d <= 5;
                int in, jn, dn, dir;
                // SW
                if (j == 1) {
                    dn = dmd[d - 1][SW - 1];
                    if (north) { in = i; jn = n1; dir = NE; }
                    else { in = n1; jn = n1 + 1 - i; dir = SE; }
                } else { in = i; jn = j - 1; dn = d; dir = NE; }
                T.edge_rgn[T.E(r, SW)] = id(dn, in, jn);
                T.edge_dir[T.E(r, SW)] = dir;
                // NW
                if (i == 1) {
                    dn = dmd[d - 1][NW - 1];
                    if (north) { in = n1 + 1 - j; jn = n1; dir = NE; }
                    else { in = n1; jn = j; dir = SE; }
                } else { in = i - 1; jn = j; dn = d; dir = SE; }
                T.edge_rgn[T.E(r, NW)] = id(dn, in, jn);
                T.edge_dir[T.E(r, NW)] = dir;
                // NE
                if (j == n1) {
                    dn = dmd[d - 1][NE - 1];
                    if (north) { in = 1; jn = n1 + 1 - i; dir = NW; }
                    else { in = i; jn = 1; dir = SW; }
                } else { in = i; jn = j + 1; dn = d; dir = SW; }
                T.edge_rgn[T.E(r, NE)] = id(dn, in, jn);
                T.edge_dir[T.E(r, NE)] = dir;
                // SE
                if (i == n1) {
                    dn = dmd[d - 1][SE - 1];
                    if (north) { in = 1; jn = j; dir = NW; }
                    else { in = n1 + 1 - j; jn = 1; dir = SW; }
                } else { in = i + 1; jn = j; dn = d; dir = NW; }
                T.edge_rgn[T.E(r, SE)] = id(dn, in, jn);
                T.edge_dir[T.E(r, SE)] = dir;
            }
    // mod_adm.f90:808-894: walk around each corner of each region
    T.vert_num.assign(T.rall * 4, -1);
    T.vert_rgn.assign(T.rall * 4 * (kVlink + 1), -1);
    T.vert_dir.assign(T.rall * 4 * (kVlink + 1), -1);
    for (int r = 1; r <= T.rall; ++r)
        for (int v = VW; v <= VS; ++v) {
            int rg = r, dir = (v == VW) ? SW : (v == VN) ? NW : (v == VE) ? NE : SE;
            int n = 0;
            while (true) {
                const int rn = T.edge_rgn[T.E(rg, dir)];
                int dn = T.edge_dir[T.E(rg, dir)] - 1;
                if (dn == 0) dn = 4;
                ++n;
                if (n > kVlink) break;  // cannot happen on the icosahedron
                T.vert_rgn[T.VT(r, v, n)] = rg;
                T.vert_dir[T.VT(r, v, n)] = dir;
                rg = rn;
                dir = dn;
                if (rg == r) break;
            }
            T.vert_num[T.V(r, v)] = n;
        }
    for (int p = 0; p < 2; ++p) {
        const int vv = p == 0 ? VN : VS;
        for (int r = 1; r <= T.rall; ++r)
            if (T.vert_num[T.V(r, vv)] == kVlink) {
                for (int n = 1; n <= kVlink; ++n) {
                    T.pl_rgn[p][n] = T.vert_rgn[T.VT(r, vv, n)];
                    T.pl_dir[p][n] = T.vert_dir[T.VT(r, vv, n)];
                }
                break;
            }
    }
    return T;
}

}  // namespace

// Which corners of the 10 diamonds are the same icosahedron vertex.  Two
// diamonds that share an edge share its two end points; the halo copy rules
// of mod_comm.f90:259-348 (read at rlevel 0, where a region is a diamond) say
// which ends meet: across a SW edge seen as the neighbour's NE edge, our W is
// its N and our S its E; seen as its SE edge, our W is its E and our S its S;
// across a NW edge seen as its SE edge, our W is its S and our N its E; seen
// as its NE edge, our W is its E and our N its N.  cls[d][c] (corners W, N,
// E, S) gets a class id 0..n-1; returns n (12 on the icosahedron).
int diamond_corner_classes(int cls[10][4])
{
    const Topology T = build_topology(0);
    enum { W = 0, N = 1, E = 2, S = 3 };
    int parent[40];
    for (int i = 0; i < 40; ++i) parent[i] = i;
    auto find = [&](int x) {
        while (parent[x] != x) x = parent[x] = parent[parent[x]];
        return x;
    };
    auto unite = [&](int d1, int c1, int d2, int c2) {
        const int a = find(d1 * 4 + c1), b = find(d2 * 4 + c2);
        if (a != b) parent[a] = b;
    };
    for (int d = 1; d <= 10; ++d) {
        const int rs = T.edge_rgn[T.E(d, SW)], ds = T.edge_dir[T.E(d, SW)];
        if (ds == NE) { unite(d - 1, W, rs - 1, N); unite(d - 1, S, rs - 1, E); }
        else { unite(d - 1, W, rs - 1, E); unite(d - 1, S, rs - 1, S); }  // SE
        const int rn = T.edge_rgn[T.E(d, NW)], dn = T.edge_dir[T.E(d, NW)];
        if (dn == SE) { unite(d - 1, W, rn - 1, S); unite(d - 1, N, rn - 1, E); }
        else { unite(d - 1, W, rn - 1, E); unite(d - 1, N, rn - 1, N); }  // NE
    }
    int id[40], n = 0;
    for (int i = 0; i < 40; ++i) id[i] = -1;
    for (int i = 0; i < 40; ++i) {
        const int r = find(i);
        if (id[r] < 0) id[r] = n++;
        cls[i / 4][i % 4] = id[r];
    }
    return n;
}

int build_halo_map(int rlevel, int gall_1d, int gall_pl, std::vector<HaloPoint> &dst, std::vector<HaloPoint> &src,
                   std::vector<int> &have_sgp, HaloPoint pole_src[2])
{
    if (rlevel < 0 || rlevel > 8) return fail(NICAM_ERR_ARG, "rlevel must be 0..8");
    if (gall_1d < 4) return fail(NICAM_ERR_ARG, "gall_1d must be >= 4");
    if (gall_pl != kVlink + 1) return fail(NICAM_ERR_ARG, "the icosahedral pole ring needs gall_pl = 6");
    const Topology T = build_topology(rlevel);
    const int gmin = 2, gmax = gall_1d - 1, ginnar = gmax - gmin + 1;  // 1-based, as the reference
    auto suf = [&](int i, int j) { return (j - 1) * gall_1d + i; };
    std::map<HaloPoint, HaloPoint> m;  // destination -> source, later entries win
    auto put = [&](HaloPoint d, HaloPoint s) { m[d] = s; };
    auto R = [&](int rgn, int g1) { return HaloPoint{0, rgn - 1, g1 - 1}; };
    auto P = [&](int pl, int v1) { return HaloPoint{1, pl - 1, v1 - 1}; };

    // mod_comm.f90:246-596 region -> region (one process owns every region: l = rgnid)
    for (int r = 1; r <= T.rall; ++r) {
        const int dSW = T.edge_dir[T.E(r, SW)], rSW = T.edge_rgn[T.E(r, SW)];
        const int dNW = T.edge_dir[T.E(r, NW)], rNW = T.edge_rgn[T.E(r, NW)];
        const int dNE = T.edge_dir[T.E(r, NE)], rNE = T.edge_rgn[T.E(r, NE)];
        const int dSE = T.edge_dir[T.E(r, SE)], rSE = T.edge_rgn[T.E(r, SE)];
        for (int n = 1; n <= ginnar; ++n) {
            if (dSW == NE) put(R(r, suf(gmin - 1 + n, gmin - 1)), R(rSW, suf(gmin - 1 + n, gmax)));
            if (dSW == SE) put(R(r, suf(gmin - 1 + n, gmin - 1)), R(rSW, suf(gmax, gmax + 1 - n)));
            if (dNW == SE) put(R(r, suf(gmin - 1, gmin - 1 + n)), R(rNW, suf(gmax, gmin - 1 + n)));
            if (dNW == NE) put(R(r, suf(gmin - 1, gmin - 1 + n)), R(rNW, suf(gmax + 1 - n, gmax)));
            if (dNE == SW) put(R(r, suf(gmin - 1 + n, gmax + 1)), R(rNE, suf(gmin - 1 + n, gmin)));
            if (dNE == NW) put(R(r, suf(gmin + n, gmax + 1)), R(rNE, suf(gmin, gmax + 1 - n)));
            if (dSE == NW) put(R(r, suf(gmax + 1, gmin - 1 + n)), R(rSE, suf(gmin, gmin - 1 + n)));
            if (dSE == SW) put(R(r, suf(gmax + 1, gmin + n)), R(rSE, suf(gmax + 1 - n, gmin)));
        }
        auto corner = [&](int dir, int &i, int &j) {
            if (dir == VN) { i = gmin; j = gmax; }
            else if (dir == VE) { i = gmax; j = gmax; }
            else if (dir == VS) { i = gmax; j = gmin; }
        };
        if (T.vert_num[T.V(r, VW)] == 4) {  // :444-469
            const int rr = T.vert_rgn[T.VT(r, VW, 3)], dd = T.vert_dir[T.VT(r, VW, 3)];
            int i = -1, j = -1;
            corner(dd, i, j);
            if (i > 0) put(R(r, suf(gmin - 1, gmin - 1)), R(rr, suf(i, j)));
        }
        if (T.vert_num[T.V(r, VN)] == 4) {  // :472-516
            const int rr = T.vert_rgn[T.VT(r, VN, 3)], dd = T.vert_dir[T.VT(r, VN, 3)];
            if (dd == VW) put(R(r, suf(gmin, gmax + 1)), R(rr, suf(gmin, gmin)));
            int i = -1, j = -1;
            if (dd == VW) { i = gmin; j = gmin + 1; }
            else corner(dd, i, j);
            if (i > 0) put(R(r, suf(gmin - 1, gmax + 1)), R(rr, suf(i, j)));
        }
        if (T.vert_num[T.V(r, VE)] == 4) {  // :519-539
            const int rr = T.vert_rgn[T.VT(r, VE, 3)], dd = T.vert_dir[T.VT(r, VE, 3)];
            if (dd == VW) put(R(r, suf(gmax + 1, gmax + 1)), R(rr, suf(gmin, gmin)));
        }
        if (T.vert_num[T.V(r, VS)] == 4) {  // :542-586
            const int rr = T.vert_rgn[T.VT(r, VS, 3)], dd = T.vert_dir[T.VT(r, VS, 3)];
            if (dd == VW) put(R(r, suf(gmax + 1, gmin)), R(rr, suf(gmin, gmin)));
            int i = -1, j = -1;
            if (dd == VW) { i = gmin + 1; j = gmin; }
            else corner(dd, i, j);
            if (i > 0) put(R(r, suf(gmax + 1, gmin - 1)), R(rr, suf(i, j)));
        }
    }
    // mod_comm.f90:1000-1095 region inner point -> pole ring, pole centre -> region halo
    for (int r = 1; r <= T.rall; ++r)
        for (int pl = 1; pl <= 2; ++pl) {
            const int vv = pl == 1 ? VN : VS;
            if (T.vert_num[T.V(r, vv)] != kVlink) continue;
            const int i_from = pl == 1 ? gmin : gmax, j_from = pl == 1 ? gmax : gmin;
            const int i_to = pl == 1 ? gmin : gmax + 1, j_to = pl == 1 ? gmax + 1 : gmin;
            int pl_to = -1;
            for (int v = 1; v <= kVlink; ++v)
                if (T.pl_rgn[pl - 1][v] == r) {
                    pl_to = v;
                    if (pl_to < 2) pl_to = kVlink + 1;  // ADM_gmin_pl / ADM_gmax_pl
                    break;
                }
            if (pl_to > 0) put(P(pl, pl_to), R(r, suf(i_from, j_from)));
            put(R(r, suf(i_to, j_to)), P(pl, 1));
        }
    // mod_comm.f90:1409-1455 singular points, halo to halo, after the exchange
    have_sgp.assign(T.rall, 0);
    for (int r = 1; r <= T.rall; ++r) {
        if (T.vert_num[T.V(r, VW)] == 3) have_sgp[r - 1] = 1;  // mod_adm.f90:262-264
        std::vector<std::pair<HaloPoint, HaloPoint>> cp;
        if (T.vert_num[T.V(r, VW)] == 3) cp.push_back({R(r, suf(gmin - 1, gmin - 1)), R(r, suf(gmin, gmin - 1))});
        if (T.vert_num[T.V(r, VN)] != 4) cp.push_back({R(r, suf(gmin - 1, gmax + 1)), R(r, suf(gmin - 1, gmax))});
        if (T.vert_num[T.V(r, VS)] != 4) cp.push_back({R(r, suf(gmax + 1, gmin - 1)), R(r, suf(gmax + 1, gmin))});
        for (auto &c : cp) {
            auto it = m.find(c.second);
            if (it == m.end()) return fail(NICAM_ERR_ARG, "singular-point source is not a filled halo point");
            m[c.first] = it->second;
        }
    }
    // mod_comm.f90:1895-2101 (COMM_var): before its exchange, each pole centre
    // takes the value of the pole-adjacent halo point of RGNMNG_rgn4pl, the
    // first region whose north (south) vertex is the pole (mod_adm.f90:393-405)
    for (int pl = 1; pl <= 2; ++pl) {
        const int vv = pl == 1 ? VN : VS;
        for (int r = 1; r <= T.rall; ++r)
            if (T.vert_num[T.V(r, vv)] == kVlink) {
                pole_src[pl - 1] = pl == 1 ? R(r, suf(gmin, gmax + 1)) : R(r, suf(gmax + 1, gmin));
                break;
            }
    }
    dst.clear();
    src.clear();
    for (auto &kv : m) {
        dst.push_back(kv.first);
        src.push_back(kv.second);
    }
    return NICAM_OK;
}

}  // namespace nicam

using namespace nicam;

extern "C" int nicam_halo_map(int rlevel, int gall_1d, int gall_pl, int capacity, int *n_out, int *dst_kind,
                              int *dst_l, int *dst_g, int *src_kind, int *src_l, int *src_g, int *have_sgp)
{
    if (!n_out) return fail(NICAM_ERR_ARG, "null n_out");
    std::vector<HaloPoint> d, s;
    std::vector<int> sgp;
    HaloPoint ps[2];
    const int rc = build_halo_map(rlevel, gall_1d, gall_pl, d, s, sgp, ps);
    if (rc != NICAM_OK) return rc;
    *n_out = (int)d.size();
    if (have_sgp) std::memcpy(have_sgp, sgp.data(), sizeof(int) * sgp.size());
    if (capacity <= 0) return NICAM_OK;  // size query
    if (capacity < (int)d.size()) return fail(NICAM_ERR_ARG, "capacity smaller than the halo map");
    if (!dst_kind || !dst_l || !dst_g || !src_kind || !src_l || !src_g) return fail(NICAM_ERR_ARG, "null output");
    for (size_t e = 0; e < d.size(); ++e) {
        dst_kind[e] = d[e].kind;
        dst_l[e] = d[e].l;
        dst_g[e] = d[e].g;
        src_kind[e] = s[e].kind;
        src_l[e] = s[e].l;
        src_g[e] = s[e].g;
    }
    return NICAM_OK;
}

extern "C" int nicam_pole_fill_map(int rlevel, int gall_1d, int *src_l, int *src_g)
{
    if (!src_l || !src_g) return fail(NICAM_ERR_ARG, "null output");
    std::vector<HaloPoint> d, s;
    std::vector<int> sgp;
    HaloPoint ps[2];
    const int rc = build_halo_map(rlevel, gall_1d, kVlink + 1, d, s, sgp, ps);
    if (rc != NICAM_OK) return rc;
    for (int p = 0; p < 2; ++p) {
        src_l[p] = ps[p].l;
        src_g[p] = ps[p].g;
    }
    return NICAM_OK;
}
