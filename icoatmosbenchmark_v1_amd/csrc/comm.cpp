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

// The shared faces of neighbouring regions.  Two regions sharing an edge each
// compute the flux through every arc of that edge from their own copies of the
// inputs and metrics, so the two values agree only to rounding.  For an update
// that conserves mass exactly, each such face is given one value: the one of
// the region with the lower index (regions before the poles).  For every
// inner cell whose neighbour through face m (1..6) is a halo point, the partner
// is the cell that halo point copies, and its face whose neighbour is a halo
// point copied from this cell.  Codes in tab (6 per edge cell): -1 = the
// region's own value, -2 = no flux (face 6 of a singular-point cell, which the
// flux sets to zero), else (l' * gall + g') * 8 + m' = minus the partner's
// outward transfer.  pole_tab[p * (gall_pl - 1) + v - 1]: the partner of the
// pole centre's face toward ring point v, coded the same way.
int build_arc_partners(int rlevel, int gall_1d, int gall_pl, const std::vector<HaloPoint> &dst,
                       const std::vector<HaloPoint> &src, const std::vector<int> &have_sgp, std::vector<int> &slot,
                       std::vector<int> &tab, std::vector<int> &pole_tab)
{
    const int gall = gall_1d * gall_1d, lall = 10 << (2 * rlevel), npl = gall_pl - 1;
    const int lo = 1, hi = gall_1d - 2;  // inner range, 0-based
    std::map<HaloPoint, HaloPoint> srcof;
    for (size_t e = 0; e < dst.size(); ++e) srcof[dst[e]] = src[e];
    const int di[7] = {0, 1, 1, 0, -1, -1, 0}, dj[7] = {0, 0, 1, 1, 0, -1, -1};
    auto inner = [&](int i, int j) { return i >= lo && i <= hi && j >= lo && j <= hi; };
    auto zero_face = [&](int l, int g, int m) { return m == 6 && have_sgp[l] && g == lo * gall_1d + lo; };
    // the face of region cell (l, g) whose neighbour is a halo point copied from `from`
    auto face_toward = [&](int l, int g, const HaloPoint &from, int &m_out) -> int {
        const int i = g % gall_1d, j = g / gall_1d;
        int found = 0;
        for (int m = 1; m <= 6; ++m) {
            const int ni = i + di[m], nj = j + dj[m];
            if (inner(ni, nj) || zero_face(l, g, m)) continue;
            auto it = srcof.find(HaloPoint{0, l, nj * gall_1d + ni});
            if (it == srcof.end()) continue;
            const HaloPoint &s = it->second;
            if (s.kind == from.kind && s.l == from.l && s.g == from.g) {
                m_out = m;
                ++found;
            }
        }
        return found;
    };
    slot.assign((size_t)lall * gall, -1);
    tab.clear();
    for (int l = 0; l < lall; ++l)
        for (int j = lo; j <= hi; ++j)
            for (int i = lo; i <= hi; ++i) {
                if (i != lo && i != hi && j != lo && j != hi) continue;
                const int g = j * gall_1d + i;
                int codes[6];
                for (int m = 1; m <= 6; ++m) {
                    const int ni = i + di[m], nj = j + dj[m];
                    codes[m - 1] = -1;
                    if (inner(ni, nj)) continue;
                    if (zero_face(l, g, m)) {
                        codes[m - 1] = -2;
                        continue;
                    }
                    auto it = srcof.find(HaloPoint{0, l, nj * gall_1d + ni});
                    if (it == srcof.end()) return fail(NICAM_ERR_ARG, "arc partners: a halo neighbour has no source");
                    const HaloPoint &s = it->second;
                    if (s.kind != 0 || s.l > l) continue;  // this region's value is the one both sides use
                    if (s.l == l) return fail(NICAM_ERR_ARG, "arc partners: a region borders itself");
                    int mp = 0;
                    if (face_toward(s.l, s.g, HaloPoint{0, l, g}, mp) != 1)
                        return fail(NICAM_ERR_ARG, "arc partners: no unique partner face");
                    codes[m - 1] = (s.l * gall + s.g) * 8 + mp;
                }
                slot[(size_t)l * gall + g] = (int)(tab.size() / 6);
                tab.insert(tab.end(), codes, codes + 6);
            }
    pole_tab.assign(2 * npl, -1);
    for (int p = 0; p < 2; ++p)
        for (int v = 1; v <= npl; ++v) {
            auto it = srcof.find(HaloPoint{1, p, v});
            if (it == srcof.end() || it->second.kind != 0)
                return fail(NICAM_ERR_ARG, "arc partners: a pole ring point has no region source");
            const HaloPoint &s = it->second;
            int mp = 0;
            if (face_toward(s.l, s.g, HaloPoint{1, p, 0}, mp) != 1)
                return fail(NICAM_ERR_ARG, "arc partners: no unique face toward the pole");
            pole_tab[p * npl + v - 1] = (s.l * gall + s.g) * 8 + mp;
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
