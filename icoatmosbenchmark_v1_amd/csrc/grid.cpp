// The icosahedral grid: positions of every grid point of a glevel/rlevel grid,
// with the halo and pole rings filled by the grid's own halo map.
//
// The reference reads its grid from files (GRD_x, GRD_x_pl in the metrics
// kernel's snapshot, kernels/NICAM/dyn_metrics/src/main.f90:217-218); here it
// is generated.  The icosahedron's 12 vertices are placed with the poles on
// the z axis and two rings of five at latitude +-atan(1/2); the 10 diamonds'
// corners come from the region topology (diamond_corner_classes).  Each
// diamond's lattice is refined glevel times by normalised great-circle
// midpoints (standard recursive division, no spring smoothing), split along
// the W-E diagonal because (i+1, j+1) is a neighbour in NICAM's stencil.
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"

namespace nicam {
namespace {

struct V3 {
    double x, y, z;
};
V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
V3 unit(V3 a)
{
    const double r = std::sqrt(dot(a, a));
    return {a.x / r, a.y / r, a.z / r};
}
// great-circle midpoint; symmetric in a and b, so both diamonds that share an
// edge compute the same bits
V3 mid(V3 a, V3 b) { return unit(add(a, b)); }

enum { CW = 0, CN = 1, CE = 2, CS = 3 };

// unit-sphere corners W, N, E, S of the 10 diamonds
int diamond_corners(V3 corner[10][4])
{
    int cls[10][4];
    if (diamond_corner_classes(cls) != 12) return fail(NICAM_ERR_STATE, "diamond topology does not close to 12 vertices");
    const double lat = std::atan(0.5), pi = std::acos(-1.0);
    for (int sigma : {1, -1}) {
        std::vector<V3> pos(12);
        std::vector<int> set(12, 0);
        auto put = [&](int c, V3 p) {
            pos[c] = p;
            set[c] = 1;
        };
        put(cls[0][CN], {0, 0, 1});
        put(cls[5][CS], {0, 0, -1});
        for (int d = 0; d < 5; ++d) {
            const double lw = sigma * d * 0.4 * pi, ls = lw + sigma * 0.2 * pi;
            put(cls[d][CW], {std::cos(lat) * std::cos(lw), std::cos(lat) * std::sin(lw), std::sin(lat)});
            put(cls[d][CS], {std::cos(lat) * std::cos(ls), std::cos(lat) * std::sin(ls), -std::sin(lat)});
        }
        bool ok = true;
        for (int c = 0; c < 12; ++c) ok = ok && set[c];
        if (!ok) return fail(NICAM_ERR_STATE, "diamond corners do not cover the icosahedron's vertices");
        const double edge = 1.0 / std::sqrt(5.0);  // cosine of the icosahedron's edge angle
        for (int d = 0; d < 10 && ok; ++d) {
            V3 w = pos[cls[d][CW]], n = pos[cls[d][CN]], e = pos[cls[d][CE]], s = pos[cls[d][CS]];
            const double c[5] = {dot(w, s), dot(s, e), dot(e, n), dot(n, w), dot(w, e)};
            for (double v : c) ok = ok && std::fabs(v - edge) < 1e-12;
            // i runs W -> S, j runs W -> N: right-handed about the outward normal
            ok = ok && dot(cross(sub(s, w), sub(n, w)), w) > 0.0;
        }
        if (!ok) continue;
        for (int d = 0; d < 10; ++d)
            for (int k = 0; k < 4; ++k) corner[d][k] = pos[cls[d][k]];
        return NICAM_OK;
    }
    return fail(NICAM_ERR_STATE, "no icosahedron embedding matches the diamond topology");
}

}  // namespace

}  // namespace nicam

using namespace nicam;

extern "C" int nicam_grid_icosahedral(int glevel, int rlevel, double radius, double *GRD_x, double *GRD_x_pl)
{
    if (rlevel < 0 || glevel < rlevel + 1 || glevel - rlevel > 12 || rlevel > 8)
        return fail(NICAM_ERR_ARG, "need 0 <= rlevel <= 8 and rlevel < glevel <= rlevel + 12");
    if (!GRD_x || !GRD_x_pl) return fail(NICAM_ERR_ARG, "null output");
    if (!(radius > 0.0)) return fail(NICAM_ERR_ARG, "radius must be positive");
    V3 corner[10][4];
    int rc = diamond_corners(corner);
    if (rc != NICAM_OK) return rc;
    const int N = 1 << glevel, nr = 1 << (glevel - rlevel), n1 = 1 << rlevel;
    const int gall_1d = nr + 2, gall = gall_1d * gall_1d, lall = 10 * n1 * n1, gall_pl = 6;
    const size_t nreg = (size_t)gall * lall;
    std::memset(GRD_x, 0, sizeof(double) * nreg * 3);
    std::memset(GRD_x_pl, 0, sizeof(double) * gall_pl * 2 * 3);
    std::vector<V3> X((size_t)(N + 1) * (N + 1));
    auto at = [&](int a, int b) -> V3 & { return X[(size_t)a * (N + 1) + b]; };
    for (int d = 0; d < 10; ++d) {
        at(0, 0) = corner[d][CW];
        at(N, 0) = corner[d][CS];
        at(N, N) = corner[d][CE];
        at(0, N) = corner[d][CN];
        for (int s = N; s > 1; s /= 2) {
            const int h = s / 2;
            for (int a = 0; a <= N; a += h)
                for (int b = 0; b <= N; b += h) {
                    const bool ao = a % s != 0, bo = b % s != 0;
                    if (ao && !bo) at(a, b) = mid(at(a - h, b), at(a + h, b));
                    else if (!ao && bo) at(a, b) = mid(at(a, b - h), at(a, b + h));
                    else if (ao && bo) at(a, b) = mid(at(a - h, b - h), at(a + h, b + h));
                }
        }
        // regions of this diamond (mod_adm.f90:660-700 numbering): l = d*n1^2 + (rj-1)*n1 + (ri-1)
        for (int rj = 1; rj <= n1; ++rj)
            for (int ri = 1; ri <= n1; ++ri) {
                const size_t l = (size_t)d * n1 * n1 + (size_t)(rj - 1) * n1 + (ri - 1);
                for (int j = 2; j <= nr + 1; ++j)
                    for (int i = 2; i <= nr + 1; ++i) {
                        const V3 p = at((ri - 1) * nr + (i - 2), (rj - 1) * nr + (j - 2));
                        const size_t g = (size_t)(j - 1) * gall_1d + (i - 1);
                        GRD_x[(0 * lall + l) * gall + g] = radius * p.x;
                        GRD_x[(1 * (size_t)lall + l) * gall + g] = radius * p.y;
                        GRD_x[(2 * (size_t)lall + l) * gall + g] = radius * p.z;
                    }
            }
    }
    for (int p = 0; p < 2; ++p) GRD_x_pl[(2 * 2 + p) * gall_pl] = p == 0 ? radius : -radius;  // z of the centres
    // halo and pole rings, as COMM_data_transfer(GRD_x, GRD_x_pl) fills them
    std::vector<HaloPoint> dst, src;
    std::vector<int> sgp;
    HaloPoint pole_src[2];
    rc = build_halo_map(rlevel, gall_1d, gall_pl, dst, src, sgp, pole_src);
    if (rc != NICAM_OK) return rc;
    for (size_t e = 0; e < dst.size(); ++e)
        for (int c = 0; c < 3; ++c) {
            auto ref = [&](const HaloPoint &h) -> double & {
                return h.kind == 0 ? GRD_x[((size_t)c * lall + h.l) * gall + h.g]
                                   : GRD_x_pl[((size_t)c * 2 + h.l) * gall_pl + h.g];
            };
            ref(dst[e]) = ref(src[e]);
        }
    return NICAM_OK;
}
