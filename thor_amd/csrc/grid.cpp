// Icosahedral grid builder. See grid.h for the layout contract.
//
// The arithmetic of every geometric quantity (vertex positions, spring
// dynamics, q-points, areas, control vectors, operators) follows THOR's
// Icogrid (src/ESP/grid.cu) so the arrays agree with the reference grid;
// the control flow (topological neighbour search, level-by-level lattice
// refinement, OpenMP over points) is this project's own.
#include "grid.h"

#include <cstdlib>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <unordered_map>

namespace thor {
namespace {

struct V3 {
    double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, double b) { return {a.x * b, a.y * b, a.z * b}; }
inline V3 operator*(double b, V3 a) { return {b * a.x, b * a.y, b * a.z}; }
inline V3 operator/(V3 a, double b) { return {a.x / b, a.y / b, a.z / b}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double length(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) { return a / length(a); }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline V3 ld(const double *p, int i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
inline void st(double *p, int i, V3 v) {
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}

// Rotation vector from v1 to v2 scaled by the great-circle angle (grid.cu:61-72).
inline V3 normproj(V3 v1, V3 v2) {
    double l1 = dot(v1, v2);
    V3 c = cross(v1, v2);
    double l2 = length(c);
    return c / l2 * std::atan2(l2, l1);
}

// Sum of three triangle-vertex vectors in an order that does not depend on the
// argument order: ascending latitude (z), ties broken by x
// (vector_operations.h sort_add3).
inline V3 sorted_sum3(V3 a, V3 b, V3 c) {
    V3 w1, w2, w3;
    if (a.z == b.z || a.z == c.z || b.z == c.z) {
        V3 odd, s1, s2;
        if (a.z == b.z) {
            odd = c; s1 = a; s2 = b;
        } else if (a.z == c.z) {
            odd = b; s1 = a; s2 = c;
        } else {
            odd = a; s1 = b; s2 = c;
        }
        V3 lo = (s1.x < s2.x) ? s1 : s2;
        V3 hi = (s1.x < s2.x) ? s2 : s1;
        if (odd.z < s1.z) {
            w1 = odd; w2 = lo; w3 = hi;
        } else {
            w1 = lo; w2 = hi; w3 = odd;
        }
    } else {
        V3 v[3] = {a, b, c};
        // unique z: ascending order
        if (v[0].z > v[1].z) std::swap(v[0], v[1]);
        if (v[1].z > v[2].z) std::swap(v[1], v[2]);
        if (v[0].z > v[1].z) std::swap(v[0], v[1]);
        w1 = v[0]; w2 = v[1]; w3 = v[2];
    }
    return w1 + w2 + w3;
}

// Interior angle at a of the spherical triangle (a, b, c) (grid.cu:2176-2198).
inline double corner_angle(V3 a, V3 b, V3 c) {
    V3 v0 = a, v1 = b - a, v2 = c - a;
    double t = 1.0 / dot(v0, v0);
    double f1 = dot(v0, v1) * t;
    double f2 = dot(v0, v2) * t;
    V3 w1 = v1 - f1 * v0;
    V3 w2 = v2 - f2 * v0;
    double ang = dot(w1, w2) / (length(w1) * length(w2));
    if (ang > 1) ang = 1;
    if (ang < -1) ang = -1;
    return std::acos(ang);
}

inline double sph_tri_area(V3 a, V3 b, V3 c) {
    double s = corner_angle(a, b, c) + corner_angle(b, a, c) + corner_angle(c, a, b);
    double r = length(a);
    return (s - M_PI) * std::pow(r, 2.0);
}

// Arc-length-scaled normal of the great circle through v1, v2 (grid.cu:2488-2500).
inline V3 arc_normal(V3 v1, V3 v2) {
    double vl = length(v1);
    double l = std::acos(dot(v1, v2) / (std::pow(vl, 2.0))) * vl;
    V3 nv = cross(v1, v2);
    return nv * (l / length(nv));
}

struct BitsKey {
    uint64_t a, b, c;
    bool operator==(const BitsKey &o) const { return a == o.a && b == o.b && c == o.c; }
};
struct BitsHash {
    size_t operator()(const BitsKey &k) const {
        uint64_t h = k.a * 0x9E3779B97F4A7C15ull;
        h ^= k.b + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
        h ^= k.c + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
        return (size_t)h;
    }
};
inline BitsKey key_of(V3 v) {
    BitsKey k;
    std::memcpy(&k.a, &v.x, 8);
    std::memcpy(&k.b, &v.y, 8);
    std::memcpy(&k.c, &v.z, 8);
    return k;
}

// Corner vertices of the 10 main rhombi: (a=0,b=0), (a=S,b=0), (a=0,b=S), (a=S,b=S)
// in lattice coordinates (a = local y direction, b = local x direction).
// Same table as grid.cu:404-413 (the icosahedron's rhombus decomposition).
const int kRhombusCorners[10][4] = {{2, 0, 7, 3},  {3, 0, 8, 4},  {4, 0, 9, 5},  {5, 0, 10, 6},
                                    {6, 0, 11, 2}, {7, 3, 1, 8},  {8, 4, 1, 9},  {9, 5, 1, 10},
                                    {10, 6, 1, 11}, {11, 2, 1, 7}};

void icosahedron(V3 v[12]) {
    double w = 2.0 * std::acos(1.0 / (2.0 * std::sin(M_PI / 5.0)));
    v[0] = {0.0, 0.0, 1.0};
    v[1] = {0.0, 0.0, -1.0};
    // Longitudes of the upper (2..6) and lower (7..11) rings, written in the same
    // floating-point form as grid.cu:349-402 so the vertices agree to the bit.
    const double lon_up[5] = {-M_PI / 5.0, M_PI / 5.0, 3.0 * M_PI / 5.0, M_PI, -(3.0 / 5.0) * M_PI};
    const double lon_lo[5] = {0.0, 2.0 * M_PI / 5.0, 4.0 * M_PI / 5.0, -4.0 * M_PI / 5.0,
                              -2.0 * M_PI / 5.0};
    double cu = std::cos(M_PI / 2.0 - w), zu = std::sin(M_PI / 2.0 - w);
    double cl = std::cos(w - M_PI / 2.0), zl = std::sin(w - M_PI / 2.0);
    for (int k = 0; k < 5; k++) {
        v[2 + k] = {std::cos(lon_up[k]) * cu, std::sin(lon_up[k]) * cu, zu};
        v[7 + k] = {std::cos(lon_lo[k]) * cl, std::sin(lon_lo[k]) * cl, zl};
    }
}

} // namespace

bool IcoGrid::is_pent(int i) const {
    for (int k = 0; k < 12; k++)
        if (pent_ind[k] == i) return true;
    return false;
}

static void spring_dynamics(IcoGrid &g) {
    // Tomita et al. (2001) spring dynamics with fixed pentagons (grid.cu:1670-1815).
    const int P = g.point_num;
    const double lba = 2.0 * M_PI / (10.0 * std::pow(2.0, g.glevel - 1));
    const double dbar = g.beta * lba;
    const double drag = 1;
    const double tstep = 2.0E-2;
    const double cond = 1.0E-4;
    const int lim = 100000;

    std::vector<V3> xyz(P), vel(P, V3{0, 0, 0}), F(P);
    std::vector<double> Fnet(P);
    for (int i = 0; i < P; i++) xyz[i] = ld(g.xyz.data(), i);
    V3 fixed[12];
    std::vector<char> pent(P, 0);
    for (int k = 0; k < 12; k++) {
        fixed[k] = xyz[g.pent_ind[k]];
        pent[g.pent_ind[k]] = 1;
    }
    std::vector<V3> Pold(P);
    int it = 0;
    for (it = 0; it < lim; it++) {
        Pold = xyz;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < P; i++) {
            V3 Fsum = {0, 0, 0};
            V3 Fn[6];
            int nn = pent[i] ? 5 : 6;
            for (int j = 0; j < 6; j++) {
                if (j < nn) {
                    V3 pj = Pold[g.point_local[i * 6 + j]];
                    V3 W = cross(Pold[i], pj);
                    V3 V = cross(W, Pold[i]);
                    double l = length(V);
                    double d = std::acos(dot(Pold[i], pj));
                    Fn[j] = ((d - dbar) * V) / l;
                } else {
                    Fn[j] = {0.0, 0.0, 0.0};
                }
            }
            Fsum = Fn[0] + Fn[1] + Fn[2] + Fn[3] + Fn[4] + Fn[5];
            Fnet[i] = length(Fsum) / lba;
            F[i] = Fsum - drag * vel[i];
        }
#pragma omp parallel for schedule(static)
        for (int i = 0; i < P; i++) {
            V3 D = xyz[i] + vel[i] * tstep;
            xyz[i] = normalize(D);
            V3 vv = vel[i] + F[i] * tstep;
            double H = dot(xyz[i], vv);
            vel[i] = vv - H * xyz[i];
        }
        for (int k = 0; k < 12; k++) {
            vel[g.pent_ind[k]] = {0.0, 0.0, 0.0};
            Fnet[g.pent_ind[k]] = 0.0;
            xyz[g.pent_ind[k]] = fixed[k];
        }
        double max_v = -1E-10;
        for (int i = 0; i < P; i++)
            if (Fnet[i] > max_v) max_v = Fnet[i];
        if (max_v < cond) break;
    }
    g.spring_iterations = it;
    for (int i = 0; i < P; i++) st(g.xyz.data(), i, xyz[i]);
}

static void find_qpoints(IcoGrid &g) {
    const int P = g.point_num;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P; i++) {
        int geo = g.is_pent(i) ? 5 : 6;
        const int *nb = &g.point_local[i * 6];
        V3 c = ld(g.xyz.data(), i);
        for (int j = 0; j < geo; j++) {
            int a = nb[j], b = nb[(j + 1) % geo];
            V3 pa = ld(g.xyz.data(), a), pb = ld(g.xyz.data(), b);
            V3 v1 = normproj(pa, c);
            V3 v2 = normproj(pb, pa);
            V3 v3 = normproj(c, pb);
            st(g.xyzq.data(), i * 6 + j, normalize(sorted_sum3(v1, v2, v3)));
        }
        if (geo == 5) st(g.xyzq.data(), i * 6 + 5, V3{0.0, 0.0, 0.0});
    }
}

static void relocate_centres(IcoGrid &g) {
    const int P = g.point_num;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P; i++) {
        int geo = g.is_pent(i) ? 5 : 6;
        V3 s = {0.0, 0.0, 0.0};
        for (int j = 1; j < geo; j++)
            s = s + normproj(ld(g.xyzq.data(), i * 6 + j), ld(g.xyzq.data(), i * 6 + j - 1));
        s = s + normproj(ld(g.xyzq.data(), i * 6 + 0), ld(g.xyzq.data(), i * 6 + geo - 1));
        st(g.xyz.data(), i, normalize(s));
    }
}

void build_horizontal(IcoGrid &g, int glevel, bool spring, double beta) {
    if (glevel < 4 || glevel > 8) throw std::runtime_error("glevel must be in 4..8");
    const int S = 1 << glevel;
    const int divide_face = glevel - 4;
    g.glevel = glevel;
    g.nl_region = 16;
    g.nfaces = 1 << (2 * divide_face);
    g.kxl = 1 << divide_face;
    g.point_num = 2 + 10 * S * S;
    g.nr = (g.point_num - 2) / 256;
    g.spring = spring;
    g.beta = beta;
    const int P = g.point_num, nl = g.nl_region, nl2 = nl * nl, kxl = g.kxl, nfaces = g.nfaces;
    const int L = S + 1;

    // 1. Lattice of each main rhombus by recursive great-circle bisection
    //    (grid.cu:415-632): a new point is the normalised midpoint of its two
    //    coarse neighbours along the lattice direction in which it is odd.
    V3 ico[12];
    icosahedron(ico);
    std::vector<V3> lat((size_t)10 * L * L);
    auto LA = [&](int fc, int a, int b) -> V3 & { return lat[((size_t)fc * L + a) * L + b]; };
    for (int fc = 0; fc < 10; fc++) {
        LA(fc, 0, 0) = ico[kRhombusCorners[fc][0]];
        LA(fc, S, 0) = ico[kRhombusCorners[fc][1]];
        LA(fc, 0, S) = ico[kRhombusCorners[fc][2]];
        LA(fc, S, S) = ico[kRhombusCorners[fc][3]];
        for (int grd = 0; grd < glevel; grd++) {
            int j = 1 << (glevel - grd - 1);
            for (int a = 0; a <= S; a += j)
                for (int b = 0; b <= S; b += j) {
                    int ao = (a / j) & 1, bo = (b / j) & 1;
                    if (!ao && !bo) continue;
                    V3 p1, p2;
                    if (ao && bo) {
                        p1 = LA(fc, a - j, b - j);
                        p2 = LA(fc, a + j, b + j);
                    } else if (ao) {
                        p1 = LA(fc, a - j, b);
                        p2 = LA(fc, a + j, b);
                    } else {
                        p1 = LA(fc, a, b - j);
                        p2 = LA(fc, a, b + j);
                    }
                    LA(fc, a, b) = normalize((p1 + p2) * 0.5);
                }
        }
    }

    // 2. Global point order (grid.cu:618-656).
    g.xyz.assign((size_t)3 * P, 0.0);
    std::vector<int> gid((size_t)10 * L * L, -1);
    std::unordered_map<BitsKey, int, BitsHash> where;
    where.reserve((size_t)P * 2);
    for (int fc = 0; fc < 10; fc++)
        for (int a = 0; a < S; a++)
            for (int b = 0; b < S; b++) {
                int ky = a / nl, j = a % nl, kx = b / nl, i = b % nl;
                int idx = fc * nfaces * nl2 + ky * kxl * nl2 + kx * nl2 + j * nl + i;
                gid[((size_t)fc * L + a) * L + b] = idx;
                st(g.xyz.data(), idx, LA(fc, a, b));
                where[key_of(LA(fc, a, b))] = idx;
            }
    st(g.xyz.data(), P - 2, V3{0.0, 0.0, 1.0});
    st(g.xyz.data(), P - 1, V3{0.0, 0.0, -1.0});
    where[key_of(V3{0.0, 0.0, 1.0})] = P - 2;
    where[key_of(V3{0.0, 0.0, -1.0})] = P - 1;
    for (int fc = 0; fc < 10; fc++)
        for (int a = 0; a <= S; a++)
            for (int b = 0; b <= S; b++) {
                if (a < S && b < S) continue;
                auto it = where.find(key_of(LA(fc, a, b)));
                if (it == where.end())
                    throw std::runtime_error("grid: rhombus edge point has no owner");
                gid[((size_t)fc * L + a) * L + b] = it->second;
            }
    auto G = [&](int fc, int a, int b) { return gid[((size_t)fc * L + a) * L + b]; };
    for (int k = 0; k < 10; k++) g.pent_ind[k] = k * nfaces * nl2;
    g.pent_ind[10] = P - 2;
    g.pent_ind[11] = P - 1;

    // 3. Undirected adjacency from the triangulated lattices.
    std::vector<int> adj((size_t)P * 6, -1);
    std::vector<int> nadj(P, 0);
    auto add_edge1 = [&](int u, int v) {
        for (int k = 0; k < nadj[u]; k++)
            if (adj[(size_t)u * 6 + k] == v) return;
        if (nadj[u] >= 6) throw std::runtime_error("grid: more than 6 neighbours");
        adj[(size_t)u * 6 + nadj[u]++] = v;
    };
    auto add_edge = [&](int u, int v) {
        add_edge1(u, v);
        add_edge1(v, u);
    };
    for (int fc = 0; fc < 10; fc++)
        for (int a = 0; a <= S; a++)
            for (int b = 0; b <= S; b++) {
                if (a + 1 <= S) add_edge(G(fc, a, b), G(fc, a + 1, b));
                if (b + 1 <= S) add_edge(G(fc, a, b), G(fc, a, b + 1));
                if (a + 1 <= S && b + 1 <= S) add_edge(G(fc, a, b), G(fc, a + 1, b + 1));
            }
    auto adjacent = [&](int u, int v) {
        for (int k = 0; k < nadj[u]; k++)
            if (adj[(size_t)u * 6 + k] == v) return true;
        return false;
    };

    // 4. Neighbour rings in rhombus order pt1..pt6 = (a+1,b), (a+1,b+1), (a,b+1),
    //    (a-1,b), (a-1,b-1), (a,b-1) (grid.cu:1495-1621); pentagons list five.
    g.point_local.assign((size_t)6 * P, -1);
    const int da[6] = {1, 1, 0, -1, -1, 0};
    const int db[6] = {0, 1, 1, 0, -1, -1};
    for (int fc = 0; fc < 10; fc++)
        for (int a = 0; a < S; a++)
            for (int b = 0; b < S; b++) {
                int c = G(fc, a, b);
                int n = nadj[c];
                int ring[6] = {-1, -1, -1, -1, -1, -1};
                for (int k = 0; k < 3; k++) ring[k] = G(fc, a + da[k], b + db[k]);
                for (int k = 3; k < n; k++) {
                    int found = -1;
                    for (int q = 0; q < n; q++) {
                        int cand = adj[(size_t)c * 6 + q];
                        bool used = false;
                        for (int r = 0; r < k; r++) used |= (ring[r] == cand);
                        if (!used && adjacent(cand, ring[k - 1])) {
                            if (found >= 0) throw std::runtime_error("grid: ambiguous ring");
                            found = cand;
                        }
                    }
                    if (found < 0) throw std::runtime_error("grid: broken ring");
                    ring[k] = found;
                    int aa = a + da[k], bb = b + db[k];
                    if (n == 6 && aa >= 0 && bb >= 0 && G(fc, aa, bb) != found)
                        throw std::runtime_error("grid: lattice/ring mismatch");
                }
                if (!adjacent(ring[n - 1], ring[0])) throw std::runtime_error("grid: open ring");
                for (int k = 0; k < 6; k++) g.point_local[(size_t)c * 6 + k] = ring[k];
            }
    // Poles: start at the nearest neighbour (ties to the highest index, as the
    // distance scan of grid.cu:744-772 does), then order by the signed angle of
    // grid.cu:919-1005.
    for (int pole = P - 2; pole < P; pole++) {
        if (nadj[pole] != 5) throw std::runtime_error("grid: pole is not a pentagon");
        int cand[5];
        for (int k = 0; k < 5; k++) cand[k] = adj[(size_t)pole * 6 + k];
        std::sort(cand, cand + 5);
        V3 pc = ld(g.xyz.data(), pole);
        int first = -1;
        double small = 2;
        for (int k = 0; k < 5; k++) {
            V3 q = ld(g.xyz.data(), cand[k]);
            double d = std::sqrt(std::pow(q.x - pc.x, 2) + std::pow(q.y - pc.y, 2)
                                 + std::pow(q.z - pc.z, 2));
            if (small >= d && d > 0) {
                small = d;
                first = cand[k];
            }
        }
        int loc[5];
        loc[0] = first;
        int m = 1;
        for (int k = 0; k < 5; k++)
            if (cand[k] != first) loc[m++] = cand[k];
        double x1 = 0, y1 = 0, z1 = 0;
        for (int j = 0; j < 5; j++) {
            x1 = x1 + g.xyz[loc[j] * 3 + 0] / 5;
            y1 = y1 + g.xyz[loc[j] * 3 + 1] / 5;
            z1 = z1 + g.xyz[loc[j] * 3 + 2] / 5;
        }
        double ang[5] = {0, 0, 0, 0, 0};
        double dx21 = g.xyz[loc[0] * 3 + 0] - x1, dy21 = g.xyz[loc[0] * 3 + 1] - y1,
               dz21 = g.xyz[loc[0] * 3 + 2] - z1;
        for (int j = 1; j < 5; j++) {
            double dx31 = g.xyz[loc[j] * 3 + 0] - x1, dy31 = g.xyz[loc[j] * 3 + 1] - y1,
                   dz31 = g.xyz[loc[j] * 3 + 2] - z1;
            double n21 = std::sqrt(dx21 * dx21 + dy21 * dy21 + dz21 * dz21);
            double n31 = std::sqrt(dx31 * dx31 + dy31 * dy31 + dz31 * dz31);
            double cx = dy21 * dz31 - dy31 * dz21, cy = dz21 * dx31 - dz31 * dx21,
                   cz = dx21 * dy31 - dx31 * dy21;
            double dn = cx * x1 + cy * y1 + cz * z1;
            double ac = std::acos((dx21 * dx31 + dy21 * dy31 + dz21 * dz31) / (n21 * n31));
            ang[j] = (dn <= 0) ? ac : 2.0 * M_PI - ac;
        }
        for (int j = 0; j < 5; j++)
            for (int k = 4; k >= j; k--)
                if (ang[j] > ang[k]) {
                    std::swap(ang[j], ang[k]);
                    std::swap(loc[j], loc[k]);
                }
        for (int k = 0; k < 5; k++) g.point_local[(size_t)pole * 6 + k] = loc[k];
        g.point_local[(size_t)pole * 6 + 5] = -1;
    }

    // 5. Rhombus maps with a one-point halo (grid.cu:1624-1668).
    const int nhl = nl + 2, nhl2 = nhl * nhl;
    g.maps.assign((size_t)g.nr * nhl2, -1);
    for (int r = 0; r < g.nr; r++) {
        int *m = &g.maps[(size_t)r * nhl2];
        auto pid = [&](int x, int y) { return r * nl2 + y * nl + x; };
        auto ringof = [&](int p, int k) { return g.point_local[(size_t)p * 6 + k]; };
        for (int y = 0; y < nl; y++)
            for (int x = 0; x < nl; x++) m[(y + 1) * nhl + x + 1] = pid(x, y);
        bool pent_corner = g.is_pent(pid(0, 0));
        for (int y = 0; y < nl; y++) {
            m[(y + 1) * nhl + 0] = ringof(pid(0, y), (pent_corner && y == 0) ? 4 : 5);
            m[(y + 1) * nhl + nl + 1] = ringof(pid(nl - 1, y), 2);
        }
        for (int x = 0; x < nl; x++) {
            m[0 * nhl + x + 1] = ringof(pid(x, 0), 3);
            m[(nl + 1) * nhl + x + 1] = ringof(pid(x, nl - 1), 0);
        }
        m[0] = pent_corner ? -1 : ringof(pid(0, 0), 4);
        m[(nl + 1) * nhl + nl + 1] = ringof(pid(nl - 1, nl - 1), 1);
        // (x=nl, y=-1) and (x=-1, y=nl) are not read by any stencil.
        m[0 * nhl + nl + 1] = ringof(ringof(pid(nl - 1, 0), 2), 3);
        m[(nl + 1) * nhl + 0] = ringof(ringof(pid(0, nl - 1), 5), 0);
    }

    // 6. Spring dynamics, q-points and centroid relocation (grid.cu:208-221).
    g.xyzq.assign((size_t)18 * P, 0.0);
    g.spring_iterations = 0;
    if (spring) {
        // on the GPU when there is one (THOR_GRID_HOST=1: the host loop)
        const char *host = std::getenv("THOR_GRID_HOST");
        std::string err;
        if ((host && std::atoi(host)) || !spring_dynamics_device(g, &err)) spring_dynamics(g);
        find_qpoints(g);
        relocate_centres(g);
    } else {
        find_qpoints(g);
    }
}

void finish_grid(IcoGrid &g, double A, double top, int nv) {
    const int P = g.point_num;
    g.A = A;
    g.top = top;
    g.nv = nv;
    // Scale to the planet radius (grid.cu:2130-2174).
    for (int i = 0; i < P; i++) {
        int geo = g.is_pent(i) ? 5 : 6;
        for (int j = 0; j < geo; j++) st(g.xyzq.data(), i * 6 + j, ld(g.xyzq.data(), i * 6 + j) * A);
        st(g.xyz.data(), i, ld(g.xyz.data(), i) * A);
    }
    g.func_r.assign((size_t)3 * P, 0.0);
    for (int i = 0; i < P; i++) st(g.func_r.data(), i, normalize(ld(g.xyz.data(), i)));

    // Control areas: sub-triangle areas alpha/beta/gamma, triangle areas and
    // control-volume areas, each normalised to the exact sphere area (grid.cu:2200-2421).
    g.areas.assign((size_t)18 * P, 0.0);
    g.areasTr.assign((size_t)6 * P, 0.0);
    g.areasT.assign((size_t)P, 0.0);
    const double area_exact = 4 * M_PI * std::pow(A, 2.0);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P; i++) {
        int geo = g.is_pent(i) ? 5 : 6;
        const int *nb = &g.point_local[(size_t)i * 6];
        V3 c = ld(g.xyz.data(), i);
        for (int j = 0; j < geo; j++) {
            V3 pj = ld(g.xyz.data(), nb[j]);
            V3 pj1 = ld(g.xyz.data(), nb[(j + 1) % geo]);
            V3 q = ld(g.xyzq.data(), i * 6 + j);
            double s0 = sph_tri_area(pj1, pj, q);
            double s1 = sph_tri_area(pj1, c, q);
            double s2 = sph_tri_area(c, pj, q);
            g.areas[(size_t)i * 18 + j * 3 + 0] = s0;
            g.areas[(size_t)i * 18 + j * 3 + 1] = s1;
            g.areas[(size_t)i * 18 + j * 3 + 2] = s2;
            g.areasTr[(size_t)i * 6 + j] = s0 + s1 + s2;
        }
        double at = 0.0;
        for (int j = 0; j < geo; j++) {
            V3 qa = ld(g.xyzq.data(), i * 6 + j);
            V3 qb = ld(g.xyzq.data(), i * 6 + (j + 1) % geo);
            at += sph_tri_area(c, qa, qb);
        }
        g.areasT[i] = at;
    }
    double tot_T = 0.0, tot_Tr = 0.0, tot_sub = 0.0;
    for (int i = 0; i < P; i++) {
        if (i < P - 2)
            for (int j = 0; j < 2; j++) {
                tot_Tr += g.areasTr[(size_t)i * 6 + j];
                tot_sub += (g.areas[(size_t)i * 18 + j * 3 + 0] + g.areas[(size_t)i * 18 + j * 3 + 1]
                            + g.areas[(size_t)i * 18 + j * 3 + 2]);
            }
    }
    for (int i = 0; i < P; i++) tot_T += g.areasT[i];
    for (int i = 0; i < P; i++) {
        g.areasT[i] *= area_exact / tot_T;
        for (int j = 0; j < 6; j++) {
            g.areasTr[(size_t)i * 6 + j] *= area_exact / tot_Tr;
            for (int k = 0; k < 3; k++) g.areas[(size_t)i * 18 + j * 3 + k] *= area_exact / tot_sub;
        }
    }

    // Control-volume edge normals and triangle edge normals (grid.cu:2423-2631).
    g.nvec.assign((size_t)18 * P, 0.0);
    g.nvecoa.assign((size_t)18 * P, 0.0);
    g.nvecti.assign((size_t)18 * P, 0.0);
    g.nvecte.assign((size_t)18 * P, 0.0);
    g.mvec.assign((size_t)18 * P, 0.0);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P; i++) {
        int geo = g.is_pent(i) ? 5 : 6;
        const int *nb = &g.point_local[(size_t)i * 6];
        for (int j = 0; j < geo; j++) {
            V3 v1 = ld(g.xyzq.data(), i * 6 + j);
            V3 v2 = ld(g.xyzq.data(), i * 6 + (j + 1) % geo);
            double vl = length(v1);
            double l = std::acos(dot(v1, v2) / (std::pow(vl, 2.0))) * vl;
            V3 nv = cross(v1, v2);
            V3 nvj = nv * (l / length(nv));
            st(g.nvec.data(), i * 6 + j, nvj);
            st(g.nvecoa.data(), i * 6 + j, nvj / g.areasT[i]);
            V3 mv = v1 - v2;
            st(g.mvec.data(), i * 6 + j, mv * (l / length(mv)));
        }
        for (int j = 0; j < geo; j++)
            st(g.nvecti.data(), i * 6 + j, arc_normal(ld(g.xyz.data(), nb[j]), ld(g.xyz.data(), i)));
        for (int j = 0; j < geo; j++)
            st(g.nvecte.data(), i * 6 + j,
               arc_normal(ld(g.xyz.data(), nb[j]), ld(g.xyz.data(), nb[(j + 1) % geo])));
    }

    // Altitudes: uniform layers (grid.cu:1923-1948).
    g.Altitude.assign(nv, 0.0);
    g.Altitudeh.assign(nv + 1, 0.0);
    double res_vert = top / nv;
    g.Altitudeh[0] = 0.0;
    for (int lev = 0; lev < nv; lev++) g.Altitudeh[lev + 1] = g.Altitudeh[lev] + res_vert;
    for (int lev = 0; lev < nv; lev++) g.Altitude[lev] = (g.Altitudeh[lev] + g.Altitudeh[lev + 1]) / 2.0;

    // Longitude / latitude (grid.cu:2111-2128).
    g.lonlat.assign((size_t)2 * P, 0.0);
    for (int i = 0; i < P; i++) {
        double x = g.xyz[i * 3 + 0], y = g.xyz[i * 3 + 1], z = g.xyz[i * 3 + 2];
        g.lonlat[i * 2 + 0] = std::atan2(y, x);
        g.lonlat[i * 2 + 1] = std::atan2(z, std::sqrt(std::pow(x, 2) + std::pow(y, 2)));
    }

    // Divergence and gradient operators, Satoh et al. (2008) eqs. 15-17
    // (grid.cu:2656-2831, 3015-3201). Coefficient 0 is the centre, m+1 the
    // ring neighbour m; a pentagon's seventh coefficient is zero.
    g.div.assign((size_t)21 * P, 0.0);
    g.grad.assign((size_t)21 * P, 0.0);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P; i++) {
        int n = g.is_pent(i) ? 5 : 6;
        double aq[6][3] = {{0}};
        for (int j = 0; j < n; j++) {
            const double *ar = &g.areas[(size_t)i * 18 + j * 3];
            double s = ar[0] + ar[1] + ar[2];
            for (int k = 0; k < 3; k++) aq[j][k] = ar[k] / s;
        }
        const double *nvb = &g.nvec[(size_t)i * 18];
        for (int k = 0; k < 3; k++) {
            double c0 = 0.0, nsum = 0.0;
            for (int j = 0; j < n; j++) {
                double t = nvb[j * 3 + k] * (aq[j][0] + aq[(j + 1) % n][0]);
                c0 = (j == 0) ? t : c0 + t;
                nsum = (j == 0) ? nvb[j * 3 + k] : nsum + nvb[j * 3 + k];
            }
            double *dv = &g.div[(size_t)i * 21];
            double *gr = &g.grad[(size_t)i * 21];
            dv[0 * 3 + k] = c0;
            gr[0 * 3 + k] = c0 - 2 * nsum;
            for (int m = 0; m < n; m++) {
                int mm1 = (m + n - 1) % n, mm2 = (m + n - 2) % n;
                double v = nvb[mm1 * 3 + k] * (aq[m][1] + aq[mm1][2]) + nvb[mm2 * 3 + k] * aq[mm1][2]
                           + nvb[m * 3 + k] * aq[m][1];
                dv[(m + 1) * 3 + k] = v;
                gr[(m + 1) * 3 + k] = v;
            }
            if (n == 5) {
                dv[6 * 3 + k] = 0.0;
                gr[6 * 3 + k] = 0.0;
            }
        }
        for (int j = 0; j < 21; j++) {
            g.div[(size_t)i * 21 + j] = g.div[(size_t)i * 21 + j] / (2.0 * g.areasT[i]);
            g.grad[(size_t)i * 21 + j] = g.grad[(size_t)i * 21 + j] / (2.0 * g.areasT[i]);
        }
    }
}

static const char kMagic[8] = {'T', 'H', 'O', 'R', 'G', 'R', 'D', '1'};

bool save_horizontal(const IcoGrid &g, const std::string &path) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    int hdr[4] = {g.glevel, g.spring ? 1 : 0, g.point_num, g.spring_iterations};
    bool ok = std::fwrite(kMagic, 1, 8, f) == 8 && std::fwrite(hdr, sizeof(int), 4, f) == 4
              && std::fwrite(&g.beta, sizeof(double), 1, f) == 1
              && std::fwrite(g.pent_ind, sizeof(int), 12, f) == 12
              && std::fwrite(g.xyz.data(), sizeof(double), g.xyz.size(), f) == g.xyz.size()
              && std::fwrite(g.xyzq.data(), sizeof(double), g.xyzq.size(), f) == g.xyzq.size()
              && std::fwrite(g.point_local.data(), sizeof(int), g.point_local.size(), f)
                     == g.point_local.size()
              && std::fwrite(g.maps.data(), sizeof(int), g.maps.size(), f) == g.maps.size();
    std::fclose(f);
    return ok;
}

bool load_horizontal(IcoGrid &g, const std::string &path, int glevel, bool spring, double beta) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    char magic[8];
    int hdr[4];
    double b;
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kMagic, 8) == 0
              && std::fread(hdr, sizeof(int), 4, f) == 4 && std::fread(&b, sizeof(double), 1, f) == 1;
    if (!ok || hdr[0] != glevel || hdr[1] != (spring ? 1 : 0) || b != beta) {
        std::fclose(f);
        return false;
    }
    const int S = 1 << glevel;
    g.glevel = glevel;
    g.nl_region = 16;
    g.nfaces = 1 << (2 * (glevel - 4));
    g.kxl = 1 << (glevel - 4);
    g.point_num = hdr[2];
    g.nr = (g.point_num - 2) / 256;
    g.spring = spring;
    g.beta = beta;
    g.spring_iterations = hdr[3];
    if (g.point_num != 2 + 10 * S * S) {
        std::fclose(f);
        return false;
    }
    g.xyz.resize((size_t)3 * g.point_num);
    g.xyzq.resize((size_t)18 * g.point_num);
    g.point_local.resize((size_t)6 * g.point_num);
    g.maps.resize((size_t)g.nr * 18 * 18);
    ok = std::fread(g.pent_ind, sizeof(int), 12, f) == 12
         && std::fread(g.xyz.data(), sizeof(double), g.xyz.size(), f) == g.xyz.size()
         && std::fread(g.xyzq.data(), sizeof(double), g.xyzq.size(), f) == g.xyzq.size()
         && std::fread(g.point_local.data(), sizeof(int), g.point_local.size(), f)
                == g.point_local.size()
         && std::fread(g.maps.data(), sizeof(int), g.maps.size(), f) == g.maps.size();
    std::fclose(f);
    return ok;
}

} // namespace thor
