// xdem_host.cpp -- host-side set-up of the DEM step: wall geometry, mass and inertia,
// contact properties, and seeded scene generators. Plain C++ (no HIP); the GPU library
// links it, and the CPU tests call it directly.
#include "xdem_host.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

struct V3 { double x, y, z; };
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double length(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V3 ld(const double* p) { return {p[0], p[1], p[2]}; }
inline xdem_double3 to_x(V3 a) { return {a.x, a.y, a.z}; }

// Rotation matrix of Euler parameters (xPointMass::setupTransformationMatrix, xPointMass.cpp:296-301).
struct Mat3 { double a00, a01, a02, a10, a11, a12, a20, a21, a22; };
Mat3 rot(const double e[4])
{
    Mat3 A;
    A.a00 = 2 * (e[0] * e[0] + e[1] * e[1] - 0.5); A.a01 = 2 * (e[1] * e[2] - e[0] * e[3]); A.a02 = 2 * (e[1] * e[3] + e[0] * e[2]);
    A.a10 = 2 * (e[1] * e[2] + e[0] * e[3]); A.a11 = 2 * (e[0] * e[0] + e[2] * e[2] - 0.5); A.a12 = 2 * (e[2] * e[3] - e[0] * e[1]);
    A.a20 = 2 * (e[1] * e[3] - e[0] * e[2]); A.a21 = 2 * (e[2] * e[3] + e[0] * e[1]); A.a22 = 2 * (e[0] * e[0] + e[3] * e[3] - 0.5);
    return A;
}
V3 mul(const Mat3& A, V3 v) { return {A.a00 * v.x + A.a01 * v.y + A.a02 * v.z, A.a10 * v.x + A.a11 * v.y + A.a12 * v.z, A.a20 * v.x + A.a21 * v.y + A.a22 * v.z}; }
V3 mulT(const Mat3& A, V3 v) { return {A.a00 * v.x + A.a10 * v.y + A.a20 * v.z, A.a01 * v.x + A.a11 * v.y + A.a21 * v.z, A.a02 * v.x + A.a12 * v.y + A.a22 * v.z}; }

// Random sequential insertion of non-overlapping spheres into a region, with a
// uniform cell grid for the overlap test. Radii are drawn per sphere; the position
// is redrawn until the sphere fits (bounded attempts).
struct Rsa {
    V3 lo, hi;          // bounding box of the region
    double cs;          // cell size >= 2 * largest radius
    int nx, ny, nz;
    std::vector<int> head;
    std::vector<int> next;
    std::vector<double> p;  // x, y, z, r
    Rsa(V3 lo_, V3 hi_, double rmax) : lo(lo_), hi(hi_), cs(2.0 * rmax)
    {
        nx = std::max(1, (int)std::ceil((hi.x - lo.x) / cs));
        ny = std::max(1, (int)std::ceil((hi.y - lo.y) / cs));
        nz = std::max(1, (int)std::ceil((hi.z - lo.z) / cs));
        head.assign((size_t)nx * ny * nz, -1);
    }
    void cell(V3 c, int& ix, int& iy, int& iz) const
    {
        ix = std::min(nx - 1, std::max(0, (int)std::floor((c.x - lo.x) / cs)));
        iy = std::min(ny - 1, std::max(0, (int)std::floor((c.y - lo.y) / cs)));
        iz = std::min(nz - 1, std::max(0, (int)std::floor((c.z - lo.z) / cs)));
    }
    bool fits(V3 c, double r) const
    {
        int ix, iy, iz;
        cell(c, ix, iy, iz);
        for (int z = std::max(0, iz - 1); z <= std::min(nz - 1, iz + 1); z++)
            for (int y = std::max(0, iy - 1); y <= std::min(ny - 1, iy + 1); y++)
                for (int x = std::max(0, ix - 1); x <= std::min(nx - 1, ix + 1); x++)
                    for (int k = head[((size_t)z * ny + y) * nx + x]; k >= 0; k = next[k]) {
                        V3 q = {p[4 * k], p[4 * k + 1], p[4 * k + 2]};
                        double rr = r + p[4 * k + 3];
                        V3 d = q - c;
                        if (dot(d, d) < rr * rr * (1.0 + 1e-9)) return false;
                    }
        return true;
    }
    void add(V3 c, double r)
    {
        int ix, iy, iz;
        cell(c, ix, iy, iz);
        int k = (int)(p.size() / 4);
        p.insert(p.end(), {c.x, c.y, c.z, r});
        size_t h = ((size_t)iz * ny + iy) * nx + ix;
        next.push_back(head[h]);
        head[h] = k;
    }
};

// Particles are numbered in cell order (z, then y, then x) of a grid of cell size cs,
// so that ids near each other are near each other in space.
void spatial_order(std::vector<double>& p, double cs, V3 lo)
{
    size_t n = p.size() / 4;
    std::vector<std::pair<unsigned long long, size_t>> key(n);
    for (size_t i = 0; i < n; i++) {
        unsigned long long cx = (unsigned long long)std::max(0.0, std::floor((p[4 * i] - lo.x) / cs));
        unsigned long long cy = (unsigned long long)std::max(0.0, std::floor((p[4 * i + 1] - lo.y) / cs));
        unsigned long long cz = (unsigned long long)std::max(0.0, std::floor((p[4 * i + 2] - lo.z) / cs));
        key[i] = {(cz << 42) | (cy << 21) | cx, i};
    }
    std::sort(key.begin(), key.end());
    std::vector<double> q(p.size());
    for (size_t i = 0; i < n; i++) std::memcpy(&q[4 * i], &p[4 * key[i].second], 4 * sizeof(double));
    p.swap(q);
}

template <class T> T* dup(const std::vector<T>& v)
{
    T* r = (T*)std::malloc(sizeof(T) * (v.size() ? v.size() : 1));
    if (!v.empty()) std::memcpy(r, v.data(), sizeof(T) * v.size());
    return r;
}

// Five walls of an axis-aligned box [0,L]x[0,W]x[0,H] with inward normals: floor, x=0,
// x=L, y=0, y=W (ids in that order). x_offset shifts the box along x.
void box_walls(double L, double W, double H, std::vector<xdem_plane_info>& pl, std::vector<double>& dbi)
{
    const double c[5][4][3] = {
        {{0, 0, 0}, {L, 0, 0}, {L, W, 0}, {0, W, 0}},   // floor, normal +z
        {{0, 0, 0}, {0, W, 0}, {0, W, H}, {0, 0, H}},   // x = 0, normal +x
        {{L, 0, 0}, {L, 0, H}, {L, W, H}, {L, W, 0}},   // x = L, normal -x
        {{0, 0, 0}, {0, 0, H}, {L, 0, H}, {L, 0, 0}},   // y = 0, normal +y
        {{0, W, 0}, {L, W, 0}, {L, W, H}, {0, W, H}},   // y = W, normal -y
    };
    pl.resize(5);
    dbi.resize(15 * 5);
    for (unsigned k = 0; k < 5; k++)
        xdem_make_plane(c[k][0], c[k][1], c[k][2], c[k][3], k, 0.0, &pl[k], &dbi[15 * k]);
}

void fill_common(xdem_scene* s, const std::vector<double>& p, double density)
{
    s->np = (unsigned)(p.size() / 4);
    s->pos = dup(p);
    std::vector<double> zero(3 * (size_t)s->np, 0.0);
    s->vel = dup(zero);
    s->mass = (double*)std::malloc(sizeof(double) * (s->np ? s->np : 1));
    s->inertia = (double*)std::malloc(sizeof(double) * 3 * (s->np ? s->np : 1));
    std::memset(s->mass, 0, sizeof(double) * s->np);
    xdem_mass_inertia(s->pos, s->np, density, s->mass, s->inertia);
    s->density = density;
    double rmin = 1e300, rmax = 0;
    for (unsigned i = 0; i < s->np; i++) {
        rmin = std::min(rmin, p[4 * i + 3]);
        rmax = std::max(rmax, p[4 * i + 3]);
    }
    s->min_radius = rmin;
    s->max_radius = rmax;
    // grid of the reference (xDiscreteElementMethodSimulation.cpp:94-102)
    s->world_origin[0] = s->world_origin[1] = s->world_origin[2] = -1.0;
    s->cell_size = rmax * 2.0;
}

void set_walls(xdem_scene* s, const std::vector<xdem_plane_info>& pl, const std::vector<double>& dbi,
               const xdem_contact_property& cp)
{
    s->nplane = (unsigned)pl.size();
    s->planes = dup(pl);
    s->plane_dbi = dup(dbi);
    std::vector<xdem_contact_property> cps(pl.size(), cp);
    s->plane_cp = dup(cps);
}

double shear_modulus(double E, double nu) { return E / (2.0 * (1.0 + nu)); }

}  // namespace

extern "C" {

unsigned long long xdem_rng_next(unsigned long long* state)
{
    unsigned long long z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double xdem_rng_uniform(unsigned long long* state)
{
    return (double)(xdem_rng_next(state) >> 11) * (1.0 / 9007199254740992.0);
}

int xdem_wheel_mesh(double radius, double width, unsigned int n_theta, unsigned int n_y, unsigned int n_grouser,
                    double grouser_height, double grouser_thickness, unsigned int n_ring, double* vertices,
                    unsigned int* ntriangle)
{
    if (!ntriangle || !(radius > 0.0) || !(width > 0.0) || n_theta < 3 || n_y < 1 || n_ring < 3) return XDEM_E_INVALID;
    const unsigned T = 2 * n_theta * n_y + 4 * n_ring + n_grouser * (6 * n_y + 4);
    if (!vertices) {
        *ntriangle = T;
        return XDEM_OK;
    }
    if (*ntriangle < T) return XDEM_E_INVALID;
    const double pi = 3.14159265358979323846;
    unsigned k = 0;
    auto tri = [&](V3 a, V3 b, V3 c) {
        const V3 v[3] = {a, b, c};
        for (int q = 0; q < 3; q++) {
            vertices[9 * k + 3 * q + 0] = v[q].x;
            vertices[9 * k + 3 * q + 1] = v[q].y;
            vertices[9 * k + 3 * q + 2] = v[q].z;
        }
        k++;
    };
    // a point at angle phi (about y, from +x towards +z), radius rr, axial position y
    auto at = [](double rr, double phi, double y) { return V3{rr * std::cos(phi), y, rr * std::sin(phi)}; };
    const double y0 = -0.5 * width, dy = width / n_y;
    // the rim: n_theta x n_y quads, outward normals
    for (unsigned i = 0; i < n_theta; i++) {
        const double p0 = 2 * pi * i / n_theta, p1 = 2 * pi * (i + 1) / n_theta;
        for (unsigned j = 0; j < n_y; j++) {
            const double ya = y0 + j * dy, yb = y0 + (j + 1) * dy;
            tri(at(radius, p0, ya), at(radius, p0, yb), at(radius, p1, yb));
            tri(at(radius, p0, ya), at(radius, p1, yb), at(radius, p1, ya));
        }
    }
    // side rings: annuli from radius - grouser_height to radius at y = -width/2 and +width/2
    const double ri = radius - grouser_height;
    for (int side = 0; side < 2; side++) {
        const double y = side ? 0.5 * width : -0.5 * width;
        for (unsigned i = 0; i < n_ring; i++) {
            const double p0 = 2 * pi * i / n_ring, p1 = 2 * pi * (i + 1) / n_ring;
            tri(at(ri, p0, y), at(radius, p0, y), at(radius, p1, y));
            tri(at(ri, p0, y), at(radius, p1, y), at(ri, p1, y));
        }
    }
    // grousers: bars of height grouser_height and thickness grouser_thickness on the rim
    for (unsigned g = 0; g < n_grouser; g++) {
        const double pc = 2 * pi * (g + 0.5) / n_grouser, half = 0.5 * grouser_thickness / radius;
        const double pa = pc - half, pb = pc + half, ro = radius + grouser_height;
        for (unsigned j = 0; j < n_y; j++) {
            const double ya = y0 + j * dy, yb = y0 + (j + 1) * dy;
            tri(at(ro, pa, ya), at(ro, pa, yb), at(ro, pb, yb));          // top
            tri(at(ro, pa, ya), at(ro, pb, yb), at(ro, pb, ya));
            tri(at(radius, pa, ya), at(radius, pa, yb), at(ro, pa, yb));  // leading side
            tri(at(radius, pa, ya), at(ro, pa, yb), at(ro, pa, ya));
            tri(at(radius, pb, ya), at(ro, pb, yb), at(radius, pb, yb));  // trailing side
            tri(at(radius, pb, ya), at(ro, pb, ya), at(ro, pb, yb));
        }
        for (int side = 0; side < 2; side++) {                            // end caps
            const double y = side ? 0.5 * width : -0.5 * width;
            tri(at(radius, pa, y), at(ro, pa, y), at(ro, pb, y));
            tri(at(radius, pa, y), at(ro, pb, y), at(radius, pb, y));
        }
    }
    *ntriangle = k;
    return XDEM_OK;
}

int xdem_make_plane(const double p0[3], const double p1[3], const double p2[3], const double p3[3],
                    unsigned int id, double mass, xdem_plane_info* out, double dbi[15])
{
    if (!p0 || !p1 || !p2 || !p3 || !out || !dbi) return XDEM_E_INVALID;
    // xPlaneObject::define(p0, p1, p2, p3), xPlaneObject.cpp:22-49
    V3 xw = ld(p0), w2 = ld(p1), w3 = ld(p2), w4 = ld(p3);
    V3 bpos = 0.5 * (xw + w3);
    const double e[4] = {1.0, 0.0, 0.0, 0.0};
    Mat3 A = rot(e);
    V3 lp[4] = {mulT(A, xw - bpos), mulT(A, w2 - bpos), mulT(A, w3 - bpos), mulT(A, w4 - bpos)};
    V3 pa0 = w2 - xw, pb0 = w4 - xw;
    if (length(pa0) == 0.0 || length(pb0) == 0.0) return XDEM_E_INVALID;
    // xParticlePlaneContact::update, xParticlePlaneContact.cpp:96-110: corners rebuilt from the pose
    V3 gxw = bpos + mul(A, lp[0]);
    V3 gw2 = bpos + mul(A, lp[1]);
    V3 gw3 = bpos + mul(A, lp[2]);
    V3 gw4 = bpos + mul(A, lp[3]);
    std::memset(out, 0, sizeof(*out));
    out->id = id;
    out->xw = to_x(gxw);
    out->w2 = to_x(gw2);
    out->w3 = to_x(gw3);
    out->w4 = to_x(gw4);
    V3 pa = gw2 - gxw, pb = gw4 - gxw;
    out->pa = to_x(pa);
    out->pb = to_x(pb);
    out->l1 = length(pa);
    out->l2 = length(pb);
    V3 u1 = pa / out->l1, u2 = pb / out->l2;
    out->u1 = to_x(u1);
    out->u2 = to_x(u2);
    out->uw = to_x(cross(u1, u2));
    const double b[15] = {mass, bpos.x, bpos.y, bpos.z, 0, 0, 0, e[0], e[1], e[2], e[3], 0, 0, 0, 0};
    std::memcpy(dbi, b, sizeof(b));
    return XDEM_OK;
}

void xdem_mass_inertia(const double* pos, unsigned int n, double density, double* mass, double* inertia)
{
    const double pi = 3.14159265358979323846;
    for (unsigned int i = 0; i < n; i++) {
        double r = pos[4 * i + 3];
        double m = density * (4.0 / 3.0) * pi * std::pow(r, 3.0);
        double J = (2.0 / 5.0) * m * std::pow(r, 2.0);
        mass[i] = mass[i] ? mass[i] : m;
        inertia[3 * i + 0] = J;
        inertia[3 * i + 1] = J;
        inertia[3 * i + 2] = J;
    }
}

double xdem_critical_time_step(double min_rad, double rho, double E, double p)
{
    const double pi = 3.14159265358979323846;
    double dt_raleigh = pi * min_rad * std::sqrt(rho * 2.0 * (1 + p) / E) / (0.1631 * (p + 0.8766));
    double dt_hertz = 2.87 * std::pow(std::pow(rho * (4.0 / 3.0) * pi * std::pow(min_rad, 3.0), 2.0) / (min_rad * E * E * 1.0), 0.2);
    double dt_cundall = 0.2 * pi * std::sqrt(rho * (4.0 / 3.0) * pi * min_rad * min_rad * 3.0 * (1.0 + 2.0 * p) / (E * 0.01));
    double min_dt = dt_raleigh < dt_hertz ? (dt_raleigh < dt_cundall ? dt_raleigh : dt_cundall)
                                          : (dt_hertz < dt_cundall ? dt_hertz : dt_cundall);
    return 0.1 * min_dt;
}

void xdem_contact_property_make(double Ei, double Ej, double nui, double nuj, double rest, double fric,
                                double rfric, double coh, double sratio, xdem_contact_property* out)
{
    std::memset(out, 0, sizeof(*out));
    out->Ei = Ei; out->Ej = Ej;
    out->pri = nui; out->prj = nuj;
    out->Gi = shear_modulus(Ei, nui); out->Gj = shear_modulus(Ej, nuj);
    out->rest = rest;
    out->fric = fric;
    out->s_fric = fric;
    out->rfric = rfric;
    out->coh = coh;
    out->sratio = sratio;
    out->amp = 1.0;
}

void xdem_fit_sphere_to_triangle(const double Pp[3], const double Qp[3], const double Rp[3], double ft, double out[4])
{
    V3 P = ld(Pp), Q = ld(Qp), R = ld(Rp);
    V3 V = Q - P, W = R - P;
    V3 N = cross(V, W);
    N = N / length(N);
    V3 M1 = (Q + P) / 2, M2 = (R + P) / 2;
    V3 D1 = cross(N, V), D2 = cross(N, W);
    double t = 0.0;
    if (std::fabs(D1.x * D2.y - D1.y * D2.x) > 1E-13)
        t = (D2.x * (M1.y - M2.y)) / (D1.x * D2.y - D1.y * D2.x) - (D2.y * (M1.x - M2.x)) / (D1.x * D2.y - D1.y * D2.x);
    else if (std::fabs(D1.x * D2.z - D1.z * D2.x) > 1E-13)
        t = (D2.x * (M1.z - M2.z)) / (D1.x * D2.z - D1.z * D2.x) - (D2.z * (M1.x - M2.x)) / (D1.x * D2.z - D1.z * D2.x);
    else if (std::fabs(D1.y * D2.z - D1.z * D2.y) > 1E-13)
        t = (D2.y * (M1.z - M2.z)) / (D1.y * D2.z - D1.z * D2.y) - (D2.z * (M1.y - M2.y)) / (D1.y * D2.z - D1.z * D2.y);
    V3 Ctri = M1 + t * D1;
    V3 Cs = {0, 0, 0};
    double fc = 0, r = length(P - Ctri);
    for (int it = 0; std::fabs(fc - ft) > 1e-4 && it < 1000; it++) {
        double d = ft * r;
        Cs = Ctri - (d / length(N)) * N;
        r = length(P - Cs);
        fc = d / r;
    }
    out[0] = Cs.x; out[1] = Cs.y; out[2] = Cs.z; out[3] = r;
}

void xdem_scene_free(xdem_scene* s)
{
    if (!s) return;
    std::free(s->pos); std::free(s->vel); std::free(s->mass); std::free(s->inertia);
    std::free(s->planes); std::free(s->plane_dbi); std::free(s->plane_cp);
    std::memset(s, 0, sizeof(*s));
}

int xdem_scene_build(int workload, unsigned long long seed, unsigned int np_override,
                     unsigned int variant, xdem_scene* s)
{
    if (!s) return XDEM_E_INVALID;
    std::memset(s, 0, sizeof(*s));
    unsigned long long rng = seed;
    const double pi = 3.14159265358979323846;
    std::vector<xdem_plane_info> pl;
    std::vector<double> dbi;
    s->contact_model = 1;
    s->gravity[0] = 0.0; s->gravity[1] = 0.0; s->gravity[2] = -9.81;

    if (workload == XDEM_SCENE_BOX_LATTICE) {
        // 20 x 20 x 25 lattice, 2.2 mm pitch from 2.2 mm, jitter +-0.05 mm, r = 1 mm
        unsigned n = np_override ? np_override : 10000u;
        if (n > 20u * 20u * 1000u) return XDEM_E_INVALID;
        std::vector<double> p;
        p.reserve(4 * (size_t)n);
        for (unsigned i = 0; i < n; i++) {
            unsigned ix = i % 20, iy = (i / 20) % 20, iz = i / 400;
            double jx = (xdem_rng_uniform(&rng) * 2.0 - 1.0) * 0.05e-3;
            double jy = (xdem_rng_uniform(&rng) * 2.0 - 1.0) * 0.05e-3;
            double jz = (xdem_rng_uniform(&rng) * 2.0 - 1.0) * 0.05e-3;
            p.insert(p.end(), {2.2e-3 * (ix + 1) + jx, 2.2e-3 * (iy + 1) + jy, 2.2e-3 * (iz + 1) + jz, 1.0e-3});
        }
        fill_common(s, p, 2500.0);
        xdem_contact_property_make(1e7, 1e7, 0.3, 0.3, 0.3, 0.3, 0.01, 0.0, 0.0, &s->cp_pp);
        box_walls(0.1, 0.1, std::max(0.3, 2.2e-3 * (n / 400 + 3)), pl, dbi);
        set_walls(s, pl, dbi, s->cp_pp);
        s->dt = 1e-5;
        s->nsteps = 10000;
        return XDEM_OK;
    }
    if (workload == XDEM_SCENE_DAM_BREAK) {
        unsigned n = np_override ? np_override : 1000000u;
        // the column shrinks with the particle count so the solid fraction stays ~0.29
        double scale = std::cbrt((double)n / 1.0e6);
        V3 lo = {0, 0, 0}, hi = {0.3 * scale, 0.3 * scale, 0.6 * scale};
        Rsa rsa(lo, hi, 2.0e-3);
        // radii drawn first, inserted largest first (random sequential insertion jams
        // well below a solid fraction of 0.29 when small spheres go in first)
        std::vector<double> radii(n);
        for (unsigned i = 0; i < n; i++) radii[i] = 1.0e-3 + 1.0e-3 * xdem_rng_uniform(&rng);
        std::sort(radii.begin(), radii.end(), [](double a, double b) { return a > b; });
        for (unsigned i = 0; i < n; i++) {
            double r = radii[i];
            bool ok = false;
            for (int a = 0; a < 100000 && !ok; a++) {
                V3 c = {r + (hi.x - lo.x - 2 * r) * xdem_rng_uniform(&rng),
                        r + (hi.y - lo.y - 2 * r) * xdem_rng_uniform(&rng),
                        r + (hi.z - lo.z - 2 * r) * xdem_rng_uniform(&rng)};
                if (rsa.fits(c, r)) { rsa.add(c, r); ok = true; }
            }
            if (!ok) return XDEM_E_INVALID;
        }
        spatial_order(rsa.p, 4.0e-3, lo);
        fill_common(s, rsa.p, 2600.0);
        xdem_contact_property_make(5e7, 5e7, 0.25, 0.25, 0.4, 0.5, 0.05, 0.05, 0.0, &s->cp_pp);
        xdem_contact_property wall;
        xdem_contact_property_make(5e7, 5e7, 0.25, 0.25, 0.4, 0.5, 0.05, 0.0, 0.0, &wall);
        box_walls(1.2 * scale, 0.3 * scale, 0.7 * scale, pl, dbi);
        set_walls(s, pl, dbi, wall);
        s->dt = 5e-6;
        s->nsteps = 2000;
        return XDEM_OK;
    }
    if (workload == XDEM_SCENE_WHEEL_BED) {
        // spheres r ~ U[2, 3] mm for a 1.0 x 0.4 x 0.15 m bed (500,000 by default), inserted
        // without overlap into a column of twice the bed's height (solid fraction ~0.3) over
        // the bin floor, to settle under gravity; a smaller count shrinks the bin's footprint
        unsigned n = np_override ? np_override : 500000u;
        const double scale = std::sqrt((double)n / 5.0e5);
        std::vector<double> radii(n);
        double vol = 0.0;
        for (unsigned i = 0; i < n; i++) {
            radii[i] = 2.0e-3 + 1.0e-3 * xdem_rng_uniform(&rng);
            vol += 4.0 / 3.0 * pi * radii[i] * radii[i] * radii[i];
        }
        const double L = 1.0 * scale, W = 0.4 * scale;
        const double H = vol / (0.30 * L * W);
        std::sort(radii.begin(), radii.end(), [](double a, double b) { return a > b; });
        V3 lo = {0, 0, 0}, hi = {L, W, H};
        Rsa rsa(lo, hi, 3.0e-3);
        for (unsigned i = 0; i < n; i++) {
            double r = radii[i];
            bool ok = false;
            for (int a = 0; a < 100000 && !ok; a++) {
                V3 c = {r + (L - 2 * r) * xdem_rng_uniform(&rng), r + (W - 2 * r) * xdem_rng_uniform(&rng),
                        r + (H - 2 * r) * xdem_rng_uniform(&rng)};
                if (rsa.fits(c, r)) { rsa.add(c, r); ok = true; }
            }
            if (!ok) return XDEM_E_INVALID;
        }
        spatial_order(rsa.p, 6.0e-3, lo);
        fill_common(s, rsa.p, 2500.0);
        xdem_contact_property_make(5e7, 5e7, 0.25, 0.25, 0.4, 0.5, 0.05, 0.0, 0.0, &s->cp_pp);
        box_walls(L, W, std::max(0.35, 1.2 * H), pl, dbi);
        set_walls(s, pl, dbi, s->cp_pp);
        s->dt = 1e-5;
        s->nsteps = 5000;
        return XDEM_OK;
    }
    if (workload == XDEM_SCENE_BATCH_MEMBER) {
        unsigned long long vs = seed ^ (0x9E3779B97F4A7C15ull * (variant + 1));
        xdem_rng_next(&vs);
        double lnn = std::log(2000.0) + (std::log(64000.0) - std::log(2000.0)) * xdem_rng_uniform(&vs);
        unsigned n = np_override ? np_override : (unsigned)std::floor(std::exp(lnn));
        double mu = 0.1 + 0.7 * xdem_rng_uniform(&vs);
        double rest = 0.1 + 0.8 * xdem_rng_uniform(&vs);
        double E = std::exp(std::log(1e6) + (std::log(1e8) - std::log(1e6)) * xdem_rng_uniform(&vs));
        std::vector<double> radii(n);
        double vol = 0.0;
        for (unsigned i = 0; i < n; i++) {
            radii[i] = 0.5e-3 + 1.5e-3 * xdem_rng_uniform(&vs);
            vol += 4.0 / 3.0 * pi * radii[i] * radii[i] * radii[i];
        }
        // box volume = 3 x packed volume (solid fraction 0.6), cube
        double side = std::cbrt(3.0 * vol / 0.6);
        std::sort(radii.begin(), radii.end(), [](double a, double b) { return a > b; });
        V3 lo = {0, 0, 0}, hi = {side, side, side};
        Rsa rsa(lo, hi, 2.0e-3);
        for (unsigned i = 0; i < n; i++) {
            double r = radii[i];
            bool ok = false;
            for (int a = 0; a < 100000 && !ok; a++) {
                V3 c = {r + (side - 2 * r) * xdem_rng_uniform(&vs), r + (side - 2 * r) * xdem_rng_uniform(&vs),
                        r + (side - 2 * r) * xdem_rng_uniform(&vs)};
                if (rsa.fits(c, r)) { rsa.add(c, r); ok = true; }
            }
            if (!ok) return XDEM_E_INVALID;
        }
        spatial_order(rsa.p, 4.0e-3, lo);
        fill_common(s, rsa.p, 2500.0);
        xdem_contact_property_make(E, E, 0.3, 0.3, rest, mu, 0.02, 0.0, 0.0, &s->cp_pp);
        box_walls(side, side, side, pl, dbi);
        set_walls(s, pl, dbi, s->cp_pp);
        // dt = 1e-5 s, capped at 0.2 of the Rayleigh time of the smallest sphere
        // (T_R = pi r sqrt(rho / G) / (0.1631 nu + 0.8766), G = E / (2 (1 + nu))): the
        // stiffest scenes (E near 1e8 Pa, r = 0.5 mm) diverge at 1e-5 s.
        const double rmin = radii.back(), nu = 0.3, G = E / (2.0 * (1.0 + nu));
        const double t_rayleigh = pi * rmin * std::sqrt(2500.0 / G) / (0.1631 * nu + 0.8766);
        s->dt = std::min(1e-5, 0.2 * t_rayleigh);
        s->nsteps = 1000;
        return XDEM_OK;
    }
    if (workload == XDEM_SCENE_PILE) {
        unsigned n = np_override ? np_override : 20000u;
        std::vector<double> radii(n);
        double vol = 0.0;
        for (unsigned i = 0; i < n; i++) {
            radii[i] = 1.0e-3 + 0.5e-3 * xdem_rng_uniform(&rng);
            vol += 4.0 / 3.0 * pi * radii[i] * radii[i] * radii[i];
        }
        const double R = 0.030;
        double H = vol / (0.30 * pi * R * R);
        std::sort(radii.begin(), radii.end(), [](double a, double b) { return a > b; });
        const double cx = 0.2, cy = 0.2;  // cylinder axis over the middle of a 0.4 m floor
        V3 lo = {cx - R, cy - R, 0.0}, hi = {cx + R, cy + R, H};
        Rsa rsa(lo, hi, 1.5e-3);
        for (unsigned i = 0; i < n; i++) {
            double r = radii[i];
            bool ok = false;
            for (int a = 0; a < 100000 && !ok; a++) {
                double rr = (R - r) * std::sqrt(xdem_rng_uniform(&rng));
                double th = 2.0 * pi * xdem_rng_uniform(&rng);
                V3 c = {cx + rr * std::cos(th), cy + rr * std::sin(th), r + (H - 2 * r) * xdem_rng_uniform(&rng)};
                if (rsa.fits(c, r)) { rsa.add(c, r); ok = true; }
            }
            if (!ok) return XDEM_E_INVALID;
        }
        spatial_order(rsa.p, 3.0e-3, lo);
        fill_common(s, rsa.p, 2500.0);
        xdem_contact_property_make(2e7, 2e7, 0.3, 0.3, 0.35, 0.45, 0.05, 0.0, 0.0, &s->cp_pp);
        const double f0[3] = {0, 0, 0}, f1[3] = {0.4, 0, 0}, f2[3] = {0.4, 0.4, 0}, f3[3] = {0, 0.4, 0};
        pl.resize(1);
        dbi.resize(15);
        xdem_make_plane(f0, f1, f2, f3, 0, 0.0, &pl[0], &dbi[0]);
        set_walls(s, pl, dbi, s->cp_pp);
        s->dt = 2e-5;
        s->nsteps = 4000;
        return XDEM_OK;
    }
    return XDEM_E_INVALID;
}

}  // extern "C"
