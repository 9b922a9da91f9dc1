// xdem_generators.cpp -- the reference's particle generators (xParticleManager::CreateCubeParticle,
// CreatePlaneParticle, CreateCircleParticle; source/xdynamics_manager/xParticleManager.cpp:646-880)
// with their layouts, their draws from frand() = rand() / RAND_MAX (include/xdynamics_algebra/
// xAlgebraMath.h:8) and their quirks, on an explicit generator state.
//
// rand() is the C library's of the reference's platform (a Windows DLL: windows.h is included by
// include/xdynamics_decl.h): the linear congruential generator x <- 214013 x + 2531011 (mod 2^32),
// returning bits 16..30 of x, RAND_MAX = 32767, initial state 1 (srand(s) sets the state to s).
// The reference seeds it nowhere except CreateCircleParticle's radii (srand(GetTickCount()), a
// clock); here every draw comes from the caller's state.
#include <cmath>
#include <list>
#include <vector>

#include "xdem_host.h"

namespace {

struct V3 { double x, y, z; };
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// A = [u pu qu] (columns), pu = (-u.y, u.x, u.z), qu = u x pu (xParticleManager.cpp:740-744, 832-836)
struct Frame {
    double a[3][3];
    explicit Frame(V3 u)
    {
        const V3 pu = {-u.y, u.x, u.z};
        const V3 qu = cross(u, pu);
        const double m[9] = {u.x, pu.x, qu.x, u.y, pu.y, qu.y, u.z, pu.z, qu.z};
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) a[r][c] = m[3 * r + c];
    }
    V3 apply(V3 p) const
    {
        return {a[0][0] * p.x + a[0][1] * p.y + a[0][2] * p.z, a[1][0] * p.x + a[1][1] * p.y + a[1][2] * p.z,
                a[2][0] * p.x + a[2][1] * p.y + a[2][2] * p.z};
    }
};

}  // namespace

extern "C" {

unsigned int xdem_crt_rand(unsigned int* state)
{
    *state = *state * 214013u + 2531011u;
    return (*state >> 16) & 0x7fffu;
}

double xdem_crt_frand(unsigned int* state) { return xdem_crt_rand(state) / (double)32767; }

int xdem_create_cube_particles(const xdem_cube_particle_data* d, unsigned int* rand_state, double* pos,
                               unsigned int* np)
{
    if (!d || !np || (pos && !rand_state)) return XDEM_E_INVALID;
    const double r = d->maxr, diameter = 2.0 * r;
    const double dim[3] = {d->dx, d->dy, d->dz};
    unsigned nd[3];
    for (int k = 0; k < 3; k++) {
        // (unsigned)(dimension / diameter) - 1 per axis (xParticleManager.cpp:666-671); a box
        // less than two diameters wide would wrap the unsigned count
        if (!(r > 0.0) || !(dim[k] >= 2.0 * diameter)) return XDEM_E_INVALID;
        nd[k] = static_cast<unsigned>(dim[k] / diameter) - 1u;
    }
    const unsigned long long n = (unsigned long long)nd[0] * nd[1] * nd[2];
    if (n > 0xffffffffull) return XDEM_E_INVALID;
    if (!pos) {
        *np = (unsigned)n;
        return XDEM_OK;
    }
    if (*np < n) return XDEM_E_INVALID;
    double space[3], gab[3], ran[3];
    for (int k = 0; k < 3; k++) {
        space[k] = (dim[k] - diameter * nd[k]) / (nd[k] + 1);
        gab[k] = diameter + space[k];
        ran[k] = r * space[k];   // the jitter amplitude is r * space (xParticleManager.cpp:680)
    }
    unsigned cnt = 0;
    for (unsigned z = 0; z < nd[2]; z++) {
        const double _z = d->lz + r + space[2];
        for (unsigned y = 0; y < nd[1]; y++) {
            const double _y = d->ly + r + space[1];
            for (unsigned x = 0; x < nd[0]; x++) {
                const double _x = d->lx + r + space[0];
                // the three draws of new_vector4d(x, y, z, r), its arguments evaluated right to
                // left as the reference's compiler does: z first
                const double fz = xdem_crt_frand(rand_state);
                const double fy = xdem_crt_frand(rand_state);
                const double fx = xdem_crt_frand(rand_state);
                double* p = pos + 4 * (size_t)cnt;
                p[0] = _x + x * gab[0] + ran[0] * fx;
                p[1] = _y + y * gab[1] + ran[1] * fy;
                p[2] = _z + z * gab[2] + ran[2] * fz;
                p[3] = r;
                cnt++;
            }
        }
    }
    if (d->minr != d->maxr)
        for (unsigned i = 0; i < cnt; i++) pos[4 * (size_t)i + 3] = d->minr + (d->maxr - d->minr) * xdem_crt_frand(rand_state);
    *np = cnt;
    return XDEM_OK;
}

int xdem_create_plane_particles(const xdem_plane_particle_data* d, unsigned int* rand_state, int reference_quirks,
                                double* pos, unsigned int* np)
{
    if (!d || !np || (pos && d->minr != d->maxr && !rand_state)) return XDEM_E_INVALID;
    const double mr = d->maxr;
    if (!(mr > 0.0) || !(d->dx >= 4.0 * mr) || !(d->dz >= 4.0 * mr) || !(d->dy >= 0.0)) return XDEM_E_INVALID;
    const unsigned nx = static_cast<unsigned>(d->dx / (2.0 * mr)) - 1u;
    const double space_x = (d->dx - 2.0 * nx * mr) / (nx + 1);
    const unsigned nz = static_cast<unsigned>(d->dz / (2.0 * mr)) - 1u;
    const double space_z = (d->dz - 2.0 * nz * mr) / (nz + 1);
    // dy is a number of layers: `for (unsigned i = 0; i < d.dy; i++)`, each layer the same
    // positions (xParticleManager.cpp:751-771)
    unsigned layers = 0;
    while (layers < d->dy) layers++;
    const unsigned long long n = (unsigned long long)layers * nx * nz;
    if (n > 0xffffffffull) return XDEM_E_INVALID;
    if (!pos) {
        *np = (unsigned)n;
        return XDEM_OK;
    }
    if (*np < n) return XDEM_E_INVALID;
    const Frame A({d->drx, d->dry, d->drz});
    const V3 sp = {d->lx, d->ly, d->lz};
    unsigned cnt = 0;
    for (unsigned i = 0; i < layers; i++)
        for (unsigned z = 0; z < nz; z++) {
            const double gab_z = (2.0 * mr + space_z) * z + space_z;
            for (unsigned x = 0; x < nx; x++) {
                const double x_pos = (2.0 * mr + space_x) * x + space_x;
                const V3 q = A.apply({0.0, x_pos + mr + space_x, gab_z + mr + space_z});
                double* p = pos + 4 * (size_t)cnt;
                p[0] = sp.x + q.x; p[1] = sp.y + q.y; p[2] = sp.z + q.z; p[3] = mr;
                cnt++;
            }
        }
    if (d->minr != d->maxr) {
        for (unsigned i = 0; i < cnt; i++) pos[4 * (size_t)i + 3] = d->minr + (d->maxr - d->minr) * xdem_crt_frand(rand_state);
    } else if (reference_quirks && cnt > 1) {
        // the reference's copy loop advances its index only inside `if (d.minr != d.maxr)`
        // (xParticleManager.cpp:778-782): with equal radii every position lands in pos[0],
        // the last one staying, and the others keep the allocation's zeros
        double last[4] = {pos[4 * (size_t)(cnt - 1)], pos[4 * (size_t)(cnt - 1) + 1], pos[4 * (size_t)(cnt - 1) + 2],
                          pos[4 * (size_t)(cnt - 1) + 3]};
        for (unsigned i = 0; i < 4 * cnt; i++) pos[i] = 0.0;
        for (int k = 0; k < 4; k++) pos[k] = last[k];
    }
    *np = cnt;
    return XDEM_OK;
}

int xdem_create_circle_particles(const xdem_circle_particle_data* d, unsigned int* rand_state, double* pos,
                                 unsigned int* np)
{
    if (!d || !np || (pos && d->minr != d->maxr && !rand_state)) return XDEM_E_INVALID;
    const double pi = 3.14159265358979323846;   // M_PI
    const double r = d->maxr, cr = 0.5 * d->diameter;
    if (!(r > 0.0) || !(cr >= 4.0 * r)) return XDEM_E_INVALID;
    const unsigned nr = static_cast<unsigned>(cr / (2.0 * r)) - 1u;
    const double rr = 2.0 * r * nr + r;
    const double space = (cr - rr) / (nr + 1);
    std::vector<unsigned> per_ring(nr + 1, 0);
    unsigned long long n = 0;
    for (unsigned i = 1; i <= nr; i++) {
        const double _r = i * (2.0 * r + space);
        const double dth = (2.0 * r + space) / _r;
        per_ring[i] = static_cast<unsigned>((2.0 * pi) / dth);
        n += per_ring[i];
    }
    if (n > 0xffffffffull) return XDEM_E_INVALID;
    if (!pos) {
        *np = (unsigned)n;
        return XDEM_OK;
    }
    if (*np < n) return XDEM_E_INVALID;
    const Frame A({d->dx, d->dy, d->dz});
    const V3 sp = {d->sx, d->sy, d->sz};
    unsigned cnt = 0;
    for (unsigned i = 1; i <= nr; i++) {   // xParticleManager.cpp:840-863
        const double _r = i * (2.0 * r + space);
        const unsigned npr = per_ring[i];
        const double dth = (2.0 * pi) / npr;
        for (unsigned j = 0; j < npr; j++) {
            const V3 q = A.apply({0.0, _r * std::sin(dth * j), _r * std::cos(dth * j)});
            double* p = pos + 4 * (size_t)cnt;
            p[0] = sp.x + q.x; p[1] = sp.y + q.y; p[2] = sp.z + q.z; p[3] = r;
            cnt++;
        }
    }
    if (d->minr != d->maxr)
        for (unsigned i = 0; i < cnt; i++) pos[4 * (size_t)i + 3] = d->minr + (d->maxr - d->minr) * xdem_crt_frand(rand_state);
    *np = cnt;
    return XDEM_OK;
}

}  // extern "C"
