// xdem_part.cpp -- part files of the step layer: the reference's result file of one step
// (xResultManager::export_step_data_to_file, source/xdynamics_manager/xResultManager.cpp:828-925,
// VERSION_NUMBER 2 of include/xdynamics_decl.h:18) for one scene, and the restart from it
// (xDynamicsSimulator::set_from_part_result, source/xdynamics_simulation/xDynamicSimulator.cpp:392-492).
//
// Layout, native-endian, no padding (single spheres: nparticles = nclusters = np):
//   double ct; uint32 np; uint32 nclusters;
//   double mass[np]; double pos[4 np]; vel[3 np]; acc[3 np]; ep[4 np]; ev[4 np]; ea[4 np];
//   sphere history: uint32 count[np]; uint32 id[np cap]; double tsd[2 np cap]   (cap = p2p capacity)
//   wall history (scenes with walls): count[np]; id[3 np]; tsd[6 np]
//   cylinder history (scenes with cylinders): count[np]; id[3 np]; tsd[6 np]
//   triangle history (scenes with a mesh): count[np]; id[5 np]; tsd[10 np]
// The file carries no section flags (as the reference's): the reader takes them from the
// context, and checks the file's size against them. The multibody block of the reference's
// file (written when the model has generalised coordinates) is not part of the DEM state and
// is not written; the body poses stay with the caller's integrator. Differences: the reference's
// restart reader skips the V2 mass block that its writer emits and so reads every later array
// one block early; this reader reads the layout the writer writes. The reference's restart
// leaves the contact histories as they are (xContactManager::set_from_part_result,
// xContactManager.cpp:316-322, has an empty body): that is load's default, and
// XDEM_PART_HISTORY restores them from the file as well, so that a run continued from a part
// file equals the run that wrote it.
#include <cstdio>
#include <vector>

#include "xdem_capi.h"

namespace {

struct Sections {
    unsigned np = 0, cap = 0;
    bool walls = false, cyl = false, mesh = false;
};

int sections(xdem_ctx* ctx, unsigned scene, Sections& s)
{
    xdem_ctx_scene_info si;
    int rc = xdem_ctx_scene_info_get(ctx, scene, &si);
    if (rc) return rc;
    s.np = si.np;
    s.cap = si.p2p_capacity;
    s.walls = si.nplane > 0;
    s.cyl = si.ncylinder > 0;
    s.mesh = si.has_mesh != 0;
    return XDEM_OK;
}

size_t file_bytes(const Sections& s)
{
    size_t b = sizeof(double) + 2 * sizeof(unsigned);
    b += sizeof(double) * (size_t)s.np * (1 + 4 + 3 + 3 + 4 + 4 + 4);
    auto hist = [&](unsigned cap) { return sizeof(unsigned) * (size_t)s.np * (1 + cap) + sizeof(double) * 2 * (size_t)s.np * cap; };
    b += hist(s.cap);
    if (s.walls) b += hist(XDEM_MAX_P2PL_COUNT);
    if (s.cyl) b += hist(XDEM_MAX_P2CY_COUNT);
    if (s.mesh) b += hist(XDEM_MAX_P2MS_COUNT);
    return b;
}

// history sections in the file's order (xResultManager.cpp:889-912)
constexpr int kKinds = 4;

const int kFields[7] = {XDEM_FIELD_MASS, XDEM_FIELD_POS, XDEM_FIELD_VEL, XDEM_FIELD_ACC, XDEM_FIELD_EP, XDEM_FIELD_EV,
                        XDEM_FIELD_EA};
const unsigned kWidth[7] = {1, 4, 3, 3, 4, 4, 4};

}  // namespace

extern "C" {

int xdem_ctx_save_part(xdem_ctx* ctx, unsigned int scene, const char* path, double ct)
{
    if (!ctx || !path) return XDEM_E_INVALID;
    Sections s;
    int rc = sections(ctx, scene, s);
    if (rc) return rc;
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return XDEM_E_INVALID;
    bool ok = std::fwrite(&ct, sizeof(ct), 1, f) == 1;
    ok = ok && std::fwrite(&s.np, sizeof(unsigned), 1, f) == 1 && std::fwrite(&s.np, sizeof(unsigned), 1, f) == 1;
    std::vector<double> buf;
    for (int k = 0; k < 7 && ok; k++) {
        buf.resize((size_t)kWidth[k] * s.np);
        if ((rc = xdem_ctx_read(ctx, kFields[k], scene, buf.data()))) break;
        ok = std::fwrite(buf.data(), sizeof(double), buf.size(), f) == buf.size();
    }
    const int kinds[kKinds] = {XDEM_HISTORY_P2P, XDEM_HISTORY_PLANE, XDEM_HISTORY_CYLINDER, XDEM_HISTORY_MESH};
    const unsigned caps[kKinds] = {s.cap, XDEM_MAX_P2PL_COUNT, XDEM_MAX_P2CY_COUNT, XDEM_MAX_P2MS_COUNT};
    const bool present[kKinds] = {true, s.walls, s.cyl, s.mesh};
    for (int h = 0; h < kKinds && ok && !rc; h++) {
        if (!present[h]) continue;
        std::vector<unsigned> cnt(s.np), ids((size_t)s.np * caps[h]);
        std::vector<double> tsd(2 * (size_t)s.np * caps[h]);
        if ((rc = xdem_ctx_read_history(ctx, kinds[h], scene, cnt.data(), ids.data(), tsd.data()))) break;
        ok = std::fwrite(cnt.data(), sizeof(unsigned), cnt.size(), f) == cnt.size() &&
             std::fwrite(ids.data(), sizeof(unsigned), ids.size(), f) == ids.size() &&
             std::fwrite(tsd.data(), sizeof(double), tsd.size(), f) == tsd.size();
    }
    if (std::fclose(f) != 0) ok = false;
    if (rc) return rc;
    return ok ? XDEM_OK : XDEM_E_INVALID;
}

int xdem_ctx_load_part(xdem_ctx* ctx, unsigned int scene, const char* path, int flags, double* ct)
{
    const int known = XDEM_PART_HISTORY | XDEM_PART_POSITIONS;
    if (!ctx || !path || (flags & ~known) || flags == known) return XDEM_E_INVALID;
    Sections s;
    int rc = sections(ctx, scene, s);
    if (rc) return rc;
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return XDEM_E_INVALID;
    std::fseek(f, 0, SEEK_END);
    const long size = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (size < 0 || (size_t)size != file_bytes(s)) {   // another model's file
        std::fclose(f);
        return XDEM_E_INVALID;
    }
    double t = 0.0;
    unsigned n[2] = {0, 0};
    bool ok = std::fread(&t, sizeof(t), 1, f) == 1 && std::fread(n, sizeof(unsigned), 2, f) == 2 && n[0] == s.np &&
              n[1] == s.np;
    std::vector<std::vector<double>> fields(7);
    for (int k = 0; k < 7 && ok; k++) {
        fields[k].resize((size_t)kWidth[k] * s.np);
        ok = std::fread(fields[k].data(), sizeof(double), fields[k].size(), f) == fields[k].size();
    }
    const int kinds[kKinds] = {XDEM_HISTORY_P2P, XDEM_HISTORY_PLANE, XDEM_HISTORY_CYLINDER, XDEM_HISTORY_MESH};
    const unsigned caps[kKinds] = {s.cap, XDEM_MAX_P2PL_COUNT, XDEM_MAX_P2CY_COUNT, XDEM_MAX_P2MS_COUNT};
    const bool present[kKinds] = {true, s.walls, s.cyl, s.mesh};
    std::vector<unsigned> cnt[kKinds], ids[kKinds];
    std::vector<double> tsd[kKinds];
    for (int h = 0; h < kKinds && ok; h++) {
        if (!present[h]) continue;
        cnt[h].resize(s.np);
        ids[h].resize((size_t)s.np * caps[h]);
        tsd[h].resize(2 * (size_t)s.np * caps[h]);
        ok = std::fread(cnt[h].data(), sizeof(unsigned), cnt[h].size(), f) == cnt[h].size() &&
             std::fread(ids[h].data(), sizeof(unsigned), ids[h].size(), f) == ids[h].size() &&
             std::fread(tsd[h].data(), sizeof(double), tsd[h].size(), f) == tsd[h].size();
    }
    std::fclose(f);
    if (!ok) return XDEM_E_INVALID;
    // the state (set_dem_data) and the contact histories (xContactManager::set_from_part_result);
    // XDEM_PART_POSITIONS: the positions only (xParticleManager::SetCurrentParticlesFromPartResult,
    // xParticleManager.cpp:563-644, which copies pos and nothing else for single spheres)
    for (int k = 0; k < 7; k++) {
        if ((flags & XDEM_PART_POSITIONS) && kFields[k] != XDEM_FIELD_POS) continue;
        if ((rc = xdem_ctx_write(ctx, kFields[k], scene, fields[k].data()))) return rc;
    }
    for (int h = 0; h < kKinds && (flags & XDEM_PART_HISTORY); h++) {
        if (!present[h]) continue;
        if ((rc = xdem_ctx_write_history(ctx, kinds[h], scene, cnt[h].data(), ids[h].data(), tsd[h].data()))) return rc;
    }
    if (ct) *ct = t;
    return XDEM_OK;
}

}  // extern "C"
