// xdem_autotune.cpp -- launch-shape tuner of the step layer.
//
// The reference fixes its launch shapes (128 threads per block, decl:17; one cell size of twice
// the largest radius, xDiscreteElementMethodSimulation.cpp:100). The step layer has knobs whose
// best values depend on the scenes: the blocks of the full-pass kernels (how many waves hide
// the neighbour gathers against how many blocks find an empty queue) and the bucket budget
// per sphere (grid aliasing against the size of the bucket table the rebuild scans) and the
// adaptive skin's cap (fewer list builds against more listed pairs and larger buckets). Results
// do not depend on any of them, bit for bit, so the tuner only measures: it builds a context
// per candidate on the caller's scenes, runs warm-up steps (which include list builds) and
// times trial steps, by coordinate descent (full-pass blocks, buckets per sphere, then the skin
// cap when the caller's skin is the adaptive one). The
// winner is cached in a text file keyed by a signature of the scenes (sphere and scene
// counts, radius range, precision, device), one line per key, so later runs with the same
// kind of scenes skip the trials.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "xdem_capi.h"

namespace {

struct Trial {
    unsigned full_blocks, buckets_per_sphere;
    double skin_cap;
    double ms_per_step;
};

std::string signature(const xdem_scene* scenes, unsigned nscene, const xdem_ctx_config& cfg)
{
    unsigned long long n = 0;
    double rmin = 1e300, rmax = 0.0;
    unsigned nwall = 0;
    for (unsigned s = 0; s < nscene; s++) {
        n += scenes[s].np;
        nwall += scenes[s].nplane;
        for (unsigned i = 0; i < scenes[s].np; i++) {
            const double r = scenes[s].pos[4 * i + 3];
            rmin = r < rmin ? r : rmin;
            rmax = r > rmax ? r : rmax;
        }
    }
    // sizes to two significant digits: scenes of the same kind share a key
    auto sig2 = [](double v) {
        char b[32];
        std::snprintf(b, sizeof(b), "%.1e", v);
        return std::string(b);
    };
    return "np=" + sig2((double)n) + ";nscene=" + std::to_string(nscene) + ";walls=" + std::to_string(nwall) +
           ";r=" + sig2(rmin) + ".." + sig2(rmax) + ";prec=" + std::to_string(cfg.precision) +
           ";fuse=" + std::to_string(cfg.fuse_velocity) + ";dev=" + std::to_string(cfg.device);
}

bool cache_lookup(const char* path, const std::string& key, unsigned& fb, unsigned& bps, double& cap, double& ms)
{
    if (!path) return false;
    std::FILE* f = std::fopen(path, "r");
    if (!f) return false;
    char line[1024];
    bool found = false;
    while (std::fgets(line, sizeof(line), f)) {
        char k[768];
        unsigned a = 0, b = 0;
        double c = 0.0, t = 0.0;
        if (std::sscanf(line, "%767s %u %u %lf %lf", k, &a, &b, &c, &t) == 5 && key == k) {
            fb = a; bps = b; cap = c; ms = t;
            found = true;   // the last line for a key wins
        }
    }
    std::fclose(f);
    return found;
}

void cache_store(const char* path, const std::string& key, unsigned fb, unsigned bps, double cap, double ms)
{
    if (!path) return;
    std::FILE* f = std::fopen(path, "a");
    if (!f) return;
    std::fprintf(f, "%s %u %u %.3f %.6f\n", key.c_str(), fb, bps, cap, ms);
    std::fclose(f);
}

// ms per step of one candidate, or a negative status.
double run_trial(const xdem_scene* scenes, unsigned nscene, const xdem_ctx_config& base, unsigned fb, unsigned bps,
                 double cap, unsigned warmup, unsigned steps, int* status)
{
    xdem_ctx_config cfg = base;
    cfg.full_blocks = fb;
    cfg.buckets_per_sphere = bps;
    cfg.skin_cap = cap;
    xdem_ctx* ctx = nullptr;
    int rc = xdem_ctx_create(&cfg, &ctx);
    for (unsigned s = 0; s < nscene && rc == XDEM_OK; s++) rc = xdem_ctx_add_scene(ctx, &scenes[s], nullptr);
    if (rc == XDEM_OK) rc = xdem_ctx_commit(ctx);
    if (rc == XDEM_OK) rc = xdem_ctx_step(ctx, warmup);
    if (rc == XDEM_OK) rc = xdem_ctx_synchronize(ctx);
    double ms = -1.0;
    if (rc == XDEM_OK) {
        const auto t0 = std::chrono::steady_clock::now();
        rc = xdem_ctx_step(ctx, steps);
        if (rc == XDEM_OK) rc = xdem_ctx_synchronize(ctx);
        const auto t1 = std::chrono::steady_clock::now();
        ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / steps;
    }
    if (rc == XDEM_OK) {
        unsigned err = 0;
        rc = xdem_ctx_error_word(ctx, &err, 1);
        if (rc == XDEM_OK && (err & ~(unsigned)XDEM_ERR_GRID_SYNC)) rc = XDEM_E_INVALID;   // a broken candidate
    }
    if (ctx) xdem_ctx_destroy(ctx);
    *status = rc;
    return ms;
}

}  // namespace

extern "C" int xdem_autotune(const xdem_scene* scenes, unsigned int nscene, const xdem_ctx_config* base,
                             unsigned int warmup, unsigned int steps, const char* cache_path,
                             xdem_autotune_result* out)
{
    if (!scenes || !nscene || !out || !steps) return XDEM_E_INVALID;
    xdem_ctx_config cfg = {};
    if (base) cfg = *base;
    std::memset(out, 0, sizeof(*out));
    const std::string key = signature(scenes, nscene, cfg);
    std::snprintf(out->key, sizeof(out->key), "%s", key.c_str());
    unsigned fb = 0, bps = 0;
    double ms = 0.0, cap = 0.0;
    if (cache_lookup(cache_path, key, fb, bps, cap, ms)) {
        out->full_blocks = fb;
        out->buckets_per_sphere = bps;
        out->skin_cap = cap;
        out->ms_per_step = ms;
        out->from_cache = 1;
        return XDEM_OK;
    }
    static const unsigned kBlocks[] = {256, 512, 1024, 2048};
    static const unsigned kBuckets[] = {2, 4, 8};
    static const double kCaps[] = {0.2, 0.4, 0.8};
    const double cap0 = cfg.skin_cap > 0.0 ? cfg.skin_cap : 0.4;
    std::vector<Trial> trials;
    auto trial = [&](unsigned f, unsigned b, double c) -> int {
        for (const Trial& t : trials)
            if (t.full_blocks == f && t.buckets_per_sphere == b && t.skin_cap == c) return XDEM_OK;
        int st = XDEM_OK;
        const double t = run_trial(scenes, nscene, cfg, f, b, c, warmup, steps, &st);
        if (st == XDEM_E_HIP || st == XDEM_E_NOT_READY) return st;   // the device failed: stop
        if (st == XDEM_OK) trials.push_back({f, b, c, t});
        return XDEM_OK;
    };
    // coordinate descent: the block count at the default bucket budget, then the budget, then
    // the skin cap (only for the adaptive skin)
    unsigned best_f = 1024, best_b = 4;
    for (unsigned f : kBlocks) {
        int rc = trial(f, best_b, cap0);
        if (rc) return rc;
    }
    auto best_of = [&]() {
        const Trial* b = nullptr;
        for (const Trial& t : trials)
            if (!b || t.ms_per_step < b->ms_per_step) b = &t;
        return b;
    };
    if (const Trial* b = best_of()) best_f = b->full_blocks;
    for (unsigned b : kBuckets) {
        int rc = trial(best_f, b, cap0);
        if (rc) return rc;
    }
    if (cfg.skin == 0.0) {
        if (const Trial* b = best_of()) best_b = b->buckets_per_sphere;
        for (double c : kCaps) {
            int rc = trial(best_f, best_b, c);
            if (rc) return rc;
        }
    }
    const Trial* b = best_of();
    if (!b) return XDEM_E_INVALID;
    out->full_blocks = b->full_blocks;
    out->buckets_per_sphere = b->buckets_per_sphere;
    out->skin_cap = b->skin_cap;
    out->ms_per_step = b->ms_per_step;
    out->ntrial = (unsigned)std::min<size_t>(trials.size(), XDEM_AUTOTUNE_MAX_TRIALS);
    for (unsigned k = 0; k < out->ntrial; k++) {
        out->trial_full_blocks[k] = trials[k].full_blocks;
        out->trial_buckets_per_sphere[k] = trials[k].buckets_per_sphere;
        out->trial_skin_cap[k] = trials[k].skin_cap;
        out->trial_ms_per_step[k] = trials[k].ms_per_step;
    }
    cache_store(cache_path, key, out->full_blocks, out->buckets_per_sphere, out->skin_cap, out->ms_per_step);
    return XDEM_OK;
}
