// sht_batch.cpp -- kernel-shape autotuner (cushtns_tune) and uneven batches of transforms on
// different configurations replayed from one HIP graph (cushtns_batch_*).  Neither has a
// counterpart in SHTns, whose GPU path runs one field of one configuration per call
// (sht_gpu.cu:495-569).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "sht_internal.h"

using namespace shtns_amd;

namespace {

struct DevBuf {
    void *p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

// milliseconds per call of f on stream s: one untimed call, then `reps` calls between events
template <class F>
double time_calls(hipStream_t s, int reps, F &&f) {
    f();
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return 1e30;
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return 1e30;
    }
    (void)hipEventRecord(a, s);
    for (int r = 0; r < reps; ++r) f();
    (void)hipEventRecord(b, s);
    float ms = 0.f;
    const bool ok = hipEventSynchronize(b) == hipSuccess && hipEventElapsedTime(&ms, a, b) == hipSuccess;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (!ok || shtns_last_error()) return 1e30;
    return ms / reps;
}

bool tune_verbose() {
    static const bool v = getenv("SHTNS_AMD_TUNE_VERBOSE") != nullptr;
    return v;
}

// ---- tuning file (the GPU counterpart of SHTns's SHT_LOAD_SAVE_CFG, sht_init.c:975-1061) ----
// One line per tuned (configuration, batch size), appended by cushtns_tune:
//   shtns_amd-tune 1 <gfx arch> lmax mmax mres nlat nphi grid_flags nfields  ncg nw ancg fsyn fana
// The file is $SHTNS_AMD_CFG, else "shtns_amd_cfg" in the working directory.  Lines that do
// not parse, or are for another device or build, are skipped.
const char *tune_file() {
    const char *e = getenv("SHTNS_AMD_CFG");
    return (e && *e) ? e : "shtns_amd_cfg";
}

std::string device_arch(int dev) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) return "unknown";
    std::string a = p.gcnArchName;
    const size_t colon = a.find(':');   // "gfx950:sramecc+:xnack-" -> "gfx950"
    return colon == std::string::npos ? a : a.substr(0, colon);
}

bool tune_cache_save(shtns_cfg c, int nfields) {
    FILE *f = fopen(tune_file(), "a");
    if (!f) return false;
    fprintf(f, "shtns_amd-tune 1 %s %d %d %d %u %u %d %d  %d %d %d %d %d\n", device_arch(c->device).c_str(), c->lmax,
            c->mmax, c->mres, c->nlat, c->nphi, c->grid_flags, nfields, c->tune_synth_ncg, c->tune_anal_nw,
            c->tune_anal_ncg, c->tune_fft_synth, c->tune_fft_anal);
    const bool ok = fclose(f) == 0;
    return ok;
}

}  // namespace

bool shtns_amd::tune_cache_apply(shtns_cfg c, int nfields) {
    FILE *f = fopen(tune_file(), "r");
    if (!f) return false;
    const std::string arch = device_arch(c->device);
    char line[512], tag[32], a[64];
    int found = 0, ch[5] = {0, 0, 0, 0, 0};
    while (fgets(line, sizeof line, f)) {
        int ver, lmax, mmax, mres, gflags, nf, v[5];
        unsigned nlat, nphi;
        if (sscanf(line, "%31s %d %63s %d %d %d %u %u %d %d %d %d %d %d %d", tag, &ver, a, &lmax, &mmax, &mres, &nlat,
                   &nphi, &gflags, &nf, &v[0], &v[1], &v[2], &v[3], &v[4]) != 15)
            continue;
        if (strcmp(tag, "shtns_amd-tune") != 0 || ver != 1 || arch != a) continue;
        if (lmax != c->lmax || mmax != c->mmax || mres != c->mres || nlat != c->nlat || nphi != c->nphi ||
            gflags != c->grid_flags)
            continue;
        if (nfields > 0 && nf != nfields) continue;
        std::copy(v, v + 5, ch);   // the last matching line wins
        found = 1;
    }
    fclose(f);
    if (!found || cushtns_set_tuning(c, ch) != 0) {
        clear_error();
        return false;
    }
    return true;
}

extern "C" int cushtns_tune(shtns_cfg c, int nfields) {
    clear_error();
    if (c == nullptr || c->ct == nullptr || nfields <= 0) {
        set_error(1, "cushtns_tune: grid not set or nfields <= 0");
        return -1;
    }
    if (!c->gpu_ready && cushtns_init_gpu(c) < 0) {
        set_error(3, "cushtns_tune: no usable GPU");
        return -1;
    }
    (void)hipSetDevice(c->device);
    if (c->load_save && tune_cache_apply(c, nfields)) return 0;   // tuned by an earlier run
    const hipStream_t s = c->comp_stream;
    DevBuf q, v;
    const size_t qb = (size_t)nfields * c->nlm * sizeof(cplx), vb = (size_t)nfields * c->nspat * sizeof(double);
    if (hipMalloc(&q.p, qb) != hipSuccess || hipMalloc(&v.p, vb) != hipSuccess) {
        set_error(2, "cushtns_tune: device allocation failed");
        return -1;
    }
    if (hipMemsetAsync(q.p, 0, qb, s) != hipSuccess || hipMemsetAsync(v.p, 0, vb, s) != hipSuccess) {
        set_error(2, "cushtns_tune: memset failed");
        return -1;
    }
    cplx *Q = (cplx *)q.p;
    double *V = (double *)v.p;
    const int M = c->nphi / 2;
    std::vector<int> fft_kinds;
    if (fft_wave_supported(M)) fft_kinds.push_back(1);
    if (lds_fft_specialized(M)) fft_kinds.push_back(2);
    if (fft_coop_supported(M)) {
        for (int k = 3; k <= 6; ++k) fft_kinds.push_back(k);
    }
    if (fft_kinds.size() < 2) fft_kinds = {0};

    // synthesis: field groups per wave x FFT kernel
    std::vector<int> ncgs{1};
    if (nfields > 4) ncgs.push_back(2);
    if (nfields > 8) ncgs.push_back(4);
    double best = 1e30;
    int b_ncg = 0, b_fft = 0;
    for (int ncg : ncgs)
        for (int fk : fft_kinds) {
            c->tune_synth_ncg = ncg;
            c->tune_fft_synth = fk;
            const double t = time_calls(s, 3, [&] { cu_SH_to_spat_batch(c, Q, V, c->lmax, nfields, c->nlm, c->nspat); });
            if (tune_verbose()) printf("[shtns_amd tune] synth ncg=%d fft=%d: %.4f ms\n", ncg, fk, t);
            if (t < best) { best = t; b_ncg = ncg; b_fft = fk; }
        }
    c->tune_synth_ncg = b_ncg;
    c->tune_fft_synth = b_fft;

    // analysis: waves per block x field groups per block x FFT kernel (0 = automatic shape)
    const int ntiles = (c->nlat_2 + 63) / 64;
    std::vector<int> nws{0};
    for (int d = 1; d <= 12; ++d)
        if (ntiles % d == 0 && anal_shape_ok(d, ntiles / d)) nws.push_back(d);
    best = 1e30;
    int b_nw = 0, b_ancg = 0, b_afft = 0;
    for (int nw : nws)
        for (int ncg = 1; ncg <= 2; ++ncg) {
            if (ncg == 2 && (nfields <= 4 || (nw > 0 && ntiles / nw > 3))) continue;
            for (int fk : fft_kinds) {
                c->tune_anal_nw = nw;
                c->tune_anal_ncg = ncg;
                c->tune_fft_anal = fk;
                const double t =
                    time_calls(s, 3, [&] { cu_spat_to_SH_batch(c, V, Q, c->lmax, nfields, c->nspat, c->nlm); });
                if (tune_verbose()) printf("[shtns_amd tune] anal nw=%d ncg=%d fft=%d: %.4f ms\n", nw, ncg, fk, t);
                if (t < best) { best = t; b_nw = nw; b_ancg = ncg; b_afft = fk; }
            }
        }
    c->tune_anal_nw = b_nw;
    c->tune_anal_ncg = b_ancg;
    c->tune_fft_anal = b_afft;
    if (hipStreamSynchronize(s) != hipSuccess) {
        set_error(2, "cushtns_tune: stream error");
        return -1;
    }
    clear_error();
    if (c->load_save && !tune_cache_save(c, nfields))
        fprintf(stderr, "! Warning ! shtns_amd could not save the tuning to %s\n", tune_file());
    return 0;
}

extern "C" int cushtns_tuning(shtns_cfg c, int *choices) {
    if (c == nullptr || choices == nullptr) return -1;
    choices[0] = c->tune_synth_ncg;
    choices[1] = c->tune_anal_nw;
    choices[2] = c->tune_anal_ncg;
    choices[3] = c->tune_fft_synth;
    choices[4] = c->tune_fft_anal;
    return 0;
}

extern "C" int cushtns_set_tuning(shtns_cfg c, const int *choices) {
    clear_error();
    if (c == nullptr || choices == nullptr) {
        set_error(1, "cushtns_set_tuning: null argument");
        return -1;
    }
    const int ncg = choices[0], nw = choices[1], ancg = choices[2], fs = choices[3], fa = choices[4];
    const bool ok = (ncg == 0 || ncg == 1 || ncg == 2 || ncg == 4) && nw >= 0 && nw <= 12 && ancg >= 0 &&
                    ancg <= 2 && fs >= 0 && fs <= 6 && fa >= 0 && fa <= 6;
    if (!ok) {
        set_error(1, "cushtns_set_tuning: invalid choice");
        return -1;
    }
    c->tune_synth_ncg = ncg;
    c->tune_anal_nw = nw;
    c->tune_anal_ncg = ancg;
    c->tune_fft_synth = fs;
    c->tune_fft_anal = fa;
    return 0;
}

// ------------------------------------------------------------------------------------------
// uneven batches
// ------------------------------------------------------------------------------------------
struct shtns_batch_info {
    struct Group {
        shtns_cfg src = nullptr;     // the caller's configuration
        shtns_cfg clone = nullptr;   // private clone running on `stream`
        hipStream_t stream = nullptr;
        hipEvent_t done = nullptr;
        int n = 0;
        cplx *q = nullptr;           // n * nlm
        double *v = nullptr;         // n * nspat
    };
    std::vector<Group> groups;
    std::vector<std::pair<int, int>> where;   // transform -> (group, slot)
    hipStream_t origin = nullptr;             // capture stream
    hipEvent_t fork = nullptr;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    int device = 0;
};

namespace {

void batch_free(shtns_batch b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    for (int d = 0; d < 2; ++d) {
        if (b->exec[d]) (void)hipGraphExecDestroy(b->exec[d]);
        if (b->graph[d]) (void)hipGraphDestroy(b->graph[d]);
    }
    for (auto &g : b->groups) {
        if (g.stream) (void)hipStreamSynchronize(g.stream);
        if (g.clone) shtns_destroy(g.clone);
        if (g.q) (void)hipFree(g.q);
        if (g.v) (void)hipFree(g.v);
        if (g.done) (void)hipEventDestroy(g.done);
        if (g.stream) (void)hipStreamDestroy(g.stream);
    }
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->origin) (void)hipStreamDestroy(b->origin);
    delete b;
}

// enqueue direction d (0 synthesis, 1 analysis) of every group on its own stream
void enqueue_groups(shtns_batch b, int d) {
    for (auto &g : b->groups) {
        shtns_cfg c = g.clone;
        if (d == 0) cu_SH_to_spat_batch(c, g.q, g.v, c->lmax, g.n, c->nlm, c->nspat);
        else cu_spat_to_SH_batch(c, g.v, g.q, c->lmax, g.n, c->nspat, c->nlm);
    }
}

bool capture(shtns_batch b, int d) {
    if (hipStreamBeginCapture(b->origin, hipStreamCaptureModeThreadLocal) != hipSuccess) return false;
    bool ok = hipEventRecord(b->fork, b->origin) == hipSuccess;
    for (auto &g : b->groups) ok = ok && hipStreamWaitEvent(g.stream, b->fork, 0) == hipSuccess;
    if (ok) enqueue_groups(b, d);
    ok = ok && shtns_last_error() == 0;
    for (auto &g : b->groups) {
        ok = ok && hipEventRecord(g.done, g.stream) == hipSuccess;
        ok = ok && hipStreamWaitEvent(b->origin, g.done, 0) == hipSuccess;
    }
    hipGraph_t graph = nullptr;
    const bool ended = hipStreamEndCapture(b->origin, &graph) == hipSuccess;
    if (!ok || !ended || !graph) {
        if (graph) (void)hipGraphDestroy(graph);
        return false;
    }
    b->graph[d] = graph;
    return hipGraphInstantiate(&b->exec[d], graph, nullptr, nullptr, 0) == hipSuccess;
}

}  // namespace

extern "C" shtns_batch cushtns_batch_create(int n, const shtns_cfg *cfgs, int prec, int tune) {
    clear_error();
    if (n <= 0 || cfgs == nullptr) {
        set_error(1, "cushtns_batch_create: empty batch");
        return nullptr;
    }
    if (prec < SHT_PREC_AUTO || prec > SHT_PREC_SPLIT) {
        set_error(1, "cushtns_batch_create: unknown precision %d", prec);
        return nullptr;
    }
    shtns_batch b = new shtns_batch_info();
    b->where.resize(n);
    // group by configuration, heaviest (largest lmax * nlat) first so its kernels start first
    std::vector<shtns_cfg> uniq;
    for (int i = 0; i < n; ++i) {
        shtns_cfg c = cfgs[i];
        if (c == nullptr || c->ct == nullptr) {
            set_error(1, "cushtns_batch_create: configuration %d has no grid", i);
            batch_free(b);
            return nullptr;
        }
        if (std::find(uniq.begin(), uniq.end(), c) == uniq.end()) uniq.push_back(c);
    }
    std::stable_sort(uniq.begin(), uniq.end(), [](shtns_cfg a, shtns_cfg c) {
        return (double)a->lmax * a->nlat > (double)c->lmax * c->nlat;
    });
    for (shtns_cfg c : uniq) {
        if (!c->gpu_ready && cushtns_init_gpu(c) < 0) {
            set_error(3, "cushtns_batch_create: no usable GPU");
            batch_free(b);
            return nullptr;
        }
        shtns_batch_info::Group g;
        g.src = c;
        b->groups.push_back(g);
    }
    b->device = uniq[0]->device;
    (void)hipSetDevice(b->device);
    for (int i = 0; i < n; ++i) {
        const int gi = (int)(std::find(uniq.begin(), uniq.end(), cfgs[i]) - uniq.begin());
        b->where[i] = {gi, b->groups[gi].n++};
    }
    bool ok = hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess;
    for (auto &g : b->groups) {
        if (!ok) break;
        ok = hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&g.done, hipEventDisableTiming) == hipSuccess;
        if (!ok) break;
        g.clone = cushtns_clone(g.src, g.stream, g.stream);
        ok = g.clone != nullptr;
        if (!ok) break;
        g.clone->prec = prec;
        const size_t qb = (size_t)g.n * g.src->nlm * sizeof(cplx), vb = (size_t)g.n * g.src->nspat * sizeof(double);
        ok = hipMalloc((void **)&g.q, qb) == hipSuccess && hipMalloc((void **)&g.v, vb) == hipSuccess &&
             hipMemsetAsync(g.q, 0, qb, g.stream) == hipSuccess && hipMemsetAsync(g.v, 0, vb, g.stream) == hipSuccess;
        if (ok && tune) ok = cushtns_tune(g.clone, g.n) == 0;
    }
    if (!ok) {
        if (!shtns_last_error()) set_error(2, "cushtns_batch_create: setup failed");
        batch_free(b);
        return nullptr;
    }
    // one eager pass in each direction sizes every clone's scratch (no allocation may happen
    // while a graph is being captured), then capture both directions
    enqueue_groups(b, 0);
    enqueue_groups(b, 1);
    for (auto &g : b->groups) ok = ok && hipStreamSynchronize(g.stream) == hipSuccess;
    ok = ok && shtns_last_error() == 0;
    ok = ok && capture(b, 0) && capture(b, 1);
    if (!ok) {
        if (!shtns_last_error()) set_error(2, "cushtns_batch_create: graph capture failed");
        batch_free(b);
        return nullptr;
    }
    clear_error();
    return b;
}

extern "C" int cushtns_batch_arrays(shtns_batch b, int i, cplx **Qlm, double **Vr) {
    clear_error();
    if (b == nullptr || i < 0 || i >= (int)b->where.size()) {
        set_error(1, "cushtns_batch_arrays: no transform %d", i);
        return -1;
    }
    const auto &g = b->groups[b->where[i].first];
    const int slot = b->where[i].second;
    if (Qlm) *Qlm = g.q + (size_t)slot * g.src->nlm;
    if (Vr) *Vr = g.v + (size_t)slot * g.src->nspat;
    return 0;
}

static int batch_run(shtns_batch b, int d, hipStream_t stream) {
    clear_error();
    if (b == nullptr || !b->exec[d]) {
        set_error(1, "cushtns_batch: invalid batch");
        return -1;
    }
    (void)hipSetDevice(b->device);
    if (hipGraphLaunch(b->exec[d], stream) != hipSuccess) {
        set_error(2, "cushtns_batch: graph launch failed");
        return -1;
    }
    return 0;
}

extern "C" int cushtns_batch_SH_to_spat(shtns_batch b, hipStream_t stream) { return batch_run(b, 0, stream); }
extern "C" int cushtns_batch_spat_to_SH(shtns_batch b, hipStream_t stream) { return batch_run(b, 1, stream); }
extern "C" void cushtns_batch_destroy(shtns_batch b) { batch_free(b); }
