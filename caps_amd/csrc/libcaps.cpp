// Host side of libcaps_amd: the libcaps C ABI over the gfx950 kernels.
//
// One device context per caps_t (stream, events, a growable HBM arena,
// pinned staging buffers), created on the first GPU call.  A batch of
// (ξ, m) blocks is laid out in the arena (BlockDesc offsets), the pipeline
// Mie -> K -> I -> assemble(1-M) -> trace -> LU runs asynchronously on the
// context's stream, and only the per-block scalars come back.  The
// determinant shortcuts of the reference (src/matrix.c:78-82, 144-145,
// 494-505) are applied to those scalars on the host.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "blocks.h"
#include "../../include/libcaps.h"

using capsd::BlockDesc;
using capsd::BlockOut;
using capsd::KPart;
using capsd::IG_MAX;
using capsd::IGroup;

struct capsx_mlp;
extern "C" {
void capsx_mlp_predict(capsx_mlp* h, const float* X, int n, float* Y);
void capsd_launch_mie(const BlockDesc*, char*, const double*, int, int, hipStream_t);
void capsd_launch_K(const BlockDesc*, char*, const double*, int, int, int, int, int*, hipStream_t);
void capsd_launch_I(const BlockDesc*, char*, const void*, const int*, const double*, int64_t, int*, const BlockOut*,
                    int, int, hipStream_t);
void capsd_launch_assemble(const BlockDesc*, char*, const double*, int, int, int, const BlockOut*, int, hipStream_t);
void capsd_launch_trace(const BlockDesc*, char*, BlockOut*, int, int, hipStream_t);
void capsd_launch_dominance(const BlockDesc*, char*, BlockOut*, int, int, int, hipStream_t);
void capsd_launch_frobenius(const BlockDesc*, char*, BlockOut*, int, int, hipStream_t);
void capsd_launch_fill_xi0(const BlockDesc*, char*, int64_t, int, hipStream_t);
void capsd_launch_mie_points(int, const double*, const int*, double, const double*, double*, hipStream_t);
void capsd_launch_fresnel_points(int, const double*, const double*, const double*, double*, hipStream_t);
void capsd_launch_bessel_points(int, const int*, const double*, const double*, double*, hipStream_t);
int capsd_launch_lu(const BlockDesc*, char*, const void*, BlockOut*, int, int, hipStream_t, int, int, int,
                    hipEvent_t*, int*, int);
void capsd_launch_lu_fill(const BlockDesc*, char*, int, unsigned long long, int, hipStream_t);
int capsd_lu_max_updates(int);
int capsd_lu_grid_panel_rows(int);
int capsd_lu_init(void);
int capsd_lu_max_dim(void);
int capsd_lu_nb(void);
int capsd_lu_npanels(void);
int capsd_lu_kernels(int gk, int ps, int nt, int prev[3]);
void capsd_launch_mfma_probe(const double*, const double*, double*, hipStream_t);
int capsd_noalloc_scratch(char** ptr, size_t* bytes);
void capsd_launch_noalloc_setup(char*, int64_t, int, int, int, int64_t, int64_t, int64_t, int64_t, int64_t, int64_t,
                                int64_t, int64_t, hipStream_t);
void capsd_launch_noalloc_result(const void*, double*, int, hipStream_t);
}

#define TERMINATE(cond, ...)                                                                      \
    do {                                                                                          \
        if (cond) {                                                                               \
            fprintf(stderr, "Fatal error: ");                                                     \
            fprintf(stderr, __VA_ARGS__);                                                         \
            fprintf(stderr, " (in %s, %s:%d)\n", __func__, __FILE__, __LINE__);                   \
            abort();                                                                              \
        }                                                                                         \
    } while (0)

#define HIPCHECK(x)                                                                               \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        TERMINATE(e_ != hipSuccess, "HIP error %s at %s", hipGetErrorString(e_), #x);             \
    } while (0)

namespace {

struct LuSide {
    int64_t off_panel, off_logs, off_moved, off_ut, off_inv;
};

double* g_lfac_dev = nullptr;
std::once_flag g_dev_once;

void device_init()
{
    std::call_once(g_dev_once, [] {
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        TERMINATE(e != hipSuccess || n < 1, "libcaps_amd needs a gfx950 GPU; none usable (%s)",
                  hipGetErrorString(e));
        hipDeviceProp_t prop;
        HIPCHECK(hipGetDeviceProperties(&prop, 0));
        TERMINATE(strncmp(prop.gcnArchName, "gfx950", 6) != 0, "libcaps_amd is built for gfx950, device is %s",
                  prop.gcnArchName);
        std::vector<double> lf(1024);
        for (int i = 0; i < 1024; i++) lf[i] = lgamma((double)i + 1.0);
        HIPCHECK(hipMalloc(&g_lfac_dev, 1024 * sizeof(double)));
        HIPCHECK(hipMemcpy(g_lfac_dev, lf.data(), 1024 * sizeof(double), hipMemcpyHostToDevice));
        HIPCHECK((hipError_t)capsd_lu_init());
    });
}

constexpr int MAX_LU_STREAMS = 4;
constexpr int MAX_PARTS = MAX_LU_STREAMS;   // parts of a batch (BatchPlan)

struct Ctx {
    hipStream_t stream = nullptr;
    hipEvent_t ev[7];
    // the LU of a batch is split into groups of blocks on these streams so
    // that one group's panel factorizations overlap another's updates
    hipStream_t lu_stream[MAX_LU_STREAMS];
    hipEvent_t lu_done[MAX_LU_STREAMS];
    hipEvent_t part_i_done[MAX_PARTS];   // part g's I finished
    hipEvent_t lu_ready = nullptr;        // a one-part batch's matrices assembled
    hipEvent_t head_fork = nullptr;       // a batch's start, for the Mie kernel's stream
    hipEvent_t mie_done = nullptr;        // the batch's Mie coefficients written
    char* arena = nullptr;
    size_t arena_size = 0;
    BlockDesc* d_desc = nullptr;
    LuSide* d_side = nullptr;
    BlockOut* d_out = nullptr;
    int* d_status = nullptr;
    IGroup* d_groups = nullptr;
    IGroup* h_groups = nullptr;
    int cap_blocks = 0;
    BlockDesc* h_desc = nullptr;  // pinned
    LuSide* h_side = nullptr;
    BlockOut* h_out = nullptr;
    int* h_status = nullptr;
    // staging buffer of the point APIs (capsx_*_points), grown on demand
    char* pts = nullptr;
    size_t pts_bytes = 0;
    double timing[6] = {0, 0, 0, 0, 0, 0};
    // accumulated since capsx_stats_reset: stage ms [0..5], blocks, blocks
    // factored (not skipped), algorithmic LU flops, batches, I pairs, and
    // with profiling on the trailing update's ms, flops, launches, bytes
    double acc[16] = {};
    // per-launch timing of the LU's trailing-update kernel (capsx_profile)
    int prof = 0;
    // captured launch sequences of recent batch shapes (capsx_set_graph)
    int graph = 0;
    std::vector<std::pair<std::vector<char>, hipGraphExec_t>> graphs;
    std::vector<hipEvent_t> gemm_ev;
    std::vector<int> gemm_info;
    // planner surrogate (capsx_set_surrogate) and the material features it
    // is evaluated with
    struct capsx_mlp* sur = nullptr;
    float sur_feat[2] = {0, 0};
    // surrogate against the computed blocks of its plans since the last
    // capsx_surrogate_report reset: blocks, sum of squared and largest
    // |log error|, frequencies, sum and largest |m-cutoff error|
    double sur_acc[6] = {};
    long plan_batches = 0;
    std::mutex mu;
};

// A batch runs its LU in up to MAX_LU_STREAMS stream groups, the first on
// the library's stream.  With HIP's default of 4 hardware queues per process
// two of those streams can land on one queue and then run one after the
// other (mixed-aspect batch of 64 blocks in a process's first context: LU
// 16.4 ms with 4 queues, 11.8 ms with 8).  Unless the caller chose a number,
// ask for 8 before the HIP runtime starts (it reads the variable once, at
// its first call).
__attribute__((constructor)) void request_hw_queues() { setenv("GPU_MAX_HW_QUEUES", "8", 0); }

Ctx* ctx_of(caps_t* self)
{
    device_init();
    if (!self->priv) {
        Ctx* c = new Ctx();
        HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        for (auto& e : c->ev) HIPCHECK(hipEventCreate(&e));
        for (int g = 0; g < MAX_LU_STREAMS; g++) {
            HIPCHECK(hipStreamCreateWithFlags(&c->lu_stream[g], hipStreamNonBlocking));
            HIPCHECK(hipEventCreateWithFlags(&c->lu_done[g], hipEventDisableTiming));
        }
        for (auto& e : c->part_i_done) HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        HIPCHECK(hipEventCreateWithFlags(&c->lu_ready, hipEventDisableTiming));
        HIPCHECK(hipEventCreateWithFlags(&c->head_fork, hipEventDisableTiming));
        HIPCHECK(hipEventCreateWithFlags(&c->mie_done, hipEventDisableTiming));
        HIPCHECK(hipMalloc(&c->d_status, sizeof(int)));
        HIPCHECK(hipHostMalloc((void**)&c->h_status, sizeof(int), hipHostMallocDefault));
        self->priv = c;
    }
    return (Ctx*)self->priv;
}

void ctx_free(Ctx* c)
{
    if (!c) return;
    hipStreamSynchronize(c->stream);
    if (c->arena) hipFree(c->arena);
    if (c->d_desc) hipFree(c->d_desc);
    if (c->d_side) hipFree(c->d_side);
    if (c->d_out) hipFree(c->d_out);
    if (c->d_status) hipFree(c->d_status);
    if (c->h_desc) hipHostFree(c->h_desc);
    if (c->h_side) hipHostFree(c->h_side);
    if (c->h_out) hipHostFree(c->h_out);
    if (c->h_status) hipHostFree(c->h_status);
    if (c->d_groups) hipFree(c->d_groups);
    if (c->h_groups) hipHostFree(c->h_groups);
    if (c->pts) hipFree(c->pts);
    for (auto& e : c->ev) hipEventDestroy(e);
    for (auto& e : c->gemm_ev) hipEventDestroy(e);
    for (auto& g : c->graphs) hipGraphExecDestroy(g.second);
    for (int g = 0; g < MAX_LU_STREAMS; g++) {
        hipStreamDestroy(c->lu_stream[g]);
        hipEventDestroy(c->lu_done[g]);
    }
    for (auto& e : c->part_i_done) hipEventDestroy(e);
    hipEventDestroy(c->lu_ready);
    hipEventDestroy(c->head_fork);
    hipEventDestroy(c->mie_done);
    hipStreamDestroy(c->stream);
    delete c;
}

void ensure_blocks(Ctx* c, int n)
{
    if (n <= c->cap_blocks) return;
    const int cap = std::max(n, 2 * c->cap_blocks);
    if (c->d_desc) {
        HIPCHECK(hipStreamSynchronize(c->stream));
        hipFree(c->d_desc);
        hipFree(c->d_side);
        hipFree(c->d_out);
        hipHostFree(c->h_desc);
        hipHostFree(c->h_side);
        hipHostFree(c->h_out);
        hipFree(c->d_groups);
        hipHostFree(c->h_groups);
    }
    HIPCHECK(hipMalloc(&c->d_desc, cap * sizeof(BlockDesc)));
    HIPCHECK(hipMalloc(&c->d_groups, cap * sizeof(IGroup)));
    HIPCHECK(hipHostMalloc((void**)&c->h_groups, cap * sizeof(IGroup), hipHostMallocDefault));
    HIPCHECK(hipMalloc(&c->d_side, cap * sizeof(LuSide)));
    HIPCHECK(hipMalloc(&c->d_out, cap * sizeof(BlockOut)));
    HIPCHECK(hipHostMalloc((void**)&c->h_desc, cap * sizeof(BlockDesc), hipHostMallocDefault));
    HIPCHECK(hipHostMalloc((void**)&c->h_side, cap * sizeof(LuSide), hipHostMallocDefault));
    HIPCHECK(hipHostMalloc((void**)&c->h_out, cap * sizeof(BlockOut), hipHostMallocDefault));
    c->cap_blocks = cap;
}

double env_double(const char* name, double dflt);

// HBM budget of one launch sequence (CAPS_AMD_ARENA_GB, default 160 GB):
// a batch's blocks go into sequences that fit it.
size_t arena_budget() { return (size_t)(env_double("CAPS_AMD_ARENA_GB", 160.0) * 1e9); }

// The device arena of a batch, grown to at least `bytes`: by a quarter more
// than before, so that slowly growing batches do not reallocate each time,
// but not beyond the budget unless `bytes` itself is more.
void ensure_arena(Ctx* c, size_t bytes)
{
    if (bytes <= c->arena_size) return;
    HIPCHECK(hipStreamSynchronize(c->stream));
    if (c->arena) HIPCHECK(hipFree(c->arena));
    c->arena = nullptr;
    const size_t sz = std::max(bytes, std::min((size_t)(c->arena_size * 1.25), arena_budget()));
    if (hipMalloc(&c->arena, sz) != hipSuccess) {
        size_t fr = 0, tot = 0;
        (void)hipGetLastError();
        (void)hipMemGetInfo(&fr, &tot);
        TERMINATE(true, "cannot allocate a batch arena of %.3f GB (previous %.3f GB; device memory free %.3f of %.3f GB)",
                  sz / 1e9, c->arena_size / 1e9, fr / 1e9, tot / 1e9);
    }
    if (env_double("CAPS_AMD_PROGRESS", 0) > 0) {
        size_t fr = 0, tot = 0;
        (void)hipMemGetInfo(&fr, &tot);
        fprintf(stderr, "capsx: arena %.3f GB (needed %.3f GB); device memory free %.3f of %.3f GB\n", sz / 1e9,
                bytes / 1e9, fr / 1e9, tot / 1e9);
    }
    c->arena_size = sz;
}

inline int64_t align256(int64_t x) { return (x + 255) & ~(int64_t)255; }

// The point APIs' device staging buffer, at least `bytes` long.
char* ensure_pts(Ctx* c, size_t bytes)
{
    if (bytes > c->pts_bytes) {
        HIPCHECK(hipStreamSynchronize(c->stream));
        if (c->pts) HIPCHECK(hipFree(c->pts));
        c->pts = nullptr;
        const size_t sz = std::max(bytes, (size_t)(c->pts_bytes * 1.5));
        HIPCHECK(hipMalloc(&c->pts, sz));
        c->pts_bytes = sz;
    }
    return c->pts;
}

// LU side buffers of one N x N block (leading dimension lda) from arena
// offset off: the factored panels of a step, log|u_ii|, the pivots of each
// panel and the step's U rows in row-major order (capsd_lu_npanels() * NB
// rows of lda doubles, the trailing updates' B operand); returns the offset
// after them.
int64_t lu_side_layout(LuSide& s, int64_t off, int N, int lda)
{
    const int nb = capsd_lu_nb();
    s.off_panel = off;
    off = align256(off + capsd_lu_npanels() * (int64_t)lda * nb * 8);
    s.off_logs = off;
    off = align256(off + (int64_t)N * 8);
    s.off_moved = off;
    off = align256(off + capsd_lu_npanels() * (int64_t)nb * 4);
    s.off_ut = off;
    off = align256(off + capsd_lu_npanels() * (int64_t)nb * lda * 8);
    s.off_inv = off;   // delta form: L11^{-1} and U11^{-1} of each panel
    return align256(off + capsd_lu_npanels() * 2 * (int64_t)nb * nb * 8);
}

double env_double(const char* name, double dflt)
{
    const char* s = getenv(name);
    return s ? atof(s) : dflt;
}

// ---- LU schedule configuration ---------------------------------------------
// levels: 1 = 64-column steps, 2 = 128-column steps of two panels with one
// K = 128 trailing update, 3 = 256-column steps of four panels with one
// K = 256 trailing update, 4 = 512-column steps of eight panels with one
// K = 512 trailing update; streams: LU groups of blocks on separate streams;
// xcd: trailing-update tile order (1 = all tiles of a block on one XCD);
// tile: trailing-update workgroup tile, 128 (x 128), 256 (x 128) or 64 (x 64).
struct LuCfg {
    int levels = 2, streams = 1, xcd = 0, tile = 128;
};
struct TuneEntry {
    int maxN, n;
    LuCfg cfg;
};
// ---- resume cache and per-frequency output --------------------------------
// The reference driver reads the lines "# xi*(L+R)/c=<ξ_>, logdetD=<F_ξ>, ..."
// of an earlier run's output (src/caps.c:72-101) and returns F_ξ from them
// for a frequency that matches (ξ_ equal, or |1 - ξ_cached/ξ_| < 1e-11,
// src/caps.c:316-327); it prints such a line for every F_ξ it computes
// (src/caps.c:265).  Per caps object, host only.
struct ResumeState {
    std::vector<std::pair<double, double>> cache;   // (ξ_, F_ξ)
    FILE* out = nullptr;
    double t0 = 0;
    // checkpointing of the T = 0 driver (capsx_set_checkpoint): at most
    // `nodes` quadrature nodes per batch (0: a level's new nodes at once),
    // stop before a batch once `seconds` have passed (0: no limit)
    int nodes = 0;
    double seconds = 0;
};
std::mutex g_resume_mu;
std::unordered_map<const caps_t*, ResumeState> g_resume;

double wall_seconds()
{
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

bool resume_lookup(const caps_t* self, double xi, double* v)
{
    std::lock_guard<std::mutex> lock(g_resume_mu);
    auto it = g_resume.find(self);
    if (it == g_resume.end()) return false;
    for (const auto& e : it->second.cache)
        if (xi == e.first || fabs(1 - e.first / xi) < 1e-11) {
            *v = e.second;
            return true;
        }
    return false;
}

void checkpoint_get(const caps_t* self, int* nodes, double* seconds)
{
    std::lock_guard<std::mutex> lock(g_resume_mu);
    auto it = g_resume.find(self);
    *nodes = it == g_resume.end() ? 0 : it->second.nodes;
    *seconds = it == g_resume.end() ? 0 : it->second.seconds;
}

void resume_record(const caps_t* self, double xi, double v)
{
    std::lock_guard<std::mutex> lock(g_resume_mu);
    auto it = g_resume.find(self);
    if (it == g_resume.end() || !it->second.out) return;
    if (xi == 0)
        fprintf(it->second.out, "# xi*(L+R)/c=0, logdetD=%.16g, t=%g\n", v, wall_seconds() - it->second.t0);
    else
        fprintf(it->second.out, "# xi*(L+R)/c=%.16g, logdetD=%.16g, t=%g\n", xi, v, wall_seconds() - it->second.t0);
    fflush(it->second.out);
}

std::mutex g_cfg_mu;
LuCfg g_override;                 // used when g_have_override
bool g_have_override = false;
std::vector<TuneEntry> g_tune;    // autotuned table (lu_tuning.txt)
// 1: the delta form for every column dominant block (default); 0: the
// pivoted LU for all blocks (capsx_lu_delta_mode, a test hook)
int g_delta_mode = 1;
std::string g_tune_path;

// Table lines: "maxN n levels streams xcd tile" ('#' starts a comment).
// A missing file empties the table (the defaults then apply).
int load_tuning(const char* path)
{
    FILE* f = fopen(path, "r");
    if (!f) {
        std::lock_guard<std::mutex> lock(g_cfg_mu);
        g_tune.clear();
        g_tune_path = path;
        return 0;
    }
    std::vector<TuneEntry> t;
    char line[512];
    while (fgets(line, sizeof line, f)) {
        if (line[0] == '#') continue;
        TuneEntry e;
        if (sscanf(line, "%d %d %d %d %d %d", &e.maxN, &e.n, &e.cfg.levels, &e.cfg.streams, &e.cfg.xcd,
                   &e.cfg.tile) == 6 &&
            e.maxN > 0 && e.n > 0 && e.cfg.levels >= 1 && e.cfg.levels <= 4 && e.cfg.streams >= 1 &&
            (e.cfg.tile == 64 || e.cfg.tile == 128 || e.cfg.tile == 256))
            t.push_back(e);
    }
    fclose(f);
    std::lock_guard<std::mutex> lock(g_cfg_mu);
    g_tune = t;
    g_tune_path = path;
    return (int)t.size();
}

// The table next to the library (or CAPS_AMD_TUNING), read once.
void load_default_tuning()
{
    static std::once_flag once;
    std::call_once(once, [] {
        const char* env = getenv("CAPS_AMD_TUNING");
        if (env) {
            load_tuning(env);
            return;
        }
        Dl_info info;
        if (dladdr((void*)&load_tuning, &info) && info.dli_fname) {
            std::string p(info.dli_fname);
            const size_t slash = p.rfind('/');
            p = (slash == std::string::npos ? std::string(".") : p.substr(0, slash)) + "/lu_tuning.txt";
            load_tuning(p.c_str());
        }
    });
}

// Configuration for a batch of n blocks of largest dimension maxN: the
// override if set, else environment variables (CAPS_AMD_LU_LEVELS,
// CAPS_AMD_LU_STREAMS, CAPS_AMD_GEMM_XCD, CAPS_AMD_GEMM_TILE) over the
// nearest autotuned shape in (log maxN, log n), else the defaults.
LuCfg lu_config(int maxN, int n)
{
    load_default_tuning();
    std::lock_guard<std::mutex> lock(g_cfg_mu);
    if (g_have_override) return g_override;
    LuCfg c;
    double best = INFINITY;
    for (const TuneEntry& e : g_tune) {
        const double dx = log((double)maxN / e.maxN), dy = 0.5 * log((double)std::max(n, 1) / e.n);
        const double d2 = dx * dx + dy * dy;
        if (d2 < best) {
            best = d2;
            c = e.cfg;
        }
    }
    if (getenv("CAPS_AMD_LU_LEVELS")) c.levels = std::max(1, std::min(4, (int)env_double("CAPS_AMD_LU_LEVELS", 2)));
    if (getenv("CAPS_AMD_LU_STREAMS")) c.streams = std::max(1, (int)env_double("CAPS_AMD_LU_STREAMS", 1));
    if (getenv("CAPS_AMD_GEMM_XCD")) c.xcd = env_double("CAPS_AMD_GEMM_XCD", 0) != 0 ? 1 : 0;
    if (getenv("CAPS_AMD_GEMM_TILE")) {
        const double t = env_double("CAPS_AMD_GEMM_TILE", 128);
        c.tile = t <= 64 ? 64 : t >= 256 ? 256 : 128;
    }
    return c;
}

// Quadrature grid of the K integrals of one block (see k_K_partial):
// x(v) = 1 + A² softplus(v)², A = 0.5/sqrt(α), step h, from v0 to the end of
// the tail of the largest ν (integrand below e^{-70} of its peak).
void grid_for(BlockDesc& d)
{
    const double alpha = d.alpha;
    const double h = env_double("CAPS_AMD_QUAD_H", 0.45);
    const double v0 = env_double("CAPS_AMD_QUAD_V0", -19.0);
    const double A = 0.5 / sqrt(alpha);
    const double nu = d.nu_max + 0.5;
    const double xp = sqrt(1 + (nu / alpha) * (nu / alpha));
    auto g = [&](double x) { return -alpha * x + nu * acosh(x); };
    const double gp = g(xp);
    double x = xp;
    int guard = 0;
    while (g(x) > gp - 70 && guard++ < 100000) x = 1 + (x - 1) * 1.05 + 1.0 / alpha;
    x = std::max(x, 1 + 80.0 / alpha);
    const double s = sqrt(x - 1) / A;
    const double vhi = s > 30 ? s : log(expm1(s));
    d.v0 = v0;
    d.h = h;
    d.A = A;
    d.n_nodes = (int)ceil((vhi - v0) / h) + 1;
    d.n_waves = (d.n_nodes + 63) / 64;
}

// A batch runs as `parts` parts, contiguous block ranges [part_b[g],
// part_b[g+1]) (the LU stream groups): part g's I, assembly and LU run in
// order on stream g, and part g's I starts when part g-1's I is done, so
// the I of one part runs beside the assembly and LU of the part before it
// (measured on the C2 batch: 2 parts 0.295 s per step, against 0.302 s
// with the parts' I run together; one stream computing every part's I
// while a second factored the parts in turn was slower, 0.31-0.44 s for
// 2-8 parts).
struct BatchPlan {
    int n = 0, maxN = 0, minN = 0, max_ldim = 0, max_waves = 0, max_nK = 0, max_nl = 0;
    int sum_waves = 0;   // K quadrature waves over the batch's blocks
    int64_t max_ntri = 0;
    size_t bytes = 0;
    std::vector<IGroup> groups;   // per part, ordered by class (see build_igroups)
    int parts = 1;
    int part_b[MAX_PARTS + 1] = {};
    int part_goff[MAX_PARTS] = {};      // first I group of each part
    int part_cls[MAX_PARTS][8] = {};    // I groups per class of each part
    // determinant forms the LU launches kernels for (capsd_launch_lu kinds):
    // 2 = delta form only (round-trip blocks; one that turns out not column
    // dominant is left unfactored and run again with 1 = pivoted), 6 = the
    // same in the symmetric form (1 - M of a round trip is symmetric: the
    // reference factors it as such, sym_spd in src/libcaps.c:954)
    int lu_kinds = 2;
};

// Groups of blocks [b0, b1) that share the I coefficients a_q: same caps
// object, m, ℓ-window and polarizations, at most IG_MAX per group; appended
// to p.groups by class, the class sizes in cls.
void build_igroups(caps_t* const* objs, const BlockDesc* desc, int b0, int b1, BatchPlan& p, int* cls_count)
{
    const int n = b1 - b0;
    std::vector<int> idx(n);
    for (int i = 0; i < n; i++) idx[i] = b0 + i;
    auto key = [&](int i) {
        const BlockDesc& d = desc[i];
        return std::make_tuple(1 - d.is_pr, (uintptr_t)objs[i], d.m, d.lmin, d.ldim);
    };
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return key(x) < key(y); });
    // classes: perfect reflector (TM only) or both polarizations, times at
    // most 1, 2, 4 or 8 blocks; the kernel is instantiated per class
    std::vector<IGroup> cls[8];
    // at most 8 blocks per group (k_I keeps a chunk of a_q and every sum in
    // registers: 125 VGPRs for 8 blocks with both polarizations, 4 waves
    // per SIMD); a run of r blocks with one key is cut into ceil(r/8) groups
    // of near-equal size, consecutive in the input order (in a free energy
    // neighbouring frequencies, whose sums take similar numbers of terms).
    // CAPS_AMD_IGROUP_MAX lowers the maximum.
    static const int genv = (int)env_double("CAPS_AMD_IGROUP_MAX", 0);
    const int gmax = std::max(1, std::min(IG_MAX, genv > 0 ? genv : IG_MAX));
    for (int i = 0; i < n;) {
        int r = 1;
        while (i + r < n && key(idx[i + r]) == key(idx[i])) r++;
        const int ng = (r + gmax - 1) / gmax;
        for (int k = 0, at = i; k < ng; k++) {
            IGroup g;
            memset(&g, 0, sizeof(g));
            g.count = r / ng + (k < r % ng ? 1 : 0);
            for (int t = 0; t < g.count; t++) g.blk[t] = idx[at + t];
            at += g.count;
            const int sc = g.count <= 1 ? 0 : g.count <= 2 ? 1 : g.count <= 4 ? 2 : 3;
            cls[(desc[idx[i]].is_pr ? 0 : 4) + sc].push_back(g);
        }
        i += r;
    }
    for (int k = 0; k < 8; k++) {
        cls_count[k] = (int)cls[k].size();
        p.groups.insert(p.groups.end(), cls[k].begin(), cls[k].end());
    }
}

// Fill descriptors for blocks [0, n) and their arena offsets.
BatchPlan plan_batch_multi(caps_t* const* objs, Ctx* c, const double* xi_, const int* m, int n, int with_lu)
{
    BatchPlan p;
    p.n = n;
    // CAPS_AMD_LU_SYM=0: the general (non-symmetric) delta form
    p.lu_kinds = env_double("CAPS_AMD_LU_SYM", 1) != 0 ? 6 : 2;
    const int nb = capsd_lu_nb();
    int64_t off = 0;
    double last_xi = NAN, ep = 0, es = 0;
    caps_t* last_obj = nullptr;
    for (int b = 0; b < n; b++) {
        caps_t* self = objs[b];
        const int ldim = self->ldim;
        const int N = 2 * ldim;
        const int lda = (N + 15) & ~15;
        BlockDesc& d = c->h_desc[b];
        memset(&d, 0, sizeof(d));
        d.xi_ = xi_[b];
        d.alpha = 2 * xi_[b];
        if (!(xi_[b] == last_xi) || self != last_obj) {
            ep = caps_epsilonm1_plate(self, xi_[b]);
            es = caps_epsilonm1_sphere(self, xi_[b]);
            last_xi = xi_[b];
            last_obj = self;
        }
        d.epsm1_plate = ep;
        d.epsm1_sphere = es;
        d.LbyR = self->LbyR;
        d.m = m[b];
        size_t lmin, lmax;
        caps_estimate_lminmax(self, m[b], &lmin, &lmax);
        d.lmin = (int)lmin;
        d.ldim = ldim;
        d.N = N;
        d.lda = lda;
        d.is_pr = std::isinf(caps_epsilonm1_plate(self, INFINITY)) ? 1 : 0;
        d.mu = m[b] > 0 ? 2 * m[b] : 2;
        d.lo = d.lmin - 1;
        d.nl = ldim + 2;
        d.nu_max = 2 * (d.lo + d.nl - 1);
        d.nK = d.nu_max - d.mu + 1;
        grid_for(d);
        d.off_mie = off;
        off = align256(off + (5 * (int64_t)ldim + 3 * (int64_t)(ldim + 2)) * 8);
        d.off_kpart = off;
        off = align256(off + (int64_t)d.n_waves * d.nK * sizeof(KPart));
        // K table after a zeroed guard of capsd::K_GUARD entries (k_K_combine):
        // the I recurrence's chunks read up to 2 (chunk - 1) entries below
        // index 0 for steps past qmax, whose coefficients are 0
        d.off_k = off + 256;
        off = align256(d.off_k + 2 * (int64_t)d.nK * (sizeof(capsd::KEnt) + 8));
        d.off_i = off;
        off = align256(off + 2 * (int64_t)d.nl * d.nl * 12);
        d.off_mat = off;
        off = align256(off + (int64_t)lda * N * 8);
        d.off_diag = off;
        off = align256(off + (int64_t)N * 8);
        d.off_piv = off;
        off = align256(off + (int64_t)N * 4);
        // column sums of |M| per 32-row tile of the assembly (k_assemble),
        // for the column dominance test of the LU's delta form
        d.n_rtiles = with_lu ? (ldim + 15) / 16 : 0;
        d.off_cpart = off;
        off = align256(off + (int64_t)d.n_rtiles * N * 8);
        LuSide& s = c->h_side[b];
        if (with_lu) {
            off = lu_side_layout(s, off, N, lda);
        } else
            s.off_panel = s.off_logs = s.off_moved = s.off_ut = s.off_inv = 0;
        p.maxN = std::max(p.maxN, N);
        p.minN = b == 0 ? N : std::min(p.minN, N);
        p.max_ldim = std::max(p.max_ldim, ldim);
        p.max_waves = std::max(p.max_waves, d.n_waves);
        p.sum_waves += d.n_waves;
        p.max_nK = std::max(p.max_nK, d.nK);
        p.max_ntri = std::max(p.max_ntri, (int64_t)d.nl * (d.nl + 1) / 2);
        p.max_nl = std::max(p.max_nl, d.nl);
    }
    p.bytes = (size_t)off;
    // parts = LU stream groups of the tuned schedule, at most
    // CAPS_AMD_PARTS (default 1; one without an LU, and in profiled runs,
    // which time each update launch alone).  Since the I, K and assembly
    // kernels and the LU's own kernels got faster, running the I of one
    // part beside the LU of another no longer pays: C2 step 164.5 ms with
    // one part, 165.4 ms with two, 165.9 ms with three (the kernels slow
    // each other down as much as they overlap)
    p.parts = 1;
    if (with_lu && !c->prof) {
        const int cap = std::max(1, (int)env_double("CAPS_AMD_PARTS", 1));
        p.parts = std::max(1, std::min(std::min(lu_config(p.maxN, n).streams, cap), std::min(MAX_PARTS, n / 4)));
    }
    p.groups.clear();
    // part 0 takes the fraction f0 of the blocks (its I runs alone, before
    // any LU can start), the others share the rest equally
    // (CAPS_AMD_PART0_FRAC; default: equal parts)
    const double f0 = p.parts > 1 ? env_double("CAPS_AMD_PART0_FRAC", 1.0 / p.parts) : 1.0;
    p.part_b[0] = 0;
    p.part_b[1] = std::max(1, std::min(n - (p.parts - 1), (int)std::lround(n * f0)));
    for (int g = 2; g <= p.parts; g++)
        p.part_b[g] = p.part_b[1] + (int)((int64_t)(n - p.part_b[1]) * (g - 1) / (p.parts - 1));
    for (int g = 0; g < p.parts; g++) {
        p.part_goff[g] = (int)p.groups.size();
        build_igroups(objs, c->h_desc, p.part_b[g], p.part_b[g + 1], p, p.part_cls[g]);
    }
    return p;
}

BatchPlan plan_batch(caps_t* self, Ctx* c, const double* xi_, const int* m, int n, int with_lu)
{
    std::vector<caps_t*> objs(n, self);
    return plan_batch_multi(objs.data(), c, xi_, m, n, with_lu);
}

size_t block_bytes_estimate(caps_t* self, int m)
{
    const int ldim = self->ldim, N = 2 * ldim;
    const size_t lda = (N + 15) & ~15;
    size_t lmin, lmax;
    caps_estimate_lminmax(self, m, &lmin, &lmax);
    const size_t nK = 2 * (lmin + ldim) + 4;
    return (size_t)2 * (ldim + 2) * (ldim + 2) * 12 + 256 + lda * N * 8 + lda * (size_t)(2 * capsd_lu_npanels() * capsd_lu_nb() + 2 * capsd_lu_nb()) * 8 + nK * (24 * 2 + 24 * 64) +
           (size_t)N * 32 + (size_t)((ldim + 15) / 16) * N * 8 + 65536;
}

// The LU of blocks [0, n) (descriptors already on the device) after event
// `ready` on the library's stream, in cfg.streams groups of blocks on their
// own streams; the library's stream waits for all groups.  With prof, HIP
// events around every trailing-update launch (one group).  kinds: the
// determinant forms the blocks may take (capsd_launch_lu).  Returns the
// number of update launches of the first group.
// The LU's stream groups of blocks [0, n) (descriptors in c->h_desc):
// contiguous runs of about equal LU cost (N^3; larger blocks come first in a
// batch, so a dominant block may run alone), bnd[g] .. bnd[g+1], each up to
// its own largest N, gN[g] (at most maxN).  Returns the number of groups.
int lu_groups(const Ctx* c, int n, int maxN, int streams, int bnd[MAX_LU_STREAMS + 1], int gN[MAX_LU_STREAMS])
{
    const int groups = std::max(1, std::min(streams, std::min(MAX_LU_STREAMS, n / 4)));
    std::vector<double> cum(n + 1, 0.0);
    for (int b = 0; b < n; b++) cum[b + 1] = cum[b] + std::pow((double)c->h_desc[b].N, 3);
    bnd[0] = 0;
    for (int g = 1; g < groups; g++) {
        int b = bnd[g - 1] + 1;
        while (b < n - (groups - g) && cum[b] < cum[n] * g / groups) b++;
        bnd[g] = b;
    }
    bnd[groups] = n;
    for (int g = 0; g < groups; g++) {
        gN[g] = 0;
        for (int b = bnd[g]; b < bnd[g + 1]; b++) gN[g] = std::max(gN[g], c->h_desc[b].N);
        gN[g] = std::min(gN[g], maxN);
    }
    return groups;
}

int launch_lu_groups(Ctx* c, int n, int maxN, const LuCfg& cfg, hipEvent_t ready, bool prof, int kinds)
{
    int bnd[MAX_LU_STREAMS + 1], gN[MAX_LU_STREAMS];
    const int groups = lu_groups(c, n, maxN, cfg.streams, bnd, gN);
    if (groups == 1)
        return capsd_launch_lu(c->d_desc, c->arena, c->d_side, c->d_out, n, maxN, c->stream, cfg.levels, cfg.xcd,
                               cfg.tile, prof ? c->gemm_ev.data() : nullptr, prof ? c->gemm_info.data() : nullptr,
                               kinds);
    // group 0 runs on the library's stream itself, groups 1.. on their own
    // streams, forked at `ready` and joined back: HIP put two fresh streams
    // on one hardware queue (the trace showed both groups serialised), while
    // the library's stream holds a queue of its own
    int launched = 0;
    for (int g = groups - 1; g >= 0; g--) {
        const int b0 = bnd[g], b1 = bnd[g + 1];
        hipStream_t sg = g == 0 ? c->stream : c->lu_stream[g - 1];
        if (g > 0) HIPCHECK(hipStreamWaitEvent(sg, ready, 0));
        const int k = capsd_launch_lu(c->d_desc + b0, c->arena, c->d_side + b0, c->d_out + b0, b1 - b0, gN[g], sg,
                                      cfg.levels, cfg.xcd, cfg.tile, nullptr, nullptr, kinds);
        if (g == 0) launched = k;
    }
    for (int g = 1; g < groups; g++) {
        HIPCHECK(hipEventRecord(c->lu_done[g], c->lu_stream[g - 1]));
        HIPCHECK(hipStreamWaitEvent(c->stream, c->lu_done[g], 0));
    }
    return launched;
}

// LU schedule of a planned one-part batch.  Blocks of unequal N run in
// MAX_LU_STREAMS groups of about equal cost, whose chains of dependent panel
// steps then overlap, each with the schedule of its share of n_eff =
// sum N^3 / maxN^3 blocks of the largest N (mixed-aspect batch of 64 blocks,
// N 614-5462: LU 15.9 ms in one group with the table's schedule for 64
// blocks, 12.4 ms this way).  One group when profiled: HIP events time the
// launches of one group.
LuCfg batch_lu_config(const Ctx* c, const BatchPlan& p, bool prof)
{
    LuCfg cfg = lu_config(p.maxN, p.n);
    bool fixed;
    {
        std::lock_guard<std::mutex> lock(g_cfg_mu);
        fixed = g_have_override;   // a schedule fixed by capsx_lu_config applies as it is
    }
    if (p.parts == 1 && p.minN < p.maxN && !fixed) {
        double w = 0;
        for (int b = 0; b < p.n; b++) w += std::pow((double)c->h_desc[b].N / p.maxN, 3);
        cfg = lu_config(p.maxN, std::max(1, (int)lround(w / MAX_LU_STREAMS)));
        cfg.streams = MAX_LU_STREAMS;
    }
    if (prof) cfg.streams = 1;
    return cfg;
}

// Segments of the K quadrature's ν loop per wave: a batch of few waves
// (fewer than two per SIMD: one block, or a few large ones) splits each
// wave's ν range over up to 8 workgroups, whose steps are otherwise one
// latency-bound chain per wave (CAPS_AMD_K_SEGMENTS fixes the number).
int k_segments(const BatchPlan& p)
{
    if (getenv("CAPS_AMD_K_SEGMENTS")) return std::max(1, std::min(8, (int)env_double("CAPS_AMD_K_SEGMENTS", 1)));
    const int target = 2048;
    if (p.sum_waves >= target || p.sum_waves <= 0) return 1;
    return std::min(8, (target + p.sum_waves - 1) / p.sum_waves);
}

// Run the pipeline for n blocks already planned; mode 0: LU path, 1: M only.
// The kernels of one planned batch on the library's stream (and the LU
// streams); stage events between the stages when `events`.  Returns the
// number of trailing-update launches of the first LU group.
int enqueue_batch(Ctx* c, const BatchPlan& p, int mode, int lu_mode_shortcuts, bool events)
{
    const int n = p.n;
    hipStream_t s = c->stream;
    if (events) HIPCHECK(hipEventRecord(c->ev[0], s));
    // the Mie coefficients beside the K quadrature (independent; the I
    // tables need both), on the LU stream that no LU group or part uses:
    // the K stage's time is then what it adds after the Mie kernel
    hipStream_t sm = c->lu_stream[MAX_LU_STREAMS - 1];
    HIPCHECK(hipEventRecord(c->head_fork, s));
    HIPCHECK(hipStreamWaitEvent(sm, c->head_fork, 0));
    capsd_launch_mie(c->d_desc, c->arena, g_lfac_dev, n, p.max_ldim, sm);
    if (events) HIPCHECK(hipEventRecord(c->ev[1], sm));
    HIPCHECK(hipEventRecord(c->mie_done, sm));
    capsd_launch_K(c->d_desc, c->arena, g_lfac_dev, n, p.max_waves, p.max_nK, k_segments(p), c->d_status, s);
    HIPCHECK(hipStreamWaitEvent(s, c->mie_done, 0));
    if (events) HIPCHECK(hipEventRecord(c->ev[2], s));
    int nlaunched = 0;
    // part g on stream g (part 0 on the library's stream); its I waits for
    // part g-1's I; stage events: I and assembly end with the last part's,
    // the LU with the join of all parts
    for (int g = 0; g < p.parts; g++) {
        hipStream_t sg = g == 0 ? s : c->lu_stream[g - 1];
        const int b0 = p.part_b[g], nb = p.part_b[g + 1] - b0;
        const bool last = g == p.parts - 1;
        if (g > 0) HIPCHECK(hipStreamWaitEvent(sg, c->part_i_done[g - 1], 0));
        const IGroup* gr = c->d_groups + p.part_goff[g];
        if (mode == 0) {
            // trace pass: the I pairs of the diagonal band, the diagonal of M,
            // the trace and its shortcut; the full I and matrix then only
            // for the blocks that need an LU
            capsd_launch_I(c->d_desc, c->arena, gr, p.part_cls[g], g_lfac_dev, p.max_nl, c->d_status, c->d_out, 1, 0,
                           sg);
            capsd_launch_assemble(c->d_desc + b0, c->arena, g_lfac_dev, nb, p.max_ldim, mode, c->d_out + b0, 1, sg);
            capsd_launch_trace(c->d_desc + b0, c->arena, c->d_out + b0, nb, 0, sg);
        }
        capsd_launch_I(c->d_desc, c->arena, gr, p.part_cls[g], g_lfac_dev, p.max_ntri, c->d_status, c->d_out,
                       0, mode == 0, sg);
        if (!last) HIPCHECK(hipEventRecord(c->part_i_done[g], sg));
        if (events && last) HIPCHECK(hipEventRecord(c->ev[3], sg));
        // the symmetric delta-form LU reads nothing above the diagonal outside
        // its 64 x 64 diagonal blocks: those mirror tiles are not assembled
        const bool sym_only = mode == 0 && !lu_mode_shortcuts && (p.lu_kinds & 4) && !(p.lu_kinds & 1);
        capsd_launch_assemble(c->d_desc + b0, c->arena, g_lfac_dev, nb, p.max_ldim, mode, c->d_out + b0,
                              (mode == 0 ? 2 : 0) | (sym_only ? 4 : 0), sg);
        if (mode == 0) capsd_launch_dominance(c->d_desc + b0, c->arena, c->d_out + b0, nb, p.maxN, g_delta_mode, sg);
        if (mode == 0 && lu_mode_shortcuts)
            capsd_launch_frobenius(c->d_desc + b0, c->arena, c->d_out + b0, nb, lu_mode_shortcuts == 2, sg);
        if (events && last) HIPCHECK(hipEventRecord(c->ev[4], sg));
        if (mode != 0) continue;
        const bool prof = c->prof && p.parts == 1;
        const LuCfg cfg = batch_lu_config(c, p, prof);
        if (p.parts == 1) {
            // one part: the LU in the schedule's stream groups
            HIPCHECK(hipEventRecord(c->lu_ready, s));
            nlaunched = launch_lu_groups(c, n, p.maxN, cfg, c->lu_ready, prof, p.lu_kinds);
            continue;
        }
        const int k = capsd_launch_lu(c->d_desc + b0, c->arena, c->d_side + b0, c->d_out + b0, nb, p.maxN, sg,
                                      cfg.levels, cfg.xcd, cfg.tile, prof ? c->gemm_ev.data() : nullptr,
                                      prof ? c->gemm_info.data() : nullptr, p.lu_kinds);
        if (g == 0) nlaunched = k;
    }
    for (int g = 1; g < p.parts; g++) {
        HIPCHECK(hipEventRecord(c->lu_done[g], c->lu_stream[g - 1]));
        HIPCHECK(hipStreamWaitEvent(s, c->lu_done[g], 0));
    }
    return nlaunched;
}

// Identity of a captured batch: everything the launch sequence depends on
// (grid shapes, group classes, LU schedule, buffers).  Descriptor contents
// are read by the kernels at run time and are not part of it.
std::vector<char> graph_key(Ctx* c, const BatchPlan& p, int mode, int lu_mode_shortcuts)
{
    const LuCfg cfg = batch_lu_config(c, p, c->prof && p.parts == 1);
    struct {
        int n, maxN, max_ldim, max_waves, max_nK, kseg, mode, shortcuts, lu_kinds, delta_mode;
        int64_t max_ntri;
        int parts, part_b[MAX_PARTS + 1], part_goff[MAX_PARTS], cls[MAX_PARTS][8];
        int levels, streams, xcd, tile;
        int lu_bnd[MAX_LU_STREAMS + 1], lu_gN[MAX_LU_STREAMS];   // the LU's stream groups (lu_groups)
        int lu_kernels[3], grid_panel_rows;   // LU kernel variants (capsx_lu_kernels, capsx_lu_grid_panel_rows)
        const void *arena, *desc, *side, *out, *groups;
    } k;
    memset(&k, 0, sizeof(k));
    k.n = p.n;
    k.maxN = p.maxN;
    k.max_ldim = p.max_ldim;
    k.max_waves = p.max_waves;
    k.kseg = k_segments(p);
    k.max_nK = p.max_nK;
    k.mode = mode;
    k.shortcuts = lu_mode_shortcuts;
    k.lu_kinds = p.lu_kinds;
    k.delta_mode = g_delta_mode;
    k.max_ntri = p.max_ntri;
    k.parts = p.parts;
    memcpy(k.part_b, p.part_b, sizeof(k.part_b));
    memcpy(k.part_goff, p.part_goff, sizeof(k.part_goff));
    memcpy(k.cls, p.part_cls, sizeof(k.cls));
    k.levels = cfg.levels;
    k.streams = cfg.streams;
    k.xcd = cfg.xcd;
    k.tile = cfg.tile;
    if (p.parts == 1 && mode == 0) lu_groups(c, p.n, p.maxN, cfg.streams, k.lu_bnd, k.lu_gN);
    capsd_lu_kernels(-1, -1, -1, k.lu_kernels);
    k.grid_panel_rows = capsd_lu_grid_panel_rows(0);
    k.arena = c->arena;
    k.desc = c->d_desc;
    k.side = c->d_side;
    k.out = c->d_out;
    k.groups = c->d_groups;
    const char* b = reinterpret_cast<const char*>(&k);
    return std::vector<char>(b, b + sizeof(k));
}

// Run the pipeline for n blocks already planned; mode 0: LU path, 1: M only.
// With graph mode on, each batch shape is captured once into a HIP graph and
// replayed; descriptor uploads and the status reset stay outside the graph.
void run_batch(caps_t* self, Ctx* c, const BatchPlan& p, int mode, int lu_mode_shortcuts)
{
    const int n = p.n;
    hipStream_t s = c->stream;
    ensure_arena(c, p.bytes);
    HIPCHECK(hipMemcpyAsync(c->d_desc, c->h_desc, n * sizeof(BlockDesc), hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(c->d_side, c->h_side, n * sizeof(LuSide), hipMemcpyHostToDevice, s));
    const int ng = (int)p.groups.size();
    memcpy(c->h_groups, p.groups.data(), ng * sizeof(IGroup));
    HIPCHECK(hipMemcpyAsync(c->d_groups, c->h_groups, ng * sizeof(IGroup), hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemsetAsync(c->d_status, 0, sizeof(int), s));
    const int nupd = capsd_lu_max_updates(p.maxN);
    if (c->prof && (int)c->gemm_ev.size() < 2 * nupd) {
        const size_t old = c->gemm_ev.size();
        c->gemm_ev.resize(2 * nupd);
        c->gemm_info.resize(7 * nupd);
        for (size_t i = old; i < c->gemm_ev.size(); i++) HIPCHECK(hipEventCreate(&c->gemm_ev[i]));
    }
    int nlaunched = 0;
    const bool use_graph = c->graph && !c->prof;
    if (use_graph) {
        const std::vector<char> key = graph_key(c, p, mode, lu_mode_shortcuts);
        hipGraphExec_t exec = nullptr;
        for (auto& g : c->graphs)
            if (g.first == key) exec = g.second;
        if (!exec) {
            hipGraph_t graph;
            HIPCHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            enqueue_batch(c, p, mode, lu_mode_shortcuts, false);
            HIPCHECK(hipStreamEndCapture(s, &graph));
            HIPCHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            HIPCHECK(hipGraphDestroy(graph));
            if (c->graphs.size() >= 8) {
                HIPCHECK(hipGraphExecDestroy(c->graphs.front().second));
                c->graphs.erase(c->graphs.begin());
            }
            c->graphs.push_back({key, exec});
        }
        HIPCHECK(hipEventRecord(c->ev[0], s));
        HIPCHECK(hipGraphLaunch(exec, s));
    } else {
        nlaunched = enqueue_batch(c, p, mode, lu_mode_shortcuts, true);
    }
    HIPCHECK(hipEventRecord(c->ev[5], s));
    HIPCHECK(hipGetLastError());
    if (mode == 0) HIPCHECK(hipMemcpyAsync(c->h_out, c->d_out, n * sizeof(BlockOut), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipMemcpyAsync(c->h_status, c->d_status, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    float ms;
    if (use_graph) {
        // a replayed graph is timed as a whole (no stage events inside)
        for (int i = 0; i < 5; i++) c->timing[i] = 0;
    } else {
        for (int i = 0; i < 5; i++) {
            HIPCHECK(hipEventElapsedTime(&ms, c->ev[i], c->ev[i + 1]));
            c->timing[i] = ms;
        }
    }
    HIPCHECK(hipEventElapsedTime(&ms, c->ev[0], c->ev[5]));
    c->timing[5] = ms;
    TERMINATE(*c->h_status != 0, "non-finite k-integral (flags %d) in a batch of %d blocks (xi_=%g, m=%d ...)",
              *c->h_status, n, c->h_desc[0].xi_, c->h_desc[0].m);
    for (int i = 0; i < 6; i++) c->acc[i] += c->timing[i];
    if (c->prof && mode == 0) {
        // algorithmic flops of each k_lu_gemm_dma launch: 2 x rows x columns
        // x depth; the time of the other update launches (k_lu_gemm tiles)
        for (int pi = 0; pi < nlaunched; pi++) {
            const int* g = &c->gemm_info[7 * pi];   // k, r_off, r_lim, c_off, c_lim, KK, dma
            HIPCHECK(hipEventElapsedTime(&ms, c->gemm_ev[2 * pi], c->gemm_ev[2 * pi + 1]));
            if (!g[6]) {
                c->acc[15] += ms;
                continue;
            }
            c->acc[11] += ms;
            c->acc[13] += 1;
            const bool sym = p.lu_kinds & 4;
            for (int b = 0; b < n; b++) {
                if (c->h_out[b].skip) continue;
                const double N = c->h_desc[b].N;
                const double r0 = g[0] + g[1], c0 = g[0] + g[3];
                // region rows x cols, depth kk: flops 2 x elements x kk; HBM
                // bytes: C read and written once, L and U read once.  The
                // symmetric form needs the elements on and below the
                // diagonal only (row >= column), and its U rows are L rows
                // (U12 = D11 L21^T): the operands are the L rows of the union
                // of the region's row and column ranges, read once
                const double rows = std::max(0.0, std::min(N, r0 + g[2]) - r0);
                const double cols = std::max(0.0, std::min(N, c0 + g[4]) - c0);
                const double kk = rows > 0 && cols > 0 ? g[5] : 0;
                double elems = rows * cols;
                if (sym && rows > 0 && cols > 0) {
                    // columns j in [c0, c0 + cols): rows max(r0, j) .. e - 1
                    const double e = r0 + rows, c1 = c0 + cols - 1;
                    const double ja = std::min(c1, r0);   // j <= r0: all rows
                    elems = ja >= c0 ? (ja - c0 + 1) * rows : 0;
                    const double jb0 = std::max(c0, r0 + 1), jb1 = std::min(c1, e - 1);
                    if (jb1 >= jb0) elems += (jb1 - jb0 + 1) * (e - (jb0 + jb1) / 2);
                }
                double operand_rows = rows + cols;
                if (sym && rows > 0 && cols > 0) {
                    const double lo = std::max(r0, c0), hi = std::min(r0 + rows, c0 + cols);
                    if (hi > lo) operand_rows -= hi - lo;   // rows in both ranges are read once
                }
                c->acc[12] += 2.0 * elems * kk;
                c->acc[14] += 16.0 * elems + 8.0 * operand_rows * kk;
            }
        }
    }
    c->acc[6] += n;
    c->acc[9] += 1;
    for (int b = 0; b < n; b++) {
        const BlockDesc& d = c->h_desc[b];
        c->acc[10] += (double)d.nl * (d.nl + 1) / 2 * (d.is_pr ? 1 : 2);
        if (mode == 0 && c->h_out[b].skip == 0) {
            const double N = d.N;
            c->acc[7] += 1;
            // LU: 2/3 N^3 - 1/2 N^2 + 5/6 N flops; the symmetric (LDL^T)
            // form of a dominant block: 1/3 N^3 + 1/2 N^2 - 5/6 N
            if ((p.lu_kinds & 4) && c->h_out[b].delta)
                c->acc[8] += 1.0 / 3 * N * N * N + 0.5 * N * N - 5.0 / 6 * N;
            else
                c->acc[8] += 2.0 / 3 * N * N * N - 0.5 * N * N + 5.0 / 6 * N;
        }
    }
}

// .npy file of a dim x dim matrix in the reference's byte layout
// (src/matrix.c:211-233): magic, version 1.0, header
// "{'descr': '<f8', 'fortran_order': True, 'shape': (dim, dim), }", then the
// elements with (i, j) at i*dim + j.  at(i, j) supplies element (i, j).
bool npy_write(const char* path, int dim, const std::function<double(int, int)>& at)
{
    FILE* f = fopen(path, "wb");
    if (!f) return false;
    char hdr[128];
    snprintf(hdr, sizeof(hdr), "{'descr': '<f8', 'fortran_order': True, 'shape': (%d, %d), }", dim, dim);
    const uint16_t len = (uint16_t)strlen(hdr);
    bool ok = fwrite("\x93NUMPY\x01\x00", 1, 8, f) == 8 && fwrite(&len, sizeof(len), 1, f) == 1 &&
              fwrite(hdr, 1, len, f) == len;
    std::vector<double> row(dim);
    for (int i = 0; i < dim && ok; i++) {
        for (int j = 0; j < dim; j++) row[j] = at(i, j);
        ok = fwrite(row.data(), sizeof(double), dim, f) == (size_t)dim;
    }
    return fclose(f) == 0 && ok;
}

// CAPS_DUMP: with a dense determinant algorithm the reference's
// kernel_logdet writes every matrix M it factors (after the trace shortcut)
// to the file the variable names (src/matrix.c:84-109).
const char* dump_path(detalg_t detalg)
{
    const char* p = getenv("CAPS_DUMP");
    return (p && *p && detalg != DETALG_HODLR) ? p : nullptr;
}

// log det(1-M) from the block scalars with the reference's shortcut rules.
double finish_block(const BlockOut& o, int lu_mode)
{
    const double tr = o.trace;
    if (o.skip == 3) return NAN;   // non-finite trace: the LU did not run
    if (fabs(tr) < 1e-8) return -tr;
    if (lu_mode && o.skip == 2) return -tr - o.tr2 / 2;
    const double ld = o.logdet;
    if (!lu_mode && fabs(tr) > fabs(ld)) return -tr;
    return ld;
}

// Dense determinant mode (every detalg but HODLR: the reference's
// kernel_logdet takes its dense branch with the Mercator shortcut for LU, QR
// and CHOLESKY, src/matrix.c:84-113, 489-505).  Returns 0 for HODLR, 1 for LU
// and QR, 2 for CHOLESKY (whose dense matrix the reference fills below the
// diagonal only, so its Mercator test sees the lower triangle).
int dense_mode(detalg_t a) { return a == DETALG_HODLR ? 0 : a == DETALG_CHOLESKY ? 2 : 1; }
bool is_lu_mode(detalg_t a) { return dense_mode(a) != 0; }

// Largest number of blocks per launch sequence that fits the HBM budget.
int chunk_blocks_multi(caps_t* const* objs, const int* m, int n)
{
    const size_t budget = arena_budget();
    size_t acc = 0;
    int k = 0;
    for (; k < n; k++) {
        acc += block_bytes_estimate(objs[k], m[k]);
        if (acc > budget && k > 0) break;
    }
    return std::max(k, 1);
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

caps_t* caps_init(double R, double L)
{
    if (L <= 0 || R <= 0) return NULL;
    caps_t* self = (caps_t*)calloc(1, sizeof(caps_t));
    TERMINATE(!self, "out of memory");
    const double LbyR = L / R;
    self->LbyR = LbyR;
    self->L = L;
    self->R = R;
    self->calL = L + R;
    self->y = -log(2.0) - log1p(LbyR);
    self->ldim = (int)ceil(std::max((double)CAPS_MINIMUM_LDIM, CAPS_FACTOR_LDIM / LbyR));
    self->epsilonm1_plate = caps_epsilonm1_perf;
    self->epsilonm1_sphere = caps_epsilonm1_perf;
    self->userdata_plate = NULL;
    self->userdata_sphere = NULL;
    self->epsrel = CAPS_EPSREL;
    self->detalg = DETALG_HODLR;
    self->priv = NULL;
    return self;
}

void caps_free(caps_t* self)
{
    if (!self) return;
    {
        std::lock_guard<std::mutex> lock(g_resume_mu);
        auto it = g_resume.find(self);
        if (it != g_resume.end()) {
            if (it->second.out) fclose(it->second.out);
            g_resume.erase(it);
        }
    }
    ctx_free((Ctx*)self->priv);
    free(self);
}

void caps_build(FILE* stream, const char* prefix)
{
    if (!prefix) prefix = "";
    fprintf(stream, "%sversion: 0.5 (libcaps_amd, gfx950)\n", prefix);
    fprintf(stream, "%scompiler: hipcc/clang %s\n", prefix, __clang_version__);
    fprintf(stream, "%scompile time: %s %s\n", prefix, __DATE__, __TIME__);
}

void caps_info(caps_t* self, FILE* stream, const char* prefix)
{
    const char* a;
    if (!prefix) prefix = "";
    switch (self->detalg) {
    case DETALG_HODLR: a = "HODLR"; break;
    case DETALG_LU: a = "LU"; break;
    case DETALG_QR: a = "QR"; break;
    case DETALG_CHOLESKY: a = "CHOLESKY"; break;
    default: a = "unknown";
    }
    fprintf(stream, "%sL/R    = %.16g\n", prefix, self->LbyR);
    fprintf(stream, "%sL      = %.16g\n", prefix, self->L);
    fprintf(stream, "%sR      = %.16g\n", prefix, self->R);
    fprintf(stream, "%sldim   = %d\n", prefix, self->ldim);
    fprintf(stream, "%sepsrel = %.1e\n", prefix, self->epsrel);
    fprintf(stream, "%sdetalg = %s\n", prefix, a);
}

int caps_set_epsrel(caps_t* self, double epsrel)
{
    if (epsrel <= 0) return 0;
    self->epsrel = epsrel;
    return 1;
}
double caps_get_epsrel(caps_t* self) { return self->epsrel; }

int caps_set_ldim(caps_t* self, int ldim)
{
    if (ldim < 1) return 0;
    self->ldim = ldim;
    return 1;
}
int caps_get_ldim(caps_t* self) { return self->ldim; }

int caps_set_detalg(caps_t* self, detalg_t detalg)
{
    self->detalg = detalg;
    return 1;
}
detalg_t caps_get_detalg(caps_t* self) { return self->detalg; }

void caps_set_epsilonm1(caps_t* self, double (*f)(double, void*), void* userdata)
{
    caps_set_epsilonm1_plate(self, f, userdata);
    caps_set_epsilonm1_sphere(self, f, userdata);
}
void caps_set_epsilonm1_plate(caps_t* self, double (*f)(double, void*), void* userdata)
{
    self->epsilonm1_plate = f;
    self->userdata_plate = userdata;
}
void caps_set_epsilonm1_sphere(caps_t* self, double (*f)(double, void*), void* userdata)
{
    self->epsilonm1_sphere = f;
    self->userdata_sphere = userdata;
}

double caps_epsilonm1_plate(caps_t* self, double xi_)
{
    return self->epsilonm1_plate(xi_ * CAPS_c / self->calL, self->userdata_plate);
}
double caps_epsilonm1_sphere(caps_t* self, double xi_)
{
    return self->epsilonm1_sphere(xi_ * CAPS_c / self->calL, self->userdata_sphere);
}
double caps_epsilonm1_perf(double, void*) { return INFINITY; }
double caps_epsilonm1_drude(double xi, void* userdata)
{
    const double* p = (const double*)userdata;
    return p[0] * p[0] / (xi * (xi + p[1]));
}

// ---- tabulated materials ------------------------------------------------------

// ε(iξ) - 1 from a table (linear interpolation in ξ) with Drude
// extrapolation below and above it.  File format of the reference's
// material_init (src/material.c:63-180): lines "# omegap_low = <eV>",
// "# gamma_low", "# omegap_high", "# gamma_high" (eV), other '#' lines
// ignored, data lines "ξ[rad/s] ε(iξ)" separated by blanks or tabs with ξ
// strictly ascending.
struct capsx_material {
    double calL = 0;
    double omegap_low = 0, gamma_low = 0, omegap_high = 0, gamma_high = 0;   // rad/s
    std::vector<double> xi, epsm1;
};

capsx_material* capsx_material_init(const char* filename, double calL)
{
    FILE* f = fopen(filename, "r");
    if (!f) return nullptr;
    capsx_material* mt = new capsx_material();
    mt->calL = calL;
    char line[2048];
    bool ok = true;
    auto key = [&](const char* ln, const char* k, double* v) {
        const char* p = strstr(ln, k);
        if (!p) return;
        p = strchr(p + strlen(k), '=');
        if (p) *v = strtod(p + 1, nullptr);
    };
    while (fgets(line, sizeof(line), f)) {
        char* b = line;
        while (*b == ' ' || *b == '\t') b++;
        if (*b == '#') {
            key(b, "omegap_low", &mt->omegap_low);
            key(b, "gamma_low", &mt->gamma_low);
            key(b, "omegap_high", &mt->omegap_high);
            key(b, "gamma_high", &mt->gamma_high);
            continue;
        }
        char* end = nullptr;
        const double x = strtod(b, &end);
        if (end == b || (*end != ' ' && *end != '\t')) continue;   // not a data line
        const double eps = strtod(end, nullptr);
        if (!mt->xi.empty() && !(x > mt->xi.back())) {
            ok = false;   // not ascending
            break;
        }
        mt->xi.push_back(x);
        mt->epsm1.push_back(eps - 1);
    }
    fclose(f);
    if (!ok || mt->xi.size() < 2) {
        delete mt;
        return nullptr;
    }
    const double hev = CAPS_hbar_eV;
    mt->omegap_low /= hev;
    mt->gamma_low /= hev;
    mt->omegap_high /= hev;
    mt->gamma_high /= hev;
    return mt;
}

void capsx_material_free(capsx_material* mt) { delete mt; }

// ε(iξ) - 1 at ξ (rad/s); userdata is the capsx_material (a callback for
// caps_set_epsilonm1).  Reference: src/material.c:255-293.
double capsx_material_epsilonm1(double xi, void* userdata)
{
    const capsx_material* mt = (const capsx_material*)userdata;
    if (xi < mt->xi.front()) return mt->omegap_low * mt->omegap_low / (xi * (xi + mt->gamma_low));
    if (xi > mt->xi.back()) return mt->omegap_high * mt->omegap_high / (xi * (xi + mt->gamma_high));
    size_t r = std::upper_bound(mt->xi.begin(), mt->xi.end(), xi) - mt->xi.begin();
    if (r >= mt->xi.size()) r = mt->xi.size() - 1;
    const size_t l = r - 1;
    return mt->epsm1[l] + (xi - mt->xi[l]) * (mt->epsm1[r] - mt->epsm1[l]) / (mt->xi[r] - mt->xi[l]);
}

// Drude parameters of the extrapolations (rad/s); NULL outputs skipped.
// Reference: src/material.c:200-210.
void capsx_material_get_extrapolation(const capsx_material* mt, double* omegap_low, double* gamma_low,
                                      double* omegap_high, double* gamma_high)
{
    if (omegap_low) *omegap_low = mt->omegap_low;
    if (gamma_low) *gamma_low = mt->gamma_low;
    if (omegap_high) *omegap_high = mt->omegap_high;
    if (gamma_high) *gamma_high = mt->gamma_high;
}

static double host_lfac(unsigned n)
{
    if (n < 1024) return lgamma((double)n + 1.0);
    const double C = 0.91893853320467274178;
    return (n + 0.5) * log((double)n) - n + C + (1. / 12) / n;
}

double caps_lnLambda(int l1, int l2, int m)
{
    return (log(2.0 * l1 + 1) + log(2.0 * l2 + 1) - log((double)l1) - log(l1 + 1.0) - log((double)l2) - log(l2 + 1.0)
            + host_lfac(l1 - m) + host_lfac(l2 - m) - host_lfac(l1 + m) - host_lfac(l2 + m)) / 2.0;
}

int caps_estimate_lminmax(caps_t* self, int m, size_t* lmin_p, size_t* lmax_p)
{
    const int ldim = self->ldim;
    if (m == 0) {
        *lmin_p = 1;
        *lmax_p = 1 + ldim;
        return 0;
    }
    const int l = (int)round(m / sqrt(2 * self->LbyR));
    int lmin = l - ldim / 2;
    if (lmin < m) lmin = m;
    *lmin_p = lmin;
    *lmax_p = lmin + ldim;
    return l;
}

// Extension: log|a_l|, log|b_l| of the sphere at n points (xi_[i], l[i]) in
// one launch (ε(iξ) - 1 of the sphere evaluated on the host per point).  As
// the reference's caps_mie, no argument checks (xi_ <= 0 or l < 1 give NaN
// or meaningless values).  Returns 0 on success (1: n < 0).
int capsx_mie_points(caps_t* self, int n, const double* xi_, const int* l, double* lna, double* lnb)
{
    if (n < 0) return 1;
    if (n == 0) return 0;
    Ctx* c = ctx_of(self);
    std::lock_guard<std::mutex> lock(c->mu);
    std::vector<double> h(2 * (size_t)n);   // xi_, eps
    double last_xi = NAN, eps = 0;
    for (int i = 0; i < n; i++) {
        if (!(xi_[i] == last_xi)) {
            eps = caps_epsilonm1_sphere(self, xi_[i]);
            last_xi = xi_[i];
        }
        h[i] = xi_[i];
        h[n + i] = eps;
    }
    const size_t o_l = align256(16 * (size_t)n), o_out = o_l + align256(4 * (size_t)n);
    char* d = ensure_pts(c, o_out + 16 * (size_t)n);
    std::vector<double> r(2 * (size_t)n);
    HIPCHECK(hipMemcpyAsync(d, h.data(), 16 * (size_t)n, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(d + o_l, l, 4 * (size_t)n, hipMemcpyHostToDevice, c->stream));
    capsd_launch_mie_points(n, (const double*)d, (const int*)(d + o_l), self->LbyR, (const double*)d + n,
                            (double*)(d + o_out), c->stream);
    HIPCHECK(hipMemcpyAsync(r.data(), d + o_out, 16 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    HIPCHECK(hipGetLastError());
    for (int i = 0; i < n; i++) {
        lna[i] = r[2 * i];
        lnb[i] = r[2 * i + 1];
    }
    return 0;
}

// Extension: Fresnel coefficients r_TE, r_TM of the plate at n points
// (xi_[i], k_[i]) in one launch.  Returns 0 on success (1: n < 0).
int capsx_fresnel_points(caps_t* self, int n, const double* xi_, const double* k_, double* r_TE, double* r_TM)
{
    if (n < 0) return 1;
    if (n == 0) return 0;
    Ctx* c = ctx_of(self);
    std::lock_guard<std::mutex> lock(c->mu);
    std::vector<double> h(3 * (size_t)n);   // xi_, k_, eps
    double last_xi = NAN, eps = 0;
    for (int i = 0; i < n; i++) {
        if (!(xi_[i] == last_xi)) {
            eps = caps_epsilonm1_plate(self, xi_[i]);
            last_xi = xi_[i];
        }
        h[i] = xi_[i];
        h[n + i] = k_[i];
        h[2 * (size_t)n + i] = eps;
    }
    const size_t o_out = align256(24 * (size_t)n);
    char* d = ensure_pts(c, o_out + 16 * (size_t)n);
    std::vector<double> r(2 * (size_t)n);
    const double* dd = (const double*)d;
    HIPCHECK(hipMemcpyAsync(d, h.data(), 24 * (size_t)n, hipMemcpyHostToDevice, c->stream));
    capsd_launch_fresnel_points(n, dd, dd + n, dd + 2 * (size_t)n, (double*)(d + o_out), c->stream);
    HIPCHECK(hipMemcpyAsync(r.data(), d + o_out, 16 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    HIPCHECK(hipGetLastError());
    for (int i = 0; i < n; i++) {
        r_TE[i] = r[2 * i];
        r_TM[i] = r[2 * i + 1];
    }
    return 0;
}

void caps_mie(caps_t* self, double xi_, int l, double* lna, double* lnb)
{
    capsx_mie_points(self, 1, &xi_, &l, lna, lnb);
}

void caps_mie_perf(caps_t* self, double xi_, int l, double* lna, double* lnb)
{
    double (*f)(double, void*) = self->epsilonm1_sphere;
    void* u = self->userdata_sphere;
    self->epsilonm1_sphere = caps_epsilonm1_perf;
    caps_mie(self, xi_, l, lna, lnb);
    self->epsilonm1_sphere = f;
    self->userdata_sphere = u;
}

void caps_fresnel(caps_t* self, double xi_, double k_, double* r_TE, double* r_TM)
{
    capsx_fresnel_points(self, 1, &xi_, &k_, r_TE, r_TM);
}

// Extension: log I_{n+1/2}(x), log K_{n+1/2}(x) (order[i] = n >= 0; NaN for
// order < 0) and I_ν(x)/I_{ν+1}(x) (nu[i] >= 0; NaN for nu < 0) at n points,
// the device functions of the Mie coefficients (src/bessel.c:635-660,
// 781-833), in one launch.  Returns 0 on success.
int capsx_bessel_points(int n, const int* order, const double* nu, const double* x, double* logI, double* logK,
                        double* ratio)
{
    if (n < 0) return 1;
    for (int i = 0; i < n; i++)
        if (!(x[i] > 0)) return 1;
    if (n == 0) return 0;
    caps_t* tmp = caps_init(1, 1);
    Ctx* c = ctx_of(tmp);
    const size_t o_nu = align256(4 * (size_t)n), o_x = o_nu + align256(8 * (size_t)n),
                 o_out = o_x + align256(8 * (size_t)n);
    char* d = ensure_pts(c, o_out + 24 * (size_t)n);
    std::vector<double> r(3 * (size_t)n);
    HIPCHECK(hipMemcpyAsync(d, order, 4 * (size_t)n, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(d + o_nu, nu, 8 * (size_t)n, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(d + o_x, x, 8 * (size_t)n, hipMemcpyHostToDevice, c->stream));
    capsd_launch_bessel_points(n, (const int*)d, (const double*)(d + o_nu), (const double*)(d + o_x),
                               (double*)(d + o_out), c->stream);
    HIPCHECK(hipMemcpyAsync(r.data(), d + o_out, 24 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    HIPCHECK(hipGetLastError());
    caps_free(tmp);
    for (int i = 0; i < n; i++) {
        if (logI) logI[i] = r[3 * i];
        if (logK) logK[i] = r[3 * i + 1];
        if (ratio) ratio[i] = r[3 * i + 2];
    }
    return 0;
}

// Extension: n blocks, block i of caps object objs[i] (objects may differ in
// R/L, ldim, material), in one batched launch sequence on objs[0]'s device
// context.  All objects must use the same detalg.  With trace / flags
// non-NULL, also each block's tr M and how its value was obtained: flags
// bits 0-1 the shortcut (0 LU, 1 trace |tr M| < 1e-8, 2 LU-mode Mercator,
// 3 non-finite trace), bit 2 set when the LU ran in delta form (1 - M
// column diagonally dominant: no pivot search, log1p of u_ii - 1), bit 3
// set when the HODLR-mode rule |tr M| > |log det| replaced the LU value by
// -tr M.  Returns 0 on success.
static int logdetD_multi_ordered(caps_t* host, caps_t* const* objs, const double* xi_, const int* m, int n,
                                 double* out, double* trace, int* flags);

int capsx_logdetD_multi_info(caps_t* const* objs, const double* xi_, const int* m, int n, double* out, double* trace,
                             int* flags)
{
    if (n <= 0) return 0;
    // the blocks of one caps object and m (whose I tables share the a_q,
    // k_I) are run next to each other, so that they fall into one part of
    // the batch and into full groups; larger blocks first (the LU's stream
    // groups are contiguous runs of blocks, balanced by cost): ldim
    // descending, then objects in the order of their first block, then m,
    // then the input order.  The batch runs on objs[0]'s context (device
    // buffers, statistics, graph setting) either way.
    std::vector<int> ord(n), first(n);
    {
        std::unordered_map<caps_t*, int> seen;
        for (int i = 0; i < n; i++) {
            first[i] = seen.emplace(objs[i], i).first->second;
            ord[i] = i;
        }
    }
    auto ldim_of = [&](int i) { return objs[i] ? objs[i]->ldim : 0; };
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
        return std::make_tuple(-ldim_of(a), first[a], m[a]) < std::make_tuple(-ldim_of(b), first[b], m[b]);
    });
    bool same = true;
    for (int i = 0; i < n && same; i++) same = ord[i] == i;
    if (same) return logdetD_multi_ordered(objs[0], objs, xi_, m, n, out, trace, flags);
    std::vector<caps_t*> po(n);
    std::vector<double> px(n), pout(n), ptr(trace ? n : 0);
    std::vector<int> pm(n), pf(flags ? n : 0);
    for (int i = 0; i < n; i++) {
        po[i] = objs[ord[i]];
        px[i] = xi_[ord[i]];
        pm[i] = m[ord[i]];
    }
    const int rc = logdetD_multi_ordered(objs[0], po.data(), px.data(), pm.data(), n, pout.data(),
                                         trace ? ptr.data() : nullptr, flags ? pf.data() : nullptr);
    if (rc) return rc;
    for (int i = 0; i < n; i++) {
        out[ord[i]] = pout[i];
        if (trace) trace[ord[i]] = ptr[i];
        if (flags) flags[ord[i]] = pf[i];
    }
    return 0;
}

static int logdetD_multi_ordered(caps_t* host, caps_t* const* objs, const double* xi_, const int* m, int n,
                                 double* out, double* trace, int* flags)
{
    for (int i = 0; i < n; i++) {
        if (!objs[i] || !(xi_[i] > 0) || m[i] < 0) {
            fprintf(stderr, "caps_logdetD_batch: block %d has xi_=%g, m=%d (need xi_ > 0, m >= 0)\n", i, xi_[i], m[i]);
            return 1;
        }
        TERMINATE(2 * objs[i]->ldim > capsd_lu_max_dim(), "ldim=%d exceeds the LU's maximum dimension %d",
                  objs[i]->ldim, capsd_lu_max_dim());
        if (is_lu_mode(objs[i]->detalg) != is_lu_mode(objs[0]->detalg)) {
            fprintf(stderr, "capsx_logdetD_multi: all objects must use the same determinant mode\n");
            return 1;
        }
    }
    Ctx* c = ctx_of(host);
    std::lock_guard<std::mutex> lock(c->mu);
    const int lu_mode = dense_mode(objs[0]->detalg);
    double tot[6] = {0, 0, 0, 0, 0, 0};
    // a batch of several launch sequences takes the whole budget at once:
    // HIP kept a freed arena's memory from the next allocation (device memory
    // free after growing a 92.6 GB arena to 100 GB: 115 of 309 GB), so an
    // arena grown between sequences held both
    if (chunk_blocks_multi(objs, m, n) < n) ensure_arena(c, arena_budget());
    // CAPS_AMD_PROGRESS=1: one line on stderr per launch sequence
    static const bool progress = env_double("CAPS_AMD_PROGRESS", 0) > 0;
    for (int b0 = 0; b0 < n;) {
        const double t_seq = progress ? wall_seconds() : 0;
        const int k = chunk_blocks_multi(objs + b0, m + b0, n - b0);
        ensure_blocks(c, k);
        BatchPlan p = plan_batch_multi(objs + b0, c, xi_ + b0, m + b0, k, 1);
        run_batch(host, c, p, 0, lu_mode);
        // blocks whose 1 - M is not column dominant were left unfactored by
        // the delta-form LU: run them again with the pivoted LU
        std::vector<int> again;
        for (int i = 0; i < k; i++)
            if (c->h_out[i].skip == 0 && !c->h_out[i].delta) again.push_back(b0 + i);
        auto take = [&](const BlockOut& o, int i) {
            out[i] = finish_block(o, lu_mode);
            if (trace) trace[i] = o.trace;
            if (flags) {
                const bool lu_ran = o.skip == 0;
                flags[i] = (o.skip & 3) | (lu_ran && o.delta ? 4 : 0) |
                           (lu_ran && !lu_mode && fabs(o.trace) > fabs(o.logdet) ? 8 : 0);
            }
        };
        for (int i = 0; i < k; i++) take(c->h_out[i], b0 + i);
        for (int i = 0; i < 6; i++) tot[i] += c->timing[i];
        if (!again.empty()) {
            std::vector<caps_t*> ro;
            std::vector<double> rx;
            std::vector<int> rm;
            for (int i : again) {
                ro.push_back(objs[i]);
                rx.push_back(xi_[i]);
                rm.push_back(m[i]);
            }
            BatchPlan q = plan_batch_multi(ro.data(), c, rx.data(), rm.data(), (int)again.size(), 1);
            q.lu_kinds = 1;
            run_batch(host, c, q, 0, lu_mode);
            for (size_t j = 0; j < again.size(); j++) take(c->h_out[j], again[j]);
            for (int i = 0; i < 6; i++) tot[i] += c->timing[i];
        }
        if (progress)
            fprintf(stderr, "capsx: blocks %d..%d of %d (dimension <= %d, %zu pivoted) in %.2f s\n", b0, b0 + k - 1,
                    n, p.maxN, again.size(), wall_seconds() - t_seq);
        b0 += k;
    }
    memcpy(c->timing, tot, sizeof(tot));
    return 0;
}

int capsx_logdetD_multi(caps_t* const* objs, const double* xi_, const int* m, int n, double* out)
{
    return capsx_logdetD_multi_info(objs, xi_, m, n, out, nullptr, nullptr);
}

int caps_logdetD_batch(caps_t* self, const double* xi_, const int* m, int n, double* out)
{
    std::vector<caps_t*> objs(n > 0 ? n : 1, self);
    return capsx_logdetD_multi(objs.data(), xi_, m, n, out);
}

// Host copy of a block's matrix (the reference's integration_t slot of
// caps_M_t).
// The K and I tables of one block (xi_, m) copied from the device, behind
// the integration object of caps_integrate_init.
struct IntegTables {
    int m = 0;
    bool is_pr = false;
    int mu = 0, nK = 0, lo = 0, nl = 0;
    std::vector<double> lnK;            // [2][nK]: log|K_nu^p|, nu = mu ..
    std::vector<double> logI;           // [2][nl][nl]: log|I_{l1 l2}^p|, l from lo
    std::vector<signed char> sgnI;      // [2][nl][nl]
};

struct caps_block_state {
    double* M;   // N x N, column-major, lda = N (caps_M_init)
    int N;
    IntegTables* tab;   // caps_integrate_init; NULL for a caps_M_t's state
};

double caps_logdetD(caps_t* self, double xi_, int m)
{
    TERMINATE(xi_ <= 0, "Matsubara frequency must be positive");
    if (const char* path = dump_path(self->detalg)) {
        // the matrix M of the block (built once more, without the LU)
        caps_M_t* M = caps_M_init(self, m, xi_);
        const caps_block_state* st = M->integration;
        double tr = 0, comp = 0;
        for (int i = 0; i < st->N; i++) {
            const double y = st->M[i + (size_t)i * st->N] - comp, t = tr + y;
            comp = (t - tr) - y;
            tr = t;
        }
        if (fabs(tr) >= 1e-8)
            TERMINATE(!npy_write(path, st->N, [&](int i, int j) { return st->M[i + (size_t)j * st->N]; }),
                      "cannot write %s", path);
        caps_M_free(M);
    }
    double out;
    caps_logdetD_batch(self, &xi_, &m, 1, &out);
    return out;
}

void caps_last_timing(caps_t* self, double ms[6])
{
    Ctx* c = ctx_of(self);
    for (int i = 0; i < 6; i++) ms[i] = c->timing[i];
}

// ---- round-trip matrix object ----------------------------------------------

caps_M_t* caps_M_init(caps_t* caps, int m, double xi_)
{
    TERMINATE(xi_ <= 0, "Matsubara frequency must be positive");
    Ctx* c = ctx_of(caps);
    std::lock_guard<std::mutex> lock(c->mu);
    ensure_blocks(c, 1);
    BatchPlan p = plan_batch(caps, c, &xi_, &m, 1, 0);
    run_batch(caps, c, p, 1, 0);
    const BlockDesc& d = c->h_desc[0];
    caps_M_t* self = (caps_M_t*)calloc(1, sizeof(caps_M_t));
    self->caps = caps;
    self->m = m;
    self->lmin = d.lmin;
    self->ldim = d.ldim;
    self->xi_ = xi_;
    self->integration_plasma = NULL;
    caps_block_state* st = (caps_block_state*)calloc(1, sizeof(caps_block_state));
    st->N = d.N;
    st->M = (double*)malloc((size_t)d.N * d.N * sizeof(double));
    HIPCHECK(hipMemcpy2D(st->M, d.N * sizeof(double), c->arena + d.off_mat, d.lda * sizeof(double),
                         d.N * sizeof(double), d.N, hipMemcpyDeviceToHost));
    self->al = (double*)malloc(d.ldim * sizeof(double));
    self->bl = (double*)malloc(d.ldim * sizeof(double));
    HIPCHECK(hipMemcpy(self->al, c->arena + d.off_mie, d.ldim * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(self->bl, c->arena + d.off_mie + d.ldim * 8, d.ldim * sizeof(double), hipMemcpyDeviceToHost));
    self->integration = st;
    return self;
}

double caps_M_elem(caps_M_t* self, int l1, int l2, char p1, char p2)
{
    const int i = 2 * (l1 - self->lmin) + (p1 == 'M'), j = 2 * (l2 - self->lmin) + (p2 == 'M');
    TERMINATE(i < 0 || j < 0 || i >= 2 * self->ldim || j >= 2 * self->ldim,
              "l1=%d, l2=%d outside the window [%d, %d)", l1, l2, self->lmin, self->lmin + self->ldim);
    const caps_block_state* st = self->integration;
    return st->M[i + (size_t)j * st->N];
}

double caps_kernel_M(int i, int j, void* args_)
{
    caps_M_t* a = (caps_M_t*)args_;
    const char pol[2] = {'E', 'M'};
    return caps_M_elem(a, a->lmin + i / 2, a->lmin + j / 2, pol[i % 2], pol[j % 2]);
}

void caps_M_free(caps_M_t* self)
{
    if (!self) return;
    if (self->integration) {
        free(self->integration->M);
        free(self->integration);
    }
    free(self->al);
    free(self->bl);
    free(self);
}

// ---- generic determinant ----------------------------------------------------

double kernel_logdet(int dim, double (*kernel)(int, int, void*), void* args, int sym_spd, detalg_t detalg)
{
    (void)sym_spd;
    TERMINATE(dim <= 0 || dim > capsd_lu_max_dim(), "dimension %d outside [1, %d]", dim, capsd_lu_max_dim());
    std::vector<double> diag(dim);
    for (int i = 0; i < dim; i++) diag[i] = kernel(i, i, args);
    double tr = 0, comp = 0;
    for (int i = 0; i < dim; i++) {
        const double y = diag[i] - comp, t = tr + y;
        comp = (t - tr) - y;
        tr = t;
    }
    if (fabs(tr) < 1e-8) return -tr;
    const int lda = (dim + 15) & ~15;
    const int dmode = dense_mode(detalg);
    std::vector<double> A((size_t)lda * dim, 0.0);
    for (int i = 0; i < dim; i++) A[i + (size_t)i * lda] = diag[i];
    // by minor diagonals as src/matrix.c:93-102; CHOLESKY fills the lower
    // triangle only (the reference's Cholesky reads only that half)
    for (int md = 1; md < dim; md++)
        for (int k = 0; k < dim - md; k++) {
            if (dmode != 2) A[k + (size_t)(md + k) * lda] = kernel(k, md + k, args);
            A[(md + k) + (size_t)k * lda] = kernel(md + k, k, args);
        }
    if (const char* path = dump_path(detalg))
        TERMINATE(!npy_write(path, dim, [&](int i, int j) { return A[i + (size_t)j * lda]; }), "cannot write %s",
                  path);
    const int lu_mode = is_lu_mode(detalg);
    if (lu_mode) {
        double norm = 0;
        for (int j = 0; j < dim; j++)
            for (int i = 0; i < dim; i++) norm += A[i + (size_t)j * lda] * A[i + (size_t)j * lda];
        norm = sqrt(norm);
        if (norm < 1) {
            std::vector<double> row(dim);
            for (int i = 0; i < dim; i++) {
                double s = 0;
                for (int j = 0; j < dim; j++) s += A[i + (size_t)j * lda] * A[j + (size_t)i * lda];
                row[i] = s;
            }
            double t2 = 0, c2 = 0;
            for (int i = 0; i < dim; i++) {
                const double y = row[i] - c2, t = t2 + y;
                c2 = (t - t2) - y;
                t2 = t;
            }
            const double mercator = -tr - t2 / 2;
            const double err = fabs(norm * norm / 2 + norm + log1p(-norm));
            if (fabs(err / mercator) < 1e-8) return mercator;
        }
    }
    // the symmetric matrix the reference's Cholesky factors (from the lower
    // triangle) for the LU
    if (dmode == 2)
        for (int j = 1; j < dim; j++)
            for (int i = 0; i < j; i++) A[i + (size_t)j * lda] = A[j + (size_t)i * lda];
    // -M, and 1 - M (the unit diagonal back) unless 1 - M is column
    // diagonally dominant (then the LU's delta form: no pivot search,
    // log1p of the pivots' deviations from 1)
    for (size_t i = 0; i < A.size(); i++) A[i] = -A[i];
    bool dominant = g_delta_mode != 0;
    for (int j = 0; j < dim && dominant; j++) {
        double sj = 0;
        for (int i = 0; i < dim; i++)
            if (i != j) sj += fabs(A[i + (size_t)j * lda]);
        dominant = sj < fabs(1.0 + A[j + (size_t)j * lda]);
    }
    if (!dominant)
        for (int i = 0; i < dim; i++) A[i + (size_t)i * lda] += 1;
    // one-block batch with a user matrix
    caps_t* tmp = caps_init(1, 1);
    Ctx* c = ctx_of(tmp);
    ensure_blocks(c, 1);
    BlockDesc& d = c->h_desc[0];
    memset(&d, 0, sizeof(d));
    d.N = dim;
    d.ldim = (dim + 1) / 2;
    d.lda = lda;
    int64_t off = 0;
    d.off_mat = off;
    off = align256(off + (int64_t)lda * dim * 8);
    LuSide& s = c->h_side[0];
    off = lu_side_layout(s, off, dim, lda);
    ensure_arena(c, off);
    BlockOut o;
    memset(&o, 0, sizeof(o));
    o.delta = dominant ? 1 : 0;
    HIPCHECK(hipMemcpyAsync(c->d_desc, &d, sizeof(BlockDesc), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(c->d_side, &s, sizeof(LuSide), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(c->d_out, &o, sizeof(BlockOut), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(c->arena, A.data(), A.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    {
        HIPCHECK(hipEventRecord(c->ev[6], c->stream));
        launch_lu_groups(c, 1, dim, lu_config(dim, 1), c->ev[6], false, dominant ? 2 : 1);
    }
    HIPCHECK(hipMemcpyAsync(&o, c->d_out, sizeof(BlockOut), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    caps_free(tmp);
    const double ld = o.logdet;
    if (!lu_mode && fabs(tr) > fabs(ld)) return -tr;
    return ld;
}

// ---- ξ = 0 Drude term --------------------------------------------------------

double caps_ht_drude(caps_t* caps)
{
    const double x = caps->LbyR, x2 = x * x;
    const double Z = (1 + x) * (1 - sqrt((x2 + 2 * x) / (x2 + 2 * x + 1)));
    double s1 = 0, s2 = 0;
    for (int l = 1;; l++) {
        const double a = pow(Z, 2 * l), b = a * Z;
        const double v1 = (2 * l + 1) * log1p(-b), v2 = b * (1 - a) / (1 - b);
        s1 += v1;
        s2 += v2;
        if (fabs(v1 / s1) < 1e-15 && fabs(v2 / s2) < 1e-15) break;
    }
    return 0.5 * (s1 + log1p(-(1 - Z * Z) * s2));
}

// ---- ξ = 0 blocks (perfect reflector / Drude EE and MM) ----------------------

// Closed-form ξ = 0 round-trip kernels.  Reference: src/libcaps.c:860-941
// (EE: exp((l1+l2+1) y + log (l1+l2)! - ½ Σ log(l±m)!); MM additionally
// × sqrt(l1 l2/((l1+1)(l2+1)))).  args is a caps_M_t (caps, m, lmin).
double caps_kernel_M0_EE(int i, int j, void* args_)
{
    caps_M_t* a = (caps_M_t*)args_;
    const double y = a->caps->y;
    const int l1 = i + a->lmin, l2 = j + a->lmin, m = a->m;
    return exp((l1 + l2 + 1) * y + host_lfac(l1 + l2)
               - 0.5 * (host_lfac(l1 + m) + host_lfac(l1 - m) + host_lfac(l2 + m) + host_lfac(l2 - m)));
}

double caps_kernel_M0_MM(int i, int j, void* args_)
{
    caps_M_t* a = (caps_M_t*)args_;
    const double y = a->caps->y;
    const int l1 = i + a->lmin, l2 = j + a->lmin, m = a->m;
    return exp((l1 + l2 + 1) * y + host_lfac(l1 + l2)
               - 0.5 * (host_lfac(l1 + m) + host_lfac(l1 - m) + host_lfac(l2 + m) + host_lfac(l2 - m)))
           * sqrt(((double)l1 * (double)l2) / ((l1 + 1.) * (l2 + 1.)));
}

}  // extern "C"

namespace {

// log det(1 - M) of the n x n block M_ij = exp(h_i + h_j + T_{i+j}) (the
// ξ = 0 kernels): filled, traced and factored on the GPU on this object's
// context, with kernel_logdet's shortcut rules (src/matrix.c:57-149).
double sep_logdet(caps_t* self, const std::vector<double>& h, const std::vector<double>& T)
{
    const int n = (int)h.size();
    TERMINATE(n <= 0 || n > capsd_lu_max_dim() || (int)T.size() != 2 * n - 1, "xi = 0 block of dimension %d", n);
    Ctx* c = ctx_of(self);
    std::lock_guard<std::mutex> lock(c->mu);
    hipStream_t s = c->stream;
    ensure_blocks(c, 1);
    const int lda = (n + 15) & ~15;
    BlockDesc& d = c->h_desc[0];
    memset(&d, 0, sizeof(d));
    d.N = n;
    d.ldim = (n + 1) / 2;
    d.lda = lda;
    int64_t off = 0;
    d.off_mat = off;
    off = align256(off + (int64_t)lda * n * 8);
    d.off_diag = off;
    off = align256(off + (int64_t)n * 8);
    d.n_rtiles = (n + 63) / 64;   // k_fill_xi0's 64-row tiles
    d.off_cpart = off;
    off = align256(off + (int64_t)d.n_rtiles * n * 8);
    const int64_t off_tab = off;
    off = align256(off + (int64_t)(3 * n - 1) * 8);
    LuSide& sd = c->h_side[0];
    off = lu_side_layout(sd, off, n, lda);
    ensure_arena(c, off);
    std::vector<double> tab(h);
    tab.insert(tab.end(), T.begin(), T.end());
    HIPCHECK(hipMemcpyAsync(c->d_desc, &d, sizeof(BlockDesc), hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(c->d_side, &sd, sizeof(LuSide), hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemsetAsync(c->d_out, 0, sizeof(BlockOut), s));
    HIPCHECK(hipMemcpyAsync(c->arena + off_tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, s));
    const int lu_mode = dense_mode(self->detalg);
    capsd_launch_fill_xi0(c->d_desc, c->arena, off_tab, n, s);
    capsd_launch_trace(c->d_desc, c->arena, c->d_out, 1, lu_mode, s);
    capsd_launch_dominance(c->d_desc, c->arena, c->d_out, 1, n, g_delta_mode, s);
    if (const char* path = dump_path(self->detalg)) {
        BlockOut o0;
        HIPCHECK(hipMemcpyAsync(&o0, c->d_out, sizeof(BlockOut), hipMemcpyDeviceToHost, s));
        std::vector<double> A((size_t)lda * n), dg(n);
        HIPCHECK(hipMemcpyAsync(A.data(), c->arena + d.off_mat, A.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipMemcpyAsync(dg.data(), c->arena + d.off_diag, n * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        if (fabs(o0.trace) >= 1e-8)   // M = -A off the diagonal, the exact diagonal of M on it
            TERMINATE(!npy_write(path, n, [&](int i, int j) { return i == j ? dg[i] : -A[i + (size_t)j * lda]; }),
                      "cannot write %s", path);
    }
    BlockOut o0;   // the block's form (delta or pivoted) decides the LU's kernels
    HIPCHECK(hipMemcpyAsync(&o0, c->d_out, sizeof(BlockOut), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    HIPCHECK(hipEventRecord(c->ev[6], s));
    launch_lu_groups(c, 1, n, lu_config(n, 1), c->ev[6], false, o0.delta ? 2 : 1);
    BlockOut o;
    HIPCHECK(hipMemcpyAsync(&o, c->d_out, sizeof(BlockOut), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    HIPCHECK(hipGetLastError());
    return finish_block(o, lu_mode);
}

// I_ν(x)/I_{ν+1}(x) by the continued fraction of DLMF 10.33.1,
// 2(ν+1)/x + 1/(2(ν+2)/x + 1/(...)), evaluated with the modified Lentz
// method to full precision (the reference's bessel_ratioI, src/bessel.c:
// 635-660, evaluates the same fraction; its length grows like x).
double ratio_I(double nu, double x)
{
    const double tiny = 1e-300;
    double f = 2 * (nu + 1) / x, C = f, D = 0;
    for (int k = 2; k < 100000000; k++) {
        const double b = 2 * (nu + k) / x;
        D = b + D;
        C = b + 1 / C;
        if (D == 0) D = tiny;
        if (C == 0) C = tiny;
        D = 1 / D;
        const double delta = C * D;
        f *= delta;
        if (fabs(delta - 1) < 1e-16) return f;
    }
    return f;
}

// (1/ν!) ∫_0^∞ -r_TE(t) e^{-t} t^ν dt, -r_TE = (sqrt(t² + β²) - t)/(sqrt(t² + β²) + t),
// β = 2ω: with t = e^u the integrand w(e^u) e^{(ν+1)u - e^u}/ν! is smooth,
// unimodal (peak at u = log(ν+1), width 1/sqrt(ν+1)) and analytic in the
// strip |Im u| < π/2 (w has its branch points at t = ±iβ), so the trapezoid
// rule with step h converges like e^{-π²/h}; h = min(0.2, width/4) and the
// sum runs out from the peak until the terms fall below 1e-20 of it.  The
// reference integrates the same function with QUADPACK (src/integration.c:
// 924-1019).
double plasma_integral(int nu, double omega)
{
    const double u0 = log(nu + 1.0), h = std::min(0.2, 0.25 / sqrt(nu + 1.0));
    const double lf = host_lfac((unsigned)nu);
    auto f = [&](double u) {
        const double t = exp(u), q = 2 * omega / t;
        const double bm1 = q * q / (sqrt(1 + q * q) + 1);   // sqrt(1 + q²) - 1 without cancellation
        return bm1 / (2 + bm1) * exp((nu + 1) * u - t - lf);
    };
    const double f0 = f(u0);
    double s = f0;
    for (int dir = -1; dir <= 1; dir += 2)
        for (int k = 1;; k++) {
            const double v = f(u0 + dir * k * h);
            s += v;
            if (v < 1e-20 * f0 && k > 4) break;
        }
    return s * h;
}

}  // namespace

extern "C" {

// Plasma-model ξ = 0 integrals (the reference's integration_plasma_t):
// ω = ωp (L+R)/c, α = ω/(1 + L/R), cached I(ν) and Bessel ratios.
struct caps_block_state_plasma {
    double omega, alpha;
    std::vector<double> I, ratio;   // by ν and by l, NaN = not computed
    std::mutex mu;
};

extern "C" {

// Reference: src/integration.c:897-915.
integration_plasma_t* caps_integrate_plasma_init(caps_t* caps, double omegap, double epsrel)
{
    (void)epsrel;   // the trapezoid rule above converges to full precision
    integration_plasma_t* p = new integration_plasma_t();
    p->omega = omegap * caps->calL / CAPS_c;
    p->alpha = p->omega / (1 + caps->LbyR);
    return p;
}

// I(l1 + l2) and the ratios I_{l-1/2}(α)/I_{l+1/2}(α) for l1 and l2.
// Reference: src/integration.c:955-1019.
double caps_integrate_plasma(integration_plasma_t* self, int l1, int l2, int m, double* ratio1, double* ratio2)
{
    (void)m;
    std::lock_guard<std::mutex> lock(self->mu);
    auto cached = [](std::vector<double>& v, int k, const std::function<double()>& fn) {
        if ((int)v.size() <= k) v.resize(k + 1, NAN);
        if (std::isnan(v[k])) v[k] = fn();
        return v[k];
    };
    const double a = self->alpha, w = self->omega;
    *ratio1 = cached(self->ratio, l1, [&] { return ratio_I(l1 - 0.5, a); });
    *ratio2 = cached(self->ratio, l2, [&] { return ratio_I(l2 - 0.5, a); });
    return cached(self->I, l1 + l2, [&] { return plasma_integral(l1 + l2, w); });
}

void caps_integrate_plasma_free(integration_plasma_t* self) { delete self; }

}  // extern "C"

}  // extern "C"

namespace {

// ln of the common ξ = 0 element factor, split as h(l1) + h(l2) + T(l1+l2)
// with h(l) = -½ (log (l+m)! + log (l-m)!) + extra(l) and
// T(ν) = (ν + 1) y + log ν! + extra_T(ν) (src/libcaps.c:860-906).
void xi0_tables(caps_t* self, int m, int lmin, std::vector<double>& h, std::vector<double>& T)
{
    const int n = self->ldim;
    h.resize(n);
    T.resize(2 * n - 1);
    for (int i = 0; i < n; i++) {
        const int l = lmin + i;
        h[i] = -0.5 * (host_lfac(l + m) + host_lfac(l - m));
    }
    for (int k = 0; k < 2 * n - 1; k++) {
        const int nu = 2 * lmin + k;
        T[k] = (nu + 1) * self->y + host_lfac(nu);
    }
}

}  // namespace

extern "C" {

// Plasma-model MM kernel.  Reference: src/libcaps.c:882-906.
double caps_kernel_M0_MM_plasma(int i, int j, void* args_)
{
    caps_M_t* a = (caps_M_t*)args_;
    const double y = a->caps->y;
    const int l1 = i + a->lmin, l2 = j + a->lmin, m = a->m;
    double r1, r2;
    const double I = caps_integrate_plasma(a->integration_plasma, l1, l2, m, &r1, &r2);
    const double alpha = a->integration_plasma->alpha;
    const double f1 = 1 - (2 * l1 + 1.) / (alpha * r1), f2 = 1 - (2 * l2 + 1.) / (alpha * r2);
    const double factor = sqrt(l1 * f1 / (l1 + 1.) * l2 * f2 / (l2 + 1.));
    return factor * exp(host_lfac(l1 + l2)
                        - 0.5 * (host_lfac(l1 + m) + host_lfac(l1 - m) + host_lfac(l2 + m) + host_lfac(l2 - m))
                        + (l1 + l2 + 1) * y) * I;
}

// log det of the ξ = 0 blocks EE, MM and the plasma-model MM (omegap in
// rad/s), each ldim x ldim, filled and factored on the GPU; NULL outputs
// are skipped.  Reference: src/libcaps.c:984-1017.
void caps_logdetD0(caps_t* self, int m, double omegap, double* EE, double* MM, double* MM_plasma)
{
    size_t lmin_, lmax_;
    caps_estimate_lminmax(self, m, &lmin_, &lmax_);
    const int lmin = (int)lmin_, n = self->ldim;
    std::vector<double> h0, T0;
    xi0_tables(self, m, lmin, h0, T0);
    if (EE) *EE = sep_logdet(self, h0, T0);
    if (MM) {
        std::vector<double> h(h0);
        for (int i = 0; i < n; i++) h[i] += 0.5 * log((lmin + i) / (lmin + i + 1.0));
        *MM = sep_logdet(self, h, T0);
    }
    if (MM_plasma) {
        integration_plasma_t* p = caps_integrate_plasma_init(self, omegap, self->epsrel);
        std::vector<double> h(h0), T(T0);
        for (int i = 0; i < n; i++) {
            const int l = lmin + i;
            double r1, r2;
            caps_integrate_plasma(p, l, l, m, &r1, &r2);
            const double f = 1 - (2 * l + 1.) / (p->alpha * r1);
            TERMINATE(!(f > 0), "plasma xi = 0 block: factor %g at l = %d (omegap = %g)", f, l, omegap);
            h[i] += 0.5 * log(l * f / (l + 1.));
        }
        for (int k = 0; k < 2 * n - 1; k++) {
            double r1, r2;
            const int nu = 2 * lmin + k;
            T[k] += log(caps_integrate_plasma(p, nu / 2, nu - nu / 2, m, &r1, &r2));
        }
        caps_integrate_plasma_free(p);
        *MM_plasma = sep_logdet(self, h, T);
    }
}

// ξ = 0 free energy of the plasma model in units of k_B T: the Drude EE
// term plus Σ'_m MM_plasma (m = 0 halved) until |v/Σ| < eps.  Reference:
// src/libcaps.c:1127-1145.
double caps_ht_plasma(caps_t* caps, double omegap, double eps)
{
    const double drude = caps_ht_drude(caps);
    double s = 0;
    for (int m = 0;; m++) {
        double v;
        caps_logdetD0(caps, m, omegap, NULL, NULL, &v);
        s += m == 0 ? v / 2 : v;
        if (fabs(v / s) < eps) return drude + s;
    }
}

// ξ = 0 free energy of perfect reflectors: Drude EE term plus the MM blocks,
// m = 0 from Bimonte's closed form, m >= 1 until |v/Σ| < eps.
// Reference: src/libcaps.c:1089-1116.
double caps_ht_perf(caps_t* caps, double eps)
{
    const double LbyR = caps->LbyR;
    const double drude = caps_ht_drude(caps);
    double MM = 0;
    const double Z = 1. / (1 + LbyR + sqrt(LbyR * (2 + LbyR)));
    for (int l = 0;; l++) {
        const double v = 0.5 * log1p(-pow(Z, 2 * l + 3));
        MM += v;
        if (fabs(v / MM) < eps) break;
    }
    for (int m = 1;; m++) {
        double v;
        caps_logdetD0(caps, m, 0, NULL, &v, NULL);
        MM += v;
        if (fabs(v / MM) < eps) return drude + MM;
    }
}

// ---- free energy (reference driver semantics) -------------------------------

}  // extern "C"

namespace {

// Kahan sum in index order (src/misc.c:26-40).
double kahan(const std::vector<double>& x)
{
    double s = 0, c = 0;
    for (double t : x) {
        const double y = t - c, tt = s + y;
        c = (tt - s) - y;
        s = tt;
    }
    return s;
}

// The m-sum of one Matsubara term from the values computed so far
// (v[m] for m = 0 .. v.size()-1): m runs 0, 1, 2, ... and stops at the first
// m with v == 0 or v/v_0 < cutoff, that m included; m = 0 is halved.
// Reference: src/caps.c:307-396 with a single worker (sequential
// semantics).  Returns false if the cutoff is not reached within v.
bool msum(const std::vector<double>& v, double cutoff, double* out, int* mstop)
{
    for (size_t m = 0; m < v.size(); m++) {
        if (v[m] == 0 || v[m] / v[0] < cutoff) {
            std::vector<double> t(v.begin(), v.begin() + m + 1);
            t[0] /= 2;
            *out = kahan(t);
            *mstop = (int)m;
            return true;
        }
    }
    return false;
}

struct Term {
    double xi;
    std::vector<double> v;   // logdetD for m = 0 .. v.size()-1
    bool done = false;
    int mstop = -1;
    double value = 0;
    bool cached = false;     // value taken from the resume cache (no blocks)
};

// A term whose frequency is in the resume cache: its m-sum is replaced by
// the cached value (v = {2 F_ξ, 0}: the m-sum of it is F_ξ, m = 0 halved,
// the next term below any cutoff).
void term_from_cache(const caps_t* self, Term& t)
{
    double val;
    if (!resume_lookup(self, t.xi, &val)) return;
    t.v = {2 * val, 0.0};
    t.cached = true;
}

// Surrogate predictions of log|logdetD(ξ_, m)| for m = 0..mcap at each of
// the frequencies xs (row q of P holds frequency q).
void sur_predict(caps_t* self, const std::vector<double>& xs, int mcap, std::vector<float>& P)
{
    Ctx* cx = ctx_of(self);
    const size_t per = (size_t)mcap + 1;
    std::vector<float> X(xs.size() * per * 5);
    P.assign(xs.size() * per, 0.f);
    for (size_t q = 0; q < xs.size(); q++)
        for (int m = 0; m <= mcap; m++) {
            float* x = &X[(q * per + m) * 5];
            x[0] = (float)log(1 / self->LbyR);
            x[1] = (float)log(xs[q]);
            x[2] = (float)m / self->ldim;
            x[3] = cx->sur_feat[0];
            x[4] = cx->sur_feat[1];
        }
    if (!xs.empty()) capsx_mlp_predict(cx->sur, X.data(), (int)(xs.size() * per), P.data());
}

// Predicted m-cutoff of one frequency: the first m whose predicted ratio
// logdetD(m)/logdetD(0) falls below the cutoff (mcap if none does).
int sur_mcut(const float* pr, int mcap, double cutoff)
{
    const double lcut = log(cutoff);
    for (int m = 1; m <= mcap; m++)
        if (pr[m] - pr[0] < lcut) return m;
    return mcap;
}

// Score a plan against the computed terms: log errors of every computed
// block with a prediction, m-cutoff error of the finished term.
void sur_score(Ctx* cx, const float* pr, int mcap, int mpred, const Term& t)
{
    for (size_t m = 0; m < t.v.size() && (int)m <= mcap; m++) {
        if (t.v[m] == 0 || !std::isfinite(t.v[m])) continue;
        const double e = fabs((double)pr[m] - log(fabs(t.v[m])));
        cx->sur_acc[0] += 1;
        cx->sur_acc[1] += e * e;
        cx->sur_acc[2] = std::max(cx->sur_acc[2], e);
    }
    if (t.mstop >= 0) {
        const double d = fabs((double)(mpred - t.mstop));
        cx->sur_acc[3] += 1;
        cx->sur_acc[4] += d;
        cx->sur_acc[5] = std::max(cx->sur_acc[5], d);
    }
}

// Evaluate the listed (term, m-range) requests in one batch.
void run_requests(caps_t* self, std::vector<Term>& terms, const std::vector<std::pair<int, int>>& req,
                  long* computed)
{
    std::vector<double> xs;
    std::vector<int> ms, owner;
    for (auto& r : req) {
        Term& t = terms[r.first];
        if (t.cached) continue;
        for (int m = (int)t.v.size(); m < r.second; m++) {
            xs.push_back(t.xi);
            ms.push_back(m);
            owner.push_back(r.first);
        }
    }
    if (xs.empty()) return;
    std::vector<double> out(xs.size());
    TERMINATE(caps_logdetD_batch(self, xs.data(), ms.data(), (int)xs.size(), out.data()) != 0, "batch failed");
    *computed += (long)xs.size();
    for (size_t i = 0; i < xs.size(); i++) terms[owner[i]].v.push_back(out[i]);
}

// Fourier-Chebyshev quadrature on [0, ∞) (Boyd, J. Sci. Comput. 2 (1987),
// eq. (3.2e); the reference's fcqs_semiinf, src/fcqs.c:107-175, boost L = 1):
// with M = N+1 and t_i = π i/M, i = 1..N,
//   x_i = cot²(t_i/2),
//   w_i = 8 sin t_i / ((1 - cos t_i)² M) Σ_{j odd <= N} sin(j t_i)/j.
// The nodes of M are a subset of those of 2M.
double fcqs_weight(double t, int N)
{
    double s = 0;
    for (int j = 1; j <= N; j += 2) s += sin(j * t) / j;
    const double omc = 1 - cos(t);
    return 8 * sin(t) / (omc * omc) / (N + 1) * s;
}

// T = 0: F (L+R)/(ħc) = ∫_0^∞ F_ξ(x/α) dx / (α π), α = 2L/(L+R), where
// F_ξ is the m-sum of one frequency (src/caps.c:235-266, 722-748).  For
// R/L > 10 the integrand is taken as 0 above the frequency where the PFA
// estimate of logdet falls below 1e-100 (src/caps.c:256-263).  The rule's
// levels M = 5, 10, 20, ... 2560 are refined until two successive levels
// agree to epsrel; every level's new nodes form one batch (modes 1, 2; with
// capsx_set_checkpoint, batches of at most that many nodes) or run one block
// at a time (mode 0).  Every F_ξ goes to the output file as soon as its
// batch is done.  With a time limit set by capsx_set_checkpoint, the driver
// stops before a batch (mode 0: before a node) once the limit has passed and
// returns NaN, after at least one batch per call; a later call that loads
// the output into the resume cache continues from there.
double free_energy_T0(caps_t* self, double cutoff, double epsrel, int mode, long* nblocks, long* nused)
{
    const int MMIN = 5, MMAX = 2560;
    const double LbyR = self->LbyR;
    const double alpha = 2 * LbyR / (1 + LbyR);
    const double xi_cut = LbyR < 0.1 ? -(1 + 1 / LbyR) * log(2 * LbyR * 1e-100) / 2 : INFINITY;
    long computed = 0, used = 0;
    std::vector<double> fcache(MMAX + 1, NAN);
    std::vector<int> mstop_at(MMAX + 1, -1);
    Ctx* cx = ctx_of(self);
    int ck_nodes;
    double ck_seconds;
    checkpoint_get(self, &ck_nodes, &ck_seconds);
    const double t_start = wall_seconds();
    int nbatches = 0;   // batches (mode 0: nodes) this call has run: at least one runs
    auto out_of_time = [&](int M, int done, int todo) {
        if (!(ck_seconds > 0) || nbatches == 0 || wall_seconds() - t_start < ck_seconds) return false;
        fprintf(stderr,
                "caps_free_energy: time limit of %g s reached at level M=%d after %d of its %d new nodes; "
                "the computed F_xi are in the output file\n",
                ck_seconds, M, done, todo);
        if (nblocks) *nblocks = computed;
        if (nused) *nused = used;
        return true;
    };
    // m-sums of the nodes ts[0..n) in one batch (modes 1, 2): first guess of
    // each node's m-range: with the surrogate (mode 2) its predicted
    // m-cutoff plus a margin, else the m-cutoff of the nearest known node at
    // a lower frequency (m-cutoffs fall with ξ at large ξ) plus a margin,
    // else 64; then every node is extended until its own cutoff is reached
    auto batch = [&](std::vector<Term>& ts) {
        std::vector<std::pair<int, int>> req;
        const int MCAP = std::min(self->ldim - 1, 4000), per = MCAP + 1;
        std::vector<float> P;
        std::vector<int> mpred;
        if (mode == 2) {
            TERMINATE(!cx->sur, "caps_free_energy mode 2 needs a surrogate (capsx_set_surrogate)");
            std::vector<double> xs;
            for (const Term& t : ts) xs.push_back(t.xi);
            sur_predict(self, xs, MCAP, P);
            for (size_t q = 0; q < ts.size(); q++) {
                const int ms = sur_mcut(&P[q * per], MCAP, cutoff);
                mpred.push_back(ms);
                req.push_back({(int)q, std::min(self->ldim * 8, ms + std::max(4, ms / 25) + 1)});
            }
        }
        for (size_t q = 0; q < ts.size() && mode != 2; q++) {
            int guess = -1;
            double best = -1;
            for (int idx = 1; idx <= MMAX; idx++) {
                if (mstop_at[idx] < 0) continue;
                const double t = M_PI * idx / MMAX, c = cos(t / 2) / sin(t / 2);
                const double xk = c * c / alpha;
                if (xk <= ts[q].xi && xk > best) {
                    best = xk;
                    guess = mstop_at[idx];
                }
            }
            req.push_back({(int)q, guess >= 0 ? guess + 8 + guess / 8 : 64});
        }
        run_requests(self, ts, req, &computed);
        cx->plan_batches++;
        for (size_t q = 0; q < ts.size(); q++) {
            Term& t = ts[q];
            int upto = (int)t.v.size();
            while (!msum(t.v, cutoff, &t.value, &t.mstop)) {
                TERMINATE(upto >= 100000, "sum over m did not converge at xi_=%g", t.xi);
                upto += 16;
                run_requests(self, ts, {{(int)q, upto}}, &computed);
                cx->plan_batches++;
            }
            if (mode == 2) sur_score(cx, &P[q * per], MCAP, mpred[q], t);
        }
    };
    double Ilast = NAN;
    for (int M = MMIN; M <= MMAX; M *= 2) {
        const int ratio = MMAX / M, N = M - 1;
        // new nodes of this level
        std::vector<Term> terms;
        std::vector<int> slot;
        for (int i = 1; i <= N; i++) {
            const int idx = ratio * i;
            if (!std::isnan(fcache[idx])) continue;
            const double t = M_PI * i / M;
            const double c = cos(t / 2) / sin(t / 2);
            const double xi = c * c / alpha;
            if (xi > xi_cut) {
                fcache[idx] = 0;
                mstop_at[idx] = -1;
                continue;
            }
            double cv;
            if (resume_lookup(self, xi, &cv)) {
                fcache[idx] = cv;
                mstop_at[idx] = -1;
                continue;
            }
            Term tm;
            tm.xi = xi;
            terms.push_back(tm);
            slot.push_back(idx);
        }
        auto finish = [&](size_t q) {
            fcache[slot[q]] = terms[q].value;
            mstop_at[slot[q]] = terms[q].mstop;
            used += terms[q].mstop + 1;
            resume_record(self, terms[q].xi, terms[q].value);
        };
        const int todo = (int)terms.size();
        if (mode == 0) {
            for (int q = 0; q < todo; q++) {
                if (out_of_time(M, q, todo)) return NAN;
                Term& t = terms[q];
                for (int m = 0;; m++) {
                    TERMINATE(m >= 100000, "sum over m did not converge at xi_=%g", t.xi);
                    t.v.push_back(caps_logdetD(self, t.xi, m));
                    computed++;
                    if (msum(t.v, cutoff, &t.value, &t.mstop)) break;
                }
                finish(q);
                nbatches++;
            }
        } else {
            const int step = ck_nodes > 0 ? ck_nodes : std::max(todo, 1);
            for (int q0 = 0; q0 < todo; q0 += step) {
                if (out_of_time(M, q0, todo)) return NAN;
                const int q1 = std::min(todo, q0 + step);
                std::vector<Term> ts(terms.begin() + q0, terms.begin() + q1);
                batch(ts);
                nbatches++;
                for (int q = q0; q < q1; q++) {
                    terms[q] = std::move(ts[q - q0]);
                    finish(q);
                }
            }
        }
        double I = 0;
        for (int i = 1; i <= N; i++) {
            const double f = fcache[ratio * i];
            TERMINATE(!std::isfinite(f), "integrand not finite at node %d of %d", i, N);
            I += fcqs_weight(M_PI * i / M, N) * f;
        }
        if (!std::isnan(Ilast) && fabs(1 - Ilast / I) < epsrel) {
            if (nblocks) *nblocks = computed;
            if (nused) *nused = used;
            return I / alpha / M_PI;
        }
        Ilast = I;
    }
    TERMINATE(true, "the xi-integral did not reach epsrel=%g with %d nodes", epsrel, MMAX - 1);
    return NAN;
}

}  // namespace

namespace {

double free_energy_msd(caps_t* self, double T_scaled, double cutoff, double epsrel, int mode, double v0,
                       long* nblocks, long* nused, long computed0);

// Eigenvalues below zero of the 2N x 2N symmetric tridiagonal matrix with
// zero diagonal and off-diagonal 1/sqrt((2j+1)(2j+3)), j = 1 .. 2N-1, by
// bisection on Sturm counts (each to full precision), ascending.
std::vector<double> psd_eigen_negative(int N)
{
    const int n = 2 * N;
    std::vector<double> e2(n);
    for (int j = 1; j < n; j++) e2[j] = 1.0 / ((2.0 * j + 1) * (2.0 * j + 3));
    auto count_below = [&](double x) {   // eigenvalues < x
        int c = 0;
        double q = -x;
        if (q < 0) c++;
        for (int i = 1; i < n; i++) {
            if (q == 0) q = 1e-300;
            q = -x - e2[i] / q;
            if (q < 0) c++;
        }
        return c;
    };
    std::vector<double> w(N);
    for (int k = 0; k < N; k++) {   // the (k+1)-th smallest eigenvalue
        double lo = -2, hi = 0;       // |λ| <= max row sum < 2/sqrt(15) + 1/sqrt(35) < 1
        for (int it = 0; it < 200 && hi - lo > 0; it++) {
            const double mid = 0.5 * (lo + hi);
            if (mid == lo || mid == hi) break;
            if (count_below(mid) > k) hi = mid;
            else lo = mid;
        }
        w[k] = 0.5 * (lo + hi);
    }
    return w;
}

// Residue η_j of the Padé pole ξ_j (z = -ξ_j²): the three-term recurrences
// of the numerator and denominator polynomials of the [N-1/N] Padé
// approximant, A_M = (2M+1) A_{M-1} + z A_{M-2}/4 (same for B, and for dB/dz
// by differentiation), from A_1 = 1/4, A_2 = 5/4, B_1 = 3, B_2 = 15 + z/4,
// rescaled every step; η = A/(2 dB/dz) at M = 2N (Hu, Xu, Yan, J. Chem.
// Phys. 133, 101106 (2010), eqs. (12)-(13); the reference's psd.c:20-62).
double psd_residue(int M2, double z)
{
    double a1 = 0.25, a2 = 1.25, b1 = 3, b2 = 15 + z / 4, d1 = 0, d2 = 0.25;
    for (int M = 3; M <= M2; M++) {
        const double a = (2 * M + 1) * a2 + z * a1 / 4;
        const double b = (2 * M + 1) * b2 + z * b1 / 4;
        const double d = (2 * M + 1) * d2 + b1 / 4 + z * d1 / 4;
        a1 = a2 / a;
        b1 = b2 / a;
        d1 = d2 / a;
        a2 = 1;
        b2 = b / a;
        d2 = d / a;
    }
    return a2 / (2 * d2);
}

}  // namespace

extern "C" {

// Extension: poles ξ_j (imaginary frequencies, ascending) and residues η_j
// of the Padé spectrum decomposition of order N of the Bose function,
// 1/(1 - e^{-x}) = 1/x + 1/2 + Σ_j 2 η_j x/(x² + ξ_j²) (Hu, Xu, Yan 2010):
// ξ_j = -2/w_j from the N negative eigenvalues w_j of the tridiagonal
// matrix of eq. (14).  Returns 0 (N <= 0: 1).  Reference: src/psd.c:80-136.
int capsx_psd(int N, double* xi, double* eta)
{
    if (N <= 0) return 1;
    const std::vector<double> w = psd_eigen_negative(N);
    for (int j = 0; j < N; j++) {
        xi[j] = -2 / w[j];
        eta[j] = psd_residue(2 * N, -xi[j] * xi[j]);
    }
    return 0;
}

}  // extern "C"

namespace {

// ξ = 0 term of the driver's sum (src/caps.c:270-300, 307-396): model 1
// Drude (caps_ht_drude); 2 perfect reflector and 3 plasma (omegap, rad/s):
// the Drude EE part plus Σ'_m of the MM blocks (m = 0 halved) up to the
// first m with v = 0 or v/v_0 < cutoff.
double xi0_term(caps_t* self, int model, double omegap, double cutoff, long* computed)
{
    const double drude = caps_ht_drude(self);
    if (model == 1) return drude;
    TERMINATE(model != 2 && model != 3, "unknown xi = 0 model %d", model);
    std::vector<double> v;
    for (int m = 0;; m++) {
        TERMINATE(m >= 100000, "xi = 0 sum over m did not converge");
        double x;
        if (model == 2) caps_logdetD0(self, m, 0, NULL, &x, NULL);
        else caps_logdetD0(self, m, omegap, NULL, NULL, &x);
        (*computed)++;
        v.push_back(x);
        if (x == 0 || x / v[0] < cutoff) break;
    }
    v[0] /= 2;
    return drude + kahan(v);
}

// T > 0 by the Padé spectrum decomposition (src/caps.c:811-832):
// F = T_scaled/π [v_0/2 + Σ_j η_j F_ξ(ξ_j T_scaled/(2π))]; the frequencies'
// m-sums planned as the Matsubara driver's (mode 0 plain, 1 batched with
// m-ranges from the previous pole, 2 surrogate-planned).
double free_energy_psd(caps_t* self, double T_scaled, double cutoff, int order, int mode, double v0,
                       long* nblocks, long* nused, long computed)
{
    Ctx* cx = ctx_of(self);
    std::vector<double> pole(order), eta(order);
    capsx_psd(order, pole.data(), eta.data());
    std::vector<Term> terms(order);
    for (int j = 0; j < order; j++) {
        terms[j].xi = pole[j] * T_scaled / (2 * M_PI);
        term_from_cache(self, terms[j]);
    }
    long used = 0;
    if (mode == 0) {
        for (Term& t : terms) {
            if (t.cached) {
                msum(t.v, cutoff, &t.value, &t.mstop);
                continue;
            }
            for (int m = 0;; m++) {
                TERMINATE(m >= 100000, "sum over m did not converge at xi_=%g", t.xi);
                t.v.push_back(caps_logdetD(self, t.xi, m));
                computed++;
                if (msum(t.v, cutoff, &t.value, &t.mstop)) break;
            }
        }
    } else {
        std::vector<std::pair<int, int>> req;
        const int MCAP = std::min(self->ldim - 1, 4000), per = MCAP + 1;
        std::vector<float> P;
        std::vector<int> mpred(order, 64);
        if (mode == 2) {
            TERMINATE(!cx->sur, "mode 2 needs a surrogate (capsx_set_surrogate)");
            std::vector<double> xs;
            for (const Term& t : terms) xs.push_back(t.xi);
            sur_predict(self, xs, MCAP, P);
            for (int j = 0; j < order; j++) mpred[j] = sur_mcut(&P[(size_t)j * per], MCAP, cutoff);
        }
        // mode 1: the first pole in a growing batch, then the others with
        // the m-range of the pole before (m-cutoffs fall with ξ)
        int j0 = 0;
        if (mode == 1) {
            int upto = 64;
            while (true) {
                run_requests(self, terms, {{0, upto}}, &computed);
                cx->plan_batches++;
                if (msum(terms[0].v, cutoff, &terms[0].value, &terms[0].mstop)) break;
                upto += 32;
            }
            j0 = 1;
        }
        for (int j = j0; j < order; j++) {
            const int ms = mode == 2 ? mpred[j] : terms[0].mstop;
            req.push_back({j, std::min(self->ldim * 8, ms + std::max(4, ms / 25) + 1)});
        }
        run_requests(self, terms, req, &computed);
        cx->plan_batches++;
        for (int j = 0; j < order; j++) {
            Term& t = terms[j];
            int upto = (int)t.v.size();
            while (!msum(t.v, cutoff, &t.value, &t.mstop)) {
                TERMINATE(upto >= 100000, "sum over m did not converge at xi_=%g", t.xi);
                upto += 16;
                run_requests(self, terms, {{j, upto}}, &computed);
                cx->plan_batches++;
            }
            if (mode == 2 && !t.cached) sur_score(cx, &P[(size_t)j * per], MCAP, mpred[j], t);
        }
    }
    std::vector<double> v{v0 / 2};
    for (int j = 0; j < order; j++) {
        v.push_back(eta[j] * terms[j].value);
        if (terms[j].cached) continue;
        used += terms[j].mstop + 1;
        resume_record(self, terms[j].xi, terms[j].value);
    }
    if (nblocks) *nblocks = computed;
    if (nused) *nused = used;
    return T_scaled / M_PI * kahan(v);
}

}  // namespace

extern "C" {

// Extension: the free energy with the reference driver's options.  xi0:
// ξ = 0 term (1 Drude, 2 perfect reflector, 3 plasma with omegap in rad/s);
// psd_order: 0 Matsubara sum, > 0 Padé spectrum decomposition of that
// order, < 0 the driver's automatic order ceil((1 - 1.5 log10 epsrel)/
// sqrt(T_eff)), T_eff = 4π k_B T L/(ħc) (src/caps.c:635-639).  T = 0
// integrates over ξ (xi0 and psd_order unused).
double capsx_free_energy(caps_t* self, double T, double cutoff, double epsrel, int mode, int xi0, double omegap,
                         int psd_order, long* nblocks, long* nused)
{
    TERMINATE(!(T >= 0), "the free energy needs T >= 0 (got %g)", T);
    if (T == 0) return free_energy_T0(self, cutoff, epsrel, mode, nblocks, nused);
    long computed = 0;
    double v0;
    if (!resume_lookup(self, 0, &v0)) {
        v0 = xi0_term(self, xi0, omegap, cutoff, &computed);
        resume_record(self, 0, v0);
    }
    const double T_scaled = 2 * M_PI * CAPS_kB * (self->R + self->L) * T / (CAPS_hbar * CAPS_c);
    if (psd_order == 0) return free_energy_msd(self, T_scaled, cutoff, epsrel, mode, v0, nblocks, nused, computed);
    if (psd_order < 0) {
        const double Teff = 4 * M_PI * CAPS_kB / CAPS_hbar / CAPS_c * T * self->L;
        psd_order = (int)ceil((1 - 1.5 * log10(epsrel)) / sqrt(Teff));
    }
    return free_energy_psd(self, T_scaled, cutoff, psd_order, mode, v0, nblocks, nused, computed);
}

// Extension: the resume cache of the free-energy drivers, read from an
// earlier run's output as the reference driver's --resume does
// (src/caps.c:72-101): every line holding "# xi*(L+R)/c=" and "logdetD="
// gives one (ξ_, F_ξ) pair.  The drivers then take F_ξ of a matching
// frequency from the cache instead of computing its blocks.  Appends to the
// object's cache; path NULL or "" clears it.  Returns the number of pairs
// read, -1 if the file cannot be opened, -2 if a line has the wrong format
// (the reference's two errors, src/caps.c:78, 90, 96).
int capsx_resume_load(caps_t* self, const char* path)
{
    std::vector<std::pair<double, double>> got;
    if (path && path[0]) {
        FILE* fh = fopen(path, "r");
        if (!fh) {
            fprintf(stderr, "cannot open %s for reading\n", path);
            return -1;
        }
        char line[512];
        while (fgets(line, sizeof(line), fh)) {
            char* p1 = strstr(line, "# xi*(L+R)/c=");
            char* p2 = strstr(line, "logdetD=");
            if (!p1 || !p2) continue;
            p1 += 13;
            char* e1 = strchr(p1, ',');
            p2 += 8;
            char* e2 = strchr(p2, ',');
            if (!e1 || !e2) {
                fclose(fh);
                fprintf(stderr, "%s has wrong format\n", path);
                return -2;
            }
            *e1 = '\0';
            *e2 = '\0';
            got.emplace_back(atof(p1), atof(p2));
        }
        fclose(fh);
    }
    std::lock_guard<std::mutex> lock(g_resume_mu);
    ResumeState& st = g_resume[self];
    if (!path || !path[0]) st.cache.clear();
    st.cache.insert(st.cache.end(), got.begin(), got.end());
    return (int)got.size();
}

// Extension: F_ξ of frequency xi_ from the resume cache, with the
// reference's matching rule (src/caps.c:316-327).  Returns 1 and sets
// *value if found, else 0.
int capsx_resume_lookup(caps_t* self, double xi_, double* value)
{
    double v;
    if (!resume_lookup(self, xi_, &v)) return 0;
    if (value) *value = v;
    return 1;
}

// Extension: append a line "# xi*(L+R)/c=<ξ_>, logdetD=<F_ξ>, t=<s>" for
// every F_ξ the free-energy drivers compute (src/caps.c:265) to path (opened
// for appending; the file capsx_resume_load reads back); NULL or "" closes
// it.  Returns 0, or -1 if the file cannot be opened.
int capsx_set_output(caps_t* self, const char* path)
{
    std::lock_guard<std::mutex> lock(g_resume_mu);
    ResumeState& st = g_resume[self];
    if (st.out) fclose(st.out);
    st.out = nullptr;
    if (!path || !path[0]) return 0;
    st.out = fopen(path, "a");
    if (!st.out) return -1;
    st.t0 = wall_seconds();
    return 0;
}

// Extension: checkpointing of the T = 0 free energy, for runs longer than
// one job: at most `nodes` quadrature nodes per batch (0: every new node of
// a refinement level in one batch, the default), and a time limit in
// seconds (0: none) after which caps_free_energy stops before its next batch
// and returns NaN (every call runs at least one batch).  Each F_ξ is written to the output file (capsx_set_output)
// when its batch is done; the next job loads that file with
// capsx_resume_load and continues.  Returns 0, or -1 for negative values.
int capsx_set_checkpoint(caps_t* self, int nodes, double seconds)
{
    if (nodes < 0 || !(seconds >= 0)) return -1;
    std::lock_guard<std::mutex> lock(g_resume_mu);
    ResumeState& st = g_resume[self];
    st.nodes = nodes;
    st.seconds = seconds;
    return 0;
}

double caps_free_energy(caps_t* self, double T, double cutoff, double epsrel, int mode, int drude_xi0,
                        long* nblocks, long* nused)
{
    TERMINATE(!(T >= 0), "caps_free_energy needs T >= 0 (got %g)", T);
    if (T == 0) return free_energy_T0(self, cutoff, epsrel, mode, nblocks, nused);
    TERMINATE(!drude_xi0, "caps_free_energy: for other xi = 0 models use capsx_free_energy");
    return capsx_free_energy(self, T, cutoff, epsrel, mode, 1, 0, 0, nblocks, nused);
}

}  // extern "C"

namespace {

// T > 0 by the Matsubara sum (src/caps.c:836-846) with the ξ = 0 term v0.
double free_energy_msd(caps_t* self, double T_scaled, double cutoff, double epsrel, int mode, double v0,
                       long* nblocks, long* nused, long computed0)
{
    long computed = computed0, used = 0;
    std::vector<double> v;      // Matsubara terms, v[0] = ξ = 0 term
    v.push_back(v0);
    std::vector<Term> terms(1);  // terms[n] for n >= 1
    auto term = [&](int n) -> Term& {
        while ((int)terms.size() <= n) {
            Term t;
            t.xi = terms.size() * T_scaled;
            term_from_cache(self, t);
            terms.push_back(t);
        }
        return terms[n];
    };

    if (mode == 0) {
        // plain: one block at a time in F_xi order (src/caps.c:307-396, 836-846)
        for (int n = 1;; n++) {
            Term& t = term(n);
            for (int m = 0; !t.cached; m++) {
                TERMINATE(m >= 100000, "sum over m did not converge at xi_=%g", t.xi);
                t.v.push_back(caps_logdetD(self, t.xi, m));
                computed++;
                if (msum(t.v, cutoff, &t.value, &t.mstop)) break;
            }
            if (t.cached) {
                msum(t.v, cutoff, &t.value, &t.mstop);
            } else {
                used += t.mstop + 1;
                resume_record(self, t.xi, t.value);
            }
            v.push_back(t.value);
            if (fabs(v[n] / v[0]) < epsrel) break;
            TERMINATE(n > 100000, "Matsubara sum did not converge");
        }
    } else {
        // batched (mode 1): terms 1 and 2 in growing chunks; from their
        // m-cutoffs and values, one batch with all further terms up to the
        // predicted end of the Matsubara sum.  Surrogate-planned (mode 2):
        // the MLP predicts log|logdetD(ξ_n, m)| for all candidate blocks, the
        // predicted m-cutoffs and end of the Matsubara sum give one batch
        // for the whole energy.  In both, every term is then checked against
        // the same cutoff rules in order and extended if the plan fell
        // short, so the summed blocks are those of the plain form.
        Ctx* cx = ctx_of(self);
        auto finish_term = [&](int n, int first, int step) {
            Term& t = term(n);
            int upto = first;
            while (true) {
                run_requests(self, terms, {{n, upto}}, &computed);
                cx->plan_batches++;
                if (msum(t.v, cutoff, &t.value, &t.mstop)) break;
                TERMINATE(upto >= 100000, "sum over m did not converge at xi_=%g", t.xi);
                upto += step;
            }
            t.done = true;
        };
        int n = 0;
        bool stop = false;
        auto accept = [&](int j) {
            Term& t = term(j);
            if (!t.cached) {
                used += t.mstop + 1;
                resume_record(self, t.xi, t.value);
            }
            v.push_back(t.value);
            n = j;
            stop = fabs(v[j] / v[0]) < epsrel;
        };
        if (mode == 2) {
            TERMINATE(!cx->sur, "caps_free_energy mode 2 needs a surrogate (capsx_set_surrogate)");
            const int NCAP = 48;
            const int MCAP = std::min(self->ldim - 1, 4000);
            const int per = MCAP + 1;
            std::vector<double> xs;
            for (int j = 1; j <= NCAP; j++) xs.push_back(j * T_scaled);
            std::vector<float> P;
            sur_predict(self, xs, MCAP, P);
            std::vector<std::pair<int, int>> req;
            std::vector<int> mpred(NCAP + 1, 0);
            for (int j = 1; j <= NCAP; j++) {
                const float* pr = &P[(size_t)(j - 1) * per];
                const int ms = sur_mcut(pr, MCAP, cutoff);
                mpred[j] = ms;
                double vj = exp((double)pr[0]) / 2;
                for (int m = 1; m <= ms; m++) vj += exp((double)pr[m]);
                const int margin = std::max(4, ms / 25);
                term(j);
                req.push_back({j, std::min(self->ldim * 8, ms + margin + 1)});
                if (vj / fabs(v[0]) < epsrel) break;
            }
            run_requests(self, terms, req, &computed);
            cx->plan_batches++;
            for (size_t r = 0; r < req.size() && !stop; r++) {
                const int j = req[r].first;
                Term& t = term(j);
                if (!msum(t.v, cutoff, &t.value, &t.mstop)) finish_term(j, (int)t.v.size() + 16, 16);
                if (!t.cached) sur_score(cx, &P[(size_t)(j - 1) * per], MCAP, mpred[j], t);
                accept(j);
            }
        }
        if (!stop && n < 1) {
            finish_term(1, 64, 32);
            accept(1);
        }
        // m-cutoff of term j for planning (a resumed term's is not known)
        auto mcut = [&](int j) { return term(j).cached ? 64 : term(j).mstop; };
        if (!stop && n < 2) {
            finish_term(2, mcut(1) + 12, 16);
            accept(2);
        }
        while (!stop) {
            if (term(n + 1).cached) {
                // a resumed term: taken in order, no batch planned past it
                Term& t = term(n + 1);
                msum(t.v, cutoff, &t.value, &t.mstop);
                accept(n + 1);
                continue;
            }
            // prediction: geometric decay of v_n, m-cutoff growing by at most
            // the last step, 8 blocks of margin, one extra term
            const double r = std::min(0.95, std::max(1e-6, fabs(v[n] / v[n - 1])));
            const double need = epsrel * fabs(v[0]) / fabs(v[n]);
            int more = need < 1 ? (int)ceil(log(need) / log(r)) + 1 : 1;
            more = std::max(1, std::min(more, 64));
            const int dm = std::max(0, mcut(n) - mcut(n - 1));
            const int mp = mcut(n) + dm + 8;
            std::vector<std::pair<int, int>> req;
            for (int j = n + 1; j <= n + more; j++) {
                term(j);
                req.push_back({j, mp + 1});
            }
            run_requests(self, terms, req, &computed);
            cx->plan_batches++;
            const int n0 = n;
            for (int j = n0 + 1; j <= n0 + more && !stop; j++) {
                Term& t = term(j);
                if (!msum(t.v, cutoff, &t.value, &t.mstop)) finish_term(j, (int)t.v.size() + 16, 16);
                accept(j);
            }
            TERMINATE(n > 100000, "Matsubara sum did not converge");
        }
    }
    v[0] /= 2;
    const double s = kahan(v);
    if (nblocks) *nblocks = computed;
    if (nused) *nused = used;
    return T_scaled / M_PI * s;
}

}  // namespace

extern "C" {

// ---- LU of user matrices and the MFMA probe (test hooks) --------------------

// Extension (test hook): log|det A| of n column-major dim x dim matrices
// (lda = dim) on the GPU's batched LU.  Returns 0 on success.
int capsx_lu_logdet_batch(const double* A, int dim, int n, double* out)
{
    if (dim <= 0 || dim > capsd_lu_max_dim() || n <= 0) return 1;
    caps_t* tmp = caps_init(1, 1);
    Ctx* c = ctx_of(tmp);
    ensure_blocks(c, n);
    const int lda = dim;
    int64_t off = 0;
    for (int b = 0; b < n; b++) {
        BlockDesc& d = c->h_desc[b];
        memset(&d, 0, sizeof(d));
        d.N = dim;
        d.ldim = (dim + 1) / 2;
        d.lda = lda;
        d.off_mat = off;
        off = align256(off + (int64_t)lda * dim * 8);
        LuSide& s = c->h_side[b];
        off = lu_side_layout(s, off, dim, lda);
    }
    ensure_arena(c, off);
    std::vector<BlockOut> o(n);
    memset(o.data(), 0, n * sizeof(BlockOut));
    HIPCHECK(hipMemcpyAsync(c->d_desc, c->h_desc, n * sizeof(BlockDesc), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(c->d_side, c->h_side, n * sizeof(LuSide), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(c->d_out, o.data(), n * sizeof(BlockOut), hipMemcpyHostToDevice, c->stream));
    for (int b = 0; b < n; b++)
        HIPCHECK(hipMemcpyAsync(c->arena + c->h_desc[b].off_mat, A + (size_t)b * dim * dim,
                                (size_t)dim * dim * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipEventRecord(c->ev[0], c->stream));
    launch_lu_groups(c, n, dim, lu_config(dim, n), c->ev[0], false, 1);
    HIPCHECK(hipEventRecord(c->ev[1], c->stream));
    HIPCHECK(hipMemcpyAsync(o.data(), c->d_out, n * sizeof(BlockOut), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    float ms;
    HIPCHECK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    for (int b = 0; b < n; b++) out[b] = o[b].logdet;
    caps_free(tmp);
    return 0;
}

// Extension (test hook): log|det(1 - M)| of n column-major dim x dim
// matrices M (lda = dim) through the device's column dominance test and LU:
// -M is uploaded with the off-diagonal column sums of |M|, k_dominance
// decides the form (delta[b] = 1: no pivot search, log1p of u_ii - 1; 0: the
// unit diagonal written back, partial pivoting).  No trace shortcut.
// Returns 0 on success.
int capsx_logdet_1mM_batch(const double* M, int dim, int n, double* out, int* delta)
{
    if (dim <= 0 || dim > capsd_lu_max_dim() || n <= 0) return 1;
    caps_t* tmp = caps_init(1, 1);
    Ctx* c = ctx_of(tmp);
    ensure_blocks(c, n);
    const int lda = (dim + 15) & ~15;
    int64_t off = 0;
    for (int b = 0; b < n; b++) {
        BlockDesc& d = c->h_desc[b];
        memset(&d, 0, sizeof(d));
        d.N = dim;
        d.ldim = (dim + 1) / 2;
        d.lda = lda;
        d.off_mat = off;
        off = align256(off + (int64_t)lda * dim * 8);
        d.n_rtiles = 1;
        d.off_cpart = off;
        off = align256(off + (int64_t)dim * 8);
        LuSide& s = c->h_side[b];
        off = lu_side_layout(s, off, dim, lda);
    }
    ensure_arena(c, off);
    std::vector<BlockOut> o(n);
    memset(o.data(), 0, n * sizeof(BlockOut));
    for (int b = 0; b < n; b++) o[b].delta = 1;
    HIPCHECK(hipMemcpyAsync(c->d_desc, c->h_desc, n * sizeof(BlockDesc), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(c->d_side, c->h_side, n * sizeof(LuSide), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(c->d_out, o.data(), n * sizeof(BlockOut), hipMemcpyHostToDevice, c->stream));
    std::vector<double> A((size_t)lda * dim, 0.0), cs(dim);
    for (int b = 0; b < n; b++) {
        const double* Mb = M + (size_t)b * dim * dim;
        for (int j = 0; j < dim; j++) {
            double sj = 0;
            for (int i = 0; i < dim; i++) {
                const double v = Mb[i + (size_t)j * dim];
                A[i + (size_t)j * lda] = -v;
                if (i != j) sj += fabs(v);
            }
            cs[j] = sj;
        }
        const BlockDesc& d = c->h_desc[b];
        HIPCHECK(hipMemcpy(c->arena + d.off_mat, A.data(), A.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(c->arena + d.off_cpart, cs.data(), dim * sizeof(double), hipMemcpyHostToDevice));
    }
    capsd_launch_dominance(c->d_desc, c->arena, c->d_out, n, dim, g_delta_mode, c->stream);
    HIPCHECK(hipEventRecord(c->ev[0], c->stream));
    launch_lu_groups(c, n, dim, lu_config(dim, n), c->ev[0], false, 3);
    HIPCHECK(hipMemcpyAsync(o.data(), c->d_out, n * sizeof(BlockOut), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    HIPCHECK(hipGetLastError());
    for (int b = 0; b < n; b++) {
        out[b] = o[b].logdet;
        if (delta) delta[b] = o[b].delta;
    }
    caps_free(tmp);
    return 0;
}

// Extension without a counterpart in the reference, for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops):
// capsx_lu_logdet_batch on device data, allocating nothing.  dA holds n
// column-major dim x dim matrices with leading dimension lda (matrix b at
// dA + b lda dim), d_logdet receives n values log|det A_b|, both device
// pointers; the work is enqueued on `stream` (a hipStream_t, NULL for the
// null stream) and the call returns without waiting.  The matrices are
// overwritten.  Descriptors and the LU's side buffers (panels, pivots, U
// rows: about 8 lda 64 doubles per matrix) live in a fixed 64 MiB device
// array per GPU, set up by a kernel (no host copies); calls that overlap on
// one GPU share it and race.  Returns 0; 1 for invalid arguments, 2 if the
// array's address cannot be had, 3 if the batch needs more than the array
// holds (nothing is launched then), 4 for a launch error.
int capsx_lu_logdet_batch_noalloc(double* dA, int dim, int lda, int n, double* d_logdet, void* stream)
{
    if (!dA || !d_logdet || dim <= 0 || dim > capsd_lu_max_dim() || n <= 0 || lda < dim) {
        fprintf(stderr, "capsx_lu_logdet_batch_noalloc: invalid arguments (dim=%d, lda=%d, n=%d)\n", dim, lda, n);
        return 1;
    }
    char* base = nullptr;
    size_t cap = 0;
    if (capsd_noalloc_scratch(&base, &cap) != 0) {
        fprintf(stderr, "capsx_lu_logdet_batch_noalloc: no scratch address on this device\n");
        return 2;
    }
    LuSide rel;
    const int64_t per = lu_side_layout(rel, 0, dim, lda);   // side buffers of one matrix, offsets from 0
    const int64_t off_side = align256((int64_t)n * sizeof(BlockDesc));
    const int64_t off_out = off_side + align256((int64_t)n * sizeof(LuSide));
    const int64_t off_buf = off_out + align256((int64_t)n * sizeof(BlockOut));
    // + 256: the trailing update's 16-byte operand loads may read one
    // element past the last U buffer (odd lda), which must stay in the array
    const int64_t need = off_buf + (int64_t)n * per + 256;
    if (need > (int64_t)cap) {
        fprintf(stderr,
                "capsx_lu_logdet_batch_noalloc: %d matrices of dimension %d need %lld bytes of scratch, the "
                "fixed array holds %zu\n",
                n, dim, (long long)need, cap);
        return 3;
    }
    hipStream_t s = (hipStream_t)stream;
    const int64_t a_off = (int64_t)((char*)dA - base);
    capsd_launch_noalloc_setup(base, a_off, dim, lda, n, off_side, off_out, off_buf, per, rel.off_logs,
                               rel.off_moved, rel.off_ut, rel.off_inv, s);
    const LuCfg cfg = lu_config(dim, n);
    capsd_launch_lu((const BlockDesc*)base, base, base + off_side, (BlockOut*)(base + off_out), n, dim, s, cfg.levels,
                    cfg.xcd, cfg.tile, nullptr, nullptr, 1);
    capsd_launch_noalloc_result(base + off_out, d_logdet, n, s);
    return hipGetLastError() == hipSuccess ? 0 : 4;
}

// Extension (autotuning): milliseconds of one batched LU of n synthetic
// dim x dim matrices with the given schedule (a value < 0 keeps the
// configured one), averaged over reps; the matrices are regenerated before
// each repetition, outside the timed region.  form 0: A = 1 - 0.3 X/sqrt(dim)
// (X uniform in [-1, 1)), the pivoted LU; form 1: 1 - M with M = 0.3 X/dim,
// column diagonally dominant, the delta form (the path of the round-trip
// blocks).  Returns 0 on success.
int capsx_lu_time_form(int dim, int n, int reps, int levels, int streams, int xcd, int tile, int form, double* ms_out)
{
    if (dim <= 0 || dim > capsd_lu_max_dim() || n <= 0 || reps <= 0 || form < 0 || form > 1) return 1;
    caps_t* tmp = caps_init(1, 1);
    Ctx* c = ctx_of(tmp);
    ensure_blocks(c, n);
    const int lda = (dim + 15) & ~15;
    int64_t off = 0;
    for (int b = 0; b < n; b++) {
        BlockDesc& d = c->h_desc[b];
        memset(&d, 0, sizeof(d));
        d.N = dim;
        d.ldim = (dim + 1) / 2;
        d.lda = lda;
        d.off_mat = off;
        off = align256(off + (int64_t)lda * dim * 8);
        LuSide& sd = c->h_side[b];
        off = lu_side_layout(sd, off, dim, lda);
    }
    ensure_arena(c, off);
    HIPCHECK(hipMemcpyAsync(c->d_desc, c->h_desc, n * sizeof(BlockDesc), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipMemcpyAsync(c->d_side, c->h_side, n * sizeof(LuSide), hipMemcpyHostToDevice, c->stream));
    for (int b = 0; b < n; b++) {
        memset(&c->h_out[b], 0, sizeof(BlockOut));
        c->h_out[b].delta = form;
    }
    LuCfg cfg = lu_config(dim, n);
    if (levels > 0) cfg.levels = std::min(levels, 4);
    if (streams > 0) cfg.streams = streams;
    if (xcd >= 0) cfg.xcd = xcd ? 1 : 0;
    if (tile > 0) cfg.tile = tile <= 64 ? 64 : tile >= 256 ? 256 : 128;
    double tot = 0;
    for (int r = 0; r < reps; r++) {
        HIPCHECK(hipMemcpyAsync(c->d_out, c->h_out, n * sizeof(BlockOut), hipMemcpyHostToDevice, c->stream));
        capsd_launch_lu_fill(c->d_desc, c->arena, n, 1234ull + r, form, c->stream);
        HIPCHECK(hipEventRecord(c->ev[0], c->stream));
        launch_lu_groups(c, n, dim, cfg, c->ev[0], false, form ? 2 : 1);
        HIPCHECK(hipEventRecord(c->ev[1], c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
        HIPCHECK(hipGetLastError());
        float ms;
        HIPCHECK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
        tot += ms;
    }
    *ms_out = tot / reps;
    caps_free(tmp);
    return 0;
}

int capsx_lu_time(int dim, int n, int reps, int levels, int streams, int xcd, int tile, double* ms_out)
{
    return capsx_lu_time_form(dim, n, reps, levels, streams, xcd, tile, 0, ms_out);
}

// Extension: capture each batch shape of this object into a HIP graph and
// replay it for later batches of the same shape (on != 0), or launch
// eagerly (0, the default).  Stage timings of replayed batches are 0 (the
// total is measured).
void capsx_set_graph(caps_t* self, int on)
{
    Ctx* c = ctx_of(self);
    std::lock_guard<std::mutex> lock(c->mu);
    c->graph = on ? 1 : 0;
    if (!on) {
        for (auto& g : c->graphs) HIPCHECK(hipGraphExecDestroy(g.second));
        c->graphs.clear();
    }
}

// Extension: fix the LU schedule of all later batches (levels 1|2|3|4, streams
// >= 1, xcd 0|1, tile 64|128|256); levels = 0 returns to the autotuned table.
// Returns 0 on success, 1 on invalid values.
int capsx_lu_config(int levels, int streams, int xcd, int tile)
{
    std::lock_guard<std::mutex> lock(g_cfg_mu);
    if (levels == 0) {
        g_have_override = false;
        return 0;
    }
    if (levels < 1 || levels > 4 || streams < 1 || (xcd != 0 && xcd != 1) ||
        (tile != 64 && tile != 128 && tile != 256))
        return 1;
    g_override.levels = levels;
    g_override.streams = streams;
    g_override.xcd = xcd;
    g_override.tile = tile;
    g_have_override = true;
    return 0;
}

// Extension: replace the autotuned table by the one in `path` (lines "maxN n
// levels streams xcd tile", '#' comments).  Returns the entries read.
int capsx_lu_tuning_load(const char* path)
{
    load_default_tuning();
    return load_tuning(path);
}

// Extension: panels taller than `rows` are factored column by column by
// grid-wide kernels instead of one workgroup per matrix (default 6144);
// rows <= 0 only queries.  Returns the previous value.
int capsx_lu_grid_panel_rows(int rows)
{
    device_init();
    return capsd_lu_grid_panel_rows(rows);
}

// Extension: the LU's kernel variants (trailing update, panel set, narrow
// tiles); values < 0 keep the current ones, prev receives the old ones.
int capsx_lu_kernels(int gemm_kernel, int panel_set, int narrow_tiles, int prev[3])
{
    device_init();
    return capsd_lu_kernels(gemm_kernel, panel_set, narrow_tiles, prev);
}

// Extension (test hook): 1 = the determinant's delta form for column
// dominant blocks (the default), 0 = the pivoted LU for every block.
// Returns the previous setting.
int capsx_lu_delta_mode(int mode)
{
    std::lock_guard<std::mutex> lock(g_cfg_mu);
    const int old = g_delta_mode;
    if (mode == 0 || mode == 1) g_delta_mode = mode;
    return old;
}

// Extension: the schedule used for n blocks of largest dimension maxN:
// cfg = {levels, streams, xcd, tile}.
void capsx_lu_config_get(int maxN, int n, int cfg[4])
{
    const LuCfg c = lu_config(maxN, n);
    cfg[0] = c.levels;
    cfg[1] = c.streams;
    cfg[2] = c.xcd;
    cfg[3] = c.tile;
}

// Extension (test hook): D = A(16x4) B(4x16) through one v_mfma_f64_16x16x4f64.
int capsx_mfma_probe(const double* A, const double* B, double* D)
{
    device_init();
    double *dA, *dB, *dD;
    HIPCHECK(hipMalloc(&dA, 64 * 8));
    HIPCHECK(hipMalloc(&dB, 64 * 8));
    HIPCHECK(hipMalloc(&dD, 256 * 8));
    HIPCHECK(hipMemcpy(dA, A, 64 * 8, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(dB, B, 64 * 8, hipMemcpyHostToDevice));
    capsd_launch_mfma_probe(dA, dB, dD, 0);
    HIPCHECK(hipDeviceSynchronize());
    HIPCHECK(hipMemcpy(D, dD, 256 * 8, hipMemcpyDeviceToHost));
    hipFree(dA);
    hipFree(dB);
    hipFree(dD);
    return 0;
}

// Extension (test hook): the K table (log|K_ν^p|, ν = mu..nu_max) and the I
// table (log|I|, sign) of one block, from the values the kernels compute.
// The K table (log|K|) and the full I table (log|I|, sign; the device keeps
// the lower triangle, the upper one is its mirror image) of block (xi_, m)
// as computed on the device.
static void block_tables_host(caps_t* self, double xi_, int m, IntegTables& t)
{
    Ctx* c = ctx_of(self);
    std::lock_guard<std::mutex> lock(c->mu);
    ensure_blocks(c, 1);
    BatchPlan p = plan_batch(self, c, &xi_, &m, 1, 0);
    run_batch(self, c, p, 1, 0);
    const BlockDesc& d = c->h_desc[0];
    t.m = m;
    t.is_pr = d.is_pr != 0;
    t.mu = d.mu;
    t.nK = d.nK;
    t.lo = d.lo;
    t.nl = d.nl;
    t.lnK.resize(2 * (size_t)d.nK);
    HIPCHECK(hipMemcpy(t.lnK.data(), c->arena + d.off_k + 2 * sizeof(capsd::KEnt) * (int64_t)d.nK,
                       2 * d.nK * sizeof(double), hipMemcpyDeviceToHost));
    // the table holds I = mant 2^exp; kept as log|I| and sign
    const size_t nl_ = (size_t)d.nl, nn = 2 * nl_ * nl_;
    std::vector<int> ex(nn);
    t.logI.resize(nn);
    t.sgnI.resize(nn);
    HIPCHECK(hipMemcpy(t.logI.data(), c->arena + d.off_i, nn * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(ex.data(), c->arena + d.off_i + 16 * (int64_t)d.nl * d.nl, nn * sizeof(int),
                       hipMemcpyDeviceToHost));
    for (size_t i = 0; i < nn; i++) {
        const double mt = t.logI[i];
        t.sgnI[i] = (signed char)((mt > 0) - (mt < 0));
        t.logI[i] = mt == 0 ? -INFINITY : log(fabs(mt)) + ex[i] * M_LN2;
    }
    for (size_t pol = 0; pol < 2; pol++)
        for (size_t r = 0; r < nl_; r++)
            for (size_t col = r + 1; col < nl_; col++) {
                const size_t i = pol * nl_ * nl_ + r * nl_ + col, src = pol * nl_ * nl_ + col * nl_ + r;
                t.logI[i] = t.logI[src];
                t.sgnI[i] = t.sgnI[src];
            }
}

int capsx_block_tables(caps_t* self, double xi_, int m, int* mu, int* nK, double* lnK /* 2*nK */, int* lo,
                       int* nl, double* logI /* 2*nl*nl */, signed char* sgnI /* 2*nl*nl */, int cap_nK, int cap_nl)
{
    IntegTables t;
    block_tables_host(self, xi_, m, t);
    *mu = t.mu;
    *nK = t.nK;
    *lo = t.lo;
    *nl = t.nl;
    if (t.nK > cap_nK || t.nl > cap_nl) return 1;
    if (lnK) memcpy(lnK, t.lnK.data(), t.lnK.size() * sizeof(double));
    if (logI) {
        memcpy(logI, t.logI.data(), t.logI.size() * sizeof(double));
        memcpy(sgnI, t.sgnI.data(), t.sgnI.size());
    }
    return 0;
}

// ---- the k-integrals of one block: src/include/integration.h:13-23 --------
// The reference evaluates K_ν and I_{l1 l2} on demand by quadrature and caches
// them; here caps_integrate_init computes the block's whole K and I tables on
// the device (the tables caps_logdetD builds: ν from 2m, or 2 for m = 0, up
// to 2(lmin + ldim) + 1, and l1, l2 over the block's ℓ-window with one more
// on either side), and the functions read them.  epsrel is accepted and
// unused (the device quadrature has a fixed accuracy, 1e-11 relative).
// Values outside the tables abort with a message.

// Reference: src/integration.c:695-731 (NULL for xi_ < 0, m < 0 or
// epsrel <= 0; here also for xi_ = 0, which has no device tables).
integration_t* caps_integrate_init(caps_t* caps, double xi_, int m, double epsrel)
{
    if (!caps || !(xi_ > 0) || m < 0 || !(epsrel > 0)) return NULL;
    TERMINATE(2 * caps->ldim > capsd_lu_max_dim(), "ldim=%d exceeds the maximum %d", caps->ldim,
              capsd_lu_max_dim() / 2);
    caps_block_state* st = (caps_block_state*)calloc(1, sizeof(caps_block_state));
    TERMINATE(!st, "out of memory");
    st->tab = new IntegTables();
    block_tables_host(caps, xi_, m, *st->tab);
    return st;
}

// Reference: src/integration.c:737-746.
void caps_integrate_free(integration_t* integration)
{
    if (!integration) return;
    delete integration->tab;
    free(integration->M);
    free(integration);
}

static const IntegTables& integ_tables(const integration_t* self)
{
    TERMINATE(!self || !self->tab, "not an object of caps_integrate_init");
    return *self->tab;
}

// log|K_nu^p| and its sign (TE negative, TM positive: the sign of r_p).
// Reference: src/integration.c:452-490.
double caps_integrate_K(integration_t* self, int nu, polarization_t p, sign_t* sign)
{
    const IntegTables& t = integ_tables(self);
    TERMINATE(nu < t.mu || nu >= t.mu + t.nK, "K: nu=%d outside the table [%d, %d]", nu, t.mu, t.mu + t.nK - 1);
    // perfect reflectors: r_TE = -r_TM, K_TE = -K_TM (the device computes TM)
    const int pol = (p == TE && !t.is_pr) ? 0 : 1;
    const double v = t.lnK[(size_t)pol * t.nK + (nu - t.mu)];
    *sign = std::isinf(v) && v < 0 ? 0 : (p == TE ? -1 : 1);
    return v;
}

// log|I_{l1 l2}^p| and its sign; 0 (log -inf) when l1 < m or l2 < m.
// Reference: src/integration.c:627-693.
double caps_integrate_I(integration_t* self, int l1, int l2, polarization_t p, sign_t* sign)
{
    const IntegTables& t = integ_tables(self);
    if (l1 < t.m || l2 < t.m) {
        *sign = 0;
        return -INFINITY;
    }
    TERMINATE(l1 < t.lo || l2 < t.lo || l1 >= t.lo + t.nl || l2 >= t.lo + t.nl,
              "I: (l1, l2) = (%d, %d) outside the table [%d, %d]", l1, l2, t.lo, t.lo + t.nl - 1);
    const size_t i = ((size_t)(p == TE ? 0 : 1) * t.nl + (l1 - t.lo)) * t.nl + (l2 - t.lo);
    *sign = t.sgnI[i];
    return t.logI[i];
}

namespace {
// log|Σ_i c_i s_i e^{x_i}| and its sign, for the A..D combinations of I
double log_combination(int n, const double* c, const double* x, const sign_t* s, sign_t* sign)
{
    double mx = -INFINITY;
    for (int i = 0; i < n; i++)
        if (s[i] != 0 && c[i] != 0) mx = std::max(mx, x[i]);
    if (std::isinf(mx)) {
        *sign = 0;
        return -INFINITY;
    }
    double sum = 0;
    for (int i = 0; i < n; i++)
        if (s[i] != 0 && c[i] != 0) sum += c[i] * s[i] * exp(x[i] - mx);
    *sign = (sign_t)((sum > 0) - (sum < 0));
    return sum == 0 ? -INFINITY : mx + log(fabs(sum));
}
}  // namespace

// A = m² I_{l1 l2} (0 for m = 0).  Reference: src/integration.c:762-776.
double caps_integrate_A(integration_t* self, int l1, int l2, polarization_t p, sign_t* sign)
{
    const int m = integ_tables(self).m;
    if (m == 0) {
        *sign = 0;
        return -INFINITY;
    }
    const double v = caps_integrate_I(self, l1, l2, p, sign);
    return *sign == 0 ? -INFINITY : 2 * log((double)m) + v;
}

// B = Σ over l1' = l1 ∓ 1, l2' = l2 ∓ 1 of the recurrence coefficients of
// P_l^m' times I_{l1' l2'} (m = 0: B = I_{l1 l2}).  Reference:
// src/integration.c:794-831.
double caps_integrate_B(integration_t* self, int l1, int l2, polarization_t p, sign_t* sign)
{
    const int m = integ_tables(self).m;
    if (m == 0) return caps_integrate_I(self, l1, l2, p, sign);
    // P_l^m'-recurrence weights: u_l = (l+1)(l+m)/(2l+1) on l-1, v_l = l(l-m+1)/(2l+1) on l+1
    const double u1 = (l1 + 1.0) * (l1 + m) / (2 * l1 + 1.0), v1 = l1 * (l1 - m + 1.0) / (2 * l1 + 1.0);
    const double u2 = (l2 + 1.0) * (l2 + m) / (2 * l2 + 1.0), v2 = l2 * (l2 - m + 1.0) / (2 * l2 + 1.0);
    double x[4];
    sign_t s[4];
    x[0] = caps_integrate_I(self, l1 - 1, l2 - 1, p, &s[0]);
    x[1] = caps_integrate_I(self, l1 + 1, l2 - 1, p, &s[1]);
    x[2] = caps_integrate_I(self, l1 - 1, l2 + 1, p, &s[2]);
    x[3] = caps_integrate_I(self, l1 + 1, l2 + 1, p, &s[3]);
    const double c[4] = {u1 * u2, -v1 * u2, -u1 * v2, v1 * v2};
    return log_combination(4, c, x, s, sign);
}

// C = m [-u_{l2} I_{l1, l2-1} + v_{l2} I_{l1, l2+1}] (0 for m = 0).
// Reference: src/integration.c:842-874.
double caps_integrate_C(integration_t* self, int l1, int l2, polarization_t p, sign_t* sign)
{
    const int m = integ_tables(self).m;
    if (m == 0) {
        *sign = 0;
        return -INFINITY;
    }
    const double u2 = (l2 + 1.0) * (l2 + m) / (2 * l2 + 1.0), v2 = l2 * (l2 - m + 1.0) / (2 * l2 + 1.0);
    double x[2];
    sign_t s[2];
    x[0] = caps_integrate_I(self, l1, l2 - 1, p, &s[0]);
    x[1] = caps_integrate_I(self, l1, l2 + 1, p, &s[1]);
    const double c[2] = {-u2, v2};
    const double v = log_combination(2, c, x, s, sign);
    return *sign == 0 ? -INFINITY : log((double)m) + v;
}

// D_{l1 l2} = C_{l2 l1}.  Reference: src/integration.c:885-888.
double caps_integrate_D(integration_t* self, int l1, int l2, polarization_t p, sign_t* sign)
{
    return caps_integrate_C(self, l2, l1, p, sign);
}

// Extension (test hook): copy the N x N matrix M of a caps_M_t (column-major).
int capsx_M_copy(caps_M_t* self, double* out)
{
    const caps_block_state* st = self->integration;
    memcpy(out, st->M, (size_t)st->N * st->N * sizeof(double));
    return st->N;
}

// Extension: accumulated statistics since the last reset: out[0..5] stage
// milliseconds (Mie, K, I, assembly, LU, total), out[6] blocks, out[7]
// blocks factored, out[8] algorithmic LU flops Σ(2/3 N³ - 1/2 N² + 5/6 N),
// out[9] batches, out[10] (l1 >= l2, polarization) I integrals.
// With capsx_profile on: out[11] milliseconds in the LU's trailing-update
// kernel k_lu_gemm_dma (HIP events around each launch), out[12] its
// algorithmic flops (2 rows cols depth per updated region, over factored
// blocks), out[13] its launches, out[14] its algorithmic HBM bytes (C read
// and written once, both operands read once); out[15] milliseconds of the
// other trailing-update launches (k_lu_gemm tiles).
void capsx_stats(caps_t* self, double out[16])
{
    Ctx* c = ctx_of(self);
    for (int i = 0; i < 16; i++) out[i] = c->acc[i];
}

void capsx_stats_reset(caps_t* self)
{
    Ctx* c = ctx_of(self);
    for (int i = 0; i < 16; i++) c->acc[i] = 0;
}

// Extension: per-launch events around the LU's trailing-update kernel.
void capsx_profile(caps_t* self, int on)
{
    Ctx* c = ctx_of(self);
    c->prof = on;
}

// Extension: attach a planner surrogate (capsx_mlp, borrowed) and the
// material features it is evaluated with (log ωp/eV, log γ/eV, with the
// sentinels of caps_amd/surrogate.py for perfect reflectors and plasma).
void capsx_set_surrogate(caps_t* self, struct capsx_mlp* mlp, float f_omegap, float f_gamma)
{
    Ctx* c = ctx_of(self);
    c->sur = mlp;
    c->sur_feat[0] = f_omegap;
    c->sur_feat[1] = f_gamma;
}

// Extension: write a dim x dim matrix (element (i, j) at M[i*dim + j]) as
// .npy in the reference's byte layout (matrix_save_to_file,
// src/matrix.c:211-254).  Returns 0 on success.
int capsx_npy_save(const char* path, const double* M, int dim)
{
    if (!path || !M || dim <= 0) return 1;
    return npy_write(path, dim, [&](int i, int j) { return M[(size_t)i * dim + j]; }) ? 0 : 1;
}

// Extension: read a square '<f8' .npy file with fortran_order True, as
// the reference's matrix_load_from_file does (src/matrix.c:270-335; the
// elements in file order).  Returns a malloc'd array (capsx_npy_free) and
// its dimension, or NULL for a file of another layout.
double* capsx_npy_load(const char* path, int* dim)
{
    FILE* f = fopen(path, "rb");
    if (!f) return nullptr;
    char magic[8];
    uint16_t len = 0;
    char hdr[512] = {0};
    double* out = nullptr;
    if (fread(magic, 1, 8, f) == 8 && memcmp(magic, "\x93NUMPY\x01\x00", 8) == 0 && fread(&len, 2, 1, f) == 1 &&
        len >= 2 && len < sizeof(hdr) && fread(hdr, 1, len, f) == len && strstr(hdr, "'descr': '<f8'") &&
        strstr(hdr, "'fortran_order': True")) {
        const char* p = strstr(hdr, "'shape': (");
        if (p) {
            p += 10;
            const long d1 = strtol(p, nullptr, 10);
            const char* q = strchr(p, ',');
            const long d2 = q ? strtol(q + 1, nullptr, 10) : -1;
            if (d1 > 0 && d1 == d2 && d1 <= 1000000) {
                out = (double*)malloc((size_t)d1 * d1 * sizeof(double));
                if (out && fread(out, sizeof(double), (size_t)d1 * d1, f) != (size_t)d1 * d1) {
                    free(out);
                    out = nullptr;
                }
                if (out) *dim = (int)d1;
            }
        }
    }
    fclose(f);
    return out;
}

void capsx_npy_free(double* M) { free(M); }

// Extension: accuracy of the planner surrogate against the blocks computed
// for its plans since the last reset (resets when reset != 0):
// out = {blocks compared, RMS and largest |log|logdetD| error|, frequencies
// compared, mean and largest |predicted - actual m-cutoff|}.
void capsx_surrogate_report(caps_t* self, double* out, int reset)
{
    Ctx* c = ctx_of(self);
    const double* a = c->sur_acc;
    out[0] = a[0];
    out[1] = a[0] > 0 ? sqrt(a[1] / a[0]) : 0;
    out[2] = a[2];
    out[3] = a[3];
    out[4] = a[3] > 0 ? a[4] / a[3] : 0;
    out[5] = a[5];
    if (reset)
        for (double& x : c->sur_acc) x = 0;
}

// Extension: launch sequences (batches) the free-energy driver used since
// the last reset; resets when reset != 0.
long capsx_plan_batches(caps_t* self, int reset)
{
    Ctx* c = ctx_of(self);
    const long r = c->plan_batches;
    if (reset) c->plan_batches = 0;
    return r;
}

// Extension: number of quadrature nodes of the K integrals of a block.
int capsx_block_nodes(caps_t* self, double xi_, int m)
{
    BlockDesc d;
    memset(&d, 0, sizeof(d));
    size_t lmin, lmax;
    caps_estimate_lminmax(self, m, &lmin, &lmax);
    d.alpha = 2 * xi_;
    d.mu = m > 0 ? 2 * m : 2;
    d.lo = (int)lmin - 1;
    d.nl = self->ldim + 2;
    d.nu_max = 2 * (d.lo + d.nl - 1);
    grid_for(d);
    return d.n_nodes;
}

}  // extern "C"
