// Batched domains: many independent FSLBM domains (pipes, periodic
// boxes) advanced together, one launch per kernel class and step, over a
// packed table of (domain, work item), one workgroup per item.  Domains are
// admitted
// against an HBM budget and removed when done; the device state of every
// admitted domain stays resident.  The launch sequence of a step can be
// captured in a hipGraph and replayed.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fslbm_codec.h"
#include "fslbm_internal.h"

using namespace fslbm;

namespace {

struct Domain {
    bool active = false;
    Geo g{};
    fslbm_params prm{};
    Quant q{1.f, 1.f, 1.f, 1.f, 0.f, 0.f};
    void *buf[2] = {nullptr, nullptr};
    uint8_t *mask2d = nullptr;
    void *omega = nullptr;
    int *tiles = nullptr;  // tiles holding fluid (nullptr: all)
    int ntiles = 0;
    std::vector<int> htiles;  // host copy (empty: all tiles)
    std::vector<uint8_t> hmask2d;
    float *uref = nullptr;     // streamwise reference profile (16-bit modes)
    std::vector<float> huref;  // its host copy
    size_t bytes = 0;
    int parity = 0;  // buffer holding the current state
    long steps = 0;
    bool held = false;  // resident but no longer stepped (result pending)
    int held_cur = 0;   // its current buffer
    int ncols = 0;      // z columns (tiles) of its work items, as of the last rebuild
};

constexpr int NCLASS = 4;  // (periodic, order-4 arms)
// Steps per captured graph.  Between two graph launches on a stream the GPU
// idles about 8 us; with two steps per graph that cost 11% of the 64-pipe
// batch's cohort steps (37 us launches), with 16 about 1.4%.  Calls of
// fewer steps, and the remainder of a call, launch directly.
constexpr int GRAPH_STEPS = 16;

int class_of(const Domain &d) { return (d.g.periodic ? 2 : 0) + (d.prm.blend != 0.0 ? 1 : 0); }

// One cohort of a many-step call (step_impl): its domains (slots) and its
// own work tables over the batch's descriptors, with their graphs.
struct Cohort {
    std::vector<int> doms;
    int4 *d_work[NCLASS] = {nullptr, nullptr, nullptr, nullptr};
    size_t cap[NCLASS] = {0, 0, 0, 0};
    int nwork[NCLASS] = {0, 0, 0, 0};
    std::vector<int4> host[NCLASS];  // kept until the next rebuild (async upload)
    bool nt = false;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};  // GRAPH_STEPS steps from flip 0 / 1
};

}  // namespace

struct fslbm_batch {
    int precision = FSLBM_F32;
    int d1 = 2;
    size_t budget = 0, used = 0;
    std::vector<Domain> doms;
    bool nt = false;  // non-temporal FP32 stores: the stepped state exceeds the LLC
    void *d_doms = nullptr;
    size_t d_doms_cap = 0;
    uint32_t *d_seeds = nullptr;  // two step counters per domain slot
    size_t d_seeds_cap = 0;
    int4 *d_work[NCLASS] = {nullptr, nullptr, nullptr, nullptr};
    size_t work_cap[NCLASS] = {0, 0, 0, 0};
    int nwork[NCLASS] = {0, 0, 0, 0};
    bool dirty = true;
    int flip = 0;   // steps since the descriptors were written, mod 2
    int chunk = 0;  // schedule (0: default per domain)
    bool use_graph = false;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};  // GRAPH_STEPS steps from flip 0 / 1
    hipStream_t stream = nullptr;
    hipStream_t capture = nullptr;  // cohort graphs are captured here, replayed on stream
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<Cohort> cohorts;
    bool cohorts_valid = false;  // their tables match the descriptors
    std::string err;
};

namespace {

int bfail(fslbm_batch *b, int code, const char *fmt, ...)
{
    if (b) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        b->err = buf;
    }
    return code;
}

#define BCHK(b, expr)                                                          \
    do {                                                                       \
        hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess)                                                  \
            return bfail((b), FSLBM_EHIP, "%s failed: %s", #expr,              \
                         hipGetErrorString(e_));                               \
    } while (0)

thread_local std::string g_batch_err;

StepArgs step_args(const Domain &d)
{
    StepArgs a;
    a.tblend = d.prm.tblend;
    a.pin = d.prm.pin;
    a.nu = d.prm.nu;
    a.blend = d.prm.blend;
    a.d1 = d.prm.d1;
    a.gauge = 0.0;
    a.q = d.q;
    a.seed = (uint32_t)d.steps;
    return a;
}

// Destroy the captured graphs, after their launches on the batch's stream
// have finished.
void drop_graphs(fslbm_batch *b)
{
    if (!b->gexec[0] && !b->gexec[1]) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (auto &g : b->gexec)
        if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
}

// The same for the cohorts' graphs; their tables no longer match.
void drop_cohort_graphs(fslbm_batch *b)
{
    bool any = false;
    for (const Cohort &co : b->cohorts) any |= co.gexec[0] || co.gexec[1];
    if (any && b->stream) (void)hipStreamSynchronize(b->stream);
    for (Cohort &co : b->cohorts)
        for (auto &g : co.gexec)
            if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
    b->cohorts_valid = false;
}

// One z column of one tile of a domain, still to be cut into work items.
struct Column {
    int dom, tile, z0, rem;
};

// Static split of the columns of one kernel class into work items.  One
// item length L serves the whole table: the one whose items, dealt to the
// XCDs in rounds of one workgroup per CU, minimise rounds x (L + 3 planes
// of start-up), the single-domain kernel's step-time model; ties go to the
// longer items.  Each column is then cut into near-equal chunks of at most
// L planes.  (Shrinking the length as the table fills, as guided
// scheduling does, left a single 16-tile pipe of 734 planes with 1024
// items, 489 of them 8 planes long, against 256 items of 46 planes in one
// round for the same domain stepped alone.)  The items are ordered longest
// first and dealt to the XCDs in groups (see below), so each XCD gets an
// even share.  One item per domain carries BATCH_LEAD.  The cut changes
// only the schedule: every cell is computed the same way whatever item
// holds it.
std::vector<int4> split_columns(std::vector<Column> &cols, int lmax, int cus, size_t ndom)
{
    const int lmin = std::min(4, lmax);
    int longest = 0;
    for (const Column &c : cols) longest = std::max(longest, c.rem);
    const long long slots = std::max(1, cus / 8);
    int L = std::max(1, std::min(lmax, longest));
    double best = -1.0;
    for (int len = std::min(lmax, std::max(longest, 1)); len >= lmin; len--) {
        long long items = 0;
        for (const Column &c : cols) items += (c.rem + len - 1) / len;
        const long long rounds = ((items + 7) / 8 + slots - 1) / slots;
        const double cost = (double)rounds * (len + 3);
        if (best < 0 || cost < best) {
            best = cost;
            L = len;
        }
    }
    // chunk-major across the columns (a domain's tiles are consecutive
    // columns), so equal-length items of neighbouring tiles and one z range
    // stay next to each other through the sort below
    std::vector<std::pair<int, int4>> items;
    int maxch = 0;
    for (const Column &c : cols)
        if (c.rem > 0) maxch = std::max(maxch, (c.rem + L - 1) / L);
    for (int k = 0; k < maxch; k++)
        for (const Column &c : cols) {
            if (c.rem <= 0) continue;
            const int nch = (c.rem + L - 1) / L;
            if (k >= nch) continue;
            const int q = c.rem / nch, r = c.rem - q * nch;
            const int z = c.z0 + k * q + std::min(k, r);
            const int take = q + (k < r ? 1 : 0);
            items.push_back({take, make_int4(c.dom, c.tile, z, z + take)});
        }
    std::stable_sort(items.begin(), items.end(),
                     [](const auto &x, const auto &y) { return x.first > y.first; });
    // Deal groups of G consecutive items (neighbouring tiles of one z
    // range, which share halo planes in L2) to the XCDs round robin:
    // group g goes to XCD g mod 8.  Workgroup b runs on XCD b mod 8, so
    // item k of the sorted list sits at 8 * (its index on its XCD) + XCD;
    // unused positions hold empty items (domain -1).
    constexpr int G = BATCH_GROUP;
    const size_t n = items.size();
    const size_t ngroups = (n + G - 1) / G;
    const size_t rounds = (ngroups + 7) / 8;
    std::vector<int4> out(n ? 8 * rounds * G : 0, make_int4(-1, 0, 0, 0));
    std::vector<char> led(ndom, 0);
    for (size_t k = 0; k < n; k++) {
        int4 it = items[k].second;
        if (!led[it.x]) {
            led[it.x] = 1;
            it.w |= BATCH_LEAD;
        }
        const size_t g = k / G, xcd = g % 8, seq = (g / 8) * G + k % G;
        out[8 * seq + xcd] = it;
    }
    // trailing empty positions need no workgroup
    while (!out.empty() && out.back().x < 0) out.pop_back();
    return out;
}

// Rewrite the descriptors and work tables (after add/remove/retune).
// chunk_override >= 0 replaces the batch's chunk (the longest item, in z
// planes; 0: the default of 128).
int rebuild(fslbm_batch *b, int chunk_override = -1)
{
    const size_t db = batch_dom_bytes(b->precision);
    const size_t nd = b->doms.size();
    std::vector<unsigned char> host(std::max<size_t>(nd, 1) * db, 0);
    std::vector<uint32_t> seeds(2 * std::max<size_t>(nd, 1), 0u);
    const int ch = chunk_override >= 0 ? chunk_override : b->chunk;
    std::vector<Column> cols[NCLASS];
    size_t stepped = 0;  // state bytes the step reads and writes
    for (size_t i = 0; i < nd; i++) {
        Domain &d = b->doms[i];
        if (!d.active || d.held) continue;
        stepped += 2 * state_bytes(d.g, b->precision);
        // the current parity of each domain is folded into its descriptor
        d.parity ^= b->flip;
        int split[3] = {1, 0, 0};
        const int nw = batch_fill_dom(host.data() + i * db, b->precision, d.g, step_args(d), 0,
                                      d.buf[0], d.buf[1], d.mask2d, d.omega, d.tiles,
                                      d.ntiles, d.parity, split);
        if (nw < 0) return bfail(b, FSLBM_EINVAL, "domain %zu: unsupported precision", i);
        seeds[2 * i] = (uint32_t)d.steps;  // the next launch has flip 0
        seeds[2 * i + 1] = (uint32_t)d.steps + 1u;
        auto &cl = cols[class_of(d)];
        d.ncols = d.htiles.empty() ? split[0] : (int)d.htiles.size();
        if (d.htiles.empty())
            for (int t = 0; t < split[0]; t++) cl.push_back({(int)i, t, 0, d.g.nz});
        else
            for (int t : d.htiles) cl.push_back({(int)i, t, 0, d.g.nz});
    }
    b->nt = stream_stores(stepped);
    const int cus = device_cus();
    std::vector<int4> work[NCLASS];
    for (int c = 0; c < NCLASS; c++) work[c] = split_columns(cols[c], ch > 0 ? ch : 128, cus, nd);
    b->flip = 0;
    // The captured graphs hold the current tables' addresses: let their
    // launches finish and destroy them before any table is freed or
    // rewritten, so that no graph exec outlives memory it refers to.
    BCHK(b, hipStreamSynchronize(b->stream));
    drop_graphs(b);
    drop_cohort_graphs(b);
    if (host.size() > b->d_doms_cap) {
        if (b->d_doms) (void)hipFree(b->d_doms);
        b->d_doms = nullptr;
        b->d_doms_cap = 0;
        BCHK(b, hipMalloc(&b->d_doms, host.size()));
        b->d_doms_cap = host.size();
    }
    BCHK(b, hipMemcpyAsync(b->d_doms, host.data(), host.size(), hipMemcpyHostToDevice, b->stream));
    const size_t sb = seeds.size() * sizeof(uint32_t);
    if (sb > b->d_seeds_cap) {
        if (b->d_seeds) (void)hipFree(b->d_seeds);
        b->d_seeds = nullptr;
        b->d_seeds_cap = 0;
        BCHK(b, hipMalloc((void **)&b->d_seeds, sb));
        b->d_seeds_cap = sb;
    }
    BCHK(b, hipMemcpyAsync(b->d_seeds, seeds.data(), sb, hipMemcpyHostToDevice, b->stream));
    for (int c = 0; c < NCLASS; c++) {
        const size_t need = std::max<size_t>(work[c].size(), 1) * sizeof(int4);
        if (need > b->work_cap[c]) {
            if (b->d_work[c]) (void)hipFree(b->d_work[c]);
            b->d_work[c] = nullptr;
            b->work_cap[c] = 0;
            BCHK(b, hipMalloc((void **)&b->d_work[c], need));
            b->work_cap[c] = need;
        }
        if (!work[c].empty())
            BCHK(b, hipMemcpyAsync(b->d_work[c], work[c].data(), work[c].size() * sizeof(int4),
                                   hipMemcpyHostToDevice, b->stream));
        b->nwork[c] = (int)work[c].size();
    }
    BCHK(b, hipStreamSynchronize(b->stream));
    b->dirty = false;
    return FSLBM_OK;
}

// Launch one step of every class of the given work tables with the given
// flip (no bookkeeping).
hipError_t launch_tables(const fslbm_batch *b, int4 *const *work, const int *nwork, bool nt,
                         int flip, hipStream_t s)
{
    for (int c = 0; c < NCLASS; c++) {
        if (nwork[c] == 0) continue;
        BatchLaunch L;
        L.doms = b->d_doms;
        L.seeds = b->d_seeds;
        L.work = work[c];
        L.nwork = nwork[c];
        L.flip = flip;
        L.grid = nwork[c];  // one workgroup per item
        L.periodic = (c & 2) != 0;
        L.fd2 = (c & 1) != 0;
        L.nt = nt;
        L.d1 = b->d1;
        L.precision = b->precision;
        L.stream = s;
        const hipError_t e = launch_step_batch(L);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_step_all(fslbm_batch *b, int flip)
{
    return launch_tables(b, b->d_work, b->nwork, b->nt, flip, b->stream);
}

// Capture the two graphs (GRAPH_STEPS steps each, from flip 0 and from
// flip 1).
// Either both are made or neither: a failure destroys what was made, so
// that no launch ever finds one graph of the pair missing.
int capture_graphs(fslbm_batch *b)
{
    drop_graphs(b);
    for (int start = 0; start < 2; start++) {
        hipGraph_t g = nullptr;
        hipError_t e0 = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
        if (e0 != hipSuccess) {
            drop_graphs(b);
            return bfail(b, FSLBM_EHIP, "graph capture failed: %s", hipGetErrorString(e0));
        }
        hipError_t e1 = hipSuccess;
        for (int k = 0; k < GRAPH_STEPS && e1 == hipSuccess; k++)
            e1 = launch_step_all(b, start ^ (k & 1));
        hipError_t e2 = hipSuccess;
        hipError_t e3 = hipStreamEndCapture(b->stream, &g);
        if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            drop_graphs(b);
            return bfail(b, FSLBM_EHIP, "graph capture failed: %s",
                         hipGetErrorString(e1 != hipSuccess ? e1 : e2 != hipSuccess ? e2 : e3));
        }
        hipError_t e4 = hipGraphInstantiate(&b->gexec[start], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e4 != hipSuccess) {
            b->gexec[start] = nullptr;
            drop_graphs(b);
            return bfail(b, FSLBM_EHIP, "graph instantiate failed: %s", hipGetErrorString(e4));
        }
    }
    return FSLBM_OK;
}

int step_all(fslbm_batch *b, int nsteps)
{
    if (b->dirty) {
        const int rc = rebuild(b);
        if (rc) return rc;
    }
    bool any = false;
    for (int c = 0; c < NCLASS; c++) any |= b->nwork[c] > 0;
    int s = 0;
    if (any && b->use_graph && nsteps >= GRAPH_STEPS) {
        if (!b->gexec[0] || !b->gexec[1]) {
            const int rc = capture_graphs(b);
            if (rc) return rc;
        }
        for (; s + GRAPH_STEPS <= nsteps; s += GRAPH_STEPS)
            BCHK(b, hipGraphLaunch(b->gexec[b->flip], b->stream));
    }
    for (; s < nsteps; s++) {
        if (any) BCHK(b, launch_step_all(b, b->flip));
        b->flip ^= 1;
    }
    for (auto &d : b->doms)
        if (d.active && !d.held) d.steps += nsteps;
    return FSLBM_OK;
}


// The work tables of one cohort over the batch's descriptors (written by
// the last rebuild, which covers every stepping domain), uploaded on the
// batch's stream.  The caller has waited for the launches that read the
// cohort's previous tables.
int build_cohort(fslbm_batch *b, Cohort &co)
{
    std::vector<Column> cols[NCLASS];
    size_t stepped = 0;
    for (int i : co.doms) {
        const Domain &d = b->doms[i];
        stepped += 2 * state_bytes(d.g, b->precision);
        auto &cl = cols[class_of(d)];
        if (d.htiles.empty())
            for (int t = 0; t < d.ncols; t++) cl.push_back({i, t, 0, d.g.nz});
        else
            for (int t : d.htiles) cl.push_back({i, t, 0, d.g.nz});
    }
    co.nt = stream_stores(stepped);
    const int cus = device_cus();
    for (int c = 0; c < NCLASS; c++) {
        co.host[c] = split_columns(cols[c], b->chunk > 0 ? b->chunk : 128, cus, b->doms.size());
        const size_t need = std::max<size_t>(co.host[c].size(), 1) * sizeof(int4);
        if (need > co.cap[c]) {
            if (co.d_work[c]) (void)hipFree(co.d_work[c]);
            co.d_work[c] = nullptr;
            co.cap[c] = 0;
            BCHK(b, hipMalloc((void **)&co.d_work[c], need));
            co.cap[c] = need;
        }
        if (!co.host[c].empty())
            BCHK(b, hipMemcpyAsync(co.d_work[c], co.host[c].data(),
                                   co.host[c].size() * sizeof(int4), hipMemcpyHostToDevice,
                                   b->stream));
        co.nwork[c] = (int)co.host[c].size();
    }
    return FSLBM_OK;
}

// Capture a cohort's two graphs (GRAPH_STEPS steps each, from flip 0 and
// from flip 1) on the capture stream; they are replayed on the batch's
// stream.
int capture_cohort(fslbm_batch *b, Cohort &co)
{
    // either both graphs are made or neither
    auto fail_ = [&](const char *what, hipError_t e) {
        for (auto &g : co.gexec)
            if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
        return bfail(b, FSLBM_EHIP, "%s failed: %s", what, hipGetErrorString(e));
    };
    for (int start = 0; start < 2; start++) {
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return fail_("graph capture", e);
        hipError_t el = hipSuccess;
        for (int k = 0; k < GRAPH_STEPS && el == hipSuccess; k++)
            el = launch_tables(b, co.d_work, co.nwork, co.nt, start ^ (k & 1), b->capture);
        e = hipStreamEndCapture(b->capture, &g);
        if (el != hipSuccess || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return fail_("graph capture", el != hipSuccess ? el : e);
        }
        e = hipGraphInstantiate(&co.gexec[start], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            co.gexec[start] = nullptr;
            return fail_("graph instantiate", e);
        }
    }
    return FSLBM_OK;
}

// Cohorts: a batch whose stepped state (both buffers of every stepping
// domain) exceeds the last-level cache streams it from HBM at every step,
// while a domain stepped alone whose state fits stays in that cache between
// steps (a tile-plane from cache took 1.06 us against 1.5 us from HBM).  So
// for a call of many steps such a batch is cut into cohorts of domains (in
// slot order) whose state together fits COHORT_BYTES, and each cohort runs
// all the steps before the next: every launch still covers many domains
// (which small domains alone cannot fill the GPU with) and keeps them in
// cache.  One descriptor table serves every cohort; each cohort has its own
// work tables (and graphs), written when the cohorts change and otherwise
// reused, so the cohorts' launches are queued back to back without waiting
// for the stream (the host builds the next cohort's tables while the GPU
// runs the previous one).  Only the schedule changes: each domain's
// arithmetic and step counter are those of one batch step after another.
// 15/16 of the cache: on the 64 pipes (10,000 steps) cohorts of 128, 192,
// 240, 256, 320 and 400 MB ran 12.82, 12.11, 11.59, 11.70, 14.09 and
// 15.33 s.
constexpr size_t COHORT_BYTES = LLC_BYTES / 16 * 15;
constexpr int COHORT_MIN_STEPS = 32;

int step_impl(fslbm_batch *b, int nsteps)
{
    std::vector<std::vector<int>> cohorts;
    if (nsteps >= COHORT_MIN_STEPS) {
        size_t total = 0, cur = 0;
        for (size_t i = 0; i < b->doms.size(); i++) {
            const Domain &d = b->doms[i];
            if (!d.active || d.held) continue;
            const size_t sb = 2 * state_bytes(d.g, b->precision);
            total += sb;
            if (cohorts.empty() || cur + sb > COHORT_BYTES) {
                cohorts.emplace_back();
                cur = 0;
            }
            cohorts.back().push_back((int)i);
            cur += sb;
        }
        if (!stream_stores(total)) cohorts.clear();  // fits the cache as a whole
    }
    if (cohorts.size() < 2) return step_all(b, nsteps);
    if (b->dirty) {  // descriptors of every stepping domain (waits for the stream)
        const int rc = rebuild(b);
        if (rc) return rc;
    }
    bool same = b->cohorts_valid && b->cohorts.size() == cohorts.size();
    for (size_t k = 0; same && k < cohorts.size(); k++) same = b->cohorts[k].doms == cohorts[k];
    if (!same) {
        // the launches of the last call may still read the cohort tables
        BCHK(b, hipStreamSynchronize(b->stream));
        drop_cohort_graphs(b);
        for (size_t k = cohorts.size(); k < b->cohorts.size(); k++)
            for (auto *w : b->cohorts[k].d_work) if (w) (void)hipFree(w);
        b->cohorts.resize(cohorts.size());
        for (size_t k = 0; k < cohorts.size(); k++) b->cohorts[k].doms = cohorts[k];
    }
    const int f0 = b->flip;
    for (size_t k = 0; k < cohorts.size(); k++) {
        Cohort &co = b->cohorts[k];
        if (!same) {
            const int rc = build_cohort(b, co);
            if (rc) return rc;
        }
        int s = 0;
        if (b->use_graph && nsteps >= GRAPH_STEPS) {
            if (!co.gexec[0] || !co.gexec[1]) {
                const int rc = capture_cohort(b, co);
                if (rc) return rc;
            }
            for (; s + GRAPH_STEPS <= nsteps; s += GRAPH_STEPS)
                BCHK(b, hipGraphLaunch(co.gexec[f0], b->stream));
        }
        for (; s < nsteps; s++)
            BCHK(b, launch_tables(b, co.d_work, co.nwork, co.nt, f0 ^ (s & 1), b->stream));
    }
    b->cohorts_valid = true;
    b->flip = f0 ^ (nsteps & 1);
    for (auto &d : b->doms)
        if (d.active && !d.held) d.steps += nsteps;
    return FSLBM_OK;
}

size_t domain_bytes(int precision, int nx, int ny, int nz, bool masked)
{
    const size_t N = (size_t)nx * ny * nz, nxy = (size_t)nx * ny;
    const size_t rb = precision == FSLBM_F64 ? 8 : 4;
    const size_t tiles = (size_t)((nx + 63) / 64) * ((ny + 7) / 8);  // bound on the tile list
    const size_t uref = precision == FSLBM_F16S || precision == FSLBM_BF16S ? 4 * nxy : 0;
    return 2 * 4 * N * storage_bytes(precision) + uref +
           (masked ? nxy + 3 * nxy * rb + tiles * sizeof(int) : 0);
}

}  // namespace

namespace fslbm {

// Check a domain against what a batch of this precision and d1 runs (no
// batch state is touched, so any thread may call it).  Returns FSLBM_OK, or
// an error with the reason in `why`.
int domain_check(int precision, int d1, const fslbm_params *prm, const uint8_t *mask, int nx,
                 int ny, int nz, bool have_state, std::string &why)
{
    auto bad = [&](const char *m) {
        why = m;
        return FSLBM_EINVAL;
    };
    if (!prm || !have_state) return bad("NULL argument");
    if (nx < 1 || ny < 1 || nz < 1) return bad("bad box");
    const long long N = (long long)nx * ny * nz, nxy = (long long)nx * ny;
    if (N >= (1ll << 31)) return bad("domain too large (N >= 2^31 cells)");
    if (prm->d1 != d1) {
        why = "domain d1 " + std::to_string(prm->d1) + " differs from the batch's " +
              std::to_string(d1);
        return FSLBM_EINVAL;
    }
    if (!std::isfinite(prm->tblend) || !std::isfinite(prm->pin) || !std::isfinite(prm->nu) ||
        !std::isfinite(prm->blend))
        return bad("non-finite run parameter");
    Geo g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.N = N;
    g.periodic = mask == nullptr;
    if (g.periodic && (2 * prm->d1 > nx || 2 * prm->d1 > ny))
        return bad("periodic box narrower than the corrector reach");
    if (mask) {
        for (long long i = 0; i < nxy; i++) {
            const uint8_t m = mask[i];
            if (m != 0 && m != 125) return bad("mask codes must be 0 or 125");
            const int x = (int)(i % nx), y = (int)(i / nx);
            if (m == 0 && (x == 0 || y == 0 || x == nx - 1 || y == ny - 1))
                return bad("fluid cell on the x/y face; a wall frame is required");
        }
    }
    if (!tiled_supported(g, precision, prm->d1, true))
        return bad("domain not supported by the tiled kernel (nz >= 16 required)");
    return FSLBM_OK;
}

// Encode a checked domain on the host (no device work, no batch state:
// any thread may call it, also while the batch's launches run).
void domain_encode(int precision, const fslbm_params *prm, const uint8_t *mask, int nx, int ny,
                   int nz, const double *const src[4], PreparedDomain &pd)
{
    const long long N = (long long)nx * ny * nz, nxy = (long long)nx * ny;
    Geo g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.N = N;
    g.periodic = mask == nullptr;
    pd = PreparedDomain();
    pd.g = g;
    pd.prm = *prm;
    pd.precision = precision;
    pd.bytes = domain_bytes(precision, nx, ny, nz, mask != nullptr);
    if (mask) pd.mask2d.assign(mask, mask + nxy);
    HostMask hm;
    hm.m = mask ? pd.mask2d.data() : nullptr;
    hm.nxy = nxy;
    hm.plane = true;
    if (precision != FSLBM_F32) {
        choose_quant(g, hm, prm->pin, src, pd.q);
        pd.uref = choose_uref(g, hm, src);
    }
    pd.host.resize(4 * (size_t)N * storage_bytes(precision));
    encode_state(g, hm, precision, pd.q, src, pd.host.data(),
                 pd.uref.empty() ? nullptr : pd.uref.data());
    if (mask) {
        pd.omega = wall_weights<float>(nx, ny, pd.mask2d.data());
        // planned with the rows of the domain's kernel class (batch_fill_dom
        // and the class launch take the same)
        pd.tiles = active_tiles(g, plan_rows(precision, false, prm->blend != 0.0),
                                pd.mask2d.data());
    }
}

// Validate and encode a domain on the host (no device work; safe while
// the batch's launches run).  The mask (NULL: periodic box) is one z plane.
int batch_prepare(fslbm_batch *b, const fslbm_params *prm, const uint8_t *mask, int nx,
                  int ny, int nz, const double *const src[4], PreparedDomain &pd)
{
    if (!b) return FSLBM_EINVAL;
    std::string why;
    const bool have = src[0] && src[1] && src[2] && src[3];
    const int rc = domain_check(b->precision, b->d1, prm, mask, nx, ny, nz, have, why);
    if (rc) return bfail(b, rc, "%s", why.c_str());
    domain_encode(b->precision, prm, mask, nx, ny, nz, src, pd);
    return FSLBM_OK;
}

// Upload a prepared domain and admit it.  Returns the slot, FSLBM_EBUSY
// when the budget cannot hold it, or another error.
int batch_attach(fslbm_batch *b, PreparedDomain &pd, hipStream_t s, bool sync)
{
    if (!s && b) s = b->stream;
    if (!b) return FSLBM_EINVAL;
    if (pd.precision != b->precision)
        return bfail(b, FSLBM_EINVAL, "domain was prepared for another precision mode");
    const size_t bytes = pd.bytes;
    if (b->used + bytes > b->budget)
        return bfail(b, FSLBM_EBUSY, "HBM budget: %zu bytes in use, %zu needed, budget %zu",
                     b->used, bytes, b->budget);
    const Geo &g = pd.g;
    const long long nxy = (long long)g.nx * g.ny;
    Domain d;
    d.active = true;
    d.g = g;
    d.prm = pd.prm;
    d.q = pd.q;
    d.bytes = bytes;
    d.hmask2d = std::move(pd.mask2d);
    d.htiles = pd.tiles;
    d.huref = pd.uref;
    const size_t sbytes = pd.host.size();
    auto cleanup = [&]() {
        for (void *q : d.buf) if (q) (void)hipFree(q);
        if (d.mask2d) (void)hipFree(d.mask2d);
        if (d.omega) (void)hipFree(d.omega);
        if (d.tiles) (void)hipFree(d.tiles);
        if (d.uref) (void)hipFree(d.uref);
    };
    bool ok = hipMalloc(&d.buf[0], sbytes) == hipSuccess &&
              hipMalloc(&d.buf[1], sbytes) == hipSuccess;
    if (ok && !d.huref.empty()) {
        ok = hipMalloc((void **)&d.uref, sizeof(float) * d.huref.size()) == hipSuccess &&
             hipMemcpyAsync(d.uref, pd.uref.data(), sizeof(float) * pd.uref.size(),
                            hipMemcpyHostToDevice, s) == hipSuccess;
        d.q.uref = d.uref;
    }
    if (ok && !g.periodic) {
        ok = hipMalloc((void **)&d.mask2d, nxy) == hipSuccess &&
             hipMalloc(&d.omega, pd.omega.size() * sizeof(float)) == hipSuccess &&
             hipMemcpyAsync(d.mask2d, d.hmask2d.data(), nxy, hipMemcpyHostToDevice, s) == hipSuccess &&
             hipMemcpyAsync(d.omega, pd.omega.data(), pd.omega.size() * sizeof(float), hipMemcpyHostToDevice, s) == hipSuccess;
        if (ok && !pd.tiles.empty()) {
            ok = hipMalloc((void **)&d.tiles, pd.tiles.size() * sizeof(int)) == hipSuccess &&
                 hipMemcpyAsync(d.tiles, pd.tiles.data(), pd.tiles.size() * sizeof(int),
                                hipMemcpyHostToDevice, s) == hipSuccess;
            d.ntiles = (int)pd.tiles.size();
        }
    }
    if (ok)
        ok = hipMemcpyAsync(d.buf[0], pd.host.data(), sbytes, hipMemcpyHostToDevice, s) == hipSuccess &&
             hipMemcpyAsync(d.buf[1], pd.host.data(), sbytes, hipMemcpyHostToDevice, s) == hipSuccess &&
             (!sync || hipStreamSynchronize(s) == hipSuccess);
    if (!ok) {
        cleanup();
        return bfail(b, FSLBM_ENOMEM, "device allocation for the domain failed");
    }
    if (sync) {
        pd.host = std::vector<unsigned char>();
        pd.omega = std::vector<float>();
        pd.tiles = std::vector<int>();
        pd.uref = std::vector<float>();
    }  // else the caller keeps them until it has synchronised s
    // reuse a free slot
    int slot = -1;
    for (size_t i = 0; i < b->doms.size(); i++)
        if (!b->doms[i].active) { slot = (int)i; break; }
    // a new domain starts at the batch's current flip, so that after the
    // next rebuild its parity (0 ^ flip folded) names buf[0]
    d.parity = b->flip;
    if (slot < 0) {
        slot = (int)b->doms.size();
        b->doms.push_back(std::move(d));
    } else {
        b->doms[slot] = std::move(d);
    }
    b->used += bytes;
    b->dirty = true;
    return slot;
}

int batch_add_plane(fslbm_batch *b, const fslbm_params *prm, const uint8_t *mask, int nx,
                    int ny, int nz, const double *const src[4])
{
    PreparedDomain pd;
    const int rc = batch_prepare(b, prm, mask, nx, ny, nz, src, pd);
    if (rc) return rc;
    return batch_attach(b, pd, nullptr, true);
}

}  // namespace fslbm

extern "C" {

int fslbm_batch_create(int precision, int d1, size_t hbm_budget, fslbm_batch_t *out)
{
    g_batch_err.clear();
    if (!out) return FSLBM_EINVAL;
    *out = nullptr;
    if (precision != FSLBM_F32 && precision != FSLBM_F16S && precision != FSLBM_BF16S) {
        g_batch_err = "batches take FP32 or 16-bit shifted storage";
        return FSLBM_EINVAL;
    }
    if (d1 != 1 && d1 != 2) {
        g_batch_err = "batches take d1 = 1 or 2";
        return FSLBM_EINVAL;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        g_batch_err = "no HIP device visible";
        return FSLBM_ENODEV;
    }
    fslbm_batch *b = new fslbm_batch();
    b->precision = precision;
    b->d1 = d1;
    b->budget = hbm_budget;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
        g_batch_err = "stream/event creation failed";
        fslbm_batch_destroy(b);
        return FSLBM_EHIP;
    }
    *out = b;
    return FSLBM_OK;
}

void fslbm_batch_destroy(fslbm_batch_t b)
{
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graphs(b);
    drop_cohort_graphs(b);
    for (auto &co : b->cohorts)
        for (auto *w : co.d_work) if (w) (void)hipFree(w);
    for (auto &d : b->doms) {
        for (void *p : d.buf) if (p) (void)hipFree(p);
        if (d.mask2d) (void)hipFree(d.mask2d);
        if (d.omega) (void)hipFree(d.omega);
        if (d.tiles) (void)hipFree(d.tiles);
        if (d.uref) (void)hipFree(d.uref);
    }
    if (b->d_doms) (void)hipFree(b->d_doms);
    if (b->d_seeds) (void)hipFree(b->d_seeds);
    for (auto *w : b->d_work) if (w) (void)hipFree(w);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->capture) (void)hipStreamDestroy(b->capture);
    delete b;
}

const char *fslbm_batch_last_error(fslbm_batch_t b)
{
    return b ? b->err.c_str() : g_batch_err.c_str();
}

long long fslbm_batch_domain_bytes(int precision, const uint8_t *mask, int nx, int ny, int nz)
{
    if (nx < 1 || ny < 1 || nz < 1) return FSLBM_EINVAL;
    return (long long)domain_bytes(precision, nx, ny, nz, mask != nullptr);
}

int fslbm_batch_add(fslbm_batch_t b, const fslbm_params *prm, const uint8_t *mask, int nx,
                    int ny, int nz, const double *p, const double *ux, const double *uy,
                    const double *uz)
{
    if (!b) return FSLBM_EINVAL;
    if (!prm || !p || !ux || !uy || !uz) return bfail(b, FSLBM_EINVAL, "NULL argument");
    if (nx < 1 || ny < 1 || nz < 1) return bfail(b, FSLBM_EINVAL, "bad box");
    const long long nxy = (long long)nx * ny;
    if (mask)
        for (int z = 1; z < nz; z++)
            if (std::memcmp(mask, mask + z * nxy, nxy) != 0)
                return bfail(b, FSLBM_EINVAL, "batched domains need a mask that is the same in every z plane");
    const double *src[4] = {p, ux, uy, uz};
    return batch_add_plane(b, prm, mask, nx, ny, nz, src);
}

int fslbm_batch_remove(fslbm_batch_t b, int slot)
{
    if (!b) return FSLBM_EINVAL;
    if (slot < 0 || slot >= (int)b->doms.size() || !b->doms[slot].active)
        return bfail(b, FSLBM_EINVAL, "no domain in slot %d", slot);
    BCHK(b, hipStreamSynchronize(b->stream));
    Domain &d = b->doms[slot];
    for (void *&q : d.buf) if (q) { (void)hipFree(q); q = nullptr; }
    if (d.mask2d) (void)hipFree(d.mask2d);
    if (d.omega) (void)hipFree(d.omega);
    if (d.tiles) (void)hipFree(d.tiles);
    if (d.uref) (void)hipFree(d.uref);
    d.mask2d = nullptr;
    d.omega = nullptr;
    d.tiles = nullptr;
    d.uref = nullptr;
    b->used -= d.bytes;
    d = Domain();
    b->dirty = true;
    return FSLBM_OK;
}

int fslbm_batch_step(fslbm_batch_t b, int nsteps)
{
    if (!b) return FSLBM_EINVAL;
    if (nsteps < 0) return bfail(b, FSLBM_EINVAL, "nsteps < 0");
    return step_impl(b, nsteps);
}

int fslbm_batch_get_state(fslbm_batch_t b, int slot, double *p, double *ux, double *uy,
                          double *uz)
{
    if (!b) return FSLBM_EINVAL;
    if (slot < 0 || slot >= (int)b->doms.size() || !b->doms[slot].active)
        return bfail(b, FSLBM_EINVAL, "no domain in slot %d", slot);
    if (!p || !ux || !uy || !uz) return bfail(b, FSLBM_EINVAL, "NULL state array");
    const Domain &d = b->doms[slot];
    // current buffer: parity ^ flip (a new domain holds its state in both)
    const int cur = d.held ? d.held_cur : d.parity ^ b->flip;
    const size_t sbytes = 4 * (size_t)d.g.N * storage_bytes(b->precision);
    std::vector<unsigned char> host(sbytes);
    BCHK(b, hipMemcpyAsync(host.data(), d.buf[cur], sbytes, hipMemcpyDeviceToHost, b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    HostMask hm;
    hm.m = d.g.periodic ? nullptr : d.hmask2d.data();
    hm.nxy = (long long)d.g.nx * d.g.ny;
    hm.plane = true;
    double *dst[4] = {p, ux, uy, uz};
    decode_state(d.g, hm, b->precision, d.q, host.data(), dst,
                 d.huref.empty() ? nullptr : d.huref.data());
    return FSLBM_OK;
}

int fslbm_batch_query(fslbm_batch_t b, size_t *used, size_t *budget, int *ndomains,
                      long long *cells)
{
    if (!b) return FSLBM_EINVAL;
    int n = 0;
    long long c = 0;
    for (const auto &d : b->doms)
        if (d.active) { n++; c += d.g.N; }
    if (used) *used = b->used;
    if (budget) *budget = b->budget;
    if (ndomains) *ndomains = n;
    if (cells) *cells = c;
    return FSLBM_OK;
}

int fslbm_batch_use_graph(fslbm_batch_t b, int enable)
{
    if (!b) return FSLBM_EINVAL;
    b->use_graph = enable != 0;
    drop_graphs(b);
    return FSLBM_OK;
}

int fslbm_batch_set_chunk(fslbm_batch_t b, int chunk)
{
    if (!b) return FSLBM_EINVAL;
    if (chunk < 0) return bfail(b, FSLBM_EINVAL, "chunk must be >= 0");
    b->chunk = chunk;
    b->dirty = true;
    return FSLBM_OK;
}

int fslbm_batch_autotune(fslbm_batch_t b, int trial_steps, int *best_chunk, double *best_ms)
{
    if (!b) return FSLBM_EINVAL;
    if (trial_steps < 1) return bfail(b, FSLBM_EINVAL, "trial_steps must be >= 1");
    const int cands[] = {0, 16, 24, 32, 48, 64, 96, 128, 192, 256};
    int best = 0;
    double bms = 1e300;
    // untimed warm-up (flip 0: no domain advances), so that the first
    // candidate is not timed on an idle GPU
    {
        int rc = rebuild(b, cands[0]);
        if (rc) return rc;
        for (int t = 0; t < 2 * trial_steps; t++) BCHK(b, launch_step_all(b, 0));
    }
    for (int cand : cands) {
        int rc = rebuild(b, cand);
        if (rc) return rc;
        // trial steps all use flip 0: they read the current states and
        // write the spare buffers, so no domain advances
        BCHK(b, launch_step_all(b, 0));
        BCHK(b, hipEventRecord(b->ev0, b->stream));
        for (int t = 0; t < trial_steps; t++) BCHK(b, launch_step_all(b, 0));
        BCHK(b, hipEventRecord(b->ev1, b->stream));
        BCHK(b, hipEventSynchronize(b->ev1));
        float ms = 0;
        BCHK(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
        if (ms / trial_steps < bms) { bms = ms / trial_steps; best = cand; }
    }
    b->chunk = best;
    const int rc = rebuild(b);
    if (rc) return rc;
    if (best_chunk) *best_chunk = best;
    if (best_ms) *best_ms = bms;
    return FSLBM_OK;
}

int fslbm_batch_time_steps(fslbm_batch_t b, int nsteps, double *ms)
{
    if (!b || !ms) return FSLBM_EINVAL;
    if (b->dirty) {
        const int rc = rebuild(b);
        if (rc) return rc;
    }
    if (b->use_graph && (!b->gexec[0] || !b->gexec[1])) {
        const int rc = capture_graphs(b);
        if (rc) return rc;
    }
    BCHK(b, hipEventRecord(b->ev0, b->stream));
    const int rc = step_impl(b, nsteps);
    if (rc) return rc;
    BCHK(b, hipEventRecord(b->ev1, b->stream));
    BCHK(b, hipEventSynchronize(b->ev1));
    float f = 0;
    BCHK(b, hipEventElapsedTime(&f, b->ev0, b->ev1));
    *ms = f;
    return FSLBM_OK;
}

int fslbm_batch_sync(fslbm_batch_t b)
{
    if (!b) return FSLBM_EINVAL;
    BCHK(b, hipStreamSynchronize(b->stream));
    return FSLBM_OK;
}

}  // extern "C"

namespace fslbm {

// Stop stepping a domain but keep it resident (its result is copied out
// later, while the others step).
int batch_hold(fslbm_batch *b, int slot)
{
    if (!b) return FSLBM_EINVAL;
    if (slot < 0 || slot >= (int)b->doms.size() || !b->doms[slot].active)
        return bfail(b, FSLBM_EINVAL, "no domain in slot %d", slot);
    Domain &d = b->doms[slot];
    if (d.held) return FSLBM_OK;
    d.held_cur = d.parity ^ b->flip;
    d.held = true;
    b->dirty = true;
    return FSLBM_OK;
}

// Step a held domain again from the buffer it was held in.  Its step
// counter moved only while it stepped, so its results are those of an
// uninterrupted run.
int batch_resume(fslbm_batch *b, int slot)
{
    if (!b) return FSLBM_EINVAL;
    if (slot < 0 || slot >= (int)b->doms.size() || !b->doms[slot].active)
        return bfail(b, FSLBM_EINVAL, "no domain in slot %d", slot);
    Domain &d = b->doms[slot];
    if (!d.held) return FSLBM_OK;
    d.parity = d.held_cur ^ b->flip;  // current buffer = parity ^ flip
    d.held = false;
    b->dirty = true;
    return FSLBM_OK;
}

// Enqueue the copy of a held domain's state to out.raw on stream s (the
// caller synchronises s before reading out.raw or removing the domain).
int batch_copy_out(fslbm_batch *b, int slot, HostResult &out, hipStream_t s)
{
    if (!b) return FSLBM_EINVAL;
    if (slot < 0 || slot >= (int)b->doms.size() || !b->doms[slot].active || !b->doms[slot].held)
        return bfail(b, FSLBM_EINVAL, "no held domain in slot %d", slot);
    const Domain &d = b->doms[slot];
    const size_t sbytes = 4 * (size_t)d.g.N * storage_bytes(b->precision);
    out.g = d.g;
    out.precision = b->precision;
    out.q = d.q;
    out.mask2d = d.hmask2d;
    out.uref = d.huref;
    out.steps = d.steps;
    out.raw.resize(sbytes);
    BCHK(b, hipMemcpyAsync(out.raw.data(), d.buf[d.held_cur], sbytes, hipMemcpyDeviceToHost, s));
    return FSLBM_OK;
}

// Copy a domain's current stored state to the host and remove it from the
// batch.
int batch_take(fslbm_batch *b, int slot, HostResult &out)
{
    if (!b) return FSLBM_EINVAL;
    if (slot < 0 || slot >= (int)b->doms.size() || !b->doms[slot].active)
        return bfail(b, FSLBM_EINVAL, "no domain in slot %d", slot);
    const Domain &d = b->doms[slot];
    const int cur = d.held ? d.held_cur : d.parity ^ b->flip;
    const size_t sbytes = 4 * (size_t)d.g.N * storage_bytes(b->precision);
    out.g = d.g;
    out.precision = b->precision;
    out.q = d.q;
    out.mask2d = d.hmask2d;
    out.uref = d.huref;
    out.steps = d.steps;
    out.raw.resize(sbytes);
    BCHK(b, hipMemcpyAsync(out.raw.data(), d.buf[cur], sbytes, hipMemcpyDeviceToHost, b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    return fslbm_batch_remove(b, slot);
}

}  // namespace fslbm
