// fslbm_run: drop-in replacement of the reference driver main()
// (PipeDemoStandalone/FSLBM.cu:489-909), on the C ABI of libfslbm.so.
//
// Same working-directory protocol:
//   in   param.txt  tblend pin nu d1 delta blend [path]   (FSLBM.cu:499-510)
//        pipe.txt   lx*ly*lz mask codes 0/1/125, one per line, x fastest
//                   (pipe.m:62, FSLBM.cu:478-485)
//        p3Dii ux3Dii uy3Dii uz3Dii, _*ii.txt, _avii.txt, qpii.txt:
//                   restart files (FSLBM.cu:386-468), used when present
//   out  qp.txt     q1 q2 p1 p2 f.f. probe every 50 steps (FSLBM.cu:663-666)
//        _p _ux _uy _uz _p2 _ux2 _uy2 _uz2 _uxy2 _uxz2 _uyz2 .txt, _av.txt
//                   averaged YZ plane statistics every 500 steps (:692-764)
//        p ux uy uz IO .txt (YZ plane x=lx/2), ux1 uy1 uz1 p1 .txt (XY
//                   plane z=lz/2) every 500 steps (:772-821)
//        p3Dii ... _avii.txt, qpii.txt backup every 50000 steps (:825-904)
//        fslbm_ckpt.bin  exact checkpoint written with each backup: the
//                   state and the statistics sums as doubles, the step
//                   number, the pending gauge and the context's exact
//                   snapshot (fslbm_snapshot_save).  A run started without
//                   --fresh resumes from it when present (the reference's
//                   float32 volumes and %e averages are the fallback), and
//                   continues the step numbering, so an interrupted run
//                   repeats a continuous one bitwise in every precision
//                   mode (a resume in another mode starts from the doubles).
// The geometry sizes, hard-coded in the reference (FSLBM.cu:512-514), are
// options here.  param.txt is re-read every 500 steps (FSLBM.cu:669-677).
// The statistics and plane files are read from the device at their step
// and written by a background thread while the steps go on (the text is
// the reference's %e format); IO.txt, the constant mask plane, is written
// once.
//
// usage: fslbm_run [--lx 120] [--ly 120] [--lz 320] [--d2 10]
//                  [--steps N] [--precision f32|f64|f16s|bf16s]
//                  [--monitor 50] [--stats 500] [--backup 50000]
//                  [--fresh] [--quiet]

#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fslbm.h"

namespace {

struct Opts {
    int lx = 120, ly = 120, lz = 320, d2 = 10;
    long steps = 500000000L;  // the reference's cycle count (FSLBM.cu:516)
    int precision = FSLBM_F32;
    int monitor = 50, stats = 500;
    long backup = 50000;
    bool fresh = false, quiet = false;
};

[[noreturn]] void die(const char *msg, fslbm_t ctx = nullptr)
{
    std::fprintf(stderr, "fslbm_run: %s%s%s\n", msg, ctx ? ": " : "",
                 ctx ? fslbm_last_error(ctx) : "");
    std::exit(1);
}

void check(int rc, fslbm_t ctx, const char *what)
{
    if (rc != FSLBM_OK) {
        std::fprintf(stderr, "fslbm_run: %s failed (%s): %s\n", what,
                     fslbm_error_string(rc), fslbm_last_error(ctx));
        std::exit(1);
    }
}

bool file_exists(const std::string &p)
{
    FILE *f = std::fopen(p.c_str(), "rb");
    if (!f) return false;
    std::fclose(f);
    return true;
}

// param.txt: %e %e %e %d %d %e then the rest (FSLBM.cu:502-509).
fslbm_params read_params()
{
    FILE *f = std::fopen("param.txt", "r");
    if (!f) die("cannot open param.txt");
    float tb, pin, nu, blend;
    int d1, delta;
    if (std::fscanf(f, "%e %e %e %d %d %e", &tb, &pin, &nu, &d1, &delta, &blend) != 6) {
        std::fclose(f);
        die("param.txt: expected tblend pin nu d1 delta blend");
    }
    std::fclose(f);
    fslbm_params p;
    p.tblend = tb;
    p.pin = pin;
    p.nu = nu;
    p.d1 = d1;
    p.delta = delta;
    p.blend = blend;
    return p;
}

// pipe.txt: integers, one per line (the reference's read_geom reads one
// value too many; here exactly lx*ly*lz values are required).
std::vector<uint8_t> read_mask(const Opts &o)
{
    FILE *f = std::fopen("pipe.txt", "rb");
    if (!f) die("cannot open pipe.txt");
    std::string text;
    char buf[1 << 16];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, k);
    std::fclose(f);
    const size_t n = (size_t)o.lx * o.ly * o.lz;
    std::vector<uint8_t> m(n);
    const char *p = text.data(), *end = p + text.size();
    for (size_t i = 0; i < n; i++) {
        while (p < end && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) p++;
        if (p == end) die("pipe.txt: fewer values than lx*ly*lz");
        int v = 0;
        const auto r = std::from_chars(p, end, v);
        if (r.ec != std::errc()) die("pipe.txt: not an integer");
        if (v != 0 && v != 1 && v != 125) die("pipe.txt: mask codes must be 0, 1 or 125");
        p = r.ptr;
        m[i] = (uint8_t)v;
    }
    return m;
}

// One value as printf's "%e\n" writes it (six decimals, exponent of at
// least two digits: std::to_chars rounds the same exact value the same
// way), without printf's per-call cost.
void append_e(std::string &out, double v)
{
    char b[48];
    const auto r = std::to_chars(b, b + sizeof b, v, std::chars_format::scientific, 6);
    out.append(b, r.ptr);
    out.push_back('\n');
}

void write_text(const std::string &name, const std::string &text)
{
    FILE *f = std::fopen(name.c_str(), "wt");
    if (!f || std::fwrite(text.data(), 1, text.size(), f) != text.size())
        die("cannot write output file");
    std::fclose(f);
}

// A plane as the reference writes it: float values, one "%e" per line.
void write_plane(const char *name, const std::vector<double> &v)
{
    std::string t;
    t.reserve(v.size() * 14);
    for (double x : v) append_e(t, (float)x);
    write_text(name, t);
}

// Writes one job at a time on a thread of its own; a new job waits for
// the previous one.
struct BackgroundWriter {
    std::thread t;
    void run(std::function<void()> job)
    {
        wait();
        t = std::thread(std::move(job));
    }
    void wait()
    {
        if (t.joinable()) t.join();
    }
    ~BackgroundWriter() { wait(); }
};

void copy_file(const char *src, const char *dst)
{
    FILE *a = std::fopen(src, "rb");
    if (!a) return;
    FILE *b = std::fopen(dst, "wb");
    if (!b) { std::fclose(a); return; }
    char buf[65536];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof buf, a)) > 0) std::fwrite(buf, 1, k, b);
    std::fclose(a);
    std::fclose(b);
}

constexpr char CKPT_MAGIC[8] = {'F', 'S', 'L', 'B', 'M', 'C', 'K', '1'};

struct CkptHeader {
    char magic[8];
    int32_t nx, ny, nz, precision;
    int64_t step;
    double gauge;
    int32_t stats_n, cnt, cnt2, pad;
};

void write_checkpoint(fslbm_t ctx, int nx, int ny, int nz, int precision, long step,
                      double gauge)
{
    const size_t N = (size_t)nx * ny * nz;
    std::vector<double> st[4];
    for (auto &v : st) v.resize(N);
    check(fslbm_get_state(ctx, st[0].data(), st[1].data(), st[2].data(), st[3].data()), ctx,
          "checkpoint state");
    CkptHeader h{};
    std::memcpy(h.magic, CKPT_MAGIC, 8);
    h.nx = nx;
    h.ny = ny;
    h.nz = nz;
    h.precision = precision;
    h.step = step;
    h.gauge = gauge;
    h.stats_n = fslbm_stats_points(ctx);
    std::vector<double> acc(11 * (size_t)h.stats_n);
    int cnt = 0, cnt2 = 0;
    check(fslbm_stats_read(ctx, acc.data(), &cnt, &cnt2), ctx, "checkpoint stats");
    h.cnt = cnt;
    h.cnt2 = cnt2;
    // the exact snapshot (stored state, 16-bit scales, rounding counter)
    // follows, so a resume in the same precision mode continues bitwise
    const long long sn = fslbm_snapshot_size(ctx);
    if (sn < 0) die("checkpoint snapshot size");
    std::vector<char> snap((size_t)sn);
    check(fslbm_snapshot_save(ctx, snap.data(), sn), ctx, "checkpoint snapshot");
    const int64_t sn64 = sn;
    FILE *f = std::fopen("fslbm_ckpt.bin.tmp", "wb");
    if (!f) die("cannot write fslbm_ckpt.bin");
    bool ok = std::fwrite(&h, sizeof h, 1, f) == 1;
    for (auto &v : st) ok = ok && std::fwrite(v.data(), sizeof(double), N, f) == N;
    ok = ok && std::fwrite(acc.data(), sizeof(double), acc.size(), f) == acc.size();
    ok = ok && std::fwrite(&sn64, sizeof sn64, 1, f) == 1 &&
         std::fwrite(snap.data(), 1, snap.size(), f) == snap.size();
    ok = std::fclose(f) == 0 && ok;
    if (!ok || std::rename("fslbm_ckpt.bin.tmp", "fslbm_ckpt.bin") != 0)
        die("writing fslbm_ckpt.bin failed");
}

// Returns the step number of the checkpoint (state, statistics and gauge
// loaded into ctx), or -1 when there is none.
long read_checkpoint(fslbm_t ctx, int nx, int ny, int nz, int precision)
{
    FILE *f = std::fopen("fslbm_ckpt.bin", "rb");
    if (!f) return -1;
    CkptHeader h{};
    if (std::fread(&h, sizeof h, 1, f) != 1 || std::memcmp(h.magic, CKPT_MAGIC, 8) != 0) {
        std::fclose(f);
        die("fslbm_ckpt.bin: not a checkpoint");
    }
    if (h.nx != nx || h.ny != ny || h.nz != nz) {
        std::fclose(f);
        die("fslbm_ckpt.bin: written for another geometry");
    }
    if (h.precision != precision)
        std::fprintf(stderr, "fslbm_run: checkpoint written in precision mode %d, resuming in %d\n",
                     h.precision, precision);
    const size_t N = (size_t)nx * ny * nz;
    std::vector<double> st[4];
    bool ok = true;
    for (auto &v : st) {
        v.resize(N);
        ok = ok && std::fread(v.data(), sizeof(double), N, f) == N;
    }
    if (h.stats_n != fslbm_stats_points(ctx)) ok = false;
    std::vector<double> acc(11 * (size_t)(h.stats_n > 0 ? h.stats_n : 0));
    ok = ok && std::fread(acc.data(), sizeof(double), acc.size(), f) == acc.size();
    // the exact snapshot, when the file has one for this precision mode
    std::vector<char> snap;
    int64_t sn = 0;
    if (ok && h.precision == precision && std::fread(&sn, sizeof sn, 1, f) == 1 &&
        sn == fslbm_snapshot_size(ctx)) {
        snap.resize((size_t)sn);
        if (std::fread(snap.data(), 1, snap.size(), f) != snap.size()) snap.clear();
    }
    std::fclose(f);
    if (!ok) die("fslbm_ckpt.bin: truncated");
    check(fslbm_set_state(ctx, st[0].data(), st[1].data(), st[2].data(), st[3].data()), ctx,
          "checkpoint state");
    if (!snap.empty())
        check(fslbm_snapshot_load(ctx, snap.data(), (long long)snap.size()), ctx,
              "checkpoint snapshot");
    check(fslbm_stats_load(ctx, acc.data(), h.cnt, h.cnt2), ctx, "checkpoint stats");
    if (h.gauge != 0.0) check(fslbm_set_gauge(ctx, h.gauge), ctx, "checkpoint gauge");
    return (long)h.step;
}

const char *const AVG_NAMES[11] = {"_p", "_ux", "_uy", "_uz", "_p2", "_ux2",
                                   "_uy2", "_uz2", "_uxy2", "_uxz2", "_uyz2"};

// FSLBM.cu:752-764 (and the backup copy, :886-899).
// The statistics sums of one sample step, read from the device.
struct Averages {
    int n = 0, cnt = 0, cnt2 = 0;
    std::vector<double> acc;  // 11 arrays of n sums
};

Averages read_averages(fslbm_t ctx)
{
    Averages a;
    a.n = fslbm_stats_points(ctx);
    a.acc.resize(11 * (size_t)a.n);
    check(fslbm_stats_read(ctx, a.acc.data(), &a.cnt, &a.cnt2), ctx, "stats read");
    return a;
}

// sum/cnt of each statistic, one "%e" per line, and the counts
// (FSLBM.cu:752-764).
void write_average_files(const Averages &a, const char *suffix)
{
    std::string t;
    for (int k = 0; k < 11; k++) {
        const double c = k < 4 ? (double)a.cnt : (double)a.cnt2;
        t.clear();
        t.reserve((size_t)a.n * 14);
        for (int i = 0; i < a.n; i++) append_e(t, a.acc[k * (size_t)a.n + i] / c);
        write_text(std::string(AVG_NAMES[k]) + suffix + ".txt", t);
    }
    write_text(std::string("_av") + suffix + ".txt",
               std::to_string(a.cnt) + " " + std::to_string(a.cnt2) + "\n");
}

void write_averages(fslbm_t ctx, const char *suffix)
{
    write_average_files(read_averages(ctx), suffix);
}

// FSLBM.cu:429-468: averages are restored as value * count.
void load_averages(fslbm_t ctx)
{
    FILE *f = std::fopen("_avii.txt", "r");
    if (!f) return;
    int cnt = 0, cnt2 = 0;
    if (std::fscanf(f, "%d %d", &cnt, &cnt2) != 2) { std::fclose(f); return; }
    std::fclose(f);
    const int n = fslbm_stats_points(ctx);
    std::vector<double> acc(11 * (size_t)n, 0.0);
    for (int k = 0; k < 11; k++) {
        const std::string name = std::string(AVG_NAMES[k]) + "ii.txt";
        FILE *g = std::fopen(name.c_str(), "r");
        if (!g) die("restart: missing an averages file");
        const double c = k < 4 ? (double)cnt : (double)cnt2;
        for (int i = 0; i < n; i++) {
            float v;
            if (std::fscanf(g, "%e", &v) != 1) { std::fclose(g); die("restart: short averages file"); }
            acc[k * (size_t)n + i] = (double)v * c;
        }
        std::fclose(g);
    }
    check(fslbm_stats_load(ctx, acc.data(), cnt, cnt2), ctx, "stats load");
}

}  // namespace

int main(int argc, char **argv)
{
    Opts o;
    for (int i = 1; i < argc; i++) {
        auto next = [&](const char *name) -> const char * {
            if (i + 1 >= argc) { std::fprintf(stderr, "missing value for %s\n", name); std::exit(2); }
            return argv[++i];
        };
        const std::string a = argv[i];
        if (a == "--lx") o.lx = std::atoi(next("--lx"));
        else if (a == "--ly") o.ly = std::atoi(next("--ly"));
        else if (a == "--lz") o.lz = std::atoi(next("--lz"));
        else if (a == "--d2") o.d2 = std::atoi(next("--d2"));
        else if (a == "--steps") o.steps = std::atol(next("--steps"));
        else if (a == "--monitor") o.monitor = std::atoi(next("--monitor"));
        else if (a == "--stats") o.stats = std::atoi(next("--stats"));
        else if (a == "--backup") o.backup = std::atol(next("--backup"));
        else if (a == "--fresh") o.fresh = true;
        else if (a == "--quiet") o.quiet = true;
        else if (a == "--precision") {
            const std::string p = next("--precision");
            o.precision = p == "f64" ? FSLBM_F64 : p == "f16s" ? FSLBM_F16S
                        : p == "bf16s" ? FSLBM_BF16S : FSLBM_F32;
        } else {
            std::fprintf(stderr, "unknown option %s\n", a.c_str());
            return 2;
        }
    }
    if (o.lx < 3 || o.ly < 3 || o.d2 < 1 || o.lz <= 2 * o.d2) die("bad geometry options");
    if (o.monitor < 1 || o.stats < 1 || o.stats % o.monitor != 0)
        die("--stats must be a multiple of --monitor (the reference nests them)");

    fslbm_params prm = read_params();
    const std::vector<uint8_t> padded = read_mask(o);
    const int nx = o.lx, ny = o.ly, nz = o.lz - 2 * o.d2;
    const size_t nxy = (size_t)nx * ny, N = nxy * nz;
    std::vector<uint8_t> mask(N);
    if (fslbm_fold_mask(o.lx, o.ly, o.lz, o.d2, padded.data(), mask.data()) != FSLBM_OK)
        die("pipe.txt: padding planes are not periodic images of the fluid planes");

    fslbm_t ctx = nullptr;
    if (fslbm_create(&prm, mask.data(), nx, ny, nz, o.precision, &ctx) != FSLBM_OK)
        die("create failed", nullptr), (void)0;
    const int r = (o.lx - 2 * o.d2) / 2;  // FSLBM.cu:512 (before the frame)
    long l1 = 0;                          // fluid cells per cross-section
    for (size_t c = 0; c < nxy; c++) l1 += mask[c] == 0;

    // Initial state: restart volumes if present (FSLBM.cu:610), else the
    // initphys ramp p = pin (lz-d2-z)/(lz-2 d2), u = 0 (FSLBM.cu:374-383).
    std::vector<double> st[4];
    for (auto &v : st) v.assign(N, 0.0);
    const char *vol[4] = {"p3Dii", "ux3Dii", "uy3Dii", "uz3Dii"};
    const bool exact = !o.fresh && file_exists("fslbm_ckpt.bin");
    const bool restart = !exact && !o.fresh && file_exists(vol[0]);
    if (exact) {
        copy_file("qpii.txt", "qp.txt");
    } else if (restart) {
        const size_t lxy = (size_t)o.lx * o.ly;
        std::vector<float> plane(lxy);
        for (int f = 0; f < 4; f++) {
            FILE *fp = std::fopen(vol[f], "rb");
            if (!fp) die("restart: missing a volume file");
            for (int z = 0; z < o.lz; z++) {
                if (std::fread(plane.data(), sizeof(float), lxy, fp) != lxy) die("restart: short volume file");
                if (z < o.d2 || z >= o.lz - o.d2) continue;
                for (size_t c = 0; c < lxy; c++) st[f][c + lxy * (z - o.d2)] = plane[c];
            }
            std::fclose(fp);
        }
        copy_file("qpii.txt", "qp.txt");
    } else {
        for (int z = 0; z < nz; z++)
            for (size_t c = 0; c < nxy; c++)
                if (mask[c + nxy * z] == 0)
                    st[0][c + nxy * z] = prm.pin * (double)(nz - z) / (double)nz;
    }
    check(fslbm_set_state(ctx, st[0].data(), st[1].data(), st[2].data(), st[3].data()),
          ctx, "set state");
    for (auto &v : st) std::vector<double>().swap(v);
    check(fslbm_stats_config(ctx, o.lx / 2, o.d2), ctx, "stats config");
    if (restart) load_averages(ctx);
    long c0 = 0;  // steps done before this run (exact checkpoint)
    if (exact) c0 = read_checkpoint(ctx, nx, ny, nz, o.precision);
    { FILE *q = std::fopen("qp.txt", "a"); if (q) std::fclose(q); }

    const int n_yz = ny * o.lz;
    std::vector<double> xy(nxy);
    double pending_gauge = 0.0;  // p2 set by the last monitoring step
    BackgroundWriter writer;     // statistics and plane files
    bool io_written = false;
    // FSLBM.cu:630-764: monitoring, and every `stats` steps the statistics
    // and plane files
    auto monitor_step = [&](long c) {
        // FSLBM.cu:630-666
        double m[4];
        check(fslbm_monitor(ctx, 1, m), ctx, "monitor");
        const double q1 = m[0], q2 = m[1], p1 = m[2], p2 = m[3];
        pending_gauge = p2;
        const double Re = 2.0 * r * q1 / prm.nu;
        const double ff = (p1 - p2) / q1 / q1 / 3.0 * 4.0 * r / (double)(o.lz - 2 * o.d2);
        double probe[4];
        check(fslbm_probe(ctx, o.lx / 2, o.ly / 2, o.lz / 2 - o.d2, probe), ctx, "probe");
        if (!o.quiet)
            std::printf("time step: %ld, dtddy: %e, q1=%e, q2=%e, p1 =%e, p2=%e, f.f.=%e, "
                        "Re=%e, #inl. elements=%ld, #outl. elements=%ld\n",
                        c, prm.nu - 0.16666666, q1, q2, p1, p2, ff, Re, l1, l1);
        FILE *q = std::fopen("qp.txt", "a");
        if (q) {
            std::fprintf(q, "%e %e %e %e %e %e\n", q1, q2, p1, p2, ff, probe[3]);
            std::fclose(q);
        }
        if (c % o.stats == 0) {
            // FSLBM.cu:667-764: reload parameters, accumulate, write
            prm = read_params();
            check(fslbm_set_params(ctx, &prm), ctx, "set params");
            check(fslbm_stats_accumulate(ctx, prm.delta), ctx, "stats");
            // the averages and, FSLBM.cu:772-821, the instantaneous planes and
            // the mask plane: read now, written while the steps go on
            Averages av = read_averages(ctx);
            std::vector<std::vector<double>> planes(8);
            for (int f = 0; f < 4; f++) {
                planes[f].resize(n_yz);
                check(fslbm_get_plane(ctx, 0, o.lx / 2, o.d2, f, planes[f].data()), ctx, "plane");
                planes[4 + f].resize(nxy);
                check(fslbm_get_plane(ctx, 2, o.lz / 2, o.d2, f, planes[4 + f].data()), ctx,
                      "plane");
            }
            std::string io;  // constant: written once
            if (!io_written) {
                for (int z = 0; z < o.lz; z++)
                    for (int y = 0; y < ny; y++)
                        io += std::to_string(padded[o.lx / 2 + (size_t)o.lx * (y + (size_t)ny * z)]) +
                              "\n";
                io_written = true;
            }
            writer.run([av = std::move(av), planes = std::move(planes), io = std::move(io)] {
                static const char *names[8] = {"p.txt",  "ux.txt",  "uy.txt",  "uz.txt",
                                               "p1.txt", "ux1.txt", "uy1.txt", "uz1.txt"};
                write_average_files(av, "");
                for (int k = 0; k < 8; k++) write_plane(names[k], planes[k]);
                if (!io.empty()) write_text("IO.txt", io);
            });
        }
    };
    for (long c = c0 + 1; c <= c0 + o.steps; c++) {
        check(fslbm_step(ctx, 1), ctx, "step");
        pending_gauge = 0.0;
        if (c % o.monitor == 0) monitor_step(c);
        if (o.backup > 0 && c % o.backup == 0) {
            // FSLBM.cu:825-904: raw float32 padded volumes + text averages
            for (int f = 0; f < 4; f++) {
                FILE *fp = std::fopen(vol[f], "wb");
                if (!fp) die("cannot write backup");
                std::vector<float> pl(nxy);
                for (int z = 0; z < o.lz; z++) {
                    check(fslbm_get_plane(ctx, 2, z, o.d2, f, xy.data()), ctx, "plane");
                    for (size_t k = 0; k < nxy; k++) pl[k] = (float)xy[k];
                    std::fwrite(pl.data(), sizeof(float), nxy, fp);
                }
                std::fclose(fp);
            }
            copy_file("qp.txt", "qpii.txt");
            write_averages(ctx, "ii");
            write_checkpoint(ctx, nx, ny, nz, o.precision, c, pending_gauge);
        }
    }
    writer.wait();
    check(fslbm_sync(ctx), ctx, "sync");
    fslbm_destroy(ctx);
    return 0;
}
