// sloth_term: the reference's main loop (main.rs:25-111) over the C-ABI of
// libsloth_hip.so -- terminal mode with resize, key poll and an fps cap,
// image mode, and the `-j` export -- rendering on the GPU and encoding the
// flush bytes there.
//
//   sloth_term [options] FILE...        model files (.obj / .stl), in queue order
//     --x R --y R --z R    turntable offsets: roll, pitch (+ pi), yaw (inputs.rs:131-149)
//     --speed R            radians per second (default 1.0)
//     --no-color           the plain-text flush (context.rs:58-61)
//     --image W[xH]        image mode at W x H (H = W when omitted; inputs.rs:159-169)
//     --export N           with --image: the `-j N` export of N turntable frames
//   options of this program (not in the reference), for scripted runs:
//     --size WxH           a fixed terminal size instead of the terminal's
//     --frames N           stop after N frames
//     --dt S               advance the turntable by speed * S per frame instead
//                          of by the measured frame time
//     --fps F              the frame-rate cap (default 500, main.rs:28-29)
//
// Terminal mode uses raw mode and hides the cursor (main.rs:52-53) when stdin
// is a terminal; 'q' or Ctrl-C quits (main.rs:64-70).
#include <poll.h>
#include <sys/ioctl.h>
#include <termios.h>
#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "sloth_hip.h"

namespace {

struct Options {
    std::vector<std::string> files;
    float x = 0.0f, y = 0.0f, z = 0.0f, speed = 1.0f;
    bool no_color = false, image = false;
    uint32_t img_w = 0, img_h = 0, export_frames = 0;
    uint32_t size_w = 0, size_h = 0;
    long frames = -1;
    float dt = -1.0f, fps = 500.0f;
};

bool parse_size(const char* s, uint32_t& w, uint32_t& h, bool square_ok)
{
    char* end = nullptr;
    unsigned long a = strtoul(s, &end, 10);
    if (end == s) return false;
    if (*end == 'x' || *end == 'X') {
        unsigned long b = strtoul(end + 1, &end, 10);
        if (*end) return false;
        w = (uint32_t)a;
        h = (uint32_t)b;
        return true;
    }
    if (*end || !square_ok) return false;
    w = h = (uint32_t)a;
    return true;
}

int usage()
{
    fprintf(stderr, "usage: sloth_term [--x R] [--y R] [--z R] [--speed R] [--no-color] [--image W[xH] [--export N]]\n"
                    "                  [--size WxH] [--frames N] [--dt S] [--fps F] FILE...\n");
    return 2;
}

int check(int rc, const char* what)
{
    if (rc) fprintf(stderr, "sloth_term: %s: %s\n", what, sloth_last_error());
    return rc;
}

struct Terminal {
    bool raw = false;
    termios saved{};
    void enter()
    {
        if (!isatty(STDIN_FILENO) || tcgetattr(STDIN_FILENO, &saved) != 0) return;
        termios t = saved;
        cfmakeraw(&t);
        if (tcsetattr(STDIN_FILENO, TCSANOW, &t) == 0) raw = true;
    }
    void leave()
    {
        if (raw) tcsetattr(STDIN_FILENO, TCSANOW, &saved);
        raw = false;
    }
    // Waits up to `ms` for a key; true when it asks to quit ('q' or Ctrl-C).
    bool quit_key(int ms)
    {
        if (!raw) {
            if (ms > 0) std::this_thread::sleep_for(std::chrono::milliseconds(ms));
            return false;
        }
        pollfd p{STDIN_FILENO, POLLIN, 0};
        if (poll(&p, 1, ms < 0 ? 0 : ms) <= 0) return false;
        char c = 0;
        if (read(STDIN_FILENO, &c, 1) != 1) return false;
        return c == 'q' || c == 3;
    }
};

void write_all(const char* p, size_t n)
{
    while (n) {
        ssize_t k = write(STDOUT_FILENO, p, n);
        if (k <= 0) return;
        p += k;
        n -= (size_t)k;
    }
}

}  // namespace

int main(int argc, char** argv)
{
    Options o;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> const char* { return i + 1 < argc ? argv[++i] : nullptr; };
        const char* v = nullptr;
        if (a == "--no-color") o.no_color = true;
        else if (a == "--x" && (v = next())) o.x = strtof(v, nullptr);
        else if (a == "--y" && (v = next())) o.y = strtof(v, nullptr);
        else if (a == "--z" && (v = next())) o.z = strtof(v, nullptr);
        else if (a == "--speed" && (v = next())) o.speed = strtof(v, nullptr);
        else if (a == "--image" && (v = next())) {
            o.image = true;
            if (!parse_size(v, o.img_w, o.img_h, true)) return usage();
        } else if (a == "--export" && (v = next())) o.export_frames = (uint32_t)strtoul(v, nullptr, 10);
        else if (a == "--size" && (v = next())) {
            if (!parse_size(v, o.size_w, o.size_h, false)) return usage();
        } else if (a == "--frames" && (v = next())) o.frames = strtol(v, nullptr, 10);
        else if (a == "--dt" && (v = next())) o.dt = strtof(v, nullptr);
        else if (a == "--fps" && (v = next())) o.fps = strtof(v, nullptr);
        else if (!a.empty() && a[0] == '-') return usage();
        else o.files.push_back(a);
    }
    if (o.files.empty() || (o.export_frames && !o.image)) return usage();
    if (check(sloth_init(0), "sloth_init")) return 1;

    // match_meshes (inputs.rs:93-129): every file's meshes, in order
    std::vector<sloth_mesh*> meshes;
    std::vector<float> bbox_max;
    for (const std::string& f : o.files) {
        sloth_model* model = nullptr;
        if (check(sloth_model_load(f.c_str(), &model), f.c_str())) return 1;
        uint32_t n = 0;
        sloth_model_mesh_count(model, &n);
        for (uint32_t i = 0; i < n; ++i) {
            sloth_mesh* m = nullptr;
            if (check(sloth_model_upload(model, i, &m), "sloth_model_upload")) return 1;
            float lo[3], hi[3];
            sloth_mesh_bbox(m, lo, hi);
            bbox_max.insert(bbox_max.end(), hi, hi + 3);
            meshes.push_back(m);
        }
        sloth_model_destroy(model);
    }

    // match_turntable (inputs.rs:131-149): pitch offset + pi, in binary32
    float turntable[4] = {o.x, o.y, o.z, o.speed};
    turntable[1] += (float)M_PI;
    const bool webify = o.image && o.export_frames > 0;
    const int mode = o.no_color ? SLOTH_ENCODE_TEXT : webify ? SLOTH_ENCODE_HTML : SLOTH_ENCODE_ANSI;
    const auto frame_time = std::chrono::duration<double>(1.0 / (double)o.fps);

    Terminal term;
    if (!o.image) {
        term.enter();
        write_all("\x1b[?25l", 6);   // crossterm cursor::Hide (main.rs:53)
    }
    if (webify) {
        const char head[] = "let frames = [\n";
        write_all(head, sizeof head - 1);
        turntable[3] = (2.0f * (float)M_PI) * (1.0f / (float)o.export_frames);   // main.rs:57
    }

    sloth_context* ctx = nullptr;
    uint32_t cw = 0, ch = 0;
    std::vector<char> out;
    long frame = 0;
    uint32_t export_count = 0;
    int rc = 0;
    for (;;) {
        const auto last = std::chrono::steady_clock::now();
        if (!o.image) {
            const auto wait = frame_time - (std::chrono::steady_clock::now() - last);
            if (term.quit_key((int)std::chrono::duration_cast<std::chrono::milliseconds>(wait).count())) break;
        }
        // Context::update (context.rs:93-141): the size comes from the
        // terminal in terminal mode, and the context follows it
        uint32_t w = o.img_w, h = o.img_h;
        if (!o.image) {
            w = o.size_w;
            h = o.size_h;
            winsize ws{};
            if (!w && ioctl(STDOUT_FILENO, TIOCGWINSZ, &ws) == 0 && ws.ws_col && ws.ws_row) {
                w = ws.ws_col;
                h = ws.ws_row;
            }
            if (!w || !h) {
                w = 80;
                h = 24;
            }
        }
        if (!ctx || w != cw || h != ch) {
            if (ctx) sloth_context_destroy(ctx);
            ctx = nullptr;
            if ((rc = check(sloth_context_create(w, h, o.image ? 1 : 0, &ctx), "sloth_context_create"))) break;
            cw = w;
            ch = h;
        }
        float rot[16];
        sloth_rotation_euler(turntable[0], turntable[1], turntable[2], rot);
        if ((rc = check(sloth_context_update(ctx, bbox_max.data(), (uint32_t)meshes.size()), "update"))) break;
        if ((rc = check(sloth_context_clear(ctx), "clear"))) break;
        for (sloth_mesh* m : meshes)
            if ((rc = check(sloth_draw_mesh(ctx, m, rot, nullptr), "sloth_draw_mesh"))) break;
        if (rc) break;
        if (webify) write_all("`\n", 2);
        uint64_t len = 0;
        if ((rc = check(sloth_context_encode(ctx, mode, nullptr, 0, &len), "encode"))) break;
        out.resize(len);
        if ((rc = check(sloth_context_encode(ctx, mode, out.data(), len, &len), "encode"))) break;
        write_all(out.data(), len);
        // main.rs:89-93: the turntable advances by its speed times the frame
        // time (the export: by its fixed step)
        float dt = o.dt;
        if (dt < 0.0f)
            dt = (float)std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - last)
                     .count() /
                 1000000000.0f;
        turntable[1] += webify ? turntable[3] : turntable[3] * dt;
        if (webify) {
            if (turntable[1] > 9.42477f || o.export_frames - 1 == export_count) {
                write_all("`];\n", 4);
                break;
            }
            write_all("`,\n", 3);
            ++export_count;
        }
        ++frame;
        if (o.image && !webify) break;
        if (o.frames >= 0 && frame >= o.frames) break;
    }
    if (!o.image) {
        write_all("\x1b[?25h", 6);   // cursor::Show (main.rs:67)
        term.leave();
    }
    if (ctx) sloth_context_destroy(ctx);
    for (sloth_mesh* m : meshes) sloth_mesh_destroy(m);
    return rc ? 1 : 0;
}
