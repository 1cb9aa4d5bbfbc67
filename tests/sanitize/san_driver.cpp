// san_driver.cpp -- runs the CPU code of this repository under
// AddressSanitizer and UndefinedBehaviorSanitizer (tests/test_sanitizers.py
// builds it with tests/sanitize/Makefile and runs it; test harness only):
//  * the oracle (oracle/ifos_oracle.c, float build): whole gradient shots on
//    a small grid -- elastic with and without the free surface, with three
//    relaxation mechanisms, with every DFT and the velocity histories of the
//    time-domain gradient -- and a consistent-adjoint shot;
//  * the SU file code of the library (ifos3d-dev_amd/csrc/suio_host.cpp):
//    a write / read round trip, a file cut short and a missing file.
// Any sanitizer report aborts the run with a non-zero status.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <unistd.h>

#include "ifos_oracle.h"
#include "suio_host.h"

static int failures = 0;
#define CHECK(c)                                                              \
    do {                                                                      \
        if (!(c)) {                                                           \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            ++failures;                                                       \
        }                                                                     \
    } while (0)

static og_grid grid(int free_surf)
{
    og_grid g;
    g.nx = 18; g.ny = 16; g.nz = 20; g.fw = 4;
    g.free_surf = free_surf; g.fdcoeff = 2;
    g.dx = g.dy = g.dz = 2.0f; g.dt = 2e-4f;
    g.nt = 40; g.time = (g.nt - 1) * g.dt;
    return g;
}

static std::vector<float> model(const og_grid &g, float v, int seed)
{
    std::vector<float> a(og_field_size(&g));
    unsigned s = 12345u + 77u * (unsigned)seed;
    for (auto &x : a) {
        s = s * 1664525u + 1013904223u;
        x = v * (1.0f + 0.05f * ((float)(s >> 8) / 16777216.0f - 0.5f));
    }
    return a;
}

static void shot(int free_surf, int L)
{
    const og_grid g = grid(free_surf);
    const size_t n = og_field_size(&g);
    std::vector<float> rho = model(g, 2000.0f, 1), vp = model(g, 4000.0f, 2), vs = model(g, 2300.0f, 3);
    std::vector<float> pi(n), u(n), tp(n, 0.03f), ts(n, 0.025f);
    for (size_t c = 0; c < n; ++c) {
        pi[c] = rho[c] * vp[c] * vp[c];
        u[c] = rho[c] * vs[c] * vs[c];
    }
    const int ntr = 3;
    const int rec[ntr * 3] = {6, 6, 8, 9, 6, 10, 12, 5, 12};
    const float finv[2] = {60.0f, 90.0f};
    std::vector<float> obs(3 * ntr * (g.nt + 1), 0.0f), syn(3 * ntr * (g.nt + 1));
    std::vector<float> g1(n, 0.0f), g2(n, 0.0f), g3(n, 0.0f), ff(6 * 2 * n), fa(6 * 2 * n);
    std::vector<float> hf((size_t)g.nt * 3 * n), ha((size_t)g.nt * 3 * n);
    float l2 = 0.0f;
    og_shot_io io;
    std::memset(&io, 0, sizeof(io));
    io.rho = rho.data(); io.pi = pi.data(); io.u = u.data();
    io.src = og_source{9, 8, 10, 0.0f, 50.0f, 1.0f, 1};
    io.source_shape = 1;
    io.ntr = ntr; io.rec_ijk = rec;
    io.obs_x = obs.data(); io.obs_y = obs.data() + ntr * (g.nt + 1); io.obs_z = obs.data() + 2 * ntr * (g.nt + 1);
    io.finv = finv; io.nf = 2; io.tast = 100;
    io.vppml = 4000.0f; io.fpml = 50.0f;
    io.want_adjoint = 1;
    io.syn_x = syn.data(); io.syn_y = syn.data() + ntr * (g.nt + 1); io.syn_z = syn.data() + 2 * ntr * (g.nt + 1);
    io.l2 = &l2;
    io.grad1 = g1.data(); io.grad2 = g2.data(); io.grad3 = g3.data();
    io.ffwd = ff.data(); io.fadj = fa.data();
    if (L) {
        io.taus = ts.data(); io.taup = tp.data(); io.L = L;
        for (int l = 0; l < L; ++l) io.eta[l] = 0.05f * (float)(l + 1);
    }
    io.vfwd = hf.data(); io.vadj = ha.data();
    CHECK(og_run_shot(&g, &io) == 0);
    CHECK(l2 > 0.0f && std::isfinite(l2));
    og_td_gradient(&g, hf.data(), ha.data(), rho.data(), pi.data(), u.data(), g1.data(), g2.data(), g3.data());
    double s = 0.0;
    for (size_t c = 0; c < n; ++c) s += std::fabs((double)g1[c]) + std::fabs((double)g2[c]) + std::fabs((double)g3[c]);
    CHECK(s > 0.0 && std::isfinite(s));
    std::printf("shot free_surf=%d L=%d: l2 %.4e, |grad| %.4e\n", free_surf, L, (double)l2, s);
}

static void adjoint_shot()
{
    const og_grid g = grid(0);
    const size_t n = og_field_size(&g);
    std::vector<float> rho = model(g, 2000.0f, 4), vp = model(g, 4000.0f, 5), vs = model(g, 2300.0f, 6);
    std::vector<float> pi(n), u(n);
    for (size_t c = 0; c < n; ++c) {
        pi[c] = rho[c] * vp[c] * vp[c];
        u[c] = rho[c] * vs[c] * vs[c];
    }
    const int ntr = 2;
    const int rec[ntr * 3] = {6, 6, 8, 12, 5, 12};
    std::vector<float> obs(3 * ntr * (g.nt + 1), 0.0f), syn(3 * ntr * (g.nt + 1));
    std::vector<float> g1(n, 0.0f), g2(n, 0.0f), g3(n, 0.0f);
    double J = 0.0;
    og_adj_io io;
    std::memset(&io, 0, sizeof(io));
    io.rho = rho.data(); io.pi = pi.data(); io.u = u.data();
    io.src = og_source{9, 8, 10, 0.0f, 50.0f, 1.0f, 1};
    io.source_shape = 1;
    io.ntr = ntr; io.rec_ijk = rec;
    io.obs_x = obs.data(); io.obs_y = obs.data() + ntr * (g.nt + 1); io.obs_z = obs.data() + 2 * ntr * (g.nt + 1);
    io.vppml = 4000.0f; io.fpml = 50.0f;
    io.syn_x = syn.data(); io.syn_y = syn.data() + ntr * (g.nt + 1); io.syn_z = syn.data() + 2 * ntr * (g.nt + 1);
    io.J = &J;
    io.grad1 = g1.data(); io.grad2 = g2.data(); io.grad3 = g3.data();
    CHECK(og_run_shot_adj(&g, &io) == 0);
    CHECK(J > 0.0 && std::isfinite(J));
    std::printf("consistent adjoint: J %.4e\n", J);
}

static void su_files(const char *dir)
{
    const int ntr = 4, ns = 37;
    std::vector<float> sec(ntr * (ns + 1)), back(ntr * (ns + 1));
    for (size_t c = 0; c < sec.size(); ++c) sec[c] = (float)c * 0.25f - 3.0f;
    const int rec[ntr * 3] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12};
    const int id[ntr] = {3, 1, 4, 2};
    const float src[3] = {10.0f, 4.0f, 12.0f};
    char err[256];
    const std::string path = std::string(dir) + "/san_vx.su";
    CHECK(ifos::su_write_host(path.c_str(), sec.data(), ntr, ns, rec, id, src, 2.0f, 2.0f, 2.0f, 2e-4f, 1, err,
                              sizeof(err)) == 0);
    CHECK(ifos::su_read_host(path.c_str(), ntr, ns, back.data(), err, sizeof(err)) == 0);
    for (int t = 0; t < ntr; ++t)
        for (int j = 1; j <= ns; ++j) CHECK(back[t * (ns + 1) + j] == sec[t * (ns + 1) + j]);
    // cut the file inside trace 2: the rest of trace 2 and traces 3, 4 read as 0
    const long cut = 2 * (240 + 4 * ns) + 240 + 4 * 10;
    CHECK(truncate(path.c_str(), cut) == 0);
    CHECK(ifos::su_read_host(path.c_str(), ntr, ns, back.data(), err, sizeof(err)) == 0);
    CHECK(back[2 * (ns + 1) + 10] == sec[2 * (ns + 1) + 10] && back[2 * (ns + 1) + 11] == 0.0f);
    CHECK(back[3 * (ns + 1) + 5] == 0.0f);
    const std::string missing = std::string(dir) + "/no_such_file.su";
    CHECK(ifos::su_read_host(missing.c_str(), ntr, ns, back.data(), err, sizeof(err)) < 0);
    CHECK(std::strstr(err, "cannot open") != nullptr);
    CHECK(ifos::su_write_host(path.c_str(), sec.data(), ntr, 70000, rec, id, src, 2.0f, 2.0f, 2.0f, 2e-4f, 1, err,
                              sizeof(err)) < 0);
    std::printf("SU files: round trip, short file, missing file, bad ns\n");
}

int main(int argc, char **argv)
{
    const char *dir = argc > 1 ? argv[1] : ".";
    shot(0, 0);
    shot(1, 0);
    shot(0, 3);
    adjoint_shot();
    su_files(dir);
    if (failures) {
        std::fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    std::printf("sanitized run: OK\n");
    return 0;
}
