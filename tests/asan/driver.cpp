// Host-side exercise of the CPU oracle (oracle/oracle.c) and of the
// library's host-only sources (csrc/io.cpp, csrc/host.cpp), built with
// AddressSanitizer by tests/test_host_asan.py: every routine runs on small
// grids, near its minimum sizes, so out-of-bounds reads or writes, use after
// free and leaks of the host code are reported.  Exits 0 when it gets through.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/mgad.h"
#include "../../oracle/oracle.h"

namespace {

std::vector<double> field(size_t n, unsigned seed, double lo = -1.0, double hi = 2.0) {
    std::vector<double> v(n);
    unsigned long long s = seed * 0x9E3779B97F4A7C15ull + 1;
    for (auto &x : v) {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        x = lo + (hi - lo) * static_cast<double>(s >> 11) * 0x1.0p-53;
    }
    return v;
}

int fail(const char *what) {
    std::fprintf(stderr, "asan driver: %s\n", what);
    return 1;
}

}  // namespace

int main(int argc, char **argv) {
    const char *tmp = argc > 1 ? argv[1] : "asan_driver.bin";
    const orc_scheme z1{ORC_WENO_Z, 1, 1e-6, ORC_LF_LOCAL, 0.0}, z2{ORC_WENO_Z, 2, 1e-40, ORC_LF_GLOBAL, 0.0},
        js{ORC_WENO_JS, 2, 1e-6, ORC_LF_GLOBAL, 0.0}, b5s{ORC_WENO_Z_B5S, 1, 1e-6, ORC_LF_LOCAL, 0.0};
    // faces and sides
    const auto v = field(5, 1), w = field(5, 2);
    double g[5];
    for (const orc_scheme *s : {&z1, &z2, &js, &b5s}) {
        if (!std::isfinite(orc_face(v.data(), w.data(), s))) return fail("face");
        if (!std::isfinite(orc_side_grad(v.data(), 1, s, g))) return fail("side_grad");
    }
    // drop-in sweeps, heat operators and RK on the smallest grids they take
    for (unsigned n : {7u, 9u}) {
        const unsigned nx = n, ny = n + 1, nz = n + 2;
        const size_t N = static_cast<size_t>(nx) * ny * nz;
        auto u = field(N, 3), uo = field(N, 4), Lu = field(N, 5);
        orc_adv_x(nx, ny, nz, 0.1, u.data(), Lu.data(), &z1);
        orc_adv_y(nx, ny, nz, 0.1, u.data(), Lu.data(), &z1);
        orc_adv_z(nx, ny, nz, 0.1, u.data(), Lu.data(), &z1);
        orc_diff13_b5(nx, ny, nz, 0.1, 0.1, 0.1, u.data(), Lu.data());
        orc_ssprk_b5(nx, ny, nz, 2, 0.01, u.data(), uo.data(), Lu.data());
        orc_lu3d_13(nx, ny, nz, 0.1, 0.1, 0.1, u.data(), Lu.data());
        orc_lu3d_o2(nx, ny, nz, 0.1, 0.1, 0.1, u.data(), Lu.data());
        orc_lu3d_blk(nx, ny, nz, 0.1, 0.1, 0.1, u.data(), Lu.data());
        orc_rk3d(3, nx, ny, nz, 0.01, u.data(), uo.data(), Lu.data());
        for (int axis = 0; axis < 3; axis++) orc_adv_b5s(axis, nx, ny, nz, 0.1, u.data(), Lu.data(), &b5s);
        orc_md3_lap(2, nx, nx, ny, nz, 1, nz - 1, 0.1, 0.1, 0.1, u.data(), Lu.data());
        orc_md3_lap(4, nx, nx, ny, nz, 2, nz - 2, 0.1, 0.1, 0.1, u.data(), Lu.data());
        orc_md3_rk(1, nx, nx, ny, nz, 0.01, u.data(), uo.data(), Lu.data());
        for (int axis = 0; axis < 3; axis++)
            orc_mb3_adv(axis, nx, nx, ny, nz, 3, nz - 3, 0.1, u.data(), Lu.data(), &js);
        orc_mb3_rk(1, nx, nx, ny, nz, 0.01, u.data(), uo.data(), Lu.data());
        auto h = field(N, 6, 0.0, 1.0);
        orc_heat7_fe(nx, ny, nz, 0.01, 0.01, 0.01, 0.1, 3, h.data());
    }
    // periodic Burgers step, run and adjoint (4 is the smallest periodic grid)
    {
        const unsigned nx = 4, ny = 5, nz = 6;
        const size_t N = static_cast<size_t>(nx) * ny * nz;
        auto u = field(N, 7, 0.0, 1.0), Lu = std::vector<double>(N), work = std::vector<double>(2 * N);
        orc_rhs_periodic(nx, ny, nz, 0.5, 0.4, 0.3, u.data(), Lu.data(), &z2);
        orc_burgers_step(nx, ny, nz, 0.5, 0.4, 0.3, 0.4, 0.0, 1.0, &z2, u.data(), work.data());
        double t = 0.0;
        orc_burgers_run(nx, ny, nz, 0.5, 0.4, 0.3, 0.4, &t, 0.2, -1, &js, u.data());
        auto lam = field(N, 8), ub = std::vector<double>(N);
        double l1 = 0, l2 = 0;
        orc_rhs_vjp(nx, ny, nz, 0.5, 0.4, 0.3, u.data(), lam.data(), ub.data(), &z2, &l1, &l2);
        auto u0 = field(N, 9, 0.0, 1.0), tg = field(N, 10, 0.0, 1.0), gr = std::vector<double>(N);
        int steps = 0;
        const double J = orc_burgers_adjoint(nx, ny, nz, 0.5, 0.4, 0.3, 0.4, 0.0, 0.3, &z2, u0.data(), tg.data(),
                                             gr.data(), &steps);
        if (!std::isfinite(J) || steps < 1) return fail("adjoint");
        auto r1 = field(32, 11), l1d = std::vector<double>(32);
        orc_rhs_1d_periodic(32, 0.1, r1.data(), l1d.data(), &z1);
        // WENO7 prototype
        auto w7 = field(N, 12, 0.0, 1.0), res = std::vector<double>(N);
        orc_weno7_rhs(nx, ny, nz, 0.5, 0.4, 0.3, w7.data(), res.data());
        double t7 = 0.0;
        std::vector<double> res7(N, 0.0);
        orc_weno7_run(nx, ny, nz, 0.5, 0.4, 0.3, 0.4, &t7, 0.1, 3, w7.data(), res7.data());
    }
    // host library: file format, initial conditions, error norms, slabs
    {
        const unsigned nx = 6, ny = 5, nz = 4;
        const size_t N = static_cast<size_t>(nx) * ny * nz;
        auto u = field(N, 13), back = std::vector<double>(N);
        if (mgad_save_binary3d(u.data(), nx, ny, nz, tmp) != MGAD_OK) return fail("save");
        if (mgad_load_binary3d(back.data(), nx, ny, nz, tmp) != MGAD_OK) return fail("load");
        std::remove(tmp);
        for (int ic = 1; ic <= 3; ic++) {
            mgad_init3d_burgers(ic, u.data(), 0.1, 0.1, 0.1, nx, ny, nz);
            mgad_init3d_heat(ic, u.data(), 0.1, 0.1, 0.1, nx, ny, nz);
        }
        mgad_get_last_error();
        double norms[3];
        mgad_heat_error_norms(u.data(), 0.1, 0.1, 0.1, 0.1, nx, ny, nz, norms);
        const unsigned nr = 2, Nzs = 3;  // two ranks of 3 planes, 3 ghost planes each side
        auto q = field(static_cast<size_t>(nx) * ny * (nr * Nzs + 6), 14);
        std::vector<double> sq(static_cast<size_t>(nx) * ny * (Nzs + 6));
        std::vector<double> merged(q.size());
        for (unsigned r = 0; r < nr; r++) {
            mgad_init_subdomain(q.data(), sq.data(), r, nx, ny, Nzs);
            mgad_merge_domains(sq.data(), merged.data(), r, nx, ny, Nzs);
        }
        auto f = field(static_cast<size_t>(9) * 10 * 11, 15);
        mgad_ic_fourier(f.data(), 9, 10, 11, -1.0, 1.0, 2, 8, 4, 0.5, 0.5);
        mgad_ic_gauss_noise(f.data(), 9, 10, 11, 0.1, 1e-3, 1);
        CopyToConstantMemory(0.1, 0.2, 0.3);
        for (int c = 0; c <= 6; c++)
            if (!mgad_error_string(c)) return fail("error string");
        if (!mgad_version()) return fail("version");
        mgad_get_last_error();
    }
    std::puts("asan driver ok");
    return 0;
}
