// io.cpp -- host-side file format, reference initial conditions and error
// norms, so runs of this library can be written and checked the way the
// reference's drivers do it.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/mgad.h"

namespace mgad {
void set_error(int code);  // host.cpp: the sticky error code
}
using mgad::set_error;

namespace {

bool dims_ok(unsigned nx, unsigned ny, unsigned nz) {
    return nx > 0 && ny > 0 && nz > 0 && 1ull * nx * ny * nz < (1ull << 40);
}

}  // namespace

// SaveBinary3D, SingleGPU/Burgers3d_WENO5/tools.cpp:59-87: the field as
// 32-bit floats, x fastest, no header (the reference always writes floats and
// names the file result.bin; path NULL keeps that name).
extern "C" int mgad_save_binary3d(const double *u, unsigned int nx, unsigned int ny, unsigned int nz,
                                  const char *path) {
    if (!u || !dims_ok(nx, ny, nz)) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    FILE *f = std::fopen(path ? path : "result.bin", "wb");
    if (!f) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    const size_t n = static_cast<size_t>(nx) * ny * nz, chunk = 1 << 20;
    std::vector<float> buf(std::min(n, chunk));
    size_t done = 0;
    bool ok = true;
    while (done < n && ok) {
        const size_t m = std::min(chunk, n - done);
        for (size_t i = 0; i < m; i++) buf[i] = static_cast<float>(u[done + i]);
        ok = std::fwrite(buf.data(), sizeof(float), m, f) == m;
        done += m;
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    return MGAD_OK;
}

// Reads a file written by SaveBinary3D back into doubles.
extern "C" int mgad_load_binary3d(double *u, unsigned int nx, unsigned int ny, unsigned int nz, const char *path) {
    if (!u || !path || !dims_ok(nx, ny, nz)) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    FILE *f = std::fopen(path, "rb");
    if (!f) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    const size_t n = static_cast<size_t>(nx) * ny * nz, chunk = 1 << 20;
    std::vector<float> buf(std::min(n, chunk));
    size_t done = 0;
    bool ok = true;
    while (done < n && ok) {
        const size_t m = std::min(chunk, n - done);
        ok = std::fread(buf.data(), sizeof(float), m, f) == m;
        for (size_t i = 0; ok && i < m; i++) u[done + i] = static_cast<double>(buf[i]);
        done += m;
    }
    // the file must hold exactly n values
    ok = ok && std::fgetc(f) == EOF;
    std::fclose(f);
    if (!ok) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    return MGAD_OK;
}

// Call_Init3d, SingleGPU/Burgers3d_WENO5/tools.cpp:92-140:
//   ic 1: u = 1 where 0.4 n < index < 0.6 n in all three directions, else 0;
//   ic 2: u = 0;
//   ic 3: u = exp(-((x-1)^2 + (y-1)^2 + (z-1)^2)/0.1), x = i dx (Burgers.h:41).
extern "C" int mgad_init3d_burgers(int ic, double *u0, double dx, double dy, double dz, unsigned int nx,
                                   unsigned int ny, unsigned int nz) {
    if (!u0 || !dims_ok(nx, ny, nz) || ic < 1 || ic > 3) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    const size_t xy = static_cast<size_t>(nx) * ny;
    for (unsigned k = 0; k < nz; k++)
        for (unsigned j = 0; j < ny; j++)
            for (unsigned i = 0; i < nx; i++) {
                const size_t o = i + static_cast<size_t>(nx) * j + xy * k;
                double v = 0.0;
                if (ic == 1) {
                    v = (i > 0.4 * nx && i < 0.6 * nx && j > 0.4 * ny && j < 0.6 * ny && k > 0.4 * nz &&
                         k < 0.6 * nz)
                            ? 1.0
                            : 0.0;
                } else if (ic == 3) {
                    const double x = i * dx, y = j * dy, z = k * dz;
                    v = 1.0 * std::exp(-((x - 1.0) * (x - 1.0) + (y - 1.0) * (y - 1.0) + (z - 1.0) * (z - 1.0)) / 0.1);
                }
                u0[o] = v;
            }
    return MGAD_OK;
}

namespace {

// EXP_DISTRIBUTION of SingleGPU/Diffusion3d_baselineCode/acoustics3d.h:29:
// the heat kernel centred at (5, 5, 5) at time t0 with diffusivity d.
double exp_distribution(unsigned i, unsigned j, unsigned k, double dx, double dy, double dz, double d, double t0) {
    const double x = -5 + i * dx, y = -5 + j * dy, z = -5 + k * dz;
    return std::exp(-(x * x + y * y + z * z) / (4 * d * t0));
}

}  // namespace

// Call_Init3D, SingleGPU/Diffusion3d_baselineCode/tools.cpp:92-143:
//   ic 1: 0, with the two bottom and two top planes (k = 0, 1, nz-2, nz-1) at 1;
//   ic 2: sin(pi i dx) sin(pi j dy) sin(pi k dz), boundary nodes 0;
//   ic 3: exp(-r^2/(4*1.0*0.1)) centred at (5, 5, 5), boundary nodes 0.
extern "C" int mgad_init3d_heat(int ic, double *u0, double dx, double dy, double dz, unsigned int nx,
                                unsigned int ny, unsigned int nz) {
    if (!u0 || !dims_ok(nx, ny, nz) || ic < 1 || ic > 3) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    const size_t xy = static_cast<size_t>(nx) * ny;
    for (unsigned k = 0; k < nz; k++)
        for (unsigned j = 0; j < ny; j++)
            for (unsigned i = 0; i < nx; i++) {
                const size_t o = i + static_cast<size_t>(nx) * j + xy * k;
                const bool boundary = i == 0 || i == nx - 1 || j == 0 || j == ny - 1 || k == 0 || k == nz - 1;
                double v;
                if (ic == 1) {
                    v = (k == 0 || k == 1 || k == nz - 2 || k == nz - 1) ? 1.0 : 0.0;
                } else if (ic == 2) {
                    v = boundary ? 0.0 : std::sin(M_PI * i * dx) * std::sin(M_PI * j * dy) * std::sin(M_PI * k * dz);
                } else {
                    v = boundary ? 0.0 : exp_distribution(i, j, k, dx, dy, dz, 1.0, 0.1);
                }
                u0[o] = v;
            }
    return MGAD_OK;
}

// CalcError, SingleGPU/Diffusion3d_baselineCode/tools.cpp:156-180, with the
// node index corrected (the reference compares every node against u[i]):
// err = (0.1/t)^(3/2) exp(-r^2/(4 t)) - u at node (i, j, k);
// out = {dx dy dz sum|err|, sqrt(dx dy dz sum err^2), max|err|}.
extern "C" int mgad_heat_error_norms(const double *u, double t, double dx, double dy, double dz, unsigned int nx,
                                     unsigned int ny, unsigned int nz, double *out) {
    if (!u || !out || !dims_ok(nx, ny, nz) || !(t > 0.0)) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    const size_t xy = static_cast<size_t>(nx) * ny;
    double l1 = 0.0, l2 = 0.0, linf = 0.0;
    for (unsigned k = 0; k < nz; k++)
        for (unsigned j = 0; j < ny; j++)
            for (unsigned i = 0; i < nx; i++) {
                const double err = (std::sqrt(0.1 / t) * (0.1 / t) * exp_distribution(i, j, k, dx, dy, dz, 1.0, t)) -
                                   u[i + static_cast<size_t>(nx) * j + xy * k];
                l1 += std::fabs(err);
                l2 += err * err;
                linf = std::fmax(linf, std::fabs(err));
            }
    out[0] = dx * dy * dz * l1;
    out[1] = std::sqrt(dx * dy * dz * l2);
    out[2] = linf;
    return MGAD_OK;
}

// Init_subdomain and Merge_domains, MultiGPU/Burgers3d_Baseline/Tools.c:181-223
// (the same in Diffusion3d_Baseline/Tools.c:180-222): subdomain n of a field
// split along z into slabs of _Nz own planes with 3 ghost planes each side.
// The global field is Nx x Ny x (P _Nz + 6); subdomain n holds its planes
// n _Nz .. n _Nz + _Nz + 5.  Init copies all _Nz + 6 planes out, Merge copies
// the _Nz own planes (3 .. _Nz + 2) back.  Each plane is a contiguous run of
// Nx Ny values, so both are one memcpy per slab.
extern "C" int mgad_init_subdomain(const double *h_q, double *h_s_q, unsigned int n, unsigned int Nx,
                                   unsigned int Ny, unsigned int _Nz) {
    if (!h_q || !h_s_q || _Nz == 0 || _Nz > (1u << 30) || !dims_ok(Nx, Ny, _Nz + 6)) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    const size_t xy = static_cast<size_t>(Nx) * Ny;
    std::memcpy(h_s_q, h_q + xy * (static_cast<size_t>(n) * _Nz), sizeof(double) * xy * (_Nz + 6));
    return MGAD_OK;
}

extern "C" int mgad_merge_domains(const double *h_s_q, double *h_q, unsigned int n, unsigned int Nx,
                                  unsigned int Ny, unsigned int _Nz) {
    if (!h_q || !h_s_q || _Nz == 0 || _Nz > (1u << 30) || !dims_ok(Nx, Ny, _Nz + 6)) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    const size_t xy = static_cast<size_t>(Nx) * Ny;
    std::memcpy(h_q + xy * (static_cast<size_t>(n) * _Nz + 3), h_s_q + xy * 3, sizeof(double) * xy * _Nz);
    return MGAD_OK;
}
