// phantom.cpp -- seeded synthetic tilt series: analytic parallel-beam line
// integrals of random ellipsoids, tilted about y, optionally mapped to raw
// images through the inverse of the stack-correction pre-stage
// (Correct/GCorrPatchShift.cu:74-89: raw(q) = ideal(R(axis)^-1 (q - s - c))),
// plus Gaussian noise.  Host C++, deterministic for a given seed whatever the
// thread count (every random draw is a counter-based splitmix64 value).
//
// Geometry matches the reconstruction: detector index iu <-> u = iu - ProjX/2,
// row iv <-> y = iv - ProjY/2, voxel (x, z) <-> (x - VolX/2, z - VolZ/2), and
// at tilt theta the ray through u runs along (-sin, 0, cos) from
// u * (cos, 0, sin).
#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include "atr_internal.h"

namespace {

struct Ellipsoid {
    double c[3];
    double m[3][3];  // R diag(1/a^2) R^T
    double rho;
    double r;        // bounding radius
};

double uni(uint64_t seed, uint64_t stream, uint64_t i)
{
    uint64_t r = atr::splitmix64(seed ^ (stream * 0x9E3779B97F4A7C15ull) ^
                                 atr::splitmix64(i + 0x632BE59BD9B4E019ull));
    return (double)(r >> 11) * (1.0 / 9007199254740992.0);  // [0, 1)
}

double gauss(uint64_t seed, uint64_t stream, uint64_t i)
{
    double u1 = uni(seed, stream, 2 * i), u2 = uni(seed, stream, 2 * i + 1);
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

double chord(const Ellipsoid& e, const double p0[3], const double d[3])
{
    double q[3] = {p0[0] - e.c[0], p0[1] - e.c[1], p0[2] - e.c[2]};
    double md[3], mq[3];
    for (int i = 0; i < 3; i++) {
        md[i] = e.m[i][0] * d[0] + e.m[i][1] * d[1] + e.m[i][2] * d[2];
        mq[i] = e.m[i][0] * q[0] + e.m[i][1] * q[1] + e.m[i][2] * q[2];
    }
    double a = d[0] * md[0] + d[1] * md[1] + d[2] * md[2];
    double b = d[0] * mq[0] + d[1] * mq[1] + d[2] * mq[2];
    double c = q[0] * mq[0] + q[1] * mq[1] + q[2] * mq[2] - 1.0;
    double disc = b * b - a * c;
    if (disc <= 0.0 || a <= 0.0) return 0.0;
    return 2.0 * std::sqrt(disc) / a;
}

template <class F>
void parallel_rows(int rows, int n_threads, F fn)
{
    int nt = std::max(1, n_threads);
    std::vector<std::thread> th;
    for (int w = 0; w < nt; w++)
        th.emplace_back([=]() {
            for (int r = w; r < rows; r += nt) fn(r);
        });
    for (auto& x : th) x.join();
}

}  // namespace

extern "C" atr_status atr_phantom_tilt_series(uint64_t seed, int proj_x, int proj_y, int n_tilts,
                                              const float* tilt_deg, int vol_z, int n_ell,
                                              float axis_min, float axis_max, float tilt_axis_deg,
                                              float max_shift, float noise_frac,
                                              float* shift_xy_out, float* proj_out, int n_threads)
{
    if (proj_x < 2 || proj_y < 1 || n_tilts < 1 || vol_z < 1 || n_ell < 0 || !tilt_deg ||
        !proj_out || axis_min <= 0 || axis_max < axis_min)
        return atr::fail(ATR_ERR_INVALID_ARG, "atr_phantom_tilt_series: bad argument");
    if (n_threads <= 0) n_threads = (int)std::max(1u, std::thread::hardware_concurrency());
    const bool raw_map = std::isfinite(tilt_axis_deg);
    int aln_x = proj_x, aln_y = proj_y;
    if (raw_map) atr::aligned_size(proj_x, proj_y, tilt_axis_deg, &aln_x, &aln_y);
    const double vx = aln_x / 2 * 2, vy = aln_y, vz = vol_z;

    std::vector<Ellipsoid> ell(n_ell);
    for (int k = 0; k < n_ell; k++) {
        Ellipsoid& e = ell[k];
        e.c[0] = (uni(seed, 1, 8 * k + 0) - 0.5) * vx;
        e.c[1] = (uni(seed, 1, 8 * k + 1) - 0.5) * vy;
        e.c[2] = (uni(seed, 1, 8 * k + 2) - 0.5) * vz;
        double ax[3];
        for (int i = 0; i < 3; i++) ax[i] = axis_min + (axis_max - axis_min) * uni(seed, 2, 3 * k + i);
        e.rho = 0.5 + 0.5 * uni(seed, 1, 8 * k + 3);
        double q[4], nrm = 0;
        for (int i = 0; i < 4; i++) {
            q[i] = gauss(seed, 3, 4 * k + i);
            nrm += q[i] * q[i];
        }
        nrm = std::sqrt(std::max(nrm, 1e-30));
        double w = q[0] / nrm, x = q[1] / nrm, y = q[2] / nrm, z = q[3] / nrm;
        double R[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)},
                          {2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)},
                          {2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)}};
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double s = 0;
                for (int l = 0; l < 3; l++) s += R[i][l] * R[j][l] / (ax[l] * ax[l]);
                e.m[i][j] = s;
            }
        e.r = std::max(ax[0], std::max(ax[1], ax[2]));
    }
    // sort by y so that a detector row only visits ellipsoids within reach
    std::sort(ell.begin(), ell.end(), [](const Ellipsoid& a, const Ellipsoid& b) { return a.c[1] < b.c[1]; });
    std::vector<double> ys(n_ell);
    for (int k = 0; k < n_ell; k++) ys[k] = ell[k].c[1];
    const double rmax = axis_max;

    std::vector<float> shifts(2 * (size_t)n_tilts, 0.0f);
    for (int t = 0; t < n_tilts; t++)
        for (int i = 0; i < 2; i++)
            shifts[2 * t + i] = raw_map ? (float)((2.0 * uni(seed, 4, 2 * t + i) - 1.0) * max_shift) : 0.0f;
    if (shift_xy_out)
        for (int i = 0; i < 2 * n_tilts; i++) shift_xy_out[i] = shifts[i];

    const float d2r = 3.141592654f / 180.0f;
    const float ang = raw_map ? tilt_axis_deg * d2r : 0.0f;
    const double ca = std::cos((double)ang), sa = std::sin((double)ang);
    const size_t img = (size_t)proj_x * proj_y;

    parallel_rows(n_tilts * proj_y, n_threads, [&](int row) {
        int t = row / proj_y, qy = row % proj_y;
        double th = (double)tilt_deg[t] * 0.017453292519943295;
        double ct = std::cos(th), st = std::sin(th);
        double d[3] = {-st, 0.0, ct};
        float* out = proj_out + t * img + (size_t)qy * proj_x;
        for (int qx = 0; qx < proj_x; qx++) {
            double u, v;
            if (raw_map) {
                double X = qx - (double)shifts[2 * t] - proj_x * 0.5;
                double Y = qy - (double)shifts[2 * t + 1] - proj_y * 0.5;
                double xc = X * ca + Y * sa, yc = -X * sa + Y * ca;
                u = xc - 0.5;
                v = yc - 0.5;
            } else {
                u = qx - aln_x / 2 * 2 * 0.5;
                v = qy - aln_y * 0.5;
            }
            double p0[3] = {u * ct, v, u * st};
            auto lo = std::lower_bound(ys.begin(), ys.end(), v - rmax) - ys.begin();
            double sum = 0.0;
            for (size_t k = lo; k < ell.size() && ys[k] <= v + rmax; k++) {
                const Ellipsoid& e = ell[k];
                if (std::fabs(v - e.c[1]) > e.r) continue;
                double uc = e.c[0] * ct + e.c[2] * st;
                if (std::fabs(u - uc) > e.r) continue;
                sum += e.rho * chord(e, p0, d);
            }
            out[qx] = (float)sum;
        }
    });
    if (noise_frac > 0) {
        double total = 0.0;
        for (size_t i = 0; i < img * n_tilts; i++) total += proj_out[i];
        double sigma = noise_frac * total / (double)(img * n_tilts);
        parallel_rows(n_tilts * proj_y, n_threads, [&](int row) {
            float* out = proj_out + (size_t)row * proj_x;
            for (int qx = 0; qx < proj_x; qx++)
                out[qx] += (float)(sigma * gauss(seed, 5, (uint64_t)row * proj_x + qx));
        });
    }
    return ATR_OK;
}
