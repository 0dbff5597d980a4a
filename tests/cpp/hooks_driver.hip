// hooks_driver.hip -- field callbacks through the C++ API, both kinds:
// the reference's float2 callbacks (host hooks: the library hands over
// float2 copies and synchronises) and the device-side callbacks (the
// library's own buffers on its stream).  The system is the inhomogeneous
// diffusion of the reference's example 07 (examples/07_inhomogeneous_
// diffusion/07_inhomogeneous_diffusion.cu:54-79): mirror ("even")
// boundary conditions imposed after every update by a real-space callback
// on v and on the coefficient field x.  A second system low-pass filters
// its field in Fourier space.  Each pair of runs must agree: bit for bit
// for the real-space callbacks (same transforms, same values), within
// 1e-6 for the Fourier ones (the float2 path projects the full spectrum
// back onto the half spectrum).  Both low-pass runs start from the same
// initial field, drawn once from a fixed-seed generator and written through
// real_array; it and the device run's final field are left in the working
// directory (lowpass_ic.f32, lowpass_phi.f32: raw float32, x fastest) so
// the caller can check the filtered run against the oracle.
// Usage: hooks_driver   (exit status = number of failed checks)
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../include/cupss.h"

namespace {

const int N = 32;  // the box is 2N x 2N, mirrored about x = N and y = N
const int STEPS = 200;
const int CUT = 10;  // Fourier filter: keep |n_x|, |n_y| <= CUT

// mirror image of cell (i, j) in the quadrant i <= sx/2, j <= sy/2
__device__ __forceinline__ int mirror_index(int i, int j, int sx, int sy)
{
    const int mi = i <= sx / 2 ? i : sx - i;
    const int mj = j <= sy / 2 ? j : sy - j;
    return mj * sx + mi;
}

__global__ void k_even_pair(float2 *f, int sx, int sy)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= sx || j >= sy) return;
    const int src = mirror_index(i, j, sx, sy);
    if (src != j * sx + i) f[j * sx + i].x = f[src].x;  // sources are never written
}

__global__ void k_even_real(float *f, int sx, int sy)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= sx || j >= sy) return;
    const int src = mirror_index(i, j, sx, sy);
    if (src != j * sx + i) f[j * sx + i] = f[src];
}

void even_pair(evolver *, float2 *f, int sx, int sy, int)
{
    hipLaunchKernelGGL(k_even_pair, dim3((sx + 63) / 64, sy), dim3(64), 0, 0, f, sx, sy);
}

void even_real(evolver *, float *f, int sx, int sy, int, int, hipStream_t s)
{
    hipLaunchKernelGGL(k_even_real, dim3((sx + 63) / 64, sy), dim3(64), 0, s, f, sx, sy);
}

__device__ __forceinline__ int signed_mode(int k, int s) { return k <= s / 2 ? k : k - s; }

// full C2C spectrum float2[sy][sx]
__global__ void k_lowpass_full(float2 *f, int sx, int sy)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= sx || j >= sy) return;
    if (abs(signed_mode(i, sx)) > CUT || abs(signed_mode(j, sy)) > CUT) f[j * sx + i] = make_float2(0.f, 0.f);
}

// half spectrum float2[sy][sx/2+1]
__global__ void k_lowpass_half(float2 *f, int sx, int sy)
{
    const int nc1 = sx / 2 + 1;
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= nc1 || j >= sy) return;
    if (i > CUT || abs(signed_mode(j, sy)) > CUT) f[j * nc1 + i] = make_float2(0.f, 0.f);
}

void lowpass_full(evolver *, float2 *f, int sx, int sy, int)
{
    hipLaunchKernelGGL(k_lowpass_full, dim3((sx + 63) / 64, sy), dim3(64), 0, 0, f, sx, sy);
}

void lowpass_half(evolver *, float2 *f, int sx, int sy, int, hipStream_t s)
{
    hipLaunchKernelGGL(k_lowpass_half, dim3((sx / 2 + 1 + 63) / 64, sy), dim3(64), 0, s, f, sx, sy);
}

std::vector<float> real_of(evolver &e, const std::string &name)
{
    e.copyAllDataToHost();
    const field *f = e.fieldsMap[name];
    const int n = e.getSystemSizeX() * e.getSystemSizeY();
    std::vector<float> out(n);
    for (int i = 0; i < n; i++) out[i] = f->real_array[i].x;
    return out;
}

// example 07's system; device = true uses the device-side callbacks
std::vector<float> mirrored(bool device, std::vector<float> *xout)
{
    const float L = 1.0f, dx = L / (float)N, dt = 0.001f * dx;
    evolver sys(RUN_GPU, 2 * N, 2 * N, dx, dx, dt, STEPS + 1);
    sys.createField("v", true);
    sys.createField("iqxv", false);
    sys.initializeDroplet("v", 1.0, 0.0, 8.0, 2.0, N / 2, N / 2, 0);
    sys.createField("x", false);
    for (int j = 0; j <= N; j++)
        for (int i = 0; i <= N; i++)
            sys.fieldsMap["x"]->real_array[j * 2 * N + i].x =
                1.0f + std::tanh((i - 0.9f * (N / 2 - 1)) / (0.1f * (N / 2 - 1)));
    for (const char *name : {"v", "x"}) {
        field *f = sys.fieldsMap[name];
        if (device) {
            f->hasCBDevice = true;
            f->callbackDevice = even_real;
        } else {
            f->hasCB = true;
            f->callback = even_pair;
        }
    }
    sys.addEquation("dt v + 0.5*q^2*v = iqx*x*iqxv + x*iqy^2*v");
    sys.addEquation("iqxv = iqx*v");
    sys.prepareProblem();
    sys.advanceTime();  // first step (plans, tables) outside the timing
    hipDeviceSynchronize();
    const auto t0 = std::chrono::steady_clock::now();
    for (int t = 1; t < STEPS; t++) sys.advanceTime();
    hipDeviceSynchronize();
    const double us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
    std::printf("%-34s %.1f us per advanceTime\n", device ? "mirror, device-side callbacks" : "mirror, float2 callbacks",
                us / (STEPS - 1));
    *xout = real_of(sys, "x");
    return real_of(sys, "v");
}

// i.i.d. uniform in [-0.1, 0.1] from a fixed seed (the library's
// initializeUniformNoise reseeds rand() from the clock, as the reference's
// does, so it cannot give two runs the same start)
std::vector<float> lowpass_ic()
{
    std::mt19937 gen(7);
    std::uniform_real_distribution<float> u(-0.1f, 0.1f);
    std::vector<float> x(4 * N * N);
    for (float &v : x) v = u(gen);
    return x;
}

void save(const char *path, const std::vector<float> &x)
{
    FILE *fh = std::fopen(path, "wb");
    if (!fh || std::fwrite(x.data(), sizeof(float), x.size(), fh) != x.size()) {
        std::printf("cannot write %s\n", path);
        std::exit(99);
    }
    std::fclose(fh);
}

// Cahn-Hilliard whose field is low-pass filtered after every update
std::vector<float> filtered(bool device, const std::vector<float> &ic)
{
    evolver sys(RUN_GPU, 2 * N, 2 * N, 1.0f, 1.0f, 0.01f, STEPS + 1);
    sys.createField("phi", true);
    sys.addParameter("a", -1.0f);
    sys.addParameter("b", 1.0f);
    sys.addParameter("k", 1.0f);
    sys.addEquation("dt phi + (a*q^2 + k*q^4)*phi = -b*q^2*phi^3");
    field *f = sys.fieldsMap["phi"];
    for (size_t i = 0; i < ic.size(); i++) f->real_array[i] = make_float2(ic[i], 0.0f);
    if (device) {
        f->hasCBFourierDevice = true;
        f->callbackFourierDevice = lowpass_half;
    } else {
        f->hasCBFourier = true;
        f->callbackFourier = lowpass_full;
    }
    sys.prepareProblem();
    for (int t = 0; t < STEPS; t++) sys.advanceTime();
    return real_of(sys, "phi");
}

int compare(const char *what, const std::vector<float> &a, const std::vector<float> &b, double tol)
{
    double num = 0, den = 0, maxd = 0;
    for (size_t i = 0; i < a.size(); i++) {
        const double d = (double)a[i] - (double)b[i];
        num += d * d;
        den += (double)a[i] * a[i];
        maxd = std::fmax(maxd, std::fabs(d));
    }
    const double rel = den > 0 ? std::sqrt(num / den) : std::sqrt(num);
    const bool ok = tol == 0 ? maxd == 0 : rel <= tol;
    std::printf("%-34s max|diff| %.3e  rel L2 %.3e  %s\n", what, maxd, rel, ok ? "ok" : "FAIL");
    return ok ? 0 : 1;
}

}  // namespace

int main()
{
    int bad = 0;
    std::vector<float> xa, xb;
    const std::vector<float> va = mirrored(false, &xa);
    const std::vector<float> vb = mirrored(true, &xb);
    bad += compare("mirror callbacks, v", va, vb, 0);
    bad += compare("mirror callbacks, x", xa, xb, 0);
    // the mirror condition holds after the last step
    double asym = 0;
    for (int j = 0; j < 2 * N; j++)
        for (int i = N + 1; i < 2 * N; i++) asym = std::fmax(asym, std::fabs(vb[j * 2 * N + i] - vb[j * 2 * N + 2 * N - i]));
    std::printf("%-34s %.3e\n", "mirror residual of v", asym);
    if (asym > 1e-5) bad++;
    const std::vector<float> ic = lowpass_ic();
    const std::vector<float> pa = filtered(false, ic);
    const std::vector<float> pb = filtered(true, ic);
    save("lowpass_ic.f32", ic);
    save("lowpass_phi.f32", pb);
    bad += compare("Fourier low-pass callbacks, phi", pa, pb, 1e-6);
    return bad;
}
