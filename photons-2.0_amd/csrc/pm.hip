// pm.hip -- the PM long-range force on one GPU (one rank's whole periodic box).
//
// The reference's partmesh() with its convolution (src/partmesh.c:800-1550,
// src/conv.f90:171-229) on one rank, in FP64:
//   1. CIC deposit, mesh points at cell centres (src/partmesh.c:877-944), into the
//      periodic N^3 mesh, then a scale by (N/L)^3 (:948-959);
//   2. hipFFT D2Z, multiply by the Green's function with sinc^-4 deconvolution and
//      the Gaussian long-range filter, pref = L^2/(pi N^3), gf(0) = pref, and pi as
//      the reference truncates it (src/conv.f90:142), hipFFT Z2D (unnormalised, as
//      2DECOMP&FFT's transforms);
//   3. the 4th-order central-difference gradient at every mesh point (the values the
//      reference forms at each particle's 8 CIC points, src/partmesh.c:1328-1550),
//      then one thread per particle interpolates them with the CIC weights in the
//      reference's term order.
// acc_pm is written in the caller's order, without G (the kick applies it).
// Clustered particles pile onto few cells, so the deposit first sorts the particles
// by the lower corner of their CIC cloud: a wavefront then sums the contributions of
// each run of equal corners (segmented scan over shuffles) and one lane per run adds
// the eight sums atomically.  The interpolation walks the same order, so neighbouring
// lanes read neighbouring cells.  The atomic adds make the density's last bits depend
// on their order.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cmath>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include <stdexcept>
#include <string>

#include "../../include/photons_fmm.h"

namespace {

#define PM_CHECK(x)                                                                   \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)
#define FFT_CHECK(x)                                                                      \
    do {                                                                                  \
        hipfftResult r_ = (x);                                                            \
        if (r_ != HIPFFT_SUCCESS) throw std::runtime_error(std::string(#x) + " failed (" + std::to_string((int)r_) + ")"); \
    } while (0)

constexpr double kPiTrunc = 3.1415926;  // src/conv.f90:142

thread_local std::string g_err;

int fail(const std::string &m)
{
    g_err = m;
    return -1;
}

struct Cic {
    int i, ii, j, jj, k, kk;
    double wi, win, wj, wjn, wk, wkn;
};

// src/partmesh.c:881-914, indices wrapped into [0, N)
__device__ __forceinline__ Cic cic(const double *p, int N, double norm, double delta)
{
    Cic c;
    int idx[3], nb[3];
    double w[3];
    for (int a = 0; a < 3; a++) {
        idx[a] = (int)(p[a] * norm);
        double wa = (p[a] - (idx[a] + 0.5) * delta) * norm;
        if (wa > 0) {
            nb[a] = idx[a] + 1;
        } else {
            wa = -wa;
            nb[a] = idx[a] - 1;
        }
        w[a] = wa;
        idx[a] = (idx[a] % N + N) % N;
        nb[a] = (nb[a] % N + N) % N;
    }
    c.i = idx[0], c.ii = nb[0], c.j = idx[1], c.jj = nb[1], c.k = idx[2], c.kk = nb[2];
    c.wi = w[0], c.wj = w[1], c.wk = w[2];
    c.win = 1.0 - w[0], c.wjn = 1.0 - w[1], c.wkn = 1.0 - w[2];
    return c;
}

__device__ __forceinline__ size_t at(int i, int j, int k, int N) { return ((size_t)i * N + j) * N + k; }

__global__ void k_fill(double *p, size_t n, double v)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

// lower corner of each particle's CIC cloud (x slowest), with the particle index
__global__ void k_corner_keys(int64_t n, const double *__restrict__ pos, int N, double norm, double delta,
                              uint32_t *__restrict__ key, int32_t *__restrict__ idx)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const Cic c = cic(pos + 3 * t, N, norm, delta);
    // the cloud covers {lo, lo+1} per axis, lo = i or the neighbour below (wrapped)
    const int lx = c.ii == (c.i + 1) % N ? c.i : c.ii;
    const int ly = c.jj == (c.j + 1) % N ? c.j : c.jj;
    const int lz = c.kk == (c.k + 1) % N ? c.k : c.kk;
    key[t] = (uint32_t)at(lx, ly, lz, N);
    idx[t] = (int32_t)t;
}

// particles in corner order: each wavefront sums runs of equal corners and the run's
// last lane adds the eight sums (corner terms formed as the reference forms them,
// mass * wx * wy * wz)
__global__ void __launch_bounds__(256) k_deposit_runs(int64_t n, const double *__restrict__ pos,
                                                      const uint32_t *__restrict__ key,
                                                      const int32_t *__restrict__ idx, double mass, int N,
                                                      double norm, double delta, double *__restrict__ mesh)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x % 64;
    const bool live = t < n;
    const uint32_t k = live ? key[t] : 0xffffffffu;
    double v[8];
    int cx = 0, cy = 0, cz = 0;
    if (live) {
        const Cic c = cic(pos + 3 * (int64_t)idx[t], N, norm, delta);
        // weights at lo and lo + 1 per axis
        const bool ux = c.ii == (c.i + 1) % N, uy = c.jj == (c.j + 1) % N, uz = c.kk == (c.k + 1) % N;
        const double x0 = ux ? c.win : c.wi, x1 = ux ? c.wi : c.win;
        const double y0 = uy ? c.wjn : c.wj, y1 = uy ? c.wj : c.wjn;
        const double z0 = uz ? c.wkn : c.wk, z1 = uz ? c.wk : c.wkn;
        cx = ux ? c.i : c.ii;
        cy = uy ? c.j : c.jj;
        cz = uz ? c.k : c.kk;
        v[0] = mass * x0 * y0 * z0;
        v[1] = mass * x1 * y0 * z0;
        v[2] = mass * x0 * y1 * z0;
        v[3] = mass * x1 * y1 * z0;
        v[4] = mass * x0 * y0 * z1;
        v[5] = mass * x1 * y0 * z1;
        v[6] = mass * x0 * y1 * z1;
        v[7] = mass * x1 * y1 * z1;
    } else {
        for (int q = 0; q < 8; q++) v[q] = 0.0;
    }
    // segmented inclusive scan: keys are sorted, so equal keys l-d and l mean the
    // whole stretch between them is one run
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t kd = __shfl_up(k, d, 64);
        const bool same = lane >= d && kd == k;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const double u = __shfl_up(v[q], d, 64);
            if (same) v[q] += u;
        }
    }
    const uint32_t knext = __shfl_down(k, 1, 64);
    const bool tail = live && (lane == 63 || knext != k);
    if (tail) {
        const int x1 = (cx + 1) % N, y1 = (cy + 1) % N, z1 = (cz + 1) % N;
        atomicAdd(&mesh[at(cx, cy, cz, N)], v[0]);
        atomicAdd(&mesh[at(x1, cy, cz, N)], v[1]);
        atomicAdd(&mesh[at(cx, y1, cz, N)], v[2]);
        atomicAdd(&mesh[at(x1, y1, cz, N)], v[3]);
        atomicAdd(&mesh[at(cx, cy, z1, N)], v[4]);
        atomicAdd(&mesh[at(x1, cy, z1, N)], v[5]);
        atomicAdd(&mesh[at(cx, y1, z1, N)], v[6]);
        atomicAdd(&mesh[at(x1, y1, z1, N)], v[7]);
    }
}

__global__ void k_scale(double *p, size_t n, double s)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] *= s;
}

__device__ __forceinline__ int wave(int a, int N) { return a > N / 2 ? a - N : a; }

__device__ __forceinline__ double sinc_pm(int v, int N)
{
    if (v == 0) return 1.0;
    const double x = kPiTrunc * v / N;
    return sin(x) / x;
}

// the half spectrum of a D2Z transform of a C-order N^3 array: [N][N][N/2+1]
__global__ void k_green(hipfftDoubleComplex *spec, int N, double ismth2, double pref)
{
    const int nc = N / 2 + 1;
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)N * N * nc) return;
    const int c = (int)(t % nc), b = (int)((t / nc) % N), a = (int)(t / ((size_t)nc * N));
    const int l = wave(a, N), m = wave(b, N), n = wave(c, N);
    double gf;
    if (l == 0 && m == 0 && n == 0) {
        gf = pref;
    } else {
        const double k2 = (double)(n * n + m * m + l * l);
        const double ff = 1.0 / (sinc_pm(l, N) * sinc_pm(m, N) * sinc_pm(n, N));
        gf = pref * exp(-k2 * ismth2) * ff * ff * ff * ff / k2;
    }
    spec[t].x *= gf;
    spec[t].y *= gf;
}

// 4th-order central differences of the potential along the three axes
__global__ void k_gradient(const double *__restrict__ phi, int N, double invx, double *__restrict__ gx,
                           double *__restrict__ gy, double *__restrict__ gz)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)N * N * N) return;
    const int k = (int)(t % N), j = (int)((t / N) % N), i = (int)(t / ((size_t)N * N));
    const double f1 = 4.0 / 3.0, f2 = 1.0 / 6.0;
    auto w = [N](int a) { return (a % N + N) % N; };
    double d;
    d = f1 * invx * (phi[at(w(i + 1), j, k, N)] - phi[at(w(i - 1), j, k, N)]);
    d -= f2 * invx * (phi[at(w(i + 2), j, k, N)] - phi[at(w(i - 2), j, k, N)]);
    gx[t] = d;
    d = f1 * invx * (phi[at(i, w(j + 1), k, N)] - phi[at(i, w(j - 1), k, N)]);
    d -= f2 * invx * (phi[at(i, w(j + 2), k, N)] - phi[at(i, w(j - 2), k, N)]);
    gy[t] = d;
    d = f1 * invx * (phi[at(i, j, w(k + 1), N)] - phi[at(i, j, w(k - 1), N)]);
    d -= f2 * invx * (phi[at(i, j, w(k + 2), N)] - phi[at(i, j, w(k - 2), N)]);
    gz[t] = d;
}

__device__ __forceinline__ double interp(const double *g, const Cic &c, int N)
{
    return c.win * c.wjn * c.wkn * g[at(c.i, c.j, c.k, N)] + c.wi * c.wjn * c.wkn * g[at(c.ii, c.j, c.k, N)] +
           c.win * c.wj * c.wkn * g[at(c.i, c.jj, c.k, N)] + c.wi * c.wj * c.wkn * g[at(c.ii, c.jj, c.k, N)] +
           c.win * c.wjn * c.wk * g[at(c.i, c.j, c.kk, N)] + c.wi * c.wjn * c.wk * g[at(c.ii, c.j, c.kk, N)] +
           c.win * c.wj * c.wk * g[at(c.i, c.jj, c.kk, N)] + c.wi * c.wj * c.wk * g[at(c.ii, c.jj, c.kk, N)];
}

// particles visited in corner order (idx), results written in the caller's order
__global__ void k_interpolate(int64_t n, const double *__restrict__ pos, const int32_t *__restrict__ idx, int N,
                              double norm, double delta, const double *__restrict__ gx,
                              const double *__restrict__ gy, const double *__restrict__ gz,
                              double *__restrict__ acc)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t p = idx[t];
    const Cic c = cic(pos + 3 * p, N, norm, delta);
    acc[3 * p] = interp(gx, c, N);
    acc[3 * p + 1] = interp(gy, c, N);
    acc[3 * p + 2] = interp(gz, c, N);
}

inline int blocks(size_t n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

struct photons_pm {
    int N = 0, device = 0;
    double L = 0, rs = 0, mass = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    hipfftHandle fwd = 0, inv = 0;
    double *mesh = nullptr, *grad = nullptr, *io_pos = nullptr, *io_acc = nullptr;
    hipfftDoubleComplex *spec = nullptr;
    int64_t io_cap = 0;
    // corner sort: keys and particle indices (two halves each), rocPRIM temporary
    uint32_t *key = nullptr;
    int32_t *idx = nullptr;
    void *sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    int64_t sort_cap = 0;
};

extern "C" {

const char *photons_pm_last_error(void) { return g_err.c_str(); }

int photons_pm_create(photons_pm **out, int nside, double boxsize, double rs, double mass, int device, void *stream)
{
    photons_pm *p = nullptr;
    try {
        if (!out) return fail("null handle");
        // the CIC corner keys (k_corner_keys) are 32-bit mesh indices: N^3 < 2^32
        if (nside < 8 || nside > 1624 || nside % 2)
            return fail("nside must be even, in [8, 1624] (32-bit mesh indices)");
        if (!(boxsize > 0.0) || !(rs > 0.0)) return fail("boxsize and rs must be positive");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
            return fail("no usable GPU for this context");
        PM_CHECK(hipSetDevice(device));
        p = new photons_pm();
        p->N = nside;
        p->L = boxsize;
        p->rs = rs;
        p->mass = mass;
        p->device = device;
        if (stream) {
            p->stream = reinterpret_cast<hipStream_t>(stream);
        } else {
            PM_CHECK(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
            p->own_stream = true;
        }
        const size_t n3 = (size_t)nside * nside * nside, nc = (size_t)nside * nside * (nside / 2 + 1);
        PM_CHECK(hipMalloc(&p->mesh, sizeof(double) * n3));
        PM_CHECK(hipMalloc(&p->grad, sizeof(double) * 3 * n3));
        PM_CHECK(hipMalloc(&p->spec, sizeof(hipfftDoubleComplex) * nc));
        FFT_CHECK(hipfftPlan3d(&p->fwd, nside, nside, nside, HIPFFT_D2Z));
        FFT_CHECK(hipfftPlan3d(&p->inv, nside, nside, nside, HIPFFT_Z2D));
        FFT_CHECK(hipfftSetStream(p->fwd, p->stream));
        FFT_CHECK(hipfftSetStream(p->inv, p->stream));
        *out = p;
        return 0;
    } catch (const std::exception &e) {
        if (p) photons_pm_destroy(p);
        return fail(std::string("photons_pm_create: ") + e.what());
    }
}

int photons_pm_destroy(photons_pm *p)
{
    if (!p) return 0;
    (void)hipSetDevice(p->device);
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    if (p->fwd) hipfftDestroy(p->fwd);
    if (p->inv) hipfftDestroy(p->inv);
    for (void *b : {(void *)p->mesh, (void *)p->grad, (void *)p->spec, (void *)p->io_pos, (void *)p->io_acc,
                    (void *)p->key, (void *)p->idx, p->sort_tmp})
        if (b) (void)hipFree(b);
    if (p->own_stream && p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
    return 0;
}

// the box length of the next evaluations (the mesh shape stays): one context serves
// boxes of one mesh size and different lengths in turn
int photons_pm_set_boxsize(photons_pm *p, double boxsize)
{
    if (!p) return fail("null handle");
    if (!(boxsize > 0.0)) return fail("boxsize must be positive");
    p->L = boxsize;
    return 0;
}

// positions [n][3] in [0, L) on the device -> acc_pm [n][3], caller order, queued on
// the context's stream
int photons_pm_force_device(photons_pm *p, int64_t n, const double *d_pos, double *d_acc)
{
    try {
        if (n < 0) return fail("negative particle count");
        if (n > INT32_MAX) return fail("more than 2^31 particles");
        PM_CHECK(hipSetDevice(p->device));
        if (n == 0) return 0;
        const int N = p->N;
        const size_t n3 = (size_t)N * N * N, nc = (size_t)N * N * (N / 2 + 1);
        const double norm = N / p->L, delta = 1.0 / norm;
        hipStream_t st = p->stream;
        if (n > p->sort_cap) {
            for (void *b : {(void *)p->key, (void *)p->idx}) if (b) PM_CHECK(hipFree(b));
            p->key = nullptr;
            p->idx = nullptr;
            p->sort_cap = 0;
            PM_CHECK(hipMalloc(&p->key, sizeof(uint32_t) * 2 * (size_t)n));
            PM_CHECK(hipMalloc(&p->idx, sizeof(int32_t) * 2 * (size_t)n));
            p->sort_cap = n;
        }
        k_corner_keys<<<blocks(n, 256), 256, 0, st>>>(n, d_pos, N, norm, delta, p->key, p->idx);
        rocprim::double_buffer<uint32_t> kb(p->key, p->key + p->sort_cap);
        rocprim::double_buffer<int32_t> vb(p->idx, p->idx + p->sort_cap);
        unsigned bits = 1;
        while (bits < 32 && ((uint64_t)1 << bits) < (uint64_t)n3) bits++;
        size_t tmp_bytes = 0;
        PM_CHECK(rocprim::radix_sort_pairs(nullptr, tmp_bytes, kb, vb, (size_t)n, 0, bits, st));
        if (tmp_bytes > p->sort_tmp_bytes) {
            if (p->sort_tmp) PM_CHECK(hipFree(p->sort_tmp));
            p->sort_tmp = nullptr;
            p->sort_tmp_bytes = 0;
            PM_CHECK(hipMalloc(&p->sort_tmp, tmp_bytes));
            p->sort_tmp_bytes = tmp_bytes;
        }
        PM_CHECK(rocprim::radix_sort_pairs(p->sort_tmp, tmp_bytes, kb, vb, (size_t)n, 0, bits, st));
        const uint32_t *skey = kb.current();
        const int32_t *sidx = vb.current();
        k_fill<<<blocks(n3, 256), 256, 0, st>>>(p->mesh, n3, 0.0);
        k_deposit_runs<<<blocks(n, 256), 256, 0, st>>>(n, d_pos, skey, sidx, p->mass, N, norm, delta, p->mesh);
        k_scale<<<blocks(n3, 256), 256, 0, st>>>(p->mesh, n3, norm * norm * norm);
        FFT_CHECK(hipfftExecD2Z(p->fwd, p->mesh, p->spec));
        const double ismth2 = (2 * kPiTrunc * p->rs / p->L) * (2 * kPiTrunc * p->rs / p->L);
        const double pref = p->L * p->L / (kPiTrunc * (double)N * N * N);
        k_green<<<blocks(nc, 256), 256, 0, st>>>(p->spec, N, ismth2, pref);
        FFT_CHECK(hipfftExecZ2D(p->inv, p->spec, p->mesh));
        k_gradient<<<blocks(n3, 256), 256, 0, st>>>(p->mesh, N, 0.5 * N / p->L, p->grad, p->grad + n3,
                                                    p->grad + 2 * n3);
        k_interpolate<<<blocks(n, 256), 256, 0, st>>>(n, d_pos, sidx, N, norm, delta, p->grad, p->grad + n3,
                                                      p->grad + 2 * n3, d_acc);
        PM_CHECK(hipGetLastError());
        return 0;
    } catch (const std::exception &e) {
        return fail(std::string("photons_pm_force_device: ") + e.what());
    }
}

int photons_pm_force_host(photons_pm *p, int64_t n, const double *pos, double *acc)
{
    try {
        if (n < 0) return fail("negative particle count");
        PM_CHECK(hipSetDevice(p->device));
        if (n == 0) return 0;
        if (n > p->io_cap) {
            if (p->io_pos) PM_CHECK(hipFree(p->io_pos));
            if (p->io_acc) PM_CHECK(hipFree(p->io_acc));
            p->io_pos = p->io_acc = nullptr;
            p->io_cap = 0;
            PM_CHECK(hipMalloc(&p->io_pos, sizeof(double) * 3 * (size_t)n));
            PM_CHECK(hipMalloc(&p->io_acc, sizeof(double) * 3 * (size_t)n));
            p->io_cap = n;
        }
        PM_CHECK(hipMemcpyAsync(p->io_pos, pos, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, p->stream));
        int rc = photons_pm_force_device(p, n, p->io_pos, p->io_acc);
        if (rc) return rc;
        PM_CHECK(hipMemcpyAsync(acc, p->io_acc, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToHost, p->stream));
        PM_CHECK(hipStreamSynchronize(p->stream));
        return 0;
    } catch (const std::exception &e) {
        return fail(std::string("photons_pm_force_host: ") + e.what());
    }
}

int photons_pm_sync(photons_pm *p)
{
    if (hipStreamSynchronize(p->stream) != hipSuccess) return fail("photons_pm_sync: stream error");
    return 0;
}

}  // extern "C"
