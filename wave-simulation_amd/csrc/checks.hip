// Checks the reference's shot loop makes around the time loop:
//  * wv_check_parameters: CheckParameter::checkNumericalArtefactsAndInstabilities
//    (src/CheckParameter/CheckParameter.hpp:166-246) with checkStabilityCriterion
//    (:60-112) and checkNumericalDispersion (:121-162), on the host;
//  * wv_check_finite: the per-shot test of Simulation.cpp:519 that wavefields and
//    seismograms hold no Inf/NaN, as one GPU pass over the handle's buffers;
//  * wv_check_finite_noalloc: the same test with its scratch memory in a static device
//    array (no allocation), counting the non-finite values per buffer.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "handle.hpp"

namespace {

// any non-finite value in p[0, n) sets *flag (plain stores: every writer writes 1)
__global__ __launch_bounds__(256) void nonfinite_scan(const float *__restrict__ p, long long n, unsigned *flag)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    bool bad = false;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const unsigned u = __float_as_uint(p[i]);
        bad |= (u & 0x7f800000u) == 0x7f800000u;  // exponent all ones: Inf or NaN
    }
    if (bad) *flag = 1u;
}

// one count per state buffer of wv_check_finite_noalloc; each GPU holds its own copy
__device__ __align__(256) unsigned long long scratch[WV_NOALLOC_MAX_BUFFERS];

// number of non-finite values in p[0, n), added to *count (one atomic per wavefront that saw any)
__global__ __launch_bounds__(256) void nonfinite_count(const float *__restrict__ p, long long n,
                                                       unsigned long long *count)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    unsigned c = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        c += (__float_as_uint(p[i]) & 0x7f800000u) == 0x7f800000u;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(count, (unsigned long long)c);
}

// zero the first n counts (a kernel, not hipMemsetAsync: the call may sit in a captured graph)
__global__ void scratch_clear(unsigned long long *p, int n)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// device address of `scratch` on each GPU, looked up once per device with that device current
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
unsigned long long *scratch_ptr[MAX_DEVICES] = {};

int scratch_address(unsigned long long **out)
{
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(WV_EDEVICE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(WV_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        scratch_ptr[dev] = (unsigned long long *)p;
    }
    *out = scratch_ptr[dev];
    return WV_OK;
}

// the device buffers holding the handle's state (allocates nothing)
void state_buffers(const wv_solver *s, wv::StateBuffers &b)
{
    if (s->plain)
        s->plain->buffers(b);
    else if (s->e3)
        s->e3->buffers(b);
    else {
        for (int f = 0; f < s->nfields; f++) b.add(s->field[s->cur][f], (long long)s->L.plane * s->cfg.nshots);
        if (s->d_seis && !s->acq_dirty) b.add(s->d_seis, (long long)s->cfg.NT * s->rec_total);
    }
}

constexpr const char *NONFINITE_MSG = "Infinite or NaN value in seismogram or/and velocity wavefield!";

std::string lower(const char *s)
{
    std::string r = s ? s : "";
    for (auto &c : r) c = (char)std::tolower((unsigned char)c);
    return r;
}

}  // namespace

extern "C" {

int wv_check_parameters(const char *dimension, const char *type, const wv_config *cfg, const float *m0,
                        const float *m1, const float *m2, double fc_max, int *dispersion_ok)
{
    if (dispersion_ok) *dispersion_ok = 1;
    if (!cfg || !m0) return fail(WV_EINVAL, "null argument");
    const std::string d = dimension ? dimension : "", t = lower(type);
    int D;
    if (d == "2D")
        D = 2;
    else if (d == "3D")
        D = 3;
    else
        return fail(WV_EINVAL, "Unknown dimension");
    const bool em = t == "tmem" || t == "emem" || t == "viscotmem" || t == "viscoemem";
    const bool elastic = t == "elastic" || t == "viscoelastic";
    const bool sh = t == "sh" || t == "viscosh";
    if (!em && !elastic && !sh && t != "acoustic") return fail(WV_EINVAL, "Unkown type " + t);
    if ((elastic || sh) && !m1) return fail(WV_EINVAL, "missing velocityS");
    if (em && (!m2 || !m0)) return fail(WV_EINVAL, "missing permittivity or permeability");
    // h of the stability criterion: the reference assigns 7 / 6, 149 / 120, 2161 / 1680, ...
    // in integer arithmetic, so h = 1 for every order
    float h;
    switch (cfg->spatial_fd_order) {
    case 2: h = 1; break;
    case 4: h = 7 / 6; break;
    case 6: h = 149 / 120; break;
    case 8: h = 2161 / 1680; break;
    case 10: h = 53089 / 40320; break;
    case 12: h = 1187803 / 887040; break;
    default: return fail(WV_EINVAL, "Unknown spatial FD order");
    }
    int N;  // points per minimum wavelength of the dispersion criterion
    switch (cfg->spatial_fd_order) {
    case 2: N = 12; break;
    case 4: N = 8; break;
    case 6: N = 7; break;
    case 8: N = 6; break;
    case 10: N = 5; break;
    default: N = 4; break;
    }
    const size_t n = (size_t)cfg->NX * cfg->NY * (D == 3 ? cfg->NZ : 1);
    // the model in the reference's ValueType (float): vmax from vp (vs for SH, the EM
    // velocity 1/sqrt(eps mu) for EM), vmin over the positive values of vp (acoustic),
    // vs (elastic, SH) or the EM velocity
    const float eps0 = 8.8541878176e-12f, mu0 = 1.2566370614e-6f;
    float vmax = 0.f, vmin = 3e8f;
    for (size_t i = 0; i < n; i++) {
        float vhi, vlo;
        if (em) {
            float v = 1.f / std::sqrt((m0[i] * eps0) * (m2[i] * mu0));
            if (!std::isfinite(v)) v = 0.f;  // replaceInvalid(velocity, 0)
            vhi = vlo = v;
        } else if (sh) {
            vhi = vlo = m1[i];
        } else {
            vhi = m0[i];
            vlo = elastic ? m1[i] : m0[i];
        }
        if (elastic && !(vhi / vlo >= (float)std::sqrt(2.0))) {
            const size_t xy = (size_t)cfg->NX * (D == 3 ? cfg->NZ : 1);
            std::ostringstream m;
            m << "vp/vs (" << vhi << "/" << vlo << ") < sqrt(2.0) at X,Y,Z =" << i % cfg->NX << "," << i / xy << ","
              << (D == 3 ? (i / cfg->NX) % cfg->NZ : 0);
            return fail(WV_EINVAL, m.str());
        }
        if (vhi > vmax) vmax = vhi;
        if (vlo < vmin && vlo > 0.f) vmin = vlo;
    }
    const float dt = (float)cfg->DT, DH = (float)cfg->DH;
    const double bound = DH / (h * std::sqrt(D) * vmax);
    if (dt > bound) {
        std::ostringstream m;
        m << "Courant-Friedrichs-Lewy-Criterion is not met: dt is " << dt
          << " but should be less than DH/(h*sqrt(D)*vpMax)=" << DH << "/ (" << h << "* sqrt(" << D << ") * " << vmax
          << ") =" << bound;
        return fail(WV_EINVAL, m.str());
    }
    const float fc = (float)fc_max;
    if (DH > vmin / (2 * fc * N)) {
        if (dispersion_ok) *dispersion_ok = 0;
        std::ostringstream m;
        m << "Criterion to avoid numerical dispersion is not met: DH is " << DH
          << " but should be less than vMin/(2*fcMax*N)=" << vmin << " / (2 * " << fc << " * " << N
          << ") = " << vmin / (2 * fc * N);
        wv_fail(WV_OK, m.str());  // a warning, as in the reference: the message stays readable
    }
    return WV_OK;
}

int wv_check_finite(wv_solver *s)
{
    if (!s || !s->inited) return fail(WV_ESTATE, "not initialised");
    wv::StateBuffers bufs;
    state_buffers(s, bufs);
    if (!s->d_flag) {
        int rc = s->dalloc(&s->d_flag, 1, s->allocs);
        if (rc) return rc;
    }
    HIP_TRY(hipMemsetAsync(s->d_flag, 0, sizeof(unsigned), s->st));
    for (int k = 0; k < bufs.k; k++) {
        const long long blocks = std::min<long long>((bufs.n[k] + 255) / 256, 4096);
        hipLaunchKernelGGL(nonfinite_scan, dim3((unsigned)blocks), dim3(256), 0, s->st, bufs.p[k], bufs.n[k],
                           s->d_flag);
    }
    HIP_TRY(hipGetLastError());
    unsigned h = 0;
    HIP_TRY(hipMemcpyAsync(&h, s->d_flag, sizeof(unsigned), hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    if (h) return fail(WV_ENONFINITE, NONFINITE_MSG);
    return WV_OK;
}

int wv_check_finite_noalloc(wv_solver *s, long long *nonfinite)
{
    if (nonfinite) *nonfinite = 0;
    if (!s || !s->inited) return fail(WV_ESTATE, "not initialised");
    wv::StateBuffers bufs;
    state_buffers(s, bufs);
    if (bufs.overflow)
        return fail(WV_EINVAL, "the handle has more state buffers than the static scratch holds "
                               "(WV_NOALLOC_MAX_BUFFERS)");
    unsigned long long *scr = nullptr;
    int rc = scratch_address(&scr);
    if (rc) return rc;
    if (bufs.k == 0) return WV_OK;
    hipLaunchKernelGGL(scratch_clear, dim3(1), dim3(64), 0, s->st, scr, bufs.k);
    for (int k = 0; k < bufs.k; k++) {
        const long long blocks = std::min<long long>((bufs.n[k] + 255) / 256, 4096);
        hipLaunchKernelGGL(nonfinite_count, dim3((unsigned)blocks), dim3(256), 0, s->st, bufs.p[k], bufs.n[k],
                           scr + k);
    }
    HIP_TRY(hipGetLastError());
    unsigned long long h[WV_NOALLOC_MAX_BUFFERS];
    HIP_TRY(hipMemcpyAsync(h, scr, sizeof(unsigned long long) * bufs.k, hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    unsigned long long total = 0;
    for (int k = 0; k < bufs.k; k++) total += h[k];
    if (nonfinite) *nonfinite = (long long)total;
    if (total) return fail(WV_ENONFINITE, NONFINITE_MSG);
    return WV_OK;
}

}  // extern "C"
