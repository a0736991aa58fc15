// Host driver of the 3D elastic adjoint-state gradient (no counterpart in the
// reference): checkpoints of the forward run, recomputation of each checkpoint
// segment with the forward kernels recording the derivative sums of every step,
// and the time-reversed adjoint steps of adjoint3d.hip.
//
// Memory plan (three levels, sized at set_checkpointing from a device-memory budget):
//   1. the forward run saves the state (9 fields in FP16 with one scale per field and
//      shot, or FP32, plus the CPML memory variables in FP32) every `every` steps --
//      in device memory when all checkpoints fit in half of the budget and leave room
//      for a sub-segment, otherwise in pinned host memory (refused beyond 80% of the
//      host memory available), copied out through a device staging buffer on a second
//      stream so the time loop does not wait for the transfer;
//   2. the adjoint pass walks the segments last to first; a segment of `every` steps is
//      recomputed from its checkpoint in sub-segments of K2 steps, saving the FP32
//      state at the start of each sub-segment (nsub = ceil(every / K2) states);
//   3. each sub-segment, last to first, is recomputed once more with the forward
//      kernels recording the nine derivative sums of every step (FP32, K2 steps),
//      then its adjoint steps run.
// K2 is the largest sub-segment whose record plus sub-segment states fit the rest of
// the budget; with K2 = every the middle level disappears (one recomputation).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "adjoint.hpp"
#include "adjoint3d.hpp"
#include "elastic3d.hpp"
#include "prep.hpp"

namespace wv {

namespace {
constexpr int ORG = 64;  // margin of position-indexed tables (as the forward tables)
int round_up(int v, int m) { return (v + m - 1) / m * m; }
// axis of each CPML memory variable (Cpml3D order)
constexpr int AXIS_OF[18] = {0, 1, 2, 0, 1, 2, 0, 1, 2, 0, 1, 2, 1, 0, 2, 0, 2, 1};

// Host memory this process may still take: MemAvailable of /proc/meminfo, lowered to
// what the cgroup's limit leaves (v2 memory.max - memory.current, or v1
// memory.limit_in_bytes - memory.usage_in_bytes).  < 0 when nothing is readable.
double host_bytes_available()
{
    double avail = -1;
    if (FILE *f = std::fopen("/proc/meminfo", "r")) {
        char key[64];
        unsigned long long kb;
        while (std::fscanf(f, "%63s %llu kB", key, &kb) == 2)
            if (!std::strcmp(key, "MemAvailable:")) {
                avail = (double)kb * 1024.0;
                break;
            }
        std::fclose(f);
    }
    auto read_num = [](const char *path, double &v) {
        FILE *f = std::fopen(path, "r");
        if (!f) return false;
        unsigned long long x;
        const bool ok = std::fscanf(f, "%llu", &x) == 1;  // "max" (no limit) does not parse
        std::fclose(f);
        if (ok) v = (double)x;
        return ok;
    };
    double lim, cur;
    if ((read_num("/sys/fs/cgroup/memory.max", lim) && read_num("/sys/fs/cgroup/memory.current", cur)) ||
        (read_num("/sys/fs/cgroup/memory/memory.limit_in_bytes", lim) &&
         read_num("/sys/fs/cgroup/memory/memory.usage_in_bytes", cur))) {
        const double left = std::max(0.0, lim - cur);
        avail = avail < 0 ? left : std::min(avail, left);
    }
    return avail;
}
}  // namespace

#define HIPA(expr)                                                                                           \
    do {                                                                                                     \
        hipError_t e_ = (expr);                                                                              \
        if (e_ != hipSuccess) return fail(WV_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct Elastic3D::Adj3State {
    int every = 0;
    bool fp16 = true;
    double budget_gib = -1;
    size_t fld_n = 0;     // floats of one field, all shots
    size_t psi_n = 0;     // floats of all 18 CPML memory variables
    size_t state_n = 0;   // 9 * fld_n + psi_n
    // adjoint pass
    std::vector<void *> allocs;
    Adj3D A{};
    Grad3D G{};
    float *tapT[6] = {};
    float *adj_src = nullptr;  // [NT][rec_total]
    float *rec = nullptr;      // [K2][9][fld_n]
    std::vector<float *> sub;  // sub-segment states (FP32), empty when nsub == 1
    int K2 = 0, nsub = 1;
    // checkpoints
    int nck = 0;
    bool on_host = false;
    size_t ck_fld_bytes = 0;  // fields of one checkpoint (FP16 or FP32)
    std::vector<void *> ck_fields;
    std::vector<float *> ck_psi;
    std::vector<void *> host_allocs;
    unsigned *ck_max = nullptr;  // [nck][9][nshots]
    void *stage = nullptr;       // device staging of one checkpoint when on_host
    hipStream_t cst = nullptr;
    hipEvent_t stage_free = nullptr, packed = nullptr;
    std::vector<char> have;
    bool grad_ready = false;
};

int Elastic3D::checkpoint_interval() const { return adj3 ? adj3->every : 0; }

size_t Elastic3D::psi_floats() const
{
    if (damping != 2) return 0;
    size_t n = 0;
    for (int q = 0; q < 18; q++) n += ps_n[AXIS_OF[q]];
    return n;
}

void Elastic3D::free_adjoint()
{
    if (!adj3) return;
    Adj3State *a = adj3;
    if (a->cst) (void)hipStreamSynchronize(a->cst);
    for (void *p : a->allocs) (void)hipFree(p);
    for (void *p : a->host_allocs) (void)hipHostFree(p);
    if (a->stage_free) (void)hipEventDestroy(a->stage_free);
    if (a->packed) (void)hipEventDestroy(a->packed);
    if (a->cst) (void)hipStreamDestroy(a->cst);
    delete a;
    adj3 = nullptr;
}

void Elastic3D::adjoint_plan(int *k2, int *nsub, int *on_host) const
{
    if (k2) *k2 = adj3 ? adj3->K2 : 0;
    if (nsub) *nsub = adj3 ? adj3->nsub : 0;
    if (on_host) *on_host = adj3 && adj3->on_host;
}

// the 9 fields then the 18 CPML memory variables of the batch, FP32
int Elastic3D::save_state(float *dst)
{
    const size_t fn = (size_t)L.vol * nshots;
    for (int f = 0; f < 9; f++) HIPA(hipMemcpyAsync(dst + f * fn, F[f], fn * 4, hipMemcpyDeviceToDevice, st));
    float *p = dst + 9 * fn;
    if (damping == 2)
        for (int q = 0; q < 18; q++) {
            HIPA(hipMemcpyAsync(p, ps[q], ps_n[AXIS_OF[q]] * 4, hipMemcpyDeviceToDevice, st));
            p += ps_n[AXIS_OF[q]];
        }
    return 0;
}

int Elastic3D::restore_state(const float *src)
{
    const size_t fn = (size_t)L.vol * nshots;
    for (int f = 0; f < 9; f++) HIPA(hipMemcpyAsync(F[f], src + f * fn, fn * 4, hipMemcpyDeviceToDevice, st));
    const float *p = src + 9 * fn;
    if (damping == 2)
        for (int q = 0; q < 18; q++) {
            HIPA(hipMemcpyAsync(ps[q], p, ps_n[AXIS_OF[q]] * 4, hipMemcpyDeviceToDevice, st));
            p += ps_n[AXIS_OF[q]];
        }
    return 0;
}

int Elastic3D::save_checkpoint(int k)
{
    Adj3State *a = adj3;
    if (k < 0 || k >= a->nck) return 0;
    const size_t fn = a->fld_n;
    // fields, then the memory variables, at dst (device: the checkpoint itself, or the staging buffer)
    char *dst = (char *)(a->on_host ? a->stage : a->ck_fields[k]);
    if (a->on_host) HIPA(hipStreamWaitEvent(st, a->stage_free, 0));  // the previous copy-out has read the stage
    for (int f = 0; f < 9; f++) {
        if (a->fp16)
            launch_ckpt_pack(F[f], (unsigned short *)dst + f * fn, a->ck_max + ((size_t)k * 9 + f) * nshots, L.vol,
                             nshots, st);
        else
            HIPA(hipMemcpyAsync((float *)dst + f * fn, F[f], fn * 4, hipMemcpyDeviceToDevice, st));
    }
    float *pdst = a->on_host ? (float *)(dst + a->ck_fld_bytes) : a->ck_psi[k];
    if (damping == 2)
        for (int q = 0; q < 18; q++) {
            HIPA(hipMemcpyAsync(pdst, ps[q], ps_n[AXIS_OF[q]] * 4, hipMemcpyDeviceToDevice, st));
            pdst += ps_n[AXIS_OF[q]];
        }
    if (a->on_host) {
        HIPA(hipEventRecord(a->packed, st));
        HIPA(hipStreamWaitEvent(a->cst, a->packed, 0));
        HIPA(hipMemcpyAsync(a->ck_fields[k], a->stage, a->ck_fld_bytes + a->psi_n * 4, hipMemcpyDeviceToHost,
                            a->cst));
        HIPA(hipEventRecord(a->stage_free, a->cst));
    }
    a->have[k] = 1;
    return 0;
}

int Elastic3D::restore_checkpoint(int k)
{
    Adj3State *a = adj3;
    const size_t fn = a->fld_n;
    const char *src = (const char *)a->ck_fields[k];
    if (a->on_host) {
        HIPA(hipStreamSynchronize(a->cst));
        HIPA(hipMemcpyAsync(a->stage, a->ck_fields[k], a->ck_fld_bytes + a->psi_n * 4, hipMemcpyHostToDevice, st));
        src = (const char *)a->stage;
    }
    for (int f = 0; f < 9; f++) {
        if (a->fp16)
            launch_ckpt_unpack((const unsigned short *)src + f * fn, F[f], a->ck_max + ((size_t)k * 9 + f) * nshots,
                               L.vol, nshots, st);
        else
            HIPA(hipMemcpyAsync(F[f], (const float *)src + f * fn, fn * 4, hipMemcpyDeviceToDevice, st));
    }
    const float *p = a->on_host ? (const float *)(src + a->ck_fld_bytes) : a->ck_psi[k];
    if (damping == 2)
        for (int q = 0; q < 18; q++) {
            HIPA(hipMemcpyAsync(ps[q], p, ps_n[AXIS_OF[q]] * 4, hipMemcpyDeviceToDevice, st));
            p += ps_n[AXIS_OF[q]];
        }
    return 0;
}

// nsteps forward steps from the device step counter, recording step j's derivative sums at
// rec + j * 9 * fld_n (eager launches; ends with the flush of the last step)
int Elastic3D::enqueue_rec(int nsteps, float *rec)
{
    const Fields3D Fd = fields();
    const Coef3D C = coefs();
    const Taps3D T = taps();
    const Cpml3D P = cpml();
    const Acq A = acq();
    const size_t fn = adj3->fld_n;
    for (int j = 0; j < nsteps; j++) {
        Rec3D R;
        for (int k = 0; k < 9; k++) R.r[k] = rec + ((size_t)j * 9 + k) * fn;
        launch_el3d(0, order, rz, L, nshots, nyc, Fd, C, T, P, A, j, j > 0, st, nullptr, nullptr, &R);
        launch_el3d(1, order, rz, L, nshots, nyc, Fd, C, T, P, A, j, 0, st, nullptr, nullptr, &R);
    }
    if (nsteps > 0) launch_flush3d(L, Fd, A, nsteps - 1, st);
    launch_advance(d_tbase, nsteps, st);
    HIPA(hipGetLastError());
    return 0;
}

int Elastic3D::set_checkpointing(int every, int fp16, double budget_gib)
{
    if (physics != ELASTIC || visco) return fail(WV_EINVAL, "the 3D adjoint is provided for elastic handles");
    if (free_surface || damping == 1)
        return fail(WV_EINVAL, "the adjoint supports CPML or no absorbing boundary, without free surface");
    if (every < 0 || every > NT) return fail(WV_EINVAL, "checkpoint interval out of range");
    HIPA(hipStreamSynchronize(st));
    free_adjoint();
    if (!every) return reset();
    int rc = sync_acquisition();
    if (rc) return rc;
    adj3 = new Adj3State();
    Adj3State *a = adj3;
    a->every = every;
    a->fp16 = fp16 != 0;
    a->budget_gib = budget_gib;
    a->fld_n = (size_t)L.vol * nshots;
    a->psi_n = psi_floats();
    a->state_n = 9 * a->fld_n + a->psi_n;
    a->nck = (NT + every - 1) / every;
    a->have.assign(a->nck, 0);
    auto dal = [&](void **p, size_t bytes) -> int {
        if (hipMalloc(p, std::max<size_t>(bytes, 4)) != hipSuccess) {
            (void)hipGetLastError();
            return fail(WV_ENOMEM, "hipMalloc failed (3D adjoint buffers)");
        }
        a->allocs.push_back(*p);
        HIPA(hipMemsetAsync(*p, 0, std::max<size_t>(bytes, 4), st));
        return 0;
    };
    // adjoint pass buffers first, so the budget below is what is left next to them
    for (int f = 0; f < 9; f++) {
        if ((rc = dal((void **)&a->A.l[f], a->fld_n * 4)) || (rc = dal((void **)&a->A.g[f], a->fld_n * 4))) return rc;
    }
    for (int q = 0; q < 18; q++) {
        a->A.m[q] = nullptr;
        if (damping == 2 && (rc = dal((void **)&a->A.m[q], ps_n[AXIS_OF[q]] * 4))) return rc;
    }
    for (int k = 0; k < 8; k++)
        if ((rc = dal((void **)&a->G.g[k], (size_t)L.vol * 8))) return rc;
    if ((rc = dal((void **)&a->adj_src, (size_t)NT * std::max(rec_total, 1) * 4))) return rc;
    {
        // transposed derivative tables, position-indexed with the forward tables' margins
        const int K = 2 * H + 1;
        const int NXp = round_up(NX, 64) + 64, NYp = NY + 64, NZp = round_up(NZ, 16) + 64;
        const bool stencil = ops_stencil;
        const float fDH = (float)DH, fDT = (float)DT;
        const int n[3] = {NX, NY, NZ}, np[3] = {NXp, NYp, NZp};
        for (int q = 0; q < 6; q++) {
            std::vector<float> t;
            build_taps(t, n[q / 2], np[q / 2], order, H, q % 2 == 0, stencil, fDH, fDT);
            std::vector<float> tt = transpose_taps(t, H);
            std::vector<float> ext((size_t)(np[q / 2] + 2 * ORG) * K, 0.f);
            std::copy(tt.begin(), tt.end(), ext.begin() + (size_t)ORG * K);
            float *base;
            if ((rc = dal((void **)&base, ext.size() * 4))) return rc;
            HIPA(hipMemcpyAsync(base, ext.data(), ext.size() * 4, hipMemcpyHostToDevice, st));
            HIPA(hipStreamSynchronize(st));  // ext is freed at the end of this iteration
            a->tapT[q] = base + (size_t)ORG * K;
        }
    }
    // budget for checkpoints, sub-segment states and the record
    size_t free_b = 0, total_b = 0;
    HIPA(hipMemGetInfo(&free_b, &total_b));
    const double avail = std::max(0.0, (double)free_b - 256.0e6);
    const double budget = a->budget_gib <= 0 ? 0.75 * avail : std::min(a->budget_gib * 1073741824.0, avail);
    a->ck_fld_bytes = a->fld_n * 9 * (a->fp16 ? 2 : 4);
    const double ck_bytes = (double)a->ck_fld_bytes + (double)a->psi_n * 4;
    const double all_ck = ck_bytes * a->nck;
    const double rec_step = 9.0 * a->fld_n * 4, state = (double)a->state_n * 4;
    // the largest sub-segment whose record and sub-segment states fit next to the checkpoints
    auto largest_k2 = [&](bool host) {
        const double left = budget - (host ? ck_bytes : all_ck);  // host checkpoints need one device stage
        for (int k = every; k >= 1; k--) {
            const int ns = (every + k - 1) / k;
            if (k * rec_step + (ns > 1 ? ns * state : 0.0) <= left) return k;
        }
        return 0;
    };
    a->on_host = all_ck > 0.5 * budget;
    int K2 = largest_k2(a->on_host);
    if (K2 == 0 && !a->on_host) {  // device-resident checkpoints crowd out the record: move them out
        a->on_host = true;
        K2 = largest_k2(true);
    }
    if (K2 == 0) return fail(WV_ENOMEM, "the adjoint's recomputation buffers do not fit the device-memory budget");
    if (a->on_host) {
        // pinned host memory cannot be swapped: refuse a plan the host cannot hold
        const double host_avail = host_bytes_available();
        if (host_avail >= 0 && all_ck > 0.8 * host_avail)
            return fail(WV_ENOMEM, "the checkpoints (" + std::to_string((long long)(all_ck / 1073741824.0)) +
                                       " GiB) do not fit the host memory available (" +
                                       std::to_string((long long)(host_avail / 1073741824.0)) +
                                       " GiB): checkpoint less often or run fewer shots per handle");
    }
    a->K2 = K2;
    a->nsub = (every + K2 - 1) / K2;
    if ((rc = dal((void **)&a->rec, (size_t)K2 * 9 * a->fld_n * 4))) return rc;
    if (a->nsub > 1)
        for (int j = 0; j < a->nsub; j++) {
            float *p;
            if ((rc = dal((void **)&p, a->state_n * 4))) return rc;
            a->sub.push_back(p);
        }
    if (a->fp16 && (rc = dal((void **)&a->ck_max, (size_t)a->nck * 9 * nshots * 4))) return rc;
    if (a->on_host) {
        if ((rc = dal(&a->stage, (size_t)ck_bytes))) return rc;
        HIPA(hipStreamCreateWithFlags(&a->cst, hipStreamNonBlocking));
        HIPA(hipEventCreateWithFlags(&a->stage_free, hipEventDisableTiming));
        HIPA(hipEventCreateWithFlags(&a->packed, hipEventDisableTiming));
        HIPA(hipEventRecord(a->stage_free, a->cst));
        for (int k = 0; k < a->nck; k++) {
            void *h = nullptr;
            if (hipHostMalloc(&h, (size_t)ck_bytes, hipHostMallocDefault) != hipSuccess) {
                (void)hipGetLastError();
                return fail(WV_ENOMEM, "hipHostMalloc failed (3D checkpoints in host memory)");
            }
            a->host_allocs.push_back(h);
            a->ck_fields.push_back(h);
            a->ck_psi.push_back(nullptr);
        }
    } else {
        for (int k = 0; k < a->nck; k++) {
            void *f;
            float *p = nullptr;
            if ((rc = dal(&f, a->ck_fld_bytes))) return rc;
            if (a->psi_n && (rc = dal((void **)&p, a->psi_n * 4))) return rc;
            a->ck_fields.push_back(f);
            a->ck_psi.push_back(p);
        }
    }
    drop_graph();
    return reset();
}

int Elastic3D::set_adjoint_source(int shot, const float *adj)
{
    if (!adj3) return fail(WV_ESTATE, "wv_set_adjoint_source before wv_set_checkpointing");
    int rc = sync_acquisition();
    if (rc) return rc;
    const int nr = nrec(shot);
    if (nr == 0) return 0;
    // time-major [NT][rec_total], the layout of the seismograms
    std::vector<float> h((size_t)NT * rec_total);
    HIPA(hipStreamSynchronize(st));
    HIPA(hipMemcpy(h.data(), adj3->adj_src, h.size() * 4, hipMemcpyDeviceToHost));
    const int r0 = rec_ptr_h[shot];
    for (int r = 0; r < nr; r++)
        for (int t = 0; t < NT; t++) h[(size_t)t * rec_total + r0 + r] = adj[(size_t)r * NT + t];
    HIPA(hipMemcpy(adj3->adj_src, h.data(), h.size() * 4, hipMemcpyHostToDevice));
    return 0;
}

int Elastic3D::run_adjoint()
{
    Adj3State *a = adj3;
    if (!a) return fail(WV_ESTATE, "wv_run_adjoint before wv_set_checkpointing");
    for (int k = 0; k < a->nck; k++)
        if (!a->have[k]) return fail(WV_ESTATE, "wv_run_adjoint needs a checkpointed wv_run(0, NT) first");
    int rc = sync_acquisition();
    if (rc) return rc;
    if (a->cst) HIPA(hipStreamSynchronize(a->cst));
    for (int f = 0; f < 9; f++) HIPA(hipMemsetAsync(a->A.l[f], 0, a->fld_n * 4, st));
    for (int q = 0; q < 18; q++)
        if (a->A.m[q]) HIPA(hipMemsetAsync(a->A.m[q], 0, ps_n[AXIS_OF[q]] * 4, st));
    for (int k = 0; k < 8; k++) HIPA(hipMemsetAsync(a->G.g[k], 0, (size_t)L.vol * 8, st));
    const Coef3D C = coefs();
    const Cpml3D P = cpml();
    const Acq Q = acq();
    const TapsT3D TT{a->tapT[0], a->tapT[1], a->tapT[2], a->tapT[3], a->tapT[4], a->tapT[5]};
    const int adj_nyc = std::max(1, std::min(nyc, NY / std::max(4 * H, 8)));
    for (int seg = a->nck - 1; seg >= 0; seg--) {
        const int s0 = seg * a->every, s1 = std::min(NT, s0 + a->every);
        if ((rc = restore_checkpoint(seg))) return rc;
        const int nsub = (s1 - s0 + a->K2 - 1) / a->K2;
        if (nsub > 1) {
            if ((rc = save_state(a->sub[0]))) return rc;
            for (int j = 1; j < nsub; j++) {
                if ((rc = run_segment(s0 + (j - 1) * a->K2, s0 + j * a->K2, false))) return rc;
                if ((rc = save_state(a->sub[j]))) return rc;
            }
        }
        for (int j = nsub - 1; j >= 0; j--) {
            const int t0 = s0 + j * a->K2, t1 = std::min(s1, t0 + a->K2);
            if (nsub > 1 && (rc = restore_state(a->sub[j]))) return rc;
            launch_set(d_tbase, t0, st);
            if ((rc = enqueue_rec(t1 - t0, a->rec))) return rc;
            for (int t = t1 - 1; t >= t0; t--) {
                Rec3D R;
                for (int k = 0; k < 9; k++) R.r[k] = a->rec + ((size_t)(t - t0) * 9 + k) * a->fld_n;
                launch_adj3d_step(order, L, nshots, adj_nyc, a->A, C, P, R, a->G, TT, Q, a->adj_src, t, st);
            }
            HIPA(hipGetLastError());
        }
    }
    a->grad_ready = true;
    return 0;
}

int Elastic3D::get_gradient(int param, double *out)
{
    Adj3State *a = adj3;
    if (!a || !a->grad_ready) return fail(WV_ESTATE, "no gradient computed");
    if (param < 0 || param > 10 || !out) return fail(WV_EINVAL, "param must be 0..10");
    const size_t n = (size_t)NX * NY * NZ;
    auto gi = [&](int x, int y, int z) { return (size_t)x + (size_t)z * NX + (size_t)y * NX * NZ; };
    HIPA(hipStreamSynchronize(st));
    std::vector<double> g[8];
    std::vector<double> h((size_t)L.vol);
    for (int k = 0; k < 8; k++) {
        HIPA(hipMemcpy(h.data(), a->G.g[k], h.size() * 8, hipMemcpyDeviceToHost));
        g[k].resize(n);
        for (int y = 0; y < NY; y++)
            for (int z = 0; z < NZ; z++)
                for (int x = 0; x < NX; x++) g[k][gi(x, y, z)] = h[(size_t)L.at(x, y, z)];
    }
    if (param >= 3) {
        std::memcpy(out, g[param - 3].data(), n * 8);
        return 0;
    }
    // chain rule through the averaging rules (Modelparameter.cpp:336-639): coefficients -> vp, vs, rho
    const std::vector<float> &vp = model_h[0], &vs = model_h[1], &rho = model_h[2];
    if (vp.size() != n) return fail(WV_ESTATE, "model not retained");
    std::vector<double> mu(n), gmu(g[4]), gvp(n, 0.0), gvs(n, 0.0), grho(n, 0.0);
    for (size_t i = 0; i < n; i++) mu[i] = std::max((double)rho[i] * vs[i] * vs[i], 1.0);
    const int len[3] = {NX, NY, NZ};
    static const int planes[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int pl = 0; pl < 3; pl++) {
        const int a0 = planes[pl][0], a1 = planes[pl][1];
        for (int y = 0; y < NY; y++)
            for (int z = 0; z < NZ; z++)
                for (int x = 0; x < NX; x++) {
                    const int pos[3] = {x, y, z};
                    const bool in0 = pos[a0] + 1 < len[a0], in1 = pos[a1] + 1 < len[a1];
                    size_t pts[4];
                    int np = 0;
                    pts[np++] = gi(x, y, z);
                    int p[3] = {x, y, z};
                    if (in0) {
                        p[a0]++;
                        pts[np++] = gi(p[0], p[1], p[2]);
                        p[a0]--;
                    }
                    if (in1) {
                        p[a1]++;
                        pts[np++] = gi(p[0], p[1], p[2]);
                        p[a1]--;
                    }
                    if (in0 && in1) {
                        p[a0]++;
                        p[a1]++;
                        pts[np++] = gi(p[0], p[1], p[2]);
                    }
                    const double w = 1.0 / np;
                    double avg = 0;
                    for (int q = 0; q < np; q++) avg += 1.0 / mu[pts[q]];
                    const double m = 1.0 / (w * avg);
                    if (m < 4.0) continue;  // the average is set to 0: no dependence
                    const double gj = g[5 + pl][pts[0]];
                    for (int q = 0; q < np; q++) gmu[pts[q]] += gj * m * m * w / (mu[pts[q]] * mu[pts[q]]);
                }
    }
    for (size_t i = 0; i < n; i++) {
        const double r = rho[i], p = vp[i], v = vs[i];
        gvp[i] = g[3][i] * 2.0 * r * p;
        grho[i] += g[3][i] * p * p;
        if (r * v * v >= 1.0) {  // mu not clamped
            gvs[i] = gmu[i] * 2.0 * r * v;
            grho[i] += gmu[i] * v * v;
        }
    }
    for (int ax = 0; ax < 3; ax++)
        for (int y = 0; y < NY; y++)
            for (int z = 0; z < NZ; z++)
                for (int x = 0; x < NX; x++) {
                    const int pos[3] = {x, y, z};
                    const size_t i = gi(x, y, z);
                    const double gk = g[ax][i];
                    if (pos[ax] + 1 < len[ax]) {
                        int p[3] = {x, y, z};
                        p[ax]++;
                        const size_t j = gi(p[0], p[1], p[2]);
                        const double avg = 0.5 * ((double)rho[i] + rho[j]);
                        const double d = -0.5 / (avg * avg);
                        grho[i] += gk * d;
                        grho[j] += gk * d;
                    } else {
                        grho[i] += gk * (-1.0 / ((double)rho[i] * rho[i]));
                    }
                }
    const std::vector<double> &res = param == 0 ? gvp : (param == 1 ? gvs : grho);
    std::memcpy(out, res.data(), n * 8);
    return 0;
}

}  // namespace wv
