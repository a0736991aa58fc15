// Adjoint-state gradient of the 2D elastic forward solver (no counterpart in
// the reference: WAVE-Simulation has no adjoint).
//
// Forward run with checkpointing: before every `every`-th step the state (the
// five wavefields, FP16 with one scale per field and shot, or FP32, plus the
// CPML memory variables in FP32) is stored in HBM.  The backward pass walks
// the checkpoints from the last to the first: it restores one, recomputes its
// segment forward while the forward kernels record the per-step quantities
// the gradient needs (the five CPML-modified derivative sums), then runs the
// adjoint steps of the segment in reverse order.
// When the FP16 record of every step of the run fits in the device-memory
// budget (wv_set_adjoint_store) and the handle runs the row-strip kernels,
// the forward run records the sums of every step directly (FP16, one scale
// per row of 64 cells, field and shot) and the backward pass needs neither
// checkpoints nor recomputation.
// Gradients with respect to 1/rho_x, 1/rho_y, M, mu, mu_xy accumulate in FP64
// on the GPU; the chain rule to vp, vs, rho follows the averaging rules of the
// coefficient preparation (host, FP64).
#include <cmath>
#include <cstring>

#include "adjoint.hpp"
#include "handle.hpp"
#include "prep.hpp"

struct AdjointState {
    int every = 0;
    bool fp16 = true;
    int nck = 0;
    std::vector<char> have;
    int next_t = -1;  // step the checkpointed forward run has reached (-1: none)
    bool valid_run = false;  // the current wv_run continues a checkpointed run from step 0
    bool forward_done = false;
    bool grad_ready = false;
    size_t fld_n = 0;  // floats of one field of all shots
    std::vector<void *> ck_fields;
    std::vector<unsigned *> ck_max;
    std::vector<float *> ck_psi;
    float *segH = nullptr;  // recorded derivative sums of a segment: [step][5][fld_n]
    int seg_slots = 0;
    float *lam[5] = {};
    float *mu[2][8] = {};
    double *grad[5] = {};
    float *adj_src = nullptr;
    size_t adj_src_n = 0;
    std::vector<std::vector<float>> adj_host;
    int *zero_ptr = nullptr;
    size_t zero_n = 0;
    float *tapT[4] = {};
    std::vector<void *> allocs, ck_allocs;
    double store_gib = -1.0;  // budget of the per-step record (< 0: 3/4 of the free device memory)
    bool store = false;       // this checkpointing setup records every step instead of checkpointing
    wv::Rec2D rec{};
};

namespace {

void free_all(std::vector<void *> &l)
{
    for (void *p : l) (void)hipFree(p);
    l.clear();
}

template <typename T>
int adalloc(std::vector<void *> &l, T **p, size_t n, hipStream_t st)
{
    void *q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess)
        return fail(WV_ENOMEM, "hipMalloc failed (adjoint buffers)");
    l.push_back(q);
    if (hipMemsetAsync(q, 0, std::max<size_t>(n, 1) * sizeof(T), st) != hipSuccess)
        return fail(WV_EDEVICE, "hipMemset failed");
    *p = (T *)q;
    return 0;
}

// psi arrays of the phased 2D kernels in a fixed order: psx0 psx1 psy0 psy1 | psx2 psx3 psy2 psy3
void psi_list(const wv_solver *s, float *p[8], size_t n[8])
{
    for (int q = 0; q < 4; q++) {
        p[q] = s->psv[0][q];
        p[4 + q] = s->pss[q];
        n[q] = n[4 + q] = (q < 2 ? s->psx_n : s->psy_n);
    }
}

size_t psi_total(const wv_solver *s)
{
    return s->cfg.damping_boundary == 2 ? 4 * s->psx_n + 4 * s->psy_n : 0;
}

// start/stop events attached to timed dispatches, collected in batches
struct KernelTimer {
    wv_solver *s;
    std::vector<hipEvent_t> ev;
    std::vector<int> kind;
    size_t used = 0;
    explicit KernelTimer(wv_solver *sv) : s(sv) {}
    ~KernelTimer()
    {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
    int init(size_t pairs)
    {
        ev.resize(2 * pairs);
        kind.resize(pairs);
        for (auto &e : ev) HIP_TRY(hipEventCreate(&e));
        return 0;
    }
    int flush()
    {
        if (!used) return 0;
        HIP_TRY(hipEventSynchronize(ev[2 * used - 1]));
        for (size_t i = 0; i < used; i++) {
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
            s->kms[kind[i]] += ms;
            s->klaunch[kind[i]] += 1;
        }
        used = 0;
        return 0;
    }
    // events for the next dispatch of kind k (nullptr when timing is off)
    int next(int k, hipEvent_t *e0, hipEvent_t *e1)
    {
        *e0 = *e1 = nullptr;
        if (ev.empty()) return 0;
        if (used == kind.size()) {
            int rc = flush();
            if (rc) return rc;
        }
        kind[used] = k;
        *e0 = ev[2 * used];
        *e1 = ev[2 * used + 1];
        used++;
        return 0;
    }
};

}  // namespace

void adjoint_destroy(AdjointState *a)
{
    if (!a) return;
    free_all(a->ck_allocs);
    free_all(a->allocs);
    delete a;
}

int adjoint_every(const wv_solver *s) { return s->adj ? s->adj->every : 0; }

int adjoint_fp16(const wv_solver *s) { return s->adj && s->adj->fp16 ? 1 : 0; }

void adjoint_invalidate_layout(wv_solver *s)
{
    if (!s->adj) return;
    s->adj->every = 0;
    s->adj->forward_done = s->adj->grad_ready = false;
    s->adj->next_t = -1;
}

const wv::Rec2D *adjoint_record(const wv_solver *s)
{
    const AdjointState *a = s->adj;
    // only for the strip height the record was laid out for (wv_run_adjoint rejects a mismatch)
    if (!a || !a->every || !a->store || s->variant != V_STRIP || s->tile_y / 4 != a->rec.srows) return nullptr;
    return &a->rec;
}

int adjoint_checkpoint(wv_solver *s, int t)
{
    AdjointState *a = s->adj;
    if (!a || !a->every || t % a->every || !a->valid_run) return 0;
    if (a->store) {
        a->have[t / a->every] = 1;  // the record replaces the checkpoint
        return 0;
    }
    const int k = t / a->every;
    if (k >= a->nck) return 0;
    for (int f = 0; f < 5; f++) {
        if (a->fp16) {
            unsigned short *h = (unsigned short *)a->ck_fields[k] + (size_t)f * a->fld_n;
            wv::launch_ckpt_pack(s->field[0][f], h, a->ck_max[k] + (size_t)f * s->cfg.nshots, s->L.plane,
                                 s->cfg.nshots, s->st);
        } else {
            float *d = (float *)a->ck_fields[k] + (size_t)f * a->fld_n;
            HIP_TRY(hipMemcpyAsync(d, s->field[0][f], a->fld_n * sizeof(float), hipMemcpyDeviceToDevice, s->st));
        }
    }
    if (psi_total(s)) {
        float *p[8];
        size_t n[8];
        psi_list(s, p, n);
        float *d = a->ck_psi[k];
        for (int q = 0; q < 8; q++) {
            HIP_TRY(hipMemcpyAsync(d, p[q], n[q] * sizeof(float), hipMemcpyDeviceToDevice, s->st));
            d += n[q];
        }
    }
    HIP_TRY(hipGetLastError());
    a->have[k] = 1;
    return 0;
}

void adjoint_run_begin(wv_solver *s, int t_start)
{
    AdjointState *a = s->adj;
    if (!a || !a->every) return;
    if (t_start == 0) std::fill(a->have.begin(), a->have.end(), 0);
    a->valid_run = t_start == 0 || a->next_t == t_start;
    a->forward_done = false;
    a->grad_ready = false;
}

void adjoint_run_done(wv_solver *s, int t_end)
{
    AdjointState *a = s->adj;
    if (!a || !a->every) return;
    a->next_t = a->valid_run ? t_end : -1;
    a->forward_done = a->valid_run && t_end == s->cfg.NT;
    a->valid_run = false;
}

extern "C" {

int wv_set_checkpointing(wv_solver *s, int every, int fp16)
{
    if (!s || !s->inited) return fail(WV_ESTATE, "wv_set_checkpointing before wv_init");
    if (s->e3) {
        const int rc = s->e3->set_checkpointing(every, fp16, s->adj ? s->adj->store_gib : -1.0);
        return rc ? fail(rc, s->e3->error) : WV_OK;
    }
    if (s->dim != 2 || s->physics != wv::ELASTIC || s->plain || s->visco)
        return fail(WV_EINVAL, "the adjoint is provided for the tiled 2D elastic solver");
    if (s->cfg.free_surface || s->cfg.damping_boundary == 1)
        return fail(WV_EINVAL, "the adjoint supports CPML or no absorbing boundary, without free surface");
    if (every < 0 || every > s->cfg.NT) return fail(WV_EINVAL, "checkpoint interval out of range");
    // the adjoint kernels address one field plane with 32-bit byte offsets
    if (!wv::strip2d_fits(s->L)) return fail(WV_EINVAL, "grid too large for the adjoint kernels (field plane of 2 GiB or more)");
    if (!s->adj) s->adj = new AdjointState();
    AdjointState *a = s->adj;
    HIP_TRY(hipStreamSynchronize(s->st));
    if (every && every == a->every && (fp16 != 0) == a->fp16 && s->variant != V_SINGLE) {
        // same layout: keep the checkpoint buffers
        a->forward_done = a->grad_ready = false;
        a->next_t = -1;
        std::fill(a->have.begin(), a->have.end(), 0);
        return wv_reset(s);
    }
    free_all(a->ck_allocs);
    a->ck_fields.clear();
    a->ck_max.clear();
    a->ck_psi.clear();
    a->store = false;
    a->rec = wv::Rec2D{};
    a->every = every;
    a->fp16 = fp16 != 0;
    a->forward_done = false;
    a->grad_ready = false;
    a->next_t = -1;
    if (!every) return WV_OK;
    // the recomputation needs a two-launch kernel pair (tiled or row-strip: both record
    // the derivative sums); the single-pass kernel is replaced by the tiled pair
    if (s->variant == V_SINGLE) {
        s->variant = V_PHASED;
        s->tile_x = wv::TILE_X;
        s->tile_y = wv::TILE_Y;
        s->cur = 0;
        s->acq_dirty = true;
    }
    s->drop_graph();
    s->cfg.steps_per_graph = every;
    a->fld_n = (size_t)s->L.plane * s->cfg.nshots;
    a->nck = (s->cfg.NT + every - 1) / every;
    a->have.assign(a->nck, 0);
    // record every step instead when it fits (row-strip kernels only)
    if (s->variant == V_STRIP) {
        const int nbx = (s->L.NX + wv::TILE_X - 1) / wv::TILE_X, srows = s->tile_y / 4;
        const int nty = (s->L.NY + srows - 1) / srows;
        const size_t fss = (size_t)s->cfg.NT * 5 * s->cfg.nshots;  // (step, field, shot) planes
        const size_t strips = fss * nty * nbx, words = strips * (srows / 2) * wv::TILE_X;
        const double need = (double)words * 4 + (double)strips * 4;
        size_t free_b = 0, total_b = 0;
        HIP_TRY(hipMemGetInfo(&free_b, &total_b));
        // what wv_run_adjoint allocates later must still fit next to the record: the five
        // adjoint fields, two sets of CPML memory variables, the FP64 gradient planes, the
        // transposed tap tables and the time-major adjoint sources
        const double later = (double)a->fld_n * 5 * 4 + 2.0 * (double)psi_total(s) * 4 +
                             (double)s->L.plane * 5 * 8 + (double)s->cfg.NT * (s->rec_total + 1) * 4 + 64.0e6;
        const double avail = std::max(0.0, (double)free_b - later);
        const double budget = a->store_gib < 0 ? 0.75 * avail : std::min(a->store_gib * 1073741824.0, avail);
        if (need <= budget) {
            // every value is written by the forward run before the adjoint reads it: no clearing
            void *h = nullptr, *sc = nullptr;
            if (hipMalloc(&h, words * 4) == hipSuccess) a->ck_allocs.push_back(h);
            if (h && hipMalloc(&sc, strips * 4) == hipSuccess) a->ck_allocs.push_back(sc);
            if (h && sc) {
                a->store = true;
                a->rec.h = (unsigned *)h;
                a->rec.scale = (float *)sc;
                a->rec.nbx = nbx;
                a->rec.nty = nty;
                a->rec.srows = srows;
                return wv_reset(s);
            }
            (void)hipGetLastError();  // did not fit after all: checkpoints and recomputation
            free_all(a->ck_allocs);
        }
    }
    const size_t np = psi_total(s);
    for (int k = 0; k < a->nck; k++) {
        int rc;
        if (a->fp16) {
            unsigned short *h;
            unsigned *m;
            if ((rc = adalloc(a->ck_allocs, &h, 5 * a->fld_n, s->st)) ||
                (rc = adalloc(a->ck_allocs, &m, (size_t)5 * s->cfg.nshots, s->st)))
                return rc;
            a->ck_fields.push_back(h);
            a->ck_max.push_back(m);
        } else {
            float *f;
            if ((rc = adalloc(a->ck_allocs, &f, 5 * a->fld_n, s->st))) return rc;
            a->ck_fields.push_back(f);
            a->ck_max.push_back(nullptr);
        }
        float *p = nullptr;
        if (np && (rc = adalloc(a->ck_allocs, &p, np, s->st))) return rc;
        a->ck_psi.push_back(p);
    }
    return wv_reset(s);
}

int wv_set_adjoint_store(wv_solver *s, double max_gib)
{
    if (!s || !s->inited) return fail(WV_ESTATE, "wv_set_adjoint_store before wv_init");
    if (!(max_gib == max_gib)) return fail(WV_EINVAL, "bad budget");
    if (!s->adj) s->adj = new AdjointState();
    s->adj->store_gib = max_gib;
    s->adj->every = 0;  // the next wv_set_checkpointing chooses the memory plan again
    return WV_OK;
}

int wv_set_adjoint_source(wv_solver *s, int shot, const float *adj)
{
    if (s && s->inited && s->e3) {
        if (shot < 0 || shot >= s->cfg.nshots || !adj) return fail(WV_EINVAL, "bad shot or adjoint source");
        const int rc = s->e3->set_adjoint_source(shot, adj);
        return rc ? fail(rc, s->e3->error) : WV_OK;
    }
    if (!s || !s->inited || !s->adj || !s->adj->every) return fail(WV_ESTATE, "checkpointing is not enabled");
    if (shot < 0 || shot >= s->cfg.nshots || !adj) return fail(WV_EINVAL, "bad shot or adjoint source");
    AdjointState *a = s->adj;
    a->adj_host.resize(s->cfg.nshots);
    const size_t n = s->shots[shot].ridx.size() * (size_t)s->cfg.NT;
    a->adj_host[shot].assign(adj, adj + n);
    return WV_OK;
}

int wv_run_adjoint(wv_solver *s)
{
    if (s && s->inited && s->e3) {
        HIP_TRY(hipEventRecord(s->ev0, s->st));
        const int rc = s->e3->run_adjoint();
        if (rc) return fail(rc, s->e3->error);
        HIP_TRY(hipEventRecord(s->ev1, s->st));
        s->ran = true;
        return WV_OK;
    }
    if (!s || !s->inited || !s->adj || !s->adj->every) return fail(WV_ESTATE, "checkpointing is not enabled");
    AdjointState *a = s->adj;
    if (!a->forward_done) return fail(WV_ESTATE, "wv_run_adjoint needs a checkpointed forward run over [0, NT)");
    for (int k = 0; k < a->nck; k++)
        if (!a->have[k]) return fail(WV_ESTATE, "missing checkpoint");
    if (a->store && (s->variant != V_STRIP || s->tile_y / 4 != a->rec.srows))
        return fail(WV_ESTATE, "the kernel variant changed after wv_set_checkpointing: repeat it and the forward run");
    int rc;
    if ((rc = s->sync_acquisition())) return rc;
    const wv::Layout2D &L = s->L;
    const int NT = s->cfg.NT, K = a->every, ns = s->cfg.nshots, H = s->H;
    const size_t plane = (size_t)L.plane;
    const size_t np = psi_total(s);
    hipStream_t st = s->st;

    // buffers (allocated once per handle)
    if (!a->lam[0]) {
        for (int f = 0; f < 5; f++)
            if ((rc = adalloc(a->allocs, &a->lam[f], a->fld_n, st))) return rc;
        for (int f = 0; f < 5; f++)
            if ((rc = adalloc(a->allocs, &a->grad[f], plane, st))) return rc;
        if (np) {
            float *p[8];
            size_t n[8];
            psi_list(s, p, n);
            for (int b = 0; b < 2; b++)
                for (int q = 0; q < 8; q++)
                    if ((rc = adalloc(a->allocs, &a->mu[b][q], n[q], st))) return rc;
        }
        // transposed derivative tables
        const int N = s->cfg.spatial_fd_order;
        const int NXp = round_up(s->cfg.NX, 64) + 64, NYp = round_up(s->cfg.NY, 32) + 32;
        const bool stencil = s->cfg.ops_mode == WV_OPS_STENCIL;
        const float DH = (float)s->cfg.DH, DT = (float)s->cfg.DT;
        std::vector<float> t[4];
        wv::build_taps(t[0], s->cfg.NX, NXp, N, H, true, stencil, DH, DT);
        wv::build_taps(t[1], s->cfg.NX, NXp, N, H, false, stencil, DH, DT);
        wv::build_taps(t[2], s->cfg.NY, NYp, N, H, true, stencil, DH, DT);
        wv::build_taps(t[3], s->cfg.NY, NYp, N, H, false, stencil, DH, DT);
        for (int q = 0; q < 4; q++) {
            std::vector<float> tt = wv::transpose_taps(t[q], H);
            std::vector<float> ext((size_t)(tt.size() / (2 * H + 1) + 2 * ORG) * (2 * H + 1), 0.f);
            std::copy(tt.begin(), tt.end(), ext.begin() + (size_t)ORG * (2 * H + 1));
            float *base;
            if ((rc = adalloc(a->allocs, &base, ext.size(), st))) return rc;
            HIP_TRY(hipMemcpyAsync(base, ext.data(), ext.size() * sizeof(float), hipMemcpyHostToDevice, st));
            HIP_TRY(hipStreamSynchronize(st));  // ext is freed at the end of this iteration
            a->tapT[q] = base + (size_t)ORG * (2 * H + 1);
        }
    }
    if (!a->store && a->seg_slots < K) {
        if ((rc = adalloc(a->allocs, &a->segH, (size_t)K * 5 * a->fld_n, st))) return rc;
        a->seg_slots = K;
    }
    const size_t nsb = (size_t)ns * s->nbx * s->nby + 1;
    if (a->zero_n < nsb) {
        if ((rc = adalloc(a->allocs, &a->zero_ptr, nsb, st))) return rc;
        a->zero_n = nsb;
    }
    // adjoint sources, time-major [NT][rec_total]
    const size_t asn = (size_t)NT * std::max(s->rec_total, 1);
    if (a->adj_src_n < asn) {
        if ((rc = adalloc(a->allocs, &a->adj_src, asn, st))) return rc;
        a->adj_src_n = asn;
    }
    {
        std::vector<float> h(asn, 0.f);
        a->adj_host.resize(ns);
        for (int sh = 0; sh < ns; sh++) {
            const int nr = (int)s->shots[sh].ridx.size(), r0 = s->rec_ptr_h[sh];
            if (a->adj_host[sh].size() != (size_t)nr * NT) continue;
            for (int r = 0; r < nr; r++)
                for (int t = 0; t < NT; t++) h[(size_t)t * s->rec_total + r0 + r] = a->adj_host[sh][(size_t)r * NT + t];
        }
        HIP_TRY(hipMemcpyAsync(a->adj_src, h.data(), asn * sizeof(float), hipMemcpyHostToDevice, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    for (int f = 0; f < 5; f++) {
        HIP_TRY(hipMemsetAsync(a->lam[f], 0, a->fld_n * sizeof(float), st));
        HIP_TRY(hipMemsetAsync(a->grad[f], 0, plane * sizeof(double), st));
    }
    float *psi[8];
    size_t pn[8];
    psi_list(s, psi, pn);
    if (np)
        for (int b = 0; b < 2; b++)
            for (int q = 0; q < 8; q++) HIP_TRY(hipMemsetAsync(a->mu[b][q], 0, pn[q] * sizeof(float), st));

    const wv::Coef2D C = s->coef2d();
    const wv::Taps T = s->taps();
    const wv::TapsT TT{a->tapT[0], a->tapT[1], a->tapT[2], a->tapT[3]};
    const wv::Cpml2D P = s->cpml2d(0);
    const wv::Fields2D F = s->fields2d(0);
    const wv::Acq Aq = s->acq();
    wv::Acq Arec = Aq;  // recompute: sources as in the forward run, no receivers recorded
    Arec.brec_ptr = a->zero_ptr;
    const wv::AdjFields AF{a->lam[0], a->lam[1], a->lam[2], a->lam[3], a->lam[4]};
    wv::Grad G;
    for (int f = 0; f < 5; f++) G.g[f] = a->grad[f];
    int cv = 0, cs = 0;  // ping-pong buffers of the velocity-set / stress-set adjoint memory variables
    const int order = s->cfg.spatial_fd_order;
    KernelTimer kt(s);
    if (s->ktiming && (rc = kt.init(64))) return rc;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventRecord(s->ev0, st));

    for (int k = a->nck - 1; k >= 0; k--) {
        const int t0 = k * K, n = std::min(K, NT - t0);
        if (!a->store) {
            // restore the checkpoint
            for (int f = 0; f < 5; f++) {
                if (a->fp16)
                    wv::launch_ckpt_unpack((const unsigned short *)a->ck_fields[k] + (size_t)f * a->fld_n,
                                           s->field[0][f], a->ck_max[k] + (size_t)f * ns, L.plane, ns, st);
                else
                    HIP_TRY(hipMemcpyAsync(s->field[0][f], (const float *)a->ck_fields[k] + (size_t)f * a->fld_n,
                                           a->fld_n * sizeof(float), hipMemcpyDeviceToDevice, st));
            }
            if (np) {
                const float *src = a->ck_psi[k];
                for (int q = 0; q < 8; q++) {
                    HIP_TRY(hipMemcpyAsync(psi[q], src, pn[q] * sizeof(float), hipMemcpyDeviceToDevice, st));
                    src += pn[q];
                }
            }
            wv::launch_set(s->d_tbase, t0, st);
            // recompute the segment; the forward kernels record what the adjoint steps read
            for (int j = 0; j < n; j++) {
                wv::Snap2D snap{};
                for (int f = 0; f < 5; f++) snap.h[f] = a->segH + ((size_t)j * 5 + f) * a->fld_n;
                const bool strip = s->variant == V_STRIP;
                if ((rc = kt.next(3, &e0, &e1))) return rc;
                if (strip)
                    wv::launch_strip2d(order, s->tile_y, L, ns, F, C, T, P, Arec, j, j > 0, st, 0, e0, e1, &snap);
                else
                    wv::launch_step2d(wv::ELASTIC, order, s->tile_y, L, ns, F, C, T, P, Arec, j, j > 0, st, 0, e0, e1, &snap);
                if ((rc = kt.next(4, &e0, &e1))) return rc;
                if (strip)
                    wv::launch_strip2d(order, s->tile_y, L, ns, F, C, T, P, Arec, j, 0, st, 1, e0, e1, &snap);
                else
                    wv::launch_step2d(wv::ELASTIC, order, s->tile_y, L, ns, F, C, T, P, Arec, j, 0, st, 1, e0, e1, &snap);
            }
        }
        // adjoint steps of the segment, last to first
        for (int j = n - 1; j >= 0; j--) {
            const int t = t0 + j;
            wv::FwdSnap SV{}, SS{};  // stress-phase and velocity-phase sums of step t
            if (a->store) {
                SV.hs = SS.hs = a->rec.h;
                SV.ss = SS.ss = a->rec.scale;
                SV.t = SS.t = t;
                SV.nbx = SS.nbx = a->rec.nbx;
                SV.nty = SS.nty = a->rec.nty;
                SV.srows = SS.srows = a->rec.srows;
                SV.nshots = SS.nshots = ns;
            } else {
                for (int f = 0; f < 5; f++) SV.h[f] = SS.h[f] = a->segH + ((size_t)j * 5 + f) * a->fld_n;
            }
            wv::launch_adj_inject(L, AF, Aq, a->adj_src, t, st);
            wv::AdjPsi M{};
            for (int q = 0; q < 4; q++) {
                // x-frame arrays are psi slots 0,1 (psx0, psx1) and 4,5 (psx2, psx3);
                // y-frame arrays are slots 2,3 and 6,7
                const int sx = q < 2 ? q : 2 + q, sy = q < 2 ? 2 + q : 4 + q;
                const int bx = q < 2 ? cv : cs, by = q < 2 ? cv : cs;
                M.mx[q] = np ? a->mu[bx][sx] : nullptr;
                M.my[q] = np ? a->mu[by][sy] : nullptr;
                M.mxn[q] = np ? a->mu[bx ^ 1][sx] : nullptr;
                M.myn[q] = np ? a->mu[by ^ 1][sy] : nullptr;
            }
            if ((rc = kt.next(5, &e0, &e1))) return rc;
            wv::launch_adj_stress(order, s->tile_y == 16 ? 16 : 8, L, ns, AF, M, SV, C, T, TT, P, G, st, e0, e1);
            cs ^= 1;
            for (int q = 2; q < 4; q++) {
                const int sx = 2 + q, sy = 4 + q;
                M.mx[q] = np ? a->mu[cs][sx] : nullptr;
                M.my[q] = np ? a->mu[cs][sy] : nullptr;
                M.mxn[q] = np ? a->mu[cs ^ 1][sx] : nullptr;
                M.myn[q] = np ? a->mu[cs ^ 1][sy] : nullptr;
            }
            if ((rc = kt.next(6, &e0, &e1))) return rc;
            wv::launch_adj_velocity(order, s->tile_y == 16 ? 16 : 8, L, ns, AF, M, SS, C, T, TT, P, G, st, e0, e1);
            cv ^= 1;
        }
    }
    HIP_TRY(hipGetLastError());
    if ((rc = kt.flush())) return rc;
    HIP_TRY(hipEventRecord(s->ev1, st));
    s->ran = true;
    a->grad_ready = true;
    return WV_OK;
}

int wv_adjoint_plan(wv_solver *s, int *k2, int *nsub, int *on_host)
{
    if (!s || !s->inited || !s->e3) return fail(WV_EINVAL, "the memory plan is reported for 3D elastic handles");
    s->e3->adjoint_plan(k2, nsub, on_host);
    return WV_OK;
}

int wv_get_gradient(wv_solver *s, int param, double *out)
{
    if (s && s->inited && s->e3) {
        const int rc = s->e3->get_gradient(param, out);
        return rc ? fail(rc, s->e3->error) : WV_OK;
    }
    if (!s || !s->inited || !s->adj || !s->adj->grad_ready) return fail(WV_ESTATE, "no gradient computed");
    if (param < 0 || param > 7 || !out) return fail(WV_EINVAL, "param must be 0..7");
    AdjointState *a = s->adj;
    const wv::Layout2D &L = s->L;
    const int NX = L.NX, NY = L.NY;
    const size_t n = (size_t)NX * NY;
    HIP_TRY(hipStreamSynchronize(s->st));
    std::vector<double> g[5];
    std::vector<double> h((size_t)L.plane);
    for (int f = 0; f < 5; f++) {
        HIP_TRY(hipMemcpy(h.data(), a->grad[f], h.size() * sizeof(double), hipMemcpyDeviceToHost));
        g[f].resize(n);
        for (int y = 0; y < NY; y++)
            for (int x = 0; x < NX; x++) g[f][(size_t)y * NX + x] = h[(size_t)L.at(x, y)];
    }
    if (param >= 3) {
        std::memcpy(out, g[param - 3].data(), n * sizeof(double));
        return WV_OK;
    }
    // chain rule: coefficients (Modelparameter.cpp:336-639 rules) -> vp, vs, rho
    const std::vector<float> &vp = s->model_h[0], &vs = s->model_h[1], &rho = s->model_h[2];
    if (vp.size() != n) return fail(WV_ESTATE, "model not retained");
    auto gi = [&](int x, int y) { return (size_t)y * NX + x; };
    std::vector<double> mu(n), gmu(g[3]), gvp(n, 0.0), gvs(n, 0.0), grho(n, 0.0);
    for (size_t i = 0; i < n; i++) mu[i] = std::max((double)rho[i] * vs[i] * vs[i], 1.0);
    for (int y = 0; y < NY; y++)
        for (int x = 0; x < NX; x++) {
            const size_t j = gi(x, y);
            const bool inx = x + 1 < NX, iny = y + 1 < NY;
            size_t pts[4];
            int npt = 0;
            double w;
            pts[npt++] = j;
            if (inx && iny) {
                pts[npt++] = gi(x + 1, y);
                pts[npt++] = gi(x, y + 1);
                pts[npt++] = gi(x + 1, y + 1);
                w = 0.25;
            } else if (inx) {
                pts[npt++] = gi(x + 1, y);
                w = 0.5;
            } else if (iny) {
                pts[npt++] = gi(x, y + 1);
                w = 0.5;
            } else {
                w = 1.0;
            }
            double avg = 0;
            for (int q = 0; q < npt; q++) avg += 1.0 / mu[pts[q]];
            const double mxy = 1.0 / (w * avg);
            if (mxy < 4.0) continue;  // mu_xy set to 0: no dependence
            for (int q = 0; q < npt; q++) gmu[pts[q]] += g[4][j] * mxy * mxy * w / (mu[pts[q]] * mu[pts[q]]);
        }
    for (int y = 0; y < NY; y++)
        for (int x = 0; x < NX; x++) {
            const size_t i = gi(x, y);
            const double r = rho[i], p = vp[i], v = vs[i];
            gvp[i] = g[2][i] * 2.0 * r * p;
            grho[i] += g[2][i] * p * p;
            if (r * v * v >= 1.0) {
                gvs[i] = gmu[i] * 2.0 * r * v;
                grho[i] += gmu[i] * v * v;
            }
            if (x + 1 < NX) {
                const double ix = 2.0 / (r + (double)rho[gi(x + 1, y)]);
                grho[i] += g[0][i] * (-0.5 * ix * ix);
                grho[gi(x + 1, y)] += g[0][i] * (-0.5 * ix * ix);
            } else {
                grho[i] += g[0][i] * (-1.0 / (r * r));
            }
            if (y + 1 < NY) {
                const double iy = 2.0 / (r + (double)rho[gi(x, y + 1)]);
                grho[i] += g[1][i] * (-0.5 * iy * iy);
                grho[gi(x, y + 1)] += g[1][i] * (-0.5 * iy * iy);
            } else {
                grho[i] += g[1][i] * (-1.0 / (r * r));
            }
        }
    const std::vector<double> &res = param == 0 ? gvp : (param == 1 ? gvs : grho);
    std::memcpy(out, res.data(), n * sizeof(double));
    return WV_OK;
}

}  // extern "C"
