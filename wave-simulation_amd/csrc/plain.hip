// Reference-sequence GPU path (see plain.hpp).  Operation order follows
// ForwardSolver2Dacoustic.cpp:91-194, ForwardSolver2Delastic.cpp:117-294 and
// ForwardSolver2Dtmem.cpp:108-171 line by line; each LAMA call is one launch.
#include "plain.hpp"

#include <algorithm>
#include <cmath>

#include "prep.hpp"
#include "wave_internal.hpp"

namespace wv {

namespace {

constexpr int TPB = 256;
inline int blocks(long long n) { return (int)std::min<long long>((n + TPB - 1) / TPB, 65535); }

__global__ void k_spmv(int n, const int *__restrict__ rp, const int *__restrict__ col, const float *__restrict__ val,
                       const float *__restrict__ x, float *__restrict__ y)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float t = 0.f;
        for (int k = rp[i]; k < rp[i + 1]; k++) t += val[k] * x[col[k]];
        y[i] = t;
    }
}

// dense vector operations
enum VOp { V_COPY, V_ADD, V_SUB, V_MUL, V_SUB2, V_NEG };
__global__ void k_vec(int n, int op, float *__restrict__ d, const float *__restrict__ s)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        switch (op) {
        case V_COPY: d[i] = s[i]; break;
        case V_ADD: d[i] = d[i] + s[i]; break;
        case V_SUB: d[i] = d[i] - s[i]; break;
        case V_MUL: d[i] = d[i] * s[i]; break;
        case V_SUB2: d[i] = d[i] - 2.f * s[i]; break;
        case V_NEG: d[i] = -s[i]; break;
        }
    }
}
// d = a * b (elementwise, new vector)
__global__ void k_vmul3(int n, float *__restrict__ d, const float *__restrict__ a, const float *__restrict__ b)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) d[i] = a[i] * b[i];
}
__global__ void k_vadd3(int n, float *__restrict__ d, const float *__restrict__ a, const float *__restrict__ b)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) d[i] = a[i] + b[i];
}

// sparse-vector passes over a pattern
__global__ void k_sp_copy(int nnz, float *__restrict__ d, const float *__restrict__ s)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += gridDim.x * blockDim.x) d[k] = s[k];
}
__global__ void k_sp_mul(int nnz, float *__restrict__ d, const float *__restrict__ s)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += gridDim.x * blockDim.x) d[k] = d[k] * s[k];
}
__global__ void k_sp_mul_dense(int nnz, float *__restrict__ d, const int *__restrict__ idx, const float *__restrict__ v)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += gridDim.x * blockDim.x) d[k] = d[k] * v[idx[k]];
}
__global__ void k_sp_add(int nnz, float *__restrict__ d, const float *__restrict__ s)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += gridDim.x * blockDim.x) d[k] = d[k] + s[k];
}
// dense[idx] op= sparse
__global__ void k_dense_add_sp(int nnz, float *__restrict__ v, const int *__restrict__ idx, const float *__restrict__ s,
                               float sign)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += gridDim.x * blockDim.x)
        v[idx[k]] = v[idx[k]] + sign * s[k];
}
__global__ void k_dense_mul_sp(int nnz, float *__restrict__ v, const int *__restrict__ idx, const float *__restrict__ s)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += gridDim.x * blockDim.x)
        v[idx[k]] = v[idx[k]] * s[k];
}
// DenseMatrix::getColumn / setColumn, DenseVector::scatter / gatherInto
__global__ void k_get_column(int n, int NT, int t, const float *__restrict__ m, float *__restrict__ col)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) col[k] = m[(long long)k * NT + t];
}
__global__ void k_set_column(int n, int NT, int t, float *__restrict__ m, const float *__restrict__ col)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) m[(long long)k * NT + t] = col[k];
}
__global__ void k_scatter_add(int n, float *__restrict__ v, const int *__restrict__ idx, const float *__restrict__ s)
{
    // sequential: sources sharing a cell add in list order
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (int k = 0; k < n; k++) v[idx[k]] += s[k];
}
__global__ void k_gather(int n, float *__restrict__ d, const float *__restrict__ v, const int *__restrict__ idx, int add)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
        d[k] = add ? d[k] + v[idx[k]] : v[idx[k]];
}
__global__ void k_scale(int n, float *__restrict__ d, float s)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) d[k] = d[k] * s;
}

}  // namespace

PlainSolver::~PlainSolver()
{
    if (st_) (void)hipStreamSynchronize(st_);
    for (void *p : acq_allocs_) (void)hipFree(p);
    for (void *p : allocs_) (void)hipFree(p);
}

int PlainSolver::alloc(void **p, size_t bytes, std::vector<void *> &list)
{
    if (bytes == 0) bytes = 4;
    if (hipMalloc(p, bytes) != hipSuccess) {
        error = "hipMalloc failed";
        return WV_ENOMEM;
    }
    (void)hipMemsetAsync(*p, 0, bytes, st_);
    list.push_back(*p);
    return 0;
}

// assemble a CSR matrix from per-position taps along axis 0 (x) or 1 (y)
int PlainSolver::upload_csr(DevCsr &d, const std::vector<float> &taps, int H, int axis)
{
    const int NX = cfg_.NX, NY = cfg_.NY, K = 2 * H + 1;
    std::vector<int> rp(1, 0), col;
    std::vector<float> val;
    for (int y = 0; y < NY; y++)
        for (int x = 0; x < NX; x++) {
            const int p = axis == 0 ? x : y;
            for (int k = 0; k < K; k++) {
                const float w = taps[(size_t)p * K + k];
                if (w == 0.f) continue;
                const int q = p + k - H;
                col.push_back(axis == 0 ? q + y * NX : x + q * NX);
                val.push_back(w);
            }
            rp.push_back((int)col.size());
        }
    d.n = n_;
    d.nnz = (int)col.size();
    int rc;
    if ((rc = alloc((void **)&d.rowptr, rp.size() * 4, allocs_)) || (rc = alloc((void **)&d.col, col.size() * 4, allocs_)) ||
        (rc = alloc((void **)&d.val, val.size() * 4, allocs_)))
        return rc;
    (void)hipMemcpyAsync(d.rowptr, rp.data(), rp.size() * 4, hipMemcpyHostToDevice, st_);
    (void)hipMemcpyAsync(d.col, col.data(), col.size() * 4, hipMemcpyHostToDevice, st_);
    (void)hipMemcpyAsync(d.val, val.data(), val.size() * 4, hipMemcpyHostToDevice, st_);
    return hipStreamSynchronize(st_) == hipSuccess ? 0 : WV_EDEVICE;
}

int PlainSolver::upload_sparse(DevSparse &d, const std::vector<int> &idx, const std::vector<float> &val)
{
    d.nnz = (int)idx.size();
    int rc;
    if (!d.idx && ((rc = alloc((void **)&d.idx, idx.size() * 4, allocs_)) ||
                   (rc = alloc((void **)&d.val, val.size() * 4, allocs_))))
        return rc;
    (void)hipMemcpyAsync(d.idx, idx.data(), idx.size() * 4, hipMemcpyHostToDevice, st_);
    (void)hipMemcpyAsync(d.val, val.data(), val.size() * 4, hipMemcpyHostToDevice, st_);
    return hipStreamSynchronize(st_) == hipSuccess ? 0 : WV_EDEVICE;
}

int PlainSolver::init(int physics, const wv_config &cfg, hipStream_t st)
{
    physics_ = physics;
    cfg_ = cfg;
    st_ = st;
    n_ = cfg.NX * cfg.NY;
    const int N = cfg.spatial_fd_order, H = N / 2;
    const bool stencil = cfg.ops_mode == WV_OPS_STENCIL;
    const float DH = (float)cfg.DH, DT = (float)cfg.DT;
    std::vector<float> t;
    int rc;
    build_taps(t, cfg.NX, cfg.NX, N, H, true, stencil, DH, DT);
    if ((rc = upload_csr(Dxf_, t, H, 0))) return rc;
    build_taps(t, cfg.NX, cfg.NX, N, H, false, stencil, DH, DT);
    if ((rc = upload_csr(Dxb_, t, H, 0))) return rc;
    build_taps(t, cfg.NY, cfg.NY, N, H, true, stencil, DH, DT);
    if ((rc = upload_csr(Dyf_, t, H, 1))) return rc;
    build_taps(t, cfg.NY, cfg.NY, N, H, false, stencil, DH, DT);
    if ((rc = upload_csr(Dyb_, t, H, 1))) return rc;
    if (cfg.free_surface == 1) {
        build_taps_free_surface(t, cfg.NY, cfg.NY, N, H, true, DH, DT);
        if ((rc = upload_csr(DyfFS_, t, H, 1))) return rc;
        build_taps_free_surface(t, cfg.NY, cfg.NY, N, H, false, DH, DT);
        if ((rc = upload_csr(DybFS_, t, H, 1))) return rc;
    }
    const int nf = physics == ELASTIC ? 5 : 3;
    for (int k = 0; k < nf; k++)
        if ((rc = alloc((void **)&f_[k], (size_t)n_ * 4, allocs_))) return rc;
    if ((rc = alloc((void **)&upd_, (size_t)n_ * 4, allocs_)) || (rc = alloc((void **)&tmp_, (size_t)n_ * 4, allocs_)) ||
        (rc = alloc((void **)&vxx_, (size_t)n_ * 4, allocs_)) || (rc = alloc((void **)&vyy_, (size_t)n_ * 4, allocs_)))
        return rc;
    for (int k = 0; k < 5; k++)
        if ((rc = alloc((void **)&c_[k], (size_t)n_ * 4, allocs_))) return rc;

    if (cfg.damping_boundary == 2) {
        const int W = cfg.boundary_width;
        std::vector<float> a, b, ah, bh;
        std::vector<int> idx;
        // CPML2D.cpp:228-270 -- sparse vectors over the x and y frames
        for (int axis = 0; axis < 2; axis++) {
            const int L = axis == 0 ? cfg.NX : cfg.NY;
            cpml_axis(L, L, W, axis == 1 && cfg.free_surface != 0, cfg.npower, cfg.fc_cpml, cfg.vmax_cpml, cfg.DT,
                      cfg.DH, a, b, ah, bh, idx);
            std::vector<int> pat;
            std::vector<float> va, vb, vah, vbh;
            for (int y = 0; y < cfg.NY; y++)
                for (int x = 0; x < cfg.NX; x++) {
                    const int p = axis == 0 ? x : y;
                    if (idx[p] < 0) continue;
                    pat.push_back(x + y * cfg.NX);
                    va.push_back(a[p]);
                    vb.push_back(b[p]);
                    vah.push_back(ah[p]);
                    vbh.push_back(bh[p]);
                }
            DevSparse *A = axis == 0 ? &ax_ : &ay_, *B = axis == 0 ? &bx_ : &by_;
            DevSparse *AH = axis == 0 ? &axh_ : &ayh_, *BH = axis == 0 ? &bxh_ : &byh_;
            if ((rc = upload_sparse(*A, pat, va)) || (rc = upload_sparse(*B, pat, vb)) ||
                (rc = upload_sparse(*AH, pat, vah)) || (rc = upload_sparse(*BH, pat, vbh)))
                return rc;
        }
        const int maxnnz = std::max(ax_.nnz, ay_.nnz);
        for (int k = 0; k < 8; k++)
            if ((rc = alloc((void **)&psi_[k], (size_t)maxnnz * 4, allocs_))) return rc;
    }
    // scratch for sparse-vector temporaries (CPML frames, free-surface row)
    if ((rc = alloc((void **)&stemp_, (size_t)std::max({ax_.nnz, ay_.nnz, cfg.NX}) * 4, allocs_))) return rc;
    if (cfg.damping_boundary == 1) {
        const int W = cfg.boundary_width;
        std::vector<double> prof = abs_profile(W, cfg.damping_coeff);
        std::vector<int> pat;
        std::vector<float> val;
        for (int y = 0; y < cfg.NY; y++)
            for (int x = 0; x < cfg.NX; x++) {
                const int dx = std::min(x, cfg.NX - 1 - x), dy = std::min(y, cfg.NY - 1 - y), mn = std::min(dx, dy);
                int d = -1;
                if (cfg.free_surface == 0) {
                    if (mn < W) d = mn;
                } else if (y < W) {
                    if (dx < W) d = dx;
                } else if (mn < W) {
                    d = mn;
                }
                if (d < 0) continue;
                pat.push_back(x + y * cfg.NX);
                val.push_back((float)prof[d]);
            }
        if ((rc = upload_sparse(damp_, pat, val))) return rc;
    }
    if (cfg.free_surface == 1) {
        std::vector<int> pat;
        std::vector<float> zero;
        for (int x = 0; x < cfg.NX; x++) {
            pat.push_back(x);
            zero.push_back(0.f);
        }
        if ((rc = upload_sparse(fsZero_, pat, zero))) return rc;
    }
    return 0;
}

int PlainSolver::set_model(const float *m0, const float *m1, const float *m2)
{
    const int NX = cfg_.NX, NY = cfg_.NY;
    std::vector<std::vector<float>> c(5, std::vector<float>((size_t)n_, 0.f));
    auto gi = [&](int x, int y) { return (size_t)x + (size_t)y * NX; };
    // one-time coefficient preparation, as the fused path (FP64, rounded once)
    auto inv_avg = [&](const std::vector<double> &m, std::vector<float> &out, int ax) {
        for (int y = 0; y < NY; y++)
            for (int x = 0; x < NX; x++) {
                const bool last = ax == 0 ? (x + 1 >= NX) : (y + 1 >= NY);
                const double a = m[gi(x, y)];
                const double avg = last ? a : 0.5 * (a + m[ax == 0 ? gi(x + 1, y) : gi(x, y + 1)]);
                double v = 1.0 / avg;
                if (!std::isfinite(v)) v = 0.0;
                out[gi(x, y)] = (float)v;
            }
    };
    std::vector<double> rho(n_), M(n_), mu(n_);
    if (physics_ != TMEM) {
        for (int i = 0; i < n_; i++) {
            rho[i] = m2[i];
            M[i] = (double)m2[i] * m0[i] * m0[i];
            if (physics_ == ELASTIC) mu[i] = std::max((double)m2[i] * m1[i] * m1[i], 1.0);
            c[0][i] = (float)M[i];
        }
        if (physics_ == ACOUSTIC) {
            inv_avg(rho, c[1], 0);
            inv_avg(rho, c[2], 1);
        } else {
            for (int i = 0; i < n_; i++) c[1][i] = (float)mu[i];
            inv_avg(rho, c[2], 0);
            inv_avg(rho, c[3], 1);
            for (int y = 0; y < NY; y++)
                for (int x = 0; x < NX; x++) {
                    const bool inx = x + 1 < NX, iny = y + 1 < NY;
                    const double i00 = 1.0 / mu[gi(x, y)];
                    double avg;
                    if (inx && iny)
                        avg = 0.25 * (i00 + 1.0 / mu[gi(x + 1, y)] + 1.0 / mu[gi(x, y + 1)] + 1.0 / mu[gi(x + 1, y + 1)]);
                    else if (inx)
                        avg = 0.5 * (i00 + 1.0 / mu[gi(x + 1, y)]);
                    else if (iny)
                        avg = 0.5 * (i00 + 1.0 / mu[gi(x, y + 1)]);
                    else
                        avg = i00;
                    double v = 1.0 / avg;
                    if (v < 4.0) v = 0.0;
                    c[4][gi(x, y)] = (float)v;
                }
        }
    } else {
        const double mu0 = 1.2566370614e-6;
        const float eps0f = 8.8541878176e-12f;
        std::vector<double> perm(n_);
        for (int i = 0; i < n_; i++) {
            perm[i] = (double)m2[i] * mu0;
            em_ca_cb(m0[i] * eps0f, m1[i], (float)cfg_.DT, c[0][i], c[1][i]);
        }
        inv_avg(perm, c[2], 0);
        inv_avg(perm, c[3], 1);
    }
    for (int k = 0; k < 5; k++) (void)hipMemcpyAsync(c_[k], c[k].data(), (size_t)n_ * 4, hipMemcpyHostToDevice, st_);
    if (physics_ == ELASTIC && cfg_.free_surface == 1) {
        std::vector<int> pat;
        std::vector<float> h, v;
        for (int x = 0; x < NX; x++) {
            const double t = M[gi(x, 0)] - 2.0 * mu[gi(x, 0)];
            pat.push_back(x);
            v.push_back((float)t);
            h.push_back((float)(-(t * t) / M[gi(x, 0)]));
        }
        int rc;
        if ((rc = upload_sparse(fsH_, pat, h)) || (rc = upload_sparse(fsV_, pat, v))) return rc;
    }
    model_ok_ = true;
    return hipStreamSynchronize(st_) == hipSuccess ? 0 : WV_EDEVICE;
}

int PlainSolver::set_acquisition(const std::vector<long long> &sidx, const std::vector<int> &stype,
                                 const std::vector<float> &ssig, const std::vector<long long> &ridx,
                                 const std::vector<int> &rtype)
{
    (void)hipStreamSynchronize(st_);
    for (void *p : acq_allocs_) (void)hipFree(p);
    acq_allocs_.clear();
    const int NT = cfg_.NT;
    nrec_total_ = (int)ridx.size();
    for (int c = 1; c <= 4; c++) {
        src_[c] = Comp();
        rec_[c] = Comp();
        std::vector<int> si, ri;
        std::vector<float> sg;
        for (size_t k = 0; k < sidx.size(); k++)
            if (stype[k] == c) {
                si.push_back((int)sidx[k]);
                sg.insert(sg.end(), ssig.begin() + (long)k * NT, ssig.begin() + (long)(k + 1) * NT);
            }
        for (size_t k = 0; k < ridx.size(); k++)
            if (rtype[k] == c) {
                ri.push_back((int)ridx[k]);
                rec_[c].order.push_back((int)k);
            }
        int rc;
        src_[c].n = (int)si.size();
        rec_[c].n = (int)ri.size();
        if (src_[c].n) {
            if ((rc = alloc((void **)&src_[c].idx, si.size() * 4, acq_allocs_)) ||
                (rc = alloc((void **)&src_[c].sig, sg.size() * 4, acq_allocs_)) ||
                (rc = alloc((void **)&src_[c].col, si.size() * 4, acq_allocs_)))
                return rc;
            (void)hipMemcpyAsync(src_[c].idx, si.data(), si.size() * 4, hipMemcpyHostToDevice, st_);
            (void)hipMemcpyAsync(src_[c].sig, sg.data(), sg.size() * 4, hipMemcpyHostToDevice, st_);
        }
        if (rec_[c].n) {
            if ((rc = alloc((void **)&rec_[c].idx, ri.size() * 4, acq_allocs_)) ||
                (rc = alloc((void **)&rec_[c].col, ri.size() * 4, acq_allocs_)) ||
                (rc = alloc((void **)&rec_[c].seis, ri.size() * 4 * (size_t)NT, acq_allocs_)))
                return rc;
            (void)hipMemcpyAsync(rec_[c].idx, ri.data(), ri.size() * 4, hipMemcpyHostToDevice, st_);
        }
        // si, sg and ri are freed at the end of this iteration: their copies complete first
        if (hipStreamSynchronize(st_) != hipSuccess) return WV_EDEVICE;
    }
    return hipStreamSynchronize(st_) == hipSuccess ? 0 : WV_EDEVICE;
}

int PlainSolver::reset()
{
    const int nf = physics_ == ELASTIC ? 5 : 3;
    for (int k = 0; k < nf; k++) (void)hipMemsetAsync(f_[k], 0, (size_t)n_ * 4, st_);
    const int maxnnz = std::max(ax_.nnz, ay_.nnz);
    for (int k = 0; k < 8; k++)
        if (psi_[k]) (void)hipMemsetAsync(psi_[k], 0, (size_t)maxnnz * 4, st_);
    for (int c = 1; c <= 4; c++)
        if (rec_[c].n) (void)hipMemsetAsync(rec_[c].seis, 0, (size_t)rec_[c].n * cfg_.NT * 4, st_);
    return 0;
}

void PlainSolver::spmv(const DevCsr &A, const float *x, float *y)
{
    k_spmv<<<blocks(A.n), TPB, 0, st_>>>(A.n, A.rowptr, A.col, A.val, x, y);
}

// CPML.cpp:84-95: temp = a; psi *= b; temp *= vec; psi += temp; vec += psi
void PlainSolver::cpml(float *vec, float *psi, const DevSparse &a, const DevSparse &b)
{
    const int nnz = a.nnz;
    if (!nnz) return;
    k_sp_copy<<<blocks(nnz), TPB, 0, st_>>>(nnz, stemp_, a.val);
    k_sp_mul<<<blocks(nnz), TPB, 0, st_>>>(nnz, psi, b.val);
    k_sp_mul_dense<<<blocks(nnz), TPB, 0, st_>>>(nnz, stemp_, a.idx, vec);
    k_sp_add<<<blocks(nnz), TPB, 0, st_>>>(nnz, psi, stemp_);
    k_dense_add_sp<<<blocks(nnz), TPB, 0, st_>>>(nnz, vec, a.idx, psi, 1.f);
}

void PlainSolver::step(int t)
{
    const int n = n_, gb = blocks(n);
    const bool pml = cfg_.damping_boundary == 2;
    const bool fs = cfg_.free_surface == 1;
    auto V = [&](int op, float *d, const float *s) { k_vec<<<gb, TPB, 0, st_>>>(n, op, d, s); };
    if (physics_ == ELASTIC) {
        float *vx = f_[0], *vy = f_[1], *sxx = f_[2], *syy = f_[3], *sxy = f_[4];
        spmv(Dxf_, sxx, upd_);
        if (pml) cpml(upd_, psi_[0], axh_, bxh_);
        spmv(fs ? DybFS_ : Dyb_, sxy, tmp_);
        if (pml) cpml(tmp_, psi_[1], ay_, by_);
        V(V_ADD, upd_, tmp_);
        V(V_MUL, upd_, c_[2]);
        V(V_ADD, vx, upd_);
        spmv(Dxb_, sxy, upd_);
        if (pml) cpml(upd_, psi_[2], ax_, bx_);
        spmv(fs ? DyfFS_ : Dyf_, syy, tmp_);
        if (pml) cpml(tmp_, psi_[3], ayh_, byh_);
        V(V_ADD, upd_, tmp_);
        V(V_MUL, upd_, c_[3]);
        V(V_ADD, vy, upd_);
        spmv(Dxb_, vx, vxx_);
        spmv(Dyb_, vy, vyy_);
        if (pml) {
            cpml(vxx_, psi_[4], ax_, bx_);
            cpml(vyy_, psi_[5], ay_, by_);
        }
        V(V_COPY, upd_, vxx_);
        V(V_ADD, upd_, vyy_);
        V(V_MUL, upd_, c_[0]);
        V(V_ADD, sxx, upd_);
        V(V_ADD, syy, upd_);
        V(V_COPY, upd_, vyy_);
        V(V_MUL, upd_, c_[1]);
        V(V_SUB2, sxx, upd_);
        V(V_COPY, upd_, vxx_);
        V(V_MUL, upd_, c_[1]);
        V(V_SUB2, syy, upd_);
        spmv(Dyf_, vx, upd_);
        if (pml) cpml(upd_, psi_[6], ayh_, byh_);
        spmv(Dxf_, vy, tmp_);
        if (pml) cpml(tmp_, psi_[7], axh_, bxh_);
        V(V_ADD, upd_, tmp_);
        V(V_MUL, upd_, c_[4]);
        V(V_ADD, sxy, upd_);
        if (fs) {
            // FreeSurface2Delastic.cpp:14-45 then FreeSurface.cpp:14-21
            const int m = fsH_.nnz;
            k_sp_copy<<<blocks(m), TPB, 0, st_>>>(m, stemp_, fsH_.val);
            k_sp_mul_dense<<<blocks(m), TPB, 0, st_>>>(m, stemp_, fsH_.idx, vxx_);
            k_dense_add_sp<<<blocks(m), TPB, 0, st_>>>(m, sxx, fsH_.idx, stemp_, 1.f);
            k_sp_copy<<<blocks(m), TPB, 0, st_>>>(m, stemp_, fsV_.val);
            k_sp_mul_dense<<<blocks(m), TPB, 0, st_>>>(m, stemp_, fsV_.idx, vyy_);
            k_dense_add_sp<<<blocks(m), TPB, 0, st_>>>(m, sxx, fsV_.idx, stemp_, -1.f);
            k_dense_mul_sp<<<blocks(m), TPB, 0, st_>>>(m, syy, fsZero_.idx, fsZero_.val);
        }
        if (cfg_.damping_boundary == 1) {
            float *all[5] = {sxx, syy, sxy, vx, vy};
            for (float *v : all) k_dense_mul_sp<<<blocks(damp_.nnz), TPB, 0, st_>>>(damp_.nnz, v, damp_.idx, damp_.val);
        }
    } else if (physics_ == ACOUSTIC) {
        float *vx = f_[0], *vy = f_[1], *p = f_[2];
        spmv(Dxf_, p, upd_);
        if (pml) cpml(upd_, psi_[0], axh_, bxh_);
        V(V_MUL, upd_, c_[1]);
        V(V_ADD, vx, upd_);
        spmv(fs ? DyfFS_ : Dyf_, p, upd_);
        if (pml) cpml(upd_, psi_[1], ayh_, byh_);
        V(V_MUL, upd_, c_[2]);
        V(V_ADD, vy, upd_);
        spmv(Dxb_, vx, upd_);
        if (pml) cpml(upd_, psi_[2], ax_, bx_);
        spmv(Dyb_, vy, tmp_);
        if (pml) cpml(tmp_, psi_[3], ay_, by_);
        V(V_ADD, upd_, tmp_);
        V(V_MUL, upd_, c_[0]);
        V(V_ADD, p, upd_);
        if (cfg_.damping_boundary == 1) {
            float *all[3] = {p, vx, vy};
            for (float *v : all) k_dense_mul_sp<<<blocks(damp_.nnz), TPB, 0, st_>>>(damp_.nnz, v, damp_.idx, damp_.val);
        }
        if (fs) k_dense_mul_sp<<<blocks(fsZero_.nnz), TPB, 0, st_>>>(fsZero_.nnz, p, fsZero_.idx, fsZero_.val);
    } else {
        float *hx = f_[0], *hy = f_[1], *ez = f_[2];
        spmv(Dyf_, ez, upd_);
        if (pml) cpml(upd_, psi_[0], ayh_, byh_);
        V(V_MUL, upd_, c_[3]);
        V(V_SUB, hx, upd_);
        spmv(Dxf_, ez, tmp_);
        if (pml) cpml(tmp_, psi_[1], axh_, bxh_);
        V(V_NEG, upd_, tmp_);
        V(V_MUL, upd_, c_[2]);
        V(V_SUB, hy, upd_);
        spmv(Dxb_, hy, upd_);
        spmv(Dyb_, hx, tmp_);
        if (pml) {
            cpml(upd_, psi_[2], ax_, bx_);
            cpml(tmp_, psi_[3], ay_, by_);
        }
        V(V_SUB, upd_, tmp_);
        V(V_MUL, upd_, c_[1]);
        k_vmul3<<<gb, TPB, 0, st_>>>(n, tmp_, c_[0], ez);
        k_vadd3<<<gb, TPB, 0, st_>>>(n, ez, tmp_, upd_);
        if (cfg_.damping_boundary == 1) {
            float *all[3] = {ez, hx, hy};
            for (float *v : all) k_dense_mul_sp<<<blocks(damp_.nnz), TPB, 0, st_>>>(damp_.nnz, v, damp_.idx, damp_.val);
        }
    }
    // sources (SourceReceiverImplSeismic.cpp:89-105; SourceReceiverImpl.cpp:28-37)
    for (int c = 1; c <= 4; c++) {
        Comp &S = src_[c];
        if (!S.n) continue;
        k_get_column<<<blocks(S.n), TPB, 0, st_>>>(S.n, cfg_.NT, t, S.sig, S.col);
        if (physics_ == ELASTIC && c == 1) {
            k_scatter_add<<<1, 64, 0, st_>>>(S.n, f_[2], S.idx, S.col);
            k_scatter_add<<<1, 64, 0, st_>>>(S.n, f_[3], S.idx, S.col);
        } else {
            float *dst = (physics_ == TMEM || c == 1) ? f_[2] : f_[c - 2];
            k_scatter_add<<<1, 64, 0, st_>>>(S.n, dst, S.idx, S.col);
        }
    }
    // receivers (SourceReceiverImplSeismic.cpp:34-47; FDTD2Delastic.cpp:12-27)
    for (int c = 1; c <= 4; c++) {
        Comp &R = rec_[c];
        if (!R.n) continue;
        if (physics_ == ELASTIC && c == 1) {
            k_gather<<<blocks(R.n), TPB, 0, st_>>>(R.n, R.col, f_[2], R.idx, 0);
            k_gather<<<blocks(R.n), TPB, 0, st_>>>(R.n, R.col, f_[3], R.idx, 1);
            k_scale<<<blocks(R.n), TPB, 0, st_>>>(R.n, R.col, 0.5f);
        } else {
            const float *srcf = (physics_ == TMEM || c == 1) ? f_[2] : f_[c - 2];
            k_gather<<<blocks(R.n), TPB, 0, st_>>>(R.n, R.col, srcf, R.idx, 0);
        }
        k_set_column<<<blocks(R.n), TPB, 0, st_>>>(R.n, cfg_.NT, t, R.seis, R.col);
    }
}

int PlainSolver::run(int t0, int t1)
{
    if (!model_ok_) {
        error = "model not set";
        return WV_ESTATE;
    }
    for (int t = t0; t < t1; t++) step(t);
    if (hipGetLastError() != hipSuccess) {
        error = "kernel launch failed";
        return WV_EDEVICE;
    }
    return 0;
}

int PlainSolver::get_seismograms(float *out)
{
    if (hipStreamSynchronize(st_) != hipSuccess) {
        error = "stream error";
        return WV_EDEVICE;
    }
    const int NT = cfg_.NT;
    std::vector<float> h;
    for (int c = 1; c <= 4; c++) {
        Comp &R = rec_[c];
        if (!R.n) continue;
        h.resize((size_t)R.n * NT);
        (void)hipMemcpy(h.data(), R.seis, h.size() * 4, hipMemcpyDeviceToHost);
        for (int k = 0; k < R.n; k++)
            std::copy(h.begin() + (size_t)k * NT, h.begin() + (size_t)(k + 1) * NT, out + (size_t)R.order[k] * NT);
    }
    return 0;
}

void PlainSolver::buffers(wv::StateBuffers &out) const
{
    const int nf = physics_ == ELASTIC ? 5 : 3;
    for (int f = 0; f < nf; f++) out.add(f_[f], (long long)n_);
    for (int c = 0; c < 5; c++)
        if (rec_[c].n) out.add(rec_[c].seis, (long long)rec_[c].n * cfg_.NT);
}

int PlainSolver::field_io(int field, float *out, const float *in)
{
    const int nf = physics_ == ELASTIC ? 5 : 3;
    if (field < 0 || field >= nf) {
        error = "bad field index";
        return WV_EINVAL;
    }
    (void)hipStreamSynchronize(st_);
    if (out) (void)hipMemcpy(out, f_[field], (size_t)n_ * 4, hipMemcpyDeviceToHost);
    if (in) (void)hipMemcpy(f_[field], in, (size_t)n_ * 4, hipMemcpyHostToDevice);
    return 0;
}

}  // namespace wv
