// Reference-sequence GPU path: one HIP kernel per LAMA operation of the
// reference's run() (CSR SpMV per derivative, one dense-vector pass per vector
// operation, CPML and boundary conditions as sparse-vector passes over their
// pattern, source/receiver as column extract + scatter / gather + set column).
// It exists as the single-shot baseline the fused batched kernels are compared
// against, and as a second GPU restatement for parity checks.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/wave_abi.h"
#include "wave_internal.hpp"

namespace wv {

struct DevCsr {
    int n = 0, nnz = 0;
    int *rowptr = nullptr, *col = nullptr;
    float *val = nullptr;
};

struct DevSparse {  // LAMA SparseVector: pattern + values
    int nnz = 0;
    int *idx = nullptr;
    float *val = nullptr;
};

class PlainSolver {
  public:
    std::string error;
    ~PlainSolver();
    int init(int physics, const wv_config &cfg, hipStream_t st);
    int set_model(const float *m0, const float *m1, const float *m2);
    int set_acquisition(const std::vector<long long> &sidx, const std::vector<int> &stype,
                        const std::vector<float> &ssig, const std::vector<long long> &ridx,
                        const std::vector<int> &rtype);
    int reset();
    int run(int t0, int t1);
    int get_seismograms(float *out);
    int field_io(int field, float *out, const float *in);
    // device buffers of the state (wavefields, seismograms) for wv_check_finite
    void buffers(wv::StateBuffers &out) const;

  private:
    int physics_ = 0;
    wv_config cfg_{};
    hipStream_t st_ = nullptr;
    int n_ = 0;
    std::vector<void *> allocs_, acq_allocs_;
    DevCsr Dxf_, Dxb_, Dyf_, Dyb_, DyfFS_, DybFS_;
    float *f_[5] = {};
    float *upd_ = nullptr, *tmp_ = nullptr, *vxx_ = nullptr, *vyy_ = nullptr;
    float *c_[5] = {};
    // CPML: coefficient sparse vectors (pattern = frame) and memory variables
    DevSparse ax_, bx_, axh_, bxh_, ay_, by_, ayh_, byh_;
    float *psi_[8] = {};   // values on the pattern of the matching a/b vector
    float *stemp_ = nullptr;  // sparse temp
    DevSparse damp_;       // ABS (zero element 1)
    DevSparse fsH_, fsV_, fsZero_;
    bool model_ok_ = false;
    // acquisition, per component 1..4
    struct Comp {
        int n = 0;
        int *idx = nullptr;
        float *sig = nullptr;  // [n][NT]
        float *col = nullptr;  // temp column [n]
        float *seis = nullptr; // [n][NT]
        std::vector<int> order;  // position in the caller's list
    };
    Comp src_[5], rec_[5];
    int nrec_total_ = 0;

    int alloc(void **p, size_t bytes, std::vector<void *> &list);
    int upload_csr(DevCsr &d, const std::vector<float> &taps, int H, int axis);
    int upload_sparse(DevSparse &d, const std::vector<int> &idx, const std::vector<float> &val);
    void spmv(const DevCsr &A, const float *x, float *y);
    void cpml(float *vec, float *psi, const DevSparse &a, const DevSparse &b);
    void step(int t);
};

}  // namespace wv
