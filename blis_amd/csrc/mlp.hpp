// mlp.hpp — the tile selector's network (mlp.hip) and its host-side handle.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace blis_amd {

constexpr int GEMM_NSCHED = 3;                  // schedules per tile (select.hip)
constexpr int GEMM_NTILE = 14;                  // tiles per data type (gemm_tiles.hpp)
constexpr int GEMM_NCFG = GEMM_NTILE * GEMM_NSCHED;

constexpr int MLP_IN = 16;
constexpr int MLP_H = 64;
constexpr int MLP_OUT = GEMM_NCFG;
constexpr int MLP_NPARAM = MLP_H * MLP_IN + MLP_H + MLP_H * MLP_H + MLP_H + MLP_OUT * MLP_H + MLP_OUT;

// Host copy of a network for one shape at a time (the selector's product
// path: no device memory, no launch, no wait). Same arithmetic, in the same
// order, as the device forward pass, so both give the same scores.
struct MlpHost {
    std::vector<float> p;      // parameters (layout of MlpNet)
    std::vector<float> musd;   // mu[IN], sd[IN]
    bool ok() const { return !p.empty(); }
    // argmin over the scores of one raw feature row (first index on ties).
    int argmin(const float* feats) const;
};

// Read a weights file written by MlpNet::save.
bool mlp_read(const char* path, std::vector<float>& params, std::vector<float>& musd);

// Device-resident network: parameters, Adam moments, gradient and the
// feature normalisation (mu, sd), plus batch scratch buffers.
class MlpNet {
public:
    MlpNet();
    ~MlpNet();
    MlpNet(const MlpNet&) = delete;
    MlpNet& operator=(const MlpNet&) = delete;

    void init(unsigned seed);
    void set_params(const float* host_params, const float* host_musd);
    void get_params(float* host_params, float* host_musd) const;
    // Mean squared error over the n rows and the gradient (host copy).
    double loss_and_grad(const float* feats, const float* targets, int n, float* grad_host);
    // Minibatch Adam; returns the last epoch's mean loss.
    double train(const float* feats, const float* targets, int n, int epochs, int batch, float lr,
                 unsigned seed);
    // Scores (n x MLP_OUT) and argmin per row; either output may be null.
    void predict(const float* feats, int n, float* scores_host, int* best_host);
    bool save(const char* path) const;
    bool load(const char* path);

private:
    void reserve(int n);
    void forward_dev(int n);
    void grad_dev(int n);

    float* P = nullptr;     // parameters
    float* M = nullptr;     // Adam first moment
    float* V = nullptr;     // Adam second moment
    float* G = nullptr;     // gradient
    float* musd = nullptr;  // mu[IN], sd[IN]
    int step = 0;
    int cap = 0;
    float* buf = nullptr;
    float *F = nullptr, *X = nullptr, *H1 = nullptr, *H2 = nullptr, *DH1 = nullptr,
          *DH2 = nullptr, *Y = nullptr, *T = nullptr, *DY = nullptr, *L = nullptr;
    int* best = nullptr;
};

} // namespace blis_amd
