// mlp.hip — the learned GEMM tile selector's network: a small FP32
// multilayer perceptron (16 features -> 64 -> 64 -> one score per tile /
// schedule configuration), trained and evaluated on the GPU by the kernels
// below.
//
// The reference has no counterpart. It picks blocksizes from a fixed
// per-architecture table (config/zen3/bli_cntx_init_zen3.c:204-220). Here the
// network predicts, for each configuration c, y_c = log2(t_c / t_best), the
// slowdown of c against the fastest configuration for the shape. The selector
// takes argmin_c y_c. The network only chooses among configurations that
// all compute bit-identical results, so it never changes a GEMM's output.
//
// Parameters (one FP32 array):
//   W1[H][IN] b1[H] W2[H][H] b2[H] W3[OUT][H] b3[OUT].
// Layers are h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), y = W3 h2 + b3.
// The loss is mean((y - t)^2) over batch and outputs, optimised with Adam.
// Every reduction runs in a fixed order, so the gradients and the trained
// weights are deterministic.
#include "internal.hpp"
#include "mlp.hpp"

#include <cmath>
#include <cstring>
#include <numeric>
#include <random>

namespace blis_amd {

namespace detail_mlp {

constexpr int IN = MLP_IN, H = MLP_H, OUT = MLP_OUT;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + H * IN, OFF_W2 = OFF_B1 + H,
              OFF_B2 = OFF_W2 + H * H, OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + OUT * H,
              NPARAM = OFF_B3 + OUT;
static_assert(NPARAM == MLP_NPARAM, "parameter layout");
static_assert(OUT <= H, "one thread per output in a 64-thread block");

// Normalise raw features: x = (f - mu) / sd.
__global__ void normalise_kernel(const float* __restrict__ f, const float* __restrict__ musd,
                                 float* __restrict__ x, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * IN) return;
    const int i = t % IN;
    x[t] = (f[t] - musd[i]) / musd[IN + i];
}

// One block of 64 threads per sample; thread j owns hidden unit j.
__global__ __launch_bounds__(64) void forward_kernel(const float* __restrict__ P,
                                                     const float* __restrict__ X, float* __restrict__ H1,
                                                     float* __restrict__ H2, float* __restrict__ Y,
                                                     int n) {
    __shared__ float xs[IN], h1s[H], h2s[H];
    const int b = blockIdx.x, j = threadIdx.x;
    if (b >= n) return;
    if (j < IN) xs[j] = X[b * IN + j];
    __syncthreads();
    float a = P[OFF_B1 + j];
    for (int i = 0; i < IN; ++i) a = fmaf(P[OFF_W1 + j * IN + i], xs[i], a);
    a = a > 0.f ? a : 0.f;
    h1s[j] = a;
    H1[b * H + j] = a;
    __syncthreads();
    a = P[OFF_B2 + j];
    for (int i = 0; i < H; ++i) a = fmaf(P[OFF_W2 + j * H + i], h1s[i], a);
    a = a > 0.f ? a : 0.f;
    h2s[j] = a;
    H2[b * H + j] = a;
    __syncthreads();
    if (j < OUT) {
        a = P[OFF_B3 + j];
        for (int i = 0; i < H; ++i) a = fmaf(P[OFF_W3 + j * H + i], h2s[i], a);
        Y[b * OUT + j] = a;
    }
}

// dY = 2 (y - t) / (n * OUT); per-sample squared error into L[b].
__global__ __launch_bounds__(64) void loss_kernel(const float* __restrict__ Y,
                                                  const float* __restrict__ T, float* __restrict__ DY,
                                                  float* __restrict__ L, int n) {
    __shared__ float part[64];
    const int b = blockIdx.x, o = threadIdx.x;
    if (b >= n) return;
    float e = 0.f;
    if (o < OUT) {
        const float d = Y[b * OUT + o] - T[b * OUT + o];
        DY[b * OUT + o] = 2.f * d / (float)(n * OUT);
        e = d * d;
    }
    part[o] = e;
    __syncthreads();
    if (o == 0) {
        float s = 0.f;
        for (int i = 0; i < 64; ++i) s += part[i];
        L[b] = s;
    }
}

// Back-propagate through the two hidden layers, per sample.
__global__ __launch_bounds__(64) void backward_kernel(const float* __restrict__ P,
                                                      const float* __restrict__ H1,
                                                      const float* __restrict__ H2,
                                                      const float* __restrict__ DY,
                                                      float* __restrict__ DH1,
                                                      float* __restrict__ DH2, int n) {
    __shared__ float dys[OUT], d2s[H];
    const int b = blockIdx.x, j = threadIdx.x;
    if (b >= n) return;
    if (j < OUT) dys[j] = DY[b * OUT + j];
    __syncthreads();
    float a = 0.f;
    for (int o = 0; o < OUT; ++o) a = fmaf(P[OFF_W3 + o * H + j], dys[o], a);
    a = H2[b * H + j] > 0.f ? a : 0.f;
    d2s[j] = a;
    DH2[b * H + j] = a;
    __syncthreads();
    a = 0.f;
    for (int i = 0; i < H; ++i) a = fmaf(P[OFF_W2 + i * H + j], d2s[i], a);
    a = H1[b * H + j] > 0.f ? a : 0.f;
    DH1[b * H + j] = a;
}

// One thread per parameter: the gradient summed over the batch in sample order.
__global__ __launch_bounds__(256) void weight_grad_kernel(const float* __restrict__ X,
                                                          const float* __restrict__ H1,
                                                          const float* __restrict__ H2,
                                                          const float* __restrict__ DY,
                                                          const float* __restrict__ DH1,
                                                          const float* __restrict__ DH2,
                                                          float* __restrict__ G, int n) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPARAM) return;
    const float *u, *v = nullptr;   // gradient = sum_b u[b*us + ui] * v[b*vs + vi]
    int us, ui, vs = 0, vi = 0;
    if (p < OFF_B1) { const int r = p / IN, c = p % IN; u = DH1; us = H; ui = r; v = X; vs = IN; vi = c; }
    else if (p < OFF_W2) { u = DH1; us = H; ui = p - OFF_B1; }
    else if (p < OFF_B2) { const int q = p - OFF_W2, r = q / H, c = q % H; u = DH2; us = H; ui = r; v = H1; vs = H; vi = c; }
    else if (p < OFF_W3) { u = DH2; us = H; ui = p - OFF_B2; }
    else if (p < OFF_B3) { const int q = p - OFF_W3, r = q / H, c = q % H; u = DY; us = OUT; ui = r; v = H2; vs = H; vi = c; }
    else { u = DY; us = OUT; ui = p - OFF_B3; }
    float g = 0.f;
    if (v) for (int b = 0; b < n; ++b) g = fmaf(u[b * us + ui], v[b * vs + vi], g);
    else for (int b = 0; b < n; ++b) g += u[b * us + ui];
    G[p] = g;
}

__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ P, float* __restrict__ M,
                                                   float* __restrict__ V, const float* __restrict__ G,
                                                   float lr, float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPARAM) return;
    const float g = G[p];
    const float m = 0.9f * M[p] + 0.1f * g;
    const float v = 0.999f * V[p] + 0.001f * g * g;
    M[p] = m;
    V[p] = v;
    P[p] -= lr * (m / c1) / (sqrtf(v / c2) + 1e-8f);
}

// Rows idx[b0 .. b0+nb) of the training set into the batch buffers.
__global__ void gather_kernel(const float* __restrict__ sF, const float* __restrict__ sT,
                              const int* __restrict__ idx, int b0, int nb, float* __restrict__ F,
                              float* __restrict__ T) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int per = IN + OUT;
    if (t >= nb * per) return;
    const int r = t / per, c = t % per;
    const int64_t src = idx[b0 + r];
    if (c < IN) F[(int64_t)r * IN + c] = sF[src * IN + c];
    else T[(int64_t)r * OUT + (c - IN)] = sT[src * OUT + (c - IN)];
}

// argmin over the OUT scores of each sample (first index on ties).
__global__ void argmin_kernel(const float* __restrict__ Y, int* __restrict__ best, int n) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    int bi = 0;
    float bv = Y[b * OUT];
    for (int o = 1; o < OUT; ++o)
        if (Y[b * OUT + o] < bv) { bv = Y[b * OUT + o]; bi = o; }
    best[b] = bi;
}

inline unsigned blocks(int threads, int tpb) { return (unsigned)((threads + tpb - 1) / tpb); }

} // namespace detail_mlp
using namespace detail_mlp;

MlpNet::MlpNet() {
    HIP_CHECK(hipMalloc(&P, sizeof(float) * NPARAM * 4));
    M = P + NPARAM;
    V = M + NPARAM;
    G = V + NPARAM;
    HIP_CHECK(hipMalloc(&musd, sizeof(float) * 2 * IN));
    HIP_CHECK(hipMemset(P, 0, sizeof(float) * NPARAM * 4));
    std::vector<float> ms(2 * IN, 0.f);
    for (int i = 0; i < IN; ++i) ms[IN + i] = 1.f;
    HIP_CHECK(hipMemcpy(musd, ms.data(), sizeof(float) * 2 * IN, hipMemcpyHostToDevice));
}

MlpNet::~MlpNet() {
    (void)hipFree(P);
    (void)hipFree(musd);
    (void)hipFree(buf);
}

void MlpNet::reserve(int n) {
    if (n <= cap) return;
    if (buf) { HIP_CHECK(hipStreamSynchronize(stream())); HIP_CHECK(hipFree(buf)); }
    // F, X: n*IN; H1, H2, DH1, DH2: n*H; Y, T, DY: n*OUT; L: n; best: n ints
    const size_t floats = (size_t)n * (2 * IN + 4 * H + 3 * OUT + 2);
    HIP_CHECK(hipMalloc(&buf, sizeof(float) * floats));
    cap = n;
    F = buf;
    X = F + (size_t)n * IN;
    H1 = X + (size_t)n * IN;
    H2 = H1 + (size_t)n * H;
    DH1 = H2 + (size_t)n * H;
    DH2 = DH1 + (size_t)n * H;
    Y = DH2 + (size_t)n * H;
    T = Y + (size_t)n * OUT;
    DY = T + (size_t)n * OUT;
    L = DY + (size_t)n * OUT;
    best = reinterpret_cast<int*>(L + n);
}

// Xavier-uniform weights from a seeded generator, zero biases, fresh Adam state.
void MlpNet::init(unsigned seed) {
    std::vector<float> p(NPARAM, 0.f);
    std::mt19937 rng(seed);
    auto fill = [&](int off, int rows, int cols) {
        const float lim = std::sqrt(6.f / (float)(rows + cols));
        std::uniform_real_distribution<float> u(-lim, lim);
        for (int i = 0; i < rows * cols; ++i) p[off + i] = u(rng);
    };
    fill(OFF_W1, H, IN);
    fill(OFF_W2, H, H);
    fill(OFF_W3, OUT, H);
    set_params(p.data(), nullptr);
}

void MlpNet::set_params(const float* host_params, const float* host_musd) {
    hipStream_t st = stream();
    HIP_CHECK(hipMemcpyAsync(P, host_params, sizeof(float) * NPARAM, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemsetAsync(M, 0, sizeof(float) * NPARAM * 2, st));
    if (host_musd)
        HIP_CHECK(hipMemcpyAsync(musd, host_musd, sizeof(float) * 2 * IN, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipStreamSynchronize(st));
    step = 0;
}

void MlpNet::get_params(float* host_params, float* host_musd) const {
    hipStream_t st = stream();
    if (host_params)
        HIP_CHECK(hipMemcpyAsync(host_params, P, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, st));
    if (host_musd)
        HIP_CHECK(hipMemcpyAsync(host_musd, musd, sizeof(float) * 2 * IN, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
}

// Forward over rows [0, n) of F (already on the device); scores in Y.
void MlpNet::forward_dev(int n) {
    hipStream_t st = stream();
    hipLaunchKernelGGL(normalise_kernel, dim3(blocks(n * IN, 256)), dim3(256), 0, st, F, musd, X, n);
    hipLaunchKernelGGL(forward_kernel, dim3(n), dim3(64), 0, st, P, X, H1, H2, Y, n);
    HIP_CHECK(hipGetLastError());
}

// Loss and gradient (into G) of the batch in F/T (device), n rows.
void MlpNet::grad_dev(int n) {
    hipStream_t st = stream();
    forward_dev(n);
    hipLaunchKernelGGL(loss_kernel, dim3(n), dim3(64), 0, st, Y, T, DY, L, n);
    hipLaunchKernelGGL(backward_kernel, dim3(n), dim3(64), 0, st, P, H1, H2, DY, DH1, DH2, n);
    hipLaunchKernelGGL(weight_grad_kernel, dim3(blocks(NPARAM, 256)), dim3(256), 0, st, X, H1, H2,
                       DY, DH1, DH2, G, n);
    HIP_CHECK(hipGetLastError());
}

double MlpNet::loss_and_grad(const float* feats, const float* targets, int n, float* grad_host) {
    reserve(n);
    hipStream_t st = stream();
    HIP_CHECK(hipMemcpyAsync(F, feats, sizeof(float) * n * IN, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(T, targets, sizeof(float) * n * OUT, hipMemcpyHostToDevice, st));
    grad_dev(n);
    std::vector<float> l(n);
    HIP_CHECK(hipMemcpyAsync(l.data(), L, sizeof(float) * n, hipMemcpyDeviceToHost, st));
    if (grad_host)
        HIP_CHECK(hipMemcpyAsync(grad_host, G, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    double s = 0.0;
    for (float v : l) s += v;
    return s / ((double)n * OUT);
}

double MlpNet::train(const float* feats, const float* targets, int n, int epochs, int batch,
                     float lr, unsigned seed) {
    if (n <= 0 || batch <= 0) return 0.0;
    // The whole training set lives on the device; minibatches are gathered
    // by a host-side permutation (deterministic for a given seed).
    float *dF = nullptr, *dT = nullptr;
    int* dI = nullptr;
    HIP_CHECK(hipMalloc(&dF, sizeof(float) * (size_t)n * IN));
    HIP_CHECK(hipMalloc(&dT, sizeof(float) * (size_t)n * OUT));
    HIP_CHECK(hipMalloc(&dI, sizeof(int) * (size_t)n));
    hipStream_t st = stream();
    HIP_CHECK(hipMemcpyAsync(dF, feats, sizeof(float) * (size_t)n * IN, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(dT, targets, sizeof(float) * (size_t)n * OUT, hipMemcpyHostToDevice, st));
    reserve(batch);
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::mt19937 rng(seed);
    double last = 0.0;
    std::vector<float> l(batch);
    for (int e = 0; e < epochs; ++e) {
        std::shuffle(perm.begin(), perm.end(), rng);
        HIP_CHECK(hipMemcpyAsync(dI, perm.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice,
                                 st));
        // perm is rewritten next epoch: make sure this copy has been consumed.
        HIP_CHECK(hipStreamSynchronize(st));
        double tot = 0.0;
        for (int b0 = 0; b0 < n; b0 += batch) {
            const int nb = std::min(batch, n - b0);
            hipLaunchKernelGGL(gather_kernel, dim3(blocks(nb * (IN + OUT), 256)), dim3(256), 0, st,
                               dF, dT, dI, b0, nb, F, T);
            grad_dev(nb);
            ++step;
            const float c1 = 1.f - std::pow(0.9f, (float)step);
            const float c2 = 1.f - std::pow(0.999f, (float)step);
            hipLaunchKernelGGL(adam_kernel, dim3(blocks(NPARAM, 256)), dim3(256), 0, st, P, M, V, G,
                               lr, c1, c2);
            HIP_CHECK(hipGetLastError());
            if (e == epochs - 1) {
                HIP_CHECK(hipMemcpyAsync(l.data(), L, sizeof(float) * nb, hipMemcpyDeviceToHost, st));
                HIP_CHECK(hipStreamSynchronize(st));
                for (int r = 0; r < nb; ++r) tot += l[r];
            }
        }
        if (e == epochs - 1) last = tot / ((double)n * OUT);
    }
    HIP_CHECK(hipStreamSynchronize(st));
    HIP_CHECK(hipFree(dF));
    HIP_CHECK(hipFree(dT));
    HIP_CHECK(hipFree(dI));
    return last;
}

void MlpNet::predict(const float* feats, int n, float* scores_host, int* best_host) {
    if (n <= 0) return;
    reserve(n);
    hipStream_t st = stream();
    HIP_CHECK(hipMemcpyAsync(F, feats, sizeof(float) * (size_t)n * IN, hipMemcpyHostToDevice, st));
    forward_dev(n);
    hipLaunchKernelGGL(argmin_kernel, dim3(blocks(n, 256)), dim3(256), 0, st, Y, best, n);
    HIP_CHECK(hipGetLastError());
    if (scores_host)
        HIP_CHECK(hipMemcpyAsync(scores_host, Y, sizeof(float) * (size_t)n * OUT,
                                 hipMemcpyDeviceToHost, st));
    if (best_host)
        HIP_CHECK(hipMemcpyAsync(best_host, best, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
}

// File format: "BLAMDMLP", int32 version 1, IN, H, OUT, then mu[IN], sd[IN],
// params[NPARAM] as little-endian float32.
bool MlpNet::save(const char* path) const {
    std::vector<float> p(NPARAM), ms(2 * IN);
    get_params(p.data(), ms.data());
    FILE* f = std::fopen(path, "wb");
    if (!f) return false;
    const int32_t hdr[4] = {1, IN, H, OUT};
    bool ok = std::fwrite("BLAMDMLP", 1, 8, f) == 8 && std::fwrite(hdr, sizeof(hdr), 1, f) == 1 &&
              std::fwrite(ms.data(), sizeof(float), ms.size(), f) == ms.size() &&
              std::fwrite(p.data(), sizeof(float), p.size(), f) == p.size();
    ok &= std::fclose(f) == 0;
    return ok;
}

bool mlp_read(const char* path, std::vector<float>& params, std::vector<float>& musd) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    char magic[8];
    int32_t hdr[4];
    std::vector<float> p(NPARAM), ms(2 * IN);
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "BLAMDMLP", 8) == 0 &&
              std::fread(hdr, sizeof(hdr), 1, f) == 1 && hdr[0] == 1 && hdr[1] == IN &&
              hdr[2] == H && hdr[3] == OUT &&
              std::fread(ms.data(), sizeof(float), ms.size(), f) == ms.size() &&
              std::fread(p.data(), sizeof(float), p.size(), f) == p.size();
    std::fclose(f);
    if (ok) {
        params.swap(p);
        musd.swap(ms);
    }
    return ok;
}

bool MlpNet::load(const char* path) {
    std::vector<float> p, ms;
    if (!mlp_read(path, p, ms)) return false;
    set_params(p.data(), ms.data());
    return true;
}

// The loops of normalise_kernel, forward_kernel and argmin_kernel for one
// row, with the same fused multiply-adds in the same order.
int MlpHost::argmin(const float* feats) const {
    const float* P = p.data();
    float x[IN], h1[H], h2[H];
    for (int i = 0; i < IN; ++i) x[i] = (feats[i] - musd[i]) / musd[IN + i];
    for (int j = 0; j < H; ++j) {
        float a = P[OFF_B1 + j];
        for (int i = 0; i < IN; ++i) a = std::fmaf(P[OFF_W1 + j * IN + i], x[i], a);
        h1[j] = a > 0.f ? a : 0.f;
    }
    for (int j = 0; j < H; ++j) {
        float a = P[OFF_B2 + j];
        for (int i = 0; i < H; ++i) a = std::fmaf(P[OFF_W2 + j * H + i], h1[i], a);
        h2[j] = a > 0.f ? a : 0.f;
    }
    int bi = 0;
    float bv = 0.f;
    for (int o = 0; o < OUT; ++o) {
        float a = P[OFF_B3 + o];
        for (int i = 0; i < H; ++i) a = std::fmaf(P[OFF_W3 + o * H + i], h2[i], a);
        if (o == 0 || a < bv) { bv = a; bi = o; }
    }
    return bi;
}

} // namespace blis_amd

using namespace blis_amd;

extern "C" {

__attribute__((visibility("default"))) int bli_amd_mlp_param_count(void) { return MLP_NPARAM; }

__attribute__((visibility("default"))) void* bli_amd_mlp_create(unsigned seed) {
    init_once();
    std::lock_guard<std::mutex> lock(api_mutex());
    MlpNet* n = new MlpNet();
    n->init(seed);
    return n;
}

__attribute__((visibility("default"))) void bli_amd_mlp_destroy(void* net) {
    std::lock_guard<std::mutex> lock(api_mutex());
    delete static_cast<MlpNet*>(net);
}

__attribute__((visibility("default")))
void bli_amd_mlp_set_params(void* net, const float* params, const float* mu_sd) {
    std::lock_guard<std::mutex> lock(api_mutex());
    static_cast<MlpNet*>(net)->set_params(params, mu_sd);
}

__attribute__((visibility("default")))
void bli_amd_mlp_get_params(void* net, float* params, float* mu_sd) {
    std::lock_guard<std::mutex> lock(api_mutex());
    static_cast<MlpNet*>(net)->get_params(params, mu_sd);
}

__attribute__((visibility("default")))
double bli_amd_mlp_loss_and_grad(void* net, const float* feats, const float* targets, int n,
                                 float* grad) {
    std::lock_guard<std::mutex> lock(api_mutex());
    return static_cast<MlpNet*>(net)->loss_and_grad(feats, targets, n, grad);
}

__attribute__((visibility("default")))
double bli_amd_mlp_train(void* net, const float* feats, const float* targets, int n, int epochs,
                         int batch, float lr, unsigned seed) {
    std::lock_guard<std::mutex> lock(api_mutex());
    return static_cast<MlpNet*>(net)->train(feats, targets, n, epochs, batch, lr, seed);
}

__attribute__((visibility("default")))
void bli_amd_mlp_predict(void* net, const float* feats, int n, float* scores, int* best) {
    std::lock_guard<std::mutex> lock(api_mutex());
    static_cast<MlpNet*>(net)->predict(feats, n, scores, best);
}

__attribute__((visibility("default"))) int bli_amd_mlp_save(void* net, const char* path) {
    std::lock_guard<std::mutex> lock(api_mutex());
    return static_cast<MlpNet*>(net)->save(path) ? 0 : -1;
}

__attribute__((visibility("default"))) void* bli_amd_mlp_load(const char* path) {
    init_once();
    std::lock_guard<std::mutex> lock(api_mutex());
    MlpNet* n = new MlpNet();
    if (!n->load(path)) { delete n; return nullptr; }
    return n;
}

} // extern "C"
