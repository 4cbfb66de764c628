// Forward, backward and Adam kernels of the planner's cost model (mlp.h).
// The network is tiny (hidden 64) and the batches are a few hundred rows,
// so every kernel maps one thread to one output value and sums in a fixed
// order: results do not depend on the launch shape.
#include "mlp.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>

#include "common.h"

namespace qvd {

namespace {

constexpr unsigned THREADS = 256;

unsigned blocks_for(uint64_t n) { return (unsigned)std::max<uint64_t>(1, (n + THREADS - 1) / THREADS); }

// out[r, o] = act(b[o] + sum_i W[o, i] in[r, i])
__global__ void k_layer(const float *__restrict__ in, uint32_t rows, uint32_t I, const float *__restrict__ W,
                        const float *__restrict__ b, uint32_t O, int relu, float *__restrict__ out) {
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (uint64_t)rows * O) return;
    const uint32_t r = (uint32_t)(idx / O), o = (uint32_t)(idx % O);
    const float *x = in + (uint64_t)r * I;
    const float *w = W + (uint64_t)o * I;
    float acc = b[o];
    for (uint32_t i = 0; i < I; i++) acc = fmaf(w[i], x[i], acc);
    out[idx] = relu ? fmaxf(acc, 0.0f) : acc;
}

// dW[o, i] = sum_r dout[r, o] in[r, i]; db[o] = sum_r dout[r, o]
// (thread (o, I) computes db[o]).
__global__ void k_layer_grad(const float *__restrict__ dout, const float *__restrict__ in, uint32_t rows, uint32_t I,
                             uint32_t O, float *__restrict__ dW, float *__restrict__ db) {
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (uint64_t)O * (I + 1)) return;
    const uint32_t o = (uint32_t)(idx / (I + 1)), i = (uint32_t)(idx % (I + 1));
    float acc = 0;
    if (i == I) {
        for (uint32_t r = 0; r < rows; r++) acc += dout[(uint64_t)r * O + o];
        db[o] = acc;
    } else {
        for (uint32_t r = 0; r < rows; r++) acc = fmaf(dout[(uint64_t)r * O + o], in[(uint64_t)r * I + i], acc);
        dW[(uint64_t)o * I + i] = acc;
    }
}

// din[r, i] = (sum_o dout[r, o] W[o, i]) where act[r, i] > 0 (ReLU), else 0
__global__ void k_layer_back(const float *__restrict__ dout, const float *__restrict__ W,
                             const float *__restrict__ act, uint32_t rows, uint32_t I, uint32_t O,
                             float *__restrict__ din) {
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (uint64_t)rows * I) return;
    const uint32_t r = (uint32_t)(idx / I), i = (uint32_t)(idx % I);
    float acc = 0;
    if (act[idx] > 0.0f)
        for (uint32_t o = 0; o < O; o++) acc = fmaf(dout[(uint64_t)r * O + o], W[(uint64_t)o * I + i], acc);
    din[idx] = acc;
}

// Mean squared error: dy[r] = 2 (y[r] - t[r]) / rows, sq[r] = (y[r] - t[r])^2
__global__ void k_mse(const float *__restrict__ y, const float *__restrict__ t, uint32_t rows, float *__restrict__ dy,
                      float *__restrict__ sq) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const float e = y[r] - t[r];
    dy[r] = 2.0f * e / (float)rows;
    sq[r] = e * e;
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, uint64_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_gather(const float *__restrict__ X, const float *__restrict__ T, const uint32_t *__restrict__ idx,
                         uint32_t rows, uint32_t I, float *__restrict__ xb, float *__restrict__ tb) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (uint64_t)rows * (I + 1)) return;
    const uint32_t r = (uint32_t)(k / (I + 1)), i = (uint32_t)(k % (I + 1));
    const uint64_t src = idx[r];
    if (i == I) tb[r] = T[src];
    else xb[(uint64_t)r * I + i] = X[src * I + i];
}

}  // namespace

void Mlp::Buf::need(size_t b) {
    if (bytes >= b) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    QVD_HIP(hipMalloc(&p, b));
    bytes = b;
}

Mlp::Buf::~Buf() {
    if (p) (void)hipFree(p);
}

Mlp::Mlp(uint32_t inputs, uint32_t hidden) {
    require(inputs >= 1 && inputs <= 4096 && hidden >= 1 && hidden <= 4096, "network sizes out of range");
    shape_.inputs = inputs;
    shape_.hidden = hidden;
    QVD_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    const size_t pb = shape_.params() * sizeof(float);
    params_.need(pb);
    grad_.need(pb);
    m_.need(pb);
    v_.need(pb);
    std::vector<float> zero(shape_.params(), 0.0f);
    QVD_HIP(hipMemcpy(params_.p, zero.data(), pb, hipMemcpyHostToDevice));
}

Mlp::~Mlp() {
    if (stream_) {
        (void)hipStreamSynchronize(stream_);
        (void)hipStreamDestroy(stream_);
    }
}

void Mlp::set_params(const float *p) {
    QVD_HIP(hipMemcpy(params_.p, p, shape_.params() * sizeof(float), hipMemcpyHostToDevice));
}

void Mlp::get_params(float *p) const {
    QVD_HIP(hipMemcpy(p, params_.p, shape_.params() * sizeof(float), hipMemcpyDeviceToHost));
}

void Mlp::forward_dev(const float *x, uint64_t rows) {
    const uint32_t I = shape_.inputs, H = shape_.hidden;
    h1_.need(rows * H * sizeof(float));
    h2_.need(rows * H * sizeof(float));
    y_.need(std::max<uint64_t>(rows, 1) * sizeof(float));
    const float *P = params_.f();
    k_layer<<<blocks_for(rows * H), THREADS, 0, stream_>>>(x, (uint32_t)rows, I, P + shape_.w1(), P + shape_.b1(), H, 1,
                                                           h1_.f());
    k_layer<<<blocks_for(rows * H), THREADS, 0, stream_>>>(h1_.f(), (uint32_t)rows, H, P + shape_.w2(), P + shape_.b2(),
                                                           H, 1, h2_.f());
    k_layer<<<blocks_for(rows), THREADS, 0, stream_>>>(h2_.f(), (uint32_t)rows, H, P + shape_.w3(), P + shape_.b3(), 1, 0,
                                                       y_.f());
    QVD_HIP(hipGetLastError());
}

double Mlp::backward_dev(const float *x, const float *t, uint64_t rows, bool want_loss) {
    const uint32_t I = shape_.inputs, H = shape_.hidden;
    dy_.need(std::max<uint64_t>(rows, 1) * sizeof(float));
    sq_.need(std::max<uint64_t>(rows, 1) * sizeof(float));
    dh1_.need(rows * H * sizeof(float));
    dh2_.need(rows * H * sizeof(float));
    const float *P = params_.f();
    float *G = grad_.f();
    const uint32_t R = (uint32_t)rows;
    k_mse<<<blocks_for(rows), THREADS, 0, stream_>>>(y_.f(), t, R, dy_.f(), sq_.f());
    k_layer_grad<<<blocks_for((uint64_t)1 * (H + 1)), THREADS, 0, stream_>>>(dy_.f(), h2_.f(), R, H, 1, G + shape_.w3(),
                                                                             G + shape_.b3());
    k_layer_back<<<blocks_for(rows * H), THREADS, 0, stream_>>>(dy_.f(), P + shape_.w3(), h2_.f(), R, H, 1, dh2_.f());
    k_layer_grad<<<blocks_for((uint64_t)H * (H + 1)), THREADS, 0, stream_>>>(dh2_.f(), h1_.f(), R, H, H, G + shape_.w2(),
                                                                             G + shape_.b2());
    k_layer_back<<<blocks_for(rows * H), THREADS, 0, stream_>>>(dh2_.f(), P + shape_.w2(), h1_.f(), R, H, H, dh1_.f());
    k_layer_grad<<<blocks_for((uint64_t)H * (I + 1)), THREADS, 0, stream_>>>(dh1_.f(), x, R, I, H, G + shape_.w1(),
                                                                             G + shape_.b1());
    QVD_HIP(hipGetLastError());
    if (!want_loss) return 0;
    std::vector<float> sq(rows);
    QVD_HIP(hipMemcpyAsync(sq.data(), sq_.p, rows * sizeof(float), hipMemcpyDeviceToHost, stream_));
    QVD_HIP(hipStreamSynchronize(stream_));
    double s = 0;
    for (float e : sq) s += e;
    return s / (double)rows;
}

void Mlp::forward(const float *x, uint64_t rows, float *y) {
    if (rows == 0) return;
    const uint32_t I = shape_.inputs;
    xbuf_.need(rows * I * sizeof(float));
    QVD_HIP(hipMemcpyAsync(xbuf_.p, x, rows * I * sizeof(float), hipMemcpyHostToDevice, stream_));
    forward_dev(xbuf_.f(), rows);
    QVD_HIP(hipMemcpyAsync(y, y_.p, rows * sizeof(float), hipMemcpyDeviceToHost, stream_));
    QVD_HIP(hipStreamSynchronize(stream_));
}

double Mlp::gradient(const float *x, const float *t, uint64_t rows, float *grad) {
    require(rows >= 1, "gradient over no rows");
    const uint32_t I = shape_.inputs;
    xbuf_.need(rows * I * sizeof(float));
    tbuf_.need(rows * sizeof(float));
    QVD_HIP(hipMemcpyAsync(xbuf_.p, x, rows * I * sizeof(float), hipMemcpyHostToDevice, stream_));
    QVD_HIP(hipMemcpyAsync(tbuf_.p, t, rows * sizeof(float), hipMemcpyHostToDevice, stream_));
    forward_dev(xbuf_.f(), rows);
    const double loss = backward_dev(xbuf_.f(), tbuf_.f(), rows, true);
    QVD_HIP(hipMemcpyAsync(grad, grad_.p, shape_.params() * sizeof(float), hipMemcpyDeviceToHost, stream_));
    QVD_HIP(hipStreamSynchronize(stream_));
    return loss;
}

double Mlp::train(const float *x, const float *t, uint64_t rows, uint32_t epochs, uint32_t batch, float lr,
                  uint64_t seed) {
    require(rows >= 1 && batch >= 1, "training needs rows and a batch size");
    const uint32_t I = shape_.inputs;
    batch = (uint32_t)std::min<uint64_t>(batch, rows);
    xbuf_.need(rows * I * sizeof(float));
    tbuf_.need(rows * sizeof(float));
    xb_.need((uint64_t)batch * I * sizeof(float));
    tb_.need((uint64_t)batch * sizeof(float));
    idx_.need(rows * sizeof(uint32_t));
    QVD_HIP(hipMemcpyAsync(xbuf_.p, x, rows * I * sizeof(float), hipMemcpyHostToDevice, stream_));
    QVD_HIP(hipMemcpyAsync(tbuf_.p, t, rows * sizeof(float), hipMemcpyHostToDevice, stream_));
    const uint64_t np = shape_.params();
    std::vector<float> zero(np, 0.0f);
    QVD_HIP(hipMemcpyAsync(m_.p, zero.data(), np * sizeof(float), hipMemcpyHostToDevice, stream_));
    QVD_HIP(hipMemcpyAsync(v_.p, zero.data(), np * sizeof(float), hipMemcpyHostToDevice, stream_));
    std::vector<uint32_t> order(rows);
    std::iota(order.begin(), order.end(), 0u);
    std::mt19937_64 rng(seed);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    uint64_t step = 0;
    for (uint32_t e = 0; e < epochs; e++) {
        std::shuffle(order.begin(), order.end(), rng);
        QVD_HIP(hipMemcpyAsync(idx_.p, order.data(), rows * sizeof(uint32_t), hipMemcpyHostToDevice, stream_));
        // the host order array must outlive the copy
        QVD_HIP(hipStreamSynchronize(stream_));
        for (uint64_t start = 0; start + batch <= rows; start += batch) {
            k_gather<<<blocks_for((uint64_t)batch * (I + 1)), THREADS, 0, stream_>>>(
                xbuf_.f(), tbuf_.f(), static_cast<const uint32_t *>(idx_.p) + start, batch, I, xb_.f(), tb_.f());
            forward_dev(xb_.f(), batch);
            backward_dev(xb_.f(), tb_.f(), batch, false);
            step++;
            const float c1 = 1.0f - std::pow(b1, (float)step), c2 = 1.0f - std::pow(b2, (float)step);
            k_adam<<<blocks_for(np), THREADS, 0, stream_>>>(params_.f(), grad_.f(), m_.f(), v_.f(), np, lr, b1, b2, eps,
                                                            c1, c2);
            QVD_HIP(hipGetLastError());
        }
    }
    forward_dev(xbuf_.f(), rows);
    return backward_dev(xbuf_.f(), tbuf_.f(), rows, true);
}

}  // namespace qvd
