// The planner's cost model: a 3-layer perceptron (inputs -> hidden -> hidden
// -> 1, ReLU between layers, FP32) with hand-written forward, backward and
// Adam kernels (mlp.hip). Parameters are one contiguous FP32 array:
//   W1[hidden x inputs], b1[hidden], W2[hidden x hidden], b2[hidden],
//   W3[hidden], b3[1]   (W row-major: W[o * in + i]).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace qvd {

struct MlpShape {
    uint32_t inputs = 0, hidden = 0;
    uint64_t params() const { return (uint64_t)hidden * inputs + hidden + (uint64_t)hidden * hidden + hidden + hidden + 1; }
    uint64_t w1() const { return 0; }
    uint64_t b1() const { return (uint64_t)hidden * inputs; }
    uint64_t w2() const { return b1() + hidden; }
    uint64_t b2() const { return w2() + (uint64_t)hidden * hidden; }
    uint64_t w3() const { return b2() + hidden; }
    uint64_t b3() const { return w3() + hidden; }
};

// Device-resident network with its training state. Every call is
// synchronous on `stream`.
class Mlp {
  public:
    Mlp(uint32_t inputs, uint32_t hidden);
    ~Mlp();
    Mlp(const Mlp &) = delete;
    Mlp &operator=(const Mlp &) = delete;

    const MlpShape &shape() const { return shape_; }
    void set_params(const float *p);
    void get_params(float *p) const;
    // y[r] for rows of x (host arrays).
    void forward(const float *x, uint64_t rows, float *y);
    // Gradient of the mean squared error over the rows with respect to every
    // parameter (host arrays); returns the loss.
    double gradient(const float *x, const float *t, uint64_t rows, float *grad);
    // Adam on minibatches of shuffled rows; returns the final full-data loss.
    double train(const float *x, const float *t, uint64_t rows, uint32_t epochs, uint32_t batch, float lr,
                 uint64_t seed);

  private:
    struct Buf {
        void *p = nullptr;
        size_t bytes = 0;
        void need(size_t b);
        ~Buf();
        float *f() const { return static_cast<float *>(p); }
    };
    // Forward over device rows; activations stay in h1_, h2_, y_.
    void forward_dev(const float *x, uint64_t rows);
    // Loss gradient of device rows (after forward_dev) into grad_; returns
    // the loss when want_loss.
    double backward_dev(const float *x, const float *t, uint64_t rows, bool want_loss);

    MlpShape shape_;
    hipStream_t stream_ = nullptr;
    Buf params_, grad_, m_, v_, h1_, h2_, y_, dy_, dh1_, dh2_, sq_, xbuf_, tbuf_, xb_, tb_, idx_;
};

}  // namespace qvd
