// mlp.hpp -- the candidate-proposer network (6 -> 128 -> 128 -> 4, BF16 MFMA,
// FP32 accumulation) and its GPU training state.
#pragma once

#include <cstdint>
#include <cstring>

#include "common.hpp"

namespace axis_hip::mlp {

constexpr int kIn = 6, kInP = 16, kH = 128, kOut = 4, kOutP = 16;
constexpr int kTile = 32;  // samples per workgroup

// offsets in the flat FP32 parameter vector (weights are [out][in])
constexpr int oW1 = 0;
constexpr int oB1 = oW1 + kH * kIn;
constexpr int oW2 = oB1 + kH;
constexpr int oB2 = oW2 + kH * kH;
constexpr int oW3 = oB2 + kH;
constexpr int oB3 = oW3 + kOut * kH;
constexpr int kParams = oB3 + kOut;  // 17,924

/// FP32 -> BF16, round to nearest even (finite inputs).
__host__ __device__ inline std::uint16_t round_bf16(float f) {
    std::uint32_t u;
#ifdef __HIP_DEVICE_COMPILE__
    u = __float_as_uint(f);
#else
    std::memcpy(&u, &f, sizeof u);
#endif
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<std::uint16_t>(u >> 16);
}

/// Seeded initial parameters (He-uniform hidden layers, small output layer, zero biases).
void init_params(std::uint64_t seed, float *out);

class Model {
   public:
    /// params_host: kParams FP32 values; feature_scale: the length unit of
    /// the inputs' extents and of the outputs (radians).
    Model(const float *params_host, float feature_scale);
    float feature_scale() const { return feature_scale_; }
    void set_params(const float *params_host, hipStream_t s);
    void get_params(float *params_host, hipStream_t s) const;
    /// out[n][4] = network(x[n][6]) (device pointers).
    void forward(const float *x, index_t n, float *out, hipStream_t s) const;
    /// Mean-loss gradient over x[n][6], t[n][4] into grad[kParams] (device); returns the mean loss.
    float gradient(const float *x, const float *t, index_t n, float wc, float tau, float *grad, hipStream_t s);
    /// One pass of Adam over the samples in order, mini-batches of `batch`; returns the mean loss.
    float train(const float *x, const float *t, index_t n, index_t batch, float lr, float wc, float tau, hipStream_t s);
    const std::uint16_t *packed() const { return wb_.get(); }
    const float *params() const { return p_.get(); }

   private:
    float feature_scale_;
    DeviceBuffer<float> p_, m_, v_;
    DeviceBuffer<std::uint16_t> wb_;
    DeviceBuffer<float> part_, loss_;
    long step_ = 0;
};

}  // namespace axis_hip::mlp
