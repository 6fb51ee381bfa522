// nsg_cnn.hpp -- the learned subgrid-stress network (shared between nsg_cnn.hip and
// nsg_closure.hip; not part of the C ABI).
//
// Network: four periodic 3x3x3 convolutions 9 -> 32 -> 32 -> 32 -> 6 channels, exact-erf GELU
// after the first three.  BF16 weights and activations, FP32 accumulation on MFMA
// (v_mfma_f32_16x16x32_bf16), FP32 pre-activations and output.  Activations are channels-last
// [batch][N^3][C] on an N^3 periodic grid, point index (x * N + y) * N + z.  Channel counts are
// padded to 16 or 32 on the device (9 -> 16 for the input, 6 -> 16 for the output); padded
// channels hold zeros and padded weights stay zero.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace nsgcnn {

constexpr int NLAYER = 4;
constexpr int CREAL[NLAYER + 1] = {9, 32, 32, 32, 6};
constexpr int CPAD[NLAYER + 1] = {16, 32, 32, 32, 16};

// master parameters of layer l (FP32): W [CPAD[l+1]][27 * CPAD[l]] (row = output channel,
// column = tap * CIN + input channel, tap = (a * 3 + b) * 3 + c for the offsets (a-1, b-1, c-1)
// along (x, y, z)), then the bias [CPAD[l+1]]
constexpr long long wcount(int l) { return (long long)CPAD[l + 1] * 27 * CPAD[l]; }
constexpr long long lcount(int l) { return wcount(l) + CPAD[l + 1]; }
constexpr long long loffset(int l) { return l == 0 ? 0 : loffset(l - 1) + lcount(l - 1); }
constexpr long long MASTER_COUNT = loffset(NLAYER);
// parameters in the exchange order (torch conv3d order, unpadded): per layer W[co][ci][a][b][c], b[co]
constexpr long long rcount(int l) { return (long long)CREAL[l + 1] * CREAL[l] * 27 + CREAL[l + 1]; }
constexpr long long PARAM_COUNT = rcount(0) + rcount(1) + rcount(2) + rcount(3);

}  // namespace nsgcnn

struct nsg_cnn {
    int device = 0, N = 0, maxB = 1;
    long long N3 = 0;
    hipStream_t s = nullptr;
    std::string err;
    // parameters
    float* master = nullptr;  // MASTER_COUNT
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    float* grad = nullptr;                   // MASTER_COUNT (same layout)
    uint16_t* wf[nsgcnn::NLAYER] = {};       // forward weights, bf16 [COUT][27 CIN]
    uint16_t* wb[nsgcnn::NLAYER] = {};       // flipped, transposed, bf16 [CIN][27 COUT] (input gradients)
    long long adam_t = 0;
    // activations [maxB][N^3][C]
    uint16_t* act[nsgcnn::NLAYER] = {};      // act[0] = network input (16 ch), act[1..3] = 32 ch
    float* pre[nsgcnn::NLAYER - 1] = {};     // pre-activations of layers 1..3 (32 ch)
    float* out = nullptr;                    // output, 16 ch (6 used)
    uint16_t* dp[2] = {};                    // dL/d(pre-activation), ping-pong, 32 ch
    float* target = nullptr;                 // [maxB][N^3][6] training targets of the current batch
    float* partial = nullptr;                // per-block weight-gradient partial sums
    size_t partial_n = 0;
    double* lpart = nullptr;                 // loss partial sums
    int wgrad_blocks = 256;
    int ncu = 256;                           // compute units of the device
    // data set (normalised): inputs bf16 [N^3][16], targets FP32 [N^3][6]
    std::vector<uint16_t*> xs;
    std::vector<float*> ys;
    double sig_in = 1.0, sig_out[6] = {1, 1, 1, 1, 1, 1};
    hipEvent_t ev[2] = {nullptr, nullptr};
};

namespace nsgcnn {
// forward pass of `B` samples whose normalised bf16 inputs are in c->act[0]; the output lands in
// c->out (normalised).  train = 1 keeps the pre-activations for a backward pass.
int forward_dev(nsg_cnn* c, hipStream_t s, int B, bool train);
}  // namespace nsgcnn
