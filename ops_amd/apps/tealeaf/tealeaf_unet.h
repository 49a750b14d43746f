// Learned warm start for the TeaLeaf implicit solve (no reference counterpart): a 4-level
// 2-D U-Net (channels 16/32/48/64, 3 x 3 convolutions, BF16 weights and activations, FP32
// accumulation on MFMA) trained on the GPU by its own forward and backward kernels.
// Parameters are one FP32 vector: per layer the weights [Cout][3][3][Cin] then the biases.
#ifndef TEALEAF_UNET_H
#define TEALEAF_UNET_H

#ifdef __cplusplus
extern "C" {
#endif

int tea_unet_init(unsigned long long seed);
long tea_unet_num_params(void);
int tea_unet_get_params(float *host);
int tea_unet_set_params(const float *host);
// in: [B][H][W][16] (channels 4..15 zero), out: [B][H][W] prediction; H, W multiples of 8
int tea_unet_forward_host(int B, int H, int W, const float *in, float *out);
// parameter gradients for d loss / d prediction = grad_out ([B][H][W])
int tea_unet_backward_host(int B, int H, int W, const float *in, const float *grad_out, float *gparams);
// device dataset of (4 features, target) pairs at H x W, and Adam training over it
// (tealeaf_capture_next fills samples from TeaLeaf solves directly on the device)
int tea_unet_dataset_reserve(int count, int H, int W);
int tea_unet_dataset_set(int index, const float *in4, const float *target);
// loss: error norm e^T (I + beta C) e relative to the target's, C the conduction part of
// the sample's operator (beta = 0: relative mean squared error; beta = 1: energy norm)
int tea_unet_train(int epochs, int batch, float lr, float beta, unsigned long long seed, float *loss_per_epoch);
int tea_unet_free(void);

#ifdef __cplusplus
}
#endif

#endif
