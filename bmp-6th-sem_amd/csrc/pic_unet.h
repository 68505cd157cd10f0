// pic_unet.h — the learned initial guess (U-Net) as used by the Poisson stage.
#pragma once

#include "../../include/bmp_pic.h"

struct pic_ctx;

namespace pic {

// The network's prediction for a solve on the N x N node image from rho (the
// right-hand side after the Dirichlet overwrite) and phi (the previous
// solution): the change of phi is out[y N + x] * max|r0| / diag on unknown
// nodes of the image, diag = 2(1/dx^2 + 1/dy^2); scl[0] holds the bits of
// max|r0| (0 stands for 1).  Device pointers, valid until the next call.
struct UnetPrediction {
    const float* out;
    const unsigned long long* scl;
    int N;
};
int unet_predict(pic_ctx* c, pic_unet* u, UnetPrediction* o);

// One Adam step on the pair (rho, phi_old) -> phi (the new solution).
int unet_train_pair(pic_ctx* c, pic_unet* u, const double* phi_old, float lr, float* loss);

} // namespace pic
