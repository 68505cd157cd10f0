// pic_unet.h — the learned initial guess (U-Net) as used by the Poisson stage.
#pragma once

#include "../../include/bmp_pic.h"

struct pic_ctx;

namespace pic {

// guess <- phi + U-Net(rho, phi) on the N x N node image (phi = the previous
// solution, rho = the right-hand side after the Dirichlet overwrite).
int unet_guess(pic_ctx* c, pic_unet* u, double* guess);

// One Adam step on the pair (rho, phi_old) -> phi (the new solution).
int unet_train_pair(pic_ctx* c, pic_unet* u, const double* phi_old, float lr, float* loss);

} // namespace pic
