// Learned-planner MLP entry points (fcc_mlp.hip).
#pragma once
#include "fcc_index.h"

namespace fcc {

constexpr int MLP_IN = 6, MLP_HID = 64, MLP_OUT = 2;
constexpr int MLP_PARAMS = MLP_IN * MLP_HID + MLP_HID + MLP_HID * MLP_HID + MLP_HID + MLP_HID * MLP_OUT + MLP_OUT +
                           2 * MLP_IN + 2 * MLP_OUT;

// feat: natiles x MLP_IN
void mlp_tile_features(const Index &ix, uint64_t spanB, float *feat, hipStream_t s);
void mlp_infer(const float *dparams, const float *feat, uint32_t n, float *pred_nnz, float *pred_cols, hipStream_t s);
// grad (trainable prefix) and loss are zeroed, then filled; pred (n x MLP_OUT, normalised) optional
void mlp_fwd_bwd(const float *dparams, const float *feat, const float *label, const uint32_t *perm, uint32_t n,
                 float *grad, float *loss, float *pred, hipStream_t s);
void mlp_adam(float *dparams, const float *grad, float *m, float *v, float lr, int step, hipStream_t s);
int mlp_trainable_count();

}  // namespace fcc
