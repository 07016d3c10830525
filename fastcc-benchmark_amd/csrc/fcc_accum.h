// Planner / accumulate entry points (fcc_accum.hip).
#pragma once
#include "fcc_index.h"

namespace fcc {

struct PlanParams {
    int force;           // 0 or a forced KIND_*
    uint32_t TL;         // rows per left tile
    uint64_t spanB;      // output columns
    uint32_t Wd;         // dense column window when TL*spanB exceeds the LDS tile
    float margin;        // learned planner: relative safety margin on predictions
    float dense_ratio;   // dense when estimated outputs >= dense_ratio * TL * spanB
    uint64_t run;        // products per unit of consecutive DIRECT tiles (0: one tile per unit)
    uint64_t group_rows; // DIRECT runs break at multiples of this row (0: no groups)
    int mirror;          // mirror mode: windows wholly left of a tile's first row are skipped
    const uint32_t *atile_id;  // tile index of each active tile (mirror mode)
    int logTL;
};

void launch_plan_tiles(const Index &ix, const float *pred, const PlanParams &P, uint32_t *t_nwin, uint32_t *t_W,
                       uint8_t *t_kind, uint32_t *t_w0, hipStream_t s);
void launch_emit_units(const uint32_t *t_nwin, const uint32_t *t_W, const uint8_t *t_kind, const uint32_t *t_ubase,
                       const uint32_t *t_w0, uint32_t natiles, uint64_t spanB, const float *pred, Units &U,
                       hipStream_t s);
void launch_refine_kinds(const unsigned long long *unit_count, const uint32_t *u_W, uint32_t n, int force, int mirror,
                         uint32_t TL, uint8_t *u_kind, unsigned long long *gtotal, hipStream_t s);
void launch_count(const AccArgs &A, unsigned grid, hipStream_t s);
// window cuts of every fiber for dense windows of width W (n = windows + 1)
void launch_fiber_cuts(const Index &ix, uint32_t n, uint32_t W, uint32_t *fcut, hipStream_t s);
void launch_numeric(const AccArgs &A, unsigned grid, bool bitmap, bool sink, int mode, hipStream_t s);
void launch_gdrain(const AccArgs &A, bool sink, int mode, hipStream_t s);

}  // namespace fcc
