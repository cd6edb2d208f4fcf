// scatterlayer: plane waves through a stack of homogeneous layers (see cwx_layer.hip).
#pragma once

#include "cwx_common.h"

namespace cwx {

// Solves the stack for P incident waves.  d_ri: L complex indices (device), d_z: L - 1 boundaries
// (device), d_inc: P records {E0[3], k[3]}; d_refl / d_trans receive P x (L-1) records.  Returns the
// number of systems with an exactly zero pivot.
int layer_solve_dev(int L, const zc* d_ri, const double* d_z, double lambda, long P, const zc* d_inc, zc* d_refl,
                    zc* d_trans, hipStream_t s);

}  // namespace cwx
