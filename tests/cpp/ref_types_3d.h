// The 3-D `geometry` members of the reference's live projector
// (3d/mat_mult.h:164-232; per-MP coordinate arrays), for the overload-layer tests.
#pragma once
#include "ref_types_common.h"

struct geometry {
  int numMPs = 0, numPaths = 0, num_emitters = 0, num_receivers = 0;
  const int *use_path = nullptr;           // [e * R + r], 0 = skip, k = row k - 1
  int **x_emitters = nullptr, **y_emitters = nullptr, **z_emitters = nullptr;
  int **x_receivers = nullptr, **y_receivers = nullptr, **z_receivers = nullptr;
  int rv_x = 0, rv_y = 0, rv_z = 0;
  float scale_factor = 1.0f;
};
