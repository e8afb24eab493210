// The 2-D `geometry` members of the reference's compiled-out projector
// (2d/mat_mult.h:135-158), for the overload-layer tests.
#pragma once
#include "ref_types_common.h"

struct geometry {
  int num_emitters = 0, num_receivers = 0;
  const bool *use_path = nullptr;          // [e * R + r]
  const int *x_emitters = nullptr, *y_emitters = nullptr;
  const int *x_receivers = nullptr, *y_receivers = nullptr;
  int ld = 0;                              // voxel x + y * ld
  float scale_factor = 1.0f;
};
