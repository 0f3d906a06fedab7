// C++ facade of the MI355X c-proj hot path: the reference's class API
// (SpatialGrid, AngularGrid, OrientationProjectionMap,
// OrientationProjectionTransform, Convolution, get_delta_rho, get_delta_f),
// the fused excess evaluation, the full functional with its minimiser and
// the batched server, over the C ABI of mdft_hip.h.  Header only; link libmdft_hip.so and the
// HIP runtime.
#pragma once

#include "Core.hpp"
#include "Grid.hpp"
#include "OrientationProjectionTransform.hpp"
#include "Convolution.hpp"
#include "Solvent.hpp"
#include "Functional.hpp"
#include "Server.hpp"
